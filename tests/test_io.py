"""Langmuir's file formats (langmuir_amd.io) and the seeded world set-up
(langmuir_amd.world) on the CPU.

Pinned by: the reference's own files (examples/*.inp, the LangmuirPython test
run's out.chk / out.dat / out.parm, copied as data into tests/golden/ref_files);
hand-derived layouts of TypedVariable::keyValue and FluxWriter (widths and
order from variable.h:195-240, writer.cpp:18-43); the MT19937 known answer; and
the oracle's World construction (same carriers, defects, potential and engine
window after set-up).
"""
import gzip
import os

import numpy as np
import pytest

from langmuir_amd import io as L
from langmuir_amd import world as W
from _util import params_pair

REF = os.path.join(os.path.dirname(__file__), "golden", "ref_files")


def ref(name):
    return os.path.join(REF, name)


# ----------------------------------------------------------------------------- parsing
def test_reads_reference_checkpoint_fixture():
    ck = L.read_checkpoint(ref("system0_part0_out.chk"))
    assert ck.electrons == list(range(10))
    assert ck.traps == list(range(10))
    assert ck.trap_potentials == [1.0] * 10
    assert ck.holes is None and ck.defects is None and ck.flux is None and ck.random_state is None
    kv = ck.parameters
    assert (kv["iterations.print"], kv["current.step"], kv["voltage.right"]) == (1, 10, 0.0)
    assert (kv["grid.x"], kv["grid.y"], kv["grid.z"]) == (128, 128, 1)
    assert kv["simulation.type"] == "transistor"      # default kept


def test_reads_reference_parameter_files():
    p = L.read_checkpoint(ref("system0_part0_out.parm")).parameters
    assert (p["current.step"], p["grid.x"]) == (10, 128)
    s = L.read_checkpoint(ref("solar.inp")).parameters
    assert s["simulation.type"] == "solarcell"
    assert (s["grid.x"], s["grid.y"], s["grid.z"]) == (96, 96, 1)
    assert s["coulomb.carriers"] is True and s["output.is.on"] is False
    assert (s["electron.percentage"], s["hole.percentage"], s["voltage.right"]) == (0.10, 0.10, 7.50)
    assert (s["recombination.rate"], s["source.rate"], s["work.size"]) == (1e-4, 0.01, 256)
    # the fully commented sample: every value after its '#' comment is stripped
    a = L.read_checkpoint(ref("sample.inp")).parameters
    assert a["random.seed"] == 0 and a["output.coulomb"] == 0 and a["output.width"] == 23
    assert (a["iterations.print"], a["iterations.real"]) == (100, 1000)
    assert (a["trap.percentage"], a["seed.percentage"], a["voltage.right"]) == (0.50, 0.05, 9.00)
    assert a["source.scale.area"] == 65536.0 and a["e.drain.l.rate"] == -1.0
    a.check()
    t = L.read_checkpoint(ref("transistor.inp")).parameters
    t.check()
    assert t["simulation.type"] == "transistor"


def test_parse_rules():
    kv = L.KeyValues()
    kv.parse_line('  Grid.X = 64   # comment = with equal sign')
    kv.parse_line('output.stub = "run"')
    kv.parse_line('voltage.left =    ')            # empty value: skipped
    kv.parse_line('# only a comment')
    kv.parse_line('electrostatic.cutoff = 9')      # constant: ignored
    kv.parse_line('coulomb.carriers = 1')
    kv.parse_line('use.opencl = FALSE')
    assert kv["grid.x"] == 64 and kv["output.stub"] == "run" and kv["voltage.left"] == 0.0
    assert kv["coulomb.carriers"] is True and kv["use.opencl"] is False
    assert kv.ignored_constants == ["electrostatic.cutoff"]
    assert "electrostatic.cutoff" not in L.SimulationParameters().__dict__ or \
        kv.to_parameters().electrostatic_cutoff == 50
    for bad in ("grid.x 64", "grid.x = 1 = 2", " = 4", "grid.w = 4", "grid.x = 4.5", "grid.x = 1e3",
                "random.seed = -1", "current.step = 4294967296", "voltage.left = 1_0",
                "coulomb.carriers = yes"):
        with pytest.raises(L.ParseError):
            L.KeyValues().parse_line(bad)


def test_last_parameter_line_needs_a_newline():
    """std::getline at the end of the file: an unterminated last line is not parsed."""
    ck = L.parse_checkpoint("[Parameters]\ngrid.x = 16\ngrid.y = 8")
    assert ck.parameters["grid.x"] == 16 and ck.parameters["grid.y"] == 128
    ck = L.parse_checkpoint("[Parameters]\ngrid.x = 16\ngrid.y = 8\n")
    assert ck.parameters["grid.y"] == 8
    ck = L.parse_checkpoint("[Parameters]\ngrid.x = 16\nend\n[Electrons]\n2\n5 7\n")
    assert ck.parameters["grid.x"] == 16 and ck.electrons == [5, 7]


def test_section_errors():
    with pytest.raises(L.ParseError):
        L.parse_checkpoint("[Electrons]\n3\n1\n2\n")           # short list
    with pytest.raises(L.ParseError):
        L.parse_checkpoint("[Nonsense]\n0\n")
    with pytest.raises(L.ParseError):
        L.parse_checkpoint("[Holes]\n1\n-4\n")
    with pytest.raises(L.ParseError):
        L.parse_checkpoint("[RandomState]\n" + " ".join(["1"] * 623) + "\n")
    # the seed comes first (rand.cpp:366-381): 624 integers leave the twister one word short
    with pytest.raises(L.ParseError):
        L.parse_checkpoint("[RandomState]\n" + " ".join(["1"] * 624) + "\n[Parameters]\n")


# ----------------------------------------------------------------------------- writing
def test_key_value_layout():
    """keyValue: key left-justified in 30, '=' in a field of 3 (2 before a leading '-'),
    reals %.15e, bools True/False, an empty string as a bare '='."""
    kv = L.KeyValues()
    kv["voltage.left"] = -1.0
    kv["output.stub"] = ""
    lines = kv.format().split("\n")
    assert lines[0] == ""
    got = dict((ln[:30].strip(), ln) for ln in lines[1:])
    assert got["voltage.left"] == "voltage.left" + " " * 18 + "= -1.000000000000000e+00"
    assert got["voltage.right"] == "voltage.right" + " " * 17 + "=  0.000000000000000e+00"
    assert got["trap.potential"] == "trap.potential" + " " * 16 + "=  1.000000000000000e-01"
    assert got["output.is.on"] == "output.is.on" + " " * 18 + "=  True"
    assert got["output.stub"] == "output.stub" + " " * 19 + "="
    assert got["max.threads"] == "max.threads" + " " * 19 + "= -1"
    assert got["source.scale.area"].endswith("=  6.553600000000000e+04")
    assert len(lines) - 1 == sum(1 for *_, c in L.PARAMETER_TABLE if not c) == 69
    assert [ln[:30].strip() for ln in lines[1:4]] == ["simulation.type", "current.step", "iterations.real"]


def test_checkpoint_round_trip(tmp_path):
    rng = np.random.default_rng(7)
    kv = L.KeyValues()
    for line in ("grid.x = 16", "grid.y = 12", "grid.z = 3", "output.chk.trap.potential = true",
                 "voltage.right = -0.25", "current.step = 40"):
        kv.parse_line(line)
    ck = L.Checkpoint(parameters=kv, electrons=[3, 1, 99], holes=[], defects=[5],
                      traps=[7, 8], trap_potentials=[0.1, 1 / 3], flux=list(range(20)),
                      random_state=[int(w) for w in rng.integers(0, 2 ** 32, 624)], random_seed=424242)
    text = L.format_checkpoint(ck)
    assert text.startswith("[Electrons]\n3\n3\n1\n99\n[Holes]\n0\n[Defects]\n1\n5\n[Traps]\n2\n7\n8\n"
                           "[TrapPotentials]\n2\n1.00000000000000006e-01\n3.33333333333333315e-01\n"
                           "[FluxState]\n20\n0\n1\n")
    assert "\n[RandomState]\n424242 " + " ".join(map(str, ck.random_state)) + "\n[Parameters]\n" in text
    assert not text.endswith("\n")
    path = tmp_path / "a.chk"
    L.write_checkpoint(str(path), ck)
    back = L.read_checkpoint(str(path))
    for f in ("electrons", "holes", "defects", "traps", "trap_potentials", "flux", "random_state",
              "random_seed"):
        assert getattr(back, f) == getattr(ck, f), f
    # max.threads is the last line and has no newline after it: the reference drops it too
    assert {k: v for k, v in back.parameters.values.items() if k != "max.threads"} == \
        {k: v for k, v in kv.values.items() if k != "max.threads"}
    gz = tmp_path / "a.chk.gz"
    gz.write_bytes(gzip.compress(path.read_bytes()))
    assert L.read_checkpoint(str(gz)).random_state == ck.random_state


def test_flux_table_layout(tmp_path):
    """Header and rows of out.dat byte for byte against the reference run's file."""
    with open(ref("system0_part0_out.dat")) as f:
        ref_lines = f.read().split("\n")
    head, rows = L.read_flux_table(ref("system0_part0_out.dat"))
    kv = L.KeyValues()
    path = tmp_path / "out.dat"
    w = L.FluxWriter(str(path), kv)
    for r in rows:
        w.write(dict(current_step=r[0], flux_attempts=r[1:21:2], flux_successes=r[2:21:2],
                     e_site=[0] * r[21], h_site=[0] * r[22]), real_ms=r[23])
    w.close()
    assert path.read_text().split("\n") == ref_lines
    assert head[:3] == ["simulation:time", "eSourceL:attempt", "eSourceL:success"] and len(head) == 24


def test_coulomb_layout(tmp_path):
    dims = (3, 2, 2)
    v = np.arange(12, dtype=np.float64) * -0.5 + 1.0
    path = tmp_path / "m.coulomb"
    L.write_coulomb(str(path), dims, v, precision=4)
    lines = path.read_text().split("\n")
    assert lines[0] == "s x y z v" and lines[-1] == ""
    assert lines[1] == "0 0 0 0 1.0000e+00" and lines[2] == "6 0 0 1 -2.0000e+00"
    assert lines[3] == "3 0 1 0 -5.0000e-01"
    s, vals = L.read_coulomb(str(path))
    assert sorted(s.tolist()) == list(range(12))
    assert np.array_equal(vals, v[s])


def test_output_names():
    kv = L.KeyValues()
    kv["current.step"] = 250
    assert L.output_name("%stub-%step.coulomb", kv) == "out-250.coulomb"
    assert L.output_name("%STUB.chk", kv) == "out.chk"
    kv["output.stub"] = ""
    assert L.output_name("%stub-%step.coulomb", kv) == "250.coulomb"
    kv["output.stub"] = "a--b-"
    assert L.output_name("%stub-%step.coulomb", kv, "/tmp/x") == "/tmp/x/a-b-250.coulomb"
    with pytest.raises(ValueError):
        kv["output.stub"] = "-"
        L.output_name("%stub.dat", kv)


def test_parameter_checks():
    kv = L.KeyValues()
    kv.check()
    for key, val in (("iterations.real", 1005), ("trap.percentage", 1.5), ("hopping.range", 3),
                     ("recombination.rate", 0.1), ("balance.charges", True), ("output.xyz.mode", 2),
                     ("e.drain.l.rate", 1.5)):
        bad = L.KeyValues()
        bad[key] = val
        with pytest.raises(L.ParseError):
            bad.check()


# ----------------------------------------------------------------------------- engine
def test_host_engine_known_answer_and_oracle(oracle):
    mt = W.MT19937(5489)
    out = [mt() for _ in range(10000)]
    assert out[-1] == 4123659995                      # C++11 [rand.predef]
    assert np.array_equal(np.array(out[:3000], np.uint32), oracle.mt_outputs(5489, 3000))
    assert mt.consumed == 10000
    # the window is x[k .. k+623]: a new engine started from it continues the stream
    again = W.MT19937(window=mt.window())
    assert [again() for _ in range(700)] == oracle.mt_outputs(5489, 700, skip=10000).tolist()
    u = W.MT19937(11)
    assert [u.integer(0, 6) for _ in range(500)] == oracle.uniform_int(11, 0, 6, 500).tolist()
    u = W.MT19937(12)
    assert [u.integer(-3, 65535) for _ in range(500)] == oracle.uniform_int(12, -3, 65535, 500).tolist()
    assert W.MT19937(3).integer(4, 4) == 4


# ----------------------------------------------------------------------------- world set-up
WORLDS = {
    "transistor_seeded": dict(grid_x=32, grid_y=24, grid_z=1, hole_percentage=0.05, electron_percentage=0.0,
                              seed_charges=0.7, e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9,
                              random_seed=5),
    "solar_grown_traps": dict(simulation_type=1, grid_x=20, grid_y=16, grid_z=6, electron_percentage=0.02,
                              hole_percentage=0.03, seed_charges=1.0, trap_percentage=0.3,
                              seed_percentage=0.05, trap_potential=0.5, voltage_right=1.5, random_seed=77),
    "defects_and_traps": dict(grid_x=24, grid_y=24, grid_z=2, electron_percentage=0.04, seed_charges=0.5,
                              defect_percentage=0.03, trap_percentage=0.1, seed_percentage=1.0,
                              coulomb_carriers=1, defects_charge=-1, random_seed=901),
    "gaussian_trap_depths": dict(grid_x=16, grid_y=12, grid_z=3, electron_percentage=0.03, seed_charges=1.0,
                                 trap_percentage=0.15, seed_percentage=0.3, trap_potential=0.3,
                                 gaussian_stdev=0.1, random_seed=4242),
}


@pytest.mark.parametrize("name", sorted(WORLDS))
def test_world_setup_matches_oracle(oracle, name):
    op, sp = params_pair(**WORLDS[name])
    ow = oracle.World(op)
    setup = W.initialize(sp)
    assert setup.electrons == ow.carriers(0)["site"].tolist()
    assert setup.holes == ow.carriers(1)["site"].tolist()
    assert setup.defects == ow.defects().tolist()
    assert np.array_equal(setup.potential, ow.potential())
    win, k = ow.rng()
    assert setup.rng_consumed == k and np.array_equal(setup.rng_window, win)
    assert len(setup.traps) == int(sp.trap_percentage * sp.volume)


def test_world_setup_with_listed_sites(oracle):
    kw = dict(simulation_type=1, grid_x=16, grid_y=16, grid_z=2, electron_percentage=0.05, hole_percentage=0.05,
              seed_charges=1.0, trap_percentage=0.2, seed_percentage=0.5, defect_percentage=0.02,
              random_seed=31)
    op, sp = params_pair(**kw)
    e, h, d, t = [3, 40, 41], [3, 100], [7, 9], [11, 12, 200]
    tp = [0.2, 0.3, 0.4]
    ow = oracle.World(op, electrons=e, holes=h, defects=d, traps=t, trap_potentials=tp)
    setup = W.initialize(sp, electrons=e, holes=h, defects=d, traps=t, trap_potentials=tp)
    assert setup.electrons[:3] == e and setup.holes[:2] == h and setup.defects[:2] == d
    assert setup.electrons == ow.carriers(0)["site"].tolist()
    assert setup.holes == ow.carriers(1)["site"].tolist()
    assert setup.defects == ow.defects().tolist()
    assert np.array_equal(setup.potential, ow.potential())
    win, k = ow.rng()
    assert setup.rng_consumed == k and np.array_equal(setup.rng_window, win)


def test_world_setup_refusals():
    _, sp = params_pair(grid_x=8, grid_y=8, electron_percentage=0.05, random_seed=2)
    with pytest.raises(ValueError):
        W.initialize(sp, electrons=[1, 2, 3, 4])          # int(0.05 * 64) = 3 allowed
    with pytest.raises(ValueError):
        W.initialize(sp, electrons=[5, 5])
    _, sp = params_pair(grid_x=8, grid_y=8, hole_percentage=0.1, electron_percentage=0.0, random_seed=2,
                        e_source_l_rate=0.0, h_source_l_rate=0.9)
    with pytest.raises(ValueError, match="hole drain"):
        W.initialize(sp)


def test_setup_of_checkpoint_resumes_engine():
    """A [RandomState] replaces random.seed; listed carriers come first."""
    kv = L.KeyValues()
    for line in ("grid.x = 16", "grid.y = 16", "electron.percentage = 0.05", "seed.charges = 1.0",
                 "random.seed = 9"):
        kv.parse_line(line)
    mt = W.MT19937(1234)
    for _ in range(1000):
        mt()
    ck = L.Checkpoint(parameters=kv, electrons=[0, 1], random_state=mt.window().tolist(), random_seed=1234)
    setup = L.setup_of(ck)
    assert setup.rng_seed == 1234
    assert setup.electrons[:2] == [0, 1] and len(setup.electrons) == int(0.05 * 256)
    # the seeded sites are the engine's draws from the restored state
    again = W.MT19937(window=ck.random_state)
    occ = {0, 1}
    expect = [0, 1]
    while len(expect) < len(setup.electrons):
        s = again.integer(0, 255)
        if s not in occ:
            occ.add(s)
            expect.append(s)
    assert setup.electrons == expect
    kw = L.simulation_kwargs(ck, setup)
    assert np.array_equal(kw["rng_window"], again.window()) and kw["rng_consumed"] == again.consumed


def reference_random_state(seed, blocks):
    """[RandomState] text in the reference's layout, built without this package: the
    seed (Random::operator<<, rand.cpp:359-364), then the twister's text state, which
    the C++ standard fixes as the last 624 raw words, oldest first.  numpy's legacy
    MT19937 regenerates its 624-word key in place like Boost, so right after it has
    handed out a whole number of blocks (pos == 624) its key is exactly those words."""
    rs = np.random.RandomState(seed)
    rs.randint(0, 2 ** 32, size=624 * blocks, dtype=np.uint64)
    name, key, pos = rs.get_state()[:3]
    assert name == "MT19937" and pos == 624
    text = f"[RandomState]\n{seed} " + " ".join(str(int(w)) for w in key) + "\n"
    return text, key.astype(np.uint32), rs


@pytest.mark.parametrize("seed,blocks", [(7, 0), (2024, 1), (4294967295, 3)])
def test_reference_random_state_resumes_the_stream(seed, blocks):
    """A 625-integer [RandomState] written the reference's way parses into the seed
    and the engine window, and the engine resumed from it continues numpy's stream;
    the checkpoint written back carries the same text."""
    text, key, rs = reference_random_state(seed, blocks)
    kv = L.KeyValues()
    for line in ("grid.x = 16", "grid.y = 8", "hole.percentage = 0.05", "seed.charges = 1.0",
                 "random.seed = 99", "e.source.l.rate = 0.0", "h.source.l.rate = 0.9", "h.drain.r.rate = 0.9"):
        kv.parse_line(line)
    ck = L.parse_checkpoint(text + "[Parameters]\n" + "".join(f"{ln}\n" for ln in
                            ("grid.x = 16", "grid.y = 8", "random.seed = 99")))
    assert ck.random_seed == seed and np.array_equal(np.asarray(ck.random_state, np.uint32), key)
    mt = W.MT19937(window=ck.random_state)
    expect = rs.randint(0, 2 ** 32, size=1500, dtype=np.uint64)
    assert [mt() for _ in range(1500)] == [int(v) for v in expect]
    # written back: the same section, seed first
    ck2 = L.Checkpoint(parameters=kv, random_state=ck.random_state, random_seed=ck.random_seed)
    out = L.format_checkpoint(ck2)
    assert text.rstrip("\n") + "\n[Parameters]" in out
    # a set-up from the file keeps the file's seed, not random.seed
    ck3 = L.Checkpoint(parameters=kv, random_state=ck.random_state, random_seed=ck.random_seed)
    assert L.setup_of(ck3).rng_seed == seed


def test_checkpoint_of_simulation_writes_the_seed():
    """Without a [RandomState] the engine is seeded from random.seed, and m_seed, the
    word written before the twister, is random.seed cast to 32 bits (rand.cpp:36-47)."""
    kv = L.KeyValues()
    for line in ("grid.x = 16", "grid.y = 8", "hole.percentage = 0.05", "seed.charges = 1.0",
                 "random.seed = 4294967301", "e.source.l.rate = 0.0", "h.source.l.rate = 0.9",
                 "h.drain.r.rate = 0.9"):
        kv.parse_line(line)
    setup = L.setup_of(L.Checkpoint(parameters=kv))
    assert setup.rng_seed == 5
    state = dict(e_site=np.zeros(0, np.int32), h_site=np.asarray(setup.holes, np.int32),
                 flux_attempts=np.zeros(10, np.uint64), flux_successes=np.zeros(10, np.uint64),
                 rng_window=setup.rng_window, current_step=0)
    text = L.format_checkpoint(L.checkpoint_of(state, kv, setup))
    words = text.split("[RandomState]\n")[1].split("\n")[0].split()
    assert len(words) == 625 and words[0] == "5"
    assert [int(w) for w in words[1:]] == [int(w) for w in setup.rng_window]


def test_gaussian_trap_depths_box_muller():
    """gaussian.stdev > 0: after the drawn traps are placed, each gets trap.potential plus
    one normal deviate, in list order (potential.cpp:211-220).  Each deviate is a fresh
    Boost (1.44-1.55) normal_distribution: two uniform_01 draws r1, r2 (u32 * 2^-32) and
    sqrt(-2 ln(1 - r2)) cos(2 pi r1) sigma.  Hand-derived from numpy's MT19937 words: the
    engine position after the placement is known from a run without the deviation."""
    kw = dict(grid_x=10, grid_y=10, grid_z=1, electron_percentage=0.02, seed_charges=1.0,
              trap_percentage=0.08, seed_percentage=1.0, trap_potential=0.25, random_seed=31)
    _, flat = params_pair(**kw)
    _, gauss = params_pair(**{**kw, "gaussian_stdev": 0.07})
    a, b = W.initialize(flat), W.initialize(gauss)
    assert a.traps == b.traps and len(a.traps) == 8
    words = np.random.RandomState(31).randint(0, 2 ** 32, size=a.rng_consumed + 16, dtype=np.uint64)
    u = words[a.rng_consumed:].astype(np.float64) / 2.0 ** 32
    expect = [0.25 + np.sqrt(-2.0 * np.log(1.0 - u[2 * i + 1])) * np.cos(2.0 * np.pi * u[2 * i]) * 0.07
              for i in range(8)]
    assert np.allclose(b.trap_potentials, expect, rtol=0, atol=1e-15)
    assert b.rng_consumed == a.rng_consumed + 16
    assert not np.allclose(b.trap_potentials, 0.25)


@pytest.mark.parametrize("X", [48, 96, 33, 100, 128])
def test_background_potential_rounds_like_the_reference(oracle, X):
    """setPotentialLinear + setPotentialGate (potential.cpp:27-72): v = m (i + 0.5) + b with
    m = (VR - VL) / X, then + slope.z (k + 0.5), each a product and a sum rounded
    separately (no fused multiply-add).  The host's map and the oracle's are those bits."""
    _, sp = params_pair(grid_x=X, grid_y=3, grid_z=5, electron_percentage=0.0, hole_percentage=0.05,
                        seed_charges=1.0, voltage_left=1.0, voltage_right=-0.3, slope_z=0.1,
                        e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9, random_seed=3)
    m = (-0.3 - 1.0) / float(X)
    expect = np.zeros((5, 3, X))
    for i in range(X):
        for k in range(5):
            expect[k, :, i] = (0.0 + (m * (i + 0.5) + 1.0)) + 0.1 * (k + 0.5)
    from langmuir_amd.host import background_potential
    assert np.array_equal(background_potential(sp), expect.ravel())
    op, _ = params_pair(grid_x=X, grid_y=3, grid_z=5, electron_percentage=0.0, hole_percentage=0.05,
                        seed_charges=1.0, voltage_left=1.0, voltage_right=-0.3, slope_z=0.1,
                        e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9, random_seed=3)
    assert np.array_equal(oracle.World(op).potential(), expect.ravel())
