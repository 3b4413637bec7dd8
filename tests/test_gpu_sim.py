"""Device KMC sweep against the FP64 oracle.

* Proposals (chooseFuture) are bit-exact: same neighbour order, same MT19937
  words, same Boost bucket rule.
* dE agrees with the oracle within 1e-5 |dE| + 2e-7 eV (FP32 compensated sums).
* Steps: after every device step the full state (lists in order, lifetimes,
  path lengths, flux counters, engine position and window, step counter) equals
  the oracle's.  A step may differ only if the oracle saw a Metropolis decision
  whose margin |c exp(-beta pd) - u| lies inside the band that the dE tolerance
  opens (beta * c * tol(dE)); the device is then re-synchronised from the oracle
  and the count of such steps is reported.
"""
import numpy as np
import pytest

from _util import compare_state, de_tolerance, gpu_from_oracle, params_pair, resync

pytestmark = pytest.mark.gpu

CONFIGS = {
    "transistor2d": dict(grid_x=64, grid_y=64, grid_z=1, coulomb_carriers=1, electron_percentage=0.0,
                         hole_percentage=0.02, seed_charges=1.0, voltage_left=1.0,
                         e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9,
                         random_seed=11),
    "solar3d": dict(simulation_type=1, grid_x=16, grid_y=16, grid_z=8, coulomb_carriers=1,
                    electron_percentage=0.02, hole_percentage=0.02, seed_charges=0.5,
                    generation_rate=0.5, recombination_rate=0.3, recombination_range=1,
                    exciton_binding=0.1, electrostatic_cutoff=30, random_seed=12),
    "gauss3d": dict(simulation_type=1, grid_x=12, grid_y=12, grid_z=12, coulomb_carriers=1,
                    coulomb_gaussian_sigma=1.0, electron_percentage=0.02, hole_percentage=0.02,
                    seed_charges=1.0, generation_rate=0.3, random_seed=13),
    "defects": dict(grid_x=24, grid_y=16, grid_z=4, coulomb_carriers=1, electron_percentage=0.0,
                    hole_percentage=0.03, seed_charges=1.0, defect_percentage=0.02,
                    defects_charge=-1, voltage_left=0.5, e_source_l_rate=0.0,
                    h_source_l_rate=0.9, h_drain_r_rate=0.9, random_seed=14),
    "hop2": dict(grid_x=20, grid_y=14, grid_z=6, coulomb_carriers=1, hopping_range=2,
                 electron_percentage=0.0, hole_percentage=0.03, seed_charges=1.0,
                 voltage_left=1.0, e_source_l_rate=0.0, h_source_l_rate=0.9,
                 h_drain_r_rate=0.9, random_seed=15, electrostatic_cutoff=12),
    "source_coulomb": dict(grid_x=32, grid_y=16, grid_z=2, coulomb_carriers=1,
                           electron_percentage=0.02, hole_percentage=0.02, seed_charges=0.5,
                           voltage_left=0.3, e_source_l_rate=0.5, e_source_r_rate=0.4,
                           h_source_l_rate=0.6, h_source_r_rate=0.3, e_drain_l_rate=0.5,
                           e_drain_r_rate=0.9, h_drain_l_rate=0.7, h_drain_r_rate=0.9,
                           source_metropolis=1, source_coulomb=1, random_seed=16),
    "balance": dict(simulation_type=1, grid_x=20, grid_y=10, grid_z=4, coulomb_carriers=1,
                    electron_percentage=0.03, hole_percentage=0.03, seed_charges=0.3,
                    generation_rate=0.4, balance_charges=1, voltage_right=0.2,
                    e_source_l_rate=0.5, e_source_r_rate=0.5, h_source_l_rate=0.5,
                    h_source_r_rate=0.5, e_drain_l_rate=0.9, h_drain_l_rate=0.2,
                    e_drain_r_rate=0.2, h_drain_r_rate=0.9, random_seed=17),
    "coulomb_off": dict(grid_x=40, grid_y=40, grid_z=1, coulomb_carriers=0, electron_percentage=0.0,
                        hole_percentage=0.05, seed_charges=1.0, voltage_left=1.0,
                        e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9,
                        random_seed=18),
}


def make(oracle, gpu, name, de_path="auto", **over):
    kw = dict(CONFIGS[name])
    kw.update(over)
    op, sp = params_pair(**kw)
    w = oracle.World(op)
    sim = gpu_from_oracle(gpu, w, sp, de_path=de_path)
    sim.set_synchronous(True)
    if de_path != "auto":
        assert sim.de_path == (de_path if kw.get("coulomb_carriers") else "pairs")
    return w, sim, sp


# every configuration through both dE paths: all pairs, and the culled brick
# lists (forced here on lattices the automatic choice would run all-pairs)
PATHS = ("pairs", "culled")


@pytest.mark.parametrize("de_path", PATHS)
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_proposals_and_delta_e(gpu, oracle, name, de_path):
    w, sim, sp = make(oracle, gpu, name, de_path)
    fe, fh, de, dh = sim.delta_e()
    w.propose()
    w.coulomb()
    e, h = w.carriers(0), w.carriers(1)
    assert np.array_equal(fe, e["fsite"]) and np.array_equal(fh, h["fsite"])
    for got, ref in ((de, e["de"]), (dh, h["de"])):
        err = np.abs(got - ref)
        assert np.all(err <= de_tolerance(ref)), f"max dE err {err.max():.3e}"


def run_parity(w, sim, steps, beta, threads=1):
    mismatched = 0
    for t in range(steps):
        w.propose()
        w.coulomb(threads)
        ne, nh = w.info()["n_e"], w.info()["n_h"]
        de_e, de_h = w.carriers(0)["de"], w.carriers(1)["de"]
        w.finish()
        sim.step(1)
        diffs = compare_state(sim.state(), w)
        if diffs:
            near = False
            for which, n, de in ((0, ne, de_e), (1, nh, de_h)):
                u, m, pd = w.last_decisions(which, n)
                band = beta * (1.0 / 3.0) * de_tolerance(de) + 1e-12
                near |= bool(np.any((u >= 0) & (np.abs(m) <= band)))
            assert near, f"step {t}: device state differs in {diffs} with no near-threshold decision"
            mismatched += 1
            resync(sim, w)
    return mismatched


@pytest.mark.parametrize("de_path", PATHS)
@pytest.mark.parametrize("name,steps", [("transistor2d", 200), ("solar3d", 200), ("gauss3d", 80),
                                        ("defects", 150), ("hop2", 120), ("source_coulomb", 150),
                                        ("balance", 150), ("coulomb_off", 200)])
def test_step_parity(gpu, oracle, name, steps, de_path):
    w, sim, sp = make(oracle, gpu, name, de_path)
    mism = run_parity(w, sim, steps, sp.inverse_kt)
    print(f"{name}: {steps} steps, {mism} near-threshold resyncs")
    assert mism <= max(2, steps // 50)
    assert sim.device_error() == 0


@pytest.mark.parametrize("de_path", PATHS)
def test_transistor_full_size_parity(gpu, oracle, de_path):
    """The 256x256x1 transistor (1310 holes, 1 V, cutoff 50) for 25 steps."""
    w, sim, sp = make(oracle, gpu, "transistor2d", de_path, grid_x=256, grid_y=256, random_seed=2024)
    assert w.info()["n_h"] == 1310
    mism = run_parity(w, sim, 25, sp.inverse_kt)
    assert mism <= 2


@pytest.mark.parametrize("name,over", [
    # axes longer than twice the cutoff: x only (bricks ordered x-major, y and z fastest)
    ("defects", dict(electrostatic_cutoff=10)),
    # x and y (a Hilbert curve over 5 x 3 bricks, inside an 8 x 8 square)
    ("defects", dict(grid_x=40, grid_y=24, electrostatic_cutoff=8)),
    # all three (Morton order)
    ("solar3d", dict(grid_x=24, grid_y=24, grid_z=24, electrostatic_cutoff=8)),
])
def test_culled_brick_orders(gpu, oracle, name, over):
    """The culled dE path orders its bricks, and so its target groups, by the
    lattice's shape against the cutoff; every order gives the oracle's dE and
    trajectory."""
    w, sim, sp = make(oracle, gpu, name, "culled", **over)
    fe, fh, de, dh = sim.delta_e()
    w.propose()
    w.coulomb()
    e, h = w.carriers(0), w.carriers(1)
    assert np.array_equal(fe, e["fsite"]) and np.array_equal(fh, h["fsite"])
    for got, ref in ((de, e["de"]), (dh, h["de"])):
        assert np.all(np.abs(got - ref) <= de_tolerance(ref))
    w, sim, sp = make(oracle, gpu, name, "culled", **over)
    assert run_parity(w, sim, 40, sp.inverse_kt) <= 2
    assert sim.device_error() == 0


@pytest.mark.parametrize("name,cutoff", [
    # the largest cutoff whose pair-function table fits (c^2 + 1 = 3970 entries)
    ("transistor2d", 63), ("gauss3d", 63),
    # one more: the culled kernel evaluates 1/r (and erf) itself
    ("transistor2d", 64), ("gauss3d", 64),
])
def test_culled_table_limit(gpu, oracle, name, cutoff):
    """The culled dE kernel looks the pair function up in an LDS table indexed by
    r^2 for cutoffs up to 63 and evaluates it in the kernel beyond; both forms,
    Coulomb and erf-screened, give the oracle's dE and trajectory."""
    w, sim, sp = make(oracle, gpu, name, "culled", electrostatic_cutoff=cutoff)
    fe, fh, de, dh = sim.delta_e()
    w.propose()
    w.coulomb()
    e, h = w.carriers(0), w.carriers(1)
    assert np.array_equal(fe, e["fsite"]) and np.array_equal(fh, h["fsite"])
    for got, ref in ((de, e["de"]), (dh, h["de"])):
        assert np.all(np.abs(got - ref) <= de_tolerance(ref))
    w, sim, sp = make(oracle, gpu, name, "culled", electrostatic_cutoff=cutoff)
    assert run_parity(w, sim, 40, sp.inverse_kt) <= 2
    assert sim.device_error() == 0


def test_solar_full_size_delta_e(gpu, oracle):
    """128x128x32 with 5242 electrons + 5242 holes, cutoff 30: proposals bit-exact,
    dE within tolerance for all 10,484 carriers."""
    w, sim, sp = make(oracle, gpu, "solar3d", grid_x=128, grid_y=128, grid_z=32,
                      electron_percentage=0.01, hole_percentage=0.01, seed_charges=1.0,
                      recombination_rate=0.0, generation_rate=0.524288, random_seed=77)
    assert sim.de_path == "culled"   # the automatic choice at this size
    fe, fh, de, dh = sim.delta_e()
    w.propose()
    w.coulomb(threads=8)
    e, h = w.carriers(0), w.carriers(1)
    assert len(fe) == 5242 and len(fh) == 5242
    assert np.array_equal(fe, e["fsite"]) and np.array_equal(fh, h["fsite"])
    ref = np.concatenate([e["de"], h["de"]])
    got = np.concatenate([de, dh])
    err = np.abs(got - ref)
    rel = err / np.maximum(np.abs(ref), 1e-12)
    print(f"solar full size: max |dE err| {err.max():.3e} eV, median rel {np.median(rel):.3e}")
    assert np.all(err <= de_tolerance(ref))


def test_solar_full_size_invariants(gpu, oracle):
    """Size-independent properties at the 128x128x32 size: determinism of two
    runs, unique sites, conservation of carriers through the flux counters."""
    kw = dict(grid_x=128, grid_y=128, grid_z=32, electron_percentage=0.01, hole_percentage=0.01,
              seed_charges=1.0, recombination_rate=0.0, generation_rate=0.524288, random_seed=78)
    w, a, sp = make(oracle, gpu, "solar3d", **kw)
    b = gpu_from_oracle(gpu, w, sp)
    s0 = a.state()
    a.step(30)
    b.step(30)
    sa, sb = a.state(), b.state()
    for k in ("e_site", "h_site", "rng_window", "flux_attempts", "flux_successes"):
        assert np.array_equal(sa[k], sb[k])
    assert len(np.unique(sa["e_site"])) == len(sa["e_site"])
    assert len(np.unique(sa["h_site"])) == len(sa["h_site"])
    ds = (sa["flux_successes"] - s0["flux_successes"]).astype(np.int64)
    assert len(sa["e_site"]) == len(s0["e_site"]) + ds[4] - ds[5] - ds[6] - ds[9]
    assert len(sa["h_site"]) == len(s0["h_site"]) + ds[4] - ds[7] - ds[8] - ds[9]
    assert sa["current_step"] == 30
    assert a.device_error() == 0


@pytest.mark.parametrize("name,de_path", [("solar3d", "culled"), ("transistor2d", "pairs"), ("defects", "culled")])
def test_graph_replay_matches_eager(gpu, oracle, name, de_path):
    """lmg_sim_step replays one captured step as a HIP graph (the default); kernel-by-kernel
    launches give the same state after every call, with host reads and eager work between
    the replays."""
    w, a, sp = make(oracle, gpu, name, de_path)
    b = gpu_from_oracle(gpu, w, sp, de_path=de_path)
    a.set_synchronous(False)
    b.set_graphs(False)
    for n in (1, 3, 1, 6, 2, 9):
        a.step(n)
        b.step(n)
        sa, sb = a.state(), b.state()
        for k in ("e_site", "h_site", "e_lifetime", "h_lifetime", "e_pathlength", "h_pathlength",
                  "rng_window", "flux_attempts", "flux_successes"):
            assert np.array_equal(sa[k], sb[k]), f"{name} after {n} more steps: {k}"
        assert sa["rng_consumed"] == sb["rng_consumed"] and sa["current_step"] == sb["current_step"]
        a.delta_e()   # eager dE-only work between replays
    assert a.device_error() == 0 and b.device_error() == 0


@pytest.mark.parametrize("name", ["solar3d", "defects", "gauss3d"])
def test_all_pairs_de_independent_of_chunking(gpu, oracle, name):
    """The all-pairs dE kernel adds exact fixed-point sums of the same 32-charge blocks
    whatever its charge-chunk count, so the autotuner's choice changes no bit."""
    kw = dict(CONFIGS[name])
    op, sp = params_pair(**kw)
    w = oracle.World(op)
    res = []
    for chunks in (1, 3, 7, 64, 0):
        sim = gpu_from_oracle(gpu, w, sp, de_path="pairs", de_chunks=chunks)
        res.append(sim.delta_e()[2:])
        t = sim.tuning
        assert t["path"] == "pairs" and t["charge_chunks"] >= 1
        if chunks:
            assert t["charge_chunks"] == chunks
        else:   # autotuned: a power of two with its time measured
            assert t["charge_chunks"] & (t["charge_chunks"] - 1) == 0 and t["de_kernel_us"] > 0
    for de, dh in res[1:]:
        assert np.array_equal(de, res[0][0]) and np.array_equal(dh, res[0][1])


@pytest.mark.parametrize("name", ["solar3d", "defects"])
def test_culled_de_independent_of_workgroups(gpu, oracle, name):
    """The culled dE kernel adds exact fixed-point sums of the same list units
    whatever its workgroup count, so a forced shape changes no bit."""
    op, sp = params_pair(**CONFIGS[name])
    w = oracle.World(op)
    res = []
    for wgs in (1, 5, 300, 0):
        sim = gpu_from_oracle(gpu, w, sp, de_path="culled", de_chunks=wgs)
        res.append(sim.delta_e()[2:])
        t = sim.tuning
        assert t["path"] == "culled"
        if wgs:   # four work items (chunks) per workgroup; one workgroup per target group at least
            assert t["work_items"] >= 4 * wgs and t["work_items"] % 4 == 0
        if wgs == 300:   # more than these lattices have target groups
            assert t["work_items"] == 1200
    for de, dh in res[1:]:
        assert np.array_equal(de, res[0][0]) and np.array_equal(dh, res[0][1])


def test_load_rejects_bad_state(gpu):
    sp = gpu.SimulationParameters(grid_x=8, grid_y=8, grid_z=1, coulomb_carriers=True,
                                  electron_percentage=0.0, hole_percentage=0.1,
                                  e_source_l_rate=0.0, random_seed=1)
    with pytest.raises(gpu.LmgError):
        gpu.Simulation(sp, holes=[3, 3])            # two holes on one site
    with pytest.raises(gpu.LmgError):
        gpu.Simulation(sp, holes=[64])              # outside the lattice
    with pytest.raises(gpu.LmgError):
        gpu.Simulation(sp, holes=list(range(7)))    # above max holes (6)
    sim = gpu.Simulation(sp, holes=[0, 9, 18])
    sim.step(5)
    assert sim.state()["current_step"] == 5
