"""The measured configurations at their real shape against the FP64 oracle.

* The 128x128x32 solar cell exactly as the benchmark times it: the two-phase
  morphology (sigma 4, split at the median, one phase 0.5 eV traps), 5242 electrons
  and 5242 holes at seeded sites, exciton generation 0.524288 per step, cutoff 30 --
  30 device steps (the culled dE path, as chosen at this size), the full state
  compared with the oracle's after every step.
* A batch of 16 replicas drawn from the benchmark's replica distribution (lattices
  32^3 .. 128x128x16, 100 .. 5000 holes, 0 .. 10 % traps with depths
  clip(N(0.3, 0.1), 0.1, 0.5) eV, cutoffs 20 .. 50, own seeds), including the largest
  lattice and 5000 holes: 10 steps, every replica after every step equal to its own
  single-context run on the dE path the batch gave it and to the oracle -- once with
  the paths the autotuner picks, once with every replica culled.

A step may differ from the oracle only through a Metropolis decision inside the band
the dE tolerance opens (as in test_gpu_sim); the solar run re-synchronises then, a
batch replica leaves the oracle comparison (its single-context comparison goes on).
"""
import os

import numpy as np
import pytest

from _util import compare_state, de_tolerance, gpu_from_oracle, oracle_state, params_pair

pytestmark = pytest.mark.gpu


def threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 1


def test_solar_bench_cell_30_steps(gpu, oracle):
    from langmuir_amd.morphology import two_phase_traps
    from test_gpu_sim import run_parity
    dims = (128, 128, 32)
    V = 128 * 128 * 32
    traps = two_phase_traps(dims, 4.0, 2024)[0]
    rng = np.random.default_rng(2024)
    e = rng.choice(V, 5242, replace=False).astype(np.int32)
    h = rng.choice(V, 5242, replace=False).astype(np.int32)
    op, sp = params_pair(simulation_type=1, grid_x=128, grid_y=128, grid_z=32, coulomb_carriers=1,
                         electron_percentage=0.01, hole_percentage=0.01, seed_charges=0.0,
                         generation_rate=0.524288, source_scale_area=0.0, electrostatic_cutoff=30,
                         trap_percentage=(len(traps) + 0.5) / V, trap_potential=0.5, random_seed=12345)
    print(f"solar bench cell: {len(traps)} traps, 5242 e + 5242 h", flush=True)
    w = oracle.World(op, electrons=e, holes=h, traps=traps, trap_potentials=np.full(len(traps), 0.5))
    assert w.rng()[1] == 0 and w.info()["n_traps"] == len(traps)
    pot = gpu.background_potential(sp, traps, np.full(len(traps), 0.5))
    assert np.array_equal(pot, w.potential())
    sim = gpu.Simulation(sp, electrons=e, holes=h, potential=pot)
    sim.set_synchronous(True)
    assert sim.de_path == "culled"
    mism = run_parity(w, sim, 30, sp.inverse_kt, threads=threads())
    print(f"solar bench cell: 30 steps, {mism} near-threshold resyncs", flush=True)
    assert mism <= 2
    assert sim.device_error() == 0
    s = sim.state()
    assert len(s["e_site"]) + len(s["h_site"]) > 10000


LATTICES = ((32, 32, 32), (48, 48, 24), (64, 64, 16), (64, 64, 32), (96, 96, 16), (128, 128, 16))


def bench_like_replicas(n=16, seed=4):
    """Replicas from the benchmark's distribution; replica 0 is the largest lattice with
    5000 holes and 10 % traps at cutoff 50, replica 1 the smallest lattice with 100 holes."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        dims = LATTICES[rng.integers(len(LATTICES))]
        V = dims[0] * dims[1] * dims[2]
        nh = int(min(rng.integers(100, 5001), V // 10))
        cutoff = int(rng.integers(20, 51))
        tfrac = rng.uniform(0.0, 0.1)
        if r == 0:
            dims, nh, cutoff, tfrac = (128, 128, 16), 5000, 50, 0.1
        elif r == 1:
            dims, nh, cutoff, tfrac = (32, 32, 32), 100, 20, 0.0
        V = dims[0] * dims[1] * dims[2]
        nt = int(tfrac * V)
        traps = rng.choice(V, nt, replace=False).astype(np.int32)
        depth = np.clip(rng.normal(0.3, 0.1, nt), 0.1, 0.5)
        holes = rng.choice(V, nh, replace=False).astype(np.int32)
        out.append(dict(dims=dims, holes=holes, traps=traps, depth=depth, cutoff=cutoff,
                        seed=int(rng.integers(1, 2 ** 31))))
    return out


@pytest.mark.parametrize("de_path", ["auto", "culled"])
def test_batch_of_bench_replicas_10_steps(gpu, oracle, de_path):
    from test_gpu_sim import run_parity  # noqa: F401  (same near-threshold rule below)
    reps = bench_like_replicas()
    worlds, states, batch_in = [], [], []
    for r, rp in enumerate(reps):
        X, Y, Z = rp["dims"]
        V = X * Y * Z
        op, sp = params_pair(grid_x=X, grid_y=Y, grid_z=Z, coulomb_carriers=1, electron_percentage=0.0,
                             hole_percentage=(len(rp["holes"]) + 0.5) / V, seed_charges=0.0, voltage_left=1.0,
                             e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9,
                             electrostatic_cutoff=rp["cutoff"], random_seed=rp["seed"],
                             trap_percentage=(len(rp["traps"]) + 0.5) / V if len(rp["traps"]) else 0.0)
        w = oracle.World(op, holes=rp["holes"], traps=rp["traps"], trap_potentials=rp["depth"])
        pot = gpu.background_potential(sp, rp["traps"], rp["depth"])
        assert np.array_equal(pot, w.potential())
        st = oracle_state(w)
        worlds.append((w, sp))
        states.append(st)
        batch_in.append(dict(params=sp, **st))
    print(f"batch of {len(reps)}: lattices {[r['dims'] for r in reps[:4]]} ..., "
          f"holes {sum(len(r['holes']) for r in reps)}", flush=True)
    batch = gpu.Batch(batch_in, de_path=de_path)
    paths = batch.de_paths
    print(f"batch dE paths {paths}", flush=True)
    if de_path == "culled":
        assert paths[0] == "culled" and paths.count("culled") == len(paths)
    # each replica's single context on the path the batch gave it (same unit sums, same bits)
    singles = [gpu.Simulation(sp, de_path=paths[r], **st) for r, ((_, sp), st) in enumerate(zip(worlds, states))]
    diverged = set()
    for t in range(10):
        batch.step(1)
        for r, ((w, sp), single) in enumerate(zip(worlds, singles)):
            single.step(1)
            bs, ss = batch.state(r), single.state()
            for k in ("e_site", "h_site", "e_lifetime", "h_lifetime", "e_pathlength", "h_pathlength",
                      "rng_window", "flux_attempts", "flux_successes"):
                assert np.array_equal(bs[k], ss[k]), f"replica {r} step {t}: {k} differs from its single run"
            assert bs["rng_consumed"] == ss["rng_consumed"]
            if r in diverged:
                continue
            w.propose()
            w.coulomb(threads())
            nh, de_h = w.info()["n_h"], w.carriers(1)["de"]
            w.finish()
            if compare_state(bs, w):
                u, m, pd = w.last_decisions(1, nh)
                band = sp.inverse_kt * (1.0 / 3.0) * de_tolerance(de_h) + 1e-12
                assert np.any((u >= 0) & (np.abs(m) <= band)), f"replica {r} step {t}: differs from the oracle"
                diverged.add(r)
    print(f"batch: 10 steps, replicas leaving the oracle comparison at a near-threshold decision: "
          f"{sorted(diverged)}", flush=True)
    assert len(diverged) <= 2
    batch.close()
