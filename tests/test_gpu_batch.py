"""Batched replicas (lmg_batch_*): each replica of a heterogeneous batch (own
lattice, cutoff, carrier count, traps and seed) must follow exactly the trajectory
it follows alone, and the oracle's."""
import numpy as np
import pytest

from _util import compare_state, de_tolerance, oracle_state, params_pair

pytestmark = pytest.mark.gpu

REPLICAS = [
    dict(grid_x=24, grid_y=20, grid_z=1, hole_percentage=0.05, electrostatic_cutoff=20, random_seed=31),
    dict(grid_x=16, grid_y=16, grid_z=8, hole_percentage=0.02, electrostatic_cutoff=12, random_seed=32),
    dict(grid_x=40, grid_y=12, grid_z=3, hole_percentage=0.04, electrostatic_cutoff=50, random_seed=33),
    dict(simulation_type=1, grid_x=12, grid_y=12, grid_z=12, electron_percentage=0.02,
         hole_percentage=0.02, seed_charges=0.5, generation_rate=0.4, electrostatic_cutoff=9,
         recombination_rate=0.2, recombination_range=1, random_seed=34),
    dict(grid_x=32, grid_y=32, grid_z=1, hole_percentage=0.08, electrostatic_cutoff=30, random_seed=35),
    dict(grid_x=8, grid_y=8, grid_z=8, hole_percentage=0.05, electrostatic_cutoff=6, random_seed=36),
]
BASE = dict(coulomb_carriers=1, electron_percentage=0.0, seed_charges=1.0, voltage_left=1.0,
            e_source_l_rate=0.0, h_source_l_rate=0.9, h_drain_r_rate=0.9)


def build(oracle, gpu, de_path="auto", replicas=REPLICAS):
    worlds, reps = [], []
    for kw in replicas:
        k = dict(BASE)
        if kw.get("simulation_type") == 1:
            k = dict(coulomb_carriers=1)
        k.update(kw)
        op, sp = params_pair(**k)
        w = oracle.World(op)
        st = oracle_state(w)
        worlds.append((w, sp, st))
        reps.append(dict(params=sp, **st))
    return worlds, gpu.Batch(reps, de_path=de_path)


@pytest.mark.parametrize("de_path", ["auto", "culled"])
def test_batch_matches_single_contexts_and_oracle(gpu, oracle, de_path):
    """40 steps of every replica against its single context on the same dE path and
    against the oracle; "culled" puts every replica through the batched culled kernels
    (one launch for all of them, workgroups mapped to replicas)."""
    worlds, batch = build(oracle, gpu, de_path)
    assert len(batch) == len(REPLICAS)
    paths = batch.de_paths
    print(f"batch dE paths {paths}, tuning {batch.tuning}", flush=True)
    if de_path == "culled":
        assert paths == ["culled"] * len(REPLICAS)
    else:   # none of these small lattices passes the geometric test
        assert paths == ["pairs"] * len(REPLICAS)
    singles = [gpu.Simulation(sp, de_path=paths[r], **st) for r, (_, sp, st) in enumerate(worlds)]
    for t in range(40):
        batch.step(1)
        for s in singles:
            s.step(1)
        for r, (w, sp, _) in enumerate(worlds):
            w.step(1)
            bs = batch.state(r)
            ss = singles[r].state()
            for k in ("e_site", "h_site", "e_pathlength", "h_pathlength", "rng_window",
                      "flux_attempts", "flux_successes"):
                assert np.array_equal(bs[k], ss[k]), f"replica {r} step {t}: {k} differs from its single run"
            assert bs["rng_consumed"] == ss["rng_consumed"]
            assert not compare_state(bs, w), f"replica {r} step {t}: differs from the oracle"


def run_against_singles(gpu, worlds, batch, steps):
    paths = batch.de_paths
    singles = [gpu.Simulation(sp, de_path=paths[r], **st) for r, (_, sp, st) in enumerate(worlds)]
    for t in range(steps):
        batch.step(1)
        for s in singles:
            s.step(1)
        for r, (w, sp, _) in enumerate(worlds):
            w.step(1)
            bs, ss = batch.state(r), singles[r].state()
            for k in ("e_site", "h_site", "e_pathlength", "h_pathlength", "rng_window"):
                assert np.array_equal(bs[k], ss[k]), f"replica {r} step {t}: {k} differs from its single run"
            assert not compare_state(bs, w), f"replica {r} step {t}: differs from the oracle"


@pytest.mark.parametrize("de_path", ["auto", "culled"])
def test_batch_with_and_without_tables(gpu, oracle, de_path):
    """A replica whose cutoff (64) is past the pair-function table's reach runs the dE
    kernels' evaluated form; in a batch it takes a launch of its own (all pairs), and
    every replica still follows its single context and the oracle."""
    reps = REPLICAS[:3] + [dict(grid_x=20, grid_y=16, grid_z=4, hole_percentage=0.04,
                                electrostatic_cutoff=64, random_seed=37)]
    worlds, batch = build(oracle, gpu, de_path, reps)
    paths = batch.de_paths
    print(f"batch dE paths {paths}", flush=True)
    if de_path == "culled":   # the table-less replica is not culled beside replicas with tables
        assert paths == ["culled"] * 3 + ["pairs"]
    else:
        assert paths == ["pairs"] * 4
    run_against_singles(gpu, worlds, batch, 30)


def test_batch_autotuned_chunks(gpu, oracle):
    _, batch = build(oracle, gpu)
    t = batch.tuning
    assert t["charge_chunks"] >= 1 and t["charge_chunks"] & (t["charge_chunks"] - 1) == 0
    assert t["de_kernel_us"] > 0
    cands = [S for S, _ in t["pairs_candidates"]]
    assert cands == [2 ** k for k in range(len(cands))] and len(cands) >= 2
    assert min(us for _, us in t["pairs_candidates"]) == pytest.approx(t["de_phase_us"])
    _, culled = build(oracle, gpu, "culled")
    t = culled.tuning
    assert t["charge_chunks"] == 0 and t["culled_replicas"] == len(REPLICAS)
    assert all(q >= 1 for q, _ in t["cull_candidates"]) and t["cull_workgroups"] >= 1
    assert t["de_phase_us"] == pytest.approx(min(us for _, us in t["cull_candidates"]))
    print(f"culled batch tuning {t}", flush=True)


def test_batch_determinism_and_invariants(gpu, oracle):
    _, a = build(oracle, gpu)
    _, b = build(oracle, gpu)
    a.step(60)
    b.step(60)
    for r in range(len(REPLICAS)):
        sa, sb = a.state(r), b.state(r)
        assert np.array_equal(sa["h_site"], sb["h_site"]) and np.array_equal(sa["e_site"], sb["e_site"])
        assert len(np.unique(sa["h_site"])) == len(sa["h_site"])
        assert sa["current_step"] == 60


def test_batch_graph_replay_matches_eager(gpu, oracle):
    """The batch step replayed as a captured HIP graph (the default) equals kernel-by-kernel
    launches, call after call, with host reads of the state between the replays."""
    _, a = build(oracle, gpu)
    _, b = build(oracle, gpu)
    b.set_graphs(False)
    for n in (1, 4, 2, 7, 1, 5):
        a.step(n)
        b.step(n)
        for r in range(len(REPLICAS)):
            sa, sb = a.state(r), b.state(r)
            for k in ("e_site", "h_site", "e_lifetime", "h_lifetime", "rng_window", "flux_attempts",
                      "flux_successes"):
                assert np.array_equal(sa[k], sb[k]), f"replica {r} after {n} more steps: {k}"
            assert sa["rng_consumed"] == sb["rng_consumed"]
