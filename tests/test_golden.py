"""The committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py): the oracle must keep reproducing them (CPU), and
the device must reproduce the trajectories, proposals and pair sums (GPU)."""
import os

import numpy as np
import pytest

from _util import de_tolerance, gpu_from_oracle, params_pair

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "golden.npz"))


def _traj_cfg():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    return mg.TRAJ, mg.STEPS


TRAJ, STEPS = _traj_cfg()


@pytest.mark.parametrize("name", sorted(TRAJ))
def test_oracle_reproduces_golden(oracle, name):
    w = oracle.World(oracle.OrcParams.make(**TRAJ[name]))
    assert np.array_equal(w.carriers(0)["site"], G[f"{name}/init_e"])
    done = 0
    for s in STEPS:
        w.step(s - done)
        done = s
        assert np.array_equal(w.carriers(0)["site"], G[f"{name}/s{s}/e_site"])
        assert np.array_equal(w.carriers(1)["site"], G[f"{name}/s{s}/h_site"])
        assert w.rng()[1] == int(G[f"{name}/s{s}/rng_consumed"][0])


def test_oracle_reproduces_golden_pairs(oracle):
    dims = tuple(int(v) for v in G["pairs/dims"])
    cs, cq, ts = G["pairs/charge_site"], G["pairs/charge_q"], G["pairs/target_site"]
    assert np.array_equal(oracle.pair_potential(dims, 9, cs, cq, ts), G["pairs/coulomb_c9"])
    assert np.array_equal(oracle.pair_potential(dims, 9, cs, cq, ts, flags=oracle.IMAGE), G["pairs/image_c9"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(TRAJ))
def test_device_reproduces_golden_trajectory(gpu, oracle, name):
    op, sp = params_pair(**TRAJ[name])
    w = oracle.World(op)          # initial configuration only
    sim = gpu_from_oracle(gpu, w, sp)
    done = 0
    for s in STEPS:
        sim.step(s - done)
        done = s
        st = sim.state()
        assert np.array_equal(st["e_site"], G[f"{name}/s{s}/e_site"]), f"electrons differ at step {s}"
        assert np.array_equal(st["h_site"], G[f"{name}/s{s}/h_site"]), f"holes differ at step {s}"
        assert np.array_equal(st["e_pathlength"], G[f"{name}/s{s}/e_path"])
        assert st["rng_consumed"] == int(G[f"{name}/s{s}/rng_consumed"][0])
        assert np.array_equal(st["rng_window"][:8], G[f"{name}/s{s}/rng_head"])
        assert np.array_equal(np.stack([st["flux_attempts"], st["flux_successes"]]), G[f"{name}/s{s}/flux"])
    fe, fh, de, dh = sim.delta_e()
    assert np.array_equal(np.concatenate([fe, fh]), G[f"{name}/next_fsite"])
    ref = G[f"{name}/next_de"]
    assert np.all(np.abs(np.concatenate([de, dh]) - ref) <= de_tolerance(ref))


@pytest.mark.gpu
def test_device_reproduces_golden_pairs(gpu):
    dims = tuple(int(v) for v in G["pairs/dims"])
    cs, cq, ts = G["pairs/charge_site"], G["pairs/charge_q"], G["pairs/target_site"]
    ctx = gpu.Context(dims, cutoff=9, gaussian_sigma=1.5)
    for key, flags in (("coulomb_c9", 0), ("image_c9", gpu.LMG_IMAGE_LEFT), ("gauss_c9_s1p5", gpu.LMG_GAUSS)):
        ref = G[f"pairs/{key}"]
        got = ctx.pair_potential(cs, cq, ts, flags)
        assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-6 * np.abs(ref).max()), key
    for key, flags in (("map_c9", 0), ("map_nocut", gpu.LMG_NO_CUTOFF)):
        ref = G[f"pairs/{key}"]
        got = ctx.potential_map(cs, cq, flags).astype(np.float64)
        assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 1e-6 * np.abs(ref).max()), key
