"""Boost uniform_int rejections in the proposal draws, forced on the device.

A proposal draws uniform_int(0, n - 1) over the carrier's neighbour list; Boost
rejects a word w >= n * bucket(n - 1) and draws again (rand.cpp:68-73), which shifts
every later carrier's word.  Rejections are about one in 10^8 draws, so the tests
build the engine state instead: the 624-word window is edited so that chosen engine
outputs of the first step come out as 0xFFFFFFFF, inside every non-power-of-two
list's rejection zone (output j is temper(x[j + 397] ^ twist(x[j], x[j + 1])), and
x[j + 397] is free for j < 227).  The device must make those carriers' proposals
(and all later ones) exactly as the oracle does, on both dE paths, also when more
rejection-zone outputs than the device keeps track of arrive at once.
"""
import numpy as np
import pytest

from _util import compare_state, gpu_from_oracle, params_pair
from test_gpu_sim import CONFIGS, run_parity

pytestmark = pytest.mark.gpu


def temper(y):
    y ^= y >> 11
    y ^= (y << 7) & 0x9D2C5680
    y ^= (y << 15) & 0xEFC60000
    y ^= y >> 18
    return y & 0xFFFFFFFF


def untemper(y):
    y ^= y >> 18
    y ^= (y << 15) & 0xEFC60000
    x = y
    for _ in range(4):
        x = y ^ ((x << 7) & 0x9D2C5680)
    y = x & 0xFFFFFFFF
    y ^= (y >> 11) ^ (y >> 22)
    return y & 0xFFFFFFFF


def twist(x0, x1):
    y = (x0 & 0x80000000) | (x1 & 0x7FFFFFFF)
    return (y >> 1) ^ (0x9908B0DF if y & 1 else 0)


def test_untemper_inverts_temper():
    rng = np.random.default_rng(3)
    for v in [0, 1, 0xFFFFFFFF, 0x80000000] + [int(x) for x in rng.integers(0, 2 ** 32, 200)]:
        assert temper(untemper(v)) == v and untemper(temper(v)) == v


def force_outputs(window, outputs, value=0xFFFFFFFF):
    w = [int(x) for x in window]
    for j in outputs:
        assert j < 227
        w[397 + j] = untemper(value) ^ twist(w[j], w[j + 1])
    return np.asarray(w, np.uint32)


def rejectable(oracle, w, dims, hop):
    """List indices (electrons, then holes) whose neighbour count is not a power of two."""
    V = dims[0] * dims[1] * dims[2]
    sites = np.concatenate([w.carriers(0)["site"], w.carriers(1)["site"]])
    out = []
    for i, s in enumerate(sites):
        n = len(oracle.neighbors(dims, hop, int(s), V, V + 1))
        if n & (n - 1):
            out.append(i)
    return out


@pytest.mark.parametrize("name,de_path,count", [("transistor2d", "pairs", 3), ("solar3d", "pairs", 3),
                                                ("solar3d", "culled", 3), ("solar3d", "culled", 20),
                                                ("hop2", "culled", 2)])
def test_forced_proposal_rejections(gpu, oracle, name, de_path, count):
    kw = dict(CONFIGS[name])
    op, sp = params_pair(**kw)
    w = oracle.World(op)
    dims = (sp.grid_x, sp.grid_y, sp.grid_z)
    win, k = w.rng()
    cand = rejectable(oracle, w, dims, sp.hopping_range)
    picks = [i for i in cand if i < 227][:count]
    if count > 3:   # more rejection-zone outputs than the device lists: fill with later outputs
        picks = sorted(set(picks) | set(range(100, 100 + count - len(picks))))
    assert len(picks) >= min(count, 2), f"{name}: too few carriers with a rejectable list"
    w.set_rng(force_outputs(win, picks), k)
    print(f"{name} {de_path}: outputs {picks} forced into the rejection zone", flush=True)
    sim = gpu_from_oracle(gpu, w, sp, de_path=de_path)
    sim.set_synchronous(True)
    fe, fh, _, _ = sim.delta_e()
    w.propose()
    e, h = w.carriers(0), w.carriers(1)
    assert np.array_equal(fe, e["fsite"]) and np.array_equal(fh, h["fsite"])
    w2 = oracle.World(op)
    w2.set_rng(force_outputs(win, picks), k)
    assert run_parity(w2, sim, 12, sp.inverse_kt) <= 1
    assert sim.device_error() == 0
