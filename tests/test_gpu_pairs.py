"""Device pair sums and potential map against the FP64 oracle.

Tolerance (stated): pair potentials and maps are FP32 terms with compensated
accumulation; every output agrees with the FP64 oracle to
    |gpu - ref| <= 1e-5 |ref| + 5e-8 S,   S = pref * sum_j |q_j| f(r_j)
(S, the sum of the magnitudes of the terms, is the condition scale of a sum with
cancellation; 5e-8 is 0.84 FP32 unit roundoff).
"""
import numpy as np
import pytest

from _util import params_pair

pytestmark = pytest.mark.gpu

PREF = 0.41141839736460467


def tol_check(got, ref, sabs):
    err = np.abs(got - ref)
    bound = 1e-5 * np.abs(ref) + 5e-8 * np.abs(sabs) + 1e-15
    bad = np.nonzero(err > bound)[0]
    assert len(bad) == 0, (f"{len(bad)} outputs out of tolerance, worst {err.max():.3e}, "
                           f"worst ratio {np.max(err / bound):.2f}")
    rel = err / np.maximum(np.abs(ref), 1e-30)
    return float(err.max()), float(np.sqrt(np.mean(rel ** 2))), float(np.max(err / bound))


@pytest.mark.parametrize("dims,n,nt,cutoff,flags,sigma", [
    ((64, 64, 1), 800, 1600, 50, 0, 0.0),
    ((32, 32, 16), 3000, 5000, 30, 0, 0.0),
    ((20, 20, 20), 500, 700, 8, 1, 1.5),
    ((40, 24, 8), 1200, 900, 50, 2, 0.0),
    ((128, 128, 32), 10484, 300, 30, 0, 0.0),
])
def test_pair_potential_matches_oracle(gpu, oracle, dims, n, nt, cutoff, flags, sigma):
    rng = np.random.default_rng(n + nt)
    V = dims[0] * dims[1] * dims[2]
    cs = rng.choice(V, n, replace=False).astype(np.int32)
    cq = rng.choice([-1, 1], n).astype(np.int32)
    ts = rng.integers(0, V, nt).astype(np.int32)
    ctx = gpu.Context(dims, cutoff=cutoff, gaussian_sigma=sigma)
    got = ctx.pair_potential(cs, cq, ts, flags)
    ref = oracle.pair_potential(dims, cutoff, cs, cq, ts, flags=flags, sigma=sigma)
    sabs = oracle.pair_potential(dims, cutoff, cs, np.abs(cq), ts, flags=flags, sigma=sigma)
    tol_check(got, ref, sabs)


def test_pair_potential_special_slot_targets(gpu, oracle):
    """Targets V+k (drain proposals) decode exactly like kernel.cl:229-231."""
    dims = (16, 8, 4)
    V = 512
    rng = np.random.default_rng(1)
    cs = rng.choice(V, 60, replace=False)
    cq = rng.choice([-1, 1], 60)
    ts = np.array([V + 2, V + 3, V + 4, 0, 5], np.int32)
    ctx = gpu.Context(dims, cutoff=50)
    tol_check(ctx.pair_potential(cs, cq, ts), oracle.pair_potential(dims, 50, cs, cq, ts),
              oracle.pair_potential(dims, 50, cs, np.abs(cq), ts))


def test_pair_potential_edge_cases(gpu):
    ctx = gpu.Context((8, 8, 8), cutoff=3)
    assert len(ctx.pair_potential([], [], [])) == 0
    out = ctx.pair_potential([], [], [1, 2, 3])
    assert np.all(out == 0.0)
    # cutoff edge: r^2 = 8 (< 9) included, r^2 = 9 excluded
    s0 = 0
    t_in = 8 * (2 + 0) + 2       # (2,2,0): r2 = 8
    t_out = 8 * (0 + 0) + 3      # (3,0,0): r2 = 9
    out = ctx.pair_potential([s0], [1], [t_in, t_out, s0])
    assert out[0] == pytest.approx(PREF / np.sqrt(8.0), rel=1e-6)
    assert out[1] == 0.0 and out[2] == 0.0
    with pytest.raises(gpu.LmgError):
        ctx.pair_potential([512], [1], [0])          # charge outside the lattice
    with pytest.raises(gpu.LmgError):
        ctx.pair_potential([0], [1], [0], gpu.LMG_GAUSS)   # sigma = 0


def test_gpu_helper_contract(gpu, oracle):
    """launchCoulombKernel2 staging + getOutputHost/getOutputHostFuture indexing."""
    from langmuir_amd import GpuHelper
    dims = (30, 20, 4)
    V = 2400
    rng = np.random.default_rng(3)
    sites = rng.choice(V, 300, replace=False)
    e, h = sites[:120], sites[120:]
    fe, fh = rng.integers(0, V, 120), rng.integers(0, V, 180)
    ctx = gpu.Context(dims, cutoff=50)
    helper = GpuHelper(ctx)
    helper.launch_coulomb_kernel2(e, fe, h, fh)
    cs = np.concatenate([e, h])
    cq = np.concatenate([-np.ones(120), np.ones(180)]).astype(np.int32)
    ref_now = oracle.pair_potential(dims, 50, cs, cq, cs)
    ref_fut = oracle.pair_potential(dims, 50, cs, cq, np.concatenate([fe, fh]))
    for ident in (0, 57, 119, 120, 200, 299):   # electrons i, holes Ne + i
        assert helper.get_output_host(ident) == pytest.approx(ref_now[ident], rel=1e-5, abs=1e-6)
        assert helper.get_output_host_future(ident) == pytest.approx(ref_fut[ident], rel=1e-5, abs=1e-6)


MAP_PATHS = {"pairs": 8, "culled": 16}   # LMG_MAP_PAIRS, LMG_MAP_CULLED


@pytest.mark.parametrize("path", sorted(MAP_PATHS))
@pytest.mark.parametrize("dims,n,cutoff,flags,sigma,dup", [
    ((33, 17, 9), 400, 6, 0, 0.0, False),      # lattice sides not multiples of the 8-site brick
    ((64, 64, 1), 1310, 50, 0, 0.0, False),    # flat: 8 x 8 x 1 tiles
    ((24, 24, 24), 900, 0, 4, 0.0, False),     # LMG_NO_CUTOFF (no culling)
    ((16, 16, 16), 300, 5, 1, 1.2, False),     # gauss
    ((40, 24, 16), 700, 9, 0, 0.0, True),      # charges sharing sites
    ((16, 16, 16), 2600, 7, 0, 0.0, True),     # dense bricks: lists of several 64-slot units
])
def test_potential_map_matches_oracle(gpu, oracle, dims, n, cutoff, flags, sigma, dup, path):
    rng = np.random.default_rng(n)
    V = dims[0] * dims[1] * dims[2]
    cs = rng.choice(V, n, replace=dup).astype(np.int32)
    cq = rng.choice([-1, 1], n).astype(np.int32)
    ctx = gpu.Context(dims, cutoff=max(cutoff, 1), gaussian_sigma=sigma)
    got = ctx.potential_map(cs, cq, flags | MAP_PATHS[path]).astype(np.float64)
    cut = cutoff if not (flags & 4) else 0
    ref = oracle.potential_map(dims, cut, cs, cq, flags=flags & 1, sigma=sigma, threads=8)
    sabs = oracle.potential_map(dims, cut, cs, np.abs(cq), flags=flags & 1, sigma=sigma, threads=8)
    tol_check(got, ref, sabs)
    if path == "culled":   # the brick lists are rebuilt in input order on every call: same bits
        again = ctx.potential_map(cs, cq, flags | MAP_PATHS[path]).astype(np.float64)
        assert np.array_equal(got, again)


@pytest.mark.parametrize("cutoff,flags", [(50, 0), (50, 8), (50, 4)])   # automatic (culled), all pairs, no cutoff
def test_potential_map_full_size_sampled(gpu, oracle, cutoff, flags):
    """256x256x64 lattice, 50,000 charges +/-1: error statistics on 4096 sampled sites."""
    dims = (256, 256, 64)
    V = 256 * 256 * 64
    rng = np.random.default_rng(2024)
    cs = rng.choice(V, 50000, replace=False).astype(np.int32)
    cq = np.where(np.arange(50000) < 25000, 1, -1).astype(np.int32)
    ctx = gpu.Context(dims, cutoff=cutoff)
    got = ctx.potential_map(cs, cq, flags)
    sample = rng.choice(V, 4096, replace=False)
    cut = cutoff if not (flags & 4) else 100000
    ref = oracle.pair_potential(dims, cut, cs, cq, sample)
    sabs = oracle.pair_potential(dims, cut, cs, np.abs(cq), sample)
    mx, rms, ratio = tol_check(got[sample].astype(np.float64), ref, sabs)
    print(f"map 256x256x64 n=50000 flags={flags}: max abs err {mx:.3e}, rms rel err {rms:.3e}, "
          f"worst err/bound {ratio:.2f}")
