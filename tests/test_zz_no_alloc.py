"""lmg_pair_potential_noalloc_dev, the entry point that allocates nothing, against
lmg_pair_potential_dev through the C-ABI: the same bits on two calls in a row (the
second with fewer targets and charges, so stale scratch contents from the first call
would show), for plain, erf-screened and image sums; both within the FP64 oracle's
tolerance; sizes above the scratch limits are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def call(L, fn, ctx, d_cs, d_cq, nc, d_ts, nt, flags, d_out):
    from langmuir_amd import _abi
    return getattr(_abi.lib(), fn)(ctx.handle, C.c_void_p(d_cs.ptr), C.c_void_p(d_cq.ptr), nc,
                                   C.c_void_p(d_ts.ptr), nt, flags, C.c_void_p(d_out.ptr))


@pytest.mark.parametrize("flags", [0, 1, 2])
def test_noalloc_pair_potential_matches_the_allocating_entry_point(gpu, oracle, flags):
    L = gpu
    dims = (96, 80, 24)
    V = dims[0] * dims[1] * dims[2]
    rng = np.random.default_rng(11 + flags)
    ctx = L.Context(dims, cutoff=30, gaussian_sigma=1.5 if flags & 1 else 0.0)
    for n_charges, n_targets in ((60000, 40000), (7001, 913)):   # the second call is smaller
        print(f"noalloc flags {flags}: {n_charges} charges, {n_targets} targets", flush=True)
        cs = rng.choice(V, n_charges, replace=False).astype(np.int32)
        cq = rng.choice([-1, 1], n_charges).astype(np.int32)
        ts = rng.integers(0, V, n_targets).astype(np.int32)
        d_cs, d_cq, d_ts = (L.DeviceBuffer.from_array(a) for a in (cs, cq, ts))
        d_a, d_b = L.DeviceBuffer(8 * n_targets), L.DeviceBuffer(8 * n_targets)
        assert call(L, "lmg_pair_potential_dev", ctx, d_cs, d_cq, n_charges, d_ts, n_targets, flags, d_a) == 0
        assert call(L, "lmg_pair_potential_noalloc_dev", ctx, d_cs, d_cq, n_charges, d_ts, n_targets, flags,
                    d_b) == 0
        ctx.sync()
        a, b = d_a.download(np.float64, n_targets), d_b.download(np.float64, n_targets)
        assert np.array_equal(a, b)
        # against the FP64 oracle on up to 1,500 of the targets
        oflags = (oracle.GAUSS if flags & 1 else 0) | (oracle.IMAGE if flags & 2 else 0)
        pick = np.arange(0, n_targets, max(1, n_targets // 1500))
        sig = 1.5 if flags & 1 else 0.0
        ref = oracle.pair_potential(dims, 30, cs, cq, ts[pick], flags=oflags, sigma=sig)
        sabs = oracle.pair_potential(dims, 30, cs, np.abs(cq), ts[pick], flags=oflags, sigma=sig)
        assert np.all(np.abs(b[pick] - ref) <= 1e-5 * np.abs(ref) + 5e-8 * np.abs(sabs) + 1e-12)
        for buf in (d_cs, d_cq, d_ts, d_a, d_b):
            buf.free()
    ctx.close()


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    L = gpu
    from langmuir_amd import _abi
    cap = _abi.LMG_ERR_CAPACITY
    ctx = L.Context((64, 64, 32), cutoff=20)
    d = L.DeviceBuffer(8 * 70000)
    lib = _abi.lib()
    assert lib.lmg_pair_potential_noalloc_dev(ctx.handle, C.c_void_p(d.ptr), C.c_void_p(d.ptr), 10,
                                              C.c_void_p(d.ptr), 65537, 0, C.c_void_p(d.ptr)) == cap
    assert b"at most 65536 targets" in lib.lmg_last_error(ctx.handle)
    assert lib.lmg_pair_potential_noalloc_dev(ctx.handle, C.c_void_p(d.ptr), C.c_void_p(d.ptr), 262145,
                                              C.c_void_p(d.ptr), 10, 0, C.c_void_p(d.ptr)) == cap
    d.free()
    ctx.close()
