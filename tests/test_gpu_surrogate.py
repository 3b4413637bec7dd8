"""The learned map surrogate on the device (lmg_surrogate_*), against the FP64
numpy oracle (oracle/surrogate.py) and the exact map kernel.

Tolerances: forward maps within 1e-4 relative RMS of the oracle's forward pass
with the same weights (BF16 hi/lo operands with FP32 accumulation carry about
2^-16 relative error per DFT stage); gradients within 1e-3 relative norm of the
oracle's analytic gradient; a trained network within the 1e-3 relative RMS gate
of the exact map on held-out snapshots.
"""
import time

import numpy as np
import pytest

from oracle import surrogate as S

pytestmark = pytest.mark.gpu

PREF = 0.41141839736460467


def rel_rms(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def random_charges(rng, V, n):
    e = rng.choice(V, n, replace=False)
    h = rng.choice(V, n, replace=False)          # may coincide with electrons (excitons)
    return np.concatenate([e, h]).astype(np.int32), np.concatenate([-np.ones(n), np.ones(n)]).astype(np.int32)


def make(gpu, dims, cutoff, flags=0, sigma=0.0, max_batch=4):
    ctx = gpu.Context(dims, cutoff=cutoff, gaussian_sigma=sigma)
    return ctx, gpu.Surrogate(ctx, flags=flags, max_batch=max_batch)


@pytest.mark.parametrize("dims", [(24, 20, 6), (13, 7, 5), (32, 24, 1)])
def test_occupancy_matches_numpy(gpu, dims):
    rng = np.random.default_rng(1)
    V = int(np.prod(dims))
    ctx, sur = make(gpu, dims, 6)
    sites, q = random_charges(rng, V, max(2, V // 20))
    got = sur.occupancy(sites, q)
    assert np.array_equal(got, S.occupancy(dims, sites, q))


@pytest.mark.parametrize("dims,cutoff,flags,sigma", [
    ((24, 20, 6), 6, 0, 0.0),
    ((16, 12, 10), 4, 4, 0.0),       # LMG_NO_CUTOFF: padded to 2N
    ((32, 24, 1), 10, 0, 0.0),       # flat lattice
    ((14, 12, 8), 5, 1, 1.2),        # LMG_GAUSS
])
def test_forward_matches_oracle(gpu, dims, cutoff, flags, sigma):
    rng = np.random.default_rng(2)
    V = int(np.prod(dims))
    none = bool(flags & 4)
    ctx, sur = make(gpu, dims, cutoff, flags, sigma)
    assert sur.padded == S.padded_shape(dims, cutoff, none)
    W = S.ideal_weights(dims, cutoff, PREF, sigma, none)
    W = W * (1 + 0.1 * rng.normal(size=W.shape))        # generic weights near the operator
    sur.set_weights(W)
    assert np.array_equal(sur.get_weights(), W.astype(np.complex64))
    occ = np.stack([S.occupancy(dims, *random_charges(rng, V, max(2, V // 30))) for _ in range(3)])
    got = sur.infer(occ)
    ref = S.forward(W.astype(np.complex64).astype(np.complex128), occ, dims, cutoff, none)
    for b in range(3):
        err = rel_rms(got[b], ref[b])
        print(f"{dims} flags {flags}: forward rel RMS {err:.2e}")
        assert err <= 1e-4


def test_gradient_matches_oracle(gpu):
    dims, cutoff = (20, 16, 6), 5
    rng = np.random.default_rng(3)
    V = int(np.prod(dims))
    ctx, sur = make(gpu, dims, cutoff)
    W = S.ideal_weights(dims, cutoff, PREF) * (1 + 0.2 * rng.normal(size=S.spectrum_shape(dims, cutoff)))
    W = W.astype(np.complex64)
    sur.set_weights(W)
    occ = np.stack([S.occupancy(dims, *random_charges(rng, V, 30)) for _ in range(3)])
    tgt = S.forward(S.ideal_weights(dims, cutoff, PREF), occ, dims, cutoff).astype(np.float32)
    g, loss = sur.gradient(occ, tgt)
    L_ref, g_ref = S.loss_and_grad(W.astype(np.complex128), occ, tgt.astype(np.float64), dims, cutoff)
    err = float(np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref))
    print(f"gradient rel err {err:.2e}, loss {loss:.6e} vs {L_ref:.6e}")
    assert err <= 1e-3
    assert abs(loss - L_ref) <= 1e-3 * L_ref


def test_gate_accepts_ideal_weights_and_rejects_untrained(gpu, oracle):
    dims, cutoff = (32, 28, 10), 8
    rng = np.random.default_rng(4)
    V = int(np.prod(dims))
    ctx, sur = make(gpu, dims, cutoff)
    sites, q = random_charges(rng, V, 200)
    exact = ctx.potential_map(sites, q)
    ref = oracle.potential_map(dims, cutoff, sites, q, pref=PREF)
    # untrained network: the gate fails and the exact kernel's map is returned
    out, err, used = sur.map(sites, q, tolerance=1e-3, gate_sites=2048)
    assert used and err > 0.5
    assert np.array_equal(out, exact)
    # the ideal operator passes the gate and matches the FP64 map
    sur.set_weights(S.ideal_weights(dims, cutoff, PREF))
    out, err, used = sur.map(sites, q, tolerance=1e-3, gate_sites=2048)
    print(f"ideal weights: gate error {err:.2e}, full-map rel RMS {rel_rms(out, ref):.2e}")
    assert not used and err <= 1e-4
    assert rel_rms(out, ref) <= 1e-4
    # a gate tighter than the network's error sends the call to the exact map
    out, err2, used = sur.map(sites, q, tolerance=err / 10, gate_sites=2048)
    assert used and np.array_equal(out, exact)


def test_graph_replays_match_eager_launches(gpu):
    """The gated map replayed from its recorded graph, with other device work
    and synchronisations between the replays, equals the eagerly launched
    sequence every time (both for an accepted and a rejected gate)."""
    dims, cutoff = (24, 20, 8), 6
    rng = np.random.default_rng(8)
    V = int(np.prod(dims))
    ctx, sur = make(gpu, dims, cutoff)
    sites, q = random_charges(rng, V, 120)
    d_s, d_q = gpu.DeviceBuffer.from_array(sites), gpu.DeviceBuffer.from_array(q)
    d_out = gpu.DeviceBuffer(4 * V)
    other = np.stack([S.occupancy(dims, *random_charges(rng, V, 50))])
    for weights in (S.ideal_weights(dims, cutoff, PREF), None):
        sur.set_weights(weights)
        sur.set_graphs(False)
        err0, used0 = sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 1e-3)
        eager = d_out.download(np.float32, V)
        sur.set_graphs(True)
        for rep in range(4):
            d_out.upload(np.full(V, np.nan, np.float32))
            err, used = sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 1e-3)
            assert (err, used) == (err0, used0)
            assert np.array_equal(d_out.download(np.float32, V), eager), f"replay {rep}"
            sur.infer(other)                     # eager work between replays
        sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 1e-3, wait=False)
        ctx.sync()
        assert sur.last_gate() == (err0, used0)
        assert np.array_equal(d_out.download(np.float32, V), eager)


@pytest.mark.parametrize("dims,cutoff", [((28, 24, 8), 7), ((13, 7, 5), 4)])   # V = 455: grids off 4-byte alignment
def test_batched_gated_maps_match_single_calls(gpu, oracle, dims, cutoff):
    """lmg_surrogate_map_batch_dev: every map of a batch (lists of different
    lengths, one empty) equals the single gated call on its list -- the map bit
    for bit, the gate decision, and the gate error to within 1e-6 (its exact pair
    sums are split over charges differently in a batch, so they differ in FP32
    rounding) -- for accepted (ideal weights) and rejected (untrained) gates."""
    rng = np.random.default_rng(11)
    V = int(np.prod(dims))
    ctx, sur = make(gpu, dims, cutoff, max_batch=4)
    n = max(10, V // 40)
    lists = [random_charges(rng, V, n), random_charges(rng, V, n // 4), (np.zeros(0, np.int32), np.zeros(0, np.int32)),
             random_charges(rng, V, n // 2)]
    counts = [len(a) for a, _ in lists]
    d_s = gpu.DeviceBuffer.from_array(np.concatenate([a for a, _ in lists]))
    d_q = gpu.DeviceBuffer.from_array(np.concatenate([b for _, b in lists]))
    d_out = gpu.DeviceBuffer(4 * V * len(lists))
    for weights in (S.ideal_weights(dims, cutoff, PREF), None):
        sur.set_weights(weights)
        d_out.upload(np.full(V * len(lists), np.nan, np.float32))
        errs, used = sur.map_batch_dev(d_s.ptr, d_q.ptr, counts, d_out.ptr, 1e-3, gate_sites=512)
        got = d_out.download(np.float32, V * len(lists)).reshape(len(lists), V)
        for b, (sites, q) in enumerate(lists):
            out, err, u = sur.map(sites, q, tolerance=1e-3, gate_sites=512)
            assert u == used[b], b
            assert np.array_equal(got[b], out), b
            # the gate's exact sums differ in FP32 rounding (about 1e-7 relative per site)
            assert abs(err - errs[b]) <= 1e-6 + 1e-6 * err, (b, err, errs[b])
            if len(sites) and weights is not None:
                assert rel_rms(got[b], oracle.potential_map(dims, cutoff, sites, q, pref=PREF)) <= 1e-4
        expect = [False] * 4 if weights is not None else [True, True, False, True]   # an empty list's map is 0
        assert used.tolist() == expect
    sur.map_batch_dev(d_s.ptr, d_q.ptr, counts, d_out.ptr, 1e-3, gate_sites=512, wait=False)
    e2, u2 = sur.last_gates(len(lists))
    assert np.array_equal(e2, errs) and np.array_equal(u2, used)


@pytest.mark.parametrize("gauss", [False, True])
def test_fallback_takes_the_culled_map_where_the_context_would(gpu, oracle, gauss):
    """A failed gate returns the map the context's own lmg_potential_map gives,
    bit for bit, whichever kernel that call picks: 64x64x16 at cutoff 10 with
    2,000 charges culls (the cutoff box of a tile covers a quarter of the lattice),
    with 200 it sums every charge at every site.  In one batch the two kinds run
    side by side, eager and from the recorded graph, with the surrogate's output
    poisoned before every call; the exact maps are checked against the FP64 map."""
    dims, cutoff = (64, 64, 16), 10
    rng = np.random.default_rng(23)
    V = int(np.prod(dims))
    sigma = 1.5 if gauss else 0.0
    fl = gpu.LMG_GAUSS if gauss else 0
    ctx, sur = make(gpu, dims, cutoff, flags=fl, sigma=sigma, max_batch=3)
    lists = [random_charges(rng, V, 1000), random_charges(rng, V, 100), (np.zeros(0, np.int32), np.zeros(0, np.int32))]
    exact = [ctx.potential_map(s, q, flags=fl) for s, q in lists]
    # the default decision for the long list is the culled kernel, for the short one all pairs
    assert np.array_equal(exact[0], ctx.potential_map(*lists[0], flags=fl | gpu.LMG_MAP_CULLED))
    assert not np.array_equal(exact[0], ctx.potential_map(*lists[0], flags=fl | gpu.LMG_MAP_PAIRS))
    assert np.array_equal(exact[1], ctx.potential_map(*lists[1], flags=fl | gpu.LMG_MAP_PAIRS))
    for b in (0, 1):
        ref = oracle.potential_map(dims, cutoff, *lists[b], flags=fl, pref=PREF, sigma=sigma)
        assert rel_rms(exact[b], ref) <= 1e-5
    for sites, q in lists[:2]:   # untrained network: every non-empty list falls back
        out, err, used = sur.map(sites, q, tolerance=1e-3, gate_sites=1024)
        assert used and np.array_equal(out, ctx.potential_map(sites, q, flags=fl))
    counts = [len(a) for a, _ in lists]
    d_s = gpu.DeviceBuffer.from_array(np.concatenate([a for a, _ in lists]))
    d_q = gpu.DeviceBuffer.from_array(np.concatenate([b for _, b in lists]))
    d_out = gpu.DeviceBuffer(4 * V * len(lists))
    for graphs in (False, True, True, True):
        sur.set_graphs(graphs)
        d_out.upload(np.full(V * len(lists), np.nan, np.float32))
        errs, used = sur.map_batch_dev(d_s.ptr, d_q.ptr, counts, d_out.ptr, 1e-3, gate_sites=1024)
        got = d_out.download(np.float32, V * len(lists)).reshape(len(lists), V)
        assert used.tolist() == [True, True, False]
        assert np.array_equal(got[0], exact[0]) and np.array_equal(got[1], exact[1])
        assert not got[2].any()
        ctx.potential_map(*lists[1], flags=fl)   # other work on the stream between replays


def solar_snapshots(gpu, sur, dims, cutoff, count, stride, seed):
    p = gpu.SimulationParameters(
        simulation_type="solarcell", grid_x=dims[0], grid_y=dims[1], grid_z=dims[2], coulomb_carriers=True,
        electron_percentage=0.01, hole_percentage=0.01, seed_charges=1.0, generation_rate=0.2,
        electrostatic_cutoff=cutoff, random_seed=seed)
    rng = np.random.default_rng(seed)
    V = p.volume
    e = rng.choice(V, p.max_electrons(), replace=False).astype(np.int32)
    h = rng.choice(V, p.max_holes(), replace=False).astype(np.int32)
    sim = gpu.Simulation(p, electrons=e, holes=h, context=sur.ctx)
    snaps = gpu.SnapshotSet(sur, count)
    snaps.add_simulation(sim, count, stride)
    return snaps


def test_training_learns_the_map_operator(gpu, oracle):
    """Trained from zero on snapshots of a solar-cell run, the network meets the
    1e-3 gate on held-out snapshots, where the untrained one fails it."""
    dims, cutoff = (32, 32, 8), 8
    ctx, sur = make(gpu, dims, cutoff, max_batch=8)
    snaps = solar_snapshots(gpu, sur, dims, cutoff, 40, 20, seed=11)
    hist = snaps.train(batch=8, iterations=400, momentum=0.9, seed=5, first=0, count=32)
    print(f"loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[0] > 0.5 and hist[-1] < 1e-5
    for i in range(32, 40):
        occ = snaps.occupancy(i)
        got = sur.infer(occ)[0]
        err = rel_rms(got, snaps.targets(i).astype(np.float64))
        assert err <= 1e-3, f"held-out snapshot {i}: rel RMS {err:.2e}"
    # the held-out labels are the exact kernel's; check one against the FP64 oracle too
    occ = snaps.occupancy(39).astype(np.int64)
    sites = np.nonzero(occ)[0].astype(np.int32)
    ref = oracle.potential_map(dims, cutoff, sites, occ[sites].astype(np.int32), pref=PREF)
    assert rel_rms(sur.infer(occ.astype(np.int8))[0], ref) <= 1e-3
    # gated map of the same snapshot: accepted, no fallback
    out, err, used = sur.map(sites, occ[sites].astype(np.int32), tolerance=1e-3)
    assert not used and err <= 1e-3 and rel_rms(out, ref) <= 1e-3


def test_training_is_deterministic(gpu):
    dims, cutoff = (16, 16, 8), 6
    ws = []
    for _ in range(2):
        ctx, sur = make(gpu, dims, cutoff, max_batch=4)
        snaps = solar_snapshots(gpu, sur, dims, cutoff, 8, 10, seed=3)
        snaps.train(batch=4, iterations=30, seed=9)
        ws.append(sur.get_weights())
    assert np.array_equal(ws[0], ws[1])


def test_full_size_gain_over_exact_map(gpu):
    """128x128x32 solar lattice, cutoff 30: the trained surrogate (gate
    included) against the exact map kernels (all pairs, and culled) on held-out snapshots."""
    dims, cutoff = (128, 128, 32), 30
    ctx, sur = make(gpu, dims, cutoff, max_batch=8)
    assert 1.5e6 < 2 * sur.n_weights < 2.0e6
    snaps = solar_snapshots(gpu, sur, dims, cutoff, 28, 40, seed=21)
    hist = snaps.train(batch=8, iterations=800, seed=1, first=0, count=24)
    for i in range(24, 28):
        err = rel_rms(sur.infer(snaps.occupancy(i))[0], snaps.targets(i).astype(np.float64))
        assert err <= 1e-3, f"held-out snapshot {i}: rel RMS {err:.2e}"
    occ = snaps.occupancy(27).astype(np.int64)
    sites = np.nonzero(occ)[0].astype(np.int32)
    q = occ[sites].astype(np.int32)
    exact = snaps.targets(27).astype(np.float64)
    d_s, d_q = gpu.DeviceBuffer.from_array(sites), gpu.DeviceBuffer.from_array(q)
    d_out = gpu.DeviceBuffer(4 * sur.V)
    err, used = sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 1e-3)
    full = rel_rms(d_out.download(np.float32, sur.V), exact)
    print(f"loss {hist[0]:.2e} -> {hist[-1]:.2e}; gate {err:.2e}, full rel RMS {full:.2e}")
    assert not used and full <= 1e-3

    def timed(fn, reps=10):
        fn()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t) / reps

    t_sur = timed(lambda: sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 1e-3, wait=False))
    assert sur.last_gate()[1] is False
    # the exact map in the reference's form (every charge at every site), and the culled
    # exact map this lattice gets by default
    t_pairs = timed(lambda: ctx.potential_map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, gpu.LMG_MAP_PAIRS))
    t_culled = timed(lambda: ctx.potential_map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, gpu.LMG_MAP_CULLED))
    print(f"surrogate map {t_sur * 1e3:.3f} ms; exact map all pairs {t_pairs * 1e3:.3f} ms "
          f"({t_pairs / t_sur:.1f}x), culled {t_culled * 1e3:.3f} ms ({t_culled / t_sur:.1f}x)")
    assert t_pairs / t_sur >= 5.0
    assert t_culled / t_sur >= 2.0
