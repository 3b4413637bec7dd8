#!/usr/bin/env python3
"""Benchmark of the MI355X Langmuir hot path.

Default workload (the configuration the headline metric is quoted on): the 3-D
bulk-heterojunction solar cell, 128x128x32 lattice, two-phase morphology
(uniform noise smoothed by a Gaussian of sigma = 4 sites, thresholded at its
median; one phase carries 0.5 eV traps), 5242 electrons + 5242 holes (1 % of
sites each) at uniformly random sites, exciton generation 0.524288 per step
(1e-3 /site/ns at 1 ps steps over 128x128), Coulomb cutoff 30, drains 0.9 on
both x faces.  One step = one KMC sweep (proposals, dE of every carrier, decisions,
conflicts, removals, injection) on the device, with all state resident in HBM.

Prints ONE JSON line.  ``value`` is KMC steps/s; ``pair_evals_per_s`` counts
2 N (N + Nd) pair evaluations per step (the reference's all-pairs loop).
``roofline`` is for the dominant kernel (the pair-sum kernel: cull_de_kernel
on the culled path, de_kernel on the all-pairs path), FP32 VALU bound: 11 flops
per in-cutoff pair term, terms counted exactly on the timed state, over the
kernel's own event-timed duration.
``map`` times the full-lattice potential map (256x256x64, 50,000 charges): cutoff 50
(culled, and all pairs) and no cutoff.
``batch`` advances 256 independent replicas of uneven size together.
``surrogate`` trains the learned map surrogate on snapshots of the solar run
and times its gated map against the exact map on held-out charges: the culled
exact map (the default at that size, ``speedup_vs_exact``) and the all-pairs
form of the reference (``speedup_vs_exact_all_pairs``).
``cpu_baseline`` times the FP64 oracle (a restatement of the reference CPU path,
OpenMP over carriers like QtConcurrent::map) on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "KMC steps/s + Coulomb pair-evals/s; surrogate vs exact maps/s at ≤1e-3 err"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# synthetic inputs
# ---------------------------------------------------------------------------
def morphology_traps(dims, sigma=4.0, seed=2024):
    """Two-phase morphology made the LangmuirPython way (uniform noise in [-0.5, 0.5)
    convolved with a normalised Gaussian, sigma 4 truncated at 3 sigma, mode 'same'),
    split at its median; returns the site ids of one phase (50 % of the lattice)."""
    from langmuir_amd.morphology import two_phase_traps
    return two_phase_traps(dims, sigma, seed)[0]


def solar_workload(seed=2024):
    from langmuir_amd.host import SimulationParameters, background_potential
    p = SimulationParameters(
        simulation_type="solarcell", grid_x=128, grid_y=128, grid_z=32, coulomb_carriers=True,
        electron_percentage=0.01, hole_percentage=0.01, seed_charges=1.0,
        generation_rate=0.524288, source_scale_area=0.0, electrostatic_cutoff=30,
        trap_potential=0.5, random_seed=12345)
    V = p.volume
    rng = np.random.default_rng(seed)
    n = p.max_electrons()
    e = rng.choice(V, n, replace=False).astype(np.int32)
    h = rng.choice(V, n, replace=False).astype(np.int32)
    traps = morphology_traps((128, 128, 32), seed=seed)
    pot = background_potential(p, traps, np.full(len(traps), 0.5))
    return p, e, h, traps, pot


def transistor_workload(seed=2024):
    from langmuir_amd.host import SimulationParameters
    p = SimulationParameters(
        grid_x=256, grid_y=256, grid_z=1, coulomb_carriers=True, electron_percentage=0.0,
        hole_percentage=0.02, seed_charges=1.0, voltage_left=1.0, e_source_l_rate=0.0,
        h_source_l_rate=0.9, h_drain_r_rate=0.9, random_seed=12345)
    rng = np.random.default_rng(seed)
    h = rng.choice(p.volume, p.max_holes(), replace=False).astype(np.int32)
    return p, np.zeros(0, np.int32), h, np.zeros(0, np.int32), None


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
WARM_S = 2.0   # seconds of untimed steps before every timed window (shader clocks settle)


def site_xyz(dims, sites):
    """Site ids -> (x, y, z) by the reference's decode (cubicgrid.cpp:132-165), drain
    slots V+k included (they decode past the lattice, as kernel.cl's staging does)."""
    X, Y, Z = dims
    s = np.asarray(sites, dtype=np.int64)
    z = s // (X * Y)
    return np.stack([s % X, s // X - z * Y, z], 1).astype(np.float64)


def in_cutoff_terms(dims, cutoff, targets, charges):
    """Exact count of (target, charge) pairs with 0 < r^2 < c^2 (integer r^2: r^2 <= c^2 - 1,
    so a pair-count radius of sqrt(c^2 - 1/2) and a same-site radius of 1/2 are exact)."""
    from scipy.spatial import cKDTree
    if len(targets) == 0 or len(charges) == 0:
        return 0
    t, c = cKDTree(site_xyz(dims, targets)), cKDTree(site_xyz(dims, charges))
    return int(t.count_neighbors(c, np.sqrt(cutoff * cutoff - 0.5))) - int(t.count_neighbors(c, 0.5))


def map_in_cutoff_terms(dims, cutoff, sites):
    """Exact count of (lattice site, charge) pairs with 0 < r^2 < c^2 for charges at distinct
    sites: per charge, the lattice rows (dy, dz) inside the sphere, clipped to the box."""
    X, Y, Z = dims
    xyz = site_xyz(dims, sites).astype(np.int32)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    c2 = cutoff * cutoff
    isqrt = np.array([int(np.floor(np.sqrt(v))) for v in range(c2)], dtype=np.int32)
    isqrt = np.array([w - 1 if w * w > v else (w + 1 if (w + 1) ** 2 <= v else w) for v, w in enumerate(isqrt)],
                     dtype=np.int32)
    dy = np.arange(-(cutoff - 1), cutoff, dtype=np.int32)
    total = 0
    for dz in range(-(cutoff - 1), cutoff):
        rem = c2 - 1 - dy * dy - dz * dz
        keep = rem >= 0
        zz = z + dz
        zok = (zz >= 0) & (zz < Z)
        if not keep.any() or not zok.any():
            continue
        d, w = dy[keep], isqrt[rem[keep]]
        xs, ys = x[zok], y[zok]
        yy = ys[None, :] + d[:, None]
        lo = np.maximum(xs[None, :] - w[:, None], 0)
        hi = np.minimum(xs[None, :] + w[:, None], X - 1)
        total += int(np.where((yy >= 0) & (yy < Y), hi - lo + 1, 0).sum(dtype=np.int64))
    return total - len(sites)   # the charge's own site (r = 0)


def warm_until(step, sync, seconds, chunk):
    """Untimed steps for a fixed wall time."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        step(chunk)
        sync()
        n += chunk
    return n


def cpu_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return dict(nproc=os.cpu_count(), affinity=affinity, model=model)


def rel_errors(got, ref, floor=0.0):
    """max |err|, RMS relative error ||err|| / ||ref||, max |err| / max(|ref|, floor)."""
    got, ref = np.asarray(got, np.float64), np.asarray(ref, np.float64)
    err = np.abs(got - ref)
    return dict(max_abs=float(err.max()) if err.size else 0.0,
                rms_rel=float(np.linalg.norm(err) / max(np.linalg.norm(ref), 1e-300)),
                max_rel=float((err / np.maximum(np.abs(ref), floor if floor > 0 else 1e-300)).max()) if err.size else 0.0)


class HipTimer:
    """HIP events on an explicit stream (the context's), via libamdhip64."""

    def __init__(self, stream):
        self.hip = C.CDLL("libamdhip64.so")
        self.stream = C.c_void_p(stream)
        self.a, self.b = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.a)) == 0
        assert self.hip.hipEventCreate(C.byref(self.b)) == 0

    def start(self):
        assert self.hip.hipEventRecord(self.a, self.stream) == 0

    def stop_ms(self):
        assert self.hip.hipEventRecord(self.b, self.stream) == 0
        assert self.hip.hipEventSynchronize(self.b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), self.a, self.b) == 0
        return float(ms.value)


MAP_DIMS = (256, 256, 64)
MAP_SAMPLE = 4096


def map_workload():
    """256x256x64 lattice, 50,000 distinct uniform-random sites, 25,000 at +1 and 25,000 at -1."""
    V = MAP_DIMS[0] * MAP_DIMS[1] * MAP_DIMS[2]
    rng = np.random.default_rng(7)
    n = 50000
    cs = rng.choice(V, n, replace=False).astype(np.int32)
    cq = np.where(np.arange(n) < n // 2, 1, -1).astype(np.int32)
    sample = np.sort(rng.choice(V, MAP_SAMPLE, replace=False)).astype(np.int32)
    return cs, cq, sample


def bench_map(reps=3):
    """Full-lattice map, 256x256x64, 50,000 charges, cutoff 50 (culled and all pairs) and
    no cutoff.  Roofline per variant: 11 FP32 flops per in-cutoff (site, charge) term over
    the map's event-timed duration.  The maps' values at MAP_SAMPLE seeded sites are kept
    for the FP64 comparison in the CPU leg."""
    import langmuir_amd as L
    dims = MAP_DIMS
    V = dims[0] * dims[1] * dims[2]
    cs, cq, sample = map_workload()
    n = len(cs)
    t0 = time.perf_counter()
    terms50 = map_in_cutoff_terms(dims, 50, cs)
    terms_all = V * n - n
    log(f"map: in-cutoff terms {terms50:.4e} at cutoff 50, {terms_all:.4e} without ({time.perf_counter() - t0:.1f} s)")
    out, samples = {}, {}
    ctx = L.Context(dims, cutoff=50)
    d_cs, d_cq = L.DeviceBuffer.from_array(cs), L.DeviceBuffer.from_array(cq)
    d_out = L.DeviceBuffer(4 * V)
    timer = HipTimer(ctx.stream)
    for name, flags, terms in (("cutoff50", 0, terms50), ("cutoff50_all_pairs", L.LMG_MAP_PAIRS, terms50),
                               ("nocutoff", L.LMG_NO_CUTOFF, terms_all)):
        warm_until(lambda k: [ctx.potential_map_dev(d_cs.ptr, d_cq.ptr, n, d_out.ptr, flags) for _ in range(k)],
                   ctx.sync, 0.5, 1)
        timer.start()
        for _ in range(reps):
            ctx.potential_map_dev(d_cs.ptr, d_cq.ptr, n, d_out.ptr, flags)
        ms = timer.stop_ms() / reps
        v = d_out.download(np.float32, V)
        samples[name] = v[sample].copy()
        tflops = 11.0 * terms / (ms * 1e-3) / 1e12
        out[name] = dict(ms_per_map=round(ms, 4), maps_per_s=round(1e3 / ms, 3),
                         pair_evals_per_s=V * n / (ms * 1e-3),
                         roofline={"bound": "valu", "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS,
                                   "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 5),
                                   "per_launch": f"11 flops x {terms} in-cutoff terms"})
        log(f"map {name}: {ms:.3f} ms/map, {V * n / (ms * 1e-3):.3e} pair-evals/s, {tflops:.2f} TFLOP/s")
    for a in (d_cs, d_cq, d_out):
        a.free()
    ctx.close()
    return out, dict(sample=sample, values=samples)


BATCH_LATTICES = ((32, 32, 32), (48, 48, 24), (64, 64, 16), (64, 64, 32), (96, 96, 16), (128, 128, 16))


def batch_workload(n_rep=256, seed=4):
    """Independent transistor-type replicas: lattice drawn from 32^3 .. 128x128x16,
    1-V bias along x, holes N ~ U[100, 5000] (at most 10 % of sites) at random
    sites, trap fraction ~ U[0, 0.1] with depths clip(N(0.3, 0.1), 0.1, 0.5) eV,
    cutoff ~ U{20..50}, per-replica seeds."""
    from langmuir_amd.host import SimulationParameters, background_potential
    rng = np.random.default_rng(seed)
    reps = []
    for r in range(n_rep):
        X, Y, Z = BATCH_LATTICES[rng.integers(len(BATCH_LATTICES))]
        V = X * Y * Z
        n = int(min(rng.integers(100, 5001), V // 10))
        p = SimulationParameters(
            grid_x=X, grid_y=Y, grid_z=Z, coulomb_carriers=True, electron_percentage=0.0,
            hole_percentage=(n + 0.5) / V, seed_charges=1.0, voltage_left=1.0, e_source_l_rate=0.0,
            h_source_l_rate=0.9, h_drain_r_rate=0.9, electrostatic_cutoff=int(rng.integers(20, 51)),
            random_seed=int(rng.integers(1, 2 ** 31)))
        nt = int(rng.uniform(0.0, 0.1) * V)
        traps = rng.choice(V, nt, replace=False)
        depth = np.clip(rng.normal(0.3, 0.1, nt), 0.1, 0.5)
        holes = rng.choice(V, n, replace=False).astype(np.int32)
        reps.append(dict(params=p, holes=holes, potential=background_potential(p, traps, depth)))
    return reps


def bench_batch(steps=200, warmup=20, n_rep=256, de_path="auto", graphs=True):
    """256 independent replicas stepped together.  value: replica-steps/s over `steps` graph
    replays after a fixed-time warm-up; roofline of the replica-wide dE kernel from a
    kernel-by-kernel run with events around it and exact in-cutoff term counts of every
    replica's pending step."""
    import langmuir_amd as L
    reps = batch_workload(n_rep)
    t0 = time.perf_counter()
    batch = L.Batch(reps, de_path=de_path)
    batch.set_graphs(graphs)
    log(f"batch of {n_rep} replicas created in {time.perf_counter() - t0:.1f} s")
    # The replicas lose carriers through their drains as they run, so the clocks are warmed
    # on a second copy of the batch and the timed one starts from the configured state.
    heater = L.Batch(reps, de_path=de_path)
    heater.set_graphs(graphs)
    warm_until(heater.step, heater.sync, WARM_S, 20)
    heater.close()
    batch.step(warmup)
    batch.sync()
    t0 = time.perf_counter()
    ms = batch.step_timed(steps)
    wall = time.perf_counter() - t0
    ns, terms = [], 0
    for r in range(n_rep):
        st = batch.state(r)
        fe, fh, _, _ = batch.delta_e(r)
        p = reps[r]["params"]
        sites = np.concatenate([st["e_site"], st["h_site"]])
        ns.append(len(sites))
        terms += in_cutoff_terms((p.grid_x, p.grid_y, p.grid_z), p.electrostatic_cutoff,
                                 np.concatenate([sites, fe, fh]), sites)
    paths = batch.de_paths
    n_ph = 50
    de_ms, rest_ms, span_ms = batch.step_phases(n_ph)
    de_us = de_ms * 1e3 / n_ph
    tflops = 11.0 * terms / (de_us * 1e-6) / 1e12
    pair_evals = sum(2.0 * n * n for n in ns)
    out = dict(replicas=n_rep, steps=steps, ms_per_step=round(wall * 1e3 / steps, 4),
               event_ms_per_step=round(ms / steps, 4),
               replica_steps_per_s=round(n_rep * steps / wall, 1),
               pair_evals_per_s=pair_evals * steps / wall,
               carriers_total=int(sum(ns)), graph_replay=graphs, autotuned=batch.tuning,
               culled_replicas=paths.count("culled"),
               kernel_us={"de_phase": round(de_us, 2), "rest_of_step": round(rest_ms * 1e3 / n_ph, 2),
                          "event_span_per_step": round(span_ms * 1e3 / n_ph, 2),
                          "source": f"{n_ph} steps kernel by kernel, HIP events around the dE phase"},
               roofline={"bound": "valu",
                         "kernel": "dE phase of all replicas (culled replicas: cull_emit_kernel + cull_de_kernel, "
                                   "one launch each; the rest: de_kernel, one launch)",
                         "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / FP32_PEAK_TFLOPS, 5),
                         "per_launch": f"11 flops x {terms} in-cutoff terms over all replicas"})
    log(f"batch: {out['ms_per_step']} ms/step for {n_rep} replicas -> "
        f"{out['replica_steps_per_s']:.0f} replica-steps/s ({out['culled_replicas']} culled, "
        f"{out['autotuned']['cull_workgroups']} culled workgroups, all-pairs chunks "
        f"{out['autotuned']['charge_chunks']}); dE phase {de_us:.0f} us, {terms:.3e} terms, {tflops:.2f} TFLOP/s")
    batch.close()
    return out


BF16_PEAK_TFLOPS = 2500.0


def surrogate_gemm_shapes(dims, padded):
    """(M, N, K, batch) of the six DFT GEMMs of one surrogate map."""
    X, Y, Z = dims
    px, py, pz = padded
    kx2 = 2 * (px // 2 + 1)
    return [(Z * Y, kx2, X, 1), (py, kx2, 2 * Y, Z), (pz, py * kx2, 2 * Z, 1),
            (Z, py * kx2, 2 * pz, 1), (Y, kx2, 2 * py, Z), (Z * Y, X, kx2, 1)]


def bench_surrogate(workload, n_snap=64, stride=25, iters=1200, reps=50):
    """Learned map surrogate on the solar lattice (128x128x32, cutoff 30):
    snapshots of the KMC run on the device labelled by the exact map kernel,
    training on the device, then held-out accuracy and gated-map throughput
    against the exact map on the same charges."""
    return _bench_surrogate(workload, n_snap, stride, iters, reps)


def _bench_surrogate(workload, n_snap, stride, iters, reps):
    import langmuir_amd as L
    p, e, h, traps, pot = workload
    dims = (p.grid_x, p.grid_y, p.grid_z)
    ctx = L.Context.from_parameters(p)
    sur = L.Surrogate(ctx, max_batch=8)
    sim = L.Simulation(p, electrons=e, holes=h, potential=pot, context=ctx)
    sim.step(200)
    snaps = L.SnapshotSet(sur, n_snap)
    t0 = time.perf_counter()
    snaps.add_simulation(sim, n_snap, stride)
    t_data = time.perf_counter() - t0
    n_train = n_snap - 8
    t0 = time.perf_counter()
    hist = snaps.train(batch=8, iterations=iters, momentum=0.9, seed=1, first=0, count=n_train)
    t_train = time.perf_counter() - t0
    errs = []
    for i in range(n_train, n_snap):
        got = sur.infer(snaps.occupancy(i))[0]
        t = snaps.targets(i).astype(np.float64)
        errs.append(float(np.linalg.norm(got - t) / np.linalg.norm(t)))
    occ = snaps.occupancy(n_snap - 1).astype(np.int64)
    sites = np.nonzero(occ)[0].astype(np.int32)
    q = occ[sites].astype(np.int32)
    d_s, d_q = L.DeviceBuffer.from_array(sites), L.DeviceBuffer.from_array(q)
    d_out = L.DeviceBuffer(4 * sur.V)
    d_occ = L.DeviceBuffer(sur.V)
    sur.occupancy_dev(d_s.ptr, d_q.ptr, len(sites), d_occ.ptr)
    timer = HipTimer(ctx.stream)

    def timed(fn, n):
        fn()
        ctx.sync()
        timer.start()
        for _ in range(n):
            fn()
        return timer.stop_ms() / n

    ms_exact = timed(lambda: ctx.potential_map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 0), 5)   # culled here
    ms_pairs = timed(lambda: ctx.potential_map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, L.LMG_MAP_PAIRS), 5)
    ms_gated = timed(lambda: sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 1e-3, wait=False), reps)
    gate_err, used = sur.last_gate()
    full_err = float(np.linalg.norm(d_out.download(np.float32, sur.V) - snaps.targets(n_snap - 1))
                     / np.linalg.norm(snaps.targets(n_snap - 1)))
    ms_infer = timed(lambda: sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 0.0, wait=False), 5)   # gate fails
    _, fb_used = sur.last_gate()
    ms_fallback, ms_infer = ms_infer, timed(lambda: sur.infer_dev(d_occ.ptr, 1, d_out.ptr), reps)
    # the eight held-out snapshots as one batch of gated maps
    nb = n_snap - n_train
    lists = []
    for i in range(n_train, n_snap):
        o = snaps.occupancy(i).astype(np.int64)
        si = np.nonzero(o)[0].astype(np.int32)
        lists.append((si, o[si].astype(np.int32)))
    counts = [len(a) for a, _ in lists]
    b_s = L.DeviceBuffer.from_array(np.concatenate([a for a, _ in lists]))
    b_q = L.DeviceBuffer.from_array(np.concatenate([c for _, c in lists]))
    b_out = L.DeviceBuffer(4 * sur.V * nb)
    ms_batch = timed(lambda: sur.map_batch_dev(b_s.ptr, b_q.ptr, counts, b_out.ptr, 1e-3, wait=False), reps)
    b_err, b_used = sur.last_gates(nb)
    b_maps = b_out.download(np.float32, sur.V * nb).reshape(nb, sur.V)
    b_full = [float(np.linalg.norm(b_maps[k] - snaps.targets(n_train + k)) / np.linalg.norm(snaps.targets(n_train + k)))
              for k in range(nb)]
    us_per_map = ms_batch * 1e3 / nb
    log(f"surrogate batch of {nb} gated maps: {ms_batch * 1e3:.1f} us ({us_per_map:.1f} us per map, gate errors <= "
        f"{b_err.max():.2e}, exact used {int(b_used.sum())}/{nb}, full-map rel RMS <= {max(b_full):.2e}) -> "
        f"{ms_exact * 1e3 / us_per_map:.1f}x the culled exact map, {ms_pairs * 1e3 / us_per_map:.1f}x all pairs")
    # the same pipeline with the exact operator's spectrum (computed, not learned)
    sur.set_analytic_weights()
    a_errs = []
    for i in range(n_train, n_snap):
        got = sur.infer(snaps.occupancy(i))[0]
        t = snaps.targets(i).astype(np.float64)
        a_errs.append(float(np.linalg.norm(got - t) / np.linalg.norm(t)))
    ms_analytic = timed(lambda: sur.map_dev(d_s.ptr, d_q.ptr, len(sites), d_out.ptr, 1e-3, wait=False), reps)
    a_gate, a_used = sur.last_gate()
    log(f"surrogate with the analytic spectrum: held-out rel RMS max {max(a_errs):.2e} (trained {max(errs):.2e}); "
        f"gated map {ms_analytic * 1e3:.1f} us (gate {a_gate:.2e}, exact used {a_used})")
    shapes = surrogate_gemm_shapes(dims, sur.padded)
    alg = sum(2.0 * m * n * k * b for m, n, k, b in shapes)
    executed = alg * 3 - 2.0 * shapes[0][0] * shapes[0][1] * shapes[0][2]   # int8 stage: two products
    tflops = alg / (ms_infer * 1e-3) / 1e12                  # algorithmic flops: one product per MAC
    tflops_batch = alg * nb / (ms_batch * 1e-3) / 1e12       # the batch of gated maps (gates included)
    pz, py, kxn = sur.spectrum_shape
    free = (pz // 2 + 1) * (py // 2 + 1) * kxn
    log(f"surrogate: {n_snap} snapshots in {t_data:.1f} s, {iters} steps in {t_train:.1f} s (loss "
        f"{hist[0]:.2e} -> {hist[-1]:.2e}); held-out rel RMS max {max(errs):.2e}; gated map {ms_gated * 1e3:.1f} us "
        f"(gate {gate_err:.2e}, exact used {used}), infer {ms_infer * 1e3:.1f} us; exact map culled "
        f"{ms_exact * 1e3:.1f} us -> {ms_exact / ms_gated:.1f}x, all pairs {ms_pairs * 1e3:.1f} us -> "
        f"{ms_pairs / ms_gated:.1f}x; a failed gate (exact fallback) {ms_fallback * 1e3:.1f} us")
    return {
        "maps_per_s": round(1e3 / ms_gated, 2), "exact_maps_per_s": round(1e3 / ms_exact, 2),
        "speedup_vs_exact": round(ms_exact / ms_gated, 2),
        "speedup_vs_exact_all_pairs": round(ms_pairs / ms_gated, 2),
        "gated_map_us": round(ms_gated * 1e3, 2), "infer_us": round(ms_infer * 1e3, 2),
        "exact_map_us": round(ms_exact * 1e3, 2), "exact_map_path": "culled",
        "exact_map_all_pairs_us": round(ms_pairs * 1e3, 2),
        "rel_rms_heldout_max": max(errs), "gate_error": gate_err, "gate_sites": 1024, "used_exact": used,
        "rel_rms_gated_map": full_err,
        "failed_gate_map_us": round(ms_fallback * 1e3, 2), "failed_gate_used_exact": fb_used,
        "batch": {"maps": nb, "gated_batch_us": round(ms_batch * 1e3, 2), "us_per_map": round(us_per_map, 2),
                  "maps_per_s": round(1e6 / us_per_map, 1),
                  "speedup_vs_exact": round(ms_exact * 1e3 / us_per_map, 2),
                  "speedup_vs_exact_all_pairs": round(ms_pairs * 1e3 / us_per_map, 2),
                  "gate_error_max": float(b_err.max()), "used_exact": int(b_used.sum()),
                  "rel_rms_max": max(b_full)},
        "analytic": {"rel_rms_heldout_max": max(a_errs), "gated_map_us": round(ms_analytic * 1e3, 2),
                     "gate_error": a_gate, "used_exact": a_used,
                     "note": "the same pipeline with the rfftn of the map kernel as weights: the operator is "
                             "linear and known, so training reproduces it from snapshots but does not beat it"},
        "weights_stored": 2 * sur.n_weights, "weights_free": free, "padded": list(sur.padded),
        "snapshots": n_snap, "snapshot_stride_steps": stride, "train_snapshots": n_train,
        "train_iterations": iters, "train_s": round(t_train, 3), "loss_first_last": [float(hist[0]), float(hist[-1])],
        "mfma": {"bound": "mfma", "algorithmic_gflop_per_map": round(alg / 1e9, 4),
                 "executed_bf16_gflop_per_map": round(executed / 1e9, 4),
                 "achieved_tflops": round(tflops, 2), "peak_tflops": BF16_PEAK_TFLOPS,
                 "frac": round(tflops / BF16_PEAK_TFLOPS, 5),
                 "batch_achieved_tflops": round(tflops_batch, 2),
                 "batch_frac": round(tflops_batch / BF16_PEAK_TFLOPS, 5),
                 "note": "algorithmic flops (2 M N K per DFT GEMM) over one inference pass; the BF16 "
                         "hi/lo split executes about 2.9x as many MFMA flops"},
    }


def measured_traffic(kernel, workload="solar"):
    """HBM bytes per launch of a kernel from the committed counter summary (measured on the
    solar workload), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            t = json.load(f)
        if t.get("workload", "solar") != workload:
            return None
        return t["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def oracle_world(workload):
    """The FP64 oracle's world of a bench workload (same lists, traps and engine seed)."""
    from oracle import pyoracle as O
    p, e, h, traps, pot = workload
    V = p.volume
    kw = dict(simulation_type=1 if p.is_solarcell else 0, grid_x=p.grid_x, grid_y=p.grid_y,
              grid_z=p.grid_z, coulomb_carriers=1, electron_percentage=p.electron_percentage,
              hole_percentage=p.hole_percentage, seed_charges=0.0, generation_rate=p.generation_rate,
              source_scale_area=p.source_scale_area, electrostatic_cutoff=p.electrostatic_cutoff,
              voltage_left=p.voltage_left, voltage_right=p.voltage_right,
              e_source_l_rate=p.e_source_l_rate, h_source_l_rate=p.h_source_l_rate,
              h_drain_r_rate=p.h_drain_r_rate, random_seed=p.random_seed,
              trap_percentage=(len(traps) + 0.5) / V if len(traps) else 0.0, trap_potential=0.5)
    return O.World(O.OrcParams.make(**kw), electrons=e, holes=h, traps=traps,
                   trap_potentials=np.full(len(traps), 0.5))


def cpu_baseline(workload, budget_s=5.0, runs=3):
    """The FP64 oracle (restated reference CPU path, OpenMP over carriers as QtConcurrent::map
    spreads them) on the same workload from the same initial state, every core in the
    affinity mask; median of `runs` samples of about budget_s seconds each."""
    info = cpu_info()
    threads = max(1, info["affinity"])
    p, e, h, _, _ = workload
    w = oracle_world(workload)
    t0 = time.perf_counter()
    w.step(1, threads)
    per = time.perf_counter() - t0
    steps = int(max(2, min(5000, budget_s / max(per, 1e-6))))
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        w.step(steps, threads)
        rates.append(steps / (time.perf_counter() - t0))
    n = len(e) + len(h)
    med = float(np.median(rates))
    return dict(value=round(med, 4), unit="KMC steps/s", cores=threads, kind="port",
                sample=f"median of {runs} runs of {steps} steps each, continuing from the workload's initial "
                       f"state (N={n} carriers at the start), FP64 oracle, OpenMP over carriers on {threads} "
                       f"threads",
                runs=[round(r, 4) for r in rates], nproc=info["nproc"], cpu_model=info["model"],
                pair_evals_per_s=2.0 * n * n * med)


def de_vs_fp64(workload, threads):
    """dE of every carrier's first proposal on the device against the FP64 oracle from the
    same state and engine (proposals must agree exactly)."""
    import langmuir_amd as L
    p, e, h, _, pot = workload
    sim = L.Simulation(p, electrons=e, holes=h, potential=pot)
    fe, fh, de_e, de_h = sim.delta_e()
    sim.ctx.close()
    w = oracle_world(workload)
    w.propose()
    w.coulomb(threads)
    ce, ch = w.carriers(0), w.carriers(1)
    assert np.array_equal(fe, ce["fsite"]) and np.array_equal(fh, ch["fsite"]), "proposals differ"
    got, ref = np.concatenate([de_e, de_h]), np.concatenate([ce["de"], ch["de"]])
    r = rel_errors(got, ref)
    tol = 1e-5 * np.abs(ref) + 2e-7
    r.update(carriers=int(len(ref)), within_tolerance=bool(np.all(np.abs(got - ref) <= tol)),
             tolerance="1e-5 |dE| + 2e-7 eV")
    return r


def maps_vs_fp64(samples, threads):
    """The device maps at the sampled sites against the FP64 oracle's pair sums there."""
    from oracle import pyoracle as O
    cs, cq, sample = map_workload()
    out = {}
    # no cutoff: any cutoff beyond the lattice diagonal (sqrt(255^2 + 255^2 + 63^2) < 366)
    for name, cutoff in (("cutoff50", 50), ("nocutoff", 400)):
        ref = O.pair_potential(MAP_DIMS, cutoff, cs, cq, sample)
        r = rel_errors(samples["values"][name], ref)
        r["sites"] = int(len(sample))
        out[name] = r
    return out


# ---------------------------------------------------------------------------
def step_leg(workload, steps, warmup, de_path, breakdown_steps, graphs=True, de_shape=0):
    """One simulation: `warmup` steps plus WARM_S seconds of untimed steps, then `steps`
    timed graph replays between device synchronisations; then a kernel-by-kernel run of
    `breakdown_steps` steps with HIP events for the phase and pair-sum-kernel times, with
    exact in-cutoff term counts of the states entering and leaving it."""
    import langmuir_amd as L
    p, e, h, traps, pot = workload
    dims = (p.grid_x, p.grid_y, p.grid_z)
    sim = L.Simulation(p, electrons=e, holes=h, potential=pot, de_path=de_path, de_chunks=de_shape)
    sim.set_graphs(graphs)
    sim.step(warmup)
    sim.ctx.sync()
    warmed = warm_until(sim.step, sim.ctx.sync, WARM_S, 200)
    st = sim.state()
    n_start = len(st["e_site"]) + len(st["h_site"])
    t0 = time.perf_counter()
    sim.step(steps)
    sim.ctx.sync()
    wall = time.perf_counter() - t0
    assert sim.device_error() == 0
    st = sim.state()
    n_end = len(st["e_site"]) + len(st["h_site"])

    def pending_terms():
        s = sim.state()
        fe, fh, _, _ = sim.delta_e()
        sites = np.concatenate([s["e_site"], s["h_site"]])
        return in_cutoff_terms(dims, p.electrostatic_cutoff, np.concatenate([sites, fe, fh]), sites)

    fe0, fh0, _, _ = sim.delta_e()    # the breakdown starts from this state
    s0 = sim.state()
    de_ms, commit_ms, span_ms, pair_ms = sim.step_timed(breakdown_steps)
    assert sim.device_error() == 0
    sites0 = np.concatenate([s0["e_site"], s0["h_site"]])
    terms = 0.5 * (in_cutoff_terms(dims, p.electrostatic_cutoff, np.concatenate([sites0, fe0, fh0]), sites0)
                   + pending_terms())
    kname = "cull_de_kernel" if sim.de_path == "culled" else "de_kernel"
    pair_us = pair_ms * 1e3 / breakdown_steps
    tflops = 11.0 * terms / (pair_us * 1e-6) / 1e12
    n_avg = 0.5 * (n_start + n_end)
    out = dict(steps_per_s=steps / wall, ms_per_step=wall * 1e3 / steps, warm_steps=int(warmup + warmed),
               carriers_start=int(n_start), carriers_end=int(n_end), de_path=sim.de_path, de_shape=sim.tuning,
               pair_evals_per_s=2.0 * n_avg * n_avg * steps / wall,
               kernel_us={"de_phase": round(de_ms * 1e3 / breakdown_steps, 3), kname: round(pair_us, 3),
                          "commit_phase": round(commit_ms * 1e3 / breakdown_steps, 3),
                          "event_span_per_step": round(span_ms * 1e3 / breakdown_steps, 3),
                          "source": f"{breakdown_steps} steps right after the timed window, kernel by kernel, "
                                    "HIP events between the phases"},
               roofline={"bound": "valu", "kernel": kname, "achieved": round(tflops, 4),
                         "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 5),
                         "traffic": measured_traffic(kname, "solar" if p.is_solarcell else "transistor"),
                         "traffic_source": "profiles/hbm_traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                           "scaled by a 256 MiB copy probe)",
                         "per_launch": f"11 flops x {int(terms)} in-cutoff terms (mean of the states entering "
                                       f"and leaving the breakdown run)"})
    sim.ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", choices=("solar", "transistor"), default="solar")
    ap.add_argument("--de-path", choices=("auto", "pairs", "culled"), default="auto",
                    help="dE path of the stepped simulation (auto: culled when it pays)")
    ap.add_argument("--breakdown-steps", type=int, default=1000)
    ap.add_argument("--de-shape", type=int, default=0, metavar="N",
                    help="fix the main leg's dE kernel shape instead of autotuning it: charge chunks (all "
                         "pairs) or workgroups (culled); for counter runs, whose timings mislead the tuner")
    ap.add_argument("--eager", action="store_true",
                    help="launch the steps kernel by kernel instead of replaying a captured HIP graph "
                         "(for profilers that do not follow graph launches)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-transistor", action="store_true")
    ap.add_argument("--no-map", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-de-path", choices=("auto", "pairs", "culled"), default="auto",
                    help="dE path of the batch's replicas (auto: the batch autotuner decides)")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--surrogate-iters", type=int, default=1200,
                    help="training steps of the surrogate (fewer for profiler runs)")
    ap.add_argument("--surrogate-snapshots", type=int, default=64,
                    help="KMC snapshots for the surrogate (8 of them held out)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (calibrates FETCH_SIZE/WRITE_SIZE "
                         "counters against a transfer of known size)")
    ap.add_argument("--cpu-budget", type=float, default=5.0, help="seconds per CPU baseline sample")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1 "
                                   "(replicas shard with no communication; see the design notes)"}))
        sys.exit(2)

    import langmuir_amd as L
    if args.copy_probe > 0:
        nbytes = args.copy_probe << 20
        a, b = L.DeviceBuffer(nbytes), L.DeviceBuffer(nbytes)
        from langmuir_amd.device import D2D, hip
        assert hip().hipMemcpy(C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(nbytes), D2D) == 0
        a.free()
        b.free()
        log(f"copy probe: {args.copy_probe} MiB device to device")
    workload = solar_workload() if args.workload == "solar" else transistor_workload()
    p = workload[0]
    log(f"workload {args.workload}: {p.grid_x}x{p.grid_y}x{p.grid_z}, {len(workload[1])} e + {len(workload[2])} h, "
        f"cutoff {p.electrostatic_cutoff}")
    main_leg = step_leg(workload, args.steps, args.warmup, args.de_path, max(args.breakdown_steps, 1),
                        graphs=not args.eager, de_shape=args.de_shape)
    k = main_leg["kernel_us"]
    kname = main_leg["roofline"]["kernel"]
    log(f"{main_leg['steps_per_s']:.1f} steps/s ({main_leg['ms_per_step'] * 1e3:.1f} us/step, graph replay, after "
        f"{main_leg['warm_steps']} warm-up steps); breakdown: dE phase {k['de_phase']:.1f} us ({kname} "
        f"{k[kname]:.1f} us), commit {k['commit_phase']:.1f} us; {main_leg['roofline']['achieved']:.2f} TFLOP/s")

    res = {
        "metric": METRIC,
        "value": round(main_leg["steps_per_s"], 3),
        "unit": "KMC steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_leg["ms_per_step"], 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": ("solar BHJ 128x128x32, 5242 e + 5242 h seeded, two-phase sigma=4 morphology "
                         "with 0.5 eV traps, generation 0.524288/step, cutoff 30"
                         if args.workload == "solar" else
                         "transistor 256x256x1, 1310 holes, 1 V bias, cutoff 50"),
            "carriers_start": main_leg["carriers_start"], "carriers_end": main_leg["carriers_end"],
            "de_path": main_leg["de_path"], "de_shape": main_leg["de_shape"],
            "warm_steps_before_timing": main_leg["warm_steps"],
            "parallelism": "single GPU",
        },
        "pair_evals_per_s": main_leg["pair_evals_per_s"],
        "pair_evals_note": "2 N^2 per step, the reference's all-pairs count; the culled path evaluates fewer",
        "kernel_us": main_leg["kernel_us"],
        "roofline": main_leg["roofline"],
    }
    if not args.no_transistor and args.workload == "solar":
        tr = step_leg(transistor_workload(), 3000, 200, "auto", 1000, graphs=not args.eager)
        tr["steps_per_s"] = round(tr["steps_per_s"], 2)
        res["transistor"] = tr
        log(f"transistor: {tr['steps_per_s']:.1f} steps/s, {tr['roofline']['kernel']} "
            f"{tr['kernel_us'][tr['roofline']['kernel']]:.1f} us, {tr['roofline']['achieved']:.2f} TFLOP/s")
    map_samples = None
    if not args.no_map:
        res["map"], map_samples = bench_map()
    if not args.no_batch:
        res["batch"] = bench_batch(de_path=args.batch_de_path, graphs=not args.eager)
    if not args.no_surrogate and args.workload == "solar":
        res["surrogate"] = bench_surrogate(workload, n_snap=max(16, args.surrogate_snapshots),
                                           iters=args.surrogate_iters)
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        cb = cpu_baseline(workload, args.cpu_budget)
        res["cpu_baseline"] = cb
        res["speedup_vs_cpu"] = round(main_leg["steps_per_s"] / cb["value"], 2)
        log(f"cpu baseline: {cb['value']:.3f} steps/s on {cb['cores']} threads ({cb['cpu_model']}), "
            f"runs {cb['runs']}")
        fp64 = {"dE": de_vs_fp64(workload, cb["cores"])}
        if "transistor" in res:
            tcb = cpu_baseline(transistor_workload(), max(1.0, args.cpu_budget / 2))
            res["transistor"]["cpu_baseline"] = tcb
            res["transistor"]["speedup_vs_cpu"] = round(res["transistor"]["steps_per_s"] / tcb["value"], 2)
            fp64["dE_transistor"] = de_vs_fp64(transistor_workload(), cb["cores"])
        if map_samples is not None:
            fp64["map"] = maps_vs_fp64(map_samples, cb["cores"])
        res["error_vs_fp64"] = fp64
        log(f"FP64 comparison: dE rms rel {fp64['dE']['rms_rel']:.2e}, max abs {fp64['dE']['max_abs']:.2e} eV"
            + ("" if map_samples is None else
               f"; map cutoff50 rms rel {fp64['map']['cutoff50']['rms_rel']:.2e}, nocutoff "
               f"{fp64['map']['nocutoff']['rms_rel']:.2e}") + f" ({time.perf_counter() - t0:.0f} s for the CPU leg)")
    print(json.dumps(res))


if __name__ == "__main__":
    main()
