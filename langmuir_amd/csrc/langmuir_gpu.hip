// langmuir_gpu.hip -- MI355X (gfx950) kernels and C-ABI for Langmuir's hot path.
//
// Kernels
//   pair_kernel / pair_reduce   pair sums at an explicit target list
//                               (kernel.cl coulomb2/gauss2, Potential::coulomb*)
//   map_kernel                  potential at every lattice site (coulomb1/gauss1)
//   de_kernel .. inject_kernel  one KMC sweep (simulation.cpp:29-111), see lmg_sim.h
//
// Pair sums: FP32 terms on integer-exact squared distances (0 < r^2 < c^2 is
// decided exactly), charges staged through LDS as float4 (x, y, z, q) tiles
// read as wave-wide broadcasts, plain FP32 accumulation over 32 charges and
// Kahan compensation across those blocks, split over charge chunks for
// occupancy and reduced in FP64 in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "langmuir_gpu.h"
#include "lmg_device.h"
#include "lmg_pairsum.h"
#include "lmg_sim.h"
#include "lmg_mapcull.h"

using namespace lmg;

// Working memory of lmg_pair_potential_noalloc_dev, for callers that cannot allocate:
// the float4 staging of the targets and charges and pair_kernel's per-chunk partials,
// sized for at most kNoAllocTargets targets and kNoAllocCharges charges.  split_count
// keeps S * nt <= (4096 + ceil(nt / 64) - 1) * 64, which bounds the partials.  Each GPU
// holds one copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for that device (g_scratch).
constexpr int kNoAllocTargets = 65536;
constexpr int kNoAllocCharges = 262144;
constexpr size_t kNoAllocPartials = size_t(4096 + kNoAllocTargets / 64) * 64;
constexpr size_t kScratchTgt = 0;
constexpr size_t kScratchChg = kScratchTgt + size_t(kNoAllocTargets) * sizeof(float4);
constexpr size_t kScratchPart = kScratchChg + size_t(kNoAllocCharges) * sizeof(float4);
constexpr size_t kScratchBytes = kScratchPart + kNoAllocPartials * sizeof(float2);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};   // per device: the address of `scratch` there

__global__ void fill_u64_kernel(unsigned long long* p, int n, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ===========================================================================
// Host side
// ===========================================================================
thread_local std::string g_create_error;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct lmg_ctx {
    lmg_params p{};
    Dims d{};
    int dev = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<DevBuf*> owned;
    // scratch for the stateless calls
    DevBuf b_tgt, b_chg, b_part, b_site, b_q, b_tsite, b_out, b_map;
    DevBuf b_mcnt, b_mstart, b_mslot, b_mtmp, b_mtmpi, b_msrc;   // culled map
    // simulation
    bool sim_loaded = false;
    bool sync_mode = false;
    int de_mode = 0;   // lmg_sim_set_de_path
    bool use_graphs = true;             // lmg_sim_set_graphs
    bool autotune = true;               // tune de_kernel at load (a batch tunes its replicas together)
    int forced_chunks = 0;              // lmg_sim_set_de_chunks (0: autotune)
    int tuned_chunks = 0;               // autotuned charge chunks of de_kernel (all-pairs path)
    double tuned_us = -1.0;             // its measured dE-kernel time (culled path: list building + sums)
    int cull_resident = 0;              // cull_de_kernel workgroups resident on the device at once
    int cull_quads_alloc = 0;           // work-item quads the culled path's item array holds
    hipGraphExec_t step_exec = nullptr; // one step, captured at the first lmg_sim_step after a load
    lmg_sim_config cfg{};
    SimDev sim{};
    SimDev* d_sim = nullptr;   // device copy the kernels read
    SimCtl* ctl = nullptr;
    std::vector<void*> sim_allocs;
    std::vector<hipEvent_t> events;
    void* scratch = nullptr;   // this device's copy of `scratch` (lmg_pair_potential_noalloc_dev)
};

namespace {

lmg_status fail(lmg_ctx* c, lmg_status code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail((ctx), LMG_ERR_HIP,                                            \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

lmg_status ensure(lmg_ctx* c, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes) return LMG_OK;
    if (b.p) HIPCHK(c, hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
    size_t want = std::max<size_t>(bytes, 256);
    HIPCHK(c, hipMalloc(&b.p, want));
    b.bytes = want;
    return LMG_OK;
}

lmg_status use_device(lmg_ctx* c) {
    HIPCHK(c, hipSetDevice(c->dev));
    return LMG_OK;
}

// c^2 - 1 as float bits for the 0 < r^2 < c^2 test
lmg_status launch_check(lmg_ctx* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, LMG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LMG_OK;
}

lmg_status check_sites(lmg_ctx* c, const int32_t* s, int n, int limit, const char* what) {
    for (int i = 0; i < n; ++i)
        if (s[i] < 0 || s[i] >= limit)
            return fail(c, LMG_ERR_ARG, std::string(what) + ": site out of range");
    return LMG_OK;
}

// pair sums with the staging and partials in caller-provided device memory (tgt: nt,
// chg: nc float4; part: split_count(nt, nc) * nt float2)
lmg_status pair_launch(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int nc, const int32_t* d_ts, int nt,
                       int flags, double* d_out, float4* tgt, float4* chg, float2* part) {
    const int S = split_count(nt, nc);
    if (nc > 0) hipLaunchKernelGGL(to_float4_kernel, dim3((nc + 255) / 256), dim3(256), 0, c->stream, c->d, d_cs, d_cq, nc, chg);
    hipLaunchKernelGGL(to_float4_kernel, dim3((nt + 255) / 256), dim3(256), 0, c->stream, c->d, d_ts, nullptr, nt, tgt);
    const uint32_t c2m1 = cutoff_bits(c->p.cutoff, flags & LMG_NO_CUTOFF);
    const float gk = gauss_k(c->p.gaussian_sigma);
    dim3 grid((nt + kBlock - 1) / kBlock, S);
    const bool gauss = flags & LMG_GAUSS, image = flags & LMG_IMAGE_LEFT;
    if (gauss && image)
        hipLaunchKernelGGL((pair_kernel<true, true>), grid, dim3(kBlock), 0, c->stream, tgt, nt, chg, nc, c2m1, gk, part, nt);
    else if (gauss)
        hipLaunchKernelGGL((pair_kernel<true, false>), grid, dim3(kBlock), 0, c->stream, tgt, nt, chg, nc, c2m1, gk, part, nt);
    else if (image)
        hipLaunchKernelGGL((pair_kernel<false, true>), grid, dim3(kBlock), 0, c->stream, tgt, nt, chg, nc, c2m1, gk, part, nt);
    else
        hipLaunchKernelGGL((pair_kernel<false, false>), grid, dim3(kBlock), 0, c->stream, tgt, nt, chg, nc, c2m1, gk, part, nt);
    const double scale = (image ? -1.0 : 1.0) * c->p.prefactor;
    hipLaunchKernelGGL(pair_reduce_kernel, dim3((nt + kBlock / kRedLanes - 1) / (kBlock / kRedLanes)), dim3(kBlock), 0,
                       c->stream, part, nt, S, nt, scale, d_out);
    return launch_check(c, "pair_kernel");
}

lmg_status pair_core(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int nc, const int32_t* d_ts,
                     int nt, int flags, double* d_out) {
    lmg_status st;
    if ((st = ensure(c, c->b_tgt, size_t(nt) * sizeof(float4))) != LMG_OK) return st;
    if ((st = ensure(c, c->b_chg, size_t(std::max(nc, 1)) * sizeof(float4))) != LMG_OK) return st;
    const int S = split_count(nt, nc);
    if ((st = ensure(c, c->b_part, size_t(S) * nt * sizeof(float2))) != LMG_OK) return st;
    return pair_launch(c, d_cs, d_cq, nc, d_ts, nt, flags, d_out, static_cast<float4*>(c->b_tgt.p),
                       static_cast<float4*>(c->b_chg.p), static_cast<float2*>(c->b_part.p));
}

// the address of `scratch` on the context's device, looked up once per device with
// that device current
lmg_status scratch_setup(lmg_ctx* c) {
    if (c->dev >= kMaxDevices) return fail(c, LMG_ERR_ARG, "device ordinal above the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[c->dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(c, LMG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[c->dev] = p;
    }
    c->scratch = g_scratch[c->dev];
    return LMG_OK;
}

// the culled map pays when the cutoff box of a 64-site tile covers well under
// the lattice and there is enough work to matter
bool map_culled(const lmg_ctx* c, int nc, int flags) {
    if ((flags & LMG_NO_CUTOFF) || nc == 0) return false;
    if (flags & LMG_MAP_CULLED) return true;
    if (flags & LMG_MAP_PAIRS) return false;
    return map_cull_pays(c->d, c->p.cutoff, nc);
}

lmg_status map_cull_core(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int nc, int flags, float* d_out,
                         const int32_t* d_run) {
    lmg_status st;
    MapBricks m = map_bricks_shape(c->d);
    const size_t cap = std::max<size_t>(nc, 1);
    void* before = c->b_mcnt.p;
    if ((st = ensure(c, c->b_mcnt, size_t(m.nb) * 4)) || (st = ensure(c, c->b_mstart, size_t(m.nb + 1) * 4)) ||
        (st = ensure(c, c->b_mslot, size_t(nc) * sizeof(int2))) || (st = ensure(c, c->b_mtmp, cap * sizeof(float4))) ||
        (st = ensure(c, c->b_mtmpi, cap * 4)) || (st = ensure(c, c->b_msrc, cap * sizeof(float4))))
        return st;
    if (c->b_mcnt.p != before) HIPCHK(c, hipMemsetAsync(c->b_mcnt.p, 0, size_t(m.nb) * 4, c->stream));
    m.cnt = static_cast<int*>(c->b_mcnt.p);
    m.start = static_cast<int*>(c->b_mstart.p);
    m.slot = static_cast<int2*>(c->b_mslot.p);
    m.tmp = static_cast<float4*>(c->b_mtmp.p);
    m.tmpi = static_cast<int*>(c->b_mtmpi.p);
    m.src = static_cast<float4*>(c->b_msrc.p);
    enqueue_map_cull(c->stream, c->d, m, c->p.cutoff, gauss_k(c->p.gaussian_sigma), float(c->p.prefactor),
                     flags & LMG_GAUSS, d_cs, d_cq, nc, 1, MapBatch{nullptr, d_run, nullptr}, d_out);
    return launch_check(c, "map_cull_kernel");
}

lmg_status map_core(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int nc, int flags, float* d_out,
                    const int32_t* d_run = nullptr) {
    if (map_culled(c, nc, flags)) return map_cull_core(c, d_cs, d_cq, nc, flags, d_out, d_run);
    lmg_status st;
    if ((st = ensure(c, c->b_chg, size_t(std::max(nc, 1)) * sizeof(float4))) != LMG_OK) return st;
    float4* chg = static_cast<float4*>(c->b_chg.p);
    if (nc > 0)
        hipLaunchKernelGGL(to_float4_kernel, dim3((nc + 255) / 256), dim3(256), 0, c->stream, c->d, d_cs, d_cq, nc,
                           chg, d_run);
    int tx = 1;
    while (tx < 64 && tx * kMapT < c->d.X) tx <<= 1;
    const int ty = kBlock / tx;
    dim3 grid((c->d.X + tx * kMapT - 1) / (tx * kMapT), (c->d.Y + ty - 1) / ty, c->d.Z);
    const uint32_t c2m1 = cutoff_bits(c->p.cutoff, flags & LMG_NO_CUTOFF);
    const float gk = gauss_k(c->p.gaussian_sigma);
    const float pref = float(c->p.prefactor);
    if (flags & LMG_GAUSS)
        hipLaunchKernelGGL(map_kernel<true>, grid, dim3(kBlock), 0, c->stream, c->d, chg, nc, c2m1, gk, pref, tx,
                           d_run, d_out);
    else
        hipLaunchKernelGGL(map_kernel<false>, grid, dim3(kBlock), 0, c->stream, c->d, chg, nc, c2m1, gk, pref, tx,
                           d_run, d_out);
    return launch_check(c, "map_kernel");
}

void sim_free(lmg_ctx* c) {
    if (c->step_exec) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipGraphExecDestroy(c->step_exec);
        c->step_exec = nullptr;
    }
    for (void* p : c->sim_allocs) (void)hipFree(p);
    c->sim_allocs.clear();
    c->sim_loaded = false;
    c->ctl = nullptr;
    c->d_sim = nullptr;
}

template <class T>
lmg_status sim_alloc(lmg_ctx* c, T*& p, size_t n) {
    void* q = nullptr;
    HIPCHK(c, hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    c->sim_allocs.push_back(q);
    p = static_cast<T*>(q);
    return LMG_OK;
}

unsigned long long next_pow2(unsigned long long v) {
    unsigned long long p = 1;
    while (p < v) p <<= 1;
    return p;
}

// The all-pairs dE kernel over a descriptor array (grid z: its length): the
// table form when every descriptor has a table, else f evaluated in the kernel
void launch_de(dim3 grid, hipStream_t st, const SimDev* P, bool coulomb, bool table, bool gauss) {
    const dim3 blk(kSimBlock);
    if (!coulomb)
        hipLaunchKernelGGL((de_kernel<false, false, false>), grid, blk, kDeLdsBytes, st, P);
    else if (table)
        hipLaunchKernelGGL((de_kernel<false, true, true>), grid, blk, kDeLdsBytes, st, P);
    else if (gauss)
        hipLaunchKernelGGL((de_kernel<true, true, false>), grid, blk, kDeLdsBytes, st, P);
    else
        hipLaunchKernelGGL((de_kernel<false, true, false>), grid, blk, kDeLdsBytes, st, P);
}

// The culled dE kernel over a descriptor array: the table form (dynamic LDS:
// the largest table) when every culled descriptor has a table, else f is
// evaluated in the kernel (1/r, or the erf-screened form).
void launch_cull_de(dim3 grid, hipStream_t st, const SimDev* P, const int2* map, bool table, bool gauss,
                    int max_rtab_n) {
    const dim3 blk(kSimBlock);
    if (table)
        hipLaunchKernelGGL((cull_de_kernel<false, true>), grid, blk, cull_lds_bytes(max_rtab_n), st, P, map);
    else if (gauss)
        hipLaunchKernelGGL((cull_de_kernel<true, false>), grid, blk, cull_lds_bytes(0), st, P, map);
    else
        hipLaunchKernelGGL((cull_de_kernel<false, false>), grid, blk, cull_lds_bytes(0), st, P, map);
}

// dE pair sums + decision flags + random-stream refill (always launched: with
// Coulomb off only the flags and the refill run).  `mid`, when given, is
// recorded before the pair-sum kernel itself (after the culled path's list
// building).
lmg_status sim_launch_de(lmg_ctx* c, hipEvent_t mid = nullptr) {
    const SimDev& D = c->sim;
    const SimDev* P = c->d_sim;
    if (D.cull) {
        const int2* none = nullptr;   // one replica: no workgroup map
        hipLaunchKernelGGL(cull_emit_kernel, dim3((D.nbrick + 3) / 4 + D.nblk + 1), dim3(kSimBlock), 0, c->stream,
                           P, none);
        if (mid) HIPCHK(c, hipEventRecord(mid, c->stream));
        // work items, the refill workgroup, then the decisions' inputs (decide_prep)
        const dim3 grid((D.items_cap + 3) / 4 + 1 + D.nblk);
        launch_cull_de(grid, c->stream, P, none, D.rtab_n > 0, D.gauss != 0, D.rtab_n);
        return launch_check(c, "cull_de_kernel");
    }
    if (mid) HIPCHK(c, hipEventRecord(mid, c->stream));
    launch_de(dim3(D.nblk + 1, D.coulomb ? D.S : 1), c->stream, P, D.coulomb != 0, D.rtab_n > 0, D.gauss != 0);
    return launch_check(c, "de_kernel");
}

// decisions, recombination, completeTick/compaction, injection and the next
// step's set-up (the proposals themselves are made by the next dE phase)
lmg_status sim_launch_commit(lmg_ctx* c) {
    const SimDev& D = c->sim;
    const SimDev* P = c->d_sim;
    hipLaunchKernelGGL(decide_kernel, dim3(D.nblk), dim3(kSimBlock), 0, c->stream, P, 0);
    if (D.sim_type == 1 && D.rate[LMG_FLUX_XD] > 0)
        hipLaunchKernelGGL(recomb_kernel, dim3(1), dim3(kOneBlock), 0, c->stream, P);
    hipLaunchKernelGGL(tick_kernel, dim3(D.nblk), dim3(kSimBlock), 0, c->stream, P);
    hipLaunchKernelGGL(inject_kernel, dim3(1), dim3(kOneBlock), 0, c->stream, P, 0);
    return launch_check(c, "step kernels");
}

// ---------------------------------------------------------------------------
// Autotuner of the all-pairs dE kernel's shape: the number of charge chunks S
// (workgroups along y per block of 256 targets).  Each candidate is timed on
// the loaded state (one warm-up, then three launches between HIP events).  A
// dE launch only rewrites the partials, the decision flags and the refilled
// part of the random ring, which the next step recomputes, and its sums are
// exact fixed-point totals of the same 32-charge blocks for every S, so the
// choice changes no result.
// ---------------------------------------------------------------------------
constexpr int kMaxChunks = 128;

// mean time of `reps` launches after one untimed launch
template <class F>
double time_launch_us(hipStream_t st, F&& launch, int reps = 3) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1e30;
    launch();
    (void)hipEventRecord(a, st);
    for (int k = 0; k < reps; ++k) launch();
    (void)hipEventRecord(b, st);
    float ms = 0.f;
    const bool ok = hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess &&
                    hipGetLastError() == hipSuccess;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ok ? 1e3 * double(ms) / double(reps) : 1e30;
}

// Autotuning of the culled path's work-item count: cull_de_kernel workgroups
// (four chunks each) in multiples of what the device holds at once, k
// resident waves of workgroups for k = 1 .. kCullTuneWaves; more workgroups
// make smaller chunks, which balance better and cost more list scans.
constexpr int kCullTuneWaves = 3;
int cull_quads(int k, int resident, int ngrp) { return std::max(k * resident - 1, ngrp); }

// candidates: powers of two up to kMaxChunks with at least one 32-charge block each
std::vector<int> chunk_candidates(int n_charges) {
    std::vector<int> v;
    const int blocks = std::max(1, (n_charges + 31) / 32);
    for (int S = 1; S <= kMaxChunks && S <= blocks; S *= 2) v.push_back(S);
    return v;
}

// The culled path's work-item count: candidates timed twice each (the second
// round sees the clocks the first one raised), the fastest kept.  The list
// building and the sums only rewrite per-step scratch, and the totals are
// exact fixed-point sums of the same units for any item count.
lmg_status sim_tune_cull(lmg_ctx* c) {
    SimDev& D = c->sim;
    c->tuned_us = -1.0;
    if (c->forced_chunks > 0 || !c->autotune) return LMG_OK;
    double best = 1e30;
    int best_q = D.items_cap / 4;
    for (int round = 0; round < 2; ++round)
        for (int k = 1; k <= kCullTuneWaves; ++k) {
            D.items_cap = 4 * cull_quads(k, c->cull_resident, D.ngrp);
            HIPCHK(c, hipMemcpy(c->d_sim, &D, sizeof(SimDev), hipMemcpyHostToDevice));
            const double us = time_launch_us(c->stream, [&] { (void)sim_launch_de(c); }, 5);
            if (us < best) {
                best = us;
                best_q = D.items_cap / 4;
            }
        }
    D.items_cap = 4 * best_q;
    HIPCHK(c, hipMemcpy(c->d_sim, &D, sizeof(SimDev), hipMemcpyHostToDevice));
    c->tuned_us = best;
    return launch_check(c, "culled dE autotune");
}

lmg_status sim_tune_chunks(lmg_ctx* c) {
    SimDev& D = c->sim;
    if (D.cull) return sim_tune_cull(c);
    if (!D.coulomb) return LMG_OK;
    if (c->forced_chunks > 0) {
        D.S = std::min(c->forced_chunks, kMaxChunks);
        HIPCHK(c, hipMemcpy(c->d_sim, &D, sizeof(SimDev), hipMemcpyHostToDevice));
        c->tuned_chunks = D.S;
        c->tuned_us = -1.0;
        return LMG_OK;
    }
    if (!c->autotune) return LMG_OK;
    const int cap = D.cap_e + D.cap_h;
    double best = 1e30;
    int best_S = D.S;
    for (int S : chunk_candidates(cap + D.n_d_eff)) {
        D.S = S;
        HIPCHK(c, hipMemcpy(c->d_sim, &D, sizeof(SimDev), hipMemcpyHostToDevice));
        const double us = time_launch_us(c->stream, [&] { (void)sim_launch_de(c); });
        if (us < best) {
            best = us;
            best_S = S;
        }
    }
    D.S = best_S;
    HIPCHK(c, hipMemcpy(c->d_sim, &D, sizeof(SimDev), hipMemcpyHostToDevice));
    c->tuned_chunks = best_S;
    c->tuned_us = best;
    return launch_check(c, "dE autotune");
}

// One KMC step as a HIP graph: every kernel reads the device descriptor, so
// the captured launches stay valid until the next load (sim_free drops it).
// The step records kernels only (no memset or copy nodes).
template <class F>
lmg_status capture_step(hipStream_t st, hipGraphExec_t* exec, F&& enqueue, std::string& err) {
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
        const lmg_status s = enqueue();
        e = hipStreamEndCapture(st, &graph);
        if (s != LMG_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return s;
        }
    }
    if (e == hipSuccess) {
        e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
    }
    if (e != hipSuccess) {
        *exec = nullptr;
        err = std::string("step graph: ") + hipGetErrorString(e);
        return LMG_ERR_HIP;
    }
    return LMG_OK;
}

lmg_status sim_launch_step(lmg_ctx* c) {
    lmg_status st;
    if ((st = sim_launch_de(c)) != LMG_OK) return st;
    return sim_launch_commit(c);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* lmg_last_error(const lmg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int32_t lmg_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

lmg_status lmg_create(const lmg_params* params, lmg_ctx** out) {
    if (!params || !out) return fail(nullptr, LMG_ERR_ARG, "null argument");
    *out = nullptr;
    const lmg_params& p = *params;
    if (p.grid_x < 1 || p.grid_y < 1 || p.grid_z < 1)
        return fail(nullptr, LMG_ERR_ARG, "grid sizes must be >= 1");
    if (int64_t(p.grid_x) * p.grid_y * p.grid_z + kSpecialSlots >= (int64_t(1) << 31))
        return fail(nullptr, LMG_ERR_ARG, "lattice too large for int32 site ids");
    if (p.grid_x > 16384 || p.grid_y > 16384 || p.grid_z > 16384)
        return fail(nullptr, LMG_ERR_ARG, "grid dimension above 16384 (float-exact distances)");
    if (p.cutoff < 1 || p.cutoff > 16384) return fail(nullptr, LMG_ERR_ARG, "cutoff must be in [1, 16384]");
    if (p.hopping_range < 1 || p.hopping_range > 2)
        return fail(nullptr, LMG_ERR_ARG, "hopping_range must be 1 or 2");
    // the dE kernels' fixed-point block sums are exact for |q| <= 31 e
    if (p.defects_charge < -31 || p.defects_charge > 31)
        return fail(nullptr, LMG_ERR_ARG, "defects_charge must be in [-31, 31]");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(nullptr, LMG_ERR_NO_DEVICE, "no HIP device available");
    if (p.device_id < 0 || p.device_id >= n) return fail(nullptr, LMG_ERR_ARG, "device_id out of range");
    lmg_ctx* c = new lmg_ctx;
    c->p = p;
    c->d = make_dims(p.grid_x, p.grid_y, p.grid_z);
    c->dev = p.device_id;
    hipError_t e = hipSetDevice(c->dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    auto drop = [c]() {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
    };
    if (e != hipSuccess) {
        g_create_error = std::string("HIP initialisation failed: ") + hipGetErrorString(e);
        drop();
        return LMG_ERR_HIP;
    }
    if (scratch_setup(c) != LMG_OK) {
        g_create_error = c->err;
        drop();
        return LMG_ERR_HIP;
    }
    *out = c;
    return LMG_OK;
}

void lmg_destroy(lmg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->dev);
    (void)hipStreamSynchronize(c->stream);
    sim_free(c);
    for (DevBuf* b : {&c->b_tgt, &c->b_chg, &c->b_part, &c->b_site, &c->b_q, &c->b_tsite, &c->b_out, &c->b_map,
                      &c->b_mcnt, &c->b_mstart, &c->b_mslot, &c->b_mtmp, &c->b_mtmpi, &c->b_msrc})
        if (b->p) (void)hipFree(b->p);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

void* lmg_stream(lmg_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

lmg_status lmg_get_params(const lmg_ctx* c, lmg_params* out) {
    if (!c || !out) return LMG_ERR_ARG;
    *out = c->p;
    return LMG_OK;
}

lmg_status lmg_sync(lmg_ctx* c) {
    if (!c) return LMG_ERR_ARG;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LMG_OK;
}

lmg_status lmg_pair_potential_dev(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int32_t nc,
                                  const int32_t* d_ts, int32_t nt, int32_t flags, double* d_out) {
    if (!c) return LMG_ERR_ARG;
    if (nc < 0 || nt < 0 || (nc > 0 && (!d_cs || !d_cq)) || (nt > 0 && (!d_ts || !d_out)))
        return fail(c, LMG_ERR_ARG, "invalid pair_potential arguments");
    if (flags & ~(LMG_GAUSS | LMG_IMAGE_LEFT | LMG_NO_CUTOFF)) return fail(c, LMG_ERR_ARG, "unsupported flags");
    if ((flags & LMG_GAUSS) && !(c->p.gaussian_sigma > 0))
        return fail(c, LMG_ERR_ARG, "LMG_GAUSS needs gaussian_sigma > 0");
    if (nt == 0) return LMG_OK;
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    return pair_core(c, d_cs, d_cq, nc, d_ts, nt, flags, d_out);
}

lmg_status lmg_pair_potential_noalloc_dev(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int32_t nc,
                                          const int32_t* d_ts, int32_t nt, int32_t flags, double* d_out) {
    if (!c) return LMG_ERR_ARG;
    if (nc < 0 || nt < 0 || (nc > 0 && (!d_cs || !d_cq)) || (nt > 0 && (!d_ts || !d_out)))
        return fail(c, LMG_ERR_ARG, "invalid pair_potential arguments");
    if (flags & ~(LMG_GAUSS | LMG_IMAGE_LEFT | LMG_NO_CUTOFF)) return fail(c, LMG_ERR_ARG, "unsupported flags");
    if ((flags & LMG_GAUSS) && !(c->p.gaussian_sigma > 0))
        return fail(c, LMG_ERR_ARG, "LMG_GAUSS needs gaussian_sigma > 0");
    if (nt > kNoAllocTargets || nc > kNoAllocCharges)
        return fail(c, LMG_ERR_CAPACITY, "lmg_pair_potential_noalloc_dev takes at most " +
                                             std::to_string(kNoAllocTargets) + " targets and " +
                                             std::to_string(kNoAllocCharges) + " charges");
    if (nt == 0) return LMG_OK;
    if (size_t(split_count(nt, nc)) * size_t(nt) > kNoAllocPartials)
        return fail(c, LMG_ERR_CAPACITY, "partials exceed the scratch array");
    if (!c->scratch) return fail(c, LMG_ERR_STATE, "no scratch address for this device");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    unsigned char* base = static_cast<unsigned char*>(c->scratch);
    return pair_launch(c, d_cs, d_cq, nc, d_ts, nt, flags, d_out, reinterpret_cast<float4*>(base + kScratchTgt),
                       reinterpret_cast<float4*>(base + kScratchChg), reinterpret_cast<float2*>(base + kScratchPart));
}

lmg_status lmg_pair_potential(lmg_ctx* c, const int32_t* cs, const int32_t* cq, int32_t nc,
                              const int32_t* ts, int32_t nt, int32_t flags, double* out) {
    if (!c) return LMG_ERR_ARG;
    if (nc < 0 || nt < 0 || (nc > 0 && (!cs || !cq)) || (nt > 0 && (!ts || !out)))
        return fail(c, LMG_ERR_ARG, "invalid pair_potential arguments");
    if (nt == 0) return LMG_OK;
    lmg_status st;
    // charges are lattice sites; targets may also be special slots (drain
    // proposals are evaluated at V+k exactly like the reference)
    if ((st = check_sites(c, cs, nc, c->d.V, "charge")) != LMG_OK) return st;
    if ((st = check_sites(c, ts, nt, c->d.V + kSpecialSlots, "target")) != LMG_OK) return st;
    if ((st = use_device(c)) != LMG_OK) return st;
    if ((st = ensure(c, c->b_site, size_t(std::max(nc, 1)) * 4)) != LMG_OK) return st;
    if ((st = ensure(c, c->b_q, size_t(std::max(nc, 1)) * 4)) != LMG_OK) return st;
    if ((st = ensure(c, c->b_tsite, size_t(nt) * 4)) != LMG_OK) return st;
    if ((st = ensure(c, c->b_out, size_t(nt) * 8)) != LMG_OK) return st;
    if (nc > 0) {
        HIPCHK(c, hipMemcpyAsync(c->b_site.p, cs, size_t(nc) * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->b_q.p, cq, size_t(nc) * 4, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->b_tsite.p, ts, size_t(nt) * 4, hipMemcpyHostToDevice, c->stream));
    st = lmg_pair_potential_dev(c, static_cast<int32_t*>(c->b_site.p), static_cast<int32_t*>(c->b_q.p), nc,
                                static_cast<int32_t*>(c->b_tsite.p), nt, flags, static_cast<double*>(c->b_out.p));
    if (st != LMG_OK) return st;
    HIPCHK(c, hipMemcpyAsync(out, c->b_out.p, size_t(nt) * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LMG_OK;
}

lmg_status lmg_potential_map_dev(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int32_t nc,
                                 int32_t flags, float* d_out) {
    if (!c) return LMG_ERR_ARG;
    if (nc < 0 || (nc > 0 && (!d_cs || !d_cq)) || !d_out) return fail(c, LMG_ERR_ARG, "invalid map arguments");
    if (flags & ~(LMG_GAUSS | LMG_NO_CUTOFF | LMG_MAP_PAIRS | LMG_MAP_CULLED)) return fail(c, LMG_ERR_ARG, "unsupported flags");
    if ((flags & LMG_GAUSS) && !(c->p.gaussian_sigma > 0))
        return fail(c, LMG_ERR_ARG, "LMG_GAUSS needs gaussian_sigma > 0");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    return map_core(c, d_cs, d_cq, nc, flags, d_out);
}

lmg_status lmg_potential_map_if_dev(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int32_t nc,
                                    int32_t flags, const int32_t* d_run, float* d_out) {
    if (!c) return LMG_ERR_ARG;
    if (nc < 0 || (nc > 0 && (!d_cs || !d_cq)) || !d_out || !d_run)
        return fail(c, LMG_ERR_ARG, "invalid map arguments");
    if (flags & ~(LMG_GAUSS | LMG_NO_CUTOFF | LMG_MAP_PAIRS | LMG_MAP_CULLED)) return fail(c, LMG_ERR_ARG, "unsupported flags");
    if ((flags & LMG_GAUSS) && !(c->p.gaussian_sigma > 0))
        return fail(c, LMG_ERR_ARG, "LMG_GAUSS needs gaussian_sigma > 0");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    return map_core(c, d_cs, d_cq, nc, flags, d_out, d_run);
}

lmg_status lmg_potential_map(lmg_ctx* c, const int32_t* cs, const int32_t* cq, int32_t nc, int32_t flags,
                             float* out) {
    if (!c) return LMG_ERR_ARG;
    if (nc < 0 || (nc > 0 && (!cs || !cq)) || !out) return fail(c, LMG_ERR_ARG, "invalid map arguments");
    lmg_status st;
    if ((st = check_sites(c, cs, nc, c->d.V, "charge")) != LMG_OK) return st;
    if ((st = use_device(c)) != LMG_OK) return st;
    if ((st = ensure(c, c->b_site, size_t(std::max(nc, 1)) * 4)) != LMG_OK) return st;
    if ((st = ensure(c, c->b_q, size_t(std::max(nc, 1)) * 4)) != LMG_OK) return st;
    if ((st = ensure(c, c->b_map, size_t(c->d.V) * 4)) != LMG_OK) return st;
    if (nc > 0) {
        HIPCHK(c, hipMemcpyAsync(c->b_site.p, cs, size_t(nc) * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->b_q.p, cq, size_t(nc) * 4, hipMemcpyHostToDevice, c->stream));
    }
    st = lmg_potential_map_dev(c, static_cast<int32_t*>(c->b_site.p), static_cast<int32_t*>(c->b_q.p), nc, flags,
                               static_cast<float*>(c->b_map.p));
    if (st != LMG_OK) return st;
    HIPCHK(c, hipMemcpyAsync(out, c->b_map.p, size_t(c->d.V) * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LMG_OK;
}

lmg_status lmg_sim_load(lmg_ctx* c, const lmg_sim_config* cfg, const lmg_sim_state* s) {
    if (!c) return LMG_ERR_ARG;
    if (!cfg || !s) return fail(c, LMG_ERR_ARG, "null config/state");
    const Dims d = c->d;
    const int Vt = d.V + kSpecialSlots;
    if (cfg->simulation_type != 0 && cfg->simulation_type != 1) return fail(c, LMG_ERR_ARG, "simulation_type");
    if (s->n_electrons < 0 || s->n_holes < 0 || s->n_defects < 0) return fail(c, LMG_ERR_ARG, "negative count");
    if (s->n_electrons > std::max(cfg->max_electrons, 0) || s->n_holes > std::max(cfg->max_holes, 0))
        return fail(c, LMG_ERR_CAPACITY, "carrier count above max_electrons/max_holes");
    for (int k = 0; k < 2; ++k) {
        if (cfg->e_drain_slot[k] >= 0 && (cfg->e_drain_slot[k] < d.V || cfg->e_drain_slot[k] >= Vt))
            return fail(c, LMG_ERR_ARG, "electron drain slot must be in [V, V+35)");
        if (cfg->h_drain_slot[k] >= 0 && (cfg->h_drain_slot[k] < d.V || cfg->h_drain_slot[k] >= Vt))
            return fail(c, LMG_ERR_ARG, "hole drain slot must be in [V, V+35)");
    }
    if (cfg->recombination_range < 0 || cfg->recombination_range > 2)
        return fail(c, LMG_ERR_ARG, "recombination_range must be 0, 1 or 2");
    lmg_status st;
    if ((st = check_sites(c, s->e_site, s->n_electrons, d.V, "electron")) != LMG_OK) return st;
    if ((st = check_sites(c, s->h_site, s->n_holes, d.V, "hole")) != LMG_OK) return st;
    if ((st = check_sites(c, s->defect_site, s->n_defects, d.V, "defect")) != LMG_OK) return st;
    // types (Grid::registerAgent / registerDefect / registerSpecialAgent)
    std::vector<uint8_t> et(Vt, kEmpty), ht(Vt, kEmpty);
    for (int i = 0; i < s->n_defects; ++i) {
        const int q = s->defect_site[i];
        if (et[q] != kEmpty) return fail(c, LMG_ERR_ARG, "duplicate defect site");
        et[q] = kDefect; ht[q] = kDefect;
    }
    for (int i = 0; i < s->n_electrons; ++i) {
        const int q = s->e_site[i];
        if (et[q] != kEmpty) return fail(c, LMG_ERR_ARG, "electron on an occupied site");
        et[q] = kElectron;
    }
    for (int i = 0; i < s->n_holes; ++i) {
        const int q = s->h_site[i];
        if (ht[q] != kEmpty) return fail(c, LMG_ERR_ARG, "hole on an occupied site");
        ht[q] = kHole;
    }
    for (int k = 0; k < 2; ++k) {
        if (cfg->e_drain_slot[k] >= 0) et[cfg->e_drain_slot[k]] = kDrain;
        if (cfg->h_drain_slot[k] >= 0) ht[cfg->h_drain_slot[k]] = kDrain;
    }
    std::vector<double> vbg(Vt, 0.0);
    if (s->potential) std::copy(s->potential, s->potential + d.V, vbg.begin());

    if ((st = use_device(c)) != LMG_OK) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    sim_free(c);
    c->cfg = *cfg;
    SimDev& D = c->sim;
    D = SimDev{};
    D.d = d;
    D.hop = c->p.hopping_range;
    D.sim_type = cfg->simulation_type;
    D.coulomb = cfg->coulomb_carriers != 0;
    D.src_metro = cfg->source_metropolis != 0;
    D.src_coul = cfg->source_coulomb != 0;
    D.balance = cfg->balance_charges != 0;
    D.rec_range = cfg->recombination_range;
    D.gauss = c->p.gaussian_sigma > 0;
    D.max_e = cfg->max_electrons;
    D.max_h = cfg->max_holes;
    D.cap_e = std::max(1, std::max(cfg->max_electrons, s->n_electrons));
    D.cap_h = std::max(1, std::max(cfg->max_holes, s->n_holes));
    D.dq = c->p.defects_charge;
    D.n_d_eff = D.dq != 0 ? s->n_defects : 0;
    for (int k = 0; k < 2; ++k) { D.e_drain[k] = cfg->e_drain_slot[k]; D.h_drain[k] = cfg->h_drain_slot[k]; }
    D.c2m1 = cutoff_bits(c->p.cutoff, false);
    D.gk = gauss_k(c->p.gaussian_sigma);
    const double gk_d = c->p.gaussian_sigma > 0 ? 1.0 / (std::sqrt(2.0) * c->p.gaussian_sigma) : 0.0;
    D.vl = cfg->voltage_left;
    D.vr = cfg->voltage_right;
    D.beta = c->p.inverse_kt;
    D.pref = c->p.prefactor;
    // sI[1][0][0] = pref * iR[1][0][0] * eR[1][0][0] (potential.cpp:316-324)
    D.self_int = c->p.prefactor * 1.0 * (c->p.gaussian_sigma > 0 ? std::erf(gk_d * 1.0) : 1.0);
    D.binding = c->p.exciton_binding;
    std::memcpy(D.rate, cfg->rate, sizeof(D.rate));
    std::memcpy(D.coupling, cfg->coupling, sizeof(D.coupling));
    const int cap = D.cap_e + D.cap_h;
    D.S_pairs = split_count(cap, cap + D.n_d_eff);
    D.S = D.S_pairs;
    D.ld = cap;
    D.lookahead = 4ull * cap + 8192ull;
    const unsigned long long ring = next_pow2(2ull * (cap + D.lookahead + 4096ull));
    D.ring_mask = ring - 1ull;
    D.nblk = (cap + kSimBlock - 1) / kSimBlock;
    // drains on both x faces give every site at least two neighbours
    D.uniform = (D.e_drain[0] >= 0 && D.e_drain[1] >= 0 && D.h_drain[0] >= 0 && D.h_drain[1] >= 0) ? 1 : 0;
    // culled dE when the cutoff box around a group of targets covers well under
    // the whole lattice (estimated from a 16^3 group plus the cutoff)
    D.cull = 0;
    D.nbx = (d.X + kBrick - 1) / kBrick;
    D.nby = (d.Y + kBrick - 1) / kBrick;
    D.nbz = (d.Z + kBrick - 1) / kBrick;
    D.nbrick = D.nbx * D.nby * D.nbz;
    // (the culled kernel addresses list slots in 24 bits)
    if (D.coulomb && D.nbrick <= kMaxCullBricks &&
        size_t(cap) + size_t(D.n_d_eff) + 3 * size_t(D.nbrick) < (size_t(1) << 24)) {
        auto frac = [&](int n) { return std::min(1.0, double(16 + 2 * c->p.cutoff) / n); };
        const bool worth = frac(d.X) * frac(d.Y) * frac(d.Z) < 0.6 && double(cap) * (cap + D.n_d_eff) >= 8e6;
        D.cull = c->de_mode == 2 || (c->de_mode == 0 && worth);
    }
    // the pair-function table (LDS of the dE kernels) for cutoffs up to 63
    {
        const long long c2 = (long long)c->p.cutoff * c->p.cutoff;
        D.rtab_n = D.coulomb && c2 + 1 <= kTabMax ? int(c2 + 1) : 0;
        D.rtab_lim = D.rtab_n > 0 ? uint32_t(4 * c2) : 0u;
    }
    if (D.cull) {
        D.cut = c->p.cutoff;
        D.ngrp = (cap + kCullGroup - 1) / kCullGroup;
        int ncu = 0, per_cu = 0;
        HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->dev));
        HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                      &per_cu,
                      D.rtab_n > 0 ? reinterpret_cast<const void*>(&cull_de_kernel<false, true>)
                                   : reinterpret_cast<const void*>(&cull_de_kernel<false, false>),
                      kSimBlock, cull_lds_bytes(D.rtab_n)));
        // work items (workgroups of four chunks of one target group): to start
        // with, as many as keep every workgroup of cull_de_kernel resident at
        // once (one wave of workgroups, no tail), at least one per group;
        // items_cap counts waves.  The autotuner tries up to kCullTuneWaves
        // waves of workgroups; lmg_sim_set_de_chunks fixes the count instead.
        c->cull_resident = std::max(1, ncu * per_cu);
        const int wgs = c->forced_chunks > 0 ? std::min(c->forced_chunks, kMaxCullWorkgroups)
                                             : cull_quads(1, c->cull_resident, D.ngrp);
        D.items_cap = 4 * std::max(wgs, D.ngrp);
        c->cull_quads_alloc = std::max(D.items_cap / 4, cull_quads(kCullTuneWaves, c->cull_resident, D.ngrp));
    }

    for (int b = 0; b < 2; ++b) {
        if ((st = sim_alloc(c, D.e_site[b], D.cap_e)) || (st = sim_alloc(c, D.e_life[b], D.cap_e)) ||
            (st = sim_alloc(c, D.e_path[b], D.cap_e)) || (st = sim_alloc(c, D.h_site[b], D.cap_h)) ||
            (st = sim_alloc(c, D.h_life[b], D.cap_h)) || (st = sim_alloc(c, D.h_path[b], D.cap_h)))
            return st;
    }
    if ((st = sim_alloc(c, D.e_fsite, D.cap_e)) || (st = sim_alloc(c, D.h_fsite, D.cap_h)) ||
        (st = sim_alloc(c, D.etype, Vt)) || (st = sim_alloc(c, D.htype, Vt)) || (st = sim_alloc(c, D.vbg, Vt)) ||
        (st = sim_alloc(c, D.claim_e, d.V)) || (st = sim_alloc(c, D.claim_h, d.V)) ||
        (st = sim_alloc(c, D.h_at, d.V)) || (st = sim_alloc(c, D.code, cap)) ||
        (st = sim_alloc(c, D.removed, cap)) || (st = sim_alloc(c, D.flag, cap)) ||
        (st = sim_alloc(c, D.tgt1, cap)) || (st = sim_alloc(c, D.tgt2, cap)) ||
        (st = sim_alloc(c, D.chg, D.n_d_eff)) ||
        (st = sim_alloc(c, D.part, size_t(std::max(kMaxChunks, D.S_pairs)) * D.ld)) ||
        (st = sim_alloc(c, D.blk_draw, D.nblk)) || (st = sim_alloc(c, D.blk_keep, 2 * D.nblk)) ||
        (st = sim_alloc(c, D.de, cap)) || (st = sim_alloc(c, D.ring, ring)) || (st = sim_alloc(c, c->ctl, 1)) ||
        (st = sim_alloc(c, c->d_sim, 1)))
        return st;
    if (D.cull) {
        const int nb = D.nbrick;
        int *bperm = nullptr, *brank = nullptr, *dstart = nullptr;
        float4* dsrc = nullptr;
        if ((st = sim_alloc(c, D.bmask, size_t(nb) * 32)) || (st = sim_alloc(c, D.scnt, nb)) ||
            (st = sim_alloc(c, D.sstart, nb + 1)) || (st = sim_alloc(c, D.tstart, nb)) ||
            (st = sim_alloc(c, D.e_who, d.V)) || (st = sim_alloc(c, D.h_who, d.V)) ||
            (st = sim_alloc(c, bperm, nb)) || (st = sim_alloc(c, brank, nb)) || (st = sim_alloc(c, dstart, nb + 1)) ||
            (st = sim_alloc(c, dsrc, D.n_d_eff)) || (st = sim_alloc(c, D.src, size_t(cap) + D.n_d_eff + 3 * size_t(nb))) ||
            (st = sim_alloc(c, D.torder, cap)) || (st = sim_alloc(c, D.tgrp, D.ngrp)) ||
            (st = sim_alloc(c, D.items, 4 * size_t(c->cull_quads_alloc))) || (st = sim_alloc(c, D.nitems, 1)) ||
            (st = sim_alloc(c, D.isum, cap)) || (st = sim_alloc(c, D.drec, cap)))
            return st;
        HIPCHK(c, hipMemset(D.bmask, 0, size_t(nb) * 32 * 4));
        HIPCHK(c, hipMemset(D.scnt, 0, size_t(nb) * 4));
        // a target slot that no carrier fills (two of one species on a site, a
        // state the device flags) keeps a valid carrier index
        HIPCHK(c, hipMemset(D.torder, 0, size_t(cap) * 4));
        HIPCHK(c, hipMemset(D.e_who, 0, size_t(d.V) * 4));
        HIPCHK(c, hipMemset(D.h_who, 0, size_t(d.V) * 4));
        {
            const std::vector<int> ones(D.ngrp, 1);   // no culled pass yet: equal shares
            HIPCHK(c, hipMemcpy(D.tgrp, ones.data(), ones.size() * 4, hipMemcpyHostToDevice));
        }
        // Brick order, which sets the target groups (runs of 64 targets): the
        // axes no longer than twice the cutoff (a film's thickness) vary
        // fastest, since a target's cutoff sphere spans them anyway; over the
        // long axes the bricks follow a Hilbert curve (two long axes: its
        // consecutive cells are always adjacent) or Morton order (three), so
        // that a run of targets keeps a small footprint.  On the 128x128x32
        // solar cell at cutoff 30 this evaluates about 20 % fewer pairs than
        // Morton order over all three axes.  The sums do not depend on it.
        const int ext[3] = {d.X, d.Y, d.Z}, nbs[3] = {D.nbx, D.nby, D.nbz};
        int wide[3], nw = 0;
        for (int a = 0; a < 3; ++a)
            if (ext[a] > 2 * c->p.cutoff) wide[nw++] = a;
        if (nw == 0)
            for (int a = 0; a < 3; ++a) wide[nw++] = a;
        auto spread = [nw](uint64_t v) {
            uint64_t r = 0;
            for (int k = 0; k < 21; ++k) r |= ((v >> k) & 1ull) << (nw * k);
            return r;
        };
        // distance along the Hilbert curve of an n x n grid (n a power of two):
        // per level, the quadrant's rank, then the cell turned into that
        // quadrant's frame
        auto hilbert = [](uint64_t n, uint64_t x, uint64_t y) {
            uint64_t dist = 0;
            for (uint64_t h = n / 2; h > 0; h /= 2) {
                const uint64_t rx = (x & h) ? 1 : 0, ry = (y & h) ? 1 : 0;
                dist += h * h * ((3 * rx) ^ ry);
                if (ry == 0) {
                    if (rx == 1) {
                        x = h - 1 - x;
                        y = h - 1 - y;
                    }
                    std::swap(x, y);
                }
            }
            return dist;
        };
        uint64_t side = 1;
        while (nw == 2 && side < uint64_t(std::max(nbs[wide[0]], nbs[wide[1]]))) side *= 2;
        uint64_t thin = 1;
        for (int a = 0; a < 3; ++a)
            if (std::find(wide, wide + nw, a) == wide + nw) thin *= uint64_t(nbs[a]);
        std::vector<std::pair<uint64_t, int>> key(nb);
        for (int b = 0; b < nb; ++b) {
            const int bc[3] = {b % D.nbx, (b / D.nbx) % D.nby, b / (D.nbx * D.nby)};
            uint64_t m = 0, t = 0;
            if (nw == 2) m = hilbert(side, uint64_t(bc[wide[0]]), uint64_t(bc[wide[1]]));
            else
                for (int k = 0; k < nw; ++k) m |= spread(uint64_t(bc[wide[k]])) << k;
            for (int a = 2; a >= 0; --a)
                if (std::find(wide, wide + nw, a) == wide + nw) t = t * uint64_t(nbs[a]) + uint64_t(bc[a]);
            key[b] = {m * thin + t, b};
        }
        std::sort(key.begin(), key.end());
        std::vector<int> perm(nb);
        for (int r = 0; r < nb; ++r) perm[r] = key[r].second;
        HIPCHK(c, hipMemcpy(bperm, perm.data(), size_t(nb) * 4, hipMemcpyHostToDevice));
        std::vector<int> rank(nb);
        for (int r = 0; r < nb; ++r) rank[perm[r]] = r;
        HIPCHK(c, hipMemcpy(brank, rank.data(), size_t(nb) * 4, hipMemcpyHostToDevice));
        // charged defects by brick, then site, then list order
        std::vector<std::pair<std::pair<int, int>, int>> dk(D.n_d_eff);
        for (int k = 0; k < D.n_d_eff; ++k) {
            int x, y, z;
            decode(d, s->defect_site[k], x, y, z);
            const int b = x / kBrick + D.nbx * (y / kBrick + D.nby * (z / kBrick));
            dk[k] = {{b, (x & 7) + 8 * (y & 7) + 64 * (z & 7)}, k};
        }
        std::sort(dk.begin(), dk.end());
        std::vector<int> ds(size_t(nb) + 1, 0);
        std::vector<float4> dv(D.n_d_eff);
        for (int k = 0; k < D.n_d_eff; ++k) {
            ds[dk[k].first.first + 1]++;
            int x, y, z;
            decode(d, s->defect_site[dk[k].second], x, y, z);
            dv[k] = make_float4(float(x), float(y), float(z), float(D.dq));
        }
        for (int b = 0; b < nb; ++b) ds[b + 1] += ds[b];
        HIPCHK(c, hipMemcpy(dstart, ds.data(), ds.size() * 4, hipMemcpyHostToDevice));
        if (D.n_d_eff > 0) HIPCHK(c, hipMemcpy(dsrc, dv.data(), dv.size() * sizeof(float4), hipMemcpyHostToDevice));
        D.bperm = bperm;
        D.brank = brank;
        D.dstart = dstart;
        D.dsrc = dsrc;
    }
    if (D.rtab_n > 0) {
        // f(r) = 1/r or erf(r / (sqrt2 sigma)) / r in FP64, rounded once to FP32
        // (potential.cpp:351-382, 430-470); 0 at r = 0 and at r^2 = c^2
        std::vector<float> tv(size_t(D.rtab_n), 0.f);
        for (int q = 1; q + 1 < D.rtab_n; ++q) {
            const double rr = std::sqrt(double(q));
            tv[q] = float(c->p.gaussian_sigma > 0 ? std::erf(gk_d * rr) / rr : 1.0 / rr);
        }
        float* rt = nullptr;
        if ((st = sim_alloc(c, rt, tv.size()))) return st;
        HIPCHK(c, hipMemcpy(rt, tv.data(), tv.size() * 4, hipMemcpyHostToDevice));
        D.rtab = rt;
    }
    D.ctl = c->ctl;
    HIPCHK(c, hipMemcpy(c->d_sim, &D, sizeof(SimDev), hipMemcpyHostToDevice));

    // uploads
    HIPCHK(c, hipMemcpy(D.etype, et.data(), Vt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(D.htype, ht.data(), Vt, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(D.vbg, vbg.data(), Vt * sizeof(double), hipMemcpyHostToDevice));
    auto up = [&](int* dst, const int32_t* src, int n) -> lmg_status {
        if (n > 0) {
            if (src) HIPCHK(c, hipMemcpy(dst, src, size_t(n) * 4, hipMemcpyHostToDevice));
            else HIPCHK(c, hipMemset(dst, 0, size_t(n) * 4));
        }
        return LMG_OK;
    };
    if ((st = up(D.e_site[0], s->e_site, s->n_electrons)) || (st = up(D.e_life[0], s->e_lifetime, s->n_electrons)) ||
        (st = up(D.e_path[0], s->e_pathlength, s->n_electrons)) || (st = up(D.h_site[0], s->h_site, s->n_holes)) ||
        (st = up(D.h_life[0], s->h_lifetime, s->n_holes)) || (st = up(D.h_path[0], s->h_pathlength, s->n_holes)))
        return st;
    if (D.n_d_eff > 0) {
        std::vector<float4> df(D.n_d_eff);
        for (int i = 0; i < D.n_d_eff; ++i) {
            int x, y, z;
            decode(d, s->defect_site[i], x, y, z);
            df[i] = make_float4(float(x), float(y), float(z), float(D.dq));
        }
        HIPCHK(c, hipMemcpy(D.chg, df.data(), df.size() * sizeof(float4), hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(fill_u64_kernel, dim3((d.V + 255) / 256), dim3(256), 0, c->stream, D.claim_e, d.V, 0ull);
    hipLaunchKernelGGL(fill_u64_kernel, dim3((d.V + 255) / 256), dim3(256), 0, c->stream, D.claim_h, d.V, 0ull);
    // engine ring: x[k .. k+623] at positions k ..
    std::vector<uint32_t> ringh(ring, 0u);
    const unsigned long long k = s->rng_consumed;
    for (int i = 0; i < 624; ++i) ringh[(k + i) & D.ring_mask] = s->rng_window[i];
    HIPCHK(c, hipMemcpy(D.ring, ringh.data(), ring * 4, hipMemcpyHostToDevice));
    SimCtl C{};
    C.n_e = s->n_electrons;
    C.n_h = s->n_holes;
    C.cur = 0;
    C.error = 0;
    C.rng_pos = k;
    C.rng_step0 = k;
    C.gen_pos = k + 624ull;
    C.ring_base = k;
    for (int i = 0; i < 10; ++i) { C.att[i] = s->flux_attempts[i]; C.suc[i] = s->flux_successes[i]; }
    C.step = s->current_step;
    HIPCHK(c, hipMemcpy(c->ctl, &C, sizeof(C), hipMemcpyHostToDevice));
    HIPCHK(c, hipDeviceSynchronize());
    hipLaunchKernelGGL(inject_kernel, dim3(1), dim3(kOneBlock), 0, c->stream, c->d_sim, 1);
    if ((st = launch_check(c, "inject_kernel(load)")) != LMG_OK) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((st = sim_tune_chunks(c)) != LMG_OK) return st;
    c->sim_loaded = true;
    int32_t code = 0;
    if ((st = lmg_sim_device_error(c, &code)) != LMG_OK) return st;
    if (code) return fail(c, LMG_ERR_DEVICE, "device error " + std::to_string(code) + " while preparing");
    return LMG_OK;
}

lmg_status lmg_sim_set_de_chunks(lmg_ctx* c, int32_t chunks) {
    if (!c) return LMG_ERR_ARG;
    if (chunks < 0) return fail(c, LMG_ERR_ARG, "negative chunk count");
    c->forced_chunks = chunks;
    return LMG_OK;
}

lmg_status lmg_sim_get_tuning(const lmg_ctx* c, int32_t* shape, double* kernel_us) {
    if (!c || !shape || !kernel_us) return LMG_ERR_ARG;
    if (!c->sim_loaded) return LMG_ERR_STATE;
    *shape = c->sim.cull ? c->sim.items_cap : c->sim.S;
    *kernel_us = c->tuned_us;
    return LMG_OK;
}

lmg_status lmg_sim_set_graphs(lmg_ctx* c, int32_t on) {
    if (!c) return LMG_ERR_ARG;
    c->use_graphs = on != 0;
    return LMG_OK;
}

lmg_status lmg_sim_set_de_path(lmg_ctx* c, int32_t mode) {
    if (!c) return LMG_ERR_ARG;
    if (mode < 0 || mode > 2) return fail(c, LMG_ERR_ARG, "dE path must be 0 (automatic), 1 (all pairs) or 2 (culled)");
    c->de_mode = mode;
    return LMG_OK;
}

lmg_status lmg_sim_get_de_path(const lmg_ctx* c, int32_t* mode) {
    if (!c || !mode) return LMG_ERR_ARG;
    if (!c->sim_loaded) return LMG_ERR_STATE;
    *mode = c->sim.cull ? 2 : 1;
    return LMG_OK;
}

lmg_status lmg_sim_set_sync(lmg_ctx* c, int32_t synchronous) {
    if (!c) return LMG_ERR_ARG;
    c->sync_mode = synchronous != 0;
    return LMG_OK;
}

lmg_status lmg_sim_step(lmg_ctx* c, int32_t n) {
    if (!c) return LMG_ERR_ARG;
    if (!c->sim_loaded) return fail(c, LMG_ERR_STATE, "lmg_sim_step before lmg_sim_load");
    if (n < 0) return fail(c, LMG_ERR_ARG, "negative step count");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    if (c->use_graphs && n > 0) {
        if (!c->step_exec) {
            std::string err;
            if ((st = capture_step(c->stream, &c->step_exec, [&] { return sim_launch_step(c); }, err)) != LMG_OK)
                return err.empty() ? st : fail(c, st, err);
        }
        for (int i = 0; i < n; ++i) HIPCHK(c, hipGraphLaunch(c->step_exec, c->stream));
    } else {
        for (int i = 0; i < n; ++i)
            if ((st = sim_launch_step(c)) != LMG_OK) return st;
    }
    if (c->sync_mode) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int32_t code = 0;
        if ((st = lmg_sim_device_error(c, &code)) != LMG_OK) return st;
        if (code) return fail(c, LMG_ERR_DEVICE, "device error " + std::to_string(code));
    }
    return LMG_OK;
}

lmg_status lmg_sim_step_timed(lmg_ctx* c, int32_t n, double* kernel_ms) {
    if (!c || !kernel_ms) return LMG_ERR_ARG;
    if (!c->sim_loaded) return fail(c, LMG_ERR_STATE, "lmg_sim_step_timed before lmg_sim_load");
    if (n < 1) return fail(c, LMG_ERR_ARG, "step count must be >= 1");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    const size_t need = 3 * size_t(n) + 1;
    while (c->events.size() < need) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->events.push_back(e);
    }
    // per step: [3i] dE phase starts, [3i + 1] pair-sum kernel starts,
    // [3i + 2] dE phase ends (commit starts), [3i + 3] commit ends
    HIPCHK(c, hipEventRecord(c->events[0], c->stream));
    for (int i = 0; i < n; ++i) {
        if ((st = sim_launch_de(c, c->events[3 * i + 1])) != LMG_OK) return st;
        HIPCHK(c, hipEventRecord(c->events[3 * i + 2], c->stream));
        if ((st = sim_launch_commit(c)) != LMG_OK) return st;
        HIPCHK(c, hipEventRecord(c->events[3 * i + 3], c->stream));
    }
    HIPCHK(c, hipEventSynchronize(c->events[3 * n]));
    double de = 0.0, cm = 0.0, pk = 0.0;
    for (int i = 0; i < n; ++i) {
        float a = 0.f, b = 0.f, p = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, c->events[3 * i], c->events[3 * i + 2]));
        HIPCHK(c, hipEventElapsedTime(&p, c->events[3 * i + 1], c->events[3 * i + 2]));
        HIPCHK(c, hipEventElapsedTime(&b, c->events[3 * i + 2], c->events[3 * i + 3]));
        de += a;
        pk += p;
        cm += b;
    }
    float span = 0.f;
    HIPCHK(c, hipEventElapsedTime(&span, c->events[0], c->events[3 * n]));
    kernel_ms[0] = de;
    kernel_ms[1] = cm;
    kernel_ms[2] = span;
    kernel_ms[3] = pk;
    int32_t code = 0;
    if ((st = lmg_sim_device_error(c, &code)) != LMG_OK) return st;
    if (code) return fail(c, LMG_ERR_DEVICE, "device error " + std::to_string(code));
    return LMG_OK;
}

lmg_status lmg_sim_device_error(lmg_ctx* c, int32_t* code) {
    if (!c || !code) return LMG_ERR_ARG;
    if (!c->ctl) return fail(c, LMG_ERR_STATE, "no simulation loaded");
    SimCtl C;
    HIPCHK(c, hipSetDevice(c->dev));
    HIPCHK(c, hipMemcpyAsync(&C, c->ctl, sizeof(C), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *code = C.error;
    return LMG_OK;
}

lmg_status lmg_sim_store(lmg_ctx* c, lmg_sim_state* s) {
    if (!c || !s) return LMG_ERR_ARG;
    if (!c->sim_loaded) return fail(c, LMG_ERR_STATE, "no simulation loaded");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    SimCtl C;
    HIPCHK(c, hipMemcpy(&C, c->ctl, sizeof(C), hipMemcpyDeviceToHost));
    if (C.error) return fail(c, LMG_ERR_DEVICE, "device error " + std::to_string(C.error));
    const SimDev& D = c->sim;
    s->n_electrons = C.n_e;
    s->n_holes = C.n_h;
    auto down = [&](int32_t* dst, const int* src, int n) -> lmg_status {
        if (dst && n > 0) HIPCHK(c, hipMemcpy(dst, src, size_t(n) * 4, hipMemcpyDeviceToHost));
        return LMG_OK;
    };
    if ((st = down(s->e_site, D.e_site[C.cur], C.n_e)) || (st = down(s->e_lifetime, D.e_life[C.cur], C.n_e)) ||
        (st = down(s->e_pathlength, D.e_path[C.cur], C.n_e)) || (st = down(s->h_site, D.h_site[C.cur], C.n_h)) ||
        (st = down(s->h_lifetime, D.h_life[C.cur], C.n_h)) || (st = down(s->h_pathlength, D.h_path[C.cur], C.n_h)))
        return st;
    const unsigned long long k = C.rng_step0;
    const unsigned long long ring = D.ring_mask + 1ull, start = k & D.ring_mask;
    const unsigned long long first = std::min<unsigned long long>(624ull, ring - start);
    HIPCHK(c, hipMemcpy(s->rng_window, D.ring + start, first * 4, hipMemcpyDeviceToHost));
    if (first < 624ull)
        HIPCHK(c, hipMemcpy(s->rng_window + first, D.ring, (624ull - first) * 4, hipMemcpyDeviceToHost));
    s->rng_consumed = k;
    for (int i = 0; i < 10; ++i) { s->flux_attempts[i] = C.att[i]; s->flux_successes[i] = C.suc[i]; }
    s->current_step = C.step;
    return LMG_OK;
}

lmg_status lmg_sim_delta_e(lmg_ctx* c, int32_t* fsite_e, int32_t* fsite_h, double* de_e, double* de_h) {
    if (!c) return LMG_ERR_ARG;
    if (!c->sim_loaded) return fail(c, LMG_ERR_STATE, "no simulation loaded");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    if ((st = sim_launch_de(c)) != LMG_OK) return st;
    hipLaunchKernelGGL(decide_kernel, dim3(c->sim.nblk), dim3(kSimBlock), 0, c->stream, c->d_sim, 1);
    if ((st = launch_check(c, "decide_kernel(dE)")) != LMG_OK) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    SimCtl C;
    HIPCHK(c, hipMemcpy(&C, c->ctl, sizeof(C), hipMemcpyDeviceToHost));
    if (C.error) return fail(c, LMG_ERR_DEVICE, "device error " + std::to_string(C.error));
    const SimDev& D = c->sim;
    if (fsite_e && C.n_e) HIPCHK(c, hipMemcpy(fsite_e, D.e_fsite, size_t(C.n_e) * 4, hipMemcpyDeviceToHost));
    if (fsite_h && C.n_h) HIPCHK(c, hipMemcpy(fsite_h, D.h_fsite, size_t(C.n_h) * 4, hipMemcpyDeviceToHost));
    if (de_e && C.n_e) HIPCHK(c, hipMemcpy(de_e, D.de, size_t(C.n_e) * 8, hipMemcpyDeviceToHost));
    if (de_h && C.n_h) HIPCHK(c, hipMemcpy(de_h, D.de + C.n_e, size_t(C.n_h) * 8, hipMemcpyDeviceToHost));
    return LMG_OK;
}

}  // extern "C"

// ===========================================================================
// Batched replicas: many independent simulations, each kernel of a step
// launched once for all of them (blockIdx.z = replica).
// ===========================================================================
struct lmg_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    // the culled replicas' dE kernels run on a second stream, forked from and
    // joined to the batch stream, beside the all-pairs replicas' de_kernel
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::string err;
    std::vector<lmg_ctx*> reps;
    SimDev* d_sims = nullptr;   // every replica (the commit kernels)
    int max_nblk = 0;
    // all-pairs replicas: their descriptors, de_kernel's grid
    SimDev* d_pairs = nullptr;   // those with a pair-function table first
    int n_pairs = 0, n_pairs_tab = 0, max_nblk_p = 0, max_S_p = 0;
    // culled replicas: their descriptors and the workgroup maps of the two
    // culled kernels (launch workgroup -> (culled replica, its workgroup))
    SimDev* d_cull = nullptr;
    int n_cull = 0;
    int2* d_emit_map = nullptr;
    int2* d_de_map = nullptr;
    int emit_wgs = 0, de_wgs = 0, de_map_cap = 0;
    std::vector<int> paths;   // per replica: 1 all pairs, 2 culled
    int tuned_S = 0;          // autotuned common charge-chunk count of the all-pairs replicas
    double tuned_us = -1.0;   // the dE phase of every replica at the chosen shapes
    std::string report;       // every timed candidate (JSON)
    bool coulomb = false, gauss = false, any_recomb = false;
    bool cull_table = false;  // the culled replicas' dE kernel in table form
    int cull_tab_n = 0;       // ... and its table entries (the largest replica's)
    bool use_graphs = true;
    hipGraphExec_t step_exec = nullptr;   // one step of every replica, captured at the first step
    std::vector<hipEvent_t> events;
};

namespace {

thread_local std::string g_batch_error;

lmg_status bfail(lmg_batch* b, lmg_status code, const std::string& msg) {
    if (b) b->err = msg;
    else g_batch_error = msg;
    return code;
}

#define BHIPCHK(b, call)                                                               \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return bfail((b), LMG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// the dE phase of every replica (pair sums, decision flags, random-stream
// refill): the culled replicas' list building and sums and de_kernel over the
// all-pairs ones, concurrently (the culled ones on the side stream) when the
// batch has both
void batch_launch_de(lmg_batch* b) {
    const dim3 blk(kSimBlock);
    if (b->n_cull) {
        const bool both = b->n_pairs > 0;
        hipStream_t st = both ? b->side : b->stream;
        if (both) {
            (void)hipEventRecord(b->fork, b->stream);
            (void)hipStreamWaitEvent(b->side, b->fork, 0);
        }
        hipLaunchKernelGGL(cull_emit_kernel, dim3(b->emit_wgs), blk, 0, st, b->d_cull,
                           (const int2*)b->d_emit_map);
        launch_cull_de(dim3(b->de_wgs), st, b->d_cull, (const int2*)b->d_de_map, b->cull_table, b->gauss,
                       b->cull_tab_n);
    }
    if (b->n_pairs) {
        // the replicas with a pair-function table first, then the others
        const int nt = std::min(b->n_pairs_tab, b->n_pairs), na = b->n_pairs - nt;
        const int gy = b->coulomb ? b->max_S_p : 1;
        if (nt) launch_de(dim3(b->max_nblk_p + 1, gy, nt), b->stream, b->d_pairs, b->coulomb, true, b->gauss);
        if (na) launch_de(dim3(b->max_nblk_p + 1, gy, na), b->stream, b->d_pairs + nt, b->coulomb, false, b->gauss);
        if (b->n_cull) {   // join the side stream
            (void)hipEventRecord(b->join, b->side);
            (void)hipStreamWaitEvent(b->stream, b->join, 0);
        }
    }
}

// decisions, recombination, ticks, injection and the next proposals of every replica
lmg_status batch_launch_commit(lmg_batch* b) {
    const dim3 blk(kSimBlock);
    const int R = int(b->reps.size());
    const dim3 g_grid(b->max_nblk, 1, R), g_one(1, 1, R);
    hipLaunchKernelGGL(decide_kernel, g_grid, blk, 0, b->stream, b->d_sims, 0);
    if (b->any_recomb) hipLaunchKernelGGL(recomb_kernel, g_one, dim3(kOneBlock), 0, b->stream, b->d_sims);
    hipLaunchKernelGGL(tick_kernel, g_grid, blk, 0, b->stream, b->d_sims);
    hipLaunchKernelGGL(inject_kernel, g_one, dim3(kOneBlock), 0, b->stream, b->d_sims, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, LMG_ERR_HIP, std::string("batch step: ") + hipGetErrorString(e));
    return LMG_OK;
}

lmg_status batch_launch_step(lmg_batch* b) {
    batch_launch_de(b);
    return batch_launch_commit(b);
}

// n steps of every replica: graph replays of one captured step, or kernel by kernel
lmg_status batch_steps(lmg_batch* b, int n) {
    if (b->use_graphs && n > 0) {
        if (!b->step_exec) {
            std::string err;
            const lmg_status st = capture_step(b->stream, &b->step_exec, [&] { return batch_launch_step(b); }, err);
            if (st != LMG_OK) return err.empty() ? st : bfail(b, st, err);
        }
        for (int i = 0; i < n; ++i) BHIPCHK(b, hipGraphLaunch(b->step_exec, b->stream));
        return LMG_OK;
    }
    for (int i = 0; i < n; ++i) {
        const lmg_status st = batch_launch_step(b);
        if (st != LMG_OK) return st;
    }
    return LMG_OK;
}


// the single context's geometric test for the culled path: the cutoff box
// around a group of targets (about 16 sites wide) covers well under the lattice
bool cull_candidate(const lmg_ctx* c) {
    auto frac = [&](int n) { return std::min(1.0, double(16 + 2 * c->p.cutoff) / n); };
    return frac(c->d.X) * frac(c->d.Y) * frac(c->d.Z) < 0.6;
}

// Puts a plan into the batch's descriptors and workgroup maps: culled[r] for the
// replicas on the culled path (with quads[r] work-item quads), the others on the
// all-pairs path with S charge chunks (capped per replica).  The replicas' own
// descriptors follow, so that their single-context calls (lmg_batch_delta_e)
// take the same path.
lmg_status batch_apply(lmg_batch* b, const std::vector<char>& culled, int S, const std::vector<int>& quads) {
    const int n = int(b->reps.size());
    std::vector<SimDev> all(n), pairs, pairs_ar, cull;
    std::vector<int> cull_rep;
    b->max_nblk_p = 0;
    b->max_S_p = 1;
    for (int r = 0; r < n; ++r) {
        lmg_ctx* c = b->reps[r];
        SimDev& D = c->sim;
        D.cull = culled[r] ? 1 : 0;
        if (D.cull) {
            D.items_cap = 4 * quads[r];
            cull.push_back(D);
            cull_rep.push_back(r);
        } else {
            D.S = std::max(1, std::min(D.S_pairs, S));
            (D.coulomb && D.rtab_n == 0 ? pairs_ar : pairs).push_back(D);
            b->max_nblk_p = std::max(b->max_nblk_p, D.nblk);
            b->max_S_p = std::max(b->max_S_p, D.S);
        }
        all[r] = D;
        b->paths[r] = D.cull ? 2 : 1;
        BHIPCHK(b, hipMemcpy(c->d_sim, &D, sizeof(SimDev), hipMemcpyHostToDevice));
    }
    b->n_pairs_tab = int(pairs.size());
    pairs.insert(pairs.end(), pairs_ar.begin(), pairs_ar.end());
    b->n_pairs = int(pairs.size());
    b->n_cull = int(cull.size());
    b->cull_table = b->n_cull > 0;
    b->cull_tab_n = 0;
    for (const SimDev& D : cull) {
        b->cull_table = b->cull_table && D.rtab_n > 0;
        b->cull_tab_n = std::max(b->cull_tab_n, D.rtab_n);
    }
    BHIPCHK(b, hipMemcpy(b->d_sims, all.data(), sizeof(SimDev) * n, hipMemcpyHostToDevice));
    if (b->n_pairs) BHIPCHK(b, hipMemcpy(b->d_pairs, pairs.data(), sizeof(SimDev) * b->n_pairs, hipMemcpyHostToDevice));
    if (!b->n_cull) return LMG_OK;
    BHIPCHK(b, hipMemcpy(b->d_cull, cull.data(), sizeof(SimDev) * b->n_cull, hipMemcpyHostToDevice));
    // cull_emit_kernel: replica by replica, ceil(nbrick / 4) + nblk + 1 workgroups each
    std::vector<int2> emit, de;
    for (int j = 0; j < b->n_cull; ++j) {
        const SimDev& D = cull[j];
        for (int x = 0, e = (D.nbrick + 3) / 4 + D.nblk + 1; x < e; ++x) emit.push_back(make_int2(j, x));
    }
    // cull_de_kernel: every replica's refill workgroup first (they run the
    // longest serial loop), then the work items, the replicas with the most
    // work per item first, so that the last workgroups to start are short ones
    std::vector<std::pair<double, int>> order;
    for (int j = 0; j < b->n_cull; ++j) {
        const SimDev& D = cull[j];
        de.push_back(make_int2(j, D.items_cap / 4));
        const double cap = D.cap_e + D.cap_h + D.n_d_eff;
        order.push_back({-cap * cap / double(D.items_cap / 4), j});
    }
    std::sort(order.begin(), order.end());
    for (const auto& o : order)
        for (int x = 0; x < cull[o.second].items_cap / 4; ++x) de.push_back(make_int2(o.second, x));
    // and last each replica's decide_prep workgroups (short: they fill slots
    // that the work items free)
    for (int j = 0; j < b->n_cull; ++j) {
        const SimDev& D = cull[j];
        for (int x = 0; x < D.nblk; ++x) de.push_back(make_int2(j, D.items_cap / 4 + 1 + x));
    }
    const int need = int(std::max(emit.size(), de.size()));
    if (need > b->de_map_cap) {
        if (b->d_emit_map) (void)hipFree(b->d_emit_map);
        if (b->d_de_map) (void)hipFree(b->d_de_map);
        b->d_emit_map = b->d_de_map = nullptr;
        BHIPCHK(b, hipMalloc(&b->d_emit_map, sizeof(int2) * need));
        BHIPCHK(b, hipMalloc(&b->d_de_map, sizeof(int2) * need));
        b->de_map_cap = need;
    }
    BHIPCHK(b, hipMemcpy(b->d_emit_map, emit.data(), sizeof(int2) * emit.size(), hipMemcpyHostToDevice));
    BHIPCHK(b, hipMemcpy(b->d_de_map, de.data(), sizeof(int2) * de.size(), hipMemcpyHostToDevice));
    b->emit_wgs = int(emit.size());
    b->de_wgs = int(de.size());
    return LMG_OK;
}

// The batch's autotuner.  Candidates, each timed twice (the second round sees
// the clocks the first one raised) and scored by its faster time (the chosen
// split is then timed as it runs, its two parts concurrently):
//   culled replicas: the total number of cull_de_kernel workgroups (one per
//     target group at least, then 1 .. 4 device-resident waves of workgroups,
//     or 1.5 and 2 per target group), shared out in proportion to each
//     replica's estimated pair count;
//   all-pairs replicas: the common charge-chunk count S (powers of two);
//   automatic path: the same S search with every replica on the all-pairs
//     path, against the best culled split.
// A dE launch only rewrites per-step scratch, and neither the path nor the
// shapes change a replica's results beyond the FP32 rounding of its unit sums
// (bit-identical to its single context on the same path).
lmg_status batch_plan(lmg_batch* b, int de_path) {
    const int n = int(b->reps.size());
    b->paths.assign(n, 1);
    std::vector<char> capable(n, 0);
    int max_charges = 1;
    for (int r = 0; r < n; ++r) {
        const SimDev& D = b->reps[r]->sim;
        capable[r] = D.cull ? 1 : 0;
        b->max_nblk = std::max(b->max_nblk, D.nblk);
        b->any_recomb |= (D.sim_type == 1 && D.rate[LMG_FLUX_XD] > 0);
        max_charges = std::max(max_charges, D.cap_e + D.cap_h + D.n_d_eff);
    }
    // one form of the culled kernel serves every culled replica: when some of the
    // cullable replicas have a pair-function table, only those are culled, so
    // each culled replica sums as its own context does
    bool any_tab = false;
    for (int r = 0; r < n; ++r) any_tab = any_tab || (capable[r] && b->reps[r]->sim.rtab_n > 0);
    if (any_tab)
        for (int r = 0; r < n; ++r)
            if (b->reps[r]->sim.rtab_n == 0) capable[r] = 0;
    BHIPCHK(b, hipMalloc(&b->d_sims, sizeof(SimDev) * n));
    BHIPCHK(b, hipMalloc(&b->d_pairs, sizeof(SimDev) * n));
    BHIPCHK(b, hipMalloc(&b->d_cull, sizeof(SimDev) * n));
    const std::vector<char> none(n, 0);
    const bool any_cull = std::find(capable.begin(), capable.end(), 1) != capable.end();
    std::ostringstream js;
    js << "{";
    auto timed = [&](double& best) {
        const double us = !b->coulomb ? 0.0 : time_launch_us(b->stream, [&] { batch_launch_de(b); }, 5);
        best = std::min(best, us);
        return us;
    };

    // culled replicas: total work-item quads
    std::vector<int> quads(n, 0);
    double cull_us = 1e30;
    if (any_cull) {
        long long groups = 0, alloc = 0, ncap = 0;
        double wsum = 0.0;
        std::vector<double> w(n, 0.0);
        int resident = 1;
        for (int r = 0; r < n; ++r) {
            if (!capable[r]) continue;
            const lmg_ctx* c = b->reps[r];
            const SimDev& D = c->sim;
            groups += D.ngrp;
            alloc += c->cull_quads_alloc;
            ++ncap;
            resident = std::max(resident, c->cull_resident);
            auto frac = [&](int e) { return std::min(1.0, double(16 + 2 * c->p.cutoff) / e); };
            const double cap = D.cap_e + D.cap_h + D.n_d_eff;
            w[r] = cap * cap * frac(c->d.X) * frac(c->d.Y) * frac(c->d.Z);
            wsum += w[r];
        }
        std::vector<long long> cand;
        for (int k = 1; k <= 4; ++k) cand.push_back(std::max(groups, (long long)k * resident - ncap));
        cand.push_back((3 * groups + 1) / 2);
        cand.push_back(2 * groups);
        for (long long& t : cand) t = std::min(t, alloc);
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        auto share = [&](long long total, std::vector<int>& q) {
            const double extra = double(total - groups);
            for (int r = 0; r < n; ++r) {
                if (!capable[r]) continue;
                const lmg_ctx* c = b->reps[r];
                const int want = c->sim.ngrp + int(extra * w[r] / std::max(wsum, 1e-30));
                q[r] = std::max(c->sim.ngrp, std::min(want, c->cull_quads_alloc));
            }
        };
        std::vector<double> best(cand.size(), 1e30);
        for (int round = 0; round < 2; ++round)
            for (size_t k = 0; k < cand.size(); ++k) {
                share(cand[k], quads);
                lmg_status st = batch_apply(b, capable, 1, quads);
                if (st != LMG_OK) return st;
                b->n_pairs = 0;   // the culled replicas alone
                timed(best[k]);
            }
        size_t kb = 0;
        js << "\"cull_candidates\": [";
        for (size_t k = 0; k < cand.size(); ++k) {
            js << (k ? ", " : "") << "[" << cand[k] << ", " << best[k] << "]";
            if (best[k] < best[kb]) kb = k;
        }
        js << "], ";
        share(cand[kb], quads);
        cull_us = best[kb];
    }

    // all-pairs replicas: S, with the cullable ones culled (split) or not (all pairs)
    auto tune_s = [&](const std::vector<char>& culled, const char* name, int& best_S) {
        const std::vector<int> cands = chunk_candidates(std::min(max_charges, 32 * 64));   // up to 64 chunks
        std::vector<double> best(cands.size(), 1e30);
        for (int round = 0; round < 2; ++round)
            for (size_t k = 0; k < cands.size(); ++k) {
                if (batch_apply(b, culled, cands[k], quads) != LMG_OK) return -1.0;
                b->n_cull = 0;   // the all-pairs replicas alone
                timed(best[k]);
                if (!b->coulomb) break;
            }
        size_t kb = 0;
        js << "\"" << name << "\": [";
        for (size_t k = 0; k < cands.size(); ++k) {
            js << (k ? ", " : "") << "[" << cands[k] << ", " << best[k] << "]";
            if (best[k] < best[kb]) kb = k;
        }
        js << "], ";
        best_S = cands[kb];
        return best[kb];
    };
    std::vector<char> plan = de_path == 1 ? none : capable;
    int S_split = 1, S_all = 1;
    double split_us = any_cull && de_path != 1 ? cull_us : 0.0;
    const bool rest = std::find(plan.begin(), plan.end(), 0) != plan.end();
    if (rest) {
        const double t = tune_s(plan, "pairs_candidates", S_split);
        if (t < 0) return b->err.empty() ? bfail(b, LMG_ERR_HIP, "batch autotune failed") : LMG_ERR_HIP;
        split_us += t;
        if (any_cull && de_path != 1) {   // both parts as they run, concurrently
            const lmg_status st = batch_apply(b, plan, S_split, quads);
            if (st != LMG_OK) return st;
            double best = 1e30;
            for (int round = 0; round < 2; ++round) timed(best);
            js << "\"split_us\": " << best << ", ";
            split_us = best;
        }
    }
    int S = S_split;
    if (de_path == 0 && any_cull) {   // the same replicas all on the all-pairs path
        const double all_us = tune_s(none, "all_pairs_candidates", S_all);
        if (all_us < 0) return b->err.empty() ? bfail(b, LMG_ERR_HIP, "batch autotune failed") : LMG_ERR_HIP;
        if (all_us < split_us) {
            plan = none;
            S = S_all;
            split_us = all_us;
        }
    }
    const lmg_status st = batch_apply(b, plan, S, quads);
    if (st != LMG_OK) return st;
    b->tuned_S = b->n_pairs ? S : 0;
    b->tuned_us = b->coulomb ? split_us : -1.0;
    long long qt = 0;
    for (int r = 0; r < n; ++r) qt += plan[r] ? quads[r] : 0;
    js << "\"culled_replicas\": " << b->n_cull << ", \"cull_workgroups\": " << qt
       << ", \"charge_chunks\": " << b->tuned_S << ", \"de_phase_us\": " << b->tuned_us << "}";
    b->report = js.str();
    return LMG_OK;
}

}  // namespace

extern "C" {

const char* lmg_batch_last_error(const lmg_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

void lmg_batch_destroy(lmg_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->step_exec) (void)hipGraphExecDestroy(b->step_exec);
    for (lmg_ctx* c : b->reps) lmg_destroy(c);
    for (void* p : {(void*)b->d_sims, (void*)b->d_pairs, (void*)b->d_cull, (void*)b->d_emit_map, (void*)b->d_de_map})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

lmg_status lmg_batch_create(int32_t n, const lmg_params* params, const lmg_sim_config* configs,
                            const lmg_sim_state* states, int32_t device_id, lmg_batch** out) {
    return lmg_batch_create_with_path(n, params, configs, states, device_id, 0, out);
}

lmg_status lmg_batch_create_with_path(int32_t n, const lmg_params* params, const lmg_sim_config* configs,
                                      const lmg_sim_state* states, int32_t device_id, int32_t de_path,
                                      lmg_batch** out) {
    if (!out || n < 1 || !params || !configs || !states) return bfail(nullptr, LMG_ERR_ARG, "invalid batch arguments");
    if (de_path < 0 || de_path > 2)
        return bfail(nullptr, LMG_ERR_ARG, "dE path must be 0 (automatic), 1 (all pairs) or 2 (culled)");
    *out = nullptr;
    lmg_batch* b = new lmg_batch;
    b->dev = device_id;
    auto give_up = [&](lmg_status st, const std::string& msg) {
        bfail(nullptr, st, msg);
        lmg_batch_destroy(b);
        return st;
    };
    for (int r = 0; r < n; ++r) {
        lmg_params p = params[r];
        p.device_id = device_id;
        lmg_ctx* c = nullptr;
        lmg_status st = lmg_create(&p, &c);
        if (st != LMG_OK) return give_up(st, "replica " + std::to_string(r) + ": " + lmg_last_error(nullptr));
        b->reps.push_back(c);
        c->autotune = false;
        // the culled path for the replicas that can use it (de_path 2), or that
        // the single context's geometric test picks (automatic)
        c->de_mode = de_path == 2 || (de_path == 0 && cull_candidate(c)) ? 2 : 1;
        st = lmg_sim_load(c, &configs[r], &states[r]);
        if (st != LMG_OK) return give_up(st, "replica " + std::to_string(r) + ": " + lmg_last_error(c));
        const SimDev& D = c->sim;
        if (r == 0) {
            b->coulomb = D.coulomb;
            b->gauss = D.gauss;
        } else if (b->coulomb != bool(D.coulomb) || (D.coulomb && b->gauss != bool(D.gauss))) {
            return give_up(LMG_ERR_ARG, "replicas must agree on coulomb.carriers and Gaussian screening");
        }
    }
    if (hipSetDevice(device_id) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
        return give_up(LMG_ERR_HIP, "batch stream creation failed");
    const lmg_status st = batch_plan(b, de_path);
    if (st != LMG_OK) {
        const std::string msg = b->err;
        return give_up(st, msg);
    }
    *out = b;
    return LMG_OK;
}

lmg_status lmg_batch_get_tuning(const lmg_batch* b, int32_t* chunks, double* kernel_us) {
    if (!b || !chunks || !kernel_us) return LMG_ERR_ARG;
    *chunks = b->tuned_S;
    *kernel_us = b->tuned_us;
    return LMG_OK;
}

lmg_status lmg_batch_get_de_paths(const lmg_batch* b, int32_t* paths) {
    if (!b || !paths) return LMG_ERR_ARG;
    for (size_t r = 0; r < b->paths.size(); ++r) paths[r] = b->paths[r];
    return LMG_OK;
}

const char* lmg_batch_tuning_report(const lmg_batch* b) { return b ? b->report.c_str() : ""; }

lmg_status lmg_batch_step(lmg_batch* b, int32_t n) {
    if (!b) return LMG_ERR_ARG;
    if (n < 0) return bfail(b, LMG_ERR_ARG, "negative step count");
    BHIPCHK(b, hipSetDevice(b->dev));
    return batch_steps(b, n);
}

lmg_status lmg_batch_set_graphs(lmg_batch* b, int32_t on) {
    if (!b) return LMG_ERR_ARG;
    b->use_graphs = on != 0;
    return LMG_OK;
}

lmg_status lmg_batch_step_timed(lmg_batch* b, int32_t n, double* kernel_ms) {
    if (!b || !kernel_ms || n < 1) return LMG_ERR_ARG;
    BHIPCHK(b, hipSetDevice(b->dev));
    while (b->events.size() < 2) {
        hipEvent_t e;
        BHIPCHK(b, hipEventCreate(&e));
        b->events.push_back(e);
    }
    BHIPCHK(b, hipEventRecord(b->events[0], b->stream));
    const lmg_status st = batch_steps(b, n);
    if (st != LMG_OK) return st;
    BHIPCHK(b, hipEventRecord(b->events[1], b->stream));
    BHIPCHK(b, hipEventSynchronize(b->events[1]));
    float ms = 0.f;
    BHIPCHK(b, hipEventElapsedTime(&ms, b->events[0], b->events[1]));
    kernel_ms[0] = ms;
    return LMG_OK;
}

lmg_status lmg_batch_step_phases(lmg_batch* b, int32_t n, double* kernel_ms) {
    if (!b || !kernel_ms || n < 1) return bfail(b, LMG_ERR_ARG, "invalid lmg_batch_step_phases arguments");
    BHIPCHK(b, hipSetDevice(b->dev));
    const size_t need = 2 * size_t(n) + 1;
    while (b->events.size() < need) {
        hipEvent_t e;
        BHIPCHK(b, hipEventCreate(&e));
        b->events.push_back(e);
    }
    // per step: [2i] dE kernel starts, [2i + 1] dE kernel ends (commit starts), [2i + 2] commit ends
    BHIPCHK(b, hipEventRecord(b->events[0], b->stream));
    for (int i = 0; i < n; ++i) {
        batch_launch_de(b);
        BHIPCHK(b, hipEventRecord(b->events[2 * i + 1], b->stream));
        const lmg_status st = batch_launch_commit(b);
        if (st != LMG_OK) return st;
        BHIPCHK(b, hipEventRecord(b->events[2 * i + 2], b->stream));
    }
    BHIPCHK(b, hipEventSynchronize(b->events[2 * n]));
    double de = 0.0, cm = 0.0;
    for (int i = 0; i < n; ++i) {
        float a = 0.f, c = 0.f;
        BHIPCHK(b, hipEventElapsedTime(&a, b->events[2 * i], b->events[2 * i + 1]));
        BHIPCHK(b, hipEventElapsedTime(&c, b->events[2 * i + 1], b->events[2 * i + 2]));
        de += a;
        cm += c;
    }
    float span = 0.f;
    BHIPCHK(b, hipEventElapsedTime(&span, b->events[0], b->events[2 * n]));
    kernel_ms[0] = de;
    kernel_ms[1] = cm;
    kernel_ms[2] = span;
    return LMG_OK;
}

lmg_status lmg_batch_sync(lmg_batch* b) {
    if (!b) return LMG_ERR_ARG;
    BHIPCHK(b, hipSetDevice(b->dev));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    return LMG_OK;
}

lmg_status lmg_batch_store(lmg_batch* b, int32_t replica, lmg_sim_state* state) {
    if (!b || !state || replica < 0 || replica >= int(b->reps.size()))
        return bfail(b, LMG_ERR_ARG, "invalid replica index");
    BHIPCHK(b, hipSetDevice(b->dev));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    lmg_status st = lmg_sim_store(b->reps[replica], state);
    if (st != LMG_OK) return bfail(b, st, lmg_last_error(b->reps[replica]));
    return LMG_OK;
}

lmg_status lmg_batch_delta_e(lmg_batch* b, int32_t replica, int32_t* fsite_e, int32_t* fsite_h, double* de_e,
                             double* de_h) {
    if (!b || replica < 0 || replica >= int(b->reps.size())) return bfail(b, LMG_ERR_ARG, "invalid replica index");
    BHIPCHK(b, hipSetDevice(b->dev));
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    lmg_status st = lmg_sim_delta_e(b->reps[replica], fsite_e, fsite_h, de_e, de_h);
    if (st != LMG_OK) return bfail(b, st, lmg_last_error(b->reps[replica]));
    return LMG_OK;
}

int32_t lmg_batch_size(const lmg_batch* b) { return b ? int32_t(b->reps.size()) : 0; }

}  // extern "C"
