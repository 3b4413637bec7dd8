// lmg_surrogate.hip -- learned spectral surrogate of the full-lattice potential
// map (the coulomb1/gauss1 map of lmg_potential_map, kernel.cl:56-220 in the
// reference), trained on the GPU by this file's own forward and backward passes.
//
// The map is a linear, translation-invariant operator on the charge grid,
// map = K * q with K(d) = pref f(|d|) for 0 < |d|^2 < c^2.  Zero-padded to
// P >= N + c - 1 sites per axis (P >= 2N - 1 without a cutoff) the aperiodic
// convolution equals a circular one, which the DFT diagonalises, so the network
// is one Fourier layer that keeps every mode of the padded grid:
//
//     map = window( irfft3( W . rfft3(pad(q)) ) ),   W: Pz x Py x (Px/2 + 1) complex
//
// The six DFT stages are GEMMs with constant DFT matrices on the matrix cores
// (lmg_gemm.h, BF16 hi/lo operands, FP32 accumulation).  Training: momentum SGD
// on L = sum_b |map_b - target_b|^2 / sum_b |target_b|^2 with the gradient from
// the adjoint chain (the transposed GEMM stages).  The updates keep W real and
// even in ky and kz -- the map commutes with the lattice reflections, so the
// exact operator has that symmetry, and tying the weights to it is what lets
// strongly correlated KMC snapshots pin the operator down -- and are
// preconditioned per frequency by the inverse mean spectral power of the
// training charges; the step size comes from a power iteration on the
// preconditioned curvature.  Inference behind an error gate: the exact
// pair sum (lmg_pair_potential_dev) at sampled sites judges the result and the
// exact map (lmg_potential_map_dev) replaces it when the sampled relative RMS
// error exceeds the caller's tolerance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "langmuir_gpu.h"
#include "lmg_device.h"
#include "lmg_gemm.h"
#include "lmg_mapcull.h"
#include "lmg_pairsum.h"

namespace {

std::string g_surr_create_error;

constexpr int kRedBlocks = 512;   // blocks of the deterministic partial sums
constexpr int kRedThreads = 256;
constexpr int kPowerIters = 12;   // power iterations for the curvature estimate
constexpr double kStepFactor = 0.9;
constexpr int kPrecondSnapshots = 256;   // occupancies averaged into the preconditioner
constexpr double kPrecondFloor = 1e-3;   // relative floor of the spectral power
constexpr int kGateMaxSplit = 128;       // charge chunks of the gate's pair sums (split_count's cap)
constexpr int kMaxGateSites = 8192;      // gate sites per map

// Spectrum layout (one sample): complex values (re, im) at index
// k = (kx Py + ky) Pz + kz, i.e. rows (kx, ky) of Pz interleaved pairs.
struct SpecDims {
    int py, pz, kxn;
};

// Mirror images of k under ky -> -ky and kz -> -kz (the lattice reflections
// y -> -y, z -> -z; x -> -x is the complex conjugate, which realness of the
// weights covers).
__device__ __forceinline__ void spec_images(long long k, SpecDims sd, long long img[4]) {
    const int kz = int(k % sd.pz);
    const long long r = k / sd.pz;
    const int ky = int(r % sd.py);
    const long long kx = r / sd.py;
    const int my = ky ? sd.py - ky : 0, mz = kz ? sd.pz - kz : 0;
    img[0] = k;
    img[1] = (kx * sd.py + my) * sd.pz + kz;
    img[2] = (kx * sd.py + ky) * sd.pz + mz;
    img[3] = (kx * sd.py + my) * sd.pz + mz;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (t < o) sh[t] += sh[t + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ void fill_u32_kernel(uint32_t* p, long long n, uint32_t v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// Adds q to the int8 occupancy of each site; a 32-bit compare-and-swap on the
// word holding the byte keeps coincident sites (an electron and a hole) exact.
__global__ void occ_scatter_kernel(const int32_t* __restrict__ site, const int32_t* __restrict__ q, int n,
                                   long long V, int8_t* occ, int* bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = site[i];
    const int qi = q[i];
    if (s < 0 || s >= V || qi < -127 || qi > 127) {
        atomicOr(bad, 1);
        return;
    }
    unsigned int* w = reinterpret_cast<unsigned int*>(occ + (s & ~3));
    const int sh = (s & 3) * 8;
    unsigned int old = *w, assumed;
    do {
        assumed = old;
        const int b = int(int8_t((assumed >> sh) & 0xffu)) + qi;
        const unsigned int nv = (assumed & ~(0xffu << sh)) | (unsigned(uint8_t(int8_t(b))) << sh);
        old = atomicCAS(w, assumed, nv);
    } while (old != assumed);
}

// map := 2 (map - target); part[2 blk] = sum (map - target)^2, part[2 blk + 1] =
// sum target^2 (target NULL means zero)
__global__ void residual_kernel(float* __restrict__ map, const float* __restrict__ tgt, long long n,
                                double* __restrict__ part) {
    __shared__ double sh[kRedThreads];
    double rr = 0.0, tt = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float t = tgt ? tgt[i] : 0.f;
        const float r = map[i] - t;
        rr += double(r) * r;
        tt += double(t) * t;
        map[i] = 2.f * r;
    }
    rr = block_sum(rr, sh);
    tt = block_sum(tt, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = rr;
        part[2 * blockIdx.x + 1] = tt;
    }
}

// part[blk] = sum of v^2 over the block's grid-stride share
__global__ void sumsq_kernel(const float* __restrict__ v, long long n, double* __restrict__ part) {
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        s += double(v[i]) * v[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[c] = sum_blk part[ncomp blk + c], in a fixed order
__global__ void reduce_kernel(const double* __restrict__ part, int nblk, int ncomp, double* __restrict__ out) {
    __shared__ double sh[kRedThreads];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(long long)b * ncomp + c];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) out[c] = s;
    }
}

// grad[k] = sum_b g[b][k] conj(x[b][k])
__global__ void grad_kernel(const float2* __restrict__ g, const float2* __restrict__ x, long long nw, int nb,
                            float2* __restrict__ grad) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nw; k += (long long)gridDim.x * blockDim.x) {
        float re = 0.f, im = 0.f;
        for (int b = 0; b < nb; ++b) {
            const float2 a = g[b * nw + k], c = x[b * nw + k];
            re += a.x * c.x + a.y * c.y;
            im += a.y * c.x - a.x * c.y;
        }
        grad[k] = make_float2(re, im);
    }
}

// dst in the weight interchange layout ([kz][ky][kx] interleaved re, im) =
// scale / den[0] * src (spectrum layout)
__global__ void to_interchange_kernel(const float2* __restrict__ src, float2* __restrict__ dst, SpecDims sd,
                                      long long nw, const double* __restrict__ den) {
    const float sc = den ? float(1.0 / den[0]) : 1.f;
    for (long long a = blockIdx.x * (long long)blockDim.x + threadIdx.x; a < nw; a += (long long)gridDim.x * blockDim.x) {
        const int kx = int(a % sd.kxn);
        const long long t = a / sd.kxn;
        const int ky = int(t % sd.py), kz = int(t / sd.py);
        const float2 v = src[((long long)kx * sd.py + ky) * sd.pz + kz];
        dst[a] = make_float2(sc * v.x, sc * v.y);
    }
}

// out = src * scale / den[0] (den NULL: 1)
__global__ void scale_kernel(float* __restrict__ out, const float* __restrict__ src, long long n, float scale,
                             const double* __restrict__ den) {
    const float s = den ? float(double(scale) / den[0]) : scale;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        out[i] = src[i] * s;
}

// h[k] += sum_b |x[b][k]|^2
__global__ void power_accum_kernel(const float2* __restrict__ x, long long nw, int nb, double* __restrict__ h) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nw; k += (long long)gridDim.x * blockDim.x) {
        double a = 0.0;
        for (int b = 0; b < nb; ++b) {
            const float2 v = x[b * nw + k];
            a += double(v.x) * v.x + double(v.y) * v.y;
        }
        h[k] += a;
    }
}

// part[blk] = sum of v over the block's share
__global__ void sum_kernel(const double* __restrict__ v, long long n, double* __restrict__ part) {
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        s += v[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// pk[k] = 1 / (sym(h)[k] + floor * mean(h)), sym = average over the mirror
// images; *total = sum_k h[k] (so mean = total / nw)
__global__ void precond_kernel(const double* __restrict__ h, long long nw, SpecDims sd, const double* __restrict__ total,
                               double floor_rel, float* __restrict__ pk) {
    const double fl = floor_rel * total[0] / double(nw);
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nw; k += (long long)gridDim.x * blockDim.x) {
        long long img[4];
        spec_images(k, sd, img);
        const double hs = 0.25 * (h[img[0]] + h[img[1]] + h[img[2]] + h[img[3]]);
        pk[k] = float(1.0 / (hs + fl));
    }
}

__device__ __forceinline__ float sym_re(const float2* __restrict__ g, long long k, SpecDims sd) {
    long long img[4];
    spec_images(k, sd, img);
    return 0.25f * ((g[img[0]].x + g[img[1]].x) + (g[img[2]].x + g[img[3]].x));
}

// out[k] = pk[k] * (Re g averaged over the mirror images), 0): the projection
// onto real weights even in ky and kz, preconditioned (pk NULL: 1)
__global__ void sym_kernel(const float2* __restrict__ g, float2* __restrict__ out, long long nw, SpecDims sd,
                           const float* __restrict__ pk) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nw; k += (long long)gridDim.x * blockDim.x) {
        float v = sym_re(g, k, sd);
        if (pk) v *= pk[k];
        out[k] = make_float2(v, 0.f);
    }
}

// preconditioned heavy-ball step on the symmetric weights:
// mom = mu mom + pk sym(grad) / den; w -= lr mom; loss[it] = red[0] / red[1]
__global__ void sgd_sym_kernel(float2* __restrict__ w, float* __restrict__ mom, const float2* __restrict__ grad,
                               const float* __restrict__ pk, long long nw, SpecDims sd, const double* __restrict__ red,
                               float lr, float mu, float* __restrict__ loss) {
    const float inv = float(1.0 / red[1]);
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nw; k += (long long)gridDim.x * blockDim.x) {
        const float m = mu * mom[k] + pk[k] * sym_re(grad, k, sd) * inv;
        mom[k] = m;
        w[k] = make_float2(w[k].x - lr * m, 0.f);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss) *loss = float(red[0] / red[1]);
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__global__ void gate_site_kernel(int32_t* site, int n, long long V, unsigned long long seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) site[i] = int32_t(splitmix64(seed * 0x100000001b3ull + (unsigned long long)i) % (unsigned long long)V);
}

// Occupancy grid words and the bad-input flag cleared in one launch.
__global__ void occ_clear_kernel(uint32_t* __restrict__ occ, long long words, int* __restrict__ bad) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *bad = 0;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < words; i += (long long)gridDim.x * blockDim.x)
        occ[i] = 0u;
}

// occ_scatter_kernel for a batch of charge lists, plus the (x, y, z, q)
// coordinates of every charge for the exact pair sums (the work of
// to_float4_kernel) in the same pass: charge i belongs to list b with
// off[b] <= i < off[b+1] and goes into grid occ + b * V.
__global__ void occ_scatter_batch_kernel(const int32_t* __restrict__ site, const int32_t* __restrict__ q,
                                         const int* __restrict__ off, int batch, lmg::Dims d, int8_t* occ, int* bad,
                                         float4* __restrict__ chg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= off[batch]) return;
    int lo = 0, hi = batch;   // the last b with off[b] <= i
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    const int s = site[i];
    const int qi = q[i];
    if (s < 0 || s >= d.V || qi < -127 || qi > 127) {
        atomicOr(bad, 1);
        chg[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    int x, y, z;
    lmg::decode(d, s, x, y, z);
    chg[i] = make_float4(float(x), float(y), float(z), float(qi));
    // grid lo starts at byte lo V, which need not be 4-byte aligned: the word
    // holding the site's byte is found from its absolute offset
    const long long a = (long long)lo * d.V + s;
    unsigned int* w = reinterpret_cast<unsigned int*>(occ + (a & ~3ll));
    const int sh = int(a & 3) * 8;
    unsigned int old = *w, assumed;
    do {
        assumed = old;
        const int b = int(int8_t((assumed >> sh) & 0xffu)) + qi;
        const unsigned int nv = (assumed & ~(0xffu << sh)) | (unsigned(uint8_t(int8_t(b))) << sh);
        old = atomicCAS(w, assumed, nv);
    } while (old != assumed);
}

// The gate in one launch: the exact potential at each gate site (the S chunk
// partials of pair_kernel summed in FP64 like pair_reduce_kernel, 16 lanes per
// site), its squared difference from the map, block sums, and in the last block
// to finish the total over blocks in block order: err = sqrt(sum d^2 / sum e^2)
// and *run = err > tol.  The finishing counter returns to zero for the next call.
// Map b = blockIdx.y of a batch uses part + b S nt, map + b V, bpart + 2 b
// gridDim.x and entry b of counter, err and run.
constexpr int kGateLanes = 16;
__global__ __launch_bounds__(256) void gate_reduce_kernel(const float2* __restrict__ part, int nt, int S, double scale,
                                                          const float* __restrict__ map, long long V,
                                                          const int32_t* __restrict__ site, double* __restrict__ bpart,
                                                          unsigned int* __restrict__ counter, double tol,
                                                          double* __restrict__ err, int32_t* __restrict__ run) {
    __shared__ double sh[256];
    __shared__ bool last;
    {
        const int b = blockIdx.y;
        part += (long long)b * S * nt;
        map += (long long)b * V;
        bpart += 2ll * b * gridDim.x;
        counter += b;
        err += b;
        run += b;
    }
    const int j = threadIdx.x % kGateLanes;
    const int t = blockIdx.x * (256 / kGateLanes) + threadIdx.x / kGateLanes;
    double v = 0.0;
    if (t < nt)
        for (int c = j; c < S; c += kGateLanes) {
            const float2 p = part[(long long)c * nt + t];
            v += double(p.x) + double(p.y);
        }
    sh[threadIdx.x] = v;
    __syncthreads();
    double dd = 0.0, ee = 0.0;
    if (j == 0 && t < nt) {
        double a = 0.0;
        for (int i = 0; i < kGateLanes; ++i) a += sh[threadIdx.x + i];
        const double e = scale * a, d = double(map[site[t]]) - e;
        dd = d * d;
        ee = e * e;
    }
    __syncthreads();
    dd = block_sum(dd, sh);
    ee = block_sum(ee, sh);
    if (threadIdx.x == 0) {
        bpart[2 * blockIdx.x] = dd;
        bpart[2 * blockIdx.x + 1] = ee;
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double a = 0.0, b = 0.0;
    const volatile double* vp = bpart;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
        a += vp[2 * i];
        b += vp[2 * i + 1];
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        const double e = b > 0 ? sqrt(a / b) : (a > 0 ? INFINITY : 0.0);
        *err = e;
        *run = (e <= tol) ? 0 : 1;
        *counter = 0u;
    }
}

int grid_for(long long n) { return int(std::min<long long>((n + 255) / 256, 4096)); }

}  // namespace

struct lmg_surrogate {
    lmg_ctx* ctx = nullptr;
    lmg_params p{};
    int flags = 0;
    int dev = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int nx = 0, ny = 0, nz = 0, px = 0, py = 0, pz = 0, kx = 0;
    long long V = 0, nw = 0;       // lattice sites; complex weights
    int cap = 0;                   // batch capacity
    // DFT matrices: forward x, y, z; inverse z, y, x; adjoint of the inverse x
    lmg::ConstOperand c_x{}, c_y{}, c_z{}, c_zi{}, c_yi{}, c_xi{}, c_xadj{};
    float *w = nullptr, *mom = nullptr, *grad = nullptr, *vec = nullptr;
    float* pk = nullptr;           // preconditioner (nw)
    double* hpow = nullptr;        // spectral power accumulator (nw)
    int8_t* occ = nullptr;
    float *s1 = nullptr, *s2 = nullptr, *s3 = nullptr, *y3 = nullptr, *map = nullptr, *tgt = nullptr;
    double* part = nullptr;        // kRedBlocks * 2 partial sums
    double* red = nullptr;         // reduction results (8)
    int* bad = nullptr;
    float* loss = nullptr;
    int loss_cap = 0;
    // error gate and exact fallback (own buffers, so a captured graph never
    // points into memory another call may reallocate)
    lmg::Dims dims{};
    int32_t* gate_site = nullptr;  // sampled sites
    float4* gate_tgt = nullptr;    // their coordinates
    float2* gate_part = nullptr;   // pair-sum partials: S x gate sites per map of the batch
    long long gate_part_cap = 0;   // float2 entries
    double* gate_bpart = nullptr;  // gate_reduce_kernel's block sums, 2 per block per map
    long long gate_bpart_cap = 0;
    int gate_cap = 0;
    int gate_n = 0;                // sites currently in gate_site, and their seed
    uint64_t gate_seed = 0;
    float4* chg = nullptr;         // charges as (x, y, z, q), the batch's lists one after another
    long long chg_cap = 0;
    int* off = nullptr;            // device: list offsets of the batch (max_batch + 1)
    std::vector<int> off_host;     // what off holds
    std::vector<int32_t> mode_host;   // what map_mode holds
    unsigned int* gate_counter = nullptr;   // per map: blocks finished in gate_reduce_kernel (0 between calls)
    double* gate_err = nullptr;    // device, per map: latest gate error
    int32_t* gate_run = nullptr;   // device, per map: latest decision (1 = exact map used)
    // the culled exact map of the fallback, for maps where the context's map
    // would cull: brick counts (zero between calls), offsets, per-charge arrays
    lmg::MapBricks mc{};           // counts and offsets for max_batch maps
    long long mc_cap = 0;          // charges the per-charge arrays hold
    int32_t* map_mode = nullptr;   // device, per map: 1 = culled exact map, 0 = all pairs (from off_host)
    // the gated map as a replayable graph, valid while its key holds
    bool use_graphs = true;
    hipGraphExec_t gexec = nullptr;
    unsigned buf_gen = 0;          // bumped whenever a gate/charge buffer is reallocated
    struct Key {
        const void *site, *q, *map;
        std::vector<int> off;
        int gate_sites;
        double tol;
        unsigned gen;
        bool operator==(const Key& o) const {
            return site == o.site && q == o.q && map == o.map && off == o.off && gate_sites == o.gate_sites &&
                   tol == o.tol && gen == o.gen;
        }
    } gkey{};
    std::vector<void*> allocs;
};

namespace {

lmg_status sfail(lmg_surrogate* s, lmg_status code, const std::string& msg) {
    if (s) s->err = msg;
    else g_surr_create_error = msg;
    return code;
}

#define SHIP(s, call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return sfail((s), LMG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

lmg_status salloc(lmg_surrogate* s, void** p, size_t bytes) {
    SHIP(s, hipMalloc(p, std::max<size_t>(bytes, 256)));
    s->allocs.push_back(*p);
    return LMG_OK;
}

// At least N + c - 1 sites (2N - 1 without a cutoff), so the circular
// convolution equals the aperiodic one; x rounded up to 6 mod 8 (Px / 2 + 1 a
// multiple of 4), y and z to a multiple of 4, for the 16-byte stores of the
// transposed stage outputs.
int padded(int n, int cutoff, bool none, bool x_axis) {
    if (n <= 1) return 1;
    const long long m = none ? 2LL * n - 1 : std::min<long long>(2LL * n - 1, (long long)n + cutoff - 1);
    return int(x_axis ? m + ((6 - m) % 8 + 8) % 8 : m + ((-m) % 4 + 4) % 4);
}

lmg::OutLayout row_major(int ldc) { return lmg::OutLayout{0, ldc, 1, 0, 1, 0, 0, 0}; }

// C[row][col] stored at (row / inner) s_ro + row % inner + (col % cmod) s_cm + (col / cmod) s_cd
lmg::OutLayout transposed(int inner, int s_ro, int cmod, int s_cm, int s_cd, long long s_batch) {
    const bool v4 = inner % 4 == 0 && s_ro % 4 == 0 && s_cm % 4 == 0 && s_cd % 4 == 0 && s_batch % 4 == 0;
    return lmg::OutLayout{1, 0, inner, s_ro, cmod, s_cm, s_cd, v4 ? 1 : 0};
}

// C[b] (M x N) = D[b] (M x K, k contiguous) * Bc, Bc stored transposed
template <typename TD>
void gemm(hipStream_t st, const TD* d, int ldd, long long sd, const lmg::ConstOperand& c, float* out, long long sc,
          int M, int N, int K, int batch, const lmg::OutLayout& o, const float* w = nullptr, int ldw = 0) {
    lmg::GemmArgs g{d, w, ldw, c, out, M, N, K, ldd, sd, sc, 1.f, o};
    dim3 grid((N + lmg::kTN - 1) / lmg::kTN, (M + lmg::kTM - 1) / lmg::kTM, batch);
    if (w)
        hipLaunchKernelGGL((lmg::gemm_bf16x3_kernel<TD, true>), grid, dim3(256), 0, st, g);
    else
        hipLaunchKernelGGL((lmg::gemm_bf16x3_kernel<TD, false>), grid, dim3(256), 0, st, g);
}

// Stage buffers per sample (split complex: a row holds [Re ... | Im ...]):
//   s1: T1 [kx][z][ri][y]  (x-stage out, y-stage in), later T6 [z][y][ri][kx]
//   s2: T2 [kx][ky][ri][z] (y-stage out, z-stage in), later T5 [z][kx][ri][ky]
//   spectrum [kx][ky][kz] complex (re, im)
long long n1_of(const lmg_surrogate* s) { return (long long)s->kx * s->nz * 2 * s->ny; }
long long n2_of(const lmg_surrogate* s) { return (long long)s->kx * s->py * 2 * s->nz; }

// rfft3 of nb occupancy grids (int8) or real grids (float) into spec; the
// x-stage matrix is the forward DFT for the forward pass and the adjoint of the
// inverse x-stage for the backward pass.
template <typename TA>
void transform_in(lmg_surrogate* s, const TA* a, const lmg::ConstOperand& xmat, int nb, float* spec) {
    const int nx = s->nx, ny = s->ny, nz = s->nz, py = s->py, pz = s->pz, kx = s->kx;
    const long long n1 = n1_of(s), n2 = n2_of(s);
    // x: rows (z, y), k = x -> T1 [kx][z][ri][y]
    gemm<TA>(s->stream, a, nx, s->V, xmat, s->s1, n1, nz * ny, 2 * kx, nx, nb,
             transposed(ny, 2 * ny, kx, nz * 2 * ny, ny, n1));
    // y: rows (kx, z), k = (ri, y) -> T2 [kx][ky][ri][z]
    gemm<float>(s->stream, s->s1, 2 * ny, n1, s->c_y, s->s2, n2, kx * nz, 2 * py, 2 * ny, nb,
                transposed(nz, py * 2 * nz, py, 2 * nz, nz, n2));
    // z: rows (kx, ky), k = (ri, z) -> spectrum [kx][ky][kz] (re, im)
    gemm<float>(s->stream, s->s2, 2 * nz, n2, s->c_z, spec, 2 * s->nw, kx * py, 2 * pz, 2 * nz, nb, row_major(2 * pz));
}

// inverse: nb spectra, each multiplied by the weights w on load -> nb maps (the
// window of the padded grid)
void transform_out(lmg_surrogate* s, const float* spec, const float* w, int nb, float* out) {
    const int nx = s->nx, ny = s->ny, nz = s->nz, py = s->py, pz = s->pz, kx = s->kx;
    const long long n1 = n1_of(s), n2 = n2_of(s);
    // inverse z of W . X: rows (kx, ky), k = (kz, re/im) -> T5 [z][kx][ri][ky]
    gemm<float>(s->stream, spec, 2 * pz, 2 * s->nw, s->c_zi, s->s2, n2, kx * py, 2 * nz, 2 * pz, nb,
                transposed(py, 2 * py, nz, kx * 2 * py, py, n2), w, 2 * pz);
    // inverse y: rows (z, kx), k = (ri, ky) -> T6 [z][y][ri][kx]
    gemm<float>(s->stream, s->s2, 2 * py, n2, s->c_yi, s->s1, n1, nz * kx, 2 * ny, 2 * py, nb,
                transposed(kx, ny * 2 * kx, ny, 2 * kx, kx, n1));
    // inverse x: rows (z, y), k = (ri, kx) -> map [z][y][x]
    gemm<float>(s->stream, s->s1, 2 * kx, n1, s->c_xi, out, s->V, nz * ny, nx, 2 * kx, nb, row_major(nx));
}

void forward(lmg_surrogate* s, const float* w, const int8_t* occ, int nb, float* out) {
    transform_in<int8_t>(s, occ, s->c_x, nb, s->s3);
    transform_out(s, s->s3, w, nb, out);
}

// gradient of sum_b |map_b - t_b|^2 (unnormalised) from g = 2 (map - t) held in
// s->map, using the spectra s->s3 of the preceding forward pass
void backward(lmg_surrogate* s, int nb) {
    transform_in<float>(s, s->map, s->c_xadj, nb, s->y3);
    hipLaunchKernelGGL(grad_kernel, dim3(grid_for(s->nw)), dim3(256), 0, s->stream,
                       reinterpret_cast<const float2*>(s->y3), reinterpret_cast<const float2*>(s->s3), s->nw, nb,
                       reinterpret_cast<float2*>(s->grad));
}

// forward + residual + loss sums (red[0] = sum r^2, red[1] = sum t^2) + backward
void loss_grad(lmg_surrogate* s, const float* w, const int8_t* occ, const float* tgt, int nb) {
    forward(s, w, occ, nb, s->map);
    hipLaunchKernelGGL(residual_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s->stream, s->map, tgt,
                       s->V * nb, s->part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kRedThreads), 0, s->stream, s->part, kRedBlocks, 2, s->red);
    backward(s, nb);
}

lmg_status sum_squares(lmg_surrogate* s, const float* v, long long n, double* host) {
    hipLaunchKernelGGL(sumsq_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s->stream, v, n, s->part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kRedThreads), 0, s->stream, s->part, kRedBlocks, 1, s->red + 4);
    SHIP(s, hipMemcpyAsync(host, s->red + 4, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    return LMG_OK;
}

lmg_status use_dev(lmg_surrogate* s) {
    SHIP(s, hipSetDevice(s->dev));
    return LMG_OK;
}

lmg_status launch_ok(lmg_surrogate* s, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return sfail(s, LMG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LMG_OK;
}

// Dense matrix in its stored orientation ([row][k], k contiguous), FP64.
struct HostMat {
    int rows = 0, k = 0;
    std::vector<double> v;
    HostMat(int r, int kk) : rows(r), k(kk), v(size_t(r) * kk, 0.0) {}
    double& at(int r, int kk) { return v[size_t(r) * k + kk]; }
};

// The DFT matrices, each stored as Bc^T [output column][k] (angles reduced
// exactly in integers, theta = 2 pi (o i mod P) / P):
//   x    [ri kx][x]          forward x: Re = cos, Im = -sin (real input)
//   y, z [ri ky][ri' y]      complex e^{-i theta} acting on split-complex rows
//                            (z writes interleaved (re, im) pairs: the spectrum)
//   zi, yi                   complex e^{+i theta}, spatial outputs (zi reads
//                            the interleaved spectrum)
//   xi   [x][ri kx]          inverse x: c_k (cos, -sin) / (Px Py Pz), c_k = 1 at
//                            kx = 0 and Px / 2, else 2 (Hermitian half spectrum)
//   xadj [ri kx][x]          transpose of xi (its adjoint)
void build_matrices(const lmg_surrogate* s, std::vector<HostMat>& out) {
    const double two_pi = 2.0 * M_PI;
    auto ang = [&](long long k, long long x, int p) { return two_pi * double((k * x) % p) / double(p); };
    const int kx = s->kx, nx = s->nx, ny = s->ny, nz = s->nz, px = s->px, py = s->py, pz = s->pz;
    const double inv_total = 1.0 / (double(px) * py * pz);
    HostMat fx(2 * kx, nx), xi(nx, 2 * kx), xadj(2 * kx, nx);
    for (int x = 0; x < nx; ++x)
        for (int k = 0; k < kx; ++k) {
            const double a = ang(k, x, px);
            const double ck = (k == 0 || (px % 2 == 0 && k == px / 2)) ? 1.0 : 2.0;
            fx.at(k, x) = std::cos(a);
            fx.at(kx + k, x) = -std::sin(a);
            const double hc = ck * std::cos(a) * inv_total, hs = -ck * std::sin(a) * inv_total;
            xi.at(x, k) = hc;
            xi.at(x, kx + k) = hs;
            xadj.at(k, x) = hc;
            xadj.at(kx + k, x) = hs;
        }
    // complex block acting on split [Re i | Im i] or interleaved (re, im) pairs
    // of the input, giving split or interleaved outputs, o < no, i < ni
    auto block = [&](int no, int ni, int p, double sign, bool inter_in, bool inter_out) {
        HostMat m(2 * no, 2 * ni);
        auto io = [&](int ri, int o) { return inter_out ? 2 * o + ri : ri * no + o; };
        auto ii = [&](int ri, int i) { return inter_in ? 2 * i + ri : ri * ni + i; };
        for (int o = 0; o < no; ++o)
            for (int i = 0; i < ni; ++i) {
                const double a = ang(o, i, p), fr = std::cos(a), fi = sign * std::sin(a);
                m.at(io(0, o), ii(0, i)) = fr;    // Re out <- Re in
                m.at(io(0, o), ii(1, i)) = -fi;   // Re out <- Im in
                m.at(io(1, o), ii(0, i)) = fi;    // Im out <- Re in
                m.at(io(1, o), ii(1, i)) = fr;    // Im out <- Im in
            }
        return m;
    };
    out.clear();
    out.push_back(fx);
    out.push_back(block(py, ny, py, -1.0, false, false));
    out.push_back(block(pz, nz, pz, -1.0, false, true));    // into the interleaved spectrum
    out.push_back(block(nz, pz, pz, +1.0, true, false));    // out of it
    out.push_back(block(ny, py, py, +1.0, false, false));
    out.push_back(xi);
    out.push_back(xadj);
}

uint16_t host_bf16(double x) {   // round to nearest even, from the FP32 value
    float f = float(x);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return uint16_t((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

double host_bf16_value(uint16_t b) {
    const uint32_t u = uint32_t(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// pad rows to a multiple of 64 and k to a multiple of 32, split into hi/lo
lmg_status upload_const(lmg_surrogate* s, const HostMat& m, lmg::ConstOperand* c) {
    const int rows = (m.rows + 63) / 64 * 64, ld = (m.k + 31) / 32 * 32;
    std::vector<uint16_t> hi(size_t(rows) * ld, 0), lo(size_t(rows) * ld, 0);
    for (int r = 0; r < m.rows; ++r)
        for (int k = 0; k < m.k; ++k) {
            const double x = m.v[size_t(r) * m.k + k];
            const uint16_t h = host_bf16(x);
            hi[size_t(r) * ld + k] = h;
            lo[size_t(r) * ld + k] = host_bf16(x - host_bf16_value(h));
        }
    void* dh = nullptr;
    void* dl = nullptr;
    lmg_status st;
    if ((st = salloc(s, &dh, hi.size() * 2)) || (st = salloc(s, &dl, lo.size() * 2))) return st;
    SHIP(s, hipMemcpy(dh, hi.data(), hi.size() * 2, hipMemcpyHostToDevice));
    SHIP(s, hipMemcpy(dl, lo.data(), lo.size() * 2, hipMemcpyHostToDevice));
    *c = lmg::ConstOperand{static_cast<const uint16_t*>(dh), static_cast<const uint16_t*>(dl), ld};
    return LMG_OK;
}

// Weights between the interchange layout of the C-ABI ([kz][ky][kx] interleaved
// re, im) and the spectrum layout ([kx][ky][kz] interleaved re, im).
void interchange(const lmg_surrogate* s, float* spectrum, float* api, bool to_spectrum) {
    for (int kz = 0; kz < s->pz; ++kz)
        for (int ky = 0; ky < s->py; ++ky)
            for (int kx = 0; kx < s->kx; ++kx) {
                const size_t a = ((size_t(kz) * s->py + ky) * s->kx + kx) * 2;
                const size_t r = ((size_t(kx) * s->py + ky) * s->pz + kz) * 2;
                if (to_spectrum) {
                    spectrum[r] = api[a];
                    spectrum[r + 1] = api[a + 1];
                } else {
                    api[a] = spectrum[r];
                    api[a + 1] = spectrum[r + 1];
                }
            }
}

uint64_t host_splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

lmg_status lmg_surrogate_create(lmg_ctx* ctx, int32_t flags, int32_t max_batch, lmg_surrogate** out) {
    if (!out) return sfail(nullptr, LMG_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!ctx) return sfail(nullptr, LMG_ERR_ARG, "ctx is NULL");
    if (flags & ~(LMG_GAUSS | LMG_NO_CUTOFF)) return sfail(nullptr, LMG_ERR_ARG, "unsupported flags");
    if (max_batch < 1 || max_batch > 4096) return sfail(nullptr, LMG_ERR_ARG, "max_batch must be in [1, 4096]");
    lmg_params p;
    if (lmg_get_params(ctx, &p) != LMG_OK) return sfail(nullptr, LMG_ERR_ARG, "invalid ctx");
    if ((flags & LMG_GAUSS) && !(p.gaussian_sigma > 0))
        return sfail(nullptr, LMG_ERR_ARG, "LMG_GAUSS needs gaussian_sigma > 0");
    auto* s = new lmg_surrogate();
    s->ctx = ctx;
    s->p = p;
    s->flags = flags;
    s->dev = p.device_id;
    s->stream = static_cast<hipStream_t>(lmg_stream(ctx));
    s->nx = p.grid_x;
    s->ny = p.grid_y;
    s->nz = p.grid_z;
    const bool none = flags & LMG_NO_CUTOFF;
    s->px = padded(s->nx, p.cutoff, none, true);
    s->py = padded(s->ny, p.cutoff, none, false);
    s->pz = padded(s->nz, p.cutoff, none, false);
    s->kx = s->px / 2 + 1;
    s->V = (long long)s->nx * s->ny * s->nz;
    s->nw = (long long)s->pz * s->py * s->kx;
    s->dims = lmg::make_dims(s->nx, s->ny, s->nz);
    s->cap = max_batch;
    const long long n1 = n1_of(s), n2 = n2_of(s);
    const long long max_rows = std::max({(long long)s->nz * s->ny, (long long)s->kx * s->nz, (long long)s->kx * s->py});
    if (s->nw * 2 * max_batch > (1LL << 31) || s->V * max_batch > (1LL << 31) ||
        (max_rows + lmg::kTM - 1) / lmg::kTM > 65535 || max_batch > 65535) {
        delete s;
        return sfail(nullptr, LMG_ERR_CAPACITY, "lattice x max_batch too large for one surrogate");
    }
    lmg_status st = LMG_OK;
    auto fail_create = [&](lmg_status code) {
        g_surr_create_error = s->err;
        lmg_surrogate_destroy(s);
        return code;
    };
    if (hipSetDevice(s->dev) != hipSuccess) return fail_create(sfail(s, LMG_ERR_NO_DEVICE, "hipSetDevice failed"));
    {
        std::vector<HostMat> mats;
        build_matrices(s, mats);
        lmg::ConstOperand* dst[7] = {&s->c_x, &s->c_y, &s->c_z, &s->c_zi, &s->c_yi, &s->c_xi, &s->c_xadj};
        for (int i = 0; i < 7; ++i)
            if ((st = upload_const(s, mats[i], dst[i])) != LMG_OK) return fail_create(st);
    }
    const size_t wbytes = size_t(s->nw) * 2 * sizeof(float);
    const size_t occ_bytes = (size_t(s->V) * max_batch + 3) / 4 * 4;
    if ((st = salloc(s, (void**)&s->w, wbytes)) || (st = salloc(s, (void**)&s->mom, wbytes)) ||
        (st = salloc(s, (void**)&s->grad, wbytes)) || (st = salloc(s, (void**)&s->vec, wbytes)) ||
        (st = salloc(s, (void**)&s->pk, size_t(s->nw) * sizeof(float))) ||
        (st = salloc(s, (void**)&s->hpow, size_t(s->nw) * sizeof(double))) ||
        (st = salloc(s, (void**)&s->occ, occ_bytes)) ||
        (st = salloc(s, (void**)&s->s1, size_t(n1) * max_batch * 4)) ||
        (st = salloc(s, (void**)&s->s2, size_t(n2) * max_batch * 4)) ||
        (st = salloc(s, (void**)&s->s3, wbytes * max_batch)) || (st = salloc(s, (void**)&s->y3, wbytes * max_batch)) ||
        (st = salloc(s, (void**)&s->map, size_t(s->V) * max_batch * 4)) ||
        (st = salloc(s, (void**)&s->tgt, size_t(s->V) * max_batch * 4)) ||
        (st = salloc(s, (void**)&s->part, kRedBlocks * 2 * sizeof(double))) ||
        (st = salloc(s, (void**)&s->red, 8 * sizeof(double))) || (st = salloc(s, (void**)&s->bad, sizeof(int))) ||
        (st = salloc(s, (void**)&s->gate_err, size_t(max_batch) * sizeof(double))) ||
        (st = salloc(s, (void**)&s->gate_counter, size_t(max_batch) * sizeof(unsigned int))) ||
        (st = salloc(s, (void**)&s->gate_run, size_t(max_batch) * sizeof(int32_t))) ||
        (st = salloc(s, (void**)&s->off, size_t(max_batch + 1) * sizeof(int))))
        return fail_create(st);
    s->mc = lmg::map_bricks_shape(s->dims);
    if ((st = salloc(s, (void**)&s->mc.cnt, size_t(s->mc.nb) * max_batch * 4)) ||
        (st = salloc(s, (void**)&s->mc.start, size_t(s->mc.nb + 1) * max_batch * 4)) ||
        (st = salloc(s, (void**)&s->map_mode, size_t(max_batch) * 4)))
        return fail_create(st);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for((long long)s->mc.nb * max_batch)), dim3(256), 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->mc.cnt), (long long)s->mc.nb * max_batch, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(s->nw * 2)), dim3(256), 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->w), s->nw * 2, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(s->nw * 2)), dim3(256), 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->mom), s->nw * 2, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s->stream, s->gate_counter, (long long)max_batch, 0u);
    if ((st = launch_ok(s, "surrogate init")) != LMG_OK) return fail_create(st);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail_create(sfail(s, LMG_ERR_HIP, "init sync"));
    *out = s;
    return LMG_OK;
}

void lmg_surrogate_destroy(lmg_surrogate* s) {
    if (!s) return;
    if (hipSetDevice(s->dev) == hipSuccess) {
        (void)hipStreamSynchronize(s->stream);
        if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
        for (void* p : s->allocs) (void)hipFree(p);
    }
    delete s;
}

lmg_status lmg_surrogate_set_graphs(lmg_surrogate* s, int32_t on) {
    if (!s) return LMG_ERR_ARG;
    s->use_graphs = on != 0;
    return LMG_OK;
}

const char* lmg_surrogate_last_error(const lmg_surrogate* s) {
    return s ? s->err.c_str() : g_surr_create_error.c_str();
}

lmg_status lmg_surrogate_shape(const lmg_surrogate* s, int32_t* padded_xyz, int64_t* n_weights) {
    if (!s || !padded_xyz || !n_weights) return LMG_ERR_ARG;
    padded_xyz[0] = s->px;
    padded_xyz[1] = s->py;
    padded_xyz[2] = s->pz;
    *n_weights = s->nw;
    return LMG_OK;
}

lmg_status lmg_surrogate_get_weights(lmg_surrogate* s, float* w) {
    if (!s) return LMG_ERR_ARG;
    if (!w) return sfail(s, LMG_ERR_ARG, "weights pointer is NULL");
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    std::vector<float> in(size_t(s->nw) * 2);
    SHIP(s, hipMemcpyAsync(in.data(), s->w, in.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    interchange(s, in.data(), w, false);
    return LMG_OK;
}

lmg_status lmg_surrogate_set_weights(lmg_surrogate* s, const float* w) {
    if (!s) return LMG_ERR_ARG;
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    if (w) {
        std::vector<float> in(size_t(s->nw) * 2);
        interchange(s, in.data(), const_cast<float*>(w), true);
        SHIP(s, hipMemcpyAsync(s->w, in.data(), in.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
        SHIP(s, hipStreamSynchronize(s->stream));
    } else {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(s->nw * 2)), dim3(256), 0, s->stream,
                           reinterpret_cast<uint32_t*>(s->w), s->nw * 2, 0u);
    }
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(s->nw * 2)), dim3(256), 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->mom), s->nw * 2, 0u);
    SHIP(s, hipStreamSynchronize(s->stream));
    return launch_ok(s, "set_weights");
}

lmg_status lmg_surrogate_occupancy_dev(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q, int32_t n,
                                       int8_t* d_occ) {
    if (!s) return LMG_ERR_ARG;
    if (n < 0 || !d_occ || (n > 0 && (!d_site || !d_q))) return sfail(s, LMG_ERR_ARG, "invalid occupancy arguments");
    if (reinterpret_cast<uintptr_t>(d_occ) % 4) return sfail(s, LMG_ERR_ARG, "occupancy grid must be 4-byte aligned");
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    // the grid holds ceil(V / 4) whole words (the scatter updates bytes through their words)
    const long long words = (s->V + 3) / 4;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(words)), dim3(256), 0, s->stream,
                       reinterpret_cast<uint32_t*>(d_occ), words, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s->stream, reinterpret_cast<uint32_t*>(s->bad), 1LL, 0u);
    if (n > 0)
        hipLaunchKernelGGL(occ_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s->stream, d_site, d_q, n, s->V,
                           d_occ, s->bad);
    return launch_ok(s, "occupancy");
}

lmg_status lmg_surrogate_infer_dev(lmg_surrogate* s, const int8_t* d_occ, int32_t batch, float* d_map) {
    if (!s) return LMG_ERR_ARG;
    if (batch < 1 || batch > s->cap || !d_occ || !d_map) return sfail(s, LMG_ERR_ARG, "invalid infer arguments");
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    forward(s, s->w, d_occ, batch, d_map);
    return launch_ok(s, "infer");
}

lmg_status lmg_surrogate_gradient_dev(lmg_surrogate* s, const int8_t* d_occ, const float* d_target, int32_t batch,
                                      float* d_grad, double* loss) {
    if (!s) return LMG_ERR_ARG;
    if (batch < 1 || batch > s->cap || !d_occ || !d_target || !d_grad)
        return sfail(s, LMG_ERR_ARG, "invalid gradient arguments");
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    loss_grad(s, s->w, d_occ, d_target, batch);
    const SpecDims sd{s->py, s->pz, s->kx};
    hipLaunchKernelGGL(to_interchange_kernel, dim3(grid_for(s->nw)), dim3(256), 0, s->stream,
                       reinterpret_cast<const float2*>(s->grad), reinterpret_cast<float2*>(d_grad), sd, s->nw,
                       s->red + 1);
    if ((st = launch_ok(s, "gradient")) != LMG_OK) return st;
    double r[2];
    SHIP(s, hipMemcpyAsync(r, s->red, sizeof(r), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    if (loss) *loss = r[1] > 0 ? r[0] / r[1] : r[0];
    return LMG_OK;
}

lmg_status lmg_surrogate_train_dev(lmg_surrogate* s, const int8_t* d_occ, const float* d_target, int32_t n_snapshots,
                                   int32_t batch, int32_t iterations, double momentum, uint64_t seed,
                                   float* loss_history) {
    if (!s) return LMG_ERR_ARG;
    if (!d_occ || !d_target || n_snapshots < 1 || batch < 1 || batch > s->cap || batch > n_snapshots ||
        iterations < 0 || !(momentum >= 0.0 && momentum < 1.0))
        return sfail(s, LMG_ERR_ARG, "invalid train arguments");
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    if (iterations > s->loss_cap) {
        if (s->loss) {
            (void)hipFree(s->loss);
            s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), static_cast<void*>(s->loss)));
            s->loss = nullptr;
        }
        if ((st = salloc(s, (void**)&s->loss, size_t(iterations) * sizeof(float))) != LMG_OK) return st;
        s->loss_cap = iterations;
    }
    uint64_t rs = seed;
    std::vector<int> perm(n_snapshots);
    for (int i = 0; i < n_snapshots; ++i) perm[i] = i;
    auto gather = [&]() -> lmg_status {   // a random minibatch (without replacement) into occ/tgt
        for (int j = 0; j < batch; ++j) {
            const int r = j + int(host_splitmix(rs) % uint64_t(n_snapshots - j));
            std::swap(perm[j], perm[r]);
            const long long src = perm[j];
            SHIP(s, hipMemcpyAsync(s->occ + j * s->V, d_occ + src * s->V, size_t(s->V), hipMemcpyDeviceToDevice,
                                   s->stream));
            SHIP(s, hipMemcpyAsync(s->tgt + j * s->V, d_target + src * s->V, size_t(s->V) * 4,
                                   hipMemcpyDeviceToDevice, s->stream));
        }
        return LMG_OK;
    };
    const SpecDims sd{s->py, s->pz, s->kx};
    const long long nw = s->nw;
    const dim3 gw(grid_for(nw)), b256(256);
    // preconditioner: inverse of the mean spectral power of the training
    // occupancies (first kPrecondSnapshots), averaged over the mirror images
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(nw * 2)), b256, 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->hpow), nw * 2, 0u);
    const int n_pre = std::min(n_snapshots, kPrecondSnapshots);
    for (int b0 = 0; b0 < n_pre; b0 += s->cap) {
        const int nb = std::min(s->cap, n_pre - b0);
        SHIP(s, hipMemcpyAsync(s->occ, d_occ + (long long)b0 * s->V, size_t(s->V) * nb, hipMemcpyDeviceToDevice,
                               s->stream));
        transform_in<int8_t>(s, s->occ, s->c_x, nb, s->s3);
        hipLaunchKernelGGL(power_accum_kernel, gw, b256, 0, s->stream, reinterpret_cast<const float2*>(s->s3), nw, nb,
                           s->hpow);
    }
    hipLaunchKernelGGL(sum_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s->stream, s->hpow, nw, s->part);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(kRedThreads), 0, s->stream, s->part, kRedBlocks, 1, s->red + 5);
    hipLaunchKernelGGL(precond_kernel, gw, b256, 0, s->stream, s->hpow, nw, sd, s->red + 5, kPrecondFloor, s->pk);
    // curvature of the preconditioned loss: power iteration on P H restricted to
    // the symmetric weights, H v = gradient at weights v with zero targets
    // scaled by the batch's target norm
    if ((st = gather()) != LMG_OK) return st;
    double den = 0.0;
    if ((st = sum_squares(s, s->tgt, s->V * batch, &den)) != LMG_OK) return st;
    if (!(den > 0)) return sfail(s, LMG_ERR_ARG, "training targets are all zero");
    {
        std::vector<float> v(size_t(nw) * 2);
        uint64_t vs = seed ^ 0x5bd1e995ull;
        for (auto& e : v) e = float(double(host_splitmix(vs) >> 11) * 0x1.0p-53 - 0.5);
        SHIP(s, hipMemcpyAsync(s->grad, v.data(), v.size() * 4, hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(sym_kernel, gw, b256, 0, s->stream, reinterpret_cast<const float2*>(s->grad),
                           reinterpret_cast<float2*>(s->mom), nw, sd, nullptr);
        double v2 = 0.0;
        if ((st = sum_squares(s, s->mom, nw * 2, &v2)) != LMG_OK) return st;
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(nw * 2)), b256, 0, s->stream, s->vec, s->mom, nw * 2,
                           float(1.0 / std::sqrt(v2)), nullptr);
    }
    double lambda = 0.0;
    for (int it = 0; it < kPowerIters; ++it) {
        loss_grad(s, s->vec, s->occ, nullptr, batch);
        hipLaunchKernelGGL(sym_kernel, gw, b256, 0, s->stream, reinterpret_cast<const float2*>(s->grad),
                           reinterpret_cast<float2*>(s->mom), nw, sd, s->pk);
        double g2 = 0.0;
        if ((st = sum_squares(s, s->mom, nw * 2, &g2)) != LMG_OK) return st;
        if (!(g2 > 0)) return sfail(s, LMG_ERR_STATE, "zero curvature: the training occupancies are empty");
        lambda = std::sqrt(g2) / den;   // |P H v| with |v| = 1
        hipLaunchKernelGGL(scale_kernel, dim3(grid_for(nw * 2)), b256, 0, s->stream, s->vec, s->mom, nw * 2,
                           float(1.0 / std::sqrt(g2)), nullptr);
    }
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(nw * 2)), b256, 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->mom), nw * 2, 0u);
    const float lr = float(kStepFactor * (1.0 - momentum) / lambda);
    const float mu = float(momentum);
    for (int it = 0; it < iterations; ++it) {
        if (it > 0 && (st = gather()) != LMG_OK) return st;
        loss_grad(s, s->w, s->occ, s->tgt, batch);
        hipLaunchKernelGGL(sgd_sym_kernel, gw, b256, 0, s->stream, reinterpret_cast<float2*>(s->w), s->mom,
                           reinterpret_cast<const float2*>(s->grad), s->pk, nw, sd, s->red, lr, mu, s->loss + it);
    }
    if ((st = launch_ok(s, "train")) != LMG_OK) return st;
    if (loss_history && iterations > 0)
        SHIP(s, hipMemcpyAsync(loss_history, s->loss, size_t(iterations) * 4, hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    return LMG_OK;
}

namespace {

// whether the exact map of a list of n charges culls, as the context's own map decides
bool map_culled(const lmg_surrogate* s, int n) {
    return !(s->flags & LMG_NO_CUTOFF) && lmg::map_cull_pays(s->dims, s->p.cutoff, n);
}

// Every launch of a batch of gated maps, on s->stream (also what a graph
// captures): occupancy grids and charge coordinates, the forward pass, the
// exact pair sums at the gate sites for every map, the gates, and the exact
// map of every map whose gate failed.  Map b's charges are entries
// [off[b], off[b+1]) of the lists; S is the gate's charge split.
void enqueue_gated(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q, int batch, long long total,
                   float* d_map, double tolerance, int gate_sites, int S) {
    const hipStream_t st = s->stream;
    const long long words = (s->V * batch + 3) / 4;
    hipLaunchKernelGGL(occ_clear_kernel, dim3(grid_for(words)), dim3(256), 0, st, reinterpret_cast<uint32_t*>(s->occ),
                       words, s->bad);
    if (total > 0)
        hipLaunchKernelGGL(occ_scatter_batch_kernel, dim3(int((total + 255) / 256)), dim3(256), 0, st, d_site, d_q,
                           s->off, batch, s->dims, s->occ, s->bad, s->chg);
    forward(s, s->w, s->occ, batch, d_map);
    const bool gauss = s->flags & LMG_GAUSS;
    const uint32_t c2m1 = lmg::cutoff_bits(s->p.cutoff, s->flags & LMG_NO_CUTOFF);
    const float gk = lmg::gauss_k(s->p.gaussian_sigma);
    const dim3 pgrid((gate_sites + lmg::kBlock - 1) / lmg::kBlock, S, batch);
    if (gauss)
        hipLaunchKernelGGL((lmg::pair_batch_kernel<true>), pgrid, dim3(lmg::kBlock), 0, st, s->gate_tgt, gate_sites,
                           s->chg, s->off, c2m1, gk, s->gate_part, gate_sites);
    else
        hipLaunchKernelGGL((lmg::pair_batch_kernel<false>), pgrid, dim3(lmg::kBlock), 0, st, s->gate_tgt, gate_sites,
                           s->chg, s->off, c2m1, gk, s->gate_part, gate_sites);
    const int per = 256 / kGateLanes;
    hipLaunchKernelGGL(gate_reduce_kernel, dim3((gate_sites + per - 1) / per, batch), dim3(256), 0, st, s->gate_part,
                       gate_sites, S, s->p.prefactor, d_map, s->V, s->gate_site, s->gate_bpart, s->gate_counter,
                       tolerance, s->gate_err, s->gate_run);
    // exact maps over the surrogate's where the gate failed (decided on the device),
    // each by the kernel the context's own map uses for that list: culled where
    // it pays (map_mode, uploaded with the offsets), else every charge at every site
    int culled = 0, max_n = 0;
    for (int b = 0; b < batch; ++b) {
        const int n = s->off_host[b + 1] - s->off_host[b];
        culled += map_culled(s, n) ? 1 : 0;
        max_n = std::max(max_n, n);
    }
    const float pref = float(s->p.prefactor);
    if (culled > 0)
        lmg::enqueue_map_cull(st, s->dims, s->mc, s->p.cutoff, gk, pref, gauss, d_site, d_q, max_n, batch,
                              lmg::MapBatch{s->off, s->gate_run, s->map_mode}, d_map);
    if (culled == batch) return;
    int tx = 1;
    while (tx < 64 && tx * lmg::kMapT < s->nx) tx <<= 1;
    const int ty = lmg::kBlock / tx;
    const dim3 mgrid((s->nx + tx * lmg::kMapT - 1) / (tx * lmg::kMapT), (s->ny + ty - 1) / ty, s->nz * batch);
    const int32_t* skip = culled > 0 ? s->map_mode : nullptr;
    if (gauss)
        hipLaunchKernelGGL(lmg::map_batch_kernel<true>, mgrid, dim3(lmg::kBlock), 0, st, s->dims, s->chg, s->off, c2m1,
                           gk, pref, tx, s->gate_run, skip, d_map);
    else
        hipLaunchKernelGGL(lmg::map_batch_kernel<false>, mgrid, dim3(lmg::kBlock), 0, st, s->dims, s->chg, s->off,
                           c2m1, gk, pref, tx, s->gate_run, skip, d_map);
}

// The gate's charge split: about 4096 waves over the batch, >= 64 charges per
// chunk of the longest list, at most kGateMaxSplit.
int gate_split(int gate_sites, int max_n, int batch) {
    const long long waves = (long long)((gate_sites + 63) / 64) * batch;
    long long S = (4096 + waves - 1) / waves;
    S = std::min<long long>(S, std::max(1, max_n / 64));
    return int(std::max<long long>(1, std::min<long long>(S, kGateMaxSplit)));
}

lmg_status gate_buffers(lmg_surrogate* s, long long total, int gate_sites, uint64_t gate_seed, int batch, int S) {
    lmg_status st;
    auto regrow = [&](void** p, size_t bytes) -> lmg_status {
        if (*p) {
            (void)hipStreamSynchronize(s->stream);
            (void)hipFree(*p);
            s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), *p));
            *p = nullptr;
        }
        ++s->buf_gen;
        return salloc(s, p, bytes);
    };
    if (gate_sites > s->gate_cap) {
        if ((st = regrow((void**)&s->gate_site, size_t(gate_sites) * 4)) ||
            (st = regrow((void**)&s->gate_tgt, size_t(gate_sites) * sizeof(float4))))
            return st;
        s->gate_cap = gate_sites;
        s->gate_n = 0;
    }
    const long long parts = (long long)S * gate_sites * batch;
    if (parts > s->gate_part_cap) {
        if ((st = regrow((void**)&s->gate_part, size_t(parts) * sizeof(float2))) != LMG_OK) return st;
        s->gate_part_cap = parts;
    }
    const long long bparts = 2ll * ((gate_sites + 256 / kGateLanes - 1) / (256 / kGateLanes)) * batch;
    if (bparts > s->gate_bpart_cap) {
        if ((st = regrow((void**)&s->gate_bpart, size_t(bparts) * sizeof(double))) != LMG_OK) return st;
        s->gate_bpart_cap = bparts;
    }
    if (total > s->mc_cap) {
        const long long cap = std::max(total, 2 * s->mc_cap);
        if ((st = regrow((void**)&s->mc.slot, size_t(cap) * sizeof(int2))) ||
            (st = regrow((void**)&s->mc.tmp, size_t(cap) * sizeof(float4))) ||
            (st = regrow((void**)&s->mc.tmpi, size_t(cap) * sizeof(int))) ||
            (st = regrow((void**)&s->mc.src, size_t(cap) * sizeof(float4))))
            return st;
        s->mc_cap = cap;
    }
    if (total > s->chg_cap) {
        const long long cap = std::max(total, 2 * s->chg_cap);
        if ((st = regrow((void**)&s->chg, size_t(cap) * sizeof(float4))) != LMG_OK) return st;
        s->chg_cap = cap;
    }
    if (s->gate_n != gate_sites || s->gate_seed != gate_seed) {
        hipLaunchKernelGGL(gate_site_kernel, dim3((gate_sites + 255) / 256), dim3(256), 0, s->stream, s->gate_site,
                           gate_sites, s->V, (unsigned long long)gate_seed);
        hipLaunchKernelGGL(lmg::to_float4_kernel, dim3((gate_sites + 255) / 256), dim3(256), 0, s->stream, s->dims,
                           s->gate_site, nullptr, gate_sites, s->gate_tgt, nullptr);
        s->gate_n = gate_sites;
        s->gate_seed = gate_seed;
    }
    return launch_ok(s, "gate set-up");
}

lmg_status read_gates(lmg_surrogate* s, int batch, double* gate_error, int32_t* used_exact) {
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    std::vector<double> e(batch);
    std::vector<int32_t> run(batch);
    int32_t bad = 0;
    SHIP(s, hipMemcpyAsync(e.data(), s->gate_err, batch * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(run.data(), s->gate_run, batch * sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(&bad, s->bad, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    if (bad) return sfail(s, LMG_ERR_ARG, "charge site outside the lattice or |q| > 127");
    if (gate_error) std::copy(e.begin(), e.end(), gate_error);
    if (used_exact) std::copy(run.begin(), run.end(), used_exact);
    return LMG_OK;
}

}  // namespace

lmg_status lmg_surrogate_map_batch_dev(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q,
                                       const int32_t* counts, int32_t batch, float* d_map, double tolerance,
                                       int32_t gate_sites, uint64_t gate_seed, double* gate_errors,
                                       int32_t* used_exact) {
    if (!s) return LMG_ERR_ARG;
    if (batch < 1 || batch > s->cap || !counts || !d_map || !(tolerance >= 0))
        return sfail(s, LMG_ERR_ARG, "invalid map arguments (batch must be in [1, max_batch])");
    if ((long long)s->nz * batch > 65535) return sfail(s, LMG_ERR_CAPACITY, "grid.z x batch above 65535");
    if (gate_sites < 1 || gate_sites > kMaxGateSites)
        return sfail(s, LMG_ERR_ARG, "gate_sites must be in [1, " + std::to_string(kMaxGateSites) + "]");
    std::vector<int> off(batch + 1, 0);
    long long total = 0;
    int max_n = 0;
    for (int b = 0; b < batch; ++b) {
        if (counts[b] < 0) return sfail(s, LMG_ERR_ARG, "negative charge count");
        total += counts[b];
        if (total > INT32_MAX) return sfail(s, LMG_ERR_CAPACITY, "more than 2^31 - 1 charges in a batch");
        off[b + 1] = int(total);
        max_n = std::max(max_n, int(counts[b]));
    }
    if (total > 0 && (!d_site || !d_q)) return sfail(s, LMG_ERR_ARG, "invalid map arguments");
    const int S = gate_split(gate_sites, max_n, batch);
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    if ((st = gate_buffers(s, total, gate_sites, gate_seed, batch, S)) != LMG_OK) return st;
    if (off != s->off_host) {   // the offsets are read by the kernels (and by replays) from device memory
        SHIP(s, hipStreamSynchronize(s->stream));
        s->off_host = off;
        s->mode_host.resize(batch);
        for (int b = 0; b < batch; ++b) s->mode_host[b] = map_culled(s, off[b + 1] - off[b]) ? 1 : 0;
        SHIP(s, hipMemcpyAsync(s->off, s->off_host.data(), (batch + 1) * sizeof(int), hipMemcpyHostToDevice,
                               s->stream));
        SHIP(s, hipMemcpyAsync(s->map_mode, s->mode_host.data(), batch * sizeof(int32_t), hipMemcpyHostToDevice,
                               s->stream));
    }
    if (!s->use_graphs) {
        enqueue_gated(s, d_site, d_q, batch, total, d_map, tolerance, gate_sites, S);
        if ((st = launch_ok(s, "surrogate map")) != LMG_OK) return st;
    } else {
        lmg_surrogate::Key key{d_site, d_q, d_map, off, gate_sites, tolerance, s->buf_gen};
        if (!s->gexec || !(key == s->gkey)) {
            if (s->gexec) {
                SHIP(s, hipStreamSynchronize(s->stream));
                SHIP(s, hipGraphExecDestroy(s->gexec));
                s->gexec = nullptr;
            }
            hipGraph_t graph = nullptr;
            SHIP(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
            enqueue_gated(s, d_site, d_q, batch, total, d_map, tolerance, gate_sites, S);
            const hipError_t cap = hipStreamEndCapture(s->stream, &graph);
            if (cap != hipSuccess) return sfail(s, LMG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(cap));
            const hipError_t inst = hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (inst != hipSuccess) {
                s->gexec = nullptr;
                return sfail(s, LMG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(inst));
            }
            s->gkey = key;
        }
        SHIP(s, hipGraphLaunch(s->gexec, s->stream));
    }
    if (!gate_errors && !used_exact) return LMG_OK;
    return read_gates(s, batch, gate_errors, used_exact);
}

lmg_status lmg_surrogate_map_dev(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q, int32_t n, float* d_map,
                                 double tolerance, int32_t gate_sites, uint64_t gate_seed, double* gate_error,
                                 int32_t* used_exact) {
    if (!s) return LMG_ERR_ARG;
    if (n < 0) return sfail(s, LMG_ERR_ARG, "invalid map arguments");
    return lmg_surrogate_map_batch_dev(s, d_site, d_q, &n, 1, d_map, tolerance, gate_sites, gate_seed, gate_error,
                                       used_exact);
}

lmg_status lmg_surrogate_last_gate(lmg_surrogate* s, double* gate_error, int32_t* used_exact) {
    if (!s) return LMG_ERR_ARG;
    return read_gates(s, 1, gate_error, used_exact);
}

lmg_status lmg_surrogate_last_gates(lmg_surrogate* s, int32_t batch, double* gate_errors, int32_t* used_exact) {
    if (!s) return LMG_ERR_ARG;
    if (batch < 1 || batch > s->cap) return sfail(s, LMG_ERR_ARG, "batch must be in [1, max_batch]");
    return read_gates(s, batch, gate_errors, used_exact);
}

}  // extern "C"
