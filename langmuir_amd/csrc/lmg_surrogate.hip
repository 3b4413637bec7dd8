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

// Mirror images of spectrum index k = (kz Py + ky) Kx + kx under ky -> -ky and
// kz -> -kz (the lattice reflections y -> -y, z -> -z; x -> -x is the complex
// conjugate, which realness of the weights covers).
struct SpecDims {
    int py, pz, kxn;
};
__device__ __forceinline__ void spec_images(long long k, SpecDims sd, long long img[4]) {
    const int kx = int(k % sd.kxn);
    const long long r = k / sd.kxn;
    const int ky = int(r % sd.py), kz = int(r / sd.py);
    const int my = ky ? sd.py - ky : 0, mz = kz ? sd.pz - kz : 0;
    img[0] = k;
    img[1] = ((long long)kz * sd.py + my) * sd.kxn + kx;
    img[2] = ((long long)mz * sd.py + ky) * sd.kxn + kx;
    img[3] = ((long long)mz * sd.py + my) * sd.kxn + kx;
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

// y[b][k] = w[k] x[b][k] on interleaved complex values
__global__ void spec_mult_kernel(const float2* __restrict__ w, const float2* __restrict__ x, float2* __restrict__ y,
                                 long long nw, int nb) {
    const long long n = nw * nb;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float2 a = w[i % nw], b = x[i];
        y[i] = make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
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

// out[k] = pk[k] * (Re g averaged over the mirror images), 0): the projection
// onto real weights even in ky and kz, preconditioned (pk NULL: 1)
__global__ void sym_kernel(const float2* __restrict__ g, float2* __restrict__ out, long long nw, SpecDims sd,
                           const float* __restrict__ pk) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nw; k += (long long)gridDim.x * blockDim.x) {
        long long img[4];
        spec_images(k, sd, img);
        float v = 0.25f * ((g[img[0]].x + g[img[1]].x) + (g[img[2]].x + g[img[3]].x));
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
        long long img[4];
        spec_images(k, sd, img);
        const float g = 0.25f * ((grad[img[0]].x + grad[img[1]].x) + (grad[img[2]].x + grad[img[3]].x));
        const float m = mu * mom[k] + pk[k] * g * inv;
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

// part[2 blk] = sum (map[site] - exact)^2, part[2 blk + 1] = sum exact^2 over
// the block's grid-stride share of the gate sites
__global__ void gate_part_kernel(const float* __restrict__ map, const int32_t* __restrict__ site,
                                 const double* __restrict__ exact, int n, double* __restrict__ part) {
    __shared__ double sh[kRedThreads];
    double a = 0.0, b = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double e = exact[i];
        const double d = double(map[site[i]]) - e;
        a += d * d;
        b += e * e;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

// err = sqrt(sum d^2 / sum e^2) over the gate sites; *run = err > tol (the
// exact map must replace the surrogate's)
__global__ void gate_final_kernel(const double* __restrict__ part, int nblk, double tol, double* __restrict__ err,
                                  int32_t* __restrict__ run) {
    __shared__ double sh[kRedThreads];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        const double e = b > 0 ? sqrt(a / b) : (a > 0 ? INFINITY : 0.0);
        *err = e;
        *run = (e <= tol) ? 0 : 1;
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
    lmg::ConstOperand c_fx{}, c_fy{}, c_fz{}, c_gz{}, c_gy{}, c_hx{}, c_hxt{};   // DFT matrices
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
    double* gate_exact = nullptr;  // exact potential there
    float2* gate_part = nullptr;   // pair-sum partials, kGateMaxSplit x gate_cap
    int gate_cap = 0;
    int gate_n = 0;                // sites currently in gate_site, and their seed
    uint64_t gate_seed = 0;
    float4* chg = nullptr;         // charges as (x, y, z, q)
    int chg_cap = 0;
    double* gate_err = nullptr;    // device: latest gate error
    int32_t* gate_run = nullptr;   // device: latest decision (1 = exact map used)
    // the gated map as a replayable graph, valid while its key holds
    bool use_graphs = true;
    hipGraphExec_t gexec = nullptr;
    unsigned buf_gen = 0;          // bumped whenever a gate/charge buffer is reallocated
    struct Key {
        const void *site, *q, *map;
        int n, gate_sites;
        double tol;
        unsigned gen;
        bool operator==(const Key& o) const {
            return site == o.site && q == o.q && map == o.map && n == o.n && gate_sites == o.gate_sites &&
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

int padded(int n, int cutoff, bool none) {
    if (n <= 1) return 1;
    long long m = none ? 2LL * n - 1 : std::min<long long>(2LL * n - 1, (long long)n + cutoff - 1);
    return int(m + (m & 1));
}

// C (M x N) = D (M x K) * Bc: data rows with k contiguous, constant Bc stored [n][k]
template <typename TD>
void gemm_data_a(hipStream_t st, const TD* d, int ldd, const lmg::ConstOperand& c, float* out, int ldc, int M, int N,
                 int K) {
    lmg::GemmArgs g{d, c, out, M, N, K, K, ldd, ldc, 0, 0, 1.f};
    dim3 grid((N + lmg::kTN - 1) / lmg::kTN, (M + lmg::kTM - 1) / lmg::kTM, 1);
    hipLaunchKernelGGL((lmg::gemm_bf16x3_kernel<0, TD, false>), grid, dim3(256), 0, st, g);
}

// C[b] (M x N) = Ac (M x 2 Kb) * [[D[b]], [i D[b]]]: complex left-multiplication
// of interleaved complex columns, constant Ac = [Re F | Im F] stored [m][k]
void gemm_cplx_left(hipStream_t st, const lmg::ConstOperand& c, const float* d, int ldd, long long sd, float* out,
                    int ldc, long long sc, int M, int N, int Kb, int batch) {
    lmg::GemmArgs g{d, c, out, M, N, 2 * Kb, Kb, ldd, ldc, sd, sc, 1.f};
    dim3 grid((N + lmg::kTN - 1) / lmg::kTN, (M + lmg::kTM - 1) / lmg::kTM, batch);
    hipLaunchKernelGGL((lmg::gemm_bf16x3_kernel<1, float, true>), grid, dim3(256), 0, st, g);
}

// rfft3 of nb occupancy grids (int8) or real grids (float) into spec; the
// x-stage matrix is the forward DFT for the forward pass and the transpose of
// the inverse x-stage for the adjoint pass.
template <typename TA>
void transform_in(lmg_surrogate* s, const TA* a, const lmg::ConstOperand& xmat, int nb, float* spec) {
    const int two_kx = 2 * s->kx;
    const long long row_y = (long long)s->py * two_kx;
    // x: [b z y][x] -> [b z y][kx]
    gemm_data_a<TA>(s->stream, a, s->nx, xmat, s->s1, two_kx, nb * s->nz * s->ny, two_kx, s->nx);
    // y: per (b, z), [y][kx] -> [ky][kx]
    gemm_cplx_left(s->stream, s->c_fy, s->s1, two_kx, (long long)s->ny * two_kx, s->s2, two_kx, row_y, s->py, two_kx,
                   s->ny, nb * s->nz);
    // z: per b, [z][ky kx] -> [kz][ky kx]
    gemm_cplx_left(s->stream, s->c_fz, s->s2, int(row_y), s->nz * row_y, spec, int(row_y), 2 * s->nw, s->pz,
                   int(row_y), s->nz, nb);
}

// inverse: spectrum y3 [b][kz][ky][kx] -> nb maps (window of the padded grid)
void transform_out(lmg_surrogate* s, int nb, float* out) {
    const int two_kx = 2 * s->kx;
    const long long row_y = (long long)s->py * two_kx;
    gemm_cplx_left(s->stream, s->c_gz, s->y3, int(row_y), 2 * s->nw, s->s2, int(row_y), s->nz * row_y, s->nz,
                   int(row_y), s->pz, nb);
    gemm_cplx_left(s->stream, s->c_gy, s->s2, two_kx, row_y, s->s1, two_kx, (long long)s->ny * two_kx, s->ny, two_kx,
                   s->py, nb * s->nz);
    gemm_data_a<float>(s->stream, s->s1, two_kx, s->c_hx, out, s->nx, nb * s->nz * s->ny, s->nx, two_kx);
}

void forward(lmg_surrogate* s, const float* w, const int8_t* occ, int nb, float* out) {
    transform_in<int8_t>(s, occ, s->c_fx, nb, s->s3);
    hipLaunchKernelGGL(spec_mult_kernel, dim3(grid_for(s->nw * nb)), dim3(256), 0, s->stream,
                       reinterpret_cast<const float2*>(w), reinterpret_cast<const float2*>(s->s3),
                       reinterpret_cast<float2*>(s->y3), s->nw, nb);
    transform_out(s, nb, out);
}

// gradient of sum_b |map_b - t_b|^2 (unnormalised) from g = 2 (map - t) held in
// s->map, using the spectra s->s3 of the preceding forward pass
void backward(lmg_surrogate* s, int nb) {
    transform_in<float>(s, s->map, s->c_hxt, nb, s->y3);
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

// The DFT matrices of the six stages (angles reduced exactly in integers):
//   fx  [c][x]        forward x (real -> interleaved half spectrum), Bc^T
//   fy  [ky][y | y]   [cos | -sin] of e^{-2 pi i ky y / Py}; fz likewise
//   gz  [z][kz | kz]  [cos | +sin] of e^{+2 pi i kz z / Pz}; gy likewise
//   hx  [x][c]        inverse x (half spectrum -> real, Hermitian weights c_k,
//                     1 / (Px Py Pz) folded in), Bc^T
//   hxt [c][x]        transpose of the inverse x-stage (its adjoint), Bc^T
void build_matrices(const lmg_surrogate* s, std::vector<HostMat>& out) {
    const double two_pi = 2.0 * M_PI;
    auto ang = [&](long long k, long long x, int p) { return two_pi * double((k * x) % p) / double(p); };
    const int kx = s->kx, nx = s->nx, ny = s->ny, nz = s->nz, px = s->px, py = s->py, pz = s->pz;
    const double inv_total = 1.0 / (double(px) * py * pz);
    HostMat fx(2 * kx, nx), hx(nx, 2 * kx), hxt(2 * kx, nx);
    for (int x = 0; x < nx; ++x)
        for (int k = 0; k < kx; ++k) {
            const double a = ang(k, x, px);
            const double ck = (k == 0 || (px % 2 == 0 && k == px / 2)) ? 1.0 : 2.0;
            fx.at(2 * k, x) = std::cos(a);
            fx.at(2 * k + 1, x) = -std::sin(a);
            const double hc = ck * std::cos(a) * inv_total, hs = -ck * std::sin(a) * inv_total;
            hx.at(x, 2 * k) = hc;
            hx.at(x, 2 * k + 1) = hs;
            hxt.at(2 * k, x) = hc;
            hxt.at(2 * k + 1, x) = hs;
        }
    auto fwd = [&](int p, int n) {   // p x [cos | -sin] (n each)
        HostMat m(p, 2 * n);
        for (int k = 0; k < p; ++k)
            for (int x = 0; x < n; ++x) {
                const double a = ang(k, x, p);
                m.at(k, x) = std::cos(a);
                m.at(k, n + x) = -std::sin(a);
            }
        return m;
    };
    auto inv = [&](int n, int p) {   // n x [cos | +sin] (p each)
        HostMat m(n, 2 * p);
        for (int x = 0; x < n; ++x)
            for (int k = 0; k < p; ++k) {
                const double a = ang(k, x, p);
                m.at(x, k) = std::cos(a);
                m.at(x, p + k) = std::sin(a);
            }
        return m;
    };
    out.clear();
    out.push_back(fx);
    out.push_back(fwd(py, ny));
    out.push_back(fwd(pz, nz));
    out.push_back(inv(nz, pz));
    out.push_back(inv(ny, py));
    out.push_back(hx);
    out.push_back(hxt);
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
    s->px = padded(s->nx, p.cutoff, none);
    s->py = padded(s->ny, p.cutoff, none);
    s->pz = padded(s->nz, p.cutoff, none);
    s->kx = s->px / 2 + 1;
    s->V = (long long)s->nx * s->ny * s->nz;
    s->nw = (long long)s->pz * s->py * s->kx;
    s->dims = lmg::make_dims(s->nx, s->ny, s->nz);
    s->cap = max_batch;
    const long long two_kx = 2 * s->kx;
    const long long n1 = (long long)s->nz * s->ny * two_kx, n2 = (long long)s->nz * s->py * two_kx;
    const long long row_tiles = ((long long)max_batch * s->nz * s->ny + lmg::kTM - 1) / lmg::kTM;
    if (s->nw * 2 * max_batch > (1LL << 31) || s->V * max_batch > (1LL << 31) || row_tiles > 65535 ||
        (long long)max_batch * s->nz > 65535) {
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
        lmg::ConstOperand* dst[7] = {&s->c_fx, &s->c_fy, &s->c_fz, &s->c_gz, &s->c_gy, &s->c_hx, &s->c_hxt};
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
        (st = salloc(s, (void**)&s->gate_err, sizeof(double))) ||
        (st = salloc(s, (void**)&s->gate_run, sizeof(int32_t))))
        return fail_create(st);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(s->nw * 2)), dim3(256), 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->w), s->nw * 2, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(s->nw * 2)), dim3(256), 0, s->stream,
                       reinterpret_cast<uint32_t*>(s->mom), s->nw * 2, 0u);
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
    SHIP(s, hipMemcpyAsync(w, s->w, size_t(s->nw) * 2 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    return LMG_OK;
}

lmg_status lmg_surrogate_set_weights(lmg_surrogate* s, const float* w) {
    if (!s) return LMG_ERR_ARG;
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    if (w) {
        SHIP(s, hipMemcpyAsync(s->w, w, size_t(s->nw) * 2 * sizeof(float), hipMemcpyHostToDevice, s->stream));
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
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(s->nw * 2)), dim3(256), 0, s->stream, d_grad, s->grad, s->nw * 2,
                       1.f, s->red + 1);
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
        transform_in<int8_t>(s, s->occ, s->c_fx, nb, s->s3);
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

// Every launch of one gated map, on s->stream (also what a graph captures).
void enqueue_gated(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q, int n, float* d_map,
                   double tolerance, int gate_sites) {
    const hipStream_t st = s->stream;
    // occupancy grid (slot 0 of s->occ)
    const long long words = (s->V + 3) / 4;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(words)), dim3(256), 0, st, reinterpret_cast<uint32_t*>(s->occ),
                       words, 0u);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<uint32_t*>(s->bad), 1LL, 0u);
    if (n > 0)
        hipLaunchKernelGGL(occ_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_site, d_q, n, s->V, s->occ,
                           s->bad);
    forward(s, s->w, s->occ, 1, d_map);
    // exact pair sums at the gate sites
    const bool gauss = s->flags & LMG_GAUSS;
    const uint32_t c2m1 = lmg::cutoff_bits(s->p.cutoff, s->flags & LMG_NO_CUTOFF);
    const float gk = lmg::gauss_k(s->p.gaussian_sigma);
    if (n > 0)
        hipLaunchKernelGGL(lmg::to_float4_kernel, dim3((n + 255) / 256), dim3(256), 0, st, s->dims, d_site, d_q, n,
                           s->chg, nullptr);
    const int S = lmg::split_count(gate_sites, n);
    const dim3 pgrid((gate_sites + lmg::kBlock - 1) / lmg::kBlock, S);
    if (gauss)
        hipLaunchKernelGGL((lmg::pair_kernel<true, false>), pgrid, dim3(lmg::kBlock), 0, st, s->gate_tgt, gate_sites,
                           s->chg, n, c2m1, gk, s->gate_part, gate_sites);
    else
        hipLaunchKernelGGL((lmg::pair_kernel<false, false>), pgrid, dim3(lmg::kBlock), 0, st, s->gate_tgt, gate_sites,
                           s->chg, n, c2m1, gk, s->gate_part, gate_sites);
    const int per = lmg::kBlock / lmg::kRedLanes;
    hipLaunchKernelGGL(lmg::pair_reduce_kernel, dim3((gate_sites + per - 1) / per), dim3(lmg::kBlock), 0, st,
                       s->gate_part, gate_sites, S, gate_sites, s->p.prefactor, s->gate_exact);
    const int nblk = std::min(kRedBlocks, (gate_sites + kRedThreads - 1) / kRedThreads);
    hipLaunchKernelGGL(gate_part_kernel, dim3(nblk), dim3(kRedThreads), 0, st, d_map, s->gate_site, s->gate_exact,
                       gate_sites, s->part);
    hipLaunchKernelGGL(gate_final_kernel, dim3(1), dim3(kRedThreads), 0, st, s->part, nblk, tolerance, s->gate_err,
                       s->gate_run);
    // exact map over the surrogate's when the gate failed (decided on the device)
    int tx = 1;
    while (tx < 64 && tx * lmg::kMapT < s->nx) tx <<= 1;
    const int ty = lmg::kBlock / tx;
    const dim3 mgrid((s->nx + tx * lmg::kMapT - 1) / (tx * lmg::kMapT), (s->ny + ty - 1) / ty, s->nz);
    const float pref = float(s->p.prefactor);
    if (gauss)
        hipLaunchKernelGGL(lmg::map_kernel<true>, mgrid, dim3(lmg::kBlock), 0, st, s->dims, s->chg, n, c2m1, gk, pref,
                           tx, s->gate_run, d_map);
    else
        hipLaunchKernelGGL(lmg::map_kernel<false>, mgrid, dim3(lmg::kBlock), 0, st, s->dims, s->chg, n, c2m1, gk,
                           pref, tx, s->gate_run, d_map);
}

lmg_status gate_buffers(lmg_surrogate* s, int n, int gate_sites, uint64_t gate_seed) {
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
            (st = regrow((void**)&s->gate_tgt, size_t(gate_sites) * sizeof(float4))) ||
            (st = regrow((void**)&s->gate_exact, size_t(gate_sites) * 8)) ||
            (st = regrow((void**)&s->gate_part, size_t(kGateMaxSplit) * gate_sites * sizeof(float2))))
            return st;
        s->gate_cap = gate_sites;
        s->gate_n = 0;
    }
    if (n > s->chg_cap) {
        const int cap = std::max(n, 2 * s->chg_cap);
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

}  // namespace

lmg_status lmg_surrogate_map_dev(lmg_surrogate* s, const int32_t* d_site, const int32_t* d_q, int32_t n, float* d_map,
                                 double tolerance, int32_t gate_sites, uint64_t gate_seed, double* gate_error,
                                 int32_t* used_exact) {
    if (!s) return LMG_ERR_ARG;
    if (n < 0 || !d_map || (n > 0 && (!d_site || !d_q)) || gate_sites < 1 || !(tolerance >= 0))
        return sfail(s, LMG_ERR_ARG, "invalid map arguments");
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    if ((st = gate_buffers(s, n, gate_sites, gate_seed)) != LMG_OK) return st;
    if (!s->use_graphs) {
        enqueue_gated(s, d_site, d_q, n, d_map, tolerance, gate_sites);
        if ((st = launch_ok(s, "surrogate map")) != LMG_OK) return st;
    } else {
        const lmg_surrogate::Key key{d_site, d_q, d_map, n, gate_sites, tolerance, s->buf_gen};
        if (!s->gexec || !(key == s->gkey)) {
            if (s->gexec) {
                SHIP(s, hipStreamSynchronize(s->stream));
                SHIP(s, hipGraphExecDestroy(s->gexec));
                s->gexec = nullptr;
            }
            hipGraph_t graph = nullptr;
            SHIP(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
            enqueue_gated(s, d_site, d_q, n, d_map, tolerance, gate_sites);
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
    if (!gate_error && !used_exact) return LMG_OK;
    return lmg_surrogate_last_gate(s, gate_error, used_exact);
}

lmg_status lmg_surrogate_last_gate(lmg_surrogate* s, double* gate_error, int32_t* used_exact) {
    if (!s) return LMG_ERR_ARG;
    lmg_status st;
    if ((st = use_dev(s)) != LMG_OK) return st;
    double e = 0.0;
    int32_t run = 0, bad = 0;
    SHIP(s, hipMemcpyAsync(&e, s->gate_err, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(&run, s->gate_run, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipMemcpyAsync(&bad, s->bad, sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    SHIP(s, hipStreamSynchronize(s->stream));
    if (bad) return sfail(s, LMG_ERR_ARG, "charge site outside the lattice or |q| > 127");
    if (gate_error) *gate_error = e;
    if (used_exact) *used_exact = run;
    return LMG_OK;
}

}  // extern "C"
