// lmg_pairsum.h -- the exact pair-sum kernels: potential at explicit targets
// (kernel.cl coulomb2/gauss2, Potential::coulomb*/coulombImage*) and at every
// lattice site (kernel.cl coulomb1/gauss1), shared by the context entry points
// and by the surrogate's error gate and fallback.
//
// FP32 terms on integer-exact squared distances (0 < r^2 < c^2 is decided
// exactly), charges staged through LDS as float4 (x, y, z, q) tiles read as
// wave-wide broadcasts, plain FP32 accumulation over 32 charges and Kahan
// compensation across those blocks; target sums split over charge chunks and
// reduced in FP64 in a fixed order (deterministic).
//
// Everything here has internal linkage: each translation unit that includes
// the header gets its own copy of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "lmg_device.h"

namespace lmg {
namespace {

constexpr int kBlock = 256;      // pair / dE / map workgroup
constexpr int kTile = 256;       // charges per LDS tile
constexpr int kBlk = 32;         // terms summed in plain FP32 before a compensated add
constexpr int kMapT = 4;         // lattice sites per map thread (along x)

__device__ __forceinline__ float rsq(float r2) { return __builtin_amdgcn_rsqf(r2); }
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// 0 < r2 < c2 for non-negative floats, as one unsigned compare on the bits
__device__ __forceinline__ bool in_cut(float r2, uint32_t c2bits_m1) {
    return (__float_as_uint(r2) - 1u) < c2bits_m1;
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ===========================================================================
// (site, q) -> float4 (x, y, z, q)
// ===========================================================================
__global__ void to_float4_kernel(Dims d, const int* __restrict__ site, const int* __restrict__ q,
                                 int n, float4* __restrict__ out, const int32_t* __restrict__ run = nullptr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (run && *run == 0)) return;
    int x, y, z;
    decode(d, site[i], x, y, z);
    out[i] = make_float4(float(x), float(y), float(z), q ? float(q[i]) : 0.f);
}

// ===========================================================================
// Pair sums at explicit targets.  grid = (ceil(nt / 256), S); chunk s covers
// charges [s*chunk, (s+1)*chunk).  part[s*ld + t] = (sum, compensation).
// ===========================================================================
template <bool kGauss, bool kImage>
__device__ __forceinline__ void pair_body(const float4* __restrict__ tgt, int nt, const float4* __restrict__ chg, int nc,
                                          uint32_t c2m1, float gk, float2* __restrict__ part, int ld, float4* tile) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int S = gridDim.y;
    const int chunk = (nc + S - 1) / S;
    const int j0 = blockIdx.y * chunk, j1 = min(nc, j0 + chunk);
    const float4 me = t < nt ? tgt[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    float s = 0.f, c = 0.f;
    for (int base = j0; base < j1; base += kTile) {
        const int n = min(kTile, j1 - base);
        __syncthreads();
        if (threadIdx.x < n) tile[threadIdx.x] = chg[base + threadIdx.x];
        __syncthreads();
        for (int k0 = 0; k0 < n; k0 += kBlk) {
            const int k1 = min(n, k0 + kBlk);
            float acc = 0.f;
            for (int k = k0; k < k1; ++k) {
                const float4 q = tile[k];
                const float dx = kImage ? (me.x + q.x + 1.f) : (me.x - q.x);
                const float dy = me.y - q.y, dz = me.z - q.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                const float a = rsq(r2);
                float f = a;
                if (kGauss) f = erff(gk * (r2 * a)) * a;
                acc = fmaf(q.w, in_cut(r2, c2m1) ? f : 0.f, acc);
            }
            kahan_add(s, c, acc);
        }
    }
    if (t < nt) part[blockIdx.y * ld + t] = make_float2(s, -c);
}

template <bool kGauss, bool kImage>
__global__ __launch_bounds__(kBlock) void pair_kernel(const float4* __restrict__ tgt, int nt,
                                                      const float4* __restrict__ chg, int nc,
                                                      uint32_t c2m1, float gk,
                                                      float2* __restrict__ part, int ld) {
    __shared__ float4 tile[kTile];
    pair_body<kGauss, kImage>(tgt, nt, chg, nc, c2m1, gk, part, ld, tile);
}

// The same targets for a batch of charge sets: set b = blockIdx.z holds
// charges [off[b], off[b+1]) and writes its partials at part + b * S * ld.
template <bool kGauss>
__global__ __launch_bounds__(kBlock) void pair_batch_kernel(const float4* __restrict__ tgt, int nt,
                                                            const float4* __restrict__ chg,
                                                            const int* __restrict__ off, uint32_t c2m1, float gk,
                                                            float2* __restrict__ part, int ld) {
    __shared__ float4 tile[kTile];
    const int b = blockIdx.z, o = off[b];
    pair_body<kGauss, false>(tgt, nt, chg + o, off[b + 1] - o, c2m1, gk,
                             part + (long long)b * gridDim.y * ld, ld, tile);
}

// out[t] = scale * sum_s part[s][t]: 16 lanes per target, lane j summing chunks
// s = j, j + 16, ... in order, then the 16 lane sums in lane order (FP64).
constexpr int kRedLanes = 16;
__global__ void pair_reduce_kernel(const float2* __restrict__ part, int nt, int S, int ld,
                                   double scale, double* __restrict__ out) {
    __shared__ double sh[kBlock];
    const int j = threadIdx.x % kRedLanes;
    const int t = blockIdx.x * (kBlock / kRedLanes) + threadIdx.x / kRedLanes;
    double v = 0.0;
    if (t < nt)
        for (int s = j; s < S; s += kRedLanes) {
            const float2 p = part[(long long)s * ld + t];
            v += double(p.x) + double(p.y);
        }
    sh[threadIdx.x] = v;
    __syncthreads();
    if (j == 0 && t < nt) {
        double a = 0.0;
        for (int i = 0; i < kRedLanes; ++i) a += sh[threadIdx.x + i];
        out[t] = scale * a;
    }
}

// ===========================================================================
// Full-lattice map.  A thread owns kMapT consecutive x sites of one (y, z)
// row; a workgroup owns (tx * kMapT) x (256 / tx) sites of one z plane.
// ===========================================================================
template <bool kGauss>
__device__ __forceinline__ void map_body(Dims d, const float4* __restrict__ chg, int nc, uint32_t c2m1, float gk,
                                         float pref, int tx_n, int z, float* __restrict__ out, float4* tile) {
    const int tx = threadIdx.x % tx_n, ty = threadIdx.x / tx_n;
    const int ty_n = kBlock / tx_n;
    const int x0 = (blockIdx.x * tx_n + tx) * kMapT;
    const int y = blockIdx.y * ty_n + ty;
    const float fy = float(y), fz = float(z);
    float s[kMapT], c[kMapT];
#pragma unroll
    for (int u = 0; u < kMapT; ++u) { s[u] = 0.f; c[u] = 0.f; }
    for (int base = 0; base < nc; base += kTile) {
        const int n = min(kTile, nc - base);
        __syncthreads();
        if (threadIdx.x < n) tile[threadIdx.x] = chg[base + threadIdx.x];
        __syncthreads();
        for (int k0 = 0; k0 < n; k0 += kBlk) {
            const int k1 = min(n, k0 + kBlk);
            float acc[kMapT];
#pragma unroll
            for (int u = 0; u < kMapT; ++u) acc[u] = 0.f;
            for (int k = k0; k < k1; ++k) {
                const float4 q = tile[k];
                const float dy = fy - q.y, dz = fz - q.z;
                const float dyz = dy * dy + dz * dz;
                const float dx0 = float(x0) - q.x;
#pragma unroll
                for (int u = 0; u < kMapT; ++u) {
                    const float dx = dx0 + float(u);
                    const float r2 = fmaf(dx, dx, dyz);
                    const float a = rsq(r2);
                    float f = a;
                    if (kGauss) f = erff(gk * (r2 * a)) * a;
                    acc[u] = fmaf(q.w, in_cut(r2, c2m1) ? f : 0.f, acc[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < kMapT; ++u) kahan_add(s[u], c[u], acc[u]);
        }
    }
    if (y < d.Y) {
#pragma unroll
        for (int u = 0; u < kMapT; ++u) {
            const int x = x0 + u;
            if (x < d.X) out[encode(d, x, y, z)] = pref * (s[u] - c[u]);
        }
    }
}

template <bool kGauss>
__global__ __launch_bounds__(kBlock) void map_kernel(Dims d, const float4* __restrict__ chg, int nc,
                                                     uint32_t c2m1, float gk, float pref, int tx_n,
                                                     const int32_t* __restrict__ run, float* __restrict__ out) {
    if (run && *run == 0) return;   // conditional map whose device-side condition is false
    __shared__ float4 tile[kTile];
    map_body<kGauss>(d, chg, nc, c2m1, gk, pref, tx_n, blockIdx.z, out, tile);
}

// A batch of conditional maps: grid z = Z * batch; map b = blockIdx.z / Z of
// charges [off[b], off[b+1]) into out + b * V when run[b] is set and skip
// (if not NULL) has skip[b] == 0.
template <bool kGauss>
__global__ __launch_bounds__(kBlock) void map_batch_kernel(Dims d, const float4* __restrict__ chg,
                                                           const int* __restrict__ off, uint32_t c2m1, float gk,
                                                           float pref, int tx_n, const int32_t* __restrict__ run,
                                                           const int32_t* __restrict__ skip, float* __restrict__ out) {
    const int b = blockIdx.z / d.Z;
    if (run[b] == 0 || (skip && skip[b] != 0)) return;   // gate passed, or map b taken by another kernel
    __shared__ float4 tile[kTile];
    const int o = off[b];
    map_body<kGauss>(d, chg + o, off[b + 1] - o, c2m1, gk, pref, tx_n, blockIdx.z - b * d.Z,
                     out + (long long)b * d.V, tile);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
uint32_t cutoff_bits(int cutoff, bool none) {
    float c2 = none ? INFINITY : float(double(cutoff) * double(cutoff));
    uint32_t b;
    std::memcpy(&b, &c2, 4);
    return b - 1u;
}

float gauss_k(double sigma) { return sigma > 0 ? float(1.0 / (std::sqrt(2.0) * sigma)) : 0.f; }

int split_count(int nt, int nc) {
    // enough waves to cover 256 CUs several times over, >= 64 charges per chunk
    const int waves = std::max(1, (nt + 63) / 64);
    int S = (4096 + waves - 1) / waves;
    S = std::min(S, std::max(1, nc / 64));
    return std::max(1, std::min(S, 128));
}

}  // namespace
}  // namespace lmg
