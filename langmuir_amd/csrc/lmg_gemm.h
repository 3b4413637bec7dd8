// lmg_gemm.h -- batched GEMM on the gfx950 matrix cores for the DFT stages of
// the spectral surrogate: one operand is data (FP32, or int8 lattice
// occupancy) and the other a constant DFT matrix.  Operands are split into
// BF16 hi + lo parts and multiplied as hi*hi + hi*lo + lo*hi with FP32
// accumulation (about 2^-16 relative error per product).
//
//   mode 0 (data = A):  C[b] (M x N) = alpha * D[b] (M x K, k contiguous) * Bc (K x N)
//   mode 1 (data = B):  C[b] (M x N) = alpha * Ac (M x K) * D'[b] (K x N, n contiguous)
//
// In mode 1 with kRot, K = 2 Kb and rows Kb..2Kb-1 of D' are rows 0..Kb-1 of D
// times i on interleaved complex pairs ((re, im) -> (-im, re)), which turns a
// complex left-multiplication into one real GEMM:
//   (Fr + i Fi) X = [Fr | Fi] * [[X], [i X]].
//
// The constant operand is stored pre-split in global memory (hi and lo BF16
// planes), k contiguous: Ac as [m][k] for mode 1, Bc transposed as [n][k] for
// mode 0, rows padded to a multiple of 64 and k to a multiple of 32 with zeros.
// Each wave reads its constant fragments straight from global memory (they stay
// in L2); the data tile goes through LDS, k contiguous, double-buffered.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lmg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kTM = 64, kTN = 64, kTK = 32;   // block tile; a K tile is kSlabs slabs of 32
constexpr int kSlabs = kTK / 32;
constexpr int kLds = kTK + 8;                 // LDS row stride (bf16): 80 bytes, 16-byte aligned rows

// round-to-nearest-even FP32 -> BF16 bits
__device__ __forceinline__ uint32_t bf16_bits(float x) {
    uint32_t u = __float_as_uint(x);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_float(uint32_t b) { return __uint_as_float(b << 16); }

// hi and lo BF16 bits of two values, packed two per 32-bit word
__device__ __forceinline__ void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
    const uint32_t ha = bf16_bits(a), hb = bf16_bits(b);
    const uint32_t la = bf16_bits(a - bf16_float(ha)), lb = bf16_bits(b - bf16_float(hb));
    hi = ha | (hb << 16);
    lo = la | (lb << 16);
}

struct ConstOperand {
    const uint16_t* hi;
    const uint16_t* lo;
    int ld;   // row length (padded K)
};

struct GemmArgs {
    const void* D;     // data operand: float, or int8_t (mode 0 only)
    ConstOperand c;    // constant operand
    float* C;
    int M, N, K, Kb;   // K = 2 Kb with kRot, else K = Kb
    int ldd, ldc;
    long long sD, sC;  // batch strides (elements)
    float alpha;
};

// One data tile's share of a thread: 4 float pairs per slab of 32 k.
struct DataRegs {
    float2 v[kSlabs][4];
};

template <int kMode, typename TD, bool kRot>
__device__ __forceinline__ void load_slab(const GemmArgs& g, const TD* __restrict__ D, int m0, int n0, int k0,
                                          float2 (&r)[4]) {
    const int t = threadIdx.x;
    if (kMode == 0) {
        // 64 rows (m) x 32 k: pair p = t + 256 e -> row p / 16, k pair p % 16
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int p = t + 256 * e, row = p >> 4, k = k0 + 2 * (p & 15), m = m0 + row;
            float2 v = make_float2(0.f, 0.f);
            if (m < g.M) {
                const TD* src = D + (long long)m * g.ldd + k;
                if (k + 1 < g.K) v = make_float2(float(src[0]), float(src[1]));
                else if (k < g.K) v.x = float(src[0]);
            }
            r[e] = v;
        }
    } else {
        // 32 k x 64 n: thread -> column pair t % 32, k row t / 32 + 8 e
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int np = t & 31, kr = (t >> 5) + 8 * e, n = n0 + 2 * np, k = k0 + kr;
            float2 v = make_float2(0.f, 0.f);
            if (k < g.K && n < g.N) {
                if (!kRot || k < g.Kb) {
                    v = *reinterpret_cast<const float2*>(D + (long long)k * g.ldd + n);
                } else {
                    const float2 u = *reinterpret_cast<const float2*>(D + (long long)(k - g.Kb) * g.ldd + n);
                    v = make_float2(-u.y, u.x);
                }
            }
            r[e] = v;
        }
    }
}

template <int kMode, typename TD, bool kRot>
__device__ __forceinline__ void load_data(const GemmArgs& g, const TD* __restrict__ D, int m0, int n0, int k0,
                                          DataRegs& r) {
#pragma unroll
    for (int sl = 0; sl < kSlabs; ++sl)
        if (k0 + 32 * sl < g.K) load_slab<kMode, TD, kRot>(g, D, m0, n0, k0 + 32 * sl, r.v[sl]);
        else
#pragma unroll
            for (int e = 0; e < 4; ++e) r.v[sl][e] = make_float2(0.f, 0.f);
}

// registers -> LDS (hi, lo planes), k contiguous per row; mode 1 transposes
template <int kMode>
__device__ __forceinline__ void store_data(const DataRegs& r, uint16_t (*hi)[kLds], uint16_t (*lo)[kLds]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int sl = 0; sl < kSlabs; ++sl) {
        if (kMode == 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int p = t + 256 * e, row = p >> 4, kk = 32 * sl + 2 * (p & 15);
                uint32_t h, l;
                split2(r.v[sl][e].x, r.v[sl][e].y, h, l);
                *reinterpret_cast<uint32_t*>(&hi[row][kk]) = h;
                *reinterpret_cast<uint32_t*>(&lo[row][kk]) = l;
            }
        } else {
            const int np = t & 31;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int kr = 32 * sl + (t >> 5) + 8 * e;
                uint32_t h, l;
                split2(r.v[sl][e].x, r.v[sl][e].y, h, l);   // (col 2 np, col 2 np + 1) at row kr
                hi[2 * np][kr] = uint16_t(h & 0xffffu);
                hi[2 * np + 1][kr] = uint16_t(h >> 16);
                lo[2 * np][kr] = uint16_t(l & 0xffffu);
                lo[2 * np + 1][kr] = uint16_t(l >> 16);
            }
        }
    }
}

constexpr int kSteps = kTK / 16;   // MFMA k-steps per K tile

struct ConstFrag {
    bf16x8_t hi[kSteps], lo[kSteps];
};

// the tile's constant fragments; steps past the padded row length are zero
__device__ __forceinline__ void load_const(const ConstOperand& c, int row, int k0, ConstFrag& f) {
    const int lane = threadIdx.x & 63;
    const long long base = (long long)row * c.ld + k0 + 8 * (lane >> 5);
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
        if (k0 + 16 * s < c.ld) {
            f.hi[s] = *reinterpret_cast<const bf16x8_t*>(c.hi + base + 16 * s);
            f.lo[s] = *reinterpret_cast<const bf16x8_t*>(c.lo + base + 16 * s);
        } else {
            f.hi[s] = bf16x8_t{};
            f.lo[s] = bf16x8_t{};
        }
    }
}

// Block: 4 waves in 2 x 2, 64 x 64 tile of C; wave w computes the 32 x 32 sub-tile
// (w / 2, w % 2) with __builtin_amdgcn_mfma_f32_32x32x16_bf16.  Lane layout of
// that instruction: A row = lane % 32, k = 8 (lane / 32) + 0..7; B col =
// lane % 32, same k; D element i of lane l is row 8 (i / 4) + 4 (l / 32) + i % 4,
// col l % 32.  An int8_t data operand is exact in BF16: its lo product is skipped.
template <int kMode, typename TD, bool kRot>
__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(GemmArgs g) {
    constexpr bool kExact = sizeof(TD) == 1;
    __shared__ __attribute__((aligned(16))) uint16_t s_hi[2][64][kLds];
    __shared__ __attribute__((aligned(16))) uint16_t s_lo[2][64][kLds];
    const TD* D = static_cast<const TD*>(g.D) + blockIdx.z * g.sD;
    float* C = g.C + blockIdx.z * g.sC;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    // constant rows this wave multiplies: m (mode 1) or n (mode 0)
    const int crow = kMode == 1 ? m0 + wm + (lane & 31) : n0 + wn + (lane & 31);
    // data rows in LDS this wave reads
    const int drow = (kMode == 1 ? wn : wm) + (lane & 31);
    const int kq = 8 * (lane >> 5);

    f32x16_t acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;

    const int tiles = (g.K + kTK - 1) / kTK;
    DataRegs dr;
    ConstFrag cf;
    load_data<kMode, TD, kRot>(g, D, m0, n0, 0, dr);
    store_data<kMode>(dr, s_hi[0], s_lo[0]);
    load_const(g.c, crow, 0, cf);
    __syncthreads();
    for (int t = 0; t < tiles; ++t) {
        const int buf = t & 1;
        const bool more = t + 1 < tiles;
        if (more) load_data<kMode, TD, kRot>(g, D, m0, n0, (t + 1) * kTK, dr);
        const ConstFrag cur = cf;
        if (more) load_const(g.c, crow, (t + 1) * kTK, cf);
#pragma unroll
        for (int s = 0; s < kSteps; ++s) {
            if (t * kTK + 16 * s >= g.K) break;
            const bf16x8_t dh = *reinterpret_cast<const bf16x8_t*>(&s_hi[buf][drow][16 * s + kq]);
            const bf16x8_t dl = *reinterpret_cast<const bf16x8_t*>(&s_lo[buf][drow][16 * s + kq]);
            if (kMode == 0) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dh, cur.hi[s], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dh, cur.lo[s], acc, 0, 0, 0);
                if (!kExact) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dl, cur.hi[s], acc, 0, 0, 0);
            } else {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.hi[s], dh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.lo[s], dh, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(cur.hi[s], dl, acc, 0, 0, 0);
            }
        }
        if (more) store_data<kMode>(dr, s_hi[buf ^ 1], s_lo[buf ^ 1]);
        __syncthreads();
    }
    const int col = n0 + wn + (lane & 31);
    if (col < g.N) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            if (row < g.M) C[(long long)row * g.ldc + col] = g.alpha * acc[i];
        }
    }
}

}  // namespace lmg
