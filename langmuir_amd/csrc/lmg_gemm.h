// lmg_gemm.h -- batched GEMM on the gfx950 matrix cores for the DFT stages of
// the spectral surrogate:
//
//   C[b] (M x N) = alpha * D[b] (M x K) * Bc (K x N)
//
// D is data (FP32, or int8 lattice occupancy) with k contiguous in every row,
// optionally multiplied on load by a complex row-wise factor W (the spectral
// weights, so the product never goes through HBM); Bc is a constant DFT
// matrix.  Operands are split into BF16 hi + lo parts and multiplied as
// hi*hi + hi*lo + lo*hi with FP32 accumulation (about 2^-16 relative error per
// product).  Bc is stored pre-split in global memory (hi
// and lo BF16 planes) transposed, as Bc^T[n][k], rows padded to a multiple of
// 64 and k to a multiple of 32 with zeros; each wave reads its fragments of it
// straight from global memory (they stay in L2).  The data tiles go through
// LDS, k contiguous, double-buffered.
//
// Output: row-major C[row * ldc + col], or transposed for the next stage of a
// separable transform: C at
//   (row / inner) * s_ro + row % inner + (col % cmod) * s_cm + (col / cmod) * s_cd
// so that consecutive rows (the next transform axis) land in consecutive
// floats; four rows of a lane go out as one 16-byte store when inner and the
// strides are multiples of 4.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lmg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kTM = 64, kTN = 64, kTK = 32;   // block tile
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

struct OutLayout {
    int trans;                      // 0: row-major with ldc; 1: transposed as above
    int ldc;
    int inner, s_ro, cmod, s_cm, s_cd;
    int vec4;                       // transposed stores may use float4
};

struct GemmArgs {
    const void* D;      // float, or int8_t
    const float* W;     // kScale: D rows times W rows, complex, on interleaved (re, im) pairs
    int ldw;
    ConstOperand c;
    float* C;
    int M, N, K;
    int ldd;
    long long sD, sC;   // batch strides (elements); batch = gridDim.z
    float alpha;
    OutLayout out;
};

// 64 rows x 32 k per tile: thread t loads pairs p = t + 256 e (row p / 16, k pair p % 16).
// With kScale the pair (k, k + 1) is a complex number, multiplied by W's.
template <typename TD, bool kScale>
__device__ __forceinline__ void load_tile(const GemmArgs& g, const TD* __restrict__ D, int m0, int k0,
                                          float2 (&r)[4]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int p = t + 256 * e, row = p >> 4, k = k0 + 2 * (p & 15), m = m0 + row;
        float2 v = make_float2(0.f, 0.f);
        if (m < g.M) {
            const TD* src = D + (long long)m * g.ldd + k;
            if (k + 1 < g.K) v = make_float2(float(src[0]), float(src[1]));
            else if (k < g.K) v.x = float(src[0]);
            if (kScale && k < g.K) {
                const float2 w = *reinterpret_cast<const float2*>(g.W + (long long)m * g.ldw + k);
                v = make_float2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            }
        }
        r[e] = v;
    }
}

__device__ __forceinline__ void store_tile(const float2 (&r)[4], uint16_t (*hi)[kLds], uint16_t (*lo)[kLds]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int p = t + 256 * e, row = p >> 4, kk = 2 * (p & 15);
        uint32_t h, l;
        split2(r[e].x, r[e].y, h, l);
        *reinterpret_cast<uint32_t*>(&hi[row][kk]) = h;
        *reinterpret_cast<uint32_t*>(&lo[row][kk]) = l;
    }
}

struct ConstFrag {
    bf16x8_t hi[2], lo[2];   // the tile's two k-steps of 16
};

__device__ __forceinline__ void load_const(const ConstOperand& c, int row, int k0, ConstFrag& f) {
    const int lane = threadIdx.x & 63;
    const long long base = (long long)row * c.ld + k0 + 8 * (lane >> 5);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        f.hi[s] = *reinterpret_cast<const bf16x8_t*>(c.hi + base + 16 * s);
        f.lo[s] = *reinterpret_cast<const bf16x8_t*>(c.lo + base + 16 * s);
    }
}

__device__ __forceinline__ long long out_index(const OutLayout& o, int row, int col) {
    if (!o.trans) return (long long)row * o.ldc + col;
    return (long long)(row / o.inner) * o.s_ro + row % o.inner + (long long)(col % o.cmod) * o.s_cm +
           (long long)(col / o.cmod) * o.s_cd;
}

// Block: 4 waves in 2 x 2, 64 x 64 tile of C; wave w computes the 32 x 32 sub-tile
// (w / 2, w % 2) with __builtin_amdgcn_mfma_f32_32x32x16_bf16.  Lane layout of
// that instruction: A row = lane % 32, k = 8 (lane / 32) + 0..7; B col =
// lane % 32, same k; D element i of lane l is row 8 (i / 4) + 4 (l / 32) + i % 4,
// col l % 32.  An int8_t data operand is exact in BF16: its lo product is skipped.
template <typename TD, bool kScale>
__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(GemmArgs g) {
    constexpr bool kExact = sizeof(TD) == 1;
    __shared__ __attribute__((aligned(16))) uint16_t s_hi[2][64][kLds];
    __shared__ __attribute__((aligned(16))) uint16_t s_lo[2][64][kLds];
    const TD* D = static_cast<const TD*>(g.D) + blockIdx.z * g.sD;
    float* C = g.C + blockIdx.z * g.sC;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int crow = n0 + wn + (lane & 31);   // constant row (an output column)
    const int drow = wm + (lane & 31);        // data row in LDS
    const int kq = 8 * (lane >> 5);

    f32x16_t acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;

    // One K tile at a time: its data and constant fragments are loaded, the
    // data stored to LDS (double-buffered, so one barrier per tile suffices:
    // every wave has finished tile t - 1 before any passes tile t's barrier and
    // overwrites tile t - 1's buffer at t + 1) and multiplied out.  No loads are
    // kept in flight across tiles: that keeps the kernel near 80 VGPRs, and the
    // six waves per SIMD this allows hide the memory latency better than deeper
    // per-wave prefetching did (the stages are short: K = 64 to 320).
    const int tiles = (g.K + kTK - 1) / kTK;
    for (int t = 0; t < tiles; ++t) {
        float2 dr[4];
        ConstFrag cf;
        load_tile<TD, kScale>(g, D, m0, t * kTK, dr);
        load_const(g.c, crow, t * kTK, cf);
        store_tile(dr, s_hi[t & 1], s_lo[t & 1]);
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8_t dh = *reinterpret_cast<const bf16x8_t*>(&s_hi[t & 1][drow][16 * s + kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dh, cf.hi[s], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dh, cf.lo[s], acc, 0, 0, 0);
            if (!kExact) {
                const bf16x8_t dl = *reinterpret_cast<const bf16x8_t*>(&s_lo[t & 1][drow][16 * s + kq]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(dl, cf.hi[s], acc, 0, 0, 0);
            }
        }
    }
    const int col = n0 + wn + (lane & 31);
    if (col >= g.N) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r0 = m0 + wm + 8 * q + 4 * (lane >> 5);
        if (g.out.trans && g.out.vec4 && r0 + 3 < g.M) {
            const float4 v = make_float4(g.alpha * acc[4 * q], g.alpha * acc[4 * q + 1], g.alpha * acc[4 * q + 2],
                                         g.alpha * acc[4 * q + 3]);
            *reinterpret_cast<float4*>(C + out_index(g.out, r0, col)) = v;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (r0 + i < g.M) C[out_index(g.out, r0 + i, col)] = g.alpha * acc[4 * q + i];
        }
    }
}

}  // namespace lmg
