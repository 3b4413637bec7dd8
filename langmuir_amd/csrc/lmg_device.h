// lmg_device.h -- device-side building blocks shared by the Langmuir HIP kernels
// (gfx950, wave64).  Lattice indexing, the MT19937 stream, Boost's uniform_int
// bucket rule and the neighbour order of the reference grid.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lmg {

constexpr int kSpecialSlots = 35;   // Grid::m_specialAgentReserve (cubicgrid.cpp:22)

// Agent::Type (include/agent.h:33-55)
constexpr uint8_t kEmpty = 0, kElectron = 1, kHole = 2, kDefect = 3, kSource = 4, kDrain = 5;

struct Dims {
    int X, Y, Z, XY, V;
    double inv_x, inv_xy;   // 1/X and 1/XY for the division-free decode
};

inline Dims make_dims(int X, int Y, int Z) {
    return Dims{X, Y, Z, X * Y, X * Y * Z, 1.0 / double(X), 1.0 / double(X * Y)};
}

// floor(n / d) for 0 <= n < 2^31: FP64 estimate (off by at most one) + correction
__host__ __device__ inline int div_fast(int n, int d, double inv) {
    int q = int(double(n) * inv);
    int r = n - q * d;
    if (r < 0) --q;
    else if (r >= d) ++q;
    return q;
}

// cubicgrid.cpp:132-145: x = s % X, y = s / X - (s / XY) * Y, z = s / XY.  Also
// used on special slots s >= V, exactly like kernel.cl:229-231.
__host__ __device__ inline void decode(const Dims& d, int s, int& x, int& y, int& z) {
    z = div_fast(s, d.XY, d.inv_xy);
    const int r = s - z * d.XY;          // s/X - z*Y == (s - z*XY)/X for s >= 0
    y = div_fast(r, d.X, d.inv_x);
    x = r - y * d.X;
}
__host__ __device__ inline int encode(const Dims& d, int x, int y, int z) {
    return d.X * (y + z * d.Y) + x;   // cubicgrid.cpp:162-165
}

// ---------------------------------------------------------------------------
// MT19937 as a word sequence: x[n+624] = x[n+397] ^ twist(x[n], x[n+1]);
// engine output k (k = 0, 1, ...) is temper(x[k + 624]).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t mt_temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}
__host__ __device__ inline uint32_t mt_next(uint32_t x0, uint32_t x1, uint32_t x397) {
    uint32_t y = (x0 & 0x80000000u) | (x1 & 0x7fffffffu);
    return x397 ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

// uniform_01<double> on a 32-bit engine: u32 * 2^-32 (rand.cpp:49-52)
__host__ __device__ inline double u01(uint32_t w) { return double(w) * (1.0 / 4294967296.0); }

// Boost >= 1.47 uniform_int bucket size for range = hi - lo in [1, 2^32 - 2]
__host__ __device__ inline uint32_t uint_bucket(uint32_t range) {
    uint32_t b = 0xffffffffu / (range + 1u);
    if (0xffffffffu % (range + 1u) == range) ++b;
    return b;
}

// ---------------------------------------------------------------------------
// Neighbour lists (Grid::neighborsSite, cubicgrid.cpp:207-471).  The order is
// part of the contract: it is indexed by the proposal draw.
// ---------------------------------------------------------------------------
struct Coord {
    int x, y, z;
};

// hopping.range = 1 without building a list: count, and the k-th entry.
__device__ inline int nbr1_count(const Dims& d, int x, int y, int z, int dl, int dr) {
    int n = (x > 0) + (x < d.X - 1) + (y > 0) + (y < d.Y - 1) + (z > 0) + (z < d.Z - 1);
    n += (x == 0 && dl >= 0) + (x == d.X - 1 && dr >= 0);
    return n;
}
__device__ inline int nbr1_at(const Dims& d, int x, int y, int z, int dl, int dr, int k) {
    // candidates in reference order: W E S N down up drainL drainR
    if (x > 0) { if (k == 0) return encode(d, x - 1, y, z); --k; }
    if (x < d.X - 1) { if (k == 0) return encode(d, x + 1, y, z); --k; }
    if (y > 0) { if (k == 0) return encode(d, x, y - 1, z); --k; }
    if (y < d.Y - 1) { if (k == 0) return encode(d, x, y + 1, z); --k; }
    if (z > 0) { if (k == 0) return encode(d, x, y, z - 1); --k; }
    if (z < d.Z - 1) { if (k == 0) return encode(d, x, y, z + 1); --k; }
    if (x == 0 && dl >= 0) { if (k == 0) return dl; --k; }
    return dr;
}

// k-th range-1 neighbour with its coordinates (drain slots decode like the
// reference's kernel.cl:229-231)
__device__ inline int nbr1_pick(const Dims& d, int x, int y, int z, int dl, int dr, int k, int& fx, int& fy,
                                int& fz) {
    fx = x; fy = y; fz = z;
    if (x > 0) { if (k == 0) { fx = x - 1; return encode(d, fx, y, z); } --k; }
    if (x < d.X - 1) { if (k == 0) { fx = x + 1; return encode(d, fx, y, z); } --k; }
    if (y > 0) { if (k == 0) { fy = y - 1; return encode(d, x, fy, z); } --k; }
    if (y < d.Y - 1) { if (k == 0) { fy = y + 1; return encode(d, x, fy, z); } --k; }
    if (z > 0) { if (k == 0) { fz = z - 1; return encode(d, x, y, fz); } --k; }
    if (z < d.Z - 1) { if (k == 0) { fz = z + 1; return encode(d, x, y, fz); } --k; }
    const int s = (x == 0 && dl >= 0 && k == 0) ? dl : dr;
    decode(d, s, fx, fy, fz);
    return s;
}

// Calls emit(site) for every neighbour in list order, hopping.range 1 or 2,
// without materialising the list (no private arrays, hence no scratch).
template <class F>
__device__ inline void nbr_enum(const Dims& d, int h, int site, int dl, int dr, F&& emit) {
    int x, y, z;
    decode(d, site, x, y, z);
#define LMG_P(a, b, c) emit(encode(d, (a), (b), (c)))
    if (h == 1) {
        if (x > 0) LMG_P(x - 1, y, z);
        if (x < d.X - 1) LMG_P(x + 1, y, z);
        if (y > 0) LMG_P(x, y - 1, z);
        if (y < d.Y - 1) LMG_P(x, y + 1, z);
        if (z > 0) LMG_P(x, y, z - 1);
        if (z < d.Z - 1) LMG_P(x, y, z + 1);
    } else if (d.Z == 1) {
        if (x > 0) {
            LMG_P(x - 1, y, z);
            if (y > 0) LMG_P(x - 1, y - 1, z);
            if (y < d.Y - 1) LMG_P(x - 1, y + 1, z);
            if (x > 1) LMG_P(x - 2, y, z);
        }
        if (x < d.X - 1) {
            LMG_P(x + 1, y, z);
            if (y > 0) LMG_P(x + 1, y - 1, z);
            if (y < d.Y - 1) LMG_P(x + 1, y + 1, z);
            if (x < d.X - 2) LMG_P(x + 2, y, z);
        }
        if (y > 0) {
            LMG_P(x, y - 1, z);
            if (y > 1) LMG_P(x, y - 2, z);
        }
        if (y < d.Y - 1) {
            LMG_P(x, y + 1, z);
            if (y < d.Y - 2) LMG_P(x, y + 2, z);
        }
    } else {
        for (int side = -1; side <= 1; side += 2) {
            bool ok = side < 0 ? (x > 0) : (x < d.X - 1);
            if (!ok) continue;
            int xx = x + side;
            LMG_P(xx, y, z);
            if (y > 0) {
                LMG_P(xx, y - 1, z);
                if (z > 0) LMG_P(xx, y - 1, z - 1);
                if (z < d.Z - 1) LMG_P(xx, y - 1, z + 1);
            }
            if (y < d.Y - 1) {
                LMG_P(xx, y + 1, z);
                if (z > 0) LMG_P(xx, y + 1, z - 1);
                if (z < d.Z - 1) LMG_P(xx, y + 1, z + 1);
            }
            if (z > 0) LMG_P(xx, y, z - 1);
            if (z < d.Z - 1) LMG_P(xx, y, z + 1);
            if (side < 0 ? (x > 1) : (x < d.X - 2)) LMG_P(x + 2 * side, y, z);
        }
        if (z > 0) {
            LMG_P(x, y, z - 1);
            if (y > 0) LMG_P(x, y - 1, z - 1);
            if (y < d.Y - 1) LMG_P(x, y + 1, z - 1);
            if (z > 1) LMG_P(x, y, z - 2);
        }
        if (z < d.Z - 1) {
            LMG_P(x, y, z + 1);
            if (y > 0) LMG_P(x, y - 1, z + 1);
            if (y < d.Y - 1) LMG_P(x, y + 1, z + 1);
            if (z < d.Z - 2) LMG_P(x, y, z + 2);
        }
        if (y > 0) {
            LMG_P(x, y - 1, z);
            if (y > 1) LMG_P(x, y - 2, z);
        }
        if (y < d.Y - 1) {
            LMG_P(x, y + 1, z);
            if (y < d.Y - 2) LMG_P(x, y + 2, z);
        }
    }
#undef LMG_P
    if (x == 0 && dl >= 0) emit(dl);        // cubicgrid.cpp:447-457
    if (x == d.X - 1 && dr >= 0) emit(dr);  // cubicgrid.cpp:459-468
}

__device__ inline int nbr_count(const Dims& d, int h, int site, int dl, int dr) {
    int n = 0;
    nbr_enum(d, h, site, dl, dr, [&](int) { ++n; });
    return n;
}

__device__ inline int nbr_at(const Dims& d, int h, int site, int dl, int dr, int k) {
    int n = 0, r = -1;
    nbr_enum(d, h, site, dl, dr, [&](int s) {
        if (n == k) r = s;
        ++n;
    });
    return r;
}

// Wave-wide prefix sums and reductions with DPP moves (no LDS traffic):
// row_shr:1,2,4,8 scan each row of 16 lanes, row_bcast:15 and row_bcast:31
// carry the row totals into the rows above.  Lanes a move has no source for
// (and rows its row mask leaves out) take the identity.
template <class Op>
__device__ __forceinline__ int dpp_scan(int x, int id, Op op) {
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x111, 0xf, 0xf, false));   // row_shr:1
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x112, 0xf, 0xf, false));   // row_shr:2
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x114, 0xf, 0xf, false));   // row_shr:4
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x118, 0xf, 0xf, false));   // row_shr:8
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x142, 0xa, 0xf, false));   // row_bcast:15 -> rows 1, 3
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x143, 0xc, 0xf, false));   // row_bcast:31 -> rows 2, 3
    return x;
}
__device__ __forceinline__ int wave_incl_sum(int x) {
    return dpp_scan(x, 0, [](int a, int b) { return a + b; });
}
// the total, wave-uniform (lane 63 of the inclusive scan)
__device__ __forceinline__ int wave_total(int incl) { return __builtin_amdgcn_readlane(incl, 63); }
__device__ __forceinline__ float wave_fmin(float v) {
    const int r = dpp_scan(__float_as_int(v), __float_as_int(3.0e38f),
                           [](int a, int b) { return __float_as_int(fminf(__int_as_float(a), __int_as_float(b))); });
    return __int_as_float(__builtin_amdgcn_readlane(r, 63));
}
__device__ __forceinline__ float wave_fmax(float v) {
    const int r = dpp_scan(__float_as_int(v), __float_as_int(-3.0e38f),
                           [](int a, int b) { return __float_as_int(fmaxf(__int_as_float(a), __int_as_float(b))); });
    return __int_as_float(__builtin_amdgcn_readlane(r, 63));
}
// integer minimum and maximum over the wave (one DPP min/max per step)
__device__ __forceinline__ int wave_imin(int v) {
    return __builtin_amdgcn_readlane(dpp_scan(v, 0x7fffffff, [](int a, int b) { return min(a, b); }), 63);
}
__device__ __forceinline__ int wave_imax(int v) {
    return __builtin_amdgcn_readlane(dpp_scan(v, int(0x80000000), [](int a, int b) { return max(a, b); }), 63);
}

// A pointer read out of a device struct is a generic (flat) address; flat loads
// count against the LDS/scalar counter as well, so any LDS or scalar wait also
// waits for them.  Loads through gmem() are global loads, tracked by the vector
// memory counter alone.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
// The same for stores and atomics: a flat store holds up every later LDS wait
// (a workgroup barrier included) until it completes; a global one does not.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gst(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

}  // namespace lmg
