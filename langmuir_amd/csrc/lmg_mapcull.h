// lmg_mapcull.h -- the full-lattice potential map with cell culling, for
// cutoff maps on lattices much wider than the cutoff (the coulomb1/gauss1 map
// of kernel.cl:56-220, launchCoulombKernel1 openclhelper.cpp:220-370, with the
// same terms and the same 0 < r^2 < c^2 test as map_kernel).
//
// Per call the charges are bucketed into bricks of 8 x 8 x 8 sites:
//   mapc_count_kernel    each charge's brick and its arrival slot (atomic);
//   mapc_scan_kernel     list offsets;
//   mapc_scatter_kernel  charges to their arrival slots;
//   mapc_sort_kernel     one wave per brick puts its charges in input order
//                        (rank by input index), so that the lists, and every
//                        sum below, are the same bits on every call.
// map_cull_kernel: one wave per tile of sites (4 x 4 x 16 with 4 sites per lane
// along z, 4 x 4 x 4 on lattices thinner than 16, 8 x 8 x 1 on lattices
// thinner than 4) visits the bricks of the tile box widened by the
// cutoff whose distance to the box is below the cutoff, in brick order.  Each
// lane copies one brick's list into the wave's LDS list (64-byte rows); every
// lane then sums the list over broadcast reads, in FP32 per 64-slot unit of a
// brick's list, with Kahan compensation across units.  A charge in a skipped
// brick has r >= c from every site of the tile, so its term is 0.
#pragma once

#include <algorithm>
#include <climits>

#include "lmg_pairsum.h"

namespace lmg {
namespace {

constexpr int kMapBrick = 8;
constexpr int kMapList = 512;   // LDS charge slots per wave
constexpr int kMapUnit = 64;    // FP32 summation unit: 64 slots of one brick's list

struct MapBricks {
    int nbx, nby, nbz, nb;
    int* cnt;       // [nb] charges per brick; zero between calls
    int* start;     // [nb + 1] list offsets
    int2* slot;     // [n] per charge: brick, arrival slot (brick -1: site outside the lattice)
    float4* tmp;    // [n] charges at their arrival slots
    int* tmpi;      // [n] their input indices
    float4* src;    // [n] brick lists in input order
};

__device__ __forceinline__ int map_brick(const MapBricks& m, int x, int y, int z) {
    return (x >> 3) + m.nbx * ((y >> 3) + m.nby * (z >> 3));
}

// A batch of maps: map b = blockIdx.y owns charges [off[b], off[b + 1]) of the
// lists (with off NULL, a single map of n charges), counts and offsets
// [b nb, (b + 1) nb) and [b (nb + 1), ...), and runs only if run[b] != 0 and
// (mode NULL or) mode[b] != 0 -- both decided on the device or uploaded
// beforehand.  On return m and n describe map b alone, whose charges start at
// entry base of the lists.
struct MapBatch {
    const int* off;
    const int32_t* run;
    const int32_t* mode;
};

__device__ __forceinline__ bool map_slice(const MapBatch& mb, MapBricks& m, int& base, int& n) {
    const int b = blockIdx.y;
    if ((mb.run && mb.run[b] == 0) || (mb.mode && mb.mode[b] == 0)) return false;
    base = 0;
    if (mb.off) {
        base = mb.off[b];
        n = mb.off[b + 1] - base;
    }
    m.cnt += b * m.nb;
    m.start += b * (m.nb + 1);
    m.slot += base;
    m.tmp += base;
    m.tmpi += base;
    m.src += base;
    return true;
}

__global__ void mapc_count_kernel(Dims d, MapBricks m, const int* __restrict__ site, int n, MapBatch mb) {
    int base;
    if (!map_slice(mb, m, base, n)) return;
    site += base;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = site[i];
    if (s < 0 || s >= d.V) {   // not a lattice site: left out (the host entry points reject it)
        m.slot[i] = make_int2(-1, 0);
        return;
    }
    int x, y, z;
    decode(d, s, x, y, z);
    const int b = map_brick(m, x, y, z);
    m.slot[i] = make_int2(b, atomicAdd(&m.cnt[b], 1));
}

// one workgroup of 1024: list offsets (thread t scans a contiguous range of
// bricks, its first eight counts loaded together)
__global__ __launch_bounds__(1024) void mapc_scan_kernel(MapBricks m, MapBatch mb) {
    int base, n = 0;
    if (!map_slice(mb, m, base, n)) return;
    __shared__ int wsum[16];
    constexpr int kPer = 8;
    const int nb = m.nb, per = (nb + 1023) / 1024;
    const int b0 = min(nb, int(threadIdx.x) * per), b1 = min(nb, b0 + per);
    int v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) v[u] = m.cnt[min(b0 + u, nb - 1)];
    int r = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        v[u] = b0 + u < b1 ? v[u] : 0;
        r += v[u];
    }
    for (int b = b0 + kPer; b < b1; ++b) r += m.cnt[b];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int incl = wave_incl_sum(r);
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int o = incl - r, total = 0;
    for (int k = 0; k < 16; ++k) {
        if (k < w) o += wsum[k];
        total += wsum[k];
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u)
        if (b0 + u < b1) {
            m.start[b0 + u] = o;
            o += v[u];
        }
    for (int b = b0 + kPer; b < b1; ++b) {
        m.start[b] = o;
        o += m.cnt[b];
    }
    if (threadIdx.x == 0) m.start[nb] = total;
}

__global__ void mapc_scatter_kernel(Dims d, MapBricks m, const int* __restrict__ site, const int* __restrict__ q,
                                    int n, MapBatch mb) {
    int base;
    if (!map_slice(mb, m, base, n)) return;
    site += base;
    if (q) q += base;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 bs = m.slot[i];
    if (bs.x < 0) return;
    const int k = m.start[bs.x] + bs.y;
    int x, y, z;
    decode(d, site[i], x, y, z);
    m.tmp[k] = make_float4(float(x), float(y), float(z), q ? float(q[i]) : 0.f);
    m.tmpi[k] = i;
}

// grid = ceil(nb / 4), one wave per brick
__global__ __launch_bounds__(256) void mapc_sort_kernel(MapBricks m, MapBatch mb) {
    int base, nmap = 0;
    if (!map_slice(mb, m, base, nmap)) return;
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= m.nb) return;
    const int lane = threadIdx.x & 63;
    const int k0 = m.start[b], n = m.cnt[b];
    for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        const int mine = e < n ? m.tmpi[k0 + e] : INT_MAX;
        const float4 val = e < n ? m.tmp[k0 + e] : make_float4(0.f, 0.f, 0.f, 0.f);
        int rank = 0;   // charges of the brick with a smaller input index (indices are distinct)
        for (int f0 = 0; f0 < n; f0 += 64) {
            const int other = f0 + lane < n ? m.tmpi[k0 + f0 + lane] : INT_MAX;
            const int cnt = min(64, n - f0);
            for (int t = 0; t < cnt; ++t) rank += __builtin_amdgcn_readlane(other, t) < mine ? 1 : 0;
        }
        if (e < n) m.src[k0 + rank] = val;
    }
    if (lane == 0) m.cnt[b] = 0;   // every lane has read it: zero again for the next call
}

// Tile shapes: kZc = 1: 4 x 4 x 4 sites (flat = 0) or 8 x 8 x 1 (flat = 1);
// kZc = 4: 4 x 4 x 16, each lane owning a column of 4 sites along z (the x and
// y differences of a charge are shared by them).
__host__ __device__ inline void map_tile_shape(int kzc, int flat, int& tsx, int& tsz) {
    tsx = flat ? 8 : 4;
    tsz = flat ? 1 : 4 * kzc;
}

// grid = ceil(tiles / 4), one wave per tile
template <bool kGauss, int kZc>
__global__ __launch_bounds__(256) void map_cull_kernel(Dims d, MapBricks m, int cut, uint32_t c2m1, float gk,
                                                       float pref, int flat, MapBatch mb, float* __restrict__ out) {
    {
        int base, n = 0;
        if (!map_slice(mb, m, base, n)) return;
        out += (long long)blockIdx.y * d.V;
    }
    __shared__ float4 lists[4][kMapList];
    __shared__ int uends[4][kMapList];   // a unit holds at least one slot
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int tsx, tsz;
    map_tile_shape(kZc, flat, tsx, tsz);
    const int ntx = (d.X + tsx - 1) / tsx, nty = (d.Y + tsx - 1) / tsx, ntz = (d.Z + tsz - 1) / tsz;
    const int tile = blockIdx.x * 4 + wv;
    if (tile >= ntx * nty * ntz) return;   // whole wave; nothing below synchronises the workgroup
    const int tx = tile % ntx, tyz = tile / ntx, ty = tyz % nty, tz = tyz / nty;
    const int x0 = tx * tsx, y0 = ty * tsx, z0 = tz * tsz;
    const int x1 = min(d.X - 1, x0 + tsx - 1), y1 = min(d.Y - 1, y0 + tsx - 1), z1 = min(d.Z - 1, z0 + tsz - 1);
    const int x = x0 + (flat ? (lane & 7) : (lane & 3));
    const int y = y0 + (flat ? (lane >> 3) : ((lane >> 2) & 3));
    const int zb = z0 + (flat ? 0 : kZc * (lane >> 4));   // first of this lane's kZc sites
    const float px = float(x), py = float(y);
    const int c = cut, c2 = c * c;
    const int bx0 = max(0, x0 - c + 1) >> 3, bx1 = min(d.X - 1, x1 + c - 1) >> 3;
    const int by0 = max(0, y0 - c + 1) >> 3, by1 = min(d.Y - 1, y1 + c - 1) >> 3;
    const int bz0 = max(0, z0 - c + 1) >> 3, bz1 = min(d.Z - 1, z1 + c - 1) >> 3;
    const int nx = bx1 - bx0 + 1, ny = by1 - by0 + 1, ncand = nx * ny * (bz1 - bz0 + 1);
    const float inx = 1.f / float(nx), iny = 1.f / float(ny);
    auto fdivi = [](int t, int dv, float inv, int& rr) {
        int qq = int(float(t) * inv);
        rr = t - qq * dv;
        if (rr < 0) { --qq; rr += dv; } else if (rr >= dv) { ++qq; rr -= dv; }
        return qq;
    };
    typedef float v4f __attribute__((ext_vector_type(4)));
    float sum[kZc], comp[kZc];
#pragma unroll
    for (int k = 0; k < kZc; ++k) { sum[k] = 0.f; comp[k] = 0.f; }
    int fill = 0, nu = 0;   // list slots and units in use
    auto flush = [&]() {
        __builtin_amdgcn_wave_barrier();
        int t = 0;
        for (int u = 0; u < nu; ++u) {
            const int e = __builtin_amdgcn_readfirstlane(uends[wv][u]);
            float acc[kZc];
#pragma unroll
            for (int k = 0; k < kZc; ++k) acc[k] = 0.f;
            auto term = [&](const float4& qc) {
                const float dx = px - qc.x, dy = py - qc.y;
                const float dxy = fmaf(dx, dx, dy * dy);
#pragma unroll
                for (int k = 0; k < kZc; ++k) {
                    const float dz = float(zb + k) - qc.z;
                    const float r2 = fmaf(dz, dz, dxy);
                    const float a = rsq(r2);
                    float f = a;
                    if (kGauss) f = erff(gk * (r2 * a)) * a;
                    acc[k] = fmaf(qc.w, in_cut(r2, c2m1) ? f : 0.f, acc[k]);
                }
            };
            for (; t + 2 <= e; t += 2) {
                const float4 q0 = lists[wv][t], q1 = lists[wv][t + 1];
                term(q0);
                term(q1);
            }
            if (t < e) term(lists[wv][t++]);
#pragma unroll
            for (int k = 0; k < kZc; ++k) kahan_add(sum[k], comp[k], acc[k]);
        }
        __builtin_amdgcn_wave_barrier();
        fill = 0;
        nu = 0;
    };
    for (int cb = 0; cb < ncand; cb += 64) {
        // this lane's candidate brick: kept if within the cutoff of the tile box
        const int ci = cb + lane;
        int k0 = 0, todo = 0;
        if (ci < ncand) {
            int rx, ry;
            const int ryz = fdivi(ci, nx, inx, rx);
            const int rz = fdivi(ryz, ny, iny, ry);
            const int bx = bx0 + rx, by = by0 + ry, bz = bz0 + rz;
            const int gx = max(0, max(kMapBrick * bx - x1, x0 - (kMapBrick * bx + kMapBrick - 1)));
            const int gy = max(0, max(kMapBrick * by - y1, y0 - (kMapBrick * by + kMapBrick - 1)));
            const int gz = max(0, max(kMapBrick * bz - z1, z0 - (kMapBrick * bz + kMapBrick - 1)));
            if (gx * gx + gy * gy + gz * gz < c2) {
                const int b = bx + m.nbx * (by + m.nby * bz);
                k0 = m.start[b];
                todo = m.start[b + 1] - k0;
            }
        }
        // into the list in lane order, whole units at a time, summing whenever it is full
        for (;;) {
            const int incl = wave_incl_sum(todo);
            const int total = wave_total(incl);
            if (total == 0) break;
            if (fill + total > kMapList && fill > 0) flush();
            const int excl = incl - todo;
            int take = min(todo, max(0, kMapList - fill - excl));   // a lane's first unit fits after a flush
            if (take < todo) take &= ~(kMapUnit - 1);
            const int nut = (take + kMapUnit - 1) / kMapUnit;
            const int uincl = wave_incl_sum(nut);
            const v4f* from = reinterpret_cast<const v4f*>(m.src) + k0;
            v4f* to = reinterpret_cast<v4f*>(&lists[wv][0]) + (fill + excl);
            int u = 0;
            for (; u + 8 <= take; u += 8) {
                const v4f a0 = from[u], a1 = from[u + 1], a2 = from[u + 2], a3 = from[u + 3];
                const v4f a4 = from[u + 4], a5 = from[u + 5], a6 = from[u + 6], a7 = from[u + 7];
                to[u] = a0; to[u + 1] = a1; to[u + 2] = a2; to[u + 3] = a3;
                to[u + 4] = a4; to[u + 5] = a5; to[u + 6] = a6; to[u + 7] = a7;
            }
            for (; u < take; ++u) to[u] = from[u];
            const int uo = nu + uincl - nut;
            for (int j = 0; j < nut; ++j) uends[wv][uo + j] = fill + excl + min(kMapUnit * (j + 1), take);
            fill += wave_total(wave_incl_sum(take));
            nu += wave_total(uincl);
            k0 += take;
            todo -= take;
        }
    }
    flush();
    if (x < d.X && y < d.Y) {
#pragma unroll
        for (int k = 0; k < kZc; ++k)
            if (zb + k < d.Z) out[encode(d, x, y, zb + k)] = pref * (sum[k] - comp[k]);
    }
}

// ---------------------------------------------------------------------------
// host helpers, shared by the context's map and the surrogate's exact fallback
// ---------------------------------------------------------------------------

// The culled map pays when the cutoff box of a 64-site tile covers well under
// the lattice and there is enough work to matter (V x n >= 1e8).
inline bool map_cull_pays(const Dims& d, int cutoff, long long nc) {
    auto frac = [&](int n) { return std::min(1.0, double(2 * cutoff + 12) / n); };
    return nc > 0 && frac(d.X) * frac(d.Y) * frac(d.Z) < 0.5 && double(d.V) * double(nc) >= 1e8;
}

inline MapBricks map_bricks_shape(const Dims& d) {
    MapBricks m{};
    m.nbx = (d.X + kMapBrick - 1) / kMapBrick;
    m.nby = (d.Y + kMapBrick - 1) / kMapBrick;
    m.nbz = (d.Z + kMapBrick - 1) / kMapBrick;
    m.nb = m.nbx * m.nby * m.nbz;
    return m;
}

// Every launch of `batch` culled maps on stream st (map b written at
// d_out + b V).  m's arrays hold batch x nb (cnt, zero on entry and again on
// exit), batch x (nb + 1) (start) and one entry per charge of the lists (the
// rest).  mb as for map_slice; with mb.off NULL, batch is 1 and the list has nc
// charges, else nc is the longest list.
inline void enqueue_map_cull(hipStream_t st, const Dims& d, const MapBricks& m, int cutoff, float gk, float pref,
                             bool gauss, const int32_t* d_cs, const int32_t* d_cq, int nc, int batch,
                             const MapBatch& mb, float* d_out) {
    const dim3 cg((nc + 255) / 256, batch);
    hipLaunchKernelGGL(mapc_count_kernel, cg, dim3(256), 0, st, d, m, d_cs, nc, mb);
    hipLaunchKernelGGL(mapc_scan_kernel, dim3(1, batch), dim3(1024), 0, st, m, mb);
    hipLaunchKernelGGL(mapc_scatter_kernel, cg, dim3(256), 0, st, d, m, d_cs, d_cq, nc, mb);
    hipLaunchKernelGGL(mapc_sort_kernel, dim3((m.nb + 3) / 4, batch), dim3(256), 0, st, m, mb);
    const int flat = d.Z < 4 ? 1 : 0;
    const int zc = d.Z >= 16 ? 4 : 1;   // sites per lane along z
    int ts, tz;
    map_tile_shape(zc, flat, ts, tz);
    const long long tiles = (long long)((d.X + ts - 1) / ts) * ((d.Y + ts - 1) / ts) * ((d.Z + tz - 1) / tz);
    const dim3 mg(unsigned((tiles + 3) / 4), batch);
    const uint32_t c2m1 = cutoff_bits(cutoff, false);
#define LMG_MAP_CULL(G, Z) \
    hipLaunchKernelGGL((map_cull_kernel<G, Z>), mg, dim3(256), 0, st, d, m, cutoff, c2m1, gk, pref, flat, mb, d_out)
    if (zc == 4) {
        if (gauss) LMG_MAP_CULL(true, 4);
        else LMG_MAP_CULL(false, 4);
    } else {
        if (gauss) LMG_MAP_CULL(true, 1);
        else LMG_MAP_CULL(false, 1);
    }
#undef LMG_MAP_CULL
}

}  // namespace
}  // namespace lmg
