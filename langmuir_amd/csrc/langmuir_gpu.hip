// langmuir_gpu.hip -- MI355X (gfx950) kernels and C-ABI for Langmuir's hot path.
//
// Kernels
//   pair_kernel / pair_reduce   pair sums at an explicit target list
//                               (kernel.cl coulomb2/gauss2, Potential::coulomb*)
//   map_kernel                  potential at every lattice site (coulomb1/gauss1)
//   de_kernel                   dE of every carrier's proposed hop, fused
//                               current/future pair sums (coulombCPU/GPU)
//   commit_kernel               one workgroup: Metropolis decisions, drains,
//                               recombination, completeTick conflict rule,
//                               order-preserving compaction, injection, next
//                               proposals and MT19937 refill (simulation.cpp:29-111)
//
// Pair sums: FP32 terms on integer-exact squared distances (0 < r^2 < c^2 is
// decided exactly), charges staged through LDS as float4 (x, y, z, q) tiles
// read as wave-wide broadcasts, plain FP32 accumulation inside a 256-charge
// tile and Kahan compensation across tiles, split over charge chunks for
// occupancy and reduced in FP64 in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "langmuir_gpu.h"
#include "lmg_device.h"

using namespace lmg;

namespace {

constexpr int kBlock = 256;      // pair / dE / map workgroup
constexpr int kTile = 256;       // charges per LDS tile
constexpr int kCommitBlock = 1024;
constexpr int kMapT = 4;         // lattice sites per map thread (along x)

// Error codes raised on the device (lmg_sim_device_error)
constexpr int kErrRing = 1, kErrCapacity = 2, kErrSite = 3, kErrBalance = 4;

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
                                 int n, float4* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int x, y, z;
    decode(d, site[i], x, y, z);
    out[i] = make_float4(float(x), float(y), float(z), q ? float(q[i]) : 0.f);
}

// ===========================================================================
// Pair sums at explicit targets.  grid = (ceil(nt / 256), S); chunk s covers
// charges [s*chunk, (s+1)*chunk).  part[s*ld + t] = (sum, compensation).
// ===========================================================================
template <bool kGauss, bool kImage>
__global__ __launch_bounds__(kBlock) void pair_kernel(const float4* __restrict__ tgt, int nt,
                                                      const float4* __restrict__ chg, int nc,
                                                      uint32_t c2m1, float gk,
                                                      float2* __restrict__ part, int ld) {
    __shared__ float4 tile[kTile];
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
        float acc = 0.f;
        for (int k = 0; k < n; ++k) {
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
    if (t < nt) part[blockIdx.y * ld + t] = make_float2(s, -c);
}

__global__ void pair_reduce_kernel(const float2* __restrict__ part, int nt, int S, int ld,
                                   double scale, double* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double v = 0.0;
    for (int s = 0; s < S; ++s) {
        float2 p = part[s * ld + t];
        v += double(p.x) + double(p.y);
    }
    out[t] = scale * v;
}

// ===========================================================================
// Full-lattice map.  A thread owns kMapT consecutive x sites of one (y, z)
// row; a workgroup owns (tx * kMapT) x (256 / tx) sites of one z plane.
// ===========================================================================
template <bool kGauss>
__global__ __launch_bounds__(kBlock) void map_kernel(Dims d, const float4* __restrict__ chg, int nc,
                                                     uint32_t c2m1, float gk, float pref, int tx_n,
                                                     float* __restrict__ out) {
    __shared__ float4 tile[kTile];
    const int tx = threadIdx.x % tx_n, ty = threadIdx.x / tx_n;
    const int ty_n = kBlock / tx_n;
    const int x0 = (blockIdx.x * tx_n + tx) * kMapT;
    const int y = blockIdx.y * ty_n + ty;
    const int z = blockIdx.z;
    const float fy = float(y), fz = float(z);
    float s[kMapT], c[kMapT];
#pragma unroll
    for (int u = 0; u < kMapT; ++u) { s[u] = 0.f; c[u] = 0.f; }
    for (int base = 0; base < nc; base += kTile) {
        const int n = min(kTile, nc - base);
        __syncthreads();
        if (threadIdx.x < n) tile[threadIdx.x] = chg[base + threadIdx.x];
        __syncthreads();
        float acc[kMapT];
#pragma unroll
        for (int u = 0; u < kMapT; ++u) acc[u] = 0.f;
        for (int k = 0; k < n; ++k) {
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
    if (y < d.Y) {
#pragma unroll
        for (int u = 0; u < kMapT; ++u) {
            const int x = x0 + u;
            if (x < d.X) out[encode(d, x, y, z)] = pref * (s[u] - c[u]);
        }
    }
}

// ===========================================================================
// dE of every carrier's proposed hop.  Target i has current site tgt1[i] and
// proposed site tgt2[i]; for every charge j the kernel accumulates
//   q_j * ( [0<r2<c2] f(r2) - [0<r1<c2] f(r1) ),
// the bracket of  p2 - p1  in ChargeAgent::coulombCPU (chargeagent.cpp:215-271).
// For 1/r the difference of two in-cutoff terms is formed without cancellation:
//   1/sqrt(r2) - 1/sqrt(r1) = (r1 - r2) (ab)^2 / (a + b),  a = r1^-1/2, b = r2^-1/2,
// with r1 - r2 an exact integer.  Counts (targets, charges) are read on the
// device so the launch needs no host synchronisation.
// ===========================================================================
template <bool kGauss>
__global__ __launch_bounds__(kBlock) void de_kernel(const float4* __restrict__ tgt1,
                                                    const float4* __restrict__ tgt2,
                                                    const int* __restrict__ counts,
                                                    const float4* __restrict__ chg, uint32_t c2m1,
                                                    float gk, float2* __restrict__ part, int ld) {
    __shared__ float4 tile[kTile];
    const int nt = counts[0], nc = counts[1];
    if (int(blockIdx.x) * kBlock >= nt) return;     // uniform per workgroup
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int S = gridDim.y;
    const int chunk = (nc + S - 1) / S;
    const int j0 = blockIdx.y * chunk, j1 = min(nc, j0 + chunk);
    const float4 p1 = i < nt ? tgt1[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 p2 = i < nt ? tgt2[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float hx = p2.x - p1.x, hy = p2.y - p1.y, hz = p2.z - p1.z;
    const float hh = hx * hx + hy * hy + hz * hz;
    float s = 0.f, c = 0.f;
    for (int base = j0; base < j1; base += kTile) {
        const int n = min(kTile, j1 - base);
        __syncthreads();
        if (threadIdx.x < n) tile[threadIdx.x] = chg[base + threadIdx.x];
        __syncthreads();
        float acc = 0.f;
        for (int k = 0; k < n; ++k) {
            const float4 q = tile[k];
            const float dx = p1.x - q.x, dy = p1.y - q.y, dz = p1.z - q.z;
            const float r1 = dx * dx + dy * dy + dz * dz;
            // |p2 - q|^2 = r1 + 2 h.(p1 - q) + |h|^2, exact in integers
            const float r2 = r1 + (2.f * (hx * dx + hy * dy + hz * dz) + hh);
            const bool in1 = in_cut(r1, c2m1), in2 = in_cut(r2, c2m1);
            const float a = rsq(r1), b = rsq(r2);
            float term;
            if (kGauss) {
                const float fa = in1 ? erff(gk * (r1 * a)) * a : 0.f;
                const float fb = in2 ? erff(gk * (r2 * b)) * b : 0.f;
                term = fb - fa;
            } else {
                const float ab = a * b;
                const float both = (r1 - r2) * (ab * ab) * rcp(a + b);
                term = (in1 && in2) ? both : ((in2 ? b : 0.f) - (in1 ? a : 0.f));
            }
            acc = fmaf(q.w, term, acc);
        }
        kahan_add(s, c, acc);
    }
    if (i < nt) part[blockIdx.y * ld + i] = make_float2(s, -c);
}

// ===========================================================================
// Device state of a simulation
// ===========================================================================
struct SimCtl {
    int n_e, n_h;
    int cur;                        // list buffer holding the current lists
    int error;
    unsigned long long rng_pos;     // engine outputs consumed
    unsigned long long rng_step0;   // rng_pos before the pending step's proposals
    unsigned long long gen_pos;     // raw words x[0 .. gen_pos) generated
    unsigned long long att[10], suc[10];
    long long step;
    int counts[2];                  // de_kernel inputs: targets, charges
};

struct SimDev {
    Dims d;
    int hop, sim_type, coulomb, src_metro, src_coul, balance, rec_range, gauss;
    int max_e, max_h, cap_e, cap_h, n_d_eff, dq;
    int e_drain[2], h_drain[2];
    uint32_t c2m1;
    float gk;
    double vl, vr, beta, pref, self_int, binding, gk_d;
    double rate[10];
    double coupling[27];
    int S, ld;
    unsigned long long lookahead;
    int* e_site[2];
    int* e_life[2];
    int* e_path[2];
    int* h_site[2];
    int* h_life[2];
    int* h_path[2];
    int* e_fsite;
    int* h_fsite;
    uint8_t* etype;
    uint8_t* htype;
    double* vbg;
    int* claim_e;
    int* claim_h;
    int* h_at;
    int* code;      // per carrier: 0 stay, 1 accepted to an empty site, 2 accepted into a drain
    int* removed;   // per carrier: recombined this step
    int* flag;      // per carrier scratch
    float4* tgt1;
    float4* tgt2;
    float4* chg;    // [defects (if charged) | electrons | holes]
    const float2* part;
    double* de;
    uint32_t* ring;
    unsigned long long ring_mask;
    SimCtl* ctl;
};

struct CommitShared {
    int wsum[32];
    unsigned long long pos;     // serial RNG position (thread 0 writes)
    int n_e, n_h;
    int cur;
    int first_bad;
    int total;
    int site, valid, ok, need_energy;
    int bal_go, bal_src;
    unsigned int att[10], suc[10];
    double red[32];
    uint32_t mt_tail[1024];
};

__device__ __forceinline__ void set_error(const SimDev& D, int code) {
    atomicCAS(&D.ctl->error, 0, code);
}

// exclusive scan of one int per thread over the workgroup (1024 threads)
__device__ int block_scan(int v, int* wsum, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int w = lane < nw ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (lane < nw) wsum[lane] = w;
    }
    __syncthreads();
    const int excl = x - v + (wid > 0 ? wsum[wid - 1] : 0);
    total = wsum[nw - 1];
    __syncthreads();
    return excl;
}

__device__ double block_sum_d(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += red[w];   // fixed order
    __syncthreads();
    return t;
}

// engine output k, and the serial draw helpers used by thread 0
__device__ __forceinline__ uint32_t rng_word(const SimDev& D, unsigned long long k) {
    return mt_temper(D.ring[(k + 624ull) & D.ring_mask]);
}
struct Serial {
    const SimDev& D;
    unsigned long long& pos;
    __device__ uint32_t word() {
        if (pos + 624ull >= D.ctl->gen_pos) set_error(D, kErrRing);
        return rng_word(D, pos++);
    }
    __device__ double random() { return u01(word()); }
    __device__ int integer(int lo, int hi) {      // Boost uniform_int (rand.cpp:68-73)
        const uint32_t range = uint32_t(hi) - uint32_t(lo);
        if (range == 0) return lo;
        if (range == 0xffffffffu) return int(uint32_t(lo) + word());
        const uint32_t b = uint_bucket(range);
        for (;;) {
            uint32_t r = word() / b;
            if (r <= range) return int(uint32_t(lo) + r);
        }
    }
    __device__ bool metropolis_coupling(double de, double beta, double cpl) {   // rand.cpp:91-106
        const double u = random();
        if (de > 0.0) return cpl * exp(-de * beta) > u;
        return cpl > u;
    }
    __device__ bool choose_yes(double p) { return p > random(); }   // rand.cpp:117-124
};

// potential of all charges at one site, direct minus left image, x prefactor
// (the sum SourceAgent::energyChange adds for source.coulomb, sourceagent.cpp:125-161)
__device__ double block_site_potential(const SimDev& D, CommitShared& sh, int site, int ncharge) {
    int x, y, z;
    decode(D.d, site, x, y, z);
    const float fx = float(x), fy = float(y), fz = float(z);
    float s = 0.f, c = 0.f;
    for (int j = threadIdx.x; j < ncharge; j += blockDim.x) {
        const float4 q = D.chg[j];
        const float dy = fy - q.y, dz = fz - q.z;
        const float dx = fx - q.x, dxi = fx + q.x + 1.f;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float ri = dxi * dxi + dy * dy + dz * dz;
        const float t = (in_cut(r2, D.c2m1) ? rsq(r2) : 0.f) - (in_cut(ri, D.c2m1) ? rsq(ri) : 0.f);
        kahan_add(s, c, q.w * t);
    }
    return D.pref * block_sum_d(double(s) - double(c), sh.red);
}

// (x, y, z, q) of carrier list entry into the charge array; also keeps
// the per-carrier target arrays
__device__ __forceinline__ float4 site_f4(const Dims& d, int s, float q) {
    int x, y, z;
    decode(d, s, x, y, z);
    return make_float4(float(x), float(y), float(z), q);
}

// ---------------------------------------------------------------------------
// decisions (chargeagent.cpp:90-161) for every carrier, electrons then holes
// ---------------------------------------------------------------------------
__device__ void phase_decide(const SimDev& D, CommitShared& sh, bool only_de) {
    const int ne = sh.n_e, nh = sh.n_h, N = ne + nh, cur = sh.cur;
    const int K = (N + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * K, i1 = min(N, i0 + K);
    int cnt = 0;
    for (int i = i0; i < i1; ++i) {
        const bool isE = i < ne;
        const int li = isE ? i : i - ne;
        const int fs = isE ? D.e_fsite[li] : D.h_fsite[li];
        const uint8_t tf = isE ? D.etype[fs] : D.htype[fs];
        const int draw = (tf == kEmpty || tf == kDrain);
        D.flag[i] = draw;
        cnt += draw;
    }
    int total;
    int off = block_scan(cnt, sh.wsum, total);
    const unsigned long long base = sh.pos;
    if (threadIdx.x == 0 && !only_de && base + total + 624ull > D.ctl->gen_pos) set_error(D, kErrRing);
    for (int i = i0; i < i1; ++i) {
        const bool isE = i < ne;
        const int li = isE ? i : i - ne;
        const int q = isE ? -1 : 1;
        const int site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
        int fs = isE ? D.e_fsite[li] : D.h_fsite[li];
        // dE (chargeagent.cpp:215-271); zero when Coulomb is off (chooseFuture sets m_de = 0)
        double de = 0.0;
        if (D.coulomb) {
            double sde = 0.0;
            for (int s = 0; s < D.S; ++s) {
                const float2 p = D.part[s * D.ld + i];
                sde += double(p.x) + double(p.y);
            }
            // bindingPotential (chargeagent.cpp:463-481)
            double bf, bs;
            if (isE) {
                bf = D.htype[fs] == kHole ? D.self_int + D.binding : 0.0;
                bs = D.htype[site] == kHole ? D.self_int + D.binding : 0.0;
            } else {
                bf = D.etype[fs] == kElectron ? D.self_int - D.binding : 0.0;
                bs = D.etype[site] == kElectron ? D.self_int - D.binding : 0.0;
            }
            de = q * (D.pref * sde - D.self_int * q + bf - bs);
        }
        D.de[i] = de;
        if (only_de) continue;
        const uint8_t tf = isE ? D.etype[fs] : D.htype[fs];
        int code = 0;
        if (tf == kEmpty) {
            const double u = u01(rng_word(D, base + off));
            ++off;
            double pd = D.vbg[fs] - D.vbg[site];
            pd *= q;
            pd += de;
            int x1, y1, z1, x2, y2, z2;
            decode(D.d, site, x1, y1, z1);
            decode(D.d, fs, x2, y2, z2);
            const int h1 = D.hop + 1;
            const double cpl = D.coupling[(abs(x1 - x2) * h1 + abs(y1 - y2)) * h1 + abs(z1 - z2)];
            const bool acc = (pd > 0.0) ? (cpl * exp(-pd * D.beta) > u) : (cpl > u);
            code = acc ? 1 : 0;
        } else if (tf == kDrain) {
            const double u = u01(rng_word(D, base + off));
            ++off;
            const int* ds = isE ? D.e_drain : D.h_drain;
            const int flux = isE ? (fs == ds[0] ? LMG_FLUX_EDL : LMG_FLUX_EDR)
                                 : (fs == ds[0] ? LMG_FLUX_HDL : LMG_FLUX_HDR);
            atomicAdd(&sh.att[flux], 1u);
            if (D.rate[flux] > u) {
                atomicAdd(&sh.suc[flux], 1u);
                code = 2;
            }
        }
        int* life = isE ? D.e_life[cur] : D.h_life[cur];
        int* path = isE ? D.e_path[cur] : D.h_path[cur];
        life[li] += 1;
        if (code) path[li] += 1;
        else fs = site;
        if (isE) D.e_fsite[li] = fs; else D.h_fsite[li] = fs;
        D.code[i] = code;
        D.removed[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && !only_de) sh.pos = base + total;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// RecombinationAgent::tryToAccept over the electrons (drainagent.cpp:113-214)
// ---------------------------------------------------------------------------
__device__ void phase_recombine(const SimDev& D, CommitShared& sh) {
    const int ne = sh.n_e, nh = sh.n_h, cur = sh.cur;
    for (int j = threadIdx.x; j < nh; j += blockDim.x) D.h_at[D.h_site[cur][j]] = j;
    __syncthreads();
    if (threadIdx.x == 0) {
        Serial R{D, sh.pos};
        int nb[40];
        int cand[34];
        for (int i = 0; i < ne; ++i) {
            const int site = D.e_site[cur][i];
            int n = 0;
            if (D.htype[site] == kHole) nb[n++] = site;
            if (D.rec_range != 0) {
                const int m = nbr_list(D.d, D.rec_range, site, D.e_drain[0], D.e_drain[1], cand);
                for (int k = 0; k < m; ++k)
                    if (D.htype[cand[k]] == kHole) nb[n++] = cand[k];
            }
            if (n == 0) continue;
            const int rs = nb[R.integer(0, n - 1)];
            const int other = D.h_at[rs];
            if (D.removed[ne + other]) continue;
            sh.att[LMG_FLUX_XD] += 1;
            if (R.choose_yes(D.rate[LMG_FLUX_XD])) {
                sh.suc[LMG_FLUX_XD] += 1;
                D.removed[i] = 1;
                D.removed[ne + other] = 1;
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// nextTick / completeTick (simulation.cpp:355-414, chargeagent.cpp:163-199):
// in list order, the first mover claiming an empty site wins; order-preserving
// removal of drained and recombined carriers into the other list buffer.
// ---------------------------------------------------------------------------
__device__ void phase_tick(const SimDev& D, CommitShared& sh) {
    const int ne = sh.n_e, nh = sh.n_h, N = ne + nh, cur = sh.cur, nxt = cur ^ 1;
    const int K = (N + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * K, i1 = min(N, i0 + K);
    for (int i = i0; i < i1; ++i) {
        if (D.code[i] == 1 && !D.removed[i]) {
            const bool isE = i < ne;
            const int li = isE ? i : i - ne;
            const int fs = isE ? D.e_fsite[li] : D.h_fsite[li];
            atomicMin(isE ? &D.claim_e[fs] : &D.claim_h[fs], li);
        }
    }
    __syncthreads();
    int keep_e = 0, keep_h = 0;
    for (int i = i0; i < i1; ++i) {
        const bool isE = i < ne;
        const int li = isE ? i : i - ne;
        const int site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
        uint8_t* ty = isE ? D.etype : D.htype;
        const int code = D.code[i];
        int keep = 1, dest = site;
        if (D.removed[i]) {
            ty[site] = kEmpty;
            keep = 0;
        } else if (code == 1) {
            const int fs = isE ? D.e_fsite[li] : D.h_fsite[li];
            if ((isE ? D.claim_e[fs] : D.claim_h[fs]) == li) {
                ty[site] = kEmpty;
                ty[fs] = isE ? kElectron : kHole;
                dest = fs;
            }
        } else if (code == 2) {
            ty[site] = kEmpty;
            keep = 0;
        }
        D.flag[i] = keep ? dest : -1;
        if (isE) keep_e += keep; else keep_h += keep;
    }
    int tot_e, tot_h;
    const int off_e = block_scan(keep_e, sh.wsum, tot_e);
    const int off_h = block_scan(keep_h, sh.wsum, tot_h);
    int we = off_e, wh = off_h;
    for (int i = i0; i < i1; ++i) {
        const bool isE = i < ne;
        const int li = isE ? i : i - ne;
        const int dest = D.flag[i];
        if (dest >= 0) {
            if (isE) {
                D.e_site[nxt][we] = dest;
                D.e_life[nxt][we] = D.e_life[cur][li];
                D.e_path[nxt][we] = D.e_path[cur][li];
                ++we;
            } else {
                D.h_site[nxt][wh] = dest;
                D.h_life[nxt][wh] = D.h_life[cur][li];
                D.h_path[nxt][wh] = D.h_path[cur][li];
                ++wh;
            }
        }
        if (D.code[i] == 1) {   // release claims (all reads are behind the scans' barriers)
            const int fs = isE ? D.e_fsite[li] : D.h_fsite[li];
            if (isE) D.claim_e[fs] = 0x7fffffff; else D.claim_h[fs] = 0x7fffffff;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sh.n_e = tot_e;
        sh.n_h = tot_h;
        sh.cur = nxt;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Injection (simulation.cpp:233-317, sourceagent.cpp:92-247).  All threads
// call these; thread 0 owns the RNG and the lists, the block helps with the
// source.coulomb pair sum.
// ---------------------------------------------------------------------------
__device__ void append_carrier(const SimDev& D, CommitShared& sh, bool isE, int site) {
    const int cur = sh.cur;
    if (isE) {
        const int n = sh.n_e;
        if (n >= D.cap_e) { set_error(D, kErrCapacity); return; }
        D.e_site[cur][n] = site; D.e_life[cur][n] = 0; D.e_path[cur][n] = 0;
        D.etype[site] = kElectron;
        sh.n_e = n + 1;
    } else {
        const int n = sh.n_h;
        if (n >= D.cap_h) { set_error(D, kErrCapacity); return; }
        D.h_site[cur][n] = site; D.h_life[cur][n] = 0; D.h_path[cur][n] = 0;
        D.htype[site] = kHole;
        sh.n_h = n + 1;
    }
}

// charges array for source.coulomb sums: rebuilt from the current lists
__device__ void refresh_charges(const SimDev& D, CommitShared& sh) {
    const int ne = sh.n_e, nh = sh.n_h, cur = sh.cur, nd = D.n_d_eff;
    for (int i = threadIdx.x; i < ne + nh; i += blockDim.x) {
        const bool isE = i < ne;
        const int s = isE ? D.e_site[cur][i] : D.h_site[cur][i - ne];
        D.chg[nd + i] = site_f4(D.d, s, isE ? -1.f : 1.f);
    }
    __syncthreads();
}

__device__ bool inject_face(const SimDev& D, CommitShared& sh, int flux) {
    const bool isE = (flux == LMG_FLUX_ESL || flux == LMG_FLUX_ESR);
    const bool left = (flux == LMG_FLUX_ESL || flux == LMG_FLUX_HSL);
    if (threadIdx.x == 0) {
        Serial R{D, sh.pos};
        sh.att[flux] += 1;
        const int yz = D.d.Y * D.d.Z;
        // neighborsFace: the sorted x = 0 (or X-1) slice; entry m is X*m (+ X-1)
        const int m = R.integer(0, yz - 1);
        const int site = D.d.X * m + (left ? 0 : D.d.X - 1);
        const int n = isE ? sh.n_e : sh.n_h;
        const int mx = isE ? D.max_e : D.max_h;
        const uint8_t t = isE ? D.etype[site] : D.htype[site];
        sh.site = site;
        sh.valid = !(n >= mx || D.rate[flux] <= 0 || t != kEmpty);
        sh.need_energy = sh.valid && D.src_metro && D.src_coul;
    }
    __syncthreads();
    double coul = 0.0;
    if (sh.need_energy) {
        refresh_charges(D, sh);
        coul = block_site_potential(D, sh, sh.site, D.n_d_eff + sh.n_e + sh.n_h);
    }
    if (threadIdx.x == 0) {
        sh.ok = 0;
        if (sh.valid) {
            Serial R{D, sh.pos};
            bool go;
            if (D.src_metro) {
                const double p2 = D.vbg[sh.site] + coul;   // m_potential of a source is 0
                const double dE = isE ? (0.0 - p2) : (p2 - 0.0);
                go = R.metropolis_coupling(dE, D.beta, D.rate[flux]);
            } else {
                go = R.choose_yes(D.rate[flux]);
            }
            if (go) {
                append_carrier(D, sh, isE, sh.site);
                sh.suc[flux] += 1;
                sh.ok = 1;
            }
        }
    }
    __syncthreads();
    return sh.ok != 0;
}

__device__ void phase_inject(const SimDev& D, CommitShared& sh) {
    if (D.sim_type == 1) {
        if (threadIdx.x == 0) {
            Serial R{D, sh.pos};
            sh.att[LMG_FLUX_XS] += 1;
            const int site = R.integer(0, D.d.V - 1);
            const bool valid = !(sh.n_e >= D.max_e || sh.n_h >= D.max_h || D.rate[LMG_FLUX_XS] <= 0 ||
                                 D.etype[site] != kEmpty || D.htype[site] != kEmpty);
            if (valid && R.choose_yes(D.rate[LMG_FLUX_XS])) {
                append_carrier(D, sh, true, site);
                append_carrier(D, sh, false, site);
                sh.suc[LMG_FLUX_XS] += 1;
            }
        }
        __syncthreads();
        if (D.balance) {   // simulation.cpp:253-317
            for (int pass = 0; pass < 2; ++pass) {
                int tries = 0;
                for (;;) {
                    if (threadIdx.x == 0) {
                        const int ne = sh.n_e, nh = sh.n_h;
                        const bool go = pass == 0 ? (ne - nh < 0 && !(ne >= D.max_e))
                                                  : (nh - ne < 0 && !(nh >= D.max_h));
                        sh.bal_go = go;
                        if (go) {
                            int src;
                            if (D.vr > D.vl) src = pass == 0 ? LMG_FLUX_ESL : LMG_FLUX_HSR;
                            else if (D.vl > D.vr) src = pass == 0 ? LMG_FLUX_ESR : LMG_FLUX_HSL;
                            else {
                                Serial R{D, sh.pos};
                                const bool hi = R.random() > 0.5;
                                src = pass == 0 ? (hi ? LMG_FLUX_ESL : LMG_FLUX_ESR)
                                                : (hi ? LMG_FLUX_HSR : LMG_FLUX_HSL);
                            }
                            sh.bal_src = src;
                        }
                    }
                    __syncthreads();
                    if (!sh.bal_go) break;
                    inject_face(D, sh, sh.bal_src);
                    if (++tries > 1024) {
                        if (threadIdx.x == 0) set_error(D, kErrBalance);
                        break;
                    }
                }
                __syncthreads();
            }
        }
    } else {
        inject_face(D, sh, LMG_FLUX_ESL);
        inject_face(D, sh, LMG_FLUX_ESR);
        inject_face(D, sh, LMG_FLUX_HSL);
        inject_face(D, sh, LMG_FLUX_HSR);
    }
}

// ---------------------------------------------------------------------------
// chooseFuture for every carrier (chargeagent.cpp:73-78): one uniform_int draw
// over the neighbour list, in list order.  Offsets come from a scan (a
// one-entry list draws nothing); the rare Boost rejection falls back to a
// serial pass from the first rejecting carrier.  Also stages the float4
// targets/charges for de_kernel.
// ---------------------------------------------------------------------------
__device__ int proposal_site(const SimDev& D, bool isE, int site, int k) {
    const int dl = isE ? D.e_drain[0] : D.h_drain[0];
    const int dr = isE ? D.e_drain[1] : D.h_drain[1];
    if (D.hop == 1) {
        int x, y, z;
        decode(D.d, site, x, y, z);
        return nbr1_at(D.d, x, y, z, dl, dr, k);
    }
    int nb[40];
    nbr_list(D.d, D.hop, site, dl, dr, nb);
    return nb[k];
}
__device__ int proposal_count(const SimDev& D, bool isE, int site) {
    const int dl = isE ? D.e_drain[0] : D.h_drain[0];
    const int dr = isE ? D.e_drain[1] : D.h_drain[1];
    if (D.hop == 1) {
        int x, y, z;
        decode(D.d, site, x, y, z);
        return nbr1_count(D.d, x, y, z, dl, dr);
    }
    int nb[40];
    return nbr_list(D.d, D.hop, site, dl, dr, nb);
}

__device__ void phase_propose(const SimDev& D, CommitShared& sh) {
    const int ne = sh.n_e, nh = sh.n_h, N = ne + nh, cur = sh.cur, nd = D.n_d_eff;
    const int K = (N + blockDim.x - 1) / blockDim.x;
    const int i0 = threadIdx.x * K, i1 = min(N, i0 + K);
    if (threadIdx.x == 0) sh.first_bad = 0x7fffffff;
    int cnt = 0;
    for (int i = i0; i < i1; ++i) {
        const bool isE = i < ne;
        const int li = isE ? i : i - ne;
        const int site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
        if (site < 0 || site >= D.d.V) set_error(D, kErrSite);
        const int n = proposal_count(D, isE, site);
        D.flag[i] = n;
        cnt += (n >= 2);
    }
    int total;
    int off = block_scan(cnt, sh.wsum, total);
    const unsigned long long base = sh.pos;
    if (threadIdx.x == 0 && base + total + 624ull > D.ctl->gen_pos) set_error(D, kErrRing);
    for (int i = i0; i < i1; ++i) {
        const bool isE = i < ne;
        const int li = isE ? i : i - ne;
        const int site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
        const int n = D.flag[i];
        int k = 0;
        if (n >= 2) {
            const uint32_t range = uint32_t(n - 1);
            const uint32_t r = rng_word(D, base + off) / uint_bucket(range);
            ++off;
            if (r > range) atomicMin(&sh.first_bad, i);
            k = min(int(r), n - 1);
        }
        const int fs = proposal_site(D, isE, site, k);
        if (isE) D.e_fsite[li] = fs; else D.h_fsite[li] = fs;
        const float q = isE ? -1.f : 1.f;
        const float4 p1 = site_f4(D.d, site, q);
        D.tgt1[i] = p1;
        D.tgt2[i] = site_f4(D.d, fs, q);
        D.chg[nd + i] = p1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long end = base + total;
        if (sh.first_bad < N) {
            // recompute the draw offset of the first rejecting carrier, then go serially
            unsigned long long pos = base;
            for (int i = 0; i < sh.first_bad; ++i) pos += (D.flag[i] >= 2);
            Serial R{D, pos};
            for (int i = sh.first_bad; i < N; ++i) {
                const bool isE = i < ne;
                const int li = isE ? i : i - ne;
                const int site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
                const int kk = R.integer(0, D.flag[i] - 1);
                const int fs = proposal_site(D, isE, site, kk);
                if (isE) D.e_fsite[li] = fs; else D.h_fsite[li] = fs;
                D.tgt2[i] = site_f4(D.d, fs, isE ? -1.f : 1.f);
            }
            end = pos;
        }
        sh.pos = end;
        D.ctl->counts[0] = N;
        D.ctl->counts[1] = nd + N;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// MT19937 refill: x[m] = x[m-227] ^ twist(x[m-624], x[m-623]) lets 227
// consecutive words be produced in parallel; the last 1024 words live in LDS.
// ---------------------------------------------------------------------------
__device__ void phase_refill(const SimDev& D, CommitShared& sh) {
    __syncthreads();
    const unsigned long long target = sh.pos + 624ull + D.lookahead;
    unsigned long long gen = D.ctl->gen_pos;
    if (gen >= target) return;
    for (int t = threadIdx.x; t < 624; t += blockDim.x) {
        const unsigned long long m = gen - 624ull + t;
        sh.mt_tail[m & 1023ull] = D.ring[m & D.ring_mask];
    }
    __syncthreads();
    while (gen < target) {
        if (threadIdx.x < 227) {
            const unsigned long long m = gen + threadIdx.x;
            const uint32_t v = mt_next(sh.mt_tail[(m - 624ull) & 1023ull], sh.mt_tail[(m - 623ull) & 1023ull],
                                       sh.mt_tail[(m - 227ull) & 1023ull]);
            sh.mt_tail[m & 1023ull] = v;
            D.ring[m & D.ring_mask] = v;
        }
        gen += 227ull;
        __syncthreads();
    }
    if (threadIdx.x == 0) D.ctl->gen_pos = gen;
    __syncthreads();
}

// mode 0: finish the pending step and propose the next one
// mode 1: propose only (after a state load)
// mode 2: dE of the pending proposals only (no RNG use)
__global__ __launch_bounds__(kCommitBlock) void commit_kernel(SimDev D, int mode) {
    __shared__ CommitShared sh;
    if (threadIdx.x == 0) {
        SimCtl& C = *D.ctl;
        sh.n_e = C.n_e;
        sh.n_h = C.n_h;
        sh.cur = C.cur;
        sh.pos = C.rng_pos;
    }
    if (threadIdx.x < 10) {
        sh.att[threadIdx.x] = 0;
        sh.suc[threadIdx.x] = 0;
    }
    __syncthreads();
    if (mode == 2) {
        phase_decide(D, sh, true);
        return;
    }
    if (mode == 0) {
        phase_decide(D, sh, false);
        if (D.sim_type == 1 && D.rate[LMG_FLUX_XD] > 0) phase_recombine(D, sh);
        phase_tick(D, sh);
        phase_inject(D, sh);
    }
    if (mode == 1) phase_refill(D, sh);   // words for the first proposals
    if (threadIdx.x == 0) D.ctl->rng_step0 = sh.pos;
    phase_propose(D, sh);
    phase_refill(D, sh);
    if (threadIdx.x == 0) {
        SimCtl& C = *D.ctl;
        C.n_e = sh.n_e;
        C.n_h = sh.n_h;
        C.cur = sh.cur;
        C.rng_pos = sh.pos;
        if (mode == 0) C.step += 1;
        if (sh.pos + 624ull > C.gen_pos) set_error(D, kErrRing);
    }
    if (threadIdx.x < 10) {
        D.ctl->att[threadIdx.x] += sh.att[threadIdx.x];
        D.ctl->suc[threadIdx.x] += sh.suc[threadIdx.x];
    }
}

__global__ void fill_i32_kernel(int* p, int n, int v) {
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
    // simulation
    bool sim_loaded = false;
    bool sync_mode = false;
    lmg_sim_config cfg{};
    SimDev sim{};
    SimCtl* ctl = nullptr;
    int de_grid_x = 0;
    std::vector<void*> sim_allocs;
    std::vector<hipEvent_t> events;
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

lmg_status pair_core(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int nc, const int32_t* d_ts,
                     int nt, int flags, double* d_out) {
    lmg_status st;
    if ((st = ensure(c, c->b_tgt, size_t(nt) * sizeof(float4))) != LMG_OK) return st;
    if ((st = ensure(c, c->b_chg, size_t(std::max(nc, 1)) * sizeof(float4))) != LMG_OK) return st;
    const int S = split_count(nt, nc);
    if ((st = ensure(c, c->b_part, size_t(S) * nt * sizeof(float2))) != LMG_OK) return st;
    float4* tgt = static_cast<float4*>(c->b_tgt.p);
    float4* chg = static_cast<float4*>(c->b_chg.p);
    float2* part = static_cast<float2*>(c->b_part.p);
    if (nc > 0) hipLaunchKernelGGL(to_float4_kernel, dim3((nc + 255) / 256), dim3(256), 0, c->stream, c->d, d_cs, d_cq, nc, chg);
    hipLaunchKernelGGL(to_float4_kernel, dim3((nt + 255) / 256), dim3(256), 0, c->stream, c->d, d_ts, nullptr, nt, tgt);
    const uint32_t c2m1 = cutoff_bits(c->p.cutoff, false);
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
    hipLaunchKernelGGL(pair_reduce_kernel, dim3((nt + 255) / 256), dim3(256), 0, c->stream, part, nt, S, nt, scale, d_out);
    return launch_check(c, "pair_kernel");
}

lmg_status map_core(lmg_ctx* c, const int32_t* d_cs, const int32_t* d_cq, int nc, int flags, float* d_out) {
    lmg_status st;
    if ((st = ensure(c, c->b_chg, size_t(std::max(nc, 1)) * sizeof(float4))) != LMG_OK) return st;
    float4* chg = static_cast<float4*>(c->b_chg.p);
    if (nc > 0) hipLaunchKernelGGL(to_float4_kernel, dim3((nc + 255) / 256), dim3(256), 0, c->stream, c->d, d_cs, d_cq, nc, chg);
    int tx = 1;
    while (tx < 64 && tx * kMapT < c->d.X) tx <<= 1;
    const int ty = kBlock / tx;
    dim3 grid((c->d.X + tx * kMapT - 1) / (tx * kMapT), (c->d.Y + ty - 1) / ty, c->d.Z);
    const uint32_t c2m1 = cutoff_bits(c->p.cutoff, flags & LMG_NO_CUTOFF);
    const float gk = gauss_k(c->p.gaussian_sigma);
    const float pref = float(c->p.prefactor);
    if (flags & LMG_GAUSS)
        hipLaunchKernelGGL(map_kernel<true>, grid, dim3(kBlock), 0, c->stream, c->d, chg, nc, c2m1, gk, pref, tx, d_out);
    else
        hipLaunchKernelGGL(map_kernel<false>, grid, dim3(kBlock), 0, c->stream, c->d, chg, nc, c2m1, gk, pref, tx, d_out);
    return launch_check(c, "map_kernel");
}

void sim_free(lmg_ctx* c) {
    for (void* p : c->sim_allocs) (void)hipFree(p);
    c->sim_allocs.clear();
    c->sim_loaded = false;
    c->ctl = nullptr;
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

lmg_status sim_launch_de(lmg_ctx* c) {
    SimDev& D = c->sim;
    dim3 grid(c->de_grid_x, D.S);
    if (D.gauss)
        hipLaunchKernelGGL(de_kernel<true>, grid, dim3(kBlock), 0, c->stream, D.tgt1, D.tgt2, c->ctl->counts, D.chg, D.c2m1, D.gk, const_cast<float2*>(D.part), D.ld);
    else
        hipLaunchKernelGGL(de_kernel<false>, grid, dim3(kBlock), 0, c->stream, D.tgt1, D.tgt2, c->ctl->counts, D.chg, D.c2m1, D.gk, const_cast<float2*>(D.part), D.ld);
    return launch_check(c, "de_kernel");
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
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(nullptr, LMG_ERR_NO_DEVICE, "no HIP device available");
    if (p.device_id < 0 || p.device_id >= n) return fail(nullptr, LMG_ERR_ARG, "device_id out of range");
    lmg_ctx* c = new lmg_ctx;
    c->p = p;
    c->d = Dims{p.grid_x, p.grid_y, p.grid_z, p.grid_x * p.grid_y, p.grid_x * p.grid_y * p.grid_z};
    c->dev = p.device_id;
    hipError_t e = hipSetDevice(c->dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_create_error = std::string("HIP initialisation failed: ") + hipGetErrorString(e);
        delete c;
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
    for (DevBuf* b : {&c->b_tgt, &c->b_chg, &c->b_part, &c->b_site, &c->b_q, &c->b_tsite, &c->b_out, &c->b_map})
        if (b->p) (void)hipFree(b->p);
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

void* lmg_stream(lmg_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

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
    if (flags & ~(LMG_GAUSS | LMG_IMAGE_LEFT)) return fail(c, LMG_ERR_ARG, "unsupported flags");
    if ((flags & LMG_GAUSS) && !(c->p.gaussian_sigma > 0))
        return fail(c, LMG_ERR_ARG, "LMG_GAUSS needs gaussian_sigma > 0");
    if (nt == 0) return LMG_OK;
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    return pair_core(c, d_cs, d_cq, nc, d_ts, nt, flags, d_out);
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
    if (flags & ~(LMG_GAUSS | LMG_NO_CUTOFF)) return fail(c, LMG_ERR_ARG, "unsupported flags");
    if ((flags & LMG_GAUSS) && !(c->p.gaussian_sigma > 0))
        return fail(c, LMG_ERR_ARG, "LMG_GAUSS needs gaussian_sigma > 0");
    lmg_status st;
    if ((st = use_device(c)) != LMG_OK) return st;
    return map_core(c, d_cs, d_cq, nc, flags, d_out);
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
    D.gk_d = c->p.gaussian_sigma > 0 ? 1.0 / (std::sqrt(2.0) * c->p.gaussian_sigma) : 0.0;
    D.vl = cfg->voltage_left;
    D.vr = cfg->voltage_right;
    D.beta = c->p.inverse_kt;
    D.pref = c->p.prefactor;
    // sI[1][0][0] = pref * iR[1][0][0] * eR[1][0][0] (potential.cpp:316-324)
    D.self_int = c->p.prefactor * 1.0 * (c->p.gaussian_sigma > 0 ? std::erf(D.gk_d * 1.0) : 1.0);
    D.binding = c->p.exciton_binding;
    std::memcpy(D.rate, cfg->rate, sizeof(D.rate));
    std::memcpy(D.coupling, cfg->coupling, sizeof(D.coupling));
    const int cap = D.cap_e + D.cap_h;
    D.S = split_count(cap, cap + D.n_d_eff);
    D.ld = cap;
    D.lookahead = 4ull * cap + 8192ull;
    const unsigned long long ring = next_pow2(2ull * (cap + D.lookahead + 4096ull));
    D.ring_mask = ring - 1ull;
    c->de_grid_x = (cap + kBlock - 1) / kBlock;

    for (int b = 0; b < 2; ++b) {
        if ((st = sim_alloc(c, D.e_site[b], D.cap_e)) || (st = sim_alloc(c, D.e_life[b], D.cap_e)) ||
            (st = sim_alloc(c, D.e_path[b], D.cap_e)) || (st = sim_alloc(c, D.h_site[b], D.cap_h)) ||
            (st = sim_alloc(c, D.h_life[b], D.cap_h)) || (st = sim_alloc(c, D.h_path[b], D.cap_h)))
            return st;
    }
    float2* part = nullptr;
    if ((st = sim_alloc(c, D.e_fsite, D.cap_e)) || (st = sim_alloc(c, D.h_fsite, D.cap_h)) ||
        (st = sim_alloc(c, D.etype, Vt)) || (st = sim_alloc(c, D.htype, Vt)) || (st = sim_alloc(c, D.vbg, Vt)) ||
        (st = sim_alloc(c, D.claim_e, d.V)) || (st = sim_alloc(c, D.claim_h, d.V)) ||
        (st = sim_alloc(c, D.h_at, d.V)) || (st = sim_alloc(c, D.code, cap)) ||
        (st = sim_alloc(c, D.removed, cap)) || (st = sim_alloc(c, D.flag, cap)) ||
        (st = sim_alloc(c, D.tgt1, cap)) || (st = sim_alloc(c, D.tgt2, cap)) ||
        (st = sim_alloc(c, D.chg, cap + D.n_d_eff)) || (st = sim_alloc(c, part, size_t(D.S) * D.ld)) ||
        (st = sim_alloc(c, D.de, cap)) || (st = sim_alloc(c, D.ring, ring)) || (st = sim_alloc(c, c->ctl, 1)))
        return st;
    D.part = part;
    D.ctl = c->ctl;

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
    hipLaunchKernelGGL(fill_i32_kernel, dim3((d.V + 255) / 256), dim3(256), 0, c->stream, D.claim_e, d.V, 0x7fffffff);
    hipLaunchKernelGGL(fill_i32_kernel, dim3((d.V + 255) / 256), dim3(256), 0, c->stream, D.claim_h, d.V, 0x7fffffff);
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
    for (int i = 0; i < 10; ++i) { C.att[i] = s->flux_attempts[i]; C.suc[i] = s->flux_successes[i]; }
    C.step = s->current_step;
    HIPCHK(c, hipMemcpy(c->ctl, &C, sizeof(C), hipMemcpyHostToDevice));
    HIPCHK(c, hipDeviceSynchronize());
    hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(kCommitBlock), 0, c->stream, D, 1);
    if ((st = launch_check(c, "commit_kernel(prepare)")) != LMG_OK) return st;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->sim_loaded = true;
    int32_t code = 0;
    if ((st = lmg_sim_device_error(c, &code)) != LMG_OK) return st;
    if (code) return fail(c, LMG_ERR_DEVICE, "device error " + std::to_string(code) + " while preparing");
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
    for (int i = 0; i < n; ++i) {
        if (c->sim.coulomb && (st = sim_launch_de(c)) != LMG_OK) return st;
        hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(kCommitBlock), 0, c->stream, c->sim, 0);
    }
    if ((st = launch_check(c, "commit_kernel")) != LMG_OK) return st;
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
    const size_t need = 2 * size_t(n) + 1;
    while (c->events.size() < need) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreate(&e));
        c->events.push_back(e);
    }
    HIPCHK(c, hipEventRecord(c->events[0], c->stream));
    for (int i = 0; i < n; ++i) {
        if (c->sim.coulomb && (st = sim_launch_de(c)) != LMG_OK) return st;
        HIPCHK(c, hipEventRecord(c->events[2 * i + 1], c->stream));
        hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(kCommitBlock), 0, c->stream, c->sim, 0);
        HIPCHK(c, hipEventRecord(c->events[2 * i + 2], c->stream));
    }
    if ((st = launch_check(c, "commit_kernel")) != LMG_OK) return st;
    HIPCHK(c, hipEventSynchronize(c->events[2 * n]));
    double de = 0.0, cm = 0.0;
    for (int i = 0; i < n; ++i) {
        float a = 0.f, b = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, c->events[2 * i], c->events[2 * i + 1]));
        HIPCHK(c, hipEventElapsedTime(&b, c->events[2 * i + 1], c->events[2 * i + 2]));
        de += a;
        cm += b;
    }
    float span = 0.f;
    HIPCHK(c, hipEventElapsedTime(&span, c->events[0], c->events[2 * n]));
    kernel_ms[0] = de;
    kernel_ms[1] = cm;
    kernel_ms[2] = span;
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
    if (c->sim.coulomb && (st = sim_launch_de(c)) != LMG_OK) return st;
    hipLaunchKernelGGL(commit_kernel, dim3(1), dim3(kCommitBlock), 0, c->stream, c->sim, 2);
    if ((st = launch_check(c, "commit_kernel(dE)")) != LMG_OK) return st;
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
