// lmg_sim.h -- device side of one KMC sweep (Simulation::performIterations,
// src/langmuirCore/simulation.cpp:29-111) as a short chain of kernels on one
// stream, every carrier-parallel phase spread over the whole GPU:
//
//   de_kernel      (grid)   every carrier's proposal (one engine word each),
//                           the dE pair sums of its hop, the decision-draw
//                           flags, and (one extra workgroup) the MT19937
//                           refill for later steps
//   decide_kernel  (grid)   draw offsets (block prefix + in-block scan), dE
//                           assembly, Metropolis / drain decisions, claims
//   recomb_kernel  (1 WG)   RecombinationAgent sweep (solar cell, rate > 0)
//   tick_kernel    (grid)   completeTick conflict rule and order-preserving
//                           compaction of the lists
//   inject_kernel  (1 WG)   injection (serial in the reference), then the
//                           first proposal draw Boost's uniform_int rejects
//                           and the serial proposals from there (rare)
//
// On the culled path cull_emit_kernel makes the proposals and cull_de_kernel
// the pair sums (below).
//
// Offsets into the random stream are computed so that every draw uses exactly
// the word the reference's serial loop would use.
#pragma once

#include "lmg_device.h"

namespace lmg {

constexpr int kSimBlock = 256;   // carriers per workgroup in the grid phases
constexpr int kOneBlock = 1024;  // single-workgroup phases
constexpr int kMaxCullWorkgroups = 65536;   // cull_de_kernel workgroups a caller may ask for

// device error codes (lmg_sim_device_error)
constexpr int kErrRing = 1, kErrCapacity = 2, kErrSite = 3, kErrBalance = 4;

struct SimCtl {
    int n_e, n_h;                 // list sizes of the pending step
    int cur;                      // buffer that holds the lists
    int error;
    int n_e_next, n_h_next;       // sizes after compaction (tick_kernel)
    int counts[2];                // de_kernel: targets, charges
    int first_bad;
    int danger_n;                 // engine outputs in Boost's rejection zone found by mt_refill
    unsigned long long danger[16];// ... the last 16 of them (output indices)
    unsigned long long rng_pos;   // engine outputs consumed before the pending decisions
    unsigned long long rng_step0; // ... before the pending step's proposals
    unsigned long long gen_pos;   // raw words x[0 .. gen_pos) generated
    unsigned long long ring_base; // the first word the ring holds (the loaded window's)
    unsigned long long draws_dec; // decision draws of the pending step
    unsigned long long draws_rec; // recombination draws of the pending step
    unsigned long long att[10], suc[10];
    long long step;
};

// The decision inputs of one carrier that do not depend on its dE, made by
// cull_de_kernel's last workgroups while the pair sums run (decide_prep)
struct DecideRec {
    double u;          // the decision draw's uniform (draw set)
    double pd0;        // q (V_bg(fs) - V_bg(site))
    double cpl;        // coupling of the hop
    double bf, bs;     // binding terms at fs and at site
    int site, fs, life, path;
    int flags;         // bit 0: draws, bits 8..15: own-grid type at fs
    int pad;
};

struct SimDev {
    Dims d;
    int hop, sim_type, coulomb, src_metro, src_coul, balance, rec_range, gauss;
    int max_e, max_h, cap_e, cap_h, n_d_eff, dq;
    int e_drain[2], h_drain[2];
    int uniform;                  // every site has >= 2 neighbours: one proposal draw each
    int nblk;                     // workgroups of kSimBlock carriers at capacity
    uint32_t c2m1;
    float gk;
    // culled dE: the pair function f by r^2 (1/r, or erf(r / (sqrt2 sigma)) / r),
    // FP64 values rounded to FP32, for r^2 = 0 .. c^2 with f = 0 at r^2 = 0 and
    // at c^2; null when c^2 + 1 > kTabMax (the kernel then evaluates f itself)
    const float* rtab;
    uint32_t rtab_lim;            // 4 c^2: byte offset of the last (zero) entry
    int rtab_n;                   // c^2 + 1 entries
    double vl, vr, beta, pref, self_int, binding;
    double rate[10];
    double coupling[27];
    int S, ld;
    int S_pairs;                  // charge chunks of the all-pairs dE kernel
    // culled dE (cull_*_kernel): 8^3-site bricks, targets in groups of 64
    int cull;
    int cut, ngrp;
    int nbx, nby, nbz, nbrick;
    uint32_t* bmask;              // [brick][electrons, holes][16] site bits (cleared by decide_kernel)
    int* scnt;                    // carriers per brick (cull_mark; cleared by decide_kernel)
    int* e_who;                   // [V] list index of the electron on a site (valid where its bit is set)
    int* h_who;                   // [V] the same for holes
    int* sstart;                  // [nbrick + 1] charge list offsets
    int* tstart;                  // [nbrick] first target slot of each brick (bricks in bperm order;
                                  // written only for more than kEmitScanBricks bricks)
    const int* bperm;             // brick order of the target groups (static, set at load)
    const int* brank;             // [nbrick] each brick's rank in bperm order (bperm's inverse)
    const int* dstart;            // [nbrick + 1] static charged-defect list offsets
    const float4* dsrc;           // static charged defects, by brick then site
    float4* src;                  // charge lists of the bricks
    int* torder;                  // targets grouped by brick
    int* tgrp;                    // slots of each target group in the last culled pass (>= 1)
    int4* items;                  // cull_de_kernel's work items: (group, chunk, chunks of the group)
    int* nitems;
    int items_cap;
    long long* isum;              // per target: fixed-point pair sum (int64 atomics)
    DecideRec* drec;              // per carrier: decide_kernel's inputs but dE
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
    unsigned long long* claim_e;  // per site: (step tag << 32) | ~list index, max wins
    unsigned long long* claim_h;
    int* h_at;
    int* code;      // per carrier: 0 stay, 1 accepted to an empty site, 2 accepted into a drain
    int* removed;   // per carrier: recombined this step
    int* flag;      // per carrier: decision draws (0/1)
    int* blk_draw;  // per workgroup: decision draws
    int* blk_keep;  // per workgroup: kept electrons, kept holes
    float4* tgt1;
    float4* tgt2;
    float4* chg;    // the charged defects (carriers are read from their site lists)
    long long* part;              // de_kernel's chunk partials, [S][ld], int64 fixed point
    double* de;
    uint32_t* ring;
    unsigned long long ring_mask;
    SimCtl* ctl;
};

__device__ __forceinline__ void set_error(const SimDev& D, int code) { atomicCAS(&D.ctl->error, 0, code); }

__device__ __forceinline__ uint32_t rng_word(const SimDev& D, unsigned long long k) {
    return mt_temper(D.ring[(k + 624ull) & D.ring_mask]);
}

__device__ __forceinline__ float4 site_f4(const Dims& d, int s, float q) {
    int x, y, z;
    decode(d, s, x, y, z);
    return make_float4(float(x), float(y), float(z), q);
}

__device__ __forceinline__ float rsqf(float r2) { return __builtin_amdgcn_rsqf(r2); }
__device__ __forceinline__ bool cut_in(float r2, uint32_t c2m1) { return (__float_as_uint(r2) - 1u) < c2m1; }
__device__ __forceinline__ void kahan(float& s, float& c, float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// exclusive scan of one int per thread over a workgroup of nw waves (DPP
// wave scans; wsum holds nw ints)
template <int kThreads>
__device__ int wg_scan(int v, int* wsum, int& total) {
    constexpr int nw = kThreads / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int x = wave_incl_sum(v);
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        const int w = wave_incl_sum(lane < nw ? wsum[lane] : 0);
        if (lane < nw) wsum[lane] = w;
    }
    __syncthreads();
    const int excl = x - v + (wid > 0 ? wsum[wid - 1] : 0);
    total = wsum[nw - 1];
    __syncthreads();
    return excl;
}

// two exclusive scans at once (wsum holds 2 nw ints)
template <int kThreads>
__device__ void wg_scan2(int a, int b, int* wsum, int& ea, int& eb, int& ta, int& tb) {
    constexpr int nw = kThreads / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int xa = wave_incl_sum(a), xb = wave_incl_sum(b);
    if (lane == 63) {
        wsum[wid] = xa;
        wsum[nw + wid] = xb;
    }
    __syncthreads();
    if (wid == 0) {
        const int wa = wave_incl_sum(lane < nw ? wsum[lane] : 0);
        const int wb = wave_incl_sum(lane < nw ? wsum[nw + lane] : 0);
        if (lane < nw) {
            wsum[lane] = wa;
            wsum[nw + lane] = wb;
        }
    }
    __syncthreads();
    ea = xa - a + (wid > 0 ? wsum[wid - 1] : 0);
    eb = xb - b + (wid > 0 ? wsum[nw + wid - 1] : 0);
    ta = wsum[nw - 1];
    tb = wsum[2 * nw - 1];
    __syncthreads();
}

// Workgroup prefix over per-workgroup counts, sum of arr[stride * b + off]
// for b < nb, in two halves so that a kernel can issue its per-carrier loads
// between them: the first wave's lane sums (loads only, no barrier), then
// their total, shared with the workgroup (a barrier)
__device__ __forceinline__ long long wg_prefix_loads(const int* arr, int nb, int stride, int off) {
    long long s = 0;
    if (threadIdx.x < 64)
        for (int b = threadIdx.x; b < nb; b += 64) s += gmem(arr)[stride * b + off];
    return s;
}
// (the counts are carriers: their sums fit an int)
__device__ __forceinline__ long long wg_prefix_total(long long s, long long* slot) {
    if (threadIdx.x < 64) {
        const int t = wave_total(wave_incl_sum(int(s)));
        if (threadIdx.x == 0) *slot = t;
    }
    __syncthreads();
    return *slot;
}

// ---------------------------------------------------------------------------
// MT19937 refill: x[m] = x[m-227] ^ twist(x[m-624], x[m-623]), so 227
// consecutive words can be produced at once from the last 624 (623 with the
// recurrence expanded, below).
// ---------------------------------------------------------------------------
// Boost's uniform_int(0, r) rejects a word w iff w >= (r + 1) * bucket(r), which is at
// least 2^32 - 1 - r.  A neighbour list holds at most 34 entries (hopping.range 2 in
// three dimensions, 32, plus two drain slots), so no proposal draw below this can be
// rejected; mt_refill lists the engine outputs at or above it.
constexpr uint32_t kRejectZone = 0xffffffffu - 33u;
constexpr int kDangerSlots = 16;

// The recurrence x[m] = x[m-227] ^ tw(x[m-624], x[m-623]) expanded twice,
//   x[m] = x[m-681] ^ tw(x[m-1078], x[m-1077]) ^ tw(x[m-851], x[m-850]) ^ tw(x[m-624], x[m-623]),
// reaches back at least 623 words, so a round makes 623 words at once (454 with
// one expansion, 227 with none, while the ring holds less history).
//
// The last 2048 words sit in LDS twice over: word m at slots m mod 2048 and
// m mod 2048 + 2048.  Every read of a round, m - k for 227 <= k <= 1078, is then
// slot (m mod 2048) + 2048 - k, one base address per word and fixed offsets
// (no wrap, no per-read address arithmetic).  Within a round the slots written
// and the slots read are never the same: the words are less than 2048 apart.
__device__ __forceinline__ uint32_t mt_tw(uint32_t x0, uint32_t x1) {
    const uint32_t y = (x0 & 0x80000000u) | (x1 & 0x7fffffffu);
    return (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}
constexpr uint32_t kTailMask = 2047u;
constexpr int kTailWords = 4096;   // LDS words of the refill's window (two copies)
// A workgroup barrier that orders LDS only: __syncthreads() also waits for
// every outstanding global store (its release fence covers global memory),
// which would cost each round a store round trip for the ring words.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// One round of kF + 1 expansions (kF = 0: 227 words, 1: 454, 2: 623) by a
// workgroup of kThreads: each thread issues the LDS reads of all its words
// before any arithmetic, so that a round costs one LDS latency for its reads
// and one for its writes.  g = gen mod 2^32 (the ring holds fewer than 2^32
// words, so ring and tail indices follow from the low 32 bits).
typedef __attribute__((address_space(1))) uint32_t gu32;
template <int kThreads, int kF>
__device__ __forceinline__ void mt_round(const SimDev& D, uint32_t* tail, uint32_t g, gu32* ring, uint32_t rmask) {
    constexpr int kW = kF == 2 ? 623 : (kF == 1 ? 454 : 227);
    constexpr int kPer = (kW + kThreads - 1) / kThreads;
    constexpr int kR = 3 + 2 * kF;   // LDS reads per word
    constexpr int kOff[7] = {624, 623, kF == 2 ? 681 : (kF == 1 ? 454 : 227), 851, 850, 1078, 1077};
    // (unconditional: the slots of a thread past the round's words are inside the
    // window too, and their values go unused; one basic block, so that every
    // read is issued before the first wait)
    uint32_t a[kPer][7];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const uint32_t* w = tail + ((g + threadIdx.x + uint32_t(kThreads * k)) & kTailMask) + 2048;
#pragma unroll
        for (int r = 0; r < kR; ++r) a[k][r] = w[-kOff[r]];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int t = int(threadIdx.x) + kThreads * k;
        if (t < kW) {
            const uint32_t m = g + uint32_t(t);
            uint32_t v = mt_tw(a[k][0], a[k][1]) ^ a[k][2];
            if (kF >= 1) v ^= mt_tw(a[k][3], a[k][4]);
            if (kF == 2) v ^= mt_tw(a[k][5], a[k][6]);
            tail[m & kTailMask] = v;
            tail[(m & kTailMask) + 2048] = v;
            ring[m & rmask] = v;   // a global store: a flat one would hold up every LDS wait
            if (mt_temper(v) >= kRejectZone) {   // engine output m - 624 (about one in 10^8)
                const int slot = atomicAdd(&D.ctl->danger_n, 1);
                // (the full index: gen_pos - (gen_pos - m) mod 2^32)
                const unsigned long long gp = D.ctl->gen_pos;
                D.ctl->danger[slot % kDangerSlots] = gp + (unsigned long long)(uint32_t(m - uint32_t(gp))) - 624ull;
            }
        }
    }
}
// Refill up to `target` with rounds of at most kMaxF + 1 expansions: a
// workgroup that shares its CU with LDS-bound pair sums runs the 227-word
// rounds (three LDS reads a word) faster than the wide ones (seven).
template <int kThreads, int kMaxF>
__device__ void mt_refill(const SimDev& D, uint32_t* tail /* LDS [kTailWords] */, unsigned long long target) {
    unsigned long long gen = D.ctl->gen_pos;
    if (gen >= target) return;
    const unsigned long long base = D.ctl->ring_base;
    gu32* ring = (gu32*)D.ring;   // (address-space cast)
    const uint32_t rmask = uint32_t(D.ring_mask);
    for (unsigned long long m = gen - min(gen - base, 1078ull) + threadIdx.x; m < gen; m += kThreads) {
        const uint32_t v = ring[uint32_t(m) & rmask];
        tail[m & kTailMask] = v;
        tail[(m & kTailMask) + 2048] = v;
    }
    lds_barrier();
    while (gen < target) {
        const unsigned long long hist = gen - base;
        if (kMaxF >= 2 && hist >= 1078ull) {
            mt_round<kThreads, 2>(D, tail, uint32_t(gen), ring, rmask);
            gen += 623ull;
        } else if (kMaxF >= 1 && hist >= 851ull) {
            mt_round<kThreads, 1>(D, tail, uint32_t(gen), ring, rmask);
            gen += 454ull;
        } else {
            mt_round<kThreads, 0>(D, tail, uint32_t(gen), ring, rmask);
            gen += 227ull;
        }
        lds_barrier();
    }
    if (threadIdx.x == 0) D.ctl->gen_pos = gen;
}

// ===========================================================================
// The pending step's proposals (chooseFuture, chargeagent.cpp:73-78), made by
// the first kernel of the step that needs them.  With one draw per carrier
// (every site has two or more neighbours), carrier i (electrons, then holes)
// draws engine output rng_step0 + i unless a Boost uniform_int rejection
// happened before it; inject_kernel finds the first rejecting carrier
// (first_bad) and makes the proposals from there on serially, into fsite.
// ===========================================================================
struct Pending {
    int site, fs;
    float4 p1, p2;   // (x, y, z, q) at the current and the proposed site
};

__device__ __forceinline__ Pending pending_proposal(const SimDev& D, int i, int ne, int cur,
                                                    unsigned long long step0, int first_bad) {
    const bool isE = i < ne;
    const int li = isE ? i : i - ne;
    const float q = isE ? -1.f : 1.f;
    Pending P;
    P.site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
    if (P.site < 0 || P.site >= D.d.V) {   // never written by the step kernels: a corrupt state
        set_error(D, kErrSite);
        P.site = 0;
    }
    int x, y, z, fx, fy, fz;
    decode(D.d, P.site, x, y, z);
    if (i < first_bad) {
        const int dl = isE ? D.e_drain[0] : D.h_drain[0];
        const int dr = isE ? D.e_drain[1] : D.h_drain[1];
        const uint32_t w = rng_word(D, step0 + (unsigned long long)i);
        if (D.hop == 1) {
            const int n = nbr1_count(D.d, x, y, z, dl, dr);
            const int k = min(int(w / uint_bucket(uint32_t(n - 1))), n - 1);
            P.fs = nbr1_pick(D.d, x, y, z, dl, dr, k, fx, fy, fz);
        } else {
            const int n = nbr_count(D.d, D.hop, P.site, dl, dr);
            const int k = min(int(w / uint_bucket(uint32_t(n - 1))), n - 1);
            P.fs = nbr_at(D.d, D.hop, P.site, dl, dr, k);
            decode(D.d, P.fs, fx, fy, fz);
        }
    } else {
        P.fs = isE ? D.e_fsite[li] : D.h_fsite[li];
        decode(D.d, P.fs, fx, fy, fz);   // a drain slot V + k decodes past the lattice (kernel.cl:229-231)
    }
    P.p1 = make_float4(float(x), float(y), float(z), q);
    P.p2 = make_float4(float(fx), float(fy), float(fz), q);
    return P;
}

// charge j of the all-pairs list [charged defects | electrons | holes]
__device__ __forceinline__ float4 charge_f4(const SimDev& D, int j, int nd, int ne, int cur) {
    if (j < nd) return D.chg[j];
    j -= nd;
    return j < ne ? site_f4(D.d, D.e_site[cur][j], -1.f) : site_f4(D.d, D.h_site[cur][j - ne], 1.f);
}

// ===========================================================================
// de_kernel: grid = (nblk + 1, S, replicas) (the largest replica's extents;
// smaller replicas exit surplus workgroups).  Workgroup (x, s) sums charge
// chunk s for carriers [256x, 256x + 256); chunk-0 workgroups also flag the
// carriers whose decision draws (own-grid type Empty or Drain at the proposed
// site, chargeagent.cpp:95-157).  Workgroup (nblk, 0) refills the random
// stream.
//
// For each charge j it accumulates q_j ( [0<r2<c2] f(r2) - [0<r1<c2] f(r1) ),
// the pair part of p2 - p1 in ChargeAgent::coulombCPU (chargeagent.cpp:215-271).
// For 1/r, two in-cutoff terms are differenced without cancellation:
//   1/sqrt(r2) - 1/sqrt(r1) = (r1 - r2) (ab)^2 / (a + b),  a = r1^-1/2, b = r2^-1/2,
// with r1 - r2 an exact integer.
// ===========================================================================
// The hop h = p2 - p1 of a target enters a term through
//   |p2 - q|^2 = r1 + t,  t = 2 h.(p1 - q) + |h|^2 = K - 2 h.q,  K = 2 h.p1 + |h|^2,
// all exact integers in FP32; HopTerm holds -2h and K for one target.
struct HopTerm {
    float ux, uy, uz, k;
};
__device__ __forceinline__ HopTerm hop_term(const float4& p1, const float4& p2) {
    const float hx = p2.x - p1.x, hy = p2.y - p1.y, hz = p2.z - p1.z;
    const float k = 2.f * (hx * p1.x + hy * p1.y + hz * p1.z) + (hx * hx + hy * hy + hz * hz);
    return HopTerm{-2.f * hx, -2.f * hy, -2.f * hz, k};
}

// [0<r2<c2] f(r2) - [0<r1<c2] f(r1) for one charge q, target pair (p1, p1+h)
template <bool kGauss>
__device__ __forceinline__ float de_term(const float4& p1, const float4& q, const HopTerm& h, uint32_t c2m1,
                                         float gk) {
    const float dx = p1.x - q.x, dy = p1.y - q.y, dz = p1.z - q.z;
    const float r1 = dx * dx + dy * dy + dz * dz;
    const float t = fmaf(h.ux, q.x, fmaf(h.uy, q.y, fmaf(h.uz, q.z, h.k)));
    const float r2 = r1 + t;
    const bool in1 = cut_in(r1, c2m1), in2 = cut_in(r2, c2m1);
    const float a = rsqf(r1), b = rsqf(r2);
    if (kGauss) {
        const float fa = erff(gk * (r1 * a)) * a;
        const float fb = erff(gk * (r2 * b)) * b;
        return (in2 ? fb : 0.f) - (in1 ? fa : 0.f);
    }
    // b - a = (r1 - r2) / (sqrt(r1) sqrt(r2) (sqrt(r1) + sqrt(r2))) = -t / (r1 r2 (a + b));
    // one term in cutoff: b or -a (three selects, no subtraction)
    const float both = -t * __builtin_amdgcn_rcpf((r1 * r2) * (a + b));
    const float one = in2 ? b : (in1 ? -a : 0.f);
    return (in1 & in2) ? both : one;
}

// Fixed point of the pair sums: every FP32 sum of a summation unit (a block of
// 32 charges here, a 64-slot piece of a brick list in cull_de_kernel) is
// scaled by 2^40, rounded to an integer and added exactly, so a carrier's total
// does not depend on how the units are split over workgroups or in what order
// they are added (at most 2^-41 per unit, about 1e-10 eV per carrier).
//
// The rounding is one FP64 fma: acc 2^40 + 1.5 2^52 lies in [2^52, 2^53) for
// |acc| < 2^11, where the double's spacing is 1, so its bit pattern is
// 0x4338000000000000 + round(acc 2^40).  A unit holds at most 64 terms of
// magnitude <= |q| (both distances >= 1), so |acc| < 2^11 for charges up to
// 31 e (lmg_create refuses larger defect charges).
constexpr double kFixScale = 1099511627776.0;   // 2^40
constexpr double kFixInv = 1.0 / kFixScale;
constexpr double kFixMagic = 6755399441055744.0;   // 1.5 2^52
__device__ __forceinline__ long long fix_round(float acc) {
    const unsigned long long b = __double_as_longlong(fma(double(acc), kFixScale, kFixMagic));
    return (long long)(b - 0x4338000000000000ull);
}

// ---------------------------------------------------------------------------
// Table form of the dE terms (both dE kernels).  Every squared distance on the lattice
// is an integer, so f(r) is a table lookup by r^2: the in-cutoff test, the
// r = 0 exclusion and (for the erf-screened pair function) erf all fold into
// the table, whose entries are 0 at r^2 = 0 and at c^2, and indices beyond
// c^2 are clamped to c^2.  The workgroup copies the table into LDS.
//
// 4 r^2 comes out exact from integer dot products.  Coordinates are taken
// relative to an origin o (the centre of the group's box); a charge q is
// staged as the record
//   w0 = (-8 qx, -8 qy), w1 = (-8 qz, 4 |q|^2) as int16 pairs, w2 = 4 |q|^2,
//   w3 = its charge (FP32 bits)
// and a target p holds (px, py), (pz, 1), (pz, 0) as int16 pairs and 4 |p|^2:
//   4 |p - q|^2 = 4|p|^2 + 4|q|^2 - 8 p.q
//              = dot2((px, py), w0, dot2((pz, 1), w1, 4|p|^2))      if 4|q|^2 < 2^15,
//              = dot2((px, py), w0, dot2((pz, 0), w1, w2 + 4|p|^2)) always
// (v_dot2_i32_i16: two int16 products plus an int32).  Both forms give the same
// integer; the short one needs every kept charge within 90 sites of o.  A term
// is then f[4r2^2] - f[4r1^2] (two LDS reads, byte offsets), times q.
// The table entries are FP64 values rounded once, so a term's error is at most
// that of two rounded values; over a 64-term unit the FP32 accumulation
// dominates, as in the arithmetic form.
// ---------------------------------------------------------------------------
constexpr int kTabMax = 4096;   // table entries (16 KB of LDS): cutoffs up to 63
typedef short v2s16 __attribute__((ext_vector_type(2)));
struct TabTarget {
    v2s16 xy, z1, z0;
    int n4;
};
__device__ __forceinline__ v2s16 pk16(int lo, int hi) {
    return __builtin_bit_cast(v2s16, (uint32_t(lo) & 0xffffu) | (uint32_t(hi) << 16));
}
// n4 = 4 |p|^2 plus the table's LDS byte address, so that the dot products give
// the entry's address directly
__device__ __forceinline__ TabTarget tab_target(const float4& p, int ox, int oy, int oz, uint32_t base) {
    const int x = int(p.x) - ox, y = int(p.y) - oy, z = int(p.z) - oz;
    return TabTarget{pk16(x, y), pk16(z, 1), pk16(z, 0), int(base) + 4 * (x * x + y * y + z * z)};
}
// v_dot2_i32_i16 in its three-operand (VOP3P) form, d = a.lo b.lo + a.hi b.hi + c.
// (__builtin_amdgcn_sdot2 is selected as the accumulate-in-place form, which
// costs a copy of a loop-invariant accumulator per use.)  An ALU instruction
// only: no memory access.
__device__ __forceinline__ int sdot2(v2s16 a, int b, int c) {
    int d;
    asm("v_dot2_i32_i16 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}
template <bool kShort>
__device__ __forceinline__ int tab_r2x4(const TabTarget& p, int w0, int w1, int w2) {
    const int zq = kShort ? sdot2(p.z1, w1, p.n4) : sdot2(p.z0, w1, w2 + p.n4);
    return sdot2(p.xy, w0, zq);
}
// A staged record (w0, w1, w2, charge bits) read from LDS in narrow pieces: w0 and w1
// as 64 bits and the charge as 32 (the short form has no use for w2), or two
// 64-bit halves.  Every lane reads the same record; among the table reads
// (about 2.4 CU cycles per 32-bit read at this spread of addresses) such a
// 128-bit broadcast read costs about 6 CU cycles on gfx950 and the 64 + 32-bit
// pair about 3.
typedef int rec_v4i __attribute__((ext_vector_type(4)));
template <bool kShort>
__device__ __forceinline__ rec_v4i rec_load(const rec_v4i* L, int t) {
    typedef int v2i __attribute__((ext_vector_type(2)));
    const v2i a = reinterpret_cast<const v2i*>(L + t)[0];
    if (kShort) return rec_v4i{a.x, a.y, 0, reinterpret_cast<const int*>(L + t)[3]};
    const v2i b = reinterpret_cast<const v2i*>(L + t)[1];
    return rec_v4i{a.x, a.y, b.x, b.y};
}
// f at LDS byte address a = base + 4 r^2 (beyond the cutoff, the last entry at lim = base + 4 c^2)
__device__ __forceinline__ float tab_f(int a, uint32_t lim) {
    typedef const __attribute__((address_space(3))) float* lds_cfp;
    return *reinterpret_cast<lds_cfp>(uintptr_t(min(uint32_t(a), lim)));
}

// kTable: the pair function from the LDS table of f by r^2 (every replica of the
// launch has one), 4 r^2 from the charge records as in cull_de_kernel (the long
// form, coordinates from the lattice origin); otherwise 1/r or the erf-screened
// form evaluated here.  Dynamic LDS: kDeLdsBytes, the table, or the refill's
// window in the refill workgroup.
constexpr int kDeLdsBytes = kTabMax * 4;
static_assert(kDeLdsBytes >= kTailWords * 4, "the refill's window fits de_kernel's dynamic LDS");
template <bool kGauss, bool kPairs, bool kTable>
__global__ __launch_bounds__(kSimBlock) void de_kernel(const SimDev* __restrict__ Dp) {
    const SimDev& D = Dp[blockIdx.z];     // replica
    __shared__ float4 tile[256];
    extern __shared__ float de_lds[];
    if (int(blockIdx.x) >= D.nblk || int(blockIdx.y) >= D.S) {
        if (int(blockIdx.x) == D.nblk && blockIdx.y == 0)
            mt_refill<kSimBlock, 2>(D, reinterpret_cast<uint32_t*>(de_lds), D.ctl->rng_pos + 624ull + D.lookahead);
        return;
    }
    const SimCtl& C = *D.ctl;
    const int ne = C.n_e, nt = C.counts[0], nc = C.counts[1], cur = C.cur, nd = D.n_d_eff;
    if (int(blockIdx.x) * kSimBlock >= nt) return;   // uniform per workgroup
    const int i = blockIdx.x * kSimBlock + threadIdx.x;
    // the target's proposal (every charge chunk makes it; chunk 0 records it)
    const int fb = C.first_bad;
    Pending P;
    P.p1 = P.p2 = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < nt) P = pending_proposal(D, i, ne, cur, C.rng_step0, fb);
    if (blockIdx.y == 0) {
        int draw = 0;
        if (i < nt) {
            const bool isE = i < ne;
            if (i < fb) (isE ? D.e_fsite : D.h_fsite)[isE ? i : i - ne] = P.fs;
            const uint8_t tf = isE ? D.etype[P.fs] : D.htype[P.fs];
            draw = (tf == kEmpty || tf == kDrain);
            D.flag[i] = draw;
        }
        const int cnt = __syncthreads_count(draw);
        if (threadIdx.x == 0) D.blk_draw[blockIdx.x] = cnt;
    }
    if (!kPairs) return;
    // chunk s: whole blocks of 32 charges [b0, b1) out of nb32
    const int S = D.S, nb32 = (nc + 31) / 32;
    const int j0 = 32 * int((long long)blockIdx.y * nb32 / S);
    const int j1 = min(nc, 32 * int((long long)(blockIdx.y + 1) * nb32 / S));
    const float4 p1 = P.p1;
    const HopTerm ht = hop_term(P.p1, P.p2);
    const uint32_t c2m1 = D.c2m1;
    const float gk = D.gk;
    // table form: the table into LDS (published by the first tile's barrier)
    typedef __attribute__((address_space(3))) float* lds_fp;
    const uint32_t tbase = uint32_t(uintptr_t((lds_fp)de_lds));   // (address-space cast)
    const uint32_t tlim = tbase + D.rtab_lim;
    const TabTarget ta1 = tab_target(P.p1, 0, 0, 0, tbase), ta2 = tab_target(P.p2, 0, 0, 0, tbase);
    if (kTable)
        for (int k = threadIdx.x; k < D.rtab_n; k += kSimBlock) de_lds[k] = gmem(D.rtab)[k];
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i* rec = reinterpret_cast<const v4i*>(tile);
    long long acc64 = 0;
    for (int base = j0; base < j1; base += 256) {
        const int n = min(256, j1 - base);
        __syncthreads();
        if (threadIdx.x < n) {
            const float4 q = charge_f4(D, base + threadIdx.x, nd, ne, cur);
            if (kTable) {   // the record of cull_de_kernel's table form, from the origin
                const int qx = int(q.x), qy = int(q.y), qz = int(q.z);
                const int n4 = 4 * (qx * qx + qy * qy + qz * qz);
                v4i w;
                w.x = int((uint32_t(-8 * qx) & 0xffffu) | (uint32_t(-8 * qy) << 16));
                w.y = int(uint32_t(-8 * qz) & 0xffffu);
                w.z = n4;
                w.w = __float_as_int(q.w);
                reinterpret_cast<v4i*>(tile)[threadIdx.x] = w;
            } else {
                tile[threadIdx.x] = q;
            }
        }
        __syncthreads();
        int k0 = 0;
        for (; k0 + 32 <= n; k0 += 32) {   // full blocks: unrolled, branch-free
            float acc = 0.f;
#pragma unroll 8
            for (int k = k0; k < k0 + 32; ++k) {
                if (kTable) {
                    const v4i w = rec_load<false>(rec, k);
                    acc = fmaf(__int_as_float(w.w), tab_f(tab_r2x4<false>(ta2, w.x, w.y, w.z), tlim) -
                                                        tab_f(tab_r2x4<false>(ta1, w.x, w.y, w.z), tlim), acc);
                } else {
                    acc = fmaf(tile[k].w, de_term<kGauss>(p1, tile[k], ht, c2m1, gk), acc);
                }
            }
            acc64 += fix_round(acc);
        }
        if (k0 < n) {   // the list's last, partial block
            float acc = 0.f;
            for (int k = k0; k < n; ++k) {
                if (kTable) {
                    const v4i w = rec_load<false>(rec, k);
                    acc = fmaf(__int_as_float(w.w), tab_f(tab_r2x4<false>(ta2, w.x, w.y, w.z), tlim) -
                                                        tab_f(tab_r2x4<false>(ta1, w.x, w.y, w.z), tlim), acc);
                } else {
                    acc = fmaf(tile[k].w, de_term<kGauss>(p1, tile[k], ht, c2m1, gk), acc);
                }
            }
            acc64 += fix_round(acc);
        }
    }
    if (i < nt) D.part[blockIdx.y * D.ld + i] = acc64;
}

// ===========================================================================
// Culled dE (cutoff well inside the lattice): the same pair sums as de_kernel
// over only the charges near each target.
//
// The lattice is cut into bricks of 8 x 8 x 8 sites (partial at the far faces).
// Every step the charges are listed brick by brick in a fixed order -- in a
// brick: electrons, holes (each by site, from occupancy bitmasks), then the
// static charged defects.  Targets are grouped 64 to a wave in the order of
// their bricks (bperm: the lattice's thin axes fastest, a Hilbert curve or
// Morton order over the long ones), so a group's current and proposed sites
// fill a small box.
//
//   cull_mark (tick_kernel for the kept carriers at their new sites,
//     inject_kernel for the injected ones, or for all after a load): each
//     carrier's bitmask bit, its brick's carrier count and its list index at
//     its site.
//   inject_kernel: the counts become list offsets (charges in linear brick
//     order, targets in bperm order).
//   cull_emit_kernel: one wave per brick writes its charge list and puts the
//     brick's carriers (e_who/h_who of each set bit) into its target slots in
//     the same order (electrons, then holes, each by site); the other
//     workgroups make the proposals and set the decision-draw flags; the last
//     one writes cull_de_kernel's work items.
//   cull_de_kernel: the pair sums of every target over the lists of the bricks
//     near its group (below), as int64 fixed-point totals.
//   decide_kernel (committing pass) clears the bitmask words and brick
//     counters each carrier set.
// ===========================================================================
constexpr int kBrick = 8;          // brick edge (sites)
constexpr int kCullGroup = 64;     // targets per group: one per lane of a wave
constexpr int kMaxCullBricks = 1 << 16;

__device__ __forceinline__ int cull_brick(const SimDev& D, int x, int y, int z) {
    return (x >> 3) + D.nbx * ((y >> 3) + D.nby * (z >> 3));
}
// bitmask word of a carrier (species 0 electrons, 1 holes) and its bit
__device__ __forceinline__ int cull_word(const SimDev& D, int x, int y, int z, int sp, uint32_t& bit) {
    const int j = (x & 7) + 8 * (y & 7) + 64 * (z & 7);
    bit = 1u << (j & 31);
    return (cull_brick(D, x, y, z) * 2 + sp) * 16 + (j >> 5);
}

__device__ __forceinline__ int load_agent(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// a carrier (list index li) at its current site into its brick: bitmask bit,
// carrier count (neither atomic returns a value) and the site's list index
__device__ __forceinline__ void cull_mark(const SimDev& D, bool isE, int li, int site) {
    int x, y, z;
    decode(D.d, site, x, y, z);
    uint32_t bit;
    __hip_atomic_fetch_or(gst(&D.bmask[cull_word(D, x, y, z, isE ? 0 : 1, bit)]), bit, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(gst(&D.scnt[cull_brick(D, x, y, z)]), 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *gst(&(isE ? D.e_who : D.h_who)[site]) = li;
}

// Offsets of the charge lists (linear brick order; sst[0 .. nb]) and of the
// targets (bricks in bperm order; tst[brick]) by one workgroup of kThreads;
// thread t scans a contiguous range of bricks, its first kOffPer bricks' static
// data loaded together (cull_offsets_load).  A brick's carriers (its targets) are the marks of its
// counter plus the carriers listed in inj_b (not yet counted there); its
// charges add the static defects.  With cnt (LDS, nbrick <= kStageBricks) the
// counters are read once, coalesced, and the injected carriers added there, so
// that the bperm-ordered scan reads LDS.  kAgent reads the counters at agent
// scope (needed when this workgroup marked them).
constexpr int kStageBricks = 8192;
constexpr int kOffPer = 8;
// Lattices of up to kEmitScanBricks bricks take their offsets in cull_emit_kernel
// instead: each of its brick workgroups sums the counters before its own bricks
// (kEmitScanBricks / kSimBlock counters and ranks per thread, all loads in
// flight together), so the serial scan leaves inject_kernel's one workgroup.
constexpr int kEmitScanBricks = 2048;
// The loads of the offsets, issued before the injections (no barrier): the
// thread's first kOffPer bricks' bperm entries and defect counts, and with cnt
// the counters into LDS.
// The loads are global (not flat) and the brick counters stay in registers
// until cull_offsets stages them in LDS, so that a workgroup barrier between
// the two (inject_kernel's injections) does not wait for them.
struct OffsetLoads {
    int bp[kOffPer], dn[kOffPer];
    int cv[kStageBricks / kOneBlock];   // brick counters b = threadIdx.x + kOneBlock u (cnt staging)
};
template <int kThreads, bool kAgent>
__device__ void cull_offsets_load(const SimDev& D, int* cnt, OffsetLoads& L) {
    static_assert(kThreads == kOneBlock, "one counter slot per kOneBlock bricks");
    const int nb = D.nbrick, per = (nb + kThreads - 1) / kThreads;
    const int b0 = min(nb, int(threadIdx.x) * per);
#pragma unroll
    for (int u = 0; u < kOffPer; ++u) {   // unconditional (clamped index): all in flight together
        const int b = min(b0 + u, nb - 1);
        L.bp[u] = gmem(D.bperm)[b];
        L.dn[u] = gmem(D.dstart)[b + 1] - gmem(D.dstart)[b];
    }
    if (cnt) {
#pragma unroll
        for (int u = 0; u < kStageBricks / kThreads; ++u) {
            const int b = int(threadIdx.x) + kThreads * u;
            L.cv[u] = b < nb ? (kAgent ? load_agent(&D.scnt[b]) : gmem(D.scnt)[b]) : 0;
        }
    }
}
template <int kThreads, bool kAgent>
__device__ void cull_offsets(const SimDev& D, int* sst, int* tst, int* wsum, const int* inj_b, int inj_n,
                             int* cnt, const OffsetLoads& L) {
    constexpr int kPer = kOffPer;
    const int nb = D.nbrick, per = (nb + kThreads - 1) / kThreads;
    const int b0 = min(nb, int(threadIdx.x) * per), b1 = min(nb, b0 + per);
    auto ld = [&](const int* p) { return kAgent ? load_agent(p) : *p; };
    const int* bp = L.bp;
    const int* dn = L.dn;
    if (cnt) {
#pragma unroll
        for (int u = 0; u < kStageBricks / kThreads; ++u) {
            const int b = int(threadIdx.x) + kThreads * u;
            if (b < nb) cnt[b] = L.cv[u];
        }
        __syncthreads();   // the staged counters
        for (int k = threadIdx.x; k < inj_n; k += kThreads) atomicAdd(&cnt[inj_b[k]], 1);
        __syncthreads();
    }
    auto carriers = [&](int b) {
        if (cnt) return cnt[b];
        int n = ld(&D.scnt[b]);
        for (int k = 0; k < inj_n; ++k) n += inj_b[k] == b;
        return n;
    };
    int sc[kPer], tc[kPer];
    int srun = 0, trun = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        sc[u] = carriers(min(b0 + u, nb - 1)) + dn[u];
        tc[u] = carriers(bp[u]);
        if (b0 + u >= b1) sc[u] = tc[u] = 0;
        srun += sc[u];
        trun += tc[u];
    }
    auto scount = [&](int b) { return carriers(b) + D.dstart[b + 1] - D.dstart[b]; };
    for (int b = b0 + kPer; b < b1; ++b) {
        srun += scount(b);
        trun += carriers(D.bperm[b]);
    }
    int so, to, soff, toff;
    wg_scan2<kThreads>(srun, trun, wsum, so, to, soff, toff);
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int b = b0 + u;
        if (b < b1) {
            sst[b] = so;
            tst[bp[u]] = to;
            so += sc[u];
            to += tc[u];
        }
    }
    for (int b = b0 + kPer; b < b1; ++b) {
        const int q = D.bperm[b];
        sst[b] = so;
        tst[q] = to;
        so += scount(b);
        to += carriers(q);
    }
    if (threadIdx.x == 0) sst[nb] = soff;
}

// Work items of cull_de_kernel, one workgroup of kSimBlock: group g gets
// Q_g = 1 + floor(K C_{g+1} / W) - floor(K C_g / W) workgroups of four chunks,
// C_g the sum of T over the groups before g, T_g its slots in the previous
// culled pass (the groups' regions move by about a site per step), W the sum of
// all T_g and K the workgroups left after one per group: every item slot is
// used, and each group's share is K T_g / W rounded up or down.  Items are
// (g, quad, Q_g) in group order.  Any Q_g gives the same sums: it only
// balances the work.
__device__ void cull_items(const SimDev& D, int* wsum) {
    const int ng = (D.ctl->counts[0] + kCullGroup - 1) / kCullGroup;
    const int per = (ng + kSimBlock - 1) / kSimBlock;
    const int g0 = min(ng, int(threadIdx.x) * per), g1 = min(ng, g0 + per);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    long long wl = 0;
    for (int g = g0; g < g1; ++g) wl += D.tgrp[g];
    // exclusive scan of the threads' sums over the workgroup (64-bit)
    long long x = wl;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    __shared__ long long wtot[4];
    if (lane == 63) wtot[wid] = x;
    __syncthreads();
    long long before = 0;
    for (int w = 0; w < wid; ++w) before += wtot[w];
    const long long W = max(1ll, wtot[0] + wtot[1] + wtot[2] + wtot[3]);
    const long long K = max(0, D.items_cap / 4 - ng);
    const long long c0 = before + x - wl;
    auto cut = [&](long long c) { return (K * c) / W; };
    int run = 0;
    long long c = c0;
    for (int g = g0; g < g1; ++g) {
        const long long t = D.tgrp[g];
        run += 1 + int(cut(c + t) - cut(c));
        c += t;
    }
    int total;
    int k = wg_scan<kSimBlock>(run, wsum, total);
    c = c0;
    for (int g = g0; g < g1; ++g) {
        const long long t = D.tgrp[g];
        const int qg = 1 + int(cut(c + t) - cut(c));
        c += t;
        for (int q = 0; q < qg; ++q) D.items[k + q] = make_int4(g, q, qg, 0);
        k += qg;
    }
    // item slots past the last (none while items_cap / 4 >= the group count)
    // tell their workgroups to exit
    for (int q = total + int(threadIdx.x); q < D.items_cap / 4; q += kSimBlock) D.items[q] = make_int4(-1, 0, 1, 0);
    if (threadIdx.x == 0) *D.nitems = total;
}

// The culled kernels of many replicas run as one launch: workgroup x of the
// launch is workgroup map[x].y of replica map[x].x (descriptor Dp[map[x].x]).
// A single context passes no map (replica 0, workgroup blockIdx.x).
__device__ __forceinline__ int2 cull_workgroup(const int2* __restrict__ map) {
    return map ? map[blockIdx.x] : make_int2(0, int(blockIdx.x));
}

// per replica ceil(nbrick / 4) + nblk + 1 workgroups (the last one: the work items);
// a brick workgroup's four waves copy four bricks' carriers and defects into
// their charge lists and their carriers into the target order
__global__ __launch_bounds__(kSimBlock) void cull_emit_kernel(const SimDev* __restrict__ Dp,
                                                              const int2* __restrict__ map) {
    const int2 wg = cull_workgroup(map);
    const SimDev& D = Dp[wg.x];
    const int wx = wg.y;
    const int* sst = D.sstart;
    const int* tst = D.tstart;
    const int nbw = (D.nbrick + 3) / 4;
    if (wx == nbw + D.nblk) {   // cull_de_kernel's work items
        __shared__ int wsum[4];
        cull_items(D, wsum);
        return;
    }
    if (wx >= nbw) {   // targets: proposals, decision-draw flags, their brick's slots
        const SimCtl& C = *D.ctl;
        const int ne = C.n_e, nt = C.counts[0];
        const int i = (wx - nbw) * kSimBlock + threadIdx.x;
        int draw = 0;
        if (i < nt) {
            const bool isE = i < ne;
            const int li = isE ? i : i - ne;
            const Pending P = pending_proposal(D, i, ne, C.cur, C.rng_step0, C.first_bad);
            if (i < C.first_bad) (isE ? D.e_fsite : D.h_fsite)[li] = P.fs;
            D.tgt1[i] = P.p1;
            D.tgt2[i] = P.p2;
            D.isum[i] = 0;
            // decision draws: own-grid type Empty or Drain at the proposed site (chargeagent.cpp:95-157)
            const uint8_t tf = isE ? D.etype[P.fs] : D.htype[P.fs];
            draw = (tf == kEmpty || tf == kDrain);
            D.flag[i] = draw;
        }
        const int cnt = __syncthreads_count(draw);
        if (threadIdx.x == 0) D.blk_draw[wx - nbw] = cnt;
        return;
    }
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = wx * 4 + wid, nb = D.nbrick;
    int base, end, t0;
    if (nb <= kEmitScanBricks) {
        // this workgroup's list offsets: the carriers of the bricks before each
        // of its four in linear order (charge lists, after the defects of
        // dstart) and in bperm order (targets), from the counters that
        // tick_kernel and inject_kernel marked
        constexpr int kPer = kEmitScanBricks / kSimBlock;
        const __attribute__((address_space(1))) int* cg = gmem(D.scnt);
        const __attribute__((address_space(1))) int* rg = gmem(D.brank);
        int rk[4], c[kPer], r[kPer];
#pragma unroll
        for (int w = 0; w < 4; ++w) rk[w] = rg[min(wx * 4 + w, nb - 1)];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {   // clamped indices: every load unconditional
            const int bb = int(threadIdx.x) + kSimBlock * u;
            c[u] = cg[min(bb, nb - 1)];
            r[u] = rg[min(bb, nb - 1)];
            if (bb >= nb) c[u] = 0;
        }
        const int nmark = cg[min(b, nb - 1)];
        const int d0 = gmem(D.dstart)[min(b, nb - 1)], d1 = gmem(D.dstart)[min(b, nb - 1) + 1];
        int sum[8];
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            sum[w] = sum[4 + w] = 0;
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int bb = int(threadIdx.x) + kSimBlock * u;
                sum[w] += bb < wx * 4 + w ? c[u] : 0;
                sum[4 + w] += r[u] < rk[w] ? c[u] : 0;
            }
        }
        __shared__ int part[4][8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int t = wave_total(wave_incl_sum(sum[k]));
            if (lane == 0) part[wid][k] = t;
        }
        __syncthreads();
        if (b >= nb) return;
        const int lin = part[0][wid] + part[1][wid] + part[2][wid] + part[3][wid];
        t0 = part[0][4 + wid] + part[1][4 + wid] + part[2][4 + wid] + part[3][4 + wid];
        base = d0 + lin;
        end = d1 + lin + nmark;
        if (lane == 0) {
            gst(D.sstart)[b] = base;   // cull_de_kernel's list bounds
            if (b == nb - 1) gst(D.sstart)[nb] = end;
        }
    } else {
        if (b >= nb) return;
        base = sst[b];
        end = sst[b + 1];
        t0 = tst[b];
    }
    const uint32_t v = lane < 32 ? D.bmask[b * 32 + lane] : 0u;
    const int cnt = __popc(v);
    const int incl = wave_incl_sum(cnt);
    const int neh = __builtin_amdgcn_readlane(incl, 31);
    const int nmark = D.scnt[b];
    if (lane == 0 && neh != nmark) set_error(D, kErrSite);   // two of one species on a site
    const int ne = D.ctl->n_e;
    const int bx = b % D.nbx, byz = b / D.nbx, by = byz % D.nby, bz = byz / D.nby;
    const bool isE = lane < 16;
    const float q = isE ? -1.f : 1.f;
    const int* who = isE ? D.e_who : D.h_who;
    int k = base + incl - cnt;
    for (uint32_t m = v; m && k < end; m &= m - 1u, ++k) {
        const int j = 32 * (lane & 15) + __builtin_ctz(m);
        const int x = kBrick * bx + (j & 7), y = kBrick * by + ((j >> 3) & 7), z = kBrick * bz + (j >> 6);
        D.src[k] = make_float4(float(x), float(y), float(z), q);
        // the carrier on this site is the brick's target of the same rank
        if (k - base < nmark) {
            const int li = who[x + D.d.X * (y + D.d.Y * z)];
            D.torder[t0 + (k - base)] = isE ? li : ne + li;
        }
    }
    const int d0 = D.dstart[b], nd = D.dstart[b + 1] - d0;
    for (int u = lane; u < nd && base + neh + u < end; u += 64) D.src[base + neh + u] = D.dsrc[d0 + u];
    // slots left when two carriers of one species share a site (flagged above):
    // zero charges far away (their terms are exactly 0)
    for (int u = base + neh + nd + lane; u < end; u += 64) D.src[u] = make_float4(-1e4f, -1e4f, -1e4f, 0.f);
}

// t / d and t % d for 0 <= t < 2^24 with inv = 1 / d (float estimate, corrected)
__device__ __forceinline__ int fdiv(int t, int d, float inv, int& r) {
    int q = int(float(t) * inv);
    r = t - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
    return q;
}

// decideFuture's inputs but dE (chargeagent.cpp:90-161) for carrier block blk
// (256 carriers), in the same order and arithmetic as decide_kernel's own
// path: the block's draws start at rng_pos plus the draws of the blocks before
// it, and a drawing carrier takes the next engine output.  Reads only state
// that the step's dE phase leaves alone, so a repeated dE phase rewrites the
// same records.  wsum: 4 ints and a long long of LDS.
__device__ void decide_prep(const SimDev& D, int blk, int* wsum, long long* pre) {
    const SimCtl& C = *D.ctl;
    const int ne = C.n_e, N = ne + C.n_h, cur = C.cur;
    if (blk * kSimBlock >= N) return;   // whole workgroup
    const int i = blk * kSimBlock + int(threadIdx.x);
    const bool valid = i < N;
    const bool isE = i < ne;
    const int li = isE ? i : i - ne;
    const int q = isE ? -1 : 1;
    const long long draws_lane = wg_prefix_loads(D.blk_draw, blk, 1, 0);
    DecideRec r{};
    int draw = 0;
    uint8_t tf = kEmpty;
    double vb_fs = 0.0, vb_site = 0.0;
    if (valid) {
        r.site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
        r.fs = isE ? D.e_fsite[li] : D.h_fsite[li];
        r.life = isE ? D.e_life[cur][li] : D.h_life[cur][li];
        r.path = isE ? D.e_path[cur][li] : D.h_path[cur][li];
        draw = D.flag[i];
        tf = isE ? D.etype[r.fs] : D.htype[r.fs];
        vb_fs = D.vbg[r.fs];
        vb_site = D.vbg[r.site];
        // bindingPotential (chargeagent.cpp:463-481)
        if (isE) {
            r.bf = D.htype[r.fs] == kHole ? D.self_int + D.binding : 0.0;
            r.bs = D.htype[r.site] == kHole ? D.self_int + D.binding : 0.0;
        } else {
            r.bf = D.etype[r.fs] == kElectron ? D.self_int - D.binding : 0.0;
            r.bs = D.etype[r.site] == kElectron ? D.self_int - D.binding : 0.0;
        }
    }
    const long long before = wg_prefix_total(draws_lane, pre);
    int total;
    const int off = wg_scan<kSimBlock>(draw, wsum, total);
    if (!valid) return;
    if (draw) {
        r.u = u01(rng_word(D, C.rng_pos + (unsigned long long)(before + off)));
        if (tf == kEmpty) {
            double pd = vb_fs - vb_site;
            pd *= q;
            r.pd0 = pd;
            int x1, y1, z1, x2, y2, z2;
            decode(D.d, r.site, x1, y1, z1);
            decode(D.d, r.fs, x2, y2, z2);
            const int h1 = D.hop + 1;
            r.cpl = D.coupling[(abs(x1 - x2) * h1 + abs(y1 - y2)) * h1 + abs(z1 - z2)];
        }
    }
    r.flags = draw | (int(tf) << 8);
    D.drec[i] = r;
}

// one record as four 16-byte global loads
__device__ __forceinline__ DecideRec load_rec(const DecideRec* p) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    static_assert(sizeof(DecideRec) == 4 * sizeof(v4i), "record of four 16-byte words");
    const __attribute__((address_space(1))) v4i* w = gmem(reinterpret_cast<const v4i*>(p));
    const v4i a[4] = {w[0], w[1], w[2], w[3]};
    DecideRec r;
    __builtin_memcpy(&r, a, sizeof(r));
    return r;
}

constexpr int kList = 256;   // LDS charge slots per wave (24 KB of LDS per workgroup with a
                             // cutoff-30 table: six workgroups per CU)
constexpr int kGather = 8;   // units whose loads a wave keeps in flight together
constexpr int kMaxCullS = 32;   // chunks per target group
constexpr int kUnit = 64;    // summation unit: a 64-slot piece of a brick's list
// cull_de_kernel's dynamic LDS: a table of rtab_n entries (0: no table), then
// four wave lists of kList float4 slots and their kList-entry unit tables
inline size_t cull_lds_bytes(int rtab_n) {
    return 16 * size_t((rtab_n + 3) / 4) + 4 * size_t(kList) * (sizeof(float4) + sizeof(int));
}

// Per replica ceil(items_cap / 4) + 1 + nblk workgroups: workgroup x takes
// work item x (group g, quad q of Q) and its wave w chunk s = 4 q + w of
// S = 4 Q; the next one completes the random-stream refill; the last nblk
// prepare the decisions' inputs (decide_prep).
//
// Every wave of group g finds the same work: the bricks of the box widened by
// the cutoff whose distance to the box is below the cutoff, their lists cut
// into units (64-slot pieces), T slots in all.  For boxes of up to 512
// candidate bricks the four waves test them together (two per thread) into a
// table in LDS; larger boxes are tested by each wave, 64 at a time.  Chunk s
// takes the units whose first slot p (in candidate order) has
// floor(p S / T) = s, so every chunk gets T / S slots give or take a unit:
//   gather: each lane copies its bricks' units of the chunk into the wave's
//     LDS list (its own 64-byte rows, all lanes at once) at the exclusive scan
//     of the counts; the list is summed whenever the next bricks do not fit;
//   sum: every lane runs over the list (broadcast reads): an FP32 sum per unit,
//     taken to fixed point (x 2^40, truncated) and added in int64; the four
//     chunks' totals are added in LDS and go into the target's isum with one
//     64-bit integer atomic.
// The units and their FP32 sums are the same for a target in any group,
// and the units a group skips hold only terms that are exactly 0 for its
// targets; integer addition is exact, so each target's total is the same
// whatever its group, chunk count or order.
// (table form: registers for six waves per SIMD, 80 VGPRs, which the LDS of six
// workgroups allows; more waves hide more of the chain and the table reads)
template <bool kGauss, bool kTable>
__global__ __launch_bounds__(kSimBlock, kTable ? 6 : 4) void cull_de_kernel(const SimDev* __restrict__ Dp,
                                                            const int2* __restrict__ map) {
    const int2 wg = cull_workgroup(map);
    const SimDev& D = Dp[wg.x];
    const int wx = wg.y;
    // LDS (dynamic, cull_lds_bytes): [pair-function table | 4 wave lists | their
    // unit tables]; the table at address 0, so that a lookup's address is the
    // clamped index itself
    extern __shared__ float4 cull_lds[];
    const int tab16 = kTable ? (D.rtab_n + 3) >> 2 : 0;
    float4 (*lists)[kList] = reinterpret_cast<float4 (*)[kList]>(cull_lds + tab16);
    int (*uends)[kList] = reinterpret_cast<int (*)[kList]>(cull_lds + tab16 + 4 * kList);
    if (wx == (D.items_cap + 3) / 4) {
        // The random-stream refill.  Its CU is LDS bound (the pair sums), where
        // the 227-word rounds (three LDS reads a word) run faster than the wide
        // ones; its serial rounds issue first.
        static_assert(4 * kList * sizeof(float4) >= kTailWords * sizeof(uint32_t), "refill window in the lists");
        __builtin_amdgcn_s_setprio(3);
        mt_refill<kSimBlock, 0>(D, reinterpret_cast<uint32_t*>(&lists[0][0]), D.ctl->rng_pos + 624ull + D.lookahead);
        return;
    }
    if (wx > (D.items_cap + 3) / 4) {   // the last nblk workgroups: the decisions' inputs
        decide_prep(D, wx - (D.items_cap + 3) / 4 - 1, &uends[0][0], reinterpret_cast<long long*>(&uends[1][0]));
        return;
    }
    // Every wave starts with a chain of dependent loads (its box, candidate
    // bricks and list units) while other waves of its SIMD may already be
    // summing: the list building runs at a raised issue priority, the sums at
    // the normal one, so that a wave's latency-bound part is not held back by
    // the others' arithmetic.
    __builtin_amdgcn_s_setprio(1);
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    // the target count and the item are independent loads
    const int nt = *gmem(&D.ctl->counts[0]);
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i it = gmem(reinterpret_cast<const v4i*>(D.items))[wx];
    if (it.x < 0 || nt <= 0) return;   // whole workgroup: no item (testing both issues both loads)
    const int g = it.x, s = 4 * it.y + wv, S = 4 * it.z;
    const int r = g * kCullGroup + lane;
    const bool valid = r < nt;
    // a group holds at least one target: the lanes past the last load its sites
    // (no branch between the loads of the chain)
    const int i = gmem(D.torder)[min(r, nt - 1)];
    // the pair-function table: the loads of its first 4 kSimBlock entries go out
    // with the chain's, and the values are written to LDS once the target sites
    // are in (read after the first workgroup barrier)
    float* rtab_lds = reinterpret_cast<float*>(cull_lds);
    const __attribute__((address_space(1))) float* rt = gmem(D.rtab);
    float tv[4];
    if (kTable) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = int(threadIdx.x) + kSimBlock * j;
            tv[j] = k < D.rtab_n ? rt[k] : 0.f;
        }
    }
    const v4f ta = gmem(reinterpret_cast<const v4f*>(D.tgt1))[i];
    const v4f tb = gmem(reinterpret_cast<const v4f*>(D.tgt2))[i];
    const float4 p1 = make_float4(ta.x, ta.y, ta.z, ta.w), p2 = make_float4(tb.x, tb.y, tb.z, tb.w);
    // the box is wave-uniform (scalar registers); the coordinates are
    // non-negative, so their float bit patterns order as integers
    auto lo_of = [&](float a, float b) { return valid ? __float_as_int(fminf(a, b)) : 0x7fffffff; };
    auto hi_of = [&](float a, float b) { return valid ? __float_as_int(fmaxf(a, b)) : int(0x80000000); };
    const int x0 = int(__int_as_float(wave_imin(lo_of(p1.x, p2.x))));
    const int y0 = int(__int_as_float(wave_imin(lo_of(p1.y, p2.y))));
    const int z0 = int(__int_as_float(wave_imin(lo_of(p1.z, p2.z))));
    const int x1 = int(__int_as_float(wave_imax(hi_of(p1.x, p2.x))));
    const int y1 = int(__int_as_float(wave_imax(hi_of(p1.y, p2.y))));
    const int z1 = int(__int_as_float(wave_imax(hi_of(p1.z, p2.z))));
    const int c = D.cut, c2 = c * c;
    const int bx0 = max(0, x0 - c + 1) >> 3, bx1 = min(D.d.X - 1, x1 + c - 1) >> 3;
    const int by0 = max(0, y0 - c + 1) >> 3, by1 = min(D.d.Y - 1, y1 + c - 1) >> 3;
    const int bz0 = max(0, z0 - c + 1) >> 3, bz1 = min(D.d.Z - 1, z1 + c - 1) >> 3;
    const int nx = bx1 - bx0 + 1, ny = by1 - by0 + 1, ncand = nx * ny * (bz1 - bz0 + 1);
    if (kTable) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = int(threadIdx.x) + kSimBlock * j;
            if (k < D.rtab_n) rtab_lds[k] = tv[j];
        }
        for (int k = int(threadIdx.x) + 4 * kSimBlock; k < D.rtab_n; k += kSimBlock) rtab_lds[k] = rt[k];
    }
    // estimates for fdiv, which corrects a quotient that is off by one
    const float inx = __builtin_amdgcn_rcpf(float(nx)), iny = __builtin_amdgcn_rcpf(float(ny));
    const HopTerm ht = hop_term(p1, p2);
    const uint32_t c2m1 = D.c2m1;
    const float gk = D.gk;
    // table form: coordinates relative to the box's centre o; a kept charge lies
    // within c + |H| of o (H: the box's half extents), so the short form of
    // 4 r^2 holds when |H| <= 90 - c
    const int ox = (x0 + x1) >> 1, oy = (y0 + y1) >> 1, oz = (z0 + z1) >> 1;
    typedef __attribute__((address_space(3))) float* lds_fp;
    const uint32_t tbase = uint32_t(uintptr_t((lds_fp)rtab_lds));   // (address-space cast)
    const TabTarget ta1 = tab_target(p1, ox, oy, oz, tbase), ta2 = tab_target(p2, ox, oy, oz, tbase);
    const int hx = max(x1 - ox, ox - x0), hy = max(y1 - oy, oy - y0), hz = max(z1 - oz, oz - z0);
    const bool short_r2 = c <= 90 && hx * hx + hy * hy + hz * hz <= (90 - c) * (90 - c);
    const uint32_t tlim = tbase + D.rtab_lim;
    // a charge as the wave's list holds it (the record above in table form)
    auto stage = [&](const v4f& a) -> v4f {
        if (!kTable) return a;
        const int qx = int(a.x) - ox, qy = int(a.y) - oy, qz = int(a.z) - oz;
        const int n4 = 4 * (qx * qx + qy * qy + qz * qz);
        v4i w;
        w.x = int((uint32_t(-8 * qx) & 0xffffu) | (uint32_t(-8 * qy) << 16));
        w.y = int((uint32_t(-8 * qz) & 0xffffu) | (uint32_t(n4) << 16));
        w.z = n4;
        w.w = __float_as_int(a.w);
        return __builtin_bit_cast(v4f, w);
    };
    // the FP32 sum of list slots [t, e) in table form: q (f[4 r2^2] - f[4 r1^2]),
    // four slots' loads at a time, added in slot order
    auto tab_sum = [&](auto kshort, int t, int e) {
        constexpr bool kS = decltype(kshort)::value;
        const v4i* L = reinterpret_cast<const v4i*>(&lists[wv][0]);
        float acc = 0.f;
        for (; t + 4 <= e; t += 4) {
            v4i w[4];
            float d[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) w[j] = rec_load<kS>(L, t + j);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                d[j] = tab_f(tab_r2x4<kS>(ta2, w[j].x, w[j].y, w[j].z), tlim) -
                       tab_f(tab_r2x4<kS>(ta1, w[j].x, w[j].y, w[j].z), tlim);
#pragma unroll
            for (int j = 0; j < 4; ++j) acc = fmaf(__int_as_float(w[j].w), d[j], acc);
        }
        for (; t < e; ++t) {
            const v4i w = rec_load<kS>(L, t);
            const float d = tab_f(tab_r2x4<kS>(ta2, w.x, w.y, w.z), tlim) -
                            tab_f(tab_r2x4<kS>(ta1, w.x, w.y, w.z), tlim);
            acc = fmaf(__int_as_float(w.w), d, acc);
        }
        return acc;
    };
    const __attribute__((address_space(1))) int* sstart = gmem(D.sstart);
    // candidate ci -> its list [k0, k0 + n) when kept, else n = 0
    auto brick = [&](int ci, int& k0) -> int {
        k0 = 0;
        if (ci >= ncand) return 0;
        int rx, ry;
        const int ryz = fdiv(ci, nx, inx, rx);
        const int rz = fdiv(ryz, ny, iny, ry);
        const int bx = bx0 + rx, by = by0 + ry, bz = bz0 + rz;
        const int gx = max(0, max(kBrick * bx - x1, x0 - (kBrick * bx + kBrick - 1)));
        const int gy = max(0, max(kBrick * by - y1, y0 - (kBrick * by + kBrick - 1)));
        const int gz = max(0, max(kBrick * bz - z1, z0 - (kBrick * bz + kBrick - 1)));
        if (gx * gx + gy * gy + gz * gz >= c2) return 0;
        const int b = bx + D.nbx * (by + D.nby * bz);
        k0 = sstart[b];
        return sstart[b + 1] - k0;
    };
    // this chunk's units of a brick whose first slot has position pex: slots
    // [from, from + todo) of its list (unit j starts at pex + 64 j)
    int lo = 0, hi = 0;
    auto owned = [&](int n, int pex, int& from) -> int {
        const int nun = (n + kUnit - 1) / kUnit;
        const int u0 = min(nun, max(0, lo - pex + kUnit - 1) / kUnit);
        const int u1 = min(nun, max(0, hi - pex + kUnit - 1) / kUnit);
        from = kUnit * u0;
        return n > 0 ? max(0, min(n, kUnit * u1) - from) : 0;
    };

    long long acc64 = 0;
    int fill = 0, nu = 0;   // list slots and units in use
    auto flush = [&]() {
        __builtin_amdgcn_s_setprio(0);
        __builtin_amdgcn_wave_barrier();
        if (kTable) {
            for (int q = 0; q < nu; ++q) {
                const int se = __builtin_amdgcn_readfirstlane(uends[wv][q]);
                const float acc = short_r2 ? tab_sum(std::true_type{}, se >> 16, se & 0xffff)
                                           : tab_sum(std::false_type{}, se >> 16, se & 0xffff);
                acc64 += fix_round(acc);
            }
        }
        for (int q = 0; q < (kTable ? 0 : nu); ++q) {
            const int se = __builtin_amdgcn_readfirstlane(uends[wv][q]);
            int t = se >> 16;
            const int e = se & 0xffff;
            float acc = 0.f;
            for (; t + 2 <= e; t += 2) {
                const float4 q0 = lists[wv][t], q1 = lists[wv][t + 1];
                acc = fmaf(q0.w, de_term<kGauss>(p1, q0, ht, c2m1, gk), acc);
                acc = fmaf(q1.w, de_term<kGauss>(p1, q1, ht, c2m1, gk), acc);
            }
            if (t < e) {
                const float4 q0 = lists[wv][t++];
                acc = fmaf(q0.w, de_term<kGauss>(p1, q0, ht, c2m1, gk), acc);
            }
            acc64 += fix_round(acc);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_setprio(1);
        fill = 0;
        nu = 0;
    };
    // A charge at distance >= c from the group's box is out of cutoff for both
    // sites of every target of the group: its terms are exactly 0 and it is not
    // copied (leaving exact zeros out of an FP32 sum changes no bit of it, so
    // the unit sums stay those of the whole list).
    const float fx0 = float(x0), fx1 = float(x1), fy0 = float(y0), fy1 = float(y1);
    const float fz0 = float(z0), fz1 = float(z1), c2f = float(c2);
    auto near = [&](const v4f& a) {
        const float gx = fmaxf(fmaxf(a.x - fx1, fx0 - a.x), 0.f);
        const float gy = fmaxf(fmaxf(a.y - fy1, fy0 - a.y), 0.f);
        const float gz = fmaxf(fmaxf(a.z - fz1, fz0 - a.z), 0.f);
        return gx * gx + gy * gy + gz * gz < c2f;
    };
    // copy `take` slots (whole units, or up to the brick's end) from list slot
    // k of the charge lists to the wave's list from slot o, each unit's near
    // charges packed at the start of its 64 slots; unit uo + j's slots are
    // [start, end) with (start << 16) | end in the unit table
    auto copy = [&](int k, int take, int o, int uo) {
        const __attribute__((address_space(1))) v4f* src = gmem(reinterpret_cast<const v4f*>(D.src)) + k;
        v4f* dst = reinterpret_cast<v4f*>(&lists[wv][0]) + o;
        for (int j = 0; j * kUnit < take; ++j) {
            const int u0 = kUnit * j, u1 = min(take, u0 + kUnit);
            int w = u0, u = u0;
            for (; u + 8 <= u1; u += 8) {
                v4f a[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) a[i] = src[u + i];
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (near(a[i])) dst[w++] = stage(a[i]);
            }
            for (; u < u1; ++u) {
                const v4f a = src[u];
                if (near(a)) dst[w++] = stage(a);
            }
            uends[wv][uo + j] = ((o + u0) << 16) | (o + w);
        }
    };
    // one brick segment per lane, [k, k + todo), in rounds of whole units
    auto segment = [&](int k, int todo) {
        for (;;) {
            const int incl = wave_incl_sum(todo);
            const int total = wave_total(incl);
            if (total == 0) break;
            if (fill + total > kList && fill > 0) flush();
            const int excl = incl - todo;
            int take = min(todo, max(0, kList - fill - excl));   // a lane's first unit fits after a flush
            if (take < todo) take &= ~(kUnit - 1);
            const int nut = (take + kUnit - 1) / kUnit;
            const int uincl = wave_incl_sum(nut);
            copy(k, take, fill + excl, nu + uincl - nut);
            fill += wave_total(wave_incl_sum(take));
            nu += wave_total(uincl);
            k += take;
            todo -= take;
        }
    };

    // chunk s of S owns the slots [lo, hi) (T < 2^24 slots, so s T fits an int)
    auto bounds = [&](int T) {
        if (s == 0 && lane == 0) D.tgrp[g] = max(1, T);
        lo = int(unsigned(s * T + S - 1) / unsigned(S));
        hi = int(unsigned((s + 1) * T + S - 1) / unsigned(S));
    };
    // the chunk from the candidates themselves, 64 at a time, in list segments
    auto by_bricks = [&]() {
        int pos = 0;
        for (int cb = 0; cb < ncand && pos < hi; cb += 64) {
            int k0, from;
            const int n = brick(cb + lane, k0);
            const int incl = wave_incl_sum(n);
            const int todo = owned(n, pos + incl - n, from);
            pos += wave_total(incl);
            segment(k0 + from, todo);
        }
    };
    if (ncand <= 2 * kSimBlock) {   // the same for the four waves (one group)
        // The workgroup tests the candidates once, two per thread, into a table of
        // list lengths, list starts and positions (slots of the candidates before),
        // in the lists of waves 2 and 3, which the gathers only overwrite after a barrier.
        int* tab_n = reinterpret_cast<int*>(&lists[2][0]);
        int* tab_k = tab_n + 2 * kSimBlock;
        int* tab_p = tab_k + 2 * kSimBlock;
        int* wtot = tab_p + 2 * kSimBlock;   // the waves' totals
        static_assert(6 * kSimBlock + 4 <= 8 * kList, "candidate table in two waves' lists");
        int ka, kb;
        const int c0 = 2 * int(threadIdx.x);
        const int na = brick(c0, ka), nb = brick(c0 + 1, kb);
        const int incl = wave_incl_sum(na + nb);
        if (lane == 63) wtot[wv] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wv; ++w) before += wtot[w];
        const int T = wtot[0] + wtot[1] + wtot[2] + wtot[3];
        const int pa = before + incl - (na + nb);
        reinterpret_cast<int2*>(tab_n)[threadIdx.x] = make_int2(na, nb);
        reinterpret_cast<int2*>(tab_k)[threadIdx.x] = make_int2(ka, kb);
        reinterpret_cast<int2*>(tab_p)[threadIdx.x] = make_int2(pa, pa + na);
        __syncthreads();
        bounds(T);
        // the first candidate whose list ends after lo (the ends p + n never
        // decrease): first the block of 8, then the candidate within it
        auto end_at = [&](int ci) { return tab_p[ci] + tab_n[ci]; };
        const int blk = __popcll(__ballot(end_at(8 * lane + 7) <= lo));
        int cf = 8 * blk;
        if (blk < 64) cf += __popcll(__ballot(lane < 8 && end_at(8 * blk + (lane & 7)) <= lo));
        // the chunk's units, (first slot << 7) | length, in candidate order; a
        // candidate's units start at p + 64 m and the chunk owns those in [lo, hi)
        int total = 0;
        for (int cb = cf; cb < 2 * kSimBlock; cb += 64) {
            const int ci = cb + lane;
            int n = 0, p = T, k0 = 0;
            if (ci < 2 * kSimBlock) {
                n = tab_n[ci];
                p = tab_p[ci];
                k0 = tab_k[ci];
            }
            const int m0 = (max(0, lo - p) + kUnit - 1) / kUnit;
            const int m1 = (max(0, min(hi, p + n) - p) + kUnit - 1) / kUnit;
            const int cnt = max(0, m1 - m0);
            const int slots = cnt > 0 ? min(n, kUnit * m1) - kUnit * m0 : 0;
            const int ui = wave_incl_sum(cnt), si = wave_incl_sum(slots);
            int o = nu + ui - cnt;
            for (int m = m0; m < m1; ++m, ++o)
                if (o < kList) uends[wv][o] = ((k0 + kUnit * m) << 7) | min(kUnit, n - kUnit * m);
            nu += wave_total(ui);
            total += wave_total(si);
            if (__builtin_amdgcn_readlane(p + n, 63) >= hi) break;   // the next batch starts at or past hi
        }
        __syncthreads();   // every wave is done with the table
        if (total <= kList) {   // the whole chunk in one gather
            __builtin_amdgcn_wave_barrier();
            // every lane copies one slot of a unit, kGather units at a time (their
            // loads in flight together); the near charges of each unit are packed
            // in order (ballot prefix)
            const __attribute__((address_space(1))) v4f* src = gmem(reinterpret_cast<const v4f*>(D.src));
            v4f* dst = reinterpret_cast<v4f*>(&lists[wv][0]);
            int o = 0;
            for (int u0 = 0; u0 < nu; u0 += 64) {
                const int nb64 = min(64, nu - u0);
                const int dsc = lane < nb64 ? uends[wv][u0 + lane] : 0;   // unit u0 + lane
                for (int u = 0; u < nb64; u += kGather) {
                    v4f a[kGather];
                    bool in[kGather];
                    // every lane loads (lanes past a unit's length, or past the
                    // last unit with descriptor 0, re-read a slot of the unit or
                    // slot 0) and keeps only its own slot
#pragma unroll
                    for (int j = 0; j < kGather; ++j) {
                        const int d = __builtin_amdgcn_readlane(dsc, (u + j) & 63);
                        const int len = d & 127;
                        in[j] = u + j < nb64 && lane < len;
                        a[j] = src[unsigned(d >> 7) + unsigned(max(0, min(lane, len - 1)))];
                    }
#pragma unroll
                    for (int j = 0; j < kGather; ++j) {
                        if (u + j < nb64) {   // (no early exit: the unrolled loads stay in registers)
                            const bool k = in[j] && near(a[j]);
                            const uint64_t m = __ballot(k);
                            if (k) dst[o + __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0))] = stage(a[j]);
                            const int n = __popcll(m);
                            uends[wv][u0 + u + j] = (o << 16) | (o + n);
                            o += n;
                        }
                    }
                }
            }
            fill = o;
        } else {
            nu = 0;
            by_bricks();
        }
    } else {
        // large boxes: T first, then the candidates again 64 at a time (the four
        // waves share the box, so they all come here; the barrier publishes the table)
        if (kTable) __syncthreads();
        int tl = 0;
        for (int cb = 0; cb < ncand; cb += 64) {
            int k0;
            tl += brick(cb + lane, k0);
        }
        bounds(wave_total(wave_incl_sum(tl)));
        by_bricks();
    }
    flush();
    // the workgroup's four chunks belong to the same 64 targets: their totals are
    // added in LDS (exact int64) and each target takes one atomic
    __syncthreads();   // every wave is done with its list
    long long* part = reinterpret_cast<long long*>(&lists[0][0]);
    part[threadIdx.x] = acc64;
    __syncthreads();
    if (wv == 0) {
        const long long t = (part[lane] + part[64 + lane]) + (part[128 + lane] + part[192 + lane]);
        if (valid && t != 0) atomicAdd(reinterpret_cast<unsigned long long*>(&D.isum[i]), (unsigned long long)t);
    }
}

// ===========================================================================
// decide_kernel: decideFuture (chargeagent.cpp:90-161) for every carrier.
// The draw of carrier i is word rng_pos + (draws of carriers before i).
// only_de = 1 assembles dE (chargeagent.cpp:215-296) and stops.
// ===========================================================================
__global__ __launch_bounds__(kSimBlock) void decide_kernel(const SimDev* __restrict__ Dp, int only_de) {
    const SimDev& D = Dp[blockIdx.z];   // replica
    __shared__ int wsum[4];
    __shared__ long long pre;
    __shared__ unsigned int att[4], suc[4];
    __shared__ uint32_t win[kSimBlock];   // the workgroup's decision words
    const SimCtl& C = *D.ctl;
    const int ne = C.n_e, nh = C.n_h, N = ne + nh, cur = C.cur;
    const int nb = (N + kSimBlock - 1) / kSimBlock;
    if (blockIdx.x == 0 && threadIdx.x < 64 && !only_de) {
        int s = 0;
        for (int b = threadIdx.x; b < nb; b += 64) s += D.blk_draw[b];
        s = wave_total(wave_incl_sum(s));
        if (threadIdx.x == 0) {
            D.ctl->draws_dec = (unsigned long long)s;
            D.ctl->draws_rec = 0ull;
        }
    }
    if (int(blockIdx.x) * kSimBlock >= N) return;
    const int i = blockIdx.x * kSimBlock + threadIdx.x;
    const bool valid = i < N;
    const bool isE = i < ne;
    const int li = isE ? i : i - ne;
    const int q = isE ? -1 : 1;
    if (D.cull && !only_de) {
        // the culled dE phase made every input but dE (decide_prep): the record
        // and the pair sum are the only loads, and no barrier comes before the
        // decisions
        if (threadIdx.x < 4) { att[threadIdx.x] = 0; suc[threadIdx.x] = 0; }
        __syncthreads();
        int code = 0, keep = 1;
        if (valid) {
            const DecideRec r = load_rec(D.drec + i);
            const long long is = *gmem(&D.isum[i]);
            {   // the step is done with this carrier's bitmask word and brick counters
                int x, y, z;
                decode(D.d, r.site, x, y, z);
                uint32_t bit;
                D.bmask[cull_word(D, x, y, z, isE ? 0 : 1, bit)] = 0u;
                D.scnt[cull_brick(D, x, y, z)] = 0;
            }
            const double sde = double(is) * kFixInv;
            const double de = q * (D.pref * sde - D.self_int * q + r.bf - r.bs);
            D.de[i] = de;
            int fs = r.fs;
            if (r.flags & 1) {
                const double u = r.u;
                if ((r.flags >> 8) == kEmpty) {
                    const double pd = r.pd0 + de;
                    // metropolisWithCoupling (rand.cpp:91-106)
                    const bool acc = (pd > 0.0) ? (r.cpl * exp(-pd * D.beta) > u) : (r.cpl > u);
                    code = acc ? 1 : 0;
                } else {   // drain: DrainAgent::tryToAccept -> chooseYes(rate) (drainagent.cpp:83-92)
                    const int* ds = isE ? D.e_drain : D.h_drain;
                    const int k = (fs == ds[0]) ? 0 : 1;
                    const int flux = isE ? (LMG_FLUX_EDL + k) : (LMG_FLUX_HDL + k);
                    const int slot = (isE ? 0 : 2) + k;
                    atomicAdd(&att[slot], 1u);
                    if (D.rate[flux] > u) {
                        atomicAdd(&suc[slot], 1u);
                        code = 2;
                    }
                }
            }
            (isE ? D.e_life[cur] : D.h_life[cur])[li] = r.life + 1;
            if (code) (isE ? D.e_path[cur] : D.h_path[cur])[li] = r.path + 1;
            else fs = r.site;
            if (isE) D.e_fsite[li] = fs; else D.h_fsite[li] = fs;
            D.code[i] = code;
            D.removed[i] = 0;
            keep = (code != 2);
            if (code == 1) {
                const unsigned long long key = ((unsigned long long)(uint32_t)(C.step + 1) << 32) |
                                               (unsigned long long)(0xffffffffu - uint32_t(li));
                atomicMax(isE ? &D.claim_e[fs] : &D.claim_h[fs], key);
            }
        }
        const int ke = __syncthreads_count(valid && isE && keep);
        const int kh = __syncthreads_count(valid && !isE && keep);
        if (threadIdx.x == 0) {
            D.blk_keep[2 * blockIdx.x] = ke;
            D.blk_keep[2 * blockIdx.x + 1] = kh;
        }
        if (threadIdx.x < 4) {
            const int flux = threadIdx.x < 2 ? LMG_FLUX_EDL + threadIdx.x : LMG_FLUX_HDL + (threadIdx.x - 2);
            if (att[threadIdx.x]) atomicAdd(&D.ctl->att[flux], (unsigned long long)att[threadIdx.x]);
            if (suc[threadIdx.x]) atomicAdd(&D.ctl->suc[flux], (unsigned long long)suc[threadIdx.x]);
        }
        return;
    }
    // draws of the earlier workgroups: loads issued here, summed after the
    // carrier's own loads
    const long long draws_lane = only_de ? 0 : wg_prefix_loads(D.blk_draw, blockIdx.x, 1, 0);
    int site = 0, fs = 0, draw = 0, life = 0, path = 0;
    uint8_t tf = kEmpty;
    double vb_fs = 0.0, vb_site = 0.0;
    if (valid) {
        site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
        fs = isE ? D.e_fsite[li] : D.h_fsite[li];
        // the decision's inputs (fs may be a drain slot: V + k is inside the
        // type and background grids)
        draw = D.flag[i];
        tf = isE ? D.etype[fs] : D.htype[fs];
        vb_fs = D.vbg[fs];
        vb_site = D.vbg[site];
        life = isE ? D.e_life[cur][li] : D.h_life[cur][li];
        path = isE ? D.e_path[cur][li] : D.h_path[cur][li];
        if (D.cull && !only_de) {   // the step is done with this carrier's bitmask word and brick counters
            int x, y, z;
            decode(D.d, site, x, y, z);
            uint32_t bit;
            D.bmask[cull_word(D, x, y, z, isE ? 0 : 1, bit)] = 0u;
            D.scnt[cull_brick(D, x, y, z)] = 0;
        }
    }
    // the draws before this workgroup (a barrier: every load above is in), then
    // the engine outputs its draws can take, fetched while dE is assembled
    const long long before = only_de ? 0 : wg_prefix_total(draws_lane, &pre);
    const uint32_t word = only_de ? 0u : rng_word(D, C.rng_pos + (unsigned long long)(before + threadIdx.x));
    double de = 0.0;
    if (valid && D.coulomb) {
        double sde;
        if (D.cull) {
            sde = double(D.isum[i]) * kFixInv;   // cull_de_kernel's exact fixed-point total
        } else {   // de_kernel's S fixed-point chunk partials, 16 loads in flight
            const long long* ip = D.part;
            long long tot = 0;
            for (int q0 = 0; q0 < D.S; q0 += 16) {
                long long v[16];   // unconditional loads (clamped index)
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = ip[min(q0 + u, D.S - 1) * D.ld + i];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (q0 + u < D.S) tot += v[u];
            }
            sde = double(tot) * kFixInv;
        }
        double bf, bs;   // bindingPotential (chargeagent.cpp:463-481)
        if (isE) {
            bf = D.htype[fs] == kHole ? D.self_int + D.binding : 0.0;
            bs = D.htype[site] == kHole ? D.self_int + D.binding : 0.0;
        } else {
            bf = D.etype[fs] == kElectron ? D.self_int - D.binding : 0.0;
            bs = D.etype[site] == kElectron ? D.self_int - D.binding : 0.0;
        }
        de = q * (D.pref * sde - D.self_int * q + bf - bs);
    }
    if (valid) D.de[i] = de;
    if (only_de) return;

    if (threadIdx.x < 4) { att[threadIdx.x] = 0; suc[threadIdx.x] = 0; }
    win[threadIdx.x] = word;
    int total;
    const int off = wg_scan<kSimBlock>(draw, wsum, total);   // (its barriers publish win)
    int code = 0, keep = 1;
    if (valid) {
        if (draw) {
            const double u = u01(win[off]);   // engine output rng_pos + before + off
            if (tf == kEmpty) {
                double pd = vb_fs - vb_site;
                pd *= q;
                pd += de;
                int x1, y1, z1, x2, y2, z2;
                decode(D.d, site, x1, y1, z1);
                decode(D.d, fs, x2, y2, z2);
                const int h1 = D.hop + 1;
                const double cpl = D.coupling[(abs(x1 - x2) * h1 + abs(y1 - y2)) * h1 + abs(z1 - z2)];
                // metropolisWithCoupling (rand.cpp:91-106)
                const bool acc = (pd > 0.0) ? (cpl * exp(-pd * D.beta) > u) : (cpl > u);
                code = acc ? 1 : 0;
            } else {   // drain: DrainAgent::tryToAccept -> chooseYes(rate) (drainagent.cpp:83-92)
                const int* ds = isE ? D.e_drain : D.h_drain;
                const int k = (fs == ds[0]) ? 0 : 1;
                const int flux = isE ? (LMG_FLUX_EDL + k) : (LMG_FLUX_HDL + k);
                const int slot = (isE ? 0 : 2) + k;
                atomicAdd(&att[slot], 1u);
                if (D.rate[flux] > u) {
                    atomicAdd(&suc[slot], 1u);
                    code = 2;
                }
            }
        }
        (isE ? D.e_life[cur] : D.h_life[cur])[li] = life + 1;
        if (code) (isE ? D.e_path[cur] : D.h_path[cur])[li] = path + 1;
        else fs = site;
        if (isE) D.e_fsite[li] = fs; else D.h_fsite[li] = fs;
        D.code[i] = code;
        D.removed[i] = 0;
        keep = (code != 2);
        if (code == 1) {
            const unsigned long long key = ((unsigned long long)(uint32_t)(C.step + 1) << 32) |
                                           (unsigned long long)(0xffffffffu - uint32_t(li));
            atomicMax(isE ? &D.claim_e[fs] : &D.claim_h[fs], key);
        }
    }
    const int ke = __syncthreads_count(valid && isE && keep);
    const int kh = __syncthreads_count(valid && !isE && keep);
    if (threadIdx.x == 0) {
        D.blk_keep[2 * blockIdx.x] = ke;
        D.blk_keep[2 * blockIdx.x + 1] = kh;
    }
    if (threadIdx.x < 4) {
        const int flux = threadIdx.x < 2 ? LMG_FLUX_EDL + threadIdx.x : LMG_FLUX_HDL + (threadIdx.x - 2);
        if (att[threadIdx.x]) atomicAdd(&D.ctl->att[flux], (unsigned long long)att[threadIdx.x]);
        if (suc[threadIdx.x]) atomicAdd(&D.ctl->suc[flux], (unsigned long long)suc[threadIdx.x]);
    }
}

// ===========================================================================
// Serial helpers of the single-workgroup phases (thread 0 owns the stream)
// ===========================================================================
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

// carriers one injection pass can add: the two face or exciton passes, and
// balanceCharges' two loops of at most 1025 tries each
constexpr int kInjMax = 2 * 1025 + 8;

struct OneShared {
    int wsum[32];
    unsigned long long pos;
    int n_e, n_h, cur;
    int first_bad;
    int site, valid, ok, need_energy;
    int bal_go, bal_src;
    unsigned int att[10], suc[10];
    double red[16];
    uint32_t tail[kTailWords];
    // carriers injected this step on the culled path: their bricks
    int inj_n;
    int inj_b[kInjMax];
    int cnt[kStageBricks];   // inject_kernel's list offsets: per-brick carriers
};

__device__ double wg_sum_d(double v, double* red) {
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

// ===========================================================================
// recomb_kernel (one workgroup): RecombinationAgent::tryToAccept over the
// electrons in list order (drainagent.cpp:113-214); sequential through
// other->removed(), so thread 0 walks the list.  Re-counts the kept carriers.
// ===========================================================================
__global__ __launch_bounds__(kOneBlock) void recomb_kernel(const SimDev* __restrict__ Dp) {
    const SimDev& D = Dp[blockIdx.z];   // replica
    if (!(D.sim_type == 1 && D.rate[LMG_FLUX_XD] > 0)) return;
    __shared__ OneShared sh;
    SimCtl& C = *D.ctl;
    const int ne = C.n_e, nh = C.n_h, N = ne + nh, cur = C.cur;
    const unsigned long long start = C.rng_pos + C.draws_dec;
    for (int j = threadIdx.x; j < nh; j += blockDim.x) D.h_at[D.h_site[cur][j]] = j;
    for (int b = threadIdx.x; b < 2 * D.nblk; b += blockDim.x) D.blk_keep[b] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        sh.pos = start;
        Serial R{D, sh.pos};
        unsigned long long att = 0, suc = 0;
        for (int i = 0; i < ne; ++i) {
            const int site = D.e_site[cur][i];
            // candidates: the own site, then the neighbour list (range
            // recombination.range), keeping the sites that hold a hole
            const bool self = D.htype[site] == kHole;
            int n = self;
            if (D.rec_range != 0)
                nbr_enum(D.d, D.rec_range, site, D.e_drain[0], D.e_drain[1],
                         [&](int s) { n += (D.htype[s] == kHole); });
            if (n == 0) continue;
            int k = R.integer(0, n - 1);
            int rs = site;
            if (!(self && k == 0)) {
                k -= self;
                int m = 0;
                nbr_enum(D.d, D.rec_range, site, D.e_drain[0], D.e_drain[1], [&](int s) {
                    if (D.htype[s] == kHole) {
                        if (m == k) rs = s;
                        ++m;
                    }
                });
            }
            const int other = D.h_at[rs];
            if (D.removed[ne + other]) continue;
            att += 1;
            if (R.choose_yes(D.rate[LMG_FLUX_XD])) {
                suc += 1;
                D.removed[i] = 1;
                D.removed[ne + other] = 1;
            }
        }
        C.att[LMG_FLUX_XD] += att;
        C.suc[LMG_FLUX_XD] += suc;
        C.draws_rec = sh.pos - start;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        const int keep = !(D.removed[i] || D.code[i] == 2);
        if (keep) atomicAdd(&D.blk_keep[2 * (i / kSimBlock) + (i < ne ? 0 : 1)], 1);
    }
}

// ===========================================================================
// tick_kernel: nextTick / completeTick (simulation.cpp:355-414,
// chargeagent.cpp:163-199).  Among movers claiming one empty site the first in
// list order wins (the later ones find it occupied and abort); drained and
// recombined carriers leave; survivors are compacted in order.
// ===========================================================================
__global__ __launch_bounds__(kSimBlock) void tick_kernel(const SimDev* __restrict__ Dp) {
    const SimDev& D = Dp[blockIdx.z];   // replica
    __shared__ int wsum[8];
    __shared__ long long pre_e, pre_h;
    const __attribute__((address_space(1))) SimCtl& C = *gmem(D.ctl);
    const int ne = C.n_e, nh = C.n_h, N = ne + nh, cur = C.cur, nxt = cur ^ 1;
    const int nb = (N + kSimBlock - 1) / kSimBlock;
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        int se = 0, sh_ = 0;
        for (int b = threadIdx.x; b < nb; b += 64) {
            se += gmem(D.blk_keep)[2 * b];
            sh_ += gmem(D.blk_keep)[2 * b + 1];
        }
        se = wave_total(wave_incl_sum(se));
        sh_ = wave_total(wave_incl_sum(sh_));
        if (threadIdx.x == 0) {
            gst(D.ctl)->n_e_next = int(se);
            gst(D.ctl)->n_h_next = int(sh_);
        }
    }
    if (int(blockIdx.x) * kSimBlock >= N) return;
    const int i = blockIdx.x * kSimBlock + threadIdx.x;
    const bool valid = i < N;
    const bool isE = i < ne;
    const int li = isE ? i : i - ne;
    // kept carriers of the earlier workgroups: loads issued here, summed
    // after the carrier's own work
    const long long keep_e_lane = wg_prefix_loads(D.blk_keep, blockIdx.x, 2, 0);
    const long long keep_h_lane = wg_prefix_loads(D.blk_keep, blockIdx.x, 2, 1);
    int keep_e = 0, keep_h = 0, dest = 0, life = 0, path = 0;
    if (valid) {
        const int site = *gmem(isE ? &D.e_site[cur][li] : &D.h_site[cur][li]);
        const int code = gmem(D.code)[i];
        const bool rem = gmem(D.removed)[i] != 0;
        life = *gmem(isE ? &D.e_life[cur][li] : &D.h_life[cur][li]);
        path = *gmem(isE ? &D.e_path[cur][li] : &D.h_path[cur][li]);
        __attribute__((address_space(1))) uint8_t* ty = gst(isE ? D.etype : D.htype);
        dest = site;
        if (rem || code == 2) {
            ty[site] = kEmpty;
        } else if (code == 1) {
            const int fs = *gmem(isE ? &D.e_fsite[li] : &D.h_fsite[li]);
            const unsigned long long key = ((unsigned long long)(uint32_t)(C.step + 1) << 32) |
                                           (unsigned long long)(0xffffffffu - uint32_t(li));
            if (*gmem(isE ? &D.claim_e[fs] : &D.claim_h[fs]) == key) {
                ty[site] = kEmpty;
                ty[fs] = isE ? kElectron : kHole;
                dest = fs;
            }
        }
        const int keep = !(rem || code == 2);
        keep_e = isE && keep;
        keep_h = !isE && keep;
    }
    const long long base_e = wg_prefix_total(keep_e_lane, &pre_e);
    const long long base_h = wg_prefix_total(keep_h_lane, &pre_h);
    int oe, oh, te, th;
    wg_scan2<kSimBlock>(keep_e, keep_h, wsum, oe, oh, te, th);
    if (keep_e) {
        const int w = int(base_e) + oe;
        gst(D.e_site[nxt])[w] = dest;
        gst(D.e_life[nxt])[w] = life;
        gst(D.e_path[nxt])[w] = path;
        if (D.cull) cull_mark(D, true, w, dest);   // the next step's brick lists
    } else if (keep_h) {
        const int w = int(base_h) + oh;
        gst(D.h_site[nxt])[w] = dest;
        gst(D.h_life[nxt])[w] = life;
        gst(D.h_path[nxt])[w] = path;
        if (D.cull) cull_mark(D, false, w, dest);
    }
}

// ===========================================================================
// Injection (simulation.cpp:233-317, sourceagent.cpp:92-247) and proposals.
// ===========================================================================
// an injected carrier (list index li) on the culled path: its bitmask bit and
// the site's list index now; its brick goes into the list that inject_kernel's
// offsets add and then into the brick's counter (cull_mark's other half)
__device__ void inject_mark(const SimDev& D, OneShared& sh, bool isE, int li, int site) {
    int x, y, z;
    decode(D.d, site, x, y, z);
    uint32_t bit;
    atomicOr(&D.bmask[cull_word(D, x, y, z, isE ? 0 : 1, bit)], bit);
    (isE ? D.e_who : D.h_who)[site] = li;
    const int k = sh.inj_n;
    if (k >= kInjMax) {
        set_error(D, kErrCapacity);
        return;
    }
    sh.inj_b[k] = cull_brick(D, x, y, z);
    sh.inj_n = k + 1;
}

__device__ void append_carrier(const SimDev& D, OneShared& sh, bool isE, int site) {
    const int cur = sh.cur;
    if (isE) {
        const int n = sh.n_e;
        if (n >= D.cap_e) { set_error(D, kErrCapacity); return; }
        D.e_site[cur][n] = site; D.e_life[cur][n] = 0; D.e_path[cur][n] = 0;
        D.etype[site] = kElectron;
        if (D.cull) inject_mark(D, sh, true, n, site);
        sh.n_e = n + 1;
    } else {
        const int n = sh.n_h;
        if (n >= D.cap_h) { set_error(D, kErrCapacity); return; }
        D.h_site[cur][n] = site; D.h_life[cur][n] = 0; D.h_path[cur][n] = 0;
        D.htype[site] = kHole;
        if (D.cull) inject_mark(D, sh, false, n, site);
        sh.n_h = n + 1;
    }
}

// potential of all charges at one site, direct minus left image, times the
// prefactor: the sum SourceAgent::energyChange adds for source.coulomb
// (sourceagent.cpp:125-161; coulombE/H/D + coulombImageE/H/D)
__device__ double wg_site_potential(const SimDev& D, OneShared& sh, int site) {
    const int ne = sh.n_e, nh = sh.n_h, cur = sh.cur, nd = D.n_d_eff;
    int x, y, z;
    decode(D.d, site, x, y, z);
    const float fx = float(x), fy = float(y), fz = float(z);
    float s = 0.f, c = 0.f;
    for (int j = threadIdx.x; j < nd + ne + nh; j += blockDim.x) {
        float4 q;
        if (j < nd) q = D.chg[j];
        else if (j < nd + ne) q = site_f4(D.d, D.e_site[cur][j - nd], -1.f);
        else q = site_f4(D.d, D.h_site[cur][j - nd - ne], 1.f);
        const float dy = fy - q.y, dz = fz - q.z;
        const float dx = fx - q.x, dxi = fx + q.x + 1.f;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float ri = dxi * dxi + dy * dy + dz * dz;
        const float t = (cut_in(r2, D.c2m1) ? rsqf(r2) : 0.f) - (cut_in(ri, D.c2m1) ? rsqf(ri) : 0.f);
        kahan(s, c, q.w * t);
    }
    return D.pref * wg_sum_d(double(s) - double(c), sh.red);
}

__device__ bool inject_face(const SimDev& D, OneShared& sh, int flux) {
    const bool isE = (flux == LMG_FLUX_ESL || flux == LMG_FLUX_ESR);
    const bool left = (flux == LMG_FLUX_ESL || flux == LMG_FLUX_HSL);
    if (threadIdx.x == 0) {
        Serial R{D, sh.pos};
        sh.att[flux] += 1;
        // neighborsFace: the sorted x = 0 (or X-1) slice; entry m is X*m (+ X-1)
        const int m = R.integer(0, D.d.Y * D.d.Z - 1);
        const int site = D.d.X * m + (left ? 0 : D.d.X - 1);
        const int n = isE ? sh.n_e : sh.n_h;
        const int mx = isE ? D.max_e : D.max_h;
        const uint8_t t = isE ? D.etype[site] : D.htype[site];
        sh.site = site;
        sh.valid = !(n >= mx || D.rate[flux] <= 0 || t != kEmpty);   // validToInject
        sh.need_energy = sh.valid && D.src_metro && D.src_coul;
    }
    __syncthreads();
    double coul = 0.0;
    if (sh.need_energy) coul = wg_site_potential(D, sh, sh.site);
    if (threadIdx.x == 0) {
        sh.ok = 0;
        if (sh.valid) {
            Serial R{D, sh.pos};
            bool go;
            if (D.src_metro) {
                const double p2 = D.vbg[sh.site] + coul;   // m_potential of a source is never set: 0
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

__device__ void inject_all(const SimDev& D, OneShared& sh) {
    if (D.sim_type == 1) {
        if (threadIdx.x == 0) {   // ExcitonSourceAgent::tryToInject
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
        if (D.balance) {   // Simulation::balanceCharges (simulation.cpp:253-317)
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

// chooseFuture (chargeagent.cpp:73-78): neighbour k = integer(0, n-1)
__device__ int proposal_site(const SimDev& D, bool isE, int site, int k) {
    const int dl = isE ? D.e_drain[0] : D.h_drain[0];
    const int dr = isE ? D.e_drain[1] : D.h_drain[1];
    if (D.hop == 1) {
        int x, y, z;
        decode(D.d, site, x, y, z);
        return nbr1_at(D.d, x, y, z, dl, dr, k);
    }
    return nbr_at(D.d, D.hop, site, dl, dr, k);
}
__device__ int proposal_count(const SimDev& D, bool isE, int site) {
    const int dl = isE ? D.e_drain[0] : D.h_drain[0];
    const int dr = isE ? D.e_drain[1] : D.h_drain[1];
    if (D.hop == 1) {
        int x, y, z;
        decode(D.d, site, x, y, z);
        return nbr1_count(D.d, x, y, z, dl, dr);
    }
    return nbr_count(D.d, D.hop, site, dl, dr);
}

// Serial chooseFuture from carrier i0 on, starting at engine output pos
// (thread 0 only): the rare path after a Boost rejection, and lattices where
// some site has fewer than two neighbours (a one-entry list draws nothing).
__device__ unsigned long long propose_serial(const SimDev& D, int i0, int ne, int N, int cur,
                                             unsigned long long pos) {
    Serial R{D, pos};
    for (int i = i0; i < N; ++i) {
        const bool isE = i < ne;
        const int li = isE ? i : i - ne;
        const int site = isE ? D.e_site[cur][li] : D.h_site[cur][li];
        if (site < 0 || site >= D.d.V) { set_error(D, kErrSite); continue; }
        const int k = R.integer(0, proposal_count(D, isE, site) - 1);
        const int fs = proposal_site(D, isE, site, k);
        if (isE) D.e_fsite[li] = fs; else D.h_fsite[li] = fs;
    }
    return pos;
}

// mode 0: injection for the pending step (after tick_kernel)
// mode 1: refill the random stream after a state load
// Then the next step's set-up: on the culled path the brick counters (mode 1
// marks every carrier; in mode 0 tick_kernel and the injections marked them)
// become list offsets here for lattices of more than kEmitScanBricks bricks
// (smaller ones in cull_emit_kernel); the proposal draws are checked for a Boost rejection
// (first_bad), from which the proposals are made serially here -- all of them
// when some site has fewer than two neighbours (one-entry lists draw nothing).
// Leaves the new lists' sizes, the proposal start rng_step0, first_bad, the
// decisions' start rng_pos and the counts in ctl.
__global__ __launch_bounds__(kOneBlock) void inject_kernel(const SimDev* __restrict__ Dp, int mode) {
    const SimDev& D = Dp[blockIdx.z];   // replica
    __shared__ OneShared sh;
    SimCtl& C = *D.ctl;
    if (threadIdx.x == 0) {
        if (mode == 0) {
            sh.n_e = C.n_e_next;
            sh.n_h = C.n_h_next;
            sh.cur = C.cur ^ 1;
            sh.pos = C.rng_pos + C.draws_dec + C.draws_rec;
        } else {
            sh.n_e = C.n_e;
            sh.n_h = C.n_h;
            sh.cur = C.cur;
            sh.pos = C.rng_pos;
        }
    }
    if (threadIdx.x < 10) {
        sh.att[threadIdx.x] = 0;
        sh.suc[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) {
        sh.first_bad = 0x7fffffff;
        sh.inj_n = 0;
    }
    // after a step: the offsets' loads and the listed engine outputs before the
    // injections (neither changes there)
    // (lattices of up to kEmitScanBricks bricks: cull_emit_kernel's workgroups take their own offsets)
    const bool offs = D.cull && D.nbrick > kEmitScanBricks;
    int* cnt = offs && D.nbrick <= kStageBricks ? sh.cnt : nullptr;
    OffsetLoads ol;
    int nd = 0;
    unsigned long long dl[kDangerSlots];
    if (mode == 0) {
        if (offs) cull_offsets_load<kOneBlock, false>(D, cnt, ol);
        if (threadIdx.x == 0) {   // (global loads: the barrier below does not wait for them)
            nd = gmem(D.ctl)->danger_n;
#pragma unroll
            for (int k = 0; k < kDangerSlots; ++k) dl[k] = gmem(D.ctl)->danger[k];
        }
    }
    __syncthreads();
    if (mode == 1) mt_refill<kOneBlock, 2>(D, sh.tail, sh.pos + 624ull + D.lookahead);
    else inject_all(D, sh);
    __syncthreads();
    const int ne = sh.n_e, N = ne + sh.n_h, cur = sh.cur;
    const unsigned long long step0 = sh.pos;
    // Carrier i's proposal draw is output step0 + i; the first one Boost rejects.  Only
    // outputs that mt_refill listed can be rejected: thread 0 checks those in range.
    // Should the list have wrapped past one of them, every carrier is checked.
    if (D.uniform && N > 0 && threadIdx.x == 0) {
        if (mode == 1) {
            nd = C.danger_n;
#pragma unroll
            for (int k = 0; k < kDangerSlots; ++k) dl[k] = C.danger[k];
        }
        unsigned long long oldest = ~0ull;
        for (int k = 0; k < min(nd, kDangerSlots); ++k) {
            const unsigned long long p = dl[k];
            oldest = p < oldest ? p : oldest;
            if (p < step0 || p >= step0 + (unsigned long long)N) continue;
            const int i = int(p - step0);
            const bool isE = i < ne;
            const int site = isE ? D.e_site[cur][i] : D.h_site[cur][i - ne];
            if (site < 0 || site >= D.d.V) continue;   // reported by the proposal below
            const uint32_t range = uint32_t(proposal_count(D, isE, site) - 1);
            if (rng_word(D, p) / uint_bucket(range) > range) sh.first_bad = min(sh.first_bad, i);
        }
        sh.valid = nd > kDangerSlots && oldest > step0;   // the full check is needed
    }
    __syncthreads();
    if (D.uniform && N > 0 && sh.valid) {
        constexpr int kB = 8;      // carriers per thread per round, their loads in flight together
        for (int i0 = threadIdx.x; i0 < N; i0 += kB * kOneBlock) {
            int site[kB];
            uint32_t w[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int i = min(i0 + u * kOneBlock, N - 1);
                site[u] = i < ne ? D.e_site[cur][i] : D.h_site[cur][i - ne];
                w[u] = rng_word(D, step0 + (unsigned long long)i);
            }
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int i = i0 + u * kOneBlock;
                if (i >= N) continue;
                if (site[u] < 0 || site[u] >= D.d.V) {
                    set_error(D, kErrSite);
                    continue;
                }
                const uint32_t range = uint32_t(proposal_count(D, i < ne, site[u]) - 1);
                if (w[u] / uint_bucket(range) > range) atomicMin(&sh.first_bad, i);
            }
        }
    }
    if (D.cull) {
        if (mode == 1) {   // after a load: every carrier, then the offsets over counters marked here
            for (int i = threadIdx.x; i < N; i += blockDim.x) {
                const bool isE = i < ne;
                const int li = isE ? i : i - ne;
                cull_mark(D, isE, li, isE ? D.e_site[cur][li] : D.h_site[cur][li]);
            }
            if (offs) {
                __syncthreads();
                cull_offsets_load<kOneBlock, true>(D, cnt, ol);
                cull_offsets<kOneBlock, true>(D, D.sstart, D.tstart, sh.wsum, sh.inj_b, 0, cnt, ol);
            }
        } else {   // tick_kernel's marks plus this step's injected carriers
            if (offs) cull_offsets<kOneBlock, false>(D, D.sstart, D.tstart, sh.wsum, sh.inj_b, sh.inj_n, cnt, ol);
            for (int k = threadIdx.x; k < sh.inj_n; k += blockDim.x) atomicAdd(&D.scnt[sh.inj_b[k]], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int fb = D.uniform ? min(sh.first_bad, N) : 0;
        if (step0 + (unsigned long long)N + 624ull > C.gen_pos) set_error(D, kErrRing);
        C.n_e = ne;
        C.n_h = sh.n_h;
        C.cur = cur;
        C.rng_step0 = step0;
        C.first_bad = fb;
        C.counts[0] = N;
        C.counts[1] = D.n_d_eff + N;
        if (mode == 0) C.step += 1;
        C.rng_pos = fb < N ? propose_serial(D, fb, ne, N, cur, step0 + (unsigned long long)fb)
                           : step0 + (unsigned long long)N;
    }
    if (threadIdx.x < 10 && mode == 0) {
        C.att[threadIdx.x] += sh.att[threadIdx.x];
        C.suc[threadIdx.x] += sh.suc[threadIdx.x];
    }
}

}  // namespace lmg
