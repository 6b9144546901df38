// h4hip.hip -- libh4hip.so: HIP kernels and C ABI of the MI355X Hermite-4
// block-timestep backend (declarations and contracts in include/h4hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "h4hip.h"
#include "h4_kernels.hpp"

using namespace h4;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define H4H_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess)                                                \
            return fail(H4H_EHIP, std::string(#call) + ": " +                \
                                      hipGetErrorString(e_));                \
    } while (0)

#define H4H_LAUNCHED()                                                       \
    do {                                                                     \
        hipError_t e_ = hipGetLastError();                                   \
        if (e_ != hipSuccess)                                                \
            return fail(H4H_EHIP, std::string("kernel launch: ") +           \
                                      hipGetErrorString(e_));                \
    } while (0)

// FP32 partial sums of the mixed kernel are folded into FP64 every
// H4_MIXED_FOLD pairs per lane.
#ifndef H4_MIXED_FOLD
#define H4_MIXED_FOLD 32
#endif
// Unroll of the mixed inner loop: 2 keeps the kernel at 108 VGPRs (4 waves
// per SIMD); 4 needs 132 (3 waves) and measured 1-8 % slower.
#ifndef H4_MIXED_UNROLL
#define H4_MIXED_UNROLL 2
#endif
// Waves of the one-workgroup step's pair loop for active sets under 64.
#ifndef H4_SMALL_NARROW_WAVES
#define H4_SMALL_NARROW_WAVES 8
#endif

constexpr uint32_t DEFAULT_GRID = 2048;       // persistent workgroups
constexpr uint32_t DEFAULT_MAX_SPLIT = 1024;  // largest j split
constexpr uint32_t ENERGY_GRID = 1024;

// ---------------------------------------------------------------------------
// Cluster geometry helpers
// ---------------------------------------------------------------------------
//
// A context holds K independent clusters (K = 1 for a single system).
// Cluster c occupies device indices off[c] .. off[c] + cnt[c] - 1, padded
// to a multiple of JT with zero-mass particles, so every JT block belongs
// to exactly one cluster (bcl[block]) and j tiles never straddle clusters.
// Host arrays hold the clusters back to back (hoff[c]).

__device__ __forceinline__ uint32_t cluster_of_block(const Ens &e, uint32_t b) { return e.bcl[b]; }

__device__ __forceinline__ uint32_t ntiles(const Ens &e, uint32_t c) {
    return (e.cnt[c] + JT - 1) / JT;
}

// Index of the cluster whose [pre[c], pre[c+1]) range holds x (pre has
// K+1 entries, non-decreasing, pre[0] = 0, x < pre[K]).
__device__ __forceinline__ uint32_t find_cluster(const unsigned long long *pre, uint32_t K,
                                                 unsigned long long x) {
    uint32_t lo = 0, hi = K;  // answer in [lo, hi)
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= x)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Layout kernels (one thread per padded device index)
// ---------------------------------------------------------------------------

struct Cols {
    double *c[8];
    uint32_t stride[8];  // doubles between consecutive particles (0 means 1)
};

__device__ __forceinline__ size_t col_at(const Cols &c, uint32_t k, uint32_t i) {
    return (size_t)i * (c.stride[k] ? c.stride[k] : 1u);
}

// dst.c[k][i] = src[h*ncols + k] for the host index h of device index i.
__global__ __launch_bounds__(BS) void k_aos_to_soa(const double *src, uint32_t ncols, Cols dst,
                                                  Ens e) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= e.nblk * JT) return;
    uint32_t c = cluster_of_block(e, i / JT);
    uint32_t local = i - e.off[c];
    if (local >= e.cnt[c]) return;
    size_t h = (size_t)e.hoff[c] + local;
    for (uint32_t k = 0; k < ncols; ++k)
        if (dst.c[k]) dst.c[k][col_at(dst, k, i)] = src[h * ncols + k];
}

// dst[h*ncols + k] = src.c[k][i] (NULL columns write 0).
__global__ __launch_bounds__(BS) void k_soa_to_aos(Cols src, uint32_t ncols, double *dst, Ens e) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= e.nblk * JT) return;
    uint32_t c = cluster_of_block(e, i / JT);
    uint32_t local = i - e.off[c];
    if (local >= e.cnt[c]) return;
    size_t h = (size_t)e.hoff[c] + local;
    for (uint32_t k = 0; k < ncols; ++k)
        dst[h * ncols + k] = src.c[k] ? src.c[k][col_at(src, k, i)] : 0.0;
}

// Same for one array of scalars (t, dt) and for the u32 active list of a
// single cluster context.
__global__ __launch_bounds__(BS) void k_scalar_to_soa(const double *src, double *dst, Ens e) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= e.nblk * JT) return;
    uint32_t c = cluster_of_block(e, i / JT);
    uint32_t local = i - e.off[c];
    if (local < e.cnt[c]) dst[i] = src[e.hoff[c] + local];
}

__global__ __launch_bounds__(BS) void k_soa_to_scalar(const double *src, double *dst, Ens e) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= e.nblk * JT) return;
    uint32_t c = cluster_of_block(e, i / JT);
    uint32_t local = i - e.off[c];
    if (local < e.cnt[c]) dst[e.hoff[c] + local] = src[i];
}

// Predictor at t_i (Hermite4::init_data, Hermite4.cpp:194-240): position
// and velocity as corrected, mass rounded to float.  Padding slots get zero
// mass far away and never become active (t = 0, dt = +inf).
__global__ __launch_bounds__(BS) void k_init_predictor(Soa s, Ens e) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= e.nblk * JT) return;
    uint32_t c = cluster_of_block(e, i / JT);
    if (i - e.off[c] < e.cnt[c]) {
        const float mf = (float)s.m[i];
        s.pp[i] = make_double4(s.r[0][i], s.r[1][i], s.r[2][i], (double)mf);
        s.pv[i] = make_double4(s.v[0][i], s.v[1][i], s.v[2][i], mass_bits(mf));
    } else {
        s.pp[i] = make_double4(PAD_POS, PAD_POS, PAD_POS, 0.0);
        s.pv[i] = make_double4(0.0, 0.0, 0.0, mass_bits(0.0f));
        s.dt[i] = INFINITY;
    }
}

// Active list of every cluster = all its particles (init_acc_jrk).
__global__ __launch_bounds__(BS) void k_identity_active(uint32_t *move, Ens e) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= e.nblk * JT) return;
    uint32_t c = cluster_of_block(e, i / JT);
    uint32_t local = i - e.off[c];
    if (local < e.cnt[c]) move[i] = i;
    if (local == 0) e.cs[c].nact = e.cnt[c];
}

// ---------------------------------------------------------------------------
// Work plan of a step (one workgroup)
// ---------------------------------------------------------------------------

// Per cluster: lane-group width, j split and work items from its active
// count; the update workgroups are shared in proportion to nact * n.
// Exclusive prefix sums over clusters go to items_pre / fin_pre /
// part_pre, totals to ctl.  `exact`: one lane per i, no j split.
// Work plan of cluster c for nact active particles; W = sum over the
// clusters of nact * n (the update workgroups are shared in proportion).
__device__ __forceinline__ ClPlan plan_cluster(const Ens &e, uint32_t c, uint32_t nact, double W,
                                               uint32_t exact) {
    ClPlan pl;
    pl.cmaj = 0;
    if (exact) {
        pl.iw = BS;
        pl.S = 1;
        pl.tpc = ntiles(e, c);
        pl.n_it = (nact + BS - 1) / BS;
        pl.rows = pl.n_it * BS;
        pl.lanes = 1;
        pl.grid = e.grid;
    } else {
        double share = W > 0.0 ? (double)nact * (double)e.cnt[c] / W : 0.0;
        uint32_t g = (uint32_t)(share * e.grid);
        pl.grid = g < 1 ? 1 : g;
        pl.iw = pick_iw(nact ? nact : 1, e.iw_fill);
        if (e.tuned && e.K == 1 && nact) {
            uint32_t bucket = 31u - (uint32_t)__clz(nact);
            if (e.tune[bucket]) pl.grid = min(e.tune[bucket], e.grid);
            if (e.tune[32 + bucket]) pl.iw = e.tune[32 + bucket];
        }
        Decomp d = decompose(nact, pl.iw, ntiles(e, c), pl.grid, e.max_split);
        pl.S = d.S;
        pl.tpc = d.tiles_per_chunk;
        pl.n_it = d.n_it;
        pl.rows = d.rows;
        pl.lanes = finish_lanes(d.S, nact);
        pl.cmaj = (d.S > 1 && pl.lanes >= 16 && d.rows >= 64) ? 1u : 0u;
    }
    return pl;
}

// Plan of a single-cluster context, made by every workgroup of the update
// kernel from the active count (so the device step needs no k_plan
// launch); workgroup 0 publishes it for the finish pass.
__device__ __forceinline__ ClPlan plan_single(const Ens &e, uint32_t exact,
                                              unsigned long long &items) {
    const uint32_t nact = e.cs[0].nact;
    const ClPlan pl = plan_cluster(e, 0, nact, (double)nact * (double)e.cnt[0], exact);
    items = nact ? (unsigned long long)pl.n_it * pl.S : 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t per = BS / pl.lanes;
        const unsigned long long units = (nact + per - 1) / per;
        e.plan[0] = pl;
        e.items_pre[0] = e.fin_pre[0] = e.part_pre[0] = 0ull;
        e.items_pre[1] = items;
        e.fin_pre[1] = units;
        e.part_pre[1] = (nact && pl.S > 1) ? 6ull * pl.S * pl.rows : 0ull;
        e.ctl->items = items;
        e.ctl->fin_units = units;
    }
    return pl;
}

__device__ void plan_all(const Ens &e, uint32_t exact) {
    __shared__ double wsum[BS / 64];
    double w = 0.0;
    for (uint32_t c = threadIdx.x; c < e.K; c += BS)
        w += (double)e.cs[c].nact * (double)e.cnt[c];
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
    if (__lane_id() == 0) wsum[threadIdx.x >> 6] = w;
    __syncthreads();
    double W = 0.0;
    for (uint32_t k = 0; k < BS / 64; ++k) W += wsum[k];
    for (uint32_t c = threadIdx.x; c < e.K; c += BS) e.plan[c] = plan_cluster(e, c, e.cs[c].nact, W, exact);
    __syncthreads();
    // Exclusive prefix sums of work items, finish units and partial doubles
    // over the clusters, BS clusters at a time.
    __shared__ unsigned long long wtot[3][BS / 64];
    unsigned long long carry[3] = {0ull, 0ull, 0ull};
    for (uint32_t base = 0; base < e.K; base += BS) {
        const uint32_t c = base + threadIdx.x;
        unsigned long long v[3] = {0ull, 0ull, 0ull};
        if (c < e.K && e.cs[c].nact) {
            const ClPlan pl = e.plan[c];
            const uint32_t nact = e.cs[c].nact;
            const uint32_t per = BS / pl.lanes;
            v[0] = (unsigned long long)pl.n_it * pl.S;
            v[1] = (nact + per - 1) / per;
            v[2] = pl.S > 1 ? 6ull * pl.S * pl.rows : 0ull;
        }
        unsigned long long inc[3] = {v[0], v[1], v[2]};
        for (uint32_t o = 1; o < 64; o <<= 1)
            for (int q = 0; q < 3; ++q) {
                unsigned long long y = __shfl_up(inc[q], o);
                if (__lane_id() >= o) inc[q] += y;
            }
        const uint32_t w = threadIdx.x >> 6;
        if (__lane_id() == 63)
            for (int q = 0; q < 3; ++q) wtot[q][w] = inc[q];
        __syncthreads();
        unsigned long long tot[3];
        for (int q = 0; q < 3; ++q) {
            unsigned long long before = 0ull, all = 0ull;
            for (uint32_t k = 0; k < BS / 64; ++k) {
                if (k < w) before += wtot[q][k];
                all += wtot[q][k];
            }
            inc[q] += before + carry[q];  // inclusive, with earlier chunks
            tot[q] = all;
        }
        if (c < e.K) {
            e.items_pre[c] = inc[0] - v[0];
            e.fin_pre[c] = inc[1] - v[1];
            e.part_pre[c] = inc[2] - v[2];
        }
        for (int q = 0; q < 3; ++q) carry[q] += tot[q];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        e.items_pre[e.K] = carry[0];
        e.fin_pre[e.K] = carry[1];
        e.part_pre[e.K] = carry[2];
        e.ctl->items = carry[0];
        e.ctl->fin_units = carry[1];
    }
}

__global__ __launch_bounds__(BS) void k_plan(Ens e, uint32_t exact) { plan_all(e, exact); }

// ---------------------------------------------------------------------------
// Predictor / active selection / save old (one workgroup per JT block)
// ---------------------------------------------------------------------------

enum : uint32_t {
    PS_PREDICT = 1,   // Hermite4CPU::predicted_pos_vel (Hermite4CPU.cpp:340-365)
    PS_SELECT = 2,    // Hermite4::find_particles_to_move (Hermite4.cpp:63-84)
    PS_SAVE_OLD = 4,  // Hermite4::save_old_acc_jrk (Hermite4.cpp:145-153)
    PS_DEVTIME = 8,   // block time = the cluster's ATIME (device step loop)
    PS_DIRTY = 16,    // record blocks holding active particles
    PS_TAIL = 32,     // make the loop tail the previous step left (device loop)
    PS_COMPACT = 64,  // the active particles' records in Ens::arec (with
                      // PS_PREDICT | PS_SELECT | PS_SAVE_OLD)
};

// The loop tail of cluster c's block step number `steps` (1-based in this
// loop) made at block time T with nact active particles, the next block
// time being mn (Hermite4CPU_integration.cpp:81-124): ITIME and ATIME, the
// step and interaction counters, the step history of cluster 0, and the
// stop.  One thread.
__device__ __forceinline__ void record_tail(const Ens &e, uint32_t c, double T, uint32_t nact,
                                            unsigned long long steps, double mn, bool stop) {
    ClState *cs = &e.cs[c];
    if (c == 0) {
        e.hist_t[(steps - 1) % e.hist_cap] = T;
        e.hist_n[(steps - 1) % e.hist_cap] = nact;
    }
    cs->itime = T;
    cs->atime = mn;
    cs->steps = steps;
    cs->interactions += (unsigned long long)nact * e.cnt[c];
    cs->nsteps += nact;
    if (stop) {
        cs->done = 1;
        atomicAdd(&e.ctl->ndone, 1u);
    }
}

// The active set of a cluster is appended to its segment of move[] with one
// atomic per workgroup that has active particles (ballot per wave, prefix
// over the waves in LDS).  Clusters whose loop has stopped are skipped.
// The work plan is made by k_plan afterwards: a last-workgroup pattern
// here would need a device-scope fence per workgroup, i.e. an L2
// write-back on every workgroup of this multi-XCD part.
// Hermite4CPU::predicted_pos_vel (Hermite4CPU.cpp:340-365) of one particle
// (position r, velocity v, forces f, float mass mf) over dt = itime - t,
// as the two predictor records.
__device__ __forceinline__ void predict_vals(const double r[3], const double v[3],
                                             const double f[6], double dt, float mf,
                                             double4 &a, double4 &b) {
#pragma clang fp contract(off)
    double dt2 = (dt * dt);
    double dt3 = (dt2 * dt);
    double dt2c = dt2 / 2.0;
    double dt3c = dt3 / 6.0;
    double pr[3], pvel[3];
    for (int k = 0; k < 3; ++k) {
        pr[k] = (dt3c * f[3 + k]) + (dt2c * f[k]) + (dt * v[k]) + r[k];
        pvel[k] = (dt2c * f[3 + k]) + (dt * f[k]) + v[k];
    }
    a = make_double4(pr[0], pr[1], pr[2], (double)mf);
    b = make_double4(pvel[0], pvel[1], pvel[2], mass_bits(mf));
}

// PS_TAIL (device loop): each cluster's last block step left its loop tail
// in its ClState (tail_*); every workgroup of the cluster takes the next
// block time as the minimum of the cluster's block minima and makes the
// stop decision itself, the cluster's first workgroup records the tail, and
// the step proceeds at that time unless the cluster's loop stops.  This
// launch reads the tail records and the block minima and writes only fields
// no workgroup of it reads (the stop flag aside: a workgroup that sees it
// set would have stopped anyway).  Blocks with active particles go to the
// dirty list `par` (the step's slot in its captured chunk, mod 2), and
// workgroup 0 clears the other list's count for the next step.
__global__ __launch_bounds__(BS) void k_select(Soa s, Ens e, uint32_t flags, double itime_arg,
                                              uint32_t par, uint32_t *move) {
#pragma clang fp contract(off)
    __shared__ uint32_t wave_off[BS / 64 + 1];
    __shared__ double wmin[BS / 64];
    __shared__ double s_time;
    __shared__ uint32_t s_go;
    const uint32_t b = blockIdx.x;
    const uint32_t i = b * JT + threadIdx.x;  // inside the padded arrays
    const bool single = e.K == 1;
    const uint32_t c = single ? 0u : cluster_of_block(e, b);
    ClState *cs = &e.cs[c];
    const bool devloop = (flags & (PS_TAIL | PS_DEVTIME)) != 0;
    // The particle's state does not depend on the block time, so its loads
    // go out before the loop tail's minimum is known and their latency
    // overlaps that of the cluster record and the block minima.  A single
    // system loads it at once (the step slots of a captured chunk that find
    // the loop stopped load it for nothing); an ensemble, whose stopped
    // clusters can be most of it, first reads whether the cluster stopped.
    const bool load = single || !devloop || cs->done == 0;
    const bool need_f = (flags & (PS_PREDICT | PS_SAVE_OLD)) != 0;
    double ti = 0.0, dti = 0.0, f[6], r[3], v[3];
    float mf = 0.f;
    for (int k = 0; k < 6; ++k) f[k] = 0.0;
    for (int k = 0; k < 3; ++k) r[k] = v[k] = 0.0;
    if (load) {
        if (flags & (PS_PREDICT | PS_SELECT)) ti = s.t[i];
        if (flags & PS_SELECT) dti = s.dt[i];
        if (need_f)
            for (int k = 0; k < 6; ++k) f[k] = s.f[k][i];
        if (flags & PS_PREDICT) {
            for (int k = 0; k < 3; ++k) {
                r[k] = s.r[k][i];
                v[k] = s.v[k][i];
            }
            mf = (float)s.m[i];
        }
    }
    bool live;
    double itime;
    if (flags & PS_TAIL) {
        // the cluster record, read ahead of the barrier (this launch writes
        // none of these fields that another workgroup reads)
        const uint32_t done = cs->done, pend = cs->tail_pend;
        const double tail_t = cs->tail_t, atime = cs->atime;
        const unsigned long long tail_steps = cs->tail_steps;
        const uint32_t tail_nact = cs->tail_nact;
        const uint32_t b0 = single ? 0u : e.off[c] / JT, nb = single ? e.nblk : ntiles(e, c);
        double mn = INFINITY;
        for (uint32_t k = threadIdx.x; k < nb; k += BS) mn = fmin(mn, e.bmin[b0 + k]);
        for (int o = 32; o > 0; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o));
        if (__lane_id() == 0) wmin[threadIdx.x >> 6] = mn;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t k = 1; k < BS / 64; ++k) mn = fmin(mn, wmin[k]);
            uint32_t go = 0;
            double T = 0.0;
            if (done) {
            } else if (pend) {
                const unsigned long long steps = tail_steps + 1;
                const bool stop = loop_stops(*e.ctl, tail_t, steps, mn);
                if (b == b0) record_tail(e, c, tail_t, tail_nact, steps, mn, stop);
                go = !stop;
                T = mn;
            } else {
                go = 1;
                T = atime;
            }
            if (b == 0) e.ctl->ndirty[par ^ 1] = 0;
            s_go = go;
            s_time = T;
        }
        __syncthreads();
        live = s_go != 0;
        itime = s_time;
    } else {
        live = (flags & PS_DEVTIME) ? cs->done == 0 : true;
        itime = (flags & PS_DEVTIME) ? cs->atime : itime_arg;
    }
    const bool in = live && (i - e.off[c]) < e.cnt[c];
    bool act = false;
    double4 pa = {}, pb = {};
    if (in) {
        if (flags & PS_PREDICT) {
            predict_vals(r, v, f, itime - ti, mf, pa, pb);
            s.pp[i] = pa;
            s.pv[i] = pb;
        }
        if (flags & PS_SELECT) {
            double tmp_time = ti + dti;
            act = fabs(itime - tmp_time) < 2 * DBL_EPSILON;
        }
    }
    if (flags & PS_SELECT) {
        unsigned long long mask = __ballot(act);
        uint32_t lane = __lane_id();
        uint32_t w = threadIdx.x >> 6;
        if (lane == 0) wave_off[w] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t total = 0;
            for (uint32_t k = 0; k < BS / 64; ++k) {
                uint32_t n = wave_off[k];
                wave_off[k] = total;
                total += n;
            }
            wave_off[BS / 64] = total ? atomicAdd(&cs->nact, total) : 0;
            if (total && (flags & PS_DIRTY)) {
                const uint32_t d = atomicAdd(&e.ctl->ndirty[par], 1u);
                uint32_t *dl = e.dirty + (size_t)par * 3 * e.nblk;
                dl[3 * d] = b;
                dl[3 * d + 1] = wave_off[BS / 64];
                dl[3 * d + 2] = total;
            }
        }
        __syncthreads();
        if (act) {
            unsigned long long below = mask & ((1ull << lane) - 1ull);
            const uint32_t slot =
                e.off[c] + wave_off[BS / 64] + wave_off[w] + (uint32_t)__popcll(below);
            move[slot] = i;
            if (flags & PS_SAVE_OLD)
                for (int k = 0; k < 6; ++k) s.old[k][i] = f[k];
            if (flags & PS_COMPACT) {
                const size_t np = (size_t)e.nblk * JT;
                e.arec[slot] = pa;
                e.arec[np + slot] = pb;
                e.arec[2 * np + slot] = make_double4(f[0], f[1], f[2], dti);
                e.arec[3 * np + slot] = make_double4(f[3], f[4], f[5], 0.0);
            }
        }
    }
}

__global__ __launch_bounds__(BS) void k_save_old(Soa s, Ens e, const uint32_t *move) {
    for (uint32_t c = 0; c < e.K; ++c) {
        uint32_t nact = e.cs[c].nact;
        for (uint32_t k = blockIdx.x * BS + threadIdx.x; k < nact; k += gridDim.x * BS) {
            uint32_t i = move[e.off[c] + k];
            for (int q = 0; q < 6; ++q) s.old[q][i] = s.f[q][i];
        }
    }
}

// ---------------------------------------------------------------------------
// Pairwise acceleration + jerk
// ---------------------------------------------------------------------------

// Loads JT consecutive predictor records starting at j0 into an LDS tile;
// each of the BS (= JT) threads moves one record (two coalesced double4
// loads).  ES = 1: pp records at [0, JT), pv records at [JT, 2 JT);
// ES = 2: record t at [2t] (pp) and [2t + 1] (pv).
template <uint32_t ES = 1>
__device__ __forceinline__ void load_jtile(const Soa &s, uint32_t j0, double4 *tile) {
    tile[ES * threadIdx.x] = s.pp[j0 + threadIdx.x];
    tile[ES == 2 ? 2 * threadIdx.x + 1 : JT + threadIdx.x] = s.pv[j0 + threadIdx.x];
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) const void global_void_t;

// Copies predictor records j0 .. j0 + JT - 1 into an LDS tile (pp records
// at [0, JT), pv records at [JT, 2 JT)) with global_load_lds_dwordx4: the
// data goes from memory to LDS without registers, so the copy of the next
// tile runs while the workgroup computes on the current one; the barrier
// after the computation waits for it (vmcnt).  Each wave moves 64
// consecutive 16-byte units per instruction to consecutive LDS units.
__device__ __forceinline__ void dma_jtile(const Soa &s, uint32_t j0, double4 *dst) {
    const double2 *gp = reinterpret_cast<const double2 *>(s.pp + j0);
    const double2 *gv = reinterpret_cast<const double2 *>(s.pv + j0);
    double2 *lp = reinterpret_cast<double2 *>(dst);
    double2 *lv = reinterpret_cast<double2 *>(dst + JT);
    const uint32_t w = threadIdx.x & ~63u;  // first unit of this wave
    static_assert(2 * JT == 2 * BS, "a tile is 2 x BS 16-byte units per array");
    __builtin_amdgcn_global_load_lds((global_void_t *)(gp + threadIdx.x), (lds_void_t *)(lp + w), 16,
                                     0, 0);
    __builtin_amdgcn_global_load_lds((global_void_t *)(gp + BS + threadIdx.x),
                                     (lds_void_t *)(lp + BS + w), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((global_void_t *)(gv + threadIdx.x), (lds_void_t *)(lv + w), 16,
                                     0, 0);
    __builtin_amdgcn_global_load_lds((global_void_t *)(gv + BS + threadIdx.x),
                                     (lds_void_t *)(lv + BS + w), 16, 0, 0);
}

// Predicted position and velocity of particle i.
__device__ __forceinline__ void load_predictor(const Soa &s, uint32_t i, double pi[6]) {
    const double4 a = s.pp[i], b = s.pv[i];
    pi[0] = a.x;
    pi[1] = a.y;
    pi[2] = a.z;
    pi[3] = b.x;
    pi[4] = b.y;
    pi[5] = b.z;
}

__device__ __forceinline__ float mass_f32(const double4 &b) {
    return __builtin_bit_cast(float, (uint32_t)__builtin_bit_cast(unsigned long long, b.w));
}

// Inner loop of one lane over one LDS tile: the lane's j sub-group js
// takes tile entries js, js + G, ..., G = BS/IW (all JT entries when
// IW = BS).  Entry jj is sp[ES jj] (position, mass) and sv[ES jj]
// (velocity): ES = 1 for a tile stored as two arrays, 2 for one stored as
// interleaved record pairs.  SELF: the tile may hold the lane's own i, whose term is
// skipped (mass 0, and r2 kept finite when e2 == 0).  MIXED: FP32 pair
// terms summed in FP32 over runs of at most 32 pairs, each run then added
// to the FP64 accumulators.
template <int IW, bool SELF, bool MIXED, uint32_t ES>
__device__ __forceinline__ void tile_pairs(const double pi[6], uint32_t i,
                                           uint32_t j0, uint32_t js,
                                           const double4 *sp, const double4 *sv,
                                           double e2, double fo[6]) {
    constexpr uint32_t G = BS / IW;
    constexpr uint32_t NQ = JT / G;
    if constexpr (!MIXED) {
#pragma unroll 4
        for (uint32_t q = 0; q < NQ; ++q) {
            uint32_t jj = js + q * G;
            double4 a = sp[ES * jj];
            double4 b = sv[ES * jj];
            if (SELF) {
                bool me = (j0 + jj) == i;
                double mj = me ? 0.0 : a.w;
                double x = me ? pi[0] + 1.0 : a.x;
                pair_fp64(pi, x, a.y, a.z, b.x, b.y, b.z, mj, e2, fo);
            } else {
                pair_fp64(pi, a.x, a.y, a.z, b.x, b.y, b.z, a.w, e2, fo);
            }
        }
    } else {
        constexpr uint32_t FOLD = NQ < H4_MIXED_FOLD ? NQ : H4_MIXED_FOLD;
        const float e2f = (float)e2;
        for (uint32_t q0 = 0; q0 < NQ; q0 += FOLD) {
            float t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll H4_MIXED_UNROLL
            for (uint32_t q = q0; q < q0 + FOLD; ++q) {
                uint32_t jj = js + q * G;
                double4 a = sp[ES * jj];
                double4 b = sv[ES * jj];
                if (SELF) {
                    bool me = (j0 + jj) == i;
                    float mj = me ? 0.f : mass_f32(b);
                    double x = me ? pi[0] + 1.0 : a.x;
                    pair_mixed(pi, x, a.y, a.z, b.x, b.y, b.z, mj, e2f, t);
                } else {
                    pair_mixed(pi, a.x, a.y, a.z, b.x, b.y, b.z, mass_f32(b), e2f, t);
                }
            }
            for (int c = 0; c < 6; ++c) fo[c] += (double)t[c];
        }
    }
}

// Sums fo over the BS/IW j sub-groups of each i slot: shuffles across the
// sub-groups inside a wave (lane bits >= log2(IW)), then LDS across the
// BS/64 waves in wave order.  The fixed tree makes it deterministic.  On
// return lanes threadIdx.x < IW hold the sums of slot threadIdx.x.
template <int IW>
__device__ __forceinline__ void reduce_subgroups(double fo[6], double *red) {
    if constexpr (IW < 64) {
#pragma unroll
        for (int o = IW; o < 64; o <<= 1)
            for (int c = 0; c < 6; ++c) fo[c] += __shfl_xor(fo[c], o);
    }
    if constexpr (IW < BS) {
        constexpr uint32_t W = BS / 64;            // waves
        constexpr uint32_t SL = IW < 64 ? IW : 64;  // slots held per wave
        uint32_t lane = __lane_id();
        uint32_t w = threadIdx.x >> 6;
        __syncthreads();
        if (lane < SL) {
            // wave w holds slots (w*64 .. w*64+63) % IW when IW >= 64
            uint32_t slot = (IW >= 64) ? (threadIdx.x % IW) : lane;
            uint32_t grp = (IW >= 64) ? (threadIdx.x / IW) : w;
            for (int c = 0; c < 6; ++c) red[(grp * IW + slot) * 6 + c] = fo[c];
        }
        __syncthreads();
        if (threadIdx.x < IW) {
            constexpr uint32_t NG = (IW >= 64) ? BS / IW : W;
            for (int c = 0; c < 6; ++c) {
                double acc = red[threadIdx.x * 6 + c];
                for (uint32_t g = 1; g < NG; ++g) acc += red[(g * IW + threadIdx.x) * 6 + c];
                fo[c] = acc;
            }
        }
    }
}

// FP64_EXACT: one lane per active i, j ascending over the cluster, j == i
// skipped, exactly Hermite4CPU::update_acc_jrk (Hermite4CPU.cpp:309-336).
// Writes f[i] directly.  Workgroups stride over the plan's i tiles.
__global__ __launch_bounds__(BS) void k_update_exact(Soa s, Ens e, const uint32_t *move,
                                                    double e2) {
    __shared__ double4 tile[2 * JT];
    const double4 *sp = tile, *sv = tile + JT;
    const bool single = e.K == 1;
    unsigned long long total = 0;
    if (single)
        (void)plan_single(e, 1u, total);
    else
        total = e.ctl->items;
    for (unsigned long long item = blockIdx.x; item < total; item += gridDim.x) {
        const uint32_t c = single ? 0u : find_cluster(e.items_pre, e.K, item);
        const uint32_t it = (uint32_t)(item - (single ? 0ull : e.items_pre[c]));
        const uint32_t nact = e.cs[c].nact;
        const uint32_t *mv = move + e.off[c];
        const uint32_t j_begin = e.off[c], j_end = e.off[c] + e.cnt[c];
        uint32_t k = it * BS + threadIdx.x;
        bool act = k < nact;
        uint32_t i = act ? mv[k] : mv[0];
        double pi[6];
        load_predictor(s, i, pi);
        double fo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (uint32_t j0 = j_begin; j0 < j_end; j0 += JT) {
            __syncthreads();
            load_jtile(s, j0, tile);
            __syncthreads();
            uint32_t jn = min((uint32_t)JT, j_end - j0);
            for (uint32_t jj = 0; jj < jn; ++jj) {
                double4 a = sp[jj];
                double4 b = sv[jj];
                if (j0 + jj != i)
                    pair_exact(pi, a.x, a.y, a.z, b.x, b.y, b.z, a.w, e2, fo);
            }
        }
        if (act)
            for (int q = 0; q < 6; ++q) s.f[q][i] = fo[q];
    }
}

// One work item of the FP64 / mixed update: i tile `it` of a cluster
// against j chunk `sc`.  With one chunk the result goes to f[i]; with
// several, the chunk sums go to partial[partial_at(pl, comp, sc, k)] and
// the finish pass adds them in chunk order.
template <int IW, bool MIXED>
__device__ __forceinline__ void update_item(const Soa &s, const ClPlan &pl, uint32_t jt_begin,
                                            uint32_t jt_end, uint32_t it, uint32_t sc,
                                            const uint32_t *mv, const double4 *cr, size_t np,
                                            uint32_t nact, double *partial, double e2,
                                            double4 *lA, double4 *lB, double *red) {
    const uint32_t il = threadIdx.x % IW;
    const uint32_t js = threadIdx.x / IW;
    uint32_t k = it * IW + il;
    bool act = k < nact;
    uint32_t t0 = jt_begin + sc * pl.tpc;
    uint32_t t1 = min(jt_end, t0 + pl.tpc);
    // FP64 wide groups keep their tile as interleaved (position, velocity)
    // record pairs: every LDS read of the pair loop then fits the
    // instructions' offset field, with no extra address arithmetic (0.5 %
    // on the full step; the mixed kernel runs as fast with two arrays).
    constexpr uint32_t ES = (IW >= 64 && !MIXED) ? 2u : 1u;
    // The first j tile is requested before the i particles, whose loads wait
    // for their indices, so the two memory latencies overlap: narrow groups
    // start its copy into lA (once the previous item is done with lA: tiles
    // or red), wide groups load it into registers.
    [[maybe_unused]] double4 ja, jb;
    if constexpr (IW < 64) {
        __syncthreads();
        dma_jtile(s, t0 * JT, lA);
    } else {
        ja = s.pp[t0 * JT + threadIdx.x];
        jb = s.pv[t0 * JT + threadIdx.x];
    }
    // The i particles' predictors: from the selection's records in
    // active-list order when it wrote them (cr; then their loads wait for
    // no index), else through the active list.
    uint32_t i = act ? mv[k] : mv[0];
    double pi[6];
    if (cr) {
        const uint32_t kk = act ? k : 0u;
        const double4 a = cr[kk], b = cr[np + kk];
        pi[0] = a.x;
        pi[1] = a.y;
        pi[2] = a.z;
        pi[3] = b.x;
        pi[4] = b.y;
        pi[5] = b.z;
    } else {
        load_predictor(s, i, pi);
    }
    double fo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    auto tile = [&](const double4 *buf, uint32_t tt) {
        const uint32_t j0 = tt * JT;
        const bool self = (i - j0) < (uint32_t)JT;
        const double4 *sv = ES == 2 ? buf + 1 : buf + JT;
        if (__ballot(self))
            tile_pairs<IW, true, MIXED, ES>(pi, i, j0, js, buf, sv, e2, fo);
        else
            tile_pairs<IW, false, MIXED, ES>(pi, i, j0, js, buf, sv, e2, fo);
    };
    // Narrow lane groups (few pairs per lane and tile): double-buffered
    // tiles, the copy of tile tt + 1 in flight while tile tt is computed, one
    // barrier per tile.  The loop is written out for the two buffers so that
    // the compiler sees the copy's target and the tile being read as
    // different LDS objects and does not wait for the copy before each
    // read.  red overlays lA, whose last reader has passed a barrier before
    // reduce_subgroups.
    if constexpr (IW < 64) {
        __syncthreads();  // the first tile's copy has landed
        for (uint32_t tt = t0; tt < t1; tt += 2) {
            if (tt + 1 < t1) dma_jtile(s, (tt + 1) * JT, lB);
            tile(lA, tt);
            __syncthreads();
            if (tt + 1 >= t1) break;
            if (tt + 2 < t1) dma_jtile(s, (tt + 2) * JT, lA);
            tile(lB, tt + 1);
            __syncthreads();
        }
    } else {
        // Wide lane groups do 64-256 pairs per lane and tile: the copy is
        // a small share of the time and one buffer keeps registers free.
        __syncthreads();  // the previous item is done with lA (tiles or red)
        lA[ES * threadIdx.x] = ja;
        lA[ES == 2 ? 2 * threadIdx.x + 1 : JT + threadIdx.x] = jb;
        __syncthreads();
        tile(lA, t0);
        for (uint32_t tt = t0 + 1; tt < t1; ++tt) {
            __syncthreads();
            load_jtile<ES>(s, tt * JT, lA);
            __syncthreads();
            tile(lA, tt);
        }
    }
    reduce_subgroups<IW>(fo, red);
    if (threadIdx.x < IW && act) {
        constexpr double scale = MIXED ? 1.0 : PAIR_SCALE;  // pair_fp64's 3 sqrt(3)
        if (pl.S == 1) {
            for (int q = 0; q < 6; ++q) s.f[q][i] = fo[q] * scale;
        } else {
            // partial_at(pl, q, sc, k), with the layout's strides
            const size_t kst = pl.cmaj ? pl.S : 1u, cst = pl.cmaj ? 1u : pl.rows;
            double *pq = partial + (size_t)sc * cst + (size_t)k * kst;
            const size_t qs = (size_t)pl.S * pl.rows;
            for (int q = 0; q < 6; ++q) pq[q * qs] = fo[q] * scale;
        }
    }
}

// FP64 and mixed precision: persistent workgroups over the plan's work
// items of all clusters; each item's lane-group width comes from the plan,
// so one launch serves every active-set size.
template <bool MIXED>
__device__ __forceinline__ void update_items(const Soa &s, const Ens &e, const uint32_t *move,
                                             double *partial, double e2, uint32_t compact) {
    __shared__ double4 lA[2 * JT];
    __shared__ double4 lB[2 * JT];
    static_assert(BS * 6 * sizeof(double) <= sizeof(lA), "red overlays lA");
    double *red = reinterpret_cast<double *>(lA);
    const bool single = e.K == 1;
    unsigned long long total = 0;
    const ClPlan pl0 = single ? plan_single(e, 0u, total) : ClPlan{};
    if (!single) total = e.ctl->items;
    // The first gridDim.x items go one to each workgroup; a workgroup that
    // has finished one takes the next unclaimed item from the queue counter,
    // so faster CUs and XCDs take more.  The last workgroup to leave the
    // queue zeroes it for the next launch.
    __shared__ unsigned long long claimed;
    unsigned long long item = blockIdx.x;
    while (item < total) {
        const uint32_t c = single ? 0u : find_cluster(e.items_pre, e.K, item);
        const ClPlan pl = single ? pl0 : e.plan[c];
        const uint32_t li = (uint32_t)(item - (single ? 0ull : e.items_pre[c]));
        // j-chunk major: items taken at about the same time share their j
        // tiles (one chunk per XCD L2 rather than S of them)
        const uint32_t sc = li / pl.n_it;
        const uint32_t it = li - sc * pl.n_it;
        const uint32_t nact = e.cs[c].nact;
        const uint32_t jt_begin = e.off[c] / JT;
        const uint32_t jt_end = jt_begin + ntiles(e, c);
        const uint32_t *mv = move + e.off[c];
        const double4 *cr = compact ? e.arec + e.off[c] : nullptr;
        const size_t np = (size_t)e.nblk * JT;
        double *part = partial + (single ? 0ull : e.part_pre[c]);
        switch (pl.iw) {
            case 4:
                update_item<4, MIXED>(s, pl, jt_begin, jt_end, it, sc, mv, cr, np, nact, part, e2,
                                      lA, lB, red);
                break;
            case 16:
                update_item<16, MIXED>(s, pl, jt_begin, jt_end, it, sc, mv, cr, np, nact, part, e2,
                                       lA, lB, red);
                break;
            case 64:
                update_item<64, MIXED>(s, pl, jt_begin, jt_end, it, sc, mv, cr, np, nact, part, e2,
                                       lA, lB, red);
                break;
            default:
                update_item<256, MIXED>(s, pl, jt_begin, jt_end, it, sc, mv, cr, np, nact, part, e2,
                                        lA, lB, red);
                break;
        }
        if (total <= gridDim.x) break;
        __syncthreads();  // every lane has read the previous claim
        if (threadIdx.x == 0) claimed = gridDim.x + atomicAdd(&e.ctl->qnext, 1ull);
        __syncthreads();
        item = claimed;
    }
    // Every workgroup claims until it draws an index >= total and then
    // counts itself out; when all have, no claim of this launch remains.
    if (total > gridDim.x && threadIdx.x == 0 &&
        atomicAdd(&e.ctl->qdone, 1u) == gridDim.x - 1) {
        atomicExch(&e.ctl->qnext, 0ull);
        atomicExch(&e.ctl->qdone, 0u);
    }
}

// The four lane-group variants in one kernel would take 144 VGPRs (3 waves
// per SIMD); each alone needs <= 126, so the FP64 kernel is held to 128
// (4 waves per SIMD; a few per-item values spill to scratch, none inside
// the tile or pair loops).  The mixed kernel fits in 120 by itself.
// compact: the step's selection wrote the active particles' records
// (k_select PS_COMPACT).
__global__ __launch_bounds__(BS, 4) void k_update_fp64(Soa s, Ens e, const uint32_t *move,
                                                      double *partial, double e2,
                                                      uint32_t compact) {
    update_items<false>(s, e, move, partial, e2, compact);
}

__global__ __launch_bounds__(BS) void k_update_mixed(Soa s, Ens e, const uint32_t *move,
                                                    double *partial, double e2,
                                                    uint32_t compact) {
    update_items<true>(s, e, move, partial, e2, compact);
}

// ---------------------------------------------------------------------------
// Corrector + Aarseth timestep + block quantisation
// ---------------------------------------------------------------------------

// NbodyUtils::get_magnitude (NbodyUtils.cpp:331-334).
__device__ __forceinline__ double magnitude(double x, double y, double z) {
#pragma clang fp contract(off)
    return sqrt(x * x + y * y + z * z);
}

// NbodyUtils::normalize_dt (NbodyUtils.cpp:371-424), bounds as arguments.
__device__ __forceinline__ double normalize_dt(double new_dt, double old_dt,
                                               double t, double dt_min,
                                               double dt_max) {
#pragma clang fp contract(off)
    if (new_dt <= old_dt / 8) {
        new_dt = dt_min;
    } else if (old_dt / 8 < new_dt && new_dt <= old_dt / 4) {
        new_dt = old_dt / 8;
    } else if (old_dt / 4 < new_dt && new_dt <= old_dt / 2) {
        new_dt = old_dt / 4;
    } else if (old_dt / 2 < new_dt && new_dt <= old_dt) {
        new_dt = old_dt / 2;
    } else if (old_dt < new_dt && new_dt <= old_dt * 2) {
        new_dt = old_dt;
    } else if (2 * old_dt < new_dt) {
        double val = t / (2 * old_dt);
        new_dt = (ceil(val) == val) ? 2.0 * old_dt : old_dt;
    } else {
        new_dt = old_dt;
    }
    if (new_dt < dt_min)
        new_dt = dt_min;
    else if (new_dt > dt_max)
        new_dt = dt_max;
    return new_dt;
}

// Hermite4CPU::correction_pos_vel (Hermite4CPU.cpp:370-411) with
// NbodyUtils::get_timestep_normal (NbodyUtils.cpp:337-365) of one particle
// at block time itime: old forces o, new forces f, predictor pr (position,
// velocity), step dt1, ETA (the reference's float, widened exactly).  Writes
// the corrected position rn and velocity vn and a2, a3; returns the new step.
__device__ __forceinline__ double correct_vals(const double o[6], const double f[6],
                                               const double pr[6], double dt1, double itime,
                                               double eta, double dt_min, double dt_max,
                                               double rn[3], double vn[3], double a2[3],
                                               double a3[3]) {
#pragma clang fp contract(off)
    double dt2 = dt1 * dt1;
    double dt3 = dt2 * dt1;
    double dt4 = dt2 * dt2;
    double dt5 = dt4 * dt1;
    for (int c = 0; c < 3; ++c)
        a2[c] = (-6 * (o[c] - f[c]) - dt1 * (4 * o[3 + c] + 2 * f[3 + c])) / dt2;
    for (int c = 0; c < 3; ++c)
        a3[c] = (12 * (o[c] - f[c]) + 6 * dt1 * (o[3 + c] + f[3 + c])) / dt3;
    for (int c = 0; c < 3; ++c) {
        rn[c] = pr[c] + (dt4 / 24) * a2[c] + (dt5 / 120) * a3[c];
        vn[c] = pr[3 + c] + (dt3 / 6) * a2[c] + (dt4 / 24) * a3[c];
    }
    double ax1_2 = a2[0] + dt1 * a3[0];
    double ay1_2 = a2[1] + dt1 * a3[1];
    double az1_2 = a2[2] + dt1 * a3[2];
    double abs_a1 = magnitude(f[0], f[1], f[2]);
    double abs_j1 = magnitude(f[3], f[4], f[5]);
    double abs_j12 = abs_j1 * abs_j1;
    double abs_a1_3 = magnitude(a3[0], a3[1], a3[2]);
    double abs_a1_2 = magnitude(ax1_2, ay1_2, az1_2);
    double abs_a1_22 = abs_a1_2 * abs_a1_2;
    double normal_dt = sqrt(eta * ((abs_a1 * abs_a1_2 + abs_j12) /
                                   (abs_j1 * abs_a1_3 + abs_a1_22)));
    return normalize_dt(normal_dt, dt1, itime, dt_min, dt_max);
}

// The corrector of particle i at block time itime, given its new forces
// f[6] (old forces in s.old, predictor in s.pp / s.pv), on the state in
// HBM.  Returns the new dt.
__device__ __forceinline__ double correct_store(const Soa &s, uint32_t i, const double o[6],
                                                const double pr[6], double dt1, const double f[6],
                                                double itime, float eta, double dt_min,
                                                double dt_max) {
    double rn[3], vn[3], a2[3], a3[3];
    const double dt_new =
        correct_vals(o, f, pr, dt1, itime, eta, dt_min, dt_max, rn, vn, a2, a3);
    for (int c = 0; c < 3; ++c) {
        s.r[c][i] = rn[c];
        s.v[c][i] = vn[c];
        s.a2[c][i] = a2[c];
        s.a3[c][i] = a3[c];
    }
    s.t[i] = itime;
    s.dt[i] = dt_new;
    return dt_new;
}

__device__ __forceinline__ double correct_one(const Soa &s, uint32_t i, const double f[6],
                                              double itime, float eta, double dt_min,
                                              double dt_max) {
    double o[6], pr[6];
    for (int c = 0; c < 6; ++c) o[c] = s.old[c][i];
    load_predictor(s, i, pr);
    return correct_store(s, i, o, pr, s.dt[i], f, itime, eta, dt_min, dt_max);
}

// The corrector's inputs of the particle at place k of the step's active
// list from the selection's records (Ens::arec): old forces o, predictor
// pr, step dt1.
__device__ __forceinline__ void load_arec(const Ens &e, size_t k, double o[6], double pr[6],
                                          double &dt1) {
    const size_t np = (size_t)e.nblk * JT;
    const double4 p = e.arec[k], q = e.arec[np + k], a = e.arec[2 * np + k],
                  j = e.arec[3 * np + k];
    pr[0] = p.x;
    pr[1] = p.y;
    pr[2] = p.z;
    pr[3] = q.x;
    pr[4] = q.y;
    pr[5] = q.z;
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
    o[3] = j.x;
    o[4] = j.y;
    o[5] = j.z;
    dt1 = a.w;
}

// Standalone corrector of cluster 0's active list at itime (the ABI call
// h4h_correction_pos_vel): forces from s.f.
__global__ __launch_bounds__(BS) void k_correct(Soa s, Ens e, const uint32_t *move, double itime,
                                               float eta, double dt_min, double dt_max) {
    const uint32_t nact = e.cs[0].nact;
    for (uint32_t k = blockIdx.x * BS + threadIdx.x; k < nact; k += gridDim.x * BS) {
        uint32_t i = move[k];
        double f[6];
        for (int c = 0; c < 6; ++c) f[c] = s.f[c][i];
        correct_one(s, i, f, itime, eta, dt_min, dt_max);
    }
}

enum : uint32_t { FIN_WRITE_F = 0, FIN_CORRECT = 1 };

// Sum of active particle k's S j-chunk partials (plan pl, partials of its
// cluster at part) on the group of L = pl.lanes consecutive lanes of one
// wave that share it: lane sl adds chunks sl, sl + L, ... in that order,
// then a fixed xor butterfly over the group, so the order is fixed for a
// given plan and results are deterministic.  With one chunk the update
// stored the forces in f[i].  Every lane of the group gets the sums.
__device__ __forceinline__ void chunk_sums(const Soa &s, const ClPlan &pl, const double *part,
                                           uint32_t k, uint32_t i, bool valid, double f[6]) {
    const uint32_t L = pl.lanes;
    const uint32_t sl = threadIdx.x % L;
    if (pl.S == 1) {
        for (int q = 0; q < 6; ++q) f[q] = valid ? s.f[q][i] : 0.0;
        return;
    }
    // The loads of a round of U chunks of all six components are issued
    // together (a few memory latencies per round instead of one per chunk;
    // S <= 1024 takes two rounds at L = 64).
    constexpr uint32_t U = 8;
    // stride between consecutive chunks and between components (partial_at)
    const size_t cs = pl.cmaj ? 1 : pl.rows;
    const size_t qs = (size_t)pl.S * pl.rows;
    const double *pk = part + (pl.cmaj ? (size_t)k * pl.S : (size_t)k);
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (valid) {
        for (uint32_t s0 = sl; s0 < pl.S; s0 += U * L) {
            double v[6][U];
#pragma unroll
            for (int q = 0; q < 6; ++q)
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t sc = s0 + u * L;
                    v[q][u] = sc < pl.S ? pk[q * qs + (size_t)sc * cs] : 0.0;
                }
#pragma unroll
            for (int q = 0; q < 6; ++q)
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    if (s0 + u * L < pl.S) acc[q] += v[q][u];
        }
    }
    for (int q = 0; q < 6; ++q) {
        for (uint32_t o = 1; o < L; o <<= 1) acc[q] += __shfl_xor(acc[q], o);
        f[q] = acc[q];
    }
}

// Finish pass of an update: adds the j-chunk partials of each active
// particle (chunk_sums) and either stores the forces (FIN_WRITE_F) or
// applies the corrector at the cluster's block time ATIME (FIN_CORRECT,
// after a selection that wrote the active particles' records, whose loads
// go out with those of the partials).  A particle's chunk sums are spread
// over L = pl.lanes lanes, so a workgroup finishes BS/L particles with no
// barrier.
__global__ __launch_bounds__(BS) void k_finish(Soa s, Ens e, const uint32_t *move,
                                              const double *partial, uint32_t mode, float eta,
                                              double dt_min, double dt_max) {
    const unsigned long long total = e.ctl->fin_units;
    for (unsigned long long u = blockIdx.x; u < total; u += gridDim.x) {
        const uint32_t c = find_cluster(e.fin_pre, e.K, u);
        const ClPlan pl = e.plan[c];
        const uint32_t L = pl.lanes;
        const uint32_t lu = (uint32_t)(u - e.fin_pre[c]);
        const uint32_t nact = e.cs[c].nact;
        const double T = e.cs[c].atime;
        const uint32_t k = lu * (BS / L) + threadIdx.x / L;
        const bool valid = k < nact;
        const uint32_t i = valid ? move[e.off[c] + k] : 0u;
        double o[6], pr[6], dt1 = 0.0;
        if (mode == FIN_CORRECT && valid) load_arec(e, (size_t)e.off[c] + k, o, pr, dt1);
        double f[6];
        chunk_sums(s, pl, partial + e.part_pre[c], k, i, valid, f);
        if (valid && threadIdx.x % L == 0) {
            if (mode == FIN_CORRECT) correct_store(s, i, o, pr, dt1, f, T, eta, dt_min, dt_max);
            if (pl.S > 1)
                for (int q = 0; q < 6; ++q) s.f[q][i] = f[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Block times
// ---------------------------------------------------------------------------

__device__ __forceinline__ double block_min(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    uint32_t w = threadIdx.x >> 6;
    if (__lane_id() == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < BS / 64; ++k) v = fmin(v, red[k]);
    }
    return v;  // valid in thread 0
}

// min(t + dt) of one JT block (padding has dt = +inf).
__device__ __forceinline__ double block_tmin(const Soa &s, uint32_t b, double *red) {
    uint32_t i = b * JT + threadIdx.x;
    double v = s.t[i] + s.dt[i];
    double m = block_min(v, red);
    return m;  // valid in thread 0
}

// Recomputes the per-block minima of the blocks that had active particles
// this step (dirty list `par`; the others did not change), and records each
// live cluster's step for its loop tail, which the next launch that needs
// the cluster's next block time makes (k_select PS_TAIL, k_small_run); the
// cluster's active count of the next step starts from zero.
__global__ __launch_bounds__(BS) void k_blockmin_dirty(Soa s, Ens e, uint32_t par) {
    __shared__ double red[BS / 64];
    for (uint32_t c = blockIdx.x * BS + threadIdx.x; c < e.K; c += gridDim.x * BS) {
        ClState *cs = &e.cs[c];
        if (cs->done) continue;
        cs->tail_t = cs->atime;
        cs->tail_nact = cs->nact;
        cs->tail_steps = cs->steps;
        cs->tail_pend = 1;
        cs->nact = 0;
    }
    // this workgroup's first list entry is read with the count (the list
    // has nblk entries and the grid is at most nblk)
    const uint32_t *dl = e.dirty + (size_t)par * 3 * e.nblk;
    uint32_t b = dl[3 * blockIdx.x];
    const uint32_t nd = e.ctl->ndirty[par];
    for (uint32_t w = blockIdx.x; w < nd; w += gridDim.x) {
        if (w != blockIdx.x) b = dl[3 * w];
        __syncthreads();
        double m = block_tmin(s, b, red);
        if (threadIdx.x == 0) e.bmin[b] = m;
    }
}

// Finish pass of a block step of a single system, one block with active
// particles per workgroup (the dirty list k_select made for the step's
// parity: block, first place in the active list, count): the chunk sums
// and the corrector with
// NbodyUtils::get_timestep_normal and normalize_dt (Hermite4CPU.cpp:370-411,
// NbodyUtils.cpp:337-424) of the block's active particles, then the block's
// min(t + dt), from which the loop tail takes the next block time
// (Hermite4::next_integration_time, Hermite4.cpp:90-102): one launch for
// what k_finish and k_blockmin_dirty do in two.  The sums are chunk_sums',
// so the forces are k_finish's bit for bit.  Workgroup 0 records the step
// for its loop tail (ClState tail_*), which the next launch that needs the
// next block time makes.  (An ensemble keeps the two
// launches: its many blocks with one or two active particles each would
// put one corrector per workgroup instead of up to BS/L.)
__global__ __launch_bounds__(BS) void k_finish_blocks(Soa s, Ens e, const uint32_t *move,
                                                     const double *partial, float eta,
                                                     double dt_min, double dt_max, uint32_t par) {
    __shared__ double red[BS / 64];
    __shared__ double tnew[JT];  // t + dt of the block's corrected particles (NaN: not active)
    ClState *cs = &e.cs[0];
    const uint32_t *dl = e.dirty + (size_t)par * 3 * e.nblk;
    // What the first block needs is requested at once: the loop's stop
    // flag, the list's count, this workgroup's list entry (read before the
    // count is known: the list has nblk entries and the grid is at most
    // nblk; an entry beyond the count is not used), the plan and the block
    // time.
    const uint32_t done = cs->done;
    const uint32_t nd = e.ctl->ndirty[par];
    uint32_t w = blockIdx.x;
    uint32_t b = dl[3 * w], k0 = dl[3 * w + 1], cnt = dl[3 * w + 2];
    const ClPlan pl = e.plan[0];
    const double T = cs->atime;
    if (done) return;  // the loop stopped at this step's selection
    const uint32_t L = pl.lanes, per = BS / L;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // this step's loop tail, for the next step's selection (k_select
        // PS_TAIL), k_small_run or k_tail to make; the active count of the
        // next step starts from zero
        cs->tail_t = T;
        cs->tail_nact = cs->nact;
        cs->tail_steps = cs->steps;
        cs->tail_pend = 1;
        cs->nact = 0;
    }
    for (; w < nd; w += gridDim.x) {
        if (w != blockIdx.x) {
            b = dl[3 * w];
            k0 = dl[3 * w + 1];
            cnt = dl[3 * w + 2];
        }
        // t + dt of the block as it stands, loaded ahead of the corrector
        const uint32_t j = b * JT + threadIdx.x;
        const double tdt = s.t[j] + s.dt[j];
        __syncthreads();  // red and tnew of the previous block have been read
        tnew[threadIdx.x] = NAN;
        __syncthreads();
        for (uint32_t base = 0; base < cnt; base += per) {
            const uint32_t kk = base + threadIdx.x / L;
            const bool valid = kk < cnt;
            const uint32_t k = k0 + (valid ? kk : 0u);
            const uint32_t i = move[k];
            // the corrector's inputs from the selection's records: their
            // loads go out with those of the partials
            double o[6], pr[6], dt1 = 0.0;
            if (valid) load_arec(e, k, o, pr, dt1);
            double f[6];
            chunk_sums(s, pl, partial, k, i, valid, f);
            if (valid && threadIdx.x % L == 0) {
                const double dtn = correct_store(s, i, o, pr, dt1, f, T, eta, dt_min, dt_max);
                if (pl.S > 1)
                    for (int q = 0; q < 6; ++q) s.f[q][i] = f[q];
                tnew[i - b * JT] = T + dtn;
            }
        }
        __syncthreads();
        const double tn = tnew[threadIdx.x];
        const double m = block_min(isnan(tn) ? tdt : tn, red);
        if (threadIdx.x == 0) e.bmin[b] = m;
    }
}

__global__ __launch_bounds__(BS) void k_blockmin_all(Soa s, Ens e) {
    __shared__ double red[BS / 64];
    for (uint32_t b = blockIdx.x; b < e.nblk; b += gridDim.x) {
        __syncthreads();
        double m = block_tmin(s, b, red);
        if (threadIdx.x == 0) e.bmin[b] = m;
    }
}

// Hermite4::next_integration_time (Hermite4.cpp:90-102) per cluster:
// min(t_i + dt_i) into the cluster's tmin_key (one workgroup per block).
__global__ __launch_bounds__(BS) void k_next_time(Soa s, Ens e) {
    __shared__ double red[BS / 64];
    for (uint32_t b = blockIdx.x; b < e.nblk; b += gridDim.x) {
        __syncthreads();
        double m = block_tmin(s, b, red);
        if (threadIdx.x == 0) atomicMin(&e.cs[cluster_of_block(e, b)].tmin_key, dkey(m));
    }
}

// Hermite4::init_dt (Hermite4.cpp:108-140) per particle, plus the cluster
// minimum of dt in tmin_key.  Non-positive, infinite or NaN ETA|a|/|j|
// map to dt_min (the reference's x86 outcome of its undefined int cast).
__global__ __launch_bounds__(BS) void k_init_dt(Soa s, Ens e, float eta, double itime,
                                               double dt_min, double dt_max) {
#pragma clang fp contract(off)
    __shared__ double red[BS / 64];
    for (uint32_t b = blockIdx.x; b < e.nblk; b += gridDim.x) {
        const uint32_t c = cluster_of_block(e, b);
        const uint32_t i = b * JT + threadIdx.x;
        double mn = INFINITY;
        if (i - e.off[c] < e.cnt[c]) {
            double a2 = magnitude(s.f[0][i], s.f[1][i], s.f[2][i]);
            double j2 = magnitude(s.f[3][i], s.f[4][i], s.f[5][i]);
            double tmp_dt = eta * (a2 / j2);
            double d;
            if (!(tmp_dt > 0.0) || isinf(tmp_dt)) {
                d = dt_min;
            } else {
                int ex = (int)(ceil(log(tmp_dt) / log(2.0)) - 1);
                d = ldexp(1.0, ex);
                if (d < dt_min)
                    d = dt_min;
                else if (d > dt_max)
                    d = dt_max;
            }
            s.dt[i] = d;
            s.t[i] = itime;
            mn = d;
        }
        __syncthreads();
        mn = block_min(mn, red);
        if (threadIdx.x == 0) atomicMin(&e.cs[c].tmin_key, dkey(mn));
    }
}

// ---------------------------------------------------------------------------
// Step-loop control
// ---------------------------------------------------------------------------

// Arms the device step loop of every cluster.  set0: cluster 0 takes
// (itime0, atime0) first.  A cluster whose ITIME is not < t_end is done
// at once (the reference's loop test, Hermite4CPU_integration.cpp:81).
__global__ __launch_bounds__(BS) void k_run_setup(Ens e, double t_end, double t_stop,
                                                 unsigned long long max_steps, uint32_t set0,
                                                 double itime0, double atime0) {
    __shared__ uint32_t cnt[BS / 64];
    uint32_t nd = 0;
    for (uint32_t c = threadIdx.x; c < e.K; c += BS) {
        ClState &cs = e.cs[c];
        if (set0 && c == 0) {
            cs.itime = itime0;
            cs.atime = atime0;
        }
        cs.steps = 0;
        cs.interactions = 0;
        cs.nsteps = 0;
        cs.small_steps = 0;
        cs.nact = 0;
        cs.tail_pend = 0;
        cs.done = (cs.itime < t_end) ? 0u : 1u;
        nd += cs.done;
    }
    for (int o = 32; o > 0; o >>= 1) nd += __shfl_xor(nd, o);
    if (__lane_id() == 0) cnt[threadIdx.x >> 6] = nd;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < BS / 64; ++k) t += cnt[k];
        RunCtl &ctl = *e.ctl;
        ctl.t_end = t_end;
        ctl.t_stop = t_stop;
        ctl.max_steps = max_steps;
        ctl.ndirty[0] = ctl.ndirty[1] = 0;
        ctl.ndone = t;
        ctl.items = 0;
        ctl.fin_units = 0;
    }
}

// The loop tail a single system's last block step left (k_finish_blocks),
// when no further step of the loop makes it: after one h4h_step.
__global__ __launch_bounds__(BS) void k_tail(Ens e) {
    __shared__ double red[BS / 64];
    ClState *cs = &e.cs[0];
    if (cs->done || !cs->tail_pend) return;
    double mn = INFINITY;
    for (uint32_t b = threadIdx.x; b < e.nblk; b += BS) mn = fmin(mn, e.bmin[b]);
    mn = block_min(mn, red);
    if (threadIdx.x == 0) {
        const unsigned long long steps = cs->tail_steps + 1;
        record_tail(e, 0, cs->tail_t, cs->tail_nact, steps, mn,
                    loop_stops(*e.ctl, cs->tail_t, steps, mn));
        cs->tail_pend = 0;
    }
}

// ---------------------------------------------------------------------------
// Small systems: whole block steps in one workgroup
// ---------------------------------------------------------------------------

constexpr uint32_t SB = 1024;         // threads of k_small_run, and the largest N it takes
constexpr uint32_t SMALL_MAX_ACT = 256;  // largest active set k_small_run takes
static_assert(SB % JT == 0 && SB / JT <= SB / 64, "block minima from the wave minima");
static_assert(H4_SMALL_NARROW_WAVES >= 1 && H4_SMALL_NARROW_WAVES <= SB / 64,
              "pair waves of the one-workgroup step");

// p (a global-memory pointer), as a value the compiler cannot hoist out of
// a loop: stores through it inside the loop address SGPR base + VGPR offset
// instead of holding one 64-bit VGPR address per array for the whole loop.
// Typed as a global pointer so that they are global (not flat) stores: a
// flat store also counts against the LDS wait of the next barrier.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *in_loop(T *p) {
    asm volatile("" : "+s"(p));
    return (__attribute__((address_space(1))) T *)p;
}

// Number of threads of the workgroup with `flag` set, and (for those) the
// count below them in thread order.  Ends with a barrier.
__device__ __forceinline__ uint32_t small_count(bool flag, uint32_t *wcnt, uint32_t &below) {
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const unsigned long long mask = __ballot(flag);
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t total = 0;
    below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    for (uint32_t k = 0; k < SB / 64; ++k) {
        const uint32_t c = wcnt[k];
        below += k < w ? c : 0u;
        total += c;
    }
    return total;
}

// Minimum of v over the workgroup (every thread gets it).  Ends after a
// barrier; wmin is free again after the next one.
__device__ __forceinline__ double small_min(double v, double *wmin) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    if (__lane_id() == 0) wmin[threadIdx.x >> 6] = v;
    __syncthreads();
    double m = wmin[0];
    for (uint32_t k = 1; k < SB / 64; ++k) m = fmin(m, wmin[k]);
    return m;
}

// Block steps of systems of n <= SB particles, workgroup c of SB threads for
// cluster c (a larger cluster's workgroup returns at once), up to max_run
// steps per launch; clusters are independent, so the workgroups never wait
// for each other.  Thread i owns particle i of its cluster and
// keeps its position, velocity, forces, time and step in registers for the
// whole launch; the threads exchange the predictor records and the active
// list through LDS, and the state goes to HBM as it changes (stores
// that nothing in the launch reads back).  Each step is the device loop's
// step: predictor of all particles at the block time T, the particles with
// t + dt = T (Hermite4::find_particles_to_move), their old forces, the pair
// sums, the corrector and new step, the next block time min(t + dt) and the
// loop tail (the record_tail of each step).  The active set of the next step is counted
// before its predictor; when it is empty or larger than max_act the launch
// ends and leaves that step to the multi-workgroup kernels.
//
// Pair sums: P = nact rounded up to a power of two i slots over PT threads
// (all SB for P >= 64, H4_SMALL_NARROW_WAVES waves below), G = PT / P lanes
// per slot (lane g of slot p takes j = g, g + G, ...; P >= 64 puts one j per
// wave, an LDS broadcast), reduced by shuffles inside a wave and in wave
// order through LDS: a fixed order for a given nact.  The mixed path folds
// its FP32 runs into FP64 every 32 pairs, as the main kernel does.  On exit
// the predictor records of the last step and the per-block minima of t + dt
// are stored for the kernels that follow.
template <bool MIXED>
__global__ __launch_bounds__(SB) void k_small_run(Soa s, Ens e, uint32_t max_run,
                                                  uint32_t max_act, double e2, double eta,
                                                  double dt_min, double dt_max,
                                                  uint32_t self_zero) {
#pragma clang fp contract(off)
    __shared__ double4 lpp[SB], lpv[SB];  // predictor records
    __shared__ __align__(16) double red[6 * SB];  // pair sums per (group, slot); corrector output
    __shared__ double stg[7 * SMALL_MAX_ACT];     // old forces and step per active slot
    __shared__ uint32_t lact[SB];          // active list
    __shared__ uint32_t wcnt[SB / 64];
    __shared__ double wmin[SB / 64];

    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    const uint32_t c = blockIdx.x;
    ClState *cs = &e.cs[c];
    const uint32_t n = e.cnt[c];
    if (n > SB) return;
    const uint32_t off = e.off[c];  // first particle of the cluster
    const bool in = tid < n;
    if (max_act > SMALL_MAX_ACT) max_act = SMALL_MAX_ACT;
    if (cs->done) return;
    const double t_end = e.ctl->t_end, t_stop = e.ctl->t_stop;
    const unsigned long long max_steps = e.ctl->max_steps;
    unsigned long long steps = cs->steps;
    uint32_t hk = (uint32_t)(steps % e.hist_cap);  // history slot of the next step
    double T = cs->atime;
    // self_zero (set by the host, self_pair_is_zero): a particle's pair with
    // itself (r = 0, v = 0) has a finite weight m / e2^1.5 for every mass of
    // the context, so it adds exact zeros and the j loop needs no test for it.

    double r[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
    double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double ti = 0.0, dti = 0.0;
    float mf = 0.0f;
    if (in) {
        const uint32_t gi = off + tid;
        for (int k = 0; k < 3; ++k) {
            r[k] = s.r[k][gi];
            v[k] = s.v[k][gi];
        }
        for (int k = 0; k < 6; ++k) f[k] = s.f[k][gi];
        ti = s.t[gi];
        dti = s.dt[gi];
        mf = (float)s.m[gi];
    }
    // The loop tail the cluster's last multi-workgroup step left comes
    // first; its next block time is the minimum of t + dt, which these
    // threads hold.  (The dirty lists are cleared for the steps that follow
    // this launch, whose parities need not continue the last step's.)
    if (cs->tail_pend) {
        const double Tp = cs->tail_t;
        const uint32_t np = cs->tail_nact;
        steps = cs->tail_steps + 1;
        const double mn = small_min(in ? ti + dti : (double)INFINITY, wmin);
        const bool stop = loop_stops(*e.ctl, Tp, steps, mn);
        if (tid == 0) {
            record_tail(e, c, Tp, np, steps, mn, stop);
            cs->tail_pend = 0;
            e.ctl->ndirty[0] = e.ctl->ndirty[1] = 0;
        }
        if (stop) return;  // (the block minima are the finish pass's)
        hk = (uint32_t)(steps % e.hist_cap);
        T = mn;
    }
    bool act = in && fabs(T - (ti + dti)) < 2 * DBL_EPSILON;
    uint32_t k_act;  // this particle's place in the active list
    uint32_t nact = small_count(act, wcnt, k_act);

    uint32_t it = 0;
    for (; it < max_run && nact >= 1 && nact <= max_act; ++it) {
        // Predictor of every particle at T; active list.
        if (in) {
            double4 a, b;
            predict_vals(r, v, f, T - ti, mf, a, b);
            lpp[tid] = a;
            lpv[tid] = b;
        }
        if (act) {
            lact[k_act] = tid;
            for (int k = 0; k < 6; ++k) stg[k_act * 7 + k] = f[k];
            stg[k_act * 7 + 6] = dti;
        }
        __syncthreads();

        // Pair sums of the active particles on PT threads: all SB when
        // P >= 64 (one slot per lane), H4_SMALL_NARROW_WAVES waves when
        // P < 64 (the same FP64 work per SIMD with fewer cross-lane and
        // cross-wave sums).
        const uint32_t lp = nact > 1 ? 32u - (uint32_t)__clz(nact - 1) : 0u;
        const uint32_t P = 1u << lp;
        const uint32_t PT = P < 64 ? 64u * H4_SMALL_NARROW_WAVES : SB;
        const uint32_t G = PT >> lp;
        const uint32_t slot = tid & (P - 1), g = tid >> lp;
        const bool pt = tid < PT;  // (PT is a multiple of 64: wave-uniform)
        const uint32_t i = lact[slot < nact ? slot : nact - 1];
        double pi[6];
        {
            const double4 a = lpp[i], b = lpv[i];
            pi[0] = a.x;
            pi[1] = a.y;
            pi[2] = a.z;
            pi[3] = b.x;
            pi[4] = b.y;
            pi[5] = b.z;
        }
        double fo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (!pt) {
        } else if constexpr (!MIXED) {
            if (self_zero) {
                for (uint32_t j = g; j < n; j += G) {
                    const double4 a = lpp[j], b = lpv[j];
                    pair_fp64(pi, a.x, a.y, a.z, b.x, b.y, b.z, a.w, e2, fo);
                }
            } else {
                for (uint32_t j = g; j < n; j += G) {
                    const double4 a = lpp[j], b = lpv[j];
                    const bool me = j == i;
                    pair_fp64(pi, me ? pi[0] + 1.0 : a.x, a.y, a.z, b.x, b.y, b.z,
                              me ? 0.0 : a.w, e2, fo);
                }
            }
        } else {
            const float e2f = (float)e2;
            for (uint32_t j0 = g; j0 < n; j0 += 32 * G) {
                const uint32_t j1 = min(n, j0 + 32 * G);
                float t6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                if (self_zero) {
                    for (uint32_t j = j0; j < j1; j += G) {
                        const double4 a = lpp[j], b = lpv[j];
                        pair_mixed(pi, a.x, a.y, a.z, b.x, b.y, b.z, (float)a.w, e2f, t6);
                    }
                } else {
                    for (uint32_t j = j0; j < j1; j += G) {
                        const double4 a = lpp[j], b = lpv[j];
                        const bool me = j == i;
                        pair_mixed(pi, me ? pi[0] + 1.0 : a.x, a.y, a.z, b.x, b.y, b.z,
                                   me ? 0.f : (float)a.w, e2f, t6);
                    }
                }
                for (int k = 0; k < 6; ++k) fo[k] += (double)t6[k];
            }
        }
        if (P < 64 && pt)
            for (uint32_t o = P; o < 64; o <<= 1)
                for (int k = 0; k < 6; ++k) fo[k] += __shfl_xor(fo[k], (int)o);
        // P <= 64: wave w holds the sums of slots 0..P-1 in lanes 0..P-1
        // (group w); P > 64: every lane holds one slot, group g.
        const uint32_t ngrp = P <= 64 ? PT / 64 : G;
        if (pt && (P > 64 || lane < P)) {
            const uint32_t grp = P <= 64 ? w : g;
            for (int k = 0; k < 6; ++k) red[(grp * P + slot) * 6 + k] = fo[k];
        }
        __syncthreads();

        // Corrector of active slot tid (threads 0 .. nact-1, the first waves:
        // each wave that holds an active slot runs the whole corrector).
        double fn[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double rn[3], vn[3], dtn = 0.0;
        if (tid < nact) {
            // groups in order; ngrp <= SB / 64 (unrolled: the loads issue together)
#pragma unroll
            for (uint32_t q = 0; q < SB / 64; ++q) {
                const bool use = q < ngrp;
                const double2 *rp =
                    reinterpret_cast<const double2 *>(red + ((use ? q : 0u) * P + tid) * 6);
                const double2 x0 = rp[0], x1 = rp[1], x2 = rp[2];
                fn[0] = use ? fn[0] + x0.x : fn[0];
                fn[1] = use ? fn[1] + x0.y : fn[1];
                fn[2] = use ? fn[2] + x1.x : fn[2];
                fn[3] = use ? fn[3] + x1.y : fn[3];
                fn[4] = use ? fn[4] + x2.x : fn[4];
                fn[5] = use ? fn[5] + x2.y : fn[5];
            }
            if (!MIXED)
                for (int k = 0; k < 6; ++k) fn[k] *= PAIR_SCALE;
            const uint32_t ia = lact[tid];  // within the cluster
            const uint32_t gi = off + ia;
            double o[6], pr[6], a2[3], a3[3];
            for (int k = 0; k < 6; ++k) o[k] = stg[tid * 7 + k];
            {
                const double4 a = lpp[ia], b = lpv[ia];
                pr[0] = a.x;
                pr[1] = a.y;
                pr[2] = a.z;
                pr[3] = b.x;
                pr[4] = b.y;
                pr[5] = b.z;
            }
            dtn = correct_vals(o, fn, pr, stg[tid * 7 + 6], T, eta, dt_min, dt_max, rn, vn, a2,
                               a3);
            for (int k = 0; k < 6; ++k) {
                in_loop(s.old[k])[gi] = o[k];
                in_loop(s.f[k])[gi] = fn[k];
            }
            for (int k = 0; k < 3; ++k) {
                in_loop(s.r[k])[gi] = rn[k];
                in_loop(s.v[k])[gi] = vn[k];
                in_loop(s.a2[k])[gi] = a2[k];
                in_loop(s.a3[k])[gi] = a3[k];
            }
            in_loop(s.t)[gi] = T;
            in_loop(s.dt)[gi] = dtn;
        }
        __syncthreads();  // red is read: the corrected state goes there for the owners
        if (tid < nact) {
            double *out = red + tid * 13;
            for (int k = 0; k < 6; ++k) out[k] = fn[k];
            for (int k = 0; k < 3; ++k) {
                out[6 + k] = rn[k];
                out[9 + k] = vn[k];
            }
            out[12] = dtn;
        }
        __syncthreads();
        if (act) {
            const double *in_ = red + k_act * 13;
            for (int k = 0; k < 6; ++k) f[k] = in_[k];
            for (int k = 0; k < 3; ++k) {
                r[k] = in_[6 + k];
                v[k] = in_[9 + k];
            }
            ti = T;
            dti = in_[12];
        }

        // Next block time (Hermite4::next_integration_time) and loop tail.
        const double mn = small_min(in ? ti + dti : (double)INFINITY, wmin);
        steps += 1;
        const bool stop = T >= t_stop || steps >= max_steps || !(T < t_end) || !(mn > T);
        if (tid == 0) {
            if (c == 0) {
                e.hist_t[hk] = T;
                e.hist_n[hk] = nact;
            }
            cs->itime = T;
            cs->atime = mn;
            cs->steps = steps;
            cs->interactions += (unsigned long long)nact * n;
            cs->nsteps += nact;
            cs->small_steps += 1;
            if (stop) {
                cs->done = 1;
                atomicAdd(&e.ctl->ndone, 1u);
            }
        }
        T = mn;
        if (++hk == e.hist_cap) hk = 0;
        if (stop) {
            ++it;
            break;
        }
        act = in && fabs(T - (ti + dti)) < 2 * DBL_EPSILON;
        nact = small_count(act, wcnt, k_act);
    }

    // Predictor records of the last step made; per-block minima of t + dt.
    if (it > 0 && in) {
        s.pp[off + tid] = lpp[tid];
        s.pv[off + tid] = lpv[tid];
    }
    __syncthreads();  // wmin is rewritten below
    double vb = in ? ti + dti : (double)INFINITY;
    for (int o = 32; o > 0; o >>= 1) vb = fmin(vb, __shfl_xor(vb, o));
    if (lane == 0) wmin[w] = vb;
    __syncthreads();
    constexpr uint32_t WPB = JT / 64;  // waves per block
    if (tid < ntiles(e, c)) {
        double m = wmin[tid * WPB];
        for (uint32_t k = 1; k < WPB; ++k) m = fmin(m, wmin[tid * WPB + k]);
        e.bmin[off / JT + tid] = m;
    }
}

// Single-cluster ABI helpers.
__global__ void k_set_nact0(Ens e, uint32_t nact) { e.cs[0].nact = nact; }
__global__ void k_reset_tmin(Ens e) {
    for (uint32_t c = 0; c < e.K; ++c) e.cs[c].tmin_key = dkey(INFINITY);
}
__global__ void k_reset_select(Ens e) {
    for (uint32_t c = 0; c < e.K; ++c) e.cs[c].nact = 0;
}
// After init_dt: every cluster at t = itime with ATIME = its minimum dt.
__global__ void k_set_times_from_tmin(Ens e, double itime) {
    for (uint32_t c = 0; c < e.K; ++c) {
        e.cs[c].itime = itime;
        e.cs[c].atime = dkey_inv(e.cs[c].tmin_key);
    }
}

// ---------------------------------------------------------------------------
// Energy: sum 1/2 m v^2 - sum_{i<j} m_i m_j / |r_ij| (unsoftened) of one
// cluster, NbodyUtils::get_energy (NbodyUtils.cpp:427-529).  1/|r_ij| from
// v_rsq_f64 and one Newton step (relative error below 2^-47 and of one
// sign, against the reference's division by sqrt): a lane sums
// m_j / |r_ij| over one j tile in plain FP64 (at most 256 terms of one
// sign, so at most 256 roundings of the partial sum: 3e-14 relative),
// multiplies by m_i and adds that into its Neumaier-compensated sum; then
// per workgroup, then over workgroups in index order (deterministic for a
// fixed grid).  A coincident pair (r = 0) makes the Newton step NaN; such
// a tile is summed again with 1/0 = inf, so coincident particles give -inf
// as in the reference.
// ---------------------------------------------------------------------------

// Neumaier-compensated sum += x.  Once the sum is infinite (coincident
// particles in the unsoftened energy) the compensation is left alone, so
// the result is the infinity a plain sum gives rather than inf - inf.
__device__ __forceinline__ void neumaier(double &sum, double &comp, double x) {
#pragma clang fp contract(off)
    double t = sum + x;
    if (isfinite(t)) {
        if (fabs(sum) >= fabs(x))
            comp += (sum - t) + x;
        else
            comp += (x - t) + sum;
    }
    sum = t;
}

__device__ __forceinline__ void block_sum2(double &a, double &b, double *red) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    uint32_t w = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) {
        red[2 * w] = a;
        red[2 * w + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < BS / 64; ++k) {
            a += red[2 * k];
            b += red[2 * k + 1];
        }
    }
}

// Items are pairs of (i tile, j tile) with j tile >= i tile of the cluster
// at device offset off, numbered row by row;
// out[4*block] = {epot, epot_comp, ekin, ekin_comp}.
// Cluster c0 + blockIdx.y of the cluster tables offs / cnts; workgroup
// blockIdx.x takes the tile pairs blockIdx.x, + stride, + 2 stride, ...
// (stride >= gridDim.x: a launch with fewer workgroups than stride has the
// same per-workgroup sums as one with stride workgroups, the others' being
// zero), and writes its four sums at out[4 (blockIdx.y gridDim.x +
// blockIdx.x)].
__global__ __launch_bounds__(BS) void k_energy(Soa s, const uint32_t *offs, const uint32_t *cnts,
                                              uint32_t c0, uint32_t stride, double *out) {
#pragma clang fp contract(off)
    __shared__ double4 sr[JT];
    __shared__ double red[2 * (BS / 64)];
    const uint32_t off = offs[c0 + blockIdx.y], n = cnts[c0 + blockIdx.y];
    const uint32_t nt = (n + BS - 1) / BS;
    const uint64_t items = (uint64_t)nt * (nt + 1) / 2;
    double ep = 0.0, epc = 0.0, ek = 0.0, ekc = 0.0;
    // Row `it` holds nt - it items.  (it, rem) locate the current item and
    // advance by stride per iteration, so the walk is O(nt) in total.
    uint32_t it = 0;
    uint64_t rem = blockIdx.x;
    for (uint64_t item = blockIdx.x; item < items; item += stride) {
        while (rem >= (uint64_t)(nt - it)) {
            rem -= (nt - it);
            ++it;
        }
        uint32_t jt = it + (uint32_t)rem;
        rem += stride;
        uint32_t li = it * BS + threadIdx.x;
        uint32_t lj0 = jt * JT;
        __syncthreads();
        {
            uint32_t j = off + lj0 + threadIdx.x;
            if (lj0 + threadIdx.x < n)
                sr[threadIdx.x] = make_double4(s.r[0][j], s.r[1][j], s.r[2][j], s.m[j]);
        }
        __syncthreads();
        if (li < n) {
            uint32_t i = off + li;
            double xi = s.r[0][i], yi = s.r[1][i], zi = s.r[2][i], mi = s.m[i];
            uint32_t jn = min((uint32_t)JT, n - lj0);
            uint32_t jb = (jt == it) ? threadIdx.x + 1 : 0;
            auto tile_sum = [&](bool guard) {
                double sum = 0.0;
                for (uint32_t jj = jb; jj < jn; ++jj) {
                    double4 q = sr[jj];
                    double rx = q.x - xi;
                    double ry = q.y - yi;
                    double rz = q.z - zi;
                    double r2 = __builtin_fma(rx, rx, __builtin_fma(ry, ry, rz * rz));
                    double y = __builtin_amdgcn_rsq(r2);
                    double e = __builtin_fma(-r2, y * y, 1.0);
                    double rinv = __builtin_fma(y * e, 0.5, y);
                    if (guard && !(r2 > 0.0)) rinv = INFINITY;
                    sum = __builtin_fma(q.w, rinv, sum);
                }
                return sum;
            };
            double tile = tile_sum(false);
            if (isnan(tile)) tile = tile_sum(true);
            if (jb < jn) neumaier(ep, epc, -(mi * tile));
            if (jt == it) {
                double vx = s.v[0][i], vy = s.v[1][i], vz = s.v[2][i];
                double vv2 = vx * vx + vy * vy + vz * vz;
                neumaier(ek, ekc, 0.5 * mi * vv2);
            }
        }
    }
    block_sum2(ep, epc, red);
    __syncthreads();
    block_sum2(ek, ekc, red);
    if (threadIdx.x == 0) {
        double *o = out + 4 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
        o[0] = ep;
        o[1] = epc;
        o[2] = ek;
        o[3] = ekc;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

namespace {
constexpr uint32_t RUN_CHUNK = 32;    // block steps per captured graph
constexpr uint32_t SMALL_RUN = 255;   // k_small_run steps per launch
constexpr uint32_t SMALL_SLOTS = 8;   // (k_small_run + block step) pairs per captured graph
static_assert(RUN_CHUNK % 2 == 0 && SMALL_SLOTS % 2 == 0,
              "a step's dirty list is its slot in the chunk mod 2; chunks follow each other");
constexpr uint32_t SMALL_ACT = 64;    // default largest active set k_small_run takes (one system)
constexpr uint32_t HIST_CAP = 2048;   // per-step history ring of cluster 0
static_assert(HIST_CAP >= RUN_CHUNK && HIST_CAP >= SMALL_SLOTS * (SMALL_RUN + 1),
              "history ring shorter than the steps of one graph");
constexpr int N_SOA = 35;           // doubles per particle in `state`
}  // namespace

struct h4h_ctx {
    uint32_t K = 0;       // clusters
    uint32_t n = 0;       // particles (all clusters)
    uint32_t n_dev = 0;   // padded device particles
    uint32_t nblk = 0;    // JT blocks
    std::vector<uint32_t> off, cnt, hoff;
    double e2 = 0.0;
    float eta = 0.0f;
    double dt_min = 0.0, dt_max = 0.0;
    int precision = H4H_FP64;
    int device = 0;
    uint32_t grid = DEFAULT_GRID;      // requested persistent workgroups
    uint32_t max_split = DEFAULT_MAX_SPLIT;
    uint32_t iw_fill = DEFAULT_IW_FILL;  // pick_iw fill rule (sixteenths)
    bool uploaded = false;
    bool timing = false;
    bool bm_valid = false;   // per-block minima match t + dt
    bool small_ok = false;   // the device can launch k_small_run
    uint32_t small_act = SMALL_ACT;  // k_small_run takes active sets up to this (0: off)
    uint32_t nact_host = 0;  // size of cluster 0's active list (single-system calls)
    double mass_max = 0.0;   // largest |float mass| of the uploaded state (+inf: not finite)

    hipStream_t stream = nullptr;
    double *state = nullptr;     // all SoA arrays
    double *staging = nullptr;   // n * 8 doubles
    double *partial = nullptr;   // j-split partials
    size_t partial_len = 0;
    double *eout = nullptr;      // energy per-workgroup sums
    uint32_t *move = nullptr;
    // cluster tables
    uint32_t *d_off = nullptr, *d_cnt = nullptr, *d_hoff = nullptr, *d_bcl = nullptr;
    ClState *d_cs = nullptr;
    ClPlan *d_plan = nullptr;
    unsigned long long *d_pre = nullptr;  // items_pre | fin_pre | part_pre
    RunCtl *d_ctl = nullptr;
    uint32_t *d_dirty = nullptr;
    double *d_bmin = nullptr;
    double4 *d_arec = nullptr;   // the step's active particles' records (Ens::arec)
    uint32_t *d_tune = nullptr;  // [64] tuning table
    double *hist_t = nullptr;
    uint32_t *hist_n = nullptr;
    // pinned host mirrors
    RunCtl *ctl_host = nullptr;
    ClState *cs_host = nullptr;
    double *hist_t_host = nullptr;
    uint32_t *hist_n_host = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool graphs = true;          // device loops replay captured chunks
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    Soa s{};
    Ens e{};
    h4h_stats stats{};
};

namespace {

int check_ctx(h4h_ctx *ctx, bool need_upload = true) {
    if (!ctx) return fail(H4H_EINVAL, "null context");
    if (need_upload && !ctx->uploaded)
        return fail(H4H_ESTATE, "no state uploaded (call h4h_upload first)");
    return H4H_OK;
}

int check_single(h4h_ctx *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->K != 1)
        return fail(H4H_EINVAL, "this call acts on one system; use the h4h_ensemble_* calls");
    return H4H_OK;
}

uint32_t blocks_for(size_t n) { return (uint32_t)((n + BS - 1) / BS); }

// Largest number of update work items any active-set size can produce for
// one cluster of n particles (the persistent grid of a single-system
// context is capped by it, so small systems do not launch idle
// workgroups).
uint32_t max_items(uint32_t n, uint32_t grid, uint32_t max_split, uint32_t fill16) {
    uint32_t n_jt = (n + JT - 1) / JT;
    uint32_t best = 1;
    for (uint32_t nact = 1;; nact = (nact < 64 ? nact + 1 : nact * 2)) {
        uint32_t k = nact > n ? n : nact;
        Decomp d = decompose(k, pick_iw(k, fill16), n_jt, grid, max_split);
        best = std::max(best, d.n_items);
        if (nact >= n) break;
    }
    return best;
}

void drop_graph(h4h_ctx *ctx) {
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
}

// (Re)sizes the j-split partials for the current grid and updates the
// launch fields of the cluster table.
int configure_launch(h4h_ctx *ctx) {
    uint32_t g = ctx->grid;
    if (ctx->K == 1)
        g = std::min(g, max_items(ctx->cnt[0], ctx->grid, ctx->max_split, ctx->iw_fill));
    ctx->e.grid = g;
    // queue target of the plans: an upper bound of their work items
    const uint32_t target = ctx->grid * QUEUE_FACTOR;
    ctx->e.max_split = ctx->max_split;
    ctx->e.iw_fill = ctx->iw_fill;
    drop_graph(ctx);
    // j-split partials for any plan of any step (partial_bound).
    size_t need;
    if (ctx->K == 1) {
        need = partial_bound(ctx->cnt[0], target);
    } else {
        unsigned long long n_total = 0, per_cluster = 0;
        for (uint32_t c = 0; c < ctx->K; ++c) {
            n_total += ctx->cnt[c];
            per_cluster += partial_bound(ctx->cnt[c], target);
        }
        need = std::min(per_cluster, ensemble_partial_bound(ctx->K, n_total, target));
    }
    if (ctx->partial && ctx->partial_len >= need) return H4H_OK;
    if (ctx->partial) H4H_HIP(hipFree(ctx->partial));
    ctx->partial = nullptr;
    H4H_HIP(hipMalloc(&ctx->partial, need * sizeof(double)));
    ctx->partial_len = need;
    return H4H_OK;
}

uint32_t finish_grid(h4h_ctx *ctx) { return std::min<uint32_t>(1024, std::max<uint32_t>(64, ctx->nblk)); }

// Pair kernel of the current precision over the plan's items.  compact:
// the step's selection wrote the active particles' records.
int launch_update(h4h_ctx *ctx, bool compact = false) {
    if (ctx->precision == H4H_FP64_EXACT) {
        hipLaunchKernelGGL(k_update_exact, dim3(ctx->e.grid), dim3(BS), 0, ctx->stream, ctx->s,
                           ctx->e, ctx->move, ctx->e2);
    } else {
        auto *kern = ctx->precision == H4H_MIXED ? k_update_mixed : k_update_fp64;
        hipLaunchKernelGGL(kern, dim3(ctx->e.grid), dim3(BS), 0, ctx->stream, ctx->s, ctx->e,
                           ctx->move, ctx->partial, ctx->e2, compact ? 1u : 0u);
    }
    H4H_LAUNCHED();
    return H4H_OK;
}

int launch_finish(h4h_ctx *ctx, uint32_t mode) {
    hipLaunchKernelGGL(k_finish, dim3(finish_grid(ctx)), dim3(BS), 0, ctx->stream, ctx->s, ctx->e,
                       ctx->move, ctx->partial, mode, ctx->eta, ctx->dt_min, ctx->dt_max);
    H4H_LAUNCHED();
    return H4H_OK;
}

int launch_select(h4h_ctx *ctx, uint32_t flags, double itime, uint32_t par = 0) {
    hipLaunchKernelGGL(k_select, dim3(ctx->nblk), dim3(BS), 0, ctx->stream, ctx->s, ctx->e,
                       flags, itime, par, ctx->move);
    H4H_LAUNCHED();
    return H4H_OK;
}

int launch_plan(h4h_ctx *ctx) {
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(BS), 0, ctx->stream, ctx->e,
                       ctx->precision == H4H_FP64_EXACT ? 1u : 0u);
    H4H_LAUNCHED();
    return H4H_OK;
}

// Whether the device can launch k_small_run (SB threads, 130 KiB of LDS).
bool small_fits(int device) {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return false;
    for (const void *k : {reinterpret_cast<const void *>(&k_small_run<false>),
                          reinterpret_cast<const void *>(&k_small_run<true>)}) {
        hipFuncAttributes fa{};
        if (hipFuncGetAttributes(&fa, k) != hipSuccess) return false;
        if (fa.sharedSizeBytes > (size_t)lds || fa.maxThreadsPerBlock < (int)SB) return false;
    }
    return true;
}

// Device loops of this context start each step slot with k_small_run: every
// cluster has at most SB particles.  (With larger clusters beside them, the
// small clusters' runs of up to SMALL_RUN steps would hold up the larger
// clusters' steps, which wait for the launch to end.)
bool small_path(const h4h_ctx *ctx) {
    return ctx->small_ok && ctx->small_act > 0 && ctx->precision != H4H_FP64_EXACT &&
           *std::max_element(ctx->cnt.begin(), ctx->cnt.end()) <= SB;
}

// Whether the pair loop of `precision` may add a particle's pair with itself
// (r = 0, v = 0) untested: its weight 3 sqrt(3) m / r^3 at r^2 = e2 (FP64;
// m / r^3 in FP32 for the mixed kernel) must be finite for the largest
// float mass mmax of the context, and e2 inside the range where the
// hardware rsqrt (refined in FP64) keeps its accuracy.  Then the pair's terms are
// finite times zero, exact zeros.  mmax is +inf when a mass is not finite.
bool self_pair_is_zero(int precision, double e2, double mmax) {
    if (precision == H4H_MIXED) {
        const float e2f = (float)e2;
        return e2f >= 0x1p-60f && mmax * std::pow((double)e2f, -1.5) < 0x1p120;
    }
    return e2 >= 0x1p-600 && mmax * std::pow(e2, -1.5) < 0x1p1016;
}

// One workgroup per cluster.
int launch_small(h4h_ctx *ctx, uint32_t max_run) {
    auto *kern = ctx->precision == H4H_MIXED ? k_small_run<true> : k_small_run<false>;
    hipLaunchKernelGGL(kern, dim3(ctx->K), dim3(SB), 0, ctx->stream, ctx->s, ctx->e, max_run,
                       ctx->small_act, ctx->e2, (double)ctx->eta, ctx->dt_min, ctx->dt_max,
                       self_pair_is_zero(ctx->precision, ctx->e2, ctx->mass_max) ? 1u : 0u);
    H4H_LAUNCHED();
    return H4H_OK;
}

int ensure_blockmin(h4h_ctx *ctx) {
    if (ctx->bm_valid) return H4H_OK;
    hipLaunchKernelGGL(k_blockmin_all, dim3(std::min<uint32_t>(ctx->nblk, 1024)), dim3(BS), 0,
                       ctx->stream, ctx->s, ctx->e);
    H4H_LAUNCHED();
    ctx->bm_valid = true;
    return H4H_OK;
}

// One block step of every live cluster, device-driven: selection with the
// fused predictor and save-old (after the loop tail of each cluster's last
// step), work plan, update, finish (chunk sums + corrector) and the
// per-block minima of the touched blocks; each cluster's loop tail is made
// by the next launch that needs its next block time (the next step's
// selection, k_small_run, or k_tail after h4h_step), so in a device loop
// that ends, the step after the last one finds the loop stopped.  One
// system: three launches (the minima in the finish pass); an ensemble: five.
// par: the step's slot in its captured chunk, mod 2 (its dirty list).
// Optional HIP events bracket the update kernel.
int enqueue_step(h4h_ctx *ctx, bool events, uint32_t par) {
    int rc;
    if ((rc = launch_select(ctx,
                            PS_PREDICT | PS_SELECT | PS_SAVE_OLD | PS_DEVTIME | PS_DIRTY | PS_TAIL |
                                PS_COMPACT,
                            0.0, par)))
        return rc;
    // (a single-system context's update kernel makes its own plan)
    if (ctx->K > 1 && (rc = launch_plan(ctx))) return rc;
    if (events) H4H_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    if ((rc = launch_update(ctx, true))) return rc;
    if (events) H4H_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    if (ctx->K == 1) {
        hipLaunchKernelGGL(k_finish_blocks, dim3(std::min<uint32_t>(ctx->nblk, 1024)), dim3(BS), 0,
                           ctx->stream, ctx->s, ctx->e, ctx->move, ctx->partial, ctx->eta,
                           ctx->dt_min, ctx->dt_max, par);
        H4H_LAUNCHED();
    } else {
        if ((rc = launch_finish(ctx, FIN_CORRECT))) return rc;
        hipLaunchKernelGGL(k_blockmin_dirty, dim3(std::min<uint32_t>(ctx->nblk, 1024)), dim3(BS), 0,
                           ctx->stream, ctx->s, ctx->e, par);
        H4H_LAUNCHED();
    }
    return H4H_OK;
}

// One chunk of the device loop: RUN_CHUNK block steps, or for a small
// system SMALL_SLOTS slots of k_small_run (up to SMALL_RUN steps with small
// active sets) followed by one block step of the multi-workgroup kernels
// (the step k_small_run left, if any).  At most HIST_CAP steps.
int enqueue_chunk(h4h_ctx *ctx) {
    int rc = H4H_OK;
    if (small_path(ctx)) {
        for (uint32_t k = 0; k < SMALL_SLOTS && rc == H4H_OK; ++k)
            if ((rc = launch_small(ctx, SMALL_RUN)) == H4H_OK) rc = enqueue_step(ctx, false, k & 1);
    } else {
        for (uint32_t k = 0; k < RUN_CHUNK && rc == H4H_OK; ++k) rc = enqueue_step(ctx, false, k & 1);
    }
    return rc;
}

// Captures one chunk into a graph (once per configuration).  The graph
// holds kernel nodes only.
int ensure_graph(h4h_ctx *ctx) {
    if (ctx->graph_exec) return H4H_OK;
    H4H_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_chunk(ctx);
    hipGraph_t g = nullptr;
    hipError_t err = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (err != hipSuccess)
        return fail(H4H_EHIP, std::string("graph capture: ") + hipGetErrorString(err));
    ctx->graph = g;
    H4H_HIP(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
    return H4H_OK;
}

int read_state(h4h_ctx *ctx) {
    H4H_HIP(hipMemcpyAsync(ctx->ctl_host, ctx->d_ctl, sizeof(RunCtl), hipMemcpyDeviceToHost,
                           ctx->stream));
    H4H_HIP(hipMemcpyAsync(ctx->cs_host, ctx->d_cs, ctx->K * sizeof(ClState),
                           hipMemcpyDeviceToHost, ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

// Update pass of the single-method calls (plan already on the device).
int update_pass(h4h_ctx *ctx, uint32_t fin_mode) {
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    int rc = launch_update(ctx);
    if (rc) return rc;
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    if ((rc = launch_finish(ctx, fin_mode))) return rc;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->timing) {
        float ms = 0.f;
        H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->stats.update_ms += ms;
        ctx->stats.update_launches += 1;
    }
    return H4H_OK;
}

int upload_cols(h4h_ctx *ctx, const double *src, uint32_t ncols, Cols dst) {
    H4H_HIP(hipMemcpyAsync(ctx->staging, src, (size_t)ctx->n * ncols * sizeof(double),
                           hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(blocks_for(ctx->n_dev)), dim3(BS), 0, ctx->stream,
                       ctx->staging, ncols, dst, ctx->e);
    H4H_LAUNCHED();
    return H4H_OK;
}

int download_cols(h4h_ctx *ctx, double *dst, uint32_t ncols, Cols src) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(blocks_for(ctx->n_dev)), dim3(BS), 0, ctx->stream, src,
                       ncols, ctx->staging, ctx->e);
    H4H_LAUNCHED();
    H4H_HIP(hipMemcpyAsync(dst, ctx->staging, (size_t)ctx->n * ncols * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int upload_scalar(h4h_ctx *ctx, const double *src, double *dst) {
    H4H_HIP(hipMemcpyAsync(ctx->staging, src, (size_t)ctx->n * sizeof(double),
                           hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_scalar_to_soa, dim3(blocks_for(ctx->n_dev)), dim3(BS), 0, ctx->stream,
                       ctx->staging, dst, ctx->e);
    H4H_LAUNCHED();
    return H4H_OK;
}

int download_scalar(h4h_ctx *ctx, const double *src, double *dst) {
    hipLaunchKernelGGL(k_soa_to_scalar, dim3(blocks_for(ctx->n_dev)), dim3(BS), 0, ctx->stream,
                       src, ctx->staging, ctx->e);
    H4H_LAUNCHED();
    H4H_HIP(hipMemcpyAsync(dst, ctx->staging, (size_t)ctx->n * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

bool valid_precision(int p) { return p == H4H_FP64 || p == H4H_FP64_EXACT || p == H4H_MIXED; }

// Energies of clusters c0 .. c0 + count - 1, in launches of as many
// clusters as fit the ENERGY_GRID x 4 sums of eout: per cluster the
// workgroups that have tile pairs (at most ENERGY_GRID, item stride
// ENERGY_GRID as in a launch of ENERGY_GRID workgroups for that cluster
// alone), then a host Neumaier sum of its workgroup sums in order.
int energies_of(h4h_ctx *ctx, uint32_t c0, uint32_t count, double *ekin, double *epot) {
    std::vector<double> h(ENERGY_GRID * 4);
    // Neumaier sum of the per-workgroup (sum, compensation) pairs (an
    // infinite sum leaves the compensation alone, as on the device).
    auto add = [](double &s, double &c2, double x) {
        double t = s + x;
        if (std::isfinite(t)) {
            if (std::fabs(s) >= std::fabs(x))
                c2 += (s - t) + x;
            else
                c2 += (x - t) + s;
        }
        s = t;
    };
    auto wgs = [&](uint32_t c) {
        const uint64_t nt = (ctx->cnt[c] + BS - 1) / BS;
        return (uint32_t)std::min<uint64_t>(ENERGY_GRID, nt * (nt + 1) / 2);
    };
    for (uint32_t b0 = c0; b0 < c0 + count;) {
        uint32_t gx = wgs(b0), kb = 1;
        while (b0 + kb < c0 + count) {
            const uint32_t g2 = std::max(gx, wgs(b0 + kb));
            if ((uint64_t)(kb + 1) * g2 > ENERGY_GRID) break;
            gx = g2;
            ++kb;
        }
        hipLaunchKernelGGL(k_energy, dim3(gx, kb), dim3(BS), 0, ctx->stream, ctx->s, ctx->d_off,
                           ctx->d_cnt, b0, (uint32_t)ENERGY_GRID, ctx->eout);
        H4H_LAUNCHED();
        H4H_HIP(hipMemcpyAsync(h.data(), ctx->eout, (size_t)kb * gx * 4 * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
        H4H_HIP(hipStreamSynchronize(ctx->stream));
        for (uint32_t k = 0; k < kb; ++k) {
            double sp = 0.0, cp = 0.0, sk = 0.0, ck = 0.0;
            const double *hk = h.data() + (size_t)k * gx * 4;
            for (uint32_t b = 0; b < gx; ++b) {
                add(sp, cp, hk[4 * b + 0]);
                add(sp, cp, hk[4 * b + 1]);
                add(sk, ck, hk[4 * b + 2]);
                add(sk, ck, hk[4 * b + 3]);
            }
            if (ekin) ekin[b0 - c0 + k] = sk + ck;
            if (epot) epot[b0 - c0 + k] = sp + cp;
        }
        b0 += kb;
    }
    return H4H_OK;
}

int create_impl(h4h_ctx **out, uint32_t K, const uint32_t *sizes, double e2, float eta,
                double dt_min, double dt_max, int precision) {
    if (!out) return fail(H4H_EINVAL, "out is null");
    *out = nullptr;
    if (K < 1 || !sizes) return fail(H4H_EINVAL, "need at least one cluster");
    if (K > (1u << 20)) return fail(H4H_EINVAL, "too many clusters");
    uint64_t total = 0, total_dev = 0;
    for (uint32_t c = 0; c < K; ++c) {
        if (sizes[c] < 1) return fail(H4H_EINVAL, "every cluster needs n >= 1");
        total += sizes[c];
        total_dev += (sizes[c] + JT - 1) / JT * JT;
    }
    if (total_dev > (1u << 30)) return fail(H4H_EINVAL, "n too large");
    if (!(e2 >= 0.0) || !std::isfinite(e2)) return fail(H4H_EINVAL, "e2 must be finite and >= 0");
    if (!(eta > 0.0f) || !std::isfinite(eta)) return fail(H4H_EINVAL, "eta must be > 0");
    if (!(dt_min > 0.0) || !(dt_max >= dt_min) || !std::isfinite(dt_max))
        return fail(H4H_EINVAL, "need 0 < dt_min <= dt_max");
    if (!valid_precision(precision)) return fail(H4H_EINVAL, "unknown precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(H4H_ENODEV, "no HIP device available");

    h4h_ctx *ctx = new (std::nothrow) h4h_ctx();
    if (!ctx) return fail(H4H_ENOMEM, "host allocation failed");
    ctx->K = K;
    ctx->n = (uint32_t)total;
    ctx->n_dev = (uint32_t)total_dev;
    ctx->nblk = ctx->n_dev / JT;
    ctx->e2 = e2;
    ctx->eta = eta;
    ctx->dt_min = dt_min;
    ctx->dt_max = dt_max;
    ctx->precision = precision;
    ctx->off.resize(K);
    ctx->cnt.resize(K);
    ctx->hoff.resize(K);
    std::vector<uint32_t> bcl(ctx->nblk);
    uint32_t o = 0, h = 0;
    for (uint32_t c = 0; c < K; ++c) {
        ctx->off[c] = o;
        ctx->cnt[c] = sizes[c];
        ctx->hoff[c] = h;
        uint32_t nb = (sizes[c] + JT - 1) / JT;
        for (uint32_t b = 0; b < nb; ++b) bcl[o / JT + b] = c;
        o += nb * JT;
        h += sizes[c];
    }

    auto cleanup = [&](int rc) {
        std::string keep = g_err;
        h4h_destroy(ctx);
        g_err = keep;
        return rc;
    };
    if (hipGetDevice(&ctx->device) != hipSuccess) return cleanup(fail(H4H_EHIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(H4H_EHIP, "hipStreamCreate failed"));
    size_t np = ctx->n_dev;
    bool ok = hipMalloc(&ctx->state, np * N_SOA * sizeof(double)) == hipSuccess &&
              hipMalloc(&ctx->staging, (size_t)ctx->n * 8 * sizeof(double)) == hipSuccess &&
              hipMalloc(&ctx->eout, ENERGY_GRID * 4 * sizeof(double)) == hipSuccess &&
              hipMalloc(&ctx->move, np * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&ctx->d_off, K * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&ctx->d_cnt, K * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&ctx->d_hoff, K * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&ctx->d_bcl, ctx->nblk * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&ctx->d_cs, K * sizeof(ClState)) == hipSuccess &&
              hipMalloc(&ctx->d_plan, K * sizeof(ClPlan)) == hipSuccess &&
              hipMalloc(&ctx->d_pre, 3 * (size_t)(K + 1) * sizeof(unsigned long long)) == hipSuccess &&
              hipMalloc(&ctx->d_ctl, sizeof(RunCtl)) == hipSuccess &&
              hipMalloc(&ctx->d_dirty, 6 * (size_t)ctx->nblk * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&ctx->d_bmin, ctx->nblk * sizeof(double)) == hipSuccess &&
              hipMalloc(&ctx->d_arec, 4 * np * sizeof(double4)) == hipSuccess &&
              hipMalloc(&ctx->d_tune, 64 * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&ctx->hist_t, HIST_CAP * sizeof(double)) == hipSuccess &&
              hipMalloc(&ctx->hist_n, HIST_CAP * sizeof(uint32_t)) == hipSuccess;
    if (!ok) return cleanup(fail(H4H_ENOMEM, "device allocation failed"));
    ok = hipHostMalloc(&ctx->ctl_host, sizeof(RunCtl), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&ctx->cs_host, K * sizeof(ClState), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&ctx->hist_t_host, HIST_CAP * sizeof(double), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc(&ctx->hist_n_host, HIST_CAP * sizeof(uint32_t), hipHostMallocDefault) == hipSuccess;
    if (!ok) return cleanup(fail(H4H_ENOMEM, "pinned allocation failed"));
    for (int k = 0; k < 4; ++k)
        if (hipEventCreate(&ctx->ev[k]) != hipSuccess)
            return cleanup(fail(H4H_EHIP, "hipEventCreate failed"));

    Ens &e = ctx->e;
    e.K = K;
    e.nblk = ctx->nblk;
    e.off = ctx->d_off;
    e.cnt = ctx->d_cnt;
    e.hoff = ctx->d_hoff;
    e.bcl = ctx->d_bcl;
    e.cs = ctx->d_cs;
    e.plan = ctx->d_plan;
    e.items_pre = ctx->d_pre;
    e.fin_pre = ctx->d_pre + (K + 1);
    e.part_pre = ctx->d_pre + 2 * (size_t)(K + 1);
    e.ctl = ctx->d_ctl;
    e.dirty = ctx->d_dirty;
    e.bmin = ctx->d_bmin;
    e.arec = ctx->d_arec;
    e.hist_t = ctx->hist_t;
    e.hist_n = ctx->hist_n;
    e.hist_cap = HIST_CAP;
    e.tuned = 0;
    e.tune = ctx->d_tune;
    int rc = configure_launch(ctx);
    if (rc) return cleanup(rc);
    ctx->small_ok = small_fits(ctx->device);
    // An ensemble's other clusters keep the rest of the chip busy, so a
    // larger step in a cluster's own workgroup costs the others nothing.
    ctx->small_act = K > 1 ? SMALL_MAX_ACT : SMALL_ACT;

    double *b = ctx->state;
    Soa &s = ctx->s;
    auto take = [&]() {
        double *p = b;
        b += np;
        return p;
    };
    for (int c = 0; c < 3; ++c) s.r[c] = take();
    s.m = take();
    for (int c = 0; c < 3; ++c) s.v[c] = take();
    for (int c = 0; c < 6; ++c) s.f[c] = take();
    for (int c = 0; c < 6; ++c) s.old[c] = take();
    for (int c = 0; c < 3; ++c) s.a2[c] = take();
    for (int c = 0; c < 3; ++c) s.a3[c] = take();
    s.t = take();
    s.dt = take();
    s.pp = reinterpret_cast<double4 *>(take());  // 4 columns each
    for (int c = 0; c < 3; ++c) (void)take();
    s.pv = reinterpret_cast<double4 *>(take());
    for (int c = 0; c < 3; ++c) (void)take();
    if (b != ctx->state + (size_t)np * N_SOA) return cleanup(fail(H4H_EINVAL, "state layout"));
    ok = hipMemcpyAsync(ctx->d_off, ctx->off.data(), K * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
         hipMemcpyAsync(ctx->d_cnt, ctx->cnt.data(), K * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
         hipMemcpyAsync(ctx->d_hoff, ctx->hoff.data(), K * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
         hipMemcpyAsync(ctx->d_bcl, bcl.data(), ctx->nblk * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->state, 0, np * N_SOA * sizeof(double), ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->d_cs, 0, K * sizeof(ClState), ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->d_ctl, 0, sizeof(RunCtl), ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->d_tune, 0, 64 * sizeof(uint32_t), ctx->stream) == hipSuccess &&
         hipMemsetAsync(ctx->d_pre, 0, 3 * (size_t)(K + 1) * sizeof(unsigned long long), ctx->stream) == hipSuccess &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return cleanup(fail(H4H_EHIP, "initialisation failed"));
    *out = ctx;
    return H4H_OK;
}

// ---------------------------------------------------------------------------
// Context-free update with a static device buffer (h4h_update_acc_jrk_static)
// ---------------------------------------------------------------------------
//
// Everything the update and finish kernels need beyond the caller's inputs
// and outputs lives in `scratch`, one file-scope device array per GPU (the
// loader gives each device its own copy): tables of one cluster, the
// active list (range-checked copy), the predictor records and the forces in
// the kernels' layouts, and the j-split partial sums.  Offsets in doubles,
// each a multiple of 32 (256 bytes).
//
// Partial sums: for n <= STATIC_MAX_N with the default launch (grid <= 2048
// workgroups, queue target 16 items per workgroup, at most 256 i slots per
// item), a plan with S > 1 chunks has S <= max(s_fill, s_tiles) where
// s_fill = ceil(grid / n_it) and s_tiles = floor(n_jt / CHUNK_TILES) <= 16
// (n_jt <= 256, so the large-set chunk rule does not apply):
// S n_it iw <= (grid + n_it - 1) iw <= grid BS + n + BS - 1 in the first
// case and S rows <= 16 (n + BS - 1) in the second, so 6 S rows <=
// ST_PART doubles.  The call also checks its own plan against ST_PART.
constexpr uint32_t STATIC_MAX_N = 65536;
static_assert(STATIC_MAX_N == H4H_STATIC_MAX_N, "header and library agree");
constexpr size_t st_round32(size_t x) { return (x + 31) / 32 * 32; }
constexpr size_t ST_PART_ROWS_FILL = (size_t)DEFAULT_GRID * BS + STATIC_MAX_N + BS - 1;
constexpr size_t ST_PART_ROWS_TILES =
    (size_t)(STATIC_MAX_N / JT / CHUNK_TILES) * (STATIC_MAX_N + BS - 1);
constexpr size_t ST_PART =
    st_round32(6 * (ST_PART_ROWS_FILL > ST_PART_ROWS_TILES ? ST_PART_ROWS_FILL : ST_PART_ROWS_TILES));
constexpr size_t ST_CTL = 0, ST_CS = 32, ST_PLAN = 64, ST_PRE = 80, ST_TAB = 96;  // header
constexpr size_t ST_MOVE = 128;                                   // u32 x STATIC_MAX_N
constexpr size_t ST_PP = ST_MOVE + STATIC_MAX_N / 2;              // double4 x STATIC_MAX_N
constexpr size_t ST_PV = ST_PP + 4 * (size_t)STATIC_MAX_N;
constexpr size_t ST_F = ST_PV + 4 * (size_t)STATIC_MAX_N;         // 6 columns
constexpr size_t ST_PARTIAL = ST_F + 6 * (size_t)STATIC_MAX_N;
constexpr size_t ST_TOTAL = ST_PARTIAL + ST_PART;
static_assert(ST_TOTAL * sizeof(double) == H4H_STATIC_SCRATCH_BYTES, "header states the size");
static_assert(sizeof(RunCtl) <= 32 * sizeof(double) && sizeof(ClState) <= 32 * sizeof(double) &&
                  sizeof(ClPlan) <= 16 * sizeof(double) &&
                  6 * sizeof(unsigned long long) <= 16 * sizeof(double),
              "header block of the static buffer");
static_assert(ST_MOVE % 32 == 0 && ST_PP % 32 == 0 && ST_PV % 32 == 0 && ST_F % 32 == 0 &&
                  ST_PARTIAL % 32 == 0,
              "256-byte sections");

// ST_TAB holds off, cnt, hoff of the one cluster and the bad-index flag.
enum : uint32_t { TAB_OFF = 0, TAB_CNT = 1, TAB_HOFF = 2, TAB_BAD = 3 };

// Tables of one cluster of n particles with nact active, and the work-queue
// counters at zero (a fill kernel, so the call can be captured in a graph).
__global__ void k_static_reset(Ens e, uint32_t *tab, uint32_t n, uint32_t nact) {
    ClState c{};
    c.nact = nact;
    e.cs[0] = c;
    RunCtl r{};
    *e.ctl = r;
    tab[TAB_OFF] = 0;
    tab[TAB_CNT] = n;
    tab[TAB_HOFF] = 0;
    tab[TAB_BAD] = 0;
}

// Predictor records of p6 (n x 6: r, v) and pm (float masses) in the tile
// layout, zero-mass padding up to n_pad; the active list copied with every
// index >= n replaced by 0 and flagged in tab[TAB_BAD].
__global__ __launch_bounds__(BS) void k_static_pack(const double *p6, const float *pm, uint32_t n,
                                                   uint32_t n_pad, const uint32_t *move_in,
                                                   uint32_t nact, uint32_t *move, Soa s,
                                                   uint32_t *tab) {
    const uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i < n) {
        const float mf = pm[i];
        const double *p = p6 + (size_t)i * 6;
        s.pp[i] = make_double4(p[0], p[1], p[2], (double)mf);
        s.pv[i] = make_double4(p[3], p[4], p[5], mass_bits(mf));
    } else if (i < n_pad) {
        s.pp[i] = make_double4(PAD_POS, PAD_POS, PAD_POS, 0.0);
        s.pv[i] = make_double4(0.0, 0.0, 0.0, mass_bits(0.0f));
    }
    if (i < nact) {
        uint32_t j = move_in[i];
        if (j >= n) {
            atomicOr(&tab[TAB_BAD], 1u);
            j = 0;
        }
        move[i] = j;
    }
}

// Rows move_in[k] of f6 (n x 6) from the force columns; every row NaN if
// the list held an index >= n (those indices are skipped).
__global__ __launch_bounds__(BS) void k_static_out(Soa s, const uint32_t *move_in, uint32_t n,
                                                  uint32_t nact, const uint32_t *tab, double *f6) {
    const uint32_t k = blockIdx.x * BS + threadIdx.x;
    if (k >= nact) return;
    const uint32_t i = move_in[k];
    if (i >= n) return;
    const bool bad = tab[TAB_BAD] != 0;
    double *row = f6 + (size_t)i * 6;
    for (int q = 0; q < 6; ++q) row[q] = bad ? (double)NAN : s.f[q][i];
}

}  // namespace

// The static buffer (one per GPU).  Host code uses the device address
// from hipGetSymbolAddress, never this name.
static __device__ __align__(256) double scratch[ST_TOTAL];

namespace {

std::mutex g_static_mu;
constexpr int STATIC_MAX_DEVICES = 256;
double *g_static_ptr[STATIC_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up once per
// device (the first call there).
int static_scratch(double **out) {
    int dev = 0;
    H4H_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= STATIC_MAX_DEVICES) return fail(H4H_ENODEV, "device index out of range");
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_static_ptr[dev]) {
        void *p = nullptr;
        H4H_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(H4H_EHIP, "hipGetSymbolAddress returned null");
        g_static_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_static_ptr[dev];
    return H4H_OK;
}

}  // namespace

extern "C" {

int h4h_abi_version(void) { return 107; }

const char *h4h_last_error(void) { return g_err.c_str(); }

int h4h_create(h4h_ctx **out, uint32_t n, double e2, float eta, double dt_min, double dt_max,
               int precision) {
    if (n < 1) {
        if (out) *out = nullptr;
        return fail(H4H_EINVAL, "n must be >= 1");
    }
    return create_impl(out, 1, &n, e2, eta, dt_min, dt_max, precision);
}

int h4h_create_ensemble(h4h_ctx **out, uint32_t nclusters, const uint32_t *sizes, double e2,
                        float eta, double dt_min, double dt_max, int precision) {
    return create_impl(out, nclusters, sizes, e2, eta, dt_min, dt_max, precision);
}

int h4h_destroy(h4h_ctx *ctx) {
    if (!ctx) return H4H_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_graph(ctx);
    for (int k = 0; k < 4; ++k)
        if (ctx->ev[k]) (void)hipEventDestroy(ctx->ev[k]);
    void *dev[] = {ctx->state, ctx->staging, ctx->partial, ctx->eout, ctx->move, ctx->d_off,
                   ctx->d_cnt, ctx->d_hoff, ctx->d_bcl, ctx->d_cs, ctx->d_plan, ctx->d_pre,
                   ctx->d_ctl, ctx->d_dirty, ctx->d_bmin, ctx->d_arec, ctx->d_tune, ctx->hist_t,
                   ctx->hist_n};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    void *host[] = {ctx->ctl_host, ctx->cs_host, ctx->hist_t_host, ctx->hist_n_host};
    for (void *p : host)
        if (p) (void)hipHostFree(p);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return H4H_OK;
}

int h4h_set_precision(h4h_ctx *ctx, int precision) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (!valid_precision(precision)) return fail(H4H_EINVAL, "unknown precision");
    if (precision != ctx->precision) drop_graph(ctx);
    ctx->precision = precision;
    return H4H_OK;
}

int h4h_set_launch(h4h_ctx *ctx, uint32_t workgroups, uint32_t max_jsplit) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (workgroups > 65536) return fail(H4H_EINVAL, "workgroups > 65536");
    ctx->grid = workgroups ? workgroups : DEFAULT_GRID;
    ctx->max_split = max_jsplit ? max_jsplit : DEFAULT_MAX_SPLIT;
    return configure_launch(ctx);
}

int h4h_upload(h4h_ctx *ctx, const double *r4, const double *v4, const double *f6,
               const double *t, const double *dt) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (!r4 || !v4) return fail(H4H_EINVAL, "r4 and v4 are required");
    const Soa &s = ctx->s;
    // Zero everything first (init_data semantics), then fill.
    H4H_HIP(hipMemsetAsync(ctx->state, 0, (size_t)ctx->n_dev * N_SOA * sizeof(double), ctx->stream));
    // Largest pair weight's mass (Predictor::m is the float mass): it decides
    // whether the one-workgroup step may skip its self-pair test, a kernel
    // argument of the captured graphs.
    double mmax = 0.0;
    for (uint32_t i = 0; i < ctx->n; ++i) {
        const double m = std::fabs((double)(float)r4[4 * (size_t)i + 3]);
        mmax = std::isfinite(m) ? std::max(mmax, m) : (double)INFINITY;
    }
    if (self_pair_is_zero(ctx->precision, ctx->e2, mmax) !=
        self_pair_is_zero(ctx->precision, ctx->e2, ctx->mass_max))
        drop_graph(ctx);
    ctx->mass_max = mmax;
    Cols c{};
    c.c[0] = s.r[0]; c.c[1] = s.r[1]; c.c[2] = s.r[2]; c.c[3] = s.m;
    if ((rc = upload_cols(ctx, r4, 4, c))) return rc;
    Cols cv{};
    cv.c[0] = s.v[0]; cv.c[1] = s.v[1]; cv.c[2] = s.v[2];
    if ((rc = upload_cols(ctx, v4, 4, cv))) return rc;
    if (f6) {
        Cols cf{};
        for (int k = 0; k < 6; ++k) cf.c[k] = s.f[k];
        if ((rc = upload_cols(ctx, f6, 6, cf))) return rc;
    }
    if (t && (rc = upload_scalar(ctx, t, s.t))) return rc;
    if (dt && (rc = upload_scalar(ctx, dt, s.dt))) return rc;
    hipLaunchKernelGGL(k_init_predictor, dim3(blocks_for(ctx->n_dev)), dim3(BS), 0, ctx->stream,
                       ctx->s, ctx->e);
    H4H_LAUNCHED();
    H4H_HIP(hipMemsetAsync(ctx->d_cs, 0, ctx->K * sizeof(ClState), ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nact_host = 0;
    ctx->bm_valid = false;
    ctx->uploaded = true;
    return H4H_OK;
}

int h4h_download(h4h_ctx *ctx, double *r4, double *v4, double *f6, double *t, double *dt,
                 double *old6, double *a2, double *a3) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const Soa &s = ctx->s;
    if (r4) {
        Cols c{};
        c.c[0] = s.r[0]; c.c[1] = s.r[1]; c.c[2] = s.r[2]; c.c[3] = s.m;
        if ((rc = download_cols(ctx, r4, 4, c))) return rc;
    }
    if (v4) {
        Cols c{};
        c.c[0] = s.v[0]; c.c[1] = s.v[1]; c.c[2] = s.v[2]; c.c[3] = nullptr;
        if ((rc = download_cols(ctx, v4, 4, c))) return rc;
    }
    if (f6) {
        Cols c{};
        for (int k = 0; k < 6; ++k) c.c[k] = s.f[k];
        if ((rc = download_cols(ctx, f6, 6, c))) return rc;
    }
    if (old6) {
        Cols c{};
        for (int k = 0; k < 6; ++k) c.c[k] = s.old[k];
        if ((rc = download_cols(ctx, old6, 6, c))) return rc;
    }
    if (a2) {
        Cols c{};
        for (int k = 0; k < 3; ++k) c.c[k] = s.a2[k];
        if ((rc = download_cols(ctx, a2, 3, c))) return rc;
    }
    if (a3) {
        Cols c{};
        for (int k = 0; k < 3; ++k) c.c[k] = s.a3[k];
        if ((rc = download_cols(ctx, a3, 3, c))) return rc;
    }
    if (t && (rc = download_scalar(ctx, s.t, t))) return rc;
    if (dt && (rc = download_scalar(ctx, s.dt, dt))) return rc;
    return H4H_OK;
}

int h4h_download_predictor(h4h_ctx *ctx, double *p6, float *pm) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const Soa &s = ctx->s;
    if (p6) {
        Cols c{};
        for (int k = 0; k < 3; ++k) {
            c.c[k] = &s.pp[0].x + k;
            c.c[3 + k] = &s.pv[0].x + k;
            c.stride[k] = c.stride[3 + k] = 4;
        }
        if ((rc = download_cols(ctx, p6, 6, c))) return rc;
    }
    if (pm) {
        std::vector<double> tmp(ctx->n);
        Cols c{};
        c.c[0] = &s.pp[0].w;
        c.stride[0] = 4;
        if ((rc = download_cols(ctx, tmp.data(), 1, c))) return rc;
        for (uint32_t i = 0; i < ctx->n; ++i) pm[i] = (float)tmp[i];
    }
    return H4H_OK;
}

int h4h_upload_predictor(h4h_ctx *ctx, const double *p6, const float *pm) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const Soa &s = ctx->s;
    if (p6) {
        Cols c{};
        for (int k = 0; k < 3; ++k) {
            c.c[k] = &s.pp[0].x + k;
            c.c[3 + k] = &s.pv[0].x + k;
            c.stride[k] = c.stride[3 + k] = 4;
        }
        if ((rc = upload_cols(ctx, p6, 6, c))) return rc;
    }
    if (pm) {
        // pp.w = the mass as a double, pv.w = its float bit pattern
        std::vector<double> tmp(2 * (size_t)ctx->n);
        for (uint32_t i = 0; i < ctx->n; ++i) {
            tmp[2 * i] = (double)pm[i];
            tmp[2 * i + 1] = mass_bits(pm[i]);
        }
        Cols c{};
        c.c[0] = &s.pp[0].w;
        c.c[1] = &s.pv[0].w;
        c.stride[0] = c.stride[1] = 4;
        if ((rc = upload_cols(ctx, tmp.data(), 2, c))) return rc;
    }
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_set_active(h4h_ctx *ctx, const uint32_t *move, uint32_t nact) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (nact > ctx->n) return fail(H4H_EINVAL, "nact > n");
    if (nact && !move) return fail(H4H_EINVAL, "move is null");
    for (uint32_t k = 0; k < nact; ++k)
        if (move[k] >= ctx->n) return fail(H4H_EINVAL, "active index out of range");
    if (nact)
        H4H_HIP(hipMemcpyAsync(ctx->move, move, nact * sizeof(uint32_t), hipMemcpyHostToDevice,
                               ctx->stream));
    hipLaunchKernelGGL(k_set_nact0, dim3(1), dim3(1), 0, ctx->stream, ctx->e, nact);
    H4H_LAUNCHED();
    if ((rc = launch_plan(ctx))) return rc;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nact_host = nact;
    return H4H_OK;
}

int h4h_get_active(h4h_ctx *ctx, uint32_t *move, uint32_t *nact) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (!nact) return fail(H4H_EINVAL, "nact is null");
    *nact = ctx->nact_host;
    if (move && ctx->nact_host) {
        H4H_HIP(hipMemcpyAsync(move, ctx->move, ctx->nact_host * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, ctx->stream));
        H4H_HIP(hipStreamSynchronize(ctx->stream));
        std::sort(move, move + ctx->nact_host);
    }
    return H4H_OK;
}

// Initial forces of every cluster: each active list = the whole cluster.
int h4h_init_acc_jrk(h4h_ctx *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_identity_active, dim3(blocks_for(ctx->n_dev)), dim3(BS), 0, ctx->stream,
                       ctx->move, ctx->e);
    H4H_LAUNCHED();
    if ((rc = launch_plan(ctx))) return rc;
    if ((rc = update_pass(ctx, FIN_WRITE_F))) return rc;
    ctx->nact_host = ctx->K == 1 ? ctx->cnt[0] : 0;
    return H4H_OK;
}

static int init_dt_impl(h4h_ctx *ctx, float eta, double itime) {
    hipLaunchKernelGGL(k_reset_tmin, dim3(1), dim3(1), 0, ctx->stream, ctx->e);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_init_dt, dim3(std::min<uint32_t>(ctx->nblk, 1024)), dim3(BS), 0,
                       ctx->stream, ctx->s, ctx->e, eta, itime, ctx->dt_min, ctx->dt_max);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_set_times_from_tmin, dim3(1), dim3(1), 0, ctx->stream, ctx->e, itime);
    H4H_LAUNCHED();
    ctx->bm_valid = false;
    return read_state(ctx);
}

int h4h_init_dt(h4h_ctx *ctx, double *ctime, float eta, double itime) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (!ctime) return fail(H4H_EINVAL, "ctime is null");
    if ((rc = init_dt_impl(ctx, eta, itime))) return rc;
    double mn = dkey_inv(ctx->cs_host[0].tmin_key);
    if (mn < *ctime) *ctime = mn;
    return H4H_OK;
}

int h4h_find_particles_to_move(h4h_ctx *ctx, double itime, uint32_t *nact) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (!nact) return fail(H4H_EINVAL, "nact is null");
    hipLaunchKernelGGL(k_reset_select, dim3(1), dim3(1), 0, ctx->stream, ctx->e);
    H4H_LAUNCHED();
    if ((rc = launch_select(ctx, PS_SELECT, itime))) return rc;
    if ((rc = launch_plan(ctx))) return rc;
    if ((rc = read_state(ctx))) return rc;
    ctx->nact_host = ctx->cs_host[0].nact;
    *nact = ctx->nact_host;
    return H4H_OK;
}

int h4h_save_old_acc_jrk(h4h_ctx *ctx) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (!ctx->nact_host) return H4H_OK;
    uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(blocks_for(ctx->nact_host), 1024));
    hipLaunchKernelGGL(k_save_old, dim3(grid), dim3(BS), 0, ctx->stream, ctx->s, ctx->e, ctx->move);
    H4H_LAUNCHED();
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_predicted_pos_vel(h4h_ctx *ctx, double itime) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if ((rc = launch_select(ctx, PS_PREDICT, itime))) return rc;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_update_acc_jrk(h4h_ctx *ctx, uint32_t nact) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (nact != ctx->nact_host)
        return fail(H4H_EINVAL, "nact differs from the current active list size");
    if (nact && (rc = update_pass(ctx, FIN_WRITE_F))) return rc;
    ctx->stats.block_steps += 1;
    ctx->stats.interactions += (uint64_t)nact * ctx->n;
    ctx->stats.nsteps += nact;
    return H4H_OK;
}

int h4h_correction_pos_vel(h4h_ctx *ctx, double itime, uint32_t nact) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (nact != ctx->nact_host)
        return fail(H4H_EINVAL, "nact differs from the current active list size");
    uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(blocks_for(nact), 1024));
    hipLaunchKernelGGL(k_correct, dim3(grid), dim3(BS), 0, ctx->stream, ctx->s, ctx->e, ctx->move,
                       itime, ctx->eta, ctx->dt_min, ctx->dt_max);
    H4H_LAUNCHED();
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    ctx->bm_valid = false;
    return H4H_OK;
}

int h4h_next_integration_time(h4h_ctx *ctx, double *atime) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (!atime) return fail(H4H_EINVAL, "atime is null");
    hipLaunchKernelGGL(k_reset_tmin, dim3(1), dim3(1), 0, ctx->stream, ctx->e);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_next_time, dim3(std::min<uint32_t>(ctx->nblk, 1024)), dim3(BS), 0,
                       ctx->stream, ctx->s, ctx->e);
    H4H_LAUNCHED();
    if ((rc = read_state(ctx))) return rc;
    *atime = dkey_inv(ctx->cs_host[0].tmin_key);
    return H4H_OK;
}

int h4h_step(h4h_ctx *ctx, double itime, uint32_t *nact, double *atime) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if ((rc = ensure_blockmin(ctx))) return rc;
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    // One-step loop: ATIME = itime, stop after the step.
    hipLaunchKernelGGL(k_run_setup, dim3(1), dim3(BS), 0, ctx->stream, ctx->e, (double)INFINITY,
                       -INFINITY, 1ull, 1u, -INFINITY, itime);
    H4H_LAUNCHED();
    // The step the device loop would make: k_small_run when it takes it.
    if (small_path(ctx) && (rc = launch_small(ctx, 1))) return rc;
    if ((rc = enqueue_step(ctx, ctx->timing, 0))) return rc;
    hipLaunchKernelGGL(k_tail, dim3(1), dim3(BS), 0, ctx->stream, ctx->e);
    H4H_LAUNCHED();
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    if ((rc = read_state(ctx))) return rc;
    if (ctx->timing) {
        float ms = 0.f;
        H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        ctx->stats.step_ms += ms;
        H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->stats.update_ms += ms;
        ctx->stats.update_launches += 1;
    }
    const ClState &cs = ctx->cs_host[0];
    uint32_t k = (uint32_t)cs.nsteps;  // the step's nact (counters were reset)
    ctx->nact_host = 0;
    ctx->stats.block_steps += 1;
    ctx->stats.interactions += cs.interactions;
    ctx->stats.nsteps += k;
    ctx->stats.small_steps += cs.small_steps;
    if (nact) *nact = k;
    if (atime) *atime = cs.atime;
    return H4H_OK;
}

// Replays captured chunks of RUN_CHUNK steps until every cluster's loop has
// stopped; copies cluster 0's per-step history when asked.
static int run_chunks(h4h_ctx *ctx, double *step_time, uint32_t *step_nact, uint64_t cap) {
    int rc = ctx->graphs ? ensure_graph(ctx) : H4H_OK;
    if (rc) return rc;
    uint64_t seen = 0;
    bool hist = (step_time || step_nact) && ctx->K == 1;
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    for (;;) {
        if (ctx->graphs) {
            H4H_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
        } else if ((rc = enqueue_chunk(ctx))) {
            return rc;
        }
        H4H_HIP(hipMemcpyAsync(ctx->ctl_host, ctx->d_ctl, sizeof(RunCtl), hipMemcpyDeviceToHost,
                               ctx->stream));
        H4H_HIP(hipMemcpyAsync(ctx->cs_host, ctx->d_cs, sizeof(ClState), hipMemcpyDeviceToHost,
                               ctx->stream));
        if (hist) {
            H4H_HIP(hipMemcpyAsync(ctx->hist_t_host, ctx->hist_t, HIST_CAP * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
            H4H_HIP(hipMemcpyAsync(ctx->hist_n_host, ctx->hist_n, HIST_CAP * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
        }
        H4H_HIP(hipStreamSynchronize(ctx->stream));
        if (hist) {
            uint64_t steps = ctx->cs_host[0].steps;
            for (uint64_t k = seen; k < steps && k < cap; ++k) {
                if (step_time) step_time[k] = ctx->hist_t_host[k % HIST_CAP];
                if (step_nact) step_nact[k] = ctx->hist_n_host[k % HIST_CAP];
            }
            seen = steps;
        }
        if (ctx->ctl_host->ndone >= ctx->K) break;
    }
    if (ctx->timing) {
        H4H_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
        H4H_HIP(hipEventSynchronize(ctx->ev[3]));
        float ms = 0.f;
        H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        ctx->stats.step_ms += ms;
    }
    return read_state(ctx);
}

int h4h_run(h4h_ctx *ctx, double t_end, double t_stop, uint64_t max_steps, h4h_run_info *io,
            double *step_time, uint32_t *step_nact) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (!io) return fail(H4H_EINVAL, "io is null");
    if (max_steps == 0) return fail(H4H_EINVAL, "max_steps must be >= 1");
    if ((rc = ensure_blockmin(ctx))) return rc;
    hipLaunchKernelGGL(k_run_setup, dim3(1), dim3(BS), 0, ctx->stream, ctx->e, t_end, t_stop,
                       (unsigned long long)max_steps, 1u, io->itime, io->atime);
    H4H_LAUNCHED();
    if ((rc = run_chunks(ctx, step_time, step_nact, max_steps))) return rc;
    const ClState &d = ctx->cs_host[0];
    io->itime = d.itime;
    io->atime = d.atime;
    io->steps = d.steps;
    io->interactions = d.interactions;
    io->nsteps = d.nsteps;
    uint32_t last = 0;
    if (d.steps) {
        // nact of the last step: from the history ring
        H4H_HIP(hipMemcpyAsync(ctx->hist_n_host, ctx->hist_n, HIST_CAP * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, ctx->stream));
        H4H_HIP(hipStreamSynchronize(ctx->stream));
        last = ctx->hist_n_host[(d.steps - 1) % HIST_CAP];
    }
    io->last_nact = last;
    if (d.steps && d.itime >= t_stop)
        io->stopped = H4H_STOP_OUTPUT;
    else if (d.steps >= max_steps)
        io->stopped = H4H_STOP_MAXSTEPS;
    else
        io->stopped = H4H_STOP_END;
    ctx->nact_host = 0;
    ctx->stats.block_steps += d.steps;
    ctx->stats.interactions += d.interactions;
    ctx->stats.nsteps += d.nsteps;
    ctx->stats.small_steps += d.small_steps;
    return H4H_OK;
}

int h4h_get_energy(h4h_ctx *ctx, double *ekin, double *epot) {
    int rc = check_single(ctx);
    if (rc) return rc;
    return energies_of(ctx, 0, 1, ekin, epot);
}

int h4h_ensemble_init(h4h_ctx *ctx, float eta_s) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if ((rc = h4h_init_acc_jrk(ctx))) return rc;
    return init_dt_impl(ctx, eta_s, 0.0);
}

int h4h_ensemble_run(h4h_ctx *ctx, double t_end, double t_stop, uint64_t max_steps,
                     h4h_cluster_info *info) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (max_steps == 0) return fail(H4H_EINVAL, "max_steps must be >= 1");
    if ((rc = ensure_blockmin(ctx))) return rc;
    hipLaunchKernelGGL(k_run_setup, dim3(1), dim3(BS), 0, ctx->stream, ctx->e, t_end, t_stop,
                       (unsigned long long)max_steps, 0u, 0.0, 0.0);
    H4H_LAUNCHED();
    if ((rc = run_chunks(ctx, nullptr, nullptr, 0))) return rc;
    for (uint32_t c = 0; c < ctx->K; ++c) {
        const ClState &d = ctx->cs_host[c];
        ctx->stats.block_steps += d.steps;
        ctx->stats.interactions += d.interactions;
        ctx->stats.nsteps += d.nsteps;
        ctx->stats.small_steps += d.small_steps;
        if (info) {
            info[c].itime = d.itime;
            info[c].atime = d.atime;
            info[c].steps = d.steps;
            info[c].interactions = d.interactions;
            info[c].nsteps = d.nsteps;
            info[c].stopped = d.steps && d.itime >= t_stop ? H4H_STOP_OUTPUT
                              : d.steps >= max_steps    ? H4H_STOP_MAXSTEPS
                                                        : H4H_STOP_END;
        }
    }
    ctx->nact_host = 0;
    return H4H_OK;
}

int h4h_ensemble_energy(h4h_ctx *ctx, double *ekin, double *epot) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    return energies_of(ctx, 0, ctx->K, ekin, epot);
}

int h4h_ensemble_sizes(h4h_ctx *ctx, uint32_t *nclusters, uint32_t *sizes) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (nclusters) *nclusters = ctx->K;
    if (sizes)
        for (uint32_t c = 0; c < ctx->K; ++c) sizes[c] = ctx->cnt[c];
    return H4H_OK;
}

int h4h_set_tuning(h4h_ctx *ctx, const uint32_t *grid32, const uint32_t *iw32) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    uint32_t table[64] = {0};
    for (int b = 0; b < 32; ++b) {
        uint32_t g = grid32 ? grid32[b] : 0u;
        uint32_t w = iw32 ? iw32[b] : 0u;
        if (w != 0 && w != 4 && w != 16 && w != 64 && w != 256)
            return fail(H4H_EINVAL, "lane-group width must be 0, 4, 16, 64 or 256");
        if (g > 65536) return fail(H4H_EINVAL, "workgroup share > 65536");
        table[b] = g;
        table[32 + b] = w;
    }
    H4H_HIP(hipMemcpyAsync(ctx->d_tune, table, sizeof(table), hipMemcpyHostToDevice, ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    ctx->e.tuned = (grid32 || iw32) ? 1u : 0u;
    drop_graph(ctx);
    return H4H_OK;
}

// Measures the update kernel for nact = 2^b (b = 0 .. log2 n, plus n) with
// evenly spaced active particles, for every candidate workgroup share and
// lane-group width, and installs the fastest per bucket.  The active list
// is left empty.
int h4h_autotune(h4h_ctx *ctx, uint32_t reps, uint32_t *grid32, uint32_t *iw32) {
    int rc = check_single(ctx);
    if (rc) return rc;
    if (reps < 1) reps = 1;
    const uint32_t n = ctx->n;
    const uint32_t gmax = ctx->e.grid;
    uint32_t best_g[32] = {0}, best_w[32] = {0};
    std::vector<uint32_t> move;
    for (uint32_t b = 0; b < 32; ++b) {
        uint64_t want = 1ull << b;
        if (want > n) break;
        uint32_t nact = (2 * want > n) ? n : (uint32_t)want;  // top bucket: all n
        move.resize(nact);
        for (uint32_t k = 0; k < nact; ++k) move[k] = (uint32_t)((uint64_t)k * n / nact);
        if ((rc = h4h_set_active(ctx, move.data(), nact))) return rc;
        std::vector<uint32_t> gs, ws;
        for (uint32_t g = 64; g <= gmax; g *= 2) gs.push_back(g);
        if (gs.empty() || gs.back() != gmax) gs.push_back(gmax);
        const uint32_t wopt[4] = {4, 16, 64, 256};
        for (uint32_t w : wopt)
            if (w <= 4 * std::max<uint32_t>(nact, 1) || w == 256) ws.push_back(w);
        float best = 1e30f;
        for (uint32_t w : ws) {
            for (uint32_t g : gs) {
                uint32_t gt[32] = {0}, wt[32] = {0};
                std::copy(best_g, best_g + 32, gt);
                std::copy(best_w, best_w + 32, wt);
                gt[b] = g;
                wt[b] = w;
                if ((rc = h4h_set_tuning(ctx, gt, wt))) return rc;
                if ((rc = launch_plan(ctx))) return rc;
                if ((rc = launch_update(ctx))) return rc;  // warm
                H4H_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
                for (uint32_t r = 0; r < reps; ++r)
                    if ((rc = launch_update(ctx))) return rc;
                H4H_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
                H4H_HIP(hipEventSynchronize(ctx->ev[1]));
                float ms = 0.f;
                H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
                if (ms < best) {
                    best = ms;
                    best_g[b] = g;
                    best_w[b] = w;
                }
            }
        }
    }
    if ((rc = h4h_set_tuning(ctx, best_g, best_w))) return rc;
    if (grid32) std::copy(best_g, best_g + 32, grid32);
    if (iw32) std::copy(best_w, best_w + 32, iw32);
    hipLaunchKernelGGL(k_set_nact0, dim3(1), dim3(1), 0, ctx->stream, ctx->e, 0u);
    H4H_LAUNCHED();
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nact_host = 0;
    return H4H_OK;
}

// Launch autotuner: every candidate (persistent workgroups x largest j
// split x lane-group fill rule) runs the same `steps` device block steps
// from the context's current state, which is saved once and restored
// before each trial and at the end; the fastest is installed.
int h4h_tune_launch(h4h_ctx *ctx, uint32_t steps, h4h_launch_choice *choice) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (steps < 1 || steps > (1u << 20)) return fail(H4H_EINVAL, "steps must be in [1, 2^20]");
    const size_t sbytes = (size_t)ctx->n_dev * N_SOA * sizeof(double);
    const size_t cbytes = (size_t)ctx->K * sizeof(ClState);
    char *save = nullptr;
    if (hipMalloc(&save, sbytes + cbytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(H4H_ENOMEM, "tune_launch: no device memory for the state copy");
    }
    auto copy_state = [&](bool store) -> int {
        if (store) {
            H4H_HIP(hipMemcpyAsync(save, ctx->state, sbytes, hipMemcpyDeviceToDevice, ctx->stream));
            H4H_HIP(hipMemcpyAsync(save + sbytes, ctx->d_cs, cbytes, hipMemcpyDeviceToDevice,
                                   ctx->stream));
        } else {
            H4H_HIP(hipMemcpyAsync(ctx->state, save, sbytes, hipMemcpyDeviceToDevice, ctx->stream));
            H4H_HIP(hipMemcpyAsync(ctx->d_cs, save + sbytes, cbytes, hipMemcpyDeviceToDevice,
                                   ctx->stream));
            ctx->bm_valid = false;
        }
        return H4H_OK;
    };
    const h4h_stats stats0 = ctx->stats;
    const uint32_t g0 = ctx->grid, s0 = ctx->max_split, f0 = ctx->iw_fill;
    const uint32_t grids[5] = {256, 512, 1024, 2048, 4096};
    const uint32_t splits[4] = {16, 64, 256, 1024};
    const uint32_t fills[3] = {8, 12, 15};
    float best = 1e30f;
    uint32_t bg = g0, bs = s0, bf = f0;
    rc = copy_state(true);
    for (uint32_t gi = 0; gi < 5 && !rc; ++gi)
        for (uint32_t si = 0; si < 4 && !rc; ++si)
            for (uint32_t fi = 0; fi < 3 && !rc; ++fi) {
                ctx->grid = grids[gi];
                ctx->max_split = splits[si];
                ctx->iw_fill = fills[fi];
                if ((rc = configure_launch(ctx))) break;
                // untimed run: graph capture and first-touch costs
                if ((rc = copy_state(false))) break;
                if ((rc = h4h_ensemble_run(ctx, INFINITY, INFINITY, std::min<uint32_t>(steps, RUN_CHUNK),
                                           nullptr)))
                    break;
                if ((rc = copy_state(false))) break;
                if (hipEventRecord(ctx->ev[0], ctx->stream) != hipSuccess) { rc = fail(H4H_EHIP, "event"); break; }
                if ((rc = h4h_ensemble_run(ctx, INFINITY, INFINITY, steps, nullptr))) break;
                float ms = 0.f;
                if (hipEventRecord(ctx->ev[1], ctx->stream) != hipSuccess ||
                    hipEventSynchronize(ctx->ev[1]) != hipSuccess ||
                    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]) != hipSuccess) {
                    rc = fail(H4H_EHIP, "tune_launch: event timing failed");
                    break;
                }
                if (ms < best) {
                    best = ms;
                    bg = grids[gi];
                    bs = splits[si];
                    bf = fills[fi];
                }
            }
    ctx->grid = rc ? g0 : bg;
    ctx->max_split = rc ? s0 : bs;
    ctx->iw_fill = rc ? f0 : bf;
    int rc2 = configure_launch(ctx);
    if (!rc2) rc2 = copy_state(false);
    if (!rc2 && hipStreamSynchronize(ctx->stream) != hipSuccess) rc2 = fail(H4H_EHIP, "sync");
    (void)hipFree(save);
    ctx->stats = stats0;
    if (rc) return rc;
    if (rc2) return rc2;
    if (choice) {
        choice->workgroups = bg;
        choice->max_jsplit = bs;
        choice->iw_fill16 = bf;
        choice->ms_per_step = best / (float)steps;
    }
    return H4H_OK;
}

int h4h_set_launch_choice(h4h_ctx *ctx, const h4h_launch_choice *choice) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (!choice) return fail(H4H_EINVAL, "choice is null");
    if (choice->workgroups < 1 || choice->workgroups > 65536)
        return fail(H4H_EINVAL, "workgroups must be in [1, 65536]");
    if (choice->max_jsplit < 1) return fail(H4H_EINVAL, "max_jsplit must be >= 1");
    if (choice->iw_fill16 < 1 || choice->iw_fill16 > 16)
        return fail(H4H_EINVAL, "iw_fill16 must be in [1, 16]");
    ctx->grid = choice->workgroups;
    ctx->max_split = choice->max_jsplit;
    ctx->iw_fill = choice->iw_fill16;
    return configure_launch(ctx);
}

int h4h_set_graphs(h4h_ctx *ctx, int enable) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    ctx->graphs = enable != 0;
    drop_graph(ctx);
    return H4H_OK;
}

int h4h_set_small_path(h4h_ctx *ctx, uint32_t max_active) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    ctx->small_act = max_active;
    drop_graph(ctx);
    return H4H_OK;
}

int h4h_set_timing(h4h_ctx *ctx, int enable) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    ctx->timing = enable != 0;
    return H4H_OK;
}

int h4h_get_stats(h4h_ctx *ctx, h4h_stats *out, int reset) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (out) *out = ctx->stats;
    if (reset) ctx->stats = h4h_stats{};
    return H4H_OK;
}

int h4h_synchronize(h4h_ctx *ctx) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_update_acc_jrk_static(const double *d_p6, const float *d_pm, uint32_t n,
                              const uint32_t *d_move, uint32_t nact, double e2, int precision,
                              double *d_f6, void *stream) {
    if (n > STATIC_MAX_N) return fail(H4H_EINVAL, "n exceeds H4H_STATIC_MAX_N (65536)");
    if (nact > n) return fail(H4H_EINVAL, "nact > n");
    if (!(e2 >= 0.0) || !std::isfinite(e2)) return fail(H4H_EINVAL, "e2 must be finite and >= 0");
    if (!valid_precision(precision)) return fail(H4H_EINVAL, "unknown precision");
    if (nact && (!d_p6 || !d_pm || !d_move || !d_f6))
        return fail(H4H_EINVAL, "null device pointer");
    double *base = nullptr;
    int rc = static_scratch(&base);
    if (rc) return rc;
    if (nact == 0) return H4H_OK;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint32_t n_pad = (n + JT - 1) / JT * JT;

    // The plan of a default single-system context of n particles.
    const uint32_t grid = std::min(DEFAULT_GRID, max_items(n, DEFAULT_GRID, DEFAULT_MAX_SPLIT,
                                                           DEFAULT_IW_FILL));
    if (precision != H4H_FP64_EXACT) {
        const Decomp d = decompose(nact, pick_iw(nact, DEFAULT_IW_FILL), n_pad / JT, grid,
                                   DEFAULT_MAX_SPLIT);
        if (d.S > 1 && 6ull * d.S * d.rows > ST_PART)
            return fail(H4H_EINVAL, "work plan exceeds the static buffer");
    }
    Soa s{};
    s.pp = reinterpret_cast<double4 *>(base + ST_PP);
    s.pv = reinterpret_cast<double4 *>(base + ST_PV);
    for (int q = 0; q < 6; ++q) s.f[q] = base + ST_F + (size_t)q * STATIC_MAX_N;
    uint32_t *tab = reinterpret_cast<uint32_t *>(base + ST_TAB);
    uint32_t *move = reinterpret_cast<uint32_t *>(base + ST_MOVE);
    unsigned long long *pre = reinterpret_cast<unsigned long long *>(base + ST_PRE);
    double *partial = base + ST_PARTIAL;
    Ens e{};
    e.K = 1;
    e.nblk = n_pad / JT;
    e.off = tab + TAB_OFF;
    e.cnt = tab + TAB_CNT;
    e.hoff = tab + TAB_HOFF;
    e.cs = reinterpret_cast<ClState *>(base + ST_CS);
    e.plan = reinterpret_cast<ClPlan *>(base + ST_PLAN);
    e.items_pre = pre;
    e.fin_pre = pre + 2;
    e.part_pre = pre + 4;
    e.ctl = reinterpret_cast<RunCtl *>(base + ST_CTL);
    e.grid = grid;
    e.max_split = DEFAULT_MAX_SPLIT;
    e.iw_fill = DEFAULT_IW_FILL;
    e.tuned = 0;

    hipLaunchKernelGGL(k_static_reset, dim3(1), dim3(1), 0, st, e, tab, n, nact);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_static_pack, dim3(n_pad / BS), dim3(BS), 0, st, d_p6, d_pm, n, n_pad,
                       d_move, nact, move, s, tab);
    H4H_LAUNCHED();
    if (precision == H4H_FP64_EXACT) {
        hipLaunchKernelGGL(k_update_exact, dim3(grid), dim3(BS), 0, st, s, e, move, e2);
    } else {
        auto *kern = precision == H4H_MIXED ? k_update_mixed : k_update_fp64;
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BS), 0, st, s, e, move, partial, e2, 0u);
    }
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_finish, dim3(std::min<uint32_t>(1024, std::max<uint32_t>(64, e.nblk))),
                       dim3(BS), 0, st, s, e, move, partial, FIN_WRITE_F, 0.0f, 0.0, 0.0);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_static_out, dim3((nact + BS - 1) / BS), dim3(BS), 0, st, s, d_move, n,
                       nact, tab, d_f6);
    H4H_LAUNCHED();
    return H4H_OK;
}

}  // extern "C"
