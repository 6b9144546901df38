// h4_kernels.hpp -- gfx950 device code of the Hermite-4 block-timestep path.
//
// State lives in HBM as structure-of-arrays (one double array per
// component) so every kernel's per-particle loads are coalesced 8-byte
// lanes; the predictor is kept as two double4 arrays in the layout of an
// LDS j tile.  The pairwise kernels stage j particles through LDS in tiles
// of JT and read them back as wave-uniform (broadcast) LDS loads.
//
// Rounding rules.  Predictor, corrector, timestep and init_dt are compiled
// with FP contraction off and use IEEE division and sqrt, so each
// operation rounds exactly as the reference's x86-64 build does
// (Hermite4CPU.cpp:340-411, NbodyUtils.cpp:337-424, Hermite4.cpp:108-140);
// the energy sum keeps the reference's terms and rounds within 1e-13.
// The FP64_EXACT pair kernel does the same for the pair terms and keeps the
// ascending-j order of Hermite4CPU::update_acc_jrk.  The FP64 pair kernel
// contracts to FMA, uses a refined v_rsq_f64 and may split the j range.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace h4 {

constexpr int BS = 256;           // threads per workgroup: 4 wave64
constexpr int JT = 256;           // j particles per LDS tile (== BS)
constexpr double PAD_POS = 1.0e18;  // position of zero-mass padding particles

// Pointer table of the SoA state (all arrays n_pad long).
struct Soa {
    double *r[3];    // corrected position
    double *m;       // mass as uploaded (double; energy uses it)
    double *v[3];    // corrected velocity
    double *f[6];    // acceleration a[3], jerk a1[3]
    double *old[6];  // forces at the start of the step
    double *a2[3];   // 2nd derivative from the Hermite interpolant
    double *a3[3];   // 3rd derivative
    double *t;       // particle time
    double *dt;      // particle block timestep
    // Predictor (Predictor, common.hpp:213-221) as two double4 arrays, the
    // layout of an LDS j tile: pp = (x, y, z, m) with m the float-rounded
    // mass as a double, pv = (vx, vy, vz, bits) with the float mass's bit
    // pattern in the low 32 bits of w (read by the mixed kernel).
    double4 *pp;
    double4 *pv;
};

// Float mass bit pattern stored in pv[i].w.
__host__ __device__ inline double mass_bits(float m) {
    unsigned long long u = __builtin_bit_cast(uint32_t, m);
    return __builtin_bit_cast(double, u);
}

// Device-side scalars of the block-step loop of one cluster (one
// independent N-body system; a plain context is one cluster).  Every step
// kernel reads the block time and the active count from here, so a step
// (or a captured chain of steps) needs no host round trip.
struct ClState {
    double itime;                     // ITIME: block time of the last step
    double atime;                     // ATIME: next block time
    uint32_t nact;                    // active count (compaction target)
    uint32_t done;                    // 1: this cluster's loop has stopped
    unsigned long long tmin_key;      // order-preserving key of min(t+dt)
    unsigned long long steps;         // block steps since the last setup
    unsigned long long interactions;  // sum of nact * n
    unsigned long long nsteps;        // sum of nact
    unsigned long long small_steps;   // block steps made by the one-workgroup kernel
    // Loop tail of the cluster's last block step, left by the step's last
    // launch for the next launch that needs the next block time to make:
    double tail_t;                    // the step's block time
    unsigned long long tail_steps;    // steps done before it
    uint32_t tail_nact;               // its active count
    uint32_t tail_pend;               // 1: that tail is still to be made
};

// Work plan of one step for one cluster (computed on the device after the
// active sets are known).
struct ClPlan {
    uint32_t iw;        // i slots per workgroup (lane-group width)
    uint32_t S;         // j chunks
    uint32_t tpc;       // JT tiles per chunk
    uint32_t n_it;      // i tiles
    uint32_t rows;      // n_it * iw
    uint32_t lanes;     // finish lanes per particle (finish_lanes(S, nact))
    uint32_t grid;      // workgroup share given to decompose()
    uint32_t cmaj;      // 1: partials chunk-major (partial_at)
};

// Place of the j-chunk partial sum of component q, chunk sc, active
// particle k in a plan's partial buffer (6 S rows doubles).  Component and
// chunk major, particle minor (a work item writes its particles' sums of
// one chunk as contiguous runs), unless the finish pass gives each
// particle 16 or more lanes and the plan has at least 64 i slots: then
// chunk minor, so that those lanes read whole 128-byte lines of one
// particle's chunks rather than one double of each of 16 or more lines
// (the items' scattered writes then cost less than the finish saves; with
// fewer i slots they cost more).  Either way the finish adds the chunks in
// the same order, so the sums do not depend on the layout.
__host__ __device__ inline size_t partial_at(const ClPlan &pl, uint32_t q, uint32_t sc,
                                             uint32_t k) {
    return pl.cmaj ? ((size_t)q * pl.rows + k) * pl.S + sc
                   : ((size_t)q * pl.S + sc) * pl.rows + k;
}

// Loop control shared by all clusters.
struct RunCtl {
    double t_end;                  // steps run while itime < t_end
    double t_stop;                 // stop after a step with itime >= t_stop
    unsigned long long max_steps;  // stop after this many steps
    uint32_t ndirty[2];            // blocks with active particles, by step parity
    uint32_t ndone;                // clusters that have stopped
    uint32_t qdone;                // update workgroups done with the work queue
    unsigned long long items;      // update work items this step
    unsigned long long fin_units;  // finish workgroup units this step
    unsigned long long qnext;      // update work queue: items claimed beyond the
                                   // first gridDim.x (zero between launches)
};

// The stop test of the reference's loop after a block step made at time T,
// the next block time being mn (Hermite4CPU_integration.cpp:81,102): an
// output time reached (t_stop), the step budget spent, ITIME >= t_end, or
// no particle left ahead of T.
__host__ __device__ inline bool loop_stops(const RunCtl &ctl, double T, unsigned long long steps,
                                           double mn) {
    return T >= ctl.t_stop || steps >= ctl.max_steps || !(T < ctl.t_end) || !(mn > T);
}

// Cluster tables and per-step scratch, by pointer (device memory).
struct Ens {
    uint32_t K;              // clusters
    uint32_t nblk;           // JT blocks of the padded state
    const uint32_t *off;     // [K] first device index (multiple of JT)
    const uint32_t *cnt;     // [K] particles
    const uint32_t *hoff;    // [K] first host index (clusters back to back)
    const uint32_t *bcl;     // [nblk] cluster of each block
    ClState *cs;             // [K]
    ClPlan *plan;            // [K]
    unsigned long long *items_pre;  // [K+1] exclusive prefix of work items
    unsigned long long *fin_pre;    // [K+1] exclusive prefix of finish units
    unsigned long long *part_pre;   // [K+1] exclusive prefix of partial doubles
    RunCtl *ctl;
    uint32_t *dirty;         // [2][nblk][3] blocks with active particles (block, first
                             // place in the cluster's active list, count), by step
                             // parity (the step's slot in its captured chunk, mod 2)
    double *bmin;            // [nblk] min(t + dt) per block
    double4 *arec;           // [4][nblk * JT] the step's active particles in the order
                             // of the active list (index off[c] + k): predictor
                             // (pp, pv), then old forces and step ((a, dt), (jerk, 0));
                             // written by the selection of a device-loop step
    double *hist_t;          // per-step history of cluster 0 (ring)
    uint32_t *hist_n;
    uint32_t hist_cap;
    uint32_t grid;           // persistent workgroups of the update kernel
    uint32_t max_split;
    uint32_t iw_fill;        // pick_iw fill rule, sixteenths
    uint32_t tuned;          // 1: tune[] holds a table for single-cluster plans
    const uint32_t *tune;    // [64]: workgroup share [0..31], lane-group width
                             // [32..63], by floor(log2(nact)); 0 = default
};

// Order-preserving map double -> uint64 for atomicMin.
__host__ __device__ inline unsigned long long dkey(double x) {
    unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline double dkey_inv(unsigned long long k) {
    unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __builtin_bit_cast(double, u);
}

// Work split of the pairwise update.  A workgroup handles IW active
// particles ("i slots") at a time; its BS lanes are IW slots x BS/IW j
// sub-groups.  The nact active particles form n_it = ceil(nact/IW) i tiles
// and the j range (n_jt tiles of JT) is cut into S chunks, giving n_it * S
// work items.  S is large enough that every one of `grid` workgroups gets
// an item when nact is small, and beyond that up to QUEUE_FACTOR items per
// workgroup (taken from the update kernel's work queue, which evens out
// faster and slower CUs and XCDs) as long as a chunk keeps CHUNK_TILES
// tiles (more chunks would mostly add partial sums to write and read).  A
// chunk of a narrow group (IW 4 or 16) holds at least 4 tiles, 4K or 16K
// pairs: shorter items cost more in partial sums than they gain (IW 16),
// longer ones leave small steps waiting on one item's tile copies (IW 4).
// Large active sets of large systems (nact >= 8192 over more than 256
// tiles) keep chunks of CHUNK_TILES_LARGE tiles: 8 to 32 items per resident
// workgroup even out the end of the launch, and their finish pass reads
// whole lines of partials (finish_lanes: 4 lanes per particle); measured at
// N = 262144, nact 8192 to 32768: 2-5 % less update time.
constexpr uint32_t QUEUE_FACTOR = 16;
constexpr uint32_t CHUNK_TILES = 16;
constexpr uint32_t CHUNK_TILES_LARGE = 4;

struct Decomp {
    uint32_t n_it;             // i tiles
    uint32_t S;                // j chunks
    uint32_t tiles_per_chunk;  // JT tiles per chunk
    uint32_t n_items;          // n_it * S
    uint32_t rows;             // n_it * IW: stride of one chunk in partials
};

__host__ __device__ inline Decomp decompose(uint32_t nact, uint32_t iw,
                                            uint32_t n_jt, uint32_t grid,
                                            uint32_t max_split) {
    Decomp d;
    d.n_it = (nact + iw - 1) / iw;
    uint32_t min_tiles = iw >= 64 ? 1u : 4u;
    uint32_t S = 1;
    const uint32_t target = grid * QUEUE_FACTOR;
    if (d.n_it > 0 && d.n_it < target) {
        const uint32_t s_fill = d.n_it < grid ? (grid + d.n_it - 1) / d.n_it : 1u;
        const uint32_t s_queue = (target + d.n_it - 1) / d.n_it;
        const uint32_t ct = (nact >= 8192 && n_jt > 256) ? CHUNK_TILES_LARGE : CHUNK_TILES;
        const uint32_t s_tiles = n_jt / ct;
        const uint32_t s_most = s_fill > s_tiles ? s_fill : s_tiles;
        S = s_queue < s_most ? s_queue : s_most;
    }
    if (S > max_split) S = max_split;
    uint32_t s_cap = (n_jt + min_tiles - 1) / min_tiles;
    if (S > s_cap) S = s_cap;
    if (S < 1) S = 1;
    d.tiles_per_chunk = (n_jt + S - 1) / S;
    d.S = d.tiles_per_chunk ? (n_jt + d.tiles_per_chunk - 1) / d.tiles_per_chunk : 1;
    d.n_items = d.n_it * d.S;
    d.rows = d.n_it * iw;
    return d;
}

// Upper bound of the j-split partials (doubles) of one cluster of n
// particles over every plan decompose() can make for a workgroup share of
// at most g / QUEUE_FACTOR: any active count, any lane-group width
// iw <= BS, any split cap.  With S > 1 chunks, S <= ceil(g / n_it) (the
// queue target), so S n_it <= g + n_it - 1
// and 6 S rows = 6 S n_it iw <= 6 (g iw + n + iw); also S <= n_jt and
// rows <= n + iw, so 6 S rows <= 6 n_jt (n + iw).
__host__ __device__ inline unsigned long long partial_bound(uint32_t n, uint32_t g) {
    const unsigned long long iw = BS, n_jt = (n + JT - 1) / JT;
    const unsigned long long a = (unsigned long long)g * iw + n + iw;
    const unsigned long long b = n_jt * ((unsigned long long)n + iw);
    return 6ull * (a < b ? a : b);
}

// The same for K clusters of total size n_total whose workgroup shares
// w_c = max(1, floor(share_c * w)) (shares summing to at most 1) give
// queue targets g_c = QUEUE_FACTOR w_c adding up to at most g + QUEUE_FACTOR K,
// g = QUEUE_FACTOR w: 6 sum_c (g_c BS + n_c + BS) <= 6 (BS (g + (QUEUE_FACTOR + 1) K) + n_total).
__host__ __device__ inline unsigned long long ensemble_partial_bound(uint32_t K,
                                                                     unsigned long long n_total,
                                                                     uint32_t g) {
    return 6ull * ((unsigned long long)BS * ((unsigned long long)g + (QUEUE_FACTOR + 1ull) * K) +
                   n_total);
}

// Lanes of the finish pass that share one particle's S chunk sums: each
// lane adds at most 8 of them (1, 4, 16 or 64 lanes), except that many
// active particles (nact >= 8192, enough to fill the part) take at most 4
// lanes each: a wave then reads whole 128-byte lines of consecutive
// particles' partials rather than one double per line.
__host__ __device__ inline uint32_t finish_lanes(uint32_t S, uint32_t nact) {
    if (S <= 8) return 1u;
    if (S <= 32 || nact >= 8192) return 4u;
    return S <= 128 ? 16u : 64u;
}

// Lane-group width for nact active particles: the widest of 256/64/16/4
// that keeps at least fill16/16 of the i slots busy (default 12/16), else
// the least padded.
constexpr uint32_t DEFAULT_IW_FILL = 12;
__host__ __device__ inline uint32_t pick_iw(uint32_t nact, uint32_t fill16 = DEFAULT_IW_FILL) {
    const uint32_t opts[4] = {256, 64, 16, 4};
    uint32_t best = 256, best_pad = 0xffffffffu;
    for (int k = 0; k < 4; ++k) {
        uint32_t iw = opts[k];
        uint32_t slots = (nact + iw - 1) / iw * iw;
        if (16ull * nact >= (unsigned long long)fill16 * slots) return iw;
        if (slots < best_pad) {
            best_pad = slots;
            best = iw;
        }
    }
    return best;
}

// ---------------------------------------------------------------------------
// Pair terms
// ---------------------------------------------------------------------------

// Hermite4CPU::force_calculation (Hermite4CPU.cpp:55-82), operation for
// operation: no contraction, IEEE sqrt and division, same association.
__device__ __forceinline__ void pair_exact(const double pi[6], double xj,
                                           double yj, double zj, double vxj,
                                           double vyj, double vzj, double mj,
                                           double e2, double fo[6]) {
#pragma clang fp contract(off)
    double rx = xj - pi[0];
    double ry = yj - pi[1];
    double rz = zj - pi[2];
    double vx = vxj - pi[3];
    double vy = vyj - pi[4];
    double vz = vzj - pi[5];

    double r2 = rx * rx + ry * ry + rz * rz + e2;
    double rinv = 1.0 / sqrt(r2);
    double r2inv = rinv * rinv;
    double r3inv = r2inv * rinv;
    double r5inv = r2inv * r3inv;
    double mr3inv = r3inv * mj;
    double mr5inv = r5inv * mj;

    double rv = rx * vx + ry * vy + rz * vz;

    fo[0] += (rx * mr3inv);
    fo[1] += (ry * mr3inv);
    fo[2] += (rz * mr3inv);
    fo[3] += (vx * mr3inv - (3 * rv) * rx * mr5inv);
    fo[4] += (vy * mr3inv - (3 * rv) * ry * mr5inv);
    fo[5] += (vz * mr3inv - (3 * rv) * rz * mr5inv);
}

// The FP64 pair loop works with s = sqrt(3) / r and q = s^2 = 3 / r^2:
// the jerk coefficient 3 (r.v) / r^2 is then (r.v) q, with no multiply by
// 3, and m s q = 3 sqrt(3) m / r^3, so every pair term it sums is
// 3 sqrt(3) times the true one.  The sums are multiplied by
// PAIR_SCALE = 1 / (3 sqrt 3) once, when the update writes them out.
constexpr double PAIR_SCALE = 0.19245008972987526;

// sqrt(3) / sqrt(x) for x > 0: hardware estimate y = (1 + d) / sqrt(x)
// (v_rsq_f64; |d| < 2^-24 over the FP64 range, measured on the part over 4M inputs),
// then one Newton step sqrt(3) y (1 + e/2), e = 1 - x y^2, which leaves a
// relative error of about -1.5 d^2, below 2^-47.
__device__ __forceinline__ double rsqrt3_fp64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x, y * y, 1.0);
    return y * fma(e, 0.8660254037844386, 1.7320508075688772);  // sqrt3 / 2, sqrt3
}

// Same pair as pair_exact with FMA contraction and rsqrt3_fp64, scaled by
// 3 sqrt(3) (see PAIR_SCALE):
//   a  += m r / |r|^3
//   j  += m/|r|^3 (v - 3 (r.v)/|r|^2 r)
__device__ __forceinline__ void pair_fp64(const double pi[6], double xj,
                                          double yj, double zj, double vxj,
                                          double vyj, double vzj, double mj,
                                          double e2, double fo[6]) {
    double rx = xj - pi[0];
    double ry = yj - pi[1];
    double rz = zj - pi[2];
    double vx = vxj - pi[3];
    double vy = vyj - pi[4];
    double vz = vzj - pi[5];
    double r2 = fma(rx, rx, fma(ry, ry, fma(rz, rz, e2)));
    double s = rsqrt3_fp64(r2);  // sqrt(3) / r
    double q = s * s;            // 3 / r^2
    double mr3 = (mj * s) * q;   // 3 sqrt(3) m / r^3
    double rv = fma(rx, vx, fma(ry, vy, rz * vz));
    double c = rv * q;           // 3 (r.v) / r^2
    fo[0] = fma(rx, mr3, fo[0]);
    fo[1] = fma(ry, mr3, fo[1]);
    fo[2] = fma(rz, mr3, fo[2]);
    fo[3] = fma(mr3, fma(-c, rx, vx), fo[3]);
    fo[4] = fma(mr3, fma(-c, ry, vy), fo[4]);
    fo[5] = fma(mr3, fma(-c, rz, vz), fo[5]);
}

// FP32 pair terms from FP32 differences r = x_j - x_i, v = v_j - v_i:
// r^2, 1/r, the force and jerk terms, added into the FP32 partial sums
// t[6].  The caller folds t into FP64 accumulators every few pairs.
// 1/r is v_rsq_f32 itself (within 1 ulp): a Newton step on it (4 more
// instructions per pair) changed the N = 262144 Plummer forces' RMS
// relative error from 2.7e-8 to 1.5e-8 (acceleration) and 2.5e-7 to
// 2.4e-7 (jerk), both inside the mixed contract, the largest errors not at
// all (they come from FP32 cancellation in v - c r), and cost 8 % of the
// kernel's time.  (No sqrt(3) scaling either: the FP32 value of sqrt(3) is
// 1.8e-8 off, a bias every pair term would share.)
__device__ __forceinline__ void pair_terms_f32(float rx, float ry, float rz, float vx,
                                               float vy, float vz, float mj, float e2,
                                               float t[6]) {
    float r2 = fmaf(rx, rx, fmaf(ry, ry, fmaf(rz, rz, e2)));
    float rinv = __builtin_amdgcn_rsqf(r2);
    float rinv2 = rinv * rinv;
    float mr3 = (mj * rinv) * rinv2;
    float rv = fmaf(rx, vx, fmaf(ry, vy, rz * vz));
    float c = (3.0f * rv) * rinv2;
    t[0] = fmaf(rx, mr3, t[0]);
    t[1] = fmaf(ry, mr3, t[1]);
    t[2] = fmaf(rz, mr3, t[2]);
    t[3] = fmaf(mr3, fmaf(-c, rx, vx), t[3]);
    t[4] = fmaf(mr3, fmaf(-c, ry, vy), t[4]);
    t[5] = fmaf(mr3, fmaf(-c, rz, vz), t[5]);
}

// Mixed-precision pair: position and velocity differences in FP64 (they
// cancel for close pairs), rounded once to FP32, then pair_terms_f32.
__device__ __forceinline__ void pair_mixed(const double pi[6], double xj,
                                           double yj, double zj, double vxj,
                                           double vyj, double vzj, float mj,
                                           float e2, float t[6]) {
    pair_terms_f32((float)(xj - pi[0]), (float)(yj - pi[1]), (float)(zj - pi[2]),
                   (float)(vxj - pi[3]), (float)(vyj - pi[4]), (float)(vzj - pi[5]), mj, e2, t);
}

}  // namespace h4
