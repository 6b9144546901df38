// h4hip.hip -- libh4hip.so: HIP kernels and C ABI of the MI355X Hermite-4
// block-timestep backend (declarations and contracts in include/h4hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
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

constexpr uint32_t DEFAULT_GRID = 2048;       // persistent workgroups
constexpr uint32_t DEFAULT_MAX_SPLIT = 1024;  // largest j split
constexpr uint32_t ENERGY_GRID = 1024;

// ---------------------------------------------------------------------------
// Layout kernels
// ---------------------------------------------------------------------------

struct Cols {
    double *c[8];
};

// dst.c[k][i] = src[i*ncols + k] for k < ncols (NULL columns skipped).
__global__ __launch_bounds__(BS) void k_aos_to_soa(const double *src,
                                                  uint32_t ncols, Cols dst,
                                                  uint32_t n) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    for (uint32_t k = 0; k < ncols; ++k)
        if (dst.c[k]) dst.c[k][i] = src[(size_t)i * ncols + k];
}

// dst[i*ncols + k] = src.c[k][i] (NULL columns write 0).
__global__ __launch_bounds__(BS) void k_soa_to_aos(Cols src, uint32_t ncols,
                                                  double *dst, uint32_t n) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    for (uint32_t k = 0; k < ncols; ++k)
        dst[(size_t)i * ncols + k] = src.c[k] ? src.c[k][i] : 0.0;
}

// Predictor at t_i (Hermite4::init_data, Hermite4.cpp:194-240): p = r, v;
// pm = (float)m.  Padding slots n..n_pad get zero mass far away.
__global__ __launch_bounds__(BS) void k_init_predictor(Soa s, uint32_t n,
                                                      uint32_t n_pad) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i >= n_pad) return;
    if (i < n) {
        for (int c = 0; c < 3; ++c) {
            s.p[c][i] = s.r[c][i];
            s.p[3 + c][i] = s.v[c][i];
        }
        s.pm[i] = (double)(float)s.m[i];
    } else {
        for (int c = 0; c < 3; ++c) {
            s.p[c][i] = PAD_POS;
            s.p[3 + c][i] = 0.0;
        }
        s.pm[i] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Step-loop control (one thread)
// ---------------------------------------------------------------------------

// Arms the device step loop (h4h_run / h4h_step).
__global__ void k_run_setup(DevInfo *info, double itime, double atime, double t_end,
                            double t_stop, unsigned long long max_steps) {
    info->done = 0;
    info->itime = itime;
    info->atime = atime;
    info->t_end = t_end;
    info->t_stop = t_stop;
    info->steps = 0;
    info->max_steps = max_steps;
    info->interactions = 0;
    info->nsteps = 0;
}

// Loop head of Hermite4CPU::integration (Hermite4CPU_integration.cpp:81-83):
// stop unless ITIME < t_end, then ITIME = ATIME; reset the active count
// and the next-time minimum.
__global__ void k_step_begin(DevInfo *info) {
    if (info->done) return;
    if (!(info->itime < info->t_end)) {
        info->done = 1;
        return;
    }
    info->itime = info->atime;
    info->nact = 0;
    info->tmin_key = dkey(INFINITY);
}

// Loop tail: ATIME = min(t+dt), counters, per-step history, and the stop
// test of the output interval (Hermite4CPU_integration.cpp:96-102).
__global__ void k_step_end(DevInfo *info, uint32_t n, double *hist_t, uint32_t *hist_n,
                           uint32_t hist_cap) {
    if (info->done) return;
    info->atime = dkey_inv(info->tmin_key);
    unsigned long long k = info->steps;
    hist_t[k % hist_cap] = info->itime;
    hist_n[k % hist_cap] = info->nact;
    info->steps = k + 1;
    info->interactions += (unsigned long long)info->nact * n;
    info->nsteps += info->nact;
    if (info->itime >= info->t_stop || info->steps >= info->max_steps) info->done = 1;
}

// Single-method ABI calls: set the block time / reset counters.
__global__ void k_set_time(DevInfo *info, double itime) {
    info->done = 0;
    info->itime = itime;
}

__global__ void k_begin_select(DevInfo *info) {
    info->done = 0;
    info->nact = 0;
}

__global__ void k_reset_tmin(DevInfo *info) {
    info->done = 0;
    info->tmin_key = dkey(INFINITY);
}

__global__ void k_set_nact(DevInfo *info, uint32_t nact) {
    info->done = 0;
    info->nact = nact;
}

__global__ __launch_bounds__(BS) void k_identity_active(uint32_t *move,
                                                       uint32_t n) {
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    if (i < n) move[i] = i;
}

// ---------------------------------------------------------------------------
// Predictor / active selection / save old (all n, O(n))
// ---------------------------------------------------------------------------

enum : uint32_t { PS_PREDICT = 1, PS_SELECT = 2, PS_SAVE_OLD = 4 };

// PS_PREDICT: Hermite4CPU::predicted_pos_vel (Hermite4CPU.cpp:340-365).
// PS_SELECT:  Hermite4::find_particles_to_move (Hermite4.cpp:63-84): the
//             active set is appended to move[] with one atomic per
//             workgroup (ballot per wave, prefix over the waves in LDS).
// PS_SAVE_OLD: Hermite4::save_old_acc_jrk (Hermite4.cpp:145-153).
// The block time is info->itime.
__global__ __launch_bounds__(BS) void k_predict_select(Soa s, uint32_t n, uint32_t flags,
                                                      uint32_t *move, DevInfo *info) {
#pragma clang fp contract(off)
    __shared__ uint32_t wave_off[BS / 64 + 1];
    if (info->done) return;
    const double itime = info->itime;
    uint32_t i = blockIdx.x * BS + threadIdx.x;
    bool in = i < n;
    bool act = false;
    if (in) {
        double ti = s.t[i];
        if (flags & PS_PREDICT) {
            double f[6], r[3], v[3];
            for (int c = 0; c < 6; ++c) f[c] = s.f[c][i];
            for (int c = 0; c < 3; ++c) {
                r[c] = s.r[c][i];
                v[c] = s.v[c][i];
            }
            double dt = itime - ti;
            double dt2 = (dt * dt);
            double dt3 = (dt2 * dt);
            double dt2c = dt2 / 2.0;
            double dt3c = dt3 / 6.0;
            for (int c = 0; c < 3; ++c) {
                s.p[c][i] = (dt3c * f[3 + c]) + (dt2c * f[c]) + (dt * v[c]) + r[c];
                s.p[3 + c][i] = (dt2c * f[3 + c]) + (dt * f[c]) + v[c];
            }
            s.pm[i] = (double)(float)s.m[i];
        }
        if (flags & PS_SELECT) {
            double tmp_time = ti + s.dt[i];
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
                uint32_t c = wave_off[k];
                wave_off[k] = total;
                total += c;
            }
            wave_off[BS / 64] = total ? atomicAdd(&info->nact, total) : 0;
        }
        __syncthreads();
        if (act) {
            unsigned long long below = mask & ((1ull << lane) - 1ull);
            move[wave_off[BS / 64] + wave_off[w] + (uint32_t)__popcll(below)] = i;
            if (flags & PS_SAVE_OLD)
                for (int c = 0; c < 6; ++c) s.old[c][i] = s.f[c][i];
        }
    }
}

__global__ __launch_bounds__(BS) void k_save_old(Soa s, const uint32_t *move,
                                                const DevInfo *info) {
    uint32_t nact = info->nact;
    for (uint32_t k = blockIdx.x * BS + threadIdx.x; k < nact;
         k += gridDim.x * BS) {
        uint32_t i = move[k];
        for (int c = 0; c < 6; ++c) s.old[c][i] = s.f[c][i];
    }
}

// ---------------------------------------------------------------------------
// Pairwise acceleration + jerk
// ---------------------------------------------------------------------------

// Loads JT consecutive predictor records starting at j0 into LDS; each of
// the BS (= JT) threads moves one record with coalesced SoA loads.
// The fourth slot of sv carries the mass as a float bit pattern in its low
// 32 bits, so the mixed-precision kernel reads it without a conversion.
__device__ __forceinline__ void load_jtile(const Soa &s, uint32_t j0,
                                           double4 *sp, double4 *sv) {
    uint32_t j = j0 + threadIdx.x;
    double m = s.pm[j];
    unsigned long long mf = __builtin_bit_cast(uint32_t, (float)m);
    sp[threadIdx.x] = make_double4(s.p[0][j], s.p[1][j], s.p[2][j], m);
    sv[threadIdx.x] = make_double4(s.p[3][j], s.p[4][j], s.p[5][j],
                                   __builtin_bit_cast(double, mf));
}

__device__ __forceinline__ float mass_f32(const double4 &b) {
    return __builtin_bit_cast(float, (uint32_t)__builtin_bit_cast(unsigned long long, b.w));
}

// FP64_EXACT: one lane per active i, j = 0..n-1 ascending, j == i skipped,
// exactly Hermite4CPU::update_acc_jrk (Hermite4CPU.cpp:309-336).  Writes
// f[i] directly (no j split).  Workgroups stride over i tiles.
__global__ __launch_bounds__(BS) void k_update_exact(Soa s,
                                                    const uint32_t *move,
                                                    const DevInfo *info,
                                                    uint32_t n, double e2) {
    __shared__ double4 sp[JT];
    __shared__ double4 sv[JT];
    if (info->done) return;
    const uint32_t nact = info->nact;
    const uint32_t n_it = (nact + BS - 1) / BS;
    for (uint32_t it = blockIdx.x; it < n_it; it += gridDim.x) {
        uint32_t k = it * BS + threadIdx.x;
        bool act = k < nact;
        uint32_t i = act ? move[k] : 0;
        double pi[6];
        for (int c = 0; c < 6; ++c) pi[c] = s.p[c][i];
        double fo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (uint32_t j0 = 0; j0 < n; j0 += JT) {
            __syncthreads();
            load_jtile(s, j0, sp, sv);
            __syncthreads();
            uint32_t jn = min((uint32_t)JT, n - j0);
            for (uint32_t jj = 0; jj < jn; ++jj) {
                double4 a = sp[jj];
                double4 b = sv[jj];
                if (j0 + jj != i)
                    pair_exact(pi, a.x, a.y, a.z, b.x, b.y, b.z, a.w, e2, fo);
            }
        }
        if (act)
            for (int c = 0; c < 6; ++c) s.f[c][i] = fo[c];
    }
}

// Inner loop of one lane over one LDS tile: the lane's j sub-group js
// takes tile entries js, js + G, ..., G = BS/IW (all JT entries when
// IW = BS).  SELF: the tile may hold the lane's own i, whose term is
// skipped (mass 0, and r2 kept finite when e2 == 0).  MIXED: FP32 pair
// terms summed in FP32 over runs of at most 32 pairs, each run then added
// to the FP64 accumulators.
template <int IW, bool SELF, bool MIXED>
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
            double4 a = sp[jj];
            double4 b = sv[jj];
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
#pragma unroll 4
            for (uint32_t q = q0; q < q0 + FOLD; ++q) {
                uint32_t jj = js + q * G;
                double4 a = sp[jj];
                double4 b = sv[jj];
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

// FP64 and mixed precision: persistent workgroups over (i tile, j chunk)
// items.  With one chunk the result goes to f[i]; with several, chunk
// partial sums go to partial[c][chunk][k] and k_reduce_partials adds them
// in chunk order.
template <int IW, bool MIXED>
__device__ __forceinline__ void update_body(const Soa &s, const uint32_t *move, uint32_t nact,
                                            double *partial, uint32_t n_jt,
                                            uint32_t max_split, double e2, double4 *sp,
                                            double4 *sv, double *red) {
    const Decomp d = decompose(nact, IW, n_jt, gridDim.x, max_split);
    const uint32_t il = threadIdx.x % IW;
    const uint32_t js = threadIdx.x / IW;
    for (uint32_t item = blockIdx.x; item < d.n_items; item += gridDim.x) {
        uint32_t it = item / d.S;
        uint32_t sc = item - it * d.S;
        uint32_t k = it * IW + il;
        bool act = k < nact;
        uint32_t i = act ? move[k] : move[0];
        double pi[6];
        for (int c = 0; c < 6; ++c) pi[c] = s.p[c][i];
        double fo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        uint32_t t0 = sc * d.tiles_per_chunk;
        uint32_t t1 = min(n_jt, t0 + d.tiles_per_chunk);
        for (uint32_t tt = t0; tt < t1; ++tt) {
            uint32_t j0 = tt * JT;
            __syncthreads();
            load_jtile(s, j0, sp, sv);
            __syncthreads();
            bool self = (i - j0) < (uint32_t)JT;
            if (__ballot(self))
                tile_pairs<IW, true, MIXED>(pi, i, j0, js, sp, sv, e2, fo);
            else
                tile_pairs<IW, false, MIXED>(pi, i, j0, js, sp, sv, e2, fo);
        }
        reduce_subgroups<IW>(fo, red);
        if (threadIdx.x < IW && act) {
            if (d.S == 1) {
                for (int c = 0; c < 6; ++c) s.f[c][i] = fo[c];
            } else {
                size_t plane = (size_t)d.S * d.rows;
                size_t off = (size_t)sc * d.rows + k;
                for (int c = 0; c < 6; ++c) partial[c * plane + off] = fo[c];
            }
        }
    }
}

// One kernel for every active-set size: the lane-group width IW is picked
// on the device from nact (pick_iw), so the launch needs no host knowledge
// of nact.
template <bool MIXED>
__global__ __launch_bounds__(BS) void k_update_fp64(Soa s, const uint32_t *move,
                                                   const DevInfo *info, double *partial,
                                                   uint32_t n_jt, uint32_t max_split,
                                                   double e2) {
    __shared__ double4 sp[JT];
    __shared__ double4 sv[JT];
    __shared__ double red[BS * 6];
    if (info->done) return;
    const uint32_t nact = info->nact;
    if (nact == 0) return;
    switch (pick_iw(nact)) {
        case 4:
            update_body<4, MIXED>(s, move, nact, partial, n_jt, max_split, e2, sp, sv, red);
            break;
        case 16:
            update_body<16, MIXED>(s, move, nact, partial, n_jt, max_split, e2, sp, sv, red);
            break;
        case 64:
            update_body<64, MIXED>(s, move, nact, partial, n_jt, max_split, e2, sp, sv, red);
            break;
        default:
            update_body<256, MIXED>(s, move, nact, partial, n_jt, max_split, e2, sp, sv, red);
            break;
    }
}

// Adds the S chunk partials of every active particle in chunk order.
// Few chunks: one lane per particle.  Many chunks: one workgroup per
// particle, lanes striding over chunks, then a fixed-order tree.
__global__ __launch_bounds__(BS) void k_reduce_partials(Soa s,
                                                       const uint32_t *move,
                                                       const DevInfo *info,
                                                       const double *partial,
                                                       uint32_t n_jt,
                                                       uint32_t grid,
                                                       uint32_t max_split) {
    __shared__ double red[BS / 64];
    if (info->done) return;
    const uint32_t nact = info->nact;
    if (nact == 0) return;
    const Decomp d = decompose(nact, pick_iw(nact), n_jt, grid, max_split);
    if (d.S == 1) return;
    size_t plane = (size_t)d.S * d.rows;
    if (d.S <= 32) {
        for (uint32_t k = blockIdx.x * BS + threadIdx.x; k < nact; k += gridDim.x * BS) {
            uint32_t i = move[k];
            for (int c = 0; c < 6; ++c) {
                double acc = 0.0;
                for (uint32_t sc = 0; sc < d.S; ++sc)
                    acc += partial[c * plane + (size_t)sc * d.rows + k];
                s.f[c][i] = acc;
            }
        }
        return;
    }
    for (uint32_t k = blockIdx.x; k < nact; k += gridDim.x) {
        uint32_t i = move[k];
        for (int c = 0; c < 6; ++c) {
            double acc = 0.0;
            for (uint32_t sc = threadIdx.x; sc < d.S; sc += BS)
                acc += partial[c * plane + (size_t)sc * d.rows + k];
            for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
            __syncthreads();
            if (__lane_id() == 0) red[threadIdx.x >> 6] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                double t = red[0];
                for (uint32_t w = 1; w < BS / 64; ++w) t += red[w];
                s.f[c][i] = t;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Corrector + Aarseth timestep + block quantisation (active only)
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
// NbodyUtils::get_timestep_normal (NbodyUtils.cpp:337-365).
__global__ __launch_bounds__(BS) void k_correct(Soa s, const uint32_t *move,
                                               const DevInfo *info, float eta,
                                               double dt_min, double dt_max) {
#pragma clang fp contract(off)
    if (info->done) return;
    const double itime = info->itime;
    const uint32_t nact = info->nact;
    for (uint32_t k = blockIdx.x * BS + threadIdx.x; k < nact;
         k += gridDim.x * BS) {
        uint32_t i = move[k];
        double f[6], o[6], a2[3], a3[3];
        for (int c = 0; c < 6; ++c) {
            f[c] = s.f[c][i];
            o[c] = s.old[c][i];
        }
        double dt1 = s.dt[i];
        double dt2 = dt1 * dt1;
        double dt3 = dt2 * dt1;
        double dt4 = dt2 * dt2;
        double dt5 = dt4 * dt1;
        for (int c = 0; c < 3; ++c)
            a2[c] = (-6 * (o[c] - f[c]) - dt1 * (4 * o[3 + c] + 2 * f[3 + c])) / dt2;
        for (int c = 0; c < 3; ++c)
            a3[c] = (12 * (o[c] - f[c]) + 6 * dt1 * (o[3 + c] + f[3 + c])) / dt3;
        for (int c = 0; c < 3; ++c) {
            s.r[c][i] = s.p[c][i] + (dt4 / 24) * a2[c] + (dt5 / 120) * a3[c];
            s.v[c][i] = s.p[3 + c][i] + (dt3 / 6) * a2[c] + (dt4 / 24) * a3[c];
            s.a2[c][i] = a2[c];
            s.a3[c][i] = a3[c];
        }
        s.t[i] = itime;

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
        s.dt[i] = normalize_dt(normal_dt, dt1, itime, dt_min, dt_max);
    }
}

// ---------------------------------------------------------------------------
// Block-time reductions
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

// Hermite4::next_integration_time (Hermite4.cpp:90-102): min(t_i + dt_i).
__global__ __launch_bounds__(BS) void k_next_time(const double *t,
                                                 const double *dt, uint32_t n,
                                                 DevInfo *info) {
    __shared__ double red[BS / 64];
    if (info->done) return;
    double mn = INFINITY;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < n;
         i += gridDim.x * BS)
        mn = fmin(mn, t[i] + dt[i]);
    mn = block_min(mn, red);
    if (threadIdx.x == 0) atomicMin(&info->tmin_key, dkey(mn));
}

// Hermite4::init_dt (Hermite4.cpp:108-140) per particle, plus min(dt).
// Non-positive, infinite or NaN ETA|a|/|j| map to dt_min (the reference's
// x86 outcome of its undefined (int) cast).
__global__ __launch_bounds__(BS) void k_init_dt(Soa s, uint32_t n,
                                               float eta, double itime,
                                               double dt_min, double dt_max,
                                               DevInfo *info) {
#pragma clang fp contract(off)
    __shared__ double red[BS / 64];
    double mn = INFINITY;
    for (uint32_t i = blockIdx.x * BS + threadIdx.x; i < n;
         i += gridDim.x * BS) {
        double a2 = magnitude(s.f[0][i], s.f[1][i], s.f[2][i]);
        double j2 = magnitude(s.f[3][i], s.f[4][i], s.f[5][i]);
        double tmp_dt = eta * (a2 / j2);
        double d;
        if (!(tmp_dt > 0.0) || isinf(tmp_dt)) {
            d = dt_min;
        } else {
            int e = (int)(ceil(log(tmp_dt) / log(2.0)) - 1);
            d = ldexp(1.0, e);
            if (d < dt_min)
                d = dt_min;
            else if (d > dt_max)
                d = dt_max;
        }
        s.dt[i] = d;
        s.t[i] = itime;
        mn = fmin(mn, d);
    }
    mn = block_min(mn, red);
    if (threadIdx.x == 0) atomicMin(&info->tmin_key, dkey(mn));
}

// ---------------------------------------------------------------------------
// Energy: sum 1/2 m v^2 - sum_{i<j} m_i m_j / |r_ij| (unsoftened),
// NbodyUtils::get_energy (NbodyUtils.cpp:427-529).  Pair terms round as
// there; sums are per-lane Neumaier-compensated, then per workgroup, then
// over workgroups in index order (deterministic for a fixed grid).
// ---------------------------------------------------------------------------

__device__ __forceinline__ void neumaier(double &sum, double &comp, double x) {
#pragma clang fp contract(off)
    double t = sum + x;
    if (fabs(sum) >= fabs(x))
        comp += (sum - t) + x;
    else
        comp += (x - t) + sum;
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

// Items are pairs of (i tile, j tile) with j tile >= i tile, numbered row
// by row; out[4*block] = {epot, epot_comp, ekin, ekin_comp}.
__global__ __launch_bounds__(BS) void k_energy(Soa s, uint32_t n,
                                              double *out) {
#pragma clang fp contract(off)
    __shared__ double4 sr[JT];
    __shared__ double red[2 * (BS / 64)];
    const uint32_t nt = (n + BS - 1) / BS;
    const uint64_t items = (uint64_t)nt * (nt + 1) / 2;
    double ep = 0.0, epc = 0.0, ek = 0.0, ekc = 0.0;
    // Row `it` holds nt - it items.  (it, rem) locate the current item and
    // advance by gridDim.x per iteration, so the walk is O(nt) in total.
    uint32_t it = 0;
    uint64_t rem = blockIdx.x;
    for (uint64_t item = blockIdx.x; item < items; item += gridDim.x) {
        while (rem >= (uint64_t)(nt - it)) {
            rem -= (nt - it);
            ++it;
        }
        uint32_t jt = it + (uint32_t)rem;
        rem += gridDim.x;
        uint32_t i = it * BS + threadIdx.x;
        uint32_t j0 = jt * JT;
        __syncthreads();
        {
            uint32_t j = j0 + threadIdx.x;
            if (j < n)
                sr[threadIdx.x] = make_double4(s.r[0][j], s.r[1][j], s.r[2][j], s.m[j]);
        }
        __syncthreads();
        if (i < n) {
            double xi = s.r[0][i], yi = s.r[1][i], zi = s.r[2][i], mi = s.m[i];
            uint32_t jn = min((uint32_t)JT, n - j0);
            uint32_t jb = (jt == it) ? threadIdx.x + 1 : 0;
            for (uint32_t jj = jb; jj < jn; ++jj) {
                double4 q = sr[jj];
                double rx = q.x - xi;
                double ry = q.y - yi;
                double rz = q.z - zi;
                double r2 = rx * rx + ry * ry + rz * rz;
                neumaier(ep, epc, -((mi * q.w) / sqrt(r2)));
            }
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
        out[4 * blockIdx.x + 0] = ep;
        out[4 * blockIdx.x + 1] = epc;
        out[4 * blockIdx.x + 2] = ek;
        out[4 * blockIdx.x + 3] = ekc;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

namespace {
constexpr uint32_t HIST_CAP = 64;      // per-step history ring (>= RUN_CHUNK)
constexpr uint32_t RUN_CHUNK = 32;     // block steps per captured graph
constexpr int N_SOA = 35;              // doubles per particle in `state`
}  // namespace

struct h4h_ctx {
    uint32_t n = 0;
    uint32_t n_pad = 0;
    uint32_t n_jt = 0;
    double e2 = 0.0;
    float eta = 0.0f;
    double dt_min = 0.0, dt_max = 0.0;
    int precision = H4H_FP64;
    int device = 0;
    uint32_t grid = DEFAULT_GRID;     // requested persistent workgroups
    uint32_t grid_eff = DEFAULT_GRID; // used (<= grid, bounded by the work n allows)
    uint32_t max_split = DEFAULT_MAX_SPLIT;
    bool uploaded = false;
    bool timing = false;
    uint32_t nact_host = 0;  // size of the current device active list

    hipStream_t stream = nullptr;
    double *state = nullptr;     // all SoA arrays
    double *staging = nullptr;   // n_pad * 8 doubles
    double *partial = nullptr;   // j-split partials
    size_t partial_len = 0;
    double *eout = nullptr;      // energy per-workgroup sums
    uint32_t *move = nullptr;
    DevInfo *info = nullptr;
    DevInfo *info_host = nullptr;  // pinned
    double *hist_t = nullptr;
    uint32_t *hist_n = nullptr;
    double *hist_t_host = nullptr;   // pinned
    uint32_t *hist_n_host = nullptr; // pinned
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    Soa s{};
    h4h_stats stats{};
};

namespace {

int check_ctx(h4h_ctx *ctx, bool need_upload = true) {
    if (!ctx) return fail(H4H_EINVAL, "null context");
    if (need_upload && !ctx->uploaded)
        return fail(H4H_ESTATE, "no state uploaded (call h4h_upload first)");
    return H4H_OK;
}

uint32_t blocks_for(uint32_t n) { return (n + BS - 1) / BS; }

// Largest number of update work items any active-set size can produce for
// this n (the persistent grid is capped by it, so small systems do not
// launch thousands of idle workgroups).
uint32_t max_items(uint32_t n, uint32_t n_jt, uint32_t grid, uint32_t max_split) {
    uint32_t best = 1;
    for (uint32_t nact = 1;; nact = (nact < 64 ? nact + 1 : nact * 2)) {
        uint32_t k = nact > n ? n : nact;
        Decomp d = decompose(k, pick_iw(k), n_jt, grid, max_split);
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

int ensure_partial(h4h_ctx *ctx) {
    size_t need = (size_t)6 * 2 * ctx->grid * BS;
    ctx->grid_eff = std::min(ctx->grid, max_items(ctx->n, ctx->n_jt, ctx->grid, ctx->max_split));
    drop_graph(ctx);
    if (ctx->partial && ctx->partial_len >= need) return H4H_OK;
    if (ctx->partial) H4H_HIP(hipFree(ctx->partial));
    ctx->partial = nullptr;
    H4H_HIP(hipMalloc(&ctx->partial, need * sizeof(double)));
    ctx->partial_len = need;
    return H4H_OK;
}

// The update pass: pair kernel + chunk reduction.  Both read nact and the
// stop flag from the device.
int launch_update(h4h_ctx *ctx) {
    if (ctx->precision == H4H_FP64_EXACT) {
        hipLaunchKernelGGL(k_update_exact, dim3(ctx->grid_eff), dim3(BS), 0, ctx->stream, ctx->s,
                           ctx->move, ctx->info, ctx->n, ctx->e2);
        H4H_LAUNCHED();
        return H4H_OK;
    }
    auto *kern = ctx->precision == H4H_MIXED ? k_update_fp64<true> : k_update_fp64<false>;
    hipLaunchKernelGGL(kern, dim3(ctx->grid_eff), dim3(BS), 0, ctx->stream, ctx->s, ctx->move,
                       ctx->info, ctx->partial, ctx->n_jt, ctx->max_split, ctx->e2);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_reduce_partials, dim3(1024), dim3(BS), 0, ctx->stream, ctx->s, ctx->move,
                       ctx->info, ctx->partial, ctx->n_jt, ctx->grid_eff, ctx->max_split);
    H4H_LAUNCHED();
    return H4H_OK;
}

int launch_predict_select(h4h_ctx *ctx, uint32_t flags) {
    hipLaunchKernelGGL(k_predict_select, dim3(blocks_for(ctx->n)), dim3(BS), 0, ctx->stream,
                       ctx->s, ctx->n, flags, ctx->move, ctx->info);
    H4H_LAUNCHED();
    return H4H_OK;
}

int launch_correct(h4h_ctx *ctx) {
    uint32_t grid = std::min<uint32_t>(blocks_for(ctx->n), 1024);
    hipLaunchKernelGGL(k_correct, dim3(grid), dim3(BS), 0, ctx->stream, ctx->s, ctx->move,
                       ctx->info, ctx->eta, ctx->dt_min, ctx->dt_max);
    H4H_LAUNCHED();
    return H4H_OK;
}

int launch_next_time(h4h_ctx *ctx) {
    uint32_t grid = std::min<uint32_t>(blocks_for(ctx->n), 512);
    hipLaunchKernelGGL(k_next_time, dim3(grid), dim3(BS), 0, ctx->stream, ctx->s.t, ctx->s.dt,
                       ctx->n, ctx->info);
    H4H_LAUNCHED();
    return H4H_OK;
}

// One block step on the stream, entirely device-driven: loop head,
// predictor + selection + save-old, update, corrector, next time, loop tail.
int enqueue_step(h4h_ctx *ctx) {
    int rc;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, ctx->stream, ctx->info);
    H4H_LAUNCHED();
    if ((rc = launch_predict_select(ctx, PS_PREDICT | PS_SELECT | PS_SAVE_OLD))) return rc;
    if ((rc = launch_update(ctx))) return rc;
    if ((rc = launch_correct(ctx))) return rc;
    if ((rc = launch_next_time(ctx))) return rc;
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, ctx->stream, ctx->info, ctx->n,
                       ctx->hist_t, ctx->hist_n, HIST_CAP);
    H4H_LAUNCHED();
    return H4H_OK;
}

// Captures RUN_CHUNK steps into one graph (once per context and
// configuration).  The graph holds kernel nodes only.
int ensure_graph(h4h_ctx *ctx) {
    if (ctx->graph_exec) return H4H_OK;
    H4H_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = H4H_OK;
    for (uint32_t k = 0; k < RUN_CHUNK && rc == H4H_OK; ++k) rc = enqueue_step(ctx);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(H4H_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    ctx->graph = g;
    H4H_HIP(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
    return H4H_OK;
}

int read_info(h4h_ctx *ctx) {
    H4H_HIP(hipMemcpyAsync(ctx->info_host, ctx->info, sizeof(DevInfo), hipMemcpyDeviceToHost,
                           ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

// The update pass with optional HIP-event timing (bench and stats).
int update_pass(h4h_ctx *ctx, bool sync_for_timing) {
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    int rc = launch_update(ctx);
    if (rc) return rc;
    if (ctx->timing) {
        H4H_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
        if (sync_for_timing) {
            H4H_HIP(hipEventSynchronize(ctx->ev[1]));
            float ms = 0.f;
            H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
            ctx->stats.update_ms += ms;
            ctx->stats.update_launches += 1;
        }
    }
    return H4H_OK;
}

int upload_cols(h4h_ctx *ctx, const double *src, uint32_t ncols, Cols dst) {
    H4H_HIP(hipMemcpyAsync(ctx->staging, src, (size_t)ctx->n * ncols * sizeof(double),
                           hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(blocks_for(ctx->n)), dim3(BS), 0, ctx->stream,
                       ctx->staging, ncols, dst, ctx->n);
    H4H_LAUNCHED();
    return H4H_OK;
}

int download_cols(h4h_ctx *ctx, double *dst, uint32_t ncols, Cols src) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(blocks_for(ctx->n)), dim3(BS), 0, ctx->stream, src,
                       ncols, ctx->staging, ctx->n);
    H4H_LAUNCHED();
    H4H_HIP(hipMemcpyAsync(dst, ctx->staging, (size_t)ctx->n * ncols * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

bool valid_precision(int p) { return p == H4H_FP64 || p == H4H_FP64_EXACT || p == H4H_MIXED; }

}  // namespace

extern "C" {

int h4h_abi_version(void) { return 101; }

const char *h4h_last_error(void) { return g_err.c_str(); }

int h4h_create(h4h_ctx **out, uint32_t n, double e2, float eta, double dt_min,
               double dt_max, int precision) {
    if (!out) return fail(H4H_EINVAL, "out is null");
    *out = nullptr;
    if (n < 1) return fail(H4H_EINVAL, "n must be >= 1");
    if (n > (1u << 30)) return fail(H4H_EINVAL, "n too large");
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
    ctx->n = n;
    ctx->n_pad = (n + JT - 1) / JT * JT;
    ctx->n_jt = ctx->n_pad / JT;
    ctx->e2 = e2;
    ctx->eta = eta;
    ctx->dt_min = dt_min;
    ctx->dt_max = dt_max;
    ctx->precision = precision;

    auto cleanup = [&](int rc) {
        std::string keep = g_err;
        h4h_destroy(ctx);
        g_err = keep;
        return rc;
    };
    if (hipGetDevice(&ctx->device) != hipSuccess) return cleanup(fail(H4H_EHIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(H4H_EHIP, "hipStreamCreate failed"));
    size_t np = ctx->n_pad;
    if (hipMalloc(&ctx->state, np * N_SOA * sizeof(double)) != hipSuccess)
        return cleanup(fail(H4H_ENOMEM, "device allocation of the state failed"));
    if (hipMalloc(&ctx->staging, np * 8 * sizeof(double)) != hipSuccess)
        return cleanup(fail(H4H_ENOMEM, "device allocation of staging failed"));
    if (hipMalloc(&ctx->eout, ENERGY_GRID * 4 * sizeof(double)) != hipSuccess ||
        hipMalloc(&ctx->move, np * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&ctx->info, sizeof(DevInfo)) != hipSuccess ||
        hipMalloc(&ctx->hist_t, HIST_CAP * sizeof(double)) != hipSuccess ||
        hipMalloc(&ctx->hist_n, HIST_CAP * sizeof(uint32_t)) != hipSuccess)
        return cleanup(fail(H4H_ENOMEM, "device allocation failed"));
    if (hipHostMalloc(&ctx->info_host, sizeof(DevInfo), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&ctx->hist_t_host, HIST_CAP * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&ctx->hist_n_host, HIST_CAP * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(H4H_ENOMEM, "pinned allocation failed"));
    for (int k = 0; k < 4; ++k)
        if (hipEventCreate(&ctx->ev[k]) != hipSuccess)
            return cleanup(fail(H4H_EHIP, "hipEventCreate failed"));
    int rc = ensure_partial(ctx);
    if (rc) return cleanup(rc);

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
    for (int c = 0; c < 6; ++c) s.p[c] = take();
    s.pm = take();
    if (hipMemsetAsync(ctx->state, 0, np * N_SOA * sizeof(double), ctx->stream) != hipSuccess ||
        hipMemsetAsync(ctx->info, 0, sizeof(DevInfo), ctx->stream) != hipSuccess)
        return cleanup(fail(H4H_EHIP, "hipMemset failed"));
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
        return cleanup(fail(H4H_EHIP, "initialisation failed"));
    *out = ctx;
    return H4H_OK;
}

int h4h_destroy(h4h_ctx *ctx) {
    if (!ctx) return H4H_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_graph(ctx);
    for (int k = 0; k < 4; ++k)
        if (ctx->ev[k]) (void)hipEventDestroy(ctx->ev[k]);
    if (ctx->state) (void)hipFree(ctx->state);
    if (ctx->staging) (void)hipFree(ctx->staging);
    if (ctx->partial) (void)hipFree(ctx->partial);
    if (ctx->eout) (void)hipFree(ctx->eout);
    if (ctx->move) (void)hipFree(ctx->move);
    if (ctx->info) (void)hipFree(ctx->info);
    if (ctx->hist_t) (void)hipFree(ctx->hist_t);
    if (ctx->hist_n) (void)hipFree(ctx->hist_n);
    if (ctx->info_host) (void)hipHostFree(ctx->info_host);
    if (ctx->hist_t_host) (void)hipHostFree(ctx->hist_t_host);
    if (ctx->hist_n_host) (void)hipHostFree(ctx->hist_n_host);
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
    return ensure_partial(ctx);
}

int h4h_upload(h4h_ctx *ctx, const double *r4, const double *v4, const double *f6,
               const double *t, const double *dt) {
    int rc = check_ctx(ctx, false);
    if (rc) return rc;
    if (!r4 || !v4) return fail(H4H_EINVAL, "r4 and v4 are required");
    const Soa &s = ctx->s;
    size_t nb = (size_t)ctx->n_pad * sizeof(double);
    // Zero everything first (init_data semantics), then fill.
    H4H_HIP(hipMemsetAsync(ctx->state, 0, nb * N_SOA, ctx->stream));
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
    if (t) H4H_HIP(hipMemcpyAsync(s.t, t, ctx->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (dt) H4H_HIP(hipMemcpyAsync(s.dt, dt, ctx->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_init_predictor, dim3(blocks_for(ctx->n_pad)), dim3(BS), 0, ctx->stream,
                       ctx->s, ctx->n, ctx->n_pad);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_set_nact, dim3(1), dim3(1), 0, ctx->stream, ctx->info, 0u);
    H4H_LAUNCHED();
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nact_host = 0;
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
    if (t) H4H_HIP(hipMemcpyAsync(t, s.t, ctx->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (dt) H4H_HIP(hipMemcpyAsync(dt, s.dt, ctx->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_download_predictor(h4h_ctx *ctx, double *p6, float *pm) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const Soa &s = ctx->s;
    if (p6) {
        Cols c{};
        for (int k = 0; k < 6; ++k) c.c[k] = s.p[k];
        if ((rc = download_cols(ctx, p6, 6, c))) return rc;
    }
    if (pm) {
        std::vector<double> tmp(ctx->n);
        H4H_HIP(hipMemcpyAsync(tmp.data(), s.pm, ctx->n * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
        H4H_HIP(hipStreamSynchronize(ctx->stream));
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
        for (int k = 0; k < 6; ++k) c.c[k] = s.p[k];
        if ((rc = upload_cols(ctx, p6, 6, c))) return rc;
    }
    if (pm) {
        std::vector<double> tmp(ctx->n);
        for (uint32_t i = 0; i < ctx->n; ++i) tmp[i] = (double)pm[i];
        H4H_HIP(hipMemcpyAsync(s.pm, tmp.data(), ctx->n * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
        H4H_HIP(hipStreamSynchronize(ctx->stream));
    }
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_set_active(h4h_ctx *ctx, const uint32_t *move, uint32_t nact) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (nact > ctx->n) return fail(H4H_EINVAL, "nact > n");
    if (nact && !move) return fail(H4H_EINVAL, "move is null");
    for (uint32_t k = 0; k < nact; ++k)
        if (move[k] >= ctx->n) return fail(H4H_EINVAL, "active index out of range");
    if (nact)
        H4H_HIP(hipMemcpyAsync(ctx->move, move, nact * sizeof(uint32_t), hipMemcpyHostToDevice,
                               ctx->stream));
    hipLaunchKernelGGL(k_set_nact, dim3(1), dim3(1), 0, ctx->stream, ctx->info, nact);
    H4H_LAUNCHED();
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nact_host = nact;
    return H4H_OK;
}

int h4h_get_active(h4h_ctx *ctx, uint32_t *move, uint32_t *nact) {
    int rc = check_ctx(ctx);
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

int h4h_init_acc_jrk(h4h_ctx *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_identity_active, dim3(blocks_for(ctx->n)), dim3(BS), 0, ctx->stream,
                       ctx->move, ctx->n);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_set_nact, dim3(1), dim3(1), 0, ctx->stream, ctx->info, ctx->n);
    H4H_LAUNCHED();
    ctx->nact_host = ctx->n;
    if ((rc = update_pass(ctx, true))) return rc;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_init_dt(h4h_ctx *ctx, double *ctime, float eta, double itime) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctime) return fail(H4H_EINVAL, "ctime is null");
    hipLaunchKernelGGL(k_reset_tmin, dim3(1), dim3(1), 0, ctx->stream, ctx->info);
    H4H_LAUNCHED();
    uint32_t grid = std::min<uint32_t>(blocks_for(ctx->n), 1024);
    hipLaunchKernelGGL(k_init_dt, dim3(grid), dim3(BS), 0, ctx->stream, ctx->s, ctx->n, eta, itime,
                       ctx->dt_min, ctx->dt_max, ctx->info);
    H4H_LAUNCHED();
    if ((rc = read_info(ctx))) return rc;
    double mn = dkey_inv(ctx->info_host->tmin_key);
    if (mn < *ctime) *ctime = mn;
    return H4H_OK;
}

int h4h_find_particles_to_move(h4h_ctx *ctx, double itime, uint32_t *nact) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!nact) return fail(H4H_EINVAL, "nact is null");
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, ctx->stream, ctx->info, itime);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_begin_select, dim3(1), dim3(1), 0, ctx->stream, ctx->info);
    H4H_LAUNCHED();
    if ((rc = launch_predict_select(ctx, PS_SELECT))) return rc;
    if ((rc = read_info(ctx))) return rc;
    ctx->nact_host = ctx->info_host->nact;
    *nact = ctx->nact_host;
    return H4H_OK;
}

int h4h_save_old_acc_jrk(h4h_ctx *ctx) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->nact_host) return H4H_OK;
    uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(blocks_for(ctx->nact_host), 1024));
    hipLaunchKernelGGL(k_save_old, dim3(grid), dim3(BS), 0, ctx->stream, ctx->s, ctx->move,
                       ctx->info);
    H4H_LAUNCHED();
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_predicted_pos_vel(h4h_ctx *ctx, double itime) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, ctx->stream, ctx->info, itime);
    H4H_LAUNCHED();
    if ((rc = launch_predict_select(ctx, PS_PREDICT))) return rc;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_update_acc_jrk(h4h_ctx *ctx, uint32_t nact) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (nact != ctx->nact_host)
        return fail(H4H_EINVAL, "nact differs from the current active list size");
    if (nact) {
        if ((rc = update_pass(ctx, true))) return rc;
    }
    ctx->stats.block_steps += 1;
    ctx->stats.interactions += (uint64_t)nact * ctx->n;
    ctx->stats.nsteps += nact;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_correction_pos_vel(h4h_ctx *ctx, double itime, uint32_t nact) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (nact != ctx->nact_host)
        return fail(H4H_EINVAL, "nact differs from the current active list size");
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, ctx->stream, ctx->info, itime);
    H4H_LAUNCHED();
    if ((rc = launch_correct(ctx))) return rc;
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    return H4H_OK;
}

int h4h_next_integration_time(h4h_ctx *ctx, double *atime) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!atime) return fail(H4H_EINVAL, "atime is null");
    hipLaunchKernelGGL(k_reset_tmin, dim3(1), dim3(1), 0, ctx->stream, ctx->info);
    H4H_LAUNCHED();
    if ((rc = launch_next_time(ctx))) return rc;
    if ((rc = read_info(ctx))) return rc;
    *atime = dkey_inv(ctx->info_host->tmin_key);
    return H4H_OK;
}

int h4h_step(h4h_ctx *ctx, double itime, uint32_t *nact, double *atime) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    // One-step loop: ITIME = ATIME = itime, stop after the step.
    hipLaunchKernelGGL(k_run_setup, dim3(1), dim3(1), 0, ctx->stream, ctx->info, -INFINITY, itime,
                       (double)INFINITY, -INFINITY, 1ull);
    H4H_LAUNCHED();
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, ctx->stream, ctx->info);
    H4H_LAUNCHED();
    if ((rc = launch_predict_select(ctx, PS_PREDICT | PS_SELECT | PS_SAVE_OLD))) return rc;
    if ((rc = update_pass(ctx, false))) return rc;
    if ((rc = launch_correct(ctx))) return rc;
    if ((rc = launch_next_time(ctx))) return rc;
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, ctx->stream, ctx->info, ctx->n,
                       ctx->hist_t, ctx->hist_n, HIST_CAP);
    H4H_LAUNCHED();
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    if ((rc = read_info(ctx))) return rc;
    if (ctx->timing) {
        float ms = 0.f;
        H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        ctx->stats.step_ms += ms;
        H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        ctx->stats.update_ms += ms;
        ctx->stats.update_launches += 1;
    }
    ctx->nact_host = ctx->info_host->nact;
    ctx->stats.block_steps += 1;
    ctx->stats.interactions += (uint64_t)ctx->nact_host * ctx->n;
    ctx->stats.nsteps += ctx->nact_host;
    if (nact) *nact = ctx->nact_host;
    if (atime) *atime = ctx->info_host->atime;
    return H4H_OK;
}

int h4h_run(h4h_ctx *ctx, double t_end, double t_stop, uint64_t max_steps, h4h_run_info *io,
            double *step_time, uint32_t *step_nact) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!io) return fail(H4H_EINVAL, "io is null");
    if (max_steps == 0) return fail(H4H_EINVAL, "max_steps must be >= 1");
    if ((rc = ensure_graph(ctx))) return rc;
    hipLaunchKernelGGL(k_run_setup, dim3(1), dim3(1), 0, ctx->stream, ctx->info, io->itime,
                       io->atime, t_end, t_stop, (unsigned long long)max_steps);
    H4H_LAUNCHED();
    uint64_t seen = 0;
    if (ctx->timing) H4H_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    for (;;) {
        H4H_HIP(hipGraphLaunch(ctx->graph_exec, ctx->stream));
        H4H_HIP(hipMemcpyAsync(ctx->info_host, ctx->info, sizeof(DevInfo), hipMemcpyDeviceToHost,
                               ctx->stream));
        if (step_time || step_nact) {
            H4H_HIP(hipMemcpyAsync(ctx->hist_t_host, ctx->hist_t, HIST_CAP * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
            H4H_HIP(hipMemcpyAsync(ctx->hist_n_host, ctx->hist_n, HIST_CAP * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
        }
        H4H_HIP(hipStreamSynchronize(ctx->stream));
        const DevInfo &d = *ctx->info_host;
        for (uint64_t k = seen; k < d.steps; ++k) {
            if (step_time) step_time[k] = ctx->hist_t_host[k % HIST_CAP];
            if (step_nact) step_nact[k] = ctx->hist_n_host[k % HIST_CAP];
        }
        seen = d.steps;
        if (d.done) break;
    }
    if (ctx->timing) {
        H4H_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
        H4H_HIP(hipEventSynchronize(ctx->ev[3]));
        float ms = 0.f;
        H4H_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
        ctx->stats.step_ms += ms;
    }
    const DevInfo &d = *ctx->info_host;
    io->itime = d.itime;
    io->atime = d.atime;
    io->steps = d.steps;
    io->interactions = d.interactions;
    io->nsteps = d.nsteps;
    io->last_nact = d.steps ? d.nact : 0;
    if (d.steps && d.itime >= t_stop)
        io->stopped = H4H_STOP_OUTPUT;
    else if (d.steps >= max_steps)
        io->stopped = H4H_STOP_MAXSTEPS;
    else
        io->stopped = H4H_STOP_END;
    ctx->nact_host = io->last_nact;
    ctx->stats.block_steps += d.steps;
    ctx->stats.interactions += d.interactions;
    ctx->stats.nsteps += d.nsteps;
    return H4H_OK;
}

int h4h_get_energy(h4h_ctx *ctx, double *ekin, double *epot) {
    int rc = check_ctx(ctx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_energy, dim3(ENERGY_GRID), dim3(BS), 0, ctx->stream, ctx->s, ctx->n,
                       ctx->eout);
    H4H_LAUNCHED();
    std::vector<double> h(ENERGY_GRID * 4);
    H4H_HIP(hipMemcpyAsync(h.data(), ctx->eout, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    H4H_HIP(hipStreamSynchronize(ctx->stream));
    // Neumaier sum of the per-workgroup (sum, compensation) pairs.
    double sp = 0.0, cp = 0.0, sk = 0.0, ck = 0.0;
    auto add = [](double &s, double &c, double x) {
        double t = s + x;
        if (std::fabs(s) >= std::fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    };
    for (uint32_t b = 0; b < ENERGY_GRID; ++b) {
        add(sp, cp, h[4 * b + 0]);
        add(sp, cp, h[4 * b + 1]);
        add(sk, ck, h[4 * b + 2]);
        add(sk, ck, h[4 * b + 3]);
    }
    if (ekin) *ekin = sk + ck;
    if (epot) *epot = sp + cp;
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

}  // extern "C"
