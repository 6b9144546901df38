/*
 * h4hip.h -- C ABI of libh4hip.so, the MI355X (gfx950) backend for
 * GraviDy's 4th-order Hermite block-timestep integrator.
 *
 * One opaque context owns one N-body system resident in HBM (SoA arrays)
 * and one HIP stream on the current device.  Every entry point returns 0
 * on success or a negative H4H_E* code; h4h_last_error() gives the text.
 * No entry point calls exit() or throws.  Calls on one context must not
 * run concurrently (one host thread per context, as in the reference,
 * where one host thread drives the integrator).
 *
 * Host array layouts are the reference's own structs, so a caller can pass
 * NbodySystem's arrays directly:
 *   r4  : n x 4 double  (x, y, z, m)      NbodySystem::h_r (double4)
 *   v4  : n x 4 double  (vx, vy, vz, w)   NbodySystem::h_v (double4)
 *   f6  : n x 6 double  (a[3], a1[3])     Forces (common.hpp:228-231)
 *   a3v : n x 3 double                     double3 (h_a2, h_a3)
 *   t, dt : n double
 *   move  : n uint32 (active particle indices)
 * Masses are rounded to float on upload for the pairwise kernel, as the
 * reference's Predictor::m is float (common.hpp:216); the energy uses
 * the double mass (NbodyUtils.cpp:446).
 *
 * References below are to cmaureir/gravidy, src/include/.
 */
#ifndef H4HIP_H
#define H4HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct h4h_ctx h4h_ctx;

/* Error codes. */
#define H4H_OK            0
#define H4H_EINVAL       -1   /* bad argument (null, size, precision...) */
#define H4H_EHIP         -2   /* a HIP runtime call failed               */
#define H4H_ENOMEM       -3   /* device or host allocation failed        */
#define H4H_ESTATE       -4   /* call out of order (e.g. no upload yet)  */
#define H4H_ENODEV       -5   /* no HIP device                           */

/* Pairwise-kernel precision.  The predictor, corrector and block-timestep
 * selection are always FP64 with the reference's rounding; the energy is
 * always FP64 and within 1e-13 of the reference's. */
#define H4H_FP64        0  /* FP64, rsqrt + refinement, FMA, j-split     */
#define H4H_FP64_EXACT  1  /* FP64 bit-identical to Hermite4CPU:
                              1.0/sqrt, no FMA, ascending j, no j-split   */
#define H4H_MIXED       2  /* FP64 differences, FP32 pair terms,
                              FP64 accumulation across j tiles            */

/* Counters kept by a context (64-bit; the reference's "interactions"
 * counter is 32-bit products, Hermite4CPU_integration.cpp:96). */
typedef struct h4h_stats {
    uint64_t block_steps;      /* h4h_step / update calls               */
    uint64_t interactions;     /* sum of nact * n over updates          */
    uint64_t nsteps;           /* sum of nact                           */
    uint64_t update_launches;  /* timed update passes                   */
    double   update_ms;        /* HIP-event time of the timed passes    */
    double   step_ms;          /* HIP-event time of timed h4h_step calls */
    uint64_t small_steps;      /* block steps made in one workgroup     */
} h4h_stats;

/* Version of this ABI (major*100 + minor). */
int h4h_abi_version(void);

/* Text of the last error on the calling thread ("" if none). */
const char *h4h_last_error(void);

/* Creates a context for n particles on the current HIP device.
 * e2: softening squared as the reference computes it, i.e. the float
 *     square of the float softening (NbodySystem.cpp:50);
 * eta: ETA_N / ns->eta used by the corrector's timestep (float,
 *     NbodySystem.hpp:67; Hermite4CPU.cpp:405);
 * dt_min, dt_max: block-timestep bounds (D_TIME_MIN/D_TIME_MAX,
 *     common.hpp:122,129; a run-time value here so 2^-30 is possible);
 * precision: H4H_FP64, H4H_FP64_EXACT or H4H_MIXED.
 * Replaces: Hermite4GPU::Hermite4GPU + alloc_arrays_device
 * (gpu/Hermite4GPU.cu:43-116,127-165), one device, no host staging arrays. */
int h4h_create(h4h_ctx **out, uint32_t n, double e2, float eta,
               double dt_min, double dt_max, int precision);

/* Frees all device memory of the context.
 * Replaces: Hermite4GPU::~Hermite4GPU / free_arrays_device
 * (gpu/Hermite4GPU.cu:119-122,169-194). */
int h4h_destroy(h4h_ctx *ctx);

/* Switches the pairwise-kernel precision of an existing context. */
int h4h_set_precision(h4h_ctx *ctx, int precision);

/* Tuning knobs of the FP64/mixed update kernel: persistent workgroups
 * (0 = default 2048; a step's work items beyond one per workgroup, up to
 * 16 per workgroup, are handed out by the kernel's work queue) and the
 * largest j-split factor (0 = default 1024). */
int h4h_set_launch(h4h_ctx *ctx, uint32_t workgroups, uint32_t max_jsplit);

/* Copies the state to the device.  r4 and v4 are required; f6, t, dt may
 * be NULL (then zero, as Hermite4::init_data sets them,
 * Hermite4.cpp:194-240).  The predictor is set to r, v as init_data does.
 * Replaces: Hermite4::init_data + the cudaMemcpy H2D of
 * Hermite4GPU::init_acc_jrk (gpu/Hermite4GPU.cu:333-375). */
int h4h_upload(h4h_ctx *ctx, const double *r4, const double *v4,
               const double *f6, const double *t, const double *dt);

/* Copies the state back; any pointer may be NULL to skip that array.
 * old6 is Hermite4::save_old_acc_jrk's h_old, a2/a3 the corrector's
 * h_a2/h_a3.  Call before any Logger / NbodyUtils use of the host arrays
 * (Hermite4CPU_integration.cpp:102-117 reads them at output times). */
int h4h_download(h4h_ctx *ctx, double *r4, double *v4, double *f6,
                 double *t, double *dt, double *old6, double *a2,
                 double *a3);

/* Copies the current predictor back: p6 is n x 6 (r[3], v[3]) and pm the
 * n float masses (Predictor, common.hpp:213-217).  Either may be NULL. */
int h4h_download_predictor(h4h_ctx *ctx, double *p6, float *pm);

/* Overwrites the predictor with host values (state injection for tests
 * and for callers that predict on their own). */
int h4h_upload_predictor(h4h_ctx *ctx, const double *p6, const float *pm);

/* N^2 initial acceleration and jerk from the current predictor, j != i.
 * Replaces: Hermite4CPU::init_acc_jrk (cpu/Hermite4CPU.cpp:289-304) and
 * Hermite4GPU::init_acc_jrk + k_init_acc_jrk (gpu/Hermite4GPU.cu:333-375,
 * Hermite4GPU_kernels.cu:47-87). */
int h4h_init_acc_jrk(h4h_ctx *ctx);

/* Initial block timesteps from the current forces; sets t_i = itime and
 * lowers *ctime to the smallest dt, as the reference does.
 * eta is ETA_S (float).
 * Replaces: Hermite4::init_dt (Hermite4.cpp:108-140). */
int h4h_init_dt(h4h_ctx *ctx, double *ctime, float eta, double itime);

/* Builds the active list {i : |itime - (t_i + dt_i)| < 2 DBL_EPSILON} on
 * the device and returns its size.  The list is a set: its order on the
 * device is unspecified (results do not depend on it); h4h_get_active
 * returns it sorted ascending as the reference's h_move is.
 * Replaces: Hermite4::find_particles_to_move (Hermite4.cpp:63-84). */
int h4h_find_particles_to_move(h4h_ctx *ctx, double itime, uint32_t *nact);

/* Replaces the active list with host indices (state injection). */
int h4h_set_active(h4h_ctx *ctx, const uint32_t *move, uint32_t nact);

/* Copies the active list back, sorted ascending; move needs n entries. */
int h4h_get_active(h4h_ctx *ctx, uint32_t *move, uint32_t *nact);

/* old_i = f_i for the active list.
 * Replaces: Hermite4::save_old_acc_jrk (Hermite4.cpp:145-153). */
int h4h_save_old_acc_jrk(h4h_ctx *ctx);

/* Taylor predictor of all n particles to itime.
 * Replaces: Hermite4CPU::predicted_pos_vel (cpu/Hermite4CPU.cpp:340-365)
 * and Hermite4GPU::predicted_pos_vel + k_prediction
 * (gpu/Hermite4GPU.cu:201-268, Hermite4GPU_kernels.cu:127-158). */
int h4h_predicted_pos_vel(h4h_ctx *ctx, double itime);

/* Acceleration and jerk of the nact active particles over all n.
 * nact must equal the size of the current active list.
 * Replaces: Hermite4CPU::update_acc_jrk (cpu/Hermite4CPU.cpp:309-336) and
 * Hermite4GPU::update_acc_jrk + k_update + k_reduce
 * (gpu/Hermite4GPU.cu:385-535, Hermite4GPU_kernels.cu:165-248). */
int h4h_update_acc_jrk(h4h_ctx *ctx, uint32_t nact);

/* Hermite corrector, Aarseth timestep and block quantisation of the
 * active particles; t_i = itime.
 * Replaces: Hermite4CPU::correction_pos_vel (cpu/Hermite4CPU.cpp:370-411)
 * with NbodyUtils::get_timestep_normal and normalize_dt
 * (utils/NbodyUtils.cpp:337-424). */
int h4h_correction_pos_vel(h4h_ctx *ctx, double itime, uint32_t nact);

/* *atime = min_i (t_i + dt_i).
 * Replaces: Hermite4::next_integration_time (Hermite4.cpp:90-102). */
int h4h_next_integration_time(h4h_ctx *ctx, double *atime);

/* One fused block step at itime: find active + save old + predict,
 * update, correct, next time.  Returns nact and the next time.
 * Replaces: the body of Hermite4CPU::integration's loop
 * (cpu/Hermite4CPU_integration.cpp:83-99). */
int h4h_step(h4h_ctx *ctx, double itime, uint32_t *nact, double *atime);

/* Result of h4h_run. */
typedef struct h4h_run_info {
    double   itime;         /* in: current block time (ITIME);
                               out: block time of the last step done       */
    double   atime;         /* in: next block time (ATIME); out: same      */
    uint64_t steps;         /* out: block steps done                       */
    uint64_t interactions;  /* out: sum of nact * n over those steps       */
    uint64_t nsteps;        /* out: sum of nact over those steps           */
    uint32_t last_nact;     /* out: nact of the last step                  */
    int      stopped;       /* out: H4H_STOP_* (why the loop stopped)      */
} h4h_run_info;

#define H4H_STOP_OUTPUT   1  /* a step with itime >= t_stop was done      */
#define H4H_STOP_END      2  /* itime >= t_end at the loop test           */
#define H4H_STOP_MAXSTEPS 3  /* max_steps steps were done                 */

/* Runs block steps on the device the way the reference's integration loop
 * does (cpu/Hermite4CPU_integration.cpp:81-124):
 *   while (ITIME < t_end) { ITIME = ATIME; step(ITIME); ... }
 * starting from io->itime / io->atime, until a step whose ITIME >= t_stop
 * has been done (the loop's output test, Hermite4CPU_integration.cpp:102,
 * with t_stop = interval_time * output_factor), or ITIME >= t_end, or
 * max_steps steps.  Steps are replayed as captured HIP graphs with no host
 * round trip per step.  step_time / step_nact (each max_steps long, or
 * NULL) receive ITIME and nact of every step. */
int h4h_run(h4h_ctx *ctx, double t_end, double t_stop, uint64_t max_steps,
            h4h_run_info *io, double *step_time, uint32_t *step_nact);

/* ---- Ensembles -------------------------------------------------------
 * One context can hold K independent clusters (each its own N-body system
 * with its own block times), integrated together: every device step
 * advances each live cluster by one block step, in the same launches.
 * The reference integrates one system per process; these calls have no
 * counterpart there and serve batched ensembles of uneven sizes. */

/* Per-cluster result of h4h_ensemble_run. */
typedef struct h4h_cluster_info {
    double   itime;         /* block time of the cluster's last step   */
    double   atime;         /* its next block time                     */
    uint64_t steps;         /* block steps done in this call           */
    uint64_t interactions;  /* sum of nact * n                         */
    uint64_t nsteps;        /* sum of nact                             */
    int      stopped;       /* H4H_STOP_*                              */
} h4h_cluster_info;

/* Context for nclusters clusters of sizes[c] particles.  Host arrays of
 * h4h_upload / h4h_download hold the clusters back to back in this order.
 * Other arguments as h4h_create (same e2, eta, bounds for all clusters). */
int h4h_create_ensemble(h4h_ctx **out, uint32_t nclusters, const uint32_t *sizes,
                        double e2, float eta, double dt_min, double dt_max,
                        int precision);

/* Number of clusters and their sizes (sizes may be NULL). */
int h4h_ensemble_sizes(h4h_ctx *ctx, uint32_t *nclusters, uint32_t *sizes);

/* Initial forces and timesteps of every cluster at t = 0, each cluster's
 * next block time its smallest dt (Hermite4CPU_integration.cpp:53-54). */
int h4h_ensemble_init(h4h_ctx *ctx, float eta_s);

/* Runs the reference's loop for every cluster at once (as h4h_run does for
 * one) until each cluster has done a step with ITIME >= t_stop, reached
 * ITIME >= t_end, or done max_steps steps.  info (nclusters entries, or
 * NULL) receives each cluster's state. */
int h4h_ensemble_run(h4h_ctx *ctx, double t_end, double t_stop, uint64_t max_steps,
                     h4h_cluster_info *info);

/* Kinetic and potential energy of every cluster (nclusters each). */
int h4h_ensemble_energy(h4h_ctx *ctx, double *ekin, double *epot);

/* Kinetic and (unsoftened) potential energy of the corrected state, in
 * FP64, within 1e-13 relative of the reference's sums (1/r from a refined
 * hardware rsqrt, compensated sums; coincident particles give -inf).
 * Replaces: NbodyUtils::get_energy (utils/NbodyUtils.cpp:427-529) and
 * Hermite4GPU::get_energy_gpu + k_energy (gpu/Hermite4GPU.cu:540-592,
 * Hermite4GPU_kernels.cu:250-288). */
int h4h_get_energy(h4h_ctx *ctx, double *ekin, double *epot);

/* Tuning table of the update kernel's work split, by bucket
 * b = floor(log2(nact)) of a single-system context: grid32[b] is the
 * number of workgroups the j split aims to fill (0 = the launch grid) and
 * iw32[b] the active particles per workgroup, 4/16/64/256 (0 = default
 * rule).  Either may be NULL; both NULL clears the table. */
int h4h_set_tuning(h4h_ctx *ctx, const uint32_t *grid32, const uint32_t *iw32);

/* Autotuner: times the update kernel on the context's own state for each
 * bucket (nact = 2^b, evenly spaced active particles) over the candidate
 * workgroup shares and lane-group widths, installs the fastest per bucket
 * and returns the table (grid32, iw32: 32 entries each, may be NULL).
 * Leaves the active list empty; call before the integration starts. */
int h4h_autotune(h4h_ctx *ctx, uint32_t reps, uint32_t *grid32, uint32_t *iw32);

/* Result of h4h_tune_launch. */
typedef struct h4h_launch_choice {
    uint32_t workgroups;   /* persistent workgroups of the update kernel     */
    uint32_t max_jsplit;   /* largest j split                                */
    uint32_t iw_fill16;    /* lane-group rule: widest i-slot count keeping at
                              least iw_fill16/16 of the slots busy           */
    float    ms_per_step;  /* device block step time of the chosen launch    */
} h4h_launch_choice;

/* Launch autotuner for ensembles (and single systems): times `steps`
 * device block steps (as h4h_ensemble_run) from the context's current
 * state for every candidate persistent-workgroup count x largest j split x
 * lane-group fill rule, restoring the state before each trial and at the
 * end, and installs the fastest.  Call after h4h_ensemble_init (or
 * h4h_init_dt).  Needs device memory for one copy of the state.  No
 * counterpart in the reference, which has no tunable launch. */
int h4h_tune_launch(h4h_ctx *ctx, uint32_t steps, h4h_launch_choice *choice);

/* Installs a launch choice (e.g. one h4h_tune_launch returned in an earlier
 * run); ms_per_step is ignored.  No counterpart in the reference. */
int h4h_set_launch_choice(h4h_ctx *ctx, const h4h_launch_choice *choice);

/* Small systems (of at most 1024 particles, FP64 or mixed precision):
 * h4h_step and the device loops of h4h_run and h4h_ensemble_run make every
 * block step whose active set has 1 to max_active particles in one
 * workgroup (each particle's state in one thread's registers, the predictor
 * records and the active list in LDS), up to 255 steps per launch; larger
 * steps go through the multi-workgroup kernels.  An ensemble whose clusters
 * all have at most 1024 particles gives each cluster a workgroup of its own
 * (one with a larger cluster keeps the multi-workgroup kernels for all:
 * the larger clusters would wait for the small ones' runs).  Default
 * max_active 64 for
 * one system, 256 for an ensemble; values above 256 act as 256; 0 turns the
 * path off.  Same algorithm, same block-step schedule; the pair sums are
 * added in another fixed order.  Ignored where the device cannot launch a
 * 1024-thread workgroup with 130 KiB of LDS.  No counterpart in the
 * reference (its GPU path launches kernels per step). */
int h4h_set_small_path(h4h_ctx *ctx, uint32_t max_active);

/* Enables HIP-event timing of update passes and steps (stats.update_ms,
 * stats.step_ms).  Off by default. */
int h4h_set_timing(h4h_ctx *ctx, int enable);

/* Device loops (h4h_run, h4h_ensemble_run) replay captured HIP graphs of
 * their step chunks (default, enable = 1) or enqueue the same kernels
 * directly (enable = 0: for tools that trace kernels one launch at a time).
 * Results are identical either way. */
int h4h_set_graphs(h4h_ctx *ctx, int enable);

/* Reads and optionally resets the counters. */
int h4h_get_stats(h4h_ctx *ctx, h4h_stats *out, int reset);

/* Blocks until all work queued on the context's stream is done. */
int h4h_synchronize(h4h_ctx *ctx);

/* ---- Update without a context and without allocation -----------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The call makes no hipMalloc, hipFree or
 * other allocating call; the work plan, work-queue counters, j-split
 * partial sums and the predictor and forces in the kernels' layouts live in
 * one static device array of H4H_STATIC_SCRATCH_BYTES that every GPU the
 * library runs on holds until the process ends.  Every call on a GPU
 * overwrites that array, so calls that overlap on one GPU (several host
 * threads, or streams that run at the same time) race. */
#define H4H_STATIC_MAX_N          65536u     /* largest n of the call      */
#define H4H_STATIC_SCRATCH_BYTES  58130688u  /* the static array, per GPU  */

/* h4h_update_acc_jrk on device arrays of the current device, with no
 * context: acceleration and jerk of the nact particles d_move[0..nact)
 * over all n particles, j != i.
 *   d_p6   : n x 6 double, predicted r[3], v[3] (Predictor, common.hpp:213-217)
 *   d_pm   : n float, Predictor::m
 *   d_move : nact uint32 indices < n
 *   d_f6   : n x 6 double (Forces, common.hpp:228-231): rows d_move[k]
 *            receive a[3], a1[3]; other rows are left as they are
 * e2 and precision as h4h_create; n <= H4H_STATIC_MAX_N, nact <= n (larger
 * sizes are refused with H4H_EINVAL).  The work is queued on `stream` (a
 * hipStream_t; NULL is the null stream) and not waited for.  The results
 * are bit for bit those of h4h_upload_predictor + h4h_set_active +
 * h4h_update_acc_jrk on a context of n particles with the default launch.
 * An index >= n in d_move is caught on the device: every row d_move[k]
 * < n of d_f6 is then NaN.  The first call on a device looks up the static
 * array's address there; a call with nact = 0 does only that, so a caller
 * that captures graphs can make it before capturing. */
int h4h_update_acc_jrk_static(const double *d_p6, const float *d_pm, uint32_t n,
                              const uint32_t *d_move, uint32_t nact, double e2,
                              int precision, double *d_f6, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* H4HIP_H */
