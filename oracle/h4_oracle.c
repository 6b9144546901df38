/*
 * h4_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of GraviDy's FP64 Hermite-4 block-timestep path
 * (class Hermite4CPU plus the Hermite4 / NbodyUtils helpers it calls).
 * It is the correctness oracle for the HIP backend and the CPU baseline
 * that bench.py times.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library; the product path
 * (gravidy_amd, libh4hip.so, gravidy-hip) never links or calls it.
 *
 * Arithmetic follows the reference expression trees operation by
 * operation (same association, same float-typed constants, sequential
 * ascending j sums).  It must be compiled with -ffp-contract=off and
 * without -ffast-math / -march so no FMA contraction or reassociation
 * happens, which is how the reference's Makefile builds on x86-64
 * (src/Makefile:40-54: -O3 -fopenmp, no -march).
 *
 * Data layouts (row-major, contiguous), same bytes as the reference types:
 *   r   : n x 4 double  (x, y, z, m)     NbodySystem::h_r  (double4)
 *   v   : n x 4 double  (vx, vy, vz, w)  NbodySystem::h_v  (double4)
 *   f   : n x 6 double  (a[3], a1[3])    Forces            (common.hpp:228-231)
 *   p   : n x 6 double  (r[3], v[3])     Predictor without its mass
 *   pm  : n float                         Predictor::m      (common.hpp:213-217)
 *   a2, a3 : n x 3 double                 NbodySystem::h_a2 / h_a3 (double3)
 *
 * Parity pins (tests/test_oracle.py): closed-form single pair, Newton's
 * third law, a = -grad(phi) and jerk = da/dt by finite differences,
 * 40-digit Decimal brute-force sums, exact Taylor predictor on a cubic,
 * exact Hermite corrector on a quintic trajectory, hand-computed Aarseth
 * step, the normalize_dt branch table, the init_dt rounding table, a hand
 * schedule for find/next-time, two-body and cube-lattice energies, and
 * Kepler / Plummer integrations.  No function here is "parity unpinned".
 * The reference itself cannot be compiled in this image (every
 * translation unit includes boost/program_options.hpp, which is absent),
 * so no oracle/_ref build exists.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* common.hpp:122 and :129 -- compile-time bounds in the reference; here
 * they are arguments so dt_min = 2^-30 configurations can be expressed. */
#define H4O_D_TIME_MIN 1.1920928955078125e-07
#define H4O_D_TIME_MAX 0.125

/* Hermite4CPU::force_calculation, src/include/cpu/Hermite4CPU.cpp:55-82.
 * pi, pj: 6 doubles (r[3], v[3]); mj: Predictor::m (float); fi += pair. */
void h4o_force_calculation(const double *pi, const double *pj, float mj,
                           double e2, double *fi)
{
    double rx = pj[0] - pi[0];
    double ry = pj[1] - pi[1];
    double rz = pj[2] - pi[2];

    double vx = pj[3] - pi[3];
    double vy = pj[4] - pi[4];
    double vz = pj[5] - pi[5];

    double r2     = rx*rx + ry*ry + rz*rz + e2;
    double rinv   = 1.0/sqrt(r2);
    double r2inv  = rinv  * rinv;
    double r3inv  = r2inv * rinv;
    double r5inv  = r2inv * r3inv;
    double mr3inv = r3inv * mj;
    double mr5inv = r5inv * mj;

    double rv = rx*vx + ry*vy + rz*vz;

    fi[0] += (rx * mr3inv);
    fi[1] += (ry * mr3inv);
    fi[2] += (rz * mr3inv);

    fi[3] += (vx * mr3inv - (3 * rv) * rx * mr5inv);
    fi[4] += (vy * mr3inv - (3 * rv) * ry * mr5inv);
    fi[5] += (vz * mr3inv - (3 * rv) * rz * mr5inv);
}

/* Hermite4CPU::init_acc_jrk, Hermite4CPU.cpp:289-304.  f is summed into
 * (the reference relies on Hermite4::init_data zeroing it, Hermite4.cpp:
 * 194-240); each i sums j = 0..n-1 in order, skipping j == i. */
void h4o_init_acc_jrk(uint32_t n, const double *p, const float *pm,
                      double e2, double *f)
{
    int64_t i;
    #pragma omp parallel for schedule(dynamic, 16)
    for (i = 0; i < (int64_t)n; i++) {
        uint32_t j;
        for (j = 0; j < n; j++) {
            if ((uint32_t)i == j) continue;
            h4o_force_calculation(&p[6*i], &p[6*(size_t)j], pm[j], e2,
                                  &f[6*i]);
        }
    }
}

/* Hermite4CPU::update_acc_jrk, Hermite4CPU.cpp:309-336.  The nested inner
 * "omp parallel for" of the reference is inert (nesting disabled), so each
 * active i is a sequential ascending-j sum, zeroed first. */
void h4o_update_acc_jrk(uint32_t n, uint32_t nact, const uint32_t *move,
                        const double *p, const float *pm, double e2,
                        double *f)
{
    int64_t k;
    #pragma omp parallel for schedule(dynamic, 4)
    for (k = 0; k < (int64_t)nact; k++) {
        uint32_t i = move[k];
        uint32_t j;
        double *fi = &f[6*(size_t)i];
        fi[0] = 0.0; fi[1] = 0.0; fi[2] = 0.0;
        fi[3] = 0.0; fi[4] = 0.0; fi[5] = 0.0;
        for (j = 0; j < n; j++) {
            if (i == j) continue;
            h4o_force_calculation(&p[6*(size_t)i], &p[6*(size_t)j], pm[j],
                                  e2, fi);
        }
    }
}

/* Hermite4CPU::predicted_pos_vel, Hermite4CPU.cpp:340-365 (all n). */
void h4o_predicted_pos_vel(uint32_t n, double itime, const double *t,
                           const double *r, const double *v,
                           const double *f, double *p, float *pm)
{
    uint32_t i;
    for (i = 0; i < n; i++) {
        const double *ri = &r[4*(size_t)i];
        const double *vi = &v[4*(size_t)i];
        const double *fi = &f[6*(size_t)i];
        double *pi = &p[6*(size_t)i];

        double dt  = itime - t[i];
        double dt2 = (dt  * dt);
        double dt3 = (dt2 * dt);

        double dt2c = dt2/2.0;
        double dt3c = dt3/6.0;

        pi[0] = (dt3c * fi[3]) + (dt2c * fi[0]) + (dt * vi[0]) + ri[0];
        pi[1] = (dt3c * fi[4]) + (dt2c * fi[1]) + (dt * vi[1]) + ri[1];
        pi[2] = (dt3c * fi[5]) + (dt2c * fi[2]) + (dt * vi[2]) + ri[2];

        pi[3] = (dt2c * fi[3]) + (dt * fi[0]) + vi[0];
        pi[4] = (dt2c * fi[4]) + (dt * fi[1]) + vi[1];
        pi[5] = (dt2c * fi[5]) + (dt * fi[2]) + vi[2];

        pm[i] = (float)ri[3];
    }
}

/* Hermite4::find_particles_to_move, Hermite4.cpp:63-84: ascending i with
 * |ITIME - (t_i + dt_i)| < 2 * DBL_EPSILON.  (The max_mass side effect of
 * the reference is unused and omitted.) */
uint32_t h4o_find_particles_to_move(uint32_t n, double itime,
                                    const double *t, const double *dt,
                                    uint32_t *move)
{
    uint32_t i, j = 0;
    for (i = 0; i < n; i++) {
        double tmp_time = t[i] + dt[i];
        if (fabs(itime - tmp_time) < 2*DBL_EPSILON) {
            move[j] = i;
            j++;
        }
    }
    return j;
}

/* Hermite4::next_integration_time, Hermite4.cpp:90-102. */
double h4o_next_integration_time(uint32_t n, const double *t,
                                 const double *dt)
{
    double ctime = t[0] + dt[0];
    uint32_t i;
    for (i = 1; i < n; i++) {
        double time = t[i] + dt[i];
        if (time < ctime) ctime = time;
    }
    return ctime;
}

/* NbodyUtils::get_magnitude, NbodyUtils.cpp:331-334. */
static double h4o_magnitude(double x, double y, double z)
{
    return sqrt(x*x + y*y + z*z);
}

/* Exponent step of Hermite4::init_dt, Hermite4.cpp:117-131, for one value
 * tmp_dt = ETA*(|a|/|j|).  The reference casts ceil(log(x)/log(2))-1 to
 * int; for x <= 0, inf or NaN that cast is undefined (x86 yields INT_MIN,
 * pow(2, INT_MIN) = 0, clamped to D_TIME_MIN).  That outcome is made
 * explicit here instead of relying on the undefined cast. */
double h4o_init_dt_value(double tmp_dt, double dt_min, double dt_max)
{
    if (!(tmp_dt > 0.0) || isinf(tmp_dt))
        return dt_min;
    {
        int e = (int)(ceil(log(tmp_dt)/log(2.0))-1);
        double d = pow(2, e);
        if (d < dt_min)
            d = dt_min;
        else if (d > dt_max)
            d = dt_max;
        return d;
    }
}

/* Hermite4::init_dt, Hermite4.cpp:108-140.  ctime follows the reference:
 * it is lowered to the smallest dt (not t+dt). */
void h4o_init_dt(uint32_t n, double *ctime, float eta, double itime,
                 const double *f, double *t, double *dt, double dt_min,
                 double dt_max)
{
    uint32_t i;
    for (i = 0; i < n; i++) {
        const double *fi = &f[6*(size_t)i];
        double a2 = h4o_magnitude(fi[0], fi[1], fi[2]);
        double j2 = h4o_magnitude(fi[3], fi[4], fi[5]);
        double tmp_dt = eta * (a2/j2);
        tmp_dt = h4o_init_dt_value(tmp_dt, dt_min, dt_max);
        dt[i] = tmp_dt;
        t[i] = itime;
        if (tmp_dt < *ctime) *ctime = tmp_dt;
    }
}

/* Hermite4::save_old_acc_jrk, Hermite4.cpp:145-153. */
void h4o_save_old_acc_jrk(uint32_t nact, const uint32_t *move,
                          const double *f, double *old)
{
    uint32_t k;
    for (k = 0; k < nact; k++) {
        uint32_t i = move[k];
        memcpy(&old[6*(size_t)i], &f[6*(size_t)i], 6*sizeof(double));
    }
}

/* NbodyUtils::get_timestep_normal, NbodyUtils.cpp:337-365.  a2i, a3i are
 * the particle's h_a2/h_a3, dti its (old) h_dt, fi its new Forces. */
double h4o_get_timestep_normal(const double *a2i, const double *a3i,
                               double dti, const double *fi, float eta)
{
    double ax1_2 = a2i[0] + dti * a3i[0];
    double ay1_2 = a2i[1] + dti * a3i[1];
    double az1_2 = a2i[2] + dti * a3i[2];

    double abs_a1 = h4o_magnitude(fi[0], fi[1], fi[2]);
    double abs_j1 = h4o_magnitude(fi[3], fi[4], fi[5]);
    double abs_j12 = abs_j1 * abs_j1;
    double abs_a1_3 = h4o_magnitude(a3i[0], a3i[1], a3i[2]);
    double abs_a1_2 = h4o_magnitude(ax1_2, ay1_2, az1_2);
    double abs_a1_22 = abs_a1_2 * abs_a1_2;

    double normal_dt = sqrt(eta * ((abs_a1 * abs_a1_2 + abs_j12) /
                                   (abs_j1 * abs_a1_3 + abs_a1_22)));
    return normal_dt;
}

/* NbodyUtils::normalize_dt, NbodyUtils.cpp:371-424, with D_TIME_MIN and
 * D_TIME_MAX (common.hpp:122,129) passed in. */
double h4o_normalize_dt(double new_dt, double old_dt, double t,
                        double dt_min, double dt_max)
{
    if (new_dt <= old_dt/8) {
        new_dt = dt_min;
    } else if (old_dt/8 < new_dt && new_dt <= old_dt/4) {
        new_dt = old_dt / 8;
    } else if (old_dt/4 < new_dt && new_dt <= old_dt/2) {
        new_dt = old_dt / 4;
    } else if (old_dt/2 < new_dt && new_dt <= old_dt) {
        new_dt = old_dt / 2;
    } else if (old_dt < new_dt && new_dt <= old_dt * 2) {
        new_dt = old_dt;
    } else if (2 * old_dt < new_dt) {
        double val = t/(2 * old_dt);
        if (ceil(val) == val)
            new_dt = 2.0 * old_dt;
        else
            new_dt = old_dt;
    } else {
        new_dt = old_dt;
    }

    if (new_dt < dt_min)
        new_dt = dt_min;
    else if (new_dt > dt_max)
        new_dt = dt_max;
    return new_dt;
}

/* Hermite4CPU::correction_pos_vel, Hermite4CPU.cpp:370-411. */
void h4o_correction_pos_vel(uint32_t n, double itime, uint32_t nact,
                            const uint32_t *move, const double *p,
                            const double *f, const double *old,
                            double *a2, double *a3, double *r, double *v,
                            double *t, double *dt, float eta,
                            double dt_min, double dt_max)
{
    uint32_t k;
    (void)n;
    for (k = 0; k < nact; k++) {
        uint32_t i = move[k];
        const double *fi = &f[6*(size_t)i];
        const double *oi = &old[6*(size_t)i];
        const double *pi = &p[6*(size_t)i];
        double *a2i = &a2[3*(size_t)i];
        double *a3i = &a3[3*(size_t)i];
        double *ri = &r[4*(size_t)i];
        double *vi = &v[4*(size_t)i];
        int c;

        double dt1 = dt[i];
        double dt2 = dt1 * dt1;
        double dt3 = dt2 * dt1;
        double dt4 = dt2 * dt2;
        double dt5 = dt4 * dt1;
        double normal_dt;

        for (c = 0; c < 3; c++)
            a2i[c] = (-6 * (oi[c] - fi[c]) - dt1 * (4 * oi[3+c] + 2 * fi[3+c])) / dt2;
        for (c = 0; c < 3; c++)
            a3i[c] = (12 * (oi[c] - fi[c]) + 6 * dt1 * (oi[3+c] + fi[3+c])) / dt3;
        for (c = 0; c < 3; c++)
            ri[c] = pi[c] + (dt4/24)*a2i[c] + (dt5/120)*a3i[c];
        for (c = 0; c < 3; c++)
            vi[c] = pi[3+c] + (dt3/6)*a2i[c] + (dt4/24)*a3i[c];

        t[i] = itime;
        normal_dt = h4o_get_timestep_normal(a2i, a3i, dt[i], fi, eta);
        normal_dt = h4o_normalize_dt(normal_dt, dt[i], t[i], dt_min, dt_max);
        dt[i] = normal_dt;
    }
}

/* NbodyUtils::get_energy, NbodyUtils.cpp:427-529 (PN terms excluded, as
 * in the default build).  Unsoftened potential over j > i.  The reference
 * adds the per-i partials with "omp atomic" in thread-completion order;
 * here they are added in ascending i, i.e. the one-thread order. */
void h4o_get_energy(uint32_t n, const double *r, const double *v,
                    double *ekin_out, double *epot_out)
{
    double *epot_i = (double *)malloc(sizeof(double) * (n ? n : 1));
    double ekin = 0.0, epot = 0.0;
    int64_t i;
    #pragma omp parallel for schedule(dynamic, 16)
    for (i = 0; i < (int64_t)n; i++) {
        double epot_tmp = 0.0;
        uint32_t j;
        const double *ri = &r[4*i];
        for (j = (uint32_t)i + 1; j < n; j++) {
            const double *rj = &r[4*(size_t)j];
            double rx = rj[0] - ri[0];
            double ry = rj[1] - ri[1];
            double rz = rj[2] - ri[2];
            double r2 = rx*rx + ry*ry + rz*rz;
            epot_tmp -= (ri[3] * rj[3]) / sqrt(r2);
        }
        epot_i[i] = epot_tmp;
    }
    for (i = 0; i < (int64_t)n; i++) {
        const double *vi = &v[4*i];
        double vvx = vi[0] * vi[0];
        double vvy = vi[1] * vi[1];
        double vvz = vi[2] * vi[2];
        double vv2 = vvx + vvy + vvz;
        double ekin_tmp = 0.5 * r[4*i+3] * vv2;
        ekin += ekin_tmp;
        epot += epot_i[i];
    }
    free(epot_i);
    *ekin_out = ekin;
    *epot_out = epot;
}

/* Hermite4CPU::integration, Hermite4CPU_integration.cpp:38-130, without
 * the Logger/NbodyUtils output side effects.  Records, per block step,
 * ITIME and nact; and at t = 0 and every output interval, ITIME and the
 * total energy.  integration_time and interval_time are float in the
 * reference (NbodySystem.hpp:69-70); they are passed as float here and
 * the interval test "ITIME >= interval_time * output_factor" is evaluated
 * in float exactly as there (float * unsigned -> float).
 *
 * r, v are updated in place; f, t, dt, a2, a3 receive the final state.
 * Returns the number of block steps taken, or -1 if max_steps was hit
 * first (the recorded prefix is valid).
 *
 * h4o_integrate_at starts at ITIME = itime0 instead of 0: the reference's
 * resumed run (ITIME = ns->snapshot_time, Hermite4CPU_integration.cpp:43),
 * which re-initialises forces and steps at that time (init_dt sets
 * t_i = ITIME and leaves ATIME at the smallest dt, so the first block step
 * of a resumed run, at that ATIME, has no active particle) and counts
 * output times from output_factor = 1 again. */
typedef struct {
    double *step_time;      /* [max_steps] */
    uint32_t *step_nact;    /* [max_steps] */
    double *energy_time;    /* [max_energy] */
    double *energy;         /* [max_energy] */
    uint64_t max_steps;
    uint32_t max_energy;
    uint32_t n_energy;      /* out */
    uint64_t interactions;  /* out: sum of nact*n in 64 bits */
    uint64_t nsteps;        /* out: sum of nact */
} h4o_record;

int64_t h4o_integrate_at(uint32_t n, double *r, double *v, double *f,
                         double *t, double *dt, double *a2, double *a3,
                         double e2, float eta_s, float eta_n, double itime0,
                         float integration_time, float interval_time,
                         double dt_min, double dt_max, h4o_record *rec)
{
    double *p = (double *)calloc((size_t)n * 6, sizeof(double));
    float *pm = (float *)calloc(n ? n : 1, sizeof(float));
    double *old = (double *)calloc((size_t)n * 6, sizeof(double));
    uint32_t *move = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    double atime = 1.0e+10;
    double itime = itime0;
    unsigned int output_factor = 1;
    int64_t steps = 0;
    double ekin, epot;
    uint32_t i;

    /* Hermite4::init_data, Hermite4.cpp:194-240 */
    for (i = 0; i < n; i++) {
        memcpy(&p[6*(size_t)i], &r[4*(size_t)i], 3*sizeof(double));
        memcpy(&p[6*(size_t)i+3], &v[4*(size_t)i], 3*sizeof(double));
        pm[i] = (float)r[4*(size_t)i+3];
    }
    memset(f, 0, sizeof(double) * 6 * (size_t)n);
    memset(a2, 0, sizeof(double) * 3 * (size_t)n);
    memset(a3, 0, sizeof(double) * 3 * (size_t)n);

    rec->n_energy = 0;
    rec->interactions = 0;
    rec->nsteps = 0;

    h4o_init_acc_jrk(n, p, pm, e2, f);
    h4o_init_dt(n, &atime, eta_s, itime, f, t, dt, dt_min, dt_max);

    h4o_get_energy(n, r, v, &ekin, &epot);
    if (rec->n_energy < rec->max_energy) {
        rec->energy_time[rec->n_energy] = itime;
        rec->energy[rec->n_energy] = ekin + epot;
        rec->n_energy++;
    }

    while (itime < integration_time) {
        uint32_t nact;
        if ((uint64_t)steps >= rec->max_steps) {
            steps = -1;
            break;
        }
        itime = atime;
        nact = h4o_find_particles_to_move(n, itime, t, dt, move);
        h4o_save_old_acc_jrk(nact, move, f, old);
        h4o_predicted_pos_vel(n, itime, t, r, v, f, p, pm);
        h4o_update_acc_jrk(n, nact, move, p, pm, e2, f);
        h4o_correction_pos_vel(n, itime, nact, move, p, f, old, a2, a3,
                               r, v, t, dt, eta_n, dt_min, dt_max);
        rec->interactions += (uint64_t)nact * n;
        atime = h4o_next_integration_time(n, t, dt);

        rec->step_time[steps] = itime;
        rec->step_nact[steps] = nact;
        steps++;

        if (itime >= (double)(interval_time * (float)output_factor)) {
            h4o_get_energy(n, r, v, &ekin, &epot);
            if (rec->n_energy < rec->max_energy) {
                rec->energy_time[rec->n_energy] = itime;
                rec->energy[rec->n_energy] = ekin + epot;
                rec->n_energy++;
            }
            output_factor += 1;
        }
        rec->nsteps += nact;
    }

    free(p);
    free(pm);
    free(old);
    free(move);
    return steps;
}

/* The body of Hermite4CPU::integration's loop
 * (Hermite4CPU_integration.cpp:83-99) for `steps` block steps from a given
 * state: ITIME = *atime; find, save old, predict, update, correct; ATIME =
 * the next block time.  r, v, f, t, dt, a2, a3 are updated in place and
 * *atime returns the next block time.  Records ITIME and nact of each step
 * (rec->max_steps >= steps entries; energies are not computed).  Used to
 * time the CPU path over the same block steps as the GPU from a state
 * whose initial forces were computed elsewhere. */
int64_t h4o_block_steps(uint32_t n, double *r, double *v, double *f,
                        double *t, double *dt, double *a2, double *a3,
                        double e2, float eta_n, double *atime, uint64_t steps,
                        double dt_min, double dt_max, h4o_record *rec)
{
    double *p = (double *)calloc((size_t)n * 6, sizeof(double));
    float *pm = (float *)calloc(n ? n : 1, sizeof(float));
    double *old = (double *)calloc((size_t)n * 6, sizeof(double));
    uint32_t *move = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint64_t k;
    rec->n_energy = 0;
    rec->interactions = 0;
    rec->nsteps = 0;
    if (steps > rec->max_steps)
        steps = rec->max_steps;
    for (k = 0; k < steps; k++) {
        double itime = *atime;
        uint32_t nact = h4o_find_particles_to_move(n, itime, t, dt, move);
        h4o_save_old_acc_jrk(nact, move, f, old);
        h4o_predicted_pos_vel(n, itime, t, r, v, f, p, pm);
        h4o_update_acc_jrk(n, nact, move, p, pm, e2, f);
        h4o_correction_pos_vel(n, itime, nact, move, p, f, old, a2, a3,
                               r, v, t, dt, eta_n, dt_min, dt_max);
        rec->interactions += (uint64_t)nact * n;
        rec->nsteps += nact;
        *atime = h4o_next_integration_time(n, t, dt);
        rec->step_time[k] = itime;
        rec->step_nact[k] = nact;
    }
    free(p);
    free(pm);
    free(old);
    free(move);
    return (int64_t)steps;
}

int64_t h4o_integrate(uint32_t n, double *r, double *v, double *f,
                      double *t, double *dt, double *a2, double *a3,
                      double e2, float eta_s, float eta_n,
                      float integration_time, float interval_time,
                      double dt_min, double dt_max, h4o_record *rec)
{
    return h4o_integrate_at(n, r, v, f, t, dt, a2, a3, e2, eta_s, eta_n, 0.0,
                            integration_time, interval_time, dt_min, dt_max, rec);
}

/* Number of OpenMP threads the parallel loops above use (1 without
 * OpenMP); reported next to CPU baseline timings. */
int h4o_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
