/*
 * sanitize_check.c -- TEST INFRASTRUCTURE ONLY.
 *
 * h4o-check: drives every function of the oracle (h4_oracle.c) on small
 * systems and on degenerate inputs, for the CPU sanitizer build
 * (AddressSanitizer + UndefinedBehaviorSanitizer, `make -C oracle
 * sanitize`).  It checks memory safety and the absence of undefined
 * behaviour (out-of-range float-to-int casts included), not numbers: those
 * are pinned by tests/test_oracle.py.  Exit status 0 when every case ran.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double *step_time;
    uint32_t *step_nact;
    double *energy_time;
    double *energy;
    uint64_t max_steps;
    uint32_t max_energy;
    uint32_t n_energy;
    uint64_t interactions;
    uint64_t nsteps;
} h4o_record;

void h4o_init_acc_jrk(uint32_t n, const double *p, const float *pm, double e2, double *f);
uint32_t h4o_find_particles_to_move(uint32_t n, double itime, const double *t,
                                    const double *dt, uint32_t *move);
double h4o_next_integration_time(uint32_t n, const double *t, const double *dt);
double h4o_init_dt_value(double tmp_dt, double dt_min, double dt_max);
double h4o_normalize_dt(double new_dt, double old_dt, double t, double dt_min, double dt_max);
void h4o_get_energy(uint32_t n, const double *r, const double *v, double *ekin, double *epot);
int64_t h4o_integrate_at(uint32_t n, double *r, double *v, double *f, double *t, double *dt,
                         double *a2, double *a3, double e2, float eta_s, float eta_n,
                         double itime0, float integration_time, float interval_time,
                         double dt_min, double dt_max, h4o_record *rec);

#define DT_MIN 1.1920928955078125e-07
#define DT_MAX 0.125

static uint64_t lcg_state = 12345u;

static double uniform(void)
{
    lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(lcg_state >> 11) * (1.0 / 9007199254740992.0);
}

/* n particles of mass 1/n in the unit cube, small random velocities; a
 * whole run to t_end from itime0, with outputs every t_end / 4. */
static int run(uint32_t n, double e2, double itime0, float t_end)
{
    double *r = calloc(4 * (size_t)n, sizeof(double));
    double *v = calloc(4 * (size_t)n, sizeof(double));
    double *f = calloc(6 * (size_t)n, sizeof(double));
    double *t = calloc(n, sizeof(double));
    double *dt = calloc(n, sizeof(double));
    double *a2 = calloc(3 * (size_t)n, sizeof(double));
    double *a3 = calloc(3 * (size_t)n, sizeof(double));
    const uint64_t max_steps = 200000;
    h4o_record rec;
    uint32_t i;
    int64_t steps;
    int ok;
    memset(&rec, 0, sizeof(rec));
    rec.step_time = calloc(max_steps, sizeof(double));
    rec.step_nact = calloc(max_steps, sizeof(uint32_t));
    rec.energy_time = calloc(64, sizeof(double));
    rec.energy = calloc(64, sizeof(double));
    rec.max_steps = max_steps;
    rec.max_energy = 64;
    for (i = 0; i < n; i++) {
        int k;
        for (k = 0; k < 3; k++) {
            r[4 * (size_t)i + k] = uniform() - 0.5;
            v[4 * (size_t)i + k] = 0.1 * (uniform() - 0.5);
        }
        r[4 * (size_t)i + 3] = 1.0 / n;
    }
    steps = h4o_integrate_at(n, r, v, f, t, dt, a2, a3, e2, 0.01f, 0.01f, itime0,
                             t_end, t_end / 4, DT_MIN, DT_MAX, &rec);
    ok = steps > 0 && rec.n_energy >= 1;
    printf("run n=%u e2=%g itime0=%g: %lld block steps, %u energies, E0 %.6e\n", n, e2,
           itime0, (long long)steps, rec.n_energy, rec.n_energy ? rec.energy[0] : 0.0);
    free(r); free(v); free(f); free(t); free(dt); free(a2); free(a3);
    free(rec.step_time); free(rec.step_nact); free(rec.energy_time); free(rec.energy);
    return ok;
}

int main(void)
{
    static const uint32_t sizes[] = {1, 2, 3, 64, 257};
    const double special[] = {0.0, -1.0, 1e-300, 5e-324, 1e300, INFINITY, -INFINITY, NAN,
                              0.0009765625, 0.125, 1.0};
    size_t k;
    int ok = 1;

    /* one particle has no force: its steps stay at the smallest one */
    ok &= run(1, 9.99999905104687e-09, 0.0, 0x1p-19f);
    for (k = 1; k < sizeof(sizes) / sizeof(sizes[0]); k++)
        ok &= run(sizes[k], 9.99999905104687e-09, 0.0, 0.0625f);
    ok &= run(64, 0.0, 0.0, 0.03125f);        /* unsoftened */
    ok &= run(64, 9.99999905104687e-09, 1.0, 1.0625f);  /* a resumed run */

    /* init_dt and normalize_dt over degenerate values: no undefined cast */
    for (k = 0; k < sizeof(special) / sizeof(special[0]); k++) {
        size_t q;
        double d = h4o_init_dt_value(special[k], DT_MIN, DT_MAX);
        ok &= d >= DT_MIN && d <= DT_MAX;
        for (q = 0; q < sizeof(special) / sizeof(special[0]); q++) {
            double nd = h4o_normalize_dt(special[k], special[q], 0.5, DT_MIN, DT_MAX);
            (void)nd;
        }
    }

    /* coincident particles: -inf potential (unsoftened energy), and the
     * pair sums with e2 = 0 stay defined behaviour */
    {
        double r[8] = {0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5};
        double v[8] = {0};
        double p[12] = {0.5, 0.5, 0.5, 0, 0, 0, 0.5, 0.5, 0.5, 0, 0, 0};
        float pm[2] = {0.5f, 0.5f};
        double f[12];
        double ek, ep, tt[2] = {0, 0}, dd[2] = {INFINITY, 0.125};
        uint32_t move[2];
        h4o_get_energy(2, r, v, &ek, &ep);
        ok &= ep == -INFINITY && ek == 0.0;
        h4o_init_acc_jrk(2, p, pm, 0.0, f);
        ok &= h4o_find_particles_to_move(2, 0.125, tt, dd, move) == 1 && move[0] == 1;
        ok &= h4o_next_integration_time(2, tt, dd) == 0.125;
    }
    printf(ok ? "h4o-check ok\n" : "h4o-check FAILED\n");
    return ok ? 0 : 1;
}
