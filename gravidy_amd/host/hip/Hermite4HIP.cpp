// Hermite4HIP.cpp -- the Hermite4 backend over libh4hip.so.
#include "Hermite4HIP.hpp"

#include <cstdlib>

Hermite4HIP::Hermite4HIP(NbodySystem *ns, Logger *logger, NbodyUtils *nu, int precision,
                         double dt_min, double dt_max)
    : Hermite4(ns, logger, nu), ctx(NULL), precision_(precision), dt_min_(dt_min),
      dt_max_(dt_max) {
    check(h4h_create(&ctx, ns->n, ns->e2, ns->eta, dt_min, dt_max, precision), "h4h_create");
    const char *names[] = {"fp64", "exact", "mixed"};
    logger->log_info(std::string("HIP backend: ") + SSTR(ns->n) + " particles, precision " +
                     (precision >= 0 && precision <= 2 ? names[precision] : "?"));
    upload();
}

Hermite4HIP::~Hermite4HIP() { h4h_destroy(ctx); }

// ABI errors end the run with a logged message, as the reference's CSC
// macro does for CUDA errors (gpu/Hermite4GPU.cuh:43-50).
void Hermite4HIP::check(int rc, const char *what) {
    if (rc == H4H_OK) return;
    logger->log_error(std::string(what) + " failed (" + SSTR(rc) + "): " + h4h_last_error());
    std::exit(EXIT_FAILURE);
}

void Hermite4HIP::upload() {
    check(h4h_upload(ctx, (const double *)ns->h_r, (const double *)ns->h_v,
                     (const double *)ns->h_f, ns->h_t, ns->h_dt),
          "h4h_upload");
}

void Hermite4HIP::download() {
    check(h4h_download(ctx, (double *)ns->h_r, (double *)ns->h_v, (double *)ns->h_f, ns->h_t,
                       ns->h_dt, (double *)ns->h_old, (double *)ns->h_a2, (double *)ns->h_a3),
          "h4h_download");
}

void Hermite4HIP::init_acc_jrk() { check(h4h_init_acc_jrk(ctx), "h4h_init_acc_jrk"); }

void Hermite4HIP::init_dt(double &ATIME, float ETA, double ITIME) {
    check(h4h_init_dt(ctx, &ATIME, ETA, ITIME), "h4h_init_dt");
}

unsigned int Hermite4HIP::find_particles_to_move(double ITIME) {
    uint32_t nact = 0;
    check(h4h_find_particles_to_move(ctx, ITIME, &nact), "h4h_find_particles_to_move");
    return nact;
}

void Hermite4HIP::save_old_acc_jrk(unsigned int) {
    check(h4h_save_old_acc_jrk(ctx), "h4h_save_old_acc_jrk");
}

void Hermite4HIP::predicted_pos_vel(double ITIME) {
    check(h4h_predicted_pos_vel(ctx, ITIME), "h4h_predicted_pos_vel");
}

void Hermite4HIP::update_acc_jrk(unsigned int nact) {
    ns->gtime.update_ini = omp_get_wtime();
    check(h4h_update_acc_jrk(ctx, nact), "h4h_update_acc_jrk");
    ns->gtime.update_end += omp_get_wtime() - ns->gtime.update_ini;
}

void Hermite4HIP::correction_pos_vel(double ITIME, unsigned int nact) {
    check(h4h_correction_pos_vel(ctx, ITIME, nact), "h4h_correction_pos_vel");
}

void Hermite4HIP::next_integration_time(double &ATIME) {
    check(h4h_next_integration_time(ctx, &ATIME), "h4h_next_integration_time");
}

double Hermite4HIP::get_energy_gpu() {
    double ekin = 0.0, epot = 0.0;
    check(h4h_get_energy(ctx, &ekin, &epot), "h4h_get_energy");
    ns->en.kinetic = ekin;
    ns->en.potential = epot;
    return ekin + epot;
}

unsigned int Hermite4HIP::step(double ITIME, double &ATIME) {
    uint32_t nact = 0;
    ns->gtime.update_ini = omp_get_wtime();
    check(h4h_step(ctx, ITIME, &nact, &ATIME), "h4h_step");
    ns->gtime.update_end += omp_get_wtime() - ns->gtime.update_ini;
    return nact;
}

// Same order of operations and output points as the reference loop; the
// per-step methods are one fused device step replayed on the device, and
// the host arrays are refreshed only where the output layer reads them.
// The interaction counter is 64-bit (the reference multiplies two 32-bit
// values, Hermite4CPU_integration.cpp:96).  nbody_attributes() is not
// called: its results are written nowhere.
void Hermite4HIP::integration() {
    ns->gtime.integration_ini = omp_get_wtime();

    double ATIME = 1.0e+10;
    double ITIME = ns->snapshot_time;
    unsigned int nact = 0;
    unsigned int nsteps = 0;
    long long interactions = 0;
    unsigned int output_factor = 1;

    unsigned int snap_number = ns->snapshot_number;
    if (!restart_path.empty()) {
        // Continue a checkpointed run: state, block times and counters as
        // they were at the top of the loop after that output.
        read_restart(ITIME, ATIME, interactions, nsteps, output_factor, snap_number);
        // RelE of the next output is relative to the energy at this time.
        ns->en.tmp = get_energy_gpu();
        logger->print_info();
        logger->write_info();
    } else {
        init_acc_jrk();
        init_dt(ATIME, ETA_S, ITIME);

        ns->en.ini = get_energy_gpu();
        ns->en.tmp = ns->en.ini;

        logger->print_info();
        logger->write_info();
        logger->print_energy_log(ITIME, ns->iterations, interactions, nsteps, ns->en.ini);

        download();
        logger->write_snapshot(snap_number, ITIME);
        snap_number++;

        if (ns->ops.print_all) logger->print_all(ITIME, snap_number);
        if (ns->ops.print_lagrange) {
            nu->lagrange_radii();
            logger->print_lagrange_radii(ITIME, nu->layers_radii);
        }
    }

    // The block steps between two output times run on the device
    // (h4h_run); it stops after the step whose ITIME passes the output
    // threshold, where the reference's loop prints.  Iter and Nsteps are
    // printed before that step is counted, as in the reference.
    while (ITIME < ns->integration_time) {
        h4h_run_info io;
        io.itime = ITIME;
        io.atime = ATIME;
        double t_stop = ns->interval_time * output_factor;
        ns->gtime.update_ini = omp_get_wtime();
        check(h4h_run(ctx, ns->integration_time, t_stop, ~0ull, &io, NULL, NULL), "h4h_run");
        ns->gtime.update_end += omp_get_wtime() - ns->gtime.update_ini;
        if (io.steps == 0) break;
        ITIME = io.itime;
        ATIME = io.atime;
        nact = io.last_nact;
        interactions += (long long)io.interactions;
        ns->iterations += (long long)io.steps - 1;
        nsteps += (unsigned int)(io.nsteps - io.last_nact);

        const bool output = io.stopped == H4H_STOP_OUTPUT;
        if (output) {
            double energy = get_energy_gpu();
            logger->print_energy_log(ITIME, ns->iterations, interactions, nsteps, energy);
            download();
            if (ns->ops.print_all) logger->print_all(ITIME, snap_number);
            if (ns->ops.print_lagrange) {
                nu->lagrange_radii();
                logger->print_lagrange_radii(ITIME, nu->layers_radii);
            }
            logger->write_snapshot(snap_number, ITIME);
            snap_number++;
            output_factor += 1;
        }
        nsteps += nact;
        ns->iterations++;
        if (output && !checkpoint_path.empty())
            write_checkpoint(ITIME, ATIME, interactions, nsteps, output_factor, snap_number);
    }

    ns->gtime.integration_end = omp_get_wtime() - ns->gtime.integration_ini;
    download();
    logger->write_snapshot(snap_number, ITIME);
    logger->add_info(std::string("SnapshotNumber:"), std::string(SSTR(snap_number)));
    if (!checkpoint_path.empty())
        write_checkpoint(ITIME, ATIME, interactions, nsteps, output_factor, snap_number);
}

// State of the loop top after an output (ns->h_* freshly downloaded).
void Hermite4HIP::write_checkpoint(double ITIME, double ATIME, long long interactions,
                                   unsigned int nsteps, unsigned int output_factor,
                                   unsigned int snap_number) {
    const uint64_t n = ns->n;
    H4Checkpoint c;
    c.resize(n);
    c.e2 = ns->e2;
    c.eta = ns->eta;
    c.dt_min = dt_min_;
    c.dt_max = dt_max_;
    c.precision = precision_;
    c.itime = ITIME;
    c.atime = ATIME;
    c.interval = ns->interval_time;
    c.output_factor = output_factor;
    c.snapshot_number = snap_number;
    c.energy_ini = ns->en.ini;
    c.iterations = ns->iterations;
    c.nsteps = nsteps;
    c.interactions = interactions;
    for (uint64_t i = 0; i < n; ++i) {
        const double4 &r = ns->h_r[i], &v = ns->h_v[i];
        const double r4[4] = {r.x, r.y, r.z, r.w}, v4[4] = {v.x, v.y, v.z, v.w};
        for (int q = 0; q < 4; ++q) {
            c.r[4 * i + q] = r4[q];
            c.v[4 * i + q] = v4[q];
        }
        const Forces &f = ns->h_f[i];
        const double f6[6] = {f.a[0], f.a[1], f.a[2], f.a1[0], f.a1[1], f.a1[2]};
        for (int q = 0; q < 6; ++q) c.f[6 * i + q] = f6[q];
        c.t[i] = ns->h_t[i];
        c.dt[i] = ns->h_dt[i];
        const double3 &p = ns->h_a2[i], &q3 = ns->h_a3[i];
        c.a2[3 * i] = p.x;
        c.a2[3 * i + 1] = p.y;
        c.a2[3 * i + 2] = p.z;
        c.a3[3 * i] = q3.x;
        c.a3[3 * i + 1] = q3.y;
        c.a3[3 * i + 2] = q3.z;
    }
    std::string err;
    if (!h4_write_checkpoint(checkpoint_path, c, &err)) {
        logger->log_error("checkpoint: " + err);
        std::exit(EXIT_FAILURE);
    }
}

void Hermite4HIP::read_restart(double &ITIME, double &ATIME, long long &interactions,
                               unsigned int &nsteps, unsigned int &output_factor,
                               unsigned int &snap_number) {
    H4Checkpoint c;
    std::string err;
    if (!h4_read_checkpoint(restart_path, c, &err)) {
        logger->log_error("restart: " + err);
        std::exit(EXIT_FAILURE);
    }
    const char *what = c.n != ns->n                       ? "particle count"
                       : c.e2 != ns->e2                   ? "softening"
                       : c.eta != (double)ns->eta         ? "eta"
                       : c.dt_min != dt_min_ || c.dt_max != dt_max_ ? "timestep bounds"
                       : c.precision != precision_        ? "precision"
                       : c.interval != (double)ns->interval_time ? "output interval"
                                                          : NULL;
    if (what) {
        logger->log_error(std::string("restart: the checkpoint's ") + what +
                          " differs from this run's");
        std::exit(EXIT_FAILURE);
    }
    const uint64_t n = c.n;
    for (uint64_t i = 0; i < n; ++i) {
        ns->h_r[i].x = c.r[4 * i];
        ns->h_r[i].y = c.r[4 * i + 1];
        ns->h_r[i].z = c.r[4 * i + 2];
        ns->h_r[i].w = c.r[4 * i + 3];
        ns->h_v[i].x = c.v[4 * i];
        ns->h_v[i].y = c.v[4 * i + 1];
        ns->h_v[i].z = c.v[4 * i + 2];
        ns->h_v[i].w = c.v[4 * i + 3];
        for (int q = 0; q < 3; ++q) {
            ns->h_f[i].a[q] = c.f[6 * i + q];
            ns->h_f[i].a1[q] = c.f[6 * i + 3 + q];
        }
        ns->h_t[i] = c.t[i];
        ns->h_dt[i] = c.dt[i];
        ns->h_a2[i].x = c.a2[3 * i];
        ns->h_a2[i].y = c.a2[3 * i + 1];
        ns->h_a2[i].z = c.a2[3 * i + 2];
        ns->h_a3[i].x = c.a3[3 * i];
        ns->h_a3[i].y = c.a3[3 * i + 1];
        ns->h_a3[i].z = c.a3[3 * i + 2];
    }
    upload();
    ITIME = c.itime;
    ATIME = c.atime;
    interactions = c.interactions;
    nsteps = (unsigned int)c.nsteps;
    output_factor = (unsigned int)c.output_factor;
    snap_number = (unsigned int)c.snapshot_number;
    ns->iterations = c.iterations;
    ns->en.ini = c.energy_ini;
    logger->log_info("restarted from " + restart_path + " at t = " + SSTR(ITIME));
}
