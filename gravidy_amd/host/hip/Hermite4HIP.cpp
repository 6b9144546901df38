// Hermite4HIP.cpp -- the Hermite4 backend over libh4hip.so.
#include "Hermite4HIP.hpp"

#include <cstdlib>

Hermite4HIP::Hermite4HIP(NbodySystem *ns, Logger *logger, NbodyUtils *nu, int precision,
                         double dt_min, double dt_max)
    : Hermite4(ns, logger, nu), ctx(NULL) {
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

    init_acc_jrk();
    init_dt(ATIME, ETA_S, ITIME);

    ns->en.ini = get_energy_gpu();
    ns->en.tmp = ns->en.ini;

    logger->print_info();
    logger->write_info();
    logger->print_energy_log(ITIME, ns->iterations, interactions, nsteps, ns->en.ini);

    download();
    unsigned int snap_number = ns->snapshot_number;
    logger->write_snapshot(snap_number, ITIME);
    snap_number++;

    if (ns->ops.print_all) logger->print_all(ITIME, snap_number);
    if (ns->ops.print_lagrange) {
        nu->lagrange_radii();
        logger->print_lagrange_radii(ITIME, nu->layers_radii);
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

        if (io.stopped == H4H_STOP_OUTPUT) {
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
    }

    ns->gtime.integration_end = omp_get_wtime() - ns->gtime.integration_ini;
    download();
    logger->write_snapshot(snap_number, ITIME);
    logger->add_info(std::string("SnapshotNumber:"), std::string(SSTR(snap_number)));
}
