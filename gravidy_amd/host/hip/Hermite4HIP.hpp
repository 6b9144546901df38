// Hermite4HIP.hpp -- MI355X backend of the Hermite-4 block-timestep
// integrator, a Hermite4 subclass beside Hermite4CPU and Hermite4GPU.
//
// It uses only what the reference's Hermite4 base provides (ns, logger, nu
// and the host arrays Hermite4's constructor allocates), so this header and
// Hermite4HIP.cpp build unchanged inside GraviDy's src/include/hip/ as well
// as in this repository's host framework.  All state lives in HBM; the
// host arrays are filled only before output (download()).
#ifndef HERMITE4HIP_HPP
#define HERMITE4HIP_HPP

#include <string>

#include "../Hermite4.hpp"
#include "H4Checkpoint.hpp"
#include "h4hip.h"

class Hermite4HIP : public Hermite4 {
  public:
    // precision: H4H_FP64 (default), H4H_FP64_EXACT or H4H_MIXED;
    // dt_min/dt_max: block-timestep bounds (the reference's compile-time
    // D_TIME_MIN/D_TIME_MAX by default).
    Hermite4HIP(NbodySystem *ns, Logger *logger, NbodyUtils *nu, int precision = H4H_FP64,
                double dt_min = D_TIME_MIN, double dt_max = D_TIME_MAX);
    ~Hermite4HIP();

    // Hermite4CPU::integration's loop (cpu/Hermite4CPU_integration.cpp:38-130)
    // over the fused device block step.
    void integration();

    // The Hermite4 / Hermite4CPU methods, on the device.
    void init_acc_jrk();
    void init_dt(double &ATIME, float ETA, double ITIME);
    unsigned int find_particles_to_move(double ITIME);
    void save_old_acc_jrk(unsigned int nact);
    void predicted_pos_vel(double ITIME);
    void update_acc_jrk(unsigned int nact);
    void correction_pos_vel(double ITIME, unsigned int nact);
    void next_integration_time(double &ATIME);
    double get_energy_gpu();

    // One fused block step at ITIME; returns nact and sets the next time.
    unsigned int step(double ITIME, double &ATIME);
    // Device -> ns->h_r, h_v, h_f, h_old, h_a2, h_a3, h_t, h_dt.
    void download();
    // ns->h_r, h_v (and h_f, h_t, h_dt) -> device.
    void upload();

    h4h_ctx *context() { return ctx; }

    // Binary checkpoints (H4Checkpoint.hpp): written to `path` at every
    // output time and at the end of integration(); a run started with
    // set_restart(path) continues from one bit for bit (same e2, eta,
    // timestep bounds, precision and interval required).
    void set_checkpoint(const std::string &path) { checkpoint_path = path; }
    void set_restart(const std::string &path) { restart_path = path; }

  private:
    h4h_ctx *ctx;
    int precision_;
    double dt_min_, dt_max_;
    std::string checkpoint_path, restart_path;
    void check(int rc, const char *what);
    void write_checkpoint(double ITIME, double ATIME, long long interactions,
                          unsigned int nsteps, unsigned int output_factor,
                          unsigned int snap_number);
    void read_restart(double &ITIME, double &ATIME, long long &interactions,
                      unsigned int &nsteps, unsigned int &output_factor,
                      unsigned int &snap_number);
};

#endif
