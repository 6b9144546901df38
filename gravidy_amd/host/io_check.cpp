// io_check.cpp -- io-check: the host side of gravidy-hip without the GPU,
// for the CPU sanitizer build (ASan + UBSan): the command line (including
// the -r text resume), the initial-condition / snapshot reader, the .info,
// .log and snapshot writers, and a binary checkpoint of the read state
// written and read back.  Input files are untrusted: every reader must
// reject a bad file with a message, never read out of bounds.
//
//   io-check <checkpoint path> [gravidy-hip options]
//
// Exit status 0 when the options and input are accepted and the checkpoint
// round trip is exact, 1 otherwise.
#include <cstdio>

#include "NbodySystem.hpp"
#include "hip/H4Checkpoint.hpp"
#include "utils/Logger.hpp"

int main(int argc, char *argv[]) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: io-check <checkpoint> [gravidy-hip options]\n");
        return 2;
    }
    const std::string ckpt = argv[1];
    argv[1] = argv[0];
    OptionsParser op(argc - 1, argv + 1);
    if (!op.check_options()) return 1;

    NbodySystem ns(op);
    if (!ns.read_input_file() || ns.n == 0) {
        std::cerr << "gravidy: no particles read from "
                  << (ns.resume ? ns.snapshot_filename : ns.input_filename) << std::endl;
        return 1;
    }
    ns.alloc_base_attributes(0);
    ns.copy_input_data();

    // The output layer on the state just read, as at the start of a run.
    Logger logger(&ns);
    ns.en.ini = ns.en.tmp = -0.25;
    logger.print_info();
    logger.write_info();
    logger.print_energy_log(ns.snapshot_time, 0, 0, 0, ns.en.ini);
    logger.write_snapshot(ns.snapshot_number, ns.snapshot_time);
    logger.add_info(std::string("SnapshotNumber:"), std::string(SSTR(ns.snapshot_number + 1)));

    H4Checkpoint c;
    c.resize(ns.n);
    c.e2 = ns.e2;
    c.eta = ns.eta;
    c.dt_min = ns.dt_min;
    c.dt_max = ns.dt_max;
    c.precision = ns.precision;
    c.itime = c.atime = ns.snapshot_time;
    c.interval = ns.interval_time;
    c.output_factor = 1;
    c.snapshot_number = ns.snapshot_number;
    c.energy_ini = ns.en.ini;
    c.iterations = c.nsteps = c.interactions = 0;
    for (unsigned int i = 0; i < ns.n; ++i) {
        const double r4[4] = {ns.h_r[i].x, ns.h_r[i].y, ns.h_r[i].z, ns.h_r[i].w};
        const double v4[4] = {ns.h_v[i].x, ns.h_v[i].y, ns.h_v[i].z, ns.h_v[i].w};
        for (int q = 0; q < 4; ++q) {
            c.r[4 * (size_t)i + q] = r4[q];
            c.v[4 * (size_t)i + q] = v4[q];
        }
        c.t[i] = ns.snapshot_time;
    }
    std::string err;
    H4Checkpoint back;
    if (!h4_write_checkpoint(ckpt, c, &err) || !h4_read_checkpoint(ckpt, back, &err)) {
        std::cerr << "io-check: " << err << std::endl;
        return 1;
    }
    if (back.n != c.n || back.r != c.r || back.v != c.v || back.t != c.t ||
        back.itime != c.itime || back.e2 != c.e2) {
        std::cerr << "io-check: checkpoint round trip differs" << std::endl;
        return 1;
    }
    std::printf("n %u time %.17g snapshot %s\n", ns.n, ns.snapshot_time,
                ns.resume ? ns.snapshot_filename.c_str() : ns.input_filename.c_str());
    return 0;
}
