// ckpt_info.cpp -- h4ckpt-info: validates a binary checkpoint written by
// gravidy-hip --checkpoint (or gravidy_amd.checkpoint) and prints its
// header, one "key value" line each, plus the particle count of each array.
// Exit status 0 if the file is a valid checkpoint, 1 otherwise.
#include <cstdio>
#include <string>

#include "hip/H4Checkpoint.hpp"

int main(int argc, char *argv[]) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: h4ckpt-info <checkpoint>\n");
        return 2;
    }
    H4Checkpoint c;
    std::string err;
    if (!h4_read_checkpoint(argv[1], c, &err)) {
        std::fprintf(stderr, "h4ckpt-info: %s\n", err.c_str());
        return 1;
    }
    std::printf("n %llu\n", (unsigned long long)c.n);
    std::printf("e2 %.17g\neta %.17g\ndt_min %.17g\ndt_max %.17g\n", c.e2, c.eta, c.dt_min,
                c.dt_max);
    std::printf("precision %lld\n", (long long)c.precision);
    std::printf("itime %.17g\natime %.17g\ninterval %.17g\n", c.itime, c.atime, c.interval);
    std::printf("output_factor %lld\nsnapshot_number %lld\n", (long long)c.output_factor,
                (long long)c.snapshot_number);
    std::printf("energy_ini %.17g\n", c.energy_ini);
    std::printf("iterations %lld\nnsteps %lld\ninteractions %lld\n", (long long)c.iterations,
                (long long)c.nsteps, (long long)c.interactions);
    double sum = 0.0;
    for (size_t k = 0; k < c.r.size(); ++k) sum += c.r[k];
    std::printf("sum_r %.17g\n", sum);
    return 0;
}
