// H4Checkpoint.hpp -- binary checkpoint of a Hermite-4 block-timestep run,
// for a bit-exact restart.
//
// The reference can restart only from a text snapshot plus its .info file
// (OptionsParser.cpp:122-324, Logger.cpp:84-97): six printed digits and a
// fresh initial timestep, so a restarted run leaves the original one.  A
// checkpoint here holds every double of the state a block step reads
// (r, v with masses, a and a1, t, dt, and a2, a3 for the outputs), the
// loop's ITIME/ATIME and counters, and the run parameters that must match
// on restart.  Layout (little endian, 8-byte fields):
//
//   "H4HCKPT" + version byte (1) | n | e2 | eta | dt_min | dt_max |
//   precision | itime | atime | interval | output_factor | snapshot_number |
//   energy_ini | iterations | nsteps | interactions |
//   r[n][4] | v[n][4] | f[n][6] | t[n] | dt[n] | a2[n][3] | a3[n][3] |
//   CRC-32 (zlib's) of all preceding bytes, as uint64
//
// Integers are stored as int64, eta as the double of its float value.
// The same format is read and written by gravidy_amd/checkpoint.py.
#ifndef H4_CHECKPOINT_HPP
#define H4_CHECKPOINT_HPP

#include <stdint.h>

#include <string>
#include <vector>

struct H4Checkpoint {
    uint64_t n = 0;
    double e2 = 0.0, eta = 0.0, dt_min = 0.0, dt_max = 0.0;
    int64_t precision = 0;
    double itime = 0.0, atime = 0.0, interval = 0.0;
    int64_t output_factor = 1, snapshot_number = 0;
    double energy_ini = 0.0;
    int64_t iterations = 0, nsteps = 0, interactions = 0;
    std::vector<double> r, v, f, t, dt, a2, a3;  // n*4, n*4, n*6, n, n, n*3, n*3

    void resize(uint64_t count);
};

// Writes to `path` through a temporary file renamed into place, so an
// interrupted write never leaves a truncated checkpoint under that name.
bool h4_write_checkpoint(const std::string &path, const H4Checkpoint &c, std::string *err);

// Reads and validates (magic, version, sizes, checksum).
bool h4_read_checkpoint(const std::string &path, H4Checkpoint &c, std::string *err);

#endif
