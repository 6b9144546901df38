// common.hpp -- types and constants of the host framework.  The struct
// layouts and constant values are those the reference's integrator is
// defined by (common.hpp:67-77, 107-129, 171-231, 287-329), so arrays of
// these types can be handed to libh4hip.so and to GraviDy's own code.
#ifndef H4_HOST_COMMON_HPP
#define H4_HOST_COMMON_HPP

#include <omp.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

typedef struct double4 { double x, y, z, w; } double4;
typedef struct double3 { double x, y, z; } double3;

const int G = 1;  // gravitational constant in N-body units

const double E = 1e-4;             // default softening
const float ETA_S = 0.01;          // eta of the initial timestep
const float ETA_N = 0.01;          // eta of the corrector's timestep
const double D_TIME_MIN = 1.1920928955078125e-07;  // 2^-23
const double D_TIME_MAX = 0.125;                   // 2^-3
const double KERNEL_GFLOP = 48e-9; // flops per pair in the .log GFLOPS column

typedef struct Energy {
    double ini;
    double end;
    double tmp;
    double kinetic;
    double potential;
} Energy;

typedef struct options {
    bool print_screen;
    bool print_all;
    bool print_lagrange;
} options;

typedef struct Predictor {
    double r[3];
    double v[3];
    float m;
} Predictor;

typedef struct Forces {
    double a[3];
    double a1[3];
} Forces;

typedef struct Gtime {
    float integration_ini, integration_end;
    float prediction_ini, prediction_end;
    float update_ini, update_end;
    float correction_ini, correction_end;
    float grav_ini, grav_end;
    float reduce_ini, reduce_end;
    float reduce_forces_ini, reduce_forces_end;
    float gflops;
} Gtime;

typedef struct file_data {
    int id;
    float m;
    double r[3];
    double v[3];
} file_data;

// Integer/number -> string (the reference's SSTR macro, Logger.hpp:52-53).
template <typename T>
std::string SSTR(const T &x) {
    std::ostringstream s;
    s << std::dec << x;
    return s.str();
}

#endif
