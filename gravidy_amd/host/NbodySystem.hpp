// NbodySystem.hpp -- system state and input reader.  Member names and types
// follow the reference's NbodySystem (NbodySystem.hpp:51-142) for the
// members an integrator backend and the output layer use.
#ifndef H4_HOST_NBODYSYSTEM_HPP
#define H4_HOST_NBODYSYSTEM_HPP

#include "common.hpp"
#include "utils/OptionsParser.hpp"

class NbodySystem {
  public:
    NbodySystem(OptionsParser op);
    ~NbodySystem();

    Gtime gtime;
    std::vector<file_data> reader;
    options ops;

    int gpus;
    long long int iterations;
    unsigned int n;
    unsigned int snapshot_number;
    unsigned int length_snap;

    float eta;
    float total_mass;
    float integration_time;
    float interval_time;

    double e2;
    double snapshot_time;
    bool resume;

    // Run-time additions of this backend: pair-kernel precision and the
    // block-timestep bounds (compile-time constants in the reference).
    int precision;
    double dt_min;
    double dt_max;

    Energy en;
    float max_mass;
    double m_g;

    std::string input_filename;
    std::string output_filename;
    std::string resume_filename;
    std::string snapshot_filename;

    unsigned int *h_id;
    unsigned int *h_move;
    double *h_t;
    double *h_dt;
    double *h_ekin;
    double *h_epot;
    double4 *h_r;
    double4 *h_v;
    double3 *h_a2;
    double3 *h_a3;
    Predictor *h_p;
    Predictor *h_i;
    Forces *h_f;
    Forces *h_old;

    bool read_input_file();
    void copy_input_data();
    void alloc_base_attributes(int rank);
    void free_base_attributes();
};

#endif
