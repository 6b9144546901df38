#include "Hermite4.hpp"

Hermite4::Hermite4(NbodySystem *ns, Logger *logger, NbodyUtils *nu)
    : ns(ns), logger(logger), nu(nu) {
    alloc_arrays_host();
    init_data();
}

Hermite4::~Hermite4() { free_arrays_host(); }

void Hermite4::alloc_arrays_host() {
    ns->h_f = new Forces[ns->n];
    ns->h_a2 = new double3[ns->n];
    ns->h_a3 = new double3[ns->n];
    ns->h_old = new Forces[ns->n];
    ns->h_t = new double[ns->n];
    ns->h_dt = new double[ns->n];
    ns->h_move = new unsigned int[ns->n];
    ns->h_p = new Predictor[ns->n];
    ns->h_i = new Predictor[ns->n];
    ns->h_ekin = new double[ns->n];
    ns->h_epot = new double[ns->n];
}

void Hermite4::free_arrays_host() {
    delete[] ns->h_f;
    delete[] ns->h_a2;
    delete[] ns->h_a3;
    delete[] ns->h_old;
    delete[] ns->h_t;
    delete[] ns->h_dt;
    delete[] ns->h_move;
    delete[] ns->h_p;
    delete[] ns->h_i;
    delete[] ns->h_ekin;
    delete[] ns->h_epot;
}

// Predictors start at the initial state, everything else at zero.
void Hermite4::init_data() {
    ns->m_g = 0.0;
    for (unsigned int i = 0; i < ns->n; i++) {
        double mass = ns->h_r[i].w;
        ns->h_p[i].m = (float)mass;
        ns->h_p[i].r[0] = ns->h_r[i].x;
        ns->h_p[i].r[1] = ns->h_r[i].y;
        ns->h_p[i].r[2] = ns->h_r[i].z;
        ns->h_p[i].v[0] = ns->h_v[i].x;
        ns->h_p[i].v[1] = ns->h_v[i].y;
        ns->h_p[i].v[2] = ns->h_v[i].z;
        std::memset(&ns->h_f[i], 0, sizeof(Forces));
        std::memset(&ns->h_old[i], 0, sizeof(Forces));
        std::memset(&ns->h_a2[i], 0, sizeof(double3));
        std::memset(&ns->h_a3[i], 0, sizeof(double3));
        ns->h_t[i] = 0.0;
        ns->h_dt[i] = 0.0;
        ns->h_move[i] = 0;
        if (mass > ns->m_g) ns->m_g = mass;
    }
}
