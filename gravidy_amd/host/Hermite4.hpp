// Hermite4.hpp -- integrator base class.  Same constructor and public
// members as the reference's Hermite4 (Hermite4.hpp:48-88): it owns the
// host arrays of NbodySystem that a backend fills (forces, old forces,
// a2/a3, t, dt, active list, predictors) and initialises them as
// Hermite4::init_data does (Hermite4.cpp:194-240).
#ifndef H4_HOST_HERMITE4_HPP
#define H4_HOST_HERMITE4_HPP

#include "utils/Logger.hpp"
#include "utils/NbodyUtils.hpp"

class Hermite4 {
  public:
    Hermite4(NbodySystem *ns, Logger *logger, NbodyUtils *nu);
    virtual ~Hermite4();

    NbodySystem *ns;
    Logger *logger;
    NbodyUtils *nu;

    virtual void integration() {}

  private:
    void init_data();
    void alloc_arrays_host();
    void free_arrays_host();
};

#endif
