// NbodyUtils.hpp -- helpers the integrator and output layer call.  Only the
// members a backend uses are provided; the reference's diagnostics (virial
// and half-mass radii, centre of density, Lagrange radii) are not part of
// this build.
#ifndef H4_HOST_NBODYUTILS_HPP
#define H4_HOST_NBODYUTILS_HPP

#include "../NbodySystem.hpp"
#include "../common.hpp"

class NbodyUtils {
  public:
    NbodyUtils(NbodySystem *ns);
    ~NbodyUtils();

    NbodySystem *ns;
    std::vector<double> layers_radii;

    double get_magnitude(double x, double y, double z);
    void lagrange_radii();
};

#endif
