#include "NbodyUtils.hpp"

NbodyUtils::NbodyUtils(NbodySystem *ns) : ns(ns) {}
NbodyUtils::~NbodyUtils() {}

double NbodyUtils::get_magnitude(double x, double y, double z) {
    return std::sqrt(x * x + y * y + z * z);
}

void NbodyUtils::lagrange_radii() {
    std::cerr << "gravidy: Lagrange radii are not part of this build" << std::endl;
}
