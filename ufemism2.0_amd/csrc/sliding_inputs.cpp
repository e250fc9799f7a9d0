// Till yield stress of the yield-stress sliding laws (host, once per solve).
//
// calc_sliding_law_Coulomb / _Budd / _ZoetIverson (sliding_laws.f90:132-242,
// :353-403) all start from
//   tau_c = N * tan(pi / 180) * phi
// with phi the till friction angle in degrees (the reference's expression,
// kept as written), and then run extend_till_yield_stress_to_neighbours
// (sliding_laws.f90:643-700): an ice-free land vertex takes the smallest tau_c
// of its grounded neighbours (their values before any vertex is changed, as the
// reference reads a gathered copy), the minimum starting at 1000 * rho_i * g,
// or Hi_min * rho_i * g when no neighbour is grounded.  Neither step depends on
// the velocity, so the result is an input of the whole solve: it is what the
// solve entry points take in place of beta_sq for those laws.

#include <cmath>
#include <vector>

#include "../../include/ufe_hip.h"

namespace {
// UPSY/basic/parameters.f90:15-23
constexpr double kIceDensity = 910.0;
constexpr double kGrav = 9.81;
constexpr double kPi = 3.141592653589793;
}  // namespace

extern "C" int ufe_till_yield_stress(int nV, const int *C_ptr, const int *C_ind, const double *effective_pressure,
                                     const double *till_friction_angle, const int *mask_grounded_ice,
                                     const int *mask_icefree_land, double Hi_min, double *till_yield_stress) {
  if (nV < 0 || !C_ptr || !C_ind || !effective_pressure || !till_friction_angle || !mask_grounded_ice ||
      !mask_icefree_land || !till_yield_stress)
    return UFE_ERR_ARGUMENT;
  for (int vi = 0; vi < nV; ++vi)
    if (C_ptr[vi + 1] < C_ptr[vi]) return UFE_ERR_ARGUMENT;
  for (int k = 0; k < C_ptr[nV]; ++k)
    if (C_ind[k] < 0 || C_ind[k] >= nV) return UFE_ERR_ARGUMENT;
  const double t = std::tan(kPi / 180.0);
  // the unextended field of every vertex: what the neighbours read
  std::vector<double> tau(nV);
  for (int vi = 0; vi < nV; ++vi) tau[vi] = effective_pressure[vi] * t * till_friction_angle[vi];
  for (int vi = 0; vi < nV; ++vi) {
    till_yield_stress[vi] = tau[vi];
    if (!mask_icefree_land[vi]) continue;
    double lo = 1000.0 * kIceDensity * kGrav;
    bool found = false;
    for (int k = C_ptr[vi]; k < C_ptr[vi + 1]; ++k) {
      const int vc = C_ind[k];
      if (mask_grounded_ice[vc]) {
        lo = std::fmin(lo, tau[vc]);
        found = true;
      }
    }
    till_yield_stress[vi] = found ? lo : Hi_min * kIceDensity * kGrav;
  }
  return UFE_OK;
}
