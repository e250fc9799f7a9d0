// Basal friction coefficient at one vertex: calc_basal_friction_coefficient
// (sliding_laws.f90:25-81) and the laws it dispatches to, shared by the SSA,
// DIVA and BPA kernels.  The velocity is the a-grid (mapped) basal velocity.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/ufe_hip.h"

namespace ufe {

struct SlidingLaw {
  int law;  // UFE_SLIDING_*
  double delta_v, beta_max, weertman_m, budd_q, budd_u_th, zi_p, zi_ut;
  // [nV] alpha_sq * effective_pressure on the device in the solve's vertex
  // order; read only by TSAI2015 and SCHOOF2005 (the host checks it is set)
  const double *cap;
};

inline SlidingLaw make_sliding_law(const ufe_params &q, const double *cap) {
  SlidingLaw s;
  s.law = q.sliding_law;
  s.delta_v = q.slid_delta_v;
  s.beta_max = q.slid_beta_max;
  s.weertman_m = q.slid_weertman_m;
  s.budd_q = q.slid_budd_q_plastic;
  s.budd_u_th = q.slid_budd_u_threshold;
  s.zi_p = q.slid_zi_p;
  s.zi_ut = q.slid_zi_ut;
  s.cap = cap;
  return s;
}

inline bool sliding_law_needs_cap(int law) { return law == UFE_SLIDING_TSAI2015 || law == UFE_SLIDING_SCHOOF2005; }

// rough: beta_sq (WEERTMAN, TSAI2015, SCHOOF2005) or the till yield stress
// (BUDD, COULOMB, ZOET_IVERSON); the result is capped at slid_beta_max (:74-75).
__device__ __forceinline__ double basal_friction(const SlidingLaw &s, int vi, double rough, double ua, double va) {
  // Bueler & Brown (2009) regularisation, as every law of the reference
  const double uabs = sqrt(s.delta_v * s.delta_v + ua * ua + va * va);
  double b;
  switch (s.law) {
    case UFE_SLIDING_WEERTMAN:  // :86-130
      b = rough * pow(uabs, 1.0 / s.weertman_m - 1.0);
      break;
    case UFE_SLIDING_BUDD:  // :200-242
      b = rough * pow(uabs, s.budd_q - 1.0) / pow(s.budd_u_th, s.budd_q);
      break;
    case UFE_SLIDING_COULOMB:  // :132-198
      b = rough / uabs;
      break;
    case UFE_SLIDING_TSAI2015:  // :244-297, Asay-Davis et al. (2016) eq. 7
      b = fmin(s.cap[vi], rough * pow(uabs, 1.0 / s.weertman_m)) / uabs;
      break;
    case UFE_SLIDING_SCHOOF2005: {  // :299-351, Asay-Davis et al. (2016) eq. 11
      const double c = s.cap[vi];
      b = (rough * pow(uabs, 1.0 / s.weertman_m) * c) /
          pow(pow(rough, s.weertman_m) * uabs + pow(c, s.weertman_m), 1.0 / s.weertman_m) / uabs;
      break;
    }
    case UFE_SLIDING_ZOET_IVERSON:  // :353-403, Zoet & Iverson (2020) eq. 3 over u
      b = rough * pow(uabs, 1.0 / s.zi_p - 1.0) * pow(uabs + s.zi_ut, -1.0 / s.zi_p);
      break;
    default:
      b = 0.0;
  }
  return fmin(s.beta_max, b);
}

}  // namespace ufe
