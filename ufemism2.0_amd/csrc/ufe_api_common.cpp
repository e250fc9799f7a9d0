// Host-side helpers shared by the C ABI: defaults, error text, version.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/ufe_hip.h"
#include "ufe_internal.h"

namespace ufe {

static std::mutex g_err_mu;
static std::string g_last_error;

void set_global_error(const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_last_error = msg;
}

const char *global_error() {
  std::lock_guard<std::mutex> lk(g_err_mu);
  return g_last_error.c_str();
}

}  // namespace ufe

extern "C" int ufe_version(void) { return 10000; }

// Reference defaults (model_configuration.f90:274-349,577-601) plus the
// engine's plain-form method choices.
extern "C" void ufe_params_default(ufe_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->glens_n = 3.0;
  p->epsilon_sq_0 = 1e-8;
  p->visc_eff_min = 1e4;
  p->vel_max = 5000.0;
  p->visc_it_nit = 50;
  p->visc_it_relax = 0.2;
  p->visc_it_norm_dUV_tol = 5e-5;
  p->rtol = 1e-7;
  p->abstol = 1e-5;
  p->include_crossterms = 1;
  p->sliding_law = UFE_SLIDING_WEERTMAN;
  p->slid_weertman_m = 3.0;
  p->slid_budd_q_plastic = 0.3;
  p->slid_budd_u_threshold = 100.0;
  p->slid_delta_v = 1e-3;
  p->slid_beta_max = 1e20;
  p->do_gl_subgrid_friction = 1;
  p->subgrid_friction_exponent = 2.0;
  for (int i = 0; i < 4; ++i) { p->bc_u[i] = UFE_BC_INFINITE; p->bc_v[i] = UFE_BC_INFINITE; }
  p->precision = UFE_PREC_FP64;
  p->krylov = UFE_KRYLOV_BICGSTAB;
  p->precond = UFE_PC_BLOCK_JACOBI;
  p->gmres_restart = 30;
  p->krylov_maxit = 10000;
  p->warm_start = 0;
  p->inner_rtol = 1e-3;
  p->stop_rule = UFE_STOP_L2UV;
  p->nonlinear_rtol = 1e-6;
  p->accel_depth = 0;
  p->accel_mixing = 0.5;
  p->accel_restart = 10.0;
  p->accel_continuation = 0;
  p->precond_fp32 = 0;
  p->ismip_hom_L = 0.0;
  p->slid_zi_p = 5.0;
  p->slid_zi_ut = 200.0;
  p->amg_smoother = UFE_SMOOTHER_BLOCK_JACOBI;
  p->check_every = 8;
}

// Defaults of the thickness solver's configuration fields
// (model_configuration.f90:356-364,391-392,437).
extern "C" void ufe_thickness_params_default(ufe_thickness_params *p) {
  p->fs = 1.5;
  for (int i = 0; i < 4; ++i) p->bc_H[i] = UFE_BC_ZERO;
  p->dt_ice_max = 10.0;
  p->dt_ice_min = 0.1;
  p->Hi_min = 0.0;
  p->rtol = 1e-8;
  p->abstol = 1e-6;
  p->krylov = UFE_KRYLOV_GMRES;
  p->gmres_restart = 30;
  p->maxit = 10000;
}
