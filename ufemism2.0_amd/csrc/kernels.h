// Device data structures and kernel launch wrappers of the engine.
#pragma once

#include "../../include/ufe_hip.h"
#include "sliding.h"
#include "ufe_device.h"

namespace ufe {

constexpr double kIceDensity = 910.0;       // parameters.f90:23
constexpr double kSeawaterDensity = 1028.0; // parameters.f90:25
constexpr double kGrav = 9.81;              // parameters.f90:20
constexpr double kT0 = 273.16;              // parameters.f90:18
constexpr double kClausiusClapeyron = 8.7e-4;

// Row types of the linearised system (solve_linearised_SSA_DIVA.f90:94-147).
// ROW_COPY: the lagged copy rows of the idealised boundary conditions (identity
// row, value from the previous iterate, :543-575 and :618-635), carried in the
// Dirichlet value arrays.
enum : signed char { ROW_FREE = 0, ROW_DIRICHLET = 1, ROW_INFINITE = 2, ROW_ZERO = 3, ROW_COPY = 4 };

// Device copies of the mesh and operators (pointers into HBM).
struct DevMesh {
  int nV = 0, nT = 0, nz = 0;
  int ab_nnz = 0, ba_nnz = 0, bb_nnz = 0;
  int *ab_ptr = nullptr, *ab_ind = nullptr;
  double *ab_map = nullptr, *ab_ddx = nullptr, *ab_ddy = nullptr;
  int *ba_ptr = nullptr, *ba_ind = nullptr;
  double *ba_map = nullptr, *ba_ddx = nullptr, *ba_ddy = nullptr;
  // M2 b->b, columns sorted within each row; this is also the block pattern
  // of the system matrix
  int *bb_ptr = nullptr, *bb_ind = nullptr, *bb_row = nullptr, *bb_diag = nullptr;
  double *bb_ddx = nullptr, *bb_ddy = nullptr, *bb_d2dx2 = nullptr, *bb_d2dxdy = nullptr,
         *bb_d2dy2 = nullptr;
  int *TriC = nullptr, *Tri = nullptr;
  double *V = nullptr;
  double *zeta = nullptr;
  // DIVA column kernels with 16 lanes per column (nz <= 16); false: one lane
  // per column (same results, kept for the equivalence test)
  bool diva_columns = true;
};

// Scalars of one Picard iteration (copied by value into kernels).
struct PicardParams {
  double n, eps0, visc_eff_min, eta_max;
  SlidingLaw sl;
  double subgrid_exp;
  int do_subgrid, crossterms, no_sliding;
};

// picard kernels
void launch_vertex_ssa(const DevMesh &m, const double *x, const double *A_vav, const double *Hi,
                       const double *rough, const PicardParams &p, double *N_a, double *beta_a, hipStream_t s);
void launch_triangle(const DevMesh &m, const double *N_a, const double *beta_a, const double *beta_eff_a,
                     const double *fgr, const double *tdx, const double *tdy, const signed char *rowtype,
                     const double *bc_u, const double *bc_v, const PicardParams &p, double4 *tri_coef,
                     double *beta_basal_b, double *rhs, hipStream_t s);
// Assembles S = D^-1 A (FP64, and FP32 if S32), the diagonal blocks D, identity
// Dinv (the scaled system's block-Jacobi inverse) and b_s = D^-1 rhs.
void launch_assemble(const DevMesh &m, const signed char *rowtype, const double4 *tri_coef, int crossterms,
                     const double *rhs, double4 *S, float4 *S32, double4 *Dblk, float4 *Dblk32, double4 *Dinv,
                     float4 *Dinv32, double *rhs_s, hipStream_t s);
void launch_driving_stress(const DevMesh &m, const double *Hi, const double *Hs, double *tdx, double *tdy,
                           hipStream_t s);
void launch_vertical_average(int nV, int nz, const double *zeta, const double *f_vk, double *out, hipStream_t s);
void launch_update(int nT, const double *x, double *u, double *uprev, double vel_max, double relax,
                   const RedSlot &slot, double *out2, const double *beta_basal_b, const double4 *tri_coef,
                   const double *F2b0, int no_sliding, double *ubase, double *taub, hipStream_t s);

// Anderson mixing (kernels_anderson.hip)
constexpr int kAAMaxDepth = UFE_ACCEL_MAX_DEPTH;
struct AAGamma {
  double g[kAAMaxDepth];
};
void launch_aa_columns(int nT, double *x, const double *u, double *fprev, double *gprev, double *dF, double *dG,
                       long stride, int p, int have_prev, int ncols, double vel_max, const RedSlot &slot,
                       double *out, hipStream_t s);
void launch_aa_mix(int n, double *x, const double *u, const double *dF, const double *dG, long stride, int ncols,
                   const AAGamma &gam, double mix, hipStream_t s);
// Host side of one solve's mixing: the ring position, the normal equations
// of the least-squares problem and the restart state.
struct AndersonState {
  int cols = 0, pos = 0;
  bool have_prev = false;
  double best = 1e300;
  double H[kAAMaxDepth][kAAMaxDepth] = {};
  int n_restarts = 0;
};
// Device work arrays of the mixing: depth ring columns dF, dG and the previous
// f, G of n2 (u, v) pairs each, and the reduction slot of 2 kAAMaxDepth + 1 sums.
struct AndersonBuffers {
  double *dF, *dG, *fprev, *gprev, *part, *out;
};
// One mixing step over n2 (u, v) pairs: x holds G(u) on entry (clamped to
// vel_max here) and the next iterate on exit; u is the current iterate and
// resid the nonlinear residual at u (the history restarts when it exceeds
// restart times the best one seen).  Synchronises the stream once.
void anderson_step(AndersonState &aa, const AndersonBuffers &buf, int depth, double mixing, double restart,
                   double vel_max, long n2, double *x, const double *u, double resid, hipStream_t s);

// DIVA kernels (kernels_diva.hip)
void launch_vertex_diva(const DevMesh &m, const double *x, const double *xbase, const double *taub,
                        const double *eta3b_tk, const double *A_vk, const double *Hi, const double *rough,
                        const PicardParams &p, double *N_a, double *beta_a, double *beta_eff_a,
                        double *eta3a_vk, double *F1a_vk, double *F2a_vk, hipStream_t s);
void launch_triangle_3d(const DevMesh &m, const double *eta3a_vk, const double *F1a_vk, const double *F2a_vk,
                        double *eta3b_tk, double *F1b_tk, double *F2b0, int full_F1, hipStream_t s);
void launch_velocities_3d(int nT, int nz, const double *ubase, const double *beta_basal_b, const double *taub,
                          const double *F1b_tk, int no_sliding, double *u3_out, double *v3_out, hipStream_t s);
void launch_transpose(int rows, int cols, const double *in, double *out, hipStream_t s);
void launch_flow_factor_pow(long n, const double *A, double glen_n, double *out, hipStream_t s);

// linear algebra (kernels_linalg.hip)
struct Bsr64 {
  int nb = 0, nnzb = 0;
  const int *ptr = nullptr, *col = nullptr;
  const double4 *val = nullptr;
};
struct Bsr32 {
  int nb = 0, nnzb = 0;
  const int *ptr = nullptr, *col = nullptr;
  const float4 *val = nullptr;
};

// y = A x (FP64) and, if slot != null, the grid sums ||b - y||^2 and ||b||^2 into out2
void launch_bsr_residual64(const Bsr64 &A, const double *x, const double *b, double *r, const RedSlot *slot,
                           double *out2, hipStream_t s);
void launch_bsr_spmv64(const Bsr64 &A, const double *x, double *y, hipStream_t s);
void launch_bsr_spmv32(const Bsr32 &A, const float *x, float *y, hipStream_t s);
void launch_csr_spmv64(int m, const int *ptr, const int *ind, const double *val, const double *x, double *y,
                       hipStream_t s);

// flotation and heat columns
void launch_fraction_gr_b(const DevMesh &m, const double *Hi, const double *Hb, const double *SL, double *fgr,
                          hipStream_t s);
struct HeatDev {
  int nV, nz;
  const double *zeta, *Ti, *u, *v, *w, *uadv, *vadv, *zx, *zy, *zz, *zt, *Phi, *Hi_eff, *T_surf, *Q;
  const int *mask;
  const double *fgr;  // [nV] vertex grounded fraction, read where mask = 3 (may be null otherwise)
  double dt, Hi_min;
  const double *tri;  // [6][nz] d/dzeta and d2/dzeta2 tridiagonals (ldiag, diag, udiag)
};
void launch_heat_columns(const HeatDev &h, double *Tout, int *unstable, const RedSlot &slot, double *count,
                         hipStream_t s);

}  // namespace ufe
