// Context, Picard driver and C ABI of the engine.
//
// The Picard loop is host code that only launches kernels and reads a few
// scalars per iteration (convergence norms); every array stays in HBM.  The
// loop restates solve_SSA (SSA_main.f90:149-237) and solve_DIVA
// (DIVA_main.f90:154-254): strain rates -> viscosity -> friction -> linear
// solve -> velocity limits -> relaxation -> L2_uv, with the adaptive
// relaxation / regularisation rule of SSA_main.f90:196-213.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "krylov.h"
#include "ufe_internal.h"

using namespace ufe;

struct ufe_ctx {
  int device = 0;
  hipStream_t s = nullptr;
  ufe_params prm{};
  DevMesh m;
  std::vector<std::unique_ptr<DBuf>> own;  // mesh/operator storage
  std::vector<int> h_TriBI, h_TriC;
  // host copies (device order) for the copy points of the idealised boundary conditions
  std::vector<double> h_V, h_TriGC;
  std::vector<int> h_ba_ptr, h_ba_ind;
  // lagged copy rows: row index, and per row the triangles and weights
  DBuf copy_row, copy_ptr, copy_tri, copy_w;
  int n_copy = 0;
  // locality ordering: device index -> caller index
  std::vector<int> vperm, tperm;
  // system
  DBuf A64, A32, Dinv64, Dinv32, Dblk, Dblk32, rhs_s;  // left-scaled system S = D^-1 A, b_s = D^-1 b
  // inputs
  DBuf Hi, Hs, A_vk_vf, A_vk_lf, A_vav, rough, fgr, bcu, bcv, rowtype;
  DBuf Apow_lf;  // A^(-1/n), layer-fastest, formed at the start of each DIVA solve
  DBuf cap;      // alpha_sq * effective_pressure of TSAI2015 / SCHOOF2005, device vertex order
  bool has_cap = false;
  std::vector<signed char> h_rowtype;
  bool grounded_exists = true;
  bool inputs_ready = false;
  // iterate and work arrays
  DBuf u, uprev, x, rhs, r64, tdx, tdy, N_a, beta_a, beta_eff_a, tri_coef, beta_basal_b, out2;
  // DIVA
  DBuf eta3b, eta3a, F1a, F2a, F1b, F2b0, ubase, taub, tmp3;
  DBuf beta_save;  // beta_b kept aside while the diagnostic residual is evaluated
  bool diva_state_ready = false;
  // Anderson mixing: ring of m difference columns (dF, dG), previous f and G,
  // reduction slot of width 2 kAAMaxDepth + 1
  DBuf aa_dF, aa_dG, aa_fprev, aa_gprev, aa_part, aa_out;
  std::unique_ptr<Krylov> kry;
  std::unique_ptr<AmgHierarchy> amg;
  std::string err;
  double timings[5] = {0, 0, 0, 0, 0};
  int n_fallbacks = 0;
  std::vector<std::unique_ptr<DBuf>> heat_bufs;  // ufe_heat_columns device buffers, reused
  bool amg_fold = true;  // UFE_AMG_FOLD=0: keep the separate level-0 pre-smoothing kernel
  double bbox[4] = {0, 1, 0, 1};  // xmin, xmax, ymin, ymax of the vertices
  DBuf feat;                     // learned-guess features [nT][16] fp32 (device order)
  std::shared_ptr<void> thickness;  // state of the thickness solver (thickness.hip), set up on demand
  std::shared_ptr<void> heat_mesh;  // vertex-triangle lists of the heat-column inputs (heat_inputs.hip)
  std::shared_ptr<void> bpa;        // zeta operators and BPA solver state (bpa.hip), set up on demand
  hipEvent_t ev[8];
};

namespace {

int fail(ufe_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  set_global_error(msg);
  return code;
}

template <typename F>
int guarded(ufe_ctx *c, F &&f) {
  try {
    return f();
  } catch (const HipError &e) {
    return fail(c, UFE_ERR_DEVICE, e.what());
  } catch (const std::exception &e) {
    return fail(c, UFE_ERR_ARGUMENT, e.what());
  }
}

template <typename U>
U *dev_copy(ufe_ctx *c, const U *h, size_t n) {
  auto b = std::make_unique<DBuf>(std::max<size_t>(n, 1) * sizeof(U));
  if (n) UFE_HIP(hipMemcpy(b->p, h, n * sizeof(U), hipMemcpyHostToDevice));
  U *p = b->as<U>();
  c->own.push_back(std::move(b));
  return p;
}

void h2d(DBuf &d, const void *h, size_t bytes, hipStream_t s) {
  if (d.bytes < bytes) d.alloc(bytes);
  UFE_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
}

BsrT<double> sys64(ufe_ctx *c) {
  BsrT<double> A;
  A.nb = c->m.nT; A.nnzb = c->m.bb_nnz; A.ptr = c->m.bb_ptr; A.col = c->m.bb_ind;
  A.val = c->A64.as<double4>(); A.dinv = c->Dinv64.as<double4>();
  return A;
}

BsrT<float> sys32(ufe_ctx *c) {
  BsrT<float> A;
  A.nb = c->m.nT; A.nnzb = c->m.bb_nnz; A.ptr = c->m.bb_ptr; A.col = c->m.bb_ind;
  A.val = c->A32.as<float4>(); A.dinv = c->Dinv32.as<float4>();
  return A;
}

// Row types from border indices, BC choices and the prescribed-velocity mask.
// Vertex nearest to (px, py): the one whose Voronoi cell holds the point
// (find_containing_vertex, mesh_utilities.f90).  Brute force over a bucket
// grid of the vertices, searched ring by ring until no closer bucket remains.
struct VertexGrid {
  double x0, y0, h;
  int nx, ny;
  std::vector<int> start, items;
  const std::vector<double> *V;
  explicit VertexGrid(const std::vector<double> &Vv) : V(&Vv) {
    const int nV = (int)(Vv.size() / 2);
    double x1 = -1e300, y1 = -1e300;
    x0 = y0 = 1e300;
    for (int i = 0; i < nV; ++i) {
      x0 = std::min(x0, Vv[2 * i]); x1 = std::max(x1, Vv[2 * i]);
      y0 = std::min(y0, Vv[2 * i + 1]); y1 = std::max(y1, Vv[2 * i + 1]);
    }
    const int side = std::max(1, (int)std::sqrt((double)nV / 4.0));
    h = std::max(x1 - x0, y1 - y0) / side * (1.0 + 1e-9) + 1e-300;
    nx = (int)((x1 - x0) / h) + 1;
    ny = (int)((y1 - y0) / h) + 1;
    start.assign((size_t)nx * ny + 1, 0);
    items.resize(nV);
    for (int i = 0; i < nV; ++i) ++start[cell(Vv[2 * i], Vv[2 * i + 1]) + 1];
    for (size_t k = 1; k < start.size(); ++k) start[k] += start[k - 1];
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < nV; ++i) items[fill[cell(Vv[2 * i], Vv[2 * i + 1])]++] = i;
  }
  int cx(double x) const { return std::min(nx - 1, std::max(0, (int)((x - x0) / h))); }
  int cy(double y) const { return std::min(ny - 1, std::max(0, (int)((y - y0) / h))); }
  int cell(double x, double y) const { return cy(y) * nx + cx(x); }
  int nearest(double px, double py) const {
    const int ix = cx(px), iy = cy(py);
    int best = -1;
    double dbest = 1e300;
    for (int r = 0; r <= std::max(nx, ny); ++r) {
      for (int j = iy - r; j <= iy + r; ++j)
        for (int i = ix - r; i <= ix + r; ++i) {
          if (i < 0 || j < 0 || i >= nx || j >= ny) continue;
          if (std::max(std::abs(i - ix), std::abs(j - iy)) != r) continue;
          for (int k = start[(size_t)j * nx + i]; k < start[(size_t)j * nx + i + 1]; ++k) {
            const int v = items[k];
            const double dx = (*V)[2 * v] - px, dy = (*V)[2 * v + 1] - py, d = dx * dx + dy * dy;
            if (d < dbest || (d == dbest && v < best)) { dbest = d; best = v; }
          }
        }
      // every vertex outside ring r is at least r*h away from the point's cell
      if (best >= 0 && (double)r * h >= std::sqrt(dbest)) break;
    }
    return best;
  }
};

// Triangles and weights of the lagged copy row of border triangle ti whose BC
// choice is a copy choice: the copy point (find_ti_copy_ISMIP_HOM_periodic /
// find_ti_copy_SSA_icestream_infinite, mesh_utilities.f90:2622-2730), the
// vertex nearest to it, and 1/d^2 weights to the triangles around that vertex.
void copy_entry(ufe_ctx *c, int ti, int choice, std::unique_ptr<VertexGrid> &grid, std::vector<int> &ctri,
                std::vector<double> &cw) {
  const double gx = c->h_TriGC[2 * ti], gy = c->h_TriGC[2 * ti + 1];
  double px, py;
  if (choice == UFE_BC_PERIODIC_ISMIP_HOM) {
    const double L = c->prm.ismip_hom_L;
    px = gx > 0.0 ? gx - L / 2.0 : gx + L / 2.0;
    py = gy > 0.0 ? gy - L / 2.0 : gy + L / 2.0;
  } else {
    py = gy;
    px = gx < 0.0 ? c->bbox[0] + (c->bbox[1] - c->bbox[0]) * 1.0 / 3.0
                  : c->bbox[0] + (c->bbox[1] - c->bbox[0]) * 2.0 / 3.0;
  }
  if (!grid) grid = std::make_unique<VertexGrid>(c->h_V);
  const int vi = grid->nearest(px, py);
  const int k0 = c->h_ba_ptr[vi], k1 = c->h_ba_ptr[vi + 1];
  double sum = 0.0;
  const size_t base = cw.size();
  for (int k = k0; k < k1; ++k) {
    const int tj = c->h_ba_ind[k];
    const double dx = px - c->h_TriGC[2 * tj], dy = py - c->h_TriGC[2 * tj + 1];
    const double d = std::sqrt(dx * dx + dy * dy);
    ctri.push_back(tj);
    cw.push_back(1.0 / (d * d));
    sum += 1.0 / (d * d);
  }
  for (size_t k = base; k < cw.size(); ++k) cw[k] /= sum;
}

void build_rowtypes(ufe_ctx *c, const int *bc_mask) {
  int nT = c->m.nT;
  c->h_rowtype.assign(2 * (size_t)nT, ROW_FREE);
  std::vector<int> crow, cptr(1, 0), ctri;
  std::vector<double> cw;
  std::unique_ptr<VertexGrid> grid;
  for (int ti = 0; ti < nT; ++ti) {
    for (int uv = 0; uv < 2; ++uv) {
      signed char t = ROW_FREE;
      if (bc_mask && bc_mask[ti] == 1) t = ROW_DIRICHLET;
      else if (c->h_TriBI[ti] > 0) {
        int side = (c->h_TriBI[ti] - 1) / 2;  // 1,2 N; 3,4 E; 5,6 S; 7,8 W
        int choice = uv == 0 ? c->prm.bc_u[side] : c->prm.bc_v[side];
        if (choice == UFE_BC_PERIODIC_ISMIP_HOM || choice == UFE_BC_INFINITE_SSA_ICESTREAM) {
          t = ROW_COPY;
          copy_entry(c, ti, choice, grid, ctri, cw);
          crow.push_back(2 * ti + uv);
          cptr.push_back((int)ctri.size());
        } else {
          t = choice == UFE_BC_ZERO ? ROW_ZERO : ROW_INFINITE;
        }
      }
      c->h_rowtype[2 * ti + uv] = t;
    }
  }
  h2d(c->rowtype, c->h_rowtype.data(), c->h_rowtype.size(), c->s);
  c->n_copy = (int)crow.size();
  if (c->n_copy) {
    h2d(c->copy_row, crow.data(), crow.size() * sizeof(int), c->s);
    h2d(c->copy_ptr, cptr.data(), cptr.size() * sizeof(int), c->s);
    h2d(c->copy_tri, ctri.data(), ctri.size() * sizeof(int), c->s);
    h2d(c->copy_w, cw.data(), cw.size() * sizeof(double), c->s);
    UFE_HIP(hipStreamSynchronize(c->s));  // the host lists end with this call
  }
}

// Values of the lagged copy rows at the current iterate, written into the
// Dirichlet value arrays (the rows are identity rows with that value).
__global__ void __launch_bounds__(kBlock) k_copy_rows(int n, const int *__restrict__ row, const int *__restrict__ ptr,
                                                      const int *__restrict__ tri, const double *__restrict__ w,
                                                      const double2 *__restrict__ u, double relax,
                                                      double *__restrict__ bcu, double *__restrict__ bcv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int rw = row[r], ti = rw / 2, uv = rw % 2;
  double val = 0.0;
  for (int k = ptr[r]; k < ptr[r + 1]; ++k) {
    const double2 q = u[tri[k]];
    val += w[k] * (uv == 0 ? q.x : q.y);
  }
  const double2 self = u[ti];
  val = (relax * val) + ((1.0 - relax) * (uv == 0 ? self.x : self.y));
  (uv == 0 ? bcu : bcv)[ti] = val;
}

// The sliding-law constants of a solve; throws (UFE_ERR_ARGUMENT through
// guarded) when the law needs the Coulomb cap and none is set.
SlidingLaw ctx_sliding(const ufe_ctx *c) {
  if (sliding_law_needs_cap(c->prm.sliding_law) && !c->has_cap)
    throw std::invalid_argument("the Tsai2015 and Schoof2005 sliding laws need ufe_set_sliding_cap first");
  return make_sliding_law(c->prm, c->has_cap ? c->cap.as<double>() : nullptr);
}

PicardParams picard_params(const ufe_ctx *c, double eps0) {
  const ufe_params &q = c->prm;
  PicardParams p;
  p.n = q.glens_n;
  p.eps0 = eps0;
  p.visc_eff_min = q.visc_eff_min;
  p.eta_max = 0.5 * std::pow(1e-18, -1.0 / q.glens_n) * std::pow(eps0, (1.0 - q.glens_n) / (2.0 * q.glens_n));
  p.sl = ctx_sliding(c);
  p.subgrid_exp = q.subgrid_friction_exponent;
  p.do_subgrid = q.do_gl_subgrid_friction;
  p.crossterms = q.include_crossterms;
  p.no_sliding = q.sliding_law == UFE_SLIDING_NO_SLIDING;
  return p;
}

}  // namespace

// FP32 correction added to the FP64 iterate (refinement step).
__global__ void __launch_bounds__(kBlock) k_add_f32(int n, double *__restrict__ x, const float *__restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)d[i];
}

__global__ void __launch_bounds__(kBlock) k_copy_or_zero(int n, const double *__restrict__ a, double *__restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a ? a[i] : 0.0;
}

namespace ufe {
void launch_f64_to_f32(int n, const double *a, float *b, hipStream_t s);
}

// FP64 Krylov solve of the assembled system into c->x with the chosen method
// (the reference's PETSc default is GMRES(30), petsc_basic.f90:106-119).
static SolveStats krylov64(ufe_ctx *c, const BsrT<double> &A, double rtol, bool zero_guess) {
  const ufe_params &q = c->prm;
  if (q.krylov == UFE_KRYLOV_GMRES)
    return c->kry->gmres64(A, c->rhs_s.as<double>(), c->x.as<double>(), rtol, q.abstol, q.krylov_maxit, zero_guess,
                           q.gmres_restart, q.check_every);
  return c->kry->bicgstab64(A, c->rhs_s.as<double>(), c->x.as<double>(), rtol, q.abstol, q.krylov_maxit, zero_guess,
                            q.check_every);
}

// ---------------------------------------------------------------------------
// Linear solve of one Picard iteration: x := solution of A x = rhs.
// ---------------------------------------------------------------------------
static void linear_solve(ufe_ctx *c, int *its, int *nref) {
  const ufe_params &q = c->prm;
  int n = 2 * c->m.nT;
  Krylov &K = *c->kry;
  bool warm = q.warm_start != 0;
  hipLaunchKernelGGL(k_copy_or_zero, dim3(blocks_for(n)), dim3(kBlock), 0, c->s, n,
                     warm ? c->u.as<double>() : (const double *)nullptr, c->x.as<double>());
  UFE_LAUNCH_CHECK();
  *its = 0;
  *nref = 0;
  // All stopping tests use the block-Jacobi-scaled residual (velocity units)
  // of the assembled system S = D^-1 A, the norm PETSc's left-preconditioned
  // KSP test measures with a block-Jacobi preconditioner.
  // the assembled system is S = D^-1 A: identity diagonal blocks
  K.set_unit_diag0(c->amg_fold);
  K.set_amg_l1(q.amg_smoother == UFE_SMOOTHER_L1_BLOCK_JACOBI);
  if (q.precision == UFE_PREC_FP64) {
    BsrT<double> A = sys64(c);
    K.set_amg_fine32(q.precond_fp32 && q.precond == UFE_PC_AMG ? c->A32.as<float4>() : nullptr);
    if (q.precond == UFE_PC_AMG) K.amg_setup64(A);
    SolveStats st = krylov64(c, A, q.rtol, !warm);
    *its = st.its;
    return;
  }
  // mixed: FP64 residual, FP32 corrections
  BsrT<double> A = sys64(c);
  BsrT<float> A32 = sys32(c);
  K.set_amg_fine32(nullptr);
  if (q.precond == UFE_PC_AMG) K.amg_setup32(A32);
  double rr, bb;
  K.residual64(A, c->x.as<double>(), c->rhs_s.as<double>(), c->r64.as<double>(), &rr, &bb);
  double r0 = std::sqrt(rr);
  double target = std::max(q.rtol * r0, q.abstol);
  float *r32 = K.tmp32(8), *d32 = K.tmp32(9);
  const double rr_start = rr;
  for (int ref = 0; ref < 20 && std::sqrt(rr) > target; ++ref) {
    launch_f64_to_f32(n, c->r64.as<double>(), r32, c->s);
    double rn = std::sqrt(rr);
    // each FP32 solve reduces the residual by at most inner_rtol (well above
    // FP32 round-off); the FP64 refinement loop takes as many steps as needed
    double inner = std::max(0.5 * target / rn, q.inner_rtol);
    SolveStats st = q.krylov == UFE_KRYLOV_GMRES
                        ? K.gmres32(A32, r32, d32, inner, 0.0, q.krylov_maxit - *its, true, q.gmres_restart,
                                    q.check_every)
                        : K.bicgstab32(A32, r32, d32, inner, 0.0, q.krylov_maxit - *its, true, q.check_every);
    *its += st.its;
    hipLaunchKernelGGL(k_add_f32, dim3(blocks_for(n)), dim3(kBlock), 0, c->s, n, c->x.as<double>(), d32);
    UFE_LAUNCH_CHECK();
    K.residual64(A, c->x.as<double>(), c->rhs_s.as<double>(), c->r64.as<double>(), &rr, &bb);
    ++*nref;
    if (*its >= q.krylov_maxit) break;
  }
  if (!(rr <= rr_start) || !(std::sqrt(rr) <= std::max(target, 0.5 * r0))) {
    // refinement did not make progress (FP32 stagnation): restart from the
    // initial guess with the FP64 solver
    ++c->n_fallbacks;
    hipLaunchKernelGGL(k_copy_or_zero, dim3(blocks_for(n)), dim3(kBlock), 0, c->s, n,
                       warm ? c->u.as<double>() : (const double *)nullptr, c->x.as<double>());
    UFE_LAUNCH_CHECK();
    if (q.precond == UFE_PC_AMG) K.amg_setup64(A);
    SolveStats st = krylov64(c, A, q.rtol, !warm);
    *its += st.its;
  }
}

// Nonlinear residual at the iterate, ||D^-1 (b - A u)|| / ||D^-1 b|| with D the
// 2x2 diagonal blocks: every block row becomes a velocity equation (m/yr), so
// momentum rows and boundary rows count on the same scale.  Rows of triangles
// where the velocity limit of apply_velocity_limits
// (SSA_DIVA_utilities.f90:110-141) is active are left out: there the
// converged iterate satisfies the limit, not the momentum balance.  The
// unscaled residual b - A u is left in r64.
static double rel_residual(ufe_ctx *c, int *n_clamped = nullptr) {
  double rr, bb;
  c->kry->residual64(sys64(c), c->u.as<double>(), c->rhs_s.as<double>(), c->r64.as<double>(), &rr, &bb,
                     c->prm.vel_max, n_clamped, c->Dblk.as<double4>(), 1);
  return bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

// ||r|| and ||b|| of the unscaled system over rows whose velocity is below
// vel_max (the rows rel_residual measures), as per-workgroup partials.
__global__ void __launch_bounds__(kBlock) k_unscaled_norms(int nT, const double2 *__restrict__ r,
                                                           const double2 *__restrict__ b,
                                                           const double2 *__restrict__ u, double vmax, RedSlot slot) {
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nT; i += gridDim.x * blockDim.x) {
    const double2 uu = u[i];
    if (vmax > 0.0 && sqrt(uu.x * uu.x + uu.y * uu.y) >= vmax * (1.0 - 1e-9)) continue;
    const double2 rr = r[i], bb = b[i];
    acc[0] += rr.x * rr.x + rr.y * rr.y;
    acc[1] += bb.x * bb.x + bb.y * bb.y;
  }
  __shared__ double lds[2][kBlock / 64];
  block_sum<2>(acc, lds);
  block_partials<2>(slot, acc);
}

// ||b - A u|| / ||b|| of the unscaled system from the residual the last
// rel_residual call left in r64 and the unscaled right-hand side.
static double unscaled_rel_residual(ufe_ctx *c) {
  const int nT = c->m.nT;
  const int nb = std::min(blocks_for(nT), kMaxPartials);
  RedSlot sl = c->kry->slot(3);
  hipLaunchKernelGGL(k_unscaled_norms, dim3(nb), dim3(kBlock), 0, c->s, nT, c->r64.as<double2>(),
                     c->rhs.as<double2>(), c->u.as<double2>(), c->prm.vel_max, sl);
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_partials_out<2>, dim3(1), dim3(kBlock), 0, c->s, sl, nb, sl.result);
  UFE_LAUNCH_CHECK();
  double h[2];
  UFE_HIP(hipMemcpyAsync(h, sl.result, sizeof(h), hipMemcpyDeviceToHost, c->s));
  UFE_HIP(hipStreamSynchronize(c->s));
  return h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
}

// Everything up to the assembled linear system at the current iterate.
static void picard_state(ufe_ctx *c, bool diva, double eps0, bool need_F1) {
  const ufe_params &q = c->prm;
  PicardParams p = picard_params(c, eps0);
  hipStream_t s = c->s;
  UFE_HIP(hipEventRecord(c->ev[0], s));
  if (!diva) {
    launch_vertex_ssa(c->m, c->u.as<double>(), c->A_vav.as<double>(), c->Hi.as<double>(), c->rough.as<double>(), p,
                      c->N_a.as<double>(), c->beta_a.as<double>(), s);
  } else {
    launch_vertex_diva(c->m, c->u.as<double>(), c->ubase.as<double>(), c->taub.as<double>(), c->eta3b.as<double>(),
                       c->Apow_lf.as<double>(), c->Hi.as<double>(), c->rough.as<double>(), p, c->N_a.as<double>(),
                       c->beta_a.as<double>(), c->beta_eff_a.as<double>(), c->eta3a.as<double>(),
                       c->F1a.as<double>(), c->F2a.as<double>(), s);
    launch_triangle_3d(c->m, c->eta3a.as<double>(), c->F1a.as<double>(), c->F2a.as<double>(), c->eta3b.as<double>(),
                       c->F1b.as<double>(), c->F2b0.as<double>(), need_F1 ? 1 : 0, s);
  }
  UFE_HIP(hipEventRecord(c->ev[1], s));
  if (c->n_copy) {
    hipLaunchKernelGGL(k_copy_rows, dim3(blocks_for(c->n_copy)), dim3(kBlock), 0, s, c->n_copy,
                       c->copy_row.as<int>(), c->copy_ptr.as<int>(), c->copy_tri.as<int>(), c->copy_w.as<double>(),
                       c->u.as<double2>(), q.visc_it_relax, c->bcu.as<double>(), c->bcv.as<double>());
    UFE_LAUNCH_CHECK();
  }
  launch_triangle(c->m, c->N_a.as<double>(), c->beta_a.as<double>(), diva ? c->beta_eff_a.as<double>() : nullptr,
                  c->fgr.as<double>(), c->tdx.as<double>(), c->tdy.as<double>(), c->rowtype.as<signed char>(),
                  c->bcu.as<double>(), c->bcv.as<double>(), p, c->tri_coef.as<double4>(),
                  c->beta_basal_b.as<double>(), c->rhs.as<double>(), s);
  UFE_HIP(hipEventRecord(c->ev[2], s));
  bool mirror = q.precision == UFE_PREC_MIXED || (q.precond_fp32 && q.precond == UFE_PC_AMG);
  launch_assemble(c->m, c->rowtype.as<signed char>(), c->tri_coef.as<double4>(), q.include_crossterms,
                  c->rhs.as<double>(), c->A64.as<double4>(), mirror ? c->A32.as<float4>() : nullptr,
                  c->Dblk.as<double4>(), mirror ? c->Dblk32.as<float4>() : nullptr, c->Dinv64.as<double4>(),
                  mirror ? c->Dinv32.as<float4>() : nullptr, c->rhs_s.as<double>(), s);
  UFE_HIP(hipEventRecord(c->ev[3], s));
}

// ---------------------------------------------------------------------------
// Anderson mixing step (accel_depth > 0): x holds G(u_k) from the linear
// solve on entry and the mixed next iterate on exit.  The normal equations
// (dF^T dF) gamma = dF^T f_k are at most 8x8 and solved on the host.
// ---------------------------------------------------------------------------
static void aa_prepare(ufe_ctx *c) {
  size_t n = 2 * (size_t)c->m.nT;
  size_t cols = (size_t)kAAMaxDepth * n * sizeof(double);
  if (c->aa_dF.bytes < cols) {
    c->aa_dF.alloc(cols);
    c->aa_dG.alloc(cols);
    c->aa_fprev.alloc(n * sizeof(double));
    c->aa_gprev.alloc(n * sizeof(double));
    c->aa_part.alloc(sizeof(double) * kMaxPartials * (2 * kAAMaxDepth + 1));
    c->aa_out.alloc(sizeof(double) * (2 * kAAMaxDepth + 1));
  }
}

static void aa_step(ufe_ctx *c, AndersonState &aa, double resid) {
  const ufe_params &q = c->prm;
  const AndersonBuffers buf{c->aa_dF.as<double>(), c->aa_dG.as<double>(), c->aa_fprev.as<double>(),
                            c->aa_gprev.as<double>(), c->aa_part.as<double>(), c->aa_out.as<double>()};
  anderson_step(aa, buf, q.accel_depth, q.accel_mixing, q.accel_restart, q.vel_max, c->m.nT, c->x.as<double>(),
                c->u.as<double>(), resid, c->s);
}

static int run_picard(ufe_ctx *c, bool diva, ufe_solve_info *info) {
  const ufe_params &q = c->prm;
  auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = c->s;
  int nT = c->m.nT;
  ufe_solve_info inf{};
  if (!c->inputs_ready) return fail(c, UFE_ERR_ARGUMENT, "inputs not uploaded");
  ctx_sliding(c);  // throws when the law needs the Coulomb cap and none is set
  if (!c->grounded_exists || (!diva && q.sliding_law == UFE_SLIDING_NO_SLIDING)) {
    // solve_SSA / solve_DIVA early exit: no grounded ice -> zero velocities
    UFE_HIP(hipMemsetAsync(c->u.p, 0, c->u.bytes, s));
    if (diva) {
      UFE_HIP(hipMemsetAsync(c->ubase.p, 0, c->ubase.bytes, s));
      UFE_HIP(hipMemsetAsync(c->taub.p, 0, c->taub.bytes, s));
      UFE_HIP(hipMemsetAsync(c->F1b.p, 0, c->F1b.bytes, s));
    }
    UFE_HIP(hipStreamSynchronize(s));
    if (info) *info = inf;
    return UFE_OK;
  }
  c->kry->set_precond(q.precond);
  launch_driving_stress(c->m, c->Hi.as<double>(), c->Hs.as<double>(), c->tdx.as<double>(), c->tdy.as<double>(), s);
  if (!diva)
    launch_vertical_average(c->m.nV, c->m.nz, c->m.zeta, c->A_vk_vf.as<double>(), c->A_vav.as<double>(), s);
  else
    launch_flow_factor_pow((long)c->m.nV * c->m.nz, c->A_vk_lf.as<double>(), q.glens_n, c->Apow_lf.as<double>(), s);
  if (diva && !c->diva_state_ready) {
    UFE_HIP(hipMemsetAsync(c->eta3b.p, 0, c->eta3b.bytes, s));
    UFE_HIP(hipMemsetAsync(c->ubase.p, 0, c->ubase.bytes, s));
    UFE_HIP(hipMemsetAsync(c->taub.p, 0, c->taub.bytes, s));
    c->diva_state_ready = true;
  }
  double L2 = 1e9, L2_prev;
  int nit_div = 0, it = 0, n_axb = 0, n_ref = 0, rc = UFE_OK, n_clamped = 0;
  double relax = q.visc_it_relax, eps0 = q.epsilon_sq_0;
  double resid = -1.0;
  bool converged = false;
  AndersonState aa;
  if (q.accel_depth > 0) aa_prepare(c);
  const bool verbose = std::getenv("UFE_VERBOSE") != nullptr;
  for (double &t : c->timings) t = 0.0;
  c->n_fallbacks = 0;
  const int n_cont = q.accel_depth > 0 ? q.accel_continuation : 0;
  while (!converged) {
    ++it;
    // regularisation continuation: epsilon_sq_0 * 10^(n_cont - it + 1) for it <= n_cont
    const bool in_cont = it <= n_cont;
    const double eps_it = in_cont ? eps0 * std::pow(10.0, n_cont - it + 1) : eps0;
    if (q.accel_depth > 0 && it == n_cont + 1 && n_cont > 0) aa = AndersonState();
    picard_state(c, diva, eps_it, false);
    if (q.stop_rule == UFE_STOP_RESIDUAL) {
      resid = rel_residual(c, &n_clamped);
      if (!in_cont && resid <= q.nonlinear_rtol) { --it; converged = true; break; }
    }
    int its = 0, nref = 0;
    linear_solve(c, &its, &nref);
    n_axb += its;
    n_ref += nref;
    if (q.accel_depth > 0) aa_step(c, aa, resid);
    RedSlot sl = c->kry->slot(3);
    launch_update(nT, c->x.as<double>(), c->u.as<double>(), c->uprev.as<double>(), q.vel_max,
                  q.accel_depth > 0 ? 1.0 : relax, sl,
                  c->out2.as<double>(), c->beta_basal_b.as<double>(), c->tri_coef.as<double4>(),
                  diva ? c->F2b0.as<double>() : nullptr, q.sliding_law == UFE_SLIDING_NO_SLIDING,
                  diva ? c->ubase.as<double>() : nullptr, diva ? c->taub.as<double>() : nullptr, s);
    double h2[2];
    UFE_HIP(hipMemcpyAsync(h2, c->out2.p, sizeof(h2), hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    float ms;
    for (int k = 0; k < 3; ++k) {
      UFE_HIP(hipEventElapsedTime(&ms, c->ev[k], c->ev[k + 1]));
      c->timings[k] += ms;
    }
    L2_prev = L2;
    L2 = 2.0 * h2[0] / std::max(h2[1], 1e-8);
    // adaptive relaxation / regularisation rule of the relaxed iteration
    if (q.accel_depth == 0 && L2 > L2_prev) ++nit_div; else nit_div = 0;
    if (nit_div > 2) { nit_div = 0; relax *= 0.9; eps0 *= 1.2; }
    if (relax <= 0.05 || eps0 >= 1e-5) {
      if (relax < 0.05 || eps0 > 1e-5) {
        rc = fail(c, UFE_ERR_DIVERGED, "viscosity iteration still diverges (relaxation or regularisation limit)");
        break;
      }
    }
    if (verbose)
      std::fprintf(stderr, "[ufe] it %d resid %.3e L2 %.3e relax %.3f eps0 %.2e krylov %d refs %d clamped %d\n", it,
                   resid, L2, relax, eps0, its, nref, n_clamped);
    if (q.stop_rule == UFE_STOP_L2UV && L2 < q.visc_it_norm_dUV_tol) converged = true;
    if (it > q.visc_it_nit) { rc = UFE_WARN_NOT_CONVERGED; break; }
  }
  if (diva && rc >= 0) {
    // F1_3D_b of the last state the loop evaluated: calc_3D_velocities
    // (DIVA_main.f90:254, :636-676) combines it with that state's beta_b and
    // the u_base / tau_b formed after the last update
    launch_triangle_3d(c->m, c->eta3a.as<double>(), c->F1a.as<double>(), c->F2a.as<double>(),
                       c->tmp3.as<double>(), c->F1b.as<double>(), c->F2b0.as<double>(), 1, s);
  }
  double resid_unscaled = -1.0;
  if (rc >= 0 && (q.stop_rule != UFE_STOP_RESIDUAL || !converged)) {
    // Diagnostic residual at the returned iterate.  The evaluation overwrites
    // the DIVA state, so eta_3D_b (returned, and the next call's warm start)
    // and beta_b (read by the 3-D velocities) are kept aside and restored.
    const size_t e_bytes = sizeof(double) * (size_t)nT * c->m.nz, b_bytes = sizeof(double) * (size_t)nT;
    if (diva) {
      if (c->beta_save.bytes < b_bytes) c->beta_save.alloc(b_bytes);
      UFE_HIP(hipMemcpyAsync(c->tmp3.p, c->eta3b.p, e_bytes, hipMemcpyDeviceToDevice, s));
      UFE_HIP(hipMemcpyAsync(c->beta_save.p, c->beta_basal_b.p, b_bytes, hipMemcpyDeviceToDevice, s));
    }
    picard_state(c, diva, eps0, false);
    resid = rel_residual(c, &n_clamped);
    resid_unscaled = unscaled_rel_residual(c);
    if (diva) {
      UFE_HIP(hipMemcpyAsync(c->eta3b.p, c->tmp3.p, e_bytes, hipMemcpyDeviceToDevice, s));
      UFE_HIP(hipMemcpyAsync(c->beta_basal_b.p, c->beta_save.p, b_bytes, hipMemcpyDeviceToDevice, s));
    }
  }
  // converged by the residual rule: r64 and rhs still hold the last evaluation at u
  if (rc >= 0 && resid_unscaled < 0.0 && resid >= 0.0) resid_unscaled = unscaled_rel_residual(c);
  UFE_HIP(hipStreamSynchronize(s));
  inf.n_visc_its = it;
  inf.rel_residual_unscaled = resid_unscaled;
  inf.n_Axb_its = n_axb;
  inf.n_refinements = n_ref;
  inf.L2_uv = L2;
  inf.rel_residual = resid;
  inf.visc_it_relax_applied = relax;
  inf.epsilon_sq_0_applied = eps0;
  inf.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  inf.n_clamped = n_clamped;
  inf.n_fallbacks = c->n_fallbacks;
  if (info) *info = inf;
  return rc;
}

// ---------------------------------------------------------------------------
// Locality ordering.  Triangles and vertices are renumbered along a Morton
// (Z-order) curve of their positions, so neighbours in the mesh are close in
// memory: the gathers of the operator products and of the SpMV hit the same
// cache lines, and the greedy aggregation builds compact aggregates whose
// members have nearby indices.  The caller's numbering is restored at the
// boundary.
// ---------------------------------------------------------------------------
static uint32_t morton2(uint32_t x, uint32_t y) {
  auto spread = [](uint32_t v) {
    v &= 0xffff;
    v = (v | (v << 8)) & 0x00ff00ff;
    v = (v | (v << 4)) & 0x0f0f0f0f;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}

static std::vector<int> morton_order(int n, const double *xy) {
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (int i = 0; i < n; ++i) {
    x0 = std::min(x0, xy[2 * i]); x1 = std::max(x1, xy[2 * i]);
    y0 = std::min(y0, xy[2 * i + 1]); y1 = std::max(y1, xy[2 * i + 1]);
  }
  double sx = 65535.0 / std::max(x1 - x0, 1e-30), sy = 65535.0 / std::max(y1 - y0, 1e-30);
  double sc = std::min(sx, sy);
  std::vector<uint32_t> key(n);
  for (int i = 0; i < n; ++i)
    key[i] = morton2((uint32_t)((xy[2 * i] - x0) * sc), (uint32_t)((xy[2 * i + 1] - y0) * sc));
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  return order;
}

struct PermutedMesh {
  ufe_mesh_desc desc{};
  std::vector<int> t_n2o, t_o2n, v_n2o, v_o2n;
  std::vector<double> V, TriGC, ab_v[3], ba_v[3], bb_v[5];
  std::vector<int> Tri, TriC, TriBI, ab_ptr, ab_ind, ba_ptr, ba_ind, bb_ptr, bb_ind;
};

static void permute_csr(int rows, const int *ptr, const int *ind, const double *const *vals, int nv,
                        const std::vector<int> &row_n2o, const std::vector<int> &col_o2n, std::vector<int> &optr,
                        std::vector<int> &oind, std::vector<double> *ovals) {
  optr.assign(rows + 1, 0);
  oind.clear();
  for (int q = 0; q < nv; ++q) ovals[q].clear();
  for (int r = 0; r < rows; ++r) {
    int o = row_n2o[r];
    for (int k = ptr[o]; k < ptr[o + 1]; ++k) {
      oind.push_back(col_o2n[ind[k]]);
      for (int q = 0; q < nv; ++q) ovals[q].push_back(vals[q][k]);
    }
    optr[r + 1] = (int)oind.size();
  }
}

static void permute_mesh(const ufe_mesh_desc &m, PermutedMesh &p) {
  int nV = m.nV, nT = m.nT;
  p.v_n2o = morton_order(nV, m.V);
  p.t_n2o = morton_order(nT, m.TriGC);
  p.v_o2n.assign(nV, 0);
  p.t_o2n.assign(nT, 0);
  for (int i = 0; i < nV; ++i) p.v_o2n[p.v_n2o[i]] = i;
  for (int i = 0; i < nT; ++i) p.t_o2n[p.t_n2o[i]] = i;
  p.V.resize(2 * (size_t)nV);
  for (int i = 0; i < nV; ++i) { p.V[2 * i] = m.V[2 * p.v_n2o[i]]; p.V[2 * i + 1] = m.V[2 * p.v_n2o[i] + 1]; }
  p.Tri.resize(3 * (size_t)nT); p.TriC.resize(3 * (size_t)nT); p.TriBI.resize(nT); p.TriGC.resize(2 * (size_t)nT);
  for (int i = 0; i < nT; ++i) {
    int o = p.t_n2o[i];
    for (int q = 0; q < 3; ++q) {
      p.Tri[3 * i + q] = p.v_o2n[m.Tri[3 * o + q]];
      int tc = m.TriC[3 * o + q];
      p.TriC[3 * i + q] = tc < 0 ? -1 : p.t_o2n[tc];
    }
    p.TriBI[i] = m.TriBI[o];
    p.TriGC[2 * i] = m.TriGC[2 * o];
    p.TriGC[2 * i + 1] = m.TriGC[2 * o + 1];
  }
  const double *ab[3] = {m.ab_map, m.ab_ddx, m.ab_ddy};
  const double *ba[3] = {m.ba_map, m.ba_ddx, m.ba_ddy};
  const double *bb[5] = {m.bb_ddx, m.bb_ddy, m.bb_d2dx2, m.bb_d2dxdy, m.bb_d2dy2};
  permute_csr(nT, m.ab_ptr, m.ab_ind, ab, 3, p.t_n2o, p.v_o2n, p.ab_ptr, p.ab_ind, p.ab_v);
  permute_csr(nV, m.ba_ptr, m.ba_ind, ba, 3, p.v_n2o, p.t_o2n, p.ba_ptr, p.ba_ind, p.ba_v);
  permute_csr(nT, m.bb_ptr, m.bb_ind, bb, 5, p.t_n2o, p.t_o2n, p.bb_ptr, p.bb_ind, p.bb_v);
  ufe_mesh_desc &d = p.desc;
  d = m;
  d.V = p.V.data(); d.Tri = p.Tri.data(); d.TriC = p.TriC.data(); d.TriBI = p.TriBI.data(); d.TriGC = p.TriGC.data();
  d.ab_ptr = p.ab_ptr.data(); d.ab_ind = p.ab_ind.data();
  d.ab_map = p.ab_v[0].data(); d.ab_ddx = p.ab_v[1].data(); d.ab_ddy = p.ab_v[2].data();
  d.ba_ptr = p.ba_ptr.data(); d.ba_ind = p.ba_ind.data();
  d.ba_map = p.ba_v[0].data(); d.ba_ddx = p.ba_v[1].data(); d.ba_ddy = p.ba_v[2].data();
  d.bb_ptr = p.bb_ptr.data(); d.bb_ind = p.bb_ind.data();
  d.bb_ddx = p.bb_v[0].data(); d.bb_ddy = p.bb_v[1].data(); d.bb_d2dx2 = p.bb_v[2].data();
  d.bb_d2dxdy = p.bb_v[3].data(); d.bb_d2dy2 = p.bb_v[4].data();
}

// caller layout (n per layer, `layers` layers) -> device numbering
template <typename U>
static std::vector<U> to_dev(const std::vector<int> &n2o, const U *src, int layers = 1) {
  size_t n = n2o.size();
  std::vector<U> out(n * layers);
  for (int k = 0; k < layers; ++k)
    for (size_t i = 0; i < n; ++i) out[k * n + i] = src[k * n + n2o[i]];
  return out;
}

template <typename U>
static void from_dev(const std::vector<int> &n2o, const U *src, U *dst, int layers = 1) {
  size_t n = n2o.size();
  for (int k = 0; k < layers; ++k)
    for (size_t i = 0; i < n; ++i) dst[k * n + n2o[i]] = src[k * n + i];
}

// Accessors for the other translation units (grounded fractions, heat columns).
namespace ufe {
int ctx_device(const ufe_ctx *c) { return c->device; }
// Device buffer i of the heat columns, kept in the context between calls and
// grown when a call needs more.
void *ctx_heat_buffer(ufe_ctx *c, int i, size_t bytes) {
  if ((int)c->heat_bufs.size() <= i) c->heat_bufs.resize(i + 1);
  auto &b = c->heat_bufs[i];
  if (!b || b->bytes < bytes) b = std::make_unique<DBuf>(std::max<size_t>(bytes, 8));
  return b->p;
}
hipStream_t ctx_stream(const ufe_ctx *c) { return c->s; }
const DevMesh *ctx_mesh(const ufe_ctx *c) { return &c->m; }
Krylov *ctx_krylov(const ufe_ctx *c) { return c->kry.get(); }
int ctx_fail(ufe_ctx *c, int code, const std::string &msg) { return fail(c, code, msg); }
const std::vector<int> &ctx_vperm(const ufe_ctx *c) { return c->vperm; }
const std::vector<int> &ctx_tperm(const ufe_ctx *c) { return c->tperm; }
std::shared_ptr<void> &ctx_thickness(ufe_ctx *c) { return c->thickness; }
std::shared_ptr<void> &ctx_heat_mesh(ufe_ctx *c) { return c->heat_mesh; }
std::shared_ptr<void> &ctx_bpa(ufe_ctx *c) { return c->bpa; }
// Lagged copy lists of every border triangle whose BC choice (per component)
// is a copy choice, with no prescribed-velocity mask: rows 2*ti+uv (device
// order), their triangles and weights (CSR).
void ctx_copy_lists(ufe_ctx *c, std::vector<int> &crow, std::vector<int> &cptr, std::vector<int> &ctri,
                    std::vector<double> &cw) {
  crow.clear(); ctri.clear(); cw.clear();
  cptr.assign(1, 0);
  std::unique_ptr<VertexGrid> grid;
  for (int ti = 0; ti < c->m.nT; ++ti) {
    if (c->h_TriBI[ti] <= 0) continue;
    const int side = (c->h_TriBI[ti] - 1) / 2;
    for (int uv = 0; uv < 2; ++uv) {
      const int choice = uv == 0 ? c->prm.bc_u[side] : c->prm.bc_v[side];
      if (choice != UFE_BC_PERIODIC_ISMIP_HOM && choice != UFE_BC_INFINITE_SSA_ICESTREAM) continue;
      copy_entry(c, ti, choice, grid, ctri, cw);
      crow.push_back(2 * ti + uv);
      cptr.push_back((int)ctri.size());
    }
  }
}
const std::vector<int> &ctx_TriBI(const ufe_ctx *c) { return c->h_TriBI; }
const ufe_params &ctx_params(const ufe_ctx *c) { return c->prm; }
SlidingLaw ctx_sliding_law(const ufe_ctx *c) { return ctx_sliding(c); }
const double *ctx_velocity(const ufe_ctx *c) { return c->u.as<double>(); }
int ctx_guard(ufe_ctx *c, const std::function<int()> &f) { return guarded(c, f); }
}  // namespace ufe

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" const char *ufe_last_error(const ufe_ctx *ctx) { return ctx ? ctx->err.c_str() : global_error(); }

extern "C" int ufe_ctx_create(ufe_ctx **out, int device, const ufe_mesh_desc *md) {
  if (!out || !md) return fail(nullptr, UFE_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  auto c = std::make_unique<ufe_ctx>();
  int rc = guarded(c.get(), [&]() -> int {
    int ndev = 0;
    UFE_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(nullptr, UFE_ERR_DEVICE, "no such HIP device");
    if (md->nV <= 0 || md->nT <= 0 || md->nz < 2 || md->nz > 64)
      return fail(nullptr, UFE_ERR_ARGUMENT, "invalid mesh sizes");
    c->device = device;
    UFE_HIP(hipSetDevice(device));
    UFE_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (auto &e : c->ev) UFE_HIP(hipEventCreate(&e));
    ufe_params_default(&c->prm);
    DevMesh &m = c->m;
    m.nV = md->nV; m.nT = md->nT; m.nz = md->nz;
    int nV = m.nV, nT = m.nT, nz = m.nz;
    // validate index ranges on the host before anything reaches a kernel
    auto check_csr = [&](int rows, int ncols, const int *ptr, const int *ind, int nnz, const char *name) {
      if (ptr[0] != 0 || ptr[rows] != nnz) throw std::runtime_error(std::string(name) + ": bad row pointer");
      for (int i = 0; i < rows; ++i)
        if (ptr[i + 1] < ptr[i]) throw std::runtime_error(std::string(name) + ": row pointer not monotone");
      for (int k = 0; k < nnz; ++k)
        if (ind[k] < 0 || ind[k] >= ncols) throw std::runtime_error(std::string(name) + ": column out of range");
    };
    check_csr(nT, nV, md->ab_ptr, md->ab_ind, md->ab_nnz, "a->b operators");
    check_csr(nV, nT, md->ba_ptr, md->ba_ind, md->ba_nnz, "b->a operators");
    check_csr(nT, nT, md->bb_ptr, md->bb_ind, md->bb_nnz, "b->b operators");
    for (int k = 0; k < 3 * nT; ++k) {
      if (md->TriC[k] < -1 || md->TriC[k] >= nT) throw std::runtime_error("TriC out of range");
      if (md->Tri[k] < 0 || md->Tri[k] >= nV) throw std::runtime_error("Tri out of range");
    }
    for (int ti = 0; ti < nT; ++ti)
      if (md->TriBI[ti] < 0 || md->TriBI[ti] > 8) throw std::runtime_error("TriBI out of range (0..8)");
    PermutedMesh pm;
    permute_mesh(*md, pm);
    c->vperm = pm.v_n2o;
    c->tperm = pm.t_n2o;
    md = &pm.desc;
    c->h_TriBI.assign(md->TriBI, md->TriBI + nT);
    c->h_TriC.assign(md->TriC, md->TriC + 3 * nT);
    c->h_V.assign(md->V, md->V + 2 * (size_t)nV);
    c->h_TriGC.assign(md->TriGC, md->TriGC + 2 * (size_t)nT);
    c->h_ba_ptr.assign(md->ba_ptr, md->ba_ptr + nV + 1);
    c->h_ba_ind.assign(md->ba_ind, md->ba_ind + md->ba_nnz);
    m.ab_nnz = md->ab_nnz; m.ba_nnz = md->ba_nnz; m.bb_nnz = md->bb_nnz;
    m.ab_ptr = dev_copy(c.get(), md->ab_ptr, nT + 1);
    m.ab_ind = dev_copy(c.get(), md->ab_ind, md->ab_nnz);
    m.ab_map = dev_copy(c.get(), md->ab_map, md->ab_nnz);
    m.ab_ddx = dev_copy(c.get(), md->ab_ddx, md->ab_nnz);
    m.ab_ddy = dev_copy(c.get(), md->ab_ddy, md->ab_nnz);
    m.ba_ptr = dev_copy(c.get(), md->ba_ptr, nV + 1);
    m.ba_ind = dev_copy(c.get(), md->ba_ind, md->ba_nnz);
    m.ba_map = dev_copy(c.get(), md->ba_map, md->ba_nnz);
    m.ba_ddx = dev_copy(c.get(), md->ba_ddx, md->ba_nnz);
    m.ba_ddy = dev_copy(c.get(), md->ba_ddy, md->ba_nnz);
    // M2 rows sorted by column: the block pattern of the system matrix
    int nnz = md->bb_nnz;
    std::vector<int> ind(nnz), row(nnz), diag(nT, -1), perm(nnz);
    std::vector<double> v[5];
    for (auto &a : v) a.resize(nnz);
    const double *src[5] = {md->bb_ddx, md->bb_ddy, md->bb_d2dx2, md->bb_d2dxdy, md->bb_d2dy2};
    for (int ti = 0; ti < nT; ++ti) {
      int k0 = md->bb_ptr[ti], k1 = md->bb_ptr[ti + 1];
      for (int k = k0; k < k1; ++k) perm[k] = k;
      std::sort(perm.begin() + k0, perm.begin() + k1,
                [&](int a, int b) { return md->bb_ind[a] < md->bb_ind[b]; });
      for (int k = k0; k < k1; ++k) {
        ind[k] = md->bb_ind[perm[k]];
        row[k] = ti;
        if (k > k0 && ind[k] == ind[k - 1]) throw std::runtime_error("duplicate column in b->b operator row");
        if (ind[k] == ti) diag[ti] = k;
        for (int q = 0; q < 5; ++q) v[q][k] = src[q][perm[k]];
      }
      if (diag[ti] < 0) throw std::runtime_error("b->b operator row without diagonal");
      // the 'infinite' boundary rows use the TriC neighbours; they must lie in the stencil
      for (int q = 0; q < 3; ++q) {
        int tj = md->TriC[3 * ti + q];
        if (tj < 0) continue;
        if (!std::binary_search(ind.begin() + k0, ind.begin() + k1, tj))
          throw std::runtime_error("TriC neighbour missing from the b->b stencil");
      }
    }
    m.bb_ptr = dev_copy(c.get(), md->bb_ptr, nT + 1);
    m.bb_ind = dev_copy(c.get(), ind.data(), nnz);
    m.bb_row = dev_copy(c.get(), row.data(), nnz);
    m.bb_diag = dev_copy(c.get(), diag.data(), nT);
    m.bb_ddx = dev_copy(c.get(), v[0].data(), nnz);
    m.bb_ddy = dev_copy(c.get(), v[1].data(), nnz);
    m.bb_d2dx2 = dev_copy(c.get(), v[2].data(), nnz);
    m.bb_d2dxdy = dev_copy(c.get(), v[3].data(), nnz);
    m.bb_d2dy2 = dev_copy(c.get(), v[4].data(), nnz);
    m.TriC = dev_copy(c.get(), md->TriC, 3 * (size_t)nT);
    m.Tri = dev_copy(c.get(), md->Tri, 3 * (size_t)nT);
    m.V = dev_copy(c.get(), md->V, 2 * (size_t)nV);
    m.zeta = dev_copy(c.get(), md->zeta, nz);
    c->bbox[0] = c->bbox[2] = 1e300;
    c->bbox[1] = c->bbox[3] = -1e300;
    for (int vi = 0; vi < nV; ++vi) {
      c->bbox[0] = std::min(c->bbox[0], md->V[2 * vi]);
      c->bbox[1] = std::max(c->bbox[1], md->V[2 * vi]);
      c->bbox[2] = std::min(c->bbox[2], md->V[2 * vi + 1]);
      c->bbox[3] = std::max(c->bbox[3], md->V[2 * vi + 1]);
    }
    // system and work arrays
    c->A64.alloc(sizeof(double4) * (size_t)nnz);
    c->A32.alloc(sizeof(float4) * (size_t)nnz);
    c->Dinv64.alloc(sizeof(double4) * (size_t)nT);
    c->Dinv32.alloc(sizeof(float4) * (size_t)nT);
    c->Dblk.alloc(sizeof(double4) * (size_t)nT);
    c->Dblk32.alloc(sizeof(float4) * (size_t)nT);
    c->rhs_s.alloc(sizeof(double) * 2 * (size_t)nT);
    size_t v2 = sizeof(double) * 2 * (size_t)nT;
    for (DBuf *b : {&c->u, &c->uprev, &c->x, &c->rhs, &c->r64, &c->ubase, &c->taub}) b->alloc(v2);
    for (DBuf *b : {&c->tdx, &c->tdy, &c->beta_basal_b, &c->F2b0, &c->fgr, &c->bcu, &c->bcv})
      b->alloc(sizeof(double) * (size_t)nT);
    c->tri_coef.alloc(sizeof(double4) * (size_t)nT);
    for (DBuf *b : {&c->N_a, &c->beta_a, &c->beta_eff_a, &c->A_vav, &c->Hi, &c->Hs, &c->rough})
      b->alloc(sizeof(double) * (size_t)nV);
    for (DBuf *b : {&c->eta3a, &c->F1a, &c->F2a, &c->A_vk_vf, &c->A_vk_lf, &c->Apow_lf})
      b->alloc(sizeof(double) * (size_t)nV * nz);
    for (DBuf *b : {&c->eta3b, &c->F1b, &c->tmp3}) b->alloc(sizeof(double) * (size_t)nT * nz);
    c->out2.alloc(sizeof(double) * 4);
    UFE_HIP(hipMemset(c->u.p, 0, v2));
    c->kry = std::make_unique<Krylov>(nT, c->s);
    {
      const char *e;
      double om = (e = std::getenv("UFE_AMG_OMEGA")) ? std::atof(e) : 0.85;
      double cs = (e = std::getenv("UFE_AMG_SCALE")) ? std::atof(e) : 1.5;
      c->kry->set_amg_options(om, 0, cs);
      if ((e = std::getenv("UFE_AMG_TAIL_ROWS"))) c->kry->set_amg_tail(std::max(0, std::atoi(e)));
      if ((e = std::getenv("UFE_AMG_FOLD"))) c->amg_fold = std::atoi(e) != 0;
      if ((e = std::getenv("UFE_DIVA_COLUMNS"))) c->m.diva_columns = std::atoi(e) != 0;
      if ((e = std::getenv("UFE_GRAPHS"))) c->kry->set_graphs(std::atoi(e) != 0);
    }
    c->kry->set_diag(m.bb_diag);
    {
      const char *e = std::getenv("UFE_AMG_PASSES");
      int passes = e ? std::max(1, std::min(3, std::atoi(e))) : 1;
      c->amg = build_amg_hierarchy(nT, md->TriC, md->bb_ptr, ind.data(), 24, 16, passes);
    }
    if (c->amg) c->kry->attach_amg(c->amg.get());
    build_rowtypes(c.get(), nullptr);
    UFE_HIP(hipStreamSynchronize(c->s));
    return UFE_OK;
  });
  if (rc != UFE_OK) {
    set_global_error(c->err.empty() ? std::string(global_error()) : c->err);
    return rc;
  }
  *out = c.release();
  return UFE_OK;
}

extern "C" int ufe_ctx_destroy(ufe_ctx *c) {
  if (!c) return UFE_OK;
  (void)hipSetDevice(c->device);
  if (c->s) (void)hipStreamSynchronize(c->s);
  for (auto &e : c->ev) (void)hipEventDestroy(e);
  hipStream_t s = c->s;
  delete c;
  if (s) (void)hipStreamDestroy(s);
  return UFE_OK;
}

extern "C" int ufe_set_amg_options(ufe_ctx *c, double omega, double coarse_scale, int tail_rows) {
  if (!c) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  if (!(omega > 0.0 && omega <= 1.0) || !(coarse_scale > 0.0 && coarse_scale <= 3.0) || tail_rows < 0 ||
      tail_rows > 4096)
    return fail(c, UFE_ERR_ARGUMENT, "multigrid options out of range");
  return guarded(c, [&]() -> int {
    c->kry->set_amg_options(omega, 0, coarse_scale);
    c->kry->set_amg_tail(tail_rows);
    return UFE_OK;
  });
}

extern "C" int ufe_set_sliding_cap(ufe_ctx *c, const double *alpha_sq, const double *effective_pressure) {
  if (!c || (!alpha_sq) != (!effective_pressure)) return fail(c, UFE_ERR_ARGUMENT, "need both alpha_sq and effective_pressure, or neither");
  return guarded(c, [&]() -> int {
    if (!alpha_sq) {
      c->has_cap = false;
      return UFE_OK;
    }
    UFE_HIP(hipSetDevice(c->device));
    const int nV = c->m.nV;
    std::vector<double> h(nV);
    for (int i = 0; i < nV; ++i) h[i] = alpha_sq[c->vperm[i]] * effective_pressure[c->vperm[i]];
    h2d(c->cap, h.data(), sizeof(double) * nV, c->s);
    UFE_HIP(hipStreamSynchronize(c->s));
    c->has_cap = true;
    return UFE_OK;
  });
}

extern "C" int ufe_set_graph_launch(ufe_ctx *c, int enable) {
  if (!c) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  return guarded(c, [&]() -> int {
    c->kry->set_graphs(enable != 0);
    return UFE_OK;
  });
}

extern "C" int ufe_set_params(ufe_ctx *c, const ufe_params *p) {
  if (!c || !p) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  return guarded(c, [&]() -> int {
    if (p->precond == UFE_PC_AMG && !c->amg) return fail(c, UFE_ERR_UNSUPPORTED, "no aggregation hierarchy for this mesh");
    if (p->krylov != UFE_KRYLOV_BICGSTAB && p->krylov != UFE_KRYLOV_GMRES)
      return fail(c, UFE_ERR_ARGUMENT, "krylov must be UFE_KRYLOV_BICGSTAB or UFE_KRYLOV_GMRES");
    if (p->krylov == UFE_KRYLOV_GMRES && (p->gmres_restart < 1 || p->gmres_restart > 64))
      return fail(c, UFE_ERR_ARGUMENT, "gmres_restart must be 1..64");
    if (p->check_every < 1 || p->krylov_maxit < 1) return fail(c, UFE_ERR_ARGUMENT, "invalid Krylov controls");
    if (p->accel_depth < 0 || p->accel_depth > UFE_ACCEL_MAX_DEPTH)
      return fail(c, UFE_ERR_ARGUMENT, "accel_depth out of range");
    if (p->accel_continuation < 0 || p->accel_continuation > 12 ||
        (p->accel_continuation > 0 && p->accel_depth == 0))
      return fail(c, UFE_ERR_ARGUMENT, "accel_continuation needs accel_depth > 0 and at most 12 decades");
    for (int side = 0; side < 4; ++side) {
      if (p->bc_u[side] < UFE_BC_INFINITE || p->bc_u[side] > UFE_BC_INFINITE_SSA_ICESTREAM)
        return fail(c, UFE_ERR_ARGUMENT, "unknown choice_BC_u");
      if (p->bc_v[side] < UFE_BC_INFINITE || p->bc_v[side] > UFE_BC_PERIODIC_ISMIP_HOM)
        return fail(c, UFE_ERR_ARGUMENT, "unknown choice_BC_v");
      if ((p->bc_u[side] == UFE_BC_PERIODIC_ISMIP_HOM || p->bc_v[side] == UFE_BC_PERIODIC_ISMIP_HOM) &&
          !(p->ismip_hom_L > 0.0))
        return fail(c, UFE_ERR_ARGUMENT, "periodic_ISMIP-HOM needs ismip_hom_L > 0");
    }
    if (p->amg_smoother != UFE_SMOOTHER_BLOCK_JACOBI && p->amg_smoother != UFE_SMOOTHER_L1_BLOCK_JACOBI)
      return fail(c, UFE_ERR_ARGUMENT, "amg_smoother must be UFE_SMOOTHER_BLOCK_JACOBI or UFE_SMOOTHER_L1_BLOCK_JACOBI");
    if (p->sliding_law < UFE_SLIDING_NO_SLIDING || p->sliding_law > UFE_SLIDING_ZOET_IVERSON)
      return fail(c, UFE_ERR_ARGUMENT, "unknown choice_sliding_law");
    if (p->sliding_law == UFE_SLIDING_ZOET_IVERSON && !(p->slid_zi_p > 0.0 && p->slid_zi_ut >= 0.0))
      return fail(c, UFE_ERR_ARGUMENT, "Zoet-Iverson needs slid_ZI_p > 0 and slid_ZI_ut >= 0");
    if (p->accel_depth > 0 && (p->stop_rule != UFE_STOP_RESIDUAL || !(p->accel_mixing > 0.0) ||
                               !(p->accel_mixing <= 1.0) || !(p->accel_restart > 1.0)))
      return fail(c, UFE_ERR_ARGUMENT,
                  "Anderson acceleration needs stop_rule = RESIDUAL, 0 < accel_mixing <= 1, accel_restart > 1");
    c->prm = *p;
    UFE_HIP(hipSetDevice(c->device));
    // border row types depend on the BC choices; keep any Dirichlet rows
    std::vector<int> mask(c->m.nT, 0);
    for (int ti = 0; ti < c->m.nT; ++ti) mask[ti] = c->h_rowtype.size() ? (c->h_rowtype[2 * ti] == ROW_DIRICHLET) : 0;
    build_rowtypes(c, mask.data());
    UFE_HIP(hipStreamSynchronize(c->s));
    return UFE_OK;
  });
}

__global__ void k_transpose(int rows, int cols, const double *__restrict__ in, double *__restrict__ out);

extern "C" int ufe_upload_inputs(ufe_ctx *c, const double *Hi, const double *Hs, const double *A_flow,
                                 const double *beta_sq, const double *fgr, const int *mask_grounded,
                                 const int *bc_mask, const double *bc_u, const double *bc_v) {
  if (!c || !Hi || !Hs || !A_flow || !beta_sq || !fgr) return fail(c, UFE_ERR_ARGUMENT, "null input array");
  if ((bc_mask || bc_u || bc_v) && !(bc_mask && bc_u && bc_v))
    return fail(c, UFE_ERR_ARGUMENT, "need to provide prescribed u,v fields and mask!");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    int nV = c->m.nV, nT = c->m.nT, nz = c->m.nz;
    hipStream_t s = c->s;
    auto pHi = to_dev(c->vperm, Hi), pHs = to_dev(c->vperm, Hs), pB = to_dev(c->vperm, beta_sq);
    auto pF = to_dev(c->tperm, fgr), pA = to_dev(c->vperm, A_flow, nz);
    h2d(c->Hi, pHi.data(), sizeof(double) * nV, s);
    h2d(c->Hs, pHs.data(), sizeof(double) * nV, s);
    h2d(c->rough, pB.data(), sizeof(double) * nV, s);
    h2d(c->fgr, pF.data(), sizeof(double) * nT, s);
    h2d(c->A_vk_vf, pA.data(), sizeof(double) * (size_t)nV * nz, s);
    launch_transpose(nz, nV, c->A_vk_vf.as<double>(), c->A_vk_lf.as<double>(), s);
    c->grounded_exists = true;
    if (mask_grounded) {
      c->grounded_exists = false;
      for (int vi = 0; vi < nV; ++vi) if (mask_grounded[vi]) { c->grounded_exists = true; break; }
    }
    std::vector<double> zero(nT, 0.0);
    auto pu = bc_u ? to_dev(c->tperm, bc_u) : zero, pv = bc_v ? to_dev(c->tperm, bc_v) : zero;
    h2d(c->bcu, pu.data(), sizeof(double) * nT, s);
    h2d(c->bcv, pv.data(), sizeof(double) * nT, s);
    std::vector<int> pm;
    if (bc_mask) pm = to_dev(c->tperm, bc_mask);
    build_rowtypes(c, bc_mask ? pm.data() : nullptr);
    UFE_HIP(hipStreamSynchronize(s));
    c->inputs_ready = true;
    return UFE_OK;
  });
}

static void interleave_up(ufe_ctx *c, DBuf &dst, const double *a, const double *b) {
  int nT = c->m.nT;
  if (!a && !b) {  // zero field, set on the device
    UFE_HIP(hipMemsetAsync(dst.p, 0, 2 * (size_t)nT * sizeof(double), c->s));
    UFE_HIP(hipStreamSynchronize(c->s));
    return;
  }
  std::vector<double> h(2 * (size_t)nT);
  for (int i = 0; i < nT; ++i) {
    int o = c->tperm[i];
    h[2 * i] = a ? a[o] : 0.0;
    h[2 * i + 1] = b ? b[o] : 0.0;
  }
  UFE_HIP(hipMemcpy(dst.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
}

static void interleave_down(ufe_ctx *c, const DBuf &src, double *a, double *b) {
  int nT = c->m.nT;
  std::vector<double> h(2 * (size_t)nT);
  UFE_HIP(hipMemcpy(h.data(), src.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < nT; ++i) {
    int o = c->tperm[i];
    if (a) a[o] = h[2 * i];
    if (b) b[o] = h[2 * i + 1];
  }
}

extern "C" int ufe_upload_velocity(ufe_ctx *c, const double *u_b, const double *v_b) {
  if (!c || (!u_b) != (!v_b)) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    interleave_up(c, c->u, u_b, v_b);
    return UFE_OK;
  });
}

extern "C" int ufe_download_velocity(ufe_ctx *c, double *u_b, double *v_b) {
  if (!c || !u_b || !v_b) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    UFE_HIP(hipStreamSynchronize(c->s));
    interleave_down(c, c->u, u_b, v_b);
    return UFE_OK;
  });
}

extern "C" int ufe_upload_diva_state(ufe_ctx *c, const double *eta_3D_b, const double *tau_bx_b,
                                     const double *tau_by_b, const double *u_base_b, const double *v_base_b) {
  if (!c) return fail(c, UFE_ERR_ARGUMENT, "null context");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    int nT = c->m.nT, nz = c->m.nz;
    std::vector<double> pe;
    if (eta_3D_b) {
      pe = to_dev(c->tperm, eta_3D_b, nz);
      h2d(c->tmp3, pe.data(), sizeof(double) * (size_t)nT * nz, c->s);
      launch_transpose(nz, nT, c->tmp3.as<double>(), c->eta3b.as<double>(), c->s);
    } else {
      UFE_HIP(hipMemsetAsync(c->eta3b.p, 0, c->eta3b.bytes, c->s));
    }
    UFE_HIP(hipStreamSynchronize(c->s));
    interleave_up(c, c->taub, tau_bx_b, tau_by_b);
    interleave_up(c, c->ubase, u_base_b, v_base_b);
    c->diva_state_ready = true;
    return UFE_OK;
  });
}

extern "C" int ufe_download_diva(ufe_ctx *c, double *u_base_b, double *v_base_b, double *u_3D_b, double *v_3D_b,
                                 double *eta_3D_b, double *tau_bx_b, double *tau_by_b) {
  if (!c) return fail(c, UFE_ERR_ARGUMENT, "null context");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    int nT = c->m.nT, nz = c->m.nz;
    hipStream_t s = c->s;
    UFE_HIP(hipStreamSynchronize(s));
    if (u_base_b || v_base_b) interleave_down(c, c->ubase, u_base_b, v_base_b);
    if (tau_bx_b || tau_by_b) interleave_down(c, c->taub, tau_bx_b, tau_by_b);
    if (u_3D_b || v_3D_b) {
      DBuf u3(sizeof(double) * (size_t)nT * nz), v3(sizeof(double) * (size_t)nT * nz);
      launch_velocities_3d(nT, nz, c->ubase.as<double>(), c->beta_basal_b.as<double>(), c->taub.as<double>(),
                           c->F1b.as<double>(), c->prm.sliding_law == UFE_SLIDING_NO_SLIDING, u3.as<double>(),
                           v3.as<double>(), s);
      UFE_HIP(hipStreamSynchronize(s));
      std::vector<double> h((size_t)nT * nz);
      if (u_3D_b) {
        UFE_HIP(hipMemcpy(h.data(), u3.p, u3.bytes, hipMemcpyDeviceToHost));
        from_dev(c->tperm, h.data(), u_3D_b, nz);
      }
      if (v_3D_b) {
        UFE_HIP(hipMemcpy(h.data(), v3.p, v3.bytes, hipMemcpyDeviceToHost));
        from_dev(c->tperm, h.data(), v_3D_b, nz);
      }
    }
    if (eta_3D_b) {
      launch_transpose(nT, nz, c->eta3b.as<double>(), c->tmp3.as<double>(), s);
      UFE_HIP(hipStreamSynchronize(s));
      std::vector<double> h((size_t)nT * nz);
      UFE_HIP(hipMemcpy(h.data(), c->tmp3.p, sizeof(double) * (size_t)nT * nz, hipMemcpyDeviceToHost));
      from_dev(c->tperm, h.data(), eta_3D_b, nz);
    }
    return UFE_OK;
  });
}

extern "C" int ufe_run(ufe_ctx *c, int stress_balance, ufe_solve_info *info) {
  if (!c) return fail(c, UFE_ERR_ARGUMENT, "null context");
  if (stress_balance != UFE_STRESS_SSA && stress_balance != UFE_STRESS_DIVA)
    return fail(c, UFE_ERR_UNSUPPORTED, "stress balance must be SSA or DIVA");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    return run_picard(c, stress_balance == UFE_STRESS_DIVA, info);
  });
}

extern "C" int ufe_solve_ssa(ufe_ctx *c, const double *Hi, const double *Hs, const double *A_flow,
                             const double *beta_sq, const double *fgr, const int *mask_grounded, const int *bc_mask,
                             const double *bc_u, const double *bc_v, double *u_b, double *v_b, ufe_solve_info *info) {
  int rc = ufe_upload_inputs(c, Hi, Hs, A_flow, beta_sq, fgr, mask_grounded, bc_mask, bc_u, bc_v);
  if (rc < 0) return rc;
  rc = ufe_upload_velocity(c, u_b, v_b);
  if (rc < 0) return rc;
  int rr = ufe_run(c, UFE_STRESS_SSA, info);
  if (rr < 0) return rr;
  rc = ufe_download_velocity(c, u_b, v_b);
  return rc < 0 ? rc : rr;
}

extern "C" int ufe_solve_diva(ufe_ctx *c, const double *Hi, const double *Hs, const double *A_flow,
                              const double *beta_sq, const double *fgr, const int *mask_grounded, const int *bc_mask,
                              const double *bc_u, const double *bc_v, double *u_vav_b, double *v_vav_b,
                              double *u_base_b, double *v_base_b, double *u_3D_b, double *v_3D_b, double *eta_3D_b,
                              double *tau_bx_b, double *tau_by_b, ufe_solve_info *info) {
  int rc = ufe_upload_inputs(c, Hi, Hs, A_flow, beta_sq, fgr, mask_grounded, bc_mask, bc_u, bc_v);
  if (rc < 0) return rc;
  rc = ufe_upload_velocity(c, u_vav_b, v_vav_b);
  if (rc < 0) return rc;
  rc = ufe_upload_diva_state(c, eta_3D_b, tau_bx_b, tau_by_b, u_base_b, v_base_b);
  if (rc < 0) return rc;
  int rr = ufe_run(c, UFE_STRESS_DIVA, info);
  if (rr < 0) return rr;
  rc = ufe_download_velocity(c, u_vav_b, v_vav_b);
  if (rc < 0) return rc;
  rc = ufe_download_diva(c, u_base_b, v_base_b, u_3D_b, v_3D_b, eta_3D_b, tau_bx_b, tau_by_b);
  return rc < 0 ? rc : rr;
}

// Plain streaming copy of n doubles (16-byte vector accesses): the reference
// kernel of known traffic against which hardware byte counters are scaled.
__global__ void __launch_bounds__(kBlock) k_stream_copy(long n2, const double2 *__restrict__ in,
                                                        double2 *__restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x)
    out[i] = in[i];
}

extern "C" int ufe_profile_copy(ufe_ctx *c, long bytes, int reps, double *avg_ms) {
  if (!c || !avg_ms || reps < 1 || bytes < 16) return fail(c, UFE_ERR_ARGUMENT, "invalid argument");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    long n2 = bytes / 16;
    DBuf a, b;
    a.alloc(n2 * 16);
    b.alloc(n2 * 16);
    UFE_HIP(hipMemsetAsync(a.p, 0, n2 * 16, c->s));
    int nb = (int)std::min<long>(blocks_for(n2), 8192);
    hipEvent_t e0, e1;
    UFE_HIP(hipEventCreate(&e0));
    UFE_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(kBlock), 0, c->s, n2, a.as<double2>(), b.as<double2>());
    UFE_HIP(hipEventRecord(e0, c->s));
    for (int i = 0; i < reps; ++i)
      hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(kBlock), 0, c->s, n2, a.as<double2>(), b.as<double2>());
    UFE_LAUNCH_CHECK();
    UFE_HIP(hipEventRecord(e1, c->s));
    UFE_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    UFE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    return UFE_OK;
  });
}

extern "C" int ufe_kernel_timing(ufe_ctx *c, int enable) {
  if (!c) return fail(c, UFE_ERR_ARGUMENT, "invalid argument");
  return guarded(c, [&]() -> int {
    c->kry->set_spmv_timing(enable != 0);
    return UFE_OK;
  });
}

extern "C" int ufe_kernel_timing_read(const ufe_ctx *c, int kernel, long long *launches, double *avg_ms,
                                      double *bytes) {
  if (!c || !launches || !avg_ms || !bytes || kernel < 0 || kernel > 1) return UFE_ERR_ARGUMENT;
  double tot = 0.0;
  c->kry->spmv_timing(kernel, launches, &tot, bytes);
  *avg_ms = *launches > 0 ? tot / (double)*launches : 0.0;
  return UFE_OK;
}

extern "C" int ufe_profile_spmv(ufe_ctx *c, int reps, double *avg_ms, double *bytes) {
  if (!c || !avg_ms || reps < 1) return fail(c, UFE_ERR_ARGUMENT, "invalid argument");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    // product with the current iterate, into a scratch vector
    *avg_ms = c->kry->time_spmv64(sys64(c), c->u.as<double>(), c->kry->tmp64(7), reps, bytes);
    return UFE_OK;
  });
}

extern "C" int ufe_last_timings(const ufe_ctx *c, double *ms5) {
  if (!c || !ms5) return UFE_ERR_ARGUMENT;
  for (int i = 0; i < 5; ++i) ms5[i] = c->timings[i];
  return UFE_OK;
}

// ---------------------------------------------------------------------------
// Generic CSR boundary
// ---------------------------------------------------------------------------
extern "C" int ufe_csr_matvec(int m, int n, int nnz, const int *ptr, const int *ind, const double *val,
                              const double *x, double *y) {
  if (m < 0 || n < 0 || nnz < 0 || !ptr || (nnz && (!ind || !val)) || !x || !y)
    return fail(nullptr, UFE_ERR_ARGUMENT, "invalid CSR arguments");
  if (ptr[0] != 0 || ptr[m] != nnz) return fail(nullptr, UFE_ERR_ARGUMENT, "bad row pointer");
  for (int i = 0; i < m; ++i)
    if (ptr[i + 1] < ptr[i]) return fail(nullptr, UFE_ERR_ARGUMENT, "row pointer not monotone");
  for (int k = 0; k < nnz; ++k)
    if (ind[k] < 0 || ind[k] >= n) return fail(nullptr, UFE_ERR_ARGUMENT, "column index out of range");
  return guarded(nullptr, [&]() -> int {
    hipStream_t s;
    UFE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DBuf dp(sizeof(int) * (m + 1)), di(sizeof(int) * std::max(nnz, 1)), dv(sizeof(double) * std::max(nnz, 1)),
        dx(sizeof(double) * std::max(n, 1)), dy(sizeof(double) * std::max(m, 1));
    UFE_HIP(hipMemcpy(dp.p, ptr, sizeof(int) * (m + 1), hipMemcpyHostToDevice));
    if (nnz) {
      UFE_HIP(hipMemcpy(di.p, ind, sizeof(int) * nnz, hipMemcpyHostToDevice));
      UFE_HIP(hipMemcpy(dv.p, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
    if (n) UFE_HIP(hipMemcpy(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
    if (m) launch_csr_spmv64(m, dp.as<int>(), di.as<int>(), dv.as<double>(), dx.as<double>(), dy.as<double>(), s);
    UFE_HIP(hipStreamSynchronize(s));
    if (m) UFE_HIP(hipMemcpy(y, dy.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    UFE_HIP(hipStreamDestroy(s));
    return UFE_OK;
  });
}

// Solve A x = b for a general square CSR matrix: rows/columns are paired into
// 2x2 blocks (explicit zeros where needed) and the BSR BiCGStab is used.
extern "C" int ufe_csr_solve_ex(int n, int nnz, const int *ptr, const int *ind, const double *val, const double *b,
                                double *x, double rtol, double abstol, int maxit, int precond, int krylov,
                                int restart, int *n_its, double *rel_residual);

extern "C" int ufe_csr_solve(int n, int nnz, const int *ptr, const int *ind, const double *val, const double *b,
                             double *x, double rtol, double abstol, int maxit, int precond, int *n_its,
                             double *rel_residual) {
  return ufe_csr_solve_ex(n, nnz, ptr, ind, val, b, x, rtol, abstol, maxit, precond, UFE_KRYLOV_BICGSTAB, 30, n_its,
                          rel_residual);
}

extern "C" int ufe_csr_solve_ex(int n, int nnz, const int *ptr, const int *ind, const double *val, const double *b,
                                double *x, double rtol, double abstol, int maxit, int precond, int krylov,
                                int restart, int *n_its, double *rel_residual) {
  if (krylov != UFE_KRYLOV_BICGSTAB && krylov != UFE_KRYLOV_GMRES)
    return fail(nullptr, UFE_ERR_ARGUMENT, "krylov must be UFE_KRYLOV_BICGSTAB or UFE_KRYLOV_GMRES");
  if (krylov == UFE_KRYLOV_GMRES && (restart < 1 || restart > 64))
    return fail(nullptr, UFE_ERR_ARGUMENT, "restart must be 1..64");
  if (n <= 0 || nnz < 0 || !ptr || !ind || !val || !b || !x) return fail(nullptr, UFE_ERR_ARGUMENT, "invalid CSR arguments");
  if (ptr[0] != 0 || ptr[n] != nnz) return fail(nullptr, UFE_ERR_ARGUMENT, "bad row pointer");
  for (int i = 0; i < n; ++i)
    if (ptr[i + 1] < ptr[i]) return fail(nullptr, UFE_ERR_ARGUMENT, "row pointer not monotone");
  for (int k = 0; k < nnz; ++k)
    if (ind[k] < 0 || ind[k] >= n) return fail(nullptr, UFE_ERR_ARGUMENT, "column index out of range");
  if (precond == UFE_PC_AMG) return fail(nullptr, UFE_ERR_UNSUPPORTED, "AMG needs the mesh context");
  if (maxit < 1) return fail(nullptr, UFE_ERR_ARGUMENT, "maxit must be positive");
  return guarded(nullptr, [&]() -> int {
    int nb = (n + 1) / 2;
    std::vector<std::vector<std::pair<int, std::array<double, 4>>>> rows(nb);
    auto slot = [&](int I, int J) -> std::array<double, 4> & {
      for (auto &e : rows[I]) if (e.first == J) return e.second;
      rows[I].push_back({J, {0.0, 0.0, 0.0, 0.0}});
      return rows[I].back().second;
    };
    for (int i = 0; i < n; ++i)
      for (int k = ptr[i]; k < ptr[i + 1]; ++k) slot(i / 2, ind[k] / 2)[2 * (i % 2) + ind[k] % 2] += val[k];
    std::vector<int> bptr(nb + 1, 0), bcol, bdiag(nb, -1);
    std::vector<double> bval, dinv(4 * (size_t)nb, 0.0);
    for (int I = 0; I < nb; ++I) {
      auto &rw = rows[I];
      slot(I, I);  // diagonal block always present
      std::sort(rw.begin(), rw.end(), [](const auto &a, const auto &c) { return a.first < c.first; });
      if (n % 2 == 1 && I == nb - 1) slot(I, I)[3] = 1.0;  // padding unknown: identity row
      for (auto &e : rw) {
        if (e.first == I) bdiag[I] = (int)bcol.size();
        bcol.push_back(e.first);
        bval.insert(bval.end(), e.second.begin(), e.second.end());
      }
      bptr[I + 1] = (int)bcol.size();
      const double *d = &bval[4 * (size_t)bdiag[I]];
      double det = d[0] * d[3] - d[1] * d[2];
      if (precond == UFE_PC_BLOCK_JACOBI && det == 0.0)
        throw std::runtime_error("singular diagonal block; use a different preconditioner");
      if (precond == UFE_PC_JACOBI && (d[0] == 0.0 || d[3] == 0.0))
        throw std::runtime_error("zero diagonal entry; use a different preconditioner");
      if (det != 0.0) {
        dinv[4 * I + 0] = d[3] / det; dinv[4 * I + 1] = -d[1] / det;
        dinv[4 * I + 2] = -d[2] / det; dinv[4 * I + 3] = d[0] / det;
      }
    }
    hipStream_t s;
    UFE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<double> bb(2 * (size_t)nb, 0.0), xx(2 * (size_t)nb, 0.0);
    std::copy(b, b + n, bb.begin());
    std::copy(x, x + n, xx.begin());
    int out_rc = UFE_OK;
    {
      DBuf dptr(sizeof(int) * bptr.size()), dcol(sizeof(int) * std::max<size_t>(bcol.size(), 1)),
          ddiag(sizeof(int) * nb), dval(sizeof(double) * std::max<size_t>(bval.size(), 4)),
          ddinv(sizeof(double) * dinv.size()), db(sizeof(double) * bb.size()), dx(sizeof(double) * xx.size());
      UFE_HIP(hipMemcpy(dptr.p, bptr.data(), dptr.bytes, hipMemcpyHostToDevice));
      UFE_HIP(hipMemcpy(dcol.p, bcol.data(), sizeof(int) * bcol.size(), hipMemcpyHostToDevice));
      UFE_HIP(hipMemcpy(ddiag.p, bdiag.data(), ddiag.bytes, hipMemcpyHostToDevice));
      UFE_HIP(hipMemcpy(dval.p, bval.data(), sizeof(double) * bval.size(), hipMemcpyHostToDevice));
      UFE_HIP(hipMemcpy(ddinv.p, dinv.data(), ddinv.bytes, hipMemcpyHostToDevice));
      UFE_HIP(hipMemcpy(db.p, bb.data(), db.bytes, hipMemcpyHostToDevice));
      UFE_HIP(hipMemcpy(dx.p, xx.data(), dx.bytes, hipMemcpyHostToDevice));
      Krylov K(nb, s);
      K.set_precond(precond);
      K.set_diag(ddiag.as<int>());
      BsrT<double> A;
      A.nb = nb; A.nnzb = (int)bcol.size(); A.ptr = dptr.as<int>(); A.col = dcol.as<int>();
      A.val = dval.as<double4>(); A.dinv = ddinv.as<double4>();
      SolveStats st = krylov == UFE_KRYLOV_GMRES
                          ? K.gmres64(A, db.as<double>(), dx.as<double>(), rtol, abstol, maxit, false, restart, 8)
                          : K.bicgstab64(A, db.as<double>(), dx.as<double>(), rtol, abstol, maxit, false, 8);
      UFE_HIP(hipMemcpy(xx.data(), dx.p, dx.bytes, hipMemcpyDeviceToHost));
      if (n_its) *n_its = st.its;
      if (rel_residual) *rel_residual = st.rel;
      if (st.status != 1) out_rc = UFE_WARN_NOT_CONVERGED;
    }
    std::copy(xx.begin(), xx.begin() + n, x);
    UFE_HIP(hipStreamDestroy(s));
    return out_rc;
  });
}

// Nonlinear SSA residual at a given velocity: r = b(u) - A(u) u with the
// resident inputs (the linearised system of solve_linearised_SSA_DIVA.f90
// assembled at u), interleaved rows 2*ti+uv.
extern "C" int ufe_ssa_residual(ufe_ctx *c, const double *u_b, const double *v_b, double *r, double *b,
                                double *rel_out) {
  if (!c || !u_b || !v_b) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    if (!c->inputs_ready) return fail(c, UFE_ERR_ARGUMENT, "inputs not uploaded");
    interleave_up(c, c->u, u_b, v_b);
    launch_driving_stress(c->m, c->Hi.as<double>(), c->Hs.as<double>(), c->tdx.as<double>(), c->tdy.as<double>(),
                          c->s);
    launch_vertical_average(c->m.nV, c->m.nz, c->m.zeta, c->A_vk_vf.as<double>(), c->A_vav.as<double>(), c->s);
    picard_state(c, false, c->prm.epsilon_sq_0, false);
    double rel = rel_residual(c);
    if (rel_out) *rel_out = rel;
    int nT = c->m.nT;
    std::vector<double> h(2 * (size_t)nT);
    auto down = [&](const DBuf &src, double *dst) {
      UFE_HIP(hipMemcpy(h.data(), src.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < nT; ++i) {
        dst[2 * c->tperm[i]] = h[2 * i];
        dst[2 * c->tperm[i] + 1] = h[2 * i + 1];
      }
    };
    if (r) down(c->r64, r);
    if (b) down(c->rhs, b);
    return UFE_OK;
  });
}

extern "C" int ufe_diva_residual(ufe_ctx *c, const double *u_b, const double *v_b, const double *eta_3D_b,
                                 const double *tau_bx_b, const double *tau_by_b, const double *u_base_b,
                                 const double *v_base_b, double *r, double *b, double *rel_out, double *eta_out,
                                 double *F1_out, double *F2s_out, double *beta_out, double *beta_eff_out) {
  if (!c || !u_b || !v_b) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  if ((!tau_bx_b) != (!tau_by_b) || (!u_base_b) != (!v_base_b))
    return fail(c, UFE_ERR_ARGUMENT, "tau_b and u_base need both components or neither");
  return guarded(c, [&]() -> int {
    UFE_HIP(hipSetDevice(c->device));
    if (!c->inputs_ready) return fail(c, UFE_ERR_ARGUMENT, "inputs not uploaded");
    const int nT = c->m.nT, nz = c->m.nz;
    int rc = ufe_upload_diva_state(c, eta_3D_b, tau_bx_b, tau_by_b, u_base_b, v_base_b);
    if (rc < 0) return rc;
    interleave_up(c, c->u, u_b, v_b);
    hipStream_t s = c->s;
    launch_driving_stress(c->m, c->Hi.as<double>(), c->Hs.as<double>(), c->tdx.as<double>(), c->tdy.as<double>(), s);
    launch_flow_factor_pow((long)c->m.nV * nz, c->A_vk_lf.as<double>(), c->prm.glens_n, c->Apow_lf.as<double>(), s);
    picard_state(c, true, c->prm.epsilon_sq_0, true);
    double rel = rel_residual(c);
    if (rel_out) *rel_out = rel;
    std::vector<double> h(2 * (size_t)nT);
    auto down2 = [&](const DBuf &src, double *dst) {
      UFE_HIP(hipMemcpy(h.data(), src.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int i = 0; i < nT; ++i) {
        dst[2 * c->tperm[i]] = h[2 * i];
        dst[2 * c->tperm[i] + 1] = h[2 * i + 1];
      }
    };
    if (r) down2(c->r64, r);
    if (b) down2(c->rhs, b);
    auto down3 = [&](const DBuf &src_tk, double *dst) {  // [triangle][k] -> caller's [k][triangle]
      launch_transpose(nT, nz, src_tk.as<double>(), c->tmp3.as<double>(), s);
      UFE_HIP(hipStreamSynchronize(s));
      std::vector<double> t((size_t)nT * nz);
      UFE_HIP(hipMemcpy(t.data(), c->tmp3.p, t.size() * sizeof(double), hipMemcpyDeviceToHost));
      from_dev(c->tperm, t.data(), dst, nz);
    };
    if (eta_out) down3(c->eta3b, eta_out);
    if (F1_out) down3(c->F1b, F1_out);
    std::vector<double> t1(nT);
    if (F2s_out) {
      UFE_HIP(hipMemcpy(t1.data(), c->F2b0.p, sizeof(double) * nT, hipMemcpyDeviceToHost));
      from_dev(c->tperm, t1.data(), F2s_out);
    }
    if (beta_out) {
      UFE_HIP(hipMemcpy(t1.data(), c->beta_basal_b.p, sizeof(double) * nT, hipMemcpyDeviceToHost));
      from_dev(c->tperm, t1.data(), beta_out);
    }
    if (beta_eff_out) {
      std::vector<double4> tc(nT);
      UFE_HIP(hipMemcpy(tc.data(), c->tri_coef.p, sizeof(double4) * nT, hipMemcpyDeviceToHost));
      for (int i = 0; i < nT; ++i) t1[i] = tc[i].w;
      from_dev(c->tperm, t1.data(), beta_eff_out);
    }
    return UFE_OK;
  });
}

// ---------------------------------------------------------------------------
// Learned initial guess: per-triangle features from the uploaded inputs and
// the network's output written into the iterate (mpnn.hip runs the network).
// Features (fp32, 16 per triangle, unused ones zero): H/1e3, tau_d/1e5 (x, y),
// (log10 beta_sq - 4)/2, log10 A + 16.5, grounded fraction, border flag,
// centroid scaled to the mesh's bounding box (x, y), Hs/1e3, |tau_d|/1e5,
// log10 |tau_d| - 4, log10 H - 3, unit driving-stress direction (x, y),
// log10 of a local shallow-ice speed estimate.
// The network predicts log10 speed and the flow direction.
// ---------------------------------------------------------------------------
constexpr int kFeat = 16;

__global__ void __launch_bounds__(kBlock) k_features(int nT, const int *__restrict__ Tri, const double *__restrict__ V,
                                                     const double *__restrict__ Hi, const double *__restrict__ Hs,
                                                     const double *__restrict__ rough, const double *__restrict__ A_vav,
                                                     const double *__restrict__ fgr, const double *__restrict__ tdx,
                                                     const double *__restrict__ tdy,
                                                     const signed char *__restrict__ rowtype, double x0, double sx,
                                                     double y0, double sy, double glen_n, double weertman_m,
                                                     float *__restrict__ X) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  double H = 0, S = 0, B = 0, A = 0, cx = 0, cy = 0;
  for (int n = 0; n < 3; ++n) {
    int vi = Tri[3 * ti + n];
    H += Hi[vi]; S += Hs[vi]; B += rough[vi]; A += A_vav[vi];
    cx += V[2 * vi]; cy += V[2 * vi + 1];
  }
  H /= 3; S /= 3; B /= 3; A /= 3; cx /= 3; cy /= 3;
  float *x = X + (size_t)ti * kFeat;
  x[0] = (float)(H / 1e3);
  x[1] = (float)(tdx[ti] / 1e5);
  x[2] = (float)(tdy[ti] / 1e5);
  x[3] = (float)((log10(fmax(B, 1e-3)) - 4.0) / 2.0);
  x[4] = (float)(log10(fmax(A, 1e-30)) + 16.5);
  x[5] = (float)fgr[ti];
  x[6] = rowtype[2 * ti] != ROW_FREE ? 1.0f : 0.0f;
  x[7] = (float)((cx - x0) * sx);
  x[8] = (float)((cy - y0) * sy);
  x[9] = (float)(S / 1e3);
  const double tn = sqrt(tdx[ti] * tdx[ti] + tdy[ti] * tdy[ti]);
  x[10] = (float)(tn / 1e5);
  x[11] = (float)(log10(fmax(tn, 1.0)) - 4.0);
  x[12] = (float)(log10(fmax(H, 1.0)) - 3.0);
  x[13] = tn > 0.0 ? (float)(tdx[ti] / tn) : 0.0f;
  x[14] = tn > 0.0 ? (float)(tdy[ti] / tn) : 0.0f;
  // local shallow-ice estimate of the speed: deformation 2A/(n+1) |tau_d|^n H
  // plus Weertman sliding (|tau_d| / beta_sq)^m, as log10 m/yr
  const double u_sia = 2.0 * A / (glen_n + 1.0) * pow(tn, glen_n) * H + pow(tn / fmax(B, 1e-3), weertman_m);
  x[15] = (float)log10(fmax(u_sia, 1e-3));
}

// velocity = 10^y0 * (y1, y2) / |(y1, y2)|  (speed capped at vel_max)
__global__ void __launch_bounds__(kBlock) k_set_guess(int nT, const float *__restrict__ Y, int ystride,
                                                      double vel_max, double *__restrict__ u) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  const float *y = Y + (size_t)ti * ystride;
  double sp = fmin(pow(10.0, fmin((double)y[0], 6.0)), vel_max);
  double dx = y[1], dy = y[2], dn = sqrt(dx * dx + dy * dy);
  if (!(dn > 1e-6)) { dx = 0.0; dy = 0.0; dn = 1.0; }
  reinterpret_cast<double2 *>(u)[ti] = make_double2(sp * dx / dn, sp * dy / dn);
}

namespace ufe {
namespace mp {

// Features of the uploaded inputs into X_dev [nT][16] (device order) and the
// triangle neighbour lists into nb_dev [nT][3]; with null pointers only N is
// reported.
int ctx_features(ufe_ctx *c, float *X_dev, int *nb_dev, hipStream_t *s_out, int *N_out) {
  *N_out = c->m.nT;
  *s_out = c->s;
  if (!X_dev) return UFE_OK;
  if (!c->inputs_ready) return fail(c, UFE_ERR_ARGUMENT, "inputs not uploaded");
  UFE_HIP(hipSetDevice(c->device));
  hipStream_t s = c->s;
  int nT = c->m.nT;
  launch_driving_stress(c->m, c->Hi.as<double>(), c->Hs.as<double>(), c->tdx.as<double>(), c->tdy.as<double>(), s);
  launch_vertical_average(c->m.nV, c->m.nz, c->m.zeta, c->A_vk_vf.as<double>(), c->A_vav.as<double>(), s);
  double sx = 1.0 / std::max(c->bbox[1] - c->bbox[0], 1e-30), sy = 1.0 / std::max(c->bbox[3] - c->bbox[2], 1e-30);
  hipLaunchKernelGGL(k_features, dim3(blocks_for(nT)), dim3(kBlock), 0, s, nT, c->m.Tri, c->m.V, c->Hi.as<double>(),
                     c->Hs.as<double>(), c->rough.as<double>(), c->A_vav.as<double>(), c->fgr.as<double>(),
                     c->tdx.as<double>(), c->tdy.as<double>(), c->rowtype.as<signed char>(), c->bbox[0], sx,
                     c->bbox[2], sy, c->prm.glens_n, c->prm.slid_weertman_m, X_dev);
  UFE_LAUNCH_CHECK();
  if (nb_dev) UFE_HIP(hipMemcpyAsync(nb_dev, c->m.TriC, sizeof(int) * 3 * (size_t)nT, hipMemcpyDeviceToDevice, s));
  return UFE_OK;
}

// u := network output (km/yr -> m/yr); DIVA state cleared.
int ctx_set_guess(ufe_ctx *c, const float *Y_dev) {
  UFE_HIP(hipSetDevice(c->device));
  int nT = c->m.nT;
  hipLaunchKernelGGL(k_set_guess, dim3(blocks_for(nT)), dim3(kBlock), 0, c->s, nT, Y_dev, 16, c->prm.vel_max,
                     c->u.as<double>());
  UFE_LAUNCH_CHECK();
  UFE_HIP(hipMemsetAsync(c->eta3b.p, 0, c->eta3b.bytes, c->s));
  UFE_HIP(hipMemsetAsync(c->ubase.p, 0, c->ubase.bytes, c->s));
  UFE_HIP(hipMemsetAsync(c->taub.p, 0, c->taub.bytes, c->s));
  c->diva_state_ready = true;
  UFE_HIP(hipStreamSynchronize(c->s));
  return UFE_OK;
}

}  // namespace mp
}  // namespace ufe

extern "C" int ufe_mpnn_features(ufe_ctx *c, float *X) {
  if (!c || !X) return fail(c, UFE_ERR_ARGUMENT, "null argument");
  return guarded(c, [&]() -> int {
    int N;
    hipStream_t s;
    int nT = c->m.nT;
    DBuf xd(sizeof(float) * kFeat * (size_t)nT);
    int rc = ufe::mp::ctx_features(c, xd.as<float>(), nullptr, &s, &N);
    if (rc != UFE_OK) return rc;
    std::vector<float> h((size_t)kFeat * nT);
    UFE_HIP(hipMemcpyAsync(h.data(), xd.p, h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nT; ++i)
      std::memcpy(X + (size_t)kFeat * c->tperm[i], h.data() + (size_t)kFeat * i, sizeof(float) * kFeat);
    return UFE_OK;
  });
}
