// Per-Picard-iteration kernels of the SSA/DIVA velocity solve.
//
// Data layout in HBM (all indices 0-based, int32):
//   b->a operators   CSR, nV rows (pattern shared by map/ddx/ddy)
//   a->b operators   CSR, nT rows (pattern shared by map/ddx/ddy)
//   M2 b->b          CSR with columns sorted per row; entry e of row ti is
//                    block e of block-row ti of the system matrix, so the
//                    assembly writes block e from M2 entry e with no lookup.
//   system matrix    BSR, 2x2 blocks [uu, uv, vu, vv] per entry, unknown
//                    order 2*ti+uv (mesh_translation_tables.f90:188-198)
//   3-D fields       layer-fastest on the device ([vertex][k]) so one thread
//                    walks a column with unit stride
#include "kernels.h"

namespace ufe {

// ---------------------------------------------------------------------------
// SSA vertex kernel: one thread per vertex, one pass over its b->a row.
//   calc_horizontal_strain_rates   (SSA_DIVA_utilities.f90:59-82)
//   calc_effective_viscosity_Glen_2D + clamp + N = eta*max(0.1,H)
//                                  (SSA_main.f90:356-382, constitutive_equation.f90:25-51)
//   calc_basal_friction_coefficient (sliding_laws.f90:25-130): u_a = map_b_a(u)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_vertex_ssa(
    int nV, const int *__restrict__ ptr, const int *__restrict__ ind, const double *__restrict__ wmap,
    const double *__restrict__ wdx, const double *__restrict__ wdy, const double *__restrict__ x,
    const double *__restrict__ A_vav, const double *__restrict__ Hi, const double *__restrict__ rough,
    PicardParams p, double *__restrict__ N_a, double *__restrict__ beta_a) {
  int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi >= nV) return;
  double ux = 0, uy = 0, vx = 0, vy = 0, ua = 0, va = 0;
  for (int k = ptr[vi]; k < ptr[vi + 1]; ++k) {
    int t = ind[k];
    double2 uv = reinterpret_cast<const double2 *>(x)[t];
    double cx = wdx[k], cy = wdy[k], cm = wmap[k];
    ux += cx * uv.x; uy += cy * uv.x; vx += cx * uv.y; vy += cy * uv.y;
    ua += cm * uv.x; va += cm * uv.y;
  }
  double e2 = ux * ux + vy * vy + ux * vy + 0.25 * (uy + vx) * (uy + vx) + p.eps0;
  double eta = 0.5 * pow(A_vav[vi], -1.0 / p.n) * pow(e2, (1.0 - p.n) / (2.0 * p.n));
  eta = fmin(fmax(eta, p.visc_eff_min), p.eta_max);
  N_a[vi] = eta * fmax(0.1, Hi[vi]);
  beta_a[vi] = basal_friction(p.sl, vi, rough[vi], ua, va);
}

// ---------------------------------------------------------------------------
// Triangle kernel: one thread per triangle, one pass over its a->b row.
//   N_b, dN/dx_b, dN/dy_b = map/ddx/ddy_a_b(N_a)         (SSA_main.f90:385-387)
//   beta_b = map_a_b(beta_a) * f_gr^p                      (SSA_main.f90:418-425)
//   DIVA: beta_eff_b = map_a_b(beta_eff_a) * f_gr^p, basal beta_b unscaled
//         (DIVA_main.f90:559-569)
//   right-hand side b of both rows (solve_linearised_SSA_DIVA.f90:94-147)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_triangle(
    int nT, const int *__restrict__ ptr, const int *__restrict__ ind, const double *__restrict__ wmap,
    const double *__restrict__ wdx, const double *__restrict__ wdy, const double *__restrict__ N_a,
    const double *__restrict__ beta_a, const double *__restrict__ beta_eff_a,
    const double *__restrict__ fgr, const double *__restrict__ tdx, const double *__restrict__ tdy,
    const signed char *__restrict__ rowtype, const double *__restrict__ bc_u,
    const double *__restrict__ bc_v, PicardParams p, double4 *__restrict__ tri_coef,
    double *__restrict__ beta_basal_b, double *__restrict__ rhs) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  double N = 0, Nx = 0, Ny = 0, be = 0, bf = 0;
  for (int k = ptr[ti]; k < ptr[ti + 1]; ++k) {
    int v = ind[k];
    double n = N_a[v], cm = wmap[k];
    N += cm * n; Nx += wdx[k] * n; Ny += wdy[k] * n;
    be += cm * beta_a[v];
    if (beta_eff_a) bf += cm * beta_eff_a[v];
  }
  double scale = p.do_subgrid ? pow(fgr[ti], p.subgrid_exp) : 1.0;
  double beta_sys;
  if (beta_eff_a) {
    beta_basal_b[ti] = be;
    beta_sys = bf * scale;
  } else {
    beta_sys = be * scale;
    if (beta_basal_b) beta_basal_b[ti] = beta_sys;
  }
  tri_coef[ti] = make_double4(N, Nx, Ny, beta_sys);
  for (int uv = 0; uv < 2; ++uv) {
    int rt = rowtype[2 * ti + uv];
    double r;
    if (rt == ROW_DIRICHLET || rt == ROW_COPY) r = uv == 0 ? bc_u[ti] : bc_v[ti];
    else if (rt != ROW_FREE) r = 0.0;
    else {
      double t = uv == 0 ? tdx[ti] : tdy[ti];
      r = p.crossterms ? -t : -t / N;
    }
    rhs[2 * ti + uv] = r;
  }
}

// ---------------------------------------------------------------------------
// Assembly: one thread per block entry, atomic-free (each block has one
// writer).  Rows follow calc_SSA_DIVA_stiffness_matrix_row_free (:180-329),
// the "sans" variant (:331-479) and calc_SSA_DIVA_stiffness_matrix_row_BC
// (:481-641, 'infinite' and 'zero'), Dirichlet rows (:94-107).
//
// The system is stored left-scaled by the inverse of its diagonal blocks,
// S = D^-1 A and b_s = D^-1 b (block-Jacobi row scaling): every block row then
// has the identity on its diagonal, whatever the physical scale of the row
// (momentum rows ~ N/h^2, boundary rows ~ 1), which is what the aggregation
// coarse correction needs.  D itself is kept so that residuals are reported
// in the unscaled norm (r = D r_s).  The solution is unchanged by the scaling.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void system_block(int ti, int tj, const signed char *__restrict__ rowtype,
                                             const int *__restrict__ TriC, double4 c, int crossterms, double xx,
                                             double xy, double yy, double dx, double dy, double blk[4]) {
  const bool diag = (ti == tj);
  const double N = c.x, Nx = c.y, Ny = c.z, be = c.w;
  for (int uv = 0; uv < 2; ++uv) {
    int rt = rowtype[2 * ti + uv];
    double a_same, a_other;  // coefficient on the same / the other component of tj
    if (rt == ROW_FREE) {
      if (crossterms) {
        if (uv == 0) {
          a_same = 4.0 * N * xx + 4.0 * Nx * dx + N * yy + Ny * dy;
          a_other = 3.0 * N * xy + 2.0 * Nx * dy + Ny * dx;
        } else {
          a_same = 4.0 * N * yy + 4.0 * Ny * dy + N * xx + Nx * dx;
          a_other = 3.0 * N * xy + 2.0 * Ny * dx + Nx * dy;
        }
        if (diag) a_same -= be;
      } else {
        a_same = uv == 0 ? 4.0 * xx + yy : 4.0 * yy + xx;
        a_other = 3.0 * xy;
        if (diag) a_same -= be / N;
      }
    } else if (rt == ROW_INFINITE) {
      int nn = 0;
      bool nb = false;
      for (int q = 0; q < 3; ++q) {
        int t = TriC[3 * ti + q];
        if (t >= 0) { ++nn; nb |= (t == tj); }
      }
      a_same = diag ? -(double)nn : (nb ? 1.0 : 0.0);
      a_other = 0.0;
    } else {  // ROW_ZERO, ROW_DIRICHLET, ROW_COPY
      a_same = diag ? 1.0 : 0.0;
      a_other = 0.0;
    }
    if (uv == 0) { blk[0] = a_same; blk[1] = a_other; }
    else { blk[2] = a_other; blk[3] = a_same; }
  }
}

template <bool kMirror32>
__global__ void __launch_bounds__(kBlock) k_assemble(
    int nnzb, const int *__restrict__ row, const int *__restrict__ col, const int *__restrict__ diagpos,
    const double *__restrict__ Mx, const double *__restrict__ My, const double *__restrict__ Mxx,
    const double *__restrict__ Mxy, const double *__restrict__ Myy,
    const signed char *__restrict__ rowtype, const int *__restrict__ TriC,
    const double4 *__restrict__ tri_coef, int crossterms, const double2 *__restrict__ rhs,
    double4 *__restrict__ S, float4 *__restrict__ S32, double4 *__restrict__ Dblk, float4 *__restrict__ Dblk32,
    double4 *__restrict__ Dinv, float4 *__restrict__ Dinv32, double2 *__restrict__ rhs_s) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnzb) return;
  int ti = row[e], tj = col[e];
  double4 c = tri_coef[ti];
  double blk[4], dg[4];
  system_block(ti, tj, rowtype, TriC, c, crossterms, Mxx[e], Mxy[e], Myy[e], Mx[e], My[e], blk);
  int d = diagpos[ti];
  system_block(ti, ti, rowtype, TriC, c, crossterms, Mxx[d], Mxy[d], Myy[d], Mx[d], My[d], dg);
  double det = dg[0] * dg[3] - dg[1] * dg[2], id = 1.0 / det;
  double i0 = dg[3] * id, i1 = -dg[1] * id, i2 = -dg[2] * id, i3 = dg[0] * id;
  double s0 = i0 * blk[0] + i1 * blk[2], s1 = i0 * blk[1] + i1 * blk[3];
  double s2 = i2 * blk[0] + i3 * blk[2], s3 = i2 * blk[1] + i3 * blk[3];
  if (ti == tj) { s0 = 1.0; s1 = 0.0; s2 = 0.0; s3 = 1.0; }
  S[e] = make_double4(s0, s1, s2, s3);
  if (kMirror32) S32[e] = make_float4((float)s0, (float)s1, (float)s2, (float)s3);
  if (ti == tj) {
    Dblk[ti] = make_double4(dg[0], dg[1], dg[2], dg[3]);
    Dinv[ti] = make_double4(1.0, 0.0, 0.0, 1.0);
    if (kMirror32) {
      Dinv32[ti] = make_float4(1.0f, 0.0f, 0.0f, 1.0f);
      Dblk32[ti] = make_float4((float)dg[0], (float)dg[1], (float)dg[2], (float)dg[3]);
    }
    double2 b = rhs[ti];
    rhs_s[ti] = make_double2(i0 * b.x + i1 * b.y, i2 * b.x + i3 * b.y);
  }
}

// ---------------------------------------------------------------------------
// Driving stress (SSA_DIVA_utilities.f90:21-57), once per solve.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_driving_stress(
    int nT, const int *__restrict__ ptr, const int *__restrict__ ind, const double *__restrict__ wmap,
    const double *__restrict__ wdx, const double *__restrict__ wdy, const double *__restrict__ Hi,
    const double *__restrict__ Hs, double *__restrict__ tdx, double *__restrict__ tdy) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  double hb = 0, sx = 0, sy = 0;
  for (int k = ptr[ti]; k < ptr[ti + 1]; ++k) {
    int v = ind[k];
    hb += wmap[k] * Hi[v];
    sx += wdx[k] * Hs[v];
    sy += wdy[k] * Hs[v];
  }
  tdx[ti] = -kIceDensity * kGrav * hb * sx;
  tdy[ti] = -kIceDensity * kGrav * hb * sy;
}

// Vertical average of A (vertex-fastest input, calc_vertically_averaged_flow_parameter,
// SSA_main.f90:312-337 with vertical_average, mesh_zeta.f90:257-283).
__global__ void __launch_bounds__(kBlock) k_vertical_average(int nV, int nz, const double *__restrict__ zeta,
                                                             const double *__restrict__ f_vk,
                                                             double *__restrict__ out) {
  int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi >= nV) return;
  double s = 0.0;
  for (int k = 0; k < nz - 1; ++k)
    s += 0.5 * (f_vk[(size_t)(k + 1) * nV + vi] + f_vk[(size_t)k * nV + vi]) * (zeta[k + 1] - zeta[k]);
  out[vi] = s;
}

// ---------------------------------------------------------------------------
// Post-solve update: apply_velocity_limits (SSA_DIVA_utilities.f90:110-141),
// relax_viscosity_iterations (:84-108) and the calc_L2_norm_uv sums
// (:143-184) in one pass, FP64 reduction.  x holds the new solution
// (interleaved u,v); u is the iterate (interleaved) updated in place; uprev
// receives the previous iterate.  DIVA additionally forms the basal
// velocities and basal shear stress (DIVA_main.f90:576-634).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_update(
    int nT, const double *__restrict__ x, double *__restrict__ u, double *__restrict__ uprev,
    double vel_max, double relax, RedSlot slot, double *__restrict__ out2,
    const double *__restrict__ beta_basal_b, const double4 *__restrict__ tri_coef,
    const double *__restrict__ F2b0, int no_sliding, double *__restrict__ ubase,
    double *__restrict__ taub) {
  double acc[2] = {0.0, 0.0};
  for (int ti = blockIdx.x * blockDim.x + threadIdx.x; ti < nT; ti += gridDim.x * blockDim.x) {
    double2 xn = reinterpret_cast<const double2 *>(x)[ti];
    double2 up = reinterpret_cast<const double2 *>(u)[ti];
    double ua = sqrt(xn.x * xn.x + xn.y * xn.y);
    if (ua > vel_max) { xn.x = xn.x * vel_max / ua; xn.y = xn.y * vel_max / ua; }
    double2 un = make_double2(relax * xn.x + (1.0 - relax) * up.x, relax * xn.y + (1.0 - relax) * up.y);
    reinterpret_cast<double2 *>(u)[ti] = un;
    reinterpret_cast<double2 *>(uprev)[ti] = up;
    double dx = un.x - up.x, dy = un.y - up.y, sx = un.x + up.x, sy = un.y + up.y;
    acc[0] += dx * dx + dy * dy;
    acc[1] += sx * sx + sy * sy;
    if (ubase) {
      if (no_sliding) reinterpret_cast<double2 *>(ubase)[ti] = make_double2(0.0, 0.0);
      else {
        double d = 1.0 + beta_basal_b[ti] * F2b0[ti];
        reinterpret_cast<double2 *>(ubase)[ti] = make_double2(un.x / d, un.y / d);
      }
      double bse = tri_coef[ti].w;
      reinterpret_cast<double2 *>(taub)[ti] = make_double2(un.x * bse, un.y * bse);
    }
  }
  __shared__ double lds[2][kBlock / 64];
  block_sum<2>(acc, lds);
  block_partials<2>(slot, acc);
}

// ---------------------------------------------------------------------------
// Host wrappers
// ---------------------------------------------------------------------------
void launch_vertex_ssa(const DevMesh &m, const double *x, const double *A_vav, const double *Hi,
                       const double *rough, const PicardParams &p, double *N_a, double *beta_a,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_vertex_ssa, dim3(blocks_for(m.nV)), dim3(kBlock), 0, s, m.nV, m.ba_ptr, m.ba_ind,
                     m.ba_map, m.ba_ddx, m.ba_ddy, x, A_vav, Hi, rough, p, N_a, beta_a);
  UFE_LAUNCH_CHECK();
}

void launch_triangle(const DevMesh &m, const double *N_a, const double *beta_a, const double *beta_eff_a,
                     const double *fgr, const double *tdx, const double *tdy, const signed char *rowtype,
                     const double *bc_u, const double *bc_v, const PicardParams &p, double4 *tri_coef,
                     double *beta_basal_b, double *rhs, hipStream_t s) {
  hipLaunchKernelGGL(k_triangle, dim3(blocks_for(m.nT)), dim3(kBlock), 0, s, m.nT, m.ab_ptr, m.ab_ind,
                     m.ab_map, m.ab_ddx, m.ab_ddy, N_a, beta_a, beta_eff_a, fgr, tdx, tdy, rowtype,
                     bc_u, bc_v, p, tri_coef, beta_basal_b, rhs);
  UFE_LAUNCH_CHECK();
}

void launch_assemble(const DevMesh &m, const signed char *rowtype, const double4 *tri_coef, int crossterms,
                     const double *rhs, double4 *S, float4 *S32, double4 *Dblk, float4 *Dblk32, double4 *Dinv,
                     float4 *Dinv32, double *rhs_s, hipStream_t s) {
  const double2 *r2 = reinterpret_cast<const double2 *>(rhs);
  double2 *rs2 = reinterpret_cast<double2 *>(rhs_s);
  if (S32)
    hipLaunchKernelGGL(k_assemble<true>, dim3(blocks_for(m.bb_nnz)), dim3(kBlock), 0, s, m.bb_nnz, m.bb_row,
                       m.bb_ind, m.bb_diag, m.bb_ddx, m.bb_ddy, m.bb_d2dx2, m.bb_d2dxdy, m.bb_d2dy2, rowtype, m.TriC,
                       tri_coef, crossterms, r2, S, S32, Dblk, Dblk32, Dinv, Dinv32, rs2);
  else
    hipLaunchKernelGGL(k_assemble<false>, dim3(blocks_for(m.bb_nnz)), dim3(kBlock), 0, s, m.bb_nnz, m.bb_row,
                       m.bb_ind, m.bb_diag, m.bb_ddx, m.bb_ddy, m.bb_d2dx2, m.bb_d2dxdy, m.bb_d2dy2, rowtype, m.TriC,
                       tri_coef, crossterms, r2, S, S32, Dblk, Dblk32, Dinv, Dinv32, rs2);
  UFE_LAUNCH_CHECK();
}

void launch_driving_stress(const DevMesh &m, const double *Hi, const double *Hs, double *tdx, double *tdy,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_driving_stress, dim3(blocks_for(m.nT)), dim3(kBlock), 0, s, m.nT, m.ab_ptr, m.ab_ind,
                     m.ab_map, m.ab_ddx, m.ab_ddy, Hi, Hs, tdx, tdy);
  UFE_LAUNCH_CHECK();
}

void launch_vertical_average(int nV, int nz, const double *zeta, const double *f_vk, double *out,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_vertical_average, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, nz, zeta, f_vk, out);
  UFE_LAUNCH_CHECK();
}

void launch_update(int nT, const double *x, double *u, double *uprev, double vel_max, double relax,
                   const RedSlot &slot, double *out2, const double *beta_basal_b, const double4 *tri_coef,
                   const double *F2b0, int no_sliding, double *ubase, double *taub, hipStream_t s) {
  int nb = blocks_for(nT);
  if (nb > kMaxPartials) nb = kMaxPartials;
  hipLaunchKernelGGL(k_update, dim3(nb), dim3(kBlock), 0, s, nT, x, u, uprev, vel_max, relax, slot, out2,
                     beta_basal_b, tri_coef, F2b0, no_sliding, ubase, taub);
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_partials_out<2>, dim3(1), dim3(kBlock), 0, s, slot, nb, out2);
  UFE_LAUNCH_CHECK();
}

}  // namespace ufe
