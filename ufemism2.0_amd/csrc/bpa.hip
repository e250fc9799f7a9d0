// BPA 3-D stress balance on the device: solve_BPA (BPA_main.f90:90-252) and
// the zeta gradients it runs on (calc_zeta_gradients, zeta_gradients.f90:19-212).
//
// Layout.  Triangles and vertices are in the context's locality (Morton)
// order.  3-D fields are layer-fastest: bk index ti*nz + k, ak vi*nz + k, bks
// ti*(nz-1) + ks (the reference's tik2n/vik2n/tiks2n order,
// mesh_translation_tables.f90:120-240), so one column is contiguous.  The
// unknowns are (u, v) pairs per bk point: block row ti*nz + k of a BSR matrix
// with 2x2 blocks, the reference's tikuv2n order.
//
// The 3-D operators of calc_3D_matrix_operators_mesh
// (mesh_disc_calc_matrix_operators_3D.f90) are never stored: every product
// with them, and every stiffness entry, is formed on the fly from the 2-D
// operators (a->b, b->a, b->b 2nd order, a->a), the three-point zeta stencils
// and the zeta gradients, which are themselves formed per point from a few
// per-vertex / per-triangle thickness and surface gradients.  Block row
// (ti, k) has the pattern {b->b row of ti} x {three layers of the zeta
// stencil of k}, in that order, so entry p of a row is (b->b entry p / 3,
// layer p % 3) with no lookup table.
//
// The linear solve is GMRES(gmres_restart) (PETSc's default KSP method)
// preconditioned by an exact solve of each triangle's column: the block
// tridiagonal part (layers k-1, k, k+1 of the same triangle) of the system,
// factored once per viscosity iteration.  The vertical coupling (eta/dz^2)
// dominates the horizontal one (eta/dx^2) by (dx/dz)^2, so the column solve
// removes the stiff part and GMRES sees a 2-D problem.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "krylov.h"
#include "sliding.h"
#include "ufe_internal.h"

struct ufe_ctx;
namespace ufe {
int ctx_device(const ufe_ctx *c);
hipStream_t ctx_stream(const ufe_ctx *c);
const DevMesh *ctx_mesh(const ufe_ctx *c);
int ctx_fail(ufe_ctx *c, int code, const std::string &msg);
const std::vector<int> &ctx_vperm(const ufe_ctx *c);
const std::vector<int> &ctx_tperm(const ufe_ctx *c);
std::shared_ptr<void> &ctx_bpa(ufe_ctx *c);
const ufe_params &ctx_params(const ufe_ctx *c);
SlidingLaw ctx_sliding_law(const ufe_ctx *c);
int ctx_guard(ufe_ctx *c, const std::function<int()> &f);
void ctx_copy_lists(ufe_ctx *c, std::vector<int> &crow, std::vector<int> &cptr, std::vector<int> &ctri,
                    std::vector<double> &cw);
const std::vector<int> &ctx_TriBI(const ufe_ctx *c);

namespace {

constexpr int kBpaMaxNz = 32;

// 1-D zeta operators (mesh_zeta.f90:287-482), device tables.  Per layer k the
// three-point stencil runs over layers base(k) .. base(k) + 2 with
// base(k) = min(max(k - 1, 0), nz - 3) (one-sided at the surface and base).
struct ZOps {
  int nz;
  const double *zeta, *zs;     // [nz], [nz-1]
  const double *d, *dd;        // [nz][3] d/dzeta, d2/dzeta2 on layers base(k)+j
  const double *kkm, *kkd;     // [nz-1][2] k -> ks: map, d/dzeta (layers ks, ks+1)
  const int *ksc;              // [nz][2] ks -> k: staggered layers
  const double *ksm, *ksd;     // [nz][2] map, d/dzeta
};

__device__ __forceinline__ int zbase(int k, int nz) { return min(max(k - 1, 0), nz - 3); }

// Per-triangle thickness and surface gradients (zeta_gradients.f90:108-137):
// 0 Hi_b, 1 dHs/dx, 2 dHs/dy, 3 dHi/dx, 4 dHi/dy, 5 d2Hs/dx2, 6 d2Hs/dxdy,
// 7 d2Hs/dy2, 8 d2Hi/dx2, 9 d2Hi/dxdy, 10 d2Hi/dy2 (b grid), 11 d(Hs-Hi)/dx,
// 12 d(Hs-Hi)/dy (the ice base, calc_driving_stress of BPA_main.f90:1739-1742).
constexpr int kZt = 16;

struct ZGrad {
  double zx, zy, zz, zxx, zxy, zyy;
};

// zeta gradients at a b-grid point on level zeta (zeta_gradients.f90:160-196)
__device__ __forceinline__ ZGrad zgrad_tri(const double *__restrict__ zt, double zeta) {
  ZGrad g;
  const double h = fmax(10.0, zt[0]);
  g.zx = (1.0 / h) * (zt[1] - zeta * zt[3]);
  g.zy = (1.0 / h) * (zt[2] - zeta * zt[4]);
  g.zz = -1.0 / h;
  g.zxx = (zt[3] * -1.0 / h) * g.zx + (1.0 / h) * (zt[5] - zeta * zt[8]);
  g.zxy = (zt[4] * -1.0 / h) * g.zx + (1.0 / h) * (zt[6] - zeta * zt[9]);
  g.zyy = (zt[4] * -1.0 / h) * g.zy + (1.0 / h) * (zt[7] - zeta * zt[10]);
  return g;
}

// Glen's law with the u,v strain rates only
// (calc_effective_viscosity_Glen_3D_uv_only, constitutive_equation.f90:53-80)
__device__ __forceinline__ double glen3(double eps0, double ux, double uy, double uz, double vx, double vy,
                                        double vz, double A, double n) {
  const double e2 = ux * ux + vy * vy + ux * vy + 0.25 * (uy + vx) * (uy + vx) + 0.25 * (uz * uz + vz * vz) + eps0;
  return 0.5 * pow(A, -1.0 / n) * pow(e2, (1.0 - n) / (2.0 * n));
}

struct BpaParams {
  double n, eps0_applied, eps0_config, visc_eff_min, eta_max;
  double subgrid_exp;
  double relax_copy;  // C%visc_it_relax of the lagged copy rows
  SlidingLaw sl;
  int do_subgrid;
  int bc_u[4], bc_v[4];
};

// ---------------------------------------------------------------------------
// zeta gradients
// ---------------------------------------------------------------------------

// a->a gradients of Hs and Hi: (dHs/dx, dHs/dy, dHi/dx, dHi/dy) per vertex
__global__ void __launch_bounds__(kBlock) k_zg_vertex(int nV, const int *__restrict__ ptr, const int *__restrict__ ind,
                                                      const double *__restrict__ cx, const double *__restrict__ cy,
                                                      const double *__restrict__ Hi, const double *__restrict__ Hs,
                                                      double4 *__restrict__ zv) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi >= nV) return;
  double sx = 0, sy = 0, ix = 0, iy = 0;
  for (int e = ptr[vi]; e < ptr[vi + 1]; ++e) {
    const int vj = ind[e];
    const double a = cx[e], b = cy[e], s = Hs[vj], h = Hi[vj];
    sx += a * s; sy += b * s; ix += a * h; iy += b * h;
  }
  zv[vi] = make_double4(sx, sy, ix, iy);
}

// per-triangle gradients (kZt layout above) from a->b products of Hi, Hs and
// of the vertex gradients
__global__ void __launch_bounds__(kBlock) k_zg_tri(int nT, const int *__restrict__ ptr, const int *__restrict__ ind,
                                                   const double *__restrict__ wm, const double *__restrict__ wx,
                                                   const double *__restrict__ wy, const double *__restrict__ Hi,
                                                   const double *__restrict__ Hs, const double4 *__restrict__ zv,
                                                   double *__restrict__ zt) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  double o[13] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int e = ptr[ti]; e < ptr[ti + 1]; ++e) {
    const int vi = ind[e];
    const double m = wm[e], x = wx[e], y = wy[e];
    const double h = Hi[vi], s = Hs[vi];
    const double4 g = zv[vi];
    o[0] += m * h;
    o[1] += x * s; o[2] += y * s;
    o[3] += x * h; o[4] += y * h;
    o[5] += x * g.x; o[6] += y * g.x; o[7] += y * g.y;
    o[8] += x * g.z; o[9] += y * g.z; o[10] += y * g.w;
    const double b = s - h;
    o[11] += x * b; o[12] += y * b;
  }
  double *z = zt + (size_t)ti * kZt;
#pragma unroll
  for (int i = 0; i < 13; ++i) z[i] = o[i];
}

// b->a second derivatives of Hs and Hi (only for the ak outputs of the
// zeta-gradient entry point): (d2Hs/dx2, d2Hs/dxdy, d2Hs/dy2, d2Hi/dx2,
// d2Hi/dxdy, d2Hi/dy2) per vertex
__global__ void __launch_bounds__(kBlock) k_zg_vertex2(int nV, const int *__restrict__ ptr,
                                                       const int *__restrict__ ind, const double *__restrict__ wx,
                                                       const double *__restrict__ wy, const double *__restrict__ zt,
                                                       double *__restrict__ out6) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi >= nV) return;
  double o[6] = {0, 0, 0, 0, 0, 0};
  for (int e = ptr[vi]; e < ptr[vi + 1]; ++e) {
    const double *z = zt + (size_t)ind[e] * kZt;
    const double x = wx[e], y = wy[e];
    o[0] += x * z[1]; o[1] += y * z[1]; o[2] += y * z[2];
    o[3] += x * z[3]; o[4] += y * z[3]; o[5] += y * z[4];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) out6[(size_t)vi * 6 + i] = o[i];
}

// The 19 zeta-gradient fields on the caller's layouts (first index fastest,
// device order within a layer): ak 7 x [nz][nV], bk 6 x [nz][nT], bks 6 x [nz-1][nT]
__global__ void __launch_bounds__(kBlock) k_zg_out(int nV, int nT, ZOps z, const double *__restrict__ Hi,
                                                   const double *__restrict__ dHi_dt,
                                                   const double *__restrict__ dHs_dt,
                                                   const double4 *__restrict__ zv, const double *__restrict__ a2,
                                                   const double *__restrict__ zt, double *__restrict__ ak,
                                                   double *__restrict__ bk, double *__restrict__ bks) {
  const int nz = z.nz;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nak = (long)nV * nz, nbk = (long)nT * nz, nbks = (long)nT * (nz - 1);
  if (i < nak) {
    const int k = (int)(i / nV), vi = (int)(i % nV);
    const double h = fmax(10.0, Hi[vi]), zeta = z.zeta[k];
    const double4 g = zv[vi];
    const double *s2 = a2 + (size_t)vi * 6;
    const double zx = (1.0 / h) * (g.x - zeta * g.z), zy = (1.0 / h) * (g.y - zeta * g.w);
    ak[0 * nak + i] = (1.0 / h) * ((dHs_dt ? dHs_dt[vi] : 0.0) - zeta * (dHi_dt ? dHi_dt[vi] : 0.0));
    ak[1 * nak + i] = zx;
    ak[2 * nak + i] = zy;
    ak[3 * nak + i] = -1.0 / h;
    ak[4 * nak + i] = (g.z * -1.0 / h) * zx + (1.0 / h) * (s2[0] - zeta * s2[3]);
    ak[5 * nak + i] = (g.w * -1.0 / h) * zx + (1.0 / h) * (s2[1] - zeta * s2[4]);
    ak[6 * nak + i] = (g.w * -1.0 / h) * zy + (1.0 / h) * (s2[2] - zeta * s2[5]);
  }
  if (i < nbk) {
    const int k = (int)(i / nT), ti = (int)(i % nT);
    const ZGrad g = zgrad_tri(zt + (size_t)ti * kZt, z.zeta[k]);
    bk[0 * nbk + i] = g.zx; bk[1 * nbk + i] = g.zy; bk[2 * nbk + i] = g.zz;
    bk[3 * nbk + i] = g.zxx; bk[4 * nbk + i] = g.zxy; bk[5 * nbk + i] = g.zyy;
  }
  if (i < nbks) {
    const int k = (int)(i / nT), ti = (int)(i % nT);
    const ZGrad g = zgrad_tri(zt + (size_t)ti * kZt, z.zs[k]);
    bks[0 * nbks + i] = g.zx; bks[1 * nbks + i] = g.zy; bks[2 * nbks + i] = g.zz;
    bks[3 * nbks + i] = g.zxx; bks[4 * nbks + i] = g.zxy; bks[5 * nbks + i] = g.zyy;
  }
}

// ---------------------------------------------------------------------------
// viscosity iteration: strain rates, viscosity, friction (BPA_main.f90:1758-1978)
// ---------------------------------------------------------------------------

// One thread per (vertex, layer): horizontal strain rates through bk->ak
// (b->a operators with the zeta chain rule), vertical shear from bk->bks
// mapped bks->ak, eta_ak, and at the base layer the basal friction
// coefficient of the vertex (from map_b_a of the basal velocity).
__global__ void __launch_bounds__(kBlock) k_bpa_vertex(
    int nV, int nT, ZOps z, const int *__restrict__ ptr, const int *__restrict__ ind, const double *__restrict__ wm,
    const double *__restrict__ wx, const double *__restrict__ wy, const double *__restrict__ Hi,
    const double4 *__restrict__ zv, const double *__restrict__ zt, const double2 *__restrict__ x,
    const double *__restrict__ A_ak, const double *__restrict__ rough, BpaParams p, double4 *__restrict__ sr4,
    double *__restrict__ eta_ak, double *__restrict__ beta_a) {
  const int nz = z.nz;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nV * nz) return;
  const int vi = (int)(i / nz), k = (int)(i % nz);
  const int b0 = zbase(k, nz);
  const double zeta = z.zeta[k];
  const double h = fmax(10.0, Hi[vi]);
  const double4 g = zv[vi];
  const double zx = (1.0 / h) * (g.x - zeta * g.z), zy = (1.0 / h) * (g.y - zeta * g.w);
  const double d0 = z.d[3 * k], d1 = z.d[3 * k + 1], d2 = z.d[3 * k + 2];
  const int s0 = z.ksc[2 * k], s1 = z.ksc[2 * k + 1];
  const double m0 = z.ksm[2 * k], m1 = z.ksm[2 * k + 1];
  double ux = 0, uy = 0, vx = 0, vy = 0, uz = 0, vz = 0, ua = 0, va = 0;
  for (int e = ptr[vi]; e < ptr[vi + 1]; ++e) {
    const int ti = ind[e];
    const double cm = wm[e], cx = wx[e], cy = wy[e];
    const double2 *col = x + (size_t)ti * nz;
    const double2 xk = col[k];
    ux += cx * xk.x; uy += cy * xk.x; vx += cx * xk.y; vy += cy * xk.y;
    const double2 q0 = col[b0], q1 = col[b0 + 1], q2 = col[b0 + 2];
    const double dzu = d0 * q0.x + d1 * q1.x + d2 * q2.x, dzv = d0 * q0.y + d1 * q1.y + d2 * q2.y;
    ux += zx * cm * dzu; uy += zy * cm * dzu; vx += zx * cm * dzv; vy += zy * cm * dzv;
    // vertical shear on the staggered layers s0, s1 of this triangle, mapped to layer k
    const double zzs = -1.0 / fmax(10.0, zt[(size_t)ti * kZt]);
    const double2 a0 = col[s0], a1 = col[s0 + 1], c0 = col[s1], c1 = col[s1 + 1];
    const double us0 = zzs * (z.kkd[2 * s0] * a0.x + z.kkd[2 * s0 + 1] * a1.x);
    const double vs0 = zzs * (z.kkd[2 * s0] * a0.y + z.kkd[2 * s0 + 1] * a1.y);
    const double us1 = zzs * (z.kkd[2 * s1] * c0.x + z.kkd[2 * s1 + 1] * c1.x);
    const double vs1 = zzs * (z.kkd[2 * s1] * c0.y + z.kkd[2 * s1 + 1] * c1.y);
    uz += cm * (m0 * us0 + m1 * us1);
    vz += cm * (m0 * vs0 + m1 * vs1);
    if (k == nz - 1) {
      const double2 xb = col[nz - 1];
      ua += cm * xb.x; va += cm * xb.y;
    }
  }
  sr4[i] = make_double4(ux, uy, vx, vy);
  const double e = glen3(p.eps0_applied, ux, uy, uz, vx, vy, vz, A_ak[i], p.n);
  eta_ak[i] = fmin(fmax(e, p.visc_eff_min), p.eta_max);
  if (k == nz - 1) beta_a[vi] = basal_friction(p.sl, vi, rough[vi], ua, va);
}

// One thread per (triangle, staggered layer): strain rates mapped ak->bks,
// vertical shear from the column, eta_bks with the configured regularisation
// (BPA_main.f90:1876-1880).
__global__ void __launch_bounds__(kBlock) k_bpa_bks(int nT, ZOps z, const int *__restrict__ ptr,
                                                    const int *__restrict__ ind, const double *__restrict__ wm,
                                                    const double *__restrict__ zt, const double2 *__restrict__ x,
                                                    const double4 *__restrict__ sr4, const double *__restrict__ A_bks,
                                                    BpaParams p, double *__restrict__ eta_bks) {
  const int nz = z.nz, ns = nz - 1;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nT * ns) return;
  const int ti = (int)(i / ns), ks = (int)(i % ns);
  const double w0 = z.kkm[2 * ks], w1 = z.kkm[2 * ks + 1];
  double ux = 0, uy = 0, vx = 0, vy = 0;
  for (int e = ptr[ti]; e < ptr[ti + 1]; ++e) {
    const int vi = ind[e];
    const double m = wm[e];
    const double4 a = sr4[(size_t)vi * nz + ks], b = sr4[(size_t)vi * nz + ks + 1];
    ux += m * (w0 * a.x + w1 * b.x); uy += m * (w0 * a.y + w1 * b.y);
    vx += m * (w0 * a.z + w1 * b.z); vy += m * (w0 * a.w + w1 * b.w);
  }
  const double zz = -1.0 / fmax(10.0, zt[(size_t)ti * kZt]);
  const double2 c0 = x[(size_t)ti * nz + ks], c1 = x[(size_t)ti * nz + ks + 1];
  const double uz = zz * (z.kkd[2 * ks] * c0.x + z.kkd[2 * ks + 1] * c1.x);
  const double vz = zz * (z.kkd[2 * ks] * c0.y + z.kkd[2 * ks + 1] * c1.y);
  const double e = glen3(p.eps0_config, ux, uy, uz, vx, vy, vz, A_bks[i], p.n);
  eta_bks[i] = fmin(fmax(e, p.visc_eff_min), p.eta_max);
}

// One thread per (triangle, layer): eta_bk (shear-factor blend of the ak and
// bks viscosities), its gradients (ak->bk d/dx, d/dy; bks->bk d/dz), and at
// the surface layer the triangle's basal friction coefficient
// (map_a_b, times f_gr^p with sub-grid friction).
__global__ void __launch_bounds__(kBlock) k_bpa_coef(int nT, ZOps z, const int *__restrict__ ptr,
                                                     const int *__restrict__ ind, const double *__restrict__ wm,
                                                     const double *__restrict__ wx, const double *__restrict__ wy,
                                                     const double *__restrict__ zt, const double2 *__restrict__ x,
                                                     const double *__restrict__ eta_ak,
                                                     const double *__restrict__ eta_bks,
                                                     const double *__restrict__ beta_a, const double *__restrict__ fgr,
                                                     BpaParams p, double4 *__restrict__ coef,
                                                     double *__restrict__ beta_b) {
  const int nz = z.nz, ns = nz - 1;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nT * nz) return;
  const int ti = (int)(i / nz), k = (int)(i % nz);
  const int b0 = zbase(k, nz);
  const double *ztt = zt + (size_t)ti * kZt;
  const ZGrad g = zgrad_tri(ztt, z.zeta[k]);
  const double d0 = z.d[3 * k], d1 = z.d[3 * k + 1], d2 = z.d[3 * k + 2];
  double ex = 0, ey = 0, ea = 0, bb = 0;
  for (int e = ptr[ti]; e < ptr[ti + 1]; ++e) {
    const int vi = ind[e];
    const double m = wm[e];
    const double *col = eta_ak + (size_t)vi * nz;
    const double ek = col[k];
    const double dz = d0 * col[b0] + d1 * col[b0 + 1] + d2 * col[b0 + 2];
    ex += wx[e] * ek + g.zx * m * dz;
    ey += wy[e] * ek + g.zy * m * dz;
    ea += m * ek;
    if (k == 0) bb += m * beta_a[vi];
  }
  const double *es = eta_bks + (size_t)ti * ns;
  const int s0 = z.ksc[2 * k], s1 = z.ksc[2 * k + 1];
  const double ez = g.zz * (z.ksd[2 * k] * es[s0] + z.ksd[2 * k + 1] * es[s1]);
  const double eb = z.ksm[2 * k] * es[s0] + z.ksm[2 * k + 1] * es[s1];
  const double2 xs = x[(size_t)ti * nz], xb = x[(size_t)ti * nz + nz - 1];
  const double us = sqrt(0.1 + xs.x * xs.x + xs.y * xs.y), ub = sqrt(0.1 + xb.x * xb.x + xb.y * xb.y);
  const double R = ub / us;
  coef[i] = make_double4(R * ea + (1.0 - R) * eb, ex, ey, ez);
  if (k == 0) beta_b[ti] = p.do_subgrid ? bb * pow(fgr[ti], p.subgrid_exp) : bb;
}

// ---------------------------------------------------------------------------
// stiffness matrix (solve_BPA_linearised, BPA_main.f90:330-1723)
// ---------------------------------------------------------------------------

struct BpaSys {
  int nT, nz;
  const int *ptr, *crow;        // block rows ti*nz+k, entry -> row
  const int *bb_ptr, *bb_ind, *TriC, *TriBI;
  const double *bb_ddx, *bb_ddy, *bb_d2dx2, *bb_d2dxdy, *bb_d2dy2;
  const double *zt;
  const double4 *coef;          // eta_bk, deta/dx, deta/dy, deta/dz
  const double *eta_bks, *beta_b;
  const int *bcm;               // prescribed mask per bk point (may be null)
  const double *bcu, *bcv;
};

// Row kind of component uv of bk point (ti, k): 0 free, 1 surface, 2 base,
// 3 identity with a value (prescribed, 'zero', copy, base without sliding),
// 4 'infinite' border
__device__ __forceinline__ int bpa_row_kind(const BpaSys &s, const BpaParams &p, int ti, int k, int uv) {
  if (s.bcm && s.bcm[(size_t)ti * s.nz + k] == 1) return 3;
  const int tb = s.TriBI[ti];
  if (tb > 0) {
    const int side = (tb - 1) / 2;
    const int choice = uv == 0 ? p.bc_u[side] : p.bc_v[side];
    return choice == UFE_BC_INFINITE ? 4 : 3;
  }
  if (k == 0) return 1;
  if (k == s.nz - 1) return p.sl.law == UFE_SLIDING_NO_SLIDING ? 3 : 2;
  return 0;
}

// One thread per block entry.  The 3-D operator coefficients of the entry
// (calc_3D_matrix_operators_mesh_bk_bk, :832-1072) are formed from the b->b
// entry and the zeta stencil, then combined per row kind: free rows
// (:480-644), surface rows (:646-889), base rows (:891-1163), 'infinite'
// borders (:1165-1723) and identity rows.
__global__ void __launch_bounds__(kBlock) k_bpa_assemble(BpaSys s, ZOps z, BpaParams p, long nnzb,
                                                         double4 *__restrict__ A) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnzb) return;
  const int nz = s.nz;
  const int r = s.crow[e];
  const int ti = r / nz, k = r % nz;
  const int pos = (int)(e - s.ptr[r]);
  const int ii = pos / 3, jz = pos % 3;
  const int be = s.bb_ptr[ti] + ii;
  const int tj = s.bb_ind[be];
  const int kk = zbase(k, nz) + jz;
  const bool diag = (tj == ti && kk == k);
  double blk[4];  // (u row: u col, v col), (v row: u col, v col)
  int kind[2] = {bpa_row_kind(s, p, ti, k, 0), bpa_row_kind(s, p, ti, k, 1)};
  bool need_ops = false;
  for (int uv = 0; uv < 2; ++uv) need_ops |= kind[uv] <= 2;
  double ox = 0, oy = 0, oz = 0, oxx = 0, oxy = 0, oyy = 0, ozz = 0;
  ZGrad g{};
  if (need_ops) {
    const double hx = s.bb_ddx[be], hy = s.bb_ddy[be];
    const double hxx = s.bb_d2dx2[be], hxy = s.bb_d2dxdy[be], hyy = s.bb_d2dy2[be];
    const double cz = z.d[3 * k + jz], czz = z.dd[3 * k + jz];
    g = zgrad_tri(s.zt + (size_t)ti * kZt, z.zeta[k]);
    if (kk == k) { ox += hx; oy += hy; oxx += hxx; oxy += hxy; oyy += hyy; }
    if (tj == ti) {
      ox += g.zx * cz;
      oy += g.zy * cz;
      oz += g.zz * cz;
      oxx += g.zxx * cz + g.zx * g.zx * czz;
      oxy += g.zxy * cz + g.zx * g.zy * czz;
      oyy += g.zyy * cz + g.zy * g.zy * czz;
      ozz += g.zz * g.zz * czz;
    }
    oxx += 2.0 * g.zx * hx * cz;
    oxy += g.zx * hy * cz + g.zy * hx * cz;
    oyy += 2.0 * g.zy * hy * cz;
  }
  const double4 c = s.coef[r];
  for (int uv = 0; uv < 2; ++uv) {
    double own = 0.0, oth = 0.0;
    const int kd = kind[uv];
    if (kd == 3) {
      own = diag ? 1.0 : 0.0;
    } else if (kd == 4) {
      if (kk == k) {
        if (tj == ti) {
          int n = 0;
          for (int q = 0; q < 3; ++q) n += s.TriC[3 * ti + q] >= 0;
          own = -(double)n;
        } else {
          for (int q = 0; q < 3; ++q)
            if (s.TriC[3 * ti + q] == tj) own = 1.0;
        }
      }
    } else {
      const double os = uv == 0 ? ox : oy, oo = uv == 0 ? oy : ox;
      const double oss = uv == 0 ? oxx : oyy, ooo = uv == 0 ? oyy : oxx;
      const double ds = uv == 0 ? c.y : c.z, dox = uv == 0 ? c.z : c.y, dez = c.w;
      if (kd == 0) {
        const double eta = c.x;
        own = 4.0 * eta * oss + 4.0 * ds * os + eta * ooo + dox * oo + eta * ozz + dez * oz;
        oth = 3.0 * eta * oxy + 2.0 * ds * oo + dox * os;
      } else {
        const double *ztt = s.zt + (size_t)ti * kZt;
        double eta, c_own_x, c_own_y, c_oth_x, c_oth_y, c_k, c_kn;
        int kn;
        if (kd == 1) {
          eta = s.eta_bks[(size_t)ti * (nz - 1)];
          const double hs = uv == 0 ? ztt[1] : ztt[2], ho = uv == 0 ? ztt[2] : ztt[1];
          const double dz = z.zeta[1] - z.zeta[0];
          const double q = 2.0 * eta / dz * g.zz;
          c_own_x = 4.0 * ds - q * 4.0 * hs + 4.0 * dez * hs;
          c_own_y = dox - q * ho + dez * ho;
          c_oth_y = 2.0 * ds - q * 2.0 * hs + 2.0 * dez * hs;
          c_oth_x = dox - q * ho + dez * ho;
          c_k = -2.0 * eta / (dz * dz) * g.zz * g.zz;
          c_kn = 2.0 * eta / (dz * dz) * g.zz * g.zz;
          kn = 1;
        } else {
          eta = s.eta_bks[(size_t)ti * (nz - 1) + nz - 2];
          const double bs = uv == 0 ? ztt[11] : ztt[12], bo = uv == 0 ? ztt[12] : ztt[11];
          const double dz = z.zeta[nz - 1] - z.zeta[nz - 2];
          const double Q = 2.0 * eta / (dz * dz) * g.zz * g.zz;
          const double R = 2.0 * eta / dz * g.zz + dez;
          c_own_x = 4.0 * ds + 4.0 * R * bs;
          c_own_y = dox + R * bo;
          c_oth_y = 2.0 * ds + 2.0 * R * bs;
          c_oth_x = dox + R * bo;
          c_k = (R * s.beta_b[ti] / eta) - Q;
          c_kn = Q;
          kn = nz - 2;
        }
        own = c_own_x * os + c_own_y * oo + 4.0 * eta * oss + eta * ooo;
        if (tj == ti && kk == k) own += c_k;
        if (tj == ti && kk == kn) own += c_kn;
        oth = c_oth_y * oo + c_oth_x * os + 3.0 * eta * oxy;
      }
    }
    if (uv == 0) { blk[0] = own; blk[1] = oth; }
    else { blk[3] = own; blk[2] = oth; }
  }
  A[e] = make_double4(blk[0], blk[1], blk[2], blk[3]);
}

// Load vector per bk point: -tau_d (tau_d = -rho g grad Hs) on the momentum
// rows, prescribed values, 0 on 'zero'/'infinite' borders and on the base
// without sliding; the copy rows are filled by k_bpa_copy.
__global__ void __launch_bounds__(kBlock) k_bpa_rhs(BpaSys s, BpaParams p, double2 *__restrict__ rhs) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)s.nT * s.nz) return;
  const int ti = (int)(i / s.nz), k = (int)(i % s.nz);
  double b[2];
  for (int uv = 0; uv < 2; ++uv) {
    const int kd = bpa_row_kind(s, p, ti, k, uv);
    if (s.bcm && s.bcm[i] == 1) b[uv] = uv == 0 ? s.bcu[i] : s.bcv[i];
    else if (kd == 3 || kd == 4) b[uv] = 0.0;
    else b[uv] = kIceDensity * kGrav * s.zt[(size_t)ti * kZt + 1 + uv];
  }
  rhs[i] = make_double2(b[0], b[1]);
}

// Lagged copy rows of 'periodic_ISMIP-HOM' per layer (BPA_main.f90:1212-1228):
// relax * (weighted previous velocity around the copy point) + (1 - relax) *
// the row's own previous velocity.  Thread per (copy row, layer).
__global__ void __launch_bounds__(kBlock) k_bpa_copy(int n, int nz, const int *__restrict__ row,
                                                     const int *__restrict__ cptr, const int *__restrict__ ctri,
                                                     const double *__restrict__ cw, const int *__restrict__ bcm,
                                                     double relax, const double *__restrict__ xprev,
                                                     double *__restrict__ rhs) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * nz) return;
  const int c = (int)(i / nz), k = (int)(i % nz);
  const int r = row[c], ti = r / 2, uv = r % 2;
  const size_t own = ((size_t)ti * nz + k) * 2 + uv;
  if (bcm && bcm[(size_t)ti * nz + k] == 1) return;
  double v = 0.0;
  for (int q = cptr[c]; q < cptr[c + 1]; ++q) v += cw[q] * xprev[((size_t)ctri[q] * nz + k) * 2 + uv];
  rhs[own] = (relax * v) + ((1.0 - relax) * xprev[own]);
}

// ---------------------------------------------------------------------------
// column preconditioner: block-tridiagonal LU of each triangle's column
// ---------------------------------------------------------------------------

__device__ __forceinline__ double4 m2mul(double4 a, double4 b) {
  return make_double4(a.x * b.x + a.y * b.z, a.x * b.y + a.y * b.w, a.z * b.x + a.w * b.z, a.z * b.y + a.w * b.w);
}
__device__ __forceinline__ double4 m2sub(double4 a, double4 b) {
  return make_double4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w);
}
__device__ __forceinline__ double4 m2inv(double4 a) {
  const double det = a.x * a.w - a.y * a.z;
  if (det == 0.0 || !isfinite(det)) return make_double4(1.0, 0.0, 0.0, 1.0);
  const double r = 1.0 / det;
  return make_double4(a.w * r, -a.y * r, -a.z * r, a.x * r);
}
__device__ __forceinline__ double2 m2vec(double4 a, double2 v) {
  return make_double2(a.x * v.x + a.y * v.y, a.z * v.x + a.w * v.y);
}

// Entry of block (ti, k) -> (ti, kk) for |kk - k| <= 1 within the stencil, or -1.
__device__ __forceinline__ long col_entry(const int *__restrict__ ptr, const int *__restrict__ bb_ptr,
                                          const int *__restrict__ bb_diag, int nz, int ti, int k, int kk) {
  const int j = kk - zbase(k, nz);
  if (kk < 0 || kk >= nz || j < 0 || j > 2) return -1;
  return (long)ptr[ti * nz + k] + 3L * (bb_diag[ti] - bb_ptr[ti]) + j;
}

// Factor per triangle: D'_0 = D_0, W_k = L_k D'_{k-1}^-1, D'_k = D_k - W_k U_{k-1};
// stores D'^-1, W and U per layer in [k][ti] order.
__global__ void __launch_bounds__(kBlock) k_col_factor(int nT, int nz, const int *__restrict__ ptr,
                                                       const int *__restrict__ bb_ptr, const int *__restrict__ bb_diag,
                                                       const double4 *__restrict__ A, double4 *__restrict__ Dinv,
                                                       double4 *__restrict__ W, double4 *__restrict__ U) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  double4 dprev_inv = make_double4(1, 0, 0, 1), uprev = make_double4(0, 0, 0, 0);
  for (int k = 0; k < nz; ++k) {
    const long ed = col_entry(ptr, bb_ptr, bb_diag, nz, ti, k, k);
    const long el = col_entry(ptr, bb_ptr, bb_diag, nz, ti, k, k - 1);
    const long eu = col_entry(ptr, bb_ptr, bb_diag, nz, ti, k, k + 1);
    double4 D = A[ed];
    const double4 L = el >= 0 ? A[el] : make_double4(0, 0, 0, 0);
    const double4 Uk = eu >= 0 ? A[eu] : make_double4(0, 0, 0, 0);
    double4 w = make_double4(0, 0, 0, 0);
    if (k > 0) {
      w = m2mul(L, dprev_inv);
      D = m2sub(D, m2mul(w, uprev));
    }
    dprev_inv = m2inv(D);
    uprev = Uk;
    const size_t o = (size_t)k * nT + ti;
    Dinv[o] = dprev_inv;
    W[o] = w;
    U[o] = Uk;
  }
}

// z = M^-1 r with the column factors.  One wavefront per 64 consecutive
// triangles: their columns of r are one contiguous stretch of memory, loaded
// into LDS with coalesced reads (a triangle's column is strided by nz * 16 B
// across lanes in memory), solved per lane in LDS (forward sweep, then
// backward in place; columns padded to an odd number of doubles so the lanes
// hit different banks), and stored back coalesced.
constexpr int kColWave = 64;
constexpr int kColPad = 2 * kBpaMaxNz + 1;

__global__ void __launch_bounds__(kColWave) k_col_apply(int nT, int nz, const double4 *__restrict__ Dinv,
                                                        const double4 *__restrict__ W,
                                                        const double4 *__restrict__ U,
                                                        const double *__restrict__ r, double *__restrict__ z) {
  __shared__ double col[kColWave * kColPad];
  const int lane = threadIdx.x;
  const int t0 = blockIdx.x * kColWave;
  const int nt = min(kColWave, nT - t0);
  const int nd = 2 * nz;                       // doubles per column
  const long base = (long)t0 * nd;
  for (int q = lane; q < nt * nd; q += kColWave) col[(q / nd) * kColPad + q % nd] = r[base + q];
  __syncthreads();
  if (lane < nt) {
    const int ti = t0 + lane;
    double *c = col + lane * kColPad;
    double px = c[0], py = c[1];  // y_0 = r_0
    for (int k = 1; k < nz; ++k) {  // y_k = r_k - W_k y_{k-1}
      const double4 w = W[(size_t)k * nT + ti];
      const double nx = c[2 * k] - (w.x * px + w.y * py);
      const double ny = c[2 * k + 1] - (w.z * px + w.w * py);
      c[2 * k] = nx;
      c[2 * k + 1] = ny;
      px = nx;
      py = ny;
    }
    double4 d = Dinv[(size_t)(nz - 1) * nT + ti];
    double zx = d.x * px + d.y * py, zy = d.z * px + d.w * py;
    c[2 * (nz - 1)] = zx;
    c[2 * (nz - 1) + 1] = zy;
    for (int k = nz - 2; k >= 0; --k) {  // z_k = D'_k^-1 (y_k - U_k z_{k+1})
      const double4 u = U[(size_t)k * nT + ti];
      const double ax = c[2 * k] - (u.x * zx + u.y * zy), ay = c[2 * k + 1] - (u.z * zx + u.w * zy);
      d = Dinv[(size_t)k * nT + ti];
      zx = d.x * ax + d.y * ay;
      zy = d.z * ax + d.w * ay;
      c[2 * k] = zx;
      c[2 * k + 1] = zy;
    }
  }
  __syncthreads();
  for (int q = lane; q < nt * nd; q += kColWave) z[base + q] = col[(q / nd) * kColPad + q % nd];
}

// ---------------------------------------------------------------------------
// velocity limits, relaxation, L2 change (BPA_main.f90:1982-2084)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_bpa_update(long n, double vel_max, double relax,
                                                       double2 *__restrict__ x, const double2 *__restrict__ xprev,
                                                       RedSlot slot) {
  __shared__ double lds[2][kBlock / 64];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < n) {
    double2 a = x[i];
    const double ua = sqrt(a.x * a.x + a.y * a.y);
    if (ua > vel_max) { a.x = a.x * vel_max / ua; a.y = a.y * vel_max / ua; }
    const double2 b = xprev[i];
    a.x = relax * a.x + (1.0 - relax) * b.x;
    a.y = relax * a.y + (1.0 - relax) * b.y;
    x[i] = a;
    v[0] = (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y);
    v[1] = (a.x + b.x) * (a.x + b.x) + (a.y + b.y) * (a.y + b.y);
  }
  block_sum<2>(v, lds);
  block_partials<2>(slot, v);
}

__global__ void __launch_bounds__(kBlock) k_bks_map_A(int nT, ZOps z, const int *__restrict__ ptr,
                                                      const int *__restrict__ ind, const double *__restrict__ wm,
                                                      const double *__restrict__ A_ak, double *__restrict__ A_bks) {
  const int nz = z.nz, ns = nz - 1;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nT * ns) return;
  const int ti = (int)(i / ns), ks = (int)(i % ns);
  double a = 0.0;
  for (int e = ptr[ti]; e < ptr[ti + 1]; ++e) {
    const double *col = A_ak + (size_t)ind[e] * nz;
    a += wm[e] * (z.kkm[2 * ks] * col[ks] + z.kkm[2 * ks + 1] * col[ks + 1]);
  }
  A_bks[i] = a;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct BpaState {
  int nV = 0, nT = 0, nz = 0;
  std::vector<int> v_o2n, t_o2n;
  // zeta set-up
  DBuf aa_ptr, aa_ind, aa_ddx, aa_ddy;
  DBuf zeta, zs, zd, zdd, kkm, kkd, ksc, ksm, ksd;
  std::vector<double> h_zeta;
  // block pattern (built on the first BPA call)
  int nb = 0;
  long nnzb = 0;
  DBuf ptr, col, crow, bb_diag_copy;
  int n_copy = 0;
  DBuf copy_row, copy_ptr, copy_tri, copy_w;
  // inputs and geometry
  DBuf Hi, Hs, A_ak, A_bks, rough, fgr, bcm, bcu, bcv, zv, zt, a2, tribi;
  bool has_bc = false;
  // iteration
  DBuf x, xprev, rhs, sr4, eta_ak, eta_bks, coef, beta_a, beta_b, A, cDinv, cW, cU, part, out, r;
  std::unique_ptr<Krylov> kry;
  // stage timings of the last solve (HIP events on the context's stream)
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double t_ms[5] = {0, 0, 0, 0, 0};
  ~BpaState() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }

  ZOps zops() const {
    ZOps z;
    z.nz = nz;
    z.zeta = zeta.as<double>(); z.zs = zs.as<double>();
    z.d = zd.as<double>(); z.dd = zdd.as<double>();
    z.kkm = kkm.as<double>(); z.kkd = kkd.as<double>();
    z.ksc = ksc.as<int>(); z.ksm = ksm.as<double>(); z.ksd = ksd.as<double>();
    return z;
  }
};

void up(DBuf &d, const void *h, size_t bytes, hipStream_t s) {
  if (d.bytes < bytes) d.alloc(std::max<size_t>(bytes, 8));
  if (bytes) UFE_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
}
void ensure(DBuf &d, size_t bytes) {
  if (d.bytes < bytes) d.alloc(std::max<size_t>(bytes, 8));
}

// calc_shape_functions_1D_reg_2nd_order (shape_functions.f90:17-79), two neighbours
void shape1d(double z, const double zc[2], double *fi, double *fii, double fc[2], double fcc[2]) {
  double dx[2], w[2], A[4] = {0, 0, 0, 0};
  for (int i = 0; i < 2; ++i) {
    dx[i] = zc[i] - z;
    w[i] = 1.0 / std::pow(std::fabs(dx[i]), 1.5);
    const double w2 = w[i] * w[i];
    A[0] += w2 * dx[i] * dx[i];
    A[1] += w2 * dx[i] * 0.5 * dx[i] * dx[i];
    A[2] += w2 * 0.5 * dx[i] * dx[i] * dx[i];
    A[3] += w2 * 0.5 * dx[i] * dx[i] * 0.5 * dx[i] * dx[i];
  }
  const double det = A[0] * A[3] - A[1] * A[2];
  const double M[4] = {A[3] / det, -A[1] / det, -A[2] / det, A[0] / det};
  for (int i = 0; i < 2; ++i) {
    const double w2 = w[i] * w[i];
    fc[i] = w2 * (M[0] * dx[i] + M[1] * 0.5 * dx[i] * dx[i]);
    fcc[i] = w2 * (M[2] * dx[i] + M[3] * 0.5 * dx[i] * dx[i]);
  }
  *fi = -(fc[0] + fc[1]);
  *fii = -(fcc[0] + fcc[1]);
}

BpaState *state(ufe_ctx *c) { return static_cast<BpaState *>(ctx_bpa(c).get()); }

// Inputs into device order and the per-solve geometry (zeta-gradient fields,
// A on the staggered layers).  A_flow, bc arrays: [nz][n] first index fastest.
void bpa_inputs(ufe_ctx *c, BpaState &b, const ufe_bpa_inputs *in, bool need_bc) {
  const DevMesh &m = *ctx_mesh(c);
  hipStream_t s = ctx_stream(c);
  const int nV = b.nV, nT = b.nT, nz = b.nz;
  const std::vector<int> &vperm = ctx_vperm(c), &tperm = ctx_tperm(c);
  std::vector<double> hv(nV), hv2(nV), hr(nV), ha((size_t)nV * nz), hf(nT);
  for (int i = 0; i < nV; ++i) {
    hv[i] = in->Hi[vperm[i]];
    hv2[i] = in->Hs[vperm[i]];
    hr[i] = in->beta_sq[vperm[i]];
    for (int k = 0; k < nz; ++k) ha[(size_t)i * nz + k] = in->A_flow[(size_t)k * nV + vperm[i]];
  }
  for (int i = 0; i < nT; ++i) hf[i] = in->fraction_gr_b ? in->fraction_gr_b[tperm[i]] : 1.0;
  up(b.Hi, hv.data(), sizeof(double) * nV, s);
  up(b.Hs, hv2.data(), sizeof(double) * nV, s);
  up(b.rough, hr.data(), sizeof(double) * nV, s);
  up(b.A_ak, ha.data(), sizeof(double) * ha.size(), s);
  up(b.fgr, hf.data(), sizeof(double) * nT, s);
  b.has_bc = need_bc && in->BC_prescr_mask_bk;
  if (b.has_bc) {
    std::vector<int> hm((size_t)nT * nz);
    std::vector<double> hu((size_t)nT * nz), hw((size_t)nT * nz);
    for (int i = 0; i < nT; ++i)
      for (int k = 0; k < nz; ++k) {
        const size_t o = (size_t)k * nT + tperm[i], d = (size_t)i * nz + k;
        hm[d] = in->BC_prescr_mask_bk[o];
        hu[d] = in->BC_prescr_u_bk[o];
        hw[d] = in->BC_prescr_v_bk[o];
      }
    up(b.bcm, hm.data(), sizeof(int) * hm.size(), s);
    up(b.bcu, hu.data(), sizeof(double) * hu.size(), s);
    up(b.bcv, hw.data(), sizeof(double) * hw.size(), s);
  }
  ensure(b.zv, sizeof(double4) * nV);
  ensure(b.zt, sizeof(double) * kZt * (size_t)nT);
  ensure(b.A_bks, sizeof(double) * (size_t)nT * (nz - 1));
  hipLaunchKernelGGL(k_zg_vertex, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, b.aa_ptr.as<int>(),
                     b.aa_ind.as<int>(), b.aa_ddx.as<double>(), b.aa_ddy.as<double>(), b.Hi.as<double>(),
                     b.Hs.as<double>(), b.zv.as<double4>());
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_zg_tri, dim3(blocks_for(nT)), dim3(kBlock), 0, s, nT, m.ab_ptr, m.ab_ind, m.ab_map,
                     m.ab_ddx, m.ab_ddy, b.Hi.as<double>(), b.Hs.as<double>(), b.zv.as<double4>(),
                     b.zt.as<double>());
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bks_map_A, dim3(blocks_for((long)nT * (nz - 1))), dim3(kBlock), 0, s, nT, b.zops(), m.ab_ptr,
                     m.ab_ind, m.ab_map, b.A_ak.as<double>(), b.A_bks.as<double>());
  UFE_LAUNCH_CHECK();
  UFE_HIP(hipStreamSynchronize(s));  // host vectors end here
}

// Block pattern of the BPA system, once per context.
void bpa_pattern(ufe_ctx *c, BpaState &b) {
  if (b.nb) return;
  const DevMesh &m = *ctx_mesh(c);
  hipStream_t s = ctx_stream(c);
  const int nT = b.nT, nz = b.nz;
  std::vector<int> bp(nT + 1), bi(m.bb_nnz);
  UFE_HIP(hipMemcpy(bp.data(), m.bb_ptr, sizeof(int) * (nT + 1), hipMemcpyDeviceToHost));
  UFE_HIP(hipMemcpy(bi.data(), m.bb_ind, sizeof(int) * m.bb_nnz, hipMemcpyDeviceToHost));
  const long nnzb = 3L * nz * m.bb_nnz;
  if (nnzb >= (1L << 31)) throw std::runtime_error("BPA system too large for 32-bit block indices");
  b.nb = nT * nz;
  b.nnzb = nnzb;
  std::vector<int> ptr((size_t)b.nb + 1), col(nnzb), crow(nnzb);
  long e = 0;
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) {
      const int r = ti * nz + k;
      ptr[r] = (int)e;
      const int base = std::min(std::max(k - 1, 0), nz - 3);
      for (int q = bp[ti]; q < bp[ti + 1]; ++q)
        for (int j = 0; j < 3; ++j, ++e) {
          col[e] = bi[q] * nz + base + j;
          crow[e] = r;
        }
    }
  ptr[b.nb] = (int)e;
  up(b.ptr, ptr.data(), sizeof(int) * ptr.size(), s);
  up(b.col, col.data(), sizeof(int) * col.size(), s);
  up(b.crow, crow.data(), sizeof(int) * crow.size(), s);
  // lagged copy rows of the periodic borders
  std::vector<int> crw, cp, ct;
  std::vector<double> cw;
  ctx_copy_lists(c, crw, cp, ct, cw);
  b.n_copy = (int)crw.size();
  if (b.n_copy) {
    up(b.copy_row, crw.data(), sizeof(int) * crw.size(), s);
    up(b.copy_ptr, cp.data(), sizeof(int) * cp.size(), s);
    up(b.copy_tri, ct.data(), sizeof(int) * ct.size(), s);
    up(b.copy_w, cw.data(), sizeof(double) * cw.size(), s);
  }
  UFE_HIP(hipStreamSynchronize(s));  // host vectors end here
  const size_t nbk = (size_t)b.nb;
  ensure(b.x, sizeof(double2) * nbk);
  ensure(b.xprev, sizeof(double2) * nbk);
  ensure(b.rhs, sizeof(double2) * nbk);
  ensure(b.r, sizeof(double2) * nbk);
  ensure(b.sr4, sizeof(double4) * (size_t)b.nV * nz);
  ensure(b.eta_ak, sizeof(double) * (size_t)b.nV * nz);
  ensure(b.eta_bks, sizeof(double) * (size_t)nT * (nz - 1));
  ensure(b.coef, sizeof(double4) * nbk);
  ensure(b.beta_a, sizeof(double) * b.nV);
  ensure(b.beta_b, sizeof(double) * nT);
  ensure(b.A, sizeof(double4) * (size_t)nnzb);
  ensure(b.cDinv, sizeof(double4) * nbk);
  ensure(b.cW, sizeof(double4) * nbk);
  ensure(b.cU, sizeof(double4) * nbk);
  ensure(b.part, sizeof(double) * kMaxPartials * 2);
  ensure(b.out, sizeof(double) * 2);
  b.kry = std::make_unique<Krylov>(b.nb, s);
  for (hipEvent_t &ev : b.ev) UFE_HIP(hipEventCreate(&ev));
}

BpaParams bpa_params(const ufe_ctx *c, double eps0_applied) {
  const ufe_params &p = ctx_params(c);
  BpaParams q{};
  q.n = p.glens_n;
  q.eps0_applied = eps0_applied;
  q.eps0_config = p.epsilon_sq_0;
  q.visc_eff_min = p.visc_eff_min;
  q.eta_max = 0.5 * std::pow(1e-18, -1.0 / p.glens_n) *
              std::pow(eps0_applied, (1.0 - p.glens_n) / (2.0 * p.glens_n));  // BPA_main.f90:1831
  q.sl = ctx_sliding_law(c);
  q.subgrid_exp = p.subgrid_friction_exponent;
  q.relax_copy = p.visc_it_relax;
  q.do_subgrid = p.do_gl_subgrid_friction;
  for (int i = 0; i < 4; ++i) { q.bc_u[i] = p.bc_u[i]; q.bc_v[i] = p.bc_v[i]; }
  return q;
}

BpaSys bpa_sys(ufe_ctx *c, const BpaState &b, const int *TriBI_dev) {
  const DevMesh &m = *ctx_mesh(c);
  BpaSys s{};
  s.nT = b.nT; s.nz = b.nz;
  s.ptr = b.ptr.as<int>(); s.crow = b.crow.as<int>();
  s.bb_ptr = m.bb_ptr; s.bb_ind = m.bb_ind; s.TriC = m.TriC; s.TriBI = TriBI_dev;
  s.bb_ddx = m.bb_ddx; s.bb_ddy = m.bb_ddy; s.bb_d2dx2 = m.bb_d2dx2; s.bb_d2dxdy = m.bb_d2dxdy; s.bb_d2dy2 = m.bb_d2dy2;
  s.zt = b.zt.as<double>();
  s.coef = b.coef.as<double4>();
  s.eta_bks = b.eta_bks.as<double>();
  s.beta_b = b.beta_b.as<double>();
  s.bcm = b.has_bc ? b.bcm.as<int>() : nullptr;
  s.bcu = b.has_bc ? b.bcu.as<double>() : nullptr;
  s.bcv = b.has_bc ? b.bcv.as<double>() : nullptr;
  return s;
}

BsrT<double> bpa_bsr(const BpaState &b) {
  BsrT<double> A;
  A.nb = b.nb; A.nnzb = (int)b.nnzb; A.ptr = b.ptr.as<int>(); A.col = b.col.as<int>();
  A.val = b.A.as<double4>(); A.dinv = nullptr;
  return A;
}

// The state at the current iterate x and the system assembled there
// (calc_strain_rates, calc_effective_viscosity,
// calc_applied_basal_friction_coefficient, solve_BPA_linearised's assembly);
// the copy rows read xprev.
void bpa_state_and_system(ufe_ctx *c, BpaState &b, const BpaParams &p, const int *TriBI_dev) {
  const DevMesh &m = *ctx_mesh(c);
  hipStream_t s = ctx_stream(c);
  const int nV = b.nV, nT = b.nT, nz = b.nz;
  const ZOps z = b.zops();
  hipLaunchKernelGGL(k_bpa_vertex, dim3(blocks_for((long)nV * nz)), dim3(kBlock), 0, s, nV, nT, z, m.ba_ptr, m.ba_ind,
                     m.ba_map, m.ba_ddx, m.ba_ddy, b.Hi.as<double>(), b.zv.as<double4>(), b.zt.as<double>(),
                     b.x.as<double2>(), b.A_ak.as<double>(), b.rough.as<double>(), p, b.sr4.as<double4>(),
                     b.eta_ak.as<double>(), b.beta_a.as<double>());
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bpa_bks, dim3(blocks_for((long)nT * (nz - 1))), dim3(kBlock), 0, s, nT, z, m.ab_ptr, m.ab_ind,
                     m.ab_map, b.zt.as<double>(), b.x.as<double2>(), b.sr4.as<double4>(), b.A_bks.as<double>(), p,
                     b.eta_bks.as<double>());
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bpa_coef, dim3(blocks_for((long)nT * nz)), dim3(kBlock), 0, s, nT, z, m.ab_ptr, m.ab_ind,
                     m.ab_map, m.ab_ddx, m.ab_ddy, b.zt.as<double>(), b.x.as<double2>(), b.eta_ak.as<double>(),
                     b.eta_bks.as<double>(), b.beta_a.as<double>(), b.fgr.as<double>(), p, b.coef.as<double4>(),
                     b.beta_b.as<double>());
  UFE_LAUNCH_CHECK();
  if (b.ev[1]) UFE_HIP(hipEventRecord(b.ev[1], s));
  const BpaSys sys = bpa_sys(c, b, TriBI_dev);
  hipLaunchKernelGGL(k_bpa_assemble, dim3(blocks_for(b.nnzb)), dim3(kBlock), 0, s, sys, z, p, b.nnzb,
                     b.A.as<double4>());
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bpa_rhs, dim3(blocks_for((long)b.nb)), dim3(kBlock), 0, s, sys, p, b.rhs.as<double2>());
  UFE_LAUNCH_CHECK();
  if (b.n_copy) {
    hipLaunchKernelGGL(k_bpa_copy, dim3(blocks_for((long)b.n_copy * nz)), dim3(kBlock), 0, s, b.n_copy, nz,
                       b.copy_row.as<int>(), b.copy_ptr.as<int>(), b.copy_tri.as<int>(), b.copy_w.as<double>(),
                       sys.bcm, p.relax_copy, b.xprev.as<double>(), b.rhs.as<double>());
    UFE_LAUNCH_CHECK();
  }
}

void velocity_up(ufe_ctx *c, BpaState &b, const double *u, const double *v) {
  const std::vector<int> &tperm = ctx_tperm(c);
  const int nT = b.nT, nz = b.nz;
  std::vector<double> h(2 * (size_t)nT * nz);
  for (int i = 0; i < nT; ++i)
    for (int k = 0; k < nz; ++k) {
      const size_t o = (size_t)k * nT + tperm[i], d = ((size_t)i * nz + k) * 2;
      h[d] = u ? u[o] : 0.0;
      h[d + 1] = v ? v[o] : 0.0;
    }
  up(b.x, h.data(), sizeof(double) * h.size(), ctx_stream(c));
  UFE_HIP(hipStreamSynchronize(ctx_stream(c)));
}

void velocity_down(ufe_ctx *c, BpaState &b, double *u, double *v) {
  const std::vector<int> &tperm = ctx_tperm(c);
  const int nT = b.nT, nz = b.nz;
  std::vector<double> h(2 * (size_t)nT * nz);
  UFE_HIP(hipMemcpyAsync(h.data(), b.x.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx_stream(c)));
  UFE_HIP(hipStreamSynchronize(ctx_stream(c)));
  for (int i = 0; i < nT; ++i)
    for (int k = 0; k < nz; ++k) {
      const size_t o = (size_t)k * nT + tperm[i], d = ((size_t)i * nz + k) * 2;
      u[o] = h[d];
      v[o] = h[d + 1];
    }
}

}  // namespace
}  // namespace ufe

using namespace ufe;

namespace {

// device numbering within each layer -> caller numbering, for the outputs
void layers_down(const std::vector<int> &perm, const std::vector<double> &dev, double *out, int layers) {
  const size_t n = perm.size();
  for (int k = 0; k < layers; ++k)
    for (size_t i = 0; i < n; ++i) out[k * n + perm[i]] = dev[k * n + i];
}

int check_bpa_params(ufe_ctx *c, const ufe_params &p) {
  for (int s = 0; s < 4; ++s)
    if (p.bc_u[s] == UFE_BC_INFINITE_SSA_ICESTREAM || p.bc_v[s] == UFE_BC_INFINITE_SSA_ICESTREAM)
      return ctx_fail(c, UFE_ERR_UNSUPPORTED,
                      "the BPA borders take 'infinite', 'zero' or 'periodic_ISMIP-HOM' (BPA_main.f90:1165-1723)");
  if (p.gmres_restart < 1 || p.gmres_restart > 64) return ctx_fail(c, UFE_ERR_ARGUMENT, "gmres_restart must be 1..64");
  ctx_sliding_law(c);  // throws when the law needs the Coulomb cap and none is set
  return UFE_OK;
}

int check_inputs(ufe_ctx *c, const ufe_bpa_inputs *in) {
  if (!in || !in->Hi || !in->Hs || !in->A_flow || !in->beta_sq) return ctx_fail(c, UFE_ERR_ARGUMENT, "null input");
  const bool a = in->BC_prescr_mask_bk, b = in->BC_prescr_u_bk, d = in->BC_prescr_v_bk;
  if ((a || b || d) && !(a && b && d))
    return ctx_fail(c, UFE_ERR_ARGUMENT, "need to provide prescribed u,v fields and mask!");
  return UFE_OK;
}

void col_factor(ufe_ctx *c, BpaState &b) {
  const DevMesh &m = *ctx_mesh(c);
  hipLaunchKernelGGL(k_col_factor, dim3(blocks_for(b.nT)), dim3(kBlock), 0, ctx_stream(c), b.nT, b.nz, b.ptr.as<int>(),
                     m.bb_ptr, m.bb_diag, b.A.as<double4>(), b.cDinv.as<double4>(), b.cW.as<double4>(),
                     b.cU.as<double4>());
  UFE_LAUNCH_CHECK();
}

void set_column_pc(ufe_ctx *c, BpaState &b) {
  BpaState *bp = &b;
  hipStream_t s = ctx_stream(c);
  b.kry->set_custom_pc([bp, s](const double *r, double *z) {
    hipLaunchKernelGGL(k_col_apply, dim3(blocks_for(bp->nT, kColWave)), dim3(kColWave), 0, s, bp->nT, bp->nz,
                       bp->cDinv.as<double4>(), bp->cW.as<double4>(), bp->cU.as<double4>(), r, z);
    UFE_LAUNCH_CHECK();
  });
}

// ||b - A x|| / ||b|| of the system assembled last, at x
double bpa_rel_residual(BpaState &b) {
  double rr = 0.0, bb = 0.0;
  b.kry->residual64(bpa_bsr(b), b.x.as<double>(), b.rhs.as<double>(), b.r.as<double>(), &rr, &bb);
  return std::sqrt(rr) / std::max(std::sqrt(bb), 1e-300);
}

}  // namespace

// Set-up of the zeta gradients and the BPA: the a->a operators (device
// order), the staggered layers and the 1-D zeta operators
// (calc_vertical_operators_reg_1D / _stag_1D, mesh_zeta.f90:287-482).
extern "C" int ufe_zeta_setup(ufe_ctx *c, int aa_nnz, const int *aa_ptr, const int *aa_ind, const double *aa_ddx,
                              const double *aa_ddy, const double *zeta_stag) {
  if (!c || !aa_ptr || !aa_ind || !aa_ddx || !aa_ddy) return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    const DevMesh &m = *ctx_mesh(c);
    hipStream_t s = ctx_stream(c);
    const int nV = m.nV, nT = m.nT, nz = m.nz;
    if (nz < 3 || nz > kBpaMaxNz)
      return ctx_fail(c, UFE_ERR_UNSUPPORTED, "the zeta gradients and the BPA need 3 to 32 layers");
    if (aa_ptr[0] != 0 || aa_ptr[nV] != aa_nnz) return ctx_fail(c, UFE_ERR_ARGUMENT, "bad a->a operator pointer");
    for (int i = 0; i < nV; ++i)
      if (aa_ptr[i + 1] < aa_ptr[i]) return ctx_fail(c, UFE_ERR_ARGUMENT, "a->a operator pointer not monotone");
    for (int e = 0; e < aa_nnz; ++e)
      if (aa_ind[e] < 0 || aa_ind[e] >= nV) return ctx_fail(c, UFE_ERR_ARGUMENT, "a->a operator column out of range");
    auto b = std::make_shared<BpaState>();
    b->nV = nV; b->nT = nT; b->nz = nz;
    const std::vector<int> &vperm = ctx_vperm(c), &tperm = ctx_tperm(c);
    b->v_o2n.assign(nV, 0);
    b->t_o2n.assign(nT, 0);
    for (int i = 0; i < nV; ++i) b->v_o2n[vperm[i]] = i;
    for (int i = 0; i < nT; ++i) b->t_o2n[tperm[i]] = i;
    std::vector<int> p(nV + 1), ind(aa_nnz);
    std::vector<double> dx(aa_nnz), dy(aa_nnz);
    p[0] = 0;
    for (int i = 0; i < nV; ++i) {
      const int o = vperm[i];
      int e = p[i];
      for (int q = aa_ptr[o]; q < aa_ptr[o + 1]; ++q, ++e) {
        ind[e] = b->v_o2n[aa_ind[q]];
        dx[e] = aa_ddx[q];
        dy[e] = aa_ddy[q];
      }
      p[i + 1] = e;
    }
    up(b->aa_ptr, p.data(), sizeof(int) * p.size(), s);
    up(b->aa_ind, ind.data(), sizeof(int) * ind.size(), s);
    up(b->aa_ddx, dx.data(), sizeof(double) * dx.size(), s);
    up(b->aa_ddy, dy.data(), sizeof(double) * dy.size(), s);
    std::vector<double> zeta(nz), zs(nz - 1);
    UFE_HIP(hipMemcpy(zeta.data(), m.zeta, sizeof(double) * nz, hipMemcpyDeviceToHost));
    for (int k = 0; k < nz - 1; ++k) {
      zs[k] = zeta_stag ? zeta_stag[k] : 0.5 * (zeta[k] + zeta[k + 1]);
      if (!(zs[k] > zeta[k] && zs[k] < zeta[k + 1]))
        return ctx_fail(c, UFE_ERR_ARGUMENT, "zeta_stag must lie strictly between consecutive zeta levels");
    }
    std::vector<double> d(3 * nz), dd(3 * nz), kkm(2 * (nz - 1)), kkd(2 * (nz - 1)), ksm(2 * nz), ksd(2 * nz);
    std::vector<int> ksc(2 * nz);
    for (int k = 0; k < nz; ++k) {
      int cc[2];
      if (k == 0) { cc[0] = 1; cc[1] = 2; }
      else if (k == nz - 1) { cc[0] = nz - 3; cc[1] = nz - 2; }
      else { cc[0] = k - 1; cc[1] = k + 1; }
      const double zc[2] = {zeta[cc[0]], zeta[cc[1]]};
      double fi, fii, fc[2], fcc[2];
      shape1d(zeta[k], zc, &fi, &fii, fc, fcc);
      const int base = std::min(std::max(k - 1, 0), nz - 3);
      d[3 * k + (k - base)] = fi;
      dd[3 * k + (k - base)] = fii;
      for (int i = 0; i < 2; ++i) {
        d[3 * k + (cc[i] - base)] = fc[i];
        dd[3 * k + (cc[i] - base)] = fcc[i];
      }
    }
    for (int ks = 0; ks < nz - 1; ++ks) {
      const double dz = zeta[ks + 1] - zeta[ks];
      const double wlo = (zeta[ks + 1] - zs[ks]) / dz;
      kkm[2 * ks] = wlo; kkm[2 * ks + 1] = 1.0 - wlo;
      kkd[2 * ks] = -1.0 / dz; kkd[2 * ks + 1] = 1.0 / dz;
    }
    for (int k = 0; k < nz; ++k) {
      int lo, hi;
      if (k == 0) { lo = 0; hi = 1; }
      else if (k == nz - 1) { lo = nz - 3; hi = nz - 2; }
      else { lo = k - 1; hi = k; }
      const double dz = zs[hi] - zs[lo];
      const double wlo = (zs[hi] - zeta[k]) / dz;
      ksc[2 * k] = lo; ksc[2 * k + 1] = hi;
      ksm[2 * k] = wlo; ksm[2 * k + 1] = 1.0 - wlo;
      ksd[2 * k] = -1.0 / dz; ksd[2 * k + 1] = 1.0 / dz;
    }
    b->h_zeta = zeta;
    up(b->zeta, zeta.data(), sizeof(double) * nz, s);
    up(b->zs, zs.data(), sizeof(double) * zs.size(), s);
    up(b->zd, d.data(), sizeof(double) * d.size(), s);
    up(b->zdd, dd.data(), sizeof(double) * dd.size(), s);
    up(b->kkm, kkm.data(), sizeof(double) * kkm.size(), s);
    up(b->kkd, kkd.data(), sizeof(double) * kkd.size(), s);
    up(b->ksc, ksc.data(), sizeof(int) * ksc.size(), s);
    up(b->ksm, ksm.data(), sizeof(double) * ksm.size(), s);
    up(b->ksd, ksd.data(), sizeof(double) * ksd.size(), s);
    const std::vector<int> &tb = ctx_TriBI(c);
    up(b->tribi, tb.data(), sizeof(int) * tb.size(), s);
    UFE_HIP(hipStreamSynchronize(s));  // host vectors end here
    ctx_bpa(c) = b;
    return UFE_OK;
  });
}

extern "C" int ufe_zeta_gradients(ufe_ctx *c, const double *Hi, const double *Hs, const double *dHi_dt,
                                  const double *dHs_dt, double *const *ak, double *const *bk, double *const *bks) {
  if (!c || !Hi || !Hs) return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    BpaState *b = state(c);
    if (!b) return ctx_fail(c, UFE_ERR_ARGUMENT, "ufe_zeta_setup has not been called");
    const DevMesh &m = *ctx_mesh(c);
    hipStream_t s = ctx_stream(c);
    const int nV = b->nV, nT = b->nT, nz = b->nz;
    const std::vector<int> &vperm = ctx_vperm(c), &tperm = ctx_tperm(c);
    std::vector<double> h(nV), g(nV), ht(nV), gt(nV);
    for (int i = 0; i < nV; ++i) {
      h[i] = Hi[vperm[i]];
      g[i] = Hs[vperm[i]];
      ht[i] = dHi_dt ? dHi_dt[vperm[i]] : 0.0;
      gt[i] = dHs_dt ? dHs_dt[vperm[i]] : 0.0;
    }
    DBuf dht, dgt, a2, oak, obk, obks;
    up(b->Hi, h.data(), sizeof(double) * nV, s);
    up(b->Hs, g.data(), sizeof(double) * nV, s);
    up(dht, ht.data(), sizeof(double) * nV, s);
    up(dgt, gt.data(), sizeof(double) * nV, s);
    ensure(b->zv, sizeof(double4) * nV);
    ensure(b->zt, sizeof(double) * kZt * (size_t)nT);
    a2.alloc(sizeof(double) * 6 * (size_t)nV);
    const size_t nak = (size_t)nV * nz, nbk = (size_t)nT * nz, nbks = (size_t)nT * (nz - 1);
    oak.alloc(sizeof(double) * 7 * nak);
    obk.alloc(sizeof(double) * 6 * nbk);
    obks.alloc(sizeof(double) * 6 * std::max<size_t>(nbks, 1));
    hipLaunchKernelGGL(k_zg_vertex, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, b->aa_ptr.as<int>(),
                       b->aa_ind.as<int>(), b->aa_ddx.as<double>(), b->aa_ddy.as<double>(), b->Hi.as<double>(),
                       b->Hs.as<double>(), b->zv.as<double4>());
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_zg_tri, dim3(blocks_for(nT)), dim3(kBlock), 0, s, nT, m.ab_ptr, m.ab_ind, m.ab_map,
                       m.ab_ddx, m.ab_ddy, b->Hi.as<double>(), b->Hs.as<double>(), b->zv.as<double4>(),
                       b->zt.as<double>());
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_zg_vertex2, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, m.ba_ptr, m.ba_ind, m.ba_ddx,
                       m.ba_ddy, b->zt.as<double>(), a2.as<double>());
    UFE_LAUNCH_CHECK();
    const long n = (long)std::max(nak, nbk);
    hipLaunchKernelGGL(k_zg_out, dim3(blocks_for(n)), dim3(kBlock), 0, s, nV, nT, b->zops(), b->Hi.as<double>(),
                       dHi_dt ? dht.as<double>() : nullptr, dHs_dt ? dgt.as<double>() : nullptr, b->zv.as<double4>(),
                       a2.as<double>(), b->zt.as<double>(), oak.as<double>(), obk.as<double>(), obks.as<double>());
    UFE_LAUNCH_CHECK();
    std::vector<double> hak(7 * nak), hbk(6 * nbk), hbks(6 * nbks);
    UFE_HIP(hipMemcpyAsync(hak.data(), oak.p, sizeof(double) * hak.size(), hipMemcpyDeviceToHost, s));
    UFE_HIP(hipMemcpyAsync(hbk.data(), obk.p, sizeof(double) * hbk.size(), hipMemcpyDeviceToHost, s));
    if (nbks) UFE_HIP(hipMemcpyAsync(hbks.data(), obks.p, sizeof(double) * hbks.size(), hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    for (int f = 0; f < 7; ++f)
      if (ak && ak[f]) {
        std::vector<double> t(hak.begin() + f * nak, hak.begin() + (f + 1) * nak);
        layers_down(vperm, t, ak[f], nz);
      }
    for (int f = 0; f < 6; ++f) {
      if (bk && bk[f]) {
        std::vector<double> t(hbk.begin() + f * nbk, hbk.begin() + (f + 1) * nbk);
        layers_down(tperm, t, bk[f], nz);
      }
      if (bks && bks[f]) {
        std::vector<double> t(hbks.begin() + f * nbks, hbks.begin() + (f + 1) * nbks);
        layers_down(tperm, t, bks[f], nz - 1);
      }
    }
    return UFE_OK;
  });
}

extern "C" int ufe_solve_bpa(ufe_ctx *c, const ufe_bpa_inputs *in, double *u_bk, double *v_bk,
                             ufe_solve_info *info) {
  if (!c || !u_bk || !v_bk) return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    BpaState *bp = state(c);
    if (!bp) return ctx_fail(c, UFE_ERR_ARGUMENT, "ufe_zeta_setup has not been called");
    BpaState &b = *bp;
    const ufe_params &p = ctx_params(c);
    int rc = check_inputs(c, in);
    if (rc != UFE_OK) return rc;
    if ((rc = check_bpa_params(c, p)) != UFE_OK) return rc;
    const int nT = b.nT, nz = b.nz;
    const auto t0 = std::chrono::steady_clock::now();
    if (info) std::memset(info, 0, sizeof(*info));
    // no grounded ice: no velocities (BPA_main.f90:123-131)
    if (in->mask_grounded_ice) {
      bool any = false;
      for (int vi = 0; vi < b.nV && !any; ++vi) any = in->mask_grounded_ice[vi] != 0;
      if (!any) {
        std::fill(u_bk, u_bk + (size_t)nT * nz, 0.0);
        std::fill(v_bk, v_bk + (size_t)nT * nz, 0.0);
        return UFE_OK;
      }
    }
    hipStream_t s = ctx_stream(c);
    bpa_pattern(c, b);
    bpa_inputs(c, b, in, true);
    velocity_up(c, b, u_bk, v_bk);
    set_column_pc(c, b);
    const BsrT<double> A = bpa_bsr(b);
    const size_t nbk = (size_t)b.nb;
    double resid = 1e9, resid_prev;
    int nit_div = 0, it = 0, n_axb = 0;
    double relax = p.visc_it_relax, eps0 = p.epsilon_sq_0;
    bool converged = false;
    rc = UFE_OK;
    RedSlot slot{b.part.as<double>(), b.out.as<double>()};
    for (double &t : b.t_ms) t = 0.0;
    while (!converged) {
      ++it;
      UFE_HIP(hipMemcpyAsync(b.xprev.p, b.x.p, sizeof(double2) * nbk, hipMemcpyDeviceToDevice, s));
      UFE_HIP(hipEventRecord(b.ev[0], s));
      bpa_state_and_system(c, b, bpa_params(c, eps0), b.tribi.as<int>());
      UFE_HIP(hipEventRecord(b.ev[2], s));
      col_factor(c, b);
      UFE_HIP(hipEventRecord(b.ev[3], s));
      const SolveStats st = b.kry->gmres64(A, b.rhs.as<double>(), b.x.as<double>(), p.rtol, p.abstol,
                                           p.krylov_maxit, !p.warm_start, p.gmres_restart, std::max(1, p.check_every));
      UFE_HIP(hipEventRecord(b.ev[4], s));
      n_axb += st.its;
      // L2 change of the relaxed iterate; one pass per kMaxPartials * kBlock rows
      double r1 = 0.0, r2 = 0.0;
      for (long off = 0; off < (long)nbk; off += (long)kMaxPartials * kBlock) {
        const long n = std::min((long)nbk - off, (long)kMaxPartials * kBlock);
        const int g = blocks_for(n);
        hipLaunchKernelGGL(k_bpa_update, dim3(g), dim3(kBlock), 0, s, n, p.vel_max, relax,
                           b.x.as<double2>() + off, b.xprev.as<double2>() + off, slot);
        UFE_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_partials_out<2>, dim3(1), dim3(kBlock), 0, s, slot, g, b.out.as<double>());
        UFE_LAUNCH_CHECK();
        double h2[2];
        UFE_HIP(hipMemcpyAsync(h2, b.out.p, sizeof(h2), hipMemcpyDeviceToHost, s));
        UFE_HIP(hipStreamSynchronize(s));
        r1 += h2[0];
        r2 += h2[1];
      }
      UFE_HIP(hipEventRecord(b.ev[5], s));
      UFE_HIP(hipEventSynchronize(b.ev[5]));
      for (int q = 0; q < 5; ++q) {
        float ms = 0.0f;
        UFE_HIP(hipEventElapsedTime(&ms, b.ev[q], b.ev[q + 1]));
        b.t_ms[q] += ms;
      }
      resid_prev = resid;
      resid = 2.0 * r1 / std::max(r2, 1e-8);
      if (!std::isfinite(resid)) { rc = ctx_fail(c, UFE_ERR_DIVERGED, "BPA viscosity iteration produced NaN"); break; }
      if (resid > resid_prev) ++nit_div; else nit_div = 0;
      if (nit_div > 2) { nit_div = 0; relax *= 0.9; eps0 *= 1.2; }
      if (relax <= 0.05 || eps0 >= 1e-5) {
        if (relax < 0.05) {
          rc = ctx_fail(c, UFE_ERR_DIVERGED, "viscosity iteration still diverges even with very low relaxation factor!");
          break;
        }
        if (eps0 > 1e-5) {
          rc = ctx_fail(c, UFE_ERR_DIVERGED,
                        "viscosity iteration still diverges even with very high effective strain rate regularisation!");
          break;
        }
      }
      if (resid < p.visc_it_norm_dUV_tol) converged = true;
      if (it > p.visc_it_nit) { rc = UFE_WARN_NOT_CONVERGED; break; }
    }
    b.kry->set_custom_pc(nullptr);
    if (rc < 0) return rc;
    velocity_down(c, b, u_bk, v_bk);
    if (info) {
      UFE_HIP(hipMemcpyAsync(b.xprev.p, b.x.p, sizeof(double2) * nbk, hipMemcpyDeviceToDevice, s));
      bpa_state_and_system(c, b, bpa_params(c, eps0), b.tribi.as<int>());
      const double rel = bpa_rel_residual(b);
      info->n_visc_its = it;
      info->n_Axb_its = n_axb;
      info->L2_uv = resid;
      info->rel_residual = rel;
      info->rel_residual_unscaled = rel;
      info->visc_it_relax_applied = relax;
      info->epsilon_sq_0_applied = eps0;
      info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return rc;
  });
}

extern "C" int ufe_bpa_residual(ufe_ctx *c, const ufe_bpa_inputs *in, const double *u_bk, const double *v_bk,
                                double *r, double *bvec, double *eta_bk, double *beta_b, double *rel_residual) {
  if (!c || !u_bk || !v_bk) return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    BpaState *bp = state(c);
    if (!bp) return ctx_fail(c, UFE_ERR_ARGUMENT, "ufe_zeta_setup has not been called");
    BpaState &b = *bp;
    const ufe_params &p = ctx_params(c);
    int rc = check_inputs(c, in);
    if (rc != UFE_OK) return rc;
    if ((rc = check_bpa_params(c, p)) != UFE_OK) return rc;
    hipStream_t s = ctx_stream(c);
    const int nT = b.nT, nz = b.nz;
    const std::vector<int> &tperm = ctx_tperm(c);
    bpa_pattern(c, b);
    bpa_inputs(c, b, in, true);
    velocity_up(c, b, u_bk, v_bk);
    const size_t nbk = (size_t)b.nb;
    UFE_HIP(hipMemcpyAsync(b.xprev.p, b.x.p, sizeof(double2) * nbk, hipMemcpyDeviceToDevice, s));
    bpa_state_and_system(c, b, bpa_params(c, p.epsilon_sq_0), b.tribi.as<int>());
    const double rel = bpa_rel_residual(b);
    if (rel_residual) *rel_residual = rel;
    auto rows_down = [&](const DBuf &src, double *dst) {
      std::vector<double> h(2 * nbk);
      UFE_HIP(hipMemcpyAsync(h.data(), src.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
      UFE_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < nT; ++i)
        for (int k = 0; k < nz; ++k)
          for (int uv = 0; uv < 2; ++uv)
            dst[((size_t)tperm[i] * nz + k) * 2 + uv] = h[((size_t)i * nz + k) * 2 + uv];
    };
    if (r) rows_down(b.r, r);
    if (bvec) rows_down(b.rhs, bvec);
    if (eta_bk) {
      std::vector<double> h(4 * nbk);
      UFE_HIP(hipMemcpyAsync(h.data(), b.coef.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
      UFE_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < nT; ++i)
        for (int k = 0; k < nz; ++k) eta_bk[(size_t)k * nT + tperm[i]] = h[4 * ((size_t)i * nz + k)];
    }
    if (beta_b) {
      std::vector<double> h(nT);
      UFE_HIP(hipMemcpyAsync(h.data(), b.beta_b.p, sizeof(double) * nT, hipMemcpyDeviceToHost, s));
      UFE_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < nT; ++i) beta_b[tperm[i]] = h[i];
    }
    return UFE_OK;
  });
}

extern "C" int ufe_bpa_last_timings(const ufe_ctx *c, double *ms5) {
  if (!c || !ms5) return ctx_fail(const_cast<ufe_ctx *>(c), UFE_ERR_ARGUMENT, "null argument");
  BpaState *b = state(const_cast<ufe_ctx *>(c));
  if (!b) return ctx_fail(const_cast<ufe_ctx *>(c), UFE_ERR_ARGUMENT, "ufe_zeta_setup has not been called");
  for (int i = 0; i < 5; ++i) ms5[i] = b->t_ms[i];
  return UFE_OK;
}

extern "C" int ufe_bpa_profile_spmv(ufe_ctx *c, int reps, double *avg_ms, double *bytes) {
  if (!c || !avg_ms || !bytes || reps < 1) return ctx_fail(c, UFE_ERR_ARGUMENT, "bad argument");
  return ctx_guard(c, [&]() -> int {
    BpaState *b = state(c);
    if (!b || !b->nb) return ctx_fail(c, UFE_ERR_ARGUMENT, "no BPA system has been assembled");
    *avg_ms = b->kry->time_spmv64(bpa_bsr(*b), b->x.as<double>(), b->r.as<double>(), reps, bytes);
    return UFE_OK;
  });
}
