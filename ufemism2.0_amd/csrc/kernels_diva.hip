// DIVA kernels (Lipscomb et al. 2019 as implemented in DIVA_main.f90).
//
// Device layout of 3-D fields is layer-fastest ([vertex][k], [triangle][k]):
// one thread walks one column with unit stride and the ζ-integrals run in
// registers.  The C ABI converts from/to the reference's (n, nz) layout.
#include "kernels.h"

namespace ufe {

constexpr int kMaxNz = 64;
// Column kernels are instantiated for at most kRegNz layers (the per-column
// arrays then stay in registers: every index is static after unrolling) and
// for kMaxNz layers (arrays in per-lane scratch memory).
constexpr int kRegNz = 16;

// One thread per vertex:
//   strain rates from u_vav (b->a ddx/ddy)            calc_horizontal_strain_rates
//   du/dz, dv/dz on the b-grid from tau_b and the previous eta_3D_b, mapped
//   b->a (DIVA_main.f90:396-405), 3-D Glen viscosity with clamp (:441-455),
//   vertical average and N = eta_vav * max(0.1, Hi) (:461-469; the reference
//   writes a default-real 0.1 there), F1/F2 integrals from the base
//   (:497-511, integrate_from_zeta_is_one_to_zeta_is_zetap, mesh_zeta.f90:163-194),
//   basal friction from u_base (:540) and beta_eff = beta / (1 + beta F2(surface))
//   (:542-557).
template <int NZ>
__global__ void __launch_bounds__(kBlock) k_vertex_diva(
    int nV, int nz, const int *__restrict__ ptr, const int *__restrict__ ind, const double *__restrict__ wmap,
    const double *__restrict__ wdx, const double *__restrict__ wdy, const double *__restrict__ zeta,
    const double *__restrict__ x, const double *__restrict__ xbase, const double *__restrict__ taub,
    const double *__restrict__ eta3b, const double *__restrict__ Apow_vk, const double *__restrict__ Hi,
    const double *__restrict__ rough, PicardParams p, double *__restrict__ N_a, double *__restrict__ beta_a,
    double *__restrict__ beta_eff_a, double *__restrict__ eta3a, double *__restrict__ F1a,
    double *__restrict__ F2a) {
  int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi >= nV) return;
  double ux = 0, uy = 0, vx = 0, vy = 0, ub = 0, vb = 0;
  double duz[NZ], dvz[NZ], et[NZ];
#pragma unroll
  for (int k = 0; k < NZ; ++k) { duz[k] = 0.0; dvz[k] = 0.0; et[k] = 0.0; }
  for (int q = ptr[vi]; q < ptr[vi + 1]; ++q) {
    int t = ind[q];
    double2 uv = reinterpret_cast<const double2 *>(x)[t];
    double2 bs = reinterpret_cast<const double2 *>(xbase)[t];
    double2 tb = reinterpret_cast<const double2 *>(taub)[t];
    double cx = wdx[q], cy = wdy[q], cm = wmap[q];
    ux += cx * uv.x; uy += cy * uv.x; vx += cx * uv.y; vy += cy * uv.y;
    ub += cm * bs.x; vb += cm * bs.y;
    const double *e = eta3b + (size_t)t * nz;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      if (k < nz) {
        double inv = zeta[k] / fmax(p.visc_eff_min, e[k]);
        duz[k] += cm * tb.x * inv;
        dvz[k] += cm * tb.y * inv;
      }
    }
  }
  double *eo = eta3a + (size_t)vi * nz;
  const double *Ap = Apow_vk + (size_t)vi * nz;
  double eprev = 0.0, evav = 0.0;
#pragma unroll
  for (int k = 0; k < NZ; ++k) {
    if (k < nz) {
      double e2 = ux * ux + vy * vy + ux * vy + 0.25 * (uy + vx) * (uy + vx) +
                  0.25 * (duz[k] * duz[k] + dvz[k] * dvz[k]) + p.eps0;
      double eta = 0.5 * Ap[k] * pow(e2, (1.0 - p.n) / (2.0 * p.n));
      eta = fmin(fmax(eta, p.visc_eff_min), p.eta_max);
      eo[k] = eta;
      et[k] = eta;
      if (k > 0) evav += 0.5 * (eta + eprev) * (zeta[k] - zeta[k - 1]);
      eprev = eta;
    }
  }
  double h = fmax(0.1, Hi[vi]);
  N_a[vi] = evav * fmax((double)0.1f, Hi[vi]);
  double *f1 = F1a + (size_t)vi * nz, *f2 = F2a + (size_t)vi * nz;
  double I1 = 0.0, I2 = 0.0, f2s = 0.0;
  f1[nz - 1] = 0.0;
  f2[nz - 1] = 0.0;
#pragma unroll
  for (int k = NZ - 2; k >= 0; --k) {
    if (k <= nz - 2) {
      double z0 = zeta[k], z1 = zeta[k + 1], e0 = et[k], e1 = et[k + 1];
      I1 = I1 - 0.5 * (z1 / e1 + z0 / e0) * (z1 - z0);
      I2 = I2 - 0.5 * (z1 * z1 / e1 + z0 * z0 / e0) * (z1 - z0);
      f1[k] = -h * I1;
      f2[k] = -h * I2;
      f2s = -h * I2;
    }
  }
  const double b = basal_friction(p.sl, vi, rough[vi], ub, vb);
  beta_a[vi] = b;
  // F2 at the surface (k = 0); 0 for a single-layer column
  if (nz < 2) f2s = 0.0;
  beta_eff_a[vi] = p.no_sliding ? 1.0 / f2s : b / (1.0 + b * f2s);
}

// One thread per triangle: eta_3D_b = map_a_b_3D(eta_3D_a) (DIVA_main.f90:458),
// F2_3D_b at the surface (needed for the basal velocities) and, at the end of
// the solve, the full F1_3D_b for calc_3D_velocities (:514).
template <int NZ>
__global__ void __launch_bounds__(kBlock) k_triangle_3d(int nT, int nz, const int *__restrict__ ptr,
                                                        const int *__restrict__ ind, const double *__restrict__ wmap,
                                                        const double *__restrict__ eta3a,
                                                        const double *__restrict__ F1a,
                                                        const double *__restrict__ F2a, double *__restrict__ eta3b,
                                                        double *__restrict__ F1b, double *__restrict__ F2b0,
                                                        int full_F1) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  double e[NZ], f[NZ];
#pragma unroll
  for (int k = 0; k < NZ; ++k) { e[k] = 0.0; f[k] = 0.0; }
  double f20 = 0.0;
  for (int q = ptr[ti]; q < ptr[ti + 1]; ++q) {
    int v = ind[q];
    double cm = wmap[q];
    const double *ev = eta3a + (size_t)v * nz;
#pragma unroll
    for (int k = 0; k < NZ; ++k)
      if (k < nz) e[k] += cm * ev[k];
    f20 += cm * F2a[(size_t)v * nz];
    if (full_F1) {
      const double *fv = F1a + (size_t)v * nz;
#pragma unroll
      for (int k = 0; k < NZ; ++k)
        if (k < nz) f[k] += cm * fv[k];
    }
  }
  double *eo = eta3b + (size_t)ti * nz;
#pragma unroll
  for (int k = 0; k < NZ; ++k)
    if (k < nz) eo[k] = e[k];
  F2b0[ti] = f20;
  if (full_F1) {
    double *fo = F1b + (size_t)ti * nz;
#pragma unroll
    for (int k = 0; k < NZ; ++k)
      if (k < nz) fo[k] = f[k];
  }
}

// ---------------------------------------------------------------------------
// Column-parallel forms of the two kernels above for nz <= 16: a group of 16
// lanes works on one vertex (one triangle) and lane k holds layer k, so the
// per-layer loads (eta_3D_b of the neighbours, A^(-1/n)) and stores (eta_3D_a,
// F1, F2; eta_3D_b, F1_3D_b) of a group are contiguous and a wavefront's
// accesses coalesce, instead of 64 lanes each walking its own column with a
// stride of nz doubles.  The vertical average and the ζ-integrals are
// sequential over the layers: every lane of the group runs the same loop over
// the layers, reading layer j from lane j with a cross-lane shuffle, so the
// sums are formed in the order of the per-thread kernels above and the results
// are identical to theirs.
// ---------------------------------------------------------------------------
constexpr int kColLanes = 16;

__global__ void __launch_bounds__(kBlock) k_vertex_diva_cols(
    int nV, int nz, const int *__restrict__ ptr, const int *__restrict__ ind, const double *__restrict__ wmap,
    const double *__restrict__ wdx, const double *__restrict__ wdy, const double *__restrict__ zeta,
    const double *__restrict__ x, const double *__restrict__ xbase, const double *__restrict__ taub,
    const double *__restrict__ eta3b, const double *__restrict__ Apow_vk, const double *__restrict__ Hi,
    const double *__restrict__ rough, PicardParams p, double *__restrict__ N_a, double *__restrict__ beta_a,
    double *__restrict__ beta_eff_a, double *__restrict__ eta3a, double *__restrict__ F1a,
    double *__restrict__ F2a) {
  const int k = threadIdx.x & (kColLanes - 1);
  const int vi = (int)((blockIdx.x * (long)blockDim.x + threadIdx.x) / kColLanes);
  if (vi >= nV) return;  // the whole group leaves together
  const bool act = k < nz;
  const double zk = act ? zeta[k] : 0.0;
  const int q0 = ptr[vi], nq = ptr[vi + 1] - q0;
  double ux = 0, uy = 0, vx = 0, vy = 0, ub = 0, vb = 0, duz = 0.0, dvz = 0.0;
  // the row's entries, up to kRowChunk at a time: lane j loads entry j (its
  // triangle, weights and 2-D fields) and every lane loads its layer of each
  // entry's eta_3D_b, all before the first use, so the gathers of a row are in
  // flight together instead of one dependent chain per entry (b->a rows have
  // 3 to 9 entries on these meshes)
  constexpr int kRowChunk = 8;
  for (int c0 = 0; c0 < nq; c0 += kRowChunk) {
    const int cnt = min(kRowChunk, nq - c0);
    int tq = 0;
    double cxq = 0.0, cyq = 0.0, cmq = 0.0;
    double2 uvq = make_double2(0.0, 0.0), bsq = uvq, tbq = uvq;
    if (k < cnt) {
      const int q = q0 + c0 + k;
      tq = ind[q];
      cxq = wdx[q]; cyq = wdy[q]; cmq = wmap[q];
      uvq = reinterpret_cast<const double2 *>(x)[tq];
      bsq = reinterpret_cast<const double2 *>(xbase)[tq];
      tbq = reinterpret_cast<const double2 *>(taub)[tq];
    }
    double eq[kRowChunk];
#pragma unroll
    for (int j = 0; j < kRowChunk; ++j) {
      const int tj = __shfl(tq, j, kColLanes);
      eq[j] = (j < cnt && act) ? eta3b[(size_t)tj * nz + k] : 1.0;
    }
    // sums in row order
#pragma unroll
    for (int j = 0; j < kRowChunk; ++j) {
      if (j < cnt) {
        const double cx = __shfl(cxq, j, kColLanes), cy = __shfl(cyq, j, kColLanes);
        const double cm = __shfl(cmq, j, kColLanes);
        const double uvx = __shfl(uvq.x, j, kColLanes), uvy = __shfl(uvq.y, j, kColLanes);
        const double bsx = __shfl(bsq.x, j, kColLanes), bsy = __shfl(bsq.y, j, kColLanes);
        const double tbx = __shfl(tbq.x, j, kColLanes), tby = __shfl(tbq.y, j, kColLanes);
        ux += cx * uvx; uy += cy * uvx; vx += cx * uvy; vy += cy * uvy;
        ub += cm * bsx; vb += cm * bsy;
        if (act) {
          const double inv = zk / fmax(p.visc_eff_min, eq[j]);
          duz += cm * tbx * inv;
          dvz += cm * tby * inv;
        }
      }
    }
  }
  double eta = 0.0;
  if (act) {
    const double e2 = ux * ux + vy * vy + ux * vy + 0.25 * (uy + vx) * (uy + vx) +
                      0.25 * (duz * duz + dvz * dvz) + p.eps0;
    eta = 0.5 * Apow_vk[(size_t)vi * nz + k] * pow(e2, (1.0 - p.n) / (2.0 * p.n));
    eta = fmin(fmax(eta, p.visc_eff_min), p.eta_max);
    eta3a[(size_t)vi * nz + k] = eta;
  }
  // vertical average (trapezoid from the surface down, in layer order)
  double evav = 0.0, eprev = __shfl(eta, 0, kColLanes);
  for (int j = 1; j < nz; ++j) {
    const double ej = __shfl(eta, j, kColLanes);
    evav += 0.5 * (ej + eprev) * (zeta[j] - zeta[j - 1]);
    eprev = ej;
  }
  // F1, F2: integrals from the base (zeta = 1) up to layer k
  const double h = fmax(0.1, Hi[vi]);
  const double q1 = act ? zk / eta : 0.0, q2 = act ? zk * zk / eta : 0.0;
  double q1n = __shfl(q1, nz - 1, kColLanes), q2n = __shfl(q2, nz - 1, kColLanes);
  double I1 = 0.0, I2 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int j = nz - 2; j >= 0; --j) {
    const double q1j = __shfl(q1, j, kColLanes), q2j = __shfl(q2, j, kColLanes);
    const double z0 = zeta[j], z1 = zeta[j + 1];
    I1 = I1 - 0.5 * (q1n + q1j) * (z1 - z0);
    I2 = I2 - 0.5 * (q2n + q2j) * (z1 - z0);
    if (k == j) { f1 = -h * I1; f2 = -h * I2; }
    q1n = q1j;
    q2n = q2j;
  }
  if (act) {
    F1a[(size_t)vi * nz + k] = f1;
    F2a[(size_t)vi * nz + k] = f2;
  }
  if (k != 0) return;
  const double f2s = nz < 2 ? 0.0 : -h * I2;  // F2 at the surface (k = 0)
  N_a[vi] = evav * fmax((double)0.1f, Hi[vi]);
  const double b = basal_friction(p.sl, vi, rough[vi], ub, vb);
  beta_a[vi] = b;
  beta_eff_a[vi] = p.no_sliding ? 1.0 / f2s : b / (1.0 + b * f2s);
}

__global__ void __launch_bounds__(kBlock) k_triangle_3d_cols(int nT, int nz, const int *__restrict__ ptr,
                                                             const int *__restrict__ ind,
                                                             const double *__restrict__ wmap,
                                                             const double *__restrict__ eta3a,
                                                             const double *__restrict__ F1a,
                                                             const double *__restrict__ F2a,
                                                             double *__restrict__ eta3b, double *__restrict__ F1b,
                                                             double *__restrict__ F2b0, int full_F1) {
  const int k = threadIdx.x & (kColLanes - 1);
  const int ti = (int)((blockIdx.x * (long)blockDim.x + threadIdx.x) / kColLanes);
  if (ti >= nT) return;
  const bool act = k < nz;
  const int q0 = ptr[ti], nq = ptr[ti + 1] - q0;
  double e = 0.0, f = 0.0, f20 = 0.0;
  // entries up to kColLanes at a time: lane j loads entry j's vertex, weight
  // and surface F2; every lane then loads its layer of each entry's columns
  for (int c0 = 0; c0 < nq; c0 += kColLanes) {
    const int cnt = min(kColLanes, nq - c0);
    int vq = 0;
    double cmq = 0.0, f2q = 0.0;
    if (k < cnt) {
      vq = ind[q0 + c0 + k];
      cmq = wmap[q0 + c0 + k];
      f2q = F2a[(size_t)vq * nz];
    }
    constexpr int kE = 4;  // a->b rows have three entries; longer rows take several rounds
    for (int j0 = 0; j0 < cnt; j0 += kE) {
      double ea[kE], fa[kE];
#pragma unroll
      for (int j = 0; j < kE; ++j) {
        const int vj = __shfl(vq, j0 + j, kColLanes);
        const bool ok = j0 + j < cnt && act;
        ea[j] = ok ? eta3a[(size_t)vj * nz + k] : 0.0;
        fa[j] = ok && full_F1 ? F1a[(size_t)vj * nz + k] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < kE; ++j) {
        const double cm = __shfl(cmq, j0 + j, kColLanes), f2 = __shfl(f2q, j0 + j, kColLanes);
        if (j0 + j < cnt) {
          if (act) {
            e += cm * ea[j];
            if (full_F1) f += cm * fa[j];
          }
          f20 += cm * f2;
        }
      }
    }
  }
  if (act) {
    eta3b[(size_t)ti * nz + k] = e;
    if (full_F1) F1b[(size_t)ti * nz + k] = f;
  }
  if (k == 0) F2b0[ti] = f20;
}

// calc_3D_velocities (DIVA_main.f90:636-676); outputs in the reference's
// (nT, nz) triangle-fastest layout.
__global__ void __launch_bounds__(kBlock) k_velocities_3d(int nT, int nz, const double *__restrict__ ubase,
                                                          const double *__restrict__ beta_b,
                                                          const double *__restrict__ taub,
                                                          const double *__restrict__ F1b, int no_sliding,
                                                          double *__restrict__ u3, double *__restrict__ v3) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  double2 ub = reinterpret_cast<const double2 *>(ubase)[ti];
  double2 tb = reinterpret_cast<const double2 *>(taub)[ti];
  double bb = beta_b[ti];
  for (int k = 0; k < nz; ++k) {
    double f = F1b[(size_t)ti * nz + k];
    size_t o = (size_t)k * nT + ti;
    if (no_sliding) { u3[o] = tb.x * f; v3[o] = tb.y * f; }
    else { u3[o] = ub.x * (1.0 + bb * f); v3[o] = ub.y * (1.0 + bb * f); }
  }
}

// out[c][r] = in[r][c] for an rows x cols row-major matrix (LDS-tiled).
__global__ void __launch_bounds__(256) k_transpose(int rows, int cols, const double *__restrict__ in,
                                                   double *__restrict__ out) {
  __shared__ double tile[32][33];
  int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    int r = r0 + j, c = c0 + tx;
    if (r < rows && c < cols) tile[j][tx] = in[(size_t)r * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    int c = c0 + j, r = r0 + tx;
    if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[tx][j];
  }
}

void launch_vertex_diva(const DevMesh &m, const double *x, const double *xbase, const double *taub,
                        const double *eta3b_tk, const double *Apow_vk, const double *Hi, const double *rough,
                        const PicardParams &p, double *N_a, double *beta_a, double *beta_eff_a, double *eta3a_vk,
                        double *F1a_vk, double *F2a_vk, hipStream_t s) {
  if (m.nz <= kColLanes && m.diva_columns)
    hipLaunchKernelGGL(k_vertex_diva_cols, dim3(blocks_for((long)m.nV * kColLanes)), dim3(kBlock), 0, s, m.nV, m.nz,
                       m.ba_ptr, m.ba_ind, m.ba_map, m.ba_ddx, m.ba_ddy, m.zeta, x, xbase, taub, eta3b_tk, Apow_vk,
                       Hi, rough, p, N_a, beta_a, beta_eff_a, eta3a_vk, F1a_vk, F2a_vk);
  else if (m.nz <= kRegNz)
    hipLaunchKernelGGL(k_vertex_diva<kRegNz>, dim3(blocks_for(m.nV)), dim3(kBlock), 0, s, m.nV, m.nz, m.ba_ptr,
                       m.ba_ind, m.ba_map, m.ba_ddx, m.ba_ddy, m.zeta, x, xbase, taub, eta3b_tk, Apow_vk, Hi, rough, p,
                       N_a, beta_a, beta_eff_a, eta3a_vk, F1a_vk, F2a_vk);
  else
    hipLaunchKernelGGL(k_vertex_diva<kMaxNz>, dim3(blocks_for(m.nV)), dim3(kBlock), 0, s, m.nV, m.nz, m.ba_ptr,
                       m.ba_ind, m.ba_map, m.ba_ddx, m.ba_ddy, m.zeta, x, xbase, taub, eta3b_tk, Apow_vk, Hi, rough, p,
                       N_a, beta_a, beta_eff_a, eta3a_vk, F1a_vk, F2a_vk);
  UFE_LAUNCH_CHECK();
}

void launch_triangle_3d(const DevMesh &m, const double *eta3a_vk, const double *F1a_vk, const double *F2a_vk,
                        double *eta3b_tk, double *F1b_tk, double *F2b0, int full_F1, hipStream_t s) {
  if (m.nz <= kColLanes && m.diva_columns)
    hipLaunchKernelGGL(k_triangle_3d_cols, dim3(blocks_for((long)m.nT * kColLanes)), dim3(kBlock), 0, s, m.nT, m.nz,
                       m.ab_ptr, m.ab_ind, m.ab_map, eta3a_vk, F1a_vk, F2a_vk, eta3b_tk, F1b_tk, F2b0, full_F1);
  else if (m.nz <= kRegNz)
    hipLaunchKernelGGL(k_triangle_3d<kRegNz>, dim3(blocks_for(m.nT)), dim3(kBlock), 0, s, m.nT, m.nz, m.ab_ptr,
                       m.ab_ind, m.ab_map, eta3a_vk, F1a_vk, F2a_vk, eta3b_tk, F1b_tk, F2b0, full_F1);
  else
    hipLaunchKernelGGL(k_triangle_3d<kMaxNz>, dim3(blocks_for(m.nT)), dim3(kBlock), 0, s, m.nT, m.nz, m.ab_ptr,
                       m.ab_ind, m.ab_map, eta3a_vk, F1a_vk, F2a_vk, eta3b_tk, F1b_tk, F2b0, full_F1);
  UFE_LAUNCH_CHECK();
}

void launch_velocities_3d(int nT, int nz, const double *ubase, const double *beta_basal_b, const double *taub,
                          const double *F1b_tk, int no_sliding, double *u3_out, double *v3_out, hipStream_t s) {
  hipLaunchKernelGGL(k_velocities_3d, dim3(blocks_for(nT)), dim3(kBlock), 0, s, nT, nz, ubase, beta_basal_b, taub,
                     F1b_tk, no_sliding, u3_out, v3_out);
  UFE_LAUNCH_CHECK();
}

// A^(-1/n) of Glen's law per vertex and layer (constitutive_equation.f90:25-51),
// once per solve: the flow factor does not change between Picard iterations,
// so the vertex kernel reads this instead of evaluating the power each time.
// Same expression as the kernel used inline, so the products are unchanged.
__global__ void __launch_bounds__(kBlock) k_flow_factor_pow(long n, const double *__restrict__ A, double glen_n,
                                                            double *__restrict__ out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) out[i] = pow(A[i], -1.0 / glen_n);
}

void launch_flow_factor_pow(long n, const double *A, double glen_n, double *out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_flow_factor_pow, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, A, glen_n, out);
  UFE_LAUNCH_CHECK();
}

void launch_transpose(int rows, int cols, const double *in, double *out, hipStream_t s) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, rows, cols, in, out);
  UFE_LAUNCH_CHECK();
}

}  // namespace ufe
