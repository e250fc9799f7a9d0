// Anderson mixing of the Picard fixed-point map (engine extension; the
// reference relaxes with a fixed factor instead, SSA_DIVA_utilities.f90:84-108).
//
// With G(u) the velocity after one linearised solve and the velocity limits,
// f_k = G(u_k) - u_k, and the last m differences dF_j = f_j+1 - f_j,
// dG_j = G(u_j+1) - G(u_j) kept in a ring, the next iterate is
//   u_k+1 = (1 - mix) (u_k - sum_j gamma_j (dG_j - dF_j)) + mix (G(u_k) - sum_j gamma_j dG_j)
// with gamma the least-squares minimiser of ||f_k - dF gamma||_2 (Walker & Ni,
// SIAM J. Numer. Anal. 49 (2011), "type II").  The two kernels below do every
// pass over the n = 2 nT unknowns: one writes the new ring column and the dot
// products of the small normal equations, the other forms the mixed iterate.
#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace ufe {

constexpr int kAAW = 2 * kAAMaxDepth + 1;

// g := clamp(x) (apply_velocity_limits), f = g - u; if have_prev, ring column
// p := (f - fprev, g - gprev); fprev := f, gprev := g.  Grid sums:
// out[j] = dF_j . dF_p, out[m + j] = dF_j . f (j < ncols), out[2m] = f . f
// (m = kAAMaxDepth), written by the second-phase kernel.
__global__ void __launch_bounds__(kBlock) k_aa_columns(
    int nT, double *__restrict__ x, const double *__restrict__ u, double *__restrict__ fprev,
    double *__restrict__ gprev, double *__restrict__ dF, double *__restrict__ dG, long stride, int p,
    int have_prev, int ncols, double vel_max, RedSlot slot, double *__restrict__ out) {
  double acc[kAAW];
#pragma unroll
  for (int j = 0; j < kAAW; ++j) acc[j] = 0.0;
  for (int ti = blockIdx.x * blockDim.x + threadIdx.x; ti < nT; ti += gridDim.x * blockDim.x) {
    double2 g = reinterpret_cast<const double2 *>(x)[ti];
    double ua = sqrt(g.x * g.x + g.y * g.y);
    if (ua > vel_max) { g.x = g.x * vel_max / ua; g.y = g.y * vel_max / ua; }
    reinterpret_cast<double2 *>(x)[ti] = g;
    double2 uk = reinterpret_cast<const double2 *>(u)[ti];
    double2 f = make_double2(g.x - uk.x, g.y - uk.y);
    double2 dfp = make_double2(0.0, 0.0);
    if (have_prev) {
      double2 fp = reinterpret_cast<const double2 *>(fprev)[ti];
      double2 gp = reinterpret_cast<const double2 *>(gprev)[ti];
      dfp = make_double2(f.x - fp.x, f.y - fp.y);
      reinterpret_cast<double2 *>(dF + p * stride)[ti] = dfp;
      reinterpret_cast<double2 *>(dG + p * stride)[ti] = make_double2(g.x - gp.x, g.y - gp.y);
    }
    reinterpret_cast<double2 *>(fprev)[ti] = f;
    reinterpret_cast<double2 *>(gprev)[ti] = g;
#pragma unroll
    for (int j = 0; j < kAAMaxDepth; ++j) {
      if (j < ncols) {
        double2 d = j == p ? dfp : reinterpret_cast<const double2 *>(dF + j * stride)[ti];
        acc[j] += d.x * dfp.x + d.y * dfp.y;
        acc[kAAMaxDepth + j] += d.x * f.x + d.y * f.y;
      }
    }
    acc[2 * kAAMaxDepth] += f.x * f.x + f.y * f.y;
  }
  __shared__ double lds[kAAW][kBlock / 64];
  block_sum<kAAW>(acc, lds);
  block_partials<kAAW>(slot, acc);
}

// x := (1 - mix)(u - sum gamma_j (dG_j - dF_j)) + mix (x - sum gamma_j dG_j),
// x holding G(u) on entry.
__global__ void __launch_bounds__(kBlock) k_aa_mix(int n, double *__restrict__ x, const double *__restrict__ u,
                                                   const double *__restrict__ dF, const double *__restrict__ dG,
                                                   long stride, int ncols, AAGamma gam, double mix) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sg = 0.0, sf = 0.0;
#pragma unroll
  for (int j = 0; j < kAAMaxDepth; ++j) {
    if (j < ncols) {
      sg += gam.g[j] * dG[j * stride + i];
      sf += gam.g[j] * dF[j * stride + i];
    }
  }
  double g = x[i];
  x[i] = (1.0 - mix) * (u[i] - (sg - sf)) + mix * (g - sg);
}

void launch_aa_columns(int nT, double *x, const double *u, double *fprev, double *gprev, double *dF, double *dG,
                       long stride, int p, int have_prev, int ncols, double vel_max, const RedSlot &slot,
                       double *out, hipStream_t s) {
  // grid-stride over at most 1,024 workgroups: the one-workgroup second phase
  // then sums 17 x 1,024 partials instead of 17 x nT / 256
  int nb = std::min(blocks_for(nT), 1024);
  hipLaunchKernelGGL(k_aa_columns, dim3(nb), dim3(kBlock), 0, s, nT, x, u, fprev, gprev, dF, dG, stride, p,
                     have_prev, ncols, vel_max, slot, out);
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_partials_out<kAAW>, dim3(1), dim3(kBlock), 0, s, slot, nb, out);
  UFE_LAUNCH_CHECK();
}

void launch_aa_mix(int n, double *x, const double *u, const double *dF, const double *dG, long stride, int ncols,
                   const AAGamma &gam, double mix, hipStream_t s) {
  hipLaunchKernelGGL(k_aa_mix, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, u, dF, dG, stride, ncols, gam, mix);
  UFE_LAUNCH_CHECK();
}

// Solves (H + lambda I) g = b by Gaussian elimination with partial pivoting;
// false if the system is numerically singular.
static bool small_solve(int m, const double (&H)[kAAMaxDepth][kAAMaxDepth], const double *b, double *g) {
  double a[kAAMaxDepth][kAAMaxDepth + 1];
  double tr = 0.0;
  for (int i = 0; i < m; ++i) tr = std::max(tr, H[i][i]);
  if (!(tr > 0.0)) return false;
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < m; ++j) a[i][j] = H[i][j] + (i == j ? 1e-12 * tr : 0.0);
    a[i][m] = b[i];
  }
  for (int k = 0; k < m; ++k) {
    int piv = k;
    for (int i = k + 1; i < m; ++i)
      if (std::fabs(a[i][k]) > std::fabs(a[piv][k])) piv = i;
    if (!(std::fabs(a[piv][k]) > 1e-14 * tr)) return false;
    if (piv != k)
      for (int j = 0; j <= m; ++j) std::swap(a[k][j], a[piv][j]);
    for (int i = k + 1; i < m; ++i) {
      double f = a[i][k] / a[k][k];
      for (int j = k; j <= m; ++j) a[i][j] -= f * a[k][j];
    }
  }
  for (int i = m - 1; i >= 0; --i) {
    double t = a[i][m];
    for (int j = i + 1; j < m; ++j) t -= a[i][j] * g[j];
    g[i] = t / a[i][i];
  }
  return true;
}

void anderson_step(AndersonState &aa, const AndersonBuffers &buf, int depth, double mixing, double restart,
                   double vel_max, long n2, double *x, const double *u, double resid, hipStream_t s) {
  const int m = depth;
  const long stride = 2L * n2;
  if (resid > restart * aa.best) {
    aa.cols = 0;
    aa.pos = 0;
    ++aa.n_restarts;
  }
  aa.best = std::min(aa.best, resid);
  int p = -1;
  if (aa.have_prev) {
    p = aa.pos;
    aa.pos = (aa.pos + 1) % m;
    aa.cols = std::min(aa.cols + 1, m);
  }
  RedSlot sl;
  sl.partials = buf.part;
  sl.result = buf.out;
  launch_aa_columns((int)n2, x, u, buf.fprev, buf.gprev, buf.dF, buf.dG, stride, p < 0 ? 0 : p,
                    aa.have_prev ? 1 : 0, aa.cols, vel_max, sl, buf.out, s);
  double h[2 * kAAMaxDepth + 1];
  UFE_HIP(hipMemcpyAsync(h, buf.out, sizeof(h), hipMemcpyDeviceToHost, s));
  UFE_HIP(hipStreamSynchronize(s));
  aa.have_prev = true;
  AAGamma gam{};
  int ncols = aa.cols;
  if (p >= 0) {
    for (int j = 0; j < aa.cols; ++j) aa.H[p][j] = aa.H[j][p] = h[j];
    if (!small_solve(aa.cols, aa.H, h + kAAMaxDepth, gam.g)) {
      // degenerate history: plain mixing step, history restarts
      ncols = 0;
      aa.cols = 0;
      aa.pos = 0;
      ++aa.n_restarts;
    }
  }
  launch_aa_mix((int)stride, x, u, buf.dF, buf.dG, stride, ncols, gam, mixing, s);
}

}  // namespace ufe
