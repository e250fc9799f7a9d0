// Device Krylov solver (BiCGStab), BSR 2x2 SpMV and the aggregation
// multigrid preconditioner.  See krylov.h for the design.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include <cstring>
#include <functional>

#include "krylov.h"

namespace ufe {

void DBuf::alloc(size_t b) {
  free();
  if (b == 0) return;
  UFE_HIP(hipMalloc(&p, b));
  bytes = b;
}

void DBuf::free() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

// ===========================================================================
// BSR SpMV with fused epilogues.  LPR lanes cooperate on one block row: lane
// k of a row reads blocks ptr[row]+k, +LPR, ... so consecutive lanes read
// consecutive blocks (coalesced), then the partial sums are combined with
// cross-lane shuffles.  Reductions are accumulated in FP64.
// ===========================================================================
enum SpmvOp { OP_PLAIN = 0, OP_RESID = 1, OP_DOT1 = 2, OP_DOT2 = 3 };

constexpr int kLpr = 4;          // lanes per block row in the V-cycle kernels
// Block product shape: 4 lanes per block row, each with two blocks' loads in
// flight, 256-thread workgroups.  On a 796k-row stencil matrix of ~10 blocks
// per row (315 MB, beyond the Infinity Cache) this streams at 5.2-5.4 TB/s
// against 4.3-4.5 TB/s with 8 lanes per row (fewer idle lanes on the row's
// last pass, one shuffle stage less, twice the rows per workgroup), while a
// plain read stream of the same buffer reaches 7.0 TB/s.
constexpr int kSpmvLpr = 4;      // lanes per block row in the block product (stencil matrices)
constexpr int kSpmvBlock = 256;  // workgroup size of the block product
constexpr int kSpmvRedGrid = kMaxPartials;  // grid cap of the block product with a reduction epilogue

// Rows far longer than the 2-D stencil's (the BPA's block rows hold three
// layers of a b->b row, ~30-40 blocks) take more lanes per row
// (spmv_lanes): 8 lanes read 256 contiguous bytes of a row per pass.
template <typename T, int OP, int LPR = kSpmvLpr>
__global__ void __launch_bounds__(kSpmvBlock) k_bsr_spmv(
    int nb, const int *__restrict__ ptr, const int *__restrict__ col,
    const typename VecT<T>::V4 *__restrict__ val, const typename VecT<T>::V2 *__restrict__ x,
    typename VecT<T>::V2 *__restrict__ y, const typename VecT<T>::V2 *__restrict__ b,
    const typename VecT<T>::V2 *__restrict__ w, RedSlot slot, double *__restrict__ out,
    KrylovScalars *__restrict__ sc, double vmax, const typename VecT<T>::V4 *__restrict__ dblk,
    int write_unscaled) {
  using V2 = typename VecT<T>::V2;
  using V4 = typename VecT<T>::V4;
  if (sc && sc->done) return;
  const int lane = threadIdx.x % LPR;
  constexpr int kRowsPerPass = kSpmvBlock / LPR;
  // one row per LPR lanes; consecutive rows per workgroup, workgroups ordered
  // per XCD, grid-stride when the grid is capped
  const int stride = gridDim.x * kRowsPerPass;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int row = xcd_block() * kRowsPerPass + (int)threadIdx.x / LPR; row < nb; row += stride) {
    T y0 = 0, y1 = 0;
    if (x) {
      const int k1 = ptr[row + 1];
      // two blocks' gathers in flight per lane; the sums keep the block order
      int k = ptr[row] + lane;
      for (; k + LPR < k1; k += 2 * LPR) {
        const V4 a0 = val[k], a1 = val[k + LPR];
        const int c0 = col[k], c1 = col[k + LPR];
        const V2 xv0 = x[c0], xv1 = x[c1];
        y0 += a0.x * xv0.x + a0.y * xv0.y;
        y1 += a0.z * xv0.x + a0.w * xv0.y;
        y0 += a1.x * xv1.x + a1.y * xv1.y;
        y1 += a1.z * xv1.x + a1.w * xv1.y;
      }
      if (k < k1) {
        const V4 a = val[k];
        const V2 xv = x[col[k]];
        y0 += a.x * xv.x + a.y * xv.y;
        y1 += a.z * xv.x + a.w * xv.y;
      }
#pragma unroll
      for (int off = LPR / 2; off > 0; off >>= 1) {
        y0 += __shfl_down(y0, off, LPR);
        y1 += __shfl_down(y1, off, LPR);
      }
    }
    if (lane == 0) {
      if (OP == OP_RESID) {
        V2 bb = b[row];
        T r0 = bb.x - y0, r1 = bb.y - y1;
        V2 r;
        r.x = r0; r.y = r1;
        if (dblk && write_unscaled) {
          // write the residual of the unscaled system, r = D r_s (norms stay scaled)
          V4 d = dblk[row];
          r.x = d.x * r0 + d.y * r1;
          r.y = d.z * r0 + d.w * r1;
        }
        y[row] = r;
        // rows where the velocity limit is active are left out of the norms
        bool clamped = false;
        if (vmax > 0.0) {
          V2 xv = x[row];
          clamped = sqrt((double)xv.x * xv.x + (double)xv.y * xv.y) >= vmax * (1.0 - 1e-9);
        }
        if (clamped) acc[2] += 1.0;
        else {
          acc[0] += (double)r0 * r0 + (double)r1 * r1;
          acc[1] += (double)bb.x * bb.x + (double)bb.y * bb.y;
        }
      } else {
        V2 r;
        r.x = y0; r.y = y1;
        y[row] = r;
        if (OP == OP_DOT1) {
          V2 ww = w[row];
          acc[0] += (double)ww.x * y0 + (double)ww.y * y1;
        } else if (OP == OP_DOT2) {
          V2 ss = w[row];
          acc[0] += (double)y0 * ss.x + (double)y1 * ss.y;
          acc[1] += (double)y0 * y0 + (double)y1 * y1;
        }
      }
    }
  }
  if (OP == OP_PLAIN) return;
  __shared__ double lds[3][kSpmvBlock / 64];
  block_sum<3, kSpmvBlock>(acc, lds);
  block_partials<3>(slot, acc);
}

// Second phase of the product's reduction (one workgroup).
constexpr int kFinishBlock = 1024;  // the block product's partials: up to 8 * 1024 in one round trip
template <int OP>
__global__ void __launch_bounds__(kFinishBlock) k_spmv_finish(RedSlot slot, int nparts, double *__restrict__ out,
                                                              KrylovScalars *__restrict__ sc) {
  if (sc && sc->done) return;
  double tot[3];
  partials_sum<3, kFinishBlock>(slot, nparts, tot);
  if (threadIdx.x != 0) return;
  if (OP == OP_RESID) {
    out[0] = tot[0];
    out[1] = tot[1];
    out[2] = tot[2];
  } else if (OP == OP_DOT1) {
    // alpha = rho / (rhat, v)
    if (tot[0] == 0.0 || !std::isfinite(tot[0])) sc->done = 2;
    else sc->alpha = sc->rho / tot[0];
  } else {
    // omega = (t, s) / (t, t)
    sc->omega = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
  }
}

// lanes per block row: 8 for rows averaging 20 blocks or more, else kSpmvLpr
inline int spmv_lanes(int nb, long nnzb) { return nnzb >= 20L * nb ? 8 : kSpmvLpr; }

template <typename T, int OP>
static void spmv_launch(const BsrT<T> &A, const T *x, T *y, const T *b, const T *w, const RedSlot &slot,
                        double *out, KrylovScalars *sc, hipStream_t s, double vmax = 0.0,
                        const typename VecT<T>::V4 *dblk = nullptr, int write_unscaled = 0,
                        hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
  using V2 = typename VecT<T>::V2;
  const int lpr = spmv_lanes(A.nb, A.nnzb);
  long threads = (long)A.nb * lpr;
  // with a reduction epilogue, fewer workgroups with several row passes each
  // amortise the workgroup reduction
  int nblk = xcd_grid(blocks_for(threads, kSpmvBlock));
  if (OP != OP_PLAIN && nblk > kSpmvRedGrid) {
    // one partial sum per workgroup: above the cap, equal row passes per workgroup
    const int passes = (nblk + kSpmvRedGrid - 1) / kSpmvRedGrid;
    nblk = xcd_grid((nblk + passes - 1) / passes);
  }
  if (ev0) UFE_HIP(hipEventRecord(ev0, s));
  if (lpr == 8)
    hipLaunchKernelGGL((k_bsr_spmv<T, OP, 8>), dim3(nblk), dim3(kSpmvBlock), 0, s, A.nb, A.ptr, A.col, A.val,
                       reinterpret_cast<const V2 *>(x), reinterpret_cast<V2 *>(y), reinterpret_cast<const V2 *>(b),
                       reinterpret_cast<const V2 *>(w), slot, out, sc, vmax, dblk, write_unscaled);
  else
    hipLaunchKernelGGL((k_bsr_spmv<T, OP, kSpmvLpr>), dim3(nblk), dim3(kSpmvBlock), 0, s, A.nb, A.ptr, A.col,
                       A.val, reinterpret_cast<const V2 *>(x), reinterpret_cast<V2 *>(y),
                       reinterpret_cast<const V2 *>(b), reinterpret_cast<const V2 *>(w), slot, out, sc, vmax, dblk,
                       write_unscaled);
  UFE_LAUNCH_CHECK();
  if (ev1) UFE_HIP(hipEventRecord(ev1, s));
  if (OP != OP_PLAIN) {
    hipLaunchKernelGGL((k_spmv_finish<OP>), dim3(1), dim3(kFinishBlock), 0, s, slot, nblk, out, sc);
    UFE_LAUNCH_CHECK();
  }
}

// ===========================================================================
// Block-local preconditioners, fused into vector updates
// ===========================================================================
template <typename T>
__device__ __forceinline__ typename VecT<T>::V2 pc_local(int pc, const typename VecT<T>::V4 *__restrict__ dinv,
                                                         const typename VecT<T>::V4 *__restrict__ aval,
                                                         const int *__restrict__ diag, int i,
                                                         typename VecT<T>::V2 q) {
  typename VecT<T>::V2 z;
  if (pc == UFE_PC_BLOCK_JACOBI) {
    typename VecT<T>::V4 d = dinv[i];
    z.x = d.x * q.x + d.y * q.y;
    z.y = d.z * q.x + d.w * q.y;
  } else if (pc == UFE_PC_JACOBI) {
    typename VecT<T>::V4 a = aval[diag[i]];
    z.x = q.x / a.x;
    z.y = q.y / a.w;
  } else {
    z = q;
  }
  return z;
}

// p = r + beta (p - omega v); phat = M p   (beta from the device scalars)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_p(int nb, const typename VecT<T>::V2 *__restrict__ r,
                                                   typename VecT<T>::V2 *__restrict__ p,
                                                   const typename VecT<T>::V2 *__restrict__ v,
                                                   typename VecT<T>::V2 *__restrict__ phat, int pc,
                                                   const typename VecT<T>::V4 *__restrict__ dinv,
                                                   const typename VecT<T>::V4 *__restrict__ aval,
                                                   const int *__restrict__ diag, const KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  T beta = (T)((sc->rho / sc->rho_old) * (sc->alpha / sc->omega));
  T om = (T)sc->omega;
  typename VecT<T>::V2 rr = r[i], pp = p[i], vv = v[i];
  pp.x = rr.x + beta * (pp.x - om * vv.x);
  pp.y = rr.y + beta * (pp.y - om * vv.y);
  p[i] = pp;
  if (phat) phat[i] = pc_local<T>(pc, dinv, aval, diag, i, pp);
}

// s = r - alpha v; shat = M s
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_s(int nb, const typename VecT<T>::V2 *__restrict__ r,
                                                   const typename VecT<T>::V2 *__restrict__ v,
                                                   typename VecT<T>::V2 *__restrict__ s,
                                                   typename VecT<T>::V2 *__restrict__ shat, int pc,
                                                   const typename VecT<T>::V4 *__restrict__ dinv,
                                                   const typename VecT<T>::V4 *__restrict__ aval,
                                                   const int *__restrict__ diag, const KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  T al = (T)sc->alpha;
  typename VecT<T>::V2 rr = r[i], vv = v[i], ss;
  ss.x = rr.x - al * vv.x;
  ss.y = rr.y - al * vv.y;
  s[i] = ss;
  if (shat) shat[i] = pc_local<T>(pc, dinv, aval, diag, i, ss);
}

// x += alpha phat + omega shat; r = s - omega t; rho_new = (rhat, r); ||r||^2
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_x(int nb, typename VecT<T>::V2 *__restrict__ x,
                                                   const typename VecT<T>::V2 *__restrict__ phat,
                                                   const typename VecT<T>::V2 *__restrict__ shat,
                                                   const typename VecT<T>::V2 *__restrict__ s,
                                                   const typename VecT<T>::V2 *__restrict__ t,
                                                   typename VecT<T>::V2 *__restrict__ r,
                                                   const typename VecT<T>::V2 *__restrict__ rhat, RedSlot slot,
                                                   KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  T al = (T)sc->alpha, om = (T)sc->omega;
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
    typename VecT<T>::V2 xx = x[i], ph = phat[i], sh = shat[i], ss = s[i], tt = t[i], rh = rhat[i], rr;
    xx.x += al * ph.x + om * sh.x;
    xx.y += al * ph.y + om * sh.y;
    x[i] = xx;
    rr.x = ss.x - om * tt.x;
    rr.y = ss.y - om * tt.y;
    r[i] = rr;
    acc[0] += (double)rh.x * rr.x + (double)rh.y * rr.y;
    acc[1] += (double)rr.x * rr.x + (double)rr.y * rr.y;
  }
  __shared__ double lds[2][kBlock / 64];
  block_sum<2>(acc, lds);
  block_partials<2>(slot, acc);
}

// Second phase of k_bicg_x (one workgroup): rho, ||r||^2 and the stop test.
__global__ void __launch_bounds__(kBlock) k_bicg_x_finish(RedSlot slot, int nparts, KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  double tot[2];
  partials_sum<2>(slot, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->its += 1;
  sc->rho_old = sc->rho;
  sc->rho = tot[0];
  sc->rr = tot[1];
  if (!(tot[1] <= sc->tol2) && !std::isfinite(tot[1])) sc->done = 2;
  else if (tot[1] <= sc->tol2) sc->done = 1;
  else if (sc->omega == 0.0 || tot[0] == 0.0) sc->done = 2;
  else if (sc->its >= sc->maxit) sc->done = 3;
}

// rhat = r, p = v = 0, scalars from ||r0||^2 (rr_in)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_init(int nb, const typename VecT<T>::V2 *__restrict__ r,
                                                      typename VecT<T>::V2 *__restrict__ rhat,
                                                      typename VecT<T>::V2 *__restrict__ p,
                                                      typename VecT<T>::V2 *__restrict__ v,
                                                      const double *__restrict__ rr_in, double rtol, double abstol,
                                                      int maxit, KrylovScalars *__restrict__ sc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) {
    rhat[i] = r[i];
    typename VecT<T>::V2 z;
    z.x = 0; z.y = 0;
    p[i] = z;
    v[i] = z;
  }
  if (i == 0) {
    double rr = rr_in[0];
    sc->rho = rr;
    sc->rho_old = 1.0;
    sc->alpha = 1.0;
    sc->omega = 1.0;
    sc->rr = rr;
    sc->bb = rr_in[1];
    double tol = fmax(rtol * sqrt(rr), abstol);
    sc->tol2 = tol * tol;
    sc->its = 0;
    sc->maxit = maxit;
    sc->done = (rr <= sc->tol2) ? 1 : 0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_zero2(int nb, typename VecT<T>::V2 *__restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) {
    typename VecT<T>::V2 z;
    z.x = 0; z.y = 0;
    x[i] = z;
  }
}

// ===========================================================================
// Aggregation multigrid kernels
// ===========================================================================

// Galerkin values of level l+1: entry c = sum of level-l blocks glist[gptr[c]..].
// The thread owning a diagonal entry writes that block's 2x2 inverse.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_galerkin(int nnzc, const int *__restrict__ gptr,
                                                     const int *__restrict__ glist,
                                                     const typename VecT<T>::V4 *__restrict__ fine,
                                                     const int *__restrict__ crow, const int *__restrict__ ccol,
                                                     typename VecT<T>::V4 *__restrict__ coarse,
                                                     typename VecT<T>::V4 *__restrict__ cdinv) {
  using V4 = typename VecT<T>::V4;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nnzc) return;
  double a = 0, b = 0, cc = 0, d = 0;
  // four fine blocks' loads in flight at a time; the sums keep the list order
  int k = gptr[c];
  const int k1 = gptr[c + 1];
  for (; k + 4 <= k1; k += 4) {
    const int i0 = glist[k], i1 = glist[k + 1], i2 = glist[k + 2], i3 = glist[k + 3];
    const V4 f0 = fine[i0], f1 = fine[i1], f2 = fine[i2], f3 = fine[i3];
    a += f0.x; b += f0.y; cc += f0.z; d += f0.w;
    a += f1.x; b += f1.y; cc += f1.z; d += f1.w;
    a += f2.x; b += f2.y; cc += f2.z; d += f2.w;
    a += f3.x; b += f3.y; cc += f3.z; d += f3.w;
  }
  for (; k < k1; ++k) {
    const V4 f = fine[glist[k]];
    a += f.x; b += f.y; cc += f.z; d += f.w;
  }
  typename VecT<T>::V4 o;
  o.x = (T)a; o.y = (T)b; o.z = (T)cc; o.w = (T)d;
  coarse[c] = o;
  int r = crow[c];
  if (ccol[c] == r) {
    double det = a * d - b * cc, id = 1.0 / det;
    typename VecT<T>::V4 inv;
    inv.x = (T)(d * id); inv.y = (T)(-b * id); inv.z = (T)(-cc * id); inv.w = (T)(a * id);
    cdinv[r] = inv;
  }
}

// l1 block-Jacobi smoother (Baker, Falgout, Kolev & Yang, SIAM J. Sci. Comput.
// 33 (2011)): the inverse of D_i + diag(s0, s1), D_i the diagonal block of
// block row i and s0, s1 the sums of the absolute values of the off-diagonal
// blocks' entries in its two scalar rows.  Formed on every level after the
// Galerkin products, once per Picard iteration.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_l1_dinv(int nb, const int *__restrict__ ptr, const int *__restrict__ col,
                                                    const typename VecT<T>::V4 *__restrict__ val,
                                                    typename VecT<T>::V4 *__restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  double a = 0, b = 0, c = 0, d = 0, s0 = 0, s1 = 0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const typename VecT<T>::V4 v = val[k];
    if (col[k] == i) {
      a = v.x; b = v.y; c = v.z; d = v.w;
    } else {
      s0 += fabs((double)v.x) + fabs((double)v.y);
      s1 += fabs((double)v.z) + fabs((double)v.w);
    }
  }
  a += s0;
  d += s1;
  const double det = a * d - b * c, id = 1.0 / det;
  typename VecT<T>::V4 inv;
  inv.x = (T)(d * id); inv.y = (T)(-b * id); inv.z = (T)(-c * id); inv.w = (T)(a * id);
  dinv[i] = inv;
}

// z0 = omega D0^-1 r0  (pre-smoothing from a zero guess on the finest level)
// The V-cycle's per-row arithmetic is written with explicit fused multiply-adds
// and separately rounded products, so every kernel that inlines it (the
// per-level kernels, the one-workgroup coarse kernel, with or without the
// folded pre-smoothing) rounds identically whatever the compiler contracts.
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }

template <typename T>
__global__ void __launch_bounds__(kBlock) k_amg_presmooth(int nb, const typename VecT<T>::V2 *__restrict__ r,
                                                          const typename VecT<T>::V4 *__restrict__ dinv, T omega,
                                                          typename VecT<T>::V2 *__restrict__ z,
                                                          const KrylovScalars *__restrict__ sc) {
  if (sc && sc->done) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  typename VecT<T>::V2 q = r[i], o;
  typename VecT<T>::V4 d = dinv[i];
  o.x = mul_rn(omega, fma(d.x, q.x, mul_rn(d.y, q.y)));
  o.y = mul_rn(omega, fma(d.w, q.y, mul_rn(d.z, q.x)));
  z[i] = o;
}

// Per-row bodies of the V-cycle passes, shared by the grid-wide kernels (fine
// levels) and the one-workgroup kernel of the coarse levels.  kLpr lanes
// cooperate on one block row or aggregate; all lanes of a group take the same
// row, so the width-kLpr shuffles stay within converged lanes.


// q_i = r_i - (A zs z)_i; MT is the storage type of the matrix values (arithmetic
// in T).  zs scales z as it is read (the finest level's pre-smoothing from zero,
// z = omega r, when its diagonal blocks are the identity).
template <typename T, typename MT>
__device__ __forceinline__ void row_resid(int row, int lane, const int *__restrict__ ptr,
                                          const int *__restrict__ col, const typename VecT<MT>::V4 *__restrict__ val,
                                          const typename VecT<T>::V2 *__restrict__ r,
                                          const typename VecT<T>::V2 *__restrict__ z,
                                          typename VecT<T>::V2 *__restrict__ q, T zs = T(1)) {
  // lane 0's epilogue operand, loaded before the row's products (its
  // arithmetic is a single subtraction, so the order changes no result)
  typename VecT<T>::V2 ri{};
  if (lane == 0) ri = r[row];
  T y0 = 0, y1 = 0;
  auto add = [&](const typename VecT<MT>::V4 &a, const typename VecT<T>::V2 &zj) {
    const T z0 = mul_rn(zs, zj.x), z1 = mul_rn(zs, zj.y);
    y0 = fma((T)a.x, z0, fma((T)a.y, z1, y0));
    y1 = fma((T)a.z, z0, fma((T)a.w, z1, y1));
  };
  // two blocks' gathers in flight per lane; the FMAs keep the block order
  int k = ptr[row] + lane;
  const int k1 = ptr[row + 1];
  for (; k + kLpr < k1; k += 2 * kLpr) {
    const typename VecT<MT>::V4 a0 = val[k], a1 = val[k + kLpr];
    const int c0 = col[k], c1 = col[k + kLpr];
    const typename VecT<T>::V2 zj0 = z[c0], zj1 = z[c1];
    add(a0, zj0);
    add(a1, zj1);
  }
  if (k < k1) add(val[k], z[col[k]]);
#pragma unroll
  for (int off = kLpr / 2; off > 0; off >>= 1) {
    y0 += __shfl_down(y0, off, kLpr);
    y1 += __shfl_down(y1, off, kLpr);
  }
  if (lane == 0) {
    typename VecT<T>::V2 o;
    o.x = ri.x - y0;
    o.y = ri.y - y1;
    q[row] = o;
  }
}

// rc_I = sum of q over the members of aggregate I, and zc_I = omega Dc^-1 rc_I
// (the next level's pre-smoothing from zero) when zc is given.
template <typename T>
__device__ __forceinline__ void agg_restrict(int I, int lane, const int *__restrict__ mem_ptr,
                                             const int *__restrict__ mem_ind,
                                             const typename VecT<T>::V2 *__restrict__ q,
                                             typename VecT<T>::V2 *__restrict__ rc,
                                             const typename VecT<T>::V4 *__restrict__ cdinv, T omega,
                                             typename VecT<T>::V2 *__restrict__ zc) {
  // the coarse diagonal-block inverse is loaded before the member sums so its
  // latency overlaps them (the epilogue's rounding is explicit, so moving the
  // load changes no result)
  typename VecT<T>::V4 d{};
  if (zc && lane == 0) d = cdinv[I];
  double s0 = 0, s1 = 0;
  for (int m = mem_ptr[I] + lane; m < mem_ptr[I + 1]; m += kLpr) {
    typename VecT<T>::V2 qi = q[mem_ind[m]];
    s0 += qi.x;
    s1 += qi.y;
  }
#pragma unroll
  for (int off = kLpr / 2; off > 0; off >>= 1) {
    s0 += __shfl_down(s0, off, kLpr);
    s1 += __shfl_down(s1, off, kLpr);
  }
  if (lane == 0) {
    typename VecT<T>::V2 o;
    o.x = (T)s0; o.y = (T)s1;
    rc[I] = o;
    if (zc) {
      typename VecT<T>::V2 zz;
      zz.x = mul_rn(omega, (T)fma((double)d.x, s0, mul_rn((double)d.y, s1)));
      zz.y = mul_rn(omega, (T)fma((double)d.w, s1, mul_rn((double)d.z, s0)));
      zc[I] = zz;
    }
  }
}

// zout_i = zn_i + omega D^-1 (r - A zn)_i with zn = z + scale * P ec: the
// prolongation fused into the post-smoothing sweep.
template <typename T, typename MT>
__device__ __forceinline__ void row_prolong_smooth(int i, int lane, const int *__restrict__ ptr,
                                                   const int *__restrict__ col,
                                                   const typename VecT<MT>::V4 *__restrict__ val,
                                                   const typename VecT<T>::V4 *__restrict__ dinv,
                                                   const int *__restrict__ agg,
                                                   const typename VecT<T>::V2 *__restrict__ r,
                                                   const typename VecT<T>::V2 *__restrict__ z,
                                                   const typename VecT<T>::V2 *__restrict__ ec, T scale, T omega,
                                                   typename VecT<T>::V2 *__restrict__ zout, T zs = T(1)) {
  // lane 0's epilogue operands, loaded before the row's products (the
  // epilogue's rounding is explicit, so the order changes no result)
  typename VecT<T>::V2 ri{}, zi{}, ei{};
  typename VecT<T>::V4 di{};
  if (lane == 0) {
    ri = r[i];
    zi = z[i];
    ei = ec[agg[i]];
    if (dinv) di = dinv[i];
  }
  T y0 = 0, y1 = 0;
  auto add = [&](const typename VecT<MT>::V4 &a, const typename VecT<T>::V2 &zj, const typename VecT<T>::V2 &ej) {
    const T z0 = fma(scale, ej.x, mul_rn(zs, zj.x)), z1 = fma(scale, ej.y, mul_rn(zs, zj.y));
    y0 = fma((T)a.x, z0, fma((T)a.y, z1, y0));
    y1 = fma((T)a.z, z0, fma((T)a.w, z1, y1));
  };
  // two blocks' gathers (column, aggregate, coarse correction) in flight per
  // lane; the FMAs keep the block order
  int k = ptr[i] + lane;
  const int k1 = ptr[i + 1];
  for (; k + kLpr < k1; k += 2 * kLpr) {
    const int j0 = col[k], j1 = col[k + kLpr];
    const typename VecT<MT>::V4 a0 = val[k], a1 = val[k + kLpr];
    const int g0 = agg[j0], g1 = agg[j1];
    const typename VecT<T>::V2 zj0 = z[j0], zj1 = z[j1], ej0 = ec[g0], ej1 = ec[g1];
    add(a0, zj0, ej0);
    add(a1, zj1, ej1);
  }
  if (k < k1) {
    const int j = col[k];
    add(val[k], z[j], ec[agg[j]]);
  }
#pragma unroll
  for (int off = kLpr / 2; off > 0; off >>= 1) {
    y0 += __shfl_down(y0, off, kLpr);
    y1 += __shfl_down(y1, off, kLpr);
  }
  if (lane == 0) {
    T q0 = ri.x - y0, q1 = ri.y - y1;
    typename VecT<T>::V2 o;
    if (dinv) {  // null: identity diagonal blocks (1 * a + 0 * b == a, so both agree bit for bit)
      const typename VecT<T>::V4 d = di;
      const T t0 = fma(d.x, q0, mul_rn(d.y, q1)), t1 = fma(d.w, q1, mul_rn(d.z, q0));
      q0 = t0;
      q1 = t1;
    }
    o.x = fma(omega, q0, fma(scale, ei.x, mul_rn(zs, zi.x)));
    o.y = fma(omega, q1, fma(scale, ei.y, mul_rn(zs, zi.y)));
    zout[i] = o;
  }
}

template <typename T, typename MT = T>
__global__ void __launch_bounds__(kBlock) k_amg_resid(int nb, const int *__restrict__ ptr, const int *__restrict__ col,
                                                      const typename VecT<MT>::V4 *__restrict__ val,
                                                      const typename VecT<T>::V2 *__restrict__ r,
                                                      const typename VecT<T>::V2 *__restrict__ z,
                                                      typename VecT<T>::V2 *__restrict__ q,
                                                      const KrylovScalars *__restrict__ sc, T zs) {
  if (sc && sc->done) return;
  const int row = (xcd_block() * blockDim.x + threadIdx.x) / kLpr;
  if (row >= nb) return;
  row_resid<T, MT>(row, threadIdx.x % kLpr, ptr, col, val, r, z, q, zs);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_amg_restrict(int nc, const int *__restrict__ mem_ptr,
                                                         const int *__restrict__ mem_ind,
                                                         const typename VecT<T>::V2 *__restrict__ q,
                                                         typename VecT<T>::V2 *__restrict__ rc,
                                                         const typename VecT<T>::V4 *__restrict__ cdinv, T omega,
                                                         typename VecT<T>::V2 *__restrict__ zc,
                                                         const KrylovScalars *__restrict__ sc) {
  if (sc && sc->done) return;
  const int I = (blockIdx.x * blockDim.x + threadIdx.x) / kLpr;
  if (I >= nc) return;
  agg_restrict<T>(I, threadIdx.x % kLpr, mem_ptr, mem_ind, q, rc, cdinv, omega, zc);
}

template <typename T, typename MT = T>
__global__ void __launch_bounds__(kBlock) k_amg_prolong_smooth(
    int nb, const int *__restrict__ ptr, const int *__restrict__ col, const typename VecT<MT>::V4 *__restrict__ val,
    const typename VecT<T>::V4 *__restrict__ dinv, const int *__restrict__ agg,
    const typename VecT<T>::V2 *__restrict__ r, const typename VecT<T>::V2 *__restrict__ z,
    const typename VecT<T>::V2 *__restrict__ ec, T scale, T omega, typename VecT<T>::V2 *__restrict__ zout,
    const KrylovScalars *__restrict__ sc, T zs) {
  if (sc && sc->done) return;
  const int i = (xcd_block() * blockDim.x + threadIdx.x) / kLpr;
  if (i >= nb) return;
  row_prolong_smooth<T, MT>(i, threadIdx.x % kLpr, ptr, col, val, dinv, agg, r, z, ec, scale, omega, zout, zs);
}

// Coarsest level: an exact solve with the dense inverse of the coarsest
// Galerkin matrix (nb <= kCoarseMax blocks), formed once per Picard iteration
// by Gauss-Jordan elimination with partial pivoting in LDS.
constexpr int kCoarseMax = 24;
constexpr int kCoarseN = 2 * kCoarseMax;
constexpr int kCoarseThreads = 1024;

template <typename T>
__global__ void __launch_bounds__(256) k_coarse_invert(int nb, const int *__restrict__ ptr, const int *__restrict__ col,
                                                       const typename VecT<T>::V4 *__restrict__ val,
                                                       double *__restrict__ inv_cm) {
  __shared__ double a[kCoarseN][2 * kCoarseN + 1];
  __shared__ int piv;
  const int n = 2 * nb, tid = threadIdx.x;
  for (int i = tid; i < kCoarseN * (2 * kCoarseN + 1); i += blockDim.x) (&a[0][0])[i] = 0.0;
  __syncthreads();
  for (int i = tid; i < nb; i += blockDim.x) {
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
      int j = col[k];
      typename VecT<T>::V4 v = val[k];
      a[2 * i][2 * j] = v.x; a[2 * i][2 * j + 1] = v.y;
      a[2 * i + 1][2 * j] = v.z; a[2 * i + 1][2 * j + 1] = v.w;
    }
    a[2 * i][n + 2 * i] = 1.0;
    a[2 * i + 1][n + 2 * i + 1] = 1.0;
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      int p = k;
      double best = fabs(a[k][k]);
      for (int i = k + 1; i < n; ++i)
        if (fabs(a[i][k]) > best) { best = fabs(a[i][k]); p = i; }
      piv = p;
    }
    __syncthreads();
    int p = piv;
    if (p != k)
      for (int j = tid; j < 2 * n; j += blockDim.x) { double t = a[k][j]; a[k][j] = a[p][j]; a[p][j] = t; }
    __syncthreads();
    double d = a[k][k];
    double rd = d != 0.0 ? 1.0 / d : 0.0;
    __syncthreads();
    for (int j = tid; j < 2 * n; j += blockDim.x) a[k][j] *= rd;
    __syncthreads();
    for (int e = tid; e < n * 2 * n; e += blockDim.x) {
      int i = e / (2 * n), j = e % (2 * n);
      if (i == k || j == k) continue;
      a[i][j] -= a[i][k] * a[k][j];
    }
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x)
      if (i != k) a[i][k] = 0.0;
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += blockDim.x) {
    int i = e % n, j = e / n;
    inv_cm[(size_t)j * n + i] = a[i][n + j];  // column-major
  }
}

// z = Ainv r on the coarsest level
template <typename T>
__global__ void __launch_bounds__(64) k_coarse_apply(int nb, const double *__restrict__ inv_cm,
                                                     const typename VecT<T>::V2 *__restrict__ r,
                                                     typename VecT<T>::V2 *__restrict__ z,
                                                     const KrylovScalars *__restrict__ sc) {
  if (sc && sc->done) return;
  __shared__ double rs[kCoarseN];
  const int n = 2 * nb, i = threadIdx.x;
  if (i < nb) { rs[2 * i] = r[i].x; rs[2 * i + 1] = r[i].y; }
  __syncthreads();
  if (i < n) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += inv_cm[(size_t)j * n + i] * rs[j];
    T *zz = reinterpret_cast<T *>(z);
    zz[i] = (T)s;
  }
}

// The coarse end of the V-cycle in one workgroup: levels first..last of the
// hierarchy are small (a few thousand block rows at most), where a launch per
// pass costs more than the pass.  Same arithmetic as the grid-wide kernels,
// level after level with workgroup barriers between the passes; each level's
// final correction is left in its z vector (the finer level's prolongation
// reads it there).
constexpr int kTailMaxLevels = 8;
constexpr int kTailThreads = 1024;
constexpr int kTailMaxRows = 4096;
struct AmgTailLevel {
  int nb;
  const int *ptr, *col, *agg, *mem_ptr, *mem_ind;
  const void *val, *dinv;
  void *r, *z, *zt;
  // byte offsets of r, z, zt in the kernel's dynamic LDS, or -1 for the
  // device-memory vectors above (the top level's r and z are the kernel's
  // input and output and always live in device memory)
  int off_r, off_z, off_zt;
};
constexpr int kTailLdsMax = 48 * 1024;  // dynamic LDS of the tail kernel, at most
struct AmgTailArgs {
  int n;  // levels in the tail; lv[n-1] is the coarsest
  double omega, scale;
  const double *cinv;  // dense inverse of the coarsest level, column-major
  AmgTailLevel lv[kTailMaxLevels];
};

// A tail level's vector: in the kernel's LDS when it has an offset there,
// else in device memory.  The work-group barriers between the phases order
// both kinds of access the same way.
__device__ __forceinline__ void *tail_vec(unsigned char *lds, int off, void *dev) {
  return off >= 0 ? (void *)(lds + off) : dev;
}

template <typename T>
__global__ void __launch_bounds__(kTailThreads) k_amg_tail(AmgTailArgs a, const KrylovScalars *__restrict__ sc) {
  using V2 = typename VecT<T>::V2;
  using V4 = typename VecT<T>::V4;
  extern __shared__ __align__(16) unsigned char tail_lds[];
  if (sc && sc->done) return;
  const int lane = threadIdx.x % kLpr, grp = threadIdx.x / kLpr, ngrp = blockDim.x / kLpr;
  const T om = (T)a.omega, cs = (T)a.scale;
  for (int l = 0; l + 1 < a.n; ++l) {
    const AmgTailLevel &L = a.lv[l];
    const AmgTailLevel &C = a.lv[l + 1];
    const V2 *Lr = (const V2 *)tail_vec(tail_lds, L.off_r, L.r);
    const V2 *Lz = (const V2 *)tail_vec(tail_lds, L.off_z, L.z);
    V2 *Lzt = (V2 *)tail_vec(tail_lds, L.off_zt, L.zt);
    for (int row = grp; row < L.nb; row += ngrp)
      row_resid<T, T>(row, lane, L.ptr, L.col, (const V4 *)L.val, Lr, Lz, Lzt);
    __syncthreads();
    V2 *zc = l + 2 < a.n ? (V2 *)tail_vec(tail_lds, C.off_z, C.z) : nullptr;
    V2 *Cr = (V2 *)tail_vec(tail_lds, C.off_r, C.r);
    for (int I = grp; I < C.nb; I += ngrp)
      agg_restrict<T>(I, lane, L.mem_ptr, L.mem_ind, Lzt, Cr, (const V4 *)C.dinv, om, zc);
    __syncthreads();
  }
  {
    const AmgTailLevel &C = a.lv[a.n - 1];
    const int n = 2 * C.nb, i = threadIdx.x;
    if (i < n) {
      const T *rr = (const T *)tail_vec(tail_lds, C.off_r, C.r);
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += a.cinv[(size_t)j * n + i] * (double)rr[j];
      ((T *)tail_vec(tail_lds, C.off_z, C.z))[i] = (T)s;
    }
    __syncthreads();
  }
  for (int l = a.n - 2; l >= 0; --l) {
    const AmgTailLevel &L = a.lv[l];
    const AmgTailLevel &C = a.lv[l + 1];
    const V2 *Lr = (const V2 *)tail_vec(tail_lds, L.off_r, L.r);
    V2 *Lz = (V2 *)tail_vec(tail_lds, L.off_z, L.z);
    V2 *Lzt = (V2 *)tail_vec(tail_lds, L.off_zt, L.zt);
    const V2 *Cz = (const V2 *)tail_vec(tail_lds, C.off_z, C.z);
    for (int row = grp; row < L.nb; row += ngrp)
      row_prolong_smooth<T, T>(row, lane, L.ptr, L.col, (const V4 *)L.val, (const V4 *)L.dinv, L.agg, Lr, Lz, Cz,
                               cs, om, Lzt);
    __syncthreads();
    for (int i = threadIdx.x; i < L.nb; i += blockDim.x) Lz[i] = Lzt[i];
    __syncthreads();
  }
}

// FP32 <-> FP64 conversions and FP64 axpy used by the refinement loop
__global__ void __launch_bounds__(kBlock) k_f64_to_f32(int n, const double *__restrict__ a, float *__restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ===========================================================================
// Host side
// ===========================================================================

static std::vector<int> greedy_aggregate(int n, const std::vector<int> &aptr, const std::vector<int> &aind,
                                         int *n_agg) {
  std::vector<int> agg(n, -1);
  int na = 0;
  // pass 1: a root whose neighbours are all free takes them
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nb = true;
    for (int k = aptr[i]; k < aptr[i + 1]; ++k)
      if (agg[aind[k]] >= 0) { free_nb = false; break; }
    if (!free_nb) continue;
    agg[i] = na;
    for (int k = aptr[i]; k < aptr[i + 1]; ++k) agg[aind[k]] = na;
    ++na;
  }
  // pass 2: leftovers join a neighbouring aggregate (or form a singleton)
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    for (int k = aptr[i]; k < aptr[i + 1]; ++k)
      if (agg[aind[k]] >= 0) { agg[i] = agg[aind[k]]; break; }
    if (agg[i] < 0) agg[i] = na++;
  }
  *n_agg = na;
  return agg;
}

// Graph of aggregates: two aggregates are adjacent if any of their members are.
static void aggregate_graph(int n, const std::vector<int> &aptr, const std::vector<int> &aind,
                            const std::vector<int> &agg, int na, std::vector<int> &nptr, std::vector<int> &nind) {
  std::vector<std::vector<int>> nbrs(na);
  for (int i = 0; i < n; ++i)
    for (int k = aptr[i]; k < aptr[i + 1]; ++k) {
      int a = agg[i], b = agg[aind[k]];
      if (a != b) nbrs[a].push_back(b);
    }
  nptr.assign(na + 1, 0);
  nind.clear();
  for (int I = 0; I < na; ++I) {
    auto &v = nbrs[I];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    nind.insert(nind.end(), v.begin(), v.end());
    nptr[I + 1] = (int)nind.size();
  }
}

template <typename U>
static void upload(DBuf &d, const std::vector<U> &h) {
  d.alloc(std::max<size_t>(h.size(), 1) * sizeof(U));
  if (!h.empty()) UFE_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
}

std::unique_ptr<AmgHierarchy> build_amg_hierarchy(int nT, const int *TriC, const int *ptr0, const int *col0,
                                                  int coarse_target, int max_levels, int passes) {
  auto H = std::make_unique<AmgHierarchy>();
  coarse_target = std::min(coarse_target, kCoarseMax);
  // level-0 aggregation graph: triangle neighbours across edges
  std::vector<int> aptr(nT + 1, 0), aind;
  aind.reserve(3 * (size_t)nT);
  for (int t = 0; t < nT; ++t) {
    for (int q = 0; q < 3; ++q)
      if (TriC[3 * t + q] >= 0) aind.push_back(TriC[3 * t + q]);
    aptr[t + 1] = (int)aind.size();
  }
  std::vector<int> mptr(ptr0, ptr0 + nT + 1), mcol(col0, col0 + ptr0[nT]);
  std::vector<int> gptr, glist;  // gather lists building the current level
  int n = nT;
  for (int lev = 0; lev < max_levels; ++lev) {
    auto L = std::make_unique<AmgLevelPattern>();
    L->nb = n;
    L->nnzb = mptr[n];
    if (lev > 0) {
      std::vector<int> crow(mptr[n]);
      for (int i = 0; i < n; ++i)
        for (int k = mptr[i]; k < mptr[i + 1]; ++k) crow[k] = i;
      upload(L->ptr, mptr);
      upload(L->col, mcol);
      upload(L->crow, crow);
      upload(L->gptr, gptr);
      upload(L->glist, glist);
    }
    bool last = (n <= coarse_target) || (lev == max_levels - 1);
    int na = 0;
    std::vector<int> agg;
    if (!last) {
      agg = greedy_aggregate(n, aptr, aind, &na);
      // further passes aggregate the aggregates (coarsening ratio ~4 per pass)
      for (int pass = 1; pass < passes && na > coarse_target; ++pass) {
        std::vector<int> gp, gi;
        aggregate_graph(n, aptr, aind, agg, na, gp, gi);
        int na2 = 0;
        std::vector<int> agg2 = greedy_aggregate(na, gp, gi, &na2);
        for (int i = 0; i < n; ++i) agg[i] = agg2[agg[i]];
        na = na2;
      }
      if (na >= n || na < 1) last = true;
    }
    if (last) {
      if (n > kCoarseMax) return nullptr;  // coarsest level must fit the single-workgroup solve
      H->levels.push_back(std::move(L));
      break;
    }
    // members of each aggregate
    std::vector<int> cnt(na + 1, 0);
    for (int i = 0; i < n; ++i) ++cnt[agg[i] + 1];
    std::partial_sum(cnt.begin(), cnt.end(), cnt.begin());
    std::vector<int> mem(n), fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i) mem[fill[agg[i]]++] = i;
    upload(L->agg, agg);
    upload(L->mem_ptr, cnt);
    upload(L->mem_ind, mem);
    // next-level pattern and gather lists
    std::vector<std::vector<std::pair<int, int>>> rows(na);  // (coarse col, fine entry)
    for (int i = 0; i < n; ++i)
      for (int k = mptr[i]; k < mptr[i + 1]; ++k) rows[agg[i]].push_back({agg[mcol[k]], k});
    std::vector<int> cptr(na + 1, 0), ccol;
    gptr.assign(1, 0);
    glist.clear();
    for (int I = 0; I < na; ++I) {
      auto &rw = rows[I];
      std::sort(rw.begin(), rw.end());
      for (size_t q = 0; q < rw.size();) {
        size_t q2 = q;
        while (q2 < rw.size() && rw[q2].first == rw[q].first) glist.push_back(rw[q2++].second);
        ccol.push_back(rw[q].first);
        gptr.push_back((int)glist.size());
        q = q2;
      }
      cptr[I + 1] = (int)ccol.size();
      std::vector<std::pair<int, int>>().swap(rw);
    }
    // next-level aggregation graph
    std::vector<int> nptr, nind;
    aggregate_graph(n, aptr, aind, agg, na, nptr, nind);
    H->levels.push_back(std::move(L));
    aptr.swap(nptr);
    aind.swap(nind);
    mptr.swap(cptr);
    mcol.swap(ccol);
    n = na;
  }
  return H;
}

Krylov::Krylov(int nb, hipStream_t s) : nb_(nb), s_(s) {
  if (const char *e = std::getenv("UFE_VERBOSE")) verbose_ = std::atoi(e);
  for (auto &v : vec64_) v.alloc(sizeof(double) * 2 * (size_t)nb);
  for (auto &v : vec32_) v.alloc(sizeof(float) * 2 * (size_t)nb);
  scal_.alloc(sizeof(KrylovScalars));
  red_part_.alloc(sizeof(double) * kMaxPartials * 4 * 4);
  red_res_.alloc(sizeof(double) * 4 * 4);
  UFE_HIP(hipMemset(scal_.p, 0, sizeof(KrylovScalars)));
  UFE_HIP(hipHostMalloc((void **)&h_scal_, sizeof(KrylovScalars), hipHostMallocDefault));
}

Krylov::~Krylov() {
  for (GraphCache &g : graphs_) g.clear();
  if (h_scal_) (void)hipHostFree(h_scal_);
  for (hipEvent_t e : tev_) (void)hipEventDestroy(e);
}

double Krylov::time_spmv64(const BsrT<double> &A, const double *x, double *y, int reps, double *bytes) {
  hipEvent_t e0, e1;
  UFE_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
  UFE_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
  RedSlot sl = slot(3);
  spmv_launch<double, OP_PLAIN>(A, x, y, nullptr, nullptr, sl, nullptr, nullptr, s_);  // warm caches/clocks
  UFE_HIP(hipEventRecord(e0, s_));
  for (int i = 0; i < reps; ++i) spmv_launch<double, OP_PLAIN>(A, x, y, nullptr, nullptr, sl, nullptr, nullptr, s_);
  UFE_HIP(hipEventRecord(e1, s_));
  UFE_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  UFE_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (bytes)
    *bytes = (double)A.nnzb * (sizeof(double4) + sizeof(int)) + (double)(A.nb + 1) * sizeof(int) +
             2.0 * A.nb * sizeof(double2);
  return reps > 0 ? ms / reps : 0.0;
}

void Krylov::set_spmv_timing(bool on) {
  if (on)
    for (int i = 0; i < 2; ++i) {
      tn_[i] = 0;
      tms_[i] = 0.0;
      tbytes_[i] = 0.0;
    }
  time_spmv_ = on;
}

RedSlot Krylov::slot(int i) const {
  RedSlot r;
  r.partials = red_part_.as<double>() + (size_t)i * kMaxPartials * 4;
  r.result = red_res_.as<double>() + 4 * i;
  return r;
}

double *Krylov::tmp64(int i) { return vec64_[i].as<double>(); }
float *Krylov::tmp32(int i) { return vec32_[i].as<float>(); }

void Krylov::residual64(const BsrT<double> &A, const double *x, const double *b, double *r, double *rr,
                        double *bb, double vmax, int *n_clamped, const double4 *dblk, int write_unscaled) {
  RedSlot sl = slot(3);
  spmv_launch<double, OP_RESID>(A, x, r, b, nullptr, sl, sl.result, nullptr, s_, vmax, dblk, write_unscaled);
  double h[3];
  UFE_HIP(hipMemcpyAsync(h, sl.result, sizeof(h), hipMemcpyDeviceToHost, s_));
  UFE_HIP(hipStreamSynchronize(s_));
  *rr = h[0];
  *bb = h[1];
  if (n_clamped) *n_clamped = (int)h[2];
}

template <> AmgValues<double> &Krylov::amg_vals<double>() { return amg64_; }
template <> AmgValues<float> &Krylov::amg_vals<float>() { return amg32_; }

template <typename T>
void Krylov::amg_setup(const BsrT<T> &A0) {
  using V2 = typename VecT<T>::V2;
  using V4 = typename VecT<T>::V4;
  if (!amg_) return;
  AmgValues<T> &av = amg_vals<T>();
  std::vector<BsrT<T>> &lev = std::is_same<T, double>::value ? *reinterpret_cast<std::vector<BsrT<T>> *>(&lev64_)
                                                             : *reinterpret_cast<std::vector<BsrT<T>> *>(&lev32_);
  int L = amg_->n_levels();
  if ((int)av.val.size() != L) {
    av.val.clear(); av.dinv.clear(); av.r.clear(); av.z.clear(); av.zt.clear();
    for (int l = 0; l < L; ++l) {
      const AmgLevelPattern &P = *amg_->levels[l];
      av.val.push_back(std::make_unique<DBuf>(l > 0 ? sizeof(V4) * (size_t)P.nnzb : 0));
      av.dinv.push_back(std::make_unique<DBuf>(l > 0 ? sizeof(V4) * (size_t)P.nb : 0));
      av.r.push_back(std::make_unique<DBuf>(sizeof(V2) * (size_t)P.nb));
      av.z.push_back(std::make_unique<DBuf>(sizeof(V2) * (size_t)P.nb));
      av.zt.push_back(std::make_unique<DBuf>(sizeof(V2) * (size_t)P.nb));
    }
  }
  lev.assign(L, BsrT<T>());
  lev[0] = A0;
  if (amg_l1_) {
    if (av.dinv[0]->bytes < sizeof(V4) * (size_t)A0.nb) av.dinv[0]->alloc(sizeof(V4) * (size_t)A0.nb);
    hipLaunchKernelGGL(k_l1_dinv<T>, dim3(blocks_for(A0.nb)), dim3(kBlock), 0, s_, A0.nb, A0.ptr, A0.col, A0.val,
                       av.dinv[0]->template as<V4>());
    UFE_LAUNCH_CHECK();
    lev[0].dinv = av.dinv[0]->template as<V4>();
  }
  for (int l = 1; l < L; ++l) {
    const AmgLevelPattern &P = *amg_->levels[l];
    BsrT<T> B;
    B.nb = P.nb;
    B.nnzb = P.nnzb;
    B.ptr = P.ptr.as<int>();
    B.col = P.col.as<int>();
    B.val = av.val[l]->template as<V4>();
    B.dinv = av.dinv[l]->template as<V4>();
    hipLaunchKernelGGL(k_galerkin<T>, dim3(blocks_for(P.nnzb)), dim3(kBlock), 0, s_, P.nnzb, P.gptr.as<int>(),
                       P.glist.as<int>(), lev[l - 1].val, P.crow.as<int>(), P.col.as<int>(),
                       av.val[l]->template as<V4>(), av.dinv[l]->template as<V4>());
    UFE_LAUNCH_CHECK();
    if (amg_l1_) {
      hipLaunchKernelGGL(k_l1_dinv<T>, dim3(blocks_for(B.nb)), dim3(kBlock), 0, s_, B.nb, B.ptr, B.col, B.val,
                         av.dinv[l]->template as<V4>());
      UFE_LAUNCH_CHECK();
    }
    lev[l] = B;
  }
  const BsrT<T> &Cc = lev[L - 1];
  if (!coarse_inv_.p) coarse_inv_.alloc(sizeof(double) * kCoarseN * kCoarseN);
  hipLaunchKernelGGL(k_coarse_invert<T>, dim3(1), dim3(256), 0, s_, Cc.nb, Cc.ptr, Cc.col, Cc.val,
                     coarse_inv_.as<double>());
  UFE_LAUNCH_CHECK();
}

void Krylov::amg_setup64(const BsrT<double> &A0) { amg_setup<double>(A0); }
void Krylov::amg_setup32(const BsrT<float> &A0) { amg_setup<float>(A0); }

template <typename T>
void Krylov::precond_apply(const BsrT<T> &A, const T *r, T *z, hipEvent_t ev0, hipEvent_t ev1) {
  using V2 = typename VecT<T>::V2;
  AmgValues<T> &av = amg_vals<T>();
  std::vector<BsrT<T>> &lev = std::is_same<T, double>::value ? *reinterpret_cast<std::vector<BsrT<T>> *>(&lev64_)
                                                             : *reinterpret_cast<std::vector<BsrT<T>> *>(&lev32_);
  KrylovScalars *sc = scal_.as<KrylovScalars>();
  int L = (int)lev.size();
  T om = (T)amg_omega_, cs = (T)amg_coarse_scale_;
  // level vectors: r_l (l>=1 in av.r), z_l in av.z; level 0 input r, output z
  const V2 *rl = reinterpret_cast<const V2 *>(r);
  V2 *z0 = av.z[0]->template as<V2>();
  if (L == 1) {
    hipLaunchKernelGGL(k_coarse_apply<T>, dim3(1), dim3(64), 0, s_, lev[0].nb, coarse_inv_.as<double>(), rl,
                       reinterpret_cast<V2 *>(z), sc);
    UFE_LAUNCH_CHECK();
    return;
  }
  // finest-level pre-smoothing from zero, z0 = omega D^-1 r: with identity
  // diagonal blocks (the scaled system) the level-0 passes read omega * r
  // directly instead (bit-identical: 1 * x + 0 * y == x)
  const bool fold = unit_diag0_ && !amg_l1_;
  const V2 *zl0 = fold ? rl : z0;
  const T zs0 = fold ? om : T(1);
  if (!fold) {
    hipLaunchKernelGGL(k_amg_presmooth<T>, dim3(blocks_for(lev[0].nb)), dim3(kBlock), 0, s_, lev[0].nb, rl,
                       lev[0].dinv, om, z0, sc);
    UFE_LAUNCH_CHECK();
  }
  // levels t..L-1 run in one workgroup (k_amg_tail); t = L-1: only the coarsest
  int t = L - 1;
  for (int l = 1; l < L - 1; ++l)
    if (lev[l].nb <= std::min(amg_tail_rows_, kTailMaxRows) && L - l <= kTailMaxLevels) { t = l; break; }
  for (int l = 0; l < t; ++l) {
    const AmgLevelPattern &P = *amg_->levels[l];
    const V2 *rin = l == 0 ? rl : av.r[l]->template as<V2>();
    bool next_coarsest = (l + 1 == L - 1);
    int nc = lev[l + 1].nb;
    const dim3 gl(xcd_grid(blocks_for((long)lev[l].nb * kLpr)));
    const V2 *zl = l == 0 ? zl0 : av.z[l]->template as<V2>();
    const T zs = l == 0 ? zs0 : T(1);
    if (l == 0 && fine32_ && std::is_same<T, double>::value)
      hipLaunchKernelGGL((k_amg_resid<T, float>), gl, dim3(kBlock), 0, s_, lev[l].nb, lev[l].ptr, lev[l].col,
                         fine32_, rin, zl, av.zt[l]->template as<V2>(), sc, zs);
    else
      hipLaunchKernelGGL((k_amg_resid<T, T>), gl, dim3(kBlock), 0, s_, lev[l].nb, lev[l].ptr, lev[l].col,
                         lev[l].val, rin, zl, av.zt[l]->template as<V2>(), sc, zs);
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_amg_restrict<T>, dim3(blocks_for((long)nc * kLpr)), dim3(kBlock), 0, s_, nc,
                       P.mem_ptr.as<int>(),
                       P.mem_ind.as<int>(), av.zt[l]->template as<V2>(), av.r[l + 1]->template as<V2>(),
                       lev[l + 1].dinv, om, next_coarsest ? (V2 *)nullptr : av.z[l + 1]->template as<V2>(), sc);
    UFE_LAUNCH_CHECK();
  }
  if (t == L - 1) {
    int c = L - 1;
    hipLaunchKernelGGL(k_coarse_apply<T>, dim3(1), dim3(64), 0, s_, lev[c].nb, coarse_inv_.as<double>(),
                       av.r[c]->template as<V2>(), av.z[c]->template as<V2>(), sc);
    UFE_LAUNCH_CHECK();
  } else {
    AmgTailArgs ta{};
    ta.n = L - t;
    ta.omega = amg_omega_;
    ta.scale = amg_coarse_scale_;
    ta.cinv = coarse_inv_.as<double>();
    for (int l = t; l < L; ++l) {
      const AmgLevelPattern &P = *amg_->levels[l];
      AmgTailLevel &d = ta.lv[l - t];
      d.nb = lev[l].nb;
      d.ptr = lev[l].ptr;
      d.col = lev[l].col;
      d.val = lev[l].val;
      d.dinv = lev[l].dinv;
      d.agg = l < L - 1 ? P.agg.as<int>() : nullptr;
      d.mem_ptr = l < L - 1 ? P.mem_ptr.as<int>() : nullptr;
      d.mem_ind = l < L - 1 ? P.mem_ind.as<int>() : nullptr;
      d.r = av.r[l]->p;
      d.z = av.z[l]->p;
      d.zt = av.zt[l]->p;
      d.off_r = d.off_z = d.off_zt = -1;
    }
    // the vectors that stay inside the kernel go to LDS: every level's zt and
    // the r and z of the levels below the top one
    int lds = 0;
    auto take = [&](int nb) {
      const int o = lds;
      lds += ((int)(nb * sizeof(V2)) + 15) & ~15;
      return o;
    };
    for (int l = 0; l < ta.n; ++l) {
      AmgTailLevel &d = ta.lv[l];
      if (l + 1 < ta.n) d.off_zt = take(d.nb);
      if (l > 0) { d.off_r = take(d.nb); d.off_z = take(d.nb); }
    }
    if (lds > kTailLdsMax) {
      for (int l = 0; l < ta.n; ++l) ta.lv[l].off_r = ta.lv[l].off_z = ta.lv[l].off_zt = -1;
      lds = 0;
    }
    hipLaunchKernelGGL(k_amg_tail<T>, dim3(1), dim3(kTailThreads), lds, s_, ta, sc);
    UFE_LAUNCH_CHECK();
  }
  for (int l = t - 1; l >= 0; --l) {
    const AmgLevelPattern &P = *amg_->levels[l];
    const V2 *rin = l == 0 ? rl : av.r[l]->template as<V2>();
    V2 *zout = l == 0 ? reinterpret_cast<V2 *>(z) : av.zt[l]->template as<V2>();
    const dim3 gl(xcd_grid(blocks_for((long)lev[l].nb * kLpr)));
    const V2 *zl = l == 0 ? zl0 : av.z[l]->template as<V2>();
    const T zs = l == 0 ? zs0 : T(1);
    const typename VecT<T>::V4 *dl = (l == 0 && fold) ? nullptr : lev[l].dinv;
    if (l == 0 && ev0) UFE_HIP(hipEventRecord(ev0, s_));
    if (l == 0 && fine32_ && std::is_same<T, double>::value)
      hipLaunchKernelGGL((k_amg_prolong_smooth<T, float>), gl, dim3(kBlock), 0, s_, lev[l].nb, lev[l].ptr,
                         lev[l].col, fine32_, dl, P.agg.as<int>(), rin, zl,
                         av.z[l + 1]->template as<V2>(), cs, om, zout, sc, zs);
    else
      hipLaunchKernelGGL((k_amg_prolong_smooth<T, T>), gl, dim3(kBlock), 0, s_, lev[l].nb, lev[l].ptr,
                         lev[l].col, lev[l].val, dl, P.agg.as<int>(), rin, zl,
                         av.z[l + 1]->template as<V2>(), cs, om, zout, sc, zs);
    UFE_LAUNCH_CHECK();
    if (l == 0 && ev1) UFE_HIP(hipEventRecord(ev1, s_));
    if (l > 0) std::swap(av.z[l], av.zt[l]);
  }
}

// ---------------------------------------------------------------------------
// Captured iteration graphs.  A chunk of BiCGStab iterations is a fixed
// sequence of about 50 kernel launches per iteration whose arguments depend
// only on the solver's buffers, the matrix and the preconditioner settings, so
// it is captured once (hipStreamBeginCapture on the solver's stream) and
// replayed with one hipGraphLaunch.  The capture is redone when anything its
// launches read from the host side changes (the key below).  The V-cycle's
// level buffers alternate between the two applications of one iteration and
// are back in place after it, so replays and eager launches interleave freely.
// ---------------------------------------------------------------------------
template <typename T>
std::vector<uintptr_t> Krylov::graph_key(const BsrT<T> &A, const T *b, const T *x, int nbx) {
  auto bits = [](double d) { uintptr_t u = 0; std::memcpy(&u, &d, sizeof(d)); return u; };
  std::vector<uintptr_t> k = {(uintptr_t)A.nb, (uintptr_t)A.nnzb, (uintptr_t)A.ptr, (uintptr_t)A.col,
                              (uintptr_t)A.val, (uintptr_t)A.dinv, (uintptr_t)b, (uintptr_t)x, (uintptr_t)nbx,
                              (uintptr_t)pc_, (uintptr_t)amg_, (uintptr_t)A_diag_, (uintptr_t)fine32_,
                              (uintptr_t)amg_tail_rows_, (uintptr_t)unit_diag0_, (uintptr_t)amg_l1_,
                              bits(amg_omega_),
                              bits(amg_coarse_scale_), (uintptr_t)coarse_inv_.p, (uintptr_t)red_part_.p,
                              (uintptr_t)vec64_[0].p, (uintptr_t)vec32_[0].p, (uintptr_t)scal_.p};
  if (pc_ == UFE_PC_AMG && amg_) {
    AmgValues<T> &av = amg_vals<T>();
    const std::vector<BsrT<T>> &lev = std::is_same<T, double>::value
                                          ? *reinterpret_cast<const std::vector<BsrT<T>> *>(&lev64_)
                                          : *reinterpret_cast<const std::vector<BsrT<T>> *>(&lev32_);
    for (const BsrT<T> &l : lev)
      for (uintptr_t v : {(uintptr_t)l.nb, (uintptr_t)l.ptr, (uintptr_t)l.col, (uintptr_t)l.val, (uintptr_t)l.dinv})
        k.push_back(v);
    for (size_t l = 0; l < av.z.size(); ++l)
      for (uintptr_t v : {(uintptr_t)av.r[l]->p, (uintptr_t)av.z[l]->p, (uintptr_t)av.zt[l]->p}) k.push_back(v);
  }
  return k;
}

template <typename T>
hipGraphExec_t Krylov::iteration_graph(int n, const std::function<void(int)> &launch_iter, const BsrT<T> &A,
                                       const T *b, T *x, int nbx) {
  GraphCache &gc = graphs_[sizeof(T) == 4];
  std::vector<uintptr_t> key = graph_key<T>(A, b, x, nbx);
  if (gc.key != key) {
    gc.clear();
    gc.key = std::move(key);
  }
  const int idx = n >= 8 ? 3 : n >= 4 ? 2 : n >= 2 ? 1 : 0;
  if (!gc.exec[idx]) {
    hipGraph_t g = nullptr;
    UFE_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < (1 << idx); ++k) launch_iter(k);
    } catch (...) {
      (void)hipStreamEndCapture(s_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    UFE_HIP(hipStreamEndCapture(s_, &g));
    const hipError_t e = hipGraphInstantiate(&gc.exec[idx], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    UFE_HIP(e);
  }
  return gc.exec[idx];
}

void Krylov::GraphCache::clear() {
  for (hipGraphExec_t &e : exec) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
  key.clear();
}

template <typename T>
SolveStats Krylov::bicgstab(const BsrT<T> &A, const T *b, T *x, double rtol, double abstol, int maxit,
                            bool zero_guess, int check_every) {
  using V2 = typename VecT<T>::V2;
  T *r, *rhat, *p, *v, *phat, *s, *shat, *t;
  if (std::is_same<T, double>::value) {
    r = (T *)tmp64(0); rhat = (T *)tmp64(1); p = (T *)tmp64(2); v = (T *)tmp64(3);
    phat = (T *)tmp64(4); s = (T *)tmp64(5); shat = (T *)tmp64(6); t = (T *)tmp64(7);
  } else {
    r = (T *)tmp32(0); rhat = (T *)tmp32(1); p = (T *)tmp32(2); v = (T *)tmp32(3);
    phat = (T *)tmp32(4); s = (T *)tmp32(5); shat = (T *)tmp32(6); t = (T *)tmp32(7);
  }
  KrylovScalars *sc = scal_.as<KrylovScalars>();
  const bool amg = (pc_ == UFE_PC_AMG) && amg_;
  const int pcl = amg ? UFE_PC_NONE : pc_;
  const int nb = A.nb;
  SolveStats st;
  int total_its = 0;
  double tol_abs = 0.0;
  int nbx = std::min(blocks_for(nb), kMaxPartials);
  for (int restart = 0; restart < 4; ++restart) {
    RedSlot sl0 = slot(0), sl1 = slot(1), sl2 = slot(2);
    if (zero_guess && restart == 0) {
      hipLaunchKernelGGL(k_zero2<T>, dim3(blocks_for(nb)), dim3(kBlock), 0, s_, nb, reinterpret_cast<V2 *>(x));
      UFE_LAUNCH_CHECK();
    }
    // r and its norm (the norm of the system given: for the block-scaled
    // system, the block-Jacobi-scaled residual) for the stopping test
    spmv_launch<T, OP_RESID>(A, x, r, b, nullptr, sl0, sl0.result, nullptr, s_, 0.0);
    double h0[3];
    UFE_HIP(hipMemcpyAsync(h0, sl0.result, sizeof(h0), hipMemcpyDeviceToHost, s_));
    UFE_HIP(hipStreamSynchronize(s_));
    double rn = std::sqrt(h0[0]);
    if (restart == 0) {
      st.r0 = rn;
      tol_abs = std::max(rtol * rn, abstol);
    }
    st.rel = rn;
    if (rn <= tol_abs || !std::isfinite(rn)) {
      st.status = std::isfinite(rn) ? 1 : 2;
      break;
    }
    hipLaunchKernelGGL(k_bicg_init<T>, dim3(blocks_for(nb)), dim3(kBlock), 0, s_, nb, reinterpret_cast<V2 *>(r),
                       reinterpret_cast<V2 *>(rhat), reinterpret_cast<V2 *>(p), reinterpret_cast<V2 *>(v),
                       sl0.result, 0.0, tol_abs, maxit - total_its, sc);
    UFE_LAUNCH_CHECK();
    bool finished = false;
    const bool timing = time_spmv_ && std::is_same<T, double>::value;
    if (timing)
      while ((int)tev_.size() < 8 * check_every) {
        // timing-only events: no system-scope fence (no L2 writeback) when recorded
        hipEvent_t e;
        UFE_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        tev_.push_back(e);
      }
    int its_seen = 0;  // k_bicg_init zeroes the device counter
    // Iterations are launched in chunks and the host polls the device's
    // convergence flag after each; launches after convergence return at once
    // but still cost a dispatch each (~50 per iteration with the multigrid
    // preconditioner).  Chunks end at the previous solve's iteration count
    // (consecutive Picard systems need similar counts), then grow 1, 2, 4, ...
    // up to check_every past it.  Which iterations run is decided on the
    // device, so the chunking does not change the result.
    // One BiCGStab iteration: every kernel returns at once after convergence
    // (device flag), so a chunk may run past it.  k picks the timing events.
    auto launch_iter = [&](int k) {
      hipLaunchKernelGGL(k_bicg_p<T>, dim3(blocks_for(nb)), dim3(kBlock), 0, s_, nb,
                         reinterpret_cast<const V2 *>(r), reinterpret_cast<V2 *>(p), reinterpret_cast<const V2 *>(v),
                         reinterpret_cast<V2 *>(amg ? (T *)nullptr : phat), pcl, A.dinv, A.val, A_diag_, sc);
      UFE_LAUNCH_CHECK();
      if (amg)
        precond_apply<T>(A, p, phat, timing ? tev_[8 * k + 4] : nullptr, timing ? tev_[8 * k + 5] : nullptr);
      spmv_launch<T, OP_DOT1>(A, phat, v, nullptr, rhat, sl1, nullptr, sc, s_, 0.0, nullptr, 0,
                              timing ? tev_[8 * k] : nullptr, timing ? tev_[8 * k + 1] : nullptr);
      hipLaunchKernelGGL(k_bicg_s<T>, dim3(blocks_for(nb)), dim3(kBlock), 0, s_, nb,
                         reinterpret_cast<const V2 *>(r), reinterpret_cast<const V2 *>(v), reinterpret_cast<V2 *>(s),
                         reinterpret_cast<V2 *>(amg ? (T *)nullptr : shat), pcl, A.dinv, A.val, A_diag_, sc);
      UFE_LAUNCH_CHECK();
      if (amg)
        precond_apply<T>(A, s, shat, timing ? tev_[8 * k + 6] : nullptr, timing ? tev_[8 * k + 7] : nullptr);
      spmv_launch<T, OP_DOT2>(A, shat, t, nullptr, s, sl2, nullptr, sc, s_, 0.0, nullptr, 0,
                              timing ? tev_[8 * k + 2] : nullptr, timing ? tev_[8 * k + 3] : nullptr);
      hipLaunchKernelGGL(k_bicg_x<T>, dim3(nbx), dim3(kBlock), 0, s_, nb, reinterpret_cast<V2 *>(x),
                         reinterpret_cast<const V2 *>(phat), reinterpret_cast<const V2 *>(shat),
                         reinterpret_cast<const V2 *>(s), reinterpret_cast<const V2 *>(t), reinterpret_cast<V2 *>(r),
                         reinterpret_cast<const V2 *>(rhat), sl0, sc);
      UFE_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_bicg_x_finish, dim3(1), dim3(kBlock), 0, s_, sl0, nbx, sc);
      UFE_LAUNCH_CHECK();
    };
    const int pred = restart == 0 ? last_its_[sizeof(T) == 4] : 0;
    while (!finished) {
      const int chunk = pred > its_seen ? std::min(check_every, pred - its_seen)
                                        : std::min(check_every, std::max(1, its_seen - pred + 1));
      if (use_graphs_ && !timing) {
        // the chunk as replays of captured graphs of 8, 4, 2 and 1 iterations
        for (int left = chunk; left > 0;) {
          const int n = left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
          UFE_HIP(hipGraphLaunch(iteration_graph<T>(n, launch_iter, A, b, x, nbx), s_));
          left -= n;
        }
      } else {
        for (int k = 0; k < chunk; ++k) launch_iter(k);
      }
      UFE_HIP(hipMemcpyAsync(h_scal_, sc, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s_));
      UFE_HIP(hipStreamSynchronize(s_));
      if (verbose_ > 1)
        std::fprintf(stderr, "[ufe:krylov%s] its %d |r| %.3e tol %.3e rho %.3e omega %.3e done %d\n",
                     sizeof(T) == 4 ? "32" : "64", h_scal_->its, std::sqrt(h_scal_->rr), std::sqrt(h_scal_->tol2),
                     h_scal_->rho, h_scal_->omega, h_scal_->done);
      if (timing) {
        // launches after convergence return at once: count the iterations done
        const int real = std::max(0, std::min(chunk, h_scal_->its - its_seen));
        for (int j = 0; j < real; ++j)
          for (int q = 0; q < 4; ++q) {
            if (q >= 2 && !amg) break;
            float ms = 0.0f;
            UFE_HIP(hipEventElapsedTime(&ms, tev_[8 * j + 2 * q], tev_[8 * j + 2 * q + 1]));
            tms_[q / 2] += ms;
          }
        tn_[0] += 2 * real;
        tbytes_[0] = (double)A.nnzb * (sizeof(double4) + sizeof(int)) + (double)(A.nb + 1) * sizeof(int) +
                     3.0 * A.nb * sizeof(double2);
        if (amg) {
          // finest-level prolongation + smoothing: matrix, row pointers, aggregate
          // map, r (also the pre-smoothed z when folded), D^-1 unless folded, the
          // coarse correction, the output
          const double nc = amg_->n_levels() > 1 ? (double)amg_->levels[1]->nb : 0.0;
          const bool fold = unit_diag0_ && !amg_l1_;
          tn_[1] += 2 * real;
          tbytes_[1] = (double)A.nnzb * ((fine32_ ? sizeof(float4) : sizeof(double4)) + sizeof(int)) +
                       (double)(A.nb + 1) * sizeof(int) + (double)A.nb * sizeof(int) +
                       (fold ? 2.0 : 3.0) * A.nb * sizeof(double2) +
                       (fold ? 0.0 : (double)A.nb * sizeof(double4)) + nc * sizeof(double2);
        }
      }
      its_seen = h_scal_->its;
      if (h_scal_->done) finished = true;
    }
    total_its += h_scal_->its;
    if (restart == 0) last_its_[sizeof(T) == 4] = h_scal_->its;
    st.status = h_scal_->done;
    // verify the true residual (recurrence drift, breakdown) and restart if needed
    RedSlot sl3 = slot(3);
    spmv_launch<T, OP_RESID>(A, x, t, b, nullptr, sl3, sl3.result, nullptr, s_, 0.0);
    double hr[3];
    UFE_HIP(hipMemcpyAsync(hr, sl3.result, sizeof(hr), hipMemcpyDeviceToHost, s_));
    UFE_HIP(hipStreamSynchronize(s_));
    st.rel = std::sqrt(hr[0]);
    if (st.rel <= tol_abs * 1.0001) { st.status = 1; break; }
    if (total_its >= maxit) { st.status = 3; break; }
    zero_guess = false;
  }
  st.rel = st.r0 > 0 ? st.rel / st.r0 : st.rel;
  st.its = total_its;
  return st;
}

SolveStats Krylov::bicgstab64(const BsrT<double> &A, const double *b, double *x, double rtol, double abstol,
                              int maxit, bool zero_guess, int check_every) {
  return bicgstab<double>(A, b, x, rtol, abstol, maxit, zero_guess, check_every);
}

SolveStats Krylov::bicgstab32(const BsrT<float> &A, const float *b, float *x, double rtol, double abstol,
                              int maxit, bool zero_guess, int check_every) {
  return bicgstab<float>(A, b, x, rtol, abstol, maxit, zero_guess, check_every);
}

// ===========================================================================
// GMRES(m), right-preconditioned, on the device.
//
// PETSc's default KSP (petsc_basic.f90:106-119 keeps it through
// KSPSetFromOptions) is restarted GMRES(30) with classical Gram-Schmidt.  Here
// the preconditioner is applied on the right, so the Arnoldi residual
// estimate is the norm of b - A x itself: with the block-Jacobi-scaled system
// that is the norm PETSc's left block-Jacobi test measures, the same norm the
// BiCGStab stopping test uses.  Per Arnoldi step j (all on the stream, no host
// synchronisation): z = M^-1 v_j; w = A z; h_ij = (w, v_i), i <= j (one pass
// over the basis with w held in registers, deterministic two-phase sums);
// w -= sum h_ij v_i with ||w||^2 (second pass); then one thread applies the
// previous Givens rotations to column j, forms the new one, updates the
// residual estimate |g_j+1| and sets the done flag; v_j+1 = w / h_j+1,j.  At
// the end of a cycle one thread solves the triangular system H y = g and
// x += M^-1 (V y).  The host polls the done flag between chunks of steps, as
// for BiCGStab, and recomputes the true residual at each restart.
// ===========================================================================
constexpr int kGmresMaxRestart = 64;
constexpr int kGmresLd = kGmresMaxRestart + 1;  // leading dimension of H

struct GmresScalars {
  double H[kGmresLd * kGmresMaxRestart];  // column j at H[j * kGmresLd + i]
  double cs[kGmresMaxRestart], sn[kGmresMaxRestart];
  double g[kGmresLd], y[kGmresMaxRestart];
  double inv_h;  // 1 / h_j+1,j of the last column (normalisation of v_j+1)
  int cols;      // Arnoldi columns completed in this cycle
};

// v0 = r / ||r||, g = (||r||, 0, ...), cycle counters reset
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gmres_init(long n, T *__restrict__ v0, const double *__restrict__ rr,
                                                       double tol_abs, int maxit, KrylovScalars *__restrict__ sc,
                                                       GmresScalars *__restrict__ gs) {
  const double beta = sqrt(rr[0]);
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) v0[i] = (T)((double)v0[i] / beta);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int k = 0; k < kGmresLd; ++k) gs->g[k] = 0.0;
    gs->g[0] = beta;
    gs->cols = 0;
    gs->inv_h = 0.0;
    sc->done = 0;
    sc->its = 0;
    sc->maxit = maxit;
    sc->tol2 = tol_abs * tol_abs;
    sc->rr = beta * beta;
  }
}

// z = M_local v (none, point or block Jacobi)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_pc_apply(int nb, const typename VecT<T>::V2 *__restrict__ v,
                                                     typename VecT<T>::V2 *__restrict__ z, int pc,
                                                     const typename VecT<T>::V4 *__restrict__ dinv,
                                                     const typename VecT<T>::V4 *__restrict__ aval,
                                                     const int *__restrict__ diag, const KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) z[i] = pc_local<T>(pc, dinv, aval, diag, i, v[i]);
}

// Partial sums of h_i = (w, v_i), i < ncol: R entries of w per lane stay in
// registers while the basis vectors stream past.
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_gmres_dots(long n, int ncol, const T *__restrict__ V, long ldv,
                                                       const T *__restrict__ w, double *__restrict__ part,
                                                       const KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  const long base = (long)blockIdx.x * kBlock * R + threadIdx.x;
  double wr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const long idx = base + (long)r * kBlock;
    wr[r] = idx < n ? (double)w[idx] : 0.0;
  }
  __shared__ double lds[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int i = 0; i < ncol; ++i) {
    const T *v = V + (long)i * ldv;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long idx = base + (long)r * kBlock;
      if (idx < n) acc += wr[r] * (double)v[idx];
    }
    acc = wave_sum(acc);
    if (lane == 0) lds[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kBlock / 64; ++q) t += lds[q];
      part[(long)i * kMaxPartials + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// Sum of `ncol` partial sets (nparts each, index order) in one workgroup.
__device__ __forceinline__ double gmres_partial_sum(const double *__restrict__ part, int nparts, double (*lds)) {
  double t = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) t += part[b];
  t = wave_sum(t);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = t;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < (int)(blockDim.x / 64); ++q) s += lds[q];
  return s;
}

// All h_ij at once: one wavefront per column (columns round robin over the
// workgroup's 16 wavefronts), four independent partial sums per lane so the
// loads of a column are in flight together, a shuffle reduction per column;
// no workgroup barrier.  The summation order is fixed by nparts.
constexpr int kGmresFinishBlock = 1024;
__global__ void __launch_bounds__(kGmresFinishBlock) k_gmres_dots_finish(int ncol, int nparts,
                                                                         const double *__restrict__ part,
                                                                         GmresScalars *__restrict__ gs, int j,
                                                                         const KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int i = wv; i < ncol; i += nw) {
    const double *__restrict__ p = part + (long)i * kMaxPartials;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = lane;
    for (; b + 192 < nparts; b += 256) {
      s0 += p[b];
      s1 += p[b + 64];
      s2 += p[b + 128];
      s3 += p[b + 192];
    }
    for (; b < nparts; b += 64) s0 += p[b];
    const double t = wave_sum((s0 + s1) + (s2 + s3));
    if (lane == 0) gs->H[j * kGmresLd + i] = t;
  }
}

// w -= sum_i h_ij v_i, partial sums of ||w||^2
template <typename T, int R>
__global__ void __launch_bounds__(kBlock) k_gmres_update(long n, int ncol, const T *__restrict__ V, long ldv,
                                                         T *__restrict__ w, const GmresScalars *__restrict__ gs, int j,
                                                         double *__restrict__ part,
                                                         const KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  __shared__ double h[kGmresLd];
  for (int i = threadIdx.x; i < ncol; i += blockDim.x) h[i] = gs->H[j * kGmresLd + i];
  __syncthreads();
  const long base = (long)blockIdx.x * kBlock * R + threadIdx.x;
  double wr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const long idx = base + (long)r * kBlock;
    wr[r] = idx < n ? (double)w[idx] : 0.0;
  }
  for (int i = 0; i < ncol; ++i) {
    const T *v = V + (long)i * ldv;
    const double hi = h[i];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long idx = base + (long)r * kBlock;
      if (idx < n) wr[r] -= hi * (double)v[idx];
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const long idx = base + (long)r * kBlock;
    if (idx < n) {
      const T wt = (T)wr[r];
      w[idx] = wt;
      acc += (double)wt * (double)wt;
    }
  }
  __shared__ double lds[kBlock / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) t += lds[q];
    part[blockIdx.x] = t;
  }
}

// h_j+1,j = ||w||; rotations; residual estimate; stop flags (one workgroup)
__global__ void __launch_bounds__(kBlock) k_gmres_norm_finish(int nparts, const double *__restrict__ part,
                                                              GmresScalars *__restrict__ gs, int j,
                                                              KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  __shared__ double lds[kBlock / 64];
  const double hn = sqrt(gmres_partial_sum(part, nparts, lds));
  if (threadIdx.x != 0) return;
  double *H = gs->H + j * kGmresLd;
  for (int i = 0; i < j; ++i) {
    const double a = H[i], b = H[i + 1];
    H[i] = gs->cs[i] * a + gs->sn[i] * b;
    H[i + 1] = -gs->sn[i] * a + gs->cs[i] * b;
  }
  const double a = H[j];
  const double r = sqrt(a * a + hn * hn);
  const double c = r > 0.0 ? a / r : 1.0, s = r > 0.0 ? hn / r : 0.0;
  gs->cs[j] = c;
  gs->sn[j] = s;
  H[j] = r;
  H[j + 1] = 0.0;
  const double gj = gs->g[j];
  gs->g[j] = c * gj;
  gs->g[j + 1] = -s * gj;
  gs->cols = j + 1;
  gs->inv_h = hn > 0.0 ? 1.0 / hn : 0.0;
  sc->its += 1;
  const double res = fabs(gs->g[j + 1]);
  sc->rr = res * res;
  if (!std::isfinite(res) || !std::isfinite(r)) sc->done = 2;
  else if (res * res <= sc->tol2 || hn == 0.0) sc->done = 1;  // converged, or the Krylov space is invariant
  else if (r == 0.0) sc->done = 2;
  else if (sc->its >= sc->maxit) sc->done = 3;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_gmres_scale(long n, T *__restrict__ w, const GmresScalars *__restrict__ gs,
                                                        const KrylovScalars *__restrict__ sc) {
  if (sc->done) return;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) w[i] = (T)((double)w[i] * gs->inv_h);
}

// y = H^-1 g over the columns of the cycle (upper triangular after the rotations)
__global__ void k_gmres_solve_y(GmresScalars *__restrict__ gs) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int m = gs->cols;
  for (int i = m - 1; i >= 0; --i) {
    double t = gs->g[i];
    for (int k = i + 1; k < m; ++k) t -= gs->H[k * kGmresLd + i] * gs->y[k];
    const double d = gs->H[i * kGmresLd + i];
    gs->y[i] = d != 0.0 ? t / d : 0.0;
  }
}

// t = sum_i y_i v_i over the cycle's columns
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gmres_combine(long n, const T *__restrict__ V, long ldv,
                                                          const GmresScalars *__restrict__ gs, T *__restrict__ t) {
  __shared__ double y[kGmresMaxRestart];
  const int m = gs->cols;
  for (int i = threadIdx.x; i < m; i += blockDim.x) y[i] = gs->y[i];
  __syncthreads();
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int k = 0; k < m; ++k) acc += y[k] * (double)V[(long)k * ldv + i];
  t[i] = (T)acc;
}

__global__ void k_gmres_set_done(KrylovScalars *__restrict__ sc, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->done = v;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_gmres_axpy(long n, T *__restrict__ x, const T *__restrict__ z) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) x[i] += z[i];
}

template <typename T, int R>
static void gmres_dots_update(long n, int ncol, const T *V, long ldv, T *w, double *part, GmresScalars *gs, int j,
                              KrylovScalars *sc, hipStream_t s) {
  const int g = (int)((n + (long)kBlock * R - 1) / ((long)kBlock * R));
  hipLaunchKernelGGL((k_gmres_dots<T, R>), dim3(g), dim3(kBlock), 0, s, n, ncol, V, ldv, w, part, sc);
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gmres_dots_finish, dim3(1), dim3(kGmresFinishBlock), 0, s, ncol, g, part, gs, j, sc);
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_gmres_update<T, R>), dim3(g), dim3(kBlock), 0, s, n, ncol, V, ldv, w, gs, j, part, sc);
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_gmres_norm_finish, dim3(1), dim3(kBlock), 0, s, g, part, gs, j, sc);
  UFE_LAUNCH_CHECK();
}

template <typename T>
SolveStats Krylov::gmres(const BsrT<T> &A, const T *b, T *x, double rtol, double abstol, int maxit, bool zero_guess,
                         int m, int check_every) {
  using V2 = typename VecT<T>::V2;
  m = std::max(1, std::min(m, kGmresMaxRestart));
  const int nb = A.nb;
  const long n = 2L * nb;
  // registers per lane for the two basis passes: the grid must fit the partial slots
  const int R = n <= (long)kMaxPartials * kBlock * 4 ? 4 : n <= (long)kMaxPartials * kBlock * 16 ? 16 : 0;
  if (R == 0) throw std::runtime_error("GMRES: system too large for the reduction slots");
  const size_t vbytes = sizeof(T) * (size_t)n;
  if (gm_V_.bytes < vbytes * (m + 1)) gm_V_.alloc(vbytes * (m + 1));
  if (gm_z_.bytes < vbytes) { gm_z_.alloc(vbytes); gm_t_.alloc(vbytes); }
  if (!gm_sc_.p) {
    gm_sc_.alloc(sizeof(GmresScalars));
    gm_part_.alloc(sizeof(double) * (size_t)kMaxPartials * kGmresLd);
  }
  T *V = gm_V_.as<T>(), *z = gm_z_.as<T>(), *t = gm_t_.as<T>();
  GmresScalars *gs = gm_sc_.as<GmresScalars>();
  double *part = gm_part_.as<double>();
  KrylovScalars *sc = scal_.as<KrylovScalars>();
  const bool amg = (pc_ == UFE_PC_AMG) && amg_;
  const int pcl = amg ? UFE_PC_NONE : pc_;
  if (zero_guess) {
    hipLaunchKernelGGL(k_zero2<T>, dim3(blocks_for(nb)), dim3(kBlock), 0, s_, nb, reinterpret_cast<V2 *>(x));
    UFE_LAUNCH_CHECK();
  }
  SolveStats st;
  int total = 0;
  double tol_abs = 0.0;
  RedSlot sl0 = slot(0);
  const int pred0 = last_gmres_its_[sizeof(T) == 4];
  for (int cycle = 0;; ++cycle) {
    // true residual into v0 (the norm of the system given)
    spmv_launch<T, OP_RESID>(A, x, V, b, nullptr, sl0, sl0.result, nullptr, s_, 0.0);
    double h0[3];
    UFE_HIP(hipMemcpyAsync(h0, sl0.result, sizeof(h0), hipMemcpyDeviceToHost, s_));
    UFE_HIP(hipStreamSynchronize(s_));
    const double rn = std::sqrt(h0[0]);
    if (cycle == 0) {
      st.r0 = rn;
      tol_abs = std::max(rtol * rn, abstol);
    }
    st.rel = rn;
    if (!std::isfinite(rn)) { st.status = 2; break; }
    if (rn <= tol_abs * (cycle ? 1.0001 : 1.0)) { st.status = 1; break; }
    if (total >= maxit) { st.status = 3; break; }
    hipLaunchKernelGGL(k_gmres_init<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s_, n, V, sl0.result, tol_abs,
                       maxit - total, sc, gs);
    UFE_LAUNCH_CHECK();
    int j = 0, seen = 0;
    // first poll where the previous solve stopped, then every check_every steps
    int next_poll = cycle == 0 && pred0 > 0 ? std::min(pred0, m) : std::min(check_every, m);
    while (j < m) {
      for (; j < next_poll; ++j) {
        T *vj = V + (long)j * n, *w = V + (long)(j + 1) * n;
        if (custom_pc_ && sizeof(T) == sizeof(double)) {
          custom_pc_(reinterpret_cast<const double *>(vj), reinterpret_cast<double *>(z));
        } else if (amg) {
          precond_apply<T>(A, vj, z);
        } else {
          hipLaunchKernelGGL(k_pc_apply<T>, dim3(blocks_for(nb)), dim3(kBlock), 0, s_, nb,
                             reinterpret_cast<const V2 *>(vj), reinterpret_cast<V2 *>(z), pcl, A.dinv, A.val,
                             A_diag_, sc);
          UFE_LAUNCH_CHECK();
        }
        spmv_launch<T, OP_PLAIN>(A, z, w, nullptr, nullptr, sl0, nullptr, sc, s_);
        if (R == 4) gmres_dots_update<T, 4>(n, j + 1, V, n, w, part, gs, j, sc, s_);
        else gmres_dots_update<T, 16>(n, j + 1, V, n, w, part, gs, j, sc, s_);
        hipLaunchKernelGGL(k_gmres_scale<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s_, n, w, gs, sc);
        UFE_LAUNCH_CHECK();
      }
      UFE_HIP(hipMemcpyAsync(h_scal_, sc, sizeof(KrylovScalars), hipMemcpyDeviceToHost, s_));
      UFE_HIP(hipStreamSynchronize(s_));
      if (verbose_ > 1)
        std::fprintf(stderr, "[ufe:gmres%s] cycle %d its %d |r| %.3e tol %.3e done %d\n", sizeof(T) == 4 ? "32" : "64",
                     cycle, total + h_scal_->its, std::sqrt(h_scal_->rr), tol_abs, h_scal_->done);
      seen = h_scal_->its;
      if (h_scal_->done) break;
      next_poll = std::min(m, j + check_every);
    }
    total += seen;
    const int status = h_scal_->done;
    // x += M^-1 V y over the columns of this cycle
    hipLaunchKernelGGL(k_gmres_solve_y, dim3(1), dim3(1), 0, s_, gs);
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gmres_combine<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s_, n, V, n, gs, t);
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gmres_set_done, dim3(1), dim3(1), 0, s_, sc, 0);
    UFE_LAUNCH_CHECK();
    if (custom_pc_ && sizeof(T) == sizeof(double)) {
      custom_pc_(reinterpret_cast<const double *>(t), reinterpret_cast<double *>(z));
    } else if (amg) {
      precond_apply<T>(A, t, z);
    } else {
      hipLaunchKernelGGL(k_pc_apply<T>, dim3(blocks_for(nb)), dim3(kBlock), 0, s_, nb, reinterpret_cast<const V2 *>(t),
                         reinterpret_cast<V2 *>(z), pcl, A.dinv, A.val, A_diag_, sc);
      UFE_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_gmres_axpy<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s_, n, x, z);
    UFE_LAUNCH_CHECK();
    if (cycle == 0) last_gmres_its_[sizeof(T) == 4] = seen;
    if (status == 2 && seen == 0) { st.status = 2; break; }
  }
  st.its = total;
  st.rel = st.r0 > 0 ? st.rel / st.r0 : st.rel;
  return st;
}

SolveStats Krylov::gmres64(const BsrT<double> &A, const double *b, double *x, double rtol, double abstol, int maxit,
                           bool zero_guess, int restart, int check_every) {
  return gmres<double>(A, b, x, rtol, abstol, maxit, zero_guess, restart, check_every);
}

SolveStats Krylov::gmres32(const BsrT<float> &A, const float *b, float *x, double rtol, double abstol, int maxit,
                           bool zero_guess, int restart, int check_every) {
  return gmres<float>(A, b, x, rtol, abstol, maxit, zero_guess, restart, check_every);
}

// Generic wrappers declared in kernels.h
void launch_bsr_spmv64(const Bsr64 &A, const double *x, double *y, hipStream_t s) {
  BsrT<double> B;
  B.nb = A.nb; B.nnzb = A.nnzb; B.ptr = A.ptr; B.col = A.col; B.val = A.val;
  RedSlot dummy{};
  spmv_launch<double, OP_PLAIN>(B, x, y, nullptr, nullptr, dummy, nullptr, nullptr, s);
}

void launch_bsr_spmv32(const Bsr32 &A, const float *x, float *y, hipStream_t s) {
  BsrT<float> B;
  B.nb = A.nb; B.nnzb = A.nnzb; B.ptr = A.ptr; B.col = A.col; B.val = A.val;
  RedSlot dummy{};
  spmv_launch<float, OP_PLAIN>(B, x, y, nullptr, nullptr, dummy, nullptr, nullptr, s);
}

void launch_bsr_residual64(const Bsr64 &A, const double *x, const double *b, double *r, const RedSlot *slot,
                           double *out2, hipStream_t s) {
  BsrT<double> B;
  B.nb = A.nb; B.nnzb = A.nnzb; B.ptr = A.ptr; B.col = A.col; B.val = A.val;
  spmv_launch<double, OP_RESID>(B, x, r, b, nullptr, *slot, out2, nullptr, s);
}

void launch_f64_to_f32(int n, const double *a, float *b, hipStream_t s) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, a, b);
  UFE_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Scalar CSR product (generic boundary, MatMult)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_csr_spmv(int m, const int *__restrict__ ptr, const int *__restrict__ ind,
                                                     const double *__restrict__ val, const double *__restrict__ x,
                                                     double *__restrict__ y) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[ind[k]];
  y[i] = s;
}

void launch_csr_spmv64(int m, const int *ptr, const int *ind, const double *val, const double *x, double *y,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_csr_spmv, dim3(blocks_for(m)), dim3(kBlock), 0, s, m, ptr, ind, val, x, y);
  UFE_LAUNCH_CHECK();
}

}  // namespace ufe
