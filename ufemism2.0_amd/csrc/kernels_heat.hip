// Grounded fractions on the triangles and the vertical heat columns, with
// their C ABI entry points.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "krylov.h"
#include "ufe_internal.h"

namespace ufe {

// ---------------------------------------------------------------------------
// calc_grounded_fractions_bilin_interp_TAF_b (subgrid_grounded_fractions_bilin_TAF.f90:182-250)
// with calc_grounded_area_triangle (:252-370); TAF per vertex from
// thickness_above_floatation (ice_geometry_basics.f90:43-56).  FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tri_area(double2 p, double2 q, double2 r) {
  return fabs((q.x - p.x) * (r.y - p.y) - (q.y - p.y) * (r.x - p.x)) / 2.0;
}

__device__ __forceinline__ double sub_area(double2 a, double2 b, double2 c, double ta, double tb, double tc) {
  double lab = ta / (ta - tb), lac = ta / (ta - tc);
  double2 pab = make_double2(a.x * (1.0 - lab) + b.x * lab, a.y * (1.0 - lab) + b.y * lab);
  double2 pac = make_double2(a.x * (1.0 - lac) + c.x * lac, a.y * (1.0 - lac) + c.y * lac);
  return tri_area(a, pab, pac);
}

__global__ void __launch_bounds__(kBlock) k_fraction_gr_b(int nT, const int *__restrict__ Tri,
                                                          const double2 *__restrict__ V, const double *__restrict__ Hi,
                                                          const double *__restrict__ Hb, const double *__restrict__ SL,
                                                          double *__restrict__ fgr) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nT) return;
  int vi[3];
  double T[3];
  for (int n = 0; n < 3; ++n) {
    vi[n] = Tri[3 * ti + n];
    T[n] = Hi[vi[n]] - fmax(0.0, (SL[vi[n]] - Hb[vi[n]]) * (kSeawaterDensity / kIceDensity));
  }
  double tmax = fmax(T[0], fmax(T[1], T[2])), tmin = fmin(T[0], fmin(T[1], T[2]));
  if (tmin >= 0.0) { fgr[ti] = 1.0; return; }
  if (tmax <= 0.0) { fgr[ti] = 0.0; return; }
  const double tol = 1e-9;
  double2 va = V[vi[0]], vb = V[vi[1]], vc = V[vi[2]];
  double At = tri_area(va, vb, vc), Ag;
  double a = T[0] >= 0 ? fmax(tol, T[0]) : fmin(-tol, T[0]);
  double b = T[1] >= 0 ? fmax(tol, T[1]) : fmin(-tol, T[1]);
  double c = T[2] >= 0 ? fmax(tol, T[2]) : fmin(-tol, T[2]);
  if (a >= 0 && b >= 0 && c >= 0) Ag = At;
  else if (a <= 0 && b <= 0 && c <= 0) Ag = 0.0;
  else if (a >= 0 && b <= 0 && c <= 0) Ag = sub_area(va, vb, vc, a, b, c);
  else if (a <= 0 && b >= 0 && c <= 0) Ag = sub_area(vb, vc, va, b, c, a);
  else if (a <= 0 && b <= 0 && c >= 0) Ag = sub_area(vc, va, vb, c, a, b);
  else if (a <= 0 && b >= 0 && c >= 0) Ag = At - sub_area(va, vb, vc, a, b, c);
  else if (a >= 0 && b <= 0 && c >= 0) Ag = At - sub_area(vb, vc, va, b, c, a);
  else Ag = At - sub_area(vc, va, vb, c, a, b);
  fgr[ti] = fmin(1.0, fmax(0.0, Ag / At));
}

void launch_fraction_gr_b(const DevMesh &m, const double *Hi, const double *Hb, const double *SL, double *fgr,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_fraction_gr_b, dim3(blocks_for(m.nT)), dim3(kBlock), 0, s, m.nT, m.Tri,
                     reinterpret_cast<const double2 *>(m.V), Hi, Hb, SL, fgr);
  UFE_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Heat columns: one thread per vertex column.  Inputs are vertex-fastest
// (nV, nz), so at each layer k the threads of a wavefront read consecutive
// addresses.  The column (<= kHeatMaxNz layers) is held in registers.
//   solve_3D_heat_equation per-vertex loop (thermodynamics_3D_heat_equation.f90:133-252)
//   solve_1D_heat_equation (:309-479), Thomas (tridiagonal_solver.f90:13-51)
//   Cp (Pounder 1965), K (Ritz 1987), T_pmp from the input T and H
//   (thermodynamics_utilities.f90:130-229)
// ---------------------------------------------------------------------------
constexpr int kHeatMaxNz = 32;

// The kernel is instantiated for NZ >= nz layers; every array index is static
// once the NZ-loops are unrolled (the layer count enters only through the
// k < nz and k == nz - 1 tests), so for NZ <= 16 a column stays in registers.
// Only the column state, the Thomas pivots and the super-diagonal are held:
// each implicit step streams the layer's inputs from memory and forms that
// row's coefficients inside the forward sweep, which keeps the register
// footprint low enough for several wavefronts per SIMD (the inputs of a
// block's columns stay in L2 for the rare extra steps of dt halving).

// One implicit step of column vi (solve_1D_heat_equation,
// thermodynamics_3D_heat_equation.f90:309-479): rows 1..nz-2 from the
// advection/diffusion coefficients, surface row T = min(T_surf, T0), basal row
// grounded T = min(T_pmp, T(nz-2) - dzeta Q / (dzeta/dz K)) with the old
// T(nz-2) (explicit), floating T = T_pmp; Thomas (tridiagonal_solver.f90:13-51);
// clamp to T_pmp.  Conductivity and heat capacity come from the input
// temperature Ti of the call (the reference forms them once before the
// column loop, :90-112); T is the column at the start of the step (on the
// first step, first = true, it is read from Ti as the sweep passes).  The
// forward sweep stores its right-hand side in T; the back substitution
// overwrites it with the new column.
// mode 1 grounded, 2 floating, 3 both, kept as fgr * grounded + (1 - fgr) *
// floating (grounding line, :171-203).  The two systems differ only in the
// basal row, which has no sub-diagonal, so they share the forward sweep; the
// back substitution carries both recursions and blends the clamped profiles.
template <int NZ>
__device__ __forceinline__ void heat_step(const HeatDev &h, const double *__restrict__ tri,
                                          const double *__restrict__ zeta, int vi, bool first, double (&T)[NZ],
                                          double T_surf, double Hpmp, double dzb, int mode, double fgr, double Q,
                                          double dt) {
  const int nV = h.nV, nz = h.nz;
  // The layer loads are the same on every sub-step; an empty asm on the
  // column index makes them opaque to loop-invariant code motion, which
  // would otherwise lift all of them (nz * 11 values) out of the dt-halving
  // loops and spill.  It emits no instruction.
  size_t col = (size_t)vi;
  asm volatile("" : "+v"(col));
  // rb: reciprocals of the Thomas pivots, formed once and used by both sweeps
  double rb[NZ], cl[NZ];
  rb[0] = 1.0;
  const double idt = 1.0 / dt;
  cl[0] = 0.0;
  if (first) T[0] = h.Ti[col];
  double told = T[0];        // the step's starting value of the previous layer
  T[0] = fmin(T_surf, kT0);  // right-hand side of the surface row
  double cprev = 0.0;        // super-diagonal of the previous row (0 for the surface row)
  double dfl = 0.0;          // basal entry of the floating system (mode 3)
#pragma unroll
  for (int k = 1; k < NZ; ++k) {
    cl[k] = 0.0;
    rb[k] = 1.0;
    if (k < nz) {
      const size_t o = (size_t)k * nV + col;
      const double Tin = h.Ti[o];
      const double K = 3.101e8 * exp(-0.0057 * Tin);
      const double zz = h.zz[o];
      const double tk = first ? Tin : T[k];
      double a, b, dk, c = 0.0;
      if (k < nz - 1) {
        const double Cp = 2115.3 + 7.79293 * (Tin - kT0);
        const double c1 = h.zt[o] + h.u[o] * h.zx[o] + h.v[o] * h.zy[o] + h.w[o] * zz;
        const double rcp = 1.0 / (kIceDensity * Cp);
        const double c2 = -K * rcp * zz * zz;
        a = c1 * tri[0 * kHeatMaxNz + k - 1] + c2 * tri[3 * kHeatMaxNz + k - 1];
        const double bl = c1 * tri[1 * kHeatMaxNz + k] + c2 * tri[4 * kHeatMaxNz + k];
        c = c1 * tri[2 * kHeatMaxNz + k] + c2 * tri[5 * kHeatMaxNz + k];
        const double r0 = -h.uadv[o] - h.vadv[o] + h.Phi[o] * rcp;
        b = idt + bl;
        dk = tk * idt + r0;
      } else {
        // floating: the basal temperature is the pressure-melting point here
        const double tpmp = kT0 - kClausiusClapeyron * Hpmp * zeta[k];
        a = 0.0;
        b = 1.0;
        dk = mode != 2 ? fmin(tpmp, told - dzb * Q / (zz * K)) : tpmp;
        dfl = tpmp;
      }
      const double w = a * rb[k - 1];
      rb[k] = 1.0 / (b - w * cprev);
      const double dprev = T[k - 1];
      told = tk;
      T[k] = dk - w * dprev;
      if (k == nz - 1) dfl = dfl - w * dprev;
      cl[k] = c;
      cprev = c;
    }
  }
  if (mode != 3) {
#pragma unroll
    for (int k = NZ - 1; k >= 0; --k) {
      if (k == nz - 1) T[k] = T[k] * rb[k];
      else if (k < nz - 1) T[k] = (T[k] - cl[k] * T[k + 1]) * rb[k];
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k)
      if (k < nz) T[k] = fmin(T[k], kT0 - kClausiusClapeyron * Hpmp * zeta[k]);
  } else {
    double tg = 0.0, tf = 0.0;  // unclamped solutions at layer k + 1
#pragma unroll
    for (int k = NZ - 1; k >= 0; --k) {
      if (k < nz) {
        const double dk = T[k];
        tg = (dk - cl[k] * tg) * rb[k];
        tf = ((k == nz - 1 ? dfl : dk) - cl[k] * tf) * rb[k];
        const double tpmp = kT0 - kClausiusClapeyron * Hpmp * zeta[k];
        T[k] = fgr * fmin(tg, tpmp) + (1.0 - fgr) * fmin(tf, tpmp);
      }
    }
  }
}

// Occupancy target per column height (a minimum: the compiler may fit more):
// the register budget it leaves lets the scheduler prefetch a few layers
// ahead without spilling.  Measured on MI355X (399k columns, 100k for
// nz = 30): a target of 2 wavefronts per SIMD was fastest for nz = 12 and 30,
// 4 for nz = 8; with the reciprocal form the kernels take 101, 104 and 63 us
// (nz = 12 then fits 3 wavefronts).
template <int NZ>
constexpr int heat_waves() { return NZ <= 8 ? 4 : 2; }

template <int NZ>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(heat_waves<NZ>())))
k_heat_columns(HeatDev h, double *__restrict__ Tout,
                                                         int *__restrict__ unstable, RedSlot slot,
                                                         double *__restrict__ count) {
  __shared__ double tri[6 * kHeatMaxNz];
  __shared__ double zeta[kHeatMaxNz];
  for (int i = threadIdx.x; i < 6 * kHeatMaxNz; i += blockDim.x) tri[i] = h.tri[i];
  for (int i = threadIdx.x; i < h.nz; i += blockDim.x) zeta[i] = h.zeta[i];
  __syncthreads();
  const int nV = h.nV, nz = h.nz;
  const double dzb = zeta[nz - 1] - zeta[nz - 2];
  double nun = 0.0, nbad = 0.0;
#pragma unroll 1
  for (int vi = blockIdx.x * blockDim.x + threadIdx.x; vi < nV; vi += gridDim.x * blockDim.x) {
    unstable[vi] = 0;
    const double H = h.Hi_eff[vi];
    const double ts = h.T_surf[vi];
    if (H < h.Hi_min) {
      for (int k = 0; k < nz; ++k) Tout[(size_t)k * nV + vi] = ts;
      continue;
    }
    const int mode = h.mask[vi];
    if (mode < 1 || mode > 3 || (mode == 3 && !h.fgr)) {
      // thick ice neither grounded nor floating: the call reports an error
      nbad += 1.0;
      continue;
    }
    const double fgr = mode == 3 ? h.fgr[vi] : 1.0;
    const double Q = h.Q[vi];
    // dt halving without resetting the column between attempts
    // (thermodynamics_3D_heat_equation.f90:161-239)
    double Tc[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) Tc[k] = 0.0;
    bool stable = false;
#pragma unroll 1
    for (int it = 1; it <= 10 && !stable; ++it) {
      const double dta = ldexp(h.dt, 1 - it);  // dt / 2^(it-1), exact
      const int nsub = 1 << (it - 1);
#pragma unroll 1
      for (int sstep = 0; sstep < nsub; ++sstep)
        heat_step<NZ>(h, tri, zeta, vi, it == 1, Tc, ts, H, dzb, mode, fgr, Q, dta);
      stable = true;
#pragma unroll
      for (int k = 0; k < NZ; ++k)
        if (k < nz && (Tc[k] != Tc[k] || Tc[k] < 180.0 || Tc[k] > kT0)) stable = false;
    }
    if (!stable) { unstable[vi] = 1; nun += 1.0; }
#pragma unroll
    for (int k = 0; k < NZ; ++k)
      if (k < nz) Tout[(size_t)k * nV + vi] = Tc[k];
  }
  double acc[2] = {nun, nbad};
  __shared__ double lds[2][kBlock / 64];
  block_sum<2>(acc, lds);
  block_partials<2>(slot, acc);
}

void launch_heat_columns(const HeatDev &h, double *Tout, int *unstable, const RedSlot &slot, double *count,
                         hipStream_t s) {
  int nb = std::min(blocks_for(h.nV), kMaxPartials);
  if (h.nz <= 8)
    hipLaunchKernelGGL(k_heat_columns<8>, dim3(nb), dim3(kBlock), 0, s, h, Tout, unstable, slot, count);
  else if (h.nz <= 12)
    hipLaunchKernelGGL(k_heat_columns<12>, dim3(nb), dim3(kBlock), 0, s, h, Tout, unstable, slot, count);
  else if (h.nz <= 16)
    hipLaunchKernelGGL(k_heat_columns<16>, dim3(nb), dim3(kBlock), 0, s, h, Tout, unstable, slot, count);
  else
    hipLaunchKernelGGL(k_heat_columns<kHeatMaxNz>, dim3(nb), dim3(kBlock), 0, s, h, Tout, unstable, slot, count);
  UFE_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_partials_out<2>, dim3(1), dim3(kBlock), 0, s, slot, nb, count);
  UFE_LAUNCH_CHECK();
}

// Robin solution of one flagged column (replace_Ti_with_robin_solution,
// thermodynamics_utilities.f90:256-337).  One thread per vertex; unflagged
// columns are left alone.
constexpr double kSecPerYear = 31556943.36;  // parameters.f90:16
__global__ void __launch_bounds__(kBlock) k_heat_robin(int nV, int nz, const double *__restrict__ zeta,
                                                       const double *__restrict__ Hi_eff,
                                                       const double *__restrict__ T_surf,
                                                       const double *__restrict__ SMB,
                                                       const double *__restrict__ G, const int *__restrict__ mask,
                                                       const int *__restrict__ flag, double Hi_min,
                                                       double *__restrict__ Ti) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi >= nV || flag[vi] != 1) return;
  // conductivity at T0 (Ritz 1987), diffusivity with c_0 = 2127.5 J/kg/K
  const double Kr = 9.828 * kSecPerYear * exp(-0.0057 * kT0);
  const double kappa = Kr / (kIceDensity * 2127.5);
  const double grad_b = -G[vi] / Kr;
  const double Ts = fmin(kT0, T_surf[vi]);
  const double H = Hi_eff[vi];
  for (int k = 0; k < nz; ++k) {
    double T = Ti[(size_t)k * nV + vi];
    if (H > Hi_min) {
      if (mask[vi] == 1 || mask[vi] == 3) {
        if (SMB[vi] > 0.0) {
          const double Lt = sqrt(2.0 * kappa * H / SMB[vi]);
          const double d = (1.0 - zeta[k]) * H;
          T = Ts + sqrt(M_PI) / 2.0 * Lt * grad_b * (erf(d / Lt) - erf(H / Lt));
        } else {
          T = Ts + ((kT0 - kClausiusClapeyron * H) - Ts) * zeta[k];
        }
      } else if (mask[vi] == 2) {
        T = Ts + zeta[k] * ((kT0 - kClausiusClapeyron * H * zeta[nz - 1]) - Ts);
      }
    } else {
      T = Ts;
    }
    Ti[(size_t)k * nV + vi] = fmin(T, kT0 - kClausiusClapeyron * H * zeta[k]);
  }
}

// zeta tridiagonals: calc_vertical_operators_reg_1D (mesh_zeta.f90:287-365)
// with 1-D 2nd-order shape functions (shape_functions.f90:17-79), then the
// tridiagonal extraction of calc_zeta_operators_tridiagonal (:487-605).
void zeta_tridiagonals(int nz, const double *zeta, double *tri /* [6][kHeatMaxNz] */) {
  for (int i = 0; i < 6 * kHeatMaxNz; ++i) tri[i] = 0.0;
  for (int k = 0; k < nz; ++k) {
    int cidx[2];
    if (k == 0) { cidx[0] = 1; cidx[1] = 2; }
    else if (k == nz - 1) { cidx[0] = nz - 3; cidx[1] = nz - 2; }
    else { cidx[0] = k - 1; cidx[1] = k + 1; }
    double dx[2], w2[2], A[4] = {0, 0, 0, 0};
    for (int i = 0; i < 2; ++i) {
      dx[i] = zeta[cidx[i]] - zeta[k];
      double w = 1.0 / std::pow(std::fabs(dx[i]), 1.5);
      w2[i] = w * w;
      A[0] += w2[i] * dx[i] * dx[i];
      A[1] += w2[i] * dx[i] * 0.5 * dx[i] * dx[i];
      A[2] += w2[i] * 0.5 * dx[i] * dx[i] * dx[i];
      A[3] += w2[i] * 0.5 * dx[i] * dx[i] * 0.5 * dx[i] * dx[i];
    }
    double det = A[0] * A[3] - A[1] * A[2];
    double M[4] = {A[3] / det, -A[1] / det, -A[2] / det, A[0] / det};
    double fx[2], fxx[2];
    for (int i = 0; i < 2; ++i) {
      fx[i] = w2[i] * (M[0] * dx[i] + M[1] * 0.5 * dx[i] * dx[i]);
      fxx[i] = w2[i] * (M[2] * dx[i] + M[3] * 0.5 * dx[i] * dx[i]);
    }
    tri[1 * kHeatMaxNz + k] = -(fx[0] + fx[1]);
    tri[4 * kHeatMaxNz + k] = -(fxx[0] + fxx[1]);
    for (int i = 0; i < 2; ++i) {
      if (cidx[i] == k - 1) { tri[0 * kHeatMaxNz + k - 1] = fx[i]; tri[3 * kHeatMaxNz + k - 1] = fxx[i]; }
      if (cidx[i] == k + 1) { tri[2 * kHeatMaxNz + k] = fx[i]; tri[5 * kHeatMaxNz + k] = fxx[i]; }
    }
  }
}

}  // namespace ufe

using namespace ufe;

namespace {
int heat_fail(int code, const std::string &m) {
  set_global_error(m);
  return code;
}
}  // namespace

// Defined in ufe_ctx.hip
struct ufe_ctx;
namespace ufe {
int ctx_device(const ufe_ctx *c);
hipStream_t ctx_stream(const ufe_ctx *c);
void *ctx_heat_buffer(ufe_ctx *c, int i, size_t bytes);
const DevMesh *ctx_mesh(const ufe_ctx *c);
Krylov *ctx_krylov(const ufe_ctx *c);
int ctx_fail(ufe_ctx *c, int code, const std::string &msg);
const std::vector<int> &ctx_vperm(const ufe_ctx *c);
const std::vector<int> &ctx_tperm(const ufe_ctx *c);
}  // namespace ufe

extern "C" int ufe_fraction_gr_b(ufe_ctx *c, const double *Hi, const double *Hb, const double *SL, double *fgr) {
  if (!c || !Hi || !Hb || !SL || !fgr) return heat_fail(UFE_ERR_ARGUMENT, "null argument");
  try {
    UFE_HIP(hipSetDevice(ctx_device(c)));
    const DevMesh &m = *ctx_mesh(c);
    hipStream_t s = ctx_stream(c);
    DBuf dHi(sizeof(double) * m.nV), dHb(sizeof(double) * m.nV), dSL(sizeof(double) * m.nV),
        dF(sizeof(double) * m.nT);
    const std::vector<int> &vp = ctx_vperm(c), &tp = ctx_tperm(c);
    std::vector<double> h1(m.nV), h2(m.nV), h3(m.nV), hf(m.nT);
    for (int i = 0; i < m.nV; ++i) { h1[i] = Hi[vp[i]]; h2[i] = Hb[vp[i]]; h3[i] = SL[vp[i]]; }
    UFE_HIP(hipMemcpyAsync(dHi.p, h1.data(), dHi.bytes, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dHb.p, h2.data(), dHb.bytes, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dSL.p, h3.data(), dSL.bytes, hipMemcpyHostToDevice, s));
    launch_fraction_gr_b(m, dHi.as<double>(), dHb.as<double>(), dSL.as<double>(), dF.as<double>(), s);
    UFE_HIP(hipMemcpyAsync(hf.data(), dF.p, dF.bytes, hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < m.nT; ++i) fgr[tp[i]] = hf[i];
    return UFE_OK;
  } catch (const std::exception &e) {
    return ctx_fail(c, UFE_ERR_DEVICE, e.what());
  }
}

extern "C" int ufe_heat_robin(ufe_ctx *c, const ufe_heat_robin_inputs *in, const int *is_unstable, double *Ti,
                              int *n_replaced) {
  if (!in || !is_unstable || !Ti || !n_replaced) return heat_fail(UFE_ERR_ARGUMENT, "null argument");
  if (in->nz < 2 || in->nV < 0 || !in->zeta || !in->Hi_eff || !in->T_surf || !in->SMB || !in->geothermal_flux ||
      !in->mask)
    return heat_fail(UFE_ERR_ARGUMENT, "invalid Robin inputs");
  const int nV = in->nV, nz = in->nz;
  int n = 0;
  for (int vi = 0; vi < nV; ++vi) n += is_unstable[vi] == 1;
  *n_replaced = 0;
  // thermodynamics_3D_heat_equation.f90:256: more than 1% unstable stops the model
  if (n >= (nV + 99) / 100)
    return heat_fail(UFE_ERR_DIVERGED, "heat equation unstable in " + std::to_string(n) + " of " +
                                           std::to_string(nV) + " columns (at most 1% may be replaced)");
  if (n == 0) return UFE_OK;
  try {
    int dev = c ? ctx_device(c) : 0;
    UFE_HIP(hipSetDevice(dev));
    hipStream_t s;
    if (c) s = ctx_stream(c);
    else UFE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const size_t n3 = sizeof(double) * (size_t)nV * nz, n1 = sizeof(double) * (size_t)nV;
    DBuf dz(sizeof(double) * nz), dH(n1), dTs(n1), dS(n1), dG(n1), dm(sizeof(int) * nV), df(sizeof(int) * nV),
        dT(n3);
    UFE_HIP(hipMemcpyAsync(dz.p, in->zeta, dz.bytes, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dH.p, in->Hi_eff, n1, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dTs.p, in->T_surf, n1, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dS.p, in->SMB, n1, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dG.p, in->geothermal_flux, n1, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dm.p, in->mask, sizeof(int) * nV, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(df.p, is_unstable, sizeof(int) * nV, hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(dT.p, Ti, n3, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_heat_robin, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, nz, dz.as<double>(),
                       dH.as<double>(), dTs.as<double>(), dS.as<double>(), dG.as<double>(), dm.as<int>(),
                       df.as<int>(), in->Hi_min_thermo, dT.as<double>());
    UFE_LAUNCH_CHECK();
    UFE_HIP(hipMemcpyAsync(Ti, dT.p, n3, hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    if (!c) UFE_HIP(hipStreamDestroy(s));
    *n_replaced = n;
    return UFE_OK;
  } catch (const std::exception &e) {
    return heat_fail(UFE_ERR_DEVICE, e.what());
  }
}

// ---------------------------------------------------------------------------
// Scratch memory of the allocation-free heat step
// (ufe_heat_columns_device_noalloc): the zeta tridiagonals, the per-workgroup
// partial counts and the two totals live in one file-scope device array.
// Every GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` only names the symbol: its device address on the
// current GPU is looked up once per device (hipGetSymbolAddress) and only that
// address reaches kernels and copies.  Layout, in doubles:
//   [0, kScratchTri)                       zeta tridiagonals (6 x kHeatMaxNz)
//   [kScratchPart, kScratchPart + 2 * kMaxPartials)   partials of the two counts
//   [kScratchCnt, kScratchCnt + 2)         the totals (unstable, invalid mask)
// Nothing in it has to start at zero: each workgroup writes its partials
// before the one-workgroup pass reads them, and that pass writes both totals.
// The array keeps the last call's contents; calls that overlap on one GPU
// race on it.
// ---------------------------------------------------------------------------
constexpr size_t kScratchTri = 6 * (size_t)kHeatMaxNz;
constexpr size_t kScratchPart = kScratchTri;
constexpr size_t kScratchCnt = kScratchPart + 2 * (size_t)kMaxPartials;
constexpr size_t kScratchDoubles = kScratchCnt + 2;  // 16,578 doubles = 132,624 bytes
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kScratchMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kScratchMaxDevices] = {};

// Device address of `scratch` on device dev (which must be current).
static double *scratch_on_device(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    UFE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the heat-step scratch array");
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  return g_scratch_dev[dev];
}

namespace {
// ufe_heat_columns (host arrays), ufe_heat_columns_device (device arrays) and,
// with use_scratch, ufe_heat_columns_device_noalloc (device arrays, scratch
// in the file-scope array, the null stream, no allocation).
int heat_columns_call(ufe_ctx *c, const ufe_heat_inputs *in, double *Tnew, int *is_unstable, int *n_unstable,
                      bool device, bool use_scratch = false) {
  if (!in || !Tnew || !is_unstable || !n_unstable) return heat_fail(UFE_ERR_ARGUMENT, "null argument");
  if (in->nz < 3 || in->nz > kHeatMaxNz || in->nV < 0)
    return heat_fail(UFE_ERR_ARGUMENT, "unsupported column size: nz must be 3.." + std::to_string(kHeatMaxNz) +
                                           " and nV >= 0");
  const double *fields[] = {in->zeta, in->Ti, in->u_3D, in->v_3D, in->w_3D, in->u_dTdxp_upwind, in->v_dTdyp_upwind,
                            in->dzeta_dx, in->dzeta_dy, in->dzeta_dz, in->dzeta_dt, in->Phi, in->Hi_eff,
                            in->T_surf, in->Q_base_grnd};
  for (const double *f : fields)
    if (!f) return heat_fail(UFE_ERR_ARGUMENT, "null input field");
  if (!in->mask) return heat_fail(UFE_ERR_ARGUMENT, "null mask");
  int nV = in->nV, nz = in->nz;
  bool any_gl = device && in->fraction_gr;
  if (!device) {
    // host arrays: the reference's checks before any device work
    for (int vi = 0; vi < nV; ++vi) {
      if (in->Hi_eff[vi] < in->Hi_min_thermo) continue;
      const int m = in->mask[vi];
      if (m != 1 && m != 2 && m != 3)
        return heat_fail(UFE_ERR_ARGUMENT,
                         "Hi_eff > Hi_min_thermo, but mask_grounded_ice and mask_floating_ice are both false");
      any_gl |= m == 3;
    }
    if (any_gl && !in->fraction_gr)
      return heat_fail(UFE_ERR_ARGUMENT, "grounding-line columns (mask 3) need fraction_gr");
  }
  try {
    int dev = c ? ctx_device(c) : 0;
    if (use_scratch) {
      // the caller's current device; its copy of the scratch array
      UFE_HIP(hipGetDevice(&dev));
      if (dev < 0 || dev >= kScratchMaxDevices)
        return heat_fail(UFE_ERR_DEVICE, "device index beyond the scratch table (64 devices)");
    }
    UFE_HIP(hipSetDevice(dev));
    double *scr = use_scratch ? scratch_on_device(dev) : nullptr;
    // device arrays may come from work on any stream of the caller's
    if (device) UFE_HIP(hipDeviceSynchronize());
    // with a context: its stream and device buffers kept between calls (a
    // time-stepping caller allocates nothing after the first call); without
    // one: a stream and buffers of this call only
    hipStream_t s = nullptr;  // the null stream for the allocation-free form
    if (c) s = ctx_stream(c);
    else if (!use_scratch) UFE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    size_t n3 = sizeof(double) * (size_t)nV * nz, n1 = sizeof(double) * (size_t)nV;
    std::vector<std::unique_ptr<DBuf>> keep;
    int nbuf = 0;
    auto buf = [&](size_t bytes) -> void * {
      if (c) return ctx_heat_buffer(c, nbuf++, bytes);
      keep.push_back(std::make_unique<DBuf>(std::max<size_t>(bytes, 8)));
      return keep.back()->p;
    };
    auto up = [&](const void *h, size_t bytes) {
      void *d = buf(bytes);
      if (bytes) UFE_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
      return d;
    };
    auto in_dev = [&](const void *p, size_t bytes) { return device ? p : (const void *)up(p, bytes); };
    double zeta_h[kHeatMaxNz];
    if (device) UFE_HIP(hipMemcpy(zeta_h, in->zeta, sizeof(double) * nz, hipMemcpyDeviceToHost));
    else std::copy(in->zeta, in->zeta + nz, zeta_h);
    double tri[6 * kHeatMaxNz];
    zeta_tridiagonals(nz, zeta_h, tri);
    HeatDev h;
    h.nV = nV; h.nz = nz;
    h.zeta = (const double *)in_dev(in->zeta, sizeof(double) * nz);
    if (use_scratch) {
      UFE_HIP(hipMemcpyAsync(scr, tri, sizeof(double) * kScratchTri, hipMemcpyHostToDevice, s));
      h.tri = scr;
    } else {
      h.tri = (const double *)up(tri, sizeof(tri));
    }
    h.Ti = (const double *)in_dev(in->Ti, n3);
    h.u = (const double *)in_dev(in->u_3D, n3);
    h.v = (const double *)in_dev(in->v_3D, n3);
    h.w = (const double *)in_dev(in->w_3D, n3);
    h.uadv = (const double *)in_dev(in->u_dTdxp_upwind, n3);
    h.vadv = (const double *)in_dev(in->v_dTdyp_upwind, n3);
    h.zx = (const double *)in_dev(in->dzeta_dx, n3);
    h.zy = (const double *)in_dev(in->dzeta_dy, n3);
    h.zz = (const double *)in_dev(in->dzeta_dz, n3);
    h.zt = (const double *)in_dev(in->dzeta_dt, n3);
    h.Phi = (const double *)in_dev(in->Phi, n3);
    h.Hi_eff = (const double *)in_dev(in->Hi_eff, n1);
    h.T_surf = (const double *)in_dev(in->T_surf, n1);
    h.Q = (const double *)in_dev(in->Q_base_grnd, n1);
    h.mask = (const int *)in_dev(in->mask, sizeof(int) * nV);
    h.fgr = any_gl ? (const double *)in_dev(in->fraction_gr, n1) : nullptr;
    h.dt = in->dt;
    h.Hi_min = in->Hi_min_thermo;
    double *Tout = device ? Tnew : (double *)buf(n3);
    int *unst = device ? is_unstable : (int *)buf(sizeof(int) * std::max(nV, 1));
    double *cnt = use_scratch ? scr + kScratchCnt : (double *)buf(sizeof(double) * 2);
    double *part = use_scratch ? scr + kScratchPart : (double *)buf(sizeof(double) * 2 * kMaxPartials);
    RedSlot slot{part, cnt};
    if (nV > 0) launch_heat_columns(h, Tout, unst, slot, cnt, s);
    double hc[2] = {0.0, 0.0};
    if (nV > 0) {
      if (!device) {
        UFE_HIP(hipMemcpyAsync(Tnew, Tout, n3, hipMemcpyDeviceToHost, s));
        UFE_HIP(hipMemcpyAsync(is_unstable, unst, sizeof(int) * nV, hipMemcpyDeviceToHost, s));
      }
      UFE_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
    }
    UFE_HIP(hipStreamSynchronize(s));
    if (!c && !use_scratch) UFE_HIP(hipStreamDestroy(s));
    if (hc[1] > 0.0)
      return heat_fail(UFE_ERR_ARGUMENT, std::to_string((long)hc[1]) +
                                             " columns with Hi_eff > Hi_min_thermo are neither grounded nor floating "
                                             "(or mask 3 without fraction_gr)");
    *n_unstable = (int)hc[0];
    return UFE_OK;
  } catch (const std::exception &e) {
    return heat_fail(UFE_ERR_DEVICE, e.what());
  }
}
}  // namespace

extern "C" int ufe_heat_columns(ufe_ctx *c, const ufe_heat_inputs *in, double *Tnew, int *is_unstable,
                                int *n_unstable) {
  return heat_columns_call(c, in, Tnew, is_unstable, n_unstable, false);
}

extern "C" int ufe_heat_columns_device(ufe_ctx *c, const ufe_heat_inputs *in, double *Tnew, int *is_unstable,
                                       int *n_unstable) {
  return heat_columns_call(c, in, Tnew, is_unstable, n_unstable, true);
}

extern "C" int ufe_heat_columns_device_noalloc(const ufe_heat_inputs *in, double *Tnew, int *is_unstable,
                                               int *n_unstable) {
  return heat_columns_call(nullptr, in, Tnew, is_unstable, n_unstable, true, true);
}
