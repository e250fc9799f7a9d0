// Inputs of the vertical heat columns that the reference computes before the
// column solve (thermodynamics_utilities.f90): the upwind horizontal
// advection terms (calc_upwind_heat_flux_derivatives :339-434), the strain
// heating (calc_strain_heating :24-91) and the frictional heating
// (calc_frictional_heating :93-128).  Fields keep the reference's layouts
// ((n, nz) with the vertex / triangle index fastest) at the boundary and are
// permuted to the context's locality order on the device.
#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "krylov.h"
#include "ufe_internal.h"

struct ufe_ctx;
namespace ufe {
int ctx_device(const ufe_ctx *c);
hipStream_t ctx_stream(const ufe_ctx *c);
const DevMesh *ctx_mesh(const ufe_ctx *c);
int ctx_fail(ufe_ctx *c, int code, const std::string &msg);
const std::vector<int> &ctx_vperm(const ufe_ctx *c);
const std::vector<int> &ctx_tperm(const ufe_ctx *c);
std::shared_ptr<void> &ctx_heat_mesh(ufe_ctx *c);
const ufe_params &ctx_params(const ufe_ctx *c);
int ctx_guard(ufe_ctx *c, const std::function<int()> &f);

namespace {

struct HeatMesh {
  DBuf vt_ptr, vt_ind;  // triangles around each vertex (device numbering), in the caller's order
};

// Upwind triangle of every vertex and the advection terms of its layers.  A
// vertex whose search finds no triangle sets *fail (the reference stops).
__global__ void __launch_bounds__(kBlock) k_heat_upwind(int nV, int nT, int nz, const int *__restrict__ vt_ptr,
                                                        const int *__restrict__ vt_ind, const int *__restrict__ Tri,
                                                        const double2 *__restrict__ V, const int *__restrict__ ab_ptr,
                                                        const int *__restrict__ ab_ind,
                                                        const double *__restrict__ ab_ddx,
                                                        const double *__restrict__ ab_ddy,
                                                        const double *__restrict__ Ti, const double *__restrict__ Hi,
                                                        const double *__restrict__ uv, const double *__restrict__ vv,
                                                        const double *__restrict__ u3, const double *__restrict__ v3,
                                                        double *__restrict__ udx, double *__restrict__ vdy,
                                                        int *__restrict__ fail) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi >= nV) return;
  if (Hi[vi] < 1.0) {
    for (int k = 0; k < nz; ++k) {
      udx[(size_t)k * nV + vi] = 0.0;
      vdy[(size_t)k * nV + vi] = 0.0;
    }
    return;
  }
  const double ux = -uv[vi], uy = -vv[vi];
  const double2 p = V[vi];
  int tup = -1;
  for (int q = vt_ptr[vi]; q < vt_ptr[vi + 1]; ++q) {
    const int ti = vt_ind[q];
    const int a = Tri[3 * ti], b = Tri[3 * ti + 1], c = Tri[3 * ti + 2];
    // the triangle's other two vertices in counter-clockwise order from vi
    const int vb = a == vi ? b : b == vi ? c : a;
    const int vc = a == vi ? c : b == vi ? a : b;
    const double2 pb = V[vb], pc = V[vc];
    const double abx = pb.x - p.x, aby = pb.y - p.y, acx = pc.x - p.x, acy = pc.y - p.y;
    if (abx * uy - aby * ux >= 0.0 && ux * acy - uy * acx >= 0.0) {
      tup = ti;
      break;
    }
  }
  if (tup < 0) {
    *fail = 1;
    return;
  }
  const int q0 = ab_ptr[tup], q1 = ab_ptr[tup + 1];
  for (int k = 0; k < nz; ++k) {
    double dx = 0.0, dy = 0.0;
    for (int q = q0; q < q1; ++q) {
      const double t = Ti[(size_t)k * nV + ab_ind[q]];
      dx += ab_ddx[q] * t;
      dy += ab_ddy[q] * t;
    }
    udx[(size_t)k * nV + vi] = u3[(size_t)k * nT + tup] * dx;
    vdy[(size_t)k * nV + vi] = v3[(size_t)k * nT + tup] * dy;
  }
}

struct StrainRates {
  const double *e[9];
};

__global__ void __launch_bounds__(kBlock) k_strain_heating(long n, int nV, const double *__restrict__ Hi,
                                                           const double *__restrict__ A, StrainRates s, double glen_n,
                                                           double *__restrict__ Phi) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (Hi[i % nV] < 0.1) {
    Phi[i] = 0.0;
    return;
  }
  const double ux = s.e[0][i], uy = s.e[1][i], uz = s.e[2][i], vx = s.e[3][i], vy = s.e[4][i], vz = s.e[5][i];
  const double wx = s.e[6][i], wy = s.e[7][i], wz = s.e[8][i];
  const double D = sqrt(0.5 * (ux * ux + vy * vy + wz * wz + 0.5 * (uy + vx) * (uy + vx) +
                               0.5 * (uz + wx) * (uz + wx) + 0.5 * (vz + wy) * (vz + wy)));
  Phi[i] = 2.0 * pow(A[i], -1.0 / glen_n) * pow(D, 1.0 / glen_n + 1.0);
}

__global__ void __launch_bounds__(kBlock) k_frictional_heating(int nV, int no_sliding, const double *__restrict__ beta,
                                                               const double *__restrict__ ub,
                                                               const int *__restrict__ grounded,
                                                               double *__restrict__ Q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nV) Q[i] = (!no_sliding && grounded[i]) ? beta[i] * ub[i] : 0.0;
}

// caller layout [layer][original index] -> device buffer [layer][device index]
void up_layers(DBuf &d, const double *h, int n, int layers, const std::vector<int> &perm, hipStream_t s) {
  std::vector<double> t((size_t)n * layers);
  for (int k = 0; k < layers; ++k)
    for (int i = 0; i < n; ++i) t[(size_t)k * n + i] = h[(size_t)k * n + perm[i]];
  d.alloc(t.size() * sizeof(double));
  UFE_HIP(hipMemcpyAsync(d.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, s));
  UFE_HIP(hipStreamSynchronize(s));
}

void down_layers(const DBuf &d, double *h, int n, int layers, const std::vector<int> &perm, hipStream_t s) {
  std::vector<double> t((size_t)n * layers);
  UFE_HIP(hipMemcpyAsync(t.data(), d.p, t.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  UFE_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < layers; ++k)
    for (int i = 0; i < n; ++i) h[(size_t)k * n + perm[i]] = t[(size_t)k * n + i];
}

}  // namespace
}  // namespace ufe

using namespace ufe;

extern "C" int ufe_heat_setup(ufe_ctx *c, const int *iTri_ptr, const int *iTri_ind) {
  if (!c || !iTri_ptr || !iTri_ind) return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    UFE_HIP(hipSetDevice(ctx_device(c)));
    const DevMesh &m = *ctx_mesh(c);
    const int nV = m.nV, nT = m.nT;
    const std::vector<int> &vperm = ctx_vperm(c), &tperm = ctx_tperm(c);
    if (iTri_ptr[0] != 0) return ctx_fail(c, UFE_ERR_ARGUMENT, "iTri_ptr[0] must be 0");
    std::vector<int> t_o2n(nT), ptr(nV + 1, 0), ind;
    for (int i = 0; i < nT; ++i) t_o2n[tperm[i]] = i;
    std::vector<int> Tri(3 * (size_t)nT);
    UFE_HIP(hipMemcpy(Tri.data(), m.Tri, Tri.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < nV; ++i) {
      const int vo = vperm[i];
      if (iTri_ptr[vo + 1] < iTri_ptr[vo]) return ctx_fail(c, UFE_ERR_ARGUMENT, "iTri_ptr not monotone");
      for (int q = iTri_ptr[vo]; q < iTri_ptr[vo + 1]; ++q) {
        const int to = iTri_ind[q];
        if (to < 0 || to >= nT) return ctx_fail(c, UFE_ERR_ARGUMENT, "iTri entry out of range");
        const int t = t_o2n[to];
        if (Tri[3 * t] != i && Tri[3 * t + 1] != i && Tri[3 * t + 2] != i)
          return ctx_fail(c, UFE_ERR_ARGUMENT, "iTri lists a triangle that does not contain the vertex");
        ind.push_back(t);
      }
      ptr[i + 1] = (int)ind.size();
    }
    auto h = std::make_shared<HeatMesh>();
    h->vt_ptr.alloc(ptr.size() * sizeof(int));
    h->vt_ind.alloc(std::max<size_t>(ind.size(), 1) * sizeof(int));
    UFE_HIP(hipMemcpy(h->vt_ptr.p, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!ind.empty()) UFE_HIP(hipMemcpy(h->vt_ind.p, ind.data(), ind.size() * sizeof(int), hipMemcpyHostToDevice));
    ctx_heat_mesh(c) = h;
    return UFE_OK;
  });
}

extern "C" int ufe_heat_upwind_derivatives(ufe_ctx *c, const double *Ti, const double *Hi, const double *u_vav,
                                           const double *v_vav, const double *u_3D_b, const double *v_3D_b,
                                           double *u_times_dTdxp_upwind, double *v_times_dTdyp_upwind) {
  if (!c || !Ti || !Hi || !u_vav || !v_vav || !u_3D_b || !v_3D_b || !u_times_dTdxp_upwind || !v_times_dTdyp_upwind)
    return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    UFE_HIP(hipSetDevice(ctx_device(c)));
    auto *h = static_cast<HeatMesh *>(ctx_heat_mesh(c).get());
    if (!h) return ctx_fail(c, UFE_ERR_ARGUMENT, "ufe_heat_setup has not been called on this context");
    const DevMesh &m = *ctx_mesh(c);
    const int nV = m.nV, nT = m.nT, nz = m.nz;
    hipStream_t s = ctx_stream(c);
    const std::vector<int> &vperm = ctx_vperm(c), &tperm = ctx_tperm(c);
    DBuf dTi, dHi, du, dv, du3, dv3, dox, doy, dfail(sizeof(int));
    up_layers(dTi, Ti, nV, nz, vperm, s);
    up_layers(dHi, Hi, nV, 1, vperm, s);
    up_layers(du, u_vav, nV, 1, vperm, s);
    up_layers(dv, v_vav, nV, 1, vperm, s);
    up_layers(du3, u_3D_b, nT, nz, tperm, s);
    up_layers(dv3, v_3D_b, nT, nz, tperm, s);
    dox.alloc(sizeof(double) * (size_t)nV * nz);
    doy.alloc(sizeof(double) * (size_t)nV * nz);
    UFE_HIP(hipMemsetAsync(dfail.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_heat_upwind, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV, nT, nz, h->vt_ptr.as<int>(),
                       h->vt_ind.as<int>(), m.Tri, reinterpret_cast<const double2 *>(m.V), m.ab_ptr, m.ab_ind,
                       m.ab_ddx, m.ab_ddy, dTi.as<double>(), dHi.as<double>(), du.as<double>(), dv.as<double>(),
                       du3.as<double>(), dv3.as<double>(), dox.as<double>(), doy.as<double>(), dfail.as<int>());
    UFE_LAUNCH_CHECK();
    int failed = 0;
    UFE_HIP(hipMemcpyAsync(&failed, dfail.p, sizeof(int), hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    if (failed) return ctx_fail(c, UFE_ERR_MESH, "could not find upwind triangle!");
    down_layers(dox, u_times_dTdxp_upwind, nV, nz, vperm, s);
    down_layers(doy, v_times_dTdyp_upwind, nV, nz, vperm, s);
    return UFE_OK;
  });
}

extern "C" int ufe_heat_strain_heating(ufe_ctx *c, const double *Hi, const double *A_flow,
                                       const double *const *strain_rates, double *Phi) {
  if (!c || !Hi || !A_flow || !strain_rates || !Phi) return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  for (int q = 0; q < 9; ++q)
    if (!strain_rates[q]) return ctx_fail(c, UFE_ERR_ARGUMENT, "null strain-rate field");
  return ctx_guard(c, [&]() -> int {
    UFE_HIP(hipSetDevice(ctx_device(c)));
    const DevMesh &m = *ctx_mesh(c);
    const int nV = m.nV, nz = m.nz;
    hipStream_t s = ctx_stream(c);
    const std::vector<int> &vperm = ctx_vperm(c);
    DBuf dHi, dA, de[9], dPhi;
    up_layers(dHi, Hi, nV, 1, vperm, s);
    up_layers(dA, A_flow, nV, nz, vperm, s);
    StrainRates sr;
    for (int q = 0; q < 9; ++q) {
      up_layers(de[q], strain_rates[q], nV, nz, vperm, s);
      sr.e[q] = de[q].as<double>();
    }
    dPhi.alloc(sizeof(double) * (size_t)nV * nz);
    const long n = (long)nV * nz;
    hipLaunchKernelGGL(k_strain_heating, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, nV, dHi.as<double>(),
                       dA.as<double>(), sr, ctx_params(c).glens_n, dPhi.as<double>());
    UFE_LAUNCH_CHECK();
    down_layers(dPhi, Phi, nV, nz, vperm, s);
    return UFE_OK;
  });
}

extern "C" int ufe_heat_frictional_heating(ufe_ctx *c, const double *basal_friction_coefficient,
                                           const double *uabs_base, const int *mask_grounded_ice,
                                           double *frictional_heating) {
  if (!c || !basal_friction_coefficient || !uabs_base || !mask_grounded_ice || !frictional_heating)
    return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    UFE_HIP(hipSetDevice(ctx_device(c)));
    const int nV = ctx_mesh(c)->nV;
    hipStream_t s = ctx_stream(c);
    const std::vector<int> &vperm = ctx_vperm(c);
    DBuf db, du, dQ, dm(sizeof(int) * (size_t)nV);
    up_layers(db, basal_friction_coefficient, nV, 1, vperm, s);
    up_layers(du, uabs_base, nV, 1, vperm, s);
    std::vector<int> mk(nV);
    for (int i = 0; i < nV; ++i) mk[i] = mask_grounded_ice[vperm[i]] ? 1 : 0;
    UFE_HIP(hipMemcpyAsync(dm.p, mk.data(), mk.size() * sizeof(int), hipMemcpyHostToDevice, s));
    dQ.alloc(sizeof(double) * (size_t)nV);
    hipLaunchKernelGGL(k_frictional_heating, dim3(blocks_for(nV)), dim3(kBlock), 0, s, nV,
                       ctx_params(c).sliding_law == UFE_SLIDING_NO_SLIDING ? 1 : 0, db.as<double>(), du.as<double>(),
                       dm.as<int>(), dQ.as<double>());
    UFE_LAUNCH_CHECK();
    down_layers(dQ, frictional_heating, nV, 1, vperm, s);
    return UFE_OK;
  });
}
