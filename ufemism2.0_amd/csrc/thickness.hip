// Semi-implicit ice-thickness step on the device: calc_dHi_dt_semiimplicit
// (conservation_of_mass_semiimplicit.f90:23-159), the reference's second
// PETSc call site, with the explicit step it calls for the time-step limit
// and the border values (conservation_of_mass_explicit.f90) and the upwind
// flux-divergence matrix (conservation_of_mass_utilities.f90:17-110).
//
// Layout.  Vertices are in the context's locality (Morton) order.  Row vi of
// the flux-divergence matrix M has its diagonal first and then one entry per
// connection of vi in the caller's neighbour order; the per-connection
// geometry (neighbour, Voronoi boundary length, unit vector to the neighbour,
// the two triangles either side of the edge) is stored once per mesh.  For
// the linear solve the scalar system is folded into 2x2 blocks of
// consecutive vertices (2i, 2i+1), so the engine's block Krylov solvers and
// block-Jacobi preconditioner apply unchanged; the block pattern and the
// position of every scalar entry in it are fixed per mesh, and each step
// only scatters values.
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
std::shared_ptr<void> &ctx_thickness(ufe_ctx *c);
const double *ctx_velocity(const ufe_ctx *c);
int ctx_guard(ufe_ctx *c, const std::function<int()> &f);

namespace {

struct ThicknessState {
  int nV = 0, nC = 0, nb = 0, nnzb = 0;
  std::vector<int> v_o2n;        // caller vertex -> device vertex
  // per vertex (device order): row start of M (diagonal first), border index, cell area
  DBuf row_ptr, vbi, area;
  // per connection (device order, rows contiguous): neighbour, Voronoi boundary
  // length, unit vector to the neighbour, triangles either side (-1 = none)
  DBuf c_ind, c_cw, c_ex, c_ey, c_tri;
  // scalar entries (nV + nC): value of M and position in the block matrix
  DBuf m_val, e_blk;
  // block system
  DBuf b_ptr, b_col, b_val, b_dinv, b_diag;
  // per-step vectors (device order)
  DBuf f_in, i_in, rate, hex, hs0, hs1, nint, rhs, x, out3;
  DBuf part, scal;
  std::unique_ptr<Krylov> kry;
};

constexpr double kRhoI = 910.0, kRhoSW = 1028.0;  // parameters.f90:23,25

__device__ __forceinline__ double surface_elev(double Hi, double Hb, double SL) {
  return Hi + fmax(SL - kRhoI / kRhoSW * Hi, Hb);
}
__device__ __forceinline__ double thickness_from_surface(double Hb, double Hs, double SL) {
  const double Hi_float = fmax(0.0, (SL - Hb) * (kRhoSW / kRhoI));
  if (Hs > Hb + Hi_float) return Hs - Hb;
  return fmin(Hi_float, (Hs - SL) / (1.0 - (kRhoI / kRhoSW)));
}

// per-vertex input fields, packed [field][vertex]
enum { F_HI, F_HB, F_SL, F_SMB, F_BMB, F_LMB, F_FM, F_TGT, F_BCH, F_GUESS, F_COUNT };
enum { I_NOICE, I_BCMASK, I_COUNT };

struct ThkDev {
  int nV;
  const int *row_ptr, *vbi, *c_ind, *c_tri;
  const double *area, *c_cw, *c_ex, *c_ey;
  const double *f;   // [F_COUNT][nV]
  const int *iv;     // [I_COUNT][nV]
  const double2 *uv; // triangle velocities
};

// Row vi of M (calc_ice_flux_divergence_matrix_upwind): edge velocity = mean
// of the triangles either side (map_velocities_from_b_to_c_2D); u_perp its
// component along the connection; diagonal sum_c Cw max(0, u_perp) / A_i
// where the vertex is fully ice-filled, off-diagonal Cw min(0, u_perp) / A_i
// where the neighbour is.  divQ = M H; rate = -divQ + fm (SMB + BMB - target)
// + LMB; the flux-limited time step of calc_flux_limited_timestep as written
// (a thinning vertex's limit is Hi / max(rate, 1e-9)), min-reduced.
__global__ void __launch_bounds__(kBlock) k_thk_rows(ThkDev d, double dt_ice_max, double Hi_min,
                                                     double *__restrict__ m_val, double *__restrict__ divQ,
                                                     double *__restrict__ rate, RedSlot slot) {
  const int nV = d.nV;
  double lim_min = dt_ice_max;
  for (int vi = blockIdx.x * blockDim.x + threadIdx.x; vi < nV; vi += gridDim.x * blockDim.x) {
    const double A_i = d.area[vi];
    const bool full_i = d.f[F_FM * nV + vi] >= 1.0;
    const int k0 = d.row_ptr[vi], k1 = d.row_ptr[vi + 1];
    double diag = 0.0;
    for (int k = k0 + 1; k < k1; ++k) {
      const int c = k - vi - 1;  // connection index: rows hold one more entry (the diagonal) each
      const int vj = d.c_ind[c];
      const int tl = d.c_tri[2 * c], tr = d.c_tri[2 * c + 1];
      double uc, vc;
      if (tl < 0) { uc = d.uv[tr].x; vc = d.uv[tr].y; }
      else if (tr < 0) { uc = d.uv[tl].x; vc = d.uv[tl].y; }
      else { uc = (d.uv[tl].x + d.uv[tr].x) / 2.0; vc = (d.uv[tl].y + d.uv[tr].y) / 2.0; }
      const double u_perp = uc * d.c_ex[c] + vc * d.c_ey[c];
      const double L_c = d.c_cw[c];
      if (full_i) diag = diag + L_c * fmax(0.0, u_perp) / A_i;
      m_val[k] = d.f[F_FM * nV + vj] >= 1.0 ? L_c * fmin(0.0, u_perp) / A_i : 0.0;
    }
    m_val[k0] = diag;
    double s = diag * d.f[F_HI * nV + vi];
    for (int k = k0 + 1; k < k1; ++k) s += m_val[k] * d.f[F_HI * nV + d.c_ind[k - vi - 1]];
    divQ[vi] = s;
    const double m = d.f[F_FM * nV + vi] * (d.f[F_SMB * nV + vi] + d.f[F_BMB * nV + vi] - d.f[F_TGT * nV + vi]) +
                     d.f[F_LMB * nV + vi];
    const double r = -s + m;
    rate[vi] = r;
    const double Hi = d.f[F_HI * nV + vi];
    if (Hi > Hi_min && r < 0.0) lim_min = fmin(lim_min, Hi / fmax(r, 1e-9));
  }
  __shared__ double lds[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double v = lim_min;
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) v = fmin(v, lds[w]);
    slot.partials[blockIdx.x] = v;
  }
}

// dt_max = max(dt_ice_min, min over vertices); out[0] = dt_max, out[1] = min(dt, dt_max)
__global__ void __launch_bounds__(kBlock) k_thk_dtmax(RedSlot slot, int nparts, double dt, double dt_ice_min,
                                                      double *__restrict__ out) {
  double v = 1e300;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) v = fmin(v, slot.partials[b]);
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  __shared__ double lds[kBlock / 64];
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) v = fmin(v, lds[w]);
    const double dt_max = fmax(dt_ice_min, v);
    out[0] = dt_max;
    out[1] = fmin(dt, dt_max);
  }
}

// explicit thickness, its surface elevation and the count of interior
// (non-border, ice-allowed) neighbours
__global__ void __launch_bounds__(kBlock) k_thk_explicit(ThkDev d, const double *__restrict__ rate,
                                                         const double *__restrict__ dts, double *__restrict__ hex,
                                                         double *__restrict__ hs, int *__restrict__ nint) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  const int nV = d.nV;
  if (vi >= nV) return;
  const double H = fmax(0.0, d.f[F_HI * nV + vi] + rate[vi] * dts[1]);
  hex[vi] = H;
  hs[vi] = surface_elev(H, d.f[F_HB * nV + vi], d.f[F_SL * nV + vi]);
  int n = 0;
  for (int k = d.row_ptr[vi] + 1; k < d.row_ptr[vi + 1]; ++k) {
    const int vj = d.c_ind[k - vi - 1];
    if (d.vbi[vj] == 0 && !d.iv[I_NOICE * nV + vj]) ++n;
  }
  nint[vi] = n;
}

// One pass of apply_ice_thickness_BC_explicit over the border vertices:
// 'zero' sides get no ice; 'infinite' sides the mean surface elevation of the
// interior neighbours (pass 0, vertices with interior neighbours) or of all
// neighbours (pass 1, the others), read from the elevations as they were at
// the start of the pass (hs_in), written to hs_out.
__global__ void __launch_bounds__(kBlock) k_thk_bc_pass(ThkDev d, int pass, int4 bcH, const int *__restrict__ nint,
                                                        const double *__restrict__ hs_in, double *__restrict__ hs_out,
                                                        double *__restrict__ hex) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  const int nV = d.nV;
  if (vi >= nV) return;
  double hs = hs_in[vi];
  const int b = d.vbi[vi];
  if (b > 0) {
    const int side = (b - 1) / 2;
    const int choice = side == 0 ? bcH.x : side == 1 ? bcH.y : side == 2 ? bcH.z : bcH.w;
    if (choice == UFE_BC_ZERO) {
      hex[vi] = 0.0;
    } else if ((pass == 0) == (nint[vi] > 0)) {
      double sum = 0.0;
      int n = 0;
      for (int k = d.row_ptr[vi] + 1; k < d.row_ptr[vi + 1]; ++k) {
        const int vj = d.c_ind[k - vi - 1];
        if (pass == 1 || (d.vbi[vj] == 0 && !d.iv[I_NOICE * nV + vj])) { sum += hs_in[vj]; ++n; }
      }
      const double Hb = d.f[F_HB * nV + vi];
      hs = fmax(Hb, sum / (double)(pass == 0 ? nint[vi] : n));
      hex[vi] = thickness_from_surface(Hb, hs, d.f[F_SL * nV + vi]);
    }
  }
  hs_out[vi] = hs;
}

// prescribed thickness and the no-ice mask on the explicit thickness; surface
// elevations recomputed for the second application of the border conditions
__global__ void __launch_bounds__(kBlock) k_thk_masks(ThkDev d, double *__restrict__ hex, double *__restrict__ hs) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  const int nV = d.nV;
  if (vi >= nV) return;
  double H = hex[vi];
  if (d.iv[I_BCMASK * nV + vi] == 1) H = fmax(0.0, d.f[F_BCH * nV + vi]);
  if (d.iv[I_NOICE * nV + vi]) H = 0.0;
  hex[vi] = H;
  hs[vi] = surface_elev(H, d.f[F_HB * nV + vi], d.f[F_SL * nV + vi]);
}

__global__ void __launch_bounds__(kBlock) k_thk_zero_blocks(long n, double4 *__restrict__ v) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) v[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// Stiffness matrix I + dt fs M and load vector H - dt (1 - fs) divQ +
// max(-H, dt m) (:112-140); border, prescribed and no-ice rows become
// identity rows with the explicit thickness, the prescribed thickness and 0
// (:165-315).  Values are scattered into the 2x2 blocks of vertex pairs; the
// initial guess is H + dt dHi_dt_guess (or the fixed value).
__global__ void __launch_bounds__(kBlock) k_thk_assemble(ThkDev d, double dt, double fs,
                                                         const double *__restrict__ m_val,
                                                         const double *__restrict__ divQ,
                                                         const double *__restrict__ hex, const int *__restrict__ e_blk,
                                                         double *__restrict__ bval, double *__restrict__ rhs,
                                                         double *__restrict__ x) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  const int nV = d.nV;
  if (vi >= nV) return;
  const double H = d.f[F_HI * nV + vi];
  int fixed = 0;
  double val = 0.0;
  if (d.vbi[vi] > 0) { fixed = 1; val = hex[vi]; }
  if (d.iv[I_BCMASK * nV + vi] == 1) { fixed = 1; val = d.f[F_BCH * nV + vi]; }
  if (d.iv[I_NOICE * nV + vi]) { fixed = 1; val = 0.0; }
  const int k0 = d.row_ptr[vi], k1 = d.row_ptr[vi + 1];
  for (int k = k0; k < k1; ++k) {
    double a = fixed ? 0.0 : m_val[k] * dt * fs;
    if (k == k0) a = fixed ? 1.0 : a + 1.0;
    bval[e_blk[k]] = a;  // e_blk: scalar position inside the block array (4 per block)
  }
  if (fixed) {
    rhs[vi] = val;
    x[vi] = val;
  } else {
    const double m = d.f[F_FM * nV + vi] * (d.f[F_SMB * nV + vi] + d.f[F_BMB * nV + vi] - d.f[F_TGT * nV + vi]) +
                     d.f[F_LMB * nV + vi];
    rhs[vi] = H - (dt * (1.0 - fs) * divQ[vi]) + fmax(-1.0 * H, dt * m);
    x[vi] = H + dt * d.f[F_GUESS * nV + vi];
  }
}

// padding unknown of an odd vertex count: identity row, zero value
__global__ void k_thk_pad(int nV, const int *__restrict__ b_diag, double4 *__restrict__ bval,
                          double *__restrict__ rhs, double *__restrict__ x) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || (nV & 1) == 0) return;
  const int bi = nV / 2;
  double4 v = bval[b_diag[bi]];
  v.w = 1.0;
  bval[b_diag[bi]] = v;
  rhs[nV] = 0.0;
  x[nV] = 0.0;
}

__global__ void __launch_bounds__(kBlock) k_thk_dinv(int nb, const int *__restrict__ b_diag,
                                                     const double4 *__restrict__ bval, double4 *__restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double4 a = bval[b_diag[i]];
  const double det = a.x * a.w - a.y * a.z;
  dinv[i] = det != 0.0 ? make_double4(a.w / det, -a.y / det, -a.z / det, a.x / det) : make_double4(1.0, 0.0, 0.0, 1.0);
}

// dHi_dt = (H(t+dt) - H) / dt
__global__ void __launch_bounds__(kBlock) k_thk_rate(int nV, const double *__restrict__ f, const double *__restrict__ x,
                                                     double dt, double *__restrict__ out) {
  const int vi = blockIdx.x * blockDim.x + threadIdx.x;
  if (vi < nV) out[vi] = (x[vi] - f[F_HI * nV + vi]) / dt;
}

ThkDev thk_dev(const ThicknessState &t, const double2 *uv) {
  ThkDev d;
  d.nV = t.nV;
  d.row_ptr = t.row_ptr.as<int>(); d.vbi = t.vbi.as<int>(); d.c_ind = t.c_ind.as<int>(); d.c_tri = t.c_tri.as<int>();
  d.area = t.area.as<double>(); d.c_cw = t.c_cw.as<double>(); d.c_ex = t.c_ex.as<double>();
  d.c_ey = t.c_ey.as<double>();
  d.f = t.f_in.as<double>(); d.iv = t.i_in.as<int>();
  d.uv = uv;
  return d;
}

template <typename U>
void upload(DBuf &b, const std::vector<U> &h) {
  b.alloc(std::max<size_t>(h.size(), 1) * sizeof(U));
  if (!h.empty()) UFE_HIP(hipMemcpy(b.p, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
}

}  // namespace
}  // namespace ufe

using namespace ufe;

extern "C" int ufe_thickness_setup(ufe_ctx *c, const ufe_vertex_geometry *g) {
  if (!c || !g || !g->C_ptr || !g->C_ind || !g->VBI || !g->A || !g->Cw || !g->ETri)
    return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  return ctx_guard(c, [&]() -> int {
    const DevMesh &m = *ctx_mesh(c);
    const int nV = m.nV, nT = m.nT, nC = g->nC;
    const std::vector<int> &vperm = ctx_vperm(c), &tperm = ctx_tperm(c);
    if (g->C_ptr[0] != 0 || g->C_ptr[nV] != nC) return ctx_fail(c, UFE_ERR_ARGUMENT, "bad connection pointer");
    for (int vi = 0; vi < nV; ++vi) {
      if (g->C_ptr[vi + 1] < g->C_ptr[vi]) return ctx_fail(c, UFE_ERR_ARGUMENT, "connection pointer not monotone");
      if (g->VBI[vi] < 0 || g->VBI[vi] > 8) return ctx_fail(c, UFE_ERR_ARGUMENT, "VBI out of range (0..8)");
      if (!(g->A[vi] > 0.0)) return ctx_fail(c, UFE_ERR_ARGUMENT, "Voronoi cell areas must be positive");
    }
    for (int k = 0; k < nC; ++k) {
      if (g->C_ind[k] < 0 || g->C_ind[k] >= nV) return ctx_fail(c, UFE_ERR_ARGUMENT, "neighbour out of range");
      const int tl = g->ETri[2 * k], tr = g->ETri[2 * k + 1];
      if (tl < -1 || tl >= nT || tr < -1 || tr >= nT || (tl < 0 && tr < 0))
        return ctx_fail(c, UFE_ERR_ARGUMENT, "edge triangles out of range");
    }
    auto t = std::make_shared<ThicknessState>();
    t->nV = nV;
    t->nC = nC;
    t->v_o2n.assign(nV, 0);
    std::vector<int> t_o2n(nT);
    for (int i = 0; i < nV; ++i) t->v_o2n[vperm[i]] = i;
    for (int i = 0; i < nT; ++i) t_o2n[tperm[i]] = i;
    // vertex coordinates (device order), for the unit vectors
    std::vector<double> Vd(2 * (size_t)nV);
    UFE_HIP(hipMemcpy(Vd.data(), m.V, Vd.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> row_ptr(nV + 1), vbi(nV), c_ind(nC), c_tri(2 * (size_t)nC);
    std::vector<double> area(nV), c_cw(nC), c_ex(nC), c_ey(nC);
    int kc = 0;
    row_ptr[0] = 0;
    for (int i = 0; i < nV; ++i) {
      const int vo = vperm[i];
      vbi[i] = g->VBI[vo];
      area[i] = g->A[vo];
      for (int k = g->C_ptr[vo]; k < g->C_ptr[vo + 1]; ++k, ++kc) {
        const int j = t->v_o2n[g->C_ind[k]];
        c_ind[kc] = j;
        c_cw[kc] = g->Cw[k];
        const double dx = Vd[2 * j] - Vd[2 * i], dy = Vd[2 * j + 1] - Vd[2 * i + 1];
        const double D = std::sqrt(dx * dx + dy * dy);
        if (!(D > 0.0)) return ctx_fail(c, UFE_ERR_ARGUMENT, "coincident connected vertices");
        c_ex[kc] = dx / D;
        c_ey[kc] = dy / D;
        for (int s = 0; s < 2; ++s) {
          const int to = g->ETri[2 * k + s];
          c_tri[2 * kc + s] = to < 0 ? -1 : t_o2n[to];
        }
      }
      row_ptr[i + 1] = row_ptr[i] + 1 + (g->C_ptr[vo + 1] - g->C_ptr[vo]);
    }
    // 2x2 blocks of the vertex pairs (2i, 2i+1); pattern = union of the pair's columns
    const int nb = (nV + 1) / 2;
    std::vector<int> b_ptr(nb + 1, 0), b_col, b_diag(nb), e_blk((size_t)nV + nC);
    std::vector<int> cols;
    for (int bi = 0; bi < nb; ++bi) {
      cols.clear();
      for (int r = 2 * bi; r < std::min(nV, 2 * bi + 2); ++r) {
        cols.push_back(r / 2);
        for (int k = row_ptr[r] + 1; k < row_ptr[r + 1]; ++k) cols.push_back(c_ind[k - r - 1] / 2);
      }
      if (cols.empty()) cols.push_back(bi);
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      const int base = (int)b_col.size();
      for (int cb : cols) {
        if (cb == bi) b_diag[bi] = (int)b_col.size();
        b_col.push_back(cb);
      }
      b_ptr[bi + 1] = (int)b_col.size();
      for (int r = 2 * bi; r < std::min(nV, 2 * bi + 2); ++r)
        for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
          const int col = k == row_ptr[r] ? r : c_ind[k - r - 1];
          const int e = base + (int)(std::lower_bound(cols.begin(), cols.end(), col / 2) - cols.begin());
          e_blk[k] = 4 * e + 2 * (r & 1) + (col & 1);
        }
    }
    t->nb = nb;
    t->nnzb = (int)b_col.size();
    upload(t->row_ptr, row_ptr); upload(t->vbi, vbi); upload(t->area, area);
    upload(t->c_ind, c_ind); upload(t->c_cw, c_cw); upload(t->c_ex, c_ex); upload(t->c_ey, c_ey);
    upload(t->c_tri, c_tri);
    upload(t->b_ptr, b_ptr); upload(t->b_col, b_col); upload(t->b_diag, b_diag); upload(t->e_blk, e_blk);
    t->m_val.alloc(sizeof(double) * ((size_t)nV + nC));
    t->b_val.alloc(sizeof(double4) * (size_t)t->nnzb);
    t->b_dinv.alloc(sizeof(double4) * (size_t)nb);
    t->f_in.alloc(sizeof(double) * F_COUNT * (size_t)nV);
    t->i_in.alloc(sizeof(int) * I_COUNT * (size_t)nV);
    for (DBuf *b : {&t->rate, &t->hex, &t->hs0, &t->hs1, &t->out3}) b->alloc(sizeof(double) * 3 * (size_t)nV);
    t->nint.alloc(sizeof(int) * (size_t)nV);
    t->rhs.alloc(sizeof(double) * 2 * (size_t)nb);
    t->x.alloc(sizeof(double) * 2 * (size_t)nb);
    t->part.alloc(sizeof(double) * kMaxPartials);
    t->scal.alloc(sizeof(double) * 4);
    t->kry = std::make_unique<Krylov>(nb, ctx_stream(c));
    t->kry->set_precond(UFE_PC_BLOCK_JACOBI);
    t->kry->set_diag(t->b_diag.as<int>());
    ctx_thickness(c) = t;
    return UFE_OK;
  });
}

extern "C" int ufe_thickness_semiimplicit(ufe_ctx *c, const ufe_thickness_params *p, const ufe_thickness_inputs *in,
                                          double *Hi_next, double *dHi_dt, double *divQ, double *dt_max,
                                          int *n_its) {
  if (!c || !p || !in || !Hi_next) return ctx_fail(c, UFE_ERR_ARGUMENT, "null argument");
  const double *fields[] = {in->Hi, in->Hb, in->SL, in->SMB, in->BMB, in->LMB, in->fraction_margin,
                            in->dHi_dt_target};
  for (const double *f : fields)
    if (!f) return ctx_fail(c, UFE_ERR_ARGUMENT, "null input field");
  if (!in->mask_noice) return ctx_fail(c, UFE_ERR_ARGUMENT, "null mask_noice");
  if ((!in->BC_prescr_mask) != (!in->BC_prescr_Hi))
    return ctx_fail(c, UFE_ERR_ARGUMENT, "need to provide prescribed both Hi and mask!");
  if ((!in->u_vav_b) != (!in->v_vav_b)) return ctx_fail(c, UFE_ERR_ARGUMENT, "need both velocity components or neither");
  if (!(in->dt > 0.0)) return ctx_fail(c, UFE_ERR_ARGUMENT, "dt must be positive");
  if (p->krylov != UFE_KRYLOV_BICGSTAB && p->krylov != UFE_KRYLOV_GMRES)
    return ctx_fail(c, UFE_ERR_ARGUMENT, "krylov must be UFE_KRYLOV_BICGSTAB or UFE_KRYLOV_GMRES");
  if (p->krylov == UFE_KRYLOV_GMRES && (p->gmres_restart < 1 || p->gmres_restart > 64))
    return ctx_fail(c, UFE_ERR_ARGUMENT, "gmres_restart must be 1..64");
  for (int s = 0; s < 4; ++s)
    if (p->bc_H[s] != UFE_BC_ZERO && p->bc_H[s] != UFE_BC_INFINITE)
      return ctx_fail(c, UFE_ERR_ARGUMENT, "bc_H must be UFE_BC_ZERO or UFE_BC_INFINITE");
  return ctx_guard(c, [&]() -> int {
    UFE_HIP(hipSetDevice(ctx_device(c)));
    auto *t = static_cast<ThicknessState *>(ctx_thickness(c).get());
    if (!t) return ctx_fail(c, UFE_ERR_ARGUMENT, "ufe_thickness_setup has not been called on this context");
    const DevMesh &m = *ctx_mesh(c);
    const int nV = t->nV, nT = m.nT;
    hipStream_t s = ctx_stream(c);
    const std::vector<int> &vperm = ctx_vperm(c), &tperm = ctx_tperm(c);
    // inputs in device order
    std::vector<double> f((size_t)F_COUNT * nV, 0.0);
    std::vector<int> iv((size_t)I_COUNT * nV, 0);
    const double *src[F_COUNT] = {in->Hi, in->Hb, in->SL, in->SMB, in->BMB, in->LMB, in->fraction_margin,
                                  in->dHi_dt_target, in->BC_prescr_Hi, in->dHi_dt_guess};
    for (int q = 0; q < F_COUNT; ++q)
      if (src[q])
        for (int i = 0; i < nV; ++i) f[(size_t)q * nV + i] = src[q][vperm[i]];
    for (int i = 0; i < nV; ++i) {
      iv[i] = in->mask_noice[vperm[i]] ? 1 : 0;
      iv[(size_t)nV + i] = in->BC_prescr_mask ? in->BC_prescr_mask[vperm[i]] : 0;
    }
    UFE_HIP(hipMemcpyAsync(t->f_in.p, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, s));
    UFE_HIP(hipMemcpyAsync(t->i_in.p, iv.data(), iv.size() * sizeof(int), hipMemcpyHostToDevice, s));
    const double2 *uv = reinterpret_cast<const double2 *>(ctx_velocity(c));
    DBuf uvbuf;
    std::vector<double> uvh;
    if (in->u_vav_b) {
      uvh.resize(2 * (size_t)nT);
      for (int i = 0; i < nT; ++i) {
        uvh[2 * i] = in->u_vav_b[tperm[i]];
        uvh[2 * i + 1] = in->v_vav_b[tperm[i]];
      }
      uvbuf.alloc(uvh.size() * sizeof(double));
      UFE_HIP(hipMemcpyAsync(uvbuf.p, uvh.data(), uvh.size() * sizeof(double), hipMemcpyHostToDevice, s));
      uv = uvbuf.as<double2>();
    }
    const ThkDev d = thk_dev(*t, uv);
    double *divq = t->out3.as<double>(), *rate = t->rate.as<double>(), *hex = t->hex.as<double>();
    double *hs0 = t->hs0.as<double>(), *hs1 = t->hs1.as<double>(), *dts = t->scal.as<double>();
    const int grid = std::min(blocks_for(nV), kMaxPartials);
    RedSlot sl{t->part.as<double>(), dts};
    hipLaunchKernelGGL(k_thk_rows, dim3(grid), dim3(kBlock), 0, s, d, p->dt_ice_max, p->Hi_min,
                       t->m_val.as<double>(), divq, rate, sl);
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_thk_dtmax, dim3(1), dim3(kBlock), 0, s, sl, grid, in->dt, p->dt_ice_min, dts);
    UFE_LAUNCH_CHECK();
    const dim3 gv(blocks_for(nV));
    const int4 bcH = make_int4(p->bc_H[0], p->bc_H[1], p->bc_H[2], p->bc_H[3]);
    hipLaunchKernelGGL(k_thk_explicit, gv, dim3(kBlock), 0, s, d, rate, dts, hex, hs0, t->nint.as<int>());
    UFE_LAUNCH_CHECK();
    // calc_dHi_dt_explicit's border conditions, prescribed thickness, no-ice
    // mask; then the border conditions once more (the matrix boundary
    // conditions apply them to the explicit thickness again)
    for (int rep = 0; rep < 2; ++rep) {
      hipLaunchKernelGGL(k_thk_bc_pass, gv, dim3(kBlock), 0, s, d, 0, bcH, t->nint.as<int>(), hs0, hs1, hex);
      UFE_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_thk_bc_pass, gv, dim3(kBlock), 0, s, d, 1, bcH, t->nint.as<int>(), hs1, hs0, hex);
      UFE_LAUNCH_CHECK();
      if (rep == 0) {
        hipLaunchKernelGGL(k_thk_masks, gv, dim3(kBlock), 0, s, d, hex, hs0);
        UFE_LAUNCH_CHECK();
      }
    }
    hipLaunchKernelGGL(k_thk_zero_blocks, dim3(blocks_for(t->nnzb)), dim3(kBlock), 0, s, (long)t->nnzb,
                       t->b_val.as<double4>());
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_thk_assemble, gv, dim3(kBlock), 0, s, d, in->dt, p->fs, t->m_val.as<double>(), divq, hex,
                       t->e_blk.as<int>(), t->b_val.as<double>(), t->rhs.as<double>(), t->x.as<double>());
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_thk_pad, dim3(1), dim3(64), 0, s, nV, t->b_diag.as<int>(), t->b_val.as<double4>(),
                       t->rhs.as<double>(), t->x.as<double>());
    UFE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_thk_dinv, dim3(blocks_for(t->nb)), dim3(kBlock), 0, s, t->nb, t->b_diag.as<int>(),
                       t->b_val.as<double4>(), t->b_dinv.as<double4>());
    UFE_LAUNCH_CHECK();
    BsrT<double> A;
    A.nb = t->nb; A.nnzb = t->nnzb; A.ptr = t->b_ptr.as<int>(); A.col = t->b_col.as<int>();
    A.val = t->b_val.as<double4>(); A.dinv = t->b_dinv.as<double4>();
    SolveStats st = p->krylov == UFE_KRYLOV_GMRES
                        ? t->kry->gmres64(A, t->rhs.as<double>(), t->x.as<double>(), p->rtol, p->abstol, p->maxit,
                                          false, p->gmres_restart, 8)
                        : t->kry->bicgstab64(A, t->rhs.as<double>(), t->x.as<double>(), p->rtol, p->abstol,
                                             p->maxit, false, 8);
    hipLaunchKernelGGL(k_thk_rate, gv, dim3(kBlock), 0, s, nV, t->f_in.as<double>(), t->x.as<double>(), in->dt,
                       rate);
    UFE_LAUNCH_CHECK();
    std::vector<double> h((size_t)nV), hs2(2);
    auto down = [&](const double *dsrc, double *dst) {
      UFE_HIP(hipMemcpyAsync(h.data(), dsrc, sizeof(double) * nV, hipMemcpyDeviceToHost, s));
      UFE_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < nV; ++i) dst[vperm[i]] = h[i];
    };
    down(t->x.as<double>(), Hi_next);
    if (dHi_dt) down(rate, dHi_dt);
    if (divQ) down(divq, divQ);
    UFE_HIP(hipMemcpyAsync(hs2.data(), dts, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    UFE_HIP(hipStreamSynchronize(s));
    if (dt_max) *dt_max = hs2[0];
    if (n_its) *n_its = st.its;
    return st.status == 1 ? UFE_OK : UFE_WARN_NOT_CONVERGED;
  });
}
