// Mesh matrix operators (host, setup-time, once per mesh).
//
// Builds the operators the velocity solve reads from UFEMISM's mesh type:
//   a->b  map/ddx/ddy   (mesh_disc_calc_matrix_operators_2D.f90:198-335)
//   b->a  map/ddx/ddy   (mesh_disc_calc_matrix_operators_2D.f90:337-474)
//   b->b  2nd-order ddx/ddy/d2dx2/d2dxdy/d2dy2
//                       (mesh_disc_calc_matrix_operators_2D.f90:612-764)
//   a->a  ddx/ddy       (mesh_disc_calc_matrix_operators_2D.f90:61-196)
// using weighted least-squares shape functions with distance weights
// w = 1/|d|^q, q = 1.5 (shape_functions.f90:13, :218-364 and :366-451).
//
// Neighbourhoods grow by whole flood-fill rings, exactly like
// extend_group_single_iteration_a/_b (mesh_utilities.f90:1856-1935): every
// member of the current group adds its not-yet-listed neighbours, in order.
// Each row lists the home element first (b->b only) and then the group in the
// order it was found, as the reference does.  The operators of one family share
// one sparsity pattern, so a family is returned as one (ptr, ind) pair and one
// value array per operator.
//
// All indices are 0-based.  The entry points follow a two-call convention:
// call with null output arrays to get the sizes, then again with arrays of at
// least those sizes.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/ufe_hip.h"

namespace {

constexpr double kQ = 1.5;

// Gauss-Jordan inverse with partial pivoting of an n x n row-major matrix.
// Returns false if a pivot is exactly zero (the reference's test is
// |det| < tiny(det), i.e. an exactly singular system).
bool invert(int n, double* a, double* inv) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) inv[i * n + j] = (i == j) ? 1.0 : 0.0;
  double det = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = std::fabs(a[c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      double v = std::fabs(a[r * n + c]);
      if (v > best) { best = v; p = r; }
    }
    if (best == 0.0) return false;
    if (p != c) {
      for (int j = 0; j < n; ++j) {
        std::swap(a[c * n + j], a[p * n + j]);
        std::swap(inv[c * n + j], inv[p * n + j]);
      }
      det = -det;
    }
    double piv = a[c * n + c];
    det *= piv;
    double rp = 1.0 / piv;
    for (int j = 0; j < n; ++j) { a[c * n + j] *= rp; inv[c * n + j] *= rp; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) { a[r * n + j] -= f * a[c * n + j]; inv[r * n + j] -= f * inv[c * n + j]; }
    }
  }
  return std::fabs(det) >= 2.2250738585072014e-308;
}

// calc_shape_functions_2D_stag_1st_order (shape_functions.f90:366-451)
bool shape_stag_1st(double x, double y, int nc, const double* xc, const double* yc,
                    double* Nf, double* Nfx, double* Nfy) {
  double A[9] = {0}, M[9];
  std::vector<double> w2(nc), dx(nc), dy(nc);
  for (int i = 0; i < nc; ++i) {
    dx[i] = xc[i] - x; dy[i] = yc[i] - y;
    double w = 1.0 / std::pow(std::sqrt(dx[i] * dx[i] + dy[i] * dy[i]), kQ);
    w2[i] = w * w;
    double b[3] = {1.0, dx[i], dy[i]};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A[r * 3 + c] += w2[i] * b[r] * b[c];
  }
  if (!invert(3, A, M)) return false;
  for (int i = 0; i < nc; ++i) {
    Nf[i]  = w2[i] * (M[0] + M[1] * dx[i] + M[2] * dy[i]);
    Nfx[i] = w2[i] * (M[3] + M[4] * dx[i] + M[5] * dy[i]);
    Nfy[i] = w2[i] * (M[6] + M[7] * dx[i] + M[8] * dy[i]);
  }
  return true;
}

// calc_shape_functions_2D_reg_1st_order (shape_functions.f90:140-216):
// d/dx, d/dy of a first-order fit through the home point; home coefficients
// are minus the sums of the neighbours'.
bool shape_reg_1st(double x, double y, int nc, const double* xc, const double* yc, double* Nfx_i,
                   double* Nfy_i, double* Nfx, double* Nfy) {
  double A[4] = {0, 0, 0, 0};
  std::vector<double> w2(nc), dx(nc), dy(nc);
  for (int i = 0; i < nc; ++i) {
    dx[i] = xc[i] - x; dy[i] = yc[i] - y;
    double w = 1.0 / std::pow(std::sqrt(dx[i] * dx[i] + dy[i] * dy[i]), kQ);
    w2[i] = w * w;
    A[0] += w2[i] * dx[i] * dx[i]; A[1] += w2[i] * dx[i] * dy[i];
    A[2] += w2[i] * dy[i] * dx[i]; A[3] += w2[i] * dy[i] * dy[i];
  }
  double det = A[0] * A[3] - A[1] * A[2];
  if (std::fabs(det) < 2.2250738585072014e-308) return false;
  double M[4] = {A[3] / det, -A[1] / det, -A[2] / det, A[0] / det};
  double sx = 0.0, sy = 0.0;
  for (int i = 0; i < nc; ++i) {
    Nfx[i] = w2[i] * (M[0] * dx[i] + M[1] * dy[i]);
    Nfy[i] = w2[i] * (M[2] * dx[i] + M[3] * dy[i]);
    sx += Nfx[i]; sy += Nfy[i];
  }
  *Nfx_i = -sx; *Nfy_i = -sy;
  return true;
}

// calc_shape_functions_2D_reg_2nd_order (shape_functions.f90:218-364)
// out[k][i], k = x, y, xx, xy, yy; home coefficients out_i[k] = -sum_i out[k][i]
bool shape_reg_2nd(double x, double y, int nc, const double* xc, const double* yc,
                   double* out_i, double* const* out) {
  double A[25] = {0}, M[25];
  std::vector<double> w2(nc), bas(5 * nc);
  for (int i = 0; i < nc; ++i) {
    double dx = xc[i] - x, dy = yc[i] - y;
    double w = 1.0 / std::pow(std::sqrt(dx * dx + dy * dy), kQ);
    w2[i] = w * w;
    double* b = &bas[5 * i];
    b[0] = dx; b[1] = dy; b[2] = 0.5 * dx * dx; b[3] = dx * dy; b[4] = 0.5 * dy * dy;
    for (int r = 0; r < 5; ++r)
      for (int c = 0; c < 5; ++c) A[r * 5 + c] += w2[i] * b[r] * b[c];
  }
  if (!invert(5, A, M)) return false;
  for (int k = 0; k < 5; ++k) {
    double s = 0.0;
    for (int i = 0; i < nc; ++i) {
      const double* b = &bas[5 * i];
      double v = w2[i] * (M[k * 5 + 0] * b[0] + M[k * 5 + 1] * b[1] + M[k * 5 + 2] * b[2] +
                          M[k * 5 + 3] * b[3] + M[k * 5 + 4] * b[4]);
      out[k][i] = v;
      s += v;
    }
    out_i[k] = -s;
  }
  return true;
}

// One flood-fill ring over triangles (extend_group_single_iteration_b).
void extend_b(const int* TriC, std::vector<int>& map, std::vector<int>& stack) {
  size_t n = stack.size();
  for (size_t i = 0; i < n; ++i) {
    int ti = stack[i];
    for (int k = 0; k < 3; ++k) {
      int tj = TriC[3 * ti + k];
      if (tj < 0 || map[tj]) continue;
      map[tj] = 1;
      stack.push_back(tj);
    }
  }
}

// One flood-fill ring over vertices (extend_group_single_iteration_a).
void extend_a(const int* C_ptr, const int* C_ind, std::vector<int>& map, std::vector<int>& stack) {
  size_t n = stack.size();
  for (size_t i = 0; i < n; ++i) {
    int vi = stack[i];
    for (int c = C_ptr[vi]; c < C_ptr[vi + 1]; ++c) {
      int vj = C_ind[c];
      if (map[vj]) continue;
      map[vj] = 1;
      stack.push_back(vj);
    }
  }
}

void clear(std::vector<int>& map, const std::vector<int>& stack) {
  for (int v : stack) map[v] = 0;
}

}  // namespace

extern "C" int ufe_mesh_operators_ab(int nV, int nT, const double* V, const int* Tri,
                                     const double* TriGC, const int* C_ptr, const int* C_ind,
                                     int* nnz_out, int* ptr, int* ind,
                                     double* val_map, double* val_ddx, double* val_ddy) {
  std::vector<int> map(nV, 0), stack;
  std::vector<double> xc, yc, f, fx, fy;
  int nnz = 0;
  if (ptr) ptr[0] = 0;
  for (int ti = 0; ti < nT; ++ti) {
    clear(map, stack);
    stack.clear();
    for (int n = 0; n < 3; ++n) { int vi = Tri[3 * ti + n]; map[vi] = 1; stack.push_back(vi); }
    while (true) {
      int nc = (int)stack.size();
      xc.resize(nc); yc.resize(nc); f.resize(nc); fx.resize(nc); fy.resize(nc);
      for (int i = 0; i < nc; ++i) { xc[i] = V[2 * stack[i]]; yc[i] = V[2 * stack[i] + 1]; }
      if (shape_stag_1st(TriGC[2 * ti], TriGC[2 * ti + 1], nc, xc.data(), yc.data(), f.data(), fx.data(), fy.data())) break;
      size_t before = stack.size();
      extend_a(C_ptr, C_ind, map, stack);
      if (stack.size() == before) return -1;
    }
    int nc = (int)stack.size();
    if (ind) {
      for (int i = 0; i < nc; ++i) {
        ind[nnz + i] = stack[i];
        val_map[nnz + i] = f[i]; val_ddx[nnz + i] = fx[i]; val_ddy[nnz + i] = fy[i];
      }
    }
    nnz += nc;
    if (ptr) ptr[ti + 1] = nnz;
  }
  *nnz_out = nnz;
  return 0;
}

extern "C" int ufe_mesh_operators_ba(int nV, int nT, const double* V, const int* TriC,
                                     const double* TriGC, const int* iTri_ptr, const int* iTri_ind,
                                     int* nnz_out, int* ptr, int* ind,
                                     double* val_map, double* val_ddx, double* val_ddy) {
  const int n_min = 3;
  std::vector<int> map(nT, 0), stack;
  std::vector<double> xc, yc, f, fx, fy;
  int nnz = 0;
  if (ptr) ptr[0] = 0;
  for (int vi = 0; vi < nV; ++vi) {
    clear(map, stack);
    stack.clear();
    for (int k = iTri_ptr[vi]; k < iTri_ptr[vi + 1]; ++k) { int ti = iTri_ind[k]; map[ti] = 1; stack.push_back(ti); }
    while ((int)stack.size() < n_min) {
      size_t before = stack.size();
      extend_b(TriC, map, stack);
      if (stack.size() == before) return -1;
    }
    while (true) {
      int nc = (int)stack.size();
      xc.resize(nc); yc.resize(nc); f.resize(nc); fx.resize(nc); fy.resize(nc);
      for (int i = 0; i < nc; ++i) { xc[i] = TriGC[2 * stack[i]]; yc[i] = TriGC[2 * stack[i] + 1]; }
      if (shape_stag_1st(V[2 * vi], V[2 * vi + 1], nc, xc.data(), yc.data(), f.data(), fx.data(), fy.data())) break;
      size_t before = stack.size();
      extend_b(TriC, map, stack);
      if (stack.size() == before) return -1;
    }
    int nc = (int)stack.size();
    if (ind) {
      for (int i = 0; i < nc; ++i) {
        ind[nnz + i] = stack[i];
        val_map[nnz + i] = f[i]; val_ddx[nnz + i] = fx[i]; val_ddy[nnz + i] = fy[i];
      }
    }
    nnz += nc;
    if (ptr) ptr[vi + 1] = nnz;
  }
  *nnz_out = nnz;
  return 0;
}

extern "C" int ufe_mesh_operators_bb2(int nT, const int* TriC, const double* TriGC,
                                      int* nnz_out, int* ptr, int* ind,
                                      double* val_ddx, double* val_ddy, double* val_d2dx2,
                                      double* val_d2dxdy, double* val_d2dy2) {
  const int n_min = 5;
  std::vector<int> map(nT, 0), stack;
  std::vector<double> xc, yc, c0, c1, c2, c3, c4;
  int nnz = 0;
  if (ptr) ptr[0] = 0;
  for (int ti = 0; ti < nT; ++ti) {
    clear(map, stack);
    stack.clear();
    map[ti] = 1;
    stack.push_back(ti);
    while ((int)stack.size() - 1 < n_min) {
      size_t before = stack.size();
      extend_b(TriC, map, stack);
      if (stack.size() == before) return -1;
    }
    double ci[5];
    while (true) {
      int nc = (int)stack.size() - 1;
      xc.resize(nc); yc.resize(nc);
      c0.resize(nc); c1.resize(nc); c2.resize(nc); c3.resize(nc); c4.resize(nc);
      for (int i = 0; i < nc; ++i) { xc[i] = TriGC[2 * stack[i + 1]]; yc[i] = TriGC[2 * stack[i + 1] + 1]; }
      double* outs[5] = {c0.data(), c1.data(), c2.data(), c3.data(), c4.data()};
      if (shape_reg_2nd(TriGC[2 * ti], TriGC[2 * ti + 1], nc, xc.data(), yc.data(), ci, outs)) break;
      size_t before = stack.size();
      extend_b(TriC, map, stack);
      if (stack.size() == before) return -1;
    }
    int nc = (int)stack.size() - 1;
    if (ind) {
      ind[nnz] = ti;
      val_ddx[nnz] = ci[0]; val_ddy[nnz] = ci[1]; val_d2dx2[nnz] = ci[2];
      val_d2dxdy[nnz] = ci[3]; val_d2dy2[nnz] = ci[4];
      for (int i = 0; i < nc; ++i) {
        int k = nnz + 1 + i;
        ind[k] = stack[i + 1];
        val_ddx[k] = c0[i]; val_ddy[k] = c1[i]; val_d2dx2[k] = c2[i];
        val_d2dxdy[k] = c3[i]; val_d2dy2[k] = c4[i];
      }
    }
    nnz += nc + 1;
    if (ptr) ptr[ti + 1] = nnz;
  }
  *nnz_out = nnz;
  return 0;
}

// a->a ddx/ddy (calc_matrix_operators_mesh_a_a,
// mesh_disc_calc_matrix_operators_2D.f90:61-196): the home vertex first, then
// its flood-fill neighbourhood (at least two neighbours, grown while the fit
// is singular).
extern "C" int ufe_mesh_operators_aa(int nV, const double* V, const int* C_ptr, const int* C_ind,
                                     int* nnz_out, int* ptr, int* ind, double* val_ddx, double* val_ddy) {
  const int n_min = 2;
  std::vector<int> map(nV, 0), stack;
  std::vector<double> xc, yc, fx, fy;
  int nnz = 0;
  if (ptr) ptr[0] = 0;
  for (int vi = 0; vi < nV; ++vi) {
    clear(map, stack);
    stack.clear();
    map[vi] = 1;
    stack.push_back(vi);
    while ((int)stack.size() - 1 < n_min) {
      size_t before = stack.size();
      extend_a(C_ptr, C_ind, map, stack);
      if (stack.size() == before) return -1;
    }
    double fxi = 0.0, fyi = 0.0;
    while (true) {
      int nc = (int)stack.size() - 1;
      xc.resize(nc); yc.resize(nc); fx.resize(nc); fy.resize(nc);
      for (int i = 0; i < nc; ++i) { xc[i] = V[2 * stack[i + 1]]; yc[i] = V[2 * stack[i + 1] + 1]; }
      if (shape_reg_1st(V[2 * vi], V[2 * vi + 1], nc, xc.data(), yc.data(), &fxi, &fyi, fx.data(), fy.data())) break;
      size_t before = stack.size();
      extend_a(C_ptr, C_ind, map, stack);
      if (stack.size() == before) return -1;
    }
    int nc = (int)stack.size() - 1;
    if (ind) {
      ind[nnz] = vi;
      val_ddx[nnz] = fxi; val_ddy[nnz] = fyi;
      for (int i = 0; i < nc; ++i) {
        ind[nnz + 1 + i] = stack[i + 1];
        val_ddx[nnz + 1 + i] = fx[i]; val_ddy[nnz + 1 + i] = fy[i];
      }
    }
    nnz += nc + 1;
    if (ptr) ptr[vi + 1] = nnz;
  }
  *nnz_out = nnz;
  return 0;
}
