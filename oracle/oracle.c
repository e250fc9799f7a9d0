/*
 * oracle.c -- FP64 CPU restatement of UFEMISM's velocity solve and vertical
 * heat columns.  TEST INFRASTRUCTURE ONLY: the parity tests, smoke() and the
 * cpu_baseline leg of bench.py call it as the checker.  The product never
 * links, loads or calls anything in this directory.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the UFEMISM2.0 checkout).  The Krylov solve restates PETSc's default KSP on
 * one rank (GMRES(30), left preconditioning with ILU(0), zero initial guess,
 * convergence on the preconditioned residual norm against
 * max(rtol*||r0||, abstol)); PETSc is a third-party dependency of the
 * reference whose version is not pinned (Homebrew latest in CI), so this is a
 * restatement of its published algorithm, not of its code.
 *
 * Pinning (tests/test_oracle_*.py): CSR product against the ut_petsc 6x6 known
 * answer; Thomas solver against the ut_math_utilities known answer; GMRES
 * against scipy's direct solve; operators against polynomial exactness;
 * assembly against the operator expression of the SSA; Picard against the
 * Schoof (2006) ice-stream closed form; the Tsai2015, Schoof2005 and
 * Zoet-Iverson laws against the pinned Weertman, Coulomb and Budd laws (their
 * limits and Schoof's harmonic identity); heat columns against the steady
 * diffusion closed form and dense solves; grounded fractions against the
 * geometric closed form and Monte-Carlo areas.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/ufe_hip.h"

#define ICE_DENSITY 910.0
#define SEAWATER_DENSITY 1028.0
#define GRAV 9.81
#define T0 273.16
#define CLAUSIUS_CLAPEYRON 8.7e-4

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------------------------------------------------------------------- */
/* CSR helpers                                                             */
/* ---------------------------------------------------------------------- */

/* multiply_CSR_matrix_with_vector_1D (CSR_matrix_vector_multiplication.f90:198-334) */
/* Loops shorter than OMP_MIN_N run on one thread: a parallel region costs
 * more than such a loop, most of all on a loaded machine. */
#define OMP_MIN_N 32768

void oracle_csr_matvec(int m, const int *ptr, const int *ind, const double *val, const double *x,
                       double *y) {
#pragma omp parallel for schedule(static) if (m >= OMP_MIN_N)
  for (int i = 0; i < m; ++i) {
    double s = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[ind[k]];
    y[i] = s;
  }
}

static double dot(int n, const double *a, const double *b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (n >= OMP_MIN_N)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* ---------------------------------------------------------------------- */
/* Thomas algorithm: solve_tridiagonal_matrix_equation                     */
/* (src/UPSY/basic/math_utilities/tridiagonal_solver.f90:13-51)           */
/* a[1..n-1] sub-diagonal (a[0] unused), b diagonal, c[0..n-2] super.     */
/* a, b, c, d are overwritten like the reference's intent(inout).          */
/* ---------------------------------------------------------------------- */
void oracle_tridiag(int n, double *a, double *b, double *c, double *d, double *x) {
  for (int i = 1; i < n; ++i) {
    double w = a[i] / b[i - 1];
    b[i] -= w * c[i - 1];
    d[i] -= w * d[i - 1];
  }
  x[n - 1] = d[n - 1] / b[n - 1];
  for (int i = n - 2; i >= 0; --i) x[i] = (d[i] - c[i] * x[i + 1]) / b[i];
}

/* ---------------------------------------------------------------------- */
/* ILU(0) + GMRES(m), left preconditioned: PETSc's default KSPSolve on one */
/* rank as called by solve_matrix_equation_PETSc (petsc_basic.f90:66-141).*/
/* ---------------------------------------------------------------------- */
typedef struct {
  int n;
  int *ptr, *ind, *diag;
  double *val;  /* factors L (unit, strictly lower) and U in place */
} ilu0_t;

static int cmp_int_pair(const void *a, const void *b) {
  const int *x = a, *y = b;
  return (x[0] > y[0]) - (x[0] < y[0]);
}

/* Copy A with sorted column indices and factor ILU(0) (IKJ variant). */
static int ilu0_factor(int n, const int *ptr, const int *ind, const double *val, ilu0_t *f) {
  int nnz = ptr[n];
  f->n = n;
  f->ptr = malloc(sizeof(int) * (n + 1));
  f->ind = malloc(sizeof(int) * nnz);
  f->diag = malloc(sizeof(int) * n);
  f->val = malloc(sizeof(double) * nnz);
  int *pair = malloc(sizeof(int) * 2 * 64);
  int cap = 64;
  memcpy(f->ptr, ptr, sizeof(int) * (n + 1));
  for (int i = 0; i < n; ++i) {
    int k0 = ptr[i], len = ptr[i + 1] - k0;
    if (len > cap) { cap = 2 * len; pair = realloc(pair, sizeof(int) * 2 * cap); }
    for (int k = 0; k < len; ++k) { pair[2 * k] = ind[k0 + k]; pair[2 * k + 1] = k0 + k; }
    qsort(pair, len, 2 * sizeof(int), cmp_int_pair);
    f->diag[i] = -1;
    for (int k = 0; k < len; ++k) {
      f->ind[k0 + k] = pair[2 * k];
      f->val[k0 + k] = val[pair[2 * k + 1]];
      if (pair[2 * k] == i) f->diag[i] = k0 + k;
      if (k > 0 && pair[2 * k] == pair[2 * k - 2]) { free(pair); return -1; } /* duplicate */
    }
    if (f->diag[i] < 0) { free(pair); return -2; }
  }
  free(pair);
  int *pos = malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) pos[i] = -1;
  for (int i = 0; i < n; ++i) {
    for (int k = f->ptr[i]; k < f->ptr[i + 1]; ++k) pos[f->ind[k]] = k;
    for (int k = f->ptr[i]; k < f->diag[i]; ++k) {
      int j = f->ind[k];
      double piv = f->val[f->diag[j]];
      if (piv == 0.0) { free(pos); return -3; }
      double lij = f->val[k] / piv;
      f->val[k] = lij;
      for (int kk = f->diag[j] + 1; kk < f->ptr[j + 1]; ++kk) {
        int p = pos[f->ind[kk]];
        if (p >= 0) f->val[p] -= lij * f->val[kk];
      }
    }
    for (int k = f->ptr[i]; k < f->ptr[i + 1]; ++k) pos[f->ind[k]] = -1;
    if (f->val[f->diag[i]] == 0.0) { free(pos); return -3; }
  }
  free(pos);
  return 0;
}

static void ilu0_apply(const ilu0_t *f, const double *r, double *z) {
  int n = f->n;
  for (int i = 0; i < n; ++i) {
    double s = r[i];
    for (int k = f->ptr[i]; k < f->diag[i]; ++k) s -= f->val[k] * z[f->ind[k]];
    z[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = z[i];
    for (int k = f->diag[i] + 1; k < f->ptr[i + 1]; ++k) s -= f->val[k] * z[f->ind[k]];
    z[i] = s / f->val[f->diag[i]];
  }
}

static void ilu0_free(ilu0_t *f) {
  free(f->ptr); free(f->ind); free(f->diag); free(f->val);
}

/* GMRES(restart) with modified Gram-Schmidt and Givens rotations, left
 * preconditioned with ILU(0).  x is overwritten with the solution; when
 * zero_guess is set it is zeroed first (PETSc's default initial guess).
 * Returns the iteration count (>= 0) or a negative error. */
int oracle_gmres_ilu0(int n, const int *ptr, const int *ind, const double *val, const double *b,
                      double *x, double rtol, double abstol, int maxit, int restart, int zero_guess,
                      double *final_pres) {
  ilu0_t f;
  int rc = ilu0_factor(n, ptr, ind, val, &f);
  if (rc) return -10 + rc;
  int m = restart;
  double *V = malloc(sizeof(double) * (size_t)n * (m + 1));
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
  double *g = malloc(sizeof(double) * (m + 1));
  double *w = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
  if (zero_guess) memset(x, 0, sizeof(double) * n);
  int its = 0;
  double rnorm0 = -1.0, ttol = 0.0, rnorm = 0.0;
  int converged = 0;
  while (!converged && its < maxit) {
    /* r = M^-1 (b - A x) */
    oracle_csr_matvec(n, ptr, ind, val, x, t);
    for (int i = 0; i < n; ++i) t[i] = b[i] - t[i];
    ilu0_apply(&f, t, V);
    rnorm = sqrt(dot(n, V, V));
    if (rnorm0 < 0.0) {
      rnorm0 = rnorm;
      ttol = fmax(rtol * rnorm0, abstol);
    }
    if (rnorm <= ttol) { converged = 1; break; }
    for (int i = 0; i < n; ++i) V[i] /= rnorm;
    memset(g, 0, sizeof(double) * (m + 1));
    g[0] = rnorm;
    int k;
    for (k = 0; k < m && its < maxit; ++k) {
      ++its;
      double *vk = V + (size_t)k * n, *vk1 = V + (size_t)(k + 1) * n;
      oracle_csr_matvec(n, ptr, ind, val, vk, t);
      ilu0_apply(&f, t, w);
      for (int j = 0; j <= k; ++j) {
        double *vj = V + (size_t)j * n;
        double h = dot(n, w, vj);
        H[j * m + k] = h;
#pragma omp parallel for schedule(static) if (n >= OMP_MIN_N)
        for (int i = 0; i < n; ++i) w[i] -= h * vj[i];
      }
      double hk1 = sqrt(dot(n, w, w));
      H[(k + 1) * m + k] = hk1;
      if (hk1 != 0.0)
        for (int i = 0; i < n; ++i) vk1[i] = w[i] / hk1;
      for (int j = 0; j < k; ++j) {
        double a = H[j * m + k], bb = H[(j + 1) * m + k];
        H[j * m + k] = cs[j] * a + sn[j] * bb;
        H[(j + 1) * m + k] = -sn[j] * a + cs[j] * bb;
      }
      double a = H[k * m + k], bb = H[(k + 1) * m + k];
      double r = hypot(a, bb);
      cs[k] = (r == 0.0) ? 1.0 : a / r;
      sn[k] = (r == 0.0) ? 0.0 : bb / r;
      H[k * m + k] = r;
      H[(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      rnorm = fabs(g[k + 1]);
      if (rnorm <= ttol || hk1 == 0.0) { ++k; converged = (rnorm <= ttol); break; }
    }
    /* back substitution and update x += V y */
    int kk = k;
    double *y = calloc(kk > 0 ? kk : 1, sizeof(double));
    for (int i = kk - 1; i >= 0; --i) {
      double s = g[i];
      for (int j = i + 1; j < kk; ++j) s -= H[i * m + j] * y[j];
      y[i] = s / H[i * m + i];
    }
    for (int j = 0; j < kk; ++j) {
      double *vj = V + (size_t)j * n;
#pragma omp parallel for schedule(static) if (n >= OMP_MIN_N)
      for (int i = 0; i < n; ++i) x[i] += y[j] * vj[i];
    }
    free(y);
  }
  if (final_pres) *final_pres = rnorm;
  free(V); free(H); free(cs); free(sn); free(g); free(w); free(t);
  ilu0_free(&f);
  return its;
}

/* ---------------------------------------------------------------------- */
/* Mesh-operator applications                                              */
/* ---------------------------------------------------------------------- */

/* calc_driving_stress (SSA_DIVA_utilities.f90:21-57) */
static void driving_stress(const ufe_mesh_desc *m, const double *Hi, const double *Hs, double *tdx,
                           double *tdy) {
  for (int ti = 0; ti < m->nT; ++ti) {
    double hb = 0, dx = 0, dy = 0;
    for (int k = m->ab_ptr[ti]; k < m->ab_ptr[ti + 1]; ++k) {
      int vi = m->ab_ind[k];
      hb += m->ab_map[k] * Hi[vi];
      dx += m->ab_ddx[k] * Hs[vi];
      dy += m->ab_ddy[k] * Hs[vi];
    }
    tdx[ti] = -ICE_DENSITY * GRAV * hb * dx;
    tdy[ti] = -ICE_DENSITY * GRAV * hb * dy;
  }
}

static void apply_csr(int rows, const int *ptr, const int *ind, const double *val, const double *x,
                      double *y) {
  oracle_csr_matvec(rows, ptr, ind, val, x, y);
}

/* vertical_average (mesh_zeta.f90:257-283) */
static double vertical_average(int nz, const double *zeta, const double *f, int stride) {
  double s = 0.0;
  for (int k = 0; k < nz - 1; ++k) s += 0.5 * (f[(k + 1) * stride] + f[k * stride]) * (zeta[k + 1] - zeta[k]);
  return s;
}

/* calc_effective_viscosity_Glen_2D (constitutive_equation.f90:25-51) */
static double glen_2d(double eps0, double ux, double uy, double vx, double vy, double A, double n) {
  double e2 = ux * ux + vy * vy + ux * vy + 0.25 * (uy + vx) * (uy + vx) + eps0;
  return 0.5 * pow(A, -1.0 / n) * pow(e2, (1.0 - n) / (2.0 * n));
}

/* calc_effective_viscosity_Glen_3D_uv_only (constitutive_equation.f90:53-80) */
static double glen_3d(double eps0, double ux, double uy, double uz, double vx, double vy, double vz,
                      double A, double n) {
  double e2 = ux * ux + vy * vy + ux * vy + 0.25 * (uy + vx) * (uy + vx) + 0.25 * (uz * uz + vz * vz) + eps0;
  return 0.5 * pow(A, -1.0 / n) * pow(e2, (1.0 - n) / (2.0 * n));
}

/* Coulomb limit alpha_sq * effective_pressure of the Tsai2015 and Schoof2005
 * laws (bed_roughness%alpha_sq, ice%effective_pressure), caller-order [nV]
 * arrays that the caller keeps alive; NULL clears. */
static const double *g_alpha_sq = NULL, *g_effective_pressure = NULL;

void oracle_set_sliding_cap(const double *alpha_sq, const double *effective_pressure) {
  g_alpha_sq = alpha_sq;
  g_effective_pressure = effective_pressure;
}

/* calc_basal_friction_coefficient + the sliding laws it dispatches to
 * (sliding_laws.f90:25-81 Weertman :86-130, Budd :200-242, Coulomb :132-198,
 * Tsai2015 :244-297, Schoof2005 :299-351, Zoet-Iverson :353-403);
 * do_subgrid_friction_on_A_grid is off by default, so bed roughness is used
 * as given.  'roughness' is beta_sq (Weertman, Tsai2015, Schoof2005) or
 * till_yield_stress (Budd, Coulomb, Zoet-Iverson).  The reference's
 * x * u**(-1) is written x / u. */
static void sliding_beta(const ufe_mesh_desc *m, const ufe_params *p, const double *u_b,
                         const double *v_b, const double *roughness, double *beta_a) {
  int nV = m->nV;
  double *ua = malloc(sizeof(double) * nV), *va = malloc(sizeof(double) * nV);
  apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_map, u_b, ua);
  apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_map, v_b, va);
  for (int vi = 0; vi < nV; ++vi) {
    double uabs = sqrt(p->slid_delta_v * p->slid_delta_v + ua[vi] * ua[vi] + va[vi] * va[vi]);
    double b;
    switch (p->sliding_law) {
      case UFE_SLIDING_NO_SLIDING: b = 0.0; break;
      case UFE_SLIDING_WEERTMAN: b = roughness[vi] * pow(uabs, 1.0 / p->slid_weertman_m - 1.0); break;
      case UFE_SLIDING_BUDD:
        b = roughness[vi] * pow(uabs, p->slid_budd_q_plastic - 1.0) /
            pow(p->slid_budd_u_threshold, p->slid_budd_q_plastic);
        break;
      case UFE_SLIDING_COULOMB: b = roughness[vi] / uabs; break;
      case UFE_SLIDING_TSAI2015: {
        double cap = g_alpha_sq[vi] * g_effective_pressure[vi];
        b = fmin(cap, roughness[vi] * pow(uabs, 1.0 / p->slid_weertman_m)) / uabs;
        break;
      }
      case UFE_SLIDING_SCHOOF2005: {
        double cap = g_alpha_sq[vi] * g_effective_pressure[vi], m_ = p->slid_weertman_m;
        b = (roughness[vi] * pow(uabs, 1.0 / m_) * cap) /
            pow(pow(roughness[vi], m_) * uabs + pow(cap, m_), 1.0 / m_) / uabs;
        break;
      }
      case UFE_SLIDING_ZOET_IVERSON:
        b = roughness[vi] * pow(uabs, 1.0 / p->slid_zi_p - 1.0) * pow(uabs + p->slid_zi_ut, -1.0 / p->slid_zi_p);
        break;
      default: b = 0.0;
    }
    beta_a[vi] = fmin(p->slid_beta_max, b);
  }
  free(ua); free(va);
}

/* ---------------------------------------------------------------------- */
/* Stiffness matrix (solve_linearised_SSA_DIVA.f90:23-641)                 */
/* ---------------------------------------------------------------------- */

static int bc_side(int tribi) { return (tribi - 1) / 2; } /* 0 N, 1 E, 2 S, 3 W */

/* Lagged copy rows of the idealised boundary conditions
 * (calc_SSA_DIVA_stiffness_matrix_row_BC, solve_linearised_SSA_DIVA.f90:543-575,
 * :618-635): the velocity at a border triangle is set to a weighted mean of
 * the previous iterate around a copy point, relaxed with visc_it_relax.  The
 * copy point (find_ti_copy_ISMIP_HOM_periodic, find_ti_copy_SSA_icestream_infinite,
 * mesh_utilities.f90:2622-2730): 'periodic_ISMIP-HOM' shifts the triangle's
 * geometric centre by L/2 towards the domain centre in x and in y;
 * 'infinite_SSA_icestream' keeps y and takes x at a third (west half) or two
 * thirds (east half) of the domain.  Its weights: 1/d^2 to the geometric
 * centres of the triangles around the vertex nearest to the point (the one
 * whose Voronoi cell holds it), normalised.  copy->ptr[row] .. ptr[row+1]
 * index the triangles and weights of row 2*ti+uv (empty for other rows). */
typedef struct {
  int *ptr, *tri;
  double *w;
} copy_lists_t;

static void copy_lists_free(copy_lists_t *c) {
  free(c->ptr); free(c->tri); free(c->w);
  c->ptr = c->tri = NULL; c->w = NULL;
}

static int is_copy_choice(int choice) {
  return choice == UFE_BC_PERIODIC_ISMIP_HOM || choice == UFE_BC_INFINITE_SSA_ICESTREAM;
}

static void copy_lists_build(const ufe_mesh_desc *m, const ufe_params *p, copy_lists_t *c) {
  const int nT = m->nT, nV = m->nV;
  double xmin = 1e300, xmax = -1e300;
  for (int vi = 0; vi < nV; ++vi) { xmin = fmin(xmin, m->V[2 * vi]); xmax = fmax(xmax, m->V[2 * vi]); }
  c->ptr = calloc(2 * (size_t)nT + 1, sizeof(int));
  int cap = 64, n = 0;
  c->tri = malloc(sizeof(int) * cap);
  c->w = malloc(sizeof(double) * cap);
  for (int row = 0; row < 2 * nT; ++row) {
    const int ti = row / 2, uv = row % 2;
    c->ptr[row] = n;
    if (m->TriBI[ti] <= 0) continue;
    const int side = bc_side(m->TriBI[ti]);
    const int choice = uv == 0 ? p->bc_u[side] : p->bc_v[side];
    if (!is_copy_choice(choice)) continue;
    const double gx = m->TriGC[2 * ti], gy = m->TriGC[2 * ti + 1];
    double px, py;
    if (choice == UFE_BC_PERIODIC_ISMIP_HOM) {
      px = gx > 0.0 ? gx - p->ismip_hom_L / 2.0 : gx + p->ismip_hom_L / 2.0;
      py = gy > 0.0 ? gy - p->ismip_hom_L / 2.0 : gy + p->ismip_hom_L / 2.0;
    } else {
      py = gy;
      px = gx < 0.0 ? xmin + (xmax - xmin) * 1.0 / 3.0 : xmin + (xmax - xmin) * 2.0 / 3.0;
    }
    int vbest = 0;
    double dbest = 1e300;
    for (int vi = 0; vi < nV; ++vi) {
      const double dx = m->V[2 * vi] - px, dy = m->V[2 * vi + 1] - py, d = dx * dx + dy * dy;
      if (d < dbest) { dbest = d; vbest = vi; }
    }
    const int k0 = m->ba_ptr[vbest], k1 = m->ba_ptr[vbest + 1];
    if (n + (k1 - k0) > cap) {
      cap = 2 * (n + (k1 - k0));
      c->tri = realloc(c->tri, sizeof(int) * cap);
      c->w = realloc(c->w, sizeof(double) * cap);
    }
    double sum = 0.0;
    for (int k = k0; k < k1; ++k) {
      const int tj = m->ba_ind[k];
      const double dx = px - m->TriGC[2 * tj], dy = py - m->TriGC[2 * tj + 1];
      const double d = sqrt(dx * dx + dy * dy);
      c->tri[n + (k - k0)] = tj;
      c->w[n + (k - k0)] = 1.0 / (d * d);
      sum += 1.0 / (d * d);
    }
    for (int k = 0; k < k1 - k0; ++k) c->w[n + k] /= sum;
    n += k1 - k0;
  }
  c->ptr[2 * nT] = n;
}

static double copy_value(const copy_lists_t *c, const ufe_params *p, int row, const double *u, const double *v) {
  const double *f = row % 2 == 0 ? u : v;
  double val = 0.0;
  for (int k = c->ptr[row]; k < c->ptr[row + 1]; ++k) val += c->w[k] * f[c->tri[k]];
  return (p->visc_it_relax * val) + ((1.0 - p->visc_it_relax) * f[row / 2]);
}

/* Assemble A and b in the reference's row order 2*ti+uv and column order
 * (stencil entry k: u column then v column; BC rows: neighbours then diagonal). */
static void assemble(const ufe_mesh_desc *m, const ufe_params *p, const double *N_b,
                     const double *dNdx, const double *dNdy, const double *beta_b,
                     const double *tdx, const double *tdy, const int *bc_mask, const double *bc_u,
                     const double *bc_v, const copy_lists_t *copy, const double *ucur, const double *vcur,
                     int *ptr, int *ind, double *val, double *bb) {
  int nnz = 0;
  ptr[0] = 0;
  for (int row = 0; row < 2 * m->nT; ++row) {
    int ti = row / 2, uv = row % 2;
    if (bc_mask && bc_mask[ti] == 1) {
      ind[nnz] = row; val[nnz] = 1.0; ++nnz;
      bb[row] = uv == 0 ? bc_u[ti] : bc_v[ti];
    } else if (m->TriBI[ti] > 0) {
      int side = bc_side(m->TriBI[ti]);
      int choice = uv == 0 ? p->bc_u[side] : p->bc_v[side];
      if (choice == UFE_BC_INFINITE) {
        int nn = 0;
        for (int n = 0; n < 3; ++n) {
          int tj = m->TriC[3 * ti + n];
          if (tj < 0) continue;
          ++nn;
          ind[nnz] = 2 * tj + uv; val[nnz] = 1.0; ++nnz;
        }
        ind[nnz] = row; val[nnz] = -(double)nn; ++nnz;
        bb[row] = 0.0;
      } else if (is_copy_choice(choice)) {
        ind[nnz] = row; val[nnz] = 1.0; ++nnz;
        bb[row] = copy_value(copy, p, row, ucur, vcur);
      } else {
        ind[nnz] = row; val[nnz] = 1.0; ++nnz;
        bb[row] = 0.0;
      }
    } else if (p->include_crossterms) {
      /* calc_SSA_DIVA_stiffness_matrix_row_free (:180-329) */
      double N = N_b[ti], Nx = dNdx[ti], Ny = dNdy[ti], be = beta_b[ti];
      for (int k = m->bb_ptr[ti]; k < m->bb_ptr[ti + 1]; ++k) {
        int tj = m->bb_ind[k];
        double xx = m->bb_d2dx2[k], xy = m->bb_d2dxdy[k], yy = m->bb_d2dy2[k];
        double dx = m->bb_ddx[k], dy = m->bb_ddy[k];
        double Au, Av;
        if (uv == 0) {
          Au = 4.0 * N * xx + 4.0 * Nx * dx + N * yy + Ny * dy;
          if (tj == ti) Au -= be;
          Av = 3.0 * N * xy + 2.0 * Nx * dy + Ny * dx;
        } else {
          Av = 4.0 * N * yy + 4.0 * Ny * dy + N * xx + Nx * dx;
          if (tj == ti) Av -= be;
          Au = 3.0 * N * xy + 2.0 * Ny * dx + Nx * dy;
        }
        ind[nnz] = 2 * tj; val[nnz] = Au; ++nnz;
        ind[nnz] = 2 * tj + 1; val[nnz] = Av; ++nnz;
      }
      bb[row] = uv == 0 ? -tdx[ti] : -tdy[ti];
    } else {
      /* calc_SSA_DIVA_sans_stiffness_matrix_row_free (:331-479) */
      double N = N_b[ti], be = beta_b[ti];
      for (int k = m->bb_ptr[ti]; k < m->bb_ptr[ti + 1]; ++k) {
        int tj = m->bb_ind[k];
        double xx = m->bb_d2dx2[k], xy = m->bb_d2dxdy[k], yy = m->bb_d2dy2[k];
        double Au, Av;
        if (uv == 0) {
          Au = 4.0 * xx + yy;
          if (tj == ti) Au -= be / N;
          Av = 3.0 * xy;
        } else {
          Av = 4.0 * yy + xx;
          if (tj == ti) Av -= be / N;
          Au = 3.0 * xy;
        }
        ind[nnz] = 2 * tj; val[nnz] = Au; ++nnz;
        ind[nnz] = 2 * tj + 1; val[nnz] = Av; ++nnz;
      }
      bb[row] = uv == 0 ? -tdx[ti] / N : -tdy[ti] / N;
    }
    ptr[row + 1] = nnz;
  }
}

static int max_row_nnz(const ufe_mesh_desc *m) {
  int mx = 0;
  for (int ti = 0; ti < m->nT; ++ti) {
    int l = m->bb_ptr[ti + 1] - m->bb_ptr[ti];
    if (l > mx) mx = l;
  }
  return mx;
}

/* ---------------------------------------------------------------------- */
/* SSA / DIVA Picard drivers                                               */
/* ---------------------------------------------------------------------- */

typedef struct {
  const ufe_mesh_desc *m;
  const ufe_params *p;
  const double *Hi, *Hs, *A_flow, *roughness, *fgr_b;
  const int *bc_mask;
  const double *bc_u, *bc_v;
  int is_diva;
  /* work */
  double *tdx, *tdy, *ux, *uy, *vx, *vy, *eta_a, *N_a, *N_b, *dNdx, *dNdy;
  double *beta_a, *beta_b, *beta_eff_a, *beta_eff_b, *A_vav;
  double *eta3a, *F1a, *F2a, *F1b, *F2b, *duz_a, *dvz_a, *tmpb, *tmpa;
  int *Aptr, *Aind;
  double *Aval, *bb, *xx;
  /* the iterate the system is assembled at (read by the lagged copy rows) */
  const double *ucur, *vcur;
  copy_lists_t copy;
} picard_t;

static void vec_a_b_3d(const ufe_mesh_desc *m, const double *fa, double *fb) {
  for (int k = 0; k < m->nz; ++k)
    apply_csr(m->nT, m->ab_ptr, m->ab_ind, m->ab_map, fa + (size_t)k * m->nV, fb + (size_t)k * m->nT);
}
static void vec_b_a_3d(const ufe_mesh_desc *m, const double *fb, double *fa) {
  for (int k = 0; k < m->nz; ++k)
    apply_csr(m->nV, m->ba_ptr, m->ba_ind, m->ba_map, fb + (size_t)k * m->nT, fa + (size_t)k * m->nV);
}

/* One Picard state evaluation at (u, v): everything up to the linear system.
 * SSA: calc_horizontal_strain_rates, calc_effective_viscosity,
 *      calc_applied_basal_friction_coefficient (SSA_main.f90:169-175, :339-430).
 * DIVA: the same plus calc_vertical_shear_strain_rates, 3-D viscosity,
 *      calc_F_integrals and calc_effective_basal_friction_coefficient
 *      (DIVA_main.f90:174-186, :375-574), using the DIVA state arrays. */
static void picard_state(picard_t *s, const double *u, const double *v, double eps0,
                         double *eta3b, const double *tbx, const double *tby,
                         const double *ubase, const double *vbase) {
  const ufe_mesh_desc *m = s->m;
  const ufe_params *p = s->p;
  int nV = m->nV, nT = m->nT, nz = m->nz;
  double n = p->glens_n;
  double eta_max = 0.5 * pow(1e-18, -1.0 / n) * pow(eps0, (1.0 - n) / (2.0 * n));
  apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_ddx, u, s->ux);
  apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_ddy, u, s->uy);
  apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_ddx, v, s->vx);
  apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_ddy, v, s->vy);
  if (!s->is_diva) {
    for (int vi = 0; vi < nV; ++vi) {
      double eta = glen_2d(eps0, s->ux[vi], s->uy[vi], s->vx[vi], s->vy[vi], s->A_vav[vi], n);
      eta = fmin(fmax(eta, p->visc_eff_min), eta_max);
      s->eta_a[vi] = eta;
      s->N_a[vi] = eta * fmax(0.1, s->Hi[vi]);
    }
  } else {
    /* vertical shear strain rates, DIVA_main.f90:396-405 */
    for (int k = 0; k < nz; ++k)
      for (int ti = 0; ti < nT; ++ti) s->tmpb[(size_t)k * nT + ti] =
          tbx[ti] * m->zeta[k] / fmax(p->visc_eff_min, eta3b[(size_t)k * nT + ti]);
    vec_b_a_3d(m, s->tmpb, s->duz_a);
    for (int k = 0; k < nz; ++k)
      for (int ti = 0; ti < nT; ++ti) s->tmpb[(size_t)k * nT + ti] =
          tby[ti] * m->zeta[k] / fmax(p->visc_eff_min, eta3b[(size_t)k * nT + ti]);
    vec_b_a_3d(m, s->tmpb, s->dvz_a);
    /* 3-D viscosity, DIVA_main.f90:441-455 */
    for (int vi = 0; vi < nV; ++vi)
      for (int k = 0; k < nz; ++k) {
        size_t i = (size_t)k * nV + vi;
        double eta = glen_3d(eps0, s->ux[vi], s->uy[vi], s->duz_a[i], s->vx[vi], s->vy[vi], s->dvz_a[i],
                             s->A_flow[i], n);
        s->eta3a[i] = fmin(fmax(eta, p->visc_eff_min), eta_max);
      }
    vec_a_b_3d(m, s->eta3a, eta3b);
    /* vertical average and N, DIVA_main.f90:461-469; the reference writes
     * max(0.1, Hi) with a default-real 0.1 here */
    for (int vi = 0; vi < nV; ++vi) {
      double ev = vertical_average(nz, m->zeta, s->eta3a + vi, nV);
      s->eta_a[vi] = ev;
      s->N_a[vi] = ev * fmax((double)0.1f, s->Hi[vi]);
    }
    /* F-integrals, DIVA_main.f90:497-515 (integrate_from_zeta_is_one_to_zeta_is_zetap) */
    for (int vi = 0; vi < nV; ++vi) {
      double h = fmax(0.1, s->Hi[vi]);
      double I1 = 0.0, I2 = 0.0;
      s->F1a[(size_t)(nz - 1) * nV + vi] = 0.0;
      s->F2a[(size_t)(nz - 1) * nV + vi] = 0.0;
      for (int k = nz - 2; k >= 0; --k) {
        double z0 = m->zeta[k], z1 = m->zeta[k + 1];
        double e0 = s->eta3a[(size_t)k * nV + vi], e1 = s->eta3a[(size_t)(k + 1) * nV + vi];
        I1 = I1 - 0.5 * (z1 / e1 + z0 / e0) * (z1 - z0);
        I2 = I2 - 0.5 * (z1 * z1 / e1 + z0 * z0 / e0) * (z1 - z0);
        s->F1a[(size_t)k * nV + vi] = -h * I1;
        s->F2a[(size_t)k * nV + vi] = -h * I2;
      }
    }
    vec_a_b_3d(m, s->F1a, s->F1b);
    vec_a_b_3d(m, s->F2a, s->F2b);
  }
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_map, s->N_a, s->N_b);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddx, s->N_a, s->dNdx);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddy, s->N_a, s->dNdy);
  if (!s->is_diva) {
    sliding_beta(m, p, u, v, s->roughness, s->beta_a);
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_map, s->beta_a, s->beta_b);
    if (p->do_gl_subgrid_friction)
      for (int ti = 0; ti < nT; ++ti) s->beta_b[ti] *= pow(s->fgr_b[ti], p->subgrid_friction_exponent);
  } else {
    sliding_beta(m, p, ubase, vbase, s->roughness, s->beta_a);
    for (int vi = 0; vi < nV; ++vi) {
      if (p->sliding_law == UFE_SLIDING_NO_SLIDING) s->beta_eff_a[vi] = 1.0 / s->F2a[vi];
      else s->beta_eff_a[vi] = s->beta_a[vi] / (1.0 + s->beta_a[vi] * s->F2a[vi]);
    }
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_map, s->beta_a, s->beta_b);
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_map, s->beta_eff_a, s->beta_eff_b);
    if (p->do_gl_subgrid_friction)
      for (int ti = 0; ti < nT; ++ti) s->beta_eff_b[ti] *= pow(s->fgr_b[ti], p->subgrid_friction_exponent);
  }
}

static void picard_assemble(picard_t *s) {
  const double *beta = s->is_diva ? s->beta_eff_b : s->beta_b;
  assemble(s->m, s->p, s->N_b, s->dNdx, s->dNdy, beta, s->tdx, s->tdy, s->bc_mask, s->bc_u, s->bc_v,
           &s->copy, s->ucur, s->vcur, s->Aptr, s->Aind, s->Aval, s->bb);
}

/* Block-scaled nonlinear residual ||D^-1 (b(u) - A(u) u)|| / ||D^-1 b(u)||, D the
 * 2x2 diagonal block of each triangle's (u, v) rows, over the rows of triangles
 * where the velocity limit (apply_velocity_limits, SSA_DIVA_utilities.f90:110-141)
 * is not active.  Scaling each block row by its diagonal block makes every row a
 * velocity equation, so momentum rows and boundary rows are on one scale. */
static double scaled_rel_residual(int nT, const int *Aptr, const int *Aind, const double *Aval,
                                  const double *bvec, const double *x, double vel_max, double *unscaled) {
  int n = 2 * nT;
  double *r = malloc(sizeof(double) * n);
  oracle_csr_matvec(n, Aptr, Aind, Aval, x, r);
  double rr = 0, b2 = 0, ru = 0, bu = 0;
  for (int ti = 0; ti < nT; ++ti) {
    double u = x[2 * ti], v = x[2 * ti + 1];
    if (vel_max > 0 && sqrt(u * u + v * v) >= vel_max * (1.0 - 1e-9)) continue;
    double D[4] = {0, 0, 0, 0};
    for (int q = 0; q < 2; ++q)
      for (int k = Aptr[2 * ti + q]; k < Aptr[2 * ti + q + 1]; ++k) {
        int j = Aind[k];
        if (j == 2 * ti) D[2 * q] += Aval[k];
        if (j == 2 * ti + 1) D[2 * q + 1] += Aval[k];
      }
    double det = D[0] * D[3] - D[1] * D[2];
    double i0 = D[3] / det, i1 = -D[1] / det, i2 = -D[2] / det, i3 = D[0] / det;
    double r0 = bvec[2 * ti] - r[2 * ti], r1 = bvec[2 * ti + 1] - r[2 * ti + 1];
    double s0 = i0 * r0 + i1 * r1, s1 = i2 * r0 + i3 * r1;
    double c0 = i0 * bvec[2 * ti] + i1 * bvec[2 * ti + 1], c1 = i2 * bvec[2 * ti] + i3 * bvec[2 * ti + 1];
    rr += s0 * s0 + s1 * s1;
    b2 += c0 * c0 + c1 * c1;
    ru += r0 * r0 + r1 * r1;
    bu += bvec[2 * ti] * bvec[2 * ti] + bvec[2 * ti + 1] * bvec[2 * ti + 1];
  }
  free(r);
  /* the unscaled ||b - A u|| / ||b|| over the same rows */
  if (unscaled) *unscaled = bu > 0 ? sqrt(ru / bu) : sqrt(ru);
  return b2 > 0 ? sqrt(rr / b2) : sqrt(rr);
}

static double rel_residual(picard_t *s, const double *u, const double *v, double *unscaled) {
  for (int ti = 0; ti < s->m->nT; ++ti) { s->xx[2 * ti] = u[ti]; s->xx[2 * ti + 1] = v[ti]; }
  return scaled_rel_residual(s->m->nT, s->Aptr, s->Aind, s->Aval, s->bb, s->xx, s->p->vel_max, unscaled);
}

static int picard_alloc(picard_t *s) {
  const ufe_mesh_desc *m = s->m;
  int nV = m->nV, nT = m->nT, nz = m->nz;
  size_t nnzA = (size_t)2 * 2 * max_row_nnz(m) * nT + 8 * (size_t)nT;
#define AL(x, n) s->x = calloc((n), sizeof(*s->x))
  AL(tdx, nT); AL(tdy, nT); AL(ux, nV); AL(uy, nV); AL(vx, nV); AL(vy, nV); AL(eta_a, nV);
  AL(N_a, nV); AL(N_b, nT); AL(dNdx, nT); AL(dNdy, nT); AL(beta_a, nV); AL(beta_b, nT);
  AL(beta_eff_a, nV); AL(beta_eff_b, nT); AL(A_vav, nV);
  AL(eta3a, (size_t)nV * nz); AL(F1a, (size_t)nV * nz); AL(F2a, (size_t)nV * nz);
  AL(F1b, (size_t)nT * nz); AL(F2b, (size_t)nT * nz); AL(duz_a, (size_t)nV * nz);
  AL(dvz_a, (size_t)nV * nz); AL(tmpb, (size_t)nT * nz); AL(tmpa, (size_t)nV * nz);
  AL(Aptr, 2 * nT + 1); AL(Aind, nnzA); AL(Aval, nnzA); AL(bb, 2 * nT); AL(xx, 2 * nT);
#undef AL
  copy_lists_build(m, s->p, &s->copy);
  return 0;
}

static void picard_free(picard_t *s) {
  void *ps[] = {s->tdx, s->tdy, s->ux, s->uy, s->vx, s->vy, s->eta_a, s->N_a, s->N_b, s->dNdx,
                s->dNdy, s->beta_a, s->beta_b, s->beta_eff_a, s->beta_eff_b, s->A_vav, s->eta3a,
                s->F1a, s->F2a, s->F1b, s->F2b, s->duz_a, s->dvz_a, s->tmpb, s->tmpa, s->Aptr,
                s->Aind, s->Aval, s->bb, s->xx};
  for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); ++i) free(ps[i]);
  copy_lists_free(&s->copy);
}

/* The shared Picard loop of solve_SSA (SSA_main.f90:149-237) and solve_DIVA
 * (DIVA_main.f90:154-254). */
static int picard(picard_t *s, double *u, double *v, double *eta3b, double *tbx, double *tby,
                  double *ubase, double *vbase, ufe_solve_info *info) {
  const ufe_mesh_desc *m = s->m;
  const ufe_params *p = s->p;
  int nT = m->nT, nV = m->nV, nz = m->nz;
  double t0 = now_s();
  picard_alloc(s);
  s->ucur = u;
  s->vcur = v;
  driving_stress(m, s->Hi, s->Hs, s->tdx, s->tdy);
  if (!s->is_diva)
    for (int vi = 0; vi < nV; ++vi) s->A_vav[vi] = vertical_average(nz, m->zeta, s->A_flow + vi, nV);
  double L2 = 1e9, L2_prev;
  int nit_div = 0;
  double relax = p->visc_it_relax, eps0 = p->epsilon_sq_0;
  int n_axb = 0, it = 0, converged = 0, rc = 0;
  double *up = malloc(sizeof(double) * nT), *vp = malloc(sizeof(double) * nT);
  double resid = -1.0, resid_u = -1.0;
  while (!converged) {
    ++it;
    picard_state(s, u, v, eps0, eta3b, tbx, tby, ubase, vbase);
    picard_assemble(s);
    if (p->stop_rule == UFE_STOP_RESIDUAL) {
      resid = rel_residual(s, u, v, &resid_u);
      if (resid <= p->nonlinear_rtol) { --it; converged = 1; break; }
    }
    /* solve_SSA_DIVA_linearised: keep previous solution, solve, split u/v */
    memcpy(up, u, sizeof(double) * nT);
    memcpy(vp, v, sizeof(double) * nT);
    for (int ti = 0; ti < nT; ++ti) { s->xx[2 * ti] = u[ti]; s->xx[2 * ti + 1] = v[ti]; }
    int its = oracle_gmres_ilu0(2 * nT, s->Aptr, s->Aind, s->Aval, s->bb, s->xx, p->rtol, p->abstol,
                                p->krylov_maxit, p->gmres_restart, !p->warm_start, NULL);
    if (its < 0) { rc = UFE_ERR_ARGUMENT; break; }
    n_axb += its;
    for (int ti = 0; ti < nT; ++ti) { u[ti] = s->xx[2 * ti]; v[ti] = s->xx[2 * ti + 1]; }
    /* apply_velocity_limits (SSA_DIVA_utilities.f90:110-141) */
    for (int ti = 0; ti < nT; ++ti) {
      double ua = sqrt(u[ti] * u[ti] + v[ti] * v[ti]);
      if (ua > p->vel_max) { u[ti] = u[ti] * p->vel_max / ua; v[ti] = v[ti] * p->vel_max / ua; }
    }
    /* relax_viscosity_iterations (:84-108) */
    for (int ti = 0; ti < nT; ++ti) {
      u[ti] = relax * u[ti] + (1.0 - relax) * up[ti];
      v[ti] = relax * v[ti] + (1.0 - relax) * vp[ti];
    }
    if (s->is_diva) {
      /* calc_basal_velocities, calc_basal_shear_stress (DIVA_main.f90:576-634) */
      for (int ti = 0; ti < nT; ++ti) {
        if (p->sliding_law == UFE_SLIDING_NO_SLIDING) { ubase[ti] = 0; vbase[ti] = 0; }
        else {
          double d = 1.0 + s->beta_b[ti] * s->F2b[ti];
          ubase[ti] = u[ti] / d;
          vbase[ti] = v[ti] / d;
        }
        tbx[ti] = u[ti] * s->beta_eff_b[ti];
        tby[ti] = v[ti] * s->beta_eff_b[ti];
      }
    }
    /* calc_L2_norm_uv (:143-184) */
    L2_prev = L2;
    double r1 = 0, r2 = 0;
    for (int ti = 0; ti < nT; ++ti) {
      r1 += (u[ti] - up[ti]) * (u[ti] - up[ti]) + (v[ti] - vp[ti]) * (v[ti] - vp[ti]);
      r2 += (u[ti] + up[ti]) * (u[ti] + up[ti]) + (v[ti] + vp[ti]) * (v[ti] + vp[ti]);
    }
    L2 = 2.0 * r1 / fmax(r2, 1e-8);
    /* adaptive relaxation (SSA_main.f90:196-213) */
    if (L2 > L2_prev) ++nit_div; else nit_div = 0;
    if (nit_div > 2) { nit_div = 0; relax *= 0.9; eps0 *= 1.2; }
    if (relax <= 0.05 || eps0 >= 1e-5) {
      if (relax < 0.05 || eps0 > 1e-5) { rc = UFE_ERR_DIVERGED; break; }
    }
    if (p->stop_rule == UFE_STOP_L2UV && L2 < p->visc_it_norm_dUV_tol) converged = 1;
    if (it > p->visc_it_nit) { rc = UFE_WARN_NOT_CONVERGED; break; }
  }
  if (s->is_diva && rc >= 0) {
    /* calc_3D_velocities is applied by the caller */
  }
  if (info) {
    if (p->stop_rule != UFE_STOP_RESIDUAL || !converged) {
      /* Diagnostic residual at the returned iterate.  The evaluation
       * overwrites the DIVA state; the returned eta_3D_b and the F1_3D_b and
       * beta_b that calc_3D_velocities reads (DIVA_main.f90:254) stay those
       * of the last state the loop evaluated, as without info. */
      double *keep_e = NULL, *keep_f = NULL, *keep_b = NULL;
      size_t n3 = (size_t)nT * nz;
      if (s->is_diva) {
        keep_e = malloc(sizeof(double) * n3);
        keep_f = malloc(sizeof(double) * n3);
        keep_b = malloc(sizeof(double) * nT);
        memcpy(keep_e, eta3b, sizeof(double) * n3);
        memcpy(keep_f, s->F1b, sizeof(double) * n3);
        memcpy(keep_b, s->beta_b, sizeof(double) * nT);
      }
      picard_state(s, u, v, eps0, eta3b, tbx, tby, ubase, vbase);
      picard_assemble(s);
      resid = rel_residual(s, u, v, &resid_u);
      if (s->is_diva) {
        memcpy(eta3b, keep_e, sizeof(double) * n3);
        memcpy(s->F1b, keep_f, sizeof(double) * n3);
        memcpy(s->beta_b, keep_b, sizeof(double) * nT);
        free(keep_e); free(keep_f); free(keep_b);
      }
    }
    info->n_visc_its = it;
    info->n_Axb_its = n_axb;
    info->n_refinements = 0;
    info->L2_uv = L2;
    info->rel_residual = resid;
    info->rel_residual_unscaled = resid_u;
    info->visc_it_relax_applied = relax;
    info->epsilon_sq_0_applied = eps0;
    info->seconds = now_s() - t0;
  }
  free(up); free(vp);
  return rc;
}

static int any_grounded(int nV, const int *mask) {
  if (!mask) return 1;
  for (int vi = 0; vi < nV; ++vi) if (mask[vi]) return 1;
  return 0;
}

/* solve_SSA (SSA_main.f90:87-242) */
int oracle_solve_ssa(const ufe_mesh_desc *m, const ufe_params *p, const double *Hi, const double *Hs,
                     const double *A_flow, const double *beta_sq, const double *fgr_b,
                     const int *mask_grounded, const int *bc_mask, const double *bc_u,
                     const double *bc_v, double *u, double *v, ufe_solve_info *info) {
  if (!any_grounded(m->nV, mask_grounded) || p->sliding_law == UFE_SLIDING_NO_SLIDING) {
    memset(u, 0, sizeof(double) * m->nT);
    memset(v, 0, sizeof(double) * m->nT);
    if (info) memset(info, 0, sizeof(*info));
    return 0;
  }
  picard_t s;
  memset(&s, 0, sizeof(s));
  s.m = m; s.p = p; s.Hi = Hi; s.Hs = Hs; s.A_flow = A_flow; s.roughness = beta_sq; s.fgr_b = fgr_b;
  s.bc_mask = bc_mask; s.bc_u = bc_u; s.bc_v = bc_v; s.is_diva = 0;
  int rc = picard(&s, u, v, NULL, NULL, NULL, NULL, NULL, info);
  picard_free(&s);
  return rc;
}

/* solve_DIVA (DIVA_main.f90:88-262) including calc_3D_velocities (:636-676) */
int oracle_solve_diva(const ufe_mesh_desc *m, const ufe_params *p, const double *Hi, const double *Hs,
                      const double *A_flow, const double *beta_sq, const double *fgr_b,
                      const int *mask_grounded, const int *bc_mask, const double *bc_u,
                      const double *bc_v, double *u, double *v, double *ubase, double *vbase,
                      double *u3d, double *v3d, double *eta3b, double *tbx, double *tby,
                      ufe_solve_info *info) {
  int nT = m->nT, nz = m->nz;
  if (!any_grounded(m->nV, mask_grounded)) {
    memset(u, 0, sizeof(double) * nT); memset(v, 0, sizeof(double) * nT);
    memset(ubase, 0, sizeof(double) * nT); memset(vbase, 0, sizeof(double) * nT);
    memset(u3d, 0, sizeof(double) * nT * nz); memset(v3d, 0, sizeof(double) * nT * nz);
    if (info) memset(info, 0, sizeof(*info));
    return 0;
  }
  picard_t s;
  memset(&s, 0, sizeof(s));
  s.m = m; s.p = p; s.Hi = Hi; s.Hs = Hs; s.A_flow = A_flow; s.roughness = beta_sq; s.fgr_b = fgr_b;
  s.bc_mask = bc_mask; s.bc_u = bc_u; s.bc_v = bc_v; s.is_diva = 1;
  int rc = picard(&s, u, v, eta3b, tbx, tby, ubase, vbase, info);
  if (rc >= 0) {
    for (int k = 0; k < nz; ++k)
      for (int ti = 0; ti < nT; ++ti) {
        size_t i = (size_t)k * nT + ti;
        if (p->sliding_law == UFE_SLIDING_NO_SLIDING) {
          u3d[i] = tbx[ti] * s.F1b[i];
          v3d[i] = tby[ti] * s.F1b[i];
        } else {
          u3d[i] = ubase[ti] * (1.0 + s.beta_b[ti] * s.F1b[i]);
          v3d[i] = vbase[ti] * (1.0 + s.beta_b[ti] * s.F1b[i]);
        }
      }
  }
  picard_free(&s);
  return rc;
}

/* Nonlinear residual of the SSA at (u, v): r = b(u) - A(u) u in the
 * reference row order 2*ti+uv; returns ||r||/||b||.  Used by the parity tests
 * to check the whole per-iteration pipeline (strain rates, viscosity,
 * sliding, a->b maps, assembly, product) at one point. */
double oracle_ssa_residual(const ufe_mesh_desc *m, const ufe_params *p, const double *Hi,
                           const double *Hs, const double *A_flow, const double *beta_sq,
                           const double *fgr_b, const int *bc_mask, const double *bc_u,
                           const double *bc_v, const double *u, const double *v, double *r_out,
                           double *b_out) {
  picard_t s;
  memset(&s, 0, sizeof(s));
  s.m = m; s.p = p; s.Hi = Hi; s.Hs = Hs; s.A_flow = A_flow; s.roughness = beta_sq; s.fgr_b = fgr_b;
  s.bc_mask = bc_mask; s.bc_u = bc_u; s.bc_v = bc_v; s.is_diva = 0;
  picard_alloc(&s);
  s.ucur = u;
  s.vcur = v;
  driving_stress(m, Hi, Hs, s.tdx, s.tdy);
  for (int vi = 0; vi < m->nV; ++vi) s.A_vav[vi] = vertical_average(m->nz, m->zeta, A_flow + vi, m->nV);
  picard_state(&s, u, v, p->epsilon_sq_0, NULL, NULL, NULL, NULL, NULL);
  picard_assemble(&s);
  int n = 2 * m->nT;
  for (int ti = 0; ti < m->nT; ++ti) { s.xx[2 * ti] = u[ti]; s.xx[2 * ti + 1] = v[ti]; }
  double *r = malloc(sizeof(double) * n);
  oracle_csr_matvec(n, s.Aptr, s.Aind, s.Aval, s.xx, r);
  double rr = 0, b2 = 0;
  for (int i = 0; i < n; ++i) {
    r[i] = s.bb[i] - r[i];
    rr += r[i] * r[i];
    b2 += s.bb[i] * s.bb[i];
    if (r_out) r_out[i] = r[i];
    if (b_out) b_out[i] = s.bb[i];
  }
  free(r);
  double rel = scaled_rel_residual(m->nT, s.Aptr, s.Aind, s.Aval, s.bb, s.xx, p->vel_max, NULL);
  (void)rr; (void)b2;
  picard_free(&s);
  return rel;
}

/* Nonlinear DIVA residual at (u, v) and a given DIVA state (eta3b, tau_b,
 * u_base), one state evaluation of solve_DIVA (DIVA_main.f90:174-186, with
 * calc_vertical_shear_strain_rates :375-410, calc_effective_viscosity
 * :412-479, calc_F_integrals :481-520 and
 * calc_effective_basal_friction_coefficient :522-574) followed by the
 * assembly of solve_SSA_DIVA_linearised.  3-D arrays are [k][ti]; eta3b is
 * read as the previous state and receives the evaluated one.  Outputs may be
 * NULL.  Returns the block-scaled relative residual. */
double oracle_diva_residual(const ufe_mesh_desc *m, const ufe_params *p, const double *Hi,
                            const double *Hs, const double *A_flow, const double *beta_sq,
                            const double *fgr_b, const double *u, const double *v, double *eta3b,
                            const double *tbx, const double *tby, const double *ubase,
                            const double *vbase, double *r_out, double *b_out, double *F1b_out,
                            double *F2s_out, double *beta_b_out, double *beta_eff_b_out) {
  picard_t s;
  memset(&s, 0, sizeof(s));
  s.m = m; s.p = p; s.Hi = Hi; s.Hs = Hs; s.A_flow = A_flow; s.roughness = beta_sq; s.fgr_b = fgr_b;
  s.is_diva = 1;
  picard_alloc(&s);
  s.ucur = u;
  s.vcur = v;
  int nT = m->nT, nz = m->nz, n = 2 * nT;
  driving_stress(m, Hi, Hs, s.tdx, s.tdy);
  picard_state(&s, u, v, p->epsilon_sq_0, eta3b, tbx, tby, ubase, vbase);
  picard_assemble(&s);
  for (int ti = 0; ti < nT; ++ti) { s.xx[2 * ti] = u[ti]; s.xx[2 * ti + 1] = v[ti]; }
  double *r = malloc(sizeof(double) * n);
  oracle_csr_matvec(n, s.Aptr, s.Aind, s.Aval, s.xx, r);
  for (int i = 0; i < n; ++i) {
    if (r_out) r_out[i] = s.bb[i] - r[i];
    if (b_out) b_out[i] = s.bb[i];
  }
  free(r);
  if (F1b_out) memcpy(F1b_out, s.F1b, sizeof(double) * (size_t)nT * nz);
  if (F2s_out) memcpy(F2s_out, s.F2b, sizeof(double) * nT); /* layer k = 1 (surface) */
  if (beta_b_out) memcpy(beta_b_out, s.beta_b, sizeof(double) * nT);
  if (beta_eff_b_out) memcpy(beta_eff_b_out, s.beta_eff_b, sizeof(double) * nT);
  double rel = scaled_rel_residual(nT, s.Aptr, s.Aind, s.Aval, s.bb, s.xx, p->vel_max, NULL);
  picard_free(&s);
  return rel;
}

/* Export the assembled system at (u, v) in CSR (reference ordering); the
 * caller provides arrays of size 2*nT+1 and nnz_cap. Returns nnz. */
int oracle_ssa_system(const ufe_mesh_desc *m, const ufe_params *p, const double *Hi, const double *Hs,
                      const double *A_flow, const double *beta_sq, const double *fgr_b,
                      const double *u, const double *v, int nnz_cap, int *ptr, int *ind, double *val,
                      double *b) {
  picard_t s;
  memset(&s, 0, sizeof(s));
  s.m = m; s.p = p; s.Hi = Hi; s.Hs = Hs; s.A_flow = A_flow; s.roughness = beta_sq; s.fgr_b = fgr_b;
  picard_alloc(&s);
  s.ucur = u;
  s.vcur = v;
  driving_stress(m, Hi, Hs, s.tdx, s.tdy);
  for (int vi = 0; vi < m->nV; ++vi) s.A_vav[vi] = vertical_average(m->nz, m->zeta, A_flow + vi, m->nV);
  picard_state(&s, u, v, p->epsilon_sq_0, NULL, NULL, NULL, NULL, NULL);
  picard_assemble(&s);
  int n = 2 * m->nT, nnz = s.Aptr[n];
  if (nnz > nnz_cap) { picard_free(&s); return -nnz; }
  memcpy(ptr, s.Aptr, sizeof(int) * (n + 1));
  memcpy(ind, s.Aind, sizeof(int) * nnz);
  memcpy(val, s.Aval, sizeof(double) * nnz);
  memcpy(b, s.bb, sizeof(double) * n);
  picard_free(&s);
  return nnz;
}

/* ---------------------------------------------------------------------- */
/* Grounded fractions (subgrid_grounded_fractions_bilin_TAF.f90:182-370)   */
/* ---------------------------------------------------------------------- */

static double tri_area(const double *p, const double *q, const double *r) {
  return fabs((q[0] - p[0]) * (r[1] - p[1]) - (q[1] - p[1]) * (r[0] - p[0])) / 2.0;
}

static double sub_area(const double *va, const double *vb, const double *vc, double ta, double tb,
                       double tc) {
  double lab = ta / (ta - tb), lac = ta / (ta - tc);
  double pab[2] = {va[0] * (1 - lab) + vb[0] * lab, va[1] * (1 - lab) + vb[1] * lab};
  double pac[2] = {va[0] * (1 - lac) + vc[0] * lac, va[1] * (1 - lac) + vc[1] * lac};
  return tri_area(va, pab, pac);
}

void oracle_fraction_gr_b(const ufe_mesh_desc *m, const double *Hi, const double *Hb,
                          const double *SL, double *fgr) {
  const double tol = 1e-9;
  for (int ti = 0; ti < m->nT; ++ti) {
    const int *t = m->Tri + 3 * ti;
    double T[3];
    for (int n = 0; n < 3; ++n) {
      int vi = t[n];
      T[n] = Hi[vi] - fmax(0.0, (SL[vi] - Hb[vi]) * (SEAWATER_DENSITY / ICE_DENSITY));
    }
    double tmax = fmax(T[0], fmax(T[1], T[2])), tmin = fmin(T[0], fmin(T[1], T[2]));
    if (tmin >= 0.0) { fgr[ti] = 1.0; continue; }
    if (tmax <= 0.0) { fgr[ti] = 0.0; continue; }
    const double *va = m->V + 2 * t[0], *vb = m->V + 2 * t[1], *vc = m->V + 2 * t[2];
    double At = tri_area(va, vb, vc), Ag;
    double a = T[0] >= 0 ? fmax(tol, T[0]) : fmin(-tol, T[0]);
    double b = T[1] >= 0 ? fmax(tol, T[1]) : fmin(-tol, T[1]);
    double c = T[2] >= 0 ? fmax(tol, T[2]) : fmin(-tol, T[2]);
    if (a >= 0 && b >= 0 && c >= 0) Ag = At;
    else if (a <= 0 && b <= 0 && c <= 0) Ag = 0;
    else if (a >= 0 && b <= 0 && c <= 0) Ag = sub_area(va, vb, vc, a, b, c);
    else if (a <= 0 && b >= 0 && c <= 0) Ag = sub_area(vb, vc, va, b, c, a);
    else if (a <= 0 && b <= 0 && c >= 0) Ag = sub_area(vc, va, vb, c, a, b);
    else if (a <= 0 && b >= 0 && c >= 0) Ag = At - sub_area(va, vb, vc, a, b, c);
    else if (a >= 0 && b <= 0 && c >= 0) Ag = At - sub_area(vb, vc, va, b, c, a);
    else Ag = At - sub_area(vc, va, vb, c, a, b);
    fgr[ti] = fmin(1.0, fmax(0.0, Ag / At));
  }
}

/* ---------------------------------------------------------------------- */
/* Heat columns (thermodynamics_3D_heat_equation.f90:114-293, :309-479)    */
/* ---------------------------------------------------------------------- */

/* zeta tridiagonal operators: calc_vertical_operators_reg_1D
 * (mesh_zeta.f90:287-365) with calc_shape_functions_1D_reg_2nd_order
 * (shape_functions.f90:17-79), then calc_zeta_operators_tridiagonal (:487-605) */
void oracle_zeta_tridiagonals(int nz, const double *zeta, double *dl, double *dd, double *du,
                              double *d2l, double *d2d, double *d2u) {
  for (int k = 0; k < nz; ++k) {
    int c[2];
    if (k == 0) { c[0] = 1; c[1] = 2; }
    else if (k == nz - 1) { c[0] = nz - 3; c[1] = nz - 2; }
    else { c[0] = k - 1; c[1] = k + 1; }
    double dx[2], w[2], A[4] = {0, 0, 0, 0};
    for (int i = 0; i < 2; ++i) {
      dx[i] = zeta[c[i]] - zeta[k];
      w[i] = 1.0 / pow(fabs(dx[i]), 1.5);
      double w2 = w[i] * w[i];
      A[0] += w2 * dx[i] * dx[i];
      A[1] += w2 * dx[i] * 0.5 * dx[i] * dx[i];
      A[2] += w2 * 0.5 * dx[i] * dx[i] * dx[i];
      A[3] += w2 * 0.5 * dx[i] * dx[i] * 0.5 * dx[i] * dx[i];
    }
    double det = A[0] * A[3] - A[1] * A[2];
    double M[4] = {A[3] / det, -A[1] / det, -A[2] / det, A[0] / det};
    double fx[2], fxx[2];
    for (int i = 0; i < 2; ++i) {
      double w2 = w[i] * w[i];
      fx[i] = w2 * (M[0] * dx[i] + M[1] * 0.5 * dx[i] * dx[i]);
      fxx[i] = w2 * (M[2] * dx[i] + M[3] * 0.5 * dx[i] * dx[i]);
    }
    double fxi = -(fx[0] + fx[1]), fxxi = -(fxx[0] + fxx[1]);
    dd[k] = fxi; d2d[k] = fxxi;
    if (k > 0) { dl[k - 1] = 0; d2l[k - 1] = 0; }
    if (k < nz - 1) { du[k] = 0; d2u[k] = 0; }
    for (int i = 0; i < 2; ++i) {
      if (c[i] == k - 1) { dl[k - 1] = fx[i]; d2l[k - 1] = fxx[i]; }
      if (c[i] == k + 1) { du[k] = fx[i]; d2u[k] = fxx[i]; }
    }
  }
}

static void solve_column(int nz, const double *zeta, const double *dl, const double *dd,
                         const double *du, const double *d2l, const double *d2d, const double *d2u,
                         const double *Ti, const double *u, const double *v, const double *w,
                         const double *uadv, const double *vadv, double T_surf, const double *Tpmp,
                         const double *Ki, const double *Cpi, const double *zx, const double *zy,
                         const double *zz, const double *zt, const double *Phi, double dt,
                         int grounded, double Q, double Tbf, double *out) {
  double a[64], b[64], c[64], d[64];
  for (int k = 0; k < nz; ++k) { a[k] = 0; b[k] = 0; c[k] = 0; d[k] = 0; }
  b[0] = 1.0;
  d[0] = fmin(T_surf, T0);
  b[nz - 1] = 1.0;
  if (grounded)
    d[nz - 1] = fmin(Tpmp[nz - 1], Ti[nz - 2] - (zeta[nz - 1] - zeta[nz - 2]) * Q / (zz[nz - 1] * Ki[nz - 1]));
  else
    d[nz - 1] = fmin(Tbf, Tpmp[nz - 1]);
  for (int k = 1; k < nz - 1; ++k) {
    double c1 = zt[k] + u[k] * zx[k] + v[k] * zy[k] + w[k] * zz[k];
    double c2 = -Ki[k] / (ICE_DENSITY * Cpi[k]) * zz[k] * zz[k];
    a[k] = c1 * dl[k - 1] + c2 * d2l[k - 1];
    b[k] = 1.0 / dt + c1 * dd[k] + c2 * d2d[k];
    c[k] = c1 * du[k] + c2 * d2u[k];
    d[k] = Ti[k] / dt - uadv[k] - vadv[k] + Phi[k] / (ICE_DENSITY * Cpi[k]);
  }
  oracle_tridiag(nz, a, b, c, d, out);
  for (int k = 0; k < nz; ++k) out[k] = fmin(out[k], Tpmp[k]);
}

/* Unstable columns: the 1% rule of thermodynamics_3D_heat_equation.f90:254-279 and
 * replace_Ti_with_robin_solution (thermodynamics_utilities.f90:256-337) per flagged
 * column; Ti is [nz][nV] (vertex fastest), updated in place. */
int oracle_heat_robin(const ufe_heat_robin_inputs *in, const int *is_unstable, double *Ti, int *n_replaced) {
  const int nV = in->nV, nz = in->nz;
  int n = 0;
  for (int vi = 0; vi < nV; ++vi) n += is_unstable[vi] == 1;
  *n_replaced = 0;
  if (n >= (nV + 99) / 100) return UFE_ERR_DIVERGED;
  const double sec_per_year = 31556943.36;                              /* parameters.f90:16 */
  const double Kr = 9.828 * sec_per_year * exp(-0.0057 * T0);    /* :283-287 */
  const double kappa = Kr / (ICE_DENSITY * 2127.5);                     /* :285,288 */
  for (int vi = 0; vi < nV; ++vi) {
    if (is_unstable[vi] != 1) continue;
    const double grad_b = -in->geothermal_flux[vi] / Kr;               /* :289 */
    const double Ts = fmin(T0, in->T_surf[vi]);                  /* :291 */
    const double H = in->Hi_eff[vi];
    for (int k = 0; k < nz; ++k) {
      double *t = &Ti[(size_t)k * nV + vi];
      const double tpmp = T0 - CLAUSIUS_CLAPEYRON * H * in->zeta[k];
      if (H > in->Hi_min_thermo) {
        if (in->mask[vi] == 1 || in->mask[vi] == 3) {
          if (in->SMB[vi] > 0.0) {                                      /* :299-308 */
            const double L = sqrt(2.0 * kappa * H / in->SMB[vi]);
            const double d = (1.0 - in->zeta[k]) * H;
            *t = Ts + sqrt(acos(-1.0)) / 2.0 * L * grad_b * (erf(d / L) - erf(H / L));
          } else {                                                      /* :310-313 */
            *t = Ts + ((T0 - CLAUSIUS_CLAPEYRON * H) - Ts) * in->zeta[k];
          }
        } else if (in->mask[vi] == 2) {                                 /* :317-321 */
          const double tpmp_base = T0 - CLAUSIUS_CLAPEYRON * H * in->zeta[nz - 1];
          *t = Ts + in->zeta[k] * (tpmp_base - Ts);
        }
      } else {
        *t = Ts;                                                        /* :325-329 */
      }
      *t = fmin(*t, tpmp);                                              /* :331-334 */
    }
  }
  *n_replaced = n;
  return 0;
}

int oracle_heat_columns(const ufe_heat_inputs *in, double *Tnew, int *is_unstable, int *n_unstable) {
  int nV = in->nV, nz = in->nz;
  if (nz > 64 || nz < 3) return UFE_ERR_ARGUMENT;
  double dl[64], dd[64], du[64], d2l[64], d2d[64], d2u[64];
  oracle_zeta_tridiagonals(nz, in->zeta, dl, dd, du, d2l, d2d, d2u);
  for (int vi = 0; vi < nV; ++vi)
    if (!(in->Hi_eff[vi] < in->Hi_min_thermo) && in->mask[vi] != 1 && in->mask[vi] != 2 &&
        !(in->mask[vi] == 3 && in->fraction_gr))
      return UFE_ERR_ARGUMENT; /* reference crash: thick ice neither grounded nor floating */
  int nun = 0;
#pragma omp parallel for reduction(+ : nun) schedule(static)
  for (int vi = 0; vi < nV; ++vi) {
    double col[12][64];
    double Ti[64], Tn[64];
    is_unstable[vi] = 0;
    if (in->Hi_eff[vi] < in->Hi_min_thermo) {
      for (int k = 0; k < nz; ++k) Tnew[(size_t)k * nV + vi] = in->T_surf[vi];
      continue;
    }
    const double *src[12] = {in->u_3D, in->v_3D, in->w_3D, in->u_dTdxp_upwind, in->v_dTdyp_upwind,
                             in->dzeta_dx, in->dzeta_dy, in->dzeta_dz, in->dzeta_dt, in->Phi, NULL, NULL};
    for (int q = 0; q < 10; ++q)
      for (int k = 0; k < nz; ++k) col[q][k] = src[q][(size_t)k * nV + vi];
    double Tpmp[64], Ki[64], Cpi[64];
    for (int k = 0; k < nz; ++k) {
      double T = in->Ti[(size_t)k * nV + vi];
      Ti[k] = T;
      /* calc_pressure_melting_point, calc_thermal_conductivity (Ritz1987),
       * calc_heat_capacity (Pounder1965): thermodynamics_utilities.f90:130-229 */
      Tpmp[k] = T0 - CLAUSIUS_CLAPEYRON * in->Hi_eff[vi] * in->zeta[k];
      Ki[k] = 3.101e8 * exp(-0.0057 * T);
      Cpi[k] = 2115.3 + 7.79293 * (T - T0);
    }
    int grounded = in->mask[vi] != 2;
    int gl = in->mask[vi] == 3;
    double Tbf = Tpmp[nz - 1];
    /* The reference does not reset the column between halving attempts: each
     * attempt starts from the profile the previous attempt left
     * (thermodynamics_3D_heat_equation.f90:161-239). */
    int stable = 0, it_dt = 0;
    double Tcur[64];
    for (int k = 0; k < nz; ++k) Tcur[k] = Ti[k];
    while (!stable && it_dt < 10) {
      ++it_dt;
      double dta = in->dt * pow(0.5, it_dt - 1);
      for (int s = 0; s < (1 << (it_dt - 1)); ++s) {
        solve_column(nz, in->zeta, dl, dd, du, d2l, d2d, d2u, Tcur, col[0], col[1], col[2], col[3], col[4],
                     in->T_surf[vi], Tpmp, Ki, Cpi, col[5], col[6], col[7], col[8], col[9], dta,
                     grounded, in->Q_base_grnd[vi], Tbf, Tn);
        if (gl) {
          /* grounding line, choice_GL_temperature_BC = 'subgrid': the fully
           * floating solution from the same profile, then
           * fraction_gr * grounded + (1 - fraction_gr) * floating
           * (thermodynamics_3D_heat_equation.f90:171-203) */
          double Tf[64];
          solve_column(nz, in->zeta, dl, dd, du, d2l, d2d, d2u, Tcur, col[0], col[1], col[2], col[3], col[4],
                       in->T_surf[vi], Tpmp, Ki, Cpi, col[5], col[6], col[7], col[8], col[9], dta,
                       0, in->Q_base_grnd[vi], Tbf, Tf);
          const double f = in->fraction_gr[vi];
          for (int k = 0; k < nz; ++k) Tn[k] = f * Tn[k] + (1.0 - f) * Tf[k];
        }
        for (int k = 0; k < nz; ++k) Tcur[k] = Tn[k];
      }
      stable = 1;
      for (int k = 0; k < nz; ++k)
        if (Tcur[k] != Tcur[k] || Tcur[k] < 180.0 || Tcur[k] > T0) stable = 0;
    }
    if (!stable) { is_unstable[vi] = 1; ++nun; }
    for (int k = 0; k < nz; ++k) Tnew[(size_t)k * nV + vi] = Tcur[k];
  }
  *n_unstable = nun;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Ice thickness, semi-implicit step (conservation_of_mass_semiimplicit.f90) */
/* ---------------------------------------------------------------------- */

/* ice_surface_elevation, Hi_from_Hb_Hs_and_SL (ice_geometry_basics.f90:28-82) */
static double surface_elevation(double Hi, double Hb, double SL) {
  return Hi + fmax(SL - ICE_DENSITY / SEAWATER_DENSITY * Hi, Hb);
}
static double thickness_from_surface(double Hb, double Hs, double SL) {
  double Hi_float = fmax(0.0, (SL - Hb) * (SEAWATER_DENSITY / ICE_DENSITY));
  double Hs_float = Hb + Hi_float;
  if (Hs > Hs_float) return Hs - Hb;
  return fmin(Hi_float, (Hs - SL) / (1.0 - (ICE_DENSITY / SEAWATER_DENSITY)));
}

static int border_side(int vbi) {  /* 1,2 north; 3,4 east; 5,6 south; 7,8 west */
  return vbi > 0 ? (vbi - 1) / 2 : -1;
}

/* apply_ice_thickness_BC_explicit (conservation_of_mass_explicit.f90): 'zero'
 * borders get no ice; 'infinite' borders take the mean surface elevation of
 * their interior neighbours (vertices not on the border and not in the
 * no-ice mask), and border vertices without such neighbours, in a second
 * pass, the mean surface elevation of all their neighbours. */
static void thickness_bc_explicit(int nV, const double *Vx, const ufe_vertex_geometry *g, const int *bcH,
                                  const int *mask_noice, const double *Hb, const double *SL, double *Hi) {
  (void)Vx;
  double *Hs = malloc(sizeof(double) * nV);
  int *nint = malloc(sizeof(int) * nV);
  for (int vi = 0; vi < nV; ++vi) {
    Hs[vi] = surface_elevation(Hi[vi], Hb[vi], SL[vi]);
    nint[vi] = 0;
    for (int c = g->C_ptr[vi]; c < g->C_ptr[vi + 1]; ++c) {
      int vj = g->C_ind[c];
      if (g->VBI[vj] == 0 && !mask_noice[vj]) ++nint[vi];
    }
  }
  /* each pass reads the surface elevations as they were when it started
   * (the reference gathers them to all processes before each pass) */
  double *Hs0 = malloc(sizeof(double) * nV);
  for (int pass = 0; pass < 2; ++pass) {
    memcpy(Hs0, Hs, sizeof(double) * nV);
    for (int vi = 0; vi < nV; ++vi) {
      int side = border_side(g->VBI[vi]);
      if (side < 0) continue;
      if (bcH[side] == UFE_BC_ZERO) { Hi[vi] = 0.0; continue; }
      if ((pass == 0) != (nint[vi] > 0)) continue;
      double sum = 0.0;
      int n = 0;
      for (int c = g->C_ptr[vi]; c < g->C_ptr[vi + 1]; ++c) {
        int vj = g->C_ind[c];
        if (pass == 1 || (g->VBI[vj] == 0 && !mask_noice[vj])) { sum += Hs0[vj]; ++n; }
      }
      double Hs_av = sum / (double)(pass == 0 ? nint[vi] : n);
      Hs[vi] = fmax(Hb[vi], Hs_av);
      Hi[vi] = thickness_from_surface(Hb[vi], Hs[vi], SL[vi]);
    }
  }
  free(Hs0);
  free(Hs);
  free(nint);
}

int oracle_thickness_semiimplicit(int nV, const double *V, const ufe_vertex_geometry *g,
                                  const ufe_thickness_params *p, const ufe_thickness_inputs *in,
                                  double *Hi_next, double *dHi_dt, double *divQ, double *dt_max_out,
                                  int *n_its) {
  const int nC = g->nC;
  const double dt = in->dt;
  /* calc_ice_flux_divergence_matrix_upwind (conservation_of_mass_utilities.f90):
   * row vi = [diag, neighbours in C order] */
  int *ptr = malloc(sizeof(int) * (nV + 1));
  int *ind = malloc(sizeof(int) * (nV + nC));
  double *M = malloc(sizeof(double) * (nV + nC));
  ptr[0] = 0;
  for (int vi = 0; vi < nV; ++vi) {
    int k0 = ptr[vi];
    double diag = 0.0;
    ind[k0] = vi;
    for (int c = g->C_ptr[vi]; c < g->C_ptr[vi + 1]; ++c) {
      int vj = g->C_ind[c], k = k0 + 1 + (c - g->C_ptr[vi]);
      /* map_velocities_from_b_to_c_2D: mean of the triangles either side */
      int tl = g->ETri[2 * c], tr = g->ETri[2 * c + 1];
      double uc, vc;
      if (tl < 0) { uc = in->u_vav_b[tr]; vc = in->v_vav_b[tr]; }
      else if (tr < 0) { uc = in->u_vav_b[tl]; vc = in->v_vav_b[tl]; }
      else { uc = (in->u_vav_b[tl] + in->u_vav_b[tr]) / 2.0; vc = (in->v_vav_b[tl] + in->v_vav_b[tr]) / 2.0; }
      double Dx = V[2 * vj] - V[2 * vi], Dy = V[2 * vj + 1] - V[2 * vi + 1];
      double D = sqrt(Dx * Dx + Dy * Dy);
      double u_perp = uc * Dx / D + vc * Dy / D;
      double A_i = g->A[vi], L_c = g->Cw[c];
      if (in->fraction_margin[vi] >= 1.0) diag = diag + L_c * fmax(0.0, u_perp) / A_i;
      ind[k] = vj;
      M[k] = in->fraction_margin[vj] >= 1.0 ? L_c * fmin(0.0, u_perp) / A_i : 0.0;
    }
    M[k0] = diag;
    ptr[vi + 1] = k0 + 1 + (g->C_ptr[vi + 1] - g->C_ptr[vi]);
  }
  oracle_csr_matvec(nV, ptr, ind, M, in->Hi, divQ);
  /* calc_dHi_dt_explicit: rates, flux-limited time step
   * (calc_flux_limited_timestep restated as written: a thinning vertex's
   * limit is Hi / max(dHi_dt, 1e-9), with dHi_dt < 0 there), explicit
   * thickness and its boundary conditions */
  double *Hex = malloc(sizeof(double) * nV), *rate = malloc(sizeof(double) * nV);
  double dt_max = p->dt_ice_max;
  for (int vi = 0; vi < nV; ++vi) {
    double m = in->fraction_margin[vi] * (in->SMB[vi] + in->BMB[vi] - in->dHi_dt_target[vi]) + in->LMB[vi];
    rate[vi] = -divQ[vi] + m;
    double lim = p->dt_ice_max;
    if (in->Hi[vi] > p->Hi_min && rate[vi] < 0.0) lim = in->Hi[vi] / fmax(rate[vi], 1e-9);
    dt_max = fmin(dt_max, lim);
  }
  dt_max = fmax(p->dt_ice_min, dt_max);
  double dt_ex = fmin(dt, dt_max);
  for (int vi = 0; vi < nV; ++vi) Hex[vi] = fmax(0.0, in->Hi[vi] + rate[vi] * dt_ex);
  thickness_bc_explicit(nV, V, g, p->bc_H, in->mask_noice, in->Hb, in->SL, Hex);
  for (int vi = 0; vi < nV; ++vi) {
    if (in->BC_prescr_mask && in->BC_prescr_mask[vi] == 1) Hex[vi] = fmax(0.0, in->BC_prescr_Hi[vi]);
    if (in->mask_noice[vi]) Hex[vi] = 0.0;
  }
  /* the matrix boundary conditions apply the explicit ones again
   * (apply_ice_thickness_BC_matrix_domain_border) */
  thickness_bc_explicit(nV, V, g, p->bc_H, in->mask_noice, in->Hb, in->SL, Hex);
  /* stiffness matrix and load vector, calc_dHi_dt_semiimplicit :112-140 */
  double *A = malloc(sizeof(double) * (nV + nC)), *bb = malloc(sizeof(double) * nV);
  for (int vi = 0; vi < nV; ++vi) {
    for (int k = ptr[vi]; k < ptr[vi + 1]; ++k) A[k] = M[k] * dt * p->fs;
    A[ptr[vi]] += 1.0;
    double m = in->fraction_margin[vi] * (in->SMB[vi] + in->BMB[vi] - in->dHi_dt_target[vi]) + in->LMB[vi];
    bb[vi] = in->Hi[vi] - (dt * (1.0 - p->fs) * divQ[vi]) + fmax(-1.0 * in->Hi[vi], dt * m);
    Hi_next[vi] = in->Hi[vi] + dt * (in->dHi_dt_guess ? in->dHi_dt_guess[vi] : 0.0);
  }
  /* border, prescribed and no-ice rows (:165-315), in that order */
  for (int vi = 0; vi < nV; ++vi) {
    int fixed = 0;
    double val = 0.0;
    if (g->VBI[vi] > 0) { fixed = 1; val = Hex[vi]; }
    if (in->BC_prescr_mask && in->BC_prescr_mask[vi] == 1) { fixed = 1; val = in->BC_prescr_Hi[vi]; }
    if (in->mask_noice[vi]) { fixed = 1; val = 0.0; }
    if (!fixed) continue;
    for (int k = ptr[vi]; k < ptr[vi + 1]; ++k) A[k] = 0.0;
    A[ptr[vi]] = 1.0;
    bb[vi] = val;
    Hi_next[vi] = val;
  }
  int its = oracle_gmres_ilu0(nV, ptr, ind, A, bb, Hi_next, p->rtol, p->abstol, p->maxit, p->gmres_restart, 0, NULL);
  for (int vi = 0; vi < nV; ++vi) dHi_dt[vi] = (Hi_next[vi] - in->Hi[vi]) / dt;
  if (dt_max_out) *dt_max_out = dt_max;
  if (n_its) *n_its = its;
  free(ptr); free(ind); free(M); free(Hex); free(rate); free(A); free(bb);
  return its < 0 ? UFE_ERR_ARGUMENT : 0;
}

/* ---------------------------------------------------------------------- */
/* Heat-column inputs (thermodynamics_utilities.f90)                       */
/* ---------------------------------------------------------------------- */

static double cross2(double ax, double ay, double bx, double by) { return ax * by - ay * bx; }

/* calc_upwind_heat_flux_derivatives (thermodynamics_utilities.f90:339-434):
 * for each vertex with Hi >= 1 m, the first triangle in its (counter-
 * clockwise) iTri list whose corner at the vertex contains the upwind vector
 * (-u_vav, -v_vav); then per layer u_3D_b * ddx_a_b(Ti) and v_3D_b *
 * ddy_a_b(Ti) on that triangle.  Fields are [k][i], vertex/triangle fastest.
 * Returns -1 (the reference's crash) when no triangle qualifies. */
int oracle_heat_upwind(const ufe_mesh_desc *m, const int *iTri_ptr, const int *iTri_ind, const double *Ti,
                       const double *Hi, const double *u_vav, const double *v_vav, const double *u3, const double *v3,
                       double *u_dTdxp, double *v_dTdyp) {
  const int nV = m->nV, nT = m->nT, nz = m->nz;
  for (int vi = 0; vi < nV; ++vi) {
    if (Hi[vi] < 1.0) {
      for (int k = 0; k < nz; ++k) u_dTdxp[(size_t)k * nV + vi] = v_dTdyp[(size_t)k * nV + vi] = 0.0;
      continue;
    }
    const double ux = -u_vav[vi], uy = -v_vav[vi];
    int tup = -1;
    for (int q = iTri_ptr[vi]; q < iTri_ptr[vi + 1] && tup < 0; ++q) {
      const int ti = iTri_ind[q];
      int vb = -1, vc = -1;
      for (int n1 = 0; n1 < 3; ++n1)
        if (m->Tri[3 * ti + n1] == vi) {
          vb = m->Tri[3 * ti + (n1 + 1) % 3];
          vc = m->Tri[3 * ti + (n1 + 2) % 3];
          break;
        }
      const double abx = m->V[2 * vb] - m->V[2 * vi], aby = m->V[2 * vb + 1] - m->V[2 * vi + 1];
      const double acx = m->V[2 * vc] - m->V[2 * vi], acy = m->V[2 * vc + 1] - m->V[2 * vi + 1];
      if (cross2(abx, aby, ux, uy) >= 0.0 && cross2(ux, uy, acx, acy) >= 0.0) tup = ti;
    }
    if (tup < 0) return -1;
    for (int k = 0; k < nz; ++k) {
      double dx = 0.0, dy = 0.0;
      for (int q = m->ab_ptr[tup]; q < m->ab_ptr[tup + 1]; ++q) {
        const double t = Ti[(size_t)k * nV + m->ab_ind[q]];
        dx += m->ab_ddx[q] * t;
        dy += m->ab_ddy[q] * t;
      }
      u_dTdxp[(size_t)k * nV + vi] = u3[(size_t)k * nT + tup] * dx;
      v_dTdyp[(size_t)k * nV + vi] = v3[(size_t)k * nT + tup] * dy;
    }
  }
  return 0;
}

/* calc_strain_heating (thermodynamics_utilities.f90:24-91): Phi = 2 A^(-1/n)
 * D^(1/n + 1) with D the effective strain rate of the nine 3-D strain-rate
 * fields (order du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy,
 * dw_dz, each [k][vi]); no heating where Hi < 0.1 m. */
void oracle_strain_heating(int nV, int nz, const double *Hi, const double *A_flow, const double *const *e,
                           double glen_n, double *Phi) {
  for (int vi = 0; vi < nV; ++vi)
    for (int k = 0; k < nz; ++k) {
      const size_t i = (size_t)k * nV + vi;
      if (Hi[vi] < 0.1) { Phi[i] = 0.0; continue; }
      const double ux = e[0][i], uy = e[1][i], uz = e[2][i], vx = e[3][i], vy = e[4][i], vz = e[5][i];
      const double wx = e[6][i], wy = e[7][i], wz = e[8][i];
      const double D = sqrt(0.5 * (ux * ux + vy * vy + wz * wz + 0.5 * (uy + vx) * (uy + vx) +
                                   0.5 * (uz + wx) * (uz + wx) + 0.5 * (vz + wy) * (vz + wy)));
      Phi[i] = 2.0 * pow(A_flow[i], -1.0 / glen_n) * pow(D, 1.0 / glen_n + 1.0);
    }
}

/* calc_frictional_heating (thermodynamics_utilities.f90:93-128): the basal
 * friction coefficient times the basal speed where the ice is grounded, zero
 * elsewhere and everywhere without sliding. */
void oracle_frictional_heating(int nV, int no_sliding, const double *beta, const double *uabs_base,
                               const int *mask_grounded, double *Q) {
  for (int vi = 0; vi < nV; ++vi) Q[vi] = (!no_sliding && mask_grounded[vi]) ? beta[vi] * uabs_base[vi] : 0.0;
}

/* ====================================================================== */
/* BPA 3-D stress balance (BPA_main.f90) and what it runs on: the zeta     */
/* gradients (zeta_gradients.f90:19-212), the 1-D zeta operators           */
/* (mesh_zeta.f90:287-482) and the 3-D operators                           */
/* (mesh_disc_calc_matrix_operators_3D.f90).                               */
/* Internal 3-D layouts follow the reference's translation tables          */
/* (mesh_translation_tables.f90:120-240): vik = vi*nz+k, tik = ti*nz+k,    */
/* tiks = ti*(nz-1)+ks, tikuv = 2*tik+uv.  The public arrays are the       */
/* Fortran ones, first index fastest ([k*nV+vi], [k*nT+ti]).               */
/* ====================================================================== */

#define BPA_MAX_NZ 64

/* calc_shape_functions_1D_reg_2nd_order (shape_functions.f90:17-79) for two
 * neighbours: d/dz and d2/dz2 coefficients of the neighbours (fc, fcc) and of
 * the home point (minus their sums). */
static void shape_1d_2nd(double z, const double zc[2], double *fi, double *fii, double fc[2], double fcc[2]) {
  double dx[2], w[2], A[4] = {0, 0, 0, 0};
  for (int i = 0; i < 2; ++i) {
    dx[i] = zc[i] - z;
    w[i] = 1.0 / pow(fabs(dx[i]), 1.5);
    double w2 = w[i] * w[i];
    A[0] += w2 * dx[i] * dx[i];
    A[1] += w2 * dx[i] * 0.5 * dx[i] * dx[i];
    A[2] += w2 * 0.5 * dx[i] * dx[i] * dx[i];
    A[3] += w2 * 0.5 * dx[i] * dx[i] * 0.5 * dx[i] * dx[i];
  }
  double det = A[0] * A[3] - A[1] * A[2];
  double M[4] = {A[3] / det, -A[1] / det, -A[2] / det, A[0] / det};
  for (int i = 0; i < 2; ++i) {
    double w2 = w[i] * w[i];
    fc[i] = w2 * (M[0] * dx[i] + M[1] * 0.5 * dx[i] * dx[i]);
    fcc[i] = w2 * (M[2] * dx[i] + M[3] * 0.5 * dx[i] * dx[i]);
  }
  *fi = -(fc[0] + fc[1]);
  *fii = -(fcc[0] + fcc[1]);
}

typedef struct {
  int nz;
  const double *zeta, *zs;
  /* M_ddzeta_k_k_1D / M_d2dzeta2_k_k_1D (calc_vertical_operators_reg_1D,
   * mesh_zeta.f90:287-365): row k = home, then its two neighbours */
  int kc[BPA_MAX_NZ][3];
  double kd[BPA_MAX_NZ][3], kdd[BPA_MAX_NZ][3];
  /* M_map_k_ks_1D / M_ddzeta_k_ks_1D (:396-431): row ks, columns ks, ks+1 */
  double kkm[BPA_MAX_NZ][2], kkd[BPA_MAX_NZ][2];
  /* M_map_ks_k_1D / M_ddzeta_ks_k_1D (:433-480): row k, columns ksc[k][0..1] */
  int ksc[BPA_MAX_NZ][2];
  double ksm[BPA_MAX_NZ][2], ksd[BPA_MAX_NZ][2];
} zops_t;

static void zops_build(int nz, const double *zeta, const double *zs, zops_t *o) {
  o->nz = nz; o->zeta = zeta; o->zs = zs;
  for (int k = 0; k < nz; ++k) {
    int c[2];
    if (k == 0) { c[0] = 1; c[1] = 2; }
    else if (k == nz - 1) { c[0] = nz - 3; c[1] = nz - 2; }
    else { c[0] = k - 1; c[1] = k + 1; }
    double zc[2] = {zeta[c[0]], zeta[c[1]]}, fi, fii, fc[2], fcc[2];
    shape_1d_2nd(zeta[k], zc, &fi, &fii, fc, fcc);
    o->kc[k][0] = k; o->kd[k][0] = fi; o->kdd[k][0] = fii;
    for (int i = 0; i < 2; ++i) { o->kc[k][1 + i] = c[i]; o->kd[k][1 + i] = fc[i]; o->kdd[k][1 + i] = fcc[i]; }
  }
  for (int ks = 0; ks < nz - 1; ++ks) {
    double dz = zeta[ks + 1] - zeta[ks];
    double wlo = (zeta[ks + 1] - zs[ks]) / dz;
    o->kkm[ks][0] = wlo; o->kkm[ks][1] = 1.0 - wlo;
    o->kkd[ks][0] = -1.0 / dz; o->kkd[ks][1] = 1.0 / dz;
  }
  for (int k = 0; k < nz; ++k) {
    int lo, hi;
    if (k == 0) { lo = 0; hi = 1; }
    else if (k == nz - 1) { lo = nz - 3; hi = nz - 2; }
    else { lo = k - 1; hi = k; }
    double dz = zs[hi] - zs[lo];
    double wlo = (zs[hi] - zeta[k]) / dz;
    o->ksc[k][0] = lo; o->ksc[k][1] = hi;
    o->ksm[k][0] = wlo; o->ksm[k][1] = 1.0 - wlo;
    o->ksd[k][0] = -1.0 / dz; o->ksd[k][1] = 1.0 / dz;
  }
}

/* zeta gradients in the internal layouts */
typedef struct {
  double *ak[7];   /* dt, dx, dy, dz, d2x2, d2xy, d2y2  [vi*nz+k]      */
  double *bk[6];   /* dx, dy, dz, d2x2, d2xy, d2y2      [ti*nz+k]      */
  double *bks[6];  /* the same                          [ti*(nz-1)+ks] */
} zgrad_t;

static void zgrad_alloc(zgrad_t *g, int nV, int nT, int nz) {
  for (int i = 0; i < 7; ++i) g->ak[i] = calloc((size_t)nV * nz, sizeof(double));
  for (int i = 0; i < 6; ++i) g->bk[i] = calloc((size_t)nT * nz, sizeof(double));
  for (int i = 0; i < 6; ++i) g->bks[i] = calloc((size_t)nT * (nz - 1), sizeof(double));
}
static void zgrad_free(zgrad_t *g) {
  for (int i = 0; i < 7; ++i) free(g->ak[i]);
  for (int i = 0; i < 6; ++i) { free(g->bk[i]); free(g->bks[i]); }
}

typedef struct {
  int nnz;
  const int *ptr, *ind;
  const double *ddx, *ddy;
} aa_ops_t;

/* calc_zeta_gradients (zeta_gradients.f90:19-212) */
static void zeta_gradients(const ufe_mesh_desc *m, const aa_ops_t *aa, const zops_t *z, const double *Hi,
                           const double *Hs, const double *dHi_dt, const double *dHs_dt, zgrad_t *g) {
  const int nV = m->nV, nT = m->nT, nz = m->nz;
  const double *H[2] = {Hi, Hs};
  double *dx_a[2], *dy_a[2], *xx_a[2], *xy_a[2], *yy_a[2], *dx_b[2], *dy_b[2], *xx_b[2], *xy_b[2], *yy_b[2];
  double *Hb = malloc(sizeof(double) * nT);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_map, Hi, Hb);
  for (int f = 0; f < 2; ++f) {
    dx_a[f] = malloc(sizeof(double) * nV); dy_a[f] = malloc(sizeof(double) * nV);
    xx_a[f] = malloc(sizeof(double) * nV); xy_a[f] = malloc(sizeof(double) * nV);
    yy_a[f] = malloc(sizeof(double) * nV);
    dx_b[f] = malloc(sizeof(double) * nT); dy_b[f] = malloc(sizeof(double) * nT);
    xx_b[f] = malloc(sizeof(double) * nT); xy_b[f] = malloc(sizeof(double) * nT);
    yy_b[f] = malloc(sizeof(double) * nT);
    apply_csr(nV, aa->ptr, aa->ind, aa->ddx, H[f], dx_a[f]);
    apply_csr(nV, aa->ptr, aa->ind, aa->ddy, H[f], dy_a[f]);
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddx, H[f], dx_b[f]);
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddy, H[f], dy_b[f]);
    apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_ddx, dx_b[f], xx_a[f]);
    apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_ddy, dx_b[f], xy_a[f]);
    apply_csr(nV, m->ba_ptr, m->ba_ind, m->ba_ddy, dy_b[f], yy_a[f]);
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddx, dx_a[f], xx_b[f]);
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddy, dx_a[f], xy_b[f]);
    apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddy, dy_a[f], yy_b[f]);
  }
  for (int vi = 0; vi < nV; ++vi) {
    double h = fmax(10.0, Hi[vi]);
    for (int k = 0; k < nz; ++k) {
      double zt = z->zeta[k];
      size_t i = (size_t)vi * nz + k;
      double zx = (1.0 / h) * (dx_a[1][vi] - zt * dx_a[0][vi]);
      double zy = (1.0 / h) * (dy_a[1][vi] - zt * dy_a[0][vi]);
      g->ak[0][i] = (1.0 / h) * ((dHs_dt ? dHs_dt[vi] : 0.0) - zt * (dHi_dt ? dHi_dt[vi] : 0.0));
      g->ak[1][i] = zx;
      g->ak[2][i] = zy;
      g->ak[3][i] = -1.0 / h;
      g->ak[4][i] = (dx_a[0][vi] * -1.0 / h) * zx + (1.0 / h) * (xx_a[1][vi] - zt * xx_a[0][vi]);
      g->ak[5][i] = (dy_a[0][vi] * -1.0 / h) * zx + (1.0 / h) * (xy_a[1][vi] - zt * xy_a[0][vi]);
      g->ak[6][i] = (dy_a[0][vi] * -1.0 / h) * zy + (1.0 / h) * (yy_a[1][vi] - zt * yy_a[0][vi]);
    }
  }
  for (int ti = 0; ti < nT; ++ti) {
    double h = fmax(10.0, Hb[ti]);
    for (int pass = 0; pass < 2; ++pass) {
      int nl = pass == 0 ? nz : nz - 1;
      double **o = pass == 0 ? g->bk : g->bks;
      for (int k = 0; k < nl; ++k) {
        double zt = pass == 0 ? z->zeta[k] : z->zs[k];
        size_t i = (size_t)ti * nl + k;
        double zx = (1.0 / h) * (dx_b[1][ti] - zt * dx_b[0][ti]);
        double zy = (1.0 / h) * (dy_b[1][ti] - zt * dy_b[0][ti]);
        o[0][i] = zx;
        o[1][i] = zy;
        o[2][i] = -1.0 / h;
        o[3][i] = (dx_b[0][ti] * -1.0 / h) * zx + (1.0 / h) * (xx_b[1][ti] - zt * xx_b[0][ti]);
        o[4][i] = (dy_b[0][ti] * -1.0 / h) * zx + (1.0 / h) * (xy_b[1][ti] - zt * xy_b[0][ti]);
        o[5][i] = (dy_b[0][ti] * -1.0 / h) * zy + (1.0 / h) * (yy_b[1][ti] - zt * yy_b[0][ti]);
      }
    }
  }
  for (int f = 0; f < 2; ++f) {
    free(dx_a[f]); free(dy_a[f]); free(xx_a[f]); free(xy_a[f]); free(yy_a[f]);
    free(dx_b[f]); free(dy_b[f]); free(xx_b[f]); free(xy_b[f]); free(yy_b[f]);
  }
  free(Hb);
}

/* Growable CSR with up to 7 value arrays sharing one pattern. */
typedef struct {
  int rows, nv, nnz, cap;
  int *ptr, *ind;
  double *val[7];
} csr_t;

static void csr_init(csr_t *a, int rows, int nv, int cap) {
  a->rows = 0; a->nv = nv; a->nnz = 0; a->cap = cap > 16 ? cap : 16;
  a->ptr = malloc(sizeof(int) * ((size_t)rows + 1));
  a->ptr[0] = 0;
  a->ind = malloc(sizeof(int) * a->cap);
  for (int i = 0; i < nv; ++i) a->val[i] = malloc(sizeof(double) * a->cap);
}
static void csr_push(csr_t *a, int col, const double *v) {
  if (a->nnz == a->cap) {
    a->cap *= 2;
    a->ind = realloc(a->ind, sizeof(int) * a->cap);
    for (int i = 0; i < a->nv; ++i) a->val[i] = realloc(a->val[i], sizeof(double) * a->cap);
  }
  a->ind[a->nnz] = col;
  for (int i = 0; i < a->nv; ++i) a->val[i][a->nnz] = v[i];
  a->nnz++;
}
static void csr_endrow(csr_t *a) { a->ptr[++a->rows] = a->nnz; }
static void csr_free(csr_t *a) {
  free(a->ptr); free(a->ind);
  for (int i = 0; i < a->nv; ++i) free(a->val[i]);
  memset(a, 0, sizeof(*a));
}

/* The 3-D operators of calc_3D_matrix_operators_mesh
 * (mesh_disc_calc_matrix_operators_3D.f90:15-59) */
typedef struct {
  csr_t bk_ak;    /* ddx, ddy              (:61-215)   */
  csr_t ak_bk;    /* ddx, ddy              (:217-371)  */
  csr_t bk_bks;   /* ddz                   (:373-490)  */
  csr_t bks_bk;   /* map, ddz              (:492-618)  */
  csr_t bks_ak;   /* map                   (:620-724)  */
  csr_t ak_bks;   /* map                   (:726-830)  */
  csr_t bk_bk;    /* ddx, ddy, ddz, d2dx2, d2dxdy, d2dy2, d2dz2 (:832-1072) */
} ops3d_t;

static void ops3d_free(ops3d_t *o) {
  csr_free(&o->bk_ak); csr_free(&o->ak_bk); csr_free(&o->bk_bks); csr_free(&o->bks_bk);
  csr_free(&o->bks_ak); csr_free(&o->ak_bks); csr_free(&o->bk_bk);
}

static void ops3d_build(const ufe_mesh_desc *m, const zops_t *z, const zgrad_t *g, ops3d_t *o) {
  const int nV = m->nV, nT = m->nT, nz = m->nz;
  double v[7];
  csr_init(&o->bk_ak, nV * nz, 2, m->ba_nnz * nz * 3);
  for (int vi = 0; vi < nV; ++vi)
    for (int k = 0; k < nz; ++k) {
      size_t r = (size_t)vi * nz + k;
      double zx = g->ak[1][r], zy = g->ak[2][r];
      for (int e = m->ba_ptr[vi]; e < m->ba_ptr[vi + 1]; ++e) {
        int ti = m->ba_ind[e];
        for (int jj = 0; jj < 3; ++jj) {
          int kk = z->kc[k][jj];
          double cd = z->kd[k][jj];
          v[0] = (kk == k ? m->ba_ddx[e] : 0.0) + zx * m->ba_map[e] * cd;
          v[1] = (kk == k ? m->ba_ddy[e] : 0.0) + zy * m->ba_map[e] * cd;
          csr_push(&o->bk_ak, ti * nz + kk, v);
        }
      }
      csr_endrow(&o->bk_ak);
    }
  csr_init(&o->ak_bk, nT * nz, 2, m->ab_nnz * nz * 3);
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) {
      size_t r = (size_t)ti * nz + k;
      double zx = g->bk[0][r], zy = g->bk[1][r];
      for (int e = m->ab_ptr[ti]; e < m->ab_ptr[ti + 1]; ++e) {
        int vi = m->ab_ind[e];
        for (int jj = 0; jj < 3; ++jj) {
          int kk = z->kc[k][jj];
          double cd = z->kd[k][jj];
          v[0] = (kk == k ? m->ab_ddx[e] : 0.0) + zx * m->ab_map[e] * cd;
          v[1] = (kk == k ? m->ab_ddy[e] : 0.0) + zy * m->ab_map[e] * cd;
          csr_push(&o->ak_bk, vi * nz + kk, v);
        }
      }
      csr_endrow(&o->ak_bk);
    }
  csr_init(&o->bk_bks, nT * (nz - 1), 1, nT * (nz - 1) * 2);
  for (int ti = 0; ti < nT; ++ti)
    for (int ks = 0; ks < nz - 1; ++ks) {
      double zz = g->bks[2][(size_t)ti * (nz - 1) + ks];
      for (int jj = 0; jj < 2; ++jj) {
        v[0] = zz * z->kkd[ks][jj];
        csr_push(&o->bk_bks, ti * nz + ks + jj, v);
      }
      csr_endrow(&o->bk_bks);
    }
  csr_init(&o->bks_bk, nT * nz, 2, nT * nz * 2);
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) {
      double zz = g->bk[2][(size_t)ti * nz + k];
      for (int jj = 0; jj < 2; ++jj) {
        v[0] = z->ksm[k][jj];
        v[1] = zz * z->ksd[k][jj];
        csr_push(&o->bks_bk, ti * (nz - 1) + z->ksc[k][jj], v);
      }
      csr_endrow(&o->bks_bk);
    }
  csr_init(&o->bks_ak, nV * nz, 1, m->ba_nnz * nz * 2);
  for (int vi = 0; vi < nV; ++vi)
    for (int k = 0; k < nz; ++k) {
      for (int e = m->ba_ptr[vi]; e < m->ba_ptr[vi + 1]; ++e) {
        int ti = m->ba_ind[e];
        for (int jj = 0; jj < 2; ++jj) {
          v[0] = m->ba_map[e] * z->ksm[k][jj];
          csr_push(&o->bks_ak, ti * (nz - 1) + z->ksc[k][jj], v);
        }
      }
      csr_endrow(&o->bks_ak);
    }
  csr_init(&o->ak_bks, nT * (nz - 1), 1, m->ab_nnz * (nz - 1) * 2);
  for (int ti = 0; ti < nT; ++ti)
    for (int ks = 0; ks < nz - 1; ++ks) {
      for (int e = m->ab_ptr[ti]; e < m->ab_ptr[ti + 1]; ++e) {
        int vi = m->ab_ind[e];
        for (int jj = 0; jj < 2; ++jj) {
          v[0] = m->ab_map[e] * z->kkm[ks][jj];
          csr_push(&o->ak_bks, vi * nz + ks + jj, v);
        }
      }
      csr_endrow(&o->ak_bks);
    }
  csr_init(&o->bk_bk, nT * nz, 7, m->bb_nnz * nz * 3);
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) {
      size_t r = (size_t)ti * nz + k;
      double zx = g->bk[0][r], zy = g->bk[1][r], zz = g->bk[2][r];
      double zxx = g->bk[3][r], zxy = g->bk[4][r], zyy = g->bk[5][r];
      for (int e = m->bb_ptr[ti]; e < m->bb_ptr[ti + 1]; ++e) {
        int tj = m->bb_ind[e];
        double hx = m->bb_ddx[e], hy = m->bb_ddy[e];
        double hxx = m->bb_d2dx2[e], hxy = m->bb_d2dxdy[e], hyy = m->bb_d2dy2[e];
        for (int jj = 0; jj < 3; ++jj) {
          int kk = z->kc[k][jj];
          double cz = z->kd[k][jj], czz = z->kdd[k][jj];
          for (int i = 0; i < 7; ++i) v[i] = 0.0;
          if (kk == k) { v[0] += hx; v[1] += hy; v[3] += hxx; v[4] += hxy; v[5] += hyy; }
          if (tj == ti) {
            v[0] += zx * cz;
            v[1] += zy * cz;
            v[2] += zz * cz;
            v[3] += zxx * cz + zx * zx * czz;
            v[4] += zxy * cz + zx * zy * czz;
            v[5] += zyy * cz + zy * zy * czz;
            v[6] += zz * zz * czz;
          }
          v[3] += 2.0 * zx * hx * cz;
          v[4] += zx * hy * cz + zy * hx * cz;
          v[5] += 2.0 * zy * hy * cz;
          csr_push(&o->bk_bk, tj * nz + kk, v);
        }
      }
      csr_endrow(&o->bk_bk);
    }
}

static void csr_apply(const csr_t *a, int which, const double *x, double *y) {
  oracle_csr_matvec(a->rows, a->ptr, a->ind, a->val[which], x, y);
}

/* calc_zeta_gradients on the public layouts (first index fastest): for the
 * parity tests of the engine's zeta-gradient entry point.  ak [7][nV*nz],
 * bk [6][nT*nz], bks [6][nT*(nz-1)]; any may be NULL. */
int oracle_zeta_gradients(const ufe_mesh_desc *m, int aa_nnz, const int *aa_ptr, const int *aa_ind,
                          const double *aa_ddx, const double *aa_ddy, const double *zeta_stag, const double *Hi,
                          const double *Hs, const double *dHi_dt, const double *dHs_dt, double *const *ak,
                          double *const *bk, double *const *bks) {
  const int nV = m->nV, nT = m->nT, nz = m->nz;
  if (nz < 3 || nz > BPA_MAX_NZ) return -1;
  aa_ops_t aa = {aa_nnz, aa_ptr, aa_ind, aa_ddx, aa_ddy};
  zops_t z;
  zops_build(nz, m->zeta, zeta_stag, &z);
  zgrad_t g;
  zgrad_alloc(&g, nV, nT, nz);
  zeta_gradients(m, &aa, &z, Hi, Hs, dHi_dt, dHs_dt, &g);
  for (int f = 0; f < 7; ++f)
    if (ak && ak[f])
      for (int vi = 0; vi < nV; ++vi)
        for (int k = 0; k < nz; ++k) ak[f][(size_t)k * nV + vi] = g.ak[f][(size_t)vi * nz + k];
  for (int f = 0; f < 6; ++f) {
    if (bk && bk[f])
      for (int ti = 0; ti < nT; ++ti)
        for (int k = 0; k < nz; ++k) bk[f][(size_t)k * nT + ti] = g.bk[f][(size_t)ti * nz + k];
    if (bks && bks[f])
      for (int ti = 0; ti < nT; ++ti)
        for (int k = 0; k < nz - 1; ++k) bks[f][(size_t)k * nT + ti] = g.bks[f][(size_t)ti * (nz - 1) + k];
  }
  zgrad_free(&g);
  return 0;
}

/* The 3-D operators applied to fields, for the operator pins: with f_bk on
 * the bk grid and f_ak on the ak grid (public layouts), out_bk [7][nT*nz] =
 * bk->bk ddx, ddy, ddz, d2dx2, d2dxdy, d2dy2, d2dz2 of f_bk; out_ak [2][nV*nz]
 * = bk->ak ddx, ddy of f_bk; out_akbk [2][nT*nz] = ak->bk ddx, ddy of f_ak;
 * out_bks [1][nT*(nz-1)] = bk->bks ddz of f_bk; out_bksbk [2][nT*nz] = bks->bk
 * map, ddz of f_bks; out_map_bks_ak [nV*nz] of f_bks; out_map_ak_bks
 * [nT*(nz-1)] of f_ak.  Zeta gradients from Hi, Hs (no time derivative). */
int oracle_ops3d_apply(const ufe_mesh_desc *m, int aa_nnz, const int *aa_ptr, const int *aa_ind,
                       const double *aa_ddx, const double *aa_ddy, const double *zeta_stag, const double *Hi,
                       const double *Hs, const double *f_bk, const double *f_ak, const double *f_bks,
                       double *const *out_bk, double *const *out_ak, double *const *out_akbk, double *out_bks,
                       double *const *out_bksbk, double *out_map_bks_ak, double *out_map_ak_bks) {
  const int nV = m->nV, nT = m->nT, nz = m->nz;
  if (nz < 3 || nz > BPA_MAX_NZ) return -1;
  aa_ops_t aa = {aa_nnz, aa_ptr, aa_ind, aa_ddx, aa_ddy};
  zops_t z;
  zops_build(nz, m->zeta, zeta_stag, &z);
  zgrad_t g;
  zgrad_alloc(&g, nV, nT, nz);
  zeta_gradients(m, &aa, &z, Hi, Hs, NULL, NULL, &g);
  ops3d_t o;
  ops3d_build(m, &z, &g, &o);
  size_t nbk = (size_t)nT * nz, nak = (size_t)nV * nz, nbks = (size_t)nT * (nz - 1);
  double *xb = malloc(sizeof(double) * nbk), *xa = malloc(sizeof(double) * nak);
  double *xs = malloc(sizeof(double) * nbks);
  double *y = malloc(sizeof(double) * (nbk > nak ? nbk : nak));
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) xb[(size_t)ti * nz + k] = f_bk[(size_t)k * nT + ti];
  for (int vi = 0; vi < nV; ++vi)
    for (int k = 0; k < nz; ++k) xa[(size_t)vi * nz + k] = f_ak[(size_t)k * nV + vi];
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz - 1; ++k) xs[(size_t)ti * (nz - 1) + k] = f_bks[(size_t)k * nT + ti];
#define OUT_BK(dst) for (int ti = 0; ti < nT; ++ti) for (int k = 0; k < nz; ++k) (dst)[(size_t)k * nT + ti] = y[(size_t)ti * nz + k]
#define OUT_AK(dst) for (int vi = 0; vi < nV; ++vi) for (int k = 0; k < nz; ++k) (dst)[(size_t)k * nV + vi] = y[(size_t)vi * nz + k]
#define OUT_BKS(dst) for (int ti = 0; ti < nT; ++ti) for (int k = 0; k < nz - 1; ++k) (dst)[(size_t)k * nT + ti] = y[(size_t)ti * (nz - 1) + k]
  for (int i = 0; i < 7; ++i) { csr_apply(&o.bk_bk, i, xb, y); OUT_BK(out_bk[i]); }
  for (int i = 0; i < 2; ++i) { csr_apply(&o.bk_ak, i, xb, y); OUT_AK(out_ak[i]); }
  for (int i = 0; i < 2; ++i) { csr_apply(&o.ak_bk, i, xa, y); OUT_BK(out_akbk[i]); }
  csr_apply(&o.bk_bks, 0, xb, y); OUT_BKS(out_bks);
  for (int i = 0; i < 2; ++i) { csr_apply(&o.bks_bk, i, xs, y); OUT_BK(out_bksbk[i]); }
  csr_apply(&o.bks_ak, 0, xs, y); OUT_AK(out_map_bks_ak);
  csr_apply(&o.ak_bks, 0, xa, y); OUT_BKS(out_map_ak_bks);
#undef OUT_BK
#undef OUT_AK
#undef OUT_BKS
  free(xb); free(xa); free(xs); free(y);
  ops3d_free(&o);
  zgrad_free(&g);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* solve_BPA (BPA_main.f90:90-252)                                         */
/* ---------------------------------------------------------------------- */

typedef struct {
  const ufe_mesh_desc *m;
  const ufe_params *p;
  int nV, nT, nz;
  zops_t z;
  zgrad_t g;
  ops3d_t o;
  const double *Hi, *Hs, *rough, *fgr;
  double *A_ak, *A_bks, *Hib;
  double *dhx, *dhy, *dbx, *dby, *tdx, *tdy;
  double *sr_ak[6], *sr_bks[6];   /* du/dx, du/dy, du/dz, dv/dx, dv/dy, dv/dz */
  double *eta_ak, *eta_bks, *eta_bk, *dex, *dey, *dez, *tmp_bk;
  double *beta_a, *beta_b, *ub, *vb;
  double *u, *v, *up, *vp;        /* [ti*nz+k] */
  int *bcm;
  double *bcu, *bcv;
  copy_lists_t copy;
  csr_t A;
  double *b, *x;
} bpa_t;

static void bpa_free(bpa_t *s) {
  zgrad_free(&s->g);
  ops3d_free(&s->o);
  free(s->A_ak); free(s->A_bks); free(s->Hib);
  free(s->dhx); free(s->dhy); free(s->dbx); free(s->dby); free(s->tdx); free(s->tdy);
  for (int i = 0; i < 6; ++i) { free(s->sr_ak[i]); free(s->sr_bks[i]); }
  free(s->eta_ak); free(s->eta_bks); free(s->eta_bk); free(s->dex); free(s->dey); free(s->dez); free(s->tmp_bk);
  free(s->beta_a); free(s->beta_b); free(s->ub); free(s->vb);
  free(s->u); free(s->v); free(s->up); free(s->vp);
  free(s->bcm); free(s->bcu); free(s->bcv);
  copy_lists_free(&s->copy);
  if (s->A.ptr) csr_free(&s->A);
  free(s->b); free(s->x);
}

/* Set-up of one solve_BPA call: zeta gradients and 3-D operators for the
 * current geometry (:144-151), driving stress (calc_driving_stress,
 * :1727-1756), inputs into the internal layouts. */
static int bpa_setup(bpa_t *s, const ufe_mesh_desc *m, const ufe_params *p, const aa_ops_t *aa,
                     const double *zeta_stag, const double *Hi, const double *Hs, const double *A_flow,
                     const double *rough, const double *fgr, const int *bc_mask, const double *bc_u,
                     const double *bc_v) {
  memset(s, 0, sizeof(*s));
  const int nV = m->nV, nT = m->nT, nz = m->nz;
  if (nz < 3 || nz > BPA_MAX_NZ) return UFE_ERR_ARGUMENT;
  for (int side = 0; side < 4; ++side) {
    int cu = p->bc_u[side], cv = p->bc_v[side];
    if (cu == UFE_BC_INFINITE_SSA_ICESTREAM || cv == UFE_BC_INFINITE_SSA_ICESTREAM) return UFE_ERR_ARGUMENT;
  }
  s->m = m; s->p = p; s->nV = nV; s->nT = nT; s->nz = nz;
  s->Hi = Hi; s->Hs = Hs; s->rough = rough; s->fgr = fgr;
  size_t nak = (size_t)nV * nz, nbk = (size_t)nT * nz, nbks = (size_t)nT * (nz - 1);
  zops_build(nz, m->zeta, zeta_stag, &s->z);
  zgrad_alloc(&s->g, nV, nT, nz);
  zeta_gradients(m, aa, &s->z, Hi, Hs, NULL, NULL, &s->g);
  ops3d_build(m, &s->z, &s->g, &s->o);
  s->A_ak = malloc(sizeof(double) * nak);
  for (int vi = 0; vi < nV; ++vi)
    for (int k = 0; k < nz; ++k) s->A_ak[(size_t)vi * nz + k] = A_flow[(size_t)k * nV + vi];
  s->A_bks = malloc(sizeof(double) * nbks);
  csr_apply(&s->o.ak_bks, 0, s->A_ak, s->A_bks);
  s->Hib = malloc(sizeof(double) * nV);
  for (int vi = 0; vi < nV; ++vi) s->Hib[vi] = Hs[vi] - Hi[vi];
  s->dhx = malloc(sizeof(double) * nT); s->dhy = malloc(sizeof(double) * nT);
  s->dbx = malloc(sizeof(double) * nT); s->dby = malloc(sizeof(double) * nT);
  s->tdx = malloc(sizeof(double) * nT); s->tdy = malloc(sizeof(double) * nT);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddx, Hs, s->dhx);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddy, Hs, s->dhy);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddx, s->Hib, s->dbx);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_ddy, s->Hib, s->dby);
  for (int ti = 0; ti < nT; ++ti) {
    s->tdx[ti] = -ICE_DENSITY * GRAV * s->dhx[ti];
    s->tdy[ti] = -ICE_DENSITY * GRAV * s->dhy[ti];
  }
  for (int i = 0; i < 6; ++i) {
    s->sr_ak[i] = malloc(sizeof(double) * nak);
    s->sr_bks[i] = malloc(sizeof(double) * nbks);
  }
  s->eta_ak = malloc(sizeof(double) * nak);
  s->eta_bks = malloc(sizeof(double) * nbks);
  s->eta_bk = malloc(sizeof(double) * nbk); s->dex = malloc(sizeof(double) * nbk);
  s->dey = malloc(sizeof(double) * nbk); s->dez = malloc(sizeof(double) * nbk);
  s->tmp_bk = malloc(sizeof(double) * nbk);
  s->beta_a = malloc(sizeof(double) * nV); s->beta_b = malloc(sizeof(double) * nT);
  s->ub = malloc(sizeof(double) * nT); s->vb = malloc(sizeof(double) * nT);
  s->u = malloc(sizeof(double) * nbk); s->v = malloc(sizeof(double) * nbk);
  s->up = malloc(sizeof(double) * nbk); s->vp = malloc(sizeof(double) * nbk);
  if (bc_mask) {
    s->bcm = malloc(sizeof(int) * nbk);
    s->bcu = malloc(sizeof(double) * nbk);
    s->bcv = malloc(sizeof(double) * nbk);
    for (int ti = 0; ti < nT; ++ti)
      for (int k = 0; k < nz; ++k) {
        size_t i = (size_t)ti * nz + k, j = (size_t)k * nT + ti;
        s->bcm[i] = bc_mask[j]; s->bcu[i] = bc_u[j]; s->bcv[i] = bc_v[j];
      }
  }
  copy_lists_build(m, p, &s->copy);
  s->b = malloc(sizeof(double) * 2 * nbk);
  s->x = malloc(sizeof(double) * 2 * nbk);
  return 0;
}

/* calc_strain_rates (:1758-1804), calc_effective_viscosity (:1806-1931) and
 * calc_applied_basal_friction_coefficient (:1933-1978) at the state's u, v. */
static void bpa_state(bpa_t *s, double eps0_applied) {
  const ufe_mesh_desc *m = s->m;
  const ufe_params *p = s->p;
  const int nV = s->nV, nT = s->nT, nz = s->nz;
  size_t nak = (size_t)nV * nz, nbk = (size_t)nT * nz, nbks = (size_t)nT * (nz - 1);
  double *f[2] = {s->u, s->v};
  for (int c = 0; c < 2; ++c) {
    csr_apply(&s->o.bk_ak, 0, f[c], s->sr_ak[3 * c + 0]);
    csr_apply(&s->o.bk_ak, 1, f[c], s->sr_ak[3 * c + 1]);
    csr_apply(&s->o.bk_bks, 0, f[c], s->sr_bks[3 * c + 2]);
    csr_apply(&s->o.ak_bks, 0, s->sr_ak[3 * c + 0], s->sr_bks[3 * c + 0]);
    csr_apply(&s->o.ak_bks, 0, s->sr_ak[3 * c + 1], s->sr_bks[3 * c + 1]);
    csr_apply(&s->o.bks_ak, 0, s->sr_bks[3 * c + 2], s->sr_ak[3 * c + 2]);
  }
  const double n = p->glens_n;
  const double eta_max = 0.5 * pow(1e-18, -1.0 / n) * pow(eps0_applied, (1.0 - n) / (2.0 * n));
  for (size_t i = 0; i < nak; ++i) {
    double e = glen_3d(eps0_applied, s->sr_ak[0][i], s->sr_ak[1][i], s->sr_ak[2][i], s->sr_ak[3][i],
                       s->sr_ak[4][i], s->sr_ak[5][i], s->A_ak[i], n);
    s->eta_ak[i] = fmin(fmax(e, p->visc_eff_min), eta_max);
  }
  /* the staggered viscosity uses the configured regularisation (:1876-1880) */
  for (size_t i = 0; i < nbks; ++i) {
    double e = glen_3d(p->epsilon_sq_0, s->sr_bks[0][i], s->sr_bks[1][i], s->sr_bks[2][i], s->sr_bks[3][i],
                       s->sr_bks[4][i], s->sr_bks[5][i], s->A_bks[i], n);
    s->eta_bks[i] = fmin(fmax(e, p->visc_eff_min), eta_max);
  }
  csr_apply(&s->o.ak_bk, 0, s->eta_ak, s->dex);
  csr_apply(&s->o.ak_bk, 1, s->eta_ak, s->dey);
  csr_apply(&s->o.bks_bk, 1, s->eta_bks, s->dez);
  /* map_a_b_3D of eta_ak and map_bks_bk of eta_bks, blended by the shear factor */
  csr_apply(&s->o.bks_bk, 0, s->eta_bks, s->tmp_bk);
  for (int ti = 0; ti < nT; ++ti) {
    size_t t0 = (size_t)ti * nz, tb = t0 + nz - 1;
    double us = sqrt(0.1 + s->u[t0] * s->u[t0] + s->v[t0] * s->v[t0]);
    double ubs = sqrt(0.1 + s->u[tb] * s->u[tb] + s->v[tb] * s->v[tb]);
    double R = ubs / us;
    for (int k = 0; k < nz; ++k) {
      double ea = 0.0;
      for (int e = m->ab_ptr[ti]; e < m->ab_ptr[ti + 1]; ++e) ea += m->ab_map[e] * s->eta_ak[(size_t)m->ab_ind[e] * nz + k];
      s->eta_bk[t0 + k] = R * ea + (1.0 - R) * s->tmp_bk[t0 + k];
    }
  }
  for (int ti = 0; ti < nT; ++ti) { s->ub[ti] = s->u[(size_t)ti * nz + nz - 1]; s->vb[ti] = s->v[(size_t)ti * nz + nz - 1]; }
  sliding_beta(m, p, s->ub, s->vb, s->rough, s->beta_a);
  apply_csr(nT, m->ab_ptr, m->ab_ind, m->ab_map, s->beta_a, s->beta_b);
  if (p->do_gl_subgrid_friction)
    for (int ti = 0; ti < nT; ++ti) s->beta_b[ti] *= pow(s->fgr[ti], p->subgrid_friction_exponent);
  (void)nbk;
}

/* solve_BPA_linearised's assembly (:330-478) with the row routines
 * calc_BPA_stiffness_matrix_row_free (:480-644), _BC_surf (:646-889),
 * _BC_base (:891-1163) and _BC_west/east/south/north (:1165-1723).
 * Returns 0 or UFE_ERR_ARGUMENT for a border with no neighbour. */
static int bpa_assemble(bpa_t *s) {
  const ufe_mesh_desc *m = s->m;
  const ufe_params *p = s->p;
  const int nT = s->nT, nz = s->nz;
  const csr_t *B = &s->o.bk_bk;
  if (s->A.ptr) csr_free(&s->A);
  csr_init(&s->A, 2 * nT * nz, 1, B->nnz * 2 * 2);
  double val;
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k)
      for (int uv = 0; uv < 2; ++uv) {
        const size_t tik = (size_t)ti * nz + k;
        const int row = (int)(2 * tik + uv);
        double rhs = 0.0;
        if (s->bcm && s->bcm[tik] == 1) {
          val = 1.0; csr_push(&s->A, row, &val);
          rhs = uv == 0 ? s->bcu[tik] : s->bcv[tik];
        } else if (m->TriBI[ti] > 0) {
          const int side = bc_side(m->TriBI[ti]);
          const int choice = uv == 0 ? p->bc_u[side] : p->bc_v[side];
          if (choice == UFE_BC_INFINITE) {
            int nn = 0;
            for (int c = 0; c < 3; ++c) {
              int tj = m->TriC[3 * ti + c];
              if (tj < 0) continue;
              ++nn;
              val = 1.0; csr_push(&s->A, (int)(2 * ((size_t)tj * nz + k) + uv), &val);
            }
            if (nn == 0) return UFE_ERR_ARGUMENT;
            val = -(double)nn; csr_push(&s->A, row, &val);
          } else if (choice == UFE_BC_ZERO) {
            val = 1.0; csr_push(&s->A, row, &val);
          } else if (choice == UFE_BC_PERIODIC_ISMIP_HOM) {
            val = 1.0; csr_push(&s->A, row, &val);
            const double *f = uv == 0 ? s->up : s->vp;
            const int cr = 2 * ti + uv;
            double cv = 0.0;
            for (int e = s->copy.ptr[cr]; e < s->copy.ptr[cr + 1]; ++e)
              cv += s->copy.w[e] * f[(size_t)s->copy.tri[e] * nz + k];
            rhs = (p->visc_it_relax * cv) + ((1.0 - p->visc_it_relax) * f[tik]);
          } else {
            return UFE_ERR_ARGUMENT;
          }
        } else {
          const double tdx = s->tdx[ti], tdy = s->tdy[ti];
          double eta, dex = s->dex[tik], dey = s->dey[tik], dez = s->dez[tik];
          /* coefficients of du/dx, du/dy, dv/dx, dv/dy, d2u/dx2, d2u/dy2, d2v/dxdy
           * (x-row) or of dv/dy, dv/dx, du/dy, du/dx, d2v/dy2, d2v/dx2, d2u/dxdy
           * (y-row), plus the column terms of the own triangle */
          double c_own_x = 0, c_own_y = 0, c_own_xx = 0, c_own_yy = 0, c_oth_x = 0, c_oth_y = 0, c_oth_xy = 0;
          double c_k = 0, c_kn = 0;
          int kn = -1, free_row = 0;
          if (k == 0) {
            eta = s->eta_bks[(size_t)ti * (nz - 1)];
            const double hx = s->dhx[ti], hy = s->dhy[ti];
            const double zz = s->g.bk[2][tik], dz = m->zeta[1] - m->zeta[0];
            const double q = 2.0 * eta / dz * zz;
            const double hs = uv == 0 ? hx : hy, ho = uv == 0 ? hy : hx;   /* own / other direction */
            const double ds = uv == 0 ? dex : dey, dox = uv == 0 ? dey : dex;
            c_own_x = 4.0 * ds - q * 4.0 * hs + 4.0 * dez * hs;   /* d(own)/d(own dir) */
            c_own_y = dox - q * ho + dez * ho;                   /* d(own)/d(other dir) */
            c_oth_y = 2.0 * ds - q * 2.0 * hs + 2.0 * dez * hs;  /* d(other)/d(other dir) */
            c_oth_x = dox - q * ho + dez * ho;                   /* d(other)/d(own dir) */
            c_own_xx = 4.0 * eta; c_own_yy = eta; c_oth_xy = 3.0 * eta;
            c_k = -2.0 * eta / (dz * dz) * zz * zz;
            c_kn = 2.0 * eta / (dz * dz) * zz * zz;
            kn = 1;
          } else if (k == nz - 1) {
            if (p->sliding_law == UFE_SLIDING_NO_SLIDING) {
              val = 1.0; csr_push(&s->A, row, &val);
              s->b[row] = 0.0;
              csr_endrow(&s->A);
              continue;
            }
            eta = s->eta_bks[(size_t)ti * (nz - 1) + nz - 2];
            const double bx = s->dbx[ti], by = s->dby[ti];
            const double zz = s->g.bk[2][tik], dz = m->zeta[nz - 1] - m->zeta[nz - 2];
            const double Q = 2.0 * eta / (dz * dz) * zz * zz;
            const double R = 2.0 * eta / dz * zz + dez;
            const double bs = uv == 0 ? bx : by, bo = uv == 0 ? by : bx;
            const double ds = uv == 0 ? dex : dey, dox = uv == 0 ? dey : dex;
            c_own_x = 4.0 * ds + 4.0 * R * bs;
            c_own_y = dox + R * bo;
            c_oth_y = 2.0 * ds + 2.0 * R * bs;
            c_oth_x = dox + R * bo;
            c_own_xx = 4.0 * eta; c_own_yy = eta; c_oth_xy = 3.0 * eta;
            c_k = (R * s->beta_b[ti] / eta) - Q;
            c_kn = Q;
            kn = nz - 2;
          } else {
            eta = s->eta_bk[tik];
            free_row = 1;
          }
          for (int e = B->ptr[tik]; e < B->ptr[tik + 1]; ++e) {
            const int col = B->ind[e], tj = col / nz, kk = col % nz;
            const double ox = B->val[0][e], oy = B->val[1][e], oz = B->val[2][e];
            const double oxx = B->val[3][e], oxy = B->val[4][e], oyy = B->val[5][e], ozz = B->val[6][e];
            /* own-direction and other-direction first/second derivatives */
            const double os = uv == 0 ? ox : oy, oo = uv == 0 ? oy : ox;
            const double oss = uv == 0 ? oxx : oyy, ooo = uv == 0 ? oyy : oxx;
            double a_own, a_oth;
            if (free_row) {
              const double ds = uv == 0 ? dex : dey, dox = uv == 0 ? dey : dex;
              a_own = 4.0 * eta * oss + 4.0 * ds * os + eta * ooo + dox * oo + eta * ozz + dez * oz;
              a_oth = 3.0 * eta * oxy + 2.0 * ds * oo + dox * os;
            } else {
              a_own = c_own_x * os + c_own_y * oo + c_own_xx * oss + c_own_yy * ooo;
              if (tj == ti && kk == k) a_own += c_k;
              if (tj == ti && kk == kn) a_own += c_kn;
              a_oth = c_oth_y * oo + c_oth_x * os + c_oth_xy * oxy;
            }
            const int col_own = (int)(2 * (size_t)col + uv), col_oth = (int)(2 * (size_t)col + (1 - uv));
            csr_push(&s->A, col_own, &a_own);
            csr_push(&s->A, col_oth, &a_oth);
          }
          rhs = uv == 0 ? -tdx : -tdy;
        }
        s->b[row] = rhs;
        csr_endrow(&s->A);
      }
  return 0;
}

static double bpa_residual_norm(const bpa_t *s, double *r_out) {
  const int n = 2 * s->nT * s->nz;
  double *ax = malloc(sizeof(double) * n);
  oracle_csr_matvec(n, s->A.ptr, s->A.ind, s->A.val[0], s->x, ax);
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = s->b[i] - ax[i];
    if (r_out) r_out[i] = r;
    rr += r * r;
    bb += s->b[i] * s->b[i];
  }
  free(ax);
  return sqrt(rr) / fmax(sqrt(bb), 1e-300);
}

static void bpa_load_x(bpa_t *s) {
  size_t nbk = (size_t)s->nT * s->nz;
  for (size_t i = 0; i < nbk; ++i) { s->x[2 * i] = s->u[i]; s->x[2 * i + 1] = s->v[i]; }
}

/* solve_BPA (:90-252).  u_bk, v_bk [nT*nz] (first index fastest): in the
 * initial velocity, out the solution.  The a->a operators and zeta_stag are
 * the mesh's (calc_zeta_gradients needs them). */
int oracle_solve_bpa(const ufe_mesh_desc *m, const ufe_params *p, int aa_nnz, const int *aa_ptr,
                     const int *aa_ind, const double *aa_ddx, const double *aa_ddy, const double *zeta_stag,
                     const double *Hi, const double *Hs, const double *A_flow, const double *beta_sq,
                     const double *fgr_b, const int *mask_grounded, const int *bc_mask, const double *bc_u,
                     const double *bc_v, double *u_bk, double *v_bk, ufe_solve_info *info) {
  const int nT = m->nT, nz = m->nz;
  const size_t nbk = (size_t)nT * nz;
  if (!any_grounded(m->nV, mask_grounded)) {
    memset(u_bk, 0, sizeof(double) * nbk);
    memset(v_bk, 0, sizeof(double) * nbk);
    if (info) memset(info, 0, sizeof(*info));
    return 0;
  }
  double t0 = now_s();
  aa_ops_t aa = {aa_nnz, aa_ptr, aa_ind, aa_ddx, aa_ddy};
  bpa_t s;
  int rc = bpa_setup(&s, m, p, &aa, zeta_stag, Hi, Hs, A_flow, beta_sq, fgr_b, bc_mask, bc_u, bc_v);
  if (rc) { bpa_free(&s); return rc; }
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) {
      s.u[(size_t)ti * nz + k] = u_bk[(size_t)k * nT + ti];
      s.v[(size_t)ti * nz + k] = v_bk[(size_t)k * nT + ti];
    }
  double resid = 1e9, resid_prev;
  int nit_div = 0, it = 0, n_axb = 0, converged = 0;
  double relax = p->visc_it_relax, eps0 = p->epsilon_sq_0;
  while (!converged) {
    ++it;
    bpa_state(&s, eps0);
    /* solve_BPA_linearised: previous solution, system, solve */
    memcpy(s.up, s.u, sizeof(double) * nbk);
    memcpy(s.vp, s.v, sizeof(double) * nbk);
    rc = bpa_assemble(&s);
    if (rc) break;
    bpa_load_x(&s);
    int its = oracle_gmres_ilu0((int)(2 * nbk), s.A.ptr, s.A.ind, s.A.val[0], s.b, s.x, p->rtol, p->abstol,
                                p->krylov_maxit, p->gmres_restart, !p->warm_start, NULL);
    if (its < 0) { rc = UFE_ERR_ARGUMENT; break; }
    n_axb += its;
    for (size_t i = 0; i < nbk; ++i) { s.u[i] = s.x[2 * i]; s.v[i] = s.x[2 * i + 1]; }
    /* apply_velocity_limits (:2051-2084) */
    for (size_t i = 0; i < nbk; ++i) {
      double ua = sqrt(s.u[i] * s.u[i] + s.v[i] * s.v[i]);
      if (ua > p->vel_max) { s.u[i] = s.u[i] * p->vel_max / ua; s.v[i] = s.v[i] * p->vel_max / ua; }
    }
    /* relax_viscosity_iterations (:1982-2005) */
    for (size_t i = 0; i < nbk; ++i) {
      s.u[i] = relax * s.u[i] + (1.0 - relax) * s.up[i];
      s.v[i] = relax * s.v[i] + (1.0 - relax) * s.vp[i];
    }
    /* calc_visc_iter_UV_resid (:2007-2049) */
    resid_prev = resid;
    double r1 = 0, r2 = 0;
    for (size_t i = 0; i < nbk; ++i) {
      r1 += (s.u[i] - s.up[i]) * (s.u[i] - s.up[i]) + (s.v[i] - s.vp[i]) * (s.v[i] - s.vp[i]);
      r2 += (s.u[i] + s.up[i]) * (s.u[i] + s.up[i]) + (s.v[i] + s.vp[i]) * (s.v[i] + s.vp[i]);
    }
    resid = 2.0 * r1 / fmax(r2, 1e-8);
    if (resid > resid_prev) ++nit_div; else nit_div = 0;
    if (nit_div > 2) { nit_div = 0; relax *= 0.9; eps0 *= 1.2; }
    if (relax <= 0.05 || eps0 >= 1e-5) {
      if (relax < 0.05 || eps0 > 1e-5) { rc = UFE_ERR_DIVERGED; break; }
    }
    if (resid < p->visc_it_norm_dUV_tol) converged = 1;
    if (it > p->visc_it_nit) { rc = UFE_WARN_NOT_CONVERGED; break; }
  }
  if (rc >= 0) {
    for (int ti = 0; ti < nT; ++ti)
      for (int k = 0; k < nz; ++k) {
        u_bk[(size_t)k * nT + ti] = s.u[(size_t)ti * nz + k];
        v_bk[(size_t)k * nT + ti] = s.v[(size_t)ti * nz + k];
      }
  }
  if (info) {
    double rel = -1.0;
    if (rc >= 0) {
      memcpy(s.up, s.u, sizeof(double) * nbk);
      memcpy(s.vp, s.v, sizeof(double) * nbk);
      bpa_state(&s, eps0);
      if (bpa_assemble(&s) == 0) { bpa_load_x(&s); rel = bpa_residual_norm(&s, NULL); }
    }
    memset(info, 0, sizeof(*info));
    info->n_visc_its = it;
    info->n_Axb_its = n_axb;
    info->L2_uv = resid;
    info->rel_residual = rel;
    info->rel_residual_unscaled = rel;
    info->visc_it_relax_applied = relax;
    info->epsilon_sq_0_applied = eps0;
    info->seconds = now_s() - t0;
  }
  bpa_free(&s);
  return rc;
}

/* One evaluation of the viscosity iteration's state at (u_bk, v_bk) with the
 * linearised system assembled there and applied to (u, v): r = b - A u and b
 * in the reference row order (tikuv = 2*(ti*nz+k)+uv), eta_bk [nT*nz] (first
 * index fastest) and beta_b [nT]; returns ||r||/||b|| (negative on error).
 * The lagged copy rows read (u_bk, v_bk) as the previous iterate. */
double oracle_bpa_residual(const ufe_mesh_desc *m, const ufe_params *p, int aa_nnz, const int *aa_ptr,
                           const int *aa_ind, const double *aa_ddx, const double *aa_ddy,
                           const double *zeta_stag, const double *Hi, const double *Hs, const double *A_flow,
                           const double *beta_sq, const double *fgr_b, const int *bc_mask, const double *bc_u,
                           const double *bc_v, const double *u_bk, const double *v_bk, double *r, double *b,
                           double *eta_bk, double *beta_b) {
  const int nT = m->nT, nz = m->nz;
  const size_t nbk = (size_t)nT * nz;
  aa_ops_t aa = {aa_nnz, aa_ptr, aa_ind, aa_ddx, aa_ddy};
  bpa_t s;
  if (bpa_setup(&s, m, p, &aa, zeta_stag, Hi, Hs, A_flow, beta_sq, fgr_b, bc_mask, bc_u, bc_v)) {
    bpa_free(&s);
    return -1.0;
  }
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) {
      s.u[(size_t)ti * nz + k] = u_bk[(size_t)k * nT + ti];
      s.v[(size_t)ti * nz + k] = v_bk[(size_t)k * nT + ti];
    }
  memcpy(s.up, s.u, sizeof(double) * nbk);
  memcpy(s.vp, s.v, sizeof(double) * nbk);
  bpa_state(&s, p->epsilon_sq_0);
  double rel = -1.0;
  if (bpa_assemble(&s) == 0) {
    bpa_load_x(&s);
    rel = bpa_residual_norm(&s, r);
    if (b) memcpy(b, s.b, sizeof(double) * 2 * nbk);
    if (eta_bk)
      for (int ti = 0; ti < nT; ++ti)
        for (int k = 0; k < nz; ++k) eta_bk[(size_t)k * nT + ti] = s.eta_bk[(size_t)ti * nz + k];
    if (beta_b) memcpy(beta_b, s.beta_b, sizeof(double) * nT);
  }
  bpa_free(&s);
  return rel;
}

/* The BPA system assembled at (u_bk, v_bk) (first index fastest), in the
 * reference's row order tikuv = 2*(ti*nz+k)+uv: CSR into ptr [2*nT*nz+1],
 * ind/val [cap], b [2*nT*nz].  Returns nnz, or -1 if cap is too small or the
 * assembly fails. */
int oracle_bpa_system(const ufe_mesh_desc *m, const ufe_params *p, int aa_nnz, const int *aa_ptr,
                      const int *aa_ind, const double *aa_ddx, const double *aa_ddy, const double *zeta_stag,
                      const double *Hi, const double *Hs, const double *A_flow, const double *beta_sq,
                      const double *fgr_b, const int *bc_mask, const double *bc_u, const double *bc_v,
                      const double *u_bk, const double *v_bk, int cap, int *ptr, int *ind, double *val,
                      double *b) {
  const int nT = m->nT, nz = m->nz;
  const size_t nbk = (size_t)nT * nz;
  aa_ops_t aa = {aa_nnz, aa_ptr, aa_ind, aa_ddx, aa_ddy};
  bpa_t s;
  if (bpa_setup(&s, m, p, &aa, zeta_stag, Hi, Hs, A_flow, beta_sq, fgr_b, bc_mask, bc_u, bc_v)) {
    bpa_free(&s);
    return -1;
  }
  for (int ti = 0; ti < nT; ++ti)
    for (int k = 0; k < nz; ++k) {
      s.u[(size_t)ti * nz + k] = u_bk[(size_t)k * nT + ti];
      s.v[(size_t)ti * nz + k] = v_bk[(size_t)k * nT + ti];
    }
  memcpy(s.up, s.u, sizeof(double) * nbk);
  memcpy(s.vp, s.v, sizeof(double) * nbk);
  bpa_state(&s, p->epsilon_sq_0);
  int nnz = -1;
  if (bpa_assemble(&s) == 0 && s.A.nnz <= cap) {
    nnz = s.A.nnz;
    memcpy(ptr, s.A.ptr, sizeof(int) * (2 * nbk + 1));
    memcpy(ind, s.A.ind, sizeof(int) * nnz);
    memcpy(val, s.A.val[0], sizeof(double) * nnz);
    memcpy(b, s.b, sizeof(double) * 2 * nbk);
  }
  bpa_free(&s);
  return nnz;
}
