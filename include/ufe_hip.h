/*
 * ufe_hip.h -- C ABI of the MI355X engine for UFEMISM's velocity solve and
 * vertical heat columns (libufe_hip.so).
 *
 * Conventions
 *   - Every entry point returns int: 0 = ok, > 0 = completed with a warning
 *     (UFE_WARN_*), < 0 = error (UFE_ERR_*); ufe_last_error() gives the text.
 *   - Arrays are plain host pointers owned by the caller.  Indices are 0-based
 *     (the Fortran shim converts).  Two-dimensional (vertex or triangle, layer)
 *     fields are stored with the vertex/triangle index fastest, which is
 *     Fortran's (nV, nz) column-major layout.
 *   - Calls are synchronous.  One context per host thread; a context is bound
 *     to the device it was created on.
 *   - Units follow the reference: m, yr, Pa, K, J.
 */
#ifndef UFE_HIP_H
#define UFE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define UFE_OK                     0
#define UFE_WARN_NOT_CONVERGED     1  /* Picard hit visc_it_nit (reference warns and exits the loop) */
#define UFE_ERR_ARGUMENT          -1
#define UFE_ERR_DEVICE            -2  /* HIP runtime error */
#define UFE_ERR_DIVERGED          -3  /* Picard diverged below relax 0.05 or above eps0^2 1e-5 (reference crash) */
#define UFE_ERR_UNSUPPORTED       -4
#define UFE_ERR_UNSTABLE          -5  /* heat columns unstable for >= 1% of vertices (reference crash) */
#define UFE_ERR_MESH              -6  /* operator construction could not find a solvable neighbourhood */

/* ---- choices ----------------------------------------------------------- */
#define UFE_SLIDING_NO_SLIDING     0
#define UFE_SLIDING_WEERTMAN       1
#define UFE_SLIDING_BUDD           2  /* needs till_yield_stress in place of beta_sq */
#define UFE_SLIDING_COULOMB        3  /* needs till_yield_stress in place of beta_sq */
#define UFE_SLIDING_TSAI2015       4  /* beta_sq, plus the cap alpha_sq * N of ufe_set_sliding_cap */
#define UFE_SLIDING_SCHOOF2005     5  /* beta_sq, plus the cap alpha_sq * N of ufe_set_sliding_cap */
#define UFE_SLIDING_ZOET_IVERSON   6  /* needs till_yield_stress in place of beta_sq */
/* The reference's 'idealised' laws need no choice of their own: ISMIP-HOM C, D
 * and F prescribe a velocity-independent beta (sliding_laws.f90:499-571), which
 * is WEERTMAN with slid_weertman_m = 1 and beta_sq = that field (uabs^0 = 1
 * exactly); 'SSA_icestream' (:462-497) is COULOMB with the analytical till
 * yield stress of Schoof (2006). */

#define UFE_BC_INFINITE            0
#define UFE_BC_ZERO                1
/* Lagged copy rows of the idealised experiments (solve_linearised_SSA_DIVA.f90:543-575,
 * :618-635): the border velocity is set to visc_it_relax * (weighted mean of the
 * previous iterate around a copy point) + (1 - visc_it_relax) * its previous value.
 * 'periodic_ISMIP-HOM': copy point = geometric centre shifted by ismip_hom_L / 2 towards
 * the domain centre in x and y (u and v).  'infinite_SSA_icestream': same y, x at one
 * third (west half) or two thirds (east half) of the domain (u only, as in the
 * reference).  Weights 1/d^2 to the triangles around the vertex nearest the copy point. */
#define UFE_BC_PERIODIC_ISMIP_HOM      2
#define UFE_BC_INFINITE_SSA_ICESTREAM  3

#define UFE_PREC_FP64              0  /* plain form: FP64 Krylov, reference tolerances */
#define UFE_PREC_MIXED             1  /* FP32 Krylov inner solves, FP64 refinement residual */

#define UFE_KRYLOV_BICGSTAB        0
#define UFE_KRYLOV_GMRES           1  /* restarted GMRES(gmres_restart), right-preconditioned (PETSc's default KSP
                                         is GMRES(30); its default ILU(0) preconditioner is not provided) */

#define UFE_PC_NONE                0
#define UFE_PC_JACOBI              1  /* point Jacobi */
#define UFE_PC_BLOCK_JACOBI        2  /* 2x2 (u,v) blocks per triangle */
#define UFE_PC_AMG                 3  /* aggregation multigrid V-cycle (coarse correction + block-Jacobi smoothing) */

#define UFE_SMOOTHER_BLOCK_JACOBI     0  /* damped 2x2 block Jacobi */
#define UFE_SMOOTHER_L1_BLOCK_JACOBI  1  /* damped l1 block Jacobi: D_i + diag(row sums of |off-diagonal blocks|) */

#define UFE_STOP_L2UV              0  /* reference stop test: L2_uv < visc_it_norm_dUV_tol */
#define UFE_STOP_RESIDUAL          1  /* FP64 block-scaled nonlinear residual (see rel_residual) <= nonlinear_rtol */

#define UFE_STRESS_SSA             0
#define UFE_STRESS_DIVA            1

/* ---- mesh and operators ------------------------------------------------
 * Operator families share one sparsity pattern (ptr, ind) each:
 *   ab : map/ddx/ddy a->b  (nT rows, nV cols)  mesh%M_{map,ddx,ddy}_a_b
 *   ba : map/ddx/ddy b->a  (nV rows, nT cols)  mesh%M_{map,ddx,ddy}_b_a
 *   bb : ddx/ddy/d2dx2/d2dxdy/d2dy2 b->b, 2nd order (nT rows, nT cols)
 *        mesh%M2_*_b_b
 * (mesh_types.f90:180-192; built by mesh_disc_calc_matrix_operators_2D.f90)
 */
typedef struct ufe_mesh_desc {
  int nV, nT, nz;
  const double *V;        /* [nV*2]  x,y interleaved                 mesh%V      */
  const int    *Tri;      /* [nT*3]  CCW vertices                     mesh%Tri    */
  const int    *TriC;     /* [nT*3]  neighbour across vertex n or -1  mesh%TriC   */
  const int    *TriBI;    /* [nT]    border index 0..8                mesh%TriBI  */
  const double *TriGC;    /* [nT*2]  geometric centres                mesh%TriGC  */
  const double *zeta;     /* [nz]    scaled vertical coordinate       mesh%zeta   */
  int ab_nnz; const int *ab_ptr, *ab_ind; const double *ab_map, *ab_ddx, *ab_ddy;
  int ba_nnz; const int *ba_ptr, *ba_ind; const double *ba_map, *ba_ddx, *ba_ddy;
  int bb_nnz; const int *bb_ptr, *bb_ind;
  const double *bb_ddx, *bb_ddy, *bb_d2dx2, *bb_d2dxdy, *bb_d2dy2;
} ufe_mesh_desc;

/* Config fields of the reference's C% struct that the velocity solve reads
 * (model_configuration.f90:274-349,577-601), plus the engine's method
 * choices.  ufe_params_default() fills the reference defaults. */
typedef struct ufe_params {
  double glens_n;                    /* Glens_flow_law_exponent          3      */
  double epsilon_sq_0;               /* Glens_flow_law_epsilon_sq_0      1e-8   */
  double visc_eff_min;               /* visc_eff_min                     1e4    */
  double vel_max;                    /* vel_max                          5000   */
  int    visc_it_nit;                /* visc_it_nit                      50     */
  double visc_it_relax;              /* visc_it_relax                    0.2    */
  double visc_it_norm_dUV_tol;       /* visc_it_norm_dUV_tol             5e-5   */
  double rtol;                       /* stress_balance_PETSc_rtol        1e-7   */
  double abstol;                     /* stress_balance_PETSc_abstol      1e-5   (block-scaled residual, m/yr) */
  int    include_crossterms;         /* do_include_SSADIVA_crossterms    1      */
  int    sliding_law;                /* choice_sliding_law               WEERTMAN */
  double slid_weertman_m;            /* slid_Weertman_m                  3      */
  double slid_budd_q_plastic;        /* slid_Budd_q_plastic              0.3    */
  double slid_budd_u_threshold;      /* slid_Budd_u_threshold            100    */
  double slid_delta_v;               /* slid_delta_v                     1e-3   */
  double slid_beta_max;              /* slid_beta_max                    1e20   */
  int    do_gl_subgrid_friction;     /* do_GL_subgrid_friction           1      */
  double subgrid_friction_exponent;  /* subgrid_friction_exponent_on_B_grid 2   */
  int    bc_u[4];                    /* BC_u_{north,east,south,west}     INFINITE */
  int    bc_v[4];                    /* BC_v_{north,east,south,west}     INFINITE */
  /* engine */
  int    precision;                  /* UFE_PREC_*                       FP64   */
  int    krylov;                     /* UFE_KRYLOV_*                     BICGSTAB */
  int    precond;                    /* UFE_PC_*                         BLOCK_JACOBI */
  int    gmres_restart;              /* GMRES restart length, 1..64      30     */
  int    krylov_maxit;               /* per linear solve                 10000  */
  int    warm_start;                 /* 0: zero initial guess (PETSc default), 1: previous iterate */
  double inner_rtol;                 /* MIXED: smallest residual reduction asked of one FP32 inner solve 1e-3 */
  int    stop_rule;                  /* UFE_STOP_*                       L2UV   */
  double nonlinear_rtol;             /* for UFE_STOP_RESIDUAL            1e-6   */
  int    check_every;                /* host polls Krylov convergence at most every k iterations
                                        (polls also land on the previous solve's count) */
  /* Nonlinear acceleration.  accel_depth = 0 keeps the reference's relaxed
   * Picard iteration with its adaptive relaxation rule.  accel_depth = m > 0
   * replaces the relaxation step by Anderson mixing over the last m Picard
   * residuals f = G(u) - u (G = one linearised solve plus velocity limits)
   * with mixing factor accel_mixing; it needs stop_rule = UFE_STOP_RESIDUAL,
   * and the history restarts when the nonlinear residual exceeds
   * accel_restart times the best one seen.  m <= UFE_ACCEL_MAX_DEPTH.
   * accel_continuation = c > 0 (with accel_depth > 0) starts the
   * regularisation at epsilon_sq_0 * 10^c and lowers it tenfold per Picard
   * iteration to epsilon_sq_0 (the mixing history restarts there); the stop
   * test applies from then on. */
  int    accel_depth;                /*                                  0      */
  double accel_mixing;               /*                                  0.5    */
  double accel_restart;              /*                                  10     */
  int    accel_continuation;         /*                                  0      */
  /* 1: the multigrid preconditioner reads an FP32 copy of the finest-level
   * matrix (values rounded once per Picard iteration; vectors and arithmetic
   * stay FP64, and the Krylov operator itself is the FP64 matrix).  The
   * preconditioner need not be exact, so this changes iteration counts at most
   * slightly and halves the bytes of the V-cycle's finest-level passes. */
  int    precond_fp32;               /*                                  0      */
  double ismip_hom_L;                /* refgeo_idealised_ISMIP_HOM_L     0 m (periodic_ISMIP-HOM needs > 0) */
  double slid_zi_p;                  /* slid_ZI_p                        5      */
  double slid_zi_ut;                 /* slid_ZI_ut                       200 m/yr */
  /* smoother of the multigrid V-cycle (UFE_PC_AMG), UFE_SMOOTHER_*.  The l1
   * form is convergent where the plain block Jacobi is not (rows whose
   * off-diagonal blocks outweigh the diagonal one, as on floating ice in the
   * "sans" form); on the marine channel it cuts the Krylov count by 40%. */
  int    amg_smoother;               /*                                  BLOCK_JACOBI */
} ufe_params;

#define UFE_ACCEL_MAX_DEPTH 8

/* Per-solve diagnostics. */
typedef struct ufe_solve_info {
  int    n_visc_its;                 /* Picard iterations                                  */
  int    n_Axb_its;                  /* Krylov iterations summed over Picard iterations     */
  int    n_refinements;              /* FP64 refinement residuals evaluated (MIXED)          */
  double L2_uv;                      /* last reference convergence norm                     */
  double rel_residual;               /* FP64 ||D^-1(b(u)-A(u)u)|| / ||D^-1 b(u)|| at the returned u,
                                        D = 2x2 diagonal blocks (rows in velocity units), over
                                        rows where the velocity limit vel_max is not active      */
  double visc_it_relax_applied;
  double epsilon_sq_0_applied;
  double seconds;                    /* wall time of the Picard loop                        */
  int    n_clamped;                  /* triangles at the velocity limit (left out of rel_residual) */
  int    n_fallbacks;                /* MIXED: linear solves redone in FP64 after FP32 refinement stalled */
  double rel_residual_unscaled;      /* FP64 ||b(u)-A(u)u|| / ||b(u)|| of the assembled (unscaled) system at the
                                        returned u, over the same rows as rel_residual                 */
} ufe_solve_info;

typedef struct ufe_ctx ufe_ctx;

void        ufe_params_default(ufe_params *p);
const char *ufe_last_error(const ufe_ctx *ctx);   /* ctx may be NULL: last global error */
int         ufe_version(void);

/* Context: device copies of the mesh, operators and solver work arrays.
 * Replaces the per-call PETSc matrix/KSP set-up of
 * petsc_basic.f90:32-141 and the mesh arrays read by SSA_main.f90:87-242. */
int  ufe_ctx_create(ufe_ctx **ctx, int device, const ufe_mesh_desc *mesh);
int  ufe_ctx_destroy(ufe_ctx *ctx);
int  ufe_set_params(ufe_ctx *ctx, const ufe_params *p);
/* Coulomb limit of the Tsai2015 and Schoof2005 sliding laws
 * (sliding_laws.f90:244-297, :299-351): the context keeps
 * alpha_sq * effective_pressure per vertex ([nV] each: bed_roughness%alpha_sq,
 * ice%effective_pressure) for every later solve until it is set again;
 * (NULL, NULL) clears it.  Solves with those laws fail with UFE_ERR_ARGUMENT
 * while it is unset. */
int  ufe_set_sliding_cap(ufe_ctx *ctx, const double *alpha_sq, const double *effective_pressure);
/* Multigrid preconditioner knobs (no reference counterpart; the reference's
 * PETSc preconditioner has none of these): damping of the block-Jacobi
 * smoother (0 < omega <= 1, default 0.85), scale of the coarse-level
 * correction (0 < coarse_scale <= 3, default 1.5), and the largest level, in
 * block rows, run in one workgroup (0..4096, default 300).  Takes effect from
 * the next solve. */
int  ufe_set_amg_options(ufe_ctx *ctx, double omega, double coarse_scale, int tail_rows);
/* Launch mode of the Krylov iterations (no reference counterpart): enable != 0
 * replays chunks of BiCGStab iterations from HIP graphs captured once per
 * context (one hipGraphLaunch instead of ~50 kernel launches per iteration;
 * the same kernels in the same order, so results are identical).  Default 0
 * (the environment variable UFE_GRAPHS=1 turns it on for new contexts). */
int  ufe_set_graph_launch(ufe_ctx *ctx, int enable);

/* ---- drop-in entry points ---------------------------------------------- */

/* solve_SSA (SSA_main.f90:87-242).
 *   Hi, Hs, beta_sq, mask_grounded_ice : [nV]   (mask may be NULL = all grounded)
 *   A_flow                             : [nV*nz] ice%A_flow after calc_ice_rheology_Glen
 *   fraction_gr_b                      : [nT]
 *   BC_prescr_mask_b/_u_b/_v_b         : [nT] or all three NULL
 *   u_b, v_b                           : [nT] in: initial guess, out: solution
 *   info                               : may be NULL */
int ufe_solve_ssa(ufe_ctx *ctx, const double *Hi, const double *Hs, const double *A_flow,
                  const double *beta_sq, const double *fraction_gr_b, const int *mask_grounded_ice,
                  const int *BC_prescr_mask_b, const double *BC_prescr_u_b, const double *BC_prescr_v_b,
                  double *u_b, double *v_b, ufe_solve_info *info);

/* solve_DIVA (DIVA_main.f90:88-262).  DIVA state that the reference keeps in
 * its solver type across calls is inout: eta_3D_b [nT*nz], tau_bx_b, tau_by_b,
 * u_base_b, v_base_b [nT].  u_3D_b, v_3D_b [nT*nz] are outputs
 * (calc_3D_velocities, DIVA_main.f90:636-676). */
int ufe_solve_diva(ufe_ctx *ctx, const double *Hi, const double *Hs, const double *A_flow,
                   const double *beta_sq, const double *fraction_gr_b, const int *mask_grounded_ice,
                   const int *BC_prescr_mask_b, const double *BC_prescr_u_b, const double *BC_prescr_v_b,
                   double *u_vav_b, double *v_vav_b, double *u_base_b, double *v_base_b,
                   double *u_3D_b, double *v_3D_b, double *eta_3D_b,
                   double *tau_bx_b, double *tau_by_b, ufe_solve_info *info);

/* ---- device-resident form (inputs uploaded once, solve in HBM) -------- */
int ufe_upload_inputs(ufe_ctx *ctx, const double *Hi, const double *Hs, const double *A_flow,
                      const double *beta_sq, const double *fraction_gr_b, const int *mask_grounded_ice,
                      const int *BC_prescr_mask_b, const double *BC_prescr_u_b, const double *BC_prescr_v_b);
/* u_b, v_b both NULL: zero velocity set on the device (no host transfer) */
int ufe_upload_velocity(ufe_ctx *ctx, const double *u_b, const double *v_b);
int ufe_run(ufe_ctx *ctx, int stress_balance, ufe_solve_info *info);
int ufe_download_velocity(ufe_ctx *ctx, double *u_b, double *v_b);
/* DIVA state up/download (eta_3D_b, tau_b, u_base_b) and 3-D outputs;
 * NULL fields upload as zero */
int ufe_upload_diva_state(ufe_ctx *ctx, const double *eta_3D_b, const double *tau_bx_b,
                          const double *tau_by_b, const double *u_base_b, const double *v_base_b);
int ufe_download_diva(ufe_ctx *ctx, double *u_base_b, double *v_base_b, double *u_3D_b,
                      double *v_3D_b, double *eta_3D_b, double *tau_bx_b, double *tau_by_b);
/* Nonlinear SSA residual r = b(u) - A(u) u at a given velocity, with the
 * uploaded inputs (r, b: [2*nT] rows 2*ti+uv, either may be NULL).  This is
 * the linearised system of solve_SSA_DIVA_linearised
 * (solve_linearised_SSA_DIVA.f90:23-178) assembled at u, applied to u. */
int ufe_ssa_residual(ufe_ctx *ctx, const double *u_b, const double *v_b, double *r, double *b,
                     double *rel_residual);
/* Nonlinear DIVA residual: one evaluation of the state of solve_DIVA's
 * viscosity iteration (DIVA_main.f90:174-186: vertical shear strain rates
 * from tau_b and the previous eta_3D_b, 3-D Glen viscosity, F-integrals,
 * basal friction from u_base and beta_eff; :375-574) at u_vav and the given
 * DIVA state, with the linearised system assembled there and applied to u_vav
 * (solve_linearised_SSA_DIVA.f90:23-178), with the uploaded inputs.
 *   in : u_vav_b, v_vav_b [nT]; eta_3D_b [nT*nz] (NULL = zero), tau_bx_b,
 *        tau_by_b, u_base_b, v_base_b [nT] (NULL = zero)
 *   out (each may be NULL): r, b [2*nT] rows 2*ti+uv; eta_3D_b_new, F1_3D_b
 *        [nT*nz] of the evaluated state; F2_b_surface [nT] (F2_3D_b at k = 1);
 *        beta_b [nT] (map_a_b of the basal friction coefficient) and
 *        beta_eff_b [nT] (after the sub-grid scaling)
 * The context's velocity and DIVA state become u_vav and the evaluated state
 * (a following ufe_run starts from them). */
int ufe_diva_residual(ufe_ctx *ctx, const double *u_vav_b, const double *v_vav_b, const double *eta_3D_b,
                      const double *tau_bx_b, const double *tau_by_b, const double *u_base_b,
                      const double *v_base_b, double *r, double *b, double *rel_residual, double *eta_3D_b_new,
                      double *F1_3D_b, double *F2_b_surface, double *beta_b, double *beta_eff_b);
/* Time of the last ufe_run spent per stage, from HIP events, in ms:
 * [0] vertex/viscosity, [1] triangle/a->b, [2] assembly, [3] SpMV, [4] Krylov vector ops */
int ufe_last_timings(const ufe_ctx *ctx, double *ms5);
/* Measurement hook (no reference counterpart): average duration in ms, from
 * HIP events on the context's stream, of `reps` launches of the FP64 block
 * product y = A x with the system assembled last, and its algorithmic HBM
 * bytes per launch (blocks + column indices + row pointers + x + y). */
int ufe_profile_spmv(ufe_ctx *ctx, int reps, double *avg_ms, double *bytes);
/* Measurement hook (no reference counterpart): average duration in ms of a
 * plain device-to-device copy kernel over `bytes` bytes (read bytes + write
 * bytes each), the known-traffic reference for hardware byte counters. */
int ufe_profile_copy(ufe_ctx *ctx, long bytes, int reps, double *avg_ms);
/* Measurement hook (no reference counterpart): enable != 0 starts timing, with
 * HIP events on the context's stream, two kernels of the BiCGStab iterations
 * of subsequent solves (launches skipped after convergence are not counted),
 * and resets the counters; 0 stops it.  Kernel 0: the FP64 block products
 * with a dot epilogue (two per iteration); kernel 1: the multigrid
 * preconditioner's finest-level prolongation + smoothing (two per iteration). */
int ufe_kernel_timing(ufe_ctx *ctx, int enable);
/* Launches of `kernel` timed so far, their average duration in ms and the
 * algorithmic HBM bytes of one launch (kernel 0: blocks + column indices + row
 * pointers + x, y and the dot operand once each; kernel 1: blocks + column
 * indices + row pointers + aggregate map + residual, pre-smoothed iterate
 * (the same vector when the pre-smoothing is folded), block inverses unless
 * folded, coarse correction and output once each). */
int ufe_kernel_timing_read(const ufe_ctx *ctx, int kernel, long long *launches, double *avg_ms, double *bytes);

/* ---- linear-algebra boundary ------------------------------------------ */

/* y = A x for a CSR matrix (0-based).  Replaces MatMult as exercised by
 * multiply_PETSc_matrix_with_vector_1D (ut_petsc.f90:57-160). */
int ufe_csr_matvec(int m, int n, int nnz, const int *ptr, const int *ind, const double *val,
                   const double *x, double *y);

/* Solve A x = b for a general square CSR matrix with the engine's Krylov
 * solver (x in: initial guess if warm_start, out: solution).  Replaces
 * solve_matrix_equation_CSR_PETSc (petsc_basic.f90:32-64). */
int ufe_csr_solve(int n, int nnz, const int *ptr, const int *ind, const double *val,
                  const double *b, double *x, double rtol, double abstol, int maxit,
                  int precond, int *n_its, double *rel_residual);
/* The same with the Krylov method chosen: krylov = UFE_KRYLOV_BICGSTAB or
 * UFE_KRYLOV_GMRES with restart length `restart` (1..64; PETSc's default
 * KSPGMRES uses 30).  Replaces solve_matrix_equation_CSR_PETSc
 * (petsc_basic.f90:32-64) with the method named. */
int ufe_csr_solve_ex(int n, int nnz, const int *ptr, const int *ind, const double *val,
                     const double *b, double *x, double rtol, double abstol, int maxit,
                     int precond, int krylov, int restart, int *n_its, double *rel_residual);

/* ---- grounded fractions ------------------------------------------------
 * calc_grounded_fractions_bilin_interp_TAF_b
 * (subgrid_grounded_fractions_bilin_TAF.f90:182-370), FP64 on the device.
 * TAF = Hi - max(0, (SL - Hb) * rho_sw/rho_i) is formed per vertex
 * (ice_geometry_basics.f90:43-56). */
int ufe_fraction_gr_b(ufe_ctx *ctx, const double *Hi, const double *Hb, const double *SL,
                      double *fraction_gr_b);

/* ---- vertical heat columns --------------------------------------------
 * The per-vertex part of solve_3D_heat_equation
 * (thermodynamics_3D_heat_equation.f90:114-293) with solve_1D_heat_equation
 * (:309-479), Pounder heat capacity, Ritz conductivity and the pressure
 * melting point recomputed per layer (thermodynamics_utilities.f90:130-229).
 * Per-layer inputs are [nV*nz] vertex-fastest; per-vertex inputs [nV].
 *   mask: 0 = column below Hi_min_thermo (set to T_surf), 1 = grounded
 *   (basal flux BC), 2 = floating (T_base_float BC), 3 = grounded side of
 *   the grounding line with choice_GL_temperature_BC = 'subgrid': each
 *   sub-step solves the column with both basal conditions and keeps
 *   fraction_gr * grounded + (1 - fraction_gr) * floating
 *   (thermodynamics_3D_heat_equation.f90:171-203; 'grounded' and 'pmp'
 *   are masks 1 and 2 for those vertices)
 * Columns still unstable after 10 halvings are reported in is_unstable and
 * counted in *n_unstable; the caller applies its fallback profile. */
typedef struct ufe_heat_inputs {
  int nV, nz;
  const double *zeta;                                   /* [nz] */
  const double *Ti;                                     /* [nV*nz] temperature at t            */
  const double *u_3D, *v_3D, *w_3D;                     /* [nV*nz] velocities                  */
  const double *u_dTdxp_upwind, *v_dTdyp_upwind;        /* [nV*nz] explicit advection terms    */
  const double *dzeta_dx, *dzeta_dy, *dzeta_dz, *dzeta_dt; /* [nV*nz] zeta gradients          */
  const double *Phi;                                    /* [nV*nz] internal heating            */
  const double *Hi_eff;                                 /* [nV]                                */
  const double *T_surf;                                 /* [nV] annual mean surface T          */
  const double *Q_base_grnd;                            /* [nV] frictional + geothermal flux   */
  const int    *mask;                                   /* [nV] 0 thin / 1 grounded / 2 floating / 3 GL blend */
  double dt;
  double Hi_min_thermo;
  const double *fraction_gr;                            /* [nV] vertex grounded fraction, read where
                                                           mask = 3; may be NULL when no mask is 3 */
} ufe_heat_inputs;

int ufe_heat_columns(ufe_ctx *ctx, const ufe_heat_inputs *in, double *Ti_tplusdt, int *is_unstable,
                     int *n_unstable);

/* The same step with every array of *in, Ti_tplusdt and is_unstable in
 * device memory (the same replaced interface, solve_3D_heat_equation,
 * thermodynamics_3D_heat_equation.f90:33-307, for a caller whose column
 * fields already live on the GPU, e.g. the 3-D velocities of a DIVA solve):
 * nothing crosses PCIe but nz + 2 doubles.  The call waits for the device's
 * queued work before it starts and returns when the step is done.  Masks
 * are checked on the device: a thick column that is neither grounded nor
 * floating (or mask 3 without fraction_gr) gives UFE_ERR_ARGUMENT after the
 * launch, with Ti_tplusdt undefined for that column. */
int ufe_heat_columns_device(ufe_ctx *ctx, const ufe_heat_inputs *in, double *Ti_tplusdt, int *is_unstable,
                            int *n_unstable);

/* Allocation-free form of ufe_heat_columns_device (no counterpart in the
 * reference: for callers that cannot allocate device memory, e.g. code under
 * HIP graph capture or a serving loop that forbids hipMalloc).  Same inputs,
 * outputs, checks and results; no context.  It runs on the caller's current
 * device and the null stream (waiting for the device's queued work first) and
 * calls no hipMalloc/hipFree: its scratch memory (the zeta tridiagonals and
 * the reduction partials, 132,624 bytes) is one static device array per GPU,
 * held until the process ends.  Limit: nz <= 32 (UFE_ERR_ARGUMENT beyond);
 * nV is not limited.  Concurrent calls on one GPU share that array and race. */
int ufe_heat_columns_device_noalloc(const ufe_heat_inputs *in, double *Ti_tplusdt, int *is_unstable,
                                    int *n_unstable);

/* Unstable columns after the heat solve (thermodynamics_3D_heat_equation.f90:254-279):
 * when fewer than ceil(nV / 100) columns are flagged, each flagged column of
 * Ti [nV*nz, vertex fastest] is replaced in place with the Robin solution
 * (replace_Ti_with_robin_solution, thermodynamics_utilities.f90:256-337):
 * grounded ice in an accumulation area the steady advection-diffusion
 * profile of Cuffey & Paterson (2010) eq. 9.13-9.22, grounded ice with
 * SMB <= 0 a linear profile from T_surf to the basal pressure-melting point,
 * floating ice a linear profile from T_surf to the basal pressure-melting
 * point, thin ice T_surf; then clamped to the pressure-melting point.
 * Otherwise returns UFE_ERR_DIVERGED (the reference stops the model).
 * Units: K, m, m/yr (SMB), J/m^2/yr (geothermal heat flux). */
typedef struct ufe_heat_robin_inputs {
  int nV, nz;
  const double *zeta;             /* [nz]                                            */
  const double *Hi_eff;           /* [nV]                                            */
  const double *T_surf;           /* [nV] annual mean surface temperature            */
  const double *SMB;              /* [nV] surface mass balance                       */
  const double *geothermal_flux;  /* [nV]                                            */
  const int    *mask;             /* [nV] 0 thin / 1 or 3 grounded / 2 floating      */
  double Hi_min_thermo;
} ufe_heat_robin_inputs;
int ufe_heat_robin(ufe_ctx *ctx, const ufe_heat_robin_inputs *in, const int *is_unstable, double *Ti,
                   int *n_replaced);

/* ---- inputs of the heat columns ----------------------------------------
 * What the reference computes before the column solve
 * (thermodynamics_utilities.f90), on the device.  Fields (n, nz) keep the
 * reference's layout, vertex / triangle index fastest.
 *   ufe_heat_setup: the triangles around each vertex, counter-clockwise
 *     (mesh%iTri / mesh%niTri as CSR, 0-based); checked and kept on the device.
 *   ufe_heat_upwind_derivatives: calc_upwind_heat_flux_derivatives
 *     (:339-434): per vertex with Hi >= 1 m the first of its triangles whose
 *     corner at the vertex contains (-u_vav, -v_vav), then per layer
 *     u_3D_b * ddx_a_b(Ti) and v_3D_b * ddy_a_b(Ti) on that triangle (the
 *     context's a->b operators); 0 where Hi < 1 m; UFE_ERR_MESH with the
 *     reference's "could not find upwind triangle!" otherwise.
 *   ufe_heat_strain_heating: calc_strain_heating (:24-91),
 *     Phi = 2 A^(-1/n) D^(1/n+1) from the nine 3-D strain rates (order du_dx,
 *     du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz; each (nV, nz)),
 *     n = the context's glens_n; 0 where Hi < 0.1 m.
 *   ufe_heat_frictional_heating: calc_frictional_heating (:93-128), the basal
 *     friction coefficient times the basal speed where mask_grounded_ice, 0
 *     elsewhere and everywhere when the context's sliding law is
 *     UFE_SLIDING_NO_SLIDING. */
int ufe_heat_setup(ufe_ctx *ctx, const int *iTri_ptr, const int *iTri_ind);
int ufe_heat_upwind_derivatives(ufe_ctx *ctx, const double *Ti, const double *Hi, const double *u_vav,
                                const double *v_vav, const double *u_3D_b, const double *v_3D_b,
                                double *u_times_dTdxp_upwind, double *v_times_dTdyp_upwind);
int ufe_heat_strain_heating(ufe_ctx *ctx, const double *Hi, const double *A_flow, const double *const *strain_rates,
                            double *Phi);
int ufe_heat_frictional_heating(ufe_ctx *ctx, const double *basal_friction_coefficient, const double *uabs_base,
                                const int *mask_grounded_ice, double *frictional_heating);

/* ---- ice thickness: semi-implicit time step ---------------------------
 * calc_dHi_dt_semiimplicit (conservation_of_mass_semiimplicit.f90:23-159), the
 * reference's second PETSc call site:
 *   M    upwind flux-divergence matrix on the Voronoi cells
 *        (calc_ice_flux_divergence_matrix_upwind,
 *        conservation_of_mass_utilities.f90:17-110), edge velocities from the
 *        triangles either side (map_velocities_from_b_to_c_2D,
 *        map_velocities_to_c_grid.f90:17-69);
 *   the explicit step (calc_dHi_dt_explicit, conservation_of_mass_explicit.f90)
 *        for the flux-limited time step and the border values, with the
 *        'zero' / 'infinite' thickness boundary conditions
 *        (apply_ice_thickness_BC_explicit);
 *   (I + dt fs M) H(t+dt) = H - dt (1 - fs) M H + max(-H, dt (fm (SMB + BMB - target) + LMB))
 *        with border, prescribed-thickness and no-ice rows replaced by
 *        identity rows, solved on the device with the engine's Krylov solver.
 * The vertex geometry (the reference's mesh%C, mesh%A, mesh%Cw and the edge
 * triangles mesh%ETri(mesh%VE)) is attached to a context once. */
typedef struct ufe_vertex_geometry {
  int nC;                   /* connections, sum of mesh%nC                              */
  const int    *C_ptr;      /* [nV+1] neighbours of vertex vi: C_ind[C_ptr[vi] .. C_ptr[vi+1])  */
  const int    *C_ind;      /* [nC]   mesh%C (counter-clockwise)                         */
  const int    *VBI;        /* [nV]   vertex border index 0..8                           */
  const double *A;          /* [nV]   Voronoi cell area                  mesh%A          */
  const double *Cw;         /* [nC]   shared Voronoi boundary length      mesh%Cw         */
  const int    *ETri;       /* [nC*2] triangles left and right of the connection's edge, -1 = none */
} ufe_vertex_geometry;

typedef struct ufe_thickness_params {
  double fs;                /* dHi_semiimplicit_fs                1.5     */
  int    bc_H[4];           /* BC_H_{north,east,south,west}       ZERO    */
  double dt_ice_max;        /* dt_ice_max                         10 yr   */
  double dt_ice_min;        /* dt_ice_min                         0.1 yr  */
  double Hi_min;            /* Hi_min                             0 m     */
  double rtol;              /* dHi_PETSc_rtol                     1e-8    */
  double abstol;            /* dHi_PETSc_abstol                   1e-6    */
  int    krylov;            /* UFE_KRYLOV_*                       GMRES   */
  int    gmres_restart;     /*                                    30      */
  int    maxit;             /*                                    10000   */
} ufe_thickness_params;

/* Per-vertex fields [nV] unless stated; velocities [nT] on the triangles.
 * u_vav_b = v_vav_b = NULL takes the context's current velocity (the last
 * velocity solve's result, still on the device).  mask_noice and
 * BC_prescr_mask are 0/1; BC_prescr_mask and BC_prescr_Hi both or neither.
 * dHi_dt_guess: the initial guess of the linear solve is Hi + dt dHi_dt_guess
 * (the reference passes the previous time step's dH/dt); NULL = Hi. */
typedef struct ufe_thickness_inputs {
  const double *Hi, *Hb, *SL;
  const double *u_vav_b, *v_vav_b;
  const double *SMB, *BMB, *LMB, *fraction_margin, *dHi_dt_target;
  const int    *mask_noice;
  const int    *BC_prescr_mask;
  const double *BC_prescr_Hi;
  const double *dHi_dt_guess;
  double dt;
} ufe_thickness_inputs;

void ufe_thickness_params_default(ufe_thickness_params *p);
/* Attach the vertex geometry of the context's mesh (checked and kept on the
 * device; a second call replaces it). */
int ufe_thickness_setup(ufe_ctx *ctx, const ufe_vertex_geometry *g);
/* One semi-implicit step.  Outputs [nV]: Hi_tplusdt, dHi_dt = (Hi_tplusdt - Hi)/dt,
 * divQ = M Hi (m/yr); *dt_max: the explicit scheme's flux-limited time step
 * (informational: the semi-implicit step uses dt); *n_its Krylov iterations. */
int ufe_thickness_semiimplicit(ufe_ctx *ctx, const ufe_thickness_params *p, const ufe_thickness_inputs *in,
                               double *Hi_tplusdt, double *dHi_dt, double *divQ, double *dt_max, int *n_its);

/* ---- zeta gradients and the BPA 3-D stress balance ---------------------
 * Both run on the context's mesh and layers (3 <= nz <= 32).
 *
 * Set-up, once per context: the a->a derivative operators of the mesh
 * (mesh%M_ddx_a_a, mesh%M_ddy_a_a; nV rows, built by ufe_mesh_operators_aa)
 * and the staggered layers mesh%zeta_stag [nz-1] (NULL: the midpoints of
 * zeta, as mesh_zeta.f90:79-80 sets them for the regular grids).  The 1-D
 * zeta operators are those of calc_vertical_operators_reg_1D / _stag_1D
 * (mesh_zeta.f90:287-482). */
int ufe_zeta_setup(ufe_ctx *ctx, int aa_nnz, const int *aa_ptr, const int *aa_ind, const double *aa_ddx,
                   const double *aa_ddy, const double *zeta_stag);

/* calc_zeta_gradients (zeta_gradients.f90:19-212).  Hi, Hs, dHi_dt, dHs_dt:
 * [nV] (the time derivatives may be NULL = 0).  Outputs, first index fastest,
 * each pointer may be NULL:
 *   ak  [7][nV*nz]:    dzeta_dt, dzeta_dx, dzeta_dy, dzeta_dz, d2zeta_dx2,
 *                      d2zeta_dxdy, d2zeta_dy2 on the vertices
 *   bk  [6][nT*nz]:    dzeta_dx, dzeta_dy, dzeta_dz, d2zeta_dx2, d2zeta_dxdy,
 *                      d2zeta_dy2 on the triangles
 *   bks [6][nT*(nz-1)]: the same on the staggered layers
 * (ice%dzeta_*_ak / _bk / _bks, read by the heat columns and the BPA). */
int ufe_zeta_gradients(ufe_ctx *ctx, const double *Hi, const double *Hs, const double *dHi_dt,
                       const double *dHs_dt, double *const *ak, double *const *bk, double *const *bks);

/* Inputs of solve_BPA (BPA_main.f90:90-110): Hi, Hs, beta_sq (or
 * till_yield_stress), mask_grounded_ice [nV] (mask NULL = all grounded);
 * A_flow [nV*nz] (vertex fastest); fraction_gr_b [nT] (NULL = 1);
 * BC_prescr_mask_bk, _u_bk, _v_bk [nT*nz] (triangle fastest) or all NULL. */
typedef struct ufe_bpa_inputs {
  const double *Hi, *Hs, *A_flow, *beta_sq, *fraction_gr_b;
  const int    *mask_grounded_ice;
  const int    *BC_prescr_mask_bk;
  const double *BC_prescr_u_bk, *BC_prescr_v_bk;
} ufe_bpa_inputs;

/* solve_BPA (BPA_main.f90:90-252): the viscosity iteration of the
 * Blatter-Pattyn approximation on the bk grid.  u_bk, v_bk [nT*nz] (triangle
 * fastest, u_bk(ti,k) of the reference): in the starting velocity (BPA%u_bk),
 * out the solution.  Parameters from the context: glens_n, epsilon_sq_0,
 * visc_eff_min, vel_max, visc_it_*, rtol/abstol (stress_balance_PETSc_*),
 * sliding law, sub-grid friction, the border choices ('infinite', 'zero',
 * 'periodic_ISMIP-HOM' with ismip_hom_L), krylov_maxit, gmres_restart,
 * warm_start.  The linear solves are GMRES(gmres_restart) preconditioned by
 * an exact solve of each triangle's column (the layers' block tridiagonal
 * part of the system); precond and krylov are not read.  Returns
 * UFE_WARN_NOT_CONVERGED when visc_it_nit is reached; info as for SSA/DIVA,
 * rel_residual = ||b - A u|| / ||b|| of the system assembled at the returned
 * velocity. */
int ufe_solve_bpa(ufe_ctx *ctx, const ufe_bpa_inputs *in, double *u_bk, double *v_bk, ufe_solve_info *info);

/* One evaluation of the viscosity iteration's state at (u_bk, v_bk) [nT*nz]
 * (strain rates, viscosity on the ak, bks and bk grids, basal friction;
 * BPA_main.f90:1758-1978) with the stiffness matrix and load vector of
 * solve_BPA_linearised (:330-1723) assembled there and applied to it, at the
 * configured epsilon_sq_0; the lagged copy rows read (u_bk, v_bk) as the
 * previous iterate.  Outputs, each may be NULL: r = b - A u and b
 * [2*nT*nz] in the reference's row order 2*(ti*nz+k)+uv (mesh%tikuv2n);
 * eta_bk [nT*nz] (triangle fastest); beta_b [nT] (the applied basal friction
 * coefficient); *rel_residual = ||r|| / ||b||. */
int ufe_bpa_residual(ufe_ctx *ctx, const ufe_bpa_inputs *in, const double *u_bk, const double *v_bk, double *r,
                     double *b, double *eta_bk, double *beta_b, double *rel_residual);
/* Measurement hooks (no reference counterpart).  Time of the last
 * ufe_solve_bpa per stage, summed over its viscosity iterations, in ms (HIP
 * events on the context's stream): [0] strain rates, viscosity and basal
 * friction, [1] assembly of the system and load vector, [2] factorisation of
 * the column preconditioner, [3] GMRES, [4] velocity limits, relaxation and
 * the L2 change.  And the average duration in ms of `reps` back-to-back FP64
 * block products y = A x with the BPA system assembled last, with their
 * algorithmic HBM bytes per launch (blocks + column indices + row pointers +
 * x + y). */
int ufe_bpa_last_timings(const ufe_ctx *ctx, double *ms5);
int ufe_bpa_profile_spmv(ufe_ctx *ctx, int reps, double *avg_ms, double *bytes);

/* ---- mesh operator construction (host, once per mesh) -----------------
 * calc_matrix_operators_mesh_a_b / _b_a / _b_b_2nd_order
 * (mesh_disc_calc_matrix_operators_2D.f90:198-474,612-764).  Two-call
 * convention: pass NULL arrays to receive *nnz, then arrays of that size. */
int ufe_mesh_operators_ab(int nV, int nT, const double *V, const int *Tri, const double *TriGC,
                          const int *C_ptr, const int *C_ind, int *nnz, int *ptr, int *ind,
                          double *val_map, double *val_ddx, double *val_ddy);
int ufe_mesh_operators_ba(int nV, int nT, const double *V, const int *TriC, const double *TriGC,
                          const int *iTri_ptr, const int *iTri_ind, int *nnz, int *ptr, int *ind,
                          double *val_map, double *val_ddx, double *val_ddy);
int ufe_mesh_operators_bb2(int nT, const int *TriC, const double *TriGC, int *nnz, int *ptr, int *ind,
                           double *val_ddx, double *val_ddy, double *val_d2dx2, double *val_d2dxdy,
                           double *val_d2dy2);
/* calc_matrix_operators_mesh_a_a (mesh_disc_calc_matrix_operators_2D.f90:61-196):
 * a->a ddx/ddy (nV rows, home vertex first), read by the zeta gradients. */
int ufe_mesh_operators_aa(int nV, const double *V, const int *C_ptr, const int *C_ind, int *nnz, int *ptr,
                          int *ind, double *val_ddx, double *val_ddy);

/* Till yield stress of the COULOMB, BUDD and ZOET_IVERSON laws, which the solves
 * take in place of beta_sq (sliding_laws.f90:132-242, :353-403): tau_c =
 * effective_pressure * tan(pi / 180) * till_friction_angle (degrees; the
 * reference's expression as written), then extend_till_yield_stress_to_neighbours
 * (:643-700): an ice-free land vertex takes the smallest unextended tau_c of its
 * grounded neighbours (at most 1000 * rho_i * g), or Hi_min * rho_i * g when
 * none is grounded.  Host code; [nV] arrays, masks 0/1, C_ptr/C_ind the
 * 0-based vertex connectivity (mesh%C). */
int ufe_till_yield_stress(int nV, const int *C_ptr, const int *C_ind, const double *effective_pressure,
                          const double *till_friction_angle, const int *mask_grounded_ice,
                          const int *mask_icefree_land, double Hi_min, double *till_yield_stress);

/* ---- learned initial guess (engine extension, no reference counterpart) --
 * A message-passing network on the triangle graph (3 layers, hidden 64,
 * BF16 matrix-core products with FP32 accumulation) that predicts the
 * velocity from per-triangle features of the uploaded inputs; it seeds the
 * Picard iteration, which stays the arbiter of the result.  Weights: one fp32
 * vector of ufe_mpnn_num_weights() values (NULL at creation = zeros).
 * Graph arguments: N nodes, X [N*16] features, nb [N*3] symmetric neighbour
 * lists (-1 = none), T [N*3] targets (log10 speed in m/yr, flow direction
 * x, y), Y [N*16] outputs (columns 0..2 used). */
typedef struct ufe_mpnn ufe_mpnn;
int ufe_mpnn_num_weights(void);
int ufe_mpnn_create(ufe_mpnn **net, int device, const float *weights);
int ufe_mpnn_destroy(ufe_mpnn *net);
int ufe_mpnn_get_weights(ufe_mpnn *net, float *weights);
int ufe_mpnn_forward(ufe_mpnn *net, int N, const float *X, const int *nb, float *Y);
/* loss = sum ||Y[:, 0:3] - T||^2 / N and its gradient (same layout as the weights) */
int ufe_mpnn_loss_grad(ufe_mpnn *net, int N, const float *X, const int *nb, const float *T, float *loss,
                       float *grad);
/* `steps` full-batch Adam steps (lr; betas 0.9, 0.999); loss_history [steps] may be NULL */
int ufe_mpnn_train(ufe_mpnn *net, int N, const float *X, const int *nb, const float *T, int steps, float lr,
                   float *loss_history);
/* features [nT*16] of the context's uploaded inputs, in the caller's triangle order */
int ufe_mpnn_features(ufe_ctx *ctx, float *X);
/* sets the context's velocity to the network's prediction (DIVA state zero) */
int ufe_mpnn_initial_guess(ufe_ctx *ctx, ufe_mpnn *net);

#ifdef __cplusplus
}
#endif
#endif /* UFE_HIP_H */
