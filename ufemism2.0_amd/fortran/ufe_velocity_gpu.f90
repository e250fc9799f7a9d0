! ufe_velocity_gpu -- Fortran host interface of libufe_hip.so.
!
! ISO_C_BINDING interfaces to the C ABI (include/ufe_hip.h) and thin wrappers
! that keep the argument meaning of the reference's ice-velocity module:
!   solve_SSA_gpu  <-> solve_SSA  (SSA_main.f90:87-242)
!   solve_DIVA_gpu <-> solve_DIVA (DIVA_main.f90:88-262)
!   solve_3D_heat_equation_gpu <-> the column solve, the 1% rule and the Robin
!     replacement of solve_3D_heat_equation (thermodynamics_3D_heat_equation.f90:33-307)
!   calc_dHi_dt_semiimplicit_gpu <-> calc_dHi_dt_semiimplicit
!     (conservation_of_mass_semiimplicit.f90:23-173), after ufe_gpu_thickness_setup
!   solve_BPA_gpu <-> solve_BPA (BPA_main.f90:90-252), after ufe_gpu_zeta_setup
! Arrays are the caller's 1-based Fortran arrays; (n, nz) fields are passed
! as they lie in memory (first index fastest).  Mesh connectivity and operator
! indices are converted to 0-based once, in ufe_gpu_init.  A failing call
! stops the program with the library's message (the reference's crash); a
! Picard loop that hits visc_it_nit prints a warning and returns, as the
! reference does.
module ufe_velocity_gpu

  use iso_c_binding
  implicit none

  private
  public :: ufe_gpu_context, ufe_gpu_params, ufe_gpu_info, ufe_gpu_init, ufe_gpu_finalise, &
            ufe_gpu_set_params, ufe_gpu_default_params, solve_SSA_gpu, solve_DIVA_gpu, &
            ufe_gpu_fraction_gr_b, solve_3D_heat_equation_gpu, ufe_gpu_thickness_params, &
            ufe_gpu_default_thickness_params, ufe_gpu_thickness_setup, calc_dHi_dt_semiimplicit_gpu, &
            ufe_gpu_zeta_setup, solve_BPA_gpu, ufe_gpu_set_sliding_cap, ufe_gpu_till_yield_stress, &
            ufe_gpu_heat_setup, calc_upwind_heat_flux_derivatives_gpu, calc_strain_heating_gpu, &
            calc_frictional_heating_gpu, calc_zeta_gradients_gpu

  integer(c_int), parameter, public :: UFE_OK = 0, UFE_WARN_NOT_CONVERGED = 1

  ! struct ufe_mesh_desc
  type, bind(C) :: ufe_mesh_desc
    integer(c_int) :: nV, nT, nz
    type(c_ptr)    :: V, Tri, TriC, TriBI, TriGC, zeta
    integer(c_int) :: ab_nnz
    type(c_ptr)    :: ab_ptr, ab_ind, ab_map, ab_ddx, ab_ddy
    integer(c_int) :: ba_nnz
    type(c_ptr)    :: ba_ptr, ba_ind, ba_map, ba_ddx, ba_ddy
    integer(c_int) :: bb_nnz
    type(c_ptr)    :: bb_ptr, bb_ind, bb_ddx, bb_ddy, bb_d2dx2, bb_d2dxdy, bb_d2dy2
  end type ufe_mesh_desc

  ! struct ufe_params (field order and types as in the C header)
  type, bind(C) :: ufe_gpu_params
    real(c_double) :: glens_n, epsilon_sq_0, visc_eff_min, vel_max
    integer(c_int) :: visc_it_nit
    real(c_double) :: visc_it_relax, visc_it_norm_dUV_tol, rtol, abstol
    integer(c_int) :: include_crossterms, sliding_law
    real(c_double) :: slid_weertman_m, slid_budd_q_plastic, slid_budd_u_threshold, slid_delta_v, slid_beta_max
    integer(c_int) :: do_gl_subgrid_friction
    real(c_double) :: subgrid_friction_exponent
    integer(c_int) :: bc_u(4), bc_v(4)
    integer(c_int) :: precision, krylov, precond, gmres_restart, krylov_maxit, warm_start
    real(c_double) :: inner_rtol
    integer(c_int) :: stop_rule
    real(c_double) :: nonlinear_rtol
    integer(c_int) :: check_every
    integer(c_int) :: accel_depth
    real(c_double) :: accel_mixing, accel_restart
    integer(c_int) :: accel_continuation
    integer(c_int) :: precond_fp32
    real(c_double) :: ismip_hom_L
    real(c_double) :: slid_zi_p, slid_zi_ut
    integer(c_int) :: amg_smoother
  end type ufe_gpu_params

  ! struct ufe_solve_info
  type, bind(C) :: ufe_gpu_info
    integer(c_int) :: n_visc_its, n_Axb_its, n_refinements
    real(c_double) :: L2_uv, rel_residual, visc_it_relax_applied, epsilon_sq_0_applied, seconds
    integer(c_int) :: n_clamped, n_fallbacks
    real(c_double) :: rel_residual_unscaled
  end type ufe_gpu_info

  ! struct ufe_heat_inputs
  type, bind(C) :: ufe_heat_inputs
    integer(c_int) :: nV, nz
    type(c_ptr)    :: zeta, Ti, u_3D, v_3D, w_3D, u_dTdxp_upwind, v_dTdyp_upwind, dzeta_dx, dzeta_dy, &
                      dzeta_dz, dzeta_dt, Phi, Hi_eff, T_surf, Q_base_grnd, mask
    real(c_double) :: dt, Hi_min_thermo
    type(c_ptr)    :: fraction_gr
  end type ufe_heat_inputs

  ! struct ufe_heat_robin_inputs
  type, bind(C) :: ufe_heat_robin_inputs
    integer(c_int) :: nV, nz
    type(c_ptr)    :: zeta, Hi_eff, T_surf, SMB, geothermal_flux, mask
    real(c_double) :: Hi_min_thermo
  end type ufe_heat_robin_inputs

  ! struct ufe_vertex_geometry
  type, bind(C) :: ufe_vertex_geometry
    integer(c_int) :: nC
    type(c_ptr)    :: C_ptr, C_ind, VBI, A, Cw, ETri
  end type ufe_vertex_geometry

  ! struct ufe_thickness_params
  type, bind(C) :: ufe_gpu_thickness_params
    real(c_double) :: fs
    integer(c_int) :: bc_H(4)
    real(c_double) :: dt_ice_max, dt_ice_min, Hi_min, rtol, abstol
    integer(c_int) :: krylov, gmres_restart, maxit
  end type ufe_gpu_thickness_params

  ! struct ufe_thickness_inputs
  type, bind(C) :: ufe_thickness_inputs
    type(c_ptr)    :: Hi, Hb, SL, u_vav_b, v_vav_b, SMB, BMB, LMB, fraction_margin, dHi_dt_target, &
                      mask_noice, BC_prescr_mask, BC_prescr_Hi, dHi_dt_guess
    real(c_double) :: dt
  end type ufe_thickness_inputs

  ! struct ufe_bpa_inputs
  type, bind(C) :: ufe_bpa_inputs
    type(c_ptr) :: Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice, &
                   BC_prescr_mask_bk, BC_prescr_u_bk, BC_prescr_v_bk
  end type ufe_bpa_inputs

  ! A device context plus the 0-based copies of the mesh arrays it was built from.
  type :: ufe_gpu_context
    type(c_ptr) :: ctx = c_null_ptr
    integer     :: nV = 0, nT = 0, nz = 0
  end type ufe_gpu_context

  interface
    subroutine c_ufe_params_default(p) bind(C, name='ufe_params_default')
      import :: ufe_gpu_params
      type(ufe_gpu_params), intent(out) :: p
    end subroutine
    function c_ufe_last_error(ctx) bind(C, name='ufe_last_error') result(msg)
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: msg
    end function
    function c_ufe_ctx_create(ctx, device, mesh) bind(C, name='ufe_ctx_create') result(rc)
      import :: c_ptr, c_int, ufe_mesh_desc
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
      type(ufe_mesh_desc), intent(in) :: mesh
      integer(c_int) :: rc
    end function
    function c_ufe_ctx_destroy(ctx) bind(C, name='ufe_ctx_destroy') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function c_ufe_set_params(ctx, p) bind(C, name='ufe_set_params') result(rc)
      import :: c_ptr, c_int, ufe_gpu_params
      type(c_ptr), value :: ctx
      type(ufe_gpu_params), intent(in) :: p
      integer(c_int) :: rc
    end function
    function c_ufe_solve_ssa(ctx, Hi, Hs, A_flow, beta_sq, fgr, mask, bcm, bcu, bcv, u_b, v_b, info) &
        bind(C, name='ufe_solve_ssa') result(rc)
      import :: c_ptr, c_int, ufe_gpu_info
      type(c_ptr), value :: ctx, Hi, Hs, A_flow, beta_sq, fgr, mask, bcm, bcu, bcv, u_b, v_b
      type(ufe_gpu_info), intent(out) :: info
      integer(c_int) :: rc
    end function
    function c_ufe_solve_diva(ctx, Hi, Hs, A_flow, beta_sq, fgr, mask, bcm, bcu, bcv, u, v, ubase, vbase, &
        u3, v3, eta3, tbx, tby, info) bind(C, name='ufe_solve_diva') result(rc)
      import :: c_ptr, c_int, ufe_gpu_info
      type(c_ptr), value :: ctx, Hi, Hs, A_flow, beta_sq, fgr, mask, bcm, bcu, bcv, u, v, ubase, vbase, &
                            u3, v3, eta3, tbx, tby
      type(ufe_gpu_info), intent(out) :: info
      integer(c_int) :: rc
    end function
    function c_ufe_heat_columns(ctx, hin, Tnew, unstable, n_unstable) bind(C, name='ufe_heat_columns') result(rc)
      import :: c_ptr, c_int, ufe_heat_inputs
      type(c_ptr), value :: ctx, Tnew, unstable
      type(ufe_heat_inputs), intent(in) :: hin
      integer(c_int), intent(out) :: n_unstable
      integer(c_int) :: rc
    end function
    function c_ufe_heat_robin(ctx, rin, unstable, Ti, n_replaced) bind(C, name='ufe_heat_robin') result(rc)
      import :: c_ptr, c_int, ufe_heat_robin_inputs
      type(c_ptr), value :: ctx, unstable, Ti
      type(ufe_heat_robin_inputs), intent(in) :: rin
      integer(c_int), intent(out) :: n_replaced
      integer(c_int) :: rc
    end function
    subroutine c_ufe_thickness_params_default(p) bind(C, name='ufe_thickness_params_default')
      import :: ufe_gpu_thickness_params
      type(ufe_gpu_thickness_params), intent(out) :: p
    end subroutine
    function c_ufe_thickness_setup(ctx, geom) bind(C, name='ufe_thickness_setup') result(rc)
      import :: c_ptr, c_int, ufe_vertex_geometry
      type(c_ptr), value :: ctx
      type(ufe_vertex_geometry), intent(in) :: geom
      integer(c_int) :: rc
    end function
    function c_ufe_thickness_semiimplicit(ctx, p, inp, Hn, dHdt, divQ, dt_max, n_its) &
        bind(C, name='ufe_thickness_semiimplicit') result(rc)
      import :: c_ptr, c_int, c_double, ufe_gpu_thickness_params, ufe_thickness_inputs
      type(c_ptr), value :: ctx
      type(ufe_gpu_thickness_params), intent(in) :: p
      type(ufe_thickness_inputs), intent(in) :: inp
      type(c_ptr), value :: Hn, dHdt, divQ
      real(c_double), intent(out) :: dt_max
      integer(c_int), intent(out) :: n_its
      integer(c_int) :: rc
    end function
    function c_ufe_zeta_setup(ctx, nnz, ptr, ind, ddx, ddy, zs) bind(C, name='ufe_zeta_setup') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, ptr, ind, ddx, ddy, zs
      integer(c_int), value :: nnz
      integer(c_int) :: rc
    end function
    function c_ufe_heat_setup(ctx, ptr, ind) bind(C, name='ufe_heat_setup') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, ptr, ind
      integer(c_int) :: rc
    end function
    function c_ufe_heat_upwind(ctx, Ti, Hi, u, v, u3, v3, ox, oy) &
        bind(C, name='ufe_heat_upwind_derivatives') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, Ti, Hi, u, v, u3, v3, ox, oy
      integer(c_int) :: rc
    end function
    function c_ufe_heat_strain(ctx, Hi, A, sr, Phi) bind(C, name='ufe_heat_strain_heating') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, Hi, A, Phi
      type(c_ptr), intent(in) :: sr(9)
      integer(c_int) :: rc
    end function
    function c_ufe_heat_friction(ctx, beta, ub, mask, fh) bind(C, name='ufe_heat_frictional_heating') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, beta, ub, mask, fh
      integer(c_int) :: rc
    end function
    function c_ufe_zeta_gradients(ctx, Hi, Hs, dHi, dHs, ak, bk, bks) bind(C, name='ufe_zeta_gradients') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, Hi, Hs, dHi, dHs
      type(c_ptr), intent(in) :: ak(7), bk(6), bks(6)
      integer(c_int) :: rc
    end function
    function c_ufe_set_sliding_cap(ctx, a, n) bind(C, name='ufe_set_sliding_cap') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, a, n
      integer(c_int) :: rc
    end function
    function c_ufe_till_yield_stress(nV, cptr, cind, en, phi, mg, ml, Hi_min, tau) &
        bind(C, name='ufe_till_yield_stress') result(rc)
      import :: c_ptr, c_int, c_double
      integer(c_int), value :: nV
      type(c_ptr), value :: cptr, cind, en, phi, mg, ml, tau
      real(c_double), value :: Hi_min
      integer(c_int) :: rc
    end function
    function c_ufe_solve_bpa(ctx, inp, u, v, info) bind(C, name='ufe_solve_bpa') result(rc)
      import :: c_ptr, c_int, ufe_bpa_inputs, ufe_gpu_info
      type(c_ptr), value :: ctx, u, v
      type(ufe_bpa_inputs), intent(in) :: inp
      type(ufe_gpu_info), intent(out) :: info
      integer(c_int) :: rc
    end function
    function c_ufe_fraction_gr_b(ctx, Hi, Hb, SL, fgr) bind(C, name='ufe_fraction_gr_b') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, Hi, Hb, SL, fgr
      integer(c_int) :: rc
    end function
  end interface

contains

  ! ---- helpers ---------------------------------------------------------------

  subroutine crash_on(rc, ctx, where)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: where
    if (rc < 0) then
      write(0, '(a)') 'ufe_velocity_gpu: ' // where // ' failed: ' // c_string(c_ufe_last_error(ctx))
      error stop 1
    end if
  end subroutine crash_on

  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
      function c_strlen(q) bind(C, name='strlen') result(l)
        import :: c_ptr, c_size_t
        type(c_ptr), value :: q
        integer(c_size_t) :: l
      end function
    end interface
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    s = transfer(chars(1:n), s)
  end function c_string

  function ufe_gpu_default_params() result(p)
    type(ufe_gpu_params) :: p
    call c_ufe_params_default(p)
  end function ufe_gpu_default_params

  ! ---- context ---------------------------------------------------------------

  ! Mesh in the reference's layout (mesh_types.f90): V(nV,2), Tri/TriC(nT,3)
  ! 1-based with TriC = 0 for no neighbour, TriBI(nT), TriGC(nT,2), zeta(nz);
  ! operators as 1-based CSR (ptr(n+1), ind(nnz), val(nnz)) of
  ! M_{map,ddx,ddy}_a_b, M_{map,ddx,ddy}_b_a (one pattern each) and
  ! M2_{ddx,ddy,d2dx2,d2dxdy,d2dy2}_b_b (one pattern).
  subroutine ufe_gpu_init(g, device, V, Tri, TriC, TriBI, TriGC, zeta, &
      ab_ptr, ab_ind, ab_map, ab_ddx, ab_ddy, ba_ptr, ba_ind, ba_map, ba_ddx, ba_ddy, &
      bb_ptr, bb_ind, bb_ddx, bb_ddy, bb_d2dx2, bb_d2dxdy, bb_d2dy2)
    type(ufe_gpu_context), intent(out) :: g
    integer, intent(in) :: device
    real(c_double), intent(in) :: V(:, :), TriGC(:, :), zeta(:)
    integer(c_int), intent(in) :: Tri(:, :), TriC(:, :), TriBI(:)
    integer(c_int), intent(in) :: ab_ptr(:), ab_ind(:), ba_ptr(:), ba_ind(:), bb_ptr(:), bb_ind(:)
    real(c_double), intent(in), target :: ab_map(:), ab_ddx(:), ab_ddy(:), ba_map(:), ba_ddx(:), ba_ddy(:)
    real(c_double), intent(in), target :: bb_ddx(:), bb_ddy(:), bb_d2dx2(:), bb_d2dxdy(:), bb_d2dy2(:)
    type(ufe_mesh_desc) :: d
    real(c_double), allocatable, target :: Vi(:), GCi(:), z(:)
    integer(c_int), allocatable, target :: Tri0(:), TriC0(:), TriBI0(:)
    integer(c_int), allocatable, target :: abp(:), abi(:), bap(:), bai(:), bbp(:), bbi(:)
    integer :: nV, nT, i, n
    integer(c_int) :: rc

    nV = size(V, 1)
    nT = size(Tri, 1)
    g%nV = nV
    g%nT = nT
    g%nz = size(zeta)
    allocate(Vi(2 * nV), GCi(2 * nT), Tri0(3 * nT), TriC0(3 * nT), TriBI0(nT), z(size(zeta)))
    do i = 1, nV
      Vi(2 * i - 1) = V(i, 1)
      Vi(2 * i)     = V(i, 2)
    end do
    do i = 1, nT
      GCi(2 * i - 1) = TriGC(i, 1)
      GCi(2 * i)     = TriGC(i, 2)
      do n = 1, 3
        Tri0(3 * (i - 1) + n)  = Tri(i, n) - 1
        TriC0(3 * (i - 1) + n) = TriC(i, n) - 1
      end do
    end do
    TriBI0 = TriBI
    z = zeta
    abp = ab_ptr - 1
    abi = ab_ind - 1
    bap = ba_ptr - 1
    bai = ba_ind - 1
    bbp = bb_ptr - 1
    bbi = bb_ind - 1

    d%nV = nV
    d%nT = nT
    d%nz = size(zeta)
    d%V = c_loc(Vi)
    d%Tri = c_loc(Tri0)
    d%TriC = c_loc(TriC0)
    d%TriBI = c_loc(TriBI0)
    d%TriGC = c_loc(GCi)
    d%zeta = c_loc(z)
    d%ab_nnz = size(ab_ind)
    d%ab_ptr = c_loc(abp)
    d%ab_ind = c_loc(abi)
    d%ab_map = c_loc(ab_map)
    d%ab_ddx = c_loc(ab_ddx)
    d%ab_ddy = c_loc(ab_ddy)
    d%ba_nnz = size(ba_ind)
    d%ba_ptr = c_loc(bap)
    d%ba_ind = c_loc(bai)
    d%ba_map = c_loc(ba_map)
    d%ba_ddx = c_loc(ba_ddx)
    d%ba_ddy = c_loc(ba_ddy)
    d%bb_nnz = size(bb_ind)
    d%bb_ptr = c_loc(bbp)
    d%bb_ind = c_loc(bbi)
    d%bb_ddx = c_loc(bb_ddx)
    d%bb_ddy = c_loc(bb_ddy)
    d%bb_d2dx2 = c_loc(bb_d2dx2)
    d%bb_d2dxdy = c_loc(bb_d2dxdy)
    d%bb_d2dy2 = c_loc(bb_d2dy2)
    ! the library copies everything to the device during the call
    rc = c_ufe_ctx_create(g%ctx, int(device, c_int), d)
    call crash_on(rc, c_null_ptr, 'ufe_ctx_create')
  end subroutine ufe_gpu_init

  subroutine ufe_gpu_set_params(g, p)
    type(ufe_gpu_context), intent(in) :: g
    type(ufe_gpu_params), intent(in) :: p
    call crash_on(c_ufe_set_params(g%ctx, p), g%ctx, 'ufe_set_params')
  end subroutine ufe_gpu_set_params

  subroutine ufe_gpu_finalise(g)
    type(ufe_gpu_context), intent(inout) :: g
    integer(c_int) :: rc
    if (c_associated(g%ctx)) rc = c_ufe_ctx_destroy(g%ctx)
    g%ctx = c_null_ptr
  end subroutine ufe_gpu_finalise

  ! ---- drop-in entry points -----------------------------------------------------

  ! solve_SSA (SSA_main.f90:87-242): u_b, v_b are the initial guess on entry
  ! and the solution on exit; the prescribed-velocity boundary condition needs
  ! all three of BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b.
  subroutine solve_SSA_gpu(g, Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice, u_b, v_b, &
      n_visc_its, n_Axb_its, BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b, info)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Hi(:), Hs(:), A_flow(:, :), beta_sq(:), fraction_gr_b(:)
    logical, intent(in) :: mask_grounded_ice(:)
    real(c_double), intent(inout), target :: u_b(:), v_b(:)
    integer, intent(out) :: n_visc_its, n_Axb_its
    integer, intent(in), optional :: BC_prescr_mask_b(:)
    real(c_double), intent(in), optional, target :: BC_prescr_u_b(:), BC_prescr_v_b(:)
    type(ufe_gpu_info), intent(out), optional :: info
    integer(c_int), allocatable, target :: mask(:), bcm(:)
    type(c_ptr) :: pbm, pbu, pbv
    type(ufe_gpu_info) :: inf
    integer(c_int) :: rc

    call bc_pointers(BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b, bcm, pbm, pbu, pbv)
    mask = merge(1_c_int, 0_c_int, mask_grounded_ice)
    rc = c_ufe_solve_ssa(g%ctx, c_loc(Hi), c_loc(Hs), c_loc(A_flow), c_loc(beta_sq), c_loc(fraction_gr_b), &
                         c_loc(mask), pbm, pbu, pbv, c_loc(u_b), c_loc(v_b), inf)
    call crash_on(rc, g%ctx, 'solve_SSA_gpu')
    if (rc == UFE_WARN_NOT_CONVERGED) write(0, '(a)') 'WARNING: viscosity iteration failed to converge!'
    n_visc_its = inf%n_visc_its
    n_Axb_its = inf%n_Axb_its
    if (present(info)) info = inf
  end subroutine solve_SSA_gpu

  ! solve_DIVA (DIVA_main.f90:88-262): u_vav_b, v_vav_b and the DIVA state
  ! (u_base_b, v_base_b, eta_3D_b, tau_bx_b, tau_by_b) are inout across calls;
  ! u_3D_b, v_3D_b(nT, nz) are outputs (calc_3D_velocities).
  subroutine solve_DIVA_gpu(g, Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice, u_vav_b, v_vav_b, &
      u_base_b, v_base_b, u_3D_b, v_3D_b, eta_3D_b, tau_bx_b, tau_by_b, n_visc_its, n_Axb_its, &
      BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b, info)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Hi(:), Hs(:), A_flow(:, :), beta_sq(:), fraction_gr_b(:)
    logical, intent(in) :: mask_grounded_ice(:)
    real(c_double), intent(inout), target :: u_vav_b(:), v_vav_b(:), u_base_b(:), v_base_b(:)
    real(c_double), intent(inout), target :: u_3D_b(:, :), v_3D_b(:, :), eta_3D_b(:, :), tau_bx_b(:), tau_by_b(:)
    integer, intent(out) :: n_visc_its, n_Axb_its
    integer, intent(in), optional :: BC_prescr_mask_b(:)
    real(c_double), intent(in), optional, target :: BC_prescr_u_b(:), BC_prescr_v_b(:)
    type(ufe_gpu_info), intent(out), optional :: info
    integer(c_int), allocatable, target :: mask(:), bcm(:)
    type(c_ptr) :: pbm, pbu, pbv
    type(ufe_gpu_info) :: inf
    integer(c_int) :: rc

    call bc_pointers(BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b, bcm, pbm, pbu, pbv)
    mask = merge(1_c_int, 0_c_int, mask_grounded_ice)
    rc = c_ufe_solve_diva(g%ctx, c_loc(Hi), c_loc(Hs), c_loc(A_flow), c_loc(beta_sq), c_loc(fraction_gr_b), &
                          c_loc(mask), pbm, pbu, pbv, c_loc(u_vav_b), c_loc(v_vav_b), c_loc(u_base_b), &
                          c_loc(v_base_b), c_loc(u_3D_b), c_loc(v_3D_b), c_loc(eta_3D_b), c_loc(tau_bx_b), &
                          c_loc(tau_by_b), inf)
    call crash_on(rc, g%ctx, 'solve_DIVA_gpu')
    if (rc == UFE_WARN_NOT_CONVERGED) write(0, '(a)') 'WARNING: viscosity iteration failed to converge!'
    n_visc_its = inf%n_visc_its
    n_Axb_its = inf%n_Axb_its
    if (present(info)) info = inf
  end subroutine solve_DIVA_gpu

  ! The triangles around each vertex, mesh%niTri(nV) and mesh%iTri(nV, :)
  ! (1-based, counter-clockwise), once per context before the upwind terms.
  subroutine ufe_gpu_heat_setup(g, niTri, iTri)
    type(ufe_gpu_context), intent(in) :: g
    integer, intent(in) :: niTri(:), iTri(:, :)
    integer(c_int), allocatable, target :: p0(:), i0(:)
    integer :: nV, vi, j, k
    nV = size(niTri)
    allocate(p0(nV + 1), i0(max(1, sum(niTri))))
    k = 0
    p0(1) = 0
    do vi = 1, nV
      do j = 1, niTri(vi)
        k = k + 1
        i0(k) = iTri(vi, j) - 1
      end do
      p0(vi + 1) = k
    end do
    call crash_on(c_ufe_heat_setup(g%ctx, c_loc(p0), c_loc(i0)), g%ctx, 'ufe_gpu_heat_setup')
  end subroutine ufe_gpu_heat_setup

  ! calc_upwind_heat_flux_derivatives (thermodynamics_utilities.f90:339-434):
  ! Ti(nV, nz), Hi(nV), u_vav, v_vav(nV), u_3D_b, v_3D_b(nT, nz) in; the two
  ! upwind terms (nV, nz) out.  Stops the program where the reference does.
  subroutine calc_upwind_heat_flux_derivatives_gpu(g, Ti, Hi, u_vav, v_vav, u_3D_b, v_3D_b, &
                                                   u_times_dTdxp_upwind, v_times_dTdyp_upwind)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Ti(:, :), Hi(:), u_vav(:), v_vav(:), u_3D_b(:, :), v_3D_b(:, :)
    real(c_double), intent(out), target :: u_times_dTdxp_upwind(:, :), v_times_dTdyp_upwind(:, :)
    call crash_on(c_ufe_heat_upwind(g%ctx, c_loc(Ti), c_loc(Hi), c_loc(u_vav), c_loc(v_vav), c_loc(u_3D_b), &
                                    c_loc(v_3D_b), c_loc(u_times_dTdxp_upwind), c_loc(v_times_dTdyp_upwind)), &
                  g%ctx, 'calc_upwind_heat_flux_derivatives_gpu')
  end subroutine calc_upwind_heat_flux_derivatives_gpu

  ! calc_strain_heating (thermodynamics_utilities.f90:24-91): the nine 3-D
  ! strain rates (nV, nz) of ice%du_dx_3D ... ice%dw_dz_3D in; ice%internal_heating out.
  subroutine calc_strain_heating_gpu(g, Hi, A_flow, du_dx, du_dy, du_dz, dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, &
                                     dw_dz, internal_heating)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Hi(:), A_flow(:, :), du_dx(:, :), du_dy(:, :), du_dz(:, :), &
                                          dv_dx(:, :), dv_dy(:, :), dv_dz(:, :), dw_dx(:, :), dw_dy(:, :), &
                                          dw_dz(:, :)
    real(c_double), intent(out), target :: internal_heating(:, :)
    type(c_ptr) :: sr(9)
    sr = [c_loc(du_dx), c_loc(du_dy), c_loc(du_dz), c_loc(dv_dx), c_loc(dv_dy), c_loc(dv_dz), c_loc(dw_dx), &
          c_loc(dw_dy), c_loc(dw_dz)]
    call crash_on(c_ufe_heat_strain(g%ctx, c_loc(Hi), c_loc(A_flow), sr, c_loc(internal_heating)), g%ctx, &
                  'calc_strain_heating_gpu')
  end subroutine calc_strain_heating_gpu

  ! calc_frictional_heating (thermodynamics_utilities.f90:93-128).
  subroutine calc_frictional_heating_gpu(g, basal_friction_coefficient, uabs_base, mask_grounded_ice, &
                                         frictional_heating)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: basal_friction_coefficient(:), uabs_base(:)
    logical, intent(in) :: mask_grounded_ice(:)
    real(c_double), intent(out), target :: frictional_heating(:)
    integer(c_int), allocatable, target :: m0(:)
    m0 = merge(1, 0, mask_grounded_ice)
    call crash_on(c_ufe_heat_friction(g%ctx, c_loc(basal_friction_coefficient), c_loc(uabs_base), c_loc(m0), &
                                      c_loc(frictional_heating)), g%ctx, 'calc_frictional_heating_gpu')
  end subroutine calc_frictional_heating_gpu

  ! calc_zeta_gradients (zeta_gradients.f90:19-212), after ufe_gpu_zeta_setup:
  ! Hi, Hs, dHi_dt, dHs_dt (nV) in; ice%dzeta_*_ak (nV, nz), _bk (nT, nz)
  ! and _bks (nT, nz-1) out, in the reference's order (dt, dx, dy, dz,
  ! d2x2, d2xdy, d2y2 on ak; dx, dy, dz, d2x2, d2xdy, d2y2 on bk and bks).
  subroutine calc_zeta_gradients_gpu(g, Hi, Hs, dHi_dt, dHs_dt, ak_dt, ak_dx, ak_dy, ak_dz, ak_dx2, ak_dxdy, &
                                     ak_dy2, bk_dx, bk_dy, bk_dz, bk_dx2, bk_dxdy, bk_dy2, bks_dx, bks_dy, &
                                     bks_dz, bks_dx2, bks_dxdy, bks_dy2)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Hi(:), Hs(:), dHi_dt(:), dHs_dt(:)
    real(c_double), intent(out), target :: ak_dt(:, :), ak_dx(:, :), ak_dy(:, :), ak_dz(:, :), ak_dx2(:, :), &
                                           ak_dxdy(:, :), ak_dy2(:, :), bk_dx(:, :), bk_dy(:, :), bk_dz(:, :), &
                                           bk_dx2(:, :), bk_dxdy(:, :), bk_dy2(:, :), bks_dx(:, :), &
                                           bks_dy(:, :), bks_dz(:, :), bks_dx2(:, :), bks_dxdy(:, :), &
                                           bks_dy2(:, :)
    type(c_ptr) :: ak(7), bk(6), bks(6)
    ak = [c_loc(ak_dt), c_loc(ak_dx), c_loc(ak_dy), c_loc(ak_dz), c_loc(ak_dx2), c_loc(ak_dxdy), c_loc(ak_dy2)]
    bk = [c_loc(bk_dx), c_loc(bk_dy), c_loc(bk_dz), c_loc(bk_dx2), c_loc(bk_dxdy), c_loc(bk_dy2)]
    bks = [c_loc(bks_dx), c_loc(bks_dy), c_loc(bks_dz), c_loc(bks_dx2), c_loc(bks_dxdy), c_loc(bks_dy2)]
    call crash_on(c_ufe_zeta_gradients(g%ctx, c_loc(Hi), c_loc(Hs), c_loc(dHi_dt), c_loc(dHs_dt), ak, bk, bks), &
                  g%ctx, 'calc_zeta_gradients_gpu')
  end subroutine calc_zeta_gradients_gpu

  ! Coulomb limit of the Tsai2015 and Schoof2005 sliding laws
  ! (sliding_laws.f90:244-351): bed_roughness%alpha_sq(nV) and
  ! ice%effective_pressure(nV), kept by the context for the following solves.
  subroutine ufe_gpu_set_sliding_cap(g, alpha_sq, effective_pressure)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: alpha_sq(:), effective_pressure(:)
    call crash_on(c_ufe_set_sliding_cap(g%ctx, c_loc(alpha_sq), c_loc(effective_pressure)), g%ctx, &
                  'ufe_gpu_set_sliding_cap')
  end subroutine ufe_gpu_set_sliding_cap

  ! Till yield stress of the Coulomb, Budd and Zoet-Iverson laws, passed to the
  ! solves in place of beta_sq (sliding_laws.f90:132-242, :353-403, with
  ! extend_till_yield_stress_to_neighbours :643-700), from mesh%nC, mesh%C
  ! (1-based neighbours, C(nV, nC_mem)), ice%effective_pressure,
  ! bed_roughness%till_friction_angle, the grounded and ice-free land masks
  ! and C%Hi_min.  Host code of the library; needs no context.
  subroutine ufe_gpu_till_yield_stress(nC, C, effective_pressure, till_friction_angle, mask_grounded_ice, &
                                       mask_icefree_land, Hi_min, till_yield_stress)
    integer, intent(in) :: nC(:), C(:, :)
    real(c_double), intent(in), target :: effective_pressure(:), till_friction_angle(:)
    logical, intent(in) :: mask_grounded_ice(:), mask_icefree_land(:)
    real(c_double), intent(in) :: Hi_min
    real(c_double), intent(out), target :: till_yield_stress(:)
    integer(c_int), allocatable, target :: cp(:), ci(:), mg(:), ml(:)
    integer :: nV, vi, k, j
    nV = size(nC)
    allocate(cp(nV + 1), ci(max(1, sum(nC))), mg(nV), ml(nV))
    k = 0
    cp(1) = 0
    do vi = 1, nV
      do j = 1, nC(vi)
        k = k + 1
        ci(k) = C(vi, j) - 1
      end do
      cp(vi + 1) = k
    end do
    mg = merge(1, 0, mask_grounded_ice)
    ml = merge(1, 0, mask_icefree_land)
    call crash_on(c_ufe_till_yield_stress(int(nV, c_int), c_loc(cp), c_loc(ci), c_loc(effective_pressure), &
                                          c_loc(till_friction_angle), c_loc(mg), c_loc(ml), Hi_min, &
                                          c_loc(till_yield_stress)), c_null_ptr, 'ufe_gpu_till_yield_stress')
  end subroutine ufe_gpu_till_yield_stress

  ! Set-up of the zeta gradients and the BPA, once per context: the a->a
  ! operators M_ddx_a_a, M_ddy_a_a as 1-based CSR (one pattern; nV rows) and
  ! mesh%zeta_stag(nz-1).
  subroutine ufe_gpu_zeta_setup(g, aa_ptr, aa_ind, aa_ddx, aa_ddy, zeta_stag)
    type(ufe_gpu_context), intent(in) :: g
    integer(c_int), intent(in) :: aa_ptr(:), aa_ind(:)
    real(c_double), intent(in), target :: aa_ddx(:), aa_ddy(:), zeta_stag(:)
    integer(c_int), allocatable, target :: p0(:), i0(:)
    p0 = aa_ptr - 1
    i0 = aa_ind - 1
    call crash_on(c_ufe_zeta_setup(g%ctx, int(size(aa_ind), c_int), c_loc(p0), c_loc(i0), c_loc(aa_ddx), &
                                   c_loc(aa_ddy), c_loc(zeta_stag)), g%ctx, 'ufe_gpu_zeta_setup')
  end subroutine ufe_gpu_zeta_setup

  ! solve_BPA (BPA_main.f90:90-252): u_bk, v_bk(nT, nz) are the starting
  ! velocity on entry and the solution on exit; the prescribed-velocity
  ! boundary condition needs all three of BC_prescr_mask_bk, BC_prescr_u_bk,
  ! BC_prescr_v_bk (nT, nz).  ufe_gpu_zeta_setup must have been called.
  subroutine solve_BPA_gpu(g, Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice, u_bk, v_bk, &
      n_visc_its, n_Axb_its, BC_prescr_mask_bk, BC_prescr_u_bk, BC_prescr_v_bk, info)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Hi(:), Hs(:), A_flow(:, :), beta_sq(:), fraction_gr_b(:)
    logical, intent(in) :: mask_grounded_ice(:)
    real(c_double), intent(inout), target :: u_bk(:, :), v_bk(:, :)
    integer, intent(out) :: n_visc_its, n_Axb_its
    integer, intent(in), optional :: BC_prescr_mask_bk(:, :)
    real(c_double), intent(in), optional, target :: BC_prescr_u_bk(:, :), BC_prescr_v_bk(:, :)
    type(ufe_gpu_info), intent(out), optional :: info
    integer(c_int), allocatable, target :: mask(:), bcm(:, :)
    type(ufe_bpa_inputs) :: inp
    type(ufe_gpu_info) :: inf
    integer(c_int) :: rc

    mask = merge(1_c_int, 0_c_int, mask_grounded_ice)
    inp%Hi = c_loc(Hi)
    inp%Hs = c_loc(Hs)
    inp%A_flow = c_loc(A_flow)
    inp%beta_sq = c_loc(beta_sq)
    inp%fraction_gr_b = c_loc(fraction_gr_b)
    inp%mask_grounded_ice = c_loc(mask)
    inp%BC_prescr_mask_bk = c_null_ptr
    inp%BC_prescr_u_bk = c_null_ptr
    inp%BC_prescr_v_bk = c_null_ptr
    if (present(BC_prescr_mask_bk) .or. present(BC_prescr_u_bk) .or. present(BC_prescr_v_bk)) then
      if (.not. (present(BC_prescr_mask_bk) .and. present(BC_prescr_u_bk) .and. present(BC_prescr_v_bk))) then
        write(0, '(a)') 'solve_BPA_gpu: need to provide prescribed u,v fields and mask!'
        error stop 1
      end if
      bcm = int(BC_prescr_mask_bk, c_int)
      inp%BC_prescr_mask_bk = c_loc(bcm)
      inp%BC_prescr_u_bk = c_loc(BC_prescr_u_bk)
      inp%BC_prescr_v_bk = c_loc(BC_prescr_v_bk)
    end if
    rc = c_ufe_solve_bpa(g%ctx, inp, c_loc(u_bk), c_loc(v_bk), inf)
    call crash_on(rc, g%ctx, 'solve_BPA_gpu')
    if (rc == UFE_WARN_NOT_CONVERGED) write(0, '(a)') 'WARNING: viscosity iteration failed to converge!'
    n_visc_its = inf%n_visc_its
    n_Axb_its = inf%n_Axb_its
    if (present(info)) info = inf
  end subroutine solve_BPA_gpu

  ! calc_grounded_fractions_bilin_interp_TAF_b (subgrid_grounded_fractions_bilin_TAF.f90:182-250)
  subroutine ufe_gpu_fraction_gr_b(g, Hi, Hb, SL, fraction_gr_b)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Hi(:), Hb(:), SL(:)
    real(c_double), intent(out), target :: fraction_gr_b(:)
    call crash_on(c_ufe_fraction_gr_b(g%ctx, c_loc(Hi), c_loc(Hb), c_loc(SL), c_loc(fraction_gr_b)), g%ctx, &
                  'ufe_gpu_fraction_gr_b')
  end subroutine ufe_gpu_fraction_gr_b

  ! solve_3D_heat_equation (thermodynamics_3D_heat_equation.f90:33-307) from the
  ! column inputs onward: every column (nV, nz), vertex index fastest, is
  ! solved implicitly with dt halving; when fewer than 1% of the columns stay
  ! unstable they are replaced with the Robin solution
  ! (replace_Ti_with_robin_solution, thermodynamics_utilities.f90:256-337),
  ! otherwise the program stops as the reference's does.  Thin ice
  ! (Hi_eff < Hi_min_thermo) takes the surface temperature.  Grounding-line
  ! vertices (mask_gl_gr) follow choice_GL_temperature_BC ('grounded' by
  ! default, 'pmp', or 'subgrid', which needs the vertex fraction_gr).
  subroutine solve_3D_heat_equation_gpu(g, zeta, Ti, u_3D, v_3D, w_3D, u_dTdxp_upwind, v_dTdyp_upwind, &
      dzeta_dx, dzeta_dy, dzeta_dz, dzeta_dt, Phi, Hi_eff, T_surf, Q_base_grnd, SMB, geothermal_heat_flux, &
      mask_grounded_ice, mask_floating_ice, dt, Hi_min_thermo, Ti_tplusdt, n_unstable, &
      mask_gl_gr, fraction_gr, choice_GL_temperature_BC)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: zeta(:), Ti(:, :), u_3D(:, :), v_3D(:, :), w_3D(:, :)
    real(c_double), intent(in), target :: u_dTdxp_upwind(:, :), v_dTdyp_upwind(:, :)
    real(c_double), intent(in), target :: dzeta_dx(:, :), dzeta_dy(:, :), dzeta_dz(:, :), dzeta_dt(:, :), Phi(:, :)
    real(c_double), intent(in), target :: Hi_eff(:), T_surf(:), Q_base_grnd(:), SMB(:), geothermal_heat_flux(:)
    logical, intent(in) :: mask_grounded_ice(:), mask_floating_ice(:)
    real(c_double), intent(in) :: dt, Hi_min_thermo
    real(c_double), intent(out), target :: Ti_tplusdt(:, :)
    integer, intent(out) :: n_unstable
    logical, intent(in), optional :: mask_gl_gr(:)
    real(c_double), intent(in), optional, target :: fraction_gr(:)
    character(len=*), intent(in), optional :: choice_GL_temperature_BC
    type(ufe_heat_inputs) :: hin
    type(ufe_heat_robin_inputs) :: rin
    integer(c_int), allocatable, target :: mask(:), mask_col(:), unstable(:)
    type(c_ptr) :: pfgr
    integer(c_int) :: nun, nrep
    integer :: nV, nz
    character(len=16) :: gl_bc
    nV = size(Ti, 1)
    nz = size(Ti, 2)
    allocate(mask(nV), mask_col(nV), unstable(nV))
    ! grounded / floating as replace_Ti_with_robin_solution reads them
    mask = 0
    where (mask_grounded_ice) mask = 1
    where (mask_floating_ice .and. .not. mask_grounded_ice) mask = 2
    ! the column solve's basal condition (thermodynamics_3D_heat_equation.f90:171-224)
    mask_col = mask
    pfgr = c_null_ptr
    gl_bc = 'grounded'
    if (present(choice_GL_temperature_BC)) gl_bc = choice_GL_temperature_BC
    if (present(mask_gl_gr)) then
      select case (trim(gl_bc))
      case ('grounded')
        where (mask_gl_gr) mask_col = 1
      case ('pmp')
        where (mask_gl_gr) mask_col = 2
      case ('subgrid')
        if (.not. present(fraction_gr)) then
          write(0, '(a)') 'ufe_velocity_gpu: choice_GL_temperature_BC = "subgrid" needs fraction_gr'
          error stop 1
        end if
        where (mask_gl_gr) mask_col = 3
        pfgr = c_loc(fraction_gr)
      case default
        write(0, '(3a)') 'ufe_velocity_gpu: unknown choice_GL_temperature_BC "', trim(gl_bc), '"!'
        error stop 1
      end select
    end if
    hin = ufe_heat_inputs(int(nV, c_int), int(nz, c_int), c_loc(zeta), c_loc(Ti), c_loc(u_3D), c_loc(v_3D), &
                          c_loc(w_3D), c_loc(u_dTdxp_upwind), c_loc(v_dTdyp_upwind), c_loc(dzeta_dx), &
                          c_loc(dzeta_dy), c_loc(dzeta_dz), c_loc(dzeta_dt), c_loc(Phi), c_loc(Hi_eff), &
                          c_loc(T_surf), c_loc(Q_base_grnd), c_loc(mask_col), dt, Hi_min_thermo, pfgr)
    call crash_on(c_ufe_heat_columns(g%ctx, hin, c_loc(Ti_tplusdt), c_loc(unstable), nun), c_null_ptr, &
                  'solve_3D_heat_equation_gpu')
    rin = ufe_heat_robin_inputs(int(nV, c_int), int(nz, c_int), c_loc(zeta), c_loc(Hi_eff), c_loc(T_surf), &
                                c_loc(SMB), c_loc(geothermal_heat_flux), c_loc(mask), Hi_min_thermo)
    call crash_on(c_ufe_heat_robin(g%ctx, rin, c_loc(unstable), c_loc(Ti_tplusdt), nrep), c_null_ptr, &
                  'solve_3D_heat_equation_gpu (unstable columns)')
    n_unstable = int(nun)
  end subroutine solve_3D_heat_equation_gpu

  function ufe_gpu_default_thickness_params() result(p)
    type(ufe_gpu_thickness_params) :: p
    call c_ufe_thickness_params_default(p)
  end function ufe_gpu_default_thickness_params

  ! Attach the Voronoi geometry the thickness step needs, from the reference's
  ! mesh arrays: nC(nV), C(nV, nC_mem) (1-based neighbours), VBI(nV), A(nV),
  ! Cw(nV, nC_mem), VE(nV, nC_mem) (1-based edge of each connection) and
  ! ETri(nE, 2) (triangles either side of each edge, 0 = none).
  subroutine ufe_gpu_thickness_setup(g, nC, C, VBI, A, Cw, VE, ETri)
    type(ufe_gpu_context), intent(in) :: g
    integer, intent(in) :: nC(:), C(:, :), VBI(:), VE(:, :), ETri(:, :)
    real(c_double), intent(in) :: A(:), Cw(:, :)
    integer(c_int), allocatable, target :: c_ptr_(:), c_ind(:), vbi0(:), etri_c(:)
    real(c_double), allocatable, target :: A0(:), cw_c(:)
    type(ufe_vertex_geometry) :: geom
    integer :: nV, vi, ci, k, ei

    nV = size(nC)
    allocate(c_ptr_(nV + 1), c_ind(sum(nC)), cw_c(sum(nC)), etri_c(2 * sum(nC)), vbi0(nV), A0(nV))
    c_ptr_(1) = 0
    k = 0
    do vi = 1, nV
      do ci = 1, nC(vi)
        k = k + 1
        c_ind(k) = C(vi, ci) - 1
        cw_c(k) = Cw(vi, ci)
        ei = VE(vi, ci)
        etri_c(2 * k - 1) = ETri(ei, 1) - 1
        etri_c(2 * k) = ETri(ei, 2) - 1
      end do
      c_ptr_(vi + 1) = k
    end do
    vbi0 = VBI
    A0 = A
    geom%nC = k
    geom%C_ptr = c_loc(c_ptr_)
    geom%C_ind = c_loc(c_ind)
    geom%VBI = c_loc(vbi0)
    geom%A = c_loc(A0)
    geom%Cw = c_loc(cw_c)
    geom%ETri = c_loc(etri_c)
    call crash_on(c_ufe_thickness_setup(g%ctx, geom), g%ctx, 'ufe_gpu_thickness_setup')
  end subroutine ufe_gpu_thickness_setup

  ! calc_dHi_dt_semiimplicit (conservation_of_mass_semiimplicit.f90:23-173):
  ! dHi_dt is the previous rate on entry (the initial guess of the linear
  ! solve is Hi + dt dHi_dt, as in the reference) and the new rate on exit;
  ! AMB comes back as the reference's arithmetic leaves it (dHi_dt minus the
  ! same rate, zero).  The prescribed-thickness condition needs both
  ! BC_prescr_mask and BC_prescr_Hi.  params defaults to
  ! ufe_gpu_default_thickness_params() (dHi_semiimplicit_fs 1.5, 'zero'
  ! borders, dHi_PETSc_rtol 1e-8, dHi_PETSc_abstol 1e-6).
  subroutine calc_dHi_dt_semiimplicit_gpu(g, Hi, Hb, SL, u_vav_b, v_vav_b, SMB, BMB, LMB, AMB, fraction_margin, &
      mask_noice, dt, dHi_dt, Hi_tplusdt, divQ, dHi_dt_target, BC_prescr_mask, BC_prescr_Hi, params, dt_max, &
      n_Axb_its)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in), target :: Hi(:), Hb(:), SL(:), u_vav_b(:), v_vav_b(:), SMB(:), BMB(:), LMB(:)
    real(c_double), intent(inout) :: AMB(:)
    real(c_double), intent(in), target :: fraction_margin(:)
    logical, intent(in) :: mask_noice(:)
    real(c_double), intent(inout) :: dt
    real(c_double), intent(inout), target :: dHi_dt(:)
    real(c_double), intent(out), target :: Hi_tplusdt(:), divQ(:)
    real(c_double), intent(in), target :: dHi_dt_target(:)
    integer, intent(in), optional :: BC_prescr_mask(:)
    real(c_double), intent(in), optional, target :: BC_prescr_Hi(:)
    type(ufe_gpu_thickness_params), intent(in), optional :: params
    real(c_double), intent(out), optional :: dt_max
    integer, intent(out), optional :: n_Axb_its
    type(ufe_gpu_thickness_params) :: p
    type(ufe_thickness_inputs) :: inp
    integer(c_int), allocatable, target :: noice(:), bcm(:)
    real(c_double), allocatable, target :: guess(:), rate(:)
    real(c_double) :: dtm
    integer(c_int) :: its, rc

    if (present(BC_prescr_mask) .neqv. present(BC_prescr_Hi)) then
      write(0, '(a)') 'calc_dHi_dt_semiimplicit_gpu: need to provide prescribed both Hi and mask!'
      error stop 1
    end if
    if (present(params)) then
      p = params
    else
      p = ufe_gpu_default_thickness_params()
    end if
    noice = merge(1_c_int, 0_c_int, mask_noice)
    guess = dHi_dt
    allocate(rate(size(Hi)))
    inp%Hi = c_loc(Hi); inp%Hb = c_loc(Hb); inp%SL = c_loc(SL)
    inp%u_vav_b = c_loc(u_vav_b); inp%v_vav_b = c_loc(v_vav_b)
    inp%SMB = c_loc(SMB); inp%BMB = c_loc(BMB); inp%LMB = c_loc(LMB)
    inp%fraction_margin = c_loc(fraction_margin); inp%dHi_dt_target = c_loc(dHi_dt_target)
    inp%mask_noice = c_loc(noice)
    inp%BC_prescr_mask = c_null_ptr
    inp%BC_prescr_Hi = c_null_ptr
    if (present(BC_prescr_mask)) then
      bcm = int(BC_prescr_mask, c_int)
      inp%BC_prescr_mask = c_loc(bcm)
      inp%BC_prescr_Hi = c_loc(BC_prescr_Hi)
    end if
    inp%dHi_dt_guess = c_loc(guess)
    inp%dt = dt
    rc = c_ufe_thickness_semiimplicit(g%ctx, p, inp, c_loc(Hi_tplusdt), c_loc(rate), c_loc(divQ), dtm, its)
    call crash_on(rc, g%ctx, 'calc_dHi_dt_semiimplicit_gpu')
    if (rc == UFE_WARN_NOT_CONVERGED) write(0, '(a)') 'WARNING: thickness solve did not reach its tolerance'
    dHi_dt = rate
    AMB = dHi_dt - rate
    if (present(dt_max)) dt_max = dtm
    if (present(n_Axb_its)) n_Axb_its = its
  end subroutine calc_dHi_dt_semiimplicit_gpu

  subroutine bc_pointers(m, u, v, bcm, pbm, pbu, pbv)
    integer, intent(in), optional :: m(:)
    real(c_double), intent(in), optional, target :: u(:), v(:)
    integer(c_int), allocatable, target, intent(inout) :: bcm(:)
    type(c_ptr), intent(out) :: pbm, pbu, pbv
    pbm = c_null_ptr
    pbu = c_null_ptr
    pbv = c_null_ptr
    if (present(m) .or. present(u) .or. present(v)) then
      if (.not. (present(m) .and. present(u) .and. present(v))) then
        write(0, '(a)') 'ufe_velocity_gpu: need to provide prescribed u,v fields and mask!'
        error stop 1
      end if
      bcm = int(m, c_int)
      pbm = c_loc(bcm)
      pbu = c_loc(u)
      pbv = c_loc(v)
    end if
  end subroutine bc_pointers

end module ufe_velocity_gpu
