! ufe_fortran_driver -- runs one velocity solve through ufe_velocity_gpu from a
! case file (raw little-endian stream written by the Python host, see
! ufemism2.0_amd/fortran_case.py) and writes u_b, v_b and the iteration
! counts.  With a thickness case it then takes one semi-implicit thickness
! step with the solved velocities (the reference's order: stress balance, then
! calc_dHi_dt_semiimplicit) and writes Hi_tplusdt, dHi_dt, divQ, dt_max and
! the Krylov count.  Used by the tests to show the Fortran interface end to end.
!
!   ufe_fortran_driver CASE_FILE OUT_FILE [THICKNESS_CASE THICKNESS_OUT]
program ufe_fortran_driver

  use iso_c_binding
  use ufe_velocity_gpu
  implicit none

  character(len=4096) :: case_file, out_file
  integer :: u_in, u_out
  integer(c_int) :: hdr(7), ip(6)
  real(c_double) :: rp(5)
  integer :: nV, nT, nz, ab_nnz, ba_nnz, bb_nnz, stress
  real(c_double), allocatable :: V(:, :), TriGC(:, :), zeta(:)
  integer(c_int), allocatable :: Tri(:, :), TriC(:, :), TriBI(:)
  integer(c_int), allocatable :: ab_ptr(:), ab_ind(:), ba_ptr(:), ba_ind(:), bb_ptr(:), bb_ind(:)
  real(c_double), allocatable :: ab_map(:), ab_ddx(:), ab_ddy(:), ba_map(:), ba_ddx(:), ba_ddy(:)
  real(c_double), allocatable :: bb_ddx(:), bb_ddy(:), bb_d2dx2(:), bb_d2dxdy(:), bb_d2dy2(:)
  real(c_double), allocatable :: Hi(:), Hs(:), A_flow(:, :), beta_sq(:), fgr(:)
  integer(c_int), allocatable :: mask(:)
  real(c_double), allocatable :: u_b(:), v_b(:), ub(:), vb(:), u3(:, :), v3(:, :), eta3(:, :), tbx(:), tby(:)
  integer(c_int) :: bip(10)
  real(c_double) :: brp(5)
  ! optional sliding block (SSA/DIVA): law, nC_mem, kind (1 Coulomb cap, 2 till
  ! yield stress from N and the friction angle); Weertman m, ZI p, ZI u_t
  integer(c_int) :: sip(3)
  real(c_double) :: srp(3)
  real(c_double), allocatable :: s_alpha(:), s_en(:), s_phi(:)
  integer, allocatable :: s_land(:), s_nC(:), s_C(:, :)
  logical :: has_sliding
  integer(c_int), allocatable :: aa_ptr(:), aa_ind(:)
  real(c_double), allocatable :: aa_ddx(:), aa_ddy(:), zeta_stag(:), u_bk(:, :), v_bk(:, :)
  type(ufe_gpu_context) :: g
  type(ufe_gpu_params) :: p
  type(ufe_gpu_info) :: info
  integer :: n_visc_its, n_Axb_its

  call get_command_argument(1, case_file)
  call get_command_argument(2, out_file)
  open(newunit=u_in, file=trim(case_file), access='stream', form='unformatted', status='old', action='read')
  read(u_in) hdr
  nV = hdr(1); nT = hdr(2); nz = hdr(3); ab_nnz = hdr(4); ba_nnz = hdr(5); bb_nnz = hdr(6)
  stress = iand(hdr(7), 15)
  has_sliding = iand(hdr(7), 16) /= 0
  allocate(V(nV, 2), Tri(nT, 3), TriC(nT, 3), TriBI(nT), TriGC(nT, 2), zeta(nz))
  read(u_in) V, Tri, TriC, TriBI, TriGC, zeta
  allocate(ab_ptr(nT + 1), ab_ind(ab_nnz), ab_map(ab_nnz), ab_ddx(ab_nnz), ab_ddy(ab_nnz))
  read(u_in) ab_ptr, ab_ind, ab_map, ab_ddx, ab_ddy
  allocate(ba_ptr(nV + 1), ba_ind(ba_nnz), ba_map(ba_nnz), ba_ddx(ba_nnz), ba_ddy(ba_nnz))
  read(u_in) ba_ptr, ba_ind, ba_map, ba_ddx, ba_ddy
  allocate(bb_ptr(nT + 1), bb_ind(bb_nnz), bb_ddx(bb_nnz), bb_ddy(bb_nnz), bb_d2dx2(bb_nnz), &
           bb_d2dxdy(bb_nnz), bb_d2dy2(bb_nnz))
  read(u_in) bb_ptr, bb_ind, bb_ddx, bb_ddy, bb_d2dx2, bb_d2dxdy, bb_d2dy2
  read(u_in) ip, rp
  allocate(Hi(nV), Hs(nV), A_flow(nV, nz), beta_sq(nV), fgr(nT), mask(nV))
  read(u_in) Hi, Hs, A_flow, beta_sq, fgr, mask
  if (has_sliding) then
    read(u_in) sip, srp
    if (sip(3) == 1) then
      allocate(s_alpha(nV), s_en(nV))
      read(u_in) s_alpha, s_en
    else if (sip(3) == 2) then
      allocate(s_en(nV), s_phi(nV), s_land(nV), s_nC(nV), s_C(nV, sip(2)))
      read(u_in) s_en, s_phi, s_land, s_nC, s_C
    end if
  end if
  if (stress == 2) then
    ! BPA: sliding law, border choices, rheology and iteration controls, the
    ! a->a operators (1-based CSR) and the staggered layers
    read(u_in) bip, brp
    allocate(aa_ptr(nV + 1), aa_ind(bip(1)), aa_ddx(bip(1)), aa_ddy(bip(1)), zeta_stag(nz - 1))
    read(u_in) aa_ptr, aa_ind, aa_ddx, aa_ddy, zeta_stag
  end if
  close(u_in)

  call ufe_gpu_init(g, 0, V, Tri, TriC, TriBI, TriGC, zeta, ab_ptr, ab_ind, ab_map, ab_ddx, ab_ddy, &
                    ba_ptr, ba_ind, ba_map, ba_ddx, ba_ddy, bb_ptr, bb_ind, bb_ddx, bb_ddy, bb_d2dx2, &
                    bb_d2dxdy, bb_d2dy2)
  p = ufe_gpu_default_params()
  p%stop_rule = ip(1); p%visc_it_nit = ip(2); p%include_crossterms = ip(3)
  p%precond = ip(4); p%warm_start = ip(5); p%accel_depth = ip(6)
  p%nonlinear_rtol = rp(1); p%vel_max = rp(2); p%abstol = rp(3); p%rtol = rp(4); p%accel_mixing = rp(5)
  if (stress == 2) then
    p%sliding_law = bip(2)
    p%bc_u = bip(3:6)
    p%bc_v = bip(7:10)
    p%glens_n = brp(1); p%epsilon_sq_0 = brp(2); p%visc_it_relax = brp(3)
    p%visc_it_norm_dUV_tol = brp(4); p%ismip_hom_L = brp(5)
  end if
  if (has_sliding) then
    p%sliding_law = sip(1)
    p%slid_weertman_m = srp(1); p%slid_zi_p = srp(2); p%slid_zi_ut = srp(3)
  end if
  call ufe_gpu_set_params(g, p)
  if (has_sliding) then
    ! sliding_laws.f90: the Coulomb cap of Tsai2015 / Schoof2005, or the till
    ! yield stress of the yield-stress laws in place of beta_sq
    if (sip(3) == 1) call ufe_gpu_set_sliding_cap(g, s_alpha, s_en)
    if (sip(3) == 2) call ufe_gpu_till_yield_stress(s_nC, s_C, s_en, s_phi, mask /= 0, s_land /= 0, &
                                                    0.0_c_double, beta_sq)
  end if

  if (stress == 2) then
    call ufe_gpu_zeta_setup(g, aa_ptr, aa_ind, aa_ddx, aa_ddy, zeta_stag)
    allocate(u_bk(nT, nz), v_bk(nT, nz))
    u_bk = 0.0_c_double
    v_bk = 0.0_c_double
    call solve_BPA_gpu(g, Hi, Hs, A_flow, beta_sq, fgr, mask /= 0, u_bk, v_bk, n_visc_its, n_Axb_its, info=info)
    open(newunit=u_out, file=trim(out_file), access='stream', form='unformatted', status='replace', &
         action='write')
    write(u_out) u_bk, v_bk, int(n_visc_its, c_int), int(n_Axb_its, c_int), info%rel_residual
    close(u_out)
    write(*, '(a,i0,a,i0,a,es10.3)') 'BPA solve done: viscosity iterations ', n_visc_its, ' krylov ', &
      n_Axb_its, ' residual ', info%rel_residual
    call ufe_gpu_finalise(g)
    stop
  end if

  allocate(u_b(nT), v_b(nT))
  u_b = 0.0_c_double
  v_b = 0.0_c_double
  if (stress == 0) then
    call solve_SSA_gpu(g, Hi, Hs, A_flow, beta_sq, fgr, mask /= 0, u_b, v_b, n_visc_its, n_Axb_its, info=info)
  else
    allocate(ub(nT), vb(nT), u3(nT, nz), v3(nT, nz), eta3(nT, nz), tbx(nT), tby(nT))
    ub = 0.0_c_double; vb = 0.0_c_double; eta3 = 0.0_c_double; tbx = 0.0_c_double; tby = 0.0_c_double
    call solve_DIVA_gpu(g, Hi, Hs, A_flow, beta_sq, fgr, mask /= 0, u_b, v_b, ub, vb, u3, v3, eta3, tbx, tby, &
                        n_visc_its, n_Axb_its, info=info)
  end if

  open(newunit=u_out, file=trim(out_file), access='stream', form='unformatted', status='replace', action='write')
  write(u_out) u_b, v_b, int(n_visc_its, c_int), int(n_Axb_its, c_int), info%rel_residual
  close(u_out)
  write(*, '(a,i0,a,i0,a,es10.3)') 'solve done: picard ', n_visc_its, ' krylov ', n_Axb_its, ' residual ', &
    info%rel_residual

  if (command_argument_count() >= 4) call thickness_step(g, Hi, u_b, v_b)
  if (command_argument_count() == 3) call heat_inputs(g, Hi, Hs, A_flow)
  call ufe_gpu_finalise(g)

contains

  ! The heat columns' inputs of thermodynamics_utilities.f90 and the zeta
  ! gradients, from a third case file; all outputs to <out>.heatin.
  subroutine heat_inputs(g, Hi, Hs, A_flow)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in) :: Hi(:), Hs(:), A_flow(:, :)
    character(len=4096) :: hcase
    integer(c_int) :: hh(2)
    integer :: uin, uout, k
    integer, allocatable :: niTri(:), iTri(:, :), mg(:)
    integer(c_int), allocatable :: a_ptr(:), a_ind(:)
    real(c_double), allocatable :: Ti(:, :), uv(:), vv(:), u3b(:, :), v3b(:, :), sr(:, :, :), beta(:), uab(:)
    real(c_double), allocatable :: a_dx(:), a_dy(:), zs(:), dHi(:), dHs(:), ox(:, :), oy(:, :), Phi(:, :), fh(:)
    real(c_double), allocatable :: ak(:, :, :), bk(:, :, :), bks(:, :, :)
    call get_command_argument(3, hcase)
    open(newunit=uin, file=trim(hcase), access='stream', form='unformatted', status='old', action='read')
    read(uin) hh
    allocate(niTri(nV), iTri(nV, hh(1)), Ti(nV, nz), uv(nV), vv(nV), u3b(nT, nz), v3b(nT, nz), sr(nV, nz, 9))
    allocate(beta(nV), uab(nV), mg(nV), a_ptr(nV + 1), a_ind(hh(2)), a_dx(hh(2)), a_dy(hh(2)), zs(nz - 1))
    allocate(dHi(nV), dHs(nV))
    read(uin) niTri, iTri, Ti, uv, vv, u3b, v3b, sr, beta, uab, mg, a_ptr, a_ind, a_dx, a_dy, zs, dHi, dHs
    close(uin)
    allocate(ox(nV, nz), oy(nV, nz), Phi(nV, nz), fh(nV), ak(nV, nz, 7), bk(nT, nz, 6), bks(nT, nz - 1, 6))
    call ufe_gpu_heat_setup(g, niTri, iTri)
    call calc_upwind_heat_flux_derivatives_gpu(g, Ti, Hi, uv, vv, u3b, v3b, ox, oy)
    call calc_strain_heating_gpu(g, Hi, A_flow, sr(:, :, 1), sr(:, :, 2), sr(:, :, 3), sr(:, :, 4), sr(:, :, 5), &
                                 sr(:, :, 6), sr(:, :, 7), sr(:, :, 8), sr(:, :, 9), Phi)
    call calc_frictional_heating_gpu(g, beta, uab, mg /= 0, fh)
    call ufe_gpu_zeta_setup(g, a_ptr, a_ind, a_dx, a_dy, zs)
    call calc_zeta_gradients_gpu(g, Hi, Hs, dHi, dHs, ak(:, :, 1), ak(:, :, 2), ak(:, :, 3), ak(:, :, 4), &
                                 ak(:, :, 5), ak(:, :, 6), ak(:, :, 7), bk(:, :, 1), bk(:, :, 2), bk(:, :, 3), &
                                 bk(:, :, 4), bk(:, :, 5), bk(:, :, 6), bks(:, :, 1), bks(:, :, 2), &
                                 bks(:, :, 3), bks(:, :, 4), bks(:, :, 5), bks(:, :, 6))
    open(newunit=uout, file=trim(out_file) // '.heatin', access='stream', form='unformatted', &
         status='replace', action='write')
    write(uout) ox, oy, Phi, fh, ak, bk, bks
    close(uout)
    k = 0
    write(*, '(a)') 'heat inputs done'
  end subroutine heat_inputs

  subroutine thickness_step(g, Hi, u_b, v_b)
    type(ufe_gpu_context), intent(in) :: g
    real(c_double), intent(in) :: Hi(:), u_b(:), v_b(:)
    character(len=4096) :: tcase, tout
    integer(c_int) :: th(6)
    integer :: nC_mem, nE, n_its, uin, uout
    integer(c_int), allocatable :: nC(:), C(:, :), VE(:, :), ETri(:, :), VBI(:), noice(:)
    real(c_double), allocatable :: A(:), Cw(:, :), Hb(:), SL(:), SMB(:), BMB(:), LMB(:), fm(:), tgt(:)
    real(c_double), allocatable :: AMB(:), dHi_dt(:), Hn(:), divQ(:)
    real(c_double) :: dt, dt_max
    type(ufe_gpu_thickness_params) :: tp

    call get_command_argument(3, tcase)
    call get_command_argument(4, tout)
    open(newunit=uin, file=trim(tcase), access='stream', form='unformatted', status='old', action='read')
    read(uin) th, dt
    nC_mem = th(1); nE = th(2)
    allocate(nC(nV), C(nV, nC_mem), VE(nV, nC_mem), ETri(nE, 2), VBI(nV), noice(nV), A(nV), Cw(nV, nC_mem), &
             Hb(nV), SL(nV), SMB(nV), BMB(nV), LMB(nV), fm(nV), tgt(nV), AMB(nV), dHi_dt(nV), Hn(nV), divQ(nV))
    read(uin) nC, C, VE, ETri, VBI, A, Cw, Hb, SL, SMB, BMB, LMB, fm, noice
    close(uin)
    call ufe_gpu_thickness_setup(g, nC, C, VBI, A, Cw, VE, ETri)
    tp = ufe_gpu_default_thickness_params()
    tp%bc_H = th(3:6)
    tgt = 0.0_c_double
    AMB = 0.0_c_double
    dHi_dt = 0.0_c_double
    call calc_dHi_dt_semiimplicit_gpu(g, Hi, Hb, SL, u_b, v_b, SMB, BMB, LMB, AMB, fm, noice /= 0, dt, dHi_dt, &
                                      Hn, divQ, tgt, params=tp, dt_max=dt_max, n_Axb_its=n_its)
    open(newunit=uout, file=trim(tout), access='stream', form='unformatted', status='replace', action='write')
    write(uout) Hn, dHi_dt, divQ, dt_max, int(n_its, c_int)
    close(uout)
    write(*, '(a,i0,a,es10.3)') 'thickness step done: krylov ', n_its, ' dt_max ', dt_max
  end subroutine thickness_step

end program ufe_fortran_driver
