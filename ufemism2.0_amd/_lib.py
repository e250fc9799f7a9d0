"""ctypes view of include/ufe_hip.h and the loader for libufe_hip.so.

The library is loaded from this directory only; if it is missing the import
of any compute entry point raises, there is no fallback path.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libufe_hip.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)

UFE_OK = 0
UFE_WARN_NOT_CONVERGED = 1
UFE_ERR_ARGUMENT = -1
UFE_ERR_DEVICE = -2
UFE_ERR_DIVERGED = -3
UFE_ERR_UNSUPPORTED = -4
UFE_ERR_UNSTABLE = -5
UFE_ERR_MESH = -6

SLIDING = {"no_sliding": 0, "Weertman": 1, "Budd": 2, "Coulomb": 3, "Tsai2015": 4, "Schoof2005": 5,
           "Zoet-Iverson": 6}
BC = {"infinite": 0, "zero": 1, "periodic_ISMIP-HOM": 2, "infinite_SSA_icestream": 3}
PREC_FP64, PREC_MIXED = 0, 1
KRYLOV_BICGSTAB, KRYLOV_GMRES = 0, 1
PC_NONE, PC_JACOBI, PC_BLOCK_JACOBI, PC_AMG = 0, 1, 2, 3
SMOOTHER_BLOCK_JACOBI, SMOOTHER_L1_BLOCK_JACOBI = 0, 1
STOP_L2UV, STOP_RESIDUAL = 0, 1
STRESS_SSA, STRESS_DIVA = 0, 1


class MeshDesc(C.Structure):
    _fields_ = [
        ("nV", C.c_int), ("nT", C.c_int), ("nz", C.c_int),
        ("V", dp), ("Tri", ip), ("TriC", ip), ("TriBI", ip), ("TriGC", dp), ("zeta", dp),
        ("ab_nnz", C.c_int), ("ab_ptr", ip), ("ab_ind", ip), ("ab_map", dp), ("ab_ddx", dp), ("ab_ddy", dp),
        ("ba_nnz", C.c_int), ("ba_ptr", ip), ("ba_ind", ip), ("ba_map", dp), ("ba_ddx", dp), ("ba_ddy", dp),
        ("bb_nnz", C.c_int), ("bb_ptr", ip), ("bb_ind", ip),
        ("bb_ddx", dp), ("bb_ddy", dp), ("bb_d2dx2", dp), ("bb_d2dxdy", dp), ("bb_d2dy2", dp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("glens_n", C.c_double), ("epsilon_sq_0", C.c_double), ("visc_eff_min", C.c_double),
        ("vel_max", C.c_double), ("visc_it_nit", C.c_int), ("visc_it_relax", C.c_double),
        ("visc_it_norm_dUV_tol", C.c_double), ("rtol", C.c_double), ("abstol", C.c_double),
        ("include_crossterms", C.c_int), ("sliding_law", C.c_int), ("slid_weertman_m", C.c_double),
        ("slid_budd_q_plastic", C.c_double), ("slid_budd_u_threshold", C.c_double),
        ("slid_delta_v", C.c_double), ("slid_beta_max", C.c_double),
        ("do_gl_subgrid_friction", C.c_int), ("subgrid_friction_exponent", C.c_double),
        ("bc_u", C.c_int * 4), ("bc_v", C.c_int * 4),
        ("precision", C.c_int), ("krylov", C.c_int), ("precond", C.c_int),
        ("gmres_restart", C.c_int), ("krylov_maxit", C.c_int), ("warm_start", C.c_int),
        ("inner_rtol", C.c_double), ("stop_rule", C.c_int), ("nonlinear_rtol", C.c_double),
        ("check_every", C.c_int),
        ("accel_depth", C.c_int),
        ("accel_mixing", C.c_double),
        ("accel_restart", C.c_double),
        ("accel_continuation", C.c_int),
        ("precond_fp32", C.c_int), ("ismip_hom_L", C.c_double),
        ("slid_zi_p", C.c_double), ("slid_zi_ut", C.c_double),
        ("amg_smoother", C.c_int),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("n_visc_its", C.c_int), ("n_Axb_its", C.c_int), ("n_refinements", C.c_int),
        ("L2_uv", C.c_double), ("rel_residual", C.c_double),
        ("visc_it_relax_applied", C.c_double), ("epsilon_sq_0_applied", C.c_double),
        ("seconds", C.c_double), ("n_clamped", C.c_int), ("n_fallbacks", C.c_int),
        ("rel_residual_unscaled", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HeatInputs(C.Structure):
    _fields_ = [
        ("nV", C.c_int), ("nz", C.c_int), ("zeta", dp), ("Ti", dp),
        ("u_3D", dp), ("v_3D", dp), ("w_3D", dp), ("u_dTdxp_upwind", dp), ("v_dTdyp_upwind", dp),
        ("dzeta_dx", dp), ("dzeta_dy", dp), ("dzeta_dz", dp), ("dzeta_dt", dp), ("Phi", dp),
        ("Hi_eff", dp), ("T_surf", dp), ("Q_base_grnd", dp), ("mask", ip),
        ("dt", C.c_double), ("Hi_min_thermo", C.c_double), ("fraction_gr", dp),
    ]


class HeatRobinInputs(C.Structure):
    _fields_ = [
        ("nV", C.c_int), ("nz", C.c_int), ("zeta", dp), ("Hi_eff", dp), ("T_surf", dp), ("SMB", dp),
        ("geothermal_flux", dp), ("mask", ip), ("Hi_min_thermo", C.c_double),
    ]


class VertexGeometry(C.Structure):
    _fields_ = [("nC", C.c_int), ("C_ptr", ip), ("C_ind", ip), ("VBI", ip), ("A", dp), ("Cw", dp),
                ("ETri", ip)]


class ThicknessParams(C.Structure):
    _fields_ = [
        ("fs", C.c_double), ("bc_H", C.c_int * 4), ("dt_ice_max", C.c_double), ("dt_ice_min", C.c_double),
        ("Hi_min", C.c_double), ("rtol", C.c_double), ("abstol", C.c_double), ("krylov", C.c_int),
        ("gmres_restart", C.c_int), ("maxit", C.c_int),
    ]


class BpaInputs(C.Structure):
    _fields_ = [
        ("Hi", dp), ("Hs", dp), ("A_flow", dp), ("beta_sq", dp), ("fraction_gr_b", dp),
        ("mask_grounded_ice", ip), ("BC_prescr_mask_bk", ip), ("BC_prescr_u_bk", dp), ("BC_prescr_v_bk", dp),
    ]


class ThicknessInputs(C.Structure):
    _fields_ = [
        ("Hi", dp), ("Hb", dp), ("SL", dp), ("u_vav_b", dp), ("v_vav_b", dp), ("SMB", dp), ("BMB", dp),
        ("LMB", dp), ("fraction_margin", dp), ("dHi_dt_target", dp), ("mask_noice", ip),
        ("BC_prescr_mask", ip), ("BC_prescr_Hi", dp), ("dHi_dt_guess", dp), ("dt", C.c_double),
    ]


def ptr(a, kind=dp):
    """Pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return C.cast(None, kind)
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(kind)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


_lib = None


def _sig(L, name, restype=None, argtypes=None):
    """Declare a signature if the symbol exists (missing symbols are reported
    by the export test, not here)."""
    if not hasattr(L, name):
        return
    f = getattr(L, name)
    if restype is not None:
        f.restype = restype
    if argtypes is not None:
        f.argtypes = argtypes


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python __graft_entry__.py` (build) first")
    L = C.CDLL(LIB_PATH)
    _sig(L, 'ufe_last_error', restype=C.c_char_p)
    _sig(L, 'ufe_last_error', argtypes=[C.c_void_p])
    _sig(L, 'ufe_params_default', argtypes=[C.POINTER(Params)])
    _sig(L, 'ufe_version', restype=C.c_int)
    _sig(L, 'ufe_ctx_create', argtypes=[C.POINTER(C.c_void_p), C.c_int, C.POINTER(MeshDesc)])
    _sig(L, 'ufe_ctx_destroy', argtypes=[C.c_void_p])
    _sig(L, 'ufe_set_params', argtypes=[C.c_void_p, C.POINTER(Params)])
    _sig(L, 'ufe_solve_ssa', argtypes=[C.c_void_p, dp, dp, dp, dp, dp, ip, ip, dp, dp, dp, dp, C.POINTER(SolveInfo)])
    _sig(L, 'ufe_solve_diva', argtypes=[C.c_void_p, dp, dp, dp, dp, dp, ip, ip, dp, dp,
                                 dp, dp, dp, dp, dp, dp, dp, dp, dp, C.POINTER(SolveInfo)])
    _sig(L, 'ufe_upload_inputs', argtypes=[C.c_void_p, dp, dp, dp, dp, dp, ip, ip, dp, dp])
    _sig(L, 'ufe_upload_velocity', argtypes=[C.c_void_p, dp, dp])
    _sig(L, 'ufe_run', argtypes=[C.c_void_p, C.c_int, C.POINTER(SolveInfo)])
    _sig(L, 'ufe_download_velocity', argtypes=[C.c_void_p, dp, dp])
    _sig(L, 'ufe_upload_diva_state', argtypes=[C.c_void_p, dp, dp, dp, dp, dp])
    _sig(L, 'ufe_download_diva', argtypes=[C.c_void_p, dp, dp, dp, dp, dp, dp, dp])
    _sig(L, 'ufe_last_timings', argtypes=[C.c_void_p, dp])
    _sig(L, 'ufe_profile_spmv', argtypes=[C.c_void_p, C.c_int, dp, dp])
    _sig(L, 'ufe_profile_copy', argtypes=[C.c_void_p, C.c_long, C.c_int, dp])
    _sig(L, 'ufe_kernel_timing', argtypes=[C.c_void_p, C.c_int])
    _sig(L, 'ufe_set_amg_options', argtypes=[C.c_void_p, C.c_double, C.c_double, C.c_int])
    _sig(L, 'ufe_heat_robin', argtypes=[C.c_void_p, C.POINTER(HeatRobinInputs), ip, dp, ip])
    _sig(L, 'ufe_kernel_timing_read', argtypes=[C.c_void_p, C.c_int, C.POINTER(C.c_longlong), dp, dp])
    fp = C.POINTER(C.c_float)
    _sig(L, 'ufe_mpnn_num_weights', restype=C.c_int, argtypes=[])
    _sig(L, 'ufe_mpnn_create', argtypes=[C.POINTER(C.c_void_p), C.c_int, fp])
    _sig(L, 'ufe_mpnn_destroy', argtypes=[C.c_void_p])
    _sig(L, 'ufe_mpnn_get_weights', argtypes=[C.c_void_p, fp])
    _sig(L, 'ufe_mpnn_forward', argtypes=[C.c_void_p, C.c_int, fp, ip, fp])
    _sig(L, 'ufe_mpnn_loss_grad', argtypes=[C.c_void_p, C.c_int, fp, ip, fp, fp, fp])
    _sig(L, 'ufe_mpnn_train', argtypes=[C.c_void_p, C.c_int, fp, ip, fp, C.c_int, C.c_float, fp])
    _sig(L, 'ufe_mpnn_features', argtypes=[C.c_void_p, fp])
    _sig(L, 'ufe_mpnn_initial_guess', argtypes=[C.c_void_p, C.c_void_p])
    _sig(L, 'ufe_set_graph_launch', argtypes=[C.c_void_p, C.c_int])
    _sig(L, 'ufe_set_sliding_cap', argtypes=[C.c_void_p, dp, dp])
    _sig(L, 'ufe_till_yield_stress', argtypes=[C.c_int, ip, ip, dp, dp, ip, ip, C.c_double, dp])
    _sig(L, 'ufe_ssa_residual', argtypes=[C.c_void_p, dp, dp, dp, dp, dp])
    _sig(L, 'ufe_diva_residual', argtypes=[C.c_void_p, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                           dp])
    _sig(L, 'ufe_csr_matvec', argtypes=[C.c_int, C.c_int, C.c_int, ip, ip, dp, dp, dp])
    _sig(L, 'ufe_csr_solve', argtypes=[C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_double,
                                C.c_int, C.c_int, ip, dp])
    _sig(L, 'ufe_csr_solve_ex', argtypes=[C.c_int, C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_double,
                                          C.c_int, C.c_int, C.c_int, C.c_int, ip, dp])
    _sig(L, 'ufe_fraction_gr_b', argtypes=[C.c_void_p, dp, dp, dp, dp])
    _sig(L, 'ufe_heat_columns_device', argtypes=[C.c_void_p, C.POINTER(HeatInputs), C.c_void_p, C.c_void_p,
                                                 C.POINTER(C.c_int)])
    _sig(L, 'ufe_heat_columns', argtypes=[C.c_void_p, C.POINTER(HeatInputs), dp, ip, ip])
    _sig(L, 'ufe_heat_columns_device_noalloc', argtypes=[C.POINTER(HeatInputs), C.c_void_p, C.c_void_p,
                                                         C.POINTER(C.c_int)])
    _sig(L, 'ufe_mesh_operators_ab', argtypes=[C.c_int, C.c_int, dp, ip, dp, ip, ip, ip, ip, ip, dp, dp, dp])
    _sig(L, 'ufe_mesh_operators_ba', argtypes=[C.c_int, C.c_int, dp, ip, dp, ip, ip, ip, ip, ip, dp, dp, dp])
    _sig(L, 'ufe_mesh_operators_bb2', argtypes=[C.c_int, ip, dp, ip, ip, ip, dp, dp, dp, dp, dp])
    _sig(L, 'ufe_mesh_operators_aa', argtypes=[C.c_int, dp, ip, ip, ip, ip, ip, dp, dp])
    _sig(L, 'ufe_heat_setup', argtypes=[C.c_void_p, ip, ip])
    _sig(L, 'ufe_heat_upwind_derivatives', argtypes=[C.c_void_p, dp, dp, dp, dp, dp, dp, dp, dp])
    _sig(L, 'ufe_heat_strain_heating', argtypes=[C.c_void_p, dp, dp, C.POINTER(dp), dp])
    _sig(L, 'ufe_heat_frictional_heating', argtypes=[C.c_void_p, dp, dp, ip, dp])
    _sig(L, 'ufe_thickness_params_default', argtypes=[C.POINTER(ThicknessParams)])
    _sig(L, 'ufe_zeta_setup', argtypes=[C.c_void_p, C.c_int, ip, ip, dp, dp, dp])
    _sig(L, 'ufe_zeta_gradients', argtypes=[C.c_void_p, dp, dp, dp, dp, C.POINTER(dp), C.POINTER(dp),
                                            C.POINTER(dp)])
    _sig(L, 'ufe_solve_bpa', argtypes=[C.c_void_p, C.POINTER(BpaInputs), dp, dp, C.POINTER(SolveInfo)])
    _sig(L, 'ufe_bpa_residual', argtypes=[C.c_void_p, C.POINTER(BpaInputs), dp, dp, dp, dp, dp, dp, dp])
    _sig(L, 'ufe_bpa_last_timings', argtypes=[C.c_void_p, dp])
    _sig(L, 'ufe_bpa_profile_spmv', argtypes=[C.c_void_p, C.c_int, dp, dp])
    _sig(L, 'ufe_thickness_setup', argtypes=[C.c_void_p, C.POINTER(VertexGeometry)])
    _sig(L, 'ufe_thickness_semiimplicit', argtypes=[C.c_void_p, C.POINTER(ThicknessParams),
                                                    C.POINTER(ThicknessInputs), dp, dp, dp, dp, ip])
    _lib = L
    return L


EXPORTED = [
    "ufe_params_default", "ufe_last_error", "ufe_version", "ufe_ctx_create", "ufe_ctx_destroy",
    "ufe_set_params", "ufe_solve_ssa", "ufe_solve_diva", "ufe_upload_inputs", "ufe_upload_velocity",
    "ufe_run", "ufe_download_velocity", "ufe_upload_diva_state", "ufe_download_diva",
    "ufe_last_timings", "ufe_profile_spmv", "ufe_profile_copy", "ufe_kernel_timing",
    "ufe_kernel_timing_read", "ufe_set_amg_options", "ufe_mpnn_num_weights", "ufe_mpnn_create",
    "ufe_mpnn_destroy", "ufe_mpnn_get_weights", "ufe_mpnn_forward", "ufe_mpnn_loss_grad", "ufe_mpnn_train",
    "ufe_mpnn_features", "ufe_mpnn_initial_guess", "ufe_set_graph_launch", "ufe_ssa_residual", "ufe_diva_residual", "ufe_csr_matvec", "ufe_csr_solve", "ufe_csr_solve_ex", "ufe_fraction_gr_b", "ufe_heat_columns", "ufe_heat_columns_device", "ufe_heat_columns_device_noalloc", "ufe_heat_robin",
    "ufe_mesh_operators_ab", "ufe_mesh_operators_ba", "ufe_mesh_operators_bb2", "ufe_mesh_operators_aa",
    "ufe_thickness_params_default", "ufe_thickness_setup", "ufe_thickness_semiimplicit",
    "ufe_heat_setup", "ufe_heat_upwind_derivatives", "ufe_heat_strain_heating", "ufe_heat_frictional_heating",
    "ufe_zeta_setup", "ufe_zeta_gradients", "ufe_solve_bpa", "ufe_bpa_residual", "ufe_bpa_last_timings",
    "ufe_bpa_profile_spmv", "ufe_set_sliding_cap", "ufe_till_yield_stress",
]


def check(rc, ctx=None, allow_warn=True):
    if rc == 0 or (allow_warn and rc > 0):
        return rc
    msg = lib().ufe_last_error(ctx)
    raise RuntimeError(f"libufe_hip error {rc}: {msg.decode() if msg else ''}")


def default_thickness_params() -> ThicknessParams:
    p = ThicknessParams()
    lib().ufe_thickness_params_default(C.byref(p))
    return p


def default_params() -> Params:
    p = Params()
    lib().ufe_params_default(C.byref(p))
    return p
