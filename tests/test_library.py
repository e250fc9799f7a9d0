"""The C-ABI library: loads, exports every entry point that include/ufe_hip.h
declares, fills the reference defaults, and rejects bad arguments on the host
without touching a device."""

import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ufe_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ufe_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported(L):
    lib = C.CDLL(L.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(L.EXPORTED) | set(names)


def test_python_binding_declares_every_symbol(L):
    assert set(declared_functions()) == set(L.EXPORTED)


def test_params_default_reference_values(L):
    """model_configuration.f90 defaults of the fields the velocity solve reads."""
    p = L.default_params()
    assert (p.glens_n, p.epsilon_sq_0, p.visc_eff_min, p.vel_max) == (3.0, 1e-8, 1e4, 5000.0)
    assert (p.visc_it_nit, p.visc_it_relax, p.visc_it_norm_dUV_tol) == (50, 0.2, 5e-5)
    assert (p.rtol, p.abstol, p.include_crossterms) == (1e-7, 1e-5, 1)
    assert (p.sliding_law, p.slid_weertman_m, p.slid_delta_v, p.slid_beta_max) == (1, 3.0, 1e-3, 1e20)
    assert (p.slid_budd_q_plastic, p.slid_budd_u_threshold) == (0.3, 100.0)
    assert (p.do_gl_subgrid_friction, p.subgrid_friction_exponent) == (1, 2.0)
    assert list(p.bc_u) == [0, 0, 0, 0] and list(p.bc_v) == [0, 0, 0, 0]
    assert p.accel_depth == 0 and p.stop_rule == 0


def test_params_struct_size_matches(L):
    """ufe_params_default writes exactly sizeof(ufe_params) bytes: a sentinel
    after the ctypes structure stays untouched."""
    lib = L.lib()
    n = C.sizeof(L.Params)
    buf = (C.c_ubyte * (n + 64))()
    for i in range(n, n + 64):
        buf[i] = 0xA5
    lib.ufe_params_default(C.cast(buf, C.POINTER(L.Params)))
    assert all(buf[i] == 0xA5 for i in range(n, n + 64))
    p = C.cast(buf, C.POINTER(L.Params)).contents
    assert p.accel_restart == 10.0 and p.check_every == 8


def test_null_arguments_rejected(L):
    lib = L.lib()
    assert lib.ufe_ctx_create(None, 0, None) == L.UFE_ERR_ARGUMENT
    assert lib.ufe_set_params(None, None) == L.UFE_ERR_ARGUMENT
    assert lib.ufe_run(None, 0, None) == L.UFE_ERR_ARGUMENT
    assert lib.ufe_heat_columns(None, None, None, None, None) == L.UFE_ERR_ARGUMENT
    assert lib.ufe_csr_matvec(-1, 1, 0, None, None, None, None, None) == L.UFE_ERR_ARGUMENT
    assert len(lib.ufe_last_error(None)) > 0


def test_operator_builder_two_call_convention(L, W):
    """Host operator builder: the sizing call (NULL arrays) reports nnz, the
    second fills them; a triangle with no neighbours makes the b->b operator
    report a mesh error (mesh_disc_calc_matrix_operators_2D.f90 needs >= 5)."""
    mesh = W.uniform_mesh(0, 100e3, 0, 100e3, 25e3, seed=1)
    lib = L.lib()
    nnz = C.c_int(0)
    nul_i, nul_d = C.cast(None, L.ip), C.cast(None, L.dp)
    TriC, TriGC = L.i32(mesh.TriC), L.f64(mesh.TriGC)
    rc = lib.ufe_mesh_operators_bb2(mesh.nT, L.ptr(TriC, L.ip), L.ptr(TriGC), C.byref(nnz), nul_i, nul_i,
                                    *([nul_d] * 5))
    assert rc == 0 and nnz.value >= 6 * mesh.nT
    bad = np.full_like(TriC, -1)
    rc = lib.ufe_mesh_operators_bb2(mesh.nT, L.ptr(bad, L.ip), L.ptr(TriGC), C.byref(nnz), nul_i, nul_i,
                                    *([nul_d] * 5))
    assert rc < 0
