"""ctypes binding of oracle/_build/liboracle.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The engine never imports it.
It reuses the engine package's ctypes structures only as a data layout.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "liboracle.so")

_o = None


def _L():
    import __graft_entry__ as ge
    ge.load_package()
    from ufemism2_amd import _lib
    return _lib


def lib():
    global _o
    if _o is not None:
        return _o
    if not os.path.exists(SO):
        subprocess.run(["make", "-C", HERE], check=True, capture_output=True)
    L = _L()
    o = C.CDLL(SO)
    dp, ip = L.dp, L.ip
    o.oracle_csr_matvec.argtypes = [C.c_int, ip, ip, dp, dp, dp]
    o.oracle_tridiag.argtypes = [C.c_int, dp, dp, dp, dp, dp]
    o.oracle_gmres_ilu0.argtypes = [C.c_int, ip, ip, dp, dp, dp, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_int, dp]
    o.oracle_solve_ssa.argtypes = [C.POINTER(L.MeshDesc), C.POINTER(L.Params), dp, dp, dp, dp, dp,
                                   ip, ip, dp, dp, dp, dp, C.POINTER(L.SolveInfo)]
    o.oracle_solve_diva.argtypes = [C.POINTER(L.MeshDesc), C.POINTER(L.Params), dp, dp, dp, dp, dp,
                                    ip, ip, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                    C.POINTER(L.SolveInfo)]
    o.oracle_ssa_residual.restype = C.c_double
    o.oracle_ssa_residual.argtypes = [C.POINTER(L.MeshDesc), C.POINTER(L.Params), dp, dp, dp, dp, dp,
                                      ip, dp, dp, dp, dp, dp, dp]
    o.oracle_diva_residual.restype = C.c_double
    o.oracle_diva_residual.argtypes = [C.POINTER(L.MeshDesc), C.POINTER(L.Params), dp, dp, dp, dp, dp, dp, dp,
                                       dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    o.oracle_ssa_system.argtypes = [C.POINTER(L.MeshDesc), C.POINTER(L.Params), dp, dp, dp, dp, dp,
                                    dp, dp, C.c_int, ip, ip, dp, dp]
    o.oracle_fraction_gr_b.argtypes = [C.POINTER(L.MeshDesc), dp, dp, dp, dp]
    o.oracle_heat_columns.argtypes = [C.POINTER(L.HeatInputs), dp, ip, ip]
    o.oracle_heat_robin.argtypes = [C.POINTER(L.HeatRobinInputs), ip, dp, ip]
    o.oracle_thickness_semiimplicit.argtypes = [C.c_int, dp, C.POINTER(L.VertexGeometry),
                                                C.POINTER(L.ThicknessParams), C.POINTER(L.ThicknessInputs),
                                                dp, dp, dp, dp, ip]
    o.oracle_heat_upwind.argtypes = [C.POINTER(L.MeshDesc), ip, ip, dp, dp, dp, dp, dp, dp, dp, dp]
    o.oracle_strain_heating.argtypes = [C.c_int, C.c_int, dp, dp, C.POINTER(dp), C.c_double, dp]
    o.oracle_frictional_heating.argtypes = [C.c_int, C.c_int, dp, dp, ip, dp]
    o.oracle_zeta_tridiagonals.argtypes = [C.c_int, dp, dp, dp, dp, dp, dp, dp]
    o.oracle_set_sliding_cap.argtypes = [dp, dp]
    _o = o
    return o


_cap_keep = None


def set_sliding_cap(alpha_sq=None, effective_pressure=None):
    """Coulomb limit of the Tsai2015 / Schoof2005 laws for the oracle's later
    solves (caller-order [nV] arrays, kept alive here); no arguments clears."""
    global _cap_keep
    L = _L()
    if alpha_sq is None:
        _cap_keep = None
        lib().oracle_set_sliding_cap(None, None)
        return
    _cap_keep = (L.f64(alpha_sq).copy(), L.f64(effective_pressure).copy())
    lib().oracle_set_sliding_cap(L.ptr(_cap_keep[0]), L.ptr(_cap_keep[1]))


def csr_matvec(ptr, ind, val, x):
    L = _L()
    ptr, ind, val, x = L.i32(ptr), L.i32(ind), L.f64(val), L.f64(x)
    y = np.zeros(ptr.size - 1)
    lib().oracle_csr_matvec(ptr.size - 1, L.ptr(ptr, L.ip), L.ptr(ind, L.ip), L.ptr(val), L.ptr(x), L.ptr(y))
    return y


def tridiag(a, b, c, d):
    """Thomas solve; a[0] unused, c[n-1] unused (arrays of length n)."""
    L = _L()
    a, b, c, d = (L.f64(np.array(t, dtype=np.float64)) for t in (a, b, c, d))
    x = np.zeros(b.size)
    lib().oracle_tridiag(b.size, L.ptr(a), L.ptr(b), L.ptr(c), L.ptr(d), L.ptr(x))
    return x


def gmres_ilu0(A, b, x0=None, rtol=1e-8, abstol=0.0, maxit=10000, restart=30):
    x, its, _ = gmres_ilu0_stats(A, b, x0, rtol, abstol, maxit, restart)
    return x, its


def gmres_ilu0_stats(A, b, x0=None, rtol=1e-8, abstol=0.0, maxit=10000, restart=30):
    """(x, iterations, final preconditioned residual norm)."""
    L = _L()
    A = A.tocsr()
    ptr, ind, val = L.i32(A.indptr), L.i32(A.indices), L.f64(A.data)
    b = L.f64(b)
    x = L.f64(np.zeros_like(b) if x0 is None else x0.copy())
    pres = C.c_double(0)
    its = lib().oracle_gmres_ilu0(b.size, L.ptr(ptr, L.ip), L.ptr(ind, L.ip), L.ptr(val), L.ptr(b),
                                  L.ptr(x), rtol, abstol, maxit, restart, int(x0 is None), C.byref(pres))
    return x, its, pres.value


class Oracle:
    """Oracle solves on one mesh descriptor (MeshDescriptor from the engine package)."""

    def __init__(self, md):
        self.md = md

    def solve_ssa(self, params, inp, u0=None, v0=None):
        L = _L()
        nT = self.md.mesh.nT
        u = L.f64(np.zeros(nT) if u0 is None else u0).copy()
        v = L.f64(np.zeros(nT) if v0 is None else v0).copy()
        info = L.SolveInfo()
        bcm, bcu, bcv = inp.get("bc_mask"), inp.get("bc_u"), inp.get("bc_v")
        rc = lib().oracle_solve_ssa(self.md.byref(), C.byref(params), L.ptr(inp["Hi"]), L.ptr(inp["Hs"]),
                                    L.ptr(inp["A_flow"]), L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]),
                                    L.ptr(inp.get("mask_grounded"), L.ip), L.ptr(bcm, L.ip), L.ptr(bcu),
                                    L.ptr(bcv), L.ptr(u), L.ptr(v), C.byref(info))
        return rc, u, v, info

    def solve_diva(self, params, inp, state=None, with_info=True):
        """with_info=False passes no info structure (the oracle then skips its
        diagnostic residual); info is returned as None."""
        L = _L()
        nT, nz = self.md.mesh.nT, len(self.md.zeta)
        st = {k: np.zeros(nT) for k in ("u", "v", "ubase", "vbase", "tbx", "tby")}
        st["eta3b"] = np.zeros(nT * nz)
        if state:
            for k, a in state.items():
                st[k] = L.f64(a).copy()
        u3 = np.zeros(nT * nz); v3 = np.zeros(nT * nz)
        info = L.SolveInfo()
        rc = lib().oracle_solve_diva(self.md.byref(), C.byref(params), L.ptr(inp["Hi"]), L.ptr(inp["Hs"]),
                                     L.ptr(inp["A_flow"]), L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]),
                                     L.ptr(inp.get("mask_grounded"), L.ip), L.ptr(inp.get("bc_mask"), L.ip),
                                     L.ptr(inp.get("bc_u")), L.ptr(inp.get("bc_v")),
                                     L.ptr(st["u"]), L.ptr(st["v"]), L.ptr(st["ubase"]), L.ptr(st["vbase"]),
                                     L.ptr(u3), L.ptr(v3), L.ptr(st["eta3b"]), L.ptr(st["tbx"]),
                                     L.ptr(st["tby"]), C.byref(info) if with_info else None)
        st["u3d"], st["v3d"] = u3, v3
        return rc, st, (info if with_info else None)

    def residual(self, params, inp, u, v):
        L = _L()
        n = 2 * self.md.mesh.nT
        r = np.zeros(n); b = np.zeros(n)
        rel = lib().oracle_ssa_residual(self.md.byref(), C.byref(params), L.ptr(inp["Hi"]), L.ptr(inp["Hs"]),
                                        L.ptr(inp["A_flow"]), L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]),
                                        L.ptr(inp.get("bc_mask"), L.ip), L.ptr(inp.get("bc_u")),
                                        L.ptr(inp.get("bc_v")), L.ptr(L.f64(u)), L.ptr(L.f64(v)),
                                        L.ptr(r), L.ptr(b))
        return rel, r, b

    def diva_residual(self, params, inp, u, v, eta_3D_b=None, tau_bx_b=None, tau_by_b=None, u_base_b=None,
                      v_base_b=None):
        """One DIVA state evaluation at u_vav = (u, v) from the given state
        (3-D arrays [k][ti]); the same dict as VelocitySolver.diva_residual."""
        L = _L()
        nT, nz = self.md.mesh.nT, len(self.md.zeta)
        z = np.zeros(nT)
        eta = L.f64(np.zeros(nT * nz) if eta_3D_b is None else eta_3D_b).copy()
        st = [L.f64(z if a is None else a) for a in (tau_bx_b, tau_by_b, u_base_b, v_base_b)]
        out = {"r": np.zeros(2 * nT), "b": np.zeros(2 * nT), "F1_3D_b": np.zeros(nT * nz),
               "F2_b_surface": np.zeros(nT), "beta_b": np.zeros(nT), "beta_eff_b": np.zeros(nT)}
        out["rel_residual"] = lib().oracle_diva_residual(
            self.md.byref(), C.byref(params), L.ptr(inp["Hi"]), L.ptr(inp["Hs"]), L.ptr(inp["A_flow"]),
            L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]), L.ptr(L.f64(u)), L.ptr(L.f64(v)), L.ptr(eta),
            *(L.ptr(a) for a in st), L.ptr(out["r"]), L.ptr(out["b"]), L.ptr(out["F1_3D_b"]),
            L.ptr(out["F2_b_surface"]), L.ptr(out["beta_b"]), L.ptr(out["beta_eff_b"]))
        out["eta_3D_b"] = eta
        return out

    def system(self, params, inp, u, v):
        import scipy.sparse as sp
        L = _L()
        nT = self.md.mesh.nT
        cap = 4 * (self.md.ops.bb.nnz + 4 * nT)
        ptr = np.zeros(2 * nT + 1, np.int32); ind = np.zeros(cap, np.int32)
        val = np.zeros(cap); b = np.zeros(2 * nT)
        nnz = lib().oracle_ssa_system(self.md.byref(), C.byref(params), L.ptr(inp["Hi"]), L.ptr(inp["Hs"]),
                                      L.ptr(inp["A_flow"]), L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]),
                                      L.ptr(L.f64(u)), L.ptr(L.f64(v)), cap, L.ptr(ptr, L.ip),
                                      L.ptr(ind, L.ip), L.ptr(val), L.ptr(b))
        assert nnz > 0
        return sp.csr_matrix((val[:nnz], ind[:nnz], ptr), shape=(2 * nT, 2 * nT)), b

    def fraction_gr_b(self, Hi, Hb, SL):
        L = _L()
        out = np.zeros(self.md.mesh.nT)
        lib().oracle_fraction_gr_b(self.md.byref(), L.ptr(L.f64(Hi)), L.ptr(L.f64(Hb)), L.ptr(L.f64(SL)),
                                   L.ptr(out))
        return out


def heat_robin(rin, is_unstable, Ti):
    """rin: HeatRobinInputs; returns (rc, new Ti, number replaced)."""
    L = _L()
    T = L.f64(Ti).copy()
    flag = L.i32(is_unstable)
    n = C.c_int(0)
    rc = lib().oracle_heat_robin(C.byref(rin), L.ptr(flag, L.ip), L.ptr(T), C.byref(n))
    return rc, T, n.value


def heat_columns(hin):
    """hin: HeatInputs structure (with its arrays kept alive by the caller)."""
    L = _L()
    n = hin.nV * hin.nz
    T = np.zeros(n); unst = np.zeros(hin.nV, np.int32); nun = C.c_int(0)
    rc = lib().oracle_heat_columns(C.byref(hin), L.ptr(T), L.ptr(unst, L.ip), C.byref(nun))
    return rc, T, unst, nun.value


def thickness_semiimplicit(mesh, geom, Hi, Hb, SL, u_vav_b, v_vav_b, SMB, BMB, LMB, fraction_margin, mask_noice,
                           dt, dHi_dt_target=None, BC_prescr_mask=None, BC_prescr_Hi=None, dHi_dt_guess=None,
                           **params):
    """One semi-implicit thickness step of the oracle, same arguments and
    returns as VelocitySolver.calc_dHi_dt_semiimplicit (velocities required):
    (rc, dHi_dt, Hi_tplusdt, divQ, dt_max, n_its)."""
    L = _L()
    from ufemism2_amd import velocity as V
    nV = mesh.nV
    p = V.thickness_params(**params)
    inp, _keep = V.make_thickness_inputs(nV, Hi, Hb, SL, u_vav_b, v_vav_b, SMB, BMB, LMB, fraction_margin,
                                         mask_noice, dt, dHi_dt_target, BC_prescr_mask, BC_prescr_Hi,
                                         dHi_dt_guess)
    keep = (L.i32(mesh.C_ptr), L.i32(mesh.C_ind), L.i32(mesh.VBI), L.f64(geom.A), L.f64(geom.Cw),
            L.i32(geom.ETri), L.f64(mesh.V))
    g = L.VertexGeometry(int(keep[0][-1]), L.ptr(keep[0], L.ip), L.ptr(keep[1], L.ip), L.ptr(keep[2], L.ip),
                         L.ptr(keep[3]), L.ptr(keep[4]), L.ptr(keep[5], L.ip))
    Hn, rate, divQ = np.zeros(nV), np.zeros(nV), np.zeros(nV)
    dt_max, its = C.c_double(), C.c_int()
    rc = lib().oracle_thickness_semiimplicit(nV, L.ptr(keep[6]), C.byref(g), C.byref(p), C.byref(inp), L.ptr(Hn),
                                             L.ptr(rate), L.ptr(divQ), C.byref(dt_max), C.byref(its))
    return rc, rate, Hn, divQ, dt_max.value, its.value


def heat_upwind(md, Ti, Hi, u_vav, v_vav, u_3D_b, v_3D_b):
    """calc_upwind_heat_flux_derivatives of the oracle: (rc, u_dTdxp, v_dTdyp)."""
    L = _L()
    m = md.mesh
    nz = len(md.zeta)
    ox, oy = np.zeros(nz * m.nV), np.zeros(nz * m.nV)
    keep = [L.i32(m.iTri_ptr), L.i32(m.iTri_ind)] + [L.f64(x) for x in (Ti, Hi, u_vav, v_vav, u_3D_b, v_3D_b)]
    rc = lib().oracle_heat_upwind(md.byref(), L.ptr(keep[0], L.ip), L.ptr(keep[1], L.ip),
                                  *(L.ptr(x) for x in keep[2:]), L.ptr(ox), L.ptr(oy))
    return rc, ox, oy


def strain_heating(nV, nz, Hi, A_flow, strain_rates, glen_n=3.0):
    L = _L()
    keep = [L.f64(x) for x in strain_rates]
    arr = (L.dp * 9)(*(L.ptr(x) for x in keep))
    Phi = np.zeros(nz * nV)
    lib().oracle_strain_heating(nV, nz, L.ptr(L.f64(Hi)), L.ptr(L.f64(A_flow)), arr, glen_n, L.ptr(Phi))
    return Phi


def frictional_heating(beta, uabs_base, mask_grounded, no_sliding=False):
    L = _L()
    Q = np.zeros(len(beta))
    mk = L.i32(np.asarray(mask_grounded).astype(np.int32))
    lib().oracle_frictional_heating(len(beta), int(no_sliding), L.ptr(L.f64(beta)), L.ptr(L.f64(uabs_base)),
                                    L.ptr(mk, L.ip), L.ptr(Q))
    return Q


def _aa_args(aa, zs):
    L = _L()
    keep = [L.i32(aa.ptr), L.i32(aa.ind), L.f64(aa.vals["ddx"]), L.f64(aa.vals["ddy"]), L.f64(zs)]
    return keep, [aa.nnz, L.ptr(keep[0], L.ip), L.ptr(keep[1], L.ip), L.ptr(keep[2]), L.ptr(keep[3]),
                  L.ptr(keep[4])]


def _setup_bpa_sigs(o):
    if getattr(o, "_bpa", False):
        return
    L = _L()
    dp, ip = L.dp, L.ip
    PD = C.POINTER(dp)
    md, pr = C.POINTER(L.MeshDesc), C.POINTER(L.Params)
    aa = [C.c_int, ip, ip, dp, dp, dp]
    o.oracle_zeta_gradients.argtypes = [md, *aa, dp, dp, dp, dp, PD, PD, PD]
    o.oracle_ops3d_apply.argtypes = [md, *aa, dp, dp, dp, dp, dp, PD, PD, PD, dp, PD, dp, dp]
    o.oracle_solve_bpa.argtypes = [md, pr, *aa, dp, dp, dp, dp, dp, ip, ip, dp, dp, dp, dp,
                                   C.POINTER(L.SolveInfo)]
    o.oracle_bpa_residual.restype = C.c_double
    o.oracle_bpa_residual.argtypes = [md, pr, *aa, dp, dp, dp, dp, dp, ip, dp, dp, dp, dp, dp, dp, dp, dp]
    o._bpa = True


def _ptrs(arrs):
    L = _L()
    return (L.dp * len(arrs))(*[L.ptr(a) for a in arrs])


def zeta_gradients(md, aa, zs, Hi, Hs, dHi_dt=None, dHs_dt=None):
    """calc_zeta_gradients; returns (ak [7][nz*nV], bk [6][nz*nT], bks [6][(nz-1)*nT]) (first index
    fastest)."""
    L = _L()
    o = lib(); _setup_bpa_sigs(o)
    nV, nT, nz = md.mesh.nV, md.mesh.nT, len(md.zeta)
    ak = [np.zeros(nV * nz) for _ in range(7)]
    bk = [np.zeros(nT * nz) for _ in range(6)]
    bks = [np.zeros(nT * (nz - 1)) for _ in range(6)]
    keep, args = _aa_args(aa, zs)
    ins = [L.f64(a) if a is not None else None for a in (Hi, Hs, dHi_dt, dHs_dt)]
    rc = o.oracle_zeta_gradients(md.byref(), *args, *(L.ptr(a) for a in ins), _ptrs(ak), _ptrs(bk), _ptrs(bks))
    assert rc == 0
    return ak, bk, bks


def ops3d_apply(md, aa, zs, Hi, Hs, f_bk, f_ak, f_bks):
    """The 3-D operators of the BPA applied to fields (first index fastest):
    dict with bk_bk [7], bk_ak [2], ak_bk [2], bk_bks, bks_bk [2] (map, ddz),
    map_bks_ak, map_ak_bks."""
    L = _L()
    o = lib(); _setup_bpa_sigs(o)
    nV, nT, nz = md.mesh.nV, md.mesh.nT, len(md.zeta)
    out = {"bk_bk": [np.zeros(nT * nz) for _ in range(7)], "bk_ak": [np.zeros(nV * nz) for _ in range(2)],
           "ak_bk": [np.zeros(nT * nz) for _ in range(2)], "bk_bks": np.zeros(nT * (nz - 1)),
           "bks_bk": [np.zeros(nT * nz) for _ in range(2)], "map_bks_ak": np.zeros(nV * nz),
           "map_ak_bks": np.zeros(nT * (nz - 1))}
    keep, args = _aa_args(aa, zs)
    ins = [L.f64(a) for a in (Hi, Hs, f_bk, f_ak, f_bks)]
    rc = o.oracle_ops3d_apply(md.byref(), *args, *(L.ptr(a) for a in ins), _ptrs(out["bk_bk"]),
                              _ptrs(out["bk_ak"]), _ptrs(out["ak_bk"]), L.ptr(out["bk_bks"]),
                              _ptrs(out["bks_bk"]), L.ptr(out["map_bks_ak"]), L.ptr(out["map_ak_bks"]))
    assert rc == 0
    return out


def solve_bpa(md, aa, zs, params, inp, u0=None, v0=None):
    """solve_BPA; inp as Oracle.solve_ssa's plus bc_mask/bc_u/bc_v on the bk
    grid ([nz*nT], first index fastest).  Returns (rc, u_bk, v_bk, info)."""
    L = _L()
    o = lib(); _setup_bpa_sigs(o)
    nT, nz = md.mesh.nT, len(md.zeta)
    u = L.f64(np.zeros(nT * nz) if u0 is None else u0).copy()
    v = L.f64(np.zeros(nT * nz) if v0 is None else v0).copy()
    info = L.SolveInfo()
    keep, args = _aa_args(aa, zs)
    rc = o.oracle_solve_bpa(md.byref(), C.byref(params), *args, L.ptr(inp["Hi"]), L.ptr(inp["Hs"]),
                            L.ptr(inp["A_flow"]), L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]),
                            L.ptr(inp.get("mask_grounded"), L.ip), L.ptr(inp.get("bc_mask"), L.ip),
                            L.ptr(inp.get("bc_u")), L.ptr(inp.get("bc_v")), L.ptr(u), L.ptr(v), C.byref(info))
    return rc, u, v, info


def bpa_residual(md, aa, zs, params, inp, u, v):
    """One BPA state evaluation at (u, v) [nz*nT]: dict with rel_residual, r,
    b [2*nT*nz] (rows 2*(ti*nz+k)+uv), eta_bk [nz*nT], beta_b [nT]."""
    L = _L()
    o = lib(); _setup_bpa_sigs(o)
    nT, nz = md.mesh.nT, len(md.zeta)
    out = {"r": np.zeros(2 * nT * nz), "b": np.zeros(2 * nT * nz), "eta_bk": np.zeros(nT * nz),
           "beta_b": np.zeros(nT)}
    keep, args = _aa_args(aa, zs)
    out["rel_residual"] = o.oracle_bpa_residual(
        md.byref(), C.byref(params), *args, L.ptr(inp["Hi"]), L.ptr(inp["Hs"]), L.ptr(inp["A_flow"]),
        L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]), L.ptr(inp.get("bc_mask"), L.ip), L.ptr(inp.get("bc_u")),
        L.ptr(inp.get("bc_v")), L.ptr(L.f64(u)), L.ptr(L.f64(v)), L.ptr(out["r"]), L.ptr(out["b"]),
        L.ptr(out["eta_bk"]), L.ptr(out["beta_b"]))
    return out


def bpa_system(md, aa, zs, params, inp, u, v):
    """The BPA system assembled at (u, v): (scipy CSR A, b), rows 2*(ti*nz+k)+uv."""
    import scipy.sparse as sp
    L = _L()
    o = lib(); _setup_bpa_sigs(o)
    if not getattr(o, "_bpa_sys", False):
        dp, ip = L.dp, L.ip
        o.oracle_bpa_system.argtypes = [C.POINTER(L.MeshDesc), C.POINTER(L.Params), C.c_int, ip, ip, dp, dp, dp,
                                        dp, dp, dp, dp, dp, ip, dp, dp, dp, dp, C.c_int, ip, ip, dp, dp]
        o._bpa_sys = True
    nT, nz = md.mesh.nT, len(md.zeta)
    n = 2 * nT * nz
    cap = 2 * 3 * 2 * md.ops.bb.nnz * nz + 8 * n
    ptr = np.zeros(n + 1, np.int32); ind = np.zeros(cap, np.int32); val = np.zeros(cap); b = np.zeros(n)
    keep, args = _aa_args(aa, zs)
    nnz = o.oracle_bpa_system(md.byref(), C.byref(params), *args, L.ptr(inp["Hi"]), L.ptr(inp["Hs"]),
                              L.ptr(inp["A_flow"]), L.ptr(inp["beta_sq"]), L.ptr(inp["fgr_b"]),
                              L.ptr(inp.get("bc_mask"), L.ip), L.ptr(inp.get("bc_u")), L.ptr(inp.get("bc_v")),
                              L.ptr(L.f64(u)), L.ptr(L.f64(v)), cap, L.ptr(ptr, L.ip), L.ptr(ind, L.ip),
                              L.ptr(val), L.ptr(b))
    assert nnz > 0
    return sp.csr_matrix((val[:nnz], ind[:nnz], ptr), shape=(n, n)), b
