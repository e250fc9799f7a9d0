"""Case files for the Fortran drivers, each one raw little-endian stream in
Fortran array order.  ufe_fortran_driver: the mesh in the reference's 1-based
layout (V(nV,2), Tri/TriC(nT,3) with TriC = 0 for none, TriGC(nT,2)), the
operators as 1-based CSR, a few solver controls and the inputs; for the BPA
also its controls, the a->a operators and zeta_stag.
ufe_fortran_heat_driver: the column fields (nV,nz), vertex index fastest,
and the per-vertex fields of one heat-equation step."""

from __future__ import annotations

import os
import subprocess

import numpy as np

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "ufe_fortran_driver")
HEAT_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "ufe_fortran_heat_driver")
HEAT_COLUMN_FIELDS = ("Ti", "u_3D", "v_3D", "w_3D", "u_dTdxp_upwind", "v_dTdyp_upwind", "dzeta_dx", "dzeta_dy",
                      "dzeta_dz", "dzeta_dt", "Phi")


def write_case(path, w, stress="SSA", sliding=None, **ctl):
    """sliding (SSA / DIVA): dict(law=UFE_SLIDING_*, m=, zi_p=, zi_ut=) with
    alpha_sq and N (the Coulomb cap of Tsai2015 / Schoof2005) or N, phi and
    icefree_land (the driver forms the till yield stress from them and the
    mesh connectivity, in place of beta_sq); flagged by bit 4 of the stress
    code in the header."""
    m, ops = w.mesh, w.ops
    i32 = lambda a: np.asarray(a, dtype="<i4")  # noqa: E731
    f64 = lambda a: np.asarray(a, dtype="<f8")  # noqa: E731
    ff = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    code = {"SSA": 0, "DIVA": 1, "BPA": 2}[stress]
    if sliding is not None:
        if stress == "BPA":
            raise ValueError("the sliding block is read for SSA and DIVA cases")
        code |= 16
    parts = [i32([m.nV, m.nT, w.nz, ops.ab.nnz, ops.ba.nnz, ops.bb.nnz, code]),
             f64(ff(m.V)), i32(ff(m.Tri + 1)), i32(ff(m.TriC + 1)), i32(m.TriBI), f64(ff(m.TriGC)), f64(w.zeta)]
    for fam, names in ((ops.ab, ["map", "ddx", "ddy"]), (ops.ba, ["map", "ddx", "ddy"]),
                       (ops.bb, ["ddx", "ddy", "d2dx2", "d2dxdy", "d2dy2"])):
        parts += [i32(fam.ptr + 1), i32(fam.ind + 1)] + [f64(fam.vals[k]) for k in names]
    parts.append(i32([ctl.get("stop_rule", 1), ctl.get("visc_it_nit", 300), ctl.get("include_crossterms", 1),
                      ctl.get("precond", 2), ctl.get("warm_start", 0), ctl.get("accel_depth", 0)]))
    parts.append(f64([ctl.get("nonlinear_rtol", 1e-8), ctl.get("vel_max", 5000.0), ctl.get("abstol", 0.0),
                      ctl.get("rtol", 1e-10), ctl.get("accel_mixing", 0.5)]))
    inp = w.inputs()
    parts += [f64(inp["Hi"]), f64(inp["Hs"]), f64(inp["A_flow"]), f64(inp["beta_sq"]), f64(inp["fgr_b"]),
              i32(inp["mask_grounded"])]
    if sliding is not None:
        kind = 1 if "alpha_sq" in sliding else 2 if "phi" in sliding else 0
        cnt = np.diff(m.C_ptr)
        nC_mem = int(cnt.max()) if kind == 2 else 0
        parts.append(i32([sliding["law"], nC_mem, kind]))
        parts.append(f64([sliding.get("m", 3.0), sliding.get("zi_p", 5.0), sliding.get("zi_ut", 200.0)]))
        if kind == 1:
            parts += [f64(sliding["alpha_sq"]), f64(sliding["N"])]
        elif kind == 2:
            C = np.zeros((m.nV, nC_mem), np.int32)
            src = np.repeat(np.arange(m.nV), cnt)
            C[src, np.arange(len(src)) - np.repeat(m.C_ptr[:-1], cnt)] = m.C_ind + 1
            parts += [f64(sliding["N"]), f64(sliding["phi"]), i32(sliding["icefree_land"]), i32(cnt), i32(ff(C))]
    if stress == "BPA":
        # sliding law, borders (u then v: north, east, south, west), rheology and
        # iteration controls, the a->a operators (1-based CSR), zeta_stag
        from .operators import build_aa, zeta_stag
        aa = ctl.get("aa") or build_aa(m)
        zs = ctl.get("zeta_stag")
        zs = zeta_stag(w.zeta) if zs is None else zs
        parts.append(i32([aa.nnz, ctl.get("sliding_law", 1), *ctl.get("bc_u", (0, 0, 0, 0)),
                          *ctl.get("bc_v", (0, 0, 0, 0))]))
        parts.append(f64([ctl.get("glens_n", 3.0), ctl.get("epsilon_sq_0", 1e-8), ctl.get("visc_it_relax", 0.2),
                          ctl.get("visc_it_norm_dUV_tol", 5e-5), ctl.get("ismip_hom_L", 0.0)]))
        parts += [i32(aa.ptr + 1), i32(aa.ind + 1), f64(aa.vals["ddx"]), f64(aa.vals["ddy"]), f64(zs)]
    with open(path, "wb") as f:
        for p in parts:
            f.write(np.ascontiguousarray(p).tobytes())


def write_thickness_case(path, mesh, geom, inp, bc_H=(1, 1, 1, 1)):
    """The reference's mesh arrays of the thickness step, 1-based: nC(nV),
    C(nV, nC_mem), VE(nV, nC_mem) (edge of each connection), ETri(nE, 2)
    (triangles either side of each edge, 0 = none), VBI, A, Cw(nV, nC_mem),
    then Hb, SL, SMB, BMB, LMB, fraction_margin, mask_noice and, in the
    header, nC_mem, nE, BC_H (north, east, south, west: 0 infinite, 1 zero)
    and dt.  Hi and the velocities are those of the velocity case."""
    nV = mesh.nV
    cnt = np.diff(mesh.C_ptr)
    nC_mem = int(cnt.max())
    src = np.repeat(np.arange(nV), cnt)
    dst = mesh.C_ind.astype(np.int64)
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    keys, edge = np.unique(lo * nV + hi, return_inverse=True)
    nE = len(keys)
    ETri = np.zeros((nE, 2), np.int32)
    fwd = src < dst   # the edge's own direction: from its lower vertex
    ETri[edge[fwd]] = geom.ETri[fwd] + 1
    C = np.zeros((nV, nC_mem), np.int32)
    VE = np.zeros((nV, nC_mem), np.int32)
    Cw = np.zeros((nV, nC_mem))
    col = np.arange(len(dst)) - np.repeat(mesh.C_ptr[:-1], cnt)
    C[src, col] = dst + 1
    VE[src, col] = edge + 1
    Cw[src, col] = geom.Cw
    ff = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    i32 = lambda a: np.asarray(a, dtype="<i4")  # noqa: E731
    f64 = lambda a: np.asarray(a, dtype="<f8")  # noqa: E731
    parts = [i32([nC_mem, nE, *bc_H]), f64([inp["dt"]]), i32(cnt), i32(ff(C)), i32(ff(VE)), i32(ff(ETri)),
             i32(mesh.VBI), f64(geom.A), f64(ff(Cw))]
    parts += [f64(inp[k]) for k in ("Hb", "SL", "SMB", "BMB", "LMB", "fraction_margin")]
    parts.append(i32(inp["mask_noice"]))
    with open(path, "wb") as f:
        for p in parts:
            f.write(np.ascontiguousarray(p).tobytes())


def read_thickness_result(path, nV):
    raw = open(path, "rb").read()
    a = np.frombuffer(raw, "<f8", 3 * nV + 1, 0)
    its = int(np.frombuffer(raw, "<i4", 1, 8 * (3 * nV + 1))[0])
    return a[:nV].copy(), a[nV:2 * nV].copy(), a[2 * nV:3 * nV].copy(), float(a[3 * nV]), its


GL_BC = {"grounded": 0, "subgrid": 1, "pmp": 2}


def write_heat_case(path, nV, nz, zeta, fields, SMB, geothermal_flux, dt, Hi_min_thermo=10.0,
                    gl_bc="grounded"):
    """fields: the column arrays of HEAT_COLUMN_FIELDS (nz * nV, vertex
    fastest), Hi_eff, T_surf, Q_base_grnd, mask (1 grounded, 2 floating,
    3 grounded at the grounding line) and optionally fraction_gr; gl_bc is
    choice_GL_temperature_BC for the mask-3 vertices."""
    f64 = lambda a: np.asarray(a, dtype="<f8").reshape(-1)  # noqa: E731
    parts = [np.asarray([nV, nz, GL_BC[gl_bc]], "<i4"), f64([dt, Hi_min_thermo]), f64(zeta)]
    parts += [f64(fields[k]) for k in HEAT_COLUMN_FIELDS]
    parts += [f64(fields["Hi_eff"]), f64(fields["T_surf"]), f64(fields["Q_base_grnd"]), f64(SMB),
              f64(geothermal_flux), np.asarray(fields["mask"], "<i4"),
              f64(fields.get("fraction_gr", np.ones(nV)))]
    for p in parts:
        assert p.size in (2, 3, nz, nV, nz * nV), p.size
    with open(path, "wb") as f:
        for p in parts:
            f.write(np.ascontiguousarray(p).tobytes())


def read_heat_result(path, nV, nz):
    raw = open(path, "rb").read()
    T = np.frombuffer(raw, "<f8", nV * nz, 0)
    n_unstable = int(np.frombuffer(raw, "<i4", 1, 8 * nV * nz)[0])
    return T, n_unstable


def run_driver(case, out, timeout=300, driver=DRIVER, extra=()):
    if not os.path.exists(driver):
        raise RuntimeError(f"{driver} is missing: build it first")
    r = subprocess.run([driver, case, out, *extra], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran driver failed ({r.returncode}): {r.stdout}\n{r.stderr}")
    return r.stdout


def read_result(path, nT):
    raw = open(path, "rb").read()
    u = np.frombuffer(raw, "<f8", nT, 0)
    v = np.frombuffer(raw, "<f8", nT, 8 * nT)
    its = np.frombuffer(raw, "<i4", 2, 16 * nT)
    res = np.frombuffer(raw, "<f8", 1, 16 * nT + 8)[0]
    return u, v, int(its[0]), int(its[1]), float(res)


def write_heat_inputs_case(path, mesh, nz, f):
    """Inputs of the driver's heat-input mode (a third argument): the triangles
    around each vertex as mesh%niTri / mesh%iTri (1-based), Ti (nV, nz),
    u_vav, v_vav (nV), u_3D_b, v_3D_b (nT, nz), the nine strain rates
    (nV, nz, 9), the basal friction coefficient and speed, the grounded mask,
    the a->a operators (1-based CSR), zeta_stag and dHi_dt, dHs_dt.  Fields
    (n, nz) are given as (nz*n,) arrays, the first index fastest."""
    i32 = lambda a: np.asarray(a, dtype="<i4")  # noqa: E731
    f64 = lambda a: np.asarray(a, dtype="<f8")  # noqa: E731
    nV = mesh.nV
    cnt = np.diff(mesh.iTri_ptr)
    mem = int(cnt.max())
    it = np.zeros((nV, mem), np.int32)
    src = np.repeat(np.arange(nV), cnt)
    it[src, np.arange(len(src)) - np.repeat(mesh.iTri_ptr[:-1], cnt)] = mesh.iTri_ind + 1
    aa = f["aa"]
    parts = [i32([mem, aa.nnz]), i32(cnt), i32(np.asfortranarray(it).ravel(order="F")), f64(f["Ti"]),
             f64(f["u_vav"]), f64(f["v_vav"]), f64(f["u_3D_b"]), f64(f["v_3D_b"])]
    parts += [f64(np.concatenate([np.asarray(x, np.float64).ravel() for x in f["strain_rates"]]))]
    parts += [f64(f["beta"]), f64(f["uabs_base"]), i32(f["mask"]), i32(aa.ptr + 1), i32(aa.ind + 1),
              f64(aa.vals["ddx"]), f64(aa.vals["ddy"]), f64(f["zeta_stag"]), f64(f["dHi_dt"]), f64(f["dHs_dt"])]
    with open(path, "wb") as fh:
        for p in parts:
            fh.write(np.ascontiguousarray(p).tobytes())


def read_heat_inputs_result(path, nV, nT, nz):
    """(u_times_dTdxp_upwind, v_times_dTdyp_upwind, Phi, frictional_heating,
    ak[7], bk[6], bks[6]) as written by the driver's heat-input mode."""
    d = np.fromfile(path, dtype="<f8")
    out, o = [], 0
    for n in (nV * nz, nV * nz, nV * nz, nV):
        out.append(d[o:o + n]); o += n
    ak = [d[o + j * nV * nz:o + (j + 1) * nV * nz] for j in range(7)]; o += 7 * nV * nz
    bk = [d[o + j * nT * nz:o + (j + 1) * nT * nz] for j in range(6)]; o += 6 * nT * nz
    bks = [d[o + j * nT * (nz - 1):o + (j + 1) * nT * (nz - 1)] for j in range(6)]
    return (*out, ak, bk, bks)
