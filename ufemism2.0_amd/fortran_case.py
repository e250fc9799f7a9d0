"""Case files for the Fortran driver (ufe_fortran_driver): the mesh in the
reference's 1-based layout (V(nV,2), Tri/TriC(nT,3) with TriC = 0 for none,
TriGC(nT,2)), the operators as 1-based CSR, a few solver controls and the
inputs, written as one raw little-endian stream in Fortran array order."""

from __future__ import annotations

import os
import subprocess

import numpy as np

DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "ufe_fortran_driver")


def write_case(path, w, stress="SSA", **ctl):
    m, ops = w.mesh, w.ops
    i32 = lambda a: np.asarray(a, dtype="<i4")  # noqa: E731
    f64 = lambda a: np.asarray(a, dtype="<f8")  # noqa: E731
    ff = lambda a: np.asfortranarray(a).ravel(order="F")  # noqa: E731
    parts = [i32([m.nV, m.nT, w.nz, ops.ab.nnz, ops.ba.nnz, ops.bb.nnz, 1 if stress == "DIVA" else 0]),
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
    with open(path, "wb") as f:
        for p in parts:
            f.write(np.ascontiguousarray(p).tobytes())


def run_driver(case, out, timeout=300):
    if not os.path.exists(DRIVER):
        raise RuntimeError(f"{DRIVER} is missing: build it first")
    r = subprocess.run([DRIVER, case, out], capture_output=True, text=True, timeout=timeout)
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
