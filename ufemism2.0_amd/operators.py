"""Mesh matrix operators and the C mesh descriptor.

The operators are built once per mesh by the library's host code
(ufe_mesh_operators_*, restating mesh_disc_calc_matrix_operators_2D.f90) and
handed to the engine, exactly like the reference's solver reads them from
``mesh%M_*`` (mesh_types.f90:180-192).
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib as L
from .mesh import Mesh


@dataclasses.dataclass
class OperatorFamily:
    ptr: np.ndarray
    ind: np.ndarray
    vals: dict

    @property
    def nnz(self):
        return int(self.ptr[-1])

    def matrix(self, name, ncols):
        import scipy.sparse as sp
        n = self.ptr.size - 1
        return sp.csr_matrix((self.vals[name], self.ind, self.ptr), shape=(n, ncols))


def _family(fn, n_rows, names, *args):
    nnz = C.c_int(0)
    nul_i = C.cast(None, L.ip)
    nul_d = C.cast(None, L.dp)
    rc = fn(*args, C.byref(nnz), nul_i, nul_i, *([nul_d] * len(names)))
    if rc != 0:
        raise RuntimeError(f"operator construction failed ({rc})")
    ptr = np.zeros(n_rows + 1, np.int32)
    ind = np.zeros(nnz.value, np.int32)
    vals = {k: np.zeros(nnz.value) for k in names}
    rc = fn(*args, C.byref(nnz), L.ptr(ptr, L.ip), L.ptr(ind, L.ip), *[L.ptr(vals[k]) for k in names])
    if rc != 0:
        raise RuntimeError(f"operator construction failed ({rc})")
    return OperatorFamily(ptr, ind, vals)


@dataclasses.dataclass
class MeshOperators:
    ab: OperatorFamily
    ba: OperatorFamily
    bb: OperatorFamily


def build_operators(mesh: Mesh) -> MeshOperators:
    lib = L.lib()
    V = L.f64(mesh.V)
    Tri = L.i32(mesh.Tri)
    TriC = L.i32(mesh.TriC)
    TriGC = L.f64(mesh.TriGC)
    ab = _family(lib.ufe_mesh_operators_ab, mesh.nT, ["map", "ddx", "ddy"],
                 mesh.nV, mesh.nT, L.ptr(V), L.ptr(Tri, L.ip), L.ptr(TriGC),
                 L.ptr(L.i32(mesh.C_ptr), L.ip), L.ptr(L.i32(mesh.C_ind), L.ip))
    ba = _family(lib.ufe_mesh_operators_ba, mesh.nV, ["map", "ddx", "ddy"],
                 mesh.nV, mesh.nT, L.ptr(V), L.ptr(TriC, L.ip), L.ptr(TriGC),
                 L.ptr(L.i32(mesh.iTri_ptr), L.ip), L.ptr(L.i32(mesh.iTri_ind), L.ip))
    bb = _family(lib.ufe_mesh_operators_bb2, mesh.nT, ["ddx", "ddy", "d2dx2", "d2dxdy", "d2dy2"],
                 mesh.nT, L.ptr(TriC, L.ip), L.ptr(TriGC))
    return MeshOperators(ab, ba, bb)


def build_aa(mesh: Mesh) -> OperatorFamily:
    """a->a ddx/ddy (mesh%M_ddx_a_a, mesh%M_ddy_a_a;
    mesh_disc_calc_matrix_operators_2D.f90:61-196), read by the zeta gradients
    of the BPA and the heat columns."""
    lib = L.lib()
    V, cp, ci = L.f64(mesh.V), L.i32(mesh.C_ptr), L.i32(mesh.C_ind)
    return _family(lib.ufe_mesh_operators_aa, mesh.nV, ["ddx", "ddy"], mesh.nV, L.ptr(V), L.ptr(cp, L.ip),
                   L.ptr(ci, L.ip))


def zeta_stag(zeta: np.ndarray) -> np.ndarray:
    """Staggered vertical grid of the regular zeta grids: the layer midpoints
    (mesh_zeta.f90:79-80)."""
    zeta = np.asarray(zeta, dtype=np.float64)
    return 0.5 * (zeta[:-1] + zeta[1:])


class MeshDescriptor:
    """Owns the arrays behind a ufe_mesh_desc so the pointers stay valid."""

    def __init__(self, mesh: Mesh, ops: MeshOperators, zeta: np.ndarray):
        self.mesh, self.ops = mesh, ops
        keep = self._keep = []

        def d(a):
            a = L.f64(a); keep.append(a); return L.ptr(a)

        def i(a):
            a = L.i32(a); keep.append(a); return L.ptr(a, L.ip)

        s = L.MeshDesc()
        s.nV, s.nT, s.nz = mesh.nV, mesh.nT, int(len(zeta))
        s.V, s.Tri, s.TriC, s.TriBI = d(mesh.V), i(mesh.Tri), i(mesh.TriC), i(mesh.TriBI)
        s.TriGC, s.zeta = d(mesh.TriGC), d(zeta)
        s.ab_nnz, s.ab_ptr, s.ab_ind = ops.ab.nnz, i(ops.ab.ptr), i(ops.ab.ind)
        s.ab_map, s.ab_ddx, s.ab_ddy = d(ops.ab.vals["map"]), d(ops.ab.vals["ddx"]), d(ops.ab.vals["ddy"])
        s.ba_nnz, s.ba_ptr, s.ba_ind = ops.ba.nnz, i(ops.ba.ptr), i(ops.ba.ind)
        s.ba_map, s.ba_ddx, s.ba_ddy = d(ops.ba.vals["map"]), d(ops.ba.vals["ddx"]), d(ops.ba.vals["ddy"])
        s.bb_nnz, s.bb_ptr, s.bb_ind = ops.bb.nnz, i(ops.bb.ptr), i(ops.bb.ind)
        s.bb_ddx, s.bb_ddy = d(ops.bb.vals["ddx"]), d(ops.bb.vals["ddy"])
        s.bb_d2dx2, s.bb_d2dxdy, s.bb_d2dy2 = (d(ops.bb.vals["d2dx2"]), d(ops.bb.vals["d2dxdy"]),
                                               d(ops.bb.vals["d2dy2"]))
        self.desc = s
        self.zeta = L.f64(zeta)

    def byref(self):
        return C.byref(self.desc)
