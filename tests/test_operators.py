"""Mesh operators of the library's host builder (ufe_mesh_operators_*),
pinned by the invariants of the shape functions and by polynomial exactness
(mesh_disc_calc_matrix_operators_2D.f90, shape_functions.f90)."""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def mo(W):
    w = W.dome(h=60e3)
    return w.mesh, w.ops


def test_rows_annihilate_constants(mo):
    mesh, ops = mo
    one_v, one_t = np.ones(mesh.nV), np.ones(mesh.nT)
    for fam, ncol, names, f in ((ops.ab, mesh.nV, ["ddx", "ddy"], one_v), (ops.ba, mesh.nT, ["ddx", "ddy"], one_t),
                                (ops.bb, mesh.nT, ["ddx", "ddy", "d2dx2", "d2dxdy", "d2dy2"], one_t)):
        for k in names:
            assert np.abs(fam.matrix(k, ncol) @ f).max() < 1e-15 * 1e3
    assert np.allclose(ops.ab.matrix("map", mesh.nV) @ one_v, 1.0, atol=1e-13)
    assert np.allclose(ops.ba.matrix("map", mesh.nT) @ one_t, 1.0, atol=1e-13)


def test_exact_on_linear_functions(mo):
    mesh, ops = mo
    a, b, c = 2.5, -1.25, 7.0
    fv = a * mesh.V[:, 0] / 1e3 + b * mesh.V[:, 1] / 1e3 + c
    ft = a * mesh.TriGC[:, 0] / 1e3 + b * mesh.TriGC[:, 1] / 1e3 + c
    # a -> b
    assert np.allclose(ops.ab.matrix("map", mesh.nV) @ fv, ft, rtol=1e-11)
    assert np.allclose(ops.ab.matrix("ddx", mesh.nV) @ fv, a / 1e3, rtol=1e-9)
    assert np.allclose(ops.ab.matrix("ddy", mesh.nV) @ fv, b / 1e3, rtol=1e-9)
    # b -> a
    assert np.allclose(ops.ba.matrix("map", mesh.nT) @ ft, fv, rtol=1e-10)
    assert np.allclose(ops.ba.matrix("ddx", mesh.nT) @ ft, a / 1e3, rtol=1e-9)
    assert np.allclose(ops.ba.matrix("ddy", mesh.nT) @ ft, b / 1e3, rtol=1e-9)
    # b -> b second order: first derivatives exact, second derivatives zero
    for k, want in (("ddx", a / 1e3), ("ddy", b / 1e3)):
        assert np.allclose(ops.bb.matrix(k, mesh.nT) @ ft, want, rtol=1e-8)
    for k in ("d2dx2", "d2dxdy", "d2dy2"):
        assert np.abs(ops.bb.matrix(k, mesh.nT) @ ft).max() < 1e-14


def test_second_order_exact_on_quadratics(mo):
    """M2 b->b rows fit a full quadratic: exact on quadratics wherever the
    stencil has >= 5 neighbours (all rows, by construction)."""
    mesh, ops = mo
    x, y = mesh.TriGC[:, 0] / 1e5, mesh.TriGC[:, 1] / 1e5
    f = 0.3 * x * x - 1.1 * x * y + 0.7 * y * y + 2 * x - y + 4
    s = 1e-5
    want = {"ddx": (0.6 * x - 1.1 * y + 2) * s, "ddy": (-1.1 * x + 1.4 * y - 1) * s,
            "d2dx2": 0.6 * s * s + 0 * x, "d2dxdy": -1.1 * s * s + 0 * x, "d2dy2": 1.4 * s * s + 0 * x}
    for k, w in want.items():
        got = ops.bb.matrix(k, mesh.nT) @ f
        assert np.abs(got - w).max() <= 1e-7 * np.abs(w).max(), k


def test_bb_stencil_lstsq_brute_force(mo):
    """Independent weighted least-squares fit with numpy's lstsq for a few
    rows: the 5 Taylor coefficients (f_x, f_y, f_xx, f_xy, f_yy) are the rows
    of the weighted pseudo-inverse, with weights 1/dist^1.5
    (calc_shape_functions_2D_reg_2nd_order, shape_functions.f90:218-364,
    q = 1.5 at :13), and the centre weight is minus their sum."""
    mesh, ops = mo
    bb = ops.bb
    for ti in (0, mesh.nT // 3, mesh.nT // 2, mesh.nT - 1):
        cols = bb.ind[bb.ptr[ti]:bb.ptr[ti + 1]]
        nb = cols[cols != ti]
        p0 = mesh.TriGC[ti]
        d = mesh.TriGC[nb] - p0
        dist = np.hypot(d[:, 0], d[:, 1])
        w = 1.0 / dist ** 1.5
        M = np.stack([d[:, 0], d[:, 1], d[:, 0] ** 2 / 2, d[:, 0] * d[:, 1], d[:, 1] ** 2 / 2], axis=1)
        Wm = M * w[:, None]
        P = np.linalg.lstsq(Wm, np.diag(w), rcond=None)[0]  # 5 x nn
        for row, k in enumerate(["ddx", "ddy", "d2dx2", "d2dxdy", "d2dy2"]):
            vals = dict(zip(bb.ind[bb.ptr[ti]:bb.ptr[ti + 1]], bb.vals[k][bb.ptr[ti]:bb.ptr[ti + 1]]))
            got = np.array([vals[c] for c in nb])
            assert np.allclose(got, P[row], rtol=1e-8, atol=1e-12 * np.abs(P[row]).max()), (ti, k)
            assert abs(vals[ti] + P[row].sum()) <= 1e-9 * np.abs(P[row]).max()


def test_mesh_topology(mo):
    mesh, _ = mo
    # CCW triangles with positive area
    P = mesh.V[mesh.Tri]
    area = ((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) -
            (P[:, 1, 1] - P[:, 0, 1]) * (P[:, 2, 0] - P[:, 0, 0])) / 2
    assert (area > 0).all()
    assert np.isclose(area.sum(), (mesh.xmax - mesh.xmin) * (mesh.ymax - mesh.ymin), rtol=1e-12)
    # TriC symmetric, -1 = border
    for ti in range(mesh.nT):
        for tj in mesh.TriC[ti]:
            if tj >= 0:
                assert ti in mesh.TriC[tj]
    # Euler: nT = 2 nV - nB - 2
    nB = int((mesh.VBI > 0).sum())
    assert mesh.nT == 2 * mesh.nV - nB - 2
