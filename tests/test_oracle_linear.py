"""Oracle linear algebra pinned to the reference's known answers and to
library routines (numpy/scipy)."""

import json
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))


def petsc_6x6():
    c = GOLD["ut_petsc_6x6"]
    e = np.array(c["entries"])
    A = sp.csr_matrix((e[:, 2], (e[:, 0].astype(int) - 1, e[:, 1].astype(int) - 1)), shape=(6, 6))
    return A, np.array(c["x"]), np.array(c["y"])


def test_csr_matvec_ut_petsc_known_answer(OB):
    A, x, y = petsc_6x6()
    assert np.array_equal(OB.csr_matvec(A.indptr, A.indices, A.data, x), y)


def tridiag_case(ni):
    n = 1 + 2 ** ni
    dx = 2.0 / (n - 1)
    a = np.full(n, 1.0 / dx ** 2); b = np.full(n, -2.0 / dx ** 2); c = np.full(n, 1.0 / dx ** 2)
    d = np.full(n, 2.0)
    b[0], c[0], d[0] = 1.0, 0.0, 1.0
    a[-1], b[-1], d[-1] = 0.0, 1.0, 1.0
    return n, a, b, c, d


@pytest.mark.parametrize("ni", GOLD["tridiag_known"]["ni"])
def test_tridiag_known_answer(OB, ni):
    n, a, b, c, d = tridiag_case(ni)
    x = OB.tridiag(a, b, c, d)
    tol = GOLD["tridiag_known"]["tol"]
    assert abs(x[0] - 1) < tol and abs(x[n // 2]) < tol and abs(x[-1] - 1) < tol


def test_tridiag_vs_dense(OB):
    rng = np.random.default_rng(0)
    n = 40
    a, c = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    b = 3 + rng.uniform(0, 1, n)
    d = rng.normal(size=n)
    M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    assert np.allclose(OB.tridiag(a, b, c, d), np.linalg.solve(M, d), rtol=1e-12, atol=1e-13)


def test_gmres_ilu0_vs_direct_random(OB):
    rng = np.random.default_rng(1)
    n = 300
    A = sp.random(n, n, density=0.02, random_state=2, format="csr") + sp.eye(n) * 4.0
    A = A.tocsr(); A.sort_indices()
    b = rng.normal(size=n)
    x, its = OB.gmres_ilu0(A, b, rtol=1e-12, maxit=2000)
    assert its > 0
    assert np.allclose(x, spl.spsolve(A.tocsc(), b), rtol=1e-9, atol=1e-10)


def test_gmres_ilu0_vs_direct_ssa_system(OB, L, dome):
    """The oracle's Krylov solve on an assembled SSA system against scipy's LU."""
    from conftest import params
    md = dome.descriptor(); inp = dome.inputs()
    o = OB.Oracle(md)
    rng = np.random.default_rng(3)
    u = rng.normal(0, 50, dome.mesh.nT); v = rng.normal(0, 50, dome.mesh.nT)
    A, b = o.system(params(L), inp, u, v)
    A.sort_indices()
    x, its = OB.gmres_ilu0(A, b, rtol=1e-13, maxit=20000)
    xd = spl.spsolve(A.tocsc(), b)
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-8


def laplace_case(W, h=25e3):
    """The reference's Laplace known answer (ct_discretisation_solve_Laplace_eq.f90:142-207):
    d2f/dx2 + d2f/dy2 = c for r < r0 and f = f_ex for r >= r0 on the triangle
    centres, with f_ex = -c/4 r0^2 + c/4 r^2, c = -1e-9, r0 = 0.8 xmax, the
    rows assembled from the M2 d2dx2 + d2dy2 operators.  The M2 operators are
    exact on quadratics in the interior, so f_ex solves the discrete system."""
    mesh = W.uniform_mesh(-500e3, 500e3, -500e3, 500e3, h, seed=9)
    ops = W.build_operators(mesh)
    c, r0 = -1e-9, mesh.xmax * 0.8
    x, y = mesh.TriGC[:, 0], mesh.TriGC[:, 1]
    f_ex = -c / 4 * r0 ** 2 + c / 4 * (x ** 2 + y ** 2)
    lap = ops.bb.matrix("d2dx2", mesh.nT) + ops.bb.matrix("d2dy2", mesh.nT)
    inside = np.hypot(x, y) < r0
    A = sp.diags((~inside).astype(float)) + sp.diags(inside.astype(float)) @ lap
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    A.sort_indices()
    b = np.where(inside, c, f_ex)
    return A, b, f_ex


def test_laplace_known_answer_oracle(OB, W):
    A, b, f_ex = laplace_case(W)
    # f_ex solves the discrete system (interior rows to round-off of c = -1e-9)
    assert np.abs(A @ f_ex - b).max() < 1e-10 * 1e-9
    x, its = OB.gmres_ilu0(A, b, rtol=1e-12, maxit=20000)
    assert its > 0
    assert np.abs(x - f_ex).max() < 1e-9 * np.abs(f_ex).max()
