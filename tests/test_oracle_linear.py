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
