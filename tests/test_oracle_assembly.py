"""Oracle assembly pinned by closed forms: manufactured quadratic velocities
with n = 1 and m = 1 (constant N and beta), the Glen viscosity and Weertman
friction at linear velocity fields (exact strain rates), and the boundary
rows (solve_linearised_SSA_DIVA.f90:94-147, 480-640)."""

import numpy as np
import pytest

import manufactured as MF
from conftest import params


@pytest.fixture(scope="module")
def case(W):
    return MF.linear_case(W)


def free_rows(w):
    return np.nonzero(w.mesh.TriBI == 0)[0]


@pytest.mark.parametrize("cross", [1, 0])
def test_free_rows_manufactured_quadratic(L, OB, case, cross):
    w = case
    p = params(L, glens_n=1.0, sliding_law=L.SLIDING["Weertman"], slid_weertman_m=1.0, include_crossterms=cross)
    u, v, ax, ay, bx, by = MF.expected_rows(w, 1e-16, 1000.0, 1e3, crossterms=bool(cross))
    A, b = OB.Oracle(w.descriptor()).system(p, w.inputs(), u, v)
    x = np.empty(2 * w.mesh.nT); x[0::2] = u; x[1::2] = v
    Ax = A @ x
    fr = free_rows(w)
    sc = np.abs(ax[fr]).max()
    assert np.abs(Ax[2 * fr] - ax[fr]).max() < 1e-8 * sc
    assert np.abs(Ax[2 * fr + 1] - ay[fr]).max() < 1e-8 * sc
    assert np.allclose(b[2 * fr], bx[fr], rtol=1e-10)
    assert np.allclose(b[2 * fr + 1], by[fr], rtol=1e-10)


def test_glen_viscosity_closed_form(L, OB, case):
    """Linear u, v: strain rates exact, eta = 1/2 A^(-1/n) (e^2 + eps0)^((1-n)/2n)
    (constitutive_equation.f90:25-51); sans rows carry b = -tau_d / N."""
    w = case
    n, A, H, eps0 = 3.0, 1e-16, 1000.0, 1e-8
    p = params(L, glens_n=n, epsilon_sq_0=eps0, include_crossterms=0)
    xg, yg = w.mesh.TriGC[:, 0], w.mesh.TriGC[:, 1]
    ux, uy, vx, vy = 2e-3, -1e-3, 5e-4, -1.5e-3
    u, v = ux * xg + uy * yg, vx * xg + vy * yg
    e2 = ux * ux + vy * vy + ux * vy + 0.25 * (uy + vx) ** 2 + eps0
    eta = 0.5 * A ** (-1 / n) * e2 ** ((1 - n) / (2 * n))
    N = eta * H
    _, b = OB.Oracle(w.descriptor()).system(p, w.inputs(), u, v)
    *_, bx, by = MF.expected_rows(w, A, H, 0.0, crossterms=True)
    fr = free_rows(w)
    assert np.allclose(b[2 * fr], bx[fr] / N, rtol=1e-9)
    assert np.allclose(b[2 * fr + 1], by[fr] / N, rtol=1e-9)


@pytest.mark.parametrize("law,expo", [("Weertman", 3.0), ("Budd", 0.3), ("Coulomb", None)])
def test_sliding_closed_form(L, OB, case, law, expo):
    """Linear u, v, constant N: the free crossterm rows reduce to -beta_b u,
    beta_b the vertex mean of beta_a(u_a) (sliding_laws.f90:25-130, 520-640)."""
    w = case
    n, A, H, eps0, dv = 3.0, 1e-16, 1000.0, 1e-8, 1e-3
    kw = dict(glens_n=n, epsilon_sq_0=eps0, sliding_law=L.SLIDING[law], slid_delta_v=dv)
    if law == "Weertman":
        kw["slid_weertman_m"] = expo
    if law == "Budd":
        kw["slid_budd_q_plastic"] = expo
    p = params(L, **kw)
    xg, yg = w.mesh.TriGC[:, 0], w.mesh.TriGC[:, 1]
    ux, uy, vx, vy = 2e-4, -1e-4, 5e-5, -1.5e-4
    u, v = ux * xg + uy * yg + 50.0, vx * xg + vy * yg - 20.0
    A_, _ = OB.Oracle(w.descriptor()).system(p, w.inputs(), u, v)
    x = np.empty(2 * w.mesh.nT); x[0::2] = u; x[1::2] = v
    Ax = A_ @ x
    V = w.mesh.V
    ua, va = ux * V[:, 0] + uy * V[:, 1] + 50.0, vx * V[:, 0] + vy * V[:, 1] - 20.0
    uabs = np.sqrt(dv * dv + ua * ua + va * va)
    rough = 1e3
    if law == "Weertman":
        ba = rough * uabs ** (1 / expo - 1)
    elif law == "Budd":
        ba = rough * uabs ** (expo - 1) / 100.0 ** expo
    else:
        ba = rough / uabs
    bb = ba[w.mesh.Tri].mean(axis=1)
    fr = free_rows(w)
    sc = np.abs(bb[fr] * u[fr]).max()
    assert np.abs(Ax[2 * fr] + bb[fr] * u[fr]).max() < 1e-7 * sc
    assert np.abs(Ax[2 * fr + 1] + bb[fr] * v[fr]).max() < 1e-7 * sc


def test_boundary_rows(L, OB, case):
    w = case
    mesh = w.mesh
    # north/south u zero, east/west u infinite, v: zero everywhere
    p = params(L, bc_u=[1, 0, 1, 0], bc_v=[1, 1, 1, 1])
    rng = np.random.default_rng(0)
    u, v = rng.normal(size=mesh.nT), rng.normal(size=mesh.nT)
    mask = np.zeros(mesh.nT, np.int32); mask[5] = 1 if mesh.TriBI[5] == 0 else 0
    inp = dict(w.inputs())
    A, b = OB.Oracle(w.descriptor()).system(p, inp, u, v)
    A = A.toarray() if mesh.nT < 3000 else A
    for ti in np.nonzero(mesh.TriBI > 0)[0][:200]:
        side = (mesh.TriBI[ti] - 1) // 2
        for uv in (0, 1):
            row = A[2 * ti + uv]
            zero = (uv == 1) or side in (0, 2)
            if zero:
                assert row[2 * ti + uv] == 1.0 and np.count_nonzero(row) == 1 and b[2 * ti + uv] == 0.0
            else:
                nb = [t for t in mesh.TriC[ti] if t >= 0]
                assert row[2 * ti + uv] == -len(nb)
                assert all(row[2 * t + uv] == 1.0 for t in nb)
                assert np.count_nonzero(row) == len(nb) + 1 and b[2 * ti + uv] == 0.0
