"""Pins of the oracle's heat-column inputs (thermodynamics_utilities.f90):

  * calc_upwind_heat_flux_derivatives (:339-434): with a temperature linear
    in x and y in every layer the a->b derivative operators are exact, so
    u*dT/dx = u3(t, k) * b_k on whichever triangle is chosen; with u3 a
    different value on every triangle the output names the chosen triangle,
    and a point a short step upwind of the vertex (along -u_vav) must lie in
    it (a point-in-triangle test, independent of the cross-product cone test
    of the restatement); thin ice (< 1 m) gives zeros;
  * calc_strain_heating (:24-91): simple shear du/dz = s gives D = s/2 and
    pure extension du/dx = -dv/dy = e gives D = e, so Phi = 2 A^(-1/n)
    D^(1 + 1/n) in closed form; no heating under 0.1 m of ice;
  * calc_frictional_heating (:93-128): beta * |u_base| where grounded, zero
    elsewhere and without sliding.
"""

import numpy as np
import pytest

import __graft_entry__ as ge

ge.load_package()
from ufemism2_amd import workloads as W  # noqa: E402
from oracle import binding as OB  # noqa: E402


@pytest.fixture(scope="module")
def chan():
    return W.channel_small(target_nV=1500, stress_balance="DIVA")


def _fields(w, seed):
    rng = np.random.default_rng(seed)
    m, nz = w.mesh, w.nz
    x, y = m.V[:, 0], m.V[:, 1]
    b = rng.normal(0.0, 1e-4, nz)
    c = rng.normal(0.0, 1e-4, nz)
    a = rng.uniform(230.0, 270.0, nz)
    Ti = (a[:, None] + b[:, None] * x[None, :] + c[:, None] * y[None, :]).reshape(-1)
    Hi = rng.uniform(0.0, 3000.0, m.nV)
    Hi[rng.random(m.nV) < 0.1] = 0.5
    # outflow from the domain centre, turned by up to 60 degrees inside the
    # domain: at the border the upwind direction points inwards (an inflowing
    # border vertex has no upwind triangle; the reference stops there)
    xc, yc = 0.5 * (m.xmin + m.xmax), 0.5 * (m.ymin + m.ymax)
    ang = np.arctan2(y - yc, x - xc) + np.where(m.VBI > 0, 0.0, rng.uniform(-np.pi / 3, np.pi / 3, m.nV))
    sp = rng.uniform(1.0, 500.0, m.nV)
    return Ti, Hi, sp * np.cos(ang), sp * np.sin(ang), b, c


def test_upwind_terms_exact_on_linear_temperature(chan):
    w = chan
    m, nz = w.mesh, w.nz
    Ti, Hi, u, v, b, c = _fields(w, 1)
    u3 = np.repeat(np.linspace(50.0, 150.0, nz), m.nT)
    v3 = np.repeat(np.linspace(-20.0, 30.0, nz), m.nT)
    rc, ox, oy = OB.heat_upwind(w.descriptor(), Ti, Hi, u, v, u3, v3)
    assert rc == 0
    ox, oy = ox.reshape(nz, m.nV), oy.reshape(nz, m.nV)
    thick = Hi >= 1.0
    ex = np.linspace(50.0, 150.0, nz)[:, None] * b[:, None]
    ey = np.linspace(-20.0, 30.0, nz)[:, None] * c[:, None]
    assert np.abs(ox[:, thick] - ex).max() < 1e-9 * np.abs(ex).max()
    assert np.abs(oy[:, thick] - ey).max() < 1e-9 * np.abs(ey).max()
    assert (ox[:, ~thick] == 0).all() and (oy[:, ~thick] == 0).all()


def _in_triangle(p, a, b, c):
    def s(p1, p2, p3):
        return (p1[0] - p3[0]) * (p2[1] - p3[1]) - (p2[0] - p3[0]) * (p1[1] - p3[1])
    d1, d2, d3 = s(p, a, b), s(p, b, c), s(p, c, a)
    neg = (d1 < 0) or (d2 < 0) or (d3 < 0)
    pos = (d1 > 0) or (d2 > 0) or (d3 > 0)
    return not (neg and pos)


def test_upwind_triangle_contains_the_upwind_direction(chan):
    w = chan
    m, nz = w.mesh, w.nz
    Ti, Hi, u, v, b, c = _fields(w, 2)
    Hi[:] = 1000.0
    tag = 1.0 + np.arange(m.nT, dtype=np.float64)   # a distinct velocity per triangle
    u3 = np.tile(tag, nz)
    v3 = np.tile(tag, nz)
    rc, ox, oy = OB.heat_upwind(w.descriptor(), Ti, Hi, u, v, u3, v3)
    assert rc == 0
    chosen = np.rint(ox.reshape(nz, m.nV)[0] / b[0]).astype(int) - 1
    h = np.sqrt((m.xmax - m.xmin) * (m.ymax - m.ymin) / m.nV)
    checked = 0
    for vi in range(m.nV):
        t = chosen[vi]
        assert vi in m.Tri[t], vi
        d = np.array([-u[vi], -v[vi]])
        p = m.V[vi] + 1e-3 * h * d / np.linalg.norm(d)
        if not (m.xmin < p[0] < m.xmax and m.ymin < p[1] < m.ymax):
            continue   # the upwind step leaves the domain at border vertices
        assert _in_triangle(p, *m.V[m.Tri[t]]), vi
        checked += 1
    assert checked > 0.8 * m.nV


def test_inflow_at_the_border_has_no_upwind_triangle(chan):
    w = chan
    m = w.mesh
    Ti, Hi, u, v, b, c = _fields(w, 3)
    Hi[:] = 1000.0
    vb = np.flatnonzero(m.VBI == 3)[0]       # an eastern border vertex
    u[vb], v[vb] = -100.0, 0.0                  # flowing in: upwind points out of the domain
    rc, _, _ = OB.heat_upwind(w.descriptor(), Ti, Hi, u, v, np.zeros(m.nT * w.nz), np.zeros(m.nT * w.nz))
    assert rc == -1


def test_strain_heating_closed_forms():
    nV, nz, n, A = 6, 3, 3.0, 1e-16
    Hi = np.array([500.0, 500.0, 500.0, 0.05, 800.0, 800.0])
    z = np.zeros(nV * nz)
    e = [z.copy() for _ in range(9)]
    s, ext = 2e-3, 5e-4
    shear = np.tile([1, 1, 1, 0, 0, 0], nz).astype(bool)
    e[2][shear] = s                    # du/dz
    e[0][~shear] = ext                 # du/dx
    e[4][~shear] = -ext                # dv/dy
    Phi = OB.strain_heating(nV, nz, Hi, np.full(nV * nz, A), e, n).reshape(nz, nV)
    want_shear = 2.0 * A ** (-1 / n) * (s / 2) ** (1 + 1 / n)
    want_ext = 2.0 * A ** (-1 / n) * ext ** (1 + 1 / n)
    assert np.allclose(Phi[:, :3], want_shear, rtol=1e-13)
    assert np.allclose(Phi[:, 4:], want_ext, rtol=1e-13)
    assert (Phi[:, 3] == 0).all()


def test_frictional_heating():
    beta = np.array([1e3, 2e3, 3e3, 4e3])
    ub = np.array([10.0, 0.0, 5.0, 7.0])
    g = np.array([1, 1, 0, 1])
    assert np.array_equal(OB.frictional_heating(beta, ub, g), [1e4, 0.0, 0.0, 2.8e4])
    assert (OB.frictional_heating(beta, ub, g, no_sliding=True) == 0).all()
