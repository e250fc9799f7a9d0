"""Pins of the oracle's BPA path (BPA_main.f90:90-2084) and what it runs on:
calc_zeta_gradients (zeta_gradients.f90:19-212), the 1-D zeta operators
(mesh_zeta.f90:287-482), the 3-D operators
(mesh_disc_calc_matrix_operators_3D.f90) and the a->a operators
(mesh_disc_calc_matrix_operators_2D.f90:61-196).

What fixes the results, independent of the code under test:
  * the 3-D operators are exact on fields that are quadratic in (x, y) times
    quadratic in zeta (second-order operators) or linear times quadratic
    (first order): they are compared with derivatives of the same field taken
    in Cartesian coordinates (x, y, z) by central finite differences, which is
    not the chain rule the operators are built from;
  * the zeta gradients against finite differences of zeta = (Hs - z)/Hi; the
    reference's second derivatives keep one of the two d(1/Hi)/dx terms
    (zeta_gradients.f90:169-171), so with a sloping Hi they equal half the
    true value for a linear geometry, and the truth where Hi is uniform;
  * a Newtonian slab on a uniform slope (n = 1, linear sliding, 'infinite'
    borders) has the BPA solution u(zeta) = u_b - U'' H^2 (1 - zeta^2) / 2 with
    M [U'', V''] = (rho g / eta) grad(s), M = [[1 + 4 sx^2 + sy^2, 3 sx sy],
    [3 sx sy, 1 + 4 sy^2 + sx^2]] and u_b = -rho g H grad(s) / beta (derived
    from the BPA and its surface and base conditions); the discretisation is
    exact on it, for any zeta spacing, so the oracle must return it to the
    solver tolerance (the 4 sx^2 terms come only from the chain-rule parts of
    the operators and the boundary rows);
  * with the flow factor varying with depth, A = A0 + A1 zeta, the slab has
    u(zeta) = u_b - rho g sx H^2 / (1 + 4 sx^2) (A0 (1 - zeta^2) + 2/3 A1
    (1 - zeta^3)); the viscosity-gradient terms make the discrete solution
    second-order accurate in the layer spacing (a sign error in them gives an
    O(1) error);
  * no sliding gives u = 0 at the base; prescribed rows return their values;
    no grounded ice gives zero velocities.
"""

import numpy as np
import pytest

import __graft_entry__ as ge

ge.load_package()
from ufemism2_amd import _lib as L, mesh as M  # noqa: E402
from ufemism2_amd.operators import MeshDescriptor, build_aa, build_operators, zeta_stag  # noqa: E402
from oracle import binding as OB  # noqa: E402

RHO_G = 910.0 * 9.81
LX = 120e3


class Case:
    def __init__(self, h=8e3, zeta=None, seed=3):
        self.mesh = M.uniform_mesh(0.0, LX, 0.0, LX, h, seed=seed)
        self.ops = build_operators(self.mesh)
        self.aa = build_aa(self.mesh)
        self.zeta = np.linspace(0.0, 1.0, 7) if zeta is None else np.asarray(zeta, dtype=np.float64)
        self.zs = zeta_stag(self.zeta)
        self.md = MeshDescriptor(self.mesh, self.ops, self.zeta)

    @property
    def nz(self):
        return len(self.zeta)


@pytest.fixture(scope="module")
def case():
    return Case(zeta=np.array([0.0, 0.08, 0.2, 0.37, 0.6, 0.82, 1.0]))


def _geometry(c, Hx=0.0, H0=1000.0, sx=-0.02, sy=0.01):
    x, y = c.mesh.V[:, 0], c.mesh.V[:, 1]
    Hi = H0 + Hx * (x - LX / 2)
    Hs = 1500.0 + sx * x + sy * y
    return Hi, Hs


def _fd_cart(F, x, y, z, h):
    """Cartesian derivatives of F(x, y, z) by 5-point central differences:
    x, y, z, xx, xy, yy, zz (exact for polynomials up to degree 4)."""
    def d1(g, p, hh):
        return (-g(p + 2 * hh) + 8 * g(p + hh) - 8 * g(p - hh) + g(p - 2 * hh)) / (12 * hh)

    def d2(g, p, hh):
        return (-g(p + 2 * hh) + 16 * g(p + hh) - 30 * g(p) + 16 * g(p - hh) - g(p - 2 * hh)) / (12 * hh * hh)

    fx = d1(lambda t: F(t, y, z), x, h[0])
    fy = d1(lambda t: F(x, t, z), y, h[1])
    fz = d1(lambda t: F(x, y, t), z, h[2])
    fxx = d2(lambda t: F(t, y, z), x, h[0])
    fyy = d2(lambda t: F(x, t, z), y, h[1])
    fzz = d2(lambda t: F(x, y, t), z, h[2])
    fxy = d1(lambda t: d1(lambda s: F(s, t, z), x, h[0]), y, h[1])
    return [fx, fy, fz, fxx, fxy, fyy, fzz]


def _cart(c, Hfun, Sfun, f):
    """F(x, y, z) = f(x, y, zeta(x, y, z)) with zeta = (Hs - z)/Hi."""
    return lambda x, y, z: f(x, y, (Sfun(x, y) - z) / Hfun(x, y))


def _pts_bk(c, Hi, Hs, zlev):
    """Cartesian points of the bk grid: triangle centres, z = Hs_b - zeta Hi_b."""
    m = c.mesh
    A = c.ops.ab.matrix("map", m.nV)
    Hb, Sb = A @ Hi, A @ Hs
    x = np.tile(m.TriGC[:, 0], len(zlev))
    y = np.tile(m.TriGC[:, 1], len(zlev))
    zz = np.repeat(zlev, m.nT)
    return x, y, np.tile(Sb, len(zlev)) - zz * np.tile(Hb, len(zlev))


def test_aa_operators_exact_on_linear_fields(case):
    m = case.mesh
    x, y = m.V[:, 0], m.V[:, 1]
    f = 3.0 + 2e-3 * x - 5e-4 * y
    Dx = case.aa.matrix("ddx", m.nV)
    Dy = case.aa.matrix("ddy", m.nV)
    assert np.abs(Dx @ f - 2e-3).max() < 1e-12
    assert np.abs(Dy @ f + 5e-4).max() < 1e-12
    assert np.array_equal(case.aa.ind[case.aa.ptr[:-1]], np.arange(m.nV))   # home vertex first


def test_zeta_gradients_first_derivatives(case):
    """Sloping Hi and Hs (linear): zeta_x, zeta_y, zeta_z and zeta_t against
    finite differences of zeta(x, y, z, t)."""
    c = case
    m = c.mesh
    Hx, sx, sy, H0 = 2e-3, -0.02, 0.01, 1000.0
    Hi, Hs = _geometry(c, Hx=Hx, H0=H0, sx=sx, sy=sy)
    dHi_dt = np.full(m.nV, 0.3)
    dHs_dt = np.full(m.nV, -0.1)
    ak, bk, bks = OB.zeta_gradients(c.md, c.aa, c.zs, Hi, Hs, dHi_dt, dHs_dt)
    Hf = lambda x, y, t=0.0: H0 + Hx * (x - LX / 2) + 0.3 * t                          # noqa: E731
    Sf = lambda x, y, t=0.0: 1500.0 + sx * x + sy * y - 0.1 * t                          # noqa: E731
    for grid, pts, out, levels in (("ak", (np.tile(m.V[:, 0], c.nz), np.tile(m.V[:, 1], c.nz)), ak, c.zeta),
                                   ("bk", (np.tile(m.TriGC[:, 0], c.nz), np.tile(m.TriGC[:, 1], c.nz)), bk,
                                    c.zeta),
                                   ("bks", (np.tile(m.TriGC[:, 0], c.nz - 1), np.tile(m.TriGC[:, 1], c.nz - 1)),
                                    bks, c.zs)):
        x, y = pts
        zl = np.repeat(levels, x.size // len(levels))
        z = Sf(x, y) - zl * Hf(x, y)
        zeta = lambda x_, y_, z_, t_=0.0: (Sf(x_, y_, t_) - z_) / Hf(x_, y_, t_)     # noqa: E731
        h = 1.0
        zx = (zeta(x + h, y, z) - zeta(x - h, y, z)) / (2 * h)
        zy = (zeta(x, y + h, z) - zeta(x, y - h, z)) / (2 * h)
        zz = (zeta(x, y, z + h) - zeta(x, y, z - h)) / (2 * h)
        off = 1 if grid == "ak" else 0
        assert np.abs(out[off + 0] - zx).max() < 1e-9 * np.abs(zx).max(), grid
        assert np.abs(out[off + 1] - zy).max() < 1e-9 * np.abs(zy).max(), grid
        assert np.abs(out[off + 2] - zz).max() < 1e-9 * np.abs(zz).max(), grid
        if grid == "ak":
            zt = (zeta(x, y, z, 1e-3) - zeta(x, y, z, -1e-3)) / 2e-3
            assert np.abs(out[0] - zt).max() < 1e-7 * np.abs(zt).max()


def test_zeta_gradients_second_derivatives(case):
    """Against finite differences of zeta(x, y, z) for a planar surface and
    planar Hi.  The reference's expressions (zeta_gradients.f90:169-171)
    keep, of the two terms d(1/Hi)/dx_i * (dHs/dx_j - zeta dHi/dx_j) of the
    true second derivative, only the one with dHi/dx (xx), dHi/dy (xy) and
    dHi/dy (yy): with Hi sloping in x they give half the true xx, zero xy
    (the true xy is not zero) and the true yy; with Hi sloping in y, the true
    xx and xy and half the true yy; with uniform Hi the true (zero) values."""
    c = case
    m = c.mesh
    for Hx, Hy, fac in ((0.0, 0.0, (1, 1, 1)), (2e-3, 0.0, (0.5, 0.0, 1)), (0.0, 2e-3, (1, 1, 0.5))):
        x0, y0 = m.V[:, 0], m.V[:, 1]
        Hi = 1000.0 + Hx * (x0 - LX / 2) + Hy * (y0 - LX / 2)
        Hs = 1500.0 - 0.02 * x0 + 0.01 * y0
        ak, bk, bks = OB.zeta_gradients(c.md, c.aa, c.zs, Hi, Hs)
        Hf = lambda x_, y_: 1000.0 + Hx * (x_ - LX / 2) + Hy * (y_ - LX / 2)        # noqa: E731
        Sf = lambda x_, y_: 1500.0 - 0.02 * x_ + 0.01 * y_                          # noqa: E731
        for grid, out, x, y, levels in (("ak", ak[4:], x0, y0, c.zeta),
                                        ("bk", bk[3:], m.TriGC[:, 0], m.TriGC[:, 1], c.zeta),
                                        ("bks", bks[3:], m.TriGC[:, 0], m.TriGC[:, 1], c.zs)):
            n = len(levels)
            xx, yy = np.tile(x, n), np.tile(y, n)
            zl = np.repeat(levels, x.size)
            z = Sf(xx, yy) - zl * Hf(xx, yy)
            zeta = lambda x_, y_: (Sf(x_, y_) - z) / Hf(x_, y_)                       # noqa: E731
            h = 50.0
            zxx = (zeta(xx + h, yy) - 2 * zeta(xx, yy) + zeta(xx - h, yy)) / h ** 2
            zyy = (zeta(xx, yy + h) - 2 * zeta(xx, yy) + zeta(xx, yy - h)) / h ** 2
            zxy = (zeta(xx + h, yy + h) - zeta(xx + h, yy - h) - zeta(xx - h, yy + h) + zeta(xx - h, yy - h)) / (
                4 * h * h)
            scale = 1e-10   # 2 zeta_x dHi/dx / Hi for these slopes
            for got, want, f, name in zip(out, (zxx, zxy, zyy), fac, ("xx", "xy", "yy")):
                assert np.abs(got - f * want).max() < 1e-5 * scale, (grid, Hx, Hy, name)
                if f != 1:
                    assert np.abs(want).max() > 0.1 * scale, (grid, Hx, Hy, name)


def _poly2(x, y, t):
    """Sum of quadratics in (x, y) times quadratics in zeta."""
    X, Y = x / LX, y / LX
    return ((1.0 + 0.3 * t - 0.5 * t * t) * (2.0 + X + 0.5 * Y + 0.7 * X * X - 0.2 * X * Y + 0.4 * Y * Y)
            + t * t * (X - Y * Y))


def _poly1(x, y, t):
    """Linear in (x, y) times quadratic in zeta."""
    X, Y = x / LX, y / LX
    return (1.0 + 0.3 * t - 0.5 * t * t) * (2.0 + X - 0.5 * Y) + t * t * (0.5 * X + Y)


def _poly11(x, y, t):
    """Linear in (x, y) times linear in zeta."""
    X, Y = x / LX, y / LX
    return (1.0 + 0.6 * t) * (2.0 + X - 0.5 * Y) + t * (0.5 * X + Y)


def test_bk_bk_operators_exact(case):
    """The seven bk->bk operators of the BPA stiffness matrix against
    Cartesian derivatives, uniform Hi and a sloping surface (the zeta
    gradients are then exact)."""
    c = case
    m = c.mesh
    Hi, Hs = _geometry(c, Hx=0.0)
    x, y, z = _pts_bk(c, Hi, Hs, c.zeta)
    f_bk = _poly2(x, y, np.repeat(c.zeta, m.nT))
    out = OB.ops3d_apply(c.md, c.aa, c.zs, Hi, Hs, f_bk, np.zeros(m.nV * c.nz), np.zeros(m.nT * (c.nz - 1)))
    Hf = lambda x_, y_: 1000.0 + 0.0 * x_                                              # noqa: E731
    Sf = lambda x_, y_: 1500.0 - 0.02 * x_ + 0.01 * y_                                  # noqa: E731
    want = _fd_cart(_cart(c, Hf, Sf, _poly2), x, y, z, (200.0, 200.0, 2.0))
    names = ["ddx", "ddy", "ddz", "d2dx2", "d2dxdy", "d2dy2", "d2dz2"]
    for i, n in enumerate(names):
        scale = np.abs(want[i]).max()
        assert np.abs(out["bk_bk"][i] - want[i]).max() < 1e-7 * scale, n


def test_first_order_3d_operators_exact(case):
    """bk->ak and ak->bk d/dx, d/dy (sloping Hi and surface), bk->bks d/dz,
    and the bks->bk, bks->ak, ak->bks maps and bks->bk d/dz, on the fields
    they are exact for."""
    c = case
    m = c.mesh
    Hx = 2e-3
    Hi, Hs = _geometry(c, Hx=Hx)
    Hf = lambda x_, y_: 1000.0 + Hx * (x_ - LX / 2)                                  # noqa: E731
    Sf = lambda x_, y_: 1500.0 - 0.02 * x_ + 0.01 * y_                                  # noqa: E731
    nz, nT, nV = c.nz, m.nT, m.nV
    zb = np.repeat(c.zeta, nT)
    zsb = np.repeat(c.zs, nT)
    xb, yb = np.tile(m.TriGC[:, 0], nz), np.tile(m.TriGC[:, 1], nz)
    xa, ya = np.tile(m.V[:, 0], nz), np.tile(m.V[:, 1], nz)
    xs, ys = np.tile(m.TriGC[:, 0], nz - 1), np.tile(m.TriGC[:, 1], nz - 1)
    za = np.repeat(c.zeta, nV)
    A = c.ops.ab.matrix("map", nV)
    Hb = np.tile(A @ Hi, nz - 1)
    Sb = np.tile(A @ Hs, nz - 1)
    out = OB.ops3d_apply(c.md, c.aa, c.zs, Hi, Hs, _poly1(xb, yb, zb), _poly1(xa, ya, za), _poly11(xs, ys, zsb))
    F = _cart(c, Hf, Sf, _poly1)
    # bk -> ak at the vertices
    z = Sf(xa, ya) - za * Hf(xa, ya)
    w = _fd_cart(F, xa, ya, z, (100.0, 100.0, 1.0))
    for i in range(2):
        assert np.abs(out["bk_ak"][i] - w[i]).max() < 1e-7 * np.abs(w[i]).max(), ("bk_ak", i)
    # ak -> bk at the triangle centres (geometry from the a->b map: exact for linear)
    Hbk, Sbk = np.tile(A @ Hi, nz), np.tile(A @ Hs, nz)
    z = Sbk - zb * Hbk
    w = _fd_cart(F, xb, yb, z, (100.0, 100.0, 1.0))
    for i in range(2):
        assert np.abs(out["ak_bk"][i] - w[i]).max() < 1e-7 * np.abs(w[i]).max(), ("ak_bk", i)
    # bk -> bks d/dz at the staggered levels (midpoints)
    z = Sb - zsb * Hb
    w = _fd_cart(F, xs, ys, z, (100.0, 100.0, 1.0))
    assert np.abs(out["bk_bks"] - w[2]).max() < 1e-9 * np.abs(w[2]).max()
    # maps of bilinear fields and the bks -> bk d/dz of a field linear in zeta
    assert np.abs(out["map_bks_ak"] - _poly11(xa, ya, za)).max() < 1e-12 * 10
    assert np.abs(out["map_ak_bks"] - _poly1(xs, ys, zsb)).max() > 0    # quadratic in zeta: not exact
    lin = OB.ops3d_apply(c.md, c.aa, c.zs, Hi, Hs, _poly11(xb, yb, zb), _poly11(xa, ya, za),
                         _poly11(xs, ys, zsb))
    assert np.abs(lin["map_ak_bks"] - _poly11(xs, ys, zsb)).max() < 1e-12 * 10
    assert np.abs(lin["bks_bk"][0] - _poly11(xb, yb, zb)).max() < 1e-12 * 10
    F11 = _cart(c, Hf, Sf, _poly11)
    z = Hbk * 0 + Sbk - zb * Hbk
    w = _fd_cart(F11, xb, yb, z, (100.0, 100.0, 1.0))
    assert np.abs(lin["bks_bk"][1] - w[2]).max() < 1e-9 * np.abs(w[2]).max()


# ---------------------------------------------------------------------------
# solve_BPA
# ---------------------------------------------------------------------------

def _params(**kw):
    p = L.default_params()
    p.glens_n = 1.0
    p.sliding_law = 1   # Weertman
    p.slid_weertman_m = 1.0
    p.visc_it_relax = 1.0
    p.visc_it_norm_dUV_tol = 1e-24
    p.visc_it_nit = 4
    p.rtol = 1e-13
    p.abstol = 0.0
    p.krylov_maxit = 20000
    p.do_gl_subgrid_friction = 0
    p.vel_max = 1e6
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _slab_inputs(c, sx, sy, H=1000.0, beta=500.0, A=None):
    m = c.mesh
    x, y = m.V[:, 0], m.V[:, 1]
    A = np.full(c.nz * m.nV, 5e-7) if A is None else A
    return dict(Hi=np.full(m.nV, H), Hs=2000.0 + sx * x + sy * y, A_flow=L.f64(A), beta_sq=np.full(m.nV, beta),
                fgr_b=np.ones(m.nT), mask_grounded=np.ones(m.nV, np.int32))


def _slab_exact(c, sx, sy, H, beta, eta, zeta):
    Mx = np.array([[1 + 4 * sx * sx + sy * sy, 3 * sx * sy], [3 * sx * sy, 1 + 4 * sy * sy + sx * sx]])
    Upp, Vpp = np.linalg.solve(Mx, RHO_G / eta * np.array([sx, sy]))
    ub, vb = -RHO_G * H * sx / beta, -RHO_G * H * sy / beta
    return ub - Upp * H * H * (1 - zeta ** 2) / 2, vb - Vpp * H * H * (1 - zeta ** 2) / 2


@pytest.mark.parametrize("sx,sy", [(-0.05, 0.0), (0.0, 0.04), (-0.03, 0.04)])
def test_newtonian_slab_closed_form(sx, sy):
    c = Case(h=12e3, zeta=np.array([0.0, 0.1, 0.25, 0.45, 0.7, 0.88, 1.0]))
    H, beta, eta = 1000.0, 500.0, 0.5 / 5e-7
    inp = _slab_inputs(c, sx, sy, H=H, beta=beta)
    rc, u, v, info = OB.solve_bpa(c.md, c.aa, c.zs, _params(), inp)
    assert rc in (0, 1), rc
    ue, ve = _slab_exact(c, sx, sy, H, beta, eta, c.zeta)
    U = u.reshape(c.nz, c.mesh.nT)
    Vv = v.reshape(c.nz, c.mesh.nT)
    scale = max(np.abs(ue).max(), np.abs(ve).max())
    assert np.abs(U - ue[:, None]).max() < 1e-8 * scale
    assert np.abs(Vv - ve[:, None]).max() < 1e-8 * scale
    # the longitudinal (4 s^2) terms are visible at this slope
    plain = -RHO_G * sx / eta * H * H * (1 - c.zeta ** 2) / 2 - RHO_G * H * sx / beta
    if sx != 0.0:
        assert np.abs(plain - ue).max() > 1e3 * 1e-8 * scale
    assert info.rel_residual < 1e-10


def test_no_sliding_slab():
    c = Case(h=12e3)
    sx = -0.04
    eta = 0.5 / 5e-7
    inp = _slab_inputs(c, sx, 0.0)
    rc, u, v, info = OB.solve_bpa(c.md, c.aa, c.zs, _params(sliding_law=0), inp)
    assert rc in (0, 1)
    Upp = RHO_G * sx / (eta * (1 + 4 * sx * sx))
    ue = -Upp * 1000.0 ** 2 * (1 - c.zeta ** 2) / 2
    U = u.reshape(c.nz, c.mesh.nT)
    assert np.abs(U - ue[:, None]).max() < 1e-8 * np.abs(ue).max()
    assert (U[-1] == 0).all()


def test_layered_flow_factor_convergence():
    """A = A0 + A1 zeta (n = 1) against the closed form over nz = 6, 11, 21,
    41: the deformation profile u - u_base converges at second order in the
    layer spacing; the basal velocity at first order, because the basal row
    reads the viscosity of the lowest staggered layer (BPA_main.f90:976,
    half a layer above the base).  A sign error in a viscosity-gradient term
    would leave an O(1) error."""
    sx, H, beta = -0.05, 1000.0, 500.0
    A0, A1 = 2e-7, 6e-7
    e_def, e_base = [], []
    for nz in (6, 11, 21, 41):
        c = Case(h=14e3, zeta=np.linspace(0.0, 1.0, nz))
        A = np.repeat(A0 + A1 * c.zeta, c.mesh.nV)
        rc, u, v, info = OB.solve_bpa(c.md, c.aa, c.zs, _params(), _slab_inputs(c, sx, 0.0, H=H, beta=beta, A=A))
        assert rc in (0, 1)
        ub = -RHO_G * H * sx / beta
        t = c.zeta
        ue = ub - RHO_G * sx * H * H / (1 + 4 * sx * sx) * (A0 * (1 - t * t) + 2.0 / 3.0 * A1 * (1 - t ** 3))
        U = u.reshape(nz, c.mesh.nT)
        e_def.append(np.abs((U - U[-1]) - (ue - ub)[:, None]).max() / (ue - ub).max())
        e_base.append(np.abs(U[-1] - ub).max() / abs(ub))
        assert np.abs(v).max() < 1e-9 * abs(ub)
    assert e_def[0] < 0.03 and e_base[-1] < 0.015, (e_def, e_base)
    for a, b in zip(e_def[:-1], e_def[1:]):
        assert a / b > 3.0, e_def
    for a, b in zip(e_base[:-1], e_base[1:]):
        assert 1.8 < a / b < 3.0, e_base


def test_prescribed_rows_and_no_grounded_ice():
    c = Case(h=12e3)
    m = c.mesh
    inp = _slab_inputs(c, -0.03, 0.0)
    rng = np.random.default_rng(1)
    mask = (rng.random(m.nT * c.nz) < 0.05).astype(np.int32)
    bu, bv = rng.normal(0, 50, m.nT * c.nz), rng.normal(0, 50, m.nT * c.nz)
    inp.update(bc_mask=mask, bc_u=bu, bc_v=bv)
    rc, u, v, _ = OB.solve_bpa(c.md, c.aa, c.zs, _params(), inp)
    assert rc in (0, 1)
    sel = mask == 1
    # identity rows, solved to the Krylov tolerance
    assert np.abs(u[sel] - bu[sel]).max() < 1e-8 * np.abs(u).max()
    assert np.abs(v[sel] - bv[sel]).max() < 1e-8 * np.abs(u).max()
    inp2 = _slab_inputs(c, -0.03, 0.0)
    inp2["mask_grounded"] = np.zeros(m.nV, np.int32)
    rc, u, v, _ = OB.solve_bpa(c.md, c.aa, c.zs, _params(), inp2, u0=np.ones(m.nT * c.nz))
    assert rc == 0 and (u == 0).all() and (v == 0).all()


def test_converged_glen_solution_and_residual():
    """n = 3 over a bumpy bed with weak linear sliding: the viscosity
    iteration converges, and the residual entry point vanishes at the
    solution it returns."""
    c = Case(h=12e3, zeta=np.linspace(0.0, 1.0, 11))
    m = c.mesh
    x, y = m.V[:, 0], m.V[:, 1]
    inp = _slab_inputs(c, -0.005, 0.0, A=np.full(c.nz * m.nV, 1e-16), beta=1e4)
    inp["Hi"] = 1000.0 + 200.0 * np.sin(x / LX * np.pi) * np.sin(y / LX * np.pi)
    p = _params(glens_n=3.0, visc_it_relax=0.5, visc_it_norm_dUV_tol=1e-15, visc_it_nit=300, rtol=1e-10)
    rc, u, v, info = OB.solve_bpa(c.md, c.aa, c.zs, p, inp, u0=np.full(c.nz * m.nT, 5.0))
    assert rc == 0, rc
    assert info.rel_residual < 1e-5
    # a second solve from the solution ends at the configured regularisation
    # (the first one's adaptive rule may have raised it), as the residual
    # entry point evaluates
    rc, u, v, info = OB.solve_bpa(c.md, c.aa, c.zs, p, inp, u0=u, v0=v)
    assert rc == 0 and info.epsilon_sq_0_applied == p.epsilon_sq_0 and info.rel_residual < 1e-5
    res = OB.bpa_residual(c.md, c.aa, c.zs, p, inp, u, v)
    assert res["rel_residual"] < 1e-5
    U = u.reshape(c.nz, m.nT)
    assert (U[0] > U[-1]).mean() > 0.95 and U[-1].min() > 1.0    # sliding plus shear
