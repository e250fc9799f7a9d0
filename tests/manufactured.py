"""Closed forms for the assembled SSA rows (shared by the oracle and GPU tests).

For Glen n = 1 the viscosity is 1/(2A) whatever the strain rate, so with
uniform A and H the depth-integrated viscosity N = H/(2A) is constant; with
Weertman m = 1 the friction coefficient is beta_sq.  The free rows of the
system (solve_linearised_SSA_DIVA.f90:180-329) applied to quadratic u, v
then equal, at each triangle's geometric centre,
   x: N (4 u_xx + u_yy + 3 v_xy) - beta u
   y: N (4 v_yy + v_xx + 3 u_xy) - beta v
(the second-order b->b operators are exact on quadratics) and the load
vector is -tau_d = rho g H grad(Hs) for a planar surface.
"""

import numpy as np

RHO_I, G = 910.0, 9.81

# u = cu . (1, x, y, x^2, xy, y^2) with x, y in units of S metres
S = 1e5
CU = np.array([30.0, 4.0, -2.0, 1.5, -0.7, 0.4])
CV = np.array([-10.0, 1.0, 3.0, -0.3, 0.9, -1.2])


def quad(c, x, y):
    x, y = x / S, y / S
    return c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y


def second(c):
    """(f_xx, f_xy, f_yy) in 1/m^2."""
    return 2 * c[3] / S ** 2, c[4] / S ** 2, 2 * c[5] / S ** 2


def linear_case(W, h=60e3, A=1e-16, H=1000.0, beta=1e3, slope=(-2e-3, 1e-3)):
    """Workload on a uniform mesh with uniform A, H and a planar surface."""
    mesh = W.uniform_mesh(-300e3, 300e3, -300e3, 300e3, h, seed=7)
    nV = mesh.nV
    Hb = slope[0] * mesh.V[:, 0] + slope[1] * mesh.V[:, 1] - 100.0
    zeta = W.zeta_regular(5)
    w = W.Workload("manufactured", mesh, W.build_operators(mesh), zeta, np.full(nV, H), Hb,
                   np.full(nV, -1e4), np.full(5 * nV, A), np.full(nV, beta), 0, "SSA", {})
    return w


def expected_rows(w, A, H, beta, crossterms=True):
    """Expected (A x*)[row] and b[row] of the free rows for quadratic x*."""
    xg, yg = w.mesh.TriGC[:, 0], w.mesh.TriGC[:, 1]
    u, v = quad(CU, xg, yg), quad(CV, xg, yg)
    uxx, uxy, uyy = second(CU)
    vxx, vxy, vyy = second(CV)
    N = H / (2 * A)
    ax = N * (4 * uxx + uyy + 3 * vxy) - beta * u
    ay = N * (4 * vyy + vxx + 3 * uxy) - beta * v
    Hs = w.Hs
    # planar surface: gradient from any triangle's vertices (exact)
    P = w.mesh.V
    gx = np.linalg.lstsq(np.c_[P, np.ones(len(P))], Hs, rcond=None)[0]
    bx = np.full(w.mesh.nT, RHO_I * G * H * gx[0])
    by = np.full(w.mesh.nT, RHO_I * G * H * gx[1])
    if not crossterms:
        ax, ay, bx, by = ax / N, ay / N, bx / N, by / N
    return u, v, ax, ay, bx, by
