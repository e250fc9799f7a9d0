"""Oracle Picard iteration pinned by the Schoof (2006) ice-stream closed form
(the reference's SSA_icestream integrated test) and by the DIVA -> SSA limit."""

import json
import os

import numpy as np
import pytest

from conftest import max_rel, params

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")))


def schoof_u(y, A, H, tantheta, L, m):
    """Schoof (2006) plastic-bed ice stream, u(y) for |y| <= W = L (m+1)^(1/m)
    and 0 outside (Schoof_SSA_solution.f90:15-64, Schoof2006_icestream.m)."""
    f = -910.0 * 9.81 * H * tantheta
    B = A ** (-1.0 / 3.0)
    W = L * (m + 1) ** (1.0 / m)
    s = np.abs(y / L)
    ua = -2 * f ** 3 * L ** 4 / (B ** 3 * H ** 3)
    ub = 0.25 * ((y / L) ** 4 - (m + 1) ** (4 / m))
    uc = -3 / ((m + 1) * (m + 4)) * (s ** (m + 4) - (m + 1) ** (1 + 4 / m))
    ud = 3 / ((m + 1) ** 2 * (2 * m + 4)) * (s ** (2 * m + 4) - (m + 1) ** (2 + 4 / m))
    ue = -1 / ((m + 1) ** 3 * (3 * m + 4)) * (s ** (3 * m + 4) - (m + 1) ** (3 + 4 / m))
    return np.where(np.abs(y) > W, 0.0, ua * (ub + uc + ud + ue))


def test_schoof_closed_form_limits():
    g = GOLD["schoof_icestream"]
    y = np.linspace(-400e3, 400e3, 801)
    u = schoof_u(y, g["A"], g["H"], g["dhdx"], g["L"], g["m"])
    W = g["L"] * (g["m"] + 1) ** (1 / g["m"])
    # continuous at the margin, even, maximal on the centre line
    assert abs(schoof_u(np.array([W * (1 - 1e-9)]), g["A"], g["H"], g["dhdx"], g["L"], g["m"])[0]) < 1e-3
    assert np.allclose(u, u[::-1])
    assert u.argmax() == 400 and 2000 < u.max() < 2500


def schoof_params(L, w, g):
    p = params(L, **{k: v for k, v in w.params.items()})
    p.visc_it_norm_dUV_tol = g["visc_it_norm_dUV_tol"]
    p.visc_it_nit = g["visc_it_nit"]
    return p


def test_schoof_icestream_oracle(W, L, OB):
    """The reference's own check with its own settings (RMSE of u against the
    closed form, reference L2 stop, 'infinite_SSA_icestream' east/west): the
    error falls by more than half from 32 to 16 km and at 16 km is below the
    reference scoreboard's 16 km value (303.7 m/yr; 181 here, 562 at 32 km
    against its 400.4: the meshes differ).  With the plain 'infinite' rows
    instead (u = mean of the neighbours) the errors are 908 and 490 m/yr: the
    lagged copy rows are what the closed form's du/dx = 0 needs."""
    g = GOLD["schoof_icestream"]
    rmse = {}
    for h in (32e3, 16e3):
        w = W.schoof_icestream(h=h)
        assert list(w.params["bc_u"]) == [1, 3, 1, 3]
        rc, u, v, info = OB.Oracle(w.descriptor()).solve_ssa(schoof_params(L, w, g), w.inputs())
        assert rc == 0, rc
        yc = w.mesh.TriGC[:, 1]
        ua = schoof_u(yc, g["A"], g["H"], g["dhdx"], g["L"], g["m"])
        rmse[h] = np.sqrt(np.mean((u - ua) ** 2))
        assert np.abs(v).max() < 0.15 * np.abs(u).max()
    assert rmse[16e3] < 0.5 * rmse[32e3]
    assert rmse[16e3] < g["reference_rmse"]["16km"]


def test_copy_rows_take_the_relaxed_copy_of_the_iterate(W, L, OB):
    """The lagged copy rows (solve_linearised_SSA_DIVA.f90:562-575): in the
    assembled system at a given iterate, every 'infinite_SSA_icestream' row is
    an identity row whose value is visc_it_relax * (the 1/d^2-weighted mean of
    the iterate on the triangles around the vertex nearest the copy point) +
    (1 - visc_it_relax) * the row's own value, the copy point being at the
    same y and a third / two thirds of the domain in x.  Checked against a
    brute-force numpy evaluation of that definition."""
    w = W.schoof_icestream(h=64e3)
    m = w.mesh
    p = schoof_params(L, w, GOLD["schoof_icestream"])
    rng = np.random.default_rng(3)
    u, v = rng.normal(0, 100, m.nT), rng.normal(0, 10, m.nT)
    A, b = OB.Oracle(w.descriptor()).system(p, w.inputs(), u, v)
    ops = w.ops
    xmin, xmax = m.V[:, 0].min(), m.V[:, 0].max()
    n = 0
    for ti in np.flatnonzero(m.TriBI > 0):
        side = (m.TriBI[ti] - 1) // 2
        if p.bc_u[side] != 3:
            continue
        gx, gy = m.TriGC[ti]
        px = xmin + (xmax - xmin) * (1 / 3 if gx < 0 else 2 / 3)
        vi = int(np.argmin(np.hypot(m.V[:, 0] - px, m.V[:, 1] - gy)))
        tris = ops.ba.ind[ops.ba.ptr[vi]:ops.ba.ptr[vi + 1]]
        wts = 1.0 / np.hypot(px - m.TriGC[tris, 0], gy - m.TriGC[tris, 1]) ** 2
        want = p.visc_it_relax * (wts @ u[tris]) / wts.sum() + (1 - p.visc_it_relax) * u[ti]
        row = 2 * ti
        assert A.indptr[row + 1] - A.indptr[row] == 1 and A.indices[A.indptr[row]] == row
        assert abs(b[row] - want) < 1e-12 * max(1.0, abs(want))
        n += 1
    assert n > 10


def frictionless_slab(W):
    """Ice over the whole domain, H = 1000 + 400 cos(pi x / L) cos(pi y / L),
    flat bed, zero velocity on the border, no friction."""
    mesh = W.uniform_mesh(-300e3, 300e3, -300e3, 300e3, 40e3, seed=9)
    x, y = mesh.V[:, 0], mesh.V[:, 1]
    H = 1000.0 + 400.0 * np.cos(np.pi * x / 600e3) * np.cos(np.pi * y / 600e3)
    nz, nV = 6, mesh.nV
    return W.Workload("frictionless_slab", mesh, W.build_operators(mesh), W.zeta_regular(nz), H,
                      np.zeros(nV), np.full(nV, -1e4), np.full(nz * nV, 1e-20), np.zeros(nV), 9, "DIVA", {})


def test_diva_reduces_to_ssa_without_friction(W, L, OB):
    """With beta = 0 the basal stress vanishes, so the DIVA vertical shear is
    zero and, for depth-uniform A, the DIVA viscosity equals the SSA one: the
    two solves must agree (DIVA_main.f90:88-262 vs SSA_main.f90:87-242)."""
    w = frictionless_slab(W)
    inp = w.inputs()
    p = params(L, bc_u=[1, 1, 1, 1], bc_v=[1, 1, 1, 1], stop_rule=1, nonlinear_rtol=1e-9, visc_it_nit=600,
               rtol=1e-12, abstol=0.0, visc_it_relax=0.5)
    o = OB.Oracle(w.descriptor())
    rc1, u1, v1, i1 = o.solve_ssa(p, inp)
    rc2, st, i2 = o.solve_diva(p, inp)
    assert rc1 == 0 and rc2 == 0, (rc1, rc2)
    assert i1.rel_residual < 1e-9 and i2.rel_residual < 1e-9
    assert 1.0 < np.abs(u1).max() < 1e4
    assert max_rel(st["u"], u1) < 1e-7 and max_rel(st["v"], v1) < 1e-7
    assert np.abs(st["tbx"]).max() == 0.0
    # no vertical shear: every layer ([k][ti] layout) moves with the vertical mean
    U = st["u3d"].reshape(w.nz, w.mesh.nT)
    assert max_rel(U, np.broadcast_to(st["u"], U.shape)) < 1e-12


def periodic_slab(W, Lx=80e3, h=8e3):
    """Uniform slab on a uniform slope over the ISMIP-HOM square [-L/2, L/2]^2
    with 'periodic_ISMIP-HOM' on every side: the driving stress is uniform,
    so with a linear sliding law (Weertman m = 1, beta = beta_sq) the SSA
    solution is the uniform velocity u = tau_d / beta_sq, and the periodic
    copy rows (relaxed copies of interior values) are satisfied by it."""
    mesh = W.uniform_mesh(-Lx / 2, Lx / 2, -Lx / 2, Lx / 2, h, seed=12)
    nV, nz = mesh.nV, 6
    x = mesh.V[:, 0]
    slope = 0.005
    H = np.full(nV, 1000.0)
    w = W.Workload("periodic_slab", mesh, W.build_operators(mesh), W.zeta_regular(nz), H, -slope * x,
                   np.full(nV, -1e4), np.full(nz * nV, 1e-16), np.full(nV, 1e3), 12, "SSA",
                   {"sliding_law": 1, "slid_weertman_m": 1.0, "bc_u": [2, 2, 2, 2], "bc_v": [2, 2, 2, 2],
                    "ismip_hom_L": Lx})
    u_exact = W.RHO_I * 9.81 * 1000.0 * slope / 1e3
    return w, u_exact


def test_periodic_ismip_hom_slab_is_uniform(W, L, OB):
    w, ue = periodic_slab(W)
    # unrelaxed copies: the lagged copy rows pass the periodic information on
    # by one copy per Picard iteration (relaxed by visc_it_relax = 0.2 this
    # takes thousands of iterations on this stiff, nearly strain-free slab)
    p = params(L, **w.params, stop_rule=1, nonlinear_rtol=1e-10, visc_it_nit=600, rtol=1e-12, abstol=0.0,
               include_crossterms=1, visc_it_relax=1.0)
    rc, u, v, info = OB.Oracle(w.descriptor()).solve_ssa(p, w.inputs())
    assert rc == 0, rc
    assert np.abs(u - ue).max() < 1e-6 * ue and np.abs(v).max() < 1e-6 * ue
