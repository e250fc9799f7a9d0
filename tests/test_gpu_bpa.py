"""Parity of the engine's BPA path (through the C ABI) with the oracle:
calc_zeta_gradients (zeta_gradients.f90:19-212), one evaluation of the BPA
viscosity iteration's state and system (BPA_main.f90:330-1978) and whole
solve_BPA calls (:90-252).

Bars: the zeta gradients and the state/system evaluation to 1e-11 of each
field's magnitude (the same sums in a different order); converged solves to
1e-6 of the largest speed (both sides solve their linear systems to 1e-10 and
stop the viscosity iteration at the same tolerance); the Newtonian slab to its
closed form at 1e-8, as the oracle."""

import os
import sys
import warnings

import numpy as np
import pytest

from conftest import max_rel, params

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_bpa import Case, LX, RHO_G, _slab_exact, _slab_inputs  # noqa: E402

pytestmark = pytest.mark.gpu


def _solver(V, c, **kw):
    s = V.VelocitySolver(c.md, **kw)
    s.zeta_setup(c.aa, c.zs)
    return s


def _bumpy(c, sx=-0.005, beta=1e4, amp=200.0, A=1e-16):
    m = c.mesh
    x, y = m.V[:, 0], m.V[:, 1]
    inp = _slab_inputs(c, sx, 0.0, A=np.full(c.nz * m.nV, A), beta=beta)
    inp["Hi"] = 1000.0 + amp * np.sin(x / LX * np.pi) * np.sin(y / LX * np.pi)
    return inp


def _field(c, seed, scale=100.0):
    m = c.mesh
    rng = np.random.default_rng(seed)
    x, y = np.tile(m.TriGC[:, 0], c.nz), np.tile(m.TriGC[:, 1], c.nz)
    t = np.repeat(c.zeta, m.nT)
    a = rng.normal(size=6)
    u = scale * (1 + 0.3 * np.sin(x / 3e4 + a[0]) * np.cos(y / 4e4 + a[1]) + 0.5 * (1 - t ** 2) + 0.05 * a[2])
    v = 0.3 * scale * (np.cos(x / 5e4 + a[3]) * np.sin(y / 2e4 + a[4]) + 0.4 * t) + rng.normal(0, 1.0, x.size)
    return u, v


def _bpa_kw(inp):
    return dict(Hi=inp["Hi"], Hs=inp["Hs"], A_flow=inp["A_flow"], beta_sq=inp["beta_sq"], fraction_gr_b=inp["fgr_b"],
                BC_prescr_mask_bk=inp.get("bc_mask"), BC_prescr_u_bk=inp.get("bc_u"),
                BC_prescr_v_bk=inp.get("bc_v"))


def test_zeta_gradients_match_oracle(V, OB):
    c = Case(h=9e3, zeta=np.array([0.0, 0.1, 0.25, 0.45, 0.7, 0.88, 1.0]))
    m = c.mesh
    x, y = m.V[:, 0], m.V[:, 1]
    Hi = 900.0 + 300.0 * np.sin(x / 4e4) * np.cos(y / 3e4) + 1e-3 * x
    Hi[::17] = 3.0   # thin ice: the 10 m floor
    Hs = 1500.0 - 4e-3 * x + 200.0 * np.cos(y / 5e4)
    dHi, dHs = np.sin(x / 1e4), np.cos(y / 1e4)
    print(f"zeta gradients nV={m.nV}", flush=True)
    s = _solver(V, c)
    g = s.zeta_gradients(Hi, Hs, dHi, dHs)
    o = OB.zeta_gradients(c.md, c.aa, c.zs, Hi, Hs, dHi, dHs)
    for grid, a, b in zip(("ak", "bk", "bks"), g, o):
        for i, (p, q) in enumerate(zip(a, b)):
            assert max_rel(p, q) < 1e-11, (grid, i)
    s.close()


CASES = {
    "bumpy_weertman3": dict(glens_n=3.0, sliding_law=1, slid_weertman_m=3.0),
    "bumpy_linear": dict(glens_n=3.0, sliding_law=1, slid_weertman_m=1.0),
    "no_sliding": dict(glens_n=3.0, sliding_law=0),
    "budd_zero_borders": dict(glens_n=3.0, sliding_law=2, bc_u=[1, 0, 1, 0], bc_v=[0, 1, 0, 1]),
}


@pytest.mark.parametrize("name", list(CASES))
def test_state_and_system_match_oracle(V, OB, L, name):
    """One evaluation at a velocity field that is not a solution: r = b - A u,
    b, eta_bk and beta_b, with prescribed points and varied borders."""
    c = Case(h=9e3, zeta=np.linspace(0.0, 1.0, 9))
    m = c.mesh
    kw = dict(CASES[name])
    p = params(L, **kw)
    inp = _bumpy(c, beta=1e4)
    if name == "budd_zero_borders":
        inp["beta_sq"] = np.full(m.nV, 3e4)
    rng = np.random.default_rng(3)
    inp["fgr_b"] = np.clip(rng.uniform(0.3, 1.4, m.nT), 0, 1)
    mask = (rng.random(m.nT * c.nz) < 0.04).astype(np.int32)
    inp.update(bc_mask=mask, bc_u=rng.normal(0, 30, m.nT * c.nz), bc_v=rng.normal(0, 30, m.nT * c.nz))
    u, v = _field(c, 7)
    print(f"BPA state {name} nT={m.nT} nz={c.nz}", flush=True)
    s = _solver(V, c, **kw)
    gpu = s.bpa_residual(u, v, **_bpa_kw(inp))
    ora = OB.bpa_residual(c.md, c.aa, c.zs, p, inp, u, v)
    for k in ("b", "eta_bk", "beta_b"):
        assert max_rel(gpu[k], ora[k]) < 1e-11, k
    # r = b - A u is a difference of large terms: compare against |A u| ~ |b| + |r|
    scale = max(np.abs(ora["b"]).max(), np.abs(ora["r"]).max())
    assert np.abs(gpu["r"] - ora["r"]).max() < 1e-10 * scale
    assert abs(gpu["rel_residual"] - ora["rel_residual"]) < 1e-9 * ora["rel_residual"]
    s.close()


def test_periodic_borders_match_oracle(V, OB, L):
    """'periodic_ISMIP-HOM' borders on a square centred on the origin: the
    lagged copy rows per layer."""
    from ufemism2_amd import mesh as M
    from ufemism2_amd.operators import MeshDescriptor, build_aa, build_operators, zeta_stag

    class Sq:
        pass

    c = Sq()
    Lh = 40e3
    c.mesh = M.uniform_mesh(-Lh, Lh, -Lh, Lh, 5e3, seed=4)
    c.ops = build_operators(c.mesh)
    c.aa = build_aa(c.mesh)
    c.zeta = np.linspace(0.0, 1.0, 7)
    c.zs = zeta_stag(c.zeta)
    c.md = MeshDescriptor(c.mesh, c.ops, c.zeta)
    c.nz = 7
    m = c.mesh
    x, y = m.V[:, 0], m.V[:, 1]
    kw = dict(glens_n=3.0, sliding_law=0, bc_u=[2] * 4, bc_v=[2] * 4, ismip_hom_L=2 * Lh)
    p = params(L, **kw)
    inp = dict(Hi=1000.0 - 500.0 * np.sin(2 * np.pi * x / (2 * Lh)) * np.sin(2 * np.pi * y / (2 * Lh)),
               Hs=2000.0 - x * np.tan(0.5 * np.pi / 180), A_flow=np.full(c.nz * m.nV, 1e-16),
               beta_sq=np.full(m.nV, 1e4), fgr_b=np.ones(m.nT), mask_grounded=np.ones(m.nV, np.int32))
    u, v = _field(c, 11, scale=20.0)
    s = _solver(V, c, **kw)
    gpu = s.bpa_residual(u, v, **_bpa_kw(inp))
    ora = OB.bpa_residual(c.md, c.aa, c.zs, p, inp, u, v)
    assert max_rel(gpu["b"], ora["b"]) < 1e-11
    scale = max(np.abs(ora["b"]).max(), np.abs(ora["r"]).max())
    assert np.abs(gpu["r"] - ora["r"]).max() < 1e-10 * scale
    s.close()


@pytest.mark.parametrize("sx,sy", [(-0.05, 0.0), (-0.03, 0.04)])
def test_newtonian_slab_closed_form(V, L, sx, sy):
    c = Case(h=12e3, zeta=np.array([0.0, 0.1, 0.25, 0.45, 0.7, 0.88, 1.0]))
    H, beta, eta = 1000.0, 500.0, 0.5 / 5e-7
    inp = _slab_inputs(c, sx, sy, H=H, beta=beta)
    kw = dict(glens_n=1.0, sliding_law=1, slid_weertman_m=1.0, visc_it_relax=1.0, visc_it_norm_dUV_tol=1e-24,
              visc_it_nit=4, rtol=1e-13, abstol=0.0, krylov_maxit=20000, do_gl_subgrid_friction=0, vel_max=1e6)
    s = _solver(V, c, **kw)
    with warnings.catch_warnings():   # the linear problem may run to visc_it_nit
        warnings.simplefilter("ignore", V.ViscosityIterationWarning)
        u, v, info = s.solve_BPA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"],
                                 inp["mask_grounded"])
    ue, ve = _slab_exact(c, sx, sy, H, beta, eta, c.zeta)
    scale = max(np.abs(ue).max(), np.abs(ve).max())
    assert np.abs(u.reshape(c.nz, -1) - ue[:, None]).max() < 1e-8 * scale
    assert np.abs(v.reshape(c.nz, -1) - ve[:, None]).max() < 1e-8 * scale
    assert info.rel_residual < 1e-10
    s.close()


@pytest.mark.parametrize("sliding", [1, 0])
def test_converged_solve_matches_oracle(V, OB, L, sliding):
    """n = 3 over a bumpy bed, weak linear sliding or none, from the same start."""
    c = Case(h=12e3, zeta=np.linspace(0.0, 1.0, 11))
    m = c.mesh
    inp = _bumpy(c)
    kw = dict(glens_n=3.0, sliding_law=sliding, slid_weertman_m=1.0, visc_it_relax=0.5, visc_it_norm_dUV_tol=1e-15,
              visc_it_nit=300, rtol=1e-10, abstol=0.0, krylov_maxit=20000, do_gl_subgrid_friction=0, vel_max=1e6)
    u0 = np.full(c.nz * m.nT, 5.0)
    rc, uo, vo, io = OB.solve_bpa(c.md, c.aa, c.zs, params(L, **kw), inp, u0=u0)
    assert rc == 0
    s = _solver(V, c, **kw)
    print(f"BPA solve sliding={sliding} nT={m.nT}", flush=True)
    u, v, info = s.solve_BPA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"],
                             u_bk=u0)
    scale = np.hypot(uo, vo).max()
    assert np.abs(u - uo).max() < 1e-6 * scale and np.abs(v - vo).max() < 1e-6 * scale
    assert info.rel_residual < 1e-5
    assert abs(info.n_visc_its - io.n_visc_its) <= 2
    s.close()


def test_no_grounded_ice_and_errors(V, L):
    c = Case(h=12e3)
    m = c.mesh
    inp = _bumpy(c)
    s = _solver(V, c)
    u, v, _ = s.solve_BPA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"],
                          np.zeros(m.nV, np.int32), u_bk=np.ones(c.nz * m.nT))
    assert (u == 0).all() and (v == 0).all()
    s.set_params(bc_u=[3, 0, 0, 0])      # 'infinite_SSA_icestream' is not a BPA border
    with pytest.raises(V.VelocitySolverError):
        s.solve_BPA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"])
    s.set_params(bc_u=[0, 0, 0, 0])
    with pytest.raises(V.VelocitySolverError):
        s.solve_BPA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"],
                    BC_prescr_mask_bk=np.zeros(c.nz * m.nT, np.int32))
    s.close()


@pytest.mark.parametrize("nV", [50_000])
def test_state_and_system_at_scale(V, OB, L, W, nV):
    """One evaluation on the channel workload mesh (~50k vertices, 10 layers:
    1M bk points, 36M blocks), against the oracle."""
    from ufemism2_amd.operators import build_aa, zeta_stag
    w = W.channel(seed=2, target_nV=nV, nz=10, stress_balance="DIVA")
    m = w.mesh

    class Cw:
        pass

    c = Cw()
    c.mesh, c.md, c.zeta, c.nz = m, w.descriptor(), w.zeta, w.nz
    c.aa, c.zs = build_aa(m), zeta_stag(w.zeta)
    wi = w.inputs()
    inp = dict(Hi=wi["Hi"], Hs=wi["Hs"], A_flow=wi["A_flow"], beta_sq=wi["beta_sq"], fgr_b=wi["fgr_b"])
    u, v = _field(c, 5, scale=200.0)
    kw = dict(glens_n=3.0)
    print(f"BPA state at scale nT={m.nT} nz={c.nz}", flush=True)
    s = _solver(V, c, **kw)
    gpu = s.bpa_residual(u, v, **_bpa_kw(inp))
    ora = OB.bpa_residual(c.md, c.aa, c.zs, params(L, **kw), inp, u, v)
    for k in ("b", "eta_bk", "beta_b"):
        assert max_rel(gpu[k], ora[k]) < 1e-11, k
    scale = max(np.abs(ora["b"]).max(), np.abs(ora["r"]).max())
    assert np.abs(gpu["r"] - ora["r"]).max() < 1e-10 * scale
    s.close()


@pytest.mark.parametrize("tight", [True, False])
def test_ismip_hom_a_reference_configuration(V, OB, L, W, tight):
    """The reference's BPA integrated test, ISMIP-HOM A at L = 160 km on its
    8 km mesh with its settings (no sliding, periodic borders, relaxation 0.4
    to 5e-7, Krylov rtol 1e-6): the engine against the oracle.  Tight: both
    solve to 1e-10 / 1e-13 and agree to 1e-6 of the largest speed; at the
    reference's tolerances the two stop after nearly the same number of
    viscosity iterations on different linear-solver round-off, within 1e-3."""
    from ufemism2_amd.operators import build_aa, zeta_stag
    w = W.ismip_hom_a(h=8e3)
    kw = dict(w.params)
    if tight:
        kw.update(rtol=1e-10, abstol=0.0, visc_it_norm_dUV_tol=1e-13)
    aa, zs = build_aa(w.mesh), zeta_stag(w.zeta)
    inp = w.inputs()
    rc, uo, vo, io = OB.solve_bpa(w.descriptor(), aa, zs, params(L, **kw), inp)
    assert rc == 0
    s = V.VelocitySolver(w.descriptor(), **kw)
    s.zeta_setup(aa, zs)
    print(f"ISMIP-HOM A 160 km nT={w.mesh.nT} tight={tight}", flush=True)
    u, v, info = s.solve_BPA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    scale = np.hypot(uo, vo).max()
    err = max(np.abs(u - uo).max(), np.abs(v - vo).max()) / scale
    assert err < (1e-6 if tight else 1e-3), err
    assert abs(info.n_visc_its - io.n_visc_its) <= (2 if tight else 3)
    U = u.reshape(w.nz, -1)
    assert 1.0 < U[0].min() and U[0].max() < 120.0   # surface speeds of the experiment (m/yr)
    s.close()
