"""The Tsai2015, Schoof2005 and Zoet-Iverson sliding laws (sliding_laws.f90:
244-403) on the device, through the C ABI, against the oracle (whose laws are
pinned in test_sliding_laws.py): one SSA, DIVA and BPA evaluation each at a
velocity that is not a solution, and converged SSA solves on the 40 km dome.
The yield-stress laws take the till yield stress formed by the library's host
routine (N tan(pi/180) phi, extended to ice-free land)."""

import numpy as np
import pytest

from conftest import max_rel, params
from test_gpu_bpa import _bpa_kw, _bumpy, _field, _solver
from test_gpu_diva import plausible_state
from test_oracle_bpa import Case

pytestmark = pytest.mark.gpu

RHO_G = 910.0 * 9.81
TSAI, SCHOOF, ZI = 4, 5, 6


def law_fields(V, w, law, seed, n_frac=0.01):
    """(rough, alpha_sq, effective_pressure): beta_sq and a Coulomb cap of the
    same order as the Weertman stress for Tsai/Schoof; the till yield stress
    at a friction angle near the reference's uniform 15 degrees for
    Zoet-Iverson."""
    m = w.mesh
    rng = np.random.default_rng(seed)
    H = np.maximum(w.Hi, 1.0) if hasattr(w, "Hi") else np.full(m.nV, 1000.0)
    N = n_frac * RHO_G * H * rng.uniform(0.5, 1.5, m.nV)
    alpha = rng.uniform(0.3, 0.7, m.nV)
    if law == ZI:
        phi = rng.uniform(10.0, 20.0, m.nV)
        grounded = np.ones(m.nV, np.int32)
        land = np.zeros(m.nV, np.int32)
        rough = V.till_yield_stress(m, N, phi, grounded, land)
        return rough, None, None
    rough = rng.uniform(0.5e4, 2e4, m.nV)
    return rough, alpha, N


@pytest.mark.parametrize("law", [TSAI, SCHOOF, ZI], ids=["tsai2015", "schoof2005", "zoet_iverson"])
def test_ssa_residual_parity(V, OB, L, chan, law):
    inp = dict(chan.inputs())
    rough, alpha, N = law_fields(V, chan, law, 11)
    inp["beta_sq"] = rough
    kw = dict(include_crossterms=0, sliding_law=law, slid_delta_v=1e-3)
    s = V.VelocitySolver(chan.descriptor(), **kw)
    if alpha is not None:
        s.set_sliding_cap(alpha, N)
    s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    rng = np.random.default_rng(7)
    u = rng.normal(0, 80, chan.mesh.nT); v = rng.normal(0, 80, chan.mesh.nT)
    rel_g, r_g, b_g = s.residual(u, v)
    s.close()
    if alpha is not None:
        OB.set_sliding_cap(alpha, N)
    try:
        rel_o, r_o, b_o = OB.Oracle(chan.descriptor()).residual(params(L, **kw), inp, u, v)
    finally:
        OB.set_sliding_cap()
    assert max_rel(b_g, b_o) < 1e-12
    assert max_rel(r_g, r_o) < 1e-11
    assert abs(rel_g - rel_o) < 1e-11 * rel_o


@pytest.mark.parametrize("law", [SCHOOF, ZI], ids=["schoof2005", "zoet_iverson"])
def test_diva_evaluation_parity(V, OB, L, W, law):
    w = W.channel_small(target_nV=3000, stress_balance="DIVA")
    inp = dict(w.inputs())
    rough, alpha, N = law_fields(V, w, law, 12)
    inp["beta_sq"] = rough
    kw = dict(include_crossterms=0, vel_max=1e5, sliding_law=law)
    st = plausible_state(w, 32)
    s = V.VelocitySolver(w.descriptor(), **kw)
    if alpha is not None:
        s.set_sliding_cap(alpha, N)
    s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    g = s.diva_residual(st["u"], st["v"], st["eta_3D_b"], st["tau_bx_b"], st["tau_by_b"], st["u_base_b"],
                        st["v_base_b"])
    s.close()
    if alpha is not None:
        OB.set_sliding_cap(alpha, N)
    try:
        o = OB.Oracle(w.descriptor()).diva_residual(params(L, **kw), inp, st["u"], st["v"], st["eta_3D_b"],
                                                    st["tau_bx_b"], st["tau_by_b"], st["u_base_b"], st["v_base_b"])
    finally:
        OB.set_sliding_cap()
    for k in ("beta_b", "beta_eff_b", "b"):
        assert max_rel(g[k], o[k]) < 1e-12, k
    assert max_rel(g["r"], o["r"]) < 1e-10


@pytest.mark.parametrize("law", [TSAI, ZI], ids=["tsai2015", "zoet_iverson"])
def test_bpa_evaluation_parity(V, OB, L, law):
    c = Case(h=9e3, zeta=np.linspace(0.0, 1.0, 9))
    m = c.mesh
    inp = _bumpy(c)
    rng = np.random.default_rng(13)
    N = 0.01 * RHO_G * inp["Hi"] * rng.uniform(0.5, 1.5, m.nV)
    alpha = rng.uniform(0.3, 0.7, m.nV)
    if law == ZI:
        inp["beta_sq"] = N * np.tan(np.pi / 180) * rng.uniform(10.0, 20.0, m.nV)
    kw = dict(glens_n=3.0, sliding_law=law)
    u, v = _field(c, 9)
    s = _solver(V, c, **kw)
    if law == TSAI:
        s.set_sliding_cap(alpha, N)
        OB.set_sliding_cap(alpha, N)
    try:
        gpu = s.bpa_residual(u, v, **_bpa_kw(inp))
        ora = OB.bpa_residual(c.md, c.aa, c.zs, params(L, **kw), inp, u, v)
    finally:
        OB.set_sliding_cap()
        s.close()
    for k in ("b", "eta_bk", "beta_b"):
        assert max_rel(gpu[k], ora[k]) < 1e-11, k
    scale = max(np.abs(ora["b"]).max(), np.abs(ora["r"]).max())
    assert np.abs(gpu["r"] - ora["r"]).max() < 1e-10 * scale


@pytest.mark.parametrize("law", [TSAI, SCHOOF, ZI], ids=["tsai2015", "schoof2005", "zoet_iverson"])
def test_dome_solve_matches_oracle(V, OB, L, W, law):
    """The 40 km dome of configuration 1 with each law (a cap near the
    Weertman stress, so both branches of Tsai2015 occur; the steep margin
    reaches vel_max under the weaker laws, as the reference clamps), the
    reference's relaxed Picard iteration with tighter stops: same iteration
    count within one, velocities within 1e-5."""
    w = W.dome(h=40e3)
    inp = dict(w.inputs())
    m = w.mesh
    rng = np.random.default_rng(21)
    N = 0.02 * RHO_G * np.maximum(inp["Hi"], 500.0) * rng.uniform(0.8, 1.2, m.nV)
    alpha = np.full(m.nV, 0.5)
    if law == ZI:
        inp["beta_sq"] = V.till_yield_stress(m, N, np.full(m.nV, 15.0), inp["mask_grounded"],
                                             np.zeros(m.nV, np.int32))
    kw = dict(w.params, sliding_law=law, abstol=0.0, visc_it_nit=300, visc_it_norm_dUV_tol=1e-5, rtol=1e-10)
    s = V.VelocitySolver(w.descriptor(), **kw)
    if law != ZI:
        s.set_sliding_cap(alpha, N)
    print(f"dome solve law={law} nV={m.nV}", flush=True)
    u, v, nv, na, info = s.solve_SSA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"],
                                     inp["mask_grounded"])
    s.close()
    if law != ZI:
        OB.set_sliding_cap(alpha, N)
    try:
        rc, uo, vo, io = OB.Oracle(w.descriptor()).solve_ssa(params(L, **kw), inp)
    finally:
        OB.set_sliding_cap()
    assert rc == 0 and abs(info.n_visc_its - io.n_visc_its) <= 1
    assert max_rel(u, uo) < 1e-5 and max_rel(v, vo) < 1e-5
    assert np.hypot(u, v).max() > 1.0
    if law == TSAI:
        # both branches of min(tau_W, alpha_sq N) occur at the solution
        ba = w.ops.ba
        ua = np.array([ba.vals["map"][ba.ptr[i]:ba.ptr[i + 1]] @ u[ba.ind[ba.ptr[i]:ba.ptr[i + 1]]]
                       for i in range(m.nV)])
        va = np.array([ba.vals["map"][ba.ptr[i]:ba.ptr[i + 1]] @ v[ba.ind[ba.ptr[i]:ba.ptr[i + 1]]]
                       for i in range(m.nV)])
        tau_w = inp["beta_sq"] * np.sqrt(1e-6 + ua ** 2 + va ** 2) ** (1 / kw.get("slid_weertman_m", 3.0))
        g = inp["mask_grounded"] > 0
        assert (tau_w[g] > alpha[g] * N[g]).sum() > 10 and (tau_w[g] < alpha[g] * N[g]).sum() > 10


def test_cap_is_required(V, dome):
    inp = dome.inputs()
    s = V.VelocitySolver(dome.descriptor(), sliding_law=TSAI)
    s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    u = np.ones(dome.mesh.nT)
    with pytest.raises(V.VelocitySolverError):
        s.residual(u, u)
    nV = dome.mesh.nV
    s.set_sliding_cap(np.full(nV, 0.5), np.full(nV, 1e5))
    s.residual(u, u)
    s.set_sliding_cap()
    with pytest.raises(V.VelocitySolverError):
        s.run("SSA")
    with pytest.raises(V.VelocitySolverError):
        s.set_params(sliding_law=7)
    s.close()
