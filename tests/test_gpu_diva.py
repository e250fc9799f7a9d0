"""DIVA parity of the HIP engine (through the C ABI): one state evaluation at
the full benchmark sizes against the oracle, the Newtonian slab-on-slope
closed form, and the outputs of a solve stopped by the reference rule.  Also
the manufactured heat column with every advection and heating term."""

import numpy as np
import pytest

from conftest import max_rel, params

pytestmark = pytest.mark.gpu


def plausible_state(w, seed):
    """A DIVA state of the magnitudes a solve passes through (velocities of
    tens to hundreds of m/yr, viscosities 1e5..1e8 Pa yr, basal stresses of
    tens of kPa) at which to compare one evaluation."""
    rng = np.random.default_rng(seed)
    nT, nz = w.mesh.nT, w.nz
    return dict(u=rng.normal(80.0, 60.0, nT), v=rng.normal(0.0, 30.0, nT),
                eta_3D_b=10 ** rng.uniform(5.0, 8.0, nT * nz), tau_bx_b=rng.normal(0.0, 3e4, nT),
                tau_by_b=rng.normal(0.0, 1e4, nT), u_base_b=rng.normal(40.0, 30.0, nT),
                v_base_b=rng.normal(0.0, 10.0, nT))


@pytest.mark.parametrize("case", ["channel_200k", "continental_400k"])
def test_diva_state_evaluation_full_size(V, OB, L, W, case):
    """One DIVA evaluation (3-D viscosity, F-integrals, basal and effective
    friction, assembly and product) at the benchmark meshes equals the
    oracle's: every per-vertex / per-triangle kernel of the DIVA path at its
    real size."""
    w = W.channel(target_nV=200_000) if case == "channel_200k" else W.continental()
    print(f"{case}: nV={w.mesh.nV} nT={w.mesh.nT} nz={w.nz}", flush=True)
    inp = w.inputs()
    kw = dict(include_crossterms=0, vel_max=1e5)
    st = plausible_state(w, 31)
    s = V.VelocitySolver(w.descriptor(), **kw)
    s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    g = s.diva_residual(st["u"], st["v"], st["eta_3D_b"], st["tau_bx_b"], st["tau_by_b"], st["u_base_b"],
                        st["v_base_b"])
    s.close()
    o = OB.Oracle(w.descriptor()).diva_residual(params(L, **kw), inp, st["u"], st["v"], st["eta_3D_b"],
                                                st["tau_bx_b"], st["tau_by_b"], st["u_base_b"], st["v_base_b"])
    for k in ("eta_3D_b", "F1_3D_b", "F2_b_surface", "beta_b", "beta_eff_b", "b"):
        assert max_rel(g[k], o[k]) < 1e-12, k
    assert max_rel(g["r"], o["r"]) < 1e-10
    assert abs(g["rel_residual"] - o["rel_residual"]) < 1e-10 * o["rel_residual"]
    # the state is not degenerate: the shear and the friction both matter
    assert np.ptp(g["eta_3D_b"]) > 1e5 and np.abs(g["F1_3D_b"]).max() > 0.0
    assert (g["beta_eff_b"] < g["beta_b"] + 1e-300).all()


@pytest.mark.parametrize("sliding_law", [1, 0])
def test_diva_slab_on_slope_closed_form_gpu(V, L, W, sliding_law):
    """The Newtonian slab on a uniform slope (closed form in
    test_oracle_pins.py) solved by the engine: u_vav with the trapezoid's
    h^2/6 term, u_base, tau_b and every 3-D layer."""
    from test_oracle_pins import slab_closed_form, slab_workload
    nz = 11
    w, eta = slab_workload(W, nz)
    inp = w.inputs()
    s = V.VelocitySolver(w.descriptor(), glens_n=1.0, slid_weertman_m=1.0, sliding_law=sliding_law, stop_rule=1,
                         nonlinear_rtol=1e-11, visc_it_nit=50, visc_it_relax=1.0, rtol=1e-13, abstol=0.0)
    st, nv, na, info = s.solve_DIVA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"],
                                    inp["mask_grounded"])
    s.close()
    tau, ub, u3, uvd, uvc = slab_closed_form(nz, eta, sliding=sliding_law == 1)
    nT = w.mesh.nT
    assert info.rel_residual < 1e-10
    assert np.abs(st["u_vav_b"] - uvd).max() < 1e-9 * uvd and np.abs(st["v_vav_b"]).max() < 1e-9 * uvd
    assert np.abs(st["tau_bx_b"] - tau).max() < 1e-9 * tau
    assert np.abs(st["u_base_b"] - ub).max() < 1e-9 * uvd
    assert np.abs(st["u_3D_b"].reshape(nz, nT) - u3[:, None]).max() < 1e-9 * uvd
    assert max_rel(st["eta_3D_b"], np.full(nz * nT, eta)) < 1e-12


def test_diva_reference_rule_outputs_match_oracle(V, OB, L, W):
    """A solve stopped by the reference rule at visc_it_nit while the iterate
    still moves: the returned u_vav, u_base, tau_b, 3-D velocities and
    eta_3D_b are those of the loop's last state (calc_3D_velocities after the
    loop, DIVA_main.f90:254) on both sides, with near-exact linear solves."""
    w = W.channel_small(target_nV=1500, stress_balance="DIVA")
    kw = dict(include_crossterms=0, vel_max=1e5, visc_it_nit=12, visc_it_norm_dUV_tol=1e-30, rtol=1e-12,
              abstol=0.0, krylov_maxit=20000)
    inp = w.inputs()
    s = V.VelocitySolver(w.descriptor(), **kw)
    with pytest.warns(V.ViscosityIterationWarning):
        st, nv, na, info = s.solve_DIVA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"],
                                        inp["mask_grounded"])
    s.close()
    rc, o, oi = OB.Oracle(w.descriptor()).solve_diva(params(L, **kw), inp)
    assert rc == 1 and nv == oi.n_visc_its == 13
    assert oi.L2_uv > 1e-6  # still moving: a recomputed state would differ at this level
    pairs = (("u_vav_b", "u"), ("v_vav_b", "v"), ("u_base_b", "ubase"), ("tau_bx_b", "tbx"),
             ("u_3D_b", "u3d"), ("v_3D_b", "v3d"), ("eta_3D_b", "eta3b"))
    for gk, ok in pairs:
        assert max_rel(st[gk], o[ok]) < 1e-7, gk
    assert abs(info.rel_residual - oi.rel_residual) < 1e-6 * oi.rel_residual


@pytest.mark.parametrize("dt", [0.5, 50.0])
def test_heat_manufactured_quadratic_fixed_point_gpu(V, dt):
    """Quadratic columns with nonzero u, v, w, dzeta/dt, dzeta/dx, dzeta/dy,
    upwind terms and steady-balance heating (test_oracle_pins.py) are fixed
    points of one engine step, as they are of the oracle's."""
    from test_oracle_pins import manufactured_columns
    nV, nz, f, Ti, _ = manufactured_columns()
    hin, keep = V.make_heat_inputs(nV, nz, f["zeta"], {k: v for k, v in f.items() if k != "zeta"}, dt)
    T, unst, nun = V.heat_columns(hin)
    assert nun == 0 and np.abs(T.reshape(nz, nV) - Ti).max() < 1e-9 * 273.16
