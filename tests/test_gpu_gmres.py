"""Device GMRES(m) (the reference's PETSc default Krylov method,
petsc_basic.f90:106-119) through the C ABI: the linear-algebra boundary on the
reference's Laplace known answer and against a direct solve, restarts, and
velocity solves with GMRES inside the Picard loop against the oracle."""

import os

import numpy as np
import pytest

from conftest import max_rel, params

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("pc", ["none", "jacobi"])
def test_gmres_laplace_known_answer(V, L, W, pc):
    """ct_discretisation_solve_Laplace_eq.f90:142-207: the exact quadratic
    solves the discrete system; GMRES(30) must reach it."""
    import scipy.sparse as sp
    from test_oracle_linear import laplace_case
    A, b, f_ex = laplace_case(W)
    d = 1.0 / A.diagonal()
    As, bs = sp.csr_matrix(sp.diags(d) @ A), d * b
    x, its, rel = V.csr_solve(As, bs, rtol=1e-13, maxit=100000, krylov=L.KRYLOV_GMRES, restart=30,
                              precond=L.PC_NONE if pc == "none" else L.PC_JACOBI)
    assert its > 30 and rel <= 1e-13  # restarted at least once
    assert np.abs(x - f_ex).max() < 1e-8 * np.abs(f_ex).max()


@pytest.mark.parametrize("restart", [5, 30, 64])
def test_gmres_vs_direct(V, OB, L, dome, restart):
    import scipy.sparse.linalg as spl
    rng = np.random.default_rng(3)
    u = rng.normal(0, 50, dome.mesh.nT); v = rng.normal(0, 50, dome.mesh.nT)
    A, b = OB.Oracle(dome.descriptor()).system(params(L), dome.inputs(), u, v)
    x, its, rel = V.csr_solve(A, b, rtol=1e-12, maxit=50000, krylov=L.KRYLOV_GMRES, restart=restart)
    xd = spl.spsolve(A.tocsc(), b)
    assert its > 0 and rel <= 1e-12
    assert max_rel(x, xd) < 1e-8


def test_gmres_argument_errors(V, L, dome):
    import scipy.sparse as sp
    A = sp.identity(10, format="csr")
    with pytest.raises(V.VelocitySolverError):
        V.csr_solve(A, np.ones(10), krylov=L.KRYLOV_GMRES, restart=65)
    with pytest.raises(V.VelocitySolverError):
        V.csr_solve(A, np.ones(10), krylov=7)
    s = V.VelocitySolver(dome.descriptor())
    with pytest.raises(V.VelocitySolverError):
        s.set_params(krylov=1, gmres_restart=0)
    s.close()


def test_dome_reference_rule_with_gmres_matches_oracle(V, OB, L, dome):
    """The reference iteration with GMRES(30) on the device (block-Jacobi
    preconditioned) against the oracle's GMRES(30) + ILU(0): same Picard
    count, same velocities to the linear-solve tolerance."""
    kw = dict(visc_it_norm_dUV_tol=1e-6, rtol=1e-11, abstol=0.0, visc_it_nit=300)
    s = V.VelocitySolver(dome.descriptor(), krylov=1, gmres_restart=30, **kw)
    inp = dome.inputs()
    u, v, nv, na, info = s.solve_SSA(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"],
                                     inp["mask_grounded"])
    s.close()
    rc, uo, vo, io = OB.Oracle(dome.descriptor()).solve_ssa(params(L, **kw), inp)
    assert rc == 0 and nv == io.n_visc_its
    assert max_rel(u, uo) < 1e-6 and max_rel(v, vo) < 1e-6


@pytest.mark.parametrize("sb", ["SSA", "DIVA"])
def test_channel_gmres_amg_converged_solve_matches_oracle(V, W, sb):
    """GMRES(30) with the multigrid preconditioner inside the accelerated
    Picard loop reaches the oracle's converged channel solution (1e-5)."""
    from test_gpu_parity import CHOSEN, TIGHT, solve
    f = np.load(os.path.join(HERE, "golden", f"oracle_channel_{sb.lower()}.npz"))
    w = W.channel_small(target_nV=1500, stress_balance=sb)
    assert str(f["digest"]) == w.mesh.digest()
    st, info = solve(V, w, **TIGHT, **dict(CHOSEN, krylov=1, gmres_restart=30))
    assert info.rel_residual <= 1e-8
    sc = max(np.abs(f["u"]).max(), np.abs(f["v"]).max())
    err = max(np.abs(st["u_vav_b"] - f["u"]).max(), np.abs(st["v_vav_b"] - f["v"]).max()) / sc
    assert err < 1e-5, err


def test_mixed_precision_gmres_matches_fp64(V, dome):
    from test_gpu_parity import solve
    kw = dict(stop_rule=1, nonlinear_rtol=1e-8, visc_it_nit=300, abstol=0.0, rtol=1e-10)
    a, _ = solve(V, dome, precision=0, **kw)
    b, ib = solve(V, dome, precision=1, inner_rtol=0.1, krylov=1, gmres_restart=30, **kw)
    assert ib.n_refinements > 0
    assert max_rel(b["u_vav_b"], a["u_vav_b"]) < 1e-5 and max_rel(b["v_vav_b"], a["v_vav_b"]) < 1e-5
