"""The autotuner of the Python host: a coordinate search over solver and
multigrid settings that keeps only configurations reaching the same solution."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TIGHT = dict(include_crossterms=0, vel_max=1e5, abstol=0.0, stop_rule=1, nonlinear_rtol=1e-8, visc_it_nit=800,
             krylov_maxit=20000)
CHOSEN = dict(precision=0, precond=3, warm_start=1, rtol=0.01, accel_depth=2, accel_mixing=1.0, accel_continuation=5,
              accel_restart=5.0)


def test_autotune_keeps_the_oracle_solution(V, W):
    f = np.load(os.path.join(HERE, "golden", "oracle_channel_diva.npz"))
    w = W.channel_small(target_nV=1500, stress_balance="DIVA")
    assert str(f["digest"]) == w.mesh.digest()
    inp = w.inputs()
    s = V.VelocitySolver(w.descriptor(), **TIGHT, **CHOSEN)
    s.upload(inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    axes = (("accel_depth", (2, 3)), ("omega", (0.75, 0.85)), ("coarse_scale", (1.5, 1.7)))
    best, table = s.autotune("DIVA", axes=axes)
    for cfg, t in table:
        print(cfg, t, flush=True)
    assert len(table) == 4 and table[0][1] is not None
    times = [t for _, t in table if t is not None]
    assert best in [cfg for cfg, _ in table]
    assert dict(table[[cfg for cfg, _ in table].index(best)][0]) == best
    assert min(times) == [t for cfg, t in table if cfg == best][0]
    # the solver is left on the winner, which reaches the oracle's solution
    assert s.params.accel_depth == best["accel_depth"]
    assert s.amg_options["omega"] == best["omega"] and s.amg_options["coarse_scale"] == best["coarse_scale"]
    s.set_velocity(None, None)
    s.reset_diva_state()
    info = s.run("DIVA")
    u, v = s.get_velocity()
    s.close()
    assert info.rel_residual <= 1e-8
    sc = max(np.abs(f["u"]).max(), np.abs(f["v"]).max())
    assert max(np.abs(u - f["u"]).max(), np.abs(v - f["v"]).max()) / sc < 1e-5


def test_amg_options_are_validated(V, W):
    w = W.channel_small(target_nV=1500, stress_balance="DIVA")
    s = V.VelocitySolver(w.descriptor(), precond=3)
    with pytest.raises(V.VelocitySolverError):
        s.set_amg_options(omega=1.5)
    with pytest.raises(KeyError):
        s.set_amg_options(passes=2)
    s.set_amg_options(omega=0.8, tail_rows=0)
    assert s.amg_options == {"omega": 0.8, "coarse_scale": 1.5, "tail_rows": 0}
    s.close()
