"""The heat-column inputs on the device (ufe_heat_upwind_derivatives,
ufe_heat_strain_heating, ufe_heat_frictional_heating) against the oracle:
same sums in the same order, so to round-off (1e-12 relative), on the channel
mesh and the 399k-vertex continental mesh; the reference's error for a
vertex without an upwind triangle."""

import numpy as np
import pytest

from conftest import max_rel
from test_oracle_heat_inputs import _fields

pytestmark = pytest.mark.gpu


def _strain_fields(w, seed):
    rng = np.random.default_rng(seed)
    n = w.nz * w.mesh.nV
    e = [rng.normal(0.0, 1e-3, n) for _ in range(9)]
    A = 10 ** rng.uniform(-17.0, -15.0, n)
    return e, A


@pytest.mark.parametrize("case", ["channel", "continental_400k"])
def test_heat_inputs_match_oracle(V, OB, W, case):
    w = W.channel_small(target_nV=1500, stress_balance="DIVA") if case == "channel" else W.continental()
    m, nz = w.mesh, w.nz
    print(f"{case}: nV={m.nV} nT={m.nT} nz={nz}", flush=True)
    Ti, Hi, u, v, b, c = _fields(w, 7)
    Ti = Ti + np.random.default_rng(8).normal(0.0, 0.5, Ti.size)   # not linear: the triangle choice matters
    rng = np.random.default_rng(9)
    u3, v3 = rng.normal(50.0, 80.0, m.nT * nz), rng.normal(0.0, 40.0, m.nT * nz)
    s = V.VelocitySolver(w.descriptor(), glens_n=3.0)
    gx, gy = s.upwind_heat_flux_derivatives(Ti, Hi, u, v, u3, v3)
    rc, ox, oy = OB.heat_upwind(w.descriptor(), Ti, Hi, u, v, u3, v3)
    assert rc == 0
    assert max_rel(gx, ox) < 1e-12 and max_rel(gy, oy) < 1e-12
    assert np.abs(gx).max() > 0.0
    e, A = _strain_fields(w, 10)
    Phi = s.strain_heating(Hi, A, e)
    assert max_rel(Phi, OB.strain_heating(m.nV, nz, Hi, A, e, 3.0)) < 1e-12
    beta, ub = rng.uniform(0.0, 1e4, m.nV), rng.uniform(0.0, 300.0, m.nV)
    g = (rng.random(m.nV) < 0.7).astype(np.int32)
    assert np.array_equal(s.frictional_heating(beta, ub, g), OB.frictional_heating(beta, ub, g))
    s.close()


def test_inflow_border_vertex_is_an_error(V, W):
    w = W.channel_small(target_nV=1500, stress_balance="DIVA")
    m = w.mesh
    Ti, Hi, u, v, b, c = _fields(w, 3)
    Hi[:] = 1000.0
    vb = np.flatnonzero(m.VBI == 3)[0]
    u[vb], v[vb] = -100.0, 0.0
    s = V.VelocitySolver(w.descriptor())
    z = np.zeros(m.nT * w.nz)
    with pytest.raises(V.VelocitySolverError, match="upwind triangle"):
        s.upwind_heat_flux_derivatives(Ti, Hi, u, v, z, z)
    s.close()


def test_frictional_heating_without_sliding(V, W):
    w = W.channel_small(target_nV=1500, stress_balance="DIVA")
    s = V.VelocitySolver(w.descriptor(), sliding_law=0)
    n = w.mesh.nV
    assert (s.frictional_heating(np.ones(n), np.ones(n), np.ones(n, np.int32)) == 0).all()
    s.close()
