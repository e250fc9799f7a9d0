"""Parity of the engine's semi-implicit ice-thickness step (through the C
ABI) with the oracle's: calc_dHi_dt_semiimplicit
(conservation_of_mass_semiimplicit.f90:23-159) on the channel and
continental meshes with converged velocities, at the continental benchmark
mesh with a synthetic flow, with either border condition, both Krylov
methods, prescribed and no-ice rows, an initial guess, and the context's own
velocity.

Bars: divQ = M H and dt_max to round-off (the same sums in the same order);
H(t+dt) to 1e-9 of the largest thickness when both sides solve to a relative
residual of 1e-12, and to 1e-3 m at the reference's own tolerances (the two
sides stop on different residual norms and preconditioners, so only the
tolerance is shared there)."""

import numpy as np
import pytest

from conftest import max_rel

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-12, abstol=0.0)


def _both(V, OB, M, w, inp, **kw):
    s = V.VelocitySolver(w.descriptor())
    g = s.thickness_setup()
    gpu = s.calc_dHi_dt_semiimplicit(**inp, **kw)
    s.close()
    rc, rate, Hn, divQ, dt_max, its = OB.thickness_semiimplicit(w.mesh, g, **inp, **kw)
    assert rc == 0
    return gpu, (rate, Hn, divQ, dt_max, its)


def _check(gpu, ora, Hmax, tol_H):
    rate, Hn, divQ, dt_max, its = gpu
    rate_o, Hn_o, divQ_o, dt_max_o, _ = ora
    assert max_rel(divQ, divQ_o) < 1e-12
    assert dt_max == dt_max_o
    assert np.abs(Hn - Hn_o).max() <= tol_H * (Hmax if tol_H < 1e-6 else 1.0)
    assert its > 0
    assert (Hn >= -1e-6).all()


@pytest.fixture(scope="module")
def channel_case(W):
    w = W.channel_small(target_nV=1500, stress_balance="DIVA")
    f = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                           "oracle_channel_diva.npz"))
    assert str(f["digest"]) == w.mesh.digest()
    return w, W.thickness_step_inputs(w, f["u"], f["v"])


@pytest.fixture(scope="module")
def continental_case(W):
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_continental_small.npz"))
    w = W.continental(target_nV=int(f["target_nV"]))
    assert str(f["digest"]) == w.mesh.digest()
    return w, W.thickness_step_inputs(w, f["u"], f["v"], dt=2.0)


@pytest.mark.parametrize("krylov", [1, 0])
@pytest.mark.parametrize("bc", ["zero", "infinite"])
def test_channel_step_matches_oracle(V, OB, channel_case, bc, krylov):
    from ufemism2_amd import mesh as M
    w, inp = channel_case
    print(f"channel nV={w.mesh.nV} bc={bc} krylov={krylov}", flush=True)
    gpu, ora = _both(V, OB, M, w, inp, bc_H=[bc] * 4, krylov=krylov, **TIGHT)
    _check(gpu, ora, inp["Hi"].max(), 1e-9)
    assert np.abs(gpu[1] - inp["Hi"]).max() > 1.0   # the step moves the ice


def test_continental_step_matches_oracle(V, OB, continental_case):
    from ufemism2_amd import mesh as M
    w, inp = continental_case
    print(f"continental nV={w.mesh.nV}", flush=True)
    gpu, ora = _both(V, OB, M, w, inp, bc_H=["infinite"] * 4, **TIGHT)
    _check(gpu, ora, inp["Hi"].max(), 1e-9)


def test_reference_tolerances(V, OB, continental_case):
    from ufemism2_amd import mesh as M
    w, inp = continental_case
    gpu, ora = _both(V, OB, M, w, inp)
    _check(gpu, ora, inp["Hi"].max(), 1e-3)


def test_continental_benchmark_mesh_synthetic_flow(V, OB, W):
    """The 400k-vertex continental mesh with a radial outflow of up to
    1 km/yr (no velocity solve): the step's kernels at their full size."""
    from ufemism2_amd import mesh as M
    w = W.continental()
    c = w.mesh.V[w.mesh.Tri].mean(axis=1)
    ctr = 0.5 * (w.mesh.V.min(axis=0) + w.mesh.V.max(axis=0))
    d = c - ctr
    r = np.hypot(d[:, 0], d[:, 1])
    sp = 1000.0 * np.clip(r / r.max(), 0.0, 1.0) ** 2
    u, v = sp * d[:, 0] / np.maximum(r, 1.0), sp * d[:, 1] / np.maximum(r, 1.0)
    inp = W.thickness_step_inputs(w, u, v, dt=1.0)
    print(f"continental full nV={w.mesh.nV}", flush=True)
    gpu, ora = _both(V, OB, M, w, inp, bc_H=["infinite"] * 4, **TIGHT)
    _check(gpu, ora, inp["Hi"].max(), 1e-9)


def test_prescribed_noice_and_guess(V, OB, channel_case):
    from ufemism2_amd import mesh as M
    w, inp = channel_case
    rng = np.random.default_rng(5)
    nV = w.mesh.nV
    pm = (rng.random(nV) < 0.05).astype(np.int32)
    ph = rng.uniform(0.0, 800.0, nV)
    noice = ((rng.random(nV) < 0.03) & (pm == 0)).astype(np.int32)
    kw = dict(inp, mask_noice=noice)
    gpu, ora = _both(V, OB, M, w, kw, BC_prescr_mask=pm, BC_prescr_Hi=ph,
                     dHi_dt_guess=rng.normal(0.0, 20.0, nV), bc_H=["infinite"] * 4, fs=1.0, **TIGHT)
    _check(gpu, ora, inp["Hi"].max(), 1e-9)
    Hn = gpu[1]
    sel = (pm == 1) & (w.mesh.VBI == 0)
    assert np.array_equal(Hn[sel], ph[sel])
    assert (Hn[noice == 1] == 0.0).all()


def test_context_velocity_and_errors(V, channel_case):
    """Velocities omitted: the step uses the solver's current velocity (bit
    for bit the same as passing it); argument errors are reported."""
    w, inp = channel_case
    s = V.VelocitySolver(w.descriptor())
    s.thickness_setup()
    s.set_velocity(inp["u_vav_b"], inp["v_vav_b"])
    a = s.calc_dHi_dt_semiimplicit(**inp)
    b = s.calc_dHi_dt_semiimplicit(**dict(inp, u_vav_b=None, v_vav_b=None))
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    with pytest.raises(V.VelocitySolverError):
        s.calc_dHi_dt_semiimplicit(**dict(inp, dt=0.0))
    with pytest.raises(V.VelocitySolverError):
        s.calc_dHi_dt_semiimplicit(**inp, BC_prescr_mask=np.zeros(w.mesh.nV, np.int32))
    with pytest.raises(V.VelocitySolverError):
        s.calc_dHi_dt_semiimplicit(**inp, gmres_restart=65)
    s.close()


def test_odd_vertex_count(V, OB, W):
    """Vertices are paired into 2x2 blocks for the solve; an odd count gets a
    padding unknown that must not change the result."""
    from ufemism2_amd import mesh as M
    from ufemism2_amd.operators import MeshDescriptor, build_operators
    for h in (9e3, 8.5e3, 8e3, 7.5e3):
        m = M.uniform_mesh(0.0, 150e3, 0.0, 100e3, h, seed=3)
        if m.nV % 2 == 1:
            break
    assert m.nV % 2 == 1
    nV, nT = m.nV, m.nT
    rng = np.random.default_rng(6)

    class Wm:
        mesh = m

        @staticmethod
        def descriptor():
            return MeshDescriptor(m, build_operators(m), np.linspace(0, 1, 5))

    x = m.V[:, 0]
    inp = dict(Hi=500.0 + 300.0 * np.sin(x / 2e4), Hb=np.full(nV, 50.0), SL=np.zeros(nV),
               u_vav_b=rng.normal(200.0, 50.0, nT), v_vav_b=rng.normal(0.0, 50.0, nT), SMB=np.full(nV, 0.5),
               BMB=np.zeros(nV), LMB=np.zeros(nV), fraction_margin=np.ones(nV),
               mask_noice=np.zeros(nV, np.int32), dt=3.0)
    gpu, ora = _both(V, OB, M, Wm, inp, bc_H=["infinite", "zero", "infinite", "zero"], **TIGHT)
    _check(gpu, ora, inp["Hi"].max(), 1e-9)
