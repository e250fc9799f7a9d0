"""Pins of the Tsai2015, Schoof2005 and Zoet-Iverson sliding laws of the oracle
(sliding_laws.f90:244-403) against the Weertman, Coulomb and Budd laws, which
the closed-form tests of the assembly pin on their own:

  * Tsai2015 is min(tau_Weertman, alpha_sq N) / u: with the cap far above the
    Weertman stress everywhere it is Weertman, far below it is Coulomb with
    the cap as till yield stress (a cap that differs per vertex, so a wrong
    vertex index fails);
  * Schoof2005 obeys tau^-m = tau_Weertman^-m + (alpha_sq N)^-m at every
    vertex (exact at uniform velocity and cap, where beta_b = beta_a);
  * Zoet-Iverson tau = tau_c (u / (u + u_t))^(1/p): u_t = 0 gives Coulomb,
    p = 1 with u_t = |u| half of Coulomb, u_t >> |u| Budd with q = 1/p and
    threshold u_t;

and of the host routine that forms the till yield stress of the yield-stress
laws (N tan(pi/180) phi, extended to ice-free land, :643-700) on a hand-made
star of vertices.
"""

import numpy as np
import pytest

import manufactured as MF
from conftest import params


@pytest.fixture(scope="module")
def case(W):
    return MF.linear_case(W)


U0, V0, DV = 50.0, -20.0, 1e-3
UABS = np.sqrt(DV * DV + U0 * U0 + V0 * V0)


def beta_b(L, OB, w, rough=None, cap=None, U0=U0, V0=V0, **kw):
    """beta_b on the free rows at the uniform velocity (U0, V0): the free rows
    of the assembled system reduce to -beta_b u there."""
    p = params(L, glens_n=3.0, epsilon_sq_0=1e-8, slid_delta_v=DV, **kw)
    inp = dict(w.inputs())
    if rough is not None:
        inp["beta_sq"] = L.f64(rough)
    nT = w.mesh.nT
    u, v = np.full(nT, U0), np.full(nT, V0)
    if cap is not None:
        OB.set_sliding_cap(np.full(w.mesh.nV, 0.5), 2.0 * np.asarray(cap))
    try:
        A_, _ = OB.Oracle(w.descriptor()).system(p, inp, u, v)
    finally:
        OB.set_sliding_cap()
    x = np.empty(2 * nT); x[0::2] = u; x[1::2] = v
    Ax = A_ @ x
    fr = np.nonzero(w.mesh.TriBI == 0)[0]
    bu, bv = -Ax[2 * fr] / U0, -Ax[2 * fr + 1] / V0
    assert np.allclose(bu, bv, rtol=1e-9)
    return bu


def test_tsai_limits(L, OB, case):
    w = case
    nV = w.mesh.nV
    rng = np.random.default_rng(3)
    m = 3.0
    tau_w = 1e3 * UABS ** (1 / m)                       # Weertman stress at beta_sq = 1e3
    wee = beta_b(L, OB, w, sliding_law=L.SLIDING["Weertman"], slid_weertman_m=m)
    big = beta_b(L, OB, w, cap=np.full(nV, 1e6 * tau_w), sliding_law=L.SLIDING["Tsai2015"], slid_weertman_m=m)
    assert np.allclose(big, wee, rtol=1e-12)
    cap = tau_w * rng.uniform(1e-4, 1e-2, nV)         # far below: the cap decides, per vertex
    cou = beta_b(L, OB, w, rough=cap, sliding_law=L.SLIDING["Coulomb"])
    small = beta_b(L, OB, w, cap=cap, sliding_law=L.SLIDING["Tsai2015"], slid_weertman_m=m)
    assert np.allclose(small, cou, rtol=1e-12)
    # in between: uniform cap at 0.6 tau_W picks the cap, at 1.7 tau_W Weertman
    mid = beta_b(L, OB, w, cap=np.full(nV, 0.6 * tau_w), sliding_law=L.SLIDING["Tsai2015"], slid_weertman_m=m)
    assert np.allclose(mid, 0.6 * wee, rtol=1e-12)
    hi = beta_b(L, OB, w, cap=np.full(nV, 1.7 * tau_w), sliding_law=L.SLIDING["Tsai2015"], slid_weertman_m=m)
    assert np.allclose(hi, wee, rtol=1e-12)


@pytest.mark.parametrize("m", [3.0, 1.0, 5.0])
def test_schoof_harmonic_identity(L, OB, case, m):
    w = case
    nV = w.mesh.nV
    wee = beta_b(L, OB, w, sliding_law=L.SLIDING["Weertman"], slid_weertman_m=m)
    tau_w = 1e3 * UABS ** (1 / m)
    for f in (0.3, 1.0, 4.0):
        cap = f * tau_w
        cou = beta_b(L, OB, w, rough=np.full(nV, cap), sliding_law=L.SLIDING["Coulomb"])
        sch = beta_b(L, OB, w, cap=np.full(nV, cap), sliding_law=L.SLIDING["Schoof2005"], slid_weertman_m=m)
        assert np.allclose(sch ** -m, wee ** -m + cou ** -m, rtol=1e-11), f
        assert (sch < np.minimum(wee, cou)).all()


def test_schoof_limits_per_vertex(L, OB, case):
    w = case
    nV = w.mesh.nV
    rng = np.random.default_rng(4)
    m = 3.0
    tau_w = 1e3 * UABS ** (1 / m)
    cap = tau_w * rng.uniform(1e-6, 1e-5, nV)
    cou = beta_b(L, OB, w, rough=cap, sliding_law=L.SLIDING["Coulomb"])
    sch = beta_b(L, OB, w, cap=cap, sliding_law=L.SLIDING["Schoof2005"], slid_weertman_m=m)
    assert np.allclose(sch, cou, rtol=1e-12)
    wee = beta_b(L, OB, w, sliding_law=L.SLIDING["Weertman"], slid_weertman_m=m)
    sch = beta_b(L, OB, w, cap=tau_w * rng.uniform(1e5, 1e6, nV), sliding_law=L.SLIDING["Schoof2005"],
                 slid_weertman_m=m)
    assert np.allclose(sch, wee, rtol=1e-12)


def test_zoet_iverson_limits(L, OB, case):
    w = case
    nV = w.mesh.nV
    rng = np.random.default_rng(5)
    tau_c = rng.uniform(1e4, 1e5, nV)
    cou = beta_b(L, OB, w, rough=tau_c, sliding_law=L.SLIDING["Coulomb"])
    zi = L.SLIDING["Zoet-Iverson"]
    z0 = beta_b(L, OB, w, rough=tau_c, sliding_law=zi, slid_zi_p=5.0, slid_zi_ut=0.0)
    assert np.allclose(z0, cou, rtol=1e-13)
    z1 = beta_b(L, OB, w, rough=tau_c, sliding_law=zi, slid_zi_p=1.0, slid_zi_ut=UABS)
    assert np.allclose(z1, 0.5 * cou, rtol=1e-13)
    p, ut = 4.0, 1e10
    budd = beta_b(L, OB, w, rough=tau_c, sliding_law=L.SLIDING["Budd"], slid_budd_q_plastic=1 / p,
                  slid_budd_u_threshold=ut)
    zb = beta_b(L, OB, w, rough=tau_c, sliding_law=zi, slid_zi_p=p, slid_zi_ut=ut)
    assert np.allclose(zb, budd, rtol=1e-8)
    # the default exponent and transition velocity lie between those limits
    zd = beta_b(L, OB, w, rough=tau_c, sliding_law=zi, slid_zi_p=5.0, slid_zi_ut=200.0)
    assert np.allclose(zd, cou * (UABS / (UABS + 200.0)) ** 0.2, rtol=1e-12)


def test_ismip_hom_idealised_laws_are_weertman_m1(L, OB, W, case):
    """ISMIP-HOM C, D, F (sliding_laws.f90:499-571): the prescribed field as
    beta_sq with Weertman m = 1 gives beta_b = the vertex mean of the field,
    whatever the velocity."""
    w = case
    fr = np.nonzero(w.mesh.TriBI == 0)[0]
    for exp in ("ISMIP-HOM_C", "ISMIP-HOM_D", "ISMIP-HOM_F"):
        beta = W.ismip_hom_friction(w.mesh, exp, 400e3)
        want = beta[w.mesh.Tri].mean(axis=1)[fr]
        for u0, v0 in ((50.0, -20.0), (3.0, 700.0)):
            got = beta_b(L, OB, w, rough=beta, U0=u0, V0=v0, sliding_law=L.SLIDING["Weertman"],
                         slid_weertman_m=1.0)
            assert np.allclose(got, want, rtol=1e-11), exp
    assert np.ptp(W.ismip_hom_friction(w.mesh, "ISMIP-HOM_C", 400e3)) > 1000.0
    with pytest.raises(ValueError):
        W.ismip_hom_friction(w.mesh, "ISMIP-HOM_E", 400e3)


def test_params_defaults(L):
    p = L.Params()
    L.lib().ufe_params_default(p)
    assert (p.slid_zi_p, p.slid_zi_ut) == (5.0, 200.0)


# -- till yield stress --------------------------------------------------------

RHO_G = 910.0 * 9.81


def star():
    """Vertex 0 joined to 1..5; 6 joined to 5 only; 7 alone."""
    nbrs = {0: [1, 2, 3, 4, 5], 1: [0], 2: [0], 3: [0], 4: [0], 5: [0, 6], 6: [5], 7: []}
    ptr = np.zeros(9, np.int32)
    ind = []
    for v in range(8):
        ind += nbrs[v]
        ptr[v + 1] = len(ind)
    return ptr, np.array(ind, np.int32)


class _Mesh:
    def __init__(self):
        self.C_ptr, self.C_ind = star()
        self.nV = 8


def test_till_yield_stress_extension(V):
    m = _Mesh()
    N = np.array([1e6, 2e6, 3e6, 4e6, 5e6, 6e6, 7e6, 8e6])
    phi = np.array([10.0, 20.0, 30.0, 5.0, 40.0, 25.0, 15.0, 35.0])
    grounded = np.array([0, 1, 1, 1, 0, 1, 0, 0])
    land = np.array([1, 0, 0, 0, 0, 0, 1, 1])
    tau = N * np.tan(np.pi / 180.0) * phi
    got = V.till_yield_stress(m, N, phi, grounded, land, Hi_min=3.0)
    want = tau.copy()
    want[0] = min(tau[1], tau[2], tau[3], tau[5], 1000 * RHO_G)   # 4 is not grounded
    want[6] = min(tau[5], 1000 * RHO_G)
    want[7] = 3.0 * RHO_G                                          # no grounded neighbour
    assert np.allclose(got, want, rtol=1e-15)
    assert got[0] == tau[3] and got[6] == tau[5]
    # the 1000 m ceiling
    big = V.till_yield_stress(m, N * 1e6, phi, grounded, land)
    assert big[0] == 1000 * RHO_G and big[7] == 0.0
    # neighbours are read before any vertex is changed: a grounded ice-free
    # land vertex passes on its own unextended value
    land2 = land.copy(); land2[5] = 1
    # (vertex 5 has no grounded neighbour and becomes Hi_min rho g = 0 before
    # vertex 6 reads it)
    got2 = V.till_yield_stress(m, N, phi, grounded, land2)
    assert got2[5] == 0.0 and got2[6] == tau[5] and got2[0] == tau[3]


def test_till_yield_stress_rejects_bad_connectivity(V):
    m = _Mesh()
    m.C_ind = m.C_ind.copy(); m.C_ind[0] = 99
    z = np.zeros(8)
    with pytest.raises(RuntimeError):
        V.till_yield_stress(m, z, z, np.zeros(8, int), np.zeros(8, int))
