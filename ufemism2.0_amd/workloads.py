"""Seeded synthetic workloads for the velocity solve and the heat columns.

Each builder returns a ``Workload`` with the mesh, its operators, the zeta
grid and the per-vertex / per-triangle inputs of solve_SSA / solve_DIVA
(Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask) in the engine's layout.

Workloads (sizes are the targets; the seed and mesh digest identify a run):
  dome        parabolic dome, 2.5k vertices, SSA, Weertman m=3, beta_sq=1e4
  channel     marine channel 640x80 km, bed -100 -> -700 m, H 1500 -> 200 m
              plus a Gaussian random field, ~200k vertices, DIVA nz=10
  continental fractal bed, plastic ice profile, ~400k vertices, nz=12
  ensemble    members of the channel geometry with A and beta_sq drawn
              log-uniformly, uneven resolutions

Geometry relations follow ice_geometry_basics.f90 (is_floating :12-26,
ice_surface_elevation :28-41, thickness_above_floatation :43-56) with
rho_i = 910, rho_sw = 1028 (parameters.f90:23-25).
"""

from __future__ import annotations

import dataclasses

import numpy as np

from .mesh import Mesh, generate_mesh, uniform_mesh
from .operators import MeshDescriptor, MeshOperators, build_operators

RHO_I = 910.0
RHO_SW = 1028.0
SEC_PER_YEAR = 31556943.36


def zeta_regular(nz: int) -> np.ndarray:
    """initialise_scaled_vertical_coordinate_regular (mesh_zeta.f90:58-84)."""
    return np.arange(nz, dtype=np.float64) / (nz - 1)


def surface_elevation(Hi, Hb, SL):
    return Hi + np.maximum(SL - RHO_I / RHO_SW * Hi, Hb)


def thickness_above_floatation(Hi, Hb, SL):
    return Hi - np.maximum(0.0, (SL - Hb) * (RHO_SW / RHO_I))


def is_floating(Hi, Hb, SL):
    return Hi < (SL - Hb) * RHO_SW / RHO_I


def grounded_fraction_b(mesh: Mesh, TAF: np.ndarray) -> np.ndarray:
    """Sub-grid grounded fraction per triangle from vertex TAF (input preparation;
    same relations as subgrid_grounded_fractions_bilin_TAF.f90:182-370)."""
    T = TAF[mesh.Tri]
    P = mesh.V[mesh.Tri]
    tol = 1e-9
    Tc = np.where(T >= 0, np.maximum(tol, T), np.minimum(-tol, T))

    def area(p, q, r):
        return np.abs((q[..., 0] - p[..., 0]) * (r[..., 1] - p[..., 1]) -
                      (q[..., 1] - p[..., 1]) * (r[..., 0] - p[..., 0])) / 2.0

    At = area(P[:, 0], P[:, 1], P[:, 2])
    out = np.where(T.min(axis=1) >= 0, 1.0, 0.0)
    mixed = (T.min(axis=1) < 0) & (T.max(axis=1) > 0)
    for i in np.nonzero(mixed)[0]:
        a, b, c = Tc[i]
        va, vb, vc = P[i]
        grd = np.array([a, b, c]) >= 0
        # rotate so that the odd one out is first
        odd = int(np.nonzero(grd != (grd.sum() >= 2))[0][0])
        idx = [(odd + k) % 3 for k in range(3)]
        t = Tc[i][idx]
        p = P[i][idx]
        lab = t[0] / (t[0] - t[1])
        lac = t[0] / (t[0] - t[2])
        pab = p[0] * (1 - lab) + p[1] * lab
        pac = p[0] * (1 - lac) + p[2] * lac
        sub = area(p[0], pab, pac)
        Ag = sub if grd[odd] else At[i] - sub
        out[i] = min(1.0, max(0.0, Ag / At[i]))
    return out


@dataclasses.dataclass
class Workload:
    name: str
    mesh: Mesh
    ops: MeshOperators
    zeta: np.ndarray
    Hi: np.ndarray
    Hb: np.ndarray
    SL: np.ndarray
    A_flow: np.ndarray        # (nz*nV,) vertex fastest
    beta_sq: np.ndarray
    seed: int
    stress_balance: str = "SSA"
    params: dict = dataclasses.field(default_factory=dict)
    _md: MeshDescriptor = None

    @property
    def Hs(self):
        return surface_elevation(self.Hi, self.Hb, self.SL)

    @property
    def fgr_b(self):
        return grounded_fraction_b(self.mesh, thickness_above_floatation(self.Hi, self.Hb, self.SL))

    @property
    def mask_grounded(self):
        return ((self.Hi > 0) & ~is_floating(self.Hi, self.Hb, self.SL)).astype(np.int32)

    @property
    def nz(self):
        return int(len(self.zeta))

    def descriptor(self) -> MeshDescriptor:
        if self._md is None:
            self._md = MeshDescriptor(self.mesh, self.ops, self.zeta)
        return self._md

    def inputs(self) -> dict:
        """Arrays in the engine/oracle layout, all contiguous."""
        d = {"Hi": np.ascontiguousarray(self.Hi), "Hs": np.ascontiguousarray(self.Hs),
             "A_flow": np.ascontiguousarray(self.A_flow), "beta_sq": np.ascontiguousarray(self.beta_sq),
             "fgr_b": np.ascontiguousarray(self.fgr_b), "mask_grounded": self.mask_grounded}
        return d

    def describe(self) -> dict:
        return {"workload": self.name, "nV": self.mesh.nV, "nTri": self.mesh.nT, "nz": self.nz,
                "unknowns": 2 * self.mesh.nT, "seed": self.seed, "mesh": self.mesh.digest(),
                "stress_balance": self.stress_balance}


def _grf(rng, xmin, xmax, ymin, ymax, corr, std, dx):
    """Gaussian random field with Gaussian correlation length `corr`, sampled on a grid."""
    nx = int(np.ceil((xmax - xmin) / dx)) + 1
    ny = int(np.ceil((ymax - ymin) / dx)) + 1
    px, py = 2 * nx, 2 * ny  # pad to limit wrap-around correlation
    w = rng.standard_normal((px, py))
    kx = np.fft.fftfreq(px, d=dx) * 2 * np.pi
    ky = np.fft.fftfreq(py, d=dx) * 2 * np.pi
    K2 = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-K2 * corr ** 2 / 8.0)))[:nx, :ny]
    f = (f - f.mean()) / f.std() * std
    gx = xmin + dx * np.arange(nx)
    gy = ymin + dx * np.arange(ny)
    return gx, gy, f


def _sample(gx, gy, f, x, y):
    from scipy.interpolate import RegularGridInterpolator
    return RegularGridInterpolator((gx, gy), f, bounds_error=False, fill_value=None)(np.stack([x, y], 1))


def dome(seed: int = 1, h: float = 36e3, nz: int = 10) -> Workload:
    """Parabolic dome H = 3000 sqrt(1 - (r/750 km)^2) on a flat bed at 0 m."""
    mesh = uniform_mesh(-900e3, 900e3, -900e3, 900e3, h, seed=seed)
    r = np.hypot(mesh.V[:, 0], mesh.V[:, 1])
    Hi = 3000.0 * np.sqrt(np.clip(1.0 - (r / 750e3) ** 2, 0.0, None))
    nV = mesh.nV
    return Workload("dome_ssa", mesh, build_operators(mesh), zeta_regular(nz), Hi, np.zeros(nV),
                    np.zeros(nV), np.full(nz * nV, 1e-16), np.full(nV, 1e4), seed, "SSA",
                    {"visc_it_norm_dUV_tol": 1e-3, "rtol": 1e-8})


def channel_geometry(x, y, rng_field):
    gx, gy, f = rng_field
    Hb = -100.0 - 600.0 * x / 640e3
    Hi = 1500.0 - 1300.0 * x / 640e3 + _sample(gx, gy, f, x, y)
    return np.maximum(Hi, 0.0), Hb


def channel_gl_x():
    """Grounding-line position of the mean channel profile (no random field)."""
    x = np.linspace(0, 640e3, 64001)
    H = 1500.0 - 1300.0 * x / 640e3
    b = -100.0 - 600.0 * x / 640e3
    fl = H < -b * RHO_SW / RHO_I
    return float(x[np.argmax(fl)])


def channel(seed: int = 2, target_nV: int = 200_000, nz: int = 10, A: float = 1e-16,
            beta_sq: float = 1e4, stress_balance: str = "DIVA") -> Workload:
    """Marine channel: 640 x 80 km, refined towards the grounding line.

    The size field grows linearly away from the grounding line with slope 0.05
    between h_min and h_max = 2.5 h_min; h_min is scaled so that the mesh has
    about target_nV vertices (a 16 -> 2 km field on this domain gives only
    ~15k vertices, so the resolution is scaled to reach the stated size).
    """
    rng = np.random.default_rng(seed)
    field = _grf(rng, 0, 640e3, 0, 80e3, 20e3, 50.0, 1e3)
    x_gl = channel_gl_x()
    # vertices ~ area / h^2 for the square-cell quadtree; h_max dominates
    h_max = np.sqrt(640e3 * 80e3 / target_nV) * 1.03
    h_min = h_max / 2.5

    def size(x, y):
        return np.minimum(h_max, h_min + 0.05 * np.abs(x - x_gl))

    mesh = generate_mesh(0.0, 640e3, 0.0, 80e3, size, seed=seed)
    Hi, Hb = channel_geometry(mesh.V[:, 0], mesh.V[:, 1], field)
    nV = mesh.nV
    return Workload(f"channel_{stress_balance.lower()}", mesh, build_operators(mesh), zeta_regular(nz),
                    Hi, Hb, np.zeros(nV), np.full(nz * nV, A), np.full(nV, beta_sq), seed, stress_balance,
                    {"visc_it_norm_dUV_tol": 1e-3})


def continental(seed: int = 3, target_nV: int = 400_000, nz: int = 12, sea_level: float = -1e4,
                A: float = 1e-17, beta_sq: float = 1e5) -> Workload:
    """Fractal bed (spectral slope -2) rescaled to [-2000, 2000] m and a
    perfectly plastic ice profile (yield stress 100 kPa) clipped to the bed.
    Sea level defaults to -10 km, so the sheet and the ice-free land around it
    are grounded: with an ocean, the ice-free ocean triangles have no friction
    and a viscosity floor of 0.1 m of ice, and their velocities run to the
    velocity limit.  A = 1e-17 and beta_sq = 1e5 keep the interior speeds
    at tens to hundreds of m/yr under the 100 kPa plastic driving stress."""
    rng = np.random.default_rng(seed)
    L = 2000e3
    n = 512
    dx = 2 * L / (n - 1)
    kx = np.fft.fftfreq(n, d=dx)
    K = np.hypot(kx[:, None], kx[None, :])
    K[0, 0] = np.inf
    amp = K ** (-1.0)  # power ~ k^-2
    spec = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * amp
    bed = np.real(np.fft.ifft2(spec))
    bed = -2000.0 + 4000.0 * (bed - bed.min()) / (bed.max() - bed.min())
    g = -L + dx * np.arange(n)
    R = 1500e3
    tau_y = 1e5
    h_fine = np.sqrt((2 * L) ** 2 / target_nV) * 0.29
    h_coarse = 4.0 * h_fine

    def size(x, y):
        d = np.abs(R - np.hypot(x, y))
        return np.minimum(h_coarse, h_fine + 0.08 * d)

    mesh = generate_mesh(-L, L, -L, L, size, seed=seed)
    x, y = mesh.V[:, 0], mesh.V[:, 1]
    Hb = _sample(g, g, bed, x, y)
    d = np.clip(R - np.hypot(x, y), 0.0, None)
    Hs_pl = np.sqrt(2.0 * tau_y / (RHO_I * 9.81) * d)
    Hi = np.where(d > 0, np.maximum(0.0, Hs_pl - np.maximum(Hb, 0.0)), 0.0)
    nV = mesh.nV
    return Workload("continental_diva", mesh, build_operators(mesh), zeta_regular(nz), Hi, Hb,
                    np.full(nV, sea_level), np.full(nz * nV, A), np.full(nV, beta_sq), seed, "DIVA",
                    {"visc_it_norm_dUV_tol": 1e-3})


def ensemble_member(k: int, seed: int = 4, nV_range=(20_000, 120_000), nz: int = 10,
                    stress_balance: str = "DIVA") -> Workload:
    """Member k of the ensemble: channel geometry, uneven resolution, A and
    beta_sq drawn log-uniformly from [1e-17, 1e-16] and [1e3, 1e5]."""
    rng = np.random.default_rng([seed, k])
    nV = int(rng.integers(nV_range[0], nV_range[1] + 1))
    A = float(10 ** rng.uniform(-17, -16))
    beta = float(10 ** rng.uniform(3, 5))
    w = channel(seed=int(rng.integers(1 << 30)), target_nV=nV, nz=nz, A=A, beta_sq=beta,
                stress_balance=stress_balance)
    w.name = f"ensemble_member_{k}"
    return w


def channel_small(seed: int = 2, target_nV: int = 6000, nz: int = 10, stress_balance="SSA"):
    """A few-thousand-vertex channel for tests the oracle finishes in seconds."""
    return channel(seed=seed, target_nV=target_nV, nz=nz, stress_balance=stress_balance)


def schoof_icestream(h: float = 16e3, seed: int = 5, nz: int = 10, A: float = 1e-18, H: float = 2000.0,
                     dhdx: float = -3e-4, L: float = 150e3, m: float = 1.0) -> Workload:
    """The reference's SSA_icestream integrated test: a 2 km slab on a bed
    inclined along x over [-400, 400] km^2, Glen A = 1e-18, and a till yield
    stress tau_y = f |y/L|^m with f = -rho_i g H dhdx
    (idealised_geometries.f90:186-212, sliding_laws.f90:450-480,
    SSA_icestream/config_0*.cfg).  beta_sq carries tau_y; with the Coulomb
    law (beta = tau_y / sqrt(delta_v^2 + |u|^2)) this is the reference's
    'idealised' sliding law for this experiment.  The ice is grounded
    everywhere (sea level -10 km)."""
    mesh = uniform_mesh(-400e3, 400e3, -400e3, 400e3, h, seed=seed)
    nV = mesh.nV
    x, y = mesh.V[:, 0], mesh.V[:, 1]
    f = -RHO_I * 9.81 * H * dhdx
    tau_y = f * np.abs(y / L) ** m
    # the reference's border conditions (SSA_icestream/config_0*.cfg): u 'zero'
    # north/south and 'infinite_SSA_icestream' east/west (the lagged copy of
    # the interior velocity at the same y); v 'zero'
    params = {"sliding_law": 3, "visc_it_relax": 0.3, "slid_delta_v": 1e-3, "epsilon_sq_0": 1e-10,
              "bc_u": [1, 3, 1, 3], "bc_v": [1, 1, 1, 1]}
    return Workload(f"schoof_icestream_{int(h / 1e3)}km", mesh, build_operators(mesh), zeta_regular(nz),
                    np.full(nV, H), dhdx * x, np.full(nV, -10000.0), np.full(nz * nV, A), tau_y, seed, "SSA",
                    params)


def ismip_hom_friction(mesh: Mesh, experiment: str, L: float, A_uniform: float = 1e-16) -> np.ndarray:
    """beta of the 'idealised' sliding laws of ISMIP-HOM C, D and F
    (sliding_laws.f90:499-571) at the vertices: 1000 + 1000 sin(2 pi x / L)
    sin(2 pi y / L), 1000 + 1000 sin(2 pi x / L), and 1 / (1000 A).  These do
    not depend on the velocity, so they are the Weertman law with
    slid_weertman_m = 1 and this field as beta_sq (uabs^0 = 1 exactly);
    slid_beta_max still applies."""
    x, y = mesh.V[:, 0], mesh.V[:, 1]
    if experiment == "ISMIP-HOM_C":
        return 1000.0 + 1000.0 * np.sin(2.0 * np.pi * x / L) * np.sin(2.0 * np.pi * y / L)
    if experiment == "ISMIP-HOM_D":
        return 1000.0 + 1000.0 * np.sin(2.0 * np.pi * x / L)
    if experiment == "ISMIP-HOM_F":
        return np.full(mesh.nV, 1.0 / (A_uniform * 1000.0))
    raise ValueError(f"no idealised sliding law for {experiment!r} (ISMIP-HOM_E is not implemented in the reference)")


def ismip_hom_a(L: float = 160e3, h: float = 8e3, nz: int = 12, seed: int = 8) -> Workload:
    """ISMIP-HOM experiment A, the reference's BPA integrated test
    (automated_testing/integrated_tests/idealised/ISMIP-HOM/config_ISMIP_HOM_A_*_BPA.cfg,
    idealised_geometries.f90:243-266): Hs = 2000 - x tan(0.5 deg), Hi = 1000 -
    500 sin(2 pi x / L) sin(2 pi y / L) on [-L, L]^2, no sliding, periodic
    borders, Glen n = 3 with A = 1e-16 and epsilon_sq_0 = 1e-12, nz = 12
    regular layers, viscosity relaxation 0.4 to 5e-7, Krylov rtol 1e-6 and
    abstol 1e-4; the reference runs it on an 8 km mesh."""
    mesh = uniform_mesh(-L, L, -L, L, h, seed=seed)
    nV = mesh.nV
    x, y = mesh.V[:, 0], mesh.V[:, 1]
    Hs = 2000.0 - x * np.tan(0.5 * np.pi / 180.0)
    Hb = Hs - 1000.0 + 500.0 * np.sin(x * 2.0 * np.pi / L) * np.sin(y * 2.0 * np.pi / L)
    params = {"glens_n": 3.0, "epsilon_sq_0": 1e-12, "sliding_law": 0, "visc_it_relax": 0.4,
              "visc_it_norm_dUV_tol": 5e-7, "visc_it_nit": 5000, "rtol": 1e-6, "abstol": 1e-4,
              "bc_u": [2, 2, 2, 2], "bc_v": [2, 2, 2, 2], "ismip_hom_L": L}
    return Workload(f"ismip_hom_a_{int(L / 1e3)}km_{h / 1e3:g}km", mesh, build_operators(mesh), zeta_regular(nz),
                    Hs - Hb, Hb, np.full(nV, -10000.0), np.full(nz * nV, 1e-16), np.zeros(nV), seed, "BPA", params)


def heat_columns_inputs(nV: int, nz: int = 12, seed: int = 6, dt: float = 10.0, gl: float = 0.0) -> dict:
    """Synthetic inputs of the vertical heat columns for nV vertices, in the
    reference's [k][vi] layer-major layout (vertex fastest): a cold surface,
    a near-steady geothermal profile, horizontal velocities decaying with
    depth, upwind advection terms, strain heating, 10 % floating columns and
    5 % thin-ice columns (thermodynamics_3D_heat_equation.f90:47-160 inputs).
    With gl > 0 that share of the grounded columns is put on the grounding
    line (mask 3) with a uniform random fraction_gr.
    Units: K, m/yr, J/m^2/yr (Q), J/m^3/yr (Phi)."""
    rng = np.random.default_rng(seed)
    zeta = zeta_regular(nz)
    H = rng.uniform(200.0, 3500.0, nV)
    thin = rng.random(nV) < 0.05
    H[thin] = rng.uniform(0.0, 9.0, thin.sum())
    Ts = rng.uniform(225.0, 262.0, nV)
    Q = rng.uniform(1.2e6, 2.2e6, nV)
    K_s = 3.101e8 * np.exp(-0.0057 * Ts)
    Tpmp_b = 273.16 - 8.7e-4 * H
    Tb = np.minimum(Ts + Q * H / K_s * 0.8, Tpmp_b - 0.5)
    z = zeta[:, None]
    Ti = Ts[None, :] + (Tb - Ts)[None, :] * z ** 1.2
    Ti = np.minimum(Ti, 273.16 - 8.7e-4 * H[None, :] * z)
    us = rng.uniform(0.0, 400.0, nV)
    u3 = us[None, :] * (1.0 - 0.6 * z ** 4) * np.cos(rng.uniform(0, 2 * np.pi, nV))[None, :]
    v3 = us[None, :] * (1.0 - 0.6 * z ** 4) * np.sin(rng.uniform(0, 2 * np.pi, nV))[None, :]
    w3 = -rng.uniform(0.0, 0.3, nV)[None, :] * (1.0 - z)
    Hs = np.maximum(H, 1.0)
    f = {
        "zeta": zeta, "Ti": Ti, "u_3D": u3, "v_3D": v3, "w_3D": w3,
        "u_dTdxp_upwind": 1e-3 * rng.normal(size=(nz, nV)),
        "v_dTdyp_upwind": 1e-3 * rng.normal(size=(nz, nV)),
        "dzeta_dx": np.broadcast_to(rng.normal(0, 1e-3, nV) / Hs, (nz, nV)),
        "dzeta_dy": np.broadcast_to(rng.normal(0, 1e-3, nV) / Hs, (nz, nV)),
        "dzeta_dz": np.broadcast_to(-1.0 / Hs, (nz, nV)),
        "dzeta_dt": np.broadcast_to(rng.normal(0, 1e-5, nV), (nz, nV)),
        "Phi": rng.uniform(0.0, 2e4, (nz, nV)) * z ** 3,
        "Hi_eff": H, "T_surf": Ts, "Q_base_grnd": Q,
        "mask": np.where(rng.random(nV) < 0.1, 2, 1).astype(np.int32),
    }
    out = {k: np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1)) for k, v in f.items()
           if k not in ("mask", "zeta")}
    out["mask"] = f["mask"]
    if gl > 0.0:
        rg = np.random.default_rng(seed + 7919)
        on_gl = (out["mask"] == 1) & (rg.random(nV) < gl)
        out["mask"] = np.where(on_gl, 3, out["mask"]).astype(np.int32)
        out["fraction_gr"] = rg.uniform(0.0, 1.0, nV)
    return {"nV": nV, "nz": nz, "zeta": zeta, "fields": out, "dt": dt}


def thickness_step_inputs(w: Workload, u_vav_b: np.ndarray, v_vav_b: np.ndarray, seed: int = 8,
                          dt: float = 5.0) -> dict:
    """Inputs of one semi-implicit thickness step (calc_dHi_dt_semiimplicit,
    conservation_of_mass_semiimplicit.f90:23-40) on a workload: its Hi, Hb,
    SL, the given vertically averaged velocities, a surface mass balance of
    +0.3 m/yr on thick ice falling to -2 m/yr on thin ice, basal melt of up to
    5 m/yr under floating ice, no lateral mass balance, fraction_margin 1
    where a vertex and all its neighbours carry ice, 0.5 on the ice margin
    and 0 without ice (the reference's fraction of the cell covered), and no
    target rate.  Units: m, m/yr, yr."""
    rng = np.random.default_rng(seed)
    m = w.mesh
    nV = m.nV
    Hi = np.ascontiguousarray(w.Hi, dtype=np.float64)
    ice = Hi > 0.0
    nb_ice = np.ones(nV, bool)
    src = np.repeat(np.arange(nV), np.diff(m.C_ptr))
    np.logical_and.at(nb_ice, src, ice[m.C_ind])
    fm = np.where(ice & nb_ice, 1.0, np.where(ice, 0.5, 0.0))
    SMB = 0.3 - 2.3 * np.exp(-Hi / 300.0) + rng.normal(0.0, 0.05, nV)
    BMB = np.where(is_floating(Hi, w.Hb, w.SL) & ice, -rng.uniform(0.0, 5.0, nV), 0.0)
    return {"Hi": Hi, "Hb": np.ascontiguousarray(w.Hb, dtype=np.float64),
            "SL": np.ascontiguousarray(w.SL, dtype=np.float64),
            "u_vav_b": np.ascontiguousarray(u_vav_b, dtype=np.float64),
            "v_vav_b": np.ascontiguousarray(v_vav_b, dtype=np.float64),
            "SMB": SMB, "BMB": BMB, "LMB": np.zeros(nV), "fraction_margin": fm,
            "mask_noice": np.zeros(nV, np.int32), "dt": float(dt)}
