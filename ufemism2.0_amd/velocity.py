"""Host-side mirror of UFEMISM's ice-velocity module interface.

``VelocitySolver.solve_SSA`` / ``solve_DIVA`` keep the argument meaning and
error behaviour of solve_SSA (SSA_main.f90:87-242) and solve_DIVA
(DIVA_main.f90:88-262): velocities are inout (initial guess in, solution
out), the optional prescribed-velocity boundary condition needs all three of
mask, u and v, the Picard iteration returns n_visc_its and n_Axb_its, hitting
visc_it_nit is a warning, a diverging iteration is an error.  All compute runs
in libufe_hip.so; there is no CPU path here.
"""

from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import _lib as L
from .operators import MeshDescriptor


class ViscosityIterationWarning(UserWarning):
    pass


class VelocitySolverError(RuntimeError):
    pass


_PARAM_NAMES = [f for f, _ in L.Params._fields_]
# multigrid preconditioner knobs (ufe_set_amg_options) and their library defaults
AMG_DEFAULTS = {"omega": 0.85, "coarse_scale": 1.5, "tail_rows": 300}
# coordinate search of VelocitySolver.autotune: one axis at a time, best kept
AUTOTUNE_AXES = (("amg_smoother", (0, 1)), ("accel_depth", (2, 3)), ("omega", (0.75, 0.85, 0.95)),
                 ("coarse_scale", (1.3, 1.5, 1.7)))


class VelocitySolver:
    """One device context: mesh, operators and solver state in HBM."""

    def __init__(self, md: MeshDescriptor, device: int = 0, **params):
        self.md = md
        self.lib = L.lib()
        self._ctx = C.c_void_p()
        rc = self.lib.ufe_ctx_create(C.byref(self._ctx), device, md.byref())
        if rc != 0:
            raise VelocitySolverError(f"ufe_ctx_create failed ({rc}): {self.lib.ufe_last_error(None).decode()}")
        self.params = L.default_params()
        self.amg_options = dict(AMG_DEFAULTS)
        self.set_params(**params)

    # -- configuration -------------------------------------------------------
    def set_params(self, **kw):
        for k, v in kw.items():
            if k not in _PARAM_NAMES:
                raise KeyError(f"unknown parameter {k}")
            if k in ("bc_u", "bc_v"):
                arr = getattr(self.params, k)
                for i, x in enumerate(v):
                    arr[i] = int(x)
            else:
                setattr(self.params, k, v)
        self._check(self.lib.ufe_set_params(self._ctx, C.byref(self.params)))

    def set_graph_launch(self, enable: bool):
        """Replay chunks of Krylov iterations from captured HIP graphs."""
        self._check(self.lib.ufe_set_graph_launch(self._ctx, int(bool(enable))))

    def _check(self, rc):
        if rc < 0:
            raise VelocitySolverError(f"libufe_hip error {rc}: {self.lib.ufe_last_error(self._ctx).decode()}")
        return rc

    @staticmethod
    def _bc(bc_mask, bc_u, bc_v):
        given = [a is not None for a in (bc_mask, bc_u, bc_v)]
        if any(given) and not all(given):
            raise VelocitySolverError("need to provide prescribed u,v fields and mask!")
        if not any(given):
            return None, None, None
        return L.i32(bc_mask), L.f64(bc_u), L.f64(bc_v)

    # -- device-resident form --------------------------------------------------
    def upload(self, Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice=None,
               BC_prescr_mask_b=None, BC_prescr_u_b=None, BC_prescr_v_b=None):
        bm, bu, bv = self._bc(BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b)
        self._keep = [L.f64(Hi), L.f64(Hs), L.f64(A_flow), L.f64(beta_sq), L.f64(fraction_gr_b),
                      None if mask_grounded_ice is None else L.i32(mask_grounded_ice), bm, bu, bv]
        k = self._keep
        self._check(self.lib.ufe_upload_inputs(self._ctx, L.ptr(k[0]), L.ptr(k[1]), L.ptr(k[2]), L.ptr(k[3]),
                                               L.ptr(k[4]), L.ptr(k[5], L.ip), L.ptr(k[6], L.ip), L.ptr(k[7]),
                                               L.ptr(k[8])))

    def set_velocity(self, u_b, v_b):
        """Initial guess; u_b = v_b = None sets zero velocity on the device."""
        if u_b is None and v_b is None:
            self._check(self.lib.ufe_upload_velocity(self._ctx, None, None))
            return
        self._check(self.lib.ufe_upload_velocity(self._ctx, L.ptr(L.f64(u_b)), L.ptr(L.f64(v_b))))

    def reset_diva_state(self):
        """Zero DIVA solver state (eta_3D_b, tau_b, u_base_b) on the device."""
        self._check(self.lib.ufe_upload_diva_state(self._ctx, None, None, None, None, None))

    def get_velocity(self):
        nT = self.md.mesh.nT
        u = np.zeros(nT)
        v = np.zeros(nT)
        self._check(self.lib.ufe_download_velocity(self._ctx, L.ptr(u), L.ptr(v)))
        return u, v

    def run(self, stress_balance="SSA"):
        info = L.SolveInfo()
        sb = L.STRESS_DIVA if stress_balance.upper() == "DIVA" else L.STRESS_SSA
        rc = self._check(self.lib.ufe_run(self._ctx, sb, C.byref(info)))
        if rc == L.UFE_WARN_NOT_CONVERGED:
            warnings.warn(f"viscosity iteration failed to converge within {self.params.visc_it_nit} iterations!",
                          ViscosityIterationWarning)
        return info

    def timings(self):
        t = np.zeros(5)
        self.lib.ufe_last_timings(self._ctx, L.ptr(t))
        return {"vertex": t[0], "triangle": t[1], "assembly": t[2]}

    def profile_spmv(self, reps=200):
        """(average ms, algorithmic bytes) of the FP64 block product on the
        system assembled last, timed with HIP events on the solver stream."""
        ms, by = C.c_double(0), C.c_double(0)
        self._check(self.lib.ufe_profile_spmv(self._ctx, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def set_amg_options(self, **kw):
        """Multigrid knobs: omega (smoother damping), coarse_scale (coarse
        correction scale), tail_rows (largest level run in one workgroup)."""
        for k in kw:
            if k not in AMG_DEFAULTS:
                raise KeyError(f"unknown multigrid option {k}")
        opts = dict(self.amg_options, **kw)
        self._check(self.lib.ufe_set_amg_options(self._ctx, float(opts["omega"]), float(opts["coarse_scale"]),
                                                 int(opts["tail_rows"])))
        self.amg_options = opts

    def set_sliding_cap(self, alpha_sq=None, effective_pressure=None):
        """Coulomb limit alpha_sq * effective_pressure ([nV] each) of the
        Tsai2015 and Schoof2005 sliding laws (sliding_laws.f90:244-351), kept
        for every later solve; no arguments clears it."""
        if alpha_sq is None:
            self._check(self.lib.ufe_set_sliding_cap(self._ctx, None, None))
            return
        a, n = L.f64(alpha_sq), L.f64(effective_pressure)
        if a.size != self.md.mesh.nV or n.size != self.md.mesh.nV:
            raise ValueError("alpha_sq and effective_pressure need nV values")
        self._check(self.lib.ufe_set_sliding_cap(self._ctx, L.ptr(a), L.ptr(n)))

    def _apply(self, cfg):
        self.set_params(**{k: v for k, v in cfg.items() if k in _PARAM_NAMES})
        self.set_amg_options(**{k: v for k, v in cfg.items() if k in AMG_DEFAULTS})

    def autotune(self, stress_balance="DIVA", axes=AUTOTUNE_AXES, tol=1e-5):
        """Fastest configuration for the uploaded inputs, by coordinate search.

        Every trial solves from zero velocity (and zero DIVA state).  A trial
        that does not reach the stopping test, or whose velocities differ from
        those of the configuration the solver had on entry by more than `tol`
        (max relative difference), is rejected: the discrete equations can have
        more than one solution, and a tuned configuration must reach the same
        one.  Iteration counts move with round-off, so the gain is specific to
        these inputs; re-tune when they change much.  Leaves the solver set to
        the winner and returns (winner, [(config, seconds or None), ...])."""
        def trial(cfg):
            self._apply(cfg)
            self.set_velocity(None, None)
            if stress_balance.upper() == "DIVA":
                self.reset_diva_state()
            with warnings.catch_warnings():
                warnings.simplefilter("ignore", ViscosityIterationWarning)
                info = self.run(stress_balance)
            u, v = self.get_velocity()
            ok = info.rel_residual <= self.params.nonlinear_rtol
            return ok, info.seconds, u, v

        cur = {k: getattr(self.params, k) for k, _ in axes if k in _PARAM_NAMES}
        cur.update({k: self.amg_options[k] for k, _ in axes if k in AMG_DEFAULTS})
        ok, t_best, u0, v0 = trial(cur)
        if not ok:
            raise VelocitySolverError("the configuration on entry does not converge; nothing to tune against")
        scale = max(np.abs(u0).max(), np.abs(v0).max(), 1e-300)
        table = [(dict(cur), t_best)]
        best = dict(cur)
        for name, values in axes:
            for val in values:
                if val == best[name]:
                    continue
                cfg = dict(best, **{name: val})
                try:
                    ok, t, u, v = trial(cfg)
                except VelocitySolverError:
                    ok, t = False, None
                if ok and max(np.abs(u - u0).max(), np.abs(v - v0).max()) / scale <= tol:
                    table.append((cfg, t))
                    if t < t_best:
                        best, t_best = cfg, t
                else:
                    table.append((cfg, None))
        self._apply(best)
        return best, table

    def kernel_timing(self, enable=True):
        """Start (and reset) or stop live timing of the Krylov block products."""
        self._check(self.lib.ufe_kernel_timing(self._ctx, int(bool(enable))))

    def kernel_timing_read(self, kernel=0):
        """(launches, average ms, algorithmic bytes per launch) timed so far;
        kernel 0: Krylov block products, 1: finest-level prolongation + smoothing."""
        n, ms, by = C.c_longlong(0), C.c_double(0), C.c_double(0)
        self._check(self.lib.ufe_kernel_timing_read(self._ctx, int(kernel), C.byref(n), C.byref(ms),
                                                    C.byref(by)))
        return n.value, ms.value, by.value

    def profile_copy(self, nbytes, reps=50):
        """Average ms of a plain copy kernel over nbytes (known-traffic reference)."""
        ms = C.c_double(0)
        self._check(self.lib.ufe_profile_copy(self._ctx, int(nbytes), reps, C.byref(ms)))
        return ms.value

    def residual(self, u_b, v_b):
        """FP64 nonlinear SSA residual r = b(u) - A(u) u at (u_b, v_b)."""
        n = 2 * self.md.mesh.nT
        r = np.zeros(n)
        b = np.zeros(n)
        rel = C.c_double()
        self._check(self.lib.ufe_ssa_residual(self._ctx, L.ptr(L.f64(u_b)), L.ptr(L.f64(v_b)), L.ptr(r),
                                              L.ptr(b), C.byref(rel)))
        return rel.value, r, b

    def diva_residual(self, u_b, v_b, eta_3D_b=None, tau_bx_b=None, tau_by_b=None, u_base_b=None, v_base_b=None):
        """One DIVA state evaluation and the nonlinear residual at u_vav =
        (u_b, v_b) from the given DIVA state (eta_3D_b in the (nT, nz)
        triangle-fastest layout).  Returns a dict with rel_residual, r, b,
        eta_3D_b, F1_3D_b, F2_b_surface, beta_b, beta_eff_b."""
        nT, nz = self.md.mesh.nT, len(self.md.zeta)
        out = {"r": np.zeros(2 * nT), "b": np.zeros(2 * nT), "eta_3D_b": np.zeros(nT * nz),
               "F1_3D_b": np.zeros(nT * nz), "F2_b_surface": np.zeros(nT), "beta_b": np.zeros(nT),
               "beta_eff_b": np.zeros(nT)}
        rel = C.c_double()
        opt = [None if a is None else L.f64(a) for a in (eta_3D_b, tau_bx_b, tau_by_b, u_base_b, v_base_b)]
        self._check(self.lib.ufe_diva_residual(
            self._ctx, L.ptr(L.f64(u_b)), L.ptr(L.f64(v_b)), *(L.ptr(a) for a in opt), L.ptr(out["r"]),
            L.ptr(out["b"]), C.byref(rel), L.ptr(out["eta_3D_b"]), L.ptr(out["F1_3D_b"]),
            L.ptr(out["F2_b_surface"]), L.ptr(out["beta_b"]), L.ptr(out["beta_eff_b"])))
        out["rel_residual"] = rel.value
        return out

    # -- drop-in entry points ---------------------------------------------------
    def solve_SSA(self, Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice=None, u_b=None, v_b=None,
                  BC_prescr_mask_b=None, BC_prescr_u_b=None, BC_prescr_v_b=None):
        """Returns (u_b, v_b, n_visc_its, n_Axb_its, info)."""
        nT = self.md.mesh.nT
        u = L.f64(np.zeros(nT) if u_b is None else u_b).copy()
        v = L.f64(np.zeros(nT) if v_b is None else v_b).copy()
        self.upload(Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice,
                    BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b)
        self.set_velocity(u, v)
        info = self.run("SSA")
        u, v = self.get_velocity()
        return u, v, info.n_visc_its, info.n_Axb_its, info

    def solve_DIVA(self, Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice=None, state=None,
                   BC_prescr_mask_b=None, BC_prescr_u_b=None, BC_prescr_v_b=None):
        """state: dict of the DIVA solver fields carried between calls
        (u_vav_b, v_vav_b, u_base_b, v_base_b, eta_3D_b, tau_bx_b, tau_by_b);
        returns (state, n_visc_its, n_Axb_its, info) with u_3D_b, v_3D_b added."""
        nT, nz = self.md.mesh.nT, len(self.md.zeta)
        st = {k: np.zeros(nT) for k in ("u_vav_b", "v_vav_b", "u_base_b", "v_base_b", "tau_bx_b", "tau_by_b")}
        st["eta_3D_b"] = np.zeros(nT * nz)
        for k, a in (state or {}).items():
            if k in st:
                st[k] = L.f64(a).copy()
        self.upload(Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice,
                    BC_prescr_mask_b, BC_prescr_u_b, BC_prescr_v_b)
        self.set_velocity(st["u_vav_b"], st["v_vav_b"])
        self._check(self.lib.ufe_upload_diva_state(self._ctx, L.ptr(st["eta_3D_b"]), L.ptr(st["tau_bx_b"]),
                                                   L.ptr(st["tau_by_b"]), L.ptr(st["u_base_b"]),
                                                   L.ptr(st["v_base_b"])))
        info = self.run("DIVA")
        st["u_vav_b"], st["v_vav_b"] = self.get_velocity()
        st["u_3D_b"] = np.zeros(nT * nz)
        st["v_3D_b"] = np.zeros(nT * nz)
        self._check(self.lib.ufe_download_diva(self._ctx, L.ptr(st["u_base_b"]), L.ptr(st["v_base_b"]),
                                               L.ptr(st["u_3D_b"]), L.ptr(st["v_3D_b"]), L.ptr(st["eta_3D_b"]),
                                               L.ptr(st["tau_bx_b"]), L.ptr(st["tau_by_b"])))
        return st, info.n_visc_its, info.n_Axb_its, info

    def fraction_gr_b(self, Hi, Hb, SL):
        out = np.zeros(self.md.mesh.nT)
        self._check(self.lib.ufe_fraction_gr_b(self._ctx, L.ptr(L.f64(Hi)), L.ptr(L.f64(Hb)), L.ptr(L.f64(SL)),
                                               L.ptr(out)))
        return out

    # -- inputs of the heat columns ---------------------------------------------
    def heat_setup(self):
        """Attach the triangles around each vertex (the mesh's iTri, counter-
        clockwise) for the upwind advection terms."""
        m = self.md.mesh
        self._heat_keep = (L.i32(m.iTri_ptr), L.i32(m.iTri_ind))
        self._check(self.lib.ufe_heat_setup(self._ctx, L.ptr(self._heat_keep[0], L.ip),
                                            L.ptr(self._heat_keep[1], L.ip)))

    def upwind_heat_flux_derivatives(self, Ti, Hi, u_vav, v_vav, u_3D_b, v_3D_b):
        """calc_upwind_heat_flux_derivatives (thermodynamics_utilities.f90:339-434):
        (u_times_dTdxp_upwind, v_times_dTdyp_upwind), each (nz*nV,) with the
        vertex fastest; Ti likewise, u_3D_b / v_3D_b (nz*nT,) triangle fastest."""
        if getattr(self, "_heat_keep", None) is None:
            self.heat_setup()
        m = self.md.mesh
        nz = len(self.md.zeta)
        ox, oy = np.zeros(nz * m.nV), np.zeros(nz * m.nV)
        a = [L.f64(x) for x in (Ti, Hi, u_vav, v_vav, u_3D_b, v_3D_b)]
        self._check(self.lib.ufe_heat_upwind_derivatives(self._ctx, *(L.ptr(x) for x in a), L.ptr(ox), L.ptr(oy)))
        return ox, oy

    def strain_heating(self, Hi, A_flow, strain_rates):
        """calc_strain_heating (thermodynamics_utilities.f90:24-91): Phi (nz*nV,)
        from the nine 3-D strain-rate fields in the order du_dx, du_dy, du_dz,
        dv_dx, dv_dy, dv_dz, dw_dx, dw_dy, dw_dz, with this solver's glens_n."""
        m = self.md.mesh
        nz = len(self.md.zeta)
        if len(strain_rates) != 9:
            raise VelocitySolverError("strain_heating needs the nine 3-D strain-rate fields")
        keep = [L.f64(x) for x in strain_rates]
        arr = (L.dp * 9)(*(L.ptr(x) for x in keep))
        Phi = np.zeros(nz * m.nV)
        self._check(self.lib.ufe_heat_strain_heating(self._ctx, L.ptr(L.f64(Hi)), L.ptr(L.f64(A_flow)), arr,
                                                     L.ptr(Phi)))
        return Phi

    def frictional_heating(self, basal_friction_coefficient, uabs_base, mask_grounded_ice):
        """calc_frictional_heating (thermodynamics_utilities.f90:93-128)."""
        Q = np.zeros(self.md.mesh.nV)
        mk = L.i32(np.asarray(mask_grounded_ice).astype(np.int32))
        self._check(self.lib.ufe_heat_frictional_heating(self._ctx, L.ptr(L.f64(basal_friction_coefficient)),
                                                         L.ptr(L.f64(uabs_base)), L.ptr(mk, L.ip), L.ptr(Q)))
        return Q

    # -- zeta gradients and the BPA ---------------------------------------------
    def zeta_setup(self, aa=None, zeta_stag=None):
        """Attach the a->a operators (built from the mesh when not given) and
        the staggered layers (midpoints when not given)."""
        from .operators import build_aa
        aa = aa if aa is not None else build_aa(self.md.mesh)
        zs = None if zeta_stag is None else L.f64(zeta_stag)
        self._zeta_keep = (L.i32(aa.ptr), L.i32(aa.ind), L.f64(aa.vals["ddx"]), L.f64(aa.vals["ddy"]), zs)
        k = self._zeta_keep
        self._check(self.lib.ufe_zeta_setup(self._ctx, aa.nnz, L.ptr(k[0], L.ip), L.ptr(k[1], L.ip), L.ptr(k[2]),
                                            L.ptr(k[3]), L.ptr(zs)))

    def _need_zeta(self):
        if getattr(self, "_zeta_keep", None) is None:
            self.zeta_setup()

    def zeta_gradients(self, Hi, Hs, dHi_dt=None, dHs_dt=None):
        """calc_zeta_gradients (zeta_gradients.f90:19-212): (ak, bk, bks) lists of
        7, 6 and 6 fields (nz*nV,), (nz*nT,), ((nz-1)*nT,), first index fastest:
        ak = dzeta_dt, dzeta_dx, dzeta_dy, dzeta_dz, d2zeta_dx2, d2zeta_dxdy,
        d2zeta_dy2; bk and bks = the same without dzeta_dt."""
        self._need_zeta()
        m = self.md.mesh
        nz = len(self.md.zeta)
        ak = [np.zeros(nz * m.nV) for _ in range(7)]
        bk = [np.zeros(nz * m.nT) for _ in range(6)]
        bks = [np.zeros((nz - 1) * m.nT) for _ in range(6)]
        ins = [None if a is None else L.f64(a) for a in (Hi, Hs, dHi_dt, dHs_dt)]
        arr = [(L.dp * len(x))(*(L.ptr(a) for a in x)) for x in (ak, bk, bks)]
        self._check(self.lib.ufe_zeta_gradients(self._ctx, *(L.ptr(a) for a in ins), *arr))
        return ak, bk, bks

    def _bpa_inputs(self, Hi, Hs, A_flow, beta_sq, fraction_gr_b=None, mask_grounded_ice=None,
                    BC_prescr_mask_bk=None, BC_prescr_u_bk=None, BC_prescr_v_bk=None):
        bcm, bcu, bcv = self._bc(BC_prescr_mask_bk, BC_prescr_u_bk, BC_prescr_v_bk)
        keep = [L.f64(Hi), L.f64(Hs), L.f64(A_flow), L.f64(beta_sq),
                None if fraction_gr_b is None else L.f64(fraction_gr_b),
                None if mask_grounded_ice is None else L.i32(mask_grounded_ice), bcm, bcu, bcv]
        s = L.BpaInputs()
        s.Hi, s.Hs, s.A_flow, s.beta_sq, s.fraction_gr_b = (L.ptr(a) for a in keep[:5])
        s.mask_grounded_ice = L.ptr(keep[5], L.ip)
        s.BC_prescr_mask_bk = L.ptr(keep[6], L.ip)
        s.BC_prescr_u_bk, s.BC_prescr_v_bk = L.ptr(keep[7]), L.ptr(keep[8])
        return s, keep

    def solve_BPA(self, Hi, Hs, A_flow, beta_sq, fraction_gr_b=None, mask_grounded_ice=None, u_bk=None, v_bk=None,
                  BC_prescr_mask_bk=None, BC_prescr_u_bk=None, BC_prescr_v_bk=None):
        """solve_BPA (BPA_main.f90:90-252).  3-D arrays (nz*nT,), triangle
        fastest.  Returns (u_bk, v_bk, info); a ViscosityIterationWarning when
        visc_it_nit is reached."""
        self._need_zeta()
        nT, nz = self.md.mesh.nT, len(self.md.zeta)
        u = L.f64(np.zeros(nz * nT) if u_bk is None else u_bk).copy()
        v = L.f64(np.zeros(nz * nT) if v_bk is None else v_bk).copy()
        s, keep = self._bpa_inputs(Hi, Hs, A_flow, beta_sq, fraction_gr_b, mask_grounded_ice, BC_prescr_mask_bk,
                                   BC_prescr_u_bk, BC_prescr_v_bk)
        info = L.SolveInfo()
        rc = self._check(self.lib.ufe_solve_bpa(self._ctx, C.byref(s), L.ptr(u), L.ptr(v), C.byref(info)))
        if rc == 1:
            warnings.warn(f"BPA viscosity iteration failed to converge within {self.params.visc_it_nit} iterations",
                          ViscosityIterationWarning)
        return u, v, info

    def bpa_residual(self, u_bk, v_bk, Hi, Hs, A_flow, beta_sq, fraction_gr_b=None, BC_prescr_mask_bk=None,
                     BC_prescr_u_bk=None, BC_prescr_v_bk=None):
        """One BPA state evaluation at (u_bk, v_bk): dict with rel_residual, r,
        b (rows 2*(ti*nz+k)+uv), eta_bk (nz*nT,) and beta_b (nT,)."""
        self._need_zeta()
        nT, nz = self.md.mesh.nT, len(self.md.zeta)
        s, keep = self._bpa_inputs(Hi, Hs, A_flow, beta_sq, fraction_gr_b, None, BC_prescr_mask_bk,
                                   BC_prescr_u_bk, BC_prescr_v_bk)
        out = {"r": np.zeros(2 * nT * nz), "b": np.zeros(2 * nT * nz), "eta_bk": np.zeros(nT * nz),
               "beta_b": np.zeros(nT)}
        rel = C.c_double(0.0)
        self._check(self.lib.ufe_bpa_residual(self._ctx, C.byref(s), L.ptr(L.f64(u_bk)), L.ptr(L.f64(v_bk)),
                                              L.ptr(out["r"]), L.ptr(out["b"]), L.ptr(out["eta_bk"]),
                                              L.ptr(out["beta_b"]), C.byref(rel)))
        out["rel_residual"] = rel.value
        return out

    def bpa_timings(self):
        """Stage times (ms) of the last solve_BPA: state, assembly, column
        factorisation, GMRES, update."""
        ms = np.zeros(5)
        self._check(self.lib.ufe_bpa_last_timings(self._ctx, L.ptr(ms)))
        return dict(zip(("state", "assembly", "column_factor", "gmres", "update"), ms.tolist()))

    def bpa_profile_spmv(self, reps=100):
        """(average ms, algorithmic bytes) of the BPA block product."""
        ms, by = C.c_double(0.0), C.c_double(0.0)
        self._check(self.lib.ufe_bpa_profile_spmv(self._ctx, int(reps), C.byref(ms), C.byref(by)))
        return ms.value, by.value

    # -- ice thickness -------------------------------------------------------
    def thickness_setup(self, geometry=None):
        """Attach the Voronoi geometry of the mesh (cell areas, connection
        widths, edge triangles; computed from the mesh when not given)."""
        from .mesh import vertex_geometry
        m = self.md.mesh
        g = geometry if geometry is not None else vertex_geometry(m)
        self._thk_keep = (L.i32(m.C_ptr), L.i32(m.C_ind), L.i32(m.VBI), L.f64(g.A), L.f64(g.Cw), L.i32(g.ETri))
        cp, ci, vbi, A, Cw, et = self._thk_keep
        self._thk_geom = L.VertexGeometry(int(cp[-1]), L.ptr(cp, L.ip), L.ptr(ci, L.ip), L.ptr(vbi, L.ip),
                                          L.ptr(A), L.ptr(Cw), L.ptr(et, L.ip))
        self._check(self.lib.ufe_thickness_setup(self._ctx, C.byref(self._thk_geom)))
        return g

    def calc_dHi_dt_semiimplicit(self, Hi, Hb, SL, u_vav_b, v_vav_b, SMB, BMB, LMB, fraction_margin, mask_noice,
                                 dt, dHi_dt_target=None, BC_prescr_mask=None, BC_prescr_Hi=None,
                                 dHi_dt_guess=None, **params):
        """One semi-implicit thickness step (calc_dHi_dt_semiimplicit,
        conservation_of_mass_semiimplicit.f90:23-159): returns
        (dHi_dt, Hi_tplusdt, divQ, dt_max, n_its).  u_vav_b = v_vav_b = None
        takes the velocity currently held by this solver.  BC_prescr_mask and
        BC_prescr_Hi go together, as in the reference.  params override
        ufe_thickness_params (fs, bc_H, dt_ice_max, dt_ice_min, Hi_min, rtol,
        abstol, krylov, gmres_restart, maxit); bc_H takes 'zero'/'infinite'
        names or the integer codes."""
        if (BC_prescr_mask is None) != (BC_prescr_Hi is None):
            raise VelocitySolverError("need to provide prescribed both Hi and mask!")
        if getattr(self, "_thk_geom", None) is None:
            self.thickness_setup()
        nV = self.md.mesh.nV
        p = thickness_params(**params)
        inp, _keep = make_thickness_inputs(nV, Hi, Hb, SL, u_vav_b, v_vav_b, SMB, BMB, LMB, fraction_margin,
                                           mask_noice, dt, dHi_dt_target, BC_prescr_mask, BC_prescr_Hi,
                                           dHi_dt_guess)
        Hn, rate, divQ = np.zeros(nV), np.zeros(nV), np.zeros(nV)
        dt_max, its = C.c_double(), C.c_int()
        rc = self._check(self.lib.ufe_thickness_semiimplicit(self._ctx, C.byref(p), C.byref(inp), L.ptr(Hn),
                                                             L.ptr(rate), L.ptr(divQ), C.byref(dt_max),
                                                             C.byref(its)))
        if rc == L.UFE_WARN_NOT_CONVERGED:
            warnings.warn("thickness solve did not reach its tolerance", ViscosityIterationWarning)
        return rate, Hn, divQ, dt_max.value, its.value

    def close(self):
        if self._ctx:
            self.lib.ufe_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def till_yield_stress(mesh, effective_pressure, till_friction_angle, mask_grounded_ice, mask_icefree_land,
                      Hi_min=0.0):
    """Till yield stress of the Coulomb, Budd and Zoet-Iverson laws, the field
    the solves take in place of beta_sq for them: N tan(pi/180) phi, then
    extend_till_yield_stress_to_neighbours (sliding_laws.f90:132-242, :353-403,
    :643-700).  Host code of the library; [nV] arrays."""
    lib = L.lib()
    nV = mesh.nV
    cp, ci = L.i32(mesh.C_ptr), L.i32(mesh.C_ind)
    N, phi = L.f64(effective_pressure), L.f64(till_friction_angle)
    mg = L.i32(np.asarray(mask_grounded_ice).astype(np.int32))
    ml = L.i32(np.asarray(mask_icefree_land).astype(np.int32))
    if not (N.size == phi.size == mg.size == ml.size == nV):
        raise ValueError("till_yield_stress needs nV values per field")
    out = np.zeros(nV)
    L.check(lib.ufe_till_yield_stress(nV, L.ptr(cp, L.ip), L.ptr(ci, L.ip), L.ptr(N), L.ptr(phi),
                                      L.ptr(mg, L.ip), L.ptr(ml, L.ip), float(Hi_min), L.ptr(out)))
    return out


def thickness_params(**params) -> L.ThicknessParams:
    """ufe_thickness_params with the library defaults, overridden by params
    (bc_H takes 'zero'/'infinite' names or the integer codes)."""
    p = L.default_thickness_params()
    names = dict(L.ThicknessParams._fields_)
    for k, v in params.items():
        if k == "bc_H":
            for i, x in enumerate(v):
                p.bc_H[i] = L.BC[x] if isinstance(x, str) else int(x)
        elif k in names:
            setattr(p, k, v)
        else:
            raise KeyError(f"unknown thickness parameter {k}")
    return p


def make_thickness_inputs(nV, Hi, Hb, SL, u_vav_b, v_vav_b, SMB, BMB, LMB, fraction_margin, mask_noice, dt,
                          dHi_dt_target=None, BC_prescr_mask=None, BC_prescr_Hi=None, dHi_dt_guess=None):
    """ufe_thickness_inputs over contiguous copies of the fields; returns the
    structure and the arrays it points into (keep them alive while it is used)."""
    keep = [L.f64(a) for a in (Hi, Hb, SL, SMB, BMB, LMB, fraction_margin,
                               np.zeros(nV) if dHi_dt_target is None else dHi_dt_target)]
    for a in keep:
        if a.shape != (nV,):
            raise VelocitySolverError(f"thickness input of shape {a.shape}, expected ({nV},)")
    uv = (None, None) if u_vav_b is None else (L.f64(u_vav_b), L.f64(v_vav_b))
    noice = L.i32(np.asarray(mask_noice).astype(np.int32))
    bcm = None if BC_prescr_mask is None else L.i32(np.asarray(BC_prescr_mask).astype(np.int32))
    bch = None if BC_prescr_Hi is None else L.f64(BC_prescr_Hi)
    guess = None if dHi_dt_guess is None else L.f64(dHi_dt_guess)
    inp = L.ThicknessInputs(L.ptr(keep[0]), L.ptr(keep[1]), L.ptr(keep[2]), L.ptr(uv[0]), L.ptr(uv[1]),
                            L.ptr(keep[3]), L.ptr(keep[4]), L.ptr(keep[5]), L.ptr(keep[6]), L.ptr(keep[7]),
                            L.ptr(noice, L.ip), L.ptr(bcm, L.ip), L.ptr(bch), L.ptr(guess), float(dt))
    return inp, (keep, uv, noice, bcm, bch, guess)


def make_heat_inputs(nV, nz, zeta, fields: dict, dt: float, Hi_min_thermo: float = 10.0):
    """HeatInputs with its arrays kept alive in the returned tuple.  fields
    may hold fraction_gr (vertex grounded fraction) for mask-3 columns, the
    grounding-line blend of choice_GL_temperature_BC = 'subgrid'."""
    keep = {k: (L.i32(v) if k == "mask" else L.f64(v)) for k, v in fields.items()}
    keep["zeta"] = L.f64(zeta)
    h = L.HeatInputs()
    h.nV, h.nz = nV, nz
    for name in ("zeta", "Ti", "u_3D", "v_3D", "w_3D", "u_dTdxp_upwind", "v_dTdyp_upwind", "dzeta_dx", "dzeta_dy",
                 "dzeta_dz", "dzeta_dt", "Phi", "Hi_eff", "T_surf", "Q_base_grnd"):
        setattr(h, name, L.ptr(keep[name]))
    h.mask = L.ptr(keep["mask"], L.ip)
    if "fraction_gr" in keep:
        h.fraction_gr = L.ptr(keep["fraction_gr"])
    h.dt = dt
    h.Hi_min_thermo = Hi_min_thermo
    return h, keep


def heat_columns(hin, solver: VelocitySolver | None = None):
    """Vertical heat columns on the device; returns (Ti_next, is_unstable, n_unstable)."""
    lib = L.lib()
    n = hin.nV * hin.nz
    T = np.zeros(n)
    unst = np.zeros(hin.nV, np.int32)
    nun = C.c_int(0)
    ctx = solver._ctx if solver is not None else None
    rc = lib.ufe_heat_columns(ctx, C.byref(hin), L.ptr(T), L.ptr(unst, L.ip), C.byref(nun))
    if rc < 0:
        raise VelocitySolverError(f"ufe_heat_columns failed ({rc}): {lib.ufe_last_error(None).decode()}")
    return T, unst, nun.value


HEAT_FIELDS = ("zeta", "Ti", "u_3D", "v_3D", "w_3D", "u_dTdxp_upwind", "v_dTdyp_upwind", "dzeta_dx", "dzeta_dy",
               "dzeta_dz", "dzeta_dt", "Phi", "Hi_eff", "T_surf", "Q_base_grnd")


def make_heat_inputs_device(nV, nz, ptrs: dict, dt: float, Hi_min_thermo: float = 10.0):
    """HeatInputs whose arrays are device addresses (ints): ptrs holds one per
    name of HEAT_FIELDS, "mask" (int32) and optionally "fraction_gr"."""
    h = L.HeatInputs()
    h.nV, h.nz = nV, nz
    for name in HEAT_FIELDS:
        setattr(h, name, C.cast(C.c_void_p(ptrs[name]), L.dp))
    h.mask = C.cast(C.c_void_p(ptrs["mask"]), L.ip)
    if ptrs.get("fraction_gr"):
        h.fraction_gr = C.cast(C.c_void_p(ptrs["fraction_gr"]), L.dp)
    h.dt = dt
    h.Hi_min_thermo = Hi_min_thermo
    return h


def heat_columns_device(hin, Tout_ptr: int, unstable_ptr: int, solver: VelocitySolver | None = None) -> int:
    """ufe_heat_columns_device: inputs and outputs in device memory (addresses
    as ints); returns the number of unstable columns."""
    lib = L.lib()
    nun = C.c_int(0)
    ctx = solver._ctx if solver is not None else None
    rc = lib.ufe_heat_columns_device(ctx, C.byref(hin), C.c_void_p(Tout_ptr), C.c_void_p(unstable_ptr),
                                     C.byref(nun))
    if rc < 0:
        raise VelocitySolverError(f"ufe_heat_columns_device failed ({rc}): {lib.ufe_last_error(None).decode()}")
    return nun.value


def heat_columns_device_noalloc(hin, Tout_ptr: int, unstable_ptr: int) -> int:
    """ufe_heat_columns_device_noalloc: the device-array heat step with no
    device allocation (scratch in the library's static array) on the current
    device and the null stream; returns the number of unstable columns."""
    lib = L.lib()
    nun = C.c_int(0)
    rc = lib.ufe_heat_columns_device_noalloc(C.byref(hin), C.c_void_p(Tout_ptr), C.c_void_p(unstable_ptr),
                                             C.byref(nun))
    if rc < 0:
        raise VelocitySolverError(
            f"ufe_heat_columns_device_noalloc failed ({rc}): {lib.ufe_last_error(None).decode()}")
    return nun.value


def make_robin_inputs(nV, nz, zeta, Hi_eff, T_surf, SMB, geothermal_flux, mask, Hi_min_thermo=10.0):
    """ufe_heat_robin_inputs from host arrays; returns (struct, arrays kept alive)."""
    keep = [L.f64(zeta), L.f64(Hi_eff), L.f64(T_surf), L.f64(SMB), L.f64(geothermal_flux), L.i32(mask)]
    r = L.HeatRobinInputs(nV, nz, L.ptr(keep[0]), L.ptr(keep[1]), L.ptr(keep[2]), L.ptr(keep[3]), L.ptr(keep[4]),
                          L.ptr(keep[5], L.ip), float(Hi_min_thermo))
    return r, keep


def heat_robin(rin, is_unstable, Ti, solver: VelocitySolver | None = None):
    """Robin replacement of the flagged columns (the reference's handling of
    unstable heat columns); returns (new Ti, number replaced).  Raises when
    ceil(nV / 100) or more columns are flagged, as the reference stops."""
    lib = L.lib()
    T = L.f64(Ti).copy()
    flag = L.i32(is_unstable)
    n = C.c_int(0)
    ctx = solver._ctx if solver is not None else None
    rc = lib.ufe_heat_robin(ctx, C.byref(rin), L.ptr(flag, L.ip), L.ptr(T), C.byref(n))
    if rc < 0:
        raise VelocitySolverError(f"ufe_heat_robin failed ({rc}): {lib.ufe_last_error(None).decode()}")
    return T, n.value


def csr_matvec(A, x):
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    lib = L.lib()
    y = np.zeros(A.shape[0])
    ptr, ind, val = L.i32(A.indptr), L.i32(A.indices), L.f64(A.data)
    rc = lib.ufe_csr_matvec(A.shape[0], A.shape[1], A.nnz, L.ptr(ptr, L.ip), L.ptr(ind, L.ip), L.ptr(val),
                            L.ptr(L.f64(x)), L.ptr(y))
    if rc < 0:
        raise VelocitySolverError(lib.ufe_last_error(None).decode())
    return y


def csr_solve(A, b, x0=None, rtol=1e-8, abstol=0.0, maxit=10000, precond=L.PC_BLOCK_JACOBI, krylov=None,
              restart=30):
    """Solve A x = b on the device (ufe_csr_solve, or ufe_csr_solve_ex when a
    Krylov method is named: L.KRYLOV_BICGSTAB or L.KRYLOV_GMRES)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    lib = L.lib()
    x = L.f64(np.zeros(A.shape[0]) if x0 is None else x0).copy()
    ptr, ind, val = L.i32(A.indptr), L.i32(A.indices), L.f64(A.data)
    its = C.c_int(0)
    rel = C.c_double(0)
    if krylov is None:
        rc = lib.ufe_csr_solve(A.shape[0], A.nnz, L.ptr(ptr, L.ip), L.ptr(ind, L.ip), L.ptr(val), L.ptr(L.f64(b)),
                               L.ptr(x), rtol, abstol, maxit, precond, C.byref(its), C.byref(rel))
    else:
        rc = lib.ufe_csr_solve_ex(A.shape[0], A.nnz, L.ptr(ptr, L.ip), L.ptr(ind, L.ip), L.ptr(val),
                                  L.ptr(L.f64(b)), L.ptr(x), rtol, abstol, maxit, precond, krylov, restart,
                                  C.byref(its), C.byref(rel))
    if rc < 0:
        raise VelocitySolverError(lib.ufe_last_error(None).decode())
    return x, its.value, rel.value
