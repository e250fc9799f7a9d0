"""Writes oracle_channel_{ssa,diva}.npz: converged FP64 oracle solutions of the
1.5k-vertex marine channel (SSA and DIVA, sans form, vel_max 1e5, nonlinear
residual 1e-8), so the GPU parity tests need not rerun a multi-second CPU
solve.  The workload is regenerated from its seed; the mesh digest stored
with the fixture guards against a different mesh.

oracle_channel_seeds.npz holds the same solves for the channel's random-field
seeds 3..6 (velocities rounded to float32, ample for a 1e-5 comparison).

oracle_channel_diva_20k.npz, oracle_ensemble_members.npz and
oracle_continental_small.npz ("scale") hold converged oracle solutions (the
reference iteration with warm-started linear solves to 1e-9, nonlinear
residual 1e-8) of larger meshes: the DIVA
channel at ~20k vertices, three members of the ensemble generator at
~12k vertices and a continental-generator mesh at ~10k vertices, with the
oracle's wall time and thread count (the measured CPU solves against which the
benchmark's CPU-time model is checked).  These take tens of minutes each.

Run: python tests/golden/make_oracle_fixtures.py [main|seeds|scale:<name>]
     (main and seeds by default; scale:channel20k, scale:members, scale:continental)
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

ge.load_package()
from ufemism2_amd import _lib as L, workloads as W  # noqa: E402
from oracle import binding as OB  # noqa: E402

CHANNEL_NV = 1500
SEEDS = (3, 4, 5, 6)
SETTINGS = dict(include_crossterms=0, vel_max=1e5, abstol=0.0, rtol=1e-11, stop_rule=1, nonlinear_rtol=1e-8,
                visc_it_nit=800)


# the larger meshes: the reference's warm-started linear solves (the previous
# Picard solution as the initial guess, solve_linearised_SSA_DIVA.f90:74-84) to
# 1e-9, ample for a nonlinear residual of 1e-8 and hours faster than cold
# solves to 1e-11
SCALE_SETTINGS = dict(SETTINGS, rtol=1e-9, warm_start=1)


def params(settings=SETTINGS):
    p = L.default_params()
    for k, v in settings.items():
        setattr(p, k, v)
    return p


def solve(w, sb, settings=SETTINGS):
    o = OB.Oracle(w.descriptor())
    inp = w.inputs()
    if sb == "SSA":
        rc, u, v, info = o.solve_ssa(params(settings), inp)
        return rc, u, v, info, None
    rc, st, info = o.solve_diva(params(settings), inp)
    return rc, st["u"], st["v"], info, st


def seeds():
    out = {}
    for seed in SEEDS:
        for sb in ("SSA", "DIVA"):
            w = W.channel(seed=seed, target_nV=CHANNEL_NV, nz=10, stress_balance=sb)
            rc, u, v, info, _ = solve(w, sb)
            assert rc == 0, rc
            key = f"{sb.lower()}_{seed}"
            out[key + "_u"] = u.astype(np.float32)
            out[key + "_v"] = v.astype(np.float32)
            out[key + "_digest"] = w.mesh.digest()
            print(seed, sb, info.n_visc_its, info.rel_residual, flush=True)
    path = os.path.join(HERE, "oracle_channel_seeds.npz")
    np.savez_compressed(path, seeds=np.array(SEEDS), **out)
    print("->", path, flush=True)


def main():
    for sb in ("SSA", "DIVA"):
        w = W.channel_small(target_nV=CHANNEL_NV, stress_balance=sb)
        rc, u, v, info, st = solve(w, sb)
        extra = {}
        if sb == "DIVA":
            nz, nT = w.nz, w.mesh.nT
            u3 = st["u3d"].reshape(nz, nT)
            extra = dict(u_base=st["ubase"], v_base=st["vbase"], tau_bx=st["tbx"], tau_by=st["tby"],
                         u3d_surface=u3[0], u3d_base=u3[-1])
        assert rc == 0, rc
        path = os.path.join(HERE, f"oracle_channel_{sb.lower()}.npz")
        np.savez_compressed(path, u=u, v=v, digest=w.mesh.digest(), n_visc_its=info.n_visc_its,
                            rel_residual=info.rel_residual, **extra)
        print(sb, w.mesh.nV, w.mesh.nT, info.n_visc_its, info.rel_residual, "->", path, flush=True)


def timed_solve(w):
    import time
    t0 = time.perf_counter()
    rc, u, v, info, st = solve(w, "DIVA", SCALE_SETTINGS)
    t = time.perf_counter() - t0
    assert rc == 0, (w.name, rc)
    meta = dict(digest=w.mesh.digest(), nV=w.mesh.nV, nT=w.mesh.nT, n_visc_its=info.n_visc_its,
                n_Axb_its=info.n_Axb_its, rel_residual=info.rel_residual,
                rel_residual_unscaled=info.rel_residual_unscaled, seconds=t,
                threads=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    print(w.name, meta, flush=True)
    return u.astype(np.float32), v.astype(np.float32), meta


def scale(name):
    if name == "channel20k":
        w = W.channel(seed=2, target_nV=20000, nz=10, stress_balance="DIVA")
        u, v, meta = timed_solve(w)
        np.savez_compressed(os.path.join(HERE, "oracle_channel_diva_20k.npz"), u=u, v=v, **meta)
    elif name == "members":
        out = {}
        for k in range(3):
            w = W.ensemble_member(k, nV_range=(10000, 14000))
            u, v, meta = timed_solve(w)
            out.update({f"m{k}_u": u, f"m{k}_v": v, **{f"m{k}_{a}": b for a, b in meta.items()}})
        np.savez_compressed(os.path.join(HERE, "oracle_ensemble_members.npz"), members=np.arange(3),
                            nV_range=np.array([10000, 14000]), **out)
    elif name == "continental":
        w = W.continental(target_nV=10000)
        u, v, meta = timed_solve(w)
        np.savez_compressed(os.path.join(HERE, "oracle_continental_small.npz"), u=u, v=v, target_nV=10000, **meta)
    else:
        raise SystemExit(f"unknown scale case {name}")


if __name__ == "__main__":
    what = sys.argv[1:] or ["main", "seeds"]
    if "main" in what:
        main()
    if "seeds" in what:
        seeds()
    for a in what:
        if a.startswith("scale:"):
            scale(a.split(":", 1)[1])
