"""Writes oracle_channel_{ssa,diva}.npz: converged FP64 oracle solutions of the
1.5k-vertex marine channel (SSA and DIVA, sans form, vel_max 1e5, nonlinear
residual 1e-8), so the GPU parity tests need not rerun a multi-second CPU
solve.  The workload is regenerated from its seed; the mesh digest stored
with the fixture guards against a different mesh.

Run: python tests/golden/make_oracle_fixtures.py
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
SETTINGS = dict(include_crossterms=0, vel_max=1e5, abstol=0.0, rtol=1e-11, stop_rule=1, nonlinear_rtol=1e-8,
                visc_it_nit=800)


def params():
    p = L.default_params()
    for k, v in SETTINGS.items():
        setattr(p, k, v)
    return p


def main():
    for sb in ("SSA", "DIVA"):
        w = W.channel_small(target_nV=CHANNEL_NV, stress_balance=sb)
        o = OB.Oracle(w.descriptor())
        inp = w.inputs()
        if sb == "SSA":
            rc, u, v, info = o.solve_ssa(params(), inp)
            extra = {}
        else:
            rc, st, info = o.solve_diva(params(), inp)
            u, v = st["u"], st["v"]
            nz, nT = w.nz, w.mesh.nT
            u3 = st["u3d"].reshape(nz, nT)
            extra = dict(u_base=st["ubase"], v_base=st["vbase"], tau_bx=st["tbx"], tau_by=st["tby"],
                         u3d_surface=u3[0], u3d_base=u3[-1])
        assert rc == 0, rc
        path = os.path.join(HERE, f"oracle_channel_{sb.lower()}.npz")
        np.savez_compressed(path, u=u, v=v, digest=w.mesh.digest(), n_visc_its=info.n_visc_its,
                            rel_residual=info.rel_residual, **extra)
        print(sb, w.mesh.nV, w.mesh.nT, info.n_visc_its, info.rel_residual, "->", path, flush=True)


if __name__ == "__main__":
    main()
