"""The Fortran host interface (ufe_velocity_gpu, ISO_C_BINDING): the driver
built from it solves the same case as the Python host through the same C ABI
and must return the same velocities."""

import os

import numpy as np
import pytest

from conftest import max_rel


def test_fortran_driver_built(pkg):
    from ufemism2_amd import fortran_case as FC
    assert os.path.exists(FC.DRIVER), "build() compiles the Fortran driver with amdflang"


def test_case_file_layout(W, tmp_path):
    from ufemism2_amd import fortran_case as FC
    w = W.dome(h=120e3)
    path = str(tmp_path / "case.bin")
    FC.write_case(path, w)
    m, ops = w.mesh, w.ops
    n_int = (7 + 3 * m.nT + 3 * m.nT + m.nT + (m.nT + 1) + ops.ab.nnz + (m.nV + 1) + ops.ba.nnz + (m.nT + 1)
             + ops.bb.nnz + 6 + m.nV)
    n_dbl = (2 * m.nV + 2 * m.nT + w.nz + 3 * ops.ab.nnz + 3 * ops.ba.nnz + 5 * ops.bb.nnz + 5 + 2 * m.nV
             + w.nz * m.nV + m.nV + m.nT)
    assert os.path.getsize(path) == 4 * n_int + 8 * n_dbl


@pytest.mark.gpu
@pytest.mark.parametrize("stress", ["SSA", "DIVA"])
def test_fortran_solve_matches_c_abi(V, W, tmp_path, stress):
    from ufemism2_amd import fortran_case as FC
    w = W.dome(h=90e3) if stress == "SSA" else W.channel_small(target_nV=1500, stress_balance="DIVA")
    if stress == "SSA":
        ctl = dict(stop_rule=1, visc_it_nit=300, include_crossterms=1, precond=2, warm_start=0, accel_depth=0,
                   nonlinear_rtol=1e-8, vel_max=5000.0, abstol=0.0, rtol=1e-10)
    else:
        ctl = dict(stop_rule=1, visc_it_nit=800, include_crossterms=0, precond=3, warm_start=1, accel_depth=2,
                   accel_mixing=0.6, nonlinear_rtol=1e-8, vel_max=1e5, abstol=0.0, rtol=0.1)
    case, out = str(tmp_path / "case.bin"), str(tmp_path / "out.bin")
    FC.write_case(case, w, stress=stress, **ctl)
    print(FC.run_driver(case, out), flush=True)
    uf, vf, nvi, nax, res = FC.read_result(out, w.mesh.nT)
    s = V.VelocitySolver(w.descriptor(), **{k: v for k, v in ctl.items()})
    inp = w.inputs()
    args = (inp["Hi"], inp["Hs"], inp["A_flow"], inp["beta_sq"], inp["fgr_b"], inp["mask_grounded"])
    if stress == "SSA":
        u, v, nv2, na2, info = s.solve_SSA(*args)
    else:
        st, nv2, na2, info = s.solve_DIVA(*args)
        u, v = st["u_vav_b"], st["v_vav_b"]
    s.close()
    assert res <= 1e-8 and nvi == nv2
    assert max_rel(uf, u) < 1e-12 and max_rel(vf, v) < 1e-12
