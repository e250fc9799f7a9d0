"""Writes reference_cases.json: the known-answer data that the reference's own
tests and integrated-test configs hold for the velocity solve and the heat
columns, transcribed as data (no reference code is run or copied).

  ut_petsc_6x6      UPSY/validation/unit_tests/ut_petsc.f90:83-150
  tridiag_known     UPSY/validation/unit_tests/ut_math_utilities.f90:65-110
  schoof_icestream  automated_testing/integrated_tests/idealised/SSA_icestream/
                    config_0*.cfg (A, H, dhdx, L, m) and the RMSE values that the
                    reference scoreboard reports for its meshes
                    (analyse_integrated_test.m; 400.4/303.7/151.9/101.2 m/yr)

Run: python tests/golden/make_golden.py
"""

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

cases = {
    "ut_petsc_6x6": {
        # 1-based (row, col, value) of the 6x6 matrix, x and the expected y
        "entries": [[1, 1, 1.0], [2, 1, 2.0], [2, 2, 3.0], [3, 2, 4.0], [3, 4, 5.0], [4, 4, 6.0],
                    [4, 5, 7.0], [5, 5, 8.0], [6, 5, 9.0], [6, 6, 10.0]],
        "x": [1.0, 2.0, 3.0, 4.0, 5.0, 6.0],
        "y": [1.0, 8.0, 28.0, 59.0, 40.0, 105.0],
    },
    "tridiag_known": {
        # d2f/dx2 = 2 on [-1, 1], f(-1) = f(1) = 1 -> f = x^2, checked at the
        # ends and the midpoint to 1e-12 for n = 2^ni + 1
        "ni": list(range(3, 13)),
        "tol": 1e-12,
    },
    "schoof_icestream": {
        "A": 1e-18, "H": 2000.0, "dhdx": -3e-4, "L": 150e3, "m": 1.0, "n": 3.0,
        "epsilon_sq_0": 1e-10, "slid_delta_v": 1e-3, "visc_it_relax": 0.3,
        "visc_it_norm_dUV_tol": 5e-8, "visc_it_nit": 5000,
        "domain": [-400e3, 400e3, -400e3, 400e3],
        "bc_u": {"north": "zero", "south": "zero", "east": "infinite", "west": "infinite"},
        "bc_v": "zero",
        "reference_rmse": {"32km": 400.4, "16km": 303.7, "8km": 151.9, "4km": 101.2},
    },
}

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)
    print("wrote", os.path.join(HERE, "reference_cases.json"))
