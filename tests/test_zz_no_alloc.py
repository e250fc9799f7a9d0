"""The allocation-free heat step (ufe_heat_columns_device_noalloc, scratch in
the library's static device array) against ufe_heat_columns_device on the
same device-resident inputs: two calls in a row, the second smaller, so the
second call runs on scratch the first one left behind."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def device_case(V, W, nV, nz, seed):
    from ufemism2_amd.devmem import DeviceArray
    c = W.heat_columns_inputs(nV, nz, seed=seed, gl=0.2)
    f = dict(c["fields"])
    f["T_surf"] = f["T_surf"].copy()
    f["Hi_eff"] = f["Hi_eff"].copy()
    f["T_surf"][:3] = 150.0          # three columns pinned below 180 K: unstable
    f["Hi_eff"][:3] = 1500.0
    dev = {k: DeviceArray.from_host(v) for k, v in f.items()}
    dev["zeta"] = DeviceArray.from_host(c["zeta"])
    ptrs = {k: d.ptr for k, d in dev.items()}
    hin = V.make_heat_inputs_device(nV, nz, ptrs, c["dt"])
    return hin, dev


def run(V, hin, nV, nz, noalloc):
    from ufemism2_amd.devmem import DeviceArray
    Tout = DeviceArray(nV * nz, np.float64)
    un = DeviceArray(nV, np.int32)
    Tout.upload(np.full(nV * nz, -1.0))
    if noalloc:
        n = V.heat_columns_device_noalloc(hin, Tout.ptr, un.ptr)
    else:
        n = V.heat_columns_device(hin, Tout.ptr, un.ptr)
    return Tout.to_host(), un.to_host(), n


def test_noalloc_heat_step_equals_device_entry_point(V, W):
    for nV, nz, seed in ((60000, 12, 5), (7000, 10, 6)):
        print(f"noalloc heat step: nV={nV} nz={nz} seed={seed}", flush=True)
        hin, keep = device_case(V, W, nV, nz, seed)
        Ta, ua, na = run(V, hin, nV, nz, noalloc=False)
        Tb, ub, nb = run(V, hin, nV, nz, noalloc=True)
        assert na == nb and na >= 3
        assert np.array_equal(ua, ub)
        assert np.array_equal(Ta, Tb)


def test_noalloc_heat_step_rejects_too_many_layers(V, W):
    hin, keep = device_case(V, W, 100, 12, 7)
    hin.nz = 40
    with pytest.raises(V.VelocitySolverError, match="nz must be"):
        run(V, hin, 100, 12, noalloc=True)
