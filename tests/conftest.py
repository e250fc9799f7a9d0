"""Shared fixtures.  `gpu` tests need an MI355X and the built libufe_hip.so;
everything else runs on the CPU (oracle, host code, library loading)."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

ge.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU and the built HIP library")


@pytest.fixture(scope="session")
def pkg():
    import ufemism2_amd
    return ufemism2_amd


@pytest.fixture(scope="session")
def L():
    from ufemism2_amd import _lib
    return _lib


@pytest.fixture(scope="session")
def W():
    from ufemism2_amd import workloads
    return workloads


@pytest.fixture(scope="session")
def OB():
    from oracle import binding
    return binding


@pytest.fixture(scope="session")
def V():
    from ufemism2_amd import velocity
    return velocity


@pytest.fixture(scope="session")
def dome(W):
    return W.dome(h=60e3)


@pytest.fixture(scope="session")
def chan(W):
    return W.channel_small(target_nV=3000)


def params(L, **kw):
    p = L.default_params()
    for k, v in kw.items():
        if k in ("bc_u", "bc_v"):
            for i, x in enumerate(v):
                getattr(p, k)[i] = int(x)
        else:
            setattr(p, k, v)
    return p


def max_rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
