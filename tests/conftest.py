"""Shared fixtures.  `-m "not gpu"` runs the oracle/host/ABI checks here; `-m gpu` runs
the parity tests on an MI355X through the C ABI."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libplbm.so")


@pytest.fixture(scope="session")
def P():
    """The engine package (periodic-lbm_amd/), loaded as periodic_lbm_amd."""
    import __graft_entry__ as g
    return g.load_package()


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def ref():
    import pyoracle
    if not pyoracle.have_ref():
        pytest.skip("oracle/_ref not built (no reference tree at build time)")
    return pyoracle.RefLib()


@pytest.fixture(scope="session")
def golden_small():
    return np.load(GOLDEN / "ref_small.npz")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch.device("cuda:0")


def tgv_fields(P, n, u0, nu, t=0.0):
    """Taylor-Green initial macroscopic fields as the reference app builds them
    (app/main_taylor_green.f90:133-149): rho = p/csqr + 1."""
    k = 2 * np.pi / n
    p, ux, uy = P.TaylorGreen(n, n, k, k, u0, nu).eval(t)
    return np.asfortranarray(p / P.csqr + 1.0), ux, uy
