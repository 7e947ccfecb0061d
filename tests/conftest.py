"""Shared test setup: import paths, the `gpu` marker, and small helpers.

`-m "not gpu"` runs the checker (CPU oracle) against analytic answers and golden fixtures, the host
logic and the C-ABI symbol table; `-m gpu` runs the HIP path against the checker on a real GPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cosserat-rod-solver-cuda_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU and the built extension")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def rs():
    """The HIP package; on a GPU box a missing extension is an error, never a fallback."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test without a GPU")
    import ropesim
    return ropesim


@pytest.fixture(params=["oracle", pytest.param("hip", marks=pytest.mark.gpu),
                        pytest.param("hip-ieee", marks=pytest.mark.gpu)])
def solver(request):
    """The analytic checks run on the CPU oracle and, on a GPU, on the HIP path in both arithmetic modes
    (tests/hip_adapter.py gives it the oracle's test surface): "hip" the default fast arithmetic, "hip-ieee"
    the correctly rounded one."""
    if request.param == "oracle":
        from oracle import pyoracle
        pyoracle.lib()
        return pyoracle
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test without a GPU")
    import hip_adapter
    return hip_adapter if request.param == "hip" else hip_adapter.IEEE


def fp32(solver, prec=64):
    """True when the run computes in float32 (the device, or the oracle at prec 32)."""
    return getattr(solver, "FP32", False) or prec == 32


def to_dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")


def host(view):
    import torch
    return torch.as_tensor(view).cpu().numpy()
