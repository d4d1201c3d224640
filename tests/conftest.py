import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import _pkg  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: longer GPU parity runs")


@pytest.fixture(scope="session")
def T():
    """The product package (topopt.jl_amd)."""
    return _pkg.load_package()


@pytest.fixture(scope="session")
def O():
    """The oracle (CPU restatement); built on first use."""
    o = _pkg.load_oracle()
    o.build()
    return o


def gpu_available():
    """GPU present (asked through the library's own HIP runtime, not torch)."""
    try:
        from importlib import import_module
        T = _pkg.load_package()
        return import_module("topopt_jl_amd._lib").device_info()[0] > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    return True


def rng(seed):
    return np.random.default_rng(seed)


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
