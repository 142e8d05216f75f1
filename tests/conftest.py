import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def relmse(x, r):
    """mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2)"""
    import numpy as np
    return float(np.mean((x.astype(np.float64) - r) ** 2 / (r.astype(np.float64) ** 2 + 1e-2)))


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


@pytest.fixture(scope="session")
def c2lib():
    import crystal2_amd
    return crystal2_amd.load_library()


@pytest.fixture(scope="session")
def gpu_available():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True
