import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

U64 = np.finfo(np.float64).eps / 2  # unit roundoff 1.1e-16


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ora():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def lila():
    """the product C ABI (liblila_hip.so) with a visible GPU"""
    import torch
    import lila_amd
    assert torch.cuda.is_available(), "GPU tests need a visible HIP device"
    lib = lila_amd.load()
    return lila_amd


def to_dev(A):
    """numpy Fortran-ordered (ld x n) -> torch cuda tensor with the same
    column-major layout (shape (n, ld), row c = column c)"""
    import torch
    A = np.asfortranarray(A)
    return torch.from_numpy(np.ascontiguousarray(A.T)).cuda()


def to_host(t):
    """inverse of to_dev"""
    return np.asfortranarray(t.cpu().numpy().T)


def say(*a):
    print(*a, flush=True)
