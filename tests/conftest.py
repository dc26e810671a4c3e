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


@pytest.fixture(autouse=True)
def no_pending_hip_error(request):
    """a GPU test must leave no pending HIP error behind (one left by a
    failed runtime call would surface at the next test's first torch call)"""
    yield
    if request.node.get_closest_marker("gpu") is None:
        return
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return
    err = hip.hipGetLastError()
    assert err == 0, "pending HIP error %d after %s" % (err, request.node.nodeid)


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
    h = torch.from_numpy(np.ascontiguousarray(A.T))
    if h.numel() * 8 > (4 << 20):
        # large copies go through a pinned host buffer (allocated pinned, not
        # registered after the fact)
        p = torch.empty(h.shape, dtype=h.dtype, pin_memory=True)
        p.copy_(h)
        h = p
    return h.cuda()


def to_host(t):
    """inverse of to_dev"""
    import torch
    if t.numel() * 8 > (4 << 20):
        p = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        p.copy_(t)
        return np.asfortranarray(p.numpy().T.copy())
    return np.asfortranarray(t.cpu().numpy().T)


def say(*a):
    print(*a, flush=True)
