"""GPU: error paths and ordering of the C ABI.

* a host operand whose staging buffer cannot be allocated: the call returns
  -2, launches nothing and leaves the library usable;
* lila_set_stream orders the switch: two asynchronous device-pointer calls
  on two streams (sharing the library's workspace, panel scratch and
  barrier counters) give the one-stream results;
* the mixed-precision form on rank-deficient matrices (a zero column,
  duplicated columns), where its low-precision R0 is singular: it falls
  back to the FP64 recursion and returns a valid factorization;
* dense entry points refuse hipGraph capture (-6) instead of recording a
  graph that would replay into freed workspace.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu


def test_failed_staging_returns_minus_two(lila, ora):
    """lda = 2^31 - 1 over 200 columns asks for ~3.4 TB of staging: the
    allocation fails before anything is read from the (small) host array"""
    A = np.zeros((8, 8), order="F")
    T = np.zeros((8, 8), order="F")
    say("staging failure")
    r = lila.call("dgeqr3", 300, 200, A, 2 ** 31 - 1, T, 8)
    assert r == -2, r
    assert lila.call("lila_get_info") == 0
    # the library still works afterwards, host and device pointers alike
    m, n = 200, 20
    X = ora.fill_uniform(5, m, n, -1.0, 1.0)
    Xo, To = X.copy(order="F"), np.zeros((n, n), order="F")
    ora.call("dgeqr3", m, n, Xo, m, To, n)
    Xh, Th = X.copy(order="F"), np.zeros((n, n), order="F")
    assert lila.call("dgeqr3", m, n, Xh, m, Th, n) == 0
    assert np.linalg.norm(Xh - Xo) / np.linalg.norm(Xo) < 100 * n * U64


def _vrtq(lila, A, mt, nx, mode=0):
    m, n = A.shape
    Ad = to_dev(A)
    Td = to_dev(np.zeros((mt, n), order="F"))
    Qd = to_dev(np.zeros((m, n), order="F"))
    p = np.array([mode, 1, 1, nx], dtype=np.int32)
    return p, Ad, Td, Qd


def test_stream_switch_is_ordered(lila, ora):
    """call 1 on stream s1, switch, call 2 on stream s2 without waiting; each
    result equals the same call run alone on the default stream"""
    m, n, mt, nx = 3000, 700, 128, 64
    A1 = ora.fill_uniform(61, m, n, -1.0, 1.0)
    A2 = ora.fill_uniform(62, m, n, -1.0, 1.0)
    ref = []
    for A in (A1, A2):
        p, Ad, Td, Qd = _vrtq(lila, A, mt, nx)
        assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
        torch.cuda.synchronize()
        ref.append((to_host(Ad), to_host(Qd)))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    bufs = [_vrtq(lila, A, mt, nx) for A in (A1, A2)]
    torch.cuda.synchronize()
    say("two streams")
    try:
        for s, (p, Ad, Td, Qd) in zip((s1, s2), bufs):
            lila.call("lila_set_stream", ctypes.c_void_p(s.cuda_stream))
            assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, Ad, m, Td, mt, Qd, m, None,
                             0) == 0
    finally:
        lila.call("lila_set_stream", None)
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    for (p, Ad, Td, Qd), (Ar, Qr) in zip(bufs, ref):
        Ag, Qg = to_host(Ad), to_host(Qd)
        # the panels' atomic sums vary in the last bits between runs
        assert np.linalg.norm(Ag - Ar) / np.linalg.norm(Ar) < 100 * n * U64
        assert np.linalg.norm(Qg - Qr) / np.linalg.norm(Qr) < 100 * n * U64


@pytest.mark.parametrize("kind", ["zero_column", "duplicate_columns", "rank_half"])
def test_mixed_falls_back_on_rank_deficient_input(lila, ora, kind):
    m, n, mt, nx = 900, 400, 64, 32
    A = ora.fill_normal(71, m, n)
    if kind == "zero_column":
        A[:, 37] = 0.0
    elif kind == "duplicate_columns":
        A[:, 100:140] = A[:, 10:50]
    else:
        A = np.asfortranarray(A[:, :n // 2] @ ora.fill_normal(72, n // 2, n))
    A = np.asfortranarray(A)
    say("mixed rank deficient", kind)
    p, Ad, Td, Qd = _vrtq(lila, A, mt, nx, mode=1)
    assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    Ag, Qg = to_host(Ad), to_host(Qd)
    assert np.all(np.isfinite(Ag)) and np.all(np.isfinite(Qg))
    tol = 100 * n * U64
    assert ora.call("test_qq_orth_full", m, n, 0, Qg, m) < tol
    assert ora.call("test_qr_repres_1", m, n, 0, A, m, Qg, m, Ag, m) < tol


def test_dense_entry_refuses_graph_capture(lila):
    m, n = 256, 32
    A = torch.zeros((n, m), dtype=torch.float64, device="cuda")
    T = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    x = torch.zeros(4, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    say("graph capture refusal")
    lila.call("lila_set_stream", ctypes.c_void_p(s.cuda_stream))
    try:
        with torch.cuda.graph(g, stream=s):
            x.add_(1.0)
            r = lila.call("dgeqr3", m, n, A, m, T, n)
    finally:
        lila.call("lila_set_stream", None)
    assert r == -6, r
    g.replay()
    torch.cuda.synchronize()
    assert x[0].item() == 1.0
