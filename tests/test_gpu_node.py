"""GPU: dgeqr3's node steps for halves at most 64 wide in one launch each
(node_update64_kernel: C -= V (T1^T (V^T C)); node_join64_kernel: T12 =
-T1 (V1^T V2) T2) against the oracle's dgeqr3 and against the same
recursion with the steps as GEMM launches (lila_set_node_fused(0)).

Cases: the n = 128 node of the 64-wide panels; ragged halves (n = 65, 100);
a tall node whose workgroups take several 64-row chunks (m = 40000: 157
rows each, where the panels are 32 wide and every node of the recursion
takes the fused steps); short nodes with fewer rows than 64 workgroups
(m = 130, 200); and the recursion above (n = 512, whose nodes of 256 and
512 keep the GEMM steps).  V, R and T agree with the oracle's to multiples
of n u and the two forms agree with each other to rounding; the profile
counters confirm which kernels ran.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def host_blas(lila, ora):
    ora.use_host_blas()
    old = lila.call("lila_set_node_fused", 1)
    yield
    lila.call("lila_set_node_fused", old)
    ora.use_builtin_blas()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def counts(lila):
    buf = ctypes.create_string_buffer(1 << 16)
    lila.call("lila_prof_read", buf, len(buf))
    out = {}
    for line in buf.value.decode().strip().splitlines():
        name, n = line.split("|")[:2]
        out[name] = int(n)
    return out


def run(lila, A, fused):
    m, n = A.shape
    lila.call("lila_set_node_fused", fused)
    Ad, Td = to_dev(A), to_dev(np.zeros((n, n), order="F"))
    lila.call("lila_prof_enable", 1)
    try:
        assert lila.call("dgeqr3", m, n, Ad, m, Td, n) == 0
        torch.cuda.synchronize()
        c = counts(lila)
    finally:
        lila.call("lila_prof_enable", 0)
    assert lila.call("lila_get_info") == 0
    return to_host(Ad), to_host(Td), c


def nodes(n, width):
    """number of dgeqr3 nodes with both halves <= 64 and n > width"""
    if n <= width:
        return 0
    n1 = n // 2
    own = 1 if (n1 <= 64 and n - n1 <= 64) else 0
    return own + nodes(n1, width) + nodes(n - n1, width)


@pytest.mark.parametrize("m,n", [(16384, 128), (3000, 128), (5000, 100), (2000, 65), (40000, 128),
                                 (130, 128), (200, 96), (8192, 512)])
def test_node_steps_against_dgeqr3_and_gemm_form(lila, ora, m, n):
    say("node steps", m, n)
    A = ora.fill_normal(501 + m + n, m, n)
    To = np.zeros((n, n), order="F")
    Ao = A.copy(order="F")
    ora.call("dgeqr3", m, n, Ao, m, To, n)
    Af, Tf, cf = run(lila, A, 1)
    Ag, Tg, cg = run(lila, A, 0)
    tol = 50 * n * U64
    for Ax, Tx in ((Af, Tf), (Ag, Tg)):
        assert rel(Ax, Ao) < tol
        assert rel(np.triu(Tx), np.triu(To)) < tol
    assert rel(Af, Ag) < tol and rel(np.triu(Tf), np.triu(Tg)) < tol
    # the factorization itself, from the fused form's V, R, T
    V = np.tril(Af, -1) + np.eye(m, n)
    Qx = np.eye(m, n) - V @ (np.triu(Tf) @ V[:n].T)
    assert np.linalg.norm(Qx @ np.triu(Af[:n]) - A) / np.linalg.norm(A) < 100 * n * U64
    width = 64 if m <= 20000 else 32
    k = nodes(n, width)
    assert cf.get("node_update64_kernel", 0) == k and cf.get("node_join64_kernel", 0) == k
    assert cg.get("node_update64_kernel", 0) == 0 and cg.get("node_join64_kernel", 0) == 0

