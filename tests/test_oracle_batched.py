"""CPU: the oracle's per-problem least squares (LAPACK dgels('N') as dgeqr2 +
dorm2r + back substitution) pinned against SciPy's LAPACK dgels, the normal
equations and the residual identities that fix the answer."""
import numpy as np
import pytest
import scipy.linalg.lapack as sl

from conftest import U64


CASES = [(64, 8, 1), (100, 100, 3), (257, 31, 16), (300, 128, 5), (77, 1, 2), (64, 17, 0), (513, 64, 9)]


@pytest.mark.parametrize("m,n,k", CASES)
def test_dgels_hh_matches_lapack(ora, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A0 = np.asfortranarray(rng.standard_normal((m + 3, n)))  # lda = m + 3
    B0 = np.asfortranarray(rng.standard_normal((m + 2, max(k, 1))))
    A, B = A0.copy(order="F"), B0.copy(order="F")
    tau = np.zeros(n)
    assert ora.call("dgels_hh", m, n, k, A, m + 3, B, m + 2, tau) == 0
    # rows beyond m (the lda / ldb padding) untouched
    assert np.array_equal(A[m:], A0[m:]) and np.array_equal(B[m:], B0[m:])
    if k == 0:
        assert np.array_equal(B, B0)
    a_ref, b_ref = A0[:m].copy(order="F"), B0[:m, :max(k, 1)].copy(order="F")
    if k > 0:
        lqr, x, info = sl.dgels(a_ref, b_ref[:, :k])
        assert info == 0
        nA = np.linalg.norm(A0[:m])
        # same factor (dlarfg convention: R and V entrywise)
        assert np.max(np.abs(A[:m] - lqr)) <= 50 * n * U64 * nA
        xs = np.linalg.lstsq(A0[:m], B0[:m, :k], rcond=None)[0]
        assert np.max(np.abs(B[:n, :k] - xs)) <= 1e-9 * max(1.0, np.max(np.abs(xs)))
        assert np.max(np.abs(B[:m, :k] - x)) <= 1e-9 * max(1.0, np.max(np.abs(x)))
        # residual: rows n..m of Q^T b carry ||A x - b||; A^T r = 0
        r = A0[:m] @ B[:n, :k] - B0[:m, :k]
        assert np.allclose(np.linalg.norm(r, axis=0), np.linalg.norm(B[n:m, :k], axis=0), rtol=1e-10, atol=1e-12)
        orth = np.linalg.norm(A0[:m].T @ r) / (nA * (nA * np.linalg.norm(B[:n, :k]) + np.linalg.norm(B0[:m, :k])))
        assert orth <= 10 * n * U64
    # tau from V: tau = 2 / (1 + ||v||^2) unless the column was already reduced
    for c in range(n):
        if c + 1 >= m:  # nothing below the diagonal: H = I
            assert tau[c] == 0.0
            continue
        v2 = 1.0 + np.sum(A[c + 1:m, c] ** 2)
        assert abs(tau[c] - 2.0 / v2) <= 1e-13


def test_dgels_hh_planted_mistake_is_caught(ora):
    """the entrywise factor check above fails for a factor with one reflector skipped"""
    m, n = 90, 12
    rng = np.random.default_rng(5)
    A0 = np.asfortranarray(rng.standard_normal((m, n)))
    A = A0.copy(order="F")
    tau = np.zeros(n)
    ora.call("dgels_hh", m, n, 0, A, m, np.zeros((m, 1), order="F"), m, tau)
    lqr, _, _ = sl.dgels(A0.copy(order="F"), np.zeros((m, 1), order="F"))
    bad = A.copy(order="F")
    bad[3:, 3] *= -1.0  # a reflector with its sign flipped
    nA = np.linalg.norm(A0)
    assert np.max(np.abs(A - lqr)) <= 50 * n * U64 * nA
    assert np.max(np.abs(bad - lqr)) > 50 * n * U64 * nA
