"""CPU: the oracle's restatements of the recursive QR library of qr3/src
(qr2_dgeqr3_R, qr2_dlarft3, qr2_dorgqr3) pinned against LAPACK dgeqrt and
dorgqr and the identities A = QR, Q^T Q = I."""
import numpy as np
import pytest
import scipy.linalg.lapack as sl

from conftest import U64

SIZES = [(40, 1), (50, 7), (300, 37), (128, 64), (64, 64), (500, 100)]


@pytest.mark.parametrize("m,n", SIZES)
def test_qr2_dgeqr3_R_matches_dgeqrt(ora, m, n):
    rng = np.random.default_rng(m + n)
    A0 = np.asfortranarray(rng.standard_normal((m, n)))
    A = A0.copy(order="F")
    T = np.zeros((n, n), order="F")
    R = np.zeros((n, n), order="F")
    assert ora.call("qr2_dgeqr3_R", m, n, A, m, T, n, R, n) == 0
    a, t, info = sl.dgeqrt(n, A0.copy(order="F"))
    assert info == 0
    nA = np.linalg.norm(A0)
    tol = 50 * n * U64
    # V below the diagonal, R apart (A's diagonal and upper part hold the
    # entries as they stood before each column's reflector: not an output)
    assert np.linalg.norm(np.tril(A, -1) - np.tril(a, -1)) <= tol * max(1.0, np.linalg.norm(np.tril(a, -1)))
    assert np.linalg.norm(np.triu(R) - np.triu(a[:n])) <= tol * nA
    # the first row of A is never written (A11's leaf keeps alpha, A12 is read)
    assert np.array_equal(A[0], A0[0])
    assert np.linalg.norm(np.triu(T) - np.triu(t)) <= tol * max(1.0, np.linalg.norm(t))


@pytest.mark.parametrize("m,n", [(300, 37), (64, 64)])
def test_qr2_dgeqr3_R_aliased_is_dgeqr3(ora, m, n):
    rng = np.random.default_rng(3 * m + n)
    A0 = np.asfortranarray(rng.standard_normal((m, n)))
    A = A0.copy(order="F")
    T = np.zeros((n, n), order="F")
    ora.call("qr2_dgeqr3_R", m, n, A, m, T, n, A, m)
    B = A0.copy(order="F")
    T2 = np.zeros((n, n), order="F")
    ora.call("dgeqr3", m, n, B, m, T2, n)
    assert np.linalg.norm(A - B) <= 50 * n * U64 * np.linalg.norm(B)
    assert np.linalg.norm(np.triu(T) - np.triu(T2)) <= 50 * n * U64 * np.linalg.norm(T2)


@pytest.mark.parametrize("m,n", SIZES)
def test_qr2_dlarft3_matches_dgeqrt_T(ora, m, n):
    rng = np.random.default_rng(7 * m + n)
    a, t, info = sl.dgeqrt(n, np.asfortranarray(rng.standard_normal((m, n))))
    tau = np.ascontiguousarray(np.diag(t))
    T = np.zeros((n, n), order="F")
    assert ora.call("qr2_dlarft3", m, n, np.asfortranarray(a), m, T, n, tau) == 0
    assert np.linalg.norm(np.triu(T) - np.triu(t)) <= 50 * n * U64 * max(1.0, np.linalg.norm(t))


@pytest.mark.parametrize("m,n", SIZES)
def test_qr2_dorgqr3_matches_dorgqr(ora, m, n):
    rng = np.random.default_rng(11 * m + n)
    A0 = np.asfortranarray(rng.standard_normal((m, n)))
    a, t, info = sl.dgeqrt(n, A0.copy(order="F"))
    tau = np.ascontiguousarray(np.diag(t))
    Q = np.asfortranarray(a.copy())
    assert ora.call("qr2_dorgqr3", m, n, Q, m, np.asfortranarray(t), n, tau) == 0
    q, _, info = sl.dorgqr(np.asfortranarray(a.copy()), tau)
    assert info == 0
    assert np.linalg.norm(Q - q) <= 50 * n * U64 * np.sqrt(n)
    # closed forms: Q^T Q = I and Q R = A
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) <= 50 * n * U64
    assert np.linalg.norm(Q @ np.triu(a[:n]) - A0) <= 50 * n * U64 * np.linalg.norm(A0)


def test_qr2_dorgqr3_planted_transposed_T_fails(ora):
    """the Q check above rejects a T given transposed (the lower triangle)"""
    m, n = 200, 24
    rng = np.random.default_rng(5)
    a, t, _ = sl.dgeqrt(n, np.asfortranarray(rng.standard_normal((m, n))))
    tau = np.ascontiguousarray(np.diag(t))
    Q = np.asfortranarray(a.copy())
    ora.call("qr2_dorgqr3", m, n, Q, m, np.asfortranarray(t.T.copy()), n, tau)
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) > 1e-3
