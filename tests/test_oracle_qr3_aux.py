"""CPU: the oracle's restatements of the UT / ISW / V2T / VT2Q / VS2Q forms
and the aux routines of qr3/src, pinned against LAPACK (dgeqrt's V and T,
dorgqr, dlarf), against each other where two routes reach the same result
(S^-1 = T, the V2T and recursive spine blocks, dV2N + dN2T = dlarft), and
against closed forms of the triangular products on ragged sizes."""
import numpy as np
import pytest
import scipy.linalg.lapack as sl

from conftest import U64

SIZES = [(40, 1), (50, 7), (300, 37), (128, 64), (500, 100)]


def spine(n):
    """(offset, size) of the diagonal blocks the ISW recursion writes"""
    out, off = [], 0
    while n > 1:
        n1 = n // 2
        out.append((off, n1))
        off += n1
        n -= n1
    if n == 1:
        out.append((off, 1))
    return out


def spine_mask(n):
    M = np.zeros((n, n), dtype=bool)
    for off, b in spine(n):
        M[off:off + b, off:off + b] = np.triu(np.ones((b, b), dtype=bool))
    return M


def lapack_vt(A0):
    """V (unit lower, in a), T and tau of LAPACK dgeqrt on A0"""
    m, n = A0.shape
    a, t, info = sl.dgeqrt(n, A0.copy(order="F"))
    assert info == 0
    return np.asfortranarray(a), np.asfortranarray(t), np.diag(t).copy()


def unit_lower(a, n):
    V = np.tril(a[:, :n], -1)
    V[:n, :n] += np.eye(n)
    return V


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def gaussian(seed, m, n):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal((m, n)))


# ---- dgeqr3_R forms ---------------------------------------------------------
@pytest.mark.parametrize("m,n", SIZES)
def test_dgeqr3_R_UT_is_inverse_T(ora, m, n):
    A0 = gaussian(m + 2 * n, m, n)
    a, t, _ = lapack_vt(A0)
    A, S, R = A0.copy(order="F"), np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
    assert ora.call("qr2_dgeqr3_R_UT", m, n, A, m, S, n, R, n) == 0
    tol = 100 * n * U64
    assert rel(np.tril(A, -1), np.tril(a, -1)) < tol if n > 1 or m > 1 else True
    assert rel(np.triu(R), np.triu(a[:n])) < tol
    assert rel(np.linalg.inv(np.triu(S)), np.triu(t)) < 10 * tol


@pytest.mark.parametrize("ut", [0, 1])
@pytest.mark.parametrize("m,n", SIZES)
def test_dgeqr3_R_ISW_spine(ora, ut, m, n):
    A0 = gaussian(7 * m + n, m, n)
    a, t, _ = lapack_vt(A0)
    sentinel = 123.25
    A, T, R = A0.copy(order="F"), np.full((n, n), sentinel, order="F"), np.zeros((n, n), order="F")
    name = "qr2_dgeqr3_R_UT_ISW" if ut else "qr2_dgeqr3_R_ISW"
    assert ora.call(name, m, n, A, m, T, n, R, n) == 0
    tol = 100 * n * U64
    assert rel(np.triu(R), np.triu(a[:n])) < tol
    M = spine_mask(n)
    for off, b in spine(n):
        blk = np.triu(T[off:off + b, off:off + b])
        ref = np.triu(t[off:off + b, off:off + b])
        got = np.linalg.inv(blk) if ut else blk
        assert rel(got, ref) < 10 * tol, (off, b)
    # with R apart from A nothing outside the spine blocks is written
    assert np.all(T[~M] == sentinel)


def test_dgeqr3_R_forms_agree_square(ora):
    """m == n: the last reflector is the identity (tau = 0), S's last diagonal
    entry is +inf, and V, R agree between the forms"""
    m = n = 33
    A0 = gaussian(5, m, n)
    A1, T1, R1 = A0.copy(order="F"), np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
    A2, S2, R2 = A0.copy(order="F"), np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
    ora.call("qr2_dgeqr3_R", m, n, A1, m, T1, n, R1, n)
    ora.call("qr2_dgeqr3_R_UT", m, n, A2, m, S2, n, R2, n)
    assert T1[n - 1, n - 1] == 0.0 and np.isposinf(S2[n - 1, n - 1])
    tol = 100 * n * U64
    assert rel(np.tril(A2, -1), np.tril(A1, -1)) < tol and rel(np.triu(R2), np.triu(R1)) < tol
    k = n - 1
    assert rel(np.linalg.inv(np.triu(S2[:k, :k])), np.triu(T1[:k, :k])) < 10 * tol


# ---- dlarft3 forms ----------------------------------------------------------
@pytest.mark.parametrize("m,n", SIZES)
def test_dlarft3_UT_is_inverse_T(ora, m, n):
    a, t, tau = lapack_vt(gaussian(11 * m + n, m, n))
    S = np.zeros((n, n), order="F")
    assert ora.call("qr2_dlarft3_UT", m, n, a.copy(order="F"), m, S, n, tau.copy()) == 0
    assert rel(np.linalg.inv(np.triu(S)), np.triu(t)) < 1000 * n * U64


@pytest.mark.parametrize("name,ut", [("qr2_dlarft3_ISW", 0), ("qr2_dlarft3_UT_ISW", 1),
                                     ("qr2_dlarft3_ISW_V2T", 0), ("qr2_dlarft3_ISW_V2T_UT", 1)])
@pytest.mark.parametrize("m,n", SIZES)
def test_dlarft3_ISW_spine(ora, name, ut, m, n):
    a, t, tau = lapack_vt(gaussian(13 * m + n, m, n))
    sentinel = -7.5
    T = np.full((n, n), sentinel, order="F")
    assert ora.call(name, m, n, a.copy(order="F"), m, T, n, tau.copy()) == 0
    M = spine_mask(n)
    assert np.all(T[~M] == sentinel)
    for off, b in spine(n):
        blk = np.triu(T[off:off + b, off:off + b])
        got = np.linalg.inv(blk) if ut else blk
        assert rel(got, np.triu(t[off:off + b, off:off + b])) < 1000 * n * U64, (off, b)


# ---- dorgqr3 forms and dVT2Q / dVS2Q -----------------------------------------
@pytest.mark.parametrize("m,n", SIZES + [(64, 64)])
def test_dorgqr3_forms_match_dorgqr(ora, m, n):
    a, t, tau = lapack_vt(gaussian(17 * m + n, m, n))
    qref, _, info = sl.dorgqr(a.copy(order="F"), tau)
    assert info == 0
    tol = 100 * n * U64
    S = np.zeros((n, n), order="F")
    ora.call("qr2_dlarft3_UT", m, n, a.copy(order="F"), m, S, n, tau.copy())
    for name, TT in (("qr2_dorgqr3_VT2Q", t), ("qr2_dorgqr3_UT", S), ("qr2_dorgqr3_VS2Q", S)):
        if name != "qr2_dorgqr3_VT2Q" and m == n:
            continue  # S(n-1, n-1) = inf: covered by the square test below
        Q = a.copy(order="F")
        assert ora.call(name, m, n, Q, m, TT.copy(order="F"), n, tau.copy()) == 0
        assert rel(Q, qref) < tol, name


def test_dorgqr3_UT_square(ora):
    m = n = 40
    a, t, tau = lapack_vt(gaussian(3, m, n))
    qref, _, _ = sl.dorgqr(a.copy(order="F"), tau)
    S = np.zeros((n, n), order="F")
    ora.call("qr2_dlarft3_UT", m, n, a.copy(order="F"), m, S, n, tau.copy())
    assert np.isposinf(S[n - 1, n - 1])
    for name in ("qr2_dorgqr3_UT", "qr2_dorgqr3_VS2Q"):
        Q = a.copy(order="F")
        ora.call(name, m, n, Q, m, S.copy(order="F"), n, tau.copy())
        assert rel(Q, qref) < 100 * n * U64, name


@pytest.mark.parametrize("m,n", [(30, 1), (50, 7), (300, 37), (64, 64), (200, 120)])
def test_aux_VT2Q_VS2Q_match_dorgqr(ora, m, n):
    a, t, tau = lapack_vt(gaussian(19 * m + n, m, n))
    qref, _, _ = sl.dorgqr(a.copy(order="F"), tau)
    Q = a.copy(order="F")
    Q[:n, :n] = np.tril(Q[:n, :n], -1) + np.triu(t)
    assert ora.call("qr2_aux_dVT2Q", m, n, Q, m) == 0
    assert rel(Q, qref) < 100 * n * U64
    if m > n:
        Q = a.copy(order="F")
        Q[:n, :n] = np.tril(Q[:n, :n], -1) + np.triu(np.linalg.inv(np.triu(t)))
        assert ora.call("qr2_aux_dVS2Q", m, n, Q, m) == 0
        assert rel(Q, qref) < 1000 * n * U64


def test_planted_transposed_T_fails(ora):
    """the dorgqr pin sees a transposed T handed to dVT2Q"""
    m, n = 120, 24
    a, t, tau = lapack_vt(gaussian(23, m, n))
    qref, _, _ = sl.dorgqr(a.copy(order="F"), tau)
    Q = a.copy(order="F")
    Q[:n, :n] = np.tril(Q[:n, :n], -1) + np.triu(t.T)  # wrong triangle
    ora.call("qr2_aux_dVT2Q", m, n, Q, m)
    assert rel(Q, qref) > 1e-3


# ---- aux: triangular products (closed forms, ragged sizes) -------------------
@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 101])
def test_aux_mLUinA(ora, n):
    A0 = gaussian(n, n, n)
    A = A0.copy(order="F")
    assert ora.call("qr2_aux_dmLUinA", n, A, n) == 0
    L = np.tril(A0, -1) + np.eye(n)
    assert rel(A, -(L @ np.triu(A0))) < 50 * n * U64


@pytest.mark.parametrize("alias", [True, False])
@pytest.mark.parametrize("n", [1, 2, 5, 33, 90])
def test_aux_ULTinU_and_UinvLTinU(ora, n, alias):
    U0 = gaussian(2 * n + 1, n, n) + 4 * np.eye(n)
    L0 = U0 if alias else gaussian(3 * n, n, n)
    Lu = np.tril(L0, -1) + np.eye(n)
    for name, ref in (("qr2_aux_dULTinU", np.triu(U0) @ Lu.T),
                      ("qr2_aux_dUinvLTinU", np.linalg.solve(np.triu(U0), Lu.T))):
        U = U0.copy(order="F")
        L = U if alias else L0.copy(order="F")
        assert ora.call(name, n, L, n, U, n) == 0
        assert rel(np.triu(U), ref) < 100 * n * U64, name
        assert np.array_equal(np.tril(U, -1), np.tril(U0, -1)), name


def test_aux_UinvLTinU_of_S_is_ULTinU_of_T(ora):
    """S = T^-1: S^-1 V1^T = T V1^T, the step dVS2Q and dVT2Q share"""
    m, n = 200, 48
    a, t, tau = lapack_vt(gaussian(29, m, n))
    S = np.linalg.inv(np.triu(t))
    X1 = np.asfortranarray(np.tril(a[:n], -1) + np.triu(t))
    X2 = np.asfortranarray(np.tril(a[:n], -1) + np.triu(S))
    ora.call("qr2_aux_dULTinU", n, X1, n, X1, n)
    ora.call("qr2_aux_dUinvLTinU", n, X2, n, X2, n)
    assert rel(np.triu(X2), np.triu(X1)) < 1000 * n * U64


@pytest.mark.parametrize("m,n", [(1, 1), (1, 5), (6, 1), (7, 4), (4, 9), (33, 20), (64, 64)])
def test_aux_ApUBTinA(ora, m, n):
    A0, U0, B0 = gaussian(m, m, n), gaussian(m + 1, m, m), gaussian(n + 2, n, m)
    A = A0.copy(order="F")
    assert ora.call("qr2_aux_dApUBTinA", m, n, A, m, U0.copy(order="F"), m, B0.copy(order="F"), n) == 0
    assert rel(A, A0 + np.triu(U0) @ B0.T) < 50 * max(m, n) * U64


# ---- aux: T from V (dV2N, dN2T, dV2tau) ---------------------------------------
@pytest.mark.parametrize("n", [1, 2, 9, 40, 77])
def test_aux_V2N_N2T_give_dlarft_T(ora, n):
    """strict upper part <- V1^T, dV2N, the Gram of the rows below, dN2T:
    LAPACK's T (the V2T route of qr2_dlarft3_ISW_V2T on the whole block)"""
    m = 3 * n + 5
    a, t, tau = lapack_vt(gaussian(31 * n, m, n))
    V = unit_lower(a, n)
    T = np.zeros((n, n), order="F")
    T[np.triu_indices(n, 1)] = V[:n, :n].T[np.triu_indices(n, 1)]
    ora.call("qr2_aux_dV2N", n, T, n)
    assert rel(np.triu(T), np.triu(V[:n, :n].T @ V[:n, :n])) < 50 * n * U64
    T = np.asfortranarray(np.triu(T) + np.triu(V[n:].T @ V[n:]))
    ora.call("qr2_aux_dN2T", n, tau.copy(), T, n)
    assert rel(np.triu(T), np.triu(t)) < 1000 * n * U64


@pytest.mark.parametrize("m,n", [(50, 7), (300, 37), (64, 64), (10, 1)])
def test_aux_V2tau(ora, m, n):
    a, t, tau = lapack_vt(gaussian(37 * m + n, m, n))
    got = np.zeros(n)
    assert ora.call("qr2_aux_dV2tau", m, n, a.copy(order="F"), m, got) == 0
    assert np.allclose(got, tau, rtol=100 * m * U64, atol=0)
    if m == n:
        assert got[-1] == 0.0


# ---- aux: dlarf ---------------------------------------------------------------
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("incv", [1, 3])
@pytest.mark.parametrize("m,n", [(1, 1), (17, 5), (40, 60)])
def test_aux_dlarf_matches_lapack(ora, side, incv, m, n):
    rng = np.random.default_rng(m * n + incv)
    L = m if side == "L" else n
    v = rng.standard_normal(L)
    tau = 0.7
    C0 = np.asfortranarray(rng.standard_normal((m, n)))
    ref = sl.dlarf(v, tau, C0.copy(order="F"), np.zeros(max(m, n)), side=side)
    Vs = np.zeros((L - 1) * incv + 1)
    Vs[::incv] = v
    C = C0.copy(order="F")
    assert ora.call("qr2_aux_dlarf_wrapper", ord(side), m, n, Vs, incv, tau, C, m, None) == 0
    assert rel(C, ref) < 50 * max(m, n) * U64
    assert ora.call("qr2_aux_dlarf_wrapper", ord(side), m, n, Vs, 0, tau, C, m, None) == -1
