"""Pins the CPU oracle (oracle/lila_oracle.c) before it is trusted as the
checker of the GPU path.  The reference holds no golden vectors or
fixtures for this path (its drivers print residuals only), so each oracle
function is pinned by
  * the host LAPACK (SciPy's OpenBLAS build: dlarfg, dgeqrf, dgeqrt, dorgqr,
    dormqr, dlarft, dpotrf, dorhr_col, dgels) -- an independent
    implementation of the same published algorithms,
  * closed forms the mathematics fixes (A = QR, Q^T Q = I, T block-diagonal
    with each block = dlarft of its columns, LU = A D + I for dgetrf_b,
    (I - V T V^T)[I;0] S = Q for the reconstruction),
  * and planted mistakes (in the oracle's inputs, on the CPU) that must fail.
"""
import ctypes
import glob
import os

import numpy as np
import pytest
import scipy.linalg.lapack as L

from conftest import U64

OB = None


def openblas():
    global OB
    if OB is None:
        import scipy
        path = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs",
                                      "libscipy_openblas*.so"))[0]
        OB = ctypes.CDLL(path)
    return OB


def _i(v):
    return ctypes.byref(ctypes.c_int(v))


def lapack_dlarft(V, tau):
    n, k = V.shape
    V = np.asfortranarray(V)
    T = np.zeros((k, k), order="F")
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    openblas().scipy_dlarft_(ctypes.byref(ctypes.c_char(b"F")), ctypes.byref(ctypes.c_char(b"C")),
                             _i(n), _i(k), ctypes.c_void_p(V.ctypes.data), _i(n),
                             ctypes.c_void_p(tau.ctypes.data), ctypes.c_void_p(T.ctypes.data), _i(k),
                             ctypes.c_size_t(1), ctypes.c_size_t(1))
    return T


def lapack_dorhr_col(Q):
    m, n = Q.shape
    A = np.asfortranarray(Q.copy())
    T = np.zeros((n, n), order="F")
    D = np.zeros(n)
    info = ctypes.c_int(0)
    openblas().scipy_dorhr_col_(_i(m), _i(n), _i(n), ctypes.c_void_p(A.ctypes.data), _i(m),
                                ctypes.c_void_p(T.ctypes.data), _i(n), ctypes.c_void_p(D.ctypes.data),
                                ctypes.byref(info))
    assert info.value == 0
    return A, T, D


def lapack_qr(A):
    qr, tau, _, info = L.dgeqrf(np.asfortranarray(A))
    assert info == 0
    return qr, tau


def lapack_q(qr, tau, n):
    q, _, info = L.dorgqr(np.asfortranarray(qr[:, :n]), tau)
    assert info == 0
    return q


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------------------
# LAPACK subset restated in the oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("scale", [1.0, 1e-300, 1e300, 1e-160])
@pytest.mark.parametrize("n", [1, 2, 7, 100])
def test_dlarfg_matches_lapack(ora, n, scale):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) * scale
    a1, x1, t1 = L.dlarfg(n, x[0], x[1:].copy())
    alpha = ctypes.c_double(x[0])
    xo = np.ascontiguousarray(x[1:].copy()) if n > 1 else np.zeros(1)
    tau = ctypes.c_double(0)
    ora.lib().ora_dlarfg(ctypes.c_int(n), ctypes.byref(alpha), ctypes.c_void_p(xo.ctypes.data),
                         ctypes.c_int(1), ctypes.byref(tau))
    assert abs(tau.value - t1) <= 4 * U64 * abs(t1)
    assert abs(alpha.value - a1) <= 4 * U64 * abs(a1)
    if n > 1:
        assert np.max(np.abs(xo - x1)) <= 8 * U64 * np.max(np.abs(x1)) + 1e-300


def test_dlarfg_zero_tail(ora):
    xo = np.zeros(4)
    alpha, tau = ctypes.c_double(-3.0), ctypes.c_double(7.0)
    ora.lib().ora_dlarfg(ctypes.c_int(5), ctypes.byref(alpha), ctypes.c_void_p(xo.ctypes.data),
                         ctypes.c_int(1), ctypes.byref(tau))
    assert tau.value == 0.0 and alpha.value == -3.0


@pytest.mark.parametrize("m,n", [(30, 30), (80, 17), (200, 64)])
def test_dgeqr2_dorg2r_dlarft_dpotrf_match_lapack(ora, m, n):
    A = ora.fill_normal(7, m, n)
    qr, tau = lapack_qr(A)
    Ao = A.copy(order="F")
    to = np.zeros(n)
    ora.call("dgeqr2", m, n, Ao, m, to)
    assert rel(Ao, qr) < 20 * n * U64
    assert rel(to, tau) < 20 * n * U64
    Tl = lapack_dlarft(qr, tau)
    To = np.zeros((n, n), order="F")
    ora.call("dlarft_fc", m, n, qr, m, tau, To, n)
    assert rel(np.triu(To), np.triu(Tl)) < 20 * n * U64
    Qo = qr.copy(order="F")
    ora.call("dorg2r", m, n, n, Qo, m, tau)
    assert rel(Qo, lapack_q(qr, tau, n)) < 20 * n * U64
    G = np.asfortranarray(A.T @ A)
    Gc = G.copy(order="F")
    assert ora.call("dpotrf_u", n, Gc, n) == 0
    c, info = L.dpotrf(G, lower=0)
    assert rel(np.triu(Gc), np.triu(c)) < 20 * n * U64


def test_dpotrf_reports_breakdown(ora):
    G = np.asfortranarray(np.diag([1.0, 2.0, -1.0, 4.0]))
    assert ora.call("dpotrf_u", 4, G, 4) == 3


# ---------------------------------------------------------------------------
# dgeqr3 and the mt-blocked T
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,n", [(1, 1), (9, 9), (100, 33), (500, 64), (300, 128)])
def test_dgeqr3_matches_dgeqrt(ora, m, n):
    A = ora.fill_uniform(3, m, n)
    Ao = A.copy(order="F")
    To = np.zeros((n, n), order="F")
    ora.call("dgeqr3", m, n, Ao, m, To, n)
    a2, t2, info = L.dgeqrt(n, np.asfortranarray(A))
    assert info == 0
    assert rel(np.triu(Ao[:n]), np.triu(a2[:n])) < 20 * n * U64
    assert rel(np.tril(Ao, -1), np.tril(a2, -1)) < 20 * n * U64 + (m == n == 1)
    assert rel(np.triu(To), np.triu(t2)) < 20 * n * U64


@pytest.mark.parametrize("m,n,ii,mt", [(400, 100, 0, 32), (400, 100, 13, 24), (300, 60, 5, 100),
                                       (90, 50, 3, 1)])
def test_dlarft_w03_blocks_are_dlarft_of_their_columns(ora, m, n, ii, mt):
    A = ora.fill_normal(4, m, n + ii)
    qr, tau = lapack_qr(A[ii:, ii:])
    Afull = A.copy(order="F")
    Afull[ii:, ii:] = qr
    T = np.zeros((mt, n + ii), order="F")
    ora.call("lila_dlarft_w03", m, n, ii, mt, Afull, m, T, mt, tau)
    j, vb = ii, min(mt - ii % mt, n)
    while vb > 0:
        Tb = lapack_dlarft(Afull[j:, j:j + vb], tau[j - ii:j - ii + vb])
        got = T[j % mt:j % mt + vb, j:j + vb]
        assert rel(np.triu(got), np.triu(Tb)) < 20 * n * U64
        j += vb
        vb = (ii + n - j) if j + mt >= ii + n else mt
    # T2tau reads the diagonal back; V2tau recomputes tau = 2/(1+||v2||^2)
    t2 = np.zeros(n)
    ora.call("lila_dT2tau_w03", m, n, ii, mt, T, mt, t2)
    assert np.array_equal(t2, tau)
    t3 = np.zeros(n)
    ora.call("lila_dV2tau_w03", m, n, ii, mt, Afull, m, t3)
    assert rel(t3, tau) < 20 * n * U64


# ---------------------------------------------------------------------------
# reflector application and Q build
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,k,nr,mt,ii", [(300, 100, 20, 32, 0), (300, 100, 20, 7, 11), (150, 40, 5, 64, 4)])
def test_dormqrf_z03_matches_dormqr(ora, m, k, nr, mt, ii):
    A = ora.fill_normal(5, m, k + ii)
    qr, tau = lapack_qr(A[ii:, ii:])
    Af = A.copy(order="F")
    Af[ii:, ii:] = qr
    T = np.zeros((mt, k + ii), order="F")
    ora.call("lila_dlarft_w03", m, k, ii, mt, Af, m, T, mt, tau)
    B = ora.fill_normal(6, m, nr + ii)
    Bo = B.copy(order="F")
    ora.call("lila_dormqrf_z03", m, nr, k, ii, ii, mt, Af, m, T, mt, Bo, m, None, 0)
    cq, _, info = L.dormqr("L", "T", np.asfortranarray(qr), tau, np.asfortranarray(B[ii:, ii:]),
                           lwork=max(1, 64 * nr))
    assert info == 0
    assert rel(Bo[ii:, ii:], cq) < 20 * k * U64
    # planted mistakes in the oracle's input must be caught by this check:
    # (a) a dropped term -- T's coupling entries (strict upper) zeroed
    Tw = T.copy(order="F")
    for c in range(ii, ii + k):
        r0 = (c // mt) * mt if c // mt > ii // mt else ii
        Tw[r0 % mt:c % mt, c] = 0.0
    Bw = B.copy(order="F")
    ora.call("lila_dormqrf_z03", m, nr, k, ii, ii, mt, Af, m, Tw, mt, Bw, m, None, 0)
    if min(mt, k) > 1:
        assert rel(Bw[ii:, ii:], cq) > 1e-6
    # (b) one wrong sign -- the first reflector's tau negated
    Tw = T.copy(order="F")
    Tw[ii % mt, ii] = -Tw[ii % mt, ii]
    Bw = B.copy(order="F")
    ora.call("lila_dormqrf_z03", m, nr, k, ii, ii, mt, Af, m, Tw, mt, Bw, m, None, 0)
    assert rel(Bw[ii:, ii:], cq) > 1e-6


CASES = [
    (87, 53, 6, 4, 7, None, None),
    (500, 300, 0, 64, 32, None, None),
    (413, 192, 37, 10, 9, 502, 449),
    (413, 192, 37, 275, 9, 502, 449),
    (213, 200, 13, 50, 16, None, None),
]


def run_ora(ora, kind, case, leaf, panel, nb_lvl=None):
    m, n, ii, mt, nx, lda, ldq = case
    lda, ldq = lda or m, ldq or m
    A = ora.fill_uniform(1, lda, n + ii, -0.5, 0.5)
    Q = ora.fill_uniform(2, ldq, n + ii, -0.5, 0.5)
    T = np.zeros((mt, n + ii), order="F")
    Ao, Qo, To = A.copy(order="F"), Q.copy(order="F"), T.copy(order="F")
    if kind == "recursive":
        p = np.array([0, leaf, panel, nx], dtype=np.int32)
        ora.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ao, lda, To, mt, Qo, ldq, None, 0)
    else:
        p = np.array([0, leaf, panel, len(nb_lvl)], dtype=np.int32)
        nbl = np.array(nb_lvl, dtype=np.int32)
        ora.call("lila_dgeqrf_w03_levelx", p, len(nb_lvl), 0, nbl, m, n, ii, mt, Ao, lda, To, mt, Qo,
                 ldq, None, 0)
    return A, Ao, To, Qo, lda, ldq


def check_against_lapack(ora, case, A, Ao, To, Qo, lda, ldq, tolf=40, lapack_r=True):
    m, n, ii, mt, nx, _, _ = case
    sub = A[ii:m, ii:ii + n]
    qr, tau = lapack_qr(sub)
    tol = tolf * n * U64
    if lapack_r:
        # same dlarfg convention: R and V agree entrywise, no sign freedom
        assert rel(np.triu(Ao[ii:ii + n, ii:ii + n]), np.triu(qr[:n])) < tol
        assert rel(np.tril(Ao[ii:m, ii:ii + n], -1), np.tril(qr, -1)) < tol
        assert rel(Qo[ii:m, ii:ii + n], lapack_q(qr, tau, n)) < tol
        t = np.zeros(n)
        ora.call("lila_dT2tau_w03", m, n, ii, mt, To, mt, t)
        sq = (m - ii == n)  # the last reflector of a square block is the identity (tau = 0)
        assert rel(t[:n - sq], tau[:n - sq]) < tol
    # invariants (reference checkers)
    assert ora.call("test_qq_orth_full", m, n, ii, Qo, ldq) < tol
    assert ora.call("test_qq_orth_1", m, n, ii, Qo, ldq) < tol
    assert ora.call("test_qr_repres_1", m, n, ii, A, lda, Qo, ldq, Ao, lda) < tol
    # lila-c-lila_param/lila/src/lila_test_r_repres_2.c:26-33: a square block's
    # last dlarfg reflector is the identity and is left out, the reconstructed
    # (Cholesky-QR) reflectors are all used
    k = n - 1 if (m - ii == n and lapack_r) else n
    assert ora.call("test_r_repres_2", m, n, k, ii, mt, A, lda, Ao, lda, To, mt) < tol


@pytest.mark.parametrize("leaf,panel", [(1, 1), (1, 0), (0, 1), (1, 2), (2, 0)])
@pytest.mark.parametrize("case", CASES)
def test_recursive_matches_lapack(ora, case, leaf, panel):
    A, Ao, To, Qo, lda, ldq = run_ora(ora, "recursive", case, leaf, panel)
    hr = (leaf == 2 or panel == 2)
    # the Cholesky-QR paths fix R's signs by reconstruction, not by dlarfg
    check_against_lapack(ora, case, A, Ao, To, Qo, lda, ldq, tolf=(400 if hr else 40), lapack_r=not hr)


@pytest.mark.parametrize("nb_lvl", [[32], [64, 16], [100, 30, 7]])
@pytest.mark.parametrize("case", CASES[:3])
def test_levelx_matches_lapack(ora, case, nb_lvl):
    A, Ao, To, Qo, lda, ldq = run_ora(ora, "levelx", case, 1, 1, nb_lvl)
    check_against_lapack(ora, case, A, Ao, To, Qo, lda, ldq)


# ---------------------------------------------------------------------------
# Householder reconstruction
# ---------------------------------------------------------------------------
def orthonormal(ora, m, n, seed):
    A = ora.fill_normal(seed, m, n)
    qr, tau = lapack_qr(A)
    return np.asfortranarray(lapack_q(qr, tau, n)), np.asfortranarray(np.triu(qr[:n]))


@pytest.mark.parametrize("m,n", [(5, 5), (40, 8), (300, 48)])
def test_dorgh2_3_matches_dorhr_col(ora, m, n):
    Q, R = orthonormal(ora, m, n, 9)
    Ain = np.zeros((m, n), order="F")
    Ain[:n] = R
    Ain[n:] = Q[n:]
    T = np.asfortranarray(Q[:n].copy())
    S = np.zeros(n)
    Ao, Qo = Ain.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dorgh2_3", m, n, Ao, m, T, n, Qo, m, None, 0, S)
    Y, Tl, D = lapack_dorhr_col(Q)
    assert np.array_equal(S, D)
    tol = 50 * n * U64
    assert rel(np.tril(Ao, -1), np.tril(Y, -1)) < tol
    assert rel(np.triu(T), np.triu(Tl)) < tol
    # closed forms: Q S = (I - V T V^T)[I;0], (Q S)(S R) = Q R
    V = np.tril(Ao, -1) + np.eye(m, n)
    H = np.eye(m)[:, :n] - V @ np.triu(T) @ V[:n].T
    assert rel(H, Q * S) < tol
    assert rel(Qo, Q * S) == 0.0
    assert rel(np.triu(Ao[:n]), S[:, None] * R) == 0.0


@pytest.mark.parametrize("m,n,i,j,mt", [(200, 24, 0, 0, 32), (260, 16, 5, 44, 20), (300, 40, 11, 64, 64)])
def test_dorgh2_global_matches_dorhr_col(ora, m, n, i, j, mt):
    """lila_dorgh2 (global index, T block at row j % mt, Q columns flipped on
    rows [i, m)) reduces to dorhr_col on the block A(j:m, j:j+n)"""
    Qb, R = orthonormal(ora, m - j, n, 17)
    A = ora.fill_uniform(3, m, j + n, -1.0, 1.0)
    A[j:j + n, j:j + n] = R
    A[j + n:, j:j + n] = Qb[n:]
    T = ora.fill_uniform(4, mt, j + n, -1.0, 1.0)
    T[j % mt:j % mt + n, j:j + n] = Qb[:n]
    Q = ora.fill_uniform(5, m, j + n, -1.0, 1.0)
    Q[j:, j:j + n] = Qb
    S = np.zeros(j + n)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dorgh2", m, n, i, j, mt, Ao, m, To, mt, Qo, m, None, 0, S)
    Y, Tl, D = lapack_dorhr_col(Qb)
    assert np.array_equal(S[j:], D)
    tol = 50 * n * U64
    assert rel(np.tril(Ao[j:, j:j + n], -1), np.tril(Y, -1)) < tol
    assert rel(np.triu(To[j % mt:j % mt + n, j:j + n]), np.triu(Tl)) < tol
    assert np.array_equal(np.triu(Ao[j:j + n, j:j + n]), D[:, None] * R)
    # columns j.. of Q flipped by S on rows [i, m) only; the rest untouched
    assert np.array_equal(Qo[i:, j:], Q[i:, j:] * D)
    assert np.array_equal(Qo[:i], Q[:i]) and np.array_equal(Qo[:, :j], Q[:, :j])
    assert np.array_equal(Ao[:, :j], A[:, :j]) and np.array_equal(Ao[:j], A[:j])


@pytest.mark.parametrize("m,n,i,mt", [(240, 40, 0, 16), (300, 70, 9, 32), (200, 50, 3, 64)])
def test_dorghr_blocked_closed_forms(ora, m, n, i, mt):
    """blocked reconstruction (c3/lila.h:43, int S): with Q orthonormal and R
    on top, the mt-blocked (V, T) reproduce Q S and S R, and every block's
    sign follows the rule of its own diagonal"""
    Qb, R = orthonormal(ora, m - i, n, 19)
    A = np.zeros((m, i + n), order="F")
    A[i:i + n, i:] = R
    Q = np.zeros((m, i + n), order="F")
    Q[i:, i:] = Qb
    T = np.zeros((mt, i + n), order="F")
    S = np.zeros(i + n, dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dorghr", m, n, i, mt, Ao, m, To, mt, Qo, m, None, 0, S)
    Sd = S[i:].astype(np.float64)
    assert set(np.unique(Sd)) <= {-1.0, 1.0}
    tol = 100 * n * U64
    assert np.array_equal(Qo[i:, i:], Qb * Sd)
    # (V, T) applied to the matrix Q R give [S R; 0]
    As = np.zeros((m, i + n), order="F")
    As[i:, i:] = Qb @ R
    assert ora.call("test_r_repres_2", m, n, n, i, mt, As, m, Ao, m, To, mt) < tol
    assert rel(np.triu(Ao[i:i + n, i:]), Sd[:, None] * R) < tol
    # single block: the blocked driver is dorhr_col
    if n <= mt - i % mt:
        Y, Tl, D = lapack_dorhr_col(Qb)
        assert np.array_equal(Sd, D)
        assert rel(np.tril(Ao[i:, i:], -1), np.tril(Y, -1)) < tol


@pytest.mark.parametrize("m,n,ii,mt", [(300, 40, 0, 64), (300, 96, 7, 32), (400, 130, 3, 50)])
def test_mt_hr_closed_forms(ora, m, n, ii, mt):
    A = ora.fill_normal(5, m, n + ii)
    Q = np.zeros((m, n + ii), order="F")
    T = np.zeros((mt, n + ii), order="F")
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    assert ora.call("lila_dgeqrf_w03_mt_hr", m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0) == 0
    tol = 400 * n * U64
    assert ora.call("test_qq_orth_full", m, n, ii, Qo, m) < tol
    assert ora.call("test_qr_repres_1", m, n, ii, A, m, Qo, m, Ao, m) < tol
    # the (V, T) pair reproduces Q: H^T applied to A gives R
    assert ora.call("test_r_repres_2", m, n, n, ii, mt, A, m, Ao, m, To, mt) < tol
    # and R agrees with LAPACK's R up to row signs
    qr, _ = lapack_qr(A[ii:, ii:])
    Rl, Ro = np.triu(qr[:n]), np.triu(Ao[ii:ii + n, ii:ii + n])
    sg = np.sign(np.diag(Ro) / np.diag(Rl))
    assert rel(Ro, sg[:, None] * Rl) < tol


@pytest.mark.parametrize("m,n", [(4, 4), (60, 12), (300, 40)])
def test_dgetrf_b_closed_form(ora, m, n):
    Q, R = orthonormal(ora, m, n, 13)
    A = Q.copy(order="F")
    Qo, U = Q.copy(order="F"), R.copy(order="F")
    ora.call("lila_dgetrf_b", m, n, 0, 4, A, m, Qo, m, U, n * n)
    f = np.where(np.all(Qo == -Q, axis=0), -1.0, 1.0)  # columns of Q flipped
    D = -f
    Lf = np.tril(A, -1) + np.eye(m, n)
    Uf = np.triu(A[:n])
    target = Q * D + np.eye(m, n)
    assert rel(Lf @ Uf, target) < 50 * n * U64
    assert rel(U, f[:, None] * R) == 0.0
    # pivots are 1 + |q_kk| after elimination: never below 1 in magnitude
    assert np.all(np.abs(np.diag(Uf)) >= 1.0 - 1e-12)


# ---------------------------------------------------------------------------
# least squares, checkers, queries, generators
# ---------------------------------------------------------------------------
def test_ls_qr3_matches_dgels(ora):
    m, n, k = 200, 30, 4
    A = ora.fill_normal(21, m, n)
    B = ora.fill_normal(22, m, k)
    Ao, Bo = A.copy(order="F"), B.copy(order="F")
    ora.call("ls_qr3", m, n, k, Ao, m, np.zeros((n, n), order="F"), n, Bo, m, np.zeros((n, k), order="F"))
    _, x, info = L.dgels(np.asfortranarray(A), np.asfortranarray(B))
    assert info == 0
    assert rel(Bo[:n], x[:n]) < 100 * n * U64


def test_checkers_on_known_matrices(ora):
    n, m, eps = 6, 10, 1e-3
    Q = np.asfortranarray(np.eye(m, n) * (1 + eps))
    expect = (2 * eps + eps * eps)
    assert abs(ora.call("test_qq_orth_full", m, n, 0, Q, m) - np.sqrt(n) * expect) < 1e-15
    assert abs(ora.call("test_qq_orth_1", m, n, 0, Q, m) - np.sqrt(n) * expect) < 1e-15
    Q2 = np.asfortranarray(np.eye(m, n))
    Q2[0, 1] = eps  # Q^T Q - I has eps at (0,1) and (1,0): half norm sees one
    half = ora.call("test_qq_orth_1", m, n, 0, Q2, m)
    full = ora.call("test_qq_orth_full", m, n, 0, Q2, m)
    assert abs(full - np.sqrt(2 * eps * eps + eps ** 4)) < 1e-15
    assert half < full


def test_queries_closed_forms(ora):
    import lila_amd
    lib = lila_amd.load()  # queries are host code: no GPU needed
    for (m, n, k, i, mt) in [(100, 20, 30, 0, 8), (2000, 500, 500, 0, 64), (1413, 346, 346, 97, 10),
                             (50, 5, 7, 3, 100)]:
        vb0 = min(mt - i % mt, k)
        expect_f = max(n * mt, (m - i - vb0) * n)  # the first block has the longest V2
        qo = ora.call("lila_query_dormqrf_w03", m, n, k, i, i + k, mt, None, m, None, mt, None, -1)
        ql = lib.lila_query_dormqrf_w03(*[ctypes.c_int(v) for v in (m, n, k, i, i + k, mt)], None,
                                        ctypes.c_int(m), None, ctypes.c_int(mt), None, ctypes.c_int(-1))
        assert qo == ql == expect_f
        qo = ora.call("lila_query_dormqrbz_w03", m, n, k, i, i + k, mt, None, m, None, m, None, mt, None, -1)
        ql = lib.lila_query_dormqrbz_w03(*[ctypes.c_int(v) for v in (m, n, k, i, i + k, mt)], None,
                                         ctypes.c_int(m), None, ctypes.c_int(m), None, ctypes.c_int(mt),
                                         None, ctypes.c_int(-1))
        assert qo == ql
        widths = [vb0]
        while sum(widths) < k:
            widths.append(min(mt, k - sum(widths)))
        if len(widths) == 1:
            assert qo == n * k
        else:
            # the backward walk of lila-c/lila_query_dormqrbz_w03.c evaluates
            # n*vb and (ml-vb)*k for every block but the first one forward;
            # the largest V2 term is then that of the second block
            assert qo == max(n * k, n * max(widths[1:]), (m - i - widths[0] - widths[1]) * k)
    for p in ([0, 1, 1, 32], [0, 0, 0, 16], [0, 2, 0, 9], [0, 1, 2, 7]):
        pa = np.array(p, dtype=np.int32)
        for (m, n, i, mt) in [(2000, 1000, 0, 64), (413, 192, 37, 10), (87, 53, 6, 4)]:
            qo = ora.call("lila_query_dgeqrf_w03_recursive", pa, m, n, i, mt, None, m, None, mt, None, m,
                          None, -1)
            ql = lib.lila_query_dgeqrf_w03_recursive(ctypes.c_void_p(pa.ctypes.data),
                                                     *[ctypes.c_int(v) for v in (m, n, i, mt)], None,
                                                     ctypes.c_int(m), None, ctypes.c_int(mt), None,
                                                     ctypes.c_int(m), None, ctypes.c_int(-1))
            assert qo == ql > 0
            nbl = np.array([50, 10], dtype=np.int32)
            pl = np.array([0, p[1], p[2], 2], dtype=np.int32)
            qo = ora.call("lila_query_dgeqrf_w03_levelx", pl, 2, 0, nbl, m, n, i, mt, None, m, None, mt,
                          None, m, None, -1)
            ql = lib.lila_query_dgeqrf_w03_levelx(ctypes.c_void_p(pl.ctypes.data), ctypes.c_int(2),
                                                  ctypes.c_int(0), ctypes.c_void_p(nbl.ctypes.data),
                                                  *[ctypes.c_int(v) for v in (m, n, i, mt)], None,
                                                  ctypes.c_int(m), None, ctypes.c_int(mt), None,
                                                  ctypes.c_int(m), None, ctypes.c_int(-1))
            assert qo == ql > 0


def test_generators(ora):
    U = ora.fill_uniform(123, 1000, 100)
    assert U.min() >= -1.0 and U.max() < 1.0
    assert abs(U.mean()) < 0.01 and abs(U.var() - 1 / 3) < 0.01
    Z = ora.fill_normal(123, 1001, 99)
    assert abs(Z.mean()) < 0.01 and abs(Z.var() - 1.0) < 0.02
    # distinct seeds give distinct streams, the same seed the same stream
    assert not np.array_equal(ora.fill_uniform(1, 10, 10), ora.fill_uniform(2, 10, 10))
    assert np.array_equal(ora.fill_uniform(1, 10, 10), ora.fill_uniform(1, 10, 10))


def test_oracle_blas_layers_agree(ora):
    """the oracle gives the same result through its built-in BLAS loops and
    through the host OpenBLAS (used for speed by the GPU tests and the CPU
    baseline)"""
    case = (300, 200, 5, 32, 16, None, None)
    out = []
    for host in (False, True):
        if host:
            assert ora.use_host_blas() is not None
        try:
            A, Ao, To, Qo, lda, ldq = run_ora(ora, "recursive", case, 1, 1)
        finally:
            ora.use_builtin_blas()
        out.append((Ao, To, Qo))
    for x, y in zip(out[0], out[1]):
        assert rel(x, y) < 50 * 200 * U64
