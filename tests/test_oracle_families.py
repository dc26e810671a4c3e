"""Pins the oracle's VRT-only (v03), Q-from-(V,T) (q03), append-columns,
LAPACK-kernel (ker_recursive, dormqrf_w00/z00/z02) and qr2 families against
the host LAPACK and the w03 path already pinned in test_oracle.py."""
import numpy as np
import pytest

from conftest import U64
from test_oracle import CASES, check_against_lapack, lapack_dlarft, lapack_q, lapack_qr, rel

import scipy.linalg.lapack as L


def vq_run(ora, case, leaf, panel, levelx=None):
    """v03 then q03 on the same (A, T): must be the VRTQ of w03"""
    m, n, ii, mt, nx, lda, ldq = case
    lda, ldq = lda or m, ldq or m
    A = ora.fill_uniform(1, lda, n + ii, -0.5, 0.5)
    Q = ora.fill_uniform(2, ldq, n + ii, -0.5, 0.5)
    T = np.zeros((mt, n + ii), order="F")
    Ao, Qo, To = A.copy(order="F"), Q.copy(order="F"), T.copy(order="F")
    if levelx is None:
        p = np.array([0, leaf, panel, nx], dtype=np.int32)
        ora.call("lila_dgeqrf_v03_recursive", p, m, n, ii, mt, Ao, lda, To, mt, None, 0)
        Av = Ao.copy(order="F")
        ora.call("lila_dgeqrf_q03_recursive", p, m, n, ii, mt, Ao, lda, To, mt, Qo, ldq, None, 0)
    else:
        p = np.array([0, leaf, panel, len(levelx)], dtype=np.int32)
        nbl = np.array(levelx, dtype=np.int32)
        ora.call("lila_dgeqrf_v03_levelx", p, len(levelx), 0, nbl, m, n, ii, mt, Ao, lda, To, mt,
                 None, 0)
        Av = Ao.copy(order="F")
        ora.call("lila_dgeqrf_q03_levelx", p, len(levelx), 0, nbl, m, n, ii, mt, Ao, lda, To, mt,
                 Qo, ldq, None, 0)
    assert np.array_equal(Av, Ao)  # the Q build reads V and T only
    return A, Ao, To, Qo, lda, ldq


@pytest.mark.parametrize("leaf,panel", [(1, 1), (1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize("case", CASES)
def test_v03_then_q03_is_vrtq(ora, case, leaf, panel):
    A, Ao, To, Qo, lda, ldq = vq_run(ora, case, leaf, panel)
    check_against_lapack(ora, case, A, Ao, To, Qo, lda, ldq)


@pytest.mark.parametrize("nb_lvl", [[32], [64, 16]])
@pytest.mark.parametrize("case", CASES[:3])
def test_v03_q03_levelx(ora, case, nb_lvl):
    A, Ao, To, Qo, lda, ldq = vq_run(ora, case, 1, 1, nb_lvl)
    check_against_lapack(ora, case, A, Ao, To, Qo, lda, ldq)


def test_v03_leaves_q_alone_and_matches_w03(ora):
    m, n, ii, mt, nx = 300, 120, 5, 40, 16
    A = ora.fill_uniform(3, m, n + ii, -1.0, 1.0)
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    Aw, Tw, Qw = A.copy(order="F"), np.zeros((mt, n + ii), order="F"), np.zeros((m, n + ii), order="F")
    # w03 splits while n > nx, v03 while n >= nx: nx + 1 gives w03 the same tree
    pw = np.array([0, 1, 1, nx - 1], dtype=np.int32)
    ora.call("lila_dgeqrf_w03_recursive", pw, m, n, ii, mt, Aw, m, Tw, mt, Qw, m, None, 0)
    Av, Tv = A.copy(order="F"), np.zeros((mt, n + ii), order="F")
    ora.call("lila_dgeqrf_v03_recursive", p, m, n, ii, mt, Av, m, Tv, mt, None, 0)
    assert rel(Av, Aw) < 10 * n * U64
    assert rel(Tv, Tw) < 10 * n * U64


@pytest.mark.parametrize("m,n,nb,mt,nx", [(87, 53, 10, 4, 1), (300, 150, 40, 32, 16), (250, 97, 31, 64, 8)])
def test_w03_appendcols_builds_the_full_qr(ora, m, n, nb, mt, nx):
    """the Block-Arnoldi driver (lila-c/main_w03_appendcols.c): columns
    appended in chunks of nb give the QR of the whole matrix"""
    A = ora.fill_uniform(4, m, n, -0.5, 0.5)
    Ao, Qo, To = A.copy(order="F"), np.zeros((m, n), order="F"), np.zeros((mt, n), order="F")
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    j = 0
    while j < n:
        vb = min(nb, n - j)
        ora.call("lila_dgeqrf_w03_appendcols", p, m, j, vb, mt, Ao, m, Qo, m, To, mt, None, 0)
        j += vb
    check_against_lapack(ora, (m, n, 0, mt, nx, None, None), A, Ao, To, Qo, m, m)


@pytest.mark.parametrize("m,n,nb", [(87, 53, 10), (300, 150, 64)])
def test_lapack_appendcols_is_dgeqrf(ora, m, n, nb):
    A = ora.fill_uniform(5, m, n, -0.5, 0.5)
    Ao, Qo = A.copy(order="F"), np.zeros((m, n), order="F")
    tau = np.zeros(n)
    j = 0
    while j < n:
        vb = min(nb, n - j)
        ora.call("lila_dgeqrf_LAPACK_appendcols", m, j, vb, 8, Ao, m, Qo, m, tau, 1, None, 0)
        j += vb
    qr, tl = lapack_qr(A)
    tol = 40 * n * U64
    assert rel(Ao, qr) < tol and rel(tau, tl) < tol
    assert rel(Qo, lapack_q(qr, tl, n)) < tol


@pytest.mark.parametrize("m,n,nx", [(87, 53, 7), (300, 200, 16), (64, 64, 8)])
def test_ker_recursive_is_dgeqrf_dorgqr(ora, m, n, nx):
    A = ora.fill_uniform(6, m, n, -0.5, 0.5)
    Ao, Qo = A.copy(order="F"), np.zeros((m, n), order="F")
    tau = np.zeros(n)
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    ora.call("lila_dgeqrf_ker_recursive", p, m, n, 0, 1, Ao, m, tau, 1, Qo, m, None, 0)
    qr, tl = lapack_qr(A)
    tol = 40 * n * U64
    sq = (m == n)
    assert rel(Ao, qr) < tol and rel(tau[:n - sq], tl[:n - sq]) < tol
    assert rel(Qo, lapack_q(qr, tl, n)) < tol


@pytest.mark.parametrize("m,k,nr,ii", [(200, 40, 9, 0), (150, 30, 5, 7)])
def test_dormqrf_w00_z00_z02_match_dormqr(ora, m, k, nr, ii):
    A = ora.fill_normal(7, m, k + ii + nr)
    qr, tau = lapack_qr(A[ii:, ii:ii + k])
    Af = A.copy(order="F")
    Af[ii:, ii:ii + k] = qr
    B = ora.fill_normal(8, m, nr + ii)
    cq, _, info = L.dormqr("L", "T", np.asfortranarray(qr), tau, np.asfortranarray(B[ii:, ii:]),
                           lwork=max(1, 64 * nr))
    assert info == 0
    tol = 40 * k * U64
    # z00: tau rebuilt from the stored vectors, C = B
    Bz = B.copy(order="F")
    ora.call("lila_dormqrf_z00", m, nr, k, ii, ii, 8, Af, m, Bz, m, None, 1, None, 0)
    assert rel(Bz[ii:, ii:], cq) < tol
    # w00: C = A's columns right of the reflectors
    Aw = Af.copy(order="F")
    Aw[:, ii + k:ii + k + nr] = B[:, ii:ii + nr]
    ora.call("lila_dormqrf_w00", m, nr, k, ii, ii + k, 8, Aw, m, None, 1, None, 0)
    assert rel(Aw[ii:, ii + k:ii + k + nr], cq) < tol
    # z02: one block with the full-layout T at T(ii, ii)
    T = np.zeros((k + ii, k + ii), order="F")
    T[ii:, ii:] = lapack_dlarft(qr, tau)
    B2 = B.copy(order="F")
    ora.call("lila_dormqrf_z02", m, nr, k, ii, ii, 8, Af, m, T, k + ii, B2, m, None, 0)
    assert rel(B2[ii:, ii:], cq) < tol


def test_qr2_methods(ora):
    """qr2: method 0 is w03's VRTQ, method 1 leaves Q alone, method 2 builds
    Q from method 1's (V, T)"""
    m, n, ii, mt, nx = 260, 96, 0, 96, 16
    A = ora.fill_uniform(9, m, n, -0.5, 0.5)
    outs = {}
    for method in (0, 1):
        p = np.array([0, 1, 1, nx, method], dtype=np.int32)
        Ao, To, Qo = A.copy(order="F"), np.zeros((mt, n), order="F"), np.zeros((m, n), order="F")
        ora.call("lila_dgeqrf_qr2_recursive", p, m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
        outs[method] = (Ao, To, Qo)
    assert not np.any(outs[1][2])
    assert rel(outs[1][0], outs[0][0]) < 10 * n * U64 and rel(outs[1][1], outs[0][1]) < 10 * n * U64
    p2 = np.array([0, 1, 1, nx, 2], dtype=np.int32)
    Ao, To = outs[1][0].copy(order="F"), outs[1][1].copy(order="F")
    Q2 = np.zeros((m, n), order="F")
    ora.call("lila_dgeqrf_qr2_recursive", p2, m, n, ii, mt, Ao, m, To, mt, Q2, m, None, 0)
    assert rel(Q2, outs[0][2]) < 10 * n * U64
    check_against_lapack(ora, (m, n, ii, mt, nx, None, None), A, *outs[0], m, m)
