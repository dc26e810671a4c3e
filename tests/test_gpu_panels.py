"""GPU: both panel algorithms of the drivers against the oracle.

Mode 0 runs every panel column by column (dgeqr3's arithmetic); mode 1 (the
default) runs Cholesky-QR2 + Householder reconstruction in one launch and
falls back inside the kernel to the column-by-column loop when the Gram is
ill-conditioned or a reflector is degenerate.  Panels taller than the
workgroups' LDS slabs take TSQR (a tree of slab Householder QRs, the explicit
Q built back down the tree, then the sign-chosen LU reconstruction) in mode
0, and in mode 1 whenever the streamed Cholesky-QR2 kernel declines them;
degenerate reflectors (tau = 0) still take the column loop.  All must
reproduce dgeqr3's V, R, T and Q to rounding.  Tolerances are multiples of
n * u.
"""
import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def oracle_on_host_blas(ora):
    ora.use_host_blas()
    yield
    ora.use_builtin_blas()


@pytest.fixture
def panel_mode(lila):
    def set_mode(m):
        return lila.call("lila_set_panel_mode", m)
    yield set_mode
    lila.call("lila_set_panel_mode", 1)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def run_dgeqr3(lila, ora, A):
    m, n = A.shape
    To = np.zeros((n, n), order="F")
    Ao = A.copy(order="F")
    ora.call("dgeqr3", m, n, Ao, m, To, n)
    Ad, Td = to_dev(A), to_dev(np.zeros((n, n), order="F"))
    assert lila.call("dgeqr3", m, n, Ad, m, Td, n) == 0
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    return (Ao, To), (to_host(Ad), to_host(Td))


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("m,n", [(100, 7), (2000, 32), (20000, 32), (126000, 32), (250000, 16),
                                 (400000, 32), (1000000, 12)])
def test_dgeqr3_both_modes(lila, ora, panel_mode, mode, m, n):
    say("dgeqr3 mode", mode, m, n)
    panel_mode(mode)
    A = ora.fill_normal(31 + m, m, n)
    (Ao, To), (Ag, Tg) = run_dgeqr3(lila, ora, A)
    tol = 50 * n * U64
    assert rel(Ag, Ao) < tol
    assert rel(np.triu(Tg), np.triu(To)) < tol


@pytest.mark.parametrize("mode", [0, 1])
def test_recursive_both_modes(lila, ora, panel_mode, mode):
    m, n, ii, mt, nx = 3000, 1000, 11, 128, 32
    say("recursive mode", mode)
    panel_mode(mode)
    A = ora.fill_uniform(5, m, n + ii, -0.5, 0.5)
    Q = np.zeros((m, n + ii), order="F")
    T = np.zeros((mt, n + ii), order="F")
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol
    assert rel(to_host(Qd), Qo) < tol
    assert rel(to_host(Td), To) < 10 * tol


def test_tall_recursive_beyond_lds_capacity(lila, ora, panel_mode):
    """m = 300000 puts every 32-column panel past 256 LDS slabs: the panels run
    the streamed Cholesky-QR2 kernel (sub-slabs through LDS once per pass)"""
    m, n, ii, mt, nx = 300000, 96, 0, 64, 32
    say("tall recursive", m, n)
    panel_mode(1)
    A = ora.fill_normal(77, m, n)
    Q = np.zeros((m, n), order="F")
    T = np.zeros((mt, n), order="F")
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol
    assert rel(to_host(Qd), Qo) < tol
    assert rel(to_host(Td), To) < 10 * tol


def test_cqr_falls_back_on_ill_conditioned_panel(lila, ora, panel_mode):
    """columns 1..n-1 = column 0 + 1e-9 * noise: kappa ~ 1e9, the Gram is
    useless and the kernel must switch to the column-by-column loop"""
    say("cqr fallback: ill-conditioned")
    panel_mode(1)
    m, n = 5000, 24
    base = ora.fill_normal(3, m, 1)
    noise = ora.fill_normal(4, m, n)
    A = np.asfortranarray(base @ np.ones((1, n)) + 1e-9 * noise)
    (Ao, To), (Ag, Tg) = run_dgeqr3(lila, ora, A)
    # R and V of the rank-deficient-ish panel agree to kappa * u
    assert rel(np.triu(Ag[:n]), np.triu(Ao[:n])) < 1e-5


def test_cqr_falls_back_on_zero_subcolumns(lila, ora, panel_mode):
    """an upper-triangular panel: every dlarfg has x = 0, tau = 0, beta =
    alpha; the reconstruction cannot produce tau = 0, so the kernel must fall
    back and return dgeqr3's exact answer (V = 0, T = 0, R = A)"""
    say("cqr fallback: tau = 0")
    panel_mode(1)
    m, n = 3000, 20
    A = np.zeros((m, n), order="F")
    A[:n] = np.triu(ora.fill_normal(8, n, n))
    (Ao, To), (Ag, Tg) = run_dgeqr3(lila, ora, A)
    assert np.array_equal(Ag, A)
    assert np.all(np.triu(Tg) == 0.0)
    assert np.array_equal(Ao, A)


@pytest.mark.parametrize("m", [400000])
def test_streamed_cqr_falls_back_on_ill_conditioned_tall_panel(lila, ora, panel_mode, m):
    """the ill-conditioned panel above, past the LDS capacity: the streamed
    kernel has only read the panel when its Cholesky test fails, declines it,
    and the panel is factored by TSQR"""
    say("streamed cqr fallback: ill-conditioned", m)
    panel_mode(1)
    n = 24
    base = ora.fill_normal(5, m, 1)
    noise = ora.fill_normal(6, m, n)
    A = np.asfortranarray(base @ np.ones((1, n)) + 1e-9 * noise)
    (Ao, To), (Ag, Tg) = run_dgeqr3(lila, ora, A)
    assert rel(np.triu(Ag[:n]), np.triu(Ao[:n])) < 1e-5


@pytest.mark.parametrize("m", [400000])
def test_streamed_cqr_falls_back_on_zero_subcolumns(lila, ora, panel_mode, m):
    """upper-triangular tall panel: tau = 0 everywhere, the streamed kernel's
    reconstruction test fails, TSQR's Q shows |Q_kk| = 1, and the column
    loop returns the panel unchanged"""
    say("streamed cqr fallback: tau = 0", m)
    panel_mode(1)
    n = 20
    A = np.zeros((m, n), order="F")
    A[:n] = np.triu(ora.fill_normal(9, n, n))
    (Ao, To), (Ag, Tg) = run_dgeqr3(lila, ora, A)
    assert np.array_equal(Ag, A)
    assert np.all(np.triu(Tg) == 0.0)


def test_panel_modes_agree_on_large_recursive_config1(lila, ora, panel_mode):
    """config 1 shape, both modes against each other and the invariants"""
    m, n, mt, nx = 2000, 1000, 64, 32
    say("config1 both modes")
    A = ora.fill_uniform(2000, m, n, -1.0, 1.0)
    outs = []
    for mode in (0, 1):
        panel_mode(mode)
        Ad, Td, Qd = to_dev(A), to_dev(np.zeros((mt, n), order="F")), to_dev(np.zeros((m, n), order="F"))
        p = np.array([0, 1, 1, nx], dtype=np.int32)
        assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
        torch.cuda.synchronize()
        outs.append((to_host(Ad), to_host(Td), to_host(Qd)))
    for x, y in zip(outs[0], outs[1]):
        assert rel(x, y) < 200 * n * U64
    Ag, Tg, Qg = outs[1]
    assert ora.call("test_qq_orth_full", m, n, 0, Qg, m) < 50 * n * U64
    assert ora.call("test_qr_repres_1", m, n, 0, A, m, Qg, m, Ag, m) < 50 * n * U64


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("m,n,spread", [(1000000, 32, 1e12), (600000, 17, 1e14), (400000, 32, 1.0)])
def test_tsqr_tall_graded_panels(lila, ora, panel_mode, mode, m, n, spread):
    """tall panels whose columns are graded over `spread` (kappa ~ spread):
    mode 0 runs TSQR directly, mode 1 has the streamed Cholesky-QR2 kernel
    decline (its Gram cannot resolve the spread) and then TSQR; V, R, T
    against the oracle's dgeqr3 (Householder, column by column), and through
    lila_dgeqr2_w03_3 the explicit panel Q against the oracle's"""
    say("tsqr tall graded", mode, m, n, spread)
    panel_mode(mode)
    A = np.asfortranarray(ora.fill_normal(91 + n, m, n) * np.logspace(0, -np.log10(spread), n)[None, :])
    (Ao, To), (Ag, Tg) = run_dgeqr3(lila, ora, A)
    tol = 100 * n * U64
    assert rel(np.tril(Ag, -1), np.tril(Ao, -1)) < tol, "V"
    # R row by row relative to its diagonal (graded rows)
    Rg, Ro = np.triu(Ag[:n]), np.triu(Ao[:n])
    assert np.max(np.abs(Rg - Ro) / np.abs(np.diag(Ro))[:, None]) < tol, "R"
    assert rel(np.triu(Tg), np.triu(To)) < tol, "T"
    # explicit panel Q through the dgeqr2 entry (i = 0, one T block)
    T2 = np.zeros((n, n), order="F")
    Q2 = np.zeros((m, n), order="F")
    Ao2, To2, Qo2 = A.copy(order="F"), T2.copy(order="F"), Q2.copy(order="F")
    ora.call("lila_dgeqr2_w03_3", m, n, 0, n, Ao2, m, To2, n, Qo2, m, None, 0)
    Ad, Td, Qd = to_dev(A), to_dev(T2), to_dev(Q2)
    assert lila.call("lila_dgeqr2_w03_3", m, n, 0, n, Ad, m, Td, n, Qd, m, None, 0) == 0
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    assert rel(to_host(Qd), Qo2) < tol, "Q"
