"""GPU: the 64-column Cholesky-QR2 panel (panel width 64: dgeqr3 hands
widths 33..64 to one launch when its LDS slabs hold the panel).

Against the oracle's dgeqr3 (V, R and T to rounding, tolerances multiples of
n * u): ordinary panels from 64 rows up to the LDS capacity, ragged widths,
and panels past the capacity (which take the 32-wide recursion).  Panels
that fail an acceptance test take the column-by-column fallback inside the
launch (left 32 columns, their reflectors on the right ones, the right
columns, the T join): an ill-conditioned panel, one ill-conditioned only in
its right half, zero sub-columns (tau = 0, which only the column loop
reproduces exactly), a reflector with tau near 2.  The profile counters
confirm which kernel ran.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def width64(lila, ora):
    ora.use_host_blas()
    old_mode = lila.call("lila_set_panel_mode", 1)
    old = lila.call("lila_set_panel_width", 64)
    assert old >= 0
    yield
    lila.call("lila_set_panel_width", old)
    lila.call("lila_set_panel_mode", old_mode)
    ora.use_builtin_blas()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def launches(lila, cls):
    buf = ctypes.create_string_buffer(1 << 16)
    lila.call("lila_prof_read", buf, len(buf))
    for line in buf.value.decode().strip().splitlines():
        name, n = line.split("|")[:2]
        if name == cls:
            return int(n)
    return 0


def run_dgeqr3(lila, ora, A):
    m, n = A.shape
    To = np.zeros((n, n), order="F")
    Ao = A.copy(order="F")
    ora.call("dgeqr3", m, n, Ao, m, To, n)
    Ad, Td = to_dev(A), to_dev(np.zeros((n, n), order="F"))
    lila.call("lila_prof_enable", 1)
    try:
        assert lila.call("dgeqr3", m, n, Ad, m, Td, n) == 0
        torch.cuda.synchronize()
        n64 = launches(lila, "panel_cqr64_kernel")
    finally:
        lila.call("lila_prof_enable", 0)
    assert lila.call("lila_get_info") == 0
    return (Ao, To), (to_host(Ad), to_host(Td)), n64


@pytest.mark.parametrize("m,n", [(64, 33), (100, 48), (2000, 64), (5000, 50), (16384, 64), (20000, 64),
                                 (30000, 64), (3000, 128)])
def test_panel64_against_dgeqr3(lila, ora, m, n):
    say("64-wide panel", m, n)
    A = ora.fill_normal(71 + m + n, m, n)
    (Ao, To), (Ag, Tg), n64 = run_dgeqr3(lila, ora, A)
    tol = 50 * n * U64
    assert rel(Ag, Ao) < tol
    assert rel(np.triu(Tg), np.triu(To)) < tol
    if m <= 20000:
        # one launch per 33..64-column piece of the recursion
        assert n64 == (1 if n <= 64 else 2)
    else:
        assert n64 == 0  # beyond the 64-wide LDS slabs: the 32-wide recursion


def test_panel64_falls_back_on_ill_conditioned_panel(lila, ora):
    """columns = column 0 + 1e-9 * noise: kappa ~ 1e9, the Gram is useless;
    the launch must take its column-by-column fallback"""
    say("64-wide fallback: ill-conditioned")
    m, n = 5000, 56
    base = ora.fill_normal(3, m, 1)
    noise = ora.fill_normal(4, m, n)
    A = np.asfortranarray(base @ np.ones((1, n)) + 1e-9 * noise)
    (Ao, To), (Ag, Tg), n64 = run_dgeqr3(lila, ora, A)
    assert n64 == 1
    assert rel(np.triu(Ag[:n]), np.triu(Ao[:n])) < 1e-5
    # the factorization itself: Q R = A with Q = [I; 0] - V T V1^T
    V = np.tril(Ag, -1) + np.eye(m, n)
    Qx = np.eye(m, n) - V @ (np.triu(Tg) @ V[:n].T)
    assert np.linalg.norm(Qx @ np.triu(Ag[:n]) - A) / np.linalg.norm(A) < 100 * n * U64


def test_panel64_falls_back_on_right_half(lila, ora):
    """a well-conditioned left half, a rank-deficient right half (its columns
    copies of left ones): Cholesky of the 64 x 64 Gram breaks down, and the
    fallback's right-half column loop meets the near-zero columns"""
    say("64-wide fallback: right half")
    m, n = 3000, 64
    A = ora.fill_normal(9, m, n)
    A[:, 40:48] = A[:, 3:11]
    A = np.asfortranarray(A)
    (Ao, To), (Ag, Tg), n64 = run_dgeqr3(lila, ora, A)
    assert n64 == 1
    V = np.tril(Ag, -1) + np.eye(m, n)
    Qx = np.eye(m, n) - V @ (np.triu(Tg) @ V[:n].T)
    assert np.linalg.norm(Qx @ np.triu(Ag[:n]) - A) / np.linalg.norm(A) < 100 * n * U64
    assert np.linalg.norm(Qx.T @ Qx - np.eye(n)) < 100 * n * U64
    # the left 40 columns are ordinary: V, R and T as the oracle's
    assert rel(Ag[:, :40], Ao[:, :40]) < 1e4 * n * U64


def test_panel64_falls_back_on_zero_subcolumns(lila, ora):
    """an upper-triangular panel: every dlarfg has x = 0, tau = 0; only the
    column loop reproduces that, so the result must be exactly dgeqr3's
    (V = 0, T = 0, R = A)"""
    say("64-wide fallback: tau = 0")
    m, n = 3000, 48
    A = np.zeros((m, n), order="F")
    A[:n] = np.triu(ora.fill_normal(8, n, n))
    (Ao, To), (Ag, Tg), n64 = run_dgeqr3(lila, ora, A)
    assert n64 == 1
    assert np.array_equal(Ag, A)
    assert np.all(np.triu(Tg) == 0.0)
    assert np.array_equal(Ao, A)


@pytest.mark.parametrize("m,n,ii,mt,nx", [(3000, 700, 0, 256, 128), (1213, 333, 13, 100, 64),
                                          (16384, 512, 0, 512, 256)])
def test_recursive_with_panel64(lila, ora, m, n, ii, mt, nx):
    """the w03 recursion with 64-wide panels: V, R, T, Q as the oracle's"""
    from test_gpu_parity import check_vrtq, run_driver
    say("recursive with 64-wide panels", m, n, ii, mt, nx)
    A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, nx])
    check_vrtq(ora, A, o, g, m, n, ii, mt, m, m)


@pytest.mark.parametrize("width", [64, 32])
@pytest.mark.parametrize("kappa", [1e1, 1e3, 1e4])
def test_second_pass_near_identity_or_cholesky(lila, ora, width, kappa):
    """A = U diag(sigma) V^T with sigma graded from 1 to 1/kappa: after the
    first Cholesky-QR pass the Gram of A R1^-1 is I + E with ||E|| ~ kappa^2 u,
    within the near-identity bound (1e-11 entrywise) at kappa = 10, beyond it
    at 1e3 and 1e4, where the second pass takes the Cholesky routine.  Both
    sides give dgeqr3's V, R, T (differences within kappa n u) and an exact
    factorization with orthonormal Q."""
    lila.call("lila_set_panel_width", width)
    m, n = 6000, width
    U, _ = np.linalg.qr(ora.fill_normal(31, m, n))
    W, _ = np.linalg.qr(ora.fill_normal(32, n, n))
    sig = kappa ** (-np.arange(n) / (n - 1))
    A = np.asfortranarray((U * sig) @ W.T)
    say("second pass, width", width, "kappa", kappa)
    (Ao, To), (Ag, Tg), n64 = run_dgeqr3(lila, ora, A)
    assert n64 == (1 if width == 64 else 0)
    tol = 50 * n * U64 * kappa
    assert rel(Ag, Ao) < tol
    assert rel(np.triu(Tg), np.triu(To)) < tol
    V = np.tril(Ag, -1) + np.eye(m, n)
    Qx = np.eye(m, n) - V @ (np.triu(Tg) @ V[:n].T)
    assert np.linalg.norm(Qx @ np.triu(Ag[:n]) - A) / np.linalg.norm(A) < 100 * n * U64
    assert np.linalg.norm(Qx.T @ Qx - np.eye(n)) < 100 * n * U64


@pytest.mark.parametrize("width,n", [(64, 64), (64, 48), (64, 33), (32, 32), (32, 20), (32, 1)])
def test_square_panel_last_reflector(lila, ora, width, n):
    """a square panel: its last column has no rows below the diagonal, so
    dgeqr3's last reflector is H = I (tau = 0: T's last diagonal entry is
    exactly 0) and R's last diagonal entry keeps the sign of the value the
    earlier reflectors leave there.  The reconstruction takes that column's
    sign from the LU's last Schur complement (+-1) instead of the column
    loop; V, R, T as the oracle's, within kappa n u.  (The 64-wide launch
    takes 64 rows and more, so a square panel there is 64 x 64; narrower
    square matrices end in a square 32-wide panel of the recursion.)"""
    lila.call("lila_set_panel_width", width)
    A = ora.fill_normal(5 + n + width, n, n)
    kappa = np.linalg.cond(A)
    say("square panel, width", width, "n", n, "kappa %.3g" % kappa)
    (Ao, To), (Ag, Tg), n64 = run_dgeqr3(lila, ora, A)
    assert n64 == (1 if width == 64 and n == 64 else 0)
    # T's last column exactly +0 (dlarfg's tau = +0; the UT forms invert it to +inf)
    assert np.all(To[:, n - 1] == 0.0) and np.all(Tg[:, n - 1] == 0.0)
    assert not np.signbit(Tg[n - 1, n - 1]) and not np.signbit(To[n - 1, n - 1])
    tol = 50 * n * U64 * kappa
    assert rel(Ag, Ao) < tol
    assert rel(np.triu(Tg), np.triu(To)) < tol
    assert np.sign(Ag[n - 1, n - 1]) == np.sign(Ao[n - 1, n - 1])
    V = np.tril(Ag, -1) + np.eye(n)
    Qx = np.eye(n) - V @ (np.triu(Tg) @ V.T)
    assert np.linalg.norm(Qx @ np.triu(Ag) - A) / np.linalg.norm(A) < 100 * n * U64
    assert np.linalg.norm(Qx.T @ Qx - np.eye(n)) < 100 * n * U64
