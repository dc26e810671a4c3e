"""GPU parity: every hot-path entry point of liblila_hip.so, called through the
C ABI on device memory, against the CPU oracle on the same seeded inputs.

Floating-point bar: the GPU and the oracle run the same algorithm with
different summation orders, so results agree to rounding.  The tolerance is
written per test as a multiple of n * u (u = 2^-53), relative to the size of
the compared quantity; the invariants ||A - QR||/||A|| and ||I - Q^T Q|| are
checked against c * n * u as well.
"""
import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def oracle_on_host_blas(ora):
    """the oracle's BLAS-3 steps on the host OpenBLAS for speed (its pinning
    in test_oracle.py covers both BLAS layers)"""
    ora.use_host_blas()
    yield
    ora.use_builtin_blas()


def rel(a, b):
    d = np.linalg.norm(a - b)
    s = max(np.linalg.norm(b), 1e-300)
    return d / s


def sync_ok(lila):
    torch.cuda.synchronize()
    info = lila.call("lila_get_info")
    assert info == 0, "device status %d" % info


# ---------------------------------------------------------------------------
# GEMM (the MFMA kernel under every BLAS-3 step)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (128, 128, 16), (130, 70, 300),
                                   (512, 512, 512), (500, 300, 257), (33, 480, 511), (512, 513, 64),
                                   (64, 96, 20000), (1000, 513, 64), (32, 32, 100000)])
def test_dgemm(lila, ora, ta, tb, m, n, k):
    # M, N <= 512 with K <= 512 take the one-shot small-product kernel (all of
    # K staged per round of 256); the rest the tiled and direct-to-LDS kernels
    say("dgemm", ta, tb, m, n, k)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = np.asfortranarray(rng.standard_normal((k, m) if ta == "T" else (m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k) if tb == "T" else (k, n)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    ref = 1.5 * (A.T if ta == "T" else A) @ (B.T if tb == "T" else B) - 0.5 * C
    Ad, Bd, Cd = to_dev(A), to_dev(B), to_dev(C)
    import ctypes
    r = lila.call("lila_dgemm", ctypes.c_char(ta.encode()), ctypes.c_char(tb.encode()), m, n, k,
                  1.5, Ad, A.shape[0], Bd, B.shape[0], -0.5, Cd, m)
    assert r == 0
    sync_ok(lila)
    got = to_host(Cd)
    # |error| <= k u |A||B| elementwise bound, checked in norm
    bound = 8 * (k + 2) * U64 * (1.5 * np.abs(A.T if ta == "T" else A) @ np.abs(B.T if tb == "T" else B)
                                 + 0.5 * np.abs(C))
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


@pytest.mark.parametrize("backend", [0, 1])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("m,n,k", [(2048, 2048, 512), (2100, 2200, 530), (1024, 8192, 3000),
                                   (4096, 1152, 16)])
def test_dgemm_direct_lds_tiles(lila, ora, ta, m, n, k, backend):
    """shapes large enough for the 128 x 128 direct-to-LDS kernel: its
    interior (backend 0; with backend 1 the plain NN interior runs on
    rocBLAS), the ragged strips and the K tail (run by the tiled kernel) must
    add up to the product"""
    say("dgemm 128x128 interior", ta, m, n, k, "backend", backend)
    old = lila.call("lila_set_gemm_backend", backend)
    try:
        run_direct_lds(lila, ta, m, n, k)
    finally:
        lila.call("lila_set_gemm_backend", old)


def run_direct_lds(lila, ta, m, n, k):
    rng = np.random.default_rng(m + 3 * n + 7 * k)
    A = np.asfortranarray(rng.standard_normal((k, m) if ta == "T" else (m, k)))
    B = np.asfortranarray(rng.standard_normal((k, n)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    opA = A.T if ta == "T" else A
    ref = -0.75 * opA @ B + 1.25 * C
    Cd = to_dev(C)
    import ctypes
    r = lila.call("lila_dgemm", ctypes.c_char(ta.encode()), ctypes.c_char(b"N"), m, n, k, -0.75,
                  to_dev(A), A.shape[0], to_dev(B), k, 1.25, Cd, m)
    assert r == 0
    sync_ok(lila)
    got = to_host(Cd)
    bound = 8 * (k + 2) * U64 * (0.75 * np.abs(opA) @ np.abs(B) + 1.25 * np.abs(C))
    assert np.all(np.abs(got - ref) <= bound + 1e-300)


# ---------------------------------------------------------------------------
# panel: dgeqr3 (fused kernel up to 32 columns, recursion above)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,n", [(1, 1), (5, 5), (100, 7), (500, 32), (2000, 32), (20000, 32),
                                 (1000, 64), (300, 100), (120000, 16)])
def test_dgeqr3(lila, ora, m, n):
    say("dgeqr3", m, n)
    A = ora.fill_uniform(11 + m + n, m, n)
    T0 = np.zeros((n, n), order="F")
    Ao, To = A.copy(order="F"), T0.copy(order="F")
    ora.call("dgeqr3", m, n, Ao, m, To, n)
    Ad, Td = to_dev(A), to_dev(T0)
    assert lila.call("dgeqr3", m, n, Ad, m, Td, n) == 0
    sync_ok(lila)
    Ag, Tg = to_host(Ad), to_host(Td)
    tol = 50 * max(n, 8) * U64
    assert rel(Ag, Ao) < tol
    assert rel(np.triu(Tg), np.triu(To)) < tol


# ---------------------------------------------------------------------------
# VRTQ drivers
# ---------------------------------------------------------------------------
def run_driver(lila, ora, kind, m, n, ii, mt, param, nb_lvl=None, lda=None, ldq=None, seed=1,
               normal=False):
    lda = lda or m
    ldq = ldq or m
    ldt = mt
    fill = ora.fill_normal if normal else (lambda s, r, c: ora.fill_uniform(s, r, c, -0.5, 0.5))
    A = fill(seed, lda, n + ii)
    Q = fill(seed + 1, ldq, n + ii)
    T = np.zeros((ldt, n + ii), order="F")
    p = np.array(param, dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    if kind == "recursive":
        args = (m, n, ii, mt)
        ora.call("lila_dgeqrf_w03_recursive", p, *args, Ao, lda, To, ldt, Qo, ldq, None, 0)
        r = lila.call("lila_dgeqrf_w03_recursive", p, *args, Ad, lda, Td, ldt, Qd, ldq, None, 0)
    elif kind == "levelx":
        nbl = np.array(nb_lvl, dtype=np.int32)
        args = (len(nb_lvl), 0, nbl, m, n, ii, mt)
        ora.call("lila_dgeqrf_w03_levelx", p, *args, Ao, lda, To, ldt, Qo, ldq, None, 0)
        r = lila.call("lila_dgeqrf_w03_levelx", p, *args, Ad, lda, Td, ldt, Qd, ldq, None, 0)
    else:
        raise ValueError(kind)
    assert r == 0
    sync_ok(lila)
    Ag, Tg, Qg = to_host(Ad), to_host(Td), to_host(Qd)
    return A, (Ao, To, Qo), (Ag, Tg, Qg)


def check_vrtq(ora, A, o, g, m, n, ii, mt, lda, ldq, tolf=100):
    Ao, To, Qo = o
    Ag, Tg, Qg = g
    tol = tolf * n * U64
    assert rel(Ag, Ao) < tol, "A/V/R"
    assert rel(Qg, Qo) < tol, "Q"
    assert rel(Tg, To) < tol * 10, "T"
    # invariants on the GPU result, checked by the oracle's checkers
    orth = ora.call("test_qq_orth_full", m, n, ii, Qg, ldq)
    repres = ora.call("test_qr_repres_1", m, n, ii, A, lda, Qg, ldq, Ag, lda)
    assert orth < tol, orth
    assert repres < tol, repres


CASES = [
    # m, n, ii, mt, nx, lda, ldq   (reference default, config 1, ugly sweeps)
    (87, 53, 6, 4, 7, None, None),
    (2000, 1000, 0, 64, 32, None, None),
    (1413, 692, 97, 10, 9, 1902, 1809),
    (1413, 692, 97, 775, 9, 1902, 1809),
    (1213, 1200, 13, 200, 32, None, None),
    (600, 300, 0, 300, 64, None, None),
]


@pytest.mark.parametrize("leaf,panel", [(1, 1), (1, 0), (0, 1), (1, 2), (2, 1)])
@pytest.mark.parametrize("case", CASES)
def test_recursive(lila, ora, case, leaf, panel):
    m, n, ii, mt, nx, lda, ldq = case
    say("recursive", case, "leaf", leaf, "panel", panel)
    if leaf == 2 or panel == 2:
        if m > 1500:
            pytest.skip("CholQR leaf on the large ugly case is covered by the hr tests")
    A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, leaf, panel, nx], lda=lda, ldq=ldq)
    check_vrtq(ora, A, o, g, m, n, ii, mt, lda or m, ldq or m,
               tolf=(2000 if (leaf == 2 or panel == 2) else 100))


@pytest.mark.parametrize("leaf,panel", [(1, 1), (0, 1), (1, 0)])
@pytest.mark.parametrize("case", [CASES[1], CASES[2], CASES[3], CASES[4]])
def test_recursive_q_after(lila, ora, case, leaf, panel):
    """Q built after the VRT recursion by whole T blocks (lila_set_q_after):
    the same V, R, T and Q as the reference's call tree, to rounding"""
    m, n, ii, mt, nx, lda, ldq = case
    say("recursive q_after", case, leaf, panel)
    old = lila.call("lila_set_q_after", 1)
    try:
        A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, leaf, panel, nx], lda=lda,
                             ldq=ldq)
    finally:
        lila.call("lila_set_q_after", old)
    check_vrtq(ora, A, o, g, m, n, ii, mt, lda or m, ldq or m)


@pytest.mark.parametrize("backend", [0, 1])
def test_recursive_gemm_backends(lila, ora, backend):
    """the plain NN interiors of the trailing updates on the library's
    direct-to-LDS kernel (0) or on rocBLAS (1): the same V, R, T and Q"""
    m, n, ii, mt, nx = 3000, 2400, 0, 512, 256
    say("recursive, gemm backend", backend)
    old = lila.call("lila_set_gemm_backend", backend)
    try:
        A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, nx])
    finally:
        lila.call("lila_set_gemm_backend", old)
    check_vrtq(ora, A, o, g, m, n, ii, mt, m, m)


@pytest.mark.parametrize("nb_lvl", [[32], [128], [64, 16]])
@pytest.mark.parametrize("case", [(500, 300, 0, 64), (777, 500, 21, 50)])
def test_levelx(lila, ora, case, nb_lvl):
    m, n, ii, mt = case
    say("levelx", case, nb_lvl)
    A, o, g = run_driver(lila, ora, "levelx", m, n, ii, mt, [0, 1, 1, len(nb_lvl)], nb_lvl=nb_lvl,
                         normal=True)
    check_vrtq(ora, A, o, g, m, n, ii, mt, m, m)


def test_recursive_host_pointers(lila, ora):
    """true drop-in: host arrays in, host arrays out"""
    m, n, ii, mt = 300, 200, 5, 32
    say("recursive host pointers")
    A = ora.fill_uniform(3, m, n + ii, -0.5, 0.5)
    Q = np.zeros((m, n + ii), order="F")
    T = np.zeros((mt, n + ii), order="F")
    p = np.array([0, 1, 1, 16], dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    Ah, Th, Qh = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ah, m, Th, mt, Qh, m, None, 0) == 0
    tol = 100 * n * U64
    assert rel(Ah, Ao) < tol and rel(Qh, Qo) < tol and rel(Th, To) < 10 * tol


# ---------------------------------------------------------------------------
# apply Q^T to a right-hand block (config 1: 2000 x 64)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,n,ii,mt,nrhs", [(2000, 1000, 0, 64, 64), (700, 300, 17, 40, 33)])
def test_dormqrf_z03(lila, ora, m, n, ii, mt, nrhs):
    say("dormqrf_z03", m, n, ii, mt, nrhs)
    A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, 32])
    Ao, To, _ = o
    B = ora.fill_uniform(99, m, nrhs + ii)
    Bo = B.copy(order="F")
    ora.call("lila_dormqrf_z03", m, nrhs, n, ii, ii, mt, Ao, m, To, mt, Bo, m, None, 0)
    Ad, Td, Bd = to_dev(Ao), to_dev(To), to_dev(B)
    assert lila.call("lila_dormqrf_z03", m, nrhs, n, ii, ii, mt, Ad, m, Td, mt, Bd, m, None, 0) == 0
    sync_ok(lila)
    assert rel(to_host(Bd), Bo) < 100 * n * U64


# ---------------------------------------------------------------------------
# T factor pieces
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,n,ii,mt", [(400, 100, 0, 32), (400, 100, 13, 24), (300, 60, 5, 100)])
def test_dlarft_w03_and_tau(lila, ora, m, n, ii, mt):
    say("dlarft_w03", m, n, ii, mt)
    A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, 16])
    Ao, To, _ = o
    tau = np.zeros(n)
    ora.call("lila_dT2tau_w03", m, n, ii, mt, To, mt, tau)
    T1 = np.zeros((mt, n + ii), order="F")
    ora.call("lila_dlarft_w03", m, n, ii, mt, Ao, m, T1, mt, tau)
    Ad, Td, taud = to_dev(Ao), to_dev(np.zeros((mt, n + ii), order="F")), torch.from_numpy(tau).cuda()
    assert lila.call("lila_dlarft_w03", m, n, ii, mt, Ad, m, Td, mt, taud) == 0
    sync_ok(lila)
    assert rel(to_host(Td), T1) < 100 * n * U64
    # T2tau / V2tau on the GPU
    tg = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert lila.call("lila_dT2tau_w03", m, n, ii, mt, to_dev(To), mt, tg) == 0
    sync_ok(lila)
    assert np.array_equal(tg.cpu().numpy(), tau)
    tv = np.zeros(n)
    ora.call("lila_dV2tau_w03", m, n, ii, mt, Ao, m, tv)
    tg2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert lila.call("lila_dV2tau_w03", m, n, ii, mt, Ad, m, tg2) == 0
    sync_ok(lila)
    assert np.max(np.abs(tg2.cpu().numpy() - tv)) < 1e-13


def test_dlarft_connect(lila, ora):
    m, n, ii, mt = 500, 64, 0, 128
    say("connect")
    A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, 16])
    Ao, To, _ = o
    # zero the connecting block of columns [32, 64) and recompute it
    T1 = To.copy(order="F")
    T1[0:32, 32:64] = 0.0
    T2 = T1.copy(order="F")
    ora.call("lila_dlarft_connect_w03", m, 32, 32, 0, mt, Ao, m, T1, mt)
    Td = to_dev(T2)
    assert lila.call("lila_dlarft_connect_w03", m, 32, 32, 0, mt, to_dev(Ao), m, Td, mt) == 0
    sync_ok(lila)
    assert rel(to_host(Td), T1) < 100 * n * U64
    assert rel(T1, To) < 100 * n * U64


# ---------------------------------------------------------------------------
# Cholesky-QR + Householder reconstruction
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m,n,ii,mt", [(300, 40, 0, 64), (500, 96, 7, 32), (2000, 256, 0, 64)])
def test_mt_hr(lila, ora, m, n, ii, mt):
    say("mt_hr", m, n, ii, mt)
    A = ora.fill_normal(5, m, n + ii)
    Q = np.zeros((m, n + ii), order="F")
    T = np.zeros((mt, n + ii), order="F")
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_w03_mt_hr", m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_w03_mt_hr", m, n, ii, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    sync_ok(lila)
    Ag, Tg, Qg = to_host(Ad), to_host(Td), to_host(Qd)
    tol = 2000 * n * U64
    assert rel(Ag, Ao) < tol and rel(Qg, Qo) < tol and rel(Tg, To) < tol
    assert ora.call("test_qr_repres_1", m, n, ii, A, m, Qg, m, Ag, m) < tol


def test_dorgh2_3_and_getrf_b(lila, ora):
    m, n = 400, 48
    say("dorgh2_3 / dgetrf_b")
    A = ora.fill_normal(8, m, n)
    # orthonormal Q and R by the oracle's Cholesky QR
    Q = A.copy(order="F")
    R = np.zeros((n, n), order="F")
    ora.call("dsyrk", "U", "T", n, m, 1.0, Q, m, 0.0, R, n)
    assert ora.call("dpotrf_u", n, R, n) == 0
    ora.call("dtrsm", "R", "U", "N", "N", m, n, 1.0, R, n, Q, m)
    # dorgh2_3 inputs: A top = R, A below = Q2, T = Q1
    Ain = np.zeros((m, n), order="F")
    Ain[:n] = np.triu(R)
    Ain[n:] = Q[n:]
    T = np.asfortranarray(Q[:n].copy())
    S = np.zeros(n)
    Ao, To, Qo, So = Ain.copy(order="F"), T.copy(order="F"), Q.copy(order="F"), S.copy()
    ora.call("lila_dorgh2_3", m, n, Ao, m, To, n, Qo, m, None, 0, So)
    Ad, Td, Qd, Sd = to_dev(Ain), to_dev(T), to_dev(Q), torch.zeros(n, dtype=torch.float64, device="cuda")
    assert lila.call("lila_dorgh2_3", m, n, Ad, m, Td, n, Qd, m, None, 0, Sd) == 0
    sync_ok(lila)
    assert np.array_equal(Sd.cpu().numpy(), So)
    tol = 200 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Td), To) < tol and rel(to_host(Qd), Qo) < tol
    # dgetrf_b: A = copy of Q, Q flipped, work = R
    Ag_in = Q.copy(order="F")
    Ao2, Qo2, Uo = Ag_in.copy(order="F"), Q.copy(order="F"), R.copy(order="F")
    ora.call("lila_dgetrf_b", m, n, 0, 8, Ao2, m, Qo2, m, Uo, n * n)
    Ad2, Qd2, Ud = to_dev(Ag_in), to_dev(Q), to_dev(R)
    assert lila.call("lila_dgetrf_b", m, n, 0, 8, Ad2, m, Qd2, m, Ud, n * n) == 0
    sync_ok(lila)
    assert rel(to_host(Ad2), Ao2) < tol and rel(to_host(Qd2), Qo2) < tol and rel(to_host(Ud), Uo) < tol


def _orthonormal(ora, m, n, seed):
    A = ora.fill_normal(seed, m, n)
    Q = A.copy(order="F")
    R = np.zeros((n, n), order="F")
    ora.call("dsyrk", "U", "T", n, m, 1.0, Q, m, 0.0, R, n)
    assert ora.call("dpotrf_u", n, R, n) == 0
    ora.call("dtrsm", "R", "U", "N", "N", m, n, 1.0, R, n, Q, m)
    # second pass for orthogonality to rounding
    R2 = np.zeros((n, n), order="F")
    ora.call("dsyrk", "U", "T", n, m, 1.0, Q, m, 0.0, R2, n)
    assert ora.call("dpotrf_u", n, R2, n) == 0
    ora.call("dtrsm", "R", "U", "N", "N", m, n, 1.0, R2, n, Q, m)
    return Q, np.asfortranarray(np.triu(R2) @ np.triu(R))


@pytest.mark.parametrize("m,n,i,j,mt", [(300, 24, 0, 0, 32), (420, 16, 5, 44, 20), (500, 40, 11, 64, 64)])
def test_dorgh2_global(lila, ora, m, n, i, j, mt):
    say("dorgh2", m, n, i, j, mt)
    Qb, R = _orthonormal(ora, m - j, n, 21)
    A = ora.fill_uniform(3, m, j + n, -1.0, 1.0)
    A[j:j + n, j:j + n] = R
    A[j + n:, j:j + n] = Qb[n:]
    T = ora.fill_uniform(4, mt, j + n, -1.0, 1.0)
    T[j % mt:j % mt + n, j:j + n] = Qb[:n]
    Q = ora.fill_uniform(5, m, j + n, -1.0, 1.0)
    Q[j:, j:j + n] = Qb
    So = np.zeros(j + n)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dorgh2", m, n, i, j, mt, Ao, m, To, mt, Qo, m, None, 0, So)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    Sd = torch.zeros(j + n, dtype=torch.float64, device="cuda")
    assert lila.call("lila_dorgh2", m, n, i, j, mt, Ad, m, Td, mt, Qd, m, None, 0, Sd) == 0
    sync_ok(lila)
    assert np.array_equal(Sd.cpu().numpy()[j:], So[j:])
    tol = 200 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Td), To) < tol and rel(to_host(Qd), Qo) < tol


@pytest.mark.parametrize("m,n,i,mt", [(400, 40, 0, 16), (600, 70, 9, 32), (500, 50, 3, 64),
                                      (3000, 700, 13, 128)])
def test_dorghr_blocked(lila, ora, m, n, i, mt):
    say("dorghr", m, n, i, mt)
    Qb, R = _orthonormal(ora, m - i, n, 23)
    A = np.zeros((m, i + n), order="F")
    A[i:i + n, i:] = R
    Q = np.zeros((m, i + n), order="F")
    Q[i:, i:] = Qb
    T = np.zeros((mt, i + n), order="F")
    So = np.zeros(i + n, dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dorghr", m, n, i, mt, Ao, m, To, mt, Qo, m, None, 0, So)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    Sd = torch.zeros(i + n, dtype=torch.int32, device="cuda")
    assert lila.call("lila_dorghr", m, n, i, mt, Ad, m, Td, mt, Qd, m, None, 0, Sd) == 0
    sync_ok(lila)
    assert np.array_equal(Sd.cpu().numpy()[i:], So[i:])
    tol = 500 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Td), To) < tol and rel(to_host(Qd), Qo) < tol


def test_dgels_qr3(lila, ora):
    m, n, k = 800, 60, 5
    say("dgels_qr3")
    A = ora.fill_normal(21, m, n)
    B = ora.fill_normal(22, m, k)
    Ao, Bo, To = A.copy(order="F"), B.copy(order="F"), np.zeros((n, n), order="F")
    ora.call("ls_qr3", m, n, k, Ao, m, To, n, Bo, m, np.zeros((n, k), order="F"))
    Ad, Bd, Td = to_dev(A), to_dev(B), to_dev(np.zeros((n, n), order="F"))
    assert lila.call("lila_dgels_qr3", m, n, k, Ad, m, Td, n, Bd, m) == 0
    sync_ok(lila)
    X = to_host(Bd)[:n]
    assert rel(X, Bo[:n]) < 1000 * n * U64
    Xl = np.linalg.lstsq(A, B, rcond=None)[0]
    assert rel(X, Xl) < 1000 * n * U64
