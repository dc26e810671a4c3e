"""GPU: the FP32-MFMA GEMM and the mixed-precision QR (lila_param[0] == 1).

The FP32 GEMM is checked against the FP64 product with the FP32 error bound
(operands rounded to FP32, FP32 accumulation): |C - C64| <= 8 (k + 2) u32
(|A||B|).  The mixed QR must give the same outputs as the FP64 path: for a
well-conditioned matrix its V, R, T and Q agree with the oracle's to c n u;
for A = U diag(sigma) V^T with kappa in {1e2, 1e5, 1e8} its backward error
and orthogonality stay at FP64 level (below 1e-13 per column normalised
Frobenius norm), like the FP64 path's on the same matrix."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host
from test_gpu_parity import rel, sync_ok

pytestmark = pytest.mark.gpu
U32 = 2.0 ** -24


@pytest.fixture
def lowp(lila):
    def set_lowp(v):
        return lila.call("lila_set_gemm_lowp", v)
    yield set_lowp
    lila.call("lila_set_gemm_lowp", 0)


@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("m,n,k", [(17, 33, 5), (130, 70, 300), (1000, 513, 64), (64, 96, 20000)])
def test_fp32_mfma_gemm(lila, ora, lowp, ta, tb, m, n, k):
    say("fp32 gemm", ta, tb, m, n, k)
    rng = np.random.default_rng(m + n + k)
    A = np.asfortranarray(rng.standard_normal((k, m) if ta == "T" else (m, k)))
    B = np.asfortranarray(rng.standard_normal((n, k) if tb == "T" else (k, n)))
    C = np.asfortranarray(rng.standard_normal((m, n)))
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    ref = -1.5 * opA @ opB + 0.5 * C
    lowp(1)
    Cd = to_dev(C)
    r = lila.call("lila_dgemm", ctypes.c_char(ta.encode()), ctypes.c_char(tb.encode()), m, n, k, -1.5,
                  to_dev(A), A.shape[0], to_dev(B), B.shape[0], 0.5, Cd, m)
    assert r == 0
    sync_ok(lila)
    got = to_host(Cd)
    bound = 8 * (k + 2) * U32 * (1.5 * np.abs(opA) @ np.abs(opB)) + 4 * U32 * np.abs(C) + 1e-300
    assert np.all(np.abs(got - ref) <= bound)
    # and it is really FP32: the error is far above FP64 rounding
    assert np.max(np.abs(got - ref)) > 1e3 * U64 * np.max(np.abs(ref))


def run_qr(lila, A, mode, mt, nx):
    m, n = A.shape
    Ad, Td, Qd = to_dev(A), to_dev(np.zeros((mt, n), order="F")), to_dev(np.zeros((m, n), order="F"))
    p = np.array([mode, 1, 1, nx], dtype=np.int32)
    assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    sync_ok(lila)
    return to_host(Ad), to_host(Td), to_host(Qd)


def errors(A, Ao, Qo):
    n = A.shape[1]
    R = np.triu(Ao[:n])
    bwd = np.linalg.norm(A - Qo @ R) / np.linalg.norm(A)
    orth = np.linalg.norm(np.eye(n) - Qo.T @ Qo)
    return bwd, orth


def test_mixed_matches_the_fp64_outputs(lila, ora):
    """a well-conditioned matrix: the mixed path's V, R, T and Q are the FP64
    path's (the reconstruction picks dlarfg's signs)"""
    m, n, mt, nx = 1500, 700, 64, 32
    say("mixed vs oracle")
    ora.use_host_blas()
    try:
        A = ora.fill_uniform(11, m, n, -1.0, 1.0)
        Ao, To, Qo = A.copy(order="F"), np.zeros((mt, n), order="F"), np.zeros((m, n), order="F")
        p = np.array([0, 1, 1, nx], dtype=np.int32)
        ora.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, Ao, m, To, mt, Qo, m, None, 0)
        Ag, Tg, Qg = run_qr(lila, A, 1, mt, nx)
        tol = 200 * n * U64
        assert rel(Ag, Ao) < tol
        assert rel(Qg, Qo) < tol
        assert rel(Tg, To) < 10 * tol
        k = n  # every reconstructed reflector is used (lila_test_r_repres_2.c:26-33)
        assert ora.call("test_r_repres_2", m, n, k, 0, mt, A, m, Ag, m, Tg, mt) < tol
    finally:
        ora.use_builtin_blas()


@pytest.mark.parametrize("kappa", [1e2, 1e5, 1e8])
def test_mixed_accuracy_per_condition_number(lila, kappa):
    """kappa = 1e2 leaves Q1 = A R0^-1 close enough to orthonormal
    (||I - Q1^T Q1||_F <= 0.5) for one Cholesky-QR pass; kappa = 1e8 needs
    the second -- both branches must reach the FP64 path's accuracy"""
    m, n, mt, nx = 2000, 800, 64, 32
    say("mixed kappa", kappa)
    rng = np.random.default_rng(int(np.log10(kappa)))
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sig = kappa ** (-np.arange(n) / (n - 1))
    A = np.asfortranarray((U * sig) @ V.T)
    Ap, _, Qp = run_qr(lila, A, 0, mt, nx)
    Am, _, Qm = run_qr(lila, A, 1, mt, nx)
    bp, op = errors(A, Ap, Qp)
    bm, om = errors(A, Am, Qm)
    say("  plain bwd %.2e orth %.2e | mixed bwd %.2e orth %.2e" % (bp, op, bm, om))
    assert bm < 1e-13 and om < 1e-13
    assert bm < 20 * bp + 1e-15 and om < 20 * op + 1e-14


@pytest.mark.parametrize("kappa", [1e2, 1e5, 1e8])
def test_mixed_tall_200000_by_1024(lila, kappa):
    """the tall shape of the mixed-precision workload, m = 200000, n = 1024,
    A = U diag(sigma) V^T (U, V from torch's FP64 QR of seeded Gaussians):
    backward error of both forms within 1e-13; ||I - Q^T Q||_F (which grows
    with n: the plain form gives ~3 n u here) within 30 n u, LAPACK's test
    threshold on such ratios; the mixed form's errors within 20x the plain
    form's"""
    m, n, mt, nx = 200_000, 1024, 512, 32
    say("mixed tall", m, n, kappa)
    f64 = torch.float64
    g = torch.Generator(device="cuda").manual_seed(int(np.log10(kappa)))
    U, _ = torch.linalg.qr(torch.randn(m, n, dtype=f64, device="cuda", generator=g))
    V, _ = torch.linalg.qr(torch.randn(n, n, dtype=f64, device="cuda", generator=g))
    sig = kappa ** (-torch.arange(n, dtype=f64, device="cuda") / (n - 1))
    A0 = ((U * sig) @ V.T.contiguous()).T.contiguous()  # column-major storage: (n, m)
    del U, V
    nA = torch.linalg.norm(A0).item()
    res = {}
    for mode in (0, 1):
        A = A0.clone()
        Q = torch.zeros_like(A0)
        T = torch.zeros((n, mt), dtype=f64, device="cuda")
        p = np.array([mode, 1, 1, nx], dtype=np.int32)
        assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, A, m, T, mt, Q, m, None, 0) == 0
        torch.cuda.synchronize()
        assert lila.call("lila_get_info") == 0
        Qm = Q.T  # (m, n) view
        R = torch.triu(A.T[:n])
        bwd = torch.linalg.norm(A0.T - Qm @ R).item() / nA
        orth = torch.linalg.norm(Qm.T @ Qm - torch.eye(n, dtype=f64, device="cuda")).item()
        res[mode] = (bwd, orth)
        del A, Q, T, R
    (bp, op), (bm, om) = res[0], res[1]
    say("  plain bwd %.2e orth %.2e | mixed bwd %.2e orth %.2e" % (bp, op, bm, om))
    assert bp < 1e-13 and op < 30 * n * U64
    assert bm < 1e-13 and om < 30 * n * U64
    assert bm < 20 * bp + 1e-15 and om < 20 * op + 1e-14
