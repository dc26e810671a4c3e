"""The Gram matrices and the Cholesky-QR2 forms of the tall-skinny
Cholesky-QR path: lila_dgram in FP64, FP32 and split-BF16 arithmetic against
FP64 products and the numpy model of the split (tests/bf16x3.py); the hr
driver with each Gram precision against the oracle's hr restatement; and the
m = 1,000,000, n = 256 path with A = U diag(sigma) V^T, kappa in {1e1, 1e3,
1e6}, checked on the GPU through orthogonality and the backward error."""
import numpy as np
import pytest
import torch

import bf16x3
from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture
def gram_prec(lila):
    """restores the default (one FP64 pass) after a test changes it"""
    yield
    lila.call("lila_set_gram_precision", 0)


@pytest.mark.parametrize("m,n", [(8192, 256), (8191, 17), (8192, 1), (20000, 200), (100, 128)])
def test_gram_precisions(lila, m, n):
    say("lila_dgram", m, n)
    rng = np.random.default_rng(m + n)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    G64 = np.triu(A.T @ A)
    out = {}
    for prec in (1, 2, 3):
        Gd = to_dev(np.zeros((n, n)))
        assert lila.call("lila_dgram", prec, m, n, to_dev(A), m, Gd, n) == 0
        torch.cuda.synchronize()
        assert lila.call("lila_get_info") == 0
        out[prec] = np.triu(to_host(Gd))
    assert rel(out[1], G64) < 100 * m * U64
    # FP32 operands and accumulation: visibly not FP64, close to it
    assert 1e-12 < rel(out[2], G64) < 1e-5
    # split BF16: the model of the split to FP32 accumulation, and FP64 to 1e-6
    assert rel(out[3], np.triu(bf16x3.gram3(A))) < 1e-6
    assert rel(out[3], G64) < 1e-6
    if n > 1:
        # a model with the mid.mid product left out is further away
        assert rel(out[3], np.triu(bf16x3.gram3(A, drop=("mm",)))) > 1e-6


def test_gram_rejects(lila):
    G = to_dev(np.zeros((4, 4)))
    assert lila.call("lila_dgram", 4, 10, 4, to_dev(np.zeros((10, 4))), 10, G, 4) == -1
    assert lila.call("lila_set_gram_precision", 7) == -1


@pytest.mark.parametrize("prec", [1, 2, 3])
@pytest.mark.parametrize("m,n,ii,mt", [(5000, 100, 0, 128), (3000, 256, 17, 300), (700, 64, 5, 64)])
def test_hr_cholqr2_parity(lila, ora, gram_prec, prec, m, n, ii, mt):
    say("mt_hr gram precision", prec, m, n, ii, mt)
    A = ora.fill_normal(11, m, n + ii)
    Q = np.zeros((m, n + ii), order="F")
    T = np.zeros((mt, n + ii), order="F")
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_w03_mt_hr", m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    assert lila.call("lila_set_gram_precision", prec) == 0
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_w03_mt_hr", m, n, ii, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    Ag, Tg, Qg = to_host(Ad), to_host(Td), to_host(Qd)
    tol = 2000 * n * U64
    assert rel(Ag, Ao) < tol and rel(Qg, Qo) < tol and rel(Tg, To) < tol
    assert ora.call("test_qq_orth_full", m, n, ii, Qg, m) < tol
    assert ora.call("test_qr_repres_1", m, n, ii, A, m, Qg, m, Ag, m) < tol


def tall_matrix(m, n, kappa, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    U, _ = torch.linalg.qr(torch.randn(m, n, dtype=torch.float64, device="cuda", generator=g))
    V, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g))
    sig = kappa ** (-torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1))
    return (U * sig) @ V.T.contiguous()


@pytest.mark.parametrize("kappa", [1e1, 1e3, 1e6])
def test_tall_cholqr_million_by_256(lila, gram_prec, kappa):
    """m = 1e6, n = 256: each Gram precision with the second FP64 pass gives
    an orthonormal Q and A = QR to FP64 accuracy where its first Gram
    resolves sigma_min^2 (FP64: every kappa here; FP32 and split BF16, whose
    first Gram carries ~1e-7 relative error: kappa <= 1e3)"""
    m, n, mt = 1_000_000, 256, 256
    say("tall cholqr", m, n, kappa)
    A0 = tall_matrix(m, n, kappa, 7)  # row-major (m, n) tensor
    A0c = A0.T.contiguous()  # column-major storage: (n, m)
    nA = torch.linalg.norm(A0).item()
    tol = 100 * n * U64
    for prec in (1, 2, 3, 0):
        lila.call("lila_set_gram_precision", prec)
        A = A0c.clone()
        Q = torch.zeros_like(A0c)
        T = torch.zeros((n, mt), dtype=torch.float64, device="cuda")
        r = lila.call("lila_dgeqrf_w03_mt_hr", m, n, 0, mt, A, m, T, mt, Q, m, None, 0)
        torch.cuda.synchronize()
        info = lila.call("lila_get_info")
        Qm = Q.T  # (m, n) view
        R = torch.triu(A.T[:n])
        orth = torch.linalg.norm(Qm.T @ Qm - torch.eye(n, dtype=torch.float64, device="cuda")).item()
        bwd = torch.linalg.norm(A0 - Qm @ R).item() / nA
        say("  gram precision", prec, "ret", r, "info", info, "orth %.2e bwd %.2e" % (orth, bwd))
        # one pass (0): orthogonality ~ kappa^2 u; FP32 and split-BF16 first
        # passes at kappa = 1e6: a breakdown or a Q far from orthonormal, the
        # limit of a ~1e-7 Gram -- both reported only
        if prec == 1 or (prec in (2, 3) and kappa <= 1e3):
            assert r == 0 and info == 0, (prec, kappa, r, info)
            assert orth < tol and bwd < tol, (prec, kappa, orth, bwd)
        del A, Q, T
