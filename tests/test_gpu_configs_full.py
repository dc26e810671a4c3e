"""GPU tests at the full sizes of the two headline workloads.

* Mixed precision against plain FP64 at m = n = 16384 with
  A = U diag(sigma) V^T, U and V the Q factors (plain FP64 path) of seeded
  Gaussians, sigma geometric from 1 to 1/kappa, kappa in {1e2, 1e5, 1e8}:
  for both forms ||A - QR||_F/||A||_F <= 1e-13 and the per-column normalised
  orthogonality ||I - Q^T Q||_F / sqrt(n) <= 1e-13; the mixed form's errors
  within 20x the plain form's, and its R row by row against the plain R.
* The batched least-squares workload: 4096 problems, m ~ U{64..4096},
  n ~ U{8..min(128, m)}, nrhs ~ U{1..16}, N(0,1), through the plan and its
  hipGraph (one persistent launch; about 16 problems per workgroup).  A fixed
  sample of 256 problems against the oracle's per-problem LAPACK dgels
  restatement, and every problem against the criteria of
  least-squares/main_ls_qr3.c:157-168, evaluated on the GPU.
* The persistent queue with far more problems than workgroups (the grid
  capped at 7): bit-identical to the uncapped launch.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from conftest import U64, say
from test_gpu_batched import Batch, check_against_oracle, make_sizes

pytestmark = pytest.mark.gpu
f64 = torch.float64


def dgemm(lila, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    r = lila.call("lila_dgemm", ctypes.c_char(ta.encode()), ctypes.c_char(tb.encode()), m, n, k,
                  alpha, A, lda, B, ldb, beta, C, ldc)
    assert r == 0


def qr16384(lila, mode, src, A, T, Q, n, mt, nx):
    p = np.array([mode, 1, 1, nx], dtype=np.int32)
    A.copy_(src)
    r = lila.call("lila_dgeqrf_w03_recursive", p, n, n, 0, mt, A, n, T, mt, Q, n, None, 0)
    assert r == 0, r
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0


def errors(lila, A0, A, Q, n):
    """||A0 - QR||_F/||A0||_F and ||I - Q^T Q||_F for column-major (n, n)
    storage (torch row c = column c)"""
    R = torch.tril(A)  # upper triangle of the column-major A
    W = torch.empty_like(A0)
    dgemm(lila, "N", "N", n, n, n, 1.0, Q, n, R, n, 0.0, W, n)
    torch.cuda.synchronize()
    bwd = (torch.linalg.norm(W - A0) / torch.linalg.norm(A0)).item()
    W.copy_(torch.eye(n, dtype=f64, device=A0.device))
    dgemm(lila, "T", "N", n, n, n, 1.0, Q, n, Q, n, -1.0, W, n)
    torch.cuda.synchronize()
    orth = torch.linalg.norm(W).item()
    return bwd, orth


@pytest.fixture(scope="module")
def uv16384(lila):
    n, mt, nx = 16384, 1024, 256
    dev = "cuda"
    G = torch.empty((n, n), dtype=f64, device=dev)
    A = torch.empty_like(G)
    Q = torch.empty_like(G)
    T = torch.zeros((n, mt), dtype=f64, device=dev)
    UV = []
    for seed in (101, 202):
        assert lila.call("lila_fill_normal", ctypes.c_ulonglong(seed), n, n, G, n) == 0
        qr16384(lila, 0, G, A, T, Q, n, mt, nx)
        UV.append(Q.clone())
    del G
    yield UV, A, Q, T
    del UV


@pytest.mark.parametrize("kappa", [1e2, 1e5, 1e8])
def test_mixed_vs_plain_16384(lila, uv16384, kappa):
    (U, V), A, Q, T = uv16384
    n, mt, nx = 16384, 1024, 256
    say("mixed vs plain 16384, kappa", kappa)
    idx = torch.arange(n, dtype=f64, device="cuda")
    sig = kappa ** (-idx / (n - 1))
    US = U * sig[:, None]  # column j of the column-major U times sigma_j
    A0 = torch.empty_like(U)
    dgemm(lila, "N", "T", n, n, n, 1.0, US, n, V, n, 0.0, A0, n)
    del US
    res, Rs = {}, {}
    for mode in (0, 1):
        qr16384(lila, mode, A0, A, T, Q, n, mt, nx)
        res[mode] = errors(lila, A0, A, Q, n)
        Rs[mode] = torch.tril(A).clone()
    (bp, op), (bm, om) = res[0], res[1]
    # R of the two forms row by row (rows may differ in sign only), relative
    # to ||A||_F; a backward-stable R moves by up to kappa u ||A|| rowwise
    sgn = torch.sign(torch.diagonal(Rs[0])) * torch.sign(torch.diagonal(Rs[1]))
    dR = (torch.max(torch.abs(Rs[1] * sgn[:, None].T - Rs[0])) / torch.linalg.norm(A0)).item()
    del Rs, A0
    say("  plain bwd %.2e orth %.2e (/sqrt n %.2e) | mixed bwd %.2e orth %.2e (/sqrt n %.2e) | "
        "max|dR|/||A|| %.2e" % (bp, op, op / math.sqrt(n), bm, om, om / math.sqrt(n), dR))
    for b, o in ((bp, op), (bm, om)):
        assert b <= 1e-13
        assert o / math.sqrt(n) <= 1e-13
    assert bm < 20 * bp + 1e-15 and om < 20 * op + 1e-14
    assert dR <= max(100 * n * U64, 10 * kappa * U64)


def test_config5_full_batch(lila, ora):
    """4096 problems through the plan's hipGraph; at one workgroup per CU each
    workgroup takes about 16 problems from the queue"""
    ora.use_host_blas()
    nprob, seed = 4096, 20241204
    rng = np.random.default_rng(seed)
    m = rng.integers(64, 4097, size=nprob)
    n = np.array([rng.integers(8, min(128, int(mm)) + 1) for mm in m])
    k = rng.integers(1, 17, size=nprob)
    sizes = [(int(a), int(b), int(c), int(a), int(a)) for a, b, c in zip(m, n, k)]
    say("config 5 batch: building", nprob, "problems")
    b = Batch(sizes, seed + 1)
    lib = lila.load()
    arrs = b.arrays()  # kept alive: the call reads the arrays behind these pointers
    plan = lib.lila_dgels_batched_plan(*[lila.conv(a) for a in arrs])
    assert plan
    plan = ctypes.c_void_p(plan)
    try:
        say("config 5 batch: graph launch")
        assert lib.lila_dgels_batched_run(plan, 1) == 0
        assert lila.call("lila_sync") == 0
        got = b.buf.clone()
        # a second replay from the pristine input repeats bit for bit
        b.buf.copy_(b.orig)
        torch.cuda.synchronize()
        assert lib.lila_dgels_batched_run(plan, 1) == 0
        assert lila.call("lila_sync") == 0
        assert torch.equal(b.buf, got)
    finally:
        lib.lila_dgels_batched_destroy(plan)
    # every problem: least-squares/main_ls_qr3.c:157-168 on the GPU
    say("config 5 batch: criteria on every problem")
    worst_orth, worst_rep = 0.0, -1.0
    for p, (mm, nn, kk, lda, ldb) in enumerate(sizes):
        oa, ob = b.offA[p], b.offB[p]
        A0 = b.orig[oa:oa + lda * nn].view(nn, lda).T[:mm]
        B0 = b.orig[ob:ob + ldb * kk].view(kk, ldb).T[:mm]
        Xf = got[ob:ob + ldb * kk].view(kk, ldb).T[:mm]
        x, low = Xf[:nn], Xf[nn:]
        r = A0 @ x - B0
        nA, nx_, nb = torch.linalg.norm(A0), torch.linalg.norm(x), torch.linalg.norm(B0)
        orth = (torch.linalg.norm(A0.T @ r) / (nA * (nA * nx_ + nb))).item()
        rep = ((torch.linalg.norm(r) - torch.linalg.norm(low)) / (nA * nx_ + nb)).item()
        worst_orth = max(worst_orth, orth)
        worst_rep = max(worst_rep, abs(rep))
        assert orth < 10 * nn * U64, (p, mm, nn, kk, orth)
        assert abs(rep) < 10 * nn * U64, (p, mm, nn, kk, rep)
    say("  worst ||A^T r|| criterion %.2e, worst residual criterion %.2e" % (worst_orth, worst_rep))
    # a fixed sample of 256 problems (and the 8 largest) against the oracle
    samp = sorted(set(np.random.default_rng(7).choice(nprob, size=256, replace=False).tolist())
                  | set(np.argsort(m * n * n)[-8:].tolist()))
    say("config 5 batch: oracle dgels on", len(samp), "problems")
    sub = Batch([sizes[p] for p in samp], 0)
    host_orig = b.orig.cpu().numpy()
    host_got = got.cpu().numpy()
    sub.host0 = np.concatenate([np.concatenate([
        host_orig[b.offA[p]:b.offA[p] + sizes[p][3] * sizes[p][1]],
        host_orig[b.offB[p]:b.offB[p] + sizes[p][4] * max(sizes[p][2], 1)]]) for p in samp])
    out = np.concatenate([np.concatenate([
        host_got[b.offA[p]:b.offA[p] + sizes[p][3] * sizes[p][1]],
        host_got[b.offB[p]:b.offB[p] + sizes[p][4] * max(sizes[p][2], 1)]]) for p in samp])
    say("  worst factor error / tol", check_against_oracle(ora, sub, out))
    ora.use_builtin_blas()


def test_batched_queue_far_more_problems_than_workgroups(lila, ora):
    """300 problems on a grid capped at 7 workgroups: each workgroup takes
    about 43 problems from the queue, reusing its LDS across problems; the
    results are bit-identical to the uncapped launch and match the oracle"""
    sizes = make_sizes(77, 300, mmax=500)
    b = Batch(sizes, 78)
    say("batched queue: default grid")
    assert lila.call("lila_dgels_batched", *b.arrays()) == 0
    assert lila.call("lila_sync") == 0
    ref = b.buf.clone()
    b.buf.copy_(b.orig)
    torch.cuda.synchronize()
    old = lila.call("lila_set_batched_grid", 7)
    try:
        say("batched queue: grid capped at 7")
        assert lila.call("lila_dgels_batched", *b.arrays()) == 0
        assert lila.call("lila_sync") == 0
    finally:
        lila.call("lila_set_batched_grid", old)
    assert torch.equal(b.buf, ref)
    check_against_oracle(ora, b, b.buf.cpu().numpy())
