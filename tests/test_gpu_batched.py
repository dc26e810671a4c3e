"""GPU parity of the batched least-squares path (lila_dgels_batched and the
plan/graph form) against the oracle's per-problem LAPACK dgels('N')
restatement, through the C ABI of liblila_hip.so."""
import ctypes

import numpy as np
import pytest

from conftest import U64, say

pytestmark = pytest.mark.gpu


def make_sizes(seed, nprob, mmax=700):
    rng = np.random.default_rng(seed)
    sizes = []
    for _ in range(nprob):
        m = int(rng.integers(64, mmax + 1))
        n = int(rng.integers(1, min(128, m) + 1))
        k = int(rng.integers(0, 17))
        sizes.append((m, n, k, m + int(rng.integers(0, 4)), m + int(rng.integers(0, 3))))
    return sizes


EDGE = [(64, 64, 1, 64, 64), (128, 128, 16, 130, 128), (4096, 128, 16, 4096, 4096),
        (3000, 8, 1, 3001, 3000), (200, 1, 16, 200, 200), (65, 16, 0, 65, 65),
        (97, 33, 5, 100, 99), (64, 17, 16, 64, 64)]


class Batch:
    """problems carved out of one flat float64 device buffer"""

    def __init__(self, sizes, seed):
        import torch
        rng = np.random.default_rng(seed)
        self.sizes = sizes
        self.offA, self.offB, off = [], [], 0
        for (m, n, k, lda, ldb) in sizes:
            self.offA.append(off)
            off += lda * n
            self.offB.append(off)
            off += ldb * max(k, 1)
        self.host0 = rng.standard_normal(off)
        self.buf = torch.from_numpy(self.host0.copy()).cuda()
        self.orig = self.buf.clone()

    def arrays(self):
        nprob = len(self.sizes)
        base = self.buf.data_ptr()
        m = np.array([s[0] for s in self.sizes], dtype=np.int32)
        n = np.array([s[1] for s in self.sizes], dtype=np.int32)
        k = np.array([s[2] for s in self.sizes], dtype=np.int32)
        lda = np.array([s[3] for s in self.sizes], dtype=np.int32)
        ldb = np.array([s[4] for s in self.sizes], dtype=np.int32)
        pa = np.array([base + 8 * o for o in self.offA], dtype=np.uint64)
        pb = np.array([base + 8 * o for o in self.offB], dtype=np.uint64)
        assert len(m) == nprob
        return nprob, m, n, k, pa, lda, pb, ldb

    def views(self, flat, p):
        m, n, k, lda, ldb = self.sizes[p]
        A = flat[self.offA[p]:self.offA[p] + lda * n].reshape((lda, n), order="F")
        B = flat[self.offB[p]:self.offB[p] + ldb * max(k, 1)].reshape((ldb, max(k, 1)), order="F")
        return A, B


def check_against_oracle(ora, batch, out):
    worst = 0.0
    for p, (m, n, k, lda, ldb) in enumerate(batch.sizes):
        A0, B0 = batch.views(batch.host0, p)
        Ag, Bg = batch.views(out, p)
        Ao, Bo = A0.copy(order="F"), B0.copy(order="F")
        tau = np.zeros(n)
        assert ora.call("dgels_hh", m, n, k, Ao, lda, Bo, ldb, tau) == 0
        # padding rows below m untouched
        assert np.array_equal(Ag[m:], A0[m:]) and np.array_equal(Bg[m:], B0[m:]), p
        tol = 100 * n * U64
        e = np.linalg.norm(Ag[:m] - Ao[:m]) / np.linalg.norm(Ao[:m])
        assert e < tol, (p, m, n, k, e)
        worst = max(worst, e / tol)
        if k == 0:
            assert np.array_equal(Bg, B0), p
            continue
        cond = np.linalg.cond(A0[:m])
        ex = np.linalg.norm(Bg[:m, :k] - Bo[:m, :k]) / np.linalg.norm(Bo[:m, :k])
        assert ex < tol * max(1.0, cond), (p, m, n, k, ex, cond)
        x = Bg[:n, :k]
        r = A0[:m] @ x - B0[:m, :k]
        nA = np.linalg.norm(A0[:m])
        orth = np.linalg.norm(A0[:m].T @ r) / (nA * (nA * np.linalg.norm(x) + np.linalg.norm(B0[:m, :k])))
        assert orth < 10 * n * U64, (p, orth)
    return worst


def test_batched_device_pointers(ora, lila):
    sizes = make_sizes(11, 56) + EDGE
    b = Batch(sizes, 12)
    say("batched device", len(sizes), "problems")
    r = lila.call("lila_dgels_batched", *b.arrays())
    assert r == 0
    assert lila.call("lila_sync") == 0
    out = b.buf.cpu().numpy()
    say("worst factor error / tol", check_against_oracle(ora, b, out))


def test_batched_graded_columns_take_the_column_loop(ora, lila):
    """columns scaled over 8 orders of magnitude: the panels' Gram matrices
    fail the Cholesky acceptance test (diagonal of R spread beyond 1e5), so
    these panels are factored column by column with T from dlarft"""
    import torch
    sizes = [(300, 40, 3, 300, 300), (700, 128, 16, 701, 700), (96, 70, 1, 96, 96), (2000, 33, 2, 2000, 2000)]
    b = Batch(sizes, 41)
    for p, (m, n, k, lda, ldb) in enumerate(sizes):
        A, _ = b.views(b.host0, p)
        A[:, :] *= np.logspace(0, -8, n)[None, :]
    b.buf.copy_(torch.from_numpy(b.host0))
    say("batched graded", len(sizes), "problems")
    assert lila.call("lila_dgels_batched", *b.arrays()) == 0
    assert lila.call("lila_sync") == 0
    check_against_oracle(ora, b, b.buf.cpu().numpy())


def test_batched_host_pointers(ora, lila):
    sizes = make_sizes(21, 12, mmax=300) + EDGE[:2]
    b = Batch(sizes, 22)
    host = b.host0.copy()
    nprob, m, n, k, _, lda, _, ldb = b.arrays()
    base = host.ctypes.data
    pa = np.array([base + 8 * o for o in b.offA], dtype=np.uint64)
    pb = np.array([base + 8 * o for o in b.offB], dtype=np.uint64)
    say("batched host", nprob, "problems")
    assert lila.call("lila_dgels_batched", nprob, m, n, k, pa, lda, pb, ldb) == 0
    check_against_oracle(ora, b, host)


def test_batched_plan_graph_replays(ora, lila):
    import torch
    sizes = make_sizes(31, 40, mmax=1200) + EDGE[:4]
    b = Batch(sizes, 32)
    lib = lila.load()
    arrs = b.arrays()  # kept alive: the call reads the arrays behind these pointers
    plan = lib.lila_dgels_batched_plan(*[lila.conv(a) for a in arrs])
    assert plan
    plan = ctypes.c_void_p(plan)
    try:
        assert lib.lila_dgels_batched_flops(plan) > 0
        say("batched plan eager")
        assert lib.lila_dgels_batched_run(plan, 0) == 0
        assert lila.call("lila_sync") == 0
        eager = b.buf.cpu().numpy().copy()
        check_against_oracle(ora, b, eager)
        for rep in range(3):
            b.buf.copy_(b.orig)
            torch.cuda.synchronize()
            say("batched plan graph replay", rep)
            assert lib.lila_dgels_batched_run(plan, 1) == 0
            assert lila.call("lila_sync") == 0
            got = b.buf.cpu().numpy()
            assert np.array_equal(got, eager), rep
            # eager work between replays
            b.buf.copy_(b.orig)
            torch.cuda.synchronize()
            assert lib.lila_dgels_batched_run(plan, 0) == 0
            assert lila.call("lila_sync") == 0
            assert np.array_equal(b.buf.cpu().numpy(), eager), rep
    finally:
        lib.lila_dgels_batched_destroy(plan)


def test_batched_rejects_out_of_range(lila):
    import torch
    buf = torch.zeros(200 * 130 + 200 * 17, dtype=torch.float64, device="cuda")
    base = buf.data_ptr()
    pa = np.array([base], dtype=np.uint64)
    pb = np.array([base + 8 * 200 * 130], dtype=np.uint64)
    for (m, n, k) in [(200, 129, 1), (100, 120, 1), (200, 10, 17), (200, 0, 1)]:
        r = lila.call("lila_dgels_batched", 1, np.array([m], np.int32), np.array([n], np.int32),
                      np.array([k], np.int32), pa, np.array([200], np.int32), pb,
                      np.array([200], np.int32))
        assert r == -1, (m, n, k)
    assert lila.call("lila_dgels_batched", 0, None, None, None, None, None, None, None) == 0


def _srel(a, b):
    s = float(np.max(np.abs(b)))
    return np.linalg.norm(a / s - b / s) / max(np.linalg.norm(b / s), 1e-300)


@pytest.mark.parametrize("scale", [1e-200, 1e200])
def test_batched_extreme_scales(ora, lila, scale):
    """A scaled by 1e-200 or 1e200: every panel's Gram underflows to zero or
    overflows, the panels take the column loop, whose norms and dot products
    run on the column scaled by a power of two near 1 / max|x| (dlarfg's
    safe scaling); V, R and X against the oracle's dgels relative to their
    largest entries, everything finite"""
    import torch
    sizes = [(300, 40, 3, 300, 300), (700, 100, 5, 701, 700), (97, 33, 2, 97, 97)]
    b = Batch(sizes, 51)
    for p, (m, n, k, lda, ldb) in enumerate(sizes):
        A, _ = b.views(b.host0, p)
        A[:, :] *= scale
    b.buf.copy_(torch.from_numpy(b.host0))
    say("batched extreme scale", scale)
    assert lila.call("lila_dgels_batched", *b.arrays()) == 0
    assert lila.call("lila_sync") == 0
    out = b.buf.cpu().numpy()
    assert np.all(np.isfinite(out))
    for p, (m, n, k, lda, ldb) in enumerate(sizes):
        A0, B0 = b.views(b.host0, p)
        Ag, Bg = b.views(out, p)
        Ao, Bo = A0.copy(order="F"), B0.copy(order="F")
        tau = np.zeros(n)
        assert ora.call("dgels_hh", m, n, k, Ao, lda, Bo, ldb, tau) == 0
        tol = 100 * n * U64
        assert _srel(np.tril(Ag[:m], -1), np.tril(Ao[:m], -1)) < tol, ("V", p)
        assert _srel(np.triu(Ag[:n]), np.triu(Ao[:n])) < tol, ("R", p)
        cond = np.linalg.cond(A0[:m] / scale)
        assert _srel(Bg[:n, :k], Bo[:n, :k]) < tol * max(1.0, cond), ("X", p)
