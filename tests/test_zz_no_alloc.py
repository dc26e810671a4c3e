"""lila_dgels_batched_noalloc (the batched least-squares launch with its
problem table, queue order and counter in the library's static device
scratch instead of allocated memory) against lila_dgels_batched, through the
C ABI: bit-identical outputs on two calls in a row, the second on a smaller
batch (the scratch keeps the first call's table and counter, which the
second must overwrite and reset), and the size limit rejected on the host."""
import numpy as np
import pytest
import torch

from conftest import say
from test_gpu_batched import Batch, make_sizes

pytestmark = pytest.mark.gpu


def _run_noalloc(lila, b):
    arrs = b.arrays()
    r = lila.call("lila_dgels_batched_noalloc", *arrs)
    assert r == 0, r
    torch.cuda.synchronize()
    return b.buf.clone()


def _run_reference(lila, b):
    b.buf.copy_(b.orig)
    torch.cuda.synchronize()
    arrs = b.arrays()
    assert lila.call("lila_dgels_batched", *arrs) == 0
    assert lila.call("lila_sync") == 0
    return b.buf.clone()


def test_noalloc_matches_the_allocating_entry(lila):
    big = Batch(make_sizes(91, 700, mmax=900), 92)
    small = Batch(make_sizes(93, 37, mmax=300), 94)
    say("noalloc: first call,", len(big.sizes), "problems")
    got1 = _run_noalloc(lila, big)
    say("noalloc: second call,", len(small.sizes), "problems")
    got2 = _run_noalloc(lila, small)
    say("noalloc: the allocating entry on the same inputs")
    ref1 = _run_reference(lila, big)
    ref2 = _run_reference(lila, small)
    assert torch.equal(got1, ref1)
    assert torch.equal(got2, ref2)
    # and the outputs are solutions: the normal equations hold
    for b, got in ((big, got1), (small, got2)):
        out = got.cpu().numpy()
        for p in range(0, len(b.sizes), 7):
            m, n, k, lda, ldb = b.sizes[p]
            if k == 0:
                continue
            A0, B0 = b.views(b.host0, p)
            _, Bg = b.views(out, p)
            x = Bg[:n, :k]
            r = A0[:m] @ x - B0[:m, :k]
            nA = np.linalg.norm(A0[:m])
            crit = np.linalg.norm(A0[:m].T @ r) / (nA * (nA * np.linalg.norm(x) + np.linalg.norm(B0[:m, :k])))
            assert crit < 1e-14, (p, crit)


def test_noalloc_rejects_beyond_its_maximum(lila):
    """16385 problems exceed the static table: -1, nothing launched"""
    nprob = 16385
    buf = torch.zeros(64 * 8, dtype=torch.float64, device="cuda")
    base = buf.data_ptr()
    m = np.full(nprob, 64, np.int32)
    n = np.full(nprob, 8, np.int32)
    k = np.zeros(nprob, np.int32)
    pa = np.full(nprob, base, np.uint64)
    assert lila.call("lila_dgels_batched_noalloc", nprob, m, n, k, pa, m, pa, m) == -1
    assert torch.count_nonzero(buf).item() == 0
