"""GPU: lila_dgemm_emu, the FP64 product emulated on int8 MFMA by the
Chinese-remainder (Ozaki II) scheme.

* integer operands whose product is exact in FP64 (entries of AB ~2^50):
  the emulated result must be bit-identical to the exact product (any wrong
  CRT constant, residue or modulus count shows up here);
* random operands with rows and columns graded over many orders of
  magnitude, against an 80-bit extended-precision product: within the
  scheme's error bound 2^-kb (rowmax_i ||b_j||_1 + colmax_j ||a_i||_1) plus
  the final roundings; for kb = 20 the error must also exceed what FP64
  rounding alone gives (the truncation really happens);
* edge shapes: ragged sizes, K = 65536 (the int32 limit), zero rows and
  columns, beta = 0 with NaN in C, 1e+-150 scales, rejected arguments.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import say, to_dev, to_host

pytestmark = pytest.mark.gpu
U64 = np.finfo(np.float64).eps / 2


def emu(lila, ta, m, n, k, alpha, A, B, beta, C, kb=53, tb="N"):
    Cd = to_dev(C)
    r = lila.call("lila_dgemm_emu", ctypes.c_char(ta.encode()), ctypes.c_char(tb.encode()), m, n, k, alpha,
                  to_dev(A), A.shape[0], to_dev(B), B.shape[0], beta, Cd, m, kb)
    torch.cuda.synchronize()
    return r, to_host(Cd)


def bound(opA, B, C, alpha, beta, kb):
    rowmax = np.max(np.abs(opA), axis=1, initial=0.0)[:, None]
    colmax = np.max(np.abs(B), axis=0, initial=0.0)[None, :]
    trunc = 2.0 ** -kb * (rowmax * np.sum(np.abs(B), axis=0)[None, :] +
                          colmax * np.sum(np.abs(opA), axis=1)[:, None])
    mag = np.abs(alpha) * (np.abs(opA) @ np.abs(B)) + np.abs(beta) * np.abs(C)
    return 2.0 * np.abs(alpha) * trunc + 4 * U64 * mag + 1e-300


@pytest.mark.parametrize("ta,m,n,k", [("N", 300, 200, 500), ("T", 257, 129, 1000), ("N", 128, 128, 128)])
def test_emu_integer_products_are_exact(lila, ta, m, n, k):
    # entries in [2^19, 2^20), B's columns of one sign each: no cancellation,
    # so every entry of AB is ~2^50 and the scheme's absolute error (~2^42 of
    # the 2^(s+t)-scaled integer X, below half an ulp of such entries) must
    # round away
    rng = np.random.default_rng(m + n + k)
    A = np.asfortranarray(rng.integers(2 ** 19, 2 ** 20, size=(k, m) if ta == "T" else (m, k)).astype(np.float64))
    sg = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)[None, :]
    B = np.asfortranarray(sg * rng.integers(2 ** 19, 2 ** 20, size=(k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-2 ** 20, 2 ** 20, size=(m, n)).astype(np.float64))
    opA = A.T if ta == "T" else A
    exact = opA.astype(np.int64) @ B.astype(np.int64) - 3 * C.astype(np.int64)  # |.| < 2^51
    say("emu exact integers", ta, m, n, k)
    r, got = emu(lila, ta, m, n, k, 1.0, A, B, -3.0, C)
    assert r == 0
    assert np.array_equal(got, exact.astype(np.float64))


@pytest.mark.parametrize("kb", [53, 40, 20])
@pytest.mark.parametrize("ta,m,n,k", [("N", 1000, 700, 1024), ("T", 300, 900, 5000), ("N", 129, 1, 7),
                                      ("T", 5, 300, 65), ("N", 77, 333, 4097)])
def test_emu_graded_within_bound(lila, ta, m, n, k, kb):
    rng = np.random.default_rng(7 * m + n + k + kb)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((k, n))
    C = rng.standard_normal((m, n))
    if ta == "T":
        A *= np.exp(rng.uniform(-20, 20, size=(1, m)))
    else:
        A *= np.exp(rng.uniform(-20, 20, size=(m, 1)))
    A *= np.exp(rng.uniform(-3, 3, size=A.shape))  # entries within a row vary too
    B *= np.exp(rng.uniform(-20, 20, size=(1, n)))
    A, B, C = np.asfortranarray(A), np.asfortranarray(B), np.asfortranarray(C)
    opA = A.T if ta == "T" else A
    ref = (0.5 * (opA.astype(np.longdouble) @ B.astype(np.longdouble)) +
           1.5 * C.astype(np.longdouble)).astype(np.float64)
    say("emu graded", ta, m, n, k, "kb", kb)
    r, got = emu(lila, ta, m, n, k, 0.5, A, B, 1.5, C, kb=kb)
    assert r == 0
    bnd = bound(opA, B, C, 0.5, 1.5, kb)
    assert np.all(np.abs(got - ref) <= bnd)
    if kb == 20:
        # truncation to 20-bit operands is visible: far above FP64 rounding
        mag = 0.5 * (np.abs(opA) @ np.abs(B)) + 1.5 * np.abs(C)
        assert np.max(np.abs(got - ref) / mag) > 1e3 * U64


def test_emu_edges(lila):
    rng = np.random.default_rng(3)
    # K at the int32 limit of the residue products
    m, n, k = 130, 70, 65536
    A = np.asfortranarray(rng.standard_normal((m, k)))
    B = np.asfortranarray(rng.standard_normal((k, n)))
    C = np.zeros((m, n), order="F")
    say("emu K = 65536")
    r, got = emu(lila, "N", m, n, k, 1.0, A, B, 0.0, C)
    assert r == 0
    ref = (A.astype(np.longdouble) @ B.astype(np.longdouble)).astype(np.float64)
    assert np.all(np.abs(got - ref) <= bound(A, B, C, 1.0, 0.0, 53))
    # zero rows / columns, beta = 0 with NaN in C (C not read), 1e+-150 scales
    m, n, k = 200, 150, 300
    A = rng.standard_normal((m, k))
    A[17] = 0.0
    A[:100] *= 1e150
    A[100:] *= 1e-150
    B = rng.standard_normal((k, n))
    B[:, 5] = 0.0
    A, B = np.asfortranarray(A), np.asfortranarray(B)
    C = np.full((m, n), np.nan, order="F")
    say("emu zero rows, NaN C, extreme scales")
    r, got = emu(lila, "N", m, n, k, 2.0, A, B, 0.0, C)
    assert r == 0
    assert np.all(np.isfinite(got))
    assert np.all(got[17] == 0.0) and np.all(got[:, 5] == 0.0)
    ref = (2.0 * (A.astype(np.longdouble) @ B.astype(np.longdouble))).astype(np.float64)
    assert np.all(np.abs(got - ref) <= bound(A, B, np.zeros_like(C), 2.0, 0.0, 53))
    # rejected: transposed B, K beyond the int32 limit, kb out of range
    C = np.zeros((4, 4), order="F")
    assert emu(lila, "N", 4, 4, 4, 1.0, np.ones((4, 4), order="F"), np.ones((4, 4), order="F"), 0.0, C,
               tb="T")[0] == -1
    assert emu(lila, "N", 4, 4, 4, 1.0, np.ones((4, 4), order="F"), np.ones((4, 4), order="F"), 0.0, C,
               kb=60)[0] == -1
    assert lila.call("lila_dgemm_emu", ctypes.c_char(b"N"), ctypes.c_char(b"N"), 4, 4, 65537, 1.0, None, 4,
                     None, 65537, 0.0, None, 4, 53) == -1
