"""The operation counts of qr3/flops (lila_amd/csrc/flops.c) through the C
ABI (host arithmetic, no GPU): every closed form equals the step-by-step sum
of its `_check` form over ranges of sizes -- the reference's own validation
of its formulas -- flops_geqr3's closed form on powers of two (the sizes it
was derived for), plus small cases counted by hand."""
import ctypes

import pytest

import lila_amd

NAMES1 = ["flops_mLUinA", "flops_V2N", "flops_ULTinU", "flops_N2T"]
NAMES2 = ["flops_ApUBTinA", "flops_VT2Q", "flops_dlarft3", "flops_qr2_dorgqr", "flops_geqr3"]


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(lila_amd.LIB_PATH)
    for n in NAMES1 + NAMES2 + [x + "_check" for x in NAMES1 + NAMES2] + [
            "flops_gemm", "flops_trmm", "flops_syrk", "flops_larft", "flops_lapack_larfg",
            "flops_lapack_larf", "flops_lapack_geqr2", "flops_lapack_org2r",
            "flops_lapack_org2r_check"]:
        getattr(L, n).restype = ctypes.c_long
    L.flops_trmm.argtypes = [ctypes.c_char, ctypes.c_int, ctypes.c_int]
    return L


@pytest.mark.parametrize("name", NAMES1)
def test_closed_form_equals_recursion_n(lib, name):
    f, g = getattr(lib, name), getattr(lib, name + "_check")
    for n in list(range(1, 130)) + [255, 256, 1000, 1023]:
        assert f(n) == g(n), (name, n)


@pytest.mark.parametrize("name", ["flops_ApUBTinA", "flops_VT2Q", "flops_dlarft3", "flops_qr2_dorgqr"])
def test_closed_form_equals_recursion_mn(lib, name):
    f, g = getattr(lib, name), getattr(lib, name + "_check")
    for n in list(range(1, 70)) + [128, 255, 512]:
        for m in (n, n + 1, 2 * n + 3, 1000 + n):
            assert f(m, n) == g(m, n), (name, m, n)


def test_geqr3_closed_form_on_powers_of_two(lib):
    for k in range(0, 11):
        n = 2 ** k
        for m in (n, n + 7, 4 * n, 100000):
            assert lib.flops_geqr3(m, n) == lib.flops_geqr3_check(m, n), (m, n)
    # elsewhere the recursion is the count and the closed form stays close
    for n in (3, 5, 100, 1000):
        m = 4 * n
        c, r = lib.flops_geqr3(m, n), lib.flops_geqr3_check(m, n)
        assert c != r and abs(c - r) < 1e-2 * r


def test_lapack_org2r_closed_form_equals_loop(lib):
    for n in range(1, 40):
        for k in range(1, n + 1):
            for m in (n, n + 3, 200):
                assert lib.flops_lapack_org2r(m, n, k) == lib.flops_lapack_org2r_check(m, n, k)


def test_small_cases_by_hand(lib):
    assert lib.flops_gemm(3, 4, 5) == 120
    assert lib.flops_trmm(b"L", 3, 5) == 45 and lib.flops_trmm(b"R", 3, 5) == 75
    assert lib.flops_syrk(4, 10) == 160
    # one column: dlarfg of length m, 3m + 5
    assert lib.flops_geqr3(17, 1) == lib.flops_lapack_larfg(17) == 56
    # -L U of order 2: one product (2) and the two 1 x 1 triangular updates
    assert lib.flops_mLUinA(2) == 4
    # A += U B^T at one entry: a multiply and an add
    assert lib.flops_ApUBTinA_check(1, 1) == 2
    # a single reflector applied to an m x n block: 4mn - n
    assert lib.flops_lapack_larf(10, 3) == 117
    # dlarft of one column costs nothing; of two: the one product of V1^T v2
    assert lib.flops_larft(10, 1) == 0 and lib.flops_dlarft3(10, 1) == 0
    # Q of a single column: m (one scaling and the 1 - tau entry)
    assert lib.flops_qr2_dorgqr_check(9, 1) == 9
