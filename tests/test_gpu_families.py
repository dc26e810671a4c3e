"""GPU parity for the VRT-only (v03), Q-from-(V,T) (q03), append-columns,
LAPACK-kernel (ker_recursive, dormqrf_w00/z00/z02) and qr2 entry points,
through the C ABI on device memory, against the oracle (whose pinning is in
test_oracle_families.py).  Tolerances are multiples of n * u."""
import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host
from test_gpu_parity import rel, sync_ok

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def oracle_on_host_blas(ora):
    ora.use_host_blas()
    yield
    ora.use_builtin_blas()


CASES = [
    # m, n, ii, mt, nx, lda, ldq
    (87, 53, 6, 4, 7, None, None),
    (1200, 600, 0, 64, 32, None, None),
    (913, 412, 37, 10, 9, 1002, 949),
    (700, 350, 13, 200, 33, None, None),
]


@pytest.mark.parametrize("leaf,panel", [(1, 1), (0, 1), (1, 0)])
@pytest.mark.parametrize("case", CASES)
def test_v03_then_q03(lila, ora, case, leaf, panel):
    m, n, ii, mt, nx, lda, ldq = case
    lda, ldq = lda or m, ldq or m
    say("v03/q03", case, leaf, panel)
    A = ora.fill_uniform(1, lda, n + ii, -0.5, 0.5)
    Q = ora.fill_uniform(2, ldq, n + ii, -0.5, 0.5)
    T = np.zeros((mt, n + ii), order="F")
    p = np.array([0, leaf, panel, nx], dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_v03_recursive", p, m, n, ii, mt, Ao, lda, To, mt, None, 0)
    ora.call("lila_dgeqrf_q03_recursive", p, m, n, ii, mt, Ao, lda, To, mt, Qo, ldq, None, 0)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_v03_recursive", p, m, n, ii, mt, Ad, lda, Td, mt, None, 0) == 0
    Qbefore = to_host(Qd)
    assert np.array_equal(Qbefore, Q)  # VRT only: Q untouched
    assert lila.call("lila_dgeqrf_q03_recursive", p, m, n, ii, mt, Ad, lda, Td, mt, Qd, ldq, None, 0) == 0
    sync_ok(lila)
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol
    assert rel(to_host(Td), To) < 10 * tol
    assert rel(to_host(Qd), Qo) < tol
    assert ora.call("test_qq_orth_full", m, n, ii, to_host(Qd), ldq) < tol


@pytest.mark.parametrize("nb_lvl", [[64], [128, 16]])
def test_v03_q03_levelx(lila, ora, nb_lvl):
    m, n, ii, mt = 900, 400, 5, 64
    say("v03/q03 levelx", nb_lvl)
    A = ora.fill_normal(3, m, n + ii)
    Q = np.zeros((m, n + ii), order="F")
    T = np.zeros((mt, n + ii), order="F")
    p = np.array([0, 1, 1, len(nb_lvl)], dtype=np.int32)
    nbl = np.array(nb_lvl, dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_v03_levelx", p, len(nb_lvl), 0, nbl, m, n, ii, mt, Ao, m, To, mt, None, 0)
    ora.call("lila_dgeqrf_q03_levelx", p, len(nb_lvl), 0, nbl, m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_v03_levelx", p, len(nb_lvl), 0, nbl, m, n, ii, mt, Ad, m, Td, mt,
                     None, 0) == 0
    assert lila.call("lila_dgeqrf_q03_levelx", p, len(nb_lvl), 0, nbl, m, n, ii, mt, Ad, m, Td, mt,
                     Qd, m, None, 0) == 0
    sync_ok(lila)
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Qd), Qo) < tol and rel(to_host(Td), To) < 10 * tol


@pytest.mark.parametrize("m,n,nb,mt,nx", [(87, 53, 10, 4, 1), (1000, 400, 96, 64, 32)])
def test_w03_appendcols(lila, ora, m, n, nb, mt, nx):
    say("appendcols", m, n, nb, mt, nx)
    A = ora.fill_uniform(4, m, n, -0.5, 0.5)
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    Ao, Qo, To = A.copy(order="F"), np.zeros((m, n), order="F"), np.zeros((mt, n), order="F")
    Ad, Qd, Td = to_dev(A), to_dev(np.zeros((m, n), order="F")), to_dev(np.zeros((mt, n), order="F"))
    j = 0
    while j < n:
        vb = min(nb, n - j)
        ora.call("lila_dgeqrf_w03_appendcols", p, m, j, vb, mt, Ao, m, Qo, m, To, mt, None, 0)
        assert lila.call("lila_dgeqrf_w03_appendcols", p, m, j, vb, mt, Ad, m, Qd, m, Td, mt, None, 0) == 0
        j += vb
    sync_ok(lila)
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Qd), Qo) < tol and rel(to_host(Td), To) < 10 * tol


@pytest.mark.parametrize("m,n,nb", [(87, 53, 10), (1500, 300, 128)])
def test_lapack_appendcols(lila, ora, m, n, nb):
    say("LAPACK appendcols", m, n, nb)
    A = ora.fill_uniform(5, m, n, -0.5, 0.5)
    Ao, Qo, to = A.copy(order="F"), np.zeros((m, n), order="F"), np.zeros(n)
    Ad, Qd = to_dev(A), to_dev(np.zeros((m, n), order="F"))
    td = torch.zeros(n, dtype=torch.float64, device="cuda")
    j = 0
    while j < n:
        vb = min(nb, n - j)
        ora.call("lila_dgeqrf_LAPACK_appendcols", m, j, vb, 8, Ao, m, Qo, m, to, 1, None, 0)
        assert lila.call("lila_dgeqrf_LAPACK_appendcols", m, j, vb, 8, Ad, m, Qd, m, td, 1, None, 0) == 0
        j += vb
    sync_ok(lila)
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Qd), Qo) < tol
    assert rel(td.cpu().numpy(), to) < tol


@pytest.mark.parametrize("m,n,nx", [(87, 53, 7), (1200, 700, 64)])
def test_ker_recursive(lila, ora, m, n, nx):
    say("ker_recursive", m, n, nx)
    A = ora.fill_uniform(6, m, n, -0.5, 0.5)
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    Ao, Qo, to = A.copy(order="F"), np.zeros((m, n), order="F"), np.zeros(n)
    ora.call("lila_dgeqrf_ker_recursive", p, m, n, 0, 1, Ao, m, to, 1, Qo, m, None, 0)
    Ad, Qd = to_dev(A), to_dev(np.zeros((m, n), order="F"))
    td = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert lila.call("lila_dgeqrf_ker_recursive", p, m, n, 0, 1, Ad, m, td, 1, Qd, m, None, 0) == 0
    sync_ok(lila)
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Qd), Qo) < tol
    assert rel(td.cpu().numpy(), to) < tol


@pytest.mark.parametrize("m,k,nr,ii", [(300, 40, 9, 0), (2000, 130, 20, 7)])
def test_dormqrf_w00_z00_z02(lila, ora, m, k, nr, ii):
    say("dormqrf w00/z00/z02", m, k, nr, ii)
    A = ora.fill_normal(7, m, k + ii + nr)
    T = np.zeros((k + ii, k + ii), order="F")
    # a factored block with its full-layout T, via the oracle's dgeqr3
    Af = A.copy(order="F")
    Tb = np.zeros((k, k), order="F")
    blk = np.asfortranarray(Af[ii:, ii:ii + k])
    ora.call("dgeqr3", m - ii, k, blk, m - ii, Tb, k)
    Af[ii:, ii:ii + k] = blk
    T[ii:, ii:] = Tb
    B = ora.fill_normal(8, m, nr + ii)
    tol = 100 * k * U64
    # z00
    Bo = B.copy(order="F")
    ora.call("lila_dormqrf_z00", m, nr, k, ii, ii, 8, Af, m, Bo, m, None, 1, None, 0)
    Bd = to_dev(B)
    assert lila.call("lila_dormqrf_z00", m, nr, k, ii, ii, 8, to_dev(Af), m, Bd, m, None, 1, None, 0) == 0
    sync_ok(lila)
    assert rel(to_host(Bd), Bo) < tol
    # w00
    Aw = Af.copy(order="F")
    Aw[:, ii + k:ii + k + nr] = B[:, ii:ii + nr]
    Awo = Aw.copy(order="F")
    ora.call("lila_dormqrf_w00", m, nr, k, ii, ii + k, 8, Awo, m, None, 1, None, 0)
    Awd = to_dev(Aw)
    assert lila.call("lila_dormqrf_w00", m, nr, k, ii, ii + k, 8, Awd, m, None, 1, None, 0) == 0
    sync_ok(lila)
    assert rel(to_host(Awd), Awo) < tol
    # z02
    B2 = B.copy(order="F")
    ora.call("lila_dormqrf_z02", m, nr, k, ii, ii, 8, Af, m, T, k + ii, B2, m, None, 0)
    B2d = to_dev(B)
    assert lila.call("lila_dormqrf_z02", m, nr, k, ii, ii, 8, to_dev(Af), m, to_dev(T), k + ii, B2d, m,
                     None, 0) == 0
    sync_ok(lila)
    assert rel(to_host(B2d), B2) < tol


@pytest.mark.parametrize("method", [0, 1, 2])
def test_qr2_recursive(lila, ora, method):
    m, n, ii, mt, nx = 800, 256, 0, 128, 32
    say("qr2 method", method)
    A = ora.fill_uniform(9, m, n, -0.5, 0.5)
    Q = np.zeros((m, n), order="F")
    T = np.zeros((mt, n), order="F")
    if method == 2:
        # Q from the (V, T) of a VRT pass
        p1 = np.array([0, 1, 1, nx, 1], dtype=np.int32)
        ora.call("lila_dgeqrf_qr2_recursive", p1, m, n, ii, mt, A, m, T, mt, Q, m, None, 0)
    p = np.array([0, 1, 1, nx, method], dtype=np.int32)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dgeqrf_qr2_recursive", p, m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_qr2_recursive", p, m, n, ii, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    sync_ok(lila)
    tol = 100 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Td), To) < 10 * tol
    if method != 1:
        assert rel(to_host(Qd), Qo) < tol
    else:
        assert not np.any(to_host(Qd))
