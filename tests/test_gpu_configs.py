"""GPU parity at the full sizes of the blocked (levelx) workload: square
m = n = 5000 and tall m = 100000, n = 512, N(0,1) entries, one level of
blocking with nb in {32, 64, 128}, T factors in the mt layout, and Q^T
applied to a 5000 x 1000 N(0,1) block with lila_dormqrf_z03.  Each GPU
result is compared with the oracle's (its BLAS-3 steps on the host OpenBLAS)
on the same seeded inputs, to c n u relative, and the oracle's checkers
bound ||A - QR||/||A|| and ||I - Q^T Q|| of the GPU result by c n u."""
import numpy as np
import pytest

from conftest import U64, say, to_dev, to_host
from test_gpu_parity import check_vrtq, rel, run_driver, sync_ok

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def oracle_on_host_blas(ora):
    ora.use_host_blas()
    yield
    ora.use_builtin_blas()


@pytest.mark.parametrize("nb", [32, 64, 128])
@pytest.mark.parametrize("m,n", [(5000, 5000), (100000, 512)])
def test_levelx_full_size(lila, ora, m, n, nb):
    mt = 512
    say("levelx full size", m, n, "nb", nb)
    A, o, g = run_driver(lila, ora, "levelx", m, n, 0, mt, [0, 1, 1, 1], nb_lvl=[nb], normal=True,
                         seed=nb)
    check_vrtq(ora, A, o, g, m, n, 0, mt, m, m)


def test_dormqrf_z03_5000_by_1000(lila, ora):
    """Q^T B for the 5000 x 5000 factor (nb = 64) and B 5000 x 1000"""
    m = n = 5000
    mt, nrhs = 512, 1000
    say("dormqrf_z03", m, n, nrhs)
    A, o, _ = run_driver(lila, ora, "levelx", m, n, 0, mt, [0, 1, 1, 1], nb_lvl=[64], normal=True,
                         seed=5)
    Ao, To, _ = o
    B = ora.fill_normal(77, m, nrhs)
    Bo = B.copy(order="F")
    ora.call("lila_dormqrf_z03", m, nrhs, n, 0, 0, mt, Ao, m, To, mt, Bo, m, None, 0)
    Ad, Td, Bd = to_dev(Ao), to_dev(To), to_dev(B)
    assert lila.call("lila_dormqrf_z03", m, nrhs, n, 0, 0, mt, Ad, m, Td, mt, Bd, m, None, 0) == 0
    sync_ok(lila)
    Bg = to_host(Bd)
    assert rel(Bg, Bo) < 100 * n * U64
    # Q^T preserves column norms
    assert np.allclose(np.linalg.norm(Bg, axis=0), np.linalg.norm(B, axis=0), rtol=100 * n * U64)
