"""The two measures the mixed-precision driver reads from its first Gram G
to decide whether a second Cholesky-QR pass is needed (internal symbols of
liblila_hip.so): ld_orth_defect = ||I - G||_F, and
ld_sym_defect_norm2_est, a power-method estimate of ||I - G||_2 that must
be a lower bound and, with the driver's 20 steps, within its factor-3
margin of the true value.  G is symmetric with only its upper triangle
valid; the test fills the lower triangle with garbage to check that it is
never read."""
import ctypes

import numpy as np
import pytest

from conftest import say, to_dev

pytestmark = pytest.mark.gpu


def make_gram(n, eig, seed):
    rng = np.random.default_rng(seed)
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    E = (V * eig) @ V.T  # I - G
    G = np.eye(n) - E
    Gs = np.triu(G) + np.tril(rng.standard_normal((n, n)) * 1e3, -1)
    return np.asfortranarray(Gs), E


@pytest.mark.parametrize("n,kind", [(500, "spread"), (777, "clustered"), (300, "one_large"),
                                    (1, "spread")])
def test_defect_measures(lila, n, kind):
    say("defect measures", n, kind)
    rng = np.random.default_rng(n)
    if kind == "spread":
        eig = rng.uniform(-0.3, 0.3, n)
    elif kind == "clustered":
        eig = 0.05 * (1.0 + 0.01 * rng.standard_normal(n)) * rng.choice([-1.0, 1.0], n)
    else:
        eig = 1e-4 * rng.standard_normal(n)
        eig[n // 2] = -0.9
    G, E = make_gram(n, eig, n + 1)
    lib = lila.load()
    lila.call("lila_get_info")  # library state set up, as by any entry point
    lib.ld_orth_defect.restype = ctypes.c_double
    lib.ld_sym_defect_norm2_est.restype = ctypes.c_double
    Gd = to_dev(G)
    fro = lib.ld_orth_defect(n, ctypes.c_void_p(Gd.data_ptr()), n)
    est = lib.ld_sym_defect_norm2_est(n, ctypes.c_void_p(Gd.data_ptr()), n, 20)
    true2 = np.max(np.abs(np.linalg.eigvalsh(E)))
    say("  fro %.6e (numpy %.6e)  est %.6e  ||E||_2 %.6e" % (fro, np.linalg.norm(E), est, true2))
    assert abs(fro - np.linalg.norm(E)) <= 1e-12 * max(1.0, np.linalg.norm(E))
    assert est <= true2 * (1 + 1e-10) + 1e-15
    assert 3.0 * est >= true2
    # the same call twice gives the same bits (fixed start vector and order)
    assert lib.ld_sym_defect_norm2_est(n, ctypes.c_void_p(Gd.data_ptr()), n, 20) == est
