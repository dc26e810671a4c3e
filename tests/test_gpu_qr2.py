"""GPU parity of the recursive QR library of qr3/src (qr2_dgeqr3_R,
qr2_dlarft3, qr2_dorgqr3, qr2_dorgqr3_VT2Q) against the oracle, through the
C ABI of liblila_hip.so."""
import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES = [(40, 1), (300, 37), (64, 64), (1000, 96), (5000, 130)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("m,n", SIZES)
def test_qr2_dgeqr3_R(lila, ora, m, n):
    say("qr2_dgeqr3_R", m, n)
    A0 = ora.fill_normal(m + n, m, n)
    Ao, To, Ro = A0.copy(order="F"), np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
    ora.call("qr2_dgeqr3_R", m, n, Ao, m, To, n, Ro, n)
    Ad, Td, Rd = to_dev(A0), to_dev(np.zeros((n, n), order="F")), to_dev(np.zeros((n, n), order="F"))
    assert lila.call("qr2_dgeqr3_R", m, n, Ad, m, Td, n, Rd, n) == 0
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0
    Ag, Tg, Rg = to_host(Ad), to_host(Td), to_host(Rd)
    tol = 100 * n * U64
    assert rel(np.tril(Ag, -1), np.tril(Ao, -1)) < tol if n > 1 else True
    assert rel(np.triu(Rg), np.triu(Ro)) < tol
    assert rel(np.triu(Tg), np.triu(To)) < 10 * tol
    # A's diagonal and upper part keep the input values here
    assert np.array_equal(np.triu(Ag[:n]), np.triu(A0[:n]))


def test_qr2_dgeqr3_R_aliased_host(lila, ora):
    m, n = 700, 50
    say("qr2_dgeqr3_R aliased, host pointers")
    A0 = ora.fill_normal(9, m, n)
    Ao, To = A0.copy(order="F"), np.zeros((n, n), order="F")
    ora.call("dgeqr3", m, n, Ao, m, To, n)
    A, T = A0.copy(order="F"), np.zeros((n, n), order="F")
    assert lila.call("qr2_dgeqr3_R", m, n, A, m, T, n, A, m) == 0
    assert rel(A, Ao) < 100 * n * U64
    assert rel(np.triu(T), np.triu(To)) < 1000 * n * U64


@pytest.mark.parametrize("m,n", SIZES)
def test_qr2_dlarft3_and_dorgqr3(lila, ora, m, n):
    say("qr2_dlarft3 / dorgqr3", m, n)
    A0 = ora.fill_normal(3 * m + n, m, n)
    V, To = A0.copy(order="F"), np.zeros((n, n), order="F")
    ora.call("dgeqr3", m, n, V, m, To, n)
    tau = np.ascontiguousarray(np.diag(To))
    # T from (V, tau)
    Td = to_dev(np.zeros((n, n), order="F"))
    taud = torch.from_numpy(tau.copy()).cuda()
    assert lila.call("qr2_dlarft3", m, n, to_dev(V), m, Td, n, taud) == 0
    torch.cuda.synchronize()
    T1 = np.zeros((n, n), order="F")
    ora.call("qr2_dlarft3", m, n, V.copy(order="F"), m, T1, n, tau.copy())
    assert rel(np.triu(to_host(Td)), np.triu(T1)) < 100 * n * U64
    # Q from (V, T), both names
    Qo = V.copy(order="F")
    ora.call("qr2_dorgqr3", m, n, Qo, m, To.copy(order="F"), n, tau.copy())
    for name in ("qr2_dorgqr3", "qr2_dorgqr3_VT2Q"):
        Qd = to_dev(V)
        assert lila.call(name, m, n, Qd, m, to_dev(To), n, taud) == 0
        torch.cuda.synchronize()
        Qg = to_host(Qd)
        assert rel(Qg, Qo) < 100 * n * U64, name
        assert np.linalg.norm(Qg.T @ Qg - np.eye(n)) < 100 * n * U64
