"""GPU parity of the UT / ISW / V2T / VT2Q / VS2Q forms and the aux routines
of qr3/src against the oracle, through the C ABI of liblila_hip.so: seeded
inputs, ragged and square sizes (m == n puts +inf on S's last diagonal
entry, as in the reference), blocks the ISW forms must leave untouched, and
host pointers."""
import numpy as np
import pytest
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu

SIZES = [(40, 1), (300, 37), (1000, 96), (5000, 130)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def spine_mask(n):
    M = np.zeros((n, n), dtype=bool)
    off = 0
    while n - off > 1:
        b = (n - off) // 2
        M[off:off + b, off:off + b] = np.triu(np.ones((b, b), dtype=bool))
        off += b
    if n - off == 1:
        M[off, off] = True
    return M


def vt(ora, seed, m, n):
    """V (below the diagonal, R above), T and tau from the oracle's dgeqr3"""
    A = ora.fill_normal(seed, m, n)
    T = np.zeros((n, n), order="F")
    ora.call("dgeqr3", m, n, A, m, T, n)
    return A, T, np.ascontiguousarray(np.diag(T))


def sync_ok(lila):
    torch.cuda.synchronize()
    assert lila.call("lila_get_info") == 0


def cmp_with_inf(g, o, tol):
    """equal +inf positions, finite parts within tol"""
    assert np.array_equal(np.isposinf(g), np.isposinf(o))
    assert not np.any(np.isnan(g))
    f = np.isfinite(o)
    assert rel(g[f], o[f]) < tol


@pytest.mark.parametrize("m,n", SIZES + [(64, 64)])
def test_dgeqr3_R_UT(lila, ora, m, n):
    say("qr2_dgeqr3_R_UT", m, n)
    A0 = ora.fill_normal(m + 3 * n, m, n)
    Ao, So, Ro = A0.copy(order="F"), np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
    ora.call("qr2_dgeqr3_R_UT", m, n, Ao, m, So, n, Ro, n)
    Ad, Sd, Rd = to_dev(A0), to_dev(np.zeros((n, n))), to_dev(np.zeros((n, n)))
    assert lila.call("qr2_dgeqr3_R_UT", m, n, Ad, m, Sd, n, Rd, n) == 0
    sync_ok(lila)
    tol = 100 * n * U64
    if n > 1:
        assert rel(np.tril(to_host(Ad), -1), np.tril(Ao, -1)) < tol
    assert rel(np.triu(to_host(Rd)), np.triu(Ro)) < tol
    cmp_with_inf(np.triu(to_host(Sd)), np.triu(So), 10 * tol)


@pytest.mark.parametrize("ut", [0, 1])
@pytest.mark.parametrize("m,n", SIZES + [(64, 64)])
def test_dgeqr3_R_ISW(lila, ora, ut, m, n):
    name = "qr2_dgeqr3_R_UT_ISW" if ut else "qr2_dgeqr3_R_ISW"
    say(name, m, n)
    A0 = ora.fill_normal(2 * m + n, m, n)
    sentinel = 5.5
    T0 = np.full((n, n), sentinel, order="F")
    Ao, To, Ro = A0.copy(order="F"), T0.copy(order="F"), np.zeros((n, n), order="F")
    ora.call(name, m, n, Ao, m, To, n, Ro, n)
    Ad, Td, Rd = to_dev(A0), to_dev(T0), to_dev(np.zeros((n, n)))
    assert lila.call(name, m, n, Ad, m, Td, n, Rd, n) == 0
    sync_ok(lila)
    tol = 100 * n * U64
    Tg = to_host(Td)
    assert rel(np.triu(to_host(Rd)), np.triu(Ro)) < tol
    M = spine_mask(n)
    assert np.all(Tg[~M] == sentinel)
    cmp_with_inf(Tg[M], To[M], 10 * tol)


@pytest.mark.parametrize("name", ["qr2_dlarft3_UT", "qr2_dlarft3_ISW", "qr2_dlarft3_UT_ISW",
                                  "qr2_dlarft3_ISW_V2T", "qr2_dlarft3_ISW_V2T_UT"])
@pytest.mark.parametrize("m,n", SIZES + [(64, 64)])
def test_dlarft3_forms(lila, ora, name, m, n):
    say(name, m, n)
    V, _, tau = vt(ora, 5 * m + n, m, n)
    sentinel = -3.25
    To = np.full((n, n), sentinel, order="F")
    ora.call(name, m, n, V.copy(order="F"), m, To, n, tau.copy())
    Td = to_dev(np.full((n, n), sentinel))
    assert lila.call(name, m, n, to_dev(V), m, Td, n, torch.from_numpy(tau.copy()).cuda()) == 0
    sync_ok(lila)
    Tg = to_host(Td)
    untouched = ~np.triu(np.ones((n, n), dtype=bool)) if name == "qr2_dlarft3_UT" else ~spine_mask(n)
    assert np.all(Tg[untouched] == sentinel)
    cmp_with_inf(Tg[~untouched], To[~untouched], 1000 * n * U64)


@pytest.mark.parametrize("m,n", SIZES + [(64, 64)])
def test_dorgqr3_UT_VS2Q(lila, ora, m, n):
    say("qr2_dorgqr3_UT / VS2Q", m, n)
    V, _, tau = vt(ora, 7 * m + n, m, n)
    S = np.zeros((n, n), order="F")
    ora.call("qr2_dlarft3_UT", m, n, V.copy(order="F"), m, S, n, tau.copy())
    taud = torch.from_numpy(tau.copy()).cuda()
    for name in ("qr2_dorgqr3_UT", "qr2_dorgqr3_VS2Q"):
        Qo = V.copy(order="F")
        ora.call(name, m, n, Qo, m, S.copy(order="F"), n, tau.copy())
        Qd = to_dev(V)
        assert lila.call(name, m, n, Qd, m, to_dev(S), n, taud) == 0
        sync_ok(lila)
        Qg = to_host(Qd)
        assert rel(Qg, Qo) < 1000 * n * U64, name
        assert np.linalg.norm(Qg.T @ Qg - np.eye(n)) < 1000 * n * U64, name


@pytest.mark.parametrize("m,n", SIZES + [(64, 64)])
def test_aux_VT2Q_VS2Q(lila, ora, m, n):
    say("qr2_aux_dVT2Q / dVS2Q", m, n)
    V, T, tau = vt(ora, 9 * m + n, m, n)
    S = np.zeros((n, n), order="F")
    ora.call("qr2_dlarft3_UT", m, n, V.copy(order="F"), m, S, n, tau.copy())
    for name, top in (("qr2_aux_dVT2Q", T), ("qr2_aux_dVS2Q", S)):
        Q0 = V.copy(order="F")
        Q0[:n, :n] = np.tril(Q0[:n, :n], -1) + np.triu(top)
        Qo = Q0.copy(order="F")
        ora.call(name, m, n, Qo, m)
        Qd = to_dev(Q0)
        assert lila.call(name, m, n, Qd, m) == 0
        sync_ok(lila)
        assert rel(to_host(Qd), Qo) < 1000 * n * U64, name


@pytest.mark.parametrize("n", [1, 2, 17, 64, 101, 300])
def test_aux_mLUinA(lila, ora, n):
    say("qr2_aux_dmLUinA", n)
    A0 = ora.fill_normal(n + 1, n, n)
    Ao = A0.copy(order="F")
    ora.call("qr2_aux_dmLUinA", n, Ao, n)
    Ad = to_dev(A0)
    assert lila.call("qr2_aux_dmLUinA", n, Ad, n) == 0
    sync_ok(lila)
    assert rel(to_host(Ad), Ao) < 100 * n * U64


@pytest.mark.parametrize("alias", [True, False])
@pytest.mark.parametrize("n", [1, 5, 33, 90, 257])
def test_aux_ULTinU_UinvLTinU(lila, ora, n, alias):
    U0 = ora.fill_normal(3 * n, n, n) + 4 * np.eye(n)
    L0 = U0 if alias else ora.fill_normal(5 * n, n, n)
    for name in ("qr2_aux_dULTinU", "qr2_aux_dUinvLTinU"):
        say(name, n, "alias" if alias else "apart")
        Uo = U0.copy(order="F")
        Lo = Uo if alias else L0.copy(order="F")
        ora.call(name, n, Lo, n, Uo, n)
        Ud = to_dev(U0)
        Ld = Ud if alias else to_dev(L0)
        assert lila.call(name, n, Ld, n, Ud, n) == 0
        sync_ok(lila)
        Ug = to_host(Ud)
        assert rel(np.triu(Ug), np.triu(Uo)) < 100 * n * U64, name
        assert np.array_equal(np.tril(Ug, -1), np.tril(U0, -1)), name


@pytest.mark.parametrize("m,n", [(1, 1), (1, 5), (6, 1), (7, 4), (4, 9), (130, 70), (300, 513)])
def test_aux_ApUBTinA(lila, ora, m, n):
    say("qr2_aux_dApUBTinA", m, n)
    A0, U0, B0 = ora.fill_normal(m, m, n), ora.fill_normal(m + 1, m, m), ora.fill_normal(n + 2, n, m)
    Ao = A0.copy(order="F")
    ora.call("qr2_aux_dApUBTinA", m, n, Ao, m, U0.copy(order="F"), m, B0.copy(order="F"), n)
    Ad = to_dev(A0)
    assert lila.call("qr2_aux_dApUBTinA", m, n, Ad, m, to_dev(U0), m, to_dev(B0), n) == 0
    sync_ok(lila)
    assert rel(to_host(Ad), Ao) < 100 * max(m, n) * U64


@pytest.mark.parametrize("n", [1, 2, 9, 32, 33, 77, 130])
def test_aux_V2N_N2T(lila, ora, n):
    say("qr2_aux_dV2N / dN2T", n)
    m = 3 * n + 5
    V, T, tau = vt(ora, 11 * n, m, n)
    T0 = np.asfortranarray(np.tril(ora.fill_normal(n, n, n), -1) + np.triu(V[:n, :n].T, 1))
    To = T0.copy(order="F")
    ora.call("qr2_aux_dV2N", n, To, n)
    Td = to_dev(T0)
    assert lila.call("qr2_aux_dV2N", n, Td, n) == 0
    sync_ok(lila)
    Tg = to_host(Td)
    assert rel(np.triu(Tg), np.triu(To)) < 100 * n * U64
    assert np.array_equal(np.tril(Tg, -1), np.tril(T0, -1))
    # N = V^T V (strict upper used) -> T
    Vu = np.tril(V, -1)
    Vu[:n, :n] += np.eye(n)
    N0 = np.asfortranarray(np.tril(T0, -1) + np.triu(Vu.T @ Vu, 1))
    No = N0.copy(order="F")
    ora.call("qr2_aux_dN2T", n, tau.copy(), No, n)
    Nd = to_dev(N0)
    assert lila.call("qr2_aux_dN2T", n, torch.from_numpy(tau.copy()).cuda(), Nd, n) == 0
    sync_ok(lila)
    Ng = to_host(Nd)
    assert rel(np.triu(Ng), np.triu(No)) < 1000 * n * U64
    assert rel(np.triu(Ng), np.triu(T)) < 1000 * n * U64
    assert np.array_equal(np.tril(Ng, -1), np.tril(N0, -1))


@pytest.mark.parametrize("m,n", [(50, 7), (300, 37), (64, 64), (10, 1), (1, 1)])
def test_aux_V2tau_host_and_device(lila, ora, m, n):
    say("qr2_aux_dV2tau", m, n)
    V, _, _ = vt(ora, 13 * m + n, m, n)
    to = np.zeros(n)
    ora.call("qr2_aux_dV2tau", m, n, V.copy(order="F"), m, to)
    td = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert lila.call("qr2_aux_dV2tau", m, n, to_dev(V), m, td) == 0
    sync_ok(lila)
    assert rel(td.cpu().numpy(), to) < 10 * m * U64
    th = np.zeros(n)
    assert lila.call("qr2_aux_dV2tau", m, n, V.copy(order="F"), m, th) == 0
    assert rel(th, to) < 10 * m * U64
    if m == n:
        assert th[-1] == 0.0


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("incv", [1, 3])
@pytest.mark.parametrize("m,n", [(1, 1), (17, 5), (400, 600)])
def test_aux_dlarf(lila, ora, side, incv, m, n):
    say("qr2_aux_dlarf_wrapper", side, incv, m, n)
    L = m if side == "L" else n
    v = ora.fill_normal(m + n + incv, (L - 1) * incv + 1, 1)[:, 0].copy()
    C0 = ora.fill_normal(m * n, m, n)
    Co = C0.copy(order="F")
    ora.call("qr2_aux_dlarf_wrapper", ord(side), m, n, v.copy(), incv, 0.37, Co, m, None)
    Cd = to_dev(C0)
    vd = torch.from_numpy(v.copy()).cuda()
    assert lila.call("qr2_aux_dlarf_wrapper", ord(side), m, n, vd, incv, 0.37, Cd, m, None) == 0
    sync_ok(lila)
    assert rel(to_host(Cd), Co) < 100 * max(m, n) * U64
    Ch = C0.copy(order="F")  # host pointers, staged
    assert lila.call("qr2_aux_dlarf_wrapper", ord(side), m, n, v.copy(), incv, 0.37, Ch, m, None) == 0
    assert rel(Ch, Co) < 100 * max(m, n) * U64
    assert lila.call("qr2_aux_dlarf_wrapper", ord(side), m, n, vd, 0, 0.37, Cd, m, None) == -1


def test_ISW_aliased_R_host(lila, ora):
    """R == A, host pointers: V, R and the spine blocks as the oracle's (the
    reference leaves the update's scratch between the spine blocks; here they
    keep their input values)"""
    m, n = 700, 50
    say("qr2_dgeqr3_R_ISW aliased, host")
    A0 = ora.fill_normal(17, m, n)
    Ao, To = A0.copy(order="F"), np.zeros((n, n), order="F")
    ora.call("qr2_dgeqr3_R_ISW", m, n, Ao, m, To, n, Ao, m)
    A, T = A0.copy(order="F"), np.zeros((n, n), order="F")
    assert lila.call("qr2_dgeqr3_R_ISW", m, n, A, m, T, n, A, m) == 0
    assert rel(A, Ao) < 100 * n * U64
    M = spine_mask(n)
    assert rel(T[M], To[M]) < 1000 * n * U64
    assert np.all(T[~M] == 0.0)
