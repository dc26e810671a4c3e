"""GPU parity of the leaf, panel and block-reflector entry points called
directly through the C ABI (not only from inside the recursive drivers), the
way the reference's own callers compose them.

* the leaf dispatchers lila_dgeqrf_w03_mt{,_hh,_l} and the panel dispatchers
  lila_dgeqr2_w03{,_3,_l,_hr} (lila-c/lila_dgeqrf_w03_mt*.c,
  lila-c/lila_dgeqr2_w03*.c), on ragged global offsets (i % mt != 0, mt > n),
  host and device pointers, against the oracle's restatement;
* the append-columns sequence of lila-c/lila_dgeqrf_w03_appendcols.c:7-10:
  lila_dormqrf_w03, the recursive driver, lila_dlarft_connect_w03 and
  lila_dormqrbz_w03 called one by one through the ABI;
* lila_ormhr2_w03_hr (lila-c/lila_ormhr2_w03_hr.c:3-84) block by block, so
  that steps with earlier blocks (j > i) solve with their unit lower V;
* R entrywise against LAPACK's R (up to the sign of each row) on the
  reference configurations.

Tolerances are multiples of n u (u = 2^-53), written in each test.
"""
import numpy as np
import pytest
import scipy.linalg
import torch

from conftest import U64, say, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def oracle_on_host_blas(ora):
    ora.use_host_blas()
    yield
    ora.use_builtin_blas()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def sync_ok(lila):
    torch.cuda.synchronize()
    info = lila.call("lila_get_info")
    assert info == 0, "device status %d" % info


def r_entrywise(A, Ag, m, n, ii, lda):
    """max |R - R_lapack| / ||A||_F over the n x n R of A(ii:m, ii:ii+n),
    each row of R_lapack taking the sign of the GPU's diagonal"""
    X = np.asarray(A[ii:m, ii:ii + n])
    Rl = scipy.linalg.qr(X, mode="r")[0][:n, :n]
    Rg = np.triu(np.asarray(Ag[ii:ii + n, ii:ii + n]))
    sg = np.sign(np.diag(Rg)) * np.sign(np.diag(Rl))
    sg[sg == 0] = 1.0
    return np.max(np.abs(Rg - sg[:, None] * Rl)) / np.linalg.norm(X)


def _inputs(ora, m, n, ii, mt, seed, normal=True):
    fill = ora.fill_normal if normal else (lambda s, r, c: ora.fill_uniform(s, r, c, -1.0, 1.0))
    A = fill(seed, m, n + ii)
    Q = fill(seed + 7, m, n + ii)
    T = fill(seed + 9, mt, n + ii)
    return A, T, Q


def _run(lila, name, head, A, T, Q, m, n, ii, mt, where):
    """call `name` through the ABI on device copies (where='dev') or on
    host arrays (where='host'); returns the outputs as host arrays"""
    if where == "dev":
        Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
        r = lila.call(name, *head, m, n, ii, mt, Ad, m, Td, mt, Qd, m, None, 0)
        assert r == 0, (name, r)
        sync_ok(lila)
        return to_host(Ad), to_host(Td), to_host(Qd)
    Ah, Th, Qh = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    r = lila.call(name, *head, m, n, ii, mt, Ah, m, Th, mt, Qh, m, None, 0)
    assert r == 0, (name, r)
    assert lila.call("lila_get_info") == 0
    return Ah, Th, Qh


def _check(ora, A, o, g, m, n, ii, tolf):
    Ao, To, Qo = o
    Ag, Tg, Qg = g
    tol = tolf * max(n, 8) * U64
    assert rel(Ag, Ao) < tol, ("A", rel(Ag, Ao))
    assert rel(Qg, Qo) < tol, ("Q", rel(Qg, Qo))
    assert rel(Tg, To) < 10 * tol, ("T", rel(Tg, To))
    assert ora.call("test_qq_orth_full", m, n, ii, Qg, m) < tol
    assert ora.call("test_qr_repres_1", m, n, ii, A, m, Qg, m, Ag, m) < tol
    assert r_entrywise(A, Ag, m, n, ii, m) < tol


# (m, n, ii, mt): the leaves get i % mt != 0 with several mt blocks, mt > n
# and ii = 0; a panel entry takes one panel inside one T block (its callers,
# the mt leaves, split at block edges: lila-c/lila_dgeqrf_w03_mt_hh.c:8-18),
# so its shapes keep ii % mt + n <= mt
LEAF_SHAPES = [(700, 45, 13, 16), (500, 30, 0, 64), (1300, 64, 37, 50)]
PANEL_SHAPES = [(500, 30, 0, 64), (800, 40, 70, 128), (333, 17, 5, 24), (2100, 32, 32, 64)]

# name, leading arguments, tolerance factor
LEAVES = [
    ("lila_dgeqrf_w03_mt", ("p0",), 100),
    ("lila_dgeqrf_w03_mt", ("p1",), 100),
    ("lila_dgeqrf_w03_mt", ("p2",), 2000),
    ("lila_dgeqrf_w03_mt_hh", ("p1",), 100),
    ("lila_dgeqrf_w03_mt_l", (), 100),
    ("lila_dgeqr2_w03", ("q0",), 100),
    ("lila_dgeqr2_w03", ("q1",), 100),
    ("lila_dgeqr2_w03", ("q2",), 2000),
    ("lila_dgeqr2_w03_3", (), 100),
    ("lila_dgeqr2_w03_l", (), 100),
    ("lila_dgeqr2_w03_hr", (), 2000),
]

PARAMS = {
    # lila_param = {mode, leaf, panel, nx}: pK selects the leaf, qK the panel
    "p0": [0, 0, 1, 32], "p1": [0, 1, 1, 32], "p2": [0, 2, 1, 32],
    "q0": [0, 1, 0, 32], "q1": [0, 1, 1, 32], "q2": [0, 1, 2, 32],
}


CASES = [(e, s) for e in LEAVES for s in (PANEL_SHAPES if e[0].startswith("lila_dgeqr2")
                                          else LEAF_SHAPES)]


@pytest.mark.parametrize("where", ["dev", "host"])
@pytest.mark.parametrize("leaf,shape", CASES, ids=lambda e: "-".join(map(str, e if isinstance(
    e, tuple) and isinstance(e[0], int) else [e[0]] + list(e[1]))))
def test_leaf_and_panel_entries(lila, ora, leaf, shape, where):
    name, head, tolf = leaf
    m, n, ii, mt = shape
    say("entry", name, head, shape, where)
    hargs = [np.array(PARAMS[h], dtype=np.int32) for h in head]
    A, T, Q = _inputs(ora, m, n, ii, mt, 100 + m + n + ii)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call(name, *hargs, m, n, ii, mt, Ao, m, To, mt, Qo, m, None, 0)
    g = _run(lila, name, hargs, A, T, Q, m, n, ii, mt, where)
    _check(ora, A, (Ao, To, Qo), g, m, n, ii, tolf)


@pytest.mark.parametrize("where", ["dev", "host"])
@pytest.mark.parametrize("m,k,n,mt", [(900, 70, 50, 32), (1200, 256, 100, 100), (500, 1, 33, 16)])
def test_appendcols_sequence(lila, ora, m, k, n, mt, where):
    """lila-c/lila_dgeqrf_w03_appendcols.c:7-10, each call through the ABI:
    factor k columns, then append n columns: dormqrf_w03 (Q^T onto the new
    columns), recursive QR of the new block at i = k, dlarft_connect_w03 and
    dormqrbz_w03 (the old reflectors onto the new part of Q)"""
    say("appendcols sequence", m, k, n, mt, where)
    p = np.array([0, 1, 1, 32], dtype=np.int32)
    A = ora.fill_normal(31 + k, m, k + n)
    Q = np.zeros((m, k + n), order="F")
    T = np.zeros((mt, k + n), order="F")
    # first k columns by the oracle (the state a caller holds before appending)
    ora.call("lila_dgeqrf_w03_recursive", p, m, k, 0, mt, A, m, T, mt, Q, m, None, 0)
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ora.call("lila_dormqrf_w03", m, n, k, 0, k, mt, Ao, m, To, mt, None, 0)
    ora.call("lila_dgeqrf_w03_recursive", p, m, n, k, mt, Ao, m, To, mt, Qo, m, None, 0)
    ora.call("lila_dlarft_connect_w03", m, n, k, 0, mt, Ao, m, To, mt)
    ora.call("lila_dormqrbz_w03", m, n, k, 0, k, mt, Ao, m, Qo, m, To, mt, None, 0)
    if where == "dev":
        Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    else:
        Ad, Td, Qd = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    assert lila.call("lila_dormqrf_w03", m, n, k, 0, k, mt, Ad, m, Td, mt, None, 0) == 0
    assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, k, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    assert lila.call("lila_dlarft_connect_w03", m, n, k, 0, mt, Ad, m, Td, mt) == 0
    assert lila.call("lila_dormqrbz_w03", m, n, k, 0, k, mt, Ad, m, Qd, m, Td, mt, None, 0) == 0
    if where == "dev":
        sync_ok(lila)
        Ag, Tg, Qg = to_host(Ad), to_host(Td), to_host(Qd)
    else:
        assert lila.call("lila_get_info") == 0
        Ag, Tg, Qg = Ad, Td, Qd
    tol = 100 * (k + n) * U64
    assert rel(Ag, Ao) < tol and rel(Qg, Qo) < tol and rel(Tg, To) < 10 * tol
    # the whole m x (k+n) factorization: LAPACK's R and Q of the full matrix
    A0 = ora.fill_normal(31 + k, m, k + n)
    assert ora.call("test_qq_orth_full", m, k + n, 0, Qg, m) < tol
    assert ora.call("test_qr_repres_1", m, k + n, 0, A0, m, Qg, m, Ag, m) < tol
    assert r_entrywise(A0, Ag, m, k + n, 0, m) < tol


@pytest.mark.parametrize("m,n,i,mt", [(400, 70, 0, 16), (600, 90, 9, 32), (900, 130, 21, 50)])
def test_ormhr2_w03_hr_blocks(lila, ora, m, n, i, mt):
    """the blocked reconstruction called one mt block at a time through the
    ABI, as lila-c/lila_dgeqrf_w03_mt_hr.c:23-29 does: every block after the
    first (jb = j - i > 0) flips the earlier rows, solves with the earlier
    blocks' unit lower V (a plain triangular solve on this path) and
    updates before its own sign-chosen LU"""
    say("ormhr2 blocks", m, n, i, mt)
    A = ora.fill_normal(41, m, i + n)
    Q = np.zeros((m, i + n), order="F")
    T = np.zeros((mt, i + n), order="F")
    # Cholesky-QR of A(i:m, i:i+n) by the oracle: R in A's top, Q1 in Q
    Ao, To, Qo = A.copy(order="F"), T.copy(order="F"), Q.copy(order="F")
    ml = m - i
    Qo[i:, i:] = Ao[i:, i:]
    G = np.zeros((n, n), order="F")
    Qs = np.asfortranarray(Qo[i:, i:])
    ora.call("dsyrk", "U", "T", n, ml, 1.0, Qs, ml, 0.0, G, n)
    assert ora.call("dpotrf_u", n, G, n) == 0
    ora.call("dtrsm", "R", "U", "N", "N", ml, n, 1.0, G, n, Qs, ml)
    Qo[i:, i:] = Qs
    Ao[i:i + n, i:] = np.triu(G)
    Ag_in, Tg_in, Qg_in = Ao.copy(order="F"), To.copy(order="F"), Qo.copy(order="F")
    So = np.zeros(i + n)
    Ad, Td, Qd = to_dev(Ag_in), to_dev(Tg_in), to_dev(Qg_in)
    Sd = torch.zeros(i + n, dtype=torch.float64, device="cuda")
    j, vb, steps = i, min(mt - i % mt, n), 0
    while vb > 0:
        ora.call("lila_ormhr2_w03_hr", m, vb, i, j, -1, mt, Ao, m, To, mt, Qo, m, None, 0, So)
        assert lila.call("lila_ormhr2_w03_hr", m, vb, i, j, -1, mt, Ad, m, Td, mt, Qd, m, None, 0, Sd) == 0
        steps += 1
        j += vb
        vb = min(mt, i + n - j)
    assert steps >= 3
    sync_ok(lila)
    assert np.array_equal(Sd.cpu().numpy()[i:], So[i:])
    tol = 1000 * n * U64
    assert rel(to_host(Ad), Ao) < tol and rel(to_host(Td), To) < tol and rel(to_host(Qd), Qo) < tol
    # the reconstruction is the Householder QR of A(i:m, i:i+n): LAPACK's R
    assert r_entrywise(A, to_host(Ad), m, n, i, m) < tol


@pytest.mark.parametrize("m,n,ii,mt,nx", [(2000, 1000, 0, 64, 32), (1413, 692, 97, 200, 32)])
def test_recursive_r_entrywise(lila, ora, m, n, ii, mt, nx):
    """config-1 shape: R entrywise against LAPACK (up to row sign), besides
    the oracle comparison and the invariants"""
    say("recursive R entrywise", m, n, ii, mt)
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    A = ora.fill_uniform(2 + m, m, n + ii, -1.0, 1.0)
    T = np.zeros((mt, n + ii), order="F")
    Q = np.zeros((m, n + ii), order="F")
    Ad, Td, Qd = to_dev(A), to_dev(T), to_dev(Q)
    assert lila.call("lila_dgeqrf_w03_recursive", p, m, n, ii, mt, Ad, m, Td, mt, Qd, m, None, 0) == 0
    sync_ok(lila)
    Ag = to_host(Ad)
    assert r_entrywise(A, Ag, m, n, ii, m) < 100 * n * U64


def _safe_rel(a, b):
    s = float(np.max(np.abs(b)))
    if s == 0.0:
        return float(np.max(np.abs(a)))
    return np.linalg.norm(a / s - b / s) / max(np.linalg.norm(b / s), 1e-300)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("m,n,scale,bigcol", [
    (3000, 32, 1e-200, None), (3000, 32, 1e200, None), (3000, 20, 1.0, (5, 1e160)),
    (5000, 1, 1e200, None), (5000, 1, 1e-200, None), (4000, 1, 1.0, (0, 1e160)),
    (200000, 32, 1e-200, None), (200000, 16, 1e200, None)])
def test_dgeqr3_extreme_scales(lila, ora, mode, m, n, scale, bigcol):
    """dlarfg's safe rescaling (lila-c/dgeqr3.c:12 delegates to LAPACK
    dlarfg): panels whose column norms underflow or overflow when squared,
    in both panel algorithms (column-by-column, and Cholesky-QR2 whose Gram
    underflows to zero or overflows and falls back), in LDS and streamed
    (200000 rows).  V (strict lower part) and T compared as they are, R
    relative to its largest entry; everything must stay finite."""
    say("dgeqr3 extreme", mode, m, n, scale, bigcol)
    A = ora.fill_uniform(7 + m + n, m, n, -1.0, 1.0) * scale
    if bigcol is not None:
        A[:, bigcol[0]] *= bigcol[1]
    A = np.asfortranarray(A)
    T0 = np.zeros((n, n), order="F")
    Ao, To = A.copy(order="F"), T0.copy(order="F")
    ora.call("dgeqr3", m, n, Ao, m, To, n)
    old = lila.call("lila_set_panel_mode", mode)
    try:
        Ad, Td = to_dev(A), to_dev(T0)
        assert lila.call("dgeqr3", m, n, Ad, m, Td, n) == 0
        sync_ok(lila)
    finally:
        lila.call("lila_set_panel_mode", old)
    Ag, Tg = to_host(Ad), to_host(Td)
    assert np.all(np.isfinite(Ag)) and np.all(np.isfinite(Tg))
    tol = 100 * max(n, 8) * U64
    Vg, Vo = np.tril(Ag, -1), np.tril(Ao, -1)
    assert _safe_rel(Vg, Vo) < tol, "V"
    assert _safe_rel(np.triu(Ag[:n]), np.triu(Ao[:n])) < tol, "R"
    assert _safe_rel(np.triu(Tg), np.triu(To)) < tol, "T"
    tau = np.diag(Tg)
    assert np.all((tau == 0) | ((tau >= 1.0) & (tau <= 2.0))), "tau outside [1, 2]"
