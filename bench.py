"""Benchmark of the LILA Householder-QR hot path on one MI355X.

Workload (one "step"): the plain FP64 drop-in -- lila_dgeqrf_w03_recursive
(leaf = hh, panel = fused dgeqr3, VRTQ: V, R, T and the explicit Q) of one
m = n = 16384 column-major matrix with entries iid Uniform(-1,1) from the
splitmix64 generator, the matrix already resident in HBM (the factorization
overwrites its input, so each timed step factors its own copy, made before
the timed region; config.inputs says so, or that the copies did not fit and
each step copied the input inside the timing).  Rate = the reference's VRTQ
count 4mn^2 - 4/3 n^3
(lila-c/main_w03.c:200) / time.

Besides the JSON line's contract fields it reports:
  roofline     -- the dominant kernel class (most time in the first warm-up
                  step, whose launches are all timed with HIP events), timed
                  again live during the timed steps with HIP events on the
                  library stream around each of its launches: achieved = its
                  algorithmic flops / its summed launch time, against FP64
                  MFMA peak (78.6 TFLOP/s);
                  traffic is the class's HBM bytes per launch from the
                  committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
                  (profiles/*_traffic.md), null when none matches;
  kernel_classes -- per-class launches, flops, bytes and event time of the
                  first warm-up step;
  mixed        -- the mixed-precision form (lila_param[0] = 1) on the same
                  input, and the per-kappa accuracy of both forms;
  batch        -- 4096 uneven least-squares problems (m ~ U{64..4096},
                  n ~ U{8..min(128,m)}, nrhs ~ U{1..16}) solved by one
                  persistent launch replayed from a hipGraph: QR/s, and the
                  oracle's per-problem dgels on the host cores beside it;
  tall_cholqr  -- the tall-skinny Cholesky-QR path at m = 1e6, n = 256: the
                  Gram A^T A in FP64, FP32 and 3-term split-BF16 MFMA
                  arithmetic (GB/s against the 8 TB/s HBM roofline of one
                  read of A), and lila_dgeqrf_w03_mt_hr with each Gram
                  precision (time, orthogonality, backward error);
  least_squares -- one m = 100000, n = 1024, nrhs = 64 least-squares problem
                  through lila_dgels_qr3, with host LAPACK dgels beside it;
  cpu_baseline -- the CPU oracle (a C restatement of LILA's w03 recursion,
                  BLAS-3 steps through the host OpenBLAS, all its threads) on
                  a bounded sample of the same workload, rank 0 only;
  accuracy     -- ||A - QR||_F/||A||_F and ||I - Q^T Q||_F of the last step,
                  computed on the GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 matrix
HBM_PEAK = 8000.0      # GB/s
METRIC = ("QR TFLOP/s mixed vs plain FP64 at backward err<=1e-13, m=n=16384; batch QR/s")


def vrtq_flops(m, n):
    return 4.0 * m * n * n - 4.0 / 3.0 * n ** 3


def parse_prof(txt):
    rows = []
    for line in txt.strip().splitlines():
        name, launches, flops, nbytes, ms = line.split("|")
        rows.append(dict(name=name, launches=int(launches), flops=float(flops),
                         bytes=float(nbytes), ms=float(ms)))
    return rows


def qr_errors(lila_amd, torch, dev, A0, A, Q, m, n):
    """||A0 - Q R||_F / ||A0||_F and ||I - Q^T Q||_F, R = upper triangle of A"""
    f64 = torch.float64
    R = torch.tril(A)  # upper triangle of the column-major A
    QR = torch.empty_like(A0)
    lila_amd.call("lila_dgemm", ctypes.c_char(b"N"), ctypes.c_char(b"N"), m, n, n, 1.0, Q, m, R, m,
                  0.0, QR, m)
    torch.cuda.synchronize()
    bwd = (torch.linalg.norm(QR - A0) / torch.linalg.norm(A0)).item()
    del QR, R
    G = torch.eye(n, dtype=f64, device=dev)
    lila_amd.call("lila_dgemm", ctypes.c_char(b"T"), ctypes.c_char(b"N"), n, n, m, 1.0, Q, m, Q, m,
                  -1.0, G, n)
    torch.cuda.synchronize()
    orth = torch.linalg.norm(G).item()
    return bwd, orth


INPUTS_NOTE = {True: "one copy of the input per timed step, in HBM before the timed region",
               False: "the input copied from a pristine device copy inside each timed step"}


def staged_inputs(torch, A0, k):
    """k device copies of A0, made before a timed region whose steps each
    factor one in place; None when they would take more than half of the
    free device memory (the steps then copy A0 inside the timing)"""
    free, _ = torch.cuda.mem_get_info()
    if k * A0.numel() * A0.element_size() > free // 2:
        return None
    return [A0.clone() for _ in range(k)]


def mixed_leg(lila_amd, lib, torch, dev, A0, A, Q, T, m, n, mt, nx, steps, bwd_plain, orth_plain,
              dt_plain):
    """mixed precision (lila_param[0] = 1) on the same matrix, timed like the
    plain steps and counted with the same VRTQ flops; then both forms on
    A = U diag(sigma) V^T with kappa in {1e2, 1e5, 1e8} (U, V: Q factors of
    seeded Gaussians from the plain path, sigma geometric from 1 to 1/kappa)"""
    pm = np.array([1, 1, 1, nx], dtype=np.int32)
    pp = np.array([0, 1, 1, nx], dtype=np.int32)

    def run(p, src, buf=None):
        if buf is None:
            A.copy_(src)
        X = A if buf is None else buf
        r = lila_amd.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, X, m, T, mt, Q, m, None, 0)
        if r != 0:
            raise RuntimeError("lila_dgeqrf_w03_recursive (mode %d) returned %d" % (p[0], r))

    run(pm, A0)
    bufs = staged_inputs(torch, A0, steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        run(pm, A0, bufs[i] if bufs else None)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    if bufs:
        A.copy_(bufs[-1])
        del bufs
    bwd, orth = qr_errors(lila_amd, torch, dev, A0, A, Q, m, n)
    out = {"value": vrtq_flops(m, n) / dt / 1e12, "unit": "TFLOP/s", "ms_per_step": dt * 1e3,
           "speedup_vs_plain": dt_plain / dt,
           "accuracy": {"backward_error": bwd, "orthogonality": orth},
           "plain_accuracy": {"backward_error": bwd_plain, "orthogonality": orth_plain},
           "method": "FP32-MFMA Householder R0 (FP64 panels), FP64 Cholesky-QR on A R0^-1 (a second "
                     "pass unless its Gram G shows ||I - G||_2 <= 0.5), Householder "
                     "reconstruction (lila_dorghr); flops counted as plain VRTQ",
           "by_kappa": []}
    if m != n:
        return out
    # U and V from the plain path (Q of seeded Gaussians)
    f64 = torch.float64
    G = torch.empty_like(A0)
    UV = []
    for seed in (101, 202):
        lila_amd.call("lila_fill_normal", ctypes.c_ulonglong(seed), m, n, G, m)
        run(pp, G)
        UV.append(Q.clone())
    U, V = UV
    idx = torch.arange(n, dtype=f64, device=dev)
    Ak = torch.empty_like(A0)
    for kappa in (1e2, 1e5, 1e8):
        sig = kappa ** (-idx / (n - 1))
        US = U * sig[:, None]  # column j of the column-major U scaled by sigma_j
        lila_amd.call("lila_dgemm", ctypes.c_char(b"N"), ctypes.c_char(b"T"), m, n, n, 1.0, US, m, V,
                      n, 0.0, Ak, m)
        del US
        row = {"kappa": kappa}
        for name, p in (("plain", pp), ("mixed", pm)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(p, Ak)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            b, o = qr_errors(lila_amd, torch, dev, Ak, A, Q, m, n)
            row[name] = {"backward_error": b, "orthogonality": o, "ms": t * 1e3}
        out["by_kappa"].append(row)
    del U, V, G, Ak, UV
    torch.cuda.empty_cache()
    return out


def mixed_tall_leg(lila_amd, torch, dev, steps, m=200000, n=1024, mt=512, nx=32):
    """the tall shape of the mixed-precision workload: plain and mixed forms
    on A = U diag(sigma) V^T, m = 200000, n = 1024, kappa in {1e2, 1e5, 1e8}
    (U, V: Q factors of seeded Gaussians from the plain path, sigma geometric
    from 1 to 1/kappa); time per factorization (mean of `steps` after one
    untimed run of each form), TFLOP/s by the plain VRTQ count, and the
    errors of each form"""
    f64 = torch.float64
    pm = np.array([1, 1, 1, nx], dtype=np.int32)
    pp = np.array([0, 1, 1, nx], dtype=np.int32)
    A = torch.empty((n, m), dtype=f64, device=dev)  # column-major m x n
    Q = torch.empty_like(A)
    T = torch.zeros((n, mt), dtype=f64, device=dev)

    def run(p, src):
        A.copy_(src)
        r = lila_amd.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, A, m, T, mt, Q, m, None, 0)
        if r != 0:
            raise RuntimeError("lila_dgeqrf_w03_recursive (mode %d) returned %d" % (p[0], r))

    G = torch.empty_like(A)
    lila_amd.call("lila_fill_normal", ctypes.c_ulonglong(303), m, n, G, m)
    run(pp, G)
    U = Q.clone()
    Gn = torch.empty((n, n), dtype=f64, device=dev)
    lila_amd.call("lila_fill_normal", ctypes.c_ulonglong(404), n, n, Gn, n)
    An = Gn.clone()
    Qn = torch.empty_like(Gn)
    Tn = torch.zeros((n, mt), dtype=f64, device=dev)
    r = lila_amd.call("lila_dgeqrf_w03_recursive", pp, n, n, 0, mt, An, n, Tn, mt, Qn, n, None, 0)
    if r != 0:
        raise RuntimeError("lila_dgeqrf_w03_recursive returned %d" % r)
    V = Qn
    del G, Gn, An, Tn
    idx = torch.arange(n, dtype=f64, device=dev)
    Ak = torch.empty_like(A)
    out = {"m": m, "n": n, "mt": mt, "nx": nx, "flops_per_factorization": vrtq_flops(m, n),
           "by_kappa": []}
    for kappa in (1e2, 1e5, 1e8):
        sig = kappa ** (-idx / (n - 1))
        US = U * sig[:, None]
        lila_amd.call("lila_dgemm", ctypes.c_char(b"N"), ctypes.c_char(b"T"), m, n, n, 1.0, US, m, V,
                      n, 0.0, Ak, m)
        del US
        row = {"kappa": kappa}
        for name, p in (("plain", pp), ("mixed", pm)):
            run(p, Ak)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                run(p, Ak)
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) / steps
            b, o = qr_errors(lila_amd, torch, dev, Ak, A, Q, m, n)
            row[name] = {"ms": t * 1e3, "tflops": vrtq_flops(m, n) / t / 1e12,
                         "backward_error": b, "orthogonality": o}
        row["mixed_speedup_vs_plain"] = row["plain"]["ms"] / row["mixed"]["ms"]
        out["by_kappa"].append(row)
    del A, Q, T, U, V, Ak
    torch.cuda.empty_cache()
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summaries (profiles/*_traffic.md: FETCH_SIZE and WRITE_SIZE runs scaled
    by a copy of known size), or None"""
    import glob
    for path in glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md")):
        name, val = None, None
        for line in open(path):
            if line.startswith("kernel: "):
                name = line[len("kernel: "):].strip()
            elif line.startswith("traffic_bytes_per_launch: "):
                val = float(line.split(":", 1)[1])
        if name == kernel and val is not None:
            return val
    return None


def read_prof(lib):
    buf = ctypes.create_string_buffer(1 << 16)
    lib.lila_prof_read(buf, len(buf))
    return parse_prof(buf.value.decode())


def cpu_baseline(m_target, mt, nx, budget_s):
    """oracle restatement + host OpenBLAS on a bounded square sample"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    path = pyoracle.use_host_blas()
    nthreads = None
    if path:
        try:
            ob = ctypes.CDLL(path)
            nthreads = int(ob.scipy_openblas_get_num_threads())
        except Exception:
            nthreads = None

    def run(nn):
        A = pyoracle.fill_uniform(7, nn, nn, -1.0, 1.0)
        Q = np.zeros((nn, nn), order="F")
        T = np.zeros((mt, nn), order="F")
        p = np.array([0, 1, 1, nx], dtype=np.int32)
        t0 = time.perf_counter()
        pyoracle.call("lila_dgeqrf_w03_recursive", p, nn, nn, 0, mt, A, nn, T, mt, Q, nn, None, 0)
        return time.perf_counter() - t0

    # grow the sample while a run takes under a twentieth of the budget, then
    # size the timed run from the last measured rate
    nn = 1024
    t = run(nn)
    while t < budget_s / 20 and nn < m_target:
        nn = min(m_target, nn * 2)
        t = run(nn)
    rate = vrtq_flops(nn, nn) / t
    target = int((budget_s * rate / (8.0 / 3.0)) ** (1.0 / 3.0)) // 64 * 64
    nn = max(nn, min(m_target, target))
    t = run(nn)
    rate = vrtq_flops(nn, nn) / t
    pyoracle.use_builtin_blas()
    # beside it the reference's own comparison, LAPACK dgeqrf + dorgqr
    # (lila-c/main_vq_LAPACK.c:100-102) on the host LAPACK, same sample size,
    # same VRTQ flop count, after a warm-up call at 2048
    from scipy.linalg import lapack

    def lapack_vrtq(X):
        # optimal workspaces (a minimal one makes dgeqrf fall back to dgeqr2)
        lw1 = int(lapack.dgeqrf(X, lwork=-1)[2][0])
        t0 = time.perf_counter()
        qr, tau, _, info = lapack.dgeqrf(X, lwork=lw1, overwrite_a=1)
        lw2 = int(lapack.dorgqr(qr, tau, lwork=-1)[1][0])
        q, _, info2 = lapack.dorgqr(qr, tau, lwork=lw2, overwrite_a=1)
        assert info == 0 and info2 == 0
        return time.perf_counter() - t0

    lapack_vrtq(np.asfortranarray(pyoracle.fill_uniform(8, 2048, 2048, -1.0, 1.0)))
    tl = lapack_vrtq(np.asfortranarray(pyoracle.fill_uniform(7, nn, nn, -1.0, 1.0)))
    lap = {"value": vrtq_flops(nn, nn) / tl / 1e12, "unit": "TFLOP/s",
           "cores": nthreads or os.cpu_count(),
           "sample": "LAPACK dgeqrf + dorgqr (SciPy's OpenBLAS, %s threads) of the same m=n=%d "
                     "matrix, %.2f s" % (nthreads, nn, tl)}
    return {
        "value": rate / 1e12,
        "unit": "TFLOP/s",
        "cores": nthreads or os.cpu_count(),
        "kind": "port",
        "sample": ("lila_dgeqrf_w03_recursive VRTQ of one m=n=%d Uniform(-1,1) matrix, mt=%d, nx=%d, "
                   "leaf=hh, panel=dgeqr3 (C restatement of lila-c, BLAS-3 steps on host OpenBLAS "
                   "with %s threads), %.1f s" % (nn, mt, nx, nthreads, t)),
        "lapack": lap,
    }


def batch_sizes(seed, nprob):
    """the batched workload's problem sizes: m ~ U{64..4096}, n ~ U{8..min(128, m)},
    nrhs ~ U{1..16}"""
    rng = np.random.default_rng(seed)
    m = rng.integers(64, 4097, size=nprob)
    n = np.array([rng.integers(8, min(128, int(mm)) + 1) for mm in m])
    k = rng.integers(1, 17, size=nprob)
    return m.astype(np.int32), n.astype(np.int32), k.astype(np.int32)


def ls_flops(m, n, k):
    """least-squares/main_ls_qr3.c:125-129"""
    m, n, k = (np.asarray(x, dtype=np.float64) for x in (m, n, k))
    return 2.0 * m * n * n - n ** 3 / 3.0 + 4.0 * (m - n) * n * k + 4.0 * n * n * k


def _dgels_worker(args):
    """one worker process of the batched CPU baseline: host LAPACK dgels
    (least-squares/main_lapack_ref.c:100), single-threaded, on its share of
    the problems; returns (problems done, wall seconds of its loop)"""
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    from scipy.linalg import lapack
    probs, seed, budget_s = args
    data = []
    for (p, mm, nn, kk) in probs:
        rng = np.random.default_rng(1000003 * seed + p)
        data.append((np.asfortranarray(rng.standard_normal((mm, nn))),
                     np.asfortranarray(rng.standard_normal((mm, kk)))))
    done = 0
    t0 = time.perf_counter()
    for A, B in data:
        info = lapack.dgels(A, B, overwrite_a=1, overwrite_b=1)[-1]
        assert info == 0
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return done, time.perf_counter() - t0


def batch_cpu_baseline(m, n, k, budget_s, seed):
    """host LAPACK dgels per problem, one single-threaded process per core, on a
    random sample of the batch; QR/s = problems done / wall clock of the
    slowest worker's loop (inputs generated before the clocks start)"""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    workers = max(1, min(16, os.cpu_count() or 1))
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(m))
    # a sample sized to about budget_s of work at ~0.3 ms per problem per worker
    nsamp = min(len(order), workers * max(8, int(budget_s / 3e-4)))
    samp = order[:nsamp]
    shares = [[(int(p), int(m[p]), int(n[p]), int(k[p])) for p in samp[w::workers]]
              for w in range(workers)]
    ctx = mp.get_context("spawn")
    with ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as ex:
        res = list(ex.map(_dgels_worker, [(sh, seed, budget_s) for sh in shares]))
    done = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    return {
        "value": done / wall,
        "unit": "QR/s",
        "cores": workers,
        "kind": "reference",
        "sample": ("host LAPACK dgels('N') per problem (SciPy's OpenBLAS, one thread per process, "
                   "%d processes), %d of the %d problems drawn at random; rate = problems / wall "
                   "clock of the slowest process's loop, %.2f s" % (workers, done, len(order), wall)),
    }


def ls_leg(lila_amd, lib, torch, steps, no_cpu, m=100_000, n=1024, k=64):
    """One large least-squares problem through lila_dgels_qr3 (the sequence of
    least-squares/main_ls_qr3.c:94-119: dgeqr3 with T, Q^T B, R X = (Q^T B)_1),
    N(0,1) A (m x n) and B (m x k) resident in HBM, restored from pristine
    copies outside the timed calls; rate by the count of
    least-squares/main_ls_qr3.c:125-129.  Beside it, host LAPACK dgels (the
    reference's comparison, least-squares/main_lapack_ref.c:100) on the same
    problem, all host threads."""
    f64 = torch.float64
    A0 = torch.empty((n, m), dtype=f64, device="cuda")
    B0 = torch.empty((k, m), dtype=f64, device="cuda")
    lila_amd.call("lila_fill_normal", ctypes.c_ulonglong(20241206), m, n, A0, m)
    lila_amd.call("lila_fill_normal", ctypes.c_ulonglong(20241207), m, k, B0, m)
    A, B = torch.empty_like(A0), torch.empty_like(B0)
    T = torch.zeros((n, n), dtype=f64, device="cuda")
    reps = max(3, steps)
    tot = 0.0
    for r in range(reps + 1):
        A.copy_(A0)
        B.copy_(B0)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = lila_amd.call("lila_dgels_qr3", m, n, k, A, m, T, n, B, m)
        e1.record()
        torch.cuda.synchronize()
        if r > 0:
            tot += e0.elapsed_time(e1)
    ms = tot / reps
    flops = float(ls_flops(m, n, k))
    X = B[:, :n].T  # n x k
    Am, Bm = A0.T, B0.T
    res = Bm - Am @ X
    nA = torch.linalg.norm(Am).item()
    eta = (torch.linalg.norm(Am.T @ res) /
           (nA * (nA * torch.linalg.norm(X) + torch.linalg.norm(Bm)))).item()
    out = {"workload": "lila_dgels_qr3 m=%d n=%d nrhs=%d N(0,1)" % (m, n, k),
           "ms": ms, "TFLOP/s": flops / (ms * 1e-3) / 1e12, "return": rc,
           "normal_equation_residual": eta}
    if not no_cpu:
        import scipy.linalg.lapack as sl
        Ah = np.asfortranarray(A0.T.cpu().numpy())
        Bh = np.asfortranarray(B0.T.cpu().numpy())
        t0 = time.perf_counter()
        sl.dgels(Ah, Bh, overwrite_a=True, overwrite_b=True)
        tc = time.perf_counter() - t0
        out["cpu_lapack_dgels"] = {"s": tc, "TFLOP/s": flops / tc / 1e12,
                                   "threads": "host LAPACK default (all host threads)"}
    del A0, B0, A, B
    return out


def tall_leg(lila_amd, lib, torch, steps, m=1_000_000, n=256):
    """Tall-skinny Cholesky-QR path at m = 1e6, n = 256 (N(0,1) entries,
    resident in HBM): the Gram A^T A in each arithmetic (FP64 MFMA, FP32
    MFMA, 3-term split-BF16 MFMA; one read of A, so the HBM roofline is
    8 m n bytes / 8 TB/s), then lila_dgeqrf_w03_mt_hr with each Gram precision
    (Cholesky-QR once, or twice with that first Gram, then the Householder
    reconstruction of V, T and Q).  Times: events on the library stream
    around the calls only."""
    f64 = torch.float64
    A0 = torch.empty((n, m), dtype=f64, device="cuda")
    lila_amd.call("lila_fill_normal", ctypes.c_ulonglong(20241205), m, n, A0, m)
    G = torch.zeros((n, n), dtype=f64, device="cuda")
    reps = max(3, steps)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    names = {1: "fp64_mfma", 2: "fp32_mfma", 3: "bf16x3_mfma"}
    gram, Gs = {}, {}
    nbytes = 8.0 * m * n
    for prec in (1, 2, 3):
        lila_amd.call("lila_dgram", prec, m, n, A0, m, G, n)
        torch.cuda.synchronize()
        e0, e1 = ev(), ev()
        e0.record()
        for _ in range(reps):
            lila_amd.call("lila_dgram", prec, m, n, A0, m, G, n)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        Gs[prec] = torch.triu(G).clone()
        gram[names[prec]] = {"ms": ms, "GB/s": nbytes / (ms * 1e-3) / 1e9,
                             "frac_hbm": nbytes / (ms * 1e-3) / HBM_PEAK / 1e9,
                             "TFLOP/s": m * n * (n + 1) / (ms * 1e-3) / 1e12}
    for prec in (2, 3):
        gram[names[prec]]["rel_err_vs_fp64"] = (torch.linalg.norm(Gs[prec] - Gs[1]) /
                                                torch.linalg.norm(Gs[1])).item()
    # the split-BF16 kernel alone (without the split-K reduction), from the
    # launch profile
    lib.lila_prof_enable(1)
    lila_amd.call("lila_dgram", 3, m, n, A0, m, G, n)
    torch.cuda.synchronize()
    rows = {r["name"]: r for r in read_prof(lib)}
    lib.lila_prof_enable(0)
    if "gram3_kernel" in rows:
        k = rows["gram3_kernel"]
        kms = k["ms"] / k["launches"]
        gram["bf16x3_mfma"]["kernel_ms"] = kms
        gram["bf16x3_mfma"]["kernel_frac_hbm"] = nbytes / (kms * 1e-3) / HBM_PEAK / 1e9
    A = torch.empty_like(A0)
    Q = torch.empty_like(A0)
    T = torch.zeros((n, n), dtype=f64, device="cuda")
    nA = torch.linalg.norm(A0).item()
    path = {}
    for prec in (0, 1, 2, 3):
        lib.lila_set_gram_precision(prec)
        tot = 0.0
        for r in range(reps + 1):
            A.copy_(A0)
            torch.cuda.synchronize()
            e0, e1 = ev(), ev()
            e0.record()
            rc = lila_amd.call("lila_dgeqrf_w03_mt_hr", m, n, 0, n, A, m, T, n, Q, m, None, 0)
            e1.record()
            torch.cuda.synchronize()
            if r > 0:
                tot += e0.elapsed_time(e1)
        info = lila_amd.call("lila_get_info")
        Qm = Q.T
        orth = torch.linalg.norm(Qm.T @ Qm - torch.eye(n, dtype=f64, device="cuda")).item()
        bwd = torch.linalg.norm(A0.T - Qm @ torch.triu(A.T[:n])).item() / nA
        path[["one_pass_fp64", "two_pass_fp64", "two_pass_fp32", "two_pass_bf16x3"][prec]] = {
            "ms": tot / reps, "return": rc, "info": info, "orthogonality": orth,
            "backward_error": bwd}
    lib.lila_set_gram_precision(0)
    del A0, A, Q
    return {"workload": "lila_dgeqrf_w03_mt_hr (Cholesky-QR + Householder reconstruction) and "
                        "its Gram, m=%d, n=%d, N(0,1)" % (m, n),
            "gram": gram, "cholqr_path": path,
            "accuracy_per_kappa": "tests/test_gpu_gram.py (kappa 1e1, 1e3, 1e6)"}


def batch_leg(lila_amd, lib, torch, dev, steps, warmup, cpu_seconds, no_cpu, nprob=4096):
    """4096 uneven least-squares problems (m ~ U{64..4096}, n ~ U{8..128},
    m >= n, nrhs ~ U{1..16}, N(0,1) entries) in one persistent launch replayed
    from a hipGraph; QR/s = problems / time of the launch (inputs resident in
    HBM, restored from a pristine copy outside the timed region)."""
    seed = 20241204
    m, n, k = batch_sizes(seed, nprob)
    lda = m.copy()
    ldb = m.copy()
    offA = np.zeros(nprob, dtype=np.int64)
    offB = np.zeros(nprob, dtype=np.int64)
    off = 0
    for p in range(nprob):
        offA[p] = off
        off += int(m[p]) * int(n[p])
        offB[p] = off
        off += int(m[p]) * int(k[p])
    buf0 = torch.empty(off, dtype=torch.float64, device=dev)
    chunk = 1 << 28
    for c0 in range(0, off, chunk):
        c1 = min(off, c0 + chunk)
        lila_amd.call("lila_fill_normal", ctypes.c_ulonglong(seed + c0), c1 - c0, 1, buf0[c0:], c1 - c0)
    buf = buf0.clone()
    base = buf.data_ptr()
    pa = (base + 8 * offA).astype(np.uint64)
    pb = (base + 8 * offB).astype(np.uint64)
    plan = lib.lila_dgels_batched_plan(*[lila_amd.conv(a) for a in (nprob, m, n, k, pa, lda, pb, ldb)])
    if not plan:
        raise RuntimeError("lila_dgels_batched_plan failed")
    plan = ctypes.c_void_p(plan)
    flops = float(ls_flops(m, n, k).sum())
    times = []
    lib.lila_prof_only(b"batched_ls_kernel")
    for it in range(warmup + steps):
        buf.copy_(buf0)
        torch.cuda.synchronize()
        if it == warmup:
            lib.lila_prof_enable(1)
        t0 = time.perf_counter()
        if lib.lila_dgels_batched_run(plan, 1) != 0:
            raise RuntimeError("lila_dgels_batched_run failed")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if it >= warmup:
            times.append(t1 - t0)
    prof = read_prof(lib)
    lib.lila_prof_enable(0)
    lib.lila_prof_only(None)
    info = lila_amd.call("lila_get_info")
    dt = sum(times) / len(times)
    # accuracy on a sample of problems: least-squares/main_ls_qr3.c:157-168's
    # ||A^T r|| / (||A|| (||A|| ||x|| + ||b||)) and R's representation error
    out = buf.cpu().numpy()
    src = buf0.cpu().numpy()
    rng = np.random.default_rng(seed + 1)
    worst_orth, worst_rep = 0.0, 0.0
    for p in rng.choice(nprob, size=64, replace=False):
        mm, nn, kk = int(m[p]), int(n[p]), int(k[p])
        A = src[offA[p]:offA[p] + mm * nn].reshape((mm, nn), order="F")
        B = src[offB[p]:offB[p] + mm * kk].reshape((mm, kk), order="F")
        F = out[offA[p]:offA[p] + mm * nn].reshape((mm, nn), order="F")
        X = out[offB[p]:offB[p] + mm * kk].reshape((mm, kk), order="F")[:nn]
        r = A @ X - B
        nA = np.linalg.norm(A)
        worst_orth = max(worst_orth, np.linalg.norm(A.T @ r) / (nA * (nA * np.linalg.norm(X) + np.linalg.norm(B))))
        R = np.triu(F[:nn])
        # ||A^T A - R^T R|| / ||A||^2 checks R without forming Q
        worst_rep = max(worst_rep, np.linalg.norm(A.T @ A - R.T @ R) / nA ** 2)
    lib.lila_dgels_batched_destroy(plan)
    del buf, buf0
    kern = None
    if prof:
        d = prof[0]
        avg = d["ms"] / d["launches"]
        kern = {"kernel": d["name"], "launches": d["launches"], "avg_launch_ms": avg,
                "tflops": d["flops"] / (d["ms"] * 1e-3) / 1e12,
                "frac_fp64_mfma_peak": d["flops"] / (d["ms"] * 1e-3) / 1e12 / FP64_MFMA_PEAK}
    cpu = None
    if not no_cpu:
        cpu = batch_cpu_baseline(m, n, k, cpu_seconds, seed)
    return {
        "workload": ("batched uneven least squares (lila_dgels_batched plan, hipGraph replay): "
                     "%d problems, m ~ U{64..4096}, n ~ U{8..min(128,m)}, nrhs ~ U{1..16}, "
                     "N(0,1)" % nprob),
        "qr_per_s": nprob / dt,
        "ms_per_step": dt * 1e3,
        "tflops": flops / dt / 1e12,
        "flops_per_step": flops,
        "bytes_in": 8.0 * off,
        "kernel": kern,
        "accuracy": {"max_norm_orth": worst_orth, "max_gram_repres": worst_rep, "sampled": 64,
                     "device_status": info},
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=16384)
    ap.add_argument("--n", type=int, default=16384)
    # mt (T block) and nx (recursion cutoff) as lila_autotune picks them for
    # m = n = 16384 on MI355X (bench.py --autotune reproduces the choice)
    ap.add_argument("--mt", type=int, default=1024)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-prof", action="store_true")
    ap.add_argument("--gemm-tile", default=None, help="GEMM tile override, e.g. 64x64")
    ap.add_argument("--panel-mode", type=int, default=1,
                    help="1: Cholesky-QR2 + reconstruction panel (default), 0: column Householder")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched least-squares leg")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched least-squares leg")
    ap.add_argument("--no-tall", action="store_true",
                    help="skip the tall-skinny Cholesky-QR leg (m=1e6, n=256)")
    ap.add_argument("--tall-only", action="store_true", help="run only the tall-skinny leg")
    ap.add_argument("--no-ls", action="store_true",
                    help="skip the large least-squares leg (m=100000, n=1024, 64 right-hand sides)")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision legs (m=n=16384 and m=200000, n=1024) "
                         "and their per-kappa accuracy tables")
    ap.add_argument("--autotune", action="store_true",
                    help="run lila_autotune on the workload shape first (from --mt/--nx) and "
                         "measure with the settings it picks")
    args = ap.parse_args()
    if args.gemm_tile:
        os.environ["LILA_GEMM_TILE"] = args.gemm_tile
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the hot path runs on one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import lila_amd
    lib = lila_amd.load()
    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    dev = torch.device("cuda:0")
    if args.batch_only:
        print(json.dumps(batch_leg(lila_amd, lib, torch, dev, args.steps, args.warmup,
                                   args.cpu_seconds, args.no_cpu)))
        return
    if args.tall_only:
        print(json.dumps(tall_leg(lila_amd, lib, torch, args.steps)))
        return
    m, n, mt, nx = args.m, args.n, args.mt, args.nx
    tuned = None
    if args.autotune:
        lib.lila_set_panel_mode(args.panel_mode)
        t0 = time.perf_counter()
        mt, nx, report = lila_amd.autotune(m, n, reps=1, mt=mt, nx=nx)
        tuned = {"mt": mt, "nx": nx, "seconds": time.perf_counter() - t0, "report": report}
    f64 = torch.float64
    # column-major m x n storage: torch shape (n, m)
    A0 = torch.empty((n, m), dtype=f64, device=dev)
    lila_amd.call("lila_fill_uniform", ctypes.c_ulonglong(20241203), m, n, A0, m, -1.0, 1.0)
    A = torch.empty_like(A0)
    Q = torch.empty_like(A0)
    T = torch.zeros((n, mt), dtype=f64, device=dev)
    p = np.array([0, 1, 1, nx], dtype=np.int32)
    lib.lila_set_panel_mode(args.panel_mode)

    def step(buf=None):
        # each factorization overwrites its input: a staged copy, else A0 copied into A
        if buf is None:
            A.copy_(A0)
        X = A if buf is None else buf
        r = lila_amd.call("lila_dgeqrf_w03_recursive", p, m, n, 0, mt, X, m, T, mt, Q, m, None, 0)
        if r != 0:
            raise RuntimeError("lila_dgeqrf_w03_recursive returned %d" % r)

    # the first warm-up step records every kernel class (the per-class table);
    # the timed steps record only the dominant class, whose events then cost
    # a few hundred records per step instead of two per launch
    prof_all, dominant = [], None
    for w in range(args.warmup):
        if w == 0 and not args.no_prof:
            lib.lila_prof_enable(1)
        step()
        if w == 0 and not args.no_prof:
            prof_all = read_prof(lib)
            lib.lila_prof_enable(0)
            dominant = max(prof_all, key=lambda r: r["ms"])["name"]
    torch.cuda.synchronize()
    assert lila_amd.call("lila_get_info") == 0
    bufs = staged_inputs(torch, A0, args.steps)
    staged_ok = bufs is not None
    if not args.no_prof:
        lib.lila_prof_only(dominant.encode() if dominant else None)
        lib.lila_prof_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(bufs[i] if bufs else None)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if bufs:
        A.copy_(bufs[-1])  # the last step's factors, for the accuracy check
        del bufs
    info = lila_amd.call("lila_get_info")
    prof = []
    if not args.no_prof:
        prof = read_prof(lib)
        lib.lila_prof_enable(0)
        lib.lila_prof_only(None)
        if not prof_all:
            prof_all = prof
    dt = (t1 - t0) / args.steps
    flops = vrtq_flops(m, n)
    value = flops / dt / 1e12

    # accuracy of the last step, on the GPU
    R = torch.tril(A)  # upper triangle of the column-major A
    QR = torch.empty_like(A0)
    lila_amd.call("lila_dgemm", ctypes.c_char(b"N"), ctypes.c_char(b"N"), m, n, n, 1.0, Q, m, R, m,
                  0.0, QR, m)
    torch.cuda.synchronize()
    bwd = (torch.linalg.norm(QR - A0) / torch.linalg.norm(A0)).item()
    del QR, R
    G = torch.eye(n, dtype=f64, device=dev)
    lila_amd.call("lila_dgemm", ctypes.c_char(b"T"), ctypes.c_char(b"N"), n, n, m, 1.0, Q, m, Q, m,
                  -1.0, G, n)
    torch.cuda.synchronize()
    orth = torch.linalg.norm(G).item()
    del G

    mixed = None
    mixed_tall = None
    if not args.no_mixed:
        mixed = mixed_leg(lila_amd, lib, torch, dev, A0, A, Q, T, m, n, mt, nx, args.steps, bwd, orth, dt)
        mixed_tall = mixed_tall_leg(lila_amd, torch, dev, args.steps)

    ls = None
    if not args.no_ls:
        ls = ls_leg(lila_amd, lib, torch, args.steps, args.no_cpu)

    tall = None
    if not args.no_tall:
        tall = tall_leg(lila_amd, lib, torch, args.steps)

    batch = None
    if not args.no_batch:
        batch = batch_leg(lila_amd, lib, torch, dev, args.steps, args.warmup, args.cpu_seconds,
                          args.no_cpu)

    roof = None
    if prof:
        dom = max(prof, key=lambda r: r["ms"])
        bound = "mfma" if dom["name"].startswith("gemm") else "hbm"
        if bound == "mfma":
            ach = dom["flops"] / (dom["ms"] * 1e-3) / 1e12
            roof = {"bound": "mfma", "achieved": ach, "peak": FP64_MFMA_PEAK, "unit": "TFLOP/s",
                    "frac": ach / FP64_MFMA_PEAK, "traffic": measured_traffic(dom["name"]),
                    "kernel": dom["name"],
                    "launches": dom["launches"], "avg_launch_ms": dom["ms"] / dom["launches"],
                    "share_of_step": dom["ms"] / (dt * 1e3 * args.steps)}
        else:
            ach = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK, "unit": "GB/s",
                    "frac": ach / HBM_PEAK, "traffic": measured_traffic(dom["name"]),
                    "kernel": dom["name"],
                    "launches": dom["launches"], "avg_launch_ms": dom["ms"] / dom["launches"],
                    "share_of_step": dom["ms"] / (dt * 1e3 * args.steps)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, mt, nx, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain FP64 recursive Householder QR with explicit Q "
                               "(lila_dgeqrf_w03_recursive, VRTQ) m=n=%d Uniform(-1,1)" % n,
                   "m": m, "n": n, "mt": mt, "nx": nx, "leaf": "hh",
                   "panel": ("dgeqr3 via Cholesky-QR2 + Householder reconstruction"
                             if args.panel_mode else "dgeqr3 column by column"),
                   "flops_per_step": flops,
                   "inputs": INPUTS_NOTE[staged_ok]},
        "accuracy": {"backward_error": bwd, "orthogonality": orth, "device_status": info},
        "mixed": mixed,
        "mixed_tall": mixed_tall,
        "batch": batch,
        "tall_cholqr": tall,
        "least_squares": ls,
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernel_classes": prof_all,
        "autotune": tuned,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
