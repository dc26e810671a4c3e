"""ctypes access to the CPU oracle (oracle/liblila_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liblila_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liblila_oracle.so"], check=True)
    return SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        _lib = ctypes.CDLL(SO)
        for name in ("ora_test_qq_orth_1", "ora_test_qq_orth_full", "ora_test_qr_repres_1",
                     "ora_test_r_repres_2", "ora_dlange_f"):
            getattr(_lib, name).restype = ctypes.c_double
    return _lib


def conv(a):
    if a is None:
        return ctypes.c_void_p(None)
    if isinstance(a, (bool, np.bool_)):
        return ctypes.c_int(int(a))
    if isinstance(a, (int, np.integer)):
        return ctypes.c_int(int(a))
    if isinstance(a, (float, np.floating)):
        return ctypes.c_double(float(a))
    if isinstance(a, np.ndarray):
        if a.dtype not in (np.float64, np.int32):
            raise TypeError("oracle arrays must be float64 or int32, got %s" % a.dtype)
        if not (a.flags.f_contiguous or a.ndim == 1):
            raise TypeError("oracle arrays must be Fortran-ordered")
        return ctypes.c_void_p(a.ctypes.data)
    if isinstance(a, str):
        return ctypes.c_char(a.encode()[0])
    return a


def call(name, *args):
    """call ora_<name>(args); numpy arrays are passed by pointer (F order)"""
    fn = getattr(lib(), name if name.startswith("ora_") else "ora_" + name)
    return fn(*[conv(a) for a in args])


def fill_uniform(seed, m, n, lo=-1.0, hi=1.0, ld=None):
    ld = ld or m
    A = np.zeros((ld, n), order="F")
    lib().ora_fill_uniform(ctypes.c_ulonglong(seed), ctypes.c_int(m), ctypes.c_int(n),
                           ctypes.c_void_p(A.ctypes.data), ctypes.c_int(ld),
                           ctypes.c_double(lo), ctypes.c_double(hi))
    return A


def fill_normal(seed, m, n, ld=None):
    ld = ld or m
    A = np.zeros((ld, n), order="F")
    lib().ora_fill_normal(ctypes.c_ulonglong(seed), ctypes.c_int(m), ctypes.c_int(n),
                          ctypes.c_void_p(A.ctypes.data), ctypes.c_int(ld))
    return A


def use_host_blas():
    """route oracle BLAS-3 calls to the host OpenBLAS (CPU-baseline timing)"""
    import glob
    import scipy
    cands = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs",
                                   "libscipy_openblas*.so"))
    for path in cands:
        if lib().ora_use_host_blas(path.encode(), b"scipy_") == 0:
            return path
    return None


def use_builtin_blas():
    lib().ora_use_builtin_blas()
