"""lila_amd -- the LILA Householder-QR hot path on one AMD MI355X.

The product is the C ABI of lila_amd/liblila_hip.so (declared in
include/lila.h).  This package only locates, builds and loads it, and
converts Python arguments (ints, floats, numpy arrays, torch tensors) for
ctypes calls.  There is no CPU fallback: without the built library or a GPU
every call fails.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblila_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include", "lila.h")

_lib = None


def _build_script():
    """lila_amd/build.py, loaded from its file under a name of its own:
    importing it as the submodule lila_amd.build would replace the build()
    function below on the package (and `from . import build` returns that
    function, not the script)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("_lila_amd_build_script",
                                                  os.path.join(HERE, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def build(verbose=False):
    """Compile liblila_hip.so in-tree."""
    return _build_script().build(verbose=verbose)


def load():
    """Load liblila_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("lila_amd: %s is not built (run __graft_entry__.build())" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        _lib.lila_get_stream.restype = ctypes.c_void_p
        _lib.lila_dgels_batched_plan.restype = ctypes.c_void_p
        _lib.lila_dgels_batched_flops.restype = ctypes.c_double
        _lib.lila_dgels_batched_flops.argtypes = [ctypes.c_void_p]
        _lib.lila_dgels_batched_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.lila_dgels_batched_destroy.argtypes = [ctypes.c_void_p]
    return _lib


def _ptr(a):
    if a is None:
        return ctypes.c_void_p(None)
    if hasattr(a, "data_ptr"):  # torch tensor
        return ctypes.c_void_p(a.data_ptr())
    if hasattr(a, "ctypes"):  # numpy array
        return ctypes.c_void_p(a.ctypes.data)
    raise TypeError(type(a))


def conv(a):
    if a is None:
        return ctypes.c_void_p(None)
    if isinstance(a, bool):
        return ctypes.c_int(int(a))
    if isinstance(a, int):
        return ctypes.c_int(a)
    if isinstance(a, float):
        return ctypes.c_double(a)
    if isinstance(a, (ctypes._SimpleCData, ctypes.Array)):
        return a
    try:
        import numpy as np
        if isinstance(a, np.integer):
            return ctypes.c_int(int(a))
        if isinstance(a, np.floating):
            return ctypes.c_double(float(a))
    except ImportError:
        pass
    return _ptr(a)


def call(name, *args):
    """Call an exported lila.h function; returns its int result."""
    fn = getattr(load(), name)
    return fn(*[conv(a) for a in args])


def exported_names():
    """Function names declared in include/lila.h."""
    import re
    txt = open(INCLUDE).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|double|long int)[\s*]+(\w+)\s*\(", txt,
                                 flags=re.M)))


def autotune(m, n, reps=1, mt=512, nx=32):
    """Run lila_autotune on an m x n problem from (mt, nx); returns the chosen
    (mt, nx) and the report text.  The GEMM and panel settings it picks stay
    in the library for the rest of the process."""
    cmt, cnx = ctypes.c_int(mt), ctypes.c_int(nx)
    buf = ctypes.create_string_buffer(4096)
    r = load().lila_autotune(ctypes.c_int(m), ctypes.c_int(n), ctypes.c_int(reps),
                             ctypes.byref(cmt), ctypes.byref(cnx), buf, ctypes.c_int(len(buf)))
    if r != 0:
        raise RuntimeError("lila_autotune(%d, %d) failed" % (m, n))
    return cmt.value, cnx.value, buf.value.decode()
