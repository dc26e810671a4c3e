"""The C ABI library loads and exports every function include/lila.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os

import lila_amd


def test_library_exports_every_declared_symbol():
    names = lila_amd.exported_names()
    assert len(names) >= 40
    lib = ctypes.CDLL(lila_amd.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_keeps_reference_signatures():
    # the drop-in entry points named by the hot path are declared
    names = set(lila_amd.exported_names())
    for n in ["dgeqr3", "lila_dgeqrf_w03_recursive", "lila_dgeqrf_w03_levelx",
              "lila_dgeqrf_w03_mt", "lila_dgeqrf_w03_mt_hh", "lila_dgeqrf_w03_mt_l",
              "lila_dgeqrf_w03_mt_hr", "lila_dgeqr2_w03", "lila_dgeqr2_w03_l",
              "lila_dgeqr2_w03_3", "lila_dgeqr2_w03_hr", "lila_dlarft_w03",
              "lila_dlarft_connect_w03", "lila_dormqrf_w03", "lila_dormqrf_z03",
              "lila_dormqrbz_w03", "lila_ormhr2_w03_hr", "lila_dorgh2_3", "lila_dorgh2",
              "lila_dorghr", "lila_dgetrf_b", "lila_dT2tau_w03", "lila_dV2tau_w03"]:
        assert n in names, n


def test_library_has_no_cpu_path_symbols():
    # the product must not link the oracle
    out = os.popen("nm -D --defined-only %s" % lila_amd.LIB_PATH).read()
    assert "ora_" not in out


def test_build_entry_reaches_the_build_script():
    """lila_amd.build() (called by the build entry point) finds the compile
    script and stays callable afterwards (no compile here)"""
    script = lila_amd._build_script()
    assert callable(script.build)
    assert script.LIB == lila_amd.LIB_PATH
    assert callable(lila_amd.build)
