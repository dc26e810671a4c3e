"""The autotuner (lila_autotune) through the C ABI: it returns settings from
its candidate sets with a report of the trials, and the drivers stay at
parity with the oracle under the settings it keeps.  Collected after the
other GPU tests, since the tuned GEMM and panel settings stay in the library
for the rest of the process."""
import ctypes

import pytest

from conftest import say
from test_gpu_parity import check_vrtq, run_driver

pytestmark = pytest.mark.gpu


def test_autotune_choice_and_parity(lila, ora):
    say("autotune 3000 x 768 from mt 256, nx 32")
    mt, nx, report = lila.autotune(3000, 768, reps=1, mt=256, nx=32)
    say(report)
    assert mt in (256, 512, 1024) and nx in (32, 64, 128, 256, 512, 1024, 2048)
    assert "panel_nb" in report and "(plain form)" in report
    assert report.startswith("start mt 256 nx 32") and "chose mt %d nx %d" % (mt, nx) in report
    for (m, n, ii) in ((1500, 700, 0), (1213, 999, 13)):
        say("recursive after autotune", m, n, ii, mt, nx)
        A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, nx])
        check_vrtq(ora, A, o, g, m, n, ii, mt, m, m)


@pytest.mark.parametrize("nb", [16, 24, 8])
def test_panel_width_parity(lila, ora, nb):
    """the tunable panel width (dgeqr3's recursion cutoff) changes where the
    panel launches start, not the results"""
    old = lila.call("lila_set_panel_width", nb)
    try:
        for (m, n, ii, mt, nx) in ((1500, 700, 0, 64, 32), (1213, 333, 13, 100, 64)):
            say("recursive with panel width", nb, m, n, ii)
            A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, nx])
            check_vrtq(ora, A, o, g, m, n, ii, mt, m, m)
    finally:
        lila.call("lila_set_panel_width", old)
    assert lila.call("lila_set_panel_width", 33) == -1 and lila.call("lila_set_panel_width", 7) == -1


def test_autotune_mixed_form(lila, ora):
    """lila_autotune_form(1): the tuner times the mixed-precision form; the
    mixed path stays at parity with the oracle under the settings it keeps"""
    import numpy as np
    old = lila.call("lila_autotune_form", 1)
    try:
        say("autotune the mixed form 2000 x 600")
        mt, nx, report = lila.autotune(2000, 600, reps=1, mt=256, nx=32)
        say(report)
        assert "(mixed form)" in report
    finally:
        lila.call("lila_autotune_form", old)
    m, n = 1500, 700
    A, o, g = run_driver(lila, ora, "recursive", m, n, 0, mt, [1, 1, 1, nx])
    check_vrtq(ora, A, o, g, m, n, 0, mt, m, m, tolf=200)


def test_autotune_rejects(lila):
    lib = lila.load()
    mt, nx = ctypes.c_int(512), ctypes.c_int(32)
    say("autotune rejects")
    assert lib.lila_autotune(100, 200, 1, ctypes.byref(mt), ctypes.byref(nx), None, 0) == -1
    assert lib.lila_autotune(100, 50, 1, None, ctypes.byref(nx), None, 0) == -1
    bad = ctypes.c_int(0)
    assert lib.lila_autotune(100, 50, 1, ctypes.byref(bad), ctypes.byref(nx), None, 0) == -1
    assert mt.value == 512 and nx.value == 32
