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
    assert mt in (256, 512, 1024) and nx in (32, 64, 128, 256)
    assert report.startswith("start mt 256 nx 32") and "chose mt %d nx %d" % (mt, nx) in report
    for (m, n, ii) in ((1500, 700, 0), (1213, 999, 13)):
        say("recursive after autotune", m, n, ii, mt, nx)
        A, o, g = run_driver(lila, ora, "recursive", m, n, ii, mt, [0, 1, 1, nx])
        check_vrtq(ora, A, o, g, m, n, ii, mt, m, m)


def test_autotune_rejects(lila):
    lib = lila.load()
    mt, nx = ctypes.c_int(512), ctypes.c_int(32)
    say("autotune rejects")
    assert lib.lila_autotune(100, 200, 1, ctypes.byref(mt), ctypes.byref(nx), None, 0) == -1
    assert lib.lila_autotune(100, 50, 1, None, ctypes.byref(nx), None, 0) == -1
    bad = ctypes.c_int(0)
    assert lib.lila_autotune(100, 50, 1, ctypes.byref(bad), ctypes.byref(nx), None, 0) == -1
    assert mt.value == 512 and nx.value == 32
