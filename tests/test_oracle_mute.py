"""The oracle's mutes against the reference's own mute tests
(tests/tmute.nim): curve times, the exact taper values of both kernels,
fractional and sharp cuts, the time axis (t0Ms and a window's firstSample),
liveness, NaN keys and infinite bounds — every value below is one the
reference's test asserts."""
import math

import numpy as np
import pytest

import oracle as O
from processseg_amd import pipeline as PL

NAN, INF = math.nan, math.inf


def near(a, b, tol=1e-6):
    return abs(a - b) <= tol


def t(curve, offset):
    return O.mute_time_ms(PL.muteTopOp(curve), offset)


def test_curves():
    """tmute.nim:15-27."""
    assert t(PL.timeMute(100), NAN) == 100
    lmo = PL.lmoMute(velocity=2000, t0Ms=10, padMs=-5)
    assert near(t(lmo, 1000), 505) and near(t(lmo, -1000), 505)
    assert math.isnan(t(lmo, NAN))
    tab = PL.tableMute([0.0, 100, 300], [10.0, 50, 250])
    assert t(tab, 0) == 10 and t(tab, 100) == 50
    assert near(t(tab, 50), 30) and near(t(tab, -200), 150)
    assert t(tab, 1000) == 250
    assert t(PL.tableMute([200.0], [80.0]), 0) == 80
    assert math.isnan(t(tab, NAN))
    # the Python mirror agrees
    for c in (lmo, tab, PL.timeMute(100)):
        for off in (0.0, 50.0, -200.0, 1000.0):
            assert near(c.muteTimeMs(off), t(c, off), 1e-12)


def test_validation():
    """tmute.nim:29-42."""
    for f in (lambda: PL.timeMute(NAN), lambda: PL.tableMute([], []),
              lambda: PL.tableMute([0.0, 1], [1.0]), lambda: PL.tableMute([1.0, 1], [1.0, 2]),
              lambda: PL.tableMute([-1.0], [1.0]), lambda: PL.tableMute([0.0], [INF]),
              lambda: PL.lmoMute(0), lambda: PL.lmoMute(-1500),
              lambda: PL.lmoMute(1500, padMs=NAN)):
        with pytest.raises(ValueError):
            f()
    x = np.ones((1, 10), np.float32)
    for op in (PL.muteTopOp(PL.timeMute(4), taperMs=-1),
               PL.muteBottomOp(PL.timeMute(4), taperMs=NAN)):
        with pytest.raises(ValueError, match="taper"):
            O.pipeline_apply([op], x, np.array([10], np.int32), 2.0)


def test_top_kernel_exact():
    """tmute.nim:44-66."""
    x = O.mute_row("top", np.ones(10), 6, 4)
    assert list(x[:3]) == [0, 0, 0]
    for j in range(3, 6):
        assert near(x[j], math.sin(math.pi / 2 * (j - 2) / 4) ** 2, 1e-12)
    assert list(x[6:]) == [1] * 4
    assert list(O.mute_row("top", np.ones(5), 2, 0)) == [0, 0, 1, 1, 1]
    frac = O.mute_row("top", np.ones(5), 2.5, 1)
    assert frac[1] == 0 and near(frac[2], math.sin(math.pi / 4) ** 2, 1e-12) and frac[3] == 1
    assert list(O.mute_row("top", np.ones(4), 99, 3)) == [0] * 4
    assert list(O.mute_row("top", np.ones(4), -10, 3)) == [1] * 4
    assert list(O.mute_row("top", np.ones(4), NAN, 3)) == [1] * 4


def test_bottom_kernel_exact():
    """tmute.nim:68-83."""
    x = O.mute_row("bottom", np.ones(10), 3, 4)
    assert list(x[:4]) == [1] * 4
    for j in range(4, 7):
        assert near(x[j], math.sin(math.pi / 2 * (7 - j) / 4) ** 2, 1e-12)
    assert list(x[7:]) == [0] * 3
    assert list(O.mute_row("bottom", np.ones(5), 2, 0)) == [1, 1, 1, 0, 0]
    assert list(O.mute_row("bottom", np.ones(4), -1, 0)) == [0] * 4
    assert list(O.mute_row("bottom", np.ones(4), 50, 2)) == [1] * 4


def test_taper_limit_and_infinities():
    """tmute.nim:108-117, 170-193."""
    for L in (0.0, 1e-3):
        assert list(O.mute_row("top", np.ones(6), 3, L)) == [0, 0, 0, 1, 1, 1]
        assert list(O.mute_row("bottom", np.ones(6), 2, L)) == [1, 1, 1, 0, 0, 0]
    one = np.ones(8, np.float32)
    assert np.all(O.mute_row("bottom", one, -INF, INF) == 0)
    assert np.all(O.mute_row("top", one, INF, INF) == 0)
    x = O.mute_row("top", one, 5, INF)
    assert np.all(x[:5] == 0) and np.all(x[5:] == 1)
    x = O.mute_row("bottom", one, 2, INF)
    assert np.all(x[:3] == 1) and np.all(x[3:] == 0)
    assert np.all(O.mute_row("top", one, -INF, 3) == 1)


def test_lmo_top_mute_on_a_gather():
    """tmute.nim:90-106: offsets 0, 400, 800 at 2000 m/s are columns 0, 50,
    100 at 4 ms; a 20 ms taper is the 5 columns before each."""
    x = np.ones((3, 150), np.float32)
    keys = np.array([[0.0, 0.0], [0.0, 400.0], [0.0, 800.0]])
    y = O.pipeline_apply([PL.muteTopOp(PL.lmoMute(velocity=2000), taperMs=20)], x,
                         np.full(3, 150, np.int32), 4.0, keys=keys)
    for r, col in enumerate((0, 50, 100)):
        assert np.all(y[r, :max(col - 5, 0)] == 0)
        for k in range(5):
            if col - 5 + k >= 0:
                assert near(float(y[r, col - 5 + k]), math.sin(math.pi / 2 * k / 5) ** 2)
        assert np.all(y[r, col:] == 1)


def test_time_axis_follows_t0_and_the_window():
    """tmute.nim:119-136."""
    x = np.ones((2, 40), np.float32)
    keys = np.array([[100.0, 0.0], [0.0, 0.0]])
    y = O.pipeline_apply([PL.muteTopOp(PL.timeMute(140))], x, np.full(2, 40, np.int32), 4.0,
                         keys=keys)
    assert y[0, 9] == 0 and y[0, 10] == 1 and y[1, 34] == 0 and y[1, 35] == 1
    w = np.ones((1, 30), np.float32)          # a window from trace sample 20
    k1 = np.array([[0.0, 0.0]])
    y = O.pipeline_apply([PL.muteTopOp(PL.timeMute(140))], w, np.array([30], np.int32), 4.0,
                         keys=k1, first_sample=20)
    assert y[0, 14] == 0 and y[0, 15] == 1
    y = O.pipeline_apply([PL.muteBottomOp(PL.timeMute(180))], y, np.array([30], np.int32), 4.0,
                         keys=k1, first_sample=20)
    assert y[0, 25] == 1 and y[0, 26] == 0


def test_table_bottom_mute_and_liveness():
    """tmute.nim:138-158: an auxiliary row (row_len -1), a NaN offset, a
    row live to 60 samples with a sentinel past it."""
    x = np.ones((4, 100), np.float32)
    x[3, 60:] = 7
    keys = np.array([[0.0, 100.0], [0.0, 200.0], [0.0, NAN], [0.0, 400.0]])
    rl = np.array([100, -1, 100, 60], np.int32)
    tab = PL.tableMute([100.0, 400], [200.0, 160])
    y = O.pipeline_apply([PL.muteBottomOp(tab)], x, rl, 4.0, keys=keys)
    assert y[0, 50] == 1 and y[0, 51] == 0
    assert y[1, 99] == 1, "auxiliary rows are untouched"
    assert y[2, 99] == 1, "a NaN offset leaves the row alone"
    assert y[3, 40] == 1 and y[3, 41] == 0 and y[3, 59] == 0
    assert np.all(y[3, 60:] == 7), "nothing past the live end is written"
    y = O.pipeline_apply([PL.muteTopOp(PL.timeMute(400))], y, rl, 4.0, keys=keys)
    assert y[3, 60] == 7


def test_nan_t0_leaves_the_row_alone():
    """tmute.nim:160-165."""
    x = np.ones((1, 10), np.float32)
    y = O.pipeline_apply([PL.muteTopOp(PL.timeMute(20))], x, np.array([10], np.int32), 4.0,
                         keys=np.array([[NAN, 100.0]]))
    assert y[0, 0] == 1


def test_float64_gather():
    """tmute.nim:167-179 (the values; allocation is not a property here)."""
    x = np.ones((4, 1000))
    keys = np.array([[0.0, o] for o in (0.0, 500, 1000, 1500)])
    ops = [PL.muteTopOp(PL.lmoMute(1800, padMs=-20), taperMs=30),
           PL.muteBottomOp(PL.tableMute([0.0, 1500], [1200.0, 1900]), taperMs=40)]
    y = O.pipeline_apply(ops, x, np.full(4, 1000, np.int32), 2.0, keys=keys)
    assert y[3, 0] == 0 and y[0, 999] == 0
