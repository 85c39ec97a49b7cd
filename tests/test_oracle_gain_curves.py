"""The oracle's time-variant gain (gain/curves.nim) against the reference's
own tests (tests/tcurves.nim): exact values at the knots, interpolation in
dB, flat extrapolation, the power and exponential curves, the table's extent
and refusals, each row's time axis (t0Ms on and off the grid, before and
beyond the table, a window's firstSample), NaN t0, liveness, the pipeline
step's rounding off the table, zeros under an infinite factor, and heldGain.
Every value and bound below is one the reference's test asserts; where that
test compares two forms of the reference with each other, the check here is
the exact arithmetic of the form the pipeline step uses."""
import math

import numpy as np
import pytest

import oracle as O
from processseg_amd import pipeline as PL

NAN, INF = math.nan, math.inf
KNOTS = [(100.0, 0.0), (1100.0, 20.0), (2100.0, 14.0)]


def near(a, b, rel):
    return abs(a - b) <= rel * max(abs(a), abs(b))


def db_to_factor(db):
    """dbToFactor (gain/gain.nim:122-124)."""
    return math.pow(10.0, db / 20)


def test_curve_values():
    """tcurves.nim:24-52, through the C restatement and the Python mirror."""
    c = PL.knotGain([PL.GainKnot(t, d) for t, d in KNOTS])
    for at in (O.gain_db_at, lambda cv, t: cv.gainDbAt(t)):
        assert at(c, 100) == 0 and at(c, 1100) == 20 and at(c, 2100) == 14
        assert at(c, 600) == 10, "linear in dB"
        assert at(c, 1600) == 17
        assert at(c, -50) == 0 and at(c, 5000) == 14, "flat outside"
        assert math.isnan(at(c, NAN))
    for fac in (O.gain_factor_at, lambda cv, t: cv.factorAt(t)):
        assert near(fac(c, 1100), 10, 1e-15)
        assert near(fac(c, 600), math.sqrt(10.0), 1e-15)
        assert math.isnan(fac(c, NAN))
        one = PL.knotGain([(0.0, 6.0)])
        assert O.gain_db_at(one, -1) == 6 and O.gain_db_at(one, 1e6) == 6
        p = PL.powerGain(2)
        assert fac(p, 1000) == 1 and near(fac(p, 2000), 4, 1e-15)
        assert near(fac(p, 500), 0.25, 1e-15)
        assert fac(p, 0) == fac(p, 1), "t below minMs is taken as minMs"
        e = PL.exponentialGain(6)
        assert near(fac(e, 2000), db_to_factor(12), 1e-15)
    p = PL.powerGain(2)
    assert near(O.gain_db_at(p, 2000), 20 * math.log10(4.0), 1e-12)
    e = PL.exponentialGain(6)
    assert O.gain_db_at(e, 0) == 0 and O.gain_db_at(e, 2000) == 12
    assert O.gain_db_at(e, -1000) == -6
    for bad in (lambda: PL.knotGain([]), lambda: PL.knotGain([(1.0, 0.0), (1.0, 3.0)]),
                lambda: PL.knotGain([(2.0, 0.0), (1.0, 3.0)]),
                lambda: PL.knotGain([(NAN, 0.0)]), lambda: PL.knotGain([(0.0, INF)]),
                lambda: PL.powerGain(NAN), lambda: PL.powerGain(2, refMs=0),
                lambda: PL.powerGain(2, minMs=-1), lambda: PL.exponentialGain(INF)):
        with pytest.raises(ValueError):
            bad()
    # the same curves built around the mirror's checks: the C plan refuses them
    for raw in (PL.GainCurve(PL.gcKnots, knots=((1.0, 0.0), (1.0, 3.0))),
                PL.GainCurve(PL.gcKnots, knots=((0.0, INF),)),
                PL.GainCurve(PL.gcPower, power=2, refMs=0),
                PL.GainCurve(PL.gcExponential, dbPerS=NAN)):
        with pytest.raises(ValueError):
            O.gain_curve_apply(raw, np.ones(4), 2.0, 4, 0.0)


def test_plan_table():
    """tcurves.nim:54-64."""
    x = np.ones(4)
    _, n = O.gain_curve_apply(PL.knotGain(KNOTS), x, 2.0, 1001, 0.0)
    assert n == 1001
    _, n = O.gain_curve_apply(PL.powerGain(2), x.astype(np.float32), 2.0, 1001, 0.0,
                              t_min_ms=-100, t_max_ms=3000)
    assert n == 1551
    p = PL.powerGain(2)
    for kw in (dict(dt_ms=0.0, samples=10), dict(dt_ms=2.0, samples=0),
               dict(dt_ms=2.0, samples=10, t_min_ms=10, t_max_ms=5),
               dict(dt_ms=2.0, samples=10, t_min_ms=INF)):
        with pytest.raises(ValueError):
            O.gain_curve_apply(p, np.ones(0, np.float32), start_ms=0.0, **kw)


def test_knots_on_the_grid():
    """tcurves.nim:70-79: t0 0, the samples at the knot times carry exactly
    the knot gains."""
    x, _ = O.gain_curve_apply(PL.knotGain(KNOTS), np.ones(1501), 2.0, 1501, 0.0)
    assert x[50] == db_to_factor(0) and x[550] == db_to_factor(20)
    assert x[1050] == db_to_factor(14)
    assert near(x[300], db_to_factor(10), 1e-15), "half way: 10 dB"
    assert x[0] == 1 and x[10] == 1, "flat before the first knot"
    assert x[1500] == db_to_factor(14), "flat after the last"


def test_t0_shifts():
    """tcurves.nim:81-140: each factor within the documented bound of the
    table's interpolation, on and off the grid, before and past the table."""
    c = PL.knotGain(KNOTS)
    for start in (100.0, 2.0, 1001.0, 0.7, -20.0, -3.3, 1500.0, 1999.0, 5000.0, -1e7):
        x, _ = O.gain_curve_apply(c, np.ones(1001), 2.0, 1001, start)
        for j in range(1001):
            want = c.factorAt(start + j * 2.0)
            assert near(x[j], want, 3e-6), f"{start} at {j}: {x[j]} vs {want}"
        if start == 100:
            assert x[0] == 1 and x[500] == 10, "on the grid: exact table values"
    # a power curve off the grid: (ΔdB·ln10/20)²/8, ΔdB over one sample
    p = PL.powerGain(2)
    x, _ = O.gain_curve_apply(p, np.ones(2001), 2.0, 2001, 51.3)
    for j in range(2001):
        t = 51.3 + j * 2.0
        d_db = p.gainDbAt(math.ceil(t / 2) * 2) - p.gainDbAt(math.floor(t / 2) * 2)
        bound = (d_db * math.log(10.0) / 20) ** 2 / 8
        assert abs(x[j] / p.factorAt(t) - 1) <= 1.01 * bound + 1e-12, j
    # a negative power: |p(p − 1)|·(dt/t)²/8 at the interval start
    inv = PL.powerGain(-1, refMs=100)
    y, _ = O.gain_curve_apply(inv, np.ones(1001), 2.0, 1001, 21.0)
    for j in range(1001):
        t0 = math.floor((21 + j * 2.0) / 2) * 2
        bound = 2.0 * (2 / t0) ** 2 / 8
        assert abs(y[j] / inv.factorAt(21 + j * 2.0) - 1) <= 1.01 * bound, j
    # a knot between grid points: first order there (|Δs|·dt/4), second
    # order elsewhere
    k = PL.knotGain([(0.0, 0.0), (1101.0, 20.0), (2000.0, 0.0)])
    z, _ = O.gain_curve_apply(k, np.ones(1001), 2.0, 1001, 1.0)
    s1 = math.log(10.0) / 20 * 20 / 1101
    s2 = math.log(10.0) / 20 * -20 / 899
    at_kink = 0.0
    for j in range(1001):
        t = 1 + j * 2.0
        err = abs(z[j] / k.factorAt(t) - 1)
        if abs(t - 1101) < 2:
            at_kink = max(at_kink, err)
        else:
            assert err <= (max(abs(s1), abs(s2)) * 2) ** 2 / 8 * 1.01, j
    assert 1e-4 < at_kink <= abs(s1 - s2) * 2 / 4 * 1.01, at_kink
    nan_t, _ = O.gain_curve_apply(p, np.ones(10), 2.0, 2001, NAN)
    assert np.all(nan_t == 1), "a NaN start time leaves the trace alone"


def test_pipeline_step_follows_each_rows_time():
    """tcurves.nim:142-172 through the pipeline: a window from trace sample
    50, rows at t0 0, 100, 51 (off the grid), NaN, a dead row, a row live to
    150 samples of the trace (100 of the window), and a block wider than the
    plan refused."""
    rows, cols = 6, 300
    x = np.ones((rows, cols), np.float32)
    x[5, 100:] = 0                      # the window past the trace's end
    t0 = np.array([0.0, 100, 51, NAN, 0, 0])
    keys = np.stack([t0, np.zeros(rows)], axis=1)
    rl = np.array([300, 300, 300, 300, -1, 100], np.int32)
    c = PL.exponentialGain(10)
    op = PL.gainCurveOp(c)
    y = O.pipeline_apply([op], x, rl, 2.0, keys=keys, first_sample=50)
    for r in (0, 1, 2):
        for j in range(cols):
            want = c.factorAt(t0[r] + (50 + j) * 2.0)
            assert near(float(y[r, j]), want, 2e-6), (r, j)
    assert np.all(y[3] == 1), "NaN t0: untouched"
    assert np.all(y[4] == 1), "dead: untouched"
    for j in range(100):
        assert near(float(y[5, j]), c.factorAt(100 + 2.0 * j), 2e-6)
    assert np.all(y[5, 100:] == 0), "padding untouched"
    with pytest.raises(ValueError):
        O.pipeline_apply([op], np.zeros((rows, 301), np.float32), np.full(rows, 301, np.int32),
                         2.0, samples=300, keys=keys, first_sample=50)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rounding_off_the_table(dtype):
    """tcurves.nim:174-214: a table of 100 samples (0 .. 198 ms) under rows
    that lie on, off and across it.  Off the table the pipeline step rounds
    the factor to the sample type, then multiplies; on it, the table's T
    values; rows not live or with a NaN t0 untouched."""
    rows, cols, first = 12, 250, 150
    x = np.array([[1 + (r + first + j) % 5 for j in range(cols)] for r in range(rows)], dtype)
    t0 = np.zeros(rows)
    t0[4], t0[5], t0[6], t0[7] = 300, 301, -90, NAN
    rl = np.full(rows, cols, np.int32)
    rl[8] = -1
    rl[9] = 120 - first if 120 > first else 0
    keys = np.stack([t0, np.zeros(rows)], axis=1)
    knots = PL.knotGain(KNOTS)
    for c in (PL.exponentialGain(10), PL.powerGain(2), knots):
        op = PL.gainCurveOp(c, tMaxMs=198)
        y = O.pipeline_apply([op], x, rl, 2.0, keys=keys, first_sample=first)
        table = np.array([c.factorAt(k * 2.0) for k in range(100)]).astype(dtype)
        for r in range(rows):
            n = rl[r]
            if n < 0 or math.isnan(t0[r]):
                assert np.array_equal(y[r], x[r]), f"row {r} untouched"
                continue
            for j in range(n):
                t = t0[r] + (first + j) * 2.0
                u = t / 2.0
                if t < 0 or u > 99:     # off the table: T(factor), then the multiply
                    want = x[r, j] * dtype(c.factorAt(t))
                    assert y[r, j] == want, (c.kind, r, j)
                elif u == round(u):     # on the grid: the table entry
                    assert y[r, j] == x[r, j] * table[int(u)], (c.kind, r, j)
            assert np.array_equal(y[r, n:], x[r, n:])


def test_zero_stays_zero_under_an_infinite_factor():
    """tcurves.nim:216-229 and 252-268."""
    z = np.zeros((2, 300), np.float32)
    z[1] = 1
    op = PL.gainCurveOp(PL.exponentialGain(2000), tMaxMs=100)
    y = O.pipeline_apply([op], z, np.full(2, 300, np.int32), 2.0)
    assert np.all(y[0] == 0), "a zero stays zero"
    assert y[1, 299] == INF
    steep = PL.knotGain([(0.0, 0.0), (1000.0, 800.0)])
    with pytest.raises(ValueError, match="overflows"):
        O.gain_curve_apply(steep, np.zeros(1, np.float32), 2.0, 501, 0.0)
    with pytest.raises(ValueError, match="overflows"):
        O.gain_curve_apply(PL.exponentialGain(800), np.zeros(1, np.float32), 2.0, 1001, 0.0)
    x, _ = O.gain_curve_apply(steep, np.zeros(501), 2.0, 501, 0.0)
    assert np.all(x == 0)
    y, _ = O.gain_curve_apply(PL.exponentialGain(6), np.zeros(100, np.float32), 2.0, 100, 1e9)
    assert np.all(y == 0)


def test_unbuilt_curve():
    """tcurves.nim:231-239 (the dt check is the pipeline's one dt)."""
    c = PL.GainCurve()
    assert math.isnan(c.gainDbAt(10)) and math.isnan(c.factorAt(10))
    assert math.isnan(O.gain_db_at(c, 10)) and math.isnan(O.gain_factor_at(c, 10))
    with pytest.raises(ValueError, match="no knots"):
        O.gain_curve_apply(c, np.ones(4, np.float32), 2.0, 10, 0.0)


def test_held_gain():
    """tcurves.nim:252-288 (heldGain is curve construction: the mirror)."""
    expc = PL.exponentialGain(6)
    held = PL.heldGain(expc, untilMs=2000)
    assert held.kind == PL.gcKnots
    assert held.gainDbAt(0) == 0
    assert near(held.gainDbAt(2000), 12, 1e-12), "6 dB/s over 2 s"
    assert held.gainDbAt(4000) == held.gainDbAt(2000), "flat after the hold"
    assert near(held.gainDbAt(1000), expc.gainDbAt(1000), 1e-12)
    p = PL.powerGain(2, refMs=1000, minMs=10)
    hp = PL.heldGain(p, untilMs=3000)
    worst, t = 0.0, 10.0
    while t <= 3000:
        worst = max(worst, abs(hp.gainDbAt(t) - p.gainDbAt(t)))
        t *= 1.005
    assert worst <= 0.001, worst
    assert hp.gainDbAt(5000) == hp.gainDbAt(3000)
    kg = PL.knotGain([(0.0, 0.0), (1000.0, 12.0), (4000.0, 24.0)])
    hk = PL.heldGain(kg, untilMs=2000)
    assert hk.gainDbAt(0) == 0 and near(hk.gainDbAt(1000), 12, 1e-12)
    assert near(hk.gainDbAt(2000), kg.gainDbAt(2000), 1e-12)
    assert hk.gainDbAt(3000) == hk.gainDbAt(2000)
    for bad in (lambda: PL.heldGain(expc, untilMs=0), lambda: PL.heldGain(expc, untilMs=NAN),
                lambda: PL.heldGain(expc, 100, step=1.0), lambda: PL.heldGain(PL.GainCurve(), 100)):
        with pytest.raises(ValueError):
            bad()
    # a knot before 0 is kept; only a first knot at or after 0 holds the gain
    # at 0 before 0
    kb = PL.knotGain([(-100.0, 0.0), (100.0, 10.0)])
    hb = PL.heldGain(kb, untilMs=300)
    assert near(hb.gainDbAt(-50), 2.5, 1e-12)
    assert near(hb.gainDbAt(0), 5, 1e-12)
    assert near(hb.gainDbAt(100), 10, 1e-12)
    assert near(hb.gainDbAt(300), kb.gainDbAt(300), 1e-12)
    assert hb.gainDbAt(400) == hb.gainDbAt(300)
    he = PL.heldGain(PL.exponentialGain(6), untilMs=1000)
    assert he.gainDbAt(-50) == he.gainDbAt(0) and he.gainDbAt(0) == 0
    # the held curve through the C restatement agrees with the mirror
    for t in (-50.0, 0.0, 37.5, 100.0, 250.0, 300.0, 999.0):
        assert O.gain_db_at(hb, t) == hb.gainDbAt(t)
        assert O.gain_db_at(hp, t) == hp.gainDbAt(t)


def test_table_span_edges():
    """Which samples take the table (spanOf, gain/curves.nim:340-355): on a
    table of 100 entries (0 .. 198 ms) under a power curve, whose chord
    differs from the curve everywhere by more than float64 rounding, a row
    off the grid interpolates exactly the samples with both neighbours on
    the table and takes the exact factor at every other sample, at both
    ends of the table."""
    p = PL.powerGain(2)
    for start, first_on, last_on in ((0.5, 0, 98), (-5.5, 3, 101)):
        y, n = O.gain_curve_apply(p, np.ones(250), 2.0, 250, start, t_max_ms=198)
        assert n == 100
        for j in range(250):
            t = start + 2.0 * j
            exact = p.factorAt(t)
            if first_on <= j <= last_on:
                lo, hi = p.factorAt(math.floor(t / 2) * 2), p.factorAt(math.ceil(t / 2) * 2)
                assert y[j] != exact and lo <= y[j] <= hi, (start, j)
            else:
                assert y[j] == exact, (start, j)
