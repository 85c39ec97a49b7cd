"""The oracle's Butterworth biquad cascades (filter/iir.nim) against the
reference's own tests (tests/tiir.nim): the design (orders, the −6 dB zero-
phase and −3 dB minimum-phase corners, an odd order, a band-pass equal to
the spec's response), the refusals, the impulse response's spectrum against
the spec's response and against the FFT filter, causality and symmetry, the
steady-state start under DC and drift, a high cut keeping the trend, live
lengths and NaN through the pipeline, the trace ends against a long
context, the flush of a decaying state to exact zeros, and a zoom window
with requiredMargin.  Every bound is the reference test's."""
import math

import numpy as np
import pytest

import oracle as O
from processseg_amd import pipeline as PL

DT = 2.0
NS = 4000


def db(g):
    return 20 * math.log10(g)


def ricker(n, dt, t_ms, f_hz, amp=1.0):
    t = (np.arange(n) * dt - t_ms) / 1000
    a = (math.pi * f_hz * t) ** 2
    return amp * (1 - 2 * a) * np.exp(-a)


def clean(n=2001):
    return ricker(n, DT, 900, 25) + ricker(n, DT, 2400, 15, amp=-0.6)


def drift(n, dt, slope_per_s, start=0.0):
    return start + slope_per_s * np.arange(n) * dt / 1000


def test_design_and_corners():
    """tiir.nim:49-78."""
    lp = O.IirPlan(PL.butterHighCut(60, 24), DT, 100)
    assert lp.order == 2
    assert abs(db(lp.magnitudeAt(60)) - db(0.5)) <= 1e-9, "zero phase: -6 dB"
    mp = O.IirPlan(PL.butterHighCut(60, 24), DT, 100, phase=PL.ipMinimum)
    assert mp.order == 4
    assert abs(db(mp.magnitudeAt(60)) + 3.0103) <= 1e-4, "minimum phase: -3 dB"
    assert abs(mp.magnitudeAt(0) - 1) <= 1e-12 and mp.magnitudeAt(250) < 1e-12
    odd = O.IirPlan(PL.butterLowCut(4, 18), DT, 100, phase=PL.ipMinimum)
    assert odd.order == 3 and abs(db(odd.magnitudeAt(4)) + 3.0103) <= 1e-4
    assert odd.magnitudeAt(0) < 1e-12 and abs(odd.magnitudeAt(250) - 1) <= 1e-12
    spec = PL.butterBandPass(4, 60, 24, 48)
    bp = O.IirPlan(spec, DT, 100)
    assert bp.order == 6
    for f in (0.5, 2.0, 4.0, 10.0, 30.0, 60.0, 90.0, 200.0):
        assert abs(bp.magnitudeAt(f) - O.response(spec, f, DT)) <= 1e-9, f
    for bad in (lambda: O.IirPlan(PL.lowCut(1, 2), DT, 100),
                lambda: O.IirPlan(PL.butterLowCut(4).withNotch(50, 2), DT, 100),
                lambda: O.IirPlan(PL.butterLowCut(4).withZeroDc(), DT, 100),
                lambda: O.IirPlan(PL.butterLowCut(4, 18), DT, 100),
                lambda: O.IirPlan(PL.butterLowCut(4, 20), DT, 100, phase=PL.ipMinimum),
                lambda: O.IirPlan(PL.butterLowCut(4), DT, 0)):
        with pytest.raises(ValueError):
            bad()
    with pytest.raises(ValueError):
        O.IirPlan(PL.FilterSpec(kind=1, low_hz=0.0, high_hz=300.0, high_slope=24.0), DT, 100)


def spectrum(h):
    return np.abs(np.fft.rfft(h))


def test_matches_frequency_domain_magnitude():
    """tiir.nim:80-112: the zero-phase impulse response has the spec's
    response (as the FFT filter applies it); minimum phase |H| of slope/6
    poles, the square root of the response of twice the slope."""
    df = 1000 / (NS * DT)
    f = np.arange(NS // 2 + 1) * df
    for s in (PL.butterBandPass(4, 60, 24), PL.butterLowCut(2, 24), PL.butterHighCut(80, 48),
              PL.butterBandPass(8, 100, 48, 24)):
        iir = O.IirPlan(s, DT, NS, trend=PL.trNone)
        x = np.zeros(NS)
        x[NS // 2] = 1
        h = iir.filtered(x)
        want = np.array([O.response(s, v, DT) for v in f])
        assert np.max(np.abs(spectrum(h) - want)) < 1e-4, (s.low_hz, s.high_hz)
        fd = O.filter_trace(x, s, DT, NS, trend=PL.trNone)
        assert np.max(np.abs(h - fd)) < 1e-4 * np.max(np.abs(fd))
    for s, s2 in ((PL.butterBandPass(4, 60, 24), PL.butterBandPass(4, 60, 48)),
                  (PL.butterLowCut(3, 18), PL.butterLowCut(3, 36))):
        iir = O.IirPlan(s, DT, NS, phase=PL.ipMinimum, trend=PL.trNone)
        x = np.zeros(NS)
        x[iir.requiredMargin + 1] = 1
        mag = spectrum(iir.filtered(x))
        want = np.sqrt([O.response(s2, v, DT) for v in f])
        own = np.array([iir.magnitudeAt(v) for v in f])
        assert np.max(np.abs(mag - want)) < 1e-4 and np.max(np.abs(mag - own)) < 1e-4


def test_causal_and_symmetric():
    """tiir.nim:114-125."""
    x = np.zeros(2001)
    x[1000] = 1
    s = PL.butterBandPass(4, 60, 24)
    m = O.IirPlan(s, DT, 2001, phase=PL.ipMinimum, trend=PL.trNone).filtered(x)
    assert np.max(np.abs(m[:1000])) == 0, "nothing before the spike"
    assert m[1000] > 0
    z = O.IirPlan(s, DT, 2001, trend=PL.trNone).filtered(x)
    k = np.arange(1, 1001)
    assert np.max(np.abs(z[1000 + k] - z[1000 - k])) < 1e-12


def test_steady_state_start():
    """tiir.nim:127-155."""
    c = clean()
    for phase in (PL.ipZero, PL.ipMinimum):
        for tr in (PL.trNone, PL.trMean, PL.trLinear):
            plan = O.IirPlan(PL.butterLowCut(2, 24), DT, 2001, phase, tr)
            ref = plan.filtered(c)
            e = np.max(np.abs(plan.filtered(c + 1000) - ref)[:200])
            assert e < 1e-6, (phase, tr, e)
            d = np.max(np.abs(plan.filtered(c + drift(2001, DT, 20, 5)) - ref)[:200])
            if tr == PL.trLinear:
                assert d < 1e-9, (phase, d)
            elif phase == PL.ipZero:
                assert d < 0.05, (tr, d)
    for phase in (PL.ipZero, PL.ipMinimum):
        plan = O.IirPlan(PL.butterHighCut(60), DT, 2001, phase)
        y = plan.filtered(c + drift(2001, DT, 3, 7))
        no = plan.filtered(c)
        t0, t1 = y[0] - no[0], y[-1] - no[-1]
        assert abs((t1 - t0) - 3 * 4.0) <= 1e-6, "the slope is kept"
        if phase == PL.ipZero:
            assert abs(t0 - 7) <= 1e-6
        else:
            assert 7 - 3 * 0.02 < t0 < 7, t0


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-6)])
@pytest.mark.parametrize("phase", [0, 1])
def test_rows_of_varying_live_length(dtype, tol, phase):
    """tiir.nim:157-190 through the pipeline: live lengths 2001, 1500, 700, 2,
    1, 0, 1999, a dead row and a failed one (row_len -1), each row as the
    single-trace form filters it, nothing past a live end touched."""
    rows = 19
    x = np.stack([clean() + (r - 9) + drift(2001, DT, r % 5) for r in range(rows)]).astype(dtype)
    rl = np.full(rows, 2001, np.int32)
    for i, n in enumerate((2001, 1500, 700, 2, 1, 0, 1999)):
        rl[3 * i] = n
    rl[4] = rl[7] = -1
    spec = PL.butterBandPass(3, 70, 24)
    y = O.pipeline_apply([PL.iirOp(spec, phase=phase)], x, rl, DT)
    plan = O.IirPlan(spec, DT, 2001, phase)
    for r in range(rows):
        n = rl[r]
        if n < 0:
            assert np.array_equal(y[r], x[r])
            continue
        want = plan.filtered(x[r, :n])
        assert np.max(np.abs(y[r, :n].astype(np.float64) - want), initial=0) <= tol, r
        assert np.array_equal(y[r, n:], x[r, n:])


def test_nan_and_shapes():
    """tiir.nim:192-209."""
    x = np.stack([np.sin(np.arange(400) + r) for r in range(3)]).astype(np.float32)
    x[1, 200] = np.nan
    y = O.pipeline_apply([PL.iirOp(PL.butterLowCut(4))], x, np.full(3, 400, np.int32), DT)
    assert np.isnan(y[1, 0])
    assert not np.isnan(y[0]).any() and not np.isnan(y[2]).any()
    with pytest.raises(ValueError):
        O.pipeline_apply([PL.iirOp(PL.butterLowCut(4))], x, np.full(3, 400, np.int32), DT,
                         samples=300)
    with pytest.raises(ValueError):
        O.IirPlan(PL.butterLowCut(4), DT, 300).filtered(np.zeros(301, np.float32))


def test_trace_ends_match_a_long_context():
    """tiir.nim:211-230."""
    n = 2001
    for spec, limit in ((PL.butterBandPass(4, 60, 24), 0.15), (PL.butterLowCut(3, 24), 0.08)):
        L = O.IirPlan(spec, DT, n).requiredMargin
        ctx = 6 * L
        err = sig = 0.0
        for seed in range(1, 21):
            long = np.random.default_rng(seed).standard_normal(n + 2 * ctx)
            truth = O.IirPlan(spec, DT, long.size).filtered(long)
            got = O.IirPlan(spec, DT, n).filtered(long[ctx:ctx + n])
            for i in range(50):
                for j in (i, n - 1 - i):
                    err += (got[j] - truth[ctx + j]) ** 2
                    sig += truth[ctx + j] ** 2
        assert math.sqrt(err / sig) < limit, (spec.low_hz, math.sqrt(err / sig))


def test_no_subnormal_state():
    """tiir.nim:232-254: a spike, then silence, at 0.25 ms through a 100 Hz
    high cut: the tail reaches exactly 0."""
    for phase in (PL.ipZero, PL.ipMinimum):
        x = np.zeros(40_000)
        x[10] = 1
        y = O.IirPlan(PL.butterHighCut(100, 24), 0.25, x.size, phase, PL.trNone).filtered(x)
        tiny = np.abs(y) < np.finfo(np.float64).tiny
        assert np.count_nonzero(tiny & (y != 0)) < 600
        assert y[30_000] == 0 and y[-1] == 0
    g = np.zeros((30, 40_000), np.float32)
    for r in range(30):
        g[r, 10 + r] = 1
    op = PL.iirOp(PL.butterHighCut(100, 24), phase=PL.ipMinimum, trend=PL.trNone)
    y = O.pipeline_apply([op], g, np.full(30, 40_000, np.int32), 0.25)
    assert np.all(y[:, 39_999] == 0)


def test_zoom_window_with_margin():
    """tiir.nim:256-280."""
    total = 4000
    x = (np.random.default_rng(3).standard_normal(total) + ricker(total, DT, 3000, 25, 5)
         + drift(total, DT, 4, 1))
    for s in (PL.butterLowCut(2, 24), PL.butterLowCut(2, 48), PL.butterBandPass(4, 60, 48)):
        for phase in (PL.ipZero, PL.ipMinimum):
            plan = O.IirPlan(s, DT, total, phase)
            whole = plan.filtered(x)
            a, b = 1500, 2300
            m = plan.requiredMargin
            lo = max(0, a - m)
            hi = min(total - 1, b + m) if phase == PL.ipZero else b
            win = plan.filtered(x[lo:hi + 1])
            e = np.max(np.abs(win[a - lo:b - lo + 1] - whole[a:b + 1]))
            assert e / np.max(np.abs(whole[a:b + 1])) < 0.01, (phase, m)
