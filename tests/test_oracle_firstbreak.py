"""The oracle's first-break picker (qc/firstbreak.nim restated) against the
reference's own checks (tests/tfirstbreak.nim): picks within a tolerance of
the onset for every method, confidence falling with the SNR, guided and
fixed search windows, the no-pick cases, a window of the block, zeros before
the data, a windowed pick reading only its stretch and the zoom window with
the required margin, the refusals.  The shots are this build's own
synthetic generator (tests/fbshots.py): the reference's test generator is
not portable."""
import math

import numpy as np
import pytest

import oracle as O
from processseg_amd.firstbreak import FirstBreakParams, pmCoppens, pmEnergyRatio, pmStaLta

from fbshots import COLS, DT, T0, VEL, keys_of, onset_ms, shot, split_spread

GUIDED = dict(guideVelocity=VEL, guideT0Ms=T0, beforeMs=60, afterMs=80)
METHODS = (pmStaLta, pmEnergyRatio, pmCoppens)


def picks(x, offsets, params, dtype=np.float32, row_len=None, t0=0.0, first=0, cols=COLS):
    plan = O.fb_plan(params, DT, cols)
    rl = np.full(x.shape[0], x.shape[1], np.int32) if row_len is None else row_len
    return O.fb_pick(plan, x.astype(dtype), rl, keys_of(offsets, t0), first)


def test_picks_within_a_tolerance():
    """tfirstbreak.nim:50-97: at high SNR every method picks every trace
    within its tolerance, confidently; unguided Coppens picks the first
    arrival; with noise a 60 Hz minimum-phase high cut helps every method."""
    x, off = shot(0.02, 1)
    on = onset_ms(off)
    for m in METHODS:
        t, c, snr, conf = picks(x, off, FirstBreakParams(pickMethod=m, **GUIDED))
        lag = t - on
        bias = lag.mean()
        spread = np.max(np.abs(lag - bias))
        tol = 2 * DT if m == pmCoppens else 5 * DT
        assert abs(bias) + spread <= tol, (m, lag)
        assert spread <= 1.5 * DT, (m, lag)
        assert np.all(conf > 0.9) and np.all(snr > 20), (m, conf, snr)
    t, *_ = picks(x, off, FirstBreakParams(pickMethod=pmCoppens))
    assert np.max(np.abs(t - on)) <= 2 * DT
    for m in METHODS:
        counts = []
        for hc in (60.0, math.inf):
            good = 0
            for seed in (2, 20, 200):
                n, off2 = shot(0.3, seed)
                p = FirstBreakParams(pickMethod=m, highCutHz=hc,
                                     windowMs=40 if m == pmEnergyRatio else 20, **GUIDED)
                t, *_ = picks(n, off2, p)
                lag = t - onset_ms(off2)
                med = np.sort(lag)[lag.size // 2]
                assert abs(med) <= 8 * DT, (m, hc, lag)
                g = int(np.sum(np.abs(lag - med) <= 3 * DT))
                if hc < math.inf:
                    assert g >= lag.size * 85 // 100, (m, lag)
                good += g
            counts.append(good)
        assert counts[0] > counts[1], (m, counts)


def test_confidence_against_snr():
    """tfirstbreak.nim:99-122: mean confidence falls with the noise; noise
    alone reads about 0 dB and is not confident."""
    last = 2.0
    for sigma in (0.02, 0.3, 1.0):
        x, off = shot(sigma, 10)
        _, _, _, conf = picks(x, off, FirstBreakParams(**GUIDED))
        assert conf.mean() < last
        last = conf.mean()
    assert last < 0.5
    off = split_spread(24, 25, 25)
    noise = np.stack([np.random.default_rng(100 + r).normal(0, 1.0, COLS)
                      for r in range(off.size)])
    for m in METHODS:
        _, _, snr, conf = picks(noise, off, FirstBreakParams(pickMethod=m, **GUIDED))
        assert np.all(np.isnan(snr) | (np.abs(snr) < 8)), snr
        assert conf.mean() < 0.1


def test_guided_and_fixed_windows():
    """tfirstbreak.nim:124-144: a burst before the first break is picked
    without a guide and avoided with one or with a fixed window after it."""
    from fbshots import add_onset
    off = split_spread(24, 25, 200)
    x, _ = shot(0.02, 3, off, dtype=np.float64)
    for r in range(x.shape[0]):
        add_onset(x[r], DT, 24, f_hz=40, amp=3, decay_ms=5)
    on = onset_ms(off)
    t, *_ = picks(x, off, FirstBreakParams(pickMethod=pmCoppens))
    assert np.sum(t < 32) > x.shape[0] // 2
    t, *_ = picks(x, off, FirstBreakParams(**GUIDED))
    assert np.max(np.abs(t - on)) <= 2 * DT
    t, *_ = picks(x, off, FirstBreakParams(pickMethod=pmCoppens, searchStartMs=60))
    assert np.max(np.abs(t - on)) <= 2 * DT
    t, *_ = picks(x, off, FirstBreakParams(searchStartMs=600, searchEndMs=700))
    assert np.all((t >= 600) & (t <= 700))


def test_no_picks_and_windows():
    """tfirstbreak.nim:146-194: a dead row, a NaN offset under a guide, an
    all-zero row and a NaN t0 give no pick (a NaN t0 unguided: a column, no
    time); a window of the block (firstSample 10) picks the same times and
    its columns are the block's; a search window past the trace picks
    nothing."""
    x, off = shot(0.02, 4)
    off = off.copy()
    off[1] = np.nan
    x[2] = 0
    keys = keys_of(off)
    keys[3, 0] = np.nan
    rl = np.full(x.shape[0], COLS, np.int32)
    rl[0] = -1
    plan = O.fb_plan(FirstBreakParams(**GUIDED), DT, COLS)
    t, c, snr, conf = O.fb_pick(plan, x, rl, keys)
    for r in range(4):
        assert np.isnan(t[r]) and c[r] == -1 and conf[r] == 0
    plan = O.fb_plan(FirstBreakParams(pickMethod=pmCoppens), DT, COLS)
    t, c, *_ = O.fb_pick(plan, x, rl, keys)
    assert not np.isnan(t[1]) and np.isnan(t[2]) and np.isnan(t[3])
    assert abs(c[3] * DT - onset_ms(off[3])) <= 2 * DT
    w = np.ascontiguousarray(x[:, 10:410])
    tw, cw, *_ = O.fb_pick(plan, w, np.minimum(rl, 400), keys, 10)
    assert np.isnan(tw[3]) and cw[3] == c[3] - 10
    g, go = shot(0.02, 6)
    full = picks(g, go, FirstBreakParams(**GUIDED))
    plan = O.fb_plan(FirstBreakParams(**GUIDED), DT, COLS)
    tw, cw, *_ = O.fb_pick(plan, np.ascontiguousarray(g[:, 10:410]),
                           np.full(g.shape[0], 400, np.int32), keys_of(go), 10)
    assert np.all(np.abs(tw - full[0]) <= DT)
    assert np.all(cw * DT == tw - 10 * DT)
    t, *_ = picks(*shot(0.02, 5), FirstBreakParams(searchStartMs=5000))
    assert np.all(np.isnan(t))


def test_zeros_before_the_data_are_not_noise():
    """tfirstbreak.nim:362-395: 100 zeros then noise and an onset at 400 ms:
    the step from zeros is no break; zeros inside a trace give no SNR."""
    from fbshots import add_onset
    rows, edge = 40, 100
    off = np.arange(rows) * 25.0 + 25
    x = np.zeros((rows, COLS))
    for r in range(rows):
        row = np.zeros(COLS)
        add_onset(row, DT, 400)
        row += np.random.default_rng(900 + r).normal(0, 0.02, COLS)
        x[r, edge:] = row[edge:]
    for m in METHODS:
        for hc in (math.inf, 60.0):
            t, c, snr, conf = picks(x, off, FirstBreakParams(pickMethod=m, highCutHz=hc))
            assert np.all((c < 0) | (np.abs(c - edge) > 4) | (conf == 0))
            assert np.sum(np.abs(t - 400) <= 8 * DT) >= rows * 95 // 100, (m, hc)
        t, c, snr, conf = picks(x, off, FirstBreakParams(pickMethod=m, searchStartMs=196,
                                                          searchEndMs=204))
        assert np.all(c == -1) and np.all(conf == 0)
    d = np.stack([np.random.default_rng(950 + r).normal(0, 0.02, COLS) for r in range(4)])
    d[:, 60:200] = 0
    t, c, snr, conf = picks(d, np.arange(4) * 25.0, FirstBreakParams(searchStartMs=380,
                                                                     searchEndMs=420))
    assert np.all(c == 200) and np.all(np.isnan(snr)) and np.all(conf == 0)


def test_windowed_pick_reads_only_its_stretch_and_zoom_margin():
    """tfirstbreak.nim:397-461: a burst past what a guided pick reads changes
    nothing, bit for bit; a zoom window holding requiredMargin samples around
    every search window picks as the whole traces do; plans that read whole
    traces need all of them."""
    x, off = shot(0.1, 11)
    loud = x.copy()
    for r in range(x.shape[0]):
        first = int(math.ceil((onset_ms(off[r]) + 80 + 40) / DT)) + 1
        loud[r, first:] += np.random.default_rng(300 + r).normal(0, 50, COLS - first).astype(np.float32)
    for m in (pmStaLta, pmEnergyRatio):
        for lc, hc in ((0.0, math.inf), (8.0, 70.0)):
            p = FirstBreakParams(pickMethod=m, lowCutHz=lc, highCutHz=hc, **GUIDED)
            a, b = picks(loud, off, p), picks(x, off, p)
            for u, v in zip(a, b):
                assert np.array_equal(u, v, equal_nan=True), (m, lc)
    g, go = shot(0.1, 12, split_spread(12, 25, 1000))
    for m in (pmStaLta, pmEnergyRatio):
        for lc, hc in ((0.0, math.inf), (8.0, 70.0)):
            p = FirstBreakParams(pickMethod=m, lowCutHz=lc, highCutHz=hc, **GUIDED)
            plan = O.fb_plan(p, DT, COLS)
            margin = O.fb_required_margin(plan)
            full = O.fb_pick(plan, g, np.full(g.shape[0], COLS, np.int32), keys_of(go))
            a = int(math.floor((onset_ms(1000) - 60) / DT)) - margin
            b = int(math.ceil((onset_ms(1275) + 80) / DT)) + margin
            assert a > 0 and b < COLS - 1
            w = np.ascontiguousarray(g[:, a:b + 1])
            got = O.fb_pick(plan, w, np.full(g.shape[0], w.shape[1], np.int32), keys_of(go), a)
            assert np.all(full[1] >= 0)
            assert np.array_equal(got[1], full[1] - a)
            for u, v in ((got[0], full[0]), (got[2], full[2]), (got[3], full[3])):
                assert np.array_equal(u, v, equal_nan=True)
    for p in (FirstBreakParams(), FirstBreakParams(pickMethod=pmCoppens, guideVelocity=VEL)):
        assert O.fb_required_margin(O.fb_plan(p, DT, COLS)) == COLS


def test_float64_traces_pick_as_float32():
    """tfirstbreak.nim:242-260: float64 copies of float32 traces pick the same
    times with a filter (the SNR within 1e-3 dB)."""
    x, off = shot(0.02, 8)
    p = FirstBreakParams(highCutHz=60, **GUIDED)
    a = picks(x, off, p, np.float32)
    b = picks(x.astype(np.float64), off, p, np.float64)
    assert np.array_equal(a[0], b[0])
    assert np.max(np.abs(a[2] - b[2])) < 1e-3


def test_refusals():
    """tfirstbreak.nim:196-216."""
    bad = [dict(), dict(windowMs=0), dict(ltaMs=10), dict(lowDb=20, highDb=20),
           dict(searchStartMs=100, searchEndMs=50), dict(guideVelocity=-1),
           dict(stabilization=math.nan), dict(lowCutHz=5, highCutHz=300),
           dict(lowCutHz=60, highCutHz=10), dict(lowCutHz=5, slopeDbPerOctave=10)]
    for i, kw in enumerate(bad):
        with pytest.raises(ValueError):
            O.fb_plan(FirstBreakParams(**kw), 0.0 if i == 0 else DT, COLS)
    with pytest.raises(ValueError):
        O.fb_plan(FirstBreakParams(), DT, 0)


def test_library_refuses_as_the_oracle():
    """The C-ABI plan refuses the parameters the oracle refuses, with the
    same text (checked before any device memory is touched)."""
    import processseg_amd as P
    bad = [dict(windowMs=0), dict(ltaMs=10), dict(lowDb=20, highDb=20),
           dict(searchStartMs=100, searchEndMs=50), dict(guideVelocity=-1),
           dict(stabilization=math.nan), dict(lowCutHz=60, highCutHz=10),
           dict(pickMethod=3), dict(guideVelocity=VEL, beforeMs=-1),
           dict(lowDb=-math.inf)]
    for kw in bad:
        with pytest.raises(ValueError) as want:
            O.fb_plan(FirstBreakParams(**kw), DT, COLS)
        with pytest.raises(ValueError) as got:
            P.initFirstBreakPlan(np.float32, FirstBreakParams(**kw), DT, COLS)
        assert str(want.value).split(":")[-1].strip() in str(got.value), (kw, got.value)
    for dt, n in ((0.0, COLS), (DT, 0)):
        with pytest.raises(ValueError):
            P.initFirstBreakPlan(np.float32, FirstBreakParams(), dt, n)
