"""Pins the oracle's AGC, statistics and normalization (gain/gain.nim,
gain/stats.nim) against brute force and sorting, as the reference's own tests
do (tgain.nim:105-299, tstats.nim:15-163)."""
import math

import numpy as np
import pytest


def ref_agc(x, h, mode, target, epsilon):
    """Direct window sums in float64 (tgain.nim:105-121)."""
    n = x.shape[0]
    eps = epsilon * math.sqrt(float(np.sum(x * x)) / n)
    y = np.empty(n)
    for i in range(n):
        lo, e = max(0, i - h), min(n, i + h + 1)
        w = x[lo:e]
        s = float(np.sum(w * w)) if mode == 0 else float(np.sum(np.abs(w)))
        level = math.sqrt(s / (e - lo)) if mode == 0 else s / (e - lo)
        y[i] = x[i] * (target / (level + eps))
    return y


@pytest.mark.parametrize("mode", [0, 1])
def test_agc_vs_brute_force(oracle, mode):
    rng = np.random.default_rng(5)
    window, dt = 500.0, 2.0
    h = 125
    for n in [1, 2, 3, 5, 7, 13, 14, 100, 124, 125, 126, 250, 251, 252, 253, 502, 503, 777, 2001]:
        x = rng.standard_normal(n) * np.exp(np.linspace(0, -7, n))  # a 60 dB decay
        want = ref_agc(x, h, mode, 1.0, 1e-9)
        got = oracle.agc(x, window, dt, mode)
        assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300)) <= 1e-12, n
        got32 = oracle.agc(x.astype(np.float32), window, dt, mode)
        rel = np.abs(got32 - want) / np.maximum(np.abs(want), 1e-30)
        assert np.max(rel) <= 1e-5, (n, np.max(rel))


def test_agc_window_sizes(oracle):
    """h = round(windowMs / (2 dt)), half away from zero (tgain.nim:160):
    500 ms at 2 ms is h = 125 (251 samples), 250 ms is h = 63."""
    import processseg_amd as P
    info = oracle.plan_info([P.agcOp(500)], np.float32, 2.0, 2001)
    assert info["half"] == [125] and info["margin"] == 125
    assert oracle.plan_info([P.agcOp(250)], np.float32, 2.0, 3001)["half"] == [63]
    assert oracle.plan_info([P.agcOp(4)], np.float32, 2.0, 10)["half"] == [1]
    for bad in [P.agcOp(1.9), P.agcOp(0), P.agcOp(math.nan), P.agcOp(100, target=-1),
                P.agcOp(100, epsilon=0)]:
        with pytest.raises(ValueError):
            oracle.plan_info([bad], np.float32, 2.0, 10)


def test_agc_flattens_a_decay_and_keeps_120db(oracle):
    """tgain.nim:170-212: a 60 dB decay flattened to within 2 %; the tail of a
    120 dB decay keeps float64 precision (no subtraction in the sums)."""
    n = 4001
    t = np.arange(n)
    x = np.sin(2 * np.pi * t / 20) * 10 ** (-3 * t / n)
    y = oracle.agc(x, 500, 2.0)
    env = [np.max(np.abs(y[i:i + 200])) for i in range(200, n - 400, 200)]
    assert max(env) / min(env) < 1.02 * (max(env) / max(env)) + 0.02 or \
        (max(env) - min(env)) / max(env) < 0.02
    z = np.sin(2 * np.pi * t / 20) * 10 ** (-6 * t / n)
    got = oracle.agc(z, 500, 2.0)
    want = ref_agc(z, 125, 0, 1.0, 1e-9)
    tail = slice(n - 600, n)
    assert np.max(np.abs(got[tail] - want[tail]) / np.abs(want[tail]).max()) <= 1e-12


def test_agc_rows_left_alone(oracle):
    """An all-zero row stays zero; a row whose gain would overflow float32 is
    left untouched (tgain.nim:276-299); a NaN row is left as it is."""
    z = np.zeros(300, np.float32)
    assert np.all(oracle.agc(z, 100, 2.0) == 0)
    tiny = (np.random.default_rng(1).standard_normal(300) * 1e-33).astype(np.float32)
    assert np.array_equal(oracle.agc(tiny, 100, 2.0), tiny)
    nanrow = np.ones(300, np.float32)
    nanrow[5] = np.nan
    assert np.array_equal(oracle.agc(nanrow, 100, 2.0), nanrow, equal_nan=True)


def sort_percentile(v, p):
    s = np.sort(v)
    h = (len(s) - 1) * p / 100
    k = int(math.floor(h))
    if k + 1 >= len(s):
        return s[-1]
    return s[k] + (h - k) * (s[k + 1] - s[k])


def test_percentile_exact(oracle):
    """percentileAbs vs a sort and numpy's linear percentile (tstats.nim:15)."""
    rng = np.random.default_rng(11)
    for n in [1, 2, 3, 10, 101, 2001]:
        for p in [0, 1, 37.5, 50, 99, 99.9, 100]:
            x = rng.standard_normal(n)
            x[rng.integers(0, n, n // 3)] = 0.25  # ties
            want = sort_percentile(np.abs(x), p)
            assert oracle.percentile_abs(x, p) == pytest.approx(want, rel=1e-15, abs=0)
            assert abs(np.percentile(np.abs(x), p) - want) <= 1e-13 * max(1.0, want)
            x32 = x.astype(np.float32)
            assert oracle.percentile_abs(x32, p) == sort_percentile(np.abs(x32.astype(np.float64)), p)
    assert oracle.percentile_abs(np.zeros(0), 50) == 0
    y = np.ones(5)
    y[2] = np.nan
    assert math.isnan(oracle.percentile_abs(y, 50))


def test_panel_stats(oracle):
    """panelStats (tstats.nim:45-163): rms and max exact, the clip within one
    histogram bin (2^-8 relative) of a sort; NaN rows and dead rows left out."""
    rng = np.random.default_rng(2)
    rows, cols = 40, 500
    x = rng.standard_normal((rows, cols)) * np.exp(rng.uniform(-5, 5, (rows, 1)))
    row_len = rng.integers(0, cols + 1, rows).astype(np.int32)
    row_len[3] = -1
    x[7, 4] = np.nan
    used = [r for r in range(rows) if row_len[r] >= 0 and r != 7]
    vals = np.concatenate([x[r, :row_len[r]] for r in used])
    for p in (50, 90, 99, 99.9):
        s = oracle.panel_stats(x, row_len, p)
        assert s["rows"] == len(used) and s["samples"] == vals.size
        assert s["rms"] == pytest.approx(math.sqrt(np.sum(vals ** 2) / vals.size), rel=1e-13)
        assert s["max_abs"] == np.max(np.abs(vals))
        want = sort_percentile(np.abs(vals.astype(np.float32).astype(np.float64)), p)
        assert abs(s["clip"] - want) <= want / 256
    assert oracle.panel_stats(x, row_len, 100)["clip"] == np.max(np.abs(vals))
