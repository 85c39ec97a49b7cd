"""Pins the oracle's (and the library's host) filter design against the
reference's known answers and closed forms (tfilterdesign.nim:23-144,
tfilterapply.nim:35-60, bench/README.md:116-117)."""
import math

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import Corners, bandPass, butterBandPass, butterHighCut, butterLowCut, highCut, lowCut

DT = 2.0
FINE = 0.01


def db(g):
    return 20 * math.log10(g) if g > 0 else -math.inf


def near(a, b, tol):
    return abs(a - b) <= tol


@pytest.fixture(params=["oracle", "library"])
def resp(request, oracle):
    """The response function under test: the oracle's or the library's."""
    if request.param == "oracle":
        return lambda spec, f, dt: oracle.response(spec, f, dt)
    return lambda spec, f, dt: spec.response(f, dt)


@pytest.fixture(params=["oracle", "library"])
def impulse(request, oracle):
    if request.param == "oracle":
        return lambda spec, dt: oracle.impulse_length(spec, dt)
    return lambda spec, dt: spec.impulseLength(dt)


def sweep(resp, spec, lo, hi, dt=DT, steps=2000):
    g = [resp(spec, lo + (hi - lo) * i / steps, dt) for i in range(steps + 1)]
    return min(g), max(g)


def test_trapezoid(resp):
    bp = bandPass(Corners(4, 8, 60, 80))
    assert sweep(resp, bp, 8, 60) == (1.0, 1.0), "no passband ripple"
    assert sweep(resp, bp, 0, 4)[1] == 0 and sweep(resp, bp, 80, 250)[1] == 0
    assert near(db(resp(bp, 6, DT)), db(0.5), 1e-9)
    assert near(db(resp(bp, 70, DT)), db(0.5), 1e-9)
    prev = 0.0
    for i in range(401):
        g = resp(bp, 4 + 4 * i / 400, DT)
        assert g >= prev
        prev = g
    lc = lowCut(0.5, 1)
    assert resp(lc, 0, DT) == 0 and near(resp(lc, 0.75, DT), 0.5, 1e-15) and resp(lc, 400, DT) == 1
    hc = highCut(60, 80)
    assert resp(hc, 0, DT) == 1 and near(resp(hc, 70, DT), 0.5, 1e-15) and resp(hc, 80, DT) == 0


def test_butterworth(resp):
    hc = butterHighCut(50, 24)
    assert near(db(resp(hc, 50, FINE)), db(0.5), 1e-9)
    assert db(sweep(resp, hc, 0, 12.5, FINE)[0]) >= -0.0339
    a1, a2 = db(resp(hc, 400, FINE)), db(resp(hc, 800, FINE))
    assert near(a1 - a2, 24.08, 0.01)
    lc = butterLowCut(4, 36)
    assert resp(lc, 0, FINE) == 0 and near(db(resp(lc, 4, FINE)), db(0.5), 1e-9)
    assert near(db(resp(lc, 1, FINE)) - db(resp(lc, 0.5, FINE)), 36.12, 0.01)
    bp = butterBandPass(5, 80, 24, 48)
    assert near(db(resp(bp, 5, FINE)), -6.02, 0.05) and near(db(resp(bp, 80, FINE)), -6.02, 0.05)
    assert resp(bp, 20, FINE) > 0.99
    h2 = butterHighCut(60, 24)
    assert near(db(resp(h2, 60, DT)), db(0.5), 1e-9)
    assert resp(h2, 250, DT) == 0
    w = math.tan(math.pi * 100 * 0.002) / math.tan(math.pi * 60 * 0.002)
    assert near(resp(h2, 100, DT), 1 / (1 + w ** 4), 1e-12)


def test_notch(resp):
    s = bandPass(Corners(2, 4, 100, 120)).withNotch(50, 4)
    assert resp(s, 50, DT) == 0 and near(db(resp(s, 49, DT)), db(0.5), 1e-9)
    assert resp(s, 48, DT) == 1 and resp(s, 52, DT) == 1


def test_validation():
    for c in [Corners(8, 4, 60, 80), Corners(4, 4, 60, 80), Corners(4, 8, 80, 80),
              Corners(4, 70, 60, 80), Corners(math.nan, 8, 60, 80), Corners(-1, 8, 60, 80),
              Corners(1, 0, 60, 80), Corners(4, 8, math.inf, 80)]:
        with pytest.raises(ValueError):
            bandPass(c)
    for lo, hi, s in [(5.0, 5.0, 24.0), (10.0, 5.0, 24.0), (5.0, 50.0, 0.0), (5.0, 50.0, -6.0),
                      (5.0, 50.0, math.inf), (-1.0, 50.0, 24.0), (math.nan, 50.0, 24.0)]:
        with pytest.raises(ValueError):
            butterBandPass(lo, hi, s)
    with pytest.raises(ValueError):
        butterLowCut(0)
    with pytest.raises(ValueError):
        butterHighCut(math.inf)
    with pytest.raises(ValueError):
        lowCut(1, 2).withNotch(0, 1)
    with pytest.raises(ValueError):
        lowCut(1, 2).withNotch(50, 0)
    # against Nyquist (250 Hz at 2 ms): validate runs inside impulseLength
    bandPass(Corners(4, 8, 200, 250)).impulseLength(2)
    for spec in [bandPass(Corners(4, 8, 200, 251)), lowCut(200, 250), butterHighCut(250),
                 lowCut(1, 2).withNotch(250, 2)]:
        with pytest.raises(ValueError):
            spec.impulseLength(2)


def test_impulse_length_known_answers(oracle, impulse):
    """bench/README.md:116-117: 4-8-60-80 Hz at 2 ms is L = 110 (nfft 2250 on
    2001 samples); a 0.5-1 Hz low cut is L = 881 (nfft 3840)."""
    assert impulse(bandPass(Corners(4, 8, 60, 80)), 2.0) == 110
    assert impulse(lowCut(0.5, 1), 2.0) == 881
    i1 = oracle.filter_info(bandPass(Corners(4, 8, 60, 80)), 2.0, 2001)
    assert (i1["margin"], i1["pad"], i1["nfft"], i1["reflect"]) == (110, 110, 2250, False)
    i2 = oracle.filter_info(lowCut(0.5, 1), 2.0, 2001)
    assert (i2["margin"], i2["nfft"]) == (881, 3840)


def test_impulse_length_relations(impulse):
    """tfilterdesign.nim:113-144."""
    a = impulse(lowCut(1, 2), 2)
    b = impulse(lowCut(1.5, 2), 2)
    assert a > 0 and near(b / a, 2, 0.25)
    assert 100 < a < 1500
    c = impulse(lowCut(1, 2), 1)
    assert near(c / a, 2, 0.05)
    assert impulse(bandPass(Corners(0, 0, math.inf, math.inf)), 2) == 0
    assert impulse(lowCut(1, 2).withZeroDc(), 2) == a
    assert impulse(highCut(60, 80), 2) < a // 10
    bw = impulse(butterLowCut(2, 24), 2)
    assert 0 < bw < 2000
    assert impulse(butterLowCut(2, 48), 2) > bw
    assert impulse(highCut(100, 120).withNotch(50, 1), 2) > impulse(highCut(100, 120), 2)
    assert impulse(bandPass(Corners(0.5, 1, 60, 80)), 2) == impulse(lowCut(0.5, 1), 2)
    assert impulse(butterBandPass(4, 60, 48), 2) == impulse(butterLowCut(4, 48), 2)
    notch_only = bandPass(Corners(0, 0, math.inf, math.inf)).withNotch(50, 1)
    assert impulse(highCut(100, 120).withNotch(50, 1), 2) == impulse(notch_only, 2)
    assert impulse(notch_only, 2) > 500
    with pytest.raises(ValueError):
        impulse(lowCut(1, 1.00001), 0.01)


def test_impulse_length_library_equals_oracle(oracle):
    """The library's host plan compiler makes the same integer decisions."""
    specs = [bandPass(Corners(4, 8, 60, 80)), bandPass(Corners(5, 10, 70, 90)).withNotch(50, 4),
             butterHighCut(80, 24), lowCut(0.5, 1), highCut(60, 80), butterLowCut(3, 36),
             lowCut(2, 6).withNotch(50, 2).withNotch(100, 3), butterBandPass(6, 90, 24, 48)]
    for spec in specs:
        for dt in (1.0, 2.0, 4.0):
            if spec.kind == 1 and spec.high_hz < math.inf and spec.high_hz >= 500 / dt:
                continue
            if spec.f4 < math.inf and spec.f4 > 500 / dt:
                continue
            assert spec.impulseLength(dt) == oracle.impulse_length(spec, dt), (spec, dt)


def test_filter_plan_shape(oracle):
    """tfilterapply.nim:35-60: nfft >= ns + 2L (zeros) or ns + 4L (reflect);
    bins no wider than half the narrowest ramp."""
    for spec, samples in [(bandPass(Corners(4, 8, 60, 80)), 2001), (butterHighCut(80), 3001),
                          (lowCut(0.02, 0.04), 2001), (highCut(60, 80), 500)]:
        info = oracle.filter_info(spec, 2.0, samples)
        L = info["margin"]
        if info["reflect"]:
            assert info["nfft"] >= samples + 4 * L
        else:
            assert info["nfft"] >= samples + 2 * min(L, samples)
        if info["pad"] == L:
            assert 1000 / (info["nfft"] * 2.0) <= oracle.narrowest_ramp_hz(spec) / 2 + 1e-12
        assert oracle.next_fast_len(info["nfft"]) == info["nfft"]
    # an impulse longer than the trace is cut to it in zero pads
    info = oracle.filter_info(lowCut(0.02, 0.04), 2.0, 2001)
    assert info["margin"] > 2001 and info["pad"] == 2001
