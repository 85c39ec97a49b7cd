"""Pins the oracle's FFT filter (filter/apply.nim), correlation
(vibroseis/correlate.nim), sweep (vibroseis/sweep.nim, taper.nim) and
pipeline (pipeline.nim) against closed forms, independent numpy
computations, and the invariants of the reference's own tests
(tfilterapply.nim, tcorrelate.nim, tpipeline.nim)."""
import math

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import Corners, bandPass, butterHighCut, highCut, lowCut

DT = 2.0


def numpy_filter(oracle, x, spec, dt, samples, trend=2):
    """The D4 edge handling written with numpy's FFT and polyfit: remove the
    least-squares line, pad with zeros or cos²-tapered odd reflections,
    rfft * response, irfft, add the line back times the 0 Hz gain."""
    info = oracle.filter_info(spec, dt, samples, trend)
    n, p, nf = x.shape[0], info["pad"], info["nfft"]
    i = np.arange(n, dtype=np.float64)
    if trend == 0:
        a = b = 0.0
    elif trend == 1 or n == 1:
        a, b = float(np.mean(x)), 0.0
    else:
        b, a = np.polyfit(i, x, 1)
    buf = np.zeros(nf)
    buf[p:p + n] = x - (a + b * i)
    if info["reflect"]:
        k = np.arange(1, p + 1)
        fall = np.cos(np.pi / 2 * k / (p + 1)) ** 2
        m = np.minimum(k, n - 1)
        d0, dn = buf[p], buf[p + n - 1]
        buf[p - k] = (2 * d0 - buf[p + m]) * fall
        buf[p + n - 1 + k] = (2 * dn - buf[p + n - 1 - m]) * fall
    df = 1000 / (nf * dt)
    H = np.array([oracle.response(spec, kk * df, dt) for kk in range(nf // 2 + 1)])
    if spec.zero_dc:
        H[0] = 0
    y = np.fft.irfft(np.fft.rfft(buf) * H, nf)[p:p + n]
    dc = info["dc_gain"]
    gb = b * dc
    ga = -gb * (n - 1) / 2 if spec.zero_dc else a * dc
    return y + ga + gb * i


SPECS = [
    (bandPass(Corners(4, 8, 60, 80)), 2),
    (butterHighCut(80, 24), 2),
    (bandPass(Corners(5, 10, 70, 90)).withNotch(50, 4), 2),
    (lowCut(3, 6).withNotch(50, 4).withZeroDc(), 2),
    (bandPass(Corners(4, 8, 60, 80)), 1),
    (highCut(60, 80), 0),
]


@pytest.mark.parametrize("spec,trend", SPECS)
@pytest.mark.parametrize("n", [2001, 1500, 700, 3, 2, 1])
def test_filter_vs_numpy(oracle, spec, trend, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 0.3 + 0.001 * np.arange(n)
    got = oracle.filter_trace(x, spec, DT, 2001, trend)
    want = numpy_filter(oracle, x, spec, DT, 2001, trend)
    assert np.max(np.abs(got - want)) <= 1e-11 * max(1.0, np.max(np.abs(want)))


def test_filter_flat_changes_nothing(oracle):
    flat = bandPass(Corners(0, 0, math.inf, math.inf))
    x = np.random.default_rng(1).standard_normal(2001) + 5
    assert np.max(np.abs(oracle.filter_trace(x, flat, DT, 2001) - x)) <= 1e-12


def test_filter_trend_removed(oracle):
    """tfilterapply.nim:78-101: DC 1000 and a drift through a low cut give the
    clean trace's result."""
    rng = np.random.default_rng(2)
    clean = rng.standard_normal(2001)
    drift = clean + 1000 + 0.25 * np.arange(2001)
    spec = lowCut(0.5, 1)
    a = oracle.filter_trace(clean, spec, DT, 2001)
    b = oracle.filter_trace(drift, spec, DT, 2001)
    assert np.max(np.abs(a - b)) <= 1e-9 * 1000
    a32 = oracle.filter_trace(clean.astype(np.float32), spec, DT, 2001)
    assert np.max(np.abs(a32 - a)) <= 2e-4


def test_filter_linear_and_symmetric(oracle):
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(2001), rng.standard_normal(2001)
    spec = bandPass(Corners(4, 8, 60, 80))
    f = lambda v: oracle.filter_trace(v, spec, DT, 2001)  # noqa: E731
    assert np.max(np.abs(f(2 * x - 3 * y) - (2 * f(x) - 3 * f(y)))) <= 1e-12
    # a symmetric impulse response (tfilterapply.nim:157-175), trend none
    imp = np.zeros(2001)
    imp[1000] = 1
    h = oracle.filter_trace(imp, spec, DT, 2001, trend=0)
    assert np.max(np.abs(h[:1000][::-1] - h[1001:])) <= 1e-12


def test_filter_no_wraparound(oracle):
    """tfilterapply.nim:134-155: a spike on the last sample against a plan six
    times longer."""
    spec = bandPass(Corners(4, 8, 60, 80))
    x = np.zeros(2001)
    x[-1] = 1
    a = oracle.filter_trace(x, spec, DT, 2001, trend=0)
    b = oracle.filter_trace(x, spec, DT, 6 * 2001, trend=0)
    assert np.max(np.abs(a - b)) <= 1e-3 * np.max(np.abs(b))


# ------------------------------------------------------------ correlation

def direct(x, p, m):
    s = 0.0
    for j in range(p.shape[0]):
        i = j + m
        if 0 <= i < x.shape[0]:
            s += x[i] * p[j]
    return s


def test_correlation_equals_direct(oracle):
    """tcorrelate.nim:40-64."""
    pilot = oracle.linear_sweep(8, 60, 1000, 2, 100, 100)
    n = 1201
    x = np.random.default_rng(7).standard_normal(n)
    y = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, 2.0, n, normalize=False)[0]
    assert y.shape[0] == n - pilot.shape[0]
    d = np.array([direct(x, pilot, m) for m in range(y.shape[0])])
    peak = np.max(np.abs(d))
    assert np.max(np.abs(y - d)) < 1e-10 * peak
    y2 = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, 2.0, n, out_samples=n,
                          normalize=False)[0]
    for m in (n - pilot.shape[0], n - 100, n - 1):
        assert abs(y2[m] - direct(x, pilot, m)) < 1e-10 * peak


@pytest.mark.parametrize("t0", [-20.0, 20.0, -900.0, 3000.0])
def test_correlation_whole_sample_start(oracle, t0):
    """tcorrelate.nim:66-94: column m is the direct correlation at lag
    m + t0/dt."""
    pilot = oracle.linear_sweep(8, 60, 1000, 2, 100, 100)
    n = 1801
    x = np.random.default_rng(11).standard_normal(n)
    y = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, 2.0, n, out_samples=n,
                         normalize=False, pilot_t0_ms=t0)[0]
    s = int(t0 / 2)
    d = np.array([0.0 if (m + s <= -pilot.shape[0] or m + s >= n) else direct(x, pilot, m + s)
                  for m in range(n)])
    assert np.max(np.abs(y - d)) < 1e-10 * np.max(np.abs(d))


def test_correlation_klauder_and_normalization(oracle):
    """tcorrelate.nim:96-134: spikes correlate back to Klauder peaks of their
    amplitude; a reflector's amplitude and scale * sum(p^2) = 1 do not depend
    on the pilot's amplitude or length."""
    pilot = oracle.linear_sweep(10, 80, 4000, 2, 200, 200)
    n = 3501
    refl = np.zeros(n)
    refl[100], refl[400], refl[900] = 1.0, -0.5, 0.25
    x = np.convolve(refl, pilot)[:n]
    y = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, 2.0, n)[0]
    assert abs(y[100] - 1.0) < 1e-3 and abs(y[400] + 0.5) < 1e-3 and abs(y[900] - 0.25) < 1e-3
    for at, amp in ((100, 1.0), (400, -0.5), (900, 0.25)):
        for d in (-3, -2, -1, 1, 2, 3):
            assert abs(y[at + d]) < abs(amp)
    n = 6001
    refl = np.zeros(n)
    refl[300] = 0.7
    for amp, length in ((1.0, 4000.0), (3.5, 8000.0)):
        pilot = oracle.linear_sweep(10, 80, length, 2, 200, 200, 0.0, amp)
        x = np.convolve(refl, pilot)[:n]
        info = oracle.corr_info(pilot, 2.0, n)
        assert abs(info["scale"] * np.sum(pilot ** 2) - 1) < 1e-9
        y = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, 2.0, n)[0]
        assert abs(y[300] - 0.7) < 1e-6 * 0.7


def test_correlation_plan_sizes(oracle):
    """bench/README.md:369: 10001-sample records with a 12 s sweep at 2 ms
    use nfft 16200; the listen time is 4001 columns."""
    pilot = oracle.linear_sweep(6, 80, 12000, 2, 300, 300)
    assert pilot.shape[0] == 6000
    info = oracle.corr_info(pilot, 2.0, 10001)
    assert info["nfft"] == 16200 and info["out_samples"] == 4001 and info["shift"] == 0
    with pytest.raises(ValueError):
        oracle.corr_info(pilot, 2.0, 5000)           # pilot longer than the record
    with pytest.raises(ValueError):
        oracle.corr_info(pilot, 2.0, 10001, pilot_t0_ms=1e30)


def test_sweep_closed_form(oracle):
    """A linear sweep's phase 2 pi (f1 t + rate t^2 / 2) and cos² ramps
    (vibroseis/sweep.nim:292-342, taper.nim:125-179); the library's sweep is
    the oracle's."""
    x = oracle.linear_sweep(6, 80, 12000, 2, 300, 300)
    t = np.arange(6000) * 0.002
    raw = np.sin(2 * np.pi * (6 * t + 0.5 * (74 / 12) * t * t))
    w = np.ones(6000)
    ts = np.arange(6000) * 2.0
    w[ts < 300] = 0.5 - 0.5 * np.cos(np.pi * ts[ts < 300] / 300)
    te = 12000 - ts
    w[te < 300] = 0.5 - 0.5 * np.cos(np.pi * te[te < 300] / 300)
    assert np.max(np.abs(x - raw * w)) <= 1e-9
    lib = P.sweepSamples(P.linearSweep(6, 80, 12000, 2, P.initSweepTaper(300, 300)))
    assert np.max(np.abs(lib - x)) <= 1e-12


# --------------------------------------------------------------- pipeline

def test_pipeline_margins_parts(oracle):
    """tpipeline.nim:153-169, 306-331."""
    flow = P.initPipeline([P.demeanOp(), P.bandPassOp(Corners(4, 8, 60, 80)), P.agcOp(500)])
    info = oracle.plan_info(flow.ops, np.float32, 2.0, 2001)
    assert info["margin"] == 110 + 125 == flow.requiredMargin(2.0)
    assert info["parts"] == 1 and info["nfft"][1] == 2250
    for dt in (1.0, 2.0, 4.0):
        assert flow.requiredMargin(dt) == oracle.plan_info(flow.ops, np.float32, dt, 1500)["margin"]
    assert P.initPipeline([P.gainDbOp(3)]).requiredMargin(2) == 0
    panel = P.initPipeline([P.normalizeOp(P.normMax, scope=P.perPanel), P.demeanOp(),
                            P.normalizeOp(P.normRms, scope=P.perPanel), P.agcOp(100)])
    assert oracle.plan_info(panel.ops, np.float32, 2.0, 500)["parts"] == 2
    assert P.decodeWindow(range(800, 1201), 110, 2001) == range(690, 1311)
    assert P.decodeWindow(range(50, 301), 110, 2001) == range(0, 411)
    assert P.decodeWindow(range(1900, 2001), 110, 2001) == range(1790, 2001)
    assert P.decodeWindow(range(10, 21), -5, 100) == range(10, 21)
    assert P.decodeWindow(range(10, 21), 2 ** 62, 100) == range(0, 100)


def test_pipeline_equals_ops_one_by_one(oracle):
    """tpipeline.nim:171-213: plan.apply equals the ops applied one by one;
    in place equals src -> dst; non-live rows are copied unchanged."""
    from processseg_amd import synth
    x, _ = synth.gather(20, 1500, 2.0, 4, dtype=np.float64)
    row_len = np.array([1500, 1400, -1, 0, 1, 2, 1500, 900] + [1500] * 12, np.int32)
    ops = [P.demeanOp(), P.bandPassOp(Corners(4, 8, 60, 80)), P.agcOp(300),
           P.normalizeOp(P.normPercentile)]
    whole = oracle.pipeline_apply(ops, x, row_len, 2.0)
    y = x.copy()
    for op in ops:
        y = oracle.pipeline_apply([op], y, row_len, 2.0)
    assert np.max(np.abs(whole - y)) <= 1e-13 * np.max(np.abs(y))
    inplace = x.copy()
    oracle.pipeline_apply(ops, inplace, row_len, 2.0, dst=inplace)
    assert np.array_equal(inplace, whole)
    assert np.array_equal(whole[2], x[2]) and np.array_equal(whole[3], x[3])
    assert np.array_equal(whole[7, 900:], x[7, 900:])


def test_compile_errors_name_the_op(oracle):
    """pipeline.nim:384-386, 511-514: "op i (kind): ..."."""
    for flow, prefix in [
        (P.initPipeline([P.demeanOp(), P.agcOp(1)]), "op 1 (opAgc): "),
        (P.initPipeline([P.scaleOp(math.inf)]), "op 0 (opScale): "),
        (P.initPipeline([P.demeanOp(), P.demeanOp(), P.clipOp(2, 1)]), "op 2 (opClip): "),
        (P.initPipeline([P.bandPassOp(Corners(4, 8, 60, 80))]), None),
    ]:
        if prefix is None:
            continue
        with pytest.raises(ValueError) as e:
            oracle.plan_info(flow.ops, np.float32, 2.0, 1000)
        assert str(e.value).startswith(prefix), str(e.value)
        with pytest.raises(ValueError) as e:  # the library, before any device work
            flow.compile(np.float32, 2.0, 1000)
        assert str(e.value).startswith(prefix), str(e.value)
    with pytest.raises(ValueError):
        oracle.plan_info([P.agcOp(100)], np.float32, 2.0, 0)
    # a calibration after an op that changes amplitudes (pipeline.nim:495-511)
    with pytest.raises(ValueError) as e:
        P.initPipeline([P.agcOp(100), P.calibrateOp(P.initCalibration())]).compile(
            np.float32, 2.0, 1000)
    assert str(e.value).startswith("op 1 (opCalibrate): a calibration must come before"), str(e.value)
