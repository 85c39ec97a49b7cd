"""The oracle's correlation branches beyond plain correlation at whole-sample
pilot starts: sweep deconvolution and sub-sample pilot start times, pinned by
the reference's own checks (tests/tcorrelate.nim:136-209, 306-364)."""
import math

import numpy as np
import pytest

DT = 2.0
# Gauss-Legendre nodes and weights on [-1, 1] (4 points)
_GL_X = np.array([-0.8611363115940526, -0.3399810435848563, 0.3399810435848563,
                  0.8611363115940526])
_GL_W = np.array([0.3478548451374538, 0.6521451548625461, 0.6521451548625461,
                  0.3478548451374538])


def _cos2(x):
    x = np.clip(x, 0.0, 1.0)
    return 0.5 - 0.5 * np.cos(np.pi * x)


def _taper(x, dt, length_ms, start_ms, end_ms):
    """cos² ramps over t < start_ms and length_ms - t < end_ms
    (vibroseis/taper.nim:160-179)."""
    t = np.arange(x.shape[0]) * dt
    w = np.ones_like(x)
    if start_ms > 0:
        s = t < start_ms
        w[s] = _cos2(t[s] / start_ms)
    if end_ms > 0:
        e = length_ms - t < end_ms
        w[e] *= _cos2((length_ms - t[e]) / end_ms)
    return x * w


def db_per_octave_sweep(f1, f2, length_ms, dt, boost_db, taper_ms):
    """A sweep whose amplitude spectrum rises boost_db per octave: f^m linear
    in time, m = boost / (20 log10(2) / 2) + 1, its phase 2π∫f by 4-point
    Gauss-Legendre per interval (vibroseis/sweep.nim:248-330)."""
    n = int(round(length_ms / dt))
    m = boost_db / (20 * math.log10(2.0) / 2) + 1

    def freq(t_ms):
        x = np.clip(t_ms, 0.0, length_ms) / length_ms
        f = f1 * np.power(1 + x * np.expm1(m * math.log(f2 / f1)), 1 / m)
        return np.clip(f, min(f1, f2), max(f1, f2))

    phase = np.zeros(n)
    a = np.arange(n - 1) * dt
    h = 0.5 * dt
    s = sum(_GL_W[q] * freq(a + h + h * _GL_X[q]) for q in range(4))
    phase[1:] = 2 * np.pi * np.cumsum(s * h / 1000)
    return _taper(np.sin(phase), dt, length_ms, taper_ms, taper_ms)


def _rfft_db(y):
    return 20 * np.log10(np.abs(np.fft.rfft(y)))


def test_deconvolution_flattens_the_spectrum(oracle):
    """tcorrelate.nim:136-179: a spike through a +4 dB/octave sweep comes back
    as a unit peak in both modes; deconvolution's wavelet spectrum is flat
    within 1 dB over 15-80 Hz, correlation's keeps |P|² (> 15 dB of rise);
    epsPercent = 1e6 tends to correlation."""
    pilot = db_per_octave_sweep(8, 90, 6000, DT, 4, 200)
    assert pilot.shape[0] == 3000
    n = 4001
    refl = np.zeros(n)
    refl[1000] = 1.0
    x = np.convolve(refl, pilot)[:n]
    spread = {}
    for mode in (0, 1):
        y = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, DT, n, out_samples=2000,
                             mode=mode, eps_percent=0.1)[0]
        assert abs(y[1000] - 1) < 1e-6, (mode, y[1000])
        s = _rfft_db(y[744:1256])[16:83]
        spread[mode] = float(s.max() - s.min())
    assert spread[1] < 1.0, spread
    assert spread[0] > 15, spread
    big = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, DT, n, out_samples=2000,
                           mode=1, eps_percent=1e6)[0]
    corr = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, DT, n,
                            out_samples=2000)[0]
    assert np.max(np.abs(big - corr)) < 1e-4


def test_deconvolution_multiplier_closed_form(oracle):
    """The deconvolution output is the inverse transform of
    X·conj(P)/(|P|²+ε)·scale with ε = epsPercent/100·max|P|² and scale =
    n / Σ'|P|²W (correlate.nim:129-154), computed here with numpy's FFT."""
    pilot = db_per_octave_sweep(8, 90, 2000, DT, 4, 100)
    n_rec = 1601
    x = np.random.default_rng(5).standard_normal(n_rec)
    info = oracle.corr_info(pilot, DT, n_rec, out_samples=n_rec, mode=1, eps_percent=0.5)
    nf = info["nfft"]
    X = np.fft.rfft(x, nf)
    Pf = np.fft.rfft(pilot, nf)
    p2 = np.abs(Pf) ** 2
    eps = 0.5 / 100 * p2.max()
    W = 1 / (p2 + eps)
    g = p2 * W
    peak = (g[0] + g[-1] + 2 * g[1:-1].sum()) / nf
    mul = np.conj(Pf) * W / peak
    mul[-1] = mul[-1].real
    full = np.fft.irfft(X * mul, nf)
    want = np.zeros(n_rec)
    want[:n_rec] = full[:n_rec]
    y = oracle.correlate(x[None, :], np.array([n_rec], np.int32), pilot, DT, n_rec,
                         out_samples=n_rec, mode=1, eps_percent=0.5)[0]
    assert abs(info["scale"] - 1 / peak) < 1e-12 / peak
    assert np.max(np.abs(y - want)) < 1e-10 * np.max(np.abs(want))


def _linear_phase(t_ms, f1, f2, length_ms):
    t = np.clip(t_ms, 0.0, length_ms) / 1000
    T = length_ms / 1000
    return 2 * np.pi * (f1 * t + 0.5 * (f2 - f1) / T * t * t)


def _sweep_at(t, f1=10.0, f2=70.0, length=3000.0, taper=200.0):
    t = np.asarray(t, np.float64)
    w = _cos2(np.minimum(t / taper, (length - t) / taper))
    y = np.sin(_linear_phase(t, f1, f2, length)) * w
    return np.where((t < 0) | (t > length), 0.0, y)


def test_sub_sample_pilot_start(oracle):
    """tcorrelate.nim:181-209: a recorded pilot whose first sample is 0.6 ms
    after the time break; correlating with its start time puts an arrival of
    the sweep at 400 ms at sample 200 (within 0.02), without it at 200.3."""
    n = 2501
    x = _sweep_at(np.arange(n) * DT - 400)
    late = _sweep_at(0.6 + np.arange(1500) * DT)

    def arrival(t0):
        y = oracle.correlate(x[None, :], np.array([n], np.int32), late, DT, n,
                             pilot_t0_ms=t0)[0]
        b = 1 + int(np.argmax(y[1:-1]))
        return b + 0.5 * (y[b - 1] - y[b + 1]) / (y[b - 1] - 2 * y[b] + y[b + 1])

    assert abs(arrival(0.6) - 200) < 0.02
    assert abs(arrival(0.0) - 200.3) < 0.02


def test_sub_sample_start_is_a_phase_ramp(oracle):
    """A start time of whole + frac samples reads the output frac samples
    later: the multiplier's e^{2πik·frac/n} (correlate.nim:150-153) against
    numpy's FFT of the direct correlation shifted by the same ramp."""
    pilot = oracle.linear_sweep(10, 70, 400, DT, 50, 50)
    n = 700
    x = np.random.default_rng(3).standard_normal(n)
    t0 = 9.4  # 4.7 samples: whole 5, frac -0.3
    info = oracle.corr_info(pilot, DT, n, out_samples=n, normalize=False, pilot_t0_ms=t0)
    nf, whole = info["nfft"], info["shift"]
    assert whole == 5
    frac = t0 / DT - whole
    k = np.arange(nf // 2 + 1)
    mul = np.conj(np.fft.rfft(pilot, nf)) * np.exp(2j * np.pi * k * frac / nf)
    mul[-1] = mul[-1].real
    full = np.fft.irfft(np.fft.rfft(x, nf) * mul, nf)
    want = np.array([full[(m + whole) % nf] if -pilot.shape[0] < m + whole < n else 0.0
                     for m in range(n)])
    y = oracle.correlate(x[None, :], np.array([n], np.int32), pilot, DT, n, out_samples=n,
                         normalize=False, pilot_t0_ms=t0)[0]
    assert np.max(np.abs(y - want)) < 1e-10 * np.max(np.abs(want))


@pytest.mark.parametrize("t0,mode,out", [(0.0, 0, 0), (-6.0, 0, 700), (9.4, 0, 0), (0.6, 1, 0)])
def test_ragged_rows_match_one_record(oracle, t0, mode, out):
    """tcorrelate.nim:306-364: rows of unequal live lengths (junk past them)
    and dead rows, both modes and pilots starting before and after the time
    break: each row equals the one-record correlation of its live samples
    (float32 within 1e-6 of the peak; float64 within 1e-12); with a
    whole-sample start and plain correlation, the direct sum."""
    n = 700
    rows = 19
    pilot = oracle.linear_sweep(10, 70, 400, DT, 50, 50).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(40)
    src = np.full((rows, n), 1e6)
    counts = np.array([n - 37 * (r % 5) for r in range(rows)], np.int32)
    for r in range(rows):
        src[r, :counts[r]] = rng.standard_normal(counts[r]).astype(np.float32)
    row_len = counts.copy()
    row_len[1] = -1
    row_len[10] = -1
    for dtype, tol in ((np.float32, 1e-6), (np.float64, 1e-12)):
        s = src.astype(dtype)
        dst = oracle.correlate(s, row_len, pilot, DT, n, out_samples=out, mode=mode,
                               normalize=False, pilot_t0_ms=t0)
        peak = float(np.max(np.abs(dst)))
        whole = int(round(t0 / DT))
        for r in range(rows):
            if row_len[r] < 0:
                assert np.all(dst[r] == 0)
                continue
            one = oracle.correlate(np.ascontiguousarray(s[r:r + 1, :counts[r]]),
                                   np.array([counts[r]], np.int32), pilot, DT, n,
                                   out_samples=out, mode=mode, normalize=False,
                                   pilot_t0_ms=t0)[0]
            assert np.max(np.abs(one.astype(np.float64) - dst[r])) <= tol * peak, (dtype, r)
            if mode == 0 and t0 == 2 * whole:
                x64 = s[r, :counts[r]].astype(np.float64)
                full = np.correlate(np.concatenate([np.zeros(pilot.shape[0]), x64,
                                                    np.zeros(pilot.shape[0])]), pilot, "valid")
                # lag L = m + whole reads full[L + pilot_len]
                want = np.array([full[m + whole + pilot.shape[0]]
                                 if 0 <= m + whole + pilot.shape[0] < full.shape[0] else 0.0
                                 for m in range(dst.shape[1])])
                assert np.max(np.abs(dst[r] - want)) <= max(2e-5, tol) * peak, (dtype, r)
