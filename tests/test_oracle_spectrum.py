"""The oracle's amplitude spectra (fft/spectrum.nim) against the reference's
own tests (tests/tspectrum.nim): the plan's lengths and refusals, the peak
within one bin, a sinusoid on a bin reading its amplitude whatever the taper
and padding (and its phase), DC and demean, the power average over live rows
(sqrt(21/3)), rows past the live end, NaN rows left out, and short windows
scaled like full ones."""
import math

import numpy as np
import pytest

import oracle as O


def sinusoid(n, dt, f, amp, phase=0.0):
    return amp * np.sin(2 * np.pi * f * np.arange(n) * dt / 1000 + phase)


def bin_hz(p, k):
    return k * 1000 / (p["nfft"] * p["dt_ms"])


def test_plan_validation():
    """tspectrum.nim:25-43."""
    p = O.spectrum_plan(2, 501)
    assert p["nfft"] == 512 and p["bins"] == 257
    assert abs(bin_hz(p, 256) - 250) <= 1e-9
    assert O.spectrum_plan(4, 100, nfft=1000)["nfft"] == 1000
    assert O.spectrum_plan(2, 512)["nfft"] == 512 and O.spectrum_plan(2, 1000)["nfft"] == 1024
    assert O.spectrum_plan(1, 60001)["nfft"] == 65536
    for kw in (dict(dt_ms=2, window_len=0), dict(dt_ms=2, window_len=-3), dict(dt_ms=0, window_len=10),
               dict(dt_ms=math.nan, window_len=10), dict(dt_ms=math.inf, window_len=10),
               dict(dt_ms=2, window_len=10, taper_fraction=-0.1),
               dict(dt_ms=2, window_len=10, taper_fraction=0.6),
               dict(dt_ms=2, window_len=100, nfft=64), dict(dt_ms=2, window_len=100, nfft=101)):
        with pytest.raises(ValueError):
            O.spectrum_plan(**kw)


def test_peak_and_amplitude_on_a_bin():
    """tspectrum.nim:45-81."""
    for win, nfft, taper in ((501, 0, 0.1), (400, 1024, 0.5), (1000, 0, 0.0), (257, 0, 0.25)):
        p = O.spectrum_plan(2, win, taper, nfft)
        for f in (7.3, 25.0, 61.7, 180.2):
            amp, _ = O.spectrum_amplitude(p, sinusoid(win, 2, f, 3))
            assert abs(bin_hz(p, int(np.argmax(amp))) - f) <= bin_hz(p, 1)
    for taper, nfft, b in ((0.0, 1000, 40), (0.1, 0, 40), (0.5, 0, 40), (0.5, 4096, 160)):
        p = O.spectrum_plan(1, 1000, taper, nfft)
        x = sinusoid(1000, 1, bin_hz(p, b), 2.5, 0.3)
        amp, ph = O.spectrum_amplitude(p, x)
        k = int(np.argmax(amp))
        assert abs(amp[k] - 2.5) <= 0.01
        if taper == 0:
            assert abs(ph[k] - (0.3 - np.pi / 2)) <= 1e-6


def test_dc_and_demean():
    """tspectrum.nim:83-93."""
    for demean in (True, False):
        p = O.spectrum_plan(2, 200, 0.0, demean=demean)
        amp, _ = O.spectrum_amplitude(p, np.full(200, 4.0))
        assert abs(amp[0] - (0.0 if demean else 4.0)) <= 1e-9


def test_average_over_live_rows():
    """tspectrum.nim:111-153: amplitudes 1..6, row 2 dead, row 4 failed
    (row_len -1), row 5 live for 100 samples, before the window at 150."""
    x = np.stack([sinusoid(600, 2, 25.0, r + 1) for r in range(6)]).astype(np.float32)
    rl = np.array([600, 600, -1, 600, -1, 100], np.int32)
    p = O.spectrum_plan(2, 400, 0.5, 400)
    amp, used = O.spectrum_average(p, x, rl, 150)
    assert used == 3 and int(np.argmax(amp)) == 20
    assert abs(amp[20] - math.sqrt(21 / 3)) <= 1e-3
    one, u1 = O.spectrum_average(p, x[3:4], rl[3:4], 150)
    assert u1 == 1 and abs(one[20] - 4) <= 1e-3
    a1, _ = O.spectrum_amplitude(p, x[3, 150:550])
    assert np.max(np.abs(a1 - one)) <= 1e-4
    zero, u0 = O.spectrum_average(p, x[2:3], rl[2:3], 150)
    assert u0 == 0 and np.all(zero == 0)
    with pytest.raises(ValueError):
        O.spectrum_average(p, x, rl, -1)


def test_nan_rows_and_short_windows():
    """tspectrum.nim:176-230."""
    x = np.stack([sinusoid(400, 2, 25.0, 2) for _ in range(4)]).astype(np.float32)
    p = O.spectrum_plan(2, 200, 0.1, 200)
    rl = np.full(4, 400, np.int32)
    for bad in (np.nan, np.inf, -np.inf):
        x[2, 10] = bad
        amp, used = O.spectrum_average(p, x, rl, 0)
        assert used == 3 and np.all(np.isfinite(amp))
        assert abs(amp[int(np.argmax(amp))] - 2) <= 0.02
        assert O.spectrum_average(p, x, rl, 200)[1] == 4
    p = O.spectrum_plan(1, 400, 0.2, 400)
    amp, _ = O.spectrum_amplitude(p, sinusoid(300, 1, 50, 3))
    assert int(np.argmax(amp)) == 20 and abs(amp[20] - 3) <= 0.03
    g = np.stack([sinusoid(500, 1, 50, 3) for _ in range(11)])
    rl = np.array([500 if r % 2 == 0 else 300 + 100 * (r % 3) for r in range(11)], np.int32)
    avg, used = O.spectrum_average(p, g, rl, 100)
    assert used == 11 and int(np.argmax(avg)) == 20 and abs(avg[20] - 3) <= 0.05
