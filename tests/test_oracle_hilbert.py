"""The oracle's Hilbert transform and envelope (fft/hilbert.nim) against what
the definition fixes: H[cos] = sin and H[sin] = -cos inside a trace, a
modulated carrier's envelope equal to its modulation, H linear and
anti-self-inverse on band-limited content (H[H[x]] = -x inside), the plan's
lengths and refusals."""
import numpy as np
import pytest

import oracle as O


def test_cos_to_sin_and_envelope():
    n, dt = 4001, 2.0
    t = np.arange(n) * dt / 1000
    w = 2 * np.pi * 25.0
    inner = slice(500, n - 500)
    h = O.hilbert(np.cos(w * t) * np.hanning(n) ** 0, n)
    # a whole-record cosine rings at the ends only (fft/hilbert.nim:20-25)
    assert np.max(np.abs(h[inner] - np.sin(w * t)[inner])) < 0.02
    taper = np.sin(np.pi * np.arange(n) / (n - 1)) ** 2         # starts and ends at zero
    x = taper * np.cos(w * t)
    env = O.hilbert(x, n, envelope=True)
    assert np.max(np.abs(env[inner] - taper[inner])) < 1e-3
    hs = O.hilbert(taper * np.sin(w * t), n)
    assert np.max(np.abs(hs[inner] + x[inner])) < 1e-3


def test_linear_and_twice_is_minus():
    rng = np.random.default_rng(4)
    n = 2001
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    ha, hb = O.hilbert(a, n), O.hilbert(b, n)
    assert np.allclose(O.hilbert(2 * a - 3 * b, n), 2 * ha - 3 * hb, atol=1e-12)
    t = np.arange(n) * 0.002
    x = np.sin(np.pi * np.arange(n) / (n - 1)) ** 2 * np.cos(2 * np.pi * 40 * t)
    hh = O.hilbert(O.hilbert(x, n), n)
    assert np.max(np.abs(hh[300:-300] + x[300:-300])) < 1e-3


def test_plan_lengths():
    assert O.hilbert(np.ones(3), 3).shape == (3,)
    with pytest.raises(ValueError):
        O.hilbert(np.ones(3), 0)
    with pytest.raises(ValueError):
        O.hilbert(np.ones(10), 10, nfft=16)       # would wrap
    with pytest.raises(ValueError):
        O.hilbert(np.ones(10), 10, nfft=22)       # not smooth


# ------------------------------------------------------ analytic signal
# The reference's thilbert.nim analytic blocks (tests/thilbert.nim:159-297),
# restated on the oracle's analyticInto.

def _gauss(j, n):
    u = (np.asarray(j, np.float64) - (n - 1) / 2) / (n / 12)
    return np.exp(-u * u)


def test_analytic_sinusoid():
    """thilbert.nim:159-195: a Gaussian-tapered tone on a bin: envelope,
    phase and frequency inside; NaN at both ends and 60 dB down; an untapered
    tone's middle still reads the carrier; zeros give NaN phase and hz."""
    n, dt = 2001, 2.0
    nfft = O.next_fast_len(2 * n)
    f = 1000 / (nfft * dt) * 80
    j = np.arange(n)
    x = _gauss(j, n) * np.cos(2 * np.pi * f * j * dt / 1000)
    env, phase, hz = O.analytic(x, n, dt)
    assert np.isnan(hz[0]) and np.isnan(hz[-1]) and np.isnan(hz[10])
    mid = slice(n // 3, 2 * n // 3)
    assert np.max(np.abs(env[mid] - _gauss(j[mid], n))) < 1e-9
    assert np.max(np.abs(hz[mid] - f)) < 1e-4
    p = 2 * np.pi * f * j[mid] * dt / 1000
    want = np.arctan2(np.sin(p), np.cos(p)) * 180 / np.pi
    want = np.where(want <= -180, want + 360, want)
    got = np.where(phase[mid] <= -180, phase[mid] + 360, phase[mid])
    assert np.max(np.abs(got - want)) < 1e-6
    bare = np.cos(2 * np.pi * f * j * dt / 1000)
    _, _, hzb = O.analytic(bare, n, dt)
    assert np.max(np.abs(hzb[mid] - f)) < 0.01
    z_env, z_phase, z_hz = O.analytic(np.zeros(n), n, dt)
    assert np.isnan(z_phase[n // 2]) and np.isnan(z_hz[n // 2]) and z_env[n // 2] == 0
    assert np.all(np.isnan(z_hz)) and np.all(z_env == 0)


@pytest.mark.parametrize("frac", [0.1, 0.24, 0.26, 0.4, 0.49])
def test_analytic_frequency_to_nyquist(frac):
    """thilbert.nim:197-216: a tone of frac·fs reads its frequency within
    0.5 % over the middle half (a step wrapped to ±π per sample, not per two)."""
    n, dt = 2001, 4.0
    fs = 1000 / dt
    f = frac * fs
    j = np.arange(n)
    _, _, hz = O.analytic(np.cos(2 * np.pi * f * j * dt / 1000 + 0.4), n, dt)
    mid = slice(n // 4, 3 * n // 4)
    assert np.max(np.abs(hz[mid] - f)) < 0.005 * f


def test_analytic_sweep_and_float32_tone():
    """thilbert.nim:217-247: a float32 tone at 0.4 fs; a 10 -> 100 Hz sweep in
    4 s reads f(t) = 10 + 22.5 t within 1 Hz away from its ends."""
    n, dt = 2001, 4.0
    fs = 1000 / dt
    j = np.arange(n)
    x32 = np.cos(2 * np.pi * 0.4 * fs * j * dt / 1000).astype(np.float32)
    _, _, hz = O.analytic(x32.astype(np.float64), n, dt)
    mid = slice(n // 4, 3 * n // 4)
    assert np.max(np.abs(hz[mid] - 0.4 * fs)) < 0.005 * 0.4 * fs
    t = j * dt / 1000
    t1 = 4.0
    tp = 2 * np.pi * (10 * t + 22.5 * t * t / 2)
    x = np.where(t <= t1, np.sin(tp) * (0.5 - 0.5 * np.cos(2 * np.pi * np.minimum(
        np.minimum(t, t1 - t), 0.4) / 0.8)), 0.0)
    _, _, hz = O.analytic(x, n, dt)
    sel = (t > 0.3) & (t < t1 - 0.3)
    assert sel.sum() > 800
    assert np.max(np.abs(hz[sel] - (10 + 22.5 * t[sel]))) < 1.0


def test_analytic_near_nyquist_nyquist_and_two_tones():
    """thilbert.nim:249-297: a 0.49 fs tone in noise stays above fs/4 (the
    circular mean keeps both steps on one side of the wrap); exactly Nyquist
    reads +fs/2; two tones read the exact (a1²f1 + a2²f2 + a1a2(f1+f2)cosΔ)/|z|²,
    negative where the beat is at its minimum."""
    n, dt = 2001, 4.0
    fs = 1000 / dt
    j = np.arange(n)
    seed = np.uint32(12345)
    noise = np.empty(n)
    for i in range(n):
        seed = np.uint32((int(seed) * 1664525 + 1013904223) & 0xFFFFFFFF)
        noise[i] = float(int(seed) >> 8) / float(1 << 24) - 0.5
    x = _gauss(j, n) * np.cos(2 * np.pi * 0.49 * fs * j * dt / 1000) + 0.05 * _gauss(j, n) * noise
    _, _, hz = O.analytic(x, n, dt)
    mid = slice(n // 3, 2 * n // 3)
    assert np.all(hz[mid] > fs / 4)
    x = _gauss(j, n) * np.cos(2 * np.pi * 0.25 * fs * j * dt / 1000)
    _, _, hz = O.analytic(x, n, dt)
    assert np.max(np.abs(hz[mid] - fs / 4)) < 0.005 * fs / 4
    x = np.where(j % 2 == 0, 1.0, -1.0)
    _, _, hz = O.analytic(x, n, dt)
    assert np.max(np.abs(hz[mid] - fs / 2)) < 0.01 * fs
    t = j * dt / 1000
    x = np.cos(2 * np.pi * 5 * t) + 0.7 * np.cos(2 * np.pi * 10 * t)
    _, _, hz = O.analytic(x, n, dt)
    d = 2 * np.pi * 5 * t[mid]
    want = (5 + 0.49 * 10 + 0.7 * 15 * np.cos(d)) / (1 + 0.49 + 1.4 * np.cos(d))
    assert np.max(np.abs(hz[mid] - want)) < 1.0
    assert np.min(hz[mid]) < -5
