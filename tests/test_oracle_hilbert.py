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
