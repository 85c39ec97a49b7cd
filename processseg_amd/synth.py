"""Seeded synthetic panels (the reference's synth.nim: Ricker wavelets on
hyperbolic events plus Gaussian noise, synth.nim:57-222).

The reference draws its noise with Nim's std/random, which it documents as not
portable across Nim versions or C libraries (synth.nim:28-43); this generator
uses its own portable streams instead: SplitMix64, counter-based (sample k of
row r of seed s is mix(s + r * 2^32 + k + 1)), and Box-Muller in float64.
"""
from __future__ import annotations

import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniforms(seed: int, shape, stream: int = 0) -> np.ndarray:
    """Uniforms in (0, 1], float64, from a counter-based SplitMix64."""
    rows = shape[0] if len(shape) > 1 else 1
    cols = shape[-1]
    with np.errstate(over="ignore"):
        base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + np.uint64(stream) * np.uint64(0x100000000000)
        r = np.arange(rows, dtype=np.uint64)[:, None] * np.uint64(1 << 32)
        k = np.arange(cols, dtype=np.uint64)[None, :] + np.uint64(1)
        z = _mix((base + r + k) * _GAMMA)
    u = ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    return u.reshape(shape)


def gauss(seed: int, shape, sigma: float, stream: int = 0) -> np.ndarray:
    """N(0, sigma) by Box-Muller on two uniform streams."""
    u1 = uniforms(seed, shape, 2 * stream + 1)
    u2 = uniforms(seed, shape, 2 * stream + 2)
    return sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * math.pi * u2)


def ricker(tau_ms: np.ndarray, freq_hz: float) -> np.ndarray:
    """ricker (synth.nim:57-62): (1 - 2a) e^{-a}, a = (pi f tau)^2."""
    a = (math.pi * freq_hz * tau_ms / 1000) ** 2
    return (1 - 2 * a) * np.exp(-a)


def hyperbolic_ms(t0_ms: float, velocity: float, offset: np.ndarray) -> np.ndarray:
    """hyperbolicMs (synth.nim:146-150)."""
    return np.sqrt(t0_ms ** 2 + (np.abs(offset) / velocity * 1000) ** 2)


def gather(rows: int, cols: int, dt_ms: float, seed: int, *, events: int = 5,
           freq_hz: float = 25.0, noise: float = 0.05, offset_step: float = 25.0,
           live=None, dtype=np.float32, t0_range=(300.0, 3000.0),
           v_range=(1500.0, 3500.0), amp_range=(0.5, 2.0)):
    """A [rows, cols] CMP-like gather: row i at offset offset_step * i m, t0 = 0;
    `events` hyperbolic reflections (t0, velocity, amplitude uniform in the
    ranges), a Ricker wavelet cut at 3 periods (synth.nim:77-96), plus
    N(0, noise).  `live`: per-row live lengths (zeros past them).  Returns
    (panel, live lengths)."""
    ev = uniforms(seed, (1, 3 * events), stream=7)[0]
    offsets = offset_step * np.arange(rows, dtype=np.float64)
    t = np.arange(cols, dtype=np.float64) * dt_ms
    x = np.zeros((rows, cols), np.float64)
    half = 3000.0 / freq_hz
    for e in range(events):
        t0 = t0_range[0] + (t0_range[1] - t0_range[0]) * ev[3 * e]
        v = v_range[0] + (v_range[1] - v_range[0]) * ev[3 * e + 1]
        amp = amp_range[0] + (amp_range[1] - amp_range[0]) * ev[3 * e + 2]
        centre = hyperbolic_ms(t0, v, offsets)
        tau = t[None, :] - centre[:, None]
        mask = np.abs(tau) <= half
        x += np.where(mask, amp * ricker(np.where(mask, tau, 0.0), freq_hz), 0.0)
    if noise > 0:
        x += gauss(seed, (rows, cols), noise)
    if live is None:
        live = np.full(rows, cols, np.int64)
    live = np.asarray(live, np.int64)
    x[np.arange(cols)[None, :] >= live[:, None]] = 0.0
    return x.astype(dtype), live


def live_lengths(seed: int, rows: int, lo: int, hi: int) -> np.ndarray:
    """Per-row live lengths uniform in [lo, hi]."""
    u = uniforms(seed, (1, rows), stream=11)[0]
    return (lo + np.floor(u * (hi - lo + 1) - 1e-12)).astype(np.int64).clip(lo, hi)


def vibroseis_records(rows: int, cols: int, pilot: np.ndarray, seed: int, p: float = 0.01,
                      noise: float = 0.1, dtype=np.float32):
    """Records = pilot convolved with sparse reflectivity (Bernoulli p,
    N(0, 1) amplitudes) plus N(0, noise), cut to `cols` samples."""
    spikes = uniforms(seed, (rows, cols), stream=21) < p
    amps = gauss(seed, (rows, cols), 1.0, stream=22)
    refl = np.where(spikes, amps, 0.0)
    n = 1
    while n < cols + pilot.shape[0]:
        n *= 2
    R = np.fft.rfft(refl, n, axis=1)
    P = np.fft.rfft(pilot, n)
    x = np.fft.irfft(R * P[None, :], n, axis=1)[:, :cols]
    if noise > 0:
        x += gauss(seed, (rows, cols), noise, stream=23)
    return x.astype(dtype), refl
