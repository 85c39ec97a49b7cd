"""Synthetic shots for the first-break tests: a first break linear in
|offset| (a decaying 30 Hz onset), a later, stronger Ricker event, white
noise; seeded, this build's own generator."""
import numpy as np

DT = 2.0
COLS = 600
VEL = 2000.0
T0 = 20.0


def split_spread(n, dx, near):
    """Offsets of a split spread: n // 2 on each side, the nearest at ±near."""
    side = near + dx * np.arange(n // 2, dtype=np.float64)
    return np.concatenate([-side[::-1], side])


def onset_ms(offset, t0=T0, vel=VEL):
    return t0 + 1000 * np.abs(offset) / vel


def add_onset(row, dt, t_ms, f_hz=30.0, amp=1.0, decay_ms=40.0):
    t = np.arange(row.size) * dt - t_ms
    on = t >= 0
    row[on] += amp * np.sin(2 * np.pi * f_hz * t[on] / 1000) * np.exp(-t[on] / decay_ms)


def add_ricker(row, dt, t_ms, f_hz=25.0, amp=1.0):
    t = (np.arange(row.size) * dt - t_ms) / 1000
    a = (np.pi * f_hz * t) ** 2
    row += amp * (1 - 2 * a) * np.exp(-a)


def shot(sigma, seed, offsets=None, cols=COLS, dt=DT, dtype=np.float32):
    """Rows and offsets: the first break at onset_ms(offset), a Ricker of
    amplitude 2 250 ms later, N(0, sigma) noise."""
    offsets = split_spread(24, 25, 25) if offsets is None else np.asarray(offsets, np.float64)
    x = np.zeros((offsets.size, cols))
    for r, off in enumerate(offsets):
        add_onset(x[r], dt, onset_ms(off))
        add_ricker(x[r], dt, onset_ms(off) + 250, amp=2)
        x[r] += np.random.default_rng(seed * 1000 + r).normal(0, sigma, cols)
    return x.astype(dtype), offsets


def keys_of(offsets, t0=0.0):
    k = np.zeros((len(offsets), 2))
    k[:, 0] = t0
    k[:, 1] = offsets
    return k
