"""The oracle's spectrogram (fft/spectrogram.nim) against the reference's own
known answers (tests/tspectrogram.nim): a brute-force windowed DFT of every
frame of random data, a sinusoid on a bin reading its amplitude in every
frame, a trace shorter than the plan and a frame past it, a NaN confined to
the frames that hold it, and the plan's refusals.  The sweep-ridge block
needs the sweep generator, which this path does not carry: not restated."""
import math

import numpy as np
import pytest

import oracle as O


def test_brute_force_dft():
    """tspectrogram.nim:28-55."""
    n = 300
    x = np.random.default_rng(4).standard_normal(n)
    amp = O.spectrogram(x, 41, 7, 2.0, n)
    assert amp.shape == (43, 33)
    for i in range(43):
        start = i * 7 - 20
        sw = sum(math.sin(math.pi * (j + 1) / 42) ** 2 for j in range(41) if 0 <= start + j < n)
        for k in range(33):
            re = im = 0.0
            for j in range(41):
                c = start + j
                if not 0 <= c < n:
                    continue
                w = math.sin(math.pi * (j + 1) / 42) ** 2
                a = -2 * math.pi * k * j / 64
                re += w * x[c] * math.cos(a)
                im += w * x[c] * math.sin(a)
            want = math.hypot(re, im) * (1.0 if k in (0, 32) else 2.0) / sw
            assert abs(amp[i, k] - want) < 1e-10 * max(1.0, want), (i, k)


def test_sinusoid_on_a_bin():
    """tspectrogram.nim:57-76."""
    n = 2000
    f = 20 * 1000 / (128 * 2.0)
    assert abs(f - 78.125) < 1e-12
    x = 3 * np.sin(2 * np.pi * f * np.arange(n) * 0.002 + 0.4)
    amp = O.spectrogram(x, 127, 25, 2.0, n, nfft=128)
    for i in range(amp.shape[0]):
        c = i * 25
        lim = 3e-4 if (c >= 63 and c + 63 < n) else 0.15
        assert abs(amp[i, 20] - 3) < lim, (i, amp[i, 20])


def test_float32_short_trace_and_nan():
    """tspectrogram.nim:113-127."""
    x = np.sin(2 * np.pi * 125 * np.arange(600) / 1000).astype(np.float32)
    amp = O.spectrogram(x, 64, 16, 1.0, 1000)
    assert abs(amp[10, 8] - 1) < 1e-3
    assert amp[-1, 8] == 0
    bad = x.copy()
    bad[160] = np.nan
    amp = O.spectrogram(bad, 64, 16, 1.0, 1000)
    assert np.isnan(amp[10, 8]) and not np.isnan(amp[20, 8])


def test_refusals():
    """tspectrogram.nim:129-134."""
    for args, kw in (((1, 10, 1, 100), {}), ((64, 0, 1, 100), {}), ((64, 10, 0, 100), {}),
                     ((64, 10, 1, 0), {}), ((64, 10, 1, 100), {"nfft": 32}),
                     ((64, 10, 1, 100), {"nfft": 77})):
        with pytest.raises(ValueError):
            O.spectrogram(np.zeros(10), *args, **kw)
    with pytest.raises(ValueError):
        O.spectrogram(np.zeros(101), 64, 10, 1, 100)


# ------------------------------------------------- averaged spectrogram
# tspectrogram.nim:157-234, 410-446 restated on the oracle.

def _tone_rows(rows, f, amps, n=128, dt=2.0):
    j = np.arange(n)
    return np.stack([np.cos(2 * np.pi * f * j * dt / 1000) * a for a in amps[:rows]])


def test_average_drops_a_nan_row():
    """tspectrogram.nim:157-177: a NaN frame drops its trace (2 of 3 used);
    the tone of amplitude 2 on bin 4 reads 2 in the middle frame."""
    f = 4 * 1000 / (32 * 2.0)
    x = _tone_rows(3, f, [2, 2, 2])
    x[2, 10] = np.nan
    amp, used = O.spectrogram_average(x, np.full(3, 128), 32, 16, 2.0, 128, 32)
    assert used == 2
    mid = amp.shape[0] // 2
    assert abs(amp[mid, 4] - 2.0) < 1e-3


@pytest.mark.parametrize("bad", [np.inf, -np.inf])
def test_average_nonfinite_rows(bad):
    """tspectrogram.nim:179-234: an infinity in a covered sample drops its
    row; a float64 row with a NaN or an infinity between frames, or made of
    subnormals, is left out before the transform (all-zero result)."""
    f = 4 * 1000 / (32 * 2.0)
    x = _tone_rows(2, f, [2, 2])
    x[1, 20] = bad
    amp, used = O.spectrogram_average(x, np.full(2, 128), 32, 16, 2.0, 128, 32)
    assert used == 1 and abs(amp[amp.shape[0] // 2, 4] - 2.0) < 1e-3
    x32 = _tone_rows(2, f, [2, 2]).astype(np.float32)
    x32[0, 20] = np.inf
    _, used32 = O.spectrogram_average(x32, np.full(2, 128), 32, 16, 2.0, 128, 32,
                                      dtype=np.float32)
    assert used32 == 1
    gap = np.full((1, 40), 0.25)
    for hole in (np.nan, np.inf, -np.inf):
        g = gap.copy()
        g[0, 5] = hole
        amp, used = O.spectrogram_average(g, [40], 2, 10, 1.0, 40, 4)
        assert used == 0 and np.all(amp == 0)
        g = gap.copy()
        g[0, 0] = hole
        assert O.spectrogram_average(g, [40], 2, 10, 1.0, 40, 4)[1] == 0
    amp, used = O.spectrogram_average(np.full((1, 40), 1e-320), [40], 2, 10, 1.0, 40, 4)
    assert used == 0 and np.all(amp == 0)


def test_average_is_the_power_mean():
    """tspectrogram.nim:410-446: float32 rows average as the power mean of
    their own spectrograms (within 1e-12 of the peak); an infinite row is left
    out."""
    f = 4 * 1000 / (32 * 2.0)
    x = _tone_rows(3, f, [1, 2, 3]).astype(np.float32)

    def power_mean(rows):
        acc = sum(O.spectrogram(x[r].astype(np.float64), 32, 16, 2.0, 128, 32) ** 2 for r in rows)
        return np.sqrt(acc / len(rows))

    amp, used = O.spectrogram_average(x, np.full(3, 128), 32, 16, 2.0, 128, 32, dtype=np.float32)
    want = power_mean([0, 1, 2])
    assert used == 3 and np.max(np.abs(amp - want)) <= 1e-12 * want.max()
    x[1, 20] = np.inf
    amp, used = O.spectrogram_average(x, np.full(3, 128), 32, 16, 2.0, 128, 32, dtype=np.float32)
    want = power_mean([0, 2])
    assert used == 2 and np.max(np.abs(amp - want)) <= 1e-12 * want.max()
    # rows not live or with no sample are not counted
    amp, used = O.spectrogram_average(x, np.array([-1, 0, 128]), 32, 16, 2.0, 128, 32,
                                      dtype=np.float32)
    assert used == 1
