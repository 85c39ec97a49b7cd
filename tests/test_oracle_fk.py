"""The f-k amplitude (fft/fk.nim): the oracle's restatement held to the
reference's tests/tfk.nim: the dip's sign and axes, a slow DFT of tapered
gathers at ordinary and extreme scales, the NaN / Inf classes of each
output, and gathers whose trace counts leave a lane remainder."""
import math

import numpy as np
import pytest

import oracle as O


def direct_fk(data: np.ndarray, taper: float) -> np.ndarray:
    """The reference test's slow DFT (tests/tfk.nim:124-160): the tapered
    gather's DFT in time, exp(-2 pi i j k / nTime), then across traces,
    output ik reading spatial bin nK/2 - ik, the modulus by hypot; float64,
    with the angles reduced exactly (integer products mod n)."""
    traces, samples = data.shape
    n_time = O.next_fast_len(samples)
    n_k = O.next_fast_len(2 * traces)
    n_freq = n_time // 2 + 1
    tw = O._tukey_fk(samples, taper, np.float64)
    rw = O._tukey_fk(traces, taper, np.float64)
    x = data * tw[None, :] * rw[:, None]
    j = np.arange(samples)
    spec = np.zeros((traces, n_freq), complex)
    for k in range(n_freq):
        ang = -2 * np.pi * ((k * j) % n_time) / n_time
        spec[:, k] = x @ np.exp(1j * ang)
    out = np.zeros(n_k * n_freq)
    half = n_k // 2
    t = np.arange(traces)
    for ik in range(n_k):
        raw = half - ik
        sk = raw if raw >= 0 else raw + n_k
        e = np.exp(1j * (-2 * np.pi * ((sk * t) % n_k) / n_k))
        z = e @ spec
        for jf in range(n_freq):
            out[ik * n_freq + jf] = math.hypot(z[jf].real, z[jf].imag)
    return out


def test_dip_sign_and_axes():
    """tfk.nim:23-65: a cosine on time bin 8 delayed by one sample per trace
    peaks two bins to the positive side of zero; nK 16, nFreq 33."""
    rows, cols, dt = 8, 64, 2.0
    p = O.fk_plan(rows, cols, dt, 0.0)
    assert p["nK"] == 16 and p["nFreq"] == 33
    assert abs(p["dfHz"] - 1000 / (64 * dt)) <= 1e-12 and abs(p["dkPerTrace"] - 1 / 16) <= 1e-15
    f = p["dfHz"] * 8
    r = np.arange(rows)[:, None]
    j = np.arange(cols)[None, :]
    x = np.cos(2 * np.pi * f * (j - r) * dt / 1000)
    amp, used = O.fk_amplitude(p, x, np.full(rows, cols), 0)
    assert used == rows
    col = amp.reshape(16, 33)[:, 8]
    zero = 8
    assert int(np.argmax(col)) == zero + 2
    assert col[zero - 2] < 0.05 * col.max()
    for bad in ((3, 64), (8, 7)):
        with pytest.raises(ValueError):
            O.fk_plan(bad[0], bad[1], dt)


@pytest.mark.parametrize("rows,cols,seed,scale", [
    (4, 8, 1, 1.0), (4, 8, 1, 1e155), (4, 8, 1, 1e-160), (4, 8, 1, 1e-200),
    (7, 16, 2, 1.0), (9, 16, 3, 1.0), (17, 16, 4, 1.0), (24, 50, 5, 1.0)])
def test_against_the_slow_dft(rows, cols, seed, scale):
    """tfk.nim:162-213: within 1e-6 of the peak of the slow DFT, no
    infinities, not flushed to zeros, at scales whose squares overflow,
    lose digits or flush; trace counts that leave a lane remainder."""
    rng = np.random.default_rng(seed)
    x = scale * rng.uniform(-1, 1, (rows, cols))
    p = O.fk_plan(rows, cols, 2.0, 0.1)
    amp, used = O.fk_amplitude(p, x, np.full(rows, cols), 0)
    assert used == rows
    want = direct_fk(x, 0.1)
    peak = want.max()
    assert 0 < peak < math.inf
    assert not np.any(np.isinf(amp)) and np.any(amp != 0)
    ok = (amp > 0) & (want > 0)
    assert np.max(np.abs(amp[ok] - want[ok])) / peak <= 1e-6


CLASSES = [
    "NNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNN",
    "NNNNNNNNNNNNNNNNNNNNINNNNNNNNNNNNNNNNNNN",
    "NNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNN",
    "INNNINNNNNNNNNNNNNNNINNNINNNNNNNNNNNNNNN",
    "NNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNN",
    "NNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNNN"]


def class_panel(kind: int) -> np.ndarray:
    x = np.ones((4, 8))
    inf, nan = math.inf, math.nan
    for r in range(4):
        for j in range(8):
            x[r, j] = [nan, inf, [inf, nan, -inf, 1.0][r], inf if (r, j) == (0, 0) else 1.0,
                       nan if (r, j) == (0, 0) else 1.0, inf if j % 2 == 0 else nan][kind]
    return x


def classes(amp: np.ndarray) -> str:
    out = ""
    for a in amp:
        out += ("N" if math.isnan(a) else "I" if a == math.inf else "i" if a == -math.inf
                else "z" if a == 0 else "f")
    return out


@pytest.mark.parametrize("kind", range(6))
def test_nan_inf_classes(kind):
    """tfk.nim:215-245: the class of every output of a 4 x 8 float64 gather
    (taper 0.1) holding NaN and infinities, as the reference's transform
    propagates them."""
    p = O.fk_plan(4, 8, 2.0, 0.1)
    amp, _ = O.fk_amplitude(p, class_panel(kind), np.full(4, 8), 0)
    assert classes(amp) == CLASSES[kind]


def test_rows_that_enter():
    """fkRows (fft/fk.nim:204-216): live rows with a sample at firstCol or
    later enter, in order; the others are left out; the count must match
    the plan."""
    rng = np.random.default_rng(9)
    x = rng.normal(size=(10, 40))
    row_len = np.array([40, -1, 5, 40, 12, -1, 40, 40, 30, 40])
    p = O.fk_plan(7, 16, 2.0, 0.1)
    amp, used = O.fk_amplitude(p, x, row_len, 10)
    assert used == 7
    keep = [0, 3, 4, 6, 7, 8, 9]  # live, with a sample at column 10 or later
    sub = O.fk_amplitude(p, x[keep], row_len[keep], 10)[0]
    assert np.array_equal(amp, sub)
    with pytest.raises(ValueError):
        O.fk_amplitude(O.fk_plan(6, 16, 2.0), x, row_len, 10)
