"""Pins the oracle's real FFT (fft/plan.nim) against the reference's own
tests (tfftplan.nim:46-190): the nextFastLen table, a naive float64 DFT with
exactly reduced phases, numpy.fft, known spectra, round trips."""
import math

import numpy as np
import pytest


def smooth_lengths(limit):
    out = []
    for n in range(1, limit + 1):
        k = n
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            out.append(n)
    return out


def ref_dft(x):
    """A naive DFT in float64 with the phase reduced exactly (j k mod n), the
    reference test's oracle (tfftplan.nim:10-29)."""
    n = x.shape[0]
    j = np.arange(n)
    bins = n // 2 + 1
    k = np.arange(bins)
    e = (np.outer(k, j) % n).astype(np.float64)
    a = -2 * np.pi * e / n
    return (np.cos(a) @ x, np.sin(a) @ x)


def test_fast_lengths(oracle):
    assert oracle.is_fast_len(1) and oracle.is_fast_len(2) and oracle.is_fast_len(2025)
    assert oracle.is_fast_len(4096)
    assert not oracle.is_fast_len(0) and not oracle.is_fast_len(7) and not oracle.is_fast_len(2001)
    assert oracle.next_fast_len(2001) == 2048 and oracle.next_fast_len(2001, even=False) == 2025
    assert oracle.next_fast_len(0) == 2 and oracle.next_fast_len(0, even=False) == 1
    assert oracle.next_fast_len(7) == 8 and oracle.next_fast_len(11) == 12
    assert oracle.next_fast_len(13) == 16
    assert oracle.next_fast_len(14, even=False) == 15
    assert oracle.next_fast_len(4097) == 4320
    smooth = set(smooth_lengths(6000))
    for n in range(1, 5001):
        e = oracle.next_fast_len(n)
        assert e in smooth and e % 2 == 0 and e >= n
        assert not any(k in smooth and k % 2 == 0 for k in range(n, e))


def test_library_fast_lengths(oracle):
    import processseg_amd as P
    for n in list(range(0, 3000)) + [4097, 16065, 10001 + 6000 + 64]:
        assert P.nextFastLen(n) == oracle.next_fast_len(n)
        assert P.nextFastLen(n, even=False) == oracle.next_fast_len(n, even=False)


def test_factorize(oracle):
    """fft/plan.nim:120-132: 4s first, then one 2, then 3s and 5s."""
    assert oracle.factorize(1125) == [3, 3, 5, 5, 5]     # nfft 2250
    assert oracle.factorize(8100) == [4, 3, 3, 3, 3, 5, 5]  # nfft 16200
    assert oracle.factorize(1536) == [4, 4, 4, 4, 2, 3]
    assert oracle.factorize(2) == [2] and oracle.factorize(1) == []


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
def test_naive_dft(oracle, dtype, tol):
    """Every 2·3·5-smooth n <= 1200 against the naive DFT, plus the round trip
    (tfftplan.nim:64-94)."""
    rng = np.random.default_rng(17)
    worst = worst_rt = 0.0
    for n in smooth_lengths(1200):
        x64 = rng.standard_normal(n)
        x = x64.astype(dtype)
        rr, ri = ref_dft(x.astype(np.float64))
        re, im = oracle.rfft(x)
        e = np.max(np.hypot(re - rr, im - ri)) / max(np.max(np.hypot(rr, ri)), 1e-300)
        assert e <= tol, (n, e)
        back = oracle.irfft(re, im, n)
        rt = np.max(np.abs(back.astype(np.float64) - x.astype(np.float64)))
        assert rt <= tol * 4, (n, rt)
        worst, worst_rt = max(worst, e), max(worst_rt, rt)
    print(dtype, "worst DFT error", worst, "round trip", worst_rt)


@pytest.mark.parametrize("n", [2250, 3456, 3072, 16200, 4320, 60000])
def test_numpy_fft(oracle, n):
    """float64 against numpy.fft.rfft / irfft (pocketfft), and float32 within
    1e-5 of float64 (tfftplan.nim:122-173)."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    re, im = oracle.rfft(x)
    X = np.fft.rfft(x)
    peak = np.max(np.abs(X))
    assert np.max(np.abs(re + 1j * im - X)) <= 1e-12 * peak
    y = oracle.irfft(X.real.copy(), X.imag.copy(), n)
    assert np.max(np.abs(y - np.fft.irfft(X, n))) <= 1e-12
    re32, im32 = oracle.rfft(x.astype(np.float32))
    assert np.max(np.abs(re32 + 1j * im32 - X)) <= 1e-5 * peak


def test_known_spectra(oracle):
    """tfftplan.nim:96-110: an impulse is flat; cos(2 pi 3 j / 16) is 8 at
    bin 3."""
    x = np.zeros(16)
    x[0] = 1
    re, im = oracle.rfft(x)
    assert np.all(np.abs(re - 1) < 1e-15) and np.all(np.abs(im) < 1e-15)
    x = np.cos(2 * np.pi * 3 * np.arange(16) / 16)
    re, im = oracle.rfft(x)
    want = np.zeros(9)
    want[3] = 8
    assert np.all(np.abs(re - want) < 1e-13) and np.all(np.abs(im) < 1e-13)


def test_inverse_ignores_edge_imaginary(oracle):
    """tfftplan.nim:175-190: the inverse takes bins 0 and n/2 as real."""
    rng = np.random.default_rng(3)
    for n in (16, 30, 2250):
        re = rng.standard_normal(n // 2 + 1)
        im = rng.standard_normal(n // 2 + 1)
        a = oracle.irfft(re, im, n)
        im2 = im.copy()
        im2[0] = 0
        im2[-1] = 0
        b = oracle.irfft(re, im2, n)
        assert np.max(np.abs(a - b)) <= 1e-14
