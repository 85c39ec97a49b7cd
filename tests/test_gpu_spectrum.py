"""Averaged amplitude spectra on the GPU (fft/spectrum.nim averageAmplitudeInto)
against the oracle: float32 within 1e-5 relative L2 of the float64 oracle,
float64 within 1e-12; the rows used exactly; dead, failed and short rows,
windows cut by the live end, NaN and Inf rows left out, nfft powers of two
and 2·3·5-smooth lengths, and the reference's sqrt(21/3) average."""
import math

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import synth
from tests.common import rel_l2

pytestmark = pytest.mark.gpu


def run(sp, x, rl, first):
    import torch
    amp, used = sp.averageAmplitudeInto(torch.from_numpy(np.ascontiguousarray(x)).cuda(),
                                        firstCol=first, row_len=torch.from_numpy(rl).cuda())
    return amp.cpu().numpy(), used


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("win,nfft,taper,first", [(501, 0, 0.1, 250), (400, 400, 0.5, 0),
                                                  (1000, 1000, 0.0, 37), (2001, 0, 0.25, 100),
                                                  (128, 0, 0.1, 1800), (4000, 0, 0.2, 0)])
def test_average_spectrum(gpu, oracle, dtype, win, nfft, taper, first):
    rows, cols = 40, 4001
    x, _ = synth.gather(rows, cols, 2.0, win + first, dtype=dtype)
    x = x + dtype(0.3)
    rl = synth.live_lengths(win, rows, 0, cols).astype(np.int32)
    rl[0] = cols
    rl[3] = -1
    rl[5] = first            # nothing live in the window
    x[7, first + 3] = np.nan  # left out
    x[8, min(first + 20, cols - 1)] = np.inf
    rl[7] = rl[8] = cols
    sp = P.initSpectrumPlan(dtype, 2.0, win, taper, nfft)
    got, used = run(sp, x, rl, first)
    p = oracle.spectrum_plan(2.0, win, taper, nfft)
    want, used_want = oracle.spectrum_average(p, x.astype(np.float64), rl, first)
    assert used == used_want and sp.nfft == p["nfft"] and sp.bins == p["bins"]
    tol = 1e-5 if dtype == np.float32 else 1e-12
    assert rel_l2(got, want) <= tol, rel_l2(got, want)


def test_reference_average(gpu, oracle):
    """tspectrum.nim:111-153 on the GPU: amplitudes 1..6 at 25 Hz, a dead
    and a failed row, a row live before the window: sqrt(21/3) on bin 20."""
    t = np.arange(600) * 2.0 / 1000
    x = np.stack([(r + 1) * np.sin(2 * np.pi * 25.0 * t) for r in range(6)]).astype(np.float32)
    rl = np.array([600, 600, -1, 600, -1, 100], np.int32)
    sp = P.initSpectrumPlan(np.float32, 2.0, 400, 0.5, 400)
    amp, used = run(sp, x, rl, 150)
    assert used == 3 and int(np.argmax(amp)) == 20
    assert abs(amp[20] - math.sqrt(21 / 3)) <= 1e-3
    empty, u0 = run(sp, x[2:3], rl[2:3], 150)
    assert u0 == 0 and np.all(empty == 0)


def test_spectrum_refusals(gpu):
    with pytest.raises(ValueError):
        P.initSpectrumPlan(np.float32, 2.0, 100, nfft=64)
    with pytest.raises(ValueError):
        P.initSpectrumPlan(np.float32, 2.0, 100, nfft=101)
    with pytest.raises(ValueError):
        P.initSpectrumPlan(np.float32, 2.0, 10, taperFraction=0.6)
    with pytest.raises(P.UnsupportedOp):
        P.initSpectrumPlan(np.float32, 2.0, 100, nfft=125)   # odd smooth length


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("win,nfft,taper,first", [(501, 0, 0.1, 250), (400, 400, 0.2, 0),
                                                  (1000, 1000, 0.0, 37)])
def test_amplitude_per_row(gpu, oracle, dtype, win, nfft, taper, first):
    """amplitudeInto row by row: the amplitude within the bar of the oracle,
    the phase where the amplitude is significant, NaN rows NaN, rows with
    nothing live in the window untouched."""
    import torch
    rows, cols = 16, 2001
    x, _ = synth.gather(rows, cols, 2.0, win, dtype=dtype)
    x = x + dtype(0.3)
    rl = synth.live_lengths(win + 1, rows, first + 1, cols).astype(np.int32)
    rl[2] = -1
    rl[4] = first
    x[6, first + 5] = np.nan
    rl[6] = cols
    sp = P.initSpectrumPlan(dtype, 2.0, win, taper, nfft)
    amp, ph = sp.amplitudeInto(torch.from_numpy(np.ascontiguousarray(x)).cuda(), firstCol=first,
                               row_len=torch.from_numpy(rl).cuda())
    amp, ph = amp.cpu().numpy(), ph.cpu().numpy()
    p = oracle.spectrum_plan(2.0, win, taper, nfft)
    tol = 1e-5 if dtype == np.float32 else 1e-12
    for r in range(rows):
        if rl[r] <= first:
            assert np.all(amp[r] == 0) and np.all(ph[r] == 0), f"row {r} untouched"
            continue
        last = min(int(rl[r]), first + win)
        wa, wp = oracle.spectrum_amplitude(p, x[r, first:last].astype(np.float64))
        if r == 6:
            assert np.all(np.isnan(amp[r])) and np.all(np.isnan(wa))
            continue
        assert rel_l2(amp[r], wa) <= tol, (r, rel_l2(amp[r], wa))
        big = wa > 1e-3 * wa.max()
        d = np.angle(np.exp(1j * (ph[r][big].astype(np.float64) - wp[big])))
        assert np.max(np.abs(d)) <= (1e-3 if dtype == np.float32 else 1e-9), r
