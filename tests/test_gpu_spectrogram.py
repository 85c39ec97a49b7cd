"""Spectrograms on the GPU (fft/spectrogram.nim through SpectrogramPlan)
against the oracle's float64 restatement: random data with frames the
trace's ends cut, a sinusoid on a bin, traces shorter than the plan (and
empty), a NaN confined to its frames, several windows, hops and nfft (powers
of two and 2·3·5-smooth), and the plan's refusals.  float64 within 1e-10 of
max(1, want) per bin (the reference test's bound); float32 within 2e-5 of
the largest amplitude."""
import numpy as np
import pytest

import processseg_amd as P

pytestmark = pytest.mark.gpu


def run(sg, x):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    amp = sg.spectrogramInto(t)
    torch.cuda.synchronize()
    return amp.cpu().numpy()


def check(got, want, dtype, label):
    assert got.shape == want.shape, label
    if dtype == np.float64:
        bad = np.abs(got - want) > 1e-10 * np.maximum(1.0, np.abs(want))
    else:
        bad = np.abs(got - want) > 2e-5 * max(1.0, float(np.max(np.abs(want))))
    assert not bad.any(), f"{label}: {np.argwhere(bad)[:5].tolist()}"


CASES = [  # window, hop, dt, samples, nfft, trace length
    (41, 7, 2.0, 300, 0, 300),
    (127, 25, 2.0, 2000, 128, 2000),
    (64, 16, 1.0, 1000, 0, 600),
    (255, 50, 1.0, 6000, 0, 6000),
    (257, 10, 1.0, 8001, 0, 8001),
    (100, 33, 4.0, 3001, 120, 2500),
    (500, 128, 2.0, 3001, 1000, 3001),
    (2, 1, 1.0, 50, 0, 50),
    (1024, 200, 0.5, 4000, 0, 1),
    (64, 16, 1.0, 1000, 0, 0),
]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", CASES, ids=[f"w{c[0]}h{c[1]}n{c[4]}l{c[5]}" for c in CASES])
def test_spectrogram(gpu, oracle, dtype, case):
    W, hop, dt, samples, nfft, n = case
    x = np.random.default_rng(W + n).standard_normal(n).astype(dtype)
    sg = P.initSpectrogramPlan(dtype, W, hop, dt, samples, nfft)
    want = oracle.spectrogram(x.astype(np.float64), W, hop, dt, samples, nfft)
    assert (sg.frames, sg.bins) == want.shape
    check(run(sg, x), want, dtype, str(case))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_known_answers(gpu, oracle, dtype):
    """tspectrogram.nim:57-76 and 113-127 on the GPU: a sinusoid of amplitude 3
    on bin 20, a trace shorter than the plan, a NaN in its frames only."""
    n = 2000
    f = 78.125
    x = (3 * np.sin(2 * np.pi * f * np.arange(n) * 0.002 + 0.4)).astype(dtype)
    sg = P.initSpectrogramPlan(dtype, 127, 25, 2.0, n, nfft=128)
    assert abs(sg.frequency(20) - f) < 1e-12 and sg.centre(3) == 75
    amp = run(sg, x)
    tol = 3e-4 if dtype == np.float64 else 1e-3
    for i in range(sg.frames):
        c = sg.centre(i)
        lim = tol if (c >= 63 and c + 63 < n) else 0.15
        assert abs(amp[i, 20] - 3) < lim, (i, amp[i, 20])
    sg = P.initSpectrogramPlan(dtype, 64, 16, 1.0, 1000)
    y = np.sin(2 * np.pi * 125 * np.arange(600) / 1000).astype(dtype)
    amp = run(sg, y)
    assert abs(amp[10, 8] - 1) < 1e-3 and amp[-1, 8] == 0
    y[160] = np.nan
    amp = run(sg, y)
    assert np.isnan(amp[10, 8]) and not np.isnan(amp[20, 8])
    want = oracle.spectrogram(y.astype(np.float64), 64, 16, 1.0, 1000)
    fin = ~np.isnan(want)
    assert np.array_equal(np.isnan(amp), ~fin)
    check(np.where(fin, amp, 0), np.where(fin, want, 0), dtype, "nan")


def test_refusals(gpu):
    import torch
    for args, kw in (((1, 10, 1, 100), {}), ((64, 0, 1, 100), {}), ((64, 10, 0, 100), {}),
                     ((64, 10, 1, 0), {}), ((64, 10, 1, 100), {"nfft": 32}),
                     ((64, 10, 1, 100), {"nfft": 77})):
        with pytest.raises(ValueError):
            P.initSpectrogramPlan(np.float32, *args, **kw)
    sg = P.initSpectrogramPlan(np.float32, 64, 10, 1.0, 100)
    with pytest.raises(ValueError):
        sg.spectrogramInto(torch.zeros(101, device="cuda"))
    with pytest.raises(ValueError):
        sg.spectrogramInto(torch.zeros(100, device="cuda", dtype=torch.float64))
    with pytest.raises(ValueError):
        sg.spectrogramInto(torch.zeros(100, device="cuda"),
                           torch.zeros(3, device="cuda", dtype=torch.float64))
