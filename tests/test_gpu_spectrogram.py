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
from processseg_amd import synth
from tests.common import rel_l2

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


def _avg_case(dtype, rows, cols, seed):
    x, _ = synth.gather(rows, cols, 1.0, seed, dtype=dtype)
    rl = synth.live_lengths(seed, rows, 1, cols).astype(np.int32)
    rl[0] = cols
    rl[1] = -1          # not live
    rl[2] = 0           # no sample
    x[3, min(5, cols - 1)] = np.nan
    rl[3] = cols        # a NaN: left out
    return x, rl


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("win,hop,nfft,cols,rows", [(257, 10, 0, 2001, 40), (64, 64, 0, 1000, 300),
                                                    (32, 16, 32, 128, 7), (100, 7, 125, 900, 25)])
def test_average_spectrogram(gpu, oracle, dtype, win, hop, nfft, cols, rows):
    """averageSpectrogramInto (fft/spectrogram.nim:370-430) against the
    oracle: float32 within 1e-5 relative L2 of the float64 oracle, float64
    within 1e-12; the rows used exactly (dead, empty and NaN rows left out;
    short live lengths); the same bits on a second call."""
    import torch
    x, rl = _avg_case(dtype, rows, cols, win + hop)
    sg = P.initSpectrogramPlan(dtype, win, hop, 1.0, cols, nfft)
    src = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    rlt = torch.from_numpy(rl).cuda()
    amp, used = sg.averageSpectrogramInto(src, row_len=rlt)
    amp2, used2 = sg.averageSpectrogramInto(src, row_len=rlt)
    got = amp.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), amp2.cpu().numpy().view(np.uint64))
    want, uw = oracle.spectrogram_average(x, rl, win, hop, 1.0, cols, nfft, dtype=dtype)
    assert used == used2 == uw == rows - 3
    tol = 1e-5 if dtype == np.float32 else 1e-12
    assert rel_l2(got, want) <= tol, rel_l2(got, want)


def test_average_spectrogram_reference_cases(gpu, oracle):
    """tspectrogram.nim:179-234, 286-299, 410-446 on the GPU: an infinity in
    a covered sample drops its row; float64 rows with a NaN or an infinity
    between frames, or of subnormals, are left out; a float32 row of 1e30
    stays in (finite); no row used: all zeros."""
    import torch
    f = 4 * 1000 / (32 * 2.0)
    j = np.arange(128)
    base = np.cos(2 * np.pi * f * j * 2 / 1000) * 2
    for dtype in (np.float32, np.float64):
        x = np.stack([base, base, base, np.full(128, 1e30)]).astype(dtype)
        x[1, 20] = np.inf
        sg = P.initSpectrogramPlan(dtype, 32, 16, 2.0, 128, 32)
        amp, used = sg.averageSpectrogramInto(torch.from_numpy(x).cuda(),
                                              P.ChunkMeta(4, 2.0, numSamples=np.full(4, 128)))
        want, uw = oracle.spectrogram_average(x, np.full(4, 128), 32, 16, 2.0, 128, 32, dtype=dtype)
        if dtype == np.float64:
            assert uw == 3      # 1e30 is inside the float64 range
        assert used == uw
        assert rel_l2(amp.cpu().numpy(), want) <= (1e-5 if dtype == np.float32 else 1e-12)
    gap = P.initSpectrogramPlan(np.float64, 2, 10, 1.0, 40, 4)
    for hole in (np.nan, np.inf, -np.inf):
        g = np.full((1, 40), 0.25)
        g[0, 5] = hole
        amp, used = gap.averageSpectrogramInto(torch.from_numpy(g).cuda(), row_len=torch.tensor(
            [40], dtype=torch.int32, device="cuda"))
        assert used == 0 and torch.all(amp == 0)
    amp, used = gap.averageSpectrogramInto(torch.full((1, 40), 1e-320, dtype=torch.float64,
                                                      device="cuda"),
                                           row_len=torch.tensor([40], dtype=torch.int32,
                                                                device="cuda"))
    assert used == 0 and torch.all(amp == 0)
    # a float32 NaN between frames (no frame covers column 5) keeps its row
    g32 = np.full((1, 40), 0.25, np.float32)
    g32[0, 5] = np.nan
    gp32 = P.initSpectrogramPlan(np.float32, 2, 10, 1.0, 40, 4)
    amp, used = gp32.averageSpectrogramInto(torch.from_numpy(g32).cuda(), row_len=torch.tensor(
        [40], dtype=torch.int32, device="cuda"))
    want, uw = oracle.spectrogram_average(g32, [40], 2, 10, 1.0, 40, 4, dtype=np.float32)
    assert used == uw == 1 and rel_l2(amp.cpu().numpy(), want) <= 1e-5


def test_average_spectrogram_full_panel(gpu, oracle):
    """The bench's workload: the averaged spectrogram of the 2000 x 3001 f32
    panel (257-sample Hann frames every 10 samples, nfft 512, one wave per
    frame), within 1e-5 relative L2 of the float64 oracle, the same bits on a
    second call."""
    import torch
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    rl = np.minimum(live, 3001).astype(np.int32)
    sg = P.initSpectrogramPlan(np.float32, 257, 10, 2.0, 3001)
    src = torch.from_numpy(x).cuda()
    rlt = torch.from_numpy(rl).cuda()
    amp, used = sg.averageSpectrogramInto(src, row_len=rlt)
    amp2, _ = sg.averageSpectrogramInto(src, row_len=rlt)
    got = amp.cpu().numpy()
    assert np.array_equal(got.view(np.uint64), amp2.cpu().numpy().view(np.uint64))
    want, uw = oracle.spectrogram_average(x, rl, 257, 10, 2.0, 3001, 0, dtype=np.float32)
    assert used == uw == 2000
    assert rel_l2(got, want) <= 1e-5, rel_l2(got, want)
