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
                                                  (128, 0, 0.1, 1800), (4000, 0, 0.2, 0),
                                                  (100, 125, 0.1, 9), (1001, 1125, 0.2, 50),
                                                  (243, 243, 0.0, 3)])
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
    assert P.initSpectrumPlan(np.float32, 2.0, 100, nfft=125).bins == 63   # odd smooth length


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("win,nfft,taper,first", [(501, 0, 0.1, 250), (400, 400, 0.2, 0),
                                                  (1000, 1000, 0.0, 37), (100, 135, 0.1, 7),
                                                  (1001, 1125, 0.0, 20)])
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


def test_average_is_bit_reproducible(gpu, oracle):
    """The rows' powers are summed in an order fixed by the row count and the
    plan (the pool form's fixed batches, core/pool.nim:776-800): two calls
    on the 2000 x 3001 panel, and a call with a caller workspace, give the
    same bits."""
    import torch
    from processseg_amd.workloads import big_panel
    x, live = big_panel(2000, 3001, np.float32)
    src = torch.from_numpy(x).cuda()
    rl = torch.from_numpy(live.astype(np.int32)).cuda()
    sp = P.initSpectrumPlan(np.float32, 2.0, 2001, 0.1)
    a, ua = sp.averageAmplitudeInto(src, firstCol=500, row_len=rl)
    b, ub = sp.averageAmplitudeInto(src, firstCol=500, row_len=rl)
    ws = torch.empty(sp.workspaceSize(2000), dtype=torch.uint8, device="cuda")
    c, uc = sp.averageAmplitudeInto(src, firstCol=500, row_len=rl, ws=ws)
    used = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    d, ud = sp.averageAmplitudeInto(src, firstCol=500, row_len=rl, ws=ws, used=used, sync=False)
    torch.cuda.synchronize()
    assert ud is used and used.item() == 2000
    assert ua == ub == uc == 2000
    a, b, c, d = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy(), d.cpu().numpy()
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    assert np.array_equal(a.view(np.uint64), c.view(np.uint64))
    assert np.array_equal(a.view(np.uint64), d.view(np.uint64))
    p = oracle.spectrum_plan(2.0, 2001, 0.1, 0)
    rows = np.arange(0, 2000, 50)
    want, _ = oracle.spectrum_average(p, x[rows].astype(np.float64), live[rows].astype(np.int32),
                                      500)
    sub, _ = sp.averageAmplitudeInto(src[rows].contiguous(), firstCol=500, row_len=rl[rows].contiguous())
    assert rel_l2(sub.cpu().numpy(), want) <= 1e-5


def test_average_on_two_side_streams(gpu, oracle):
    """Two plans' averages queued on two side streams at once (each call its
    own workspace, `used` read after its stream): each equals the oracle."""
    import torch
    rows, cols = 300, 2001
    x, _ = synth.gather(rows, cols, 2.0, 17, dtype=np.float32)
    rl = np.full(rows, cols, np.int32)
    src = torch.from_numpy(x).cuda()
    rlt = torch.from_numpy(rl).cuda()
    plans = [P.initSpectrumPlan(np.float32, 2.0, 1001, 0.1), P.initSpectrumPlan(np.float32, 2.0, 500, 0.2)]
    firsts = [100, 900]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    outs = []
    for _ in range(3):
        for sp, f, s in zip(plans, firsts, streams):
            outs.append((sp, f, s, sp.averageAmplitudeInto(src, firstCol=f, row_len=rlt, stream=s,
                                                           sync=False)))
    for sp, f, s, (amp, used) in outs:
        s.synchronize()
        p = oracle.spectrum_plan(2.0, sp.windowLen, 0.1 if sp.windowLen == 1001 else 0.2, 0)
        want, uw = oracle.spectrum_average(p, x.astype(np.float64), rl, f)
        assert int(used.item()) == uw == rows
        assert rel_l2(amp.cpu().numpy(), want) <= 1e-5


def test_wrappers_refuse_before_launch(gpu):
    """The Python wrappers check what the kernels index without bounds checks
    (the reference's ValueError for a block that does not match its meta):
    a meta or row_len of another row count, a row_len of another dtype, a dst
    of another shape or dtype, an amp off the device."""
    import torch
    rows, cols = 8, 300
    x = torch.zeros(rows, cols, device="cuda")
    good = torch.full((rows,), cols, dtype=torch.int32, device="cuda")
    bad_rl = [torch.full((rows - 1,), cols, dtype=torch.int32, device="cuda"),
              torch.full((rows,), cols, dtype=torch.int64, device="cuda"),
              torch.full((rows,), cols, dtype=torch.int32)]
    short_meta = P.ChunkMeta(rows - 1, 2.0)
    sp = P.initSpectrumPlan(np.float32, 2.0, 128, 0.1)
    hp = P.initHilbertPlan(np.float32, cols)
    calls = []
    for rl in bad_rl:
        calls += [lambda rl=rl: sp.averageAmplitudeInto(x, row_len=rl),
                  lambda rl=rl: sp.amplitudeInto(x, row_len=rl),
                  lambda rl=rl: hp.hilbertInto(x, row_len=rl),
                  lambda rl=rl: hp.envelopeInto(x, row_len=rl)]
    calls += [lambda: sp.averageAmplitudeInto(x, short_meta),
              lambda: sp.amplitudeInto(x, short_meta),
              lambda: sp.amplitudeInto(x, P.ChunkMeta(rows, 4.0)),
              lambda: hp.hilbertInto(x, short_meta),
              lambda: hp.hilbertInto(x, torch.zeros(rows, cols - 1, device="cuda"), row_len=good),
              lambda: hp.hilbertInto(x, torch.zeros(rows, cols, dtype=torch.float64, device="cuda"),
                                     row_len=good),
              lambda: hp.envelopeInto(x, torch.zeros(rows, cols), row_len=good),
              lambda: sp.amplitudeInto(x, row_len=good, amp=torch.zeros(rows, sp.bins)),
              lambda: sp.averageAmplitudeInto(x, row_len=good, firstCol=-1),
              lambda: sp.averageAmplitudeInto(x, row_len=good,
                                              used=torch.zeros(1, dtype=torch.float64)),
              lambda: sp.averageAmplitudeInto(x, row_len=good,
                                              used=torch.zeros(1, device="cuda"))]
    for i, c in enumerate(calls):
        with pytest.raises(ValueError):
            c()
    sg = P.initSpectrogramPlan(np.float32, 64, 10, 1.0, 200)
    with pytest.raises(ValueError):
        sg.spectrogramInto(torch.zeros(200, device="cuda"), amp=torch.zeros(sg.frames, sg.bins,
                                                                            dtype=torch.float64))
    torch.cuda.synchronize()
