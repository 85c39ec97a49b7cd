"""appliedResponseInto (pipeline.nim:840-880) and its moduli
(responseModuli, pipeline.nim:774-838): the library's moduli against the
plain sine / cosine sum at the reference test's tolerance on the CPU
(tests/tpipeline.nim:447-523), then the applied response on the GPU: pass and
stop bands, the trend removal turned off, a scale's gain, a refused op, and
the moduli of the impulse response the plan gives."""
import ctypes as C
import math

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import _lib


def moduli(x, hz, dt_s):
    x = np.ascontiguousarray(x, np.float64)
    h = np.ascontiguousarray(hz, np.float64)
    g = np.zeros(h.size)
    _lib.check(_lib.lib().pseg_response_moduli(x.ctypes.data, x.size, h.ctypes.data, h.size, dt_s,
                                               g.ctypes.data))
    return g


def sine_cosine(x, hz, dt_s):
    """The reference test's sum (tests/tpipeline.nim:452-461), the angle of
    each sample computed afresh."""
    j = np.arange(len(x), dtype=np.float64)
    out = []
    for f in hz:
        p = -2 * math.pi * f * j * dt_s
        out.append(math.hypot(float(np.sum(x * np.cos(p))), float(np.sum(x * np.sin(p)))))
    return np.array(out)


def log_hz(nyq, n):
    """logHz (tests/tpipeline.nim:481-489)."""
    out = np.zeros(n)
    out[-1] = nyq
    lo, hi = nyq * 1e-6, nyq * (1 - 1e-9)
    for i in range(1, n - 1):
        out[i] = lo * (hi / lo) ** ((i - 1) / (n - 3))
    return out


@pytest.mark.parametrize("n", [64, 1000, 4001, 16384])
@pytest.mark.parametrize("dt_ms", [0.5, 2.0, 4.0])
def test_moduli_against_sine_cosine(n, dt_ms):
    """d <= 1e-12 + 1e-9 |want| over log and linear frequency grids up to
    Nyquist, for a decaying random trace of n samples."""
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) * np.exp(-np.abs(np.arange(n) - n // 2) / 40)
    nyq = 500 / dt_ms
    for hz in (log_hz(nyq, 31), np.linspace(0, nyq, 17), np.array([nyq * 0.37])):
        got = moduli(x, hz, dt_ms / 1000)
        want = sine_cosine(x, hz, dt_ms / 1000)
        assert np.all(np.abs(got - want) <= 1e-12 + 1e-9 * np.abs(want))


def test_moduli_refusals():
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().pseg_response_moduli(None, -1, None, 0, 0.002, None))


@pytest.mark.gpu
def test_applied_response_bands_and_trend(gpu):
    """tests/tpipeline.nim:403-445: a band-pass's pass and stop bands, the
    flow left as it was; a low cut applied without its trend removal (the
    kept trend would differ); scale 3 gives 3; an infinite scale is refused."""
    flow = P.initPipeline([P.bandPassOp(P.Corners(8, 12, 40, 50), trend=P.trLinear)])
    g = flow.appliedResponseInto(np.float64, 2.0, 256, [5.0, 25.0, 80.0])
    assert g[1] > 0.5 and g[0] < 0.05 and g[2] < 0.05
    assert flow[0].trend == P.trLinear
    low = P.initPipeline([P.filterOp(P.lowCut(4, 8), P.trLinear)])
    applied = low.appliedResponseInto(np.float64, 2.0, 256, [2.0, 20.0])
    kept = apply_impulse(low, np.float64, 2.0, 256, keep_trend=True)
    by_hand = sine_cosine(kept, [2.0, 20.0], 0.002)
    assert abs(applied[0] - by_hand[0]) > 1e-4
    assert abs(P.initPipeline([P.scaleOp(3)]).appliedResponseInto(np.float64, 2.0, 64, [10.0])[0]
               - 3) < 1e-9
    with pytest.raises(ValueError):
        P.initPipeline([P.scaleOp(math.inf)]).appliedResponseInto(np.float64, 2.0, 32, [1.0])


def apply_impulse(flow, T, dt_ms, samples, keep_trend=False):
    """impulseOut (tests/tpipeline.nim:463-477): the plan (trend removal off
    unless kept) applied on the GPU to a unit impulse at samples // 2."""
    import torch
    ops = [P.Op(**{**op.__dict__}) for op in flow.ops]
    if not keep_trend:
        for op in ops:
            if op.kind in (P.pipeline.opFilter, P.pipeline.opIir):
                op.trend = P.trNone
    plan = P.initPipeline(ops).compile(T, dt_ms, samples)
    x = np.zeros((1, samples), T)
    x[0, samples // 2] = 1
    t = torch.from_numpy(x).cuda()
    meta = P.ChunkMeta(1, dt_ms, numSamples=[samples], t0Ms=[0.0], offset=[0.0])
    plan.apply(t, meta)
    return t.cpu().numpy()[0].astype(np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_applied_response_is_the_moduli_of_the_impulse_response(gpu, T):
    """tests/tpipeline.nim:447-523: for a band-pass, an IIR band-pass and a
    gain with a low cut, at 0.5, 2 and 4 ms and 64 .. 4001 samples, the
    applied response equals the sine / cosine sum of the impulse response
    the plan gives, to 1e-12 + 1e-9 |want|."""
    flows = [P.initPipeline([P.bandPassOp(P.Corners(8, 12, 40, 50))]),
             P.initPipeline([P.iirOp(P.butterBandPass(4, 60, 24))]),
             P.initPipeline([P.gainDbOp(12), P.filterOp(P.lowCut(4, 8))])]
    for flow in flows:
        for dt_ms in (0.5, 2.0, 4.0):
            nyq = 500 / dt_ms
            for samples in (64, 1000, 4001):
                x = apply_impulse(flow, T, dt_ms, samples)
                for hz in (log_hz(nyq, 17), np.linspace(0, nyq, 15)):
                    got = flow.appliedResponseInto(T, dt_ms, samples, hz)
                    want = sine_cosine(x, hz, dt_ms / 1000)
                    assert np.all(np.abs(got - want) <= 1e-12 + 1e-9 * np.abs(want))
