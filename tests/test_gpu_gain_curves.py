"""Time-variant gain on the GPU (gain/curves.nim through pipeline.nim's
gainCurveOp) against the oracle: knot, power and exponential curves and a
held curve, rows on and off the table's grid, starting before it and running
past it (the factors there from the curve itself, on the device), NaN t0,
rows not live, live lengths, a window's firstSample, zeros under an infinite
factor, and a gain curve inside a filter / AGC flow.  Float32 within 1e-5
relative L2 of the float64 oracle, float64 within 1e-12; samples past a live
end and rows not live unchanged, bit for bit."""
import math

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import synth
from tests.common import rel_l2

pytestmark = pytest.mark.gpu

ROWS, COLS, DT = 48, 1500, 2.0

CURVES = {
    "knots": P.knotGain([(100.0, 0.0), (1100.0, 20.0), (2100.0, 14.0)]),
    "power": P.powerGain(2.0, refMs=1000.0, minMs=20.0),
    "inverse": P.powerGain(-1.0, refMs=100.0),
    "exponential": P.exponentialGain(12.0),
    "held": P.heldGain(P.powerGain(1.5, minMs=10.0), untilMs=2500.0),
}


def survey(dtype, seed=5):
    x, _ = synth.gather(ROWS, COLS, DT, seed, dtype=dtype)
    live = synth.live_lengths(seed, ROWS, 400, COLS).astype(np.int32)
    live[0] = COLS
    t0 = np.zeros(ROWS)
    t0[1::6] = 100.0            # on the grid, later
    t0[2::6] = 51.0             # off the grid
    t0[3::6] = -90.5            # before the table, off the grid
    t0[4::6] = 2500.0           # past the table: every factor from the curve
    t0[13] = math.nan           # the headers cannot give a t0
    status = np.zeros(ROWS, np.uint8)
    status[7] = P.tsDead
    offsets = np.linspace(0, 3000, ROWS)
    return x, live, t0, status, offsets


def run(flow, x, meta, dtype, inplace=False, record=0):
    import torch
    plan = flow.compile(dtype, DT, COLS, recordSamples=record)
    src = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    if inplace:
        plan.apply(src, meta)
        got = src
    else:
        got = torch.empty_like(src)
        plan.apply(src, got, meta)
    torch.cuda.synchronize()
    return got.cpu().numpy()


def check(got, want, x, row_len, dtype, label):
    tol = 1e-5 if dtype == np.float32 else 1e-12
    for r in range(got.shape[0]):
        n = int(row_len[r])
        if n < 0:
            assert np.array_equal(got[r], x[r]), f"{label}: row {r} not live, not copied"
            continue
        assert np.array_equal(got[r, n:], x[r, n:]), f"{label}: row {r} past its live end"
        if n == 0:
            continue
        e = rel_l2(got[r, :n], want[r, :n])
        assert e <= tol, f"{label}: row {r}: rel L2 {e:.3e}"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("curve", sorted(CURVES))
@pytest.mark.parametrize("extent", ["record", "short", "wide"])
def test_gain_curves(gpu, oracle, dtype, curve, extent):
    """One gainCurveOp: the default table (0 ms to the record's end), a short
    one (most samples off it) and a wide one (-200 ms .. 6000 ms)."""
    x, live, t0, status, offsets = survey(dtype)
    meta = P.ChunkMeta(ROWS, DT, numSamples=live, status=status, t0Ms=t0, offset=offsets)
    row_len = meta.row_len(COLS)
    kw = {"record": {}, "short": {"tMaxMs": 300.0}, "wide": {"tMinMs": -200.0, "tMaxMs": 6000.0}}
    op = P.gainCurveOp(CURVES[curve], **kw[extent])
    flow = P.initPipeline([op])
    got = run(flow, x, meta, dtype)
    want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), row_len, DT, keys=meta.keys())
    check(got, want, x, row_len, dtype, f"{curve} {extent}")
    assert np.array_equal(got[13], x[13]), "a NaN t0 leaves the row alone"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_gain_in_a_flow_and_a_window(gpu, oracle, dtype):
    """demean, band-pass, spherical divergence (t²), AGC, an exponential gain,
    on a window from trace sample 200 of traces of 2400 samples (the table to
    the record's end, recordSamples), src -> dst and in place."""
    x, live, t0, status, offsets = survey(dtype, seed=9)
    meta = P.ChunkMeta(ROWS, DT, numSamples=live + 200, status=status, t0Ms=t0, offset=offsets,
                       firstSample=200)
    row_len = meta.row_len(COLS)
    flow = P.initPipeline([
        P.demeanOp(), P.bandPassOp(P.Corners(4, 8, 60, 80)),
        P.gainCurveOp(P.powerGain(2.0)), P.agcOp(400),
        P.gainCurveOp(P.exponentialGain(-3.0))])
    got = run(flow, x, meta, dtype, record=2400)
    want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), row_len, DT,
                                 record_samples=2400, keys=meta.keys(), first_sample=200)
    check(got, want, x, row_len, dtype, "flow")
    again = run(flow, x, meta, dtype, inplace=True, record=2400)
    assert np.array_equal(got, again), "in place differs from src -> dst"


def test_zero_stays_zero_under_an_infinite_factor(gpu, oracle):
    """An exponential gain past float32 off the table: zeros stay zero,
    other samples become Inf, as the oracle has them (bit for bit where the
    factor is infinite)."""
    x = np.zeros((4, 300), np.float32)
    x[1] = 1.0
    x[2, ::2] = -2.0
    x[3, 100:] = 0.5
    meta = P.ChunkMeta(4, DT)
    flow = P.initPipeline([P.gainCurveOp(P.exponentialGain(2000), tMaxMs=100)])
    got = run(flow, x, meta, np.float32)
    want = oracle.pipeline_apply(flow.ops, x, meta.row_len(300), DT)
    assert np.all(got[0] == 0)
    assert got[1, 299] == np.inf and got[2, 298] == -np.inf
    inf = ~np.isfinite(want)
    assert np.array_equal(got[inf], want[inf])
    assert np.array_equal(got[x == 0], x[x == 0])
    fin = np.isfinite(want) & (x != 0)
    assert np.max(np.abs(got[fin] / want[fin] - 1)) <= 1e-6


def test_gain_without_keys(gpu, oracle):
    """apply(row_len=...) with no keys: every row starts at t0 0."""
    import torch
    x, live, _, _, _ = survey(np.float32)
    rl = live.astype(np.int32)
    flow = P.initPipeline([P.gainCurveOp(CURVES["knots"])])
    plan = flow.compile(np.float32, DT, COLS)
    src = torch.from_numpy(x).cuda()
    dst = torch.empty_like(src)
    plan.apply(src, dst, row_len=torch.from_numpy(rl).cuda())
    got = dst.cpu().numpy()
    want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), rl, DT)
    check(got, want, x, rl, np.float32, "no keys")


def test_gain_refusals(gpu):
    steep = P.knotGain([(0.0, 0.0), (1000.0, 800.0)])
    with pytest.raises(ValueError, match="overflows"):
        P.initPipeline([P.gainCurveOp(steep)]).compile(np.float32, DT, 501)
    P.initPipeline([P.gainCurveOp(steep)]).compile(np.float64, DT, 501)
    with pytest.raises(ValueError, match="no knots"):
        P.initPipeline([P.gainCurveOp(P.GainCurve())]).compile(np.float32, DT, 100)
    with pytest.raises(ValueError, match="before"):
        P.initPipeline([P.gainCurveOp(P.powerGain(2), tMinMs=10, tMaxMs=5)]).compile(
            np.float32, DT, 100)
    with pytest.raises(ValueError, match="strictly increase"):
        raw = P.GainCurve(P.gcKnots, knots=((1.0, 0.0), (1.0, 3.0)))
        P.initPipeline([P.gainCurveOp(raw)]).compile(np.float32, DT, 100)
