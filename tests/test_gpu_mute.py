"""Mutes on the GPU (core/mute.nim through pipeline.nim's muteTopOp /
muteBottomOp) against the oracle: every curve kind, sharp and tapered, the
rows' keys (t0Ms, offset) and a window's firstSample from the ChunkMeta,
NaN keys, rows not live, live lengths, and mutes inside a filter / AGC flow.
Float32 within 1e-5 relative L2 of the float64 oracle (the tapers' sin² is
evaluated in float64 on both sides), float64 within 1e-12; samples past a
live end and rows not live unchanged, bit for bit."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import synth
from tests.common import rel_l2

pytestmark = pytest.mark.gpu

ROWS, COLS, DT = 40, 1500, 2.0


def survey(dtype, seed=3):
    x, _ = synth.gather(ROWS, COLS, DT, seed, dtype=dtype)
    x = x + dtype(0.25)
    live = synth.live_lengths(seed, ROWS, 600, COLS).astype(np.int32)
    live[0] = COLS
    offsets = np.linspace(-2400, 2400, ROWS)
    t0 = np.where(np.arange(ROWS) % 5 == 0, 40.0, 0.0)
    status = np.zeros(ROWS, np.uint8)
    status[7] = P.tsDead
    offsets[11] = np.nan       # an offset the headers cannot give
    t0[13] = np.nan            # likewise a t0
    return x, live, offsets, t0, status


CURVES = {
    "time": P.timeMute(420.0),
    "table": P.tableMute([0.0, 800, 2000], [150.0, 600, 1300]),
    "lmo": P.lmoMute(2500.0, t0Ms=30.0, padMs=-12.0),
}


def run(flow, x, meta, dtype, inplace=False):
    import torch
    plan = flow.compile(dtype, DT, COLS)
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
@pytest.mark.parametrize("taper", [0.0, 36.0])
def test_single_mutes(gpu, oracle, dtype, curve, taper):
    x, live, offsets, t0, status = survey(dtype)
    meta = P.ChunkMeta(ROWS, DT, numSamples=live, status=status, t0Ms=t0, offset=offsets)
    row_len = meta.row_len(COLS)
    for op in (P.muteTopOp(CURVES[curve], taperMs=taper),
               P.muteBottomOp(CURVES[curve], taperMs=taper)):
        flow = P.initPipeline([op])
        got = run(flow, x, meta, dtype)
        want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), row_len, DT,
                                     keys=meta.keys())
        check(got, want, x, row_len, dtype, f"{curve} {op.kind} taper {taper}")
        if curve != "time":
            assert np.array_equal(got[11], x[11]), "a NaN offset leaves the row alone"
        assert np.array_equal(got[13], x[13]), "a NaN t0 leaves the row alone"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mutes_in_a_flow_and_a_window(gpu, oracle, dtype):
    """demean, band-pass, top LMO mute, AGC, bottom table mute on a window
    that starts at trace sample 120 (mute columns follow firstSample)."""
    x, live, offsets, t0, status = survey(dtype, seed=8)
    meta = P.ChunkMeta(ROWS, DT, numSamples=live + 120, status=status, t0Ms=t0,
                       offset=offsets, firstSample=120)
    row_len = meta.row_len(COLS)
    flow = P.initPipeline([
        P.demeanOp(), P.bandPassOp(P.Corners(4, 8, 60, 80)),
        P.muteTopOp(CURVES["lmo"], taperMs=20), P.agcOp(300),
        P.muteBottomOp(CURVES["table"], taperMs=50)])
    got = run(flow, x, meta, dtype)
    want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), row_len, DT,
                                 keys=meta.keys(), first_sample=120)
    check(got, want, x, row_len, dtype, "flow")
    again = run(flow, x, meta, dtype, inplace=True)
    assert np.array_equal(got, again), "in place differs from src -> dst"


def test_mute_without_keys(gpu, oracle):
    """apply(row_len=...) with no keys: t0Ms 0 and offset NaN, so a time
    mute applies and an offset-dependent one leaves the rows alone."""
    import torch
    x, live, _, _, _ = survey(np.float32)
    rl = live.astype(np.int32)
    src = torch.from_numpy(x).cuda()
    for curve, moved in (("time", True), ("lmo", False)):
        flow = P.initPipeline([P.muteTopOp(CURVES[curve], taperMs=10)])
        plan = flow.compile(np.float32, DT, COLS)
        dst = torch.empty_like(src)
        plan.apply(src, dst, row_len=torch.from_numpy(rl).cuda())
        got = dst.cpu().numpy()
        want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), rl, DT)
        check(got, want, x, rl, np.float32, f"no keys {curve}")
        assert (not np.array_equal(got, x)) == moved


def test_mute_refusals(gpu):
    with pytest.raises(ValueError, match="taper"):
        P.initPipeline([P.muteTopOp(P.timeMute(100), taperMs=-1)]).compile(np.float32, DT, 100)
    bad = P.Op(P.pipeline.opMuteBottom, mute=P.MuteCurve(P.mkTable, offsets=(2.0, 1.0),
                                                          timesMs=(1.0, 2.0)))
    with pytest.raises(ValueError, match="ascending"):
        P.initPipeline([bad]).compile(np.float32, DT, 100)
