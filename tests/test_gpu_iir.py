"""Butterworth biquad cascades on the GPU (filter/iir.nim through pipeline.nim's
iirOp) against the oracle's serial float64 recursion: zero and minimum phase,
low cuts, high cuts and band-passes of several orders (odd included), the
three trend removals, live lengths from 0 to the record, rows not live, the
three threads-per-row classes, in place, inside a flow, the margin the plan
reports, and the refusals.  The GPU runs each section's pass as a parallel
scan; float32 within 1e-5 relative L2 of the float64 oracle, float64 within
1e-12; samples past a live end and rows not live unchanged, bit for bit."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import synth
from tests.common import rel_l2

pytestmark = pytest.mark.gpu

DT = 2.0

ZERO = {
    "bp4-60": P.butterBandPass(4, 60, 24),
    "lc3": P.butterLowCut(3, 24),
    "hc80x48": P.butterHighCut(80, 48),
    "bp8-100": P.butterBandPass(8, 100, 48, 24),
}
MINIMUM = {
    "bp4-60": P.butterBandPass(4, 60, 24),
    "lc3x18": P.butterLowCut(3, 18),
    "hc60x30": P.butterHighCut(60, 30),
}


def panel(rows, cols, dtype, seed):
    x, _ = synth.gather(rows, cols, DT, seed, dtype=dtype)
    x = x + dtype(0.5) + (np.arange(cols) * 2e-4).astype(dtype)   # a DC level and a drift
    live = synth.live_lengths(seed, rows, 0, cols).astype(np.int32)
    live[0] = cols
    live[1] = 1
    live[2] = 2
    if rows > 5:
        live[5] = -1
    return x, live


def run(plan, x, live, inplace=False):
    import torch
    src = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    rl = torch.from_numpy(live).cuda()
    if inplace:
        plan.apply(src, row_len=rl)
        got = src
    else:
        got = torch.empty_like(src)
        plan.apply(src, got, row_len=rl)
    torch.cuda.synchronize()
    return got.cpu().numpy()


def check(got, want, x, live, dtype, label):
    tol = 1e-5 if dtype == np.float32 else 1e-12
    for r in range(got.shape[0]):
        n = int(live[r])
        if n < 0:
            assert np.array_equal(got[r], x[r]), f"{label}: row {r} not live, not copied"
            continue
        assert np.array_equal(got[r, n:], x[r, n:]), f"{label}: row {r} past its live end"
        if n == 0:
            continue
        e = rel_l2(got[r, :n], want[r, :n])
        assert e <= tol, f"{label}: row {r} ({n} live): rel L2 {e:.3e}"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("phase,name", [(P.ipZero, k) for k in ZERO] +
                         [(P.ipMinimum, k) for k in MINIMUM])
@pytest.mark.parametrize("trend", [P.trNone, P.trMean, P.trLinear])
def test_iir(gpu, oracle, dtype, phase, name, trend):
    spec = (ZERO if phase == P.ipZero else MINIMUM)[name]
    x, live = panel(24, 1500, dtype, seed=11)
    flow = P.initPipeline([P.iirOp(spec, phase=phase, trend=trend)])
    plan = flow.compile(dtype, DT, 1500)
    got = run(plan, x, live)
    want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), live, DT)
    check(got, want, x, live, dtype, f"{name} phase {phase} trend {trend}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("cols", [300, 4500, 9000])
def test_iir_thread_classes(gpu, oracle, dtype, cols):
    """Rows for 128, 256 and 512 threads per row, zero and minimum phase."""
    x, live = panel(8, cols, dtype, seed=cols)
    for phase in (P.ipZero, P.ipMinimum):
        flow = P.initPipeline([P.iirOp(P.butterBandPass(5, 70, 24), phase=phase)])
        plan = flow.compile(dtype, DT, cols)
        got = run(plan, x, live)
        want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), live, DT)
        check(got, want, x, live, dtype, f"{cols} phase {phase}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_iir_in_a_flow(gpu, oracle, dtype):
    """demean > zero-phase IIR band-pass > AGC > minimum-phase IIR high cut,
    src -> dst and in place."""
    x, live = panel(32, 2001, dtype, seed=4)
    flow = P.initPipeline([P.demeanOp(), P.iirOp(P.butterBandPass(4, 60, 24)), P.agcOp(300),
                           P.iirOp(P.butterHighCut(90, 18), phase=P.ipMinimum)])
    plan = flow.compile(dtype, DT, 2001)
    got = run(plan, x, live)
    want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), live, DT)
    check(got, want, x, live, dtype, "flow")
    again = run(plan, x, live, inplace=True)
    assert np.array_equal(got, again), "in place differs from src -> dst"


def test_iir_margin_and_refusals(gpu, oracle):
    for spec, phase in ((P.butterBandPass(4, 60, 24), P.ipZero),
                        (P.butterLowCut(3, 18), P.ipMinimum)):
        flow = P.initPipeline([P.iirOp(spec, phase=phase)])
        want = oracle.IirPlan(spec, DT, 1000, phase).requiredMargin
        assert flow.requiredMargin(DT) == want
        assert flow.compile(np.float32, DT, 1000).requiredMargin == want
    with pytest.raises(ValueError, match="Butterworth"):
        P.initPipeline([P.iirOp(P.bandPass(P.Corners(4, 8, 60, 80)))]).compile(np.float32, DT, 500)
    with pytest.raises(ValueError, match="multiple of 12"):
        P.initPipeline([P.iirOp(P.butterLowCut(4, 18))]).compile(np.float32, DT, 500)
    with pytest.raises(ValueError, match="notches"):
        P.initPipeline([P.iirOp(P.butterLowCut(4).withNotch(50, 2))]).compile(np.float32, DT, 500)
