"""Vibroseis correlation on the GPU beyond the default plan, at the full
record size (1200 records of 10001 samples, 12 s sweep): sweep
deconvolution, pilots that start before and after the time break (whole and
sub-sample), unnormalized output and an output as wide as the record (the
zero tail past the last lag), each against the float64 oracle on every row
(vibroseis/correlate.nim:82-160, 241-290, 421-464)."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import workloads as W

from common import max_abs, rel_l2

pytestmark = pytest.mark.gpu

ROWS, REC = 1200, 10001

CASES = [
    # (name, t0Ms, mode, epsPercent, normalize, outSamples)
    ("deconvolution", 0.0, 1, 1.0, True, 0),
    ("deconvolution_sub_sample_start", 0.6, 1, 0.1, True, 0),
    ("sub_sample_start", 0.6, 0, 1.0, True, 0),
    ("start_before_break", -20.0, 0, 1.0, True, 0),
    ("start_far_before_break_wide_out", -900.0, 0, 1.0, True, REC),
    ("start_after_break_wide_out", 3000.0, 0, 1.0, True, REC),
    ("unnormalized", 0.0, 0, 1.0, False, 0),
    ("wide_out", 0.0, 0, 1.0, True, REC),
]


@pytest.fixture(scope="module")
def records():
    return W.vibroseis_panel(ROWS, REC)


def _meta(rows):
    status = np.zeros(rows, np.uint8)
    status[7] = P.tsDead
    status[11] = P.tsAuxiliary
    nums = np.full(rows, REC, np.int64)
    nums[13] = 7000
    nums[17] = 0
    return P.ChunkMeta(rows, 2.0, numSamples=nums, status=status, failed=[19])


def _run(T, x, pilot, t0, mode, eps, normalize, out):
    import torch
    pil = P.Pilot(pilot.samples, pilot.dtMs, t0)
    plan = P.initCorrelationPlan(T, pil, 2.0, REC, outSamples=out, mode=mode, epsPercent=eps,
                                 normalize=normalize)
    rows = x.shape[0]
    meta = _meta(rows)
    dst = torch.full((rows, plan.outSamples), 7.0, dtype=torch.float32 if T == np.float32
                     else torch.float64, device="cuda")
    plan.correlateInto(torch.from_numpy(np.ascontiguousarray(x.astype(T))).cuda(), dst, meta)
    torch.cuda.synchronize()
    return plan, dst.cpu().numpy(), meta.row_len(REC, for_correlation=True)


@pytest.mark.parametrize("name,t0,mode,eps,normalize,out", CASES, ids=[c[0] for c in CASES])
def test_float32_every_row(gpu, oracle, records, name, t0, mode, eps, normalize, out):
    pilot = W.sweep_pilot()
    print(f"correlation {name}: f32 1200x10001 t0 {t0} mode {mode} eps {eps} "
          f"normalize {normalize} out {out}", flush=True)
    plan, got, rl = _run(np.float32, records, pilot, t0, mode, eps, normalize, out)
    info = oracle.corr_info(pilot.samples, 2.0, REC, out_samples=out, mode=mode, eps_percent=eps,
                            normalize=normalize, pilot_t0_ms=t0)
    assert (plan.nfft, plan.shift, plan.outSamples) == (info["nfft"], info["shift"],
                                                        info["out_samples"])
    want = oracle.correlate(records.astype(np.float64), rl, pilot.samples, 2.0, REC,
                            out_samples=out, mode=mode, eps_percent=eps, normalize=normalize,
                            pilot_t0_ms=t0)
    peak = np.max(np.abs(want))
    e = max_abs(got, want) / peak
    r = rel_l2(got, want)
    print(f"  nfft {plan.nfft} shift {plan.shift}: max abs / peak {e:.2e}, rel L2 {r:.2e}")
    assert r <= 1e-5 and e <= 2e-5
    for row in (7, 19):
        assert np.all(got[row] == 0)
    # columns no lag reaches are exact zeros (lagSpans, correlate.nim:265-274)
    zero_cols = np.all(want == 0, axis=0)
    assert np.all(got[:, zero_cols] == 0)
    if out == REC and plan.shift >= 0:
        assert zero_cols[REC - plan.shift:].all()


F64_CASES = [c for c in CASES if c[1] in (0.0, 0.6, -20.0)]


@pytest.mark.parametrize("name,t0,mode,eps,normalize,out", F64_CASES,
                         ids=[c[0] for c in F64_CASES])
def test_float64_rows(gpu, oracle, records, name, t0, mode, eps, normalize, out):
    pilot = W.sweep_pilot()
    rows = 128
    x = records[:rows].astype(np.float64)
    print(f"correlation {name}: f64 {rows}x10001", flush=True)
    plan, got, rl = _run(np.float64, x, pilot, t0, mode, eps, normalize, out)
    want = oracle.correlate(x, rl, pilot.samples, 2.0, REC, out_samples=out, mode=mode,
                            eps_percent=eps, normalize=normalize, pilot_t0_ms=t0)
    e = max_abs(got, want) / np.max(np.abs(want))
    r = rel_l2(got, want)
    print(f"  max abs / peak {e:.2e}, rel L2 {r:.2e}")
    assert r <= 1e-12 and e <= 1e-12
