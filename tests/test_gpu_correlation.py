"""Vibroseis correlation on the GPU at the full record size: 1200 records of
10001 samples at 2 ms against a 12 s 6-80 Hz sweep (nfft 16200, 4001 output
lags), compared with the float64 oracle on every row
(vibroseis/correlate.nim:82-160, 421-464)."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import workloads as W

from common import max_abs, rel_l2

pytestmark = pytest.mark.gpu

ROWS, REC = 1200, 10001


@pytest.fixture(scope="module")
def records():
    return W.vibroseis_panel(ROWS, REC)


def _meta(rows):
    status = np.zeros(rows, np.uint8)
    status[7] = P.tsDead
    status[11] = P.tsAuxiliary
    nums = np.full(rows, REC, np.int64)
    nums[13] = 7000            # a short record: zeros past its live end
    nums[17] = 0
    failed = [r for r in (19, 1100) if r < rows]
    return P.ChunkMeta(rows, 2.0, numSamples=nums, status=status, failed=failed)


def test_full_size_against_oracle(gpu, oracle, records):
    import torch
    pilot = W.sweep_pilot()
    plan = P.initCorrelationPlan(np.float32, pilot, 2.0, REC)
    info = oracle.corr_info(pilot.samples, 2.0, REC)
    assert (plan.nfft, plan.outSamples) == (info["nfft"], info["out_samples"]) == (16200, 4001)
    assert plan.threadsPerRecord == 512
    meta = _meta(ROWS)
    src = torch.from_numpy(records).cuda()
    dst = torch.full((ROWS, plan.outSamples), 7.0, dtype=torch.float32, device="cuda")
    plan.correlateInto(src, dst, meta)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    rl = meta.row_len(REC, for_correlation=True)
    want = oracle.correlate(records.astype(np.float64), rl, pilot.samples, 2.0, REC)
    peak = np.max(np.abs(want))
    e = max_abs(got, want) / peak
    r = rel_l2(got, want)
    print("1200x10001 correlation: max abs / peak", e, "rel L2", r)
    assert r <= 1e-5 and e <= 2e-5
    # dead and failed rows are zero; the auxiliary trace is correlated
    for row in (7, 19, 1100):
        assert np.all(got[row] == 0)
    assert np.any(got[11] != 0)


def test_batch_independence(gpu, records):
    """A record's output does not depend on the other rows of the launch."""
    import torch
    pilot = W.sweep_pilot()
    plan = P.initCorrelationPlan(np.float32, pilot, 2.0, REC)
    full = torch.from_numpy(records[:64]).cuda()
    dst = torch.empty(64, plan.outSamples, dtype=torch.float32, device="cuda")
    plan.correlateInto(full, dst, P.ChunkMeta(64, 2.0))
    one = torch.empty(1, plan.outSamples, dtype=torch.float32, device="cuda")
    plan.correlateInto(full[37:38].contiguous(), one, P.ChunkMeta(1, 2.0))
    torch.cuda.synchronize()
    assert np.array_equal(dst[37].cpu().numpy(), one[0].cpu().numpy())



def test_full_size_float64_against_oracle(gpu, oracle, records):
    """The same records in float64 (nfft 16200 in one workgroup's LDS):
    within float64 rounding of the oracle."""
    import torch
    pilot = W.sweep_pilot()
    plan = P.initCorrelationPlan(np.float64, pilot, 2.0, REC)
    assert (plan.nfft, plan.outSamples) == (16200, 4001)
    rows = 256
    x = records[:rows].astype(np.float64)
    meta = _meta(rows)
    dst = torch.full((rows, plan.outSamples), 7.0, dtype=torch.float64, device="cuda")
    plan.correlateInto(torch.from_numpy(x).cuda(), dst, meta)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    rl = meta.row_len(REC, for_correlation=True)
    want = oracle.correlate(x, rl, pilot.samples, 2.0, REC)
    e = max_abs(got, want) / np.max(np.abs(want))
    r = rel_l2(got, want)
    print("256x10001 float64 correlation: max abs / peak", e, "rel L2", r)
    assert r <= 1e-12 and e <= 1e-12
    for row in (7, 19):
        assert np.all(got[row] == 0)
