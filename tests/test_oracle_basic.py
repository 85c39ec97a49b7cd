"""Pins the oracle's basic kernels, per-trace normalization, the meta
flattening and the ProcessPool clone (core/basic.nim, gain/gain.nim:198-219,
core/types.nim:212-230, core/pool.nim; tbasic.nim, tpool.nim:68-121)."""
import math

import numpy as np
import pytest

import processseg_amd as P


def test_demean_detrend_closed_forms(oracle):
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 8, 9, 2001):
        x = rng.standard_normal(n) + 3 + 0.01 * np.arange(n)
        y = oracle.pipeline_apply([P.demeanOp()], x[None, :], np.array([n], np.int32), 2.0)[0]
        assert np.max(np.abs(y - (x - np.mean(x)))) <= 1e-13
        d = oracle.pipeline_apply([P.detrendOp()], x[None, :], np.array([n], np.int32), 2.0)[0]
        if n == 1:
            assert d[0] == 0
        else:
            b, a = np.polyfit(np.arange(n), x, 1)
            assert np.max(np.abs(d - (x - (a + b * np.arange(n))))) <= 1e-11


def test_scale_clip_polarity(oracle):
    x = np.random.default_rng(5).standard_normal((3, 100))
    rl = np.array([100, 50, -1], np.int32)
    ops = [P.scaleOp(0.5), P.clipOp(-0.3, 0.4), P.polarityOp(), P.gainDbOp(20)]
    y = oracle.pipeline_apply(ops, x, rl, 2.0)
    want = -np.clip(x * 0.5, -0.3, 0.4) * 10.0
    assert np.max(np.abs(y[0] - want[0])) <= 1e-14
    assert np.max(np.abs(y[1, :50] - want[1, :50])) <= 1e-14
    assert np.array_equal(y[1, 50:], x[1, 50:]) and np.array_equal(y[2], x[2])


@pytest.mark.parametrize("stat", [P.normMax, P.normRms, P.normPercentile])
def test_normalize_reaches_target(oracle, stat):
    rng = np.random.default_rng(6)
    x = rng.standard_normal((5, 800)) * np.array([[1e-3], [1], [50], [1e3], [7]])
    x[4] = 0
    rl = np.full(5, 800, np.int32)
    y = oracle.pipeline_apply([P.normalizeOp(stat, target=2.5, percentile=90)], x, rl, 2.0)
    for r in range(4):
        v = np.abs(y[r])
        level = {P.normMax: v.max(), P.normRms: math.sqrt(np.mean(v ** 2)),
                 P.normPercentile: np.percentile(v, 90)}[stat]
        assert level == pytest.approx(2.5, rel=1e-12)
    assert np.all(y[4] == 0), "an all-zero trace is left alone"


def test_meta_flatten_semantics(oracle):
    """isLive: status live or unknown and not failed; liveSamples: numSamples
    - firstSample clamped to 0 .. cols (core/types.nim:221-230); correlation
    uses every row but dead and failed ones (correlate.nim:442-443)."""
    nums = np.array([100, 100, 100, 100, 50, 5, 300, 100], np.int64)
    status = np.array([0, 1, 2, 3, 0, 0, 0, 0], np.uint8)
    got = oracle.meta_flatten(80, nums, status, failed=[7], first_sample=10)
    assert list(got) == [80, 80, -1, -1, 40, 0, 80, -1]
    got = oracle.meta_flatten(80, nums, status, failed=[7], first_sample=10, used_for_corr=True)
    assert list(got) == [80, 80, -1, 80, 40, 0, 80, -1]


def test_pool_equals_single_thread(oracle):
    """tpool.nim:68-121: the pool gives the single-thread bits for any worker
    count, with uneven rows (first half dead)."""
    from processseg_amd.workloads import big_flow
    from processseg_amd import synth
    x, live = synth.gather(203, 1501, 2.0, 8)
    rl = live.astype(np.int32)
    rl[:100] = -1
    flow = big_flow()
    want = oracle.pipeline_apply(flow.ops, x, rl, 2.0)
    for workers in (1, 2, 3, 5, 8):
        pool = oracle.Pool(workers)
        dst = np.empty_like(x)
        pool.apply_f32(flow.ops, x, dst, rl, 2.0)
        pool.close()
        assert np.array_equal(dst, want), workers
