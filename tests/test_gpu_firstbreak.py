"""First-break picks on the GPU (qc/firstbreak.nim) against the oracle: the
pick columns and times equal, bit for bit, for every method, with and
without the minimum-phase band-pass, guided, fixed and whole-trace search
windows, float32 and float64; the SNR within 1e-12 dB (the device's log10
against the host's) and the confidence within 1e-13; the no-pick cases
(dead rows, NaN offsets under a guide, zero rows, NaN and infinite t0, a
search window past the trace, a single live sample), windows of a block
(firstSample, short live lengths), a 2000 x 3001 panel, the meta path and
the refusals."""
import math

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd.firstbreak import FirstBreakParams, pmCoppens, pmEnergyRatio, pmStaLta
from tests.fbshots import COLS, DT, T0, VEL, add_onset, keys_of, shot, split_spread

pytestmark = pytest.mark.gpu

GUIDED = dict(guideVelocity=VEL, guideT0Ms=T0, beforeMs=60, afterMs=80)
METHODS = (pmStaLta, pmEnergyRatio, pmCoppens)
BANDS = ((0.0, math.inf), (0.0, 60.0), (8.0, 70.0))


def both(oracle, x, rl, keys, params, dtype, first=0, samples=None):
    import torch
    x = np.ascontiguousarray(x.astype(dtype))
    samples = samples or x.shape[1]
    op = oracle.fb_plan(params, DT, samples)
    want = oracle.fb_pick(op, x, rl, keys, first)
    gp = P.initFirstBreakPlan(dtype, params, DT, samples)
    assert gp.requiredMargin == oracle.fb_required_margin(op)
    src = torch.from_numpy(x).cuda()
    k = None if keys is None else torch.from_numpy(np.ascontiguousarray(keys, np.float64)).cuda()
    got = gp.pickInto(src, row_len=torch.from_numpy(np.ascontiguousarray(rl, np.int32)).cuda(),
                      keys=k, first_sample=first)
    torch.cuda.synchronize()
    return want, (got.timeMs.cpu().numpy(), got.col.cpu().numpy(), got.snrDb.cpu().numpy(),
                  got.confidence.cpu().numpy())


def check(want, got, what=""):
    assert np.array_equal(got[1], want[1]), (what, np.nonzero(got[1] != want[1]))
    assert np.array_equal(got[0], want[0], equal_nan=True), what
    assert np.array_equal(np.isnan(got[2]), np.isnan(want[2])), what
    m = np.isfinite(want[2])
    assert np.array_equal(got[2][~m], want[2][~m], equal_nan=True), what
    assert np.all(np.abs(got[2][m] - want[2][m]) <= 1e-12 * np.maximum(1, np.abs(want[2][m]))), what
    assert np.all(np.abs(got[3] - want[3]) <= 1e-13), what


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("band", BANDS)
@pytest.mark.parametrize("method", METHODS)
def test_methods_and_bands(gpu, oracle, dtype, band, method):
    for sigma, seed in ((0.02, 1), (0.3, 2)):
        x, off = shot(sigma, seed)
        rl = np.full(x.shape[0], COLS, np.int32)
        for guided in (GUIDED, {}, dict(searchStartMs=60, searchEndMs=500)):
            p = FirstBreakParams(pickMethod=method, lowCutHz=band[0], highCutHz=band[1],
                                 **guided)
            want, got = both(oracle, x, rl, keys_of(off), p, dtype)
            assert np.sum(want[1] >= 0) >= x.shape[0] // 2
            check(want, got, (sigma, guided))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("slope", [6.0, 48.0, 54.0, 96.0])
def test_band_pass_slopes(gpu, oracle, dtype, slope):
    """Band-passes of 1 to 16 biquads per edge: up to eight biquads the
    cascade runs in one pass over the row, beyond in one pass per biquad
    through the padded copy; both give the oracle's picks exactly."""
    x, off = shot(0.1, 7)
    rl = np.full(x.shape[0], COLS, np.int32)
    for method in METHODS:
        p = FirstBreakParams(pickMethod=method, lowCutHz=8.0, highCutHz=70.0,
                             slopeDbPerOctave=slope, **GUIDED)
        want, got = both(oracle, x, rl, keys_of(off), p, dtype)
        check(want, got, (slope, method))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_no_picks_and_block_windows(gpu, oracle, dtype):
    x, off = shot(0.05, 4)
    off = off.copy()
    off[1] = np.nan
    x[2] = 0
    x[5] = 0
    x[5, 300] = 1.0                   # a single live sample
    keys = keys_of(off, t0=4.0)
    keys[3, 0] = np.nan
    keys[4, 0] = np.inf
    rl = np.full(x.shape[0], COLS, np.int32)
    rl[0] = -1
    rl[6] = 0
    rl[7] = 1
    rl[8] = 250                       # a live end before the window
    for m in METHODS:
        for band in BANDS[::2]:
            for kw in (GUIDED, {}, dict(searchStartMs=5000), dict(searchStartMs=196,
                                                                  searchEndMs=204)):
                p = FirstBreakParams(pickMethod=m, lowCutHz=band[0], highCutHz=band[1], **kw)
                want, got = both(oracle, x, rl, keys, p, dtype)
                check(want, got, (m, band, kw))
                w = np.ascontiguousarray(x[:, 10:410])
                want, got = both(oracle, w, np.minimum(rl, 400), keys, p, dtype, first=10,
                                 samples=COLS)
                check(want, got, (m, band, kw, "window"))
    want, got = both(oracle, x, rl, None, FirstBreakParams(pickMethod=pmCoppens), dtype)
    check(want, got, "no keys")


def test_zeros_before_the_data(gpu, oracle):
    rows, edge = 40, 100
    off = np.arange(rows) * 25.0 + 25
    x = np.zeros((rows, COLS))
    for r in range(rows):
        row = np.zeros(COLS)
        add_onset(row, DT, 400)
        row += np.random.default_rng(900 + r).normal(0, 0.02, COLS)
        x[r, edge:] = row[edge:]
    rl = np.full(rows, COLS, np.int32)
    for m in METHODS:
        for hc in (math.inf, 60.0):
            want, got = both(oracle, x, rl, keys_of(off),
                             FirstBreakParams(pickMethod=m, highCutHz=hc), np.float32)
            check(want, got, (m, hc))


@pytest.mark.parametrize("method,band", [(pmEnergyRatio, (0.0, 60.0)), (pmCoppens, (8.0, 70.0)),
                                         (pmStaLta, (0.0, math.inf))])
def test_full_panel(gpu, oracle, method, band):
    """A 2000 x 3001 panel at 2 ms: live lengths 2400 .. 3001, a split
    spread out to 5 km (the last onsets at 2.5 s)."""
    import torch
    rows, cols = 2000, 3001
    off = split_spread(rows, 5.0, 5.0)
    x, _ = shot(0.1, 30, off, cols=cols)
    live = 2400 + (np.arange(rows) * 7919) % 602
    for r in range(rows):
        x[r, live[r]:] = 0
    p = FirstBreakParams(pickMethod=method, lowCutHz=band[0], highCutHz=band[1], **GUIDED)
    want, got = both(oracle, x, live.astype(np.int32), keys_of(off), p, np.float32)
    check(want, got, "panel")
    assert np.sum(want[1] >= 0) > rows * 9 // 10
    # the meta path reads the same keys and live lengths
    meta = P.ChunkMeta(rows, DT, numSamples=live, offset=off)
    plan = P.initFirstBreakPlan(np.float32, p, DT, cols)
    picks = plan.pickInto(torch.from_numpy(x).cuda(), meta)
    assert np.array_equal(picks.col.cpu().numpy(), want[1])


def test_refusals(gpu):
    import torch
    plan = P.initFirstBreakPlan(np.float32, FirstBreakParams(**GUIDED), DT, COLS)
    x = torch.zeros((4, COLS + 1), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError, match="columns"):
        plan.pickInto(x, row_len=torch.zeros(4, dtype=torch.int32, device="cuda"))
    x = torch.zeros((4, COLS), dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        plan.pickInto(x, P.ChunkMeta(4, 4.0))
    with pytest.raises(ValueError):
        plan.pickInto(x, P.ChunkMeta(4, DT), ws=torch.zeros(8, dtype=torch.uint8, device="cuda"))
    with pytest.raises(ValueError):
        plan.pickInto(x.double(), P.ChunkMeta(4, DT))
