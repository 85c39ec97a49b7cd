"""Calibration on the GPU: calibrateOp / uncalibrateOp as a step of the row
kernel (pseg_plan_apply_calib), against the oracle (its resolution, its
per-row scale and its pipeline) and the values of the reference's
tests/tcalibop.nim and tests/tcalibrate.nim."""
import math

import numpy as np
import pytest
import torch

import oracle as O
import processseg_amd as P
from processseg_amd import calibrate as K

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
NaN = math.nan
D1 = K.descale(1, K.duMicroVPerCount)
GEO = K.sensitivity(20, K.suVPerMPerS)
HYD = K.sensitivity(20, K.suVPerBar)
ROWS, COLS, DT = 24, 300, 2.0


def op_meta(cols=COLS):
    """tcalibop.nim:18-36: a hydrophone every fourth row, the others
    geophones; a file descale 1e-6 (1 + r mod 5) per row; row 5
    auxiliary, row 9 dead, row 13 an accelerometer without a rule."""
    st = np.zeros(ROWS, np.uint8)
    st[5] = 3
    st[9] = 2
    sensor = np.array([1 if r % 4 == 0 else 2 for r in range(ROWS)], np.uint8)
    sensor[13] = 7
    scales = [K.FileScale(1e-6 * (1 + r % 5), NaN, K.auRecorded) for r in range(ROWS)]
    return P.ChunkMeta(ROWS, DT, numSamples=[cols] * ROWS, status=st, sensorType=sensor,
                       scales=scales)


CAL = K.initCalibration([
    K.CalibrationRule(K.RuleMatch(sensorType=1), sensitivity=K.sensitivity(20, K.suVPerBar)),
    K.CalibrationRule(K.RuleMatch(sensorType=2), sensitivity=K.sensitivity(22.5, K.suVPerMPerS))])


def raw_panel(dtype=np.float32, cols=COLS):
    r = np.arange(ROWS)[:, None]
    j = np.arange(cols)[None, :]
    return (1000 * np.sin(0.05 * j * (r + 1)) + 300).astype(dtype)


def oracle_k(cal, meta, inverse=False):
    """The rows' multipliers from the oracle's resolution (calibrateImpl's k)."""
    from test_oracle_calibrate import oracle_rule
    p = O.calib_plan([oracle_rule(r) for r in cal.rules], cal.target, cal.fileFactors,
                     cal.scaleAuxiliary)
    ks = []
    failed = set(int(f) for f in meta.failed)
    for r in range(len(meta)):
        fs = meta.scales[r] if meta.scales is not None else K.FileScale()
        row = dict(status=int(meta.status[r]), failed=r in failed,
                   sensorType=int(meta.sensorType[r]), component=int(meta.component[r]),
                   channelSet=int(meta.channelSet[r]), channel=float(meta.channel[r]),
                   rcvLine=float(meta.rcvLine[r]), rcvPoint=float(meta.rcvPoint[r]),
                   rcvIndex=int(meta.rcvIndex[r]), fsDescale=fs.descale,
                   fsSensitivity=fs.sensitivity, fsUnit=fs.unit, fsPhysical=fs.physical)
        res = O.calib_resolve_row(p, row, meta.hasFileScales)
        ks.append(1 / res["factor"] if inverse and res["status"] == 0 else res["factor"])
    return np.array(ks)


def run(flow, x, meta, T=np.float32, in_place=False):
    plan = flow.compile(T, DT, x.shape[1])
    src = torch.from_numpy(x).to(DEV)
    dst = src if in_place else torch.empty_like(src)
    plan.apply(src, dst, meta)
    torch.cuda.synchronize()
    return dst.cpu().numpy(), plan


def same_bits(a, b):
    both_nan = np.isnan(a) & np.isnan(b)
    assert np.array_equal(np.where(both_nan, 0, a), np.where(both_nan, 0, b))


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("in_place", [False, True])
def test_calibrate_step_equals_the_oracle_bit_for_bit(T, in_place):
    """A calibrate-only flow: each row times T(k) over its live length, dead
    and auxiliary rows included (pipeline.nim:44-45), bit for bit the
    oracle's scale with the oracle's factors."""
    meta = op_meta()
    x = raw_panel(T)
    got, plan = run(P.initPipeline([P.calibrateOp(CAL)]), x, meta, T, in_place)
    assert plan.amplitudeUnits == 1 and plan.nCalib == 1
    k = oracle_k(CAL, meta)
    assert np.isnan(k[13]) and np.sum(np.isnan(k)) == 1
    want = O.calibrate_rows(x, k, meta.row_len(COLS, 2))
    same_bits(got, want)
    # the auxiliary row in volts from the file's descale (row 5: 1e-6 V/count)
    assert abs(got[5, 7] - x[5, 7] * 1e-6) <= 1e-6 * abs(x[5, 7] * 1e-6)
    # the dead row is converted too (a geophone: 1e-6·(1 + 9 mod 5) / 22.5)
    assert abs(got[9, 3] - x[9, 3] * 5e-6 / 22.5) <= 1e-6 * abs(got[9, 3])


def test_op_equals_kernel_then_ops():
    """tcalibop.nim:43-72: calibrate > demean equals the calibration then
    demean on the live rows (oracle), the accelerometer NaN, statistics
    physical (m/s around 1e-4)."""
    meta = op_meta()
    x = raw_panel()
    got, plan = run(P.initPipeline([P.calibrateOp(CAL), P.demeanOp()]), x, meta)
    k = oracle_k(CAL, meta)
    cal = O.calibrate_rows(x, k, meta.row_len(COLS, 2))
    want = O.pipeline_apply([P.demeanOp()], cal.astype(np.float64), meta.row_len(COLS), DT)
    live = meta.row_len(COLS) >= 0
    live[13] = False
    assert O.rel_l2(got[live].astype(np.float64), want[live]) <= 1e-6
    assert np.all(np.isnan(got[13]))
    same_bits(got[~(meta.row_len(COLS) >= 0)], cal[~(meta.row_len(COLS) >= 0)])
    rms1 = float(np.sqrt(np.mean(got[1].astype(np.float64) ** 2)))
    assert 1e-5 < rms1 < 1e-3


def test_panel_normalization_sees_physical_values():
    """tcalibop.nim:74-104: equal counts in every row, factors per row: the
    per-panel max normalization is set by the row with the largest factor."""
    meta = op_meta()
    x = np.tile((1000 * np.sin(0.1 * np.arange(COLS))).astype(np.float32), (ROWS, 1))
    got, plan = run(P.initPipeline([P.calibrateOp(CAL), P.normalizeOp(P.normMax, P.perPanel)]),
                    x, meta)
    assert plan.amplitudeUnits == 2
    rows = K.initCalibrationPlan(CAL).resolve(meta)
    live = meta.row_len(COLS) >= 0
    best, best_row = 0.0, -1
    for r in range(ROWS):
        if live[r] and rows[r].status == K.csConverted and abs(rows[r].factor) > best:
            best, best_row = abs(rows[r].factor), r
    assert abs(np.max(np.abs(got[best_row].astype(np.float64))) - 1) <= 1e-5
    other = np.max(np.abs(got[2].astype(np.float64)))
    assert abs(other - rows[2].factor / best) <= 1e-5 * other


def test_agc_after_calibration_float64():
    """tcalibop.nim:106-125: calibrate > AGC 100 ms equals the calibration
    then the AGC (oracle) on the live rows, float64."""
    meta = op_meta()
    x = raw_panel(np.float64)
    got, _ = run(P.initPipeline([P.calibrateOp(CAL), P.agcOp(100)]), x, meta, np.float64)
    cal = O.calibrate_rows(x, oracle_k(CAL, meta), meta.row_len(COLS, 2))
    want = O.pipeline_apply([P.agcOp(100)], cal, meta.row_len(COLS), DT)
    ok = ~np.isnan(want)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert np.max(np.abs(got[ok] - want[ok])) <= 1e-12 * np.max(np.abs(want[ok]))


def test_uncalibrate_in_a_flow():
    """tcalibop.nim:190-206: calibrate > band-pass > uncalibrate equals the
    band-pass alone on rows that convert (float64, 1e-9 · 1000)."""
    meta = op_meta()
    x = raw_panel(np.float64)
    band = P.bandPassOp(P.Corners(4, 8, 60, 80))
    got, plan = run(P.initPipeline([P.calibrateOp(CAL), band, P.uncalibrateOp(CAL)]), x, meta,
                    np.float64)
    assert plan.amplitudeUnits == 0 and plan.nCalib == 2
    only, _ = run(P.initPipeline([band]), x, meta, np.float64)
    for r in (1, 2, 3, 4):
        assert np.max(np.abs(got[r] - only[r])) <= 1e-9 * 1000


def test_rows_by_status_values():
    """tcalibrate.nim:193-226 on the GPU: live (converted), live
    accelerometer (NaN), dead with 10 live samples (converted there, the
    rest untouched), dead accelerometer (kept), auxiliary (volts), failed
    (untouched)."""
    cols = 16
    sensor = np.array([2, 7, 2, 7, 0, 2], np.uint8)
    meta = P.ChunkMeta(6, DT, numSamples=[cols, cols, 10, cols, cols, cols],
                       status=[0, 0, 2, 2, 3, 0], failed=[5], sensorType=sensor)
    cal = K.initCalibration([K.CalibrationRule(K.RuleMatch(sensorType=2), sensitivity=GEO),
                             K.CalibrationRule(descale=D1)])
    x = np.full((6, cols), 1000.0)
    got, _ = run(P.initPipeline([P.calibrateOp(cal)]), x, meta, np.float64, in_place=True)
    k = 1e-6 / 20 * 1000
    for c in range(cols):
        assert abs(got[0, c] - k) <= 1e-12 * k
        assert math.isnan(got[1, c])
        assert (abs(got[2, c] - k) <= 1e-12 * k) if c < 10 else got[2, c] == 1000
        assert got[3, c] == 1000 and got[5, c] == 1000
        assert abs(got[4, c] - 1e-6 * 1000) <= 1e-12
    s = K.initCalibrationPlan(cal).resolveInto(meta)
    assert (s.converted, s.unresolved, s.kept, s.noSensitivity) == (3, 1, 2, 1)


@pytest.mark.parametrize("T,tol", [(np.float64, 1e-12), (np.float32, 1e-6)])
def test_inverse_round_trip(T, tol):
    """tcalibrate.nim:287-320: calibrate then uncalibrate gives the counts
    back (float64 to 1e-12, float32 to 1e-6); the accelerometer NaN."""
    cols = 16
    meta = P.ChunkMeta(3, DT, numSamples=[cols] * 3, sensorType=[2, 1, 7], component=[2, 1, 7],
                       channel=[3, 3, 3], rcvLine=[10] * 3, rcvPoint=[201] * 3, rcvIndex=[1] * 3)
    cal = K.initCalibration([
        K.CalibrationRule(K.RuleMatch(sensorType=1), sensitivity=HYD),
        K.CalibrationRule(K.RuleMatch(sensorType=2), sensitivity=GEO,
                          descale=K.descale(0.3, K.duMicroVPerCount)),
        K.CalibrationRule(descale=D1)])
    x = np.array([[c * 37 - 200 + 0.5 for c in range(cols)]] * 3, T)
    plan = K.initCalibrationPlan(cal)
    t = torch.from_numpy(x.copy()).to(DEV)
    plan.calibrate(t, meta)
    mid = t.cpu().numpy()
    assert np.all(np.isnan(mid[2]))
    assert abs(mid[0, 3] - x[0, 3] * 0.3e-6 / 20) <= 1e-6 * abs(mid[0, 3])
    plan.uncalibrate(t, meta)
    back = t.cpu().numpy().astype(np.float64)
    for r in (0, 1):
        assert np.all(np.abs(back[r] - x[r]) <= tol * np.abs(x[r].astype(np.float64)))
    dst = torch.empty_like(t)
    src = torch.from_numpy(x.copy()).to(DEV)
    plan.calibrateInto(src, dst, meta)
    assert np.array_equal(src.cpu().numpy(), x)


def test_scale_rows():
    """tcalibrate.nim:322-340: scaleRows multiplies every row that did not
    fail over its live length, whatever its status; NaN fills a row."""
    cols = 16
    meta = P.ChunkMeta(3, DT, numSamples=[cols, 4, cols], status=[0, 3, 0], failed=[2])
    x = torch.full((3, cols), 1000.0, dtype=torch.float64, device=DEV)
    K.scaleRows(x, meta, [0.5, 3.0, 7.0])
    got = x.cpu().numpy()
    assert np.all(got[0] == 500) and np.all(got[2] == 1000)
    assert np.all(got[1, :4] == 3000) and np.all(got[1, 4:] == 1000)
    src = torch.full((3, cols), 1000.0, dtype=torch.float64, device=DEV)
    dst = torch.zeros_like(src)
    K.scaleRows(src, meta, [NaN, 2.0, 7.0], dst=dst)
    d = dst.cpu().numpy()
    assert np.all(np.isnan(d[0])) and np.all(src.cpu().numpy() == 1000)
    assert np.all(d[1, :4] == 2000) and np.all(d[1, 4:] == 1000) and np.all(d[2] == 1000)
    with pytest.raises(ValueError):
        K.scaleRows(src, meta, [1.0, math.inf, 1.0])


def test_order_rule_and_units():
    """tcalibop.nim:127-173: a calibration after an op that changes
    amplitudes, or a second one, is refused with the op's index; linear
    scale-free ops may come first; amplitudeUnits."""
    for bad in ([P.gainDbOp(6), P.calibrateOp(CAL)], [P.scaleOp(2), P.calibrateOp(CAL)],
                [P.clipOp(-1, 1), P.calibrateOp(CAL)], [P.agcOp(100), P.calibrateOp(CAL)],
                [P.normalizeOp(P.normRms), P.calibrateOp(CAL)],
                [P.gainCurveOp(P.powerGain(2)), P.calibrateOp(CAL)],
                [P.calibrateOp(CAL), P.agcOp(100), P.uncalibrateOp(CAL)]):
        with pytest.raises(ValueError) as e:
            P.initPipeline(bad).compile(np.float32, DT, COLS)
        assert "must come before every op that changes amplitudes" in str(e.value)
        assert str(e.value).startswith(f"op {len(bad) - 1}")
    with pytest.raises(ValueError) as e:
        P.initPipeline([P.calibrateOp(CAL), P.demeanOp(), P.calibrateOp(CAL)]).compile(
            np.float32, DT, COLS)
    assert str(e.value).startswith("op 2 (opCalibrate): one per pipeline: op 0")
    dup = K.initCalibration([K.CalibrationRule(descale=K.descale(1, K.duVPerCount)),
                             K.CalibrationRule(sensitivity=K.sensitivity(1, K.suVPerPa))])
    with pytest.raises(ValueError) as e:
        P.initPipeline([P.calibrateOp(dup)]).compile(np.float32, DT, COLS)
    assert str(e.value).startswith("op 0 (opCalibrate): rule 1:")
    for good in ([P.demeanOp(), P.detrendOp(), P.polarityOp(), P.calibrateOp(CAL)],
                 [P.bandPassOp(P.Corners(4, 8, 60, 80)), P.calibrateOp(CAL), P.uncalibrateOp(CAL)],
                 [P.muteTopOp(P.timeMute(100)), P.calibrateOp(CAL)]):
        P.initPipeline(good).compile(np.float32, DT, COLS)

    def units(ops):
        return P.initPipeline(ops).compile(np.float32, DT, COLS).amplitudeUnits
    assert units([]) == 0 and units([P.demeanOp(), P.agcOp(100)]) == 0
    assert units([P.calibrateOp(CAL)]) == 1
    assert units([P.calibrateOp(CAL), P.detrendOp(), P.bandPassOp(P.Corners(4, 8, 60, 80))]) == 1
    assert units([P.calibrateOp(CAL), P.gainDbOp(3)]) == 2
    assert units([P.calibrateOp(CAL), P.demeanOp(), P.uncalibrateOp(CAL)]) == 0
    assert units([P.uncalibrateOp(CAL), P.calibrateOp(CAL)]) == 1


def test_needs_the_factors():
    """A calibrating plan refuses the apply entry points that carry no
    factors, and batches, before any device work."""
    plan = P.initPipeline([P.calibrateOp(CAL)]).compile(np.float32, DT, COLS)
    src = torch.zeros((ROWS, COLS), dtype=torch.float32, device=DEV)
    rl = torch.full((ROWS,), COLS, dtype=torch.int32, device=DEV)
    with pytest.raises(ValueError, match="calibration step"):
        plan.apply(src, src, row_len=rl)
    with pytest.raises(ValueError, match="calibration step"):
        plan.applyHost(np.zeros((ROWS, COLS), np.float32), meta=op_meta())


def test_wide_rows_and_mixed_flow():
    """Rows of 5000 samples (256 threads per row) in float32: calibrate >
    band-pass > AGC against the oracle's calibration then pipeline."""
    cols = 5000
    rng = np.random.default_rng(3)
    n = 40
    st = rng.integers(0, 4, n).astype(np.uint8)
    sensor = rng.integers(0, 3, n).astype(np.uint8)
    num = rng.integers(3000, cols + 1, n)
    meta = P.ChunkMeta(n, DT, numSamples=num, status=st, sensorType=sensor, failed=[7, 8])
    cal = K.initCalibration([K.CalibrationRule(K.RuleMatch(sensorType=1), sensitivity=HYD),
                             K.CalibrationRule(K.RuleMatch(sensorType=2), sensitivity=GEO),
                             K.CalibrationRule(descale=D1)])
    x = rng.normal(0, 300, (n, cols)).astype(np.float32)
    flow = P.initPipeline([P.calibrateOp(cal), P.bandPassOp(P.Corners(5, 10, 70, 90)),
                           P.agcOp(250)])
    got, plan = run(flow, x, meta)
    assert plan.threadsPerRow == 256
    cal_x = O.calibrate_rows(x, oracle_k(cal, meta), meta.row_len(cols, 2))
    want = O.pipeline_apply(flow.ops[1:], cal_x.astype(np.float64), meta.row_len(cols), DT)
    live = (meta.row_len(cols) >= 0) & ~np.any(np.isnan(want), axis=1)
    assert O.rel_l2(got[live].astype(np.float64), want[live]) <= 1e-5
    same_bits(got[~(meta.row_len(cols) >= 0)], cal_x[~(meta.row_len(cols) >= 0)])


def test_full_panel_calibrated_flow():
    """The bench's workload: calibrateOp (a hydrophone rule and a geophone
    rule over file descales) then demean > band-pass + notch > high cut >
    AGC on the 2000 x 3001 f32 panel; every row computed, 64 rows held to the
    float64 oracle (its factors, its scale, its pipeline) within 1e-5
    relative L2."""
    from processseg_amd.workloads import big_flow, big_panel
    x, live = big_panel()
    rows, cols = x.shape
    sensor = np.where(np.arange(rows) % 4 == 0, 1, 2).astype(np.uint8)
    scales = [K.FileScale(1e-6 * (1 + r % 5), NaN, K.auRecorded) for r in range(rows)]
    meta = P.ChunkMeta(rows, DT, numSamples=live, sensorType=sensor, scales=scales)
    flow = P.initPipeline([P.calibrateOp(CAL)] + big_flow().ops)
    got, plan = run(flow, x, meta)
    assert plan.nCalib == 1 and np.all(np.isfinite(got))
    pick = np.arange(0, rows, rows // 64)[:64]
    sub = P.ChunkMeta(len(pick), DT, numSamples=live[pick], sensorType=sensor[pick],
                      scales=[scales[r] for r in pick])
    k = oracle_k(CAL, sub)
    xs = x[pick].astype(np.float64)
    rl = sub.row_len(cols)
    want = O.pipeline_apply(big_flow().ops, O.calibrate_rows(xs, k, sub.row_len(cols, 2)), rl, DT)
    e = float(np.linalg.norm(got[pick] - want) / np.linalg.norm(want))
    assert e <= 1e-5, e
