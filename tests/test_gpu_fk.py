"""The f-k amplitude on the GPU (pseg_fk_amplitude, FkPlan.fkAmplitudeInto)
against the oracle's restatement (itself held to tests/tfk.nim): the dip's
sign, random gathers with dead, failed and short rows and a window offset,
extreme float64 scales, and the NaN / Inf class of every output."""
import math

import numpy as np
import pytest
import torch

import oracle as O
import processseg_amd as P
from test_oracle_fk import CLASSES, class_panel, classes

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def gpu_fk(x, meta, first_col, samples, taper=0.1, dt=2.0, T=np.float64):
    traces = P.fkRows(meta, x.shape[1], first_col, samples)
    plan = P.initFkPlan(T, dt, traces, samples, taper)
    amp = torch.zeros(plan.nK * plan.nFreq, dtype=torch.float64, device=DEV)
    src = torch.from_numpy(np.ascontiguousarray(x, T)).to(DEV)
    used = plan.fkAmplitudeInto(src, meta, first_col, amp)
    torch.cuda.synchronize()
    return amp.cpu().numpy(), used, plan


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_dip_sign(T):
    """tfk.nim:23-65 on the GPU: the peak two bins to the positive side of
    zero at time bin 8; nK 16, nFreq 33."""
    rows, cols, dt = 8, 64, 2.0
    f = 1000 / (64 * dt) * 8
    r = np.arange(rows)[:, None]
    j = np.arange(cols)[None, :]
    x = np.cos(2 * np.pi * f * (j - r) * dt / 1000)
    meta = P.ChunkMeta(rows, dt, numSamples=[cols] * rows)
    amp, used, plan = gpu_fk(x, meta, 0, cols, taper=0.0, T=T)
    assert used == rows and plan.nK == 16 and plan.nFreq == 33
    col = amp.reshape(16, 33)[:, 8]
    assert int(np.argmax(col)) == 10 and col[6] < 0.05 * col.max()
    want, _ = O.fk_amplitude(O.fk_plan(rows, cols, dt, 0.0, T), x.astype(T), np.full(rows, cols), 0)
    tol = 1e-5 if T == np.float32 else 1e-12
    assert np.max(np.abs(amp - want)) <= tol * want.max()


@pytest.mark.parametrize("T,tol", [(np.float32, 2e-6), (np.float64, 1e-12)])
@pytest.mark.parametrize("rows,cols,first,samples", [
    (24, 300, 0, 300), (61, 1000, 120, 500), (300, 2001, 0, 2001), (1000, 500, 37, 400)])
def test_random_gathers(T, tol, rows, cols, first, samples):
    """Gathers with dead, auxiliary and failed rows, rows that end before
    the window and rows that end inside it: equal to the oracle within tol
    of the peak (float32: the transforms' rounding)."""
    rng = np.random.default_rng(rows + cols)
    x = rng.normal(0, 1, (rows, cols))
    status = rng.choice([0, 0, 0, 0, 1, 2, 3], rows).astype(np.uint8)
    num = rng.integers(first // 2, cols + 1, rows)
    failed = sorted(rng.choice(rows, max(1, rows // 20), replace=False).tolist())
    meta = P.ChunkMeta(rows, 2.0, numSamples=num, status=status, failed=failed)
    amp, used, plan = gpu_fk(x, meta, first, samples, T=T)
    p = O.fk_plan(used, samples, 2.0, 0.1, T)
    want, used2 = O.fk_amplitude(p, x.astype(T), meta.row_len(cols), first)
    assert used == used2
    assert np.max(np.abs(amp - want)) <= tol * want.max()


@pytest.mark.parametrize("scale", [1e155, 1e-160, 1e-200])
def test_extreme_float64(scale):
    """tfk.nim:186-199: squares that overflow, lose digits or flush: the
    hypot modulus, no infinities, within 1e-12 of the oracle's peak."""
    rng = np.random.default_rng(1)
    x = scale * rng.uniform(-1, 1, (4, 8))
    meta = P.ChunkMeta(4, 2.0, numSamples=[8] * 4)
    amp, _, _ = gpu_fk(x, meta, 0, 8)
    want, _ = O.fk_amplitude(O.fk_plan(4, 8, 2.0, 0.1), x, np.full(4, 8), 0)
    assert not np.any(np.isinf(amp)) and np.any(amp != 0)
    assert np.max(np.abs(amp - want)) <= 1e-12 * want.max()


@pytest.mark.parametrize("kind", range(6))
def test_nan_inf_classes(kind):
    """tfk.nim:215-245: the class of every output of a 4 x 8 float64 gather
    holding NaN and infinities is the reference's."""
    meta = P.ChunkMeta(4, 2.0, numSamples=[8] * 4)
    amp, _, _ = gpu_fk(class_panel(kind), meta, 0, 8)
    assert classes(amp) == CLASSES[kind]


def test_refusals():
    """initFkPlan and fkAmplitudeInto refuse before any device work."""
    with pytest.raises(ValueError):
        P.initFkPlan(np.float32, 2.0, 3, 64)
    with pytest.raises(ValueError):
        P.initFkPlan(np.float32, 2.0, 8, 7)
    plan = P.initFkPlan(np.float32, 2.0, 8, 64)
    meta = P.ChunkMeta(8, 2.0, numSamples=[64] * 8)
    src = torch.zeros((8, 64), dtype=torch.float32, device=DEV)
    amp = torch.zeros(plan.nK * plan.nFreq, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError):
        plan.fkAmplitudeInto(src, P.ChunkMeta(8, 4.0, numSamples=[64] * 8), 0, amp)
    with pytest.raises(ValueError):
        plan.fkAmplitudeInto(src, meta, -1, amp)
    with pytest.raises(ValueError):
        plan.fkAmplitudeInto(src, P.ChunkMeta(8, 2.0, numSamples=[64] * 8, status=[2] + [0] * 7),
                             0, amp)
    with pytest.raises(ValueError):
        plan.fkAmplitudeInto(src, meta, 0, amp[:-1])


def test_full_panel():
    """The bench's workload: the f-k amplitude of the 2000 x 3001 f32 panel
    (live lengths 2400 .. 3001, 4000 wavenumbers x 1537 frequencies), within
    2e-6 of the oracle's peak."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(2000, 2.0, numSamples=live)
    amp, used, plan = gpu_fk(x, meta, 0, 3001, T=np.float32)
    assert used == 2000 and (plan.nK, plan.nFreq) == (4000, 1537)
    want, used2 = O.fk_amplitude(O.fk_plan(2000, 3001, 2.0, 0.1, np.float32), x,
                                 meta.row_len(3001), 0)
    assert used2 == 2000
    assert np.max(np.abs(amp - want)) <= 2e-6 * want.max()
