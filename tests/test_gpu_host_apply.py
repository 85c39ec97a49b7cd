"""pseg_plan_apply_host: apply(plan, src: Tensor, dst, meta, ws) on host
buffers, staged through device memory the library keeps.  Through the C
ABI directly (plain host pointers) and through PipelinePlan.applyHost: the
result equals the device-resident apply bit for bit (src -> dst and in
place, float32 and float64, a per-panel normalization that needs a
workspace, a mute that reads the rows' keys, dead and short rows), and
the float64 oracle within the float32 tolerance; a larger block after a
smaller regrows the staging; refusals before any copy."""
import ctypes as C

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import _lib, synth
from tests.common import rel_l2

pytestmark = pytest.mark.gpu


def panel(rows, cols, dtype=np.float32, seed=3):
    x, _ = synth.gather(rows, cols, 2.0, seed, dtype=dtype)
    live = synth.live_lengths(seed, rows, cols // 2, cols)
    st = np.zeros(rows, np.uint8)
    st[3] = P.tsDead
    meta = P.ChunkMeta(rows, 2.0, numSamples=live, status=st, failed=[5],
                       t0Ms=np.where(np.arange(rows) % 4 == 0, 12.0, 0.0),
                       offset=np.linspace(-2500, 2500, rows))
    return np.ascontiguousarray(x), meta


def device_apply(plan, x, meta):
    import torch
    src = torch.from_numpy(x).cuda()
    dst = torch.empty_like(src)
    plan.apply(src, dst, meta)
    return dst.cpu().numpy()


FLOWS = {
    "bandpass_agc": lambda: P.initPipeline([P.demeanOp(), P.bandPassOp(P.Corners(5, 10, 70, 90)),
                                            P.agcOp(250)]),
    "panel_norm": lambda: P.initPipeline([P.bandPassOp(P.Corners(4, 8, 60, 80)),
                                          P.normalizeOp(P.normRms, scope=P.perPanel)]),
    "mute": lambda: P.initPipeline([P.muteTopOp(P.lmoMute(2500.0), taperMs=30),
                                    P.agcOp(400)]),
}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("name", sorted(FLOWS))
def test_host_apply_equals_device_apply(gpu, oracle, dtype, name):
    flow = FLOWS[name]()
    x, meta = panel(200, 1501, dtype)
    plan = flow.compile(dtype, 2.0, 1501)
    want = device_apply(plan, x, meta)
    dst = np.full_like(x, 7.0)
    plan.applyHost(x, dst, meta)
    assert np.array_equal(dst, want)
    inplace = x.copy()
    plan.applyHost(inplace, meta)
    assert np.array_equal(inplace, want)
    if name != "mute":
        ref = oracle.pipeline_apply(flow.ops, x.astype(np.float64), meta.row_len(1501), 2.0)
        assert rel_l2(dst, ref) <= (1e-5 if dtype == np.float32 else 1e-12)


def test_c_abi_host_pointers_and_regrowth(gpu):
    """The entry point through ctypes with numpy buffers only: a small block,
    then a larger one (the staging grows), then the small one again; after
    a release the staging is made again."""
    flow = FLOWS["bandpass_agc"]()
    plan = flow.compile(np.float32, 2.0, 3001)
    L = _lib.lib()
    for rows, cols in ((16, 1000), (600, 3001), (16, 1000)):
        x, meta = panel(rows, cols, seed=rows)
        rl = meta.row_len(cols)
        dst = np.empty_like(x)
        _lib.check(L.pseg_plan_apply_host(plan._h, x.ctypes.data, dst.ctypes.data, rows, cols,
                                          rl.ctypes.data, None, 0, None))
        assert np.array_equal(dst, device_apply(plan, x, meta)), (rows, cols)
    _lib.check(L.pseg_host_staging_release())
    x, meta = panel(32, 2000, seed=9)
    dst = np.empty_like(x)
    plan.applyHost(x, dst, meta)
    assert np.array_equal(dst, device_apply(plan, x, meta))


def test_refusals(gpu):
    flow = FLOWS["bandpass_agc"]()
    plan = flow.compile(np.float32, 2.0, 1000)
    x, meta = panel(8, 1001)
    with pytest.raises(ValueError, match="columns"):
        plan.applyHost(x, np.empty_like(x), meta)
    x, meta = panel(8, 1000)
    with pytest.raises(ValueError):
        plan.applyHost(x.astype(np.float64), None, meta)
    with pytest.raises(ValueError):
        plan.applyHost(x, np.empty_like(x), P.ChunkMeta(8, 4.0))
    L = _lib.lib()
    rl = meta.row_len(1000)
    buf = np.zeros(8 * 1000 + 10, np.float32)
    st = L.pseg_plan_apply_host(plan._h, buf.ctypes.data, buf.ctypes.data + 40, 8, 1000,
                                rl.ctypes.data, None, 0, None)
    assert st == _lib.E_INVALID
    st = L.pseg_plan_apply_host(plan._h, x.ctypes.data, x.ctypes.data, 8, 1000, None, None, 0,
                                None)
    assert st == _lib.E_INVALID
