"""SEG-Y and SEG-D sample decode on the GPU against the oracle, bit for bit:
every format (IBM float32, int32, int16, IEEE float32, int8; SEG-D 24-bit
and 32-bit integers, IEEE float32 and float64) to float32 and
float64, random words of every bit pattern (IEEE NaNs keep their payload in
float32), traces with 240-byte headers and odd sample counts, empty shapes,
the refusals, a decoded IEEE panel equal to the panel it was written from at
the full 2000 x 3001 size, and a decoded IBM panel through the headline
flow."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import workloads as W

pytestmark = pytest.mark.gpu

FMTS = [P.fmtIbm32, P.fmtInt32, P.fmtInt16, P.fmtIeee32, P.fmtInt8, P.fmtSegd8036,
        P.fmtSegd8038, P.fmtSegd8058, P.fmtSegd8080]
NB = {1: 4, 2: 4, 3: 2, 5: 4, 8: 1, 8036: 3, 8038: 4, 8058: 4, 8080: 8}


def traces(fmt, rows, ns, seed):
    rng = np.random.default_rng(seed)
    pitch = 240 + NB[fmt] * ns
    return rng.integers(0, 256, rows * pitch, dtype=np.uint8), pitch


def same(got, want):
    """Equal values, NaNs where the other has NaNs, and equal bits elsewhere."""
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint8), want[~nan].view(np.uint8))


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ns", [1, 7, 1001, 3001])
def test_decode_formats(gpu, oracle, fmt, dtype, ns):
    import torch
    rows = 9
    raw, pitch = traces(fmt, rows, ns, seed=fmt * 100 + ns)
    got = P.decodeSamples(torch.from_numpy(raw).cuda(), fmt, ns, rows=rows, dtype=dtype)
    torch.cuda.synchronize()
    want = oracle.decode_samples(raw, fmt, ns, rows, dtype=dtype)
    same(got.cpu().numpy(), want)
    if fmt == P.fmtIeee32 and dtype == np.float32:
        bits = raw.reshape(rows, pitch)[:, 240:].copy().view(">u4").astype(np.uint32)
        assert np.array_equal(got.cpu().numpy().view(np.uint32), bits), "IEEE words as they are"


def test_decode_shapes_and_refusals(gpu):
    import torch
    raw = torch.zeros(4 * 1000 + 240, dtype=torch.uint8, device="cuda")
    assert P.decodeSamples(raw, P.fmtIbm32, 0, rows=3).shape == (3, 0)
    assert P.decodeSamples(raw, P.fmtIbm32, 10, rows=0).shape == (0, 10)
    with pytest.raises(ValueError, match="multiples of 4"):
        P.decodeSamples(raw, P.fmtIbm32, 10, rows=2, pitch=282)
    with pytest.raises(ValueError, match="unknown sample format"):
        P.decodeSamples(raw, 4, 10, rows=1)
    with pytest.raises(ValueError, match="do not fit"):
        P.decodeSamples(raw, P.fmtIbm32, 1000, rows=2)
    # 2-byte words need no alignment
    got = P.decodeSamples(raw, P.fmtInt16, 100, rows=3, pitch=241, offset=1)
    assert got.shape == (3, 100) and float(got.abs().max()) == 0.0


def test_ieee_panel_round_trip(gpu):
    """The 2000 x 3001 float32 panel written as SEG-Y IEEE traces decodes to
    itself, bit for bit."""
    import torch
    x, _ = W.big_panel()
    rows, ns = x.shape
    blob = np.zeros((rows, 240 + 4 * ns), np.uint8)
    blob[:, 240:] = x.astype(">f4").view(np.uint8).reshape(rows, 4 * ns)
    got = P.decodeSamples(torch.from_numpy(blob.reshape(-1)).cuda(), P.fmtIeee32, ns, rows=rows)
    assert torch.equal(got.cpu(), torch.from_numpy(x))


def test_decode_then_flow(gpu, oracle):
    """IBM traces decoded on the device, then the headline flow on them, equal
    the oracle's decode then the oracle's flow within the float32 bar."""
    import torch
    from tests.common import rel_l2
    rows, ns = 64, 3001
    x, live = W.big_panel()
    x = x[:rows].astype(np.float64)
    # encode the panel as IBM words (truncating to 24 fraction bits)
    s = (x < 0).astype(np.uint32)
    a = np.abs(x)
    e = np.where(a > 0, np.floor(np.log2(np.maximum(a, 1e-300)) / 4).astype(np.int64) + 1, 0)
    f = np.where(a > 0, np.floor(a / 16.0 ** e * 2 ** 24), 0).astype(np.uint64)
    w = (s << 31) | (((e + 64) & 0x7F).astype(np.uint32) << 24) | (f.astype(np.uint32) & 0xFFFFFF)
    blob = np.zeros((rows, 240 + 4 * ns), np.uint8)
    blob[:, 240:] = w.astype(">u4").view(np.uint8).reshape(rows, 4 * ns)
    dec = P.decodeSamples(torch.from_numpy(blob.reshape(-1)).cuda(), P.fmtIbm32, ns, rows=rows)
    want_dec = oracle.decode_samples(blob.reshape(-1), P.fmtIbm32, ns, rows)
    same(dec.cpu().numpy(), want_dec)
    flow = W.big_flow()
    plan = flow.compile(np.float32, 2.0, ns)
    rl = torch.from_numpy(live[:rows].astype(np.int32)).cuda()
    out = torch.empty_like(dec)
    plan.apply(dec, out, row_len=rl)
    want = oracle.pipeline_apply(flow.ops, want_dec.astype(np.float64), live[:rows].astype(np.int32), 2.0)
    got = out.cpu().numpy()
    for r in range(rows):
        n = int(live[r])
        assert rel_l2(got[r, :n], want[r, :n]) <= 1e-5


def ibm_words(x):
    """x (float64) as IBM float32 words, the fraction truncated to 24 bits."""
    s = (x < 0).astype(np.uint32)
    a = np.abs(x)
    e = np.where(a > 0, np.floor(np.log2(np.maximum(a, 1e-300)) / 4).astype(np.int64) + 1, 0)
    f = np.where(a > 0, np.floor(a / 16.0 ** e * 2 ** 24), 0).astype(np.uint64)
    return (s << 31) | (((e + 64) & 0x7F).astype(np.uint32) << 24) | (f.astype(np.uint32) & 0xFFFFFF)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fmt", [P.fmtIbm32, P.fmtInt16, P.fmtIeee32, P.fmtSegd8036,
                                 P.fmtSegd8080])
def test_apply_raw_equals_decode_then_apply(gpu, oracle, dtype, fmt):
    """Plan.applyRaw on raw traces equals decodeSamples then Plan.apply:
    bit for bit, for a core flow and for a flow with a mute and a gain curve
    that read the rows' keys in a window; and the oracle within the bar."""
    import torch
    from tests.common import rel_l2
    rows, ns, first = 48, 2001, 100
    x, live = W.big_panel(rows, ns, seed=5)
    xd = x.astype(np.float64) * 1000
    if fmt == P.fmtIbm32:
        w, b = ibm_words(xd), ">u4"
    elif fmt == P.fmtInt16:
        w, b = np.clip(np.round(xd), -32768, 32767).astype(np.int64), ">i2"
    elif fmt == P.fmtSegd8036:
        w, b = np.clip(np.round(xd * 100), -2 ** 23, 2 ** 23 - 1).astype(np.int64), "i24"
    elif fmt == P.fmtSegd8080:
        w, b = xd, ">f8"
    else:
        w, b = xd.astype(np.float32), ">f4"
    nb = 3 if b == "i24" else np.dtype(b).itemsize
    blob = np.zeros((rows, 240 + nb * ns), np.uint8)
    if b == "i24":
        u = np.asarray(w) & 0xFFFFFF
        b3 = np.stack([(u >> 16) & 0xFF, (u >> 8) & 0xFF, u & 0xFF], axis=-1).astype(np.uint8)
        blob[:, 240:] = b3.reshape(rows, 3 * ns)
    else:
        blob[:, 240:] = np.asarray(w).astype(b).view(np.uint8).reshape(rows, nb * ns)
    raw = torch.from_numpy(blob.reshape(-1)).cuda()
    meta = P.ChunkMeta(rows, 2.0, numSamples=live + first, firstSample=first,
                       t0Ms=np.linspace(0, 30, rows), offset=np.linspace(100, 2500, rows))
    flows = [W.big_flow(),
             P.initPipeline([P.demeanOp(), P.bandPassOp(P.Corners(4, 8, 60, 80)),
                             P.muteTopOp(P.lmoMute(2500.0), taperMs=20),
                             P.gainCurveOp(P.powerGain(2.0)), P.agcOp(300)])]
    dec = P.decodeSamples(raw, fmt, ns, rows=rows, dtype=dtype)
    for flow in flows:
        plan = flow.compile(dtype, 2.0, ns)
        two = torch.empty_like(dec)
        plan.apply(dec, two, meta)
        one = torch.empty_like(dec)
        plan.applyRaw(raw, fmt, one, meta)
        torch.cuda.synchronize()
        assert torch.equal(one, two), "one pass differs from decode then apply"
        rl = meta.row_len(ns)
        want = oracle.pipeline_apply(flow.ops, dec.cpu().numpy().astype(np.float64), rl, 2.0,
                                     keys=meta.keys(), first_sample=first)
        got = one.cpu().numpy()
        tol = 1e-5 if dtype == np.float32 else 1e-12
        for r in range(rows):
            n = int(rl[r])
            if n > 0:
                assert rel_l2(got[r, :n], want[r, :n]) <= tol, r
    with pytest.raises(P.UnsupportedOp):
        panel = P.initPipeline([P.normalizeOp(P.normRms, P.perPanel)]).compile(dtype, 2.0, ns)
        panel.applyRaw(raw, fmt, torch.empty_like(dec), meta)
