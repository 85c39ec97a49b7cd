"""Uneven batches: many records with their own plans (sample intervals 1, 2
and 4 ms; 1000 to 6001 samples; a few rows each) as persistent launches.
Each record must equal plan.apply of that record alone bit for bit, and the
float64 oracle within the float32 tolerance; a captured graph replays to the
same bits."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import synth, workloads as W
from tests.common import rel_l2

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5


def make_records(n=24, seed=11):
    shapes = W.mixed_shapes(n, rows=(1, 9), cols=(1000, 6001), seed=seed)
    recs = []
    for i, (rows, cols, dt) in enumerate(shapes):
        x, _ = synth.gather(rows, cols, dt, seed + i)
        live = synth.live_lengths(seed + i, rows, cols // 2, cols).astype(np.int32)
        if rows > 2:
            live[1] = -1          # a dead or failed row: copied unchanged
        if rows > 3:
            live[2] = 0
        recs.append((x, live, dt))
    return recs


def plans_for(recs, flow):
    cache = {}
    out = []
    for x, _, dt in recs:
        key = (dt, x.shape[1])
        if key not in cache:
            cache[key] = flow.compile(np.float32, dt, x.shape[1])
        out.append(cache[key])
    return out


def test_batch_equals_single_applies_and_oracle(gpu, oracle):
    import torch
    recs = make_records()
    flow = W.mixed_flow()
    plans = plans_for(recs, flow)
    srcs = [torch.from_numpy(x).cuda() for x, _, _ in recs]
    dsts = [torch.full_like(s, 9.0) for s in srcs]
    lens = [torch.from_numpy(l).cuda() for _, l, _ in recs]
    batch = P.PipelineBatch([P.BatchItem(p, s, d, row_len=l)
                             for p, s, d, l in zip(plans, srcs, dsts, lens)])
    assert batch.launches == 2 and sorted(batch.threads) == [128, 256]
    batch.launch()
    alone = [torch.empty_like(s) for s in srcs]
    for p, s, a, l in zip(plans, srcs, alone, lens):
        p.apply(s, a, row_len=l)
    torch.cuda.synchronize()
    worst = 0.0
    for i, ((x, live, dt), d, a) in enumerate(zip(recs, dsts, alone)):
        got = d.cpu().numpy()
        assert np.array_equal(got, a.cpu().numpy()), f"record {i}: batch != alone"
        want = oracle.pipeline_apply(flow.ops, x.astype(np.float64), live, dt)
        e = rel_l2(got, want)
        worst = max(worst, e)
        assert e <= F32_TOL, f"record {i} {x.shape} dt {dt}: rel L2 {e}"
        if x.shape[0] > 2:
            assert np.array_equal(got[1], x[1])
    print("batch of", len(recs), "records: worst rel L2 vs FP64 oracle", worst)


def test_batch_in_place(gpu):
    import torch
    recs = make_records(8, seed=5)
    flow = W.mixed_flow()
    plans = plans_for(recs, flow)
    bufs = [torch.from_numpy(x).cuda() for x, _, _ in recs]
    ref = [torch.empty_like(b) for b in bufs]
    lens = [torch.from_numpy(l).cuda() for _, l, _ in recs]
    for p, b, r, l in zip(plans, bufs, ref, lens):
        p.apply(b, r, row_len=l)
    batch = P.PipelineBatch([P.BatchItem(p, b, row_len=l) for p, b, l in zip(plans, bufs, lens)])
    batch.launch()
    torch.cuda.synchronize()
    for b, r in zip(bufs, ref):
        assert torch.equal(b, r)


def test_batch_graph_replay(gpu):
    """The batch captured in a graph: every replay (with eager work and a
    synchronization between replays, and dst cleared first) gives the eager
    bits."""
    import torch
    recs = make_records(12, seed=7)
    flow = W.mixed_flow()
    plans = plans_for(recs, flow)
    srcs = [torch.from_numpy(x).cuda() for x, _, _ in recs]
    dsts = [torch.empty_like(s) for s in srcs]
    lens = [torch.from_numpy(l).cuda() for _, l, _ in recs]
    batch = P.PipelineBatch([P.BatchItem(p, s, d, row_len=l)
                             for p, s, d, l in zip(plans, srcs, dsts, lens)])
    batch.launch()
    torch.cuda.synchronize()
    eager = [d.clone() for d in dsts]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batch.launch()
    torch.cuda.synchronize()
    for rep in range(3):
        for d in dsts:
            d.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i, (d, e) in enumerate(zip(dsts, eager)):
            assert torch.equal(d, e), f"replay {rep}, record {i}"
        _ = sum(float(s.sum()) for s in srcs)  # eager work between replays


def test_batch_refuses_overlapping_items(gpu):
    import torch
    flow = W.mixed_flow()
    plan = flow.compile(np.float32, 2.0, 1000)
    a = torch.zeros(8, 1000, device="cuda")
    l = torch.full((8,), 1000, dtype=torch.int32, device="cuda")
    with pytest.raises(ValueError, match="overlaps"):
        P.PipelineBatch([P.BatchItem(plan, a, row_len=l), P.BatchItem(plan, a[:4], row_len=l)])


def test_batch_of_flows_reading_trace_keys(gpu):
    """Records whose flows read the rows' keys — a top LMO mute and a t²
    gain curve, an IIR band-pass and a bottom table mute — with their own
    t0Ms, offsets and windows (firstSample), in one batch: each record
    equals plan.apply of that record with its ChunkMeta, bit for bit."""
    import torch
    flows = [
        P.initPipeline([P.demeanOp(), P.muteTopOp(P.lmoMute(2000.0), taperMs=20),
                        P.gainCurveOp(P.powerGain(2.0)), P.agcOp(300)]),
        P.initPipeline([P.iirOp(P.butterBandPass(4, 60, 24)),
                        P.muteBottomOp(P.tableMute([0.0, 2000], [1500.0, 2600]), taperMs=40)]),
    ]
    recs = make_records(12, seed=23)
    items, metas, want = [], [], []
    for i, (x, live, dt) in enumerate(recs):
        rows, cols = x.shape
        first = (i % 3) * 50
        meta = P.ChunkMeta(rows, dt, numSamples=np.maximum(live, 0) + first,
                           t0Ms=np.linspace(0, 40, rows), offset=np.linspace(-1500, 2500, rows),
                           firstSample=first)
        plan = flows[i % 2].compile(np.float32, dt, cols)
        src = torch.from_numpy(x).cuda()
        ref = torch.empty_like(src)
        plan.apply(src, ref, meta)
        want.append(ref)
        dst = torch.empty_like(src)
        items.append(P.BatchItem(plan, src, dst, meta=meta))
        metas.append(meta)
    batch = P.PipelineBatch(items)
    batch.launch()
    torch.cuda.synchronize()
    for i, (it, ref) in enumerate(zip(items, want)):
        assert torch.equal(it.dst, ref), f"record {i} differs from its own apply"


def test_tuned_launch_shape_keeps_the_bits(gpu):
    """pseg_batch_tune picks each class's launch shape by timing (persistent
    grids of 1/2, 3/4 and all resident workgroups, or one workgroup per
    row); the batch then gives the same bits as before tuning, in eager
    launches and in a captured graph; an in-place batch is refused."""
    import torch
    recs = make_records(32, seed=21)
    flow = W.mixed_flow()
    plans = plans_for(recs, flow)
    srcs = [torch.from_numpy(x).cuda() for x, _, _ in recs]
    lens = [torch.from_numpy(l).cuda() for _, l, _ in recs]
    before = [torch.full_like(s, 7.0) for s in srcs]
    P.PipelineBatch([P.BatchItem(p, s, d, row_len=l)
                     for p, s, d, l in zip(plans, srcs, before, lens)]).launch()
    dsts = [torch.full_like(s, 5.0) for s in srcs]
    batch = P.PipelineBatch([P.BatchItem(p, s, d, row_len=l)
                             for p, s, d, l in zip(plans, srcs, dsts, lens)])
    t = batch.tune(repeats=2)
    assert len(t["grid"]) == batch.launches
    for g in range(batch.launches):
        assert t["best_ms"][g] > 0 and t["best_ms"][g] == min(x for x in t["candidate_ms"][g] if x > 0)
        assert (batch.grid[g] == 0) == t["direct"][g]
    for d in dsts:
        d.fill_(3.0)
    batch.launch()
    torch.cuda.synchronize()
    for a, b in zip(before, dsts):
        assert torch.equal(a, b)
    for d in dsts:
        d.fill_(1.0)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        batch.launch(stream=s)
    for _ in range(2):
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(before, dsts):
            assert torch.equal(a, b)
    inplace = P.PipelineBatch([P.BatchItem(plans[0], srcs[0].clone(), None, row_len=lens[0])])
    with pytest.raises(ValueError, match="in place"):
        inplace.tune()
