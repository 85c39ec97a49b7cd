#!/usr/bin/env python3
"""Throughput of the panel hot path on one MI355X.

Workload (the headline): demean -> band-pass (cos² 5-10-70-90 Hz with a 50 Hz
notch 4 Hz wide) -> Butterworth high cut 80 Hz 24 dB/oct -> AGC 250 ms on a
2000 x 3001 float32 panel at 2 ms, live lengths uniform in [2400, 3001] per
row, synthetic seeded data (Ricker wavelets on hyperbolic events + Gaussian
noise).  A step is one Plan.apply src -> dst of one panel; inputs are resident
in HBM and 8 distinct src/dst pairs rotate (384 MB, beyond the 256 MB
Infinity Cache), so every step reads and writes HBM.

Prints one JSON line: Msamples/s, the roofline of the fused kernel (HIP events
on its stream), the accuracy against the float64 oracle on a sample of rows,
and the CPU baseline (the oracle's ProcessPool clone on the host cores).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Vector-ALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64
# instruction = 614.4 G wave-instructions/s (78.6 TFLOP/s of non-packed fp32 FMA).
VALU_PEAK_GWINST = 256 * 4 * 2.4 / 4
ROWS, COLS, DT = 2000, 3001, 2.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_flow(n_ops=None):
    """The headline flow, or its first `n_ops` ops (the per-step profile
    runs the flow cut after each op)."""
    from processseg_amd.workloads import big_flow
    flow = big_flow()
    if n_ops is not None:
        import processseg_amd as P
        flow = P.initPipeline(flow.ops[:n_ops])
    return flow


FLOW_STEPS = ["demean", "band-pass 5-10-70-90 Hz + 50 Hz notch",
              "Butterworth high cut 80 Hz 24 dB/oct", "AGC 250 ms"]


def flow_name(n_ops=None):
    return " > ".join(FLOW_STEPS[:n_ops])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _timed_samples(fn, per_sample, warm=3, samples=15):
    """`warm` untimed samples, then `samples` timed ones, each `per_sample`
    calls of fn; returns the seconds per call of every timed sample."""
    for _ in range(warm):
        for _ in range(per_sample):
            fn()
    out = []
    for _ in range(samples):
        t0 = time.perf_counter()
        for _ in range(per_sample):
            fn()
        out.append((time.perf_counter() - t0) / per_sample)
    return out


def cpu_baseline_leg(flow, x, live, gpu_out):
    """The oracle (a C restatement of the reference) on the host: the
    ProcessPool clone over the whole panel on the host cores the process may
    use (at most 16, one GPU's share of a box), 3 warm-up and 15 timed
    samples of about half a second each, median and best; then one core the
    same way (samples of two applies); and the accuracy of the GPU output
    against the float64 oracle on 64 rows."""
    import oracle as O
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except Exception:
        avail = nproc
    cores = min(avail, 16)
    row_len = live.astype(np.int32)
    dst = np.empty_like(x)
    res = {}
    for name, threads, per in (("all", cores, None), ("one", 1, 2)):
        pool = O.Pool(threads)
        fn = lambda: pool.apply_f32(flow.ops, x, dst, row_len, DT)  # noqa: E731
        if per is None:
            t0 = time.perf_counter()
            fn()
            per = max(1, int(0.5 / max(time.perf_counter() - t0, 1e-4)))
        secs = _timed_samples(fn, per)
        pool.close()
        rates = [x.size / s / 1e6 for s in secs]
        res[name] = {"median": float(np.median(rates)), "best": float(np.max(rates)),
                     "applies": per * 18, "threads": threads}
    a, o = res["all"], res["one"]
    # accuracy vs the float64 oracle on 64 rows spread over the panel
    idx = np.linspace(0, x.shape[0] - 1, 64).astype(int)
    want = O.pipeline_apply(flow.ops, x[idx].astype(np.float64), row_len[idx], DT)
    got = gpu_out[idx].astype(np.float64)
    rel = float(np.linalg.norm(got - want) / np.linalg.norm(want))
    mabs = float(np.max(np.abs(got - want)))
    mirror = dst[idx].astype(np.float64)
    rel_cpu = float(np.linalg.norm(mirror - want) / np.linalg.norm(want))
    return ({"value": round(a["median"], 2), "unit": "Msamples/s", "cores": cores, "kind": "port",
             "best": round(a["best"], 2),
             "one_core": {"value": round(o["median"], 2), "best": round(o["best"], 2)},
             "cpu": _cpu_model(), "nproc": nproc, "affinity": avail,
             "sample": f"the whole {x.shape[0]}x{x.shape[1]} panel: 3 warm-up + 15 timed samples "
                       f"of {a['applies'] // 18} applies on {cores} threads (ProcessPool clone, "
                       f"runs of 256 KiB), median (best {a['best']:.0f}); one core: samples of "
                       f"{o['applies'] // 18} applies"},
            {"vs": "float64 oracle, 64 rows", "rel_l2": rel, "max_abs": mabs,
             "cpu_mirror_rel_l2": rel_cpu})


def kernel_name(cols, threads, dtype="float"):
    """The instantiation pseg_plan_apply launches for rows of `cols` samples
    with the plan's `threads` per row: the row lives in registers,
    ceil(cols / threads) samples per thread, rounded up to a compiled width;
    the form with demean / filter / AGC only (form 0)."""
    per = -(-cols // threads)
    widths = (8, 16, 24) if threads == 512 else (4, 8, 12, 16, 24)
    per = min(w for w in widths if w >= per)
    return f"pseg::pipeline_kernel<{dtype}, {threads}, {per}, 0>"


def copy_probe(torch, dev, args):
    """A device copy of one 2000 x 3001 f32 panel (24.008 MB read, 24.008 MB
    written) between rotating buffers: the known-size kernel that calibrates
    the FETCH_SIZE / WRITE_SIZE counters."""
    srcs = [torch.ones(ROWS, COLS, device=dev) for _ in range(args.pairs)]
    dsts = [torch.empty_like(srcs[0]) for _ in range(args.pairs)]
    for i in range(args.warmup):
        dsts[i % args.pairs].copy_(srcs[i % args.pairs])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        dsts[i % args.pairs].copy_(srcs[i % args.pairs])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000 / args.steps
    nbytes = ROWS * COLS * 4
    print(json.dumps({"probe": "device copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "ms_per_copy": round(ms, 5),
                      "GBps": round(2 * nbytes / (ms / 1000) / 1e9, 1)}), flush=True)


def correlation_leg(torch, dev, steps, warmup, pairs, with_cpu):
    """Vibroseis correlation of 1200 records x 10001 samples at 2 ms with a
    12 s sweep (nfft 16200, 4001 output lags): one correlateInto per step over
    rotating src/dst pairs.  Algorithmic bytes per record: (10001 + 4001) x 4."""
    import processseg_amd as P
    from processseg_amd.workloads import sweep_pilot, vibroseis_panel
    rows, rec = 1200, 10001
    pilot = sweep_pilot()
    x = vibroseis_panel(rows, rec)
    plan = P.initCorrelationPlan(np.float32, pilot, DT, rec)
    row_len = torch.full((rows,), rec, dtype=torch.int32, device=dev)
    srcs = [torch.from_numpy(x).to(dev) for _ in range(pairs)]
    dsts = [torch.empty(rows, plan.outSamples, device=dev) for _ in range(pairs)]
    stream = torch.cuda.current_stream()
    for i in range(warmup):
        plan.correlateInto(srcs[i % pairs], dsts[i % pairs], row_len=row_len, stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        plan.correlateInto(srcs[i % pairs], dsts[i % pairs], row_len=row_len, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # launch durations: the same launches again between events, untimed
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for i in range(steps):
        starts[i].record(stream)
        plan.correlateInto(srcs[i % pairs], dsts[i % pairs], row_len=row_len, stream=stream)
        ends[i].record(stream)
    torch.cuda.synchronize()
    avg_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    algo = rows * (rec + plan.outSamples) * 4
    out = {"workload": "vibroseis correlation 1200x10001 f32, 12 s 6-80 Hz sweep, nfft 16200, "
                       "4001 lags",
           "value": round(rows * rec / (wall / steps) / 1e6, 1), "unit": "Msamples/s (input)",
           "ms_per_step": round(wall * 1000 / steps, 5),
           "roofline": {"bound": "hbm", "achieved": round(algo / (avg_ms / 1000) / 1e9, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(algo / (avg_ms / 1000) / 1e9 / HBM_PEAK_GBS, 4),
                        "kernel": f"pseg::corr_kernel<float, {plan.threadsPerRecord}>",
                        "avg_launch_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": algo}}
    if with_cpu:
        import oracle as O
        got = dsts[(steps - 1) % pairs][:32].cpu().numpy().astype(np.float64)
        want = O.correlate(x[:32].astype(np.float64), np.full(32, rec, np.int32), pilot.samples,
                           DT, rec)
        sub = x[:400]
        t1 = time.perf_counter()
        O.correlate(sub.astype(np.float32), np.full(400, rec, np.int32), pilot.samples, DT, rec)
        el32 = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": round(400 * rec / el32 / 1e6, 2), "unit": "Msamples/s",
                               "cores": 1, "kind": "port",
                               "sample": f"400 records in float32, {el32:.2f} s on one core"}
        out["accuracy"] = {"vs": "float64 oracle, 32 records",
                           "rel_l2": float(np.linalg.norm(got - want) / np.linalg.norm(want)),
                           "max_abs_over_peak": float(np.max(np.abs(got - want))
                                                      / np.max(np.abs(want)))}
    return out


def panels_batch_leg(torch, P, plan, srcs, dsts, row_len, steps, warmup):
    """The same plan on the 8 rotating panels as one PipelineBatch (16000
    rows, one persistent launch): what a caller with many panels at hand
    gets, the ends of consecutive panels overlapping.  Reported beside the
    headline, which applies one panel per step as Plan.apply does."""
    batch = P.PipelineBatch([P.BatchItem(plan, s, d, row_len=row_len)
                             for s, d in zip(srcs, dsts)])

    def timed():
        for _ in range(warmup):
            batch.launch()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.launch()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps
    t_rule = timed()
    grid_rule = list(batch.grid)
    batch.tune()          # the launch shape by timing (the same bits)
    t = timed()
    samples = batch.samples
    return {"workload": f"{len(srcs)} panels of 2000x3001 f32, the headline flow, one batch",
            "value": round(samples / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_batch": round(t * 1000, 4), "grid": batch.grid,
            "ms_before_tuning": round(t_rule * 1000, 4), "grid_before_tuning": grid_rule,
            "achieved_GBps": round(samples * 8 / t / 1e9, 1)}


def f64_panel_leg(torch, P, flow, dev, steps, warmup):
    """The headline flow and panel in float64 (the big-panel configuration
    is float32 and float64): one Plan.apply per step, 4 buffer pairs."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel(dtype=np.float64)
    plan = flow.compile(np.float64, DT, COLS)
    srcs = [torch.from_numpy(x).to(dev) for _ in range(4)]
    dsts = [torch.empty_like(srcs[0]) for _ in range(4)]
    rl = torch.from_numpy(live.astype(np.int32)).to(dev)
    for i in range(warmup):
        plan.apply(srcs[i % 4], dsts[i % 4], row_len=rl)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        plan.apply(srcs[i % 4], dsts[i % 4], row_len=rl)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    samples = ROWS * COLS
    return {"workload": "the headline flow on the 2000x3001 panel in float64",
            "value": round(samples / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5), "threads_per_row": plan.threadsPerRow,
            "achieved_GBps": round(samples * 16 / t / 1e9, 1)}


def muted_flow_leg(torch, P, dev, steps, warmup):
    """The headline panel with mutes in the flow (the kernel form with every
    step): demean > band-pass > top LMO mute (2500 m/s, 20 ms taper) > AGC >
    bottom table mute (50 ms taper), the rows' offsets -2400 .. 2400 m and
    t0 0 from the ChunkMeta; one Plan.apply per step, 4 buffer pairs."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live, offset=np.linspace(-2400, 2400, ROWS))
    flow = P.initPipeline([
        P.demeanOp(), P.bandPassOp(P.Corners(5, 10, 70, 90)),
        P.muteTopOp(P.lmoMute(2500.0, padMs=-12.0), taperMs=20), P.agcOp(250),
        P.muteBottomOp(P.tableMute([0.0, 800, 2400], [2600.0, 3600, 5200]), taperMs=50)])
    plan = flow.compile(np.float32, DT, COLS)
    srcs = [torch.from_numpy(x).to(dev) for _ in range(4)]
    dsts = [torch.empty_like(srcs[0]) for _ in range(4)]
    for i in range(warmup):
        plan.apply(srcs[i % 4], dsts[i % 4], meta)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        plan.apply(srcs[i % 4], dsts[i % 4], meta)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    samples = ROWS * COLS
    return {"workload": "demean > band-pass 5-10-70-90 Hz > top LMO mute > AGC 250 ms > "
                        "bottom table mute, 2000x3001 f32, offsets from the meta",
            "value": round(samples / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5), "achieved_GBps": round(samples * 8 / t / 1e9, 1)}


def gain_flow_leg(torch, P, dev, steps, warmup):
    """The headline panel with a time-variant gain in the flow (the kernel
    form with every step): demean > band-pass > spherical divergence
    (powerGain 2, (t / 1 s)²) > AGC, the rows' t0 from the ChunkMeta (every
    fourth row delayed 33.3 ms, off the table's grid); one Plan.apply per
    step, 4 buffer pairs."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    t0 = np.where(np.arange(ROWS) % 4 == 3, 33.3, 0.0)
    meta = P.ChunkMeta(ROWS, DT, numSamples=live, t0Ms=t0)
    flow = P.initPipeline([
        P.demeanOp(), P.bandPassOp(P.Corners(5, 10, 70, 90)),
        P.gainCurveOp(P.powerGain(2.0)), P.agcOp(250)])
    plan = flow.compile(np.float32, DT, COLS)
    srcs = [torch.from_numpy(x).to(dev) for _ in range(4)]
    dsts = [torch.empty_like(srcs[0]) for _ in range(4)]
    for i in range(warmup):
        plan.apply(srcs[i % 4], dsts[i % 4], meta)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        plan.apply(srcs[i % 4], dsts[i % 4], meta)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t1) / steps
    samples = ROWS * COLS
    return {"workload": "demean > band-pass 5-10-70-90 Hz > powerGain 2 (t^2) > AGC 250 ms, "
                        "2000x3001 f32, t0 from the meta (a quarter of the rows off the grid)",
            "value": round(samples / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5), "achieved_GBps": round(samples * 8 / t / 1e9, 1)}


def iir_flow_leg(torch, P, dev, steps, warmup):
    """The headline panel through a time-domain Butterworth band-pass (the
    kernel form with every step): demean > zero-phase IIR band-pass 5-70 Hz
    24 dB/octave (two biquads, forward and backward: four scanned passes) >
    AGC 250 ms; one Plan.apply per step, 4 buffer pairs."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live)
    flow = P.initPipeline([P.demeanOp(), P.iirOp(P.butterBandPass(5, 70, 24)), P.agcOp(250)])
    plan = flow.compile(np.float32, DT, COLS)
    srcs = [torch.from_numpy(x).to(dev) for _ in range(4)]
    dsts = [torch.empty_like(srcs[0]) for _ in range(4)]
    for i in range(warmup):
        plan.apply(srcs[i % 4], dsts[i % 4], meta)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        plan.apply(srcs[i % 4], dsts[i % 4], meta)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t1) / steps
    samples = ROWS * COLS
    return {"workload": "demean > zero-phase IIR Butterworth band-pass 5-70 Hz 24 dB/oct > "
                        "AGC 250 ms, 2000x3001 f32",
            "value": round(samples / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5), "achieved_GBps": round(samples * 8 / t / 1e9, 1),
            "margin": plan.requiredMargin}


def decode_leg(torch, P, dev, steps, warmup):
    """SEG-Y decode of the 2000 x 3001 panel's traces (240-byte headers, IBM
    float32 samples, big-endian) to a float32 panel on the device, 8 rotating
    buffer pairs (450 MB, beyond the Infinity Cache).  Algorithmic bytes:
    4 read + 4 written per sample."""
    from processseg_amd.workloads import big_panel
    x, _ = big_panel()
    rows, ns = x.shape
    a = np.abs(x.astype(np.float64))
    e = np.where(a > 0, np.floor(np.log2(np.maximum(a, 1e-300)) / 4).astype(np.int64) + 1, 0)
    f = np.where(a > 0, np.floor(a / 16.0 ** e * 2 ** 24), 0).astype(np.uint64)
    w = (((x < 0).astype(np.uint32) << 31) | (((e + 64) & 0x7F).astype(np.uint32) << 24)
         | (f.astype(np.uint32) & 0xFFFFFF))
    blob = np.zeros((rows, 240 + 4 * ns), np.uint8)
    blob[:, 240:] = w.astype(">u4").view(np.uint8).reshape(rows, 4 * ns)
    raws = [torch.from_numpy(blob.reshape(-1)).to(dev) for _ in range(8)]
    outs = [torch.empty(rows, ns, device=dev) for _ in range(8)]
    for i in range(warmup):
        P.decodeSamples(raws[i % 8], P.fmtIbm32, ns, rows=rows, out=outs[i % 8])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        P.decodeSamples(raws[i % 8], P.fmtIbm32, ns, rows=rows, out=outs[i % 8])
    torch.cuda.synchronize()
    t = (time.perf_counter() - t1) / steps
    algo = rows * ns * 8
    return {"workload": "SEG-Y IBM float32 traces (240-byte headers) of the 2000x3001 panel to "
                        "float32 on the device",
            "value": round(rows * ns / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5),
            "roofline": {"bound": "hbm", "achieved": round(algo / t / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_launch": algo,
                         "note": "wall clock over back-to-back launches, not kernel events"}}


def host_staged_leg(torch, P, dev, steps, warmup):
    """The headline flow on a panel held in host memory (Plan.applyHost:
    pseg_plan_apply_host copies the 2000 x 3001 block to the device, runs
    the plan, copies it back): the rate a caller of the reference's
    Tensor-in-host-memory apply sees, PCIe copies included (pageable host
    arrays)."""
    from processseg_amd.workloads import big_flow, big_panel
    x, live = big_panel()
    rows, cols = x.shape
    meta = P.ChunkMeta(rows, DT, numSamples=live)
    plan = big_flow().compile(np.float32, DT, cols)
    dst = np.empty_like(x)
    for _ in range(warmup):
        plan.applyHost(x, dst, meta)
    t1 = time.perf_counter()
    for _ in range(steps):
        plan.applyHost(x, dst, meta)
    t = (time.perf_counter() - t1) / steps
    return {"workload": "the headline flow on a 2000x3001 f32 panel in host memory (copies to "
                        "and from the device included)",
            "value": round(rows * cols / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 4),
            "host_bytes_per_step": rows * cols * 8}


def raw_flow_leg(torch, P, dev, steps, warmup):
    """The headline flow from raw SEG-Y IBM traces of the 2000 x 3001 panel:
    Plan.applyRaw (each row decoded as the kernel loads it) against
    decodeSamples then Plan.apply (two passes), 4 buffer sets each."""
    from processseg_amd.workloads import big_flow, big_panel
    x, live = big_panel()
    rows, ns = x.shape
    a = np.abs(x.astype(np.float64))
    e = np.where(a > 0, np.floor(np.log2(np.maximum(a, 1e-300)) / 4).astype(np.int64) + 1, 0)
    f = np.where(a > 0, np.floor(a / 16.0 ** e * 2 ** 24), 0).astype(np.uint64)
    w = (((x < 0).astype(np.uint32) << 31) | (((e + 64) & 0x7F).astype(np.uint32) << 24)
         | (f.astype(np.uint32) & 0xFFFFFF))
    blob = np.zeros((rows, 240 + 4 * ns), np.uint8)
    blob[:, 240:] = w.astype(">u4").view(np.uint8).reshape(rows, 4 * ns)
    meta = P.ChunkMeta(rows, DT, numSamples=live)
    plan = big_flow().compile(np.float32, DT, ns)
    raws = [torch.from_numpy(blob.reshape(-1)).to(dev) for _ in range(4)]
    mids = [torch.empty(rows, ns, device=dev) for _ in range(4)]
    outs = [torch.empty(rows, ns, device=dev) for _ in range(4)]

    def one(i):
        plan.applyRaw(raws[i % 4], P.fmtIbm32, outs[i % 4], meta)

    def two(i):
        P.decodeSamples(raws[i % 4], P.fmtIbm32, ns, rows=rows, out=mids[i % 4])
        plan.apply(mids[i % 4], outs[i % 4], meta)

    res = {}
    for name, fn in (("one_pass", one), ("decode_then_apply", two)):
        for i in range(warmup):
            fn(i)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(steps):
            fn(i)
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - t1) / steps
    return {"workload": "the headline flow from SEG-Y IBM float32 traces (240-byte headers) of "
                        "the 2000x3001 panel",
            "value": round(rows * ns / res["one_pass"] / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(res["one_pass"] * 1000, 5),
            "ms_decode_then_apply": round(res["decode_then_apply"] * 1000, 5)}


def spectrum_leg(torch, P, dev, steps, warmup):
    """The RMS amplitude spectrum of the 2000 x 3001 panel's windows of 2001
    samples from column 500 (10 % Tukey taper, nfft 2048, 1025 bins): one
    averageAmplitudeInto per step.  Algorithmic bytes: the windows read."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live)
    sp = P.initSpectrumPlan(np.float32, DT, 2001)
    src = torch.from_numpy(x).to(dev)
    rl = meta.device_row_len(COLS, dev)
    amp = torch.empty(sp.bins, dtype=torch.float64, device=dev)
    ws = torch.empty(sp.workspaceSize(ROWS), dtype=torch.uint8, device=dev)
    used = torch.empty(1, dtype=torch.float64, device=dev)
    for _ in range(warmup):
        sp.averageAmplitudeInto(src, firstCol=500, amp=amp, row_len=rl, ws=ws, used=used,
                                sync=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        sp.averageAmplitudeInto(src, firstCol=500, amp=amp, row_len=rl, ws=ws, used=used,
                                sync=False)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t1) / steps
    return {"workload": "RMS amplitude spectrum of 2000 windows of 2001 samples (nfft 2048) of "
                        "the 2000x3001 f32 panel",
            "value": round(ROWS * 2001 / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5)}


def hilbert_leg(torch, P, dev, steps, warmup):
    """The envelope of every live row of the 2000 x 3001 panel (nfft 6144,
    no wrap): one envelopeInto per step into rotating outputs."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live)
    hp = P.initHilbertPlan(np.float32, COLS)
    src = torch.from_numpy(x).to(dev)
    rl = meta.device_row_len(COLS, dev)
    outs = [torch.empty_like(src) for _ in range(4)]
    for i in range(warmup):
        hp.envelopeInto(src, outs[i % 4], row_len=rl)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        hp.envelopeInto(src, outs[i % 4], row_len=rl)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t1) / steps
    return {"workload": "envelope (Hilbert transform, nfft 6144) of the 2000x3001 f32 panel",
            "value": round(ROWS * COLS / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5), "nfft": hp.nfft}


def spectrogram_leg(torch, P, dev, steps, warmup):
    """The spectrogram of one 8001-sample trace at 1 ms: 257-sample Hann
    frames every 10 samples (801 frames, nfft 512), one spectrogramInto per
    step into rotating outputs."""
    sg = P.initSpectrogramPlan(np.float32, 257, 10, 1.0, 8001)
    x = torch.from_numpy(np.random.default_rng(5).standard_normal(8001).astype(np.float32)).to(dev)
    outs = [torch.empty((sg.frames, sg.bins), dtype=torch.float64, device=dev) for _ in range(4)]
    for i in range(warmup):
        sg.spectrogramInto(x, outs[i % 4])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        sg.spectrogramInto(x, outs[i % 4])
    torch.cuda.synchronize()
    t = (time.perf_counter() - t1) / steps
    return {"workload": "spectrogram of one 8001-sample f32 trace (257-sample Hann frames every "
                        "10 samples, 801 frames, nfft 512)",
            "value": round(sg.frames / t / 1e6, 3), "unit": "Mframes/s",
            "ms_per_step": round(t * 1000, 5)}


def _timed(torch, fn, steps, warmup):
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        fn(i)
    torch.cuda.synchronize()
    return (time.perf_counter() - t1) / steps


def calibrated_flow_leg(torch, P, dev, steps, warmup):
    """calibrateOp (two sensor rules over file descales) > the headline flow
    on the 2000 x 3001 panel: Plan.apply with the meta, so each call
    resolves the rows' factors on the host (pseg_plan_calib_factors) and
    copies them with the row lengths; and the same apply with the factors
    resolved once (the calibration step of the row kernel alone)."""
    from processseg_amd import calibrate as K
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    sensor = np.where(np.arange(ROWS) % 4 == 0, 1, 2).astype(np.uint8)
    scales = [K.FileScale(1e-6 * (1 + r % 5), float("nan"), K.auRecorded) for r in range(ROWS)]
    meta = P.ChunkMeta(ROWS, DT, numSamples=live, sensorType=sensor, scales=scales)
    cal = K.initCalibration([
        K.CalibrationRule(K.RuleMatch(sensorType=1), sensitivity=K.sensitivity(20, K.suVPerBar)),
        K.CalibrationRule(K.RuleMatch(sensorType=2),
                          sensitivity=K.sensitivity(22.5, K.suVPerMPerS))])
    flow = make_flow()
    plan = P.initPipeline([P.calibrateOp(cal)] + flow.ops).compile(np.float32, DT, COLS)
    src = torch.from_numpy(x).to(dev)
    outs = [torch.empty_like(src) for _ in range(4)]
    t_meta = _timed(torch, lambda i: plan.apply(src, outs[i % 4], meta), steps, warmup)
    k = plan.calibFactors(meta)
    t_k = _timed(torch, lambda i: plan.apply(src, outs[i % 4], meta, calib_factors=k), steps,
                 warmup)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.calibFactors(meta)
    t_res = (time.perf_counter() - t0) / steps
    return {"workload": "calibrate (hydrophone / geophone rules over file descales) > "
                        + flow_name() + ", 2000x3001 f32",
            "value": round(ROWS * COLS / t_meta / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t_meta * 1000, 5),
            "ms_factors_given": round(t_k * 1000, 5),
            "ms_host_resolution": round(t_res * 1000, 5)}


def segd_decode_leg(torch, P, dev, steps, warmup):
    """SEG-D 8036 (24-bit two's-complement integers, big-endian) traces of
    the 2000 x 3001 panel's size (no headers) to float32, 8 rotating buffer
    pairs.  Algorithmic bytes: 3 read + 4 written per sample."""
    rows, ns = ROWS, COLS
    v = np.random.default_rng(3).integers(-(1 << 23), 1 << 23, (rows, ns)).astype(np.int64)
    u = (v & 0xFFFFFF).astype(np.uint32)
    blob = np.stack([(u >> 16) & 0xFF, (u >> 8) & 0xFF, u & 0xFF], -1).astype(np.uint8)
    raws = [torch.from_numpy(blob.reshape(-1).copy()).to(dev) for _ in range(8)]
    outs = [torch.empty(rows, ns, device=dev) for _ in range(8)]
    t = _timed(torch, lambda i: P.decodeSamples(raws[i % 8], P.fmtSegd8036, ns, rows=rows,
                                                offset=0, out=outs[i % 8]), steps, warmup)
    algo = rows * ns * 7
    return {"workload": "SEG-D 8036 24-bit integer traces, 2000 x 3001, to float32 on the device",
            "value": round(rows * ns / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5),
            "achieved_GBps": round(algo / t / 1e9, 1)}


def analytic_leg(torch, P, dev, steps, warmup):
    """analyticInto of every live row of the 2000 x 3001 panel (nfft 6144):
    envelope, instantaneous phase and frequency as float64 (24 B written
    per sample)."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live)
    hp = P.initHilbertPlan(np.float32, COLS)
    src = torch.from_numpy(x).to(dev)
    rl = meta.device_row_len(COLS, dev)
    outs = [tuple(torch.empty(ROWS, COLS, dtype=torch.float64, device=dev) for _ in range(3))
            for _ in range(2)]
    t = _timed(torch, lambda i: hp.analyticInto(src, DT, *outs[i % 2], row_len=rl), steps, warmup)
    return {"workload": "analytic signal (envelope, phase, frequency; nfft 6144) of the "
                        "2000x3001 f32 panel",
            "value": round(ROWS * COLS / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5)}


def spectrogram_average_leg(torch, P, dev, steps, warmup):
    """averageSpectrogramInto over the 2000 x 3001 panel: 257-sample Hann
    frames every 10 samples (301 frames per row, nfft 512), one launch over
    rows x frames, partials summed in a fixed order."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live)
    sg = P.initSpectrogramPlan(np.float32, 257, 10, DT, COLS)
    src = torch.from_numpy(x).to(dev)
    rl = meta.device_row_len(COLS, dev)
    amp = torch.empty((sg.frames, sg.bins), dtype=torch.float64, device=dev)
    ws = torch.empty(max(sg.averageWorkspace(ROWS), 1), dtype=torch.uint8, device=dev)
    t = _timed(torch, lambda i: sg.averageSpectrogramInto(src, amp=amp, row_len=rl, ws=ws,
                                                          sync=False), steps, warmup)
    return {"workload": "average spectrogram of the 2000x3001 f32 panel (257-sample Hann frames "
                        "every 10 samples, nfft 512)",
            "value": round(ROWS * sg.frames / t / 1e6, 2), "unit": "Mframes/s",
            "ms_per_step": round(t * 1000, 5), "frames_per_row": sg.frames}


def fk_leg(torch, P, dev, steps, warmup):
    """The f-k amplitude of the 2000 x 3001 panel (10 % Tukey tapers, float32
    transforms, float64 output): a time pass per trace and a spatial pass per
    frequency."""
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live)
    traces = P.fkRows(meta, COLS, 0, COLS)
    plan = P.initFkPlan(np.float32, DT, traces, COLS, 0.1)
    src = torch.from_numpy(x).to(dev)
    amp = torch.empty(plan.nK * plan.nFreq, dtype=torch.float64, device=dev)
    t = _timed(torch, lambda i: plan.fkAmplitudeInto(src, meta, 0, amp), steps, warmup)
    return {"workload": f"f-k amplitude of the 2000x3001 f32 panel ({plan.nK} wavenumbers x "
                        f"{plan.nFreq} frequencies)",
            "value": round(ROWS * COLS / t / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t * 1000, 5)}


def firstbreak_leg(torch, P, dev, steps, warmup):
    """First-break picks (energy ratio, the reference's defaults, a band-pass
    8-70 Hz) of every live row of the 2000 x 3001 panel, one thread per row
    in the reference's serial float64 order."""
    from processseg_amd.firstbreak import FirstBreakParams
    from processseg_amd.workloads import big_panel
    x, live = big_panel()
    meta = P.ChunkMeta(ROWS, DT, numSamples=live)
    gp = P.initFirstBreakPlan(np.float32, FirstBreakParams(lowCutHz=8.0, highCutHz=70.0), DT,
                              COLS)
    src = torch.from_numpy(x).to(dev)
    rl = meta.device_row_len(COLS, dev)
    t = _timed(torch, lambda i: gp.pickInto(src, row_len=rl), steps, warmup)
    return {"workload": "first-break picks (energy ratio, 8-70 Hz band-pass) of the 2000x3001 "
                        "f32 panel",
            "value": round(ROWS / t / 1e6, 4), "unit": "Mtraces/s",
            "ms_per_step": round(t * 1000, 5)}


def fold_leg(torch, P, dev, steps, warmup):
    """Fold counting of two million traces over a 500 x 500 grid at azimuth 37
    with 5 offset classes, 8 azimuth sectors and offset ranges (integer
    atomics), the map cleared between calls."""
    from processseg_amd.fold import FoldOptions, gridBinning
    n = 2_000_000
    rng = np.random.default_rng(11)
    b = gridBinning(600_000.0, 4_100_000.0, 37.0, 12.5, 25.0, inlines=500, crosslines=500)
    srcxy = rng.uniform(-2000, 14000, (n, 2)) + (600_000.0, 4_100_000.0)
    rcv = srcxy + rng.normal(0, 1500, (n, 2))
    keys = P.traceKeyTable(n, srcX=srcxy[:, 0], srcY=srcxy[:, 1], rcvX=rcv[:, 0], rcvY=rcv[:, 1],
                           cmpX=(srcxy[:, 0] + rcv[:, 0]) / 2, cmpY=(srcxy[:, 1] + rcv[:, 1]) / 2,
                           offset=rng.normal(0, 3000, n))
    opts = FoldOptions(offsetBounds=(0, 500, 1000, 2000, 4000, 8000.0), azimuthSectors=8,
                       sectorStartDeg=-10, offsetRange=True)
    m = P.initFoldMap(b, opts)
    kd = torch.from_numpy(np.ascontiguousarray(keys)).to(dev)
    rl = torch.zeros(n, dtype=torch.int32, device=dev)

    def step(i):
        m.clear()
        m.addFold(keys=kd, row_len=rl)
    t = _timed(torch, step, steps, warmup)
    return {"workload": "fold of 2M traces over a 500x500 grid (5 offset classes, 8 sectors, "
                        "offset ranges)",
            "value": round(n / t / 1e6, 1), "unit": "Mtraces/s",
            "ms_per_step": round(t * 1000, 5)}


def batch_leg(torch, dev, steps, warmup):
    """64 records of an uneven survey (rows U[240, 2400], samples U[1000,
    6001], dt 1, 2 or 4 ms; band-pass 5-10-70-90 Hz + AGC 500 ms, one plan per
    (dt, samples)) as one PipelineBatch: persistent launches replayed from a
    captured graph (launch shapes as the rule gives them, then as the
    autotuner picks them), against one Plan.apply launch per record."""
    import processseg_amd as P
    from processseg_amd import synth
    from processseg_amd.workloads import mixed_flow, mixed_shapes
    shapes = mixed_shapes(64)
    R = max(r for r, _, _ in shapes)
    Cm = max(c for _, c, _ in shapes)
    base, _ = synth.gather(R, Cm, 2.0, 9)   # records are cut from one seeded gather
    base = torch.from_numpy(base).to(dev)
    flow = mixed_flow()
    plans, srcs, dsts, lens = {}, [], [], []
    for r, c, dt in shapes:
        if (dt, c) not in plans:
            plans[(dt, c)] = flow.compile(np.float32, dt, c)
        srcs.append(base[:r, :c].contiguous())
        dsts.append(torch.empty_like(srcs[-1]))
        lens.append(torch.full((r,), c, dtype=torch.int32, device=dev))
    del base
    items = [P.BatchItem(plans[(dt, c)], s, d, row_len=l)
             for (r, c, dt), s, d, l in zip(shapes, srcs, dsts, lens)]
    batch = P.PipelineBatch(items)
    stream = torch.cuda.current_stream()

    def timed(fn, n):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    t_eager = timed(lambda: batch.launch(stream), steps)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batch.launch()
    t_rule = timed(g.replay, steps)
    grid_rule = list(batch.grid)
    # the autotuner: each launch class's shape chosen by timing (the same
    # bits), then the graph captured again
    tuning = batch.tune(stream=stream)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batch.launch()
    t_graph = timed(g.replay, steps)

    def per_record():
        for it in items:
            it.plan.apply(it.src, it.dst, row_len=it.row_len, stream=stream)
    t_loop = timed(per_record, steps)
    samples = batch.samples
    return {"workload": "64 records, rows U[240, 2400], samples U[1000, 6001], dt 1/2/4 ms, "
                        "band-pass 5-10-70-90 Hz + AGC 500 ms, f32",
            "records": 64, "samples": samples, "plans": len(plans),
            "launches": batch.launches, "threads_per_row": batch.threads,
            "grid": batch.grid, "grid_before_tuning": grid_rule,
            "tuning_candidate_ms": tuning["candidate_ms"],
            "value": round(samples / t_graph / 1e6, 1), "unit": "Msamples/s",
            "ms_per_step": round(t_graph * 1000, 4),
            "ms_before_tuning": round(t_rule * 1000, 4),
            "ms_eager": round(t_eager * 1000, 4),
            "ms_one_launch_per_record": round(t_loop * 1000, 4),
            "achieved_GBps": round(samples * 8 / t_graph / 1e9, 1)}


SERVE_COLD_S = 5.0                       # the cold-cache stream (plans compiled on use)
SERVE_SWEEP = (1000.0, 4000.0, 8000.0, 16000.0)   # offered loads above 200/s, req/s
SERVE_SWEEP_S = 1.5


def serve_leg(torch, dev, seconds, rate_hz):
    """Zoom windows of a resident 20000 x 4001 f32 survey (320 MB; a seeded
    2000 x 4001 gather tiled 10 times) under a Poisson stream of `rate_hz`
    requests/s (100-2000 traces x 200-4001 shown samples, three flows).  The
    continuously batched server against one launch per request in arrival
    order, each with: a short stream while plans compile per window width on
    first use (cold), the `seconds`-long stream with every plan cached (warm:
    precompiled plans), the same stream at higher offered loads (a load sweep,
    where the one-launch path queues), and the whole stream submitted at once
    (throughput)."""
    import gc
    import queue
    import threading
    import processseg_amd as P
    from processseg_amd import synth
    from processseg_amd.workloads import serve_flows, serve_requests
    tile, _ = synth.gather(2000, 4001, 2.0, 12)
    survey = torch.from_numpy(tile).to(dev).repeat(10, 1).contiguous()
    trace_len = torch.full((survey.shape[0],), survey.shape[1], dtype=torch.int32, device=dev)
    reqs = serve_requests(int(seconds * rate_hz), survey.shape[0], survey.shape[1], rate_hz)
    cold = reqs[:int(min(seconds, SERVE_COLD_S) * rate_hz)]
    flows = serve_flows()

    def pct(x, q):
        return round(float(np.percentile(x, q)), 3)

    def run(srv, ids, paced, reqs):
        q = queue.Queue()
        res = []

        def collect():
            while True:
                t = q.get()
                if t is None:
                    return
                r = srv.wait(t)
                res.append(r)
                srv.release(t)
        # the load generator is Python: no garbage-collector pauses and a
        # short GIL switch interval while it paces the stream, so requests
        # leave on time (a late submitter turns the Poisson stream into
        # bursts the server then has to absorb)
        gc.disable()
        switch = sys.getswitchinterval()
        sys.setswitchinterval(1e-4)
        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        for at, f, tr0, ntr, a, b in reqs:
            if paced:
                while time.perf_counter() - t0 < at:
                    time.sleep(min(0.0005, max(0.0, at - (time.perf_counter() - t0))))
            q.put(srv.submit(ids[f], tr0, ntr, a, b))
        t_sub = time.perf_counter() - t0
        q.put(None)
        th.join()
        sys.setswitchinterval(switch)
        gc.enable()
        wall = time.perf_counter() - t0
        lat = np.array([r.latency_ms for r in res])
        queued = np.array([(r.start_ns - r.submit_ns) / 1e6 for r in res])
        samples = sum(r.rows * r.window for r in res)
        return {"requests": len(res), "submit_s": round(t_sub, 3), "wall_s": round(wall, 3),
                "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                "mean_ms": round(float(lat.mean()), 3),
                "queued_p50_ms": pct(queued, 50), "queued_p99_ms": pct(queued, 99),
                "compiled_frac": round(float(np.mean([r.plan_compiled for r in res])), 3),
                "mean_batch": round(float(np.mean([r.batch_requests for r in res])), 2),
                "requests_per_s": round(len(res) / wall, 1),
                "Msamples_per_s": round(samples / wall / 1e6, 1)}

    out = {"workload": f"{len(reqs)} zoom-window requests, Poisson {rate_hz:g}/s for "
                       f"{seconds:g} s, survey {survey.shape[0]}x{survey.shape[1]} f32 resident, "
                       "3 flows",
           "stream_s": seconds, "cold_stream_s": min(seconds, SERVE_COLD_S),
           "unit": "ms (submit to result on the device; queued: submit to launch)"}
    for name, mb, inflight in (("batched", 256, 2), ("one_launch_per_request", 1, 1)):
        srv = P.Server(survey, trace_len, 2.0, arena_bytes=4 << 30, max_batch_requests=mb,
                       max_inflight=inflight)
        ids = [srv.add_flow(f) for f in flows]
        log(f"serving: {name}, cold plan cache ({len(cold)} requests)")
        out[name] = {"cold": run(srv, ids, True, cold)}
        # precompiled plans: one single-trace request per (flow, window) of the stream
        uniq = sorted({(f, a, b) for _, f, _, _, a, b in reqs})
        t0 = time.perf_counter()
        run(srv, ids, False, [(0.0, f, 0, 1, a, b) for f, a, b in uniq])
        out[name]["precompile"] = {"windows": len(uniq),
                                   "s": round(time.perf_counter() - t0, 2)}
        log(f"serving: {name}, warm plan cache ({len(reqs)} requests)")
        out[name]["warm"] = run(srv, ids, True, reqs)
        sweep = {}
        for hz in SERVE_SWEEP:
            # the warm stream's requests at a higher offered load
            sr = [(at * rate_hz / hz,) + tuple(rest)
                  for at, *rest in reqs[:int(SERVE_SWEEP_S * hz)]]
            log(f"serving: {name}, {hz:g} req/s offered ({len(sr)} requests)")
            sweep[f"{hz:g}"] = run(srv, ids, True, sr)
        out[name]["load_sweep"] = sweep
        log(f"serving: {name}, burst")
        out[name]["burst"] = run(srv, ids, False, reqs[:1600])
        out[name]["server"] = srv.info()
        srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-secondary", action="store_true",
                    help="only the headline panel (no correlation leg)")
    ap.add_argument("--legs", default="all",
                    help="comma-separated secondary legs to run (panels_as_batch, muted_flow, "
                         "gain_flow, iir_flow, decode, host_staged, raw_flow, spectrum, hilbert, spectrogram, "
                         "calibrated_flow, segd_decode, analytic, spectrogram_average, fk, firstbreak, fold, "
                         "f64_panel, correlation, uneven_batch, serving), "
                         "or all")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the serving leg's Poisson stream with precompiled plans "
                         "(the cold-cache stream is its first 5 s)")
    ap.add_argument("--flow-ops", type=int, default=None,
                    help="run only the first N ops of the headline flow (a per-step profile; "
                         "the line then names the cut flow)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="time a plain device copy of the panel instead (the calibration "
                         "kernel for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the panel path runs on a single MI355X "
              f"(rows shard across GPUs with no exchange); --gpus {args.gpus} "
              "is not run", file=sys.stderr)
        sys.exit(2)

    import torch
    import processseg_amd as P

    dev = torch.device("cuda:0")
    if args.copy_probe:
        copy_probe(torch, dev, args)
        return
    from processseg_amd.workloads import big_panel
    log("generating the 2000 x 3001 panel")
    x, live = big_panel(ROWS, COLS, np.float32)
    flow = make_flow(args.flow_ops)
    plan = flow.compile(np.float32, DT, COLS)
    row_len = torch.from_numpy(live.astype(np.int32)).to(dev)
    srcs = [torch.from_numpy(x).to(dev) for _ in range(args.pairs)]
    dsts = [torch.empty_like(srcs[0]) for _ in range(args.pairs)]
    stream = torch.cuda.current_stream()

    log(f"plan: {plan.steps}, lds {plan.lds_bytes} B")
    for i in range(args.warmup):
        plan.apply(srcs[i % args.pairs], dsts[i % args.pairs], row_len=row_len)
    torch.cuda.synchronize()

    # timed region: K steps back to back on the plan's stream, as a caller's
    # loop of Plan.apply runs them
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        k = i % args.pairs
        plan.apply(srcs[k], dsts[k], row_len=row_len, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the kernel's launch duration: the same K launches again, each between
    # two events on its stream (outside the timed region: the timestamps
    # between launches cost the stream a few microseconds per step)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    for i in range(args.steps):
        k = i % args.pairs
        starts[i].record(stream)
        plan.apply(srcs[k], dsts[k], row_len=row_len, stream=stream)
        ends[i].record(stream)
    torch.cuda.synchronize()
    per = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_ms = float(np.mean(per))
    ms_per_step = wall * 1000 / args.steps
    samples = ROWS * COLS
    value = samples / (ms_per_step / 1000) / 1e6

    algo_bytes = samples * 8  # 4 read + 4 written per sample
    achieved = algo_bytes / (avg_ms / 1000) / 1e9
    traffic = None
    valu = None
    tpath = os.path.join(ROOT, "profiles", "traffic_bandpass_agc_2000x3001_f32.json")
    if os.path.exists(tpath) and args.flow_ops is None:
        with open(tpath) as f:
            prof = json.load(f)
        traffic = prof.get("bytes_per_launch")
        if prof.get("valu_insts_per_wave") and prof.get("waves_per_launch"):
            # The kernel's second bound: vector-ALU issue.  A wave64 VALU
            # instruction occupies a 16-lane SIMD for 4 cycles; MI355X has
            # 256 CUs x 4 SIMDs at 2.4 GHz.  Instructions per launch come
            # from the committed PMC pass (they depend on the plan, not the
            # clock); the time is this run's.
            winst = prof["valu_insts_per_wave"] * prof["waves_per_launch"]
            rate = winst / (avg_ms / 1000) / 1e9
            peak = VALU_PEAK_GWINST
            valu = {"achieved": round(rate, 1), "peak": peak, "unit": "G wave-instructions/s",
                    "frac": round(rate / peak, 4),
                    "insts_per_launch": round(winst),
                    "insts_source": "SQ_INSTS_VALU of a rocprofv3 --pmc pass of this command, "
                                    "profiles/" + os.path.basename(tpath)}

    out = dsts[(args.steps - 1) % args.pairs].cpu().numpy()
    line = {
        "metric": "Msamples/s band-pass+AGC 2000x3001 f32 (%HBM roofline); p50/p99 served latency",
        "value": round(value, 1),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Ricker + hyperbolic events + Gaussian noise)",
        "config": {"workload": flow_name(args.flow_ops) + ", 2000x3001 f32, dt 2 ms, "
                               "live lengths U[2400, 3001]",
                   "rows": ROWS, "cols": COLS, "dt_ms": DT,
                   "nfft": [s["nfft"] for s in plan.steps if s["nfft"]],
                   "buffer_pairs": args.pairs},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": ("FETCH_SIZE and WRITE_SIZE of rocprofv3 --pmc passes of this "
                                        "command, scaled by a plain copy's: profiles/"
                                        + os.path.basename(tpath)) if traffic else None,
                     "kernel": kernel_name(COLS, plan.threadsPerRow),
                     "avg_launch_ms": round(avg_ms, 5),
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "valu_issue": valu},
    }
    if not args.no_cpu:
        log("CPU baseline leg")
        cpu, acc = cpu_baseline_leg(flow, x, live, out)
        line["cpu_baseline"] = cpu
        line["accuracy"] = acc
    if not args.no_secondary:
        legs = set(args.legs.split(",")) if args.legs != "all" else None

        def want(name):
            return legs is None or name in legs

        sec = line["secondary"] = {}
        if want("panels_as_batch"):
            log("panels as one batch")
            sec["panels_as_batch"] = panels_batch_leg(
                torch, P, plan, srcs, dsts, row_len, max(10, args.steps // 10), 3)
        del srcs, dsts
        torch.cuda.empty_cache()
        if want("muted_flow"):
            log("muted flow leg")
            sec["muted_flow"] = muted_flow_leg(torch, P, dev, max(10, args.steps // 4), 3)
            torch.cuda.empty_cache()
        if want("gain_flow"):
            log("gain curve flow leg")
            sec["gain_flow"] = gain_flow_leg(torch, P, dev, max(10, args.steps // 4), 3)
            torch.cuda.empty_cache()
        if want("iir_flow"):
            log("IIR flow leg")
            sec["iir_flow"] = iir_flow_leg(torch, P, dev, max(10, args.steps // 4), 3)
            torch.cuda.empty_cache()
        if want("decode"):
            log("decode leg")
            sec["decode"] = decode_leg(torch, P, dev, max(20, args.steps // 2), 5)
            torch.cuda.empty_cache()
        if want("host_staged"):
            log("host-staged leg")
            sec["host_staged"] = host_staged_leg(torch, P, dev, max(10, args.steps // 10), 3)
            torch.cuda.empty_cache()
        if want("raw_flow"):
            log("raw flow leg")
            sec["raw_flow"] = raw_flow_leg(torch, P, dev, max(10, args.steps // 4), 3)
            torch.cuda.empty_cache()
        if want("spectrum"):
            log("spectrum leg")
            sec["spectrum"] = spectrum_leg(torch, P, dev, max(10, args.steps // 4), 3)
            torch.cuda.empty_cache()
        if want("hilbert"):
            log("Hilbert leg")
            sec["hilbert"] = hilbert_leg(torch, P, dev, max(10, args.steps // 4), 3)
            torch.cuda.empty_cache()
        if want("spectrogram"):
            log("spectrogram leg")
            sec["spectrogram"] = spectrogram_leg(torch, P, dev, max(20, args.steps), 5)
        for name, fn, n_steps in (
                ("calibrated_flow", calibrated_flow_leg, max(10, args.steps // 4)),
                ("segd_decode", segd_decode_leg, max(20, args.steps // 2)),
                ("analytic", analytic_leg, max(10, args.steps // 10)),
                ("spectrogram_average", spectrogram_average_leg, max(10, args.steps // 10)),
                ("fk", fk_leg, max(10, args.steps // 10)),
                ("firstbreak", firstbreak_leg, max(5, args.steps // 20)),
                ("fold", fold_leg, max(5, args.steps // 20))):
            if want(name):
                log(f"{name} leg")
                sec[name] = fn(torch, P, dev, n_steps, 2)
                torch.cuda.empty_cache()
        if want("f64_panel"):
            log("float64 panel leg")
            sec["f64_panel"] = f64_panel_leg(torch, P, flow, dev, max(10, args.steps // 4), 3)
            torch.cuda.empty_cache()
        if want("correlation"):
            log("correlation leg")
            sec["correlation"] = correlation_leg(
                torch, dev, max(10, args.steps // 4), 3, 4, not args.no_cpu)
            torch.cuda.empty_cache()
        if want("uneven_batch"):
            log("uneven batch leg")
            sec["uneven_batch"] = batch_leg(torch, dev, max(5, args.steps // 20), 2)
            torch.cuda.empty_cache()
        if want("serving"):
            sv = serve_leg(torch, dev, args.serve_seconds, 200.0)
            sec["serving"] = sv
            line["p50_ms"] = sv["batched"]["warm"]["p50_ms"]
            line["p99_ms"] = sv["batched"]["warm"]["p99_ms"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
