"""The serving loop on the GPU: zoom windows of a device-resident survey.

A served request must equal the reference's zoomWindow on the same survey
(tests/utils/viewersim.nim:167-191): decode the window widened by the flow's
requiredMargin (decodeWindow), apply the plan compiled for that window width
(planFor, viewersim.nim:53-62), keep the shown samples.  That is computed here
with PipelinePlan.apply on a copy of the window (bit for bit) and with the
float64 oracle (within the float32 tolerance).  A request's bits must not
depend on the requests it shares a launch with (served alone == served in a
mix)."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import synth
from tests.common import rel_l2

pytestmark = pytest.mark.gpu

TRACES, SAMPLES, DT = 600, 4001, 2.0


def flows():
    return [
        P.initPipeline([P.bandPassOp(P.Corners(4, 8, 60, 80)), P.agcOp(500)]),
        P.initPipeline([P.bandPassOp(P.Corners(4, 8, 60, 80)),
                        P.normalizeOp(P.normPercentile)]),
        P.initPipeline([P.demeanOp(),
                        P.filterOp(P.bandPass(P.Corners(2, 4, 100, 120)).withNotch(50, 4)),
                        P.agcOp(250)]),
    ]


@pytest.fixture(scope="module")
def survey():
    import torch
    x, _ = synth.gather(TRACES, SAMPLES, DT, 21)
    tl = synth.live_lengths(21, TRACES, 3000, SAMPLES).astype(np.int32)
    tl[17] = -1                      # a dead trace
    tl[40] = 0
    for r in range(TRACES):
        if tl[r] >= 0:
            x[r, tl[r]:] = 0
    return x, tl, torch.from_numpy(x).cuda(), torch.from_numpy(tl).cuda()


def expected(flow, x, tl, trace0, ntraces, shown_first, shown_last, oracle=None):
    """zoomWindow with the Python mirror of the plan: returns (bits, oracle)."""
    import torch
    m = flow.requiredMargin(DT)
    w = P.decodeWindow(range(shown_first, shown_last + 1), m, SAMPLES)
    first, cols = w.start, len(w)
    win = np.ascontiguousarray(x[trace0:trace0 + ntraces, first:first + cols])
    tls = tl[trace0:trace0 + ntraces]
    row_len = np.where(tls < 0, -1, np.clip(tls - first, 0, cols)).astype(np.int32)
    plan = flow.compile(np.float32, DT, cols, recordSamples=SAMPLES)
    src = torch.from_numpy(win).cuda()
    dst = torch.empty_like(src)
    plan.apply(src, dst, row_len=torch.from_numpy(row_len).cuda())
    got = dst.cpu().numpy()[:, shown_first - first:shown_last - first + 1]
    want = None
    if oracle is not None:
        want = oracle.pipeline_apply(flow.ops, win.astype(np.float64), row_len, DT)
        want = want[:, shown_first - first:shown_last - first + 1]
    return got, want


def requests(n, seed):
    u = synth.uniforms(seed, (5, n), stream=41)
    out = []
    for i in range(n):
        ntr = 1 + int(u[0, i] * 120)
        t0 = int(u[1, i] * (TRACES - ntr))
        width = 50 + int(u[2, i] * 3000)
        s0 = int(u[3, i] * (SAMPLES - width))
        out.append((int(u[4, i] * 3), t0, ntr, s0, s0 + width - 1))
    return out


def test_served_equals_zoom_window(gpu, oracle, survey):
    x, tl, xs, tls = survey
    F = flows()
    srv = P.Server(xs, tls, DT, arena_bytes=256 << 20)
    ids = [srv.add_flow(f) for f in F]
    reqs = requests(24, 5) + [(0, 10, 20, 0, 99), (2, 590, 10, 3900, 4000), (1, 15, 5, 0, 4000)]
    tickets = [srv.submit(ids[f], t0, n, a, b) for f, t0, n, a, b in reqs]
    worst = 0.0
    for k, ((f, t0, n, a, b), t) in enumerate(zip(reqs, tickets)):
        res = srv.wait(t)
        got = srv.read(t).cpu().numpy()
        srv.release(t)
        assert (res.rows, res.cols) == (n, b - a + 1)
        bits, want = expected(F[f], x, tl, t0, n, a, b, oracle if k % 4 == 0 else None)
        assert np.array_equal(got, bits), f"request {k} {reqs[k]}"
        if want is not None:
            e = rel_l2(got, want)
            worst = max(worst, e)
            assert e <= 1e-5, f"request {k}: rel L2 {e}"
    info = srv.info()
    print("served", info, "worst rel L2 vs FP64 oracle", worst)
    assert info["batches"] < len(reqs)   # requests shared launches
    srv.close()


def test_alone_equals_in_mix(gpu, survey):
    x, tl, xs, tls = survey
    F = flows()
    probe = [(0, 100, 64, 1000, 2999), (1, 7, 33, 0, 1200), (2, 300, 100, 2500, 4000)]
    alone = P.Server(xs, tls, DT, max_batch_requests=1, max_inflight=1)
    ids = [alone.add_flow(f) for f in F]
    ref = []
    for f, t0, n, a, b in probe:
        t = alone.submit(ids[f], t0, n, a, b)
        ref.append(alone.read(t).cpu().numpy())
        assert alone.wait(t).batch_requests == 1
        alone.release(t)
    alone.close()
    mix = P.Server(xs, tls, DT)
    ids = [mix.add_flow(f) for f in F]
    others = requests(40, 9)
    tickets = []
    for i, (f, t0, n, a, b) in enumerate(others):
        tickets.append(("other", mix.submit(ids[f], t0, n, a, b)))
        if i % 13 == 0 and i // 13 < len(probe):
            f2, t2, n2, a2, b2 = probe[i // 13]
            tickets.append((i // 13, mix.submit(ids[f2], t2, n2, a2, b2)))
    shared = 0
    for tag, t in tickets:
        res = mix.wait(t)
        if tag != "other":
            got = mix.read(t).cpu().numpy()
            assert np.array_equal(got, ref[tag]), f"probe {tag} differs in the mix"
            shared = max(shared, res.batch_requests)
        mix.release(t)
    mix.close()
    assert shared > 1, "the probes never shared a launch"


def test_admission_waits_for_arena_room(gpu, survey):
    """An arena that holds one result: the second request is admitted only
    after the first result is released."""
    _, _, xs, tls = survey
    rows, cols = 50, 1000
    srv = P.Server(xs, tls, DT, arena_bytes=(rows * cols * 4 + 255) // 256 * 256)
    fid = srv.add_flow(flows()[0])
    a = srv.submit(fid, 0, rows, 0, cols - 1)
    b = srv.submit(fid, 100, rows, 0, cols - 1)
    srv.wait(a)
    with pytest.raises(TimeoutError):
        srv.wait(b, timeout_ms=300)     # PSEG_E_BUSY: not admitted yet
    srv.release(a)
    res = srv.wait(b, timeout_ms=10000)
    assert res.rows == rows
    srv.release(b)
    with pytest.raises(AssertionError):   # PSEG_E_WORKSPACE: can never fit
        srv.submit(fid, 0, rows + 1, 0, cols - 1)
    srv.close()


def test_server_argument_errors(gpu, survey):
    _, _, xs, tls = survey
    srv = P.Server(xs, tls, DT)
    fid = srv.add_flow(flows()[0])
    with pytest.raises(ValueError, match="outside the survey"):
        srv.submit(fid, TRACES - 5, 10, 0, 100)
    with pytest.raises(ValueError, match="not a window"):
        srv.submit(fid, 0, 10, 200, 100)
    with pytest.raises(ValueError, match="no flow"):
        srv.submit(7, 0, 10, 0, 100)
    with pytest.raises(P.UnsupportedOp, match="per-panel"):
        srv.add_flow(P.initPipeline([P.normalizeOp(P.normRms, scope=P.perPanel)]))
    srv.close()


def test_served_flows_read_trace_keys(gpu, survey):
    """A flow with a top LMO mute and a t² gain curve, served with the
    survey's per-trace keys (t0Ms, offset), equals Plan.apply on the same
    decoded window with those keys and the window's firstSample, bit for
    bit."""
    import torch
    x, tl, xs, tls = survey
    keys = np.stack([np.where(np.arange(TRACES) % 3 == 0, 20.0, 0.0),
                     np.linspace(-3000, 3000, TRACES)], axis=1)
    kd = torch.from_numpy(keys).cuda()
    flow = P.initPipeline([P.bandPassOp(P.Corners(4, 8, 60, 80)),
                           P.muteTopOp(P.lmoMute(2500.0), taperMs=30),
                           P.gainCurveOp(P.powerGain(2.0)), P.agcOp(400)])
    srv = P.Server(xs, tls, DT, arena_bytes=128 << 20, trace_keys=kd)
    fid = srv.add_flow(flow)
    for t0, n, a, b in ((5, 40, 0, 999), (300, 64, 1500, 3200), (590, 10, 3000, 4000)):
        t = srv.submit(fid, t0, n, a, b)
        srv.wait(t)
        got = srv.read(t).cpu().numpy()
        srv.release(t)
        m = flow.requiredMargin(DT)
        w = P.decodeWindow(range(a, b + 1), m, SAMPLES)
        first, cols = w.start, len(w)
        win = torch.from_numpy(np.ascontiguousarray(x[t0:t0 + n, first:first + cols])).cuda()
        tls_w = tl[t0:t0 + n]
        row_len = np.where(tls_w < 0, -1, np.clip(tls_w - first, 0, cols)).astype(np.int32)
        plan = flow.compile(np.float32, DT, cols, recordSamples=SAMPLES)
        dst = torch.empty_like(win)
        plan.apply(win, dst, row_len=torch.from_numpy(row_len).cuda(),
                   keys=kd[t0:t0 + n].contiguous(), first_sample=first)
        want = dst.cpu().numpy()[:, a - first:b - first + 1]
        assert np.array_equal(got, want), (t0, n, a, b)


def test_every_launch_class_in_one_batch(gpu):
    """Nine requests that need the nine launch classes (windows of 128, 256
    and 512 threads per row x the core, mute and IIR kernel forms), queued
    behind a long request so they share one batch: each equals Plan.apply on
    its decoded window, bit for bit."""
    import torch
    traces, samples = 2000, 7000
    x, _ = synth.gather(traces, samples, DT, 33)
    tl = np.full(traces, samples, np.int32)
    xs, tls = torch.from_numpy(x).cuda(), torch.from_numpy(tl).cuda()
    F = [P.initPipeline([P.bandPassOp(P.Corners(4, 8, 60, 80)), P.agcOp(300)]),
         P.initPipeline([P.muteTopOp(P.timeMute(100.0), taperMs=20), P.agcOp(300)]),
         P.initPipeline([P.iirOp(P.butterBandPass(5, 70, 24)), P.agcOp(300)])]
    srv = P.Server(xs, tls, DT, arena_bytes=1 << 30, max_inflight=1)
    ids = [srv.add_flow(f) for f in F]
    blocker = srv.submit(ids[2], 0, traces, 0, samples - 1)
    reqs = [(f, 10 + 30 * i, 25, a, a + width - 1)
            for i, (f, (a, width)) in enumerate(
                (f, w) for f in range(3) for w in ((100, 1500), (500, 4500), (0, 6900)))]
    tickets = [srv.submit(ids[f], t0, n, a, b) for f, t0, n, a, b in reqs]
    srv.wait(blocker)
    srv.release(blocker)
    shared = 0
    for (f, t0, n, a, b), t in zip(reqs, tickets):
        res = srv.wait(t)
        shared = max(shared, res.batch_requests)
        got = srv.read(t).cpu().numpy()
        srv.release(t)
        m = F[f].requiredMargin(DT)
        w = P.decodeWindow(range(a, b + 1), m, samples)
        first, cols = w.start, len(w)
        win = torch.from_numpy(np.ascontiguousarray(x[t0:t0 + n, first:first + cols])).cuda()
        plan = F[f].compile(np.float32, DT, cols, recordSamples=samples)
        dst = torch.empty_like(win)
        # the window's column 0 is trace sample `first` (the mute's time axis)
        plan.apply(win, dst, row_len=torch.full((n,), cols, dtype=torch.int32, device="cuda"),
                   first_sample=first)
        want = dst.cpu().numpy()[:, a - first:b - first + 1]
        assert np.array_equal(got, want), (f, a, b)
    print("largest batch", shared, srv.info())
    srv.close()
