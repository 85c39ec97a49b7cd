"""The named workloads: the flows and seeded panels the benchmarks and the
parity tests run (48 x 2001 plain plan; 2000 x 3001 band-pass + AGC;
1200 x 10001 vibroseis records correlated with a 12 s sweep)."""
import numpy as np

from . import synth
from .correlate import generatePilot, initSweepTaper, linearSweep
from .pipeline import (Corners, agcOp, bandPass, bandPassOp, butterHighCut, demeanOp, filterOp,
                       initPipeline)

# The big panel's filter: a cos² 5-10-70-90 Hz band-pass with a 50 Hz notch
# 4 Hz wide (-6 dB at 50 +- 1 Hz), then a Butterworth high cut at 80 Hz,
# 24 dB/octave, as a second op (a FilterSpec is a trapezoid or a Butterworth,
# not both: filter/design.nim:85-90), then AGC 250 ms (h = 63).
BIG_NOTCH = (50.0, 4.0)
BIG_HIGHCUT = (80.0, 24.0)


def plain_flow():
    """48 x 2001 at 2 ms: demean, band-pass 4-8-60-80 Hz, AGC 500 ms."""
    return initPipeline([demeanOp(), bandPassOp(Corners(4, 8, 60, 80)), agcOp(500)])


def big_flow():
    """demean, band-pass 5-10-70-90 + notch, Butterworth high cut, AGC 250 ms."""
    return initPipeline([
        demeanOp(),
        filterOp(bandPass(Corners(5, 10, 70, 90)).withNotch(*BIG_NOTCH)),
        filterOp(butterHighCut(*BIG_HIGHCUT)),
        agcOp(250),
    ])


def plain_panel(dtype=np.float32, seed=1):
    """48 x 2001, offsets 25 i m, every row live to its end."""
    return synth.gather(48, 2001, 2.0, seed, dtype=dtype)


def big_panel(rows=2000, cols=3001, dtype=np.float32, seed=2):
    """2000 x 3001, live lengths uniform in [2400, 3001], zeros past them."""
    live = synth.live_lengths(seed, rows, 2400, cols)
    return synth.gather(rows, cols, 2.0, seed, live=live, dtype=dtype)


def sweep_pilot(dt_ms=2.0):
    """Linear sweep 6 -> 80 Hz, 12 s, cos² tapers of 300 ms at both ends
    (6000 samples at 2 ms)."""
    return generatePilot(linearSweep(6, 80, 12000, dt_ms, initSweepTaper(300, 300)))


def vibroseis_panel(rows=1200, cols=10001, dtype=np.float32, seed=4):
    """Uncorrelated records: the sweep convolved with sparse reflectivity
    (Bernoulli 0.01, N(0, 1)) plus N(0, 0.1), `cols` samples at 2 ms."""
    x, _ = synth.vibroseis_records(rows, cols, sweep_pilot().samples, seed, dtype=dtype)
    return x


def mixed_flow():
    """Band-pass 5-10-70-90 Hz and AGC 500 ms: every corner is under the
    Nyquist frequency of 1, 2 and 4 ms sampling."""
    return initPipeline([bandPassOp(Corners(5, 10, 70, 90)), agcOp(500)])


def mixed_shapes(n=64, rows=(240, 2400), cols=(1000, 6001), dts=(1.0, 2.0, 4.0), seed=3):
    """Record shapes of an uneven batch: rows and cols uniform in the given
    inclusive ranges, the sample interval uniform over `dts`.  Returns a list
    of (rows, cols, dt_ms)."""
    u = synth.uniforms(seed, (3, n), stream=31)
    out = []
    for i in range(n):
        r = rows[0] + int(u[0, i] * (rows[1] - rows[0] + 1))
        c = cols[0] + int(u[1, i] * (cols[1] - cols[0] + 1))
        out.append((min(r, rows[1]), min(c, cols[1]), dts[min(int(u[2, i] * len(dts)),
                                                              len(dts) - 1)]))
    return out


def serve_flows():
    """The three flows a viewer serves: band-pass + AGC; band-pass + 99th
    percentile normalization per trace; demean + a band-pass with a 50 Hz
    notch + AGC."""
    from .pipeline import normalizeOp, normPercentile
    return [
        initPipeline([bandPassOp(Corners(4, 8, 60, 80)), agcOp(500)]),
        initPipeline([bandPassOp(Corners(4, 8, 60, 80)), normalizeOp(normPercentile)]),
        initPipeline([demeanOp(), filterOp(bandPass(Corners(2, 4, 100, 120)).withNotch(50, 4)),
                      agcOp(250)]),
    ]


def serve_requests(n, traces=20000, samples=4001, rate_hz=200.0, seed=5):
    """A Poisson stream of zoom-window requests: arrival times (s) at
    `rate_hz`, flow uniform over the three, 100-2000 traces, a shown window
    of 200-4001 samples, both placed uniformly.  Returns a list of
    (arrival_s, flow, trace0, ntraces, shown_first, shown_last)."""
    u = synth.uniforms(seed, (6, n), stream=51)
    gaps = -np.log1p(-np.minimum(u[0], 1 - 1e-12)) / rate_hz
    t = np.cumsum(gaps)
    out = []
    for i in range(n):
        ntr = 100 + int(u[1, i] * 1901)
        width = 200 + int(u[2, i] * (samples - 200 + 1))
        ntr, width = min(ntr, 2000), min(width, samples)
        t0 = int(u[3, i] * (traces - ntr + 1))
        s0 = int(u[4, i] * (samples - width + 1))
        out.append((float(t[i]), min(int(u[5, i] * 3), 2), min(t0, traces - ntr), ntr,
                    min(s0, samples - width), min(s0, samples - width) + width - 1))
    return out
