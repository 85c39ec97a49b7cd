"""The named workloads: the flows and seeded panels the benchmarks and the
parity tests run (48 x 2001 plain plan; 2000 x 3001 band-pass + AGC)."""
import numpy as np

from . import synth
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
