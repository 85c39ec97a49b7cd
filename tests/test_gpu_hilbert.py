"""The Hilbert transform and envelope on the GPU (fft/hilbert.nim) against the
oracle: float32 within 1e-5 relative L2 of the float64 oracle, float64
within 1e-12, over each row's live samples; rows not live and samples past a
live end unchanged; in place and not; the three threads-per-row classes; a float64 transform beyond one workgroup's
LDS refused."""
import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import synth
from tests.common import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("cols", [300, 3001, 12001])
@pytest.mark.parametrize("envelope", [False, True])
def test_hilbert(gpu, oracle, dtype, cols, envelope):
    import torch
    rows = 12
    x, _ = synth.gather(rows, cols, 2.0, cols, dtype=dtype)
    rl = synth.live_lengths(cols, rows, 1, cols).astype(np.int32)
    rl[0] = cols
    rl[3] = -1
    rl[4] = 0
    if dtype == np.float64 and cols == 12001:
        # nfft 24300 in float64 is 194 KB: beyond one workgroup's LDS
        with pytest.raises(P.UnsupportedOp):
            P.initHilbertPlan(dtype, cols)
        return
    hp = P.initHilbertPlan(dtype, cols)
    src = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    fn = hp.envelopeInto if envelope else hp.hilbertInto
    got = fn(src, row_len=torch.from_numpy(rl).cuda()).cpu().numpy()
    again = src.clone()
    fn(again, again, row_len=torch.from_numpy(rl).cuda())
    assert np.array_equal(again.cpu().numpy(), got), "in place differs"
    tol = 1e-5 if dtype == np.float32 else 1e-12
    for r in range(rows):
        n = int(rl[r])
        if n <= 0:
            assert np.array_equal(got[r], x[r])
            continue
        assert np.array_equal(got[r, n:], x[r, n:])
        want = oracle.hilbert(x[r, :n].astype(np.float64), cols, envelope=envelope)
        assert rel_l2(got[r, :n], want) <= tol, (r, rel_l2(got[r, :n], want))
