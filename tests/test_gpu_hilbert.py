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


def _analytic_rows(n, dt, dtype):
    """Tones from 0.1 to 0.49 of fs (Gaussian-tapered), a sweep, two tones,
    a Nyquist alternation, noise, a zero row."""
    j = np.arange(n)
    fs = 1000 / dt
    u = (j - (n - 1) / 2) / (n / 12)
    g = np.exp(-u * u)
    t = j * dt / 1000
    rows = [g * np.cos(2 * np.pi * fr * fs * t + 0.4) for fr in (0.1, 0.24, 0.26, 0.4, 0.49)]
    rows.append(np.sin(2 * np.pi * (10 * t + 22.5 * t * t / 2)) * g)
    rows.append(np.cos(2 * np.pi * 5 * t) + 0.7 * np.cos(2 * np.pi * 10 * t))
    rows.append(np.where(j % 2 == 0, 1.0, -1.0))
    rows.append(np.random.default_rng(8).standard_normal(n))
    rows.append(np.zeros(n))
    return np.stack(rows).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("cols", [2001, 4001])
def test_analytic(gpu, oracle, dtype, cols):
    """analyticInto (fft/hilbert.nim:160-215) per row against the oracle's
    restatement on the same rows: the envelope to 1e-5 relative L2 (float32;
    1e-12 float64), the phase and the frequency where the envelope stands
    clear of zero and of the floor, the NaN masks equal away from the floor;
    a short live length, a dead row and a NaN row."""
    import torch
    dt = 4.0
    x = _analytic_rows(cols, dt, dtype)
    rows = x.shape[0] + 2
    x = np.concatenate([x, x[:1], x[:1]])
    x[-1, 100] = np.nan
    rl = np.full(rows, cols, np.int32)
    rl[3] = cols - 777
    rl[-2] = -1
    hp = P.initHilbertPlan(dtype, cols)
    src = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    env, ph, hz = hp.analyticInto(src, dt, row_len=torch.from_numpy(rl).cuda())
    env, ph, hz = env.cpu().numpy(), ph.cpu().numpy(), hz.cpu().numpy()
    tol = 1e-5 if dtype == np.float32 else 1e-12
    for r in range(rows):
        n = rl[r]
        if n < 0:
            assert np.all(np.isnan(env[r])) and np.all(np.isnan(hz[r]))
            continue
        we, wp, wh = oracle.analytic(x[r, :n].astype(np.float64), cols, dt)
        assert np.all(np.isnan(env[r, n:]))
        if np.isnan(x[r, :n]).any():
            assert np.all(np.isnan(env[r, :n])) and np.all(np.isnan(hz[r, :n]))
            continue
        if not np.any(we > 0):
            assert np.all(env[r, :n] == 0) and np.all(np.isnan(ph[r, :n])) and np.all(np.isnan(hz[r, :n]))
            continue
        assert rel_l2(env[r, :n], we) <= tol, (r, rel_l2(env[r, :n], we))
        peak = we.max()
        strong = we > 1e-2 * peak
        dp = (ph[r, :n] - wp + 180) % 360 - 180
        assert np.max(np.abs(dp[strong])) <= (2e-3 if dtype == np.float32 else 1e-9), r
        thr = peak * 1e-3
        clear = np.ones(n, bool)
        near = np.abs(we - thr) < 1e-3 * thr + 1e-6 * peak
        clear[1:-1] = ~(near[:-2] | near[1:-1] | near[2:])
        assert np.array_equal(np.isnan(hz[r, :n])[clear], np.isnan(wh)[clear]), r
        both = clear & ~np.isnan(wh) & (np.minimum(np.minimum(np.roll(we, 1), we), np.roll(we, -1)) > 1e-2 * peak)
        fs = 1000 / dt
        if r == 7:
            # exactly Nyquist: H[x] is 0 to rounding, so each phase step sits
            # on the wrap at ±pi and rounding decides the reading (the
            # reference pins this case in float64 only)
            continue
        dh = np.abs(hz[r, :n] - wh)[both]
        assert dh.size == 0 or np.max(dh) <= (1e-3 * fs if dtype == np.float32 else 1e-8 * fs), r


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("cols,nfft", [(500, 1125), (1000, 2025), (61, 135)])
def test_hilbert_odd_nfft(gpu, oracle, dtype, cols, nfft):
    """An odd 2·3·5-smooth nfft (fft/hilbert.nim:44-53 takes any smooth
    length >= 2·samples): the transform, the envelope and the analytic
    signal against the oracle over each row's live samples."""
    import torch
    rows = 6
    x, _ = synth.gather(rows, cols, 2.0, nfft, dtype=dtype)
    rl = np.full(rows, cols, np.int32)
    rl[2] = cols // 2
    rl[4] = -1
    hp = P.initHilbertPlan(dtype, cols, nfft)
    assert hp.nfft == nfft
    src = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    rlt = torch.from_numpy(rl).cuda()
    h = hp.hilbertInto(src, row_len=rlt).cpu().numpy()
    env, ph, hz = hp.analyticInto(src, 2.0, row_len=rlt)
    env = env.cpu().numpy()
    tol = 1e-5 if dtype == np.float32 else 1e-12
    for r in range(rows):
        n = int(rl[r])
        if n <= 0:
            assert np.array_equal(h[r], x[r])
            continue
        want = oracle.hilbert(x[r, :n].astype(np.float64), cols, nfft)
        assert rel_l2(h[r, :n], want) <= tol, (r, rel_l2(h[r, :n], want))
        we, _, _ = oracle.analytic(x[r, :n].astype(np.float64), cols, 2.0, nfft=nfft)
        assert rel_l2(env[r, :n], we) <= tol
