"""The Hilbert transform and the envelope of traces on the GPU (fft/hilbert.nim),
through the same real FFT as the filters: each row's live samples
zero-padded to nfft >= 2·samples, -i·sign(f) on the spectrum (0 Hz and
Nyquist set to zero), inverse.

```python
hp = initHilbertPlan(np.float32, samples=4001)
h = hp.hilbertInto(panel, row_len=rl)        # H[cos] = sin
env = hp.envelopeInto(panel, row_len=rl)     # |x + i·H[x]|
```
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, _check_panel, _rows_of, _stream_handle


class HilbertPlan:
    """initHilbertPlan (fft/hilbert.nim:44-53)."""

    def __init__(self, T, samples: int, nfft: int = 0):
        self.T = np.dtype(T).type
        self.samples = int(samples)
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_hilbert_compile(0 if self.T == np.float32 else 1,
                                                    self.samples, int(nfft), C.byref(self._h)))
        n = C.c_int32()
        _lib.check(_lib.lib().pseg_hilbert_nfft(self._h, C.byref(n)))
        self.nfft = int(n.value)

    def _run(self, samples, dst, meta, row_len, envelope, stream):
        _check_panel(samples, "samples", self.T)
        rows, cols = samples.shape
        if cols > self.samples:
            raise ValueError(f"the rows have {cols} samples, the plan takes at most {self.samples}")
        row_len = _rows_of(meta, row_len, rows, cols, samples.device, None,
                           "hilbertInto / envelopeInto")
        if dst is None:
            dst = samples.clone()
        else:
            _check_panel(dst, "dst", self.T, rows=rows, cols=cols, device=samples.device)
        _lib.check(_lib.lib().pseg_hilbert_apply(
            self._h, samples.data_ptr(), dst.data_ptr(), rows, cols, row_len.data_ptr(),
            1 if envelope else 0, _stream_handle(stream)))
        return dst

    def hilbertInto(self, samples, dst=None, meta: ChunkMeta | None = None, *, row_len=None,
                    stream=None):
        """hilbertInto (fft/hilbert.nim:127-132) of every live row; dst (default:
        a copy of samples) keeps its values past the live ends."""
        return self._run(samples, dst, meta, row_len, False, stream)

    def envelopeInto(self, samples, dst=None, meta: ChunkMeta | None = None, *, row_len=None,
                     stream=None):
        """envelopeInto (fft/hilbert.nim:134-140): sqrt(x² + H[x]²)."""
        return self._run(samples, dst, meta, row_len, True, stream)

    def analyticInto(self, samples, dtMs: float, env=None, phaseDeg=None, hz=None,
                     meta: ChunkMeta | None = None, *, row_len=None, floorDb: float = -60.0,
                     stream=None):
        """analyticInto (fft/hilbert.nim:160-215) of every live row over its
        live length: (env, phaseDeg, hz), float64 [rows, cols] GPU tensors
        (created as NaN when not given; rows not live and columns past the
        live end keep what they hold).  phaseDeg in (-180, 180], NaN where
        env is 0; hz the circular mean of the two one-sample phase steps
        about a sample, NaN at the ends and below peak·10^(floorDb/20)."""
        import torch
        _check_panel(samples, "samples", self.T)
        rows, cols = samples.shape
        if cols > self.samples:
            raise ValueError(f"the rows have {cols} samples, the plan takes at most {self.samples}")
        if not (math.isfinite(dtMs) and dtMs > 0):
            raise ValueError(f"dtMs must be finite and > 0, got {dtMs}")
        if math.isnan(floorDb):
            raise ValueError("floorDb must not be NaN")
        row_len = _rows_of(meta, row_len, rows, cols, samples.device, None, "analyticInto")
        outs = []
        for t, name in ((env, "env"), (phaseDeg, "phaseDeg"), (hz, "hz")):
            if t is None:
                t = torch.full((rows, cols), float("nan"), dtype=torch.float64,
                               device=samples.device)
            else:
                _check_panel(t, name, np.float64, rows=rows, cols=cols, device=samples.device)
            outs.append(t)
        _lib.check(_lib.lib().pseg_hilbert_analytic(
            self._h, samples.data_ptr(), rows, cols, row_len.data_ptr(), float(dtMs),
            float(floorDb), outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr(),
            _stream_handle(stream)))
        return tuple(outs)

    def __del__(self):
        try:
            if self._h:
                _lib.lib().pseg_hilbert_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def initHilbertPlan(T, samples: int, nfft: int = 0) -> HilbertPlan:
    return HilbertPlan(T, samples, nfft)
