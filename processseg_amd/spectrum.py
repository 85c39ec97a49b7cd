"""Amplitude spectra for a viewer's spectrum panel on the GPU (fft/spectrum.nim):
the RMS amplitude spectrum of the windows of a block's live rows, through the
same real FFT as the filters.

```python
sp = initSpectrumPlan(np.float32, dtMs=2.0, windowLen=501)    # taper 10 % each end
amp, used = sp.averageAmplitudeInto(samples, meta, firstCol=250)  # float64 [bins] on the GPU
```
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, _stream_handle, sameInterval


class SpectrumPlan:
    """initSpectrumPlan (fft/spectrum.nim:77-101).  ValueError as the
    reference raises; UnsupportedOp for an odd nfft or one too long."""

    def __init__(self, T, dtMs: float, windowLen: int, taperFraction: float = 0.1,
                 nfft: int = 0, demean: bool = True):
        self.T = np.dtype(T).type
        self.dtMs = float(dtMs)
        self.windowLen = int(windowLen)
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_spectrum_compile(
            0 if self.T == np.float32 else 1, self.dtMs, self.windowLen, float(taperFraction),
            int(nfft), 1 if demean else 0, C.byref(self._h)))
        n, b = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().pseg_spectrum_info_get(self._h, C.byref(n), C.byref(b)))
        self.nfft, self.bins = int(n.value), int(b.value)
        self._acc = {}

    def binHz(self, k: int) -> float:
        """binHz (fft/spectrum.nim:121-123)."""
        return k * 1000 / (self.nfft * self.dtMs)

    def averageAmplitudeInto(self, samples, meta: ChunkMeta | None = None, firstCol: int = 0,
                             amp=None, *, row_len=None, stream=None, sync: bool = True):
        """averageAmplitudeInto (fft/spectrum.nim:410-433) on a [rows, cols]
        GPU tensor: returns (amp, used) — amp a float64 [bins] GPU tensor,
        used the rows averaged (read back when `sync`; else None)."""
        import torch
        if not isinstance(samples, torch.Tensor) or not samples.is_cuda or samples.dim() != 2 \
                or not samples.is_contiguous():
            raise ValueError("samples must be a C-contiguous [rows, cols] GPU tensor")
        if np.dtype(str(samples.dtype).replace("torch.", "")) != self.T:
            raise ValueError(f"samples hold {samples.dtype}, the plan {self.T}")
        rows, cols = samples.shape
        if row_len is None:
            if meta is None:
                raise ValueError("averageAmplitudeInto needs a ChunkMeta (or row_len)")
            if not sameInterval(meta.dtMs, self.dtMs):
                raise ValueError(f"the rows' interval {meta.dtMs} ms is not the plan's {self.dtMs} ms")
            row_len = meta.device_row_len(cols, samples.device)
        dev = samples.device
        acc = self._acc.get(dev)
        if acc is None:
            acc = self._acc[dev] = torch.empty(self.bins + 1, dtype=torch.float64, device=dev)
        if amp is None:
            amp = torch.empty(self.bins, dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().pseg_spectrum_average(
            self._h, samples.data_ptr(), rows, cols, row_len.data_ptr(), int(firstCol),
            acc.data_ptr(), amp.data_ptr(), _stream_handle(stream)))
        used = int(acc[self.bins].item()) if sync else None
        return amp, used

    def amplitudeInto(self, samples, meta: ChunkMeta | None = None, firstCol: int = 0,
                      amp=None, phase=None, *, row_len=None, stream=None):
        """amplitudeInto (fft/spectrum.nim:295-312) of each row's window from
        firstCol: (amp, phase), [rows, bins] GPU tensors of the plan's dtype.
        A window holding NaN or Inf gives NaN in every bin; rows with no live
        sample in the window are left as they are (amp and phase are created
        zeroed when not given)."""
        import torch
        if not isinstance(samples, torch.Tensor) or not samples.is_cuda or samples.dim() != 2 \
                or not samples.is_contiguous():
            raise ValueError("samples must be a C-contiguous [rows, cols] GPU tensor")
        if np.dtype(str(samples.dtype).replace("torch.", "")) != self.T:
            raise ValueError(f"samples hold {samples.dtype}, the plan {self.T}")
        rows, cols = samples.shape
        if row_len is None:
            if meta is None:
                raise ValueError("amplitudeInto needs a ChunkMeta (or row_len)")
            row_len = meta.device_row_len(cols, samples.device)
        for t in (amp, phase):
            if t is not None and (t.dtype != samples.dtype or tuple(t.shape) != (rows, self.bins)
                                  or not t.is_contiguous() or not t.is_cuda):
                raise ValueError(f"amp and phase must be contiguous [{rows}, {self.bins}] GPU "
                                 f"tensors of {samples.dtype}")
        if amp is None:
            amp = torch.zeros((rows, self.bins), dtype=samples.dtype, device=samples.device)
        if phase is None:
            phase = torch.zeros((rows, self.bins), dtype=samples.dtype, device=samples.device)
        _lib.check(_lib.lib().pseg_spectrum_amplitude(
            self._h, samples.data_ptr(), rows, cols, row_len.data_ptr(), int(firstCol),
            amp.data_ptr(), phase.data_ptr(), _stream_handle(stream)))
        return amp, phase

    def __del__(self):
        try:
            if self._h:
                _lib.lib().pseg_spectrum_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def initSpectrumPlan(T, dtMs: float, windowLen: int, taperFraction: float = 0.1, nfft: int = 0,
                     demean: bool = True) -> SpectrumPlan:
    return SpectrumPlan(T, dtMs, windowLen, taperFraction, nfft, demean)
