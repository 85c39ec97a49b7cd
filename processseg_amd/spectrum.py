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
from .pipeline import ChunkMeta, _check_panel, _rows_of, _stream_handle


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

    def binHz(self, k: int) -> float:
        """binHz (fft/spectrum.nim:121-123)."""
        return k * 1000 / (self.nfft * self.dtMs)

    def workspaceSize(self, rows: int) -> int:
        """Bytes of device workspace an average over `rows` rows needs (the
        workgroups' partial sums)."""
        b = C.c_size_t()
        _lib.check(_lib.lib().pseg_spectrum_workspace_size(self._h, int(rows), C.byref(b)))
        return int(b.value)

    def averageAmplitudeInto(self, samples, meta: ChunkMeta | None = None, firstCol: int = 0,
                             amp=None, *, row_len=None, ws=None, used=None, stream=None,
                             sync: bool = True):
        """averageAmplitudeInto (fft/spectrum.nim:410-433) on a [rows, cols]
        GPU tensor: returns (amp, used) — amp a float64 [bins] GPU tensor,
        used the rows averaged (read back after `stream` has finished when
        `sync`; else a one-value float64 GPU tensor written on `stream`).
        `ws`: a uint8 GPU tensor of at least workspaceSize(rows) bytes, or
        None for one allocated for this call; `used`: a one-value float64 GPU
        tensor to write the rows used into, or None for one allocated for
        this call.  The same input gives the same bits on every call."""
        import torch
        _check_panel(samples, "samples", self.T)
        rows, cols = samples.shape
        if int(firstCol) < 0:
            raise ValueError(f"firstCol must be ≥ 0, got {firstCol}")
        row_len = _rows_of(meta, row_len, rows, cols, samples.device, self.dtMs,
                           "averageAmplitudeInto")
        dev = samples.device
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        need = self.workspaceSize(rows)
        if ws is not None and (not isinstance(ws, torch.Tensor) or not ws.is_cuda
                               or ws.device != dev or not ws.is_contiguous()
                               or ws.numel() * ws.element_size() < need):
            raise ValueError(f"ws must be a contiguous GPU tensor of at least {need} bytes")
        if amp is not None and (not isinstance(amp, torch.Tensor) or amp.dtype != torch.float64
                                or amp.numel() != self.bins or not amp.is_contiguous()
                                or amp.device != dev):
            raise ValueError(f"amp must be {self.bins} contiguous float64 values on {dev}")
        if used is not None and (not isinstance(used, torch.Tensor)
                                 or used.dtype != torch.float64 or used.numel() != 1
                                 or used.device != dev):
            raise ValueError(f"used must be one float64 value on {dev}")
        if ws is None or amp is None or used is None:
            # allocations belong to the stream that writes them
            with torch.cuda.stream(s):
                if ws is None:
                    ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
                if amp is None:
                    amp = torch.empty(self.bins, dtype=torch.float64, device=dev)
                if used is None:
                    used = torch.empty(1, dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().pseg_spectrum_average(
            self._h, samples.data_ptr(), rows, cols, row_len.data_ptr(), int(firstCol),
            ws.data_ptr(), ws.numel() * ws.element_size(), amp.data_ptr(), used.data_ptr(),
            _stream_handle(s)))
        if not sync:
            return amp, used
        s.synchronize()
        return amp, int(used.item())

    def amplitudeInto(self, samples, meta: ChunkMeta | None = None, firstCol: int = 0,
                      amp=None, phase=None, *, row_len=None, stream=None):
        """amplitudeInto (fft/spectrum.nim:295-312) of each row's window from
        firstCol: (amp, phase), [rows, bins] GPU tensors of the plan's dtype.
        A window holding NaN or Inf gives NaN in every bin; rows with no live
        sample in the window are left as they are (amp and phase are created
        zeroed when not given)."""
        import torch
        _check_panel(samples, "samples", self.T)
        rows, cols = samples.shape
        if int(firstCol) < 0:
            raise ValueError(f"firstCol must be ≥ 0, got {firstCol}")
        row_len = _rows_of(meta, row_len, rows, cols, samples.device, self.dtMs, "amplitudeInto")
        for t, name in ((amp, "amp"), (phase, "phase")):
            if t is not None:
                _check_panel(t, name, self.T, rows=rows, cols=self.bins, device=samples.device)
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
