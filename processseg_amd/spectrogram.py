"""Spectrograms of traces on the GPU (fft/spectrogram.nim): one workgroup per
frame, Hann-weighted and zero-padded to nfft, through the same real FFT as
the filters; amplitudes |X[k]|·2/Σw (1/Σw at 0 Hz and Nyquist) in float64.

```python
sg = initSpectrogramPlan(np.float32, windowLen=257, hop=10, dtMs=1, samples=8001)
amp = sg.spectrogramInto(trace)          # [frames, bins] float64 on the device
```
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, _check_panel, _rows_of, _stream_handle


class SpectrogramPlan:
    """initSpectrogramPlan (fft/spectrogram.nim:63-85)."""

    def __init__(self, T, windowLen: int, hop: int, dtMs: float, samples: int, nfft: int = 0):
        self.T = np.dtype(T).type
        self.windowLen, self.hop, self.dtMs = int(windowLen), int(hop), float(dtMs)
        self.samples = int(samples)
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_spectrogram_compile(
            0 if self.T == np.float32 else 1, self.windowLen, self.hop, self.dtMs,
            self.samples, int(nfft), C.byref(self._h)))
        fr, bi, nf = C.c_int32(), C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().pseg_spectrogram_info_get(self._h, C.byref(fr), C.byref(bi),
                                                        C.byref(nf)))
        self.frames, self.bins, self.nfft = int(fr.value), int(bi.value), int(nf.value)

    def frequency(self, k: int) -> float:
        """The frequency of bin k, Hz (fft/spectrogram.nim:119-121)."""
        return k * 1000.0 / (self.nfft * self.dtMs)

    def centre(self, frame: int) -> int:
        """The trace column frame `frame` is centred on (fft/spectrogram.nim:115-117)."""
        return frame * self.hop

    def spectrogramInto(self, trace, amp=None, *, stream=None):
        """spectrogramInto (fft/spectrogram.nim:263-346) of one GPU trace of at
        most `samples` values: amp is [frames, bins] float64 on the device."""
        import torch
        if not isinstance(trace, torch.Tensor) or not trace.is_cuda or trace.dim() != 1 \
                or not trace.is_contiguous():
            raise ValueError("trace must be a contiguous 1-D GPU tensor")
        if np.dtype(str(trace.dtype).replace("torch.", "")) != self.T:
            raise ValueError(f"trace holds {trace.dtype}, the plan {self.T}")
        if amp is None:
            amp = torch.empty((self.frames, self.bins), dtype=torch.float64, device=trace.device)
        elif (not isinstance(amp, torch.Tensor) or not amp.is_cuda or amp.device != trace.device
              or amp.dtype != torch.float64 or amp.numel() != self.frames * self.bins
              or not amp.is_contiguous()):
            raise ValueError("amp needs frames·bins contiguous float64 values on the trace's GPU")
        if trace.numel() > self.samples:
            raise ValueError(f"the trace has {trace.numel()} samples, the plan takes at most "
                             f"{self.samples}")
        _lib.check(_lib.lib().pseg_spectrogram_apply(self._h, trace.data_ptr(), trace.numel(),
                                                     amp.data_ptr(), _stream_handle(stream)))
        return amp

    def averageWorkspace(self, rows: int) -> int:
        """Bytes of device workspace averageSpectrogramInto needs for `rows` rows."""
        b = C.c_size_t()
        _lib.check(_lib.lib().pseg_spectrogram_average_workspace_size(self._h, int(rows),
                                                                       C.byref(b)))
        return int(b.value)

    def averageSpectrogramInto(self, samples, meta: ChunkMeta | None = None, amp=None, *,
                               row_len=None, ws=None, stream=None, sync: bool = True):
        """averageSpectrogramInto (fft/spectrogram.nim:370-430) of the live
        rows of a [rows, cols] GPU tensor (cols <= samples), one launch over
        rows × frames: returns (amp, used) — amp a float64 [frames, bins] GPU
        tensor, used the rows averaged (read after `stream` when `sync`; else
        a one-value float64 GPU tensor).  The same input gives the same bits
        on every call."""
        import torch
        _check_panel(samples, "samples", self.T)
        rows, cols = samples.shape
        if cols > self.samples:
            raise ValueError(f"the rows have {cols} samples, the plan takes at most {self.samples}")
        row_len = _rows_of(meta, row_len, rows, cols, samples.device, self.dtMs,
                           "averageSpectrogramInto")
        dev = samples.device
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        need = self.averageWorkspace(rows)
        with torch.cuda.stream(s):
            if ws is None:
                ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
            elif (not isinstance(ws, torch.Tensor) or not ws.is_cuda or ws.device != dev
                  or not ws.is_contiguous() or ws.numel() * ws.element_size() < need):
                raise ValueError(f"ws must be a contiguous GPU tensor of at least {need} bytes")
            if amp is None:
                amp = torch.empty((self.frames, self.bins), dtype=torch.float64, device=dev)
            elif (not isinstance(amp, torch.Tensor) or amp.dtype != torch.float64
                  or amp.numel() != self.frames * self.bins or not amp.is_contiguous()
                  or amp.device != dev):
                raise ValueError("amp needs frames·bins contiguous float64 values on the "
                                 "samples' GPU")
            used = torch.empty(1, dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().pseg_spectrogram_average(
            self._h, samples.data_ptr(), rows, cols, row_len.data_ptr(), ws.data_ptr(),
            ws.numel() * ws.element_size(), amp.data_ptr(), used.data_ptr(), _stream_handle(s)))
        if not sync:
            return amp, used
        s.synchronize()
        return amp, int(used.item())

    def __del__(self):
        try:
            if self._h:
                _lib.lib().pseg_spectrogram_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def initSpectrogramPlan(T, windowLen: int, hop: int, dtMs: float, samples: int,
                        nfft: int = 0) -> SpectrogramPlan:
    return SpectrogramPlan(T, windowLen, hop, dtMs, samples, nfft)
