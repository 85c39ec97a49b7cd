"""The f-k amplitude of a gather (fft/fk.nim), on the GPU through the C ABI:
initFkPlan, fkRows and FkPlan.fkAmplitudeInto."""
from __future__ import annotations

import ctypes as C
import dataclasses
from typing import Optional

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, _check_panel, _dtype_code, _stream_handle, sameInterval


@dataclasses.dataclass(frozen=True)
class FkAxes:
    """FkAxes (fft/fk.nim:50-58): amp[ik * nFreq + jf] is wavenumber index
    ik (zero at nK // 2, positive k later) and frequency bin jf."""
    nFreq: int
    nK: int
    dfHz: float
    dkPerTrace: float


def fkRows(meta: ChunkMeta, cols: int, firstCol: int, nCols: int, rows: Optional[range] = None) -> int:
    """fkRows (fft/fk.nim:204-216): the live rows of a block `cols` wide
    with a sample at `firstCol` or later (the count initFkPlan wants)."""
    return len(_entering(meta, cols, firstCol, nCols, rows))


def _entering(meta: ChunkMeta, cols: int, firstCol: int, nCols: int, rows=None) -> np.ndarray:
    if firstCol < 0 or nCols < 1:
        return np.zeros(0, np.int32)
    rl = meta.row_len(cols)
    idx = np.nonzero(rl > firstCol)[0]
    if rows is not None:
        idx = idx[(idx >= max(rows.start, 0)) & (idx < rows.stop)]
    return idx.astype(np.int32)


class FkPlan:
    """An f-k transform of `traces` live rows by windows of `samples`
    columns (initFkPlan, fft/fk.nim:82-106).  Immutable."""

    def __init__(self, T, dtMs: float, traces: int, samples: int, taperFraction: float = 0.1):
        self.T = np.dtype(T)
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_fk_compile(_dtype_code(T), float(dtMs), int(traces),
                                              int(samples), float(taperFraction),
                                              C.byref(self._h)))
        nf, nk = C.c_int32(), C.c_int32()
        df, dk = C.c_double(), C.c_double()
        _lib.check(_lib.lib().pseg_fk_axes(self._h, C.byref(nf), C.byref(nk), C.byref(df),
                                           C.byref(dk)))
        self.axes = FkAxes(nf.value, nk.value, df.value, dk.value)
        self.dtMs = float(dtMs)
        self.traces = int(traces)
        self.samples = int(samples)
        b = C.c_size_t()
        _lib.check(_lib.lib().pseg_fk_workspace_size(self._h, C.byref(b)))
        self.workspaceSize = int(b.value)
        self._ws = {}

    @property
    def nFreq(self) -> int:
        return self.axes.nFreq

    @property
    def nK(self) -> int:
        return self.axes.nK

    @property
    def dfHz(self) -> float:
        return self.axes.dfHz

    @property
    def dkPerTrace(self) -> float:
        return self.axes.dkPerTrace

    def fkAmplitudeInto(self, samples, meta: ChunkMeta, firstCol: int, amp, rows=None,
                        stream=None) -> int:
        """fkAmplitudeInto (fft/fk.nim:438-504): the modulus of the window
        firstCol ..< firstCol + samples into amp (a float64 GPU tensor of
        nK · nFreq values, frequency fastest, unscaled).  Returns the live
        rows used.  ValueError as the reference raises it; asynchronous on
        `stream`."""
        import torch
        _check_panel(samples, "samples", self.T)
        n, cols = samples.shape
        if len(meta) != n:
            raise ValueError(f"samples have {n} rows, the meta {len(meta)}")
        if not sameInterval(meta.dtMs, self.dtMs):
            raise ValueError(f"the rows' interval {meta.dtMs} ms is not the plan's {self.dtMs} ms")
        if firstCol < 0:
            raise ValueError(f"firstCol must be ≥ 0, got {firstCol}")
        if (not isinstance(amp, torch.Tensor) or not amp.is_cuda or amp.dtype != torch.float64
                or amp.numel() != self.nK * self.nFreq or not amp.is_contiguous()
                or amp.device != samples.device):
            raise ValueError("amp needs nK·nFreq float64 values on the samples' GPU")
        live = _entering(meta, cols, firstCol, self.samples, rows)
        if len(live) != self.traces:
            raise ValueError(f"the window has {len(live)} live rows, the plan was built for "
                             f"{self.traces}")
        dev = samples.device
        key = str(dev)
        ws = self._ws.get(key)
        if ws is None:
            ws = torch.empty(self.workspaceSize, dtype=torch.uint8, device=dev)
            self._ws[key] = ws
        lr = torch.from_numpy(live).to(dev)
        rl = meta.device_row_len(cols, dev)
        self._hold = lr
        _lib.check(_lib.lib().pseg_fk_amplitude(
            self._h, samples.data_ptr(), n, cols, rl.data_ptr(), lr.data_ptr(), int(firstCol),
            amp.data_ptr(), ws.data_ptr(), ws.numel(), _stream_handle(stream)))
        return len(live)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                _lib.lib().pseg_fk_destroy(h)
            except Exception:
                pass


def initFkPlan(T, dtMs: float, traces: int, samples: int, taperFraction: float = 0.1) -> FkPlan:
    return FkPlan(T, dtMs, traces, samples, taperFraction)
