"""First-break picking on the GPU (qc/firstbreak.nim): an optional
minimum-phase Butterworth band-pass, then STA/LTA, the modified energy
ratio or the modified Coppens ratio on the trace energy, in a fixed or
LMO-guided search window, with a confidence per pick from the windowed SNR.

```python
params = FirstBreakParams(pickMethod=pmEnergyRatio, windowMs=20, highCutHz=60,
                          guideVelocity=2000, beforeMs=60, afterMs=120)
plan = initFirstBreakPlan(np.float32, params, dtMs=2.0, samples=3001)
picks = plan.pickInto(samples, meta)     # timeMs, col, snrDb, confidence per row
```

One GPU thread per row runs the reference's serial arithmetic (prefix
energies, the attribute's argmax, the filter's recursion) in float64 in the
same order as the oracle, so the pick columns are the oracle's, bit for bit.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, _check_panel, _rows_of, _stream_handle

# PickMethod (qc/firstbreak.nim:140-144)
pmStaLta, pmEnergyRatio, pmCoppens = 0, 1, 2
NaN = math.nan
Inf = math.inf


@dataclasses.dataclass
class FirstBreakParams:
    """FirstBreakParams (qc/firstbreak.nim:146-170), the reference's defaults."""
    pickMethod: int = pmEnergyRatio
    windowMs: float = 20.0
    ltaMs: float = 100.0
    stabilization: float = 0.2
    lowCutHz: float = 0.0
    highCutHz: float = Inf
    slopeDbPerOctave: float = 24.0
    searchStartMs: float = NaN
    searchEndMs: float = NaN
    guideVelocity: float = NaN
    guideT0Ms: float = 0.0
    beforeMs: float = 50.0
    afterMs: float = 100.0
    snrWindowMs: float = 40.0
    lowDb: float = 6.0
    highDb: float = 20.0


class PfbParams(C.Structure):
    _fields_ = [("method", C.c_int32), ("reserved", C.c_int32)] + [
        (k, C.c_double) for k in ("window_ms", "lta_ms", "stabilization", "low_cut_hz",
                                  "high_cut_hz", "slope_db_per_octave", "search_start_ms",
                                  "search_end_ms", "guide_velocity", "guide_t0_ms", "before_ms",
                                  "after_ms", "snr_window_ms", "low_db", "high_db")]


_NAMES = (("window_ms", "windowMs"), ("lta_ms", "ltaMs"), ("stabilization", "stabilization"),
          ("low_cut_hz", "lowCutHz"), ("high_cut_hz", "highCutHz"),
          ("slope_db_per_octave", "slopeDbPerOctave"), ("search_start_ms", "searchStartMs"),
          ("search_end_ms", "searchEndMs"), ("guide_velocity", "guideVelocity"),
          ("guide_t0_ms", "guideT0Ms"), ("before_ms", "beforeMs"), ("after_ms", "afterMs"),
          ("snr_window_ms", "snrWindowMs"), ("low_db", "lowDb"), ("high_db", "highDb"))


class FirstBreakPicks:
    """The picks of a block, one per row (FirstBreakPick, qc/firstbreak.nim:
    172-177), as GPU tensors: timeMs (float64, NaN: no pick or no time axis),
    col (int64, -1: no pick), snrDb (float64, NaN: none), confidence
    (float64, 0 without a pick)."""

    def __init__(self, timeMs, col, snrDb, confidence):
        self.timeMs, self.col, self.snrDb, self.confidence = timeMs, col, snrDb, confidence

    def __len__(self):
        return int(self.col.numel())


class FirstBreakPlan:
    """initFirstBreakPlan (qc/firstbreak.nim:195-251) on the GPU.  ValueError
    as the reference raises."""

    def __init__(self, T, params: FirstBreakParams, dtMs: float, samples: int):
        self.T = np.dtype(T).type
        self.params = params
        self.dtMs = float(dtMs)
        self.samples = int(samples)
        fp = PfbParams()
        fp.method = int(params.pickMethod)
        for c_name, name in _NAMES:
            setattr(fp, c_name, float(getattr(params, name)))
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_firstbreak_compile(
            0 if self.T == np.float32 else 1, C.byref(fp), self.dtMs, self.samples,
            C.byref(self._h)))
        m = C.c_int64()
        _lib.check(_lib.lib().pseg_firstbreak_required_margin(self._h, C.byref(m)))
        self.requiredMargin = int(m.value)

    def workspaceSize(self, rows: int, cols: int) -> int:
        b = C.c_size_t()
        _lib.check(_lib.lib().pseg_firstbreak_workspace_size(self._h, int(rows), int(cols),
                                                             C.byref(b)))
        return int(b.value)

    def pickInto(self, samples, meta: ChunkMeta | None = None, dst: FirstBreakPicks | None = None,
                 *, row_len=None, keys=None, first_sample: int = 0, ws=None, stream=None):
        """pickInto (qc/firstbreak.nim:492-541) of every live row of a
        [rows, cols] GPU tensor (cols <= samples), each to its live length on
        the meta's time axis (t0Ms and firstSample; offset for a guided
        window).  Returns the picks (dst when given)."""
        import torch
        _check_panel(samples, "samples", self.T)
        rows, cols = samples.shape
        if cols > self.samples:
            raise ValueError(f"the block has {cols} columns, the plan takes {self.samples}")
        dev = samples.device
        if row_len is None:
            row_len = _rows_of(meta, None, rows, cols, dev, self.dtMs, "pickInto")
            keys = meta.device_keys(dev)
            first_sample = meta.firstSample
        else:
            row_len = _rows_of(None, row_len, rows, cols, dev, None, "pickInto")
        if keys is not None and (not isinstance(keys, torch.Tensor) or not keys.is_cuda
                                 or keys.dtype != torch.float64
                                 or tuple(keys.shape) != (rows, 2) or not keys.is_contiguous()):
            raise ValueError("keys must be a contiguous [rows, 2] float64 GPU tensor")
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        need = self.workspaceSize(rows, cols)
        with torch.cuda.stream(s):
            if ws is None:
                ws = torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
            elif (not isinstance(ws, torch.Tensor) or not ws.is_cuda or ws.device != dev
                  or ws.numel() * ws.element_size() < need):
                raise ValueError(f"ws must be a GPU tensor of at least {need} bytes")
            if dst is None:
                dst = FirstBreakPicks(torch.empty(rows, dtype=torch.float64, device=dev),
                                      torch.empty(rows, dtype=torch.int64, device=dev),
                                      torch.empty(rows, dtype=torch.float64, device=dev),
                                      torch.empty(rows, dtype=torch.float64, device=dev))
            elif len(dst) != rows:
                raise ValueError(f"dst holds {len(dst)} picks, the meta {rows} rows")
        _lib.check(_lib.lib().pseg_firstbreak_pick(
            self._h, samples.data_ptr(), rows, cols, row_len.data_ptr(),
            keys.data_ptr() if keys is not None else None, int(first_sample), ws.data_ptr(),
            ws.numel() * ws.element_size(), dst.timeMs.data_ptr(), dst.col.data_ptr(),
            dst.snrDb.data_ptr(), dst.confidence.data_ptr(), _stream_handle(s)))
        return dst

    def __del__(self):
        try:
            if self._h:
                _lib.lib().pseg_firstbreak_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def initFirstBreakPlan(T, params: FirstBreakParams, dtMs: float, samples: int) -> FirstBreakPlan:
    return FirstBreakPlan(T, params, dtMs, samples)
