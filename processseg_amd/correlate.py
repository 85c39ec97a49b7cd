"""Vibroseis correlation on the GPU (vibroseis/correlate.nim) and the sweep
generation its pilot comes from (vibroseis/sweep.nim, taper.nim, pilot.nim).

```python
pilot = generatePilot(linearSweep(6, 80, 12000, 2, initSweepTaper(300, 300)))
plan = initCorrelationPlan(float32, pilot, dtMs=2, recordSamples=10001)
plan.correlateInto(raw, correlated, meta)     # [rows, 10001] -> [rows, outSamples]
```
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, _check_panel, _dtype_code, _rows_of, _stream_handle

cmCorrelate, cmDeconvolve = 0, 1


@dataclasses.dataclass(frozen=True)
class SweepTaper:
    """cos² (tpCos2) end ramps in ms (vibroseis/taper.nim:60-67)."""
    startMs: float = 0.0
    endMs: float = 0.0


def initSweepTaper(startMs: float, endMs: float) -> SweepTaper:
    for name, v in (("startMs", startMs), ("endMs", endMs)):
        if not math.isnan(v) and not (v >= 0 and v < math.inf):
            raise ValueError(f"a taper {name} must be ≥ 0 and finite (or NaN for none), got {v}")
    return SweepTaper(float(startMs), float(endMs))


@dataclasses.dataclass(frozen=True)
class SweepSpec:
    """A linear sweep (vibroseis/sweep.nim:81-94, 185-192)."""
    f1Hz: float
    f2Hz: float
    lengthMs: float
    dtMs: float
    taper: SweepTaper = SweepTaper()
    startPhaseDeg: float = 0.0
    amplitude: float = 1.0


def linearSweep(f1Hz, f2Hz, lengthMs, dtMs, taper: SweepTaper = SweepTaper(),
                startPhaseDeg: float = 0.0, amplitude: float = 1.0) -> SweepSpec:
    spec = SweepSpec(float(f1Hz), float(f2Hz), float(lengthMs), float(dtMs), taper,
                     float(startPhaseDeg), float(amplitude))
    sweepSamples(spec)  # validates
    return spec


def samplesInLength(lengthMs: float, dtMs: float) -> int:
    """samplesInLength (core/types.nim:147-166)."""
    if not (lengthMs > 0 and lengthMs < math.inf and dtMs > 0 and dtMs < math.inf):
        return 0
    n = lengthMs / dtMs
    if n > 2147483647:
        return 0
    r = math.floor(n + 0.5)  # round half away from zero, n > 0
    return int(r) if abs(n - r) <= 1e-6 else int(math.floor(n))


def sweepSamples(spec: SweepSpec) -> np.ndarray:
    """sweepSamples (vibroseis/sweep.nim:344-349), float64, by the library."""
    n = samplesInLength(spec.lengthMs, spec.dtMs)
    x = np.zeros(max(n, 1), np.float64)
    _lib.check(_lib.lib().pseg_linear_sweep(spec.f1Hz, spec.f2Hz, spec.lengthMs, spec.dtMs,
                                            spec.taper.startMs, spec.taper.endMs,
                                            spec.startPhaseDeg, spec.amplitude,
                                            x.ctypes.data, n))
    return x[:n]


@dataclasses.dataclass
class Pilot:
    """A pilot sweep (vibroseis/pilot.nim:55-62)."""
    samples: np.ndarray
    dtMs: float
    t0Ms: float = 0.0

    def __len__(self):
        return int(self.samples.shape[0])


def generatePilot(spec: SweepSpec) -> Pilot:
    """generatePilot (vibroseis/pilot.nim:115-119), float64 samples."""
    return Pilot(sweepSamples(spec), spec.dtMs, 0.0)


class CorrelationPlan:
    """initCorrelationPlan (vibroseis/correlate.nim:82-160) on the GPU: the
    pilot spectrum is computed in float64 on the host, the per-bin multiplier
    stays on the device."""

    def __init__(self, T, pilot: Pilot, dtMs: float, recordSamples: int, outSamples: int = 0,
                 mode: int = cmCorrelate, epsPercent: float = 1.0, normalize: bool = True):
        self.T = np.dtype(T)
        p = np.ascontiguousarray(pilot.samples, np.float64)
        self._pilot = p
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_corr_compile(_dtype_code(T), p.ctypes.data, p.shape[0],
                                                pilot.dtMs, pilot.t0Ms, dtMs, recordSamples,
                                                outSamples, mode, epsPercent,
                                                1 if normalize else 0, C.byref(self._h)))
        info = _lib.CorrInfoC()
        _lib.check(_lib.lib().pseg_corr_info_get(self._h, C.byref(info)))
        self.nfft = int(info.nfft)
        self.shift = int(info.shift)
        self.outSamples = int(info.out_samples)
        self.pilotSamples = int(info.pilot_samples)
        self.recordSamples = int(info.record_samples)
        self.dtMs = float(info.dt_ms)
        self.scale = float(info.scale)
        self.threadsPerRecord = int(info.threads_per_record)
        self.mode = mode

    def correlateInto(self, src, dst, meta: ChunkMeta = None, *, row_len=None, stream=None):
        """correlateInto (vibroseis/correlate.nim:421-464)."""
        _check_panel(src, "src", self.T)
        _check_panel(dst, "dst", self.T, rows=src.shape[0], cols=self.outSamples,
                     device=src.device)
        rows, cols = src.shape
        if cols > self.recordSamples:
            raise ValueError(f"the records have {cols} samples, the plan takes at most "
                             f"{self.recordSamples}")
        if row_len is None and meta is not None and meta.firstSample != 0:
            raise ValueError("correlation needs whole records from sample 0, got firstSample "
                             f"{meta.firstSample}")
        row_len = _rows_of(meta, row_len, rows, cols, src.device, self.dtMs, "correlateInto",
                           for_correlation=True)
        _lib.check(_lib.lib().pseg_corr_apply(self._h, src.data_ptr(), rows, cols,
                                              row_len.data_ptr(), dst.data_ptr(), dst.shape[1],
                                              _stream_handle(stream)))
        return dst

    def __del__(self):
        try:
            if self._h:
                _lib.lib().pseg_corr_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass


def initCorrelationPlan(T, pilot: Pilot, dtMs: float, recordSamples: int, outSamples: int = 0,
                        mode: int = cmCorrelate, epsPercent: float = 1.0,
                        normalize: bool = True) -> CorrelationPlan:
    return CorrelationPlan(T, pilot, dtMs, recordSamples, outSamples, mode, epsPercent, normalize)
