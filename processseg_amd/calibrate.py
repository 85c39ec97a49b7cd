"""Calibration of recorded values to volts or physical units
(core/calibrate.nim, core/units.nim), mirrored on the C ABI.

The rules are checked and indexed by the library (``pseg_calib_compile``,
initCalibrationPlan) and each row's factor is resolved there too
(``pseg_calib_resolve``, resolveRow): both are per-row metadata work on the
host.  The samples are scaled on the GPU by the row kernel's calibration step
(``calibrateOp`` / ``uncalibrateOp`` in a pipeline, or ``CalibrationPlan.
calibrate`` here, which runs that step alone).

Units (``sensitivity``, ``descale``, ``openCircuitSensitivity``) are converted
to SI here as the reference's constructors do (core/units.nim:153-349); a rule
reaches the library with its factors in SI.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np

from . import _lib

NaN = math.nan

# AmplitudeUnit (core/units.nim:57-64)
(auRecorded, auVolt, auVelocity, auAcceleration, auPressure, auForce) = range(6)
AMPLITUDE_UNIT_NAMES = ["recorded", "V", "m/s", "m/s²", "Pa", "N"]

# SensitivityUnit (core/units.nim:66-78)
(suVPerMPerS, suVPerInPerS, suVPerMPerS2, suVPerG, suVPerBar, suVPerPa, suMicroVPerPa,
 suDbReVPerMicroPa, suVPerN, suVPerKN, suVPerLbf) = range(11)
# DescaleUnit (core/units.nim:99-104)
(duVPerCount, duMilliVPerCount, duMicroVPerCount, duNanoVPerCount) = range(4)
# SensitivityForm (core/units.nim:80-84)
(snNone, snAsWired, snOpenCircuit) = range(3)

# CalibrationTarget, FileFactors (core/calibrate.nim:104-115)
(ctPhysical, ctVolts) = range(2)
(ffFallback, ffFirst, ffIgnore) = range(3)
# CalibrationStatus, FactorSource (core/calibrate.nim:155-165)
(csConverted, csUnresolved, csKept) = range(3)
(fsNone, fsRule, fsFile) = range(3)

_INCH = 0.0254
_STANDARD_G = 9.80665
_BAR = 1e5
_POUND_FORCE = 0.45359237 * _STANDARD_G


def _finite(x: float) -> bool:
    return isinstance(x, (int, float)) and math.isfinite(x)


def quantityOf(u: int) -> int:
    """quantityOf (core/units.nim:153-160): what a sensitivity in `u` measures."""
    if u in (suVPerMPerS, suVPerInPerS):
        return auVelocity
    if u in (suVPerMPerS2, suVPerG):
        return auAcceleration
    if u in (suVPerBar, suVPerPa, suMicroVPerPa, suDbReVPerMicroPa):
        return auPressure
    if u in (suVPerN, suVPerKN, suVPerLbf):
        return auForce
    raise ValueError(f"unknown sensitivity unit {u}")


def _to_si(value: float, u: int) -> float:
    # toSi (core/units.nim:162-172): volts per SI unit of the quantity
    if u in (suVPerMPerS, suVPerMPerS2, suVPerPa, suVPerN):
        return value
    if u == suVPerInPerS:
        return value / _INCH
    if u == suVPerG:
        return value / _STANDARD_G
    if u == suVPerBar:
        return value / _BAR
    if u == suMicroVPerPa:
        return value * 1e-6
    if u == suDbReVPerMicroPa:
        try:
            return math.pow(10.0, value / 20) * 1e6
        except OverflowError:
            return math.inf
    if u == suVPerKN:
        return value * 1e-3
    if u == suVPerLbf:
        return value / _POUND_FORCE
    raise ValueError(f"unknown sensitivity unit {u}")


def _checked_si(value: float, unit: int) -> float:
    # checkedSi (core/units.nim:185-194)
    if unit == suDbReVPerMicroPa:
        if not _finite(value):
            raise ValueError(f"a sensitivity in dB must be finite, got {value}")
    elif not _finite(value) or value <= 0:
        raise ValueError(f"a sensitivity must be finite and positive, got {value}")
    si = _to_si(float(value), unit)
    if not math.isfinite(si) or si <= 0:
        raise ValueError(f"a sensitivity of {value} (unit {unit}) is out of range")
    return si


@dataclasses.dataclass(frozen=True)
class Sensitivity:
    """A sensor's output per unit of what it measures, in SI
    (core/units.nim:86-97).  ``Sensitivity()`` is not set."""
    form: int = snNone
    unit: int = auRecorded
    si: float = NaN

    @property
    def voltsPerUnit(self) -> float:
        return NaN if self.form == snNone else self.si

    def isSet(self) -> bool:
        return self.form != snNone


def sensitivity(value: float, unit: int) -> Sensitivity:
    """sensitivity (core/units.nim:196-207): as wired to the recorder."""
    return Sensitivity(snAsWired, quantityOf(unit), _checked_si(value, unit))


def openCircuitSensitivity(value: float, unit: int, coilOhms: float, shuntOhms: float,
                           inputOhms: float = math.inf, inSeries: int = 1,
                           inParallel: int = 1) -> Sensitivity:
    """openCircuitSensitivity (core/units.nim:209-250):
    S = n·S0 · R_L / (n·R_c / m + R_L), R_L = shunt ∥ input."""
    s0 = _checked_si(value, unit)
    if not _finite(coilOhms) or coilOhms < 0:
        raise ValueError(f"the coil resistance must be finite and ≥ 0 Ω, got {coilOhms}")
    if math.isnan(shuntOhms) or shuntOhms <= 0:
        raise ValueError(f"the shunt resistance must be positive (Inf: none), got {shuntOhms}")
    if math.isnan(inputOhms) or inputOhms <= 0:
        raise ValueError(f"the input impedance must be positive (Inf: not counted), got {inputOhms}")
    if inSeries < 1 or inParallel < 1:
        raise ValueError("inSeries and inParallel must be ≥ 1")
    n, m = float(inSeries), float(inParallel)
    load = 1 / (1 / shuntOhms + 1 / inputOhms)
    rc = n * coilOhms / m
    divider = 1.0 if load == math.inf else load / (rc + load)
    v = n * s0 * divider
    if not math.isfinite(v) or v <= 0:
        raise ValueError(f"an open-circuit sensitivity of {value} with these resistances is out "
                         f"of range: {v} V per unit")
    return Sensitivity(snOpenCircuit, quantityOf(unit), v)


@dataclasses.dataclass(frozen=True)
class Descale:
    """Volts at the recorder's input per recorded unit (core/units.nim:
    106-111).  ``Descale()`` is not set."""
    vpc: float = 0.0

    @property
    def voltsPerCount(self) -> float:
        return self.vpc if self.vpc > 0 else NaN

    def isSet(self) -> bool:
        return self.vpc > 0


def descale(value: float, unit: int) -> Descale:
    """descale (core/units.nim:331-341)."""
    if not _finite(value) or value <= 0:
        raise ValueError(f"a descale must be finite and positive, got {value}")
    scale = {duVPerCount: 1.0, duMilliVPerCount: 1e-3, duMicroVPerCount: 1e-6,
             duNanoVPerCount: 1e-9}[unit]
    v = value * scale
    if not math.isfinite(v) or v <= 0:
        raise ValueError(f"a descale of {value} (unit {unit}) is out of range")
    return Descale(v)


@dataclasses.dataclass
class FileScale:
    """What a file's headers say about one trace's scale, in SI
    (core/units.nim:113-127); NaN where they say nothing."""
    descale: float = NaN
    sensitivity: float = NaN
    unit: int = auRecorded
    physical: bool = False


def initFileScale() -> FileScale:
    return FileScale()


@dataclasses.dataclass(frozen=True)
class StationKey:
    line: float
    point: float
    index: int = 0


def stationKey(line: float, point: float, index: int = 0) -> StationKey:
    """A receiver station (index 0: any index)."""
    return StationKey(float(line), float(point), int(index))


@dataclasses.dataclass
class RuleMatch:
    """The rows a rule applies to (core/calibrate.nim:115-131); unset
    criteria match any row."""
    sensorType: int = -1
    component: int = -1
    channelSet: int = -1
    channel: float = NaN
    station: Optional[StationKey] = None


@dataclasses.dataclass
class CalibrationRule:
    """Factors for the rows `match` selects (core/calibrate.nim:133-145)."""
    match: RuleMatch = dataclasses.field(default_factory=RuleMatch)
    descale: Descale = dataclasses.field(default_factory=Descale)
    sensitivity: Sensitivity = dataclasses.field(default_factory=Sensitivity)
    gainDb: float = NaN

    def to_c(self) -> _lib.CalibRuleC:
        r = _lib.CalibRuleC()
        m = self.match
        r.sensor_type = int(m.sensorType)
        r.component = int(m.component)
        r.channel_set = int(m.channelSet)
        r.channel = float(m.channel)
        if m.station is not None:
            r.has_station = 1
            r.station_line = m.station.line
            r.station_point = m.station.point
            r.station_index = m.station.index
        r.descale = self.descale.voltsPerCount
        r.sensitivity = self.sensitivity.voltsPerUnit
        r.sensitivity_unit = self.sensitivity.unit
        r.gain_db = float(self.gainDb)
        return r


@dataclasses.dataclass
class Calibration:
    """The rules and how to apply them (core/calibrate.nim:147-153)."""
    rules: List[CalibrationRule] = dataclasses.field(default_factory=list)
    target: int = ctPhysical
    fileFactors: int = ffFallback
    scaleAuxiliary: bool = True

    def rules_c(self):
        arr = (_lib.CalibRuleC * max(1, len(self.rules)))()
        for i, r in enumerate(self.rules):
            arr[i] = r.to_c()
        return arr


def initCalibration(rules: Sequence[CalibrationRule] = (), target: int = ctPhysical,
                    fileFactors: int = ffFallback, scaleAuxiliary: bool = True) -> Calibration:
    return Calibration(list(rules), target, fileFactors, scaleAuxiliary)


@dataclasses.dataclass
class RowCalibration:
    """How one row is calibrated (core/calibrate.nim:167-182)."""
    factor: float
    unit: int
    status: int
    descaleFrom: int
    sensitivityFrom: int
    descaleRule: int
    sensitivityRule: int
    gainRule: int
    conflict: bool
    mismatch: bool


@dataclasses.dataclass
class CalibrationSummary:
    """Counts over the rows of one resolve or calibrate (core/calibrate.nim:
    184-193)."""
    units: List[int]
    converted: int = 0
    unresolved: int = 0
    kept: int = 0
    noDescale: int = 0
    noSensitivity: int = 0
    conflicts: int = 0
    mismatches: int = 0


def _summary(s: _lib.CalibSummaryC) -> CalibrationSummary:
    return CalibrationSummary(list(s.units), s.converted, s.unresolved, s.kept, s.no_descale,
                              s.no_sensitivity, s.conflicts, s.mismatches)


def _row(r: _lib.RowCalibrationC) -> RowCalibration:
    return RowCalibration(r.factor, r.unit, r.status, r.descale_from, r.sensitivity_from,
                          r.descale_rule, r.sensitivity_rule, r.gain_rule, bool(r.conflict),
                          bool(r.mismatch))


class CalibrationPlan:
    """initCalibrationPlan (core/calibrate.nim:316-360): a validated,
    indexed Calibration; immutable."""

    def __init__(self, cal: Calibration):
        self.calibration = cal
        self._h = C.c_void_p()
        arr = cal.rules_c()
        _lib.check(_lib.lib().pseg_calib_compile(arr, len(cal.rules), int(cal.target),
                                                 int(cal.fileFactors), int(cal.scaleAuxiliary),
                                                 C.byref(self._h)))

    @property
    def target(self) -> int:
        return self.calibration.target

    def resolveInto(self, meta, rows: Optional[list] = None) -> CalibrationSummary:
        """resolveInto (core/calibrate.nim:514-527): each row's calibration
        into `rows` (one per row of `meta`, replaced in place) and the
        counts.  Touches no sample."""
        n = len(meta)
        if rows is not None and len(rows) != n:
            raise ValueError(f"need one RowCalibration per row: {n} rows, {len(rows)} given")
        cr = meta.calib_rows()
        out = (_lib.RowCalibrationC * max(1, n))()
        s = _lib.CalibSummaryC()
        _lib.check(_lib.lib().pseg_calib_resolve(self._h, cr, n, int(meta.hasFileScales), out,
                                                 C.byref(s)))
        if rows is not None:
            for i in range(n):
                rows[i] = _row(out[i])
        return _summary(s)

    def resolve(self, meta) -> List[RowCalibration]:
        rows = [None] * len(meta)
        self.resolveInto(meta, rows)
        return rows

    def _run(self, src, dst, meta, inverse: bool) -> CalibrationSummary:
        from .pipeline import calibrateOp, initPipeline, uncalibrateOp
        op = uncalibrateOp(self.calibration) if inverse else calibrateOp(self.calibration)
        plan = initPipeline([op]).compile(_np_dtype(src), meta.dtMs, src.shape[1])
        plan.apply(src, dst, meta)
        return self.resolveInto(meta)

    def calibrate(self, samples, meta) -> CalibrationSummary:
        """calibrate (core/calibrate.nim:625-637) in place on a GPU tensor."""
        return self._run(samples, samples, meta, False)

    def calibrateInto(self, src, dst, meta) -> CalibrationSummary:
        """calibrateInto (core/calibrate.nim:639-652)."""
        return self._run(src, dst, meta, False)

    def uncalibrate(self, samples, meta) -> CalibrationSummary:
        """uncalibrate (core/calibrate.nim:654-668) in place."""
        return self._run(samples, samples, meta, True)

    def uncalibrateInto(self, src, dst, meta) -> CalibrationSummary:
        """uncalibrateInto (core/calibrate.nim:670-680)."""
        return self._run(src, dst, meta, True)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value:
            try:
                _lib.lib().pseg_calib_destroy(h)
            except Exception:
                pass


def initCalibrationPlan(cal: Calibration) -> CalibrationPlan:
    return CalibrationPlan(cal)


def _np_dtype(t):
    import torch
    return {torch.float32: np.float32, torch.float64: np.float64}[t.dtype]


def ratioDbInto(a: CalibrationPlan, b: CalibrationPlan, meta, dst) -> tuple:
    """ratioDbInto (core/calibrate.nim:682-705): per row 20·log10(factorA /
    factorB) for rows both convert to the same unit, else NaN; returns
    (rows compared, minDb, maxDb)."""
    n = len(meta)
    if len(dst) != n:
        raise ValueError(f"need one value per row: {n} rows, {len(dst)} given")
    ra, rb = a.resolve(meta), b.resolve(meta)
    cnt, lo, hi = 0, NaN, NaN
    for r in range(n):
        x, y = ra[r], rb[r]
        if x.status == csConverted and y.status == csConverted and x.unit == y.unit:
            db = 20 * math.log10(abs(x.factor / y.factor))
            dst[r] = db
            lo, hi = (db, db) if cnt == 0 else (min(lo, db), max(hi, db))
            cnt += 1
        else:
            dst[r] = NaN
    return cnt, lo, hi


def scaleRows(samples, meta, factors, dst=None):
    """scaleRows / scaleRowsInto (core/calibrate.nim:583-613) on GPU
    tensors: every row that did not fail times factors[r] over its live
    length, whatever its status (a NaN factor fills the row with NaN), on the
    row kernel's calibration step."""
    f = np.ascontiguousarray(factors, np.float64)
    n = len(meta)
    if f.shape != (n,):
        raise ValueError(f"need one factor per row: {n} rows, {f.size} factors")
    for i, x in enumerate(f):
        if math.isinf(x):
            raise ValueError(f"factor {i} is infinite")
    k = f.copy()
    if meta.failed.size:
        k[meta.failed] = 1.0  # a failed row is never touched
    from .pipeline import calibrateOp, initPipeline
    plan = initPipeline([calibrateOp(initCalibration())]).compile(
        _np_dtype(samples), meta.dtMs, samples.shape[1])
    out = samples if dst is None else dst
    plan.apply(samples, out, meta, calib_factors=k.reshape(1, n))
    return out
