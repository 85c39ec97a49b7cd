"""The reference's pipeline API over the GPU plan (pipeline.nim, filter/design.nim,
gain/gain.nim, core/types.nim), mirrored name for name.

```python
flow = initPipeline([demeanOp(), bandPassOp(Corners(4, 8, 60, 80)), agcOp(windowMs=500)])
plan = flow.compile(float32, dtMs=2, samples=2001)   # ValueError if invalid
plan.apply(raw, processed, meta)                      # src -> dst, device tensors
plan.apply(samples, meta)                             # in place
```

Panels are torch tensors on the GPU ([rows, cols], C-contiguous, float32 or
float64).  A ``ChunkMeta`` holds the per-row facts the kernels read (liveness
and live length, flattened once into a device int32 array; t0Ms and offset,
which mutes and gain curves read, into a float64 pair per row).  Ops outside
the GPU path (IIR, calibration) raise ``UnsupportedOp`` at compile: keep the
CPU plan for such a flow.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import UnsupportedOp, WorkspaceError  # noqa: F401  (re-exported)

Inf = math.inf
NaN = math.nan

float32 = np.float32
float64 = np.float64

# TrendRemoval (filter/edges.nim:8-14)
trNone, trMean, trLinear = 0, 1, 2
# IirPhase (filter/iir.nim:77-80)
ipZero, ipMinimum = 0, 1
# AgcMode (gain/gain.nim:92-95)
agcRms, agcMeanAbs = 0, 1
# NormStat / NormScope (gain/gain.nim:72-81)
normMax, normRms, normPercentile = 0, 1, 2
perTrace, perPanel = 0, 1
# OpKind (pipeline.nim:92-108)
(opDemean, opDetrend, opScale, opGainDb, opClip, opPolarity, opMuteTop, opMuteBottom,
 opFilter, opIir, opAgc, opNormalize, opGainCurve, opCalibrate, opUncalibrate) = range(15)
# TraceStatus of parseSEG as ChunkMeta reads it
tsLive, tsUnknown, tsDead, tsAuxiliary = 0, 1, 2, 3

_KIND_NAMES = ["opDemean", "opDetrend", "opScale", "opGainDb", "opClip", "opPolarity",
               "opMuteTop", "opMuteBottom", "opFilter", "opIir", "opAgc", "opNormalize",
               "opGainCurve", "opCalibrate", "opUncalibrate"]


# ------------------------------------------------------------- filter specs

@dataclasses.dataclass(frozen=True)
class Corners:
    """The corners of a cos² trapezoid in Hz (filter/design.nim:63-67)."""
    f1: float
    f2: float
    f3: float
    f4: float


@dataclasses.dataclass(frozen=True)
class FilterSpec:
    """A zero-phase response (filter/design.nim:78-90): kind 0 trapezoid,
    1 Butterworth magnitude; notches as (freqHz, widthHz)."""
    kind: int = 0
    f1: float = 0.0
    f2: float = 0.0
    f3: float = Inf
    f4: float = Inf
    low_hz: float = 0.0
    low_slope: float = 24.0
    high_hz: float = Inf
    high_slope: float = 24.0
    notches: tuple = ()
    zero_dc: bool = False

    def withNotch(self, freqHz: float, widthHz: float) -> "FilterSpec":
        """withNotch (filter/design.nim:212-218)."""
        if not (freqHz > 0 and freqHz < Inf):
            raise ValueError(f"a notch frequency must be finite and > 0, got {freqHz}")
        if not (widthHz > 0 and widthHz < Inf):
            raise ValueError(f"a notch width must be finite and > 0, got {widthHz}")
        return dataclasses.replace(self, notches=self.notches + ((float(freqHz), float(widthHz)),))

    def withZeroDc(self) -> "FilterSpec":
        """withZeroDc (filter/design.nim:220-223)."""
        return dataclasses.replace(self, zero_dc=True)

    def response(self, f: float, dtMs: float) -> float:
        """response (filter/design.nim:236-255), computed by the library."""
        s = spec_c(self)
        return float(_lib.lib().pseg_response(C.byref(s), f, dtMs))

    def impulseLength(self, dtMs: float) -> int:
        """impulseLength (filter/design.nim:353-366)."""
        s = spec_c(self)
        L = C.c_int64()
        _lib.check(_lib.lib().pseg_impulse_length(C.byref(s), dtMs, C.byref(L)))
        return int(L.value)


def _check_corners(c: Corners) -> None:
    # checkCorners (filter/design.nim:103-115)
    for f in (c.f1, c.f2, c.f3, c.f4):
        if math.isnan(f):
            raise ValueError(f"a corner is NaN: {c}")
    has_low, has_high = c.f2 > 0, c.f3 < Inf
    if has_low and not (c.f1 >= 0 and c.f1 < c.f2):
        raise ValueError(f"a low cut needs 0 ≤ f1 < f2 (a sloped ramp), got {c}")
    if not has_low and (c.f1 != 0 or c.f2 != 0):
        raise ValueError(f"no low cut needs f1 = f2 = 0, got {c}")
    if has_high and not (c.f3 < c.f4 and c.f4 < Inf):
        raise ValueError(f"a high cut needs f3 < f4 < Inf (a sloped ramp), got {c}")
    if not has_high and c.f4 != Inf:
        raise ValueError(f"no high cut needs f3 = f4 = Inf, got {c}")
    if has_low and has_high and not (c.f2 <= c.f3):
        raise ValueError(f"a band-pass needs f2 ≤ f3, got {c}")


def bandPass(c: Corners) -> FilterSpec:
    """bandPass (filter/design.nim:177-181)."""
    _check_corners(c)
    return FilterSpec(kind=0, f1=float(c.f1), f2=float(c.f2), f3=float(c.f3), f4=float(c.f4))


def lowCut(f1: float, f2: float) -> FilterSpec:
    return bandPass(Corners(f1, f2, Inf, Inf))


def highCut(f3: float, f4: float) -> FilterSpec:
    return bandPass(Corners(0, 0, f3, f4))


def butterBandPass(lowHz: float, highHz: float, slopeDbPerOctave: float = 24.0,
                   highSlopeDbPerOctave: float = NaN) -> FilterSpec:
    """butterBandPass (filter/design.nim:191-200)."""
    hs = slopeDbPerOctave if math.isnan(highSlopeDbPerOctave) else highSlopeDbPerOctave
    if not (lowHz >= 0 and lowHz < Inf):
        raise ValueError(f"a low-cut corner must be finite and ≥ 0 (0: none), got {lowHz}")
    if not (highHz > 0):
        raise ValueError(f"a high-cut corner must be > 0 (Inf: none), got {highHz}")
    for on, s in ((lowHz > 0, slopeDbPerOctave), (highHz < Inf, hs)):
        if on and not (s > 0 and s < Inf):
            raise ValueError(f"a slope must be finite and > 0 dB/octave, got {s}")
    if not (lowHz < highHz):
        raise ValueError(f"the low-cut corner must be below the high-cut one, got {lowHz} and {highHz}")
    return FilterSpec(kind=1, low_hz=float(lowHz), low_slope=float(slopeDbPerOctave),
                      high_hz=float(highHz), high_slope=float(hs))


def butterLowCut(hz: float, slopeDbPerOctave: float = 24.0) -> FilterSpec:
    if not (hz > 0):
        raise ValueError(f"a low-cut corner must be > 0, got {hz}")
    return butterBandPass(hz, Inf, slopeDbPerOctave, slopeDbPerOctave)


def butterHighCut(hz: float, slopeDbPerOctave: float = 24.0) -> FilterSpec:
    if not (hz < Inf):
        raise ValueError(f"a high-cut corner must be finite, got {hz}")
    return butterBandPass(0, hz, slopeDbPerOctave, slopeDbPerOctave)


def spec_c(spec: FilterSpec) -> _lib.FilterSpecC:
    s = _lib.FilterSpecC()
    s.kind = int(spec.kind)
    s.zero_dc = int(bool(spec.zero_dc))
    if len(spec.notches) > _lib.MAX_NOTCH:
        raise ValueError(f"at most {_lib.MAX_NOTCH} notches")
    s.n_notch = len(spec.notches)
    s.f1, s.f2, s.f3, s.f4 = spec.f1, spec.f2, spec.f3, spec.f4
    s.low_hz, s.low_slope, s.high_hz, s.high_slope = (spec.low_hz, spec.low_slope,
                                                       spec.high_hz, spec.high_slope)
    for i, (f, w) in enumerate(spec.notches):
        s.notch_hz[i] = f
        s.notch_width[i] = w
    return s


# -------------------------------------------------------------------- ops

MAX_KNOTS = 32
mkTime, mkTable, mkLmo = 0, 1, 2


@dataclasses.dataclass(frozen=True)
class MuteCurve:
    """The mute time of each trace (core/mute.nim:31-52): kind mkTime (one
    time), mkTable (linear in |offset| between knots, flat beyond them) or
    mkLmo (t0Ms + |offset| msPerDist + padMs).  Build with timeMute,
    tableMute or lmoMute, which validate it."""
    kind: int
    tMs: float = 0.0
    offsets: tuple = ()
    timesMs: tuple = ()
    msPerDist: float = 0.0
    t0Ms: float = 0.0
    padMs: float = 0.0

    def muteTimeMs(self, offset: float) -> float:
        """muteTimeMs (core/mute.nim:100-118)."""
        if self.kind == mkTime:
            return self.tMs
        if self.kind == mkLmo:
            return self.t0Ms + abs(offset) * self.msPerDist + self.padMs
        x = abs(offset)
        if math.isnan(x):
            return math.nan
        o, t = self.offsets, self.timesMs
        if x <= o[0]:
            return t[0]
        if x >= o[-1]:
            return t[-1]
        lo, hi = 0, len(o) - 1
        while hi - lo > 1:
            mid = (lo + hi) >> 1
            if o[mid] <= x:
                lo = mid
            else:
                hi = mid
        u = (x - o[lo]) / (o[hi] - o[lo])
        return t[lo] + u * (t[hi] - t[lo])


def _finite(x) -> bool:
    return isinstance(x, (int, float)) and math.isfinite(x)


def timeMute(tMs: float) -> MuteCurve:
    """timeMute (core/mute.nim:58-62)."""
    if not _finite(tMs):
        raise ValueError(f"a mute time must be finite, got {tMs}")
    return MuteCurve(mkTime, tMs=float(tMs))


def tableMute(offsets, timesMs) -> MuteCurve:
    """tableMute (core/mute.nim:64-86): |offset| knots strictly ascending."""
    offsets = tuple(float(v) for v in offsets)
    timesMs = tuple(float(v) for v in timesMs)
    if len(offsets) == 0:
        raise ValueError("a mute table needs at least one knot")
    if len(offsets) != len(timesMs):
        raise ValueError(f"a mute table has {len(offsets)} offsets and {len(timesMs)} times")
    for i, (o, t) in enumerate(zip(offsets, timesMs)):
        if not _finite(o) or o < 0:
            raise ValueError(f"mute table offsets must be finite and ≥ 0, got {o}")
        if i > 0 and o <= offsets[i - 1]:
            raise ValueError(f"mute table offsets must be strictly ascending: {offsets[i - 1]} "
                             f"is followed by {o}")
        if not _finite(t):
            raise ValueError(f"mute table times must be finite, got {t}")
    return MuteCurve(mkTable, offsets=offsets, timesMs=timesMs)


def lmoMute(velocity: float, t0Ms: float = 0.0, padMs: float = 0.0) -> MuteCurve:
    """lmoMute (core/mute.nim:88-98): velocity in distance units per second."""
    if not _finite(velocity) or velocity <= 0:
        raise ValueError(f"an LMO mute velocity must be finite and positive, got {velocity}")
    if not _finite(t0Ms) or not _finite(padMs):
        raise ValueError(f"LMO mute times must be finite, got t0Ms {t0Ms}, padMs {padMs}")
    return MuteCurve(mkLmo, msPerDist=1000 / float(velocity), t0Ms=float(t0Ms), padMs=float(padMs))


# GainCurveKind (gain/curves.nim:68-72)
gcKnots, gcPower, gcExponential = 0, 1, 2


@dataclasses.dataclass(frozen=True)
class GainKnot:
    """A point of a knotGain curve (gain/curves.nim:63-66)."""
    tMs: float
    dB: float


@dataclasses.dataclass(frozen=True)
class GainCurve:
    """A gain against time (gain/curves.nim:74-84): kind gcKnots (linear in
    dB between knots, flat outside them), gcPower ((max(t, minMs)/refMs)^power)
    or gcExponential (dbPerS dB per second).  Build with knotGain, powerGain
    or exponentialGain, which validate it; ``GainCurve()`` is the unbuilt
    curve (no knots) that a plan refuses."""
    kind: int = gcKnots
    knots: tuple = ()           # ((tMs, dB), ...)
    power: float = 0.0
    refMs: float = 1000.0
    minMs: float = 1.0
    dbPerS: float = 0.0

    def gainDbAt(self, tMs: float) -> float:
        """gainDbAt (gain/curves.nim:134-153)."""
        if math.isnan(tMs):
            return math.nan
        if self.kind == gcPower:
            return 20 * self.power * math.log10(max(tMs, self.minMs) / self.refMs)
        if self.kind == gcExponential:
            return self.dbPerS * tMs / 1000
        k = self.knots
        if not k:
            return math.nan
        if tMs <= k[0][0]:
            return k[0][1]
        if tMs >= k[-1][0]:
            return k[-1][1]
        lo, hi = 0, len(k) - 1
        while hi - lo > 1:
            mid = (lo + hi) >> 1
            if k[mid][0] <= tMs:
                lo = mid
            else:
                hi = mid
        f = (tMs - k[lo][0]) / (k[hi][0] - k[lo][0])
        return k[lo][1] + f * (k[hi][1] - k[lo][1])

    def factorAt(self, tMs: float) -> float:
        """factorAt (gain/curves.nim:154-160): 10^(gainDbAt/20), computed
        directly for a power curve."""
        if math.isnan(tMs):
            return math.nan
        if self.kind == gcPower:
            return math.pow(max(tMs, self.minMs) / self.refMs, self.power)
        db = self.gainDbAt(tMs)
        return math.nan if math.isnan(db) else math.pow(10.0, db / 20)


def knotGain(knots) -> GainCurve:
    """knotGain (gain/curves.nim:100-112): at least one knot, every value
    finite, times strictly increasing.  Knots are GainKnot or (tMs, dB)."""
    ks = tuple((float(k.tMs), float(k.dB)) if isinstance(k, GainKnot) else
               (float(k[0]), float(k[1])) for k in knots)
    if not ks:
        raise ValueError("a knot gain needs at least one knot")
    for i, (t, d) in enumerate(ks):
        if not (math.isfinite(t) and math.isfinite(d)):
            raise ValueError(f"knot {i} is not finite: ({t}, {d})")
        if i > 0 and not t > ks[i - 1][0]:
            raise ValueError(f"knot times must strictly increase: {ks[i - 1][0]} ms is "
                             f"followed by {t} ms")
    return GainCurve(gcKnots, knots=ks)


def powerGain(power: float, refMs: float = 1000.0, minMs: float = 1.0) -> GainCurve:
    """powerGain (gain/curves.nim:114-125)."""
    if not _finite(power):
        raise ValueError(f"a power must be finite, got {power}")
    if not (_finite(refMs) and refMs > 0):
        raise ValueError(f"refMs must be finite and positive, got {refMs}")
    if not (_finite(minMs) and minMs > 0):
        raise ValueError(f"minMs must be finite and positive, got {minMs}")
    return GainCurve(gcPower, power=float(power), refMs=float(refMs), minMs=float(minMs))


def exponentialGain(dbPerS: float) -> GainCurve:
    """exponentialGain (gain/curves.nim:127-131)."""
    if not _finite(dbPerS):
        raise ValueError(f"dbPerS must be finite, got {dbPerS}")
    return GainCurve(gcExponential, dbPerS=float(dbPerS))


def heldGain(curve: GainCurve, untilMs: float, step: float = 1.02) -> GainCurve:
    """heldGain (gain/curves.nim:162-212): the curve up to untilMs and its
    gain at untilMs after it, as knots: an exponential is two knots from 0, a
    power curve knots from minMs every `step` in time plus one at untilMs, a
    knot curve its knots before untilMs plus the gain there."""
    if not (_finite(untilMs) and untilMs > 0):
        raise ValueError(f"untilMs must be finite and > 0, got {untilMs}")
    if not (_finite(step) and step > 1):
        raise ValueError(f"step must be finite and > 1, got {step}")
    if curve.kind == gcKnots and not curve.knots:
        raise ValueError("an unbuilt curve cannot be held")
    out = []
    if curve.kind == gcExponential:
        out = [(0.0, curve.gainDbAt(0)), (float(untilMs), curve.gainDbAt(untilMs))]
    elif curve.kind == gcPower:
        start = curve.minMs
        if untilMs <= start:
            out.append((float(untilMs), curve.gainDbAt(untilMs)))
        else:
            count = int(math.log(untilMs / start) / math.log(step)) + 2
            if count > 100_000:
                raise ValueError(f"holding this power curve would take {count} knots; widen "
                                 "step or shorten untilMs")
            t = start
            while t < untilMs:
                out.append((t, curve.gainDbAt(t)))
                nxt = t * step
                if not nxt > t:
                    break
                t = nxt
            if not out or out[-1][0] < untilMs:
                out.append((float(untilMs), curve.gainDbAt(untilMs)))
    else:
        for k in curve.knots:
            if k[0] < untilMs:
                out.append(k)
            else:
                break
        if not out or out[-1][0] < untilMs:
            out.append((float(untilMs), curve.gainDbAt(untilMs)))
    return knotGain(out)


@dataclasses.dataclass
class Op:
    """One operation as plain data (pipeline.nim:158-174).  Parameters are
    edited in place and checked at compile."""
    kind: int
    value: float = 0.0          # scale factor, or gain in dB
    lo: float = 0.0             # clip
    hi: float = 0.0
    spec: Optional[FilterSpec] = None
    trend: int = trLinear       # filter
    window_ms: float = 0.0      # agc
    mode: int = agcRms
    target: float = 1.0         # agc, normalize
    epsilon: float = 1e-9       # agc
    stat: int = normMax         # normalize
    scope: int = perTrace
    percentile: float = 99.0
    mute: Optional[MuteCurve] = None  # muteTop, muteBottom
    taper_ms: float = 0.0
    phase: int = 0              # iir: ipZero, ipMinimum
    curve: Optional[GainCurve] = None  # gainCurve, tabulated over t_min_ms .. t_max_ms
    t_min_ms: float = 0.0
    t_max_ms: float = math.nan
    calibration: Optional[object] = None  # calibrate, uncalibrate (calibrate.Calibration)

    def to_c(self) -> _lib.OpC:
        o = _lib.OpC()
        o.kind = self.kind
        o.trend = self.trend
        o.mode = self.mode
        o.stat = self.stat
        o.scope = self.scope
        o.value = self.value
        o.lo = self.lo
        o.hi = self.hi
        o.window_ms = self.window_ms
        o.target = self.target
        o.epsilon = self.epsilon
        o.percentile = self.percentile
        o.iir_phase = self.phase
        if self.spec is not None:
            o.spec = spec_c(self.spec)
        if self.mute is not None:
            mute_fields(o, self.mute, self.taper_ms)
        if self.curve is not None:
            gain_fields(o, self.curve, self.t_min_ms, self.t_max_ms)
        if self.calibration is not None:
            cal = self.calibration
            n = len(cal.rules)
            # the struct keeps the rule array it points to alive
            if n:
                arr = (_lib.CalibRuleC * n)()
                for i, r in enumerate(cal.rules):
                    arr[i] = r.to_c()
                o.calib_rules = arr
            o.calib_n_rules = n
            o.calib_target = int(cal.target)
            o.calib_file_factors = int(cal.fileFactors)
            o.calib_scale_aux = int(bool(cal.scaleAuxiliary))
        return o


def gain_fields(o, c: GainCurve, t_min_ms: float, t_max_ms: float):
    """Fill the gain curve fields of a C op; the struct keeps the knot
    arrays it points to alive."""
    o.gain_kind = c.kind
    o.gain_knots = len(c.knots)
    if c.knots:
        o.gain_knot_ms = (C.c_double * len(c.knots))(*[k[0] for k in c.knots])
        o.gain_knot_db = (C.c_double * len(c.knots))(*[k[1] for k in c.knots])
    o.gain_power = c.power
    o.gain_ref_ms = c.refMs
    o.gain_min_ms = c.minMs
    o.gain_db_per_s = c.dbPerS
    o.gain_tmin_ms = float(t_min_ms)
    o.gain_tmax_ms = float(t_max_ms)


def mute_fields(o, c: MuteCurve, taper_ms: float):
    """Fill the mute fields of a C op (pseg_op, or the oracle's)."""
    if c.kind == mkTable and len(c.offsets) > MAX_KNOTS:
        raise UnsupportedOp(f"a mute table of {len(c.offsets)} knots (at most {MAX_KNOTS})")
    o.mute_kind = c.kind
    o.mute_knots = len(c.offsets)
    o.mute_t_ms = c.t0Ms if c.kind == mkLmo else c.tMs
    o.mute_ms_per_dist = c.msPerDist
    o.mute_pad_ms = c.padMs
    o.taper_ms = float(taper_ms)
    for i, (a, b) in enumerate(zip(c.offsets, c.timesMs)):
        o.mute_offsets[i] = a
        o.mute_times[i] = b


def demeanOp() -> Op:
    return Op(opDemean)


def detrendOp() -> Op:
    return Op(opDetrend)


def scaleOp(factor: float) -> Op:
    return Op(opScale, value=float(factor))


def gainDbOp(dB: float) -> Op:
    return Op(opGainDb, value=float(dB))


def clipOp(lo: float, hi: float) -> Op:
    return Op(opClip, lo=float(lo), hi=float(hi))


def polarityOp() -> Op:
    return Op(opPolarity)


def muteTopOp(curve: MuteCurve, taperMs: float = 0.0) -> Op:
    """muteTopOp (pipeline.nim:283-285): a top mute with a taper before it."""
    return Op(opMuteTop, mute=curve, taper_ms=float(taperMs))


def muteBottomOp(curve: MuteCurve, taperMs: float = 0.0) -> Op:
    """muteBottomOp (pipeline.nim:287-289): a bottom mute with a taper after it."""
    return Op(opMuteBottom, mute=curve, taper_ms=float(taperMs))


def gainCurveOp(curve: GainCurve, tMinMs: float = 0.0, tMaxMs: float = NaN) -> Op:
    """gainCurveOp (pipeline.nim:318-322): a time-variant gain, tabulated over
    tMinMs .. tMaxMs (NaN: to the end of the record, ``recordSamples``)."""
    return Op(opGainCurve, curve=curve, t_min_ms=float(tMinMs), t_max_ms=float(tMaxMs))


def filterOp(spec: FilterSpec, trend: int = trLinear) -> Op:
    return Op(opFilter, spec=spec, trend=trend)


def bandPassOp(corners: Corners, trend: int = trLinear) -> Op:
    return filterOp(bandPass(corners), trend)


def agcOp(windowMs: float, mode: int = agcRms, target: float = 1.0,
          epsilon: float = 1e-9) -> Op:
    return Op(opAgc, window_ms=float(windowMs), mode=mode, target=float(target),
              epsilon=float(epsilon))


def normalizeOp(stat: int, scope: int = perTrace, target: float = 1.0,
                percentile: float = 99.0) -> Op:
    return Op(opNormalize, stat=stat, scope=scope, target=float(target),
              percentile=float(percentile))


def iirOp(spec: FilterSpec, phase: int = 0, trend: int = trLinear) -> Op:
    """iirOp (pipeline.nim:300-303): a Butterworth biquad cascade (a
    butterBandPass / butterLowCut / butterHighCut spec), zero phase
    (ipZero, forward-backward) or minimum phase (ipMinimum)."""
    return Op(opIir, spec=spec, trend=trend, phase=int(phase))


def calibrateOp(cal) -> Op:
    """calibrateOp (pipeline.nim:324-329): recorded values to volts or
    physical units per row (a calibrate.Calibration); it must come before
    every op that changes amplitudes."""
    return Op(opCalibrate, calibration=cal)


def uncalibrateOp(cal) -> Op:
    """uncalibrateOp (pipeline.nim:331-334): rows in `cal`'s units back to
    recorded values."""
    return Op(opUncalibrate, calibration=cal)


def _ops_c(ops: Sequence[Op]):
    arr = (_lib.OpC * max(1, len(ops)))()
    for i, op in enumerate(ops):
        arr[i] = op.to_c()
    return arr


# ------------------------------------------------------------------- meta

# pseg_calib_row (include/pseg.h) as a numpy record, field for field
_CALIB_ROW = np.dtype({
    "names": ["status", "failed", "sensor_type", "file_physical", "component", "channel_set",
              "rcv_index", "channel", "rcv_line", "rcv_point", "file_descale",
              "file_sensitivity", "file_unit", "reserved"],
    "formats": ["u1", "u1", "u1", "u1", "<i4", "<i4", "<i4", "<f8", "<f8", "<f8", "<f8", "<f8",
                "<i4", "<i4"],
    "offsets": [0, 1, 2, 3, 4, 8, 12, 16, 24, 32, 40, 48, 56, 60],
    "itemsize": 64})


class ChunkMeta:
    """The per-row facts of a block (core/types.nim:69-122): status per row
    (tsLive, tsUnknown, tsDead, tsAuxiliary), trace length numSamples per
    row, failed rows, dt and firstSample, and the trace keys mutes read
    (t0Ms, offset per row; default 0 and NaN; traceKeys, the source,
    receiver and midpoint coordinates and header numbers fold maps read).
    ``row_len(cols)`` flattens
    isLive/liveSamples to one int32 per row (-1: not live)."""

    def __init__(self, rows: int, dtMs: float, numSamples=None, status=None,
                 failed: Iterable[int] = (), firstSample: int = 0, t0Ms=None, offset=None,
                 traceKeys=None, sensorType=None, component=None, channel=None, rcvLine=None,
                 rcvPoint=None, rcvIndex=None, channelSet=None, scales=None):
        if not (isinstance(dtMs, (int, float)) and math.isfinite(dtMs) and dtMs > 0):
            raise ValueError(f"dtMs must be finite and positive, got {dtMs}")
        if firstSample < 0:
            raise ValueError(f"firstSample must be ≥ 0, got {firstSample}")
        self.rows = int(rows)
        self.dtMs = float(dtMs)
        self.firstSample = int(firstSample)
        big = np.iinfo(np.int64).max // 4
        self.numSamples = (np.full(rows, big, np.int64) if numSamples is None
                           else np.ascontiguousarray(numSamples, np.int64))
        self.status = (np.zeros(rows, np.uint8) if status is None
                       else np.ascontiguousarray(status, np.uint8))
        self.failed = np.ascontiguousarray(sorted(int(f) for f in failed), np.int64)
        if self.numSamples.shape != (rows,) or self.status.shape != (rows,):
            raise ValueError("numSamples and status need one value per row")
        self.t0Ms = (np.zeros(rows) if t0Ms is None
                     else np.ascontiguousarray(t0Ms, np.float64).reshape(rows))
        self.offset = (np.full(rows, np.nan) if offset is None
                       else np.ascontiguousarray(offset, np.float64).reshape(rows))
        # the keys fold counting reads ([rows, 10] float64, fold.KEY_COLUMNS)
        self.traceKeys = (None if traceKeys is None
                          else np.ascontiguousarray(traceKeys, np.float64))
        if self.traceKeys is not None and self.traceKeys.shape != (rows, 10):
            raise ValueError(f"traceKeys must be [rows, 10], got {self.traceKeys.shape}")
        # what a calibration reads (TraceKeys.sensorType, component, channel,
        # rcvLine / rcvPoint / rcvIndex; TraceRows.channelSetIndex; the
        # optional FileScale column, core/calibrate.nim:430-497)

        def col(v, fill, dt):
            a = np.full(rows, fill, dt) if v is None else np.ascontiguousarray(v, dt).reshape(rows)
            return a
        self.sensorType = col(sensorType, 0, np.uint8)
        self.component = col(component, 0, np.int32)
        self.channel = col(channel, np.nan, np.float64)
        self.rcvLine = col(rcvLine, np.nan, np.float64)
        self.rcvPoint = col(rcvPoint, np.nan, np.float64)
        self.rcvIndex = col(rcvIndex, 0, np.int32)
        self.channelSet = col(channelSet, -1, np.int32)
        self.scales = None if scales is None else list(scales)
        if self.scales is not None and len(self.scales) != rows:
            raise ValueError(f"scales needs one FileScale per row: {rows} rows, "
                             f"{len(self.scales)} given")
        self._cache = {}

    def __len__(self):
        return self.rows

    def row_len(self, cols: int, for_correlation=False) -> np.ndarray:
        """for_correlation: False / 0 live rows, True / 1 rows a correlation
        uses, 2 every row (the length a calibration step scales)."""
        out = np.zeros(self.rows, np.int32)
        fl = self.failed if self.failed.size else np.zeros(1, np.int64)
        _lib.check(_lib.lib().pseg_meta_flatten(
            self.rows, cols, self.firstSample, self.numSamples.ctypes.data,
            self.status.ctypes.data, fl.ctypes.data, int(self.failed.size),
            int(for_correlation), out.ctypes.data))
        return out

    @property
    def hasFileScales(self) -> bool:
        return self.scales is not None

    def isFailed(self, row: int) -> bool:
        return bool(np.any(self.failed == row))

    def calib_rows(self):
        """The rows as resolveRow reads them (a pseg_calib_row per row), built
        once per meta like its device copies."""
        arr = self._cache.get("calib_rows")
        if arr is not None:
            return arr
        n = self.rows
        a = np.zeros(max(1, n), _CALIB_ROW)
        a["status"][:n] = self.status
        if self.failed.size:
            a["failed"][self.failed[(self.failed >= 0) & (self.failed < n)]] = 1
        a["sensor_type"][:n] = self.sensorType
        a["component"][:n] = self.component
        a["channel_set"][:n] = self.channelSet
        a["rcv_index"][:n] = self.rcvIndex
        a["channel"][:n] = self.channel
        a["rcv_line"][:n] = self.rcvLine
        a["rcv_point"][:n] = self.rcvPoint
        if self.scales is not None:
            a["file_descale"][:n] = [fs.descale for fs in self.scales]
            a["file_sensitivity"][:n] = [fs.sensitivity for fs in self.scales]
            a["file_unit"][:n] = [fs.unit for fs in self.scales]
            a["file_physical"][:n] = [1 if fs.physical else 0 for fs in self.scales]
        else:
            a["file_descale"] = math.nan
            a["file_sensitivity"] = math.nan
        arr = (_lib.CalibRowC * max(1, n)).from_buffer(a)
        self._cache["calib_rows"] = arr
        return arr

    def keys(self) -> np.ndarray:
        """(t0Ms, offset) per row, [rows, 2] float64 (meta.keys(row))."""
        return np.stack([self.t0Ms, self.offset], axis=1)

    def device_keys(self, device):
        import torch
        key = ("keys", str(device))
        t = self._cache.get(key)
        if t is None:
            t = torch.from_numpy(self.keys()).to(device)
            self._cache[key] = t
        return t

    def device_row_len(self, cols: int, device, for_correlation=False):
        import torch
        key = (cols, str(device), int(for_correlation))
        t = self._cache.get(key)
        if t is None:
            t = torch.from_numpy(self.row_len(cols, for_correlation)).to(device)
            self._cache[key] = t
        return t


def sameInterval(dtMs: float, planDtMs: float) -> bool:
    """sameInterval (core/types.nim:136-145)."""
    return abs(dtMs - planDtMs) <= 1e-9 * planDtMs


def decodeWindow(shown: range, margin: int, traceLen: int) -> range:
    """decodeWindow (pipeline.nim:550-560) for an inclusive `shown` given as
    range(a, b + 1); returns range(first, last + 1)."""
    a, b = C.c_int32(), C.c_int32()
    _lib.lib().pseg_decode_window(shown.start, shown.stop - 1, int(min(margin, 2**62)),
                                  traceLen, C.byref(a), C.byref(b))
    return range(a.value, b.value + 1)


def nextFastLen(n: int, even: bool = True) -> int:
    return int(_lib.lib().pseg_next_fast_len(n, 1 if even else 0))


# --------------------------------------------------------------- pipeline

class Pipeline:
    """An ordered list of Ops (pipeline.nim:176-178, 338-380)."""

    def __init__(self, ops: Iterable[Op] = ()):
        self.ops: List[Op] = list(ops)

    def __len__(self):
        return len(self.ops)

    def __getitem__(self, i):
        return self.ops[i]

    def __setitem__(self, i, op):
        self.ops[i] = op

    def add(self, op: Op) -> None:
        self.ops.append(op)

    def compile(self, T, dtMs: float, samples: int, recordSamples: int = 0) -> "PipelinePlan":
        """compile (pipeline.nim:466-523) into a plan on the current GPU."""
        return PipelinePlan(self, T, dtMs, samples, recordSamples)

    def appliedResponseInto(self, T, dtMs: float, samples: int, hz, gain=None) -> np.ndarray:
        """appliedResponseInto (pipeline.nim:840-880): the response of the
        flow as applied (every FFT and IIR filter's trend removal off) to a
        unit impulse at sample samples // 2, |DFT| at the frequencies `hz`
        (Hz); the impulse runs through the plan on the GPU.  Fills and
        returns `gain` (float64, len(hz))."""
        h = np.ascontiguousarray(hz, np.float64)
        out = np.zeros(h.size) if gain is None else gain
        if len(out) != h.size:
            raise ValueError("hz and gain need the same length")
        g = np.zeros(h.size)
        arr = _ops_c(self.ops)
        _lib.check(_lib.lib().pseg_applied_response(arr, len(self.ops), _dtype_code(T),
                                                    float(dtMs), int(samples), h.ctypes.data,
                                                    h.size, g.ctypes.data))
        out[:] = g
        return out

    def requiredMargin(self, dtMs: float) -> int:
        """requiredMargin (pipeline.nim:525-548) without compiling."""
        m = C.c_int32()
        arr = _ops_c(self.ops)
        _lib.check(_lib.lib().pseg_required_margin(arr, len(self.ops), dtMs, C.byref(m)))
        return int(m.value)


def initPipeline(ops: Iterable[Op] = ()) -> Pipeline:
    return Pipeline(ops)


def _dtype_code(T) -> int:
    dt = np.dtype(T)
    if dt == np.float32:
        return 0
    if dt == np.float64:
        return 1
    raise TypeError(f"the sample type must be float32 or float64, got {dt}")


def _stream_handle(stream) -> Optional[int]:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _check_panel(t, name: str, T=None, rows: Optional[int] = None, cols: Optional[int] = None,
                 device=None):
    """ValueError unless `t` is a C-contiguous [rows, cols] GPU tensor (of the
    numpy dtype T, on `device`, of the given shape where those are given):
    what the kernels index without bounds checks is checked here, before any
    launch (the reference raises ValueError for a block that does not match)."""
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if t.dim() != 2:
        raise ValueError(f"{name} must be a [rows, cols] tensor, got rank {t.dim()}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous (row-major)")
    if T is not None and np.dtype(str(t.dtype).replace("torch.", "")) != np.dtype(T):
        raise ValueError(f"{name} holds {t.dtype}, the plan {np.dtype(T)}")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, the samples on {device}")
    if rows is not None and t.shape[0] != rows:
        raise ValueError(f"{name} has {t.shape[0]} rows, expected {rows}")
    if cols is not None and t.shape[1] != cols:
        raise ValueError(f"{name} has {t.shape[1]} columns, expected {cols}")


def _check_row_len(row_len, rows: int, device):
    """ValueError unless `row_len` is a contiguous int32 GPU tensor of `rows`
    values on `device` (one live length per row, -1 for a row left alone)."""
    import torch
    if (not isinstance(row_len, torch.Tensor) or not row_len.is_cuda
            or row_len.dtype != torch.int32 or not row_len.is_contiguous()):
        raise ValueError("row_len must be a contiguous int32 GPU tensor")
    if row_len.numel() != rows:
        raise ValueError(f"row_len has {row_len.numel()} values, the samples {rows} rows")
    if row_len.device != device:
        raise ValueError(f"row_len is on {row_len.device}, the samples on {device}")


def _rows_of(meta, row_len, rows: int, cols: int, device, plan_dt: Optional[float], what: str,
             for_correlation: bool = False):
    """The device live lengths of a call: from `meta` (checked against the
    block's rows and the plan's interval, as the reference's checkShape and
    sameInterval do) or a caller's `row_len` (checked as _check_row_len)."""
    if row_len is None:
        if meta is None:
            raise ValueError(f"{what} needs a ChunkMeta (or row_len)")
        if len(meta) != rows:
            raise ValueError(f"samples have {rows} rows, the meta {len(meta)}")
        if plan_dt is not None and not sameInterval(meta.dtMs, plan_dt):
            raise ValueError(f"the rows' interval {meta.dtMs} ms is not the plan's {plan_dt} ms")
        return meta.device_row_len(cols, device, for_correlation)
    _check_row_len(row_len, rows, device)
    return row_len


class PipelinePlan:
    """A compiled pipeline on the GPU (pipeline.nim:200-211).  Immutable; any
    number of streams may apply it at once."""

    def __init__(self, flow: Pipeline, T, dtMs: float, samples: int, recordSamples: int = 0):
        self.T = np.dtype(T)
        self._h = C.c_void_p()
        arr = _ops_c(flow.ops)
        _lib.check(_lib.lib().pseg_plan_compile(arr, len(flow.ops), _dtype_code(T), dtMs,
                                                samples, recordSamples, C.byref(self._h)))
        info = _lib.PlanInfoC()
        _lib.check(_lib.lib().pseg_plan_info_get(self._h, C.byref(info)))
        self.generation = int(info.generation)
        self.dtMs = float(info.dt_ms)
        self.samples = int(info.samples)
        self.recordSamples = int(info.record_samples)
        self.requiredMargin = int(info.margin)
        self.partCount = int(info.parts)
        self.lds_bytes = int(info.lds_bytes)
        self.threadsPerRow = int(info.threads_per_row)
        n = info.n_steps
        self.steps = [dict(kind=_KIND_NAMES[info.step_kind[i]], nfft=info.step_nfft[i],
                           pad=info.step_pad[i], margin=info.step_margin[i],
                           reflect=bool(info.step_reflect[i])) for i in range(n)]
        # mutes read the rows' trace keys (t0Ms, offset) from the meta
        self.readsKeys = any(op.kind in (opMuteTop, opMuteBottom, opGainCurve) for op in flow.ops)
        # amplitudeUnits (pipeline.nim:571-577): paRecorded, paCalibrated, paRelative
        self.amplitudeUnits = int(info.amplitude_units)
        self.nCalib = int(info.n_calib)
        self._ws = {}

    def __len__(self):
        return len(self.steps)

    def workspaceSize(self, rows: int, cols: int) -> int:
        b = C.c_size_t()
        _lib.check(_lib.lib().pseg_plan_workspace_size(self._h, rows, cols, C.byref(b)))
        return int(b.value)

    def _workspace(self, rows, cols, device):
        import torch
        need = self.workspaceSize(rows, cols)
        if need == 0:
            return None, 0
        key = str(device)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < need:
            ws = torch.empty(need, dtype=torch.uint8, device=device)
            self._ws[key] = ws
        return ws, need

    def calibFactors(self, meta: ChunkMeta) -> np.ndarray:
        """The multipliers of the plan's calibration steps for the rows of
        `meta`, [nCalib, rows] float64 (pseg_plan_calib_factors)."""
        n = len(meta)
        out = np.zeros((max(self.nCalib, 1), max(n, 1)), np.float64)
        cnt = C.c_int32()
        _lib.check(_lib.lib().pseg_plan_calib_factors(
            self._h, meta.calib_rows(), n, int(meta.hasFileScales), out.ctypes.data, None,
            C.byref(cnt)))
        return out[:self.nCalib, :n]

    def apply(self, src, dst=None, meta: Optional[ChunkMeta] = None, *, row_len=None,
              ws=None, stream=None, keys=None, first_sample: int = 0, calib_factors=None):
        """apply(src, dst, meta) (src -> dst) or apply(samples, meta) (in
        place), pipeline.nim:736-772.  `row_len` (a device int32 tensor) may
        replace `meta`; then `keys` ([rows, 2] float64 device tensor of
        t0Ms, offset) and `first_sample` give what mutes and gain curves read
        from a meta.  A plan with calibration steps needs the meta (each
        row's factors resolve from its keys; ``calib_factors``, [nCalib,
        rows], replaces them).
        Asynchronous on `stream` (default: torch's current)."""
        import torch
        if isinstance(dst, ChunkMeta) and meta is None:
            meta, dst = dst, None
        if dst is None:
            dst = src
        _check_panel(src, "src", self.T)
        _check_panel(dst, "dst", self.T, device=src.device)
        if src.shape != dst.shape:
            raise ValueError(f"src is {src.shape[0]} × {src.shape[1]}, dst {dst.shape[0]} × {dst.shape[1]}")
        rows, cols = src.shape
        if row_len is None:
            if meta is None:
                raise ValueError("apply needs a ChunkMeta (or row_len)")
            if len(meta) != rows:
                raise ValueError(f"samples have {rows} rows, the meta {len(meta)}")
            if not sameInterval(meta.dtMs, self.dtMs):
                raise ValueError(f"the plan is for dtMs = {self.dtMs}, the rows have {meta.dtMs}")
            row_len = meta.device_row_len(cols, src.device)
            if self.readsKeys:
                keys = meta.device_keys(src.device)
            first_sample = meta.firstSample
        else:
            _check_row_len(row_len, rows, src.device)
        if keys is not None:
            if (not isinstance(keys, torch.Tensor) or not keys.is_cuda or keys.dtype != torch.float64
                    or tuple(keys.shape) != (rows, 2) or not keys.is_contiguous()):
                raise ValueError("keys must be a contiguous [rows, 2] float64 GPU tensor")
        if ws is None:
            ws, need = self._workspace(rows, cols, src.device)
        else:
            need = ws.numel()
        if self.nCalib > 0:
            if meta is None:
                raise ValueError("a plan with a calibration step needs the ChunkMeta (the rows' "
                                 "keys give their factors)")
            k = (self.calibFactors(meta) if calib_factors is None
                 else np.ascontiguousarray(calib_factors, np.float64))
            if k.shape != (self.nCalib, rows):
                raise ValueError(f"calibration factors must be [{self.nCalib}, {rows}]")
            kd = torch.from_numpy(k.copy()).to(src.device)
            cl = meta.device_row_len(cols, src.device, 2)
            self._hold = (kd, cl)  # alive until the next apply on this plan
            _lib.check(_lib.lib().pseg_plan_apply_calib(
                self._h, src.data_ptr(), dst.data_ptr(), rows, cols, row_len.data_ptr(),
                keys.data_ptr() if keys is not None else None, int(first_sample),
                cl.data_ptr(), kd.data_ptr(), ws.data_ptr() if ws is not None else None, need,
                _stream_handle(stream)))
            return dst
        _lib.check(_lib.lib().pseg_plan_apply_keys(
            self._h, src.data_ptr(), dst.data_ptr(), rows, cols, row_len.data_ptr(),
            keys.data_ptr() if keys is not None else None, int(first_sample),
            ws.data_ptr() if ws is not None else None, need, _stream_handle(stream)))
        return dst

    def applyHost(self, src: np.ndarray, dst: Optional[np.ndarray] = None,
                  meta: Optional[ChunkMeta] = None, stream=None) -> np.ndarray:
        """apply(plan, src: Tensor[T], dst: var Tensor[T], meta, ws) and the
        in-place apply(plan, samples, meta, ws) (pipeline.nim:758-772) on
        host arrays: the block is staged through device memory the library
        keeps, the plan runs there, and the result is in dst when this
        returns.  src, dst: C-contiguous [rows, cols] numpy arrays of the
        plan's type (dst None: in place)."""
        if isinstance(dst, ChunkMeta) and meta is None:
            meta, dst = dst, None
        if dst is None:
            dst = src
        for name, a in (("src", src), ("dst", dst)):
            if not isinstance(a, np.ndarray) or a.ndim != 2 or a.dtype != self.T \
                    or not a.flags.c_contiguous:
                raise ValueError(f"{name} must be a C-contiguous [rows, cols] {self.T} array")
        if not dst.flags.writeable:
            raise ValueError("dst must be writeable")
        if src.shape != dst.shape:
            raise ValueError(f"src is {src.shape[0]} × {src.shape[1]}, dst {dst.shape[0]} × "
                             f"{dst.shape[1]}")
        if meta is None:
            raise ValueError("applyHost needs a ChunkMeta")
        rows, cols = src.shape
        if len(meta) != rows:
            raise ValueError(f"samples have {rows} rows, the meta {len(meta)}")
        if not sameInterval(meta.dtMs, self.dtMs):
            raise ValueError(f"the plan is for dtMs = {self.dtMs}, the rows have {meta.dtMs}")
        rl = meta.row_len(cols)
        keys = np.ascontiguousarray(meta.keys()) if self.readsKeys else None
        _lib.check(_lib.lib().pseg_plan_apply_host(
            self._h, src.ctypes.data, dst.ctypes.data, rows, cols, rl.ctypes.data,
            keys.ctypes.data if keys is not None else None, int(meta.firstSample),
            _stream_handle(stream)))
        return dst

    def applyRaw(self, raw, fmt: int, dst, meta: Optional[ChunkMeta] = None, *, row_len=None,
                 keys=None, first_sample: int = 0, offset: int = 240,
                 pitch: Optional[int] = None, stream=None):
        """apply with the source as raw SEG-Y traces on the device (a uint8
        tensor: a header of `offset` bytes, then big-endian samples of format
        `fmt`, traces `pitch` bytes apart): each row is decoded as the kernel
        loads it, so the traces cross HBM once into `dst` ([rows, cols] of the
        plan's dtype).  As decodeSamples then apply, in one pass."""
        import torch
        from .decode import SAMPLE_BYTES
        nb = SAMPLE_BYTES.get(fmt)
        if nb is None:
            raise ValueError(f"unknown sample format {fmt} (SEG-Y codes 1, 2, 3, 5, 8; SEG-D "
                             "8036, 8038, 8058, 8080)")
        if not isinstance(raw, torch.Tensor) or not raw.is_cuda or raw.dtype != torch.uint8 \
                or not raw.is_contiguous():
            raise ValueError("raw must be a contiguous uint8 GPU tensor")
        if not isinstance(dst, torch.Tensor) or not dst.is_cuda or dst.dim() != 2 \
                or not dst.is_contiguous():
            raise ValueError("dst must be a C-contiguous [rows, cols] GPU tensor")
        if np.dtype(str(dst.dtype).replace("torch.", "")) != self.T:
            raise ValueError(f"dst holds {dst.dtype}, the plan {self.T}")
        rows, cols = dst.shape
        if raw.device != dst.device:
            raise ValueError(f"raw is on {raw.device}, dst on {dst.device}")
        if pitch is None:
            pitch = offset + nb * cols
        if rows > 0 and (rows - 1) * pitch + offset + nb * cols > raw.numel():
            raise ValueError(f"{rows} traces of pitch {pitch} do not fit {raw.numel()} bytes")
        if row_len is None:
            if meta is None:
                raise ValueError("applyRaw needs a ChunkMeta (or row_len)")
            if len(meta) != rows:
                raise ValueError(f"dst has {rows} rows, the meta {len(meta)}")
            if not sameInterval(meta.dtMs, self.dtMs):
                raise ValueError(f"the plan is for dtMs = {self.dtMs}, the rows have {meta.dtMs}")
            row_len = meta.device_row_len(cols, dst.device)
            if self.readsKeys:
                keys = meta.device_keys(dst.device)
            first_sample = meta.firstSample
        else:
            _check_row_len(row_len, rows, dst.device)
        if keys is not None:
            if (not isinstance(keys, torch.Tensor) or not keys.is_cuda or keys.dtype != torch.float64
                    or tuple(keys.shape) != (rows, 2) or not keys.is_contiguous()):
                raise ValueError("keys must be a contiguous [rows, 2] float64 GPU tensor")
        _lib.check(_lib.lib().pseg_plan_apply_raw(
            self._h, fmt, raw.data_ptr(), pitch, offset, dst.data_ptr(), rows, cols,
            row_len.data_ptr(), keys.data_ptr() if keys is not None else None, int(first_sample),
            _stream_handle(stream)))
        return dst

    def __del__(self):
        try:
            if self._h:
                _lib.lib().pseg_plan_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass
