"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a C restatement of processSEG's panel hot path (oracle.c,
or_tmpl.h).  Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg
of bench.py use it, as the checker; the product (processseg_amd) never imports
it.  Functions take numpy arrays and op lists built with processseg_amd's op
builders (any object with the same attribute names works).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAX_NOTCH = 8


class OrSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("zero_dc", C.c_int32), ("n_notch", C.c_int32),
                ("reserved", C.c_int32),
                ("f1", C.c_double), ("f2", C.c_double), ("f3", C.c_double), ("f4", C.c_double),
                ("low_hz", C.c_double), ("low_slope", C.c_double),
                ("high_hz", C.c_double), ("high_slope", C.c_double),
                ("notch_hz", C.c_double * MAX_NOTCH), ("notch_width", C.c_double * MAX_NOTCH)]


class OrOp(C.Structure):
    _fields_ = [("kind", C.c_int32), ("trend", C.c_int32), ("mode", C.c_int32),
                ("stat", C.c_int32), ("scope", C.c_int32), ("reserved", C.c_int32),
                ("value", C.c_double), ("lo", C.c_double), ("hi", C.c_double),
                ("window_ms", C.c_double), ("target", C.c_double), ("epsilon", C.c_double),
                ("percentile", C.c_double), ("spec", OrSpec),
                ("mute_kind", C.c_int32), ("mute_knots", C.c_int32),
                ("mute_t_ms", C.c_double), ("mute_ms_per_dist", C.c_double),
                ("mute_pad_ms", C.c_double), ("taper_ms", C.c_double),
                ("mute_offsets", C.c_double * 32), ("mute_times", C.c_double * 32),
                ("gain_kind", C.c_int32), ("gain_knots", C.c_int32),
                ("gain_knot_ms", C.POINTER(C.c_double)), ("gain_knot_db", C.POINTER(C.c_double)),
                ("gain_power", C.c_double), ("gain_ref_ms", C.c_double),
                ("gain_min_ms", C.c_double), ("gain_db_per_s", C.c_double),
                ("gain_tmin_ms", C.c_double), ("gain_tmax_ms", C.c_double),
                ("iir_phase", C.c_int32), ("reserved2", C.c_int32)]


class OrBiquad(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("b0", "b1", "b2", "a1", "a2", "g")]


class OrIir(C.Structure):
    _fields_ = [("n_sections", C.c_int32), ("phase", C.c_int32), ("trend", C.c_int32),
                ("reflect", C.c_int32), ("pad", C.c_int64), ("lead", C.c_int64),
                ("tail", C.c_int64), ("samples", C.c_int64), ("dt_ms", C.c_double),
                ("dc_gain", C.c_double), ("dc_slope", C.c_double), ("sec", OrBiquad * 32)]


class OrFilterInfo(C.Structure):
    _fields_ = [("margin", C.c_int64), ("pad", C.c_int64), ("nfft", C.c_int64),
                ("reflect", C.c_int32), ("zero_dc", C.c_int32), ("dc_gain", C.c_double)]


class OrPlanInfo(C.Structure):
    _fields_ = [("margin", C.c_int64), ("parts", C.c_int32), ("n_steps", C.c_int32),
                ("nfft", C.c_int64 * 16), ("half", C.c_int64 * 16)]


class OrCorrInfo(C.Structure):
    _fields_ = [("nfft", C.c_int64), ("shift", C.c_int64), ("out_samples", C.c_int64),
                ("pilot_samples", C.c_int64), ("record_samples", C.c_int64),
                ("scale", C.c_double)]


class OrFbParams(C.Structure):
    _fields_ = [("method", C.c_int32), ("reserved", C.c_int32)] + [
        (k, C.c_double) for k in ("window_ms", "lta_ms", "stabilization", "low_cut_hz",
                                  "high_cut_hz", "slope_db_per_octave", "search_start_ms",
                                  "search_end_ms", "guide_velocity", "guide_t0_ms", "before_ms",
                                  "after_ms", "snr_window_ms", "low_db", "high_db")]


class OrFbPlan(C.Structure):
    _fields_ = [("p", OrFbParams), ("dt_ms", C.c_double), ("samples", C.c_int64),
                ("w", C.c_int64), ("lta", C.c_int64), ("snr", C.c_int64), ("before", C.c_int64),
                ("after", C.c_int64), ("filtered", C.c_int32), ("whole", C.c_int32),
                ("filter", OrIir)]


class OracleError(ValueError):
    pass


def build(force: bool = False) -> str:
    """Compile liboracle.so with make (gcc)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE],
                       check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
        E = [C.c_char_p, C.c_int]
        L.or_is_fast_len.argtypes = [C.c_int64]
        L.or_mute_time_ms.argtypes = [C.POINTER(OrOp), C.c_double]
        L.or_mute_time_ms.restype = C.c_double
        L.or_fb_init.argtypes = [C.POINTER(OrFbPlan), C.POINTER(OrFbParams), C.c_double,
                                 C.c_int64, C.c_char_p, C.c_int]
        L.or_fb_required_margin.argtypes = [C.POINTER(OrFbPlan)]
        L.or_fb_required_margin.restype = C.c_int64
        for sfx in ("f32", "f64"):
            getattr(L, f"or_fb_pick_{sfx}").argtypes = [
                C.POINTER(OrFbPlan), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_bin_of.argtypes = [C.POINTER(OrBinning), C.c_void_p]
        L.or_bin_of.restype = C.c_int64
        L.or_bin_at.argtypes = [C.POINTER(OrBinning), C.c_double, C.c_double]
        L.or_bin_at.restype = C.c_int64
        L.or_fold_add.argtypes = [C.POINTER(OrBinning), C.c_void_p, C.c_int32, C.c_int32,
                                  C.c_double, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32] + \
            [C.c_void_p] * 6
        L.or_iir_init.argtypes = [C.POINTER(OrIir), C.POINTER(OrSpec), C.c_double, C.c_int64,
                                  C.c_int32, C.c_int32, C.c_char_p, C.c_int]
        L.or_iir_magnitude_at.argtypes = [C.POINTER(OrIir), C.c_double]
        L.or_iir_magnitude_at.restype = C.c_double
        L.or_iir_order.argtypes = [C.POINTER(OrIir)]
        L.or_iir_order.restype = C.c_int32
        for fn in ("or_gain_db_at", "or_gain_factor_at"):
            getattr(L, fn).argtypes = [C.POINTER(OrOp), C.c_double]
            getattr(L, fn).restype = C.c_double
        L.or_next_fast_len.argtypes = [C.c_int64, C.c_int]
        L.or_next_fast_len.restype = C.c_int64
        L.or_factorize.argtypes = [C.c_int64, i32p]
        for t, p in (("f32", f32p), ("f64", f64p)):
            getattr(L, f"or_rfft_{t}").argtypes = [p, C.c_int64, p, p]
            getattr(L, f"or_irfft_{t}").argtypes = [p, p, C.c_int64, p]
            getattr(L, f"or_demean_{t}").argtypes = [p, C.c_int64]
            getattr(L, f"or_agc_{t}").argtypes = [p, C.c_int64, C.c_double, C.c_double,
                                                 C.c_int32, C.c_double, C.c_double]
            getattr(L, f"or_percentile_abs_{t}").argtypes = [p, C.c_int64, C.c_double]
            getattr(L, f"or_percentile_abs_{t}").restype = C.c_double
            getattr(L, f"or_panel_stats_{t}").argtypes = [p, C.c_int32, C.c_int32, i32p,
                                                         C.c_double, f64p]
            getattr(L, f"or_filter_{t}").argtypes = [p, C.c_int64, C.POINTER(OrSpec),
                                                    C.c_double, C.c_int64, C.c_int32] + E
            getattr(L, f"or_pipeline_apply_{t}").argtypes = [
                C.POINTER(OrOp), C.c_int32, C.c_double, C.c_int64, C.c_int64, p, p,
                C.c_int32, C.c_int32, i32p] + E
            getattr(L, f"or_pipeline_apply_keys_{t}").argtypes = [
                C.POINTER(OrOp), C.c_int32, C.c_double, C.c_int64, C.c_int64, p, p,
                C.c_int32, C.c_int32, i32p, C.c_void_p, C.c_int64] + E
            getattr(L, f"or_mute_top_{t}").argtypes = [p, C.c_int64, C.c_double, C.c_double]
            getattr(L, f"or_iir_apply_{t}").argtypes = [C.POINTER(OrIir), p, C.c_int64]
            getattr(L, f"or_gain_curve_apply_{t}").argtypes = [
                C.POINTER(OrOp), C.c_double, C.c_int64, p, C.c_int64, C.c_double,
                C.POINTER(C.c_int64), C.c_char_p, C.c_int]
            getattr(L, f"or_mute_bottom_{t}").argtypes = [p, C.c_int64, C.c_double, C.c_double]
            getattr(L, f"or_correlate_{t}").argtypes = [
                f64p, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_int64,
                C.c_int32, C.c_double, C.c_int32, p, C.c_int32, C.c_int32, i32p, p,
                C.c_int32] + E
        L.or_response.argtypes = [C.POINTER(OrSpec), C.c_double, C.c_double]
        L.or_response.restype = C.c_double
        L.or_narrowest_ramp_hz.argtypes = [C.POINTER(OrSpec)]
        L.or_narrowest_ramp_hz.restype = C.c_double
        L.or_impulse_length.argtypes = [C.POINTER(OrSpec), C.c_double,
                                        C.POINTER(C.c_int64)] + E
        L.or_filter_info_of.argtypes = [C.POINTER(OrSpec), C.c_double, C.c_int64, C.c_int32,
                                        C.POINTER(OrFilterInfo)] + E
        L.or_meta_flatten.argtypes = [C.c_int32, C.c_int32, C.c_int64, i64p, u8p, i64p,
                                      C.c_int32, C.c_int, i32p]
        L.or_plan_info_of.argtypes = [C.POINTER(OrOp), C.c_int32, C.c_int32, C.c_double,
                                      C.c_int64, C.c_int64, C.POINTER(OrPlanInfo)] + E
        L.or_corr_info_of.argtypes = [f64p, C.c_int64, C.c_double, C.c_double, C.c_double,
                                      C.c_int64, C.c_int64, C.c_int32, C.c_double, C.c_int32,
                                      C.POINTER(OrCorrInfo)] + E
        L.or_samples_in_length.argtypes = [C.c_double, C.c_double]
        L.or_samples_in_length.restype = C.c_int64
        L.or_linear_sweep.argtypes = [C.c_double] * 8 + [f64p, C.c_int64] + E
        L.or_pool_create.argtypes = [C.c_int32]
        L.or_pool_create.restype = C.c_void_p
        L.or_pool_destroy.argtypes = [C.c_void_p]
        L.or_pool_apply_f32.argtypes = [C.c_void_p, C.POINTER(OrOp), C.c_int32, C.c_double,
                                        C.c_int64, f32p, f32p, C.c_int32, C.c_int32, i32p,
                                        C.c_int32] + E
        _lib = L
    return _lib


def _err():
    return C.create_string_buffer(512)


def _check(rc, err):
    if rc != 0:
        raise OracleError(err.value.decode())


def _sfx(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(f"sample type must be float32 or float64, got {dt}")


# ------------------------------------------------------------------ specs --

def spec_struct(spec) -> OrSpec:
    """An OrSpec from a filter spec object (processseg_amd.FilterSpec)."""
    s = OrSpec()
    s.kind = int(spec.kind)
    s.zero_dc = int(bool(spec.zero_dc))
    notches = list(spec.notches)
    if len(notches) > MAX_NOTCH:
        raise ValueError("too many notches")
    s.n_notch = len(notches)
    s.f1, s.f2, s.f3, s.f4 = spec.f1, spec.f2, spec.f3, spec.f4
    s.low_hz, s.low_slope = spec.low_hz, spec.low_slope
    s.high_hz, s.high_slope = spec.high_hz, spec.high_slope
    for i, (f, w) in enumerate(notches):
        s.notch_hz[i] = f
        s.notch_width[i] = w
    return s


def op_struct(op) -> OrOp:
    o = OrOp()
    o.kind = int(op.kind)
    o.trend = int(getattr(op, "trend", 2))
    o.mode = int(getattr(op, "mode", 0))
    o.stat = int(getattr(op, "stat", 0))
    o.scope = int(getattr(op, "scope", 0))
    o.value = float(getattr(op, "value", 0.0))
    o.lo = float(getattr(op, "lo", 0.0))
    o.hi = float(getattr(op, "hi", 0.0))
    o.window_ms = float(getattr(op, "window_ms", 0.0))
    o.target = float(getattr(op, "target", 1.0))
    o.epsilon = float(getattr(op, "epsilon", 1e-9))
    o.percentile = float(getattr(op, "percentile", 99.0))
    spec = getattr(op, "spec", None)
    if spec is not None:
        o.spec = spec_struct(spec)
    mute = getattr(op, "mute", None)
    if mute is not None:
        o.mute_kind = int(mute.kind)
        o.mute_knots = len(mute.offsets)
        o.mute_t_ms = float(mute.t0Ms if mute.kind == 2 else mute.tMs)
        o.mute_ms_per_dist = float(mute.msPerDist)
        o.mute_pad_ms = float(mute.padMs)
        o.taper_ms = float(getattr(op, "taper_ms", 0.0))
        for i, (a, b) in enumerate(zip(mute.offsets, mute.timesMs)):
            o.mute_offsets[i] = a
            o.mute_times[i] = b
    o.iir_phase = int(getattr(op, "phase", 0))
    curve = getattr(op, "curve", None)
    if curve is not None:
        gain_fields(o, curve, getattr(op, "t_min_ms", 0.0), getattr(op, "t_max_ms", math.nan))
    return o


def gain_fields(o, curve, t_min_ms: float = 0.0, t_max_ms: float = math.nan):
    """A gain curve (kind 0 knots, 1 power, 2 exponential) into the C op; the
    struct keeps the knot arrays alive."""
    o.gain_kind = int(curve.kind)
    n = len(curve.knots)
    o.gain_knots = n
    if n:
        o.gain_knot_ms = (C.c_double * n)(*[k[0] for k in curve.knots])
        o.gain_knot_db = (C.c_double * n)(*[k[1] for k in curve.knots])
    o.gain_power = float(curve.power)
    o.gain_ref_ms = float(curve.refMs)
    o.gain_min_ms = float(curve.minMs)
    o.gain_db_per_s = float(curve.dbPerS)
    o.gain_tmin_ms = float(t_min_ms)
    o.gain_tmax_ms = float(t_max_ms)


def _ops_array(ops):
    arr = (OrOp * max(1, len(ops)))()
    for i, op in enumerate(ops):
        arr[i] = op_struct(op)
    return arr


# ----------------------------------------------------------------- lengths --

def is_fast_len(n: int) -> bool:
    return bool(lib().or_is_fast_len(n))


def next_fast_len(n: int, even: bool = True) -> int:
    return int(lib().or_next_fast_len(n, 1 if even else 0))


def factorize(n: int) -> list:
    r = np.zeros(64, np.int32)
    c = lib().or_factorize(n, r)
    return [int(v) for v in r[:c]]


# --------------------------------------------------------------------- FFT --

def rfft(x: np.ndarray):
    x = np.ascontiguousarray(x)
    n = x.shape[0]
    re = np.zeros(n // 2 + 1, x.dtype)
    im = np.zeros(n // 2 + 1, x.dtype)
    if getattr(lib(), f"or_rfft_{_sfx(x.dtype)}")(x, n, re, im):
        raise OracleError(f"{n} is not a 2·3·5-smooth length")
    return re, im


def irfft(re: np.ndarray, im: np.ndarray, n: int):
    re = np.ascontiguousarray(re)
    im = np.ascontiguousarray(im, dtype=re.dtype)
    x = np.zeros(n, re.dtype)
    if getattr(lib(), f"or_irfft_{_sfx(re.dtype)}")(re, im, n, x):
        raise OracleError(f"{n} is not a 2·3·5-smooth length")
    return x


# ------------------------------------------------------------------ design --

def response(spec, f: float, dt_ms: float) -> float:
    s = spec_struct(spec)
    return float(lib().or_response(C.byref(s), f, dt_ms))


def narrowest_ramp_hz(spec) -> float:
    s = spec_struct(spec)
    return float(lib().or_narrowest_ramp_hz(C.byref(s)))


def impulse_length(spec, dt_ms: float) -> int:
    s = spec_struct(spec)
    L = C.c_int64()
    e = _err()
    _check(lib().or_impulse_length(C.byref(s), dt_ms, C.byref(L), e, 512), e)
    return int(L.value)


def filter_info(spec, dt_ms: float, samples: int, trend: int = 2) -> dict:
    s = spec_struct(spec)
    out = OrFilterInfo()
    e = _err()
    _check(lib().or_filter_info_of(C.byref(s), dt_ms, samples, trend, C.byref(out), e, 512), e)
    return dict(margin=out.margin, pad=out.pad, nfft=out.nfft, reflect=bool(out.reflect),
                zero_dc=bool(out.zero_dc), dc_gain=out.dc_gain)


# -------------------------------------------------------------- kernels --

def filter_trace(x: np.ndarray, spec, dt_ms: float, samples: int, trend: int = 2):
    y = np.ascontiguousarray(x).copy()
    s = spec_struct(spec)
    e = _err()
    _check(getattr(lib(), f"or_filter_{_sfx(y.dtype)}")(y, y.shape[0], C.byref(s), dt_ms,
                                                         samples, trend, e, 512), e)
    return y


def demean(x: np.ndarray):
    y = np.ascontiguousarray(x).copy()
    getattr(lib(), f"or_demean_{_sfx(y.dtype)}")(y, y.shape[0])
    return y


def agc(x: np.ndarray, window_ms: float, dt_ms: float, mode: int = 0, target: float = 1.0,
        epsilon: float = 1e-9):
    y = np.ascontiguousarray(x).copy()
    rc = getattr(lib(), f"or_agc_{_sfx(y.dtype)}")(y, y.shape[0], window_ms, dt_ms, mode,
                                                    target, epsilon)
    if rc:
        raise OracleError("invalid AGC parameters")
    return y


def percentile_abs(x: np.ndarray, p: float) -> float:
    x = np.ascontiguousarray(x)
    return float(getattr(lib(), f"or_percentile_abs_{_sfx(x.dtype)}")(x, x.shape[0], p))


def panel_stats(x: np.ndarray, row_len: np.ndarray, clip_percentile: float = 99.0) -> dict:
    x = np.ascontiguousarray(x)
    out = np.zeros(5)
    getattr(lib(), f"or_panel_stats_{_sfx(x.dtype)}")(
        x, x.shape[0], x.shape[1], np.ascontiguousarray(row_len, np.int32), clip_percentile, out)
    return dict(rms=out[0], max_abs=out[1], clip=out[2], rows=int(out[3]), samples=int(out[4]))


# ---------------------------------------------------------------- meta --

def meta_flatten(cols: int, num_samples, status, failed=(), first_sample: int = 0,
                 used_for_corr: bool = False) -> np.ndarray:
    ns = np.ascontiguousarray(num_samples, np.int64)
    st = np.ascontiguousarray(status, np.uint8)
    fl = np.ascontiguousarray(np.array(sorted(failed), np.int64).reshape(-1))
    if fl.size == 0:
        fl = np.zeros(1, np.int64)
        nfl = 0
    else:
        nfl = int(fl.size)
    out = np.zeros(ns.shape[0], np.int32)
    lib().or_meta_flatten(ns.shape[0], cols, first_sample, ns, st, fl, nfl,
                          1 if used_for_corr else 0, out)
    return out


# ------------------------------------------------------------ pipeline --

def plan_info(ops, dtype, dt_ms: float, samples: int, record_samples: int = 0) -> dict:
    arr = _ops_array(ops)
    out = OrPlanInfo()
    e = _err()
    dt = 0 if np.dtype(dtype) == np.float32 else 1
    _check(lib().or_plan_info_of(arr, len(ops), dt, dt_ms, samples, record_samples,
                                 C.byref(out), e, 512), e)
    return dict(margin=out.margin, parts=out.parts, n_steps=out.n_steps,
                nfft=[out.nfft[i] for i in range(out.n_steps)],
                half=[out.half[i] for i in range(out.n_steps)])


def pipeline_apply(ops, src: np.ndarray, row_len, dt_ms: float, samples: int | None = None,
                   record_samples: int = 0, dst: np.ndarray | None = None, keys=None,
                   first_sample: int = 0) -> np.ndarray:
    """compile + apply: src (rows x cols, float32 or float64) -> a new array.
    keys: (t0Ms, offset) per row for mutes (None: t0Ms 0, offset NaN);
    first_sample: ChunkMeta.firstSample."""
    src = np.ascontiguousarray(src)
    rows, cols = src.shape
    out = np.empty_like(src) if dst is None else dst
    arr = _ops_array(ops)
    e = _err()
    rl = np.ascontiguousarray(row_len, np.int32)
    k = None if keys is None else np.ascontiguousarray(keys, np.float64).reshape(rows, 2)
    _check(getattr(lib(), f"or_pipeline_apply_keys_{_sfx(src.dtype)}")(
        arr, len(ops), dt_ms, samples if samples is not None else cols, record_samples,
        src, out, rows, cols, rl, None if k is None else k.ctypes.data, first_sample, e, 512), e)
    return out


def mute_time_ms(op, offset: float) -> float:
    """muteTimeMs of a mute op (core/mute.nim:100-118)."""
    return float(lib().or_mute_time_ms(C.byref(op_struct(op)), offset))


class IirPlan:
    """initIirPlan (filter/iir.nim:193-239) through the C restatement:
    order, phase, requiredMargin, magnitudeAt and applyFilter on one trace
    (a copy).  ValueError as the plan raises."""

    def __init__(self, spec, dt_ms: float, samples: int, phase: int = 0, trend: int = 2):
        self.p = OrIir()
        e = _err()
        _check(lib().or_iir_init(C.byref(self.p), C.byref(spec_struct(spec)), dt_ms, samples,
                                 phase, trend, e, 512), e)

    @property
    def order(self) -> int:
        return int(lib().or_iir_order(C.byref(self.p)))

    @property
    def requiredMargin(self) -> int:
        return int(self.p.pad)

    def magnitudeAt(self, f: float) -> float:
        return float(lib().or_iir_magnitude_at(C.byref(self.p), f))

    def filtered(self, x: np.ndarray) -> np.ndarray:
        y = np.ascontiguousarray(x).copy()
        if y.shape[0] > self.p.samples:
            raise ValueError("a trace longer than the plan's samples")
        getattr(lib(), f"or_iir_apply_{_sfx(y.dtype)}")(C.byref(self.p), y, y.shape[0])
        return y


def gain_db_at(curve, t_ms: float) -> float:
    """gainDbAt (gain/curves.nim:134-153)."""
    o = OrOp()
    gain_fields(o, curve)
    return float(lib().or_gain_db_at(C.byref(o), t_ms))


def gain_factor_at(curve, t_ms: float) -> float:
    """factorAt (gain/curves.nim:154-160)."""
    o = OrOp()
    gain_fields(o, curve)
    return float(lib().or_gain_factor_at(C.byref(o), t_ms))


def gain_curve_apply(curve, x: np.ndarray, dt_ms: float, samples: int, start_ms: float,
                     t_min_ms: float = 0.0, t_max_ms: float = math.nan):
    """initGainCurvePlan + the pipeline's applyGainCurve on one trace (on a
    copy): returns (trace, table length).  ValueError as the plan raises."""
    o = OrOp()
    gain_fields(o, curve, t_min_ms, t_max_ms)
    y = np.ascontiguousarray(x).copy()
    tl = C.c_int64(0)
    e = _err()
    _check(getattr(lib(), f"or_gain_curve_apply_{_sfx(y.dtype)}")(
        C.byref(o), dt_ms, samples, y, y.shape[0], start_ms, C.byref(tl), e, 512), e)
    return y, int(tl.value)


def mute_row(kind: str, x: np.ndarray, col: float, taper_cols: float) -> np.ndarray:
    """muteTop / muteBottom of one row in column units, on a copy."""
    y = np.ascontiguousarray(x).copy()
    getattr(lib(), f"or_mute_{kind}_{_sfx(y.dtype)}")(y, y.shape[0], col, taper_cols)
    return y


class Pool:
    """The ProcessPool clone: persistent worker threads."""

    def __init__(self, threads: int):
        self.threads = int(threads)
        self._p = lib().or_pool_create(self.threads)

    def apply_f32(self, ops, src, dst, row_len, dt_ms, samples=None, repeats: int = 1):
        arr = _ops_array(ops)
        e = _err()
        rows, cols = src.shape
        _check(lib().or_pool_apply_f32(self._p, arr, len(ops), dt_ms,
                                       samples if samples is not None else cols, src, dst,
                                       rows, cols, np.ascontiguousarray(row_len, np.int32),
                                       repeats, e, 512), e)
        return dst

    def close(self):
        if self._p:
            lib().or_pool_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------- correlation --

def corr_info(pilot, dt_ms, record_samples, out_samples=0, mode=0, eps_percent=1.0,
              normalize=True, pilot_dt_ms=None, pilot_t0_ms=0.0) -> dict:
    p = np.ascontiguousarray(pilot, np.float64)
    out = OrCorrInfo()
    e = _err()
    _check(lib().or_corr_info_of(p, p.shape[0], dt_ms if pilot_dt_ms is None else pilot_dt_ms,
                                 pilot_t0_ms, dt_ms, record_samples, out_samples, mode,
                                 eps_percent, 1 if normalize else 0, C.byref(out), e, 512), e)
    return dict(nfft=out.nfft, shift=out.shift, out_samples=out.out_samples,
                pilot_samples=out.pilot_samples, record_samples=out.record_samples,
                scale=out.scale)


def correlate(src: np.ndarray, row_len, pilot, dt_ms, record_samples, out_samples=0, mode=0,
              eps_percent=1.0, normalize=True, pilot_dt_ms=None, pilot_t0_ms=0.0):
    src = np.ascontiguousarray(src)
    info = corr_info(pilot, dt_ms, record_samples, out_samples, mode, eps_percent, normalize,
                     pilot_dt_ms, pilot_t0_ms)
    rows, cols = src.shape
    dst = np.zeros((rows, info["out_samples"]), src.dtype)
    p = np.ascontiguousarray(pilot, np.float64)
    e = _err()
    _check(getattr(lib(), f"or_correlate_{_sfx(src.dtype)}")(
        p, p.shape[0], dt_ms if pilot_dt_ms is None else pilot_dt_ms, pilot_t0_ms, dt_ms,
        record_samples, out_samples, mode, eps_percent, 1 if normalize else 0, src, rows, cols,
        np.ascontiguousarray(row_len, np.int32), dst, dst.shape[1], e, 512), e)
    return dst


# ----------------------------------------------------------- first breaks --

_FB_FIELDS = (("window_ms", "windowMs"), ("lta_ms", "ltaMs"), ("stabilization", "stabilization"),
              ("low_cut_hz", "lowCutHz"), ("high_cut_hz", "highCutHz"),
              ("slope_db_per_octave", "slopeDbPerOctave"), ("search_start_ms", "searchStartMs"),
              ("search_end_ms", "searchEndMs"), ("guide_velocity", "guideVelocity"),
              ("guide_t0_ms", "guideT0Ms"), ("before_ms", "beforeMs"), ("after_ms", "afterMs"),
              ("snr_window_ms", "snrWindowMs"), ("low_db", "lowDb"), ("high_db", "highDb"))


def fb_plan(params, dt_ms: float, samples: int) -> OrFbPlan:
    """initFirstBreakPlan (qc/firstbreak.nim:195-251); params carries the
    FirstBreakParams fields under the reference's names (pickMethod 0
    pmStaLta, 1 pmEnergyRatio, 2 pmCoppens)."""
    fp = OrFbParams()
    fp.method = int(params.pickMethod)
    for c_name, name in _FB_FIELDS:
        setattr(fp, c_name, float(getattr(params, name)))
    plan = OrFbPlan()
    e = _err()
    _check(lib().or_fb_init(C.byref(plan), C.byref(fp), dt_ms, samples, e, 512), e)
    return plan


def fb_required_margin(plan: OrFbPlan) -> int:
    return int(lib().or_fb_required_margin(C.byref(plan)))


def fb_pick(plan: OrFbPlan, x: np.ndarray, row_len, keys=None, first_sample: int = 0):
    """pickInto (qc/firstbreak.nim:492-541) on the rows of x [rows, cols]:
    row_len[r] < 0 a row not live; keys [rows, 2] (t0Ms, offset) or None.
    Returns (time_ms, col, snr_db, confidence) arrays."""
    x = np.ascontiguousarray(x)
    rows, cols = x.shape
    rl = np.ascontiguousarray(row_len, np.int32)
    k = None if keys is None else np.ascontiguousarray(keys, np.float64)
    t = np.empty(rows)
    c = np.empty(rows, np.int64)
    snr = np.empty(rows)
    conf = np.empty(rows)
    getattr(lib(), f"or_fb_pick_{_sfx(x.dtype)}")(
        C.byref(plan), x.ctypes.data, rows, cols, rl.ctypes.data,
        None if k is None else k.ctypes.data, int(first_sample), t.ctypes.data, c.ctypes.data,
        snr.ctypes.data, conf.ctypes.data)
    return t, c, snr, conf


# ---------------------------------------------------------- fold maps --

class OrBinning(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("inlines", C.c_int64),
                ("crosslines", C.c_int64)] + [
        (k, C.c_double) for k in ("first0", "step0", "first1", "step1", "x0", "y0", "ux", "uy",
                                  "vx", "vy", "size0", "size1", "max_off")]


def _binning(b) -> OrBinning:
    o = OrBinning()
    o.kind, o.inlines, o.crosslines = int(b.kind), int(b.inlines), int(b.crosslines)
    for f in ("first0", "step0", "first1", "step1", "x0", "y0", "ux", "uy", "vx", "vy", "size0",
              "size1"):
        setattr(o, f, float(getattr(b, f)))
    o.max_off = float(b.maxOff)
    return o


def bin_of(binning, keys: np.ndarray) -> np.ndarray:
    """binOf (stack/binning.nim:228-237) of each row of a [rows, 10] key table."""
    k = np.ascontiguousarray(keys, np.float64)
    b = _binning(binning)
    L = lib()
    return np.array([L.or_bin_of(C.byref(b), k[r].ctypes.data) for r in range(k.shape[0])],
                    np.int64)


class OrFold:
    """A fold map on the host (qc/fold.nim:293-303) counted by the oracle."""

    def __init__(self, binning, opts):
        n, m = int(binning.inlines), int(binning.crosslines)
        self.binning = binning
        self.bounds = np.ascontiguousarray(tuple(opts.offsetBounds), np.float64)
        self.sectors = int(opts.azimuthSectors)
        self.start = float(opts.sectorStartDeg)
        self.offset_range = bool(opts.offsetRange)
        classes = max(len(self.bounds) - 1, 0)
        self.fold = np.zeros((n, m), np.int32)
        self.offsetFold = np.zeros((classes, n, m) if classes else (0,), np.int32)
        self.azimuthFold = np.zeros((self.sectors, n, m) if self.sectors else (0,), np.int32)
        self.offsetMin = np.full((n, m) if self.offset_range else (0,), np.nan, np.float32)
        self.offsetMax = self.offsetMin.copy()
        self.counters = np.zeros(4, np.int64)

    def add(self, keys: np.ndarray, row_len) -> "OrFold":
        """addFold (qc/fold.nim:583-588): rows with row_len < 0 are not live."""
        k = np.ascontiguousarray(keys, np.float64)
        rl = np.ascontiguousarray(row_len, np.int32)
        b = _binning(self.binning)
        ptr = lambda a: a.ctypes.data if a.size else None  # noqa: E731
        lib().or_fold_add(C.byref(b), ptr(self.bounds), len(self.bounds), self.sectors,
                          C.c_double(self.start), 1 if self.offset_range else 0, ptr(k),
                          rl.ctypes.data, k.shape[0], self.fold.ctypes.data,
                          ptr(self.offsetFold), ptr(self.azimuthFold), ptr(self.offsetMin),
                          ptr(self.offsetMax), self.counters.ctypes.data)
        return self

    binned = property(lambda self: int(self.counters[0]))
    outside = property(lambda self: int(self.counters[1]))
    noCoordinates = property(lambda self: int(self.counters[2]))
    notLive = property(lambda self: int(self.counters[3]))


# ------------------------------------------------------------- sweeps --

def samples_in_length(length_ms: float, dt_ms: float) -> int:
    return int(lib().or_samples_in_length(length_ms, dt_ms))


def linear_sweep(f1, f2, length_ms, dt_ms, taper_start_ms=0.0, taper_end_ms=0.0,
                 start_phase_deg=0.0, amplitude=1.0) -> np.ndarray:
    n = samples_in_length(length_ms, dt_ms)
    x = np.zeros(max(n, 1), np.float64)
    e = _err()
    _check(lib().or_linear_sweep(f1, f2, length_ms, dt_ms, taper_start_ms, taper_end_ms,
                                 start_phase_deg, amplitude, x, n, e, 512), e)
    return x[:n]


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = math.sqrt(float(np.sum((a - b) ** 2)))
    r = math.sqrt(float(np.sum(b ** 2)))
    return d / r if r > 0 else d


# ---------------------------------------------------------- sample decode --

def decode_samples(raw: np.ndarray, fmt: int, samples: int, rows: int, offset: int = 240,
                   pitch: int | None = None, dtype=np.float32) -> np.ndarray:
    """SEG-Y sample decode (SEG-Y rev 2 data sample format codes, big-endian
    words), restated from the standard's definitions in numpy: IBM float32
    (-1)^s · f/2^24 · 16^(e−64), exact in float64 and rounded once to
    float32; int32, int16, int8 two's complement; IEEE float32 as is.  SEG-D
    demultiplexed codes: 8036 24-bit two's complement, 8038 int32, 8058 IEEE
    float32, 8080 IEEE float64 (rounded once to float32)."""
    nb = {1: 4, 2: 4, 3: 2, 5: 4, 8: 1, 8036: 3, 8038: 4, 8058: 4, 8080: 8}[fmt]
    if pitch is None:
        pitch = offset + nb * samples
    raw = np.ascontiguousarray(raw, np.uint8)
    out = np.empty((rows, samples), dtype)
    for r in range(rows):
        b = raw[r * pitch + offset: r * pitch + offset + nb * samples]
        if fmt == 1:
            w = b.view(">u4").astype(np.uint64)
            frac = (w & 0xFFFFFF).astype(np.float64)
            expo = ((w >> 24) & 0x7F).astype(np.int64)
            mag = np.ldexp(frac, (4 * expo - 280).astype(np.int32))
            with np.errstate(over="ignore"):
                out[r] = np.where((w >> 31) == 1, -mag, mag).astype(dtype)
        elif fmt in (2, 8038):
            out[r] = b.view(">i4").astype(dtype)
        elif fmt == 8036:
            t = b.reshape(-1, 3).astype(np.int64)
            u = (t[:, 0] << 16) | (t[:, 1] << 8) | t[:, 2]
            out[r] = np.where(u >= 1 << 23, u - (1 << 24), u).astype(dtype)
        elif fmt == 8058:
            with np.errstate(invalid="ignore"):
                out[r] = b.view(">f4").astype(dtype)
        elif fmt == 8080:
            with np.errstate(invalid="ignore", over="ignore"):
                out[r] = b.view(">f8").astype(dtype)
        elif fmt == 3:
            out[r] = b.view(">i2").astype(dtype)
        elif fmt == 5:
            with np.errstate(invalid="ignore"):
                out[r] = b.view(">f4").astype(dtype)
        else:
            out[r] = b.view("i1").astype(dtype)
    return out


# --------------------------------------------------------------- spectrum --

def _round_half_away(x: float) -> int:
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def spectrum_plan(dt_ms: float, window_len: int, taper_fraction: float = 0.1, nfft: int = 0,
                  demean: bool = True) -> dict:
    """initSpectrumPlan (fft/spectrum.nim:77-101): the Tukey taper (sin² ramps
    of round(fraction · window) samples at each end), its sum, nfft (the
    next power of two by default; a 2·3·5-smooth length ≥ the window)."""
    if window_len < 1:
        raise ValueError(f"a spectrum window needs >= 1 sample, got {window_len}")
    if not (dt_ms > 0 and dt_ms < math.inf):
        raise ValueError(f"dtMs must be finite and > 0, got {dt_ms}")
    if not (taper_fraction >= 0 and taper_fraction <= 0.5):
        raise ValueError(f"taperFraction must be in 0 .. 0.5, got {taper_fraction}")
    n = nfft if nfft else 1 << (window_len - 1).bit_length()
    if n < window_len:
        raise ValueError(f"nfft {n} is shorter than the window ({window_len})")
    if not is_fast_len(n):
        raise ValueError(f"nfft {n} is not 2·3·5-smooth")
    ramp = _round_half_away(taper_fraction * window_len)
    d = np.minimum(np.arange(window_len), window_len - 1 - np.arange(window_len))
    w = np.where(d >= ramp, 1.0, np.sin(np.pi / 2 * (d + 1) / (ramp + 1)) ** 2)
    return {"taper": w, "taper_sum": float(np.sum(w)), "nfft": n, "bins": n // 2 + 1,
            "dt_ms": dt_ms, "demean": demean, "window": window_len}


def _spectrum_window(p: dict, x: np.ndarray):
    """windowInto (fft/spectrum.nim:271-293): demeaned (a finite sum only),
    tapered — a shorter window takes min(taper[j], taper[n-1-j]) and its own
    Σw — and zero-padded to nfft; Σw is NaN for a window holding NaN or Inf."""
    W = p["window"]
    n = min(len(x), W)
    xs = np.asarray(x[:n], np.float64)
    s = float(np.sum(xs))
    finite = math.isfinite(s)
    m = s / n if (p["demean"] and n > 0 and finite) else 0.0
    t = p["taper"]
    w = t[:n] if n == W else np.minimum(t[:n], t[:n][::-1])
    y = np.zeros(p["nfft"])
    y[:n] = (xs - m) * w
    sw = p["taper_sum"] if n == W else float(np.sum(w))
    return y, (sw if finite else math.nan)


def _sides(p: dict) -> np.ndarray:
    s = np.full(p["bins"], 2.0)
    s[0] = 1.0
    if p["nfft"] % 2 == 0:
        s[-1] = 1.0
    return s


def spectrum_amplitude(p: dict, x: np.ndarray):
    """amplitudeInto (fft/spectrum.nim:295-312): |X[k]|·sides/Σw and the phase."""
    y, sw = _spectrum_window(p, x)
    X = np.fft.rfft(y)
    return np.abs(X) * _sides(p) / sw, np.angle(X)


def spectrum_average(p: dict, x: np.ndarray, row_len, first_col: int):
    """averageAmplitudeInto (fft/spectrum.nim:334-433): over live rows with
    live samples past first_col and a finite window, sqrt(mean(|X|²/Σw²))·sides;
    returns (amp, rows used)."""
    if first_col < 0:
        raise ValueError(f"firstCol must be >= 0, got {first_col}")
    acc = np.zeros(p["bins"])
    used = 0
    for r in range(x.shape[0]):
        live = int(row_len[r])
        if live <= first_col:        # not live (-1), or no live sample in the window
            continue
        last = min(live, first_col + p["window"])
        y, sw = _spectrum_window(p, x[r, first_col:last])
        if sw > 0:
            X = np.fft.rfft(y)
            acc += (X.real ** 2 + X.imag ** 2) / (sw * sw)
            used += 1
    amp = np.sqrt(acc / used) * _sides(p) if used else np.zeros(p["bins"])
    return amp, used


# ---------------------------------------------------------------- Hilbert --

def hilbert(x: np.ndarray, samples: int, nfft: int = 0, envelope: bool = False) -> np.ndarray:
    """hilbertInto / envelopeInto (fft/hilbert.nim:44-150) in float64: x
    zero-padded to nfft (nextFastLen(2·samples) by default; >= 2·samples),
    -i·sign(f) on the spectrum with 0 Hz and Nyquist set to zero, inverse."""
    if samples < 1:
        raise ValueError(f"a Hilbert plan needs samples >= 1, got {samples}")
    n = nfft if nfft else next_fast_len(2 * samples)
    if n < 2 * samples:
        raise ValueError(f"nfft {n} is below twice the samples ({2 * samples})")
    if not is_fast_len(n):
        raise ValueError(f"nfft {n} is not 2·3·5-smooth")
    xs = np.asarray(x, np.float64)
    buf = np.zeros(n)
    buf[:xs.size] = xs
    X = np.fft.rfft(buf)
    X[0] = 0
    last = X.size - 1 if n % 2 == 0 else X.size
    X[1:last] = -1j * X[1:last]
    if last < X.size:
        X[last] = 0
    h = np.fft.irfft(X, n)[:xs.size]
    return np.sqrt(xs * xs + h * h) if envelope else h


def _wrap_pi(d):
    """wrapPi (fft/hilbert.nim:152-158): d brought into (-pi, pi]."""
    two = 2 * math.pi
    r = d - two * np.floor((d + math.pi) / two)
    r = np.where(r <= -math.pi, r + two, r)
    return np.where(r > math.pi, r - two, r)


def analytic(x: np.ndarray, samples: int, dt_ms: float, floor_db: float = -60.0,
             nfft: int = 0, h: np.ndarray | None = None):
    """analyticInto (fft/hilbert.nim:160-215) in float64: the envelope
    |x + i·H[x]|, the instantaneous phase in degrees in (-180, 180] (NaN
    where the envelope is 0) and the instantaneous frequency in Hz, the
    circular mean of the two wrapped one-sample phase steps about each
    sample (NaN at both ends and where a sample or a neighbour is 0 or
    below peak·10^(floor_db/20)).  `h`: the Hilbert transform to use (default:
    this oracle's)."""
    xs = np.asarray(x, np.float64)
    n = xs.size
    if h is None:
        h = hilbert(xs, samples, nfft)
    h = np.asarray(h, np.float64)
    env = np.hypot(xs, h)
    with np.errstate(invalid="ignore"):
        deg = np.arctan2(h, xs) * (180 / math.pi)
    deg = np.where(deg <= -180, deg + 360, deg)
    phase = np.where(env > 0, deg, np.nan)
    peak = float(np.max(env[env > 0])) if np.any(env > 0) else 0.0
    thr = peak * math.pow(10.0, floor_db / 20.0)
    scale = 1000 / (4 * math.pi * dt_ms)
    hz = np.full(n, np.nan)
    if n > 2:
        e0, e1, e2 = env[:-2], env[1:-1], env[2:]
        bad = (e0 <= 0) | (e1 <= 0) | (e2 <= 0) | (e0 < thr) | (e1 < thr) | (e2 < thr)
        with np.errstate(invalid="ignore"):
            back = _wrap_pi((phase[1:-1] - phase[:-2]) * math.pi / 180)
            fwd = _wrap_pi((phase[2:] - phase[1:-1]) * math.pi / 180)
            val = _wrap_pi(back + 0.5 * _wrap_pi(fwd - back)) * (2 * scale)
        hz[1:-1] = np.where(bad, np.nan, val)
    return env, phase, hz


# ------------------------------------------------------------ spectrogram --

def spectrogram(x: np.ndarray, window_len: int, hop: int, dt_ms: float, samples: int,
                nfft: int = 0) -> np.ndarray:
    """initSpectrogramPlan + spectrogramInto (fft/spectrogram.nim:63-85,
    263-346) in float64: frame i covers columns i·hop − window_len//2 + j,
    weighted by sin²(π(j+1)/(window_len+1)) (outside the trace: zero),
    zero-padded to nfft (the next power of two of the window by default);
    amp[i, k] = |X[k]|·2/Σw (1/Σw at 0 Hz and at Nyquist), Σw the weights
    of the frame's samples inside the trace; a frame with none reads 0."""
    if window_len < 2:
        raise ValueError(f"windowLen must be >= 2, got {window_len}")
    if hop < 1:
        raise ValueError(f"hop must be >= 1, got {hop}")
    if samples < 1:
        raise ValueError(f"samples must be >= 1, got {samples}")
    if not (dt_ms > 0 and dt_ms < math.inf):
        raise ValueError(f"dtMs must be finite and > 0, got {dt_ms}")
    n = nfft if nfft else 1 << (window_len - 1).bit_length()
    if n < window_len:
        raise ValueError(f"nfft {n} is shorter than the window ({window_len})")
    if not is_fast_len(n):
        raise ValueError(f"nfft {n} is not 2·3·5-smooth")
    xs = np.asarray(x, np.float64)
    if xs.size > samples:
        raise ValueError("a trace longer than the plan's samples")
    frames = (samples - 1) // hop + 1
    w = np.sin(np.pi * np.arange(1, window_len + 1) / (window_len + 1)) ** 2
    cols = np.arange(frames)[:, None] * hop - window_len // 2 + np.arange(window_len)[None, :]
    inside = (cols >= 0) & (cols < xs.size)
    buf = np.zeros((frames, n))
    buf[:, :window_len] = np.where(inside, xs[np.clip(cols, 0, max(xs.size - 1, 0))]
                                   if xs.size else 0.0, 0.0) * w
    sw = np.sum(np.where(inside, w, 0.0), axis=1)
    X = np.fft.rfft(buf, axis=1)
    sides = np.full(n // 2 + 1, 2.0)
    sides[0] = 1.0
    if n % 2 == 0:
        sides[-1] = 1.0
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.hypot(X.real, X.imag) * sides[None, :] / sw[:, None]
    amp[sw <= 0] = 0.0
    return amp


def _leaves_float64_range(x: np.ndarray) -> bool:
    """leavesFloat64Range (fft/spectrogram.nim:137-156): a NaN, an infinity,
    a subnormal, or a nonzero magnitude below 1e-100 or above 1e140."""
    a = np.abs(np.asarray(x, np.float64))
    with np.errstate(invalid="ignore"):
        bad = ~np.isfinite(a) | (a > 1e140) | ((a != 0) & (a < 1e-100))
    return bool(np.any(bad))


def spectrogram_average(x: np.ndarray, row_len, window_len: int, hop: int, dt_ms: float,
                        samples: int, nfft: int = 0, dtype=np.float64):
    """averageSpectrogramInto (fft/spectrogram.nim:370-430): the power mean
    sqrt(Σ amp² / used) of the rows' spectrograms over their live samples
    (row_len[r] < 0: not live; 0: no sample).  A float64 row whose live
    samples leave the safe range is left out before the transform; any row
    whose spectrogram holds a NaN or an infinity is left out.  Returns
    (amp [frames, bins], used); with no row used amp is all zero."""
    frames = (samples - 1) // hop + 1
    acc = None
    used = 0
    for r in range(x.shape[0]):
        n = int(row_len[r])
        if n <= 0:
            continue
        xs = np.asarray(x[r, :n])
        if np.dtype(dtype) == np.float64 and _leaves_float64_range(xs):
            continue
        with np.errstate(over="ignore", invalid="ignore"):
            a = spectrogram(xs.astype(np.float64), window_len, hop, dt_ms, samples, nfft)
        if not np.all(np.isfinite(a)):
            continue
        acc = a * a if acc is None else acc + a * a
        used += 1
    if used == 0:
        n = nfft if nfft else 1 << (window_len - 1).bit_length()
        return np.zeros((frames, n // 2 + 1)), 0
    return np.sqrt(acc / used), used


# ------------------------------------------------------------ calibration
#
# A restatement of core/calibrate.nim's plan and per-row resolution in plain
# Python: rules as dicts with the keys sensorType, component, channelSet
# (-1: any), channel (NaN: any), station (None or (line, point, index)),
# descale (V/count, NaN: none), sensitivity (V per SI unit, NaN: none),
# unit (the sensitivity's quantity: 2 m/s, 3 m/s², 4 Pa, 5 N), gainDb (NaN:
# none).  Rows as dicts: status (0 live, 1 unknown, 2 dead, 3 aux), failed,
# sensorType, component, channelSet (-1: none), channel, rcvLine, rcvPoint,
# rcvIndex, and the FileScale fields fsDescale, fsSensitivity, fsUnit,
# fsPhysical.

_CAL_BITS = {"sensorType": 1, "component": 2, "channelSet": 4, "channel": 8, "station": 16}
_F32_MAX = float(np.finfo(np.float32).max)


def _hundredths(x):
    # round half away from zero (Nim's round), core/calibrate.nim:225
    y = x * 100
    r = math.floor(abs(y) + 0.5) * (1 if y >= 0 else -1)
    return r / 100 + 0.0


def _rule_key(r):
    """keyOf (core/calibrate.nim:256-264): (mask, the masked values)."""
    mask, vals = 0, []
    if r["sensorType"] >= 0:
        mask |= 1
        vals.append(("st", r["sensorType"]))
    if r["component"] >= 0:
        mask |= 2
        vals.append(("co", r["component"]))
    if r["channelSet"] >= 0:
        mask |= 4
        vals.append(("cs", r["channelSet"]))
    if not math.isnan(r["channel"]):
        mask |= 8
        vals.append(("ch", _hundredths(r["channel"])))
    if r["station"] is not None:
        mask |= 16
        line, point, index = r["station"]
        vals.append(("sta", _hundredths(line), _hundredths(point), int(index)))
    return mask, tuple(vals)


def calib_plan(rules, target=0, file_factors=0, scale_aux=True):
    """initCalibrationPlan (core/calibrate.nim:316-360); ValueError("rule i:
    ...") as the reference raises it."""
    p = {"target": target, "ff": file_factors, "aux": scale_aux, "catch": -1, "index": {},
         "d": [], "s": [], "g": [], "u": []}
    masks = set()
    for i, r in enumerate(rules):
        has_d, has_s, has_g = (not math.isnan(r["descale"]), not math.isnan(r["sensitivity"]),
                               not math.isnan(r["gainDb"]))
        if not (has_d or has_s or has_g):
            raise ValueError(f"rule {i}: gives no factor (set a descale, a sensitivity or gainDb)")
        lin = math.nan
        if has_g:
            if not math.isfinite(r["gainDb"]):
                raise ValueError(f"rule {i}: gainDb must be finite, got {r['gainDb']}")
            try:
                lin = 10.0 ** (r["gainDb"] / 20)
            except OverflowError:
                lin = math.inf
            if not (math.isfinite(lin) and lin > 0):
                raise ValueError(f"rule {i}: gainDb = {r['gainDb']} is out of range")
        if not -1 <= r["sensorType"] <= 255:
            raise ValueError(f"rule {i}: sensorType must be -1 (any) or 0 .. 255")
        if not -1 <= r["component"] <= 255:
            raise ValueError(f"rule {i}: component must be -1 (any) or 0 .. 255")
        if r["channelSet"] < -1:
            raise ValueError(f"rule {i}: channelSet must be -1 (any) or ≥ 0")
        if not math.isnan(r["channel"]) and not math.isfinite(r["channel"]):
            raise ValueError(f"rule {i}: channel must be finite (NaN: any)")
        if r["station"] is not None and not all(math.isfinite(v) for v in r["station"][:2]):
            raise ValueError(f"rule {i}: the station's line and point must be finite")
        key = _rule_key(r)
        if key[0] == 0:
            if p["catch"] >= 0:
                raise ValueError(f"rule {i}: has no criteria, like rule {p['catch']}: one "
                                 "catch-all only")
            p["catch"] = i
        else:
            if key in p["index"]:
                raise ValueError(f"rule {i}: matches the same rows as rule {p['index'][key]}")
            p["index"][key] = i
            masks.add(key[0])
        p["d"].append(r["descale"] if has_d else math.nan)
        p["s"].append(r["sensitivity"] if has_s else math.nan)
        p["u"].append(r.get("unit", 0) if has_s else 0)
        p["g"].append(lin)
    p["classes"] = sorted(masks, reverse=True)
    return p


def _row_key(mask, row, index=None):
    vals = []
    if mask & 1:
        vals.append(("st", int(row["sensorType"])))
    if mask & 2:
        vals.append(("co", int(row["component"])))
    if mask & 4:
        vals.append(("cs", int(row["channelSet"])))
    if mask & 8:
        vals.append(("ch", _hundredths(row["channel"])))
    if mask & 16:
        vals.append(("sta", _hundredths(row["rcvLine"]), _hundredths(row["rcvPoint"]), index))
    return mask, tuple(vals)


def _specific(p, row):
    """specificRules (core/calibrate.nim:383-410)."""
    f = {"d": -1, "s": -1, "g": -1}

    def take(r):
        for k in "dsg":
            if f[k] < 0 and not math.isnan(p[k][r]):
                f[k] = r
    for mask in p["classes"]:
        if mask & 4 and row["channelSet"] < 0:
            continue
        if mask & 8 and math.isnan(row["channel"]):
            continue
        if mask & 16:
            if math.isnan(row["rcvLine"]) or math.isnan(row["rcvPoint"]):
                continue
            idx = int(row["rcvIndex"])
            for t in ((idx, 0) if idx != 0 else (0, 1)):
                r = p["index"].get(_row_key(mask, row, t), -1)
                if r >= 0:
                    take(r)
        else:
            r = p["index"].get(_row_key(mask, row), -1)
            if r >= 0:
                take(r)
        if min(f.values()) >= 0:
            break
    return f


def calib_resolve_row(p, row, has_file_scales=False):
    """resolveRow (core/calibrate.nim:430-497): a dict with factor, unit,
    status (0 converted, 1 unresolved, 2 kept), descaleFrom /
    sensitivityFrom (0 none, 1 rule, 2 file), descaleRule, sensitivityRule,
    gainRule, conflict, mismatch."""
    res = {"factor": 1.0, "unit": 0, "status": 2, "descaleFrom": 0, "sensitivityFrom": 0,
           "descaleRule": -1, "sensitivityRule": -1, "gainRule": -1, "conflict": False,
           "mismatch": False}
    if row.get("failed", False):
        return res
    aux = row["status"] == 3
    live = row["status"] in (0, 1)
    if aux and not p["aux"]:
        return res
    use_file = p["ff"] != 2 and has_file_scales
    fd = row.get("fsDescale", math.nan) if use_file else math.nan
    fsen = row.get("fsSensitivity", math.nan) if use_file else math.nan
    fu = row.get("fsUnit", 0) if use_file else 0
    fphys = bool(row.get("fsPhysical", False)) and use_file
    spec = _specific(p, row)
    c = p["catch"]

    def rule_for(k):
        if spec[k] >= 0:
            return spec[k]
        return c if c >= 0 and not math.isnan(p[k][c]) else -1
    rd, rs, rg = rule_for("d"), rule_for("s"), rule_for("g")
    unit, factor = 0, math.nan
    phys_units = (2, 3, 4, 5)
    if fphys:
        res["conflict"] = rd >= 0 or rs >= 0 or rg >= 0
        if p["target"] == 0 and not aux and math.isfinite(fd) and fd > 0:
            factor, unit = fd, fu
            res["descaleFrom"] = res["sensitivityFrom"] = 2
    else:
        file_d = math.isfinite(fd) and fd > 0
        file_s = math.isfinite(fsen) and fsen > 0 and fu in phys_units

        def pick(file_has, file_v, si, ri, k):
            if p["ff"] == 1 and file_has:
                return file_v, 2, -1
            if si >= 0:
                return p[k][si], 1, si
            if file_has:
                return file_v, 2, -1
            if ri >= 0:
                return p[k][ri], 1, ri
            return math.nan, 0, -1
        d, res["descaleFrom"], res["descaleRule"] = pick(file_d, fd, spec["d"], rd, "d")
        want_s = not aux and p["target"] == 0
        s = math.nan
        s_unit = 0
        if want_s:
            s, res["sensitivityFrom"], res["sensitivityRule"] = pick(file_s, fsen, spec["s"], rs, "s")
            s_unit = (fu if res["sensitivityFrom"] == 2 else
                      p["u"][res["sensitivityRule"]] if res["sensitivityRule"] >= 0 else 0)

        def differs(a, b):
            return abs(a - b) > 1e-3 * max(abs(a), abs(b))
        if file_d and rd >= 0 and differs(fd, p["d"][rd]):
            res["conflict"] = True
        if want_s and file_s and rs >= 0 and (fu != p["u"][rs] or differs(fsen, p["s"][rs])):
            res["conflict"] = True
        g = p["g"][rg] if rg >= 0 else 1.0
        res["gainRule"] = rg
        if want_s and not math.isnan(s):
            st = int(row["sensorType"])
            want = 4 if st == 1 else 2 if 2 <= st <= 5 else 3 if 6 <= st <= 9 else 0
            res["mismatch"] = (want != 0 and want != s_unit) or (fu in phys_units and fu != s_unit)
        if not math.isnan(d):
            if aux or p["target"] == 1:
                factor, unit = d / g, 1
            elif not math.isnan(s):
                factor, unit = d / (g * s), s_unit
    if factor > 0 and factor <= _F32_MAX and 1 / factor <= _F32_MAX:
        res.update(factor=factor, unit=unit, status=0)
    elif live:
        res.update(factor=math.nan, status=1)
    return res


def calibrate_rows(x: np.ndarray, k, live_len) -> np.ndarray:
    """calibrateImpl's per-row scale (core/calibrate.nim:615-624, scaleRow
    :562-574): row r's first live_len[r] samples times T(k[r]); k 1 leaves
    the row as it is (a kept row is copied whole).  Returns a new array."""
    out = np.array(x, copy=True)
    for r in range(x.shape[0]):
        if k[r] == 1.0:
            continue
        n = int(live_len[r])
        with np.errstate(invalid="ignore", over="ignore"):
            out[r, :n] = out[r, :n] * x.dtype.type(k[r])
    return out


# -------------------------------------------------------------------- f-k

def _tukey_fk(n: int, fraction: float, dtype):
    """tukey (fft/fk.nim:70-80): 1 in the middle, 0.5 - 0.5 cos(pi (d + 0.5)
    / ramp) over ramp = round(fraction n) samples at each end, as T."""
    ramp = 0 if fraction <= 0 else _round_half_away(fraction * n)
    w = np.ones(n)
    for i in range(n):
        d = min(i, n - 1 - i)
        if ramp > 0 and d < ramp:
            w[i] = 0.5 - 0.5 * math.cos(math.pi * (d + 0.5) / ramp)
    return w.astype(dtype)


def fk_plan(traces: int, samples: int, dt_ms: float, taper_fraction: float = 0.1,
            dtype=np.float64) -> dict:
    """initFkPlan (fft/fk.nim:82-106): the time and wavenumber lengths, the
    axes and the tapers; ValueError as the reference raises it."""
    if traces < 4:
        raise ValueError(f"an f-k gather needs ≥ 4 live rows, got {traces}")
    if samples < 8:
        raise ValueError(f"an f-k window needs ≥ 8 samples, got {samples}")
    if not (dt_ms > 0 and math.isfinite(dt_ms)):
        raise ValueError(f"dtMs must be finite and > 0, got {dt_ms}")
    if not (0 <= taper_fraction <= 0.5):
        raise ValueError(f"taperFraction must be in 0 .. 0.5, got {taper_fraction}")
    n_time = next_fast_len(samples)
    n_k = next_fast_len(2 * traces)
    return dict(traces=traces, samples=samples, dt_ms=dt_ms, dtype=np.dtype(dtype),
                nTime=n_time, nK=n_k, nFreq=n_time // 2 + 1, dfHz=1000 / (n_time * dt_ms),
                dkPerTrace=1 / n_k, timeW=_tukey_fk(samples, taper_fraction, dtype),
                traceW=_tukey_fk(traces, taper_fraction, dtype))


def _leaves_f64_fk(x: np.ndarray) -> bool:
    """leavesFloat64Range (fft/fk.nim:164-181): NaN, infinite, subnormal,
    below 1e-100 or above 1e140 (zero allowed)."""
    a = np.abs(x.astype(np.float64))
    return bool(np.any(~np.isfinite(a) | (a > 1e140) | ((a < 1e-100) & (a != 0))))


def fk_amplitude(p: dict, x: np.ndarray, row_len, first_col: int) -> tuple:
    """fkAmplitudeInto (fft/fk.nim:438-504): returns (amp [nK * nFreq]
    float64, rows used).  row_len: the live length of each live row, -1 for
    the others; a live row enters when its live length passes first_col.
    The time pass is the oracle's real FFT of each tapered trace in T; the
    spatial pass two real FFTs (the real and imaginary parts across traces)
    combined with the sign flip, the modulus in float64 (hypot when a
    float64 sample leaves range)."""
    T = p["dtype"].type
    x = np.asarray(x)
    live = [r for r in range(x.shape[0]) if row_len[r] >= 0 and row_len[r] > first_col]
    if len(live) != p["traces"]:
        raise ValueError(f"the window has {len(live)} live rows, the plan was built for "
                         f"{p['traces']}")
    n_time, n_k, n_freq, S = p["nTime"], p["nK"], p["nFreq"], p["samples"]
    spec_re = np.zeros((len(live), n_freq), p["dtype"])
    spec_im = np.zeros((len(live), n_freq), p["dtype"])
    extreme = False
    for ti, r in enumerate(live):
        last = min(int(row_len[r]), first_col + S)
        seg = x[r, first_col:last].astype(p["dtype"])
        if p["dtype"] == np.float64 and _leaves_f64_fk(seg):
            extreme = True
        lane = np.zeros(n_time, p["dtype"])
        with np.errstate(invalid="ignore", over="ignore"):
            lane[:last - first_col] = (seg * p["timeW"][:last - first_col]) * p["traceW"][ti]
        spec_re[ti], spec_im[ti] = rfft(lane)
    amp = np.zeros(n_k * n_freq)
    half = n_k // 2
    for jf in range(n_freq):
        a = np.zeros(n_k, p["dtype"])
        b = np.zeros(n_k, p["dtype"])
        a[:len(live)] = spec_re[:, jf]
        b[:len(live)] = spec_im[:, jf]
        ar, ai = (v.astype(np.float64) for v in rfft(a))
        br, bi = (v.astype(np.float64) for v in rfft(b))
        for ik in range(n_k):
            raw = half - ik
            sk = raw if raw >= 0 else raw + n_k
            if sk <= half:
                re, im = float(ar[sk] - bi[sk]), float(ai[sk] + br[sk])
            else:
                k = n_k - sk
                re, im = float(ar[k] + bi[k]), float(br[k] - ai[k])
            amp[ik * n_freq + jf] = (math.hypot(re, im) if extreme
                                     else math.sqrt(re * re + im * im))
    return amp, len(live)
