"""ctypes binding of libpseg.so (the C ABI of include/pseg.h).

The library is built in-tree (``make -C processseg_amd``, or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback for the panel path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpseg.so")

MAX_NOTCH = 8
MAX_OPS = 16

OK, E_INVALID, E_UNSUPPORTED, E_HIP, E_WORKSPACE, E_BUSY = range(6)


class FilterSpecC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("zero_dc", C.c_int32), ("n_notch", C.c_int32),
                ("reserved", C.c_int32),
                ("f1", C.c_double), ("f2", C.c_double), ("f3", C.c_double), ("f4", C.c_double),
                ("low_hz", C.c_double), ("low_slope", C.c_double),
                ("high_hz", C.c_double), ("high_slope", C.c_double),
                ("notch_hz", C.c_double * MAX_NOTCH), ("notch_width", C.c_double * MAX_NOTCH)]


class CalibRuleC(C.Structure):
    _fields_ = [("sensor_type", C.c_int32), ("component", C.c_int32),
                ("channel_set", C.c_int32), ("has_station", C.c_int32),
                ("channel", C.c_double), ("station_line", C.c_double),
                ("station_point", C.c_double), ("station_index", C.c_int32),
                ("sensitivity_unit", C.c_int32), ("descale", C.c_double),
                ("sensitivity", C.c_double), ("gain_db", C.c_double)]


class OpC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("trend", C.c_int32), ("mode", C.c_int32),
                ("stat", C.c_int32), ("scope", C.c_int32), ("reserved", C.c_int32),
                ("value", C.c_double), ("lo", C.c_double), ("hi", C.c_double),
                ("window_ms", C.c_double), ("target", C.c_double), ("epsilon", C.c_double),
                ("percentile", C.c_double), ("spec", FilterSpecC),
                ("mute_kind", C.c_int32), ("mute_knots", C.c_int32),
                ("mute_t_ms", C.c_double), ("mute_ms_per_dist", C.c_double),
                ("mute_pad_ms", C.c_double), ("taper_ms", C.c_double),
                ("mute_offsets", C.c_double * 32), ("mute_times", C.c_double * 32),
                ("gain_kind", C.c_int32), ("gain_knots", C.c_int32),
                ("gain_knot_ms", C.POINTER(C.c_double)), ("gain_knot_db", C.POINTER(C.c_double)),
                ("gain_power", C.c_double), ("gain_ref_ms", C.c_double),
                ("gain_min_ms", C.c_double), ("gain_db_per_s", C.c_double),
                ("gain_tmin_ms", C.c_double), ("gain_tmax_ms", C.c_double),
                ("iir_phase", C.c_int32), ("reserved2", C.c_int32),
                ("calib_rules", C.POINTER(CalibRuleC)), ("calib_n_rules", C.c_int32),
                ("calib_target", C.c_int32), ("calib_file_factors", C.c_int32),
                ("calib_scale_aux", C.c_int32)]


class CalibRowC(C.Structure):
    _fields_ = [("status", C.c_uint8), ("failed", C.c_uint8), ("sensor_type", C.c_uint8),
                ("file_physical", C.c_uint8), ("component", C.c_int32),
                ("channel_set", C.c_int32), ("rcv_index", C.c_int32),
                ("channel", C.c_double), ("rcv_line", C.c_double), ("rcv_point", C.c_double),
                ("file_descale", C.c_double), ("file_sensitivity", C.c_double),
                ("file_unit", C.c_int32), ("reserved", C.c_int32)]


class RowCalibrationC(C.Structure):
    _fields_ = [("factor", C.c_double), ("unit", C.c_int32), ("status", C.c_int32),
                ("descale_from", C.c_int32), ("sensitivity_from", C.c_int32),
                ("descale_rule", C.c_int32), ("sensitivity_rule", C.c_int32),
                ("gain_rule", C.c_int32), ("conflict", C.c_int32), ("mismatch", C.c_int32)]


class CalibSummaryC(C.Structure):
    _fields_ = [("units", C.c_int32 * 6), ("converted", C.c_int32), ("unresolved", C.c_int32),
                ("kept", C.c_int32), ("no_descale", C.c_int32), ("no_sensitivity", C.c_int32),
                ("conflicts", C.c_int32), ("mismatches", C.c_int32), ("reserved", C.c_int32)]


class PlanInfoC(C.Structure):
    _fields_ = [("generation", C.c_uint64), ("dtype", C.c_int32), ("n_steps", C.c_int32),
                ("parts", C.c_int32), ("samples", C.c_int32), ("record_samples", C.c_int32),
                ("margin", C.c_int32), ("dt_ms", C.c_double),
                ("step_kind", C.c_int32 * MAX_OPS), ("step_nfft", C.c_int32 * MAX_OPS),
                ("step_pad", C.c_int32 * MAX_OPS), ("step_margin", C.c_int32 * MAX_OPS),
                ("step_reflect", C.c_int32 * MAX_OPS), ("lds_bytes", C.c_size_t),
                ("threads_per_row", C.c_int32), ("amplitude_units", C.c_int32),
                ("n_calib", C.c_int32), ("reserved", C.c_int32)]


class CorrInfoC(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("nfft", C.c_int32), ("shift", C.c_int32),
                ("out_samples", C.c_int32), ("pilot_samples", C.c_int32),
                ("record_samples", C.c_int32), ("dt_ms", C.c_double), ("scale", C.c_double),
                ("lds_bytes", C.c_size_t), ("threads_per_record", C.c_int32),
                ("reserved", C.c_int32)]


class BatchItemC(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("src", C.c_void_p), ("dst", C.c_void_p),
                ("rows", C.c_int32), ("cols", C.c_int32), ("row_len", C.c_void_p),
                ("row_keys", C.c_void_p), ("first_sample", C.c_int64)]


class BatchInfoC(C.Structure):
    _fields_ = [("items", C.c_int32), ("launches", C.c_int32), ("rows", C.c_int64),
                ("samples", C.c_int64), ("threads", C.c_int32 * 12), ("grid", C.c_int32 * 12)]


class BatchTuningC(C.Structure):
    _fields_ = [("launches", C.c_int32), ("grid", C.c_int32 * 12), ("direct", C.c_int32 * 12),
                ("best_ms", C.c_float * 12), ("candidate_ms", (C.c_float * 4) * 12)]


class ServerConfigC(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("traces", C.c_int32), ("samples", C.c_int32),
                ("max_batch_requests", C.c_int32), ("pitch", C.c_int64), ("survey", C.c_void_p),
                ("trace_len", C.c_void_p), ("dt_ms", C.c_double), ("arena_bytes", C.c_size_t),
                ("max_batch_rows", C.c_int32), ("max_inflight", C.c_int32),
                ("trace_keys", C.c_void_p)]


class ServeResultC(C.Structure):
    _fields_ = [("status", C.c_int32), ("rows", C.c_int32), ("cols", C.c_int32),
                ("first", C.c_int32), ("window", C.c_int32), ("batch_requests", C.c_int32),
                ("plan_compiled", C.c_int32), ("reserved", C.c_int32), ("samples", C.c_void_p),
                ("generation", C.c_uint64), ("submit_ns", C.c_int64), ("start_ns", C.c_int64),
                ("done_ns", C.c_int64)]


class ServerInfoC(C.Structure):
    _fields_ = [("requests", C.c_int64), ("batches", C.c_int64), ("launches", C.c_int64),
                ("plans_compiled", C.c_int64), ("arena_peak_bytes", C.c_int64)]


class UnsupportedOp(ValueError):
    """An op outside the GPU path: keep the CPU plan for that flow."""


class WorkspaceError(AssertionError):
    """A workspace missing or smaller than the plan needs (the reference's
    AssertionDefect for an unreserved Workspace)."""


class HipError(RuntimeError):
    pass


class BusyError(TimeoutError):
    """A request that is not done yet (PSEG_E_BUSY)."""


def build(force: bool = False) -> str:
    """Compile libpseg.so for gfx950 (hipcc cross-compiles without a GPU)."""
    args = ["make", "-s", "-C", HERE]
    if force:
        args.insert(2, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


_lib = None

SYMBOLS = [
    "pseg_last_error", "pseg_device_count", "pseg_meta_flatten", "pseg_next_fast_len",
    "pseg_impulse_length", "pseg_response", "pseg_plan_compile", "pseg_plan_destroy",
    "pseg_plan_info_get", "pseg_required_margin", "pseg_decode_window",
    "pseg_plan_workspace_size", "pseg_plan_apply", "pseg_plan_apply_keys", "pseg_corr_compile",
    "pseg_corr_destroy",
    "pseg_corr_info_get", "pseg_corr_apply", "pseg_linear_sweep", "pseg_batch_create",
    "pseg_batch_launch", "pseg_batch_info_get", "pseg_batch_destroy", "pseg_server_create",
    "pseg_server_add_flow", "pseg_server_submit", "pseg_server_wait", "pseg_server_read",
    "pseg_server_release", "pseg_server_info_get", "pseg_server_destroy",
    "pseg_decode_samples", "pseg_plan_apply_raw", "pseg_spectrum_compile",
    "pseg_spectrum_destroy", "pseg_spectrum_info_get", "pseg_spectrum_average",
    "pseg_spectrum_workspace_size", "pseg_spectrum_average_noalloc",
    "pseg_spectrum_amplitude", "pseg_spectrogram_average", "pseg_spectrogram_average_workspace_size", "pseg_hilbert_analytic", "pseg_hilbert_compile", "pseg_hilbert_destroy",
    "pseg_hilbert_nfft", "pseg_hilbert_apply", "pseg_spectrogram_compile",
    "pseg_spectrogram_destroy", "pseg_spectrogram_info_get", "pseg_spectrogram_apply",
    "pseg_firstbreak_compile", "pseg_firstbreak_destroy", "pseg_firstbreak_required_margin",
    "pseg_firstbreak_workspace_size", "pseg_firstbreak_pick", "pseg_fold_add", "pseg_fold_clear",
    "pseg_fold_difference", "pseg_plan_apply_host", "pseg_host_staging_release",
    "pseg_plan_apply_calib", "pseg_calib_compile", "pseg_calib_destroy", "pseg_calib_resolve",
    "pseg_plan_calib_factors", "pseg_fk_compile", "pseg_fk_destroy", "pseg_fk_axes",
    "pseg_fk_workspace_size", "pseg_fk_amplitude", "pseg_batch_tune", "pseg_response_moduli",
    "pseg_applied_response",
]


def lib():
    """The loaded library; raises ImportError when it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it first (make -C processseg_amd "
                          "or __graft_entry__.build()); the panel path has no CPU fallback")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, f64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.pseg_last_error.argtypes = [C.c_char_p, C.c_size_t]
    L.pseg_last_error.restype = C.c_size_t
    L.pseg_device_count.restype = i32
    L.pseg_meta_flatten.argtypes = [i32, i32, i64, vp, vp, vp, i32, i32, vp]
    L.pseg_next_fast_len.argtypes = [i64, i32]
    L.pseg_next_fast_len.restype = i64
    L.pseg_impulse_length.argtypes = [C.POINTER(FilterSpecC), f64, C.POINTER(i64)]
    L.pseg_response.argtypes = [C.POINTER(FilterSpecC), f64, f64]
    L.pseg_response.restype = f64
    L.pseg_plan_compile.argtypes = [C.POINTER(OpC), i32, i32, f64, i32, i32, C.POINTER(vp)]
    L.pseg_plan_destroy.argtypes = [vp]
    L.pseg_plan_info_get.argtypes = [vp, C.POINTER(PlanInfoC)]
    L.pseg_required_margin.argtypes = [C.POINTER(OpC), i32, f64, C.POINTER(i32)]
    L.pseg_decode_window.argtypes = [i32, i32, i64, i32, C.POINTER(i32), C.POINTER(i32)]
    L.pseg_decode_window.restype = None
    L.pseg_plan_workspace_size.argtypes = [vp, i32, i32, C.POINTER(C.c_size_t)]
    L.pseg_plan_apply.argtypes = [vp, vp, vp, i32, i32, vp, vp, C.c_size_t, vp]
    L.pseg_plan_apply_keys.argtypes = [vp, vp, vp, i32, i32, vp, vp, C.c_int64, vp, C.c_size_t, vp]
    L.pseg_plan_apply_calib.argtypes = [vp, vp, vp, i32, i32, vp, vp, C.c_int64, vp, vp, vp,
                                        C.c_size_t, vp]
    L.pseg_fk_compile.argtypes = [i32, f64, i32, i32, f64, C.POINTER(vp)]
    L.pseg_fk_destroy.argtypes = [vp]
    L.pseg_fk_axes.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(f64),
                               C.POINTER(f64)]
    L.pseg_fk_workspace_size.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.pseg_fk_amplitude.argtypes = [vp, vp, i32, i32, vp, vp, i32, vp, vp, C.c_size_t, vp]
    L.pseg_calib_compile.argtypes = [vp, i32, i32, i32, i32, C.POINTER(vp)]
    L.pseg_calib_destroy.argtypes = [vp]
    L.pseg_calib_resolve.argtypes = [vp, vp, i32, i32, vp, vp]
    L.pseg_plan_calib_factors.argtypes = [vp, vp, i32, i32, vp, vp, C.POINTER(i32)]
    L.pseg_decode_samples.argtypes = [i32, vp, C.c_int64, C.c_int64, i32, i32, i32, vp, vp]
    L.pseg_spectrum_compile.argtypes = [i32, f64, i32, f64, i32, i32, C.POINTER(vp)]
    L.pseg_spectrum_destroy.argtypes = [vp]
    L.pseg_spectrum_info_get.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
    L.pseg_spectrum_average.argtypes = [vp, vp, i32, i32, vp, i32, vp, C.c_size_t, vp, vp, vp]
    L.pseg_spectrum_workspace_size.argtypes = [vp, i32, C.POINTER(C.c_size_t)]
    L.pseg_spectrum_average_noalloc.argtypes = [vp, vp, i32, i32, vp, i32, vp, vp, vp]
    L.pseg_spectrum_amplitude.argtypes = [vp, vp, i32, i32, vp, i32, vp, vp, vp]
    L.pseg_hilbert_compile.argtypes = [i32, i32, i32, C.POINTER(vp)]
    L.pseg_hilbert_destroy.argtypes = [vp]
    L.pseg_hilbert_nfft.argtypes = [vp, C.POINTER(i32)]
    L.pseg_spectrogram_average.argtypes = [vp, vp, i32, i32, vp, vp, C.c_size_t, vp, vp, vp]
    L.pseg_spectrogram_average_workspace_size.argtypes = [vp, i32, C.POINTER(C.c_size_t)]
    L.pseg_hilbert_analytic.argtypes = [vp, vp, i32, i32, vp, f64, f64, vp, vp, vp, vp]
    L.pseg_hilbert_apply.argtypes = [vp, vp, vp, i32, i32, vp, i32, vp]
    L.pseg_spectrogram_compile.argtypes = [i32, i32, i32, f64, i32, i32, C.POINTER(vp)]
    L.pseg_spectrogram_destroy.argtypes = [vp]
    L.pseg_spectrogram_info_get.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.pseg_spectrogram_apply.argtypes = [vp, vp, i32, vp, vp]
    L.pseg_firstbreak_compile.argtypes = [i32, vp, f64, i64, C.POINTER(vp)]
    L.pseg_firstbreak_destroy.argtypes = [vp]
    L.pseg_firstbreak_required_margin.argtypes = [vp, C.POINTER(i64)]
    L.pseg_firstbreak_workspace_size.argtypes = [vp, i32, i32, C.POINTER(C.c_size_t)]
    L.pseg_plan_apply_host.argtypes = [vp, vp, vp, i32, i32, vp, vp, i64, vp]
    L.pseg_host_staging_release.argtypes = []
    L.pseg_fold_add.argtypes = [vp, vp, i32, i32, f64, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp,
                                vp]
    L.pseg_fold_clear.argtypes = [i64, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp]
    L.pseg_fold_difference.argtypes = [vp, vp, vp, i64, vp]
    L.pseg_firstbreak_pick.argtypes = [vp, vp, i32, i32, vp, vp, i64, vp, C.c_size_t, vp, vp, vp,
                                       vp, vp]
    L.pseg_plan_apply_raw.argtypes = [vp, i32, vp, C.c_int64, C.c_int64, vp, i32, i32, vp, vp,
                                      C.c_int64, vp]
    L.pseg_corr_compile.argtypes = [i32, vp, i32, f64, f64, f64, i32, i32, i32, f64, i32,
                                    C.POINTER(vp)]
    L.pseg_corr_destroy.argtypes = [vp]
    L.pseg_corr_info_get.argtypes = [vp, C.POINTER(CorrInfoC)]
    L.pseg_corr_apply.argtypes = [vp, vp, i32, i32, vp, vp, i32, vp]
    L.pseg_linear_sweep.argtypes = [f64] * 8 + [vp, i32]
    L.pseg_batch_create.argtypes = [C.POINTER(BatchItemC), i32, C.POINTER(vp)]
    L.pseg_batch_launch.argtypes = [vp, vp]
    L.pseg_batch_info_get.argtypes = [vp, C.POINTER(BatchInfoC)]
    L.pseg_batch_destroy.argtypes = [vp]
    L.pseg_batch_tune.argtypes = [vp, i32, vp, C.POINTER(BatchTuningC)]
    L.pseg_response_moduli.argtypes = [vp, i32, vp, i32, f64, vp]
    L.pseg_applied_response.argtypes = [C.POINTER(OpC), i32, i32, f64, i32, vp, i32, vp]
    L.pseg_server_create.argtypes = [C.POINTER(ServerConfigC), C.POINTER(vp)]
    L.pseg_server_add_flow.argtypes = [vp, C.POINTER(OpC), i32, C.POINTER(i32)]
    L.pseg_server_submit.argtypes = [vp, i32, i32, i32, i32, i32, C.POINTER(C.c_uint64)]
    L.pseg_server_wait.argtypes = [vp, C.c_uint64, i32, C.POINTER(ServeResultC)]
    L.pseg_server_read.argtypes = [vp, C.c_uint64, vp, vp]
    L.pseg_server_release.argtypes = [vp, C.c_uint64]
    L.pseg_server_info_get.argtypes = [vp, C.POINTER(ServerInfoC)]
    L.pseg_server_destroy.argtypes = [vp]
    for name in SYMBOLS:
        if not hasattr(L, name):
            raise ImportError(f"{LIB_PATH} does not export {name}")
    _lib = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().pseg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int) -> None:
    """Raise the exception the reference raises for a status."""
    if rc == OK:
        return
    msg = last_error()
    if rc == E_INVALID:
        raise ValueError(msg)
    if rc == E_UNSUPPORTED:
        raise UnsupportedOp(msg)
    if rc == E_WORKSPACE:
        raise WorkspaceError(msg)
    if rc == E_BUSY:
        raise BusyError(msg)
    raise HipError(msg)
