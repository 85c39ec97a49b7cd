"""The C ABI on the CPU: the library loads, exports every function
include/pseg.h declares, and its host logic (meta flattening, lengths,
margins, sweeps) equals the oracle's.  No kernel runs here."""
import ctypes as C
import os
import re

import numpy as np

import processseg_amd as P
from processseg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"\b(pseg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 18
    for name in names:
        assert hasattr(L, name), name
    assert set(_lib.SYMBOLS) <= set(names)


def test_device_count_never_fails():
    assert _lib.lib().pseg_device_count() >= 0


def test_meta_flatten_equals_oracle(oracle):
    rng = np.random.default_rng(0)
    for rows, cols, first in [(1, 10, 0), (50, 2001, 0), (200, 501, 100), (7, 0, 3)]:
        nums = rng.integers(-5, cols + 300, rows).astype(np.int64)
        status = rng.integers(0, 4, rows).astype(np.uint8)
        failed = sorted(set(rng.integers(0, rows, rows // 5).tolist()))
        meta = P.ChunkMeta(rows, 2.0, numSamples=nums, status=status, failed=failed,
                           firstSample=first)
        for corr in (False, True):
            got = meta.row_len(cols, corr)
            want = oracle.meta_flatten(cols, nums, status, failed, first, corr)
            assert np.array_equal(got, want)


def test_meta_flatten_rejects_bad_failed_lists():
    import pytest
    meta = P.ChunkMeta(4, 2.0)
    meta.failed = np.array([2, 1], np.int64)
    with pytest.raises(ValueError):
        meta.row_len(10)
    meta.failed = np.array([4], np.int64)
    with pytest.raises(ValueError):
        meta.row_len(10)


def test_apply_checks_before_touching_samples():
    """A NULL plan or a block wider than the plan is refused by the ABI
    itself, with the reference's messages (pipeline.nim:619-641)."""
    L = _lib.lib()
    rc = L.pseg_plan_apply(None, None, None, 1, 1, None, None, 0, None)
    assert rc == _lib.E_INVALID and "plan" in _lib.last_error()
    rc = L.pseg_corr_apply(None, None, 1, 1, None, None, 1, None)
    assert rc == _lib.E_INVALID


def test_corr_compile_validation():
    import pytest
    pilot = P.generatePilot(P.linearSweep(6, 80, 12000, 2, P.initSweepTaper(300, 300)))
    with pytest.raises(ValueError, match="not shorter than the record"):
        P.initCorrelationPlan(np.float32, pilot, 2.0, 5000)
    with pytest.raises(ValueError, match="resample"):
        P.initCorrelationPlan(np.float32, pilot, 4.0, 20000)
    zero = P.Pilot(np.zeros(100), 2.0)
    with pytest.raises(ValueError, match="all zeros"):
        P.initCorrelationPlan(np.float32, zero, 2.0, 5000)


def test_corr_float64_too_long_is_refused():
    """A 20001-sample float64 record against a 12 s sweep needs an FFT whose
    float64 buffer does not fit one workgroup's LDS: the plan says so at
    compile (PSEG_E_UNSUPPORTED) instead of failing at apply."""
    import pytest
    pilot = P.generatePilot(P.linearSweep(6, 80, 12000, 2, P.initSweepTaper(300, 300)))
    with pytest.raises(P.UnsupportedOp, match="LDS"):
        P.initCorrelationPlan(np.float64, pilot, 2.0, 20001)


def test_pipeline_lds_limit_is_unsupported():
    """A float64 low-cut at 0.1-0.2 Hz on 10001-sample rows pads to a
    19200-point FFT, more LDS than one workgroup has: compile refuses it with
    PSEG_E_UNSUPPORTED, naming the op."""
    import pytest
    flow = P.initPipeline([P.demeanOp(), P.filterOp(P.lowCut(0.1, 0.2))])
    with pytest.raises(P.UnsupportedOp, match=r"op 1 \(.*LDS"):
        flow.compile(np.float64, 2.0, 10001)
    with pytest.raises(P.UnsupportedOp, match="beyond"):
        flow.compile(np.float32, 2.0, 12289)


def test_batch_checks_before_running():
    """pseg_batch_create validates every item as pseg_plan_apply would and
    names the item (no GPU needed: nothing is allocated before the checks)."""
    L = _lib.lib()
    h = C.c_void_p()
    assert L.pseg_batch_create(None, 1, C.byref(h)) == _lib.E_INVALID
    items = (_lib.BatchItemC * 2)()
    items[0] = _lib.BatchItemC(None, 16, 16, 1, 1, 16)
    assert L.pseg_batch_create(items, 2, C.byref(h)) == _lib.E_INVALID
    assert "item 0" in _lib.last_error()
    assert L.pseg_batch_launch(None, None) == _lib.E_INVALID
    assert L.pseg_batch_destroy(None) == _lib.OK


def test_server_checks_arguments():
    """pseg_server_* refuse NULL and malformed arguments before touching a
    device."""
    L = _lib.lib()
    h = C.c_void_p()
    assert L.pseg_server_create(None, C.byref(h)) == _lib.E_INVALID
    cfg = _lib.ServerConfigC()
    cfg.dtype, cfg.traces, cfg.samples, cfg.pitch = 0, 10, 100, 50
    assert L.pseg_server_create(C.byref(cfg), C.byref(h)) == _lib.E_INVALID
    assert "pitch" in _lib.last_error()
    t = C.c_uint64()
    assert L.pseg_server_submit(None, 0, 0, 1, 0, 1, C.byref(t)) == _lib.E_INVALID
    assert L.pseg_server_destroy(None) == _lib.OK


def test_decode_checks_before_touching_memory():
    """pseg_decode_samples refuses bad arguments before any device work (the
    pointers below are never dereferenced)."""
    L = _lib.lib()
    fake_raw, fake_dst = 1 << 20, 1 << 30
    bad = [
        (4, fake_raw, 4240, 240, 2, 1000, 0, fake_dst),        # not a SEG-Y format code
        (1, fake_raw, 4240, 240, 2, 1000, 7, fake_dst),        # unknown dtype
        (1, fake_raw, 4240, 240, -1, 1000, 0, fake_dst),       # negative rows
        (1, None, 4240, 240, 2, 1000, 0, fake_dst),            # NULL raw
        (1, fake_raw, 4000, 240, 2, 1000, 0, fake_dst),        # samples past the pitch
        (1, fake_raw, 4242, 240, 2, 1000, 0, fake_dst),        # 4-byte words misaligned
        (1, fake_raw + 2, 4240, 240, 2, 1000, 0, fake_dst),    # raw misaligned
        (1, fake_raw, 4240, 240, 2, 1000, 0, fake_raw + 64),   # dst over the traces
    ]
    for args in bad:
        assert L.pseg_decode_samples(*args, None) == _lib.E_INVALID, args
    # empty shapes are no-ops
    assert L.pseg_decode_samples(1, fake_raw, 4240, 240, 0, 1000, 0, fake_dst, None) == _lib.OK
    assert L.pseg_decode_samples(1, fake_raw, 4240, 240, 2, 0, 0, fake_dst, None) == _lib.OK


def test_spectrum_compile_checks_before_device_work():
    """pseg_spectrum_compile refuses bad plans (fft/spectrum.nim:77-101) before
    it touches the device."""
    import ctypes as C
    L = _lib.lib()
    h = C.c_void_p()
    for args in [(0, 2.0, 0, 0.1, 0, 1), (0, 0.0, 10, 0.1, 0, 1), (0, float("nan"), 10, 0.1, 0, 1),
                 (0, 2.0, 10, 0.6, 0, 1), (0, 2.0, 10, -0.1, 0, 1), (0, 2.0, 100, 0.1, 64, 1),
                 (0, 2.0, 100, 0.1, 101, 1), (7, 2.0, 100, 0.1, 0, 1)]:
        assert L.pseg_spectrum_compile(*args, C.byref(h)) == _lib.E_INVALID, args
    assert L.pseg_spectrum_average(None, None, 1, 1, None, 0, None, 0, None, None, None) == _lib.E_INVALID
    assert L.pseg_spectrum_workspace_size(None, 1, None) == _lib.E_INVALID


def test_spectrogram_compile_checks_before_device_work():
    """pseg_spectrogram_compile refuses bad plans (fft/spectrogram.nim:63-85)
    before it touches the device."""
    import ctypes as C
    L = _lib.lib()
    h = C.c_void_p()
    for args in [(0, 1, 10, 1.0, 100, 0), (0, 64, 0, 1.0, 100, 0), (0, 64, 10, 0.0, 100, 0),
                 (0, 64, 10, 1.0, 0, 0), (0, 64, 10, 1.0, 100, 32), (0, 64, 10, 1.0, 100, 77),
                 (3, 64, 10, 1.0, 100, 0)]:
        assert L.pseg_spectrogram_compile(*args, C.byref(h)) == _lib.E_INVALID, args
    assert L.pseg_spectrogram_apply(None, None, 0, None, None) == _lib.E_INVALID
