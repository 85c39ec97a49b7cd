"""Serving zoom windows of a device-resident survey (the C ABI's pseg_server).

A request is the reference viewer's zoomWindow (tests/utils/viewersim.nim:
167-191): the shown samples widened by the flow's requiredMargin and clamped
by decodeWindow, the plan for that window width compiled on first use and
cached like planFor (viewersim.nim:53-62), the plan applied to the window's
rows, the shown samples returned.  The server queues requests and serves them
by continuous batching (every admitted request of the queue in one persistent
launch, two launches in flight), admitting a request when its result fits in
the server's device arena.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .pipeline import Pipeline, _dtype_code, _ops_c, _stream_handle


@dataclass
class ServeResult:
    status: int
    rows: int
    cols: int
    first: int              # the decoded window: samples first .. first + window - 1
    window: int
    batch_requests: int     # requests in the launch that served this one
    plan_compiled: bool     # the request paid for its plan's compile
    generation: int
    submit_ns: int
    start_ns: int
    done_ns: int

    @property
    def latency_ms(self) -> float:
        return (self.done_ns - self.submit_ns) / 1e6


class Server:
    def __init__(self, survey, trace_len, dtMs: float, *, arena_bytes: int = 1 << 30,
                 max_batch_requests: int = 256, max_batch_rows: int = 0,
                 max_inflight: int = 2, trace_keys=None):
        """`survey`: a [traces, samples] GPU tensor (C-contiguous rows);
        `trace_len`: int32 GPU tensor, each trace's numSamples (-1: not
        live); `trace_keys`: optional [traces, 2] float64 GPU tensor of each
        trace's (t0Ms, offset), which mutes and gain curves read (None: t0Ms
        0, offset NaN).  All must outlive the server."""
        import torch
        if not isinstance(survey, torch.Tensor) or not survey.is_cuda or survey.dim() != 2 \
                or survey.stride(1) != 1:
            raise ValueError("survey must be a [traces, samples] GPU tensor with unit column stride")
        if not isinstance(trace_len, torch.Tensor) or trace_len.dtype != torch.int32 \
                or trace_len.numel() != survey.shape[0] or not trace_len.is_cuda:
            raise ValueError("trace_len must be an int32 GPU tensor with one value per trace")
        self.T = np.dtype(str(survey.dtype).replace("torch.", ""))
        if trace_keys is not None and (
                not isinstance(trace_keys, torch.Tensor) or not trace_keys.is_cuda
                or trace_keys.dtype != torch.float64 or not trace_keys.is_contiguous()
                or tuple(trace_keys.shape) != (survey.shape[0], 2)):
            raise ValueError("trace_keys must be a contiguous [traces, 2] float64 GPU tensor")
        self._keep = (survey, trace_len, trace_keys)
        cfg = _lib.ServerConfigC()
        cfg.dtype = _dtype_code(self.T)
        cfg.traces, cfg.samples = survey.shape
        cfg.max_batch_requests = max_batch_requests
        cfg.pitch = survey.stride(0)
        cfg.survey = survey.data_ptr()
        cfg.trace_len = trace_len.data_ptr()
        cfg.dt_ms = dtMs
        cfg.arena_bytes = arena_bytes
        cfg.max_batch_rows = max_batch_rows
        cfg.max_inflight = max_inflight
        cfg.trace_keys = trace_keys.data_ptr() if trace_keys is not None else None
        self.traces, self.samples, self.dtMs = int(survey.shape[0]), int(survey.shape[1]), dtMs
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_server_create(C.byref(cfg), C.byref(self._h)))

    def add_flow(self, flow: Pipeline) -> int:
        arr = _ops_c(flow.ops)
        idx = C.c_int32()
        _lib.check(_lib.lib().pseg_server_add_flow(self._h, arr, len(flow.ops), C.byref(idx)))
        return int(idx.value)

    def submit(self, flow: int, trace0: int, ntraces: int, shown_first: int,
               shown_last: int) -> int:
        t = C.c_uint64()
        _lib.check(_lib.lib().pseg_server_submit(self._h, flow, trace0, ntraces, shown_first,
                                                 shown_last, C.byref(t)))
        return int(t.value)

    def wait(self, ticket: int, timeout_ms: int = -1) -> ServeResult:
        r = _lib.ServeResultC()
        _lib.check(_lib.lib().pseg_server_wait(self._h, ticket, timeout_ms, C.byref(r)))
        if r.status != _lib.OK:
            _lib.check(r.status)
        return ServeResult(int(r.status), int(r.rows), int(r.cols), int(r.first),
                           int(r.window), int(r.batch_requests), bool(r.plan_compiled),
                           int(r.generation), int(r.submit_ns), int(r.start_ns), int(r.done_ns))

    def read(self, ticket: int, out=None, stream=None):
        """The request's samples as a [rows, cols] GPU tensor (copied)."""
        import torch
        res = self.wait(ticket)
        if out is None:
            out = torch.empty(res.rows, res.cols, dtype=self._keep[0].dtype,
                              device=self._keep[0].device)
        _lib.check(_lib.lib().pseg_server_read(self._h, ticket, out.data_ptr(),
                                               _stream_handle(stream)))
        return out

    def release(self, ticket: int):
        _lib.check(_lib.lib().pseg_server_release(self._h, ticket))

    def info(self) -> dict:
        i = _lib.ServerInfoC()
        _lib.check(_lib.lib().pseg_server_info_get(self._h, C.byref(i)))
        return {k: int(getattr(i, k)) for k, _ in _lib.ServerInfoC._fields_}

    def close(self):
        if self._h:
            _lib.lib().pseg_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
