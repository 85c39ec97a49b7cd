"""Many records, each with its own compiled plan, as one unit of GPU work.

The reference applies a plan per record (``pool.apply(plan, src, dst, meta)``,
core/pool.nim:277-308, called in a loop over records as bench/viewersim.nim
does).  ``PipelineBatch`` takes the list of such calls once, keeps the work
table on the device, and runs it as one persistent launch per (dtype,
threads per row) class: rows are handed out heaviest first from a device
counter, so records of very different sizes and sample intervals balance
across the chip.  Every record's result equals ``plan.apply`` of that record
alone, bit for bit.  ``launch`` allocates nothing and copies nothing from the
host, so it can be captured in a CUDA/HIP graph and replayed.
"""
import ctypes as C
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, PipelinePlan, _stream_handle, sameInterval


@dataclass
class BatchItem:
    """One record: plan.apply(src, dst, meta) (dst may be src: in place).
    `row_len` (device int32, from ChunkMeta.device_row_len) may replace
    `meta`; then `keys` ([rows, 2] float64 device tensor of t0Ms, offset) and
    `first_sample` give what mutes and gain curves read from a meta."""
    plan: PipelinePlan
    src: object
    dst: object = None
    meta: Optional[ChunkMeta] = None
    row_len: object = None
    keys: object = None
    first_sample: int = 0


class PipelineBatch:
    def __init__(self, items: List[BatchItem]):
        import torch
        arr = (_lib.BatchItemC * max(len(items), 1))()
        keep = []
        for i, it in enumerate(items):
            dst = it.src if it.dst is None else it.dst
            for t, name in ((it.src, "src"), (dst, "dst")):
                if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dim() != 2 \
                        or not t.is_contiguous():
                    raise ValueError(f"item {i}: {name} must be a C-contiguous [rows, cols] "
                                     "GPU tensor")
                if np.dtype(str(t.dtype).replace("torch.", "")) != it.plan.T:
                    raise ValueError(f"item {i}: {name} holds {t.dtype}, the plan {it.plan.T}")
            if it.src.shape != dst.shape:
                raise ValueError(f"item {i}: src is {tuple(it.src.shape)}, dst {tuple(dst.shape)}")
            rows, cols = it.src.shape
            row_len = it.row_len
            keys, first = it.keys, int(it.first_sample)
            if row_len is None:
                if it.meta is None:
                    raise ValueError(f"item {i}: needs a ChunkMeta (or row_len)")
                if len(it.meta) != rows:
                    raise ValueError(f"item {i}: samples have {rows} rows, the meta "
                                     f"{len(it.meta)}")
                if not sameInterval(it.meta.dtMs, it.plan.dtMs):
                    raise ValueError(f"item {i}: the plan is for dtMs = {it.plan.dtMs}, the rows "
                                     f"have {it.meta.dtMs}")
                row_len = it.meta.device_row_len(cols, it.src.device)
                if it.plan.readsKeys:
                    keys = it.meta.device_keys(it.src.device)
                first = it.meta.firstSample
            if keys is not None and (not isinstance(keys, torch.Tensor) or not keys.is_cuda
                                     or keys.dtype != torch.float64
                                     or tuple(keys.shape) != (rows, 2)
                                     or not keys.is_contiguous()):
                raise ValueError(f"item {i}: keys must be a contiguous [rows, 2] float64 GPU tensor")
            arr[i] = _lib.BatchItemC(it.plan._h.value, it.src.data_ptr(), dst.data_ptr(), rows,
                                     cols, row_len.data_ptr(),
                                     keys.data_ptr() if keys is not None else None, first)
            keep += [it.plan, it.src, dst, row_len, keys]
        self._keep = keep  # the plans and buffers outlive the batch
        self._h = C.c_void_p()
        _lib.check(_lib.lib().pseg_batch_create(arr, len(items), C.byref(self._h)))
        info = _lib.BatchInfoC()
        _lib.check(_lib.lib().pseg_batch_info_get(self._h, C.byref(info)))
        self.items = int(info.items)
        self.launches = int(info.launches)
        self.rows = int(info.rows)
        self.samples = int(info.samples)
        self.threads = [int(info.threads[g]) for g in range(self.launches)]
        self.grid = [int(info.grid[g]) for g in range(self.launches)]

    def tune(self, repeats: int = 5, stream=None) -> dict:
        """Times each launch class with persistent grids of 1/2, 3/4 and all
        of the resident workgroups and with one workgroup per row, and keeps
        the fastest (pseg_batch_tune): the same bits, from a launch shape
        measured on this device.  Items must not run in place."""
        t = _lib.BatchTuningC()
        _lib.check(_lib.lib().pseg_batch_tune(self._h, int(repeats), _stream_handle(stream),
                                              C.byref(t)))
        info = _lib.BatchInfoC()
        _lib.check(_lib.lib().pseg_batch_info_get(self._h, C.byref(info)))
        self.grid = [int(info.grid[g]) for g in range(self.launches)]
        return {"grid": [int(t.grid[g]) for g in range(t.launches)],
                "direct": [bool(t.direct[g]) for g in range(t.launches)],
                "best_ms": [float(t.best_ms[g]) for g in range(t.launches)],
                "candidate_ms": [[float(x) for x in t.candidate_ms[g]] for g in range(t.launches)]}

    def launch(self, stream=None):
        """Run every item on `stream` (default: torch's current stream)."""
        _lib.check(_lib.lib().pseg_batch_launch(self._h, _stream_handle(stream)))

    def __del__(self):
        try:
            if self._h:
                _lib.lib().pseg_batch_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass
