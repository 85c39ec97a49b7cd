"""SEG-Y and SEG-D sample decode on the GPU: rows of big-endian samples as a file holds
them (a trace header, then the samples) to a [rows, cols] float32 or float64
panel, on the device next to the survey, before a plan runs on it.

```python
raw = torch.from_numpy(file_bytes).cuda()          # traces of 240 + 4 * ns bytes
panel = decodeSamples(raw, fmtIbm32, ns, rows=ntraces)   # float32 [ntraces, ns]
plan.apply(panel, meta)
```

The formats are SEG-Y rev 2's data sample format codes (binary header bytes
3225-3226): IBM float32, int32, int16, IEEE float32 and int8, and SEG-D's
demultiplexed codes 8036 (24-bit integer), 8038 (int32), 8058 (IEEE float32)
and 8080 (IEEE float64), big-endian.  IBM, integer and IEEE words decode
exactly, but int32 and float64 to float32, which round to nearest.  This is the decode parseSEG runs on its decode workers before the
prefetcher's transform hands a chunk to the plan (batch.nim:267-309)."""
from __future__ import annotations

import numpy as np

from . import _lib
from .pipeline import _stream_handle

fmtIbm32, fmtInt32, fmtInt16, fmtIeee32, fmtInt8 = 1, 2, 3, 5, 8
# SEG-D demultiplexed data sample formats (general header 1, bytes 3-4)
fmtSegd8036, fmtSegd8038, fmtSegd8058, fmtSegd8080 = 8036, 8038, 8058, 8080
SAMPLE_BYTES = {fmtIbm32: 4, fmtInt32: 4, fmtInt16: 2, fmtIeee32: 4, fmtInt8: 1,
                fmtSegd8036: 3, fmtSegd8038: 4, fmtSegd8058: 4, fmtSegd8080: 8}
TRACE_HEADER_BYTES = 240


def decodeSamples(raw, fmt: int, samples: int, rows: int | None = None,
                  offset: int = TRACE_HEADER_BYTES, pitch: int | None = None,
                  dtype=np.float32, out=None, stream=None):
    """Decode `rows` traces from the device byte tensor `raw` (uint8): trace
    r's sample j is the word at byte r·pitch + offset + j·bytes (pitch
    default: offset + bytes·samples, a SEG-Y file's traces).  Returns (or
    fills) a C-contiguous [rows, samples] tensor of dtype on raw's device.
    ValueError for an unknown format, a shape that does not fit `raw`, or
    misaligned 4- or 8-byte words."""
    import torch
    if not isinstance(raw, torch.Tensor) or not raw.is_cuda or raw.dtype != torch.uint8 \
            or not raw.is_contiguous():
        raise ValueError("raw must be a contiguous uint8 GPU tensor")
    if fmt not in SAMPLE_BYTES:
        raise ValueError(f"unknown sample format {fmt} (SEG-Y codes 1, 2, 3, 5, 8; SEG-D 8036, "
                         "8038, 8058, 8080)")
    b = SAMPLE_BYTES[fmt]
    if pitch is None:
        pitch = offset + b * samples
    if rows is None:
        rows = raw.numel() // pitch if pitch > 0 else 0
    if rows > 0 and (rows - 1) * pitch + offset + b * samples > raw.numel():
        raise ValueError(f"{rows} traces of pitch {pitch} do not fit {raw.numel()} bytes")
    tdt = torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
    if out is None:
        out = torch.empty((rows, samples), dtype=tdt, device=raw.device)
    elif (out.dtype != tdt or tuple(out.shape) != (rows, samples) or not out.is_contiguous()
          or not out.is_cuda):
        raise ValueError(f"out must be a contiguous [{rows}, {samples}] {tdt} GPU tensor")
    _lib.check(_lib.lib().pseg_decode_samples(
        fmt, raw.data_ptr(), pitch, offset, rows, samples, 0 if tdt == torch.float32 else 1,
        out.data_ptr(), _stream_handle(stream)))
    return out
