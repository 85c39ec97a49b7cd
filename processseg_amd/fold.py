"""CMP binnings (stack/binning.nim) and fold maps (qc/fold.nim) with the
counting on the GPU.

```python
b = gridBinning(x0, y0, azimuthDeg=30, crosslineSpacing=25, inlineSpacing=25,
                inlines=200, crosslines=300)
m = initFoldMap(b, FoldOptions(offsetBounds=(0, 300, 800, 1500), azimuthSectors=6,
                               offsetRange=True))
m.addFold(meta)              # meta.traceKeys: [rows, 10] float64, traceKeyTable(...)
m.fold                       # [inlines, crosslines] int32 on the GPU
```

One GPU thread per trace computes its bin with the reference's float64
arithmetic and adds integer counts atomically: a map is the same whatever
order its traces are added in, in one block or many.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from typing import NamedTuple, Sequence

import numpy as np

from . import _lib
from .pipeline import ChunkMeta, _check_row_len, _stream_handle

bnCdp, bnLine, bnGrid, bnInlineCrossline = 0, 1, 2, 3

# The trace keys counting reads, the columns of a traceKeys table.
KEY_COLUMNS = ("srcX", "srcY", "rcvX", "rcvY", "cmpX", "cmpY", "offset", "ensemble", "inline",
               "crossline")


def traceKeyTable(rows: int, **columns) -> np.ndarray:
    """A [rows, 10] float64 table of the keys fold counting reads (columns
    KEY_COLUMNS, NaN where not given)."""
    t = np.full((rows, len(KEY_COLUMNS)), np.nan)
    for name, v in columns.items():
        if name not in KEY_COLUMNS:
            raise ValueError(f"unknown trace key {name!r}; keys: {', '.join(KEY_COLUMNS)}")
        t[:, KEY_COLUMNS.index(name)] = np.broadcast_to(np.asarray(v, np.float64), (rows,))
    return t


class BinningC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("inlines", C.c_int64),
                ("crosslines", C.c_int64)] + [
        (k, C.c_double) for k in ("first0", "step0", "first1", "step1", "x0", "y0", "ux", "uy",
                                  "vx", "vy", "size0", "size1", "max_off")]


class BinLocation(NamedTuple):
    """What a bin stands for (stack/binning.nim:68-73); NaN where undefined."""
    cdp: float
    inline: float
    crossline: float
    x: float
    y: float


@dataclasses.dataclass(frozen=True)
class Binning:
    """A bin definition (stack/binning.nim:55-66); build it with cdpBinning,
    lineBinning, gridBinning or inlineCrosslineBinning."""
    kind: int
    inlines: int
    crosslines: int
    first0: float = 0.0
    step0: float = 0.0
    first1: float = 0.0
    step1: float = 0.0
    x0: float = 0.0
    y0: float = 0.0
    ux: float = 0.0
    uy: float = 0.0
    vx: float = 0.0
    vy: float = 0.0
    size0: float = 0.0
    size1: float = 0.0
    maxOff: float = 0.0

    def __len__(self) -> int:
        return self.inlines * self.crosslines

    @property
    def is3D(self) -> bool:
        return self.kind in (bnGrid, bnInlineCrossline)

    def location(self, bin: int) -> BinLocation:
        """location / locationAt (stack/binning.nim:256-284)."""
        if bin < 0 or bin >= len(self):
            raise ValueError(f"bin {bin} is outside 0 ..< {len(self)}")
        nan = math.nan
        if self.kind == bnCdp:
            return BinLocation(self.first0 + bin * self.step0, nan, nan, nan, nan)
        if self.kind == bnLine:
            return BinLocation(self.first0 + float(bin) * self.step0, nan, nan,
                               self.x0 + float(bin) * self.size0 * self.ux,
                               self.y0 + float(bin) * self.size0 * self.uy)
        i, j = divmod(bin, self.crosslines)
        il = self.first0 + float(i) * self.step0
        xl = self.first1 + float(j) * self.step1
        if self.kind == bnInlineCrossline:
            return BinLocation(nan, il, xl, nan, nan)
        a, c = float(j) * self.size0, float(i) * self.size1
        return BinLocation(nan, il, xl, self.x0 + a * self.ux + c * self.vx,
                           self.y0 + a * self.uy + c * self.vy)

    def _c(self) -> BinningC:
        b = BinningC()
        b.kind, b.inlines, b.crosslines = self.kind, self.inlines, self.crosslines
        for f in ("first0", "step0", "first1", "step1", "x0", "y0", "ux", "uy", "vx", "vy",
                  "size0", "size1"):
            setattr(b, f, float(getattr(self, f)))
        b.max_off = float(self.maxOff)
        return b


def _finite(x) -> bool:
    return isinstance(x, (int, float)) and math.isfinite(x)


def _check_finite(x, what):
    if not _finite(x):
        raise ValueError(f"{what} must be finite, got {x}")


def _check_positive(x, what):
    if not _finite(x) or x <= 0:
        raise ValueError(f"{what} must be finite and positive, got {x}")


def _check_count(n, what):
    if not isinstance(n, (int, np.integer)) or n <= 0:
        raise ValueError(f"{what} must be positive, got {n}")


def _unit_at(azimuth_deg: float):
    # degToRad (d * (pi / 180)), then (sin, cos): azimuths clockwise from +y
    a = azimuth_deg * (math.pi / 180.0)
    return math.sin(a), math.cos(a)


def cdpBinning(firstCdp: float, count: int, step: float = 1.0) -> Binning:
    """cdpBinning (stack/binning.nim:80-89)."""
    _check_finite(firstCdp, "firstCdp")
    _check_count(count, "count")
    _check_positive(step, "step")
    return Binning(bnCdp, 1, int(count), first0=float(firstCdp), step0=float(step))


def lineBinning(x0: float, y0: float, azimuthDeg: float, binSize: float, count: int,
                firstCdp: float = 1.0, maxOffLine: float = math.inf,
                cdpStep: float = 1.0) -> Binning:
    """lineBinning (stack/binning.nim:95-121)."""
    _check_finite(x0, "x0")
    _check_finite(y0, "y0")
    _check_finite(azimuthDeg, "azimuthDeg")
    _check_finite(firstCdp, "firstCdp")
    _check_positive(binSize, "binSize")
    _check_positive(cdpStep, "cdpStep")
    _check_count(count, "count")
    if math.isnan(maxOffLine) or maxOffLine <= 0:
        raise ValueError(f"maxOffLine must be positive, got {maxOffLine}")
    ux, uy = _unit_at(azimuthDeg)
    return Binning(bnLine, 1, int(count), first0=float(firstCdp), step0=float(cdpStep),
                   x0=float(x0), y0=float(y0), ux=ux, uy=uy, size0=float(binSize),
                   maxOff=float(maxOffLine))


def gridBinning(x0: float, y0: float, azimuthDeg: float, crosslineSpacing: float,
                inlineSpacing: float, inlines: int, crosslines: int, firstInline: float = 1.0,
                firstCrossline: float = 1.0, clockwise: bool = True, inlineStep: float = 1.0,
                crosslineStep: float = 1.0) -> Binning:
    """gridBinning (stack/binning.nim:126-160)."""
    _check_finite(x0, "x0")
    _check_finite(y0, "y0")
    _check_finite(azimuthDeg, "azimuthDeg")
    _check_finite(firstInline, "firstInline")
    _check_finite(firstCrossline, "firstCrossline")
    _check_positive(crosslineSpacing, "crosslineSpacing")
    _check_positive(inlineSpacing, "inlineSpacing")
    _check_positive(inlineStep, "inlineStep")
    _check_positive(crosslineStep, "crosslineStep")
    _check_count(inlines, "inlines")
    _check_count(crosslines, "crosslines")
    if inlines > (2**63 - 1) // crosslines:
        raise ValueError("inlines × crosslines overflows")
    ux, uy = _unit_at(azimuthDeg)
    vx, vy = _unit_at(azimuthDeg + (90.0 if clockwise else -90.0))
    return Binning(bnGrid, int(inlines), int(crosslines), first0=float(firstInline),
                   step0=float(inlineStep), first1=float(firstCrossline),
                   step1=float(crosslineStep), x0=float(x0), y0=float(y0), ux=ux, uy=uy, vx=vx,
                   vy=vy, size0=float(crosslineSpacing), size1=float(inlineSpacing))


def inlineCrosslineBinning(firstInline: float, inlines: int, firstCrossline: float,
                           crosslines: int, inlineStep: float = 1.0,
                           crosslineStep: float = 1.0) -> Binning:
    """inlineCrosslineBinning (stack/binning.nim:162-175)."""
    _check_finite(firstInline, "firstInline")
    _check_finite(firstCrossline, "firstCrossline")
    _check_positive(inlineStep, "inlineStep")
    _check_positive(crosslineStep, "crosslineStep")
    _check_count(inlines, "inlines")
    _check_count(crosslines, "crosslines")
    if inlines > (2**63 - 1) // crosslines:
        raise ValueError("inlines × crosslines overflows")
    return Binning(bnInlineCrossline, int(inlines), int(crosslines), first0=float(firstInline),
                   step0=float(inlineStep), first1=float(firstCrossline),
                   step1=float(crosslineStep))


@dataclasses.dataclass(frozen=True)
class FoldOptions:
    """FoldOptions (qc/fold.nim:285-291)."""
    offsetBounds: Sequence[float] = ()
    azimuthSectors: int = 0
    sectorStartDeg: float = 0.0
    offsetRange: bool = False


def _short(x: float) -> str:
    # six significant digits (qc/fold.nim:204-207)
    return format(x, ".6g")


def _check_bin_budget(b: Binning, maxBins: int):
    """checkBinBudget (qc/fold.nim:308-323)."""
    if maxBins < 1:
        raise ValueError(f"initFoldMap: maxBins must be ≥ 1, got {maxBins}")
    if len(b) > maxBins:
        msg = (f"initFoldMap: {b.inlines} × {b.crosslines} bins (inlines × crosslines) are more "
               f"than maxBins ({maxBins})")
        if b.kind == bnGrid and b.inlines > 1 and b.crosslines > 1:
            o, c, i = b.location(0), b.location(1), b.location(b.crosslines)
            sc, si = math.hypot(c.x - o.x, c.y - o.y), math.hypot(i.x - o.x, i.y - o.y)
            msg += (f": bins of {_short(sc)} × {_short(si)} over {_short(sc * b.crosslines)} × "
                    f"{_short(si * b.inlines)}")
        raise ValueError(msg + "; check the bin sizes, or pass a larger maxBins")


class FoldMap:
    """Source-receiver pairs per bin (qc/fold.nim:293-303), the counts as
    GPU tensors: fold [inlines, crosslines] int32; offsetFold [classes,
    inlines, crosslines] and azimuthFold [sectors, inlines, crosslines]
    (empty without their options); offsetMin / offsetMax float32 (NaN:
    empty bin; empty without offsetRange)."""

    def __init__(self, binning: Binning, opts: FoldOptions = FoldOptions(),
                 maxBins: int = 1 << 26, device=None, _empty: bool = False):
        import torch
        _check_bin_budget(binning, maxBins)
        bnd = tuple(float(e) for e in opts.offsetBounds)
        if len(bnd) == 1:
            raise ValueError("initFoldMap: offsetBounds needs at least two edges")
        for i, e in enumerate(bnd):
            if not math.isfinite(e) or (i > 0 and e <= bnd[i - 1]):
                raise ValueError(f"initFoldMap: offsetBounds must be finite and strictly "
                                 f"ascending, got {list(bnd)}")
        if opts.azimuthSectors < 0:
            raise ValueError("initFoldMap: azimuthSectors must be ≥ 0")
        if not math.isfinite(opts.sectorStartDeg):
            raise ValueError("initFoldMap: sectorStartDeg must be finite")
        self.binning = binning
        self.options = FoldOptions(bnd, int(opts.azimuthSectors), float(opts.sectorStartDeg),
                                   bool(opts.offsetRange))
        dev = torch.device(device if device is not None else "cuda")
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        n, m = binning.inlines, binning.crosslines
        self._classes = max(len(bnd) - 1, 0)
        z = dict(dtype=torch.int32, device=dev)
        self.fold = torch.empty((n, m), **z)
        self.offsetFold = torch.empty((self._classes, n, m) if self._classes else (0,), **z)
        s = self.options.azimuthSectors
        self.azimuthFold = torch.empty((s, n, m) if s else (0,), **z)
        f = dict(dtype=torch.float32, device=dev)
        self.offsetMin = torch.empty((n, m) if opts.offsetRange else (0,), **f)
        self.offsetMax = torch.empty((n, m) if opts.offsetRange else (0,), **f)
        self._bounds = torch.tensor(bnd if bnd else (0.0,), dtype=torch.float64, device=dev)
        self._counters = torch.empty(4, dtype=torch.int64, device=dev)
        if not _empty:
            self.clear()

    def _ptr(self, t):
        return t.data_ptr() if t.numel() else None

    def clear(self, stream=None):
        """clear (qc/fold.nim:416-427): every count to 0, the offset range to
        NaN; the storage is kept."""
        _lib.check(_lib.lib().pseg_fold_clear(
            len(self.binning), self._classes, self.options.azimuthSectors,
            1 if self.options.offsetRange else 0, self.fold.data_ptr(), self._ptr(self.offsetFold),
            self._ptr(self.azimuthFold), self._ptr(self.offsetMin), self._ptr(self.offsetMax),
            self._counters.data_ptr(), _stream_handle(stream)))

    def addFold(self, meta: ChunkMeta = None, *, keys=None, row_len=None, stream=None):
        """addFold(m, meta) (qc/fold.nim:583-588): the live rows of `meta`
        (status live or unknown, not failed), keys from meta.traceKeys (or
        only the header offset without one); or a [rows, 10] float64 GPU
        tensor `keys` (columns KEY_COLUMNS) with `row_len` (int32, < 0: not
        live)."""
        import torch
        if keys is None:
            if meta is None:
                raise ValueError("addFold needs a ChunkMeta (or keys and row_len)")
            rows = len(meta)
            tk = getattr(meta, "traceKeys", None)
            if tk is None:
                tk = traceKeyTable(rows, offset=meta.offset)
            keys = torch.from_numpy(np.ascontiguousarray(tk, np.float64)).to(self.device)
            row_len = meta.device_row_len(0, self.device)
        else:
            rows = int(keys.shape[0]) if hasattr(keys, "shape") else -1
        if (not isinstance(keys, torch.Tensor) or keys.device != self.device
                or keys.dtype != torch.float64 or tuple(keys.shape) != (rows, len(KEY_COLUMNS))
                or not keys.is_contiguous()):
            raise ValueError(f"keys must be a contiguous [rows, {len(KEY_COLUMNS)}] float64 "
                             "tensor on the map's device")
        _check_row_len(row_len, rows, self.device)
        _lib.check(_lib.lib().pseg_fold_add(
            C.byref(self.binning._c()), self._bounds.data_ptr(), len(self.options.offsetBounds),
            self.options.azimuthSectors, self.options.sectorStartDeg,
            1 if self.options.offsetRange else 0, keys.data_ptr(), row_len.data_ptr(), rows,
            self.fold.data_ptr(), self._ptr(self.offsetFold), self._ptr(self.azimuthFold),
            self._ptr(self.offsetMin), self._ptr(self.offsetMax), self._counters.data_ptr(),
            _stream_handle(stream)))

    def clone(self) -> "FoldMap":
        """clone (qc/fold.nim:353-361): an independent copy."""
        c = FoldMap(self.binning, self.options, maxBins=max(len(self.binning), 1),
                    device=self.device, _empty=True)
        for name in ("fold", "offsetFold", "azimuthFold", "offsetMin", "offsetMax", "_counters"):
            getattr(c, name).copy_(getattr(self, name))
        return c

    def _count(self, i) -> int:
        return int(self._counters[i].item())

    @property
    def binned(self) -> int:
        """Pairs counted in a bin."""
        return self._count(0)

    @property
    def outside(self) -> int:
        """Pairs whose midpoint is in no bin."""
        return self._count(1)

    @property
    def noCoordinates(self) -> int:
        """Traces with neither a midpoint nor source and receiver coordinates."""
        return self._count(2)

    @property
    def notLive(self) -> int:
        """Rows left out as not live (dead, auxiliary, failed)."""
        return self._count(3)


def initFoldMap(binning: Binning, opts: FoldOptions = FoldOptions(), maxBins: int = 1 << 26,
                device=None) -> FoldMap:
    """initFoldMap (qc/fold.nim:325-351): ValueError as the reference raises
    (before any count is allocated)."""
    return FoldMap(binning, opts, maxBins, device)


def _same_value(a: float, b: float) -> bool:
    return a == b or (math.isnan(a) and math.isnan(b))


def _same_bins(a: Binning, b: Binning) -> bool:
    """sameBins (qc/fold.nim:603-610): kind, counts, first and last bin."""
    if a.kind != b.kind or a.inlines != b.inlines or a.crosslines != b.crosslines:
        return False
    for bin in (0, len(a) - 1):
        p, q = a.location(bin), b.location(bin)
        if not all(_same_value(u, v) for u, v in zip(p, q)):
            return False
    return True


def foldDifference(preplot: FoldMap, actual: FoldMap, stream=None):
    """foldDifference (qc/fold.nim:612-622): pre-plot minus actual fold per
    bin, [inlines, crosslines] int32 on the GPU."""
    import torch
    if not _same_bins(preplot.binning, actual.binning):
        raise ValueError("foldDifference: the maps have different binnings")
    out = torch.empty_like(preplot.fold)
    _lib.check(_lib.lib().pseg_fold_difference(preplot.fold.data_ptr(), actual.fold.data_ptr(),
                                               out.data_ptr(), out.numel(),
                                               _stream_handle(stream)))
    return out
