"""engine: Options, Stats, contexts, sources and sinks, and the two entry
points every operation funnels into — ProcessN (internal/exec/process.go:28)
and ProcessChunked (internal/exec/chunked.go:33) — over the HIP library.

Programming errors raise EnginePanic before anything is written, with the
reference's messages (its panics). Returned values mirror Go's `error`:
None on success, a Canceled or SourceError instance otherwise.
"""
from __future__ import annotations

import ctypes
import os
import threading
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .raster import Float32Raster, RasterPanic, _is_torch, _ptr

BYTES_PER_CELL = 4


class EnginePanic(RasterPanic):
    """The reference panics: a programming error, raised before any write."""


class DeviceError(RuntimeError):
    """The HIP runtime failed (a failure mode the CPU reference does not have)."""


class Canceled(Exception):
    """context.Canceled: the context was done before every unit was written."""

    def __init__(self, msg: str = "context canceled"):
        super().__init__(msg)


class SourceError(Exception):
    """A source or sink failed; wraps the operand and tile position like
    internal/exec/chunked.go:379-393."""


# -- context -----------------------------------------------------------------

class Context:
    """A cancellable context. The library polls its flag between units of
    the plan (bands or tiles), never inside a kernel."""

    def __init__(self):
        self._flag = ctypes.c_int32(0)

    def cancel(self) -> None:
        self._flag.value = 1

    def done(self) -> bool:
        return self._flag.value != 0

    def err(self) -> Optional[Exception]:
        return Canceled() if self.done() else None

    @property
    def flag_address(self) -> int:
        return ctypes.addressof(self._flag)


def Background() -> Context:
    return Context()


def WithCancel(parent: Optional[Context] = None):
    ctx = Context()
    return ctx, ctx.cancel


# -- Options and Stats ---------------------------------------------------------

@dataclass
class Stats:
    """engine.Stats (engine/stats.go:77-104)."""
    Cells: int = 0
    Ideal: int = 0
    Tiles: int = 0
    Bands: int = 0
    SourceRead: int = 0
    SinkWritten: int = 0
    KernelRead: int = 0
    KernelWritten: int = 0

    def Total(self) -> int:
        return self.SourceRead + self.KernelRead + self.KernelWritten + self.SinkWritten

    def Amplification(self) -> float:
        return 0.0 if self.Ideal == 0 else self.Total() / self.Ideal

    def BytesPerCell(self) -> float:
        return 0.0 if self.Cells == 0 else self.Total() / self.Cells

    def Halo(self, inputs: int) -> int:
        if inputs <= 0:
            return 0
        return self.KernelRead - self.Cells * BYTES_PER_CELL * inputs

    def Add(self, o: "Stats") -> None:
        for f in ("Cells", "Ideal", "Tiles", "Bands", "SourceRead", "SinkWritten", "KernelRead",
                  "KernelWritten"):
            setattr(self, f, getattr(self, f) + getattr(o, f))

    def _add_c(self, s: _lib.Stats) -> None:
        self.Cells += s.cells
        self.Ideal += s.ideal
        self.Tiles += s.tiles
        self.Bands += s.bands
        self.SourceRead += s.source_read
        self.SinkWritten += s.sink_written
        self.KernelRead += s.kernel_read
        self.KernelWritten += s.kernel_written


@dataclass
class Options:
    """engine.Options (engine/engine.go:7-28). Workers is a scheduling hint:
    the device decides how the plan's units run; results never change."""
    TileWidth: int = 0
    TileHeight: int = 0
    Workers: int = 0
    Stats: Optional[Stats] = None


# -- kernel descriptors ------------------------------------------------------

@dataclass
class KernelDesc:
    """A resolved kernel: what the reference's kernel constructors build
    (terrain/*.go, focal/*.go, transfer/*.go), as the C ABI's strata_kernel."""
    op: int
    radius: int
    flags: int = 0
    k: Sequence[float] = ()
    taps: Optional[np.ndarray] = None
    nin: int = 1
    nout: int = 1
    name: str = ""

    def to_c(self):
        kc = _lib.Kernel()
        kc.op, kc.radius, kc.flags = self.op, self.radius, self.flags
        for i, v in enumerate(self.k):
            kc.k[i] = float(np.float32(v))
        taps = None
        if self.taps is not None and len(self.taps):
            taps = np.ascontiguousarray(self.taps, dtype=np.float32)
            kc.ntaps = len(taps)
            kc.taps = taps.ctypes.data
        else:
            kc.ntaps = 0
            kc.taps = None
        return kc, taps  # keep taps alive for the call


class FanOut:
    """exec.NewPipeline(1, stages, outs) where stage i is parts[i] reading
    the pipeline's one input and writing output i: the pipeline
    terrain.newFeatures builds (terrain/features.go:103-117). The radius
    is the largest part's; each output keeps its own ring and validity
    (internal/exec/kernel.go:94-110)."""

    def __init__(self, parts: Sequence[KernelDesc], name: str = "engine.Pipeline"):
        if not parts:
            raise EnginePanic("engine: pipeline has no stages")
        for i, p in enumerate(parts):
            if p is None:
                raise EnginePanic(f"engine: stage {i} has a nil kernel")
            if p.nin != 1 or p.nout != 1:
                raise EnginePanic(f"engine: stage {i} takes {p.nin} inputs and {p.nout} outputs; a fan-out "
                                  "stage takes one of each")
        self.parts = list(parts)
        self.radius = max(p.radius for p in self.parts)
        self.nin = 1
        self.nout = len(self.parts)
        self.name = name

    def to_c(self):
        arr = (_lib.Kernel * len(self.parts))()
        keep = []
        for i, p in enumerate(self.parts):
            kc, taps = p.to_c()
            arr[i] = kc
            keep.append(taps)
        return arr, keep


class Pipeline:
    """exec.NewPipeline(1, stages, [last]) over a chain: stage i reads stage
    i-1's value, stage 0 the pipeline's one input, and the last stage's
    value is its output (internal/exec/pipeline.go:166-187), the pipeline
    graph.Slope -> graph.Lookup -> graph.Reclass lowers to
    (graph/graph.go:250-287). Its radius is the sum of the stage radii
    (radii add along a chain), which is also its output's edge ring and
    validity erosion; the output is what the stages write as separate
    calls. The library takes the point stages after a radius-1 terrain
    stage into that stage's store."""

    def __init__(self, stages: Sequence[KernelDesc], name: str = "engine.Pipeline"):
        if not stages:
            raise EnginePanic("engine: pipeline has no stages")
        for i, k in enumerate(stages):
            if k is None:
                raise EnginePanic(f"engine: stage {i} has a nil kernel")
            if k.nin != 1 or k.nout != 1:
                raise EnginePanic(f"engine: stage {i} takes {k.nin} inputs and {k.nout} outputs; a chain stage "
                                  "takes one of each")
        self.stages = list(stages)
        self.radius = sum(k.radius for k in self.stages)
        self.nin = 1
        self.nout = 1
        self.name = name

    def to_c(self):
        arr = (_lib.Kernel * len(self.stages))()
        keep = []
        for i, k in enumerate(self.stages):
            kc, taps = k.to_c()
            arr[i] = kc
            keep.append(taps)
        return arr, keep


# -- overlap checks (internal/overlap/overlap.go) ------------------------------

def _span(r: Float32Raster) -> int:
    return (r.Height - 1) * r.Stride + r.Width


def _spans_meet(a0: int, an: int, b0: int, bn: int) -> bool:
    return a0 < b0 + bn and b0 < a0 + an


def data_spans(a: Float32Raster, b: Float32Raster) -> bool:
    return _spans_meet(_ptr(a.Data), _span(a) * 4, _ptr(b.Data), _span(b) * 4)


def _bit_addr(r: Float32Raster) -> int:
    return _ptr(r.Valid) * 8 + r.ValidOffset


def bit_spans(a: Float32Raster, b: Float32Raster) -> bool:
    if a.Valid is None or b.Valid is None:
        return False
    return _spans_meet(_bit_addr(a), _span(a), _bit_addr(b), _span(b))


DISJOINT, SAME, PARTIAL = 0, 1, 2


def _relate(a0, a_stride, b0, b_stride, w, h) -> int:
    if not _spans_meet(a0, (h - 1) * a_stride + w, b0, (h - 1) * b_stride + w):
        return DISJOINT
    if a0 == b0:
        return SAME if (a_stride == b_stride or h == 1) else PARTIAL
    if a_stride != b_stride:
        return PARTIAL
    stride = a_stride
    diff = a0 - b0
    dy, dx = int(diff / stride), diff - int(diff / stride) * stride  # Go truncating division
    if dx < 0:
        dy, dx = dy - 1, dx + stride

    def rows_meet(d):
        return -h < d < h

    if (dx < w and rows_meet(dy)) or (dx > stride - w and rows_meet(dy + 1)):
        return PARTIAL
    return DISJOINT


def data_relation(a: Float32Raster, b: Float32Raster) -> int:
    a0, b0 = _ptr(a.Data), _ptr(b.Data)
    if (a0 - b0) % 4 != 0:
        return PARTIAL if _spans_meet(a0, _span(a) * 4, b0, _span(b) * 4) else DISJOINT
    return _relate(a0 // 4, a.Stride, b0 // 4, b.Stride, a.Width, a.Height)


def bits_relation(a: Float32Raster, b: Float32Raster) -> int:
    if a.Valid is None or b.Valid is None:
        return DISJOINT
    return _relate(_bit_addr(a), a.Stride, _bit_addr(b), b.Stride, a.Width, a.Height)


# -- checks (internal/exec/process.go:37-135) --------------------------------------

def _check_kernel(ndst: int, nsrc: int, k: KernelDesc, opts: Options) -> None:
    if k is None:
        raise EnginePanic("engine: nil kernel")
    if k.radius < 0:
        raise EnginePanic(f"engine: kernel radius {k.radius} is negative")
    if nsrc != k.nin or ndst != k.nout:
        raise EnginePanic(f"engine: kernel takes {k.nin} inputs and {k.nout} outputs, got {nsrc} and {ndst}")
    if opts.TileWidth < 0 or opts.TileHeight < 0 or opts.Workers < 0:
        raise EnginePanic(f"engine: negative Options {opts}")


def _require_raster(name: str, i: int, r: Float32Raster, ref: Float32Raster, ref_name: str) -> None:
    err = r.Validate()
    if err is not None:
        raise EnginePanic(f"engine: {name}[{i}]: {err}")
    if r.Width != ref.Width or r.Height != ref.Height:
        raise EnginePanic(f"engine: {name}[{i}] is {r.Width}×{r.Height}, {ref_name} is {ref.Width}×{ref.Height}")


def check(dst: List[Float32Raster], src: List[Float32Raster], k: KernelDesc, opts: Options) -> None:
    _check_kernel(len(dst), len(src), k, opts)
    r = k.radius
    for i, d in enumerate(dst):
        _require_raster("dst", i, d, dst[0], "dst[0]")
    masked = False
    for i, s in enumerate(src):
        _require_raster("src", i, s, dst[0], "dst[0]")
        masked = masked or s.Valid is not None
    for i, d in enumerate(dst):
        if masked and d.Valid is None:
            raise EnginePanic(
                f"engine: an input has a validity mask but dst[{i}].Valid is nil; allocate it with "
                "raster.NewFloat32Like, or set Valid on its root raster before windowing")
    for i, d in enumerate(dst):
        for j in range(i + 1, len(dst)):
            meet = data_spans(d, dst[j]) if r > 0 else data_relation(d, dst[j]) != DISJOINT
            if meet:
                raise EnginePanic(f"engine: dst[{i}] and dst[{j}] share Data")
            meet = bit_spans(d, dst[j]) if r > 0 else bits_relation(d, dst[j]) != DISJOINT
            if meet:
                raise EnginePanic(f"engine: dst[{i}] and dst[{j}] share validity bits")
        for j, s in enumerate(src):
            if r > 0:
                if data_spans(d, s):
                    raise EnginePanic(f"engine: dst[{i}] and src[{j}] share Data; a kernel with radius {r} "
                                      "would read neighbours it has already overwritten")
                if bit_spans(d, s):
                    raise EnginePanic(f"engine: dst[{i}] and src[{j}] share validity bits")
                continue
            rel = data_relation(d, s)
            if rel == PARTIAL:
                raise EnginePanic(f"engine: dst[{i}] overlaps src[{j}] at a different offset or stride")
            if rel == SAME and k.nout > 1:
                raise EnginePanic(f"engine: dst[{i}] is src[{j}]; only a kernel with one output runs in place")
            if bits_relation(d, s) == PARTIAL:
                raise EnginePanic(f"engine: dst[{i}] validity bits overlap src[{j}]'s at a different offset "
                                  "or stride")


# -- the calls ----------------------------------------------------------------

def _stream_handle(rasters) -> int:
    for r in rasters:
        if r is not None and r.on_device:
            import torch
            return torch.cuda.current_stream(r.Data.device).cuda_stream
    return 0


def _c_options(opts: Options, ctx: Optional[Context], stream: int, async_: bool = False) -> _lib.Options:
    o = _lib.Options()
    o.tile_width, o.tile_height, o.workers = opts.TileWidth, opts.TileHeight, opts.Workers
    o.stream = stream or None
    o.cancel = ctx.flag_address if ctx is not None else None
    o.async_ = 1 if async_ else 0
    return o


def _raise_status(rc: int) -> None:
    msg = _lib.last_error()
    if rc == _lib.EINVAL:
        raise EnginePanic(msg)
    raise DeviceError(msg or f"strata: status {rc}")


def ProcessN(ctx: Optional[Context], dst: List[Float32Raster], src: List[Float32Raster], k: KernelDesc,
             opts: Optional[Options] = None, *, async_: bool = False) -> Optional[Exception]:
    """exec.ProcessN: run k over every cell of dst from src."""
    opts = opts or Options()
    check(dst, src, k, opts)
    L = _lib.lib()
    kc, taps = k.to_c()
    cd = (_lib.Raster * len(dst))(*[d.to_c() for d in dst])
    cs = (_lib.Raster * len(src))(*[s.to_c() for s in src])
    st = _lib.Stats()
    co = _c_options(opts, ctx if (ctx is not None and not async_) else None, _stream_handle(dst + src), async_)
    if isinstance(k, FanOut):
        rc = L.strata_process_fanout(kc, len(k.parts), cd, cs, ctypes.byref(co), ctypes.byref(st))
    elif isinstance(k, Pipeline):
        rc = L.strata_process_chain(kc, len(k.stages), cd, cs, ctypes.byref(co), ctypes.byref(st))
    else:
        rc = L.strata_process(ctypes.byref(kc), cd, len(dst), cs, len(src), ctypes.byref(co), ctypes.byref(st))
    del taps
    if rc < 0:
        _raise_status(rc)
    if opts.Stats is not None:
        opts.Stats._add_c(st)
    if rc == _lib.CANCELLED:
        return Canceled()
    return None


def PlanStats(k: KernelDesc, width: int, height: int, opts: Optional[Options] = None,
              chunked: bool = False) -> Stats:
    """The Stats a completed call adds, in closed form."""
    opts = opts or Options()
    kc, taps = k.to_c()
    co = _c_options(opts, None, 0)
    st = _lib.Stats()
    if isinstance(k, FanOut):
        rc = _lib.lib().strata_plan_stats_fanout(kc, len(k.parts), width, height, ctypes.byref(co),
                                                 1 if chunked else 0, ctypes.byref(st))
    elif isinstance(k, Pipeline):
        rc = _lib.lib().strata_plan_stats_chain(kc, len(k.stages), width, height, ctypes.byref(co),
                                                1 if chunked else 0, ctypes.byref(st))
    else:
        rc = _lib.lib().strata_plan_stats(ctypes.byref(kc), k.nout, k.nin, width, height, ctypes.byref(co),
                                          1 if chunked else 0, ctypes.byref(st))
    if rc < 0:
        _raise_status(rc)
    out = Stats()
    out._add_c(st)
    return out


# -- sources and sinks (engine/source.go) -----------------------------------------

class RasterSource:
    """engine.RasterSource: Size, Masked and ReadWindow(ctx, dst, x, y)."""

    def Size(self):
        raise NotImplementedError

    def Masked(self) -> bool:
        raise NotImplementedError

    def ReadWindow(self, ctx, dst: Float32Raster, x: int, y: int) -> Optional[Exception]:
        raise NotImplementedError


class RasterSink:
    def Size(self):
        raise NotImplementedError

    def Masked(self) -> bool:
        raise NotImplementedError

    def WriteWindow(self, ctx, src: Float32Raster, x: int, y: int) -> Optional[Exception]:
        raise NotImplementedError


def _copy_cells(dst: Float32Raster, src: Float32Raster, valid: bool) -> None:
    """Host copy of cells (and bits) between equal-size rasters (engine/source.go:152-183)."""
    from .raster import MaskSet, MaskGet
    for y in range(dst.Height):
        dst.Data[y * dst.Stride:y * dst.Stride + dst.Width] = src.Data[y * src.Stride:y * src.Stride + src.Width]
    if valid and dst.Valid is not None:
        _copy_validity(dst, src)


def _copy_validity(dst: Float32Raster, src: Float32Raster) -> None:
    from .raster import unpack_bits
    dv = dst.Valid
    for y in range(dst.Height):
        if src.Valid is None:
            bits = np.ones(dst.Width, dtype=bool)
        else:
            bits = unpack_bits(src.Valid, src.ValidOffset + y * src.Stride, src.Width)
        _put_bits(dv, dst.ValidOffset + y * dst.Stride, bits)


def _put_bits(m: np.ndarray, off: int, bits: np.ndarray) -> None:
    """Write a run of bits into a host uint64 mask, leaving other bits alone."""
    n = len(bits)
    if n == 0:
        return
    w0, w1 = off >> 6, (off + n - 1) >> 6
    words = m[w0:w1 + 1].copy()
    allb = np.unpackbits(words.view(np.uint8), bitorder="little").astype(bool)
    s = off - (w0 << 6)
    allb[s:s + n] = bits
    m[w0:w1 + 1] = np.packbits(allb.reshape(-1, 8), axis=1, bitorder="little").reshape(-1).view(np.uint64)


class MemorySource(RasterSource):
    """A source over a raster in memory (engine/source.go:76-107). A raster
    in device memory is read in place by the device tile queue."""

    def __init__(self, r: Float32Raster):
        err = r.Validate()
        if err is not None:
            raise EnginePanic(f"engine: NewMemorySource: {err}")
        self.r = r

    def Raster(self) -> Float32Raster:
        return self.r

    def Size(self):
        return self.r.Width, self.r.Height

    def Masked(self) -> bool:
        return self.r.Valid is not None

    def ReadWindow(self, ctx, dst: Float32Raster, x: int, y: int):
        win = self.r.Window(x, y, dst.Width, dst.Height)
        _copy_cells(dst, win if not win.on_device else win.to("cpu"), dst.Valid is not None)
        return None


class MemorySink(RasterSink):
    def __init__(self, r: Float32Raster):
        err = r.Validate()
        if err is not None:
            raise EnginePanic(f"engine: NewMemorySink: {err}")
        self.r = r
        self._mu = threading.Lock()

    def Raster(self) -> Float32Raster:
        return self.r

    def Size(self):
        return self.r.Width, self.r.Height

    def Masked(self) -> bool:
        return self.r.Valid is not None

    def WriteWindow(self, ctx, src: Float32Raster, x: int, y: int):
        dst = self.r.Window(x, y, src.Width, src.Height)
        _copy_cells(dst, src, False)
        if dst.Valid is not None:
            with self._mu:
                _copy_validity(dst, src)
        return None


NewMemorySource = MemorySource
NewMemorySink = MemorySink


def _plain_memory(x) -> bool:
    """A memory source or sink whose window copies are MemorySource's and
    MemorySink's own (a subclass that overrides them is a callback)."""
    if isinstance(x, MemorySource):
        return type(x).ReadWindow is MemorySource.ReadWindow
    if isinstance(x, MemorySink):
        return type(x).WriteWindow is MemorySink.WriteWindow
    return False


def _resident(x) -> Optional[Float32Raster]:
    if _plain_memory(x) and x.r.on_device:
        return x.r
    return None


def _in_memory(x) -> Optional[Float32Raster]:
    """The raster of a memory source or sink, in device or host memory: the
    library reads device rasters in place and copies windows of host ones
    itself, as MemorySource.ReadWindow and MemorySink.WriteWindow do."""
    if _plain_memory(x):
        return x.r
    return None


# -- raw files (engine/raw.go, engine/rawfile.go) ---------------------------------------

@dataclass
class RawOptions:
    """engine.RawOptions (engine/raw.go:16-26): with HasFill, a cell equal to
    Fill (as a float; a NaN Fill matches any NaN) is NoData."""
    Fill: float = 0.0
    HasFill: bool = False


class RawFile:
    """engine.RawFile (engine/rawfile.go:14-40): one file through one or more
    handles; ReadAt and WriteAt take the handles in turn. The library reads
    and writes it through the first handle's descriptor."""

    def __init__(self, fds: List[int]):
        self._fds = list(fds)
        self._next = 0
        self._lock = threading.Lock()

    def fileno(self) -> int:
        return self._fds[0]

    def _fd(self) -> int:
        with self._lock:
            fd = self._fds[self._next % len(self._fds)]
            self._next += 1
        return fd

    def ReadAt(self, n: int, off: int) -> bytes:
        return os.pread(self._fd(), n, off)

    def WriteAt(self, b, off: int) -> int:
        return os.pwrite(self._fd(), b, off)

    def Handles(self) -> int:
        return len(self._fds)

    def Sync(self) -> None:
        os.fsync(self._fds[0])

    def Close(self) -> None:
        for fd in self._fds:
            os.close(fd)
        self._fds = []


def OpenRawFile(name: str, flag: int, perm: int = 0o644, handles: int = 1) -> RawFile:
    """engine.OpenRawFile (engine/rawfile.go:42-65): only the first handle
    creates or truncates; O_APPEND is rejected, WriteAt cannot use it."""
    if flag & os.O_APPEND:
        raise OSError(f"open {name}: engine: OpenRawFile: O_APPEND does not allow WriteAt")
    fds = []
    try:
        for i in range(max(1, handles)):
            if i == 1:
                flag &= ~(os.O_CREAT | os.O_EXCL | os.O_TRUNC)
            fds.append(os.open(name, flag, perm))
    except OSError:
        for fd in fds:
            os.close(fd)
        raise
    return RawFile(fds)


def CreateRawFile(name: str, size: int, perm: int = 0o644, handles: int = 1) -> RawFile:
    """engine.CreateRawFile (engine/rawfile.go:67-110), without the memory
    mapping: a new file of size bytes, its blocks allocated at once."""
    f = OpenRawFile(name, os.O_RDWR | os.O_CREAT | os.O_TRUNC, perm, handles)
    try:
        os.ftruncate(f.fileno(), size)
    except OSError:
        f.Close()
        raise
    return f


def _require_raw_size(fn: str, nil: bool, width: int, height: int) -> None:
    # engine/raw.go:275-285
    if nil:
        raise EnginePanic(f"engine: {fn}: nil file")
    if width <= 0 or height <= 0:
        raise EnginePanic(f"engine: {fn}: dimensions must be positive, got {width}×{height}")
    if width > (2 ** 63 - 1) // 4 // height:
        raise EnginePanic(f"engine: {fn}: {width}×{height} cells do not fit in a file")


def _require_region(fn: str, name: str, r: Float32Raster, x: int, y: int, w: int, h: int) -> None:
    # engine/source.go:193-201
    err = r.Validate()
    if err is not None:
        raise EnginePanic(f"engine: {fn}: {name}: {err}")
    if x < 0 or y < 0 or x > w - r.Width or y > h - r.Height:
        raise EnginePanic(f"engine: {fn}: {r.Width}×{r.Height} region at ({x}, {y}) outside {w}×{h} raster")


def _read_at(f, n: int, off: int) -> bytes:
    if hasattr(f, "ReadAt"):
        return f.ReadAt(n, off)
    return os.pread(f.fileno(), n, off)


def _write_at(f, b, off: int) -> int:
    if hasattr(f, "WriteAt"):
        return f.WriteAt(b, off)
    return os.pwrite(f.fileno(), b, off)


def _fill_validity(cells: np.ndarray, fill: float) -> np.ndarray:
    """vec.ValidBits (internal/vec/validity.go:11-19) for one row."""
    f = np.float32(fill)
    if f != f:
        return ~np.isnan(cells)
    return cells != f


class RawSource(RasterSource):
    """engine.RawSource (engine/raw.go:28-118): a headerless file of
    width·height little-endian float32 cells, row-major. A chunked run over
    a file with a descriptor has the library read it and derive validity
    from Fill on the device."""

    def __init__(self, f, width: int, height: int, opts: RawOptions = RawOptions()):
        _require_raw_size("NewRawSource", f is None, width, height)
        self.f, self.w, self.h, self.opts = f, width, height, opts

    def Size(self):
        return self.w, self.h

    def Masked(self) -> bool:
        return self.opts.HasFill

    def ReadWindow(self, ctx, dst: Float32Raster, x: int, y: int):
        _require_region("RawSource.ReadWindow", "dst", dst, x, y, self.w, self.h)
        if self.opts.HasFill and dst.Valid is None:
            raise EnginePanic("engine: RawSource.ReadWindow: the source has a fill value and dst has no validity mask")
        if ctx is not None and ctx.done():
            return ctx.err()
        for row in range(dst.Height):
            n = 4 * dst.Width
            b = _read_at(self.f, n, 4 * ((y + row) * self.w + x))
            if len(b) < n:
                return SourceError(f"engine: raw source: reading row {y + row}: unexpected EOF")
            cells = np.frombuffer(b, dtype="<f4").astype(np.float32)
            dst.Data[row * dst.Stride:row * dst.Stride + dst.Width] = cells
            if dst.Valid is not None:
                bits = _fill_validity(cells, self.opts.Fill) if self.opts.HasFill else np.ones(dst.Width, bool)
                _put_bits(dst.Valid, dst.ValidOffset + row * dst.Stride, bits)
        return None


class RawSink(RasterSink):
    """engine.RawSink (engine/raw.go:164-245): writes rows into the file;
    with HasFill, Fill goes into every invalid cell first (in the source
    raster itself, as fillInvalid does)."""

    def __init__(self, f, width: int, height: int, opts: RawOptions = RawOptions()):
        _require_raw_size("NewRawSink", f is None, width, height)
        self.f, self.w, self.h, self.opts = f, width, height, opts

    def Size(self):
        return self.w, self.h

    def Masked(self) -> bool:
        return self.opts.HasFill

    def WriteWindow(self, ctx, src: Float32Raster, x: int, y: int):
        _require_region("RawSink.WriteWindow", "src", src, x, y, self.w, self.h)
        if src.Valid is not None and not self.opts.HasFill:
            raise EnginePanic("engine: RawSink.WriteWindow: src has a validity mask and the sink has no fill value")
        if ctx is not None and ctx.done():
            return ctx.err()
        from .raster import unpack_bits
        for row in range(src.Height):
            s0 = row * src.Stride
            cells = src.Data[s0:s0 + src.Width]
            if src.Valid is not None:
                bad = ~unpack_bits(src.Valid, src.ValidOffset + s0, src.Width)
                cells[bad] = np.float32(self.opts.Fill)
            b = np.ascontiguousarray(cells, dtype="<f4").tobytes()
            n = _write_at(self.f, b, 4 * ((y + row) * self.w + x))
            if n < len(b):
                return SourceError(f"engine: raw sink: writing row {y + row}: short write")
        return None


def NewRawSource(f, width: int, height: int, opts: RawOptions = RawOptions()) -> RawSource:
    return RawSource(f, width, height, opts)


def NewRawSink(f, width: int, height: int, opts: RawOptions = RawOptions()) -> RawSink:
    return RawSink(f, width, height, opts)


def _raw(x) -> Optional[_lib.Raw]:
    """The library's descriptor of a raw source or sink over a file with a
    descriptor, or None."""
    if not isinstance(x, (RawSource, RawSink)) or not hasattr(x.f, "fileno"):
        return None
    try:
        fd = x.f.fileno()
    except (OSError, ValueError):
        return None
    r = _lib.Raw()
    r.fd, r.has_fill, r.fill, r.offset = fd, 1 if x.opts.HasFill else 0, float(np.float32(x.opts.Fill)), 0
    return r


def _read_callback(ctx, s: RasterSource, j: int, errors: List[Exception]):
    """The read callback of source j: ReadWindow into the library's pinned
    buffer; an error or a panic is kept in errors and reported as a
    nonzero return."""
    def read(user, buf, ww, hh, stride, valid, voff, x, y):
        try:
            n = (hh - 1) * stride + ww
            data = np.ctypeslib.as_array(ctypes.cast(buf, ctypes.POINTER(ctypes.c_float)), shape=(n,))
            vm = None
            if valid:
                vm = np.ctypeslib.as_array(ctypes.cast(valid, ctypes.POINTER(ctypes.c_uint64)),
                                           shape=((voff + n + 63) >> 6,))
            v = Float32Raster(data, ww, hh, stride, vm, voff)
            err = s.ReadWindow(ctx, v, x, y)
            if err is not None:
                errors.append(SourceError(f"engine: reading src[{j}] at ({x}, {y}): {err}"))
                return 1
            if vm is not None and not s.Masked():
                _copy_validity(v, Float32Raster(data, ww, hh, stride, None, 0))
            return 0
        except Exception as e:  # a panicking source
            errors.append(e)
            return 1
    return _lib.READ_FN(read)


# -- statistics folded into the pass (graph/sinks.go) ------------------------------------

def _go_min(a: float, b: float) -> float:
    """Go's builtin min of two float32s: NaN if either is, -0 < +0."""
    if a != a or b != b:
        return float("nan")
    if a == b:
        return a if np.signbit(a) else b
    return min(a, b)


def _go_max(a: float, b: float) -> float:
    if a != a or b != b:
        return float("nan")
    if a == b:
        return b if np.signbit(a) else a
    return max(a, b)


def _set_limbs(limbs, x: int) -> None:
    n = len(limbs)
    x &= (1 << (64 * n)) - 1  # two's complement over n limbs
    for i in range(n):
        limbs[i] = (x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF


def _empty_fold() -> "_lib.Fold":
    f = _lib.Fold()
    f.min, f.max = float("inf"), float("-inf")
    return f


def _combine_folds(a: "_lib.Fold", b: "_lib.Fold") -> None:
    """Folds partial b into a (summary.Reducer.Combine, internal/summary/
    summary.go:62-75): counts and tallies add, the extremes take Go's min
    and max, the exact sums add as integers."""
    if b.count == 0:
        return
    if a.count == 0:
        a.min, a.max = b.min, b.max
    else:
        a.min, a.max = _go_min(a.min, b.min), _go_max(a.max, b.max)
    a.count += b.count
    a.nan += b.nan
    a.pos_inf += b.pos_inf
    a.neg_inf += b.neg_inf
    a.not_neg_zero = 1 if (a.not_neg_zero or b.not_neg_zero) else 0
    _set_limbs(a.sum, _lib.limbs_to_int(a.sum) + _lib.limbs_to_int(b.sum))
    _set_limbs(a.sumsq, _lib.limbs_to_int(a.sumsq) + _lib.limbs_to_int(b.sumsq))


STATS_FOLD = _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS  # what reduce.Stats folds


class TeeSink(RasterSink):
    """graph's teeSink (graph/sinks.go:17-71): every window goes to each of
    sinks, and is first folded into a summary of the whole output (the
    partial reduce.Stats combines; what selects the reduction, 0 for
    none). With no sinks nothing is written: statistics alone write no
    raster.

    ProcessChunked and the fan-out form fold on the device, inside the pass
    that computes the output, before any sink sees the tile; WriteWindow
    called directly folds the window with Reduce. Fold() is the combined
    partial of every window so far; reduce.Folded finishes it."""

    def __init__(self, width: int, height: int, masked: bool, sinks: Sequence[RasterSink] = (),
                 what: int = STATS_FOLD):
        if width <= 0 or height <= 0:
            raise EnginePanic(f"engine: tee sink dimensions must be positive, got {width}×{height}")
        for i, t in enumerate(sinks):
            if t is None:
                raise EnginePanic(f"engine: tee sink's sink {i} is nil")
            if t.Size() != (width, height):
                raise EnginePanic(f"engine: tee sink is {width}×{height}, its sink {i} is {t.Size()}")
            # The library hands the tee's one sink the tee's validity as is.
            if t.Masked() != bool(masked):
                raise EnginePanic(f"engine: tee sink is {'' if masked else 'not '}Masked but its sink {i} is "
                                  f"{'' if t.Masked() else 'not '}Masked")
        self.w, self.h, self.masked = width, height, bool(masked)
        self.sinks = list(sinks)
        self.what = what
        self._fold = _empty_fold()
        self._mu = threading.Lock()

    def Size(self):
        return self.w, self.h

    def Masked(self) -> bool:
        return self.masked

    def Fold(self) -> "_lib.Fold":
        return self._fold

    def _add(self, f: "_lib.Fold") -> None:
        with self._mu:
            _combine_folds(self._fold, f)

    def WriteWindow(self, ctx, src: Float32Raster, x: int, y: int):
        # Fold first: a sink may overwrite the Data of invalid cells.
        if self.what:
            f, err = Reduce(ctx, [src], self.what, Options(Workers=1))
            if err is not None:
                return err
            self._add(f)
        for t in self.sinks:
            err = t.WriteWindow(ctx, src, x, y)
            if err is not None:
                return err
        return None


class _TeeWrites(RasterSink):
    """A tee's writes alone (its fold runs on the device)."""

    def __init__(self, tee: TeeSink):
        self.tee = tee

    def Size(self):
        return self.tee.Size()

    def Masked(self) -> bool:
        return self.tee.Masked()

    def WriteWindow(self, ctx, src: Float32Raster, x: int, y: int):
        for t in self.tee.sinks:
            err = t.WriteWindow(ctx, src, x, y)
            if err is not None:
                return err
        return None


def _sink_target(d: RasterSink) -> Optional[RasterSink]:
    """What the library writes for sink d: d itself, or for a tee its one
    sink, a writer over its sinks, or None when it only folds."""
    if not isinstance(d, TeeSink):
        return d
    if len(d.sinks) == 1:
        return d.sinks[0]
    if not d.sinks:
        return None if d.what else _TeeWrites(d)
    return _TeeWrites(d)


def ProcessChunked(ctx: Optional[Context], dst: List[RasterSink], src: List[RasterSource], k: KernelDesc,
                   opts: Optional[Options] = None, *, async_: bool = False) -> Optional[Exception]:
    """exec.ProcessChunked: tiles read with their halo, computed on the
    device, written to the sinks; a memory source or sink in device memory
    is read or written in place (the device tile queue)."""
    opts = opts or Options()
    _check_kernel(len(dst), len(src), k, opts)
    for i, d in enumerate(dst):
        if d is None:
            raise EnginePanic(f"engine: dst[{i}] is a nil sink")
    for i, s in enumerate(src):
        if s is None:
            raise EnginePanic(f"engine: src[{i}] is a nil source")
    w, h = dst[0].Size()
    if w <= 0 or h <= 0:
        raise EnginePanic(f"engine: dst[0] is {w}×{h}; sizes must be positive")
    masked = False
    for i, s in enumerate(src):
        sw, sh = s.Size()
        if (sw, sh) != (w, h):
            raise EnginePanic(f"engine: src[{i}] is {sw}×{sh}, dst[0] is {w}×{h}")
        masked = masked or s.Masked()
    for i, d in enumerate(dst):
        sw, sh = d.Size()
        if (sw, sh) != (w, h):
            raise EnginePanic(f"engine: dst[{i}] is {sw}×{sh}, dst[0] is {w}×{h}")
        if masked and not d.Masked():
            raise EnginePanic(f"engine: a source is Masked but dst[{i}] is not; its validity would be lost")
    targets = [_sink_target(d) for d in dst]
    for i, d in enumerate(targets):
        if not isinstance(d, MemorySink):
            continue
        a = d.r
        for j in range(i + 1, len(dst)):
            o = targets[j]
            if isinstance(o, MemorySink) and (data_spans(a, o.r) or _words_meet(a, o.r)):
                raise EnginePanic(f"engine: memory sinks dst[{i}] and dst[{j}] share memory")
        for j, s in enumerate(src):
            if isinstance(s, MemorySource) and (data_spans(a, s.r) or _words_meet(a, s.r)):
                raise EnginePanic(
                    f"engine: memory sink dst[{i}] shares memory with memory source src[{j}]; tiles would read "
                    "cells other tiles write (use a Tiled entry point on the rasters instead)")

    keep = []  # objects that must outlive the call
    errors: List[Exception] = []

    def make_read(s: RasterSource, j: int):
        return _read_callback(ctx, s, j, errors)

    def make_write(d: RasterSink, j: int):
        def write(user, buf, ww, hh, stride, valid, voff, x, y):
            try:
                n = (hh - 1) * stride + ww
                data = np.ctypeslib.as_array(ctypes.cast(buf, ctypes.POINTER(ctypes.c_float)), shape=(n,))
                vm = None
                if valid:
                    vm = np.ctypeslib.as_array(ctypes.cast(valid, ctypes.POINTER(ctypes.c_uint64)),
                                               shape=((voff + n + 63) >> 6,))
                err = d.WriteWindow(ctx, Float32Raster(data, ww, hh, stride, vm, voff), x, y)
                if err is not None:
                    errors.append(SourceError(f"engine: writing dst[{j}] at ({x}, {y}): {err}"))
                    return 1
                return 0
            except Exception as e:
                errors.append(e)
                return 1
        return _lib.WRITE_FN(write)

    csrc = (_lib.Source * len(src))()
    for j, s in enumerate(src):
        sw, sh = s.Size()
        c = csrc[j]
        c.width, c.height, c.masked = sw, sh, 1 if s.Masked() else 0
        res = _in_memory(s)
        raw = _raw(s)
        if res is not None:
            rc_ = res.to_c()
            keep.append(rc_)
            c.resident = ctypes.pointer(rc_)
        elif raw is not None:
            keep.append(raw)
            c.raw = ctypes.pointer(raw)
        else:
            fn = make_read(s, j)
            keep.append(fn)
            c.read = fn
    cdst = (_lib.Sink * len(dst))()
    folds = {}  # dst index -> the device's partial of that tee
    for j, d in enumerate(dst):
        sw, sh = d.Size()
        c = cdst[j]
        c.width, c.height, c.masked = sw, sh, 1 if d.Masked() else 0
        if isinstance(d, TeeSink) and d.what:
            folds[j] = _lib.Fold()
            c.fold = ctypes.pointer(folds[j])
            c.fold_what = d.what
        d = targets[j]
        if d is None:
            continue
        res = _in_memory(d)
        raw = _raw(d)
        if res is not None:
            rc_ = res.to_c()
            keep.append(rc_)
            c.resident = ctypes.pointer(rc_)
        elif raw is not None:
            keep.append(raw)
            c.raw = ctypes.pointer(raw)
        else:
            fn = make_write(d, j)
            keep.append(fn)
            c.write = fn
    kc, taps = k.to_c()
    rasters = [x.r for x in list(src) + targets if isinstance(x, (MemorySource, MemorySink))]
    co = _c_options(opts, ctx if not async_ else None, _stream_handle(rasters), async_)
    st = _lib.Stats()
    if isinstance(k, FanOut):
        rc = _lib.lib().strata_process_fanout_chunked(kc, len(k.parts), cdst, csrc, ctypes.byref(co),
                                                      ctypes.byref(st))
    elif isinstance(k, Pipeline):
        rc = _lib.lib().strata_process_chain_chunked(kc, len(k.stages), cdst, csrc, ctypes.byref(co),
                                                     ctypes.byref(st))
    else:
        rc = _lib.lib().strata_process_chunked(ctypes.byref(kc), cdst, len(dst), csrc, len(src),
                                               ctypes.byref(co), ctypes.byref(st))
    del keep, taps
    if opts.Stats is not None:
        opts.Stats._add_c(st)
    if rc == 0:
        for j, f in folds.items():
            dst[j]._add(f)
    if errors and not isinstance(errors[0], SourceError):
        raise errors[0]
    if rc == _lib.EIO:
        return errors[0] if errors else SourceError(_lib.last_error())
    if rc < 0:
        _raise_status(rc)
    if rc == _lib.CANCELLED:
        return Canceled()
    return None


def _words_meet(a: Float32Raster, b: Float32Raster) -> bool:
    if a.Valid is None or b.Valid is None:
        return False

    def words(r):
        first = _bit_addr(r) >> 6
        return first, ((_bit_addr(r) + _span(r) + 63) >> 6) - first

    a0, an = words(a)
    b0, bn = words(b)
    return _spans_meet(a0, an, b0, bn)


# -- reductions (internal/exec/reduce.go, reducechunked.go) ------------------------------

def Reduce(ctx: Optional[Context], src: List[Float32Raster], what: int,
           opts: Optional[Options] = None):
    """exec.Reduce over one raster with the device fold selected by what
    (_lib.REDUCE_*): (the combined partial as a _lib.Fold, None), or
    (None, Canceled()) when ctx was done before the fold ran."""
    opts = opts or Options()
    if len(src) != 1:
        raise EnginePanic(f"engine: reducer takes 1 inputs, got {len(src)}")
    if opts.TileWidth < 0 or opts.TileHeight < 0 or opts.Workers < 0:
        raise EnginePanic(f"engine: negative Options {opts}")
    err = src[0].Validate()
    if err is not None:
        raise EnginePanic(f"engine: src[0]: {err}")
    cs = src[0].to_c()
    co = _c_options(opts, ctx, _stream_handle(src))
    out = _lib.Fold()
    st = _lib.Stats()
    rc = _lib.lib().strata_reduce(what, ctypes.byref(cs), ctypes.byref(co), ctypes.byref(out), ctypes.byref(st))
    if rc < 0:
        _raise_status(rc)
    if opts.Stats is not None:
        opts.Stats._add_c(st)
    if rc == _lib.CANCELLED:
        return None, Canceled()
    return out, None


def ReduceChunked(ctx: Optional[Context], src: List[RasterSource], what: int,
                  opts: Optional[Options] = None):
    """exec.ReduceChunked over one source: tiles read by the source's
    ReadWindow (a memory source in device memory is read in place) and
    folded on the device. (fold, None), or (None, err) for cancellation or
    a source's error."""
    opts = opts or Options()
    if len(src) != 1:
        raise EnginePanic(f"engine: reducer takes 1 inputs, got {len(src)}")
    if opts.TileWidth < 0 or opts.TileHeight < 0 or opts.Workers < 0:
        raise EnginePanic(f"engine: negative Options {opts}")
    s = src[0]
    if s is None:
        raise EnginePanic("engine: src[0] is a nil source")
    w, h = s.Size()
    if w <= 0 or h <= 0:
        raise EnginePanic(f"engine: src[0] is {w}×{h}; sizes must be positive")
    errors: List[Exception] = []
    keep = []
    c = _lib.Source()
    c.width, c.height, c.masked = w, h, 1 if s.Masked() else 0
    res = _resident(s)
    raw = _raw(s)
    if res is not None:
        rc_ = res.to_c()
        keep.append(rc_)
        c.resident = ctypes.pointer(rc_)
    elif raw is not None:
        keep.append(raw)
        c.raw = ctypes.pointer(raw)
    else:
        fn = _read_callback(ctx, s, 0, errors)
        keep.append(fn)
        c.read = fn
    co = _c_options(opts, ctx, _stream_handle([res] if res is not None else []))
    out = _lib.Fold()
    st = _lib.Stats()
    rc = _lib.lib().strata_reduce_chunked(what, ctypes.byref(c), ctypes.byref(co), ctypes.byref(out),
                                          ctypes.byref(st))
    del keep
    if opts.Stats is not None:
        opts.Stats._add_c(st)
    if errors and not isinstance(errors[0], SourceError):
        raise errors[0]
    if rc == _lib.EIO:
        return None, (errors[0] if errors else SourceError(_lib.last_error()))
    if rc < 0:
        _raise_status(rc)
    if rc == _lib.CANCELLED:
        return None, Canceled()
    return out, None
