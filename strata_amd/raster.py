"""raster: the Float32Raster data model (raster/raster.go, raster/window.go,
raster/mask.go), over numpy arrays (host memory) or torch tensors (device
memory, or host).

Cell (x, y) is data[y*Stride + x]; len(data) == (Height-1)*Stride + Width.
Valid is None (every cell valid) or a uint64 bitmap shared by a raster and
all its windows: the validity of data[i] is bit ValidOffset+i, LSB first,
set = valid. Data under a cleared bit is unspecified.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _lib

try:  # torch is the device-memory plumbing; the host path works without it
    import torch
except ImportError:  # pragma: no cover - torch is in the image
    torch = None


class RasterPanic(ValueError):
    """A programming error: the reference panics with the same message."""


def _is_torch(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor)


def _ptr(a) -> int:
    if a is None:
        return 0
    if _is_torch(a):
        return a.data_ptr()
    return a.ctypes.data


def _on_device(a) -> bool:
    return _is_torch(a) and a.is_cuda


class Float32Raster:
    """Mirrors raster.Float32Raster (raster/raster.go:41-54)."""

    __slots__ = ("Data", "Width", "Height", "Stride", "Valid", "ValidOffset")

    def __init__(self, Data, Width: int, Height: int, Stride: Optional[int] = None, Valid=None,
                 ValidOffset: int = 0):
        self.Data = Data
        self.Width = int(Width)
        self.Height = int(Height)
        self.Stride = int(Width if Stride is None else Stride)
        self.Valid = Valid
        self.ValidOffset = int(ValidOffset)

    # -- geometry ---------------------------------------------------------
    def __repr__(self) -> str:
        where = "device" if self.on_device else "host"
        return (f"Float32Raster({self.Width}x{self.Height}, stride {self.Stride}, "
                f"{'masked' if self.Valid is not None else 'unmasked'}, {where})")

    @property
    def on_device(self) -> bool:
        return _on_device(self.Data)

    def Index(self, x: int, y: int) -> int:
        if not (0 <= x < self.Width and 0 <= y < self.Height):
            raise RasterPanic(f"raster: cell ({x}, {y}) outside {self.Width}×{self.Height} raster")
        return y * self.Stride + x

    def Row(self, y: int):
        if not (0 <= y < self.Height):
            raise RasterPanic(f"raster: row {y} out of range [0, {self.Height})")
        return self.Data[y * self.Stride:y * self.Stride + self.Width]

    def Window(self, x: int, y: int, w: int, h: int) -> "Float32Raster":
        """A view of the w×h region at (x, y) (raster/window.go:12-33)."""
        if w <= 0 or h <= 0:
            raise RasterPanic(f"raster: window dimensions must be positive, got {w}×{h}")
        if x < 0 or y < 0 or x > self.Width - w or y > self.Height - h:
            raise RasterPanic(f"raster: window {w}×{h} at ({x}, {y}) outside {self.Width}×{self.Height} raster")
        start = y * self.Stride + x
        end = start + (h - 1) * self.Stride + w
        win = Float32Raster(self.Data[start:end], w, h, self.Stride, self.Valid, 0)
        if self.Valid is not None:
            win.ValidOffset = self.ValidOffset + start
        return win

    def IsValid(self, x: int, y: int) -> bool:
        i = self.Index(x, y)
        return self.Valid is None or MaskGet(self.Valid, self.ValidOffset + i)

    def SetValid(self, x: int, y: int, valid: bool) -> None:
        i = self.Index(x, y)
        if self.Valid is None:
            raise RasterPanic("raster: SetValid on a raster without a validity mask")
        MaskSet(self.Valid, self.ValidOffset + i, valid)

    def Validate(self) -> Optional[str]:
        """The error string raster.Validate returns, or None (raster/raster.go:172-203)."""
        if self.Width <= 0 or self.Height <= 0:
            return f"raster: dimensions must be positive, got {self.Width}×{self.Height}"
        if self.Stride < self.Width:
            return f"raster: stride {self.Stride} is less than width {self.Width}"
        n = (self.Height - 1) * self.Stride + self.Width
        if len(self.Data) < n:
            return f"raster: data has {len(self.Data)} cells, need {n}"
        if self.Valid is not None:
            if self.ValidOffset < 0:
                return f"raster: negative ValidOffset {self.ValidOffset}"
            bits = len(self.Valid) * 64
            if self.ValidOffset > bits - n:
                return (f"raster: mask has {bits} bits, need {self.ValidOffset + n} "
                        f"(offset {self.ValidOffset} + {n} cells)")
        return None

    # -- C ABI ------------------------------------------------------------
    def to_c(self) -> _lib.Raster:
        r = _lib.Raster()
        r.data = _ptr(self.Data)
        r.width, r.height, r.stride = self.Width, self.Height, self.Stride
        r.valid = _ptr(self.Valid) if self.Valid is not None else None
        r.valid_offset = self.ValidOffset
        r.mem = _lib.MEM_DEVICE if self.on_device else _lib.MEM_HOST
        return r

    # -- copies between host and device (tests, benchmarks) ----------------
    def to(self, device) -> "Float32Raster":
        """A copy with the same layout (stride, mask offset) on `device`:
        'cpu' gives numpy arrays, anything else torch tensors."""
        data, valid = self.Data, self.Valid
        if str(device) == "cpu" or device == "numpy":
            d = data.detach().cpu().numpy().copy() if _is_torch(data) else np.array(data, copy=True)
            v = None
            if valid is not None:
                v = (valid.detach().cpu().numpy().view(np.uint64).copy() if _is_torch(valid)
                     else np.array(valid, copy=True))
            return Float32Raster(d, self.Width, self.Height, self.Stride, v, self.ValidOffset)
        td = (data.to(device) if _is_torch(data) else torch.from_numpy(np.ascontiguousarray(data)).to(device))
        tv = None
        if valid is not None:
            tv = (valid.to(device) if _is_torch(valid)
                  else torch.from_numpy(np.ascontiguousarray(valid).view(np.int64)).to(device))
        return Float32Raster(td.contiguous(), self.Width, self.Height, self.Stride, tv, self.ValidOffset)

    def numpy_data(self) -> np.ndarray:
        d = self.Data
        return d.detach().cpu().numpy() if _is_torch(d) else np.asarray(d)

    def numpy_valid(self) -> Optional[np.ndarray]:
        v = self.Valid
        if v is None:
            return None
        return v.detach().cpu().numpy().view(np.uint64) if _is_torch(v) else np.asarray(v)


# -- constructors (raster/raster.go:56-104) --------------------------------

def _require_shape(width: int, height: int, stride: int) -> int:
    if width <= 0 or height <= 0:
        raise RasterPanic(f"raster: dimensions must be positive, got {width}×{height}")
    if stride < width:
        raise RasterPanic(f"raster: stride {stride} is less than width {width}")
    return (height - 1) * stride + width


def NewFloat32(width: int, height: int, data) -> Float32Raster:
    return NewFloat32Stride(width, height, width, data)


def NewFloat32Stride(width: int, height: int, stride: int, data) -> Float32Raster:
    n = _require_shape(width, height, stride)
    if len(data) < n:
        raise RasterPanic(f"raster: data has {len(data)} cells, need {n} for {width}×{height} with stride {stride}")
    return Float32Raster(data[:n], width, height, stride)


def NewFloat32Like(r: Float32Raster) -> Float32Raster:
    """A zeroed compact raster of r's size, in r's memory; with an all-valid
    mask when r has one."""
    n = _require_shape(r.Width, r.Height, r.Width)
    if _is_torch(r.Data):
        data = torch.zeros(n, dtype=torch.float32, device=r.Data.device)
        valid = NewMask(n, device=r.Data.device) if r.Valid is not None else None
    else:
        data = np.zeros(n, dtype=np.float32)
        valid = NewMask(n) if r.Valid is not None else None
    return Float32Raster(data, r.Width, r.Height, r.Width, valid, 0)


def NewDevice(width: int, height: int, stride: Optional[int] = None, masked: bool = False,
              device="cuda") -> Float32Raster:
    """An uninitialised raster in device memory (an all-valid mask if masked)."""
    stride = width if stride is None else stride
    n = _require_shape(width, height, stride)
    data = torch.empty(n, dtype=torch.float32, device=device)
    return Float32Raster(data, width, height, stride, NewMask(n, device=device) if masked else None, 0)


# -- masks (raster/mask.go) --------------------------------------------------

def MaskWords(n: int) -> int:
    return (n + 63) >> 6


def NewMask(n: int, device=None):
    """All n cells valid, unused bits of the last word cleared. numpy
    uint64 on the host, torch int64 (same bits) on a device."""
    if n < 0:
        raise RasterPanic(f"raster: negative mask size {n}")
    m = np.full(MaskWords(n), np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    if n & 63:
        m[-1] = np.uint64((1 << (n & 63)) - 1)
    if device is not None and str(device) != "cpu":
        return torch.from_numpy(m.view(np.int64)).to(device)
    return m


def MaskGet(m, i: int) -> bool:
    w = int(m[i >> 6]) & 0xFFFFFFFFFFFFFFFF
    return (w >> (i & 63)) & 1 != 0


def MaskSet(m, i: int, valid: bool) -> None:
    w = int(m[i >> 6]) & 0xFFFFFFFFFFFFFFFF
    b = 1 << (i & 63)
    w = (w | b) if valid else (w & ~b)
    if _is_torch(m):
        m[i >> 6] = w - (1 << 64) if w >= (1 << 63) else w
    else:
        m[i >> 6] = np.uint64(w)


def pack_bits(valid: np.ndarray, nbits: Optional[int] = None) -> np.ndarray:
    """Pack a flat bool array into LSB-first uint64 words."""
    valid = np.asarray(valid, dtype=bool).ravel()
    n = len(valid) if nbits is None else nbits
    words = MaskWords(n)
    padded = np.zeros(words * 64, dtype=bool)
    padded[:len(valid)] = valid
    b = np.packbits(padded.reshape(-1, 8), axis=1, bitorder="little").reshape(-1)
    return b.view(np.uint64).copy()


def unpack_bits(m, off: int, n: int) -> np.ndarray:
    """Bits [off, off+n) of a mask as a bool array (only the words holding
    them are read)."""
    if n <= 0:
        return np.zeros(0, dtype=bool)
    w0, w1 = off >> 6, (off + n - 1) >> 6
    if _is_torch(m):
        arr = m[w0:w1 + 1].detach().cpu().numpy().view(np.uint64)
    else:
        arr = np.asarray(m, dtype=np.uint64)[w0:w1 + 1]
    bits = np.unpackbits(np.ascontiguousarray(arr).view(np.uint8), bitorder="little").astype(bool)
    s = off - (w0 << 6)
    return bits[s:s + n]


def valid_grid(r: Float32Raster) -> np.ndarray:
    """The raster's validity as an H×W bool array (all True without a mask)."""
    if r.Valid is None:
        return np.ones((r.Height, r.Width), dtype=bool)
    n = (r.Height - 1) * r.Stride + r.Width
    flat = unpack_bits(r.Valid, r.ValidOffset, n)
    out = np.zeros((r.Height, r.Width), dtype=bool)
    for y in range(r.Height):
        out[y] = flat[y * r.Stride:y * r.Stride + r.Width]
    return out


def data_grid(r: Float32Raster) -> np.ndarray:
    """The raster's cells as an H×W float32 array (a copy)."""
    d = r.numpy_data()
    n = (r.Height - 1) * r.Stride + r.Width
    full = np.zeros(r.Height * r.Stride, dtype=np.float32)
    full[:n] = d[:n]
    return full.reshape(r.Height, r.Stride)[:, :r.Width].copy()
