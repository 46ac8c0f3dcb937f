"""Python binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the checker or the timed CPU
baseline. The product package never imports it.

The oracle is a scalar C restatement of the reference's canonical kernels
(see strata_oracle.c for file:line citations) and of the engine's per-cell
edge and validity rules. It works on host rasters (numpy arrays).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class OrcRaster(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("width", c_int64), ("height", c_int64), ("stride", c_int64),
                ("valid", c_void_p), ("valid_offset", c_int64)]


class OrcKernel(ctypes.Structure):
    _fields_ = [("op", c_int32), ("radius", c_int32), ("flags", c_int32), ("ntaps", c_int32),
                ("k", c_float * 16), ("taps", c_void_p)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_atan32.argtypes = [c_float]
        L.orc_atan32.restype = c_float
        L.orc_atan2f32.argtypes = [c_float, c_float]
        L.orc_atan2f32.restype = c_float
        L.orc_process.argtypes = [POINTER(OrcKernel), POINTER(OrcRaster), c_int, POINTER(OrcRaster), c_int]
        L.orc_process.restype = c_int
        L.orc_process_rows.argtypes = [POINTER(OrcKernel), POINTER(OrcRaster), c_int, POINTER(OrcRaster), c_int,
                                       c_int64, c_int64, c_int]
        L.orc_process_rows.restype = c_int
        L.orc_correlate_f64.argtypes = [POINTER(c_double), c_int64, POINTER(OrcRaster), POINTER(c_float), c_int]
        L.orc_correlate_f64.restype = c_int
        L.orc_tiled_avx2.argtypes = [POINTER(OrcKernel), POINTER(OrcRaster), c_int, POINTER(OrcRaster), c_int,
                                     c_int64, c_int64, c_int]
        L.orc_tiled_avx2.restype = c_int
        _lib = L
    return _lib


def _host(a):
    if a is None:
        return None
    if hasattr(a, "detach"):
        raise TypeError("the oracle takes host (numpy) rasters")
    return a


def _raster(r) -> OrcRaster:
    o = OrcRaster()
    d = _host(r.Data)
    assert d.dtype == np.float32
    o.data = d.ctypes.data
    o.width, o.height, o.stride = r.Width, r.Height, r.Stride
    v = _host(r.Valid)
    o.valid = v.ctypes.data if v is not None else None
    o.valid_offset = r.ValidOffset
    return o


def _kernel(k):
    kc = OrcKernel()
    kc.op, kc.radius, kc.flags = k.op, k.radius, k.flags
    for i, v in enumerate(k.k):
        kc.k[i] = float(np.float32(v))
    taps = None
    if k.taps is not None and len(k.taps):
        taps = np.ascontiguousarray(k.taps, dtype=np.float32)
        kc.ntaps = len(taps)
        kc.taps = taps.ctypes.data
    return kc, taps


def process(k, dst, src, y0: int = 0, y1: int = None, threads: int = 1) -> None:
    """The reference's ProcessN result, cell by cell, into host rasters dst."""
    kc, taps = _kernel(k)
    cd = (OrcRaster * len(dst))(*[_raster(d) for d in dst])
    cs = (OrcRaster * len(src))(*[_raster(s) for s in src])
    if y1 is None:
        y1 = src[0].Height
    rc = lib().orc_process_rows(ctypes.byref(kc), cd, len(dst), cs, len(src), y0, y1, threads)
    del taps
    if rc != 0:
        raise ValueError(f"oracle: bad kernel or operands for {k}")


def tiled_avx2(k, dst, src, tile_w: int = 0, tile_h: int = 0, threads: int = 1) -> None:
    """The CPU baseline: the reference's AVX2 Tiled path (radius-1
    Gradient, Slope, Hillshade) over its band plan with `threads` workers,
    into host rasters dst. Same bits as process()."""
    kc, taps = _kernel(k)
    cd = (OrcRaster * len(dst))(*[_raster(d) for d in dst])
    cs = (OrcRaster * len(src))(*[_raster(s) for s in src])
    rc = lib().orc_tiled_avx2(ctypes.byref(kc), cd, len(dst), cs, len(src), tile_w, tile_h, threads)
    del taps
    if rc == -2:
        raise ValueError(f"oracle: the AVX2 baseline does not cover {k}")
    if rc == -3:
        raise RuntimeError("oracle: this CPU has no AVX2")
    if rc != 0:
        raise ValueError(f"oracle: bad kernel or operands for {k}")


def atan32(x: float) -> float:
    return lib().orc_atan32(x)


def atan2f32(y: float, x: float) -> float:
    return lib().orc_atan2f32(y, x)


def correlate_f64(src, weights, r: int) -> np.ndarray:
    """FP64 correlation over the interior (NaN ring), H×W float64."""
    out = np.zeros((src.Height, src.Width), dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    lib().orc_correlate_f64(out.ctypes.data_as(POINTER(c_double)), src.Width, ctypes.byref(_raster(src)),
                            w.ctypes.data_as(POINTER(c_float)), r)
    return out
