"""focal: Correlate, Convolve, CorrelateSeparable, Mean, Min and Max
(focal/*.go) in plain, Tiled and Chunked forms over the HIP library.

Radius 1..MaxRadius. Weights are checked finite; Convolve rotates them by
180° once (focal/correlate.go:70-88); Gaussian taps are exp(-x²/2σ²) in
float64, normalised, rounded to float32 (focal/separable.go:53-76). Edges
are NaN and invalid; an output cell is valid iff its (2r+1)² window is.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .engine import EnginePanic, KernelDesc, Options, ProcessChunked, ProcessN
from .raster import Float32Raster

MaxRadius = 8


class FocalPanic(EnginePanic):
    pass


def checkRadius(r: int) -> None:
    if r < 1 or r > MaxRadius:
        raise FocalPanic(f"focal: Radius must be 1 to {MaxRadius}, got {r}")


def checkFinite(what: str, v) -> None:
    for i, x in enumerate(np.asarray(v, dtype=np.float32)):
        if math.isnan(x) or math.isinf(x):
            raise FocalPanic(f"focal: {what} must be finite, got {x} at index {i}")


@dataclass
class WeightsOptions:
    Radius: int = 0
    Weights: Sequence[float] = ()


@dataclass
class SeparableOptions:
    Radius: int = 0
    Row: Sequence[float] = ()
    Col: Sequence[float] = ()


@dataclass
class BoxOptions:
    Radius: int = 0


def newWeightsKernel(opts: WeightsOptions, rotate: bool, mfma: bool = False) -> KernelDesc:
    checkRadius(opts.Radius)
    k = 2 * opts.Radius + 1
    w = np.asarray(opts.Weights, dtype=np.float32)
    if len(w) != k * k:
        raise FocalPanic(f"focal: Radius {opts.Radius} needs {k}×{k} = {k * k} Weights, got {len(w)}")
    checkFinite("Weights", w)
    if rotate:
        w = w[::-1].copy()
    flags = _lib.FLAG_MFMA if mfma else 0
    return KernelDesc(_lib.OP_CORRELATE, opts.Radius, flags, (), w, 1, 1,
                      "focal.Convolve" if rotate else "focal.Correlate")


def Gaussian(radius: int, sigma: float) -> np.ndarray:
    checkRadius(radius)
    if not (sigma > 0) or math.isinf(sigma):
        raise FocalPanic(f"focal: Gaussian sigma must be positive and finite, got {sigma}")
    g = []
    s = 0.0
    for i in range(2 * radius + 1):
        x = float(i - radius)
        v = math.exp(-x * x / (2 * sigma * sigma))
        g.append(v)
        s += v
    return np.array([np.float32(v / s) for v in g], dtype=np.float32)


def newSeparableKernel(opts: SeparableOptions) -> KernelDesc:
    checkRadius(opts.Radius)
    k = 2 * opts.Radius + 1
    row = np.asarray(opts.Row, dtype=np.float32)
    col = np.asarray(opts.Col, dtype=np.float32)
    if len(row) != k or len(col) != k:
        raise FocalPanic(f"focal: Radius {opts.Radius} needs {k} Row and {k} Col taps, got {len(row)} and {len(col)}")
    checkFinite("Row taps", row)
    checkFinite("Col taps", col)
    return KernelDesc(_lib.OP_SEPARABLE, opts.Radius, 0, (), np.concatenate([row, col]), 1, 1,
                      "focal.CorrelateSeparable")


_BOX = {"mean": _lib.OP_MEAN, "min": _lib.OP_MIN, "max": _lib.OP_MAX}


def newBox(opts: BoxOptions, op: str) -> KernelDesc:
    checkRadius(opts.Radius)
    return KernelDesc(_BOX[op], opts.Radius, 0, (), None, 1, 1, "focal." + op.capitalize())


def _run(k, dst, src):
    ProcessN(None, [dst], [src], k, Options(Workers=1))


def _tiled(ctx, eopts, k, dst, src):
    return ProcessN(ctx, [dst], [src], k, eopts)


def _chunked(ctx, eopts, k, dst, src):
    return ProcessChunked(ctx, [dst], [src], k, eopts)


def Correlate(dst, src, opts: WeightsOptions) -> None:
    _run(newWeightsKernel(opts, False), dst, src)


def CorrelateTiled(ctx, dst, src, opts: WeightsOptions, eopts: Options):
    return _tiled(ctx, eopts, newWeightsKernel(opts, False), dst, src)


def CorrelateChunked(ctx, dst, src, opts: WeightsOptions, eopts: Options):
    return _chunked(ctx, eopts, newWeightsKernel(opts, False), dst, src)


def Convolve(dst, src, opts: WeightsOptions) -> None:
    _run(newWeightsKernel(opts, True), dst, src)


def ConvolveTiled(ctx, dst, src, opts: WeightsOptions, eopts: Options):
    return _tiled(ctx, eopts, newWeightsKernel(opts, True), dst, src)


def ConvolveChunked(ctx, dst, src, opts: WeightsOptions, eopts: Options):
    return _chunked(ctx, eopts, newWeightsKernel(opts, True), dst, src)


def CorrelateSeparable(dst, src, opts: SeparableOptions) -> None:
    _run(newSeparableKernel(opts), dst, src)


def CorrelateSeparableTiled(ctx, dst, src, opts: SeparableOptions, eopts: Options):
    return _tiled(ctx, eopts, newSeparableKernel(opts), dst, src)


def CorrelateSeparableChunked(ctx, dst, src, opts: SeparableOptions, eopts: Options):
    return _chunked(ctx, eopts, newSeparableKernel(opts), dst, src)


def Mean(dst, src, opts: BoxOptions) -> None:
    _run(newBox(opts, "mean"), dst, src)


def MeanTiled(ctx, dst, src, opts: BoxOptions, eopts: Options):
    return _tiled(ctx, eopts, newBox(opts, "mean"), dst, src)


def MeanChunked(ctx, dst, src, opts: BoxOptions, eopts: Options):
    return _chunked(ctx, eopts, newBox(opts, "mean"), dst, src)


def Min(dst, src, opts: BoxOptions) -> None:
    _run(newBox(opts, "min"), dst, src)


def MinTiled(ctx, dst, src, opts: BoxOptions, eopts: Options):
    return _tiled(ctx, eopts, newBox(opts, "min"), dst, src)


def MinChunked(ctx, dst, src, opts: BoxOptions, eopts: Options):
    return _chunked(ctx, eopts, newBox(opts, "min"), dst, src)


def Max(dst, src, opts: BoxOptions) -> None:
    _run(newBox(opts, "max"), dst, src)


def MaxTiled(ctx, dst, src, opts: BoxOptions, eopts: Options):
    return _tiled(ctx, eopts, newBox(opts, "max"), dst, src)


def MaxChunked(ctx, dst, src, opts: BoxOptions, eopts: Options):
    return _chunked(ctx, eopts, newBox(opts, "max"), dst, src)


OPKERNELS = {
    "focal.Correlate": lambda o: newWeightsKernel(o, False),
    "focal.Convolve": lambda o: newWeightsKernel(o, True),
    "focal.CorrelateSeparable": newSeparableKernel,
    "focal.Mean": lambda o: newBox(o, "mean"),
    "focal.Min": lambda o: newBox(o, "min"),
    "focal.Max": lambda o: newBox(o, "max"),
}


def CorrelateError(k: KernelDesc, src: Float32Raster, out: Float32Raster) -> dict:
    """Error of a Correlate output against the FP64 correlation of src,
    computed on the device over the interior cells whose window is finite:
    max |out - ref|, max |out - ref| / |ref|, max |out - ref| / sum |w x|,
    and the number of cells compared. For reporting the precision of the
    FP32 path and of the split-BF16x3 MFMA path (newWeightsKernel(...,
    mfma=True))."""
    import ctypes
    if not (src.on_device and out.on_device):
        raise FocalPanic("focal: the error report reads device rasters")
    kc, taps = k.to_c()
    s, o = src.to_c(), out.to_c()
    a, r, sc = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_int64()
    rc = _lib.lib().strata_correlate_error(ctypes.byref(kc), ctypes.byref(s), ctypes.byref(o), ctypes.byref(a),
                                           ctypes.byref(r), ctypes.byref(sc), ctypes.byref(n))
    del taps
    if rc < 0:
        raise FocalPanic(_lib.last_error())
    return {"max_abs": a.value, "max_rel": r.value, "max_scaled": sc.value, "cells": n.value}
