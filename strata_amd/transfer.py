"""transfer: Lookup (bounded piecewise-linear) and Reclass (step function)
(transfer/transfer.go, transfer/tiled.go, internal/curve/scalar.go), the
pointwise stages of the Slope → Lookup → Reclass serving chain.

NaN passes through; Lookup returns a knot's y exactly on the knot and
y0 + t·(y1 − y0) between knots, clamped to the end values outside them;
Reclass returns values[#breaks ≤ v]. Validity is the input's.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np

from . import _lib
from .engine import EnginePanic, KernelDesc, Options, ProcessChunked, ProcessN, data_relation, PARTIAL
from .raster import Float32Raster


class TransferPanic(EnginePanic):
    pass


def _check_increasing(op: str, name: str, xs: np.ndarray) -> None:
    for i, x in enumerate(xs):
        if x != x:
            raise TransferPanic(f"{op}: {name} must not hold NaN, but {name}[{i}] is NaN")
    for i in range(1, len(xs)):
        if not (xs[i - 1] < xs[i]):
            raise TransferPanic(f"{op}: {name} must be strictly increasing, but {name}[{i}] = {xs[i]} is not "
                                f"above {name}[{i - 1}] = {xs[i - 1]}")


def newReclassKernel(op: str, breaks: Sequence[float], values: Sequence[float]) -> KernelDesc:
    b = np.asarray(breaks, dtype=np.float32)
    v = np.asarray(values, dtype=np.float32)
    if len(v) != len(b) + 1:
        raise TransferPanic(f"{op}: values must hold one more element than breaks, got {len(v)} values and "
                            f"{len(b)} breaks")
    _check_increasing(op, "breaks", b)
    if 2 * len(b) + 1 > _lib.MAX_TAPS:
        raise TransferPanic(f"{op}: {len(b)} breaks exceed the device table limit of {(_lib.MAX_TAPS - 1) // 2}")
    return KernelDesc(_lib.OP_RECLASS, 0, 0, (), np.concatenate([b, v]), 1, 1, "transfer.Reclass")


def newLookupKernel(op: str, xs: Sequence[float], ys: Sequence[float]) -> KernelDesc:
    x = np.asarray(xs, dtype=np.float32)
    y = np.asarray(ys, dtype=np.float32)
    if len(x) != len(y):
        raise TransferPanic(f"{op}: xs and ys must have equal length, got {len(x)} and {len(y)}")
    if len(x) == 0:
        raise TransferPanic(op + ": the table needs at least one knot")
    _check_increasing(op, "xs", x)
    for i, v in enumerate(x):
        if v - v != 0:
            raise TransferPanic(f"{op}: xs must be finite, but xs[{i}] = {v}")
    if 2 * len(x) > _lib.MAX_TAPS:
        raise TransferPanic(f"{op}: {len(x)} knots exceed the device table limit of {_lib.MAX_TAPS // 2}")
    return KernelDesc(_lib.OP_LOOKUP, 0, 0, (), np.concatenate([x, y]), 1, 1, "transfer.Lookup")


def _check(op: str, dst: Float32Raster, src: Float32Raster) -> None:
    """pointwise.Check / CheckMasks (internal/pointwise/pointwise.go:40-67)."""
    for name, r in (("dst", dst), ("src", src)):
        err = r.Validate()
        if err is not None:
            raise TransferPanic(f"{op}: {name}: {err}")
        if r.Width != dst.Width or r.Height != dst.Height:
            raise TransferPanic(f"{op}: {name} dimensions differ from dst")
    if data_relation(dst, src) == PARTIAL:
        raise TransferPanic(f"{op}: dst overlaps src at a different offset or stride")
    if dst.Valid is None and src.Valid is not None:
        raise TransferPanic(f"{op}: an input has a validity mask but dst.Valid is nil; allocate dst with "
                            "raster.NewFloat32Like, or set Valid on its root raster before windowing")


def Reclass(dst: Float32Raster, src: Float32Raster, breaks, values) -> None:
    op = "transfer.Reclass"
    k = newReclassKernel(op, breaks, values)
    _check(op, dst, src)
    ProcessN(None, [dst], [src], k, Options(Workers=1))


def Lookup(dst: Float32Raster, src: Float32Raster, xs, ys) -> None:
    op = "transfer.Lookup"
    k = newLookupKernel(op, xs, ys)
    _check(op, dst, src)
    ProcessN(None, [dst], [src], k, Options(Workers=1))


def ReclassTiled(ctx, dst, src, breaks, values, opts: Options):
    return ProcessN(ctx, [dst], [src], newReclassKernel("transfer.ReclassTiled", breaks, values), opts)


def LookupTiled(ctx, dst, src, xs, ys, opts: Options):
    return ProcessN(ctx, [dst], [src], newLookupKernel("transfer.LookupTiled", xs, ys), opts)


def ReclassChunked(ctx, dst, src, breaks, values, opts: Options):
    return ProcessChunked(ctx, [dst], [src], newReclassKernel("transfer.ReclassChunked", breaks, values), opts)


def LookupChunked(ctx, dst, src, xs, ys, opts: Options):
    return ProcessChunked(ctx, [dst], [src], newLookupKernel("transfer.LookupChunked", xs, ys), opts)
