"""reduce: Count, MinMax, Sum and Stats of a raster's valid cells, in the
reference's three forms — plain, Tiled and Chunked — over the HIP library.

The device folds the raster (engine.Reduce / ReduceChunked): counts, the
NaN and infinity tallies, Go's min and max, and the exact sums as integers
(the bins of internal/accum). What is left is the finishing the
reference does once per call on the host (internal/accum/result.go),
restated here with Python integers:

  Sum     the exact sum, correctly rounded to float64 (big.Rat.Float64)
  Mean    the exact sum over the count, correctly rounded
  StdDev  the square root of the exact variance through 256-bit binary
          floats as math/big does: the numerator rounded to 256 bits, the
          quotient rounded to 256 bits, its square root rounded to 256
          bits, then to float64. The root is taken correctly rounded;
          math/big's Newton root agrees except within about 2^-288 of a
          256-bit rounding midpoint.

Values (reduce/doc.go): only valid cells take part; a NaN in a valid cell
is a value; -0 < +0; every NaN result is the canonical one. A cancelled or
failed reduction returns no value (zero results and the error).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass
from math import isqrt
from typing import Optional, Tuple

import numpy as np

from . import _lib
from .engine import Options, RasterSource, Reduce, ReduceChunked
from .raster import Float32Raster

# math.NaN() and float32(math.NaN()) in Go: the canonical NaNs results use.
NaN64 = struct.unpack("<d", struct.pack("<Q", 0x7FF8000000000001))[0]
NaN32 = np.uint32(0x7FC00000).view(np.float32)

SUM_BIAS = 150       # bin k of the sums weighs 2^(k-150)
SQ_BIAS = 2 * SUM_BIAS
PREC = 256           # math/big precision of StdDev (internal/accum/result.go)

_plain = Options(Workers=1)  # the plain functions: one worker, one tile (reduce/reduce.go:37-43)


def _canonical32(v: float) -> np.float32:
    v = np.float32(v)
    return NaN32 if v != v else v


# -- finishing (internal/accum/result.go) -------------------------------------------

def _special(f) -> Optional[float]:
    if f.nan > 0 or (f.pos_inf > 0 and f.neg_inf > 0):
        return NaN64
    if f.pos_inf > 0:
        return math.inf
    if f.neg_inf > 0:
        return -math.inf
    return None


def _zero(f) -> float:
    # A sum of zeros is -0 only when every value was -0.
    return -0.0 if f.count > 0 and not f.not_neg_zero else 0.0


def _sum_value(f) -> float:
    sp = _special(f)
    if sp is not None:
        return sp
    x = _lib.limbs_to_int(f.sum)
    if x == 0:
        return _zero(f)
    return x / (1 << SUM_BIAS)  # int true division: correctly rounded


def _mean_value(f) -> float:
    if f.count == 0:
        return NaN64
    sp = _special(f)
    if sp is not None:
        return sp
    x = _lib.limbs_to_int(f.sum)
    if x == 0:
        return _zero(f)
    return x / (f.count << SUM_BIAS)


def _round_int(m: int, prec: int) -> Tuple[int, int]:
    """m > 0 rounded to prec significant bits, nearest-even: (q, e) with
    the value q·2^e (big.Float.SetInt at precision prec)."""
    n = m.bit_length()
    if n <= prec:
        return m, 0
    sh = n - prec
    q, r = m >> sh, m & ((1 << sh) - 1)
    half = 1 << (sh - 1)
    if r > half or (r == half and q & 1):
        q += 1
        if q.bit_length() > prec:
            q >>= 1
            sh += 1
    return q, sh


def _round_quotient(a: int, b: int, prec: int) -> Tuple[int, int]:
    """a/b > 0 rounded to prec bits, nearest-even: (q, e), value q·2^e."""
    s = max(0, prec + 2 + b.bit_length() - a.bit_length())
    q, r = divmod(a << s, b)
    sh = q.bit_length() - prec
    keep, rem = q >> sh, q & ((1 << sh) - 1)
    half = 1 << (sh - 1)
    if rem > half or (rem == half and (r != 0 or keep & 1)):
        keep += 1
        if keep.bit_length() > prec:
            keep >>= 1
            sh += 1
    return keep, sh - s


def _round_sqrt(m: int, e: int, prec: int) -> Tuple[int, int]:
    """sqrt(m·2^e) > 0 rounded to prec bits, nearest-even: (q, e')."""
    if e & 1:
        m <<= 1
        e -= 1
    t = max(0, (2 * (prec + 2) - m.bit_length() + 1) // 2)
    big = m << (2 * t)
    s = isqrt(big)
    exact = s * s == big
    sh = s.bit_length() - prec
    keep, rem = s >> sh, s & ((1 << sh) - 1)
    half = 1 << (sh - 1)
    if rem > half or (rem == half and (not exact or keep & 1)):
        keep += 1
        if keep.bit_length() > prec:
            keep >>= 1
            sh += 1
    return keep, sh + e // 2 - t


def _stddev_value(f) -> float:
    if f.count == 0 or f.nan > 0 or f.pos_inf > 0 or f.neg_inf > 0:
        return NaN64
    n = f.count
    x1 = _lib.limbs_to_int(f.sum)
    x2 = _lib.limbs_to_int(f.sumsq)
    num = n * x2 - x1 * x1          # variance = num / den exactly
    den = (n * n) << SQ_BIAS
    if num == 0:
        return 0.0
    nm, ne = _round_int(num, PREC)
    dm, de = _round_int(den, PREC)
    qm, qe = _round_quotient(nm, dm, PREC)
    qe += ne - de
    sm, se = _round_sqrt(qm, qe, PREC)
    return math.ldexp(float(sm), se)  # int -> float rounds once, to nearest-even


# -- the reductions ---------------------------------------------------------------

def _fold(ctx, src, what, opts, chunked):
    if chunked:
        return ReduceChunked(ctx, [src], what, opts)
    return Reduce(ctx, [src], what, opts)


def CountTiled(ctx, src: Float32Raster, opts: Options):
    """reduce.CountTiled (reduce/reduce.go:56-59): (count, err)."""
    f, err = _fold(ctx, src, _lib.REDUCE_COUNT, opts, False)
    return (0, err) if err is not None else (int(f.count), None)


def CountChunked(ctx, src: RasterSource, opts: Options):
    """reduce.CountChunked (reduce/reduce.go:61-65)."""
    f, err = _fold(ctx, src, _lib.REDUCE_COUNT, opts, True)
    return (0, err) if err is not None else (int(f.count), None)


def Count(src: Float32Raster) -> int:
    """reduce.Count (reduce/reduce.go:14-20): the number of valid cells."""
    n, _ = CountTiled(None, src, _plain)
    return n


def _minmax(f):
    if f.count == 0:
        return NaN32, NaN32, 0
    return _canonical32(f.min), _canonical32(f.max), int(f.count)


def MinMaxTiled(ctx, src: Float32Raster, opts: Options):
    """reduce.MinMaxTiled (reduce/reduce.go:67-75): (mn, mx, count, err)."""
    f, err = _fold(ctx, src, _lib.REDUCE_MINMAX, opts, False)
    if err is not None:
        return np.float32(0), np.float32(0), 0, err
    return (*_minmax(f), None)


def MinMaxChunked(ctx, src: RasterSource, opts: Options):
    """reduce.MinMaxChunked (reduce/reduce.go:77-86)."""
    f, err = _fold(ctx, src, _lib.REDUCE_MINMAX, opts, True)
    if err is not None:
        return np.float32(0), np.float32(0), 0, err
    return (*_minmax(f), None)


def MinMax(src: Float32Raster):
    """reduce.MinMax (reduce/reduce.go:22-33): the smallest and largest
    valid cells with Go's min and max, and how many took part; NaN, NaN
    and 0 with no valid cells."""
    mn, mx, n, _ = MinMaxTiled(None, src, _plain)
    return mn, mx, n


def SumTiled(ctx, src: Float32Raster, opts: Options):
    """reduce.SumTiled (reduce/sum.go:49-57): (sum, count, err)."""
    f, err = _fold(ctx, src, _lib.REDUCE_SUM, opts, False)
    if err is not None:
        return 0.0, 0, err
    return _sum_value(f), int(f.count), None


def SumChunked(ctx, src: RasterSource, opts: Options):
    """reduce.SumChunked (reduce/sum.go:59-66)."""
    f, err = _fold(ctx, src, _lib.REDUCE_SUM, opts, True)
    if err is not None:
        return 0.0, 0, err
    return _sum_value(f), int(f.count), None


def Sum(src: Float32Raster):
    """reduce.Sum (reduce/sum.go:37-47): the exact sum of the valid cells,
    correctly rounded to float64, and their count."""
    s, n, _ = SumTiled(None, src, _plain)
    return s, n


@dataclass
class Summary:
    """reduce.Summary (reduce/sum.go:14-31)."""
    Count: int = 0
    Sum: float = 0.0
    Mean: float = 0.0
    StdDev: float = 0.0
    Min: np.float32 = np.float32(0)
    Max: np.float32 = np.float32(0)


def _summary(f) -> Summary:
    # internal/summary/summary.go:46-55: with no cells Count and Sum are 0
    # and the rest NaN.
    s = Summary(int(f.count), _sum_value(f), _mean_value(f), _stddev_value(f), NaN32, NaN32)
    if f.count > 0:
        s.Min, s.Max = _canonical32(f.min), _canonical32(f.max)
    return s


_STATS = _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS


def StatsTiled(ctx, src: Float32Raster, opts: Options):
    """reduce.StatsTiled (reduce/sum.go:78-85): (Summary, err)."""
    f, err = _fold(ctx, src, _STATS, opts, False)
    return (Summary(), err) if err is not None else (_summary(f), None)


def StatsChunked(ctx, src: RasterSource, opts: Options):
    """reduce.StatsChunked (reduce/sum.go:87-95)."""
    f, err = _fold(ctx, src, _STATS, opts, True)
    return (Summary(), err) if err is not None else (_summary(f), None)


def Stats(src: Float32Raster) -> Summary:
    """reduce.Stats (reduce/sum.go:68-76): Count, Sum, Mean, StdDev, Min
    and Max of the valid cells from one pass."""
    s, _ = StatsTiled(None, src, _plain)
    return s


def Folded(tee) -> Summary:
    """The summary an engine.TeeSink folded (graph/sinks.go:68-71,
    folder.summary): reduce.Stats of every window the tee was handed,
    folded in the pass that computed them. A tee whose reduction is not
    the full Stats fold gives only the fields its reduction holds."""
    return _summary(tee.Fold())
