"""CPU oracle of the reductions (reduce.Count / MinMax / Sum / Stats).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker. The product
package never imports it.

Restates, cell by cell in exact rational arithmetic:
  the fold of the valid cells      reduce/reduce.go:88-228 (countOp,
                                   minMaxOp: Go's builtin min and max,
                                   NaN absorbing, -0 < +0)
  the exact sums                   internal/accum/accum.go:192-245,
                                   moments.go:95-112 (the bins' integers:
                                   sum·2^150 and sum of squares·2^300)
  the Summary                      reduce/stats_test.go:20-84 (refStats,
                                   the reference's own checker: exact
                                   rationals, specials by hand, the
                                   root of the exact variance)
Pinned in tests/test_oracle_reduce.py against math.fsum (a correctly
rounded float sum), statistics.pvariance on exact fractions, closed
forms of small cases, and the known-answer list of
reduce/stats_test.go:155-205.
"""
from __future__ import annotations

import math
import struct
from decimal import Decimal, getcontext
from fractions import Fraction
from typing import Iterable, List, Optional

import numpy as np


def f32_bits(v) -> int:
    return int(np.float32(v).view(np.uint32))


def go_min(a: np.float32, b: np.float32) -> np.float32:
    """Go's builtin min on float32 (NaN wins, -0 < +0)."""
    if a != a:
        return a
    if b != b:
        return b
    if a == b:
        return a if f32_bits(a) >> 31 else b
    return a if a < b else b


def go_max(a: np.float32, b: np.float32) -> np.float32:
    if a != a:
        return a
    if b != b:
        return b
    if a == b:
        return b if f32_bits(a) >> 31 else a
    return a if a > b else b


def valid_values(data: np.ndarray, width: int, height: int, stride: int,
                 valid: Optional[np.ndarray], valid_offset: int) -> List[np.float32]:
    """The valid cells' values in row order."""
    out = []
    for y in range(height):
        for x in range(width):
            i = y * stride + x
            if valid is not None:
                b = valid_offset + i
                if not (int(valid[b >> 6]) >> (b & 63)) & 1:
                    continue
            out.append(np.float32(data[i]))
    return out


def fold(vals: Iterable[np.float32]) -> dict:
    """What the device's fold returns for these valid values: counts,
    tallies, Go min/max (+Inf/-Inf when empty), and the exact integers
    X1 = Σ v·2^150 and X2 = Σ v²·2^300 over the finite values."""
    count = nan = pinf = ninf = 0
    nnz = False
    mn, mx = np.float32(np.inf), np.float32(-np.inf)
    s1 = Fraction(0)
    s2 = Fraction(0)
    for v in vals:
        v = np.float32(v)
        count += 1
        nnz = nnz or f32_bits(v) != 0x80000000
        mn, mx = go_min(mn, v), go_max(mx, v)
        if v != v:
            nan += 1
        elif v == np.inf:
            pinf += 1
        elif v == -np.inf:
            ninf += 1
        else:
            q = Fraction(float(v))
            s1 += q
            s2 += q * q
    x1 = s1 * (1 << 150)
    x2 = s2 * (1 << 300)
    assert x1.denominator == 1 and x2.denominator == 1
    return dict(count=count, nan=nan, pos_inf=pinf, neg_inf=ninf, not_neg_zero=nnz, min=mn, max=mx,
                x1=int(x1), x2=int(x2))


GO_NAN64 = struct.unpack("<d", struct.pack("<Q", 0x7FF8000000000001))[0]


def summary(vals: List[np.float32]) -> dict:
    """refStats (reduce/stats_test.go:20-84): Count, Sum, Mean, StdDev,
    Min, Max. StdDev is the root of the exact variance to 120 significant
    digits, rounded to float64; the reference's test compares it within
    one ulp."""
    n = len(vals)
    f = fold(vals)
    mn = f["min"] if n else np.float32(np.nan)
    mx = f["max"] if n else np.float32(np.nan)
    s = dict(Count=n, Min=mn, Max=mx, Sum=0.0, Mean=GO_NAN64, StdDev=GO_NAN64)
    if n == 0:
        return s
    nans, pos, neg = f["nan"], f["pos_inf"], f["neg_inf"]
    if nans > 0 or (pos > 0 and neg > 0):
        s["Sum"] = s["Mean"] = GO_NAN64
        return s
    if pos > 0:
        s["Sum"] = s["Mean"] = math.inf
        return s
    if neg > 0:
        s["Sum"] = s["Mean"] = -math.inf
        return s
    q = [Fraction(float(v)) for v in vals]
    total = sum(q, Fraction(0))
    mean = total / n
    var = sum(((x - mean) ** 2 for x in q), Fraction(0)) / n
    s["Sum"] = float(total)
    s["Mean"] = float(mean)
    if total == 0 and all(f32_bits(v) == 0x80000000 for v in vals):
        s["Sum"] = s["Mean"] = -0.0
    getcontext().prec = 120
    s["StdDev"] = float((Decimal(var.numerator) / Decimal(var.denominator)).sqrt())
    return s
