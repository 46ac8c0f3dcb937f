"""The reduction oracle pinned, and the host's finishing of a fold checked
against it, on the CPU.

Pins (none retypes the oracle's formulas):
  X1 (the exact sum)      math.fsum, a correctly rounded float sum, and an
                          integer restatement through math.frexp
  X2 (sum of squares)     the same integer restatement
  Summary                 statistics.pvariance over exact fractions, closed
                          forms of small cases, and the known-answer list of
                          reduce/stats_test.go:155-205 (specials, signed
                          zeros, cancellation, beyond float32, subnormals)
Then strata_amd.reduce's finishing (internal/accum/result.go: Sum, Mean,
StdDev through 256-bit floats) runs on folds the oracle builds, with no
device involved.
"""
import math
import statistics
from fractions import Fraction

import numpy as np
import pytest

from oracle import reduce_ref as R
from strata_amd import _lib
from strata_amd import reduce as reduce_mod

F32 = np.float32
MAXF = float(np.finfo(np.float32).max)
TINY = float(np.finfo(np.float32).smallest_subnormal)
NEG0 = F32(-0.0)


def int_sums(vals):
    """X1, X2 by integer significands: v = m·2^e exactly (math.frexp of the
    float32 value as a float64), so v·2^150 and v²·2^300 are integers."""
    x1 = x2 = 0
    for v in vals:
        v = float(v)
        if not math.isfinite(v) or v == 0:
            continue
        m, e = math.frexp(v)
        mi = int(m * (1 << 24))  # the float32 significand, exactly
        e -= 24
        x1 += shift_exact(mi, e + 150)
        x2 += shift_exact(mi * mi, 2 * e + 300)
    return x1, x2


def shift_exact(m, s):
    if s >= 0:
        return m << s
    assert m % (1 << -s) == 0
    return m >> -s


def rnd(rng, n, spread):
    return [F32(v) for v in rng.standard_normal(n) * np.exp2(rng.integers(-spread, spread, n))]


@pytest.mark.parametrize("seed", range(6))
def test_exact_sums_pinned(seed):
    rng = np.random.default_rng(seed)
    vals = rnd(rng, 400, 40 if seed % 2 else 120)
    if seed == 3:
        vals += [F32(TINY), F32(-3 * TINY), F32(MAXF), F32(-MAXF), NEG0, F32(0.0)]
    f = R.fold(vals)
    x1, x2 = int_sums(vals)
    assert f["x1"] == x1 and f["x2"] == x2
    assert f["x1"] / (1 << 150) == math.fsum(float(v) for v in vals)
    assert f["count"] == len(vals)


def test_min_max_semantics():
    f = R.fold([F32(0.0), NEG0])
    assert R.f32_bits(f["min"]) == 0x80000000 and R.f32_bits(f["max"]) == 0
    f = R.fold([F32(1), F32(np.nan), F32(-5)])
    assert np.isnan(f["min"]) and np.isnan(f["max"]) and f["nan"] == 1
    f = R.fold([F32(np.inf), F32(2), F32(-np.inf)])
    assert f["min"] == -np.inf and f["max"] == np.inf and (f["pos_inf"], f["neg_inf"]) == (1, 1)


@pytest.mark.parametrize("seed", range(4))
def test_variance_pinned(seed):
    rng = np.random.default_rng(100 + seed)
    vals = [F32(v) for v in 300 + 2000 * rng.random(257)]
    s = R.summary(vals)
    var = statistics.pvariance([Fraction(float(v)) for v in vals])
    want = math.sqrt(float(var))  # within an ulp of the exact root
    assert abs(s["StdDev"] - want) <= math.ulp(want)
    assert s["Sum"] == math.fsum(float(v) for v in vals)
    x1, _ = int_sums(vals)
    assert s["Mean"] == x1 / (len(vals) << 150)  # int true division rounds once


def known_cases():
    """reduce/stats_test.go:155-205 with the values each case fixes."""
    nan, inf = F32(np.nan), F32(np.inf)
    return {
        "one cell": ([F32(7)], dict(Count=1, Sum=7.0, Mean=7.0, StdDev=0.0, Min=7, Max=7)),
        "ordinary": ([F32(v) for v in (3, 1, 4, 1, 5, 9, 2, 6)],
                     dict(Count=8, Sum=31.0, Mean=3.875, StdDev=math.sqrt(6.609375), Min=1, Max=9)),
        "negatives": ([F32(v) for v in (-3, -1, -4)],
                      dict(Count=3, Sum=-8.0, Mean=-8 / 3, StdDev=math.sqrt(14 / 9), Min=-4, Max=-1)),
        "infinities": ([inf, F32(1), -inf], dict(Count=3, Sum="nan", Mean="nan", StdDev="nan", Min=-np.inf,
                                                 Max=np.inf)),
        "one infinity": ([inf, F32(1), F32(2)], dict(Count=3, Sum=math.inf, Mean=math.inf, StdDev="nan", Min=1,
                                                     Max=np.inf)),
        "nan": ([F32(1), nan, F32(3)], dict(Count=3, Sum="nan", Mean="nan", StdDev="nan", Min="nan", Max="nan")),
        "signed zeros": ([F32(0), NEG0], dict(Count=2, Sum=0.0, Mean=0.0, StdDev=0.0, Min=NEG0, Max=F32(0))),
        "all -0": ([NEG0], dict(Count=1, Sum=-0.0, Mean=-0.0, StdDev=0.0, Min=NEG0, Max=NEG0)),
        "all invalid": ([], dict(Count=0, Sum=0.0, Mean="nan", StdDev="nan", Min="nan", Max="nan")),
        "cancellation": ([F32(MAXF), F32(1e-30), F32(-MAXF)],
                         dict(Count=3, Sum=float(F32(1e-30)), Mean=float(F32(1e-30)) / 3, Min=-MAXF, Max=MAXF)),
        "beyond float32": ([F32(MAXF), F32(MAXF)], dict(Count=2, Sum=2 * MAXF, Mean=MAXF, StdDev=0.0, Min=MAXF,
                                                        Max=MAXF)),
        "wide block": ([F32(1), F32(2.0 ** -33), F32(-1), F32(3)],
                       dict(Count=4, Sum=3 + 2.0 ** -33, Mean=(3 + 2.0 ** -33) / 4, Min=-1, Max=3)),
        "subnormals": ([F32(TINY), F32(-2 * TINY), F32(2.0 ** -130)],
                       dict(Count=3, Sum=2.0 ** -130 - TINY, Min=-2 * TINY, Max=2.0 ** -130)),
    }


def same64(a, b):
    if isinstance(b, str):
        return a != a
    return a == b and math.copysign(1, a) == math.copysign(1, b)


def same32(a, b):
    if isinstance(b, str):
        return a != a
    return R.f32_bits(a) == R.f32_bits(F32(b))


@pytest.mark.parametrize("name", list(known_cases()))
def test_oracle_known_answers(name):
    vals, want = known_cases()[name]
    s = R.summary(vals)
    for k, v in want.items():
        if k in ("Min", "Max"):
            assert same32(s[k], v), (k, s[k], v)
        elif k == "StdDev" and not isinstance(v, str):
            assert abs(s[k] - v) <= math.ulp(v) if v else s[k] == v, (k, s[k], v)
        elif k == "Count":
            assert s[k] == v
        else:
            assert same64(s[k], v), (k, s[k], v)


def fold_struct(vals):
    """A _lib.Fold holding the oracle's fold of vals, as the device returns it."""
    f = R.fold(vals)
    out = _lib.Fold()
    out.count, out.nan, out.pos_inf, out.neg_inf = f["count"], f["nan"], f["pos_inf"], f["neg_inf"]
    out.not_neg_zero = 1 if f["not_neg_zero"] else 0
    out.min, out.max = float(f["min"]), float(f["max"])
    for name, x, n in (("sum", f["x1"], 6), ("sumsq", f["x2"], 10)):
        u = x % (1 << (64 * n))
        arr = getattr(out, name)
        for i in range(n):
            arr[i] = (u >> (64 * i)) & ((1 << 64) - 1)
        assert _lib.limbs_to_int(arr) == x
    return out


@pytest.mark.parametrize("name", list(known_cases()))
def test_finishing_known_answers(name):
    vals, _ = known_cases()[name]
    want = R.summary(vals)
    got = reduce_mod._summary(fold_struct(vals))
    assert got.Count == want["Count"]
    assert same64(got.Sum, want["Sum"] if want["Sum"] == want["Sum"] else "nan")
    assert same64(got.Mean, want["Mean"] if want["Mean"] == want["Mean"] else "nan")
    sd = want["StdDev"]
    if sd != sd:
        assert got.StdDev != got.StdDev
    else:
        assert abs(got.StdDev - sd) <= math.ulp(sd) if sd else got.StdDev == sd
    for k in ("Min", "Max"):
        w = want[k]
        assert same32(getattr(got, k), "nan" if w != w else w)
    # Every NaN is the canonical one (reduce/stats_test.go:249-263).
    for v in (got.Sum, got.Mean, got.StdDev):
        if v != v:
            assert np.float64(v).view(np.uint64) == 0x7FF8000000000001
    for v in (got.Min, got.Max):
        if v != v:
            assert R.f32_bits(v) == 0x7FC00000


@pytest.mark.parametrize("seed", range(8))
def test_finishing_random(seed):
    """Sum and Mean bit for bit, StdDev within an ulp of the exact root, over
    random spreads (the 256-bit route of internal/accum/result.go)."""
    rng = np.random.default_rng(seed)
    vals = rnd(rng, int(rng.integers(1, 300)), int(rng.integers(1, 60)))
    want = R.summary(vals)
    got = reduce_mod._summary(fold_struct(vals))
    assert got.Sum == want["Sum"] and got.Mean == want["Mean"]
    assert abs(got.StdDev - want["StdDev"]) <= math.ulp(want["StdDev"])


def test_round_helpers_exact_cases():
    # sqrt of an exact square comes back exact; ties round to even.
    assert reduce_mod._round_sqrt(49, 0, 256) == (7 << 253, -253)
    q, e = reduce_mod._round_int((1 << 300) + (1 << 44), 256)  # exactly half an ulp above: tie to even
    assert (q, e) == (1 << 255, 45)
    q, e = reduce_mod._round_int((1 << 300) + (3 << 44), 256)  # one and a half: rounds up to even
    assert (q, e) == ((1 << 255) + 2, 45)
    q, e = reduce_mod._round_quotient(1, 3, 256)
    assert Fraction(q) * Fraction(2) ** e - Fraction(1, 3) < Fraction(2) ** (e - 1)
