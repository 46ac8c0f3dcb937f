"""Atan32 and Atan2F32 of the oracle against the reference's accuracy tests
(internal/stencil/stencil_test.go:9-43, 348-417) and the constants against
their exact values. CPU only."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle


def nearest_f32(x: Fraction) -> int:
    """Bits of the binary32 value nearest to the exact rational x > 0."""
    f = np.float32(float(x))
    best = None
    for cand in (np.nextafter(f, np.float32(0)), f, np.nextafter(f, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - x)
        if best is None or err < best[0]:
            best = (err, int(np.float32(cand).view(np.uint32)))
    return best[1]


def pi_fraction():
    # 60 digits of pi: far beyond what binary32 rounding can see.
    return Fraction("3.14159265358979323846264338327950288419716939937510582097494")


def test_constants_are_nearest_binary32():
    pi = pi_fraction()
    assert nearest_f32(pi / 2) == 0x3FC90FDB
    assert nearest_f32(pi / 4) == 0x3F490FDB
    assert nearest_f32(pi) == 0x40490FDB
    assert nearest_f32(Fraction(180) / pi) == 0x42652EE1
    # The device kernels use the same bit patterns.
    src = open(__file__.replace("tests/test_oracle_atan.py", "strata_amd/csrc/arith.hpp")).read()
    for bits in ("0x3fc90fdbu", "0x3f490fdbu", "0x40490fdbu", "0x42652ee1u"):
        assert bits in src


def test_atan32_accuracy_stride():
    """A stride through every non-negative finite float32 plus the
    neighbourhoods of the reduction thresholds: error <= 1.5e-7 rad."""
    bits = np.arange(0, 0x7F800000, 9973, dtype=np.uint64).astype(np.uint32)
    xs = bits.view(np.float32)
    worst = 0.0
    for x in xs[::7]:  # every 7th of the reference's stride keeps the CPU run short
        e = abs(oracle.atan32(float(x)) - math.atan(float(x)))
        worst = max(worst, e)
    assert worst <= 1.5e-7
    for th in (np.float32(0.4142135623730950), np.float32(2.414213562373095), np.float32(1)):
        x = th
        for _ in range(1000):
            x = np.nextafter(x, np.float32(0))
        for _ in range(2000):
            assert abs(oracle.atan32(float(x)) - math.atan(float(x))) <= 1.5e-7
            x = np.nextafter(x, np.float32(10))


def test_atan32_specials():
    assert np.float32(oracle.atan32(0.0)).view(np.uint32) == 0
    assert np.float32(oracle.atan32(math.inf)).view(np.uint32) == 0x3FC90FDB
    assert math.isnan(oracle.atan32(math.nan))


def test_atan2f32_special_table():
    negz = -0.0
    pi = np.float32(math.pi)
    cases = [
        (0, 0, 0.0), (negz, 0, negz), (0, negz, pi), (negz, negz, -pi), (0, -5, pi), (negz, -5, -pi),
        (5, 0, pi / np.float32(2)), (-5, negz, -pi / np.float32(2)), (math.inf, math.inf, pi / np.float32(4)),
        (-math.inf, -math.inf, -(pi - pi / np.float32(4))), (3, -math.inf, pi), (-math.inf, 3, -pi / np.float32(2)),
    ]
    for y, x, want in cases:
        got = np.float32(oracle.atan2f32(y, x))
        assert got.view(np.uint32) == np.float32(want).view(np.uint32), (y, x, got, want)


def atan2_sample(rng):
    v = np.float32(math.pow(2, rng.random() * 200 - 100))
    return -v if rng.integers(2) == 0 else v


def test_atan2f32_accuracy():
    rng = np.random.default_rng(21)
    specials = [0.0, -0.0, math.inf, -math.inf, math.nan, 1.0, -1.0, 3.4028234663852886e38,
                -3.4028234663852886e38, 1e-45, -1e-45, 1.17549435e-38]
    for y in specials:
        for x in specials:
            got = oracle.atan2f32(y, x)
            want = math.atan2(np.float32(y), np.float32(x)) if not (math.isnan(y) or math.isnan(x)) else math.nan
            if math.isnan(want):
                assert math.isnan(got)
                continue
            assert math.copysign(1, got) == math.copysign(1, want)
            assert abs(got - want) <= 3e-7
    for _ in range(20000):
        y, x = atan2_sample(rng), atan2_sample(rng)
        for xx in (x, np.nextafter(y, np.float32(0)) * np.float32(1 - 2 * int(rng.integers(2)))):
            got = oracle.atan2f32(float(y), float(xx))
            want = math.atan2(float(y), float(xx))
            assert math.copysign(1, got) == math.copysign(1, want)
            assert abs(got - want) <= 3e-7
