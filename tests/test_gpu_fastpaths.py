"""The kernels' fast paths against the full correctly rounded sequences they
stand in for, on the device: exhaustively for the square root (every
non-negative float the fast path accepts) and the small-argument
arctangent (every float in [0, tan(pi/8)]), and over 2^33 hashed operand
pairs for the quotient."""
import ctypes

import pytest

from strata_amd import _lib

pytestmark = pytest.mark.gpu


def selfcheck(kind, lo, n):
    bad, first, checked = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = _lib.lib().strata_selfcheck(kind, lo, n, ctypes.byref(bad), ctypes.byref(first), ctypes.byref(checked))
    assert rc == 0, _lib.last_error()
    return bad.value, first.value, checked.value


def test_sqrt_fast_exhaustive(gpu):
    # All patterns of +0 .. +Inf: the fast path accepts +0 and [2^-96, FLT_MAX].
    bad, first, checked = selfcheck(0, 0, 0x7F800001)
    assert checked == (0x7F7FFFFF - 0x0F800000 + 1) + 1  # [2^-96, FLT_MAX] and +0
    assert bad == 0, f"sqrt fast path differs at bits {first:#x}"


def test_atan_small_exhaustive(gpu):
    bad, first, checked = selfcheck(1, 0, 0x3ED413CD + 1)  # float32(0.41421356) bits
    assert checked == 0x3ED413CD + 1
    assert bad == 0, f"small atan differs at bits {first:#x}"


def test_slope_small_path_exhaustive(gpu):
    """Every s = gx^2 + gy^2 the fused slope test accepts: +0 and
    [2^-96, 0x3e2fb0cd], whose rounded roots are all <= tan(pi/8)."""
    bad, first, checked = selfcheck(3, 0, 0x3E2FB0CD + 1)
    assert checked == (0x3E2FB0CD - 0x0F800000 + 1) + 1
    assert bad == 0, f"slope fast path differs at bits {first:#x}"


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_div_fast_random(gpu, seed):
    bad, first, checked = selfcheck(2, seed, 1 << 31)
    assert checked == 1 << 31
    assert bad == 0, f"fast quotient differs at pair {first}"
