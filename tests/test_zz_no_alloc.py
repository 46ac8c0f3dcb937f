"""strata_reduce_noalloc against strata_reduce, through the C ABI.

The no-allocation entry point folds in the library's static device scratch;
its partial must equal strata_reduce's field for field (the exact sums are
integers, so the two agree bit for bit). Two calls in a row, the second on
a smaller raster, check that the shared scratch keeps nothing of the first
call. A host raster and a cancelled context are refused before any device
work, as the header says.
"""
import ctypes

import numpy as np
import pytest

from strata_amd import _lib
from strata_amd.raster import Float32Raster

pytestmark = pytest.mark.gpu

WHATS = [_lib.REDUCE_COUNT, _lib.REDUCE_MINMAX, _lib.REDUCE_SUM, _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS]


def device_raster(gpu, seed, w, h, stride, offset, masked):
    import torch
    rng = np.random.default_rng(seed)
    n = (h - 1) * stride + w
    data = (rng.standard_normal(n) * 1000).astype(np.float32)
    data[rng.random(n) < 0.001] = np.nan
    data[rng.random(n) < 0.001] = -0.0
    data[rng.random(n) < 0.0005] = np.inf
    mask = None
    if masked:
        nbits = offset + n
        mask = torch.from_numpy(rng.integers(-2**63, 2**63 - 1, (nbits + 63) // 64 + 1, dtype=np.int64)).to(gpu)
    return Float32Raster(torch.from_numpy(data).to(gpu), w, h, stride, mask, offset)


def fold_bytes(f: _lib.Fold) -> bytes:
    return ctypes.string_at(ctypes.addressof(f), ctypes.sizeof(f))


def both(r: Float32Raster, what: int):
    L = _lib.lib()
    c = r.to_c()
    fa, fb = _lib.Fold(), _lib.Fold()
    sa, sb = _lib.Stats(), _lib.Stats()
    opts = _lib.Options()
    assert L.strata_reduce(what, ctypes.byref(c), ctypes.byref(opts), ctypes.byref(fa), ctypes.byref(sa)) == 0, \
        _lib.last_error()
    rc = L.strata_reduce_noalloc(what, ctypes.byref(c), ctypes.byref(opts), ctypes.byref(fb), ctypes.byref(sb))
    assert rc == 0, _lib.last_error()
    return fa, fb, sa, sb


@pytest.mark.parametrize("what", WHATS)
def test_noalloc_equals_reduce_twice(gpu, what):
    big = device_raster(gpu, 1, 1531, 1207, 1600, 37, True)
    small = device_raster(gpu, 2, 211, 83, 211, 5, True)
    for r in (big, small):
        fa, fb, sa, sb = both(r, what)
        print(f"noalloc what={what:#x} {r.Width}x{r.Height}: count {fb.count}", flush=True)
        assert fold_bytes(fa) == fold_bytes(fb)
        assert ctypes.string_at(ctypes.addressof(sa), ctypes.sizeof(sa)) == \
            ctypes.string_at(ctypes.addressof(sb), ctypes.sizeof(sb))
        assert fb.count > 0


def test_noalloc_unmasked_and_windowed(gpu):
    parent = device_raster(gpu, 3, 900, 700, 900, 0, False)
    import torch
    win = Float32Raster(parent.Data[70 * 900 + 13:], 500, 400, 900, None, 0)
    for r in (parent, win):
        fa, fb, _, _ = both(r, _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS)
        assert fold_bytes(fa) == fold_bytes(fb)
        assert fb.count == r.Width * r.Height
    torch.cuda.synchronize()


def test_noalloc_refuses_host_rasters(gpu):
    L = _lib.lib()
    data = np.zeros(64, np.float32)
    r = Float32Raster(data, 8, 8, 8, None, 0).to_c()
    f, opts = _lib.Fold(), _lib.Options()
    rc = L.strata_reduce_noalloc(_lib.REDUCE_SUM, ctypes.byref(r), ctypes.byref(opts), ctypes.byref(f), None)
    assert rc == _lib.EINVAL
    assert "device rasters only" in _lib.last_error()
