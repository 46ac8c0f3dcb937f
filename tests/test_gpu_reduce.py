"""Reductions on the device against the oracle: the fold's exact integers,
tallies and extremes bit for bit, and the finished Summary.

The reference's promises (reduce/doc.go): only valid cells take part and
data under an invalid cell is never read; a NaN is a value; -0 < +0; every
result is independent of layout, tiling, workers and backend; a
cancelled reduction returns no value. Cases follow reduce/stats_test.go
(known values at widths around the vector and word sizes, a DEM-like
raster masked and windowed) plus wide exponent spreads, and at full size
the exact sums are checked against math.fsum.
"""
import math

import numpy as np
import pytest

from oracle import reduce_ref as R
from strata_amd import _lib, engine
from strata_amd import reduce as rd
from strata_amd.engine import MemorySource, Options, Stats
from strata_amd.raster import Float32Raster

from test_oracle_reduce import known_cases

pytestmark = pytest.mark.gpu

F32 = np.float32


def to_device(r: Float32Raster, dev):
    import torch
    v = None if r.Valid is None else torch.from_numpy(np.asarray(r.Valid).view(np.int64).copy()).to(dev)
    return Float32Raster(torch.from_numpy(np.asarray(r.Data).copy()).to(dev), r.Width, r.Height, r.Stride, v,
                         r.ValidOffset)


def pack(bits: np.ndarray, nwords: int) -> np.ndarray:
    out = np.zeros(nwords, dtype=np.uint64)
    idx = np.nonzero(bits)[0]
    np.bitwise_or.at(out, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    return out


def lay_out(vals, valid, w):
    """A w-wide raster holding vals in order; cells past them invalid."""
    n = len(vals)
    h = max(1, -(-n // w))
    data = np.zeros(w * h, dtype=np.float32)
    data[:n] = vals if n else data[:0]
    bits = np.zeros(w * h, dtype=bool)
    bits[:n] = True if valid is None else valid
    masked = valid is not None or w * h != n
    mask = pack(bits, (w * h + 63) // 64 + 1) if masked else None
    if mask is not None:
        data[~bits] = np.nan  # never read
    return Float32Raster(data, w, h, w, mask, 0)


def oracle_of(r: Float32Raster):
    vals = R.valid_values(np.asarray(r.Data), r.Width, r.Height, r.Stride,
                          None if r.Valid is None else np.asarray(r.Valid), r.ValidOffset)
    return R.fold(vals), R.summary(vals)


def check_fold(f: _lib.Fold, want: dict, what=_lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS):
    assert f.count == want["count"]
    if what == _lib.REDUCE_COUNT:
        return  # validity alone: no cell is read
    assert (f.nan, f.pos_inf, f.neg_inf) == (want["nan"], want["pos_inf"], want["neg_inf"])
    assert bool(f.not_neg_zero) == want["not_neg_zero"]
    if want["count"]:
        assert R.f32_bits(F32(f.min)) == R.f32_bits(want["min"]) or (np.isnan(f.min) and np.isnan(want["min"]))
        assert R.f32_bits(F32(f.max)) == R.f32_bits(want["max"]) or (np.isnan(f.max) and np.isnan(want["max"]))
    if what & (_lib.REDUCE_SUM | _lib.REDUCE_MOMENTS):
        assert _lib.limbs_to_int(f.sum) == want["x1"]
    if what & _lib.REDUCE_MOMENTS:
        assert _lib.limbs_to_int(f.sumsq) == want["x2"]


def same_summary(got: rd.Summary, want: dict):
    def s64(a, b):
        return (a != a and b != b) or (a == b and math.copysign(1, a) == math.copysign(1, b))

    def s32(a, b):
        return (a != a and b != b) or R.f32_bits(a) == R.f32_bits(b)

    assert got.Count == want["Count"]
    assert s64(got.Sum, want["Sum"]), (got.Sum, want["Sum"])
    assert s64(got.Mean, want["Mean"]), (got.Mean, want["Mean"])
    sd = want["StdDev"]
    assert (got.StdDev != got.StdDev and sd != sd) or abs(got.StdDev - sd) <= math.ulp(sd), (got.StdDev, sd)
    assert s32(got.Min, want["Min"]) and s32(got.Max, want["Max"])


@pytest.mark.parametrize("name", list(known_cases()))
@pytest.mark.parametrize("w", [0, 1, 7, 8, 9, 63, 64, 65, 130])
def test_known_values(gpu, name, w):
    vals, _ = known_cases()[name]
    valid = None
    if name == "all invalid":
        vals, valid = [F32(1), F32(2), F32(3)], [False, False, False]
    w = w or max(1, len(vals))
    if w < len(vals):
        pytest.skip("narrower than the data")
    r = lay_out(vals, valid, w)
    fw, sw = oracle_of(r)
    for where in ("host", "device"):
        rr = r if where == "host" else to_device(r, gpu)
        f, err = engine.Reduce(None, [rr], _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS)
        assert err is None
        check_fold(f, fw)
        same_summary(rd.Stats(rr), sw)


def dem_like(seed, w, h, masked):
    rng = np.random.default_rng(seed)
    data = (300 + 2000 * rng.random(w * h)).astype(np.float32)
    mask = None
    if masked:
        mask = pack(rng.integers(0, 10, w * h) != 0, (w * h + 63) // 64 + 1)
    return Float32Raster(data, w, h, w, mask, 0)


def fold_stats(W, H, tw, th, chunked):
    """Closed-form counters of a fold (internal/exec/reduce.go:242-257,
    reducechunked.go:186-205), restated from the band loop."""
    s = Stats()
    band_cells = 1 << 16

    def bands(x0, y0, w, h, tw_, th_):
        tW = min(tw_, w) if tw_ > 0 else w
        tH = min(th_, h) if th_ > 0 else h
        rows = max(1, band_cells // tW)
        for ty in range(0, h, tH):
            for tx in range(0, w, tW):
                for by in range(ty, min(ty + tH, h), rows):
                    c = (min(tx + tW, w) - tx) * (min(by + rows, ty + tH, h) - by)
                    s.Cells += c
                    s.Bands += 1
                    s.KernelRead += 4 * c

    if not chunked:
        bands(0, 0, W, H, tw, th)
    else:
        tW = min(tw, W) if tw > 0 else W
        tH = min(th, H) if th > 0 else H
        for y0 in range(0, H, tH):
            for x0 in range(0, W, tW):
                w, h = min(tW, W - x0), min(tH, H - y0)
                s.Tiles += 1
                s.SourceRead += 4 * w * h
                bands(x0, y0, w, h, 0, 0)
    s.Ideal = s.Cells * 4
    return s


@pytest.mark.parametrize("kind", ["unmasked", "masked", "window"])
def test_dem_every_path(gpu, kind):
    """TestStatsDEM (reduce/stats_test.go:207-223): every form and Options
    value gives the plain function's bits."""
    r = dem_like(21, 301, 67, kind != "unmasked")
    if kind == "window":
        r = r.Window(3, 2, 301 - 7, 67 - 5)
    fw, sw = oracle_of(r)
    d = to_device(r, gpu)
    plain = rd.Stats(d)
    same_summary(plain, sw)
    for opts in (Options(), Options(TileWidth=64, TileHeight=16), Options(TileHeight=7), Options(TileWidth=300)):
        for chunked in (False, True):
            for src in ("device", "host"):
                st = Stats()
                o = Options(opts.TileWidth, opts.TileHeight, opts.Workers, st)
                rr = d if src == "device" else r
                if chunked:
                    s, err = rd.StatsChunked(None, MemorySource(rr), o)
                else:
                    s, err = rd.StatsTiled(None, rr, o)
                assert err is None
                assert s == plain
                assert st == fold_stats(r.Width, r.Height, opts.TileWidth, opts.TileHeight, chunked)
    assert rd.Count(d) == fw["count"]
    mn, mx, n = rd.MinMax(d)
    assert (mn, mx, n) == (fw["min"], fw["max"], fw["count"])
    s, n = rd.Sum(d)
    assert s == sw["Sum"] and n == fw["count"]


@pytest.mark.parametrize("seed", range(4))
def test_wide_spread_with_specials(gpu, seed):
    """Exponents across the whole float32 range (subnormals to the largest),
    signed zeros and a few specials, under a window with a mid-word mask
    offset: the device's exact integers equal the oracle's."""
    rng = np.random.default_rng(50 + seed)
    W, H = 517, 41
    vals = (rng.standard_normal(W * H) * np.exp2(rng.integers(-140, 120, W * H))).astype(np.float32)
    vals[rng.integers(0, W * H, 40)] = -0.0
    vals[rng.integers(0, W * H, 40)] = 0.0
    if seed >= 2:
        vals[rng.integers(0, W * H, 3)] = np.inf
        vals[rng.integers(0, W * H, 2)] = np.nan
    base = 37
    mask = pack(np.concatenate([np.zeros(base, bool), rng.random(W * H) > 0.2]), (base + W * H + 63) // 64 + 1)
    r = Float32Raster(vals, W, H, W, mask, base).Window(5, 3, W - 11, H - 4)
    fw, sw = oracle_of(r)
    d = to_device(r, gpu)
    for what in (_lib.REDUCE_COUNT, _lib.REDUCE_MINMAX, _lib.REDUCE_SUM, _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS):
        f, err = engine.Reduce(None, [d], what)
        assert err is None
        check_fold(f, fw, what)
    same_summary(rd.Stats(d), sw)


def window_values(rng, n, field0, spread):
    """n float32s with random signs and full 24-bit significands whose
    exponent fields lie in [field0, field0 + spread) (field 0 is a
    subnormal), a few of them zero."""
    fields = rng.integers(field0, field0 + spread, n).astype(np.uint32)
    mant = rng.integers(0, 1 << 23, n, dtype=np.uint32)
    sign = rng.integers(0, 2, n, dtype=np.uint32) << np.uint32(31)
    vals = (sign | (fields << np.uint32(23)) | mant).view(np.float32)
    vals[rng.integers(0, n, 5)] = 0.0
    return vals


@pytest.mark.parametrize("field0", [0, 120, 244, 235])
@pytest.mark.parametrize("spread", [10, 11, 12, 19, 20, 21])
def test_exponent_windows(gpu, field0, spread):
    """Steps whose exponents fill the fold's fast window exactly (11 bins
    with squares, 20 without), or overflow it by one, at the bottom,
    middle and top of the exponent range: the exact integers are the
    oracle's either way."""
    if field0 + spread > 255:
        pytest.skip("past the largest exponent")
    rng = np.random.default_rng(1000 * field0 + spread)
    W, H = 2048, 3
    vals = window_values(rng, W * H, field0, spread)
    r = Float32Raster(vals, W, H, W, None, 0)
    fw, sw = oracle_of(r)
    d = to_device(r, gpu)
    for what in (_lib.REDUCE_SUM, _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS):
        f, err = engine.Reduce(None, [d], what)
        assert err is None
        check_fold(f, fw, what)


def test_invalid_data_never_read(gpu):
    """TestStatsInvalidDataNeverRead (reduce/stats_test.go:265-286)."""
    r = dem_like(7, 200, 30, True)
    want = rd.Stats(to_device(r, gpu))
    rng = np.random.default_rng(1)
    bits = np.unpackbits(np.asarray(r.Valid).view(np.uint8), bitorder="little")[: 200 * 30].astype(bool)
    data = np.asarray(r.Data).copy()
    data[~bits] = rng.choice(np.array([np.nan, np.inf, -np.inf, 1e30], np.float32), int((~bits).sum()))
    got = rd.Stats(to_device(Float32Raster(data, 200, 30, 200, r.Valid, 0), gpu))
    assert got == want


def test_cancelled_returns_no_value(gpu):
    """TestStatsCancelledReturnsNoValue (reduce/stats_test.go:288-308)."""
    r = to_device(dem_like(3, 64, 64, True), gpu)
    ctx, cancel = engine.WithCancel()
    cancel()
    s, n, err = rd.SumTiled(ctx, r, Options())
    assert (s, n) == (0.0, 0) and isinstance(err, engine.Canceled)
    s, n, err = rd.SumChunked(ctx, MemorySource(r), Options(TileHeight=8))
    assert (s, n) == (0.0, 0) and isinstance(err, engine.Canceled)
    s, err = rd.StatsTiled(ctx, r, Options())
    assert s == rd.Summary() and err is not None
    s, err = rd.StatsChunked(ctx, MemorySource(r), Options(TileHeight=8))
    assert s == rd.Summary() and err is not None


def test_full_size_exact_sums(gpu):
    """4096² fBm DEM with blob NoData: the exact sums against math.fsum (a
    correctly rounded sum) of the valid values and of their squares, which
    float64 holds exactly; count and extremes against numpy."""
    import torch
    from strata_amd import synth
    N = 4096
    data = synth.fbm_dem(N, N, hurst=0.8, seed=1, lo=0.0, hi=3000.0, device=gpu)
    valid, _ = synth.blob_nodata(N, N, frac=0.01, rmin=5, rmax=50, seed=2, device=gpu)
    r = Float32Raster(data, N, N, N, valid, 0)
    f, err = engine.Reduce(None, [r], _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS)
    assert err is None
    host = data.cpu().numpy()
    bits = np.unpackbits(valid.cpu().numpy().view(np.uint8), bitorder="little")[: N * N].astype(bool)
    v = host[bits].astype(np.float64)
    assert f.count == v.size
    assert _lib.limbs_to_int(f.sum) / (1 << 150) == math.fsum(v)
    assert _lib.limbs_to_int(f.sumsq) / (1 << 300) == math.fsum(v * v)
    assert F32(f.min) == host[bits].min() and F32(f.max) == host[bits].max()
    torch.cuda.synchronize()
