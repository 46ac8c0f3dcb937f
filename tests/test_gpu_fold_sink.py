"""Statistics folded into the pass that computes them (graph/sinks.go:
17-71): a tee sink folds every tile of an output before any sink sees it,
and a tee with no sinks writes nothing at all.

The expected partial is the reduction of the output as the op writes it:
the oracle computes the output on the host, and the library's plain
reduction (pinned against the oracle's exact fold in test_gpu_reduce.py)
folds it. The tee's partial must equal it field for field — the exact
sums as integers, the extremes bit for bit — whatever the tiling, the
step of rows the resident pass uses, the kind of source and sink, and
with the outputs' own validity (only valid cells fold when the sink is
masked; with no masks the NaN ring folds too). For small rasters the
finished Summary is also checked against the oracle's own Stats."""
import os

import numpy as np
import pytest

from oracle import oracle
from oracle import reduce_ref as R
from strata_amd import _lib, engine, terrain
from strata_amd import reduce as rd
from strata_amd.engine import (Canceled, MemorySink, MemorySource, Options, PlanStats, RawOptions, RawSink,
                               RawSource, Reduce, Stats, TeeSink, WithCancel, OpenRawFile, CreateRawFile)
from strata_amd.raster import Float32Raster

from test_gpu_reduce import check_fold, oracle_of, same_summary, to_device

pytestmark = pytest.mark.gpu

SLOPE = terrain.SlopeOptions(CellSize=30)


def dem(W, H, seed, masked=True, nodata=0.03):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    z = (500 + 40 * np.sin(xx / 17.0) * np.cos(yy / 11.0) + rng.normal(0, 3, (H, W))).astype(np.float32)
    valid = None
    if masked:
        bits = rng.random(W * H) >= nodata
        valid = np.zeros((W * H + 63) // 64 + 1, np.uint64)
        idx = np.nonzero(bits)[0]
        np.bitwise_or.at(valid, idx >> 6, np.left_shift(np.uint64(1), (idx & 63).astype(np.uint64)))
    return Float32Raster(z.reshape(-1), W, H, W, valid, 0)


def empty(W, H, masked, dev=None):
    import torch
    n = W * H
    words = (n + 63) // 64 + 1
    if dev is None:
        return Float32Raster(np.zeros(n, np.float32), W, H, W, np.zeros(words, np.uint64) if masked else None, 0)
    return Float32Raster(torch.zeros(n, dtype=torch.float32, device=dev), W, H, W,
                         torch.zeros(words, dtype=torch.int64, device=dev) if masked else None, 0)


def oracle_output(k, src):
    out = empty(src.Width, src.Height, src.Valid is not None)
    oracle.process(k, [out], [src])
    return out


def fold_of(r):
    f, err = Reduce(None, [r], engine.STATS_FOLD, Options())
    assert err is None
    return f


def same_fold(a: _lib.Fold, b: _lib.Fold):
    assert (a.count, a.nan, a.pos_inf, a.neg_inf, a.not_neg_zero) == (b.count, b.nan, b.pos_inf, b.neg_inf,
                                                                      b.not_neg_zero)
    assert list(a.sum) == list(b.sum) and list(a.sumsq) == list(b.sumsq)
    if a.count:
        assert np.float32(a.min).view(np.uint32) == np.float32(b.min).view(np.uint32) or (
            np.isnan(a.min) and np.isnan(b.min))
        assert np.float32(a.max).view(np.uint32) == np.float32(b.max).view(np.uint32) or (
            np.isnan(a.max) and np.isnan(b.max))


@pytest.fixture
def stripe_mb():
    """Sets the resident pass's step of rows for one test, and restores it."""
    L = _lib.lib()
    saved = _lib.Tuning()
    assert L.strata_get_tuning(saved) == 0

    def set_mb(mb):
        t = _lib.Tuning()
        assert L.strata_get_tuning(t) == 0
        t.fold_stripe_mb = mb
        assert L.strata_set_tuning(t) == 0, _lib.last_error()
    yield set_mb
    assert L.strata_set_tuning(saved) == 0


@pytest.mark.parametrize("W,H,opts,mb", [
    (301, 203, dict(), 128),
    (301, 203, dict(TileHeight=37), 128),
    (1000, 333, dict(TileWidth=300, TileHeight=50), 1),  # steps of 256 rows
    (4096, 700, dict(TileHeight=256), 1),                # steps of 64 rows, the last one short
])
def test_stats_alone_resident(gpu, stripe_mb, W, H, opts, mb):
    stripe_mb(mb)
    src = dem(W, H, seed=W + H)
    k = terrain.slopeOp(SLOPE)
    want = fold_of(oracle_output(k, src))
    tee = TeeSink(W, H, True)  # no sinks: nothing is written
    st = Stats()
    err = terrain.SlopeChunked(None, tee, MemorySource(to_device(src, gpu)), SLOPE, Options(Stats=st, **opts))
    assert err is None
    same_fold(tee.Fold(), want)
    # Stats are those of the same call with a sink: folding changes no counter.
    assert st == PlanStats(k, W, H, Options(**opts), chunked=True)
    # The kept scratch goes back to the device; the next call maps it again.
    assert _lib.lib().strata_trim() == 0, _lib.last_error()
    again = TeeSink(W, H, True)
    assert terrain.SlopeChunked(None, again, MemorySource(to_device(src, gpu)), SLOPE, Options(**opts)) is None
    same_fold(again.Fold(), want)


def test_summary_against_the_oracle(gpu):
    W, H = 120, 77
    src = dem(W, H, seed=3)
    k = terrain.slopeOp(SLOPE)
    out = oracle_output(k, src)
    _, want = oracle_of(out)
    tee = TeeSink(W, H, True)
    assert terrain.SlopeChunked(None, tee, MemorySource(to_device(src, gpu)), SLOPE, Options(TileHeight=20)) is None
    same_summary(rd.Folded(tee), want)
    wf, _ = oracle_of(out)
    check_fold(tee.Fold(), wf)


def test_tee_writes_and_folds(gpu):
    """A tee over a device memory sink: the raster is the op's, bit for bit,
    and the fold is its reduction."""
    W, H = 517, 301
    src = dem(W, H, seed=5)
    k = terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45))
    want = oracle_output(k, src)
    out = empty(W, H, True, gpu)
    tee = TeeSink(W, H, True, [MemorySink(out)])
    err = engine.ProcessChunked(None, [tee], [MemorySource(to_device(src, gpu))], k, Options(TileHeight=64))
    assert err is None
    got = out.Data.cpu().numpy()
    assert np.array_equal(got.view(np.uint32)[_valid(want)], want.Data.view(np.uint32)[_valid(want)])
    assert np.array_equal(out.Valid.cpu().numpy().view(np.uint64)[:-1], want.Valid[:-1])
    same_fold(tee.Fold(), fold_of(want))


def _valid(r):
    bits = np.unpackbits(np.asarray(r.Valid).view(np.uint8), bitorder="little")[:r.Width * r.Height]
    return bits.astype(bool)


@pytest.mark.parametrize("sinks", [0, 1, 2])
def test_host_source_tile_loop(gpu, sinks):
    """A host memory source: tiles are staged through pinned buffers and
    each tile's output is folded from its device buffer."""
    W, H = 333, 190
    src = dem(W, H, seed=11)
    k = terrain.slopeOp(SLOPE)
    want = oracle_output(k, src)
    outs = [empty(W, H, True) for _ in range(sinks)]
    tee = TeeSink(W, H, True, [MemorySink(o) for o in outs])
    err = engine.ProcessChunked(None, [tee], [MemorySource(src)], k, Options(TileWidth=100, TileHeight=40))
    assert err is None
    same_fold(tee.Fold(), fold_of(want))
    ok = _valid(want)
    for o in outs:
        assert np.array_equal(o.Data.view(np.uint32)[ok], want.Data.view(np.uint32)[ok])


def test_raw_file_tee(gpu, tmp_path):
    """File to file with a tee: the fill goes into invalid cells after the
    fold, which never reads them."""
    W, H = 257, 131
    z = dem(W, H, seed=13, masked=False)
    data = np.asarray(z.Data).copy()
    rng = np.random.default_rng(1)
    data[rng.random(W * H) < 0.03] = -9999.0
    ipath, opath = str(tmp_path / "in.raw"), str(tmp_path / "out.raw")
    data.astype("<f4").tofile(ipath)
    opts = RawOptions(Fill=-9999.0, HasFill=True)
    k = terrain.slopeOp(SLOPE)
    fi = OpenRawFile(ipath, os.O_RDONLY)
    host_src = Float32Raster(np.zeros(W * H, np.float32), W, H, W, np.zeros((W * H + 63) // 64 + 1, np.uint64), 0)
    assert RawSource(fi, W, H, opts).ReadWindow(None, host_src, 0, 0) is None
    want = oracle_output(k, host_src)
    fo = CreateRawFile(opath, 4 * W * H)
    tee = TeeSink(W, H, True, [RawSink(fo, W, H, opts)])
    err = terrain.SlopeChunked(None, tee, RawSource(fi, W, H, opts), SLOPE, Options(TileHeight=30))
    fi.Close()
    fo.Close()
    assert err is None
    same_fold(tee.Fold(), fold_of(want))
    got = np.fromfile(opath, dtype="<f4")
    ok = _valid(want)
    assert np.array_equal(got.view(np.uint32)[ok], want.Data.view(np.uint32)[ok])
    assert np.all(got[~ok] == -9999.0)


def test_unmasked_ring_folds_as_nan(gpu):
    """No masks anywhere: every cell folds, the NaN ring included, so Sum
    and Mean are NaN and Count is every cell (exec.Reduce over an
    unmasked window)."""
    W, H = 200, 90
    src = dem(W, H, seed=17, masked=False)
    k = terrain.slopeOp(SLOPE)
    want = oracle_output(k, src)
    tee = TeeSink(W, H, False)
    assert terrain.SlopeChunked(None, tee, MemorySource(to_device(src, gpu)), SLOPE, Options()) is None
    f = tee.Fold()
    assert f.count == W * H and f.nan == 2 * W + 2 * H - 4
    same_fold(f, fold_of(want))
    s = rd.Folded(tee)
    assert s.Count == W * H and np.isnan(s.Sum) and np.isnan(s.Mean)


def test_features_with_tees(gpu, stripe_mb):
    """The fan-out form: some outputs only folded, some written and folded,
    one only written; each fold is its own output's reduction."""
    stripe_mb(1)
    W, H = 1024, 600
    src = dem(W, H, seed=19)
    T = terrain
    ops = [T.SlopeOptions(CellSize=30), T.AspectOptions(CellSize=30, FitRadius=2),
           T.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45),
           T.CurvatureOptions(CellSize=30, Type=T.CurvatureMean)]
    wants = [oracle_output(T._feature_kernel(i, op), src) for i, op in enumerate(ops)]
    written = [empty(W, H, True, gpu) for _ in ops]
    sinks = [TeeSink(W, H, True),                                 # folded only
             TeeSink(W, H, True, [MemorySink(written[1])]),       # written and folded
             MemorySink(written[2]),                              # written only
             TeeSink(W, H, True)]                                 # folded only
    err = T.FeaturesChunked(None, [T.FeatureSink(op, s) for op, s in zip(ops, sinks)],
                            MemorySource(to_device(src, gpu)), Options(TileHeight=128))
    assert err is None
    for i in (0, 1, 3):
        same_fold(sinks[i].Fold(), fold_of(wants[i]))
    for i in (1, 2):
        ok = _valid(wants[i])
        got = written[i].Data.cpu().numpy()
        assert np.array_equal(got.view(np.uint32)[ok], wants[i].Data.view(np.uint32)[ok]), i


def test_cancelled_folds_nothing(gpu):
    W, H = 128, 128
    src = dem(W, H, seed=23)
    ctx, cancel = WithCancel(None)
    cancel()
    tee = TeeSink(W, H, True)
    err = terrain.SlopeChunked(ctx, tee, MemorySource(src), SLOPE, Options(TileHeight=16))
    assert isinstance(err, Canceled)
    assert tee.Fold().count == 0 and rd.Folded(tee).Count == 0


def test_write_window_folds_each_window(gpu):
    """WriteWindow called directly folds its window with the library's
    reduction; windows combine into the reduction of the whole."""
    W, H = 90, 61
    r = oracle_output(terrain.slopeOp(SLOPE), dem(W, H, seed=29))
    tee = TeeSink(W, H, True)
    for y0, y1 in ((0, 20), (20, 21), (21, 61)):
        assert tee.WriteWindow(None, r.Window(0, y0, W, y1 - y0), 0, y0) is None
    same_fold(tee.Fold(), fold_of(r))
