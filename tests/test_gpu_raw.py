"""File-to-file Chunked on the device against the oracle.

A raw file (engine/raw.go) is read by the library tile by tile with its
halo; the fill value's NoData is derived on the device; the op runs; the
fill goes into every invalid output cell on the device; the tile is written
back. The expected output file is the oracle's result over the raster that
RawSource.ReadWindow gives on the host, written through RawSink.WriteWindow:
the reference's definition of the same run with in-memory operands.
"""
import os

import numpy as np
import pytest

from oracle import oracle
from strata_amd import engine, terrain
from strata_amd import reduce as rd
from strata_amd.engine import (MemorySink, Options, OpenRawFile, CreateRawFile, PlanStats, RawOptions, RawSink,
                               RawSource, SourceError, Stats)
from strata_amd.raster import Float32Raster

pytestmark = pytest.mark.gpu

FILL = -9999.0


def dem_file(path, W, H, seed, fill=FILL, nodata=0.03):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    z = (500 + 40 * np.sin(xx / 17.0) * np.cos(yy / 11.0) + rng.normal(0, 3, (H, W))).astype(np.float32)
    z[rng.random((H, W)) < nodata] = fill
    z.astype("<f4").tofile(path)
    return z


def host_read(path, W, H, opts):
    """The whole file as RawSource.ReadWindow gives it."""
    f = OpenRawFile(path, os.O_RDONLY)
    r = Float32Raster(np.zeros(W * H, np.float32), W, H, W,
                      np.zeros((W * H + 63) // 64 + 1, np.uint64) if opts.HasFill else None, 0)
    assert RawSource(f, W, H, opts).ReadWindow(None, r, 0, 0) is None
    f.Close()
    return r


def expected_file(k, src, W, H, fill_opts, path):
    """The oracle's output over src, written through RawSink."""
    masked = src.Valid is not None or fill_opts.HasFill
    out = Float32Raster(np.zeros(W * H, np.float32), W, H, W,
                        np.zeros((W * H + 63) // 64 + 1, np.uint64) if masked else None, 0)
    oracle.process(k, [out], [src])
    f = CreateRawFile(path, 4 * W * H)
    assert RawSink(f, W, H, fill_opts).WriteWindow(None, out, 0, 0) is None
    f.Close()
    return np.fromfile(path, dtype="<f4")


def same_bits(a, b):
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("tw,th", [(0, 37), (100, 64), (0, 0)])
def test_slope_file_to_file(gpu, tmp_path, tw, th):
    W, H = 301, 203
    ipath, opath, epath = (str(tmp_path / n) for n in ("in.raw", "out.raw", "want.raw"))
    dem_file(ipath, W, H, seed=1)
    opts = RawOptions(Fill=FILL, HasFill=True)
    k = terrain.slopeOp(terrain.SlopeOptions(CellSize=30))
    want = expected_file(k, host_read(ipath, W, H, opts), W, H, opts, epath)
    fi, fo = OpenRawFile(ipath, os.O_RDONLY), CreateRawFile(opath, 4 * W * H)
    st = Stats()
    err = terrain.SlopeChunked(None, RawSink(fo, W, H, opts), RawSource(fi, W, H, opts),
                               terrain.SlopeOptions(CellSize=30), Options(TileWidth=tw, TileHeight=th, Stats=st))
    fi.Close()
    fo.Close()
    assert err is None
    assert same_bits(np.fromfile(opath, dtype="<f4"), want)
    assert st == PlanStats(k, W, H, Options(TileWidth=tw, TileHeight=th), chunked=True)


def test_features_file_to_files(gpu, tmp_path):
    """Mixed radii from one raw DEM to four raw outputs, each with its own
    ring and validity."""
    W, H = 257, 131
    ipath = str(tmp_path / "in.raw")
    dem_file(ipath, W, H, seed=2)
    opts = RawOptions(Fill=FILL, HasFill=True)
    ops = [terrain.SlopeOptions(CellSize=30), terrain.AspectOptions(CellSize=30, FitRadius=3),
           terrain.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45),
           terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureMean)]
    src_host = host_read(ipath, W, H, opts)
    wants = [expected_file(terrain._feature_kernel(i, op), src_host, W, H, opts, str(tmp_path / f"w{i}.raw"))
             for i, op in enumerate(ops)]
    fi = OpenRawFile(ipath, os.O_RDONLY)
    fos = [CreateRawFile(str(tmp_path / f"o{i}.raw"), 4 * W * H) for i in range(len(ops))]
    err = terrain.FeaturesChunked(None, [terrain.FeatureSink(op, RawSink(fos[i], W, H, opts))
                                         for i, op in enumerate(ops)],
                                  RawSource(fi, W, H, opts), Options(TileHeight=50))
    assert err is None
    for f in fos + [fi]:
        f.Close()
    for i in range(len(ops)):
        assert same_bits(np.fromfile(str(tmp_path / f"o{i}.raw"), dtype="<f4"), wants[i]), i


def test_nan_fill_and_unmasked(gpu, tmp_path):
    """A NaN fill (every NaN is NoData), and a file without a fill into a
    device memory sink (no validity, every cell valid)."""
    import torch
    W, H = 190, 77
    ipath = str(tmp_path / "in.raw")
    dem_file(ipath, W, H, seed=3, fill=np.nan)
    k = terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=10, Azimuth=200, Altitude=30))
    nan_opts = RawOptions(Fill=float("nan"), HasFill=True)
    want = expected_file(k, host_read(ipath, W, H, nan_opts), W, H, nan_opts, str(tmp_path / "w.raw"))
    fi, fo = OpenRawFile(ipath, os.O_RDONLY), CreateRawFile(str(tmp_path / "o.raw"), 4 * W * H)
    err = engine.ProcessChunked(None, [RawSink(fo, W, H, nan_opts)], [RawSource(fi, W, H, nan_opts)], k,
                                Options(TileHeight=16))
    fo.Close()
    assert err is None
    assert same_bits(np.fromfile(str(tmp_path / "o.raw"), dtype="<f4"), want)
    # No fill: the NaNs are values.
    src = host_read(ipath, W, H, RawOptions())
    ref = Float32Raster(np.zeros(W * H, np.float32), W, H, W, None, 0)
    oracle.process(k, [ref], [src])
    out = Float32Raster(torch.zeros(W * H, dtype=torch.float32, device=gpu), W, H, W, None, 0)
    err = engine.ProcessChunked(None, [MemorySink(out)], [RawSource(fi, W, H)], k, Options(TileHeight=16))
    fi.Close()
    assert err is None
    got = out.Data.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(ref.Data))
    ok = ~np.isnan(got)
    assert same_bits(got[ok], ref.Data[ok])


def test_stats_of_a_raw_file(gpu, tmp_path):
    W, H = 333, 91
    ipath = str(tmp_path / "in.raw")
    dem_file(ipath, W, H, seed=4)
    opts = RawOptions(Fill=FILL, HasFill=True)
    want = rd.Stats(host_read(ipath, W, H, opts))
    fi = OpenRawFile(ipath, os.O_RDONLY)
    got, err = rd.StatsChunked(None, RawSource(fi, W, H, opts), Options(TileHeight=20))
    fi.Close()
    assert err is None and got == want


def test_short_file_writes_a_prefix(gpu, tmp_path):
    """A file that ends early: the tiles before the failing one are written,
    and the error names the row (engine/raw.go:86-92 wrapped as
    internal/exec/chunked.go:379-393)."""
    W, H = 64, 40
    ipath, opath = str(tmp_path / "in.raw"), str(tmp_path / "out.raw")
    z = dem_file(ipath, W, H, seed=5)
    with open(ipath, "r+b") as fh:
        fh.truncate(4 * W * 30)
    opts = RawOptions(Fill=FILL, HasFill=True)
    fi, fo = OpenRawFile(ipath, os.O_RDONLY), CreateRawFile(opath, 4 * W * H)
    err = terrain.SlopeChunked(None, RawSink(fo, W, H, opts), RawSource(fi, W, H, opts),
                               terrain.SlopeOptions(CellSize=30), Options(TileHeight=10))
    fi.Close()
    fo.Close()
    assert isinstance(err, SourceError)
    assert "reading src[0]" in str(err) and "reading row 30: unexpected EOF" in str(err)
    got = np.fromfile(opath, dtype="<f4").reshape(H, W)
    assert np.all(got[:20] != 0) and np.all(got[20:] == 0)  # tiles 0 and 1 written, the rest untouched


def test_large_tiles_split_across_io_threads(gpu, tmp_path):
    """Tiles of 8 MiB are read and written in parts by the library's IO
    threads (full-width rows: one pread/pwrite per part); the file is the
    one the row-by-row reference writes."""
    W, H = 2048, 1500
    ipath, opath, epath = (str(tmp_path / n) for n in ("in.raw", "out.raw", "want.raw"))
    dem_file(ipath, W, H, seed=6)
    opts = RawOptions(Fill=FILL, HasFill=True)
    k = terrain.slopeOp(terrain.SlopeOptions(CellSize=30))
    want = expected_file(k, host_read(ipath, W, H, opts), W, H, opts, epath)
    fi, fo = OpenRawFile(ipath, os.O_RDONLY), CreateRawFile(opath, 4 * W * H)
    err = terrain.SlopeChunked(None, RawSink(fo, W, H, opts), RawSource(fi, W, H, opts),
                               terrain.SlopeOptions(CellSize=30), Options(TileHeight=1024))
    fi.Close()
    fo.Close()
    assert err is None
    assert same_bits(np.fromfile(opath, dtype="<f4"), want)


def test_large_tile_short_file_names_first_missing_row(gpu, tmp_path):
    """A file that ends inside a tile read in parts: the error names the
    first row that is missing, as a row-by-row read would."""
    W, H = 2048, 1200
    ipath, opath = str(tmp_path / "in.raw"), str(tmp_path / "out.raw")
    dem_file(ipath, W, H, seed=7)
    with open(ipath, "r+b") as fh:
        fh.truncate(4 * W * 700 + 12)  # row 700 only partly there
    opts = RawOptions(Fill=FILL, HasFill=True)
    fi, fo = OpenRawFile(ipath, os.O_RDONLY), CreateRawFile(opath, 4 * W * H)
    err = terrain.SlopeChunked(None, RawSink(fo, W, H, opts), RawSource(fi, W, H, opts),
                               terrain.SlopeOptions(CellSize=30), Options(TileHeight=1024))
    fi.Close()
    fo.Close()
    assert isinstance(err, SourceError)
    assert "reading src[0]" in str(err) and "reading row 700: unexpected EOF" in str(err), str(err)
