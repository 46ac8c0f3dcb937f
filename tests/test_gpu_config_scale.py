"""Parity at the benchmark configurations' own scale, against the oracle.

Every case runs on the 20000x20000 fBm DEM (Hurst 0.8, 0..3000 m, 0.5%
NoData in circular blobs of radius 5..50) that the terrain-suite,
batching and serving benchmarks use, built on the device from fixed
seeds:

- the six terrain-suite ops (Slope, Aspect, Hillshade at azimuth 315 and
  altitude 45, profile, plan and mean Curvature; CellSize 30) as
  ProcessChunked with TileHeight 256 over device-resident operands, the
  top and bottom 2000 rows compared bit for bit (data and validity) with
  the oracle run over those rows of the whole raster;
- 64 batch tiles with sides drawn log-uniformly from 256 to 2048 at random
  offsets, ops drawn from {Slope, Hillshade, mean Curvature, focal Mean
  r=3}, run as one batch, each tile against the oracle over its window;
- served requests of 2048x2048 and 4096x4096 (and 512x512) windows through
  the Slope -> Lookup -> Reclass chain, each against the oracle's three
  calls over its window.

The reference computes each window as a raster of its own (edges at the
window's borders), which is how both the batch and the served requests
are specified (internal/exec/process.go:28, terrain/terrain.go).
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import focal, synth, terrain, transfer
from strata_amd.batch import Batch, Task
from strata_amd.engine import MemorySink, MemorySource, Options, ProcessChunked
from strata_amd.raster import Float32Raster
from strata_amd.serve import Server

from helpers import same_floats

pytestmark = pytest.mark.gpu

N = 20000
STRIP = 2000
THREADS = 16
LOOKUP = transfer.newLookupKernel("transfer.Lookup", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])
RECLASS = transfer.newReclassKernel("transfer.Reclass", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5])


def suite_ops():
    T = terrain
    return [
        ("slope", T.slopeOp(T.SlopeOptions(CellSize=30))),
        ("aspect", T.aspectOp(T.AspectOptions(CellSize=30))),
        ("hillshade", T.hillshadeOp(T.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45, ZFactor=1))),
        ("curvature-profile", T.curvatureOp(T.CurvatureOptions(CellSize=30, Type=T.CurvatureProfile))),
        ("curvature-plan", T.curvatureOp(T.CurvatureOptions(CellSize=30, Type=T.CurvaturePlan))),
        ("curvature-mean", T.curvatureOp(T.CurvatureOptions(CellSize=30, Type=T.CurvatureMean))),
    ]


@pytest.fixture(scope="module")
def dem(gpu):
    """The DEM on the device and its host copy."""
    import torch
    data = synth.fbm_dem(N, N, hurst=0.8, seed=1, lo=0.0, hi=3000.0, device=gpu)
    valid, _ = synth.blob_nodata(N, N, frac=0.005, rmin=5, rmax=50, seed=2, device=gpu)
    torch.cuda.synchronize()
    dev = Float32Raster(data, N, N, N, valid, 0)
    host = Float32Raster(data.cpu().numpy(), N, N, N, valid.cpu().numpy().view(np.uint64).copy(), 0)
    yield dev, host
    del dev, data, valid
    torch.cuda.empty_cache()


def window(r: Float32Raster, x: int, y: int, w: int, h: int) -> Float32Raster:
    return r.Window(x, y, w, h)


def compact(w: int, h: int, device=None) -> Float32Raster:
    words = (w * h + 63) // 64
    if device is None:
        return Float32Raster(np.zeros(w * h, np.float32), w, h, w, np.zeros(words, np.uint64), 0)
    import torch
    return Float32Raster(torch.zeros(w * h, dtype=torch.float32, device=device), w, h, w,
                         torch.zeros(words, dtype=torch.int64, device=device), 0)


def check_rows(got_data, got_words, want: Float32Raster, y0: int, y1: int, what: str):
    a, b = y0 * N, y1 * N
    eq = same_floats(got_data, want.Data[a:b])
    assert eq.all(), (f"{what}: rows {y0}..{y1}: {int((~eq).sum())} cells differ, first at row "
                      f"{y0 + int(np.argmin(eq)) // N} col {int(np.argmin(eq)) % N}")
    # Rows of N cells start on word boundaries here (N * STRIP is a multiple of 64).
    assert a % 64 == 0 and b % 64 == 0
    want_words = want.Valid[a // 64:b // 64]
    bad = np.nonzero(got_words != want_words)[0]
    assert len(bad) == 0, f"{what}: rows {y0}..{y1}: {len(bad)} mask words differ, first at word {a // 64 + bad[0]}"


@pytest.mark.parametrize("name", [n for n, _ in suite_ops()])
def test_terrain_suite_chunked_256_strips_match_oracle(dem, gpu, name):
    import torch
    dev, host = dem
    k = dict(suite_ops())[name]
    out = compact(N, N, gpu)
    assert ProcessChunked(None, [MemorySink(out)], [MemorySource(dev)], k, Options(TileHeight=256)) is None
    torch.cuda.synchronize()
    want = compact(N, N)
    for y0 in (0, N - STRIP):
        y1 = y0 + STRIP
        oracle.process(k, [want], [host], y0, y1, threads=THREADS)
        got = out.Data[y0 * N:y1 * N].cpu().numpy()
        words = out.Valid[y0 * N // 64:y1 * N // 64].cpu().numpy().view(np.uint64)
        check_rows(got, words, want, y0, y1, f"{name} Chunked TileHeight 256")
    del out
    torch.cuda.empty_cache()


def test_batch_of_64_config_tiles_matches_oracle(dem, gpu):
    import torch
    dev, host = dem
    rng = np.random.default_rng(4096)
    ops = [terrain.slopeOp(terrain.SlopeOptions(CellSize=30)),
           terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45)),
           terrain.curvatureOp(terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureMean)),
           focal.newBox(focal.BoxOptions(3), "mean")]
    tiles = []
    for i in range(64):
        w, h = (int(v) for v in np.exp(rng.uniform(np.log(256), np.log(2048), 2)))
        x, y = int(rng.integers(0, N - w + 1)), int(rng.integers(0, N - h + 1))
        tiles.append((ops[int(rng.integers(0, len(ops)))], x, y, w, h))
    tasks = [Task(k, window(dev, x, y, w, h), compact(w, h, gpu)) for k, x, y, w, h in tiles]
    with Batch(tasks, graph=True) as b:
        b.run()
        b.sync()
    torch.cuda.synchronize()
    for i, ((k, x, y, w, h), t) in enumerate(zip(tiles, tasks)):
        want = compact(w, h)
        oracle.process(k, [want], [window(host, x, y, w, h)], threads=THREADS)
        got = t.dst.Data.cpu().numpy()
        eq = same_floats(got, want.Data)
        assert eq.all(), f"tile {i} {w}x{h} at ({x}, {y}): {int((~eq).sum())} cells differ"
        gw = t.dst.Valid.cpu().numpy().view(np.uint64)
        assert np.array_equal(gw, want.Valid), f"tile {i} {w}x{h} at ({x}, {y}): mask differs"


def test_served_config_windows_match_oracle(dem, gpu):
    dev, host = dem
    slope = terrain.slopeOp(terrain.SlopeOptions(CellSize=30))
    rng = np.random.default_rng(5000)
    wins = []
    for side in (4096, 2048, 512, 2048, 4096, 512, 512):
        wins.append((int(rng.integers(0, N - side + 1)), int(rng.integers(0, N - side + 1)), side, side))
    wins.append((N - 4096, N - 4096, 4096, 4096))  # the DEM's far corner
    with Server(dev, slope, (LOOKUP, RECLASS), slots=4, hbm_budget=2 << 30) as srv:
        ids = [srv.submit(*w) for w in wins]
        for rid in ids:
            assert srv.wait(rid) >= 0.0
        for rid, (x, y, w, h) in zip(ids, wins):
            data, mask = srv.fetch(rid, w, h)
            srv.release(rid)
            want = compact(w, h)
            oracle.process(slope, [want], [window(host, x, y, w, h)], threads=THREADS)
            oracle.process(LOOKUP, [want], [want], threads=THREADS)
            oracle.process(RECLASS, [want], [want], threads=THREADS)
            eq = same_floats(data, want.Data)
            assert eq.all(), f"request {w}x{h} at ({x}, {y}): {int((~eq).sum())} cells differ"
            bits_got = np.unpackbits(mask.view(np.uint8), bitorder="little")[: w * h]
            bits_want = np.unpackbits(want.Valid.view(np.uint8), bitorder="little")[: w * h]
            assert np.array_equal(bits_got, bits_want), f"request {w}x{h} at ({x}, {y}): mask differs"
