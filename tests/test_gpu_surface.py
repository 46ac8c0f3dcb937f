"""terrain.Surface against the standalone ops, bit for bit.

The reference's claim (terrain/surface.go:53-60) is that each product
Surface writes is exactly what Gradient, Slope, Aspect or Hillshade write
with the same settings. The oracle computes each product with its
standalone op; the device computes any subset in one pass. Whole parent
buffers are compared, so writes outside the outputs fail too.
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import terrain
from strata_amd.engine import MemorySink, MemorySource, Options

from helpers import Scene, assert_same_parents, host_copy

pytestmark = pytest.mark.gpu

NAMES = ("Dx", "Dy", "Slope", "Aspect", "Hillshade")


def standalone(opts: terrain.SurfaceOptions):
    g = terrain.gradientOp(terrain.GradientOptions(opts.CellSize, opts.CellSizeY, opts.ZFactor, opts.FitRadius))
    return {
        "Slope": terrain.slopeOp(terrain.SlopeOptions(opts.CellSize, opts.CellSizeY, opts.ZFactor, opts.FitRadius,
                                                      opts.Units)),
        "Aspect": terrain.aspectOp(terrain.AspectOptions(opts.CellSize, opts.CellSizeY, opts.ZFactor, opts.FitRadius,
                                                         opts.ZeroForFlat, opts.Trigonometric)),
        "Hillshade": terrain.hillshadeOp(terrain.HillshadeOptions(opts.CellSize, opts.CellSizeY, opts.ZFactor,
                                                                  opts.FitRadius, opts.Azimuth, opts.Altitude)),
        "Gradient": g,
    }


def scene(kind, want, seed, W, H):
    rng = np.random.default_rng(seed)
    sc = Scene()
    masked = kind != "compact"
    if kind == "windows":
        sc.add_parent("src", rng, W + 13, H + 7, masked=True, valid_base=41)
        sc.add_window("s", "src", 5, 3, W, H)
    else:
        sc.add_parent("src", rng, W, H, masked=masked)
        sc.add_window("s", "src", 0, 0, W, H)
    for i, name in enumerate(NAMES):
        if not want[i]:
            continue
        if kind == "windows":
            sc.add_parent(name, rng, W + 3 + 8 * i, H + 2, masked=True, valid_base=7 + 11 * i)
            sc.add_window(name, name, 2, 1, W, H)
        else:
            sc.add_parent(name, rng, W, H, masked=masked)
            sc.add_window(name, name, 0, 0, W, H)
    return sc


def expected(sc, want, opts):
    host, bufs = sc.materialize(None)
    ks = standalone(opts)
    if want[0] or want[1]:
        # Gradient writes both; a scratch copy stands in for an unwanted one.
        scratch = {n: host[n] if want[i] else None for i, n in enumerate(("Dx", "Dy"))}
        ref = host["s"]
        tmp = []
        for n in ("Dx", "Dy"):
            if scratch[n] is None:
                from strata_amd.raster import Float32Raster
                tmp.append(Float32Raster(np.zeros(ref.Width * ref.Height, np.float32), ref.Width, ref.Height,
                                         ref.Width, None if ref.Valid is None else
                                         np.zeros((ref.Width * ref.Height + 63) // 64, np.uint64), 0))
            else:
                tmp.append(scratch[n])
        oracle.process(ks["Gradient"], tmp, [host["s"]])
    for i, n in enumerate(("Slope", "Aspect", "Hillshade")):
        if want[2 + i]:
            oracle.process(ks[n], [host[n]], [host["s"]])
    return bufs


SUBSETS = [(1, 1, 1, 1, 1), (0, 0, 1, 0, 0), (1, 0, 0, 1, 0), (0, 1, 1, 0, 1), (0, 0, 0, 0, 1)]
OPTS = [
    terrain.SurfaceOptions(CellSize=30, Azimuth=315, Altitude=45),
    terrain.SurfaceOptions(CellSize=3, CellSizeY=4, ZFactor=2, Units=terrain.SlopePercent, ZeroForFlat=True,
                           Trigonometric=True, Azimuth=-117.5, Altitude=23),
    terrain.SurfaceOptions(CellSize=30, FitRadius=2, Azimuth=315, Altitude=45),
    terrain.SurfaceOptions(CellSize=10, FitRadius=5, Units=terrain.SlopeRadians, Azimuth=90, Altitude=60),
]


@pytest.mark.parametrize("oi", range(len(OPTS)))
@pytest.mark.parametrize("want", SUBSETS)
@pytest.mark.parametrize("kind", ["compact", "masked", "windows"])
def test_surface_matches_standalone(gpu, oi, want, kind):
    import torch
    opts = OPTS[oi]
    sc = scene(kind, want, seed=17 + oi, W=203, H=97)
    want_b = expected(sc, want, opts)
    dev, dbufs = sc.materialize(gpu)
    out = terrain.SurfaceOutputs(**{n: dev[n] for i, n in enumerate(NAMES) if want[i]})
    terrain.Surface(out, dev["s"], opts)
    torch.cuda.synchronize()
    assert_same_parents(host_copy(dbufs), want_b, f"Surface {want} opts {oi} {kind}")


def test_surface_chunked(gpu):
    import torch
    opts = OPTS[0]
    want = (1, 1, 1, 1, 1)
    sc = scene("masked", want, seed=5, W=611, H=301)
    want_b = expected(sc, want, opts)
    dev, dbufs = sc.materialize(gpu)
    sinks = terrain.SurfaceSinks(**{n: MemorySink(dev[n]) for n in NAMES})
    err = terrain.SurfaceChunked(None, sinks, MemorySource(dev["s"]), opts, Options(TileHeight=64))
    assert err is None
    torch.cuda.synchronize()
    assert_same_parents(host_copy(dbufs), want_b, "SurfaceChunked")


def test_surface_needs_an_output(gpu):
    with pytest.raises(terrain.TerrainPanic, match="no outputs"):
        terrain.newSurface([False] * 5, OPTS[0])
