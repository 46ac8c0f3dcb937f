"""terrain.Features against the standalone ops, bit for bit.

The reference's claim (terrain/features.go:52-59): each output Features
writes is bit for bit what its own function writes with the same options,
Data and validity — its border is its own window's radius and its
validity is the DEM's mask eroded by its own window, not by the largest
one. The oracle computes every output with its standalone op; the device
computes the whole stack in one call (radius-1 Horn ops in one shared
sweep, fit ops beside it). Whole parent buffers are compared, so writes
outside the outputs fail too, and the call's Stats must equal the closed
form of the reference's band and tile loops for the pipeline.
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import terrain
from strata_amd.engine import MemorySink, MemorySource, Options, PlanStats, Stats

from helpers import Scene, assert_same_parents, host_copy

pytestmark = pytest.mark.gpu

T = terrain
STACKS = {
    # Every radius-1 Horn op, one shared sweep.
    "horn4": [T.SlopeOptions(CellSize=30), T.AspectOptions(CellSize=30),
              T.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45),
              T.CurvatureOptions(CellSize=30, Type=T.CurvatureProfile)],
    # Mixed radii: rings 1..4 side by side, validity eroded per output.
    "mixed": [T.SlopeOptions(CellSize=3, CellSizeY=4, ZFactor=2, Units=T.SlopePercent),
              T.SlopeOptions(CellSize=30, FitRadius=2),
              T.AspectOptions(CellSize=10, FitRadius=4, ZeroForFlat=True, Trigonometric=True),
              T.CurvatureOptions(CellSize=30, Type=T.CurvatureMean),
              T.HillshadeOptions(CellSize=30, FitRadius=1, Azimuth=90, Altitude=60)],
    # Seven radius-1 ops in one sweep.
    "horn7": [T.SlopeOptions(CellSize=30), T.SlopeOptions(CellSize=30, Units=T.SlopeRadians),
              T.SlopeOptions(CellSize=5, CellSizeY=7, Units=T.SlopePercent),
              T.AspectOptions(CellSize=30), T.AspectOptions(CellSize=2, Trigonometric=True, ZeroForFlat=True),
              T.HillshadeOptions(CellSize=1, ZFactor=3, Azimuth=-117.5, Altitude=23),
              T.CurvatureOptions(CellSize=30, Type=T.CurvaturePlan)],
    # Ten: more than one sweep's worth of outputs (8 per sweep).
    "horn10": [T.SlopeOptions(CellSize=30), T.SlopeOptions(CellSize=30, Units=T.SlopeRadians),
               T.AspectOptions(CellSize=30), T.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45),
               T.CurvatureOptions(CellSize=30, Type=T.CurvatureProfile),
               T.CurvatureOptions(CellSize=30, Type=T.CurvaturePlan),
               T.CurvatureOptions(CellSize=30, Type=T.CurvatureMean),
               T.SlopeOptions(CellSize=7, Units=T.SlopePercent),
               T.HillshadeOptions(CellSize=4, CellSizeY=3, Azimuth=10, Altitude=80),
               T.AspectOptions(CellSize=30, FitRadius=2)],
    # One fit feature alone.
    "fit1": [T.CurvatureOptions(CellSize=30, FitRadius=3, Type=T.CurvaturePlan)],
}


def scene(kind, n, seed, W, H):
    rng = np.random.default_rng(seed)
    sc = Scene()
    masked = kind != "compact"
    if kind == "windows":
        sc.add_parent("src", rng, W + 13, H + 7, masked=True, valid_base=41)
        sc.add_window("s", "src", 5, 3, W, H)
    else:
        sc.add_parent("src", rng, W, H, masked=masked)
        sc.add_window("s", "src", 0, 0, W, H)
    for i in range(n):
        name = f"o{i}"
        if kind == "windows":
            sc.add_parent(name, rng, W + 3 + 8 * i, H + 2, masked=True, valid_base=7 + 11 * i)
            sc.add_window(name, name, 2, 1, W, H)
        else:
            sc.add_parent(name, rng, W, H, masked=masked)
            sc.add_window(name, name, 0, 0, W, H)
    return sc


def expected(sc, ops):
    host, bufs = sc.materialize(None)
    for i, op in enumerate(ops):
        oracle.process(T._feature_kernel(i, op), [host[f"o{i}"]], [host["s"]])
    return bufs


@pytest.mark.parametrize("stack", list(STACKS))
@pytest.mark.parametrize("kind", ["compact", "masked", "windows"])
def test_features_match_standalone(gpu, stack, kind):
    import torch
    ops = STACKS[stack]
    sc = scene(kind, len(ops), seed=len(stack) + len(kind), W=203, H=97)
    want = expected(sc, ops)
    dev, dbufs = sc.materialize(gpu)
    T.Features([T.Feature(op, dev[f"o{i}"]) for i, op in enumerate(ops)], dev["s"])
    torch.cuda.synchronize()
    assert_same_parents(host_copy(dbufs), want, f"Features {stack} {kind}")


@pytest.mark.parametrize("stack", ["mixed", "horn10"])
@pytest.mark.parametrize("tw,th", [(64, 16), (0, 7), (37, 0)])
def test_features_tiled_stats(gpu, stack, tw, th):
    import torch
    ops = STACKS[stack]
    sc = scene("masked", len(ops), seed=3, W=150, H=71)
    want = expected(sc, ops)
    dev, dbufs = sc.materialize(gpu)
    st = Stats()
    opts = Options(TileWidth=tw, TileHeight=th, Stats=st)
    err = T.FeaturesTiled(None, [T.Feature(op, dev[f"o{i}"]) for i, op in enumerate(ops)], dev["s"], opts)
    assert err is None
    torch.cuda.synchronize()
    assert_same_parents(host_copy(dbufs), want, f"FeaturesTiled {stack} {tw}x{th}")
    assert st == PlanStats(T.newFeatures(ops), 150, 71, Options(TileWidth=tw, TileHeight=th))


@pytest.mark.parametrize("stack", ["mixed", "horn4"])
@pytest.mark.parametrize("where", ["device", "host"])
def test_features_chunked(gpu, stack, where):
    """Device memory sources and sinks run as the device tile queue; host
    ones go through the read/write callbacks tile by tile, each tile read
    grown by the largest radius."""
    import torch
    ops = STACKS[stack]
    W, H = 311, 133
    sc = scene("masked", len(ops), seed=11, W=W, H=H)
    want = expected(sc, ops)
    rasters, bufs = sc.materialize(gpu if where == "device" else None)
    st = Stats()
    opts = Options(TileWidth=128, TileHeight=40, Stats=st)
    sinks = [T.FeatureSink(op, MemorySink(rasters[f"o{i}"])) for i, op in enumerate(ops)]
    err = T.FeaturesChunked(None, sinks, MemorySource(rasters["s"]), opts)
    assert err is None
    if where == "device":
        torch.cuda.synchronize()
    assert_same_parents(host_copy(bufs), want, f"FeaturesChunked {stack} {where}")
    assert st == PlanStats(T.newFeatures(ops), W, H, Options(TileWidth=128, TileHeight=40), chunked=True)


def test_features_full_sweep_shape(gpu):
    """A raster wider than one sweep block and taller than one strip, with
    every radius-1 op: the shared sweep's pair path and its edges."""
    import torch
    ops = STACKS["horn4"]
    sc = scene("masked", len(ops), seed=29, W=1031, H=203)
    want = expected(sc, ops)
    dev, dbufs = sc.materialize(gpu)
    T.Features([T.Feature(op, dev[f"o{i}"]) for i, op in enumerate(ops)], dev["s"])
    torch.cuda.synchronize()
    assert_same_parents(host_copy(dbufs), want, "Features 1031x203")
