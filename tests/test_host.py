"""Host-side logic on the CPU: the C-ABI library loads and exports every
declared symbol, the closed-form Stats equal an independent restatement of
the reference's band and tile loops, option resolution gives the
reference's constants, and programming errors raise before anything
runs, with the reference's message prefixes."""
import math
import os
import re

import numpy as np
import pytest

from strata_amd import _lib, engine, focal, terrain, transfer
from strata_amd.engine import EnginePanic, Options, PlanStats
from strata_amd.raster import Float32Raster, NewFloat32, NewFloat32Like, NewMask

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "strata_hip.h")).read()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(strata_\w+)\s*\(", header, re.M))
    assert {"strata_process", "strata_process_chunked", "strata_plan_stats"} <= declared
    L = _lib.lib()
    for name in declared:
        assert hasattr(L, name), f"{name} declared in include/strata_hip.h but not exported"
    assert set(_lib.EXPORTS) == declared
    assert _lib.lib().strata_version().decode().startswith("strata-hip")


# -- Stats: an independent restatement of internal/exec/tile.go, halo.go, chunked.go -------

BAND_CELLS = 1 << 16


def cdiv(a, b):
    return (a + b - 1) // b


def plan_bands(w, h, tw, th):
    tileW = min(tw, w) if tw > 0 else w
    tileH = min(th, h) if th > 0 else h
    band_rows = max(1, BAND_CELLS // tileW)
    out = []
    for ty in range(0, h, tileH):
        for tx in range(0, w, tileW):
            for by in range(ty, min(ty + tileH, h), band_rows):
                out.append((tx, by, min(tx + tileW, w), min(by + band_rows, ty + tileH, h)))
    return out


def band_stats(s, r, W, H, nin, nout, b, rmin=None):
    # job.band (internal/exec/halo.go:21-57): the kernel runs outside the
    # narrowest edge ring, rmin, and reads its window grown by r.
    lo = r if rmin is None else rmin
    x0, y0, x1, y1 = b
    cells = (x1 - x0) * (y1 - y0)
    ix0, ix1, iy0, iy1 = max(x0, lo), min(x1, W - lo), max(y0, lo), min(y1, H - lo)
    interior = ix0 < ix1 and iy0 < iy1
    inner = (ix1 - ix0) * (iy1 - iy0) if interior else 0
    s.Cells += cells
    s.Bands += 1
    s.KernelWritten += (cells - inner) * 4 * nout
    if interior:
        s.KernelRead += (ix1 - ix0 + 2 * r) * (iy1 - iy0 + 2 * r) * 4 * nin
        s.KernelWritten += inner * 4 * nout


def ref_stats(k, W, H, tw, th, chunked, rmin=None):
    s = engine.Stats()
    r, nin, nout = k.radius, k.nin, k.nout
    if not chunked:
        for b in plan_bands(W, H, tw, th):
            band_stats(s, r, W, H, nin, nout, b, rmin)
    else:
        tileW = min(tw, W) if tw > 0 else W
        tileH = min(th, H) if th > 0 else H
        for y0 in range(0, H, tileH):
            for x0 in range(0, W, tileW):
                x1, y1 = min(x0 + tileW, W), min(y0 + tileH, H)
                rx0, ry0, rx1, ry1 = max(0, x0 - r), max(0, y0 - r), min(W, x1 + r), min(H, y1 + r)
                s.Tiles += 1
                s.SourceRead += (rx1 - rx0) * (ry1 - ry0) * 4 * nin
                for b in plan_bands(x1 - x0, y1 - y0, 0, 0):
                    band_stats(s, r, W, H, nin, nout, (b[0] + x0, b[1] + y0, b[2] + x0, b[3] + y0), rmin)
                s.SinkWritten += (x1 - x0) * (y1 - y0) * 4 * nout
    s.Ideal = s.Cells * 4 * (nin + nout)
    return s


KERNELS = [terrain.slopeOp(terrain.SlopeOptions(CellSize=30)),
           terrain.gradientOp(terrain.GradientOptions(CellSize=1)),
           terrain.slopeOp(terrain.SlopeOptions(CellSize=30, FitRadius=3)),
           focal.newBox(focal.BoxOptions(5), "mean"),
           transfer.newReclassKernel("t", [1], [1, 2])]


@pytest.mark.parametrize("k", KERNELS, ids=lambda k: f"{k.name}-r{k.radius}")
@pytest.mark.parametrize("W,H", [(1, 1), (2, 9), (17, 5), (300, 200), (70000, 3), (4096, 4096)])
@pytest.mark.parametrize("tw,th", [(0, 0), (1, 0), (7, 5), (64, 1), (256, 256), (0, 256), (100000, 100000)])
@pytest.mark.parametrize("chunked", [False, True])
def test_plan_stats_match_reference_loops(k, W, H, tw, th, chunked):
    if W * H > 1 << 22 and (tw, th) in ((1, 0), (7, 5), (64, 1)):
        pytest.skip("too many units for the Python restatement")
    got = PlanStats(k, W, H, Options(TileWidth=tw, TileHeight=th), chunked=chunked)
    assert got == ref_stats(k, W, H, tw, th, chunked)


def test_stats_known_values():
    """engine/stats.go: KernelWritten is exactly Cells*4*outputs; a Chunked
    pointwise call is about 2.0 Amplification; Halo is 0 at radius 0."""
    k = transfer.newReclassKernel("t", [1], [1, 2])
    s = PlanStats(k, 1000, 1000, Options(TileHeight=100), chunked=True)
    assert s.KernelWritten == s.Cells * 4 and s.Halo(1) == 0
    assert abs(s.Amplification() - 2.0) < 1e-9
    k = terrain.slopeOp(terrain.SlopeOptions(CellSize=1))
    s = PlanStats(k, 20000, 20000, Options(TileHeight=256), chunked=True)
    assert s.Tiles == 79 and s.Cells == 400_000_000
    assert s.KernelWritten == s.Cells * 4


FEATURE_STACKS = [
    [terrain.SlopeOptions(CellSize=30), terrain.AspectOptions(CellSize=30)],
    [terrain.SlopeOptions(CellSize=30), terrain.SlopeOptions(CellSize=30, FitRadius=4),
     terrain.CurvatureOptions(CellSize=30, FitRadius=2)],
    [terrain.HillshadeOptions(CellSize=30, FitRadius=3, Azimuth=315, Altitude=45)],
]


@pytest.mark.parametrize("si", range(len(FEATURE_STACKS)))
@pytest.mark.parametrize("W,H", [(1, 1), (5, 3), (17, 5), (300, 200), (70000, 3)])
@pytest.mark.parametrize("tw,th", [(0, 0), (7, 5), (64, 1), (256, 256)])
@pytest.mark.parametrize("chunked", [False, True])
def test_fanout_plan_stats_match_reference_loops(si, W, H, tw, th, chunked):
    """A Features pipeline (terrain/features.go:103-117) has the largest
    stage radius; every output's ring is its own radius, so the kernel runs
    outside the narrowest one (internal/exec/tile.go:136-184)."""
    if W * H > 1 << 18 and (tw, th) in ((7, 5), (64, 1)):
        pytest.skip("too many units for the Python restatement")
    k = terrain.newFeatures(FEATURE_STACKS[si])
    rmin = min(p.radius for p in k.parts)
    got = PlanStats(k, W, H, Options(TileWidth=tw, TileHeight=th), chunked=chunked)
    assert got == ref_stats(k, W, H, tw, th, chunked, rmin)


def test_features_panics():
    """terrain.newFeatures (terrain/features.go:106-113) and each op's own
    option checks, before anything runs."""
    with pytest.raises(terrain.TerrainPanic, match="^terrain: Features has no outputs"):
        terrain.newFeatures([])
    with pytest.raises(terrain.TerrainPanic, match="^terrain: Features output 1 has no Op"):
        terrain.newFeatures([terrain.SlopeOptions(CellSize=1), None])
    with pytest.raises(terrain.TerrainPanic, match="^terrain: Features output 0 has Op GradientOptions"):
        terrain.newFeatures([terrain.GradientOptions(CellSize=1)])
    with pytest.raises(terrain.TerrainPanic, match="^terrain: FitRadius"):
        terrain.newFeatures([terrain.SlopeOptions(CellSize=1, FitRadius=9)])
    with pytest.raises(EnginePanic, match="^engine: stage 0 takes 1 inputs and 2 outputs"):
        engine.FanOut([terrain.gradientOp(terrain.GradientOptions(CellSize=1))])
    k = terrain.newFeatures([terrain.SlopeOptions(CellSize=1), terrain.AspectOptions(CellSize=1, FitRadius=3)])
    assert (k.radius, k.nin, k.nout) == (3, 1, 2)
    def ras(w, h):
        return NewFloat32(w, h, np.zeros(w * h, np.float32))

    dem = ras(8, 8)
    with pytest.raises(EnginePanic, match="kernel takes 1 inputs and 2 outputs, got 1 and 1"):
        engine.check([ras(8, 8)], [dem], k, Options())
    with pytest.raises(EnginePanic, match="dst\\[1\\] is 7×8"):
        engine.check([ras(8, 8), ras(7, 8)], [dem], k, Options())


# -- option resolution --------------------------------------------------------------

def f32bits(x):
    return int(np.float32(x).view(np.uint32))


def test_horn_scales_and_slope_scale():
    kx, ky = terrain.cellSizes(30, 0, 0)
    assert f32bits(kx) == f32bits(1 / 240) and kx == ky
    kx, ky = terrain.cellSizes(10, 5, 2)
    assert f32bits(kx) == f32bits(2 / 80) and f32bits(ky) == f32bits(2 / 40)
    assert f32bits(terrain.slopeScale(terrain.SlopeDegrees)[0]) == 0x42652EE1
    assert terrain.slopeScale(terrain.SlopeRadians) == (1.0, True)
    assert terrain.slopeScale(terrain.SlopePercent) == (100.0, False)


def test_hillshade_light_defaults():
    c, bx, by = terrain.hillshadeLight(0, 0)  # 315 and 45
    s = math.sqrt(0.5)
    assert f32bits(c) == f32bits(255 * math.sin(math.radians(45)))
    assert abs(bx - 255 * s * s) < 1e-3 and abs(by - 255 * s * s) < 1e-3
    assert terrain.hillshadeLight(1e308, 45) == terrain.hillshadeLight(math.fmod(1e308, 360), 45)


def test_zt_and_wood_scales():
    kp, kq, kr, kt, ks = terrain.ZTScales(3.0, 4.0, 1.0)
    assert (f32bits(kp), f32bits(kq), f32bits(kr), f32bits(kt), f32bits(ks)) == (
        f32bits(1 / 6), f32bits(1 / 8), f32bits(1 / 9), f32bits(1 / 16), f32bits(1 / 48))
    # r = 1: K = 3, S = 2, Q = 6 (taps -1, 2, -1 squared)
    kp, kq, kr, kt, ks = terrain.woodFit(1, 1, 0, 0)
    assert (kp, kq, kr, kt, ks) == tuple(float(np.float32(v)) for v in (1 / 6, 1 / 6, 1 / 3, 1 / 3, 1 / 4))


@pytest.mark.parametrize("fn,prefix", [
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=0)), "terrain: CellSize"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=math.inf)), "terrain: CellSize"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1, ZFactor=math.nan)), "terrain: ZFactor"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1, Units=7)), "terrain: unknown SlopeUnits"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1e-40)), "terrain: ZFactor/(8"),
    (lambda: terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=1, Altitude=91)), "terrain: Altitude"),
    (lambda: terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=1, Azimuth=math.inf)), "terrain: Azimuth"),
    (lambda: terrain.curvatureOp(terrain.CurvatureOptions(CellSize=1, Type=3)), "terrain: unknown CurvatureType"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1, FitRadius=9)), "terrain: FitRadius"),
    (lambda: focal.newBox(focal.BoxOptions(0), "mean"), "focal: Radius"),
    (lambda: focal.newWeightsKernel(focal.WeightsOptions(1, [1] * 8), False), "focal: Radius 1 needs"),
    (lambda: focal.newWeightsKernel(focal.WeightsOptions(1, [1] * 8 + [math.nan]), False), "focal: Weights"),
    (lambda: focal.newSeparableKernel(focal.SeparableOptions(1, [1, 1], [1, 1, 1])), "focal: Radius 1 needs"),
    (lambda: transfer.newLookupKernel("transfer.Lookup", [], []), "transfer.Lookup: the table"),
    (lambda: transfer.newLookupKernel("transfer.Lookup", [1, 1], [1, 2]), "transfer.Lookup: xs must be strictly"),
    (lambda: transfer.newReclassKernel("transfer.Reclass", [1], [1]), "transfer.Reclass: values must"),
])
def test_option_panics(fn, prefix):
    with pytest.raises(EnginePanic) as e:
        fn()
    assert str(e.value).startswith(prefix)


@pytest.mark.parametrize("fn,prefix", [
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1, FitRadius=-1)), "terrain: FitRadius"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1, FitRadius=9)), "terrain: FitRadius"),
    (lambda: terrain.newSurface((False, False, True, False, False), terrain.SurfaceOptions(CellSize=1, FitRadius=-1)),
     "terrain: FitRadius"),
    (lambda: terrain.newSurface((False, False, True, False, False), terrain.SurfaceOptions(CellSize=1, FitRadius=9)),
     "terrain: FitRadius"),
    (lambda: terrain.curvatureOp(terrain.CurvatureOptions(FitRadius=2)), "terrain: CellSize"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1, FitRadius=2, Units=7)), "terrain: unknown SlopeUnits"),
    (lambda: terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=1, FitRadius=2, Altitude=95)), "terrain: Altitude"),
    (lambda: terrain.slopeOp(terrain.SlopeOptions(CellSize=1e30, ZFactor=1e-30, FitRadius=2)),
     "terrain: the quadratic fit's factors"),
])
def test_fit_panics(fn, prefix):
    """terrain/fit_test.go:250-264: out-of-range fit radii (ops and Surface),
    a fit without a cell size, bad units or altitude with a fit, and fit
    factors that underflow float32."""
    with pytest.raises(terrain.TerrainPanic) as e:
        fn()
    assert str(e.value).startswith(prefix)


def host(w, h, masked=False, stride=None):
    stride = stride or w
    n = (h - 1) * stride + w
    return Float32Raster(np.zeros(n, np.float32), w, h, stride, NewMask(n) if masked else None, 0)


def test_operand_panics_before_any_work():
    k = terrain.slopeOp(terrain.SlopeOptions(CellSize=1))
    with pytest.raises(EnginePanic, match="^engine: an input has a validity mask"):
        engine.ProcessN(None, [host(5, 5)], [host(5, 5, masked=True)], k)
    with pytest.raises(EnginePanic, match=r"^engine: src\[0\] is 5×4, dst\[0\] is 5×5"):
        engine.ProcessN(None, [host(5, 5)], [host(5, 4)], k)
    with pytest.raises(EnginePanic, match="^engine: kernel takes 1 inputs and 1 outputs"):
        engine.ProcessN(None, [host(5, 5), host(5, 5)], [host(5, 5)], k)
    with pytest.raises(EnginePanic, match="^engine: negative Options"):
        engine.ProcessN(None, [host(5, 5)], [host(5, 5)], k, Options(TileWidth=-1))
    parent = host(10, 10)
    with pytest.raises(EnginePanic, match=r"^engine: dst\[0\] and src\[0\] share Data"):
        engine.ProcessN(None, [parent.Window(0, 0, 5, 5)], [parent.Window(2, 2, 5, 5)], k)
    with pytest.raises(EnginePanic, match=r"^engine: dst\[0\]: raster: stride"):
        engine.ProcessN(None, [Float32Raster(np.zeros(25, np.float32), 5, 5, 4)], [host(5, 5)], k)
    g = terrain.gradientOp(terrain.GradientOptions(CellSize=1))
    d = host(5, 5)
    with pytest.raises(EnginePanic, match=r"^engine: dst\[0\] and dst\[1\] share Data"):
        engine.ProcessN(None, [d, d], [host(5, 5)], g)
    # A radius-0 kernel may run in place; at a different offset it may not.
    r0 = transfer.newLookupKernel("t", [0], [1])
    p = host(10, 10)
    with pytest.raises(EnginePanic, match="^engine: dst.0. overlaps src.0. at a different offset"):
        engine.ProcessN(None, [p.Window(1, 0, 5, 5)], [p.Window(0, 0, 5, 5)], r0)


def test_chunked_panics():
    k = terrain.slopeOp(terrain.SlopeOptions(CellSize=1))
    src = engine.MemorySource(host(6, 6, masked=True))
    with pytest.raises(EnginePanic, match="^engine: a source is Masked but dst.0. is not"):
        engine.ProcessChunked(None, [engine.MemorySink(host(6, 6))], [src], k)
    with pytest.raises(EnginePanic, match=r"^engine: src\[0\] is 6×6, dst\[0\] is 6×5"):
        engine.ProcessChunked(None, [engine.MemorySink(host(6, 5, masked=True))], [src], k)
    shared = host(6, 6, masked=True)
    with pytest.raises(EnginePanic, match="^engine: memory sink dst.0. shares memory with memory source"):
        engine.ProcessChunked(None, [engine.MemorySink(shared)], [engine.MemorySource(shared)], k)


def test_raster_model():
    r = NewFloat32(5, 4, np.arange(20, dtype=np.float32))
    r.Valid = NewMask(20)
    w = r.Window(1, 1, 3, 2)
    assert w.Stride == 5 and w.ValidOffset == 6 and list(w.Row(1)) == [11, 12, 13]
    w.SetValid(2, 1, False)
    assert not r.IsValid(3, 2) and r.IsValid(2, 2)
    assert w.Validate() is None
    assert Float32Raster(np.zeros(3, np.float32), 2, 2).Validate().startswith("raster: data has 3 cells")
    with pytest.raises(EnginePanic.__mro__[1]):
        r.Window(3, 3, 3, 3)


CHAIN_STACKS = [
    [terrain.slopeOp(terrain.SlopeOptions(CellSize=30)), transfer.newLookupKernel("t", [0, 5], [1, 2]),
     transfer.newReclassKernel("t", [1.5], [1, 2])],
    [focal.newBox(focal.BoxOptions(2), "mean"), terrain.slopeOp(terrain.SlopeOptions(CellSize=30))],
    [focal.newBox(focal.BoxOptions(1), "max"), focal.newBox(focal.BoxOptions(2), "mean"),
     terrain.curvatureOp(terrain.CurvatureOptions(CellSize=10))],
]


@pytest.mark.parametrize("si", range(len(CHAIN_STACKS)))
@pytest.mark.parametrize("W,H", [(1, 1), (5, 3), (17, 5), (300, 200), (70000, 3)])
@pytest.mark.parametrize("tw,th", [(0, 0), (7, 5), (64, 1), (256, 256)])
@pytest.mark.parametrize("chunked", [False, True])
def test_chain_plan_stats_match_reference_loops(si, W, H, tw, th, chunked):
    """A chain is one kernel whose radius is the sum of its stages' radii
    (internal/exec/pipeline.go:32-41): its Stats are that kernel's."""
    pipe = engine.Pipeline(CHAIN_STACKS[si])
    assert pipe.radius == sum(k.radius for k in CHAIN_STACKS[si])
    got = PlanStats(pipe, W, H, Options(TileWidth=tw, TileHeight=th), chunked=chunked)
    assert got == ref_stats(pipe, W, H, tw, th, chunked)


def test_chain_checks():
    with pytest.raises(EnginePanic):
        engine.Pipeline([])
    with pytest.raises(EnginePanic):
        engine.Pipeline([terrain.gradientOp(terrain.GradientOptions(CellSize=1))])
    too_far = engine.Pipeline([focal.newBox(focal.BoxOptions(5), "mean"), focal.newBox(focal.BoxOptions(4), "max")])
    with pytest.raises(EnginePanic, match="reaches 9"):
        PlanStats(too_far, 100, 100)


@pytest.mark.parametrize("op,xs,rest,msg", [
    (_lib.OP_LOOKUP, [0, float("nan"), 2], [1, 2, 3], "transfer.Lookup: xs must not hold NaN, but xs[1] is NaN"),
    (_lib.OP_LOOKUP, [0, 2, 2], [1, 2, 3], "transfer.Lookup: xs must be strictly increasing, but xs[2] = 2 is not "
                                           "above xs[1] = 2"),
    (_lib.OP_LOOKUP, [1, 0], [1, 2], "transfer.Lookup: xs must be strictly increasing, but xs[1] = 0 is not above"),
    (_lib.OP_LOOKUP, [-0.0, 0.0], [1, 2], "transfer.Lookup: xs must be strictly increasing"),
    (_lib.OP_LOOKUP, [0, float("inf")], [1, 2], "transfer.Lookup: xs must be finite, but xs[1] = inf"),
    (_lib.OP_LOOKUP, [float("-inf")], [1], "transfer.Lookup: xs must be finite, but xs[0] = -inf"),
    (_lib.OP_RECLASS, [float("nan")], [1, 2], "transfer.Reclass: breaks must not hold NaN, but breaks[0] is NaN"),
    (_lib.OP_RECLASS, [3, 1], [1, 2, 3], "transfer.Reclass: breaks must be strictly increasing, but breaks[1] = 1"),
])
def test_c_abi_refuses_tables_the_reference_refuses(op, xs, rest, msg):
    """The C ABI checks Lookup and Reclass tables as the reference's
    constructors do (transfer/transfer.go:176-221): the device kernels count
    the entries at or below a value, which equals the reference's scan only
    for a strictly increasing table free of NaN."""
    taps = np.asarray(list(xs) + list(rest), np.float32)
    k = _lib.Kernel()
    k.op, k.radius, k.ntaps, k.taps = op, 0, len(taps), taps.ctypes.data
    st = _lib.Stats()
    rc = _lib.lib().strata_plan_stats(ctypes_ref(k), 1, 1, 10, 10, None, 0, ctypes_ref(st))
    assert rc == _lib.EINVAL
    assert _lib.last_error().startswith(msg)


@pytest.mark.parametrize("op,taps", [
    (_lib.OP_LOOKUP, [0, 1]),                                   # one knot
    (_lib.OP_LOOKUP, [-1e30, 1e30, float("nan"), float("inf")]),  # ys are not checked
    (_lib.OP_RECLASS, [5]),                                     # no breaks
    (_lib.OP_RECLASS, [float("-inf"), float("inf"), 1, 2, 3]),  # infinite breaks are ordered values
])
def test_c_abi_accepts_tables_the_reference_accepts(op, taps):
    taps = np.asarray(taps, np.float32)
    k = _lib.Kernel()
    k.op, k.radius, k.ntaps, k.taps = op, 0, len(taps), taps.ctypes.data
    st = _lib.Stats()
    assert _lib.lib().strata_plan_stats(ctypes_ref(k), 1, 1, 10, 10, None, 0, ctypes_ref(st)) == 0


def ctypes_ref(x):
    import ctypes
    return ctypes.byref(x)
