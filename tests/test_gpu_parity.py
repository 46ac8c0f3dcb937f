"""GPU parity: the HIP library against the oracle, bit for bit.

Every case builds the same operands twice — host copies for the oracle,
device copies for the library — runs both, and compares the whole parent
buffers: cells and mask words of the outputs (any NaN matches any NaN,
+0 != -0) and everything around them (row padding, a parent's other cells
and bits), which must be left untouched.
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import engine, focal, terrain, transfer
from strata_amd.engine import Context, MemorySink, MemorySource, Options, ProcessChunked, ProcessN

from helpers import Scene, assert_same_parents, host_copy, plane_values

pytestmark = pytest.mark.gpu


def terrain_kernels():
    ks = []
    for fit in (0, 1, 2, 5, 8):
        ks.append(("gradient", terrain.gradientOp(terrain.GradientOptions(CellSize=30, FitRadius=fit))))
        for u in (terrain.SlopeDegrees, terrain.SlopePercent, terrain.SlopeRadians):
            ks.append((f"slope{int(u)}", terrain.slopeOp(terrain.SlopeOptions(CellSize=30, FitRadius=fit, Units=u))))
        ks.append(("aspect", terrain.aspectOp(terrain.AspectOptions(CellSize=3, CellSizeY=4, FitRadius=fit))))
        ks.append(("aspect-trig0", terrain.aspectOp(terrain.AspectOptions(CellSize=5, FitRadius=fit,
                                                                          Trigonometric=True, ZeroForFlat=True))))
        ks.append(("hillshade", terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30, FitRadius=fit))))
        ks.append(("hillshade-az", terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=2, CellSizeY=7, ZFactor=3,
                                                                                FitRadius=fit, Azimuth=-117.5,
                                                                                Altitude=23))))
        for t in (terrain.CurvatureProfile, terrain.CurvaturePlan, terrain.CurvatureMean):
            ks.append((f"curv{int(t)}", terrain.curvatureOp(terrain.CurvatureOptions(CellSize=10, FitRadius=fit,
                                                                                      Type=t))))
    return ks


def focal_kernels():
    rng = np.random.default_rng(7)
    ks = []
    for r in range(1, 9):
        n = 2 * r + 1
        w = rng.uniform(-1, 1, n * n).astype(np.float32)
        w[rng.random(n * n) < 0.1] = 0  # zero weights are applied too
        ks.append((f"correlate{r}", focal.newWeightsKernel(focal.WeightsOptions(r, w), False)))
        ks.append((f"convolve{r}", focal.newWeightsKernel(focal.WeightsOptions(r, w), True)))
        g = focal.Gaussian(r, r / 2)
        ks.append((f"gauss{r}", focal.newSeparableKernel(focal.SeparableOptions(r, g, g))))
        ks.append((f"sep{r}", focal.newSeparableKernel(focal.SeparableOptions(
            r, rng.uniform(-2, 2, n).astype(np.float32), rng.uniform(-2, 2, n).astype(np.float32)))))
        for op in ("mean", "min", "max"):
            ks.append((f"{op}{r}", focal.newBox(focal.BoxOptions(r), op)))
    return ks


def transfer_kernels():
    return [
        ("lookup", transfer.newLookupKernel("t", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])),
        ("reclass", transfer.newReclassKernel("t", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5])),
        ("lookup1", transfer.newLookupKernel("t", [500], [7])),
        ("lookup-nan", transfer.newLookupKernel("t", [400, 450, 500, 600], [np.nan, -0.0, np.inf, 2])),
        ("reclass0", transfer.newReclassKernel("t", [], [3])),
    ]


ALL = terrain_kernels() + focal_kernels() + transfer_kernels()


def build_scene(kind: str, k, seed: int, W: int, H: int):
    """Operands for one case: `kind` picks the layout."""
    rng = np.random.default_rng(seed)
    sc = Scene()
    if kind == "compact":
        sc.add_parent("src", rng, W, H, masked=False)
        sc.add_window("s", "src", 0, 0, W, H)
        for o in range(k.nout):
            sc.add_parent(f"dst{o}", rng, W, H, masked=False)
            sc.add_window(f"d{o}", f"dst{o}", 0, 0, W, H)
    elif kind == "masked":
        sc.add_parent("src", rng, W, H, masked=True)
        sc.add_window("s", "src", 0, 0, W, H)
        for o in range(k.nout):
            sc.add_parent(f"dst{o}", rng, W, H, masked=True)
            sc.add_window(f"d{o}", f"dst{o}", 0, 0, W, H)
    elif kind == "windows":
        # A source window in a wider parent at a mid-word bit offset, and
        # output windows in parents with other strides and bit bases.
        sw, sh = W + 19, H + 11
        sc.add_parent("src", rng, sw, sh, masked=True, valid_base=37)
        sc.add_window("s", "src", 7, 5, W, H)
        for o in range(k.nout):
            dw, dh = W + 5 + 64 * o, H + 3
            sc.add_parent(f"dst{o}", rng, dw, dh, masked=True, valid_base=11 + 40 * o)
            sc.add_window(f"d{o}", f"dst{o}", 3, 2, W, H)
    elif kind == "unmasked-in":
        sc.add_parent("src", rng, W + 3, H, masked=False)
        sc.add_window("s", "src", 1, 0, W, H)
        for o in range(k.nout):
            sc.add_parent(f"dst{o}", rng, W + 1, H + 1, masked=True, valid_base=63)
            sc.add_window(f"d{o}", f"dst{o}", 1, 1, W, H)
    else:
        raise ValueError(kind)
    return sc


def run_case(gpu, k, sc: Scene, opts=None, ctx=None):
    host, hbufs = sc.materialize(None)
    dst_h = [host[f"d{o}"] for o in range(k.nout)]
    oracle.process(k, dst_h, [host["s"]])
    dev, dbufs = sc.materialize(gpu)
    dst_d = [dev[f"d{o}"] for o in range(k.nout)]
    err = ProcessN(ctx, dst_d, [dev["s"]], k, opts or Options())
    assert err is None
    import torch
    torch.cuda.synchronize()
    return host_copy(dbufs), hbufs


@pytest.mark.parametrize("name,k", ALL, ids=[n + f"-r{k.radius}" + ("fit" if k.flags & 0x100 else "")
                                              for n, k in ALL])
@pytest.mark.parametrize("kind", ["compact", "masked", "windows", "unmasked-in"])
def test_parity(gpu, name, k, kind):
    W, H = (300, 70) if k.radius <= 2 else (150, 75)
    sc = build_scene(kind, k, hash((name, kind)) & 0xFFFF, W, H)
    got, want = run_case(gpu, k, sc)
    assert_same_parents(got, want, f"{name} {kind}")


SIZES = [(1, 1), (1, 5), (5, 1), (2, 2), (3, 3), (4, 3), (17, 17), (63, 9), (64, 4), (65, 6), (127, 3),
         (129, 5), (255, 67), (256, 64), (257, 65), (513, 129)]


@pytest.mark.parametrize("W,H", SIZES)
@pytest.mark.parametrize("name", ["slope0", "curv2", "gradient", "mean3", "correlate2", "lookup", "gauss8"])
def test_sizes(gpu, W, H, name):
    k = dict((n, kk) for n, kk in ALL if not (kk.flags & 0x100))[name] if name != "gradient" else \
        terrain.gradientOp(terrain.GradientOptions(CellSize=1))
    for kind in ("masked", "windows"):
        sc = build_scene(kind, k, W * 1000 + H, W, H)
        got, want = run_case(gpu, k, sc)
        assert_same_parents(got, want, f"{name} {W}x{H} {kind}")


@pytest.mark.parametrize("tw,th", [(0, 0), (1, 0), (7, 5), (64, 1), (256, 256), (1000, 3)])
def test_tiled_with_cancel_flag_matches(gpu, tw, th):
    """Tiled forms with a live context run unit by unit; bits are unchanged."""
    for name in ("slope0", "max2", "reclass"):
        k = dict(ALL)[name]
        sc = build_scene("windows", k, 99, 301, 97)
        st = engine.Stats()
        got, want = run_case(gpu, k, sc, Options(TileWidth=tw, TileHeight=th, Stats=st), Context())
        assert_same_parents(got, want, f"{name} tiled {tw}x{th}")
        assert st.Cells == 301 * 97


def test_cancelled_before_start_writes_nothing(gpu):
    k = dict(ALL)["slope0"]
    sc = build_scene("masked", k, 5, 200, 50)
    dev, dbufs = sc.materialize(gpu)
    before = host_copy(dbufs)
    ctx = Context()
    ctx.cancel()
    st = engine.Stats()
    err = ProcessN(ctx, [dev["d0"]], [dev["s"]], k, Options(TileHeight=8, Stats=st))
    assert isinstance(err, engine.Canceled)
    assert_same_parents(host_copy(dbufs), before, "cancelled")
    assert st.Cells == 0 and st.Bands == 0


class CountingSource(MemorySource):
    """A host memory source that cancels its context after n reads."""

    def __init__(self, r, ctx, n):
        super().__init__(r)
        self.ctx, self.n, self.reads = ctx, n, 0

    def ReadWindow(self, ctx, dst, x, y):
        self.reads += 1
        if self.reads == self.n:
            self.ctx.cancel()
        return super().ReadWindow(ctx, dst, x, y)


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("name", ["slope0", "hillshade", "gradient", "gauss3", "reclass", "curv1"])
def test_chunked_matches(gpu, name, resident):
    k = dict(ALL)[name]
    for tw, th in ((0, 16), (50, 7), (0, 0), (333, 1)):
        sc = build_scene("windows", k, 3, 211, 83)
        host, hbufs = sc.materialize(None)
        oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
        if resident:
            dev, dbufs = sc.materialize(gpu)
            src = MemorySource(dev["s"])
            sinks = [MemorySink(dev[f"d{o}"]) for o in range(k.nout)]
        else:
            dev, dbufs = sc.materialize(None)
            src = MemorySource(dev["s"])
            sinks = [MemorySink(dev[f"d{o}"]) for o in range(k.nout)]
        st = engine.Stats()
        err = ProcessChunked(Context(), sinks, [src], k, Options(TileWidth=tw, TileHeight=th, Stats=st))
        assert err is None
        assert_same_parents(host_copy(dbufs), hbufs, f"chunked {name} {tw}x{th} resident={resident}")
        want = engine.PlanStats(k, 211, 83, Options(TileWidth=tw, TileHeight=th), chunked=True)
        assert st == want


def test_chunked_cancel_leaves_tile_prefix(gpu):
    k = dict(ALL)["slope0"]
    W, H, th = 120, 90, 10
    sc = build_scene("masked", k, 11, W, H)
    host, hbufs = sc.materialize(None)
    oracle.process(k, [host["d0"]], [host["s"]])
    dev, dbufs = sc.materialize(None)
    before = host_copy(dbufs)
    ctx = Context()
    src = CountingSource(dev["s"], ctx, 4)
    st = engine.Stats()
    err = ProcessChunked(ctx, [MemorySink(dev["d0"])], [src], k, Options(TileHeight=th, Stats=st))
    assert isinstance(err, engine.Canceled)
    assert st.Tiles == 4
    got = host_copy(dbufs)["dst0"].data.reshape(H, W)
    want = hbufs["dst0"].data.reshape(H, W)
    old = before["dst0"].data.reshape(H, W)
    from helpers import same_floats
    assert same_floats(got[:4 * th], want[:4 * th]).all()
    assert same_floats(got[4 * th:], old[4 * th:]).all()


def test_host_memory_operands(gpu):
    """Host rasters go through the same kernels, staged by the library."""
    for name in ("aspect", "correlate3", "lookup", "gradient"):
        k = dict(ALL)[name]
        sc = build_scene("windows", k, 21, 97, 41)
        host, hbufs = sc.materialize(None)
        oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
        h2, b2 = sc.materialize(None)
        assert ProcessN(None, [h2[f"d{o}"] for o in range(k.nout)], [h2["s"]], k) is None
        assert_same_parents(host_copy(b2), hbufs, f"host operands {name}")


def test_in_place_pointwise(gpu):
    """A radius-0 kernel may write its own input (internal/exec/doc.go)."""
    import torch
    k = dict(ALL)["lookup"]
    rng = np.random.default_rng(3)
    sc = Scene()
    sc.add_parent("p", rng, 70, 30, masked=True, valid_base=5,
                  values=rng.uniform(-5, 45, 70 * 30).astype(np.float32))
    sc.add_window("w", "p", 2, 1, 60, 25)
    host, hbufs = sc.materialize(None)
    ref = host["w"]
    from strata_amd.raster import Float32Raster
    out = Float32Raster(np.zeros(60 * 25, np.float32), 60, 25, 60, np.zeros(30, np.uint64), 0)
    oracle.process(k, [out], [ref])
    dev, dbufs = sc.materialize(gpu)
    assert ProcessN(None, [dev["w"]], [dev["w"]], k) is None
    torch.cuda.synchronize()
    got = host_copy(dbufs)["p"].data.reshape(30, 70)[1:26, 2:62]
    from helpers import same_floats
    assert same_floats(got, out.Data.reshape(25, 60)).all()


def test_known_answer_plane_on_gpu(gpu):
    """terrain/terrain_test.go:57-107 on the device: Horn is exact on a plane."""
    import math
    from strata_amd.raster import NewFloat32, NewFloat32Like, data_grid
    d = plane_values(9, 7, 0.3, -0.4, 10.0, 5.0)
    dem = NewFloat32(9, 7, d).to(gpu)
    for units, want, tol in ((terrain.SlopeDegrees, math.degrees(math.atan(0.5)), 1e-3),
                             (terrain.SlopeRadians, math.atan(0.5), 2e-5), (terrain.SlopePercent, 50.0, 1e-3)):
        dst = NewFloat32Like(dem)
        terrain.Slope(dst, dem, terrain.SlopeOptions(CellSize=10, CellSizeY=5, Units=units))
        g = data_grid(dst)
        assert np.isnan(g[0]).all() and np.isnan(g[-1]).all() and np.isnan(g[:, 0]).all()
        assert np.abs(g[1:-1, 1:-1] - want).max() <= tol


@pytest.mark.parametrize("sx,sw_extra,dx,dw_extra", [(1, 0, 0, 0), (0, 1, 1, 0), (1, 1, 1, 1), (3, 2, 2, 3),
                                                     (0, 0, 1, 1)])
def test_unaligned_blocks_every_parity(gpu, sx, sw_extra, dx, dw_extra):
    """Whole 512-column blocks over operands that allow no pair accesses at
    even offsets: the source and output windows start at odd or even
    columns of parents with odd or even strides, so a row's parity can
    alternate from one row to the next. Every radius-1 terrain op, against
    the oracle, whole parent buffers compared."""
    W, H = 1100, 40
    ks = [(n, k) for n, k in ALL if k.radius == 1 and not (k.flags & 0x100) and n.split("-")[0].rstrip("0123456789")
          in ("gradient", "slope", "aspect", "hillshade", "curv")]
    assert len(ks) >= 6
    for name, k in ks:
        rng = np.random.default_rng(hash((name, sx, dx)) & 0xFFFF)
        sc = Scene()
        sc.add_parent("src", rng, W + 8 + sw_extra, H + 2, masked=True, valid_base=5)
        sc.add_window("s", "src", sx, 1, W, H)
        for o in range(k.nout):
            sc.add_parent(f"dst{o}", rng, W + 4 + dw_extra, H + 1, masked=True, valid_base=17 + 8 * o)
            sc.add_window(f"d{o}", f"dst{o}", dx, 1, W, H)
        got, want = run_case(gpu, k, sc)
        assert_same_parents(got, want, f"{name} sx {sx} dx {dx} strides +{sw_extra}/+{dw_extra}")
