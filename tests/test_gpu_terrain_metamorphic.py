"""Metamorphic relations of the radius-1 terrain ops on the device, after
the reference's terrain/metamorphic_test.go:17-60. Horn's and
Zevenbergen–Thorne's sums have a fixed order, so these hold bit for bit
(a negated gradient as a value: a zero sum is +0 in either order):

- a mirror of the columns negates dx and keeps dy; Slope and the three
  curvatures are unchanged, and so is Hillshade with the azimuth negated
  (sin is odd); a mirror of the rows negates dy and keeps the curvatures;
- a transposition with the cell sizes swapped swaps dx and dy and keeps
  Slope;
- adding an integer constant to a small integer DEM changes no result;
- scaling the DEM by a power of two and ZFactor by its inverse changes no
  result;
- negating the DEM and ZFactor changes no result.

Each side runs through its own path (plain, Tiled or Chunked over memory
sources and sinks) on a DEM of a few million cells; inputs are made and
results compared on the host. Scaling and negation also run over the
20000² DEM, compared on the device.
"""
import numpy as np
import pytest
import torch

from strata_amd import terrain as T
from strata_amd.engine import MemorySink, MemorySource, Options, ProcessChunked, ProcessN
from strata_amd.raster import Float32Raster
from strata_amd.synth import pack_mask

pytestmark = pytest.mark.gpu

W, H = 2003, 1501
CURV = (T.CurvatureProfile, T.CurvaturePlan, T.CurvatureMean)


def dem_values(rng, w, h, kind):
    if kind == "int":
        return rng.integers(-1000, 1001, (h, w)).astype(np.float32)
    x = np.arange(w, dtype=np.float64)[None, :]
    y = np.arange(h, dtype=np.float64)[:, None]
    v = (800 + 120 * np.sin(x / 37) * np.cos(y / 23) + rng.standard_normal((h, w)) * 15).astype(np.float32)
    if kind == "hazards":
        hz = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 3e38, -3e38, 1e-40], np.float32)
        pick = rng.random((h, w)) < 0.01
        v[pick] = rng.choice(hz, int(pick.sum()))
    return v


def run(k, nout, data, valid, path):
    """k over (data, valid) on the device through path 0 (plain), 1 (Tiled)
    or 2 (Chunked); returns [(data, valid)] per output, on the host."""
    h, w = data.shape
    n = w * h
    dev = torch.device("cuda:0")
    src = Float32Raster(torch.from_numpy(np.ascontiguousarray(data).reshape(-1)).to(dev), w, h, w,
                        pack_mask(torch.from_numpy(np.ascontiguousarray(valid).reshape(-1)).to(dev)), 0)
    outs = [Float32Raster(torch.full((n,), 7.0, dtype=torch.float32, device=dev), w, h, w,
                          pack_mask(torch.zeros(n, dtype=torch.bool, device=dev)), 0) for _ in range(nout)]
    if path == 0:
        assert ProcessN(None, outs, [src], k, Options()) is None
    elif path == 1:
        assert ProcessN(None, outs, [src], k, Options(TileWidth=700, TileHeight=333, Workers=2)) is None
    else:
        assert ProcessChunked(None, [MemorySink(o) for o in outs], [MemorySource(src)], k,
                              Options(TileHeight=190)) is None
    torch.cuda.synchronize()
    res = []
    for o in outs:
        bits = np.unpackbits(o.Valid.cpu().numpy().view(np.uint8), bitorder="little")[:n].astype(bool)
        res.append((o.Data.cpu().numpy().reshape(h, w), bits.reshape(h, w)))
    return res


def require(got, want, f=lambda v: v, same_value=False):
    """requireSameCell (terrain/metamorphic_test.go:436-458): the same
    validity, and on valid cells the same bits of f(want), or with
    same_value the same value (+0 matching -0, as where a zero sum is +0
    in either order, so negating a result's inputs does not negate it);
    NaN matches NaN."""
    (gd, gv), (wd, wv) = got, want
    assert (gv == wv).all(), "validity differs"
    g, fw = gd[wv], f(wd[wv]).astype(np.float32)
    ok = (g == fw) if same_value else (g.view(np.uint32) == fw.view(np.uint32))
    bad = ~(ok | (np.isnan(g) & np.isnan(fw)))
    assert not bad.any(), f"{int(bad.sum())} cells differ, first {g[bad][:4]} want {fw[bad][:4]}"


def source(seed, kind, w=W, h=H):
    rng = np.random.default_rng(seed)
    return dem_values(rng, w, h, kind), rng.random((h, w)) >= 0.01


def ops(cs=30.0, csy=20.0, z=1.5, az=117.5):
    """(name, kernel, outputs) for every radius-1 op at these settings."""
    out = [("gradient", T.gradientOp(T.GradientOptions(CellSize=cs, CellSizeY=csy, ZFactor=z)), 2),
           ("slope", T.slopeOp(T.SlopeOptions(CellSize=cs, CellSizeY=csy, ZFactor=z)), 1),
           ("aspect", T.aspectOp(T.AspectOptions(CellSize=cs, CellSizeY=csy, ZFactor=z)), 1),
           ("hillshade", T.hillshadeOp(T.HillshadeOptions(CellSize=cs, CellSizeY=csy, ZFactor=z, Azimuth=az,
                                                          Altitude=35)), 1)]
    for ct in CURV:
        out.append((f"curvature{int(ct)}", T.curvatureOp(T.CurvatureOptions(CellSize=cs, CellSizeY=csy, ZFactor=z,
                                                                           Type=ct)), 1))
    return out


def test_mirror_columns():
    data, valid = source(1, "hazards")
    m = lambda a: np.ascontiguousarray(a[:, ::-1])
    md, mv = m(data), m(valid)
    for i, (name, k, nout) in enumerate(ops()):
        if name == "aspect":
            continue
        want = run(k, nout, data, valid, i % 3)
        kk = T.hillshadeOp(T.HillshadeOptions(CellSize=30.0, CellSizeY=20.0, ZFactor=1.5, Azimuth=-117.5,
                                              Altitude=35)) if name == "hillshade" else k
        got = run(kk, nout, md, mv, (i + 1) % 3)
        if name == "gradient":
            require(got[0], (m(want[0][0]), m(want[0][1])), lambda v: -v, same_value=True)
            require(got[1], (m(want[1][0]), m(want[1][1])))
        else:
            require(got[0], (m(want[0][0]), m(want[0][1])))


def test_mirror_rows():
    data, valid = source(2, "hazards")
    m = lambda a: np.ascontiguousarray(a[::-1, :])
    md, mv = m(data), m(valid)
    for i, (name, k, nout) in enumerate(ops()):
        if name not in ("gradient", "slope") and not name.startswith("curvature"):
            continue
        want = run(k, nout, data, valid, (i + 2) % 3)
        got = run(k, nout, md, mv, i % 3)
        if name == "gradient":
            require(got[0], (m(want[0][0]), m(want[0][1])))
            require(got[1], (m(want[1][0]), m(want[1][1])), lambda v: -v, same_value=True)
        else:
            require(got[0], (m(want[0][0]), m(want[0][1])))


def test_transpose_with_swapped_cell_sizes():
    data, valid = source(3, "hazards")
    t = lambda a: np.ascontiguousarray(a.T)
    for path in range(3):
        g = run(T.gradientOp(T.GradientOptions(CellSize=30.0, CellSizeY=20.0, ZFactor=1.5)), 2, data, valid, path)
        gt = run(T.gradientOp(T.GradientOptions(CellSize=20.0, CellSizeY=30.0, ZFactor=1.5)), 2, t(data), t(valid),
                 (path + 1) % 3)
        require(gt[0], (t(g[1][0]), t(g[1][1])))
        require(gt[1], (t(g[0][0]), t(g[0][1])))
        s = run(T.slopeOp(T.SlopeOptions(CellSize=30.0, CellSizeY=20.0, ZFactor=1.5)), 1, data, valid, path)
        st = run(T.slopeOp(T.SlopeOptions(CellSize=20.0, CellSizeY=30.0, ZFactor=1.5)), 1, t(data), t(valid),
                 (path + 2) % 3)
        require(st[0], (t(s[0][0]), t(s[0][1])))


@pytest.mark.parametrize("c", [-977, 4096])
def test_integer_shift(c):
    data, valid = source(4, "int")
    shifted = data + np.float32(c)
    for i, (name, k, nout) in enumerate(ops()):
        want = run(k, nout, data, valid, i % 3)
        got = run(k, nout, shifted, valid, (i + 1) % 3)
        for o in range(nout):
            require(got[o], want[o])


@pytest.mark.parametrize("e", [-5, 3])
def test_power_of_two_with_zfactor(e):
    data, valid = source(5, "smooth")
    f = np.float32(2.0 ** e)
    scaled = data * f
    base, inv = ops(z=1.5), ops(z=1.5 / float(f))
    for i, ((name, k, nout), (_, kk, _)) in enumerate(zip(base, inv)):
        want = run(k, nout, data, valid, i % 3)
        got = run(kk, nout, scaled, valid, (i + 2) % 3)
        for o in range(nout):
            require(got[o], want[o])


def test_negation_is_negative_zfactor():
    data, valid = source(6, "hazards")
    base, neg = ops(z=1.5), ops(z=-1.5)
    for i, ((name, k, nout), (_, kk, _)) in enumerate(zip(base, neg)):
        want = run(k, nout, data, valid, i % 3)
        got = run(kk, nout, -data, valid, (i + 1) % 3)
        for o in range(nout):
            require(got[o], want[o], same_value=name == "gradient")


def test_scaling_and_negation_at_20000():
    """Power-of-two scaling against ZFactor and negation against a negative
    ZFactor, for Slope, Aspect, Hillshade and the three curvatures over the
    20000² fBm DEM with 0.5% NoData blobs (the terrain-suite workload),
    compared bit for bit on the device. Every op sees the elevations only
    through differences scaled by constants linear in ZFactor, so scaling
    the DEM by 2^k and ZFactor by 2^-k (or negating both) leaves each
    derivative, and so each output, unchanged."""
    from strata_amd import synth

    N = 20000
    dev = torch.device("cuda:0")
    data = synth.fbm_dem(N, N, hurst=0.8, seed=1, lo=0.0, hi=3000.0, device=dev)
    valid, _ = synth.blob_nodata(N, N, frac=0.005, rmin=5, rmax=50, seed=2, device=dev)

    def go(k, d):
        out = Float32Raster(torch.empty(N * N, dtype=torch.float32, device=dev), N, N, N, torch.zeros_like(valid), 0)
        assert ProcessChunked(None, [MemorySink(out)], [MemorySource(Float32Raster(d, N, N, N, valid, 0))], k,
                              Options(TileHeight=256)) is None
        return out

    bits = ((valid.view(-1, 1) >> torch.arange(64, device=dev)) & 1).reshape(-1)[:N * N].bool()
    curv = lambda t: lambda z: T.curvatureOp(T.CurvatureOptions(CellSize=30, ZFactor=z, Type=t))
    for name, mk in (("slope", lambda z: T.slopeOp(T.SlopeOptions(CellSize=30, ZFactor=z))),
                     ("aspect", lambda z: T.aspectOp(T.AspectOptions(CellSize=30, ZFactor=z))),
                     ("hillshade", lambda z: T.hillshadeOp(T.HillshadeOptions(CellSize=30, ZFactor=z))),
                     ("curvature-profile", curv(T.CurvatureProfile)),
                     ("curvature-plan", curv(T.CurvaturePlan)),
                     ("curvature-mean", curv(T.CurvatureMean))):
        base = go(mk(1.0), data)
        ok = base.Valid
        want = base.Data[bits]
        for d, z in ((data * 0.125, 8.0), (data * 4.0, 0.25), (-data, -1.0)):
            got = go(mk(z), d)
            assert torch.equal(got.Valid, ok), (name, z)
            assert torch.equal(got.Data[bits].view(torch.int32), want.view(torch.int32)), (name, z)
            del got
        del base, want
        torch.cuda.empty_cache()
