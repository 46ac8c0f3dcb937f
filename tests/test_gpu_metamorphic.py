"""Metamorphic relations of the focal ops on the device, after the
reference's focalRelations (focal/metamorphic_test.go:66-215), which its
fuzz corpus (focal/testdata/fuzz/FuzzFocalRelations) drives:

- crop: a window's interior outputs are the full raster's;
- locality: changing one input cell changes only outputs within r of it,
  and invalidating one cell invalidates exactly the valid outputs within r;
- data under invalid cells changes nothing;
- scaling the input or the weights by a power of two scales the result
  exactly (smooth data, weights of a few bits: no overflow, no subnormals);
- negation: Correlate, Convolve, CorrelateSeparable and Mean of -x are the
  negated results up to the sign of a zero, and Min(-x) = -Max(x) exactly;
- Min(x + c) = Min(x) + c and likewise Max;
- on exact data (small integers and integer weights), every op commutes
  with the eight symmetries of the grid (weights mapped too), the separable
  form is Correlate with the outer product of its taps (up to the sign of a
  zero) and Mean is Correlate with unit weights divided by (2r+1)².

These hold at any size, so they run here on rasters far larger than the
oracle's, each side through its own path: plain, Tiled or Chunked over
memory sources and sinks. Inputs are made and results compared on the host;
one test runs two relations at 16384² entirely on the device.
"""
import numpy as np
import pytest
import torch

from strata_amd import focal
from strata_amd.engine import MemorySink, MemorySource, Options, ProcessChunked, ProcessN
from strata_amd.raster import Float32Raster
from strata_amd.synth import pack_mask

pytestmark = pytest.mark.gpu

KINDS = ("correlate", "convolve", "separable", "mean", "min", "max")
W, H = 1531, 1207  # odd sizes: every tile and unit shape meets a ragged edge
SYM_W, SYM_H = 701, 433


# -- operands and paths ------------------------------------------------------

def make_spec(rng, kind, r, weights):
    """newSpec / relSpec (focal/helpers_test.go:59-84,
    focal/metamorphic_test.go:47-57)."""
    k = 2 * r + 1

    def draw(n):
        if weights == "int":
            return rng.integers(-3, 4, n).astype(np.float32)
        if weights == "moderate":
            return (rng.integers(-1000, 1001, n) / 64).astype(np.float32)
        v = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32).view(np.float32).copy()
        bad = ~np.isfinite(v)
        v[bad] = (rng.integers(-100, 101, int(bad.sum())) / 8).astype(np.float32)
        return v

    s = {"kind": kind, "r": r}
    if kind in ("correlate", "convolve"):
        s["w"] = draw(k * k)
    elif kind == "separable":
        s["row"], s["col"] = draw(k), draw(k)
    return s


def kernel(s):
    r = s["r"]
    if s["kind"] in ("correlate", "convolve"):
        return focal.newWeightsKernel(focal.WeightsOptions(r, list(map(float, s["w"]))), s["kind"] == "convolve")
    if s["kind"] == "separable":
        return focal.newSeparableKernel(focal.SeparableOptions(r, list(map(float, s["row"])),
                                                               list(map(float, s["col"]))))
    return focal.newBox(focal.BoxOptions(r), s["kind"])


def values(rng, w, h, kind):
    """newSrc's cell values (focal/helpers_test.go:299-315)."""
    n = w * h
    if kind == "any":
        v = (rng.integers(-4000, 4001, n).astype(np.float32) / np.float32(7)).astype(np.float32)
        pick = rng.integers(0, 3, n) == 0
        v[pick] = rng.integers(0, 2 ** 32, int(pick.sum()), dtype=np.uint64).astype(np.uint32).view(np.float32)
        return v.reshape(h, w)
    if kind == "smooth":
        x = np.arange(w, dtype=np.float64)[None, :]
        y = np.arange(h, dtype=np.float64)[:, None]
        noise = rng.integers(-100, 101, (h, w)) / 10
        return (500 + 40 * np.sin(x / 5) + 30 * np.cos(y / 3) + noise).astype(np.float32)
    return rng.integers(-50, 51, (h, w)).astype(np.float32)


def source(rng, w, h, kind, masked):
    data = values(rng, w, h, kind)
    valid = rng.integers(0, 10, (h, w)) != 0 if masked else None
    return data, valid


def run(s, data, valid, path):
    """Run s over (data, valid) on the device through path 0 (plain), 1
    (Tiled) or 2 (Chunked over memory sources and sinks); the result comes
    back to the host as (data, valid)."""
    h, w = data.shape
    n = w * h
    dev = torch.device("cuda:0")
    src = Float32Raster(torch.from_numpy(np.ascontiguousarray(data).reshape(-1)).to(dev), w, h, w,
                        pack_mask(torch.from_numpy(valid.reshape(-1)).to(dev)) if valid is not None else None, 0)
    out = Float32Raster(torch.full((n,), 12345.0, dtype=torch.float32, device=dev), w, h, w,
                        pack_mask(torch.zeros(n, dtype=torch.bool, device=dev)) if valid is not None else None, 0)
    k = kernel(s)
    if path == 0:
        assert ProcessN(None, [out], [src], k, Options()) is None
    elif path == 1:
        assert ProcessN(None, [out], [src], k, Options(TileWidth=333, TileHeight=97, Workers=3)) is None
    else:
        assert ProcessChunked(None, [MemorySink(out)], [MemorySource(src)], k, Options(TileHeight=130)) is None
    torch.cuda.synchronize()
    got = out.Data.cpu().numpy().reshape(h, w)
    if valid is None:
        return got, None
    words = out.Valid.cpu().numpy().view(np.uint64)
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(bool)
    return got, bits.reshape(h, w)


def require(got, want, f=lambda v: v, same_value=False, where=None):
    """requireCell over cells (focal/metamorphic_test.go:253-277): the same
    validity, and on valid cells the same bits (or the same value, so ±0
    match) of f(want); NaN matches NaN."""
    gd, gv = got
    wd, wv = want
    if where is None:
        where = np.ones(wd.shape, bool)
    if wv is not None:
        assert (gv[where] == wv[where]).all(), "validity differs"
        where = where & wv
    g = gd[where]
    fw = f(wd[where]).astype(np.float32)
    both_nan = np.isnan(g) & np.isnan(fw)
    ok = (g == fw) if same_value else (g.view(np.uint32) == fw.view(np.uint32))
    bad = ~(ok | both_nan)
    assert not bad.any(), f"{int(bad.sum())} cells differ, first {g[bad][:4]} want {fw[bad][:4]}"


def near(h, w, px, py, r):
    m = np.zeros((h, w), bool)
    m[max(0, py - r):py + r + 1, max(0, px - r):px + r + 1] = True
    return m


CASES = [(kind, r, seed) for seed, (kind, r) in enumerate((kd, r) for kd in KINDS for r in (1, 3))]


# -- the relations ----------------------------------------------------------

@pytest.mark.parametrize("kind,r,seed", CASES)
def test_crop(kind, r, seed):
    rng = np.random.default_rng(100 + seed)
    s = make_spec(rng, kind, r, "float")
    data, valid = source(rng, W, H, "any", bool(seed % 3))
    full = run(s, data, valid, seed % 3)
    x0, y0 = W // 7, H // 12
    cw, ch = 2 * W // 3, 2 * H // 3
    part = run(s, data[y0:y0 + ch, x0:x0 + cw], None if valid is None else valid[y0:y0 + ch, x0:x0 + cw],
               (seed + 1) % 3)
    inner = (slice(r, ch - r), slice(r, cw - r))
    want = (full[0][y0:y0 + ch, x0:x0 + cw][inner], None if valid is None else full[1][y0:y0 + ch, x0:x0 + cw][inner])
    require((part[0][inner], None if valid is None else part[1][inner]), want)


@pytest.mark.parametrize("kind,r,seed", CASES)
def test_one_changed_cell(kind, r, seed):
    rng = np.random.default_rng(200 + seed)
    s = make_spec(rng, kind, r, "float")
    data, valid = source(rng, W, H, "any", bool(seed % 2))
    px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
    changed = data.copy()
    changed[py, px] = np.float32(rng.standard_normal() * 1e3)
    a, b = run(s, data, valid, seed % 3), run(s, changed, valid, (seed + 2) % 3)
    require(b, a, where=~near(H, W, px, py, r))


@pytest.mark.parametrize("kind,r,seed", CASES)
def test_one_more_invalid_cell(kind, r, seed):
    rng = np.random.default_rng(300 + seed)
    s = make_spec(rng, kind, r, "float")
    data, valid = source(rng, W, H, "any", True)
    px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
    less = valid.copy()
    less[py, px] = False
    a, b = run(s, data, valid, seed % 3), run(s, data, less, (seed + 1) % 3)
    nb = near(H, W, px, py, r)
    assert (b[1] == (a[1] & ~nb)).all()
    require(b, a, where=a[1] & ~nb)


@pytest.mark.parametrize("kind,r,seed", CASES)
def test_data_under_invalid_cells(kind, r, seed):
    rng = np.random.default_rng(400 + seed)
    s = make_spec(rng, kind, r, "float")
    data, valid = source(rng, W, H, "any", True)
    scrambled = data.copy()
    bad = ~valid
    scrambled[bad] = rng.integers(0, 2 ** 32, int(bad.sum()), dtype=np.uint64).astype(np.uint32).view(np.float32)
    require(run(s, scrambled, valid, seed % 3), run(s, data, valid, (seed + 2) % 3))


@pytest.mark.parametrize("kind,r,seed", CASES)
def test_power_of_two_scaling(kind, r, seed):
    rng = np.random.default_rng(500 + seed)
    s = make_spec(rng, kind, r, "moderate")
    data, valid = source(rng, W, H, "smooth", bool(seed % 2))
    e = int(rng.integers(-6, 7))
    f = np.float32(2.0 ** e)
    want = run(s, data, valid, seed % 3)
    if seed % 2 or kind in ("mean", "min", "max"):
        got = run(s, (data * f).astype(np.float32), valid, (seed + 1) % 3)
    else:
        ss = dict(s)
        if "w" in s:
            ss["w"] = (s["w"] * f).astype(np.float32)
        else:
            ss["row"] = (s["row"] * f).astype(np.float32)
        got = run(ss, data, valid, (seed + 1) % 3)
    require(got, want, lambda v: v * f)


@pytest.mark.parametrize("kind,r,seed", CASES)
def test_negation(kind, r, seed):
    rng = np.random.default_rng(600 + seed)
    s = make_spec(rng, kind, r, "float")
    data, valid = source(rng, W, H, "any", bool(seed % 2))
    neg = -data
    got = run(s, neg, valid, seed % 3)
    if kind in ("min", "max"):
        other = dict(s, kind="max" if kind == "min" else "min")
        require(got, run(other, data, valid, (seed + 1) % 3), lambda v: -v)
        return
    want = run(s, data, valid, (seed + 1) % 3)
    require(got, want, lambda v: -v, same_value=True)
    if kind != "mean":
        ns = dict(s)
        for key in ("w", "row"):
            if key in s:
                ns[key] = -s[key]
        require(run(ns, data, valid, (seed + 2) % 3), want, lambda v: -v, same_value=True)


@pytest.mark.filterwarnings("ignore::RuntimeWarning")  # Inf + (-Inf) cells of the expected side
@pytest.mark.parametrize("kind,r,seed", [(k, r, i) for i, (k, r) in enumerate(
    [("min", 1), ("max", 2), ("min", 4), ("max", 3)])])
def test_shifted_min_max(kind, r, seed):
    rng = np.random.default_rng(700 + seed)
    data, valid = source(rng, W, H, "any", bool(seed % 2))
    c = np.float32(int(rng.integers(-1000, 1001)) / 8)
    s = {"kind": kind, "r": r}
    require(run(s, (data + c).astype(np.float32), valid, seed % 3), run(s, data, valid, (seed + 1) % 3),
            lambda v: (v + c).astype(np.float32))


def dihedral(a, swap, flipx, flipy):
    """The result of a symmetry: transpose first, then mirror the result's
    axes (rastertest's Dihedral, as focal/metamorphic_test.go:224-251 uses it)."""
    if swap:
        a = a.T
    if flipx:
        a = a[:, ::-1]
    if flipy:
        a = a[::-1, :]
    return np.ascontiguousarray(a)


def transform_spec(s, swap, flipx, flipy):
    """transformSpec (focal/metamorphic_test.go:224-251)."""
    out = dict(s)
    k = 2 * s["r"] + 1
    if "w" in s:
        out["w"] = dihedral(s["w"].reshape(k, k), swap, flipx, flipy).reshape(-1)
    if "row" in s:
        row, col = (s["col"], s["row"]) if swap else (s["row"], s["col"])
        out["row"] = row[::-1].copy() if flipx else row
        out["col"] = col[::-1].copy() if flipy else col
    return out


SYMS = [(sw, fx, fy) for sw in (False, True) for fx in (False, True) for fy in (False, True)]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("r", [1, 2])
def test_symmetries_of_the_grid(kind, r):
    rng = np.random.default_rng(800 + KINDS.index(kind) * 10 + r)
    s = make_spec(rng, kind, r, "int")
    data, valid = source(rng, SYM_W, SYM_H, "int", True)
    want = run(s, data, valid, 0)
    for i, (sw, fx, fy) in enumerate(SYMS):
        got = run(transform_spec(s, sw, fx, fy), dihedral(data, sw, fx, fy), dihedral(valid, sw, fx, fy), i % 3)
        require(got, (dihedral(want[0], sw, fx, fy), dihedral(want[1], sw, fx, fy)),
                same_value=kind == "separable")


@pytest.mark.parametrize("r", [1, 2, 3])
def test_separable_is_the_outer_product(r):
    rng = np.random.default_rng(900 + r)
    s = make_spec(rng, "separable", r, "int")
    data, valid = source(rng, W, H, "int", True)
    outer = {"kind": "correlate", "r": r, "w": np.outer(s["col"], s["row"]).astype(np.float32).reshape(-1)}
    require(run(s, data, valid, r % 3), run(outer, data, valid, (r + 1) % 3), same_value=True)


@pytest.mark.parametrize("r", [1, 3, 5])
def test_mean_is_unit_correlate_over_n(r):
    rng = np.random.default_rng(950 + r)
    data, valid = source(rng, W, H, "int", bool(r % 2))
    k = 2 * r + 1
    ones = {"kind": "correlate", "r": r, "w": np.ones(k * k, np.float32)}
    n = np.float32(k * k)
    require(run({"kind": "mean", "r": r}, data, valid, r % 3), run(ones, data, valid, (r + 2) % 3),
            lambda v: (v / n).astype(np.float32))


# -- the C3 size, on the device ----------------------------------------------

def test_scaling_and_negation_at_16384():
    """Power-of-two scaling and negation of Mean r=3 and Correlate r=2 over
    the 16384² focal workload, compared bit for bit on the device."""
    S = 16384
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(16384)
    x = torch.arange(S, device=dev, dtype=torch.float64)
    data = (500 + 40 * torch.sin(x / 5)[None, :] + 30 * torch.cos(x / 3)[:, None]
            + torch.randint(-100, 101, (S, S), generator=gen, device=dev) / 10).to(torch.float32).reshape(-1)
    valid = pack_mask(torch.rand(S * S, generator=gen, device=dev) >= 0.02)
    w5 = (torch.randint(-1000, 1001, (25,), generator=gen, device=dev) / 64).to(torch.float32).cpu().tolist()
    ops = {"mean3": focal.newBox(focal.BoxOptions(3), "mean"),
           "correlate2": focal.newWeightsKernel(focal.WeightsOptions(2, w5), False)}

    def go(k, d):
        src = Float32Raster(d, S, S, S, valid, 0)
        out = Float32Raster(torch.empty(S * S, dtype=torch.float32, device=dev), S, S, S, torch.zeros_like(valid), 0)
        assert ProcessN(None, [out], [src], k, Options()) is None
        return out

    for name, k in ops.items():
        base = go(k, data)
        ok = base.Valid
        bits = ((ok.view(-1, 1) >> torch.arange(64, device=dev)) & 1).reshape(-1)[:S * S].bool()
        for f in (0.25, 8.0):
            got = go(k, data * f)
            assert torch.equal(got.Valid, ok), name
            a, b = got.Data[bits], base.Data[bits] * f
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (name, f)
            del got, a, b
        got = go(k, -data)
        assert torch.equal(got.Valid, ok), name
        a, b = got.Data[bits], -base.Data[bits]
        assert bool((a == b).all()), name  # the same values; a zero's sign may differ
        del got, a, b, base, bits
        torch.cuda.empty_cache()
