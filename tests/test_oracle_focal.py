"""The oracle's focal and transfer kernels and its validity rule against
independent brute-force definitions (focal/helpers_test.go naiveCell and
naiveValid, restated in numpy float32 scalars), the reference's
hand-computed cases (focal/focal_test.go:163-243) and its worked examples
(transfer/example_test.go). CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle
from strata_amd import focal, terrain, transfer
from strata_amd.raster import Float32Raster, NewFloat32, NewFloat32Like, data_grid, pack_bits, valid_grid

from helpers import Scene, hazard_values

F = np.float32
pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def go_min(a, b):
    if a != a:
        return a
    if b != b:
        return b
    if a == b:
        return a if math.copysign(1, a) < 0 else b
    return a if a < b else b


def go_max(a, b):
    if a != a:
        return a
    if b != b:
        return b
    if a == b:
        return b if math.copysign(1, a) < 0 else a
    return a if a > b else b


def naive_cell(kind, r, w, row, col, src, x, y):
    """focal/helpers_test.go:144-199, one float32 operation at a time."""
    k = 2 * r + 1

    def v(c, j):
        return F(src[y + j - r, x + c - r])

    if kind == "correlate":
        acc = F(F(w[0]) * v(0, 0))
        for t in range(1, k * k):
            acc = F(acc + F(F(w[t]) * v(t % k, t // k)))
        return acc
    if kind == "separable":
        acc = F(0)
        for c in range(k):
            cv = F(F(col[0]) * v(c, 0))
            for j in range(1, k):
                cv = F(cv + F(F(col[j]) * v(c, j)))
            p = F(F(row[c]) * cv)
            acc = p if c == 0 else F(acc + p)
        return acc
    if kind == "mean":
        acc = F(0)
        for c in range(k):
            cv = v(c, 0)
            for j in range(1, k):
                cv = F(cv + v(c, j))
            acc = cv if c == 0 else F(acc + cv)
        return F(acc / F(k * k))
    acc = v(0, 0)
    for j in range(k):
        for c in range(k):
            acc = go_min(acc, v(c, j)) if kind == "min" else go_max(acc, v(c, j))
    return F(acc)


def same(a, b):
    a, b = F(a), F(b)
    return (a != a and b != b) or a.view(np.uint32) == b.view(np.uint32)


def kernel_for(kind, r, w, row, col):
    if kind == "correlate":
        return focal.newWeightsKernel(focal.WeightsOptions(r, w), False)
    if kind == "separable":
        return focal.newSeparableKernel(focal.SeparableOptions(r, row, col))
    return focal.newBox(focal.BoxOptions(r), kind)


@pytest.mark.parametrize("kind", ["correlate", "separable", "mean", "min", "max"])
@pytest.mark.parametrize("r", [1, 2, 3, 5, 8])
def test_matches_naive(kind, r):
    rng = np.random.default_rng(r * 10 + len(kind))
    W, H = 2 * r + 9, 2 * r + 6
    k = 2 * r + 1
    src_vals = hazard_values(rng, W * H, base=10, relief=50, hazard_frac=0.03)
    valid = rng.random(W * H) > 0.05
    src = Float32Raster(src_vals, W, H, W, pack_bits(valid), 0)
    w = rng.uniform(-1, 1, k * k).astype(F)
    row, col = rng.uniform(-2, 2, k).astype(F), rng.uniform(-2, 2, k).astype(F)
    dst = Float32Raster(np.zeros(W * H, F), W, H, W, pack_bits(np.zeros(W * H, bool)), 0)
    oracle.process(kernel_for(kind, r, w, row, col), [dst], [src])
    g, gv = data_grid(dst), valid_grid(dst)
    s2 = src_vals.reshape(H, W)
    v2 = valid.reshape(H, W)
    for y in range(H):
        for x in range(W):
            inside = r <= x < W - r and r <= y < H - r
            want_valid = inside and v2[y - r:y + r + 1, x - r:x + r + 1].all() if inside else False
            assert gv[y, x] == want_valid
            if not inside:
                assert g[y, x] != g[y, x]
                continue
            assert same(g[y, x], naive_cell(kind, r, w, row, col, s2, x, y)), (kind, r, x, y)


def test_hand_computed():
    src = NewFloat32(3, 3, np.array([0, 1, 2, 10, 11, 12, 20, 21, 22], F))
    w = np.array([0, 0, 1, 0, 0, 0, 0, 0, 0], F)
    cases = [
        (focal.newWeightsKernel(focal.WeightsOptions(1, w), False), 2),
        (focal.newWeightsKernel(focal.WeightsOptions(1, w), True), 20),
        (focal.newSeparableKernel(focal.SeparableOptions(1, [0, 0, 1], [1, 0, 0])), 2),
        (focal.newSeparableKernel(focal.SeparableOptions(1, [1, 1, 1], [1, 2, 1])),
         0 + 1 + 2 + 2 * (10 + 11 + 12) + 20 + 21 + 22),
        (focal.newBox(focal.BoxOptions(1), "mean"), 11),
        (focal.newBox(focal.BoxOptions(1), "min"), 0),
        (focal.newBox(focal.BoxOptions(1), "max"), 22),
    ]
    for k, want in cases:
        dst = NewFloat32(3, 3, np.zeros(9, F))
        oracle.process(k, [dst], [src])
        assert dst.Data[4] == want
        assert all(v != v for i, v in enumerate(dst.Data) if i != 4)


def test_convolve_is_rotated_correlate():
    rng = np.random.default_rng(5)
    for r in (1, 2, 4, 8):
        k = 2 * r + 1
        src = Float32Raster(hazard_values(rng, 40 * 17), 40, 17, 40, pack_bits(rng.random(680) > 0.1), 0)
        w = rng.uniform(-1, 1, k * k).astype(F)
        a, b = NewFloat32Like(src), NewFloat32Like(src)
        oracle.process(focal.newWeightsKernel(focal.WeightsOptions(r, w), True), [a], [src])
        oracle.process(focal.newWeightsKernel(focal.WeightsOptions(r, w[::-1].copy()), False), [b], [src])
        assert (a.Data.view(np.uint32) == b.Data.view(np.uint32)).all()
        assert (a.Valid == b.Valid).all()


def test_signed_zero():
    negz = F(-0.0)
    src = NewFloat32(40, 3, np.full(120, negz, F))
    ones = np.ones(9, F)
    for k in (focal.newWeightsKernel(focal.WeightsOptions(1, ones), False),
              focal.newSeparableKernel(focal.SeparableOptions(1, ones[:3], ones[:3])),
              focal.newBox(focal.BoxOptions(1), "mean"), focal.newBox(focal.BoxOptions(1), "min"),
              focal.newBox(focal.BoxOptions(1), "max")):
        dst = NewFloat32Like(src)
        oracle.process(k, [dst], [src])
        assert (dst.Data[41:79].view(np.uint32) == 0x80000000).all()


def test_gaussian():
    for r in (1, 3, 8):
        for sigma in (0.5, 1, 2.5, 100):
            g = focal.Gaussian(r, sigma)
            assert len(g) == 2 * r + 1
            assert (g == g[::-1]).all()
            assert all(g[i] >= g[i - 1] for i in range(1, r + 1))
            assert abs(float(np.sum(g.astype(np.float64))) - 1) <= 1e-6
    for sigma in (0, -1, math.nan, math.inf):
        with pytest.raises(focal.FocalPanic, match="^focal: "):
            focal.Gaussian(2, sigma)
    with pytest.raises(focal.FocalPanic):
        focal.Gaussian(0, 1)


# -- transfer (transfer/example_test.go) -------------------------------------------

def run_point(k, vals):
    src = NewFloat32(len(vals), 1, np.array(vals, F))
    dst = NewFloat32Like(src)
    oracle.process(k, [dst], [src])
    return dst.Data


def test_reclass_example():
    k = transfer.newReclassKernel("t", [1.2, 5.9, 11.2, 21.3], [1, 2, 3, 4, 5])
    assert list(run_point(k, [0.4, 1.2, 8, 11.2, 21.3, 35])) == [1, 2, 3, 4, 5, 5]


def test_lookup_examples():
    k = transfer.newLookupKernel("t", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])
    got = run_point(k, [0, 7.5, 25, 40, 75])
    assert [float(v) for v in got] == [float(F(v)) for v in (1, 1.2, 2.45, 4.5, 4.5)]
    k = transfer.newLookupKernel("t", [0, 90, 180, 270, 360], [0.8, 1, 1.25, 1, 0.8])
    got = run_point(k, [0, 90, 180, 270, 360])
    assert [float(v) for v in got] == [float(F(v)) for v in (0.8, 1, 1.25, 1, 0.8)]


def test_lookup_reclass_nan_and_knots():
    k = transfer.newLookupKernel("t", [0, 1, 2], [np.nan, -0.0, 5])
    got = run_point(k, [np.nan, -1, 0, 1, 1.5, 2, 9])
    assert got[0] != got[0] and got[1] != got[1] and got[2] != got[2]  # NaN input, clamp to NaN, knot NaN
    assert got[3].view(np.uint32) == 0x80000000  # a knot returns its y exactly (-0)
    assert got[4] == F(2.5) and got[5] == 5 and got[6] == 5
    k = transfer.newReclassKernel("t", [1], [7, 8])
    got = run_point(k, [np.nan, -0.0, 1, 2])
    assert got[0] != got[0] and list(got[1:]) == [7, 8, 8]


def test_fire_risk_chain_example():
    """transfer/example_test.go:90-129: slope 21.8 degrees, class 3 in the
    interior, NaN carried from the border."""
    w, h = 5, 3
    elev = np.array([100 - 4 * x for y in range(h) for x in range(w)], F)
    dem = NewFloat32(w, h, elev)
    slope = NewFloat32Like(dem)
    oracle.process(terrain.slopeOp(terrain.SlopeOptions(CellSize=10)), [slope], [dem])
    assert round(float(slope.Data[7]), 1) == 21.8
    f = NewFloat32Like(dem)
    oracle.process(transfer.newLookupKernel("t", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5]), [f], [slope])
    cls = NewFloat32Like(dem)
    oracle.process(transfer.newReclassKernel("t", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5]), [cls], [f])
    # The reference multiplies by an aspect factor of 1 (aspect 90) before
    # classing; slope factor 1.9 + (21.8-20)/10 * 1.1 = 2.1 is class 3.
    assert cls.Data[7] == 3
    assert cls.Data[0] != cls.Data[0]


# -- validity against the naive definition at offsets and strides ----------------------------

@pytest.mark.parametrize("r", [1, 3])
def test_validity_windows_and_offsets(r):
    rng = np.random.default_rng(9 + r)
    sc = Scene()
    sc.add_parent("src", rng, 91, 30, masked=True, valid_base=37, invalid_frac=0.03)
    sc.add_window("s", "src", 5, 3, 70, 22)
    sc.add_parent("dst", rng, 75, 25, masked=True, valid_base=13)
    sc.add_window("d", "dst", 2, 1, 70, 22)
    ras, bufs = sc.materialize(None)
    before = bufs["dst"].valid.copy()
    oracle.process(focal.newBox(focal.BoxOptions(r), "max"), [ras["d"]], [ras["s"]])
    sv = valid_grid(ras["s"])
    dv = valid_grid(ras["d"])
    for y in range(22):
        for x in range(70):
            inside = r <= x < 70 - r and r <= y < 22 - r
            want = inside and sv[y - r:y + r + 1, x - r:x + r + 1].all()
            assert dv[y, x] == want
    # Bits outside the window are untouched.
    from strata_amd.raster import unpack_bits
    allbits_before = unpack_bits(before, 0, len(before) * 64)
    allbits_after = unpack_bits(bufs["dst"].valid, 0, len(before) * 64)
    own = np.zeros(len(allbits_before), bool)
    d = ras["d"]
    for y in range(22):
        own[d.ValidOffset + y * d.Stride:d.ValidOffset + y * d.Stride + 70] = True
    assert (allbits_before[~own] == allbits_after[~own]).all()
