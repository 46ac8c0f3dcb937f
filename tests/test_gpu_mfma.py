"""The split-BF16x3 MFMA Correlate path and the FP64 error report.

The MFMA path is approximate by design: each product w*x is formed from
two-part bf16 splits, so its error against the exact product is within
about 3*2^-18 of |w*x|, and the fp32 accumulation adds at most a few ulps of
sum |w x|. The tests bound max |out - ref| / sum |w x| against the FP64
correlation, check that it is exact where the exact path is required
(windows holding NaN or Inf get the FP32 path's value, bit for bit), and
that masks and the NaN edge ring are the FP32 path's. The FP64 error kernel
itself is checked against the CPU oracle's FP64 correlation.
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import focal
from strata_amd.engine import Options, ProcessN
from strata_amd.raster import Float32Raster

from helpers import Scene, assert_same_parents, host_copy, same_floats

pytestmark = pytest.mark.gpu

# Each product within 3*2^-18 |w x|; fp32 sums of up to 289 terms add at
# most 289*2^-24 of sum |w x|. The bound below leaves a factor of two.
SCALED_BOUND = 2 * (3 * 2.0 ** -18 + 289 * 2.0 ** -24)


def weights(r, seed):
    rng = np.random.default_rng(seed)
    n = 2 * r + 1
    return rng.uniform(-1, 1, n * n).astype(np.float32)


def device_raster(gpu, values, W, H, stride=None, valid=None, off=0):
    import torch
    stride = stride or W
    v = torch.from_numpy(np.ascontiguousarray(values, dtype=np.float32)).to(gpu)
    m = None if valid is None else torch.from_numpy(valid.view(np.int64).copy()).to(gpu)
    return Float32Raster(v, W, H, stride, m, off)


def f64_error(src_h, out_h, w, r):
    """max |out - ref| / sum|w x| and max |out - ref| over the interior, on the CPU."""
    ref = oracle.correlate_f64(src_h, w, r)
    W, H = src_h.Width, src_h.Height
    d = np.asarray(src_h.Data, dtype=np.float64)[: (H - 1) * src_h.Stride + W]
    grid = np.lib.stride_tricks.as_strided(d, (H, W), (src_h.Stride * 8, 8))
    out = np.lib.stride_tricks.as_strided(np.asarray(out_h.Data, np.float32), (H, W), (out_h.Stride * 4, 4))
    n = 2 * r + 1
    mag = np.zeros((H - 2 * r, W - 2 * r))
    for j in range(n):
        for c in range(n):
            mag += np.abs(float(w[j * n + c]) * grid[j:j + H - 2 * r, c:c + W - 2 * r])
    err = np.abs(out[r:H - r, r:W - r].astype(np.float64) - ref[r:H - r, r:W - r])
    return float((err / mag).max()), float(err.max())


@pytest.mark.parametrize("r", [1, 2, 4, 5, 8])
def test_mfma_within_bound_of_f64(gpu, r):
    import torch
    rng = np.random.default_rng(r)
    W, H = 301, 187
    vals = rng.uniform(0, 1000, W * H).astype(np.float32)
    w = weights(r, 10 + r)
    src = device_raster(gpu, vals, W, H)
    out = device_raster(gpu, np.zeros(W * H), W, H)
    k = focal.newWeightsKernel(focal.WeightsOptions(r, w), False, mfma=True)
    assert ProcessN(None, [out], [src], k, Options()) is None
    torch.cuda.synchronize()
    host_src = Float32Raster(vals, W, H, W)
    host_out = Float32Raster(out.Data.cpu().numpy(), W, H, W)
    scaled, _ = f64_error(host_src, host_out, w, r)
    assert scaled <= SCALED_BOUND, f"r={r}: max |err|/sum|wx| = {scaled:.3g} > {SCALED_BOUND:.3g}"
    rep = focal.CorrelateError(k, src, out)
    assert rep["cells"] == (W - 2 * r) * (H - 2 * r)
    assert rep["max_scaled"] == pytest.approx(scaled, rel=1e-9, abs=1e-15)
    # The ring is NaN, exactly as on the FP32 path.
    o = host_out.Data.reshape(H, W)
    assert np.isnan(o[:r]).all() and np.isnan(o[-r:]).all() and np.isnan(o[:, :r]).all() and np.isnan(o[:, -r:]).all()


def test_f64_report_matches_cpu_oracle(gpu):
    """The device FP64 correlation against the oracle's, on the FP32 path's output."""
    import torch
    r = 4
    rng = np.random.default_rng(3)
    W, H = 257, 131
    vals = (rng.standard_normal(W * H) * 100).astype(np.float32)
    w = weights(r, 4)
    src = device_raster(gpu, vals, W, H)
    out = device_raster(gpu, np.zeros(W * H), W, H)
    k = focal.newWeightsKernel(focal.WeightsOptions(r, w), False)
    assert ProcessN(None, [out], [src], k, Options()) is None
    torch.cuda.synchronize()
    host_src = Float32Raster(vals, W, H, W)
    host_out = Float32Raster(out.Data.cpu().numpy(), W, H, W)
    scaled, mx = f64_error(host_src, host_out, w, r)
    rep = focal.CorrelateError(k, src, out)
    assert rep["max_abs"] == pytest.approx(mx, rel=1e-12, abs=1e-12)
    assert rep["max_scaled"] == pytest.approx(scaled, rel=1e-9)
    assert rep["max_scaled"] < 289 * 2.0 ** -24  # the exact FP32 order is within its own rounding


@pytest.mark.parametrize("r,kind", [(4, "masked"), (5, "windows"), (8, "windows"), (2, "unmasked-in")])
def test_mfma_hazards_masks_and_layouts(gpu, r, kind):
    """Hazard inputs (NaN, ±Inf, ±MaxFloat32, subnormals) in windowed,
    masked operands: masks and untouched memory are bit-exact with the
    oracle; cells whose window holds a non-finite value are bit-exact with
    the FP32 path; the rest are within the bound."""
    import torch
    from test_gpu_parity import build_scene

    w = weights(r, 20 + r)
    k32 = focal.newWeightsKernel(focal.WeightsOptions(r, w), False)
    kmf = focal.newWeightsKernel(focal.WeightsOptions(r, w), False, mfma=True)
    sc = build_scene(kind, k32, seed=40 + r, W=173, H=97)
    host, hbufs = sc.materialize(None)
    oracle.process(k32, [host["d0"]], [host["s"]])
    dev, dbufs = sc.materialize(gpu)
    assert ProcessN(None, [dev["d0"]], [dev["s"]], kmf, Options()) is None
    torch.cuda.synchronize()
    got = host_copy(dbufs)
    # Masks and everything outside the output cells: bit-exact.
    win = sc.wins["d0"]
    for name in hbufs:
        if name != win.parent:
            assert_same_parents({name: got[name]}, {name: hbufs[name]}, "untouched parent")
    gp, wp = got[win.parent], hbufs[win.parent]
    if wp.valid is not None:
        assert np.array_equal(gp.valid, wp.valid), "mask differs"
    # Cells: exact where the window is non-finite (and on the ring), bounded elsewhere.
    W, H = win.w, win.h
    g = np.array([gp.data[(win.y + y) * win.width + win.x: (win.y + y) * win.width + win.x + W] for y in range(H)])
    e = np.array([wp.data[(win.y + y) * win.width + win.x: (win.y + y) * win.width + win.x + W] for y in range(H)])
    sw = sc.wins["s"]
    sp = sc.parents[sw.parent].data
    src = np.array([sp[(sw.y + y) * sw.width + sw.x: (sw.y + y) * sw.width + sw.x + W] for y in range(H)],
                   dtype=np.float64)
    n = 2 * r + 1
    finite = np.isfinite(src)
    win_finite = np.ones((H - 2 * r, W - 2 * r), bool)
    mag = np.zeros((H - 2 * r, W - 2 * r))
    ref = np.zeros((H - 2 * r, W - 2 * r))
    with np.errstate(invalid="ignore", over="ignore"):
        for j in range(n):
            for c in range(n):
                blk = src[j:j + H - 2 * r, c:c + W - 2 * r]
                win_finite &= finite[j:j + H - 2 * r, c:c + W - 2 * r]
                t = float(w[j * n + c]) * blk
                ref += t
                mag += np.abs(t)
    gi, ei = g[r:H - r, r:W - r], e[r:H - r, r:W - r]
    exact = same_floats(gi, ei)
    assert exact[~win_finite].all(), "a window with a non-finite value differs from the FP32 path"
    with np.errstate(invalid="ignore"):
        scaled = np.abs(gi.astype(np.float64) - ref) / np.where(mag > 0, mag, 1.0)
    ok = win_finite & (mag < 1e37)  # no fp32 overflow in the sums
    assert (scaled[ok] <= SCALED_BOUND).all(), f"max scaled error {scaled[ok].max():.3g}"
    ring = np.ones((H, W), bool)
    ring[r:H - r, r:W - r] = False
    assert np.isnan(g[ring]).all()
