"""The oracle's Wood quadratic fit (FitRadius 1..8, terrain/fit.go:29-226)
against the closed forms the reference's own fit tests hold
(terrain/fit_test.go:35-132): on integer planes the fitted gradient is the
plane's and every curvature is +0; on a quadratic surface, which least
squares fits exactly, the gradient and the three curvatures are the
surface's own, to the float32 rounding bound those tests derive; a fitted
output's validity is the radius-r box rule (terrain/fit_test.go:230-248).
CPU only."""

import numpy as np
import pytest

from oracle import oracle
from strata_amd import terrain
from strata_amd.raster import NewFloat32, NewFloat32Like, data_grid

from helpers import plane_values

MAX_RADIUS = 8
CURVATURES = (terrain.CurvatureProfile, terrain.CurvaturePlan, terrain.CurvatureMean)


def quadratic(w, h, cx, cy, c):
    """terrain/fit_test.go:24-33: c0 x² + c1 y² + c2 xy + c3 x + c4 y + c5 at
    x = col·cx, y = row·cy, rounded once to float32."""
    d = np.zeros(w * h, np.float32)
    for row in range(h):
        for col in range(w):
            x, y = col * cx, row * cy
            d[row * w + col] = np.float32(c[0] * x * x + c[1] * y * y + c[2] * x * y + c[3] * x + c[4] * y + c[5])
    return NewFloat32(w, h, d)


def border_mask(h, w, r):
    m = np.zeros((h, w), bool)
    m[:r, :] = m[-r:, :] = True
    m[:, :r] = m[:, -r:] = True
    return m


@pytest.mark.parametrize("ab", [(0, 0), (1, 0), (0, -3), (2, 5), (-7, 4)])
def test_fit_plane(ab):
    """terrain/fit_test.go:39-75 at every radius."""
    a, b = ab
    for r in range(1, MAX_RADIUS + 1):
        dem = NewFloat32(2 * r + 5, 2 * r + 4, plane_values(2 * r + 5, 2 * r + 4, a, b, 1, 1))
        dx, dy = NewFloat32Like(dem), NewFloat32Like(dem)
        oracle.process(terrain.gradientOp(terrain.GradientOptions(CellSize=1, FitRadius=r)), [dx, dy], [dem])
        border = border_mask(dem.Height, dem.Width, r)
        gx, gy = data_grid(dx), data_grid(dy)
        assert np.isnan(gx[border]).all() and np.isnan(gy[border]).all()
        assert (np.abs(gx[~border].astype(np.float64) - a) <= 1e-6 * abs(a)).all()
        assert (np.abs(gy[~border].astype(np.float64) - b) <= 1e-6 * abs(b)).all()
        for ct in CURVATURES:
            cv = NewFloat32Like(dem)
            oracle.process(terrain.curvatureOp(terrain.CurvatureOptions(CellSize=1, Type=ct, FitRadius=r)), [cv],
                           [dem])
            g = data_grid(cv)
            assert np.isnan(g[border]).all()
            assert (g[~border].view(np.uint32) == 0).all(), (ab, r, ct)


@pytest.mark.parametrize("r", [1, 2, 4, 8])
def test_fit_quadratic(r):
    """terrain/fit_test.go:77-132: rectangular cells and a ZFactor."""
    w, h = 40, 36
    cx, cy, zf = 2.0, 3.0, 1.5
    coef = (0.03, -0.02, 0.01, 0.4, -0.25, 50)
    dem = quadratic(w, h, cx, cy, coef)
    zmax = float(np.abs(dem.Data.astype(np.float64)).max())
    dx, dy = NewFloat32Like(dem), NewFloat32Like(dem)
    oracle.process(terrain.gradientOp(terrain.GradientOptions(CellSize=cx, CellSizeY=cy, ZFactor=zf, FitRadius=r)),
                   [dx, dy], [dem])
    curv = {}
    for ct in CURVATURES:
        curv[ct] = NewFloat32Like(dem)
        oracle.process(terrain.curvatureOp(terrain.CurvatureOptions(CellSize=cx, CellSizeY=cy, ZFactor=zf, Type=ct,
                                                                    FitRadius=r)), [curv[ct]], [dem])
    gxs, gys = data_grid(dx), data_grid(dy)
    cgs = {ct: data_grid(c) for ct, c in curv.items()}
    # The elevations' float32 rounding (up to zmax·2⁻²⁴ per cell) and the
    # sums' (2r+1)² further roundings, over the fit's divisors.
    k = float(2 * r + 1)
    e = 4 * k * k * zmax * 2.0 ** -24
    s2 = r * (r + 1) * (2 * r + 1) / 3
    gtol_x = zf * e * r * k / (k * s2 * cx)
    gtol_y = zf * e * r * k / (k * s2 * cy)
    c0, c1, c2, c3, c4, _ = coef
    rr, tt, ss = zf * 2 * c0, zf * 2 * c1, zf * c2
    checked = 0
    for row in range(r, h - r):
        for col in range(r, w - r):
            x, y = col * cx, row * cy
            p = zf * (2 * c0 * x + c2 * y + c3)
            q = zf * (2 * c1 * y + c2 * x + c4)
            gx, gy = float(gxs[row, col]), float(gys[row, col])
            assert abs(gx - p) <= gtol_x + 1e-6 * abs(p), (r, col, row, gx, p)
            assert abs(gy - q) <= gtol_y + 1e-6 * abs(q), (r, col, row, gy, q)
            g = p * p + q * q
            want = {
                terrain.CurvatureProfile: -(p * p * rr + 2 * p * q * ss + q * q * tt) / (g * (1 + g) ** 1.5),
                terrain.CurvaturePlan: -(q * q * rr - 2 * p * q * ss + p * p * tt) / g ** 1.5,
                terrain.CurvatureMean: -((1 + q * q) * rr - 2 * p * q * ss + (1 + p * p) * tt) / (2 * (1 + g) ** 1.5),
            }
            for ct, v in want.items():
                got = float(cgs[ct][row, col])
                assert abs(got - v) <= 2e-3 * abs(v) + 1e-6, (r, ct, col, row, got, v)
            checked += 1
    assert checked == (w - 2 * r) * (h - 2 * r)
    border = border_mask(h, w, r)
    assert np.isnan(gxs[border]).all() and all(np.isnan(c[border]).all() for c in cgs.values())


def test_fit_quadratic_fails_a_dropped_term():
    """The quadratic check is sharp enough to catch a plausible mistake: the
    same surface without its xy term gives a plan curvature the closed form
    of the full surface rejects."""
    w, h, r = 40, 36, 2
    coef = (0.03, -0.02, 0.0, 0.4, -0.25, 50)
    dem = quadratic(w, h, 2.0, 3.0, coef)
    cv = NewFloat32Like(dem)
    oracle.process(terrain.curvatureOp(terrain.CurvatureOptions(CellSize=2.0, CellSizeY=3.0, ZFactor=1.5,
                                                                Type=terrain.CurvaturePlan, FitRadius=r)), [cv], [dem])
    got = data_grid(cv)
    zf, (c0, c1, c2, c3, c4) = 1.5, (0.03, -0.02, 0.01, 0.4, -0.25)
    misses = 0
    for row in range(r, h - r):
        for col in range(r, w - r):
            x, y = col * 2.0, row * 3.0
            p, q = zf * (2 * c0 * x + c2 * y + c3), zf * (2 * c1 * y + c2 * x + c4)
            rr, tt, ss = zf * 2 * c0, zf * 2 * c1, zf * c2
            g = p * p + q * q
            v = -(q * q * rr - 2 * p * q * ss + p * p * tt) / g ** 1.5
            misses += abs(float(got[row, col]) - v) > 2e-3 * abs(v) + 1e-6
    assert misses > 0


@pytest.mark.parametrize("r", [1, 3, 8])
def test_fit_validity_is_the_radius_r_box(r):
    """terrain/fit_test.go:230-248: a fitted output is valid exactly where
    its (2r+1)² window lies inside the raster and holds only valid inputs
    (the reference compares with a radius-r box op's validity; here the box
    rule itself, brute force)."""
    from helpers import hazard_values
    from strata_amd.raster import Float32Raster, pack_bits, valid_grid

    rng = np.random.default_rng(83 + r)
    w, h = 70, 2 * r + 11
    bits = rng.random(w * h) >= 0.02
    dem = Float32Raster(hazard_values(rng, w * h), w, h, w, pack_bits(bits), 0)
    sv = bits.reshape(h, w)
    want = np.zeros((h, w), bool)
    for y in range(r, h - r):
        for x in range(r, w - r):
            want[y, x] = sv[y - r:y + r + 1, x - r:x + r + 1].all()
    for k in (terrain.slopeOp(terrain.SlopeOptions(CellSize=1, FitRadius=r)),
              terrain.curvatureOp(terrain.CurvatureOptions(CellSize=1, FitRadius=r))):
        got = NewFloat32Like(dem)
        oracle.process(k, [got], [dem])
        assert (valid_grid(got) == want).all(), k.name
