"""The oracle's terrain kernels against the reference's own known answers
(terrain/terrain_test.go, terrain/shade_test.go, terrain/curvature_test.go,
internal/stencil/stencil_test.go) and the mathematics: closed forms on
planes and quadrics, float64 evaluations with derived tolerances, the
Esri worked examples. CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle
from strata_amd import terrain
from strata_amd.raster import Float32Raster, NewFloat32, NewFloat32Like, NewMask, data_grid, valid_grid

from helpers import plane_values


def run1(k, dem):
    dst = NewFloat32Like(dem)
    oracle.process(k, [dst], [dem])
    return dst


def check_border_nan(g):
    assert np.isnan(g[0]).all() and np.isnan(g[-1]).all() and np.isnan(g[:, 0]).all() and np.isnan(g[:, -1]).all()


def plane(w, h, ax, ay, cs, csy):
    return NewFloat32(w, h, plane_values(w, h, ax, ay, cs, csy))


# -- terrain/terrain_test.go -------------------------------------------------

def test_tilted_plane():
    ax, ay, cs, csy = 0.3, -0.4, 10.0, 5.0
    dem = plane(9, 7, ax, ay, cs, csy)
    dx, dy = NewFloat32Like(dem), NewFloat32Like(dem)
    oracle.process(terrain.gradientOp(terrain.GradientOptions(CellSize=cs, CellSizeY=csy)), [dx, dy], [dem])
    for g, want in ((data_grid(dx), ax), (data_grid(dy), ay)):
        check_border_nan(g)
        assert np.abs(g[1:-1, 1:-1] - want).max() <= 1e-5
    m = math.hypot(ax, ay)
    for units, want, tol in ((terrain.SlopeDegrees, math.degrees(math.atan(m)), 1e-3),
                             (terrain.SlopeRadians, math.atan(m), 2e-5), (terrain.SlopePercent, 100 * m, 1e-3)):
        g = data_grid(run1(terrain.slopeOp(terrain.SlopeOptions(CellSize=cs, CellSizeY=csy, Units=units)), dem))
        check_border_nan(g)
        assert np.abs(g[1:-1, 1:-1] - want).max() <= tol
    g = data_grid(run1(terrain.slopeOp(terrain.SlopeOptions(CellSize=cs, CellSizeY=csy, ZFactor=2,
                                                            Units=terrain.SlopePercent)), dem))
    assert np.abs(g[1:-1, 1:-1] - 200 * m).max() <= 2e-3


def test_flat_is_plus_zero():
    dem = NewFloat32(12, 5, np.full(60, 1234.5, np.float32))
    for u in (terrain.SlopeDegrees, terrain.SlopePercent, terrain.SlopeRadians):
        g = data_grid(run1(terrain.slopeOp(terrain.SlopeOptions(CellSize=30, Units=u)), dem))
        assert (g[1:-1, 1:-1].view(np.uint32) == 0).all()


def test_esri_slope_hand_computed():
    dem = NewFloat32(3, 3, np.array([50, 45, 50, 30, 30, 30, 8, 10, 10], np.float32))
    dx, dy = NewFloat32Like(dem), NewFloat32Like(dem)
    oracle.process(terrain.gradientOp(terrain.GradientOptions(CellSize=5)), [dx, dy], [dem])
    assert abs(dx.Data[4] - 0.05) <= 1e-6 and abs(dy.Data[4] + 3.8) <= 1e-6
    rr = math.hypot(0.05, 3.8)
    for units, want, tol in ((terrain.SlopeDegrees, 75.25762, 1e-4), (terrain.SlopePercent, 100 * rr, 1e-3),
                             (terrain.SlopeRadians, math.atan(rr), 1e-6)):
        got = run1(terrain.slopeOp(terrain.SlopeOptions(CellSize=5, Units=units)), dem).Data[4]
        assert abs(got - want) <= tol


def test_gradient_signs():
    dem = plane(5, 5, 1, 2, 1, 1)
    dx, dy = NewFloat32Like(dem), NewFloat32Like(dem)
    oracle.process(terrain.gradientOp(terrain.GradientOptions(CellSize=1)), [dx, dy], [dem])
    assert dx.Data[12] == 1 and dy.Data[12] == 2


# -- terrain/shade_test.go ---------------------------------------------------

def angle_diff(a, b):
    d = math.fmod(abs(a - b), 360)
    return min(d, 360 - d)


@pytest.mark.parametrize("name,ax,ay,bearing,trig", [
    ("N", 0, 0.5, 0, 90), ("NE", -0.3, 0.3, 45, 45), ("E", -0.5, 0, 90, 0), ("SE", -0.2, -0.2, 135, 315),
    ("S", 0, -0.5, 180, 270), ("SW", 0.4, -0.4, 225, 225), ("W", 0.5, 0, 270, 180), ("NW", 0.1, 0.1, 315, 135),
    ("NNE", -0.25, 0.25 * math.sqrt(3), 30, 60)])
def test_aspect_planes(name, ax, ay, bearing, trig):
    dem = plane(9, 6, ax, ay, 10, 10)
    for t, want in ((False, bearing), (True, trig)):
        g = data_grid(run1(terrain.aspectOp(terrain.AspectOptions(CellSize=10, Trigonometric=t)), dem))
        check_border_nan(g)
        for v in g[1:-1, 1:-1].ravel():
            assert 0 <= v < 360 and angle_diff(float(v), want) <= 1e-4


def test_aspect_anisotropic():
    dem = plane(7, 7, 0.05, 0.1, 20, 10)
    g = data_grid(run1(terrain.aspectOp(terrain.AspectOptions(CellSize=20, CellSizeY=10)), dem))
    want = 360 + math.degrees(math.atan2(-0.05, 0.1))
    assert max(angle_diff(float(v), want) for v in g[1:-1, 1:-1].ravel()) <= 1e-4


def test_aspect_exact_axes():
    for ax, ay, want in ((0, 1, 0), (-1, 0, 90), (0, -1, 180), (1, 0, 270)):
        g = data_grid(run1(terrain.aspectOp(terrain.AspectOptions(CellSize=1)), plane(5, 5, ax, ay, 1, 1)))
        inner = g[1:-1, 1:-1]
        assert np.abs(inner - want).max() <= 2e-5
        assert not (inner.view(np.uint32) == 0x80000000).any()
    d = np.array([0, 0, 0, 0, 0, 0, 1e6, 1e6, 1e6 + 0.0625], np.float32)
    for ax in (0.0625, -0.0625):
        d[8] = np.float32(1e6 + ax)
        v = run1(terrain.aspectOp(terrain.AspectOptions(CellSize=1)), NewFloat32(3, 3, d.copy())).Data[4]
        assert 0 <= v < 1e-4 and not math.copysign(1, v) < 0


def test_aspect_flat():
    dem = NewFloat32(10, 6, np.full(60, 812.25, np.float32))
    dem.Valid = NewMask(60)
    for opts, want in ((terrain.AspectOptions(CellSize=5), -1.0),
                       (terrain.AspectOptions(CellSize=5, Trigonometric=True), -1.0),
                       (terrain.AspectOptions(CellSize=5, ZeroForFlat=True), 0.0)):
        dst = run1(terrain.aspectOp(opts), dem)
        g, v = data_grid(dst), valid_grid(dst)
        check_border_nan(g)
        assert (g[1:-1, 1:-1].view(np.uint32) == np.float32(want).view(np.uint32)).all()
        assert v[1:-1, 1:-1].all() and not v[0].any()
    g = data_grid(run1(terrain.aspectOp(terrain.AspectOptions(CellSize=1)), plane(5, 5, 0, 1e-3, 1, 1)))
    assert g[2, 2] == 0


def test_aspect_esri_hand_computed():
    dem = NewFloat32(3, 3, np.array([101, 92, 85, 101, 92, 85, 101, 91, 84], np.float32))
    got = run1(terrain.aspectOp(terrain.AspectOptions(CellSize=5)), dem).Data[4]
    want = 90 + math.degrees(math.atan(0.375 / 8.125))
    assert abs(got - want) <= 1e-4 and abs(want - 92.64) <= 5e-3


def horn64(d, n, x, cs, csy):
    def z(dx, dy):
        return float(d[(1 + dy) * n + x + dx])
    gx = ((z(1, -1) + 2 * z(1, 0) + z(1, 1)) - (z(-1, -1) + 2 * z(-1, 0) + z(-1, 1))) / (8 * cs)
    gy = ((z(-1, 1) + 2 * z(0, 1) + z(1, 1)) - (z(-1, -1) + 2 * z(0, -1) + z(1, -1))) / (8 * csy)
    return gx, gy


def test_aspect_random_windows():
    rng = np.random.default_rng(31)
    n = 200
    d = (rng.standard_normal(n * 3) * 10.0 ** rng.integers(0, 5, n * 3)).astype(np.float32)
    dem = NewFloat32(n, 3, d)
    for trig in (False, True):
        g = data_grid(run1(terrain.aspectOp(terrain.AspectOptions(CellSize=3, CellSizeY=4, Trigonometric=trig)), dem))
        for x in range(1, n - 1):
            gx, gy = horn64(d, n, x, 3, 4)
            want = math.degrees(math.atan2(gy, -gx) if trig else math.atan2(-gx, gy))
            v = float(g[1, x])
            assert 0 <= v < 360 and angle_diff(v, want) <= 1e-3


@pytest.mark.parametrize("az,alt,want_alt", [(0, 0, 45), (90, 30, 30), (360, 90, 90), (0, 1e-3, 1e-3)])
def test_hillshade_flat(az, alt, want_alt):
    dem = NewFloat32(8, 5, np.full(40, -42, np.float32))
    g = data_grid(run1(terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30, Azimuth=az, Altitude=alt)), dem))
    check_border_nan(g)
    assert np.abs(g[1:-1, 1:-1] - 255 * math.sin(math.radians(want_alt))).max() <= 1e-4


def test_hillshade_huge_azimuth():
    dem = plane(6, 5, 3, -1, 1, 1)
    for az in (1e308, -1e308, 360 * float(1 << 60), 1e20 + 45):
        red = math.fmod(az, 360) or 360
        got = data_grid(run1(terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=1, Azimuth=az)), dem))
        want = data_grid(run1(terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=1, Azimuth=red)), dem))
        assert (got[1:-1, 1:-1] == want[1:-1, 1:-1]).all()


@pytest.mark.parametrize("az,alt", [(315, 45), (0, 45), (90, 30), (200, 60), (45, 80), (270, 10)])
def test_hillshade_facing_light(az, alt):
    azr, altr = math.radians(az), math.radians(alt)
    fall = math.tan(math.pi / 2 - altr)
    ax, ay = -fall * math.sin(azr), fall * math.cos(azr)
    opts = terrain.HillshadeOptions(CellSize=2, Azimuth=360 if az == 0 else az, Altitude=alt)
    g = data_grid(run1(terrain.hillshadeOp(opts), plane(9, 7, ax, ay, 2, 2)))[1:-1, 1:-1]
    assert np.abs(g - 255).max() <= 5e-3 and (g <= 255).all()
    g = data_grid(run1(terrain.hillshadeOp(opts), plane(9, 7, -ax, -ay, 2, 2)))[1:-1, 1:-1]
    away = max(0, 255 * math.sin(2 * altr - math.pi / 2))
    assert np.abs(g - away).max() <= 5e-3
    if alt < 45:
        assert (g.view(np.uint32) == 0).all()


def test_hillshade_hand_computed():
    esri = NewFloat32(3, 3, np.array([101, 92, 85, 101, 92, 85, 101, 91, 84], np.float32))
    assert run1(terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=5)), esri).Data[4] == 0
    got = run1(terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=5, Azimuth=90)), esri).Data[4]
    s = math.atan(math.hypot(1.625, 0.075))
    asp = math.radians(90 + math.degrees(math.atan(0.075 / 1.625)))
    want = 255 * (math.sin(math.pi / 4) * math.cos(s) + math.cos(math.pi / 4) * math.sin(s) * math.cos(math.pi / 2 - asp))
    assert abs(got - want) <= 1e-3 and abs(want - 247.874) <= 1e-3
    rng = np.random.default_rng(41)
    n = 300
    d = (rng.standard_normal(n * 3) * 20).astype(np.float32)
    dem = NewFloat32(n, 3, d)
    for _ in range(20):
        az, alt = rng.random() * 720 - 360, 0.5 + rng.random() * 89.5
        cs, csy = 1 + rng.random() * 30, 1 + rng.random() * 30
        g = data_grid(run1(terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=cs, CellSizeY=csy, Azimuth=az,
                                                                        Altitude=alt)), dem))
        for x in range(1, n - 1):
            gx, gy = horn64(d, n, x, cs, csy)
            slope = math.atan(math.hypot(gx, gy))
            aspect = math.atan2(-gx, gy)
            altr, azr = math.radians(alt), math.radians(az)
            want = max(0.0, 255 * (math.sin(altr) * math.cos(slope) + math.cos(altr) * math.sin(slope) *
                                   math.cos(azr - aspect)))
            v = float(g[1, x])
            assert 0 <= v <= 255 and abs(v - want) <= 2e-3


# -- terrain/curvature_test.go -------------------------------------------------

TYPES = (terrain.CurvatureProfile, terrain.CurvaturePlan, terrain.CurvatureMean)


def curvature64(ct, d):
    p, q, r, s, t = d
    g = p * p + q * q
    w = 1 + g
    if ct == terrain.CurvatureProfile:
        return 0.0 if g == 0 else -(p * p * r + 2 * p * q * s + q * q * t) / (g * w * math.sqrt(w))
    if ct == terrain.CurvaturePlan:
        return 0.0 if g == 0 else -(q * q * r - 2 * p * q * s + p * p * t) / (g * math.sqrt(g))
    return -((1 + q * q) * r - 2 * p * q * s + (1 + p * p) * t) / (2 * w * math.sqrt(w))


def test_curvature_hand_computed():
    dem = NewFloat32(3, 3, np.array([10, 12, 15, 11, 13, 17, 13, 16, 20], np.float32))
    for ct, want in ((terrain.CurvatureProfile, -32 / (13 * 14 ** 1.5)), (terrain.CurvaturePlan, -20 / 13 ** 1.5),
                     (terrain.CurvatureMean, -12 / 14 ** 1.5)):
        got = run1(terrain.curvatureOp(terrain.CurvatureOptions(CellSize=1, Type=ct)), dem).Data[4]
        assert abs(got - want) <= 1e-6 * abs(want)


@pytest.mark.parametrize("a,b,c,d,e,cx,cy,z", [
    (3, -2, 1, 5, -7, 2, 4, 1), (-1, 0, -1, 0, 0, 1, 1, 1), (1, 0, -1, 0.5, 0, 0.5, 0.5, 1),
    (2, 3, -4, -1, 2, 8, 2, -0.25)])
def test_curvature_quadric(a, b, c, d, e, cx, cy, z):
    """Every elevation and ZT derivative is exact here, so only the
    formula's float32 rounding separates the result from the analytic one."""
    w, h = 9, 7
    vals = np.zeros(w * h, np.float32)
    for y in range(h):
        for x in range(w):
            X, Y = (x - w // 2) * cx, (y - h // 2) * cy
            vals[y * w + x] = np.float32(a * X * X + b * X * Y + c * Y * Y + d * X + e * Y)
    dem = NewFloat32(w, h, vals)
    for ct in TYPES:
        g = data_grid(run1(terrain.curvatureOp(terrain.CurvatureOptions(CellSize=cx, CellSizeY=cy, ZFactor=z,
                                                                        Type=ct)), dem))
        check_border_nan(g)
        for y in range(1, h - 1):
            for x in range(1, w - 1):
                X, Y = (x - w // 2) * cx, (y - h // 2) * cy
                der = (z * (2 * a * X + b * Y + d), z * (b * X + 2 * c * Y + e), z * 2 * a, z * b, z * 2 * c)
                want = curvature64(ct, der)
                # 16 ulp of the magnitudes involved: a dropped term or a
                # wrong sign moves the result by far more.
                # The reference's scale (terrain/curvature_test.go
                # curvatureTol): the terms' magnitudes made to add up.
                mags = [abs(v) for v in der]
                if ct != terrain.CurvatureProfile:
                    mags[3] = -mags[3]
                scale = abs(curvature64(ct, tuple(mags))) + abs(want)
                assert abs(float(g[y, x]) - want) <= 16 * 2 ** -24 * scale + 1e-30


def test_curvature_flat_and_plane():
    for dem in (NewFloat32(8, 6, np.full(48, 812.25, np.float32)), plane(8, 6, 3, -2, 1, 1)):
        dem.Valid = NewMask(48)
        for ct in TYPES:
            for o in (terrain.CurvatureOptions(CellSize=1, Type=ct),
                      terrain.CurvatureOptions(CellSize=3, CellSizeY=0.5, ZFactor=-2, Type=ct)):
                dst = run1(terrain.curvatureOp(o), dem)
                g, v = data_grid(dst), valid_grid(dst)
                assert (g[1:-1, 1:-1].view(np.uint32) == 0).all() and v[1:-1, 1:-1].all()


def test_curvature_signs():
    def surface(f):
        d = np.zeros(9, np.float32)
        for y in range(3):
            for x in range(3):
                d[y * 3 + x] = np.float32(f(float(x + 3), float(y - 1)))
        return NewFloat32(3, 3, d)

    def at(dem, ct):
        return run1(terrain.curvatureOp(terrain.CurvatureOptions(CellSize=1, Type=ct)), dem).Data[4]

    dome = surface(lambda X, Y: 100 - X * X - Y * Y)
    bowl = surface(lambda X, Y: X * X + Y * Y)
    saddle = surface(lambda X, Y: X * X - Y * Y)
    for ct in TYPES:
        assert at(dome, ct) > 0 and at(bowl, ct) < 0
    assert at(saddle, terrain.CurvatureProfile) < 0 and at(saddle, terrain.CurvaturePlan) > 0


def test_curvature_nan_centre():
    for c in (np.nan, np.inf, -np.inf):
        dem = NewFloat32(3, 3, np.array([5, 5, 5, 5, c, 5, 5, 5, 5], np.float32))
        for ct in TYPES:
            v = run1(terrain.curvatureOp(terrain.CurvatureOptions(CellSize=1, Type=ct)), dem).Data[4]
            if ct == terrain.CurvatureMean and c == c:
                assert math.isinf(v)
            else:
                assert v != v


# -- internal/stencil/stencil_test.go ------------------------------------------------------

def test_horn_gradient_near_flat():
    """Each neighbour difference is exact (Sterbenz), so dx and dy are
    within one ulp of the sum of the magnitudes of their terms."""
    rng = np.random.default_rng(31)
    n = 37
    for base in (500, 1000, 4000, 8800):
        for relief in (0.01, 1, 100):
            rows = (base + (rng.random((3, n + 2)) * 2 - 1) * relief).astype(np.float32)
            dem = NewFloat32(n + 2, 3, rows.ravel().copy())
            dx, dy = NewFloat32Like(dem), NewFloat32Like(dem)
            oracle.process(terrain.gradientOp(terrain.GradientOptions(CellSize=0.125)), [dx, dy], [dem])
            z = rows.astype(np.float64)
            for i in range(n):
                for got, a, b, d in ((dx.Data[n + 2 + i + 1], z[0, i + 2] - z[0, i], z[2, i + 2] - z[2, i],
                                      z[1, i + 2] - z[1, i]),
                                     (dy.Data[n + 2 + i + 1], z[2, i] - z[0, i], z[2, i + 2] - z[0, i + 2],
                                      z[2, i + 1] - z[0, i + 1])):
                    want = a + b + 2 * d
                    s = abs(a) + abs(b) + 2 * abs(d)
                    tol = 0 if s == 0 else 2.0 ** (math.frexp(s)[1] - 24)
                    assert abs(float(got) - want) <= tol
