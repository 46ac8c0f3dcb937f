"""terrain: Gradient, Slope, Aspect, Hillshade and Curvature in the
reference's three forms — plain, Tiled and Chunked — over the HIP library.

Option resolution follows the reference exactly: constants are computed in
float64 and rounded once to float32, as Go does, and handed to the device
in the kernel descriptor:
  resolveCells / cellSizes / HornScales   terrain/stencil.go:14-47,
                                          internal/stencil/horn.go:114-116
  slopeScale                              terrain/slope.go:84-94
  aspectFlat                              terrain/aspect.go:97-102
  hillshadeLight                          terrain/hillshade.go:94-117
  ZTScales / newCurvatureKernel           internal/stencil/curvature.go:29-33,
                                          terrain/curvature.go:113-140
  newWoodFit                              terrain/fit.go:36-70
Edges are NaN and invalid; an output cell is valid iff its 3×3 (or fit)
window is valid (terrain/doc.go).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from enum import IntEnum
from typing import List, Optional

import numpy as np

from . import _lib
from .engine import (Background, Context, EnginePanic, FanOut, KernelDesc, Options, ProcessChunked, ProcessN,
                     RasterSink, RasterSource)
from .raster import Float32Raster

MaxRadius = 8
AspectFlat = -1.0


class TerrainPanic(EnginePanic):
    pass


def _f32(x: float) -> float:
    """float32(x) as Go rounds it: overflow gives ±Inf, which the usable
    checks then reject."""
    with np.errstate(over="ignore"):
        return float(np.float32(x))


def _usable(k: float) -> bool:
    return k != 0 and not math.isinf(k)


def resolveCells(cellSize: float, cellSizeY: float, zFactor: float):
    if not (cellSize > 0) or math.isinf(cellSize):
        raise TerrainPanic(f"terrain: CellSize must be positive and finite, got {cellSize}")
    if cellSizeY == 0:
        cellSizeY = cellSize
    if not (cellSizeY > 0) or math.isinf(cellSizeY):
        raise TerrainPanic(f"terrain: CellSizeY must be positive and finite (or 0 for CellSize), got {cellSizeY}")
    if zFactor == 0:
        zFactor = 1.0
    if math.isnan(zFactor) or math.isinf(zFactor):
        raise TerrainPanic(f"terrain: ZFactor must be finite, got {zFactor}")
    return float(cellSize), float(cellSizeY), float(zFactor)


def HornScales(cx: float, cy: float, z: float):
    return _f32(z / (8 * cx)), _f32(z / (8 * cy))


def cellSizes(cellSize: float, cellSizeY: float, zFactor: float):
    cx, cy, z = resolveCells(cellSize, cellSizeY, zFactor)
    kx, ky = HornScales(cx, cy, z)
    if not _usable(kx) or not _usable(ky):
        raise TerrainPanic(
            "terrain: ZFactor/(8·CellSize) and ZFactor/(8·CellSizeY) must be finite and non-zero as float32, "
            f"got {kx} and {ky} for CellSize {cellSize}, CellSizeY {cellSizeY} and ZFactor {zFactor}")
    return kx, ky


class SlopeUnits(IntEnum):
    SlopeDegrees = 0
    SlopePercent = 1
    SlopeRadians = 2


SlopeDegrees, SlopePercent, SlopeRadians = SlopeUnits.SlopeDegrees, SlopeUnits.SlopePercent, SlopeUnits.SlopeRadians


def slopeScale(u):
    if u == SlopeDegrees:
        return _f32(180 / math.pi), True
    if u == SlopeRadians:
        return 1.0, True
    if u == SlopePercent:
        return 100.0, False
    raise TerrainPanic(f"terrain: unknown SlopeUnits {int(u)}")


def hillshadeLight(azimuth: float, altitude: float):
    az, alt = float(azimuth), float(altitude)
    if math.isnan(az) or math.isinf(az):
        raise TerrainPanic(f"terrain: Azimuth must be finite, got {az}")
    if az == 0:
        az = 315.0
    if alt == 0:
        alt = 45.0
    if not (0 < alt <= 90):
        raise TerrainPanic(f"terrain: Altitude must be in (0, 90] (or 0 for 45), got {altitude}")
    az = math.fmod(az, 360)
    az, alt = az * math.pi / 180, alt * math.pi / 180
    c = _f32(255 * math.sin(alt))
    bx = _f32(-255 * math.cos(alt) * math.sin(az))
    by = _f32(255 * math.cos(alt) * math.cos(az))
    return c, bx, by


def ZTScales(cx: float, cy: float, z: float):
    return (_f32(z / (2 * cx)), _f32(z / (2 * cy)), _f32(z / (cx * cx)), _f32(z / (cy * cy)),
            _f32(z / (4 * cx * cy)))


def woodFit(fitRadius: int, cellSize: float, cellSizeY: float, zFactor: float):
    """Wood's quadratic fit factors kp, kq, kr, kt, ks (terrain/fit.go:36-70)."""
    if fitRadius < 1 or fitRadius > MaxRadius:
        raise TerrainPanic(f"terrain: FitRadius must be 0 to {MaxRadius}, got {fitRadius}")
    cx, cy, z = resolveCells(cellSize, cellSizeY, zFactor)
    r = fitRadius
    q = 0.0
    for i in range(-r, r + 1):
        w = 3 * i * i - r * (r + 1)
        q += float(w * w)
    k = float(2 * r + 1)
    s = float(r * (r + 1) * (2 * r + 1)) / 3
    kp = _f32(z / (k * s * cx))
    kq = _f32(z / (k * s * cy))
    kr = _f32(6 * z / (k * q * cx * cx))
    kt = _f32(6 * z / (k * q * cy * cy))
    ks = _f32(z / (s * s * cx * cy))
    for v in (kp, kq, kr, kt, ks):
        if not _usable(v):
            raise TerrainPanic(
                "terrain: the quadratic fit's factors ZFactor/(K·S·CellSize), 6·ZFactor/(K·Q·CellSize²) and "
                "ZFactor/(S²·CellSize·CellSizeY), and the same for CellSizeY, must be finite and non-zero as "
                f"float32, got {kp}, {kq}, {kr}, {kt} and {ks} for CellSize {cellSize}, CellSizeY {cellSizeY}, "
                f"ZFactor {zFactor} and FitRadius {fitRadius}")
    return kp, kq, kr, kt, ks


def _check_fit(fitRadius: int) -> None:
    if fitRadius < 0 or fitRadius > MaxRadius:
        raise TerrainPanic(f"terrain: FitRadius must be 0 to {MaxRadius}, got {fitRadius}")


# -- options -------------------------------------------------------------------

@dataclass
class GradientOptions:
    CellSize: float = 0.0
    CellSizeY: float = 0.0
    ZFactor: float = 0.0
    FitRadius: int = 0


@dataclass
class SlopeOptions:
    CellSize: float = 0.0
    CellSizeY: float = 0.0
    ZFactor: float = 0.0
    FitRadius: int = 0
    Units: SlopeUnits = SlopeDegrees


@dataclass
class AspectOptions:
    CellSize: float = 0.0
    CellSizeY: float = 0.0
    ZFactor: float = 0.0
    FitRadius: int = 0
    ZeroForFlat: bool = False
    Trigonometric: bool = False


@dataclass
class HillshadeOptions:
    CellSize: float = 0.0
    CellSizeY: float = 0.0
    ZFactor: float = 0.0
    FitRadius: int = 0
    Azimuth: float = 0.0
    Altitude: float = 0.0


class CurvatureType(IntEnum):
    CurvatureProfile = 0
    CurvaturePlan = 1
    CurvatureMean = 2


CurvatureProfile, CurvaturePlan, CurvatureMean = (CurvatureType.CurvatureProfile, CurvatureType.CurvaturePlan,
                                                  CurvatureType.CurvatureMean)


@dataclass
class CurvatureOptions:
    CellSize: float = 0.0
    CellSizeY: float = 0.0
    ZFactor: float = 0.0
    FitRadius: int = 0
    Type: CurvatureType = CurvatureProfile


# -- kernel constructors (terrain/fit.go:187-226 gradientOp..curvatureOp) -------------

def _fit_desc(op: int, o, product_k=(), flags: int = 0, nout: int = 1, name: str = "") -> KernelDesc:
    f = woodFit(o.FitRadius, o.CellSize, o.CellSizeY, o.ZFactor)
    k = list(f) + list(product_k)
    return KernelDesc(op, o.FitRadius, flags | _lib.FLAG_FIT, k, None, 1, nout, name)


def gradientOp(o: GradientOptions) -> KernelDesc:
    _check_fit(o.FitRadius)
    if o.FitRadius:
        return _fit_desc(_lib.OP_GRADIENT, o, nout=2, name="terrain.Gradient")
    kx, ky = cellSizes(o.CellSize, o.CellSizeY, o.ZFactor)
    return KernelDesc(_lib.OP_GRADIENT, 1, 0, [kx, ky], None, 1, 2, "terrain.Gradient")


def slopeOp(o: SlopeOptions) -> KernelDesc:
    _check_fit(o.FitRadius)
    if o.FitRadius:
        woodFit(o.FitRadius, o.CellSize, o.CellSizeY, o.ZFactor)
        scale, atan = slopeScale(o.Units)
        return _fit_desc(_lib.OP_SLOPE, o, [scale], _lib.FLAG_ATAN if atan else 0, name="terrain.Slope")
    kx, ky = cellSizes(o.CellSize, o.CellSizeY, o.ZFactor)
    scale, atan = slopeScale(o.Units)
    return KernelDesc(_lib.OP_SLOPE, 1, _lib.FLAG_ATAN if atan else 0, [kx, ky, scale], None, 1, 1, "terrain.Slope")


def aspectFlat(zeroForFlat: bool) -> float:
    return 0.0 if zeroForFlat else AspectFlat


def aspectOp(o: AspectOptions) -> KernelDesc:
    _check_fit(o.FitRadius)
    flags = _lib.FLAG_TRIG if o.Trigonometric else 0
    if o.FitRadius:
        return _fit_desc(_lib.OP_ASPECT, o, [aspectFlat(o.ZeroForFlat)], flags, name="terrain.Aspect")
    kx, ky = cellSizes(o.CellSize, o.CellSizeY, o.ZFactor)
    return KernelDesc(_lib.OP_ASPECT, 1, flags, [kx, ky, aspectFlat(o.ZeroForFlat)], None, 1, 1, "terrain.Aspect")


def hillshadeOp(o: HillshadeOptions) -> KernelDesc:
    _check_fit(o.FitRadius)
    if o.FitRadius:
        woodFit(o.FitRadius, o.CellSize, o.CellSizeY, o.ZFactor)
        return _fit_desc(_lib.OP_HILLSHADE, o, list(hillshadeLight(o.Azimuth, o.Altitude)),
                         name="terrain.Hillshade")
    kx, ky = cellSizes(o.CellSize, o.CellSizeY, o.ZFactor)
    c, bx, by = hillshadeLight(o.Azimuth, o.Altitude)
    return KernelDesc(_lib.OP_HILLSHADE, 1, 0, [kx, ky, c, bx, by], None, 1, 1, "terrain.Hillshade")


def curvatureKind(t) -> int:
    if t in (CurvatureProfile, CurvaturePlan, CurvatureMean):
        return int(t)
    raise TerrainPanic(f"terrain: unknown CurvatureType {int(t)}")


def curvatureOp(o: CurvatureOptions) -> KernelDesc:
    _check_fit(o.FitRadius)
    if o.FitRadius:
        woodFit(o.FitRadius, o.CellSize, o.CellSizeY, o.ZFactor)
        kind = curvatureKind(o.Type)
        return _fit_desc(_lib.OP_CURVATURE, o, (), kind << _lib.CURV_SHIFT, name="terrain.Curvature")
    cx, cy, z = resolveCells(o.CellSize, o.CellSizeY, o.ZFactor)
    kind = curvatureKind(o.Type)
    ks = ZTScales(cx, cy, z)
    for v in ks:
        if not _usable(v):
            raise TerrainPanic(
                "terrain: ZFactor/(2·CellSize), ZFactor/CellSize² and ZFactor/(4·CellSize·CellSizeY), and the "
                "same for CellSizeY, must be finite and non-zero as float32, got "
                f"{ks[0]}, {ks[1]}, {ks[2]}, {ks[3]} and {ks[4]} for CellSize {o.CellSize}, "
                f"CellSizeY {o.CellSizeY} and ZFactor {o.ZFactor}")
    return KernelDesc(_lib.OP_CURVATURE, 1, kind << _lib.CURV_SHIFT, list(ks), None, 1, 1, "terrain.Curvature")


# -- the three forms of each operation --------------------------------------------------

def _run(k: KernelDesc, dem: Float32Raster, *outs: Float32Raster) -> None:
    # The plain functions: one call over whole rasters (terrain/stencil.go:56-60).
    ProcessN(None, list(outs), [dem], k, Options(Workers=1))


def _run_tiled(ctx, eopts, k, dem, *outs):
    return ProcessN(ctx, list(outs), [dem], k, eopts)


def _run_chunked(ctx, eopts, k, dem: RasterSource, *outs: RasterSink):
    return ProcessChunked(ctx, list(outs), [dem], k, eopts)


def Gradient(dx, dy, dem, opts: GradientOptions) -> None:
    _run(gradientOp(opts), dem, dx, dy)


def GradientTiled(ctx, dx, dy, dem, opts: GradientOptions, eopts: Options):
    return _run_tiled(ctx, eopts, gradientOp(opts), dem, dx, dy)


def GradientChunked(ctx, dx, dy, dem, opts: GradientOptions, eopts: Options):
    return _run_chunked(ctx, eopts, gradientOp(opts), dem, dx, dy)


def Slope(dst, dem, opts: SlopeOptions) -> None:
    _run(slopeOp(opts), dem, dst)


def SlopeTiled(ctx, dst, dem, opts: SlopeOptions, eopts: Options):
    return _run_tiled(ctx, eopts, slopeOp(opts), dem, dst)


def SlopeChunked(ctx, dst, dem, opts: SlopeOptions, eopts: Options):
    return _run_chunked(ctx, eopts, slopeOp(opts), dem, dst)


def Aspect(dst, dem, opts: AspectOptions) -> None:
    _run(aspectOp(opts), dem, dst)


def AspectTiled(ctx, dst, dem, opts: AspectOptions, eopts: Options):
    return _run_tiled(ctx, eopts, aspectOp(opts), dem, dst)


def AspectChunked(ctx, dst, dem, opts: AspectOptions, eopts: Options):
    return _run_chunked(ctx, eopts, aspectOp(opts), dem, dst)


def Hillshade(dst, dem, opts: HillshadeOptions) -> None:
    _run(hillshadeOp(opts), dem, dst)


def HillshadeTiled(ctx, dst, dem, opts: HillshadeOptions, eopts: Options):
    return _run_tiled(ctx, eopts, hillshadeOp(opts), dem, dst)


def HillshadeChunked(ctx, dst, dem, opts: HillshadeOptions, eopts: Options):
    return _run_chunked(ctx, eopts, hillshadeOp(opts), dem, dst)


def Curvature(dst, dem, opts: CurvatureOptions) -> None:
    _run(curvatureOp(opts), dem, dst)


def CurvatureTiled(ctx, dst, dem, opts: CurvatureOptions, eopts: Options):
    return _run_tiled(ctx, eopts, curvatureOp(opts), dem, dst)


def CurvatureChunked(ctx, dst, dem, opts: CurvatureOptions, eopts: Options):
    return _run_chunked(ctx, eopts, curvatureOp(opts), dem, dst)


# Name -> kernel constructor, as the reference's opkernel registry
# (terrain/opkernel.go:14-43).
OPKERNELS = {
    "terrain.Gradient": gradientOp,
    "terrain.Slope": slopeOp,
    "terrain.Aspect": aspectOp,
    "terrain.Hillshade": hillshadeOp,
    "terrain.Curvature": curvatureOp,
}


# -- Surface: several products from one read and one gradient ------------------------------

@dataclass
class SurfaceOptions:
    """terrain.SurfaceOptions (terrain/surface.go:12-32)."""
    CellSize: float = 0.0
    CellSizeY: float = 0.0
    ZFactor: float = 0.0
    FitRadius: int = 0
    Units: int = SlopeDegrees
    ZeroForFlat: bool = False
    Trigonometric: bool = False
    Azimuth: float = 0.0
    Altitude: float = 0.0


@dataclass
class SurfaceOutputs:
    """terrain.SurfaceOutputs (terrain/surface.go:34-44): None is a product
    not wanted; at least one must be set."""
    Dx: Optional[Float32Raster] = None
    Dy: Optional[Float32Raster] = None
    Slope: Optional[Float32Raster] = None
    Aspect: Optional[Float32Raster] = None
    Hillshade: Optional[Float32Raster] = None


@dataclass
class SurfaceSinks:
    """terrain.SurfaceSinks (terrain/surface.go:46-52) for SurfaceChunked."""
    Dx: Optional[RasterSink] = None
    Dy: Optional[RasterSink] = None
    Slope: Optional[RasterSink] = None
    Aspect: Optional[RasterSink] = None
    Hillshade: Optional[RasterSink] = None


_SURF_BITS = (_lib.SURF_DX, _lib.SURF_DY, _lib.SURF_SLOPE, _lib.SURF_ASPECT, _lib.SURF_HILLSHADE)


def newSurface(want, o: SurfaceOptions) -> KernelDesc:
    """The kernel of terrain.newSurface (terrain/surface.go:104-146): the
    wanted products of one gradient, in SurfaceOutputs order. Every
    product's options are checked, wanted or not, as the reference does."""
    _check_fit(o.FitRadius)
    scale, atan = slopeScale(o.Units)
    flat = aspectFlat(o.ZeroForFlat)
    c, bx, by = hillshadeLight(o.Azimuth, o.Altitude)
    flags = 0
    for w, bit in zip(want, _SURF_BITS):
        if w:
            flags |= bit
    if not flags:
        raise TerrainPanic("terrain: Surface has no outputs; set at least one of Dx, Dy, Slope, Aspect and Hillshade")
    if atan:
        flags |= _lib.FLAG_ATAN
    if o.Trigonometric:
        flags |= _lib.FLAG_TRIG
    nout = sum(1 for w in want if w)
    if o.FitRadius:
        f = woodFit(o.FitRadius, o.CellSize, o.CellSizeY, o.ZFactor)
        return KernelDesc(_lib.OP_SURFACE, o.FitRadius, flags | _lib.FLAG_FIT, list(f) + [scale, flat, c, bx, by],
                          None, 1, nout, "terrain.Surface")
    kx, ky = cellSizes(o.CellSize, o.CellSizeY, o.ZFactor)
    return KernelDesc(_lib.OP_SURFACE, 1, flags, [kx, ky, scale, flat, c, bx, by], None, 1, nout,
                      "terrain.Surface")


def _surface_outs(out):
    rasters = [out.Dx, out.Dy, out.Slope, out.Aspect, out.Hillshade]
    return [r is not None for r in rasters], [r for r in rasters if r is not None]


def Surface(out: SurfaceOutputs, dem, opts: SurfaceOptions) -> None:
    """terrain.Surface (terrain/surface.go:53-68): any of dem's gradient,
    slope, aspect and hillshade in one pass, each bit for bit what the
    standalone op writes."""
    want, outs = _surface_outs(out)
    _run(newSurface(want, opts), dem, *outs)


def SurfaceTiled(ctx, out: SurfaceOutputs, dem, opts: SurfaceOptions, eopts: Options):
    want, outs = _surface_outs(out)
    return _run_tiled(ctx, eopts, newSurface(want, opts), dem, *outs)


def SurfaceChunked(ctx, out: SurfaceSinks, dem, opts: SurfaceOptions, eopts: Options):
    want, outs = _surface_outs(out)
    return _run_chunked(ctx, eopts, newSurface(want, opts), dem, *outs)


# -- Features: a feature stack, each output its own op, from one read ------------------------

@dataclass
class Feature:
    """terrain.Feature (terrain/features.go:29-34): an operation — one of
    SlopeOptions, AspectOptions, HillshadeOptions and CurvatureOptions,
    each meaning what it means to its own function — and the raster its
    result goes to."""
    Op: object
    Dst: Float32Raster


@dataclass
class FeatureSink:
    """terrain.FeatureSink (terrain/features.go:36-40) for FeaturesChunked."""
    Op: object
    Dst: RasterSink


def _feature_kernel(i: int, op) -> KernelDesc:
    # FeatureOp.featureKernel (terrain/features.go:21-27). HeatLoad,
    # Orientation and Ruggedness are operations this library does not run.
    if op is None:
        raise TerrainPanic(f"terrain: Features output {i} has no Op")
    if isinstance(op, SlopeOptions):
        return slopeOp(op)
    if isinstance(op, AspectOptions):
        return aspectOp(op)
    if isinstance(op, HillshadeOptions):
        return hillshadeOp(op)
    if isinstance(op, CurvatureOptions):
        return curvatureOp(op)
    raise TerrainPanic(f"terrain: Features output {i} has Op {type(op).__name__}, which is not a FeatureOp this "
                       "library runs (Slope, Aspect, Hillshade and Curvature options are)")


def newFeatures(ops) -> FanOut:
    """terrain.newFeatures (terrain/features.go:103-117): the pipeline that
    fans the DEM out to one stage per feature; output i is stage i."""
    if len(ops) == 0:
        raise TerrainPanic("terrain: Features has no outputs")
    return FanOut([_feature_kernel(i, op) for i, op in enumerate(ops)], "terrain.Features")


def Features(out: List[Feature], dem) -> None:
    """terrain.Features (terrain/features.go:42-71): several terrain
    features of one DEM in one pass over it, each bit for bit what its own
    function writes with the same options: its own edge ring and its
    validity eroded by its own window."""
    FeaturesTiled(None, out, dem, Options(Workers=1))


def FeaturesTiled(ctx, out: List[Feature], dem, eopts: Options):
    """terrain.FeaturesTiled (terrain/features.go:73-86)."""
    k = newFeatures([f.Op for f in out])
    return ProcessN(ctx, [f.Dst for f in out], [dem], k, eopts)


def FeaturesChunked(ctx, out: List[FeatureSink], dem: RasterSource, eopts: Options):
    """terrain.FeaturesChunked (terrain/features.go:88-101): tiles read
    grown by the largest radius asked for."""
    k = newFeatures([f.Op for f in out])
    return ProcessChunked(ctx, [f.Dst for f in out], [dem], k, eopts)
