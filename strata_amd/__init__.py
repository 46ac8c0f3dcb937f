"""strata_amd — MI355X-native (gfx950) raster stencil path of strata.

Terrain Gradient/Slope/Aspect/Hillshade/Curvature, focal Correlate/
Convolve/CorrelateSeparable/Mean/Min/Max and transfer Lookup/Reclass with
validity-bitmap NoData, in the reference's plain, Tiled and Chunked forms,
computed by hand-written HIP kernels behind the C ABI in
include/strata_hip.h (libstrata_hip.so, built in-tree by build.py).
"""
from . import _lib, engine, focal, raster, terrain, transfer  # noqa: F401
from .engine import (Background, Canceled, Context, EnginePanic, MemorySink, MemorySource, Options,  # noqa: F401
                     ProcessChunked, ProcessN, Stats, WithCancel)
from .raster import Float32Raster, NewFloat32, NewFloat32Like, NewFloat32Stride, NewMask  # noqa: F401


def library_path() -> str:
    return _lib.LIB_PATH


def version() -> str:
    return _lib.lib().strata_version().decode()
