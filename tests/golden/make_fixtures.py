"""Write the golden fixtures: seeded hazard-laden inputs (NaN, ±Inf, ±0,
±MaxFloat32, subnormals, random NoData, a window at a mid-word bit offset)
and the oracle's outputs for every op of the path, as small .npz files.

The reference ships no output goldens (its testdata holds fuzz inputs
only), so these are regression vectors of this build's oracle, which is
itself pinned by the reference's known-answer tests. Run from the repo
root: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle  # noqa: E402
from strata_amd import focal, terrain, transfer  # noqa: E402
from strata_amd.raster import Float32Raster, pack_bits  # noqa: E402

W, H = 45, 31           # the window's size
PW, PH = 53, 36          # the parent buffer holding it
WX, WY, BASE = 5, 3, 29  # the window's position and the parent's first bit


def kernels():
    rng = np.random.default_rng(1234)
    w3 = rng.uniform(-1, 1, 49).astype(np.float32)
    return {
        "gradient": terrain.gradientOp(terrain.GradientOptions(CellSize=30)),
        "slope_deg": terrain.slopeOp(terrain.SlopeOptions(CellSize=30)),
        "slope_pct": terrain.slopeOp(terrain.SlopeOptions(CellSize=10, CellSizeY=7, ZFactor=2,
                                                          Units=terrain.SlopePercent)),
        "aspect": terrain.aspectOp(terrain.AspectOptions(CellSize=30)),
        "aspect_trig": terrain.aspectOp(terrain.AspectOptions(CellSize=30, Trigonometric=True, ZeroForFlat=True)),
        "hillshade": terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30)),
        "curv_profile": terrain.curvatureOp(terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureProfile)),
        "curv_plan": terrain.curvatureOp(terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvaturePlan)),
        "curv_mean": terrain.curvatureOp(terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureMean)),
        "fit2_slope": terrain.slopeOp(terrain.SlopeOptions(CellSize=30, FitRadius=2)),
        "fit3_curv": terrain.curvatureOp(terrain.CurvatureOptions(CellSize=30, FitRadius=3,
                                                                  Type=terrain.CurvatureMean)),
        "correlate3": focal.newWeightsKernel(focal.WeightsOptions(3, w3), False),
        "convolve3": focal.newWeightsKernel(focal.WeightsOptions(3, w3), True),
        "gauss4": focal.newSeparableKernel(focal.SeparableOptions(4, focal.Gaussian(4, 2), focal.Gaussian(4, 2))),
        "mean3": focal.newBox(focal.BoxOptions(3), "mean"),
        "min2": focal.newBox(focal.BoxOptions(2), "min"),
        "max5": focal.newBox(focal.BoxOptions(5), "max"),
        "lookup": transfer.newLookupKernel("t", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5]),
        "reclass": transfer.newReclassKernel("t", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5]),
    }


def inputs():
    from helpers import hazard_values
    rng = np.random.default_rng(20261203)
    data = hazard_values(rng, PW * PH, base=20, relief=15, hazard_frac=0.03)
    valid = pack_bits(rng.random(BASE + PW * PH + 64) >= 0.04)
    return data, valid


def window(data, valid):
    start = WY * PW + WX
    return Float32Raster(data[start:start + (H - 1) * PW + W], W, H, PW, valid, BASE + start)


def compute(k, data, valid):
    src = window(data, valid)
    outs = []
    for _ in range(k.nout):
        outs.append(Float32Raster(np.zeros(W * H, np.float32), W, H, W, np.zeros((W * H + 63) // 64, np.uint64), 0))
    oracle.process(k, outs, [src])
    return outs


def main():
    data, valid = inputs()
    np.savez(os.path.join(HERE, "inputs.npz"), data=data, valid=valid,
             geom=np.array([W, H, PW, PH, WX, WY, BASE], np.int64))
    for name, k in kernels().items():
        outs = compute(k, data, valid)
        arrays = {}
        for i, o in enumerate(outs):
            arrays[f"data{i}"] = o.Data
            arrays[f"valid{i}"] = o.Valid
        np.savez(os.path.join(HERE, f"{name}.npz"), **arrays)
        print(name)


if __name__ == "__main__":
    main()
