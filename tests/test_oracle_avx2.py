"""The CPU baseline (the reference's AVX2 Tiled path restated in the oracle,
oracle/strata_oracle.c orc_tiled_avx2) against the per-cell oracle, bit
for bit: data (any NaN matches any NaN, +0 != -0) and the whole parent
buffers around the outputs, mask words included.

The reference holds its AVX2 lanes to its scalar rows the same way
(internal/stencil/simd_test.go:36-300, terrain/terrain_test.go:271-342).
Widths cross the 8-cell lane boundary and the 64-bit mask words; windows
sit at odd offsets with mid-word validity; plans vary the tile shape so
bands share mask words across threads; hazards (NaN, infinities, signed
zeros, extremes, subnormals) reach every lane position.
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import terrain

from helpers import assert_same_parents, host_copy
from test_gpu_parity import build_scene

KERNELS = {
    "slope-deg": terrain.slopeOp(terrain.SlopeOptions(CellSize=30)),
    "slope-pct": terrain.slopeOp(terrain.SlopeOptions(CellSize=3, Units=terrain.SlopePercent)),
    "slope-rad": terrain.slopeOp(terrain.SlopeOptions(CellSize=1, CellSizeY=2, Units=terrain.SlopeRadians)),
    "hillshade": terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30)),
    "hillshade-az": terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=2, CellSizeY=7, ZFactor=3,
                                                                  Azimuth=-117.5, Altitude=23)),
    "gradient": terrain.gradientOp(terrain.GradientOptions(CellSize=1)),
}


def run_both(k, kind, seed, W, H, tw, th, threads):
    sc = build_scene(kind, k, seed, W, H)
    host, want = sc.materialize(None)
    oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
    got_r, got = sc.materialize(None)
    oracle.tiled_avx2(k, [got_r[f"d{o}"] for o in range(k.nout)], [got_r["s"]], tw, th, threads)
    return host_copy(got), want


@pytest.mark.parametrize("name", list(KERNELS))
@pytest.mark.parametrize("W", [1, 2, 3, 9, 10, 11, 17, 64, 65, 66, 131])
@pytest.mark.parametrize("kind", ["masked", "windows", "unmasked-in", "compact"])
def test_avx2_rows_equal_oracle(name, W, kind):
    k = KERNELS[name]
    got, want = run_both(k, kind, W * 31 + len(kind), W, 13, 0, 0, 1)
    assert_same_parents(got, want, f"{name} {kind} W={W}")


@pytest.mark.parametrize("tw,th", [(0, 0), (7, 5), (64, 1), (300, 20), (1, 0), (0, 3)])
@pytest.mark.parametrize("threads", [1, 4])
@pytest.mark.parametrize("name", ["slope-deg", "hillshade", "gradient"])
def test_avx2_plans_and_threads_equal_oracle(name, tw, th, threads):
    k = KERNELS[name]
    got, want = run_both(k, "windows", tw * 7 + th, 301, 97, tw, th, threads)
    assert_same_parents(got, want, f"{name} plan {tw}x{th} threads {threads}")


def test_avx2_wide_rows_many_bands():
    """Bands of max(1, 65536/W) rows: at W = 3000 each band is 21 rows and
    eight threads claim them; the rows' ends are not word-aligned."""
    k = KERNELS["slope-deg"]
    got, want = run_both(k, "windows", 5, 3000, 200, 0, 0, 8)
    assert_same_parents(got, want, "slope 3000x200, 8 threads")


def test_avx2_refuses_ops_it_does_not_cover():
    k = terrain.aspectOp(terrain.AspectOptions(CellSize=30))
    sc = build_scene("masked", k, 1, 20, 20)
    r, _ = sc.materialize(None)
    with pytest.raises(ValueError):
        oracle.tiled_avx2(k, [r["d0"]], [r["s"]])
