"""Chains through strata_process / strata_process_chunked against the
oracle's separate calls, bit for bit.

exec.NewPipeline over a chain (internal/exec/pipeline.go:166-187, what
graph.Slope -> graph.Lookup -> graph.Reclass lowers to, graph/graph.go:
250-287) computes what its stages compute as separate calls, each on the
one before: the output's edge ring and validity erosion are the chain's
reach (the sum of the radii), and every op here carries NaN through, so
the ring holds NaN either way (pipeline.go:83-114). The cases cover the
serving chain (a radius-1 stage with its point stages applied in
registers), stencils after stencils (staged through device scratch), a
point stage first, fit and focal stages followed by point stages, windows
at mid-word offsets, Tiled, cancellable and Chunked plans. Stats are one
kernel's of the chain's radius.
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import engine, focal, terrain, transfer
from strata_amd.engine import Context, MemorySink, MemorySource, Options, Pipeline, ProcessChunked, ProcessN
from strata_amd.raster import Float32Raster

from helpers import assert_same_parents, host_copy
from test_gpu_parity import build_scene

pytestmark = pytest.mark.gpu

SLOPE = terrain.slopeOp(terrain.SlopeOptions(CellSize=30))
LOOKUP = transfer.newLookupKernel("risk", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])
RECLASS = transfer.newReclassKernel("risk", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5])
HILL = terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30))
RNG = np.random.default_rng(11)

CHAINS = {
    "slope-lookup-reclass": [SLOPE, LOOKUP, RECLASS],
    "slope-lookup": [SLOPE, LOOKUP],
    "hillshade-reclass-lookup": [HILL, transfer.newReclassKernel("t", [100, 200], [0, 10, 40]),
                                 transfer.newLookupKernel("t", [0, 40], [-1, 1])],
    "aspect-reclass": [terrain.aspectOp(terrain.AspectOptions(CellSize=30)),
                       transfer.newReclassKernel("t", [90, 180, 270], [1, 2, 3, 4])],
    "curvature-lookup": [terrain.curvatureOp(terrain.CurvatureOptions(CellSize=10, Type=terrain.CurvatureMean)),
                         transfer.newLookupKernel("t", [-0.01, 0, 0.01], [0, 1, 2])],
    "mean2-slope": [focal.newBox(focal.BoxOptions(2), "mean"), SLOPE],
    "slope-max1": [SLOPE, focal.newBox(focal.BoxOptions(1), "max")],
    "lookup-slope": [transfer.newLookupKernel("t", [0, 1000], [0, 2000]), SLOPE],
    "correlate1-reclass": [focal.newWeightsKernel(focal.WeightsOptions(1, RNG.uniform(-1, 1, 9).astype(np.float32)),
                                                  False),
                           transfer.newReclassKernel("t", [0], [0, 1])],
    "slopefit2-lookup": [terrain.slopeOp(terrain.SlopeOptions(CellSize=30, FitRadius=2)), LOOKUP],
    "max1-mean2-curvature": [focal.newBox(focal.BoxOptions(1), "max"), focal.newBox(focal.BoxOptions(2), "mean"),
                             terrain.curvatureOp(terrain.CurvatureOptions(CellSize=10))],
    "slope-lookup-reclass-lookup-reclass": [SLOPE, LOOKUP, RECLASS, transfer.newLookupKernel("t", [1, 5], [5, 1]),
                                            transfer.newReclassKernel("t", [3], [0, 1])],
}


def oracle_chain(stages, src, out):
    """The stages as separate calls, each on the one before, the last into
    out. The intermediates are compact and masked when the output is: a
    pipeline's output validity is its inputs' eroded by the reach, and with
    no masked input the interior of the reach (internal/exec/pipeline.go:
    83-100, doc.go:76-108), which is what masked intermediates carry."""
    cur = src
    W, H = src.Width, src.Height
    for i, k in enumerate(stages):
        if i == len(stages) - 1:
            nxt = out
        else:
            valid = None if out.Valid is None else np.zeros((W * H + 63) // 64 + 1, np.uint64)
            nxt = Float32Raster(np.zeros(W * H, np.float32), W, H, W, valid, 0)
        oracle.process(k, [nxt], [cur])
        cur = nxt


def case(name, kind, seed, W=301, H=97):
    stages = CHAINS[name]
    pipe = Pipeline(stages)
    sc = build_scene(kind, pipe, seed, W, H)
    host, want = sc.materialize(None)
    oracle_chain(stages, host["s"], host["d0"])
    return pipe, sc, want


@pytest.mark.parametrize("kind", ["masked", "windows", "unmasked-in", "compact"])
@pytest.mark.parametrize("name", list(CHAINS))
def test_chain_equals_separate_calls(gpu, name, kind):
    pipe, sc, want = case(name, kind, hash((name, kind)) & 0xFFFF)
    for tw, th, live in ((0, 0, False), (7, 5, True), (64, 1, False), (0, 16, True)):
        dev, dbufs = sc.materialize(gpu)
        st = engine.Stats()
        err = ProcessN(Context() if live else None, [dev["d0"]], [dev["s"]], pipe,
                       Options(TileWidth=tw, TileHeight=th, Stats=st))
        assert err is None
        assert_same_parents(host_copy(dbufs), want, f"{name} {kind} plan {tw}x{th}")
        assert st == engine.PlanStats(pipe, 301, 97, Options(TileWidth=tw, TileHeight=th))


@pytest.mark.parametrize("resident", [True, False])
@pytest.mark.parametrize("name", ["slope-lookup-reclass", "mean2-slope", "max1-mean2-curvature", "lookup-slope"])
def test_chain_chunked(gpu, name, resident):
    pipe, sc, want = case(name, "windows", 5)
    for tw, th in ((0, 16), (50, 7), (0, 0)):
        dev, dbufs = sc.materialize(gpu if resident else None)
        st = engine.Stats()
        err = ProcessChunked(Context(), [MemorySink(dev["d0"])], [MemorySource(dev["s"])], pipe,
                             Options(TileWidth=tw, TileHeight=th, Stats=st))
        assert err is None
        assert_same_parents(host_copy(dbufs), want, f"chunked {name} {tw}x{th} resident={resident}")
        assert st == engine.PlanStats(pipe, 301, 97, Options(TileWidth=tw, TileHeight=th), chunked=True)


def test_serving_chain_on_a_large_window(gpu):
    """The serving chain over a 2048x1500 window of a wider masked parent,
    with host operands (staged) and device operands."""
    pipe, sc, want = case("slope-lookup-reclass", "windows", 9, 2048, 1500)
    for where in (gpu, None):
        r, bufs = sc.materialize(where)
        assert ProcessN(None, [r["d0"]], [r["s"]], pipe, Options()) is None
        assert_same_parents(host_copy(bufs), want, f"serving chain 2048x1500 on {'device' if where else 'host'}")
