"""Batches against the oracle, bit for bit.

A batch runs many tasks (a stencil op on its own tile, then point ops on
the output in place) in one persistent launch. Each task must give exactly
what the separate calls give: the oracle runs every task as those calls
(the stencil op, then each point op over the output), and the test
compares whole parent buffers, so writes outside a task's output (its row
padding, a parent's other cells and mask bits) also fail it.
"""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import focal, terrain, transfer
from strata_amd.batch import Batch, Task
from strata_amd.engine import EnginePanic

from helpers import Scene, assert_same_parents, host_copy

pytestmark = pytest.mark.gpu


def stencil_ops():
    ops = [
        ("slope", terrain.slopeOp(terrain.SlopeOptions(CellSize=30))),
        ("slope-pct", terrain.slopeOp(terrain.SlopeOptions(CellSize=10, Units=terrain.SlopePercent))),
        ("slope-rad", terrain.slopeOp(terrain.SlopeOptions(CellSize=2, CellSizeY=3, Units=terrain.SlopeRadians))),
        ("aspect", terrain.aspectOp(terrain.AspectOptions(CellSize=3, CellSizeY=4))),
        ("hillshade", terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30))),
    ]
    for t in (terrain.CurvatureProfile, terrain.CurvaturePlan, terrain.CurvatureMean):
        ops.append((f"curv{int(t)}", terrain.curvatureOp(terrain.CurvatureOptions(CellSize=10, Type=t))))
    for r in (1, 2, 3):
        for op in ("mean", "min", "max"):
            ops.append((f"{op}{r}", focal.newBox(focal.BoxOptions(r), op)))
    return ops


LOOKUP = transfer.newLookupKernel("t", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])
RECLASS = transfer.newReclassKernel("t", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5])
CHAINS = [(), (LOOKUP,), (RECLASS,), (LOOKUP, RECLASS)]


def build(seed: int, n: int, focal_only=None):
    """A scene of n tasks: windows of two source parents (one masked at a
    mid-word bit base, one unmasked), each task writing a window of its own
    output parent. focal_only: the one focal op (by name) among the ops."""
    rng = np.random.default_rng(seed)
    ops = stencil_ops()
    if focal_only is not None:
        ops = [(nm, k) for nm, k in ops if k.radius == 1 and nm[:4] in ("slop", "aspe", "hill", "curv")
               or nm == focal_only]
    sc = Scene()
    SW, SH = 1300, 900
    sc.add_parent("srcm", rng, SW, SH, masked=True, valid_base=29)
    sc.add_parent("srcu", rng, SW + 3, SH, masked=False)
    shapes = [(1, 1), (1, 40), (40, 1), (2, 2), (3, 3), (513, 7), (64, 33), (700, 500), (1024, 2)]
    tasks = []
    for i in range(n):
        if i < len(shapes):
            w, h = shapes[i]
        else:
            w, h = (int(v) for v in np.exp(rng.uniform(np.log(8), np.log(640), 2)))
        name, k = ops[i % len(ops)]
        chain = CHAINS[i % len(CHAINS)] if k.radius == 1 and name[:4] in ("slop", "aspe", "hill", "curv") else ()
        masked_src = i % 3 != 2
        parent = "srcm" if masked_src else "srcu"
        x = int(rng.integers(0, SW - w + 1))
        y = int(rng.integers(0, SH - h + 1))
        sc.add_window(f"s{i}", parent, x, y, w, h)
        masked_dst = masked_src or i % 2 == 0
        dw = w + int(rng.integers(0, 9))
        sc.add_parent(f"dst{i}", rng, dw, h + 2, masked=masked_dst, valid_base=int(rng.integers(0, 64)))
        sc.add_window(f"d{i}", f"dst{i}", dw - w, 1, w, h)
        tasks.append((name, k, chain))
    return sc, tasks


def expected(sc, tasks):
    host, bufs = sc.materialize(None)
    for i, (_, k, chain) in enumerate(tasks):
        oracle.process(k, [host[f"d{i}"]], [host[f"s{i}"]])
        for p in chain:
            oracle.process(p, [host[f"d{i}"]], [host[f"d{i}"]])
    return bufs


def device_tasks(sc, tasks, gpu):
    import torch
    dev, dbufs = sc.materialize(gpu)
    torch.cuda.synchronize()  # the batch runs on a stream of its own
    return [Task(k, dev[f"s{i}"], dev[f"d{i}"], chain) for i, (_, k, chain) in enumerate(tasks)], dbufs


def reset(dbufs, sc):
    """Put every output parent back to its stale starting contents."""
    import torch
    for name, p in sc.parents.items():
        if name.startswith("dst"):
            dbufs[name].data.copy_(torch.from_numpy(p.data))
            if p.valid is not None:
                dbufs[name].valid.copy_(torch.from_numpy(p.valid.view(np.int64)))


@pytest.mark.parametrize("seed,n", [(1, 40), (2, 97)])
def test_batch_matches_separate_calls(gpu, seed, n):
    import torch
    sc, tasks = build(seed, n)
    want = expected(sc, tasks)
    dt, dbufs = device_tasks(sc, tasks, gpu)
    with Batch(dt) as b:
        b.run()
        b.sync()
        torch.cuda.synchronize()
        assert_same_parents(host_copy(dbufs), want, f"batch seed {seed}")
        info = b.info()
        assert info["tasks"] == n
        assert info["cells"] == sum(t.src.Width * t.src.Height for t in dt)


@pytest.mark.parametrize("focal", ["mean3", "min1", "max2"])
def test_batch_one_focal_kind_matches_separate_calls(gpu, focal):
    """Every focal task of one kind and radius, beside the terrain ops: the
    batch runs its focal list through the kernel built for that op alone."""
    import torch
    sc, tasks = build(7, 36, focal_only=focal)
    assert {nm for nm, k, _ in tasks if nm not in [o[0] for o in stencil_ops()[:8]]} == {focal}
    want = expected(sc, tasks)
    dt, dbufs = device_tasks(sc, tasks, gpu)
    with Batch(dt) as b:
        b.run()
        b.sync()
        torch.cuda.synchronize()
        assert_same_parents(host_copy(dbufs), want, f"batch of {focal}")


@pytest.mark.parametrize("rows,bpc", [(4, 1), (16, 2), (128, 0), (1000, 0)])
def test_batch_shapes_do_not_change_results(gpu, rows, bpc):
    import torch
    sc, tasks = build(3, 30)
    want = expected(sc, tasks)
    dt, dbufs = device_tasks(sc, tasks, gpu)
    with Batch(dt, rows=rows, blocks_per_cu=bpc) as b:
        b.run()
        b.sync()
        torch.cuda.synchronize()
        assert_same_parents(host_copy(dbufs), want, f"rows {rows} blocks/CU {bpc}")


def test_batch_graph_replays(gpu):
    """The captured graph, replayed with eager device work (resets) and
    synchronisations between replays; every replay is checked."""
    import torch
    sc, tasks = build(4, 50)
    want = expected(sc, tasks)
    dt, dbufs = device_tasks(sc, tasks, gpu)
    with Batch(dt, graph=True) as b:
        assert b.info()["graph"]
        for rep in range(3):
            reset(dbufs, sc)
            torch.cuda.synchronize()
            b.run()
            b.sync()
            torch.cuda.synchronize()
            assert_same_parents(host_copy(dbufs), want, f"graph replay {rep}")


def test_batch_tune_and_per_tile(gpu):
    import torch
    sc, tasks = build(5, 45)
    want = expected(sc, tasks)
    dt, dbufs = device_tasks(sc, tasks, gpu)
    with Batch(dt, graph=True) as b:
        chosen = b.tune(reps=1)
        assert chosen["rows"] in (16, 32, 64, 128) and chosen["graph"]
        reset(dbufs, sc)
        torch.cuda.synchronize()
        b.run()
        b.sync()
        torch.cuda.synchronize()
        assert_same_parents(host_copy(dbufs), want, "tuned")
        reset(dbufs, sc)
        torch.cuda.synchronize()
        b.run_per_tile()
        b.sync()
        torch.cuda.synchronize()
        assert_same_parents(host_copy(dbufs), want, "per tile")


def test_batch_rejects(gpu):
    sc, tasks = build(6, 3)
    dt, _ = device_tasks(sc, tasks, gpu)
    w = np.ones(9, np.float32)
    bad = Task(focal.newWeightsKernel(focal.WeightsOptions(1, w), False), dt[0].src, dt[0].dst)
    with pytest.raises(EnginePanic, match="does not run in a batch"):
        Batch([bad])
    host, _ = sc.materialize(None)
    with pytest.raises(EnginePanic, match="device memory"):
        Batch([Task(tasks[0][1], host["s0"], host["d0"])])
    fit = terrain.slopeOp(terrain.SlopeOptions(CellSize=30, FitRadius=2))
    with pytest.raises(EnginePanic, match="does not run in a batch"):
        Batch([Task(fit, dt[0].src, dt[0].dst)])
