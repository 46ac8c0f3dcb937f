"""Concurrent calls from many host threads. The reference's entry points may
be called from many goroutines at once, and tiling and workers never change
a result (internal/exec/kernel.go:14-19, engine/doc.go:32-41). Here eight
threads, half of them on streams of their own, make a shuffled mix of plain
and Chunked calls over device and host operands, and each result must equal
the same call made alone. ctypes
releases the GIL around each library call, so the calls overlap inside the
library."""
import threading

import numpy as np
import pytest
import torch

from strata_amd import focal, terrain
from strata_amd.engine import MemorySink, MemorySource, Options, ProcessChunked, ProcessN
from strata_amd.raster import Float32Raster, NewMask, pack_bits

pytestmark = pytest.mark.gpu

W, H = 700, 523
WHERE = ("device", "host", "chunked-device", "chunked-host")


def jobs():
    return [
        ("slope", terrain.slopeOp(terrain.SlopeOptions(CellSize=30)), 1),
        ("gradient", terrain.gradientOp(terrain.GradientOptions(CellSize=10, CellSizeY=12)), 2),
        ("curvature", terrain.curvatureOp(terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureMean)), 1),
        ("aspect-fit2", terrain.aspectOp(terrain.AspectOptions(CellSize=30, FitRadius=2)), 1),
        ("mean3", focal.newBox(focal.BoxOptions(3), "mean"), 1),
        ("correlate2", focal.newWeightsKernel(focal.WeightsOptions(2, list(np.linspace(-1, 1, 25))), False), 1),
        ("mfma4", focal.newWeightsKernel(focal.WeightsOptions(4, list(np.linspace(-1, 1, 81))), False, mfma=True),
         1),
    ]


def run(k, nout, data, valid, where):
    """k over (data, valid) through one path; results as host arrays."""
    n = W * H
    if where.endswith("device"):
        dev = torch.device("cuda:0")
        src = Float32Raster(torch.from_numpy(data).to(dev), W, H, W, torch.from_numpy(valid.view(np.int64)).to(dev), 0)
        outs = [Float32Raster(torch.zeros(n, dtype=torch.float32, device=dev), W, H, W, NewMask(n, device=dev), 0)
                for _ in range(nout)]
    else:
        src = Float32Raster(data.copy(), W, H, W, valid.copy(), 0)
        outs = [Float32Raster(np.zeros(n, np.float32), W, H, W, NewMask(n), 0) for _ in range(nout)]
    if where.startswith("chunked"):
        assert ProcessChunked(None, [MemorySink(o) for o in outs], [MemorySource(src)], k,
                              Options(TileHeight=97)) is None
    else:
        assert ProcessN(None, outs, [src], k, Options()) is None
    res = []
    for o in outs:
        d, v = o.Data, o.Valid
        if hasattr(d, "cpu"):
            torch.cuda.synchronize()
            d, v = d.cpu().numpy(), v.cpu().numpy()
        res.append((np.asarray(d).copy(), np.asarray(v).view(np.uint64).copy()))
    return res


def same(a, b):
    return all(np.array_equal(x[0].view(np.uint32), y[0].view(np.uint32)) and np.array_equal(x[1], y[1])
               for x, y in zip(a, b))


def test_threads_match_calls_made_alone():
    rng = np.random.default_rng(2024)
    x = np.arange(W, dtype=np.float64)[None, :]
    y = np.arange(H, dtype=np.float64)[:, None]
    data = (700 + 90 * np.sin(x / 31) * np.cos(y / 17) + rng.standard_normal((H, W)) * 5).astype(np.float32)
    data = data.reshape(-1)
    valid = pack_bits(rng.random(W * H) >= 0.01)
    js = jobs()
    items = [(j, w) for j in range(len(js)) for w in WHERE]
    alone = {(j, w): run(js[j][1], js[j][2], data, valid, w) for j, w in items}
    # Every path gives the same result for one job.
    for j in range(len(js)):
        for w in WHERE[1:]:
            assert same(alone[(j, w)], alone[(j, WHERE[0])]), (js[j][0], w)

    failures, errors = [], []

    def calls(seed):
        order = np.random.default_rng(seed)
        for _ in range(3):
            for i in order.permutation(len(items)):
                j, w = items[i]
                if not same(run(js[j][1], js[j][2], data, valid, w), alone[(j, w)]):
                    failures.append((seed, js[j][0], w))

    def worker(seed):
        # Odd threads share the default stream; even ones each launch on a
        # stream of their own (default priority), where device operands are
        # also made and read back.
        try:
            if seed % 2:
                calls(seed)
            else:
                with torch.cuda.stream(torch.cuda.Stream(device=torch.device("cuda:0"))):
                    calls(seed)
        except Exception as e:  # reported below; a thread's exception does not fail the test by itself
            errors.append((seed, repr(e)))

    threads = [threading.Thread(target=worker, args=(s,)) for s in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors[:4]
    assert not failures, failures[:8]
