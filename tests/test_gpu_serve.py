"""The request loop: every result equals the chain run on its window as a
raster of its own (the oracle: Slope, then Lookup and Reclass over the
output in place), whether the request runs alone or among others, under
budgets and slot counts that force queueing and admission waits."""
import numpy as np
import pytest

from oracle import oracle
from strata_amd import terrain, transfer
from strata_amd.engine import EnginePanic
from strata_amd.raster import Float32Raster
from strata_amd.serve import Server, poisson_schedule

from helpers import Scene, same_floats

pytestmark = pytest.mark.gpu

SLOPE = terrain.slopeOp(terrain.SlopeOptions(CellSize=30))
LOOKUP = transfer.newLookupKernel("t", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])
RECLASS = transfer.newReclassKernel("t", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5])
W, H = 1500, 1100


def scene(seed=11):
    rng = np.random.default_rng(seed)
    sc = Scene()
    # Terrain-like values (steep enough that every class occurs) plus hazards.
    n = (W + 3) * H
    yy, xx = np.divmod(np.arange(n), W + 3)
    base = (np.sin(xx / 37.0) * 300 + np.cos(yy / 23.0) * 250 + rng.standard_normal(n) * 8).astype(np.float32)
    hz = rng.random(n) < 0.002
    base[hz] = np.nan
    sc.add_parent("dem", rng, W + 3, H, masked=True, valid_base=37, values=base, invalid_frac=0.01)
    sc.add_window("src", "dem", 0, 0, W, H)
    return sc


def expected(host_src: Float32Raster, x, y, w, h):
    start = y * host_src.Stride + x
    win = Float32Raster(host_src.Data[start:start + (h - 1) * host_src.Stride + w], w, h, host_src.Stride,
                        host_src.Valid, host_src.ValidOffset + start)
    dst = Float32Raster(np.zeros(w * h, np.float32), w, h, w, np.zeros((w * h + 63) // 64 + 1, np.uint64), 0)
    oracle.process(SLOPE, [dst], [win])
    oracle.process(LOOKUP, [dst], [dst])
    oracle.process(RECLASS, [dst], [dst])
    return dst


def assert_result(srv, rid, host_src, win, what):
    x, y, w, h = win
    data, mask = srv.fetch(rid, w, h)
    want = expected(host_src, x, y, w, h)
    eq = same_floats(data, want.Data)
    assert eq.all(), f"{what}: window {win}: {int((~eq).sum())} cells differ, first at {int(np.argmin(eq))}"
    bits_got = np.unpackbits(mask.view(np.uint8), bitorder="little")[: w * h]
    bits_want = np.unpackbits(want.Valid.view(np.uint8), bitorder="little")[: w * h]
    assert np.array_equal(bits_got, bits_want), f"{what}: window {win}: mask differs"


@pytest.fixture(scope="module")
def operands(gpu):
    import torch
    sc = scene()
    host, _ = sc.materialize(None)
    dev, bufs = sc.materialize(gpu)
    torch.cuda.synchronize()
    return host["src"], dev["src"], bufs


WINDOWS = [(0, 0, W, H), (0, 0, 1, 1), (5, 7, 2, 3), (W - 3, H - 3, 3, 3), (100, 200, 512, 512),
           (1, 1, 700, 500), (333, 17, 1, 600), (17, 333, 600, 1), (900, 600, 600, 500), (64, 64, 513, 63)]


def test_alone_then_in_mix(operands):
    host_src, dev_src, _ = operands
    with Server(dev_src, SLOPE, (LOOKUP, RECLASS), slots=3, max_batch=16) as srv:
        for win in WINDOWS[:4]:
            rid = srv.submit(*win)
            assert srv.wait(rid) >= 0.0
            assert_result(srv, rid, host_src, win, "alone")
            srv.release(rid)
        rng = np.random.default_rng(3)
        mix = list(WINDOWS)
        for _ in range(60):
            w, h = (int(v) for v in rng.integers(1, 400, 2))
            mix.append((int(rng.integers(0, W - w + 1)), int(rng.integers(0, H - h + 1)), w, h))
        ids = [srv.submit(*win) for win in mix]
        for rid in ids:
            srv.wait(rid)
        for rid, win in zip(ids, mix):
            assert_result(srv, rid, host_src, win, "in mix")
            srv.release(rid)
        st = srv.stats()
        assert st["requests"] == 4 + len(mix)
        assert st["batches"] >= 1


def test_tight_budget_queues_in_order(operands):
    """A budget that holds about one large result at a time and two slots:
    requests wait for memory, still in arrival order, and results are
    unchanged."""
    host_src, dev_src, _ = operands
    big = WINDOWS[5]
    budget = 700 * 500 * 4 + 64 * 1024 + 700 * 500 // 8 + 4096
    mix = [big, WINDOWS[4], big, WINDOWS[2], WINDOWS[8], big, WINDOWS[1]]
    with Server(dev_src, SLOPE, (LOOKUP, RECLASS), slots=2, max_batch=4, hbm_budget=budget) as srv:
        ids = []
        for win in mix:
            rid = srv.submit(*win)
            ids.append(rid)
        # Release each result once checked, so the queue can drain.
        for rid, win in zip(ids, mix):
            srv.wait(rid)
            assert_result(srv, rid, host_src, win, "tight budget")
            srv.release(rid)
        st = srv.stats()
        assert st["admission_waits"] >= 1
        assert st["peak_bytes"] <= budget


def test_replay_poisson(operands):
    _, dev_src, _ = operands
    windows, arrivals = poisson_schedule(2000.0, 80, W, H, seed=5, mix=((0.8, 128), (0.15, 512), (0.05, 1024)))
    with Server(dev_src, SLOPE, (LOOKUP, RECLASS), slots=4) as srv:
        lat = srv.replay(windows, arrivals)
        assert lat.shape == (80,)
        assert np.all(np.isfinite(lat)) and np.all(lat > 0)
        st = srv.stats()
        assert st["requests"] == 80
        assert st["cells"] == int((windows[:, 2] * windows[:, 3]).sum())


def test_rejects(operands):
    _, dev_src, _ = operands
    with Server(dev_src, SLOPE, (LOOKUP, RECLASS), hbm_budget=1 << 20) as srv:
        with pytest.raises(EnginePanic, match="outside"):
            srv.submit(W - 10, 0, 20, 5)
        with pytest.raises(EnginePanic, match="budget"):
            srv.submit(0, 0, 1000, 1000)
    with pytest.raises(EnginePanic):
        Server(dev_src, terrain.gradientOp(terrain.GradientOptions(CellSize=30)))
