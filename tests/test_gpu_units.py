"""strata_process over host operands and with a live context.

- Host, device and mixed operands give the oracle's bits; a cancellable
  call runs as units of the plan claimed in order (four in flight on the
  call's stream), here at small units that split rows between units and
  share mask words across them.
- A context done before the call writes nothing, host operands included
  (internal/exec/doc.go:165-178: "A context that is done before the call
  writes nothing").
- Cancelled mid-run, the cells and bits written are exactly those of a
  prefix of the plan's bands, everything else is untouched, and Stats are
  that prefix's (internal/exec/worker.go:144-181, doc.go:165-183). A
  watcher thread reads a device output cell of an early unit from its own
  stream and cancels as soon as it sees the cell written, so the cancel
  lands mid-run; Gradient's second output in host memory shows that a host
  output keeps the same prefix. The same holds, deterministically, when the
  library's test knob stops a call after a set number of units, and when
  another thread cancels a call while it runs (whether it lands mid-run or
  after the end).
"""
import contextlib
import threading
import time

import numpy as np
import pytest

from oracle import oracle
from strata_amd import _lib, engine, focal, terrain, transfer
from strata_amd.engine import Context, Options, ProcessN

from helpers import assert_same_parents, host_copy, same_floats
from test_gpu_parity import build_scene
from test_host import band_stats, plan_bands

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def unit_kcells(n):
    L = _lib.lib()
    t = _lib.Tuning()
    assert L.strata_get_tuning(t) == 0
    old = t.unit_kcells
    t.unit_kcells = n
    assert L.strata_set_tuning(t) == 0, _lib.last_error()
    try:
        yield
    finally:
        t.unit_kcells = old
        assert L.strata_set_tuning(t) == 0


@contextlib.contextmanager
def stop_after_units(n):
    L = _lib.lib()
    t = _lib.Tuning()
    assert L.strata_get_tuning(t) == 0
    old = t.stop_after_units
    t.stop_after_units = n
    assert L.strata_set_tuning(t) == 0, _lib.last_error()
    try:
        yield
    finally:
        t.stop_after_units = old
        assert L.strata_set_tuning(t) == 0


def kernels():
    rng = np.random.default_rng(3)
    w = rng.uniform(-1, 1, 25).astype(np.float32)
    g = focal.Gaussian(8, 4.0)
    return {
        "slope": terrain.slopeOp(terrain.SlopeOptions(CellSize=30)),
        "gradient": terrain.gradientOp(terrain.GradientOptions(CellSize=30)),
        "aspect": terrain.aspectOp(terrain.AspectOptions(CellSize=3, CellSizeY=4)),
        "curv-mean": terrain.curvatureOp(terrain.CurvatureOptions(CellSize=10, Type=terrain.CurvatureMean)),
        "slope-fit3": terrain.slopeOp(terrain.SlopeOptions(CellSize=30, FitRadius=3)),
        "mean3": focal.newBox(focal.BoxOptions(3), "mean"),
        "max1": focal.newBox(focal.BoxOptions(1), "max"),
        "correlate2": focal.newWeightsKernel(focal.WeightsOptions(2, w), False),
        "gauss8": focal.newSeparableKernel(focal.SeparableOptions(8, g, g)),
        "reclass": transfer.newReclassKernel("t", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5]),
    }


K = kernels()


def mixed(sc, gpu, src_where, dst_where, nout):
    """Operands with the source on one side and the outputs on another;
    returns the rasters and the parent buffers they live in."""
    host, hbufs = sc.materialize(None)
    dev, dbufs = sc.materialize(gpu)
    src = (host if src_where == "host" else dev)["s"]
    dsts, bufs = [], {}
    for o in range(nout):
        side = host if dst_where[o] == "host" else dev
        dsts.append(side[f"d{o}"])
        bufs[f"dst{o}"] = (hbufs if dst_where[o] == "host" else dbufs)[f"dst{o}"]
    return src, dsts, bufs


@pytest.mark.parametrize("where", ["host-host", "device-host", "host-device"])
@pytest.mark.parametrize("kc", [0, 1])
@pytest.mark.parametrize("kind", ["masked", "windows", "unmasked-in", "compact"])
@pytest.mark.parametrize("name", list(K))
def test_host_operands_match_oracle(gpu, name, kind, kc, where):
    k = K[name]
    W, H = (301, 71) if k.radius <= 3 else (157, 83)
    sc = build_scene(kind, k, hash((name, kind)) & 0xFFFF, W, H)
    host, want = sc.materialize(None)
    oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
    sw, dw = where.split("-")
    src, dsts, bufs = mixed(sc, gpu, sw, [dw] * k.nout, k.nout)
    st = engine.Stats()
    with unit_kcells(kc):
        for tw, th in ((0, 0), (7, 5), (64, 1)):
            err = ProcessN(Context() if tw else None, dsts, [src], k, Options(TileWidth=tw, TileHeight=th, Stats=st))
            assert err is None
    import torch
    torch.cuda.synchronize()
    got = host_copy(bufs)
    assert_same_parents(got, {n: want[n] for n in bufs}, f"{name} {kind} unit {kc}K {where}")
    one = engine.PlanStats(k, W, H, Options())
    assert st.Cells == 3 * one.Cells and st.KernelWritten == 3 * one.KernelWritten


@pytest.mark.parametrize("where", ["host-host", "device-host", "host-device"])
def test_cancelled_before_start_writes_nothing(gpu, where):
    k = K["slope"]
    sc = build_scene("windows", k, 5, 200, 50)
    sw, dw = where.split("-")
    src, dsts, bufs = mixed(sc, gpu, sw, [dw], 1)
    before = host_copy(bufs)
    ctx = Context()
    ctx.cancel()
    st = engine.Stats()
    err = ProcessN(ctx, dsts, [src], k, Options(TileHeight=8, Stats=st))
    assert isinstance(err, engine.Canceled)
    assert_same_parents(host_copy(bufs), before, f"cancelled before start, {where}")
    assert st == engine.Stats()


def expected_prefix(sc, k, W, H, tw, th, nbands, before, want):
    """The parents after a call that finished the first nbands bands: the
    oracle's cells and bits inside them, the old ones everywhere else."""
    out = {n: type(p)(p.data.copy(), None if p.valid is None else p.valid.copy()) for n, p in before.items()}
    bands = plan_bands(W, H, tw, th)[:nbands]
    for o in range(k.nout):
        win = sc.wins[f"d{o}"]
        name = win.parent
        if name not in out:
            continue
        for (x0, y0, x1, y1) in bands:
            ys, xs = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
            a = ((win.y + ys) * win.width + win.x + xs).ravel()
            out[name].data[a] = want[name].data[a]
            if out[name].valid is not None:
                bits = (win.valid_base + a).astype(np.int64)
                sh = (bits & 63).astype(np.uint64)
                m = np.uint64(1) << sh
                wv = ((want[name].valid[bits >> 6] >> sh) & np.uint64(1)) << sh
                cur = out[name].valid
                np.bitwise_and.at(cur, bits >> 6, ~m)
                np.bitwise_or.at(cur, bits >> 6, wv)
    return out


def prefix_stats(k, W, H, tw, th, nbands):
    s = engine.Stats()
    for b in plan_bands(W, H, tw, th)[:nbands]:
        band_stats(s, k.radius, W, H, k.nin, k.nout, b)
    s.Ideal = s.Cells * 4 * (k.nin + k.nout)
    return s


@pytest.mark.parametrize("plan", [(0, 0), (300, 100), (0, 37)])
@pytest.mark.parametrize("where", ["device", "mixed-device-src", "mixed-host-src"])
def test_cancel_mid_run_leaves_band_prefix(gpu, plan, where):
    """"device": Slope with every operand in device memory; "mixed-*":
    Gradient with dx in device memory (the watched output) and dy in host
    memory, the source on the device or the host."""
    import torch
    k = K["slope"] if where == "device" else K["gradient"]
    W, H = 4000, 4000
    tw, th = plan
    sc = build_scene("windows", k, 17, W, H)
    host, want = sc.materialize(None)
    oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
    sw = "host" if where == "mixed-host-src" else "device"
    dw = ["device"] if where == "device" else ["device", "host"]
    src, dsts, bufs = mixed(sc, gpu, sw, dw, k.nout)
    before = host_copy(bufs)
    watched = dsts[0].Data  # a device output
    win = sc.wins["d0"]
    # A cell the call changes, in the second band group or tile row.
    y = 150
    a = (win.y + y) * win.width + win.x
    cols = np.nonzero(~same_floats(before[win.parent].data[a:a + W], want[win.parent].data[a:a + W]))[0]
    x = int(cols[len(cols) // 2])
    idx = y * dsts[0].Stride + x
    old_bits = before[win.parent].data[a + x:a + x + 1].view(np.uint32)[0]
    ctx = Context()
    stop = threading.Event()

    def watch():
        side = torch.cuda.Stream(device=gpu)
        with torch.cuda.stream(side):
            while not stop.is_set():
                v = watched[idx:idx + 1].to("cpu").numpy()
                if v.view(np.uint32)[0] != old_bits:
                    ctx.cancel()
                    return
                time.sleep(0.0002)

    th_w = threading.Thread(target=watch)
    st = engine.Stats()
    with unit_kcells(16):
        th_w.start()
        try:
            err = ProcessN(ctx, dsts, [src], k, Options(TileWidth=tw, TileHeight=th, Stats=st))
        finally:
            stop.set()
            th_w.join()
    torch.cuda.synchronize()
    total = len(plan_bands(W, H, tw, th))
    print(f"cancel {where} plan {plan}: {st.Bands} of {total} bands finished", flush=True)
    assert isinstance(err, engine.Canceled)
    assert 0 < st.Bands < total
    assert st == prefix_stats(k, W, H, tw, th, st.Bands)
    exp = expected_prefix(sc, k, W, H, tw, th, st.Bands, before, want)
    assert_same_parents(host_copy(bufs), exp, f"cancelled {where} {plan}")


@pytest.mark.parametrize("plan", [(0, 0), (256, 256), (1000, 3)])
def test_cancellable_device_call_in_many_units(gpu, plan):
    """Device operands with a live context run as units on the call's
    stream; at units of 16384 cells a 1531x1207 window is hundreds of
    launches, each unit's bits the oracle's."""
    for name in ("slope", "mean3", "correlate2"):
        k = K[name]
        sc = build_scene("windows", k, 23, 1531, 1207)
        host, want = sc.materialize(None)
        oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
        dev, dbufs = sc.materialize(gpu)
        st = engine.Stats()
        with unit_kcells(16):
            err = ProcessN(Context(), [dev[f"d{o}"] for o in range(k.nout)], [dev["s"]], k,
                           Options(TileWidth=plan[0], TileHeight=plan[1], Stats=st))
        assert err is None
        assert_same_parents(host_copy(dbufs), want, f"{name} in units, plan {plan}")
        assert st == engine.PlanStats(k, 1531, 1207, Options(TileWidth=plan[0], TileHeight=plan[1]))


@pytest.mark.parametrize("plan", [(0, 0), (300, 100), (0, 37)])
@pytest.mark.parametrize("where", ["device", "mixed-host-src"])
def test_stop_after_units_leaves_band_prefix(gpu, plan, where):
    """The call stops after three units (the library's test knob, the same
    path as a flag seen set before the fourth claim): exactly the bands of
    those units are written and counted."""
    import torch
    k = K["slope"] if where == "device" else K["gradient"]
    W, H = 1500, 1300
    tw, th = plan
    sc = build_scene("windows", k, 19, W, H)
    host, want = sc.materialize(None)
    oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
    sw = "host" if where == "mixed-host-src" else "device"
    dw = ["device"] if where == "device" else ["device", "host"]
    src, dsts, bufs = mixed(sc, gpu, sw, dw, k.nout)
    before = host_copy(bufs)
    st = engine.Stats()
    with unit_kcells(16), stop_after_units(3):
        err = ProcessN(Context(), dsts, [src], k, Options(TileWidth=tw, TileHeight=th, Stats=st))
    torch.cuda.synchronize()
    total = len(plan_bands(W, H, tw, th))
    assert isinstance(err, engine.Canceled)
    assert 0 < st.Bands < total
    assert st == prefix_stats(k, W, H, tw, th, st.Bands)
    exp = expected_prefix(sc, k, W, H, tw, th, st.Bands, before, want)
    assert_same_parents(host_copy(bufs), exp, f"stopped after three units, {where} {plan}")


@pytest.mark.parametrize("name", ["slope", "mean3"])
def test_cancel_from_another_thread(gpu, name):
    """A call cancelled from another thread while it runs either finishes
    first (every band written) or ends cancelled with exactly a prefix of
    the plan's bands written, the rest untouched, and that prefix's Stats.
    Both are the reference's contract; which one happens depends on
    timing."""
    import torch
    k = K[name]
    W, H = 3001, 2003
    sc = build_scene("windows", k, 29, W, H)
    host, want = sc.materialize(None)
    oracle.process(k, [host[f"d{o}"] for o in range(k.nout)], [host["s"]])
    src, dsts, bufs = mixed(sc, gpu, "device", ["device"], k.nout)
    before = host_copy(bufs)
    ctx = Context()
    started = threading.Event()

    def cancel_soon():
        started.wait()
        ctx.cancel()

    th_c = threading.Thread(target=cancel_soon)
    th_c.start()
    st = engine.Stats()
    with unit_kcells(64):
        started.set()
        err = ProcessN(ctx, dsts, [src], k, Options(Stats=st))
    th_c.join()
    torch.cuda.synchronize()
    total = len(plan_bands(W, H, 0, 0))
    print(f"{name}: {'cancelled' if err else 'finished'} with {st.Bands} of {total} bands", flush=True)
    if err is None:
        assert st.Bands == total
    else:
        assert isinstance(err, engine.Canceled)
        assert 0 <= st.Bands < total
    assert st == prefix_stats(k, W, H, 0, 0, st.Bands)
    exp = expected_prefix(sc, k, W, H, 0, 0, st.Bands, before, want)
    assert_same_parents(host_copy(bufs), exp, f"{name} cancelled from a thread")
