#!/usr/bin/env python3
"""Benchmark of the raster stencil path on one MI355X.

Workload (the headline): terrain.SlopeChunked over a device-resident
20000×20000 float32 DEM — seeded diamond-square fBm (Hurst 0.8) scaled to
0..3000 m, 0.5% NoData in circular blobs of radius 5..50 — Horn 3×3,
CellSize 30, degrees, engine.Options{TileHeight: 256}. A step is one Slope
over the whole raster (data and validity kernels). `value` is Gcells/s
over the timed steps, inputs already resident in HBM.

Also reported: the roofline of the dominant kernel (the Slope data
kernel) from HIP events the library records around each of its launches
on the launch stream, the rest of the terrain suite (Aspect, Hillshade,
profile/plan/mean Curvature, Gradient) on the same DEM, a batch of 4096
uneven mixed-op tiles run as one persistent graph-captured launch against
one launch set per tile, a serving sweep (Poisson arrivals at fixed QPS,
p50/p99 latency of the chain Slope -> Lookup -> Reclass on window
requests), and the CPU baseline: the reference's AVX2 Tiled path restated
in C (oracle/strata_oracle.c), timed on up to 16 host cores over the whole
DEM for Slope and Hillshade and on one core for the reference's plain
4096x4096 Slope, whose outputs are compared bit for bit with the GPU's.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Gcells/s + % HBM peak (Slope 20000²); p50/p99 tile latency at fixed QPS"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=20000, help="DEM width and height")
    ap.add_argument("--tile-height", type=int, default=256)
    ap.add_argument("--suite", choices=["terrain", "slope"], default="terrain",
                    help="terrain: Slope then the rest of the suite; slope: Slope only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU work (thread-seconds) to aim for in each timed baseline op")
    ap.add_argument("--sweep-rows", type=int, default=0,
                    help="rows per block of the radius-1 kernels (0: the library's default)")
    ap.add_argument("--batch-tiles", type=int, default=4096,
                    help="tiles in the batched-launch comparison (0: skip it)")
    ap.add_argument("--batch-reps", type=int, default=10, help="timed runs of the batch")
    ap.add_argument("--serve-seconds", type=float, default=1.0,
                    help="arrival time per QPS level of the serving sweep (0: skip it)")
    ap.add_argument("--serve-qps", type=str, default="100,200,500,1000,2000,5000,10000,20000,50000,100000",
                    help="QPS levels of the serving sweep")
    ap.add_argument("--serve-burst", type=int, default=2048,
                    help="requests of the closed-loop serving run (all arriving at once) and of the per-tile "
                         "comparison on the same request mix (0: skip both)")
    ap.add_argument("--serve-rows", type=int, default=0,
                    help="rows per sweep unit when serving (0: chosen per request; short units run a "
                         "small request's rows in parallel)")
    ap.add_argument("--focal-size", type=int, default=16384,
                    help="side of the focal-suite raster (0: skip the focal suite)")
    ap.add_argument("--focal-steps", type=int, default=5, help="timed runs of each focal op")
    ap.add_argument("--reduce-steps", type=int, default=10,
                    help="timed runs of each reduction over the DEM (0: skip)")
    ap.add_argument("--raw-size", type=int, default=20000,
                    help="side of the file-to-file SlopeChunked run (raw float32 files; 0: skip)")
    ap.add_argument("--no-paths", dest="paths", action="store_false",
                    help="skip the cancellable, chain and host-memory Slope runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one device copy of the DEM (for HBM counter calibration)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure the Slope kernel's HBM traffic")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-op", default="slope",
                    choices=("slope", "mean3", "correlate1", "gauss2", "correlate8-mfma", "correlate4-mfma"),
                    help="with --pmc-child: the op run twice before the calibration copy (focal ops over the "
                         "--focal-size raster of the focal workload)")
    return ap.parse_args()


def main() -> int:
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) is not supported. "
              "The path shards by row strips with a radius-r halo exchange; see DESIGN.md.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from strata_amd import _lib, synth, terrain
    from strata_amd.engine import MemorySink, MemorySource, Options, ProcessChunked, Stats
    from strata_amd.raster import Float32Raster

    if args.pmc_child:
        return pmc_child(args)

    dev = torch.device("cuda:0")
    L = _lib.lib()
    assert L.strata_init(0) == 0, _lib.last_error()
    tuning = _lib.Tuning()
    L.strata_get_tuning(tuning)
    if args.sweep_rows:
        tuning.sweep_rows = args.sweep_rows
        assert L.strata_set_tuning(tuning) == 0, _lib.last_error()
    N = args.size
    cells = N * N

    t0 = time.time()
    dem_data = synth.fbm_dem(N, N, hurst=0.8, seed=1, lo=0.0, hi=3000.0, device=dev)
    valid, covered = synth.blob_nodata(N, N, frac=0.005, rmin=5, rmax=50, seed=2, device=dev)
    torch.cuda.synchronize()
    log(f"bench: DEM {N}x{N} generated in {time.time() - t0:.1f}s, NoData {covered * 100:.3f}%")

    dem = Float32Raster(dem_data, N, N, N, valid, 0)
    out = Float32Raster(torch.empty(cells, dtype=torch.float32, device=dev), N, N, N,
                        torch.zeros_like(valid), 0)
    src, sink = MemorySource(dem), MemorySink(out)
    eopts = Options(TileHeight=args.tile_height)

    def run_op(k, sinks):
        err = ProcessChunked(None, sinks, [src], k, eopts, async_=True)
        assert err is None

    def time_op(k, sinks, steps, warmup):
        """The steps timed as a caller runs them; then the kernels' own
        durations over as many steps again, with the library's timing events
        around each launch (which add a marker between launches, so they
        stay out of the timed steps)."""
        for _ in range(warmup):
            run_op(k, sinks)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            run_op(k, sinks)
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        timing = _lib.KernelTiming().begin()
        for _ in range(steps):
            run_op(k, sinks)
        torch.cuda.synchronize()
        timing.end()
        return el, timing

    slope = terrain.slopeOp(terrain.SlopeOptions(CellSize=30, Units=terrain.SlopeDegrees))
    el, timing = time_op(slope, [sink], args.steps, args.warmup)
    ms_step = el * 1e3 / args.steps
    value = cells * args.steps / el / 1e9
    kern_ms = timing.data_ms / max(1, timing.data_launches)
    vkern_ms = timing.validity_ms / max(1, timing.validity_launches)
    # The dominant kernel: the sweep, which also writes the validity when no
    # validity launch of its own was recorded (4 B read + 4 B written per
    # cell, plus 1/8 B of mask in and out when it writes the validity).
    fused = timing.validity_launches == 0
    data_bytes = (8.25 if fused else 8) * cells
    kernel_name = "horn_sweep_valid<SLOPE> (data and validity)" if fused else "horn_sweep<SLOPE> (data kernel)"
    achieved = data_bytes / (kern_ms * 1e-3) / 1e9
    op_bytes = 8.25 * cells  # plus 1/8 B mask in and out
    log(f"bench: Slope {ms_step:.3f} ms/step, {value:.1f} Gcells/s; {kernel_name} {kern_ms:.3f} ms "
        f"({achieved:.0f} GB/s), separate validity kernel {vkern_ms:.3f} ms")

    # Statistics of the configuration, in the reference's closed form.
    st = Stats()
    run_op_stats = ProcessChunked(None, [sink], [src], slope, Options(TileHeight=args.tile_height, Stats=st))
    assert run_op_stats is None

    # The CPU baseline, and the GPU Slope output checked against it, before
    # the rest of the suite reuses the output raster.
    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(args, dem, out, slope, N, log)

    suite = {}
    if args.suite == "terrain":
        out2 = Float32Raster(torch.empty(cells, dtype=torch.float32, device=dev), N, N, N,
                             torch.zeros_like(valid), 0)
        ops = [
            ("Aspect", terrain.aspectOp(terrain.AspectOptions(CellSize=30)), [sink]),
            ("Hillshade", terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45,
                                                                       ZFactor=1)), [sink]),
            ("CurvatureProfile", terrain.curvatureOp(terrain.CurvatureOptions(
                CellSize=30, Type=terrain.CurvatureProfile)), [sink]),
            ("CurvaturePlan", terrain.curvatureOp(terrain.CurvatureOptions(
                CellSize=30, Type=terrain.CurvaturePlan)), [sink]),
            ("CurvatureMean", terrain.curvatureOp(terrain.CurvatureOptions(
                CellSize=30, Type=terrain.CurvatureMean)), [sink]),
            ("Gradient", terrain.gradientOp(terrain.GradientOptions(CellSize=30)), [sink, MemorySink(out2)]),
        ]
        # terrain.Surface: Dx, Dy, Slope, Aspect and Hillshade from one read
        # (five outputs, 24.75 B/cell with masks).
        extra = [Float32Raster(torch.empty(cells, dtype=torch.float32, device=dev), N, N, N,
                               torch.zeros_like(valid), 0) for _ in range(3)]
        ops.append(("Surface (Dx, Dy, Slope, Aspect, Hillshade)",
                    terrain.newSurface([True] * 5, terrain.SurfaceOptions(CellSize=30, Azimuth=315, Altitude=45)),
                    [sink, MemorySink(out2)] + [MemorySink(e) for e in extra]))
        # terrain.Features over the whole suite (Slope, Aspect, Hillshade and
        # the three curvatures): six outputs from one read of the DEM.
        extra.append(Float32Raster(torch.empty(cells, dtype=torch.float32, device=dev), N, N, N,
                                   torch.zeros_like(valid), 0))
        stack = [terrain.SlopeOptions(CellSize=30), terrain.AspectOptions(CellSize=30),
                 terrain.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45, ZFactor=1),
                 terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureProfile),
                 terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvaturePlan),
                 terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureMean)]
        ops.append(("Features (Slope, Aspect, Hillshade, Curvature profile, plan, mean)",
                    terrain.newFeatures(stack), [sink, MemorySink(out2)] + [MemorySink(e) for e in extra]))
        suite["Slope"] = {"gcells_s": round(value, 2), "ms": round(ms_step, 4),
                          "op_hbm_frac": round(op_bytes / (el / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
                          "data_kernel_ms": round(kern_ms, 4), "validity_kernel_ms": round(vkern_ms, 4)}
        n_steps = max(5, args.steps // 2)
        for name, k, sinks in ops:
            e2, t2 = time_op(k, sinks, n_steps, 2)
            nout = len(sinks)
            b = (4 + 4 * nout + 0.125 * (1 + nout)) * cells
            suite[name] = {"gcells_s": round(cells * n_steps / e2 / 1e9, 2), "ms": round(e2 * 1e3 / n_steps, 4),
                           "op_hbm_frac": round(b / (e2 / n_steps) / 1e9 / HBM_PEAK_GBS, 4),
                           "data_kernel_ms": round(t2.data_ms / max(1, t2.data_launches), 4)}
            log(f"bench: {name} {suite[name]}")
        # The six ops one call each against the one-pass stack.
        sep = sum(suite[n]["ms"] for n in ("Slope", "Aspect", "Hillshade", "CurvatureProfile", "CurvaturePlan",
                                           "CurvatureMean"))
        fused = [v for n, v in suite.items() if n.startswith("Features")][0]["ms"]
        suite["Features vs six calls"] = {"six_calls_ms": round(sep, 4), "features_ms": fused,
                                          "speedup": round(sep / fused, 3)}
        del out2, extra

    paths = None
    if args.paths:
        paths = paths_bench(args, dem, out, src, sink, slope, N, dev, log)

    traffic = None
    if not args.no_traffic:
        traffic = pmc_traffic(args, N, log)

    reduce_res = None
    if args.reduce_steps > 0:
        reduce_res = reduce_bench(args, dem, N, log)
        del out, sink
        torch.cuda.empty_cache()
        reduce_res["folded_stats"] = folded_bench(args, dem, N, dev, log)

    raw_res = None
    if args.raw_size > 0:
        raw_res = raw_bench(args, dev, log)

    batch = None
    if args.batch_tiles > 0:
        out = sink = None
        torch.cuda.empty_cache()
        batch = batch_bench(args, dem, N, dev, log)

    focal_res = None
    if args.focal_size > 0:
        focal_res = focal_bench(args, dev, log)

    serving = None
    if args.serve_seconds > 0:
        serving = serving_bench(args, dem, N, log)

    if args.calibrate_copy:
        tmp = torch.empty_like(dem_data)
        tmp.copy_(dem_data)
        torch.cuda.synchronize()
        del tmp

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": (f"synthetic: seeded diamond-square fBm {N}x{N} (Hurst 0.8, 0-3000 m), "
                 f"{covered * 100:.2f}% NoData in circular blobs of radius 5-50"),
        "config": {"workload": f"terrain.SlopeChunked {N}x{N} float32 DEM, device-resident, "
                               f"TileHeight {args.tile_height}, Horn 3x3, CellSize 30, degrees",
                   "width": N, "height": N, "tile_height": args.tile_height, "sweep_rows": tuning.sweep_rows,
                   "stats": {"cells": st.Cells, "tiles": st.Tiles, "bands": st.Bands,
                             "amplification": round(st.Amplification(), 4)}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_launch"] if traffic and "bytes_per_launch" in traffic else None,
                     "traffic_detail": traffic,
                     "kernel": kernel_name, "bytes_per_launch": data_bytes,
                     "kernel_ms": round(kern_ms, 4), "launches": timing.data_launches,
                     # engine.Stats of one op (the reference's byte accounting, in closed
                     # form) beside the bytes the device moved (counters, data kernel).
                     "stats_check": {"ideal_bytes": st.Ideal, "engine_total_bytes": st.Total(),
                                     "engine_kernel_read": st.KernelRead,
                                     "engine_kernel_written": st.KernelWritten,
                                     "device_hbm_bytes": (traffic["bytes_per_launch"]
                                                          if traffic and "bytes_per_launch" in traffic else None),
                                     "device_over_ideal": (round(traffic["bytes_per_launch"] / st.Ideal, 4)
                                                           if traffic and "bytes_per_launch" in traffic else None)}},
        "op_roofline": {"bytes_per_step": op_bytes, "frac": round(op_bytes / (el / args.steps) / 1e9 /
                                                                   HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "parity_sample": parity,
        "suite": suite,
        "batch": batch,
        "serving": serving,
        "focal": focal_res,
        "reduce": reduce_res,
        "file_to_file": raw_res,
        "paths": paths,
    }
    if serving and serving["levels"]:
        at = min(serving["levels"], key=lambda lv: abs(lv["qps"] - 1000.0))
        line["latency"] = {"qps": at["qps"], "p50_ms": at["p50_ms"], "p99_ms": at["p99_ms"],
                           "what": "serving chain Slope -> Lookup -> Reclass per window request, "
                                   "arrival to completion"}
    print(json.dumps(line), flush=True)
    return 0


def paths_bench(args, dem, out, src, sink, slope, N, dev, log):
    """The other ways a caller runs the headline op over the same DEM:

    - cancellable: ProcessN with a live context (plan units of at least
      2^30 cells claimed in order, so the 20000x20000 DEM is one unit)
      against the same call with none, in alternating blocks of calls;
    - chain: Slope -> Lookup -> Reclass (the serving chain's tables) as one
      exec.Pipeline call, SlopeChunked-style over the resident DEM, against
      the three calls it replaces, with the outputs compared bit for bit;
    - host memory: the DEM and the output in host (pageable) memory, as a
      Go caller hands them over, through ProcessN (the library streams the
      rasters through its tile pipeline) and through ProcessChunked with
      TileHeight 256
      (MemorySource/MemorySink windows copied by the library into pinned
      tile buffers), in Gcells/s and in GB/s over PCIe; the output is
      compared with the device-resident result bit for bit."""
    import numpy as np
    import torch

    from strata_amd import transfer
    from strata_amd.engine import Context, MemorySink, MemorySource, Options, Pipeline, ProcessChunked, ProcessN
    from strata_amd.raster import Float32Raster

    cells = N * N
    res = {}
    steps = max(5, args.steps // 2)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    # Alternating blocks of five calls, so the card's clock, which moves
    # over back-to-back runs, weighs on both forms alike.
    def plain():
        ProcessN(None, [out], [dem], slope, Options())

    def live():
        ProcessN(Context(), [out], [dem], slope, Options())

    plain_s = live_s = 0.0
    blocks = max(4, steps // 5)
    for _ in range(blocks):
        plain_s += timed(plain, 5)
        live_s += timed(live, 5)
    plain_s /= blocks
    live_s /= blocks
    res["cancellable"] = {"plain_ms": round(plain_s * 1e3, 4), "cancellable_ms": round(live_s * 1e3, 4),
                          "overhead": round(live_s / plain_s - 1, 4),
                          "what": "ProcessN Slope over the resident DEM, synchronous calls, with and without a "
                                  "live context"}
    log(f"bench: cancellable {res['cancellable']}")

    lookup = transfer.newLookupKernel("risk", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])
    reclass = transfer.newReclassKernel("risk", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5])
    pipe = Pipeline([slope, lookup, reclass])
    tmp = Float32Raster(torch.empty(cells, dtype=torch.float32, device=dev), N, N, N,
                        torch.zeros_like(dem.Valid), 0)
    eopts = Options(TileHeight=args.tile_height)
    tsink = MemorySink(tmp)

    def separate():
        ProcessChunked(None, [sink], [src], slope, eopts, async_=True)
        ProcessChunked(None, [tsink], [MemorySource(out)], lookup, eopts, async_=True)
        ProcessChunked(None, [sink], [MemorySource(tmp)], reclass, eopts, async_=True)

    sep_s = timed(separate, steps)
    want = (out.Data.clone(), out.Valid.clone())
    chain_s = timed(lambda: ProcessChunked(None, [sink], [src], pipe, eopts, async_=True), steps)
    same = bool(torch.equal(out.Data.view(torch.int32), want[0].view(torch.int32)) and
                torch.equal(out.Valid, want[1]))
    res["chain"] = {"chain_ms": round(chain_s * 1e3, 4), "gcells_s": round(cells / chain_s / 1e9, 2),
                    "three_calls_ms": round(sep_s * 1e3, 4), "speedup": round(sep_s / chain_s, 3),
                    "identical": same,
                    "what": f"Slope -> Lookup -> Reclass over the resident {N}x{N} DEM, TileHeight "
                            f"{args.tile_height}, one pipeline call against three calls"}
    log(f"bench: chain {res['chain']}")
    del tmp, tsink, want

    # Host memory. The device result of Slope is the reference bits.
    ProcessChunked(None, [sink], [src], slope, eopts)
    want_d = out.Data.cpu().numpy()
    want_v = out.Valid.cpu().numpy().view(np.uint64)
    hsrc = Float32Raster(dem.Data.cpu().numpy(), N, N, N, dem.Valid.cpu().numpy().view(np.uint64).copy(), 0)
    hout = Float32Raster(np.zeros(cells, np.float32), N, N, N, np.zeros(len(want_v), np.uint64), 0)
    mask_b = 8 * len(want_v)

    def same_host():
        return bool(np.array_equal(hout.Data.view(np.uint32), want_d.view(np.uint32)) and
                    np.array_equal(hout.Valid, want_v))

    reps = 3
    h_plain = timed(lambda: ProcessN(None, [hout], [hsrc], slope, Options()), reps)
    ok_plain = same_host()
    hout.Data[:] = 0
    hout.Valid[:] = 0
    h_chunk = timed(lambda: ProcessChunked(None, [MemorySink(hout)], [MemorySource(hsrc)], slope, eopts), reps)
    ok_chunk = same_host()
    # Bytes over PCIe: ProcessN moves the source and its mask in, the output's
    # mask in and out (its words are shared with cells outside it in general)
    # and the output's cells out; the chunked tiles move each tile's window
    # with its halo rows in and the tile out, with their masks.
    plain_b = 4 * cells * 2 + 3 * mask_b
    chunk_b = 4 * cells * 2 + 2 * mask_b
    res["host_memory"] = {
        "plain": {"ms": round(h_plain * 1e3, 2), "gcells_s": round(cells / h_plain / 1e9, 3),
                  "pcie_gb_s": round(plain_b / h_plain / 1e9, 2), "identical": ok_plain},
        "chunked": {"ms": round(h_chunk * 1e3, 2), "gcells_s": round(cells / h_chunk / 1e9, 3),
                    "pcie_gb_s": round(chunk_b / h_chunk / 1e9, 2), "identical": ok_chunk,
                    "tile_height": args.tile_height},
        "what": f"terrain.Slope {N}x{N} with the DEM and the output in host (pageable) memory: ProcessN "
                f"(streamed by the library through its tile pipeline, tiles of about 16 MiB) and ProcessChunked "
                f"(MemorySource/MemorySink "
                f"windows copied by the library into pinned tile buffers)"}
    log(f"bench: host memory {res['host_memory']}")
    return res


def raw_bench(args, dev, log):
    """File-to-file SlopeChunked: an fBm DEM written as a raw float32 file
    with a fill value (0.5% blob NoData), read tile by tile by the library,
    its NoData derived on the device, the slope written to a second raw file
    with the fill in its invalid cells. Wall time of the whole call (file
    reads and writes in the page cache, PCIe both ways, kernels)."""
    import tempfile

    import numpy as np
    import torch

    from strata_amd import synth, terrain
    from strata_amd.engine import CreateRawFile, OpenRawFile, Options, RawOptions, RawSink, RawSource

    N = args.raw_size
    fill = -9999.0
    data = synth.fbm_dem(N, N, hurst=0.8, seed=1, lo=0.0, hi=3000.0, device=dev)
    valid, _ = synth.blob_nodata(N, N, frac=0.005, rmin=5, rmax=50, seed=2, device=dev)
    idx = torch.arange(N * N, device=dev)
    bits = ((valid[idx >> 6] >> (idx & 63)) & 1).bool()
    data = torch.where(bits, data, torch.tensor(fill, device=dev))
    opts = RawOptions(Fill=fill, HasFill=True)
    res = {}
    with tempfile.TemporaryDirectory(prefix="strata_raw_") as d:
        ipath, opath = os.path.join(d, "dem.raw"), os.path.join(d, "slope.raw")
        data.cpu().numpy().astype("<f4").tofile(ipath)
        del data, bits, idx
        fi = OpenRawFile(ipath, os.O_RDONLY)
        fo = CreateRawFile(opath, 4 * N * N)
        k = terrain.SlopeOptions(CellSize=30)
        runs = []
        for rep in range(3):
            t = time.perf_counter()
            err = terrain.SlopeChunked(None, RawSink(fo, N, N, opts), RawSource(fi, N, N, opts), k,
                                       Options(TileHeight=args.tile_height))
            runs.append(time.perf_counter() - t)
            assert err is None, err
        fi.Close()
        fo.Close()
        el = min(runs)
        res = {"size": N, "tile_height": args.tile_height, "s": round(el, 4),
               "gcells_s": round(N * N / el / 1e9, 3), "file_gb_s": round(8 * N * N / el / 1e9, 3),
               "runs_s": [round(r, 4) for r in runs],
               # benchmarks/chunked/RESULTS.md:346 (Slope 20000^2, Chunked strips 256, raw
               # file to file, 12 workers on a Ryzen 9 3900X): another machine's figure.
               "reference_published_gcells_s": 0.832 if N == 20000 else None}
    log(f"bench: file-to-file {res}")
    return res


def reduce_bench(args, dem, N, log):
    """reduce.Count / MinMax / Sum / Stats over the device-resident DEM: one
    fold launch each (plus the one-block combine and the read-back of the
    bins); whole-call wall time, and the HBM share of the bytes a fold must
    read (4 B data + 1/8 B mask per cell; the mask alone for Count)."""
    import torch

    from strata_amd import _lib, engine
    from strata_amd import reduce as rd

    cells = N * N
    res = {}
    for name, what in (("Count", _lib.REDUCE_COUNT), ("MinMax", _lib.REDUCE_MINMAX), ("Sum", _lib.REDUCE_SUM),
                       ("Stats", _lib.REDUCE_MINMAX | _lib.REDUCE_MOMENTS)):
        for _ in range(2):
            engine.Reduce(None, [dem], what)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.reduce_steps):
            f, err = engine.Reduce(None, [dem], what)
            assert err is None
        el = (time.perf_counter() - t) / args.reduce_steps
        b = (0.125 if what == _lib.REDUCE_COUNT else 4.125) * cells
        res[name] = {"gcells_s": round(cells / el / 1e9, 2), "ms": round(el * 1e3, 4),
                     "hbm_frac": round(b / el / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"bench: reduce {name} {res[name]}")
    s = rd.Stats(dem)
    res["summary"] = {"count": s.Count, "sum": s.Sum, "mean": s.Mean, "stddev": s.StdDev, "min": float(s.Min),
                      "max": float(s.Max)}
    return res


def folded_bench(args, dem, N, dev, log):
    """Slope's Stats folded into the Slope pass through a tee sink
    (graph/sinks.go): alone (no raster written), and with the raster
    written too, against the op followed by reduce.Stats of its output.
    Whole-call wall times; the resident pass's step of rows is swept for
    the alone form."""
    import torch

    from strata_amd import _lib, terrain
    from strata_amd import reduce as rd
    from strata_amd.engine import MemorySink, MemorySource, Options, ProcessChunked, TeeSink
    from strata_amd.raster import Float32Raster

    cells = N * N
    slope = terrain.slopeOp(terrain.SlopeOptions(CellSize=30))
    out = Float32Raster(torch.empty(cells, dtype=torch.float32, device=dev), N, N, N,
                        torch.zeros((cells + 63) // 64 + 1, dtype=torch.int64, device=dev), 0)
    opts = Options(TileHeight=args.tile_height)

    def alone():
        tee = TeeSink(N, N, True)
        assert ProcessChunked(None, [tee], [MemorySource(dem)], slope, opts) is None
        return tee

    def written():
        tee = TeeSink(N, N, True, [MemorySink(out)])
        assert ProcessChunked(None, [tee], [MemorySource(dem)], slope, opts) is None
        return tee

    def separate():
        assert ProcessChunked(None, [MemorySink(out)], [MemorySource(dem)], slope, opts) is None
        return rd.Stats(out)

    def timed(fn):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.reduce_steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.reduce_steps

    L = _lib.lib()
    saved = _lib.Tuning()
    L.strata_get_tuning(saved)
    sweep = {}
    for mb in (64, 256, 1024, 4096):
        t = _lib.Tuning()
        L.strata_get_tuning(t)
        t.fold_stripe_mb = mb
        assert L.strata_set_tuning(t) == 0, _lib.last_error()
        sweep[mb] = round(timed(alone) * 1e3, 4)
    L.strata_set_tuning(saved)
    res = {"step_mb_sweep_ms": sweep}
    for name, fn in (("alone", alone), ("written", written), ("separate", separate)):
        el = timed(fn)
        b = (4.125 if name == "alone" else 8.25) * cells
        res[name] = {"ms": round(el * 1e3, 4), "gcells_s": round(cells / el / 1e9, 2),
                     "hbm_frac": round(b / el / 1e9 / HBM_PEAK_GBS, 4)}
    want = separate()
    res["identical"] = rd.Folded(alone()) == want and rd.Folded(written()) == want
    res["speedup_alone"] = round(res["separate"]["ms"] / res["alone"]["ms"], 3)
    log(f"bench: folded stats {res}")
    del out
    torch.cuda.empty_cache()
    return res


def pmc_child(args) -> int:
    """The workload of one counter pass: the op twice on the same input,
    then one device copy of the input (the calibration kernel). Slope runs
    over the DEM as ProcessChunked with --tile-height; the focal ops over
    the focal workload's raster as ProcessN."""
    import torch

    from strata_amd import _lib, focal, synth, terrain
    from strata_amd.engine import MemorySink, MemorySource, Options, ProcessChunked, ProcessN
    from strata_amd.raster import Float32Raster

    dev = torch.device("cuda:0")
    L = _lib.lib()
    assert L.strata_init(0) == 0, _lib.last_error()
    if args.sweep_rows:
        t = _lib.Tuning()
        L.strata_get_tuning(t)
        t.sweep_rows = args.sweep_rows
        assert L.strata_set_tuning(t) == 0, _lib.last_error()
    if args.pmc_op != "slope":
        import numpy as np
        S = args.focal_size
        data = synth.uniform_raster(S * S, 0.0, 1000.0, seed=3, device=dev)
        valid = synth.random_nodata(S * S, 0.02, seed=4, device=dev)
        src = Float32Raster(data, S, S, S, valid, 0)
        dst = Float32Raster(torch.empty(S * S, dtype=torch.float32, device=dev), S, S, S, torch.zeros_like(valid), 0)
        r = {"mean3": 3, "correlate1": 1, "gauss2": 2, "correlate8-mfma": 8, "correlate4-mfma": 4}[args.pmc_op]
        n = 2 * r + 1
        if args.pmc_op == "mean3":
            k = focal.newBox(focal.BoxOptions(3), "mean")
        elif args.pmc_op == "gauss2":
            g = focal.Gaussian(r, r / 2)
            k = focal.newSeparableKernel(focal.SeparableOptions(r, g, g))
        else:
            w = np.random.default_rng(5).uniform(-1, 1, n * n).astype(np.float32)
            k = focal.newWeightsKernel(focal.WeightsOptions(r, w), False, mfma=args.pmc_op.endswith("mfma"))
        for _ in range(2):
            assert ProcessN(None, [dst], [src], k, Options()) is None
        tmp = torch.empty_like(data)
        tmp.copy_(data)
        torch.cuda.synchronize()
        return 0
    N = args.size
    data = synth.fbm_dem(N, N, hurst=0.8, seed=1, lo=0.0, hi=3000.0, device=dev)
    valid, _ = synth.blob_nodata(N, N, frac=0.005, rmin=5, rmax=50, seed=2, device=dev)
    dem = Float32Raster(data, N, N, N, valid, 0)
    out = Float32Raster(torch.empty(N * N, dtype=torch.float32, device=dev), N, N, N, torch.zeros_like(valid), 0)
    slope = terrain.slopeOp(terrain.SlopeOptions(CellSize=30, Units=terrain.SlopeDegrees))
    for _ in range(2):
        assert ProcessChunked(None, [MemorySink(out)], [MemorySource(dem)], slope,
                              Options(TileHeight=args.tile_height)) is None
    tmp = torch.empty_like(data)
    tmp.copy_(data)
    torch.cuda.synchronize()
    return 0


SIZED_READS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum")


def pmc_traffic(args, N, log):
    """HBM bytes per launch of the Slope kernel from two rocprofv3
    counter passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each
    scaled by what the same counter reads for a plain device copy of the
    DEM (4·N² bytes read and written) in the same run. A third pass counts
    the read requests by size (32, 64 and 128 bytes): FETCH_SIZE counts a
    128-byte request as 64 bytes on this part, so the copy's factor holds
    for a kernel only if its requests have the copy's sizes; the sized
    count needs no factor, and the copy checks it."""
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(exe):
        return {"error": "rocprofv3 not found"}
    copy_kib = 4.0 * N * N / 1024.0
    got = {}
    import csv
    with tempfile.TemporaryDirectory(prefix="strata_pmc_") as d:
        for name, ctrs in (("FETCH_SIZE", ["FETCH_SIZE"]), ("WRITE_SIZE", ["WRITE_SIZE"]), ("SIZED", list(SIZED_READS))):
            out = os.path.join(d, name)
            cmd = [exe, "--pmc", *ctrs, "--output-format", "csv", "-d", out, "-o", "run", "--",
                   sys.executable, os.path.abspath(__file__), "--pmc-child", "--size", str(N),
                   "--tile-height", str(args.tile_height)]
            if args.sweep_rows:
                cmd += ["--sweep-rows", str(args.sweep_rows)]
            log(f"bench: counter pass {' '.join(ctrs)}")
            try:
                p = subprocess.run(cmd, capture_output=True, text=True, timeout=150)
            except subprocess.TimeoutExpired:
                return {"error": f"rocprofv3 --pmc {name} timed out"}
            files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
            if p.returncode != 0 or not files:
                return {"error": f"rocprofv3 --pmc {name} failed ({p.returncode})"}
            # value per dispatch: the counters' weighted sum (bytes for SIZED, KiB otherwise)
            weight = {"TCC_EA0_RDREQ_32B_sum": 32.0, "TCC_EA0_RDREQ_64B_sum": 64.0, "TCC_EA0_RDREQ_128B_sum": 128.0,
                      "FETCH_SIZE": 1.0, "WRITE_SIZE": 1.0}
            slope, copy = {}, {}
            with open(files[0]) as f:
                for row in csv.DictReader(f):
                    kname = row.get("Kernel_Name", "")
                    v = float(row["Counter_Value"]) * weight[row["Counter_Name"]]
                    # The aligned Slope sweep, with or without the validity waves.
                    if any(k in kname for k in ("horn_sweep<2, false>", "horn_sweep<2>", "horn_sweep_valid<2, false>",
                                                "horn_sweep_valid<2>")):
                        slope[row["Dispatch_Id"]] = slope.get(row["Dispatch_Id"], 0.0) + v
                    elif "copyBuffer" in kname:
                        copy[row["Dispatch_Id"]] = copy.get(row["Dispatch_Id"], 0.0) + v
            if not slope or not copy:
                return {"error": f"rocprofv3 --pmc {name}: kernels not found"}
            got[name] = (sum(slope.values()) / len(slope), max(copy.values()))
    ff = copy_kib / got["FETCH_SIZE"][1]
    wf = copy_kib / got["WRITE_SIZE"][1]
    fetch = got["FETCH_SIZE"][0] * ff * 1024
    write = got["WRITE_SIZE"][0] * wf * 1024
    sized_read, sized_copy = got["SIZED"]
    res = {"bytes_per_launch": round(fetch + write), "read_bytes": round(fetch), "write_bytes": round(write),
           "fetch_kib_raw": round(got["FETCH_SIZE"][0], 1), "write_kib_raw": round(got["WRITE_SIZE"][0], 1),
           "fetch_factor": round(ff, 4), "write_factor": round(wf, 4),
           "read_bytes_by_request_size": round(sized_read),
           "copy_read_bytes_by_request_size_over_actual": round(sized_copy / (copy_kib * 1024), 4),
           "calibration": f"device copy of the {N}x{N} DEM ({copy_kib:.0f} KiB read and written), same run",
           "algorithmic_bytes": 8.25 * N * N, "algorithmic_note": "4 B in and out per cell and 1/8 B of mask "
                                                                   "in and out (the sweep writes the validity)"}
    log(f"bench: traffic {res['bytes_per_launch'] / 1e9:.3f} GB/launch (read {fetch / 1e9:.3f}, write "
        f"{write / 1e9:.3f}; factors {ff:.3f}/{wf:.3f}); reads by request size {sized_read / 1e9:.3f} GB "
        f"(copy check {res['copy_read_bytes_by_request_size_over_actual']})")
    return res


def batch_tiles(n, N, seed):
    """n tiles with sides log-uniform in [256, 2048], at uniform offsets
    inside the N×N DEM, each with an op drawn uniformly from {Slope,
    Hillshade, mean Curvature, focal Mean r=3}."""
    import numpy as np

    rng = np.random.default_rng(seed)
    lo, hi = np.log(256), np.log(2048)
    w = np.minimum(np.exp(rng.uniform(lo, hi, n)).astype(np.int64), N)
    h = np.minimum(np.exp(rng.uniform(lo, hi, n)).astype(np.int64), N)
    x = (rng.random(n) * (N - w + 1)).astype(np.int64)
    y = (rng.random(n) * (N - h + 1)).astype(np.int64)
    op = rng.integers(0, 4, n)
    return w, h, x, y, op


def batch_bench(args, dem, N, dev, log):
    """Many uneven tiles with mixed ops: one persistent, graph-captured,
    tuned launch against one plain-path launch set per tile. Both write the
    same output buffers; their outputs are compared bit for bit."""
    import torch

    from strata_amd import focal, terrain
    from strata_amd.batch import Batch, Task
    from strata_amd.raster import Float32Raster

    ops = [terrain.slopeOp(terrain.SlopeOptions(CellSize=30)),
           terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45, ZFactor=1)),
           terrain.curvatureOp(terrain.CurvatureOptions(CellSize=30, Type=terrain.CurvatureMean)),
           focal.newBox(focal.BoxOptions(3), "mean")]
    w, h, x, y, op = batch_tiles(args.batch_tiles, N, seed=4)
    cells = int((w * h).sum())
    out = torch.empty(cells, dtype=torch.float32, device=dev)
    words = (cells + 63) // 64 + 1
    mask = torch.zeros(words, dtype=torch.int64, device=dev)
    tasks = []
    off = 0
    for i in range(len(w)):
        wi, hi_, xi, yi = int(w[i]), int(h[i]), int(x[i]), int(y[i])
        start = yi * N + xi
        src = Float32Raster(dem.Data[start:start + (hi_ - 1) * N + wi], wi, hi_, N, dem.Valid, start)
        dst = Float32Raster(out[off:off + wi * hi_], wi, hi_, wi, mask, off)
        tasks.append(Task(ops[int(op[i])], src, dst))
        off += wi * hi_
    torch.cuda.synchronize()

    b = Batch(tasks, graph=True)
    shape = b.tune(reps=2)

    def timed(fn, reps):
        fn()
        b.sync()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        b.sync()
        return (time.perf_counter() - t) / reps

    per_tile_s = timed(b.run_per_tile, max(1, args.batch_reps // 5))
    ref = out.clone()
    ref_mask = mask.clone()
    batch_s = timed(b.run, args.batch_reps)
    identical = bool(torch.equal(out.view(torch.int32), ref.view(torch.int32)) and torch.equal(mask, ref_mask))
    info = b.info()
    b.close()
    del ref, ref_mask
    n = len(tasks)
    res = {
        "workload": f"{n} tiles, sides log-uniform 256-2048, ops uniform over Slope, Hillshade, mean Curvature, "
                    f"focal Mean r=3, windows of the {N}x{N} DEM, each with its own output and mask",
        "tiles": n, "cells": cells,
        "batched": {"ms": round(batch_s * 1e3, 3), "tiles_per_s": round(n / batch_s, 1),
                    "gcells_s": round(cells / batch_s / 1e9, 2),
                    "hbm_frac": round(8.25 * cells / batch_s / 1e9 / HBM_PEAK_GBS, 4),
                    "launch": "one persistent kernel in a captured HIP graph", "shape": shape,
                    "work_units": info["data_units"] + info["validity_units"]},
        "per_tile": {"ms": round(per_tile_s * 1e3, 3), "tiles_per_s": round(n / per_tile_s, 1),
                     "gcells_s": round(cells / per_tile_s / 1e9, 2),
                     "launch": "strata_process per tile (data and validity kernels)"},
        "speedup": round(per_tile_s / batch_s, 2),
        "identical": identical,
    }
    log(f"bench: batch {n} tiles {cells / 1e9:.2f} Gcells: batched {batch_s * 1e3:.2f} ms "
        f"({res['batched']['gcells_s']} Gcells/s, shape {shape}), per tile {per_tile_s * 1e3:.2f} ms, "
        f"speedup {res['speedup']}x, identical {identical}")
    del out, mask
    torch.cuda.empty_cache()
    return res


def focal_bench(args, dev, log):
    """Focal ops over an S×S raster of seeded uniform values in [0, 1000)
    with 2% i.i.d. NoData: Correlate and Convolve at r = 1, 2, 4, 8 with
    seeded weights U[-1, 1], CorrelateSeparable with Gaussian taps
    (sigma = r/2) at r = 1, 2, 4, 8, and Mean, Min, Max at r = 3; plain
    ProcessN over the device-resident raster."""
    import numpy as np
    import torch

    from strata_amd import _lib, focal, synth
    from strata_amd.engine import Options, ProcessN
    from strata_amd.raster import Float32Raster

    S = args.focal_size
    cells = S * S
    data = synth.uniform_raster(cells, 0.0, 1000.0, seed=3, device=dev)
    valid = synth.random_nodata(cells, 0.02, seed=4, device=dev)
    src = Float32Raster(data, S, S, S, valid, 0)
    dst = Float32Raster(torch.empty(cells, dtype=torch.float32, device=dev), S, S, S, torch.zeros_like(valid), 0)
    rng = np.random.default_rng(5)
    ops = []
    for r in (1, 2, 4, 8):
        n = 2 * r + 1
        w = rng.uniform(-1, 1, n * n).astype(np.float32)
        ops.append((f"Correlate r={r}", focal.newWeightsKernel(focal.WeightsOptions(r, w), False), 2 * n * n - 1))
        ops.append((f"Convolve r={r}", focal.newWeightsKernel(focal.WeightsOptions(r, w), True), 2 * n * n - 1))
    errs = {}
    for r in (4, 8):
        n = 2 * r + 1
        w = np.random.default_rng(50 + r).uniform(-1, 1, n * n).astype(np.float32)
        k32 = focal.newWeightsKernel(focal.WeightsOptions(r, w), False)
        kmf = focal.newWeightsKernel(focal.WeightsOptions(r, w), False, mfma=True)
        ops.append((f"Correlate r={r} MFMA split-BF16x3", kmf, 2 * n * n - 1))
        errs[r] = (k32, kmf)
    # Executed MFMA flops per cell: per weight row, three 16x16x32 products
    # (16*16*32*2 flops each) per 16x16 output block.
    mfma_flops = {f"Correlate r={r} MFMA split-BF16x3": (2 * r + 1) * 3 * 16 * 16 * 32 * 2 // 256 for r in (4, 8)}
    for r in (1, 2, 4, 8):
        g = focal.Gaussian(r, r / 2)
        ops.append((f"CorrelateSeparable Gaussian r={r}", focal.newSeparableKernel(focal.SeparableOptions(r, g, g)),
                    2 * 2 * (2 * r + 1) - 2))
    for op in ("mean", "min", "max"):
        ops.append((f"{op.capitalize()} r=3", focal.newBox(focal.BoxOptions(3), op), 13 if op == "mean" else 12))
    res = {"workload": f"{S}x{S} float32 uniform [0,1000) seeded, 2% i.i.d. NoData; plain ProcessN, "
                       "device-resident", "ops": {}}
    for name, k, flops in ops:
        assert ProcessN(None, [dst], [src], k, Options()) is None  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.focal_steps):
            assert ProcessN(None, [dst], [src], k, Options(), async_=True) is None
        torch.cuda.synchronize()
        el = (time.perf_counter() - t) / args.focal_steps
        # The kernels' durations, in separate steps (the timing events stay
        # out of the timed ones).
        timing = _lib.KernelTiming().begin()
        for _ in range(args.focal_steps):
            assert ProcessN(None, [dst], [src], k, Options(), async_=True) is None
        torch.cuda.synchronize()
        timing.end()
        kms = timing.data_ms / max(1, timing.data_launches)
        res["ops"][name] = {"gcells_s": round(cells / el / 1e9, 2), "ms": round(el * 1e3, 4),
                            "data_kernel_ms": round(kms, 4),
                            "hbm_frac": round(8.25 * cells / el / 1e9 / HBM_PEAK_GBS, 4),
                            "fp32_tflops": round(flops * cells / (kms * 1e-3) / 1e12, 2)}
        if name in mfma_flops:
            res["ops"][name]["mfma_executed_flops_per_cell"] = mfma_flops[name]
            res["ops"][name]["mfma_executed_tflops"] = round(mfma_flops[name] * cells / (kms * 1e-3) / 1e12, 1)
        log(f"bench: focal {name}: {res['ops'][name]}")
    # Precision of both Correlate paths against the FP64 correlation (on the
    # device), over every interior cell.
    res["error_vs_f64"] = {}
    for r, (k32, kmf) in errs.items():
        for label, k in (("fp32", k32), ("mfma_bf16x3", kmf)):
            assert ProcessN(None, [dst], [src], k, Options()) is None
            torch.cuda.synchronize()
            e = focal.CorrelateError(k, src, dst)
            res["error_vs_f64"][f"r={r} {label}"] = {"max_abs": e["max_abs"], "max_rel": e["max_rel"],
                                                     "max_abs_over_sum_abs_wx": e["max_scaled"],
                                                     "cells": e["cells"]}
            log(f"bench: focal error r={r} {label}: {res['error_vs_f64'][f'r={r} {label}']}")
    del src, dst, data, valid
    torch.cuda.empty_cache()
    res["bf16_gemm"] = bf16_gemm_rate(dev, log)
    return res


def bf16_gemm_rate(dev, log):
    """The dense bf16 rate the vendor library reaches on this GPU: an
    8192x8192x8192 torch.mm of contiguous bf16 operands (hipBLASLt), the
    best of five timed runs after a warm-up."""
    import torch

    n = 8192
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    a = torch.randn(n, n, generator=g, device=dev, dtype=torch.float32).to(torch.bfloat16).contiguous()
    b = torch.randn(n, n, generator=g, device=dev, dtype=torch.float32).to(torch.bfloat16).contiguous()
    c = torch.mm(a, b)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            torch.mm(a, b, out=c)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) / 10)
    out = {"shape": [n, n, n], "ms": round(best, 4), "tflops": round(2 * n ** 3 / (best * 1e-3) / 1e12, 1),
           "what": "torch.mm bf16 (hipBLASLt) on contiguous operands, best of 5 runs of 10"}
    log(f"bench: bf16 GEMM {out}")
    del a, b, c
    torch.cuda.empty_cache()
    return out


def serving_bench(args, dem, N, log):
    """Open-loop Poisson arrivals at each QPS level against the resident
    DEM; each request is the chain Slope -> Lookup -> Reclass on a window
    (80% 512², 15% 2048², 5% 4096²). Latency runs from each request's
    scheduled arrival to its completion."""
    import numpy as np

    from strata_amd import terrain, transfer
    from strata_amd.serve import Server, poisson_schedule

    slope = terrain.slopeOp(terrain.SlopeOptions(CellSize=30))
    lookup = transfer.newLookupKernel("risk", [0, 5, 10, 20, 30, 40], [1, 1.1, 1.3, 1.9, 3, 4.5])
    reclass = transfer.newReclassKernel("class", [1.2, 1.6, 2.2, 3], [1, 2, 3, 4, 5])
    budget = 8 << 30
    levels = []
    with Server(dem, slope, (lookup, reclass), slots=4, hbm_budget=budget, max_batch=256,
                rows=args.serve_rows, auto_release=True) as srv:
        w, a = poisson_schedule(500.0, 100, N, N, seed=99)
        srv.replay(w, a)  # warm-up
        for i, qps in enumerate(float(q) for q in args.serve_qps.split(",")):
            n = int(max(150, qps * args.serve_seconds))
            windows, arrivals = poisson_schedule(qps, n, N, N, seed=100 + i)
            before = srv.stats()
            lat = srv.replay(windows, arrivals)
            after = srv.stats()
            cells = int((windows[:, 2] * windows[:, 3]).sum())
            span = float(arrivals[-1] + lat[-1] / 1e3) if n else 0.0
            batches = after["batches"] - before["batches"]
            lv = {"qps": qps, "requests": n, "p50_ms": round(float(np.percentile(lat, 50)), 4),
                  "p99_ms": round(float(np.percentile(lat, 99)), 4), "mean_ms": round(float(lat.mean()), 4),
                  "max_ms": round(float(lat.max()), 4), "achieved_rps": round(n / span, 1),
                  "gcells_s": round(cells / span / 1e9, 3), "batches": batches,
                  "mean_batch": round(n / max(1, batches), 2),
                  "admission_waits": after["admission_waits"] - before["admission_waits"],
                  "mean_queue_ms": round((after["queue_ms"] - before["queue_ms"]) / n, 4),
                  "mean_service_ms": round((after["service_ms"] - before["service_ms"]) / n, 4)}
            levels.append(lv)
            log(f"bench: serve {qps:.0f} QPS: p50 {lv['p50_ms']} ms, p99 {lv['p99_ms']} ms, "
                f"{lv['achieved_rps']} req/s, mean batch {lv['mean_batch']}, queue {lv['mean_queue_ms']} ms, "
                f"service {lv['mean_service_ms']} ms")
        burst = None
        if args.serve_burst > 0:
            # Closed loop past the knee: every request of the mix arrives at
            # once, so the server runs at its own pace; its throughput is
            # the served form's capacity.
            bw, _ = poisson_schedule(1000.0, args.serve_burst, N, N, seed=300)
            zero = np.zeros(len(bw))
            srv.replay(bw, zero)  # warm-up of the same shape
            before = srv.stats()
            lat = srv.replay(bw, zero)
            after = srv.stats()
            span = float(lat.max()) / 1e3
            batches = after["batches"] - before["batches"]
            cells = int((bw[:, 2] * bw[:, 3]).sum())
            burst = {"requests": len(bw), "cells": cells, "makespan_ms": round(span * 1e3, 3),
                     "tiles_per_s": round(len(bw) / span, 1), "gcells_s": round(cells / span / 1e9, 2),
                     "batches": batches, "mean_batch": round(len(bw) / max(1, batches), 2),
                     "p50_ms": round(float(np.percentile(lat, 50)), 4),
                     "p99_ms": round(float(np.percentile(lat, 99)), 4),
                     "admission_waits": after["admission_waits"] - before["admission_waits"]}
            log(f"bench: serve closed loop: {burst}")
        st = srv.stats()
    res = {"workload": f"open-loop Poisson arrivals; windows 80% 512², 15% 2048², 5% 4096² of the resident "
                       f"{N}x{N} DEM; chain Slope (CellSize 30, degrees) -> Lookup (0,5,10,20,30,40 -> "
                       f"1,1.1,1.3,1.9,3,4.5) -> Reclass (1.2,1.6,2.2,3 -> 1..5) fused in one pass",
           "slots": 4, "hbm_budget_bytes": budget, "max_batch": 256, "rows": args.serve_rows, "levels": levels,
           "peak_result_bytes": st["peak_bytes"], "max_in_flight": st["max_in_flight"]}
    if burst is not None:
        res["closed_loop"] = burst
        res["per_tile"] = serving_per_tile(bw, dem, N, (slope, lookup, reclass), log)
        res["served_over_per_tile"] = round(burst["tiles_per_s"] / res["per_tile"]["tiles_per_s"], 2)
        log(f"bench: served / per-tile tiles/s: {res['served_over_per_tile']}")
    return res


def serving_per_tile(windows, dem, N, chain, log):
    """The same requests through the plain per-tile path: for each window,
    strata_process of Slope, then of Lookup and Reclass over its output in
    place, launched back to back on one stream (the batch API's per-tile
    runner; per window this is faster than one strata_process_chain call),
    each request with its own output raster."""
    import numpy as np
    import torch

    from strata_amd.batch import Batch, Task
    from strata_amd.raster import Float32Raster

    slope, lookup, reclass = chain
    dev = dem.Data.device
    w = windows[:, 2].astype(np.int64)
    h = windows[:, 3].astype(np.int64)
    cells = int((w * h).sum())
    out = torch.empty(cells, dtype=torch.float32, device=dev)
    mask = torch.zeros((cells + 63) // 64 + 1, dtype=torch.int64, device=dev)
    tasks, off = [], 0
    for (x, y, wi, hi_) in np.asarray(windows, dtype=np.int64):
        x, y, wi, hi_ = int(x), int(y), int(wi), int(hi_)
        start = y * N + x
        src = Float32Raster(dem.Data[start:start + (hi_ - 1) * N + wi], wi, hi_, N, dem.Valid, start)
        dst = Float32Raster(out[off:off + wi * hi_], wi, hi_, wi, mask, off)
        tasks.append(Task(slope, src, dst, (lookup, reclass)))
        off += wi * hi_
    b = Batch(tasks)
    b.run_per_tile()
    b.sync()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        b.run_per_tile()
    b.sync()
    s = (time.perf_counter() - t) / reps
    b.close()
    del out, mask
    torch.cuda.empty_cache()
    res = {"requests": len(tasks), "ms": round(s * 1e3, 3), "tiles_per_s": round(len(tasks) / s, 1),
           "gcells_s": round(cells / s / 1e9, 2),
           "launch": "strata_process per request and stage (Slope, then Lookup and Reclass in place), back to back "
                     "on one stream"}
    log(f"bench: serve per-tile path: {res}")
    return res


def cpu_baseline(args, dem, out, slope, N, log):
    """The CPU baseline: the reference's AVX2 Tiled path restated in C
    (oracle/strata_oracle.c orc_tiled_avx2: 8-cell lanes in
    internal/stencil/simd_amd64.go's operation order, bands of
    max(1, 65536/W) rows claimed in order by worker threads as
    internal/exec/worker.go:144-181), timed on the host cores over the
    whole DEM for Slope and Hillshade. The GPU's Slope output is compared
    with it bit for bit over the whole raster."""
    import numpy as np
    import torch

    from oracle import oracle
    from strata_amd import terrain
    from strata_amd.raster import Float32Raster

    # First, while no worker threads are still running after a parallel run
    # (they would take the cores the launching thread needs).
    plain = plain_4096(slope, log)

    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    n = N * N
    src = Float32Raster(dem.Data.cpu().numpy(), N, N, N, dem.Valid.cpu().numpy().view(np.uint64), 0)
    dst = Float32Raster(np.empty(n, np.float32), N, N, N, np.zeros(len(src.Valid), np.uint64), 0)

    def timed(k):
        t = time.perf_counter()
        oracle.tiled_avx2(k, [dst], [src], 0, 0, threads)
        one = time.perf_counter() - t
        reps = max(1, min(int(args.cpu_seconds / max(one * threads, 1e-6)), 50))
        t = time.perf_counter()
        for _ in range(reps):
            oracle.tiled_avx2(k, [dst], [src], 0, 0, threads)
        el = (time.perf_counter() - t) / reps
        return el, reps

    hill = terrain.hillshadeOp(terrain.HillshadeOptions(CellSize=30, Azimuth=315, Altitude=45))
    h_s, h_reps = timed(hill)
    s_s, s_reps = timed(slope)  # last, so dst holds Slope for the comparison
    log(f"bench: CPU AVX2 Tiled {threads} threads, {N}x{N}: Slope {s_s * 1e3:.1f} ms ({n / s_s / 1e9:.3f} Gcells/s), "
        f"Hillshade {h_s * 1e3:.1f} ms ({n / h_s / 1e9:.3f} Gcells/s)")

    got = out.Data.cpu().numpy()
    same = (got.view(np.uint32) == dst.Data.view(np.uint32)) | (np.isnan(got) & np.isnan(dst.Data))
    gv = out.Valid.cpu().numpy().view(np.uint64)
    parity = {"cells": int(n), "data_equal": bool(same.all()), "valid_equal": bool(np.array_equal(gv, dst.Valid)),
              "against": "the CPU baseline's Slope of the whole DEM"}
    cpu = {"value": round(n / s_s / 1e9, 4), "unit": "Gcells/s", "cores": threads, "kind": "port",
           "sample": f"the whole {N}x{N} DEM, terrain.Slope (CellSize 30, degrees) through a C restatement of the "
                     f"reference's AVX2 Tiled path, {threads} threads, {s_reps} timed passes",
           "hillshade": {"value": round(n / h_s / 1e9, 4), "unit": "Gcells/s", "ms": round(h_s * 1e3, 2),
                         "passes": h_reps},
           "slope_ms": round(s_s * 1e3, 2)}
    cpu["plain_4096"] = plain
    return cpu, parity


def plain_4096(slope, log):
    """The reference's CPU-runnable configuration: plain terrain.Slope
    (CellSize 30, degrees) on a 4096x4096 fBm DEM with 1% of cells NoData
    i.i.d. The CPU baseline (the reference's AVX2 path restated) runs it on
    one thread, as the plain path does, beside the reference's published
    one-core figure; the device runs the same call (strata_process); the
    two outputs, data and validity, are compared bit for bit over the whole
    raster."""
    import numpy as np
    import torch

    from oracle import oracle
    from strata_amd import synth, terrain
    from strata_amd.raster import Float32Raster

    n = 4096
    dev = torch.device("cuda:0")
    data = synth.fbm_dem(n, n, hurst=0.8, seed=3, lo=0.0, hi=3000.0, device=dev)
    valid = synth.random_nodata(n * n, 0.01, seed=4, device=dev)
    words = valid.numel()
    dem = Float32Raster(data, n, n, n, valid, 0)
    out = Float32Raster(torch.empty(n * n, dtype=torch.float32, device=dev), n, n, n,
                        torch.zeros(words, dtype=torch.int64, device=dev), 0)
    opts = terrain.SlopeOptions(CellSize=30)
    for _ in range(3):
        terrain.Slope(out, dem, opts)
    torch.cuda.synchronize()
    reps = 50
    t = time.perf_counter()
    for _ in range(reps):
        terrain.Slope(out, dem, opts)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t) / reps

    src = Float32Raster(data.cpu().numpy().copy(), n, n, n, valid.cpu().numpy().view(np.uint64).copy(), 0)
    dst = Float32Raster(np.zeros(n * n, np.float32), n, n, n, np.zeros(words, np.uint64), 0)
    oracle.tiled_avx2(slope, [dst], [src], 0, 0, 1)
    cpu_reps = 10
    t = time.perf_counter()
    for _ in range(cpu_reps):
        oracle.tiled_avx2(slope, [dst], [src], 0, 0, 1)
    cpu_s = (time.perf_counter() - t) / cpu_reps

    got = out.Data.cpu().numpy()
    same = (got.view(np.uint32) == dst.Data.view(np.uint32)) | (np.isnan(got) & np.isnan(dst.Data))
    gv = out.Valid.cpu().numpy().view(np.uint64)
    res = {"workload": "plain terrain.Slope, 4096x4096 fBm DEM (Hurst 0.8, 0-3000 m), 1% NoData i.i.d., CellSize 30",
           "cpu_1_thread_s": round(cpu_s, 4), "cpu_1_thread_gcells_s": round(n * n / cpu_s / 1e9, 4),
           "cpu_kind": "the reference's AVX2 path restated in C, one thread",
           "reference_published_1_core_gcells_s": 0.688,
           "gpu_ms": round(gpu_s * 1e3, 4), "gpu_gcells_s": round(n * n / gpu_s / 1e9, 2),
           "gpu_over_cpu": round(cpu_s / gpu_s, 1),
           "identical": bool(same.all() and np.array_equal(gv, dst.Valid))}
    log(f"bench: plain Slope 4096^2: CPU AVX2 1 thread {cpu_s * 1e3:.1f} ms ({n * n / cpu_s / 1e9:.3f} Gcells/s), "
        f"GPU {gpu_s * 1e3:.3f} ms, identical {res['identical']}")
    return res

if __name__ == "__main__":
    sys.exit(main())
