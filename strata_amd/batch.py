"""Batches: many tiles with mixed ops in one persistent device launch.

Each task is what exec.ProcessN (internal/exec/process.go:28) does for one
tile on its own rasters (its own NaN edge ring, terrain/doc.go:89-90),
followed by point ops (transfer.Lookup, transfer.Reclass) applied to the
task's output in place, as the reference's chain examples write them
(transfer/example_test.go:90-129). The batch changes only the scheduling:
every output is bit-identical to those calls (tests/test_gpu_batch.py).

The whole batch runs as one persistent kernel that claims work units from a
device counter, optionally captured in a HIP graph; `run_per_tile` runs the
same tasks through the plain path, one launch set per tile, as the baseline
the batch is measured against.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

from . import _lib
from .engine import EnginePanic, KernelDesc, Options, _raise_status, _stream_handle, check
from .raster import Float32Raster

# Ops a batch runs (the C ABI checks the same and reports the rest).
STENCIL_OPS = (_lib.OP_SLOPE, _lib.OP_ASPECT, _lib.OP_HILLSHADE, _lib.OP_CURVATURE, _lib.OP_MEAN,
               _lib.OP_MIN, _lib.OP_MAX)
POINT_OPS = (_lib.OP_LOOKUP, _lib.OP_RECLASS)


@dataclass
class Task:
    """One tile: k over src into dst, then each post op over dst in place."""
    k: KernelDesc
    src: Float32Raster
    dst: Float32Raster
    post: Sequence[KernelDesc] = field(default_factory=tuple)


class Batch:
    """A device work list over tasks. The caller keeps the rasters alive
    while the batch is used; the kernels' tables are copied."""

    def __init__(self, tasks: Sequence[Task], *, rows: int = 0, blocks_per_cu: int = 0, graph: bool = False,
                 stream: Optional[int] = None):
        if not tasks:
            raise EnginePanic("strata: a batch needs at least one task")
        for t in tasks:
            check([t.dst], [t.src], t.k, Options())
            for p in t.post:
                check([t.dst], [t.dst], p, Options())
            if not (t.src.on_device and t.dst.on_device):
                raise EnginePanic("strata: batch rasters must be in device memory")
        self.tasks = list(tasks)
        arr = (_lib.Task * len(tasks))()
        keep = []
        for i, t in enumerate(tasks):
            kc, taps = t.k.to_c()
            keep.append(taps)
            arr[i].k = kc
            if len(t.post) > 2:
                raise EnginePanic("strata: a task takes at most two point ops")
            for j, p in enumerate(t.post):
                pc, ptaps = p.to_c()
                keep.append(ptaps)
                arr[i].post[j] = pc
            arr[i].npost = len(t.post)
            arr[i].src = t.src.to_c()
            arr[i].dst = t.dst.to_c()
        o = _lib.BatchOptions()
        o.rows, o.blocks_per_cu, o.graph = rows, blocks_per_cu, 1 if graph else 0
        if stream is None:
            stream = _stream_handle([tasks[0].src])
        o.stream = stream or None
        h = ctypes.c_void_p()
        rc = _lib.lib().strata_batch_create(arr, len(tasks), ctypes.byref(o), ctypes.byref(h))
        del keep
        if rc < 0:
            _raise_status(rc)
        self._h = h

    def _call(self, name: str, *args) -> None:
        if self._h is None:
            raise EnginePanic("strata: batch is closed")
        rc = getattr(_lib.lib(), name)(self._h, *args)
        if rc < 0:
            _raise_status(rc)

    def run(self) -> None:
        """Enqueue the whole batch (one launch, or its graph)."""
        self._call("strata_batch_run")

    def run_per_tile(self) -> None:
        """Enqueue the same tasks through the plain path, call by call."""
        self._call("strata_batch_run_per_tile")

    def sync(self) -> None:
        self._call("strata_batch_sync")

    def tune(self, reps: int = 3) -> dict:
        """Time each candidate launch shape and keep the fastest."""
        o = _lib.BatchOptions()
        self._call("strata_batch_tune", reps, ctypes.byref(o))
        return {"rows": o.rows, "blocks_per_cu": o.blocks_per_cu, "graph": bool(o.graph)}

    def info(self) -> dict:
        i = _lib.BatchInfo()
        self._call("strata_batch_info_get", ctypes.byref(i))
        return {"tasks": i.tasks, "cells": i.cells, "data_units": i.data_units,
                "validity_units": i.validity_units, "rows": i.rows, "blocks": i.blocks, "graph": bool(i.graph)}

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().strata_batch_destroy(self._h)
            self._h = None

    def __enter__(self) -> "Batch":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ProcessBatch(tasks: Sequence[Task], **kw) -> None:
    """Run tasks as one batch and wait for it."""
    with Batch(tasks, **kw) as b:
        b.run()
        b.sync()
