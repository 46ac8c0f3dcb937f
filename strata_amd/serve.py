"""Serving: window requests against a device-resident raster.

A request is a window of the resident raster; its result is the chain (a
stencil op, exec.ProcessN on the window as a raster of its own,
internal/exec/process.go:28, then point ops such as transfer.Lookup and
transfer.Reclass) on that window. The library's scheduler admits requests
in arrival order into batches, one persistent launch each, as batch slots
and result memory (an HBM budget) free up, with several batches in flight
on their own streams. Results are bit-identical whether a request runs
alone or among others (tests/test_gpu_serve.py).
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .engine import EnginePanic, KernelDesc, _raise_status
from .raster import Float32Raster


class Server:
    def __init__(self, src: Float32Raster, k: KernelDesc, post: Sequence[KernelDesc] = (), *, slots: int = 4,
                 hbm_budget: int = 4 << 30, max_batch: int = 256, rows: int = 0, auto_release: bool = False):
        if not src.on_device:
            raise EnginePanic("strata: the served raster must be in device memory")
        if len(post) > 2:
            raise EnginePanic("strata: a chain takes at most two point ops")
        self.src = src
        c = _lib.ServeConfig()
        c.src = src.to_c()
        kc, t0 = k.to_c()
        c.k = kc
        keep = [t0]
        for j, p in enumerate(post):
            pc, pt = p.to_c()
            keep.append(pt)
            c.post[j] = pc
        c.npost = len(post)
        c.slots, c.hbm_budget, c.max_batch, c.rows = slots, hbm_budget, max_batch, rows
        c.auto_release = 1 if auto_release else 0
        h = ctypes.c_void_p()
        rc = _lib.lib().strata_serve_create(ctypes.byref(c), ctypes.byref(h))
        del keep
        if rc < 0:
            _raise_status(rc)
        self._h = h

    def _check(self, rc: int) -> None:
        if rc < 0:
            _raise_status(rc)

    def submit(self, x: int, y: int, w: int, h: int) -> int:
        win = _lib.Window(x, y, w, h)
        rid = ctypes.c_int64()
        self._check(_lib.lib().strata_serve_submit(self._h, ctypes.byref(win), ctypes.byref(rid)))
        return rid.value

    def wait(self, rid: int) -> float:
        """Block until request rid completes; its latency in ms."""
        ms = ctypes.c_double()
        self._check(_lib.lib().strata_serve_wait(self._h, rid, ctypes.byref(ms)))
        return ms.value

    def fetch(self, rid: int, w: int, h: int) -> Tuple[np.ndarray, np.ndarray]:
        """A completed result: (w*h float32 cells, mask words; cell i at bit i)."""
        data = np.empty(w * h, np.float32)
        mask = np.empty((w * h + 63) // 64, np.uint64)
        self._check(_lib.lib().strata_serve_fetch(self._h, rid, data.ctypes.data, mask.ctypes.data))
        return data, mask

    def release(self, rid: int) -> None:
        self._check(_lib.lib().strata_serve_release(self._h, rid))

    def replay(self, windows: np.ndarray, arrivals: np.ndarray) -> np.ndarray:
        """Open loop: submit windows[i] (x, y, w, h) at arrivals[i] seconds
        after the call starts; returns each request's latency in ms, from
        its scheduled arrival."""
        n = len(windows)
        wa = (_lib.Window * n)()
        for i, (x, y, w, h) in enumerate(np.asarray(windows, dtype=np.int64)):
            wa[i] = _lib.Window(int(x), int(y), int(w), int(h))
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.zeros(n, np.float64)
        self._check(_lib.lib().strata_serve_replay(
            self._h, wa, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
            lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return lat

    def stats(self) -> Dict[str, float]:
        s = _lib.ServeStats()
        self._check(_lib.lib().strata_serve_stats_get(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in _lib.ServeStats._fields_ if f != "reserved"}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _lib.lib().strata_serve_destroy(self._h)
            self._h = None

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_schedule(qps: float, n: int, src_w: int, src_h: int, seed: int,
                     mix=((0.80, 512), (0.15, 2048), (0.05, 4096))) -> Tuple[np.ndarray, np.ndarray]:
    """n requests arriving as a seeded Poisson process at qps: windows
    (x, y, w, h) with sides drawn from mix (probability, side), at uniform
    offsets inside the raster, and arrival times in seconds."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    probs = np.array([p for p, _ in mix])
    sides = np.array([s for _, s in mix])
    side = sides[rng.choice(len(mix), size=n, p=probs / probs.sum())]
    w = np.minimum(side, src_w)
    h = np.minimum(side, src_h)
    x = (rng.random(n) * (src_w - w + 1)).astype(np.int64)
    y = (rng.random(n) * (src_h - h + 1)).astype(np.int64)
    return np.stack([x, y, w, h], axis=1), arrivals
