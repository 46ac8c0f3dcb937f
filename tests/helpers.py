"""Operand builders and comparisons shared by the tests.

A Scene is a set of parent buffers (host numpy arrays) and windows into
them, so the same operands can be materialised on the host (for the
oracle) and on the device (for the HIP library) with identical layouts:
strides wider than rows, windows at odd offsets, validity offsets in the
middle of words, stale bits and poisoned padding around the outputs.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from strata_amd.raster import Float32Raster, MaskWords, pack_bits

F32_MAX = np.float32(3.4028234663852886e38)
SUBNORMAL = np.float32(1e-40)


def hazard_values(rng: np.random.Generator, n: int, base: float = 500.0, relief: float = 50.0,
                  hazard_frac: float = 0.02) -> np.ndarray:
    """Terrain-like values with a sprinkle of hazards: NaN, ±Inf, ±0,
    ±MaxFloat32 and subnormals."""
    v = (base + rng.standard_normal(n) * relief).astype(np.float32)
    hz = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, F32_MAX, -F32_MAX, SUBNORMAL, -SUBNORMAL, 1e30, -1e-30],
                  dtype=np.float32)
    pick = rng.random(n) < hazard_frac
    v[pick] = rng.choice(hz, size=int(pick.sum()))
    return v


def plane_values(w: int, h: int, ax: float, ay: float, cs: float, csy: float) -> np.ndarray:
    """terrain/terrain_test.go:28-36 plane()."""
    d = np.zeros(w * h, dtype=np.float32)
    for y in range(h):
        for x in range(w):
            d[y * w + x] = np.float32(ax * float(x) * cs + ay * float(y) * csy + 100)
    return d


@dataclass
class Parent:
    data: np.ndarray                   # float32, flat
    valid: Optional[np.ndarray] = None  # uint64 words


@dataclass
class Win:
    parent: str
    x: int
    y: int
    w: int
    h: int
    width: int  # the parent's logical width (its stride)
    valid_base: int = 0  # ValidOffset of the parent's cell (0, 0)


@dataclass
class Scene:
    parents: Dict[str, Parent] = field(default_factory=dict)
    wins: Dict[str, Win] = field(default_factory=dict)

    def add_parent(self, name: str, rng: np.random.Generator, stride: int, rows: int, masked: bool,
                   valid_base: int = 0, invalid_frac: float = 0.05, values=None, stale: bool = True) -> None:
        n = stride * rows
        data = values if values is not None else hazard_values(rng, n)
        data = np.ascontiguousarray(data, dtype=np.float32).copy()
        valid = None
        if masked:
            bits = rng.random(valid_base + n + 64) >= invalid_frac
            if not stale:
                bits[:] = True
            valid = pack_bits(bits, valid_base + n + 64)
        self.parents[name] = Parent(data, valid)
        self._bases = getattr(self, "_bases", {})
        self._bases[name] = (stride, valid_base)

    def add_window(self, name: str, parent: str, x: int, y: int, w: int, h: int) -> None:
        stride, base = self._bases[parent]
        self.wins[name] = Win(parent, x, y, w, h, stride, base)

    def materialize(self, device=None, masked: Optional[Dict[str, bool]] = None):
        """(rasters by window name, parent buffers by name) on the host
        (device None) or on a torch device."""
        bufs = {}
        for name, p in self.parents.items():
            if device is None:
                bufs[name] = Parent(p.data.copy(), None if p.valid is None else p.valid.copy())
            else:
                import torch
                bufs[name] = Parent(torch.from_numpy(p.data.copy()).to(device),
                                    None if p.valid is None else torch.from_numpy(p.valid.view(np.int64).copy()).to(device))
        rasters = {}
        for name, w in self.wins.items():
            p = bufs[w.parent]
            start = w.y * w.width + w.x
            end = start + (w.h - 1) * w.width + w.w
            valid = p.valid
            use_mask = valid is not None and (masked is None or masked.get(name, True))
            rasters[name] = Float32Raster(p.data[start:end], w.w, w.h, w.width, valid if use_mask else None,
                                          (w.valid_base + start) if use_mask else 0)
        return rasters, bufs


def host_copy(bufs: Dict[str, Parent]) -> Dict[str, Parent]:
    out = {}
    for k, p in bufs.items():
        if hasattr(p.data, "detach"):
            out[k] = Parent(p.data.detach().cpu().numpy().copy(),
                            None if p.valid is None else p.valid.detach().cpu().numpy().view(np.uint64).copy())
        else:
            out[k] = Parent(p.data.copy(), None if p.valid is None else p.valid.copy())
    return out


def same_floats(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Elementwise bit equality, any NaN matching any NaN (+0 != -0)."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint32) == b.view(np.uint32)) | both_nan


def assert_same_parents(got: Dict[str, Parent], want: Dict[str, Parent], what: str = "") -> None:
    for name in want:
        g, w = got[name], want[name]
        eq = same_floats(g.data, w.data)
        if not eq.all():
            i = int(np.argmin(eq))
            raise AssertionError(f"{what}: parent {name} data differs at flat index {i}: "
                                 f"got {g.data[i]!r} ({g.data[i:i+1].view(np.uint32)[0]:#x}), "
                                 f"want {w.data[i]!r} ({w.data[i:i+1].view(np.uint32)[0]:#x}); "
                                 f"{int((~eq).sum())} cells differ")
        if w.valid is not None:
            if not np.array_equal(g.valid, w.valid):
                diff = np.nonzero(g.valid != w.valid)[0]
                k = int(diff[0])
                raise AssertionError(f"{what}: parent {name} mask word {k} differs: got {int(g.valid[k]):#018x}, "
                                     f"want {int(w.valid[k]):#018x}; {len(diff)} words differ")


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in float32 ulps (ordered integer representation); NaN pairs 0."""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    d = np.abs(ia - ib)
    both_nan = np.isnan(a) & np.isnan(b)
    return np.where(both_nan, 0, d)
