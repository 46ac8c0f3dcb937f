"""Synthetic inputs of the benchmark configurations (seeded, on any torch
device): diamond-square fBm DEMs, i.i.d. and blob NoData masks, and
uniform focal rasters.

The reference benchmarks its kernels on sinusoid DEMs
(benchmarks/internal/suite/suite.go:322-342); these configurations ask for
fractal terrain instead. Diamond-square on a (2^n+1)^2 grid: each level
sets square centres to the mean of their four corners and edge midpoints
to the mean of their (three or four) neighbours, plus a uniform
displacement in [-a, a) whose amplitude a shrinks by 2^-H per level
(H = Hurst exponent). The grid is cropped to the requested size and
scaled affinely to [lo, hi].
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np
import torch


def _uniform(shape, gen, device, amp: float) -> torch.Tensor:
    return (torch.rand(shape, generator=gen, device=device, dtype=torch.float32) * 2 - 1) * amp


def fbm_grid(levels: int, hurst: float, seed: int, device="cuda") -> torch.Tensor:
    """Diamond-square fBm on a (2^levels + 1)^2 float32 grid."""
    n = (1 << levels) + 1
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    g = torch.zeros((n, n), dtype=torch.float32, device=device)
    step = n - 1
    g[0::step, 0::step] = _uniform((2, 2), gen, device, 1.0)
    amp = 1.0
    decay = 2.0 ** (-hurst)
    while step > 1:
        half = step // 2
        amp *= decay
        # Diamond step: centres of squares.
        c = (g[0:-1:step, 0:-1:step] + g[0:-1:step, step::step] + g[step::step, 0:-1:step] +
             g[step::step, step::step]) * 0.25
        g[half::step, half::step] = c + _uniform(c.shape, gen, device, amp)
        centres = g[half::step, half::step]          # (m-1) x (m-1), m = corners per side
        # Square step (a): rows on the corner grid, columns at midpoints.
        left = g[0::step, 0:-1:step]
        right = g[0::step, step::step]
        s = left + right
        cnt = torch.full_like(s, 2.0)
        s[1:] += centres
        cnt[1:] += 1
        s[:-1] += centres
        cnt[:-1] += 1
        g[0::step, half::step] = s / cnt + _uniform(s.shape, gen, device, amp)
        # Square step (b): rows at midpoints, columns on the corner grid.
        up = g[0:-1:step, 0::step]
        down = g[step::step, 0::step]
        s = up + down
        cnt = torch.full_like(s, 2.0)
        s[:, 1:] += centres
        cnt[:, 1:] += 1
        s[:, :-1] += centres
        cnt[:, :-1] += 1
        g[half::step, 0::step] = s / cnt + _uniform(s.shape, gen, device, amp)
        step = half
    return g


def fbm_dem(width: int, height: int, hurst: float = 0.8, seed: int = 1, lo: float = 0.0, hi: float = 3000.0,
            device="cuda") -> torch.Tensor:
    """A width×height float32 fBm DEM (compact, flat), scaled to [lo, hi]."""
    levels = max(1, math.ceil(math.log2(max(width, height) - 1))) if max(width, height) > 2 else 1
    g = fbm_grid(levels, hurst, seed, device)[:height, :width]
    mn, mx = torch.aminmax(g)
    scale = (hi - lo) / float(mx - mn) if float(mx - mn) > 0 else 0.0
    out = ((g - mn) * scale + lo).contiguous().reshape(-1)
    del g
    return out


def pack_mask(valid: torch.Tensor) -> torch.Tensor:
    """Pack a flat bool tensor into LSB-first 64-bit words (int64 storage)."""
    n = valid.numel()
    words = (n + 63) // 64
    pad = torch.zeros(words * 64, dtype=torch.bool, device=valid.device)
    pad[:n] = valid.reshape(-1)
    b = pad.view(words, 64).to(torch.int64)
    shifts = torch.arange(64, device=valid.device, dtype=torch.int64)
    return (b << shifts).sum(dim=1)


def random_nodata(n: int, frac: float, seed: int, device="cuda") -> torch.Tensor:
    """Validity words with each of n cells invalid with probability frac."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    valid = torch.rand(n, generator=gen, device=device) >= frac
    return pack_mask(valid)


def blob_nodata(width: int, height: int, frac: float = 0.005, rmin: float = 5, rmax: float = 50, seed: int = 2,
                device="cuda") -> Tuple[torch.Tensor, float]:
    """Validity words with circular NoData blobs of radius U[rmin, rmax] at
    uniform centres, added until at least `frac` of the cells are covered.
    Returns (words, covered fraction)."""
    rng = np.random.default_rng(seed)
    invalid = torch.zeros((height, width), dtype=torch.bool, device=device)
    target = frac * width * height
    covered = 0
    mean_area = math.pi * (rmin * rmin + rmin * rmax + rmax * rmax) / 3
    while covered < target:
        batch = int(min(64, max(1, (target - covered) / mean_area)))
        for _ in range(batch):
            r = rng.uniform(rmin, rmax)
            cx, cy = rng.uniform(0, width), rng.uniform(0, height)
            x0, x1 = max(0, int(math.floor(cx - r))), min(width, int(math.ceil(cx + r)) + 1)
            y0, y1 = max(0, int(math.floor(cy - r))), min(height, int(math.ceil(cy + r)) + 1)
            if x0 >= x1 or y0 >= y1:
                continue
            ys = torch.arange(y0, y1, device=device, dtype=torch.float32).view(-1, 1) + 0.5
            xs = torch.arange(x0, x1, device=device, dtype=torch.float32).view(1, -1) + 0.5
            disk = (xs - cx) ** 2 + (ys - cy) ** 2 <= r * r
            invalid[y0:y1, x0:x1] |= disk
        covered = int(invalid.sum())
    words = pack_mask(~invalid.reshape(-1))
    return words, covered / (width * height)


def uniform_raster(n: int, lo: float, hi: float, seed: int, device="cuda") -> torch.Tensor:
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    return torch.rand(n, generator=gen, device=device, dtype=torch.float32) * (hi - lo) + lo
