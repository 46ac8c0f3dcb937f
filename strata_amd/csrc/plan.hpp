// plan.hpp — the engine's plan of bands and tiles, its Stats in closed form,
// and the units a cancellable call claims: plain C++ (no HIP), shared by the
// runtime and the host tests.
//
// Reference behaviour restated here:
//   plan and bands   internal/exec/tile.go:298-359
//   band accounting  internal/exec/halo.go:21-57,97-105; worker.go:59-68
//   chunked tiles    internal/exec/chunked.go:159-174,335-374
#pragma once

#include <stdint.h>

#include <algorithm>

#include "strata_hip.h"

namespace strata {

constexpr int64_t kBytesPerCell = 4;
constexpr int64_t kBandCells = 1 << 16;  // internal/exec/tile.go:27

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct Plan {
    int64_t w = 0, h = 0, tileW = 0, tileH = 0, bandRows = 0, tilesX = 0, tileRows = 0, perTile = 0,
            perLastTile = 0, bands = 0;
};

inline Plan new_plan(int64_t w, int64_t h, int64_t tileW, int64_t tileH)
{
    Plan p;
    p.w = w;
    p.h = h;
    p.tileW = tileW > 0 ? std::min(tileW, w) : w;
    p.tileH = tileH > 0 ? std::min(tileH, h) : h;
    if (w == 0 || h == 0) return p;
    p.bandRows = std::max<int64_t>(1, kBandCells / p.tileW);
    p.tilesX = cdiv(w, p.tileW);
    p.tileRows = cdiv(h, p.tileH);
    p.perTile = cdiv(p.tileH, p.bandRows);
    p.perLastTile = cdiv(h - (p.tileRows - 1) * p.tileH, p.bandRows);
    p.bands = p.tilesX * ((p.tileRows - 1) * p.perTile + p.perLastTile);
    return p;
}

struct Rect {
    int64_t x0, y0, x1, y1;
};

// Band i of the plan: its tile row, tile and band within the tile.
struct BandPos {
    int64_t tr, tx, b, per;
};

inline BandPos band_pos(const Plan &p, int64_t i)
{
    const int64_t perRow = p.tilesX * p.perTile;
    BandPos q{i / perRow, 0, 0, p.perTile};
    int64_t j = i % perRow;
    if (q.tr >= p.tileRows - 1) {
        q.tr = p.tileRows - 1;
        j = i - (p.tileRows - 1) * perRow;
        q.per = p.perLastTile;
    }
    q.tx = j / q.per;
    q.b = j % q.per;
    return q;
}

inline Rect plan_band(const Plan &p, int64_t i)
{
    const BandPos q = band_pos(p, i);
    Rect r;
    r.x0 = q.tx * p.tileW;
    r.x1 = std::min(r.x0 + p.tileW, p.w);
    const int64_t ty = q.tr * p.tileH;
    r.y0 = ty + q.b * p.bandRows;
    r.y1 = std::min(std::min(r.y0 + p.bandRows, ty + p.tileH), p.h);
    return r;
}

// A unit of a cancellable call: bands [b0, b1) of the plan, whose union is
// the rectangle rc. Units are claimed in order, so the bands of the units a
// call finishes are a prefix of the plan (internal/exec/doc.go:165-183).
struct Unit {
    Rect rc;
    int64_t b0, b1;
};

// The unit that starts at band i, of about `target` cells or fewer bands
// when the plan's shape ends it sooner: at a tile's first band, whole tiles
// of that tile row (and, from a row's first tile, whole tile rows below once
// the row is taken); past a tile's first band, or when one tile alone holds
// `target` cells, the tile's next bands top to bottom. Every unit is a
// rectangle made of consecutive bands.
inline Unit plan_unit(const Plan &p, int64_t i, int64_t target)
{
    const BandPos q = band_pos(p, i);
    const int64_t ty = q.tr * p.tileH, th = std::min(p.tileH, p.h - ty);
    auto tile_w = [&](int64_t tx) { return std::min(p.tileW, p.w - tx * p.tileW); };
    Unit u;
    u.b0 = i;
    if (q.b == 0 && tile_w(q.tx) * th < target) {
        int64_t k = 0, cells = 0;
        while (q.tx + k < p.tilesX && cells < target) cells += tile_w(q.tx + k++) * th;
        u.rc = Rect{q.tx * p.tileW, ty, std::min((q.tx + k) * p.tileW, p.w), ty + th};
        u.b1 = i + k * q.per;
        if (q.tx == 0 && k == p.tilesX) {
            for (int64_t tr = q.tr + 1; tr < p.tileRows && cells < target; ++tr) {
                const int64_t y = tr * p.tileH, h = std::min(p.tileH, p.h - y);
                cells += p.w * h;
                u.rc.y1 = y + h;
                u.b1 += p.tilesX * (tr == p.tileRows - 1 ? p.perLastTile : p.perTile);
            }
        }
        return u;
    }
    u.rc = plan_band(p, i);
    u.b1 = i + 1;
    for (int64_t b = q.b + 1; b < q.per && (u.rc.x1 - u.rc.x0) * (u.rc.y1 - u.rc.y0) < target; ++b)
        u.rc.y1 = plan_band(p, u.b1++).y1;
    return u;
}

// One band's counters (internal/exec/halo.go:22-39, 97-105), for the
// band [x0, x1) x [y0, y1) of a W x H raster: the kernel runs over the
// cells outside the narrowest edge ring (rmin) and reads its window grown
// by the radius r.
inline void band_stats(strata_stats &s, int r, int rmin, int64_t W, int64_t H, int nin, int nout, const Rect &b)
{
    const int64_t cells = (b.x1 - b.x0) * (b.y1 - b.y0);
    const int64_t ix0 = std::max<int64_t>(b.x0, rmin), ix1 = std::min<int64_t>(b.x1, W - rmin);
    const int64_t iy0 = std::max<int64_t>(b.y0, rmin), iy1 = std::min<int64_t>(b.y1, H - rmin);
    const bool interior = ix0 < ix1 && iy0 < iy1;
    const int64_t inner = interior ? (ix1 - ix0) * (iy1 - iy0) : 0;
    s.cells += cells;
    s.bands += 1;
    s.kernel_written += (cells - inner) * kBytesPerCell * nout;
    if (interior) {
        s.kernel_read += (ix1 - ix0 + 2 * r) * (iy1 - iy0 + 2 * r) * kBytesPerCell * nin;
        s.kernel_written += inner * kBytesPerCell * nout;
    }
}

inline void add_stats(strata_stats *out, const strata_stats &s, int nin, int nout)
{
    if (!out) return;
    out->cells += s.cells;
    out->ideal += s.cells * kBytesPerCell * (nin + nout);
    out->tiles += s.tiles;
    out->bands += s.bands;
    out->source_read += s.source_read;
    out->sink_written += s.sink_written;
    out->kernel_read += s.kernel_read;
    out->kernel_written += s.kernel_written;
}

// The chunked tile grid (internal/exec/chunked.go:159-174).
struct TileGrid {
    int64_t W, H, tileW, tileH, tilesX, tiles;
};

inline TileGrid tile_grid(int64_t W, int64_t H, const strata_options *o)
{
    TileGrid g;
    g.W = W;
    g.H = H;
    g.tileW = (o && o->tile_width > 0) ? std::min(o->tile_width, W) : W;
    g.tileH = (o && o->tile_height > 0) ? std::min(o->tile_height, H) : H;
    g.tilesX = cdiv(W, g.tileW);
    g.tiles = g.tilesX * cdiv(H, g.tileH);
    return g;
}

inline Rect tile_rect(const TileGrid &g, int64_t i)
{
    Rect t;
    t.x0 = (i % g.tilesX) * g.tileW;
    t.y0 = (i / g.tilesX) * g.tileH;
    t.x1 = std::min(t.x0 + g.tileW, g.W);
    t.y1 = std::min(t.y0 + g.tileH, g.H);
    return t;
}

// One chunked tile's counters (chunked.go:335-374): the source window,
// the sink tile, and the tile's own plan of bands.
inline void tile_stats(strata_stats &s, int r, int rmin, const TileGrid &g, int nin, int nout, const Rect &t)
{
    const int64_t rx0 = std::max<int64_t>(0, t.x0 - r), ry0 = std::max<int64_t>(0, t.y0 - r);
    const int64_t rx1 = std::min<int64_t>(g.W, t.x1 + r), ry1 = std::min<int64_t>(g.H, t.y1 + r);
    s.tiles += 1;
    s.source_read += (rx1 - rx0) * (ry1 - ry0) * kBytesPerCell * nin;
    const Plan p = new_plan(t.x1 - t.x0, t.y1 - t.y0, 0, 0);
    for (int64_t b = 0; b < p.bands; ++b) {
        Rect band = plan_band(p, b);
        band.x0 += t.x0;
        band.x1 += t.x0;
        band.y0 += t.y0;
        band.y1 += t.y0;
        band_stats(s, r, rmin, g.W, g.H, nin, nout, band);
    }
    s.sink_written += (t.x1 - t.x0) * (t.y1 - t.y0) * kBytesPerCell * nout;
}

} // namespace strata
