// kernels.hip — gfx950 kernels of the raster stencil path.
//
// Three data kernels and one validity kernel:
//
//   horn_sweep   radius-1 terrain ops (Horn gradient, slope, aspect,
//                hillshade; Zevenbergen-Thorne curvature). One lane per
//                column, each wave walks down a strip of rows holding the
//                3x3 window in registers; the left and right neighbours
//                come from the adjacent lanes through DPP wave shifts, the
//                wave's two outer columns from one extra load. Every input
//                cell is loaded once per strip (plus two halo rows).
//   focal_tile   radius-r ops (focal Correlate/Convolve, CorrelateSeparable,
//                Mean, Min, Max; terrain ops through Wood's fit). The tile
//                and its halo are staged in LDS; separable forms run the
//                column pass into LDS and the row pass from it.
//   point_rows   radius-0 transfer ops (Lookup, Reclass).
//   validity     output validity, word by word: AND of the (2r+1) source
//                rows, horizontal erosion by shifts, interior mask; whole
//                words are stored, words shared with cells outside the
//                region are updated with 64-bit atomics on the region's
//                own bits only.
//
// The arithmetic is arith.hpp's, in the reference's order.
#include "arith.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace strata {

// ---------------------------------------------------------------------------
// Shared helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ float load_cell(const SrcView &s, int64_t x, int64_t y)
{
    bool in = x >= s.lx0 && x < s.lx1 && y >= s.ly0 && y < s.ly1;
    return in ? s.p[(y - s.sy) * s.stride + (x - s.sx)] : 0.0f;
}

// v from lane-1; lane 0 gets `edge` (the wave's left halo column).
__device__ __forceinline__ float from_left(float v, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), 0x138 /* wave_shr:1 */,
                                                                 0xf, 0xf, false));
}
// v from lane+1; lane 63 gets `edge` (the wave's right halo column).
__device__ __forceinline__ float from_right(float v, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), 0x130 /* wave_shl:1 */,
                                                                 0xf, 0xf, false));
}

// ---------------------------------------------------------------------------
// horn_sweep: radius-1 terrain kernels
// ---------------------------------------------------------------------------

constexpr int kSweepThreads = 256;
constexpr int kSweepRows = 64;

struct HornArgs {
    float k[8];
    int32_t flags;
};

template <int OP>
__device__ __forceinline__ void horn_cell(const HornArgs &a, float z1, float z2, float z3, float z4, float z5,
                                          float z6, float z7, float z8, float z9, float &o0, float &o1)
{
    if constexpr (OP == STRATA_OP_CURVATURE) {
        o0 = zt_curvature(z1, z2, z3, z4, z5, z6, z7, z8, z9, a.k, (a.flags >> 4) & 3);
    } else {
        float gx = horn_dx(z1, z3, z4, z6, z7, z9) * a.k[0];
        float gy = horn_dy(z1, z2, z3, z7, z8, z9) * a.k[1];
        if constexpr (OP == STRATA_OP_GRADIENT) {
            o0 = gx;
            o1 = gy;
        } else if constexpr (OP == STRATA_OP_SLOPE) {
            o0 = slope_of(gx, gy, a.k[2], (a.flags & 1) != 0);
        } else if constexpr (OP == STRATA_OP_ASPECT) {
            o0 = aspect_of(gx, gy, a.k[2], (a.flags & 2) != 0);
        } else {
            o0 = shade_of(gx, gy, a.k[2], a.k[3], a.k[4]);
        }
    }
}

template <int OP>
__global__ __launch_bounds__(kSweepThreads) void horn_sweep(HornArgs a, Region rg, SrcView s, DstView d0,
                                                            DstView d1)
{
    const int lane = threadIdx.x & 63;
    const int64_t x = rg.ox + (int64_t)blockIdx.x * kSweepThreads + threadIdx.x;
    const int64_t hx = lane == 0 ? x - 1 : x + 1;
    const bool halo_lane = lane == 0 || lane == 63;
    const int64_t y0 = rg.oy + (int64_t)blockIdx.y * kSweepRows;
    const int64_t y1 = min(y0 + (int64_t)kSweepRows, rg.oy + rg.h);
    const bool col_in = x >= rg.ox && x < rg.ox + rg.w;
    const bool col_interior = x >= 1 && x < rg.W - 1;

    // Rows y-1, y, y+1 of the window: centre value, left and right.
    float cm = load_cell(s, x, y0 - 1);
    float hm = halo_lane ? load_cell(s, hx, y0 - 1) : 0.0f;
    float c0 = load_cell(s, x, y0);
    float h0 = halo_lane ? load_cell(s, hx, y0) : 0.0f;
    float cn = load_cell(s, x, y0 + 1);
    float hn = halo_lane ? load_cell(s, hx, y0 + 1) : 0.0f;
    float lm = from_left(cm, hm), rm = from_right(cm, hm);
    float l0 = from_left(c0, h0), r0 = from_right(c0, h0);

    for (int64_t y = y0; y < y1; ++y) {
        const float cp = cn, hp = hn;
        // Issue the next row's loads before this row's arithmetic.
        if (y + 2 <= y1) {
            cn = load_cell(s, x, y + 2);
            hn = halo_lane ? load_cell(s, hx, y + 2) : 0.0f;
        }
        const float lp = from_left(cp, hp), rp = from_right(cp, hp);
        float o0 = edge_nan(), o1 = edge_nan();
        if (col_interior && y >= 1 && y < rg.H - 1) {
            horn_cell<OP>(a, lm, cm, rm, l0, c0, r0, lp, cp, rp, o0, o1);
        }
        if (col_in) {
            d0.p[(y - rg.oy) * d0.stride + (x - rg.ox)] = o0;
            if constexpr (OP == STRATA_OP_GRADIENT) d1.p[(y - rg.oy) * d1.stride + (x - rg.ox)] = o1;
        }
        lm = l0; cm = c0; rm = r0;
        l0 = lp; c0 = cp; r0 = rp;
    }
}

// ---------------------------------------------------------------------------
// focal_tile: radius-r kernels through LDS
// ---------------------------------------------------------------------------

constexpr int kTileX = 64;
constexpr int kTileY = 32;
constexpr int kTileThreads = 256;
constexpr int kRowsPerWave = kTileY / (kTileThreads / 64);

// Kinds of focal_tile work (host picks one per op).
enum FocalKind {
    FK_CORRELATE = 0,
    FK_SEPARABLE = 1,
    FK_MEAN = 2,
    FK_MIN = 3,
    FK_MAX = 4,
    FK_FIT = 5,
};

template <int KIND, int R>
__global__ __launch_bounds__(kTileThreads) void focal_tile(const Params P, Region rg, SrcView s, DstView d0,
                                                           DstView d1)
{
    constexpr int N = 2 * R + 1;
    constexpr int LW = kTileX + 2 * R;  // staged columns
    constexpr int LH = kTileY + 2 * R;  // staged rows
    constexpr int NCOL = (KIND == FK_CORRELATE) ? 1 : (KIND == FK_FIT ? 3 : 1);
    __shared__ float tile[LH][LW];
    __shared__ float colv[NCOL][(KIND == FK_CORRELATE) ? 1 : kTileY][LW];

    const int tid = threadIdx.x;
    const int64_t tx0 = rg.ox + (int64_t)blockIdx.x * kTileX;
    const int64_t ty0 = rg.oy + (int64_t)blockIdx.y * kTileY;

    for (int i = tid; i < LH * LW; i += kTileThreads) {
        int row = i / LW, col = i - row * LW;
        tile[row][col] = load_cell(s, tx0 - R + col, ty0 - R + row);
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int64_t x = tx0 + lane;
    const int rbase = wave * kRowsPerWave;
    float out0[kRowsPerWave], out1[kRowsPerWave];

    if constexpr (KIND == FK_CORRELATE) {
        // Per output: w[0]*v(0,0), then acc + w[t]*v(t) over the weights in
        // row-major order (internal/focalrow/focalrow.go:217-230). Each LDS
        // value is read once and applied to every output row that uses it;
        // for each output the terms still arrive in row-major order.
        float acc[kRowsPerWave];
#pragma unroll
        for (int o = 0; o < kRowsPerWave; ++o) acc[o] = 0.0f;
        for (int row = 0; row < kRowsPerWave + 2 * R; ++row) {
#pragma unroll
            for (int c = 0; c < N; ++c) {
                const float v = tile[rbase + row][lane + c];
#pragma unroll
                for (int o = 0; o < kRowsPerWave; ++o) {
                    const int j = row - o;
                    if (j >= 0 && j < N) {
                        const float term = P.taps[j * N + c] * v;
                        acc[o] = (j == 0 && c == 0) ? term : acc[o] + term;
                    }
                }
            }
        }
#pragma unroll
        for (int o = 0; o < kRowsPerWave; ++o) out0[o] = acc[o];
    } else {
        // Column pass for the tile's rows over the staged columns.
        for (int i = tid; i < kTileY * LW; i += kTileThreads) {
            const int rr = i / LW, cc = i - rr * LW;
            if constexpr (KIND == FK_SEPARABLE) {
                // ColumnCorrelate with the col taps (focalrow.go:232-242).
                const float *col = P.taps + N;
                float a = col[0] * tile[rr][cc];
#pragma unroll
                for (int j = 1; j < N; ++j) a = a + col[j] * tile[rr + j][cc];
                colv[0][rr][cc] = a;
            } else if constexpr (KIND == FK_MEAN) {
                float a = tile[rr][cc];  // ColumnSum (focalrow.go:244-250)
#pragma unroll
                for (int j = 1; j < N; ++j) a += tile[rr + j][cc];
                colv[0][rr][cc] = a;
            } else if constexpr (KIND == FK_MIN) {
                float a = tile[rr][cc];
#pragma unroll
                for (int j = 1; j < N; ++j) a = go_min(a, tile[rr + j][cc]);
                colv[0][rr][cc] = a;
            } else if constexpr (KIND == FK_MAX) {
                float a = tile[rr][cc];
#pragma unroll
                for (int j = 1; j < N; ++j) a = go_max(a, tile[rr + j][cc]);
                colv[0][rr][cc] = a;
            } else {
                // Wood fit (terrain/fit.go:137-145): plain, i-weighted and
                // t2-weighted column sums; ti = j - R, t2 = 3(j-R)^2 - R(R+1).
                float cs = tile[rr][cc];
#pragma unroll
                for (int j = 1; j < N; ++j) cs += tile[rr + j][cc];
                float cj = (float)(0 - R) * tile[rr][cc];
#pragma unroll
                for (int j = 1; j < N; ++j) cj = cj + (float)(j - R) * tile[rr + j][cc];
                colv[0][rr][cc] = cs;
                colv[1][rr][cc] = cj;
                if (P.op == STRATA_OP_CURVATURE) {
                    float ct = (float)(3 * R * R - R * (R + 1)) * tile[rr][cc];
#pragma unroll
                    for (int j = 1; j < N; ++j)
                        ct = ct + (float)(3 * (j - R) * (j - R) - R * (R + 1)) * tile[rr + j][cc];
                    colv[2][rr][cc] = ct;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int o = 0; o < kRowsPerWave; ++o) {
            const int rr = rbase + o;
            out1[o] = 0.0f;
            if constexpr (KIND == FK_SEPARABLE) {
                const float *row = P.taps;  // RowCorrelate (focalrow.go:275-284)
                float a = row[0] * colv[0][rr][lane];
#pragma unroll
                for (int c = 1; c < N; ++c) a = a + row[c] * colv[0][rr][lane + c];
                out0[o] = a;
            } else if constexpr (KIND == FK_MEAN) {
                float a = colv[0][rr][lane];  // RowMean (focalrow.go:286-293)
#pragma unroll
                for (int c = 1; c < N; ++c) a += colv[0][rr][lane + c];
                out0[o] = a / (float)(N * N);
            } else if constexpr (KIND == FK_MIN) {
                float a = colv[0][rr][lane];
#pragma unroll
                for (int c = 1; c < N; ++c) a = go_min(a, colv[0][rr][lane + c]);
                out0[o] = a;
            } else if constexpr (KIND == FK_MAX) {
                float a = colv[0][rr][lane];
#pragma unroll
                for (int c = 1; c < N; ++c) a = go_max(a, colv[0][rr][lane + c]);
                out0[o] = a;
            } else {
                // Row passes and factors (terrain/fit.go:140-182).
                float p = (float)(0 - R) * colv[0][rr][lane];
#pragma unroll
                for (int c = 1; c < N; ++c) p = p + (float)(c - R) * colv[0][rr][lane + c];
                float q = 1.0f * colv[1][rr][lane];
#pragma unroll
                for (int c = 1; c < N; ++c) q = q + 1.0f * colv[1][rr][lane + c];
                p = p * P.k[0];
                q = q * P.k[1];
                switch (P.op) {
                case STRATA_OP_GRADIENT:
                    out0[o] = p;
                    out1[o] = q;
                    break;
                case STRATA_OP_SLOPE:
                    out0[o] = slope_of(p, q, P.k[5], (P.flags & 1) != 0);
                    break;
                case STRATA_OP_ASPECT:
                    out0[o] = aspect_of(p, q, P.k[5], (P.flags & 2) != 0);
                    break;
                case STRATA_OP_HILLSHADE:
                    out0[o] = shade_of(p, q, P.k[5], P.k[6], P.k[7]);
                    break;
                default: {
                    float r2 = (float)(3 * R * R - R * (R + 1)) * colv[0][rr][lane];
#pragma unroll
                    for (int c = 1; c < N; ++c)
                        r2 = r2 + (float)(3 * (c - R) * (c - R) - R * (R + 1)) * colv[0][rr][lane + c];
                    float t2 = 1.0f * colv[2][rr][lane];
#pragma unroll
                    for (int c = 1; c < N; ++c) t2 = t2 + 1.0f * colv[2][rr][lane + c];
                    float s2 = (float)(0 - R) * colv[1][rr][lane];
#pragma unroll
                    for (int c = 1; c < N; ++c) s2 = s2 + (float)(c - R) * colv[1][rr][lane + c];
                    r2 = r2 * P.k[2];
                    t2 = t2 * P.k[3];
                    s2 = s2 * P.k[4];
                    out0[o] = curvature_of(p, q, r2, s2, t2, (P.flags >> 4) & 3);
                }
                }
            }
        }
    }

    // Store: kernel values over interior cells, NaN over the edge ring.
    const bool col_in = x >= rg.ox && x < rg.ox + rg.w;
    const bool col_interior = x >= R && x < rg.W - R;
#pragma unroll
    for (int o = 0; o < kRowsPerWave; ++o) {
        const int64_t y = ty0 + rbase + o;
        if (!col_in || y >= rg.oy + rg.h) continue;
        const bool interior = col_interior && y >= R && y < rg.H - R;
        d0.p[(y - rg.oy) * d0.stride + (x - rg.ox)] = interior ? out0[o] : edge_nan();
        if constexpr (KIND == FK_FIT) {
            if (P.op == STRATA_OP_GRADIENT)
                d1.p[(y - rg.oy) * d1.stride + (x - rg.ox)] = interior ? out1[o] : edge_nan();
        }
    }
}

// ---------------------------------------------------------------------------
// point_rows: radius-0 transfer kernels
// ---------------------------------------------------------------------------

constexpr int kPointRows = 16;

template <int OP>
__global__ __launch_bounds__(256) void point_rows(const Params P, Region rg, SrcView s, DstView d0)
{
    const int64_t x = rg.ox + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (x >= rg.ox + rg.w) return;
    const int64_t y0 = rg.oy + (int64_t)blockIdx.y * kPointRows;
    const int64_t y1 = min(y0 + (int64_t)kPointRows, rg.oy + rg.h);
    for (int64_t y = y0; y < y1; ++y) {
        const float v = s.p[(y - s.sy) * s.stride + (x - s.sx)];
        float o;
        if constexpr (OP == STRATA_OP_LOOKUP) {
            const int n = P.ntaps / 2;
            o = lookup_of(v, P.taps, P.taps + n, n);
        } else {
            const int nb = (P.ntaps - 1) / 2;
            o = reclass_of(v, P.taps, nb, P.taps + nb);
        }
        d0.p[(y - rg.oy) * d0.stride + (x - rg.ox)] = o;
    }
}

// ---------------------------------------------------------------------------
// validity: output bits, one 64-bit word per thread
// ---------------------------------------------------------------------------

// Bits [pos, pos+64) of m, reading only words up to index `last`.
__device__ __forceinline__ uint64_t bits_at(const uint64_t *m, int64_t pos, int64_t last)
{
    const int64_t w = pos >> 6;
    const int sh = (int)(pos & 63);
    uint64_t v = m[w] >> sh;
    if (sh != 0 && w + 1 <= last) v |= m[w + 1] << (64 - sh);
    return v;
}

__global__ __launch_bounds__(256) void validity_kernel(int r, Region rg, MaskView src, int has_src,
                                                       MaskView dst, int64_t row0)
{
    const int64_t y = row0 + blockIdx.y;          // raster row
    const int64_t ob = dst.off + (y - dst.y0) * dst.stride + (rg.ox - dst.x0);  // bit of cell (ox, y)
    const int64_t wfirst = ob >> 6, wlast = (ob + rg.w - 1) >> 6;
    const int64_t wi = wfirst + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (wi > wlast) return;

    // Bits of this word that belong to the region's row.
    const int64_t wb = wi << 6;
    const int64_t own_lo = max(wb, ob), own_hi = min(wb + 64, ob + rg.w);
    uint64_t own = (own_hi - own_lo == 64) ? ~0ull : (((1ull << (own_hi - own_lo)) - 1) << (own_lo - wb));

    // Interior cells of the row in this word.
    uint64_t val = 0;
    if (y >= r && y < rg.H - r) {
        const int64_t lo = max(own_lo, ob + (r - rg.ox));
        const int64_t hi = min(own_hi, ob + (rg.W - r - rg.ox));
        if (lo < hi) {
            const int n = (int)(hi - lo);
            uint64_t acc = (n == 64) ? ~0ull : ((1ull << n) - 1);
            if (has_src) {
                const int64_t xlo = rg.ox + (lo - ob);  // first interior cell
                for (int dy = -r; dy <= r; ++dy) {
                    const int64_t a = src.off + (y + dy - src.y0) * src.stride + (xlo - r - src.x0);
                    const int64_t last = (a + n + 2 * r - 1) >> 6;
                    uint64_t rowv = ~0ull;
                    for (int c = 0; c <= 2 * r; ++c) rowv &= bits_at(src.bits, a + c, last);
                    acc &= rowv;
                }
            }
            val = acc << (lo - wb);
        }
    }
    uint64_t *word = dst.bits + wi;
    if (own == ~0ull) {
        *word = val;
    } else {
        atomicAnd((unsigned long long *)word, (unsigned long long)(~own | val));
        atomicOr((unsigned long long *)word, (unsigned long long)val);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
constexpr int64_t kMaxGridY = 65535;

template <int OP>
static hipError_t run_horn(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                           hipStream_t st)
{
    HornArgs a;
    for (int i = 0; i < 8; ++i) a.k[i] = P.k[i];
    a.flags = P.flags;
    const DstView d1 = nout > 1 ? d[1] : d[0];
    const int64_t gx = cdiv(rg.w, kSweepThreads);
    for (int64_t yb = 0; yb < cdiv(rg.h, kSweepRows); yb += kMaxGridY) {
        Region sub = rg;
        sub.oy = rg.oy + yb * kSweepRows;
        sub.h = min(rg.h - yb * kSweepRows, kMaxGridY * kSweepRows);
        DstView e0 = d[0], e1 = d1;
        e0.p += yb * kSweepRows * e0.stride;
        e1.p += yb * kSweepRows * e1.stride;
        dim3 grid((unsigned)gx, (unsigned)cdiv(sub.h, kSweepRows));
        hipLaunchKernelGGL(horn_sweep<OP>, grid, dim3(kSweepThreads), 0, st, a, sub, s, e0, e1);
    }
    return hipGetLastError();
}

template <int KIND, int R>
static hipError_t run_focal_r(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                              hipStream_t st)
{
    const DstView d1 = nout > 1 ? d[1] : d[0];
    const int64_t gx = cdiv(rg.w, kTileX);
    for (int64_t yb = 0; yb < cdiv(rg.h, kTileY); yb += kMaxGridY) {
        Region sub = rg;
        sub.oy = rg.oy + yb * kTileY;
        sub.h = min(rg.h - yb * kTileY, kMaxGridY * kTileY);
        DstView e0 = d[0], e1 = d1;
        e0.p += yb * kTileY * e0.stride;
        e1.p += yb * kTileY * e1.stride;
        dim3 grid((unsigned)gx, (unsigned)cdiv(sub.h, kTileY));
        hipLaunchKernelGGL((focal_tile<KIND, R>), grid, dim3(kTileThreads), 0, st, P, sub, s, e0, e1);
    }
    return hipGetLastError();
}

template <int KIND>
static hipError_t run_focal(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                            hipStream_t st)
{
    switch (P.radius) {
    case 1: return run_focal_r<KIND, 1>(P, rg, s, d, nout, st);
    case 2: return run_focal_r<KIND, 2>(P, rg, s, d, nout, st);
    case 3: return run_focal_r<KIND, 3>(P, rg, s, d, nout, st);
    case 4: return run_focal_r<KIND, 4>(P, rg, s, d, nout, st);
    case 5: return run_focal_r<KIND, 5>(P, rg, s, d, nout, st);
    case 6: return run_focal_r<KIND, 6>(P, rg, s, d, nout, st);
    case 7: return run_focal_r<KIND, 7>(P, rg, s, d, nout, st);
    case 8: return run_focal_r<KIND, 8>(P, rg, s, d, nout, st);
    }
    return hipErrorInvalidValue;
}

template <int OP>
static hipError_t run_point(const Params &P, const Region &rg, const SrcView &s, const DstView *d, hipStream_t st)
{
    const int64_t gx = cdiv(rg.w, 256);
    for (int64_t yb = 0; yb < cdiv(rg.h, kPointRows); yb += kMaxGridY) {
        Region sub = rg;
        sub.oy = rg.oy + yb * kPointRows;
        sub.h = min(rg.h - yb * kPointRows, kMaxGridY * kPointRows);
        DstView e0 = d[0];
        e0.p += yb * kPointRows * e0.stride;
        dim3 grid((unsigned)gx, (unsigned)cdiv(sub.h, kPointRows));
        hipLaunchKernelGGL(point_rows<OP>, grid, dim3(256), 0, st, P, sub, s, e0);
    }
    return hipGetLastError();
}

hipError_t launch_data(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                       hipStream_t st)
{
    if (rg.w <= 0 || rg.h <= 0) return hipSuccess;
    const bool fit = (P.flags & 0x100) != 0;
    if (fit && P.op >= STRATA_OP_GRADIENT && P.op <= STRATA_OP_CURVATURE)
        return run_focal<FK_FIT>(P, rg, s, d, nout, st);
    switch (P.op) {
    case STRATA_OP_GRADIENT: return run_horn<STRATA_OP_GRADIENT>(P, rg, s, d, nout, st);
    case STRATA_OP_SLOPE: return run_horn<STRATA_OP_SLOPE>(P, rg, s, d, nout, st);
    case STRATA_OP_ASPECT: return run_horn<STRATA_OP_ASPECT>(P, rg, s, d, nout, st);
    case STRATA_OP_HILLSHADE: return run_horn<STRATA_OP_HILLSHADE>(P, rg, s, d, nout, st);
    case STRATA_OP_CURVATURE: return run_horn<STRATA_OP_CURVATURE>(P, rg, s, d, nout, st);
    case STRATA_OP_CORRELATE: return run_focal<FK_CORRELATE>(P, rg, s, d, nout, st);
    case STRATA_OP_SEPARABLE: return run_focal<FK_SEPARABLE>(P, rg, s, d, nout, st);
    case STRATA_OP_MEAN: return run_focal<FK_MEAN>(P, rg, s, d, nout, st);
    case STRATA_OP_MIN: return run_focal<FK_MIN>(P, rg, s, d, nout, st);
    case STRATA_OP_MAX: return run_focal<FK_MAX>(P, rg, s, d, nout, st);
    case STRATA_OP_LOOKUP: return run_point<STRATA_OP_LOOKUP>(P, rg, s, d, st);
    case STRATA_OP_RECLASS: return run_point<STRATA_OP_RECLASS>(P, rg, s, d, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_validity(int r, const Region &rg, const MaskView &src, int has_src, const MaskView &dst,
                           hipStream_t st)
{
    if (rg.w <= 0 || rg.h <= 0) return hipSuccess;
    // Words a row of the region can touch: at most w/64 + 2.
    const int64_t words = rg.w / 64 + 2;
    const int64_t gx = cdiv(words, 256);
    for (int64_t yb = 0; yb < rg.h; yb += kMaxGridY) {
        const int64_t rows = min(rg.h - yb, kMaxGridY);
        dim3 grid((unsigned)gx, (unsigned)rows);
        hipLaunchKernelGGL(validity_kernel, grid, dim3(256), 0, st, r, rg, src, has_src, dst, rg.oy + yb);
    }
    return hipGetLastError();
}

} // namespace strata
