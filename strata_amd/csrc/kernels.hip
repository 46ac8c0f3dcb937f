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
//
// Layout: a block of 256 lanes covers 512 columns, two adjacent columns per
// lane, and walks down a strip of kSweepRows rows. Each input row is loaded
// once per strip (a dwordx2 per lane when rows are 8-byte aligned); the cell
// left of a lane's pair comes from lane-1 and the cell right of it from
// lane+1 through DPP wave shifts, and each wave's two outer columns from one
// extra load by lanes 0 and 63. Rows live in a ring of four register sets,
// unrolled so no row ever moves between registers, with the next row's
// loads in flight while the current one is computed.
//
// Blocks whose columns and rows (with the halo) lie inside the buffer take
// the unchecked path: uniform row addresses plus 32-bit lane offsets.

constexpr int kSweepThreads = 256;
constexpr int kSweepCols = kSweepThreads * 2;
constexpr int kSweepRows = 64;

struct HornArgs {
    float k[8];
    int32_t flags;
    int32_t aligned;  // source and outputs allow 8-byte pair accesses at even offsets
};

// An input row as one lane holds it: its pair of cells, the cells left and
// right of the pair, and the pair's horizontal differences (z3 - z1 for
// each of the two columns), which three output rows share.
struct Row {
    float c0, c1, l, r, hd0, hd1;
};

// Correctly rounded sqrtf for x = +0 or x in [2^-96, FLT_MAX]: the hardware
// root and the one-ulp correction the compiler's full sequence applies,
// without its denormal scaling and special-value fix-up (not needed there).
__device__ __forceinline__ bool sqrt_fast_ok(float x) { return x == 0.0f || (x >= 0x1p-96f && x <= 3.40282347e38f); }
__device__ __forceinline__ float sqrt_fast(float x)
{
    const float y = __builtin_amdgcn_sqrtf(x);
    const float ydn = u2f(f2u(y) - 1u), yup = u2f(f2u(y) + 1u);
    const float rdn = __builtin_fmaf(-ydn, y, x), rup = __builtin_fmaf(-yup, y, x);
    float r = rdn <= 0.0f ? ydn : y;
    return rup > 0.0f ? yup : r;
}

// Correctly rounded n / d for |n| in [2^-30, 2^30] and d in [1, 2^30]: the
// compiler's refinement sequence (reciprocal refined by one Newton step, two
// residual corrections) without the range scaling and fix-up, which never
// apply there.
__device__ __forceinline__ bool div_fast_ok(float n, float d)
{
    const float an = u2f(f2u(n) & 0x7fffffffu);
    return an >= 0x1p-30f && an <= 0x1p30f && d >= 1.0f && d <= 0x1p30f;
}
__device__ __forceinline__ float div_fast(float n, float d)
{
    float y = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = n * y;
    float r = __builtin_fmaf(-d, q, n);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-d, q, n);
    return __builtin_fmaf(r, y, q);
}

// Atan32 for x in [0, tan(pi/8)]: the reduction picks num = x, den = 1 and
// y0 = 0 there, so t = x exactly, and y0 + v == v because v is never -0.
__device__ __forceinline__ float atan32_small(float t)
{
    const float z = t * t;
    float p = kAtanC4 * z - kAtanC3;
    p = p * z + kAtanC2;
    p = p * z - kAtanC1;
    return (p * z) * t + t;
}

// Slope of a pair of cells; the fast paths are taken by whole waves.
__device__ __forceinline__ void slope_pair(float gx0, float gy0, float gx1, float gy1, float scale, bool atan_,
                                           float &o0, float &o1)
{
    const float s0 = gx0 * gx0 + gy0 * gy0, s1 = gx1 * gx1 + gy1 * gy1;
    float m0, m1;
    if (__all(sqrt_fast_ok(s0) && sqrt_fast_ok(s1))) {
        m0 = sqrt_fast(s0);
        m1 = sqrt_fast(s1);
    } else {
        m0 = __builtin_sqrtf(s0);
        m1 = __builtin_sqrtf(s1);
    }
    if (!atan_) {
        o0 = m0 * scale;
        o1 = m1 * scale;
    } else if (__all(m0 <= kTanPi8 && m1 <= kTanPi8)) {
        o0 = atan32_small(m0) * scale;
        o1 = atan32_small(m1) * scale;
    } else {
        o0 = atan32(m0) * scale;
        o1 = atan32(m1) * scale;
    }
}

// Hillshade of a pair (shade_of with the fast root and quotient when the
// whole wave is in their ranges).
__device__ __forceinline__ void shade_pair(float gx0, float gy0, float gx1, float gy1, float c, float bx, float by,
                                           float &o0, float &o1)
{
    const float n0 = c + (bx * gx0 + by * gy0), n1 = c + (bx * gx1 + by * gy1);
    const float w0 = 1.0f + (gx0 * gx0 + gy0 * gy0), w1 = 1.0f + (gx1 * gx1 + gy1 * gy1);
    float v0, v1;
    if (__all(sqrt_fast_ok(w0) && sqrt_fast_ok(w1))) {
        const float d0 = sqrt_fast(w0), d1 = sqrt_fast(w1);
        if (__all(div_fast_ok(n0, d0) && div_fast_ok(n1, d1))) {
            v0 = div_fast(n0, d0);
            v1 = div_fast(n1, d1);
        } else {
            v0 = n0 / d0;
            v1 = n1 / d1;
        }
    } else {
        v0 = n0 / __builtin_sqrtf(w0);
        v1 = n1 / __builtin_sqrtf(w1);
    }
    if (v0 < 0.0f) v0 = 0.0f;
    if (v0 > 255.0f) v0 = 255.0f;
    if (v1 < 0.0f) v1 = 0.0f;
    if (v1 > 255.0f) v1 = 255.0f;
    o0 = v0;
    o1 = v1;
}

// The op over the two cells of a lane from rows A (y-1), B (y), C (y+1):
// outputs (o00, o01) for the first output and (o10, o11) for the second.
template <int OP>
__device__ __forceinline__ void horn_pair(const HornArgs &a, const Row &A, const Row &B, const Row &C, float &o00,
                                          float &o01, float &o10, float &o11)
{
    if constexpr (OP == STRATA_OP_CURVATURE) {
        const int kind = (a.flags >> 4) & 3;
        o00 = zt_curvature(A.l, A.c0, A.c1, B.l, B.c0, B.c1, C.l, C.c0, C.c1, a.k, kind);
        o01 = zt_curvature(A.c0, A.c1, A.r, B.c0, B.c1, B.r, C.c0, C.c1, C.r, a.k, kind);
    } else {
        // Horn (internal/stencil/horn.go:172-180), grouped as the reference
        // groups it; the differences are shared between rows and columns.
        const float dx0 = (A.hd0 + C.hd0) + (B.hd0 + B.hd0);
        const float dx1 = (A.hd1 + C.hd1) + (B.hd1 + B.hd1);
        const float gl = C.l - A.l, g0 = C.c0 - A.c0, g1 = C.c1 - A.c1, gr = C.r - A.r;
        const float dy0 = (gl + g1) + (g0 + g0);
        const float dy1 = (g0 + gr) + (g1 + g1);
        const float gx0 = dx0 * a.k[0], gy0 = dy0 * a.k[1];
        const float gx1 = dx1 * a.k[0], gy1 = dy1 * a.k[1];
        if constexpr (OP == STRATA_OP_GRADIENT) {
            o00 = gx0;
            o01 = gx1;
            o10 = gy0;
            o11 = gy1;
        } else if constexpr (OP == STRATA_OP_SLOPE) {
            slope_pair(gx0, gy0, gx1, gy1, a.k[2], (a.flags & 1) != 0, o00, o01);
        } else if constexpr (OP == STRATA_OP_ASPECT) {
            o00 = aspect_of(gx0, gy0, a.k[2], (a.flags & 2) != 0);
            o01 = aspect_of(gx1, gy1, a.k[2], (a.flags & 2) != 0);
        } else {
            shade_pair(gx0, gy0, gx1, gy1, a.k[2], a.k[3], a.k[4], o00, o01);
        }
    }
}

template <int OP>
__device__ __forceinline__ void sweep(const HornArgs &a, const Region &rg, const SrcView &s, const DstView &d0,
                                      const DstView &d1, int64_t xb, int64_t y0, int64_t y1)
{
    constexpr bool kTwo = OP == STRATA_OP_GRADIENT;
    const int lane = threadIdx.x & 63;
    const int32_t off = 2 * (int32_t)threadIdx.x;  // the pair's first column, relative to xb
    const int64_t x = xb + off;
    const bool halo_lane = lane == 0 || lane == 63;
    // Column offsets of the cells this lane loads, clamped into the buffer:
    // a clamped cell is read only for edge outputs, which become NaN.
    auto col = [&](int64_t c) { return (int32_t)(min(max(c, s.lx0), s.lx1 - 1) - s.sx); };
    const int32_t o0 = col(x), o1 = col(x + 1), oh = col(lane == 0 ? x - 1 : x + 2);
    const bool ci0 = x >= 1 && x < rg.W - 1;
    const bool ci1 = x + 1 >= 1 && x + 1 < rg.W - 1;
    const bool in0 = x >= rg.ox && x < rg.ox + rg.w;
    const bool in1 = x + 1 >= rg.ox && x + 1 < rg.ox + rg.w;
    // Block-uniform: every pair is inside the buffer (loads) and the region
    // (stores), at 8-byte-aligned addresses.
    const bool pair_in = a.aligned && xb >= s.lx0 && xb + kSweepCols <= s.lx1;
    const bool pair_out = a.aligned && xb >= rg.ox && xb + kSweepCols <= rg.ox + rg.w;
    const int32_t poff = (int32_t)(xb - s.sx) + off;

    auto load = [&](Row &row, int64_t y) {
        const float *rp = s.p + (min(max(y, s.ly0), s.ly1 - 1) - s.sy) * s.stride;
        float c0, c1, h = 0.0f;
        if (pair_in) {
            const float2 v = *reinterpret_cast<const float2 *>(rp + poff);
            c0 = v.x;
            c1 = v.y;
        } else {
            c0 = rp[o0];
            c1 = rp[o1];
        }
        if (halo_lane) h = rp[oh];
        row.c0 = c0;
        row.c1 = c1;
        row.l = h;  // the halo value until finish() forms l and r
    };
    auto finish = [&](Row &row) {
        const float h = row.l;
        row.l = from_left(row.c1, h);
        row.r = from_right(row.c0, h);
        row.hd0 = row.c1 - row.l;
        row.hd1 = row.r - row.c0;
    };
    auto store = [&](const DstView &d, int64_t y, float v0, float v1) {
        float *rp = d.p + (y - rg.oy) * d.stride + (xb - rg.ox);
        if (pair_out) {
            *reinterpret_cast<float2 *>(rp + off) = make_float2(v0, v1);
        } else {
            if (in0) rp[off] = v0;
            if (in1) rp[off + 1] = v1;
        }
    };

    Row R[4];
    load(R[0], y0 - 1);
    load(R[1], y0);
    load(R[2], y0 + 1);
    finish(R[0]);
    finish(R[1]);
    for (int64_t y = y0; y < y1; y += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t yy = y + u;
            if (yy < y1) {
                Row &A = R[u & 3];
                Row &B = R[(u + 1) & 3];
                Row &C = R[(u + 2) & 3];
                Row &D = R[(u + 3) & 3];
                if (yy + 2 <= y1) load(D, yy + 2);
                finish(C);
                float o00 = 0.0f, o01 = 0.0f, o10 = 0.0f, o11 = 0.0f;
                horn_pair<OP>(a, A, B, C, o00, o01, o10, o11);
                const bool ri = yy >= 1 && yy < rg.H - 1;
                const float nan = edge_nan();
                store(d0, yy, (ri && ci0) ? o00 : nan, (ri && ci1) ? o01 : nan);
                if constexpr (kTwo) store(d1, yy, (ri && ci0) ? o10 : nan, (ri && ci1) ? o11 : nan);
            }
        }
    }
}

template <int OP>
__global__ __launch_bounds__(kSweepThreads) void horn_sweep(HornArgs a, Region rg, SrcView s, DstView d0,
                                                            DstView d1)
{
    const int64_t xb = rg.ox + (int64_t)blockIdx.x * kSweepCols;
    const int64_t y0 = rg.oy + (int64_t)blockIdx.y * kSweepRows;
    const int64_t y1 = min(y0 + (int64_t)kSweepRows, rg.oy + rg.h);
    sweep<OP>(a, rg, s, d0, d1, xb, y0, y1);
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

// The op's constants; taps come through an accessor (kernel arguments for a
// standalone launch, a device buffer for a batch).
struct FocalArgs {
    int32_t op, flags;
    float k[8];
};
struct ArgTaps {
    const Params &P;
    __device__ float operator[](int i) const { return P.taps[i]; }
};
struct PtrTaps {
    const float *p;
    __device__ float operator[](int i) const { return p[i]; }
};

// LDS floats one focal_block<KIND, R> uses.
template <int KIND, int R>
constexpr int focal_lds_floats()
{
    return (kTileY + 2 * R) * (kTileX + 2 * R) +
           ((KIND == FK_CORRELATE) ? 0 : (KIND == FK_FIT ? 3 : 1) * kTileY * (kTileX + 2 * R));
}

// One kTileX x kTileY output tile at (tx0, ty0) of the region.
template <int KIND, int R, class TAPS>
__device__ __forceinline__ void focal_block(const FocalArgs &P, const TAPS &taps, const Region &rg,
                                            const SrcView &s, const DstView &d0, const DstView &d1, int64_t tx0,
                                            int64_t ty0, float *lds)
{
    constexpr int N = 2 * R + 1;
    constexpr int LW = kTileX + 2 * R;  // staged columns
    constexpr int LH = kTileY + 2 * R;  // staged rows
    constexpr int CY = (KIND == FK_CORRELATE) ? 1 : kTileY;
    float(*tile)[LW] = reinterpret_cast<float(*)[LW]>(lds);
    float(*colv)[CY][LW] = reinterpret_cast<float(*)[CY][LW]>(lds + LH * LW);

    const int tid = threadIdx.x;
    __syncthreads();  // the LDS may still be read by the previous tile of a batch
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
                        const float term = taps[j * N + c] * v;
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
                float a = taps[N] * tile[rr][cc];
#pragma unroll
                for (int j = 1; j < N; ++j) a = a + taps[N + j] * tile[rr + j][cc];
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
                // RowCorrelate with the row taps (focalrow.go:275-284).
                float a = taps[0] * colv[0][rr][lane];
#pragma unroll
                for (int c = 1; c < N; ++c) a = a + taps[c] * colv[0][rr][lane + c];
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

template <int KIND, int R>
__global__ __launch_bounds__(kTileThreads) void focal_tile(const Params P, Region rg, SrcView s, DstView d0,
                                                           DstView d1)
{
    __shared__ float lds[focal_lds_floats<KIND, R>()];
    FocalArgs a;
    a.op = P.op;
    a.flags = P.flags;
#pragma unroll
    for (int i = 0; i < 8; ++i) a.k[i] = P.k[i];
    const int64_t tx0 = rg.ox + (int64_t)blockIdx.x * kTileX;
    const int64_t ty0 = rg.oy + (int64_t)blockIdx.y * kTileY;
    focal_block<KIND, R>(a, ArgTaps{P}, rg, s, d0, d1, tx0, ty0, lds);
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

// Bits [o, o+64) of the 192-bit little-endian value (w0, w1, w2), o < 128.
__device__ __forceinline__ uint64_t bits3(uint64_t w0, uint64_t w1, uint64_t w2, int o)
{
    if (o >= 64) {
        o -= 64;
        w0 = w1;
        w1 = w2;
    }
    return o == 0 ? w0 : (w0 >> o) | (w1 << (64 - o));
}

// One thread per output mask word of one row of the region. For the
// word's interior cells [lo, hi): each of the 2R+1 source rows is read
// as the bits of cells [lo-R, hi+R) (at most three words, all loaded up
// front), the rows are ANDed, and the result is eroded horizontally by
// ANDing its 2R+1 shifts. Words only partly owned by the region's row
// are updated with atomics on the row's own bits.
template <int R>
__global__ __launch_bounds__(256) void validity_kernel(Region rg, MaskView src, int has_src, MaskView dst,
                                                       int64_t row0)
{
    const int64_t y = row0 + blockIdx.y;
    const int64_t ob = dst.off + (y - dst.y0) * dst.stride + (rg.ox - dst.x0);  // bit of cell (ox, y)
    const int64_t wfirst = ob >> 6, wlast = (ob + rg.w - 1) >> 6;
    const int64_t wi = wfirst + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (wi > wlast) return;

    const int64_t wb = wi << 6;
    const int64_t own_lo = max(wb, ob), own_hi = min(wb + 64, ob + rg.w);
    const uint64_t own = (own_hi - own_lo == 64) ? ~0ull : (((1ull << (own_hi - own_lo)) - 1) << (own_lo - wb));

    uint64_t val = 0;
    if (y >= R && y < rg.H - R) {
        const int64_t lo = max(own_lo, ob + (R - rg.ox));
        const int64_t hi = min(own_hi, ob + (rg.W - R - rg.ox));
        if (lo < hi) {
            const int n = (int)(hi - lo);
            uint64_t acc = (n == 64) ? ~0ull : ((1ull << n) - 1);
            if (has_src) {
                const int64_t xlo = rg.ox + (lo - ob);  // first interior cell of the word
                uint64_t w[2 * R + 1][3];
                int sh[2 * R + 1];
#pragma unroll
                for (int j = 0; j <= 2 * R; ++j) {
                    const int64_t a = src.off + (y - R + j - src.y0) * src.stride + (xlo - R - src.x0);
                    const int64_t base = a >> 6, last = (a + n + 2 * R - 1) >> 6;
                    sh[j] = (int)(a & 63);
#pragma unroll
                    for (int k = 0; k < 3; ++k) w[j][k] = (base + k <= last) ? src.bits[base + k] : 0ull;
                }
                // Rows aligned to cell xlo-R: bits [0, 64) and [64, 64+2R).
                uint64_t v0 = ~0ull, v1 = ~0ull;
#pragma unroll
                for (int j = 0; j <= 2 * R; ++j) {
                    v0 &= bits3(w[j][0], w[j][1], w[j][2], sh[j]);
                    if (R > 0) v1 &= bits3(w[j][0], w[j][1], w[j][2], sh[j] + 64);
                }
                uint64_t e = v0;
#pragma unroll
                for (int c = 1; c <= 2 * R; ++c) e &= (v0 >> c) | (v1 << (64 - c));
                acc &= e;
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
    // Pair (8-byte) accesses need every pair's first column at an even
    // offset from 8-byte-aligned row starts.
    auto al8 = [](const void *p) { return ((uintptr_t)p & 7) == 0; };
    bool aligned = al8(s.p) && (s.stride & 1) == 0 && ((rg.ox - s.sx) & 1) == 0;
    for (int o = 0; o < nout; ++o) aligned = aligned && al8(d[o].p) && (d[o].stride & 1) == 0;
    a.aligned = aligned ? 1 : 0;
    const int64_t gx = cdiv(rg.w, kSweepCols);
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
        switch (r) {
#define STRATA_VALIDITY_CASE(RR)                                                                            \
    case RR:                                                                                                \
        hipLaunchKernelGGL(validity_kernel<RR>, grid, dim3(256), 0, st, rg, src, has_src, dst, rg.oy + yb); \
        break;
            STRATA_VALIDITY_CASE(0)
            STRATA_VALIDITY_CASE(1)
            STRATA_VALIDITY_CASE(2)
            STRATA_VALIDITY_CASE(3)
            STRATA_VALIDITY_CASE(4)
            STRATA_VALIDITY_CASE(5)
            STRATA_VALIDITY_CASE(6)
            STRATA_VALIDITY_CASE(7)
            STRATA_VALIDITY_CASE(8)
#undef STRATA_VALIDITY_CASE
        default:
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

} // namespace strata

// ---------------------------------------------------------------------------
// Self-checks of the fast paths against the full sequences (run by tests)
// ---------------------------------------------------------------------------

namespace strata {

__device__ __forceinline__ uint32_t mix32(uint64_t v)
{
    v ^= v >> 33;
    v *= 0xff51afd7ed558ccdull;
    v ^= v >> 33;
    v *= 0xc4ceb9fe1a85ec53ull;
    v ^= v >> 33;
    return (uint32_t)v;
}

// kind 0: sqrt_fast vs sqrtf over bit patterns [lo, lo+n) in its range.
// kind 1: atan32_small vs atan32 over bit patterns [lo, lo+n) up to tan(pi/8).
// kind 2: div_fast vs '/' over n hashed pairs (seed lo) in its range.
__global__ void selfcheck_kernel(int kind, uint64_t lo, uint64_t n, unsigned long long *bad,
                                 unsigned long long *first, unsigned long long *checked)
{
    unsigned long long mine = 0, seen = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        bool differ = false, ran = false;
        if (kind == 0) {
            const float x = u2f((uint32_t)(lo + i));
            if (sqrt_fast_ok(x)) {
                ran = true;
                differ = f2u(sqrt_fast(x)) != f2u(__builtin_sqrtf(x));
            }
        } else if (kind == 1) {
            const float x = u2f((uint32_t)(lo + i));
            if (x >= 0.0f && x <= kTanPi8) {
                ran = true;
                differ = f2u(atan32_small(x)) != f2u(atan32(x));
            }
        } else {
            // |n| in [2^-30, 2^30) with random sign and mantissa, d in [1, 2^30).
            const uint32_t h1 = mix32(lo * 0x9e3779b97f4a7c15ull + 2 * i), h2 = mix32(lo + 2 * i + 1);
            const uint32_t en = 127 - 30 + (h1 >> 26) % 60, ed = 127 + (h2 >> 27) % 30;
            const float num = u2f((h1 & 0x80000000u) | (en << 23) | (h1 & 0x7fffffu));
            const float den = u2f((ed << 23) | (h2 & 0x7fffffu));
            if (div_fast_ok(num, den)) {
                ran = true;
                differ = f2u(div_fast(num, den)) != f2u(num / den);
            }
        }
        seen += ran;
        if (differ) {
            ++mine;
            atomicMin(first, (unsigned long long)i);
        }
    }
    if (mine) atomicAdd(bad, mine);
    atomicAdd(checked, seen);
}

hipError_t launch_selfcheck(int kind, uint64_t lo, uint64_t n, unsigned long long *dev3, hipStream_t st)
{
    hipLaunchKernelGGL(selfcheck_kernel, dim3(8192), dim3(256), 0, st, kind, lo, n, dev3, dev3 + 1, dev3 + 2);
    return hipGetLastError();
}

} // namespace strata
