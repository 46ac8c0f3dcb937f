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
//   validity     output validity, one 32-bit mask word per lane: AND of the
//                (2r+1) source rows, horizontal erosion by funnel shifts
//                with the neighbouring lanes' words, interior mask; whole
//                words are stored, words shared with cells outside the
//                region are updated with atomics on the region's own bits.
//
// The arithmetic is arith.hpp's, in the reference's order.
#include <algorithm>

#include "arith.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace strata {

// ---------------------------------------------------------------------------
// Shared helpers
// ---------------------------------------------------------------------------

// Rasters, masks and work lists live in global memory. Pointers that reach
// a kernel inside a structure (a work list) are generic to the compiler,
// which would then use flat loads and stores (waiting on both the vector
// and the scalar memory counters); accessing through these global-typed
// pointers gives global instructions, as for pointers in kernel arguments.
#define STRATA_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ STRATA_GLOBAL T *gp(T *p)
{
    return (STRATA_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ const STRATA_GLOBAL T *gp(const T *p)
{
    return (const STRATA_GLOBAL T *)p;
}
// Constant memory: device memory that no access of the launch writes. Loads
// through such a pointer are scalar and, unlike global loads, are not
// ordered against the launch's stores, so the compiler keeps what it loaded
// across them instead of loading it again after each one.
#define STRATA_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const STRATA_CONST T *cp(const T *p)
{
    return (const STRATA_CONST T *)p;
}
typedef float pair_t __attribute__((ext_vector_type(2)));

// The tile workgroup b of nb takes. Workgroups are dealt out to the eight
// XCDs in turn (b runs on XCD b % 8), each with its own L2; numbered this
// way, each XCD takes one contiguous run of tiles in row-major order, so
// tiles that share halo lines (a row's neighbours, and the next tile row
// soon after) read them through the same L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb)
{
    const uint32_t xcd = b % kXcds, q = b / kXcds, per = nb / kXcds, extra = nb % kXcds;
    return xcd * per + min(xcd, extra) + q;
}

__device__ __forceinline__ float load_cell(const SrcView &s, int64_t x, int64_t y)
{
    bool in = x >= s.lx0 && x < s.lx1 && y >= s.ly0 && y < s.ly1;
    return in ? gp(s.p)[(y - s.sy) * s.stride + (x - s.sx)] : 0.0f;
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
// lane, and walks down a strip of rows (the tuned strip height). Each input row is loaded
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

struct HornArgs {
    float k[8];
    int32_t flags;
    int32_t aligned;  // source and outputs allow 8-byte pair accesses at even offsets
    int32_t rows;     // rows per strip (one block's height)
    uint32_t blocks_x;  // column blocks per strip
};

// An input row as one lane holds it: its pair of cells, the cells left and
// right of the pair, and the pair's horizontal differences (z3 - z1 for
// each of the two columns), which three output rows share.
struct Row {
    float c0, c1, l, r, hd0, hd1;
    bool odd;  // kSweepAnyParity: the row's cells at x - 1, x were loaded as a pair
};

// Correctly rounded sqrtf for x = +0 or x in [2^-96, FLT_MAX]: the hardware
// root and the one-ulp correction the compiler's full sequence applies,
// without its denormal scaling and special-value fix-up (not needed there).
__device__ __forceinline__ unsigned sqrt_fast_ok(float x)
{
    const uint32_t u = f2u(x);  // +0, or [2^-96, FLT_MAX] (0x0f800000 .. 0x7f7fffff)
    return (u == 0u) | ((u - 0x0f800000u) <= (0x7f7fffffu - 0x0f800000u));
}
__device__ __forceinline__ float sqrt_fast(float x)
{
    const float y = __builtin_amdgcn_sqrtf(x);
    const float ydn = u2f(f2u(y) - 1u), yup = u2f(f2u(y) + 1u);
    const float rdn = __builtin_fmaf(-ydn, y, x), rup = __builtin_fmaf(-yup, y, x);
    float r = rdn <= 0.0f ? ydn : y;
    return rup > 0.0f ? yup : r;
}

// Correctly rounded n / d for |n| and d in [2^-30, 2^30]: the compiler's
// refinement sequence (reciprocal refined by one Newton step, two residual
// corrections) without its range scaling and fix-up, which apply only to
// operands or quotients near overflow, underflow or the denormals (the
// quotient here lies in [2^-60, 2^60]).
__device__ __forceinline__ unsigned div_fast_ok(float n, float d)
{
    const float an = u2f(f2u(n) & 0x7fffffffu);
    return (an >= 0x1p-30f) & (an <= 0x1p30f) & (d >= 0x1p-30f) & (d <= 0x1p30f);
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

// True on every active lane (a wave-uniform vote without exec branching).
// The tests it combines are 0/1 values joined with &, which keeps them
// free of short-circuit branches.
__device__ __forceinline__ bool wave_all(unsigned p) { return __builtin_amdgcn_ballot_w64(p == 0u) == 0; }

// s = gx^2 + gy^2 whose rounded root m satisfies both fast paths: s = +0 or
// s in [2^-96, S*], where S* (0x3e2fb0cd) is the largest float whose
// correctly rounded root is <= tan(pi/8) as float32. Then m = sqrt_fast(s)
// and Atan32(m) = atan32_small(m).
constexpr uint32_t kSlopeSmallMax = 0x3e2fb0cdu;
__device__ __forceinline__ unsigned slope_small_ok(float s)
{
    const uint32_t u = f2u(s);
    return (u == 0u) | ((u - 0x0f800000u) <= (kSlopeSmallMax - 0x0f800000u));
}

// Slope of a pair of cells; the fast paths are taken by whole waves.
__device__ __forceinline__ void slope_pair(float gx0, float gy0, float gx1, float gy1, float scale, bool atan_,
                                           float &o0, float &o1)
{
    const float s0 = gx0 * gx0 + gy0 * gy0, s1 = gx1 * gx1 + gy1 * gy1;
    if (atan_ && wave_all(slope_small_ok(s0) & slope_small_ok(s1))) {
        o0 = atan32_small(sqrt_fast(s0)) * scale;
        o1 = atan32_small(sqrt_fast(s1)) * scale;
        return;
    }
    float m0, m1;
    if (wave_all(sqrt_fast_ok(s0) & sqrt_fast_ok(s1))) {
        m0 = sqrt_fast(s0);
        m1 = sqrt_fast(s1);
    } else {
        m0 = __builtin_sqrtf(s0);
        m1 = __builtin_sqrtf(s1);
    }
    if (!atan_) {
        o0 = m0 * scale;
        o1 = m1 * scale;
    } else if (wave_all((m0 <= kTanPi8) & (m1 <= kTanPi8))) {
        o0 = atan32_small(m0) * scale;
        o1 = atan32_small(m1) * scale;
    } else {
        o0 = atan32(m0) * scale;
        o1 = atan32(m1) * scale;
    }
}

// n / d for a pair, with the fast quotient when the whole wave is in its
// range.
__device__ __forceinline__ void div_pair(float n0, float d0, float n1, float d1, float &q0, float &q1)
{
    if (wave_all(div_fast_ok(n0, d0) & div_fast_ok(n1, d1))) {
        q0 = div_fast(n0, d0);
        q1 = div_fast(n1, d1);
    } else {
        q0 = n0 / d0;
        q1 = n1 / d1;
    }
}

// atan32 of a pair (x >= 0, +Inf or NaN): the same reduction, its quotient
// through div_pair. Below tan(pi/8) the reduction divides by 1, which is
// exact, so t = x there and that lane lends the quotient a 1 / 1.
__device__ __forceinline__ void atan32_pair(float x0, float x1, float &a0, float &a1)
{
    auto reduce = [](float x, float &num, float &den, float &y0, bool &small) {
        num = x;
        den = 1.0f;
        y0 = 0.0f;
        const bool mid = x > kTanPi8, hi = x > kTan3Pi8;
        if (mid) {
            num = x - 1.0f;
            den = x + 1.0f;
            y0 = u2f(kPi4Bits);
        }
        if (hi) {
            num = -1.0f;
            den = x;
            y0 = u2f(kPi2Bits);
        }
        small = !mid;  // x <= tan(pi/8), or NaN
    };
    float n0, d0, y00, n1, d1, y01;
    bool s0, s1;
    reduce(x0, n0, d0, y00, s0);
    reduce(x1, n1, d1, y01, s1);
    float t0, t1;
    div_pair(s0 ? 1.0f : n0, s0 ? 1.0f : d0, s1 ? 1.0f : n1, s1 ? 1.0f : d1, t0, t1);
    t0 = s0 ? n0 : t0;  // x / 1 == x (x is a quotient, so a NaN is already quiet)
    t1 = s1 ? n1 : t1;
    auto poly = [](float t, float y0) {
        const float z = t * t;
        float p = kAtanC4 * z - kAtanC3;
        p = p * z + kAtanC2;
        p = p * z - kAtanC1;
        return y0 + ((p * z) * t + t);
    };
    a0 = poly(t0, y00);
    a1 = poly(t1, y01);
}

// aspect_of for a pair: Atan2F32's quotient |y| / |x| and atan32's through
// the pair forms, the rest as aspect_of (internal/stencil/aspect.go:36-48,
// 94-123).
__device__ __forceinline__ void aspect_pair(float gx0, float gy0, float gx1, float gy1, float flat, bool trig,
                                            float &o0, float &o1)
{
    const float y0 = trig ? gy0 : 0.0f - gx0, x0 = trig ? 0.0f - gx0 : gy0;
    const float y1 = trig ? gy1 : 0.0f - gx1, x1 = trig ? 0.0f - gx1 : gy1;
    const float ax0 = u2f(f2u(x0) & 0x7fffffffu), ay0 = u2f(f2u(y0) & 0x7fffffffu);
    const float ax1 = u2f(f2u(x1) & 0x7fffffffu), ay1 = u2f(f2u(y1) & 0x7fffffffu);
    float q0, q1, a0, a1;
    div_pair(ay0, ax0, ay1, ax1, q0, q1);
    atan32_pair(q0, q1, a0, a1);
    auto finish = [&](float a, float x, float y, float ax, float ay) {
        if (ay == 0.0f && ax == 0.0f) a = 0.0f;
        if (ay == ax && ay > 3.40282346638528859811704183484516925440e+38f) a = u2f(kPi4Bits);
        if (f2u(x) & 0x80000000u) a = u2f(kPiBits) - a;
        a = u2f(f2u(a) | (f2u(y) & 0x80000000u));
        float d = a * u2f(kRadToDegBits);
        float off = 0.0f;
        if (d < 0.0f) off = 360.0f;
        d += off;
        if (d >= 360.0f) d = 0.0f;
        if (y == 0.0f && x == 0.0f) d = flat;
        return d;
    };
    o0 = finish(a0, x0, y0, ax0, ay0);
    o1 = finish(a1, x1, y1, ax1, ay1);
}

// Hillshade of a pair (shade_of with the fast root and quotient when the
// whole wave is in their ranges).
__device__ __forceinline__ void shade_pair(float gx0, float gy0, float gx1, float gy1, float c, float bx, float by,
                                           float &o0, float &o1)
{
    const float n0 = c + (bx * gx0 + by * gy0), n1 = c + (bx * gx1 + by * gy1);
    const float w0 = 1.0f + (gx0 * gx0 + gy0 * gy0), w1 = 1.0f + (gx1 * gx1 + gy1 * gy1);
    float v0, v1;
    if (wave_all(sqrt_fast_ok(w0) & sqrt_fast_ok(w1))) {
        const float d0 = sqrt_fast(w0), d1 = sqrt_fast(w1);
        if (wave_all(div_fast_ok(n0, d0) & div_fast_ok(n1, d1))) {
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

// Horn's scaled gradient (internal/stencil/horn.go:172-180) of a lane's two
// cells from rows A (y-1), B (y), C (y+1), grouped as the reference groups
// it; the differences are shared between rows and columns.
__device__ __forceinline__ void horn_grad(const HornArgs &a, const Row &A, const Row &B, const Row &C, float &gx0,
                                          float &gy0, float &gx1, float &gy1)
{
    const float dx0 = (A.hd0 + C.hd0) + (B.hd0 + B.hd0);
    const float dx1 = (A.hd1 + C.hd1) + (B.hd1 + B.hd1);
    const float gl = C.l - A.l, g0 = C.c0 - A.c0, g1 = C.c1 - A.c1, gr = C.r - A.r;
    const float dy0 = (gl + g1) + (g0 + g0);
    const float dy1 = (g0 + gr) + (g1 + g1);
    gx0 = dx0 * a.k[0];
    gy0 = dy0 * a.k[1];
    gx1 = dx1 * a.k[0];
    gy1 = dy1 * a.k[1];
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
        float gx0, gy0, gx1, gy1;
        horn_grad(a, A, B, C, gx0, gy0, gx1, gy1);
        if constexpr (OP == STRATA_OP_GRADIENT) {
            o00 = gx0;
            o01 = gx1;
            o10 = gy0;
            o11 = gy1;
        } else if constexpr (OP == STRATA_OP_SLOPE) {
            slope_pair(gx0, gy0, gx1, gy1, a.k[2], (a.flags & 1) != 0, o00, o01);
        } else if constexpr (OP == STRATA_OP_ASPECT) {
            aspect_pair(gx0, gy0, gx1, gy1, a.k[2], (a.flags & 2) != 0, o00, o01);
        } else {
            shade_pair(gx0, gy0, gx1, gy1, a.k[2], a.k[3], a.k[4], o00, o01);
        }
    }
}

// The identity on a sweep's output values (a lane's two).
struct NoPost {
    __device__ void operator()(float &, float &) const {}
};

// LDS memory, for pointers that reach a function through a structure.
#define STRATA_LDS __attribute__((address_space(3)))

// A point op's table in LDS for the counting forms of Lookup and Reclass
// (arith.hpp): a Lookup's knots as (x, y) pairs, a Reclass's nb + 1 class
// values. Each copy takes an even number of floats, so copies placed one
// after another from an 8-byte-aligned start keep their pairs aligned.
__device__ __forceinline__ int lds_table_floats(int op, int cnt)
{
    return op == STRATA_OP_LOOKUP ? cnt : (((cnt - 1) / 2 + 2) & ~1);
}
// Copies op's table (cnt floats at t in the C ABI's layout: xs then ys, or
// breaks then values) into lt. Every thread of the block calls it, before
// a barrier; tid is the thread's index among nthreads.
template <class T>
__device__ __forceinline__ void lds_table(STRATA_LDS float *lt, int op, T t, int cnt, int tid, int nthreads)
{
    if (op == STRATA_OP_LOOKUP) {
        const int n = cnt / 2;
        for (int m = tid; m < cnt; m += nthreads) lt[m] = t[(m & 1) ? n + (m >> 1) : (m >> 1)];
    } else {
        const int nb = (cnt - 1) / 2;
        for (int m = tid; m <= nb; m += nthreads) lt[m] = t[nb + m];
    }
}
// op's values of v, in place (internal/curve/scalar.go:39-114): the
// table's x side is read at uniform indices from t, the entries picked are
// gathered from its LDS copy lt.
template <int N, class T>
__device__ __forceinline__ void point_values(float (&v)[N], int op, T t, int cnt, const STRATA_LDS float *lt)
{
    if (op == STRATA_OP_LOOKUP) {
        const int n = cnt / 2;
        lookup_count(v, t, t + n, n, reinterpret_cast<const STRATA_LDS pair_t *>(lt));
    } else {
        const int nb = (cnt - 1) / 2;
        reclass_count(v, t, nb, lt);
    }
}

// A chain's point stages on a sweep's output values, in order, so a stencil
// followed by point stages writes the chain's value straight from
// registers. The tables are in device memory that the launch only reads
// (x sides read as uniform values) and in the block's LDS, one copy after
// another in stage order (lds_table).
constexpr int kPostLdsFloats = kPostTaps + 2 * kMaxPost;
struct LdsPost {
    const STRATA_CONST PostArgs *a;
    const STRATA_LDS float *lt;
    // A lane's two values.
    __device__ void operator()(float &v0, float &v1) const
    {
        float v[2] = {v0, v1};
        int lo = 0;
        for (int i = 0; i < a->n; ++i) {
            const int op = a->op[i], c = a->cnt[i];
            point_values(v, op, a->taps + a->off[i], c, lt + lo);
            lo += lds_table_floats(op, c);
        }
        v0 = v[0];
        v1 = v[1];
    }
};

// MODE (wave-uniform, chosen per block or unit):
//   kSweepChecked   every cell loaded and stored on its own, columns clamped
//                   into the buffer and stores kept inside the region;
//   kSweepPairs     every pair of the block lies inside the buffer and the
//                   region, at 8-byte-aligned addresses, so rows move as
//                   dwordx2 loads and stores;
//   kSweepAnyParity the block's columns lie inside the buffer and the
//                   region, at any alignment: each row (uniform per row)
//                   loads and stores aligned pairs, one cell to the left
//                   when its first cell is at an odd 4-byte address, and
//                   the lanes' cells are formed with the same DPP moves.
//                   A lane 0 pair that starts one cell before the buffer
//                   lies in the 8-byte block of the buffer's first cell
//                   and feeds only an edge output; likewise lane 63's pair
//                   that ends one cell past it.
// `post`
// maps every stored value (edge NaNs included) of the first output.
// tid: the lane's index among the lanes sharing columns from xb (a block's
// 256 or a wave's 64); each lane takes columns xb + 2*tid and xb + 2*tid + 1.
// For STRATA_OP_SURFACE, dset holds one view per wanted product (a.flags'
// STRATA_SURF_* bits, in output order) and d0, d1 are unused.
// For kOpHornFan, dset holds output o's view and fan its op.
constexpr int kSweepChecked = 0, kSweepPairs = 1, kSweepAnyParity = 2;

template <int OP, int MODE, class Post = NoPost>
__device__ __forceinline__ void sweep(const HornArgs &a, const Region &rg, const SrcView &s, const DstView &d0,
                                      const DstView &d1, int64_t xb, int64_t y0, int64_t y1, uint32_t tid,
                                      const Post &post = Post(), const DstView *dset = nullptr,
                                      const FanArgs *fan = nullptr)
{
    constexpr bool kTwo = OP == STRATA_OP_GRADIENT;
    const int lane = (int)(tid & 63);
    const uint32_t off = 2u * tid;  // the pair's first column, relative to xb
    const int64_t x = xb + off;
    const bool halo_lane = (uint32_t)(lane - 1) >= 62u;  // lanes 0 and 63
    // Byte offsets from a row's start of the cells this lane loads, clamped
    // into the buffer: a clamped cell is read only for edge outputs, which
    // become NaN. They are never negative (the buffer starts at or after sx)
    // and are 32-bit, so each access is a uniform row address plus a lane
    // offset, with no per-lane 64-bit address arithmetic.
    auto col = [&](int64_t c) { return 4u * (uint32_t)(min(max(c, s.lx0), s.lx1 - 1) - s.sx); };
    const uint32_t o0 = col(x), o1 = col(x + 1), oh = col(lane == 0 ? x - 1 : x + 2);
    // Every lane issues the halo load of a row, lanes other than 0 and 63 at
    // their own first cell (a cache hit): with the same loads on every lane
    // and path, the row loaded ahead stays in flight while the current row
    // is computed (a load under a divergent branch makes the wave wait for
    // every outstanding access, stores included, before each row).
    const uint32_t oho = halo_lane ? oh : o0;
    const uint32_t so = 4u * ((uint32_t)(xb - rg.ox) + off);  // store byte offset in a region row
    const bool ci0 = x >= 1 && x < rg.W - 1;
    const bool ci1 = x + 1 >= 1 && x + 1 < rg.W - 1;
    const bool in0 = x >= rg.ox && x < rg.ox + rg.w;
    const bool in1 = x + 1 >= rg.ox && x + 1 < rg.ox + rg.w;
    typedef const STRATA_GLOBAL char *crow_t;
    typedef STRATA_GLOBAL char *row_t;

    // Source rows are clamped into the buffer in 32 bits relative to y0.
    const int32_t rlo = (int32_t)min(max(s.ly0 - y0, (int64_t)-2), (int64_t)0);
    const int32_t rhi = (int32_t)min(max(s.ly1 - 1 - y0, (int64_t)-2), (int64_t)4097);
    // kSweepAnyParity: byte offsets of the block's first cell from a row's
    // start, in the source and in an output (uniform).
    const uint32_t lbase = 4u * (uint32_t)(xb - s.sx), sbase = 4u * (uint32_t)(xb - rg.ox);
    auto load = [&](Row &row, int32_t u) {  // row y0 + u
        const int32_t uc = min(max(u, rlo), rhi);
        const crow_t rp = reinterpret_cast<crow_t>(gp(s.p) + (y0 + uc - s.sy) * s.stride);
        float c0, c1, h = 0.0f, h2 = 0.0f;
        row.odd = false;
        if constexpr (MODE == kSweepPairs) {
            const pair_t v = *reinterpret_cast<const STRATA_GLOBAL pair_t *>(rp + o0);
            c0 = v.x;
            c1 = v.y;
        } else if constexpr (MODE == kSweepAnyParity) {
            row.odd = ((reinterpret_cast<uintptr_t>(rp) + lbase) & 4u) != 0u;
            if (!row.odd) {
                const pair_t v = *reinterpret_cast<const STRATA_GLOBAL pair_t *>(rp + o0);
                c0 = v.x;
                c1 = v.y;
                h = *reinterpret_cast<const STRATA_GLOBAL float *>(rp + oho);
            } else {
                // Cells x - 1, x; lane 63 also x + 1, x + 2 (its right
                // neighbour's pair), which the other lanes load as their own
                // pair again.
                const pair_t v = *reinterpret_cast<const STRATA_GLOBAL pair_t *>(rp + o0 - 4u);
                c0 = v.x;
                c1 = v.y;
                // The step is signed and added to the address: o0 is 0 for
                // lane 0 at the buffer's first cell, and o0 - 4 in 32 bits
                // would wrap to 4 GiB past the row.
                const int32_t es = lane == 63 ? 4 : -4;
                const pair_t e = *reinterpret_cast<const STRATA_GLOBAL pair_t *>(rp + o0 + es);
                h = e.x;
                h2 = e.y;
            }
        } else {
            c0 = *reinterpret_cast<const STRATA_GLOBAL float *>(rp + o0);
            c1 = *reinterpret_cast<const STRATA_GLOBAL float *>(rp + o1);
        }
        if constexpr (MODE != kSweepAnyParity) h = *reinterpret_cast<const STRATA_GLOBAL float *>(rp + oho);
        row.c0 = c0;
        row.c1 = c1;
        row.l = h;  // the halo values until finish() forms l and r
        row.r = h2;
    };
    auto finish = [&](Row &row) {
        if (MODE == kSweepAnyParity && row.odd) {
            // Loaded: c0 = cell x - 1, c1 = cell x; lane 63: l, r = cells
            // x + 1, x + 2. Lane + 1's pair holds cells x + 1, x + 2.
            const float xm1 = row.c0, x0 = row.c1;
            const float x1 = from_right(xm1, row.l);
            const float x2 = from_right(x0, row.r);
            row.l = xm1;
            row.c0 = x0;
            row.c1 = x1;
            row.r = x2;
        } else {
            const float h = row.l;
            row.l = from_left(row.c1, h);
            row.r = from_right(row.c0, h);
        }
        row.hd0 = row.c1 - row.l;
        row.hd1 = row.r - row.c0;
    };
    auto store = [&](const DstView &d, int64_t y, float v0, float v1) {
        const row_t rp = reinterpret_cast<row_t>(gp(d.p) + (y - rg.oy) * d.stride);
        if constexpr (MODE == kSweepPairs) {
            // Outputs are not read again by this launch: stream them.
            __builtin_nontemporal_store(pair_t{v0, v1}, reinterpret_cast<STRATA_GLOBAL pair_t *>(rp + so));
        } else if constexpr (MODE == kSweepAnyParity) {
            if (((reinterpret_cast<uintptr_t>(rp) + sbase) & 4u) == 0u) {
                __builtin_nontemporal_store(pair_t{v0, v1}, reinterpret_cast<STRATA_GLOBAL pair_t *>(rp + so));
            } else {
                // Odd 4-byte offset: each lane's two cells as two 4-byte
                // stores (every lane's cells lie inside the region here);
                // the cache joins the two halves of each line.
                __builtin_nontemporal_store(v0, reinterpret_cast<STRATA_GLOBAL float *>(rp + so));
                __builtin_nontemporal_store(v1, reinterpret_cast<STRATA_GLOBAL float *>(rp + so + 4u));
            }
        } else {
            if (in0) *reinterpret_cast<STRATA_GLOBAL float *>(rp + so) = v0;
            if (in1) *reinterpret_cast<STRATA_GLOBAL float *>(rp + so + 4u) = v1;
        }
    };

    // Rows are counted from y0 in 32 bits (a strip is at most 4096 rows), so
    // the loop tests and the interior test are scalar 32-bit compares.
    const int32_t nrows = (int32_t)(y1 - y0);
    const int32_t ilo = (int32_t)min(max(1 - y0, (int64_t)-1), (int64_t)nrows + 1);          // rows >= 1
    const int32_t ihi = (int32_t)min(max(rg.H - 1 - y0, (int64_t)-1), (int64_t)nrows + 1);   // rows < H-1
    Row R[4];
    load(R[0], -1);
    load(R[1], 0);
    load(R[2], 1);
    finish(R[0]);
    finish(R[1]);
    // One output row from ring slot u (rows A, B, C; D takes the next row).
    auto step = [&](const int u, const int32_t uu) __attribute__((always_inline)) {
        const int64_t yy = y0 + uu;
        Row &A = R[u & 3];
        Row &B = R[(u + 1) & 3];
        Row &C = R[(u + 2) & 3];
        Row &D = R[(u + 3) & 3];
        // Always issued (past the strip it reloads the strip's last row, a
        // cache hit), so no wait has to cover a load that may not exist.
        load(D, min(uu + 2, nrows));
        finish(C);
        const bool ri = uu >= ilo && uu < ihi;
        const float nan = edge_nan();
        if constexpr (OP == STRATA_OP_SURFACE) {
            // terrain.Surface (terrain/surface.go:52-102): every wanted
            // product from one gradient, each with its op's arithmetic.
            float gx0, gy0, gx1, gy1;
            horn_grad(a, A, B, C, gx0, gy0, gx1, gy1);
            auto put = [&](int o, float v0, float v1) {
                store(dset[o], yy, (ri && ci0) ? v0 : nan, (ri && ci1) ? v1 : nan);
            };
            const int f = a.flags;
            int o = 0;
            if (f & STRATA_SURF_DX) put(o++, gx0, gx1);
            if (f & STRATA_SURF_DY) put(o++, gy0, gy1);
            if (f & STRATA_SURF_SLOPE) {
                float v0, v1;
                slope_pair(gx0, gy0, gx1, gy1, a.k[2], (f & STRATA_FLAG_ATAN) != 0, v0, v1);
                put(o++, v0, v1);
            }
            if (f & STRATA_SURF_ASPECT) {
                float v0, v1;
                aspect_pair(gx0, gy0, gx1, gy1, a.k[3], (f & STRATA_FLAG_TRIG) != 0, v0, v1);
                put(o++, v0, v1);
            }
            if (f & STRATA_SURF_HILLSHADE) {
                float v0, v1;
                shade_pair(gx0, gy0, gx1, gy1, a.k[4], a.k[5], a.k[6], v0, v1);
                put(o++, v0, v1);
            }
        } else if constexpr (OP == kOpHornFan) {
            // terrain.Features (terrain/features.go:68-117): each output
            // is its own op with its own constants. The unscaled Horn
            // differences are the same for every op; each scales them
            // by its own kx, ky, as the op alone does.
            const float dx0 = (A.hd0 + C.hd0) + (B.hd0 + B.hd0);
            const float dx1 = (A.hd1 + C.hd1) + (B.hd1 + B.hd1);
            const float gl = C.l - A.l, g0 = C.c0 - A.c0, g1 = C.c1 - A.c1, gr = C.r - A.r;
            const float dy0 = (gl + g1) + (g0 + g0);
            const float dy1 = (g0 + gr) + (g1 + g1);
            // Rolled: one copy of each op's code however many outputs
            // (unrolled, the kernel outgrows the instruction cache).
#pragma unroll 1
            for (int o = 0; o < kMaxOutputs; ++o) {
                if (o < fan->n) {
                    const float *k = fan->k[o];
                    const int f = fan->flags[o];
                    const float gx0 = dx0 * k[0], gy0 = dy0 * k[1];
                    const float gx1 = dx1 * k[0], gy1 = dy1 * k[1];
                    float v0, v1;
                    switch (fan->op[o]) {
                    case STRATA_OP_SLOPE:
                        slope_pair(gx0, gy0, gx1, gy1, k[2], (f & STRATA_FLAG_ATAN) != 0, v0, v1);
                        break;
                    case STRATA_OP_ASPECT:  // the short form: this loop's code stays small
                        v0 = aspect_of(gx0, gy0, k[2], (f & STRATA_FLAG_TRIG) != 0);
                        v1 = aspect_of(gx1, gy1, k[2], (f & STRATA_FLAG_TRIG) != 0);
                        break;
                    case STRATA_OP_HILLSHADE:
                        shade_pair(gx0, gy0, gx1, gy1, k[2], k[3], k[4], v0, v1);
                        break;
                    default: {  // STRATA_OP_CURVATURE
                        const int kind = (f >> STRATA_CURV_SHIFT) & 3;
                        v0 = zt_curvature(A.l, A.c0, A.c1, B.l, B.c0, B.c1, C.l, C.c0, C.c1, k, kind);
                        v1 = zt_curvature(A.c0, A.c1, A.r, B.c0, B.c1, B.r, C.c0, C.c1, C.r, k, kind);
                    }
                    }
                    store(dset[o], yy, (ri && ci0) ? v0 : nan, (ri && ci1) ? v1 : nan);
                }
            }
        } else {
            float o00 = 0.0f, o01 = 0.0f, o10 = 0.0f, o11 = 0.0f;
            horn_pair<OP>(a, A, B, C, o00, o01, o10, o11);
            float p0 = (ri && ci0) ? o00 : nan, p1 = (ri && ci1) ? o01 : nan;
            post(p0, p1);
            store(d0, yy, p0, p1);
            if constexpr (kTwo) store(d1, yy, (ri && ci0) ? o10 : nan, (ri && ci1) ? o11 : nan);
        }    };
    // Whole groups of four rows with no condition inside, so the ring keeps
    // fixed registers and every step issues the same memory operations; then
    // the strip's last rows.
    int32_t ub = 0;
    for (; ub + 4 <= nrows; ub += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) step(u, ub + u);
    }
#pragma unroll
    for (int u = 0; u < 3; ++u)
        if (ub + u < nrows) step(u, ub + u);
}

// UNALIGNED: the operands do not allow pair accesses at even offsets, so
// blocks inside the buffer and the region take kSweepAnyParity. A separate
// instantiation: the aligned kernel keeps only the paths it runs (with the
// third path beside them, the aligned sweep ran 2% slower).
template <int OP, bool UNALIGNED>
__global__ __launch_bounds__(kSweepThreads) void horn_sweep(HornArgs a, Region rg, SrcView s, DstView d0,
                                                            DstView d1)
{
    // Workgroup ids are dealt round-robin to the chip's XCDs, each with its
    // own L2. Renumber so each XCD takes a contiguous run of tiles in raster
    // order: side-by-side tiles share the cache lines of their halo columns,
    // and tiles one above the other their halo rows, which then stay in the
    // same L2 instead of being fetched from HBM once per XCD.
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t xb = rg.ox + (int64_t)(t % a.blocks_x) * kSweepCols;
    const int64_t y0 = rg.oy + (int64_t)(t / a.blocks_x) * a.rows;
    const int64_t y1 = min(y0 + (int64_t)a.rows, rg.oy + rg.h);
    const bool inside = xb >= s.lx0 && xb + kSweepCols <= s.lx1 && xb >= rg.ox && xb + kSweepCols <= rg.ox + rg.w;
    if constexpr (UNALIGNED) {
        if (inside)
            sweep<OP, kSweepAnyParity>(a, rg, s, d0, d1, xb, y0, y1, threadIdx.x);
        else
            sweep<OP, kSweepChecked>(a, rg, s, d0, d1, xb, y0, y1, threadIdx.x);
    } else {
        if (inside && a.aligned)
            sweep<OP, kSweepPairs>(a, rg, s, d0, d1, xb, y0, y1, threadIdx.x);
        else
            sweep<OP, kSweepChecked>(a, rg, s, d0, d1, xb, y0, y1, threadIdx.x);
    }
}

// Copies a chain's point stages into device memory.
__global__ __launch_bounds__(256) void stage_post(PostArgs pa, PostArgs *out)
{
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&pa);
    uint32_t *dst = reinterpret_cast<uint32_t *>(out);
    for (uint32_t i = threadIdx.x; i < sizeof(PostArgs) / 4; i += 256) dst[i] = src[i];
}

hipError_t launch_stage_post(const PostArgs &pa, PostArgs *dev, hipStream_t st)
{
    static_assert(sizeof(PostArgs) % 4 == 0, "PostArgs is copied in 32-bit words");
    hipLaunchKernelGGL(stage_post, dim3(1), dim3(256), 0, st, pa, dev);
    return hipGetLastError();
}

// horn_sweep with a chain's point stages (in device memory) applied to
// every stored value (single-output ops).
template <int OP, bool UNALIGNED>
__global__ __launch_bounds__(kSweepThreads) void horn_sweep_post(HornArgs a, Region rg, SrcView s, DstView d0,
                                                                 const PostArgs *__restrict__ pa)
{
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t xb = rg.ox + (int64_t)(t % a.blocks_x) * kSweepCols;
    const int64_t y0 = rg.oy + (int64_t)(t / a.blocks_x) * a.rows;
    const int64_t y1 = min(y0 + (int64_t)a.rows, rg.oy + rg.h);
    const bool inside = xb >= s.lx0 && xb + kSweepCols <= s.lx1 && xb >= rg.ox && xb + kSweepCols <= rg.ox + rg.w;
    __shared__ __align__(16) float lt[kPostLdsFloats];
    const STRATA_CONST PostArgs *pc = cp(pa);
    int lo = 0;
    for (int i = 0; i < pc->n; ++i) {
        const int op = pc->op[i], c = pc->cnt[i];
        lds_table((STRATA_LDS float *)lt + lo, op, pc->taps + pc->off[i], c, (int)threadIdx.x, kSweepThreads);
        lo += lds_table_floats(op, c);
    }
    __syncthreads();
    const LdsPost post{pc, (const STRATA_LDS float *)lt};
    if constexpr (UNALIGNED) {
        if (inside)
            sweep<OP, kSweepAnyParity>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x, post);
        else
            sweep<OP, kSweepChecked>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x, post);
    } else {
        if (inside && a.aligned)
            sweep<OP, kSweepPairs>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x, post);
        else
            sweep<OP, kSweepChecked>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x, post);
    }
}

// The output views of one launch, by value in the kernel arguments.
struct DstSet {
    DstView v[kMaxOutputs];
};

// terrain.Surface at radius 1: horn_sweep's layout, every wanted product.
__global__ __launch_bounds__(kSweepThreads) void horn_surface(HornArgs a, Region rg, SrcView s, DstSet d)
{
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t xb = rg.ox + (int64_t)(t % a.blocks_x) * kSweepCols;
    const int64_t y0 = rg.oy + (int64_t)(t / a.blocks_x) * a.rows;
    const int64_t y1 = min(y0 + (int64_t)a.rows, rg.oy + rg.h);
    const bool pairs = a.aligned && xb >= s.lx0 && xb + kSweepCols <= s.lx1 && xb >= rg.ox &&
                       xb + kSweepCols <= rg.ox + rg.w;
    if (pairs)
        sweep<STRATA_OP_SURFACE, kSweepPairs>(a, rg, s, d.v[0], d.v[0], xb, y0, y1, threadIdx.x, NoPost(), d.v);
    else
        sweep<STRATA_OP_SURFACE, kSweepChecked>(a, rg, s, d.v[0], d.v[0], xb, y0, y1, threadIdx.x, NoPost(), d.v);
}

// A fan-out of radius-1 terrain ops: horn_sweep's layout, output o is
// f.op[o] over the shared rows.
__global__ __launch_bounds__(kSweepThreads) void horn_fan(HornArgs a, FanArgs f, Region rg, SrcView s, DstSet d)
{
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t xb = rg.ox + (int64_t)(t % a.blocks_x) * kSweepCols;
    const int64_t y0 = rg.oy + (int64_t)(t / a.blocks_x) * a.rows;
    const int64_t y1 = min(y0 + (int64_t)a.rows, rg.oy + rg.h);
    const bool pairs = a.aligned && xb >= s.lx0 && xb + kSweepCols <= s.lx1 && xb >= rg.ox &&
                       xb + kSweepCols <= rg.ox + rg.w;
    if (pairs)
        sweep<kOpHornFan, kSweepPairs>(a, rg, s, d.v[0], d.v[0], xb, y0, y1, threadIdx.x, NoPost(), d.v, &f);
    else
        sweep<kOpHornFan, kSweepChecked>(a, rg, s, d.v[0], d.v[0], xb, y0, y1, threadIdx.x, NoPost(), d.v, &f);
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
    float k[12];
};
struct ArgTaps {
    const Params &P;
    __device__ float operator[](int i) const { return P.taps[i]; }
};
struct PtrTaps {
    const float *p;
    __device__ float operator[](int i) const { return p[i]; }
};

// Correlate weights in LDS: rows of 2R+1 padded to a multiple of four, so a
// weight row is a run of aligned 16-byte broadcast reads.
template <int R>
constexpr int weight_pitch()
{
    return (2 * R + 1 + 3) & ~3;
}

// LDS floats one focal_block<KIND, R> uses.
template <int KIND, int R>
constexpr int focal_lds_floats()
{
    return (kTileY + 2 * R) * (kTileX + 2 * R) +
           ((KIND == FK_CORRELATE) ? (2 * R + 1) * weight_pitch<R>()
                                   : (KIND == FK_FIT ? 3 : 1) * kTileY * (kTileX + 2 * R));
}

// One kTileX x kTileY output tile at (tx0, ty0) of the region.
// For terrain.Surface through the fit (P.op STRATA_OP_SURFACE), dset holds
// one view per wanted product, in output order.
// UNROLL (a kernel per op, not the batch's kernel, which inlines every
// kind): stage cell by cell and run the column pass with compile-time trip
// counts, several LDS reads in flight together. Inlined beside the other
// kinds inside the batch's tile loop, that form needs several times the
// registers; the batch stages row by row and keeps the column loop rolled.
// IN_LOOP (a caller that runs tiles in a loop): the thread index passes
// through an empty asm, so the staging offsets derived from it are formed
// per tile instead of hoisted out of the loop and held in registers.
template <int KIND, int R, class TAPS, bool UNROLL = true, bool IN_LOOP = false>
__device__ __forceinline__ void focal_block(const FocalArgs &P, const TAPS &taps, const Region &rg,
                                            const SrcView &s, const DstView &d0, const DstView &d1, int64_t tx0,
                                            int64_t ty0, float *lds, const DstView *dset = nullptr)
{
    constexpr int N = 2 * R + 1;
    constexpr int LW = kTileX + 2 * R;  // staged columns
    constexpr int LH = kTileY + 2 * R;  // staged rows
    constexpr int CY = (KIND == FK_CORRELATE) ? 1 : kTileY;
    float(*tile)[LW] = reinterpret_cast<float(*)[LW]>(lds);
    float(*colv)[CY][LW] = reinterpret_cast<float(*)[CY][LW]>(lds + LH * LW);

    int tid = threadIdx.x;
    if constexpr (IN_LOOP) __asm__ volatile("" : "+v"(tid));
    // Staging. Both forms issue every load of the thread before any LDS
    // write. UNROLL (a kernel per op): cells in thread order, unchecked
    // loads when the staged window lies inside the buffer. Otherwise (the
    // batch's kernel) row by row: wave w takes staged rows w, w + 4, ...,
    // lane l columns l and 64 + l (LW <= 80), with uniform row addresses
    // and no per-cell address arithmetic for the compiler to hoist out of
    // the batch's tile loop; rows and columns outside the buffer are
    // clamped into it and read as zeros.
    if constexpr (UNROLL) {
        constexpr int kStage = (LH * LW + kTileThreads - 1) / kTileThreads;
        float sv[kStage];
        const int64_t bx = tx0 - R, by = ty0 - R;
        if (bx >= s.lx0 && bx + LW <= s.lx1 && by >= s.ly0 && by + LH <= s.ly1 &&
            (int64_t)(LH - 1) * s.stride + LW < ((int64_t)1 << 29)) {
            // Inside the buffer, with the window's byte offsets in 32 bits:
            // every load at a 32-bit offset from one uniform base, the
            // offsets stepped from cell to cell (256 cells on: DR rows and
            // DC columns), so no load waits on address arithmetic. Cells
            // past the window read its last cell, which is not staged.
            constexpr int DR = kTileThreads / LW, DC = kTileThreads % LW;
            const char *base = reinterpret_cast<const char *>(s.p + ((by - s.sy) * s.stride + (bx - s.sx)));
            const uint32_t st4 = (uint32_t)s.stride * 4u;
            const uint32_t last = (uint32_t)(LH - 1) * st4 + 4u * (uint32_t)(LW - 1);
            int row = tid / LW, col = tid - (tid / LW) * LW;
            uint32_t off = (uint32_t)row * st4 + 4u * (uint32_t)col;
#pragma unroll
            for (int k = 0; k < kStage; ++k) {
                sv[k] = *reinterpret_cast<const STRATA_GLOBAL float *>(gp(base) + (row < LH ? off : last));
                row += DR;
                col += DC;
                off += DR * st4 + 4u * DC;
                if (col >= LW) {
                    col -= LW;
                    row += 1;
                    off += st4 - 4u * LW;
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kStage; ++k) {
                const int i = min(tid + k * kTileThreads, LH * LW - 1);
                const int row = i / LW, col = i - row * LW;
                sv[k] = load_cell(s, bx + col, by + row);
            }
        }
        __syncthreads();  // the LDS may still be read by the previous tile
#pragma unroll
        for (int k = 0; k < kStage; ++k) {
            const int i = tid + k * kTileThreads;
            if (i < LH * LW) lds[i] = sv[k];
        }
    } else {
        constexpr int kStageRows = (LH + 3) / 4;
        const int sw = tid >> 6, sl = tid & 63;
        const int64_t bx = tx0 - R, by = ty0 - R;
        const bool cols_in = bx >= s.lx0 && bx + LW <= s.lx1;  // uniform
        const int64_t ca = min(max(bx + sl, s.lx0), s.lx1 - 1), cb = min(max(bx + 64 + sl, s.lx0), s.lx1 - 1);
        const bool in_a = cols_in || ca == bx + sl;
        const bool in_b = sl + 64 < LW && (cols_in || cb == bx + 64 + sl);
        const int64_t oa = ca - s.sx, ob = cb - s.sx;
        float va[kStageRows], vb[kStageRows];
#pragma unroll
        for (int k = 0; k < kStageRows; ++k) {
            const int row = sw + 4 * k;
            const int64_t yy = by + row;
            const bool rok = row < LH && yy >= s.ly0 && yy < s.ly1;  // uniform
            const STRATA_GLOBAL float *rp = gp(s.p) + (min(max(yy, s.ly0), s.ly1 - 1) - s.sy) * s.stride;
            va[k] = rp[oa];
            vb[k] = rp[ob];
            va[k] = (rok && in_a) ? va[k] : 0.0f;
            vb[k] = (rok && in_b) ? vb[k] : 0.0f;
        }
        __syncthreads();  // the LDS may still be read by the previous tile of a batch
#pragma unroll
        for (int k = 0; k < kStageRows; ++k) {
            const int row = sw + 4 * k;
            if (row < LH) {
                tile[row][sl] = va[k];
                if (sl + 64 < LW) tile[row][64 + sl] = vb[k];
            }
        }
    }
    if constexpr (KIND == FK_CORRELATE && R > 2) {
        constexpr int WP = weight_pitch<R>();
        float *wts = lds + LH * LW;
        for (int i = tid; i < N * WP; i += kTileThreads) {
            const int j = i / WP, c = i - j * WP;
            wts[i] = c < N ? taps[j * N + c] : 0.0f;
        }
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const int64_t x = tx0 + lane;
    const int rbase = wave * kRowsPerWave;
    float out0[kRowsPerWave], out1[kRowsPerWave];

    if constexpr (KIND == FK_CORRELATE) {
        // Per output: w[0]*v(0,0), then acc + w[t]*v(t) over the weights in
        // row-major order (internal/focalrow/focalrow.go:217-230). Each input
        // row's 2R+1 values are read from LDS once into registers and
        // applied to every output row that uses that input row; for each
        // output the terms still arrive in row-major order. The sums start
        // at -0: -0 + t == t exactly for every t (signed zeros and NaN
        // included), so the first product starts the sum as in the
        // reference, without a select.
        constexpr int WP = weight_pitch<R>();
        const float *wts = lds + LH * LW;
        float acc[kRowsPerWave];
#pragma unroll
        for (int o = 0; o < kRowsPerWave; ++o) acc[o] = -0.0f;
        if constexpr (R <= 2) {
            // Few weights: the row loop unrolls and the weights are
            // constants read from the kernel arguments.
#pragma unroll
            for (int row = 0; row < kRowsPerWave + 2 * R; ++row) {
#pragma unroll
                for (int c = 0; c < N; ++c) {
                    const float v = tile[rbase + row][lane + c];
#pragma unroll
                    for (int o = 0; o < kRowsPerWave; ++o) {
                        const int j = row - o;
                        if (j >= 0 && j < N) acc[o] = acc[o] + taps[j * N + c] * v;
                    }
                }
            }
        } else {
            // Not unrolled: unrolled, the weights of every row would be held
            // in registers at once. Weight rows come from LDS.
#pragma unroll 1
            for (int row = 0; row < kRowsPerWave + 2 * R; ++row) {
                float v[N];
#pragma unroll
                for (int c = 0; c < N; ++c) v[c] = tile[rbase + row][lane + c];
#pragma unroll
                for (int o = 0; o < kRowsPerWave; ++o) {
                    const int j = row - o;
                    if (j >= 0 && j < N) {  // wave-uniform
                        const float *wr = wts + j * WP;
#pragma unroll
                        for (int c = 0; c < N; ++c) acc[o] = acc[o] + wr[c] * v[c];
                    }
                }
            }
        }
#pragma unroll
        for (int o = 0; o < kRowsPerWave; ++o) out0[o] = acc[o];
    } else {
        // Column pass for the tile's rows over the staged columns. Unrolled
        // (a compile-time trip count), the LDS reads of several cells are in
        // flight together.
        // The column value of cell i of the tile's rows (every kind but the fit).
        // Cell i of the tile's rows is tile[i / LW][i % LW], at lds[i]; the
        // cells below it are LW floats on.
        auto col_value = [&](int i) {
            const float *t = lds + i;
            float a;
            if constexpr (KIND == FK_SEPARABLE) {
                // ColumnCorrelate with the col taps (focalrow.go:232-242).
                a = taps[N] * t[0];
#pragma unroll
                for (int j = 1; j < N; ++j) a = a + taps[N + j] * t[j * LW];
            } else if constexpr (KIND == FK_MEAN) {
                a = t[0];  // ColumnSum (focalrow.go:244-250)
#pragma unroll
                for (int j = 1; j < N; ++j) a += t[j * LW];
            } else if constexpr (KIND == FK_MIN) {
                a = t[0];
#pragma unroll
                for (int j = 1; j < N; ++j) a = go_min(a, t[j * LW]);
            } else {
                a = t[0];
#pragma unroll
                for (int j = 1; j < N; ++j) a = go_max(a, t[j * LW]);
            }
            return a;
        };
        auto col_cell = [&](int i) {
            const int rr = i / LW, cc = i - rr * LW;
            if constexpr (KIND != FK_FIT) {
                colv[0][rr][cc] = col_value(i);
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
        };
        if constexpr (UNROLL && KIND != FK_FIT) {
            // Every cell of the thread computed before any is stored: the
            // stores go to the same LDS array as the reads, so a store
            // between two cells' reads would keep the later reads waiting.
            constexpr int kColIters = (kTileY * LW + kTileThreads - 1) / kTileThreads;
            float cv[kColIters];
#pragma unroll
            for (int k = 0; k < kColIters; ++k) {
                const int i = min(tid + k * kTileThreads, kTileY * LW - 1);
                cv[k] = col_value(i);
            }
            float *cw = lds + LH * LW;  // colv[0][i / LW][i % LW]
#pragma unroll
            for (int k = 0; k < kColIters; ++k) {
                const int i = tid + k * kTileThreads;
                if (i < kTileY * LW) cw[i] = cv[k];
            }
        } else if constexpr (UNROLL) {
            constexpr int kColIters = (kTileY * LW + kTileThreads - 1) / kTileThreads;
#pragma unroll
            for (int k = 0; k < kColIters; ++k) {
                const int i = tid + k * kTileThreads;
                if (i < kTileY * LW) col_cell(i);
            }
        } else {
            for (int i = tid; i < kTileY * LW; i += kTileThreads) col_cell(i);
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
                case STRATA_OP_SURFACE:  // the products follow from p and q at the store
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

    if constexpr (UNROLL && KIND != FK_FIT) {
        // A tile inside the region and the interior stores every value,
        // at 32-bit offsets from the wave's first row, with no per-cell
        // tests.
        const bool whole = tx0 + kTileX <= rg.ox + rg.w && ty0 + kTileY <= rg.oy + rg.h && tx0 >= R &&
                           tx0 + kTileX <= rg.W - R && ty0 >= R && ty0 + kTileY <= rg.H - R &&
                           (int64_t)kTileY * d0.stride < ((int64_t)1 << 29);
        if (whole) {
            char *ob = reinterpret_cast<char *>(d0.p + ((ty0 + rbase - rg.oy) * d0.stride + (tx0 - rg.ox)));
            const uint32_t st4 = (uint32_t)d0.stride * 4u;
            uint32_t off = 4u * (uint32_t)lane;
#pragma unroll
            for (int o = 0; o < kRowsPerWave; ++o) {
                *reinterpret_cast<STRATA_GLOBAL float *>(gp(ob) + off) = out0[o];
                off += st4;
            }
            return;
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
        if constexpr (KIND == FK_FIT) {
            if (P.op == STRATA_OP_SURFACE) {
                // Each wanted product from the fit's gradient (p, q), with
                // the k layout kp..ks, slope scale, aspect flat, c, bx, by.
                const float p = out0[o], q = out1[o], nan = edge_nan();
                auto put = [&](int i, float v) {
                    gp(dset[i].p)[(y - rg.oy) * dset[i].stride + (x - rg.ox)] = interior ? v : nan;
                };
                int i = 0;
                if (P.flags & STRATA_SURF_DX) put(i++, p);
                if (P.flags & STRATA_SURF_DY) put(i++, q);
                if (P.flags & STRATA_SURF_SLOPE) put(i++, slope_of(p, q, P.k[5], (P.flags & STRATA_FLAG_ATAN) != 0));
                if (P.flags & STRATA_SURF_ASPECT) put(i++, aspect_of(p, q, P.k[6], (P.flags & STRATA_FLAG_TRIG) != 0));
                if (P.flags & STRATA_SURF_HILLSHADE) put(i++, shade_of(p, q, P.k[7], P.k[8], P.k[9]));
                continue;
            }
        }
        gp(d0.p)[(y - rg.oy) * d0.stride + (x - rg.ox)] = interior ? out0[o] : edge_nan();
        if constexpr (KIND == FK_FIT) {
            if (P.op == STRATA_OP_GRADIENT)
                gp(d1.p)[(y - rg.oy) * d1.stride + (x - rg.ox)] = interior ? out1[o] : edge_nan();
        }
    }
}

template <int KIND, int R>
__global__ __launch_bounds__(kTileThreads) void focal_tile(const Params P, Region rg, SrcView s, DstSet d)
{
    __shared__ float lds[focal_lds_floats<KIND, R>()];
    FocalArgs a;
    a.op = P.op;
    a.flags = P.flags;
#pragma unroll
    for (int i = 0; i < 12; ++i) a.k[i] = P.k[i];
    const uint32_t gx = (uint32_t)((rg.w + kTileX - 1) / kTileX), t = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t tx0 = rg.ox + (int64_t)(t % gx) * kTileX;
    const int64_t ty0 = rg.oy + (int64_t)(t / gx) * kTileY;
    focal_block<KIND, R>(a, ArgTaps{P}, rg, s, d.v[0], d.v[1], tx0, ty0, lds, d.v);
}

// ---------------------------------------------------------------------------
// correlate_mfma_pipe: focal Correlate/Convolve as split-BF16x3 matrix products
// ---------------------------------------------------------------------------
//
// For a strip of 16 output columns at x0, weight row i contributes
// out[y][x0+n] += sum_k X_i[y][k] * T_i[k][n], with X_i[y][k] the input at
// row y+i-r, column x0-r+k, and T_i[k][n] = w[i][k-n] for 0 <= k-n <= 2r
// (else 0): a (rows x 32) by (32 x 16) product, since 16+2r <= 32. Each
// operand is split into two bf16 parts, x = hi + lo (hi = bf16(x), lo =
// bf16(x - hi)), and each 16x16 output block accumulates hi*hi + hi*lo +
// lo*hi over the 2r+1 weight rows with v_mfma_f32_16x16x32_bf16 in fp32.
// The dropped lo*lo term and the two bf16 roundings bound each product's
// error by about 3*2^-18 of |w*x|: this path is approximate, unlike the
// FP32 path, which follows the reference's operation order exactly.
//
// Each 64x64 output tile is computed by four waves: wave w takes columns
// 16w..16w+15 and four 16-row blocks. The input tile and halo are staged
// in LDS as the hi and lo planes (rows padded so each lane's 8 operand
// values are one aligned 16-byte read). Operand layouts of the 16x16x32 bf16 MFMA: lane l
// holds A[m = l%16][k = 8(l/16) .. +7], B[k = 8(l/16) .. +7][n = l%16], and
// C[m = 4(l/16) .. +3][n = l%16].
//
// Non-finite inputs cannot be split (Inf - Inf is NaN) and would leak into
// neighbouring outputs through the zeros of T_i, so they are staged as 0
// and every output whose window holds one is recomputed in the exact FP32
// order, which gives the FP32 path's value there.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaTile = 64;
constexpr uint32_t kMfmaStrip = 16;  // tiles per strip of correlate_mfma_pipe's order

// v = h + l with h = bf16(v) and l = bf16(v - h), for finite v. Finite
// values above the largest bf16 (0x7f7f) would round to infinity; their
// high part is truncated instead, which keeps v - h finite.
__device__ __forceinline__ void split_bf16(float v, __bf16 &h, __bf16 &l)
{
    __bf16 t = (__bf16)v;
    if (__builtin_isinf((float)t)) t = __builtin_bit_cast(__bf16, (uint16_t)(f2u(v) >> 16));
    h = t;
    l = (__bf16)(v - (float)t);
}

// bf16 per staged row: the A operands read columns 0..79 (a wave's 16 + 2R
// <= 32 columns at 16-column steps). At 80 (40 dwords) the sixteen rows one
// operand read touches fall in distinct LDS banks; 88 cost bank conflicts
// on every read (half the LDS's busy cycles).
template <int R>
constexpr int mfma_pitch()
{
    return 80;
}

// Exact FP32 correlate of one cell, in the reference's order; also reports
// whether the window holds a non-finite value.
template <int R>
__device__ __forceinline__ float correlate_cell(const float *taps, const SrcView &s, int64_t x, int64_t y,
                                                bool &nonfinite)
{
    constexpr int N = 2 * R + 1;
    float acc = 0.0f;
    nonfinite = false;
    // The rare path: kept rolled so it does not raise the kernel's registers.
#pragma unroll 1
    for (int j = 0; j < N; ++j)
#pragma unroll 1
        for (int c = 0; c < N; ++c) {
            const float v = load_cell(s, x - R + c, y - R + j);
            nonfinite = nonfinite || !__builtin_isfinite(v);
            const float term = taps[j * N + c] * v;
            acc = (j == 0 && c == 0) ? term : acc + term;
        }
    return acc;
}

// The Toeplitz operands of every weight row, split: entry (i, lane) holds
// the lane's B operand of T_i (k = 8(lane/16) .. +7, n = lane % 16) as its
// bf16 high and low parts. Built once per call by N waves; every workgroup
// of correlate_mfma_pipe then reads it (from L2) instead of building it.
template <int R>
__global__ __launch_bounds__(64) void mfma_toeplitz(const Params P, bf16x8 *tab)
{
    constexpr int N = 2 * R + 1;
    const int i = blockIdx.x, lane = threadIdx.x, n = lane & 15, kg = lane >> 4;
    bf16x8 bh, bl;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int j = 8 * kg + e - n;
        const float w = (j >= 0 && j < N) ? P.taps[i * N + j] : 0.0f;
        __bf16 h, l;
        split_bf16(w, h, l);
        bh[e] = h;
        bl[e] = l;
    }
    gp(tab)[(i * 64 + lane) * 2] = bh;
    gp(tab)[(i * 64 + lane) * 2 + 1] = bl;
}

// Two staged values as their bf16 high parts (hp) and low parts (lp),
// packed two to a word (the first in the low half), for finite values whose
// magnitude is below the largest bf16: the high part is the value rounded
// to bf16, the low part the rest rounded to bf16.
__device__ __forceinline__ void split_pair_fast(float a, float b, uint32_t &hp, uint32_t &lp)
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 h{(__bf16)a, (__bf16)b};
    hp = __builtin_bit_cast(uint32_t, h);
    const float ha = __builtin_bit_cast(float, hp << 16), hb = __builtin_bit_cast(float, hp & 0xffff0000u);
    const bf16x2 l{(__bf16)(a - ha), (__bf16)(b - hb)};
    lp = __builtin_bit_cast(uint32_t, l);
}
// The same for any two values: a non-finite value is staged as 0 (and
// reported), a finite one above the bf16 range through split_bf16.
__device__ __forceinline__ void split_pair_any(float a, float b, uint32_t &hp, uint32_t &lp, bool &any)
{
    if (!__builtin_isfinite(a)) {
        any = true;
        a = 0.0f;
    }
    if (!__builtin_isfinite(b)) {
        any = true;
        b = 0.0f;
    }
    __bf16 ha, la, hb, lb;
    split_bf16(a, ha, la);
    split_bf16(b, hb, lb);
    hp = (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
    lp = (uint32_t)__builtin_bit_cast(uint16_t, la) | ((uint32_t)__builtin_bit_cast(uint16_t, lb) << 16);
}
// Below the largest bf16 (0x7f7f), so rounding to bf16 stays finite.
constexpr float kBf16Safe = 3.38e38f;

// Stages one tile's window (the tile at tx0, ty0 and its halo) into the
// planes hi and lo, by 256 threads (p = 0..255). Returns whether the
// thread's values held a non-finite one.
template <int R>
__device__ __forceinline__ bool mfma_stage(const SrcView &s, int64_t tx0, int64_t ty0, int p, __bf16 *hi, __bf16 *lo)
{
    constexpr int LH = kMfmaTile + 2 * R, LW = kMfmaTile + 2 * R, PITCH = mfma_pitch<R>();
    // Staged columns: the A operands read columns 0..79 of a staged row;
    // those at LW and past it meet zeros of T_i and are staged as 0.
    constexpr int PPR = 40;  // column pairs per staged row
    constexpr int NP = LH * PPR;
    constexpr int KST = (NP + 255) / 256;
    static_assert(2 * PPR <= PITCH, "staged pairs fit the LDS row");
    // Staging in pairs of columns, every load of the thread issued before
    // any split (the trip count is a compile-time constant). A tile whose
    // staged window lies inside the buffer, with row offsets that fit 32
    // bits, takes unchecked loads at 32-bit byte offsets from one uniform
    // base. Pairs are split two values at a time; a pass whose values are
    // all finite and inside the bf16 range (a wave-uniform test) takes the
    // short split. Each pair's high and low parts go to LDS as one 32-bit
    // store each.
    const int64_t bx = tx0 - R, by = ty0 - R;
    const bool inside = bx >= s.lx0 && bx + LW <= s.lx1 && by >= s.ly0 && by + LH <= s.ly1 &&
                        (int64_t)(LH - 1) * s.stride + LW < ((int64_t)1 << 29);
    float va[KST], vb[KST];
    // The thread's first pair, then 256 pairs on: 6 rows and 16 pairs.
    int row = p / PPR, pc = p - row * PPR;
    if (inside) {
        const char *base = reinterpret_cast<const char *>(s.p + ((by - s.sy) * s.stride + (bx - s.sx)));
        const uint32_t st4 = (uint32_t)s.stride * 4u;
        // A pair is one 8-byte load when every pair starts 8-byte
        // aligned (LW is even, so a pair is wholly inside or outside).
        const bool pair8 = (reinterpret_cast<uintptr_t>(base) & 7u) == 0 && (st4 & 7u) == 0;
        uint32_t off = (uint32_t)row * st4 + 8u * (uint32_t)pc;
#pragma unroll
        for (int k = 0; k < KST; ++k) {
            const int col = 2 * pc;
            // Past the last row or column the load is clamped to the
            // window's first cell and its value replaced by 0.
            const bool ok0 = row < LH && col < LW, ok1 = row < LH && col + 1 < LW;
            float a, b;
            if (pair8) {
                typedef float f32x2 __attribute__((ext_vector_type(2)));
                const f32x2 ab = *reinterpret_cast<const STRATA_GLOBAL f32x2 *>(gp(base) + (ok0 ? off : 0u));
                a = ab.x;
                b = ab.y;
            } else {
                a = *reinterpret_cast<const STRATA_GLOBAL float *>(gp(base) + (ok0 ? off : 0u));
                b = *reinterpret_cast<const STRATA_GLOBAL float *>(gp(base) + (ok1 ? off + 4u : 0u));
            }
            va[k] = ok0 ? a : 0.0f;
            vb[k] = ok1 ? b : 0.0f;
            pc += 16;
            row += 6;
            off += 6u * st4 + 128u;
            if (pc >= PPR) {
                pc -= PPR;
                row += 1;
                off += st4 - 8u * PPR;
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < KST; ++k) {
            const int col = 2 * pc;
            const bool ok = row < LH;
            va[k] = ok && col < LW ? load_cell(s, bx + col, by + row) : 0.0f;
            vb[k] = ok && col + 1 < LW ? load_cell(s, bx + col + 1, by + row) : 0.0f;
            pc += 16;
            row += 6;
            if (pc >= PPR) {
                pc -= PPR;
                row += 1;
            }
        }
    }
    bool any = false;
    uint32_t *hi32 = reinterpret_cast<uint32_t *>(hi), *lo32 = reinterpret_cast<uint32_t *>(lo);
    row = p / PPR;
    pc = p - row * PPR;
#pragma unroll
    for (int k = 0; k < KST; ++k) {
        const bool in = row < LH;
        const bool plain = __builtin_fabsf(va[k]) < kBf16Safe && __builtin_fabsf(vb[k]) < kBf16Safe;
        uint32_t hp, lp;
        if (__builtin_amdgcn_ballot_w64(in && !plain) == 0)
            split_pair_fast(va[k], vb[k], hp, lp);
        else
            split_pair_any(va[k], vb[k], hp, lp, any);
        if (in) {
            hi32[row * (PITCH / 2) + pc] = hp;
            lo32[row * (PITCH / 2) + pc] = lp;
        }
        pc += 16;
        row += 6;
        if (pc >= PPR) {
            pc -= PPR;
            row += 1;
        }
    }

    return any;
}

// One wave's products of a staged tile: columns 16*wave.. of the tile, its
// four 16-row blocks into acc.
template <int R>
__device__ __forceinline__ void mfma_products(const __bf16 *hi, const __bf16 *lo, const bf16x8 *tab, int lane,
                                              int wave, f32x4 (&acc)[4])
{
    constexpr int N = 2 * R + 1, PITCH = mfma_pitch<R>();
    const int n = lane & 15, kg = lane >> 4;
    const int c0 = wave * 16;
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // Each block accumulates hi*hi, hi*lo, lo*hi per weight row, in that
    // order; the four blocks' products interleave.
    const STRATA_GLOBAL bf16x8 *tp = gp(tab) + lane * 2;
    const __bf16 *ahp = &hi[(n)*PITCH + c0 + 8 * kg];
    const __bf16 *alp = &lo[(n)*PITCH + c0 + 8 * kg];
    auto row = [&](int i, const bf16x8 &bh, const bf16x8 &bl) __attribute__((always_inline)) {
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            ah[b] = *reinterpret_cast<const bf16x8 *>(ahp + (16 * b + i) * PITCH);  // A: rows 16b+i.., columns 8kg..
            al[b] = *reinterpret_cast<const bf16x8 *>(alp + (16 * b + i) * PITCH);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[b], bh, acc[b], 0, 0, 0);
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[b], bl, acc[b], 0, 0, 0);
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[b], bh, acc[b], 0, 0, 0);
    };
    // Weight rows in pairs with two operand sets, so each row's B operands
    // load a row ahead of their products with no copies between sets; the
    // scheduling barriers keep each load ahead of the products it must
    // overlap. Row 0's operands pass through an empty asm: with a load on
    // both sides of the loop's phi, the compiler would turn the phi into
    // one load at the top of each iteration, waited for right away.
    bf16x8 b0h = tp[0], b0l = tp[1];
    __asm__ volatile("" : "+v"(b0h), "+v"(b0l));
    int i = 0;
#pragma unroll 1
    for (; i + 1 < N; i += 2) {
        const bf16x8 b1h = tp[(i + 1) * 128], b1l = tp[(i + 1) * 128 + 1];
        __builtin_amdgcn_sched_barrier(0);
        row(i, b0h, b0l);
        const int i2 = i + 2 < N ? i + 2 : N - 1;
        b0h = tp[i2 * 128];
        b0l = tp[i2 * 128 + 1];
        __builtin_amdgcn_sched_barrier(0);
        row(i + 1, b1h, b1l);
    }
    if (i < N) row(i, b0h, b0l);  // the last row of an odd count
}

// A whole tile: inside the region and the raster's interior, with nothing
// to recompute and row offsets within 32 bits.
template <int R>
__device__ __forceinline__ bool mfma_whole(const Region &rg, const DstView &d0, int64_t tx0, int64_t ty0, bool fix)
{
    return !fix && tx0 + kMfmaTile <= rg.ox + rg.w && ty0 + kMfmaTile <= rg.oy + rg.h && tx0 >= R &&
           tx0 + kMfmaTile <= rg.W - R && ty0 >= R && ty0 + kMfmaTile <= rg.H - R &&
           (int64_t)kMfmaTile * d0.stride < ((int64_t)1 << 29);
}

// Stores one wave's products from its registers: a whole tile without
// conditions at 32-bit offsets; any other tile cell by cell, with the edge
// ring and, where the window held a non-finite value (fix), the exact FP32
// value.
template <int R>
__device__ __forceinline__ void mfma_store(const float *wts, const Region &rg, const SrcView &s, const DstView &d0,
                                           int64_t tx0, int64_t ty0, int lane, int wave, bool fix,
                                           const f32x4 (&acc)[4])
{
    const int n = lane & 15, kg = lane >> 4;
    const int c0 = wave * 16;
    const int64_t x = tx0 + c0 + n;
    if (mfma_whole<R>(rg, d0, tx0, ty0, fix)) {
        char *ob = reinterpret_cast<char *>(d0.p + ((ty0 - rg.oy) * d0.stride + (tx0 - rg.ox)));
        const uint32_t st4 = (uint32_t)d0.stride * 4u;
        const uint32_t o0 = (uint32_t)(4 * kg) * st4 + 4u * (uint32_t)(c0 + n);
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                *reinterpret_cast<STRATA_GLOBAL float *>(gp(ob) + o0 + (uint32_t)(16 * b + e) * st4) = acc[b][e];
        return;
    }
    const bool col_in = x < rg.ox + rg.w;
    const bool col_interior = x >= R && x < rg.W - R;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t y = ty0 + 16 * b + 4 * kg + e;
            if (!col_in || y >= rg.oy + rg.h) continue;
            const bool interior = col_interior && y >= R && y < rg.H - R;
            float v = acc[b][e];
            if (fix && interior) {
                bool nf;
                const float exact = correlate_cell<R>(wts, s, x, y, nf);
                if (nf) v = exact;
            }
            gp(d0.p)[(y - rg.oy) * d0.stride + (x - rg.ox)] = interior ? v : edge_nan();
        }
}

// correlate_mfma_pipe: a workgroup of eight waves stays resident and loops
// over tiles, so the staging runs beside the products: waves
// 4..7 stage the next tile into one of two LDS buffers while waves 0..3 run
// the current tile's products from the other and store them, and one
// barrier per tile hands the buffers over. Every thread of the workgroup
// runs the same number of iterations (the count depends on the workgroup
// only), so every barrier is reached by all of them. The grid is a
// multiple of the XCD count; workgroup b serves XCD b % 8, which owns one
// band of tile rows, and the band's workgroups take its tiles in steps.
// Six waves per SIMD (three workgroups a CU, which their LDS allows at
// every radius) bound the registers.
template <int R>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(6))) void correlate_mfma_pipe(
    const Params P, Region rg, SrcView s, DstView d0, const bf16x8 *__restrict__ tab)
{
    constexpr int N = 2 * R + 1;
    constexpr int LH = kMfmaTile + 2 * R, PITCH = mfma_pitch<R>(), PLANES = 2 * LH * PITCH;
    __shared__ __attribute__((aligned(16))) __bf16 planes[2 * PLANES];
    __shared__ float wts[N * N];
    // Per buffer: the iteration whose tile, staged there, held a
    // non-finite value (-1: none yet).
    __shared__ int nonfinite[2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t gx = (uint32_t)((rg.w + kMfmaTile - 1) / kMfmaTile);
    const uint32_t gy = (uint32_t)((rg.h + kMfmaTile - 1) / kMfmaTile);
    const uint32_t S = gridDim.x / kXcds, xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    // The XCD's band of tile rows, taken in strips kMfmaStrip tiles wide,
    // each strip row by row: the S tiles in flight on the XCD form a block
    // about kMfmaStrip wide, so a tile's neighbours on every side, whose
    // windows share its halo, are staged close to it in time and the
    // shared lines come from the XCD's L2.
    const uint32_t per = gy / kXcds, extra = gy % kXcds;
    const uint32_t ry0 = xcd * per + min(xcd, extra), nrows = per + (xcd < extra ? 1u : 0u);
    const uint32_t cnt = nrows * gx;
    // This workgroup's tiles: XCD-local numbers slot + k*S for k < K.
    const int K = slot < cnt ? (int)((cnt - slot + S - 1) / S) : 0;
    if (K == 0) return;  // the whole workgroup, before any barrier
    auto origin = [&](int k, int64_t &tx0, int64_t &ty0) {
        const uint32_t j = slot + (uint32_t)k * S;
        const uint32_t strip = j / (kMfmaStrip * nrows), jj = j - strip * kMfmaStrip * nrows;
        const uint32_t ws = min(kMfmaStrip, gx - strip * kMfmaStrip);
        tx0 = rg.ox + (int64_t)(strip * kMfmaStrip + jj % ws) * kMfmaTile;
        ty0 = rg.oy + (int64_t)(ry0 + jj / ws) * kMfmaTile;
    };
    if (tid < 2) nonfinite[tid] = -1;
    for (int i = tid; i < N * N; i += 512) wts[i] = P.taps[i];
    __syncthreads();
    const bool producer = wave >= 4;
    if (producer) {
        int64_t tx0, ty0;
        origin(0, tx0, ty0);
        if (mfma_stage<R>(s, tx0, ty0, tid - 256, planes, planes + LH * PITCH)) nonfinite[0] = 0;
    }
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        if (producer) {
            if (k + 1 < K) {
                int64_t tx0, ty0;
                origin(k + 1, tx0, ty0);
                __bf16 *pl = planes + ((k + 1) & 1) * PLANES;
                // The thread's staging offsets are derived anew each tile
                // from an opaque copy of its index: hoisted out of the
                // loop, they would stay live through it and spill.
                int p = tid - 256;
                __asm__ volatile("" : "+v"(p));
                if (mfma_stage<R>(s, tx0, ty0, p, pl, pl + LH * PITCH)) nonfinite[(k + 1) & 1] = k + 1;
            }
        } else {
            int64_t tx0, ty0;
            origin(k, tx0, ty0);
            const __bf16 *pl = planes + (k & 1) * PLANES;
            int ln = lane;  // opaque, as the producers' index
            __asm__ volatile("" : "+v"(ln));
            f32x4 acc[4];
            mfma_products<R>(pl, pl + LH * PITCH, tab, ln, wave, acc);
            mfma_store<R>(wts, rg, s, d0, tx0, ty0, ln, wave, nonfinite[k & 1] == k, acc);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// correlate_error_f64: a Correlate output against the FP64 correlation
// ---------------------------------------------------------------------------
//
// Measurement only: for every interior cell of the region whose window is
// finite, the FP64 correlation (products and sum in double) and the error
// of out there: max |out - ref|, max |out - ref| / |ref| (ref != 0) and max
// |out - ref| / sum |w x|; non-negative doubles order like their bit
// patterns, so the maxima are unsigned 64-bit atomics of one value per
// workgroup. acc: {max abs, max rel, max scaled, cells counted}.
template <int R>
__global__ __launch_bounds__(256) void correlate_error_f64(const Params P, Region rg, SrcView s, const float *out,
                                                           int64_t out_stride, unsigned long long *acc)
{
    constexpr int N = 2 * R + 1;
    __shared__ double red[3][256];
    __shared__ unsigned long long cnt[256];
    const int64_t x = rg.ox + (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t y = rg.oy + blockIdx.y;
    double ea = 0.0, er = 0.0, es = 0.0;
    unsigned long long c = 0;
    if (x < rg.ox + rg.w && x >= R && x < rg.W - R && y >= R && y < rg.H - R) {
        double sum = 0.0, mag = 0.0;
        bool finite = true;
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                const double v = (double)load_cell(s, x - R + k, y - R + j);
                finite = finite && __builtin_isfinite(v);
                const double t = (double)P.taps[j * N + k] * v;
                sum += t;
                mag += __builtin_fabs(t);
            }
        const double o = (double)gp(out)[(y - rg.oy) * out_stride + (x - rg.ox)];
        if (finite) {
            const double d = __builtin_fabs(o - sum);
            ea = d;
            er = sum != 0.0 ? d / __builtin_fabs(sum) : (d == 0.0 ? 0.0 : d);
            es = mag > 0.0 ? d / mag : d;
            c = 1;
        }
    }
    red[0][threadIdx.x] = ea;
    red[1][threadIdx.x] = er;
    red[2][threadIdx.x] = es;
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = max(red[q][threadIdx.x], red[q][threadIdx.x + st]);
            cnt[threadIdx.x] += cnt[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int q = 0; q < 3; ++q) atomicMax(acc + q, (unsigned long long)__builtin_bit_cast(uint64_t, red[q][0]));
        atomicAdd(acc + 3, cnt[0]);
    }
}


// ---------------------------------------------------------------------------
// point_rows: radius-0 transfer kernels
// ---------------------------------------------------------------------------

constexpr int kPointRows = 16;

// The table is read at uniform indices from the kernel arguments and
// gathered from the block's LDS copy.
template <int OP>
__global__ __launch_bounds__(256) void point_rows(const Params P, Region rg, SrcView s, DstView d0)
{
    __shared__ __align__(16) float lt[kMaxTaps + 2];
    lds_table((STRATA_LDS float *)lt, OP, P.taps, P.ntaps, (int)threadIdx.x, 256);
    __syncthreads();
    const int64_t x = rg.ox + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (x >= rg.ox + rg.w) return;
    const int64_t y0 = rg.oy + (int64_t)blockIdx.y * kPointRows;
    const int64_t y1 = min(y0 + (int64_t)kPointRows, rg.oy + rg.h);
    const STRATA_LDS float *tl = (const STRATA_LDS float *)lt;
    if (y1 - y0 == kPointRows) {
        // The column's cells all loaded, then one pass over the table.
        float v[kPointRows];
#pragma unroll
        for (int i = 0; i < kPointRows; ++i) v[i] = gp(s.p)[(y0 + i - s.sy) * s.stride + (x - s.sx)];
        point_values(v, OP, P.taps, P.ntaps, tl);
#pragma unroll
        for (int i = 0; i < kPointRows; ++i) gp(d0.p)[(y0 + i - rg.oy) * d0.stride + (x - rg.ox)] = v[i];
        return;
    }
    for (int64_t y = y0; y < y1; ++y) {
        float v[1] = {gp(s.p)[(y - s.sy) * s.stride + (x - s.sx)]};
        point_values(v, OP, P.taps, P.ntaps, tl);
        gp(d0.p)[(y - rg.oy) * d0.stride + (x - rg.ox)] = v[0];
    }
}

// ---------------------------------------------------------------------------
// validity: output bits, one 32-bit mask word per lane
// ---------------------------------------------------------------------------
//
// Masks are LSB-first, so a []uint64 bitmap read as 32-bit words has bit b in
// word b>>5 at position b&31, and every funnel shift of two 32-bit words is
// one v_alignbit.
//
// A wave takes kVsRows rows of a region and 64 consecutive 32-bit "cell
// words" of them: lane L holds cells [xc + 32L, xc + 32L + 32) of each row.
// It first loads, one coalesced word per lane, every source row the strip
// needs (kVsRows + 2R rows, all in flight at once) and aligns each to the
// cell grid with the next lane's word (DPP) and one v_alignbit. The AND of
// 2R+1 rows, eroded horizontally by R cells with the neighbouring lanes'
// words, is a row's validity in cell space; the output word, whose alignment
// to the cell grid varies from row to row, is one more funnel shift of the
// lane's and its left neighbour's cell words. Lanes 2..61 store output words;
// lanes 0, 1, 62 and 63 only supply neighbours (lane 1's erosion needs lane
// 0's cells, lane 61's needs lane 62's, whose alignment needs lane 63's word).
// Words only partly owned by the region's row are updated with atomics on
// the row's own bits; all reads stay inside the source buffer's rows.
constexpr int kVsOut = 60;   // output words per wave and row

__device__ __forceinline__ uint32_t lane_up(uint32_t v)  // v of lane+1 (lane 63: 0)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t lane_down(uint32_t v)  // v of lane-1 (lane 0: 0)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}

// Bits [lo, hi) of a 32-bit word whose first bit is at 0, lo and hi clamped.
__device__ __forceinline__ uint32_t bit_range(int32_t lo, int32_t hi)
{
    lo = max(lo, 0);
    hi = min(hi, 32);
    if (lo >= hi) return 0u;
    const int n = hi - lo;
    return (n == 32 ? ~0u : ((1u << n) - 1u)) << lo;
}

__device__ __forceinline__ int32_t clamp32(int64_t v, int32_t lo, int32_t hi)
{
    return (int32_t)min(max(v, (int64_t)lo), (int64_t)hi);
}

// Validity of rows [y0, y1) (y1 - y0 <= kVsRows) and output words
// [wx*kVsOut, wx*kVsOut + kVsOut) of each of the region's rows. wx, y0 and
// y1 are wave-uniform: every row and word offset below is scalar, advanced
// by one stride per row, and lanes only add their own 32-bit word index.
// cv, cown carry the bits of a row's last word into the next row's first
// (see chain below), also into the wave's next call when it continues with
// the rows from y1 up to yend.
// Waves whose words lie inside the source rows and the output rows (all but
// the first and last wave of a row, and strips at the buffer's top and
// bottom) take unchecked loads and whole-word stores.
template <int R>
__device__ __forceinline__ void validity_strip(const Region &rg, const MaskView &src, int has_src,
                                               const MaskView &dst, int64_t wx, int64_t y0, int64_t y1,
                                               int64_t yend, uint32_t &cv, uint32_t &cown)
{
    constexpr int NL = kVsRows + 2 * R;  // source rows y0-R .. y0+kVsRows-1+R
    const int lane = (int)(threadIdx.x & 63);
    const int64_t xc = rg.ox + wx * (32 * kVsOut) - 64;  // lane 2's cells start at the wave's first word
    const uint32_t *sbits = reinterpret_cast<const uint32_t *>(src.bits);
    uint32_t *dbits = reinterpret_cast<uint32_t *>(dst.bits);

    // Interior cells [R, W-R) of the lane's cell word; the same in every row.
    uint32_t imask = ~0u;
    if (xc < R || xc + 64 * 32 > rg.W - R) {
        const int32_t u0 = clamp32(R - xc, -4096, 4096), u1 = clamp32(rg.W - R - xc, -4096, 4096);
        imask = bit_range(u0 - 32 * lane, u1 - 32 * lane);
    }

    uint32_t A[NL];
    if (has_src) {
        uint32_t Wd[NL];
        uint32_t sh[NL];
        int64_t p = src.off + (y0 - R - src.y0) * src.stride + (xc - src.x0);  // bit of cell (xc, y0-R)
        if (xc >= src.lx0 && xc + 64 * 32 <= src.lx1 && y0 - R >= src.ly0 && y0 + kVsRows + R <= src.ly1) {
#pragma unroll
            for (int j = 0; j < NL; ++j) {
                const STRATA_GLOBAL uint32_t *rowp = gp(sbits) + (p >> 5);
                Wd[j] = rowp[lane];
                sh[j] = (uint32_t)(p & 31);
                p += src.stride;
            }
        } else {
            // Rows the buffer does not hold read as zeros (cells outside the
            // raster); lanes outside the row's own words read the nearest
            // one and discard it, rows outside the buffer a word of its first
            // row, so every load stays inside the buffer.
            const int64_t safe = (src.off + (src.ly0 - src.y0) * src.stride + (src.lx0 - src.x0)) >> 5;
            const int64_t dlo = src.lx0 - xc, dhi = src.lx1 - 1 - xc;  // row's cells relative to xc
#pragma unroll
            for (int j = 0; j < NL; ++j) {
                const int64_t yy = y0 - R + j;
                const int64_t base = p >> 5;  // word of lane 0
                const int64_t lo = ((p + dlo) >> 5) - base, hi = ((p + dhi) >> 5) - base;
                const bool ok = yy >= src.ly0 && yy < src.ly1 && lo <= 63 && hi >= 0;
                const STRATA_GLOBAL uint32_t *rowp = gp(sbits) + (ok ? base : safe);
                const int32_t lo32 = ok ? clamp32(lo, 0, 63) : 0, hi32 = ok ? clamp32(hi, 0, 63) : 0;
                const int32_t lc = min(max(lane, lo32), hi32);
                const uint32_t w = rowp[lc];
                Wd[j] = (ok && lc == lane) ? w : 0u;
                sh[j] = (uint32_t)(p & 31);
                p += src.stride;
            }
        }
#pragma unroll
        for (int j = 0; j < NL; ++j) A[j] = __builtin_amdgcn_alignbit(lane_up(Wd[j]), Wd[j], sh[j]);
    }

    // Output rows: lane l stores word wb + l of the row (lanes 2..61, up to
    // the row's last word), whose cells start s = ob & 31 cells before the
    // lane's own cell word. A wave past the first and before the last
    // kVsOut + 2 words of every row never meets a partly owned word.
    int64_t ob = dst.off + (y0 - dst.y0) * dst.stride + (rg.ox - dst.x0);  // bit of cell (ox, y0)
    const bool whole = wx > 0 && ((rg.w - 1) >> 5) >= wx * kVsOut + kVsOut;
    const bool store_lane = lane >= 2 && lane < 2 + kVsOut;
    // Rows that follow each other in the output's bits (stride = width), all
    // of each row in this wave: a row's last word is the next row's first,
    // so its bits are carried (cv, cown) into that word, stored whole once.
    const bool chain = dst.stride == rg.w && wx == 0 && ((rg.w + 30) >> 5) + 2 < 2 + kVsOut;
#pragma unroll
    for (int i = 0; i < kVsRows; ++i) {
        const int64_t y = y0 + i;
        if (y >= y1) break;
        const bool row_in = y >= R && y < rg.H - R;
        uint32_t v;
        if (has_src) {
            uint32_t a = A[i];
#pragma unroll
            for (int j = 1; j <= 2 * R; ++j) a &= A[i + j];
            const uint32_t lo = lane_down(a), hi = lane_up(a);
            uint32_t e = a;
#pragma unroll
            for (int c = 1; c <= R; ++c) {
                e &= __builtin_amdgcn_alignbit(hi, a, (uint32_t)c);        // cell + c
                e &= __builtin_amdgcn_alignbit(a, lo, (uint32_t)(32 - c));  // cell - c
            }
            v = row_in ? e : 0u;
        } else {
            v = row_in ? ~0u : 0u;
        }
        v &= imask;
        const uint32_t s = (uint32_t)(ob & 31);
        const int64_t wb = (ob >> 5) + wx * kVsOut - 2;
        STRATA_GLOBAL uint32_t *rowd = gp(dbits) + wb;
        uint32_t *rowd_flat = dbits + wb;  // for the atomics
        const uint32_t prev = lane_down(v);
        const uint32_t out = s == 0 ? v : __builtin_amdgcn_alignbit(v, prev, 32u - s);
        if (whole) {
            if (store_lane) rowd[lane] = out;
        } else {
            const int64_t oe = ob + rg.w;  // bit after the row
            const int32_t last = clamp32(((oe - 1) >> 5) - wb, -1, 64);  // lane of the row's last word
            const uint32_t e = (uint32_t)(oe & 31);
            uint32_t own = ~0u, val = out;
            if (wx == 0 && lane == 2) own &= ~0u << s;             // the row's first word
            if (e != 0 && lane == last) own &= (1u << e) - 1u;     // the row's last word
            if (lane == 2 && cown != 0u) {                         // the previous row's last bits
                val = (out & own) | cv;
                own |= cown;
            }
            // The row's last word waits for the next row's first bits.
            const bool carry = chain && e != 0 && y + 1 < yend;
            if (store_lane && lane <= last && !(carry && lane == last)) {
                if (own == ~0u) {
                    rowd[lane] = val;
                } else {
                    atomicAnd(rowd_flat + lane, ~own | (val & own));
                    atomicOr(rowd_flat + lane, val & own);
                }
            }
            cv = carry ? (uint32_t)__builtin_amdgcn_readlane((int)(val & own), last) : 0u;
            cown = carry ? (uint32_t)__builtin_amdgcn_readlane((int)own, last) : 0u;
        }
        ob += dst.stride;
    }
}

// One wave per (word group, strip): wave t of the launch takes word group
// t % waves_x of strip t / waves_x.
template <int R>
__global__ __launch_bounds__(256) void validity_kernel(Region rg, MaskView src, int has_src, MaskView dst,
                                                       uint32_t waves_x)
{
    const uint32_t wv = blockIdx.x * 4 + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t wx = wv % waves_x, strip = wv / waves_x;
    const int64_t y0 = rg.oy + (int64_t)strip * kVsRows;
    const int64_t y1 = min(y0 + (int64_t)kVsRows, rg.oy + rg.h);
    if (y0 >= y1) return;  // past the last strip
    uint32_t cv = 0u, cown = 0u;
    validity_strip<R>(rg, src, has_src, dst, wx, y0, y1, y1, cv, cown);
}

// A radius-1 Horn op with one masked output, data and validity in one
// launch: blocks [0, data_blocks) sweep as horn_sweep does, the blocks after
// them run validity_strip waves (four per block), so the validity's short
// work fills the sweep's tail instead of a launch of its own.
template <int OP, bool UNALIGNED>
__global__ __launch_bounds__(kSweepThreads) void horn_sweep_valid(HornArgs a, Region rg, SrcView s, DstView d0,
                                                                  uint32_t data_blocks, MaskView vsrc, int has_src,
                                                                  MaskView vdst, uint32_t waves_x)
{
    const uint32_t b = blockIdx.x;
    if (b >= data_blocks) {
        const uint32_t wv = (b - data_blocks) * 4 + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint32_t wx = wv % waves_x, strip = wv / waves_x;
        const int64_t y0 = rg.oy + (int64_t)strip * kVsRows;
        const int64_t y1 = min(y0 + (int64_t)kVsRows, rg.oy + rg.h);
        if (y0 >= y1) return;  // past the last strip
        uint32_t cv = 0u, cown = 0u;
        validity_strip<1>(rg, vsrc, has_src, vdst, wx, y0, y1, y1, cv, cown);
        return;
    }
    const uint32_t t = xcd_tile(b, data_blocks);
    const int64_t xb = rg.ox + (int64_t)(t % a.blocks_x) * kSweepCols;
    const int64_t y0 = rg.oy + (int64_t)(t / a.blocks_x) * a.rows;
    const int64_t y1 = min(y0 + (int64_t)a.rows, rg.oy + rg.h);
    const bool inside = xb >= s.lx0 && xb + kSweepCols <= s.lx1 && xb >= rg.ox && xb + kSweepCols <= rg.ox + rg.w;
    if constexpr (UNALIGNED) {
        if (inside)
            sweep<OP, kSweepAnyParity>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x);
        else
            sweep<OP, kSweepChecked>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x);
    } else {
        if (inside && a.aligned)
            sweep<OP, kSweepPairs>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x);
        else
            sweep<OP, kSweepChecked>(a, rg, s, d0, d0, xb, y0, y1, threadIdx.x);
    }
}

// validity_strip's rule for a narrow region, whose row (its output words
// and four helper lanes) fits in G <= 32 lanes: the wave's lanes form
// 64 / G groups of G lanes, group g running through the region's rows
// [y0 + run g, y0 + run (g + 1)) below y1 in chunks of kVgRows, this call
// taking chunk c, its lane l standing for
// validity_strip's lane l of the region's first word group (cell word
// xc + 32 l with xc = rg.ox - 64). Lanes past the last group only compute.
// Row and word offsets differ between groups, so they are per lane; every
// load is clamped into the source rows as on validity_strip's checked path,
// and a row's first and last output words update the row's own bits with
// atomics (a narrow output's rows share words).
template <int R>
__device__ __forceinline__ void validity_groups(const Region &rg, const MaskView &src, int has_src,
                                                const MaskView &dst, int64_t y0, int64_t y1, int G, int64_t run,
                                                int c, uint32_t &cv, uint32_t &cown)
{
    constexpr int NL = kVgRows + 2 * R;
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / G, l = lane - g * G;
    const int64_t ys = y0 + (int64_t)g * run + (int64_t)c * kVgRows;  // the chunk's first row
    const int64_t rend = min(y0 + (int64_t)(g + 1) * run, y1);       // the group's run ends here
    const int64_t ye = g < 64 / G ? max(ys, min(ys + kVgRows, rend)) : ys;
    const int64_t xc = rg.ox - 64;
    const uint32_t *sbits = reinterpret_cast<const uint32_t *>(src.bits);
    uint32_t *dbits = reinterpret_cast<uint32_t *>(dst.bits);

    const int32_t u0 = clamp32(R - xc, -4096, 4096), u1 = clamp32(rg.W - R - xc, -4096, 4096);
    const uint32_t imask = bit_range(u0 - 32 * l, u1 - 32 * l);

    uint32_t A[NL];
    if (has_src) {
        const int64_t p0 = src.off + (ys - R - src.y0) * src.stride + (xc - src.x0);  // bit of cell (xc, ys - R)
        const int64_t safe = (src.off + (src.ly0 - src.y0) * src.stride + (src.lx0 - src.x0)) >> 5;
        const int64_t dlo = src.lx0 - xc, dhi = src.lx1 - 1 - xc;
        int64_t p = p0;
#pragma unroll
        for (int j = 0; j < NL; ++j) {
            const int64_t yy = ys - R + j;
            const int64_t base = p >> 5;  // the word of the group's lane 0
            const int64_t lo = ((p + dlo) >> 5) - base, hi = ((p + dhi) >> 5) - base;
            const bool ok = yy >= src.ly0 && yy < src.ly1 && lo <= G - 1 && hi >= 0;
            const int32_t lo32 = ok ? clamp32(lo, 0, G - 1) : 0, hi32 = ok ? clamp32(hi, 0, G - 1) : 0;
            const int32_t lc = min(max(l, lo32), hi32);
            const STRATA_GLOBAL uint32_t *rowp = gp(sbits) + (ok ? base + lc : safe);
            const uint32_t w = *rowp;
            A[j] = (ok && lc == l) ? w : 0u;
            p += src.stride;
        }
        // Each row's word aligned to the cell grid (its shift recomputed
        // from the first row's bit rather than kept per row).
        const uint32_t p32 = (uint32_t)p0, st32 = (uint32_t)src.stride;
#pragma unroll
        for (int j = 0; j < NL; ++j)
            A[j] = __builtin_amdgcn_alignbit(lane_up(A[j]), A[j], (p32 + (uint32_t)j * st32) & 31u);
    }

    int64_t ob = dst.off + (ys - dst.y0) * dst.stride + (rg.ox - dst.x0);  // bit of cell (ox, ys)
    // Rows that follow each other in the output's bits (stride = width): a
    // row's last word is the next row's first, so the group carries its
    // bits (cv, cown, at the group's lane 2) into that word, stored whole
    // once, through the group's run (into its next chunk's call).
    const bool chain = dst.stride == rg.w;
#pragma unroll
    for (int i = 0; i < kVgRows; ++i) {
        const int64_t y = ys + i;
        const bool row_in = y >= R && y < rg.H - R;
        uint32_t v;
        if (has_src) {
            uint32_t a = A[i];
#pragma unroll
            for (int j = 1; j <= 2 * R; ++j) a &= A[i + j];
            const uint32_t lo = lane_down(a), hi = lane_up(a);
            uint32_t e = a;
#pragma unroll
            for (int c = 1; c <= R; ++c) {
                e &= __builtin_amdgcn_alignbit(hi, a, (uint32_t)c);        // cell + c
                e &= __builtin_amdgcn_alignbit(a, lo, (uint32_t)(32 - c));  // cell - c
            }
            v = row_in ? e : 0u;
        } else {
            v = row_in ? ~0u : 0u;
        }
        v &= imask;
        const uint32_t s = (uint32_t)(ob & 31);
        const int64_t wb = (ob >> 5) - 2;  // the output word of the group's lane 0
        const uint32_t prev = lane_down(v);
        const uint32_t out = s == 0 ? v : __builtin_amdgcn_alignbit(v, prev, 32u - s);
        const int64_t oe = ob + rg.w;  // bit after the row
        const int32_t last = clamp32(((oe - 1) >> 5) - wb, -1, 64);  // lane of the row's last word
        const uint32_t e = (uint32_t)(oe & 31);
        uint32_t own = ~0u, val = out;
        if (l == 2) own &= ~0u << s;                     // the row's first word
        if (e != 0 && l == last) own &= (1u << e) - 1u;  // the row's last word
        if (l == 2 && cown != 0u) {                      // the previous row's last bits
            val = (out & own) | cv;
            own |= cown;
        }
        // The row's last word waits for the next row's first bits.
        const bool carry = chain && e != 0 && y + 1 < rend;
        if (y < ye && l >= 2 && l <= last && !(carry && l == last)) {
            uint32_t *word = dbits + wb + l;
            if (own == ~0u) {
                *gp(word) = val;
            } else {
                atomicAnd(word, ~own | (val & own));
                atomicOr(word, val & own);
            }
        }
        // Lane 2 of each group takes its group's last word (group lanes stay
        // inside the wave: g * G + last < 64 whenever a row is carried).
        const int from = 4 * min(max(lane - l + last, 0), 63);
        const uint32_t tv = (uint32_t)__builtin_amdgcn_ds_bpermute(from, (int)(val & own));
        const uint32_t to = (uint32_t)__builtin_amdgcn_ds_bpermute(from, (int)own);
        cv = carry ? tv : 0u;
        cown = carry ? to : 0u;
        ob += dst.stride;
    }
}

// ---------------------------------------------------------------------------
// batch_kernel: many tiles and mixed ops, persistent launches over work lists
// ---------------------------------------------------------------------------
//
// A batch's work is split into three lists: radius-1 sweep units, focal
// tile units and validity units. Each list is run by its own persistent
// kernel (batch_kernel<FAMILY>), so each compiles to the registers its own
// work needs and keeps its occupancy; the three run side by side. Within a
// list, workgroups claim units from a counter in order until the list is
// empty, so uneven tiles balance across the chip. The next claim is issued
// while the current unit runs, and the unit's task is found by binary
// search over the tasks' first units unless it is the previous unit's task.
// Unit indices are read back through readfirstlane, so the search and the
// task's fields are scalar.

// The point ops of a batch op, applied to a stencil op's output values. The
// op table is read-only for the launch and read as constant memory, so the
// sweep's stores do not make the compiler load it again after each one.
// A batch op's point stages on a sweep's output values (the counting forms
// of arith.hpp): the x sides read at uniform indices from the op in device
// memory, the entries picked gathered from the wave's own LDS copy of the
// op's tables (lds_table's layout, stage after stage).
constexpr int kWavePostFloats = 2 * kBatchPostTaps;
struct ChainPost {
    const STRATA_CONST BatchOp *op;
    const STRATA_LDS float *lt;
    __device__ void operator()(float &v0, float &v1) const
    {
        float v[2] = {v0, v1};
        int lo = 0;
        for (int i = 0; i < op->npost; ++i) {
            const int o = op->post_op[i], c = op->post_n[i];
            point_values(v, o, op->post_taps[i], c, lt + lo);
            lo += lds_table_floats(o, c);
        }
        v0 = v[0];
        v1 = v[1];
    }
};

// Copies op's point tables into the wave's LDS copy; every lane of the wave
// calls it. LDS operations of one wave complete in order, so the wave's
// later reads see these writes, and writes made after its earlier reads
// cannot overtake them.
__device__ __forceinline__ void stage_wave_post(const BatchOp &op, STRATA_LDS float *lt, int lane)
{
    const STRATA_CONST BatchOp *o = cp(&op);
    int lo = 0;
    for (int i = 0; i < o->npost; ++i) {
        lds_table(lt + lo, o->post_op[i], o->post_taps[i], o->post_n[i], lane, 64);
        lo += lds_table_floats(o->post_op[i], o->post_n[i]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kBatchMaxFocalR = 3;
constexpr int kBatchLdsFloats = focal_lds_floats<FK_MEAN, kBatchMaxFocalR>();
constexpr int kFocalTilesPerUnit = 4;  // focal tiles stacked in one unit

// One wave's unit: 128 columns (two per lane) of `rows` rows.
constexpr int kWaveCols = 128;

template <int OP>
__device__ __forceinline__ void batch_horn(const BatchOp &op, const BatchTask &t, int64_t lu,
                                           const STRATA_LDS float *lt)
{
    HornArgs a;
#pragma unroll
    for (int i = 0; i < 8; ++i) a.k[i] = op.k[i];
    a.flags = op.flags;
    a.aligned = t.aligned;
    a.rows = (int32_t)t.rows;
    a.blocks_x = t.blocks_x;
    const int64_t xb = t.rg.ox + (lu % t.blocks_x) * kWaveCols;
    const int64_t y0 = t.rg.oy + (lu / t.blocks_x) * t.rows;
    const int64_t y1 = min(y0 + (int64_t)t.rows, t.rg.oy + t.rg.h);
    // Units inside the source buffer and the region move pairs: aligned
    // ones directly, the others row by row at their own parity.
    const bool inside = xb >= t.s.lx0 && xb + kWaveCols <= t.s.lx1 && xb >= t.rg.ox &&
                        xb + kWaveCols <= t.rg.ox + t.rg.w;
    const ChainPost post{cp(&op), lt};
    const uint32_t lane = threadIdx.x & 63;
    if (inside && t.aligned)
        sweep<OP, kSweepPairs>(a, t.rg, t.s, t.d, t.d, xb, y0, y1, lane, post);
    else if (inside)
        sweep<OP, kSweepAnyParity>(a, t.rg, t.s, t.d, t.d, xb, y0, y1, lane, post);
    else
        sweep<OP, kSweepChecked>(a, t.rg, t.s, t.d, t.d, xb, y0, y1, lane, post);
}

template <int KIND, int R>
__device__ __forceinline__ void batch_focal_tiles(const FocalArgs &a, const BatchTask &t, int64_t lu, float *lds)
{
    const PtrTaps taps{nullptr};
    const int64_t tx0 = t.rg.ox + (lu % t.blocks_x) * kTileX;
    const int64_t ty = t.rg.oy + (lu / t.blocks_x) * t.rows;
#pragma unroll 1
    for (int k = 0; k < kFocalTilesPerUnit; ++k) {
        const int64_t ty0 = ty + k * kTileY;
        if (ty0 >= t.rg.oy + t.rg.h) break;
        focal_block<KIND, R, PtrTaps, false>(a, taps, t.rg, t.s, t.d, t.d, tx0, ty0, lds);
    }
}

template <int R>
__device__ __forceinline__ void batch_focal(const BatchOp &op, const BatchTask &t, int64_t lu, float *lds)
{
    FocalArgs a;
    a.op = op.op;
    a.flags = op.flags;
#pragma unroll
    for (int i = 0; i < 12; ++i) a.k[i] = op.k[i];
    switch (op.kind) {
    case BK_MEAN: batch_focal_tiles<FK_MEAN, R>(a, t, lu, lds); break;
    case BK_MIN: batch_focal_tiles<FK_MIN, R>(a, t, lu, lds); break;
    default: batch_focal_tiles<FK_MAX, R>(a, t, lu, lds); break;
    }
}

// One validity wave of a task per unit.
__device__ __forceinline__ void batch_validity(const BatchOp &op, const BatchTask &t, int64_t w)
{
    const int64_t end = t.rg.oy + t.rg.h;
    uint32_t cv = 0u, cown = 0u;  // a row's last bits, carried into the next row's first word
    if (t.vgroup) {
        // 64 / vgroup groups, each through a run of kVUnitChunks chunks of
        // kVgRows rows, one chunk per call.
        const int G = (int)t.vgroup;
        const int64_t run = (int64_t)kVgRows * kVUnitChunks;
        const int64_t y0 = t.rg.oy + w * (64 / G) * run;
        for (int c = 0; c < kVUnitChunks; ++c) {
            if (y0 + (int64_t)c * kVgRows >= end) break;
            switch (op.radius) {
            case 1: validity_groups<1>(t.rg, t.sm, t.has_src, t.dm, y0, end, G, run, c, cv, cown); break;
            case 2: validity_groups<2>(t.rg, t.sm, t.has_src, t.dm, y0, end, G, run, c, cv, cown); break;
            default: validity_groups<3>(t.rg, t.sm, t.has_src, t.dm, y0, end, G, run, c, cv, cown); break;
            }
        }
        return;
    }
    // kVUnitChunks strips of kVsRows rows of the unit's word column wx.
    const int64_t wx = w % t.vwaves_x, strip = (w / t.vwaves_x) * kVUnitChunks;
    const int64_t yend = min(t.rg.oy + (strip + kVUnitChunks) * kVsRows, end);
    for (int c = 0; c < kVUnitChunks; ++c) {
        const int64_t y0 = t.rg.oy + (strip + c) * kVsRows;
        if (y0 >= end) break;
        const int64_t y1 = min(y0 + (int64_t)kVsRows, end);
        switch (op.radius) {
        case 1: validity_strip<1>(t.rg, t.sm, t.has_src, t.dm, wx, y0, y1, yend, cv, cown); break;
        case 2: validity_strip<2>(t.rg, t.sm, t.has_src, t.dm, wx, y0, y1, yend, cv, cown); break;
        default: validity_strip<3>(t.rg, t.sm, t.has_src, t.dm, wx, y0, y1, yend, cv, cown); break;
        }
    }
}

// The last task whose first unit is <= u.
__device__ __forceinline__ int64_t batch_find(const BatchTask *tasks, int64_t n, int64_t u)
{
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (tasks[mid].u0 <= u)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Each list's counter sits on its own 256-byte line.
constexpr int kCounterStride = 64;  // unsigned words

__global__ __launch_bounds__(64) void batch_clear(unsigned *counters)
{
    if (threadIdx.x < 2) counters[threadIdx.x * kCounterStride] = 0u;
}

// Sweep units, then validity units, claimed four at a time by a
// workgroup: its four waves take four consecutive units, which for a
// radius-1 task are 512 adjacent columns of the same rows (the shape of
// horn_sweep's workgroups, so rows stream as whole DRAM pages and the
// waves share their boundary lines in L1). The next claim is made before
// the current units run, into the other of two LDS slots. The lists and
// ops are read-only for the whole launch and never alias the rasters the
// units write (__restrict__), so their fields stay in registers across
// the units' stores instead of being reloaded after each one.
__global__ __launch_bounds__(256) void batch_wave_kernel(const BatchHeader *__restrict__ hdr,
                                                         const BatchTask *__restrict__ sweeps,
                                                         const BatchTask *__restrict__ valids,
                                                         const BatchOp *__restrict__ ops, unsigned *counter)
{
    __shared__ unsigned s_claim[2];
    __shared__ __align__(16) float s_post[4][kWavePostFloats];  // each wave's copy of its op's point tables
    const int64_t nh = hdr[BF_HORN].nunits, nv = hdr[BF_VALIDITY].nunits;
    const int64_t th = hdr[BF_HORN].ntasks, tv = hdr[BF_VALIDITY].ntasks;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int64_t ti = 0, tu0 = 0, tu1 = 0;  // the current task and its units [tu0, tu1) in its list
    bool in_valid = false;            // which list the cached task belongs to
    int32_t staged = -1;              // the op whose point tables the wave's LDS copy holds
    STRATA_LDS float *lt = (STRATA_LDS float *)s_post[wv];
    if (threadIdx.x == 0) s_claim[0] = atomicAdd(counter, 4u);
    __syncthreads();
    int slot = 0;
    int64_t base = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane(s_claim[0]);
    while (base < nh + nv) {
        if (threadIdx.x == 0) s_claim[slot ^ 1] = atomicAdd(counter, 4u);  // consumed after these units
        const int64_t u = base + wv;
        if (u < nh + nv) {
        const bool valid = u >= nh;
        const BatchTask *list = valid ? valids : sweeps;
        const int64_t lu_all = valid ? u - nh : u;
        if (valid != in_valid || lu_all < tu0 || lu_all >= tu1) {
            ti = batch_find(list, valid ? tv : th, lu_all);
            tu0 = list[ti].u0;
            tu1 = list[ti].u1;
            in_valid = valid;
        }
        const BatchTask &t = list[ti];
        const BatchOp &op = ops[t.op];
        const int64_t lu = lu_all - tu0;
        if (!valid) {
            if (op.npost > 0 && t.op != staged) {
                stage_wave_post(op, lt, (int)(threadIdx.x & 63));
                staged = t.op;
            }
            switch (op.op) {
            case STRATA_OP_SLOPE: batch_horn<STRATA_OP_SLOPE>(op, t, lu, lt); break;
            case STRATA_OP_ASPECT: batch_horn<STRATA_OP_ASPECT>(op, t, lu, lt); break;
            case STRATA_OP_HILLSHADE: batch_horn<STRATA_OP_HILLSHADE>(op, t, lu, lt); break;
            default: batch_horn<STRATA_OP_CURVATURE>(op, t, lu, lt); break;
            }
        } else {
            batch_validity(op, t, lu);
        }
        }
        __syncthreads();  // the next claim is visible, and every wave is done with these units
        slot ^= 1;
        base = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane(s_claim[slot]);
    }
}

// Focal units use the workgroup's LDS tile: the workgroup claims them
// together.
__global__ __launch_bounds__(256) void batch_focal_kernel(const BatchHeader *__restrict__ hdr,
                                                          const BatchTask *__restrict__ tasks,
                                                          const BatchOp *__restrict__ ops, unsigned *counter)
{
    __shared__ float lds[kBatchLdsFloats];
    __shared__ unsigned claimed;
    const int64_t ntasks = hdr[BF_FOCAL].ntasks, nunits = hdr[BF_FOCAL].nunits;
    int64_t ti = 0, tu0 = 0, tu1 = 0;
    if (threadIdx.x == 0) claimed = atomicAdd(counter, 1u);
    __syncthreads();
    int64_t u = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane(claimed);
    while (u < nunits) {
        unsigned next = 0u;
        if (threadIdx.x == 0) next = atomicAdd(counter, 1u);  // consumed after this unit
        if (u < tu0 || u >= tu1) {
            ti = batch_find(tasks, ntasks, u);
            tu0 = tasks[ti].u0;
            tu1 = tasks[ti].u1;
        }
        const BatchTask &t = tasks[ti];
        const BatchOp &op = ops[t.op];
        switch (op.radius) {
        case 1: batch_focal<1>(op, t, u - tu0, lds); break;
        case 2: batch_focal<2>(op, t, u - tu0, lds); break;
        default: batch_focal<3>(op, t, u - tu0, lds); break;
        }
        __syncthreads();  // the unit is done (and the LDS free) on every wave
        if (threadIdx.x == 0) claimed = next;
        __syncthreads();
        u = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane(claimed);
    }
}

// The focal list when every focal task has the same kind and radius: each
// tile is one focal_block of that op alone, with the per-op kernels'
// cell-by-cell staging (every load of the thread in flight at 32-bit
// offsets), which the kernel that inlines every kind cannot afford in
// registers.
template <int KIND, int R>
__global__ __launch_bounds__(256) void batch_focal_one(const BatchHeader *__restrict__ hdr,
                                                       const BatchTask *__restrict__ tasks,
                                                       const BatchOp *__restrict__ ops, unsigned *counter)
{
    __shared__ float lds[kBatchLdsFloats];
    static_assert(focal_lds_floats<KIND, R>() <= kBatchLdsFloats, "the tile fits the batch's LDS");
    __shared__ unsigned claimed;
    const int64_t ntasks = hdr[BF_FOCAL].ntasks, nunits = hdr[BF_FOCAL].nunits;
    int64_t ti = 0, tu0 = 0, tu1 = 0;
    if (threadIdx.x == 0) claimed = atomicAdd(counter, 1u);
    __syncthreads();
    int64_t u = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane(claimed);
    while (u < nunits) {
        unsigned next = 0u;
        if (threadIdx.x == 0) next = atomicAdd(counter, 1u);  // consumed after this unit
        if (u < tu0 || u >= tu1) {
            ti = batch_find(tasks, ntasks, u);
            tu0 = tasks[ti].u0;
            tu1 = tasks[ti].u1;
        }
        const BatchTask &t = tasks[ti];
        const BatchOp &op = ops[t.op];
        FocalArgs a;
        a.op = op.op;
        a.flags = op.flags;
#pragma unroll
        for (int i = 0; i < 12; ++i) a.k[i] = op.k[i];
        const PtrTaps taps{nullptr};  // box kinds read no taps
        const int64_t lu = u - tu0;
        const int64_t tx0 = t.rg.ox + (lu % t.blocks_x) * kTileX;
        const int64_t ty = t.rg.oy + (lu / t.blocks_x) * t.rows;
#pragma unroll 1
        for (int k = 0; k < kFocalTilesPerUnit; ++k) {
            const int64_t ty0 = ty + k * kTileY;
            if (ty0 >= t.rg.oy + t.rg.h) break;
            focal_block<KIND, R, PtrTaps, true, true>(a, taps, t.rg, t.s, t.d, t.d, tx0, ty0, lds);
        }
        __syncthreads();  // the unit is done (and the LDS free) on every wave
        if (threadIdx.x == 0) claimed = next;
        __syncthreads();
        u = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane(claimed);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
constexpr int64_t kMaxGridY = 65535;
// Workgroups of a 1-D tile launch: gridDim.x times 256 threads stays below 2^32.
constexpr int64_t kMaxTileBlocks = int64_t(1) << 23;

// Launch-shape tuning (never changes a result).
static Tuning g_tuning{16, 4096, 0, 0, {0, 0, 0, 0}};
static int sweep_rows() { return g_tuning.sweep_rows; }

Tuning get_tuning() { return g_tuning; }

bool set_tuning(const Tuning &t)
{
    if (t.sweep_rows < 4 || t.sweep_rows > 4096) return false;
    if (t.fold_stripe_mb < 1 || t.fold_stripe_mb > 4096) return false;
    if (t.unit_kcells < 0 || t.unit_kcells > (1 << 20) || t.stop_after_units < 0) return false;
    g_tuning = t;
    return true;
}

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
    a.rows = sweep_rows();
    const int64_t gx = cdiv(rg.w, kSweepCols);
    a.blocks_x = (uint32_t)gx;
    // One 1-D launch per run of strips that keeps the block count in 32 bits.
    const int64_t strips_per_launch = std::max<int64_t>(1, ((int64_t)1 << 30) / gx);
    for (int64_t sb = 0; sb < cdiv(rg.h, a.rows); sb += strips_per_launch) {
        Region sub = rg;
        sub.oy = rg.oy + sb * a.rows;
        sub.h = std::min<int64_t>(rg.h - sb * a.rows, strips_per_launch * a.rows);
        DstView e0 = d[0], e1 = d1;
        e0.p += sb * a.rows * e0.stride;
        e1.p += sb * a.rows * e1.stride;
        const dim3 grid((unsigned)(gx * cdiv(sub.h, a.rows)));
        if constexpr (OP != STRATA_OP_GRADIENT) {
            if (P.post.n > 0) {
                if (!P.post_dev) return hipErrorInvalidValue;
                if (aligned)
                    hipLaunchKernelGGL((horn_sweep_post<OP, false>), grid, dim3(kSweepThreads), 0, st, a, sub, s, e0,
                                       P.post_dev);
                else
                    hipLaunchKernelGGL((horn_sweep_post<OP, true>), grid, dim3(kSweepThreads), 0, st, a, sub, s, e0,
                                       P.post_dev);
                continue;
            }
        }
        if (aligned)
            hipLaunchKernelGGL((horn_sweep<OP, false>), grid, dim3(kSweepThreads), 0, st, a, sub, s, e0, e1);
        else
            hipLaunchKernelGGL((horn_sweep<OP, true>), grid, dim3(kSweepThreads), 0, st, a, sub, s, e0, e1);
    }
    return hipGetLastError();
}

// Surface (fan == false) or a fan-out of radius-1 ops (fan == true): one
// sweep writing every output view of d.
static hipError_t run_surface(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                              hipStream_t st, bool fan = false)
{
    HornArgs a;
    for (int i = 0; i < 8; ++i) a.k[i] = P.k[i];
    a.flags = P.flags;
    auto al8 = [](const void *p) { return ((uintptr_t)p & 7) == 0; };
    bool aligned = al8(s.p) && (s.stride & 1) == 0 && ((rg.ox - s.sx) & 1) == 0;
    for (int o = 0; o < nout; ++o) aligned = aligned && al8(d[o].p) && (d[o].stride & 1) == 0;
    a.aligned = aligned ? 1 : 0;
    a.rows = sweep_rows();
    const int64_t gx = cdiv(rg.w, kSweepCols);
    a.blocks_x = (uint32_t)gx;
    const int64_t strips_per_launch = std::max<int64_t>(1, ((int64_t)1 << 30) / gx);
    for (int64_t sb = 0; sb < cdiv(rg.h, a.rows); sb += strips_per_launch) {
        Region sub = rg;
        sub.oy = rg.oy + sb * a.rows;
        sub.h = std::min<int64_t>(rg.h - sb * a.rows, strips_per_launch * a.rows);
        DstSet e;
        for (int o = 0; o < kMaxOutputs; ++o) {
            e.v[o] = d[o < nout ? o : 0];
            e.v[o].p += sb * a.rows * e.v[o].stride;
        }
        const dim3 grid((unsigned)(gx * cdiv(sub.h, a.rows)));
        if (fan)
            hipLaunchKernelGGL(horn_fan, grid, dim3(kSweepThreads), 0, st, a, P.fan, sub, s, e);
        else
            hipLaunchKernelGGL(horn_surface, grid, dim3(kSweepThreads), 0, st, a, sub, s, e);
    }
    return hipGetLastError();
}

template <int KIND, int R>
static hipError_t run_focal_r(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                              hipStream_t st)
{
    // A 1-D grid of tiles (xcd_tile numbers them), at most kMaxTileBlocks
    // workgroups per launch.
    const int64_t gx = cdiv(rg.w, kTileX), step = std::max<int64_t>(1, kMaxTileBlocks / gx);
    for (int64_t yb = 0; yb < cdiv(rg.h, kTileY); yb += step) {
        Region sub = rg;
        sub.oy = rg.oy + yb * kTileY;
        sub.h = min(rg.h - yb * kTileY, step * kTileY);
        DstSet e;
        for (int o = 0; o < kMaxOutputs; ++o) {
            e.v[o] = d[o < nout ? o : 0];
            e.v[o].p += yb * kTileY * e.v[o].stride;
        }
        const dim3 grid((unsigned)(gx * cdiv(sub.h, kTileY)));
        hipLaunchKernelGGL((focal_tile<KIND, R>), grid, dim3(kTileThreads), 0, st, P, sub, s, e);
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

template <int R>
static hipError_t run_mfma_r(const Params &P, const Region &rg, const SrcView &s, const DstView *d, hipStream_t st)
{
    constexpr int N = 2 * R + 1;
    bf16x8 *tab = nullptr;
    hipError_t e = scratch_alloc_async((void **)&tab, sizeof(bf16x8) * 2 * 64 * N, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mfma_toeplitz<R>, dim3(N), dim3(64), 0, st, P, tab);
    // Resident workgroups per XCD: the device's CUs times the kernel's
    // occupancy, over the XCDs.
    int dev = 0, cus = 0, per_cu = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess ||
        (e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, correlate_mfma_pipe<R>, 512, 0)) != hipSuccess) {
        (void)hipFreeAsync(tab, st);
        return e;
    }
    const int64_t resident = std::max<int64_t>(1, (int64_t)cus * std::max(per_cu, 1) / kXcds);
    const int64_t gx = cdiv(rg.w, kMfmaTile), step = std::max<int64_t>(1, kMaxTileBlocks / gx);
    for (int64_t yb = 0; yb < cdiv(rg.h, kMfmaTile); yb += step) {
        Region sub = rg;
        sub.oy = rg.oy + yb * kMfmaTile;
        sub.h = min(rg.h - yb * kMfmaTile, step * kMfmaTile);
        DstView e0 = d[0];
        e0.p += yb * kMfmaTile * e0.stride;
        // As many resident workgroups as the chip holds, in whole multiples
        // of the XCD count, and no more than the tiles need.
        const int64_t xcd_tiles = gx * cdiv(cdiv(sub.h, kMfmaTile), kXcds);  // the largest band's tiles
        const int64_t grid = std::max<int64_t>(kXcds, std::min<int64_t>(resident, xcd_tiles) * kXcds);
        hipLaunchKernelGGL(correlate_mfma_pipe<R>, dim3((unsigned)grid), dim3(512), 0, st, P, sub, s, e0, tab);
    }
    e = hipGetLastError();
    const hipError_t f = hipFreeAsync(tab, st);
    return e != hipSuccess ? e : f;
}

static hipError_t run_mfma(const Params &P, const Region &rg, const SrcView &s, const DstView *d, hipStream_t st)
{
    switch (P.radius) {
    case 1: return run_mfma_r<1>(P, rg, s, d, st);
    case 2: return run_mfma_r<2>(P, rg, s, d, st);
    case 3: return run_mfma_r<3>(P, rg, s, d, st);
    case 4: return run_mfma_r<4>(P, rg, s, d, st);
    case 5: return run_mfma_r<5>(P, rg, s, d, st);
    case 6: return run_mfma_r<6>(P, rg, s, d, st);
    case 7: return run_mfma_r<7>(P, rg, s, d, st);
    case 8: return run_mfma_r<8>(P, rg, s, d, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_correlate_error(const Params &P, const Region &rg, const SrcView &s, const float *out,
                                  int64_t out_stride, unsigned long long *acc, hipStream_t st)
{
    if (rg.w <= 0 || rg.h <= 0) return hipSuccess;
    const int64_t gx = cdiv(rg.w, 256);
    for (int64_t yb = 0; yb < rg.h; yb += kMaxGridY) {
        Region sub = rg;
        sub.oy = rg.oy + yb;
        sub.h = min(rg.h - yb, kMaxGridY);
        const float *o = out + yb * out_stride;
        const dim3 grid((unsigned)gx, (unsigned)sub.h);
        switch (P.radius) {
#define STRATA_ERR_CASE(RR)                                                                                   \
    case RR:                                                                                                  \
        hipLaunchKernelGGL(correlate_error_f64<RR>, grid, dim3(256), 0, st, P, sub, s, o, out_stride, acc);    \
        break;
            STRATA_ERR_CASE(1)
            STRATA_ERR_CASE(2)
            STRATA_ERR_CASE(3)
            STRATA_ERR_CASE(4)
            STRATA_ERR_CASE(5)
            STRATA_ERR_CASE(6)
            STRATA_ERR_CASE(7)
            STRATA_ERR_CASE(8)
#undef STRATA_ERR_CASE
        default:
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
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
    if (fit && P.op >= STRATA_OP_GRADIENT && P.op <= STRATA_OP_SURFACE)
        return run_focal<FK_FIT>(P, rg, s, d, nout, st);
    switch (P.op) {
    case STRATA_OP_SURFACE: return run_surface(P, rg, s, d, nout, st);
    case kOpHornFan: return run_surface(P, rg, s, d, nout, st, true);
    case STRATA_OP_GRADIENT: return run_horn<STRATA_OP_GRADIENT>(P, rg, s, d, nout, st);
    case STRATA_OP_SLOPE: return run_horn<STRATA_OP_SLOPE>(P, rg, s, d, nout, st);
    case STRATA_OP_ASPECT: return run_horn<STRATA_OP_ASPECT>(P, rg, s, d, nout, st);
    case STRATA_OP_HILLSHADE: return run_horn<STRATA_OP_HILLSHADE>(P, rg, s, d, nout, st);
    case STRATA_OP_CURVATURE: return run_horn<STRATA_OP_CURVATURE>(P, rg, s, d, nout, st);
    case STRATA_OP_CORRELATE:
        if (P.flags & STRATA_FLAG_MFMA) return run_mfma(P, rg, s, d, st);
        return run_focal<FK_CORRELATE>(P, rg, s, d, nout, st);
    case STRATA_OP_SEPARABLE: return run_focal<FK_SEPARABLE>(P, rg, s, d, nout, st);
    case STRATA_OP_MEAN: return run_focal<FK_MEAN>(P, rg, s, d, nout, st);
    case STRATA_OP_MIN: return run_focal<FK_MIN>(P, rg, s, d, nout, st);
    case STRATA_OP_MAX: return run_focal<FK_MAX>(P, rg, s, d, nout, st);
    case STRATA_OP_LOOKUP: return run_point<STRATA_OP_LOOKUP>(P, rg, s, d, st);
    case STRATA_OP_RECLASS: return run_point<STRATA_OP_RECLASS>(P, rg, s, d, st);
    }
    return hipErrorInvalidValue;
}

template <int OP>
static hipError_t run_horn_valid(const Params &P, const Region &rg, const SrcView &s, const DstView &d,
                                 const MaskView &vsrc, int has_src, const MaskView &vdst, hipStream_t st, bool *fused)
{
    HornArgs a;
    for (int i = 0; i < 8; ++i) a.k[i] = P.k[i];
    a.flags = P.flags;
    auto al8 = [](const void *p) { return ((uintptr_t)p & 7) == 0; };
    const bool aligned = al8(s.p) && (s.stride & 1) == 0 && ((rg.ox - s.sx) & 1) == 0 && al8(d.p) &&
                         (d.stride & 1) == 0;
    a.aligned = aligned ? 1 : 0;
    a.rows = sweep_rows();
    const int64_t gx = cdiv(rg.w, kSweepCols);
    a.blocks_x = (uint32_t)gx;
    const int64_t data_blocks = gx * cdiv(rg.h, a.rows);
    const int64_t waves_x = cdiv(rg.w / 32 + 2, kVsOut);
    const int64_t vblocks = cdiv(waves_x * cdiv(rg.h, kVsRows), 4);
    // One launch of at most 2^31 blocks; larger regions take the two kernels.
    if (data_blocks + vblocks > ((int64_t)1 << 31) - 1) {
        *fused = false;
        return hipSuccess;
    }
    const dim3 grid((unsigned)(data_blocks + vblocks));
    if (aligned)
        hipLaunchKernelGGL((horn_sweep_valid<OP, false>), grid, dim3(kSweepThreads), 0, st, a, rg, s, d,
                           (uint32_t)data_blocks, vsrc, has_src, vdst, (uint32_t)waves_x);
    else
        hipLaunchKernelGGL((horn_sweep_valid<OP, true>), grid, dim3(kSweepThreads), 0, st, a, rg, s, d,
                           (uint32_t)data_blocks, vsrc, has_src, vdst, (uint32_t)waves_x);
    *fused = true;
    return hipGetLastError();
}

hipError_t launch_data_valid(const Params &P, const Region &rg, const SrcView &s, const DstView &d,
                             const MaskView &vsrc, int has_src, const MaskView &vdst, hipStream_t st, bool *fused)
{
    *fused = false;
    if (rg.w <= 0 || rg.h <= 0) return hipSuccess;
    if ((P.flags & STRATA_FLAG_FIT) || P.radius != 1 || P.post.n > 0) return hipSuccess;
    switch (P.op) {
    case STRATA_OP_SLOPE: return run_horn_valid<STRATA_OP_SLOPE>(P, rg, s, d, vsrc, has_src, vdst, st, fused);
    case STRATA_OP_ASPECT: return run_horn_valid<STRATA_OP_ASPECT>(P, rg, s, d, vsrc, has_src, vdst, st, fused);
    case STRATA_OP_HILLSHADE:
        return run_horn_valid<STRATA_OP_HILLSHADE>(P, rg, s, d, vsrc, has_src, vdst, st, fused);
    case STRATA_OP_CURVATURE:
        return run_horn_valid<STRATA_OP_CURVATURE>(P, rg, s, d, vsrc, has_src, vdst, st, fused);
    }
    return hipSuccess;
}

hipError_t launch_validity(int r, const Region &rg, const MaskView &src, int has_src, const MaskView &dst,
                           hipStream_t st)
{
    if (rg.w <= 0 || rg.h <= 0) return hipSuccess;
    // Waves per strip: the row's words (at most w/32 + 2), kVsOut per wave.
    const int64_t waves_x = cdiv(rg.w / 32 + 2, kVsOut);
    const int64_t strips_per_launch = std::max<int64_t>(1, ((int64_t)1 << 30) / waves_x);
    for (int64_t sb = 0; sb < cdiv(rg.h, kVsRows); sb += strips_per_launch) {
        Region sub = rg;
        sub.oy = rg.oy + sb * kVsRows;
        sub.h = std::min<int64_t>(rg.h - sb * kVsRows, strips_per_launch * kVsRows);
        const int64_t waves = waves_x * cdiv(sub.h, kVsRows);
        const dim3 grid((unsigned)cdiv(waves, 4));
        switch (r) {
#define STRATA_VALIDITY_CASE(RR)                                                                             \
    case RR:                                                                                                 \
        hipLaunchKernelGGL(validity_kernel<RR>, grid, dim3(256), 0, st, sub, src, has_src, dst,               \
                           (uint32_t)waves_x);                                                               \
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

void batch_units(const BatchOp &op, BatchTask &t, int rows)
{
    if (op.kind == BK_HORN) {
        t.rows = (uint32_t)rows;
        t.blocks_x = (uint32_t)cdiv(t.rg.w, kWaveCols);
    } else {
        t.rows = kTileY * kFocalTilesPerUnit;
        t.blocks_x = (uint32_t)cdiv(t.rg.w, kTileX);
    }
    t.vwaves_x = (uint32_t)cdiv(t.rg.w / 32 + 2, kVsOut);
    // A narrow region's rows (output words plus four helper lanes) fit in
    // a group of lanes: several row blocks share a wave.
    const int64_t g = cdiv(t.rg.w, 32) + 5;
    t.vgroup = g <= 32 ? (uint32_t)g : 0u;
    if (t.vgroup) t.vwaves_x = 1;
}

hipError_t launch_batch(const BatchLaunch &L, hipStream_t st)
{
    if (L.clear) {
        hipLaunchKernelGGL(batch_clear, dim3(1), dim3(64), 0, st, L.counters);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // The focal list, then the wave lists, one after the other on st. With
    // the focal kernel on a second stream beside the wave kernel, the two
    // shared the chip and the bandwidth, and the mixed 4096-tile batch took
    // 7.52 ms against 7.24 ms one after the other.
    if (L.nunits[BF_FOCAL] > 0) {
        const dim3 g((unsigned)L.blocks_focal);
        const BatchTask *ft = L.tasks[BF_FOCAL];
        unsigned *fc = L.counters + kCounterStride;
        switch (L.focal_one) {
#define STRATA_FOCAL_ONE(K, KIND, R)                                                                          \
    case batch_focal_one_code(K, R):                                                                          \
        hipLaunchKernelGGL((batch_focal_one<KIND, R>), g, dim3(256), 0, st, L.hdr, ft, L.ops, fc);          \
        break;
            STRATA_FOCAL_ONE(BK_MEAN, FK_MEAN, 1)
            STRATA_FOCAL_ONE(BK_MEAN, FK_MEAN, 2)
            STRATA_FOCAL_ONE(BK_MEAN, FK_MEAN, 3)
            STRATA_FOCAL_ONE(BK_MIN, FK_MIN, 1)
            STRATA_FOCAL_ONE(BK_MIN, FK_MIN, 2)
            STRATA_FOCAL_ONE(BK_MIN, FK_MIN, 3)
            STRATA_FOCAL_ONE(BK_MAX, FK_MAX, 1)
            STRATA_FOCAL_ONE(BK_MAX, FK_MAX, 2)
            STRATA_FOCAL_ONE(BK_MAX, FK_MAX, 3)
#undef STRATA_FOCAL_ONE
        default:
            hipLaunchKernelGGL(batch_focal_kernel, g, dim3(256), 0, st, L.hdr, ft, L.ops, fc);
        }
    }
    if (L.nunits[BF_HORN] + L.nunits[BF_VALIDITY] > 0)
        hipLaunchKernelGGL(batch_wave_kernel, dim3((unsigned)L.blocks_wave), dim3(256), 0, st, L.hdr,
                           L.tasks[BF_HORN], L.tasks[BF_VALIDITY], L.ops, L.counters);
    return hipGetLastError();
}

int batch_blocks_per_cu(int family, int focal_one)
{
    int n = 0;
    hipError_t e = hipSuccess;
    if (family != BF_FOCAL) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_wave_kernel, 256, 0);
    } else {
        switch (focal_one) {
#define STRATA_FOCAL_ONE(K, KIND, R)                                                                          \
    case batch_focal_one_code(K, R):                                                                          \
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_focal_one<KIND, R>, 256, 0);               \
        break;
            STRATA_FOCAL_ONE(BK_MEAN, FK_MEAN, 1)
            STRATA_FOCAL_ONE(BK_MEAN, FK_MEAN, 2)
            STRATA_FOCAL_ONE(BK_MEAN, FK_MEAN, 3)
            STRATA_FOCAL_ONE(BK_MIN, FK_MIN, 1)
            STRATA_FOCAL_ONE(BK_MIN, FK_MIN, 2)
            STRATA_FOCAL_ONE(BK_MIN, FK_MIN, 3)
            STRATA_FOCAL_ONE(BK_MAX, FK_MAX, 1)
            STRATA_FOCAL_ONE(BK_MAX, FK_MAX, 2)
            STRATA_FOCAL_ONE(BK_MAX, FK_MAX, 3)
#undef STRATA_FOCAL_ONE
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_focal_kernel, 256, 0);
        }
    }
    return (e != hipSuccess || n < 1) ? 1 : n;
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
        } else if (kind == 3) {
            // The slope fast path on s = gx^2 + gy^2 against the full one.
            const float x = u2f((uint32_t)(lo + i));
            if (slope_small_ok(x)) {
                ran = true;
                differ = f2u(atan32_small(sqrt_fast(x))) != f2u(atan32(__builtin_sqrtf(x)));
            }
        } else {
            // |n| and d in [2^-30, 2^30), n with a random sign, random mantissas.
            const uint32_t h1 = mix32(lo * 0x9e3779b97f4a7c15ull + 2 * i), h2 = mix32(lo + 2 * i + 1);
            const uint32_t en = 127 - 30 + (h1 >> 26) % 60, ed = 127 - 30 + (h2 >> 26) % 60;
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
