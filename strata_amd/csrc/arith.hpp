// arith.hpp — per-cell float32 arithmetic of the stencil path, device side.
//
// Every function is the reference's scalar (canonical) kernel arithmetic in
// the same operation order, each operator rounding once to binary32. The
// library is built with -ffp-contract=off and without fast-math, so no
// multiply-add is fused, `/` is the correctly rounded quotient and sqrtf the
// correctly rounded root (equal to float32(math.Sqrt(float64(x))) in Go).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace strata {

__device__ __forceinline__ uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }

// float32(math.NaN()): the value the engine writes into edge cells
// (internal/exec/tile.go setup, terrain/stencil.go window3.Edge).
constexpr uint32_t kEdgeNaNBits = 0x7fc00000u;
__device__ __forceinline__ float edge_nan() { return u2f(kEdgeNaNBits); }

// Cephes atanf constants (internal/stencil/horn.go:226-238). The pi
// multiples and 180/pi are the nearest binary32 values of the exact
// constants, as Go's constant conversion gives.
constexpr float kTan3Pi8 = 2.414213562373095f;
constexpr float kTanPi8 = 0.4142135623730950f;
constexpr float kAtanC4 = 8.05374449538e-2f;
constexpr float kAtanC3 = 1.38776856032e-1f;
constexpr float kAtanC2 = 1.99777106478e-1f;
constexpr float kAtanC1 = 3.33329491539e-1f;
constexpr uint32_t kPi2Bits = 0x3fc90fdbu;
constexpr uint32_t kPi4Bits = 0x3f490fdbu;
constexpr uint32_t kPiBits = 0x40490fdbu;
constexpr uint32_t kRadToDegBits = 0x42652ee1u;

// Atan32 (internal/stencil/horn.go:250-262): for x >= 0, +Inf or NaN.
__device__ __forceinline__ float atan32(float x)
{
    float num = x, den = 1.0f, y0 = 0.0f;
    bool mid = x > kTanPi8;
    bool hi = x > kTan3Pi8;
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
    float t = num / den;
    float z = t * t;
    float p = kAtanC4 * z - kAtanC3;
    p = p * z + kAtanC2;
    p = p * z - kAtanC1;
    return y0 + ((p * z) * t + t);
}

// Atan2F32 (internal/stencil/aspect.go:36-48).
__device__ __forceinline__ float atan2f32(float y, float x)
{
    float ax = u2f(f2u(x) & 0x7fffffffu);
    float ay = u2f(f2u(y) & 0x7fffffffu);
    float a = atan32(ay / ax);
    if (ay == 0.0f && ax == 0.0f) a = 0.0f;
    if (ay == ax && ay > 3.40282346638528859811704183484516925440e+38f) a = u2f(kPi4Bits);
    if (f2u(x) & 0x80000000u) a = u2f(kPiBits) - a;
    return u2f(f2u(a) | (f2u(y) & 0x80000000u));
}

// Horn weighted differences (internal/stencil/horn.go:172-180).
__device__ __forceinline__ float horn_dx(float z1, float z3, float z4, float z6, float z7, float z9)
{
    float d = z6 - z4;
    return ((z3 - z1) + (z9 - z7)) + (d + d);
}
__device__ __forceinline__ float horn_dy(float z1, float z2, float z3, float z7, float z8, float z9)
{
    float d = z8 - z2;
    return ((z7 - z1) + (z9 - z3)) + (d + d);
}

// magnitude (horn.go:226-228).
__device__ __forceinline__ float magnitude(float gx, float gy) { return __builtin_sqrtf(gx * gx + gy * gy); }

// Slope from a scaled gradient (horn.go:200-222, gradrow.go:44-58).
__device__ __forceinline__ float slope_of(float gx, float gy, float scale, bool atan_)
{
    float m = magnitude(gx, gy);
    if (atan_) return atan32(m) * scale;
    return m * scale;
}

// aspectArgs + aspectDegrees (internal/stencil/aspect.go:94-123).
__device__ __forceinline__ float aspect_of(float gx, float gy, float flat, bool trig)
{
    float y, x;
    if (trig) {
        y = gy;
        x = 0.0f - gx;
    } else {
        y = 0.0f - gx;
        x = gy;
    }
    float d = atan2f32(y, x) * u2f(kRadToDegBits);
    float off = 0.0f;
    if (d < 0.0f) off = 360.0f;
    d += off;
    if (d >= 360.0f) d = 0.0f;
    if (y == 0.0f && x == 0.0f) d = flat;
    return d;
}

// shade (internal/stencil/aspect.go:161-173). Written so NaN falls
// through both clamps unchanged, as Go's `if v < 0` / `if v > 255` do.
__device__ __forceinline__ float shade_of(float gx, float gy, float c, float bx, float by)
{
    float num = c + (bx * gx + by * gy);
    float den = __builtin_sqrtf(1.0f + (gx * gx + gy * gy));
    float v = num / den;
    if (v < 0.0f) v = 0.0f;
    if (v > 255.0f) v = 255.0f;
    return v;
}

// Curvature from the five derivatives (internal/stencil/curvature.go:94-129,
// 138-183). kind: 0 profile, 1 plan, 2 mean.
__device__ __forceinline__ float curvature_of(float p, float q, float r, float s, float t, int kind)
{
    float p2 = p * p, q2 = q * q;
    float pq = p * q;
    if (kind == 0) {
        float g = p2 + q2;
        float w = 1.0f + g;
        float num = (p2 * r + (pq + pq) * s) + q2 * t;
        float v = 0.0f - num / g / (w * __builtin_sqrtf(w));
        if (g == 0.0f) v = num + 0.0f;
        return v;
    }
    if (kind == 1) {
        float g = p2 + q2;
        float num = (q2 * r - (pq + pq) * s) + p2 * t;
        float v = 0.0f - num / g / __builtin_sqrtf(g);
        if (g == 0.0f) v = num + 0.0f;
        return v;
    }
    float w = 1.0f + (p2 + q2);
    float num = ((1.0f + q2) * r - (pq + pq) * s) + (1.0f + p2) * t;
    float d = w * __builtin_sqrtf(w);
    return 0.0f - num / (d + d);
}

// Zevenbergen-Thorne derivatives of one 3x3 window (curvature.go:79-88).
__device__ __forceinline__ float zt_curvature(float z1, float z2, float z3, float z4, float z5, float z6, float z7,
                                              float z8, float z9, const float *kc, int kind)
{
    float p = (z6 - z4) * kc[0];
    float q = (z8 - z2) * kc[1];
    float c = z5 + z5;
    float r = ((z4 + z6) - c) * kc[2];
    float t = ((z2 + z8) - c) * kc[3];
    float s = ((z1 + z9) - (z3 + z7)) * kc[4];
    return curvature_of(p, q, r, s, t, kind);
}

// Go's builtin min and max on float32: NaN wins, -0 < +0. These are the
// IEEE 754-2019 minimum/maximum operations (v_minimum3/v_maximum3 on gfx950).
__device__ __forceinline__ float go_min(float a, float b) { return __builtin_elementwise_minimum(a, b); }
__device__ __forceinline__ float go_max(float a, float b) { return __builtin_elementwise_maximum(a, b); }

// internal/curve/scalar.go:39-56 (Reclass) and 84-114 (Lookup). The scan
// for the first knot above v runs over every knot with a uniform index and
// keeps the first hit by selects: the tables are read with wave-uniform
// indices only (scalar loads, no per-lane gathers and no divergent loop),
// and the value is computed from the same knots in the same operations.
// T is the tables' pointer type (a generic, global or constant pointer to
// const float).
template <class T>
__device__ __forceinline__ float reclass_of(float v, T breaks, int nb, T values)
{
    float out = values[nb];
    bool found = false;
    for (int j = 0; j < nb; j++) {
        const bool hit = !found && v < breaks[j];
        out = hit ? values[j] : out;
        found = found || hit;
    }
    return v != v ? v : out;
}
template <class T>
__device__ __forceinline__ float lookup_of(float v, T xs, T ys, int n)
{
    // state 2: no knot above v (k == n); 0: the first knot (k == 0); 1:
    // between knots k-1 and k, which x0, y0, x1, y1 hold.
    int state = 2;
    float x0 = 0.0f, y0 = 0.0f, x1 = 1.0f, y1 = 0.0f, px = 0.0f, py = 0.0f;
    for (int j = 0; j < n; j++) {
        const float xj = xs[j], yj = ys[j];
        const bool hit = state == 2 && v < xj;
        if (hit) {
            state = j == 0 ? 0 : 1;
            x0 = px;
            y0 = py;
            x1 = xj;
            y1 = yj;
        }
        px = xj;
        py = yj;
    }
    const float t = (v - x0) / (x1 - x0);
    const float mid = v == x0 ? y0 : y0 + t * (y1 - y0);
    const float out = state == 0 ? ys[0] : (state == 2 ? ys[n - 1] : mid);
    return v != v ? v : out;
}

// The counting forms of the same two ops. For a table that strictly
// increases and holds no NaN (the reference's constructors panic on any
// other, transfer/transfer.go:204-221, and the library's checks refuse
// them), the entries above a value that is not NaN are a suffix of the
// table, so the scan's k (the first entry above v) is the number of
// entries at or below v: one compare and one add per entry, with no
// selects of the entries' values. The entries k picks are then gathered
// per lane from a copy of the table in LDS. A NaN v counts every entry
// (every v < x is false), which keeps k inside the table, and gives v.
// Each form takes N values, which share every read of the table.
template <int N, class T>
__device__ __forceinline__ void count_at_or_below(const float (&v)[N], T xs, int n, int (&k)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i) k[i] = 0;
    for (int j = 0; j < n; j++) {
        const float x = xs[j];
#pragma unroll
        for (int i = 0; i < N; ++i) k[i] += v[i] < x ? 0 : 1;
    }
}
// values: the nb + 1 class values in LDS.
template <int N, class T, class L>
__device__ __forceinline__ void reclass_count(float (&v)[N], T breaks, int nb, L values)
{
    int k[N];
    count_at_or_below(v, breaks, nb, k);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const float out = values[k[i]];
        v[i] = v[i] != v[i] ? v[i] : out;
    }
}
// knots: (xs[j], ys[j]) as 2-vectors in LDS, one per knot.
template <int N, class T, class L2>
__device__ __forceinline__ void lookup_count(float (&v)[N], T xs, T ys, int n, L2 knots)
{
    if (n == 1) {  // every value is below, on or above the one knot
        const float y = ys[0];
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = v[i] != v[i] ? v[i] : y;
        return;
    }
    int k[N];
    count_at_or_below(v, xs, n, k);
    const float ylo = ys[0], yhi = ys[n - 1];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const int kk = min(max(k[i], 1), n - 1);  // the segment for 0 < k < n; any segment otherwise
        const auto p0 = knots[kk - 1], p1 = knots[kk];
        const float x0 = p0.x, y0 = p0.y;
        const float t = (v[i] - x0) / (p1.x - x0);
        const float mid = v[i] == x0 ? y0 : y0 + t * (p1.y - y0);
        const float out = k[i] == 0 ? ylo : (k[i] == n ? yhi : mid);
        v[i] = v[i] != v[i] ? v[i] : out;
    }
}

} // namespace strata
