/*
 * strata_oracle.c — scalar CPU restatement of the reference's hot path.
 *
 * TEST INFRASTRUCTURE: the checker for the HIP library and the timed CPU
 * baseline. Never linked into the product.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (see oracle/Makefile).
 * x86-64 evaluates float expressions in binary32 (FLT_EVAL_METHOD 0), so
 * each operator below rounds once, as a Go float32 operation does.
 *
 * Parity pins (see tests/test_oracle_*.py): the reference's known-answer
 * tests (tilted planes, Esri worked examples, flat cells, aspect axes,
 * focal hand-computed 3x3, transfer examples), Atan32/Atan2F32 accuracy
 * bounds and special-case table, brute-force validity, and the
 * mathematics (closed forms on planes and quadrics).
 */
#include "strata_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* The edge value every kernel of the path writes: float32(math.NaN()),
 * which is the quiet NaN 0x7fc00000 (internal/exec/tile.go setup,
 * terrain/stencil.go window3.Edge). */
static inline float edge_nan(void) { return u2f(0x7fc00000u); }

/* ---------------------------------------------------------------------
 * Arctangents — internal/stencil/horn.go:224-262, aspect.go:8-56.
 * Cephes atanf (S. L. Moshier): reduce around 0, pi/4, pi/2 at tan(pi/8)
 * and tan(3pi/8), then ((c4 z - c3) z + c2) z - c1.
 * ------------------------------------------------------------------- */
#define ATAN_TAN3PI8 2.414213562373095f   /* float32(2.414213562373095) */
#define ATAN_TANPI8 0.4142135623730950f   /* float32(0.4142135623730950) */
#define ATAN_C4 8.05374449538e-2f
#define ATAN_C3 1.38776856032e-1f
#define ATAN_C2 1.99777106478e-1f
#define ATAN_C1 3.33329491539e-1f
/* float32(pi/2), float32(pi/4), float32(pi), float32(180/pi): the
 * nearest binary32 values to the exact constants (checked in tests). */
#define ATAN_PI2_BITS 0x3fc90fdbu
#define ATAN_PI4_BITS 0x3f490fdbu
#define ATAN_PI_BITS 0x40490fdbu
#define RAD_TO_DEG_BITS 0x42652ee1u

float orc_atan32(float x) /* horn.go:250-262 */
{
    float num = x, den = 1.0f, y0 = 0.0f;
    if (x > ATAN_TANPI8) {
        num = x - 1.0f;
        den = x + 1.0f;
        y0 = u2f(ATAN_PI4_BITS);
    }
    if (x > ATAN_TAN3PI8) {
        num = -1.0f;
        den = x;
        y0 = u2f(ATAN_PI2_BITS);
    }
    float t = num / den;
    float z = t * t;
    float p = ATAN_C4 * z - ATAN_C3;
    p = p * z + ATAN_C2;
    p = p * z - ATAN_C1;
    return y0 + ((p * z) * t + t);
}

float orc_atan2f32(float y, float x) /* aspect.go:36-48 */
{
    float ax = u2f(f2u(x) & 0x7fffffffu);
    float ay = u2f(f2u(y) & 0x7fffffffu);
    float a = orc_atan32(ay / ax);
    if (ay == 0.0f && ax == 0.0f) a = 0.0f;
    if (ay == ax && ay > 3.40282346638528859811704183484516925440e+38f) a = u2f(ATAN_PI4_BITS);
    if (f2u(x) & 0x80000000u) a = u2f(ATAN_PI_BITS) - a;
    return u2f(f2u(a) | (f2u(y) & 0x80000000u));
}

/* ---------------------------------------------------------------------
 * Per-cell terrain arithmetic.
 * ------------------------------------------------------------------- */
/* horn.go:172-180: opposite cells first, then the doubled centre pair. */
static inline float horn_dx(float z1, float z3, float z4, float z6, float z7, float z9)
{
    float d = z6 - z4;
    return ((z3 - z1) + (z9 - z7)) + (d + d);
}
static inline float horn_dy(float z1, float z2, float z3, float z7, float z8, float z9)
{
    float d = z8 - z2;
    return ((z7 - z1) + (z9 - z3)) + (d + d);
}
/* horn.go:226-228: float32(sqrt(float64(float32(gx*gx)+float32(gy*gy)))) */
static inline float magnitude(float gx, float gy)
{
    return (float)sqrt((double)(gx * gx + gy * gy));
}
/* gradrow.go:44-58 and horn.go:200-222: slope of a scaled gradient. */
static inline float slope_of(float gx, float gy, float scale, int atan_)
{
    if (atan_) {
        float m = magnitude(gx, gy);
        return orc_atan32(m) * scale;
    }
    return magnitude(gx, gy) * scale;
}
/* aspect.go:104-123 */
static inline float aspect_of(float gx, float gy, float flat, int trig)
{
    float y, x;
    if (trig) { y = gy; x = 0.0f - gx; }
    else { y = 0.0f - gx; x = gy; }
    float d = orc_atan2f32(y, x) * u2f(RAD_TO_DEG_BITS);
    float off = 0.0f;
    if (d < 0.0f) off = 360.0f;
    d += off;
    if (d >= 360.0f) d = 0.0f;
    if (y == 0.0f && x == 0.0f) d = flat;
    return d;
}
/* aspect.go:161-173 */
static inline float shade_of(float gx, float gy, float c, float bx, float by)
{
    float num = c + (bx * gx + by * gy);
    float den = (float)sqrt((double)(1.0f + (gx * gx + gy * gy)));
    float v = num / den;
    if (v < 0.0f) v = 0.0f;
    if (v > 255.0f) v = 255.0f;
    return v;
}
static inline float sqrt32(float x) { return (float)sqrt((double)x); }
/* curvature.go:94-129 and 138-183 (the from-derivatives half). */
static inline float curvature_of(float p, float q, float r, float s, float t, int kind)
{
    float p2 = p * p, q2 = q * q;
    float pq = p * q;
    if (kind == 0) {
        float g = p2 + q2;
        float w = 1.0f + g;
        float num = (p2 * r + (pq + pq) * s) + q2 * t;
        float v = 0.0f - num / g / (w * sqrt32(w));
        if (g == 0.0f) v = num + 0.0f;
        return v;
    }
    if (kind == 1) {
        float g = p2 + q2;
        float num = (q2 * r - (pq + pq) * s) + p2 * t;
        float v = 0.0f - num / g / sqrt32(g);
        if (g == 0.0f) v = num + 0.0f;
        return v;
    }
    float w = 1.0f + (p2 + q2);
    float num = ((1.0f + q2) * r - (pq + pq) * s) + (1.0f + p2) * t;
    float d = w * sqrt32(w);
    return 0.0f - num / (d + d);
}

/* Go's builtin min/max on float32: NaN wins, -0 < +0
 * (focal/doc.go:49-51, internal/focalrow/focalrow.go:167-211). */
static inline float go_min(float a, float b)
{
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return (f2u(a) & 0x80000000u) ? a : b;
    return a < b ? a : b;
}
static inline float go_max(float a, float b)
{
    if (a != a) return a;
    if (b != b) return b;
    if (a == b) return (f2u(a) & 0x80000000u) ? b : a;
    return a > b ? a : b;
}

/* ---------------------------------------------------------------------
 * Per-cell kernels. win(c, j) is the source cell at column offset c and
 * row offset j in [0, 2r], top-left of the window at (0, 0).
 * ------------------------------------------------------------------- */
typedef struct {
    const float *base; /* cell (x-r, y-r) */
    int64_t stride;
} window_t;
#define WIN(w, c, j) ((w).base[(int64_t)(j) * (w).stride + (c)])

/* terrain/stencil.go + horn.go:160-170, slope.go:102-107,
 * aspect.go:125-130, hillshade.go:124-129, curvature.go:148-153.
 * out[0..1] receive the values; returns the number written. */
static void horn_cell(const orc_kernel *k, window_t w, float *out)
{
    float z1 = WIN(w, 0, 0), z2 = WIN(w, 1, 0), z3 = WIN(w, 2, 0);
    float z4 = WIN(w, 0, 1), z5 = WIN(w, 1, 1), z6 = WIN(w, 2, 1);
    float z7 = WIN(w, 0, 2), z8 = WIN(w, 1, 2), z9 = WIN(w, 2, 2);
    if (k->op == ORC_OP_CURVATURE) {
        /* curvature.go:79-88 ztDerivs */
        float kp = k->k[0], kq = k->k[1], kr = k->k[2], kt = k->k[3], ks = k->k[4];
        float p = (z6 - z4) * kp;
        float q = (z8 - z2) * kq;
        float c = z5 + z5;
        float r = ((z4 + z6) - c) * kr;
        float t = ((z2 + z8) - c) * kt;
        float s = ((z1 + z9) - (z3 + z7)) * ks;
        out[0] = curvature_of(p, q, r, s, t, (k->flags >> ORC_CURV_SHIFT) & 3);
        return;
    }
    float gx = horn_dx(z1, z3, z4, z6, z7, z9) * k->k[0];
    float gy = horn_dy(z1, z2, z3, z7, z8, z9) * k->k[1];
    switch (k->op) {
    case ORC_OP_GRADIENT:
        out[0] = gx;
        out[1] = gy;
        break;
    case ORC_OP_SLOPE:
        out[0] = slope_of(gx, gy, k->k[2], k->flags & ORC_FLAG_ATAN);
        break;
    case ORC_OP_ASPECT:
        out[0] = aspect_of(gx, gy, k->k[2], (k->flags & ORC_FLAG_TRIG) != 0);
        break;
    case ORC_OP_HILLSHADE:
        out[0] = shade_of(gx, gy, k->k[2], k->k[3], k->k[4]);
        break;
    }
}

/* terrain/fit.go:113-182: Wood's quadratic fit through the focalrow
 * column and row passes, per cell. The taps: ti = i, t2 = 3i^2 - r(r+1),
 * ones, for i in -r..r (fit.go:42-51). */
static void fit_cell(const orc_kernel *k, window_t w, float *out)
{
    int r = k->radius, n = 2 * r + 1;
    float ti[17], t2[17];
    float cs[17], cj[17], ct[17];
    for (int i = -r; i <= r; i++) {
        ti[i + r] = (float)i;
        t2[i + r] = (float)(3 * i * i - r * (r + 1));
    }
    int curv = k->op == ORC_OP_CURVATURE;
    for (int c = 0; c < n; c++) {
        /* ColumnSum: copy row 0, then add rows 1..n-1 (focalrow.go:244-250) */
        float s = WIN(w, c, 0);
        for (int j = 1; j < n; j++) s += WIN(w, c, j);
        cs[c] = s;
        /* ColumnCorrelate with ti (focalrow.go:232-242) */
        float a = ti[0] * WIN(w, c, 0);
        for (int j = 1; j < n; j++) a = a + ti[j] * WIN(w, c, j);
        cj[c] = a;
        if (curv) {
            float b = t2[0] * WIN(w, c, 0);
            for (int j = 1; j < n; j++) b = b + t2[j] * WIN(w, c, j);
            ct[c] = b;
        }
    }
    /* RowCorrelate (focalrow.go:275-284), then scaleRow (fit.go:184-190) */
    float p = ti[0] * cs[0];
    for (int c = 1; c < n; c++) p = p + ti[c] * cs[c];
    float q = 1.0f * cj[0];
    for (int c = 1; c < n; c++) q = q + 1.0f * cj[c];
    p = p * k->k[0];
    q = q * k->k[1];
    switch (k->op) {
    case ORC_OP_GRADIENT:
        out[0] = p;
        out[1] = q;
        return;
    case ORC_OP_SLOPE:
        out[0] = slope_of(p, q, k->k[5], k->flags & ORC_FLAG_ATAN);
        return;
    case ORC_OP_ASPECT:
        out[0] = aspect_of(p, q, k->k[5], (k->flags & ORC_FLAG_TRIG) != 0);
        return;
    case ORC_OP_HILLSHADE:
        out[0] = shade_of(p, q, k->k[5], k->k[6], k->k[7]);
        return;
    default: {
        float rr = t2[0] * cs[0];
        for (int c = 1; c < n; c++) rr = rr + t2[c] * cs[c];
        float tt = 1.0f * ct[0];
        for (int c = 1; c < n; c++) tt = tt + 1.0f * ct[c];
        float ss = ti[0] * cj[0];
        for (int c = 1; c < n; c++) ss = ss + ti[c] * cj[c];
        rr = rr * k->k[2];
        tt = tt * k->k[3];
        ss = ss * k->k[4];
        out[0] = curvature_of(p, q, rr, ss, tt, (k->flags >> ORC_CURV_SHIFT) & 3);
    }
    }
}

/* focal: correlate.go:95-100 -> focalrow.go:217-230 (row-major terms,
 * first product then acc + product); separable.go:104-111 -> column
 * then row pass; box.go:97-116 -> ColumnSum/RowMean, ColumnMin/RowMin,
 * ColumnMax/RowMax. */
static float focal_cell(const orc_kernel *k, window_t w)
{
    int r = k->radius, n = 2 * r + 1;
    switch (k->op) {
    case ORC_OP_CORRELATE: {
        const float *wt = k->taps;
        float acc = wt[0] * WIN(w, 0, 0);
        for (int j = 0; j < n; j++)
            for (int c = 0; c < n; c++) {
                if (j == 0 && c == 0) continue;
                acc = acc + wt[j * n + c] * WIN(w, c, j);
            }
        return acc;
    }
    case ORC_OP_SEPARABLE: {
        const float *row = k->taps, *col = k->taps + n;
        float cv[17];
        for (int c = 0; c < n; c++) {
            float a = col[0] * WIN(w, c, 0);
            for (int j = 1; j < n; j++) a = a + col[j] * WIN(w, c, j);
            cv[c] = a;
        }
        float acc = row[0] * cv[0];
        for (int c = 1; c < n; c++) acc = acc + row[c] * cv[c];
        return acc;
    }
    case ORC_OP_MEAN: {
        float cv[17];
        for (int c = 0; c < n; c++) {
            float a = WIN(w, c, 0);
            for (int j = 1; j < n; j++) a += WIN(w, c, j);
            cv[c] = a;
        }
        float acc = cv[0];
        for (int c = 1; c < n; c++) acc += cv[c];
        return acc / (float)(n * n);
    }
    case ORC_OP_MIN:
    case ORC_OP_MAX: {
        int mn = k->op == ORC_OP_MIN;
        float cv[17];
        for (int c = 0; c < n; c++) {
            float a = WIN(w, c, 0);
            for (int j = 1; j < n; j++) a = mn ? go_min(a, WIN(w, c, j)) : go_max(a, WIN(w, c, j));
            cv[c] = a;
        }
        float acc = cv[0];
        for (int c = 1; c < n; c++) acc = mn ? go_min(acc, cv[c]) : go_max(acc, cv[c]);
        return acc;
    }
    }
    return edge_nan();
}

/* internal/curve/scalar.go:39-56 */
static float reclass_cell(float v, const float *breaks, int nb, const float *values)
{
    if (v != v) return v;
    int kk = nb;
    for (int j = 0; j < nb; j++)
        if (v < breaks[j]) { kk = j; break; }
    return values[kk];
}
/* internal/curve/scalar.go:84-114 */
static float lookup_cell(float v, const float *xs, const float *ys, int n)
{
    if (v != v) return v;
    int kk = n;
    for (int j = 0; j < n; j++)
        if (v < xs[j]) { kk = j; break; }
    if (kk == 0) return ys[0];
    if (kk == n) return ys[n - 1];
    float x0 = xs[kk - 1], y0 = ys[kk - 1];
    if (v == x0) return y0;
    float t = (v - x0) / (xs[kk] - x0);
    return y0 + t * (ys[kk] - y0);
}

/* ---------------------------------------------------------------------
 * Engine semantics, per cell (internal/exec/doc.go:46-108,
 * internal/exec/halo.go:21-290): the kernel over cells at least r from
 * the edge of the rasters passed in, the edge value (NaN) over the ring,
 * output bit = not edge AND every masked input's (2r+1)^2 box valid;
 * with no masked input a masked output gets its interior set and ring
 * cleared. Only the outputs' own cells and bits are written.
 * ------------------------------------------------------------------- */
static inline int get_bit(const uint64_t *m, int64_t i) { return (int)((m[i >> 6] >> (i & 63)) & 1u); }
/* Atomic, so threads owning neighbouring rows may share a word. */
static inline void put_bit(uint64_t *m, int64_t i, int v)
{
    uint64_t b = (uint64_t)1 << (i & 63);
    if (v) __atomic_fetch_or(&m[i >> 6], b, __ATOMIC_RELAXED);
    else __atomic_fetch_and(&m[i >> 6], ~b, __ATOMIC_RELAXED);
}

static int n_outputs(const orc_kernel *k) { return k->op == ORC_OP_GRADIENT ? 2 : 1; }

static int check_kernel(const orc_kernel *k, int ndst, int nsrc)
{
    if (nsrc != 1 || ndst != n_outputs(k)) return -1;
    switch (k->op) {
    case ORC_OP_GRADIENT: case ORC_OP_SLOPE: case ORC_OP_ASPECT:
    case ORC_OP_HILLSHADE: case ORC_OP_CURVATURE:
        if (k->flags & ORC_FLAG_FIT) return (k->radius >= 1 && k->radius <= 8) ? 0 : -1;
        return k->radius == 1 ? 0 : -1;
    case ORC_OP_CORRELATE: case ORC_OP_SEPARABLE: case ORC_OP_MEAN:
    case ORC_OP_MIN: case ORC_OP_MAX:
        return (k->radius >= 1 && k->radius <= 8) ? 0 : -1;
    case ORC_OP_LOOKUP: case ORC_OP_RECLASS:
        return k->radius == 0 ? 0 : -1;
    }
    return -1;
}

typedef struct {
    const orc_kernel *k;
    orc_raster *dst;
    int ndst;
    const orc_raster *src;
    int64_t y0, y1;
} rows_job;

/* Data of the rows; validity_rows writes their bits. */
static void data_rows(const rows_job *j)
{
    const orc_kernel *k = j->k;
    const orc_raster *s = &j->src[0];
    int64_t W = s->width, H = s->height;
    int r = k->radius;
    int fit = (k->flags & ORC_FLAG_FIT) != 0;
    for (int64_t y = j->y0; y < j->y1; y++) {
        for (int64_t x = 0; x < W; x++) {
            int edge = x < r || y < r || x >= W - r || y >= H - r;
            float out[2] = {edge_nan(), edge_nan()};
            if (!edge) {
                if (r == 0) {
                    float v = s->data[y * s->stride + x];
                    if (k->op == ORC_OP_LOOKUP) {
                        int n = k->ntaps / 2;
                        out[0] = lookup_cell(v, k->taps, k->taps + n, n);
                    } else {
                        int nb = (k->ntaps - 1) / 2;
                        out[0] = reclass_cell(v, k->taps, nb, k->taps + nb);
                    }
                } else {
                    window_t w = {s->data + (y - r) * s->stride + (x - r), s->stride};
                    if (k->op >= ORC_OP_CORRELATE) out[0] = focal_cell(k, w);
                    else if (fit) fit_cell(k, w, out);
                    else horn_cell(k, w, out);
                }
            }
            for (int o = 0; o < j->ndst; o++) j->dst[o].data[y * j->dst[o].stride + x] = out[o];
        }
    }
}

static void validity_rows(const rows_job *j)
{
    const orc_kernel *k = j->k;
    const orc_raster *s = &j->src[0];
    int64_t W = s->width, H = s->height;
    int r = k->radius;
    for (int o = 0; o < j->ndst; o++) {
        orc_raster *d = &j->dst[o];
        if (!d->valid) continue;
        for (int64_t y = j->y0; y < j->y1; y++)
            for (int64_t x = 0; x < W; x++) {
                int edge = x < r || y < r || x >= W - r || y >= H - r;
                int v = !edge;
                if (v && s->valid) {
                    for (int dy = -r; dy <= r && v; dy++)
                        for (int dx = -r; dx <= r; dx++)
                            if (!get_bit(s->valid, s->valid_offset + (y + dy) * s->stride + (x + dx))) {
                                v = 0;
                                break;
                            }
                }
                put_bit(d->valid, d->valid_offset + y * d->stride + x, v);
            }
    }
}

static void *data_thread(void *arg)
{
    data_rows((const rows_job *)arg);
    validity_rows((const rows_job *)arg);
    return NULL;
}

int orc_process_rows(const orc_kernel *k, orc_raster *dst, int ndst, const orc_raster *src, int nsrc,
                     int64_t y0, int64_t y1, int nthreads)
{
    if (!k || !dst || !src || check_kernel(k, ndst, nsrc) != 0) return -1;
    for (int o = 0; o < ndst; o++)
        if (dst[o].width != src[0].width || dst[o].height != src[0].height) return -1;
    if (src[0].valid)
        for (int o = 0; o < ndst; o++)
            if (!dst[o].valid) return -1;
    if (y0 < 0) y0 = 0;
    if (y1 > src[0].height) y1 = src[0].height;
    if (y1 <= y0) return 0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads == 1) {
        rows_job j = {k, dst, ndst, src, y0, y1};
        data_thread(&j);
    } else {
        pthread_t th[256];
        rows_job jobs[256];
        if (nthreads > 256) nthreads = 256;
        int64_t n = y1 - y0;
        int used = 0;
        for (int t = 0; t < nthreads; t++) {
            int64_t a = y0 + n * t / nthreads, b = y0 + n * (t + 1) / nthreads;
            if (b <= a) continue;
            jobs[used] = (rows_job){k, dst, ndst, src, a, b};
            if (pthread_create(&th[used], NULL, data_thread, &jobs[used]) != 0) {
                data_thread(&jobs[used]);
                continue;
            }
            used++;
        }
        for (int t = 0; t < used; t++) pthread_join(th[t], NULL);
    }
    return 0;
}

int orc_process(const orc_kernel *k, orc_raster *dst, int ndst, const orc_raster *src, int nsrc)
{
    if (!src || nsrc < 1) return -1;
    return orc_process_rows(k, dst, ndst, src, nsrc, 0, src[0].height, 1);
}

int orc_correlate_f64(double *dst, int64_t dst_stride, const orc_raster *s, const float *w, int r)
{
    int n = 2 * r + 1;
    int64_t W = s->width, H = s->height;
    for (int64_t y = 0; y < H; y++)
        for (int64_t x = 0; x < W; x++) {
            if (x < r || y < r || x >= W - r || y >= H - r) {
                dst[y * dst_stride + x] = (double)edge_nan();
                continue;
            }
            double acc = 0.0;
            for (int j = 0; j < n; j++)
                for (int c = 0; c < n; c++)
                    acc += (double)w[j * n + c] * (double)s->data[(y - r + j) * s->stride + (x - r + c)];
            dst[y * dst_stride + x] = acc;
        }
    return 0;
}

/* =====================================================================
 * CPU baseline: the reference's AVX2 Tiled path, restated.
 *
 * Lanes (8 cells per 256-bit vector), internal/stencil/simd_amd64.go:
 *   hornDiff8 :74-84, hornGradientLanes :92-106, hornMagnitude8 and
 *   magnitude8 :119-128, hornSlopeAtanLanes :136-148, atan8 :179-192
 *   (both reductions computed, selected in Atan32's order),
 *   hornHillshadeLanes and shade8 :284-305; each lane runs the scalar
 *   formula operation for operation (IEEE sqrt and divide, no fused
 *   multiply-add), the row's last n % 8 cells run the scalar kernel.
 * Schedule: the band plan of internal/exec/tile.go:298-359 (tiles row
 *   major, bands of max(1, 65536/tileW) rows), claimed in order by
 *   worker threads from one counter (internal/exec/worker.go:144-181).
 * Validity, word by word as internal/stencil/mask.go:55-160 (ErodeBox at
 *   radius 1: the AND of the three source rows, shrunk by two cells),
 *   over the band's interior; the engine's edge ring gets cleared bits
 *   (internal/exec/halo.go:161-290). Words a band shares with another
 *   band are updated atomically on its own bits.
 * Radius-1 Gradient, Slope and Hillshade. The results equal the per-cell
 * functions above bit for bit (tests/test_oracle_avx2.py).
 * ===================================================================== */
#include <immintrin.h>

#define ORC_AVX2 __attribute__((target("avx2")))

ORC_AVX2 static inline __m256 atan8(__m256 x)
{
    const __m256 one = _mm256_set1_ps(1.0f);
    const __m256 mid = _mm256_cmp_ps(x, _mm256_set1_ps(ATAN_TANPI8), _CMP_GT_OQ);
    const __m256 big = _mm256_cmp_ps(x, _mm256_set1_ps(ATAN_TAN3PI8), _CMP_GT_OQ);
    __m256 num = _mm256_blendv_ps(x, _mm256_sub_ps(x, one), mid);
    num = _mm256_blendv_ps(num, _mm256_set1_ps(-1.0f), big);
    __m256 den = _mm256_blendv_ps(one, _mm256_add_ps(x, one), mid);
    den = _mm256_blendv_ps(den, x, big);
    __m256 y0 = _mm256_blendv_ps(_mm256_setzero_ps(), _mm256_set1_ps(u2f(ATAN_PI4_BITS)), mid);
    y0 = _mm256_blendv_ps(y0, _mm256_set1_ps(u2f(ATAN_PI2_BITS)), big);
    const __m256 t = _mm256_div_ps(num, den);
    const __m256 z = _mm256_mul_ps(t, t);
    __m256 p = _mm256_sub_ps(_mm256_mul_ps(_mm256_set1_ps(ATAN_C4), z), _mm256_set1_ps(ATAN_C3));
    p = _mm256_add_ps(_mm256_mul_ps(p, z), _mm256_set1_ps(ATAN_C2));
    p = _mm256_sub_ps(_mm256_mul_ps(p, z), _mm256_set1_ps(ATAN_C1));
    return _mm256_add_ps(y0, _mm256_add_ps(_mm256_mul_ps(_mm256_mul_ps(p, z), t), t));
}

/* Horn's unscaled differences of eight adjacent cells; rows r0..r2 start
 * at the column left of the first cell. */
ORC_AVX2 static inline void horn_diff8(const float *r0, const float *r1, const float *r2, __m256 *dx, __m256 *dy)
{
    const __m256 z1 = _mm256_loadu_ps(r0), z2 = _mm256_loadu_ps(r0 + 1), z3 = _mm256_loadu_ps(r0 + 2);
    const __m256 z4 = _mm256_loadu_ps(r1), z6 = _mm256_loadu_ps(r1 + 2);
    const __m256 z7 = _mm256_loadu_ps(r2), z8 = _mm256_loadu_ps(r2 + 1), z9 = _mm256_loadu_ps(r2 + 2);
    __m256 d = _mm256_sub_ps(z6, z4);
    *dx = _mm256_add_ps(_mm256_add_ps(_mm256_sub_ps(z3, z1), _mm256_sub_ps(z9, z7)), _mm256_add_ps(d, d));
    d = _mm256_sub_ps(z8, z2);
    *dy = _mm256_add_ps(_mm256_add_ps(_mm256_sub_ps(z7, z1), _mm256_sub_ps(z9, z3)), _mm256_add_ps(d, d));
}

/* The op over cells [0, n) of a row (window rows r0..r2 from the column
 * left of cell 0) in whole lanes; returns the cells written. */
ORC_AVX2 static int64_t horn_lanes(const orc_kernel *k, float *o0, float *o1, const float *r0, const float *r1,
                                   const float *r2, int64_t n)
{
    const __m256 kx = _mm256_set1_ps(k->k[0]), ky = _mm256_set1_ps(k->k[1]);
    int64_t i = 0;
    __m256 dx, dy;
    switch (k->op) {
    case ORC_OP_GRADIENT:
        for (; i + 8 <= n; i += 8) {
            horn_diff8(r0 + i, r1 + i, r2 + i, &dx, &dy);
            _mm256_storeu_ps(o0 + i, _mm256_mul_ps(dx, kx));
            _mm256_storeu_ps(o1 + i, _mm256_mul_ps(dy, ky));
        }
        break;
    case ORC_OP_SLOPE: {
        const __m256 scale = _mm256_set1_ps(k->k[2]);
        const int atan_ = (k->flags & ORC_FLAG_ATAN) != 0;
        for (; i + 8 <= n; i += 8) {
            horn_diff8(r0 + i, r1 + i, r2 + i, &dx, &dy);
            const __m256 gx = _mm256_mul_ps(dx, kx), gy = _mm256_mul_ps(dy, ky);
            __m256 m = _mm256_sqrt_ps(_mm256_add_ps(_mm256_mul_ps(gx, gx), _mm256_mul_ps(gy, gy)));
            if (atan_) m = atan8(m);
            _mm256_storeu_ps(o0 + i, _mm256_mul_ps(m, scale));
        }
        break;
    }
    case ORC_OP_HILLSHADE: {
        const __m256 c = _mm256_set1_ps(k->k[2]), bx = _mm256_set1_ps(k->k[3]), by = _mm256_set1_ps(k->k[4]);
        const __m256 zero = _mm256_setzero_ps(), hi = _mm256_set1_ps(255.0f), one = _mm256_set1_ps(1.0f);
        for (; i + 8 <= n; i += 8) {
            horn_diff8(r0 + i, r1 + i, r2 + i, &dx, &dy);
            const __m256 gx = _mm256_mul_ps(dx, kx), gy = _mm256_mul_ps(dy, ky);
            const __m256 num = _mm256_add_ps(c, _mm256_add_ps(_mm256_mul_ps(bx, gx), _mm256_mul_ps(by, gy)));
            const __m256 den =
                _mm256_sqrt_ps(_mm256_add_ps(one, _mm256_add_ps(_mm256_mul_ps(gx, gx), _mm256_mul_ps(gy, gy))));
            __m256 v = _mm256_div_ps(num, den);
            v = _mm256_blendv_ps(v, zero, _mm256_cmp_ps(v, zero, _CMP_LT_OQ));
            v = _mm256_blendv_ps(v, hi, _mm256_cmp_ps(v, hi, _CMP_GT_OQ));
            _mm256_storeu_ps(o0 + i, v);
        }
        break;
    }
    }
    return i;
}

/* Word-level bit helpers of the baseline's validity (mask.go:183-303). */
static void bits_extract(uint64_t *acc, const uint64_t *m, int64_t from, int64_t n)
{
    const int64_t nw = (n + 63) >> 6, w0 = from >> 6, last = (from + n - 1) >> 6;
    const int sh = (int)(from & 63);
    for (int64_t i = 0; i < nw; i++) {
        uint64_t v = m[w0 + i] >> sh;
        if (sh && w0 + i + 1 <= last) v |= m[w0 + i + 1] << (64 - sh);
        acc[i] = v;
    }
}
static void bits_and(uint64_t *acc, const uint64_t *m, int64_t from, int64_t n)
{
    const int64_t nw = (n + 63) >> 6, w0 = from >> 6, last = (from + n - 1) >> 6;
    const int sh = (int)(from & 63);
    for (int64_t i = 0; i < nw; i++) {
        uint64_t v = m[w0 + i] >> sh;
        if (sh && w0 + i + 1 <= last) v |= m[w0 + i + 1] << (64 - sh);
        acc[i] &= v;
    }
}
/* Cell i becomes the AND of cells i, i+1, i+2 (shrink2Row). */
static void bits_shrink2(uint64_t *acc, int64_t nw)
{
    for (int64_t i = 0; i < nw; i++) {
        const uint64_t a = acc[i], b = i + 1 < nw ? acc[i + 1] : 0;
        acc[i] = a & ((a >> 1) | (b << 63)) & ((a >> 2) | (b << 62));
    }
}
/* Bits [to, to+n) of m from acc (acc == NULL: all set to v); words only
 * partly in the range are updated atomically. */
static void bits_deposit(uint64_t *m, int64_t to, const uint64_t *acc, int64_t n, int v)
{
    if (n <= 0) return;
    const int64_t w0 = to >> 6, w1 = (to + n - 1) >> 6, nsw = (n + 63) >> 6;
    for (int64_t w = w0; w <= w1; w++) {
        const int64_t lo = (to > 64 * w ? to : 64 * w) - 64 * w;
        const int64_t hi = (to + n < 64 * w + 64 ? to + n : 64 * w + 64) - 64 * w;
        const uint64_t own = hi - lo == 64 ? ~(uint64_t)0 : ((((uint64_t)1 << (hi - lo)) - 1) << lo);
        uint64_t bits;
        if (!acc) {
            bits = v ? ~(uint64_t)0 : 0;
        } else {
            const int64_t sb = 64 * w - to;
            if (sb < 0) {
                bits = acc[0] << (-sb);
            } else {
                const int64_t q = sb >> 6;
                const int s = (int)(sb & 63);
                bits = acc[q] >> s;
                if (s && q + 1 < nsw) bits |= acc[q + 1] << (64 - s);
            }
        }
        bits &= own;
        if (own == ~(uint64_t)0) {
            m[w] = bits;
        } else {
            __atomic_fetch_and(&m[w], ~own | bits, __ATOMIC_RELAXED);
            __atomic_fetch_or(&m[w], bits, __ATOMIC_RELAXED);
        }
    }
}

typedef struct {
    const orc_kernel *k;
    orc_raster *dst;
    int ndst;
    const orc_raster *src;
    int64_t tileW, tileH, bandRows, tilesX, tileRows, perTile, perLastTile, bands;
    int64_t next; /* claim counter */
} tiled_job;

/* plan.band (internal/exec/tile.go:343-359) */
static void tiled_band(const tiled_job *j, int64_t i, int64_t *x0, int64_t *y0, int64_t *x1, int64_t *y1)
{
    const int64_t W = j->src->width, H = j->src->height;
    const int64_t perRow = j->tilesX * j->perTile;
    int64_t tr = i / perRow, q = i % perRow, per = j->perTile;
    if (tr >= j->tileRows - 1) {
        tr = j->tileRows - 1;
        q = i - (j->tileRows - 1) * perRow;
        per = j->perLastTile;
    }
    const int64_t tx = q / per, b = q % per, ty = tr * j->tileH;
    *x0 = tx * j->tileW;
    *x1 = *x0 + j->tileW < W ? *x0 + j->tileW : W;
    *y0 = ty + b * j->bandRows;
    int64_t e = *y0 + j->bandRows;
    if (e > ty + j->tileH) e = ty + j->tileH;
    *y1 = e < H ? e : H;
}

ORC_AVX2 static void tiled_band_run(const tiled_job *j, int64_t x0, int64_t y0, int64_t x1, int64_t y1, uint64_t *acc)
{
    const orc_kernel *k = j->k;
    const orc_raster *s = j->src;
    const int64_t W = s->width, H = s->height;
    const int64_t ix0 = x0 > 1 ? x0 : 1, ix1 = x1 < W - 1 ? x1 : W - 1;
    const float nan = edge_nan();
    for (int64_t y = y0; y < y1; y++) {
        float *o0 = j->dst[0].data + y * j->dst[0].stride;
        float *o1 = j->ndst > 1 ? j->dst[1].data + y * j->dst[1].stride : NULL;
        const int row_in = y >= 1 && y < H - 1 && ix0 < ix1;
        for (int64_t x = x0; x < x1; x++) {
            if (row_in && x >= ix0 && x < ix1) {
                x = ix1 - 1;
                continue;
            }
            o0[x] = nan;
            if (o1) o1[x] = nan;
        }
        if (!row_in) continue;
        const float *r0 = s->data + (y - 1) * s->stride + (ix0 - 1);
        const float *r1 = r0 + s->stride, *r2 = r1 + s->stride;
        const int64_t n = ix1 - ix0;
        int64_t i = horn_lanes(k, o0 + ix0, o1 ? o1 + ix0 : NULL, r0, r1, r2, n);
        for (; i < n; i++) {
            window_t w = {r0 + i, s->stride};
            float out[2];
            horn_cell(k, w, out);
            o0[ix0 + i] = out[0];
            if (o1) o1[ix0 + i] = out[1];
        }
    }
    for (int o = 0; o < j->ndst; o++) {
        orc_raster *d = &j->dst[o];
        if (!d->valid) continue;
        for (int64_t y = y0; y < y1; y++) {
            const int64_t row = d->valid_offset + y * d->stride;
            if (y < 1 || y >= H - 1 || ix0 >= ix1) {
                bits_deposit(d->valid, row + x0, NULL, x1 - x0, 0);
                continue;
            }
            if (x0 < ix0) bits_deposit(d->valid, row + x0, NULL, ix0 - x0, 0);
            if (ix1 < x1) bits_deposit(d->valid, row + ix1, NULL, x1 - ix1, 0);
            const int64_t n = ix1 - ix0;
            if (!s->valid) {
                bits_deposit(d->valid, row + ix0, NULL, n, 1);
                continue;
            }
            const int64_t sb = s->valid_offset + (y - 1) * s->stride + (ix0 - 1);
            bits_extract(acc, s->valid, sb, n + 2);
            bits_and(acc, s->valid, sb + s->stride, n + 2);
            bits_and(acc, s->valid, sb + 2 * s->stride, n + 2);
            bits_shrink2(acc, (n + 2 + 63) >> 6);
            bits_deposit(d->valid, row + ix0, acc, n, 0);
        }
    }
}

static void *tiled_worker(void *arg)
{
    tiled_job *j = (tiled_job *)arg;
    uint64_t *acc = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)((j->src->width + 2 + 63) / 64 + 1));
    if (!acc) return NULL;
    for (;;) {
        const int64_t i = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
        if (i >= j->bands) break;
        int64_t x0, y0, x1, y1;
        tiled_band(j, i, &x0, &y0, &x1, &y1);
        tiled_band_run(j, x0, y0, x1, y1, acc);
    }
    free(acc);
    return NULL;
}

int orc_tiled_avx2(const orc_kernel *k, orc_raster *dst, int ndst, const orc_raster *src, int nsrc, int64_t tile_w,
                   int64_t tile_h, int nthreads)
{
    if (!k || !dst || !src || check_kernel(k, ndst, nsrc) != 0) return -1;
    if (k->radius != 1 || (k->flags & ORC_FLAG_FIT) ||
        (k->op != ORC_OP_GRADIENT && k->op != ORC_OP_SLOPE && k->op != ORC_OP_HILLSHADE))
        return -2;
    if (!__builtin_cpu_supports("avx2")) return -3;
    const int64_t W = src[0].width, H = src[0].height;
    for (int o = 0; o < ndst; o++)
        if (dst[o].width != W || dst[o].height != H) return -1;
    if (src[0].valid)
        for (int o = 0; o < ndst; o++)
            if (!dst[o].valid) return -1;
    tiled_job j = {k, dst, ndst, src, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    j.tileW = tile_w > 0 && tile_w < W ? tile_w : W;
    j.tileH = tile_h > 0 && tile_h < H ? tile_h : H;
    j.bandRows = 65536 / j.tileW > 1 ? 65536 / j.tileW : 1;
    j.tilesX = (W + j.tileW - 1) / j.tileW;
    j.tileRows = (H + j.tileH - 1) / j.tileH;
    j.perTile = (j.tileH + j.bandRows - 1) / j.bandRows;
    j.perLastTile = (H - (j.tileRows - 1) * j.tileH + j.bandRows - 1) / j.bandRows;
    j.bands = j.tilesX * ((j.tileRows - 1) * j.perTile + j.perLastTile);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    int used = 0;
    for (int t = 1; t < nthreads; t++)
        if (pthread_create(&th[used], NULL, tiled_worker, &j) == 0) used++;
    tiled_worker(&j);
    for (int t = 0; t < used; t++) pthread_join(th[t], NULL);
    return 0;
}
