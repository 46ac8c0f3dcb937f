/*
 * strata_oracle.h — CPU oracle for the raster stencil hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar C restatement of the
 * reference's canonical (scalar) kernels and of its engine's per-cell
 * edge and validity rules. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or
 * the timed CPU baseline; the product library never links it.
 *
 * Every function follows the reference file:line cited in
 * strata_oracle.c. Arithmetic is IEEE binary32 in the reference's order,
 * built with -ffp-contract=off so no multiply-add is fused, which is what
 * Go guarantees for the explicit float32(...) conversions it wraps every
 * product in.
 */
#ifndef STRATA_ORACLE_H
#define STRATA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors raster.Float32Raster (raster/raster.go:41-54). valid == NULL
 * means every cell is valid; bit valid_offset + y*stride + x (LSB first)
 * is the validity of cell (x, y). */
typedef struct orc_raster {
    float *data;
    int64_t width, height, stride;
    uint64_t *valid;
    int64_t valid_offset;
} orc_raster;

/* Operation ids: the same numbering the product's C ABI uses. */
enum {
    ORC_OP_GRADIENT = 1,
    ORC_OP_SLOPE = 2,
    ORC_OP_ASPECT = 3,
    ORC_OP_HILLSHADE = 4,
    ORC_OP_CURVATURE = 5,
    ORC_OP_CORRELATE = 10,
    ORC_OP_SEPARABLE = 11,
    ORC_OP_MEAN = 12,
    ORC_OP_MIN = 13,
    ORC_OP_MAX = 14,
    ORC_OP_LOOKUP = 20,
    ORC_OP_RECLASS = 21,
};

/* flags */
#define ORC_FLAG_ATAN 0x1        /* slope: take the arctangent */
#define ORC_FLAG_TRIG 0x2        /* aspect: trigonometric angle */
#define ORC_FLAG_FIT 0x100       /* terrain op through Wood's fit, radius = FitRadius */
#define ORC_CURV_SHIFT 4         /* curvature kind in bits 4..5 */

typedef struct orc_kernel {
    int32_t op;
    int32_t radius;
    int32_t flags;
    int32_t ntaps;
    float k[16];
    const float *taps;
} orc_kernel;

float orc_atan32(float x);
float orc_atan2f32(float y, float x);

/* The whole ProcessN result, cell by cell: the kernel over interior
 * cells, NaN and cleared bits in the edge ring, and output validity as
 * the AND of every masked input's (2r+1)^2 box. Returns 0, or -1 on a
 * bad descriptor. */
int orc_process(const orc_kernel *k, orc_raster *dst, int ndst, const orc_raster *src, int nsrc);

/* Same, for the output rows [y0, y1) only, with nthreads worker threads
 * splitting the rows; used to time the CPU baseline on a bounded
 * sample. */
int orc_process_rows(const orc_kernel *k, orc_raster *dst, int ndst, const orc_raster *src, int nsrc,
                     int64_t y0, int64_t y1, int nthreads);

/* The CPU baseline: the reference's AVX2 Tiled path for radius-1
 * Gradient, Slope and Hillshade (no fit), its band plan for tile_w x
 * tile_h tiles (0: the raster's size) claimed in order by nthreads
 * threads. Same results as orc_process. Returns 0, -1 on a bad call, -2
 * for an op it does not cover, -3 without AVX2. */
int orc_tiled_avx2(const orc_kernel *k, orc_raster *dst, int ndst, const orc_raster *src, int nsrc, int64_t tile_w,
                   int64_t tile_h, int nthreads);

/* FP64 correlation used as the accuracy yardstick for approximate paths:
 * dst[i] = sum over the (2r+1)^2 window of w*v in double. */
int orc_correlate_f64(double *dst, int64_t dst_stride, const orc_raster *src, const float *w, int r);

#ifdef __cplusplus
}
#endif
#endif
