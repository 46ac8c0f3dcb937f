/*
 * strata_hip.h — C ABI of the MI355X (gfx950) raster stencil library.
 *
 * The seam is the reference engine's two entry points, which every
 * terrain.*, focal.* and transfer.* function funnels into:
 *
 *   exec.ProcessN(ctx, dst, src []raster.Float32Raster, k Kernel, opts)
 *       internal/exec/process.go:28           -> strata_process
 *   exec.ProcessChunked(ctx, dst []RasterSink, src []RasterSource, k, opts)
 *       internal/exec/chunked.go:33           -> strata_process_chunked
 *
 * The Kernel value is replaced by a plain descriptor (strata_kernel) that
 * carries the constants the caller already resolved in float32
 * (terrain/stencil.go:36-47 cellSizes, terrain/slope.go:84-94
 * slopeScale, terrain/hillshade.go:94-117 hillshadeLight,
 * terrain/curvature.go:113-127, terrain/fit.go:36-70,
 * focal/correlate.go:70-88, focal/separable.go:53-89), so the library
 * never restates Go's math.Sin/Cos/Mod/Exp.
 *
 * Plain pointers and sizes only. Rasters may live in device memory
 * (STRATA_MEM_DEVICE, the fast path: inputs already resident in HBM) or
 * host memory (STRATA_MEM_HOST: staged through pinned buffers). Only the
 * outputs' own cells and validity bits are written (internal/exec/doc.go:
 * 107-108); row padding and a parent's other bits are left alone.
 *
 * Programming errors (the reference's panics) return STRATA_EINVAL
 * before anything is written; strata_last_error() gives the message,
 * prefixed like the reference's ("engine: ...").
 */
#ifndef STRATA_HIP_H
#define STRATA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
#define STRATA_OK 0
#define STRATA_CANCELLED 1   /* ctx.Err(): a prefix of the plan was written */
#define STRATA_EINVAL (-1)   /* programming error (the reference panics)  */
#define STRATA_EDEVICE (-2)  /* HIP runtime or device failure             */
#define STRATA_EIO (-3)      /* a source or sink callback failed          */
#define STRATA_ENOMEM (-4)   /* device or pinned allocation failed        */

/* ------------------------------------------------------------------ */
/* Rasters: mirrors raster.Float32Raster (raster/raster.go:41-54)      */
/* ------------------------------------------------------------------ */
#define STRATA_MEM_DEVICE 0
#define STRATA_MEM_HOST 1

typedef struct strata_raster {
    float *data;          /* cell (x, y) is data[y*stride + x]           */
    int64_t width;
    int64_t height;
    int64_t stride;       /* >= width                                    */
    uint64_t *valid;      /* NULL: every cell valid; else LSB-first bits  */
    int64_t valid_offset; /* bit index of data[0] in valid                */
    int32_t mem;          /* STRATA_MEM_DEVICE or STRATA_MEM_HOST         */
    int32_t reserved;
} strata_raster;

/* ------------------------------------------------------------------ */
/* Kernel descriptor: the reference's name-registered kernels          */
/* (internal/opkernel/opkernel.go:32-52, terrain/opkernel.go:14-43,    */
/* focal/opkernel.go:12-19)                                            */
/* ------------------------------------------------------------------ */
enum strata_op {
    STRATA_OP_GRADIENT = 1,   /* terrain.Gradient: 1 input, 2 outputs (dx, dy) */
    STRATA_OP_SLOPE = 2,      /* terrain.Slope                                 */
    STRATA_OP_ASPECT = 3,     /* terrain.Aspect                                */
    STRATA_OP_HILLSHADE = 4,  /* terrain.Hillshade                             */
    STRATA_OP_CURVATURE = 5,  /* terrain.Curvature                             */
    STRATA_OP_SURFACE = 6,    /* terrain.Surface: the STRATA_SURF_* products    */
                              /* of one gradient, outputs in that order        */
    STRATA_OP_CORRELATE = 10, /* focal.Correlate / Convolve (weights pre-rotated) */
    STRATA_OP_SEPARABLE = 11, /* focal.CorrelateSeparable                      */
    STRATA_OP_MEAN = 12,      /* focal.Mean                                    */
    STRATA_OP_MIN = 13,       /* focal.Min                                     */
    STRATA_OP_MAX = 14,       /* focal.Max                                     */
    STRATA_OP_LOOKUP = 20,    /* transfer.Lookup (radius 0)                    */
    STRATA_OP_RECLASS = 21,   /* transfer.Reclass (radius 0)                   */
};

#define STRATA_FLAG_ATAN 0x1   /* slope: atan(m)*scale (degrees, radians)     */
#define STRATA_FLAG_TRIG 0x2   /* aspect: trigonometric angle                 */
#define STRATA_FLAG_FIT 0x100  /* terrain op through Wood's quadratic fit,    */
                               /* radius = FitRadius (terrain/fit.go:187-226) */
#define STRATA_CURV_SHIFT 4    /* curvature kind (0 profile, 1 plan, 2 mean)  */
/* terrain.Surface products (terrain/surface.go:34-44), in output order */
#define STRATA_SURF_DX 0x1000
#define STRATA_SURF_DY 0x2000
#define STRATA_SURF_SLOPE 0x4000
#define STRATA_SURF_ASPECT 0x8000
#define STRATA_SURF_HILLSHADE 0x10000
#define STRATA_SURF_ALL 0x1f000
#define STRATA_FLAG_MFMA 0x200 /* Correlate: split-BF16x3 MFMA path (meant for */
                               /* r >= 4). Approximate: each product within    */
                               /* about 3*2^-18 of |w*x|; windows holding a    */
                               /* non-finite value get the exact FP32 result.  */

/*
 * k[] layout by op (all already-resolved float32 constants):
 *   Horn ops (radius 1):  k0 kx, k1 ky; slope k2 scale; aspect k2 flat;
 *                         hillshade k2 c, k3 bx, k4 by
 *   ZT curvature:         k0..k4 = kp, kq, kr, kt, ks
 *   FIT (any terrain op): k0..k4 = Wood factors kp, kq, kr, kt, ks;
 *                         slope k5 scale; aspect k5 flat;
 *                         hillshade k5 c, k6 bx, k7 by
 *   Surface (radius 1):   k0 kx, k1 ky, k2 slope scale, k3 aspect flat,
 *                         k4 c, k5 bx, k6 by (ATAN, TRIG as for the ops)
 *   Surface with FIT:     k0..k4 Wood factors, k5 slope scale, k6 aspect
 *                         flat, k7 c, k8 bx, k9 by
 * taps (host pointer, copied during the call):
 *   CORRELATE: (2r+1)^2 weights, row-major; SEPARABLE: row[2r+1] then
 *   col[2r+1]; LOOKUP: xs[n] then ys[n] (ntaps = 2n);
 *   RECLASS: breaks[n] then values[n+1] (ntaps = 2n+1).
 */
typedef struct strata_kernel {
    int32_t op;
    int32_t radius;
    int32_t flags;
    int32_t ntaps;
    float k[16];
    const float *taps;
} strata_kernel;

/* ------------------------------------------------------------------ */
/* Options: engine.Options (engine/engine.go:7-28) plus device plumbing */
/* ------------------------------------------------------------------ */
typedef struct strata_options {
    int64_t tile_width;  /* 0: raster width                              */
    int64_t tile_height; /* 0: raster height                             */
    int64_t workers;     /* scheduling hint only; never changes results  */
    void *stream;        /* hipStream_t to run on; NULL: the null stream  */
    const volatile int32_t *cancel; /* host flag, nonzero = ctx done;
                                       polled between units of the plan */
    int32_t async;       /* 1: return once enqueued (device operands,
                            cancel == NULL only); 0: return when done    */
    int32_t reserved;
} strata_options;

/* engine.Stats (engine/stats.go:77-104), field for field, in the
 * closed form the reference's band and tile loops produce. */
typedef struct strata_stats {
    int64_t cells;
    int64_t ideal;
    int64_t tiles;
    int64_t bands;
    int64_t source_read;
    int64_t sink_written;
    int64_t kernel_read;
    int64_t kernel_written;
} strata_stats;

/* ------------------------------------------------------------------ */
/* Library                                                             */
/* ------------------------------------------------------------------ */

/* Selects the HIP device and loads the kernels. Optional: every entry
 * point initialises the current device on first use. */
int strata_init(int device);

/* Message of the calling thread's last error, or "" after success. */
const char *strata_last_error(void);

/* Version string of the library build. */
const char *strata_version(void);

/* Returns the scratch that chunked calls keep for reuse (the device
 * buffers of fold-only sinks) to the device. No counterpart in the
 * reference. Not to be called while a call runs. */
int strata_trim(void);

/*
 * exec.ProcessN (internal/exec/process.go:28): run k over every cell of
 * dst[0..ndst) from src[0..nsrc). Same checks (process.go:37-135, minus
 * Go-slice overlap that a caller checks before crossing the boundary),
 * same edge policy (NaN ring of width radius), same validity rules
 * (internal/exec/halo.go:161-290) and the same Stats for every Options
 * value (internal/exec/halo.go:32-39,103-104, worker.go:59-68).
 * Cancellation: units of the plan (bands) are claimed in order and
 * always finished, so after STRATA_CANCELLED the written bands are a
 * prefix of the plan (internal/exec/doc.go:165-183).
 * Returns STRATA_OK, STRATA_CANCELLED, or a negative error. Stats, when
 * not NULL, are added to (engine/stats.go:63-70).
 */
int strata_process(const strata_kernel *k, const strata_raster *dst, int ndst, const strata_raster *src,
                   int nsrc, const strata_options *opts, strata_stats *stats);

/* ------------------------------------------------------------------ */
/* Chunked: engine.RasterSource / RasterSink (engine/source.go:19-71)  */
/* ------------------------------------------------------------------ */

/* ReadWindow (engine/source.go:40): fill the w x h cells of buf (host
 * memory, row stride `stride`, validity at bit valid_offset of valid when
 * valid != NULL) with the region at (x, y). Return 0, or nonzero for an
 * IO error. */
typedef int (*strata_read_fn)(void *user, float *buf, int64_t w, int64_t h, int64_t stride, uint64_t *valid,
                              int64_t valid_offset, int64_t x, int64_t y);
/* WriteWindow (engine/source.go:67). */
typedef int (*strata_write_fn)(void *user, const float *buf, int64_t w, int64_t h, int64_t stride,
                               const uint64_t *valid, int64_t valid_offset, int64_t x, int64_t y);

/* A source is either resident (a raster the library reads windows of
 * itself, engine.MemorySource: engine/source.go:76-107; in device memory
 * read in place, in host memory copied window by window into pinned tile
 * buffers by the library's IO threads), a raw file, or a callback reader.
 * A resident sink likewise (engine.MemorySink). */
/* engine.RawSource / RawSink (engine/raw.go:28-245) over a file the
 * library reads and writes itself: width*height little-endian float32
 * cells, row-major, from byte `offset`. With has_fill a source is Masked
 * and a cell equal to fill as a float (-0 == 0; a NaN fill matches any
 * NaN) is invalid, its validity computed on the device as tiles arrive; a
 * sink with has_fill gets fill in every invalid cell (engine/raw.go:
 * 247-273). Short reads and failed writes are STRATA_EIO. */
typedef struct strata_raw {
    int32_t fd;       /* open file descriptor; pread/pwrite only        */
    int32_t has_fill;
    float fill;
    int32_t reserved;
    int64_t offset;   /* byte offset of cell (0, 0)                     */
} strata_raw;

typedef struct strata_source {
    int64_t width, height;
    int32_t masked;          /* Masked() (a raw source: has_fill)        */
    int32_t reserved;
    strata_read_fn read;     /* used when resident and raw are NULL      */
    void *user;
    const strata_raster *resident;
    const strata_raw *raw;   /* used when resident is NULL               */
} strata_source;

/* A sink may also fold what it is given, as graph's teeSink does
 * (graph/sinks.go:17-71): with fold != NULL, every tile of the output is
 * folded on the device with the reduction fold_what (STRATA_REDUCE_*,
 * below) before the sink receives it, and the combined partial is stored
 * into *fold when the call completes (zeroed otherwise). Only the tile's
 * valid cells take part when the sink is masked, every cell (the NaN ring
 * included) when it is not: exec.Reduce over the window the sink is
 * handed. A sink with a fold and no resident raster, raw file or write
 * callback receives nothing else: statistics alone write no raster. */
struct strata_fold;

typedef struct strata_sink {
    int64_t width, height;
    int32_t masked;
    int32_t fold_what;       /* STRATA_REDUCE_* bits when fold != NULL   */
    strata_write_fn write;   /* used when resident and raw are NULL      */
    void *user;
    const strata_raster *resident;
    const strata_raw *raw;   /* used when resident is NULL               */
    struct strata_fold *fold;
} strata_sink;

/*
 * exec.ProcessChunked (internal/exec/chunked.go:33): tiles in row-major
 * order, each read with its halo clipped to the raster, computed exactly
 * as strata_process computes those cells, and written to every sink.
 * Device-resident sources and sinks form a device tile queue (no copies);
 * host-resident, raw and callback ones are staged through pinned host
 * buffers, a tile's copies overlapping the previous tile's sink writes. Tiles are claimed in order and finished
 * once claimed, so a cancelled or failed call leaves a prefix of whole
 * tiles (engine/doc.go:188-215). Stats as chunked.go:302-374.
 */
int strata_process_chunked(const strata_kernel *k, const strata_sink *dst, int ndst, const strata_source *src,
                           int nsrc, const strata_options *opts, strata_stats *stats);

/* Closed-form Stats of a call that completes (no device work): what
 * strata_process (chunked = 0) or strata_process_chunked (chunked = 1)
 * adds for a width x height raster. */
int strata_plan_stats(const strata_kernel *k, int ndst, int nsrc, int64_t width, int64_t height,
                      const strata_options *opts, int chunked, strata_stats *stats);

/* ------------------------------------------------------------------ */
/* Fan-out: one input, several single-output kernels, one pass         */
/* ------------------------------------------------------------------ */
/*
 * exec.ProcessN / exec.ProcessChunked with the kernel exec.NewPipeline(1,
 * stages, outs) where stage i is parts[i] reading the one input and
 * writing output i: the pipeline terrain.Features builds
 * (terrain/features.go:103-117, called at features.go:76-98). Output i is
 * bit for bit what parts[i] alone writes: its own edge ring (its radius)
 * and its validity eroded by its own radius (internal/exec/kernel.go:
 * 94-110, internal/exec/tile.go:155-184). The call's radius is the
 * largest part radius; Stats follow the reference's band loop with the
 * kernel run outside the narrowest ring (internal/exec/halo.go:21-57).
 * Radius-1 Horn ops (Slope, Aspect, Hillshade, Curvature without FIT)
 * share one read of the source. Each part must have one input and one
 * output; at most 16 parts. Errors as strata_process.
 */
int strata_process_fanout(const strata_kernel *parts, int nparts, const strata_raster *dst,
                          const strata_raster *src, const strata_options *opts, strata_stats *stats);
int strata_process_fanout_chunked(const strata_kernel *parts, int nparts, const strata_sink *dst,
                                  const strata_source *src, const strata_options *opts, strata_stats *stats);
/* Closed-form Stats of a fan-out call, as strata_plan_stats. */
int strata_plan_stats_fanout(const strata_kernel *parts, int nparts, int64_t width, int64_t height,
                             const strata_options *opts, int chunked, strata_stats *stats);

/* ------------------------------------------------------------------ */
/* Chains: a pipeline of stages, each reading the one before it        */
/* ------------------------------------------------------------------ */
/*
 * exec.ProcessN / exec.ProcessChunked with the kernel exec.NewPipeline(1,
 * stages, [last]) where stage i reads stage i-1's value and stage 0 the
 * input (internal/exec/pipeline.go:166-187, :539; the chains graph.Slope,
 * graph.Lookup, graph.Reclass lower to, graph/graph.go:250-287). The
 * output is bit for bit what the stages write as separate calls, each on
 * the one before: its edge ring and its validity erosion are the chain's
 * reach, the sum of the stage radii (pipeline.go:32-41, 83-114), and its
 * Stats are those of one kernel of that radius. Point stages (Lookup,
 * Reclass) after a radius-1 Horn stage are applied to its values in
 * registers before the store; any other stage writes its values once, into
 * device scratch the next stage reads. Every stage has one input and one
 * output; at most 8 stages and a reach of at most 8. Errors as
 * strata_process.
 */
int strata_process_chain(const strata_kernel *stages, int nstages, const strata_raster *dst, const strata_raster *src,
                         const strata_options *opts, strata_stats *stats);
int strata_process_chain_chunked(const strata_kernel *stages, int nstages, const strata_sink *dst,
                                 const strata_source *src, const strata_options *opts, strata_stats *stats);
/* Closed-form Stats of a chain call, as strata_plan_stats. */
int strata_plan_stats_chain(const strata_kernel *stages, int nstages, int64_t width, int64_t height,
                            const strata_options *opts, int chunked, strata_stats *stats);

/* ------------------------------------------------------------------ */
/* Reductions: exec.Reduce / exec.ReduceChunked with the reduce        */
/* package's reducers                                                  */
/* ------------------------------------------------------------------ */
/*
 * exec.Reduce (internal/exec/reduce.go:103) and exec.ReduceChunked
 * (internal/exec/reducechunked.go:14) over one raster, with the fold of
 * reduce.Count / MinMax / Sum / Stats (reduce/reduce.go:88-228,
 * reduce/sum.go:96-114, internal/summary/summary.go:59-104). Only valid
 * cells take part; data under an invalid cell is never read.
 *
 * The result is the combined partial, not the finished float64 values:
 * the exact sums come back as integers (what the reference's bins hold,
 * internal/accum/accum.go:49-76, carried into one number), so the caller
 * finishes Sum, Mean and StdDev with its own exact arithmetic
 * (internal/accum/result.go), as the reference does.
 *
 * `what` selects the work: COUNT alone folds only validity (no cell is
 * read; only count is filled); MINMAX adds the extremes and the NaN and
 * infinity tallies; SUM the exact sum; MOMENTS the sum and the sum of
 * squares.
 *
 * Stats as the reference's fold loops: per band Cells, Bands and
 * KernelRead (reduce.go:242-257), per chunked tile Tiles and SourceRead
 * (reducechunked.go:186-205); Ideal = Cells*4. A cancelled call returns
 * STRATA_CANCELLED and no value (reduce.go:88-101).
 */
#define STRATA_REDUCE_COUNT 0x1
#define STRATA_REDUCE_MINMAX 0x2
#define STRATA_REDUCE_SUM 0x4
#define STRATA_REDUCE_MOMENTS 0x8

typedef struct strata_fold {
    int64_t count;         /* valid cells                                  */
    int64_t nan;           /* valid NaNs                                   */
    int64_t pos_inf;       /* valid +Inf                                   */
    int64_t neg_inf;       /* valid -Inf                                   */
    int32_t not_neg_zero;  /* 1 once a valid value other than -0 was seen  */
    float min;             /* Go min / max of the valid cells (NaN         */
    float max;             /* absorbing, -0 < +0); +Inf/-Inf when count 0  */
    int32_t reserved;
    /* Sum of the finite valid cells times 2^150, an exact integer:
     * two's complement, 64-bit limbs, least significant first. */
    uint64_t sum[6];
    /* Sum of their squares times 2^300, likewise (MOMENTS only). */
    uint64_t sumsq[10];
} strata_fold;

int strata_reduce(int what, const strata_raster *src, const strata_options *opts, strata_fold *out,
                  strata_stats *stats);
int strata_reduce_chunked(int what, const strata_source *src, const strata_options *opts, strata_fold *out,
                          strata_stats *stats);

/*
 * strata_reduce over a device raster without allocating anything. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The fold's scratch (the exponent bins,
 * the running total and one partial per workgroup, under 128 KiB) is a
 * static device array that the library holds on each GPU it runs on until
 * the process ends. Results, Stats and cancellation are strata_reduce's.
 * src must be in device memory and hold at most 2^40 cells (STRATA_EINVAL
 * otherwise). Calls that overlap on one GPU share its scratch and race:
 * the caller serializes them.
 */
int strata_reduce_noalloc(int what, const strata_raster *src, const strata_options *opts, strata_fold *out,
                          strata_stats *stats);

/* ------------------------------------------------------------------ */
/* Launch tuning (no counterpart in the reference): shapes only, never  */
/* results                                                              */
/* ------------------------------------------------------------------ */
typedef struct strata_tuning {
    int32_t sweep_rows;  /* rows per block of the radius-1 terrain kernels (4..4096) */
    /* MiB of output per pass step when a chunked call over device-resident
     * operands folds an output it writes nowhere (a fold-only sink): the
     * rows go through a device buffer of about this size, kept for reuse
     * (strata_trim). Smaller steps hold less memory but launch more,
     * shorter passes (1..4096, default 4096). */
    int32_t fold_stripe_mb;
    /* Cells per unit, in 1024s, of a call that can be cancelled: the work
     * claimed between two polls of the cancel flag (four units are in
     * flight; a call of one unit runs as one launch). 0: the library's
     * choice, a sixteenth of the raster and at least 1 Gi cells. 0..2^20. */
    int32_t unit_kcells;
    /* For tests of cancellation: when > 0, a strata_process call with a
     * cancel flag claims at most this many units and then ends as if its
     * context had been cancelled (no effect on calls without a flag). 0: off. */
    int32_t stop_after_units;
    int32_t reserved[4];
} strata_tuning;

int strata_get_tuning(strata_tuning *t);
int strata_set_tuning(const strata_tuning *t);

/* ------------------------------------------------------------------ */
/* Self-checks (no counterpart in the reference)                       */
/* ------------------------------------------------------------------ */

/* Compare a device fast path with the full correctly rounded sequence it
 * stands in for: kind 0 the square root over the float bit patterns
 * [lo, lo+n), kind 1 the small-argument arctangent over [lo, lo+n),
 * kind 2 the quotient over n hashed operand pairs (seed lo), kind 3 the
 * fused small-slope path (root and arctangent of s = gx^2 + gy^2) over the
 * bit patterns [lo, lo+n). Reports the
 * number of differing results, the first differing index and how many
 * inputs were in the fast path's range. */
int strata_selfcheck(int kind, uint64_t lo, uint64_t n, uint64_t *mismatches, uint64_t *first, uint64_t *checked);

/* ------------------------------------------------------------------ */
/* Measurement (no counterpart in the reference, which has only Stats)  */
/* ------------------------------------------------------------------ */

/* Error of a Correlate output against the FP64 correlation of src (products
 * and sums in double, on the device) over the interior cells whose window
 * is finite: max |out - ref|, max |out - ref| / |ref| and max
 * |out - ref| / sum |w x|, and the number of cells compared. Used to report
 * the precision of the FP32 path and of the split-BF16x3 MFMA path
 * (STRATA_FLAG_MFMA). Device rasters only. */
int strata_correlate_error(const strata_kernel *k, const strata_raster *src, const strata_raster *out,
                           double *max_abs, double *max_rel, double *max_scaled, int64_t *cells);

/* Start recording HIP events around every data and validity kernel launch
 * (on the launch's own stream), discarding earlier records. */
int strata_timing_begin(void);
/* Stop recording; wait for the recorded launches and return the summed
 * kernel durations and launch counts. */
int strata_timing_end(double *data_ms, int64_t *data_launches, double *validity_ms, int64_t *validity_launches);

/* ------------------------------------------------------------------ */
/* Batches: many tiles, mixed ops, one persistent launch               */
/* ------------------------------------------------------------------ */
/*
 * No single counterpart in the reference: a batch is many independent
 * exec.ProcessN calls (internal/exec/process.go:28), one per task, each
 * on its own rasters (its own edge ring, terrain/doc.go:89-90), followed
 * by point ops applied to the task's output in place, as a chain
 * Slope -> transfer.Lookup -> transfer.Reclass written as separate calls
 * (transfer/example_test.go:90-129) or through exec.Pipeline
 * (internal/exec/pipeline.go:11-187). Every output is bit-identical to
 * those calls.
 *
 * Stencil ops a batch runs: Slope, Aspect, Hillshade, Curvature (radius
 * 1, no fit) and focal Mean, Min, Max at radius 1..3. Point ops: Lookup
 * and Reclass with at most 64 taps each. All rasters live in device
 * memory; tasks must not write memory another task reads.
 */
typedef struct strata_task {
    strata_kernel k;       /* the stencil op                              */
    strata_kernel post[2]; /* point ops applied to its output, in order    */
    int32_t npost;         /* 0..2                                         */
    int32_t reserved;
    strata_raster src;     /* the tile, e.g. a window of a resident raster */
    strata_raster dst;     /* its output, the same width and height        */
} strata_task;

typedef struct strata_batch_options {
    int32_t rows;          /* rows per work unit of radius-1 ops; 0: 64    */
    int32_t blocks_per_cu; /* persistent workgroups per CU; 0: occupancy   */
    int32_t graph;         /* 1: run as a captured HIP graph               */
    int32_t reserved;
    void *stream;          /* hipStream_t; NULL: a stream of the batch's own */
} strata_batch_options;

typedef struct strata_batch_info {
    int64_t tasks;
    int64_t cells;
    int64_t data_units;
    int64_t validity_units;
    int32_t rows;
    int32_t blocks;
    int32_t graph;
    int32_t reserved;
} strata_batch_info;

typedef struct strata_batch strata_batch;

/* Checks the tasks (STRATA_EINVAL before anything runs) and builds the
 * device work list. The caller keeps the rasters alive while the batch
 * is used; kernel taps are copied. */
int strata_batch_create(const strata_task *tasks, int ntasks, const strata_batch_options *opts,
                        strata_batch **out);
/* Enqueues the whole batch (one persistent launch, or its graph) on the
 * batch's stream and returns. */
int strata_batch_run(strata_batch *b);
/* Enqueues the same tasks through the plain path: one strata_process call
 * per task and per point op (the per-tile launch baseline; for point ops
 * this is faster than one strata_process_chain call per task, whose fused
 * sweep stages its tables and runs the validity pass in launches of their
 * own). */
int strata_batch_run_per_tile(strata_batch *b);
/* Times the batch at each candidate launch shape (rows per unit, blocks
 * per CU), reps runs each, and keeps the fastest; results never depend
 * on the shape. The chosen shape is written to chosen when not NULL. */
int strata_batch_tune(strata_batch *b, int reps, strata_batch_options *chosen);
/* Waits for the batch's stream. */
int strata_batch_sync(strata_batch *b);
int strata_batch_info_get(const strata_batch *b, strata_batch_info *out);
void strata_batch_destroy(strata_batch *b);

/* ------------------------------------------------------------------ */
/* Serving: window requests against a device-resident raster          */
/* ------------------------------------------------------------------ */
/*
 * No counterpart in the reference (a library, not a server). A request is
 * a window (x, y, w, h) of the resident raster; its result is what the
 * chain gives on that window as a raster of its own: the stencil op
 * (exec.ProcessN, internal/exec/process.go:28, NaN ring at the window's
 * edge, terrain/doc.go:89-90) and then the point ops, bit for bit the same
 * whether the request runs alone or among others.
 *
 * A scheduler thread admits pending requests in arrival order into
 * batches (one persistent launch each, the batch kernel) as soon as a
 * batch slot is free and their result memory fits the HBM budget; up to
 * `slots` batches run at once on their own streams, so requests that
 * arrive while others run join the next batch (continuous batching).
 * Latency runs from a request's arrival to the moment its batch is seen
 * complete.
 */
typedef struct strata_serve_config {
    strata_raster src;     /* device-resident raster requests read        */
    strata_kernel k;       /* the stencil op (as in a batch)              */
    strata_kernel post[2]; /* point ops applied after it, in order        */
    int32_t npost;
    int32_t slots;         /* batches in flight at once; 0: 4             */
    int64_t hbm_budget;    /* bytes of result memory; 0: 4 GiB            */
    int32_t max_batch;     /* requests per batch; 0: 256                  */
    int32_t rows;          /* rows per work unit of radius-1 ops; 0: per
                              request, short units for small windows      */
    int32_t auto_release;  /* 1: drop each result when it completes       */
    int32_t reserved;
} strata_serve_config;

typedef struct strata_window {
    int64_t x, y, w, h;
} strata_window;

typedef struct strata_serve_stats {
    int64_t requests;      /* completed                                   */
    int64_t batches;
    int64_t cells;
    int64_t max_in_flight; /* requests admitted and not yet complete      */
    int64_t admission_waits; /* times the queue head waited for memory    */
    int64_t peak_bytes;    /* peak result memory in use                   */
    double busy_ms;        /* wall time with at least one batch in flight */
    double queue_ms;       /* sum over requests of arrival -> launch      */
    double service_ms;     /* sum over requests of launch -> completion   */
} strata_serve_stats;

typedef struct strata_server strata_server;

/* Checks the configuration (STRATA_EINVAL), allocates the result arena
 * and starts the scheduler. */
int strata_serve_create(const strata_serve_config *cfg, strata_server **out);
/* Queues a request and returns at once with its id. A window outside the
 * raster, or one whose result can never fit the budget, is STRATA_EINVAL. */
int strata_serve_submit(strata_server *s, const strata_window *w, int64_t *id);
/* Waits for request id; its latency in milliseconds when not NULL. */
int strata_serve_wait(strata_server *s, int64_t id, double *latency_ms);
/* Copies a completed, unreleased result out: w*h cells row by row and the
 * validity of cell i at bit i of mask (ceil(w*h/64) words), both host or
 * device memory; either may be NULL. */
int strata_serve_fetch(strata_server *s, int64_t id, float *data, uint64_t *mask);
/* Frees a completed result's memory. */
int strata_serve_release(strata_server *s, int64_t id);
/* Open-loop replay: submits windows[i] at arrival_s[i] seconds after the
 * call starts (non-decreasing), waits for all, and writes each latency
 * (from its scheduled arrival) to latency_ms[i]. */
int strata_serve_replay(strata_server *s, const strata_window *windows, const double *arrival_s, int n,
                        double *latency_ms);
int strata_serve_stats_get(strata_server *s, strata_serve_stats *out);
void strata_serve_destroy(strata_server *s);

#ifdef __cplusplus
}
#endif
#endif
