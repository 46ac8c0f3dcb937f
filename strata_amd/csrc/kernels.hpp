// kernels.hpp — launch interface between the host runtime and the gfx950
// kernels. Coordinates are raster cells; every view says where its memory
// sits in raster coordinates, so one kernel serves the plain call (views at
// the raster origin), a band or tile of the plan (a sub-region) and a
// chunked tile buffer (a view placed at the tile's position).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "strata_hip.h"

namespace strata {

constexpr int kMaxTaps = 320;   // 289 correlate weights (r = 8) fit comfortably
constexpr int kMaxRadius = 8;   // focal.MaxRadius (focal/focal.go) and terrain.MaxRadius
constexpr int kMaxOutputs = 8;  // outputs of one launch: Surface's 5 products, or a fan-out sweep's ops

// The output region [ox, ox+w) x [oy, oy+h) of a W x H raster.
struct Region {
    int64_t W, H;
    int64_t ox, oy, w, h;
};

// Input data: raster cell (x, y) lives at p[(y - sy)*stride + (x - sx)] and
// may be read only inside [lx0, lx1) x [ly0, ly1) (the buffer's extent).
struct SrcView {
    const float *p;
    int64_t stride;
    int64_t sx, sy;
    int64_t lx0, lx1, ly0, ly1;
};

// Output data: region cell (x, y) lives at p[(y - oy)*stride + (x - ox)].
struct DstView {
    float *p;
    int64_t stride;
};

// Validity bits: cell (x, y) is bit off + (y - y0)*stride + (x - x0), held
// for cells [lx0, lx1) x [ly0, ly1) (the buffer's extent in the raster).
struct MaskView {
    uint64_t *bits;
    int64_t off;
    int64_t stride;
    int64_t x0, y0;
    int64_t lx0, lx1, ly0, ly1;
};

// Radius-1 terrain ops reading one source in one pass (a fan-out of
// single-output ops, as terrain.Features builds, terrain/features.go:
// 103-117): output o is op[o] (Slope, Aspect, Hillshade or Curvature)
// with its own flags[o] and constants k[o][] (the strata_kernel layout).
struct FanArgs {
    int32_t n;
    int32_t op[kMaxOutputs];
    int32_t flags[kMaxOutputs];
    float k[kMaxOutputs][8];
};

// Internal op id of a FanArgs launch (never crosses the C ABI).
constexpr int32_t kOpHornFan = 100;

// Point stages (Lookup, Reclass) applied in order to every value a
// radius-1 Horn sweep stores: the trailing stages of a chain, fused into
// the stencil's store (a chain of exec.Pipeline, internal/exec/
// pipeline.go:187, run as one kernel). Stage i is op[i] with its table of
// cnt[i] floats at taps[off[i]].
constexpr int kMaxPost = 4;
constexpr int kPostTaps = 256;
struct PostArgs {
    int32_t n;
    int32_t op[kMaxPost];
    int32_t off[kMaxPost];
    int32_t cnt[kMaxPost];
    float taps[kPostTaps];
};

// Resolved per-call constants: k[] as the C ABI's strata_kernel, taps
// copied by value into the kernel arguments (read through the scalar
// cache as uniform values); fan for kOpHornFan; post for a radius-1 Horn
// op whose store takes a chain's point stages.
struct Params {
    int32_t op, radius, flags, ntaps;
    float k[16];
    float taps[kMaxTaps];
    FanArgs fan;
    PostArgs post;
    const PostArgs *post_dev;  // post in device memory (launch_stage_post), read by the fused sweep
};

// Writes pa into device memory at dev (a kernel copies it from its own
// arguments, in stream order).
hipError_t launch_stage_post(const PostArgs &pa, PostArgs *dev, hipStream_t st);

// Launch-shape knobs (results never depend on them).
struct Tuning {
    int32_t sweep_rows;      // rows per block of the radius-1 sweep
    int32_t fold_stripe_mb;  // output MiB per step of a resident fold-only pass
    int32_t unit_kcells;     // cells per unit of a cancellable call, in 1024s (0: the default)
    int32_t stop_after_units;  // tests: a cancellable call stops after this many units (0: off)
    int32_t reserved[4];
};
Tuning get_tuning();
bool set_tuning(const Tuning &t);

// Stream-ordered device memory from the library's pool, which keeps what
// calls free for later calls (freed with hipFreeAsync on the same stream).
hipError_t scratch_alloc_async(void **p, size_t bytes, hipStream_t st);

// Data of every cell in the region: the op over interior cells, NaN over
// the edge ring. nout is 1, 2 (Gradient) or up to 5 (Surface). Returns a
// HIP error code.
hipError_t launch_data(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                       hipStream_t st);

// Data and validity of a radius-1 Horn op (Slope, Aspect, Hillshade,
// Curvature; no fit, no fused point stages) with one masked output in one
// launch: the validity waves run as extra workgroups after the sweep's.
// *fused tells whether it launched; when not, the caller launches
// launch_data and launch_validity.
hipError_t launch_data_valid(const Params &P, const Region &rg, const SrcView &s, const DstView &d,
                             const MaskView &vsrc, int has_src, const MaskView &vdst, hipStream_t st, bool *fused);

// Validity of the region's cells of one output: interior AND every cell of
// the (2r+1)^2 box valid in the source (has_src), or interior alone.
hipError_t launch_validity(int r, const Region &rg, const MaskView &src, int has_src, const MaskView &dst,
                           hipStream_t st);

// Error of a Correlate output (out, region-relative rows of out_stride)
// against the FP64 correlation over the region's interior cells with a
// finite window; acc (device, zeroed) receives {max abs, max rel, max
// |err| / sum|w x|} as double bit patterns and the count of cells.
hipError_t launch_correlate_error(const Params &P, const Region &rg, const SrcView &s, const float *out,
                                  int64_t out_stride, unsigned long long *acc, hipStream_t st);

// Self-check of a fast path against the full sequence over n inputs (see
// selfcheck_kernel); dev3 holds {mismatches, first mismatch index, checked}.
hipError_t launch_selfcheck(int kind, uint64_t lo, uint64_t n, unsigned long long *dev3, hipStream_t st);

// ---------------------------------------------------------------------------
// Reductions: one raster folded to counts, extremes and exact sums
// ---------------------------------------------------------------------------

// Exact sums as integers binned by float32 exponent field (the layout of
// internal/accum/accum.go:49-76): a finite value with field f (e = max(f,
// 1)) contributes its signed significand at weight 2^e, and its squared
// significand's low and high 24 bits at weights 2^(2e) and 2^(2e + 24).
// Bin k of sum weighs 2^(k - 150), bin j of sq 2^(j - 300). A bin may also
// take a partial sum at a lower weight, split into 32-bit pieces at bins k
// and k + 32, which is what the 48 bins above the exponents are for; no
// bin grows by 2^32 per 512 values, so int64 bins hold 2^40 values.
constexpr int kFoldSumBins = 256 + 48;
constexpr int kFoldSqBins = 2 * 255 + 24 + 48 + 1;
struct FoldBins {
    long long sum[kFoldSumBins];
    long long sq[kFoldSqBins];
};

// What is not binned: counts of the valid cells, of NaNs and of each
// infinity; nnz nonzero once a valid value other than -0 was folded; the
// Go min and max of the valid cells (+Inf and -Inf when there are none).
struct FoldPart {
    unsigned long long count, nan, pinf, ninf;
    unsigned nnz;
    float mn, mx;
    unsigned pad;
};

// One raster region to fold: cell (x, y) at p[y*stride + x] for x < w,
// y < h, valid when bits is null or bit off + y*mstride + x is set.
struct FoldArgs {
    const float *p;
    int64_t stride, w, h;
    const uint64_t *bits;
    int64_t off, mstride;
    int32_t what;  // STRATA_REDUCE_* flags
    int32_t pad;
};

constexpr int kFoldMaxBlocks = 2048;  // partials one launch writes at most

// Folds a's valid cells into bins (atomically) and the running total,
// using parts (kFoldMaxBlocks entries) as scratch; clear: start bins and
// total from empty first.
hipError_t launch_fold(const FoldArgs &a, FoldBins *bins, FoldPart *parts, FoldPart *total, bool clear,
                       hipStream_t st);

// ---------------------------------------------------------------------------
// Raw files: fill-value NoData of a tile buffer
// ---------------------------------------------------------------------------

// A chunked tile buffer: cell (x, y) at p[y*stride + x], its validity bit
// y*stride + x of bits; stride is a multiple of 64.
struct RawTile {
    float *p;
    int64_t stride, w, h;
    uint64_t *bits;
};

// Writes the validity words of the tile's rows from the fill value: a cell
// is invalid where it equals fill as a float (a NaN fill: where it is a NaN).
hipError_t launch_raw_validity(const RawTile &t, float fill, hipStream_t st);
// Stores fill into every invalid cell of the tile.
hipError_t launch_raw_fill(const RawTile &t, float fill, hipStream_t st);

// ---------------------------------------------------------------------------
// Batches: many tiles, mixed ops, one persistent launch over a work list
// ---------------------------------------------------------------------------

constexpr int kBatchPostTaps = 64;  // floats per point-op table in a batch op

// Families of stencil work a batch runs.
enum BatchKind : int32_t {
    BK_HORN = 0,  // radius-1 terrain ops (horn_sweep's work)
    BK_MEAN = 1,  // focal Mean, Min, Max at radius 1..3
    BK_MIN = 2,
    BK_MAX = 3,
};

// One op of a batch: the stencil op and the point ops (Lookup, Reclass)
// applied to its output in order, which is the standalone chain of ops
// cell for cell because point ops read only their own cell.
struct BatchOp {
    int32_t op, radius, flags, kind;
    float k[16];
    int32_t npost, pad;
    int32_t post_op[2];
    int32_t post_n[2];
    float post_taps[2][kBatchPostTaps];
};

// One tile: an op over its own raster (its edge ring is its own), written to
// its own output. A task sits in two work lists: its data list (units of one
// wave's 128 columns of `rows` rows for radius-1 ops, or a workgroup's column
// of focal tiles) and, when its output has a mask, the validity list (units
// of one validity wave). [u0, u1) are its units in the list the copy belongs
// to.
// Rows per wave of the validity kernel (8 against 16: 0.038 against
// 0.043 ms on the 20000² Slope mask; 32 took 0.060 ms).
constexpr int kVsRows = 8;
// Rows per row group of a narrow region's validity units (eight: sixteen
// raised the wave kernel's registers).
constexpr int kVgRows = 8;
// A batch's validity unit: this many consecutive row chunks (of kVsRows
// rows, or of a narrow region's row groups) of one task, run one after the
// other by the wave that claims the unit, so the claim, the workgroup's
// barrier and the task's fields are paid once for all of them.
constexpr int kVUnitChunks = 4;

struct BatchTask {
    Region rg;
    SrcView s;
    DstView d;
    MaskView sm, dm;
    int32_t op;       // index into the batch's op table
    int32_t has_src;  // the source has a mask
    int32_t vunits;   // validity waves (0 when the output has no mask)
    int32_t aligned;  // pair (8-byte) accesses allowed
    uint32_t blocks_x;  // data units across
    uint32_t rows;      // rows per data unit
    uint32_t vwaves_x;  // validity waves across
    uint32_t vgroup;    // lanes per row group of a narrow region's validity units (0: one row per wave)
    int64_t u0, u1;
};

// Work lists of a batch. The radius-1 sweep list and the validity list are
// run by one persistent kernel (sweep units first), the focal list by a
// second one beside it.
enum BatchFamily : int {
    BF_HORN = 0,      // radius-1 terrain sweeps
    BF_FOCAL = 1,     // focal tiles
    BF_VALIDITY = 2,  // validity of every masked output
};
constexpr int kBatchFamilies = 3;

struct BatchHeader {
    int64_t ntasks, nunits;
};

// Device pointers of one batch's lists (the three headers, each list's
// tasks sorted by u0, two claim counters 256 bytes apart: sweeps and
// validity, then focal) and the launch shape.
struct BatchLaunch {
    const BatchOp *ops;
    const BatchHeader *hdr;  // kBatchFamilies headers
    const BatchTask *tasks[kBatchFamilies];
    unsigned *counters;
    int64_t nunits[kBatchFamilies];
    int blocks_wave, blocks_focal;
    bool clear;  // zero the counters first (they are not uploaded zeroed)
    // When every focal task has one kind and radius: 4 * kind + radius
    // (batch_focal_one_code), and the focal list runs through a kernel
    // built for that op alone; -1 otherwise.
    int focal_one;
};

// The code of a focal kind (BK_MEAN, BK_MIN, BK_MAX) and radius (1..3).
constexpr int batch_focal_one_code(int kind, int radius) { return 4 * kind + radius; }

// Fills a task's unit shape for `rows` rows per radius-1 unit: rows,
// blocks_x (data units across) and vwaves_x.
void batch_units(const BatchOp &op, BatchTask &t, int rows);

// Runs the lists' persistent kernels on st: the focal list, then the wave
// lists.
hipError_t launch_batch(const BatchLaunch &L, hipStream_t st);

// Workgroups of a list's kernel resident per CU (occupancy), for sizing.
int batch_blocks_per_cu(int family, int focal_one = -1);

} // namespace strata
