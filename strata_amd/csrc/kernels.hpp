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

// Resolved per-call constants: k[] as the C ABI's strata_kernel, taps
// copied by value into the kernel arguments (read through the scalar
// cache as uniform values).
struct Params {
    int32_t op, radius, flags, ntaps;
    float k[16];
    float taps[kMaxTaps];
};

// Launch-shape knobs (results never depend on them).
struct Tuning {
    int32_t sweep_rows;  // rows per block of the radius-1 sweep
    int32_t reserved[7];
};
Tuning get_tuning();
bool set_tuning(const Tuning &t);

// Data of every cell in the region: the op over interior cells, NaN over
// the edge ring. nout is 1 or 2 (Gradient). Returns a HIP error code.
hipError_t launch_data(const Params &P, const Region &rg, const SrcView &s, const DstView *d, int nout,
                       hipStream_t st);

// Validity of the region's cells of one output: interior AND every cell of
// the (2r+1)^2 box valid in the source (has_src), or interior alone.
hipError_t launch_validity(int r, const Region &rg, const MaskView &src, int has_src, const MaskView &dst,
                           hipStream_t st);

// Self-check of a fast path against the full sequence over n inputs (see
// selfcheck_kernel); dev3 holds {mismatches, first mismatch index, checked}.
hipError_t launch_selfcheck(int kind, uint64_t lo, uint64_t n, unsigned long long *dev3, hipStream_t st);

} // namespace strata
