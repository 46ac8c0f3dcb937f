// reduce.hip — the fold of a raster's valid cells on the device: counts, the
// NaN and infinity tallies, Go's min and max, and the exact sums of
// reduce.Sum and reduce.Stats as integers binned by exponent.
//
// Reference behaviour restated here:
//   countOp, minMaxOp   reduce/reduce.go:88-228 (NaN absorbing, -0 < +0)
//   sumOp, summary      reduce/sum.go:96-114, internal/summary/summary.go:59-104
//   exact bins          internal/accum/accum.go:192-245 (split, addSum1,
//                       addSpecial), internal/accum/moments.go:95-112
//                       (addFinite: the square's 24-bit halves)
// The bins are integers, so the order cells are added in never changes
// them: waves, workgroups and launches may split the raster any way.
//
// Layout: a unit is one wave's run of up to 2048 cells of one row (64
// lanes, 32 steps). Each lane keeps a running sum for the exponent it saw
// last and adds it to the workgroup's LDS bins only when the exponent
// changes, so smooth data touches LDS rarely; the LDS bins go to the
// global bins (one atomic per nonzero bin) when the workgroup ends.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "arith.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace strata {

namespace {

constexpr int kFoldThreads = 256;
constexpr int kFoldWaves = kFoldThreads / 64;
constexpr int kUnitCells = 2048;

// One lane's running fold.
struct LaneFold {
    uint32_t count, nan, pinf, ninf, nnz;
    float mn, mx;
    int cur;  // exponent bin of the running sums, -1 for none
    long long acc, alo, ahi;
};

__device__ __forceinline__ void lds_add(long long *p, long long v)
{
    atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}

template <bool SQ>
__device__ __forceinline__ void flush(LaneFold &L, long long *s_sum, long long *s_sq)
{
    if (L.cur >= 0) {
        lds_add(&s_sum[L.cur], L.acc);
        if constexpr (SQ) {
            lds_add(&s_sq[2 * L.cur], L.alo);
            lds_add(&s_sq[2 * L.cur + 24], L.ahi);
        }
    }
    L.acc = L.alo = L.ahi = 0;
}

// One valid cell (addSum1 / addFinite / addSpecial and the min/max fold).
template <bool BIN, bool SQ>
__device__ __forceinline__ void fold_cell(LaneFold &L, float v, long long *s_sum, long long *s_sq)
{
    const uint32_t u = f2u(v);
    L.count += 1;
    L.nnz |= u ^ 0x80000000u;
    L.mn = go_min(L.mn, v);
    L.mx = go_max(L.mx, v);
    const uint32_t field = (u >> 23) & 0xffu;
    if (field == 0xffu) {
        const uint32_t isnan = (u & 0x7fffffu) != 0u;
        L.nan += isnan;
        L.pinf += (isnan ^ 1u) & ((u >> 31) ^ 1u);
        L.ninf += (isnan ^ 1u) & (u >> 31);
        return;
    }
    if constexpr (BIN) {
        const int e = field ? (int)field : 1;
        const uint32_t mag = (u & 0x7fffffu) | (field ? 0x800000u : 0u);
        if (e != L.cur) {
            flush<SQ>(L, s_sum, s_sq);
            L.cur = e;
        }
        L.acc += (u >> 31) ? -(long long)mag : (long long)mag;
        if constexpr (SQ) {
            const uint64_t s = (uint64_t)mag * mag;
            L.alo += (long long)(s & 0xffffffu);
            L.ahi += (long long)(s >> 24);
        }
    }
}

__device__ __forceinline__ FoldPart part_identity()
{
    FoldPart p;
    p.count = p.nan = p.pinf = p.ninf = 0;
    p.nnz = 0;
    p.mn = __builtin_inff();
    p.mx = -__builtin_inff();
    p.pad = 0;
    return p;
}

// minMaxOp.Combine / summary.Reducer.Combine: the identities are the empty
// partial's, so no emptiness test is needed.
__device__ __forceinline__ FoldPart combine(FoldPart a, const FoldPart &b)
{
    a.count += b.count;
    a.nan += b.nan;
    a.pinf += b.pinf;
    a.ninf += b.ninf;
    a.nnz |= b.nnz;
    a.mn = go_min(a.mn, b.mn);
    a.mx = go_max(a.mx, b.mx);
    return a;
}

__device__ __forceinline__ FoldPart shfl_part(const FoldPart &p, int src)
{
    FoldPart q;
    q.count = __shfl(p.count, src);
    q.nan = __shfl(p.nan, src);
    q.pinf = __shfl(p.pinf, src);
    q.ninf = __shfl(p.ninf, src);
    q.nnz = __shfl(p.nnz, src);
    q.mn = __shfl(p.mn, src);
    q.mx = __shfl(p.mx, src);
    q.pad = 0;
    return q;
}

// The workgroup's partial, in LDS slot 0 for thread 0; every thread calls.
__device__ __forceinline__ FoldPart block_combine(FoldPart p, FoldPart *s_part)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int d = 32; d >= 1; d >>= 1) p = combine(p, shfl_part(p, lane ^ d));
    if (lane == 0) s_part[wv] = p;
    __syncthreads();
    FoldPart r = part_identity();
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r = combine(r, s_part[i]);
    return r;
}

template <bool MASKED, bool BIN, bool SQ>
__global__ __launch_bounds__(kFoldThreads) void fold_kernel(FoldArgs a, int64_t units_x, int64_t units,
                                                           FoldBins *bins, FoldPart *parts)
{
    __shared__ long long s_sum[kFoldSumBins];
    __shared__ long long s_sq[kFoldSqBins];
    __shared__ FoldPart s_part[kFoldWaves];
    if constexpr (BIN) {
        for (int i = threadIdx.x; i < kFoldSumBins; i += kFoldThreads) s_sum[i] = 0;
        if constexpr (SQ)
            for (int i = threadIdx.x; i < kFoldSqBins; i += kFoldThreads) s_sq[i] = 0;
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    LaneFold L;
    L.count = L.nan = L.pinf = L.ninf = L.nnz = 0;
    L.mn = __builtin_inff();
    L.mx = -__builtin_inff();
    L.cur = -1;
    L.acc = L.alo = L.ahi = 0;

    // Units are claimed grid-stride, one per wave at a time.
    for (int64_t u = (int64_t)blockIdx.x * kFoldWaves + wv; u < units; u += (int64_t)gridDim.x * kFoldWaves) {
        const int64_t y = u / units_x;
        const int64_t x0 = (u - y * units_x) * kUnitCells;
        const int32_t n = (int32_t)min((int64_t)kUnitCells, a.w - x0);
        const float *row = a.p + y * a.stride + x0;
        const int64_t b0 = a.off + y * a.mstride + x0;  // validity bit of cell x0
#pragma unroll 4
        for (int32_t i = lane; i < n; i += 64) {
            bool valid = true;
            if constexpr (MASKED) {
                const int64_t b = b0 + i;
                valid = ((a.bits[b >> 6] >> (b & 63)) & 1u) != 0;
            }
            if (valid) fold_cell<BIN, SQ>(L, row[i], s_sum, s_sq);
        }
    }
    if constexpr (BIN) {
        flush<SQ>(L, s_sum, s_sq);
        __syncthreads();
        // Nonzero LDS bins into the global bins.
        for (int i = threadIdx.x; i < kFoldSumBins; i += kFoldThreads)
            if (s_sum[i] != 0) atomicAdd(reinterpret_cast<unsigned long long *>(&bins->sum[i]),
                                         (unsigned long long)s_sum[i]);
        if constexpr (SQ)
            for (int i = threadIdx.x; i < kFoldSqBins; i += kFoldThreads)
                if (s_sq[i] != 0) atomicAdd(reinterpret_cast<unsigned long long *>(&bins->sq[i]),
                                            (unsigned long long)s_sq[i]);
    }
    FoldPart p;
    p.count = L.count;
    p.nan = L.nan;
    p.pinf = L.pinf;
    p.ninf = L.ninf;
    p.nnz = L.nnz;
    p.mn = L.mn;
    p.mx = L.mx;
    p.pad = 0;
    p = block_combine(p, s_part);
    if (threadIdx.x == 0) parts[blockIdx.x] = p;
}

// Combines parts[0..n) into *total.
__global__ __launch_bounds__(kFoldThreads) void fold_finish(const FoldPart *parts, int n, FoldPart *total)
{
    __shared__ FoldPart s_part[kFoldWaves];
    FoldPart p = part_identity();
    for (int i = threadIdx.x; i < n; i += kFoldThreads) p = combine(p, parts[i]);
    p = block_combine(p, s_part);
    if (threadIdx.x == 0) *total = combine(*total, p);
}

// Empties the bins and the total (a kernel, not a memset, so the call can
// sit in a captured graph).
__global__ __launch_bounds__(kFoldThreads) void fold_clear(FoldBins *bins, FoldPart *total)
{
    for (int i = threadIdx.x; i < kFoldSumBins; i += kFoldThreads) bins->sum[i] = 0;
    for (int i = threadIdx.x; i < kFoldSqBins; i += kFoldThreads) bins->sq[i] = 0;
    if (threadIdx.x == 0) *total = part_identity();
}

template <bool MASKED, bool BIN, bool SQ>
void launch_fold_t(const FoldArgs &a, int64_t units_x, int64_t units, int grid, FoldBins *bins, FoldPart *parts,
                   hipStream_t st)
{
    hipLaunchKernelGGL((fold_kernel<MASKED, BIN, SQ>), dim3(grid), dim3(kFoldThreads), 0, st, a, units_x, units,
                       bins, parts);
}

} // namespace

hipError_t launch_fold(const FoldArgs &a, FoldBins *bins, FoldPart *parts, FoldPart *total, bool clear,
                       hipStream_t st)
{
    if (clear) hipLaunchKernelGGL(fold_clear, dim3(1), dim3(kFoldThreads), 0, st, bins, total);
    if (a.w <= 0 || a.h <= 0) return hipGetLastError();
    const int64_t units_x = (a.w + kUnitCells - 1) / kUnitCells;
    const int64_t units = units_x * a.h;
    const int grid = (int)std::min<int64_t>(kFoldMaxBlocks, (units + kFoldWaves - 1) / kFoldWaves);
    const bool masked = a.bits != nullptr;
    const bool sq = (a.what & STRATA_REDUCE_MOMENTS) != 0;
    const bool bin = sq || (a.what & STRATA_REDUCE_SUM) != 0;
    if (masked) {
        if (sq) launch_fold_t<true, true, true>(a, units_x, units, grid, bins, parts, st);
        else if (bin) launch_fold_t<true, true, false>(a, units_x, units, grid, bins, parts, st);
        else launch_fold_t<true, false, false>(a, units_x, units, grid, bins, parts, st);
    } else {
        if (sq) launch_fold_t<false, true, true>(a, units_x, units, grid, bins, parts, st);
        else if (bin) launch_fold_t<false, true, false>(a, units_x, units, grid, bins, parts, st);
        else launch_fold_t<false, false, false>(a, units_x, units, grid, bins, parts, st);
    }
    hipLaunchKernelGGL(fold_finish, dim3(1), dim3(kFoldThreads), 0, st, parts, grid, total);
    return hipGetLastError();
}

} // namespace strata
