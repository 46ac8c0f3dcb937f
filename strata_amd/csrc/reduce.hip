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
// lanes, 4 steps of 8 loads each). For each load the wave sums the
// significands of its lanes by exponent bin (one DPP sum per distinct
// exponent, few for real data) and one lane adds each sum to the
// workgroup's LDS bins; the LDS bins go to the global bins (one atomic per
// nonzero bin) when the workgroup ends.
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
constexpr int kStepLoads = 8;  // loads per lane in flight in one wave step

// One lane's running fold of what is not binned.
struct LaneFold {
    uint32_t count, nan, pinf, ninf, nnz;
    float mn, mx;
};

__device__ __forceinline__ void lds_add(long long *p, long long v)
{
    atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}

// Sum over the wave of x (|sum| < 2^31), returned to every lane. DPP row
// shifts give each row's sum in its lane 15, the row broadcasts carry them
// to lane 63.
__device__ __forceinline__ int wave_sum(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xe, true);  // row_shr:4, lanes 4-15
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xc, true);  // row_shr:8, lanes 8-15
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, true);  // row_bcast:15, rows 1 and 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, true);  // row_bcast:31, rows 2 and 3
    return __builtin_amdgcn_readlane(x, 63);
}

// Maximum over the wave of x, returned to every lane: wave_sum's pattern
// with max (0, what disabled lanes read, is its identity).
__device__ __forceinline__ uint32_t wave_max(uint32_t x)
{
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xe, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xc, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, true));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, true));
    return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// True on every lane of the wave (a ballot; p is 0 or 1).
__device__ __forceinline__ bool wave_all(uint32_t p) { return __builtin_amdgcn_ballot_w64(p == 0u) == 0; }

// v·2^k added to bins k (the low 32 bits, as a non-negative value) and
// k + 32 (the rest, signed): v = (v >> 32)·2^32 + (v & 0xffffffff).
__device__ __forceinline__ void lds_add_split(long long *bins, int k, long long v)
{
    lds_add(&bins[k], v & 0xffffffffll);
    lds_add(&bins[k + 32], v >> 32);
}

// Sum over the wave of a 64-bit value below 2^45 in magnitude, exactly:
// its low 24 bits and the rest summed apart (each sum stays below 2^31).
__device__ __forceinline__ long long wave_sum64(long long x)
{
    const int lo = wave_sum((int)(x & 0xffffff));
    const int hi = wave_sum((int)(x >> 24));
    return ((long long)hi << 24) + lo;
}

// One valid cell's tallies and extremes (addSpecial and the min/max fold);
// returns its exponent bin, or 0 for NaN and the infinities, which are
// counted, not binned.
__device__ __forceinline__ uint32_t tally(LaneFold &L, float v)
{
    const uint32_t u = f2u(v);
    L.count += 1;
    L.nnz |= u ^ 0x80000000u;
    L.mn = go_min(L.mn, v);
    L.mx = go_max(L.mx, v);
    const uint32_t field = (u >> 23) & 0xffu;
    const uint32_t special = field == 0xffu ? 1u : 0u;
    const uint32_t isnan = special & ((u & 0x7fffffu) != 0u ? 1u : 0u);
    L.nan += isnan;
    L.pinf += (special ^ isnan) & ((u >> 31) ^ 1u);
    L.ninf += (special ^ isnan) & (u >> 31);
    return special ? 0u : (field ? field : 1u);
}

// The wave's cells of one load slot into the LDS bins (split, addSum1 and
// addFinite): every distinct exponent bin among the lanes' keys (0: no
// cell) adds the wave's sum of signed significands, and for squares the
// sums of the squared significands' low and high 24 bits, once.
template <bool SQ>
__device__ __forceinline__ void bin_wave(uint32_t key, uint32_t u, long long *s_sum, long long *s_sq)
{
    const uint32_t mag = (u & 0x7fffffu) | (((u >> 23) & 0xffu) ? 0x800000u : 0u);
    const int sig = (u >> 31) ? -(int)mag : (int)mag;
    int lo = 0, hi = 0;
    if constexpr (SQ) {
        const uint64_t sq = (uint64_t)mag * mag;
        lo = (int)(sq & 0xffffffu);
        hi = (int)(sq >> 24);
    }
    const int lane = (int)(threadIdx.x & 63);
    for (;;) {
        const uint64_t pending = __builtin_amdgcn_ballot_w64(key != 0u);
        if (pending == 0) break;
        const uint32_t k0 = __builtin_amdgcn_readlane(key, __builtin_ctzll(pending));
        const bool mine = key == k0;
        const int ssum = wave_sum(mine ? sig : 0);
        int slo = 0, shi = 0;
        if constexpr (SQ) {
            slo = wave_sum(mine ? lo : 0);
            shi = wave_sum(mine ? hi : 0);
        }
        if (lane == 0) {
            lds_add(&s_sum[k0], ssum);
            if constexpr (SQ) {
                lds_add(&s_sq[2 * k0], slo);
                lds_add(&s_sq[2 * k0 + 24], shi);
            }
        }
        key = mine ? 0u : key;
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

// DATA false: a count of valid cells only (reduce.Count), no cell read.
template <bool MASKED, bool DATA, bool BIN, bool SQ>
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
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform in the wave
    LaneFold L;
    L.count = L.nan = L.pinf = L.ninf = L.nnz = 0;
    L.mn = __builtin_inff();
    L.mx = -__builtin_inff();

    // Units are claimed grid-stride, one per wave at a time. A wave step
    // covers 512 cells: eight loads per lane in flight, and the validity
    // words of the step (at most nine) loaded once for the wave.
    typedef const __attribute__((address_space(1))) float *grow_t;
    typedef const __attribute__((address_space(1))) uint64_t *gbits_t;
    for (int64_t u = (int64_t)blockIdx.x * kFoldWaves + wv; u < units; u += (int64_t)gridDim.x * kFoldWaves) {
        const int64_t y = u / units_x;
        const int64_t x0 = (u - y * units_x) * kUnitCells;
        const int32_t n = (int32_t)min((int64_t)kUnitCells, a.w - x0);
        const grow_t row = (grow_t)(a.p + y * a.stride + x0);
        const int64_t b0 = a.off + y * a.mstride + x0;  // validity bit of cell x0
        if constexpr (!DATA) {
            // reduce.Count: whole validity words, the unit's ends masked off.
            if constexpr (MASKED) {
                const int64_t wf = b0 >> 6, wl = (b0 + n - 1) >> 6;
                for (int64_t k = wf + lane; k <= wl; k += 64) {
                    uint64_t m = ((gbits_t)a.bits)[k];
                    if (k == wf) m &= ~0ull << (b0 & 63);
                    if (k == wl) m &= ~0ull >> (63 - ((b0 + n - 1) & 63));
                    L.count += (uint32_t)__builtin_popcountll(m);
                }
            } else {
                L.count += lane == 0 ? (uint32_t)n : 0u;
            }
            continue;
        }
        for (int32_t i0 = 0; i0 < n; i0 += kStepLoads * 64) {
            float v[kStepLoads];
            uint32_t ok[kStepLoads];
#pragma unroll
            for (int j = 0; j < kStepLoads; ++j) {
                const int32_t i = i0 + 64 * j + lane;
                ok[j] = i < n ? 1u : 0u;
                v[j] = ok[j] ? row[i] : 0.0f;
            }
            if constexpr (MASKED) {
                // Words [bb >> 6, last] hold the step's bits; lane l of
                // load j reads bit sh + 64 j + l of them.
                const int64_t bb = b0 + i0;
                const int32_t cnt = min(n - i0, kStepLoads * 64);
                const int sh = (int)(bb & 63);
                const int last = (sh + cnt - 1) >> 6;  // 0..kStepLoads
                const gbits_t wp = (gbits_t)(a.bits + (bb >> 6));
                // The 64 bits of load j, for the whole wave: words j and
                // j + 1 funnel-shifted right by sh (uniform values).
                uint64_t prev = wp[0];
#pragma unroll
                for (int j = 0; j < kStepLoads; ++j) {
                    const uint64_t next = j + 1 <= last ? wp[j + 1] : 0ull;
                    const uint64_t t = sh ? (prev >> sh) | (next << (64 - sh)) : prev;
                    ok[j] &= (uint32_t)((t >> lane) & 1u);
                    prev = next;
                }
            }
            uint32_t key[kStepLoads];
#pragma unroll
            for (int j = 0; j < kStepLoads; ++j) key[j] = ok[j] ? tally(L, v[j]) : 0u;
            if constexpr (BIN) {
                // A window of kWin exponent bins (base, the wave's largest
                // key] below the step's largest key. When every key of the
                // step lies in it, a lane adds sig·2^(e - base) (and the
                // square halves at 2^(2(e - base))) in 64 bits and the wave
                // sums those once. A lane's sums stay below 2^47: 8 loads of
                // 2^24·2^19 (sums alone, kWin 20), or of 2^24·2^20 (square
                // halves, kWin 11), so each wave_sum64 half stays below 2^31.
                constexpr uint32_t kWin = SQ ? 11u : 20u;
                uint32_t k1 = 0;
#pragma unroll
                for (int j = 0; j < kStepLoads; ++j) k1 = max(k1, key[j]);
                const uint32_t kmax = wave_max(k1);
                if (kmax != 0u) {
                    const uint32_t base = (kmax > kWin ? kmax : kWin) - (kWin - 1u);
                    uint32_t inwin = 1u;
#pragma unroll
                    for (int j = 0; j < kStepLoads; ++j) inwin &= (key[j] == 0u || key[j] >= base) ? 1u : 0u;
                    if (wave_all(inwin)) {
                        long long S = 0, Lq = 0, Hq = 0;
#pragma unroll
                        for (int j = 0; j < kStepLoads; ++j) {
                            if (key[j] == 0u) continue;
                            const uint32_t u = f2u(v[j]);
                            const uint32_t mag = (u & 0x7fffffu) | (((u >> 23) & 0xffu) ? 0x800000u : 0u);
                            const uint32_t d = key[j] - base;
                            const long long m = (long long)mag << d;
                            S += (u >> 31) ? -m : m;
                            if constexpr (SQ) {
                                const uint64_t q = (uint64_t)mag * mag;
                                Lq += (long long)(q & 0xffffffu) << (2 * d);
                                Hq += (long long)(q >> 24) << (2 * d);
                            }
                        }
                        const long long sw = wave_sum64(S);
                        long long lw = 0, hw = 0;
                        if constexpr (SQ) {
                            lw = wave_sum64(Lq);
                            hw = wave_sum64(Hq);
                        }
                        if (lane == 0) {
                            lds_add_split(s_sum, (int)base, sw);
                            if constexpr (SQ) {
                                lds_add_split(s_sq, 2 * (int)base, lw);
                                lds_add_split(s_sq, 2 * (int)base + 24, hw);
                            }
                        }
                    } else {
#pragma unroll
                        for (int j = 0; j < kStepLoads; ++j) bin_wave<SQ>(key[j], f2u(v[j]), s_sum, s_sq);
                    }
                }
            }
        }
    }
    if constexpr (BIN) {
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

template <bool MASKED, bool DATA, bool BIN, bool SQ>
void launch_fold_t(const FoldArgs &a, int64_t units_x, int64_t units, int grid, FoldBins *bins, FoldPart *parts,
                   hipStream_t st)
{
    hipLaunchKernelGGL((fold_kernel<MASKED, DATA, BIN, SQ>), dim3(grid), dim3(kFoldThreads), 0, st, a, units_x,
                       units, bins, parts);
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
    const bool data = (a.what & ~STRATA_REDUCE_COUNT) != 0;
    if (masked) {
        if (sq) launch_fold_t<true, true, true, true>(a, units_x, units, grid, bins, parts, st);
        else if (bin) launch_fold_t<true, true, true, false>(a, units_x, units, grid, bins, parts, st);
        else if (data) launch_fold_t<true, true, false, false>(a, units_x, units, grid, bins, parts, st);
        else launch_fold_t<true, false, false, false>(a, units_x, units, grid, bins, parts, st);
    } else {
        if (sq) launch_fold_t<false, true, true, true>(a, units_x, units, grid, bins, parts, st);
        else if (bin) launch_fold_t<false, true, true, false>(a, units_x, units, grid, bins, parts, st);
        else if (data) launch_fold_t<false, true, false, false>(a, units_x, units, grid, bins, parts, st);
        else launch_fold_t<false, false, false, false>(a, units_x, units, grid, bins, parts, st);
    }
    hipLaunchKernelGGL(fold_finish, dim3(1), dim3(kFoldThreads), 0, st, parts, grid, total);
    return hipGetLastError();
}

} // namespace strata
