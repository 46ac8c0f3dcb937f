// rawio.hip — the fill-value NoData of raw files, on the device: the
// validity of a tile as it arrives (engine/raw.go:139-160 validity,
// internal/vec/validity.go:11-19 ValidBits: a cell is invalid where it
// equals the fill as a float, so -0 matches 0, and with a NaN fill where it
// is any NaN), and the fill written into the invalid cells of an output
// tile before it leaves (engine/raw.go:247-273 fillInvalid).
//
// Tiles live in the chunked engine's buffers: rows of `stride` cells with
// stride a multiple of 64, so every row's validity starts a word. One wave
// covers 64 cells of a row; a ballot is the row's validity word.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "arith.hpp"
#include "kernels.hpp"

namespace strata {

namespace {

constexpr int kRawThreads = 256;
constexpr int kRawMaxGridY = 65535;

__global__ __launch_bounds__(kRawThreads) void raw_validity_kernel(RawTile t, float fill, int fill_nan)
{
    const int lane = threadIdx.x & 63;
    const int64_t x = (int64_t)blockIdx.x * kRawThreads + threadIdx.x;
    for (int64_t y = blockIdx.y; y < t.h; y += gridDim.y) {
        bool valid = false;
        if (x < t.w) {
            const float v = t.p[y * t.stride + x];
            valid = fill_nan ? !(v != v) : !(v == fill);
        }
        const uint64_t word = __builtin_amdgcn_ballot_w64(valid);
        if (lane == 0 && x < t.w) t.bits[(y * t.stride + x) >> 6] = word;
    }
}

__global__ __launch_bounds__(kRawThreads) void raw_fill_kernel(RawTile t, float fill)
{
    const int64_t x = (int64_t)blockIdx.x * kRawThreads + threadIdx.x;
    if (x >= t.w) return;
    for (int64_t y = blockIdx.y; y < t.h; y += gridDim.y) {
        const int64_t i = y * t.stride + x;
        if (!((t.bits[i >> 6] >> (i & 63)) & 1u)) t.p[i] = fill;
    }
}

dim3 raw_grid(const RawTile &t)
{
    return dim3((unsigned)((t.w + kRawThreads - 1) / kRawThreads), (unsigned)std::min<int64_t>(t.h, kRawMaxGridY));
}

} // namespace

hipError_t launch_raw_validity(const RawTile &t, float fill, hipStream_t st)
{
    if (t.w <= 0 || t.h <= 0) return hipSuccess;
    const int fill_nan = fill != fill ? 1 : 0;
    hipLaunchKernelGGL(raw_validity_kernel, raw_grid(t), dim3(kRawThreads), 0, st, t, fill, fill_nan);
    return hipGetLastError();
}

hipError_t launch_raw_fill(const RawTile &t, float fill, hipStream_t st)
{
    if (t.w <= 0 || t.h <= 0) return hipSuccess;
    hipLaunchKernelGGL(raw_fill_kernel, raw_grid(t), dim3(kRawThreads), 0, st, t, fill);
    return hipGetLastError();
}

} // namespace strata
