// host.hpp — host-side helpers shared by the C ABI's source files: error
// reporting (the calling thread's last error), kernel descriptor and raster
// checks, and the resolved per-call constants.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "strata_hip.h"

namespace strata {

// Records the message as the calling thread's last error; returns code.
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

int n_outputs(const strata_kernel *k);
int check_kernel(const strata_kernel *k, int ndst, int nsrc);
int check_raster(const char *name, int i, const strata_raster &r);
Params make_params(const strata_kernel *k);

// Batches (batch.hip).
int64_t cdiv64(int64_t a, int64_t b);
int batch_kind(const strata_kernel &k);  // BatchKind, or -1 if a batch cannot run k
int check_task(const strata_task &t, int i);
BatchOp make_op(const strata_task &t);
BatchTask make_task(const strata_task &t, int32_t op);
// A batch's work lists as one buffer: the three headers at 0, the ops at
// `ops`, each list's tasks at tasks[f], the claim counters at `counters`.
struct ListLayout {
    size_t ops = 0, tasks[kBatchFamilies] = {0, 0, 0}, counters = 0, bytes = 0;
    BatchHeader h[kBatchFamilies] = {};
    int focal_one = -1;  // every focal task's kind and radius (batch_focal_one_code), or -1
};
// Bytes of the buffer for nops ops and ntasks tasks (an upper bound).
size_t list_bytes(int nops, size_t ntasks);
// Writes the lists of n tasks (made by make_task, in work-list order;
// masked[i]: task i's output has a mask) into out, for `rows` rows per
// radius-1 unit, or (rows 0) rows chosen per task for latency.
ListLayout write_lists(const BatchOp *ops, int nops, const BatchTask *tasks, const bool *masked, size_t n,
                       int rows, char *out);
// The launch of lists uploaded to dev, with at most the given workgroups
// for the wave kernel (sweeps and validity) and the focal kernel; clear:
// zero the counters with a kernel first (replays of uploaded lists).
BatchLaunch launch_view(const ListLayout &L, char *dev, int max_wave_blocks, int max_focal_blocks, bool clear);

} // namespace strata

#define HIP_TRY(expr, what)                                                                                  \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return ::strata::hip_fail(e_, what);                                           \
    } while (0)
