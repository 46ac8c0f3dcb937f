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

} // namespace strata

#define HIP_TRY(expr, what)                                                                                  \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return ::strata::hip_fail(e_, what);                                           \
    } while (0)
