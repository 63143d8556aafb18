// szx_rate.hpp -- fixed-rate compression on the exact codec (szx_rate.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szx_amd.h"

namespace szxamd {

uint64_t fixed_rate_workspace(uint64_t n, int w, int bs);

// Solve for the bound and leave its stream in d_out (d_result as for
// compress_launch).  d_params: the rate model (learned method only).
// Returns 0, or -1 with *err set on a device error.  Synchronizes `st`
// after every pass.
int fixed_rate_solve(const void *d_in, uint64_t n, int w, int bs, float target, float tol, int method,
                     const float *d_params, uint8_t *d_out, uint8_t *d_ws, uint64_t *d_result,
                     uint64_t *d_index, szx_amd_rate_info *info, hipStream_t st, const char **err);

}  // namespace szxamd
