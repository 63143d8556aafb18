// szx_probe.hip -- the device copy rate that bounds the codec kernels: a
// plain copy of `bytes` bytes, one 16-byte non-temporal load and one 16-byte
// non-temporal store per lane, one workgroup per 4 KiB.  Of the patterns
// measured on 2 x 512 MiB (four or eight 16-byte loads per lane, grid-stride
// loops of 1024..4096 workgroups, plain or non-temporal loads and stores) this
// one ran fastest: 6.47 TB/s, against 5.83 TB/s for four loads per lane with
// plain loads and 5.25-5.34 TB/s for torch's device copy.  Timed by the same
// kernel timer as the codec kernels (kind 3).  No counterpart in the
// reference: a measurement helper, reading and writing exactly 2 x bytes.
#include <hip/hip_runtime.h>

#include "szx_amd.h"
#include "szx_internal.hpp"

namespace szxamd {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, uint64_t n16)
{
    const uint64_t j = blockIdx.x * 256ull + threadIdx.x;
    if (j < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[j]), &dst[j]);
}

}  // namespace szxamd

using namespace szxamd;

extern "C" int szx_amd_copy_dev(const void *d_src, void *d_dst, size_t bytes, void *hipStream)
{
    api_clear_error();
    if (!d_src || !d_dst || bytes == 0) {
        api_error("NULL pointer or bytes == 0");
        return SZx_NSCS;
    }
    if (((uintptr_t)d_src & 15) || ((uintptr_t)d_dst & 15) || (bytes & 15)) {
        api_error("pointers and bytes must be multiples of 16");
        return SZx_NSCS;
    }
    const uint64_t n16 = bytes / 16;
    const uint64_t grid = (n16 + 255) / 256;
    if (grid > 0x7FFFFFFFull) {
        api_error("copy of %zu bytes is too large for one launch", bytes);
        return SZx_NSCS;
    }
    hipStream_t st = static_cast<hipStream_t>(hipStream);
    timing_mark_begin(KK_COPY, st);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(256), 0, st, static_cast<const u32x4 *>(d_src),
                       static_cast<u32x4 *>(d_dst), n16);
    timing_mark_end(KK_COPY, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        api_error("launch: %s", hipGetErrorString(e));
        return SZx_NSCS;
    }
    return SZx_SCES;
}
