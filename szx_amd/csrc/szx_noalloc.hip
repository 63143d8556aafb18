// szx_noalloc.hip -- szx_amd_compress_dev_noalloc: the device compress with
// no workspace argument and no allocation, for callers that cannot call
// hipMalloc (code under HIP graph capture, serving loops that forbid it).
//
// The compressor's workspace (look-back words, chunk states, totals, REL
// range partials, bitmap and constant-median scratch, record-size table) lives
// in a file-scope device array, `scratch`, of SZX_AMD_NOALLOC_SCRATCH_BYTES.
// Every GPU the library runs on holds its own copy at its own address until
// the process ends.  In host code the array's name is only a stand-in, so its
// device address is looked up once per device (hipGetSymbolAddress, with that
// device current) on the entry point's first call there, and only that
// address is handed to kernels.
//
// Calls on one GPU share its array: the compress clears the words that must
// start at zero with a kernel (k_zero_u64, part of every compress launch),
// and two calls that overlap in time on one GPU race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "szx_amd.h"
#include "szx_internal.hpp"

namespace szxamd {

static __device__ __align__(256) uint8_t scratch[SZX_AMD_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per
// device); nullptr with the error text set on failure.
void *scratch_on_current_device()
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        api_error("hipGetDevice: %s", hipGetErrorString(e));
        return nullptr;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        api_error("device %d is beyond the %d devices the scratch table holds", dev, kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            api_error("hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

}  // namespace szxamd

using namespace szxamd;

extern "C" int szx_amd_compress_dev_noalloc(const void *d_in, size_t n, int dataType, int errBoundMode,
                                            float absErrBound, float relBoundRatio, int blockSize, void *d_out,
                                            size_t outCapacity, uint64_t *d_result, uint64_t *d_index,
                                            void *hipStream)
{
    api_clear_error();
    const int w = dataType == SZx_FLOAT ? 4 : dataType == SZx_DOUBLE ? 8 : 0;
    if (!w) {
        api_error("unsupported dataType %d", dataType);
        return SZx_TERR;
    }
    if (!bs_valid(blockSize)) {
        api_error("unsupported blockSize %d", blockSize);
        return SZx_NSCS;
    }
    if (errBoundMode != ABS && errBoundMode != REL) {
        api_error("bad errBoundMode");
        return SZx_BERR;
    }
    if (!d_in || !d_out || !d_result || n == 0) {
        api_error("NULL argument or n == 0");
        return SZx_NSCS;
    }
    if (((uintptr_t)d_in & 15) != 0) {
        api_error("d_in must be 16-byte aligned");
        return SZx_NSCS;
    }
    const CompressLayout L = compress_layout(n, w, blockSize);
    if (L.ws_bytes > SZX_AMD_NOALLOC_SCRATCH_BYTES) {
        api_error("%zu values at blockSize %d need %llu bytes of workspace, more than the %u-byte scratch "
                  "array; use szx_amd_compress_dev with a workspace",
                  n, blockSize, (unsigned long long)L.ws_bytes, (unsigned)SZX_AMD_NOALLOC_SCRATCH_BYTES);
        return SZx_MERR;
    }
    if (outCapacity < L.capacity) {
        api_error("outCapacity < %llu", (unsigned long long)L.capacity);
        return SZx_NSCS;
    }
    uint8_t *ws = static_cast<uint8_t *>(scratch_on_current_device());
    if (!ws) return SZx_NSCS;
    // the HIP events of the benchmark timer are created on demand: keep them
    // out of this call
    const bool timing = timing_set(false);
    const int rc = compress_launch(d_in, n, w, errBoundMode == REL, absErrBound, relBoundRatio, blockSize,
                                   (uint8_t *)d_out, ws, d_result, d_index, (hipStream_t)hipStream);
    timing_set(timing);
    if (rc) {
        api_error("compress launch failed");
        return SZx_NSCS;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        api_error("launch: %s", hipGetErrorString(e));
        return SZx_NSCS;
    }
    return SZx_SCES;
}
