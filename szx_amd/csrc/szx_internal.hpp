// szx_internal.hpp -- host-side interface between the C ABI (szx_api.cpp)
// and the kernel launchers (szx_compress.hip, szx_decompress.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace szxamd {

// Device memory layout of one compress call.
struct CompressLayout {
    uint64_t nb;          // blocks
    uint64_t nbit;        // bitmap bytes
    uint64_t region;      // bytes reserved in d_out in front of the payload
    uint64_t payload_max; // largest possible payload
    uint64_t capacity;    // bytes d_out must hold
    uint64_t ntiles;
    uint64_t ws_tiles, ws_misc, ws_parts, ws_bitmap, ws_cmed, ws_bytes;
};

CompressLayout compress_layout(uint64_t n, int w, int bs);

// Enqueue a whole compress on `st`.  d_result receives
// {stream offset in d_out, stream bytes, bound bits}.  Returns <0 for an
// unsupported block size.
int compress_launch(const void *d_in, uint64_t n, int w, int rel_mode, float abs_eb, float rel,
                    int bs, uint8_t *d_out, uint8_t *d_ws, uint64_t *d_result, hipStream_t st);

// Device memory layout of one decompress call.
struct DecompressLayout {
    uint64_t nb_max;      // blocks for the smallest supported block size
    uint64_t ws_hdr, ws_info, ws_bytes;
};

DecompressLayout decompress_layout(uint64_t n, int w, int bs_hint, uint64_t stream_bytes);

// Enqueue a whole decompress on `st`; d_status receives 0 or a negative code.
int decompress_launch(const uint8_t *d_stream, uint64_t stream_bytes, uint64_t n, int w,
                      void *d_out, uint8_t *d_ws, uint64_t ws_bytes, int exact, int32_t *d_status,
                      hipStream_t st);

}  // namespace szxamd
