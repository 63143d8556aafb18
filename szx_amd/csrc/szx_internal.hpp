// szx_internal.hpp -- host-side interface between the C ABI (szx_api.cpp)
// and the kernel launchers (szx_compress.hip, szx_decompress.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace szxamd {

// Values per tile (one workgroup of the compressor, and the unit of the
// record index shared by the compressor and the decompressor).
constexpr int TILE_VALUES_F32 = 4096;
constexpr int TILE_VALUES_F64 = 2048;
inline uint64_t tile_values(int w) { return w == 4 ? TILE_VALUES_F32 : TILE_VALUES_F64; }
// The compressor encodes several index tiles per workgroup (longer tiles
// amortise its look-back wait): INDEX_PER_TILE_WIDE for blocks of 16..128
// values (16384 fp32 / 8192 fp64 values, 64 KiB), INDEX_PER_TILE otherwise
// and in the batched launch; its record index stays per index tile.
constexpr int INDEX_PER_TILE = 2;
constexpr int INDEX_PER_TILE_WIDE = 4;
constexpr bool compress_wide(int w, int bs) { return w > 0 && bs >= 16 && bs <= 128; }
inline uint64_t compress_tile_values(int w, int bs)
{
    return (compress_wide(w, bs) ? INDEX_PER_TILE_WIDE : INDEX_PER_TILE) * tile_values(w);
}
// Record index: per tile {payload offset of its first record, constant
// blocks before it} plus a closing pair (u64), then per block a u16 -- the
// record's start inside its tile's payload bytes (non-constant) or its rank
// among the tile's constant blocks.
inline uint64_t index_tiles(uint64_t n, int w)
{
    const uint64_t t = (n + tile_values(w) - 1) / tile_values(w);
    return t ? t : 1;
}
// Block sizes: the tiled kernels take the powers of two 4..256; every other
// size 1..255 runs on the per-block kernels (szx_generic.hip).
inline bool bs_valid(int bs) { return bs >= 1 && bs <= 256; }
inline bool bs_tiled(int bs) { return bs >= 4 && bs <= 256 && (bs & (bs - 1)) == 0; }
// For the other block sizes the record index is per block: {payload offset
// of its record (or of the next record, for a constant block), constant
// blocks before it}, plus the closing pair {payload bytes, constant blocks}.
inline uint64_t index_words(uint64_t n, int w, int bs)
{
    const uint64_t nb = (n + bs - 1) / bs;
    if (!bs_tiled(bs)) return 2 * (nb + 1);
    return 2 * (index_tiles(n, w) + 1) + (nb + 3) / 4;
}

// Optional kernel timing: when enabled on this thread, the launchers bracket
// each main kernel (compress, decode, index build) with a pair of HIP events
// on the launch stream.
enum KernelKind { KK_COMPRESS = 0, KK_DECODE = 1, KK_INDEX = 2, KK_COPY = 3, KK_COUNT = 4 };
// Error text of the calling thread (szx_amd_last_error).
void api_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void api_clear_error();

// Turn this thread's kernel timing on or off; returns the previous setting.
bool timing_set(bool on);
void timing_mark_begin(int kind, hipStream_t st);
void timing_mark_end(int kind, hipStream_t st);

// Device memory layout of one compress call.
struct CompressLayout {
    uint64_t nb;          // blocks
    uint64_t nbit;        // bitmap bytes
    uint64_t region;      // bytes reserved in d_out in front of the payload
    uint64_t payload_max; // largest possible payload
    uint64_t capacity;    // bytes d_out must hold
    uint64_t ntiles;      // compressor tiles
    uint64_t nidx;        // index tiles
    uint64_t ws_tiles, ws_chunks, ws_cacc, ws_misc, ws_parts, ws_bitmap, ws_cmed, ws_bytes;
    // per-block kernels only: record sizes, scan chunk sums, per-block pairs
    uint64_t ws_gsize, ws_gparts, ws_gidx;
    uint64_t ws_osz;      // random-access container: u16 record sizes
};

CompressLayout compress_layout(uint64_t n, int w, int bs);

// Enqueue a whole compress on `st`.  d_result receives
// {stream offset in d_out, stream bytes, bound bits}; d_index (optional)
// receives the record index.  fast_mode 2 writes the blocked stream, 3 and 4
// the random-access container (tiled block sizes only).  Returns <0 for an
// unsupported block size or mode.
int compress_launch(const void *d_in, uint64_t n, int w, int rel_mode, float abs_eb, float rel,
                    int bs, uint8_t *d_out, uint8_t *d_ws, uint64_t *d_result, uint64_t *d_index,
                    hipStream_t st, int fast_mode = 2);

// Value range (computeValueRange_*, szx_float.cc:478-499): d_minmax receives
// {min, max} as T; d_ws holds range_workspace_bytes().
uint64_t range_workspace_bytes();
void range_launch(const void *d_in, uint64_t n, int w, uint8_t *d_ws, void *d_minmax, hipStream_t st);
// The same pass with the REL bound fl(range * rel) in *eb_out as well.
void range_bound_launch(const void *d_in, uint64_t n, int w, float rel, uint8_t *d_ws, float *eb_out,
                        void *d_minmax, hipStream_t st);

// Non-blocked codec (fastMode 1, szx_nonblocked.hip): one record over all
// values against the global median.
// Reader of the reference's CUDA codec (cuSZx) stream, fp32 (szx_cuszx.hip).
uint64_t cuszx_workspace(uint64_t n);
void cuszx_decompress_launch(const uint8_t *d_stream, uint64_t len, uint64_t n, float *d_out, uint8_t *d_ws,
                             int32_t *d_status, hipStream_t st);
// The writer: bs a multiple of 32 in 32..224; d_out holds cuszx_capacity(n,
// bs) bytes; d_result receives {0, stream bytes, bound bits, 1 if a block
// needs more than 32 significant bits (no stream) else 0}.
uint64_t cuszx_capacity(uint64_t n, int bs);
uint64_t cuszx_compress_workspace(uint64_t n, int bs);
void cuszx_compress_launch(const float *d_in, uint64_t n, float eb, int bs, uint8_t *d_out, uint8_t *d_ws,
                           uint64_t *d_result, hipStream_t st);

uint64_t nb_capacity(uint64_t n, int w);
uint64_t nb_workspace(uint64_t n);
void nb_compress_launch(const void *d_in, uint64_t n, int w, int rel_mode, float abs_eb, float rel, uint8_t *d_out,
                        uint8_t *d_ws, uint64_t *d_result, hipStream_t st);
void nb_decompress_launch(const uint8_t *d_stream, uint64_t len, uint64_t n, int w, void *d_out, uint8_t *d_ws,
                          int exact, int32_t *d_status, hipStream_t st);
// d_x[i] += d_pred[i] (float), the prediction step of SZx_fast_decompress_pred.
void add_pred_launch(float *d_x, const float *d_pred, uint64_t n, hipStream_t st);

// Checked header of a stream (k_check_header): block size, blocks, bitmap
// bytes, constant blocks, first payload byte.
struct DecodeHeader {
    uint64_t bs, nb, nbit, nconst, q0;
    int64_t rc;   // the header check's own status (0 = well-formed)
    uint64_t qbit, qmed;   // stream offsets of the bitmap and the constant medians
    int64_t medle;         // 1: medians in native (little-endian) order
};

// Device memory layout of one decompress call.
struct DecompressLayout {
    uint64_t ws_hdr, ws_index, ws_build, ws_bytes;
};

DecompressLayout decompress_layout(uint64_t n, int w, int bs, uint64_t stream_bytes);

// Per-block kernels for the block sizes the tiled kernels do not take
// (szx_generic.hip).
struct GenericCompress {
    const void *x;
    uint64_t n, nb;
    int bs;
    float eb;
    const float *eb_dev;  // REL: bound computed on the device
    uint8_t *payload;     // records
    uint8_t *bitmap;
    uint32_t *cmed;       // constant medians, big-endian, by rank
    uint64_t *totals;     // {payload bytes, constant blocks}
    uint32_t *size;       // per block: record bytes, 0 = constant
    uint64_t *idx;        // per-block pairs and the closing pair
    uint64_t *parts;      // 2 words per scan chunk
    uint64_t nparts;
};
uint64_t generic_parts(uint64_t nb);
void generic_compress_launch(const GenericCompress &g, int w, hipStream_t st);
void generic_scan_launch(const GenericCompress &g, hipStream_t st);

struct GenericDecode {
    const uint8_t *s;
    uint64_t len, n;
    const DecodeHeader *h;
    const uint64_t *idx;  // per-block pairs and the closing pair
    int32_t *status;
    void *out;
    int exact;
};
void generic_decode_launch(const GenericDecode &p, int w, int bs, uint64_t nb, hipStream_t st);

// Enqueue a whole decompress on `st`; d_status receives 0 or a negative code.
// d_index (optional) is the record index the compressor wrote; without it the
// index is rebuilt from the stream.
// fast_mode 2 reads the blocked stream, 3 and 4 the random-access container
// (its size table gives the record index; tiled block sizes only).
int decompress_launch(const uint8_t *d_stream, uint64_t stream_bytes, uint64_t n, int w, int bs,
                      void *d_out, uint8_t *d_ws, uint64_t ws_bytes, const uint64_t *d_index,
                      int exact, int32_t *d_status, hipStream_t st, int fast_mode = 2);

}  // namespace szxamd
