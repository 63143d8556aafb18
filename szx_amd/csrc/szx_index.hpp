// szx_index.hpp -- parallel record-index build from a stream (szx_index.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szx_internal.hpp"

namespace szxamd {

constexpr uint32_t IX_CHUNK = 4096;   // payload bytes per chunk
constexpr uint32_t IX_RMAX = 2117;    // largest record: 5 + 64 + 256 * 8

struct IxParams {
    const uint8_t *s;
    uint64_t len, n;
    int bs, w, nl;
    int no_spec;        // speculative chains off (every chunk repaired in order)
    uint32_t rmax, maxr;
    uint64_t nch, ntiles, tile_blocks;
    const DecodeHeader *h;
    int32_t *status;
    uint32_t *spec;     // per chunk: speculative chain start | its exit << 16
    uint32_t *tex;      // per chunk: exit from its entry (bit 31: differs from the speculative one)
    uint32_t *nfail;    // chunks whose exit differs
    uint32_t *cnt;      // per chunk: records found
    uint64_t *cbase;    // per chunk: records before it in its group of 1024 chunks
    uint64_t *csum;     // per 1024 chunks: records before them
    uint16_t *pos;      // per chunk: record starts
    uint64_t *rec;      // per record: payload offset
    uint64_t *tbase;    // per tile: non-constant blocks before it in its group of 1024 tiles
    uint64_t *tsum;     // per 1024 tiles: non-constant blocks before them
    uint64_t *found;    // records found in total
    uint64_t *pend;     // payload end
    uint64_t *index;    // output: tile pairs
    uint16_t *bidx;     // output: per-block entries
};

struct IxLayout {
    uint64_t nch, ntiles, nl;
    uint32_t rmax, maxr;
    uint64_t o_spec, o_tex, o_cnt, o_cbase, o_pos, o_rec, o_tbase, o_tsum, o_csum, o_misc, o_osz, o_bsum, bytes;
};

IxLayout ix_layout(uint64_t n, int w, int bs, uint64_t len);

// Build the record index of a stream of len bytes (header already checked
// into *h).  ws holds ix_layout(...).bytes.
void index_build_launch(const uint8_t *s, uint64_t len, uint64_t n, int w, int bs, const DecodeHeader *h,
                        uint64_t *index, uint16_t *bidx, uint8_t *ws, int32_t *status, hipStream_t st);

// The same index for a random-access container, from its record size table.
void index_from_sizes_launch(const uint8_t *s, uint64_t len, uint64_t n, int w, int bs, const DecodeHeader *h,
                             uint64_t *index, uint16_t *bidx, uint8_t *ws, int32_t *status, hipStream_t st);

}  // namespace szxamd
