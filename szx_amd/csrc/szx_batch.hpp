// szx_batch.hpp -- parameter blocks of the compressor (one field) and of the
// batched launch (many fields), shared by the kernels and the host planner.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace szxamd {

struct ChunkState;   // look-back chunk words (szx_common.hpp)

// One field's compressor parameters (device pointers).
struct CompressParams {
    const void *x;
    uint64_t n;
    float eb;             // absolute bound (ABS)
    const float *eb_dev;  // REL: bound computed on the device, else nullptr
    uint8_t *payload;     // first record byte
    uint8_t *bitmap;      // scratch bitmap, nbit bytes
    uint32_t *cmed;       // scratch constant medians, big-endian words
    unsigned long long *tiles;   // per-tile look-back words
    ChunkState *chunks;          // per-64-tile look-back words
    unsigned long long *cacc;    // per-64-tile aggregate accumulators
    uint64_t *totals;     // [0] payload bytes, [1] constant blocks
    uint64_t *index;      // optional record index: per tile {payload offset, constants before}
    uint16_t *bidx;       // its per-block part (start in tile, or constant rank), or nullptr
    int *err;             // set when a look-back wait overran its bound
    uint64_t ntiles;
    uint64_t nbit;
    uint64_t nb;
    uint64_t nidx;        // index tiles
    uint16_t *osz;        // random-access container: record size by non-constant rank, or nullptr
    int fm4;              // fastMode-4 fp64 block statistics (radius rounded to float first)
    // REL with the range pass fused into the launch: the pass's per-workgroup
    // {min, max} (RANGE_LB_PARTS pairs of the field's type), folded into the
    // bound by every workgroup; nullptr otherwise
    const void *rparts;
    float rel;            // the bound ratio
    float *eb_out;        // where tile 0 leaves the bound for the assembly
};

// One field's header / bitmap / median assembly (device pointers).
struct AssembleParams {
    uint8_t *out;
    uint64_t region;      // bytes reserved in front of the payload
    const uint8_t *bitmap;
    const uint32_t *cmed;
    const uint64_t *totals;
    uint64_t nbit;
    int bs;
    const float *eb_dev;
    float eb;
    uint64_t *result;     // [0] stream offset, [1] stream size, [2] bound bits, [3] error
    const int *err;
    int ra;               // 0: fastMode-2 header; 3 / 4: random-access container
    uint64_t nb;          // blocks (random-access container)
    const uint16_t *osz;  // random-access container: record sizes
};

// Workgroups (of 1024 lanes) of the single-field REL range pass whose
// partials the compressor folds itself.
constexpr int RANGE_LB_PARTS = 256;

// Values per REL range item of the batched range pass.
constexpr uint64_t RANGE_ITEM = 1ull << 20;
// Words per block of the batched clear; bytes per thread of the batched
// header/bitmap/median assembly.
constexpr uint64_t ZERO_ITEM = 1024;
constexpr uint64_t ASSEMBLE_BYTES_PER_THREAD = 8;
constexpr uint64_t ASSEMBLE_MAX_BLOCKS = 512;

// One field of a batched launch.  The look-back state is sized for the
// smallest tiles (one index tile per compressor tile); ntiles[k] is the
// compressor tile count with 1, 2 or 4 index tiles per tile (k = 0, 1, 2).
struct BatchField {
    CompressParams p;
    AssembleParams a;
    ChunkState *chunks;
    uint64_t ntiles[3];
    uint64_t zwords;   // words cleared before a compress, from p.tiles (look-back words, totals, error)
    float relv;   // REL ratio (rel != 0)
    int rel;
    int w;        // 4 or 8
};

// Batched kernels (szx_compress.hip), all reading device copies of the
// field table.  tbase: first global tile of each field (nf + 1 entries) for
// the chosen tile shape; rbase: first range item of each field (nf + 1).
void batch_range_launch(const BatchField *bf, const uint64_t *rbase, int nf, uint64_t nitems, double *part,
                        hipStream_t st);
// tmap: the field of every tile of the global tile list (total).
int batch_compress_launch(int bs, int lv, int ipt, const BatchField *bf, const uint64_t *tbase,
                          const uint32_t *tmap, uint64_t total, hipStream_t st);
// zbase / abase: first block of each field in the flattened clear / assembly
// grids (nf + 1 entries; the last is the grid size).
void batch_zero_launch(const BatchField *bf, const uint64_t *zbase, int nf, uint64_t nitems, hipStream_t st);
void batch_assemble_launch(const BatchField *bf, const uint64_t *abase, int nf, uint64_t nitems, hipStream_t st);

struct DecodeHeader;

// One field of a batched decode (the stream sits at out + result[0]).
struct BatchDecodeField {
    const uint8_t *out;
    const uint64_t *result;
    uint64_t n;
    const uint64_t *index;   // record index written by the batched compress
    void *dst;
    int32_t *status;
    DecodeHeader *h;         // workspace
    uint64_t ntidx;          // index tiles
    int w;
};

// Batched decode (szx_decompress.hip): header checks, every tile's bytes
// (into recs, batch_decode_records_bytes(total) bytes), then every index tile
// of every field.  tbase: first global index tile of each field (nf + 1);
// tmap: the field of every global index tile (total).
uint64_t batch_decode_records_bytes(uint64_t total);
int batch_decode_launch(int bs, const BatchDecodeField *f, const uint64_t *tbase, const uint32_t *tmap,
                        void *recs, int nf, uint64_t total, int exact, hipStream_t st);

}  // namespace szxamd
