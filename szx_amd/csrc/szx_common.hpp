// szx_common.hpp -- device/host helpers shared by the gfx950 codec kernels.
//
// The codec is exact integer and bit work around a handful of IEEE float
// operations; everything here is written so that the GPU performs the same
// operations, in the same precision and rounding, as the reference CPU codec
// (no FMA contraction: the library is compiled with -ffp-contract=off; float
// denormals are preserved: -fno-gpu-flush-denormals-to-zero).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace szxamd {

// Word traits: the codec works on the raw bits of each value.
template <typename T> struct Word;
template <> struct Word<float> {
    using U = uint32_t;
    static constexpr int W = 4;
    static constexpr int LEN_LO = 9;     // szx_float.cc:1711-1712
    static constexpr int LEN_BASE = 10;  // 9 + radExpo - reqExpo + 1, szx_float.cc:1710
};
template <> struct Word<double> {
    using U = uint64_t;
    static constexpr int W = 8;
    static constexpr int LEN_LO = 12;    // szx_double.cc:2024-2025
    static constexpr int LEN_BASE = 12;  // szx_double.cc:2023
};

__host__ __device__ inline uint32_t f32_bits(float v) { return __builtin_bit_cast(uint32_t, v); }
__host__ __device__ inline float bits_f32(uint32_t u) { return __builtin_bit_cast(float, u); }
__host__ __device__ inline uint64_t f64_bits(double v) { return __builtin_bit_cast(uint64_t, v); }
__host__ __device__ inline double bits_f64(uint64_t u) { return __builtin_bit_cast(double, u); }

template <typename T> __host__ __device__ inline typename Word<T>::U to_bits(T v);
template <> __host__ __device__ inline uint32_t to_bits<float>(float v) { return f32_bits(v); }
template <> __host__ __device__ inline uint64_t to_bits<double>(double v) { return f64_bits(v); }
template <typename T> __host__ __device__ inline T from_bits(typename Word<T>::U u);
template <> __host__ __device__ inline float from_bits<float>(uint32_t u) { return bits_f32(u); }
template <> __host__ __device__ inline double from_bits<double>(uint64_t u) { return bits_f64(u); }

__host__ __device__ inline uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xFF00u) | ((v << 8) & 0xFF0000u) | (v << 24);
}

// Unbiased exponent field of the float radius (getExponent_float,
// szx_BytesToolkit.cc:335-346): a plain bit extract, so subnormals give -127.
__host__ __device__ inline int rad_expo(float r) { return (int)((f32_bits(r) >> 23) & 0xFF) - 127; }

// Unbiased exponent field of the error bound widened to double
// (getPrecisionReqLength_double, szx_BytesToolkit.cc:375-387), computed from
// the float bits with integer operations only so that no float->double
// conversion (and no denormal mode) is involved.  A float subnormal m*2^-149
// is a normal double with exponent floor(log2 m) - 149.
__host__ __device__ inline int req_expo(float eb)
{
    uint32_t u = f32_bits(eb) & 0x7FFFFFFFu;
    uint32_t e = u >> 23;
    if (e == 0xFF) return 1024;
    if (e != 0) return (int)e - 127;
    if (u == 0) return -1023;
    int lz = __builtin_clz(u);
    return (31 - lz) - 149;
}

// reqLength of a non-constant block (szx_float.cc:1708-1717,
// szx_double.cc:2020-2031).  clamped = 1 when the length hit 8*W; the stored
// median is then 0.
template <typename T>
__host__ __device__ inline int req_length(int radexp, int reqexp, int &clamped)
{
    int len = Word<T>::LEN_BASE + radexp - reqexp;
    clamped = 0;
    if (len < Word<T>::LEN_LO) len = Word<T>::LEN_LO;
    if (len > 8 * Word<T>::W) { len = 8 * Word<T>::W; clamped = 1; }
    return len;
}

// Stream layout constants (szx_float.cc:1091-1104).
constexpr int HEADER_BYTES = 13;

__host__ __device__ inline size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// Decoupled look-back state for single-pass tile scans.
//
// Every published word is self-contained (a 2-bit flag -- aggregate or
// inclusive prefix -- packed with its value), so words are written and read
// with RELAXED agent-scope atomics only: on gfx950 an agent-scope release
// costs an L2 writeback (buffer_wbl2) and an acquire an L2 invalidate
// (buffer_inv), because the 8 XCDs have separate L2s.  Where two words
// describe one entry, a reader accepts the entry only when both carry the
// same flag.
// ---------------------------------------------------------------------------

constexpr unsigned long long FLAG_AGG = 1ull << 62;
constexpr unsigned long long FLAG_INC = 2ull << 62;
constexpr unsigned long long FLAG_MASK = 3ull << 62;
constexpr unsigned long long VAL_MASK = (1ull << 62) - 1;

__device__ inline unsigned long long load_relaxed(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store_relaxed(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// Two-level decoupled look-back (compressor tiles over one field).
//
// Tiles are grouped in chunks of LB_CHUNK.  As soon as a tile knows its
// aggregate (record bytes, constant blocks) it publishes it in one 64-bit
// word -- flag | constants << 40 | bytes -- and adds it, with a count, to its
// chunk's accumulator; the tile that completes the count publishes the chunk
// aggregate.  A tile's exclusive prefix is the sum of the aggregates before
// it in its chunk (one window of at most 63 words) plus its chunk's prefix,
// found by scanning the chunk words backwards to the nearest inclusive value;
// the first tile of each chunk then publishes that prefix as the previous
// chunk's inclusive value.
//
// No tile waits for another tile's look-back, only for aggregates, which
// every tile publishes before its own look-back, so a caller can issue the
// first reads (lookback_poll) right after publishing, do other work, and only
// then finish (lookback_finish): the round trips to the other XCDs overlap
// that work.  A tile waits only for lower-numbered tiles, which start first
// (in-order dispatch), so every wait ends; the waits are bounded all the same
// (LB_SPIN_LIMIT polls) and report failure instead of hanging.
// A chunk's aggregate and its inclusive prefix live in separate word pairs,
// each written once: the inclusive prefix of chunk j can be known (by the
// first tile of chunk j + 1, from chunk j's tile words) before the tile that
// completes chunk j has published the aggregate, and one shared pair could
// then end with one word of each kind.
struct ChunkState {
    unsigned long long agg[2];   // FLAG_AGG | bytes, FLAG_AGG | constants
    unsigned long long inc[2];   // FLAG_INC | bytes, FLAG_INC | constants
};

__device__ inline void publish_pair(unsigned long long *w, unsigned long long flag, unsigned long long a,
                                    unsigned long long b)
{
    store_relaxed(&w[1], flag | b);
    store_relaxed(&w[0], flag | a);
}

constexpr int LB_CHUNK = 64;
constexpr unsigned long long T_BYTES = (1ull << 40) - 1;

// Window accept/sum step over the 64 lanes' words; returns true when
// accepted.  `stop` is the lowest lane (nearest predecessor) holding an
// inclusive value, or 64 if none.
__device__ inline bool window_ready(bool ready, bool isinc, int &stop)
{
    const uint64_t nr = __ballot(!ready);
    const uint64_t inc = __ballot(ready && isinc);
    const uint64_t low = inc & (~inc + 1);
    const uint64_t upto = inc ? (low | (low - 1)) : ~0ull;
    stop = inc ? __builtin_ctzll(inc) : 64;
    return (nr & upto) == 0;
}

constexpr int LB_SPIN_LIMIT = 1 << 19;

// Back-off between polls of look-back words that are not ready yet: every
// poll is a memory-side round trip, and thousands of waiting waves polling
// flat out would compete with the data traffic.  False once the wait has
// overrun its bound.
__device__ inline bool poll_backoff(int &round)
{
    if (round < 2) __builtin_amdgcn_s_sleep(2);
    else if (round < 6) __builtin_amdgcn_s_sleep(16);
    else __builtin_amdgcn_s_sleep(64);
    return ++round <= LB_SPIN_LIMIT;
}

__device__ inline unsigned long long wave_sum(unsigned long long v)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

struct LookbackRefs {
    unsigned long long *tiles;   // per tile: flag | constants << 40 | bytes
    ChunkState *chunks;          // per chunk: aggregate and inclusive prefix
    unsigned long long *cacc;    // per chunk: bytes | constants << 32 | tiles << 56
    uint64_t ntiles;
};

// Publish a tile's aggregate (one lane) and add it to its chunk's
// accumulator.  Returns the accumulator's previous value, which the caller
// hands to lookback_chunk_done once convenient: waiting for the atomic's
// round trip right away would stall the workgroup.  Chunk sums fit their
// fields: a chunk holds 64 tiles of at most 2^16 record bytes and 2^11
// constant blocks each.
__device__ inline unsigned long long lookback_publish(const LookbackRefs &L, uint64_t tile, uint32_t bytes,
                                                      uint32_t consts)
{
    store_relaxed(&L.tiles[tile], FLAG_AGG | ((unsigned long long)consts << 40) | bytes);
    const unsigned long long mine =
        (unsigned long long)bytes | ((unsigned long long)consts << 32) | (1ull << 56);
    return atomicAdd(&L.cacc[tile / LB_CHUNK], mine);
}

// The tile that completes its chunk's count publishes the chunk aggregate
// (one lane; `old` from lookback_publish for the same tile and aggregate).
__device__ inline void lookback_chunk_done(const LookbackRefs &L, uint64_t tile, unsigned long long old,
                                           uint32_t bytes, uint32_t consts)
{
    const uint64_t c = tile / LB_CHUNK;
    const uint64_t left = L.ntiles - c * LB_CHUNK;
    const uint64_t count = left < (uint64_t)LB_CHUNK ? left : (uint64_t)LB_CHUNK;
    if ((old >> 56) + 1 == count) {
        const unsigned long long sum =
            old + ((unsigned long long)bytes | ((unsigned long long)consts << 32));
        publish_pair(L.chunks[c].agg, FLAG_AGG, sum & 0xFFFFFFFFull, (sum >> 32) & 0xFFFFFFull);
    }
}

// The words a tile's look-back reads, loaded by the 64 lanes of a wave: lane
// l holds tile (chunk start + l) for l below the tile's place in its chunk;
// the tile words of the LB_PREV chunks before it (tile l of each); and, for
// l < LB_WIN, chunk (c - 1 - LB_PREV - l) (an inclusive zero before chunk 0).
//
// The preceding chunks are read tile by tile because a chunk's own aggregate
// appears only after its last tile's accumulator atomic has come back and
// that tile has published it: for the chunks just before a tile that is
// several microseconds after their tiles' aggregates, while each tile word is
// visible about one memory round trip after its tile published it.  Chunk
// words are used from LB_PREV + 1 chunks back, where they are long published.
constexpr int LB_PREV = 4;

struct LookbackPoll {
    unsigned long long t, ca, cb, ia, ib;
    unsigned long long p[LB_PREV];
};

// Chunk words are read LB_WIN at a time (the nearest inclusive value is
// normally a few chunks back); lanes past the window hold zeros that neither
// stop nor add to the scan.
constexpr int LB_WIN = 16;

__device__ inline void lookback_chunk_words(const LookbackRefs &L, int64_t j, LookbackPoll &q)
{
    const bool in = __lane_id() < LB_WIN;
    q.ca = !in ? FLAG_AGG : j >= 0 ? load_relaxed(&L.chunks[j].agg[0]) : 0ull;
    q.cb = !in ? FLAG_AGG : j >= 0 ? load_relaxed(&L.chunks[j].agg[1]) : 0ull;
    q.ia = !in ? 0ull : j >= 0 ? load_relaxed(&L.chunks[j].inc[0]) : FLAG_INC;
    q.ib = !in ? 0ull : j >= 0 ? load_relaxed(&L.chunks[j].inc[1]) : FLAG_INC;
}

// Tile word `lane` of chunk j (a ready zero for j < 0).
__device__ inline unsigned long long prev_tile_word(const LookbackRefs &L, int64_t j, int lane)
{
    return j >= 0 ? load_relaxed(&L.tiles[(uint64_t)j * LB_CHUNK + lane]) : FLAG_AGG;
}

// Issue a tile's look-back reads (one wave).  Calling it again later refreshes
// them: a word that was ready stays ready, with the same value or a later
// inclusive one.
__device__ inline LookbackPoll lookback_poll(const LookbackRefs &L, uint64_t tile)
{
    const int lane = __lane_id();
    const uint64_t c = tile / LB_CHUNK;
    const int li = (int)(tile % LB_CHUNK);
    LookbackPoll q;
    q.t = lane < li ? load_relaxed(&L.tiles[c * LB_CHUNK + lane]) : FLAG_AGG;
#pragma unroll
    for (int k = 0; k < LB_PREV; k++) q.p[k] = prev_tile_word(L, (int64_t)c - 1 - k, lane);
    lookback_chunk_words(L, (int64_t)c - 1 - LB_PREV - lane, q);
    return q;
}

// Finish a tile's look-back from its polled words (all 64 lanes of a wave;
// every lane returns the same prefix).  `publish`: this wave records the
// chunk prefix of a chunk-leading tile.  False if a wait overran its bound.
__device__ inline bool lookback_finish(const LookbackRefs &L, uint64_t tile, LookbackPoll q,
                                       bool publish, unsigned long long &ex_bytes,
                                       unsigned long long &ex_cnt)
{
    const int lane = __lane_id();
    const uint64_t c = tile / LB_CHUNK;
    const int li = (int)(tile % LB_CHUNK);
    const uint64_t c0 = c * LB_CHUNK;

    // 1. the aggregates before the tile in its chunk and in the LB_PREV
    //    chunks before that
    int round = 0;
    while (true) {
        bool ready = (q.t & FLAG_MASK) != 0;
#pragma unroll
        for (int k = 0; k < LB_PREV; k++) ready = ready && (q.p[k] & FLAG_MASK) != 0;
        if (__ballot(!ready) == 0) break;
        if (!poll_backoff(round)) return false;
        if ((q.t & FLAG_MASK) == 0) q.t = load_relaxed(&L.tiles[c0 + lane]);
#pragma unroll
        for (int k = 0; k < LB_PREV; k++)
            if ((q.p[k] & FLAG_MASK) == 0) q.p[k] = prev_tile_word(L, (int64_t)c - 1 - k, lane);
    }
    unsigned long long lb = 0, lc = 0;
    const unsigned long long own = lane < li ? (q.t & VAL_MASK) : 0ull;
    lb += own & T_BYTES;
    lc += own >> 40;
#pragma unroll
    for (int k = 0; k < LB_PREV; k++) {
        const unsigned long long w = q.p[k] & VAL_MASK;
        lb += w & T_BYTES;
        lc += w >> 40;
    }
    const unsigned long long in_b = wave_sum(lb);
    const unsigned long long in_c = wave_sum(lc);

    // 2. the prefix of the chunks before those
    unsigned long long cx_b = 0, cx_c = 0;
    int64_t base = (int64_t)c - 1 - LB_PREV;
    round = 0;
    while (base >= 0) {
        const bool inc = (q.ia & FLAG_MASK) == FLAG_INC && (q.ib & FLAG_MASK) == FLAG_INC;
        const bool agg = (q.ca & FLAG_MASK) == FLAG_AGG && (q.cb & FLAG_MASK) == FLAG_AGG;
        int stop;
        if (!window_ready(inc || agg, inc, stop)) {
            if (!poll_backoff(round)) return false;
            lookback_chunk_words(L, base - lane, q);
            continue;
        }
        const unsigned long long vb = inc ? q.ia : q.ca, vc = inc ? q.ib : q.cb;
        cx_b += wave_sum(lane <= stop ? (vb & VAL_MASK) : 0ull);
        cx_c += wave_sum(lane <= stop ? (vc & VAL_MASK) : 0ull);
        if (stop < 64) break;
        base -= LB_WIN;
        lookback_chunk_words(L, base - lane, q);
    }
    ex_bytes = cx_b + in_b;
    ex_cnt = cx_c + in_c;
    // a chunk-leading tile's exclusive prefix is the previous chunk's
    // inclusive one
    if (publish && li == 0 && c > 0 && lane == 0) publish_pair(L.chunks[c - 1].inc, FLAG_INC, ex_bytes, ex_cnt);
    return true;
}

// ---- segment helpers shared by the per-block decoders ----

// Inclusive sum over the S lanes of a segment.
template <int S> __device__ inline uint32_t seg_incl_sum(uint32_t v, int l)
{
#pragma unroll
    for (int d = 1; d < S; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, S);
        if (l >= d) v += o;
    }
    return v;
}

// Carry of the top three bytes of the reconstructed words across the lanes
// of a segment: bytes 0..2 hold the bytes W-1, W-2, W-3 of the last word that
// wrote them, bits 24..26 say which were written.  combine(a, b) = b's
// written bytes over a's.
__device__ inline uint32_t carry_combine(uint32_t a, uint32_t b)
{
    const uint32_t f = b >> 24;
    const uint32_t keep = ((f & 1u) ? 0xFFu : 0u) | ((f & 2u) ? 0xFF00u : 0u) | ((f & 4u) ? 0xFF0000u : 0u);
    return ((b & keep) | (a & ~keep & 0xFFFFFFu)) | ((a | b) & 0x7000000u);
}

// Up to 8 staged bytes starting at LDS byte pos, little-endian.
__device__ inline uint64_t lds_window8(const uint32_t *w, uint32_t pos)
{
    const uint32_t i = pos >> 2, sh = pos & 3;
    const uint32_t a = w[i], b = w[i + 1], c = w[i + 2];
    return ((uint64_t)__builtin_amdgcn_alignbyte(c, b, sh) << 32) | __builtin_amdgcn_alignbyte(b, a, sh);
}
__device__ inline uint32_t lds_window4(const uint32_t *w, uint32_t pos)
{
    const uint32_t i = pos >> 2;
    return __builtin_amdgcn_alignbyte(w[i + 1], w[i], pos & 3);
}

}  // namespace szxamd
