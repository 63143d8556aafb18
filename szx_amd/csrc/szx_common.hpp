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

struct ChunkState;
__device__ inline void publish_pair(ChunkState *t, unsigned long long flag, unsigned long long a,
                                    unsigned long long b);

// Two-level decoupled look-back, called by the 64 lanes of one wave.
//
// Tiles are grouped in chunks of 64.  A tile publishes its aggregate in one
// 64-bit word -- flag | constants (22 bits) << 40 | bytes (40 bits); both fit
// because a chunk holds 64 tiles -- and scans only the tiles before it in its
// own chunk (one window, no chain).  The chunk's last tile publishes the
// chunk total; every tile then scans the chunk totals before its chunk, again
// one window of 64 in the common case.  Inclusive prefixes therefore never
// have to travel through long chains of tiles, which matters because every
// poll is a round trip across the XCDs (~1 us).
//
// Tiles must start in index order (the hardware dispatches a grid's
// workgroups in increasing order), so every tile waited on has been
// dispatched and the loops end.  The waits are still bounded (LB_SPIN_LIMIT
// polls, a second or more): on overrun the look-back reports failure and the
// caller flags the call as failed instead of hanging.
struct ChunkState {
    unsigned long long w[2];   // flag | bytes, flag | constants
};

__device__ inline void publish_pair(ChunkState *t, unsigned long long flag, unsigned long long a,
                                    unsigned long long b)
{
    store_relaxed(&t->w[1], flag | b);
    store_relaxed(&t->w[0], flag | a);
}

constexpr int LB_CHUNK = 64;
constexpr unsigned long long T_BYTES = (1ull << 40) - 1;

// Window accept/sum step over NL lanes' words; returns true when accepted.
// `stop` is the lowest lane (nearest predecessor) holding an inclusive prefix
// or 64 if none.
__device__ inline bool window_ready(bool ready, bool isinc, int lane, int &stop)
{
    const uint64_t nr = __ballot(!ready);
    const uint64_t inc = __ballot(ready && isinc);
    const uint64_t low = inc & (~inc + 1);
    const uint64_t upto = inc ? (low | (low - 1)) : ~0ull;
    stop = inc ? __builtin_ctzll(inc) : 64;
    (void)lane;
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

__device__ inline bool lookback_2level(unsigned long long *tiles, ChunkState *chunks, uint64_t tile,
                                       uint64_t ntiles, unsigned long long agg_bytes,
                                       unsigned long long agg_cnt, unsigned long long &ex_bytes,
                                       unsigned long long &ex_cnt)
{
    const int lane = __lane_id();
    const uint64_t chunk = tile / LB_CHUNK;
    const int li = (int)(tile % LB_CHUNK);
    const uint64_t c0 = chunk * LB_CHUNK;

    // 1. within the chunk
    unsigned long long in_b = 0, in_c = 0;
    if (li == 0) {
        if (lane == 0) store_relaxed(&tiles[tile], FLAG_INC | (agg_cnt << 40) | agg_bytes);
    } else {
        if (lane == 0) store_relaxed(&tiles[tile], FLAG_AGG | (agg_cnt << 40) | agg_bytes);
        int round = 0;
        while (true) {
            // lane l looks at tile (tile - 1 - l); lanes beyond the chunk start see
            // a virtual inclusive prefix of zero
            const int j = li - 1 - lane;
            const unsigned long long w = j >= 0 ? load_relaxed(&tiles[c0 + j]) : FLAG_INC;
            const unsigned long long f = w & FLAG_MASK;
            int stop;
            if (!window_ready(f != 0, f == FLAG_INC, lane, stop)) {
                if (!poll_backoff(round)) return false;
                continue;
            }
            const unsigned long long v = lane <= stop ? (w & VAL_MASK) : 0ull;
            in_b = wave_sum(v & T_BYTES);
            in_c = wave_sum(v >> 40);
            break;
        }
        if (lane == 0)
            store_relaxed(&tiles[tile], FLAG_INC | ((in_c + agg_cnt) << 40) | (in_b + agg_bytes));
    }
    const bool last_in_chunk = li == LB_CHUNK - 1 || tile == ntiles - 1;

    // 2. across chunks
    unsigned long long cx_b = 0, cx_c = 0;
    if (chunk == 0) {
        if (last_in_chunk && lane == 0)
            publish_pair(&chunks[0], FLAG_INC, in_b + agg_bytes, in_c + agg_cnt);
    } else {
        if (last_in_chunk && lane == 0)
            publish_pair(&chunks[chunk], FLAG_AGG, in_b + agg_bytes, in_c + agg_cnt);
        int64_t base = (int64_t)chunk - 1;
        int round = 0;
        while (true) {
            const int64_t j = base - lane;
            unsigned long long wa = FLAG_INC, wb = FLAG_INC;
            if (j >= 0) {
                wa = load_relaxed(&chunks[j].w[0]);
                wb = load_relaxed(&chunks[j].w[1]);
            }
            const unsigned long long fa = wa & FLAG_MASK, fb = wb & FLAG_MASK;
            int stop;
            if (!window_ready(fa != 0 && fa == fb, fa == FLAG_INC && fb == FLAG_INC, lane, stop)) {
                if (!poll_backoff(round)) return false;
                continue;
            }
            cx_b += wave_sum(lane <= stop ? (wa & VAL_MASK) : 0ull);
            cx_c += wave_sum(lane <= stop ? (wb & VAL_MASK) : 0ull);
            if (stop < 64) break;
            base -= 64;
        }
        if (last_in_chunk && lane == 0)
            publish_pair(&chunks[chunk], FLAG_INC, cx_b + in_b + agg_bytes, cx_c + in_c + agg_cnt);
    }
    ex_bytes = cx_b + in_b;
    ex_cnt = cx_c + in_c;
    return true;
}

}  // namespace szxamd
