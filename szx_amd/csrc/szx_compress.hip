// szx_compress.hip -- single-pass blockwise fast compressor for gfx950.
//
// One launch of k_compress turns the whole input into the records of the
// fastMode-2 stream (szx_float.cc:1059-1139, szx_double.cc:1160-1238):
//
//   * a workgroup (256 lanes, 4 wave64s) owns a tile of 16384 fp32 / 8192
//     fp64 values for blocks of 16..128 values, else 8192 fp32 / 4096 fp64
//     values (four or two record-index tiles: a longer tile amortises the
//     look-back wait over more data); each lane holds LV consecutive values
//     (8, or 4 for blocks of 4) per slab of 256 x LV values, so a block of BS
//     values is an aligned group of BS/LV lanes and every per-block reduction
//     / scan is a DPP or swizzle step inside that group;
//   * block statistics (min/max -> radius, median, constant test) and the
//     XOR leading-byte encoding are computed in registers; the records of the
//     tile are staged in LDS;
//   * the tile's record bytes and constant-block count are published through
//     a two-level decoupled look-back (tiles within 64-tile chunks, whose
//     aggregates build up in one atomic accumulator per chunk; the tile order
//     is the launch order, blockIdx.x), whose reads are issued before the
//     records are staged so that staging overlaps their round trip;
//   * records are written at payload + (bytes of all earlier records).
//
// The payload start does not depend on the number of constant blocks: the
// caller's buffer reserves the largest possible header+bitmap+median region
// in front of it, and k_assemble writes the header, bitmap and constant
// medians immediately before the payload once the totals are known.  The
// stream therefore starts at a data-dependent offset inside the output
// buffer, which is returned with the size.  This keeps HBM traffic at one read
// of the input plus one write of the stream.
#include <type_traits>

#include "szx_common.hpp"
#include "szx_batch.hpp"
#include "szx_internal.hpp"
#include "szx_wave.hpp"

namespace szxamd {


template <typename T> struct SlabLoad;
template <> struct SlabLoad<float> {
    __device__ static void load4_full(const float *x, float *v)
    {
        const float4 f = *reinterpret_cast<const float4 *>(x);
        v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
    }
    __device__ static void load4(const float *x, uint64_t e0, uint64_t n, float *v)
    {
        if (e0 + 4 <= n) {
            float4 f = *reinterpret_cast<const float4 *>(x + e0);
            v[0] = f.x; v[1] = f.y; v[2] = f.z; v[3] = f.w;
        } else {
            for (int k = 0; k < 4; k++) v[k] = e0 + k < n ? x[e0 + k] : 0.0f;
        }
    }
};
template <> struct SlabLoad<double> {
    __device__ static void load4_full(const double *x, double *v)
    {
        const double2 a = *reinterpret_cast<const double2 *>(x);
        const double2 b = *reinterpret_cast<const double2 *>(x + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
    __device__ static void load4(const double *x, uint64_t e0, uint64_t n, double *v)
    {
        if (e0 + 4 <= n) {
            double2 a = *reinterpret_cast<const double2 *>(x + e0);
            double2 b = *reinterpret_cast<const double2 *>(x + e0 + 2);
            v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
        } else {
            for (int k = 0; k < 4; k++) v[k] = e0 + k < n ? x[e0 + k] : 0.0;
        }
    }
};

template <typename T> __device__ inline T pos_inf();
template <> __device__ inline float pos_inf<float>() { return __builtin_huge_valf(); }
template <> __device__ inline double pos_inf<double>() { return __builtin_huge_val(); }

// Block-wide exclusive scan of NBLK packed words held in LDS, done by wave 0.
// Returns the total in every lane of wave 0 (broadcast later through LDS).
template <int NBLK>
__device__ inline uint32_t wave0_exclusive_scan(uint32_t *a, int lane)
{
    constexpr int PER = NBLK >= 64 ? NBLK / 64 : 1;
    uint32_t loc[PER];
    uint32_t sum = 0;
    const int base = lane * PER;
#pragma unroll
    for (int i = 0; i < PER; i++) {
        uint32_t v = base + i < NBLK ? a[base + i] : 0u;
        loc[i] = sum;
        sum += v;
    }
    uint32_t inc = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    const uint32_t excl = inc - sum;
#pragma unroll
    for (int i = 0; i < PER; i++)
        if (base + i < NBLK) a[base + i] = excl + loc[i];
    return __shfl(inc, 63, 64);
}

// What a lane keeps of one slab between encoding and staging.
//   mid   = per value the word with its mid bytes as the low bytes
//   meta  = mid-byte offset in the record (12) | rb (4) << 12 |
//           non-constant << 16 | reqLength (7) << 17
//   code  = the lane's 2-bit lead codes (LV/4 bytes) | lead bytes of the
//           block (7) << 16
//   mb    = per value 8 x (mid bytes), one byte each
// fp32 keeps the mid bytes of two neighbouring values as one run (PAIRS):
// value 2p's bytes, then value 2p+1's, as the low bytes of a 64-bit word
// (zero above), so staging handles a pair at a time.
template <typename T, int LV> struct SlabState {
    static constexpr bool PAIRS = sizeof(T) == 4;
    static constexpr int NM = PAIRS ? LV / 2 : LV;
    static_assert(PAIRS ? NM <= 8 : NM * 4 <= 32, "the mid counts fit one or two words");
    static_assert(!PAIRS || LV % 2 == 0, "values pair up");
    static_assert(LV <= 16, "the lead codes fit one word");
    using Mid = typename std::conditional<PAIRS, uint64_t, typename Word<T>::U>::type;
    using MB = typename std::conditional<(PAIRS && NM > 4), uint64_t, uint32_t>::type;
    Mid mid[NM];
    // PAIRS: mid bits of run r (0..64) in byte r; else mid bytes of value k
    // (0..8) in bits 4k..4k+3
    MB mb;
    // meta: lane's mid-byte offset in its block (bits 0-11), rb (12-15),
    // non-constant (16), reqLength (17-23), lead-code bytes of the block
    // (24-30); code: the lane's lead codes, 2 bits per value, first value in
    // the top bits of byte 0
    uint32_t meta, code;
    float hmed;
};

template <typename T> __device__ inline T vmin(T a, T b);
template <typename T> __device__ inline T vmax(T a, T b);
template <> __device__ inline float vmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <> __device__ inline float vmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }
template <> __device__ inline double vmin<double>(double a, double b) { return __builtin_fmin(a, b); }
template <> __device__ inline double vmax<double>(double a, double b) { return __builtin_fmax(a, b); }

// 8 x the leading bytes (capped at 3) that s shares with the previous word:
// the reference tests xr >> 8(W-3), >> 8(W-2), >> 8(W-1) against zero
// (szx_float.cc:96-101, szx_double.cc:80-85), i.e. min(clz(xr) / 8, 3) on the
// top 32 bits of the word.
template <typename U> __device__ inline int lead8(U xr)
{
    const uint32_t top = sizeof(U) == 4 ? (uint32_t)xr : (uint32_t)((uint64_t)xr >> 32);
    const int lz = __builtin_clzg(top, 32);
    return (lz < 24 ? lz : 24) & ~7;
}

// Exact block min/max with the reference's NaN behaviour: NaNs never win a
// comparison (szx_float.cc:327-333), so fmin/fmax over the values, except that
// a NaN first value makes both NaN.
template <typename T, int LV, int SEG, bool FULL>
__device__ inline void minmax_float(const T (&v)[LV], const bool (&valid)[LV], T &mn, T &mx)
{
    using U = typename Word<T>::U;
    mn = pos_inf<T>();
    mx = -pos_inf<T>();
#pragma unroll
    for (int k = 0; k < LV; k++) {
        if (FULL || valid[k]) {
            mn = vmin(mn, v[k]);
            mx = vmax(mx, v[k]);
        }
    }
    if constexpr (sizeof(T) == 4) {
        // across lanes on order-preserving integer keys (negative floats'
        // magnitude bits flipped), so each DPP step is one integer min / max;
        // the lane extremes are numbers here (the per-lane fmin / fmax above
        // already dropped the NaNs) and the map is its own inverse
        auto key = [](uint32_t u) { return u ^ ((uint32_t)((int32_t)u >> 31) & 0x7FFFFFFFu); };
        mn = from_bits<T>(key(seg_reduce<SEG>(key(to_bits<T>(mn)), [](uint32_t a, uint32_t b) {
            return (uint32_t)((int32_t)a < (int32_t)b ? (int32_t)a : (int32_t)b);
        })));
        mx = from_bits<T>(key(seg_reduce<SEG>(key(to_bits<T>(mx)), [](uint32_t a, uint32_t b) {
            return (uint32_t)((int32_t)a > (int32_t)b ? (int32_t)a : (int32_t)b);
        })));
    } else {
        mn = from_bits<T>(seg_reduce<SEG>(to_bits<T>(mn), [](U a, U b) {
            return to_bits<T>(vmin(from_bits<T>(a), from_bits<T>(b)));
        }));
        mx = from_bits<T>(seg_reduce<SEG>(to_bits<T>(mx), [](U a, U b) {
            return to_bits<T>(vmax(from_bits<T>(a), from_bits<T>(b)));
        }));
    }
    const T first = from_bits<T>(seg_first<SEG>(to_bits<T>(v[0])));
    if (first != first) { mn = first; mx = first; }
}

// Encode one slab: block statistics, reqLength, shifted words, lead codes and
// mid-byte counts, all in registers.  FULL = every value of the slab exists.
// recb = record bytes of the lane's block (0 if constant), isconst, med = the
// block median as float.
template <typename T, int BS, int LV, bool FULL>
__device__ __forceinline__ void encode_slab(const T (&v)[LV], uint64_t e0, uint64_t n, int lane,
                                            int lseg, int reqexp, float eb, int fm4, SlabState<T, LV> &st,
                                            uint32_t &recb, uint32_t &isconst, float &med_out)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    constexpr int SEG = BS / LV;
    bool exists = true;
    int nel = BS;
    bool valid[LV];
#pragma unroll
    for (int k = 0; k < LV; k++) valid[k] = true;
    if constexpr (!FULL) {
        const uint64_t bstart = e0 - (uint64_t)LV * (uint64_t)lseg;
        exists = bstart < n;
        const uint64_t rem = exists ? n - bstart : 0;
        nel = rem >= (uint64_t)BS ? BS : (int)rem;
#pragma unroll
        for (int k = 0; k < LV; k++) valid[k] = e0 + k < n;
    }

    // fmin/fmax ignore NaNs (quiet or signalling: the compiler quiets its
    // operands first), as the reference's comparisons do; a NaN first value
    // makes both NaN (minmax_float).
    T mn, mx;
    minmax_float<T, LV, SEG, FULL>(v, valid, mn, mx);

    // radius, median, constant test (szx_float.cc:335-347,
    // szx_double.cc:990-1003): median and radius are kept as float.  The
    // random-access OpenMP compressor's fp64 statistics round the radius to
    // float first and use it for the median and the test
    // (szx_double.cc:1120-1135).
    const T rad = (mx - mn) * (T)0.5;
    const float rad_f = (float)rad;
    const bool f4 = W == 8 && fm4;
    const T med = mn + (f4 ? (T)rad_f : rad);
    const float med_f = (float)med;
    const bool nonconst = exists && !(f4 ? (rad_f <= eb) : (rad <= (T)eb));

    int clamped;
    const int len = req_length<T>(rad_expo(rad_f), reqexp, clamped);
    const int rb = (len + 7) >> 3;
    const int shift = (-len) & 7;
    const float hmed = clamped ? 0.0f : med_f;
    // fp32 subtracts the stored median; fp64 keeps the unclamped one
    // (szx_float.cc:60,90 vs szx_double.cc:38,78).
    const T msub = W == 4 ? (T)hmed : (T)med_f;

    U w[LV];
#pragma unroll
    for (int k = 0; k < LV; k++) w[k] = to_bits<T>(v[k] - msub) >> shift;
    U prev = seg_prev<SEG>(w[LV - 1], lane);

    // Per value: lead bytes ld (shared with the previous word), mid bytes
    // m = rb - ld (>= 0), and the word shifted so that its mid bytes
    // [lo, lo + m), lo = W - rb, are its low bytes (only those are staged).
    const int rb8 = 8 * rb, lo8 = 8 * W - rb8;
    // lead codes of values 8h..8h+7 (x 8) in codex[h]
    uint32_t codex[(LV + 7) / 8] = {}, tot8 = 0;
    typename SlabState<T, LV>::MB mb = 0;
    U mid_even = 0;      // PAIRS: the pair's first value, its mid bits
    uint32_t m8_even = 0;
    // NARROW: no lane of the wave needs 4 bytes, so every m8 is at most 24
    // and one bit-field extract gives the mid bits (4 required bytes with no
    // lead byte, m8 = 32, is the width the extract cannot encode)
    auto values = [&](auto narrow) {
        constexpr bool NARROW = decltype(narrow)::value;
#pragma unroll
        for (int k = 0; k < LV; k++) {
            const int ld8 = lead8<U>(w[k] ^ prev);
            // max(rb - ld, 0) mid bytes (x 8): one clamped subtract
            uint32_t m8 = __builtin_elementwise_sub_sat((uint32_t)rb8, (uint32_t)ld8);
            // the value's mid bytes as the low bytes of `mid`; fp32 keeps exactly
            // those m bytes (the bytes above them are zero), so the stage can OR
            // the word in without a mask
            U mid;
            if constexpr (W == 4)
                mid = (NARROW || m8 < 32) ? __builtin_amdgcn_ubfe(w[k], (uint32_t)lo8, m8) : w[k];
            else
                mid = w[k] >> lo8;
            uint32_t ldv = (uint32_t)ld8;
            if constexpr (!FULL) {
                if (!valid[k]) { m8 = 0; ldv = 0; mid = 0; }
            }
            // element k -> byte k/4 of the lane's lead codes, bits 7-6 first
            // (8 ld here; the sum is divided by 8 below)
            codex[k >> 3] |= ldv << (8 * ((k & 7) >> 2) + 6 - 2 * (k & 3));
            if constexpr (SlabState<T, LV>::PAIRS) {
                if ((k & 1) == 0) {
                    mid_even = mid;
                    m8_even = m8;
                } else {
                    st.mid[k >> 1] = (uint64_t)(uint32_t)mid_even | ((uint64_t)(uint32_t)mid << m8_even);
                    const uint32_t mp = m8_even + m8;   // the pair's mid bits
                    mb |= (typename SlabState<T, LV>::MB)mp << (8 * (k >> 1));
                    tot8 += mp;
                }
            } else {
                st.mid[k] = mid;
                mb |= (m8 >> 3) << (4 * k);
                tot8 += m8;
            }
            prev = w[k];
        }
    };
    if (W == 4 && __builtin_amdgcn_ballot_w64(rb == 4) == 0)
        values(std::true_type{});
    else
        values(std::false_type{});
    uint32_t code = codex[0] >> 3;
    if constexpr (LV > 8) code |= (codex[1] >> 3) << 16;
    const uint32_t tot = tot8 >> 3;
    const uint32_t inc = seg_scan_incl<SEG>(tot, lane);
    const uint32_t midtot = seg_last<SEG>(inc);
    const uint32_t nlead = (uint32_t)(nel + 3) >> 2;

    st.meta = (inc - tot) | ((uint32_t)rb << 12) | ((nonconst ? 1u : 0u) << 16) |
              ((uint32_t)len << 17) | (nlead << 24);
    st.code = code;
    st.mb = mb;
    st.hmed = hmed;
    recb = nonconst ? 5u + nlead + midtot : 0u;
    isconst = exists && !nonconst ? 1u : 0u;
    med_out = med_f;
}

// Write the m low bytes of t to LDS at p (every payload byte of a tile is
// written exactly once, by the lane that owns it, so the stage needs no
// clearing).
template <typename U>
__device__ inline void stage_bytes(uint8_t *p, U t, int m)
{
#pragma unroll
    for (int j = 0; j < (int)sizeof(U); j++)
        if (j < m) p[j] = (uint8_t)(t >> (8 * j));
}

// Record bytes a tile stages in LDS.  A tile whose records are larger writes
// them straight to global memory once its prefix is known (a second encode of
// the tile, several times slower), so the cap trades resident workgroups per
// CU against how compressible a tile must be to stay staged: 20 KiB for the
// 32 KiB tiles (staged above about 1.6x compression), 48 KiB for the 64 KiB
// tiles (above about 1.33x) -- 44 KiB for 64 KiB tiles of 512 or more blocks
// (fp32 block sizes 16 and 32, fp64 16), whose per-block scan words would
// otherwise push a workgroup past the LDS share of three per CU (above about
// 1.45x).
constexpr int STAGE_CAP_BYTES = 20480;
constexpr int STAGE_CAP_WIDE_BYTES = 49152;
constexpr int STAGE_CAP_WIDE_MANY_BLOCKS_BYTES = 45056;

// Shape of a compressor tile: T, block size, values per lane, slabs.
template <typename T, int BS, int LV, int SLABS> struct TileShape {
    static constexpr int W = Word<T>::W;
    static constexpr int SLAB = 256 * LV;           // values per slab
    static constexpr int TILE = SLAB * SLABS;       // values per tile
    static constexpr int NBLK = TILE / BS;          // blocks per tile (multiple of 8)
    static constexpr int STAGE = TILE * W + NBLK * (5 + BS / 4);   // largest record bytes
    static constexpr int CAP = TILE * W < 65536 ? STAGE_CAP_BYTES
                               : NBLK >= 512    ? STAGE_CAP_WIDE_MANY_BLOCKS_BYTES
                                                : STAGE_CAP_WIDE_BYTES;
    static constexpr int STAGE_CAP = STAGE < CAP ? STAGE : CAP;
    static constexpr int STAGE_DW = (STAGE_CAP + 32 + 15) / 16 * 4;
};

// A tile's block words (scan, medians, bitmap bits, total) and its staged
// records.
template <typename T, int BS, int LV, int SLABS> struct TileMeta {
    using S = TileShape<T, BS, LV, SLABS>;
    uint32_t pack[S::NBLK];       // record bytes | constant<<20 -> exclusive prefix
    float med[S::NBLK];           // medians (constant blocks)
    uint8_t state[S::NBLK];       // 1 = non-constant (bitmap bit)
    uint32_t total;               // the tile's (bytes | constants << 20)
};
template <typename T, int BS, int LV, int SLABS> struct Stage {
    __attribute__((aligned(16))) uint32_t w[TileShape<T, BS, LV, SLABS>::STAGE_DW];   // staged records
};
template <typename T, int BS, int LV, int SLABS> struct CompressLds {
    Stage<T, BS, LV, SLABS> stage;
    TileMeta<T, BS, LV, SLABS> meta;
    unsigned long long excl[2];   // the tile's exclusive (bytes, constants)
};

// Issue every load of a tile (values past n read as zero and are masked);
// FULL: the whole tile exists (plain 16-byte loads).
template <typename T, int LV, int SLABS, bool FULL>
__device__ __forceinline__ void load_tile(const CompressParams &p, uint64_t tile, T (&v)[SLABS][LV])
{
    constexpr int SLAB = 256 * LV;
    const T *x = static_cast<const T *>(p.x);
    const uint64_t t0 = tile * (uint64_t)(SLAB * SLABS);
    if constexpr (FULL) {
        const T *xt = x + t0 + (uint64_t)LV * threadIdx.x;
#pragma unroll
        for (int u = 0; u < SLABS; u++)
#pragma unroll
            for (int q = 0; q < LV; q += 4) SlabLoad<T>::load4_full(xt + u * SLAB + q, &v[u][q]);
    } else {
#pragma unroll
        for (int u = 0; u < SLABS; u++)
#pragma unroll
            for (int q = 0; q < LV; q += 4)
                SlabLoad<T>::load4(x, t0 + (uint64_t)u * SLAB + (uint64_t)LV * threadIdx.x + q, p.n, &v[u][q]);
    }
}

// Block statistics and records of a tile whose values are in v, in
// registers (st); each block's record bytes, constant flag and median go to
// M for the scan.
template <typename T, int BS, int LV, int SLABS>
__device__ __forceinline__ void encode_slabs(const CompressParams &p, uint64_t tile, const T (&v)[SLABS][LV],
                                             SlabState<T, LV> (&st)[SLABS], TileMeta<T, BS, LV, SLABS> &M,
                                             float eb)
{
    using S = TileShape<T, BS, LV, SLABS>;
    constexpr int SEG = BS / LV;          // lanes per block
    constexpr int SLAB = S::SLAB;
    constexpr int TILE = S::TILE;
    constexpr int BPS = SLAB / BS;        // blocks per slab
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int lseg = tid & (SEG - 1);
    const int reqexp = req_expo(eb);
    const uint64_t n = p.n;
    const uint64_t t0 = tile * TILE;
    const bool full = t0 + TILE <= n;

#pragma unroll
    for (int u = 0; u < SLABS; u++) {
        const uint64_t e0 = t0 + (uint64_t)u * SLAB + (uint64_t)LV * tid;
        uint32_t recb, isconst;
        float med;
        if (full)
            encode_slab<T, BS, LV, true>(v[u], e0, n, lane, lseg, reqexp, eb, p.fm4, st[u], recb, isconst, med);
        else
            encode_slab<T, BS, LV, false>(v[u], e0, n, lane, lseg, reqexp, eb, p.fm4, st[u], recb, isconst, med);
        if (lseg == 0) {
            const int lb = u * BPS + tid / SEG;
            M.pack[lb] = recb | (isconst << 20);
            M.state[lb] = recb ? 1 : 0;
            M.med[lb] = med;
        }
    }
}

// Encode a tile whose values are in v (all 256 lanes): block statistics and
// records in registers (st), the exclusive scan of record sizes over the
// tile's blocks (M), and the aggregate published for the look-back.  `pend`
// (lane 0 of wave 0) receives the chunk accumulator's previous value for
// lookback_chunk_done.
template <typename T, int BS, int LV, int SLABS, int IPT>
__device__ __forceinline__ void encode_scan(const CompressParams &p, const LookbackRefs &L, uint64_t tile,
                                            const T (&v)[SLABS][LV], SlabState<T, LV> (&st)[SLABS],
                                            TileMeta<T, BS, LV, SLABS> &M, unsigned long long &pend, float eb)
{
    using S = TileShape<T, BS, LV, SLABS>;
    constexpr int NBLK = S::NBLK;
    static_assert(BS >= LV && BS <= 256 && (BS & (BS - 1)) == 0, "power-of-two block size");
    static_assert(NBLK % 8 == 0, "tiles own whole bitmap bytes");
    static_assert((NBLK / IPT) % 8 == 0, "index tiles own whole bitmap bytes");
    static_assert(S::STAGE / IPT < (1 << 16), "record starts in an index tile fit the 16-bit block index");
    const int lane = threadIdx.x & 63;
    encode_slabs<T, BS, LV, SLABS>(p, tile, v, st, M, eb);
    __syncthreads();

    if (threadIdx.x < 64) {
        const uint32_t tot = wave0_exclusive_scan<NBLK>(M.pack, lane);
        if (lane == 0) {
            M.total = tot;
            pend = lookback_publish(L, tile, tot & 0xFFFFFu, tot >> 20);
        }
    }
    __syncthreads();
}

// Write an encoded tile's records at s_stage + (tile-relative byte offset):
// the LDS stage, or the payload itself for a tile too large for the stage.
// Each byte is written once, by the lane that owns it.  M must hold the
// tile's scan (encode_scan, then a barrier).
template <typename T, int BS, int LV, int SLABS, bool LDS>
__device__ __forceinline__ void stage_slab(const SlabState<T, LV> &st, int u, const TileMeta<T, BS, LV, SLABS> &M,
                                           uint8_t *s_stage)
{
    using U = typename Word<T>::U;
    constexpr int SEG = BS / LV;
    constexpr int NLB = LV / 4;           // lead-code bytes per lane
    constexpr int BPS = 256 * LV / BS;    // blocks per slab
    const int tid = threadIdx.x;
    const int lseg = tid & (SEG - 1);
    const uint32_t meta = st.meta;
    if (!((meta >> 16) & 1)) return;
    const int lb = u * BPS + tid / SEG;
    const uint32_t base = M.pack[lb] & 0xFFFFFu;
    const uint32_t nlead = (meta >> 24) & 0x7Fu;
    if (lseg == 0) {   // reqLength byte + big-endian median
        const uint32_t hb = bswap32(f32_bits(st.hmed));
        s_stage[base] = (uint8_t)((meta >> 17) & 0x7Fu);
        s_stage[base + 1] = (uint8_t)hb;
        s_stage[base + 2] = (uint8_t)(hb >> 8);
        s_stage[base + 3] = (uint8_t)(hb >> 16);
        s_stage[base + 4] = (uint8_t)(hb >> 24);
    }
#pragma unroll
    for (int b = 0; b < NLB; b++)
        if ((uint32_t)(lseg * NLB + b) < nlead)
            s_stage[base + 5 + lseg * NLB + b] = (uint8_t)(st.code >> (8 * b));
    if constexpr (LDS && sizeof(U) == 4) {
        // Mid bytes OR-ed into the zeroed stage as aligned 8-byte words: a
        // run's bytes start at bit offset o8 of the stage.  LDS atomics are the expensive part here
        // (measured: two unconditional ORs per value ran 0.28 ms against 0.22
        // ms for predicated ones at 512^3), so the lanes with nothing to add
        // sit out.  OR is idempotent and every stage byte has one writer, so
        // the words a run shares with its neighbours need no care.
        static_assert(TileShape<T, BS, LV, SLABS>::STAGE_DW % 2 == 0, "the stage is whole 8-byte words");
        uint64_t *sq = reinterpret_cast<uint64_t *>(s_stage);
        uint32_t o8 = 8u * (base + 5 + nlead + (meta & 0xFFFu));
#pragma unroll
        for (int r = 0; r < LV / 2; r++) {
            // two neighbouring values as one run of mp <= 64 bits (zero
            // above), OR-ed into the one or two 8-byte stage words it touches
            const uint32_t mp = (st.mb >> (8 * r)) & 0xFFu;
            if (mp) {
                const uint64_t run = st.mid[r];
                const uint32_t s = o8 & 63u;
                uint64_t *pq = sq + (o8 >> 6);
                __hip_atomic_fetch_or(pq, run << s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                // bits past the first word (only when s > 0): the run's top s
                if (s + mp > 64u)
                    __hip_atomic_fetch_or(pq + 1, run >> (64u - s), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            o8 += mp;
        }
    } else {
        uint8_t *q = s_stage + base + 5 + nlead + (meta & 0xFFFu);
#pragma unroll
        for (int r = 0; r < SlabState<T, LV>::NM; r++) {
            const int m = SlabState<T, LV>::PAIRS ? (int)((st.mb >> (8 * r)) & 0xFFu) >> 3
                                                  : (int)(st.mb >> (4 * r)) & 0xF;
            stage_bytes<typename SlabState<T, LV>::Mid>(q, st.mid[r], m);
            q += m;
        }
    }
}

// Stage every slab's records; after(u) runs after slab u is staged.
template <typename T, int BS, int LV, int SLABS, typename After>
__device__ __forceinline__ void stage_records(const SlabState<T, LV> (&st)[SLABS],
                                              const TileMeta<T, BS, LV, SLABS> &M, uint8_t *s_stage, After after)
{
#pragma unroll
    for (int u = 0; u < SLABS; u++) {
        stage_slab<T, BS, LV, SLABS, true>(st[u], u, M, s_stage);
        after(u);
    }
}

// Write out a tile once its exclusive prefix (ebytes, econst) is known (all
// 256 lanes; B's staging must be complete): the record index entries, the
// payload when it was staged (16-byte stores inside, bytes at the edges, whose
// other bytes belong to the neighbouring tiles), constant medians and bitmap
// bytes.
template <typename T, int BS, int LV, int SLABS, int IPT>
__device__ __forceinline__ void copy_out(const CompressParams &p, uint64_t tile, const TileMeta<T, BS, LV, SLABS> &B,
                                         const Stage<T, BS, LV, SLABS> &G, uint64_t ebytes, uint64_t econst,
                                         bool staged)
{
    using S = TileShape<T, BS, LV, SLABS>;
    constexpr int NBLK = S::NBLK;
    constexpr int IT = IPT;               // index tiles per tile
    constexpr int NBLKI = NBLK / IT;      // blocks per index tile
    const int tid = threadIdx.x;
    const uint32_t tw = B.total;
    const uint32_t tbytes = tw & 0xFFFFFu;
    const uint8_t *s_stage = reinterpret_cast<const uint8_t *>(G.w);

    // record index: one pair per index tile, from the tile's prefix and the
    // scanned word at the index tile's first block
    if (p.index && tid < IT) {
        const uint64_t it = tile * IT + tid;
        if (it < p.nidx) {
            const uint32_t pk = B.pack[tid * NBLKI];
            p.index[2 * it] = ebytes + (pk & 0xFFFFFu);
            p.index[2 * it + 1] = econst + (pk >> 20);
        }
    }
    if (tid == 0 && tile == p.ntiles - 1) {
        p.totals[0] = ebytes + tbytes;
        p.totals[1] = econst + (tw >> 20);
        if (p.index) {
            p.index[2 * p.nidx] = ebytes + tbytes;
            p.index[2 * p.nidx + 1] = econst + (tw >> 20);
        }
    }

    if (staged) {
        uint8_t *dst = p.payload + ebytes;
        const uint32_t head = (uint32_t)((16 - ((uintptr_t)dst & 15)) & 15);
        const uint32_t h = head < tbytes ? head : tbytes;
        if (tid < (int)h) dst[tid] = s_stage[tid];
        const uint32_t nchunk = (tbytes - h) >> 4;
        const int r = h & 3;
        for (uint32_t c = tid; c < nchunk; c += 256) {
            const uint32_t o = h + 16 * c;             // stage offset of the chunk
            const uint32_t *sw = G.w + (o >> 2);
            const uint32_t a0 = sw[0], a1 = sw[1], a2 = sw[2], a3 = sw[3];
            const uint32_t a4 = r ? sw[4] : 0u;
            uint4 val;
            if (r) {
                val.x = __builtin_amdgcn_alignbyte(a1, a0, r);
                val.y = __builtin_amdgcn_alignbyte(a2, a1, r);
                val.z = __builtin_amdgcn_alignbyte(a3, a2, r);
                val.w = __builtin_amdgcn_alignbyte(a4, a3, r);
            } else {
                val = make_uint4(a0, a1, a2, a3);
            }
            // non-temporal: the stream is not read again by this launch, and
            // keeping it out of the caches leaves them to the input (REL
            // step 0.305 -> 0.294 ms at 512^3 fp32)
            typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
            const u32x4_t wv = {val.x, val.y, val.z, val.w};
            __builtin_nontemporal_store(wv, reinterpret_cast<u32x4_t *>(dst + o));
        }
        const uint32_t tail0 = h + 16 * nchunk;
        if (tid < (int)(tbytes - tail0)) dst[tail0 + tid] = s_stage[tail0 + tid];
    }

    // Constant medians in block order: the exclusive constant count of a
    // block sits in the high bits of its scanned word, so a block is
    // constant exactly when the next block's count is one higher.
    for (int lb = tid; lb < NBLK; lb += 256) {
        const uint32_t cexcl = B.pack[lb] >> 20;
        const uint32_t cnext = lb + 1 < NBLK ? (B.pack[lb + 1] >> 20) : (tw >> 20);
        if (cnext != cexcl)
            p.cmed[econst + cexcl] = bswap32(f32_bits(B.med[lb]));
    }
    for (int i = tid; i < NBLK / 8; i += 256) {
        const uint64_t gbyte = tile * (NBLK / 8) + i;
        if (gbyte < p.nbit) {
            uint32_t b = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) b |= (uint32_t)B.state[8 * i + k] << (7 - k);
            p.bitmap[gbyte] = (uint8_t)b;
        }
    }
    // random-access container: every record's size at its rank among the
    // non-constant blocks (szx_float.cc:1464)
    if (p.osz) {
        for (int lb = tid; lb < NBLK; lb += 256) {
            const uint64_t b = tile * NBLK + lb;
            if (b < p.nb && B.state[lb]) {
                const uint32_t cur = B.pack[lb];
                const uint32_t nxt = lb + 1 < NBLK ? B.pack[lb + 1] : tw;
                p.osz[b - (econst + (cur >> 20))] = (uint16_t)((nxt - cur) & 0xFFFFFu);
            }
        }
    }
    // per-block entries of the record index: the record's start inside its
    // index tile's payload bytes, or the constant block's rank among the
    // index tile's constant blocks (both fit 16 bits)
    if (p.bidx) {
        for (int lb = tid; lb < NBLK; lb += 256) {
            const uint64_t b = tile * NBLK + lb;
            if (b < p.nb) {
                const uint32_t base = B.pack[(lb / NBLKI) * NBLKI];
                p.bidx[b] = (uint16_t)(B.state[lb] ? ((B.pack[lb] - base) & 0xFFFFFu)
                                                   : ((B.pack[lb] >> 20) - (base >> 20)));
            }
        }
    }
}

// REL with the range pass fused into the launch: every wave folds the pass's
// RANGE_LB_PARTS {min, max} pairs (four per lane) into the bound itself, so
// no barrier is needed.  The loads are issued before the tile's and the
// fold follows them in the same block, so it waits for its own loads only.
template <typename T> struct RelParts {
    static constexpr int PL = RANGE_LB_PARTS / 64;   // partials per lane
    static_assert(RANGE_LB_PARTS % 128 == 0, "whole 16-byte loads per lane");
    T mn[PL], mx[PL];
    T first;
    __device__ void load(const CompressParams &p)
    {
        const T *q = static_cast<const T *>(p.rparts) + 2 * PL * (threadIdx.x & 63);
        T t[2 * PL];
#pragma unroll
        for (int j = 0; j < 2 * PL; j += 4) SlabLoad<T>::load4_full(q + j, t + j);
#pragma unroll
        for (int k = 0; k < PL; k++) {
            mn[k] = t[2 * k];
            mx[k] = t[2 * k + 1];
        }
        first = static_cast<const T *>(p.x)[0];
    }
    // Min and max skip NaNs (the partials hold none), a NaN x[0] makes the
    // range NaN, and the bound is rounded as the reference rounds it:
    // fl32(range * rel) for fp32, (float)(range * (double)rel) for fp64
    // (szx.cc:249-253, 303-308).
    __device__ float fold(float rel) const
    {
        using U = typename Word<T>::U;
        T a = mn[0], b = mx[0];
#pragma unroll
        for (int k = 1; k < PL; k++) {
            a = mn[k] < a ? mn[k] : a;
            b = mx[k] > b ? mx[k] : b;
        }
        a = from_bits<T>(seg_reduce<64>(to_bits<T>(a), [](U x, U y) {
            const T fx = from_bits<T>(x), fy = from_bits<T>(y);
            return fy < fx ? y : x;
        }));
        b = from_bits<T>(seg_reduce<64>(to_bits<T>(b), [](U x, U y) {
            const T fx = from_bits<T>(x), fy = from_bits<T>(y);
            return fy > fx ? y : x;
        }));
        if (first != first) {
            a = first;
            b = first;
        }
        const T range = b - a;
        return (float)(range * (T)rel);
    }
};

// Finish a tile's look-back from its polled words (wave 0) and leave its
// exclusive prefix in ex[0..1] for the workgroup (read after a barrier).  On a
// wait that overran its bound the call is flagged failed and the prefix is
// zero, which keeps every write in bounds.
__device__ __forceinline__ void tile_prefix(const CompressParams &p, const LookbackRefs &L, uint64_t tile,
                                            const LookbackPoll &q, unsigned long long *ex)
{
    if (threadIdx.x < 64) {
        unsigned long long eb = 0, ec = 0;
        if (!lookback_finish(L, tile, q, true, eb, ec)) {
            eb = 0;
            ec = 0;
            if (threadIdx.x == 0) atomicExch(p.err, 1);
        }
        if (threadIdx.x == 0) {
            ex[0] = eb;
            ex[1] = ec;
        }
    }
}

// One tile per workgroup.  The tile publishes its aggregate right after
// its scan; staging the records then overlaps the look-back.
template <typename T, int BS, int LV, int SLABS, int IPT>
__device__ __forceinline__ void compress_tile(const CompressParams &p, uint64_t tile, CompressLds<T, BS, LV, SLABS> &lds)
{
    const LookbackRefs L{p.tiles, p.chunks, p.cacc, p.ntiles};
    T v[SLABS][LV];
    unsigned long long pend = 0;
    float eb = p.eb;
    {
        constexpr int TILE = 256 * LV * SLABS;
        RelParts<T> rp;
        if (tile * (uint64_t)TILE + TILE <= p.n) {
            if (p.rparts) rp.load(p);
            load_tile<T, LV, SLABS, true>(p, tile, v);
            if (p.rparts) eb = rp.fold(p.rel);
        } else {
            if (p.rparts) rp.load(p);
            load_tile<T, LV, SLABS, false>(p, tile, v);
            if (p.rparts) eb = rp.fold(p.rel);
        }
    }
    if (p.rparts) {
        if (tile == 0 && threadIdx.x == 0) *p.eb_out = eb;
    } else if (p.eb_dev) {
        eb = *p.eb_dev;
    }
    if constexpr (sizeof(T) == 4) {   // the fp32 stage is OR-ed into (stage_slab)
        constexpr int N4 = TileShape<T, BS, LV, SLABS>::STAGE_DW / 4;
        for (int i = threadIdx.x; i < N4; i += 256) reinterpret_cast<uint4 *>(lds.stage.w)[i] = make_uint4(0, 0, 0, 0);
    }
    SlabState<T, LV> st[SLABS];
    encode_scan<T, BS, LV, SLABS, IPT>(p, L, tile, v, st, lds.meta, pend, eb);
    // wave 0 issues the look-back reads halfway through staging: the words it
    // reads appear about a memory round trip after their tiles publish them,
    // and the rest of staging covers the round trip
    LookbackPoll q{};
    const bool staged = (lds.meta.total & 0xFFFFFu) <= (uint32_t)TileShape<T, BS, LV, SLABS>::STAGE_CAP;
    if (staged) {
        stage_records<T, BS, LV, SLABS>(st, lds.meta, reinterpret_cast<uint8_t *>(lds.stage.w), [&](int u) {
            if (u == (SLABS - 1) / 2 && threadIdx.x < 64) q = lookback_poll(L, tile);
        });
    } else {
        if (threadIdx.x < 64) q = lookback_poll(L, tile);
    }
    tile_prefix(p, L, tile, q, lds.excl);
    __syncthreads();
    if (!staged) {
        // records too large for the stage: encode the tile again, one slab at
        // a time (nothing held across the look-back), and write them at their
        // final offsets
        constexpr int SLAB = 256 * LV;
        const uint64_t t0 = tile * (uint64_t)(SLAB * SLABS);
        const int reqexp = req_expo(eb);
        const int lseg = threadIdx.x & (BS / LV - 1);
        for (int u = 0; u < SLABS; u++) {
            T vu[LV];
            const uint64_t e0 = t0 + (uint64_t)u * SLAB + (uint64_t)LV * threadIdx.x;
#pragma unroll
            for (int q = 0; q < LV; q += 4) SlabLoad<T>::load4(static_cast<const T *>(p.x), e0 + q, p.n, &vu[q]);
            SlabState<T, LV> su;
            uint32_t recb, isconst;
            float med;
            encode_slab<T, BS, LV, false>(vu, e0, p.n, threadIdx.x & 63, lseg, reqexp, eb, p.fm4, su, recb, isconst,
                                          med);
            stage_slab<T, BS, LV, SLABS, false>(su, u, lds.meta, p.payload + lds.excl[0]);
        }
    }
    copy_out<T, BS, LV, SLABS, IPT>(p, tile, lds.meta, lds.stage, lds.excl[0], lds.excl[1], staged);
    // the chunk aggregate is the look-back's fallback for chunks far back
    // (nearer ones are read tile by tile, and chunk inclusive prefixes appear
    // sooner): publishing it last keeps the accumulator atomic's round trip
    // off the tile's path
    if (threadIdx.x == 0) lookback_chunk_done(L, tile, pend, lds.meta.total & 0xFFFFFu, lds.meta.total >> 20);
}

// Three workgroups per CU (the LDS stage allows three): ask the register
// allocator for the 168 registers that three waves per SIMD leave each.
template <typename T, int BS, int LV, int SLABS, int IPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_compress(CompressParams p)
{
    __shared__ CompressLds<T, BS, LV, SLABS> lds;
    compress_tile<T, BS, LV, SLABS, IPT>(p, blockIdx.x, lds);
}

// Zero the look-back state and totals (a kernel, not a memset, so the
// whole sequence can be captured in a HIP graph).
__global__ void k_zero_u64(unsigned long long *p, uint64_t nwords)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nwords;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Header, bitmap and constant medians in front of the payload, by thread gid
// of nthr.  fastMode 2 (szx_float.cc:1091-1115): [1,1,1,0], (u8)blockSize,
// big-endian constant count, bitmap, big-endian medians.  Random-access
// container (szx_float.cc:1438-1478, 1260-1375): [1,1,1,1], big-endian u64
// blockSize and constant count, the u16 record sizes, bitmap, medians
// (big-endian for fastMode 3, native for fastMode 4).
__device__ inline void assemble_part(const AssembleParams &a, uint64_t gid, uint64_t nthr)
{
    const uint64_t nconst = a.totals[1];
    const uint64_t pbytes = a.totals[0];
    const uint64_t nrec = a.ra ? a.nb - nconst : 0;
    const uint64_t hdr = a.ra ? 20 + 2 * nrec : HEADER_BYTES;   // bitmap offset
    const uint64_t front = hdr + a.nbit + 4 * nconst;
    const uint64_t start = a.region - front;
    uint8_t *s = a.out + start;
    if (gid == 0) {
        if (a.ra) {
            s[0] = 1; s[1] = 1; s[2] = 1; s[3] = 1;
            for (int k = 0; k < 8; k++) s[4 + k] = (uint8_t)((uint64_t)a.bs >> (56 - 8 * k));
            for (int k = 0; k < 8; k++) s[12 + k] = (uint8_t)(nconst >> (56 - 8 * k));
        } else {
            s[0] = 1; s[1] = 1; s[2] = 1; s[3] = 0;
            s[4] = (uint8_t)a.bs;                           // (u8)blockSize, :1095
            for (int k = 0; k < 8; k++) s[5 + k] = (uint8_t)(nconst >> (56 - 8 * k));
        }
        a.result[0] = start;
        a.result[1] = front + pbytes;
        a.result[2] = f32_bits(a.eb_dev ? *a.eb_dev : a.eb);
        a.result[3] = (uint64_t)*a.err;
    }
    if (a.ra) {
        const uint8_t *ob = reinterpret_cast<const uint8_t *>(a.osz);   // little-endian u16
        for (uint64_t i = gid; i < 2 * nrec; i += nthr) s[20 + i] = ob[i];
    }
    for (uint64_t i = gid; i < a.nbit; i += nthr) s[hdr + i] = a.bitmap[i];
    const uint8_t *cm = reinterpret_cast<const uint8_t *>(a.cmed);   // big-endian medians
    if (a.ra == 4) {
        for (uint64_t i = gid; i < 4 * nconst; i += nthr) s[hdr + a.nbit + i] = cm[(i & ~(uint64_t)3) | (3 - (i & 3))];
    } else {
        for (uint64_t i = gid; i < 4 * nconst; i += nthr) s[hdr + a.nbit + i] = cm[i];
    }
}

__global__ void k_assemble(AssembleParams a)
{
    assemble_part(a, blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, (uint64_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// REL: global value range and bound (szx.cc:249-253, 303-308).
// ---------------------------------------------------------------------------
// The grid sweeps the array from its end to its start: the front of the
// array, which the compressor reads first, is then the part most recently
// read, so the compressor finds up to the last-level cache's worth of it
// there (REL step 0.317 -> 0.305 ms at 512^3 fp32).  Min and max do not
// depend on the order (NaNs lose every comparison either way).
// NT lanes per workgroup; `zero` (if set) has its nzero words cleared along
// the way (the compressor's look-back state, when the pass is part of a
// compress launch).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_range_partial(const T *x, uint64_t n, T *part,
                                                      unsigned long long *zero, uint64_t nzero)
{
    T mn = pos_inf<T>(), mx = -pos_inf<T>();
    const uint64_t nq = n / 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    for (; q + 3 * stride < nq; q += 4 * stride) {
        T v[4][4];
#pragma unroll
        for (int u = 0; u < 4; u++) SlabLoad<T>::load4(x, 4 * (nq - 1 - (q + u * stride)), n, v[u]);
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                mn = v[u][k] < mn ? v[u][k] : mn;
                mx = v[u][k] > mx ? v[u][k] : mx;
            }
    }
    for (; q < nq; q += stride) {
        T v[4];
        SlabLoad<T>::load4(x, 4 * (nq - 1 - q), n, v);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            mn = v[k] < mn ? v[k] : mn;
            mx = v[k] > mx ? v[k] : mx;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {
        const T v = x[4 * nq + threadIdx.x];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    for (int m = 1; m < 64; m <<= 1) {
        T o = __shfl_xor(mn, m, 64);
        mn = o < mn ? o : mn;
        o = __shfl_xor(mx, m, 64);
        mx = o > mx ? o : mx;
    }
    if (zero)
        for (uint64_t i = blockIdx.x * (uint64_t)NT + threadIdx.x; i < nzero; i += (uint64_t)gridDim.x * NT)
            zero[i] = 0;
    __shared__ T sm[2][NT / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[0][w] = mn; sm[1][w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < NT / 64; k++) {
            mn = sm[0][k] < mn ? sm[0][k] : mn;
            mx = sm[1][k] > mx ? sm[1][k] : mx;
        }
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}

// Folds the partials.  eb_out (if set) receives the REL bound; mm_out (if
// set) receives {min, max} with the reference's semantics (NaN when x[0] is
// NaN, NaNs elsewhere skipped).
template <typename T>
__global__ __launch_bounds__(256) void k_range_final(const T *x, const T *part, int nparts,
                                                    float rel, float *eb_out, T *mm_out)
{
    T mn = pos_inf<T>(), mx = -pos_inf<T>();
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        mn = part[2 * i] < mn ? part[2 * i] : mn;
        mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
    }
    for (int m = 1; m < 64; m <<= 1) {
        T o = __shfl_xor(mn, m, 64);
        mn = o < mn ? o : mn;
        o = __shfl_xor(mx, m, 64);
        mx = o > mx ? o : mx;
    }
    __shared__ T sm[2][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[0][w] = mn; sm[1][w] = mx; }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
        mn = sm[0][k] < mn ? sm[0][k] : mn;
        mx = sm[1][k] > mx ? sm[1][k] : mx;
    }
    const T first = x[0];
    if (first != first) { mn = first; mx = first; }
    const T range = mx - mn;
    // fp32: fl32(range * rel); fp64: (float)(range * (double)rel)
    if (eb_out) *eb_out = (float)(range * (T)rel);
    if (mm_out) {
        mm_out[0] = mn;
        mm_out[1] = mx;
    }
}

// ---------------------------------------------------------------------------
// Batched compress: many fields (fp32 and fp64 mixed, any sizes) in one
// launch.  Every field keeps its own look-back state; the compressor tiles of
// all fields form one global list (field by field), one workgroup per tile.
// A tile waits only on earlier tiles of its own field, which have lower
// workgroup numbers and so have started before it.
// ---------------------------------------------------------------------------

// The field owning flattened item `item`: the last f with base[f] <= item
// (base has nf + 1 entries; fields with no items are skipped over).
__device__ inline int field_of(const uint64_t *base, int nf, uint64_t item)
{
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (base[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// REL range over the batch: items of RANGE_ITEM values, field by field.
template <typename T>
__device__ inline void range_item(const T *x, uint64_t a, uint64_t b, double *slot)
{
    // a is a multiple of RANGE_ITEM and x 16-byte aligned, so every quad of
    // values starts on a 16-byte boundary; RU quads per thread are in flight
    constexpr int RU = 4;
    T mn = pos_inf<T>(), mx = -pos_inf<T>();
    const uint64_t sweep = 4ull * blockDim.x;
    uint64_t i = a + 4ull * threadIdx.x;
    for (; i + (RU - 1) * sweep + 4 <= b; i += RU * sweep) {
        T v[RU][4];
#pragma unroll
        for (int u = 0; u < RU; u++) SlabLoad<T>::load4(x, i + u * sweep, b, v[u]);
#pragma unroll
        for (int u = 0; u < RU; u++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                mn = v[u][k] < mn ? v[u][k] : mn;
                mx = v[u][k] > mx ? v[u][k] : mx;
            }
    }
    for (; i < b; i += sweep)
        for (int k = 0; k < 4 && i + k < b; k++) {
            const T v = x[i + k];
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
    for (int m = 1; m < 64; m <<= 1) {
        T o = __shfl_xor(mn, m, 64);
        mn = o < mn ? o : mn;
        o = __shfl_xor(mx, m, 64);
        mx = o > mx ? o : mx;
    }
    __shared__ T sm[2][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sm[0][w] = mn;
        sm[1][w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            mn = sm[0][k] < mn ? sm[0][k] : mn;
            mx = sm[1][k] > mx ? sm[1][k] : mx;
        }
        slot[0] = (double)mn;   // exact for fp32 and fp64
        slot[1] = (double)mx;
    }
}

__global__ __launch_bounds__(256) void k_batch_range(const BatchField *bf, const uint64_t *rbase, int nf,
                                                     double *part)
{
    const uint64_t item = blockIdx.x;
    const int fi = field_of(rbase, nf, item);
    const BatchField &f = bf[fi];
    const uint64_t k = item - rbase[fi];
    const uint64_t a = k * RANGE_ITEM, b = a + RANGE_ITEM < f.p.n ? a + RANGE_ITEM : f.p.n;
    if (f.w == 4) range_item(static_cast<const float *>(f.p.x), a, b, part + 2 * item);
    else range_item(static_cast<const double *>(f.p.x), a, b, part + 2 * item);
}

// Fold a REL field's items into its bound (the reference's rounding).
__global__ void k_batch_range_final(BatchField *bf, const uint64_t *rbase, int nf, const double *part)
{
    const int fi = blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= nf) return;
    const BatchField &f = bf[fi];
    if (!f.rel) return;
    double mn = __builtin_inf(), mx = -__builtin_inf();
    for (uint64_t i = rbase[fi]; i < rbase[fi + 1]; i++) {
        mn = part[2 * i] < mn ? part[2 * i] : mn;
        mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
    }
    float *eb = const_cast<float *>(f.p.eb_dev);
    if (f.w == 4) {
        const float first = static_cast<const float *>(f.p.x)[0];
        float lo = (float)mn, hi = (float)mx;
        if (first != first) lo = hi = first;
        *eb = (float)((hi - lo) * f.relv);                // fl32(range * rel)
    } else {
        const double first = static_cast<const double *>(f.p.x)[0];
        double lo = mn, hi = mx;
        if (first != first) lo = hi = first;
        *eb = (float)((hi - lo) * (double)f.relv);        // (float)(range * (double)rel)
    }
}

// Many fields: one workgroup per tile of the batch's global tile list
// (fields in order, tbase[f] = first global tile of field f).  Workgroups
// start in index order, so every tile a look-back waits for has started.
// LVF / LVD: values per lane of the fp32 / fp64 fields' tiles.
template <int BS, int LVF, int LVD, int IPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_compress_batch(const BatchField *bf, const uint64_t *tbase,
                                                                                                const uint32_t *tmap)
{
    constexpr int SF = IPT * TILE_VALUES_F32 / (256 * LVF);
    constexpr int SD = IPT * TILE_VALUES_F64 / (256 * LVD);
    __shared__ union Lds {
        CompressLds<float, BS, LVF, SF> f;
        CompressLds<double, BS, LVD, SD> d;
    } lds;
    const uint64_t g = blockIdx.x;
    const uint32_t f = tmap[g];   // the field of tile g (a table built with the plan)
    CompressParams p = bf[f].p;
    p.ntiles = bf[f].ntiles[IPT == 4 ? 2 : IPT - 1];
    if (bf[f].w == 4) compress_tile<float, BS, LVF, SF, IPT>(p, g - tbase[f], lds.f);
    else compress_tile<double, BS, LVD, SD, IPT>(p, g - tbase[f], lds.d);
}

// Clear every field's look-back words, totals and error flag (one block per
// ZERO_ITEM words, flattened over the fields).
__global__ __launch_bounds__(256) void k_zero_batch(const BatchField *bf, const uint64_t *zbase, int nf)
{
    const int f = field_of(zbase, nf, blockIdx.x);
    const uint64_t a = (blockIdx.x - zbase[f]) * ZERO_ITEM;
    const uint64_t b = a + ZERO_ITEM < bf[f].zwords ? a + ZERO_ITEM : bf[f].zwords;
    unsigned long long *z = bf[f].p.tiles;
    for (uint64_t i = a + threadIdx.x; i < b; i += blockDim.x) z[i] = 0;
}

// Header, bitmap and medians of every field; abase gives each field's
// first block of this flattened grid.
__global__ __launch_bounds__(256) void k_assemble_batch(const BatchField *bf, const uint64_t *abase, int nf)
{
    const int f = field_of(abase, nf, blockIdx.x);
    assemble_part(bf[f].a, (blockIdx.x - abase[f]) * (uint64_t)blockDim.x + threadIdx.x,
                  (abase[f + 1] - abase[f]) * blockDim.x);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static constexpr int kRangeParts = 1024;

CompressLayout compress_layout(uint64_t n, int w, int bs)
{
    CompressLayout L{};
    L.nb = ceil_div(n, (uint64_t)bs);
    L.nbit = ceil_div(L.nb, 8);
    // front region: the larger (random-access) header, the size table and the
    // medians take at most 20 + 2 nb' + 4 nc <= 20 + 4 nb bytes with the bitmap
    L.region = (20 + L.nbit + 4 * L.nb + 15) & ~(uint64_t)15;
    const uint64_t rec_max = 5 + ceil_div((uint64_t)bs, 4) + (uint64_t)bs * w;
    L.payload_max = L.nb * rec_max;
    L.capacity = L.region + L.payload_max + 16;
    L.ntiles = ceil_div(n, compress_tile_values(w, bs));
    if (L.ntiles == 0) L.ntiles = 1;
    L.nidx = index_tiles(n, w);
    // workspace: tile words | chunk words | chunk accumulators | misc (totals,
    // eb) | range parts | bitmap | cmed.
    // The look-back words are sized for one index tile per compressor tile,
    // the smallest tile any launch shape uses (the batched launch shares this
    // layout).
    uint64_t off = 0;
    L.ws_tiles = off;   off += L.nidx * sizeof(unsigned long long);
    L.ws_chunks = off;  off += ceil_div(L.nidx, LB_CHUNK) * sizeof(ChunkState);
    L.ws_cacc = off;    off += ceil_div(L.nidx, LB_CHUNK) * sizeof(unsigned long long);
    L.ws_misc = off;    off += 64 * sizeof(uint64_t);   // totals at +128 B, eb at +256 B, error at +384 B
    off = (off + 255) & ~(uint64_t)255;   // the range partials are read as 16-byte vectors
    L.ws_parts = off;   off += 2 * kRangeParts * sizeof(double);
    L.ws_bitmap = off;  off += (L.nbit + 15) & ~(uint64_t)15;
    L.ws_cmed = off;    off += (4 * L.nb + 15) & ~(uint64_t)15;
    L.ws_osz = off;     off += (2 * L.nb + 15) & ~(uint64_t)15;
    if (!bs_tiled(bs)) {
        off = (off + 255) & ~(uint64_t)255;
        L.ws_gsize = off;   off += (4 * L.nb + 255) & ~(uint64_t)255;
        L.ws_gparts = off;  off += 16 * generic_parts(L.nb);
        L.ws_gidx = off;    off += 8 * index_words(n, w, bs);
    }
    L.ws_bytes = off;
    return L;
}

template <typename T, int BS>
static void launch_main(const CompressParams &p, uint64_t ntiles, hipStream_t st)
{
    // 16 values per lane for the wide fp32 tiles (half the per-slab block
    // statistics and scans of 8), else 8 where a block spans at least 8
    // values, else 4
    constexpr bool WIDE = compress_wide((int)sizeof(T), BS);
    constexpr int LV = (WIDE && sizeof(T) == 4) ? 16 : (BS >= 8 ? 8 : 4);
    constexpr int IPT = WIDE ? INDEX_PER_TILE_WIDE : INDEX_PER_TILE;
    constexpr int TILE = IPT * (sizeof(T) == 4 ? TILE_VALUES_F32 : TILE_VALUES_F64);
    constexpr int SLABS = TILE / (256 * LV);
    hipLaunchKernelGGL((k_compress<T, BS, LV, SLABS, IPT>), dim3((unsigned)ntiles), dim3(256), 0, st, p);
}

template <typename T>
static int dispatch_bs(int bs, const CompressParams &p, uint64_t ntiles, hipStream_t st)
{
    switch (bs) {
    case 4: launch_main<T, 4>(p, ntiles, st); break;
    case 8: launch_main<T, 8>(p, ntiles, st); break;
    case 16: launch_main<T, 16>(p, ntiles, st); break;
    case 32: launch_main<T, 32>(p, ntiles, st); break;
    case 64: launch_main<T, 64>(p, ntiles, st); break;
    case 128: launch_main<T, 128>(p, ntiles, st); break;
    case 256: launch_main<T, 256>(p, ntiles, st); break;
    default: return -1;
    }
    return 0;
}

uint64_t range_workspace_bytes() { return 2 * kRangeParts * sizeof(double) + 256; }

void range_bound_launch(const void *d_in, uint64_t n, int w, float rel, uint8_t *d_ws, float *eb_out,
                        void *d_minmax, hipStream_t st)
{
    if (w == 4) {
        float *part = reinterpret_cast<float *>(d_ws);
        hipLaunchKernelGGL((k_range_partial<float, 256>), dim3(kRangeParts), dim3(256), 0, st,
                           static_cast<const float *>(d_in), n, part, nullptr, 0);
        hipLaunchKernelGGL(k_range_final<float>, dim3(1), dim3(256), 0, st, static_cast<const float *>(d_in),
                           part, kRangeParts, rel, eb_out, static_cast<float *>(d_minmax));
    } else {
        double *part = reinterpret_cast<double *>(d_ws);
        hipLaunchKernelGGL((k_range_partial<double, 256>), dim3(kRangeParts), dim3(256), 0, st,
                           static_cast<const double *>(d_in), n, part, nullptr, 0);
        hipLaunchKernelGGL(k_range_final<double>, dim3(1), dim3(256), 0, st, static_cast<const double *>(d_in),
                           part, kRangeParts, rel, eb_out, static_cast<double *>(d_minmax));
    }
}

void range_launch(const void *d_in, uint64_t n, int w, uint8_t *d_ws, void *d_minmax, hipStream_t st)
{
    if (w == 4) {
        float *part = reinterpret_cast<float *>(d_ws);
        hipLaunchKernelGGL((k_range_partial<float, 256>), dim3(kRangeParts), dim3(256), 0, st,
                           static_cast<const float *>(d_in), n, part, nullptr, 0);
        hipLaunchKernelGGL(k_range_final<float>, dim3(1), dim3(256), 0, st,
                           static_cast<const float *>(d_in), part, kRangeParts, 0.0f,
                           (float *)nullptr, static_cast<float *>(d_minmax));
    } else {
        double *part = reinterpret_cast<double *>(d_ws);
        hipLaunchKernelGGL((k_range_partial<double, 256>), dim3(kRangeParts), dim3(256), 0, st,
                           static_cast<const double *>(d_in), n, part, nullptr, 0);
        hipLaunchKernelGGL(k_range_final<double>, dim3(1), dim3(256), 0, st,
                           static_cast<const double *>(d_in), part, kRangeParts, 0.0f,
                           (float *)nullptr, static_cast<double *>(d_minmax));
    }
}

int compress_launch(const void *d_in, uint64_t n, int w, int rel_mode, float abs_eb, float rel,
                    int bs, uint8_t *d_out, uint8_t *d_ws, uint64_t *d_result, uint64_t *d_index,
                    hipStream_t st, int fast_mode)
{
    const bool ra = fast_mode == 3 || fast_mode == 4;
    if (!(fast_mode == 2 || ra) || (ra && !bs_tiled(bs))) return -1;
    const CompressLayout L = compress_layout(n, w, bs);
    unsigned long long *misc = reinterpret_cast<unsigned long long *>(d_ws + L.ws_misc);
    const uint64_t zero_words = (L.ws_misc + 64 * sizeof(uint64_t)) / 8;
    unsigned long long *zero = reinterpret_cast<unsigned long long *>(d_ws);
    float *eb_dev = nullptr;        // REL: the bound, computed on the device
    const void *rparts = nullptr;   // REL, tiled: the range partials k_compress folds
    if (rel_mode && bs_tiled(bs)) {
        // one launch for the range partials and the clear; every compressor
        // workgroup folds the partials into the bound itself (fp64: the fold
        // of 256 pairs of doubles costs its kernel 3-4 %, the two launches
        // it saves more: REL call 0.292 -> 0.280 ms at 8192^2)
        eb_dev = reinterpret_cast<float *>(misc + 32);
        rparts = d_ws + L.ws_parts;
        if (w == 4)
            hipLaunchKernelGGL((k_range_partial<float, 1024>), dim3(RANGE_LB_PARTS), dim3(1024), 0, st,
                               static_cast<const float *>(d_in), n, (float *)(d_ws + L.ws_parts), zero, zero_words);
        else
            hipLaunchKernelGGL((k_range_partial<double, 1024>), dim3(RANGE_LB_PARTS), dim3(1024), 0, st,
                               static_cast<const double *>(d_in), n, (double *)(d_ws + L.ws_parts), zero, zero_words);
    } else {
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((zero_words + 255) / 256 < 1024 ? (zero_words + 255) / 256 : 1024)),
                           dim3(256), 0, st, zero, zero_words);
    }
    if (rel_mode && !rparts) {
        eb_dev = reinterpret_cast<float *>(misc + 32);
        if (w == 4) {
            float *part = reinterpret_cast<float *>(d_ws + L.ws_parts);
            hipLaunchKernelGGL((k_range_partial<float, 256>), dim3(kRangeParts), dim3(256), 0, st,
                               static_cast<const float *>(d_in), n, part, nullptr, 0);
            hipLaunchKernelGGL(k_range_final<float>, dim3(1), dim3(256), 0, st,
                               static_cast<const float *>(d_in), part, kRangeParts, rel, eb_dev,
                               (float *)nullptr);
        } else {
            double *part = reinterpret_cast<double *>(d_ws + L.ws_parts);
            hipLaunchKernelGGL((k_range_partial<double, 256>), dim3(kRangeParts), dim3(256), 0, st,
                               static_cast<const double *>(d_in), n, part, nullptr, 0);
            hipLaunchKernelGGL(k_range_final<double>, dim3(1), dim3(256), 0, st,
                               static_cast<const double *>(d_in), part, kRangeParts, rel, eb_dev,
                               (double *)nullptr);
        }
    }
    CompressParams p{};
    p.x = d_in;
    p.n = n;
    p.eb = abs_eb;
    p.eb_dev = rparts ? nullptr : eb_dev;
    p.rparts = rparts;
    p.rel = rel;
    p.eb_out = rparts ? eb_dev : nullptr;
    p.payload = d_out + L.region;
    p.bitmap = d_ws + L.ws_bitmap;
    p.cmed = reinterpret_cast<uint32_t *>(d_ws + L.ws_cmed);
    p.tiles = reinterpret_cast<unsigned long long *>(d_ws + L.ws_tiles);
    p.chunks = reinterpret_cast<ChunkState *>(d_ws + L.ws_chunks);
    p.cacc = reinterpret_cast<unsigned long long *>(d_ws + L.ws_cacc);
    p.totals = reinterpret_cast<uint64_t *>(misc + 16);
    p.err = reinterpret_cast<int *>(misc + 48);
    p.index = d_index;
    p.bidx = d_index ? reinterpret_cast<uint16_t *>(d_index + 2 * (L.nidx + 1)) : nullptr;
    p.nidx = L.nidx;
    p.nb = L.nb;
    p.ntiles = L.ntiles;
    p.nbit = L.nbit;
    p.osz = ra ? reinterpret_cast<uint16_t *>(d_ws + L.ws_osz) : nullptr;
    p.fm4 = fast_mode == 4;
    int rc = 0;
    timing_mark_begin(KK_COMPRESS, st);
    if (bs_tiled(bs)) {
        rc = w == 4 ? dispatch_bs<float>(bs, p, L.ntiles, st) : dispatch_bs<double>(bs, p, L.ntiles, st);
    } else if (bs_valid(bs)) {
        GenericCompress g{};
        g.x = d_in;
        g.n = n;
        g.nb = L.nb;
        g.bs = bs;
        g.eb = abs_eb;
        g.eb_dev = eb_dev;
        g.payload = p.payload;
        g.bitmap = p.bitmap;
        g.cmed = p.cmed;
        g.totals = p.totals;
        g.size = reinterpret_cast<uint32_t *>(d_ws + L.ws_gsize);
        g.idx = d_index ? d_index : reinterpret_cast<uint64_t *>(d_ws + L.ws_gidx);
        g.parts = reinterpret_cast<uint64_t *>(d_ws + L.ws_gparts);
        g.nparts = generic_parts(L.nb);
        generic_compress_launch(g, w, st);
    } else {
        rc = -1;
    }
    timing_mark_end(KK_COMPRESS, st);
    if (rc) return rc;
    AssembleParams a{};
    a.out = d_out;
    a.region = L.region;
    a.bitmap = p.bitmap;
    a.cmed = p.cmed;
    a.totals = p.totals;
    a.nbit = L.nbit;
    a.bs = bs;
    a.eb_dev = eb_dev;
    a.eb = abs_eb;
    a.result = d_result;
    a.err = p.err;
    a.ra = ra ? fast_mode : 0;
    a.nb = L.nb;
    a.osz = p.osz;
    const uint64_t work = L.nbit > 4 * L.nb ? L.nbit : 4 * L.nb;
    uint64_t blocks = (work + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_assemble, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return 0;
}

// ---------------------------------------------------------------------------
// batched launchers
// ---------------------------------------------------------------------------

void batch_range_launch(const BatchField *bf, const uint64_t *rbase, int nf, uint64_t nitems, double *part,
                        hipStream_t st)
{
    if (nitems) hipLaunchKernelGGL(k_batch_range, dim3((unsigned)nitems), dim3(256), 0, st, bf, rbase, nf, part);
    hipLaunchKernelGGL(k_batch_range_final, dim3((nf + 255) / 256), dim3(256), 0, st,
                       const_cast<BatchField *>(bf), rbase, nf, part);
}

// Persistent grid: every workgroup resident at once (the round-robin walk
// relies on it), at most one per tile.
template <int BS, int LVF, int LVD, int IPT>
static void launch_batch(const BatchField *bf, const uint64_t *tbase, const uint32_t *tmap, uint64_t total,
                         hipStream_t st)
{
    if (total)
        hipLaunchKernelGGL((k_compress_batch<BS, LVF, LVD, IPT>), dim3((unsigned)total), dim3(256), 0, st, bf, tbase,
                           tmap);
}

template <int BS>
static int batch_shape(int lv, int ipt, const BatchField *bf, const uint64_t *tbase, const uint32_t *tmap,
                       uint64_t total, hipStream_t st)
{
    constexpr int L8 = BS >= 8 ? 8 : 4;
    if (lv == 16) {   // the fp32 fields' wide tiles take 16 values per lane, fp64 8
        if (ipt != 4) return -1;
        if constexpr (BS >= 16 && BS <= 128) launch_batch<BS, 16, 8, 4>(bf, tbase, tmap, total, st);
        else return -1;
        return 0;
    }
    if (lv == 8 && BS >= 8) {
        if (ipt == 4) {
            if constexpr (BS >= 16 && BS <= 128) launch_batch<BS, 8, 8, 4>(bf, tbase, tmap, total, st);
            else return -1;
        } else if (ipt == 2) {
            launch_batch<BS, L8, L8, 2>(bf, tbase, tmap, total, st);
        } else {
            launch_batch<BS, L8, L8, 1>(bf, tbase, tmap, total, st);
        }
        return 0;
    }
    if (lv == 4) {
        if (ipt == 4) return -1;
        if (ipt == 2) launch_batch<BS, 4, 4, 2>(bf, tbase, tmap, total, st);
        else launch_batch<BS, 4, 4, 1>(bf, tbase, tmap, total, st);
        return 0;
    }
    return -1;
}

int batch_compress_launch(int bs, int lv, int ipt, const BatchField *bf, const uint64_t *tbase,
                          const uint32_t *tmap, uint64_t total, hipStream_t st)
{
    if (ipt != 1 && ipt != 2 && ipt != 4) return -1;
    switch (bs) {
    case 4: return batch_shape<4>(lv, ipt, bf, tbase, tmap, total, st);
    case 8: return batch_shape<8>(lv, ipt, bf, tbase, tmap, total, st);
    case 16: return batch_shape<16>(lv, ipt, bf, tbase, tmap, total, st);
    case 32: return batch_shape<32>(lv, ipt, bf, tbase, tmap, total, st);
    case 64: return batch_shape<64>(lv, ipt, bf, tbase, tmap, total, st);
    case 128: return batch_shape<128>(lv, ipt, bf, tbase, tmap, total, st);
    case 256: return batch_shape<256>(lv, ipt, bf, tbase, tmap, total, st);
    default: return -1;
    }
}

void batch_zero_launch(const BatchField *bf, const uint64_t *zbase, int nf, uint64_t nitems, hipStream_t st)
{
    if (nitems) hipLaunchKernelGGL(k_zero_batch, dim3((unsigned)nitems), dim3(256), 0, st, bf, zbase, nf);
}

void batch_assemble_launch(const BatchField *bf, const uint64_t *abase, int nf, uint64_t nitems, hipStream_t st)
{
    if (nitems) hipLaunchKernelGGL(k_assemble_batch, dim3((unsigned)nitems), dim3(256), 0, st, bf, abase, nf);
}

}  // namespace szxamd
