// szx_decompress.hip -- blockwise fast decompressor for gfx950.
//
// A fastMode-2 stream has no block index: the start of non-constant record k
// is the sum of the sizes of records 0..k-1, and a size is only known by
// parsing that record's reqLength byte and lead codes (szxd_float.cc:276-283).
// Decoding therefore works from a record index kept outside the stream:
// per compressor tile, the payload offset of the tile's first record and the
// constant blocks before it, and per block a 16-bit entry (the record's start
// inside the tile, or the constant block's rank inside the tile).  This
// library's compressor writes it as a by-product of its scan; k_index_serial
// rebuilds it for any other stream.  With the index every tile decodes
// independently:
//
//   1. stage the tile's record bytes (16-byte loads) and its constant medians
//      in LDS;
//   2. decode every block in parallel, 4 values per lane, a block being an
//      aligned group of bs/4 lanes (the mirror of the compressor's layout);
//   3. check the index against the records: a scan of the decoded record
//      sizes must reproduce every start, and the tile's records must end
//      exactly at the next tile's first record; a stream or index that fails
//      is rejected (status < 0), and no access leaves the staged bytes.
//
// The decode of one record mirrors szxd_float.cc:34-244 / szxd_double.cc:25-449:
// value i takes its top min(lead_i, rb) bytes from value i-1 and the rest of
// its rb bytes from the stream.  The "from value i-1" chain is resolved in
// parallel: each lane folds its 4 values into a last-writer record of the
// (at most 3) copyable byte positions, and an exclusive segmented scan with
// the last-writer-wins operator gives every lane the bytes it inherits.
#include "szx_common.hpp"
#include "szx_batch.hpp"
#include "szx_index.hpp"
#include "szx_internal.hpp"
#include "szx_wave.hpp"

namespace szxamd {

// Header check: block size, counts and the bitmap/median region must fit.

__device__ inline void check_header(const uint8_t *s, uint64_t len, uint64_t n, int bs_expect,
                                    DecodeHeader *h, int32_t *status)
{
    int32_t rc = 0;
    uint64_t bs = 0, nb = 0, nbit = 0, nconst = 0;
    if (len < HEADER_BYTES) {
        rc = -10;
    } else {
        bs = s[4] ? s[4] : 256;   // byte 0 stands for 256 ((u8)256 in the encoder)
        nb = ceil_div(n, bs);
        nbit = ceil_div(nb, 8);
        for (int k = 0; k < 8; k++) nconst = (nconst << 8) | s[5 + k];
        if ((int)bs != bs_expect) rc = -11;
        else if (nconst > nb) rc = -12;
        else if (HEADER_BYTES + nbit + 4 * nconst > len) rc = -13;
    }
    h->bs = bs; h->nb = nb; h->nbit = nbit; h->nconst = nconst;
    h->q0 = HEADER_BYTES + nbit + 4 * nconst;
    h->qbit = HEADER_BYTES;
    h->qmed = HEADER_BYTES + nbit;
    h->medle = 0;
    h->rc = rc;
    *status = rc;
}

// Random-access container (szxd_float.cc:313-476): [1,1,1,1], big-endian u64
// blockSize and constant count, a u16 size per record, bitmap, medians
// (native order for fastMode 4), records.
__device__ inline void check_header_ra(const uint8_t *s, uint64_t len, uint64_t n, int bs_expect, int fm,
                                       DecodeHeader *h, int32_t *status)
{
    int32_t rc = 0;
    uint64_t bs = 0, nb = 0, nbit = 0, nconst = 0, qbit = 20;
    if (len < 20) {
        rc = -10;
    } else {
        for (int k = 0; k < 8; k++) bs = (bs << 8) | s[4 + k];
        for (int k = 0; k < 8; k++) nconst = (nconst << 8) | s[12 + k];
        if (bs != (uint64_t)bs_expect) {
            rc = -11;
        } else {
            nb = ceil_div(n, bs);
            nbit = ceil_div(nb, 8);
            if (nconst > nb) rc = -12;
            else {
                qbit = 20 + 2 * (nb - nconst);
                if (qbit + nbit + 4 * nconst > len) rc = -13;
            }
        }
    }
    if (rc) bs = (uint64_t)bs_expect;
    h->bs = bs; h->nb = nb; h->nbit = nbit; h->nconst = nconst;
    h->qbit = qbit;
    h->qmed = qbit + nbit;
    h->q0 = qbit + nbit + 4 * nconst;
    h->medle = fm == 4;
    h->rc = rc;
    *status = rc;
}

__global__ void k_check_header(const uint8_t *s, uint64_t len, uint64_t n, int bs_expect, int fm,
                               DecodeHeader *h, int32_t *status)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (fm == 3 || fm == 4) check_header_ra(s, len, n, bs_expect, fm, h, status);
        else check_header(s, len, n, bs_expect, h, status);
    }
}

// Whole-vector fp32 output stores are non-temporal: the decoded array is
// written once and not read back by this kernel, so it streams past the
// caches (k_decode_tiles 0.178 -> 0.136 ms at 512^3 fp32).
typedef float f32x4_t __attribute__((ext_vector_type(4)));
template <typename T> struct Store4;
template <> struct Store4<float> {
    __device__ static void store_full(float *o, uint64_t e0, const float v[4])
    {
        f32x4_t w = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(w, reinterpret_cast<f32x4_t *>(o + e0));
    }
    __device__ static void store(float *o, uint64_t e0, uint64_t n, const float v[4])
    {
        if (e0 + 4 <= n) {
            *reinterpret_cast<float4 *>(o + e0) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
            for (int k = 0; k < 4; k++)
                if (e0 + k < n) o[e0 + k] = v[k];
        }
    }
};
template <> struct Store4<double> {
    // plain stores (the fp64 tile decoder writes whole tiles through LDS
    // instead; non-temporal stores here ran at 1.3-1.4 TB/s against 2.2-2.5)
    __device__ static void store_full(double *o, uint64_t e0, const double v[4])
    {
        *reinterpret_cast<double2 *>(o + e0) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2 *>(o + e0 + 2) = make_double2(v[2], v[3]);
    }
    __device__ static void store(double *o, uint64_t e0, uint64_t n, const double v[4])
    {
        if (e0 + 4 <= n) {
            *reinterpret_cast<double2 *>(o + e0) = make_double2(v[0], v[1]);
            *reinterpret_cast<double2 *>(o + e0 + 2) = make_double2(v[2], v[3]);
        } else {
            for (int k = 0; k < 4; k++)
                if (e0 + k < n) o[e0 + k] = v[k];
        }
    }
};

__device__ inline float be_f32(const uint8_t *p)
{
    return bits_f32(((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) |
                    (uint32_t)p[3]);
}

// Exclusive prefix sum over the 64 lanes of a wave.
__device__ inline uint32_t wave_excl_scan(uint32_t v, int lane)
{
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    return inc - v;
}

// Copy bytes [src, src + nbytes) of global memory into LDS at dst + (src & 15)
// (dst 16-byte aligned): 16-byte loads for the aligned chunks that lie wholly
// inside [lo, hi) -- the caller's buffer -- and single bytes for the rest.
__device__ inline void stage_bytes(uint8_t *dst, const uint8_t *src, uint32_t nbytes,
                                   const uint8_t *lo, const uint8_t *hi, int tid)
{
    const uintptr_t a = (uintptr_t)src & ~(uintptr_t)15;
    const uint32_t head = (uint32_t)((uintptr_t)src - a);
    const uint32_t nchunk = (head + nbytes + 15) / 16;
    for (uint32_t c = tid; c < nchunk; c += 256) {
        const uint8_t *g = reinterpret_cast<const uint8_t *>(a + 16 * (uintptr_t)c);
        if (g >= lo && g + 16 <= hi)
            *reinterpret_cast<uint4 *>(dst + 16 * c) = *reinterpret_cast<const uint4 *>(g);
    }
    // the first and last chunks may be partial: one byte per thread
    if (tid < 32 && nchunk) {
        const uint32_t c = tid < 16 ? 0 : nchunk - 1;
        const uint32_t k = tid & 15;
        const uint8_t *g = reinterpret_cast<const uint8_t *>(a + 16 * (uintptr_t)c) + k;
        if (g >= lo && g < hi) dst[16 * c + k] = *g;
    }
}

// Everything the slab decoder reads from the tile's LDS staging.
struct TileLds {
    const uint8_t *st;          // staged record bytes (16-byte aligned base)
    uint32_t rdel;              // offset of the tile's first record byte in st
    const uint8_t *meds;        // the tile's constant medians (big-endian)
    uint32_t nct, tb;           // constant blocks, record bytes of the tile
    const uint16_t *start;      // per block: record start / constant rank
    uint16_t *size;             // per block: decoded record size
    const uint8_t *state;       // per block: non-constant
    int *err;
    int medle;                  // medians in native (little-endian) order
};

// Decode the LV values of one lane: lane lseg of the BS/LV lanes holding
// block lb, first value e0.  FULL: the whole block lies inside the array.
// KEEP: leave the values in vout for the caller to store (FULL only).
template <typename T, int BS, int LV, bool FULL, bool KEEP = false>
__device__ inline void decode_lane(const TileLds &L, uint64_t e0, uint64_t n, int lb, int lane, int lseg,
                                   T *out, int exact, T *vout = nullptr)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    constexpr int SEG = BS / LV;
    constexpr int NLB = LV / 4;          // lead-code bytes per lane
    static_assert(!KEEP || FULL, "kept values are whole blocks");
    T v[LV];
    if (!L.state[lb]) {
        const uint32_t rank = L.start[lb];
        float med = 0.0f;
        if (rank < L.nct) {
            const uint8_t *p = L.meds + 4 * rank;
            const uint32_t be = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) |
                                (uint32_t)p[3];
            med = bits_f32(L.medle ? bswap32(be) : be);
        } else if (lseg == 0) {
            *L.err = 1;
        }
#pragma unroll
        for (int k = 0; k < LV; k++) v[k] = (T)med;
    } else {
        const uint64_t bstart = e0 - (uint64_t)LV * lseg;
        const int nel = FULL ? BS : (n - bstart >= (uint64_t)BS ? BS : (int)(n - bstart));
        const int nlead = (nel + 3) >> 2;
        const uint32_t start = L.start[lb];
        const bool head_ok = start + 5u + (uint32_t)nlead <= L.tb;   // fixed part staged
        const uint8_t *rec = L.st + L.rdel + (head_ok ? start : 0u);
        const int rl = head_ok ? rec[0] : 1;
        const bool rl_ok = head_ok && rl >= 1 && rl <= 8 * W;
        const int rb = rl_ok ? (rl + 7) >> 3 : 1;
        const int shift = (8 - (rl & 7)) & 7;
        const int lo = W - rb;
        const float med = bits_f32(((uint32_t)rec[1] << 24) | ((uint32_t)rec[2] << 16) |
                                   ((uint32_t)rec[3] << 8) | (uint32_t)rec[4]);
        uint32_t code[NLB];
#pragma unroll
        for (int b = 0; b < NLB; b++)
            code[b] = (rl_ok && (lseg * NLB + b) < nlead) ? rec[5 + lseg * NLB + b] : 0u;
        // lead bytes lk = min(code, rb) and mid bytes m = rb - lk per value
        // (past the array end: no bytes)
        uint32_t lk[LV], mk[LV], tot = 0;
#pragma unroll
        for (int k = 0; k < LV; k++) {
            const uint32_t ld = (code[k >> 2] >> (6 - 2 * (k & 3))) & 3u;
            lk[k] = ld < (uint32_t)rb ? ld : (uint32_t)rb;
            mk[k] = (FULL || e0 + k < n) ? (uint32_t)rb - lk[k] : 0u;
            tot += mk[k];
        }
        const uint32_t inc = seg_scan_incl<SEG>(tot, lane);
        const uint32_t rsize = 5u + (uint32_t)nlead + seg_last<SEG>(inc);
        const bool ok = rl_ok && start + rsize <= L.tb;
        if (lseg == 0) {
            if (ok) L.size[lb] = (uint16_t)rsize;
            else *L.err = 1;
        }
        // Value k's m bytes start at stage byte p (p runs through the lane's
        // bytes; a bad record reads from byte 0 on and is flagged above).  They
        // land on word bytes [W - rb, W - lk): the gathered word shifted up by
        // lo = W - rb bytes supplies bytes [0, W - lk) (zeros below lo), value
        // k-1 the top lk bytes -- one bit-field insert.  Bits the final shift
        // drops are never masked.  Pass 1 runs with a zero carry-in: t0[k] is
        // value k's word if nothing came from the previous lane; inh[k] marks
        // the top bytes value k still inherits from before the lane.
        // (two or three aligned LDS words and a byte funnel per value:
        // unaligned LDS reads are slow)
        const uint32_t *stw = reinterpret_cast<const uint32_t *>(L.st);
        uint32_t p = ok ? (uint32_t)(rec - L.st) + 5u + (uint32_t)nlead + (inc - tot) : 0u;
        const int lo8 = 8 * lo;
        U t0[LV], inh[LV];
        U tp = 0, ih = ~(U)0;
#pragma unroll
        for (int k = 0; k < LV; k++) {
            const uint32_t *w = stw + (p >> 2);
            const uint32_t sh = p & 3u;
            U xk;
            if constexpr (W == 4) {
                xk = __builtin_amdgcn_alignbyte(w[1], w[0], sh);
            } else {
                const uint32_t d0 = w[0], d1 = w[1], d2 = w[2];
                xk = (U)__builtin_amdgcn_alignbyte(d1, d0, sh) | ((U)__builtin_amdgcn_alignbyte(d2, d1, sh) << 32);
            }
            U xs = xk << lo8;
            if constexpr (W == 8) {
                if (!exact && rb == 5 && lk[k] == 0 && mk[k])   // szxd_double.cc:244 reads q[3]
                    xs = (xs & ~((U)0xFF << 56)) | (((xk >> 24) & 0xFF) << 56);
            }
            const U take = ~(U)0 >> (lk[k] << 3);   // bytes below W - lk
            tp = (xs & take) | (tp & ~take);
            ih &= ~take;
            t0[k] = tp;
            inh[k] = ih;
            p += mk[k];   // at most 8 W bytes past the start, inside the stage
        }
        // the lane's copyable bytes W-3..W-1: the ones it wrote (no longer
        // inherited) and their values; a last-writer scan over the segment
        // then gives every lane what it inherits from the lanes before it
        const uint32_t wm = ~(uint32_t)(ih >> (8 * (W - 3))) & 0xFFFFFFu;
        uint32_t cd = (uint32_t)(tp >> (8 * (W - 3))) & wm, cm = wm;
        seg_scan_last_writer<SEG>(cd, cm, lane);
        const uint32_t cin = seg_prev<SEG>(cd, lane);
        const U cw = (U)cin << (8 * (W - 3));
        const T mT = (T)med;
#pragma unroll
        for (int k = 0; k < LV; k++) v[k] = from_bits<T>((t0[k] | (cw & inh[k])) << shift) + mT;
    }
    if (KEEP) {
#pragma unroll
        for (int k = 0; k < LV; k++) vout[k] = v[k];
    } else if (FULL) {
#pragma unroll
        for (int k = 0; k < LV; k += 4) Store4<T>::store_full(out, e0 + k, v + k);
    } else {
#pragma unroll
        for (int k = 0; k < LV; k++)
            if (e0 + k < n) out[e0 + k] = v[k];
    }
}

// Everything one tile decode needs (one stream; the batch keeps one per field).
struct DecodeParams {
    const uint8_t *s;           // stream
    uint64_t len, n;            // stream bytes, values
    const DecodeHeader *h;      // checked header
    const uint64_t *index;      // record index: tile pairs
    const uint16_t *bidx;       // record index: per-block entries
    int32_t *status;
    void *out;
    int exact;
};

// Where one index tile's bytes are, from the checked header and the record
// index: what a tile decode needs before it can stage its records.  flags:
// 0 ready, 1 the header check failed (nothing to do), 2 the index entry is
// out of range (status -20).
struct TileSrc {
    const uint8_t *rsrc;        // the tile's first record byte
    const uint8_t *msrc;        // its first constant median
    const uint8_t *lo, *hi;     // the stream's bytes
    const uint8_t *bmp;         // the bitmap byte of its first block
    const uint16_t *bidx;       // the record-index entry of its first block
    uint32_t tb, nct;           // record bytes, constant blocks
    uint32_t nvb;               // blocks of the tile inside the array
    uint32_t flags;
    uint32_t field, tile;       // batched decode: the field and its tile
    uint32_t medle, pad;
};

template <typename T, int BS, int LV> struct DecodeShape {
    static constexpr int W = Word<T>::W;
    static constexpr int TILE = sizeof(T) == 4 ? TILE_VALUES_F32 : TILE_VALUES_F64;
    static constexpr int NBLK = TILE / BS;
    static constexpr int STAGE = TILE * W + NBLK * (5 + BS / 4);
};

template <typename T, int BS, int LV> struct DecodeLds {
    using S = DecodeShape<T, BS, LV>;
    __attribute__((aligned(16))) uint8_t st[S::STAGE + 64];
    __attribute__((aligned(16))) uint8_t med[4 * S::NBLK + 48];
    uint16_t start[S::NBLK];   // record start, or constant rank within the tile
    uint16_t size[S::NBLK];    // decoded record size (0: constant block)
    uint8_t state[S::NBLK];
    int err;
};

// Decode index tile `tile` of one stream (TILE values, NBLK blocks); each
// lane decodes LV consecutive values per slab of 256 * LV values.
// The bytes of index tile `tile` of one stream (NBLK blocks of the tile shape).
template <int NBLK, int STAGE>
__device__ inline TileSrc tile_src(const DecodeParams &p, uint64_t tile)
{
    TileSrc t{};
    const DecodeHeader *h = p.h;
    // gate on the header check, which an earlier kernel wrote
    if (h->rc != 0) {
        t.flags = 1;
        return t;
    }
    const uint64_t *index = p.index;
    const uint64_t nb = h->nb, q0 = h->q0, nconst = h->nconst;
    const uint64_t gb0 = tile * NBLK;
    const uint64_t off = q0 + index[2 * tile];
    const uint64_t end = q0 + index[2 * tile + 2];
    const uint64_t cbase = index[2 * tile + 1];
    const uint64_t cend = index[2 * tile + 3];
    if (end < off || end > p.len || end - off > (uint64_t)STAGE || cend < cbase ||
        cend - cbase > (uint64_t)NBLK || cend > nconst) {
        t.flags = 2;
        return t;
    }
    t.rsrc = p.s + off;
    t.msrc = p.s + h->qmed + 4 * cbase;   // inside [0, q0) <= len
    t.lo = p.s;
    t.hi = p.s + p.len;
    t.bmp = p.s + h->qbit + (gb0 >> 3);   // NBLK is a multiple of 8
    t.bidx = p.bidx + gb0;
    t.tb = (uint32_t)(end - off);
    t.nct = (uint32_t)(cend - cbase);
    t.nvb = gb0 >= nb ? 0u : (nb - gb0 < (uint64_t)NBLK ? (uint32_t)(nb - gb0) : (uint32_t)NBLK);
    t.medle = (uint32_t)h->medle;
    t.flags = 0;
    return t;
}

template <typename T, int BS, int LV>
__device__ inline void decode_tile_core(const DecodeParams &p, uint64_t tile, const TileSrc &ts,
                                        DecodeLds<T, BS, LV> &lds);

template <typename T, int BS, int LV>
__device__ inline void decode_tile(const DecodeParams &p, uint64_t tile, DecodeLds<T, BS, LV> &lds)
{
    using S = DecodeShape<T, BS, LV>;
    decode_tile_core<T, BS, LV>(p, tile, tile_src<S::NBLK, S::STAGE>(p, tile), lds);
}

// Decode index tile `tile` from its bytes ts (tile_src).
template <typename T, int BS, int LV>
__device__ inline void decode_tile_core(const DecodeParams &p, uint64_t tile, const TileSrc &ts,
                                        DecodeLds<T, BS, LV> &lds)
{
    using S = DecodeShape<T, BS, LV>;
    constexpr int SEG = BS / LV;
    constexpr int TILE = S::TILE;
    constexpr int SLAB = 256 * LV;
    constexpr int SLABS = TILE / SLAB;
    constexpr int BPS = SLAB / BS;
    constexpr int NBLK = S::NBLK;
    constexpr int STAGE = S::STAGE;
    constexpr int PER_LANE = (NBLK + 63) / 64;   // blocks per lane in the check scan
    static_assert(SLABS >= 1 && TILE % SLAB == 0, "slabs tile the tile");

    uint8_t *st = lds.st;
    uint8_t *s_med = lds.med;
    uint16_t *s_start = lds.start;
    uint16_t *s_size = lds.size;
    uint8_t *s_state = lds.state;
    int &s_err = lds.err;
    const uint64_t n = p.n;
    int32_t *status = p.status;
    T *out = static_cast<T *>(p.out);
    const int exact = p.exact;

    // The flags come from the header check and the index, which earlier
    // kernels wrote, so every thread of the workgroup takes the same branch
    // (the status word itself may change while this kernel runs: other
    // tiles report into it).
    if (ts.flags == 1) return;
    const int tid = threadIdx.x;
    if (ts.flags != 0) {
        if (tid == 0) atomicMin(status, -20);
        return;
    }
    const int lane = tid & 63;
    const int lseg = tid % SEG;
    const uint32_t tb = ts.tb, nct = ts.nct, nvb = ts.nvb;
    const uint8_t *rsrc = ts.rsrc;
    const uint8_t *msrc = ts.msrc;
    const uint32_t rdel = (uint32_t)((uintptr_t)rsrc & 15), mdel = (uint32_t)((uintptr_t)msrc & 15);
    if (tid == 0) s_err = 0;
    stage_bytes(st, rsrc, tb, ts.lo, ts.hi, tid);
    stage_bytes(s_med, msrc, 4 * nct, ts.lo, ts.hi, tid);
    for (int lb = tid; lb < NBLK; lb += 256) {
        s_state[lb] = (uint32_t)lb < nvb ? (ts.bmp[lb >> 3] >> (7 - (lb & 7))) & 1 : 0;
        s_start[lb] = (uint32_t)lb < nvb ? ts.bidx[lb] : 0;
        s_size[lb] = 0;
    }
    __syncthreads();

    const TileLds L{st, rdel, s_med + mdel, nct, tb, s_start, s_size, s_state, &s_err, (int)ts.medle};
    const bool full = (tile + 1) * (uint64_t)TILE <= n;
    if (full && sizeof(T) == 8) {
        // fp64: a lane's values are 8 LV consecutive bytes, and its own four
        // 16-byte stores per 8 values leave every line partly written per
        // instruction.  The tile's values go through the stage instead (free
        // once every lane has decoded) and out in whole 16-byte chunks
        // across the wave, non-temporal (8192^2 fp64: 0.214 -> 0.118 ms;
        // fp32, whose lanes write 32 bytes each, gains nothing from it).
        T vv[SLABS][LV];
#pragma unroll
        for (int u = 0; u < SLABS; u++) {
            const uint64_t e0 = tile * TILE + (uint64_t)u * SLAB + (uint64_t)LV * tid;
            decode_lane<T, BS, LV, true, true>(L, e0, n, u * BPS + tid / SEG, lane, lseg, out, exact, vv[u]);
        }
        __syncthreads();
        static_assert(TILE * (int)sizeof(T) <= STAGE, "the tile's values fit the stage");
        typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
        T *so = reinterpret_cast<T *>(st);
#pragma unroll
        for (int u = 0; u < SLABS; u++)
#pragma unroll
            for (int k = 0; k < LV; k += 16 / (int)sizeof(T)) {
                u32x4_t w;
                __builtin_memcpy(&w, &vv[u][k], 16);
                *reinterpret_cast<u32x4_t *>(so + u * SLAB + LV * tid + k) = w;
            }
        __syncthreads();
        const u32x4_t *sv = reinterpret_cast<const u32x4_t *>(so);
        u32x4_t *ov = reinterpret_cast<u32x4_t *>(out + tile * TILE);
        constexpr int CHUNKS = TILE * (int)sizeof(T) / 16;
#pragma unroll
        for (int c = tid; c < CHUNKS; c += 256) __builtin_nontemporal_store(sv[c], ov + c);
    } else {
#pragma unroll
        for (int u = 0; u < SLABS; u++) {
            const uint64_t e0 = tile * TILE + (uint64_t)u * SLAB + (uint64_t)LV * tid;
            const int lb = u * BPS + tid / SEG;
            if (full) {
                decode_lane<T, BS, LV, true>(L, e0, n, lb, lane, lseg, out, exact);
            } else {
                if (e0 - (uint64_t)LV * lseg >= n) continue;   // the whole block is past the end
                decode_lane<T, BS, LV, false>(L, e0, n, lb, lane, lseg, out, exact);
            }
        }
    }
    __syncthreads();

    // the index must be what the records say: starts are the exclusive scan
    // of the record sizes, constant ranks count the constant blocks, and the
    // tile's records and medians end where the next tile's begin
    if (tid < 64) {
        uint32_t bytes = 0, consts = 0;
#pragma unroll
        for (int k = 0; k < PER_LANE; k++) {
            const int lb = lane * PER_LANE + k;
            if (lb < NBLK && (uint32_t)lb < nvb) {
                if (s_state[lb]) bytes += s_size[lb];
                else consts++;
            }
        }
        uint32_t eb = wave_excl_scan(bytes, lane), ec = wave_excl_scan(consts, lane);
        int bad = 0;
#pragma unroll
        for (int k = 0; k < PER_LANE; k++) {
            const int lb = lane * PER_LANE + k;
            if (lb < NBLK && (uint32_t)lb < nvb) {
                if (s_state[lb]) {
                    bad |= s_start[lb] != eb;
                    eb += s_size[lb];
                } else {
                    bad |= s_start[lb] != ec;
                    ec++;
                }
            }
        }
        if (lane == 63) bad |= eb != tb || ec != nct;
        if (__ballot(bad) && lane == 0) s_err = 1;
    }
    __syncthreads();
    if (s_err && tid == 0) atomicMin(status, -21);
}

template <typename T, int BS, int LV>
__global__ __launch_bounds__(256) void k_decode_tiles(DecodeParams p)
{
    __shared__ DecodeLds<T, BS, LV> lds;
    decode_tile<T, BS, LV>(p, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// Batched decode: every field of a batched compress in one launch, one
// workgroup per index tile of every field (the field found by a binary search
// of the tile bases), so tiles of all fields overlap on the CUs exactly as a
// single field's tiles do (the field of each tile comes from a table built
// with the plan).  The stream of a field is found on the device from
// the compressor's result words, so a compress and its decode can be enqueued
// (or captured) together.
// ---------------------------------------------------------------------------

__global__ void k_decode_header_batch(const BatchDecodeField *f, int nf, int bs)
{
    const int fi = blockIdx.x * blockDim.x + threadIdx.x;
    if (fi >= nf) return;
    const BatchDecodeField &d = f[fi];
    const uint8_t *s = d.out + d.result[0];
    check_header(s, d.result[1], d.n, bs, d.h, d.status);
}

// Every tile's bytes (tile_src), one thread per tile, so that a decode
// workgroup starts staging its records after one load (the field's result
// words, header and index entries are resolved here, off its path).
template <int BS, int LV>
__global__ __launch_bounds__(256) void k_decode_setup_batch(const BatchDecodeField *f, const uint64_t *tbase,
                                                            const uint32_t *tmap, TileSrc *recs, uint64_t total)
{
    using S32 = DecodeShape<float, BS, LV>;
    using S64 = DecodeShape<double, BS, LV>;
    const uint64_t g = blockIdx.x * 256ull + threadIdx.x;
    if (g >= total) return;
    const uint32_t fi = tmap[g];   // the field holding tile g
    const BatchDecodeField &d = f[fi];
    const DecodeParams p{d.out + d.result[0], d.result[1], d.n, d.h, d.index,
                         reinterpret_cast<const uint16_t *>(d.index + 2 * (d.ntidx + 1)), d.status, d.dst, 0};
    const uint64_t t = g - tbase[fi];
    TileSrc r = d.w == 4 ? tile_src<S32::NBLK, S32::STAGE>(p, t) : tile_src<S64::NBLK, S64::STAGE>(p, t);
    r.field = fi;
    r.tile = (uint32_t)t;
    recs[g] = r;
}

template <int BS, int LV>
__global__ __launch_bounds__(256) void k_decode_batch(const BatchDecodeField *f, const TileSrc *recs, int exact)
{
    __shared__ union Lds {
        DecodeLds<float, BS, LV> f;
        DecodeLds<double, BS, LV> d;
    } lds;
    const TileSrc r = recs[blockIdx.x];
    const BatchDecodeField &d = f[r.field];
    // the core reads only the values count, status, output and mode of p
    const DecodeParams p{nullptr, 0, d.n, nullptr, nullptr, nullptr, d.status, d.dst, exact};
    if (d.w == 4) decode_tile_core<float, BS, LV>(p, r.tile, r, lds.f);
    else decode_tile_core<double, BS, LV>(p, r.tile, r, lds.d);
}

template <int BS>
static void launch_decode_batch(const BatchDecodeField *f, const uint64_t *tbase, const uint32_t *tmap,
                                TileSrc *recs, uint64_t total, int exact, hipStream_t st)
{
    constexpr int LV = BS >= 8 ? 8 : 4;
    if (!total) return;
    hipLaunchKernelGGL((k_decode_setup_batch<BS, LV>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, f,
                       tbase, tmap, recs, total);
    hipLaunchKernelGGL((k_decode_batch<BS, LV>), dim3((unsigned)total), dim3(256), 0, st, f, recs, exact);
}

uint64_t batch_decode_records_bytes(uint64_t total) { return total * sizeof(TileSrc); }

int batch_decode_launch(int bs, const BatchDecodeField *f, const uint64_t *tbase, const uint32_t *tmap,
                        void *recs, int nf, uint64_t total, int exact, hipStream_t st)
{
    TileSrc *r = static_cast<TileSrc *>(recs);
    hipLaunchKernelGGL(k_decode_header_batch, dim3((nf + 255) / 256), dim3(256), 0, st, f, nf, bs);
    switch (bs) {
    case 4: launch_decode_batch<4>(f, tbase, tmap, r, total, exact, st); break;
    case 8: launch_decode_batch<8>(f, tbase, tmap, r, total, exact, st); break;
    case 16: launch_decode_batch<16>(f, tbase, tmap, r, total, exact, st); break;
    case 32: launch_decode_batch<32>(f, tbase, tmap, r, total, exact, st); break;
    case 64: launch_decode_batch<64>(f, tbase, tmap, r, total, exact, st); break;
    case 128: launch_decode_batch<128>(f, tbase, tmap, r, total, exact, st); break;
    case 256: launch_decode_batch<256>(f, tbase, tmap, r, total, exact, st); break;
    default: return -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

DecompressLayout decompress_layout(uint64_t n, int w, int bs, uint64_t stream_bytes)
{
    DecompressLayout L{};
    L.ws_hdr = 0;
    L.ws_index = 256;
    L.ws_build = (L.ws_index + 8 * index_words(n, w, bs) + 255) & ~255ull;
    L.ws_bytes = L.ws_build + ix_layout(n, w, bs, stream_bytes).bytes;
    return L;
}

template <typename T, int BS>
static void launch_decode(const uint8_t *s, uint64_t len, uint64_t n, const DecodeHeader *h,
                          const uint64_t *index, int32_t *status, void *out, int exact,
                          uint64_t ntiles, hipStream_t st)
{
    // 8 values per lane where a block spans at least 8 values (fp64 tiles of
    // 2048 values take one slab of 256 x 8), else 4
    constexpr int LV = BS >= 8 ? 8 : 4;
    const DecodeParams p{s, len, n, h, index,
                         reinterpret_cast<const uint16_t *>(index + 2 * (ntiles + 1)), status, out,
                         exact};
    hipLaunchKernelGGL((k_decode_tiles<T, BS, LV>), dim3((unsigned)ntiles), dim3(256), 0, st, p);
}

template <typename T>
static int dispatch_decode(int bs, const uint8_t *s, uint64_t len, uint64_t n,
                           const DecodeHeader *h, const uint64_t *index, int32_t *status,
                           void *out, int exact, uint64_t ntiles, hipStream_t st)
{
    switch (bs) {
    case 4: launch_decode<T, 4>(s, len, n, h, index, status, out, exact, ntiles, st); break;
    case 8: launch_decode<T, 8>(s, len, n, h, index, status, out, exact, ntiles, st); break;
    case 16: launch_decode<T, 16>(s, len, n, h, index, status, out, exact, ntiles, st); break;
    case 32: launch_decode<T, 32>(s, len, n, h, index, status, out, exact, ntiles, st); break;
    case 64: launch_decode<T, 64>(s, len, n, h, index, status, out, exact, ntiles, st); break;
    case 128: launch_decode<T, 128>(s, len, n, h, index, status, out, exact, ntiles, st); break;
    case 256: launch_decode<T, 256>(s, len, n, h, index, status, out, exact, ntiles, st); break;
    default: return -1;
    }
    return 0;
}

int decompress_launch(const uint8_t *d_stream, uint64_t stream_bytes, uint64_t n, int w, int bs,
                      void *d_out, uint8_t *d_ws, uint64_t ws_bytes, const uint64_t *d_index,
                      int exact, int32_t *d_status, hipStream_t st, int fast_mode)
{
    const bool ra = fast_mode == 3 || fast_mode == 4;
    if (!bs_valid(bs) || !(fast_mode == 2 || ra) || (ra && !bs_tiled(bs))) return -1;
    const DecompressLayout L = decompress_layout(n, w, bs, stream_bytes);
    if (ws_bytes < (d_index ? L.ws_index : L.ws_bytes)) return -1;   // with an index only the header is kept
    DecodeHeader *h = reinterpret_cast<DecodeHeader *>(d_ws + L.ws_hdr);
    const uint64_t tv = tile_values(w);
    const uint64_t ntiles = ceil_div(n, tv);
    hipLaunchKernelGGL(k_check_header, dim3(1), dim3(64), 0, st, d_stream, stream_bytes, n, bs, fast_mode, h,
                       d_status);
    const uint64_t *index = d_index;
    if (!index) {
        uint64_t *built = reinterpret_cast<uint64_t *>(d_ws + L.ws_index);
        uint16_t *bbuilt =
            bs_tiled(bs) ? reinterpret_cast<uint16_t *>(built + 2 * (index_tiles(n, w) + 1)) : nullptr;
        timing_mark_begin(KK_INDEX, st);
        if (ra)
            index_from_sizes_launch(d_stream, stream_bytes, n, w, bs, h, built, bbuilt, d_ws + L.ws_build,
                                    d_status, st);
        else
            index_build_launch(d_stream, stream_bytes, n, w, bs, h, built, bbuilt, d_ws + L.ws_build,
                               d_status, st);
        timing_mark_end(KK_INDEX, st);
        index = built;
    }
    if (!bs_tiled(bs)) {
        const GenericDecode g{d_stream, stream_bytes, n, h, index, d_status, d_out, exact};
        timing_mark_begin(KK_DECODE, st);
        generic_decode_launch(g, w, bs, ceil_div(n, (uint64_t)bs), st);
        timing_mark_end(KK_DECODE, st);
        return 0;
    }
    if (ntiles == 0) return 0;
    timing_mark_begin(KK_DECODE, st);
    const int rc = w == 4 ? dispatch_decode<float>(bs, d_stream, stream_bytes, n, h, index,
                                                   d_status, d_out, exact, ntiles, st)
                          : dispatch_decode<double>(bs, d_stream, stream_bytes, n, h, index,
                                                    d_status, d_out, exact, ntiles, st);
    timing_mark_end(KK_DECODE, st);
    return rc;
}

}  // namespace szxamd
