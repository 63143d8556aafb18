// szx_cuszx.hip -- reader and writer of the stream the reference's CUDA
// codec (cuSZx, fp32) writes (the writer is at the end of the file).
//
// Stream (cuszx_entry.cu:20-70):
//   [1, 1, 1, 0, (u8)bs], big-endian u64 constant-block count, the state
//   bitmap over the nb = n / bs full blocks (MSB first, 1 = non-constant), the
//   n % bs trailing values raw, a native f32 median per constant block, a
//   native i16 mid-byte count per non-constant block, then per non-constant
//   block a record: native f32 median, reqLength byte, bs/4 lead-code bytes
//   (2 bits a value, MSB first), mid bytes.
//
// The reference decodes on the host and the GPU in turn
// (cuszx_entry.cu:129-209): the host walks the records into 4 bs-byte slots,
// one CUDA block of bs threads decodes a slot (cuszxd_float.cu:161-294), and
// the host places constant and decoded blocks.  Here the whole path runs on
// the device:
//   k_cz_header   one lane: the section offsets, checks, status
//   k_cz_part / k_cz_top / k_cz_apply
//                 two exclusive scans side by side: non-constant blocks
//                 before every 64-block bitmap word, and record offsets from
//                 the i16 counts
//   k_cz_decode   a workgroup unit of 256 / S blocks, a block a segment of S
//                 lanes (S = the smallest power of two with 4 S >= bs), 4
//                 values a lane; the unit's records (contiguous) staged in LDS
//   k_cz_tail     the trailing values
//
// The reference's kernel takes each value's inherited bytes from the last
// value at or before it that carried the byte itself (its per-byte max scan
// of source positions, cuszxd_float.cu:125-159, 271-287), so the segment
// carries those bytes with the last-writer scan of the fastMode-2 decoder.
// Two corners of that kernel are kept: with one significant byte every value
// reads its own top byte (the source position is set for every value,
// cuszxd_float.cu:235-239), and a slot past its record reads zeros (the slot
// buffer is cleared before the record is copied in, cuszx_entry.cu:148-149).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szx_common.hpp"
#include "szx_internal.hpp"

namespace szxamd {

struct CzHeader {
    uint64_t n, nb, nconst, nc, rem, nw;
    uint64_t q_bits, q_rem, q_med, q_off, q_rec;
    int32_t bs, msize, rc, pad;
};

// Status codes (the first error wins through atomicMin).
constexpr int32_t CZ_BAD_HEADER = -41;   // short stream, block size, constant count
constexpr int32_t CZ_BAD_RECORD = -42;   // a record length outside [0, 4 bs]
constexpr int32_t CZ_BAD_BITMAP = -43;   // bitmap ones differ from the non-constant count
constexpr int32_t CZ_TRUNCATED = -44;    // records past the end of the stream
constexpr int32_t CZ_BAD_READ = -45;     // a mid-byte position outside its block

constexpr int CZ_PER = 16;               // scan items per lane
constexpr int CZ_CH = 256 * CZ_PER;      // scan items per chunk
constexpr int CZ_SPAN = 4096;            // record bytes of one decode unit: (256 / S) * 4 bs <= 16 * 256

struct CzParams {
    const uint8_t *s;
    uint64_t len, n;
    CzHeader *h;
    uint64_t *wbase;   // per bitmap word of 64 blocks: non-constant blocks before it; nw + 1
    uint64_t *roff;    // per non-constant block: record offset after q_rec; nc + 1
    uint64_t *parts;   // per mode and chunk: scan partials
    uint64_t nparts;
    float *out;
    int32_t *status;
};

__global__ void k_cz_header(const uint8_t *s, uint64_t len, uint64_t n, CzHeader *h, int32_t *status)
{
    if (threadIdx.x != 0) return;
    CzHeader H = {};
    H.n = n;
    int32_t rc = 0;
    if (len < 13) {
        rc = CZ_BAD_HEADER;
    } else {
        const int bs = s[4];   // cuszx_entry.cu:136
        if (bs == 0 || bs % 32 != 0 || bs > 224) {
            rc = CZ_BAD_HEADER;
        } else {
            uint64_t nconst = 0;
            for (int k = 0; k < 8; k++) nconst = (nconst << 8) | s[5 + k];
            H.bs = bs;
            H.msize = 5 + bs / 4;
            H.nb = n / (uint64_t)bs;
            H.rem = n % (uint64_t)bs;
            H.nw = (H.nb + 63) / 64;
            if (nconst > H.nb) {
                rc = CZ_BAD_HEADER;
            } else {
                H.nconst = nconst;
                H.nc = H.nb - nconst;
                H.q_bits = 13;
                H.q_rem = H.q_bits + (H.nb + 7) / 8;
                H.q_med = H.q_rem + 4 * H.rem;
                H.q_off = H.q_med + 4 * nconst;
                H.q_rec = H.q_off + 2 * H.nc;
                if (H.q_rec > len) rc = CZ_BAD_HEADER;
            }
        }
    }
    H.rc = rc;
    *h = H;
    *status = rc;
}

// Bitmap bits [b0, b0 + cnt) of the stream (cnt <= 64), popcount.
__device__ inline uint32_t cz_ones(const uint8_t *s, const CzHeader &H, uint64_t b0, int cnt)
{
    uint32_t c = 0;
    for (int k = 0; k < cnt;) {
        const uint64_t b = b0 + k;
        const int bit = (int)(b & 7);
        const int take = 8 - bit < cnt - k ? 8 - bit : cnt - k;
        const uint32_t v = s[H.q_bits + (b >> 3)];
        const uint32_t m = ((0xFFu >> bit) & (0xFFu << (8 - bit - take))) & 0xFFu;
        c += __builtin_popcount(v & m);
        k += take;
    }
    return c;
}

// Scan item i of mode 0 (ones of bitmap word i) or mode 1 (length of record
// i: its i16 count plus the fixed part, cuszx_entry.cu:163).
__device__ inline uint32_t cz_item(const CzParams &p, const CzHeader &H, int mode, uint64_t i)
{
    if (mode == 0) {
        const uint64_t b0 = 64 * i;
        const uint64_t left = H.nb - b0;
        return cz_ones(p.s, H, b0, left < 64 ? (int)left : 64);
    }
    const uint8_t *c = p.s + H.q_off + 2 * i;
    const int leng = (int)(int16_t)((uint16_t)c[0] | (uint16_t)c[1] << 8) + H.msize;
    if (leng < 0 || leng > 4 * H.bs) {   // cuszx_entry.cu:165-169 stops the program
        atomicMin(p.status, CZ_BAD_RECORD);
        return 0;
    }
    return (uint32_t)leng;
}

__device__ inline uint64_t cz_count(const CzHeader &H, int mode) { return mode == 0 ? H.nw : H.nc; }

// Workgroup-wide (256 lanes) exclusive scan of one u64; total in t.
__device__ inline uint64_t cz_block_scan(uint64_t a, uint64_t &t)
{
    __shared__ uint64_t ws[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t ia = a;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t o = __shfl_up(ia, d, 64);
        if (lane >= d) ia += o;
    }
    if (lane == 63) ws[wv] = ia;
    __syncthreads();
    uint64_t pre = 0;
    t = 0;
    for (int k = 0; k < 4; k++) {
        if (k < wv) pre += ws[k];
        t += ws[k];
    }
    __syncthreads();
    return pre + ia - a;
}

__device__ inline uint64_t cz_lane_sum(const CzParams &p, const CzHeader &H, int mode, uint64_t i0)
{
    const uint64_t cnt = cz_count(H, mode);
    uint64_t a = 0;
    for (int k = 0; k < CZ_PER; k++)
        if (i0 + k < cnt) a += cz_item(p, H, mode, i0 + k);
    return a;
}

// grid (nparts, 2): chunk sums of both scans.
__global__ __launch_bounds__(256) void k_cz_part(CzParams p)
{
    const CzHeader H = *p.h;
    if (H.rc != 0) return;
    const int mode = blockIdx.y;
    const uint64_t i0 = blockIdx.x * (uint64_t)CZ_CH + threadIdx.x * (uint64_t)CZ_PER;
    uint64_t t;
    cz_block_scan(cz_lane_sum(p, H, mode, i0), t);
    if (threadIdx.x == 0) p.parts[mode * p.nparts + blockIdx.x] = t;
}

// grid (1, 2): exclusive scan of the chunk sums, in place.
__global__ __launch_bounds__(256) void k_cz_top(CzParams p)
{
    if (p.h->rc != 0) return;
    uint64_t *q = p.parts + blockIdx.y * p.nparts;
    uint64_t carry = 0;
    for (uint64_t base = 0; base < p.nparts; base += 256) {
        const uint64_t i = base + threadIdx.x;
        uint64_t t;
        const uint64_t e = cz_block_scan(i < p.nparts ? q[i] : 0, t);
        if (i < p.nparts) q[i] = carry + e;
        carry += t;
    }
}

// grid (nparts, 2): per-item exclusive offsets, the closing entry, and the
// checks on the totals.
__global__ __launch_bounds__(256) void k_cz_apply(CzParams p)
{
    const CzHeader H = *p.h;
    if (H.rc != 0) return;
    const int mode = blockIdx.y;
    const uint64_t cnt = cz_count(H, mode);
    const uint64_t i0 = blockIdx.x * (uint64_t)CZ_CH + threadIdx.x * (uint64_t)CZ_PER;
    uint64_t t;
    uint64_t a = cz_block_scan(cz_lane_sum(p, H, mode, i0), t) + p.parts[mode * p.nparts + blockIdx.x];
    uint64_t *dst = mode == 0 ? p.wbase : p.roff;
    for (int k = 0; k < CZ_PER; k++) {
        const uint64_t i = i0 + k;
        if (i >= cnt) break;
        dst[i] = a;
        a += cz_item(p, H, mode, i);
    }
    // the lane holding the last item (or lane 0 of chunk 0 when there is
    // none) closes the array
    const bool closer = cnt == 0 ? (blockIdx.x == 0 && threadIdx.x == 0) : (i0 < cnt && cnt <= i0 + CZ_PER);
    if (!closer) return;
    dst[cnt] = a;
    if (mode == 0 && a != H.nc) atomicMin(p.status, CZ_BAD_BITMAP);
    if (mode == 1 && a > p.len - H.q_rec) atomicMin(p.status, CZ_TRUNCATED);
}

// 4 record bytes at q.. of a block slot (record of leng bytes at LDS byte r0,
// zeros after it).
__device__ inline uint32_t cz_slot_word(const uint32_t *stage, uint32_t r0, int q, int leng)
{
    if (q >= leng) return 0u;
    const uint32_t v = lds_window4(stage, r0 + (uint32_t)q);
    return leng - q >= 4 ? v : v & ((1u << (8 * (leng - q))) - 1u);
}

// Bitmap word wi (blocks 64 wi .. 64 wi + 63) with block 64 wi + k at bit
// 63 - k; bytes past the bitmap read as zero.
__device__ inline uint64_t cz_word(const uint8_t *s, const CzHeader &H, uint64_t wi)
{
    const uint64_t nbit = (H.nb + 7) / 8;
    uint64_t w = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const uint64_t i = 8 * wi + t;
        w = (w << 8) | (i < nbit ? s[H.q_bits + i] : 0u);
    }
    return w;
}

// Set bits among the first k (0..64) of a bitmap word.
__device__ inline uint32_t cz_first(uint64_t w, int k) { return k ? __builtin_popcountll(w >> (64 - k)) : 0u; }

// Decode units of 256 / S consecutive blocks.  256 / S divides 64, so a unit
// lies inside one bitmap word: its blocks' states and ranks come from that
// word and the word's scanned base.
template <int S>
__device__ inline void cz_decode_units(const CzParams &p, const CzHeader &H, uint32_t *stage, int aligned)
{
    constexpr int BPW = 256 / S;
    const int bs = H.bs, nl = bs / 4;
    const uint64_t nb = H.nb, units = (nb + BPW - 1) / BPW;
    const int seg = threadIdx.x / S, l = threadIdx.x & (S - 1);
    const bool act = l < nl;
    const uint8_t *lds = reinterpret_cast<const uint8_t *>(stage);
    for (uint64_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint64_t b0 = u * BPW, bend = b0 + BPW < nb ? b0 + BPW : nb;
        const uint64_t wi = b0 >> 6;
        const uint64_t word = cz_word(p.s, H, wi), base = p.wbase[wi];
        const int k0 = (int)(b0 & 63);
        const uint64_t j0 = base + cz_first(word, k0), j1 = base + cz_first(word, k0 + (int)(bend - b0));
        const uint64_t P0 = p.roff[j0], P1 = p.roff[j1];
        // stage record bytes [P0, P1): aligned words inside the stream, a
        // partial last word byte by byte
        const uint8_t *src = p.s + H.q_rec + P0;
        const uint8_t *end = p.s + p.len;
        const uint32_t head = (uint32_t)((uintptr_t)src & 3);
        const uint32_t *sw = reinterpret_cast<const uint32_t *>(src - head);
        const uint32_t nw = (head + (uint32_t)(P1 - P0) + 3) / 4;
        __syncthreads();   // the previous unit's reads of stage are done
        for (uint32_t i = threadIdx.x; i < nw; i += 256) {
            const uint8_t *a = reinterpret_cast<const uint8_t *>(sw + i);
            if (a + 4 <= end) {
                stage[i] = sw[i];
            } else {
                uint32_t v = 0;
                for (int k = 0; k < 4; k++)
                    if (a + k < end) v |= (uint32_t)a[k] << (8 * k);
                stage[i] = v;
            }
        }
        __syncthreads();
        const uint64_t b = b0 + seg;
        if (b >= bend) continue;
        const int kb = k0 + seg;
        const uint64_t rank = base + cz_first(word, kb);   // non-constant blocks before b
        float *y = p.out + b * (uint64_t)bs + 4 * l;
        float v[4];
        if (!((word >> (63 - kb)) & 1)) {
            // constant block: its native median (cuszx_entry.cu:157-159, 191-196)
            const uint8_t *c = p.s + H.q_med + 4 * (b - rank);
            const float med = bits_f32((uint32_t)c[0] | (uint32_t)c[1] << 8 | (uint32_t)c[2] << 16 |
                                       (uint32_t)c[3] << 24);
            v[0] = v[1] = v[2] = v[3] = med;
        } else {
            const uint32_t r0 = head + (uint32_t)(p.roff[rank] - P0);
            const int leng = (int)(p.roff[rank + 1] - p.roff[rank]);
            // cuszxd_float.cu:185-197
            const float med = bits_f32(cz_slot_word(stage, r0, 0, leng));
            const int rl = 4 < leng ? lds[r0 + 4] : 0;
            const int rb = rl % 8 ? rl / 8 + 1 : rl / 8;
            const int shift = rl % 8 ? 8 - rl % 8 : 0;
            const bool bytes_ok = rb >= 1 && rb <= 4;
            const uint32_t code = act && 5 + l < leng ? lds[r0 + 5 + l] : 0u;
            // word bytes at or above 4 - rb (rb >= 1)
            const uint32_t sig = rb >= 4 ? ~0u : rb >= 1 ? ~0u << (8 * (4 - rb)) : 0u;
            int lk[4];
            int lsum = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                lk[k] = (int)(code >> (6 - 2 * k)) & 3;
                if (act) lsum += rb - lk[k];
            }
            // prefix sums of the mid-byte counts (cuszxd_float.cu:197-198)
            const int at0 = (int)seg_incl_sum<S>((uint32_t)lsum, l) - lsum;
            int at = at0;
            uint32_t wv[4], wr[4];
            uint32_t mine = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int m = rb - lk[k];
                at += m;
                uint32_t cur = 0;
                wr[k] = 0;
                if (act && bytes_ok && m >= 1) {
                    // the value's own bytes 4-rb .. 4-rb+m-1, in that order,
                    // ending at slot byte msize + at (cuszxd_float.cu:206-267)
                    const int p0 = H.msize + at - m;
                    if (p0 < 0 || p0 + m > 4 * bs) {
                        atomicMin(p.status, CZ_BAD_READ);
                    } else {
                        const uint32_t raw = cz_slot_word(stage, r0, p0, leng);
                        cur = (m == 4 ? raw : raw & ((1u << (8 * m)) - 1u)) << (8 * (4 - rb));
                    }
                    if (rb >= 2) {
                        // carried bytes: the value's own among word bytes 1..3
                        const uint32_t pm = sig & (~0u >> (8 * lk[k])) & 0xFFFFFF00u;
                        const uint32_t fl = __builtin_bswap32(pm);
                        wr[k] = (__builtin_bswap32(cur & pm) & 0xFFFFFFu) |
                                (((fl & 1u) | ((fl >> 7) & 2u) | ((fl >> 14) & 4u)) << 24);
                    }
                    mine = carry_combine(mine, wr[k]);
                }
                wv[k] = cur;
            }
            uint32_t acc = mine;
#pragma unroll
            for (int d = 1; d < S; d <<= 1) {
                const uint32_t o = __shfl_up(acc, d, S);
                if (l >= d) acc = carry_combine(o, acc);
            }
            uint32_t cin = __shfl_up(acc, 1, S);
            if (l == 0) cin = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                // inherited word bytes: the top `lead`, from the last value
                // that carried them; none with one significant byte
                const uint32_t inh = lk[k] ? ~0u << (8 * (4 - lk[k])) : 0u;
                const uint32_t from = rb >= 2 ? __builtin_bswap32(cin) & 0xFFFFFF00u : 0u;
                const uint32_t w = (wv[k] & ~inh) | (from & inh);
                cin = carry_combine(cin, wr[k]);
                v[k] = bits_f32(w << shift) + med;   // cuszxd_float.cu:290-292
            }
        }
        if (act) {
            if (aligned) {
                typedef float f32x4_t __attribute__((ext_vector_type(4)));
                const f32x4_t w4 = {v[0], v[1], v[2], v[3]};
                __builtin_nontemporal_store(w4, reinterpret_cast<f32x4_t *>(y));   // streamed past the caches
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) y[k] = v[k];
            }
        }
    }
}

// A grid-stride loop over the decode units, so the grid does not depend on
// the block size in the stream; the segment width is chosen here.
__global__ __launch_bounds__(256) void k_cz_decode(CzParams p, int aligned)
{
    __shared__ uint32_t stage[CZ_SPAN / 4 + 4];
    const CzHeader H = *p.h;
    if (H.rc != 0 || *p.status != 0) return;
    const int nl = H.bs / 4;
    if (nl <= 8) cz_decode_units<8>(p, H, stage, aligned);
    else if (nl <= 16) cz_decode_units<16>(p, H, stage, aligned);
    else if (nl <= 32) cz_decode_units<32>(p, H, stage, aligned);
    else cz_decode_units<64>(p, H, stage, aligned);
}

// The n % bs trailing values, stored raw (cuszx_entry.cu:155).
__global__ __launch_bounds__(256) void k_cz_tail(CzParams p)
{
    const CzHeader H = *p.h;
    if (H.rc != 0 || *p.status != 0) return;
    for (uint64_t r = threadIdx.x; r < H.rem; r += 256) {
        const uint8_t *c = p.s + H.q_rem + 4 * r;
        p.out[H.nb * (uint64_t)H.bs + r] =
            bits_f32((uint32_t)c[0] | (uint32_t)c[1] << 8 | (uint32_t)c[2] << 16 | (uint32_t)c[3] << 24);
    }
}

// ---- writer ----
//
// The stream cuSZx_fast_compress_args_unpredictable_blocked_float writes
// (cuszx_float.cu:222-298 per block, assembled by cuszx_entry.cu:20-70), in
// two passes over the input: k_czw_pass<false> sizes every record (0 for a
// constant block), the generic codec's scan turns the sizes into record
// offsets and constant ranks and writes the bitmap, and k_czw_pass<true>
// writes every median, count and record at its place; k_czw_finish writes
// the header and the trailing values.

struct CzwParams {
    const float *x;
    uint64_t n, nb, rem;
    int bs, msize, S;
    float eb;
    uint8_t *out;
    uint32_t *size;          // per block: record bytes, 0 = constant
    const uint64_t *idx;     // per block: {record offset, constant blocks before}
    const uint64_t *totals;  // {record bytes, constant blocks}
    uint64_t *result;        // {0, stream bytes, bound bits, 1 = a block needs > 32 bits}
};

// One block on a segment of S lanes, 4 values a lane: the statistics
// (cuszx_float.cu:238-278), reqLength (:53-60), the shifted words (:138),
// lead counts against the previous word (:142-167: the equal leading bytes,
// at most the significant ones, so an equal word at 4 bytes counts 4) and the
// segment's mid-byte prefix sums (:170-171).
struct CzwBlock {
    float med;
    bool cst, bad;
    int rl, rb;
    uint32_t w[4];
    int ld[4];
    uint32_t lsum, incl, total;
};

template <int S>
__device__ inline void czw_block(const CzwParams &p, uint64_t b, int l, CzwBlock &k)
{
    const bool act = l < p.bs / 4;
    const float *xb = p.x + b * (uint64_t)p.bs + 4 * l;
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = act ? xb[i] : 0.0f;
    float mn = act ? fminf(fminf(v[0], v[1]), fminf(v[2], v[3])) : __builtin_huge_valf();
    float mx = act ? fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3])) : -__builtin_huge_valf();
#pragma unroll
    for (int d = 1; d < S; d <<= 1) {
        mn = fminf(mn, __shfl_xor(mn, d, S));
        mx = fmaxf(mx, __shfl_xor(mx, d, S));
    }
    const float radius = (mx - mn) / 2.0f;
    k.med = mn + radius;
    k.cst = radius <= p.eb;
    k.bad = false;
    k.lsum = k.incl = k.total = 0;
    if (k.cst) return;
    const int radExpo = (int)((__builtin_bit_cast(uint32_t, radius) & 0x7F800000u) >> 23) - 127;
    const int reqExpo = (int)((__builtin_bit_cast(uint32_t, p.eb) & 0x7F800000u) >> 23) - 127;
    k.rl = 9 + radExpo - reqExpo + 1;
    k.bad = k.rl > 32;
    if (k.bad) return;
    k.rb = k.rl % 8 ? k.rl / 8 + 1 : k.rl / 8;
    const int shift = k.rl % 8 ? 8 - k.rl % 8 : 0;
#pragma unroll
    for (int i = 0; i < 4; i++) k.w[i] = __builtin_bit_cast(uint32_t, v[i] - k.med) >> shift;
    uint32_t prev = __shfl_up(k.w[3], 1, S);
    if (l == 0) prev = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t x = k.w[i] ^ prev;
        const int lz = x ? __builtin_clz(x) >> 3 : 4;
        k.ld[i] = lz < k.rb ? lz : k.rb;
        if (act) k.lsum += (uint32_t)(k.rb - k.ld[i]);
        prev = k.w[i];
    }
    k.incl = seg_incl_sum<S>(k.lsum, l);
    k.total = __shfl(k.incl, S - 1, S);
}

template <int S, bool WRITE>
__global__ __launch_bounds__(256) void k_czw_pass(CzwParams p)
{
    constexpr int BPW = 256 / S;
    const int l = threadIdx.x & (S - 1);
    const uint64_t b = blockIdx.x * (uint64_t)BPW + threadIdx.x / S;
    if (b >= p.nb) return;   // whole segments
    CzwBlock k;
    czw_block<S>(p, b, l, k);
    if (!WRITE) {
        if (l == 0) {
            p.size[b] = k.cst ? 0u : (uint32_t)p.msize + k.total;
            if (k.bad) p.result[3] = 1;
        }
        return;
    }
    if (k.bad) return;
    const uint64_t nconst = p.totals[1];
    const uint64_t q_med = 13 + (p.nb + 7) / 8 + 4 * p.rem;
    const uint64_t q_off = q_med + 4 * nconst, q_rec = q_off + 2 * (p.nb - nconst);
    const uint64_t roff = p.idx[2 * b], cbefore = p.idx[2 * b + 1];
    const uint32_t mb = __builtin_bit_cast(uint32_t, k.med);
    if (k.cst) {
        // the constant block's native median, by rank (cuszx_entry.cu:56-58)
        if (l < 4) p.out[q_med + 4 * cbefore + l] = (uint8_t)(mb >> (8 * l));
        return;
    }
    const uint64_t j = b - cbefore;
    if (l < 2) p.out[q_off + 2 * j + l] = (uint8_t)(k.total >> (8 * l));   // native i16 (cuszx_entry.cu:60)
    uint8_t *r = p.out + q_rec + roff;
    if (l < 4) r[l] = (uint8_t)(mb >> (8 * l));
    if (l == 4) r[4] = (uint8_t)k.rl;
    if (l < p.bs / 4) {
        // the four lead counts OR-ed into one byte, a count of 4 spilling into
        // the neighbour's field (cuszx_float.cu:38-51)
        uint32_t code = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) code |= (uint32_t)k.ld[i] << (2 * (3 - i));
        r[5 + l] = (uint8_t)code;
        // each value's bytes 4-rb .. 4-rb+m-1 in that order (cuszx_float.cu:173-214)
        uint32_t pos = (uint32_t)p.msize + k.incl - k.lsum;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = k.rb - k.ld[i];
            for (int q = 0; q < m; q++) r[pos++] = (uint8_t)(k.w[i] >> (8 * (4 - k.rb + q)));
        }
    }
}

__global__ void k_czw_init(uint64_t *result)
{
    if (threadIdx.x == 0) result[3] = 0;
}

// Header, stream size and the trailing values (cuszx_entry.cu:39-50).
__global__ __launch_bounds__(256) void k_czw_finish(CzwParams p)
{
    const uint64_t rec = p.nb ? p.totals[0] : 0, nconst = p.nb ? p.totals[1] : 0;
    const uint64_t q_rem = 13 + (p.nb + 7) / 8;
    const uint64_t q_rec = q_rem + 4 * p.rem + 4 * nconst + 2 * (p.nb - nconst);
    if (threadIdx.x < 13) {
        const int i = threadIdx.x;
        const uint8_t h[5] = {1, 1, 1, 0, (uint8_t)p.bs};
        p.out[i] = i < 5 ? h[i] : (uint8_t)(nconst >> (8 * (12 - i)));
    }
    if (threadIdx.x == 0) {
        p.result[0] = 0;
        p.result[1] = q_rec + rec;
        p.result[2] = __builtin_bit_cast(uint32_t, p.eb);
    }
    const uint8_t *tail = reinterpret_cast<const uint8_t *>(p.x + p.nb * (uint64_t)p.bs);
    for (uint64_t i = threadIdx.x; i < 4 * p.rem; i += 256) p.out[q_rem + i] = tail[i];
}

// ---- host side ----

// Workspace for n values: sized for the smallest block size (32).
struct CzWsLayout {
    uint64_t nbmax, nwmax, nparts, o_wbase, o_roff, o_parts, bytes;
};

static CzWsLayout cz_ws_layout(uint64_t n)
{
    auto align_up = [](uint64_t v, uint64_t a) { return (v + a - 1) / a * a; };
    CzWsLayout L;
    L.nbmax = n / 32;
    L.nwmax = (L.nbmax + 63) / 64;
    const uint64_t items = L.nbmax > L.nwmax ? L.nbmax : L.nwmax;
    L.nparts = ceil_div(items ? items : 1, (uint64_t)CZ_CH);
    L.o_wbase = 256;
    L.o_roff = L.o_wbase + align_up(8 * (L.nwmax + 1), 256);
    L.o_parts = L.o_roff + align_up(8 * (L.nbmax + 1), 256);
    L.bytes = L.o_parts + align_up(16 * L.nparts, 256);
    return L;
}

uint64_t cuszx_workspace(uint64_t n) { return cz_ws_layout(n).bytes; }

void cuszx_decompress_launch(const uint8_t *d_stream, uint64_t len, uint64_t n, float *d_out, uint8_t *d_ws,
                             int32_t *d_status, hipStream_t st)
{
    const CzWsLayout L = cz_ws_layout(n);
    CzParams p;
    p.s = d_stream;
    p.len = len;
    p.n = n;
    p.h = reinterpret_cast<CzHeader *>(d_ws);
    p.wbase = reinterpret_cast<uint64_t *>(d_ws + L.o_wbase);
    p.roff = reinterpret_cast<uint64_t *>(d_ws + L.o_roff);
    p.parts = reinterpret_cast<uint64_t *>(d_ws + L.o_parts);
    p.nparts = L.nparts;
    p.out = d_out;
    p.status = d_status;
    hipLaunchKernelGGL(k_cz_header, dim3(1), dim3(64), 0, st, d_stream, len, n, p.h, d_status);
    timing_mark_begin(KK_DECODE, st);
    hipLaunchKernelGGL(k_cz_part, dim3((unsigned)L.nparts, 2), dim3(256), 0, st, p);
    hipLaunchKernelGGL(k_cz_top, dim3(1, 2), dim3(256), 0, st, p);
    hipLaunchKernelGGL(k_cz_apply, dim3((unsigned)L.nparts, 2), dim3(256), 0, st, p);
    // units of >= 640 values (256 / S blocks of bs <= 4 S values, bs >= 2.5 S)
    const uint64_t units = ceil_div(n ? n : 1, (uint64_t)640);
    const unsigned grid = (unsigned)(units < 65536 ? units : 65536);
    const int aligned = ((uintptr_t)d_out & 15) == 0;
    hipLaunchKernelGGL(k_cz_decode, dim3(grid), dim3(256), 0, st, p, aligned);
    hipLaunchKernelGGL(k_cz_tail, dim3(1), dim3(256), 0, st, p);
    timing_mark_end(KK_DECODE, st);
}

// ---- writer, host side ----

static int czw_lanes(int bs)
{
    int S = 8;
    while (S < bs / 4) S <<= 1;
    return S;
}

template <int S>
static void czw_pass_s(const CzwParams &p, bool write, unsigned grid, hipStream_t st)
{
    if (write) hipLaunchKernelGGL((k_czw_pass<S, true>), dim3(grid), dim3(256), 0, st, p);
    else hipLaunchKernelGGL((k_czw_pass<S, false>), dim3(grid), dim3(256), 0, st, p);
}

static void czw_pass(const CzwParams &p, bool write, unsigned grid, hipStream_t st)
{
    switch (p.S) {
    case 8: czw_pass_s<8>(p, write, grid, st); break;
    case 16: czw_pass_s<16>(p, write, grid, st); break;
    case 32: czw_pass_s<32>(p, write, grid, st); break;
    default: czw_pass_s<64>(p, write, grid, st); break;
    }
}

uint64_t cuszx_capacity(uint64_t n, int bs)
{
    const uint64_t nb = n / (uint64_t)bs;
    return 13 + (nb + 7) / 8 + 4 * (n % (uint64_t)bs) + nb * (2 + 5 + (uint64_t)bs / 4 + 4 * (uint64_t)bs);
}

struct CzwLayout {
    uint64_t nb, nparts, o_size, o_idx, o_parts, o_tot, bytes;
};

static CzwLayout czw_layout(uint64_t n, int bs)
{
    auto al = [](uint64_t v) { return (v + 255) & ~255ull; };
    CzwLayout L;
    L.nb = n / (uint64_t)bs;
    L.nparts = generic_parts(L.nb);
    L.o_size = 0;
    L.o_idx = al(4 * (L.nb + 1));
    L.o_parts = L.o_idx + al(16 * (L.nb + 1));
    L.o_tot = L.o_parts + al(16 * L.nparts);
    L.bytes = L.o_tot + 256;
    return L;
}

uint64_t cuszx_compress_workspace(uint64_t n, int bs) { return czw_layout(n, bs).bytes; }

void cuszx_compress_launch(const float *d_in, uint64_t n, float eb, int bs, uint8_t *d_out, uint8_t *d_ws,
                           uint64_t *d_result, hipStream_t st)
{
    const CzwLayout L = czw_layout(n, bs);
    CzwParams p;
    p.x = d_in;
    p.n = n;
    p.nb = L.nb;
    p.rem = n % (uint64_t)bs;
    p.bs = bs;
    p.msize = 5 + bs / 4;
    p.S = czw_lanes(bs);
    p.eb = eb;
    p.out = d_out;
    p.size = reinterpret_cast<uint32_t *>(d_ws + L.o_size);
    p.idx = reinterpret_cast<const uint64_t *>(d_ws + L.o_idx);
    p.totals = reinterpret_cast<const uint64_t *>(d_ws + L.o_tot);
    p.result = d_result;
    hipLaunchKernelGGL(k_czw_init, dim3(1), dim3(64), 0, st, d_result);
    timing_mark_begin(KK_COMPRESS, st);
    if (L.nb) {
        const unsigned grid = (unsigned)ceil_div(L.nb, (uint64_t)(256 / p.S));
        czw_pass(p, false, grid, st);
        GenericCompress g{};
        g.nb = L.nb;
        g.bs = bs;
        g.size = p.size;
        g.idx = reinterpret_cast<uint64_t *>(d_ws + L.o_idx);
        g.parts = reinterpret_cast<uint64_t *>(d_ws + L.o_parts);
        g.nparts = L.nparts;
        g.totals = reinterpret_cast<uint64_t *>(d_ws + L.o_tot);
        g.bitmap = d_out + 13;
        generic_scan_launch(g, st);
        czw_pass(p, true, grid, st);
    }
    hipLaunchKernelGGL(k_czw_finish, dim3(1), dim3(256), 0, st, p);
    timing_mark_end(KK_COMPRESS, st);
}

}  // namespace szxamd
