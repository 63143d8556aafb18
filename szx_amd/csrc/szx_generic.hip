// szx_generic.hip -- the blocked codec for block sizes the tiled kernels do
// not take.
//
// The tiled kernels (szx_compress.hip, szx_decompress.hip) map a block onto an
// aligned group of lanes and need a power-of-two block size of at least 4.
// The reference's blocked API takes any block size that fits its one-byte
// header field (szx_float.cc:1095), so every other size 1..255 runs here.
// A block is a segment of S lanes (S = the smallest power of two with
// 4 S >= bs), each lane holding 4 consecutive values; segment shuffles give
// the block statistics, the byte offsets and the cross-lane carries of the
// reference's per-block loops (szx_float.cc:314-371 and 49-252 for
// compression, szxd_float.cc:34-244 for decompression; fp64 in szx_double.cc
// and szxd_double.cc).  The streams are the same bytes.
//
// Compression (two passes over the input, no per-block staging):
//   k_gen_encode<T, false>  per block: statistics and record size (0 = constant)
//   k_gen_part / k_gen_top / k_gen_apply
//                           exclusive scan of (record bytes, constant blocks)
//                           -> per-block pairs {record offset in the payload,
//                           constant blocks before}, a closing pair {payload
//                           bytes, constant blocks} and the totals
//   k_gen_encode<T, true>   per block again: the record at its offset, or the
//                           constant median at its rank
//   k_gen_bitmap            one lane per bitmap byte
// The per-block pairs are the record index for these block sizes; the
// decoder's index builder (szx_index.hip) writes the same pairs.
//
// Decompression: k_gen_decode, one segment per block, checks every record
// against its pair and the next one.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szx_common.hpp"
#include "szx_internal.hpp"

namespace szxamd {

// Lead bytes (capped at 3) a shifted word shares with the previous one
// (szx_float.cc:96-101, szx_double.cc:80-85).
template <typename U> __device__ inline int gen_lead(U xr)
{
    constexpr int W = sizeof(U);
    if ((xr >> (8 * (W - 3))) == 0) return 3;
    if ((xr >> (8 * (W - 2))) == 0) return 2;
    if ((xr >> (8 * (W - 1))) == 0) return 1;
    return 0;
}

__device__ inline bool gen_nonconst(const uint8_t *bitmap, uint64_t b)
{
    return (bitmap[b >> 3] >> (7 - (b & 7))) & 1;
}

template <typename T> __device__ inline T gmin(T a, T b);
template <typename T> __device__ inline T gmax(T a, T b);
template <> __device__ inline float gmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <> __device__ inline float gmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }
template <> __device__ inline double gmin<double>(double a, double b) { return __builtin_fmin(a, b); }
template <> __device__ inline double gmax<double>(double a, double b) { return __builtin_fmax(a, b); }

// Payload bytes the blocks of one workgroup (256 / S of them) can span:
// records of at most 5 + S + 4 S W bytes (bs <= 4 S), plus a 4-byte alignment
// head; LDS words with 3 words of slack for the decoder's 8-byte windows.
template <int S, int W> struct Span {
    static constexpr int BYTES = (256 / S) * (5 + S + 4 * S * W) + 4;
    static constexpr int WORDS = (BYTES + 3) / 4 + 3;
};

// Values of one workgroup's blocks: at most 256 / S * bs <= 1024, staged
// through LDS in aligned 16-byte chunks (the value array is 16-byte aligned),
// plus one chunk of alignment slack at each end.
template <typename T> struct Vals {
    static constexpr int CE = 16 / (int)sizeof(T);   // values per chunk
    static constexpr int N = 1024 + 2 * CE;
};

// Load values [e0, e1) of x (n values) into xs, chunk-aligned: xs[i] holds
// x[c0 * CE + i], c0 = e0 / CE.  Chunks past the array end go value by value.
template <typename T>
__device__ inline void stage_values(const T *x, uint64_t n, uint64_t e0, uint64_t e1, T *xs)
{
    constexpr int CE = Vals<T>::CE;
    const uint64_t c0 = e0 / CE, c1 = (e1 + CE - 1) / CE;
    for (uint64_t c = c0 + threadIdx.x; c < c1; c += 256) {
        T *d = xs + (c - c0) * CE;
        if ((c + 1) * CE <= n) {
            *reinterpret_cast<uint4 *>(d) = *reinterpret_cast<const uint4 *>(x + c * CE);
        } else {
            for (int j = 0; j < CE; j++)
                if (c * CE + j < n) d[j] = x[c * CE + j];
        }
    }
}

// Store values [e0, e1) from xs (same layout) to x: chunks wholly inside the
// range as 16-byte stores, the edge chunks (shared with the neighbouring
// workgroups) value by value.
template <typename T>
__device__ inline void unstage_values(T *x, uint64_t e0, uint64_t e1, const T *xs)
{
    constexpr int CE = Vals<T>::CE;
    const uint64_t c0 = e0 / CE, c1 = (e1 + CE - 1) / CE;
    for (uint64_t c = c0 + threadIdx.x; c < c1; c += 256) {
        const T *s = xs + (c - c0) * CE;
        if (c * CE >= e0 && (c + 1) * CE <= e1) {
            *reinterpret_cast<uint4 *>(x + c * CE) = *reinterpret_cast<const uint4 *>(s);
        } else {
            for (int j = 0; j < CE; j++) {
                const uint64_t e = c * CE + j;
                if (e >= e0 && e < e1) x[e] = s[j];
            }
        }
    }
}

// One block's encode state, per lane of its segment.
template <typename T> struct GenBlock {
    using U = typename Word<T>::U;
    bool constant;
    int m, len, rb;
    float med_f, hmed;
    bool val[4];
    U sw[4];
    int ld[4];
    uint32_t lsum, incl, total;
};

// One block per segment of S lanes, 4 consecutive values per lane (S = the
// smallest power of two with 4 S >= bs): statistics, reqLength, shifted
// words, lead codes and the lane's byte offset in the block's mid bytes.
// x = the block's values (staged in LDS).
template <typename T, int S>
__device__ inline void gen_block(const GenericCompress &g, uint64_t b, int l, const T *x, GenBlock<T> &k_)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    const float eb = g.eb_dev ? *g.eb_dev : g.eb;
    const uint64_t rest = g.n - b * (uint64_t)g.bs;
    const int m = rest < (uint64_t)g.bs ? (int)rest : g.bs;
    k_.m = m;
    T v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        k_.val[k] = 4 * l + k < m;
        v[k] = k_.val[k] ? x[4 * l + k] : (T)0;
    }
    // block statistics (szx_float.cc:327-333, szx_double.cc:985-991): NaNs
    // never win a comparison, so fmin / fmax over the values, except that a
    // NaN first value makes both NaN.  Which of two zeros is kept cannot change
    // the radius or the median.
    T lo = v[0], hi = v[0];
#pragma unroll
    for (int k = 1; k < 4; k++)
        if (k_.val[k]) {
            lo = gmin(lo, v[k]);
            hi = gmax(hi, v[k]);
        }
    if (4 * l >= m) {   // lanes past the block's end
        lo = __builtin_huge_valf();
        hi = -__builtin_huge_valf();
    }
#pragma unroll
    for (int d = 1; d < S; d <<= 1) {
        lo = gmin(lo, __shfl_xor(lo, d, S));
        hi = gmax(hi, __shfl_xor(hi, d, S));
    }
    const T first = __shfl(v[0], 0, S);
    if (first != first) {
        lo = first;
        hi = first;
    }
    const T rad = (hi - lo) / (T)2;
    const T med = lo + rad;
    k_.med_f = (float)med;
    const float rad_f = (float)rad;
    k_.constant = W == 4 ? (rad <= (T)eb) : ((double)rad <= (double)eb);
    if (k_.constant) return;   // segment-uniform
    int clamped;
    k_.len = req_length<T>(rad_expo(rad_f), req_expo(eb), clamped);
    k_.rb = (k_.len + 7) >> 3;
    const int shift = (-k_.len) & 7;
    k_.hmed = clamped ? 0.0f : k_.med_f;
    // fp32 encodes against the stored median; fp64 against the median before
    // the clamp (szx_float.cc:60,90; szx_double.cc:38,78)
    const T msub = W == 4 ? (T)k_.hmed : (T)k_.med_f;
#pragma unroll
    for (int k = 0; k < 4; k++) k_.sw[k] = to_bits<T>(v[k] - msub) >> shift;
    // the previous lane's last word (every lane before the last valid one is full)
    U prev = __shfl_up(k_.sw[3], 1, S);
    if (l == 0) prev = 0;
    k_.lsum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        k_.ld[k] = gen_lead<U>(k_.sw[k] ^ prev);
        prev = k_.sw[k];
        if (k_.val[k] && k_.ld[k] < k_.rb) k_.lsum += k_.rb - k_.ld[k];
    }
    k_.incl = seg_incl_sum<S>(k_.lsum, l);
    k_.total = __shfl(k_.incl, S - 1, S);
}

// WRITE = false: record size per block (0 = constant).  WRITE = true: the
// workgroup's records are built in LDS (they are contiguous in the payload)
// and written out with aligned word stores; constant medians go to their rank.
template <typename T, int S, bool WRITE>
__global__ __launch_bounds__(256) void k_gen_encode(GenericCompress g)
{
    constexpr int W = Word<T>::W;
    constexpr int BPW = 256 / S;
    __shared__ __align__(16) T xs[Vals<T>::N];
    const int l = threadIdx.x & (S - 1);
    const uint64_t b0 = blockIdx.x * (uint64_t)BPW;
    const uint64_t b = b0 + threadIdx.x / S;
    const bool live = b < g.nb;   // whole segments
    const uint64_t bend = b0 + BPW < g.nb ? b0 + BPW : g.nb;
    const uint64_t e0 = b0 * (uint64_t)g.bs, e1 = bend * (uint64_t)g.bs < g.n ? bend * (uint64_t)g.bs : g.n;
    stage_values<T>(static_cast<const T *>(g.x), g.n, e0, e1, xs);
    __syncthreads();
    GenBlock<T> k_;
    if (live) gen_block<T, S>(g, b, l, xs + (b * (uint64_t)g.bs - e0 / Vals<T>::CE * Vals<T>::CE), k_);
    if (!WRITE) {
        if (live && l == 0) g.size[b] = k_.constant ? 0u : 5u + (uint32_t)((k_.m + 3) >> 2) + k_.total;
        return;
    }
    __shared__ uint32_t stage[WRITE ? Span<S, W>::WORDS : 1];
    uint8_t *lds = reinterpret_cast<uint8_t *>(stage);
    const uint64_t P0 = g.idx[2 * b0], P1 = g.idx[2 * bend];
    uint8_t *dst0 = g.payload + P0;
    const uint32_t head = (uint32_t)((uintptr_t)dst0 & 3);
    if (live && k_.constant && l == 0) g.cmed[g.idx[2 * b + 1]] = bswap32(f32_bits(k_.med_f));
    if (live && !k_.constant) {
        const int nlead = (k_.m + 3) >> 2;
        uint8_t *rec = lds + head + (uint32_t)(g.idx[2 * b] - P0);
        if (l == 0) {
            const uint32_t hb = f32_bits(k_.hmed);
            rec[0] = (uint8_t)k_.len;
            rec[1] = (uint8_t)(hb >> 24);
            rec[2] = (uint8_t)(hb >> 16);
            rec[3] = (uint8_t)(hb >> 8);
            rec[4] = (uint8_t)hb;
        }
        if (4 * l < k_.m) {
            uint32_t code = 0;
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (k_.val[k]) code |= (uint32_t)k_.ld[k] << (6 - 2 * k);
            rec[5 + l] = (uint8_t)code;   // szx_TypeManager.cc:198-273
        }
        uint8_t *q = rec + 5 + nlead + (k_.incl - k_.lsum);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k_.val[k])
                for (int t = W - k_.rb; t < W - k_.ld[k]; t++) *q++ = (uint8_t)(k_.sw[k] >> (8 * t));
    }
    __syncthreads();
    // copy out: words wholly inside [P0, P1) as words, the two edge words
    // (shared with the neighbouring workgroups' spans) byte by byte
    const uint32_t span = (uint32_t)(P1 - P0);
    const uint32_t nw = (head + span + 3) / 4;
    uint32_t *dw = reinterpret_cast<uint32_t *>(dst0 - head);
    for (uint32_t i = threadIdx.x; i < nw; i += 256) {
        const uint32_t lo = 4 * i, hi = lo + 4;
        if (lo >= head && hi <= head + span) {
            dw[i] = stage[i];
        } else {
            uint8_t *db = reinterpret_cast<uint8_t *>(dw);
            for (uint32_t j = lo < head ? head : lo; j < hi && j < head + span; j++) db[j] = lds[j];
        }
    }
}

// ---- exclusive scan of (record bytes, constant blocks) over the blocks ----

constexpr int GEN_PER = 16;                 // blocks per lane
constexpr int GEN_CH = 256 * GEN_PER;       // blocks per scan chunk

// Workgroup-wide (256 lanes) exclusive scan of two u64 sums; returns the
// workgroup totals in ta, tc.
__device__ inline void gen_block_scan(uint64_t &a, uint64_t &c, uint64_t &ta, uint64_t &tc)
{
    __shared__ uint64_t wa[4], wc[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t ia = a, ic = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t oa = __shfl_up(ia, d, 64), oc = __shfl_up(ic, d, 64);
        if (lane >= d) {
            ia += oa;
            ic += oc;
        }
    }
    if (lane == 63) {
        wa[wv] = ia;
        wc[wv] = ic;
    }
    __syncthreads();
    uint64_t pa = 0, pc = 0;
    ta = 0;
    tc = 0;
    for (int k = 0; k < 4; k++) {
        if (k < wv) {
            pa += wa[k];
            pc += wc[k];
        }
        ta += wa[k];
        tc += wc[k];
    }
    a = pa + ia - a;
    c = pc + ic - c;
    __syncthreads();
}

__device__ inline void gen_sums(const uint32_t *size, uint64_t nb, uint64_t b0, uint64_t &a, uint64_t &c)
{
    a = 0;
    c = 0;
    for (int k = 0; k < GEN_PER; k++) {
        const uint64_t b = b0 + k;
        if (b < nb) {
            const uint32_t s = size[b];
            a += s;
            c += s == 0;
        }
    }
}

__global__ __launch_bounds__(256) void k_gen_part(GenericCompress g)
{
    uint64_t a, c, ta, tc;
    gen_sums(g.size, g.nb, blockIdx.x * (uint64_t)GEN_CH + threadIdx.x * (uint64_t)GEN_PER, a, c);
    gen_block_scan(a, c, ta, tc);
    if (threadIdx.x == 0) {
        g.parts[2 * blockIdx.x] = ta;
        g.parts[2 * blockIdx.x + 1] = tc;
    }
}

// One workgroup: exclusive scan of the chunk sums, in place.
__global__ __launch_bounds__(256) void k_gen_top(GenericCompress g)
{
    uint64_t ca = 0, cc = 0;
    for (uint64_t base = 0; base < g.nparts; base += 256) {
        const uint64_t i = base + threadIdx.x;
        uint64_t a = i < g.nparts ? g.parts[2 * i] : 0, c = i < g.nparts ? g.parts[2 * i + 1] : 0;
        uint64_t ta, tc;
        gen_block_scan(a, c, ta, tc);
        if (i < g.nparts) {
            g.parts[2 * i] = ca + a;
            g.parts[2 * i + 1] = cc + c;
        }
        ca += ta;
        cc += tc;
    }
}

__global__ __launch_bounds__(256) void k_gen_apply(GenericCompress g)
{
    const uint64_t b0 = blockIdx.x * (uint64_t)GEN_CH + threadIdx.x * (uint64_t)GEN_PER;
    uint64_t a, c, ta, tc;
    gen_sums(g.size, g.nb, b0, a, c);
    gen_block_scan(a, c, ta, tc);
    a += g.parts[2 * blockIdx.x];
    c += g.parts[2 * blockIdx.x + 1];
    for (int k = 0; k < GEN_PER; k++) {
        const uint64_t b = b0 + k;
        if (b >= g.nb) break;
        g.idx[2 * b] = a;
        g.idx[2 * b + 1] = c;
        const uint32_t s = g.size[b];
        a += s;
        c += s == 0;
        if (b + 1 == g.nb) {
            g.idx[2 * g.nb] = a;
            g.idx[2 * g.nb + 1] = c;
            g.totals[0] = a;
            g.totals[1] = c;
        }
    }
}

__global__ __launch_bounds__(256) void k_gen_bitmap(GenericCompress g)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= (g.nb + 7) / 8) return;
    uint32_t v = 0;
    for (int k = 0; k < 8; k++) {
        const uint64_t b = 8 * i + k;
        if (b < g.nb && g.size[b] != 0) v |= 0x80u >> k;
    }
    g.bitmap[i] = (uint8_t)v;
}

// ---- decompression ----

// One block per segment of S lanes, 4 consecutive values per lane.  The
// workgroup's records (contiguous in the payload) are staged in LDS with
// aligned word loads, and its values leave through LDS with aligned 16-byte
// stores.  The record's lead codes give every value's stream byte count; a
// segment scan places each lane's bytes, and a second scan carries the
// inherited top bytes (szxd_float.cc:78-199: a word takes its first
// min(lead, rb) bytes from the previous word) from lane to lane.  Decodes
// block b into out (its values in LDS); a block that fails a check sets the
// status and leaves its values unwritten.
template <typename T, int S>
__device__ inline void decode_segment(const GenericDecode &p, uint64_t b, uint64_t bend, uint64_t P0,
                                      uint64_t P1, uint32_t head, const uint32_t *stage, T *out)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    const DecodeHeader *h = p.h;
    const uint64_t bs = h->bs;
    const int l = threadIdx.x & (S - 1);
    if (b >= bend) return;   // whole segments
    const uint64_t rest = p.n - b * bs;
    const int m = rest < bs ? (int)rest : (int)bs;
    bool val[4];
#pragma unroll
    for (int k = 0; k < 4; k++) val[k] = 4 * l + k < m;
    const uint64_t off = p.idx[2 * b], rank = p.idx[2 * b + 1];
    const uint64_t noff = p.idx[2 * b + 2], nrank = p.idx[2 * b + 3];
    if (!gen_nonconst(p.s + h->qbit, b)) {
        // constant block: its median, widened for fp64 (szxd_float.cc:49-56)
        if (rank >= h->nconst || nrank != rank + 1 || noff != off) {
            if (l == 0) atomicMin(p.status, -21);
            return;
        }
        const uint8_t *c = p.s + h->qmed + 4 * rank;
        const uint32_t be = ((uint32_t)c[0] << 24) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 8) | (uint32_t)c[3];
        const T med = (T)bits_f32(h->medle ? bswap32(be) : be);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (val[k]) out[4 * l + k] = med;
        return;
    }
    const int nlead = (m + 3) >> 2;
    if (nrank != rank || off < P0 || noff > P1 || noff < off || noff - off < 5u + (uint64_t)nlead) {
        if (l == 0) atomicMin(p.status, -21);
        return;
    }
    const uint8_t *lds = reinterpret_cast<const uint8_t *>(stage);
    const uint32_t r0 = head + (uint32_t)(off - P0);   // record start in LDS
    const uint8_t *rec = lds + r0;
    const int rl = rec[0];
    if (rl < 1 || rl > 8 * W) {
        if (l == 0) atomicMin(p.status, -21);
        return;
    }
    const int rb = (rl + 7) >> 3;
    const int shift = (-rl) & 7;
    const T med = (T)bits_f32(((uint32_t)rec[1] << 24) | ((uint32_t)rec[2] << 16) | ((uint32_t)rec[3] << 8) |
                              (uint32_t)rec[4]);
    const uint32_t code = 4 * l < m ? rec[5 + l] : 0u;
    int lk[4];
    uint32_t lsum = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int ld = (int)(code >> (6 - 2 * k)) & 3;
        lk[k] = ld < rb ? ld : rb;
        if (val[k]) lsum += rb - lk[k];
    }
    const uint32_t incl = seg_incl_sum<S>(lsum, l);
    const uint32_t total = __shfl(incl, S - 1, S);
    if (5u + (uint64_t)nlead + total != noff - off) {   // segment-uniform
        if (l == 0) atomicMin(p.status, -21);
        return;
    }
    // the lane's words from the stream alone (inherited bytes still zero),
    // masked to the low 8W - shift bits
    const U mask = ~(U)0 >> shift;
    uint32_t pos = r0 + 5 + (uint32_t)nlead + (incl - lsum);
    U wv[4];
    uint32_t mine = 0;   // this lane's carry contribution
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int cnt = rb - lk[k];
        U cur = 0;
        if (val[k] && cnt > 0) {
            const U raw = W == 4 ? (U)lds_window4(stage, pos) : (U)lds_window8(stage, pos);
            cur = (raw & (~(U)0 >> (8 * (W - cnt)))) << (8 * (W - rb));
            if (W == 8 && !p.exact && rb == 5 && lk[k] == 0)   // szxd_double.cc:240-245: byte 7 <- 4th byte
                cur = (cur & (~(U)0 >> 8)) | (((raw >> 24) & (U)0xFF) << (8 * (W - 1)));
            pos += cnt;
        }
        cur &= mask;
        if (val[k]) {
            // bytes W-1-i (i < 3) this word writes: those below W - lk
            uint32_t wr = 0;
#pragma unroll
            for (int i = 0; i < 3; i++)
                if (lk[k] <= i) wr |= ((uint32_t)(cur >> (8 * (W - 1 - i))) & 0xFFu) << (8 * i) | (1u << (24 + i));
            mine = carry_combine(mine, wr);
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
    if (l == 0) cin = 0;   // the first word of a block follows a zero word
    U prev = ((U)(cin & 0xFFu) << (8 * (W - 1))) | ((U)((cin >> 8) & 0xFFu) << (8 * (W - 2))) |
             ((U)((cin >> 16) & 0xFFu) << (8 * (W - 3)));
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const U top = lk[k] ? ~(U)0 << (8 * (W - lk[k])) : (U)0;
        const U cur = wv[k] | (prev & top);
        if (val[k]) out[4 * l + k] = from_bits<T>(cur << shift) + med;
        prev = cur;
    }
}

template <typename T, int S>
__global__ __launch_bounds__(256) void k_gen_decode(GenericDecode p)
{
    constexpr int W = Word<T>::W;
    constexpr int BPW = 256 / S;
    __shared__ uint32_t stage[Span<S, W>::WORDS];
    __shared__ __align__(16) T xs[Vals<T>::N];
    const DecodeHeader *h = p.h;
    if (h->rc != 0) return;
    const uint64_t nb = h->nb, bs = h->bs, q0 = h->q0;
    const uint64_t b0 = blockIdx.x * (uint64_t)BPW;
    if (b0 >= nb) return;
    const uint64_t bend = b0 + BPW < nb ? b0 + BPW : nb;
    const uint64_t P0 = p.idx[2 * b0], P1 = p.idx[2 * bend];
    const uint64_t rmax = 5 + (bs + 3) / 4 + bs * W;
    if (P1 < P0 || P1 - P0 > (bend - b0) * rmax || P1 > p.len - q0) {   // workgroup-uniform
        if (threadIdx.x == 0) atomicMin(p.status, -21);
        return;
    }
    // stage payload bytes [P0, P1): aligned words inside the stream, the last
    // partial word byte by byte (q0 >= 13, so the first word starts inside)
    const uint8_t *src = p.s + q0 + P0;
    const uint8_t *end = p.s + p.len;
    const uint32_t head = (uint32_t)((uintptr_t)src & 3);
    const uint32_t *sw = reinterpret_cast<const uint32_t *>(src - head);
    const uint32_t span = (uint32_t)(P1 - P0);
    const uint32_t nw = (head + span + 3) / 4;
    for (uint32_t i = threadIdx.x; i < nw; i += 256) {
        const uint8_t *a = reinterpret_cast<const uint8_t *>(sw + i);
        if (a + 4 <= end) {
            stage[i] = sw[i];
        } else {
            uint32_t v = 0;
            for (int j = 0; j < 4; j++)
                if (a + j < end) v |= (uint32_t)a[j] << (8 * j);
            stage[i] = v;
        }
    }
    __syncthreads();
    const uint64_t e0 = b0 * bs, e1 = bend * bs < p.n ? bend * bs : p.n;
    decode_segment<T, S>(p, b0 + threadIdx.x / S, bend, P0, P1, head, stage,
                         xs + (b0 + threadIdx.x / S) * bs - e0 / Vals<T>::CE * Vals<T>::CE);
    __syncthreads();
    unstage_values<T>(static_cast<T *>(p.out), e0, e1, xs);
}

// ---- host side ----

uint64_t generic_parts(uint64_t nb) { return ceil_div(nb ? nb : 1, (uint64_t)GEN_CH); }

// Lanes per block: the smallest power of two S with 4 S >= bs.
static int seg_lanes(int bs)
{
    int s = 1;
    while (4 * s < bs) s <<= 1;
    return s;
}

template <typename T, int S>
static void encode_pass(const GenericCompress &g, bool write, hipStream_t st)
{
    const unsigned gb = (unsigned)ceil_div(g.nb * S, 256);
    if (write) hipLaunchKernelGGL((k_gen_encode<T, S, true>), dim3(gb), dim3(256), 0, st, g);
    else hipLaunchKernelGGL((k_gen_encode<T, S, false>), dim3(gb), dim3(256), 0, st, g);
}

template <typename T>
static void encode_dispatch(const GenericCompress &g, bool write, hipStream_t st)
{
    switch (seg_lanes(g.bs)) {
    case 1: encode_pass<T, 1>(g, write, st); break;
    case 2: encode_pass<T, 2>(g, write, st); break;
    case 4: encode_pass<T, 4>(g, write, st); break;
    case 8: encode_pass<T, 8>(g, write, st); break;
    case 16: encode_pass<T, 16>(g, write, st); break;
    case 32: encode_pass<T, 32>(g, write, st); break;
    default: encode_pass<T, 64>(g, write, st); break;
    }
}

// The scan and the bitmap alone, from g.size (callers with their own
// per-block passes, e.g. the cuSZx writer).
void generic_scan_launch(const GenericCompress &g, hipStream_t st)
{
    const unsigned np = (unsigned)g.nparts;
    hipLaunchKernelGGL(k_gen_part, dim3(np), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_top, dim3(1), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_apply, dim3(np), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_bitmap, dim3((unsigned)ceil_div(ceil_div(g.nb, 8), 256)), dim3(256), 0, st, g);
}

void generic_compress_launch(const GenericCompress &g, int w, hipStream_t st)
{
    const unsigned np = (unsigned)g.nparts;
    if (w == 4) encode_dispatch<float>(g, false, st);
    else encode_dispatch<double>(g, false, st);
    hipLaunchKernelGGL(k_gen_part, dim3(np), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_top, dim3(1), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_apply, dim3(np), dim3(256), 0, st, g);
    if (w == 4) encode_dispatch<float>(g, true, st);
    else encode_dispatch<double>(g, true, st);
    hipLaunchKernelGGL(k_gen_bitmap, dim3((unsigned)ceil_div(ceil_div(g.nb, 8), 256)), dim3(256), 0, st, g);
}

template <typename T, int S>
static void decode_pass(const GenericDecode &p, uint64_t nb, hipStream_t st)
{
    hipLaunchKernelGGL((k_gen_decode<T, S>), dim3((unsigned)ceil_div(nb * S, 256)), dim3(256), 0, st, p);
}

template <typename T>
static void decode_dispatch(const GenericDecode &p, int bs, uint64_t nb, hipStream_t st)
{
    switch (seg_lanes(bs)) {
    case 1: decode_pass<T, 1>(p, nb, st); break;
    case 2: decode_pass<T, 2>(p, nb, st); break;
    case 4: decode_pass<T, 4>(p, nb, st); break;
    case 8: decode_pass<T, 8>(p, nb, st); break;
    case 16: decode_pass<T, 16>(p, nb, st); break;
    case 32: decode_pass<T, 32>(p, nb, st); break;
    default: decode_pass<T, 64>(p, nb, st); break;
    }
}

void generic_decode_launch(const GenericDecode &p, int w, int bs, uint64_t nb, hipStream_t st)
{
    if (!nb) return;
    if (w == 4) decode_dispatch<float>(p, bs, nb, st);
    else decode_dispatch<double>(p, bs, nb, st);
}

}  // namespace szxamd
