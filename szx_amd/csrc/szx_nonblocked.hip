// szx_nonblocked.hip -- the non-blocked codec (fastMode 1): one record over
// all n values against the global median.
//
// Stream (szx_float.cc:1492-1699, szx_double.cc:1593-2010): [1,1,1,0], the
// reqLength byte, the big-endian f32 median, the big-endian u64 size of the
// lead-code array (ceil(n / 4)), the lead codes, then the mid bytes of every
// value in order.  The median and radius come from the value range
// (computeValueRange_*: for fp64 the radius is rounded to float before the
// median is formed); reqLength, the shift, the lead codes and the mid bytes
// follow the blocked codec's rules with the whole array as one block.
//
// Every value's lead code depends only on its predecessor, so the stream is
// built in tiles of 4096 values (16 consecutive values per lane):
//   compress:   k_nb_setup (median, reqLength), k_nb_pass<T, false> (lead
//               codes straight into the stream, mid-byte count per tile), a
//               scan over the tiles, k_nb_pass<T, true> (mid bytes at their
//               offsets), k_nb_finish (header, sizes)
//   decompress: k_nb_header, k_nb_dcount (mid-byte count per tile from the
//               lead codes), a scan, k_nb_carry (per tile: the last bytes it
//               writes at the three positions a later value can inherit),
//               k_nb_carry_scan (the bytes every tile inherits, last writer
//               wins), k_nb_decode.  The decoder stages each tile's mid bytes
//               (contiguous in the stream) in LDS with aligned word loads.
// A value inherits its first min(lead, rb) bytes from the previous value
// (szxd_float.cc:78-199), possibly across many tiles; the carry scan resolves
// that chain with the same last-writer operator the blocked decoders use.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "szx_common.hpp"
#include "szx_internal.hpp"

namespace szxamd {

constexpr int NB_PER = 16;                 // values per lane
constexpr int NB_TILE = 256 * NB_PER;      // values per tile
constexpr uint64_t NB_HDR = 17;            // header bytes before the lead codes

// Codec parameters decided on the device (k_nb_setup / k_nb_header).
struct NbParams {
    int32_t len, rb, shift, ok;
    uint32_t hdr_bits;     // stored median (f32 bits)
    float eb;
    double msub;           // median subtracted before the shift (fp32: the stored one)
    float med;             // decoder: the stored median
    uint64_t total;        // mid bytes
};

template <typename U> __device__ inline int nb_lead(U xr)
{
    constexpr int W = sizeof(U);
    if ((xr >> (8 * (W - 3))) == 0) return 3;
    if ((xr >> (8 * (W - 2))) == 0) return 2;
    if ((xr >> (8 * (W - 1))) == 0) return 1;
    return 0;
}

// Exclusive sum over the 256 lanes of a workgroup; *total receives the sum.
__device__ inline uint32_t nb_block_excl(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t s_w[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int k = 0; k < 4; k++) {
        if (k < wv) pre += s_w[k];
        tot += s_w[k];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

// Last-writer record of the three inheritable byte positions W-1, W-2, W-3:
// bytes 0..2 hold them, bits 24..26 say which were written; b's written bytes
// win over a's.
__device__ inline uint32_t nb_combine(uint32_t a, uint32_t b)
{
    const uint32_t f = b >> 24;
    const uint32_t keep = ((f & 1u) ? 0xFFu : 0u) | ((f & 2u) ? 0xFF00u : 0u) | ((f & 4u) ? 0xFF0000u : 0u);
    return ((b & keep) | (a & ~keep & 0xFFFFFFu)) | ((a | b) & 0x7000000u);
}

// Exclusive last-writer scan over the 256 lanes (lane 0 gets 0).
__device__ inline uint32_t nb_block_excl_lw(uint32_t v, uint32_t *total)
{
    __shared__ uint32_t s_w[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc = nb_combine(o, inc);
    }
    uint32_t ex = __shfl_up(inc, 1, 64);
    if (lane == 0) ex = 0;
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
    for (int k = 0; k < 4; k++) {
        if (k < wv) pre = nb_combine(pre, s_w[k]);
        tot = nb_combine(tot, s_w[k]);
    }
    __syncthreads();
    *total = tot;
    return nb_combine(pre, ex);
}

// ---- compression ----

// Median, radius and reqLength from the range (szx_float.cc:494-495,
// szx_double.cc:1153-1154; computeReqLength_*, szx_float.cc:1708-1717).
template <typename T>
__global__ void k_nb_setup(const T *mm, const float *eb_dev, float eb_abs, NbParams *P)
{
    if (threadIdx.x != 0) return;
    const float eb = eb_dev ? *eb_dev : eb_abs;
    const T lo = mm[0], hi = mm[1];
    float rad_f, med_f;
    if (sizeof(T) == 4) {
        const T rad = (hi - lo) / (T)2;
        rad_f = (float)rad;
        med_f = (float)(lo + rad);
    } else {
        rad_f = (float)((hi - lo) / (T)2);
        med_f = (float)(lo + (T)rad_f);
    }
    int clamped;
    const int len = req_length<T>(rad_expo(rad_f), req_expo(eb), clamped);
    const float hdr = clamped ? 0.0f : med_f;
    P->len = len;
    P->rb = (len + 7) >> 3;
    P->shift = (-len) & 7;
    P->ok = 1;
    P->hdr_bits = f32_bits(hdr);
    P->eb = eb;
    P->msub = sizeof(T) == 4 ? (double)hdr : (double)med_f;   // szx_float.cc:1516 / szx_double.cc:1597
}

template <typename T>
__device__ inline typename Word<T>::U nb_word(T v, const NbParams &P)
{
    return to_bits<T>(v - (T)P.msub) >> P.shift;
}

// Lead codes of the lane's 16 values and their mid-byte counts.  WRITE =
// false stores the lead-code bytes (4 per lane, at their fixed place) and the
// tile's mid-byte count; WRITE = true stores the mid bytes at their offsets.
template <typename T, bool WRITE>
__global__ __launch_bounds__(256) void k_nb_pass(const T *x, uint64_t n, const NbParams *Pp, uint8_t *out,
                                                 uint32_t *tsum, const uint64_t *toff)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    const NbParams P = *Pp;
    const uint64_t e0 = blockIdx.x * (uint64_t)NB_TILE + threadIdx.x * (uint64_t)NB_PER;
    U s[NB_PER];
    int ld[NB_PER];
    uint32_t cnt = 0;
    U prev = e0 > 0 && e0 <= n ? nb_word<T>(x[e0 - 1], P) : (U)0;
#pragma unroll
    for (int k = 0; k < NB_PER; k++) {
        const bool valid = e0 + k < n;
        s[k] = valid ? nb_word<T>(x[e0 + k], P) : (U)0;
        ld[k] = valid ? nb_lead<U>(s[k] ^ prev) : 0;
        if (valid && ld[k] < P.rb) cnt += P.rb - ld[k];
        prev = s[k];
    }
    uint32_t total;
    const uint32_t ex = nb_block_excl(cnt, &total);
    if (!WRITE) {
        // lead codes, MSB-first four to a byte (szx_TypeManager.cc:198-273)
        const uint64_t nlead = (n + 3) / 4;
#pragma unroll
        for (int j = 0; j < NB_PER / 4; j++) {
            const uint64_t lb = e0 / 4 + j;
            if (lb < nlead)
                out[NB_HDR + lb] = (uint8_t)((ld[4 * j] << 6) | (ld[4 * j + 1] << 4) | (ld[4 * j + 2] << 2) |
                                             ld[4 * j + 3]);
        }
        if (threadIdx.x == 0) tsum[blockIdx.x] = total;
        return;
    }
    uint8_t *q = out + NB_HDR + (n + 3) / 4 + toff[blockIdx.x] + ex;
#pragma unroll
    for (int k = 0; k < NB_PER; k++)
        if (e0 + k < n)
            for (int t = W - P.rb; t < W - ld[k]; t++) *q++ = (uint8_t)(s[k] >> (8 * t));
}

// One workgroup: exclusive scan of the tile sums (toff), total in P->total.
__global__ __launch_bounds__(1024) void k_nb_scan(const uint32_t *tsum, uint64_t ntiles, uint64_t *toff,
                                                  NbParams *P)
{
    __shared__ uint64_t s[1024];
    const uint64_t per = (ntiles + 1023) / 1024;
    const uint64_t a = threadIdx.x * per, b = a + per < ntiles ? a + per : ntiles;
    uint64_t sum = 0;
    for (uint64_t i = a; i < b; i++) sum += tsum[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint64_t v = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint64_t run = s[threadIdx.x] - sum;
    for (uint64_t i = a; i < b; i++) {
        toff[i] = run;
        run += tsum[i];
    }
    if (threadIdx.x == 1023) P->total = s[1023];
}

// Header (szx_float.cc:1678-1688) and the results.
__global__ void k_nb_finish(uint64_t n, const NbParams *P, uint8_t *out, uint64_t *result)
{
    if (threadIdx.x != 0) return;
    const uint64_t nlead = (n + 3) / 4;
    out[0] = 1; out[1] = 1; out[2] = 1; out[3] = 0;
    out[4] = (uint8_t)P->len;
    for (int k = 0; k < 4; k++) out[5 + k] = (uint8_t)(P->hdr_bits >> (24 - 8 * k));
    for (int k = 0; k < 8; k++) out[9 + k] = (uint8_t)(nlead >> (56 - 8 * k));
    result[0] = 0;
    result[1] = NB_HDR + nlead + P->total;
    result[2] = f32_bits(P->eb);
    result[3] = 0;
}

// ---- decompression ----

// szxd_float.cc:478-513: reqLength, median, lead-array size.
__global__ void k_nb_header(const uint8_t *s, uint64_t len, uint64_t n, int w, NbParams *P, int32_t *status)
{
    if (threadIdx.x != 0) return;
    int32_t rc = 0;
    const uint64_t nlead = (n + 3) / 4;
    int rl = 1;
    uint32_t mb = 0;
    if (len < NB_HDR) {
        rc = -10;
    } else {
        rl = s[4];
        uint64_t ls = 0;
        for (int k = 0; k < 8; k++) ls = (ls << 8) | s[9 + k];
        for (int k = 0; k < 4; k++) mb = (mb << 8) | s[5 + k];
        if (rl < 1 || rl > 8 * w) rc = -11;
        else if (ls != nlead) rc = -12;
        else if (NB_HDR + nlead > len) rc = -13;
    }
    if (rc) rl = 1;
    P->len = rl;
    P->rb = (rl + 7) >> 3;
    P->shift = (-rl) & 7;
    P->ok = rc == 0;
    P->med = bits_f32(mb);
    *status = rc;
}

// Lead code k of the lane's values from the stream.
__device__ inline int nb_code(const uint8_t *s, uint64_t e)
{
    return (s[NB_HDR + (e >> 2)] >> (6 - 2 * (int)(e & 3))) & 3;
}

__global__ __launch_bounds__(256) void k_nb_dcount(const uint8_t *s, uint64_t n, const NbParams *Pp, uint32_t *tsum)
{
    const NbParams P = *Pp;
    if (!P.ok) return;
    const uint64_t e0 = blockIdx.x * (uint64_t)NB_TILE + threadIdx.x * (uint64_t)NB_PER;
    uint32_t cnt = 0;
    for (int k = 0; k < NB_PER; k++) {
        if (e0 + k >= n) break;
        const int ld = nb_code(s, e0 + k);
        cnt += P.rb - (ld < P.rb ? ld : P.rb);
    }
    uint32_t total;
    nb_block_excl(cnt, &total);
    if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// The payload must hold the mid bytes the lead codes ask for.
__global__ void k_nb_check(uint64_t len, uint64_t n, NbParams *P, int32_t *status)
{
    if (threadIdx.x != 0 || !P->ok) return;
    if (P->total > len - NB_HDR - (n + 3) / 4) {
        P->ok = 0;
        *status = -33;
    }
}

// The lane's words from the stream alone (inherited bytes still zero),
// masked to the low 8W - shift bits, and its last-writer record.
template <typename T>
__device__ inline uint32_t nb_words(const uint8_t *s, const uint8_t *mids, uint64_t n, uint64_t e0, const NbParams &P,
                                    uint32_t q, int exact, typename Word<T>::U (&wv)[NB_PER], int (&lk)[NB_PER])
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    const U mask = ~(U)0 >> P.shift;
    uint32_t mine = 0;
    for (int k = 0; k < NB_PER; k++) {
        U cur = 0;
        lk[k] = 0;
        if (e0 + k < n) {
            const int ld = nb_code(s, e0 + k);
            lk[k] = ld < P.rb ? ld : P.rb;
            for (int t = W - P.rb; t < W - lk[k]; t++) {
                uint32_t src = (uint32_t)(t - (W - P.rb));
                if (W == 8 && !exact && P.rb == 5 && lk[k] == 0 && t == 7) src = 3;   // szxd_double.cc:906
                cur |= (U)mids[q + src] << (8 * t);
            }
            q += P.rb - lk[k];
            cur &= mask;
            uint32_t wr = 0;
            for (int i = 0; i < 3; i++)
                if (lk[k] <= i) wr |= ((uint32_t)(cur >> (8 * (W - 1 - i))) & 0xFFu) << (8 * i) | (1u << (24 + i));
            mine = nb_combine(mine, wr);
        }
        wv[k] = cur;
    }
    return mine;
}

// Stage the tile's mid bytes [src, src + cnt) of the stream (ending at end)
// into LDS words at byte offset (src & 3); returns that offset.
__device__ inline uint32_t nb_stage(uint32_t *bw, const uint8_t *src, uint32_t cnt, const uint8_t *end)
{
    const uint32_t head = (uint32_t)((uintptr_t)src & 3);
    const uint32_t *sw = reinterpret_cast<const uint32_t *>(src - head);
    const uint32_t nw = (head + cnt + 3) / 4;
    for (uint32_t i = threadIdx.x; i < nw; i += 256) {
        const uint8_t *a = reinterpret_cast<const uint8_t *>(sw + i);
        if (a + 4 <= end) {
            bw[i] = sw[i];
        } else {
            uint32_t v = 0;
            for (int j = 0; j < 4; j++)
                if (a + j < end) v |= (uint32_t)a[j] << (8 * j);
            bw[i] = v;
        }
    }
    __syncthreads();
    return head;
}

template <typename T>
__global__ __launch_bounds__(256) void k_nb_carry(const uint8_t *s, uint64_t len, uint64_t n, const NbParams *Pp,
                                                  const uint64_t *toff, uint32_t *tcar, int exact)
{
    using U = typename Word<T>::U;
    __shared__ uint32_t bw[(NB_TILE * sizeof(T)) / 4 + 2];
    const NbParams P = *Pp;
    if (!P.ok) return;
    const uint64_t e0 = blockIdx.x * (uint64_t)NB_TILE + threadIdx.x * (uint64_t)NB_PER;
    uint32_t cnt = 0;
    for (int k = 0; k < NB_PER; k++) {
        if (e0 + k >= n) break;
        const int ld = nb_code(s, e0 + k);
        cnt += P.rb - (ld < P.rb ? ld : P.rb);
    }
    uint32_t total;
    const uint32_t ex = nb_block_excl(cnt, &total);
    const uint32_t head = nb_stage(bw, s + NB_HDR + (n + 3) / 4 + toff[blockIdx.x], total, s + len);
    U wv[NB_PER];
    int lk[NB_PER];
    const uint32_t mine = nb_words<T>(s, reinterpret_cast<const uint8_t *>(bw), n, e0, P, head + ex, exact, wv, lk);
    uint32_t tot;
    nb_block_excl_lw(mine, &tot);
    if (threadIdx.x == 0) tcar[blockIdx.x] = tot;
}

// One workgroup: exclusive last-writer scan of the tile records (in place).
__global__ __launch_bounds__(1024) void k_nb_carry_scan(uint32_t *tcar, uint64_t ntiles, const NbParams *P)
{
    __shared__ uint32_t s[1024];
    if (!P->ok) return;
    const uint64_t per = (ntiles + 1023) / 1024;
    const uint64_t a = threadIdx.x * per, b = a + per < ntiles ? a + per : ntiles;
    uint32_t acc = 0;
    for (uint64_t i = a; i < b; i++) acc = nb_combine(acc, tcar[i]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0u;
        __syncthreads();
        if (threadIdx.x >= (unsigned)d) s[threadIdx.x] = nb_combine(v, s[threadIdx.x]);
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? s[threadIdx.x - 1] : 0u;
    for (uint64_t i = a; i < b; i++) {
        const uint32_t t = tcar[i];
        tcar[i] = run;
        run = nb_combine(run, t);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_nb_decode(const uint8_t *s, uint64_t len, uint64_t n, const NbParams *Pp,
                                                   const uint64_t *toff, const uint32_t *tcar, T *out, int exact)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    __shared__ uint32_t bw[(NB_TILE * sizeof(T)) / 4 + 2];
    const NbParams P = *Pp;
    if (!P.ok) return;
    const uint64_t e0 = blockIdx.x * (uint64_t)NB_TILE + threadIdx.x * (uint64_t)NB_PER;
    uint32_t cnt = 0;
    for (int k = 0; k < NB_PER; k++) {
        if (e0 + k >= n) break;
        const int ld = nb_code(s, e0 + k);
        cnt += P.rb - (ld < P.rb ? ld : P.rb);
    }
    uint32_t total;
    const uint32_t ex = nb_block_excl(cnt, &total);
    const uint32_t head = nb_stage(bw, s + NB_HDR + (n + 3) / 4 + toff[blockIdx.x], total, s + len);
    U wv[NB_PER];
    int lk[NB_PER];
    const uint32_t mine = nb_words<T>(s, reinterpret_cast<const uint8_t *>(bw), n, e0, P, head + ex, exact, wv, lk);
    uint32_t tot;
    const uint32_t cin = nb_combine(tcar[blockIdx.x], nb_block_excl_lw(mine, &tot));
    U prev = ((U)(cin & 0xFFu) << (8 * (W - 1))) | ((U)((cin >> 8) & 0xFFu) << (8 * (W - 2))) |
             ((U)((cin >> 16) & 0xFFu) << (8 * (W - 3)));
    const T med = (T)P.med;
    for (int k = 0; k < NB_PER; k++) {
        if (e0 + k >= n) break;
        const U top = lk[k] ? ~(U)0 << (8 * (W - lk[k])) : (U)0;
        const U cur = wv[k] | (prev & top);
        out[e0 + k] = from_bits<T>(cur << P.shift) + med;
        prev = cur;
    }
}

// ---- host side ----

static uint64_t nb_tiles(uint64_t n) { return n ? (n + NB_TILE - 1) / NB_TILE : 1; }

uint64_t nb_capacity(uint64_t n, int w) { return NB_HDR + (n + 3) / 4 + n * (uint64_t)w + 16; }

// params | range partials | tile sums (u32) | tile offsets (u64) | tile carries (u32)
uint64_t nb_workspace(uint64_t n)
{
    const uint64_t t = nb_tiles(n);
    auto al = [](uint64_t v) { return (v + 255) & ~255ull; };
    return 256 + al(range_workspace_bytes()) + al(4 * t) + al(8 * t) + al(4 * t);
}

struct NbWs {
    NbParams *P;
    float *eb;
    void *mm;
    uint8_t *range;
    uint32_t *tsum;
    uint64_t *toff;
    uint32_t *tcar;
};

static NbWs nb_ws(uint8_t *ws, uint64_t n)
{
    const uint64_t t = nb_tiles(n);
    auto al = [](uint64_t v) { return (v + 255) & ~255ull; };
    NbWs r;
    r.P = reinterpret_cast<NbParams *>(ws);
    r.eb = reinterpret_cast<float *>(ws + 128);
    r.mm = ws + 192;
    uint64_t off = 256;
    r.range = ws + off; off += al(range_workspace_bytes());
    r.tsum = reinterpret_cast<uint32_t *>(ws + off); off += al(4 * t);
    r.toff = reinterpret_cast<uint64_t *>(ws + off); off += al(8 * t);
    r.tcar = reinterpret_cast<uint32_t *>(ws + off);
    return r;
}

void nb_compress_launch(const void *d_in, uint64_t n, int w, int rel_mode, float abs_eb, float rel, uint8_t *d_out,
                        uint8_t *d_ws, uint64_t *d_result, hipStream_t st)
{
    const NbWs S = nb_ws(d_ws, n);
    const uint64_t t = nb_tiles(n);
    range_bound_launch(d_in, n, w, rel, S.range, rel_mode ? S.eb : nullptr, S.mm, st);
    const float *ebd = rel_mode ? S.eb : nullptr;
    timing_mark_begin(KK_COMPRESS, st);
    if (w == 4) {
        hipLaunchKernelGGL(k_nb_setup<float>, dim3(1), dim3(64), 0, st, static_cast<const float *>(S.mm), ebd, abs_eb,
                           S.P);
        hipLaunchKernelGGL((k_nb_pass<float, false>), dim3((unsigned)t), dim3(256), 0, st,
                           static_cast<const float *>(d_in), n, S.P, d_out, S.tsum, S.toff);
    } else {
        hipLaunchKernelGGL(k_nb_setup<double>, dim3(1), dim3(64), 0, st, static_cast<const double *>(S.mm), ebd,
                           abs_eb, S.P);
        hipLaunchKernelGGL((k_nb_pass<double, false>), dim3((unsigned)t), dim3(256), 0, st,
                           static_cast<const double *>(d_in), n, S.P, d_out, S.tsum, S.toff);
    }
    hipLaunchKernelGGL(k_nb_scan, dim3(1), dim3(1024), 0, st, S.tsum, t, S.toff, S.P);
    if (w == 4)
        hipLaunchKernelGGL((k_nb_pass<float, true>), dim3((unsigned)t), dim3(256), 0, st,
                           static_cast<const float *>(d_in), n, S.P, d_out, S.tsum, S.toff);
    else
        hipLaunchKernelGGL((k_nb_pass<double, true>), dim3((unsigned)t), dim3(256), 0, st,
                           static_cast<const double *>(d_in), n, S.P, d_out, S.tsum, S.toff);
    timing_mark_end(KK_COMPRESS, st);
    hipLaunchKernelGGL(k_nb_finish, dim3(1), dim3(64), 0, st, n, S.P, d_out, d_result);
}

void nb_decompress_launch(const uint8_t *d_stream, uint64_t len, uint64_t n, int w, void *d_out, uint8_t *d_ws,
                          int exact, int32_t *d_status, hipStream_t st)
{
    const NbWs S = nb_ws(d_ws, n);
    const uint64_t t = nb_tiles(n);
    hipLaunchKernelGGL(k_nb_header, dim3(1), dim3(64), 0, st, d_stream, len, n, w, S.P, d_status);
    timing_mark_begin(KK_DECODE, st);
    hipLaunchKernelGGL(k_nb_dcount, dim3((unsigned)t), dim3(256), 0, st, d_stream, n, S.P, S.tsum);
    hipLaunchKernelGGL(k_nb_scan, dim3(1), dim3(1024), 0, st, S.tsum, t, S.toff, S.P);
    hipLaunchKernelGGL(k_nb_check, dim3(1), dim3(64), 0, st, len, n, S.P, d_status);
    if (w == 4) {
        hipLaunchKernelGGL(k_nb_carry<float>, dim3((unsigned)t), dim3(256), 0, st, d_stream, len, n, S.P, S.toff,
                           S.tcar, exact);
        hipLaunchKernelGGL(k_nb_carry_scan, dim3(1), dim3(1024), 0, st, S.tcar, t, S.P);
        hipLaunchKernelGGL(k_nb_decode<float>, dim3((unsigned)t), dim3(256), 0, st, d_stream, len, n, S.P, S.toff,
                           S.tcar, static_cast<float *>(d_out), exact);
    } else {
        hipLaunchKernelGGL(k_nb_carry<double>, dim3((unsigned)t), dim3(256), 0, st, d_stream, len, n, S.P, S.toff,
                           S.tcar, exact);
        hipLaunchKernelGGL(k_nb_carry_scan, dim3(1), dim3(1024), 0, st, S.tcar, t, S.P);
        hipLaunchKernelGGL(k_nb_decode<double>, dim3((unsigned)t), dim3(256), 0, st, d_stream, len, n, S.P, S.toff,
                           S.tcar, static_cast<double *>(d_out), exact);
    }
    timing_mark_end(KK_DECODE, st);
}

// SZx_fast_decompress_pred (szxd_float.cc:25-32): x[i] += pred[i] after a
// non-blocked decode, one IEEE float addition per value (no contraction).
__global__ __launch_bounds__(256) void k_add_pred(float *x, const float *pred, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) x[i] = x[i] + pred[i];
}

void add_pred_launch(float *d_x, const float *d_pred, uint64_t n, hipStream_t st)
{
    uint64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    if (g) hipLaunchKernelGGL(k_add_pred, dim3((unsigned)g), dim3(256), 0, st, d_x, d_pred, n);
}

}  // namespace szxamd
