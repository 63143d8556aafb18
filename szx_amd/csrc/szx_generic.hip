// szx_generic.hip -- the blocked codec for block sizes the tiled kernels do
// not take.
//
// The tiled kernels (szx_compress.hip, szx_decompress.hip) map a block onto an
// aligned group of lanes and need a power-of-two block size of at least 4.
// The reference's blocked API takes any block size that fits its one-byte
// header field (szx_float.cc:1095), so every other size 1..255 runs here: one
// lane per block, walking the block's values in order exactly as the
// reference's per-block loops do (szx_float.cc:314-371 and 49-252 for
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
// Decompression: k_gen_decode, one lane per block, checks every record
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

template <typename T, bool WRITE>
__global__ __launch_bounds__(256) void k_gen_encode(GenericCompress g)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= g.nb) return;
    const float eb = g.eb_dev ? *g.eb_dev : g.eb;
    const T *x = static_cast<const T *>(g.x) + b * (uint64_t)g.bs;
    const uint64_t rest = g.n - b * (uint64_t)g.bs;
    const int m = rest < (uint64_t)g.bs ? (int)rest : g.bs;

    // block statistics with the reference's compare order (szx_float.cc:327-333,
    // szx_double.cc:985-991): NaNs never replace, a NaN first value stays, and
    // the first of two equal-comparing zeros is kept
    T lo = x[0], hi = x[0];
    for (int j = 1; j < m; j++) {
        const T v = x[j];
        if (lo > v) lo = v;
        else if (hi < v) hi = v;
    }
    const T rad = (hi - lo) / (T)2;
    const T med = lo + rad;
    const float med_f = (float)med, rad_f = (float)rad;
    const bool constant = W == 4 ? (rad <= (T)eb) : ((double)rad <= (double)eb);
    if (constant) {
        if (WRITE) g.cmed[g.idx[2 * b + 1]] = bswap32(f32_bits(med_f));
        else g.size[b] = 0;
        return;
    }
    int clamped;
    const int len = req_length<T>(rad_expo(rad_f), req_expo(eb), clamped);
    const int rb = (len + 7) >> 3;
    const int shift = (-len) & 7;
    const float hmed = clamped ? 0.0f : med_f;
    // fp32 encodes against the stored median; fp64 against the median before
    // the clamp (szx_float.cc:60,90; szx_double.cc:38,78)
    const T msub = W == 4 ? (T)hmed : (T)med_f;
    const int nlead = (m + 3) >> 2;

    uint8_t *rec = WRITE ? g.payload + g.idx[2 * b] : nullptr;
    if (WRITE) {
        const uint32_t hb = f32_bits(hmed);
        rec[0] = (uint8_t)len;
        rec[1] = (uint8_t)(hb >> 24);
        rec[2] = (uint8_t)(hb >> 16);
        rec[3] = (uint8_t)(hb >> 8);
        rec[4] = (uint8_t)hb;
    }
    uint32_t pos = 5 + (uint32_t)nlead;
    uint32_t code = 0;
    U prev = 0;
    for (int i = 0; i < m; i++) {
        const U s = to_bits<T>(x[i] - msub) >> shift;
        const int ld = gen_lead<U>(s ^ prev);
        prev = s;
        if (WRITE) {
            code |= (uint32_t)ld << (6 - 2 * (i & 3));
            if ((i & 3) == 3 || i == m - 1) {
                rec[5 + (i >> 2)] = (uint8_t)code;
                code = 0;
            }
            for (int k = W - rb; k < W - ld; k++) rec[pos++] = (uint8_t)(s >> (8 * k));
        } else if (ld < rb) {
            pos += rb - ld;
        }
    }
    if (!WRITE) g.size[b] = pos;
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

template <typename T>
__global__ __launch_bounds__(256) void k_gen_decode(GenericDecode p)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    const DecodeHeader *h = p.h;
    if (h->rc != 0) return;
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nb = h->nb, bs = h->bs, q0 = h->q0;
    if (b >= nb) return;
    const uint64_t rest = p.n - b * bs;
    const int m = rest < bs ? (int)rest : (int)bs;
    T *out = static_cast<T *>(p.out) + b * bs;
    const uint64_t off = p.idx[2 * b], rank = p.idx[2 * b + 1];
    const uint64_t noff = p.idx[2 * b + 2], nrank = p.idx[2 * b + 3];
    if (!gen_nonconst(p.s + HEADER_BYTES, b)) {
        // constant block: its median, widened for fp64 (szxd_float.cc:49-56)
        if (rank >= h->nconst || nrank != rank + 1 || noff != off) {
            atomicMin(p.status, -21);
            return;
        }
        const uint8_t *c = p.s + HEADER_BYTES + h->nbit + 4 * rank;
        const float med = bits_f32(((uint32_t)c[0] << 24) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 8) |
                                   (uint32_t)c[3]);
        for (int j = 0; j < m; j++) out[j] = (T)med;
        return;
    }
    const int nlead = (m + 3) >> 2;
    if (nrank != rank || noff < off || noff - off < 5u + (uint64_t)nlead || noff > p.len - q0) {
        atomicMin(p.status, -21);
        return;
    }
    const uint8_t *rec = p.s + q0 + off;
    const uint8_t *end = p.s + q0 + noff;
    const int rl = rec[0];
    if (rl < 1 || rl > 8 * W) {
        atomicMin(p.status, -21);
        return;
    }
    // szxd_float.cc:78-199: lead bytes from the previous word, the rest from
    // the stream, masked to the low 8W - shift bits, shifted back, median added
    const int rb = (rl + 7) >> 3;
    const int shift = (-rl) & 7;
    const T med = (T)bits_f32(((uint32_t)rec[1] << 24) | ((uint32_t)rec[2] << 16) | ((uint32_t)rec[3] << 8) |
                              (uint32_t)rec[4]);
    const uint8_t *q = rec + 5 + nlead;
    const U mask = ~(U)0 >> shift;
    U prev = 0;
    for (int i = 0; i < m; i++) {
        const int ld = (rec[5 + (i >> 2)] >> (6 - 2 * (i & 3))) & 3;
        const int l = ld < rb ? ld : rb;
        if (q + (rb - l) > end) {
            atomicMin(p.status, -21);
            return;
        }
        U cur = l ? (prev & (~(U)0 << (8 * (W - l)))) : (U)0;
        for (int k = W - rb; k < W - l; k++) {
            int src = k - (W - rb);
            if (W == 8 && !p.exact && rb == 5 && l == 0 && k == 7) src = 3;   // szxd_double.cc:240-245
            cur |= (U)q[src] << (8 * k);
        }
        q += rb - l;
        cur &= mask;
        out[i] = from_bits<T>(cur << shift) + med;
        prev = cur;
    }
    if (q != end) atomicMin(p.status, -21);
}

// ---- host side ----

uint64_t generic_parts(uint64_t nb) { return ceil_div(nb ? nb : 1, (uint64_t)GEN_CH); }

void generic_compress_launch(const GenericCompress &g, int w, hipStream_t st)
{
    const unsigned gb = (unsigned)ceil_div(g.nb, 256);
    const unsigned np = (unsigned)g.nparts;
    if (w == 4) hipLaunchKernelGGL((k_gen_encode<float, false>), dim3(gb), dim3(256), 0, st, g);
    else hipLaunchKernelGGL((k_gen_encode<double, false>), dim3(gb), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_part, dim3(np), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_top, dim3(1), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_apply, dim3(np), dim3(256), 0, st, g);
    if (w == 4) hipLaunchKernelGGL((k_gen_encode<float, true>), dim3(gb), dim3(256), 0, st, g);
    else hipLaunchKernelGGL((k_gen_encode<double, true>), dim3(gb), dim3(256), 0, st, g);
    hipLaunchKernelGGL(k_gen_bitmap, dim3((unsigned)ceil_div(ceil_div(g.nb, 8), 256)), dim3(256), 0, st, g);
}

void generic_decode_launch(const GenericDecode &p, int w, uint64_t nb, hipStream_t st)
{
    const unsigned gb = (unsigned)ceil_div(nb, 256);
    if (!gb) return;
    if (w == 4) hipLaunchKernelGGL(k_gen_decode<float>, dim3(gb), dim3(256), 0, st, p);
    else hipLaunchKernelGGL(k_gen_decode<double>, dim3(gb), dim3(256), 0, st, p);
}

}  // namespace szxamd
