// szx_decompress.hip -- blockwise fast decompressor for gfx950.
//
// A fastMode-2 stream has no block index: the start of non-constant record k
// is the sum of the sizes of records 0..k-1, and a size is only known by
// parsing that record's reqLength byte and lead codes (szxd_float.cc:276-283).
// Decoding is therefore two steps:
//
//   1. locate: find every record's byte offset (and every constant block's
//      median index) -> one 64-bit info word per block;
//   2. decode: every block in parallel, 4 values per lane, a block being an
//      aligned group of bs/4 lanes (the mirror of the compressor's layout).
//
// The decode of one record mirrors szxd_float.cc:34-244 / szxd_double.cc:25-449:
// value i takes its top min(lead_i, rb) bytes from value i-1 and the rest of
// its rb bytes from the stream.  The "from value i-1" chain is resolved in
// parallel: each lane folds its 4 values into a last-writer record of the
// (at most 3) copyable byte positions, and an exclusive segmented scan with
// the last-writer-wins operator gives every lane the bytes it inherits.
#include "szx_common.hpp"
#include "szx_internal.hpp"

namespace szxamd {

struct DecodeHeader {
    uint64_t bs, nb, nbit, nconst;
};

constexpr uint64_t INFO_CONST = 1ull << 63;

// Serial record walk (one lane).  Validates every length against the
// stream size so that the decode kernel never reads out of bounds.
__global__ void k_locate_serial(const uint8_t *s, uint64_t len, uint64_t n, int w,
                                DecodeHeader *h, uint64_t *info, uint64_t info_cap,
                                int32_t *status)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int32_t rc = 0;
    uint64_t bs = 0, nb = 0, nbit = 0, nconst = 0;
    if (len < HEADER_BYTES) { rc = -10; goto done; }
    bs = s[4] ? s[4] : 256;   // byte 0 stands for 256 ((u8)256 in the encoder)
    if (bs < 4 || (bs & (bs - 1))) { rc = -11; goto done; }
    nb = ceil_div(n, bs);
    nbit = ceil_div(nb, 8);
    for (int k = 0; k < 8; k++) nconst = (nconst << 8) | s[5 + k];
    if (nconst > nb) { rc = -12; goto done; }
    if (HEADER_BYTES + nbit + 4 * nconst > len) { rc = -13; goto done; }
    if (nb > info_cap) { rc = -14; goto done; }
    {
        uint64_t q = HEADER_BYTES + nbit + 4 * nconst, k0 = 0;
        for (uint64_t b = 0; b < nb; b++) {
            const int bit = (s[HEADER_BYTES + (b >> 3)] >> (7 - (b & 7))) & 1;
            if (!bit) {
                if (k0 >= nconst) { rc = -18; goto done; }
                info[b] = INFO_CONST | k0++;
                continue;
            }
            const uint64_t nel = (b + 1) * bs <= n ? bs : n - b * bs;
            const uint64_t nl = (nel + 3) >> 2;
            if (q + 5 + nl > len) { rc = -15; goto done; }
            const int rl = s[q];
            if (rl == 0 || rl > 8 * w) { rc = -16; goto done; }
            const int rb = (rl + 7) >> 3;
            uint64_t sz = 5 + nl;
            for (uint64_t i = 0; i < nel; i++) {
                const int ld = (s[q + 5 + (i >> 2)] >> (6 - 2 * (int)(i & 3))) & 3;
                if (ld < rb) sz += (uint64_t)(rb - ld);
            }
            if (q + sz > len) { rc = -17; goto done; }
            info[b] = q;
            q += sz;
        }
        if (k0 != nconst) rc = -18;
    }
done:
    h->bs = bs; h->nb = nb; h->nbit = nbit; h->nconst = nconst;
    *status = rc;
}

// last-writer-wins combine of the packed copyable-byte record:
// bytes 0..2 = values of positions W-3, W-2, W-1; bits 24..26 = written flags.
__device__ inline uint32_t lw_combine(uint32_t left, uint32_t right)
{
    const uint32_t hm = (right >> 24) & 7u;
    const uint32_t m = ((hm & 1u) ? 0xFFu : 0u) | ((hm & 2u) ? 0xFF00u : 0u) | ((hm & 4u) ? 0xFF0000u : 0u);
    return ((right & m) | (left & ~m & 0xFFFFFFu)) | ((left | right) & 0x7000000u);
}

template <typename T> struct Store4;
template <> struct Store4<float> {
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

template <typename T>
__global__ __launch_bounds__(256) void k_decode(const uint8_t *s, uint64_t n, const DecodeHeader *h,
                                               const uint64_t *info, const int32_t *status, T *out,
                                               int exact)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    if (*status != 0) return;
    const int bs = (int)h->bs;
    const int seg = bs >> 2;
    const uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t e0 = 4 * g;
    const int lseg = (int)(g % (uint64_t)seg);
    const uint64_t bstart = e0 - 4u * (uint64_t)lseg;
    const bool exists = bstart < n;
    const uint64_t gb = bstart / (uint64_t)bs;
    const uint64_t word = exists ? info[gb] : INFO_CONST;
    T v[4] = {0, 0, 0, 0};

    if (word & INFO_CONST) {
        if (exists) {
            const float med = be_f32(s + HEADER_BYTES + h->nbit + 4 * (word & ~INFO_CONST));
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = (T)med;
            Store4<T>::store(out, e0, n, v);
        }
        return;
    }

    // non-constant block: every lane of the segment takes this path
    const uint8_t *rec = s + word;
    const uint64_t rem = n - bstart;
    const int nel = rem >= (uint64_t)bs ? bs : (int)rem;
    const int nlead = (nel + 3) >> 2;
    const int rl = rec[0];
    const int rb = (rl + 7) >> 3;
    const int shift = (8 - (rl & 7)) & 7;
    const int lo = W - rb;
    const float med = be_f32(rec + 1);
    const uint32_t code = lseg < nlead ? rec[5 + lseg] : 0u;

    int lk[4];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int ld = (code >> (6 - 2 * k)) & 3;
        lk[k] = ld < rb ? ld : rb;
        if (e0 + k < n) tot += (uint32_t)(rb - lk[k]);
    }
    uint32_t inc = tot;
    for (int d = 1; d < seg; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, seg);
        if (lseg >= d) inc += o;
    }
    const uint8_t *src = rec + 5 + nlead + (inc - tot);

    const U wmask = ~(U)0 >> shift;
    U e[4];
    uint32_t carry = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        U ek = 0;
        if (e0 + k < n) {
            const int m = rb - lk[k];
            for (int j = 0; j < m; j++) ek |= (U)src[j] << (8 * (lo + j));
            if constexpr (W == 8) {
                if (!exact && rb == 5 && lk[k] == 0)   // szxd_double.cc:244 reads q[3]
                    ek = (ek & ~((U)0xFF << 56)) | ((U)src[3] << 56);
            }
            src += m;
            ek &= wmask;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int pos = W - 3 + q;
                if (pos >= lo && pos < W - lk[k]) {
                    carry = (carry & ~(0xFFu << (8 * q))) | ((uint32_t)((ek >> (8 * pos)) & 0xFF) << (8 * q));
                    carry |= 1u << (24 + q);
                }
            }
        }
        e[k] = ek;
    }
    // exclusive segmented last-writer scan of the carries
    uint32_t cinc = carry;
    for (int d = 1; d < seg; d <<= 1) {
        const uint32_t o = __shfl_up(cinc, d, seg);
        if (lseg >= d) cinc = lw_combine(o, cinc);
    }
    uint32_t cin = __shfl_up(cinc, 1, seg);
    if (lseg == 0) cin = 0;
    U prev = ((U)(cin & 0xFF) << (8 * (W - 3))) | ((U)((cin >> 8) & 0xFF) << (8 * (W - 2))) |
             ((U)((cin >> 16) & 0xFF) << (8 * (W - 1)));
    const T mT = (T)med;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const U cm = lk[k] == 0 ? (U)0 : (~(U)0 << (8 * (W - lk[k])));
        const U t = e[k] | (prev & cm);
        v[k] = from_bits<T>(t << shift) + mT;
        prev = t;
    }
    Store4<T>::store(out, e0, n, v);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

DecompressLayout decompress_layout(uint64_t n, int w, int bs_hint, uint64_t stream_bytes)
{
    (void)w;
    (void)stream_bytes;
    DecompressLayout L{};
    const uint64_t bs = bs_hint >= 4 ? (uint64_t)bs_hint : 4;
    L.nb_max = ceil_div(n, bs);
    L.ws_hdr = 0;
    L.ws_info = 256;
    L.ws_bytes = L.ws_info + 8 * (L.nb_max + 1);
    return L;
}

int decompress_launch(const uint8_t *d_stream, uint64_t stream_bytes, uint64_t n, int w,
                      void *d_out, uint8_t *d_ws, uint64_t ws_bytes, int exact, int32_t *d_status,
                      hipStream_t st)
{
    if (ws_bytes < 256 + 8) return -1;
    DecodeHeader *h = reinterpret_cast<DecodeHeader *>(d_ws);
    uint64_t *info = reinterpret_cast<uint64_t *>(d_ws + 256);
    const uint64_t cap = (ws_bytes - 256) / 8;
    hipLaunchKernelGGL(k_locate_serial, dim3(1), dim3(64), 0, st, d_stream, stream_bytes, n, w, h,
                       info, cap, d_status);
    const uint64_t lanes = ceil_div(n, 4);
    const uint64_t blocks = ceil_div(lanes, 256);
    if (blocks == 0) return 0;
    if (w == 4)
        hipLaunchKernelGGL(k_decode<float>, dim3((unsigned)blocks), dim3(256), 0, st, d_stream, n,
                           h, info, d_status, static_cast<float *>(d_out), exact);
    else
        hipLaunchKernelGGL(k_decode<double>, dim3((unsigned)blocks), dim3(256), 0, st, d_stream, n,
                           h, info, d_status, static_cast<double *>(d_out), exact);
    return 0;
}

}  // namespace szxamd
