// szx_decompress.hip -- blockwise fast decompressor for gfx950.
//
// A fastMode-2 stream has no block index: the start of non-constant record k
// is the sum of the sizes of records 0..k-1, and a size is only known by
// parsing that record's reqLength byte and lead codes (szxd_float.cc:276-283).
// Decoding therefore works from a record index -- per compressor tile, the
// payload offset of the tile's first record and the number of constant
// blocks before it -- which this library's compressor emits as a by-product
// of its look-back scan, and which k_index_* rebuilds for any other stream.
// With the index every tile decodes independently:
//
//   1. stage the tile's records in LDS and walk their sizes (a few dozen
//      records, one wave);
//   2. decode every block in parallel, 4 values per lane, a block being an
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

// Header check: block size, counts and the bitmap/median region must fit.
struct DecodeHeader {
    uint64_t bs, nb, nbit, nconst, q0;
};

__global__ void k_check_header(const uint8_t *s, uint64_t len, uint64_t n, int bs_expect,
                               DecodeHeader *h, int32_t *status)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
    *status = rc;
}

// Record index of a foreign stream by the reference's serial walk
// (szxd_float.cc:276-283), one lane: per tile of `tile_blocks` blocks, the
// payload offset of its first record and the constant blocks before it.
__global__ void k_index_serial(const uint8_t *s, uint64_t len, uint64_t n, int w,
                               uint64_t tile_blocks, const DecodeHeader *h, uint64_t *index,
                               int32_t *status)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || *status != 0) return;
    const uint64_t bs = h->bs, nb = h->nb, q0 = h->q0;
    uint64_t q = q0, k0 = 0;
    int32_t rc = 0;
    for (uint64_t b = 0; b < nb; b++) {
        if (b % tile_blocks == 0) {
            index[2 * (b / tile_blocks)] = q - q0;
            index[2 * (b / tile_blocks) + 1] = k0;
        }
        if (!((s[HEADER_BYTES + (b >> 3)] >> (7 - (b & 7))) & 1)) {
            k0++;
            continue;
        }
        const uint64_t nel = (b + 1) * bs <= n ? bs : n - b * bs;
        const uint64_t nl = (nel + 3) >> 2;
        if (q + 5 + nl > len) { rc = -15; break; }
        const int rl = s[q];
        if (rl == 0 || rl > 8 * w) { rc = -16; break; }
        const int rb = (rl + 7) >> 3;
        uint64_t sz = 5 + nl;
        for (uint64_t i = 0; i < nel; i++) {
            const int ld = (s[q + 5 + (i >> 2)] >> (6 - 2 * (int)(i & 3))) & 3;
            if (ld < rb) sz += (uint64_t)(rb - ld);
        }
        if (q + sz > len) { rc = -17; break; }
        q += sz;
    }
    if (!rc && k0 != h->nconst) rc = -18;
    const uint64_t nt = (nb + tile_blocks - 1) / tile_blocks;
    index[2 * (nt ? nt : 1)] = q - q0;
    index[2 * (nt ? nt : 1) + 1] = k0;
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

// Decode one tile (the compressor's tile: TILE values, NBLK blocks).
//   1. stage the tile's record bytes [index[t], index[t+1]) in LDS;
//   2. wave 0 walks the tile's records (sizes from the lead codes, 64 lanes
//      summing one record's codes per step) -> record start per block;
//   3. every lane decodes 4 values per slab, a block being BS/4 lanes.
// Every record must lie inside the staged span and the walk must end exactly
// at the next tile's first record; otherwise the stream (or index) is
// rejected and nothing reads out of bounds.
template <typename T, int BS, int SLABS>
__global__ __launch_bounds__(256) void k_decode_tiles(const uint8_t *s, uint64_t len, uint64_t n,
                                                     const DecodeHeader *h, const uint64_t *index,
                                                     int32_t *status, T *out, int exact)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    constexpr int SEG = BS / 4;
    constexpr int SLAB = 1024;
    constexpr int TILE = SLAB * SLABS;
    constexpr int BPS = SLAB / BS;
    constexpr int NBLK = TILE / BS;
    constexpr int STAGE = TILE * W + NBLK * (5 + BS / 4);

    __shared__ uint8_t st[STAGE + 16];
    __shared__ uint32_t s_start[NBLK];   // record start, or constant rank within the tile
    __shared__ uint8_t s_state[NBLK];
    __shared__ int s_err;

    if (*status != 0) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int lseg = tid % SEG;
    const uint64_t tile = blockIdx.x;
    const uint64_t nb = h->nb, q0 = h->q0, nconst = h->nconst, nbit = h->nbit;
    const uint64_t gb0 = tile * NBLK;
    const uint64_t off = q0 + index[2 * tile];
    const uint64_t end = q0 + index[2 * tile + 2];
    const uint64_t cbase = index[2 * tile + 1];
    if (end < off || end > len || end - off > (uint64_t)STAGE || cbase > nconst) {
        if (tid == 0) atomicMin(status, -20);
        return;
    }
    const uint32_t tb = (uint32_t)(end - off);
    if (tid == 0) s_err = 0;
    for (uint32_t i = tid; i < tb; i += 256) st[i] = s[off + i];
    for (int lb = tid; lb < NBLK; lb += 256) {
        const uint64_t b = gb0 + lb;
        s_state[lb] = b < nb ? (s[HEADER_BYTES + (b >> 3)] >> (7 - (b & 7))) & 1 : 0;
    }
    __syncthreads();

    if (tid < 64) {
        uint32_t pos = 0, crank = 0;
        int err = 0;
        for (int lb = 0; lb < NBLK; lb++) {
            const uint64_t b = gb0 + lb;
            if (b >= nb) break;
            if (!s_state[lb]) {
                if (lane == 0) s_start[lb] = crank;
                crank++;
                continue;
            }
            const uint64_t rem = n - b * BS;
            const int nel = rem >= (uint64_t)BS ? BS : (int)rem;
            const int nlead = (nel + 3) >> 2;
            if (pos + 5 + (uint32_t)nlead > tb) { err = 1; break; }
            const int rl = st[pos];
            if (rl == 0 || rl > 8 * W) { err = 1; break; }
            const int rb = (rl + 7) >> 3;
            uint32_t part = 0;
            for (int j = lane; j < nlead; j += 64) {
                const uint32_t c = st[pos + 5 + j];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const int ld = (c >> (6 - 2 * k)) & 3;
                    if (4 * j + k < nel && ld < rb) part += (uint32_t)(rb - ld);
                }
            }
#pragma unroll
            for (int m = 1; m < 64; m <<= 1) part += __shfl_xor(part, m, 64);
            if (lane == 0) s_start[lb] = pos;
            pos += 5 + (uint32_t)nlead + part;
            if (pos > tb) { err = 1; break; }
        }
        if (!err && pos != tb) err = 1;
        if (!err && cbase + crank > nconst) err = 1;
        if (lane == 0 && err) s_err = 1;
    }
    __syncthreads();
    if (s_err) {
        if (tid == 0) atomicMin(status, -21);
        return;
    }

    const uint8_t *cmed = s + HEADER_BYTES + nbit;
#pragma unroll
    for (int u = 0; u < SLABS; u++) {
        const uint64_t e0 = tile * TILE + (uint64_t)u * SLAB + 4u * (uint64_t)tid;
        const int lb = u * BPS + tid / SEG;
        const uint64_t bstart = e0 - 4u * (uint64_t)lseg;
        if (bstart >= n) continue;   // whole segment
        T v[4];
        if (!s_state[lb]) {
            const float med = be_f32(cmed + 4 * (cbase + s_start[lb]));
#pragma unroll
            for (int k = 0; k < 4; k++) v[k] = (T)med;
            Store4<T>::store(out, e0, n, v);
            continue;
        }
        const uint8_t *rec = st + s_start[lb];
        const uint64_t rem = n - bstart;
        const int nel = rem >= (uint64_t)BS ? BS : (int)rem;
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
#pragma unroll
        for (int d = 1; d < SEG; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, SEG);
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
                        carry = (carry & ~(0xFFu << (8 * q))) |
                                ((uint32_t)((ek >> (8 * pos)) & 0xFF) << (8 * q));
                        carry |= 1u << (24 + q);
                    }
                }
            }
            e[k] = ek;
        }
        uint32_t cinc = carry;
#pragma unroll
        for (int d = 1; d < SEG; d <<= 1) {
            const uint32_t o = __shfl_up(cinc, d, SEG);
            if (lseg >= d) cinc = lw_combine(o, cinc);
        }
        uint32_t cin = __shfl_up(cinc, 1, SEG);
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
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

DecompressLayout decompress_layout(uint64_t n, int w, int bs, uint64_t stream_bytes)
{
    (void)bs;
    (void)stream_bytes;
    DecompressLayout L{};
    L.ws_hdr = 0;
    L.ws_index = 256;
    L.ws_bytes = L.ws_index + 8 * index_words(n, w);
    return L;
}

template <typename T, int BS>
static void launch_decode(const uint8_t *s, uint64_t len, uint64_t n, const DecodeHeader *h,
                          const uint64_t *index, int32_t *status, void *out, int exact,
                          uint64_t ntiles, hipStream_t st)
{
    constexpr int SLABS = (sizeof(T) == 4 ? TILE_VALUES_F32 : TILE_VALUES_F64) / 1024;
    hipLaunchKernelGGL((k_decode_tiles<T, BS, SLABS>), dim3((unsigned)ntiles), dim3(256), 0, st, s,
                       len, n, h, index, status, static_cast<T *>(out), exact);
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
                      int exact, int32_t *d_status, hipStream_t st)
{
    const DecompressLayout L = decompress_layout(n, w, bs, stream_bytes);
    if (ws_bytes < L.ws_bytes) return -1;
    DecodeHeader *h = reinterpret_cast<DecodeHeader *>(d_ws + L.ws_hdr);
    const uint64_t tv = tile_values(w);
    const uint64_t ntiles = ceil_div(n, tv);
    hipLaunchKernelGGL(k_check_header, dim3(1), dim3(64), 0, st, d_stream, stream_bytes, n, bs, h,
                       d_status);
    const uint64_t *index = d_index;
    if (!index) {
        uint64_t *built = reinterpret_cast<uint64_t *>(d_ws + L.ws_index);
        timing_mark_begin(KK_INDEX, st);
        hipLaunchKernelGGL(k_index_serial, dim3(1), dim3(64), 0, st, d_stream, stream_bytes, n, w,
                           tv / (uint64_t)bs, h, built, d_status);
        timing_mark_end(KK_INDEX, st);
        index = built;
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
