// szx_index.hip -- the record index of a fastMode-2 stream, built in
// parallel from the stream alone (for streams that arrive without the index
// this library's compressor writes).
//
// The payload is a chain: a record starting at byte p has size
//   size(p) = 5 + NL + rb * bs - sum over the codes of its NL lead-code bytes
//             of min(code, rb)
// with rb = ceil(reqLength / 8), reqLength = byte p, NL = ceil(bs / 4) (every
// record but the stream's last belongs to a full block; the unused codes of
// a partly filled last lead byte are zero, szx_TypeManager.cc:246-252, and
// subtract nothing).  Following the chain from byte 0 is inherently serial,
// so it is cut into chunks of IX_C payload bytes:
//
//   A  k_ix_map     per chunk: size(p) for every byte p, then chains walked
//                   from every candidate entry (merging chains stop where they
//                   meet) give, for every possible entry offset e (the first
//                   record start at or after the chunk start lies less than one
//                   record size R into the chunk), the exit offset into the
//                   next chunk.  Chains from different entries merge as soon
//                   as they meet a common record boundary, so almost every
//                   chunk maps all valid entries to one exit ("converged") and
//                   the next chunk's entry is known without this one's.
//   B  k_ix_entry   entries: a chunk after a converged chunk starts at its
//                   exit; a run of unconverged chunks is followed through their
//                   stored exit tables by one thread.
//   C  k_ix_walk    per chunk (one wave): walk the chain from the entry, one
//                   step per record, keep the record starts, count them.
//   D  scans        records before each chunk; non-constant blocks before
//                   each compressor tile.
//   E  k_ix_scatter record starts in record order.
//   F  k_ix_tail    the last record's true size (its block may be partial) and
//                   the payload end.
//   G  k_ix_tiles   the tile pairs and the per-block entries of the index;
//      k_ix_pairs   for the block sizes the tiled kernels do not take, the
//                   per-block pairs instead (tiles of one block).
//
// Every read stays inside [stream, stream + len); a chain that leaves the
// stream or meets an impossible reqLength is marked invalid.  A stream whose
// chain does not produce exactly the records its bitmap needs is rejected
// (status -30..-33); the decoder re-checks every record against the index.
#include "szx_common.hpp"
#include "szx_index.hpp"
#include "szx_internal.hpp"

namespace szxamd {

constexpr uint32_t IX_C = IX_CHUNK;        // payload bytes per chunk
constexpr uint16_t IX_INV = 0xFFFF;        // invalid chain / position
constexpr int IX_PAD = 96;                 // staged bytes past the chunk (5 + NL <= 69)
constexpr uint32_t IX_PRE_PER = 17;                            // prefix entries per thread
constexpr uint32_t IX_PRE = 256 * IX_PRE_PER;                  // >= IX_C + 5 + 64 + 1
// staged words: 16-byte alignment head + every byte the prefix sums read
constexpr uint32_t IX_WORDS = (16 + IX_PRE) / 4 + 2;
static_assert(IX_PRE >= IX_C + IX_PAD, "prefix covers the staged chunk");
constexpr int IX_PER = IX_C / 256;                             // chunk positions per thread

__device__ inline uint32_t ld_u32_unaligned(const uint32_t *w, uint32_t byte)
{
    const uint32_t a = w[byte >> 2], b = w[(byte >> 2) + 1];
    return __builtin_amdgcn_alignbyte(b, a, byte & 3);
}

// Stage bytes [src, src + cnt) into LDS words, zero-padded, at byte offset
// (src & 15): aligned 16-byte chunks lying wholly inside the stream
// [lo, hi) are copied with one load (their bytes outside the range are
// stream bytes no step reads), the rest byte by byte; returns the offset.
__device__ inline uint32_t ix_stage(uint32_t *bw, const uint8_t *src, uint32_t cnt, uint32_t words,
                                    const uint8_t *lo, const uint8_t *hi)
{
    uint8_t *b = reinterpret_cast<uint8_t *>(bw);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) bw[i] = 0;
    __syncthreads();
    const uintptr_t a = (uintptr_t)src & ~(uintptr_t)15;
    const uint32_t head = (uint32_t)((uintptr_t)src - a);
    const uint32_t nchunk = (head + cnt + 15) / 16;
    for (uint32_t c = threadIdx.x; c < nchunk; c += blockDim.x) {
        const uint8_t *g = reinterpret_cast<const uint8_t *>(a + 16 * (uintptr_t)c);
        if (g >= lo && g + 16 <= hi)
            *reinterpret_cast<uint4 *>(b + 16 * c) = *reinterpret_cast<const uint4 *>(g);
    }
    // the first and last chunks may be partial: one byte per thread
    if (threadIdx.x < 32 && nchunk) {
        const uint32_t c = threadIdx.x < 16 ? 0 : nchunk - 1;
        const uint32_t k = threadIdx.x & 15;
        const uint8_t *g = reinterpret_cast<const uint8_t *>(a + 16 * (uintptr_t)c) + k;
        if (g >= lo && g < hi) b[16 * c + k] = *g;
    }
    __syncthreads();
    return head;
}

// Per lead-code byte c, packed in 16-bit fields: the sum over its four 2-bit
// codes of min(code, 1), of min(code, 2) and of the code itself -- what a
// record with rb = 1, 2 or >= 3 subtracts from 4 rb per lead byte.
__device__ inline uint64_t lead_counts(uint32_t c)
{
    const uint32_t lo = c & 0x55u, hi = (c >> 1) & 0x55u;
    const uint32_t f1 = __builtin_popcount(lo | hi);
    const uint32_t f3 = __builtin_popcount(lo) + 2 * __builtin_popcount(hi);
    const uint32_t f2 = f3 - __builtin_popcount(lo & hi);
    return (uint64_t)f1 | ((uint64_t)f2 << 16) | ((uint64_t)f3 << 32);
}

// Chain step for every chunk byte: J[p] = p + size(p), or IX_INV; positions
// at or past IX_C absorb.  The NL-byte window sums of lead counts come from
// an exclusive prefix sum over the staged bytes (pre, IX_PRE entries), so a
// step costs two reads.  avail: payload bytes present from the chunk start.
__device__ inline void ix_steps(const uint32_t *bw, uint32_t base, uint64_t *pre, uint16_t *J,
                                uint32_t dom, uint32_t avail, int nl, int bs, int w)
{
    const uint8_t *b = reinterpret_cast<const uint8_t *>(bw) + base;
    __shared__ uint64_t s_part[256];
    // per-thread runs of IX_PRE_PER bytes
    const uint32_t q0 = threadIdx.x * IX_PRE_PER;
    uint64_t cnt[IX_PRE_PER];   // loads first, then arithmetic (LDS stores below may alias)
#pragma unroll
    for (uint32_t i = 0; i < IX_PRE_PER; i++) cnt[i] = b[q0 + i];
    uint64_t run = 0;
#pragma unroll
    for (uint32_t i = 0; i < IX_PRE_PER; i++) {
        cnt[i] = lead_counts((uint32_t)cnt[i]);
        run += cnt[i];
    }
    s_part[threadIdx.x] = run;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const uint64_t v = threadIdx.x >= (unsigned)d ? s_part[threadIdx.x - d] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    run = s_part[threadIdx.x] - run;
#pragma unroll
    for (uint32_t i = 0; i < IX_PRE_PER; i++) {
        pre[q0 + i] = run;
        run += cnt[i];
    }
    __syncthreads();
    // chunk positions: IX_PER per thread, all loads before any store
    uint32_t rl[IX_PER];
    uint64_t lo[IX_PER], hi[IX_PER];
#pragma unroll
    for (int i = 0; i < IX_PER; i++) {
        const uint32_t p = threadIdx.x + 256 * i;
        rl[i] = b[p];
        lo[i] = pre[p + 5];
        hi[i] = pre[p + 5 + nl];
    }
#pragma unroll
    for (int i = 0; i < IX_PER; i++) {
        const uint32_t p = threadIdx.x + 256 * i;
        uint16_t j = IX_INV;
        if (p + 5 + (uint32_t)nl <= avail && rl[i] != 0 && rl[i] <= 8u * w) {
            const uint32_t rb = (rl[i] + 7) >> 3;
            const uint32_t sh = rb >= 3 ? 32 : (rb == 2 ? 16 : 0);
            const uint32_t sub = (uint32_t)((hi[i] - lo[i]) >> sh) & 0xFFFFu;
            // nonzero unused codes of a partial lead byte (a damaged stream)
            // could subtract more than the block holds: no record there
            if (sub <= rb * (uint32_t)bs) j = (uint16_t)(p + 5 + (uint32_t)nl + rb * (uint32_t)bs - sub);
        }
        J[p] = j;
    }
    for (uint32_t p = IX_C + threadIdx.x; p < dom; p += blockDim.x) J[p] = (uint16_t)p;
}

// A: exit map of every chunk.  After the step table, one chain per candidate
// entry e < rmax is walked, each position claimed by the first chain to reach
// it (an LDS compare-and-swap): a chain that meets a claimed position stops
// and links to its owner, whose chain continues from there exactly as its own
// would.  Chains merge after a few records, so the walk costs about one step
// per distinct record start; links are then resolved by pointer jumping (link
// positions increase along a link chain, so there are no cycles).
constexpr uint32_t IX_NONE = 0xFFFFFFFFu;
constexpr uint32_t IX_RES_INV = 0x40000000u;
constexpr uint32_t IX_RES_LINK = 0x80000000u;

__global__ __launch_bounds__(256) void k_ix_map(IxParams x)
{
    __shared__ uint32_t bw[IX_WORDS];
    __shared__ uint16_t J0[IX_C];
    __shared__ union {
        uint64_t pre[IX_PRE];
        struct {
            uint32_t owner[IX_C];
            uint32_t res[IX_RMAX];
        } w;
    } u;
    __shared__ unsigned s_min, s_max, s_more;
    const uint64_t c = blockIdx.x;
    const DecodeHeader &h = *x.h;
    if (*x.status != 0) return;
    const uint64_t P = x.len - h.q0;       // payload bytes present in the buffer
    const uint64_t g0 = c * IX_C;
    if (g0 >= P) {
        if (threadIdx.x == 0) {
            x.conv[c] = 1;
            x.exits[c] = IX_INV;
        }
        return;
    }
    const uint32_t avail = (uint32_t)(P - g0 < IX_C + IX_PAD ? P - g0 : IX_C + IX_PAD);
    const uint32_t base = ix_stage(bw, x.s + h.q0 + g0, avail, IX_WORDS, x.s, x.s + x.len);
    ix_steps(bw, base, u.pre, J0, IX_C, avail, x.nl, x.bs, x.w);
    __syncthreads();   // the prefix sums die here: their space holds the claims
    for (uint32_t p = threadIdx.x; p < IX_C; p += blockDim.x) u.w.owner[p] = IX_NONE;
    if (threadIdx.x == 0) {
        s_min = 0xFFFFFFFFu;
        s_max = 0;
    }
    __syncthreads();
    const uint32_t rmax = x.rmax;
    for (uint32_t e = threadIdx.x; e < rmax; e += blockDim.x) {
        uint32_t p = e, r;
        for (;;) {
            if (p >= IX_C) {
                r = p - IX_C;
                break;
            }
            const uint32_t o = atomicCAS(&u.w.owner[p], IX_NONE, e);
            if (o != IX_NONE) {
                r = IX_RES_LINK | o;
                break;
            }
            const uint16_t j = J0[p];
            if (j == IX_INV) {
                r = IX_RES_INV;
                break;
            }
            p = j;
        }
        u.w.res[e] = r;
    }
    __syncthreads();
    for (int round = 0; round < 16; round++) {   // link depth < 2^16
        if (threadIdx.x == 0) s_more = 0;
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < rmax; e += blockDim.x) {
            const uint32_t r = u.w.res[e];
            if (r & IX_RES_LINK) {
                const uint32_t t = u.w.res[r & 0xFFFFu];
                u.w.res[e] = t;
                if (t & IX_RES_LINK) s_more = 1;
            }
        }
        __syncthreads();
        const bool more = s_more != 0;
        __syncthreads();   // every lane has read the flag before it is cleared
        if (!more) break;
    }
    // entries e < rmax: exit offset into the next chunk (wave min / max, then
    // one LDS atomic per wave)
    unsigned mn = 0xFFFFFFFFu, mx = 0;
    for (uint32_t e = threadIdx.x; e < rmax; e += blockDim.x) {
        const uint32_t r = u.w.res[e];
        if (!(r & (IX_RES_INV | IX_RES_LINK))) {
            mn = r < mn ? r : mn;
            mx = r > mx ? r : mx;
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const unsigned a = __shfl_xor(mn, m, 64), b = __shfl_xor(mx, m, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if ((threadIdx.x & 63) == 0 && mn != 0xFFFFFFFFu) {
        atomicMin(&s_min, mn);
        atomicMax(&s_max, mx);
    }
    __syncthreads();
    const bool conv = s_min == s_max || s_min == 0xFFFFFFFFu;
    if (threadIdx.x == 0) {
        x.conv[c] = conv ? 1 : 0;
        x.exits[c] = s_min == 0xFFFFFFFFu ? IX_INV : (uint16_t)s_min;
    }
    if (!conv)
        for (uint32_t e = threadIdx.x; e < rmax; e += blockDim.x) {
            const uint32_t r = u.w.res[e];
            x.tables[c * rmax + e] = (r & (IX_RES_INV | IX_RES_LINK)) ? IX_INV : (uint16_t)r;
        }
}

// B: entry offset of every chunk (one thread per run of unconverged chunks).
__global__ void k_ix_entry(IxParams x)
{
    const uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (c >= x.nch || *x.status != 0) return;
    if (c != 0 && !x.conv[c - 1]) return;   // inside a run: its start thread writes it
    uint16_t e = c == 0 ? 0 : x.exits[c - 1];
    for (uint64_t k = c; k < x.nch; k++) {
        x.entry[k] = e;
        if (x.conv[k]) break;
        e = e == IX_INV || e >= x.rmax ? IX_INV : x.tables[k * x.rmax + e];
    }
}

// Stage bytes [src, src + cnt) into one wave's LDS words at byte offset
// (src & 15): aligned 16-byte chunks lying wholly inside the stream [lo, hi)
// with one load, the edge chunks byte by byte (bytes outside the stream read
// as zero); returns the offset.  Words past the staged bytes are not cleared.
__device__ inline uint32_t ix_stage_wave(uint32_t *bw, const uint8_t *src, uint32_t cnt, const uint8_t *lo,
                                         const uint8_t *hi, int lane)
{
    uint8_t *b = reinterpret_cast<uint8_t *>(bw);
    const uintptr_t a = (uintptr_t)src & ~(uintptr_t)15;
    const uint32_t head = (uint32_t)((uintptr_t)src - a);
    const uint32_t nchunk = (head + cnt + 15) / 16;
    for (uint32_t c = lane; c < nchunk; c += 64) {
        const uint8_t *g = reinterpret_cast<const uint8_t *>(a + 16 * (uintptr_t)c);
        if (g >= lo && g + 16 <= hi) {
            *reinterpret_cast<uint4 *>(b + 16 * c) = *reinterpret_cast<const uint4 *>(g);
        } else {
            for (int k = 0; k < 16; k++) b[16 * c + k] = (g + k >= lo && g + k < hi) ? g[k] : 0;
        }
    }
    return head;
}

constexpr uint32_t IXW_WORDS = (16 + IX_C + IX_PAD) / 4 + 2;   // one wave's staged chunk

// C: record starts inside every chunk, from its entry.  One wave per chunk
// follows the chain: a record's size comes from its reqLength byte and its
// lead-code bytes (one byte per lane, a wave sum of the same per-byte counts
// the map kernel's prefix sums hold), so the chain costs one step per record.
// Marks the entry and every chain position below IX_C, stopping where the
// map kernel's step table would be invalid.
__global__ __launch_bounds__(256) void k_ix_walk(IxParams x)
{
    __shared__ uint32_t bw[4][IXW_WORDS];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t c = blockIdx.x * 4ull + wv;
    if (*x.status != 0) return;   // workgroup-uniform: written by earlier kernels only
    const DecodeHeader &h = *x.h;
    const uint64_t P = x.len - h.q0;
    const uint64_t g0 = c * IX_C;
    const bool live = c < x.nch;
    const bool has = live && g0 < P;
    const uint32_t avail = has ? (uint32_t)(P - g0 < IX_C + IX_PAD ? P - g0 : IX_C + IX_PAD) : 0u;
    uint32_t base = 0;
    if (has) base = ix_stage_wave(bw[wv], x.s + h.q0 + g0, avail, x.s, x.s + x.len, lane);
    __syncthreads();
    if (!live) return;
    const uint8_t *b = reinterpret_cast<const uint8_t *>(bw[wv]) + base;
    const uint32_t nl = (uint32_t)x.nl, bsz = (uint32_t)x.bs, wmax = 8u * (uint32_t)x.w;
    uint32_t p = has ? x.entry[c] : IX_INV;
    uint32_t k = 0;
    uint16_t *pos = x.pos + c * x.maxr;
    if (p != IX_INV) {
        while (p < IX_C) {
            if (lane == 0 && k < x.maxr) pos[k] = (uint16_t)p;
            k++;
            const uint32_t rl = b[p];
            if (p + 5 + nl > avail || rl == 0 || rl > wmax) break;
            const uint32_t rb = (rl + 7) >> 3;
            const uint32_t sh = rb >= 3 ? 32 : (rb == 2 ? 16 : 0);
            uint32_t sub = (uint32_t)lane < nl ? (uint32_t)(lead_counts(b[p + 5 + lane]) >> sh) & 0xFFFFu : 0u;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) sub += __shfl_xor(sub, d, 64);
            if (sub > rb * bsz) break;
            p = p + 5 + nl + rb * bsz - sub;
        }
    }
    if (lane == 0) x.cnt[c] = k < x.maxr ? k : x.maxr;
}

// Exclusive scan of n u32 values in three steps: per 1024-value block an
// LDS scan (block sums aside), one workgroup over the block sums, then the
// block offsets are added.  *total receives the sum.
__device__ inline uint64_t block_excl_scan_1024(uint64_t v, uint64_t *s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint64_t o = threadIdx.x >= (unsigned)d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += o;
        __syncthreads();
    }
    return s[threadIdx.x] - v;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(const uint32_t *in, uint64_t n, uint64_t *out,
                                                      uint64_t *bsum)
{
    __shared__ uint64_t s[1024];
    const uint64_t i = blockIdx.x * 1024ull + threadIdx.x;
    const uint64_t v = i < n ? in[i] : 0;
    const uint64_t e = block_excl_scan_1024(v, s);
    if (i < n) out[i] = e;
    if (threadIdx.x == 1023) bsum[blockIdx.x] = s[1023];
}

__global__ __launch_bounds__(1024) void k_scan_tops(uint64_t *bsum, uint64_t nblk, uint64_t *total)
{
    __shared__ uint64_t s[1024];
    const uint64_t per = (nblk + 1023) / 1024;
    const uint64_t a = threadIdx.x * per, b = a + per < nblk ? a + per : nblk;
    uint64_t sum = 0;
    for (uint64_t i = a; i < b; i++) sum += bsum[i];
    uint64_t run = block_excl_scan_1024(sum, s);
    for (uint64_t i = a; i < b; i++) {
        const uint64_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023 && total) *total = s[1023];
}

__global__ void k_scan_add(uint64_t *out, uint64_t n, const uint64_t *bsum)
{
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / 1024];
}

static void scan_u32(const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *bsum, uint64_t *total,
                     hipStream_t st)
{
    const uint64_t nblk = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nblk), dim3(1024), 0, st, in, n, out, bsum);
    hipLaunchKernelGGL(k_scan_tops, dim3(1), dim3(1024), 0, st, bsum, nblk, total);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n, bsum);
}

// E: record starts (payload offsets) in record order.
__global__ void k_ix_scatter(IxParams x)
{
    const uint64_t c = blockIdx.x;
    if (*x.status != 0) return;
    const uint64_t base = x.cbase[c];
    const uint32_t k = x.cnt[c];
    const uint64_t K = x.h->nb - x.h->nconst;
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x)
        if (base + i < K) x.rec[base + i] = c * IX_C + x.pos[c * x.maxr + i];
}

// Non-constant blocks per compressor tile (bitmap popcount).
__global__ void k_ix_tile_counts(IxParams x)
{
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= x.ntiles || *x.status != 0) return;
    const DecodeHeader &h = *x.h;
    const uint64_t b0 = t * x.tile_blocks, b1 = b0 + x.tile_blocks < h.nb ? b0 + x.tile_blocks : h.nb;
    uint32_t c = 0;
    for (uint64_t b = b0; b < b1; b++) c += (x.s[h.qbit + (b >> 3)] >> (7 - (b & 7))) & 1;
    x.tcnt[t] = c;
}

// F: the last record's true size and the payload end; checks the record count.
__global__ void k_ix_tail(IxParams x)
{
    if (threadIdx.x != 0 || *x.status != 0) return;
    const DecodeHeader &h = *x.h;
    const uint64_t K = h.nb - h.nconst;
    const uint64_t P = x.len - h.q0;
    if (*x.found < K) {
        *x.status = -30;
        return;
    }
    if (K == 0) {
        *x.pend = 0;
        return;
    }
    // the last non-constant block
    uint64_t b = h.nb;
    while (b > 0) {
        b--;
        if ((x.s[h.qbit + (b >> 3)] >> (7 - (b & 7))) & 1) break;
    }
    const uint64_t nel = (b + 1) * (uint64_t)x.bs <= x.n ? (uint64_t)x.bs : x.n - b * (uint64_t)x.bs;
    const uint64_t nl = (nel + 3) / 4;
    const uint64_t q = x.rec[K - 1];
    if (q + 5 + nl > P) {
        *x.status = -31;
        return;
    }
    const uint8_t *r = x.s + h.q0 + q;
    const uint32_t rl = r[0];
    if (rl == 0 || rl > 8u * x.w) {
        *x.status = -32;
        return;
    }
    const uint32_t rb = (rl + 7) >> 3;
    uint64_t sz = 5 + nl;
    for (uint64_t i = 0; i < nel; i++) {
        const uint32_t ld = (r[5 + (i >> 2)] >> (6 - 2 * (i & 3))) & 3;
        if (ld < rb) sz += rb - ld;
    }
    if (q + sz > P) {
        *x.status = -33;
        return;
    }
    *x.pend = q + sz;
}

// G: tile pairs and per-block entries.  One workgroup per tile.
__global__ __launch_bounds__(256) void k_ix_tiles(IxParams x)
{
    __shared__ uint32_t s_nc[1024 + 1];   // non-constant blocks before each block of the tile
    const uint64_t t = blockIdx.x;
    if (*x.status != 0) return;
    const DecodeHeader &h = *x.h;
    const uint64_t K = h.nb - h.nconst;
    const uint64_t b0 = t * x.tile_blocks;
    const uint64_t nbt = b0 + x.tile_blocks <= h.nb ? x.tile_blocks : h.nb - b0;
    const uint64_t ord0 = x.tbase[t];
    if (threadIdx.x == 0) {
        uint32_t a = 0;
        for (uint64_t i = 0; i < nbt; i++) {
            s_nc[i] = a;
            a += (x.s[h.qbit + ((b0 + i) >> 3)] >> (7 - ((b0 + i) & 7))) & 1;
        }
        s_nc[nbt] = a;
        const uint64_t off = ord0 < K ? x.rec[ord0] : *x.pend;
        x.index[2 * t] = off;
        x.index[2 * t + 1] = b0 - ord0;
        if (t == x.ntiles - 1) {
            x.index[2 * x.ntiles] = *x.pend;
            x.index[2 * x.ntiles + 1] = h.nconst;
        }
    }
    __syncthreads();
    const uint64_t off = ord0 < K ? x.rec[ord0] : *x.pend;
    for (uint64_t i = threadIdx.x; i < nbt; i += blockDim.x) {
        const uint64_t b = b0 + i;
        const bool nc = (x.s[h.qbit + (b >> 3)] >> (7 - (b & 7))) & 1;
        uint64_t v;
        if (nc) {
            const uint64_t ord = ord0 + s_nc[i];
            v = ord < K ? x.rec[ord] - off : 0xFFFF;
        } else {
            v = i - s_nc[i];
        }
        x.bidx[b] = (uint16_t)(v > 0xFFFF ? 0xFFFF : v);
    }
}

// G (per-block index): {record offset, constant blocks before} for every
// block, then the closing pair.  tbase holds the non-constant blocks before
// each block (tiles of one block).
__global__ void k_ix_pairs(IxParams x)
{
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (*x.status != 0) return;
    const DecodeHeader &h = *x.h;
    if (b > h.nb) return;
    if (b == h.nb) {
        x.index[2 * b] = *x.pend;
        x.index[2 * b + 1] = h.nconst;
        return;
    }
    const uint64_t K = h.nb - h.nconst;
    const uint64_t ord = x.tbase[b];
    x.index[2 * b] = ord < K ? x.rec[ord] : *x.pend;
    x.index[2 * b + 1] = b - ord;
}

// Random-access container (szxd_float.cc:313-476): the record sizes are in
// the stream (u16, native order, from byte 20), so the record starts are
// their exclusive scan.  Sizes as u32, zero past the last record.
__global__ void k_ra_sizes(IxParams x, uint32_t *sz, uint64_t nmax)
{
    const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (k >= nmax) return;
    const DecodeHeader &h = *x.h;
    const uint64_t K = *x.status == 0 ? h.nb - h.nconst : 0;   // 20 + 2K <= len (header check)
    sz[k] = k < K ? (uint32_t)x.s[20 + 2 * k] | ((uint32_t)x.s[21 + 2 * k] << 8) : 0u;
}

// The records found are the size table's; the payload must fit the stream.
__global__ void k_ra_tail(IxParams x)
{
    if (threadIdx.x != 0 || *x.status != 0) return;
    const DecodeHeader &h = *x.h;
    *x.found = h.nb - h.nconst;
    if (*x.pend > x.len - h.q0) *x.status = -33;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

IxLayout ix_layout(uint64_t n, int w, int bs, uint64_t len)
{
    IxLayout L{};
    const uint64_t nb = (n + bs - 1) / bs;
    L.nch = (len + IX_C - 1) / IX_C;
    if (!L.nch) L.nch = 1;
    L.nl = (bs + 3) / 4;
    L.rmax = 5 + (uint32_t)L.nl + (uint32_t)bs * (uint32_t)w;
    L.maxr = IX_C / (5 + (uint32_t)L.nl) + 2;
    L.ntiles = bs_tiled(bs) ? index_tiles(n, w) : (nb ? nb : 1);
    auto al = [](uint64_t v) { return (v + 255) & ~255ull; };
    uint64_t off = 0;
    L.o_exits = off;  off += al(2 * L.nch);
    L.o_conv = off;   off += al(L.nch);
    L.o_entry = off;  off += al(2 * L.nch);
    L.o_cnt = off;    off += al(4 * L.nch);
    L.o_cbase = off;  off += al(8 * L.nch);
    L.o_tables = off; off += al(2 * L.nch * L.rmax);
    L.o_pos = off;    off += al(2 * L.nch * L.maxr);
    L.o_rec = off;    off += al(8 * (nb ? nb : 1));
    L.o_tcnt = off;   off += al(4 * L.ntiles);
    L.o_tbase = off;  off += al(8 * L.ntiles);
    L.o_misc = off;   off += 256;   // found, pend, scratch total
    L.o_osz = off;    off += al(4 * (nb ? nb : 1));   // random-access container sizes
    uint64_t mx = L.nch > L.ntiles ? L.nch : L.ntiles;
    mx = mx > nb ? mx : nb;
    L.o_bsum = off;   off += al(8 * (mx / 1024 + 1));
    L.bytes = off;
    return L;
}

static IxParams ix_params(const uint8_t *s, uint64_t len, uint64_t n, int w, int bs, const DecodeHeader *h,
                          uint64_t *index, uint16_t *bidx, uint8_t *ws, int32_t *status, const IxLayout &L)
{
    IxParams x{};
    x.s = s;
    x.len = len;
    x.n = n;
    x.bs = bs;
    x.w = w;
    x.nl = (int)L.nl;
    x.rmax = L.rmax;
    x.maxr = L.maxr;
    x.nch = L.nch;
    x.ntiles = L.ntiles;
    x.tile_blocks = bs_tiled(bs) ? tile_values(w) / (uint64_t)bs : 1;
    x.h = h;
    x.status = status;
    x.exits = reinterpret_cast<uint16_t *>(ws + L.o_exits);
    x.conv = ws + L.o_conv;
    x.entry = reinterpret_cast<uint16_t *>(ws + L.o_entry);
    x.cnt = reinterpret_cast<uint32_t *>(ws + L.o_cnt);
    x.cbase = reinterpret_cast<uint64_t *>(ws + L.o_cbase);
    x.tables = reinterpret_cast<uint16_t *>(ws + L.o_tables);
    x.pos = reinterpret_cast<uint16_t *>(ws + L.o_pos);
    x.rec = reinterpret_cast<uint64_t *>(ws + L.o_rec);
    x.tcnt = reinterpret_cast<uint32_t *>(ws + L.o_tcnt);
    x.tbase = reinterpret_cast<uint64_t *>(ws + L.o_tbase);
    x.found = reinterpret_cast<uint64_t *>(ws + L.o_misc);
    x.pend = reinterpret_cast<uint64_t *>(ws + L.o_misc + 8);
    x.index = index;
    x.bidx = bidx;
    return x;
}

// G: the index entries from the record starts (x.rec), the records found and
// the payload end.
static void index_tiles_launch(IxParams &x, const IxLayout &L, uint8_t *ws, bool tiled, bool chain_tail,
                               hipStream_t st)
{
    uint64_t *scratch = reinterpret_cast<uint64_t *>(ws + L.o_misc + 16);
    uint64_t *bsum = reinterpret_cast<uint64_t *>(ws + L.o_bsum);
    hipLaunchKernelGGL(k_ix_tile_counts, dim3((unsigned)((L.ntiles + 255) / 256)), dim3(256), 0, st, x);
    scan_u32(x.tcnt, L.ntiles, x.tbase, bsum, scratch, st);
    if (chain_tail) hipLaunchKernelGGL(k_ix_tail, dim3(1), dim3(64), 0, st, x);
    if (tiled)
        hipLaunchKernelGGL(k_ix_tiles, dim3((unsigned)L.ntiles), dim3(256), 0, st, x);
    else
        hipLaunchKernelGGL(k_ix_pairs, dim3((unsigned)((L.ntiles + 256) / 256)), dim3(256), 0, st, x);
}

void index_build_launch(const uint8_t *s, uint64_t len, uint64_t n, int w, int bs, const DecodeHeader *h,
                        uint64_t *index, uint16_t *bidx, uint8_t *ws, int32_t *status, hipStream_t st)
{
    const IxLayout L = ix_layout(n, w, bs, len);
    IxParams x = ix_params(s, len, n, w, bs, h, index, bidx, ws, status, L);
    uint64_t *bsum = reinterpret_cast<uint64_t *>(ws + L.o_bsum);
    const unsigned nch = (unsigned)L.nch;
    hipLaunchKernelGGL(k_ix_map, dim3(nch), dim3(256), 0, st, x);
    hipLaunchKernelGGL(k_ix_entry, dim3((nch + 255) / 256), dim3(256), 0, st, x);
    hipLaunchKernelGGL(k_ix_walk, dim3((nch + 3) / 4), dim3(256), 0, st, x);
    scan_u32(x.cnt, L.nch, x.cbase, bsum, x.found, st);
    hipLaunchKernelGGL(k_ix_scatter, dim3(nch), dim3(256), 0, st, x);
    index_tiles_launch(x, L, ws, bs_tiled(bs), true, st);
}

void index_from_sizes_launch(const uint8_t *s, uint64_t len, uint64_t n, int w, int bs, const DecodeHeader *h,
                             uint64_t *index, uint16_t *bidx, uint8_t *ws, int32_t *status, hipStream_t st)
{
    const IxLayout L = ix_layout(n, w, bs, len);
    IxParams x = ix_params(s, len, n, w, bs, h, index, bidx, ws, status, L);
    uint64_t *bsum = reinterpret_cast<uint64_t *>(ws + L.o_bsum);
    uint32_t *sz = reinterpret_cast<uint32_t *>(ws + L.o_osz);
    const uint64_t nb = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_ra_sizes, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, x, sz, nb);
    scan_u32(sz, nb, x.rec, bsum, x.pend, st);
    hipLaunchKernelGGL(k_ra_tail, dim3(1), dim3(64), 0, st, x);
    index_tiles_launch(x, L, ws, bs_tiled(bs), false, st);
}

}  // namespace szxamd
