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
//   A  k_ix_spec    per chunk (one wave): a speculative chain from the lowest
//                   candidate entry (the first record start at or after the
//                   chunk start lies less than one record size R into the
//                   chunk) whose chain gets through the chunk.  Chains from
//                   different record boundaries merge after a few records,
//                   so it almost always joins the true chain.
//   B  k_ix_verify  per chunk (one wave): from the entry the chunk before
//                   predicts (its speculative exit), walk until the chain
//                   meets the speculative one; record starts = walked ones +
//                   the speculative ones from there.  A chunk whose exit
//                   differs from its speculative exit is flagged.
//   C  k_ix_repair  (only when a chunk was flagged) one wave redoes, in chunk
//                   order, the chunks whose entries were wrong.
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
#include <stdlib.h>

#include "szx_common.hpp"
#include "szx_index.hpp"
#include "szx_internal.hpp"
#include "szx_wave.hpp"

namespace szxamd {

constexpr uint32_t IX_C = IX_CHUNK;        // payload bytes per chunk
constexpr uint16_t IX_INV = 0xFFFF;        // invalid chain / position
constexpr int IX_PAD = 96;                 // staged bytes past the chunk (5 + NL <= 69)

// Per lead-code byte c: the sum over its four 2-bit codes of min(code, 1),
// of min(code, 2) and of the code itself -- what a record with rb = 1, 2 or
// >= 3 subtracts from 4 rb per lead byte -- packed at bits 0, 10 and 21 of a
// u32.  A record's NL <= 64 lead bytes sum to at most 256 / 512 / 768 per
// field, below 2^10 / 2^11 / 2^11, so the fields of a sum never carry.
__device__ inline uint32_t lead_counts(uint32_t c)
{
    const uint32_t lo = c & 0x55u, hi = (c >> 1) & 0x55u;
    const uint32_t f1 = __builtin_popcount(lo | hi);
    const uint32_t f3 = __builtin_popcount(lo) + 2 * __builtin_popcount(hi);
    const uint32_t f2 = f3 - __builtin_popcount(lo & hi);
    return f1 | (f2 << 10) | (f3 << 21);
}

constexpr uint32_t IX_MAXR = IX_C / 6 + 2;                      // record starts per chunk, any bs
constexpr uint32_t IXW_WORDS = (16 + IX_C + IX_PAD) / 4 + 2;     // one wave's staged chunk

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

// One step of the record chain inside a chunk, for the whole wave: the
// position after the record at p (wave-uniform, < IX_C), or IX_INV where no
// record can start (the reqLength byte out of range, the record's lead codes
// past the bytes present, or lead codes subtracting more than the block
// holds).  byte(i) reads chunk byte i; lane l < NL takes lead-code byte l, the
// packed counts are summed inside each 16-lane row (DPP) and the row sums
// read out as scalars.
template <typename Byte>
__device__ inline uint32_t chain_step(const Byte &byte, uint32_t p, uint32_t avail, uint32_t nl, uint32_t bsz,
                                      uint32_t wmax, int lane)
{
    if (p + 5 + nl > avail) return IX_INV;
    const uint32_t rl = byte(p);
    uint32_t c = (uint32_t)lane < nl ? lead_counts(byte(p + 5 + (uint32_t)lane)) : 0u;
    c += partner<1>(c);
    c += partner<2>(c);
    c += partner<4>(c);
    c += partner<8>(c);
    uint32_t d = __builtin_amdgcn_readlane(c, 0);
    if (nl > 16) d += __builtin_amdgcn_readlane(c, 16);
    if (nl > 32) d += __builtin_amdgcn_readlane(c, 32) + __builtin_amdgcn_readlane(c, 48);
    const uint32_t rlu = __builtin_amdgcn_readfirstlane(rl);
    if (rlu == 0 || rlu > wmax) return IX_INV;
    const uint32_t rb = (rlu + 7) >> 3;
    const uint32_t sub = rb >= 3 ? d >> 21 : rb == 2 ? (d >> 10) & 0x7FFu : d & 0x3FFu;
    // nonzero unused codes of a partial lead byte (a damaged stream) could
    // subtract more than the block holds: no record there
    if (sub > rb * bsz) return IX_INV;
    return p + 5 + nl + rb * bsz - sub;
}

struct IxChunk {
    uint64_t c, g0;
    uint32_t avail;
    bool has;    // the chunk starts inside the payload
    bool ends;   // the payload ends inside the chunk
};

__device__ inline IxChunk ix_chunk(const IxParams &x, uint64_t c)
{
    IxChunk k;
    const uint64_t P = x.len - x.h->q0;   // payload bytes present in the buffer
    k.c = c;
    k.g0 = c * IX_C;
    k.has = k.g0 < P;
    k.ends = P - k.g0 <= IX_C;
    k.avail = k.has ? (uint32_t)(P - k.g0 < IX_C + IX_PAD ? P - k.g0 : IX_C + IX_PAD) : 0u;
    return k;
}

// Walk from `entry` to the end of the chunk, keeping every position (the
// record starts; a position whose record is invalid is kept and ends the
// walk); `stop(p)` ends the walk early at a position other knowledge covers.
// Returns the positions kept; *last is the position the walk ended at (>=
// IX_C past the chunk, IX_INV after an invalid record, or the stop position).
template <typename Byte, typename Stop>
__device__ inline uint32_t chain_walk(const Byte &byte, uint32_t entry, const IxChunk &k,
                                      const IxParams &x, uint16_t *keep, uint32_t cap, const Stop &stop,
                                      uint32_t *last, int lane)
{
    const uint32_t nl = (uint32_t)x.nl, bsz = (uint32_t)x.bs, wmax = 8u * (uint32_t)x.w;
    uint32_t p = entry, n = 0;
    while (p < IX_C) {
        if (stop(p)) break;
        if (lane == 0 && n < cap) keep[n] = (uint16_t)p;
        n++;
        p = chain_step(byte, p, k.avail, nl, bsz, wmax, lane);
    }
    *last = p;
    return n < cap ? n : cap;
}

// A: speculative chains.  One wave per chunk walks from the lowest candidate
// entry (a position below rmax with a plausible reqLength byte) whose chain
// reaches the end of the chunk, keeping its record starts.  Chains from
// different record boundaries merge after a few records, so this chain
// almost always joins the true one inside the chunk; B checks it.
// spec[c] = start | exit << 16 (IX_INV halves when no chain gets through).
// The chunk where the payload ends is not speculated (its chains end inside
// it): B walks it from its entry.
__global__ __launch_bounds__(256) void k_ix_spec(IxParams x)
{
    __shared__ uint32_t bw[4][IXW_WORDS];
    __shared__ uint32_t pfx[4][128];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) *x.nfail = 0;   // counted by B
    if (*x.status != 0) return;   // workgroup-uniform: written by earlier kernels only
    const uint64_t c = blockIdx.x * 4ull + wv;
    const bool live = c < x.nch;
    const IxChunk k = ix_chunk(x, live ? c : 0);
    uint32_t base = 0;
    if (live && k.has) base = ix_stage_wave(bw[wv], x.s + x.h->q0 + k.g0, k.avail, x.s, x.s + x.len, lane);
    __syncthreads();
    if (!live) return;
    const uint8_t *b = reinterpret_cast<const uint8_t *>(bw[wv]) + base;
    const auto byte = [b](uint32_t i) -> uint32_t { return b[i]; };
    const uint32_t nl = (uint32_t)x.nl, bsz = (uint32_t)x.bs, wmax = 8u * (uint32_t)x.w, rmax = x.rmax;
    uint16_t *pos = x.pos + c * x.maxr;
    uint32_t start = IX_INV, ex = IX_INV, n = 0;
    const auto never = [](uint32_t) { return false; };
    for (uint32_t q0 = 0; k.has && !k.ends && !x.no_spec && q0 < rmax && start == IX_INV; q0 += 64) {
        // Candidates q0 + lane, kept only if their record and the next
        // record start look valid (or the record leaves the chunk): the
        // records' lead-count window sums come from an inclusive prefix over
        // the 64 + NL bytes from q0 + 5.  Bytes past `avail` may be stale;
        // only candidates whose windows lie before it use them.
        const uint32_t q = q0 + (uint32_t)lane;
        const uint32_t ca = lead_counts(b[q0 + 5 + lane]);
        const uint32_t cb = (uint32_t)lane < nl ? lead_counts(b[q0 + 69 + lane]) : 0u;
        const uint32_t sa = seg_scan_incl<64>(ca, lane), sb = seg_scan_incl<64>(cb, lane);
        pfx[wv][lane] = sa;
        pfx[wv][64 + lane] = __builtin_amdgcn_readlane(sa, 63) + sb;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t rl = b[q];
        bool ok = q < rmax && q + 5 + nl <= k.avail && rl != 0 && rl <= wmax;
        if (ok) {
            const uint32_t d = pfx[wv][lane + nl - 1] - (lane ? pfx[wv][lane - 1] : 0u);
            const uint32_t rb = (rl + 7) >> 3;
            const uint32_t sub = rb >= 3 ? d >> 21 : rb == 2 ? (d >> 10) & 0x7FFu : d & 0x3FFu;
            const uint32_t nx = q + 5 + nl + rb * bsz - sub;
            ok = sub <= rb * bsz;
            if (ok && nx < IX_C) {
                ok = nx + 5 + nl <= k.avail;
                if (ok) {
                    const uint32_t r2 = b[nx];
                    ok = r2 != 0 && r2 <= wmax;
                }
            }
        }
        unsigned long long m = __ballot(ok);
        __builtin_amdgcn_wave_barrier();   // pfx is rewritten by the next batch
        while (m) {
            const uint32_t e = q0 + (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            uint32_t last;
            n = chain_walk(byte, e, k, x, pos, x.maxr, never, &last, lane);
            if (last != IX_INV) {
                start = e;
                ex = last - IX_C;
                break;
            }
        }
    }
    if (lane == 0) {
        x.spec[c] = start | (ex << 16);
        x.cnt[c] = start == IX_INV ? 0u : n;
    }
}

// Position p of the chunk's speculative chain (sorted starts in lst[0..n)):
// its index, found by the lanes, or -1.
__device__ inline int spec_rank(const uint16_t *lst, uint32_t n, uint32_t p, int lane)
{
    int r = -1;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const uint32_t i = i0 + (uint32_t)lane;
        const unsigned long long m = __ballot(i < n && lst[i] == p);
        if (m) {
            r = (int)(i0 + __builtin_ctzll(m));
            break;
        }
    }
    return r;
}

// B: true record starts.  One wave per chunk walks from its entry -- the
// speculative exit of the chunk before it (0 for the first) -- reading the
// stream, until the chain meets the speculative chain (almost always within
// a few records); the record starts are the walked ones followed by the
// speculative ones from the meeting point.  tex[c] = the chunk's exit for
// that entry, with bit 31 set and nfail counted when it differs from the
// speculative exit (then the entries of the chunks after it are wrong, and C
// redoes them).
__global__ __launch_bounds__(256) void k_ix_verify(IxParams x)
{
    __shared__ uint32_t bits[4][IX_C / 32];
    __shared__ uint16_t lst[4][IX_MAXR], pre[4][IX_MAXR];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t i = lane; i < IX_C / 32; i += 64) bits[wv][i] = 0;
    __syncthreads();
    if (*x.status != 0) return;
    const uint64_t c = blockIdx.x * 4ull + wv;
    if (c >= x.nch) return;
    const IxChunk k = ix_chunk(x, c);
    const uint32_t sp = x.spec[c];
    const uint32_t sex = sp >> 16;
    const uint32_t n = x.cnt[c];
    uint16_t *pos = x.pos + c * x.maxr;
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t q = pos[i];
        lst[wv][i] = (uint16_t)q;
        atomicOr(&bits[wv][q >> 5], 1u << (q & 31));
    }
    const uint32_t entry = c == 0 ? 0u : x.spec[c - 1] >> 16;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint8_t *g = x.s + x.h->q0 + k.g0;
    const auto byte = [g](uint32_t i) -> uint32_t { return g[i]; };
    const uint32_t *bw = bits[wv];
    const auto on_spec = [bw](uint32_t p) { return ((bw[p >> 5] >> (p & 31)) & 1u) != 0; };
    uint32_t last = IX_INV, m = 0;
    if (k.has && entry != IX_INV) m = chain_walk(byte, entry, k, x, pre[wv], IX_MAXR, on_spec, &last, lane);
    // the walked starts, then the speculative ones from the meeting point
    uint32_t from = n, ex;
    const int r = last < IX_C ? spec_rank(lst[wv], n, last, lane) : -1;
    if (r >= 0) {   // met the speculative chain at `last`
        from = (uint32_t)r;
        ex = sex;
    } else {
        ex = (last == IX_INV || last < IX_C) ? IX_INV : last - IX_C;
    }
    const uint32_t total = m + (n - from) < x.maxr ? m + (n - from) : x.maxr;
    for (uint32_t i = lane; i < total; i += 64) pos[i] = i < m ? pre[wv][i] : lst[wv][from + i - m];
    if (lane == 0) {
        x.cnt[c] = total;
        const bool bad = ex != sex;
        x.tex[c] = ex | (bad ? 0x80000000u : 0u);
        if (bad) atomicAdd(x.nfail, 1u);
    }
}

// C: repair after speculative exits that were wrong (rare: a chunk whose
// speculative chain never met the true one).  One wave, in chunk order: from
// the first chunk whose exit differed, each following chunk is walked in full
// from the true exit of the one before, until a chunk's true exit equals its
// speculative one (its successors' entries were right); then the next chunk
// whose exit differed, and so on.
__global__ __launch_bounds__(64) void k_ix_repair(IxParams x)
{
    __shared__ uint16_t keep[IX_MAXR];
    const int lane = threadIdx.x;
    if (*x.status != 0 || *x.nfail == 0) return;
    const auto never = [](uint32_t) { return false; };
    uint64_t c = 0;
    for (;;) {
        // next chunk whose exit differed from its speculative exit
        uint64_t f = x.nch;
        for (uint64_t i0 = c; i0 < x.nch; i0 += 64) {
            const uint64_t i = i0 + (uint64_t)lane;
            const unsigned long long m = __ballot(i < x.nch && (x.tex[i] & 0x80000000u));
            if (m) {
                f = i0 + __builtin_ctzll(m);
                break;
            }
        }
        if (f + 1 >= x.nch) return;
        uint32_t t = x.tex[f] & 0xFFFFu;   // true exit of f (its entry was right)
        c = f + 1;
        for (; c < x.nch; c++) {
            if (t == (x.spec[c - 1] >> 16)) break;   // entries right again from here
            const IxChunk k = ix_chunk(x, c);
            const uint8_t *g = x.s + x.h->q0 + k.g0;
            const auto byte = [g](uint32_t i) -> uint32_t { return g[i]; };
            uint32_t last = IX_INV, n = 0;
            if (k.has && t != IX_INV) n = chain_walk(byte, t, k, x, keep, IX_MAXR, never, &last, lane);
            const uint32_t cap = n < x.maxr ? n : x.maxr;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            uint16_t *pos = x.pos + c * x.maxr;
            for (uint32_t i = lane; i < cap; i += 64) pos[i] = keep[i];
            const uint32_t tn = last == IX_INV ? IX_INV : last - IX_C;
            if (lane == 0) {
                x.cnt[c] = cap;
                x.tex[c] = tn;
            }
            __builtin_amdgcn_wave_barrier();
            t = tn;
        }
        if (c >= x.nch) return;
    }
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

// E: record starts (payload offsets) in record order.  One wave per chunk.
__global__ __launch_bounds__(256) void k_ix_scatter(IxParams x)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t c = blockIdx.x * 4ull + wv;
    if (c >= x.nch || *x.status != 0) return;
    const uint64_t base = x.cbase[c] + x.csum[c >> 10];
    const uint32_t k = x.cnt[c];
    const uint64_t K = x.h->nb - x.h->nconst;
    for (uint32_t i = lane; i < k; i += 64)
        if (base + i < K) x.rec[base + i] = c * IX_C + x.pos[c * x.maxr + i];
}

// Bitmap byte j of a block range [b0, b1) with b0 a multiple of 8: the bits of
// blocks at or past b1 cleared (the bitmap is MSB first).
__device__ inline uint32_t bitmap_byte(const IxParams &x, const DecodeHeader &h, uint64_t j, uint64_t b1)
{
    uint32_t v = x.s[h.qbit + j];
    if (8 * j + 8 > b1) v &= (0xFF00u >> (b1 - 8 * j)) & 0xFFu;
    return v;
}

// Non-constant blocks per index tile (popcount of its bitmap bytes: a tile
// holds a multiple of 8 blocks, or is one block).
__device__ inline uint32_t tile_count(const IxParams &x, const DecodeHeader &h, uint64_t t)
{
    const uint64_t b0 = t * x.tile_blocks, b1 = b0 + x.tile_blocks < h.nb ? b0 + x.tile_blocks : h.nb;
    uint32_t c = 0;
    if (x.tile_blocks % 8 == 0) {
        for (uint64_t j = b0 / 8; 8 * j < b1; j++) c += __builtin_popcount(bitmap_byte(x, h, j, b1));
    } else {
        for (uint64_t b = b0; b < b1; b++) c += (x.s[h.qbit + (b >> 3)] >> (7 - (b & 7))) & 1;
    }
    return c;
}

// Non-constant blocks before each tile, in two steps like the scans above:
// per 1024 tiles the exclusive scan of their counts (tbase) and the block
// sums (tsum, then scanned by k_scan_tops); readers add tsum[t >> 10].
__global__ __launch_bounds__(1024) void k_ix_tile_scan(IxParams x)
{
    __shared__ uint64_t s[1024];
    const uint64_t t = blockIdx.x * 1024ull + threadIdx.x;
    if (*x.status != 0) return;   // workgroup-uniform
    const uint64_t v = t < x.ntiles ? tile_count(x, *x.h, t) : 0;
    const uint64_t e = block_excl_scan_1024(v, s);
    if (t < x.ntiles) x.tbase[t] = e;
    if (threadIdx.x == 1023) x.tsum[blockIdx.x] = s[1023];
}

__device__ inline uint64_t tile_base(const IxParams &x, uint64_t t) { return x.tbase[t] + x.tsum[t >> 10]; }

// F: the last record's true size and the payload end; checks the record
// count.  One wave: the bitmap is searched from its end 64 bytes at a time,
// and the last record's lead codes are summed one byte per lane.
__global__ void k_ix_tail(IxParams x)
{
    const int lane = threadIdx.x;
    if (*x.status != 0) return;
    const DecodeHeader &h = *x.h;
    const uint64_t K = h.nb - h.nconst;
    const uint64_t P = x.len - h.q0;
    int32_t st = 0;
    uint64_t pend = 0;
    if (*x.found < K) st = -30;
    if (st == 0 && K > 0) {
        // the last non-constant block (one exists: K > 0)
        const uint64_t nbyte = (h.nb + 7) / 8;
        uint64_t b = 0;
        for (uint64_t top = nbyte; top > 0; top = top > 64 ? top - 64 : 0) {
            uint32_t v = 0;
            if ((uint64_t)lane < top) {
                const uint64_t j = top - 1 - lane;
                v = x.s[h.qbit + j];
                if (8 * j + 8 > h.nb) v &= (0xFF00u >> (h.nb - 8 * j)) & 0xFFu;
            }
            const unsigned long long m = __ballot(v != 0);
            if (m) {
                const int l = __builtin_ctzll(m);   // the lowest lane holds the highest byte
                const uint32_t vb = __shfl(v, l, 64);
                b = 8 * (top - 1 - l) + 7 - __builtin_ctz(vb);
                break;
            }
        }
        const uint64_t nel = (b + 1) * (uint64_t)x.bs <= x.n ? (uint64_t)x.bs : x.n - b * (uint64_t)x.bs;
        const uint64_t nl = (nel + 3) / 4;
        const uint64_t q = x.rec[K - 1];
        if (q + 5 + nl > P) {
            st = -31;
        } else {
            const uint8_t *r = x.s + h.q0 + q;
            const uint32_t rl = r[0];
            if (rl == 0 || rl > 8u * x.w) {
                st = -32;
            } else {
                const uint32_t rb = (rl + 7) >> 3;
                uint32_t add = 0;
                for (uint64_t l = lane; l < nl; l += 64) {
                    const uint32_t by = r[5 + l];
                    for (uint32_t k = 0; k < 4 && 4 * l + k < nel; k++) {
                        const uint32_t ld = (by >> (6 - 2 * k)) & 3;
                        if (ld < rb) add += rb - ld;
                    }
                }
#pragma unroll
                for (int m = 1; m < 64; m <<= 1) add += __shfl_xor(add, m, 64);
                const uint64_t sz = 5 + nl + add;
                if (q + sz > P) st = -33;
                else pend = q + sz;
            }
        }
    }
    if (lane == 0) {
        if (st) *x.status = st;
        else *x.pend = pend;
    }
}

// G: tile pairs and per-block entries.  One wave per tile; the non-constant
// blocks before each block of the tile come from a wave scan of the
// popcounts of the tile's bitmap bytes (at most 128: 1024 blocks).
__global__ __launch_bounds__(256) void k_ix_tiles(IxParams x)
{
    __shared__ uint32_t s_byte[4][128], s_pre[4][128];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t t = blockIdx.x * 4ull + wv;
    if (t >= x.ntiles || *x.status != 0) return;
    const DecodeHeader &h = *x.h;
    const uint64_t K = h.nb - h.nconst;
    const uint64_t b0 = t * x.tile_blocks;
    const uint64_t nbt = b0 + x.tile_blocks <= h.nb ? x.tile_blocks : h.nb - b0;
    const uint64_t b1 = b0 + nbt;
    const uint32_t nbytes = (uint32_t)((nbt + 7) / 8);
    const uint64_t ord0 = tile_base(x, t);
    {
        // two bytes per lane, inclusive wave scan of their popcounts
        const uint32_t j0 = 2 * (uint32_t)lane, j1 = j0 + 1;
        const uint32_t v0 = j0 < nbytes ? bitmap_byte(x, h, b0 / 8 + j0, b1) : 0u;
        const uint32_t v1 = j1 < nbytes ? bitmap_byte(x, h, b0 / 8 + j1, b1) : 0u;
        const uint32_t c0 = __builtin_popcount(v0), c1 = __builtin_popcount(v1);
        const uint32_t inc = seg_scan_incl<64>(c0 + c1, lane);
        const uint32_t ex = inc - c0 - c1;
        s_byte[wv][j0] = v0;
        s_byte[wv][j1] = v1;
        s_pre[wv][j0] = ex;
        s_pre[wv][j1] = ex + c0;
    }
    const uint64_t off = ord0 < K ? x.rec[ord0] : *x.pend;
    if (lane == 0) {
        x.index[2 * t] = off;
        x.index[2 * t + 1] = b0 - ord0;
        if (t == x.ntiles - 1) {
            x.index[2 * x.ntiles] = *x.pend;
            x.index[2 * x.ntiles + 1] = h.nconst;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint64_t i = lane; i < nbt; i += 64) {
        const uint32_t j = (uint32_t)(i >> 3), k = (uint32_t)(i & 7);
        const uint32_t by = s_byte[wv][j];
        const uint32_t before = s_pre[wv][j] + (k ? (uint32_t)__builtin_popcount(by >> (8 - k)) : 0u);
        const bool nc = (by >> (7 - k)) & 1;
        uint64_t v;
        if (nc) {
            const uint64_t ord = ord0 + before;
            v = ord < K ? x.rec[ord] - off : 0xFFFF;
        } else {
            v = i - before;
        }
        x.bidx[b0 + i] = (uint16_t)(v > 0xFFFF ? 0xFFFF : v);
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
    const uint64_t ord = tile_base(x, b);
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
    L.o_spec = off;   off += al(4 * L.nch);
    L.o_tex = off;    off += al(4 * L.nch);
    L.o_cnt = off;    off += al(4 * L.nch);
    L.o_cbase = off;  off += al(8 * L.nch);
    L.o_pos = off;    off += al(2 * L.nch * L.maxr);
    L.o_rec = off;    off += al(8 * (nb ? nb : 1));
    L.o_tbase = off;  off += al(8 * L.ntiles);
    L.o_tsum = off;   off += al(8 * (L.ntiles / 1024 + 1));
    L.o_csum = off;   off += al(8 * (L.nch / 1024 + 1));
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
    x.spec = reinterpret_cast<uint32_t *>(ws + L.o_spec);
    x.tex = reinterpret_cast<uint32_t *>(ws + L.o_tex);
    x.cnt = reinterpret_cast<uint32_t *>(ws + L.o_cnt);
    x.cbase = reinterpret_cast<uint64_t *>(ws + L.o_cbase);
    x.pos = reinterpret_cast<uint16_t *>(ws + L.o_pos);
    x.rec = reinterpret_cast<uint64_t *>(ws + L.o_rec);
    x.tbase = reinterpret_cast<uint64_t *>(ws + L.o_tbase);
    x.tsum = reinterpret_cast<uint64_t *>(ws + L.o_tsum);
    x.csum = reinterpret_cast<uint64_t *>(ws + L.o_csum);
    x.found = reinterpret_cast<uint64_t *>(ws + L.o_misc);
    x.pend = reinterpret_cast<uint64_t *>(ws + L.o_misc + 8);
    x.nfail = reinterpret_cast<uint32_t *>(ws + L.o_misc + 24);
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
    const uint64_t nblk = (L.ntiles + 1023) / 1024;
    hipLaunchKernelGGL(k_ix_tile_scan, dim3((unsigned)nblk), dim3(1024), 0, st, x);
    hipLaunchKernelGGL(k_scan_tops, dim3(1), dim3(1024), 0, st, x.tsum, nblk, scratch);
    if (chain_tail) hipLaunchKernelGGL(k_ix_tail, dim3(1), dim3(64), 0, st, x);
    if (tiled)
        hipLaunchKernelGGL(k_ix_tiles, dim3((unsigned)((L.ntiles + 3) / 4)), dim3(256), 0, st, x);
    else
        hipLaunchKernelGGL(k_ix_pairs, dim3((unsigned)((L.ntiles + 256) / 256)), dim3(256), 0, st, x);
}

void index_build_launch(const uint8_t *s, uint64_t len, uint64_t n, int w, int bs, const DecodeHeader *h,
                        uint64_t *index, uint16_t *bidx, uint8_t *ws, int32_t *status, hipStream_t st)
{
    const IxLayout L = ix_layout(n, w, bs, len);
    IxParams x = ix_params(s, len, n, w, bs, h, index, bidx, ws, status, L);
    // SZX_AMD_INDEX_NO_SPECULATION=1 (for tests): no speculative chains, so
    // every chunk after the first goes through the in-order repair
    const char *ns = getenv("SZX_AMD_INDEX_NO_SPECULATION");
    x.no_spec = ns && *ns && *ns != '0';
    const unsigned nch = (unsigned)L.nch;
    hipLaunchKernelGGL(k_ix_spec, dim3((nch + 3) / 4), dim3(256), 0, st, x);
    hipLaunchKernelGGL(k_ix_verify, dim3((nch + 3) / 4), dim3(256), 0, st, x);
    hipLaunchKernelGGL(k_ix_repair, dim3(1), dim3(64), 0, st, x);
    // records before each chunk: cbase + csum[c >> 10] (read by the scatter)
    const uint64_t nblk = (L.nch + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nblk), dim3(1024), 0, st, x.cnt, L.nch, x.cbase, x.csum);
    hipLaunchKernelGGL(k_scan_tops, dim3(1), dim3(1024), 0, st, x.csum, nblk, x.found);
    hipLaunchKernelGGL(k_ix_scatter, dim3((nch + 3) / 4), dim3(256), 0, st, x);
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
