// szx_compress.hip -- single-pass blockwise fast compressor for gfx950.
//
// One launch of k_compress turns the whole input into the records of the
// fastMode-2 stream (szx_float.cc:1059-1139, szx_double.cc:1160-1238):
//
//   * a workgroup (256 lanes, 4 wave64s) owns a tile of SLABS x 1024 values;
//     each lane holds 4 consecutive values (one 128-bit load for fp32, two for
//     fp64), so a block of BS values is an aligned group of BS/4 lanes and
//     every per-block reduction / scan is a butterfly inside that group;
//   * block statistics (min/max -> radius, median, constant test) and the
//     XOR leading-byte encoding are computed in registers; the records of the
//     tile are staged in LDS;
//   * the tile's record bytes and constant-block count are published through
//     a decoupled look-back (tickets give tiles in launch order, so every
//     predecessor a tile waits on is already resident);
//   * records are written at payload + (bytes of all earlier records).
//
// The payload start does not depend on the number of constant blocks: the
// caller's buffer reserves the largest possible header+bitmap+median region
// in front of it, and k_assemble writes the header, bitmap and constant
// medians immediately before the payload once the totals are known.  The
// stream therefore starts at a data-dependent offset inside the output
// buffer, which is returned with the size.  This keeps HBM traffic at one read
// of the input plus one write of the stream.
#include "szx_common.hpp"
#include "szx_internal.hpp"

namespace szxamd {

struct CompressParams {
    const void *x;
    uint64_t n;
    float eb;             // absolute bound (ABS)
    const float *eb_dev;  // REL: bound computed on the device, else nullptr
    uint8_t *payload;     // first record byte
    uint8_t *bitmap;      // scratch bitmap, nbit bytes
    uint32_t *cmed;       // scratch constant medians, big-endian words
    TileState *tiles;
    unsigned *ticket;
    uint64_t *totals;     // [0] payload bytes, [1] constant blocks
    uint64_t *index;      // optional record index: per tile {payload offset, constants before}
    uint64_t ntiles;
    uint64_t nbit;
};

template <typename T> struct SlabLoad;
template <> struct SlabLoad<float> {
    __device__ static void load4(const float *x, uint64_t e0, uint64_t n, float v[4])
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
    __device__ static void load4(const double *x, uint64_t e0, uint64_t n, double v[4])
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

template <typename T, int BS, int SLABS>
__global__ __launch_bounds__(256) void k_compress(CompressParams p)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    constexpr int SEG = BS / 4;           // lanes per block
    constexpr int SLAB = 1024;            // values per slab (256 lanes x 4)
    constexpr int TILE = SLAB * SLABS;
    constexpr int BPS = SLAB / BS;        // blocks per slab
    constexpr int NBLK = TILE / BS;       // blocks per tile (multiple of 8)
    constexpr int STAGE = TILE * W + NBLK * (5 + BS / 4);
    static_assert(BS >= 4 && BS <= 256 && (BS & (BS - 1)) == 0, "power-of-two block size");
    static_assert(NBLK % 8 == 0, "tiles own whole bitmap bytes");
    static_assert(STAGE < (1 << 20), "record bytes fit the packed scan word");

    __shared__ uint32_t s_pack[NBLK];     // record bytes | constant<<20 -> exclusive prefix
    __shared__ uint8_t s_state[NBLK];     // 1 = non-constant (bitmap bit)
    __shared__ float s_med[NBLK];         // medians (constant blocks)
    __shared__ uint8_t s_stage[STAGE + 16];
    __shared__ unsigned long long s_excl[2];
    __shared__ uint32_t s_total;
    __shared__ unsigned s_tile;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int lseg = tid % SEG;
    if (tid == 0) s_tile = atomicAdd(p.ticket, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    const float eb = p.eb_dev ? *p.eb_dev : p.eb;
    const int reqexp = req_expo(eb);
    const uint64_t n = p.n;
    const T *x = static_cast<const T *>(p.x);

    U s[SLABS][4];
    uint32_t excl_b[SLABS];   // lane's first mid-byte within its record's mid-byte area
    uint32_t code_b[SLABS];   // lane's lead-code byte
    uint32_t mpack[SLABS];    // 4 x 8-bit mid-byte counts
    int rb_[SLABS], nlead_[SLABS], len_[SLABS];
    float hmed_[SLABS];
    bool nonconst_[SLABS];

#pragma unroll
    for (int u = 0; u < SLABS; u++) {
        const uint64_t e0 = tile * TILE + (uint64_t)u * SLAB + 4u * (uint64_t)tid;
        const uint64_t bstart = e0 - 4u * (uint64_t)lseg;
        const bool exists = bstart < n;
        const uint64_t rem = exists ? n - bstart : 0;
        const int nel = rem >= (uint64_t)BS ? BS : (int)rem;
        T v[4];
        SlabLoad<T>::load4(x, e0, n, v);
        bool valid[4];
#pragma unroll
        for (int k = 0; k < 4; k++) valid[k] = e0 + k < n;

        // min / max: NaNs never win a comparison, as in the reference's
        // sequential scan (szx_float.cc:327-333); a NaN first value poisons
        // both, exactly as it does there.
        T mn = pos_inf<T>(), mx = -pos_inf<T>();
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (valid[k]) {
                mn = v[k] < mn ? v[k] : mn;
                mx = v[k] > mx ? v[k] : mx;
            }
        }
#pragma unroll
        for (int m = 1; m < SEG; m <<= 1) {
            T o = __shfl_xor(mn, m, SEG);
            mn = o < mn ? o : mn;
            o = __shfl_xor(mx, m, SEG);
            mx = o > mx ? o : mx;
        }
        const T first = __shfl(v[0], 0, SEG);
        if (first != first) { mn = first; mx = first; }

        // radius, median, constant test (szx_float.cc:335-347,
        // szx_double.cc:990-1003): median and radius are stored as float.
        const T rad = (mx - mn) * (T)0.5;
        const T med = mn + rad;
        const float med_f = (float)med;
        const float rad_f = (float)rad;
        const bool nonconst = exists && !(rad <= (T)eb);

        int clamped;
        const int len = req_length<T>(rad_expo(rad_f), reqexp, clamped);
        const int rb = (len + 7) >> 3;
        const int shift = (8 - (len & 7)) & 7;
        const float hmed = clamped ? 0.0f : med_f;
        // fp32 subtracts the stored median; fp64 keeps the unclamped one
        // (szx_float.cc:60,90 vs szx_double.cc:38,78).
        const T msub = W == 4 ? (T)hmed : (T)med_f;

#pragma unroll
        for (int k = 0; k < 4; k++) s[u][k] = to_bits<T>(v[k] - msub) >> shift;
        U prev = __shfl_up(s[u][3], 1, SEG);
        if (lseg == 0) prev = 0;

        uint32_t code = 0, mp = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const U xr = s[u][k] ^ prev;
            int ld = 0;
            if ((xr >> (8 * (W - 3))) == 0) ld = 3;
            else if ((xr >> (8 * (W - 2))) == 0) ld = 2;
            else if ((xr >> (8 * (W - 1))) == 0) ld = 1;
            const int m = valid[k] && ld < rb ? rb - ld : 0;
            if (valid[k]) code |= (uint32_t)ld << (6 - 2 * k);
            mp |= (uint32_t)m << (8 * k);
            tot += m;
            prev = s[u][k];
        }
        // segment-inclusive scan of the lane's mid-byte count
        uint32_t inc = tot;
#pragma unroll
        for (int d = 1; d < SEG; d <<= 1) {
            uint32_t o = __shfl_up(inc, d, SEG);
            if (lseg >= d) inc += o;
        }
        const uint32_t midtot = __shfl(inc, SEG - 1, SEG);
        const int nlead = (nel + 3) >> 2;

        excl_b[u] = inc - tot;
        code_b[u] = code;
        mpack[u] = mp;
        rb_[u] = rb;
        nlead_[u] = nlead;
        len_[u] = len;
        hmed_[u] = hmed;
        nonconst_[u] = nonconst;

        if (lseg == 0) {
            const int lb = u * BPS + tid / SEG;
            const uint32_t recb = nonconst ? 5u + (uint32_t)nlead + midtot : 0u;
            const uint32_t isconst = exists && !nonconst ? 1u : 0u;
            s_pack[lb] = recb | (isconst << 20);
            s_state[lb] = nonconst ? 1 : 0;
            s_med[lb] = med_f;
        }
    }
    __syncthreads();

    if (tid < 64) {
        const uint32_t tot = wave0_exclusive_scan<NBLK>(s_pack, lane);
        if (lane == 0) s_total = tot;
    }
    __syncthreads();

    // Stage the tile's records in LDS (byte-addressed, tile-relative).
#pragma unroll
    for (int u = 0; u < SLABS; u++) {
        if (!nonconst_[u]) continue;
        const int lb = u * BPS + tid / SEG;
        const uint32_t base = s_pack[lb] & 0xFFFFFu;
        if (lseg == 0) {
            const uint32_t hb = bswap32(f32_bits(hmed_[u]));
            s_stage[base] = (uint8_t)len_[u];
            s_stage[base + 1] = (uint8_t)hb;
            s_stage[base + 2] = (uint8_t)(hb >> 8);
            s_stage[base + 3] = (uint8_t)(hb >> 16);
            s_stage[base + 4] = (uint8_t)(hb >> 24);
        }
        if (lseg < nlead_[u]) s_stage[base + 5 + lseg] = (uint8_t)code_b[u];
        uint32_t pos = base + 5 + (uint32_t)nlead_[u] + excl_b[u];
        const int lo = W - rb_[u];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int m = (mpack[u] >> (8 * k)) & 0xFF;
            for (int j = 0; j < m; j++) s_stage[pos++] = (uint8_t)(s[u][k] >> (8 * (lo + j)));
        }
    }

    // Decoupled look-back over (record bytes, constant blocks).
    if (tid == 0) {
        const uint32_t tw = s_total;
        const unsigned long long tbytes = tw & 0xFFFFFu;
        const unsigned long long tconst = tw >> 20;
        unsigned long long ebytes = 0, econst = 0;
        TileState *ts = p.tiles;
        if (tile == 0) {
            ts[0].cnt_inc = tconst;
            store_release(&ts[0].status, FLAG_INC | tbytes);
        } else {
            ts[tile].cnt_agg = tconst;
            store_release(&ts[tile].status, FLAG_AGG | tbytes);
            int64_t j = (int64_t)tile - 1;
            while (j >= 0) {
                const unsigned long long st = load_acquire(&ts[j].status);
                const unsigned long long f = st & ~VAL_MASK;
                if (f == 0) { __builtin_amdgcn_s_sleep(1); continue; }
                ebytes += st & VAL_MASK;
                if (f == FLAG_INC) { econst += ts[j].cnt_inc; break; }
                econst += ts[j].cnt_agg;
                j--;
            }
            ts[tile].cnt_inc = econst + tconst;
            store_release(&ts[tile].status, FLAG_INC | (ebytes + tbytes));
        }
        s_excl[0] = ebytes;
        s_excl[1] = econst;
        if (p.index) {
            p.index[2 * tile] = ebytes;
            p.index[2 * tile + 1] = econst;
        }
        if (tile == p.ntiles - 1) {
            p.totals[0] = ebytes + tbytes;
            p.totals[1] = econst + tconst;
            if (p.index) {
                p.index[2 * p.ntiles] = ebytes + tbytes;
                p.index[2 * p.ntiles + 1] = econst + tconst;
            }
        }
    }
    __syncthreads();

    const uint64_t ebytes = s_excl[0];
    const uint64_t econst = s_excl[1];
    const uint32_t tbytes = s_total & 0xFFFFFu;

    // Copy the staged records out: aligned dwords inside, bytes at the edges
    // (the neighbouring tiles own the other bytes of an edge dword).
    {
        uint8_t *dst = p.payload + ebytes;
        const uintptr_t d0 = (uintptr_t)dst;
        const uint32_t head = (uint32_t)((4 - (d0 & 3)) & 3);
        const uint32_t h = head < tbytes ? head : tbytes;
        if (tid < (int)h) dst[tid] = s_stage[tid];
        const uint32_t nw = (tbytes - h) >> 2;
        for (uint32_t w = tid; w < nw; w += 256) {
            const uint32_t o = h + 4 * w;
            const uint32_t v = (uint32_t)s_stage[o] | ((uint32_t)s_stage[o + 1] << 8) |
                               ((uint32_t)s_stage[o + 2] << 16) | ((uint32_t)s_stage[o + 3] << 24);
            *reinterpret_cast<uint32_t *>(dst + o) = v;
        }
        const uint32_t tail0 = h + 4 * nw;
        if (tid < (int)(tbytes - tail0)) dst[tail0 + tid] = s_stage[tail0 + tid];
    }

    // Constant medians in block order: the exclusive constant count of a
    // block sits in the high bits of its scanned word, so a block is
    // constant exactly when the next block's count is one higher.
    for (int lb = tid; lb < NBLK; lb += 256) {
        const uint32_t cexcl = s_pack[lb] >> 20;
        const uint32_t cnext = lb + 1 < NBLK ? (s_pack[lb + 1] >> 20) : (s_total >> 20);
        if (cnext != cexcl)
            p.cmed[econst + cexcl] = bswap32(f32_bits(s_med[lb]));
    }
    for (int i = tid; i < NBLK / 8; i += 256) {
        const uint64_t gbyte = tile * (NBLK / 8) + i;
        if (gbyte < p.nbit) {
            uint32_t b = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) b |= (uint32_t)s_state[8 * i + k] << (7 - k);
            p.bitmap[gbyte] = (uint8_t)b;
        }
    }
}

// Zero the look-back state, ticket and totals (a kernel, not a memset, so the
// whole sequence can be captured in a HIP graph).
__global__ void k_zero_u64(unsigned long long *p, uint64_t nwords)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nwords;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Header, bitmap and constant medians in front of the payload.
struct AssembleParams {
    uint8_t *out;
    uint64_t region;      // bytes reserved in front of the payload
    const uint8_t *bitmap;
    const uint32_t *cmed;
    const uint64_t *totals;
    uint64_t nbit;
    int bs;
    const float *eb_dev;
    float eb;
    uint64_t *result;     // [0] stream offset, [1] stream size, [2] bound bits
};

__global__ void k_assemble(AssembleParams a)
{
    const uint64_t nconst = a.totals[1];
    const uint64_t pbytes = a.totals[0];
    const uint64_t front = HEADER_BYTES + a.nbit + 4 * nconst;
    const uint64_t start = a.region - front;
    uint8_t *s = a.out + start;
    const uint64_t gid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
    if (gid == 0) {
        s[0] = 1; s[1] = 1; s[2] = 1; s[3] = 0;        // szx_float.cc:1091-1094
        s[4] = (uint8_t)a.bs;                           // (u8)blockSize, :1095
        for (int k = 0; k < 8; k++) s[5 + k] = (uint8_t)(nconst >> (56 - 8 * k));
        a.result[0] = start;
        a.result[1] = front + pbytes;
        a.result[2] = f32_bits(a.eb_dev ? *a.eb_dev : a.eb);
    }
    for (uint64_t i = gid; i < a.nbit; i += nthr) s[HEADER_BYTES + i] = a.bitmap[i];
    const uint8_t *cm = reinterpret_cast<const uint8_t *>(a.cmed);
    for (uint64_t i = gid; i < 4 * nconst; i += nthr) s[HEADER_BYTES + a.nbit + i] = cm[i];
}

// ---------------------------------------------------------------------------
// REL: global value range and bound (szx.cc:249-253, 303-308).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_range_partial(const T *x, uint64_t n, T *part)
{
    T mn = pos_inf<T>(), mx = -pos_inf<T>();
    const uint64_t nq = n / 4;
    for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nq;
         q += (uint64_t)gridDim.x * blockDim.x) {
        T v[4];
        SlabLoad<T>::load4(x, 4 * q, n, v);
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
    __shared__ T sm[2][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[0][w] = mn; sm[1][w] = mx; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            mn = sm[0][k] < mn ? sm[0][k] : mn;
            mx = sm[1][k] > mx ? sm[1][k] : mx;
        }
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}

template <typename T>
__global__ void k_range_final(const T *x, const T *part, int nparts, float rel, float *eb_out)
{
    if (threadIdx.x != 0) return;
    T mn = pos_inf<T>(), mx = -pos_inf<T>();
    for (int i = 0; i < nparts; i++) {
        mn = part[2 * i] < mn ? part[2 * i] : mn;
        mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
    }
    const T first = x[0];
    if (first != first) { mn = first; mx = first; }
    const T range = mx - mn;
    // fp32: fl32(range * rel); fp64: (float)(range * (double)rel)
    *eb_out = (float)(range * (T)rel);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static constexpr int kSlabsF32 = TILE_VALUES_F32 / 1024;
static constexpr int kSlabsF64 = TILE_VALUES_F64 / 1024;
static constexpr int kRangeParts = 1024;

CompressLayout compress_layout(uint64_t n, int w, int bs)
{
    CompressLayout L{};
    L.nb = ceil_div(n, (uint64_t)bs);
    L.nbit = ceil_div(L.nb, 8);
    L.region = (HEADER_BYTES + L.nbit + 4 * L.nb + 15) & ~(uint64_t)15;
    const uint64_t rec_max = 5 + ceil_div((uint64_t)bs, 4) + (uint64_t)bs * w;
    L.payload_max = L.nb * rec_max;
    L.capacity = L.region + L.payload_max + 16;
    L.ntiles = ceil_div(n, (uint64_t)tile_values(w));
    if (L.ntiles == 0) L.ntiles = 1;
    // workspace: tiles | ticket+totals (4 words) | eb (1 word) | range parts | bitmap | cmed
    uint64_t off = 0;
    L.ws_tiles = off;   off += L.ntiles * sizeof(TileState);
    L.ws_misc = off;    off += 8 * sizeof(uint64_t);
    L.ws_parts = off;   off += 2 * kRangeParts * sizeof(double);
    L.ws_bitmap = off;  off += (L.nbit + 15) & ~(uint64_t)15;
    L.ws_cmed = off;    off += (4 * L.nb + 15) & ~(uint64_t)15;
    L.ws_bytes = off;
    return L;
}

template <typename T, int BS>
static void launch_main(const CompressParams &p, uint64_t ntiles, hipStream_t st)
{
    constexpr int SLABS = sizeof(T) == 4 ? kSlabsF32 : kSlabsF64;
    hipLaunchKernelGGL((k_compress<T, BS, SLABS>), dim3((unsigned)ntiles), dim3(256), 0, st, p);
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

int compress_launch(const void *d_in, uint64_t n, int w, int rel_mode, float abs_eb, float rel,
                    int bs, uint8_t *d_out, uint8_t *d_ws, uint64_t *d_result, uint64_t *d_index,
                    hipStream_t st)
{
    const CompressLayout L = compress_layout(n, w, bs);
    unsigned long long *misc = reinterpret_cast<unsigned long long *>(d_ws + L.ws_misc);
    const uint64_t zero_words = (L.ws_misc + 8 * sizeof(uint64_t)) / 8;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((zero_words + 255) / 256 < 1024 ? (zero_words + 255) / 256 : 1024)),
                       dim3(256), 0, st, reinterpret_cast<unsigned long long *>(d_ws), zero_words);
    float *eb_dev = nullptr;
    if (rel_mode) {
        eb_dev = reinterpret_cast<float *>(misc + 4);
        if (w == 4) {
            float *part = reinterpret_cast<float *>(d_ws + L.ws_parts);
            hipLaunchKernelGGL(k_range_partial<float>, dim3(kRangeParts), dim3(256), 0, st,
                               static_cast<const float *>(d_in), n, part);
            hipLaunchKernelGGL(k_range_final<float>, dim3(1), dim3(64), 0, st,
                               static_cast<const float *>(d_in), part, kRangeParts, rel, eb_dev);
        } else {
            double *part = reinterpret_cast<double *>(d_ws + L.ws_parts);
            hipLaunchKernelGGL(k_range_partial<double>, dim3(kRangeParts), dim3(256), 0, st,
                               static_cast<const double *>(d_in), n, part);
            hipLaunchKernelGGL(k_range_final<double>, dim3(1), dim3(64), 0, st,
                               static_cast<const double *>(d_in), part, kRangeParts, rel, eb_dev);
        }
    }
    CompressParams p{};
    p.x = d_in;
    p.n = n;
    p.eb = abs_eb;
    p.eb_dev = eb_dev;
    p.payload = d_out + L.region;
    p.bitmap = d_ws + L.ws_bitmap;
    p.cmed = reinterpret_cast<uint32_t *>(d_ws + L.ws_cmed);
    p.tiles = reinterpret_cast<TileState *>(d_ws + L.ws_tiles);
    p.ticket = reinterpret_cast<unsigned *>(misc);
    p.totals = reinterpret_cast<uint64_t *>(misc + 2);
    p.index = d_index;
    p.ntiles = L.ntiles;
    p.nbit = L.nbit;
    timing_mark_begin(KK_COMPRESS, st);
    const int rc = w == 4 ? dispatch_bs<float>(bs, p, L.ntiles, st)
                          : dispatch_bs<double>(bs, p, L.ntiles, st);
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
    const uint64_t work = L.nbit > 4 * L.nb ? L.nbit : 4 * L.nb;
    uint64_t blocks = (work + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_assemble, dim3((unsigned)blocks), dim3(256), 0, st, a);
    return 0;
}

}  // namespace szxamd
