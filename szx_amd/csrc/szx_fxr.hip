// szx_fxr.hip -- fixed-rate (FXR) support kernels for gfx950.
//
// 1. The reference's fixed-rate bound search (szx_float.cc:254-312, 501-584,
//    684-832; fp64 szx_double.cc:437-495, 497-580, 682-829) picks the bound
//    from a 1-in-`rate` block sample and an analytic size model.  The model
//    consumes three integers per candidate bound -- constant sampled blocks,
//    the sum of reqBytes and the sum of min(lead, reqBytes) -- so the GPU
//    computes those sums exactly (k_fxr_sample, k_fxr_estimate) and the host
//    evaluates the model's float formula and runs the search, which makes the
//    chosen bound identical to the reference's.
//
// 2. Exact rate of the real codec over a whole octave of bounds
//    (k_rate_octave).  For every bound eb in [2^e, 2^(e+1)) the exponent of eb
//    is e, so a non-constant block's reqLength, median, shift, lead codes and
//    record size do not depend on where eb lies in the octave; only the
//    constant test radius <= eb does.  One pass computing every block's
//    record size at exponent e and binning the blocks whose radius lies in
//    (2^e, 2^(e+1)] by radius therefore gives the exact stream size for the
//    bounds eb_k = 2^e (1 + k / OCT_BINS), k = 0 .. OCT_BINS-1.
#include "szx_common.hpp"
#include "szx_fxr.hpp"

namespace szxamd {

// One wave per block; bs <= 256, so a lane holds at most 4 values.
constexpr int FX_MAXC = 4;

template <typename T>
__device__ inline void load_block(const T *b, int bs, int lane, T (&v)[FX_MAXC])
{
#pragma unroll
    for (int c = 0; c < FX_MAXC; c++) {
        const int j = lane + 64 * c;
        v[c] = j < bs ? b[j] : b[0];
    }
}

// Block min / max with the reference's sequential semantics
// (szx_float.cc:268-276: `if (min > v) min = v; else if (max < v) max = v;`
// starting from the block's first value).  For NaN-free blocks this is the
// plain min / max; a block holding a NaN is redone serially by lane 0.
template <typename T>
__device__ inline void block_min_max(const T *b, int bs, int lane, const T (&v)[FX_MAXC], T &lo, T &hi)
{
    lo = v[0];
    hi = v[0];
    bool nan = false;
#pragma unroll
    for (int c = 0; c < FX_MAXC; c++) {
        nan |= v[c] != v[c];
        lo = v[c] < lo ? v[c] : lo;
        hi = v[c] > hi ? v[c] : hi;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const T ol = __shfl_xor(lo, m, 64), oh = __shfl_xor(hi, m, 64);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
    }
    if (__ballot(nan)) {
        T l = b[0], h = b[0];
        if (lane == 0)
            for (int j = 1; j < bs; j++) {
                const T x = b[j];
                if (l > x) l = x;
                else if (h < x) h = x;
            }
        lo = __shfl(l, 0, 64);
        hi = __shfl(h, 0, 64);
    }
}

__device__ inline unsigned long long wave_sum64(unsigned long long v)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Sum over a block of min(lead_i, rb) for the XOR leading-byte code of
// s_i = bits(x_i - median) >> shift (szx_float.cc:540-575).
template <typename T>
__device__ inline unsigned long long block_lead_sum(const T (&v)[FX_MAXC], int bs, int lane, T med,
                                                    int shift, int rb)
{
    using U = typename Word<T>::U;
    constexpr int W = Word<T>::W;
    unsigned long long acc = 0;
    U carry = 0;   // s of the last value of the previous 64-value column
#pragma unroll
    for (int c = 0; c < FX_MAXC; c++) {
        if (64 * c >= bs) break;
        const T d = v[c] - med;
        const U s = to_bits<T>(d) >> shift;
        U prev = __shfl_up(s, 1, 64);
        if (lane == 0) prev = carry;
        carry = __shfl(s, 63, 64);
        const U xr = s ^ prev;
        int ld = 0;
        if ((xr >> (8 * (W - 3))) == 0) ld = 3;
        else if ((xr >> (8 * (W - 2))) == 0) ld = 2;
        else if ((xr >> (8 * (W - 1))) == 0) ld = 1;
        if (lane + 64 * c < bs) acc += (unsigned long long)(ld < rb ? ld : rb);
    }
    return wave_sum64(acc);
}

// ---------------------------------------------------------------------------
// 1. the reference's sampled model
// ---------------------------------------------------------------------------

// Sampled block i starts at value i * stride.  med/rad as the reference
// stores them (float; fp64 rounds), part = per-workgroup {min lo, max hi}
// over the workgroup's blocks, NaN block extremes skipped.
template <typename T>
__global__ __launch_bounds__(256) void k_fxr_sample(const T *x, uint64_t ns, uint64_t stride, int bs,
                                                    float *med, float *rad, T *part)
{
    __shared__ T s_lo[4], s_hi[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    T glo = __builtin_inf(), ghi = -__builtin_inf();
    for (uint64_t i = blockIdx.x * 4ull + wv; i < ns; i += gridDim.x * 4ull) {
        const T *b = x + i * stride;
        T v[FX_MAXC];
        load_block(b, bs, lane, v);
        T lo, hi;
        block_min_max(b, bs, lane, v, lo, hi);
        const T r = (hi - lo) / 2;
        if (lane == 0) {
            rad[i] = (float)r;
            med[i] = (float)(lo + r);
        }
        glo = lo < glo ? lo : glo;
        ghi = hi > ghi ? hi : ghi;
    }
    if (lane == 0) {
        s_lo[wv] = glo;
        s_hi[wv] = ghi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            glo = s_lo[k] < glo ? s_lo[k] : glo;
            ghi = s_hi[k] > ghi ? s_hi[k] : ghi;
        }
        part[2 * blockIdx.x] = glo;
        part[2 * blockIdx.x + 1] = ghi;
    }
}

// Range of the sample (szx_float.cc:260-290): the running extremes start at
// x[0] and take each sampled block's min / max when it compares beyond them.
template <typename T>
__global__ void k_fxr_range(const T *x, const T *part, int nparts, float *range_out)
{
    __shared__ T s_lo[256], s_hi[256];
    T lo = __builtin_inf(), hi = -__builtin_inf();
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
        lo = part[2 * k] < lo ? part[2 * k] : lo;
        hi = part[2 * k + 1] > hi ? part[2 * k + 1] : hi;
    }
    s_lo[threadIdx.x] = lo;
    s_hi[threadIdx.x] = hi;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)blockDim.x; k++) {
            lo = s_lo[k] < lo ? s_lo[k] : lo;
            hi = s_hi[k] > hi ? s_hi[k] : hi;
        }
        T glo = x[0], ghi = x[0];
        if (ghi < hi) ghi = hi;
        if (glo > lo) glo = lo;
        *range_out = (float)(ghi - glo);
    }
}

// Model sums for K candidate bounds in one pass over the sample:
// part[(wg * K + k) * 3 + {0,1,2}] = {constant blocks, sum reqBytes,
// sum min(lead, reqBytes)}.
template <typename T>
__global__ __launch_bounds__(256) void k_fxr_estimate(const T *x, uint64_t ns, uint64_t stride, int bs,
                                                      const float *med, const float *rad,
                                                      const float *ebs, int K,
                                                      unsigned long long *part)
{
    __shared__ unsigned long long s_sum[FXR_MAX_CAND * 3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < 3 * K; k += 256) s_sum[k] = 0;
    __syncthreads();
    for (uint64_t i = blockIdx.x * 4ull + wv; i < ns; i += gridDim.x * 4ull) {
        const T *b = x + i * stride;
        T v[FX_MAXC];
        load_block(b, bs, lane, v);
        const float r = rad[i], m = med[i];
        for (int k = 0; k < K; k++) {
            const float eb = ebs[k];
            if (r <= eb) {
                if (lane == 0) atomicAdd(&s_sum[3 * k], 1ull);
                continue;
            }
            int clamped;
            const int len = req_length<T>(rad_expo(r), req_expo(eb), clamped);
            const int rb = (len + 7) / 8, shift = (8 - len % 8) % 8;
            const T mm = clamped ? (T)0 : (T)m;   // the model uses the zeroed median
            const unsigned long long ls = block_lead_sum(v, bs, lane, mm, shift, rb);
            if (lane == 0) {
                atomicAdd(&s_sum[3 * k + 1], (unsigned long long)rb);
                atomicAdd(&s_sum[3 * k + 2], ls);
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * K; k += 256)
        part[(uint64_t)blockIdx.x * 3 * K + k] = s_sum[k];
}

// out[c] = sum over g of part[g * width + c]; one thread per column.
__global__ void k_sum_parts(const unsigned long long *part, int nparts, int width,
                            unsigned long long *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= width) return;
    unsigned long long s = 0;
    for (int g = 0; g < nparts; g++) s += part[(uint64_t)g * width + c];
    out[c] = s;
}

static void sum_parts(const unsigned long long *part, int nparts, int width, unsigned long long *out,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_sum_parts, dim3((width + 255) / 256), dim3(256), 0, st, part, nparts, width,
                       out);
}

// ---------------------------------------------------------------------------
// 2. exact rate over an octave of bounds
// ---------------------------------------------------------------------------

// Record bytes of a non-constant block at bound exponent reqexp
// (szx_float.cc:1059-1139 record: reqLength byte, median, lead codes,
// mid-bytes).
template <typename T>
__device__ inline uint64_t record_bytes(const T (&v)[FX_MAXC], int n, int lane, float radius,
                                        float median, int reqexp)
{
    constexpr int W = Word<T>::W;
    int clamped;
    const int len = req_length<T>(rad_expo(radius), reqexp, clamped);
    const int rb = (len + 7) / 8, shift = (8 - len % 8) % 8;
    // the fp32 payload uses the zeroed median when the length clamps; the
    // fp64 payload keeps the real one (szx_double.cc:2027-2030)
    const T med = (W == 4 && clamped) ? (T)0 : (T)median;
    const uint64_t leads = block_lead_sum(v, n, lane, med, shift, rb);
    return 5 + (uint64_t)(n + 3) / 4 + (uint64_t)rb * n - leads;
}

// One wave per codec block.  acc layout (u64):
//   [0] bytes of constant blocks and records fixed for the whole octave
//       (radius <= 2^e: always constant, 4 bytes; radius > 2^(e+1): always a
//       record)
//   [1 + 2k], [2 + 2k]: blocks with radius in bin k of (2^e, 2^(e+1)]:
//       count, record bytes.
// Per-workgroup partials; k_sum_parts folds them.
template <typename T>
__global__ __launch_bounds__(256) void k_rate_octave(const T *x, uint64_t n, int bs, int reqexp,
                                                     unsigned long long *part)
{
    constexpr int WIDTH = 1 + 2 * OCT_BINS;
    __shared__ unsigned long long s_acc[WIDTH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < WIDTH; k += 256) s_acc[k] = 0;
    __syncthreads();
    const uint64_t nb = (n + bs - 1) / bs;
    // 2^reqexp and 2^(reqexp+1) as doubles (exact for every float exponent)
    const double lo_edge = ldexp(1.0, reqexp), hi_edge = 2.0 * lo_edge;
    unsigned long long fixed = 0;
    for (uint64_t i = blockIdx.x * 4ull + wv; i < nb; i += gridDim.x * 4ull) {
        const T *b = x + i * (uint64_t)bs;
        const int cnt = (int)(n - i * (uint64_t)bs < (uint64_t)bs ? n - i * (uint64_t)bs : bs);
        T v[FX_MAXC];
        load_block(b, cnt, lane, v);
        T lo, hi;
        block_min_max(b, cnt, lane, v, lo, hi);
        const T r = (hi - lo) / 2;
        const T m = lo + r;
        // the constant test compares the block radius (double for fp64) with
        // the float bound widened to T
        const double rd = (double)r;
        if (rd <= lo_edge) {
            fixed += 4;
            continue;
        }
        const uint64_t rec = record_bytes<T>(v, cnt, lane, (float)r, (float)m, reqexp);
        if (!(rd <= hi_edge)) {   // also NaN radii: never constant
            fixed += rec;
            continue;
        }
        // bin k holds radii in (2^e (1 + k/B), 2^e (1 + (k+1)/B)]
        const double t = (rd / lo_edge - 1.0) * OCT_BINS;   // in (0, OCT_BINS]
        const int k = t >= OCT_BINS ? OCT_BINS - 1 : (t <= 1.0 ? 0 : (int)ceil(t) - 1);
        if (lane == 0) {
            atomicAdd(&s_acc[1 + 2 * k], 1ull);
            atomicAdd(&s_acc[2 + 2 * k], (unsigned long long)rec);
        }
    }
    if (lane == 0 && fixed) atomicAdd(&s_acc[0], fixed);
    __syncthreads();
    for (int k = threadIdx.x; k < WIDTH; k += 256) part[(uint64_t)blockIdx.x * WIDTH + k] = s_acc[k];
}

// ---------------------------------------------------------------------------
// 3. per-chunk statistics and the rate model's inputs
// ---------------------------------------------------------------------------

// Per workgroup partial statistics of its slice of one chunk (RateStats
// layout: doubles lo, hi, sum x, sum x^2, sum |d1|, sum d1^2, sum |d2|, then
// u32 blocks, zero-radius blocks, radius-exponent histogram [256]).
// d1 / d2 are first / second differences inside a codec block.
template <typename T>
__global__ __launch_bounds__(256) void k_rate_stats(const T *x, uint64_t n, int bs, uint64_t chunk,
                                                    int wg_per_chunk, RateStats *part)
{
    __shared__ double s_d[4][7];
    __shared__ unsigned s_hist[RATE_HIST];
    __shared__ unsigned s_cnt[2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < RATE_HIST; k += 256) s_hist[k] = 0;
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t c = blockIdx.x / wg_per_chunk, g = blockIdx.x % wg_per_chunk;
    const uint64_t c0 = c * chunk, c1 = (c0 + chunk < n) ? c0 + chunk : n;
    const uint64_t bpc = (c1 - c0 + bs - 1) / bs;   // blocks in the chunk
    double lo = __builtin_inf(), hi = -__builtin_inf(), sx = 0, sxx = 0, sa = 0, sq = 0, s2 = 0;
    for (uint64_t b = g * 4 + wv; b < bpc; b += (uint64_t)wg_per_chunk * 4) {
        const uint64_t e0 = c0 + b * (uint64_t)bs;
        const int cnt = (int)(c1 - e0 < (uint64_t)bs ? c1 - e0 : (uint64_t)bs);
        T v[FX_MAXC];
        load_block(x + e0, cnt, lane, v);
        T blo, bhi;
        block_min_max(x + e0, cnt, lane, v, blo, bhi);
        const float r = (float)((bhi - blo) / 2);
        if (lane == 0) {
            atomicAdd(&s_cnt[0], 1u);
            if (r == 0.0f) atomicAdd(&s_cnt[1], 1u);
            atomicAdd(&s_hist[(f32_bits(r) >> 23) & 0xFF], 1u);
        }
        lo = (double)blo < lo ? (double)blo : lo;
        hi = (double)bhi > hi ? (double)bhi : hi;
        double carry1 = 0, carry0 = 0;   // last d1 and value of the previous column
#pragma unroll
        for (int q = 0; q < FX_MAXC; q++) {
            if (64 * q >= cnt) break;
            const int j = lane + 64 * q;
            const double xv = (double)v[q];
            double prev = __shfl_up(xv, 1, 64);
            if (lane == 0) prev = carry0;
            const double d1 = xv - prev;
            double pd1 = __shfl_up(d1, 1, 64);
            if (lane == 0) pd1 = carry1;
            carry0 = __shfl(xv, 63, 64);
            carry1 = __shfl(d1, 63, 64);
            if (j < cnt) {
                sx += xv;
                sxx += xv * xv;
                if (j >= 1) {
                    sa += fabs(d1);
                    sq += d1 * d1;
                }
                if (j >= 2) s2 += fabs(d1 - pd1);
            }
        }
    }
    double acc[7] = {lo, hi, sx, sxx, sa, sq, s2};
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        double o = __shfl_xor(acc[0], m, 64);
        acc[0] = o < acc[0] ? o : acc[0];
        o = __shfl_xor(acc[1], m, 64);
        acc[1] = o > acc[1] ? o : acc[1];
#pragma unroll
        for (int k = 2; k < 7; k++) acc[k] += __shfl_xor(acc[k], m, 64);
    }
    if (lane == 0)
        for (int k = 0; k < 7; k++) s_d[wv][k] = acc[k];
    __syncthreads();
    RateStats &o = part[blockIdx.x];
    if (threadIdx.x == 0) {
        double a[7];
        for (int k = 0; k < 7; k++) a[k] = s_d[0][k];
        for (int w = 1; w < 4; w++) {
            a[0] = s_d[w][0] < a[0] ? s_d[w][0] : a[0];
            a[1] = s_d[w][1] > a[1] ? s_d[w][1] : a[1];
            for (int k = 2; k < 7; k++) a[k] += s_d[w][k];
        }
        for (int k = 0; k < 7; k++) o.d[k] = a[k];
        o.blocks = s_cnt[0];
        o.zero = s_cnt[1];
        o.values = c1 - c0;
    }
    for (int k = threadIdx.x; k < RATE_HIST; k += 256) o.hist[k] = s_hist[k];
}

// Fold wg_per_chunk partials per chunk (one workgroup per chunk).
__global__ void k_rate_stats_fold(const RateStats *part, int wg_per_chunk, RateStats *out)
{
    const RateStats *p = part + (uint64_t)blockIdx.x * wg_per_chunk;
    RateStats &o = out[blockIdx.x];
    for (int k = threadIdx.x; k < RATE_HIST; k += blockDim.x) {
        unsigned h = 0;
        for (int g = 0; g < wg_per_chunk; g++) h += p[g].hist[k];
        o.hist[k] = h;
    }
    if (threadIdx.x == 0) {
        double a[7];
        for (int k = 0; k < 7; k++) a[k] = p[0].d[k];
        unsigned blocks = p[0].blocks, zero = p[0].zero;
        for (int g = 1; g < wg_per_chunk; g++) {
            a[0] = p[g].d[0] < a[0] ? p[g].d[0] : a[0];
            a[1] = p[g].d[1] > a[1] ? p[g].d[1] : a[1];
            for (int k = 2; k < 7; k++) a[k] += p[g].d[k];
            blocks += p[g].blocks;
            zero += p[g].zero;
        }
        for (int k = 0; k < 7; k++) o.d[k] = a[k];
        o.blocks = blocks;
        o.zero = zero;
        o.values = p[0].values;
    }
}

__device__ inline float log2_ratio(double a, double r)
{
    const double q = a / r;
    const double lo = 8.673617379884035e-19;   // 2^-60
    return (float)(log2(q > lo ? q : lo) / 16.0);
}

// X[(k * nchunks + c)][24]: the rate model's inputs for chunk c at bound
// ebs[k] (see rate_features for the meaning of each input).
__global__ void k_rate_inputs(const RateStats *st, int nchunks, const float *ebs, int K, float *X)
{
    const uint64_t id = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (id >= (uint64_t)K * nchunks) return;
    const int k = (int)(id / nchunks), c = (int)(id % nchunks);
    const RateStats &s = st[c];
    float *f = X + id * RATE_INPUTS;
    const float eb = ebs[k];
    const int e = req_expo(eb);
    double range = s.d[1] - s.d[0];
    if (!(range > 0)) range = 1e-300;
    const double nv = (double)s.values;
    const double mean = s.d[2] / nv;
    double var = s.d[3] / nv - mean * mean;
    if (var < 0) var = 0;
    const double nd1 = nv > 1 ? nv - 1 : 1, nd2 = nv > 2 ? nv - 2 : 1;
    const int er = (int)floor(log2(range));
    const double le = log2((double)eb);
    f[0] = (float)(e - er) / 16.0f;
    f[1] = (float)(le - e);
    f[2] = log2_ratio((double)eb, range);
    f[3] = log2_ratio(sqrt(var), range);
    f[4] = log2_ratio(s.d[4] / nd1, range);
    f[5] = log2_ratio(sqrt(s.d[5] / nd1), range);
    f[6] = log2_ratio(s.d[6] / nd2, range);
    const float nb = s.blocks ? (float)s.blocks : 1.0f;
    f[7] = (float)s.zero / nb;
    // histogram of (radius exponent - bound exponent): <= -1, 0 .. 13, >= 14
    float h[16];
    for (int b = 0; b < 16; b++) h[b] = 0;
    for (int b = 0; b < RATE_HIST; b++) {
        const unsigned cnt = s.hist[b];
        if (!cnt) continue;
        const int d = (b - 127) - e;
        const int bin = d <= -1 ? 0 : (d >= 14 ? 15 : d + 1);
        h[bin] += (float)cnt;
    }
    for (int b = 0; b < 16; b++) f[8 + b] = h[b] / nb;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

static int fx_grid(uint64_t waves)
{
    uint64_t g = (waves + 3) / 4;
    if (g < 1) g = 1;
    if (g > FXR_MAX_WG) g = FXR_MAX_WG;
    return (int)g;
}

FxrLayout fxr_layout(uint64_t n, int w, int bs, int rate)
{
    FxrLayout L{};
    L.ns = n / ((uint64_t)rate * (uint64_t)bs);
    uint64_t off = 0;
    L.ws_med = off;   off += ((L.ns * 4 + 255) & ~255ull);
    L.ws_rad = off;   off += ((L.ns * 4 + 255) & ~255ull);
    L.ws_part = off;  off += (uint64_t)FXR_MAX_WG * FXR_MAX_CAND * 3 * 8;
    L.ws_ebs = off;   off += 256;
    L.ws_sums = off;  off += FXR_MAX_CAND * 3 * 8;
    L.ws_range = off; off += 256;
    (void)w;
    L.ws_bytes = off;
    return L;
}

void fxr_sample_launch(const void *x, uint64_t n, int w, int bs, int rate, uint8_t *ws, hipStream_t st)
{
    const FxrLayout L = fxr_layout(n, w, bs, rate);
    const int g = fx_grid(L.ns);
    float *med = reinterpret_cast<float *>(ws + L.ws_med);
    float *rad = reinterpret_cast<float *>(ws + L.ws_rad);
    float *range = reinterpret_cast<float *>(ws + L.ws_range);
    if (w == 4) {
        float *part = reinterpret_cast<float *>(ws + L.ws_part);
        hipLaunchKernelGGL(k_fxr_sample<float>, dim3(g), dim3(256), 0, st,
                           static_cast<const float *>(x), L.ns, (uint64_t)rate * bs, bs, med, rad, part);
        hipLaunchKernelGGL(k_fxr_range<float>, dim3(1), dim3(256), 0, st,
                           static_cast<const float *>(x), part, g, range);
    } else {
        double *part = reinterpret_cast<double *>(ws + L.ws_part);
        hipLaunchKernelGGL(k_fxr_sample<double>, dim3(g), dim3(256), 0, st,
                           static_cast<const double *>(x), L.ns, (uint64_t)rate * bs, bs, med, rad, part);
        hipLaunchKernelGGL(k_fxr_range<double>, dim3(1), dim3(256), 0, st,
                           static_cast<const double *>(x), part, g, range);
    }
}

void fxr_estimate_launch(const void *x, uint64_t n, int w, int bs, int rate, uint8_t *ws, int K,
                         hipStream_t st)
{
    const FxrLayout L = fxr_layout(n, w, bs, rate);
    const int g = fx_grid(L.ns);
    const float *med = reinterpret_cast<const float *>(ws + L.ws_med);
    const float *rad = reinterpret_cast<const float *>(ws + L.ws_rad);
    const float *ebs = reinterpret_cast<const float *>(ws + L.ws_ebs);
    auto *part = reinterpret_cast<unsigned long long *>(ws + L.ws_part);
    auto *sums = reinterpret_cast<unsigned long long *>(ws + L.ws_sums);
    if (w == 4)
        hipLaunchKernelGGL(k_fxr_estimate<float>, dim3(g), dim3(256), 0, st,
                           static_cast<const float *>(x), L.ns, (uint64_t)rate * bs, bs, med, rad, ebs,
                           K, part);
    else
        hipLaunchKernelGGL(k_fxr_estimate<double>, dim3(g), dim3(256), 0, st,
                           static_cast<const double *>(x), L.ns, (uint64_t)rate * bs, bs, med, rad, ebs,
                           K, part);
    sum_parts(part, g, 3 * K, sums, st);
}

uint64_t rate_stats_workspace_bytes(uint64_t n)
{
    const uint64_t nchunks = (n + RATE_CHUNK - 1) / RATE_CHUNK;
    return (nchunks ? nchunks : 1) * RATE_WG_PER_CHUNK * sizeof(RateStats);
}

void rate_stats_launch(const void *x, uint64_t n, int w, int bs, uint8_t *ws, RateStats *d_stats,
                       hipStream_t st)
{
    const uint64_t nchunks = (n + RATE_CHUNK - 1) / RATE_CHUNK;
    RateStats *part = reinterpret_cast<RateStats *>(ws);
    const unsigned g = (unsigned)(nchunks * RATE_WG_PER_CHUNK);
    if (w == 4)
        hipLaunchKernelGGL(k_rate_stats<float>, dim3(g), dim3(256), 0, st,
                           static_cast<const float *>(x), n, bs, (uint64_t)RATE_CHUNK,
                           RATE_WG_PER_CHUNK, part);
    else
        hipLaunchKernelGGL(k_rate_stats<double>, dim3(g), dim3(256), 0, st,
                           static_cast<const double *>(x), n, bs, (uint64_t)RATE_CHUNK,
                           RATE_WG_PER_CHUNK, part);
    hipLaunchKernelGGL(k_rate_stats_fold, dim3((unsigned)nchunks), dim3(256), 0, st, part,
                       RATE_WG_PER_CHUNK, d_stats);
}

void rate_inputs_launch(const RateStats *d_stats, int nchunks, const float *d_ebs, int K, float *d_X,
                        hipStream_t st)
{
    const uint64_t total = (uint64_t)K * nchunks;
    hipLaunchKernelGGL(k_rate_inputs, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, d_stats,
                       nchunks, d_ebs, K, d_X);
}

uint64_t octave_workspace_bytes() { return (uint64_t)OCT_WG * (1 + 2 * OCT_BINS) * 8; }

void rate_octave_launch(const void *x, uint64_t n, int w, int bs, int reqexp, uint8_t *ws,
                        unsigned long long *d_acc, hipStream_t st)
{
    auto *part = reinterpret_cast<unsigned long long *>(ws);
    const uint64_t nb = (n + bs - 1) / bs;
    uint64_t g = (nb + 3) / 4;
    if (g < 1) g = 1;
    if (g > OCT_WG) g = OCT_WG;
    if (w == 4)
        hipLaunchKernelGGL(k_rate_octave<float>, dim3((unsigned)g), dim3(256), 0, st,
                           static_cast<const float *>(x), n, bs, reqexp, part);
    else
        hipLaunchKernelGGL(k_rate_octave<double>, dim3((unsigned)g), dim3(256), 0, st,
                           static_cast<const double *>(x), n, bs, reqexp, part);
    sum_parts(part, (int)g, 1 + 2 * OCT_BINS, d_acc, st);
}

}  // namespace szxamd
