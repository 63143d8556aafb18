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

// Lanes past the block repeat its first value; columns past it are not loaded.
template <typename T>
__device__ inline void load_block(const T *b, int bs, int lane, T (&v)[FX_MAXC])
{
    v[0] = b[lane < bs ? lane : 0];
#pragma unroll
    for (int c = 1; c < FX_MAXC; c++) {
        const int j = lane + 64 * c;
        v[c] = 64 * c < bs ? b[j < bs ? j : 0] : v[0];
    }
}

// Blocks a wave loads before working on them (memory-level parallelism of
// the one-wave-per-block passes).
constexpr int PF = 4;

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
    const uint64_t step = gridDim.x * 4ull;
    for (uint64_t i0 = blockIdx.x * 4ull + wv; i0 < ns; i0 += PF * step) {
        T v[PF][FX_MAXC];
#pragma unroll
        for (int p = 0; p < PF; p++)
            if (i0 + p * step < ns) load_block(x + (i0 + p * step) * stride, bs, lane, v[p]);
#pragma unroll
        for (int p = 0; p < PF; p++) {
            const uint64_t i = i0 + p * step;
            if (i >= ns) break;
            T lo, hi;
            block_min_max(x + i * stride, bs, lane, v[p], lo, hi);
            const T r = (hi - lo) / 2;
            if (lane == 0) {
                rad[i] = (float)r;
                med[i] = (float)(lo + r);
            }
            glo = lo < glo ? lo : glo;
            ghi = hi > ghi ? hi : ghi;
        }
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
    const uint64_t step = gridDim.x * 4ull;
    for (uint64_t i0 = blockIdx.x * 4ull + wv; i0 < ns; i0 += PF * step) {
      T vp[PF][FX_MAXC];
#pragma unroll
      for (int p = 0; p < PF; p++)
          if (i0 + p * step < ns) load_block(x + (i0 + p * step) * stride, bs, lane, vp[p]);
#pragma unroll
      for (int p = 0; p < PF; p++) {
        const uint64_t i = i0 + p * step;
        if (i >= ns) break;
        const T (&v)[FX_MAXC] = vp[p];
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
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 3 * K; k += 256)
        part[(uint64_t)blockIdx.x * 3 * K + k] = s_sum[k];
}

__global__ void k_fill_u64(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// out[c] += sum over g of part[g * width + c] (out zeroed first).  A
// workgroup folds 64 columns over a slice of the rows (4 rows per step, four
// independent accumulators per thread) and adds its sums with one atomic per
// column.
constexpr int FOLD_SPLITS = 16;
__global__ __launch_bounds__(256) void k_fold_parts(const unsigned long long *part, int nparts, int width,
                                                    unsigned long long *out)
{
    __shared__ unsigned long long s[4][64];
    const int col = blockIdx.x * 64 + (threadIdx.x & 63), rg = threadIdx.x >> 6;
    unsigned long long a[4] = {0, 0, 0, 0};
    if (col < width) {
        const int rstep = FOLD_SPLITS * 4;
        int r = blockIdx.y * 4 + rg;
        for (; r + 3 * rstep < nparts; r += 4 * rstep)
#pragma unroll
            for (int u = 0; u < 4; u++) a[u] += part[(uint64_t)(r + u * rstep) * width + col];
        for (; r < nparts; r += rstep) a[0] += part[(uint64_t)r * width + col];
    }
    s[rg][threadIdx.x & 63] = a[0] + a[1] + a[2] + a[3];
    __syncthreads();
    if (rg == 0 && col < width) {
        const unsigned long long t = s[0][threadIdx.x] + s[1][threadIdx.x] + s[2][threadIdx.x] +
                                     s[3][threadIdx.x];
        if (t) atomicAdd(&out[col], t);
    }
}

static void sum_parts(const unsigned long long *part, int nparts, int width, unsigned long long *out,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_u64, dim3((width + 255) / 256), dim3(256), 0, st, out, width);
    hipLaunchKernelGGL(k_fold_parts, dim3((width + 63) / 64, FOLD_SPLITS), dim3(256), 0, st, part, nparts,
                       width, out);
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

// One wave per codec block, over NOCT adjacent octaves: octave j (0 <= j <
// NOCT) holds the bounds with exponent reqexp - j.  acc layout (u64, zeroed
// before the launch), per octave j at j * OCT_ACC:
//   [0] bytes of constant blocks and records fixed for the whole octave
//       (radius <= 2^(reqexp-j): always constant, 4 bytes; radius >
//       2^(reqexp-j+1): always a record)
//   [1 + 2k], [2 + 2k]: blocks with radius in bin k of (2^(reqexp-j),
//       2^(reqexp-j+1)]: count, record bytes.
// With EDGES, the two bounds just outside the span as well (every block's
// reqLength moves by one bit there, so the rate jumps), at NOCT * OCT_ACC:
//   [0], [1]: at 2^(reqexp+1) (exponent reqexp+1): constant blocks, record
//       bytes;
//   [2], [3]: at the float below 2^(reqexp-NOCT+1) (exponent reqexp-NOCT):
//       constant blocks, record bytes.
// The record bytes of a block depend on the exponent only, so each octave
// costs one record size per block whose radius exceeds its lower edge.
// Each workgroup bins in LDS and adds its non-zero bins to acc.
template <typename T, bool EDGES, int NOCT>
__global__ __launch_bounds__(256) void k_rate_octave(const T *x, uint64_t n, int bs, int reqexp,
                                                     unsigned long long *acc)
{
    constexpr int EDGE = NOCT * OCT_ACC;
    constexpr int WIDTH = EDGE + (EDGES ? 4 : 0);
    __shared__ unsigned long long s_acc[WIDTH];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < WIDTH; k += 256) s_acc[k] = 0;
    __syncthreads();
    const uint64_t nb = (n + bs - 1) / bs;
    // 2^(reqexp-j) as doubles (exact for every float exponent)
    double lo_edge[NOCT];
#pragma unroll
    for (int j = 0; j < NOCT; j++) lo_edge[j] = ldexp(1.0, reqexp - j);
    const double top = 2.0 * lo_edge[0];
    // the float bound just below the span (bounds are floats for both dtypes)
    const double below = (double)nextafterf(ldexpf(1.0f, reqexp - NOCT + 1), 0.0f);
    unsigned long long fixed[NOCT], a_const = 0, a_rec = 0, b_const = 0, b_rec = 0;
#pragma unroll
    for (int j = 0; j < NOCT; j++) fixed[j] = 0;
    const uint64_t step = gridDim.x * 4ull;
    for (uint64_t i0 = blockIdx.x * 4ull + wv; i0 < nb; i0 += PF * step) {
        T vp[PF][FX_MAXC];
#pragma unroll
        for (int p = 0; p < PF; p++) {
            const uint64_t i = i0 + p * step;
            if (i < nb) {
                const int cnt = (int)(n - i * (uint64_t)bs < (uint64_t)bs ? n - i * (uint64_t)bs : bs);
                load_block(x + i * (uint64_t)bs, cnt, lane, vp[p]);
            }
        }
#pragma unroll
        for (int p = 0; p < PF; p++) {
            const uint64_t i = i0 + p * step;
            if (i >= nb) break;
            const T *b = x + i * (uint64_t)bs;
            const int cnt = (int)(n - i * (uint64_t)bs < (uint64_t)bs ? n - i * (uint64_t)bs : bs);
            const T (&v)[FX_MAXC] = vp[p];
            T lo, hi;
            block_min_max(b, cnt, lane, v, lo, hi);
            const T r = (hi - lo) / 2;
            const T m = lo + r;
            // the constant test compares the block radius (double for fp64)
            // with the float bound widened to T; a NaN radius is never
            // constant
            const double rd = (double)r;
            if (EDGES) {
                if (rd <= top) a_const++;
                else a_rec += record_bytes<T>(v, cnt, lane, (float)r, (float)m, reqexp + 1);
                if (rd <= below) b_const++;
                else b_rec += record_bytes<T>(v, cnt, lane, (float)r, (float)m, reqexp - NOCT);
            }
#pragma unroll
            for (int j = 0; j < NOCT; j++) {
                if (rd <= lo_edge[j]) {
                    fixed[j] += 4;
                    continue;
                }
                const uint64_t rec = record_bytes<T>(v, cnt, lane, (float)r, (float)m, reqexp - j);
                if (!(rd <= 2.0 * lo_edge[j])) {
                    fixed[j] += rec;
                    continue;
                }
                // bin k holds radii in (2^e (1 + k/B), 2^e (1 + (k+1)/B)]
                const double t = (rd / lo_edge[j] - 1.0) * OCT_BINS;   // in (0, OCT_BINS]
                const int k = t >= OCT_BINS ? OCT_BINS - 1 : (t <= 1.0 ? 0 : (int)ceil(t) - 1);
                if (lane == 0) {
                    atomicAdd(&s_acc[j * OCT_ACC + 1 + 2 * k], 1ull);
                    atomicAdd(&s_acc[j * OCT_ACC + 2 + 2 * k], (unsigned long long)rec);
                }
            }
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < NOCT; j++)
            if (fixed[j]) atomicAdd(&s_acc[j * OCT_ACC], fixed[j]);
    }
    if (EDGES && lane == 0) {
        if (a_const) atomicAdd(&s_acc[EDGE + 0], a_const);
        if (a_rec) atomicAdd(&s_acc[EDGE + 1], a_rec);
        if (b_const) atomicAdd(&s_acc[EDGE + 2], b_const);
        if (b_rec) atomicAdd(&s_acc[EDGE + 3], b_rec);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < WIDTH; k += 256)
        if (s_acc[k]) atomicAdd(&acc[k], s_acc[k]);
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
    const uint64_t step = (uint64_t)wg_per_chunk * 4;
    for (uint64_t b0 = g * 4 + wv; b0 < bpc; b0 += PF * step) {
      T vp[PF][FX_MAXC];
#pragma unroll
      for (int p = 0; p < PF; p++) {
          const uint64_t b = b0 + p * step;
          if (b < bpc) {
              const uint64_t e0 = c0 + b * (uint64_t)bs;
              load_block(x + e0, (int)(c1 - e0 < (uint64_t)bs ? c1 - e0 : (uint64_t)bs), lane, vp[p]);
          }
      }
#pragma unroll
      for (int p = 0; p < PF; p++) {
        const uint64_t b = b0 + p * step;
        if (b >= bpc) break;
        const uint64_t e0 = c0 + b * (uint64_t)bs;
        const int cnt = (int)(c1 - e0 < (uint64_t)bs ? c1 - e0 : (uint64_t)bs);
        const T (&v)[FX_MAXC] = vp[p];
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

// The solver's ladder of candidate bounds, range * 2^(-k/4), k < K, from the
// chunk statistics (one workgroup).
__global__ void k_rate_ladder(const RateStats *st, int nchunks, int K, float *ebs)
{
    __shared__ double s_range;
    if (threadIdx.x == 0) {
        double lo = st[0].d[0], hi = st[0].d[1];
        for (int c = 1; c < nchunks; c++) {
            lo = st[c].d[0] < lo ? st[c].d[0] : lo;
            hi = st[c].d[1] > hi ? st[c].d[1] : hi;
        }
        const double r = hi - lo;
        s_range = r > 0 ? r : 1.0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x) ebs[k] = (float)(s_range * exp2(-0.25 * k));
}

// X[(k * nchunks + c)][24]: the rate model's inputs for chunk c at bound
// ebs[k].  One workgroup per chunk: the chunk's radius-exponent histogram and
// its prefix sums sit in LDS, then each thread fills the inputs of bounds.
//   0   (exponent of eb - floor(log2 range)) / 16
//   1   position of eb inside its octave, log2(eb) - exponent, in [0, 1)
//   2   log2(eb / range) / 16
//   3-6 log2(std, mean |d1|, rms d1, mean |d2| relative to the range) / 16
//   7   fraction of blocks with zero radius
//   8.. fraction of blocks per (radius exponent - eb exponent): <= -1, 0..13, >= 14
__global__ __launch_bounds__(256) void k_rate_inputs(const RateStats *st, int nchunks, const float *ebs,
                                                     int K, float *X)
{
    __shared__ unsigned s_pre[RATE_HIST + 1];   // s_pre[b] = blocks with exponent field < b
    const int c = blockIdx.x;
    const RateStats &s = st[c];
    if (threadIdx.x == 0) {
        unsigned a = 0;
        for (int b = 0; b < RATE_HIST; b++) {
            s_pre[b] = a;
            a += s.hist[b];
        }
        s_pre[RATE_HIST] = a;
    }
    __syncthreads();
    double range = s.d[1] - s.d[0];
    if (!(range > 0)) range = 1e-300;
    const double nv = (double)s.values;
    const double mean = s.d[2] / nv;
    double var = s.d[3] / nv - mean * mean;
    if (var < 0) var = 0;
    const double nd1 = nv > 1 ? nv - 1 : 1, nd2 = nv > 2 ? nv - 2 : 1;
    const int er = (int)floor(log2(range));
    const float nb = s.blocks ? (float)s.blocks : 1.0f;
    const float f3 = log2_ratio(sqrt(var), range), f4 = log2_ratio(s.d[4] / nd1, range);
    const float f5 = log2_ratio(sqrt(s.d[5] / nd1), range), f6 = log2_ratio(s.d[6] / nd2, range);
    auto below = [&](int field) {   // blocks whose exponent field is < field
        return (float)s_pre[field < 0 ? 0 : (field > RATE_HIST ? RATE_HIST : field)];
    };
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        float *f = X + ((uint64_t)k * nchunks + c) * RATE_INPUTS;
        const float eb = ebs[k];
        const int e = req_expo(eb);
        const double le = log2((double)eb);
        f[0] = (float)(e - er) / 16.0f;
        f[1] = (float)(le - e);
        f[2] = log2_ratio((double)eb, range);
        f[3] = f3;
        f[4] = f4;
        f[5] = f5;
        f[6] = f6;
        f[7] = (float)s.zero / nb;
        // exponent field b has radius exponent b - 127; d = b - 127 - e
        const int base = e + 127;   // field of d = 0
        f[8] = below(base) / nb;
        for (int d = 0; d < 14; d++) f[9 + d] = (below(base + d + 1) - below(base + d)) / nb;
        f[23] = ((float)s_pre[RATE_HIST] - below(base + 14)) / nb;
    }
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

void rate_ladder_launch(const RateStats *d_stats, int nchunks, int K, float *d_ebs, hipStream_t st)
{
    hipLaunchKernelGGL(k_rate_ladder, dim3(1), dim3(256), 0, st, d_stats, nchunks, K, d_ebs);
}

void rate_inputs_launch(const RateStats *d_stats, int nchunks, const float *d_ebs, int K, float *d_X,
                        hipStream_t st)
{
    hipLaunchKernelGGL(k_rate_inputs, dim3((unsigned)nchunks), dim3(256), 0, st, d_stats, nchunks,
                       d_ebs, K, d_X);
}

uint64_t octave_workspace_bytes() { return 256; }

void rate_octave_launch(const void *x, uint64_t n, int w, int bs, int reqexp, uint8_t *ws,
                        unsigned long long *d_acc, hipStream_t st, bool edges, int nocts)
{
    (void)ws;
    const int width = nocts * OCT_ACC + (edges || nocts == 2 ? 4 : 0);
    hipLaunchKernelGGL(k_fill_u64, dim3((width + 255) / 256), dim3(256), 0, st, d_acc, width);
    const uint64_t nb = (n + bs - 1) / bs;
    uint64_t g = (nb + 4 * PF - 1) / (4 * PF);
    if (g < 1) g = 1;
    if (g > OCT_WG) g = OCT_WG;
    const dim3 grid((unsigned)g), blk(256);
    const float *xf = static_cast<const float *>(x);
    const double *xd = static_cast<const double *>(x);
    if (nocts == 2) {   // the searches' octave pairs always carry the edges
        if (w == 4)
            hipLaunchKernelGGL((k_rate_octave<float, true, 2>), grid, blk, 0, st, xf, n, bs, reqexp, d_acc);
        else
            hipLaunchKernelGGL((k_rate_octave<double, true, 2>), grid, blk, 0, st, xd, n, bs, reqexp, d_acc);
    } else if (w == 4 && edges) {
        hipLaunchKernelGGL((k_rate_octave<float, true, 1>), grid, blk, 0, st, xf, n, bs, reqexp, d_acc);
    } else if (w == 4) {
        hipLaunchKernelGGL((k_rate_octave<float, false, 1>), grid, blk, 0, st, xf, n, bs, reqexp, d_acc);
    } else if (edges) {
        hipLaunchKernelGGL((k_rate_octave<double, true, 1>), grid, blk, 0, st, xd, n, bs, reqexp, d_acc);
    } else {
        hipLaunchKernelGGL((k_rate_octave<double, false, 1>), grid, blk, 0, st, xd, n, bs, reqexp, d_acc);
    }
}

}  // namespace szxamd
