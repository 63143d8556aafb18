// szx_api.cpp -- the C ABI of libszx_amd.so (declared in include/szx_amd.h).
//
// Host entry points keep the reference's signatures and ownership (host
// pointers in, malloc'd buffers out) and move the data through device buffers
// that are cached per thread and device, so repeated calls do not pay for
// hipMalloc.  Device entry points only enqueue kernels.
#include "szx_amd.h"

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>

#include "szx_fxr.hpp"
#include "szx_fxr_host.hpp"
#include "szx_internal.hpp"
#include "szx_mlp.hpp"
#include "szx_rate.hpp"

using namespace szxamd;

namespace {

thread_local std::string g_err;

void set_err(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    set_err("%s: %s", what, hipGetErrorString(e));
    return false;
}

int width_of(int dataType)
{
    if (dataType == SZx_FLOAT) return 4;
    if (dataType == SZx_DOUBLE) return 8;
    return 0;
}

bool bs_supported(int bs) { return bs_valid(bs); }

// Growable device scratch kept for the life of the thread (one per role).
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int dev = -1;
    bool reserve(size_t bytes)
    {
        int d = 0;
        if (!hip_ok(hipGetDevice(&d), "hipGetDevice")) return false;
        if (p && cap >= bytes && dev == d) return true;
        if (p) {
            int cur = d;
            (void)hipSetDevice(dev);
            (void)hipFree(p);
            (void)hipSetDevice(cur);
            p = nullptr;
            cap = 0;
        }
        if (!hip_ok(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc")) return false;
        cap = bytes;
        dev = d;
        return true;
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

struct HostCtx {
    DevBuf in, out, ws, res, fx;
    hipStream_t stream = nullptr;
    int dev = -1;
    bool ready()
    {
        int d = 0;
        if (!hip_ok(hipGetDevice(&d), "hipGetDevice")) return false;
        if (stream && dev == d) return true;
        if (stream) (void)hipStreamDestroy(stream);
        if (!hip_ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate"))
            return false;
        dev = d;
        return true;
    }
};

thread_local HostCtx g_ctx;

// Kernel timing (szx_amd_timing_*): events recorded around main kernels.
struct TimingState {
    bool on = false;
    struct Pair { hipEvent_t a, b; int kind; };
    Pair pairs[4096];
    int npairs = 0;
    hipEvent_t pending[KK_COUNT] = {};
    hipEvent_t pool[8192];
    int npool = 0, used = 0;
    hipEvent_t get()
    {
        if (used < npool) return pool[used++];
        if (npool >= 8192) return nullptr;
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        pool[npool++] = e;
        used++;
        return e;
    }
};
thread_local TimingState g_tim;

size_t data_length(size_t r5, size_t r4, size_t r3, size_t r2, size_t r1)
{
    // szx.cc:73-101: trailing zero dims are absent, r1 is the fastest.
    if (r1 == 0) return 0;
    if (r2 == 0) return r1;
    if (r3 == 0) return r1 * r2;
    if (r4 == 0) return r1 * r2 * r3;
    if (r5 == 0) return r1 * r2 * r3 * r4;
    return r1 * r2 * r3 * r4 * r5;
}

// The reference's fixed-rate bound (szx_float.cc:921-949 without the final
// compress) for device data: sample and model sums on the GPU, the search on
// the host.  With default_init the search starts at 1% of the sampled range,
// as the reference's fixed-rate compressor does; otherwise at init.
// Synchronizes `st` once per model pass.
constexpr int FXR_RATE = 10;   // szx_float.cc:930: one block in 10 is sampled

bool fxr_bound_device(const void *d_in, size_t n, int w, int bs, float target, float tol,
                      bool default_init, float init, uint8_t *d_ws, hipStream_t st, float *eb_out,
                      int *n_est)
{
    const FxrLayout L = fxr_layout(n, w, bs, FXR_RATE);
    fxr_sample_launch(d_in, n, w, bs, FXR_RATE, d_ws, st);
    if (default_init) {
        float range = 0.0f;
        if (!hip_ok(hipMemcpyAsync(&range, d_ws + L.ws_range, 4, hipMemcpyDeviceToHost, st), "D2H") ||
            !hip_ok(hipStreamSynchronize(st), "fixed-rate sample"))
            return false;
        init = range * 1E-2;   // szx_float.cc:941 (float * double, rounded to float)
    }
    FxrEvaluator eval = [&](const float *ebs, int K, FxrSums *out) {
        uint64_t sums[3 * FXR_MAX_CAND];
        if (!hip_ok(hipMemcpyAsync(d_ws + L.ws_ebs, ebs, 4 * (size_t)K, hipMemcpyHostToDevice, st),
                    "H2D bounds"))
            return false;
        fxr_estimate_launch(d_in, n, w, bs, FXR_RATE, d_ws, K, st);
        if (!hip_ok(hipMemcpyAsync(sums, d_ws + L.ws_sums, 24 * (size_t)K, hipMemcpyDeviceToHost, st),
                    "D2H sums") ||
            !hip_ok(hipStreamSynchronize(st), "fixed-rate model"))
            return false;
        for (int k = 0; k < K; k++) out[k] = {sums[3 * k], sums[3 * k + 1], sums[3 * k + 2]};
        return true;
    };
    int ne = 0;
    if (!fxr_search(target, tol, init, w, bs, n, L.ns, eval, eb_out, &ne)) return false;
    if (n_est) *n_est = ne;
    return true;
}

struct FxrReq {
    float target, tol;
};

// Host-in / host-out compress through the device path.  With fxr set the
// bound is the reference's fixed-rate choice for (target ratio, tolerance).
// With dst set the stream is copied there (the caller sized it) instead of
// into a malloc'd buffer.  A look-back wait that overran its bound (another
// process holding the GPU's compute units for a long time) makes the stream
// invalid; the compress is then enqueued once more before the call fails.
unsigned char *host_compress(const void *data, size_t n, int w, int rel_mode, float abs_eb,
                             float rel, int bs, size_t *outSize, const FxrReq *fxr = nullptr,
                             int fast_mode = SZx_WITH_BLOCK_FAST_CMPR, unsigned char *dst = nullptr)
{
    if (outSize) *outSize = 0;
    if (!data || n == 0) {
        set_err("empty input");
        return nullptr;
    }
    if (!bs_supported(bs)) {
        set_err("blockSize %d is not supported (1..256)", bs);
        return nullptr;
    }
    HostCtx &c = g_ctx;
    if (!c.ready()) return nullptr;
    const CompressLayout L = compress_layout(n, w, bs);
    if (!c.in.reserve(n * (size_t)w) || !c.out.reserve(L.capacity) || !c.ws.reserve(L.ws_bytes) ||
        !c.res.reserve(64))
        return nullptr;
    if (!hip_ok(hipMemcpyAsync(c.in.p, data, n * (size_t)w, hipMemcpyHostToDevice, c.stream),
                "hipMemcpyAsync H2D"))
        return nullptr;
    if (fxr) {
        if (!c.fx.reserve(fxr_layout(n, w, bs, FXR_RATE).ws_bytes) ||
            !fxr_bound_device(c.in.p, n, w, bs, fxr->target, fxr->tol, true, 0.0f,
                              (uint8_t *)c.fx.p, c.stream, &abs_eb, nullptr))
            return nullptr;
        rel_mode = 0;
    }
    uint64_t res[4] = {0, 0, 0, 0};
    for (int attempt = 0; attempt < 2; attempt++) {
        if (compress_launch(c.in.p, n, w, rel_mode, abs_eb, rel, bs, (uint8_t *)c.out.p,
                            (uint8_t *)c.ws.p, (uint64_t *)c.res.p, nullptr, c.stream, fast_mode)) {
            set_err("compress launch failed");
            return nullptr;
        }
        if (!hip_ok(hipMemcpyAsync(res, c.res.p, sizeof res, hipMemcpyDeviceToHost, c.stream),
                    "hipMemcpyAsync D2H") ||
            !hip_ok(hipStreamSynchronize(c.stream), "compress"))
            return nullptr;
        if (!res[3]) break;
    }
    if (res[3]) {
        set_err("compress: a look-back wait overran its bound twice (device busy?)");
        return nullptr;
    }
    unsigned char *out = dst ? dst : (unsigned char *)malloc(res[1] ? res[1] : 1);
    if (!out) {
        set_err("malloc(%llu) failed", (unsigned long long)res[1]);
        return nullptr;
    }
    if (!hip_ok(hipMemcpy(out, (uint8_t *)c.out.p + res[0], res[1], hipMemcpyDeviceToHost),
                "hipMemcpy D2H stream")) {
        if (!dst) free(out);
        return nullptr;
    }
    if (outSize) *outSize = res[1];
    return out;
}

void *host_decompress(const unsigned char *bytes, size_t len, size_t n, int w, int exact,
                      int fast_mode = SZx_WITH_BLOCK_FAST_CMPR)
{
    if (!bytes || n == 0) {
        set_err("empty input");
        return nullptr;
    }
    const bool ra = fast_mode != SZx_WITH_BLOCK_FAST_CMPR;
    if (len < (ra ? 20u : 13u)) {
        set_err("stream shorter than its header");
        return nullptr;
    }
    int bs = bytes[4] ? bytes[4] : 256;
    if (ra) {   // big-endian u64 blockSize (szxd_float.cc:419)
        uint64_t v = 0;
        for (int k = 0; k < 8; k++) v = (v << 8) | bytes[4 + k];
        bs = v >= 4 && v <= 256 && (v & (v - 1)) == 0 ? (int)v : 0;
    }
    if (!bs_supported(bs) || (ra && !bs_tiled(bs))) {
        set_err("blockSize %d is not supported", bs);
        return nullptr;
    }
    HostCtx &c = g_ctx;
    if (!c.ready()) return nullptr;
    const DecompressLayout L = decompress_layout(n, w, bs, len);
    if (!c.in.reserve(len) || !c.out.reserve(n * (size_t)w) || !c.ws.reserve(L.ws_bytes) ||
        !c.res.reserve(64))
        return nullptr;
    if (!hip_ok(hipMemcpyAsync(c.in.p, bytes, len, hipMemcpyHostToDevice, c.stream), "H2D"))
        return nullptr;
    if (decompress_launch((const uint8_t *)c.in.p, len, n, w, bs, c.out.p, (uint8_t *)c.ws.p,
                          L.ws_bytes, nullptr, exact, (int32_t *)c.res.p, c.stream, fast_mode)) {
        set_err("decompress launch failed");
        return nullptr;
    }
    int32_t status = 0;
    if (!hip_ok(hipMemcpyAsync(&status, c.res.p, 4, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipStreamSynchronize(c.stream), "decompress"))
        return nullptr;
    if (status != 0) {
        set_err("malformed or unsupported stream (code %d)", status);
        return nullptr;
    }
    void *out = malloc(n * (size_t)w);
    if (!out) {
        set_err("malloc failed");
        return nullptr;
    }
    if (!hip_ok(hipMemcpy(out, c.out.p, n * (size_t)w, hipMemcpyDeviceToHost), "D2H output")) {
        free(out);
        return nullptr;
    }
    return out;
}

// fastMode 1 (one record over the whole array) through the device path.
unsigned char *host_compress_nb(const void *data, size_t n, int w, int rel_mode, float abs_eb, float rel,
                                size_t *outSize)
{
    if (outSize) *outSize = 0;
    if (!data || n == 0) {
        set_err("empty input");
        return nullptr;
    }
    HostCtx &c = g_ctx;
    if (!c.ready()) return nullptr;
    if (!c.in.reserve(n * (size_t)w) || !c.out.reserve(nb_capacity(n, w)) || !c.ws.reserve(nb_workspace(n)) ||
        !c.res.reserve(64))
        return nullptr;
    if (!hip_ok(hipMemcpyAsync(c.in.p, data, n * (size_t)w, hipMemcpyHostToDevice, c.stream), "H2D"))
        return nullptr;
    nb_compress_launch(c.in.p, n, w, rel_mode, abs_eb, rel, (uint8_t *)c.out.p, (uint8_t *)c.ws.p,
                       (uint64_t *)c.res.p, c.stream);
    uint64_t res[4] = {0, 0, 0, 0};
    if (!hip_ok(hipMemcpyAsync(res, c.res.p, sizeof res, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipStreamSynchronize(c.stream), "compress"))
        return nullptr;
    unsigned char *out = (unsigned char *)malloc(res[1] ? res[1] : 1);
    if (!out) {
        set_err("malloc failed");
        return nullptr;
    }
    if (!hip_ok(hipMemcpy(out, (uint8_t *)c.out.p + res[0], res[1], hipMemcpyDeviceToHost), "D2H stream")) {
        free(out);
        return nullptr;
    }
    if (outSize) *outSize = res[1];
    return out;
}

void *host_decompress_nb(const unsigned char *bytes, size_t len, size_t n, int w)
{
    if (!bytes || n == 0) {
        set_err("empty input");
        return nullptr;
    }
    HostCtx &c = g_ctx;
    if (!c.ready()) return nullptr;
    if (!c.in.reserve(len ? len : 1) || !c.out.reserve(n * (size_t)w) || !c.ws.reserve(nb_workspace(n)) ||
        !c.res.reserve(64))
        return nullptr;
    if (!hip_ok(hipMemcpyAsync(c.in.p, bytes, len, hipMemcpyHostToDevice, c.stream), "H2D"))
        return nullptr;
    nb_decompress_launch((const uint8_t *)c.in.p, len, n, w, c.out.p, (uint8_t *)c.ws.p, 0, (int32_t *)c.res.p,
                         c.stream);
    int32_t status = 0;
    if (!hip_ok(hipMemcpyAsync(&status, c.res.p, 4, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipStreamSynchronize(c.stream), "decompress"))
        return nullptr;
    if (status != 0) {
        set_err("malformed stream (code %d)", status);
        return nullptr;
    }
    void *out = malloc(n * (size_t)w);
    if (!out) {
        set_err("malloc failed");
        return nullptr;
    }
    if (!hip_ok(hipMemcpy(out, c.out.p, n * (size_t)w, hipMemcpyDeviceToHost), "D2H output")) {
        free(out);
        return nullptr;
    }
    return out;
}

// Bytes of a cuSZx stream for n values, from its header and the i16 count
// of every record (the reference's decoder walks the same fields,
// cuszx_entry.cu:134-172, and takes no length).
bool cuszx_stream_length(const unsigned char *s, size_t n, size_t *len)
{
    const int bs = s[4];
    if (bs == 0 || bs % 32 != 0 || bs > 224) {
        set_err("cuSZx blockSize %d is not a multiple of 32 in 32..224", bs);
        return false;
    }
    const size_t nb = n / (size_t)bs;
    uint64_t nconst = 0;
    for (int k = 0; k < 8; k++) nconst = (nconst << 8) | s[5 + k];
    if (nconst > nb) {
        set_err("cuSZx stream: %llu constant blocks of %llu", (unsigned long long)nconst,
                (unsigned long long)nb);
        return false;
    }
    const size_t nc = nb - (size_t)nconst, msize = 5 + (size_t)bs / 4;
    const size_t q_off = 13 + (nb + 7) / 8 + 4 * (n % (size_t)bs) + 4 * (size_t)nconst;
    size_t total = q_off + 2 * nc;
    for (size_t j = 0; j < nc; j++) {
        const int16_t sh = (int16_t)((uint16_t)s[q_off + 2 * j] | (uint16_t)s[q_off + 2 * j + 1] << 8);
        const long leng = (long)sh + (long)msize;
        if (leng < 0 || leng > 4L * bs) {   // cuszx_entry.cu:165-169
            set_err("cuSZx stream: record %llu is %ld bytes, larger than its block",
                    (unsigned long long)j, leng);
            return false;
        }
        total += (size_t)leng;
    }
    *len = total;
    return true;
}

// Length of a fastMode-2 stream from its own structure (header, bitmap,
// constant medians, then each record's size from its lead codes).
size_t stream_bytes(const unsigned char *s, size_t nbEle, int w)
{
    const size_t bs = s[4] ? s[4] : 256;
    const size_t nb = (nbEle + bs - 1) / bs, nbit = (nb + 7) / 8;
    uint64_t nconst = 0;
    for (int k = 0; k < 8; k++) nconst = (nconst << 8) | s[5 + k];
    if (nconst > nb) {
        set_err("corrupt stream header");
        return 0;
    }
    size_t q = 13 + nbit + 4 * nconst;
    for (size_t b = 0; b < nb; b++) {
        if (!((s[13 + (b >> 3)] >> (7 - (b & 7))) & 1)) continue;
        const size_t nel = (b + 1) * bs <= nbEle ? bs : nbEle - b * bs;
        const int rl = s[q];
        if (rl == 0 || rl > 8 * w) {
            set_err("corrupt record at byte %zu", q);
            return 0;
        }
        const int rb = (rl + 7) / 8;
        size_t sz = 5 + (nel + 3) / 4;
        for (size_t i = 0; i < nel; i++) {
            const int ld = (s[q + 5 + (i >> 2)] >> (6 - 2 * (int)(i & 3))) & 3;
            if (ld < rb) sz += (size_t)(rb - ld);
        }
        q += sz;
    }
    return q;
}

// Length of a random-access container (fastMode 3 / 4) for nbEle values from
// its own header: [1,1,1,1], big-endian u64 blockSize and constant count,
// the native u16 record sizes, bitmap, constant medians, records
// (szxd_float.cc:313-366 walks the same fields).  0 on a malformed header.
size_t ra_stream_bytes(const unsigned char *s, size_t nbEle, int w)
{
    uint64_t bs = 0, nconst = 0;
    for (int k = 0; k < 8; k++) bs = (bs << 8) | s[4 + k];
    for (int k = 0; k < 8; k++) nconst = (nconst << 8) | s[12 + k];
    if (s[3] != 1 || bs < 1 || bs > 256) {
        set_err("not a random-access container, or blockSize %llu unsupported", (unsigned long long)bs);
        return 0;
    }
    const size_t nb = (nbEle + bs - 1) / bs, nbit = (nb + 7) / 8;
    if (nconst > nb) {
        set_err("corrupt container header");
        return 0;
    }
    const size_t nrec = nb - nconst;
    const size_t rec_max = 5 + (bs + 3) / 4 + (size_t)bs * w;
    size_t q = 20 + 2 * nrec + nbit + 4 * nconst;
    for (size_t j = 0; j < nrec; j++) {
        const size_t sz = (size_t)s[20 + 2 * j] | (size_t)s[21 + 2 * j] << 8;
        if (sz < 5 || sz > rec_max) {
            set_err("corrupt record size %zu at record %zu", sz, j);
            return 0;
        }
        q += sz;
    }
    return q;
}

}  // namespace

namespace szxamd {
void api_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
void api_clear_error() { g_err.clear(); }

bool timing_set(bool on)
{
    const bool was = g_tim.on;
    g_tim.on = on;
    return was;
}
void timing_mark_begin(int kind, hipStream_t st)
{
    TimingState &t = g_tim;
    if (!t.on || kind < 0 || kind >= KK_COUNT) return;
    hipEvent_t e = t.get();
    if (e && hipEventRecord(e, st) == hipSuccess) t.pending[kind] = e;
}
void timing_mark_end(int kind, hipStream_t st)
{
    TimingState &t = g_tim;
    if (!t.on || kind < 0 || kind >= KK_COUNT || !t.pending[kind] || t.npairs >= 4096) return;
    hipEvent_t e = t.get();
    if (e && hipEventRecord(e, st) == hipSuccess)
        t.pairs[t.npairs++] = {t.pending[kind], e, kind};
    t.pending[kind] = nullptr;
}
}  // namespace szxamd

extern "C" {

void szx_amd_timing_begin(void)
{
    g_tim.on = true;
    g_tim.npairs = 0;
    g_tim.used = 0;
    for (int k = 0; k < KK_COUNT; k++) g_tim.pending[k] = nullptr;
}

int szx_amd_timing_end(double *ms_sum, int *count)
{
    TimingState &t = g_tim;
    t.on = false;
    for (int k = 0; k < KK_COUNT; k++) {
        ms_sum[k] = 0.0;
        count[k] = 0;
    }
    for (int i = 0; i < t.npairs; i++) {
        if (!hip_ok(hipEventSynchronize(t.pairs[i].b), "hipEventSynchronize")) return -1;
        float ms = 0.0f;
        if (!hip_ok(hipEventElapsedTime(&ms, t.pairs[i].a, t.pairs[i].b), "hipEventElapsedTime"))
            return -1;
        ms_sum[t.pairs[i].kind] += ms;
        count[t.pairs[i].kind] += 1;
    }
    t.npairs = 0;
    return 0;
}

const char *szx_amd_last_error(void) { return g_err.c_str(); }

size_t SZx_computeDataLength(size_t r5, size_t r4, size_t r3, size_t r2, size_t r1)
{
    return data_length(r5, r4, r3, r2, r1);
}

unsigned char *SZx_fast_compress_args(int fastMode, int dataType, void *data, size_t *outSize,
                                      int errBoundMode, float absErrBound, float relBoundRatio,
                                      float compressionRatio, float tolerance, size_t r5,
                                      size_t r4, size_t r3, size_t r2, size_t r1)
{
    g_err.clear();
    if (outSize) *outSize = 0;
    const int w = width_of(dataType);
    if (!w) {
        set_err("unsupported dataType %d", dataType);
        return nullptr;
    }
    const size_t n = data_length(r5, r4, r3, r2, r1);
    // szx.cc:234: the blocked family always uses blockSize 128.
    if (errBoundMode == FXR) {
        // szx.cc:238-243, 293-296: fixed rate takes the blocked fixed-rate
        // compressor whatever fastMode says
        const FxrReq req{compressionRatio, tolerance};
        return host_compress(data, n, w, 0, 0.0f, 0.0f, 128, outSize, &req);
    }
    // szx.cc:245-268: fastMode 2 is the blocked stream, 3 and 4 the
    // random-access container (3: serial layout, 4: the OpenMP one); any other
    // fastMode takes the non-blocked codec (szx.cc:270-287)
    if (errBoundMode != ABS && errBoundMode != REL) {
        set_err("errBoundMode %d is not provided", errBoundMode);
        return nullptr;
    }
    if (fastMode != SZx_WITH_BLOCK_FAST_CMPR && fastMode != SZx_RANDOMACCESS_FAST_CMPR &&
        fastMode != SZx_OPENMP_FAST_CMPR)
        return host_compress_nb(data, n, w, errBoundMode == REL, absErrBound, relBoundRatio, outSize);
    return host_compress(data, n, w, errBoundMode == REL, absErrBound, relBoundRatio, 128, outSize,
                         nullptr, fastMode);
}

unsigned char *SZ_fast_compress_args(int fastMode, int dataType, void *data, size_t *outSize,
                                     int errBoundMode, float absErrBound, float relBoundRatio,
                                     float compressionRatio, float tolerance, size_t r5, size_t r4,
                                     size_t r3, size_t r2, size_t r1)
{
    return SZx_fast_compress_args(fastMode, dataType, data, outSize, errBoundMode, absErrBound,
                                  relBoundRatio, compressionRatio, tolerance, r5, r4, r3, r2, r1);
}

void *SZx_fast_decompress(int fastMode, int dataType, unsigned char *bytes, size_t byteLength,
                          size_t r5, size_t r4, size_t r3, size_t r2, size_t r1)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w) {
        set_err("unsupported dataType %d", dataType);
        return nullptr;
    }
    // szx.cc:391-403: fastMode 1 reads the non-blocked stream, 3 and 4 the
    // random-access container
    if (fastMode == SZx_NO_BLOCK_FAST_CMPR)
        return host_decompress_nb(bytes, byteLength, data_length(r5, r4, r3, r2, r1), w);
    if (fastMode != SZx_WITH_BLOCK_FAST_CMPR && fastMode != SZx_RANDOMACCESS_FAST_CMPR &&
        fastMode != SZx_OPENMP_FAST_CMPR) {
        set_err("fastMode %d is not provided (fastMode 1, 2, 3 and 4 are)", fastMode);
        return nullptr;
    }
    return host_decompress(bytes, byteLength, data_length(r5, r4, r3, r2, r1), w, 0, fastMode);
}

void *SZ_fast_decompress(int fastMode, int dataType, unsigned char *bytes, size_t byteLength,
                         size_t r5, size_t r4, size_t r3, size_t r2, size_t r1)
{
    return SZx_fast_decompress(fastMode, dataType, bytes, byteLength, r5, r4, r3, r2, r1);
}

unsigned char *SZx_fast_compress_args_unpredictable_blocked(void *oriData, int dataType,
                                                            size_t *outSize, float absErrBound,
                                                            size_t nbEle, int blockSize)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w) {
        set_err("unsupported dataType %d", dataType);
        return nullptr;
    }
    return host_compress(oriData, nbEle, w, 0, absErrBound, 0.0f, blockSize, outSize);
}

void SZx_fast_decompress_args_unpredictable_blocked(void *newData, int dataType, size_t nbEle,
                                                    unsigned char *cmpBytes)
{
    g_err.clear();
    void **dst = (void **)newData;
    if (!dst) {
        set_err("newData is NULL");
        return;
    }
    *dst = nullptr;
    const int w = width_of(dataType);
    if (!w || !cmpBytes) {
        set_err("unsupported dataType %d or NULL stream", dataType);
        return;
    }
    // The signature carries no length: walk the stream's own structure.
    const size_t len = stream_bytes(cmpBytes, nbEle, w);
    if (!len) return;
    *dst = host_decompress(cmpBytes, len, nbEle, w, 0);
}

double SZx_computeValueRange(void *oriData, int dataType, size_t length, float *radius,
                             float *medianValue)
{
    // szx_float.cc:478-499 / szx_double.cc:1137-1158: min / max on the device,
    // then the reference's scalar arithmetic (in T, radius and median rounded
    // to float) on the two extremes.
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !oriData || length == 0) {
        set_err("unsupported dataType %d or empty input", dataType);
        return 0.0;
    }
    HostCtx &c = g_ctx;
    if (!c.ready() || !c.in.reserve(length * (size_t)w) || !c.ws.reserve(range_workspace_bytes()) ||
        !c.res.reserve(64))
        return 0.0;
    double mm[2] = {0, 0};
    if (!hip_ok(hipMemcpyAsync(c.in.p, oriData, length * (size_t)w, hipMemcpyHostToDevice, c.stream),
                "H2D"))
        return 0.0;
    range_launch(c.in.p, length, w, (uint8_t *)c.ws.p, c.res.p, c.stream);
    if (!hip_ok(hipMemcpyAsync(mm, c.res.p, 2 * (size_t)w, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipStreamSynchronize(c.stream), "value range"))
        return 0.0;
    if (w == 4) {
        float lo, hi;
        memcpy(&lo, (const char *)mm, 4);
        memcpy(&hi, (const char *)mm + 4, 4);
        const float range = hi - lo;
        if (radius) {
            *radius = range / 2;
            *medianValue = lo + *radius;
        }
        return range;
    }
    const double range = mm[1] - mm[0];
    if (radius) {
        *radius = range / 2;
        *medianValue = mm[0] + *radius;
    }
    return range;
}

float SZx_estimateErrorBoundbasedonCR(int dataType, float targetCompressionRatio, float tolerance,
                                      void *data, float initErrorBound, int blockSize,
                                      size_t nbEle)
{
    // The reference's public entry (szx.cc:458-464) reaches an estimator that
    // ends the process in its first call (szx_float.cc:680).  This runs the
    // search its fixed-rate compressor uses (szx_float.cc:684-832) from the
    // caller's initErrorBound instead.
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !data || nbEle == 0) {
        set_err("unsupported dataType %d or empty input", dataType);
        return 0.0f;
    }
    if (blockSize < 1 || blockSize > 256) {
        set_err("blockSize %d is not supported (1..256)", blockSize);
        return 0.0f;
    }
    HostCtx &c = g_ctx;
    if (!c.ready() || !c.in.reserve(nbEle * (size_t)w) ||
        !c.fx.reserve(fxr_layout(nbEle, w, blockSize, FXR_RATE).ws_bytes))
        return 0.0f;
    if (!hip_ok(hipMemcpyAsync(c.in.p, data, nbEle * (size_t)w, hipMemcpyHostToDevice, c.stream), "H2D"))
        return 0.0f;
    float eb = 0.0f;
    if (!fxr_bound_device(c.in.p, nbEle, w, blockSize, targetCompressionRatio, tolerance, false,
                          initErrorBound, (uint8_t *)c.fx.p, c.stream, &eb, nullptr))
        return 0.0f;
    return eb;
}

size_t szx_amd_compress_bound(size_t n, int dataType, int blockSize)
{
    const int w = width_of(dataType);
    if (!w || blockSize < 1 || blockSize > 256) return 0;
    const size_t nb = (n + blockSize - 1) / blockSize;
    return 13 + (nb + 7) / 8 + nb * (5 + (blockSize + 3) / 4 + (size_t)blockSize * w);
}

size_t szx_amd_compress_capacity(size_t n, int dataType, int blockSize)
{
    const int w = width_of(dataType);
    if (!w || !bs_supported(blockSize)) return 0;
    return compress_layout(n, w, blockSize).capacity;
}

size_t szx_amd_compress_workspace(size_t n, int dataType, int blockSize)
{
    const int w = width_of(dataType);
    if (!w || !bs_supported(blockSize)) return 0;
    return compress_layout(n, w, blockSize).ws_bytes;
}

size_t szx_amd_index_words(size_t n, int dataType, int blockSize)
{
    const int w = width_of(dataType);
    return (w && bs_supported(blockSize)) ? index_words(n, w, blockSize) : 0;
}

static int compress_dev(int fastMode, const void *d_in, size_t n, int dataType, int errBoundMode,
                        float absErrBound, float relBoundRatio, int blockSize, void *d_out,
                        size_t outCapacity, void *d_work, size_t workBytes, uint64_t *d_result,
                        uint64_t *d_index, void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w) { set_err("unsupported dataType %d", dataType); return SZx_TERR; }
    if (!bs_supported(blockSize)) { set_err("unsupported blockSize %d", blockSize); return SZx_NSCS; }
    if (fastMode != SZx_WITH_BLOCK_FAST_CMPR && !bs_tiled(blockSize)) {
        set_err("the random-access container takes blockSize 4..256, powers of two");
        return SZx_NSCS;
    }
    if (errBoundMode != ABS && errBoundMode != REL) { set_err("bad errBoundMode"); return SZx_BERR; }
    if (!d_in || !d_out || !d_work || !d_result || n == 0) { set_err("NULL argument or n == 0"); return SZx_NSCS; }
    if (((uintptr_t)d_in & 15) != 0) { set_err("d_in must be 16-byte aligned"); return SZx_NSCS; }
    const CompressLayout L = compress_layout(n, w, blockSize);
    if (outCapacity < L.capacity) { set_err("outCapacity < %llu", (unsigned long long)L.capacity); return SZx_NSCS; }
    if (workBytes < L.ws_bytes) { set_err("workBytes < %llu", (unsigned long long)L.ws_bytes); return SZx_NSCS; }
    if (compress_launch(d_in, n, w, errBoundMode == REL, absErrBound, relBoundRatio, blockSize,
                        (uint8_t *)d_out, (uint8_t *)d_work, d_result, d_index,
                        (hipStream_t)hipStream, fastMode))
        return SZx_NSCS;
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

int szx_amd_compress_dev(const void *d_in, size_t n, int dataType, int errBoundMode,
                         float absErrBound, float relBoundRatio, int blockSize, void *d_out,
                         size_t outCapacity, void *d_work, size_t workBytes, uint64_t *d_result,
                         uint64_t *d_index, void *hipStream)
{
    return compress_dev(SZx_WITH_BLOCK_FAST_CMPR, d_in, n, dataType, errBoundMode, absErrBound,
                        relBoundRatio, blockSize, d_out, outCapacity, d_work, workBytes, d_result, d_index,
                        hipStream);
}

int szx_amd_compress_ra_dev(int fastMode, const void *d_in, size_t n, int dataType, int errBoundMode,
                            float absErrBound, float relBoundRatio, int blockSize, void *d_out,
                            size_t outCapacity, void *d_work, size_t workBytes, uint64_t *d_result,
                            uint64_t *d_index, void *hipStream)
{
    if (fastMode != SZx_RANDOMACCESS_FAST_CMPR && fastMode != SZx_OPENMP_FAST_CMPR) {
        g_err.clear();
        set_err("fastMode %d is not a random-access mode (3 or 4)", fastMode);
        return SZx_NSCS;
    }
    return compress_dev(fastMode, d_in, n, dataType, errBoundMode, absErrBound, relBoundRatio, blockSize,
                        d_out, outCapacity, d_work, workBytes, d_result, d_index, hipStream);
}

size_t szx_amd_decompress_workspace(size_t n, int dataType, int blockSize, size_t streamBytes)
{
    const int w = width_of(dataType);
    if (!w || !bs_supported(blockSize)) return 0;
    return decompress_layout(n, w, blockSize, streamBytes).ws_bytes;
}

static int decompress_dev(int fastMode, const void *d_stream, size_t streamBytes, size_t n, int dataType,
                          int blockSize, void *d_out, void *d_work, size_t workBytes,
                          const uint64_t *d_index, int flags, int32_t *d_status, void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w) { set_err("unsupported dataType %d", dataType); return SZx_TERR; }
    if (!bs_supported(blockSize)) { set_err("unsupported blockSize %d", blockSize); return SZx_NSCS; }
    if (fastMode != SZx_WITH_BLOCK_FAST_CMPR && !bs_tiled(blockSize)) {
        set_err("the random-access container takes blockSize 4..256, powers of two");
        return SZx_NSCS;
    }
    if (!d_stream || !d_out || !d_work || !d_status || n == 0) { set_err("NULL argument or n == 0"); return SZx_NSCS; }
    if (((uintptr_t)d_out & 15) != 0) { set_err("d_out must be 16-byte aligned"); return SZx_NSCS; }
    if (decompress_launch((const uint8_t *)d_stream, streamBytes, n, w, blockSize, d_out,
                          (uint8_t *)d_work, workBytes, d_index, flags & SZX_AMD_DECODE_EXACT,
                          d_status, (hipStream_t)hipStream, fastMode)) {
        set_err("workBytes too small or bad parameters");
        return SZx_NSCS;
    }
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

int szx_amd_decompress_dev(const void *d_stream, size_t streamBytes, size_t n, int dataType,
                           int blockSize, void *d_out, void *d_work, size_t workBytes,
                           const uint64_t *d_index, int flags, int32_t *d_status, void *hipStream)
{
    return decompress_dev(SZx_WITH_BLOCK_FAST_CMPR, d_stream, streamBytes, n, dataType, blockSize, d_out,
                          d_work, workBytes, d_index, flags, d_status, hipStream);
}

int szx_amd_decompress_ra_dev(int fastMode, const void *d_stream, size_t streamBytes, size_t n,
                              int dataType, int blockSize, void *d_out, void *d_work, size_t workBytes,
                              const uint64_t *d_index, int flags, int32_t *d_status, void *hipStream)
{
    if (fastMode != SZx_RANDOMACCESS_FAST_CMPR && fastMode != SZx_OPENMP_FAST_CMPR) {
        g_err.clear();
        set_err("fastMode %d is not a random-access mode (3 or 4)", fastMode);
        return SZx_NSCS;
    }
    return decompress_dev(fastMode, d_stream, streamBytes, n, dataType, blockSize, d_out, d_work, workBytes,
                          d_index, flags, d_status, hipStream);
}

size_t szx_amd_nb_capacity(size_t n, int dataType)
{
    const int w = width_of(dataType);
    return w && n ? nb_capacity(n, w) : 0;
}

size_t szx_amd_nb_workspace(size_t n, int dataType)
{
    const int w = width_of(dataType);
    return w && n ? nb_workspace(n) : 0;
}

int szx_amd_compress_nb_dev(const void *d_in, size_t n, int dataType, int errBoundMode, float absErrBound,
                            float relBoundRatio, void *d_out, size_t outCapacity, void *d_work, size_t workBytes,
                            uint64_t *d_result, void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w) { set_err("unsupported dataType %d", dataType); return SZx_TERR; }
    if (errBoundMode != ABS && errBoundMode != REL) { set_err("bad errBoundMode"); return SZx_BERR; }
    if (!d_in || !d_out || !d_work || !d_result || n == 0) { set_err("NULL argument or n == 0"); return SZx_NSCS; }
    if (outCapacity < nb_capacity(n, w) || workBytes < nb_workspace(n)) {
        set_err("output buffer or workspace too small");
        return SZx_NSCS;
    }
    nb_compress_launch(d_in, n, w, errBoundMode == REL, absErrBound, relBoundRatio, (uint8_t *)d_out,
                       (uint8_t *)d_work, d_result, (hipStream_t)hipStream);
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

int szx_amd_decompress_nb_dev(const void *d_stream, size_t streamBytes, size_t n, int dataType, void *d_out,
                              void *d_work, size_t workBytes, int flags, int32_t *d_status, void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w) { set_err("unsupported dataType %d", dataType); return SZx_TERR; }
    if (!d_stream || !d_out || !d_work || !d_status || n == 0) { set_err("NULL argument or n == 0"); return SZx_NSCS; }
    if (workBytes < nb_workspace(n)) { set_err("workspace too small"); return SZx_NSCS; }
    nb_decompress_launch((const uint8_t *)d_stream, streamBytes, n, w, d_out, (uint8_t *)d_work,
                         flags & SZX_AMD_DECODE_EXACT, d_status, (hipStream_t)hipStream);
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

void cuSZx_fast_decompress_args_unpredictable_blocked_float(float **newData, size_t nbEle,
                                                            unsigned char *cmpBytes)
{
    g_err.clear();
    if (!newData) {
        set_err("NULL newData");
        return;
    }
    *newData = nullptr;
    size_t len = 0;
    if (!cmpBytes || nbEle == 0) {
        set_err("empty input");
        return;
    }
    if (!cuszx_stream_length(cmpBytes, nbEle, &len)) return;
    HostCtx &c = g_ctx;
    if (!c.ready()) return;
    if (!c.in.reserve(len) || !c.out.reserve(nbEle * 4) || !c.ws.reserve(cuszx_workspace(nbEle)) ||
        !c.res.reserve(64))
        return;
    if (!hip_ok(hipMemcpyAsync(c.in.p, cmpBytes, len, hipMemcpyHostToDevice, c.stream), "H2D"))
        return;
    cuszx_decompress_launch((const uint8_t *)c.in.p, len, nbEle, (float *)c.out.p, (uint8_t *)c.ws.p,
                            (int32_t *)c.res.p, c.stream);
    int32_t status = 0;
    if (!hip_ok(hipMemcpyAsync(&status, c.res.p, 4, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipStreamSynchronize(c.stream), "decompress"))
        return;
    if (status != 0) {
        set_err("malformed cuSZx stream (code %d)", status);
        return;
    }
    float *out = (float *)malloc(nbEle * 4);
    if (!out) {
        set_err("malloc failed");
        return;
    }
    if (!hip_ok(hipMemcpy(out, c.out.p, nbEle * 4, hipMemcpyDeviceToHost), "D2H output")) {
        free(out);
        return;
    }
    *newData = out;
}

static bool cuszx_bs_ok(int bs) { return bs >= 32 && bs <= 224 && bs % 32 == 0; }

unsigned char *cuSZx_fast_compress_args_unpredictable_blocked_float(float *oriData, size_t *outSize,
                                                                     float absErrBound, size_t nbEle,
                                                                     int blockSize)
{
    g_err.clear();
    if (outSize) *outSize = 0;
    if (!oriData || nbEle == 0) {
        set_err("empty input");
        return nullptr;
    }
    if (!cuszx_bs_ok(blockSize)) {
        set_err("cuSZx blockSize %d is not a multiple of 32 in 32..224", blockSize);
        return nullptr;
    }
    HostCtx &c = g_ctx;
    if (!c.ready()) return nullptr;
    const uint64_t cap = cuszx_capacity(nbEle, blockSize);
    if (!c.in.reserve(nbEle * 4) || !c.out.reserve(cap) || !c.ws.reserve(cuszx_compress_workspace(nbEle, blockSize)) ||
        !c.res.reserve(64))
        return nullptr;
    if (!hip_ok(hipMemcpyAsync(c.in.p, oriData, nbEle * 4, hipMemcpyHostToDevice, c.stream), "H2D"))
        return nullptr;
    cuszx_compress_launch((const float *)c.in.p, nbEle, absErrBound, blockSize, (uint8_t *)c.out.p,
                          (uint8_t *)c.ws.p, (uint64_t *)c.res.p, c.stream);
    uint64_t res[4] = {0, 0, 0, 0};
    if (!hip_ok(hipMemcpyAsync(res, c.res.p, sizeof res, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipStreamSynchronize(c.stream), "compress"))
        return nullptr;
    if (res[3]) {
        set_err("cuSZx: a block needs more than 32 significant bits for this bound");
        return nullptr;
    }
    unsigned char *out = (unsigned char *)malloc(res[1]);
    if (!out) {
        set_err("malloc failed");
        return nullptr;
    }
    if (!hip_ok(hipMemcpy(out, c.out.p, res[1], hipMemcpyDeviceToHost), "D2H stream")) {
        free(out);
        return nullptr;
    }
    if (outSize) *outSize = res[1];
    return out;
}

size_t szx_amd_cuszx_capacity(size_t nbEle, int blockSize)
{
    return nbEle && cuszx_bs_ok(blockSize) ? cuszx_capacity(nbEle, blockSize) : 0;
}

size_t szx_amd_cuszx_compress_workspace(size_t nbEle, int blockSize)
{
    return nbEle && cuszx_bs_ok(blockSize) ? cuszx_compress_workspace(nbEle, blockSize) : 0;
}

int szx_amd_compress_cuszx_dev(const float *d_in, size_t nbEle, float absErrBound, int blockSize, void *d_out,
                               size_t outCapacity, void *d_work, size_t workBytes, uint64_t *d_result,
                               void *hipStream)
{
    g_err.clear();
    if (!d_in || !d_out || !d_work || !d_result || nbEle == 0) {
        set_err("NULL argument or nbEle == 0");
        return SZx_NSCS;
    }
    if (!cuszx_bs_ok(blockSize)) {
        set_err("cuSZx blockSize %d is not a multiple of 32 in 32..224", blockSize);
        return SZx_NSCS;
    }
    if (outCapacity < cuszx_capacity(nbEle, blockSize) || workBytes < cuszx_compress_workspace(nbEle, blockSize)) {
        set_err("output buffer or workspace too small");
        return SZx_NSCS;
    }
    cuszx_compress_launch(d_in, nbEle, absErrBound, blockSize, (uint8_t *)d_out, (uint8_t *)d_work, d_result,
                          (hipStream_t)hipStream);
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

size_t szx_amd_cuszx_workspace(size_t nbEle) { return nbEle ? cuszx_workspace(nbEle) : 0; }

int szx_amd_decompress_cuszx_dev(const void *d_stream, size_t streamBytes, size_t nbEle, float *d_out,
                                 void *d_work, size_t workBytes, int32_t *d_status, void *hipStream)
{
    g_err.clear();
    if (!d_stream || !d_out || !d_work || !d_status || nbEle == 0) {
        set_err("NULL argument or nbEle == 0");
        return SZx_NSCS;
    }
    if (workBytes < cuszx_workspace(nbEle)) {
        set_err("workspace too small");
        return SZx_NSCS;
    }
    cuszx_decompress_launch((const uint8_t *)d_stream, streamBytes, nbEle, d_out, (uint8_t *)d_work, d_status,
                            (hipStream_t)hipStream);
    return hip_ok(hipGetLastError(), "launch") ? SZx_SCES : SZx_NSCS;
}

size_t szx_amd_fxr_workspace(size_t n, int dataType, int blockSize)
{
    const int w = width_of(dataType);
    if (!w || blockSize < 1 || blockSize > 256) return 0;
    return fxr_layout(n, w, blockSize, FXR_RATE).ws_bytes;
}

int szx_amd_fxr_bound_dev(const void *d_in, size_t n, int dataType, int blockSize,
                          float targetCompressionRatio, float tolerance, float initErrorBound,
                          void *d_work, size_t workBytes, float *errorBound, int *modelPasses,
                          void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !d_in || n == 0 || !errorBound || blockSize < 1 || blockSize > 256) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    if (workBytes < fxr_layout(n, w, blockSize, FXR_RATE).ws_bytes) {
        set_err("workspace too small");
        return SZx_MERR;
    }
    const bool dflt = !(initErrorBound > 0.0f);
    if (!fxr_bound_device(d_in, n, w, blockSize, targetCompressionRatio, tolerance, dflt,
                          initErrorBound, (uint8_t *)d_work, (hipStream_t)hipStream, errorBound,
                          modelPasses))
        return SZx_NSCS;
    return SZx_SCES;
}

size_t szx_amd_fixed_rate_workspace(size_t n, int dataType, int blockSize)
{
    const int w = width_of(dataType);
    if (!w || !bs_tiled(blockSize)) return 0;
    return fixed_rate_workspace(n, w, blockSize);
}

int szx_amd_fixed_rate_dev(const void *d_in, size_t n, int dataType, int blockSize,
                           float targetBitsPerValue, float tolerance, int method,
                           const float *d_modelParams, void *d_out, size_t outCapacity,
                           void *d_work, size_t workBytes, uint64_t *d_result, uint64_t *d_index,
                           szx_amd_rate_info *info, void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !d_in || n == 0 || !d_out || !d_result || !bs_tiled(blockSize) ||
        !(targetBitsPerValue > 0) || !(tolerance >= 0) ||
        (method != SZX_AMD_RATE_BISECT && method != SZX_AMD_RATE_LEARNED) ||
        (method == SZX_AMD_RATE_LEARNED && !d_modelParams)) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    if (outCapacity < compress_layout(n, w, blockSize).capacity ||
        workBytes < fixed_rate_workspace(n, w, blockSize)) {
        set_err("output buffer or workspace too small");
        return SZx_MERR;
    }
    const char *err = nullptr;
    if (fixed_rate_solve(d_in, n, w, blockSize, targetBitsPerValue, tolerance, method, d_modelParams,
                         (uint8_t *)d_out, (uint8_t *)d_work, d_result, d_index, info,
                         (hipStream_t)hipStream, &err)) {
        set_err("fixed rate: %s", err ? err : "device error");
        return SZx_NSCS;
    }
    return SZx_SCES;
}

size_t szx_amd_rate_octave_workspace(void) { return octave_workspace_bytes(); }

int szx_amd_rate_octave_dev(const void *d_in, size_t n, int dataType, int blockSize, int reqExpo,
                            uint64_t *d_acc, void *d_work, size_t workBytes, void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !d_in || n == 0 || !d_acc || blockSize < 1 || blockSize > 256 ||
        workBytes < octave_workspace_bytes() || reqExpo < -1074 || reqExpo > 1023) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    rate_octave_launch(d_in, n, w, blockSize, reqExpo, (uint8_t *)d_work,
                       (unsigned long long *)d_acc, (hipStream_t)hipStream);
    return SZx_SCES;
}

int szx_amd_rate_octave_span_dev(const void *d_in, size_t n, int dataType, int blockSize, int reqExpo,
                                 int nOctaves, uint64_t *d_acc, void *d_work, size_t workBytes,
                                 void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !d_in || n == 0 || !d_acc || blockSize < 1 || blockSize > 256 ||
        workBytes < octave_workspace_bytes() || nOctaves < 1 || nOctaves > 2 ||
        reqExpo < nOctaves - 149 || reqExpo > 126) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    rate_octave_launch(d_in, n, w, blockSize, reqExpo, (uint8_t *)d_work, (unsigned long long *)d_acc,
                       (hipStream_t)hipStream, true, nOctaves);
    return SZx_SCES;
}

size_t szx_amd_rate_stats_bytes(size_t n)
{
    const size_t c = (n + RATE_CHUNK - 1) / RATE_CHUNK;
    return (c ? c : 1) * sizeof(RateStats);
}

size_t szx_amd_rate_stats_workspace(size_t n) { return rate_stats_workspace_bytes(n); }

int szx_amd_rate_stats_dev(const void *d_in, size_t n, int dataType, int blockSize, void *d_stats,
                           void *d_work, size_t workBytes, void *hipStream)
{
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !d_in || n == 0 || !d_stats || blockSize < 1 || blockSize > 256 ||
        workBytes < rate_stats_workspace_bytes(n)) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    rate_stats_launch(d_in, n, w, blockSize, (uint8_t *)d_work, (RateStats *)d_stats,
                      (hipStream_t)hipStream);
    return SZx_SCES;
}

int szx_amd_rate_inputs_dev(const void *d_stats, size_t n, const float *d_ebs, int K, float *d_X,
                            void *hipStream)
{
    g_err.clear();
    if (!d_stats || !d_ebs || !d_X || K < 1 || n == 0) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    const int nch = (int)((n + RATE_CHUNK - 1) / RATE_CHUNK);
    rate_inputs_launch((const RateStats *)d_stats, nch, d_ebs, K, d_X, (hipStream_t)hipStream);
    return SZx_SCES;
}

int szx_amd_mlp_forward_dev(const float *d_params, const float *d_X, size_t nSamples, float *d_y,
                            void *hipStream)
{
    g_err.clear();
    if (!d_params || !d_X || !d_y) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    if (nSamples) mlp_forward_launch(d_params, d_X, nSamples, d_y, (hipStream_t)hipStream);
    return SZx_SCES;
}

size_t szx_amd_mlp_grad_workspace(void) { return mlp_grad_workspace_bytes(); }

int szx_amd_mlp_grad_dev(const float *d_params, const float *d_X, const float *d_T, size_t nSamples,
                         float *d_grad, float *d_loss, void *d_work, size_t workBytes,
                         void *hipStream)
{
    g_err.clear();
    if (!d_params || !d_X || !d_T || !d_grad || !d_loss || nSamples == 0 ||
        workBytes < mlp_grad_workspace_bytes()) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    mlp_grad_launch(d_params, d_X, d_T, nSamples, d_grad, d_loss, (uint8_t *)d_work,
                    (hipStream_t)hipStream);
    return SZx_SCES;
}

int szx_amd_mlp_adam_dev(float *d_params, const float *d_grad, float *d_m, float *d_v, float lr,
                         float beta1, float beta2, float eps, int step, void *hipStream)
{
    g_err.clear();
    if (!d_params || !d_grad || !d_m || !d_v || step < 1) {
        set_err("invalid arguments");
        return SZx_NSCS;
    }
    mlp_adam_launch(d_params, d_grad, d_m, d_v, lr, beta1, beta2, eps, step, (hipStream_t)hipStream);
    return SZx_SCES;
}

// ---- the rest of szx.h (szx.h:74-98) ------------------------------------------

int SZx_computeDimension(size_t r5, size_t r4, size_t r3, size_t r2, size_t r1)
{
    // szx.cc:41-71: the number of leading non-zero dims, r1 first
    if (r1 == 0) return 0;
    if (r2 == 0) return 1;
    if (r3 == 0) return 2;
    if (r4 == 0) return 3;
    if (r5 == 0) return 4;
    return 5;
}

int SZx_filterDimension(size_t r5, size_t r4, size_t r3, size_t r2, size_t r1, size_t *correctedDimension)
{
    // szx.cc:109-229: c[] = {r1..r5}; each dim of extent 1 is dropped (tested
    // from the slowest present dim down to r1), the dims above it shifting
    // down one place.  Returns 1 if a dim was dropped, 0 if not, 2 for a 1-D
    // array of no values.  The reference's shifts are kept as they are
    // (which slot a removal clears differs between ranks).
    const int dim = SZx_computeDimension(r5, r4, r3, r2, r1);
    size_t *c = correctedDimension;
    c[0] = r1;
    c[1] = r2;
    c[2] = r3;
    c[3] = r4;
    c[4] = r5;
    const size_t r[5] = {r1, r2, r3, r4, r5};
    if (dim == 1) return r1 < 1 ? 2 : 0;
    if (dim < 2) return 0;
    int changed = 0;
    // the slowest dim of extent 1 is cleared in place
    if (r[dim - 1] == 1) {
        c[dim - 1] = 0;
        changed = 1;
    }
    // every faster dim of extent 1 is removed by shifting the slower ones down
    // over it; the vacated top slot becomes 0 (for ranks below 5 the
    // reference shifts the absent dim's 0 into it)
    for (int d = dim - 2; d >= 0; d--) {
        if (r[d] != 1) continue;
        for (int k = d; k < dim - 1; k++) c[k] = c[k + 1];
        c[dim - 1] = 0;
        changed = 1;
    }
    return changed;
}

size_t SZx_fast_compress_args_unpredictable_blocked_args(void *oriData, int dataType, unsigned char *outputBytes,
                                                         float absErrBound, size_t nbEle, int blockSize)
{
    // szx.cc:501-512 -> szx_float.cc:980-1057: the blocked stream written to
    // the caller's buffer, its size returned.  The reference's double branch
    // returns 0 ("not implemented yet", with the double call commented out);
    // here it writes the double stream (szx_double.cc:1160-1238 bytes).
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !outputBytes) {
        set_err("unsupported dataType %d or NULL output", dataType);
        return 0;
    }
    size_t out = 0;
    if (!host_compress(oriData, nbEle, w, 0, absErrBound, 0.0f, blockSize, &out, nullptr,
                       SZx_WITH_BLOCK_FAST_CMPR, outputBytes))
        return 0;
    return out;
}

unsigned char *SZx_fast_compress_args_unpredictable_blocked_randomaccess_openmp(void *oriData, int dataType,
                                                                               size_t *outSize, float absErrBound,
                                                                               size_t nbEle, int blockSize)
{
    // szx.cc:441-448 -> szx_float.cc:1218-1405 / szx_double.cc:1320-1507: the
    // fastMode-4 container at an explicit blockSize (powers of two 4..256).
    g_err.clear();
    if (outSize) *outSize = 0;
    const int w = width_of(dataType);
    if (!w) {
        set_err("unsupported dataType %d", dataType);
        return nullptr;
    }
    if (!bs_tiled(blockSize)) {
        set_err("the random-access container takes blockSize 4..256, powers of two (got %d)", blockSize);
        return nullptr;
    }
    return host_compress(oriData, nbEle, w, 0, absErrBound, 0.0f, blockSize, outSize, nullptr,
                         SZx_OPENMP_FAST_CMPR);
}

void *SZx_fast_decompress_args_unpredictable_blocked_randomaccess_openmp(int dataType, size_t nbEle,
                                                                        unsigned char *cmpBytes)
{
    // szx.cc:450-456 -> szxd_float.cc:313-399 / szxd_double.cc:518-605.  The
    // signature carries no length: it is read from the container's size table.
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !cmpBytes || nbEle == 0) {
        set_err("unsupported dataType %d, NULL stream or no values", dataType);
        return nullptr;
    }
    const size_t len = ra_stream_bytes(cmpBytes, nbEle, w);
    if (!len) return nullptr;
    return host_decompress(cmpBytes, len, nbEle, w, 0, SZx_OPENMP_FAST_CMPR);
}

void *SZx_fast_decompress_pred(int dataType, float *preData, unsigned char *curBytes, size_t byteLength,
                               size_t r5, size_t r4, size_t r3, size_t r2, size_t r1)
{
    // szx.cc:353-376 -> szxd_float.cc:25-32: decode a non-blocked (fastMode 1)
    // stream and add the prediction preData[i] to every value.  The reference
    // returns NULL for double (its branch is commented out), and so does this.
    g_err.clear();
    if (dataType != SZx_FLOAT) {
        set_err("SZx_fast_decompress_pred decodes float only (the reference returns NULL for double)");
        return nullptr;
    }
    const size_t n = data_length(r5, r4, r3, r2, r1);
    if (!preData || !curBytes || n == 0) {
        set_err("NULL prediction or stream, or no values");
        return nullptr;
    }
    HostCtx &c = g_ctx;
    if (!c.ready()) return nullptr;
    if (!c.in.reserve(byteLength ? byteLength : 1) || !c.out.reserve(n * 4) || !c.ws.reserve(nb_workspace(n)) ||
        !c.res.reserve(64) || !c.fx.reserve(n * 4))
        return nullptr;
    if (!hip_ok(hipMemcpyAsync(c.in.p, curBytes, byteLength, hipMemcpyHostToDevice, c.stream), "H2D") ||
        !hip_ok(hipMemcpyAsync(c.fx.p, preData, n * 4, hipMemcpyHostToDevice, c.stream), "H2D prediction"))
        return nullptr;
    nb_decompress_launch((const uint8_t *)c.in.p, byteLength, n, 4, c.out.p, (uint8_t *)c.ws.p, 0,
                         (int32_t *)c.res.p, c.stream);
    add_pred_launch((float *)c.out.p, (const float *)c.fx.p, n, c.stream);
    int32_t status = 0;
    if (!hip_ok(hipMemcpyAsync(&status, c.res.p, 4, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipStreamSynchronize(c.stream), "decompress"))
        return nullptr;
    if (status != 0) {
        set_err("malformed stream (code %d)", status);
        return nullptr;
    }
    float *out = (float *)malloc(n * 4);
    if (!out) {
        set_err("malloc failed");
        return nullptr;
    }
    if (!hip_ok(hipMemcpy(out, c.out.p, n * 4, hipMemcpyDeviceToHost), "D2H output")) {
        free(out);
        return nullptr;
    }
    return out;
}

float computeRadiusBuffer(void *oriData, int dataType, size_t nbEle, int samplingRate, int blockSize,
                          float **radiusArray, float **mediusArray, void *buffer)
{
    // szx.cc:475-485 -> szx_float.cc:254-312 / szx_double.cc:437-495: copy
    // every samplingRate-th block into *buffer (malloc'd), each sampled
    // block's radius and median (as float) into *radiusArray / *mediusArray
    // (malloc'd, one spare entry as in the reference), and return the range
    // of the sampled values (extremes starting from oriData[0]).  The block
    // statistics run on the GPU (k_fxr_sample, k_fxr_range); the sampled
    // blocks come back with one strided copy.
    g_err.clear();
    const int w = width_of(dataType);
    void **buf = (void **)buffer;
    if (!w || !oriData || !radiusArray || !mediusArray || !buf || nbEle == 0 || samplingRate < 1 ||
        blockSize < 1 || blockSize > 256) {
        set_err("invalid arguments");
        return 0.0f;
    }
    const size_t stride = (size_t)samplingRate * blockSize, ns = nbEle / stride;
    float *rad = (float *)malloc(sizeof(float) * (ns + 1));
    float *med = (float *)malloc(sizeof(float) * (ns + 1));
    void *smp = malloc((size_t)w * (ns + 1) * blockSize);
    if (!rad || !med || !smp) {
        free(rad);
        free(med);
        free(smp);
        set_err("malloc failed");
        return 0.0f;
    }
    *radiusArray = rad;
    *mediusArray = med;
    *buf = smp;
    if (ns == 0) return 0.0f;   // no block sampled: both extremes are oriData[0]
    HostCtx &c = g_ctx;
    const FxrLayout L = fxr_layout(nbEle, w, blockSize, samplingRate);
    float range = 0.0f;
    if (!c.ready() || !c.in.reserve(nbEle * (size_t)w) || !c.fx.reserve(L.ws_bytes) ||
        !hip_ok(hipMemcpyAsync(c.in.p, oriData, nbEle * (size_t)w, hipMemcpyHostToDevice, c.stream), "H2D"))
        return 0.0f;
    fxr_sample_launch(c.in.p, nbEle, w, blockSize, samplingRate, (uint8_t *)c.fx.p, c.stream);
    if (!hip_ok(hipMemcpyAsync(rad, (uint8_t *)c.fx.p + L.ws_rad, 4 * ns, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipMemcpyAsync(med, (uint8_t *)c.fx.p + L.ws_med, 4 * ns, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipMemcpyAsync(&range, (uint8_t *)c.fx.p + L.ws_range, 4, hipMemcpyDeviceToHost, c.stream), "D2H") ||
        !hip_ok(hipMemcpy2DAsync(smp, (size_t)w * blockSize, c.in.p, (size_t)w * stride, (size_t)w * blockSize, ns,
                                 hipMemcpyDeviceToHost, c.stream), "D2H sample") ||
        !hip_ok(hipStreamSynchronize(c.stream), "radius buffer"))
        return 0.0f;
    return range;
}

float estimateCRbasedonErrorBound_buffered(int dataType, float errorBound, void *buffer, float *medianArray,
                                           float *radiusArray, int samplingRate, int blockSize, size_t nbEle,
                                           size_t *sumReqNbBytes, size_t *sum_actual_leadNumbers)
{
    // szx.cc:487-499 -> szx_float.cc:501-584 / szx_double.cc:497-582: the
    // modelled compression ratio of the nbEle / (blockSize * samplingRate)
    // buffered blocks at errorBound.  The three sums (constant blocks, sum of
    // reqBytes, sum of min(lead, reqBytes)) are computed on the GPU by the
    // same kernel the fixed-rate search uses; the ratio formula is evaluated
    // here in the reference's float arithmetic.
    g_err.clear();
    const int w = width_of(dataType);
    if (!w || !buffer || !medianArray || !radiusArray || nbEle == 0 || samplingRate < 1 || blockSize < 1 ||
        blockSize > 256) {
        set_err("invalid arguments");
        return 0.0f;
    }
    const size_t ns = nbEle / ((size_t)blockSize * samplingRate);
    const size_t nv = ns * (size_t)blockSize;
    // the buffered blocks are contiguous: a 1-in-1 sample of ns blocks
    const FxrLayout L = fxr_layout(nv, w, blockSize, 1);
    HostCtx &c = g_ctx;
    uint64_t sums[3] = {0, 0, 0};
    if (ns) {
        if (!c.ready() || !c.in.reserve(nv * (size_t)w) || !c.fx.reserve(L.ws_bytes) ||
            !hip_ok(hipMemcpyAsync(c.in.p, buffer, nv * (size_t)w, hipMemcpyHostToDevice, c.stream), "H2D") ||
            !hip_ok(hipMemcpyAsync((uint8_t *)c.fx.p + L.ws_med, medianArray, 4 * ns, hipMemcpyHostToDevice,
                                   c.stream), "H2D") ||
            !hip_ok(hipMemcpyAsync((uint8_t *)c.fx.p + L.ws_rad, radiusArray, 4 * ns, hipMemcpyHostToDevice,
                                   c.stream), "H2D") ||
            !hip_ok(hipMemcpyAsync((uint8_t *)c.fx.p + L.ws_ebs, &errorBound, 4, hipMemcpyHostToDevice, c.stream),
                    "H2D"))
            return 0.0f;
        fxr_estimate_launch(c.in.p, nv, w, blockSize, 1, (uint8_t *)c.fx.p, 1, c.stream);
        if (!hip_ok(hipMemcpyAsync(sums, (uint8_t *)c.fx.p + L.ws_sums, 24, hipMemcpyDeviceToHost, c.stream),
                    "D2H sums") ||
            !hip_ok(hipStreamSynchronize(c.stream), "model"))
            return 0.0f;
    }
    if (sumReqNbBytes) *sumReqNbBytes = sums[1];
    if (sum_actual_leadNumbers) *sum_actual_leadNumbers = sums[2];
    return fxr_model_cr(w, blockSize, nbEle, ns, FxrSums{sums[0], sums[1], sums[2]});
}

}  // extern "C"
