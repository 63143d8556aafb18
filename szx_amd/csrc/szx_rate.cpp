// szx_rate.cpp -- fixed-rate compression on the exact codec: find a bound
// whose stream has targetBitsPerValue (8 * stream bytes / n) within a
// relative tolerance, and leave that stream in the output buffer.
//
// Two solvers, both judged by the real codec (every stream is exact codec
// output; nothing is approximated in the codec arithmetic):
//
//  * bisection: bisect log2(eb) between range * 2^-D and the range, one full
//    compress per step, until the rate is within tolerance.
//
//  * learned: one statistics pass gives per-chunk features; the rate MLP
//    (szx_mlp.hip) predicts each chunk's log2 bits/value at a ladder of
//    bounds and the predicted field rate picks a bound.  Its exponent names
//    an octave of bounds; one octave pass (k_rate_octave) gives the exact
//    stream size for OCT_BINS bounds spread over that octave, so the best
//    bound in the octave is known exactly before compressing (a prediction
//    close to an octave edge sizes the octaves on both sides of it in the
//    same pass).  If the target lies outside those sizes the neighbouring
//    octave is tried (at most OCT_TRIES octave passes), then bisection
//    takes over.
//    The final compress runs at the chosen bound and its size is checked.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "szx_amd.h"
#include "szx_fxr.hpp"
#include "szx_internal.hpp"
#include "szx_mlp.hpp"
#include "szx_rate.hpp"

namespace szxamd {

namespace {

constexpr int K_MAX = 256;      // ladder of candidate bounds (quarter octaves)
constexpr int OCT_TRIES = 3;
constexpr double OCT_PAIR = 0.1;   // log2 distance from an octave edge
constexpr int BISECT_MAX = 60;
constexpr int STREAM_HEADER = 13;   // szx_float.cc:1091-1098

struct RateLayout {
    uint64_t comp, oct, acc, stats_ws, stats, ebs, X, y, range, minmax, bytes;
    uint64_t nchunks;
};

RateLayout rate_layout(uint64_t n, int w, int bs)
{
    RateLayout L{};
    const CompressLayout C = compress_layout(n, w, bs);
    L.nchunks = (n + RATE_CHUNK - 1) / RATE_CHUNK;
    if (!L.nchunks) L.nchunks = 1;
    auto al = [](uint64_t v) { return (v + 255) & ~255ull; };
    uint64_t off = 0;
    L.comp = off;     off += al(C.ws_bytes);
    L.oct = off;      off += al(octave_workspace_bytes());
    L.acc = off;      off += al((2 * OCT_ACC + 4) * 8);
    L.stats_ws = off; off += al(rate_stats_workspace_bytes(n));
    L.stats = off;    off += al(L.nchunks * sizeof(RateStats));
    L.ebs = off;      off += al(K_MAX * 4);
    L.X = off;        off += al((uint64_t)K_MAX * L.nchunks * RATE_INPUTS * 4);
    L.y = off;        off += al((uint64_t)K_MAX * L.nchunks * 4);
    L.range = off;    off += al(range_workspace_bytes());
    L.minmax = off;   off += 256;
    L.bytes = off;
    return L;
}

struct Solver {
    const void *x;
    uint64_t n;
    int w, bs;
    float target, tol;
    uint8_t *out, *ws;
    uint64_t *result, *index;
    hipStream_t st;
    RateLayout L;
    int passes = 0;
    float last_eb = -1.0f;   // bound of the stream now in the output buffer
    double last_bpv = 0.0;
    float best_eb = 0;
    double best_bpv = -1;
    const char *err = nullptr;

    bool ok(hipError_t e, const char *what)
    {
        if (e == hipSuccess) return true;
        err = what;
        return false;
    }
    bool within(double bpv) const { return fabs(bpv - target) <= (double)tol * target; }
    void note(float eb, double bpv)
    {
        if (best_bpv < 0 || fabs(bpv - target) < fabs(best_bpv - target)) {
            best_bpv = bpv;
            best_eb = eb;
        }
    }

    // One full compress at eb; returns the exact rate.
    bool compress_at(float eb, double *bpv)
    {
        if (compress_launch(x, n, w, 0, eb, 0.0f, bs, out, ws + L.comp, result, index, st)) {
            err = "compress launch";
            return false;
        }
        uint64_t r[4];
        if (!ok(hipMemcpyAsync(r, result, sizeof r, hipMemcpyDeviceToHost, st), "D2H result") ||
            !ok(hipStreamSynchronize(st), "compress"))
            return false;
        if (r[3]) {
            err = "compress: a look-back wait overran its bound";
            return false;
        }
        passes++;
        last_eb = eb;
        *bpv = last_bpv = 8.0 * (double)r[1] / (double)n;
        note(eb, *bpv);
        return true;
    }

    // Exact stream sizes for the OCT_BINS bounds of octave e (nocts = 2:
    // also those of octave e - 1, in the same pass; bpv and ebs then run
    // from octave e - 1's lowest bound up), and for the bounds just outside
    // the span: edge_up at 2^(e+1), edge_dn at the float below the span's
    // lowest bound (eb_dn).
    double edge_up = 0.0, edge_dn = 0.0;
    float eb_up = 0.0f, eb_dn = 0.0f;
    bool octave(int e, int nocts, std::vector<double> &bpv, std::vector<float> &ebs)
    {
        // float bounds, both edges inside
        const int e_min = -149 + nocts;
        e = e < e_min ? e_min : (e > 126 ? 126 : e);
        rate_octave_launch(x, n, w, bs, e, ws + L.oct, (unsigned long long *)(ws + L.acc), st, true, nocts);
        std::vector<unsigned long long> acc((size_t)nocts * OCT_ACC + 4);
        if (!ok(hipMemcpyAsync(acc.data(), ws + L.acc, acc.size() * 8, hipMemcpyDeviceToHost, st),
                "D2H octave") ||
            !ok(hipStreamSynchronize(st), "octave pass"))
            return false;
        passes++;
        const uint64_t nb = (n + bs - 1) / bs, nbit = (nb + 7) / 8;
        bpv.assign((size_t)nocts * OCT_BINS, 0.0);
        ebs.assign((size_t)nocts * OCT_BINS, 0.0f);
        for (int j = 0; j < nocts; j++) {
            // octave e - j fills entries [(nocts - 1 - j) * OCT_BINS, ...)
            const unsigned long long *a = acc.data() + (size_t)j * OCT_ACC;
            const size_t at = (size_t)(nocts - 1 - j) * OCT_BINS;
            // size(eb_k) = header + bitmap + fixed + 4 * (count of bins < k)
            //              + records of bins >= k
            uint64_t rec_all = 0;
            for (int k = 0; k < OCT_BINS; k++) rec_all += a[2 + 2 * k];
            uint64_t const_part = 0, rec_part = rec_all;
            for (int k = 0; k < OCT_BINS; k++) {
                const uint64_t size = STREAM_HEADER + nbit + a[0] + const_part + rec_part;
                bpv[at + k] = 8.0 * (double)size / (double)n;
                ebs[at + k] = ldexpf(1.0f + (float)k / OCT_BINS, e - j);
                const_part += 4 * a[1 + 2 * k];
                rec_part -= a[2 + 2 * k];
            }
        }
        const unsigned long long *g = acc.data() + (size_t)nocts * OCT_ACC;
        edge_up = 8.0 * (double)(STREAM_HEADER + nbit + 4 * g[0] + g[1]) / (double)n;
        edge_dn = 8.0 * (double)(STREAM_HEADER + nbit + 4 * g[2] + g[3]) / (double)n;
        eb_up = ldexpf(1.0f, e + 1);
        eb_dn = nextafterf(ldexpf(1.0f, e - nocts + 1), 0.0f);
        return true;
    }

    // Octave search from the octave(s) e (nocts = 2: e and e - 1) on; on
    // success *eb is the best bound of the span that holds the target (or
    // the best seen when none does).  Later tries size one octave each, next
    // to the span tried before.  Returns false only on a device error;
    // *found says whether a span bracketed the target.
    bool octave_search(int e, int nocts, float *eb, bool *found, int *lo_e, int *hi_e)
    {
        *found = false;
        *lo_e = -100000;
        *hi_e = 100000;
        std::vector<double> bpv;
        std::vector<float> ebs;
        int last_dir = 0;           // +1: moved to looser bounds, -1: tighter
        float edge_eb = 0.0f;       // nearest bound of the previous span
        double edge_bpv = -1.0;
        for (int t = 0; t < OCT_TRIES; t++) {
            if (!octave(e, nocts, bpv, ebs)) return false;
            const int nk = (int)bpv.size();
            // bpv is non-increasing in k
            int bk = 0;
            for (int k = 1; k < nk; k++)
                if (fabs(bpv[k] - target) < fabs(bpv[bk] - target)) bk = k;
            int dir = 0;
            if (bpv[nk - 1] > target * (1.0 + tol)) dir = +1;   // bounds too tight
            else if (bpv[0] < target * (1.0 - tol)) dir = -1;   // bounds too loose
            if (dir == 0) {
                *eb = ebs[bk];
                *found = true;
                return true;
            }
            // the bound just past the span's edge in that direction: the
            // answer itself, or the far side of a jump that holds the target
            // (no bound reaches it: take the nearer side), or else the next
            // octave is needed
            const double edge = dir > 0 ? edge_up : edge_dn;
            const float eb_edge = dir > 0 ? eb_up : eb_dn;
            const double near = dir > 0 ? bpv[nk - 1] : bpv[0];
            const float eb_near = dir > 0 ? ebs[nk - 1] : ebs[0];
            if (within(edge)) {
                *eb = eb_edge;
                *found = true;
                return true;
            }
            if (dir > 0 ? edge < target * (1.0 - tol) : edge > target * (1.0 + tol)) {
                *eb = fabs(edge - target) < fabs(near - target) ? eb_edge : eb_near;
                *found = true;
                return true;
            }
            if (last_dir == -dir) {
                // the target falls in the jump between two neighbouring
                // octaves (every block's reqLength moves by one bit at the
                // octave edge): no bound reaches it; take the nearer edge
                *eb = fabs(bpv[bk] - target) <= fabs(edge_bpv - target) ? ebs[bk] : edge_eb;
                *found = true;
                return true;
            }
            edge_eb = ebs[bk];
            edge_bpv = bpv[bk];
            last_dir = dir;
            if (dir > 0) {
                *lo_e = e + 1 > *lo_e ? e + 1 : *lo_e;
                e++;
            } else {
                const int bottom = e - nocts + 1;
                *hi_e = bottom < *hi_e ? bottom : *hi_e;
                e = bottom - 1;
            }
            nocts = 1;
        }
        return true;
    }

    bool bisect(double lg_lo, double lg_hi, float *eb_out)
    {
        for (int i = 0; i < BISECT_MAX; i++) {
            const double mid = 0.5 * (lg_lo + lg_hi);
            const float eb = (float)exp2(mid);
            double bpv;
            if (!compress_at(eb, &bpv)) return false;
            if (within(bpv)) {
                *eb_out = eb;
                return true;
            }
            if (bpv > target) lg_lo = mid;   // too many bits: loosen the bound
            else lg_hi = mid;
            if (lg_hi - lg_lo < 1e-7) break;
        }
        *eb_out = best_eb;
        return true;
    }

    bool value_range(double *range)
    {
        range_launch(x, n, w, ws + L.range, ws + L.minmax, st);
        double mm[2] = {0, 0};
        if (!ok(hipMemcpyAsync(mm, ws + L.minmax, 2 * (size_t)w, hipMemcpyDeviceToHost, st), "D2H") ||
            !ok(hipStreamSynchronize(st), "range"))
            return false;
        if (w == 4) {
            float f[2];
            memcpy(f, mm, 8);
            *range = (double)f[1] - (double)f[0];
        } else {
            *range = mm[1] - mm[0];
        }
        return true;
    }
};

}  // namespace

uint64_t fixed_rate_workspace(uint64_t n, int w, int bs) { return rate_layout(n, w, bs).bytes; }

int fixed_rate_solve(const void *d_in, uint64_t n, int w, int bs, float target, float tol, int method,
                     const float *d_params, uint8_t *d_out, uint8_t *d_ws, uint64_t *d_result,
                     uint64_t *d_index, szx_amd_rate_info *info, hipStream_t st, const char **err)
{
    Solver S;
    S.x = d_in;
    S.n = n;
    S.w = w;
    S.bs = bs;
    S.target = target;
    S.tol = tol;
    S.out = d_out;
    S.ws = d_ws;
    S.result = d_result;
    S.index = d_index;
    S.st = st;
    S.L = rate_layout(n, w, bs);
    const double depth = w == 4 ? 40.0 : 64.0;   // octaves below the range searched
    float eb = 0.0f;
    double bpv = 0.0;
    int feature_passes = 0;
    bool done = false;

    if (method == SZX_AMD_RATE_LEARNED) {
        // features and the model's ladder of predictions, one synchronisation
        const int nch = (int)S.L.nchunks;
        const int K = (int)(4 * depth) < K_MAX ? (int)(4 * depth) : K_MAX;
        const RateStats *d_stats = (const RateStats *)(d_ws + S.L.stats);
        float *d_ebs = (float *)(d_ws + S.L.ebs);
        rate_stats_launch(d_in, n, w, bs, d_ws + S.L.stats_ws, (RateStats *)(d_ws + S.L.stats), st);
        rate_ladder_launch(d_stats, nch, K, d_ebs, st);
        rate_inputs_launch(d_stats, nch, d_ebs, K, (float *)(d_ws + S.L.X), st);
        mlp_forward_launch(d_params, (const float *)(d_ws + S.L.X), (uint64_t)K * nch,
                           (float *)(d_ws + S.L.y), st);
        std::vector<float> ladder(K), y((size_t)K * nch);
        if (!S.ok(hipMemcpyAsync(ladder.data(), d_ebs, 4 * (size_t)K, hipMemcpyDeviceToHost, st),
                  "D2H ladder") ||
            !S.ok(hipMemcpyAsync(y.data(), d_ws + S.L.y, y.size() * 4, hipMemcpyDeviceToHost, st),
                  "D2H predictions") ||
            !S.ok(hipStreamSynchronize(st), "rate model")) {
            *err = S.err;
            return -1;
        }
        feature_passes = 1;
        const double range = (double)ladder[0];
        // predicted field rate along the ladder (bits of the records,
        // medians and bitmap per chunk, plus the 13-byte header)
        std::vector<double> pred(K);
        for (int k = 0; k < K; k++) {
            double bits = 8.0 * STREAM_HEADER;
            for (int c = 0; c < nch; c++) {
                const uint64_t vals = n - (uint64_t)c * RATE_CHUNK < RATE_CHUNK ? n - (uint64_t)c * RATE_CHUNK
                                                                                : RATE_CHUNK;
                bits += exp2((double)y[(size_t)k * nch + c]) * (double)vals;
            }
            pred[k] = bits / (double)n;
        }
        // first rung whose predicted rate reaches the target; interpolate in
        // log2(rate) between it and the rung before
        double lg_eb = log2((double)ladder[K - 1]);
        for (int k = 0; k < K; k++) {
            if (pred[k] >= target) {
                if (k == 0) {
                    lg_eb = log2((double)ladder[0]);
                } else {
                    const double a = log2(pred[k - 1]), b = log2(pred[k]), t = log2((double)target);
                    const double f = b > a ? (t - a) / (b - a) : 1.0;
                    lg_eb = log2((double)ladder[k - 1]) - 0.25 * f;
                }
                break;
            }
        }
        // a prediction within OCT_PAIR of an octave edge sizes the octaves
        // on both sides of that edge in its first pass
        const int e0 = (int)floor(lg_eb);
        const double frac = lg_eb - e0;
        const int e_top = frac < OCT_PAIR ? e0 : (frac > 1.0 - OCT_PAIR ? e0 + 1 : e0);
        const int nocts = frac < OCT_PAIR || frac > 1.0 - OCT_PAIR ? 2 : 1;
        bool found = false;
        int lo_e, hi_e;
        if (!S.octave_search(e_top, nocts, &eb, &found, &lo_e, &hi_e)) {
            *err = S.err;
            return -1;
        }
        if (found) {
            if (!S.compress_at(eb, &bpv)) {
                *err = S.err;
                return -1;
            }
            done = true;
        } else {
            // bisection inside what the octave passes ruled out
            double glo = log2(range) - depth, ghi = log2(range) + 1;
            if (lo_e > -100000) glo = glo > lo_e ? glo : (double)lo_e;
            if (hi_e < 100000) ghi = ghi < hi_e ? ghi : (double)hi_e;
            if (!(glo < ghi)) {
                glo = log2(range) - depth;
                ghi = log2(range) + 1;
            }
            if (!S.bisect(glo, ghi, &eb)) {
                *err = S.err;
                return -1;
            }
        }
    } else {
        double range;
        if (!S.value_range(&range)) {
            *err = S.err;
            return -1;
        }
        if (!(range > 0)) range = 1.0;
        if (!S.bisect(log2(range) - depth, log2(range) + 1, &eb)) {
            *err = S.err;
            return -1;
        }
    }
    if (!done) {
        // leave the stream of the chosen bound (the best seen when no bound
        // was within tolerance) in the output buffer
        if (!S.within(S.best_bpv)) eb = S.best_eb;
        if (S.last_eb == eb) {
            bpv = S.last_bpv;
        } else if (!S.compress_at(eb, &bpv)) {
            *err = S.err;
            return -1;
        }
    }
    if (info) {
        info->errorBound = eb;
        info->bitsPerValue = bpv;
        info->codecPasses = S.passes;
        info->featurePasses = feature_passes;
        info->hit = S.within(bpv) ? 1 : 0;
    }
    return 0;
}

}  // namespace szxamd
