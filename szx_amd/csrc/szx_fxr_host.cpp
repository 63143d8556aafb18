// szx_fxr_host.cpp -- host half of the reference's fixed-rate bound search.
//
// The search (szx_float.cc:684-832, szx_double.cc:682-829) is a short,
// data-dependent control loop over model evaluations; each evaluation needs
// three integer sums over the sampled blocks, which the GPU computes
// (k_fxr_estimate).  The model formula and the loop are evaluated here with
// the reference's own types (float / double promotions), so the chosen bound
// is bit-identical to the reference's.
//
// The bracket phase's candidates are fixed by the start bound (init / 8^i or
// init * 8^i, i = 1..10), so all 21 are evaluated in one pass over the sample
// and the loop below replays the reference's order over the results; only the
// bisection steps need one pass each.
#include "szx_fxr_host.hpp"

#include <math.h>

namespace szxamd {

// szx_float.cc:576-582 / szx_double.cc:574-580.
float fxr_model_cr(int w, int bs, size_t n, size_t ns, const FxrSums &s)
{
    float metadata = 13.0 * bs / n;
    float block_cost = w == 4 ? 33.0 / 8 : 65.0 / 8;
    float avg_req = 1.0 * s.sum_req / (ns - s.nconst);
    float avg_lead = 1.0 * s.sum_lead / (ns - s.nconst);
    float p_const = 1.0 * s.nconst / ns;
    float cr = (w * bs) / (metadata + block_cost + (1 + (0.25 + avg_req) * bs - avg_lead) * (1 - p_const));
    return cr;
}

bool fxr_search(float target, float tol, float init_eb, int w, int bs, size_t n, size_t ns,
                const FxrEvaluator &eval, float *eb_out, int *n_est)
{
    const int stride = 8, iters = 10;
    // candidates of the bracket phase, in the reference's float arithmetic
    float cand[1 + 2 * 10];
    cand[0] = init_eb;
    float d = init_eb, u = init_eb;
    for (int i = 0; i < iters; i++) {
        d = d / stride;
        u = u * stride;
        cand[1 + i] = d;
        cand[1 + iters + i] = u;
    }
    FxrSums pre[1 + 2 * 10];
    if (!eval(cand, 1 + 2 * iters, pre)) return false;

    // history of evaluated bounds (the seen-guard and best-so-far use it)
    size_t h_req[32], h_lead[32];
    float h_eb[32], h_cr[32];
    int k = 0;
    auto record = [&](float e, const FxrSums &s, float cr) {
        h_req[k] = s.sum_req;
        h_lead[k] = s.sum_lead;
        h_eb[k] = e;
        h_cr[k] = cr;
        k++;
    };

    float eb = init_eb, result = 0, lo = 0, hi = 0;
    int found = 0;
    float cr = fxr_model_cr(w, bs, n, ns, pre[0]);
    record(eb, pre[0], cr);
    *n_est = 1;
    if (fabs(target - cr) < tol) {   // the first test is absolute (szx_float.cc:707)
        *eb_out = eb;
        return true;
    }
    float ttol = target * tol;
    if (target < cr) {   // bound too loose: shrink it
        lo = eb / stride;
        hi = eb;
        for (int i = 0; i < iters; i++) {
            cr = fxr_model_cr(w, bs, n, ns, pre[1 + i]);
            (*n_est)++;
            record(lo, pre[1 + i], cr);
            if (fabs(cr - target) < ttol) {
                found = 1;
                result = lo;
                break;
            }
            if (target > cr) break;
            hi = lo;
            lo = lo / stride;
        }
    } else if (target > cr) {   // bound too tight: grow it
        lo = eb;
        hi = eb * stride;
        for (int i = 0; i < iters; i++) {
            cr = fxr_model_cr(w, bs, n, ns, pre[1 + iters + i]);
            (*n_est)++;
            record(hi, pre[1 + iters + i], cr);
            if (fabs(target - cr) < ttol) {
                found = 1;
                result = hi;
                break;
            }
            if (target < cr) break;
            lo = hi;
            hi = hi * stride;
        }
    }
    if (!found) {
        for (int i = 0; i < iters; i++) {
            eb = (lo + hi) / 2;
            FxrSums s;
            if (!eval(&eb, 1, &s)) return false;
            cr = fxr_model_cr(w, bs, n, ns, s);
            (*n_est)++;
            int seen = 0;
            for (int j = 0; j < k; j++)
                if (h_req[j] == s.sum_req && h_lead[j] == s.sum_lead) {
                    seen = 1;
                    break;
                }
            float best = 10000000;
            int bi = 0;
            for (int j = 0; j < k; j++) {
                float e = fabs(h_cr[j] - target);
                if (e < best) {
                    best = e;
                    bi = j;
                }
            }
            if (seen) {
                result = h_eb[bi];
                break;
            }
            record(eb, s, cr);
            if (fabs(cr - target) < ttol) {
                result = eb;
                break;
            }
            if (target > cr) lo = eb;
            else hi = eb;
        }
    }
    *eb_out = result;
    return true;
}

}  // namespace szxamd
