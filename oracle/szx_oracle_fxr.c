/*
 * szx_oracle_fxr.c -- CPU restatement of the reference's fixed-rate (FXR)
 * error-bound search.  TEST INFRASTRUCTURE ONLY (the checker / CPU baseline);
 * never part of libszx_amd.so.
 *
 * The reference picks the error bound for a target compression ratio from a
 * 1-in-10 block sample and an analytic size model, never running the codec
 * until the final compress:
 *   sampling      szx_float.cc:254-312   (fp64: szx_double.cc:437-495)
 *   CR model      szx_float.cc:501-584   (fp64: szx_double.cc:497-580)
 *   search        szx_float.cc:684-832   (fp64: szx_double.cc:682-829)
 *   driver        szx_float.cc:921-949   (fp64: szx_double.cc:917-943)
 * Arithmetic is restated with the reference's own types (float vs double
 * promotions), since the chosen bound is order- and rounding-sensitive.
 */
#include "szx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint32_t f32_bits(float v) { uint32_t u; memcpy(&u, &v, 4); return u; }
static uint64_t f64_bits(double v) { uint64_t u; memcpy(&u, &v, 8); return u; }

/* Sample the first block of every group of `rate` blocks.  Returns the value
 * range over the sampled blocks (float).  buf receives the sampled values,
 * med/rad one float each per sampled block. */
float szxo_fxr_sample(const void *x, int w, size_t n, int rate, int bs,
                      void *buf, float *med, float *rad)
{
    size_t stride = (size_t)rate * (size_t)bs, ns = n / stride;
    if (w == 4) {
        const float *v = (const float *)x;
        float g_lo = v[0], g_hi = v[0];
        for (size_t i = 0; i < ns; i++) {
            const float *b = v + i * stride;
            memcpy((float *)buf + i * (size_t)bs, b, (size_t)bs * 4);
            float lo = b[0], hi = b[0];
            for (int j = 1; j < bs; j++) {
                if (lo > b[j]) lo = b[j];
                else if (hi < b[j]) hi = b[j];
            }
            if (g_hi < hi) g_hi = hi;
            if (g_lo > lo) g_lo = lo;
            float r = (hi - lo) / 2;
            rad[i] = r;
            med[i] = lo + r;
        }
        return g_hi - g_lo;
    } else {
        const double *v = (const double *)x;
        double g_lo = v[0], g_hi = v[0];
        for (size_t i = 0; i < ns; i++) {
            const double *b = v + i * stride;
            memcpy((double *)buf + i * (size_t)bs, b, (size_t)bs * 8);
            double lo = b[0], hi = b[0];
            for (int j = 1; j < bs; j++) {
                if (lo > b[j]) lo = b[j];
                else if (hi < b[j]) hi = b[j];
            }
            if (g_hi < hi) g_hi = hi;
            if (g_lo > lo) g_lo = lo;
            double r = (hi - lo) / 2;
            rad[i] = (float)r;
            med[i] = (float)(lo + r);
        }
        return (float)(g_hi - g_lo);
    }
}

/* Analytic compression-ratio estimate at error bound eb over the sample.
 * sum_req / sum_lead return the two integer sums the search uses to detect
 * a repeated estimate. */
float szxo_fxr_estimate_cr(float eb, const void *buf, int w, const float *med,
                           const float *rad, int rate, int bs, size_t n,
                           size_t *sum_req, size_t *sum_lead)
{
    size_t ns = n / ((size_t)bs * (size_t)rate), nconst = 0;
    float metadata = 13.0 * bs / n;
    float block_cost = w == 4 ? 33.0 / 8 : 65.0 / 8;
    *sum_req = 0;
    *sum_lead = 0;
    for (size_t i = 0; i < ns; i++) {
        float m = med[i], r = rad[i];
        if (r <= eb) { nconst++; continue; }
        int clamped;
        int len = szxo_req_length(w, eb, r, &clamped);
        if (clamped) m = 0;   /* both widths: the model uses the zeroed median */
        int rb = (len + 7) / 8, shift = (8 - len % 8) % 8;
        *sum_req += (size_t)rb;
        uint64_t prev = 0;
        size_t lead_sum = 0;
        for (int j = 0; j < bs; j++) {
            uint64_t s;
            if (w == 4) {
                float d = ((const float *)buf)[i * (size_t)bs + j] - m;
                s = f32_bits(d) >> shift;
            } else {
                double d = ((const double *)buf)[i * (size_t)bs + j] - (double)m;
                s = f64_bits(d) >> shift;
            }
            uint64_t xr = s ^ prev;
            int ld = 0;
            if ((xr >> (8 * (w - 3))) == 0) ld = 3;
            else if ((xr >> (8 * (w - 2))) == 0) ld = 2;
            else if ((xr >> (8 * (w - 1))) == 0) ld = 1;
            lead_sum += (size_t)(ld >= rb ? rb : ld);
            prev = s;
        }
        *sum_lead += lead_sum;
    }
    float avg_req = 1.0 * (*sum_req) / (ns - nconst);
    float avg_lead = 1.0 * (*sum_lead) / (ns - nconst);
    float p_const = 1.0 * nconst / ns;
    float cr = w * bs / (metadata + block_cost +
                         (1 + (0.25 + avg_req) * bs - avg_lead) * (1 - p_const));
    return cr;
}

/* Bracket-then-bisect search for the bound whose estimated ratio is within
 * tol of target.  *n_est returns the number of model evaluations. */
float szxo_fxr_search(float target, float tol, int rate, float init_eb, int bs,
                      size_t n, const void *buf, int w, const float *med,
                      const float *rad, int *n_est)
{
    const int stride = 8, iters = 10;
    size_t hist_req[24], hist_lead[24];
    float hist_eb[24], hist_cr[24];
    int k = 0, found = 0;
    size_t sr, sl;
    float eb = init_eb, result = 0, lo = 0, hi = 0;
    float cr = szxo_fxr_estimate_cr(eb, buf, w, med, rad, rate, bs, n, &sr, &sl);
    hist_req[k] = sr; hist_lead[k] = sl; hist_eb[k] = eb; hist_cr[k] = cr; k++;
    *n_est = 1;
    if (fabs(target - cr) < tol) return eb;       /* first test is absolute */
    float ttol = target * tol;
    if (target < cr) {                             /* bound too loose: shrink */
        lo = eb / stride; hi = eb;
        for (int i = 0; i < iters; i++) {
            cr = szxo_fxr_estimate_cr(lo, buf, w, med, rad, rate, bs, n, &sr, &sl);
            (*n_est)++;
            hist_req[k] = sr; hist_lead[k] = sl; hist_eb[k] = lo; hist_cr[k] = cr; k++;
            if (fabs(cr - target) < ttol) { found = 1; result = lo; break; }
            if (target > cr) break;
            hi = lo;
            lo = lo / stride;
        }
    } else if (target > cr) {                      /* bound too tight: grow */
        lo = eb; hi = eb * stride;
        for (int i = 0; i < iters; i++) {
            cr = szxo_fxr_estimate_cr(hi, buf, w, med, rad, rate, bs, n, &sr, &sl);
            (*n_est)++;
            hist_req[k] = sr; hist_lead[k] = sl; hist_eb[k] = hi; hist_cr[k] = cr; k++;
            if (fabs(target - cr) < ttol) { found = 1; result = hi; break; }
            if (target < cr) break;
            lo = hi;
            hi = hi * stride;
        }
    }
    if (!found) {
        for (int i = 0; i < iters; i++) {
            eb = (lo + hi) / 2;
            cr = szxo_fxr_estimate_cr(eb, buf, w, med, rad, rate, bs, n, &sr, &sl);
            (*n_est)++;
            int seen = 0;
            for (int j = 0; j < k; j++)
                if (hist_req[j] == sr && hist_lead[j] == sl) { seen = 1; break; }
            float best = 10000000;
            int bi = 0;
            for (int j = 0; j < k; j++) {
                float e = fabs(hist_cr[j] - target);
                if (e < best) { best = e; bi = j; }
            }
            if (seen) { result = hist_eb[bi]; break; }
            hist_req[k] = sr; hist_lead[k] = sl; hist_eb[k] = eb; hist_cr[k] = cr; k++;
            if (fabs(cr - target) < ttol) { result = eb; break; }
            if (target > cr) lo = eb;
            else hi = eb;
        }
    }
    return result;
}

/* Whole FXR bound selection for a field (szx_float.cc:921-949 without the
 * final compress): sample 1 block in 10, start at 1% of the sampled range. */
float szxo_fxr_error_bound(const void *x, int w, size_t n, float target_cr,
                           float tol, int bs, int *n_est)
{
    const int rate = 10;
    size_t ns = n / ((size_t)rate * (size_t)bs);
    void *buf = malloc((ns + 1) * (size_t)bs * (size_t)w);
    float *med = (float *)malloc((ns + 1) * sizeof(float));
    float *rad = (float *)malloc((ns + 1) * sizeof(float));
    float range = szxo_fxr_sample(x, w, n, rate, bs, buf, med, rad);
    float init = range * 1E-2;
    float eb = szxo_fxr_search(target_cr, tol, rate, init, bs, n, buf, w, med, rad, n_est);
    free(buf); free(med); free(rad);
    return eb;
}
