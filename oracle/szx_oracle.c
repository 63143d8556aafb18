/*
 * szx_oracle.c -- CPU restatement of SZx's blockwise fast codec (fastMode 2).
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never part of libszx_amd.so.
 *
 * Stream layout restated here (szx_float.cc:1091-1131):
 *   [0..4]   1, 1, 1, 0, (u8)blockSize
 *   [5..12]  big-endian u64 number of constant blocks (nc)
 *   [13..]   state bitmap, ceil(nb/8) bytes, MSB-first, 1 = non-constant
 *   then     nc big-endian float medians of the constant blocks, block order
 *   then     one record per non-constant block, block order:
 *              u8 reqLength | BE float median | ceil(n/4) lead bytes | mid-bytes
 *
 * Build: gcc -O2 -std=c99 -ffp-contract=off (no fast-math: every float
 * operation must be the IEEE operation the reference performs).
 */
#include "szx_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ---- bit views ---------------------------------------------------------- */

static uint32_t f32_bits(float v) { uint32_t u; memcpy(&u, &v, 4); return u; }
static float bits_f32(uint32_t u) { float v; memcpy(&v, &u, 4); return v; }
static uint64_t f64_bits(double v) { uint64_t u; memcpy(&u, &v, 8); return u; }
static double bits_f64(uint64_t u) { double v; memcpy(&v, &u, 8); return v; }

/* Big-endian float (floatToBytes / bytesToFloat, szx_BytesToolkit.cc:460-476). */
static void put_be_f32(uint8_t *p, float v)
{
    uint32_t u = f32_bits(v);
    p[0] = (uint8_t)(u >> 24); p[1] = (uint8_t)(u >> 16);
    p[2] = (uint8_t)(u >> 8);  p[3] = (uint8_t)u;
}
static float get_be_f32(const uint8_t *p)
{
    return bits_f32(((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) |
                    ((uint32_t)p[2] << 8) | (uint32_t)p[3]);
}
/* Big-endian u64 (longToBytes_bigEndian / bytesToLong_bigEndian,
 * szx_BytesToolkit.cc:266-317). */
static void put_be_u64(uint8_t *p, uint64_t v)
{
    for (int k = 0; k < 8; k++) p[k] = (uint8_t)(v >> (56 - 8 * k));
}
static uint64_t get_be_u64(const uint8_t *p)
{
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v = (v << 8) | p[k];
    return v;
}

static size_t ceil_div(size_t a, size_t b) { return (a + b - 1) / b; }

size_t szxo_compress_bound(size_t n, int bs, int w)
{
    size_t nb = ceil_div(n, (size_t)bs);
    /* header + bitmap + per block max(4-byte median, full record) */
    return 13 + ceil_div(nb, 8) + nb * (5 + ceil_div((size_t)bs, 4) + (size_t)bs * (size_t)w);
}

/* ---- block statistics (A3) ---------------------------------------------- */

/* Sequential min/max with the reference's compare order
 * (szx_float.cc:327-333): the running minimum is replaced only by a strictly
 * smaller value, the maximum only by a strictly larger one, and a value that
 * lowered the minimum is not tested against the maximum. */
size_t szxo_block_stats_f32(const float *x, size_t n, float eb, int bs,
                            uint8_t *state, float *median, float *radius)
{
    size_t nb = ceil_div(n, (size_t)bs), nconst = 0;
    for (size_t b = 0; b < nb; b++) {
        const float *v = x + b * (size_t)bs;
        size_t m = (b + 1) * (size_t)bs <= n ? (size_t)bs : n - b * (size_t)bs;
        float lo = v[0], hi = v[0];
        for (size_t j = 1; j < m; j++) {
            if (lo > v[j]) lo = v[j];
            else if (hi < v[j]) hi = v[j];
        }
        float rad = (hi - lo) / 2;
        float med = lo + rad;
        state[b] = rad <= eb ? 0 : 1;
        nconst += state[b] == 0;
        median[b] = med;
        radius[b] = rad;
    }
    return nconst;
}

size_t szxo_block_stats_f64(const double *x, size_t n, float eb, int bs,
                            uint8_t *state, float *median, float *radius)
{
    size_t nb = ceil_div(n, (size_t)bs), nconst = 0;
    for (size_t b = 0; b < nb; b++) {
        const double *v = x + b * (size_t)bs;
        size_t m = (b + 1) * (size_t)bs <= n ? (size_t)bs : n - b * (size_t)bs;
        double lo = v[0], hi = v[0];
        for (size_t j = 1; j < m; j++) {
            if (lo > v[j]) lo = v[j];
            else if (hi < v[j]) hi = v[j];
        }
        double rad = (hi - lo) / 2;
        double med = lo + rad;
        state[b] = rad <= (double)eb ? 0 : 1;   /* szx_double.cc:994 */
        nconst += state[b] == 0;
        median[b] = (float)med;                 /* szx_double.cc:1002-1003 */
        radius[b] = (float)rad;
    }
    return nconst;
}

/* ---- reqLength (A4) ----------------------------------------------------- */

/* radExpo: unbiased exponent field of the float radius (getExponent_float,
 * szx_BytesToolkit.cc:335-346).  reqExpo: unbiased exponent field of the
 * error bound widened to double (getPrecisionReqLength_double, :375-387).
 * fp32 length = 10 + radExpo - reqExpo in [9, 32] (szx_float.cc:1708-1717);
 * fp64 length = 12 + radExpo - reqExpo in [12, 64] (szx_double.cc:2020-2031). */
int szxo_req_length(int w, float eb, float radius, int *clamped)
{
    int rad_expo = (int)((f32_bits(radius) >> 23) & 0xFF) - 127;
    int req_expo = (int)((f64_bits((double)eb) >> 52) & 0x7FF) - 1023;
    int lo = w == 4 ? 9 : 12, hi = 8 * w;
    int len = (w == 4 ? 10 : 12) + rad_expo - req_expo;
    *clamped = 0;
    if (len < lo) len = lo;
    if (len > hi) { len = hi; *clamped = 1; }
    return len;
}

/* ---- one-block encoder (A5, A6) ----------------------------------------- */

/* Generic over the word width w: s_i = bits(x_i - median) >> shift (logical),
 * lead_i = number of leading bytes (capped at 3) that s_i shares with s_{i-1}
 * (s_{-1} = 0), 2-bit lead codes packed MSB-first four to a byte
 * (convertIntArray2ByteArray_fast_2b_args, szx_TypeManager.cc:198-273), and
 * the bytes [w-rb, w-lead_i) of s_i appended in ascending byte order when
 * lead_i < rb. */
static size_t encode_words(const uint64_t *s, size_t n, int w, int len,
                           float hdr_median, uint8_t *out)
{
    int rb = (len + 7) / 8;
    size_t nlead = ceil_div(n, 4);
    uint8_t *lead = out + 5;
    size_t pos = 5 + nlead;
    uint64_t prev = 0;
    out[0] = (uint8_t)len;
    put_be_f32(out + 1, hdr_median);
    memset(lead, 0, nlead);
    for (size_t i = 0; i < n; i++) {
        uint64_t xr = s[i] ^ prev;
        int ld = 0;
        if ((xr >> (8 * (w - 3))) == 0) ld = 3;
        else if ((xr >> (8 * (w - 2))) == 0) ld = 2;
        else if ((xr >> (8 * (w - 1))) == 0) ld = 1;
        lead[i >> 2] |= (uint8_t)(ld << (6 - 2 * (int)(i & 3)));
        for (int b = w - rb; b < w - ld; b++) out[pos++] = (uint8_t)(s[i] >> (8 * b));
        prev = s[i];
    }
    return pos;
}

size_t szxo_encode_block_f32(const float *x, size_t n, float eb, float median,
                             float radius, uint8_t *out)
{
    int clamped;
    int len = szxo_req_length(4, eb, radius, &clamped);
    if (clamped) median = 0;   /* fp32 encodes against the zeroed median */
    int shift = (8 - len % 8) % 8;
    uint64_t s[256];
    if (n > 256) return 0;   /* blockSize is stored in one byte */
    for (size_t i = 0; i < n; i++) {
        float d = x[i] - median;
        s[i] = (uint64_t)(f32_bits(d) >> shift);
    }
    return encode_words(s, n, 4, len, median, out);
}

size_t szxo_encode_block_f64(const double *x, size_t n, float eb, float median,
                             float radius, uint8_t *out)
{
    int clamped;
    int len = szxo_req_length(8, eb, radius, &clamped);
    /* szx_double.cc:38: the payload uses the median captured before the
     * clamp; only the stored header median becomes 0. */
    double med = (double)median;
    int shift = (8 - len % 8) % 8;
    uint64_t s[256];
    if (n > 256) return 0;
    for (size_t i = 0; i < n; i++) {
        double d = x[i] - med;
        s[i] = f64_bits(d) >> shift;
    }
    return encode_words(s, n, 8, len, clamped ? 0.0f : median, out);
}

/* ---- stream assembly (A7, A8) ------------------------------------------- */

static size_t compress_generic(const void *x, size_t n, float eb, int bs, int w, uint8_t *out)
{
    size_t nb = ceil_div(n, (size_t)bs);
    size_t nbit = ceil_div(nb, 8);
    uint8_t *state = (uint8_t *)malloc(nb ? nb : 1);
    float *med = (float *)malloc((nb ? nb : 1) * sizeof(float));
    float *rad = (float *)malloc((nb ? nb : 1) * sizeof(float));
    size_t nconst = w == 4
        ? szxo_block_stats_f32((const float *)x, n, eb, bs, state, med, rad)
        : szxo_block_stats_f64((const double *)x, n, eb, bs, state, med, rad);
    out[0] = 1; out[1] = 1; out[2] = 1; out[3] = 0;
    out[4] = (uint8_t)bs;                 /* (u8)blockSize: 256 -> 0 */
    put_be_u64(out + 5, nconst);
    uint8_t *bitmap = out + 13;
    memset(bitmap, 0, nbit);
    uint8_t *cmed = bitmap + nbit;
    uint8_t *rec = cmed + 4 * nconst;
    size_t k = 0;
    for (size_t b = 0; b < nb; b++) {
        size_t m = (b + 1) * (size_t)bs <= n ? (size_t)bs : n - b * (size_t)bs;
        if (state[b]) {
            bitmap[b >> 3] |= (uint8_t)(0x80 >> (b & 7));
            rec += w == 4
                ? szxo_encode_block_f32((const float *)x + b * (size_t)bs, m, eb, med[b], rad[b], rec)
                : szxo_encode_block_f64((const double *)x + b * (size_t)bs, m, eb, med[b], rad[b], rec);
        } else {
            put_be_f32(cmed + 4 * k++, med[b]);
        }
    }
    free(state); free(med); free(rad);
    return (size_t)(rec - out);
}

size_t szxo_compress_f32(const float *x, size_t n, float eb, int bs, uint8_t *out)
{
    return compress_generic(x, n, eb, bs, 4, out);
}

size_t szxo_compress_f64(const double *x, size_t n, float eb, int bs, uint8_t *out)
{
    return compress_generic(x, n, eb, bs, 8, out);
}

/* ---- decoder (A10, A11) ------------------------------------------------- */

/* Lead code i of a record (szxd_float.cc:88-90). */
static int lead_code(const uint8_t *lead, size_t i)
{
    return (lead[i >> 2] >> (6 - 2 * (int)(i & 3))) & 3;
}

size_t szxo_record_size(const uint8_t *rec, size_t n, int w)
{
    int rb = (rec[0] + 7) / 8;
    size_t nlead = ceil_div(n, 4), sz = 5 + nlead;
    (void)w;
    for (size_t i = 0; i < n; i++) {
        int ld = lead_code(rec + 5, i);
        if (ld < rb) sz += (size_t)(rb - ld);
    }
    return sz;
}

/* Reconstruct the shifted words t_i of one record.  t_i takes its bytes
 * [w-l, w) from t_{i-1} (l = min(lead_i, rb)), bytes [w-rb, w-l) from the
 * stream in ascending order, and zero below; after each element the word is
 * masked to its low 8w-shift bits, as the reference's shift-left/shift-right
 * round trip does (szxd_float.cc:197-199).  quirk5 reproduces the fp64
 * reqBytes=5, lead=0 case that fills byte 7 from the fourth stream byte
 * (szxd_double.cc:240-245) while still consuming five. */
static size_t decode_words(const uint8_t *rec, size_t n, int w, uint64_t *t,
                           int quirk5, size_t avail)
{
    int len = rec[0];
    int rb = (len + 7) / 8;
    int shift = (8 - len % 8) % 8;
    size_t nlead = ceil_div(n, 4);
    const uint8_t *q = rec + 5 + nlead;
    const uint8_t *end = rec + avail;
    uint64_t mask = shift ? (~(uint64_t)0 >> (64 - 8 * w)) >> shift
                          : (~(uint64_t)0 >> (64 - 8 * w));
    uint64_t prev = 0;
    for (size_t i = 0; i < n; i++) {
        int ld = lead_code(rec + 5, i);
        int l = ld < rb ? ld : rb;
        uint64_t cur = 0;
        int lo = w - rb;
        if (q + (rb - l) > end) return 0;
        for (int b = w - l; b < w; b++) cur |= prev & ((uint64_t)0xFF << (8 * b));
        for (int b = lo; b < w - l; b++) {
            int src = b - lo;
            if (quirk5 && w == 8 && rb == 5 && l == 0 && b == 7) src = 3;
            cur |= (uint64_t)q[src] << (8 * b);
        }
        q += rb - l;
        cur &= mask;
        t[i] = cur;
        prev = cur;
    }
    return (size_t)(q - rec);
}

typedef struct {
    size_t nb, nbit, nconst, bs;
    const uint8_t *bitmap, *cmed, *payload;
} stream_view;

static int parse_header(const uint8_t *s, size_t len, size_t n, stream_view *v)
{
    if (len < 13) return -1;
    v->bs = s[4] ? s[4] : 256;            /* 0 means 256 (documented extension) */
    v->nb = ceil_div(n, v->bs);
    v->nbit = ceil_div(v->nb, 8);
    v->nconst = (size_t)get_be_u64(s + 5);
    if (v->nconst > v->nb) return -2;
    if (13 + v->nbit + 4 * v->nconst > len) return -3;
    v->bitmap = s + 13;
    v->cmed = v->bitmap + v->nbit;
    v->payload = v->cmed + 4 * v->nconst;
    return 0;
}

static int decompress_generic(const uint8_t *s, size_t len, size_t n, int w,
                              void *out, int quirk5)
{
    stream_view v;
    int rc = parse_header(s, len, n, &v);
    if (rc) return rc;
    const uint8_t *q = v.payload, *end = s + len;
    size_t k = 0;
    uint64_t t[256];
    for (size_t b = 0; b < v.nb; b++) {
        size_t m = (b + 1) * v.bs <= n ? v.bs : n - b * v.bs;
        int nonconst = (v.bitmap[b >> 3] >> (7 - (b & 7))) & 1;
        if (!nonconst) {
            if (k >= v.nconst) return -4;
            float med = get_be_f32(v.cmed + 4 * k++);
            for (size_t j = 0; j < m; j++) {
                if (w == 4) ((float *)out)[b * v.bs + j] = med;
                else ((double *)out)[b * v.bs + j] = (double)med;
            }
            continue;
        }
        if (q + 5 + ceil_div(m, 4) > end) return -5;
        int rlen = q[0];
        if (rlen < 1 || rlen > 8 * w) return -6;
        int shift = (8 - rlen % 8) % 8;
        float med = get_be_f32(q + 1);
        size_t used = decode_words(q, m, w, t, quirk5, (size_t)(end - q));
        if (!used) return -7;
        for (size_t j = 0; j < m; j++) {
            if (w == 4) {
                float d = bits_f32((uint32_t)(t[j] << shift));
                ((float *)out)[b * v.bs + j] = d + med;
            } else {
                double d = bits_f64(t[j] << shift);
                ((double *)out)[b * v.bs + j] = d + (double)med;
            }
        }
        q += used;
    }
    if (k != v.nconst) return -8;
    return 0;
}

int szxo_decompress_f32(const uint8_t *s, size_t len, size_t n, float *out)
{
    return decompress_generic(s, len, n, 4, out, 0);
}

int szxo_decompress_f64(const uint8_t *s, size_t len, size_t n, double *out, int ref_quirk)
{
    return decompress_generic(s, len, n, 8, out, ref_quirk);
}

long szxo_record_offsets(const uint8_t *s, size_t len, size_t n, int w, uint64_t *offs)
{
    stream_view v;
    int rc = parse_header(s, len, n, &v);
    if (rc) return rc;
    const uint8_t *q = v.payload, *end = s + len;
    long k = 0;
    for (size_t b = 0; b < v.nb; b++) {
        size_t m = (b + 1) * v.bs <= n ? v.bs : n - b * v.bs;
        if (!((v.bitmap[b >> 3] >> (7 - (b & 7))) & 1)) continue;
        if (q + 5 + ceil_div(m, 4) > end) return -5;
        offs[k++] = (uint64_t)(q - s);
        q += szxo_record_size(q, m, w);
        if (q > end) return -5;
    }
    return k;
}

/* ---- random-access container (fastMode 3 and 4) ------------------------ */

size_t szxo_compress_ra_bound(size_t n, int bs, int w)
{
    size_t nb = ceil_div(n, (size_t)bs);
    return 20 + ceil_div(nb, 8) + nb * (2 + 4 + 5 + ceil_div((size_t)bs, 4) + (size_t)bs * (size_t)w);
}

/* fastMode-4 fp64 block statistics (szx_double.cc:1120-1135): the radius is
 * rounded to float first; the median and the constant test use that float. */
static size_t block_stats_f64_fm4(const double *x, size_t n, float eb, int bs,
                                  uint8_t *state, float *median, float *radius)
{
    size_t nb = ceil_div(n, (size_t)bs), nconst = 0;
    for (size_t b = 0; b < nb; b++) {
        const double *v = x + b * (size_t)bs;
        size_t m = (b + 1) * (size_t)bs <= n ? (size_t)bs : n - b * (size_t)bs;
        double lo = v[0], hi = v[0];
        for (size_t j = 1; j < m; j++) {
            if (lo > v[j]) lo = v[j];
            else if (hi < v[j]) hi = v[j];
        }
        float rad = (float)((hi - lo) / 2);
        float med = (float)(lo + (double)rad);
        state[b] = rad <= eb ? 0 : 1;
        nconst += state[b] == 0;
        median[b] = med;
        radius[b] = rad;
    }
    return nconst;
}

size_t szxo_compress_ra(const void *x, size_t n, float eb, int bs, int w, int fast_mode,
                        uint8_t *out)
{
    size_t nb = ceil_div(n, (size_t)bs);
    size_t nbit = ceil_div(nb, 8);
    uint8_t *state = (uint8_t *)malloc(nb ? nb : 1);
    float *med = (float *)malloc((nb ? nb : 1) * sizeof(float));
    float *rad = (float *)malloc((nb ? nb : 1) * sizeof(float));
    size_t nconst;
    if (w == 4) nconst = szxo_block_stats_f32((const float *)x, n, eb, bs, state, med, rad);
    else if (fast_mode == 4) nconst = block_stats_f64_fm4((const double *)x, n, eb, bs, state, med, rad);
    else nconst = szxo_block_stats_f64((const double *)x, n, eb, bs, state, med, rad);
    size_t nrec = nb - nconst;
    out[0] = 1; out[1] = 1; out[2] = 1; out[3] = 1;   /* szx_float.cc:1438-1441 */
    put_be_u64(out + 4, (uint64_t)bs);                  /* sizeToBytes, :1444-1447 */
    put_be_u64(out + 12, nconst);
    uint8_t *osz = out + 20;                            /* native u16 sizes */
    uint8_t *bitmap = osz + 2 * nrec;
    memset(bitmap, 0, nbit);
    uint8_t *cmed = bitmap + nbit;
    uint8_t *rec = cmed + 4 * nconst;
    size_t k = 0, r = 0;
    for (size_t b = 0; b < nb; b++) {
        size_t m = (b + 1) * (size_t)bs <= n ? (size_t)bs : n - b * (size_t)bs;
        if (state[b]) {
            bitmap[b >> 3] |= (uint8_t)(0x80 >> (b & 7));
            size_t sz = w == 4
                ? szxo_encode_block_f32((const float *)x + b * (size_t)bs, m, eb, med[b], rad[b], rec)
                : szxo_encode_block_f64((const double *)x + b * (size_t)bs, m, eb, med[b], rad[b], rec);
            osz[2 * r] = (uint8_t)sz;                   /* little-endian host */
            osz[2 * r + 1] = (uint8_t)(sz >> 8);
            r++;
            rec += sz;
        } else if (fast_mode == 4) {
            memcpy(cmed + 4 * k++, &med[b], 4);         /* native f32, szx_float.cc:1375 */
        } else {
            put_be_f32(cmed + 4 * k++, med[b]);         /* floatToBytes, :1466 */
        }
    }
    free(state); free(med); free(rad);
    return (size_t)(rec - out);
}

int szxo_decompress_ra(const uint8_t *s, size_t len, size_t n, int w, int fast_mode, void *out,
                       int ref_quirk)
{
    if (len < 20) return -1;
    size_t bs = (size_t)get_be_u64(s + 4);
    if (bs < 1 || bs > 256) return -2;
    size_t nb = ceil_div(n, bs), nbit = ceil_div(nb, 8);
    size_t nconst = (size_t)get_be_u64(s + 12);
    if (nconst > nb) return -2;
    size_t nrec = nb - nconst;
    const uint8_t *osz = s + 20;
    const uint8_t *bitmap = osz + 2 * nrec;
    const uint8_t *cmed = bitmap + nbit;
    const uint8_t *q = cmed + 4 * nconst, *end = s + len;
    if (q > end) return -3;
    uint64_t t[256];
    size_t k = 0, r = 0;
    for (size_t b = 0; b < nb; b++) {
        size_t m = (b + 1) * bs <= n ? bs : n - b * bs;
        if (!((bitmap[b >> 3] >> (7 - (b & 7))) & 1)) {
            if (k >= nconst) return -4;
            float med;
            if (fast_mode == 4) memcpy(&med, cmed + 4 * k, 4);
            else med = get_be_f32(cmed + 4 * k);
            k++;
            for (size_t j = 0; j < m; j++) {
                if (w == 4) ((float *)out)[b * bs + j] = med;
                else ((double *)out)[b * bs + j] = (double)med;
            }
            continue;
        }
        if (r >= nrec) return -4;
        size_t sz = (size_t)osz[2 * r] | ((size_t)osz[2 * r + 1] << 8);
        r++;
        if (q + 5 + ceil_div(m, 4) > end || q + sz > end) return -5;
        int rlen = q[0];
        if (rlen < 1 || rlen > 8 * w) return -6;
        int shift = (8 - rlen % 8) % 8;
        float med = get_be_f32(q + 1);
        if (!decode_words(q, m, w, t, ref_quirk, (size_t)(end - q))) return -7;
        for (size_t j = 0; j < m; j++) {
            if (w == 4) {
                ((float *)out)[b * bs + j] = bits_f32((uint32_t)(t[j] << shift)) + med;
            } else {
                ((double *)out)[b * bs + j] = bits_f64(t[j] << shift) + (double)med;
            }
        }
        q += sz;   /* the reference advances by the stored size (szxd_float.cc:444) */
    }
    if (k != nconst || r != nrec) return -8;
    return 0;
}

/* ---- non-blocked codec (fastMode 1) -------------------------------------- */

size_t szxo_compress_nb_bound(size_t n, int w)
{
    return 17 + ceil_div(n, 4) + n * (size_t)w;
}

size_t szxo_compress_nb(const void *x, size_t n, float eb, int w, uint8_t *out)
{
    float med, rad;
    if (w == 4) {
        const float *v = (const float *)x;
        float lo = v[0], hi = v[0];
        for (size_t i = 0; i < n; i++) {
            if (lo > v[i]) lo = v[i];
            else if (hi < v[i]) hi = v[i];
        }
        rad = (hi - lo) / 2;                 /* szx_float.cc:494-495 */
        med = lo + rad;
    } else {
        const double *v = (const double *)x;
        double lo = v[0], hi = v[0];
        for (size_t i = 0; i < n; i++) {
            if (lo > v[i]) lo = v[i];
            else if (hi < v[i]) hi = v[i];
        }
        rad = (float)((hi - lo) / 2);        /* float *radius, szx_double.cc:1153-1154 */
        med = (float)(lo + (double)rad);
    }
    int clamped;
    int len = szxo_req_length(w, eb, rad, &clamped);
    int shift = (8 - len % 8) % 8;
    /* fp32 subtracts the median after the clamp zeroed it; fp64 the one
     * captured before (szx_float.cc:1516, szx_double.cc:1597,1617) */
    float hdr = clamped ? 0.0f : med;
    uint64_t *s = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    for (size_t i = 0; i < n; i++) {
        if (w == 4) s[i] = f32_bits(((const float *)x)[i] - hdr) >> shift;
        else s[i] = f64_bits(((const double *)x)[i] - (double)med) >> shift;
    }
    /* one record written 12 bytes in: its reqLength byte and median move to
     * bytes 4..8, the big-endian lead-array size takes bytes 9..16 */
    size_t rec = encode_words(s, n, w, len, hdr, out + 12);
    free(s);
    out[0] = 1; out[1] = 1; out[2] = 1; out[3] = 0;   /* szx_float.cc:1506-1509 */
    memmove(out + 4, out + 12, 5);
    put_be_u64(out + 9, ceil_div(n, 4));
    return 12 + rec;
}

int szxo_decompress_nb(const uint8_t *s, size_t len, size_t n, int w, void *out, int ref_quirk)
{
    if (len < 17) return -1;
    int rlen = s[4];
    if (rlen < 1 || rlen > 8 * w) return -6;
    if (get_be_u64(s + 9) != ceil_div(n, 4)) return -2;
    if (17 + ceil_div(n, 4) > len) return -5;
    /* the record as fastMode 2 lays it out: reqLength, median, leads, mids */
    uint8_t *rec = (uint8_t *)malloc(len - 12);
    memcpy(rec, s + 4, 5);
    memcpy(rec + 5, s + 17, len - 17);
    uint64_t *t = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    size_t used = decode_words(rec, n, w, t, ref_quirk, len - 12);
    int shift = (8 - rlen % 8) % 8;
    float med = get_be_f32(s + 5);
    for (size_t j = 0; used && j < n; j++) {
        if (w == 4) ((float *)out)[j] = bits_f32((uint32_t)(t[j] << shift)) + med;
        else ((double *)out)[j] = bits_f64(t[j] << shift) + (double)med;
    }
    free(t);
    free(rec);
    return used ? 0 : -7;
}

/* ---- value range and REL bound (A1, A2) --------------------------------- */

float szxo_value_range_f32(const float *x, size_t n)
{
    float lo = x[0], hi = x[0];
    for (size_t i = 0; i < n; i++) {
        if (lo > x[i]) lo = x[i];
        else if (hi < x[i]) hi = x[i];
    }
    return hi - lo;
}

double szxo_value_range_f64(const double *x, size_t n)
{
    double lo = x[0], hi = x[0];
    for (size_t i = 0; i < n; i++) {
        if (lo > x[i]) lo = x[i];
        else if (hi < x[i]) hi = x[i];
    }
    return hi - lo;
}

/* fp32: eb = fl32(range * rel) (szx.cc:251-252). */
float szxo_rel_eb_f32(const float *x, size_t n, float rel)
{
    float r = szxo_value_range_f32(x, n);
    return r * rel;
}

/* fp64: eb = (float)((double)range * (double)rel) (szx.cc:306-307). */
float szxo_rel_eb_f64(const double *x, size_t n, float rel)
{
    double r = szxo_value_range_f64(x, n);
    return (float)(r * (double)rel);
}
