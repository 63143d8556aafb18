/*
 * szx_caller.c -- a plain C program written against SZx's public C API
 * (szx.h:74-98), linked to libszx_amd.so.  Built twice: against
 * include/szx_amd.h, and (where the reference's sources are present at build
 * time) against the reference's own szx.h, to show that a caller of the
 * reference compiles and links unchanged.  The GPU tests run it on golden
 * inputs and compare the files it writes with the golden outputs.
 *
 *   szx_caller codec   f32|f64 abs|rel <bound> <n> <in> <stream_out> <dec_out>
 *       SZx_fast_compress_args(fastMode 2, blockSize 128) and
 *       SZx_fast_decompress(fastMode 2)
 *   szx_caller blocked f32|f64 <eb> <bs> <n> <in> <stream_out> <dec_out>
 *       SZx_fast_compress_args_unpredictable_blocked, checked byte for byte
 *       against SZx_fast_compress_args_unpredictable_blocked_args, then
 *       SZx_fast_decompress_args_unpredictable_blocked (no length given)
 *   szx_caller ra      f32|f64 <eb> <bs> <n> <in> <stream_out> <dec_out>
 *       the randomaccess_openmp pair (fastMode-4 container, no length given)
 *   szx_caller pred    <n> <stream_in> <pred_in> <dec_out>
 *       SZx_fast_decompress_pred (non-blocked stream + prediction, float)
 *   szx_caller sample  f32|f64 <rate> <bs> <eb> <n> <in> <out_prefix>
 *       computeRadiusBuffer (writes <prefix>.buf/.rad/.med) and
 *       estimateCRbasedonErrorBound_buffered; prints
 *       "range <r> cr <cr> req <sum> lead <sum>"
 *   szx_caller dims
 *       prints SZx_computeDimension, SZx_computeDataLength and
 *       SZx_filterDimension for a fixed list of dims
 *   szx_caller union
 *       checks the byte-view unions' sizes and byte order
 *
 * Exit status 0 on success; otherwise a message on stderr.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef SZX_CALLER_REF_HEADER
#include <szx.h>
#define LAST_ERROR() ""
#else
#include "szx_amd.h"
#define LAST_ERROR() szx_amd_last_error()
#endif

static void *read_file(const char *path, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    void *p = malloc(sz > 0 ? (size_t)sz : 1);
    if (p && sz > 0 && fread(p, 1, (size_t)sz, f) != (size_t)sz) {
        free(p);
        p = NULL;
    }
    fclose(f);
    if (len) *len = (size_t)(sz > 0 ? sz : 0);
    return p;
}

static int write_file(const char *path, const void *p, size_t len)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    size_t w = fwrite(p, 1, len, f);
    fclose(f);
    return w == len ? 0 : -1;
}

static int fail(const char *what)
{
    fprintf(stderr, "szx_caller: %s failed: %s\n", what, LAST_ERROR());
    return 1;
}

static int dtype_of(const char *s, size_t *w)
{
    if (!strcmp(s, "f32")) { *w = 4; return SZx_FLOAT; }
    *w = 8;
    return SZx_DOUBLE;
}

static void *load_values(const char *path, size_t n, size_t w)
{
    size_t len = 0;
    void *x = read_file(path, &len);
    if (!x || len != n * w) {
        fprintf(stderr, "szx_caller: %s holds %zu bytes, expected %zu\n", path, len, n * w);
        free(x);
        return NULL;
    }
    return x;
}

static int run_codec(char **a)
{
    size_t w;
    const int dt = dtype_of(a[0], &w);
    const int rel = !strcmp(a[1], "rel");
    const float bound = strtof(a[2], NULL);
    const size_t n = strtoull(a[3], NULL, 10);
    void *x = load_values(a[4], n, w);
    if (!x) return 1;
    size_t outSize = 0;
    unsigned char *s = SZx_fast_compress_args(SZx_WITH_BLOCK_FAST_CMPR, dt, x, &outSize, rel ? REL : ABS,
                                              rel ? 0.0f : bound, rel ? bound : 0.0f, 0.0f, 0.0f, 0, 0, 0, 0, n);
    if (!s) return fail("SZx_fast_compress_args");
    void *d = SZx_fast_decompress(SZx_WITH_BLOCK_FAST_CMPR, dt, s, outSize, 0, 0, 0, 0, n);
    if (!d) return fail("SZx_fast_decompress");
    const int rc = write_file(a[5], s, outSize) || write_file(a[6], d, n * w);
    free(s);
    free(d);
    free(x);
    return rc;
}

static int run_blocked(char **a)
{
    size_t w;
    const int dt = dtype_of(a[0], &w);
    const float eb = strtof(a[1], NULL);
    const int bs = atoi(a[2]);
    const size_t n = strtoull(a[3], NULL, 10);
    void *x = load_values(a[4], n, w);
    if (!x) return 1;
    size_t outSize = 0;
    unsigned char *s = SZx_fast_compress_args_unpredictable_blocked(x, dt, &outSize, eb, n, bs);
    if (!s) return fail("SZx_fast_compress_args_unpredictable_blocked");
    /* the caller-buffer variant: the reference's callers size it n * W */
    const size_t cap = outSize > n * w ? outSize : n * w;
    unsigned char *s2 = (unsigned char *)malloc(cap);
    const size_t out2 = SZx_fast_compress_args_unpredictable_blocked_args(x, dt, s2, eb, n, bs);
    if (out2 != outSize || memcmp(s, s2, outSize) != 0) {
        fprintf(stderr, "szx_caller: _blocked_args wrote %zu bytes, _blocked %zu (or different bytes)\n", out2,
                outSize);
        return 1;
    }
    void *d = NULL;
    SZx_fast_decompress_args_unpredictable_blocked(&d, dt, n, s);
    if (!d) return fail("SZx_fast_decompress_args_unpredictable_blocked");
    const int rc = write_file(a[5], s, outSize) || write_file(a[6], d, n * w);
    free(s);
    free(s2);
    free(d);
    free(x);
    return rc;
}

static int run_ra(char **a)
{
    size_t w;
    const int dt = dtype_of(a[0], &w);
    const float eb = strtof(a[1], NULL);
    const int bs = atoi(a[2]);
    const size_t n = strtoull(a[3], NULL, 10);
    void *x = load_values(a[4], n, w);
    if (!x) return 1;
    size_t outSize = 0;
    unsigned char *s = SZx_fast_compress_args_unpredictable_blocked_randomaccess_openmp(x, dt, &outSize, eb, n, bs);
    if (!s) return fail("SZx_fast_compress_args_unpredictable_blocked_randomaccess_openmp");
    void *d = SZx_fast_decompress_args_unpredictable_blocked_randomaccess_openmp(dt, n, s);
    if (!d) return fail("SZx_fast_decompress_args_unpredictable_blocked_randomaccess_openmp");
    const int rc = write_file(a[5], s, outSize) || write_file(a[6], d, n * w);
    free(s);
    free(d);
    free(x);
    return rc;
}

static int run_pred(char **a)
{
    const size_t n = strtoull(a[0], NULL, 10);
    size_t len = 0;
    unsigned char *s = (unsigned char *)read_file(a[1], &len);
    float *pred = (float *)load_values(a[2], n, 4);
    if (!s || !pred) return 1;
    float *d = (float *)SZx_fast_decompress_pred(SZx_FLOAT, pred, s, len, 0, 0, 0, 0, n);
    if (!d) return fail("SZx_fast_decompress_pred");
    if (SZx_fast_decompress_pred(SZx_DOUBLE, pred, s, len, 0, 0, 0, 0, n) != NULL) {
        fprintf(stderr, "szx_caller: SZx_fast_decompress_pred decoded double (the reference returns NULL)\n");
        return 1;
    }
    const int rc = write_file(a[3], d, n * 4);
    free(s);
    free(pred);
    free(d);
    return rc;
}

static int run_sample(char **a)
{
    size_t w;
    const int dt = dtype_of(a[0], &w);
    const int rate = atoi(a[1]);
    const int bs = atoi(a[2]);
    const float eb = strtof(a[3], NULL);
    const size_t n = strtoull(a[4], NULL, 10);
    void *x = load_values(a[5], n, w);
    if (!x) return 1;
    float *rad = NULL, *med = NULL;
    void *buf = NULL;
    const float range = computeRadiusBuffer(x, dt, n, rate, bs, &rad, &med, &buf);
    if (!rad || !med || !buf) return fail("computeRadiusBuffer");
    size_t req = 0, lead = 0;
    const float cr = estimateCRbasedonErrorBound_buffered(dt, eb, buf, med, rad, rate, bs, n, &req, &lead);
    const size_t ns = n / ((size_t)rate * bs);
    char path[4096];
    int rc = 0;
    snprintf(path, sizeof path, "%s.buf", a[6]);
    rc |= write_file(path, buf, ns * bs * w);
    snprintf(path, sizeof path, "%s.rad", a[6]);
    rc |= write_file(path, rad, ns * 4);
    snprintf(path, sizeof path, "%s.med", a[6]);
    rc |= write_file(path, med, ns * 4);
    printf("range %.9g cr %.9g req %zu lead %zu\n", range, cr, req, lead);
    free(rad);
    free(med);
    free(buf);
    free(x);
    return rc;
}

static int run_dims(void)
{
    static const size_t dims[][5] = {
        {0, 0, 0, 0, 7},     {0, 0, 0, 1, 7},    {0, 0, 0, 7, 1},    {0, 0, 1, 1, 1},
        {0, 0, 100, 500, 500}, {0, 1, 9, 1, 5},  {1, 1, 1, 1, 1},    {2, 1, 3, 1, 4},
        {1, 2, 1, 3, 1},     {0, 0, 0, 0, 0},    {0, 4, 1, 1, 6},    {3, 1, 1, 2, 1},
    };
    for (size_t i = 0; i < sizeof dims / sizeof dims[0]; i++) {
        const size_t *d = dims[i];
        size_t c[5] = {0, 0, 0, 0, 0};
        const int dim = SZx_computeDimension(d[0], d[1], d[2], d[3], d[4]);
        const size_t len = SZx_computeDataLength(d[0], d[1], d[2], d[3], d[4]);
        const int ch = SZx_filterDimension(d[0], d[1], d[2], d[3], d[4], c);
        printf("%d %zu %d %zu %zu %zu %zu %zu\n", dim, len, ch, c[0], c[1], c[2], c[3], c[4]);
    }
    return 0;
}

static int run_union(void)
{
    lfloat f;
    ldouble g;
    lint16 s;
    lint32 i;
    lint64 l;
    f.value = 1.0f;
    g.value = 1.0;
    s.usvalue = 0x0102;
    i.uivalue = 0x01020304u;
    l.ulvalue = 0x0102030405060708ul;
    if (sizeof f != 4 || sizeof g != 8 || sizeof s != 2 || sizeof i != 4 || sizeof l != 8) return 1;
    if (f.ivalue != 0x3f800000u || f.byte[3] != 0x3f) return 1;
    if (g.lvalue != 0x3ff0000000000000ul || g.byte[7] != 0x3f) return 1;
    if (s.byte[0] != 2 || i.byte[0] != 4 || l.byte[0] != 8) return 1;
    printf("unions ok\n");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && !strcmp(argv[1], "dims")) return run_dims();
    if (argc >= 2 && !strcmp(argv[1], "union")) return run_union();
    if (argc == 9 && !strcmp(argv[1], "codec")) return run_codec(argv + 2);
    if (argc == 9 && !strcmp(argv[1], "blocked")) return run_blocked(argv + 2);
    if (argc == 9 && !strcmp(argv[1], "ra")) return run_ra(argv + 2);
    if (argc == 6 && !strcmp(argv[1], "pred")) return run_pred(argv + 2);
    if (argc == 9 && !strcmp(argv[1], "sample")) return run_sample(argv + 2);
    fprintf(stderr, "usage: see the head of szx_caller.c\n");
    return 2;
}
