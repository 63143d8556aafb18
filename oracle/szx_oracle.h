/*
 * szx_oracle.h -- CPU restatement of SZx's blockwise fast codec (fastMode 2).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the GPU codec is compared
 * against; it is never linked into, loaded by, or called from the product
 * library (libszx_amd.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * The code is written from the behaviour, not copied: one generic encoder and
 * decoder serve both word widths (4-byte float, 8-byte double) where the
 * reference spells out one unrolled loop per reqBytes value.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - golden streams/decodes produced by the compiled reference (tests/golden/)
 *   - the compiled reference itself (oracle/_ref) on random inputs, when built
 *   - closed forms: exhaustive reqLength table, hand-built streams, size formula,
 *     all-constant fields, round-trip error bound.
 */
#ifndef SZX_ORACLE_H
#define SZX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Worst-case stream size of the blocked codec for n elements of width w. */
size_t szxo_compress_bound(size_t n, int bs, int w);

/* Per-block min/max -> median, radius, state (1 = non-constant).
 * Returns the number of constant blocks.  Arrays hold ceil(n/bs) entries.
 * fp32: szx_float.cc:314-371.  fp64: szx_double.cc:973-1030 (median and radius
 * are rounded to float; the constant test uses the double radius). */
size_t szxo_block_stats_f32(const float *x, size_t n, float eb, int bs,
                            uint8_t *state, float *median, float *radius);
size_t szxo_block_stats_f64(const double *x, size_t n, float eb, int bs,
                            uint8_t *state, float *median, float *radius);

/* reqLength of a non-constant block (szx_float.cc:1708-1717,
 * szx_double.cc:2020-2031).  *clamped is set to 1 when the length hit the
 * upper clamp (the header median is then stored as 0). */
int szxo_req_length(int w, float eb, float radius, int *clamped);

/* Encode one non-constant block record; returns its size in bytes.
 * fp32: szx_float.cc:49-252.  fp64: szx_double.cc:34-433. */
size_t szxo_encode_block_f32(const float *x, size_t n, float eb, float median,
                             float radius, uint8_t *out);
size_t szxo_encode_block_f64(const double *x, size_t n, float eb, float median,
                             float radius, uint8_t *out);

/* Size of the record starting at rec for a block of n elements (parses the
 * reqLength byte and the lead codes).  Mirrors the byte count returned by
 * szxd_float.cc:34-244 / szxd_double.cc:25-449. */
size_t szxo_record_size(const uint8_t *rec, size_t n, int w);

/* Whole-stream compress (szx_float.cc:1059-1139 / szx_double.cc:1160-1238).
 * out must hold szxo_compress_bound() bytes.  Returns the stream size. */
size_t szxo_compress_f32(const float *x, size_t n, float eb, int bs, uint8_t *out);
size_t szxo_compress_f64(const double *x, size_t n, float eb, int bs, uint8_t *out);

/* Whole-stream decompress (szxd_float.cc:247-311 / szxd_double.cc:452-516).
 * Returns 0 on success, <0 on a malformed/truncated stream.
 * ref_quirk (fp64 only): 1 reproduces the reference decoder's reqBytes=5,
 * lead=0 byte-7 read (szxd_double.cc:244 reads q[3] for byte 7);
 * 0 decodes the bytes the encoder actually wrote. */
int szxo_decompress_f32(const uint8_t *s, size_t len, size_t n, float *out);
int szxo_decompress_f64(const uint8_t *s, size_t len, size_t n, double *out,
                        int ref_quirk);

/* Byte offset (from stream start) of every non-constant record, in block
 * order, found by the serial walk of szxd_float.cc:276-283.  Returns the
 * number of records written to offs, or <0 on a malformed stream. */
long szxo_record_offsets(const uint8_t *s, size_t len, size_t n, int w,
                         uint64_t *offs);

/* Random-access container (fastMode 3 and 4): [1,1,1,1], big-endian u64
 * blockSize, big-endian u64 constant-block count, a native-endian u16 size
 * per non-constant record, the state bitmap, the constant medians
 * (big-endian f32 for fastMode 3, native f32 for fastMode 4), then the same
 * records as fastMode 2.
 *   fastMode 3: szx_float.cc:1407-1489, szx_double.cc:1510-1590 (block
 *               statistics as fastMode 2).
 *   fastMode 4: szx_float.cc:1218-1405, szx_double.cc:1320-1507; its block
 *               statistics (computeStateMedianRadius_*2, szx_float.cc:461-476,
 *               szx_double.cc:1120-1135) differ for fp64: the radius is rounded
 *               to float before the constant test (float compare) and before
 *               the median (median = (float)(min + (double)radius_f32)).
 *               Min/max equal the sequential ones for NaN-free data.
 * out must hold szxo_compress_ra_bound() bytes.  Returns the stream size. */
size_t szxo_compress_ra_bound(size_t n, int bs, int w);
size_t szxo_compress_ra(const void *x, size_t n, float eb, int bs, int w, int fast_mode,
                        uint8_t *out);

/* Decode a random-access stream (szxd_float.cc:313-476,
 * szxd_double.cc:518-681): records are located through the size table.
 * Returns 0 or <0 on a malformed stream.  ref_quirk as for fastMode 2. */
int szxo_decompress_ra(const uint8_t *s, size_t len, size_t n, int w, int fast_mode, void *out,
                       int ref_quirk);

/* Non-blocked codec (fastMode 1): one record over all n values against the
 * global median -- [1,1,1,0], reqLength byte, big-endian f32 median,
 * big-endian u64 lead-code bytes, lead codes, mid bytes
 * (szx_float.cc:1492-1699, szx_double.cc:1593-2010).  The median and radius
 * come from the value range (computeValueRange_*, szx_float.cc:478-499,
 * szx_double.cc:1137-1158: for fp64 the radius is rounded to float before the
 * median is formed).  Decoder: szxd_float.cc:478-662, szxd_double.cc:683-1000
 * (ref_quirk as for fastMode 2).  out must hold szxo_compress_nb_bound(). */
size_t szxo_compress_nb_bound(size_t n, int w);
size_t szxo_compress_nb(const void *x, size_t n, float eb, int w, uint8_t *out);
int szxo_decompress_nb(const uint8_t *s, size_t len, size_t n, int w, void *out, int ref_quirk);

/* The stream of the reference's CUDA codec, cuSZx (fp32;
 * szx_oracle_cuszx.c): compress (cuszx_float.cu:222-298 and
 * cuszx_entry.cu:20-70; bs a multiple of 32 in 32..224; returns the size or <0) and
 * decompress (cuszxd_float.cu:161-294 and cuszx_entry.cu:129-209; returns 0
 * or <0 on a stream the reference has no defined result for). */
size_t szxo_cuszx_bound(size_t n, int bs);
long szxo_cuszx_compress(const float *x, size_t n, float eb, int bs, uint8_t *out);
int szxo_cuszx_decompress(const uint8_t *s, size_t len, size_t n, float *out);

/* Value range and REL -> ABS error bound (szx.cc:249-253, 303-308;
 * szx_float.cc:478-499, szx_double.cc:1137-1158). */
float szxo_value_range_f32(const float *x, size_t n);
double szxo_value_range_f64(const double *x, size_t n);
float szxo_rel_eb_f32(const float *x, size_t n, float rel);
float szxo_rel_eb_f64(const double *x, size_t n, float rel);

/* Fixed-rate (FXR) bound search restated from the reference
 * (szx_oracle_fxr.c).  w is the word width (4 or 8). */
float szxo_fxr_sample(const void *x, int w, size_t n, int rate, int bs,
                      void *buf, float *med, float *rad);
float szxo_fxr_estimate_cr(float eb, const void *buf, int w, const float *med,
                           const float *rad, int rate, int bs, size_t n,
                           size_t *sum_req, size_t *sum_lead);
float szxo_fxr_search(float target, float tol, int rate, float init_eb, int bs,
                      size_t n, const void *buf, int w, const float *med,
                      const float *rad, int *n_est);
float szxo_fxr_error_bound(const void *x, int w, size_t n, float target_cr,
                           float tol, int bs, int *n_est);

#ifdef __cplusplus
}
#endif
#endif
