/*
 * szx_oracle_cuszx.c -- CPU restatement of the stream written and read by the
 * reference's CUDA codec, cuSZx (fp32 only).
 *
 * TEST INFRASTRUCTURE ONLY (see szx_oracle.h): the checker for the GPU reader
 * and writer of this format; never linked into libszx_amd.so.
 *
 * Stream (host assembly, cuszx_entry.cu:20-70):
 *   [1, 1, 1, 0, (u8)bs]
 *   big-endian u64 constant-block count            (szx_BytesToolkit.cc:811-814)
 *   state bitmap over the nb = n / bs full blocks, MSB first, 1 = non-constant
 *                                                  (szx_TypeManager.cc:16-40)
 *   the n % bs trailing values, raw native f32
 *   native f32 median of every constant block, block order
 *   native i16 mid-byte count of every non-constant block
 *                                                  (szx_BytesToolkit.cc:430-443)
 *   per non-constant block: native f32 median, reqLength byte, bs/4 lead-code
 *   bytes (2 bits a value, MSB first), then its mid bytes
 *
 * The decoder below is a literal restatement of the reference's kernel
 * (cuszxd_float.cu:161-294): every value ("thread") builds its word bytes and
 * a source-position word from its lead code and the prefix sum of mid-byte
 * counts, the position words are scanned with the reference's per-byte max
 * (_compareByte, :73-123), and the inherited bytes are taken from the source
 * values' words.  The reference's kernel overwrites those words while other
 * threads may still read them (:269-288, no barrier between); a source word's
 * byte is one the source wrote itself, or a zero byte of value 0, and both are
 * the same before and after the overwrite, so reading the first-written words
 * is the reference's result on every stream.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "szx_oracle.h"

static uint32_t cz_bits(float v) { uint32_t u; memcpy(&u, &v, 4); return u; }
static float cz_float(uint32_t u) { float v; memcpy(&v, &u, 4); return v; }

static uint64_t cz_be_u64(const uint8_t *p)
{
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v = (v << 8) | p[k];
    return v;
}

/* Lead bytes of a shifted word against the previous one
 * (cuszx_float.cu:147-167).  For 4 significant bytes an equal word gives 4,
 * which the 2-bit code field cannot hold (see cz_pack_codes). */
static int cz_lead(int rb, uint32_t x)
{
    switch (rb) {
    case 1: return (x >> 24) == 0;
    case 2: return (x >> 16) == 0 ? 2 : (x >> 24) == 0 ? 1 : 0;
    case 3: return (x >> 8) == 0 ? 3 : (x >> 16) == 0 ? 2 : (x >> 24) == 0 ? 1 : 0;
    case 4: return x == 0 ? 4 : (x >> 8) == 0 ? 3 : (x >> 16) == 0 ? 2 : (x >> 24) == 0 ? 1 : 0;
    }
    return 0;
}

/* Four lead counts into one code byte (cuszx_float.cu:38-51): value j of the
 * group is shifted left by 2 (3 - j) and the four are OR-ed, then truncated to
 * a byte, so a count of 4 lands in the neighbouring field. */
static uint8_t cz_pack_codes(const int *lead4)
{
    uint32_t v = 0;
    for (int j = 0; j < 4; j++) v |= (uint32_t)lead4[j] << (2 * (3 - j));
    return (uint8_t)v;
}

size_t szxo_cuszx_bound(size_t n, int bs)
{
    if (bs <= 0) return 0;
    const size_t nb = n / (size_t)bs;
    return 13 + (nb + 7) / 8 + 4 * (n % (size_t)bs) + nb * (2 + 5 + (size_t)bs / 4 + 4 * (size_t)bs);
}

/* Compress (cuszx_float.cu:222-298 per block, _compute_oneBlock :123-220,
 * _compute_reqLength :53-60; host assembly cuszx_entry.cu:20-70).
 * bs is a multiple of 32 in 32..224 (one byte; bs/32 warps of 32 threads).
 * The reference's block reduction (cuszx_float.cu:254-265) combines the bs/32
 * warp results by halving, which covers the whole block for 32, 64 and 128
 * only; for 96, 160, 192 and 224 its result is undefined and this takes the
 * minimum and maximum of the whole block.  Returns the stream size, or <0 when
 * a block needs more than 32 significant bits (the reference then writes mid
 * bytes it never filled; its decoder refuses such a block anyway). */
long szxo_cuszx_compress(const float *x, size_t n, float eb, int bs, uint8_t *out)
{
    if (bs <= 0 || bs % 32 != 0 || bs > 224) return -1;
    const size_t nb = n / (size_t)bs, rem = n % (size_t)bs;
    const size_t nlead = (size_t)bs / 4, msize = 5 + nlead;
    uint8_t *state = (uint8_t *)calloc(nb ? nb : 1, 1);
    float *med = (float *)malloc((nb ? nb : 1) * sizeof(float));
    uint8_t *rec = (uint8_t *)malloc((nb ? nb : 1) * (msize + 4 * (size_t)bs));
    int16_t *cnt = (int16_t *)calloc(nb ? nb : 1, sizeof(int16_t));
    if (!state || !med || !rec || !cnt) {
        free(state); free(med); free(rec); free(cnt);
        return -2;
    }
    const int reqExpo = (int)((cz_bits(eb) & 0x7F800000u) >> 23) - 127;
    size_t nconst = 0;
    long rc = 0;
    for (size_t b = 0; b < nb && rc == 0; b++) {
        const float *v = x + b * (size_t)bs;
        float mn = v[0], mx = v[0];
        for (int i = 1; i < bs; i++) {
            mn = v[i] < mn ? v[i] : mn;
            mx = v[i] > mx ? v[i] : mx;
        }
        const float radius = (mx - mn) / 2;
        const float median = mn + radius;
        med[b] = median;
        state[b] = radius <= eb ? 0 : 1;
        if (!state[b]) {
            nconst++;
            continue;
        }
        const int radExpo = (int)((cz_bits(radius) & 0x7F800000u) >> 23) - 127;
        const int reqLength = 9 + radExpo - reqExpo + 1;
        if (reqLength > 32) {
            rc = -3;
            break;
        }
        const int rb = reqLength % 8 ? reqLength / 8 + 1 : reqLength / 8;
        const int shift = reqLength % 8 ? 8 - reqLength % 8 : 0;
        uint8_t *r = rec + b * (msize + 4 * (size_t)bs);
        memcpy(r, &median, 4);
        r[4] = (uint8_t)reqLength;
        uint8_t *mid = r + msize;
        size_t m_total = 0;
        uint32_t prev = 0;
        int lead4[4];
        for (int i = 0; i < bs; i++) {
            const float d = v[i] - median;
            const uint32_t cur = cz_bits(d) >> shift;
            const int ld = cz_lead(rb, cur ^ prev);
            prev = cur;
            lead4[i & 3] = ld;
            if ((i & 3) == 3) r[5 + i / 4] = cz_pack_codes(lead4);
            const int m = rb - ld;   /* 0..rb */
            for (int j = 0; j < m; j++) mid[m_total++] = (uint8_t)(cur >> (8 * (4 - rb + j)));
        }
        cnt[b] = (int16_t)m_total;
    }
    if (rc == 0) {
        uint8_t *p = out;
        p[0] = 1; p[1] = 1; p[2] = 1; p[3] = 0; p[4] = (uint8_t)bs;
        for (int k = 0; k < 8; k++) p[5 + k] = (uint8_t)((uint64_t)nconst >> (56 - 8 * k));
        p += 13;
        for (size_t i = 0; i < (nb + 7) / 8; i++) {
            uint8_t byte = 0;
            for (int j = 0; j < 8 && 8 * i + j < nb; j++) byte |= (uint8_t)(state[8 * i + j] << (7 - j));
            *p++ = byte;
        }
        memcpy(p, x + nb * (size_t)bs, 4 * rem);
        p += 4 * rem;
        for (size_t b = 0; b < nb; b++)
            if (!state[b]) {
                memcpy(p, &med[b], 4);
                p += 4;
            }
        for (size_t b = 0; b < nb; b++)
            if (state[b]) {
                memcpy(p, &cnt[b], 2);
                p += 2;
            }
        for (size_t b = 0; b < nb; b++)
            if (state[b]) {
                const size_t len = msize + (size_t)(uint16_t)cnt[b];
                memcpy(p, rec + b * (msize + 4 * (size_t)bs), len);
                p += len;
            }
        rc = (long)(p - out);
    }
    free(state); free(med); free(rec); free(cnt);
    return rc;
}

/* _compareByte (cuszxd_float.cu:73-123): per-byte maximum over the bytes the
 * word width uses; for one byte the whole words are compared as ints. */
static uint32_t cz_compare(uint32_t pre, uint32_t cur, int rb)
{
    if (rb == 1) return (int32_t)pre > (int32_t)cur ? pre : cur;
    const int nbyte = rb >= 2 && rb <= 4 ? rb : 0;
    for (int k = 0; k < nbyte; k++) {
        const uint32_t mask = 0xFFu << (8 * k);
        if ((pre & mask) > (cur & mask)) cur = (cur & ~mask) | (pre & mask);
    }
    return cur;
}

/* One non-constant block (cuszxd_float.cu:182-293).  buf holds the block's
 * 4 bs bytes: its record, zero past the record (cuszx_entry.cu:148-149,170).
 * Returns 0, or -7 when a mid-byte position falls outside the buffer. */
static int cz_decode_block(const uint8_t *buf, int bs, float *y)
{
    const int msize = 5 + bs / 4;
    float median;
    memcpy(&median, buf, 4);
    const int reqLength = buf[4];
    const int rb = reqLength % 8 ? reqLength / 8 + 1 : reqLength / 8;
    const int shift = reqLength % 8 ? 8 - reqLength % 8 : 0;
    uint8_t tmp[256][4];
    uint32_t pos[256];
    int lead[256];
    long sum = 0;
    for (int t = 0; t < bs; t++) {
        lead[t] = (buf[5 + (t >> 2)] >> (6 - ((t & 3) << 1))) & 3;
        const int m = rb - lead[t];
        sum += m;
        memset(tmp[t], 0, 4);
        pos[t] = 0;
        /* byte k of the word is tmp[t][k] (k = 0 lowest); position byte 0, 1,
         * 2, 3 names the source of word byte 3, 2, 1, 0 */
        if (rb >= 1 && rb <= 4 && m >= 1 && m <= rb) {
            for (int j = 0; j < m; j++) {
                const long at = msize + sum - m + j;
                if (at < 0 || at >= 4L * bs) return -7;
                const int k = 4 - rb + j;
                tmp[t][k] = buf[at];
                pos[t] |= (uint32_t)t << (8 * (3 - k));
            }
        }
        if (rb == 1) pos[t] |= (uint32_t)t;   /* cuszxd_float.cu:235-239 */
    }
    for (int t = 1; t < bs; t++) pos[t] = cz_compare(pos[t - 1], pos[t], rb);
    for (int t = 0; t < bs; t++) {
        uint8_t v[4];
        memcpy(v, tmp[t], 4);
        const int L = lead[t];
        /* cuszxd_float.cu:273-287: word byte 3 - i from the source named by
         * position byte i, for i < lead */
        for (int i = 0; i < L; i++) v[3 - i] = tmp[(pos[t] >> (8 * i)) & 0xFF][3 - i];
        uint32_t word = (uint32_t)v[0] | (uint32_t)v[1] << 8 | (uint32_t)v[2] << 16 | (uint32_t)v[3] << 24;
        word <<= shift;
        y[t] = cz_float(word) + median;
    }
    return 0;
}

/* Whole-stream decompress (cuszx_entry.cu:129-209).  The reference takes no
 * stream length; len bounds every read here.  Returns 0, or <0 where the
 * reference has no defined result or stops:
 *   -1 stream shorter than its fixed sections
 *   -2 block size not a multiple of 32 in 32..224 (the reference launches
 *      bs/32 warps of 32 threads and leaves other sizes partly undecoded)
 *   -3 more constant blocks than blocks
 *   -4 records past the end of the stream
 *   -5 the bitmap's constant blocks differ from the count
 *   -6 a record longer than 4 bs bytes (the reference prints a warning and
 *      exits, cuszx_entry.cu:165-169) or shorter than zero
 *   -7 a mid-byte position outside its block */
int szxo_cuszx_decompress(const uint8_t *s, size_t len, size_t n, float *out)
{
    if (len < 13) return -1;
    const int bs = s[4];
    if (bs == 0 || bs % 32 != 0 || bs > 224) return -2;
    const size_t nb = n / (size_t)bs, rem = n % (size_t)bs;
    const uint64_t nconst = cz_be_u64(s + 5);
    if (nconst > nb) return -3;
    const size_t nc = nb - (size_t)nconst;
    const size_t nbit = (nb + 7) / 8, msize = 5 + (size_t)bs / 4;
    const size_t q_bits = 13, q_rem = q_bits + nbit, q_med = q_rem + 4 * rem;
    const size_t q_off = q_med + 4 * (size_t)nconst, q_rec = q_off + 2 * nc;
    if (q_rec > len) return -1;
    size_t ones = 0;
    for (size_t b = 0; b < nb; b++) ones += (s[q_bits + b / 8] >> (7 - b % 8)) & 1;
    if (ones != nc) return -5;
    memcpy(out + nb * (size_t)bs, s + q_rem, 4 * rem);
    uint8_t buf[4 * 224];
    size_t p = q_rec, j = 0, c = 0;
    for (size_t b = 0; b < nb; b++) {
        float *y = out + b * (size_t)bs;
        if (!((s[q_bits + b / 8] >> (7 - b % 8)) & 1)) {
            float m;
            memcpy(&m, s + q_med + 4 * c++, 4);
            for (int i = 0; i < bs; i++) y[i] = m;
            continue;
        }
        int16_t sh;
        memcpy(&sh, s + q_off + 2 * j++, 2);
        const long leng = (long)sh + (long)msize;
        if (leng < 0 || leng > 4L * bs) return -6;
        if (p + (size_t)leng > len) return -4;
        memset(buf, 0, sizeof buf);
        memcpy(buf, s + p, (size_t)leng);
        p += (size_t)leng;
        const int rc = cz_decode_block(buf, bs, y);
        if (rc) return rc;
    }
    return 0;
}
