/*
 * szx_amd.h -- C ABI of libszx_amd.so, an MI355X (gfx950) implementation of
 * SZx's fast codecs: the blockwise one (fastMode = SZx_WITH_BLOCK_FAST_CMPR),
 * its random-access container (fastMode 3 / 4), the non-blocked one
 * (fastMode 1), and a reader and writer of the stream the reference's CUDA
 * codec (cuSZx) produces.
 *
 * Two layers:
 *
 *  1. Drop-in host API.  Every function szx.h:74-98 declares (and the two of
 *     cuszx_entry.h), with the same names, argument meaning and ownership,
 *     plus szx.h's byte-view unions; a caller relinks against libszx_amd.so
 *     unchanged.
 *     Host pointers in, malloc'd host buffers out (release with free()).
 *     Every fastMode (1, 2, 3, 4) runs on the GPU.  Failures return NULL / 0 and set
 *     szx_amd_last_error() instead of exiting the process.
 *
 *  2. Device-resident API (szx_amd_*_dev).  Device pointers, a caller
 *     workspace and a HIP stream; nothing is allocated, nothing is copied to
 *     the host, and every call is asynchronous on `stream`.
 *
 * Streams are byte-identical to the reference CPU codec at the same
 * parameters (szx_float.cc:1059-1139, szx_double.cc:1160-1238; the other
 * modes cited at their entry points) and decode bit-identically to the
 * reference decoder (szxd_float.cc:247-311, szxd_double.cc:452-516).
 */
#ifndef SZX_AMD_H
#define SZX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (szx_defines.h:19-21, 36-37, 54-57, 60-66) ------------------ */
#ifndef SZx_FLOAT
#define ABS 0
#define REL 1
#define FXR 2
#define SZx_FLOAT 0
#define SZx_DOUBLE 1
#define SZx_NO_BLOCK_FAST_CMPR 1
#define SZx_WITH_BLOCK_FAST_CMPR 2
#define SZx_RANDOMACCESS_FAST_CMPR 3
#define SZx_OPENMP_FAST_CMPR 4
#define SZx_SCES 0
#define SZx_NSCS -1
#define SZx_FERR -2
#define SZx_TERR -3
#define SZx_DERR -4
#define SZx_MERR -5
#define SZx_BERR -6
#endif

/* ---- byte-view unions (szx.h:39-72) -----------------------------------------
 * Same layout as the reference's; skipped when the reference's szx.h was
 * included first. */
#ifndef _SZX_H
typedef union lint16 {
    unsigned short usvalue;
    short svalue;
    unsigned char byte[2];
} lint16;

typedef union lint32 {
    int ivalue;
    unsigned int uivalue;
    unsigned char byte[4];
} lint32;

typedef union lint64 {
    long lvalue;
    unsigned long ulvalue;
    unsigned char byte[8];
} lint64;

typedef union ldouble {
    double value;
    unsigned long lvalue;
    unsigned char byte[8];
} ldouble;

typedef union lfloat {
    float value;
    unsigned int ivalue;
    unsigned char byte[4];
} lfloat;
#endif

/* ---- 1. drop-in host API ---------------------------------------------------- */

/* szx.h:74 / szx.cc:41-71.  Number of dims given: r1 first, the first zero
 * ends them (0..5). */
int SZx_computeDimension(size_t r5, size_t r4, size_t r3, size_t r2, size_t r1);

/* szx.h:76 / szx.cc:109-229.  correctedDimension[0..4] = {r1..r5} with the
 * dims of extent 1 removed the way the reference removes them; returns 1 if
 * a dim was removed, else 0. */
int SZx_filterDimension(size_t r5, size_t r4, size_t r3, size_t r2, size_t r1, size_t *correctedDimension);

/* szx.h:75 / szx.cc:73-101.  Product of the non-zero dims, r1 fastest. */
size_t SZx_computeDataLength(size_t r5, size_t r4, size_t r3, size_t r2, size_t r1);

/* szx.h:77-78 / szx.cc:231-348.  Compress a host array.  blockSize is 128,
 * as in the reference.  fastMode 2 writes the blocked stream; fastMode 3 and
 * 4 the random-access container (szx_float.cc:1407-1489 and 1218-1405: a u16
 * size per record before the bitmap; 3 stores big-endian medians, 4 native
 * ones and rounds fp64 block radii to float before the constant test, as the
 * reference's OpenMP compressor does).  errBoundMode ABS uses absErrBound, REL uses
 * relBoundRatio * (max - min) computed with the reference's float rounding,
 * FXR (fixed rate, any fastMode as in szx.cc:238-243) picks the bound the
 * reference's fixed-rate compressor picks for `compressionRatio` within
 * `tolerance` (szx_float.cc:921-949) and compresses with it, so the stream is
 * the reference's.  Returns a malloc'd stream and its size in *outSize; NULL
 * on error. */
unsigned char *SZx_fast_compress_args(int fastMode, int dataType, void *data, size_t *outSize,
                                      int errBoundMode, float absErrBound, float relBoundRatio,
                                      float compressionRatio, float tolerance, size_t r5,
                                      size_t r4, size_t r3, size_t r2, size_t r1);

/* szx.h:80 / szx.cc:378-430.  Decompress a fastMode-2 stream, or a fastMode
 * 3 / 4 random-access container (szxd_float.cc:313-476), into a malloc'd
 * host array of SZx_computeDataLength(r5..r1) values.  byteLength bounds
 * every read (the reference ignores it).  NULL on a malformed stream. */
void *SZx_fast_decompress(int fastMode, int dataType, unsigned char *bytes, size_t byteLength,
                          size_t r5, size_t r4, size_t r3, size_t r2, size_t r1);

/* szx.h:89 / szx.cc:466-473.  Blocked compressor with an explicit blockSize
 * in [1, 256] (256 is stored as the byte 0, as the reference's (u8) cast
 * does).  absErrBound is absolute. */
unsigned char *SZx_fast_compress_args_unpredictable_blocked(void *oriData, int dataType,
                                                            size_t *outSize, float absErrBound,
                                                            size_t nbEle, int blockSize);

/* szx.h:98 / szx.cc:514-524.  newData is really T**: *newData receives a
 * malloc'd array.  Unlike the reference (whose fp64 branch is a no-op) both
 * float and double are decoded.  The stream length is not passed in this
 * signature, so it is read from the stream itself (header, bitmap, records). */
void SZx_fast_decompress_args_unpredictable_blocked(void *newData, int dataType, size_t nbEle,
                                                    unsigned char *cmpBytes);

/* szx.h:81 / szx.cc:432-438.  max - min of the array (radius and median are
 * returned when the pointers are non-NULL, as float).  The extremes are found
 * on the GPU with the reference's semantics (szx_float.cc:478-499: NaN if the
 * first value is NaN, other NaNs skipped). */
double SZx_computeValueRange(void *oriData, int dataType, size_t length, float *radius,
                             float *medianValue);

/* szx.h:87 / szx.cc:458-464.  Error bound whose modelled compression ratio
 * is within tolerance of targetCompressionRatio, searched from
 * initErrorBound.  The reference's own entry reaches an estimator that exits
 * the process (szx_float.cc:680); this runs the search its fixed-rate
 * compressor uses (szx_float.cc:684-832, szx_double.cc:682-829) on a 1-in-10
 * block sample, with the model sums computed on the GPU.  blockSize 1..256.
 * Returns 0 and sets szx_amd_last_error() on failure. */
float SZx_estimateErrorBoundbasedonCR(int dataType, float targetCompressionRatio, float tolerance,
                                      void *data, float initErrorBound, int blockSize,
                                      size_t nbEle);

/* szx.h:97 / szx.cc:501-512 (szx_float.cc:980-1057).  The blocked stream
 * written into the caller's buffer outputBytes, which the caller sizes (the
 * reference writes without a bound; szx_amd_compress_bound() is the largest
 * stream); returns the stream size, 0 on error.  The reference returns 0
 * for double (its call is commented out); this writes the double stream,
 * the bytes SZx_fast_compress_args_unpredictable_blocked writes. */
size_t SZx_fast_compress_args_unpredictable_blocked_args(void *oriData, int dataType, unsigned char *outputBytes,
                                                         float absErrBound, size_t nbEle, int blockSize);

/* szx.h:82-83 / szx.cc:441-448 (szx_float.cc:1218-1405,
 * szx_double.cc:1320-1507).  The fastMode-4 random-access container at an
 * explicit blockSize: powers of two 4..256 (others return NULL). */
unsigned char *SZx_fast_compress_args_unpredictable_blocked_randomaccess_openmp(void *oriData, int dataType,
                                                                               size_t *outSize, float absErrBound,
                                                                               size_t nbEle, int blockSize);

/* szx.h:84 / szx.cc:450-456 (szxd_float.cc:313-399, szxd_double.cc:518-605).
 * Decode a fastMode-4 container of nbEle values into a malloc'd array; the
 * container's length comes from its own size table. */
void *SZx_fast_decompress_args_unpredictable_blocked_randomaccess_openmp(int dataType, size_t nbEle,
                                                                        unsigned char *cmpBytes);

/* szx.h:79 / szx.cc:353-376 (szxd_float.cc:25-32; deprecated in the
 * reference).  Decode a non-blocked (fastMode 1) float stream and add
 * preData[i] to each value; NULL for double, as in the reference. */
void *SZx_fast_decompress_pred(int dataType, float *preData, unsigned char *curBytes, size_t byteLength,
                               size_t r5, size_t r4, size_t r3, size_t r2, size_t r1);

/* szx.h:91 / szx.cc:475-485 (szx_float.cc:254-312, szx_double.cc:437-495).
 * Sample every samplingRate-th block of blockSize values: *buffer (really
 * T**) receives a malloc'd copy of the sampled blocks, *radiusArray and
 * *mediusArray malloc'd per-block radius and median (float); returns the
 * value range of the sampled blocks.  Release the three arrays with free(). */
float computeRadiusBuffer(void *oriData, int dataType, size_t nbEle, int samplingRate, int blockSize,
                          float **radiusArray, float **mediusArray, void *buffer);

/* szx.h:93-94 / szx.cc:487-499 (szx_float.cc:501-584, szx_double.cc:497-582).
 * Modelled compression ratio at errorBound of the blocks computeRadiusBuffer
 * sampled; *sumReqNbBytes and *sum_actual_leadNumbers receive the model's
 * sums.  0 on error. */
float estimateCRbasedonErrorBound_buffered(int dataType, float errorBound, void *buffer, float *medianArray,
                                           float *radiusArray, int samplingRate, int blockSize, size_t nbEle,
                                           size_t *sumReqNbBytes, size_t *sum_actual_leadNumbers);

/* Aliases under the pre-SZx names used by older callers
 * (example/testfloat_compress_fastmode2.cu:78). */
unsigned char *SZ_fast_compress_args(int fastMode, int dataType, void *data, size_t *outSize,
                                     int errBoundMode, float absErrBound, float relBoundRatio,
                                     float compressionRatio, float tolerance, size_t r5, size_t r4,
                                     size_t r3, size_t r2, size_t r1);
void *SZ_fast_decompress(int fastMode, int dataType, unsigned char *bytes, size_t byteLength,
                         size_t r5, size_t r4, size_t r3, size_t r2, size_t r1);

/* ---- 2. device-resident API ------------------------------------------------ */

/* Error text of the last failing call on this thread ("" if none). */
const char *szx_amd_last_error(void);

/* Largest stream the codec can emit for n values (header, bitmap and a full
 * record for every block). */
size_t szx_amd_compress_bound(size_t n, int dataType, int blockSize);

/* Capacity d_out must have for szx_amd_compress_dev: the stream is built in
 * place and its start floats inside the buffer (see below). */
size_t szx_amd_compress_capacity(size_t n, int dataType, int blockSize);

/* Workspace bytes for szx_amd_compress_dev. */
size_t szx_amd_compress_workspace(size_t n, int dataType, int blockSize);

/* uint64_t words of a record index for n values.  For blockSize a power of
 * two 4..256: per index tile (4096 fp32 / 2048 fp64 values) the payload
 * offset of the tile's first record and the number of constant blocks before
 * the tile, plus a closing pair; then per block a uint16_t -- the record's
 * start inside its tile's payload bytes, or a constant block's rank among its
 * tile's constant blocks.  For every other blockSize 1..255: per block the
 * payload offset of its record (of the next record, for a constant block)
 * and the constant blocks before it, plus a closing pair {payload bytes,
 * constant blocks}.  The index lives outside the stream (the stream stays
 * byte-identical); handing it to szx_amd_decompress_dev lets every block
 * decode at once without walking the records. */
size_t szx_amd_index_words(size_t n, int dataType, int blockSize);

/* Compress n device values.  errBoundMode ABS or REL (REL's bound is computed
 * on the device with the reference's float rounding).  blockSize 1..256: the
 * powers of two 4..256 run on the tiled kernels, every other size on per-block
 * kernels (same stream bytes, lower throughput).  On completion
 * d_result[0] = byte offset of the stream inside d_out, d_result[1] = stream
 * size, d_result[2] = the absolute error bound used (as float bits),
 * d_result[3] = 0, or 1 if the call failed (a bounded wait between
 * workgroups ran out; the stream is then invalid).
 * d_result is device (or host-pinned) memory of 4 uint64_t.  d_index, if not
 * NULL, receives the record index (szx_amd_index_words() words).
 * Returns 0, or a negative SZx_* code (argument errors only; the call is
 * asynchronous). */
int szx_amd_compress_dev(const void *d_in, size_t n, int dataType, int errBoundMode,
                         float absErrBound, float relBoundRatio, int blockSize, void *d_out,
                         size_t outCapacity, void *d_work, size_t workBytes, uint64_t *d_result,
                         uint64_t *d_index, void *hipStream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc):
 * szx_amd_compress_dev without a workspace argument.  The workspace lives in
 * a static device array of SZX_AMD_NOALLOC_SCRATCH_BYTES that each GPU holds
 * until the process ends; the call allocates nothing.  A compress whose
 * szx_amd_compress_workspace() exceeds the array returns SZx_MERR (16 MiB,
 * about 6 bytes per block: 2^27 values at blockSize 64..256, 2^26 at
 * 32..256, 2^24 at 8..256).  Calls on one GPU share its array: calls that
 * overlap in time on one GPU race (one at a time per GPU).  Other arguments
 * and results as szx_amd_compress_dev. */
#define SZX_AMD_NOALLOC_SCRATCH_BYTES (16u << 20)
int szx_amd_compress_dev_noalloc(const void *d_in, size_t n, int dataType, int errBoundMode,
                                 float absErrBound, float relBoundRatio, int blockSize, void *d_out,
                                 size_t outCapacity, uint64_t *d_result, uint64_t *d_index, void *hipStream);

/* Workspace bytes for szx_amd_decompress_dev (0 for an unsupported blockSize). */
size_t szx_amd_decompress_workspace(size_t n, int dataType, int blockSize, size_t streamBytes);

/* Decode flags. */
#define SZX_AMD_DECODE_REFERENCE 0   /* bit-identical to the reference decoder */
#define SZX_AMD_DECODE_EXACT 1       /* fp64: read the bytes the encoder wrote
                                        where the reference decoder misreads
                                        one (reqBytes 5, lead 0, byte 7) */

/* Decompress a device-resident stream of streamBytes bytes into n values at
 * d_out.  blockSize must match the stream's (checked on the device).  d_index
 * is the record index written by szx_amd_compress_dev, or NULL to rebuild it
 * from the stream.  d_status (device memory, one int32) receives 0 or a
 * negative code for a malformed stream or index. */
int szx_amd_decompress_dev(const void *d_stream, size_t streamBytes, size_t n, int dataType,
                           int blockSize, void *d_out, void *d_work, size_t workBytes,
                           const uint64_t *d_index, int flags, int32_t *d_status,
                           void *hipStream);

/* szx.cc:254-268 (the fastMode 3 / 4 branches of SZx_fast_compress_args) on
 * device memory: the random-access container
 * (SZx_fast_compress_args_unpredictable_blocked_randomaccess_float[_openmp],
 * szx_float.cc:1407-1489 / 1218-1405, and the double versions).  fastMode 3
 * or 4; blockSize a power of two 4..256; buffers, workspace, index and
 * results as szx_amd_compress_dev. */
int szx_amd_compress_ra_dev(int fastMode, const void *d_in, size_t n, int dataType, int errBoundMode,
                            float absErrBound, float relBoundRatio, int blockSize, void *d_out,
                            size_t outCapacity, void *d_work, size_t workBytes, uint64_t *d_result,
                            uint64_t *d_index, void *hipStream);

/* szx.cc:391-403 (SZx_fast_decompress, fastMode 3 / 4) on device memory:
 * decode a random-access container (szxd_float.cc:313-476,
 * szxd_double.cc:518-681).  Without d_index the record index comes from the
 * container's size table.  Arguments as szx_amd_decompress_dev. */
int szx_amd_decompress_ra_dev(int fastMode, const void *d_stream, size_t streamBytes, size_t n,
                              int dataType, int blockSize, void *d_out, void *d_work, size_t workBytes,
                              const uint64_t *d_index, int flags, int32_t *d_status, void *hipStream);

/* Non-blocked codec (fastMode 1, szx_float.cc:1492-1699 /
 * szx_double.cc:1593-2010 and szxd_float.cc:478-662 / szxd_double.cc:683-1000:
 * one record over all n values against the global median) on device memory.
 * Output capacity and workspace (shared by both directions) in bytes, 0 for
 * unsupported arguments. */
size_t szx_amd_nb_capacity(size_t n, int dataType);
size_t szx_amd_nb_workspace(size_t n, int dataType);

/* szx.cc:270-287 (the non-blocked branch of SZx_fast_compress_args) on the
 * device: ABS or REL.  The stream starts at d_out; d_result receives
 * {0, stream bytes, bound bits, 0}.  Asynchronous on hipStream. */
int szx_amd_compress_nb_dev(const void *d_in, size_t n, int dataType, int errBoundMode, float absErrBound,
                            float relBoundRatio, void *d_out, size_t outCapacity, void *d_work, size_t workBytes,
                            uint64_t *d_result, void *hipStream);

/* szx.cc:393-394 (SZx_fast_decompress, fastMode 1) on the device: decode a
 * non-blocked stream of streamBytes bytes into n values; flags and d_status as
 * for szx_amd_decompress_dev. */
int szx_amd_decompress_nb_dev(const void *d_stream, size_t streamBytes, size_t n, int dataType, void *d_out,
                              void *d_work, size_t workBytes, int flags, int32_t *d_status, void *hipStream);

/* Reader of the stream the reference's CUDA codec (cuSZx) writes, fp32.
 * cuszx_entry.h:11 / cuszx_entry.cu:129-209: decode nbEle values from
 * cmpBytes into a malloc'd array (*newData; free with free()).  The block size
 * comes from the stream and must be a multiple of 32 in 32..224 (the sizes the
 * reference's kernel decodes whole).  The reference takes no stream length;
 * it is found, as the reference's host loop finds it, from the header and the
 * per-record counts.  On error *newData is NULL and szx_amd_last_error() says
 * why; where the reference prints a warning and exits (a record longer than
 * 4 blockSize bytes) this returns that error. */
void cuSZx_fast_decompress_args_unpredictable_blocked_float(float **newData, size_t nbEle,
                                                            unsigned char *cmpBytes);

/* Writer of the same stream, fp32 (cuszx_entry.h:10 / cuszx_entry.cu:72-127
 * with the kernel cuszx_float.cu:222-298): compress nbEle values with the
 * absolute bound absErrBound and blockSize values a block (a multiple of 32 in
 * 32..224; the reference's block reduction covers the whole block for 32, 64
 * and 128, and for the other sizes this takes the whole block's range).
 * Returns a malloc'd stream of *outSize bytes (free with free()), or NULL
 * with szx_amd_last_error() set, also when a block would need more than 32
 * significant bits (the reference then writes mid bytes it never filled). */
unsigned char *cuSZx_fast_compress_args_unpredictable_blocked_float(float *oriData, size_t *outSize,
                                                                     float absErrBound, size_t nbEle,
                                                                     int blockSize);

/* The writer on device memory: d_out holds szx_amd_cuszx_capacity() bytes and
 * receives the stream at offset 0; d_work holds
 * szx_amd_cuszx_compress_workspace() bytes; d_result (4 x uint64) receives
 * {0, stream bytes, bound bits, 1 if a block needs more than 32 significant
 * bits (no valid stream) else 0}.  Asynchronous on hipStream. */
size_t szx_amd_cuszx_capacity(size_t nbEle, int blockSize);
size_t szx_amd_cuszx_compress_workspace(size_t nbEle, int blockSize);
int szx_amd_compress_cuszx_dev(const float *d_in, size_t nbEle, float absErrBound, int blockSize, void *d_out,
                               size_t outCapacity, void *d_work, size_t workBytes, uint64_t *d_result,
                               void *hipStream);

/* The same decode on device memory: d_stream holds streamBytes bytes,
 * d_out nbEle floats; d_work holds szx_amd_cuszx_workspace(nbEle) bytes
 * (256-byte aligned).  *d_status is 0 after the stream completes, or a
 * negative code for a malformed stream (-41 header, -42 record length, -43
 * bitmap, -44 truncated, -45 a mid-byte position outside its block).
 * Asynchronous on hipStream; allocates nothing. */
size_t szx_amd_cuszx_workspace(size_t nbEle);
int szx_amd_decompress_cuszx_dev(const void *d_stream, size_t streamBytes, size_t nbEle, float *d_out,
                                 void *d_work, size_t workBytes, int32_t *d_status, void *hipStream);

/* Workspace bytes for szx_amd_fxr_bound_dev (0 for unsupported arguments). */
size_t szx_amd_fxr_workspace(size_t n, int dataType, int blockSize);

/* The reference's fixed-rate bound choice (szx_float.cc:921-949 without the
 * final compress; fp64 szx_double.cc:917-943) for n device values: a 1-in-10
 * block sample, model sums on the GPU, the bracket-then-bisect search on the
 * host.  initErrorBound <= 0 starts the search at 1% of the sampled value
 * range, as the reference's fixed-rate compressor does.  *errorBound receives
 * the bound (bit-identical to the reference's), *modelPasses (if not NULL)
 * the number of model evaluations the search made.  Synchronous: it waits
 * on `stream` once per model pass.  Returns 0 or a negative SZx_* code. */
int szx_amd_fxr_bound_dev(const void *d_in, size_t n, int dataType, int blockSize,
                          float targetCompressionRatio, float tolerance, float initErrorBound,
                          void *d_work, size_t workBytes, float *errorBound, int *modelPasses,
                          void *hipStream);

/* ---- 3. fixed rate on the exact codec ------------------------------------
 * No counterpart in the reference: its FXR mode (above) hits a *modelled*
 * ratio from a sample; these entry points hit a target rate of the real
 * stream, 8 * stream bytes / n bits per value, within a relative tolerance,
 * and leave that stream in d_out.  Every candidate is judged by the exact
 * codec; reduced precision is used only inside the rate model. */

#define SZX_AMD_RATE_BISECT 0    /* bisection on log2(eb), one compress per step */
#define SZX_AMD_RATE_LEARNED 1   /* rate-model proposal, exact octave pass,
                                    final compress; bisection if it misses */

typedef struct {
    float errorBound;      /* bound of the stream left in d_out */
    double bitsPerValue;   /* its exact rate */
    int codecPasses;       /* full compresses + octave (exact size) passes */
    int featurePasses;     /* statistics passes (learned method) */
    int hit;               /* 1 if |rate - target| <= tolerance * target */
} szx_amd_rate_info;

size_t szx_amd_fixed_rate_workspace(size_t n, int dataType, int blockSize);

/* Fixed-rate compress of n device values (blockSize: powers of two 4..256).
 * d_out / outCapacity / d_result / d_index as for szx_amd_compress_dev;
 * d_modelParams: SZX_AMD_RATE_MODEL_PARAMS floats of the rate model (device
 * memory; the learned method only).  Synchronous (waits on `stream` after
 * each pass).  Returns 0 or a negative SZx_* code. */
int szx_amd_fixed_rate_dev(const void *d_in, size_t n, int dataType, int blockSize,
                           float targetBitsPerValue, float tolerance, int method,
                           const float *d_modelParams, void *d_out, size_t outCapacity,
                           void *d_work, size_t workBytes, uint64_t *d_result, uint64_t *d_index,
                           szx_amd_rate_info *info, void *hipStream);

/* Exact stream sizes for a whole octave of bounds [2^reqExpo, 2^(reqExpo+1)):
 * d_acc receives 1 + 2 * SZX_AMD_OCTAVE_BINS u64 -- bytes fixed over the
 * octave (always-constant blocks' medians and always-present records), then
 * per radius bin k (radius in (2^e (1 + k/B), 2^e (1 + (k+1)/B)]) the block
 * count and their record bytes.  The stream at eb_k = 2^e (1 + k/B) has
 * 13 + ceil(nb/8) + fixed + 4 * (blocks in bins < k) + (records of bins >= k)
 * bytes.  reqExpo in [-1074, 1023]. */
#define SZX_AMD_OCTAVE_BINS 512
size_t szx_amd_rate_octave_workspace(void);
int szx_amd_rate_octave_dev(const void *d_in, size_t n, int dataType, int blockSize, int reqExpo,
                            uint64_t *d_acc, void *d_work, size_t workBytes, void *hipStream);

/* The same sizes over nOctaves (1 or 2) adjacent octaves in one pass, with
 * the bounds just outside them: octave reqExpo - j (j < nOctaves) fills
 * d_acc[j * (1 + 2B) ...] as above; then 4 u64 follow -- constant blocks
 * and record bytes at 2^(reqExpo+1), and constant blocks and record bytes
 * at the float bound just below 2^(reqExpo-nOctaves+1).  The stream at such
 * an edge bound has 13 + ceil(nb/8) + 4 * constant + record bytes.
 * reqExpo in [nOctaves - 149, 126] (float bounds on both sides).  The
 * fixed-rate solver's learned search sizes a pair when the model's bound
 * lies close to an octave edge.  No counterpart in the reference. */
int szx_amd_rate_octave_span_dev(const void *d_in, size_t n, int dataType, int blockSize, int reqExpo,
                                 int nOctaves, uint64_t *d_acc, void *d_work, size_t workBytes,
                                 void *hipStream);

/* The rate model: an MLP 24 -> 64 -> 64 -> 1 (GELU) predicting a chunk's
 * log2(bits per value) at a bound from the chunk's statistics.  Parameters
 * are one FP32 array: W1[64][24], b1[64], W2[64][64] (out, in), b2[64],
 * w3[64], b3.  The kernels run the hidden layers on BF16 matrix cores with
 * FP32 accumulation. */
#define SZX_AMD_RATE_MODEL_PARAMS 5825
#define SZX_AMD_RATE_MODEL_INPUTS 24
#define SZX_AMD_RATE_CHUNK 524288        /* values per statistics chunk */

/* Per-chunk statistics (ceil(n / SZX_AMD_RATE_CHUNK) chunks). */
size_t szx_amd_rate_stats_bytes(size_t n);
size_t szx_amd_rate_stats_workspace(size_t n);
int szx_amd_rate_stats_dev(const void *d_in, size_t n, int dataType, int blockSize, void *d_stats,
                           void *d_work, size_t workBytes, void *hipStream);
/* Model inputs for every (bound k, chunk c): d_X[(k * nchunks + c) * 24 ...]. */
int szx_amd_rate_inputs_dev(const void *d_stats, size_t n, const float *d_ebs, int K, float *d_X,
                            void *hipStream);
/* d_y[i] = model(d_X[i]) for nSamples rows. */
int szx_amd_mlp_forward_dev(const float *d_params, const float *d_X, size_t nSamples, float *d_y,
                            void *hipStream);
/* Gradient of mean((model(X) - T)^2) into d_grad[SZX_AMD_RATE_MODEL_PARAMS]
 * and the loss into *d_loss. */
size_t szx_amd_mlp_grad_workspace(void);
int szx_amd_mlp_grad_dev(const float *d_params, const float *d_X, const float *d_T, size_t nSamples,
                         float *d_grad, float *d_loss, void *d_work, size_t workBytes,
                         void *hipStream);
/* One bias-corrected Adam step (step >= 1) on the FP32 parameters. */
int szx_amd_mlp_adam_dev(float *d_params, const float *d_grad, float *d_m, float *d_v, float lr,
                         float beta1, float beta2, float eps, int step, void *hipStream);

/* ---- Batched launch -------------------------------------------------------
 * No counterpart in the reference, which compresses one array per call
 * (SZx_fast_compress_args, szx.cc:223-262).  Many fields of uneven sizes,
 * fp32 and fp64 mixed, each with its own ABS bound or REL ratio, go through
 * one launch sequence per direction (a few kernels covering every field).
 * Each field's stream is byte-identical
 * to the stream SZx_fast_compress_args writes for that field alone, and is
 * placed inside the field's d_out as szx_amd_compress_dev places it. */
typedef struct szx_amd_field {
    const void *d_in;       /* n values, 16-byte aligned */
    size_t n;               /* >= 1 */
    int dataType;           /* SZx_FLOAT or SZx_DOUBLE */
    int errBoundMode;       /* ABS or REL */
    float absErrBound;      /* ABS */
    float relBoundRatio;    /* REL */
    void *d_out;            /* szx_amd_compress_capacity(n, dataType, blockSize) bytes */
    void *d_dec;            /* n values (16-byte aligned) for szx_amd_batch_decompress, or
                               NULL (for every field) when the batch is never decompressed */
} szx_amd_field;

typedef struct szx_amd_batch szx_amd_batch;   /* host-side plan */

#define SZX_AMD_BATCH_FUSED 0        /* every field in one launch sequence */
#define SZX_AMD_BATCH_PER_FIELD 1    /* the single-field launches, field by field
                                        (for comparison) */

/* Device workspace bytes of a batch (0 for invalid fields). */
size_t szx_amd_batch_workspace(const szx_amd_field *fields, int nfields, int blockSize);

/* Lay out d_work (256-byte aligned, szx_amd_batch_workspace bytes) and upload
 * the field tables; synchronizes `stream`.  The plan keeps the pointers: the
 * fields' buffers and d_work must outlive it.  NULL on error. */
szx_amd_batch *szx_amd_batch_create(const szx_amd_field *fields, int nfields, int blockSize,
                                    void *d_work, size_t workBytes, void *hipStream);
void szx_amd_batch_destroy(szx_amd_batch *batch);

/* Compress every field.  FUSED only enqueues kernels: nothing is
 * allocated, copied or waited for, so the call can be captured in a hipGraph
 * and replayed.  Result words of field i land at szx_amd_batch_results_dev()
 * + 4 i, with the meaning of szx_amd_compress_dev's d_result. */
int szx_amd_batch_compress(szx_amd_batch *batch, int mode, void *hipStream);

/* Device arrays of the plan: nfields x 4 result words, nfields decode
 * statuses (int32, as szx_amd_decompress_dev's d_status), and field i's
 * record index (szx_amd_index_words words). */
uint64_t *szx_amd_batch_results_dev(szx_amd_batch *batch);
int32_t *szx_amd_batch_status_dev(szx_amd_batch *batch);
uint64_t *szx_amd_batch_index_dev(szx_amd_batch *batch, int field);

/* Copy the result words to h_results (nfields x 4, or NULL) and keep them in
 * the plan for PER_FIELD decompression; synchronizes `stream`. */
int szx_amd_batch_results(szx_amd_batch *batch, uint64_t *h_results, void *hipStream);

/* Decompress every field's stream (found through its result words) into its
 * d_dec.  flags as szx_amd_decompress_dev.  FUSED only enqueues kernels;
 * PER_FIELD needs szx_amd_batch_results() after the compress. */
int szx_amd_batch_decompress(szx_amd_batch *batch, int flags, int mode, void *hipStream);

/* Tile shape of the fused compress: values per lane (4, 8 or 16; 8 needs
 * blockSize >= 8; 16 applies to the fp32 fields, the fp64 fields taking 8) and
 * index tiles per workgroup tile (1, 2 or 4, i.e. 4096, 8192 or 16384 fp32
 * values; 4 needs 8 or 16 values per lane and blockSize 16..128, and 16 values
 * per lane needs 4).
 * szx_amd_batch_tune times every shape on this batch
 * (reps timed passes each, after one untimed pass; synchronizes), keeps the
 * fastest and reports it with its mean milliseconds per compress.  The choice
 * is stored per batch class (GPU arch and CU count, blockSize, fp32 share of
 * the bytes, log2 of the bytes and of the field count) in a JSON file,
 * szx_amd_tune.json beside libszx_amd.so or the path in SZX_AMD_TUNE_CACHE
 * ("off" disables it): for a class already there, szx_amd_batch_tune and
 * szx_amd_batch_create take the stored shape without timing (and the tuner
 * reports the stored milliseconds).  If timing fails, the plan keeps the shape
 * it had.  The environment variable SZX_AMD_BATCH_SHAPE="lanes,tiles" sets the
 * initial shape of new plans (before the file). */
int szx_amd_batch_set_shape(szx_amd_batch *batch, int valuesPerLane, int tilesPerGroup);
int szx_amd_batch_tune(szx_amd_batch *batch, int reps, int *valuesPerLane, int *tilesPerGroup,
                       float *ms, void *hipStream);

/* Kernel timing for benchmarks: after szx_amd_timing_begin(), this thread's
 * launches bracket their main kernels with HIP events on the launch stream;
 * szx_amd_timing_end() waits for them and returns, per kernel kind
 * (0 = compress, 1 = decode, 2 = record-index build, 3 = szx_amd_copy_dev),
 * the summed milliseconds and the number of launches.  Arrays hold 4
 * entries. */
void szx_amd_timing_begin(void);
int szx_amd_timing_end(double *ms_sum, int *count);

/* Device copy of `bytes` bytes (pointers and size multiples of 16), one
 * non-temporal 16-byte load and store per lane: the copy rate that bounds
 * the codec kernels on this GPU.  No counterpart in the reference: a
 * measurement helper for benchmarks. */
int szx_amd_copy_dev(const void *d_src, void *d_dst, size_t bytes, void *hipStream);

#ifdef __cplusplus
}
#endif
#endif
