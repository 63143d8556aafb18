// szx_fxr.hpp -- host interface of the fixed-rate kernels (szx_fxr.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace szxamd {

constexpr int FXR_MAX_CAND = 32;    // candidate bounds per model pass
constexpr int FXR_MAX_WG = 1024;    // workgroups of a model pass
constexpr int OCT_BINS = 512;       // radius bins per octave of bounds
constexpr int OCT_WG = 2048;        // workgroups of an octave pass
constexpr int OCT_ACC = 1 + 2 * OCT_BINS;   // u64 accumulators of an octave pass (+4 with edges)

// Workspace of the sampled model: per sampled block median and radius,
// per-workgroup partial sums, the candidate bounds, the folded sums and the
// sampled value range.
struct FxrLayout {
    uint64_t ns;   // sampled blocks
    uint64_t ws_med, ws_rad, ws_part, ws_ebs, ws_sums, ws_range, ws_bytes;
};
FxrLayout fxr_layout(uint64_t n, int w, int bs, int rate);

// Sample the blocks and their range (float at ws + ws_range).
void fxr_sample_launch(const void *x, uint64_t n, int w, int bs, int rate, uint8_t *ws, hipStream_t st);
// Model sums for the K bounds at ws + ws_ebs -> ws + ws_sums (3 u64 each:
// constant blocks, sum reqBytes, sum min(lead, reqBytes)).
void fxr_estimate_launch(const void *x, uint64_t n, int w, int bs, int rate, uint8_t *ws, int K,
                         hipStream_t st);

// Exact rate over the octave of bounds with exponent reqexp: d_acc receives
// OCT_ACC u64 (fixed bytes, then per radius bin count and record bytes), and
// with `edges` four more: constant blocks and record bytes at 2^(reqexp+1)
// and at the float bound below 2^reqexp; ws holds octave_workspace_bytes()
// (unused scratch, kept for the interface).  nocts = 2 sizes the octaves
// reqexp and reqexp - 1 in the same pass (2 * OCT_ACC u64, reqexp's first;
// always with the edges, the lower one then below 2^(reqexp-1)).
uint64_t octave_workspace_bytes();
void rate_octave_launch(const void *x, uint64_t n, int w, int bs, int reqexp, uint8_t *ws,
                        unsigned long long *d_acc, hipStream_t st, bool edges = false, int nocts = 1);

// Per-chunk statistics feeding the rate model (RATE_CHUNK values per chunk;
// the last chunk may be shorter).  d: min, max, sum x, sum x^2, sum |d1|,
// sum d1^2, sum |d2| (differences inside codec blocks); hist: codec blocks
// per float radius exponent field (0..255).
constexpr uint64_t RATE_CHUNK = 1ull << 19;
constexpr int RATE_WG_PER_CHUNK = 32;
constexpr int RATE_HIST = 256;
constexpr int RATE_INPUTS = 24;
struct RateStats {
    double d[7];
    uint64_t values;
    unsigned blocks, zero;
    unsigned hist[RATE_HIST];
};

uint64_t rate_stats_workspace_bytes(uint64_t n);
void rate_stats_launch(const void *x, uint64_t n, int w, int bs, uint8_t *ws, RateStats *d_stats,
                       hipStream_t st);
// d_ebs[k] = field range * 2^(-k/4), k < K (range from the statistics).
void rate_ladder_launch(const RateStats *d_stats, int nchunks, int K, float *d_ebs, hipStream_t st);
// X[(k * nchunks + c)][RATE_INPUTS] for K bounds at d_ebs.
void rate_inputs_launch(const RateStats *d_stats, int nchunks, const float *d_ebs, int K, float *d_X,
                        hipStream_t st);

}  // namespace szxamd
