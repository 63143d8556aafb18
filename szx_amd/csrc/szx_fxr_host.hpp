// szx_fxr_host.hpp -- host half of the reference's fixed-rate bound search.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <functional>

namespace szxamd {

// Sums of the sampled model at one bound (as computed by k_fxr_estimate).
struct FxrSums {
    uint64_t nconst, sum_req, sum_lead;
};

// Evaluates K bounds; false on a device error.
using FxrEvaluator = std::function<bool(const float *ebs, int K, FxrSums *out)>;

// Model compression ratio of the sample at one bound.
float fxr_model_cr(int w, int bs, size_t n, size_t ns, const FxrSums &s);

// The reference's bracket-then-bisect search from init_eb; *n_est counts the
// model evaluations the reference performs.  False on a device error.
bool fxr_search(float target, float tol, float init_eb, int w, int bs, size_t n, size_t ns,
                const FxrEvaluator &eval, float *eb_out, int *n_est);

}  // namespace szxamd
