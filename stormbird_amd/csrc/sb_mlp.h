// Warm-start network (sb_mlp.hip) as seen by the lifting-line step.
#pragma once

#include <hip/hip_runtime.h>

struct sb_mlp;

namespace sb {

constexpr int MLP_FEATURES = 64;  // per-strip input width (58 used, zero padded)

// y[n] = network(X[n][MLP_FEATURES]) on stream st (device pointers): one
// kernel, no allocation (it may be captured into a graph).
hipError_t mlp_forward_on(sb_mlp* m, const float* X, long n, float* y, hipStream_t st);

}  // namespace sb
