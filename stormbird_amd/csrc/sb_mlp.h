// Warm-start network (sb_mlp.hip) as seen by the lifting-line step.
#pragma once

#include <hip/hip_runtime.h>

struct sb_mlp;

namespace sb {

constexpr int MLP_FEATURES = 64;  // per-strip input width (45 used, zero padded)

// y[n] = network(X[n][MLP_FEATURES]) on stream st (device pointers); allocates
// nothing when mlp_reserve(m, >= n) was called before.
hipError_t mlp_forward_on(sb_mlp* m, const float* X, long n, float* y, hipStream_t st);
void mlp_reserve(sb_mlp* m, long n);

}  // namespace sb
