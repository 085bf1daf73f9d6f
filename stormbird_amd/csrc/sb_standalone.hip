// Kernel-level entry points that work on caller-given arrays rather than on a
// simulation handle: the frozen-wake matrix-vector product.
#include <hip/hip_runtime.h>

#include <string>

#include "sb_device.h"
#include "sb_host.h"

using namespace sb;

// FrozenWake::update_induced_velocities_at_control_points (wake/frozen_wake.rs:268-301):
// u_r = fixed_r + (sum_c F[r][c] * gamma_c), with the x, y and z sums each started at 0 and
// taken over the columns in ascending order.  One lane per row walks its row of F, which is
// contiguous in the row-major [r][c][3] layout, so each lane streams its own cache lines.
__global__ __launch_bounds__(256) void k_frozen_matvec(const double* __restrict__ fixed,
                                                       const double* __restrict__ F,
                                                       const double* __restrict__ gamma, int n,
                                                       double* __restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double* row = F + 3L * n * r;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int c = 0; c < n; ++c) {
        const double g = gamma[c];
        sx += row[3 * c] * g;
        sy += row[3 * c + 1] * g;
        sz += row[3 * c + 2] * g;
    }
    out[3 * r] = fixed[3 * r] + sx;
    out[3 * r + 1] = fixed[3 * r + 1] + sy;
    out[3 * r + 2] = fixed[3 * r + 2] + sz;
}

extern "C" int sb_frozen_matvec(const double* fixed, const double* factors, const double* strengths, size_t n,
                                double* out) {
    if (n == 0) return 0;
    if (!fixed || !factors || !strengths || !out) return -1;
    if (n > 65536) {
        g_error = "sb_frozen_matvec: at most 65536 rows";
        return -4;
    }
    double* d = nullptr;
    hipStream_t st = nullptr;
    int rc = 0;
    try {
        const size_t nf = 3 * n * n;
        check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
        d = dalloc<double>(nf + 3 * n + n + 3 * n);
        double* d_F = d;
        double* d_fixed = d + nf;
        double* d_g = d_fixed + 3 * n;
        double* d_out = d_g + n;
        check(hipMemcpyAsync(d_F, factors, sizeof(double) * nf, hipMemcpyHostToDevice, st), "upload");
        check(hipMemcpyAsync(d_fixed, fixed, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st), "upload");
        check(hipMemcpyAsync(d_g, strengths, sizeof(double) * n, hipMemcpyHostToDevice, st), "upload");
        hipLaunchKernelGGL(k_frozen_matvec, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_fixed, d_F, d_g,
                           (int)n, d_out);
        check(hipGetLastError(), "k_frozen_matvec");
        check(hipMemcpyAsync(out, d_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st), "download");
        check(hipStreamSynchronize(st), "sync");
    } catch (const std::exception& e) {
        g_error = e.what();
        rc = -5;
    }
    if (d) (void)hipFree(d);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}
