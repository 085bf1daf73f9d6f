// Allocation-free form of the standalone Biot-Savart sum, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  The panel records it stages live in one file-scope device array,
// `scratch`, of NOALLOC_SCRATCH_BYTES per device; every call on a device reuses that
// device's array, so calls that overlap on one GPU (different streams, threads) race.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>
#include <string>

#include "sb_device.h"
#include "sb_host.h"

using namespace sb;

extern "C" long sb_noalloc_max_panels(void);

namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;  // 32 MiB
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device.
// (In host code `scratch` names a host-side stand-in, not device memory.)
unsigned char* scratch_on_current_device() {
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[dev];
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Bytes of scratch the staged records of n panels take: far-field records, corners, core lengths.
size_t staged_bytes(size_t n) {
    return align256(sizeof(PanelRec) * n) + align256(sizeof(double) * 12 * n) + align256(sizeof(double) * n);
}
}  // namespace

// Records [n][22] (4 corners, centre, normal, far-field length^2, viscous core length,
// area / (4 pi), strength) into the layout the pair kernels read.
__global__ __launch_bounds__(256) void k_stage_records(const double* __restrict__ recs, long n,
                                                       PanelRec* __restrict__ rec, double* __restrict__ corners,
                                                       double* __restrict__ core) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double* q = recs + 22 * k;
    PanelRec r;
    r.cx = q[12];
    r.cy = q[13];
    r.cz = q[14];
    r.nx = q[15];
    r.ny = q[16];
    r.nz = q[17];
    r.far2 = q[18];
    r.area_term = q[20];
    r.strength = q[21];
    r.cbase = -1.0;
    rec[k] = r;
    for (int c = 0; c < 12; ++c) corners[12 * k + c] = q[c];
    core[k] = q[19];
}

__global__ __launch_bounds__(256) void k_zero_fill(double* __restrict__ out, long n) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = 0.0;
}

extern "C" int sb_wake_induced_velocities_noalloc(const double* panel_records_device, size_t n_panels,
                                                  const double* targets_xyz_device, size_t n_targets, int symmetry,
                                                  int mode, double* out_xyz_device, void* stream) {
    if ((!panel_records_device && n_panels) || (!targets_xyz_device && n_targets) || (!out_xyz_device && n_targets))
        return -1;
    if (symmetry < 0 || symmetry > 3 || (mode != MODE_EXACT && mode != MODE_FAST)) return -4;
    if (n_targets > (size_t)INT_MAX || n_panels > (size_t)INT_MAX) {
        g_error = "sb_wake_induced_velocities_noalloc: more than 2^31 - 1 panels or targets";
        return -4;
    }
    if (staged_bytes(n_panels) > NOALLOC_SCRATCH_BYTES) {
        g_error = "sb_wake_induced_velocities_noalloc: " + std::to_string(n_panels) + " panels need " +
                  std::to_string(staged_bytes(n_panels)) + " bytes of scratch; the limit is " +
                  std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes (" + std::to_string(sb_noalloc_max_panels()) +
                  " panels)";
        return -4;
    }
    if (n_targets == 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    return guarded([&] {
        unsigned char* base = scratch_on_current_device();
        if (n_panels == 0) {  // a fill kernel, not a memset: callers may capture this call into a graph
            const long n = 3L * (long)n_targets;
            hipLaunchKernelGGL(k_zero_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out_xyz_device, n);
            check(hipGetLastError(), "k_zero_fill");
            return;
        }
        PanelRec* d_rec = reinterpret_cast<PanelRec*>(base);
        double* d_corners = reinterpret_cast<double*>(base + align256(sizeof(PanelRec) * n_panels));
        double* d_core = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(d_corners) +
                                                   align256(sizeof(double) * 12 * n_panels));
        hipLaunchKernelGGL(k_stage_records, dim3((unsigned)((n_panels + 255) / 256)), dim3(256), 0, st,
                           panel_records_device, (long)n_panels, d_rec, d_corners, d_core);
        check(hipGetLastError(), "k_stage_records");
        Inst d{};
        d.N = (int)n_panels;
        d.M = 1;
        d.P = d.N + 1;
        d.symmetry = symmetry;
        d.rec = d_rec;
        d.corners = d_corners;
        d.core = d_core;
        check(launch_induced_targets(d, targets_xyz_device, (long)n_targets, (long)n_panels, mode, out_xyz_device,
                                     nullptr, st),
              "induced");
    });
}

extern "C" long sb_noalloc_max_panels(void) {
    size_t lo = 0, hi = NOALLOC_SCRATCH_BYTES / sizeof(PanelRec) + 1;
    while (lo + 1 < hi) {  // largest n with staged_bytes(n) <= the scratch size
        size_t mid = (lo + hi) / 2;
        if (staged_bytes(mid) <= NOALLOC_SCRATCH_BYTES) lo = mid;
        else hi = mid;
    }
    return (long)lo;
}
