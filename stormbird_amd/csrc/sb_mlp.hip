// Learned warm start for the circulation solve: a per-strip MLP
// (64 -> 256 -> 256 -> 256 -> 256 -> 1, GELU) whose training and inference
// GEMMs run on the gfx950 matrix cores in bf16 with fp32 accumulation and fp32
// master weights (Adam).  No counterpart in the reference: the reference's
// solver starts from the previous circulation (or zero at the first step,
// lifting_line/simulation.rs:91); the network only supplies that initial
// guess, and the unchanged fixed-point solve finishes from it.
//
// GEMM convention: C[M][N] = A[M][K] * Bt[N][K]^T, both operands bf16 with K
// contiguous (K a multiple of 16), fp32 accumulation in
// v_mfma_f32_16x16x16_bf16.  Fragment layout of that instruction (wave64):
// lane l holds A[l%16][4(l/16)+v] and Bt[l%16][4(l/16)+v] for v = 0..3, and
// returns D[4(l/16)+v][l%16].
//
// Parameter layout (fp32, flat): for layer l = 0..4, W_l[out][in] then b_l[out],
// sizes 256x64, 256x256 (x3), 1x256.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "sb_host.h"
#include "sb_mlp.h"

namespace sb {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __host__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
    float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

struct GemmArgs {
    const uint16_t* A;   // [M][K]
    const uint16_t* Bt;  // [N][K]
    int M, N, K;
    int k_split;         // K range per z-slice (multiple of 16)
    float* C;            // fp32 out [M][N] (z-slice s writes C + s*M*N when split)
    const float* bias;   // [N] or null
    uint16_t* Hb;        // bf16 of gelu(C + bias) [M][N], or null
};

// 64x64 tile per workgroup (2x2 waves of 32x32, each 2x2 MFMA 16x16 tiles).
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int m0 = blockIdx.y * 64 + (wave / 2) * 32;
    const int n0 = blockIdx.x * 64 + (wave % 2) * 32;
    const int kb = blockIdx.z * g.k_split;
    const int ke = kb + g.k_split < g.K ? kb + g.k_split : g.K;
    const int r = lane % 16, kq = 4 * (lane / 16);
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x4 zero = {0, 0, 0, 0};
    for (int k = kb; k < ke; k += 16) {
        bf16x4 a[2], b[2];
        for (int i = 0; i < 2; ++i) {
            int row = m0 + 16 * i + r;
            a[i] = row < g.M ? *reinterpret_cast<const bf16x4*>(g.A + (long)row * g.K + k + kq) : zero;
            int col = n0 + 16 * i + r;
            b[i] = col < g.N ? *reinterpret_cast<const bf16x4*>(g.Bt + (long)col * g.K + k + kq) : zero;
        }
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    float* C = g.C + (long)blockIdx.z * g.M * g.N;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int v = 0; v < 4; ++v) {
                int row = m0 + 16 * i + kq + v;
                int col = n0 + 16 * j + r;
                if (row >= g.M || col >= g.N) continue;
                float z = acc[i][j][v];
                if (g.bias) z += g.bias[col];
                C[(long)row * g.N + col] = z;
                if (g.Hb) g.Hb[(long)row * g.N + col] = f2bf(gelu(z));
            }
}

// out[i] = sum over slices s of part[s][i] (ascending s), i < n
__global__ void k_sum_slices(const float* part, int slices, long n, float* out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int q = 0; q < slices; ++q) s += part[(long)q * n + i];
    out[i] = s;
}

__global__ void k_to_bf16(const float* x, long n, uint16_t* y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// bf16 [rows][cols] -> [cols][ld] (rows <= ld; columns rows..ld-1 zero)
__global__ void k_transpose_bf16(const uint16_t* x, int rows, int cols, int ld, uint16_t* y) {
    __shared__ uint16_t t[32][33];
    int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 256 threads: 32 x 8
    for (int k = ty; k < 32; k += 8) {
        int rr = r0 + k, cc = c0 + tx;
        t[k][tx] = (rr < rows && cc < cols) ? x[(long)rr * cols + cc] : (uint16_t)0;
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        int cc = c0 + k, rr = r0 + tx;
        if (cc < cols && rr < ld) y[(long)cc * ld + rr] = t[tx][k];
    }
}

// Output layer y = Hb3 . w4 + b4 and its loss gradient dy = 2 (y - t) / n.
__global__ void k_head(const uint16_t* H, const float* w, const float* b, long n, const float* target, float* y,
                       float* dy, float* sq) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint16_t* h = H + i * 256;
    float s = 0.f;
    for (int k = 0; k < 256; ++k) s += bf2f(h[k]) * bf2f(f2bf(w[k]));
    s += b[0];
    y[i] = s;
    if (target) {
        float e = s - target[i];
        dy[i] = 2.0f * e / (float)n;
        sq[i] = e * e;
    }
}

// dZ3 = (dy * w4) * gelu'(Z3), stored bf16; also the head gradient partials.
__global__ void k_head_back(const float* dy, const float* w, const float* Z, long n, uint16_t* dZb) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n * 256) return;
    long i = q / 256;
    int k = (int)(q % 256);
    float g = dy[i] * bf2f(f2bf(w[k]));
    dZb[q] = f2bf(g * gelu_grad(Z[q]));
}

// dZ = dH * gelu'(Z) -> bf16
__global__ void k_gelu_back(const float* dH, const float* Z, long n, uint16_t* dZb) {
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dZb[q] = f2bf(dH[q] * gelu_grad(Z[q]));
}

// Column sums of a bf16 [rows][cols] matrix (bias gradients), two passes over
// fixed row chunks so the order is deterministic.
__global__ void k_colsum_bf16(const uint16_t* x, long rows, int cols, int chunk, float* part) {
    int c = threadIdx.x + blockIdx.y * blockDim.x;
    if (c >= cols) return;
    long r0 = (long)blockIdx.x * chunk, r1 = r0 + chunk < rows ? r0 + chunk : rows;
    float s = 0.f;
    for (long r = r0; r < r1; ++r) s += bf2f(x[r * cols + c]);
    part[(long)blockIdx.x * cols + c] = s;
}

// head gradients: dw4[k] = sum_i dy_i H3[i][k], db4 = sum_i dy_i (chunked partials)
__global__ void k_head_grad(const float* dy, const uint16_t* H, long n, int chunk, float* part) {
    int k = threadIdx.x;  // 0..256 (256 = bias)
    long r0 = (long)blockIdx.x * chunk, r1 = r0 + chunk < n ? r0 + chunk : n;
    float s = 0.f;
    for (long r = r0; r < r1; ++r) s += dy[r] * (k < 256 ? bf2f(H[r * 256 + k]) : 1.0f);
    part[(long)blockIdx.x * 257 + k] = s;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of one layer's weights: W [out][in] and W^T [in][out]
__global__ void k_weights_bf16(const float* W, int out, int in, uint16_t* Wb, uint16_t* WTb) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= out * in) return;
    int o = q / in, i = q % in;
    uint16_t h = f2bf(W[q]);
    Wb[q] = h;
    WTb[(long)i * out + o] = h;
}

__global__ void k_reduce_sq(const float* sq, long n, float* out) {
    __shared__ float red[256];
    float s = 0.f;
    for (long i = threadIdx.x; i < n; i += 256) s += sq[i];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int i = 0; i < 256; ++i) t += red[i];
        out[0] = t / (float)n;
    }
}

}  // namespace sb

using namespace sb;

namespace {

constexpr int kLayers = 5;
constexpr int kIn[kLayers] = {MLP_FEATURES, 256, 256, 256, 256};
constexpr int kOut[kLayers] = {256, 256, 256, 256, 1};

long param_offset(int l) {
    long o = 0;
    for (int q = 0; q < l; ++q) o += (long)kOut[q] * kIn[q] + kOut[q];
    return o;
}

unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

struct sb_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_input = nullptr;  // orders device inputs written on the null stream
    long n_params = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* Wb[kLayers] = {};
    uint16_t* WTb[kLayers] = {};
    long step = 0;
    // activations for up to cap rows
    long cap = 0;
    uint16_t* Xb = nullptr;          // [cap][64]
    float* Z[4] = {};                // pre-activations [cap][256]
    uint16_t* Hb[4] = {};            // bf16 activations [cap][256]
    float *y = nullptr, *dy = nullptr, *sq = nullptr, *loss = nullptr;
    uint16_t *dZb = nullptr, *T0 = nullptr, *T1 = nullptr;  // dZ (bf16), transposes
    float *dH = nullptr, *part = nullptr;
    long part_cap = 0;

    ~sb_mlp() {
        (void)hipSetDevice(device);
        void* bufs[] = {p, g, m, v, Xb, y, dy, sq, loss, dZb, T0, T1, dH, part};
        for (void* q : bufs)
            if (q) (void)hipFree(q);
        for (int l = 0; l < kLayers; ++l) {
            if (Wb[l]) (void)hipFree(Wb[l]);
            if (WTb[l]) (void)hipFree(WTb[l]);
        }
        for (int l = 0; l < 4; ++l) {
            if (Z[l]) (void)hipFree(Z[l]);
            if (Hb[l]) (void)hipFree(Hb[l]);
        }
        if (ev_input) (void)hipEventDestroy(ev_input);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void create(int dev, unsigned long long seed) {
        device = dev;
        check(hipSetDevice(dev), "hipSetDevice");
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreateWithFlags(&ev_input, hipEventDisableTiming), "hipEventCreate");
        n_params = param_offset(kLayers);
        p = dalloc<float>(n_params);
        g = dalloc<float>(n_params);
        m = dalloc<float>(n_params);
        v = dalloc<float>(n_params);
        for (int l = 0; l < kLayers; ++l) {
            Wb[l] = dalloc<uint16_t>((long)kOut[l] * kIn[l]);
            WTb[l] = dalloc<uint16_t>((long)kOut[l] * kIn[l]);
        }
        // He-normal weights (seeded), zero biases
        std::mt19937_64 rng(seed);
        std::vector<float> h(n_params, 0.f);
        for (int l = 0; l < kLayers; ++l) {
            std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / kIn[l]));
            float* W = h.data() + param_offset(l);
            for (long q = 0; q < (long)kOut[l] * kIn[l]; ++q) W[q] = nd(rng) * (l == kLayers - 1 ? 0.5f : 1.0f);
        }
        set_params(h.data());
    }

    void set_params(const float* host) {
        check(hipMemcpyAsync(p, host, sizeof(float) * n_params, hipMemcpyHostToDevice, stream), "upload params");
        check(hipMemsetAsync(m, 0, sizeof(float) * n_params, stream), "memset");
        check(hipMemsetAsync(v, 0, sizeof(float) * n_params, stream), "memset");
        step = 0;
        refresh_bf16();
        check(hipStreamSynchronize(stream), "set params");
    }

    void refresh_bf16() {
        for (int l = 0; l < kLayers; ++l) {
            long n = (long)kOut[l] * kIn[l];
            hipLaunchKernelGGL(k_weights_bf16, dim3(nblk(n, 256)), dim3(256), 0, stream, p + param_offset(l), kOut[l],
                               kIn[l], Wb[l], WTb[l]);
        }
        check(hipGetLastError(), "weights bf16");
    }

    void ensure(long n) {
        if (n <= cap) return;
        long c = ((n + 1023) / 1024) * 1024;
        void* bufs[] = {Xb, y, dy, sq, dZb, T0, T1, dH};
        for (void* q : bufs)
            if (q) (void)hipFree(q);
        for (int l = 0; l < 4; ++l) {
            if (Z[l]) (void)hipFree(Z[l]);
            if (Hb[l]) (void)hipFree(Hb[l]);
        }
        Xb = dalloc<uint16_t>(c * MLP_FEATURES);
        for (int l = 0; l < 4; ++l) {
            Z[l] = dalloc<float>(c * 256);
            Hb[l] = dalloc<uint16_t>(c * 256);
        }
        y = dalloc<float>(c);
        dy = dalloc<float>(c);
        sq = dalloc<float>(c);
        if (!loss) loss = dalloc<float>(1);
        dZb = dalloc<uint16_t>(c * 256);
        T0 = dalloc<uint16_t>(c * 256);
        T1 = dalloc<uint16_t>(c * 256);
        dH = dalloc<float>(c * 256);
        cap = c;
    }

    void ensure_part(long n) {
        if (n <= part_cap) return;
        if (part) (void)hipFree(part);
        part = dalloc<float>(n);
        part_cap = n;
    }

    void gemm(const uint16_t* A, const uint16_t* Bt, int M, int N, int K, float* C, const float* bias, uint16_t* Hbo,
              int slices) {
        GemmArgs a{A, Bt, M, N, K, K, C, bias, Hbo};
        if (slices > 1) {
            int ks = ((K / slices + 15) / 16) * 16;
            slices = (K + ks - 1) / ks;
            a.k_split = ks;
            ensure_part((long)slices * M * N);
            a.C = part;
            a.bias = nullptr;
            a.Hb = nullptr;
        }
        hipLaunchKernelGGL(k_gemm_bf16, dim3(nblk(N, 64), nblk(M, 64), slices), dim3(256), 0, stream, a);
        if (slices > 1)
            hipLaunchKernelGGL(k_sum_slices, dim3(nblk((long)M * N, 256)), dim3(256), 0, stream, part, slices,
                               (long)M * N, C);
        check(hipGetLastError(), "gemm");
    }

    // forward for n rows of fp32 features (device); y (device) gets the outputs
    void forward(const float* X, long n, const float* target) {
        ensure(n);
        hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n * MLP_FEATURES, 256)), dim3(256), 0, stream, X, n * MLP_FEATURES, Xb);
        const uint16_t* in = Xb;
        for (int l = 0; l < 4; ++l) {
            gemm(in, Wb[l], (int)n, 256, kIn[l], Z[l], p + param_offset(l) + (long)kOut[l] * kIn[l], Hb[l], 1);
            in = Hb[l];
        }
        const float* w4 = p + param_offset(4);
        hipLaunchKernelGGL(k_head, dim3(nblk(n, 256)), dim3(256), 0, stream, Hb[3], w4, w4 + 256, n, target, y, dy,
                           sq);
        check(hipGetLastError(), "head");
    }

    // forward + backward into g (device); returns the mean squared error
    float gradients(const float* X, const float* target, long n) {
        forward(X, n, target);
        hipLaunchKernelGGL(k_reduce_sq, dim3(1), dim3(256), 0, stream, sq, n, loss);
        const int chunk = 1024;
        long nch = (n + chunk - 1) / chunk;
        const long n16 = ((n + 15) / 16) * 16;
        // head
        ensure_part(nch * 257);
        hipLaunchKernelGGL(k_head_grad, dim3((unsigned)nch), dim3(257), 0, stream, dy, Hb[3], n, chunk, part);
        hipLaunchKernelGGL(k_sum_slices, dim3(nblk(257, 256)), dim3(256), 0, stream, part, (int)nch, 257L,
                           g + param_offset(4));
        const float* w4 = p + param_offset(4);
        hipLaunchKernelGGL(k_head_back, dim3(nblk(n * 256, 256)), dim3(256), 0, stream, dy, w4, Z[3], n, dZb);
        for (int l = 3; l >= 0; --l) {
            const int out = kOut[l], in = kIn[l];
            const uint16_t* Hin = l > 0 ? Hb[l - 1] : Xb;
            // dW_l = dZ_l^T . H_{l-1}: A = dZ^T [out][n16], Bt = H^T [in][n16]
            hipLaunchKernelGGL(k_transpose_bf16, dim3(nblk(out, 32), nblk(n16, 32)), dim3(256), 0, stream, dZb, (int)n,
                               out, (int)n16, T0);
            hipLaunchKernelGGL(k_transpose_bf16, dim3(nblk(in, 32), nblk(n16, 32)), dim3(256), 0, stream, Hin, (int)n,
                               in, (int)n16, T1);
            int slices = (int)std::min<long>(64, std::max<long>(1, n16 / 1024));
            gemm(T0, T1, out, in, (int)n16, g + param_offset(l), nullptr, nullptr, slices);
            // db_l = column sums of dZ_l
            ensure_part(nch * out);
            hipLaunchKernelGGL(k_colsum_bf16, dim3((unsigned)nch, nblk(out, 256)), dim3(256), 0, stream, dZb, n, out,
                               chunk, part);
            hipLaunchKernelGGL(k_sum_slices, dim3(nblk(out, 256)), dim3(256), 0, stream, part, (int)nch, (long)out,
                               g + param_offset(l) + (long)out * in);
            if (l > 0) {
                // dH_{l-1} = dZ_l . W_l: A = dZ [n][out], Bt = W^T [in][out]
                gemm(dZb, WTb[l], (int)n, in, out, dH, nullptr, nullptr, 1);
                hipLaunchKernelGGL(k_gelu_back, dim3(nblk(n * in, 256)), dim3(256), 0, stream, dH, Z[l - 1], n * in,
                                   dZb);
            }
        }
        check(hipGetLastError(), "backward");
        float h = 0.f;
        check(hipMemcpyAsync(&h, loss, sizeof(float), hipMemcpyDeviceToHost, stream), "loss");
        check(hipStreamSynchronize(stream), "gradients");
        return h;
    }

    float train_step(const float* X, const float* target, long n, float lr) {
        float l = gradients(X, target, n);
        step += 1;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        hipLaunchKernelGGL(k_adam, dim3(nblk(n_params, 256)), dim3(256), 0, stream, p, g, m, v, n_params, lr, b1, b2,
                           eps, c1, c2);
        refresh_bf16();
        check(hipStreamSynchronize(stream), "train step");
        return l;
    }
};

namespace sb {
// Inference on a caller stream for the warm start inside a step: y[n] from X[n][64] (device).
hipError_t mlp_forward_on(sb_mlp* m, const float* X, long n, float* y, hipStream_t st) {
    hipStream_t keep = m->stream;
    m->stream = st;
    hipError_t e = hipSuccess;
    try {
        m->forward(X, n, nullptr);
        e = hipMemcpyAsync(y, m->y, sizeof(float) * n, hipMemcpyDeviceToDevice, st);
    } catch (const HipError& err) {
        e = (hipError_t)err.code;
    }
    m->stream = keep;
    return e;
}
void mlp_reserve(sb_mlp* m, long n) { m->ensure(n); }
}  // namespace sb

extern "C" {

sb_mlp* sb_mlp_create(int device, unsigned long long seed) {
    auto m = std::make_unique<sb_mlp>();
    int rc = guarded([&] { m->create(device, seed); });
    return rc == 0 ? m.release() : nullptr;
}

void sb_mlp_destroy(sb_mlp* m) { delete m; }

long sb_mlp_nr_params(const sb_mlp* m) { return m ? m->n_params : -1; }

int sb_mlp_get_params(const sb_mlp* m, float* out, size_t cap) {
    if (!m || !out) return -1;
    if (cap < (size_t)m->n_params) return -3;
    return guarded([&] {
        check(hipMemcpyAsync(out, m->p, sizeof(float) * m->n_params, hipMemcpyDeviceToHost, m->stream), "download");
        check(hipStreamSynchronize(m->stream), "get params");
    });
}

int sb_mlp_set_params(sb_mlp* m, const float* params, size_t n) {
    if (!m || !params) return -1;
    if (n != (size_t)m->n_params) return -3;
    return guarded([&] { m->set_params(params); });
}

int sb_mlp_forward(sb_mlp* m, const float* x_dev, long n, float* y_dev) {
    if (!m || !x_dev || !y_dev) return -1;
    if (n <= 0) return -3;
    return guarded([&] {
        check(hipSetDevice(m->device), "hipSetDevice");
        wait_for_null_stream(m->stream, m->ev_input);
        m->forward(x_dev, n, nullptr);
        check(hipMemcpyAsync(y_dev, m->y, sizeof(float) * n, hipMemcpyDeviceToDevice, m->stream), "copy");
        check(hipStreamSynchronize(m->stream), "forward");
    });
}

int sb_mlp_gradients(sb_mlp* m, const float* x_dev, const float* t_dev, long n, float* grads_host, float* loss) {
    if (!m || !x_dev || !t_dev) return -1;
    if (n <= 0) return -3;
    return guarded([&] {
        check(hipSetDevice(m->device), "hipSetDevice");
        wait_for_null_stream(m->stream, m->ev_input);
        float l = m->gradients(x_dev, t_dev, n);
        if (loss) *loss = l;
        if (grads_host) {
            check(hipMemcpyAsync(grads_host, m->g, sizeof(float) * m->n_params, hipMemcpyDeviceToHost, m->stream),
                  "download");
            check(hipStreamSynchronize(m->stream), "gradients");
        }
    });
}

int sb_mlp_train_step(sb_mlp* m, const float* x_dev, const float* t_dev, long n, float lr, float* loss) {
    if (!m || !x_dev || !t_dev) return -1;
    if (n <= 0) return -3;
    return guarded([&] {
        check(hipSetDevice(m->device), "hipSetDevice");
        wait_for_null_stream(m->stream, m->ev_input);
        float l = m->train_step(x_dev, t_dev, n, lr);
        if (loss) *loss = l;
    });
}

}  // extern "C"
