// Learned warm start for the circulation solve: a per-strip MLP
// (64 -> 256 -> 256 -> 256 -> 256 -> 1, GELU in its sigmoid form) whose training and inference
// GEMMs run on the gfx950 matrix cores in bf16 with fp32 accumulation and fp32
// master weights (Adam).  No counterpart in the reference: the reference's
// solver starts from the previous circulation (or zero at the first step,
// lifting_line/simulation.rs:91); the network only supplies that initial
// guess, and the unchanged fixed-point solve finishes from it.
//
// GEMMs on v_mfma_f32_16x16x32_bf16 (gfx950): fragment layout (wave64) lane l
// holds A[l%16][8(l/16)+v] and B[l%16][8(l/16)+v] for v = 0..7 (K = 32) and
// returns D[4(l/16)+v][l%16], v = 0..3, D = A B^T.  The kernels pass the output's
// column operand as A and its row operand as B, so that a lane's four results
// are four adjacent columns of one output row (one 16-byte store).  Kernels
// staging 32-deep K slices of 128-row operand tiles through double-buffered LDS:
// - k_gemm_nt: C[M][N] = A[M][K] * Bt[N][K]^T (K contiguous in both), with the
//   layer's epilogue fused (bias and GELU forward; GELU derivative backward);
// - k_gemm_tn: C[O][I] = sum over rows r of A[r][O] B[r][I] (the weight
//   gradient, rows contiguous in neither operand: each tile is transposed on
//   its way into LDS), split over row slices, with the bias gradient (column
//   sums of A) taken from the same staged tile.
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

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// fp32 -> bf16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// two of them in one 32-bit word, a in the low half
__device__ inline uint32_t f2bf2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
__device__ __host__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// e^x on the hardware's base-2 exponential (relative error ~1e-7; 0 below e^-87)
__device__ inline float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// GELU in the sigmoid form of Hendrycks & Gimpel (2016), x s(1.702 x) with s the
// logistic function (within 0.02 of x Phi(x)): an exponential, a reciprocal and three
// more operations, no branch; the values it feeds are rounded to bf16 right after.
__device__ inline float gelu_s(float x) { return __builtin_amdgcn_rcpf(1.0f + exp_hw(-1.702f * x)); }
__device__ inline float gelu(float x) { return x * gelu_s(x); }
__device__ inline float gelu_grad(float x) {  // s + 1.702 x s (1 - s)
    const float sg = gelu_s(x);
    return sg + 1.702f * x * (sg * (1.0f - sg));
}

__device__ inline f32x4 mfma_bf16_k32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

constexpr int GT = 128;   // output tile edge (both kernels)
constexpr int GK = 32;    // K depth of one LDS stage
constexpr int GLD = 40;   // LDS row stride in bf16 (80 B: 16-byte aligned fragment reads)

enum { EPI_PLAIN = 0, EPI_FWD = 1, EPI_BACK = 2 };

struct NtArgs {
    const uint16_t* A;   // [M][K]
    const uint16_t* Bt;  // [N][K]
    int M, N, K;         // K a multiple of 32, N a multiple of 16
    int epi;
    const float* bias;   // [N] (EPI_FWD)
    float* Z;            // fp32 [M][N]: the product (EPI_PLAIN) or bias + product (EPI_FWD; null: not kept)
    uint16_t* Hb;        // bf16 [M][N]: gelu(product + bias) (EPI_FWD), product * gelu'(Zin) (EPI_BACK)
    const float* Zin;    // [M][N] (EPI_BACK)
};

// 128x128 output tile per workgroup of 4 waves (each 64x64: 4x4 MFMA tiles).
__global__ __launch_bounds__(256) void k_gemm_nt(NtArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][GT][GLD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][GT][GLD];
    const int t = threadIdx.x, wave = t / 64, lane = t % 64;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
    const s16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    s16x8 ra[2], rb[2];
    // thread t moves 16-byte pieces t and t + 256 of each 128 x 32 operand slice
    auto load = [&](int k) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int seg = t + 256 * i, row = seg / 4, part = seg % 4;
            ra[i] = m0 + row < g.M ? *reinterpret_cast<const s16x8*>(g.A + (long)(m0 + row) * g.K + k + 8 * part)
                                   : zero;
            rb[i] = n0 + row < g.N ? *reinterpret_cast<const s16x8*>(g.Bt + (long)(n0 + row) * g.K + k + 8 * part)
                                   : zero;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int seg = t + 256 * i, row = seg / 4, part = seg % 4;
            *reinterpret_cast<s16x8*>(&As[buf][row][8 * part]) = ra[i];
            *reinterpret_cast<s16x8*>(&Bs[buf][row][8 * part]) = rb[i];
        }
    };
    // the accumulators start at the bias (EPI_FWD; zero otherwise): z = b + sum of products
    f32x4 acc[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = n0 + wn + 16 * j + 4 * (lane / 16);
        const f32x4 b0 = g.epi == EPI_FWD && col < g.N ? *reinterpret_cast<const f32x4*>(g.bias + col)
                                                        : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i][j] = b0;
    }
    const int nk = g.K / GK;
    load(0);
    store(0);
    __syncthreads();
    for (int s = 0; s < nk; ++s) {
        const int buf = s & 1;
        if (s + 1 < nk) load((s + 1) * GK);  // in flight while this stage computes
        s16x8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const s16x8*>(&As[buf][wm + 16 * i + lane % 16][8 * (lane / 16)]);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const s16x8*>(&Bs[buf][wn + 16 * j + lane % 16][8 * (lane / 16)]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = mfma_bf16_k32(b[j], a[i], acc[i][j]);
        if (s + 1 < nk) store(buf ^ 1);  // that half was last read before the previous barrier
        __syncthreads();
    }
    // lane: row m0 + wm + 16 i + lane % 16, columns n0 + wn + 16 j + 4 (lane / 16) + (0..3)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = m0 + wm + 16 * i + lane % 16;
            const int col = n0 + wn + 16 * j + 4 * (lane / 16);
            if (row >= g.M || col >= g.N) continue;  // N is a multiple of 16: all four columns or none
            const long q = (long)row * g.N + col;
            f32x4 z = acc[i][j];
            if (g.epi == EPI_BACK) {
                const f32x4 zi = *reinterpret_cast<const f32x4*>(g.Zin + q);
                *reinterpret_cast<uint2*>(g.Hb + q) = make_uint2(f2bf2(z[0] * gelu_grad(zi[0]), z[1] * gelu_grad(zi[1])),
                                                                 f2bf2(z[2] * gelu_grad(zi[2]), z[3] * gelu_grad(zi[3])));
            } else {
                if (g.Z) *reinterpret_cast<f32x4*>(g.Z + q) = z;
                if (g.epi == EPI_FWD)
                    *reinterpret_cast<uint2*>(g.Hb + q) =
                        make_uint2(f2bf2(gelu(z[0]), gelu(z[1])), f2bf2(gelu(z[2]), gelu(z[3])));
            }
        }
}

struct TnArgs {
    const uint16_t* A;    // [M][O]
    const uint16_t* B;    // [M][I]
    int M, O, I;          // I a multiple of 16
    int rows_per_slice;   // a multiple of 32; slice z covers rows [z rows_per_slice, ...)
    float* part;          // [slices][O][I]
    float* part_bias;     // [slices][O]: column sums of A over the slice (tiles of column 0)
};

// 128x128 tile of C per workgroup and row slice; each 32-row stage of A and B is
// loaded row-coalesced (lane = column) and stored transposed, rows contiguous.
__global__ __launch_bounds__(256) void k_gemm_tn(TnArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][GT][GLD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][GT][GLD];
    __shared__ float red[256];
    const int t = threadIdx.x, wave = t / 64, lane = t % 64;
    const int o0 = blockIdx.y * GT, i0 = blockIdx.x * GT;
    const long r_begin = (long)blockIdx.z * g.rows_per_slice;
    const long r_end = r_begin + g.rows_per_slice < g.M ? r_begin + g.rows_per_slice : g.M;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
    const int c = t % GT, h = t / GT;  // this thread's column and half (rows 16h .. 16h + 15 of a stage)
    const bool oc = o0 + c < g.O, ic = i0 + c < g.I;
    const bool sums = blockIdx.x == 0;
    s16x8 va[2], vb[2];
    float bsum = 0.f;
    auto load = [&](long r) {
        const long r1 = r + 16 * h;
        const int rows = (int)(r_end - r1 < 16 ? (r_end - r1 > 0 ? r_end - r1 : 0) : 16);
        const uint16_t* pa = g.A + r1 * g.O + o0 + c;
        const uint16_t* pb = g.B + r1 * g.I + i0 + c;
        if (rows == 16 && oc && ic) {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                va[q / 8][q % 8] = (short)*pa;
                vb[q / 8][q % 8] = (short)*pb;
                pa += g.O;
                pb += g.I;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                va[q / 8][q % 8] = (short)(q < rows && oc ? *pa : 0);
                vb[q / 8][q % 8] = (short)(q < rows && ic ? *pb : 0);
                pa += g.O;
                pb += g.I;
            }
        }
    };
    auto store = [&](int buf) {
        *reinterpret_cast<s16x8*>(&As[buf][c][16 * h]) = va[0];
        *reinterpret_cast<s16x8*>(&As[buf][c][16 * h + 8]) = va[1];
        *reinterpret_cast<s16x8*>(&Bs[buf][c][16 * h]) = vb[0];
        *reinterpret_cast<s16x8*>(&Bs[buf][c][16 * h + 8]) = vb[1];
        if (sums) {
#pragma unroll
            for (int q = 0; q < 16; ++q) bsum += bf2f((uint16_t)va[q / 8][q % 8]);
        }
    };
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const long nk = r_end > r_begin ? (r_end - r_begin + GK - 1) / GK : 0;
    if (nk > 0) {
        load(r_begin);
        store(0);
    }
    __syncthreads();
    for (long s = 0; s < nk; ++s) {
        const int buf = (int)(s & 1);
        if (s + 1 < nk) load(r_begin + (s + 1) * GK);
        s16x8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const s16x8*>(&As[buf][wm + 16 * i + lane % 16][8 * (lane / 16)]);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const s16x8*>(&Bs[buf][wn + 16 * j + lane % 16][8 * (lane / 16)]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = mfma_bf16_k32(b[j], a[i], acc[i][j]);
        if (s + 1 < nk) store(buf ^ 1);
        __syncthreads();
    }
    float* C = g.part + (long)blockIdx.z * g.O * g.I;
    // lane: row o0 + wm + 16 i + lane % 16 of C, columns i0 + wn + 16 j + 4 (lane / 16) + (0..3)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = o0 + wm + 16 * i + lane % 16;
            const int col = i0 + wn + 16 * j + 4 * (lane / 16);
            if (o < g.O && col < g.I) *reinterpret_cast<f32x4*>(C + (long)o * g.I + col) = acc[i][j];
        }
    if (sums) {
        red[t] = bsum;
        __syncthreads();
        if (t < GT && oc) g.part_bias[(long)blockIdx.z * g.O + o0 + t] = red[t] + red[t + GT];
    }
}

// Whole-network inference, one kernel: 128 rows per workgroup, each of its WAVES
// waves owning 128 / WAVES rows whose activations stay in LDS from layer to layer
// (bf16, as the layer-by-layer forward rounds them), the layer's weights streamed
// through LDS in 64-deep K slices shared by the waves.  Every product and epilogue is
// the one k_gemm_nt and k_head compute (same instruction, same K order), so the
// outputs equal the layer-by-layer forward bit for bit.
constexpr int IR = 128;   // rows per workgroup
constexpr int ILD = 264;  // activation row stride in bf16 (528 B)
constexpr int IBK = 64;   // K depth of one weight slice
constexpr int IWLD = 72;  // weight-slice row stride in bf16 (144 B)
constexpr int kInferWaves = 4;  // waves per inference workgroup (32 rows each; eight waves of 16 rows measured no faster)

struct InferArgs {
    const float* X;          // [n][64]
    long n;
    const uint16_t* W[4];    // bf16 [256][in_l]
    const float* b[4];       // [256]
    const float* w4;         // head weights [256] (fp32; rounded to bf16 as k_head does)
    const float* b4;         // [1]
    float* y;                // [n]
};

template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_mlp_infer(InferArgs g) {
    constexpr int NT = 64 * WAVES;       // threads
    constexpr int RT = IR / 16 / WAVES;  // 16-row MFMA tiles per wave
    __shared__ __attribute__((aligned(16))) uint16_t H[IR][ILD];
    __shared__ __attribute__((aligned(16))) uint16_t Ws[2][256][IWLD];
    __shared__ float w4s[256];  // head weights as k_head rounds them
    const int t = threadIdx.x, wave = t / 64, lane = t % 64;
    const long r0 = (long)blockIdx.x * IR;
    if (t < 256) w4s[t] = bf2f(f2bf(g.w4[t]));
    {
        // the features: 128 rows x 64 fp32 = 2048 pieces of 4 values, all loads in
        // flight before the first store
        constexpr int PER = IR * MLP_FEATURES / 4 / NT;
        float4 f[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int q = t + NT * i, row = q / (MLP_FEATURES / 4), c4 = q % (MLP_FEATURES / 4);
            f[i] = r0 + row < g.n ? *reinterpret_cast<const float4*>(g.X + (r0 + row) * MLP_FEATURES + 4 * c4)
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int q = t + NT * i, row = q / (MLP_FEATURES / 4), c4 = q % (MLP_FEATURES / 4);
            *reinterpret_cast<uint2*>(&H[row][4 * c4]) = make_uint2(f2bf2(f[i].x, f[i].y), f2bf2(f[i].z, f[i].w));
        }
    }
    const int wr = wave * 16 * RT;
    constexpr int WP = 256 * IBK / 8 / NT;  // 16-byte pieces of a weight slice per thread
    s16x8 rw[WP];
    // thread t moves 16-byte pieces t + NT i of each 256 x 64 weight slice
    auto load_w = [&](const uint16_t* W, int K, int k) {
#pragma unroll
        for (int i = 0; i < WP; ++i) {
            const int p = t + NT * i, row = p / 8, part = p % 8;
            rw[i] = *reinterpret_cast<const s16x8*>(W + (long)row * K + k + 8 * part);
        }
    };
    auto store_w = [&](int buf) {
#pragma unroll
        for (int i = 0; i < WP; ++i) {
            const int p = t + NT * i, row = p / 8, part = p % 8;
            *reinterpret_cast<s16x8*>(&Ws[buf][row][8 * part]) = rw[i];
        }
    };
    load_w(g.W[0], MLP_FEATURES, 0);
    for (int l = 0; l < 4; ++l) {
        const int K = l == 0 ? MLP_FEATURES : 256;
        const int nk = K / IBK;
        const uint16_t* W = g.W[l];
        store_w(0);  // the layer's first slice, loaded before the previous layer's epilogue
        __syncthreads();
        // the accumulators start at the bias: z = b + sum of products
        f32x4 acc[RT][16];
        const float* bias = g.b[l];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const f32x4 b0 = *reinterpret_cast<const f32x4*>(bias + 16 * j + 4 * (lane / 16));
#pragma unroll
            for (int i = 0; i < RT; ++i) acc[i][j] = b0;
        }
        for (int s = 0; s < nk; ++s) {
            const int buf = s & 1;
            if (s + 1 < nk) load_w(W, K, (s + 1) * IBK);
#pragma unroll
            for (int kk = 0; kk < IBK / GK; ++kk) {
                const int kc = IBK * s + GK * kk + 8 * (lane / 16);
                s16x8 a[RT];
#pragma unroll
                for (int i = 0; i < RT; ++i) a[i] = *reinterpret_cast<const s16x8*>(&H[wr + 16 * i + lane % 16][kc]);
                // all sixteen weight fragments of this K step in flight before the first product
                s16x8 bf[16];
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    bf[j] = *reinterpret_cast<const s16x8*>(&Ws[buf][16 * j + lane % 16][GK * kk + 8 * (lane / 16)]);
#pragma unroll
                for (int j = 0; j < 16; ++j) {
#pragma unroll
                    for (int i = 0; i < RT; ++i) acc[i][j] = mfma_bf16_k32(bf[j], a[i], acc[i][j]);
                }
                // schedule: six LDS reads ahead, then two products per further read, so that
                // each read's latency hides behind a dozen products
                static_assert(RT == 2, "the schedule below counts two row tiles");
                __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
#pragma unroll
                for (int q = 0; q < 12; ++q) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                }
                __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
            }
            if (s + 1 < nk) store_w(buf ^ 1);
            __syncthreads();
        }
        if (l < 3) load_w(g.W[l + 1], 256, 0);  // in flight during the epilogue
        // this wave's rows: GELU, back into the activations (read only by this wave)
#pragma unroll
        for (int i = 0; i < RT; ++i)
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int row = wr + 16 * i + lane % 16, col = 16 * j + 4 * (lane / 16);
                const f32x4 z = acc[i][j];
                *reinterpret_cast<uint2*>(&H[row][col]) =
                    make_uint2(f2bf2(gelu(z[0]), gelu(z[1])), f2bf2(gelu(z[2]), gelu(z[3])));
            }
    }
    __syncthreads();
    if (t < IR && r0 + t < g.n) {
        float s = 0.f;
        for (int k = 0; k < 256; ++k) s += bf2f(H[t][k]) * w4s[k];
        s += g.b4[0];
        g.y[r0 + t] = s;
    }
}

// out[i] = sum over slices s of part[s][i] (ascending s), i < n
__global__ void k_sum_slices(const float* part, int slices, long n, float* out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    int q = 0;
    for (; q + 8 <= slices; q += 8) {  // eight loads in flight, added in slice order
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[(long)(q + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; q < slices; ++q) s += part[(long)q * n + i];
    out[i] = s;
}

__global__ void k_to_bf16(const float* x, long n, uint16_t* y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
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
    uint16_t *dZb = nullptr, *dZb2 = nullptr;  // back-propagated dZ (bf16), two layers in turn
    float* part = nullptr;
    long part_cap = 0;

    ~sb_mlp() {
        (void)hipSetDevice(device);
        void* bufs[] = {p, g, m, v, Xb, y, dy, sq, loss, dZb, dZb2, part};
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
        void* bufs[] = {Xb, y, dy, sq, dZb, dZb2};
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
        dZb2 = dalloc<uint16_t>(c * 256);
        cap = c;
    }

    void ensure_part(long n) {
        if (n <= part_cap) return;
        if (part) (void)hipFree(part);
        part = dalloc<float>(n);
        part_cap = n;
    }

    // C[M][N] = A[M][K] Bt[N][K]^T with the epilogue `epi` (k_gemm_nt)
    void gemm_nt(const uint16_t* A, const uint16_t* Bt, int M, int N, int K, int epi, const float* bias, float* Zo,
                 uint16_t* Hbo, const float* Zin) {
        NtArgs a{A, Bt, M, N, K, epi, bias, Zo, Hbo, Zin};
        hipLaunchKernelGGL(k_gemm_nt, dim3(nblk(N, GT), nblk(M, GT)), dim3(256), 0, stream, a);
        check(hipGetLastError(), "gemm");
    }

    // W-gradient C[O][I] = A[M][O]^T B[M][I] and the bias gradient (column sums of A),
    // over at most 128 row slices of 512 or more rows, summed in slice order (k_gemm_tn)
    void gemm_tn(const uint16_t* A, const uint16_t* B, long M, int O, int I, float* C, float* bias_grad) {
        const int slices = (int)std::min<long>(128, std::max<long>(1, (M + 511) / 512));
        const int rows = (int)(((M + slices - 1) / slices + GK - 1) / GK * GK);
        ensure_part((long)slices * O * I + (long)slices * O);
        TnArgs a{A, B, (int)M, O, I, rows, part, part + (long)slices * O * I};
        hipLaunchKernelGGL(k_gemm_tn, dim3(nblk(I, GT), nblk(O, GT), slices), dim3(256), 0, stream, a);
        hipLaunchKernelGGL(k_sum_slices, dim3(nblk((long)O * I, 256)), dim3(256), 0, stream, part, slices,
                           (long)O * I, C);
        hipLaunchKernelGGL(k_sum_slices, dim3(nblk(O, 256)), dim3(256), 0, stream, a.part_bias, slices, (long)O,
                           bias_grad);
        check(hipGetLastError(), "gemm");
    }

    // inference: y (device) for n rows of fp32 features (device), one kernel
    void infer(const float* X, long n, float* out) {
        InferArgs a{};
        a.X = X;
        a.n = n;
        for (int l = 0; l < 4; ++l) {
            a.W[l] = Wb[l];
            a.b[l] = p + param_offset(l) + (long)kOut[l] * kIn[l];
        }
        a.w4 = p + param_offset(4);
        a.b4 = a.w4 + 256;
        a.y = out;
        hipLaunchKernelGGL(k_mlp_infer<kInferWaves>, dim3(nblk(n, IR)), dim3(64 * kInferWaves), 0, stream, a);
        check(hipGetLastError(), "inference");
    }

    // forward for n rows of fp32 features (device), keeping what the backward pass needs;
    // y (device) gets the outputs
    void forward(const float* X, long n, const float* target) {
        ensure(n);
        hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n * MLP_FEATURES, 256)), dim3(256), 0, stream, X, n * MLP_FEATURES, Xb);
        const uint16_t* in = Xb;
        for (int l = 0; l < 4; ++l) {
            gemm_nt(in, Wb[l], (int)n, 256, kIn[l], EPI_FWD, p + param_offset(l) + (long)kOut[l] * kIn[l], Z[l], Hb[l],
                    nullptr);
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
        const int chunk = 128;
        long nch = (n + chunk - 1) / chunk;
        // head
        ensure_part(nch * 257);
        hipLaunchKernelGGL(k_head_grad, dim3((unsigned)nch), dim3(257), 0, stream, dy, Hb[3], n, chunk, part);
        hipLaunchKernelGGL(k_sum_slices, dim3(nblk(257, 256)), dim3(256), 0, stream, part, (int)nch, 257L,
                           g + param_offset(4));
        const float* w4 = p + param_offset(4);
        hipLaunchKernelGGL(k_head_back, dim3(nblk(n * 256, 256)), dim3(256), 0, stream, dy, w4, Z[3], n, dZb);
        uint16_t *dz = dZb, *dz_next = dZb2;
        for (int l = 3; l >= 0; --l) {
            const int out = kOut[l], in = kIn[l];
            const uint16_t* Hin = l > 0 ? Hb[l - 1] : Xb;
            // dW_l = dZ_l^T H_{l-1} and db_l = column sums of dZ_l
            gemm_tn(dz, Hin, n, out, in, g + param_offset(l), g + param_offset(l) + (long)out * in);
            if (l > 0) {
                // dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1}): A = dZ [n][out], Bt = W^T [in][out]
                gemm_nt(dz, WTb[l], (int)n, in, out, EPI_BACK, nullptr, nullptr, dz_next, Z[l - 1]);
                std::swap(dz, dz_next);
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
        m->infer(X, n, y);
    } catch (const HipError& err) {
        e = (hipError_t)err.code;
    }
    m->stream = keep;
    return e;
}
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
        m->infer(x_dev, n, y_dev);
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
