// mlp_kernels.hip -- learned troubled-cell indicator for DG P=2 (no reference
// counterpart).
//
// Network: 192 inputs (180 features, 12 zeros) -> 64 ReLU -> 64 ReLU -> 1
// sigmoid.  Features of an updated cell: its own 4 x 9 modes and those of its
// E, W, S(+i), N(-i) neighbours after the ghost fill, each variable divided by
// s_c = |mean_c| + |mean_rho| of the cell and the cell's own means subtracted
// from the mode-0 entries.  Weights are kept in FP32 (master copy for Adam)
// with BF16 copies feeding v_mfma_f32_16x16x32_bf16 (FP32 accumulation).
//
// MFMA 16x16x32 bf16 lane layout (lane l, g = l/16, i = l%16): A holds
// A[i][8g + e], B holds B[8g + e][i] (e = 0..7), D holds D[4g + v][i].
//
// Training is a sequence of NT GEMMs C[M][N] = A[M][K] Bt[N][K]^T plus
// elementwise kernels; inference inside the DG stage is one fused kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "zk_internal.h"
#include "zk_device.h"

#include "mlp_device.h"

struct zk_mlp {
    int device = 0;
    std::string err;
    hipStream_t stream = nullptr;
    float *w = nullptr;          // FP32 master parameters (P_N)
    float *g = nullptr;          // gradients (P_N)
    float *m1 = nullptr, *m2 = nullptr;  // Adam moments
    int adam_t = 0;
    __bf16 *w1b = nullptr;       // W1 [64][192]
    __bf16 *w2b = nullptr;       // W2 [64][64] natural
    __bf16 *w2p = nullptr;       // W2 with columns permuted for the inference B layout
    __bf16 *w2t = nullptr;       // W2^T [64][64]
    // dataset (harvested on the device)
    __bf16 *X = nullptr;         // [cap][192]
    float *y = nullptr;          // [cap]
    int cap = 0;
    int *count = nullptr;        // device counter of harvested samples
    // training scratch for batch B
    int B = 0;
    __bf16 *xb = nullptr, *xbt = nullptr, *h1 = nullptr, *h1t = nullptr, *dh2 = nullptr, *dh2t = nullptr, *dh1 = nullptr;
    float *h1pre = nullptr, *h2pre = nullptr, *z = nullptr, *dz = nullptr, *yb = nullptr, *dh1pre = nullptr;
    float *loss = nullptr;
    int *idx = nullptr;
};

namespace {

using zk::MLP_H;
using zk::MLP_IN;


// C[M][N] (fp32, ldc) = A[M][K] (bf16, lda) * Bt[N][K]^T (bf16, ldb).  One
// wave per 16x16 tile of C; M, N multiples of 16, K a multiple of 32, rows of
// A and Bt 16-byte aligned.
__global__ __launch_bounds__(256) void gemm_nt_kernel(const __bf16 *__restrict__ A, int lda, const __bf16 *__restrict__ Bt,
                                                      int ldb, float *__restrict__ C, int ldc, int M, int N, int K)
{
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int tiles_n = N / 16;
    if (wave >= (M / 16) * tiles_n) return;
    const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    const int m0 = (wave / tiles_n) * 16, n0 = (wave % tiles_n) * 16;
    const __bf16 *a = A + (size_t)(m0 + i) * lda + 8 * g;
    const __bf16 *b = Bt + (size_t)(n0 + i) * ldb + 8 * g;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < K; k += 32) {
        const bf16x8 av = *reinterpret_cast<const bf16x8 *>(a + k);
        const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(b + k);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) C[(size_t)(m0 + 4 * g + v) * ldc + n0 + i] = acc[v];
}

void gemm_nt(hipStream_t st, const __bf16 *A, int lda, const __bf16 *Bt, int ldb, float *C, int ldc, int M, int N, int K)
{
    const int waves = (M / 16) * (N / 16);
    hipLaunchKernelGGL(gemm_nt_kernel, dim3((waves + 3) / 4), dim3(256), 0, st, A, lda, Bt, ldb, C, ldc, M, N, K);
}

// ---- elementwise pieces of the training step ----
__global__ void gather_kernel(const __bf16 *__restrict__ X, const float *__restrict__ y, const int *__restrict__ idx,
                              __bf16 *__restrict__ xb, __bf16 *__restrict__ xbt, float *__restrict__ yb, int B)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * MLP_IN) return;
    const int c = t / MLP_IN, k = t % MLP_IN;
    const __bf16 v = X[(size_t)idx[c] * MLP_IN + k];
    xb[(size_t)c * MLP_IN + k] = v;
    xbt[(size_t)k * B + c] = v;
    if (k == 0) yb[c] = y[idx[c]];
}

// H1 = relu(H1pre + b1) as bf16 in [h][c] and [c][h] layouts
__global__ void act1_kernel(const float *__restrict__ pre, const float *__restrict__ w, __bf16 *__restrict__ h1,
                            __bf16 *__restrict__ h1t, int B)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MLP_H * B) return;
    const int h = t / B, c = t % B;
    const float v = fmaxf(pre[t] + w[zk::P_B1 + h], 0.0f);
    const __bf16 bv = (__bf16)v;
    h1[t] = bv;
    h1t[(size_t)c * MLP_H + h] = bv;
}

// z = w3 . relu(H2pre + b2) + b3, dz = (sigmoid(z) - y)/B, BCE loss, and dH2
__global__ void head_kernel(const float *__restrict__ h2pre, const float *__restrict__ w, const float *__restrict__ y,
                            float *__restrict__ z, float *__restrict__ dz, float *__restrict__ loss, int B)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= B) return;
    float acc = 0.0f;
    for (int j = 0; j < MLP_H; ++j) acc += w[zk::P_W3 + j] * fmaxf(h2pre[(size_t)j * B + c] + w[zk::P_B2 + j], 0.0f);
    const float zz = acc + w[zk::P_B3];
    z[c] = zz;
    const float p = zk::sigmoidf(zz);
    dz[c] = (p - y[c]) / (float)B;
    // stable BCE: max(z,0) - z*y + log(1 + exp(-|z|))
    const float l = fmaxf(zz, 0.0f) - zz * y[c] + log1pf(__expf(-fabsf(zz)));
    atomicAdd(loss, l / (float)B);
}

__global__ void dh2_kernel(const float *__restrict__ h2pre, const float *__restrict__ w, const float *__restrict__ dz,
                           __bf16 *__restrict__ dh2, __bf16 *__restrict__ dh2t, int B)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MLP_H * B) return;
    const int j = t / B, c = t % B;
    const float d = (h2pre[t] + w[zk::P_B2 + j] > 0.0f) ? dz[c] * w[zk::P_W3 + j] : 0.0f;
    const __bf16 bd = (__bf16)d;
    dh2[t] = bd;
    dh2t[(size_t)c * MLP_H + j] = bd;
}

// per-row sums for the bias / output-weight gradients (deterministic: one
// thread per output, serial over the batch)
__global__ void small_grads_kernel(const float *__restrict__ h2pre, const float *__restrict__ w,
                                   const float *__restrict__ dz, const float *__restrict__ dh1pre,
                                   const float *__restrict__ h1pre, float *__restrict__ g, int B)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MLP_H) {  // dw3[j] and db2[j]
        const int j = t;
        float a = 0.0f, b = 0.0f;
        for (int c = 0; c < B; ++c) {
            const float pre = h2pre[(size_t)j * B + c] + w[zk::P_B2 + j];
            const float h2 = fmaxf(pre, 0.0f);
            a += dz[c] * h2;
            if (pre > 0.0f) b += dz[c] * w[zk::P_W3 + j];
        }
        g[zk::P_W3 + j] = a;
        g[zk::P_B2 + j] = b;
    } else if (t < 2 * MLP_H) {  // db1[h]
        const int h = t - MLP_H;
        float a = 0.0f;
        for (int c = 0; c < B; ++c)
            if (h1pre[(size_t)h * B + c] + w[zk::P_B1 + h] > 0.0f) a += dh1pre[(size_t)h * B + c];
        g[zk::P_B1 + h] = a;
    } else if (t == 2 * MLP_H) {
        float a = 0.0f;
        for (int c = 0; c < B; ++c) a += dz[c];
        g[zk::P_B3] = a;
    }
}

// dH1 = (W2^T dH2) masked by the layer-1 ReLU, as bf16 [h][c]
__global__ void dh1_kernel(const float *__restrict__ dh1raw, const float *__restrict__ h1pre,
                           const float *__restrict__ w, __bf16 *__restrict__ dh1, int B)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= MLP_H * B) return;
    const int h = t / B;
    dh1[t] = (__bf16)((h1pre[t] + w[zk::P_B1 + h] > 0.0f) ? dh1raw[t] : 0.0f);
}

// W2 natural, transposed and inference-permuted bf16 copies; W1 bf16
__global__ void refresh_kernel(const float *__restrict__ w, __bf16 *__restrict__ w1b, __bf16 *__restrict__ w2b,
                               __bf16 *__restrict__ w2t, __bf16 *__restrict__ w2p)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < MLP_H * MLP_IN) w1b[t] = (__bf16)w[zk::P_W1 + t];
    if (t < MLP_H * MLP_H) {
        const int j = t / MLP_H, h = t % MLP_H;
        const __bf16 v = (__bf16)w[zk::P_W2 + t];
        w2b[t] = v;
        w2t[(size_t)h * MLP_H + j] = v;
        // inference B layout for layer 2: K slot kk = 32 s + 8 g + e carries
        // hidden unit 32 s + 16 (e / 4) + 4 g + (e % 4)
        const int kk = h, s = kk / 32, gg = (kk % 32) / 8, e = kk % 8;
        const int hp = 32 * s + 16 * (e / 4) + 4 * gg + (e % 4);
        w2p[(size_t)j * MLP_H + kk] = (__bf16)w[zk::P_W2 + j * MLP_H + hp];
    }
}

__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                            float *__restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float a = b1 * m1[t] + (1.0f - b1) * gt;
    const float b = b2 * m2[t] + (1.0f - b2) * gt * gt;
    m1[t] = a;
    m2[t] = b;
    w[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void zero_f_kernel(float *p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0f;
}

}  // namespace

namespace {

// Features of a list of cells (for tests and for harvesting): X[t][k].
__global__ void features_kernel(const double *__restrict__ Q, const int32_t *__restrict__ src, zk::StLay L, int nx,
                                const int32_t *__restrict__ cells, int n, __bf16 *__restrict__ X)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * MLP_IN) return;
    const int c = t / MLP_IN, k = t % MLP_IN;
    double mean[4], scale[4];
    zk::mlp_scales(Q, L, (size_t)cells[c], mean, scale);
    X[t] = (__bf16)zk::mlp_feature(Q, src, L, nx, (size_t)cells[c], mean, scale, k);
}

// Scores of a list of cells through the fused path (16 cells per wave).
__global__ __launch_bounds__(256) void score_kernel(const double *__restrict__ Q, const int32_t *__restrict__ src,
                                                    zk::StLay L, int nx, const int32_t *__restrict__ cells, int n,
                                                    const __bf16 *__restrict__ w1b, const __bf16 *__restrict__ w2p,
                                                    const float *__restrict__ w, float *__restrict__ out)
{
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int base = wave * 16;
    if (base >= n) return;
    const int i = threadIdx.x & 15;
    const int t = base + i < n ? base + i : n - 1;
    const float sc = zk::mlp_score16(Q, src, L, nx, (size_t)cells[t], w1b, w2p, w);
    if ((threadIdx.x & 63) < 16 && base + i < n) out[base + i] = sc;
}

// Inference on host-provided features (n multiple of 16): the same layers as
// mlp_score16 but reading X instead of computing features.
__global__ __launch_bounds__(256) void infer_x_kernel(const __bf16 *__restrict__ X, int n, const __bf16 *__restrict__ w1b,
                                                      const __bf16 *__restrict__ w2p, const float *__restrict__ w,
                                                      float *__restrict__ out)
{
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int base = wave * 16;
    if (base >= n) return;
    const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    const __bf16 *x = X + (size_t)(base + i) * MLP_IN;
    f4 h1acc[4];
    for (int t = 0; t < 4; ++t) h1acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < MLP_IN / 32; ++s) {
        const bf16x8 xv = *reinterpret_cast<const bf16x8 *>(x + 32 * s + 8 * g);
        for (int t = 0; t < 4; ++t) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(w1b + (size_t)(16 * t + i) * MLP_IN + 32 * s + 8 * g);
            h1acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, xv, h1acc[t], 0, 0, 0);
        }
    }
    bf16x8 hv[2];
    for (int s = 0; s < 2; ++s)
        for (int e = 0; e < 8; ++e) {
            const int t = 2 * s + e / 4, v = e % 4;
            hv[s][e] = (__bf16)fmaxf(h1acc[t][v] + w[zk::P_B1 + 16 * t + 4 * g + v], 0.0f);
        }
    float part = 0.0f;
    for (int t = 0; t < 4; ++t) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < 2; ++s) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(w2p + (size_t)(16 * t + i) * MLP_H + 32 * s + 8 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, hv[s], acc, 0, 0, 0);
        }
        for (int v = 0; v < 4; ++v) {
            const int j = 16 * t + 4 * g + v;
            part += w[zk::P_W3 + j] * fmaxf(acc[v] + w[zk::P_B2 + j], 0.0f);
        }
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    if (lane < 16) out[base + i] = zk::sigmoidf(part + w[zk::P_B3]);
}

#define MLP_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) {                                                           \
            M->err = std::string(#call) + ": " + hipGetErrorString(_e);                   \
            return ZK_ERR_HIP;                                                            \
        }                                                                                 \
    } while (0)

zk_status ensure_batch(zk_mlp *M, int B)
{
    if (M->B == B) return ZK_OK;
    void *ptrs[] = {M->xb, M->xbt, M->h1, M->h1t, M->dh2, M->dh2t, M->dh1, M->h1pre, M->h2pre, M->z, M->dz, M->yb,
                    M->dh1pre, M->idx};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    const size_t b = (size_t)B;
    MLP_HIP(hipMalloc(&M->xb, b * MLP_IN * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->xbt, b * MLP_IN * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->h1, b * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->h1t, b * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->dh2, b * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->dh2t, b * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->dh1, b * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->h1pre, b * MLP_H * sizeof(float)));
    MLP_HIP(hipMalloc(&M->h2pre, b * MLP_H * sizeof(float)));
    MLP_HIP(hipMalloc(&M->dh1pre, b * MLP_H * sizeof(float)));
    MLP_HIP(hipMalloc(&M->z, b * sizeof(float)));
    MLP_HIP(hipMalloc(&M->dz, b * sizeof(float)));
    MLP_HIP(hipMalloc(&M->yb, b * sizeof(float)));
    MLP_HIP(hipMalloc(&M->idx, b * sizeof(int)));
    M->B = B;
    return ZK_OK;
}

void refresh(zk_mlp *M)
{
    hipLaunchKernelGGL(refresh_kernel, dim3((MLP_H * MLP_IN + 255) / 256), dim3(256), 0, M->stream, M->w, M->w1b,
                       M->w2b, M->w2t, M->w2p);
}

// forward + backward of one batch already gathered in xb/xbt/yb; gradients in M->g
void fwd_bwd(zk_mlp *M, int B)
{
    hipStream_t st = M->stream;
    const int nb = (MLP_H * B + 255) / 256;
    hipLaunchKernelGGL(zero_f_kernel, dim3(1), dim3(64), 0, st, M->loss, 1);
    gemm_nt(st, M->w1b, MLP_IN, M->xb, MLP_IN, M->h1pre, B, MLP_H, B, MLP_IN);
    hipLaunchKernelGGL(act1_kernel, dim3(nb), dim3(256), 0, st, M->h1pre, M->w, M->h1, M->h1t, B);
    gemm_nt(st, M->w2b, MLP_H, M->h1t, MLP_H, M->h2pre, B, MLP_H, B, MLP_H);
    hipLaunchKernelGGL(head_kernel, dim3((B + 255) / 256), dim3(256), 0, st, M->h2pre, M->w, M->yb, M->z, M->dz,
                       M->loss, B);
    hipLaunchKernelGGL(dh2_kernel, dim3(nb), dim3(256), 0, st, M->h2pre, M->w, M->dz, M->dh2, M->dh2t, B);
    // dW2 = dH2 H1^T, dH1raw = W2^T dH2, dW1 = dH1 X
    gemm_nt(st, M->dh2, B, M->h1, B, M->g + zk::P_W2, MLP_H, MLP_H, MLP_H, B);
    gemm_nt(st, M->w2t, MLP_H, M->dh2t, MLP_H, M->dh1pre, B, MLP_H, B, MLP_H);
    hipLaunchKernelGGL(dh1_kernel, dim3(nb), dim3(256), 0, st, M->dh1pre, M->h1pre, M->w, M->dh1, B);
    gemm_nt(st, M->dh1, B, M->xbt, B, M->g + zk::P_W1, MLP_IN, MLP_H, MLP_IN, B);
    hipLaunchKernelGGL(small_grads_kernel, dim3(1), dim3(256), 0, st, M->h2pre, M->w, M->dz, M->dh1pre, M->h1pre,
                       M->g, B);
}

}  // namespace

// ---- hooks used by the DG stage (dg_kernels.hip) ----
namespace zk {
const __bf16 *mlp_w1b(const zk_mlp *M) { return M->w1b; }
const __bf16 *mlp_w2p(const zk_mlp *M) { return M->w2p; }
const float *mlp_w(const zk_mlp *M) { return M->w; }
__bf16 *mlp_X(zk_mlp *M) { return M->X; }
float *mlp_y(zk_mlp *M) { return M->y; }
int *mlp_count(zk_mlp *M) { return M->count; }
int mlp_cap(const zk_mlp *M) { return M->cap; }
}  // namespace zk

extern "C" {

zk_status zk_mlp_create(int device, uint64_t seed, int capacity, zk_mlp **out)
{
    if (!out || capacity < 0) return ZK_ERR_ARG;
    zk_mlp *M = new zk_mlp();
    M->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess) {
        delete M;
        return ZK_ERR_HIP;
    }
    // He-normal initialisation from a seeded generator (splitmix64 + Box-Muller)
    std::vector<float> w(zk::P_N, 0.0f);
    uint64_t s = seed ^ 0x9E3779B97F4A7C15ull;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() {
        const double u1 = next() + 1e-300, u2 = next();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int h = 0; h < MLP_H; ++h)
        for (int k = 0; k < zk::MLP_FEAT; ++k) w[zk::P_W1 + h * MLP_IN + k] = (float)(normal() * std::sqrt(2.0 / zk::MLP_FEAT));
    for (int j = 0; j < MLP_H * MLP_H; ++j) w[zk::P_W2 + j] = (float)(normal() * std::sqrt(2.0 / MLP_H));
    for (int j = 0; j < MLP_H; ++j) w[zk::P_W3 + j] = (float)(normal() * std::sqrt(1.0 / MLP_H));
    MLP_HIP(hipMalloc(&M->w, zk::P_N * sizeof(float)));
    MLP_HIP(hipMalloc(&M->g, zk::P_N * sizeof(float)));
    MLP_HIP(hipMalloc(&M->m1, zk::P_N * sizeof(float)));
    MLP_HIP(hipMalloc(&M->m2, zk::P_N * sizeof(float)));
    MLP_HIP(hipMalloc(&M->w1b, MLP_H * MLP_IN * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->w2b, MLP_H * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->w2p, MLP_H * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->w2t, MLP_H * MLP_H * sizeof(__bf16)));
    MLP_HIP(hipMalloc(&M->loss, sizeof(float)));
    MLP_HIP(hipMalloc(&M->count, sizeof(int)));
    M->cap = capacity;
    if (capacity > 0) {
        MLP_HIP(hipMalloc(&M->X, (size_t)capacity * MLP_IN * sizeof(__bf16)));
        MLP_HIP(hipMalloc(&M->y, (size_t)capacity * sizeof(float)));
    }
    MLP_HIP(hipMemcpy(M->w, w.data(), zk::P_N * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(hipMemset(M->g, 0, zk::P_N * sizeof(float)));
    MLP_HIP(hipMemset(M->m1, 0, zk::P_N * sizeof(float)));
    MLP_HIP(hipMemset(M->m2, 0, zk::P_N * sizeof(float)));
    MLP_HIP(hipMemset(M->count, 0, sizeof(int)));
    refresh(M);
    MLP_HIP(hipStreamSynchronize(M->stream));
    *out = M;
    return ZK_OK;
}

zk_status zk_mlp_destroy(zk_mlp *M)
{
    if (!M) return ZK_OK;
    (void)hipSetDevice(M->device);
    if (M->stream) (void)hipStreamSynchronize(M->stream);
    void *ptrs[] = {M->w, M->g, M->m1, M->m2, M->w1b, M->w2b, M->w2p, M->w2t, M->X, M->y, M->count, M->xb, M->xbt,
                    M->h1, M->h1t, M->dh2, M->dh2t, M->dh1, M->h1pre, M->h2pre, M->z, M->dz, M->yb, M->dh1pre,
                    M->loss, M->idx};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
    return ZK_OK;
}

const char *zk_mlp_last_error(const zk_mlp *M) { return M ? M->err.c_str() : "null mlp"; }

int zk_mlp_param_count(void) { return zk::P_N; }

zk_status zk_mlp_get_params(zk_mlp *M, float *w)
{
    if (!M || !w) return ZK_ERR_ARG;
    MLP_HIP(hipSetDevice(M->device));
    MLP_HIP(hipMemcpyAsync(w, M->w, zk::P_N * sizeof(float), hipMemcpyDeviceToHost, M->stream));
    MLP_HIP(hipStreamSynchronize(M->stream));
    return ZK_OK;
}

zk_status zk_mlp_set_params(zk_mlp *M, const float *w)
{
    if (!M || !w) return ZK_ERR_ARG;
    MLP_HIP(hipSetDevice(M->device));
    MLP_HIP(hipMemcpyAsync(M->w, w, zk::P_N * sizeof(float), hipMemcpyHostToDevice, M->stream));
    refresh(M);
    MLP_HIP(hipStreamSynchronize(M->stream));
    return ZK_OK;
}

// Gradients and loss of host-provided samples (n a multiple of 32): X as
// bf16 bit patterns [n][192], labels y[n].  For gradient checks.
zk_status zk_mlp_grad(zk_mlp *M, const uint16_t *X, const float *y, int n, float *grad, float *loss)
{
    if (!M || !X || !y || !grad || n <= 0 || n % 32) {
        if (M) M->err = "zk_mlp_grad: n must be a positive multiple of 32";
        return ZK_ERR_ARG;
    }
    MLP_HIP(hipSetDevice(M->device));
    zk_status st = ensure_batch(M, n);
    if (st) return st;
    std::vector<__bf16> xt((size_t)n * MLP_IN);
    for (int c = 0; c < n; ++c)
        for (int k = 0; k < MLP_IN; ++k) std::memcpy(&xt[(size_t)k * n + c], &X[(size_t)c * MLP_IN + k], 2);
    MLP_HIP(hipMemcpyAsync(M->xb, X, (size_t)n * MLP_IN * 2, hipMemcpyHostToDevice, M->stream));
    MLP_HIP(hipMemcpyAsync(M->xbt, xt.data(), (size_t)n * MLP_IN * 2, hipMemcpyHostToDevice, M->stream));
    MLP_HIP(hipMemcpyAsync(M->yb, y, (size_t)n * sizeof(float), hipMemcpyHostToDevice, M->stream));
    fwd_bwd(M, n);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(grad, M->g, zk::P_N * sizeof(float), hipMemcpyDeviceToHost, M->stream));
    if (loss) MLP_HIP(hipMemcpyAsync(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost, M->stream));
    MLP_HIP(hipStreamSynchronize(M->stream));
    return ZK_OK;
}

// Scores of host-provided features (n a multiple of 16).
zk_status zk_mlp_infer(zk_mlp *M, const uint16_t *X, int n, float *scores)
{
    if (!M || !X || !scores || n <= 0 || n % 16) return ZK_ERR_ARG;
    MLP_HIP(hipSetDevice(M->device));
    __bf16 *dx = nullptr;
    float *ds = nullptr;
    MLP_HIP(hipMalloc(&dx, (size_t)n * MLP_IN * 2));
    MLP_HIP(hipMalloc(&ds, (size_t)n * sizeof(float)));
    MLP_HIP(hipMemcpyAsync(dx, X, (size_t)n * MLP_IN * 2, hipMemcpyHostToDevice, M->stream));
    hipLaunchKernelGGL(infer_x_kernel, dim3((n / 16 + 3) / 4), dim3(256), 0, M->stream, dx, n, M->w1b, M->w2p, M->w, ds);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(scores, ds, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, M->stream));
    MLP_HIP(hipStreamSynchronize(M->stream));
    (void)hipFree(dx);
    (void)hipFree(ds);
    return ZK_OK;
}

zk_status zk_mlp_dataset(zk_mlp *M, int *count, uint16_t *X, float *y, int max)
{
    if (!M) return ZK_ERR_ARG;
    MLP_HIP(hipSetDevice(M->device));
    int n = 0;
    MLP_HIP(hipMemcpyAsync(&n, M->count, sizeof(int), hipMemcpyDeviceToHost, M->stream));
    MLP_HIP(hipStreamSynchronize(M->stream));
    if (n > M->cap) n = M->cap;
    if (count) *count = n;
    const int k = n < max ? n : max;
    if (X && k > 0) MLP_HIP(hipMemcpy(X, M->X, (size_t)k * MLP_IN * 2, hipMemcpyDeviceToHost));
    if (y && k > 0) MLP_HIP(hipMemcpy(y, M->y, (size_t)k * sizeof(float), hipMemcpyDeviceToHost));
    return ZK_OK;
}

zk_status zk_mlp_clear_dataset(zk_mlp *M)
{
    if (!M) return ZK_ERR_ARG;
    MLP_HIP(hipSetDevice(M->device));
    MLP_HIP(hipMemsetAsync(M->count, 0, sizeof(int), M->stream));
    MLP_HIP(hipStreamSynchronize(M->stream));
    return ZK_OK;
}

// Adam over the device dataset: `epochs` passes in a seeded shuffled order,
// batches of B samples (a multiple of 32).  Per-epoch mean loss to loss_out.
zk_status zk_mlp_train(zk_mlp *M, int epochs, int B, float lr, uint64_t seed, float *loss_out)
{
    if (!M || epochs < 0 || B <= 0 || B % 32) return ZK_ERR_ARG;
    MLP_HIP(hipSetDevice(M->device));
    int n = 0;
    MLP_HIP(hipMemcpy(&n, M->count, sizeof(int), hipMemcpyDeviceToHost));
    if (n > M->cap) n = M->cap;
    if (n < B) {
        M->err = "zk_mlp_train: fewer samples than one batch";
        return ZK_ERR_ARG;
    }
    zk_status st = ensure_batch(M, B);
    if (st) return st;
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = k;
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return s;
    };
    const int nbatch = n / B;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int k = n - 1; k > 0; --k) std::swap(order[k], order[rnd() % (uint64_t)(k + 1)]);
        double ep_loss = 0.0;
        for (int b = 0; b < nbatch; ++b) {
            MLP_HIP(hipMemcpyAsync(M->idx, order.data() + (size_t)b * B, B * sizeof(int), hipMemcpyHostToDevice,
                                   M->stream));
            hipLaunchKernelGGL(gather_kernel, dim3((B * MLP_IN + 255) / 256), dim3(256), 0, M->stream, M->X, M->y,
                               M->idx, M->xb, M->xbt, M->yb, B);
            fwd_bwd(M, B);
            M->adam_t += 1;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f - std::pow(b1, (float)M->adam_t), c2 = 1.0f - std::pow(b2, (float)M->adam_t);
            hipLaunchKernelGGL(adam_kernel, dim3((zk::P_N + 255) / 256), dim3(256), 0, M->stream, M->w, M->g, M->m1,
                               M->m2, zk::P_N, lr, b1, b2, 1e-8f, c1, c2);
            refresh(M);
            float l = 0.0f;
            MLP_HIP(hipMemcpyAsync(&l, M->loss, sizeof(float), hipMemcpyDeviceToHost, M->stream));
            MLP_HIP(hipStreamSynchronize(M->stream));
            ep_loss += l;
        }
        if (loss_out) loss_out[ep] = (float)(ep_loss / nbatch);
    }
    MLP_HIP(hipGetLastError());
    return ZK_OK;
}

}  // extern "C"

// Features and scores of given cells of a DG P=2 handle's current state (tests).
namespace zk {
zk_status mlp_features_of(Handle &H, zk_mlp *M, const int32_t *cells, int n, uint16_t *X, float *scores)
{
    int32_t *dc = nullptr;
    __bf16 *dx = nullptr;
    float *ds = nullptr;
    if (hipMalloc(&dc, (size_t)n * 4) != hipSuccess || hipMalloc(&dx, (size_t)n * MLP_IN * 2) != hipSuccess ||
        hipMalloc(&ds, (size_t)n * 4) != hipSuccess) {
        H.err = "mlp_features_of: allocation failed";
        return ZK_ERR_HIP;
    }
    (void)hipMemcpyAsync(dc, cells, (size_t)n * 4, hipMemcpyHostToDevice, H.stream);
    hipLaunchKernelGGL(features_kernel, dim3((n * MLP_IN + 255) / 256), dim3(256), 0, H.stream, H.q, H.src, H.lay,
                       H.g.nx, dc, n, dx);
    if (M)
        hipLaunchKernelGGL(score_kernel, dim3(((n + 15) / 16 + 3) / 4), dim3(256), 0, H.stream, H.q, H.src, H.lay,
                           H.g.nx, dc, n, M->w1b, M->w2p, M->w, ds);
    if (X) (void)hipMemcpyAsync(X, dx, (size_t)n * MLP_IN * 2, hipMemcpyDeviceToHost, H.stream);
    if (M && scores) (void)hipMemcpyAsync(scores, ds, (size_t)n * 4, hipMemcpyDeviceToHost, H.stream);
    const hipError_t e = hipStreamSynchronize(H.stream);
    (void)hipFree(dc);
    (void)hipFree(dx);
    (void)hipFree(ds);
    if (e != hipSuccess) {
        H.err = std::string("mlp_features_of: ") + hipGetErrorString(e);
        return ZK_ERR_HIP;
    }
    return ZK_OK;
}
}  // namespace zk
