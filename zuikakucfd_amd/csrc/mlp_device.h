// mlp_device.h -- features and fused inference of the troubled-cell MLP
// (device side, shared by the MLP and DG kernel files).
#pragma once

#include <hip/hip_runtime.h>

#include "zk_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace zk {
constexpr int MLP_IN = 192, MLP_FEAT = 180, MLP_H = 64;
// parameter offsets in the FP32 master vector
constexpr int P_W1 = 0, P_B1 = P_W1 + MLP_H * MLP_IN, P_W2 = P_B1 + MLP_H, P_B2 = P_W2 + MLP_H * MLP_H,
              P_W3 = P_B2 + MLP_H, P_B3 = P_W3 + MLP_H, P_N = P_B3 + 1;

__device__ __forceinline__ float sigmoidf(float x) { return 1.0f / (1.0f + __expf(-x)); }
}  // namespace zk

// ---------------------------------------------------------------------------
// Features and fused inference inside the DG (P = 2) stage
// ---------------------------------------------------------------------------
namespace zk {

// Feature k of the cell whose own stage-output modes are Q (planes c*9+m) at
// cell k0: k = n*36 + c*9 + m, n = self, E (j+1), W (j-1), S (i+1), N (i-1).
__device__ __forceinline__ float mlp_feature(const double *__restrict__ Q, const int32_t *__restrict__ src,
                                             const StLay &L, int nx, size_t k0, const double mean[4], const double scale[4], int k)
{
    if (k >= MLP_FEAT) return 0.0f;
    const int n = k / 36, c = (k / 9) % 4, m = k % 9;
    const long off[5] = {0, 1, -1, (long)nx, -(long)nx};
    const size_t cell = (size_t)((long)k0 + off[n]);
    const int32_t e = src[cell];
    double v = Q[L.at(c * 9 + m, (size_t)(e >> kSrcShift))];
    if ((c == 1 && (e & 1)) || (c == 2 && (e & 2))) v = -v;
    if (m == 0) v -= mean[c];
    return (float)(v / scale[c]);
}

__device__ __forceinline__ void mlp_scales(const double *__restrict__ Q, const StLay &L, size_t k, double mean[4],
                                           double scale[4])
{
#pragma unroll
    for (int c = 0; c < 4; ++c) mean[c] = Q[L.at(c * 9, k)];
    const double r = fabs(mean[0]);
#pragma unroll
    for (int c = 0; c < 4; ++c) scale[c] = fabs(mean[c]) + r + 1e-12;
}

// Score sigmoid(MLP(features)) of 16 cells per wave.  cells[t] is the grid
// index of work item t; the wave must be fully populated (padded lanes reuse a
// valid cell).  Returns the score of the lane's cell (column i) in every lane.
__device__ __forceinline__ float mlp_score16(const double *__restrict__ Q, const int32_t *__restrict__ src, const StLay &L, int nx,
                             size_t k0, const __bf16 *__restrict__ w1b, const __bf16 *__restrict__ w2p,
                             const float *__restrict__ w)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    (void)i;
    double mean[4], scale[4];
    mlp_scales(Q, L, k0, mean, scale);
    // layer 1: H1[h][cell] = sum_k W1[h][k] X[cell][k]; B operand = this lane's features
    f4 h1acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) h1acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < MLP_IN / 32; ++s) {
        bf16x8 xv;
#pragma unroll
        for (int e = 0; e < 8; ++e) xv[e] = (__bf16)mlp_feature(Q, src, L, nx, k0, mean, scale, 32 * s + 8 * g + e);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(w1b + (size_t)(16 * t + (lane & 15)) * MLP_IN + 32 * s + 8 * g);
            h1acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, xv, h1acc[t], 0, 0, 0);
        }
    }
    // lane holds H1[16t + 4g + v][cell]; bias, ReLU, bf16 -> layer-2 B operand
    bf16x8 hv[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = 2 * s + e / 4, v = e % 4;
            const int h = 16 * t + 4 * g + v;
            hv[s][e] = (__bf16)fmaxf(h1acc[t][v] + w[P_B1 + h], 0.0f);
        }
    f4 h2acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        h2acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const bf16x8 av = *reinterpret_cast<const bf16x8 *>(w2p + (size_t)(16 * t + (lane & 15)) * MLP_H + 32 * s + 8 * g);
            h2acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, hv[s], h2acc[t], 0, 0, 0);
        }
    }
    // layer 3: partial dot over this lane's 16 hidden units, summed over g
    float part = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = 16 * t + 4 * g + v;
            part += w[P_W3 + j] * fmaxf(h2acc[t][v] + w[P_B2 + j], 0.0f);
        }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return sigmoidf(part + w[P_B3]);
}

}  // namespace zk

