// fv_tile.h -- one WENO5 stage over one TX x TY tile (shared by the
// single-grid and the ensemble kernels).
//
// The tile plus a 3-cell halo is staged in LDS as SoA FP64 with the ghost
// fill applied through the precomputed source map; every x face and then
// every y face of the tile is reconstructed and its flux computed exactly
// once; each cell forms dU = -(dt/dx (Fr-Fl) + dt/dy (Fd-Fu)) and the
// Shu-Osher combination.  Per-face arithmetic: weno5_reconstruct and hllcFlux
// of src_WENO/fluid.cpp:230-317; stage combination: :468-493.
#pragma once

#include "zk_internal.h"
#include "zk_device.h"
#include "fv_face.h"

namespace zk {

constexpr int FV_THREADS = 256;

struct FvArgs {
    const double *Un;
    const double *Uin;
    double *Uout;
    const int32_t *src;
    const uint8_t *upd;
    const StepScalars *S;
    int ny, nx;
    size_t np;
    double da, gamma;
    double ca, cb;  // Uout = ca*Un + cb*(Uf + dU) for stages 2 and 3
    int stage;
    int flux;
    int positivity;
    int tiles_x;
    int ntiles;          // blocks past ntiles combine the ghost ring instead
    const int32_t *ring; // cells outside the update range
    int nring;
};

// Face positivity check in FP64 (an option with no reference counterpart).
__device__ __forceinline__ bool face_positive(const double Q[4], double gamma)
{
    if (!(Q[0] >= 1e-12)) return false;
    const double p = (gamma - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) / Q[0]);
    return p >= 1e-12;
}

// Face sweep of one direction over the tile: every face computed once.
// X: faces between tile cells (r, c-1) and (r, c), r < TY, c <= TX; otherwise
// faces between (r-1, c) and (r, c), r <= TY, c < TX.  Flux planes are stored
// as (F0, F_rho_u, F_rho_v, F3) at s_f[v][face].
template <int TX, int TY, bool F32, bool X>
__device__ __forceinline__ void fv_sweep(const FvArgs &A, double (*s_u)[TY + 6][TX + 6], double (*s_f)[(TX + 1) * (TY + 1)],
                                         int i0, int j0)
{
    constexpr int NF = X ? (TX + 1) * TY : TX * (TY + 1);
    constexpr int W = X ? TX + 1 : TX;
    for (int f = threadIdx.x; f < NF; f += FV_THREADS) {
        const int r = f / W, c = f % W;
        int sr, sc;
        constexpr int dr = X ? 0 : 1, dc = X ? 1 : 0;
        if (X) {
            if (i0 + r >= A.ny - 3 || j0 + c - 1 >= A.nx - 3) continue;
            sr = r + 3; sc = c;
        } else {
            if (j0 + c >= A.nx - 3 || i0 + r - 1 >= A.ny - 3) continue;
            sr = r; sc = c + 3;
        }
        double QL[4], QR[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double w[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = s_u[v][sr + k * dr][sc + k * dc];
            if (F32) {
                // FP32 reconstruction of the deltas to the face's left cell
                const double rf = w[2];
                float dl, drr;
                weno5_pair_f32((float)(w[0] - rf), (float)(w[1] - rf), 0.0f, (float)(w[3] - rf), (float)(w[4] - rf),
                               (float)(w[5] - rf), dl, drr);
                QL[v] = rf + (double)dl;
                QR[v] = rf + (double)drr;
            } else {
                weno5_pair_f64(w[0], w[1], w[2], w[3], w[4], w[5], QL[v], QR[v]);
            }
        }
        if (A.positivity && !(face_positive(QL, A.gamma) && face_positive(QR, A.gamma))) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                QL[v] = s_u[v][sr + 2 * dr][sc + 2 * dc];
                QR[v] = s_u[v][sr + 3 * dr][sc + 3 * dc];
            }
        }
        // y faces: normal momentum is rho*v (component 2); the flux is called
        // with components 1 and 2 swapped and swapped back (fluid.cpp:444-451).
        double L[4], R[4], F[4];
        L[0] = QL[0]; R[0] = QR[0]; L[3] = QL[3]; R[3] = QR[3];
        L[1] = X ? QL[1] : QL[2]; L[2] = X ? QL[2] : QL[1];
        R[1] = X ? QR[1] : QR[2]; R[2] = X ? QR[2] : QR[1];
        if (A.flux == ZK_FLUX_HLL) hll_f64(L, R, A.gamma, F);
        else if (F32) hllc_fv_f32(L, R, A.gamma, F);
        else hllc_fv_fast_f64(L, R, A.gamma, F);
        s_f[0][f] = F[0];
        s_f[1][f] = X ? F[1] : F[2];
        s_f[2][f] = X ? F[2] : F[1];
        s_f[3][f] = F[3];
    }
}

// One stage over a TX x TY tile with 256 threads: stage the tile and its
// 3-cell halo (ghost fill applied) in LDS, sweep the x faces, fold them into
// per-cell accumulators, reuse the flux buffer for the y faces, then form
// dU = -(dt/dx (Fr - Fl) + dt/dy (Fd - Fu)) and the Shu-Osher combination.
template <int TX, int TY, bool F32>
__device__ __forceinline__ void fv_tile_body(const FvArgs &A, int tile, double (*s_u)[TY + 6][TX + 6],
                                             double (*s_f)[(TX + 1) * (TY + 1)])
{
    constexpr int SW = TX + 6, SH = TY + 6;
    constexpr int CPT = (TX * TY + FV_THREADS - 1) / FV_THREADS;  // cells per thread
    const int tid = threadIdx.x;
    const int tx = tile % A.tiles_x, ty = tile / A.tiles_x;
    const int j0 = 3 + tx * TX, i0 = 3 + ty * TY;
    const int ny = A.ny, nx = A.nx;
    const size_t np = A.np;

    for (int t = tid; t < SH * SW; t += FV_THREADS) {
        const int r = t / SW, c = t % SW;
        const int i = i0 - 3 + r, j = j0 - 3 + c;
        if (i < ny && j < nx) {
            const int32_t e = A.src[(size_t)i * nx + j];
            const size_t s = (size_t)(e >> kSrcShift);
            double v0 = A.Uin[s], v1 = A.Uin[np + s], v2 = A.Uin[2 * np + s], v3 = A.Uin[3 * np + s];
            if (e & 1) v1 = -v1;
            if (e & 2) v2 = -v2;
            s_u[0][r][c] = v0; s_u[1][r][c] = v1; s_u[2][r][c] = v2; s_u[3][r][c] = v3;
        }
    }
    __syncthreads();

    const double dt = A.S->dt;
    const double dtdx = dt / A.da, dtdy = dt / A.da;
    double xs[CPT][4];

    fv_sweep<TX, TY, F32, true>(A, s_u, s_f, i0, j0);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int t = tid + q * FV_THREADS;
        const int r = t / TX, c = t % TX;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            xs[q][v] = (t < TX * TY) ? dtdx * (s_f[v][r * (TX + 1) + c + 1] - s_f[v][r * (TX + 1) + c]) : 0.0;
    }
    __syncthreads();
    fv_sweep<TX, TY, F32, false>(A, s_u, s_f, i0, j0);
    __syncthreads();

#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int t = tid + q * FV_THREADS;
        if (t >= TX * TY) break;
        const int r = t / TX, c = t % TX;
        const int i = i0 + r, j = j0 + c;
        if (i >= ny - 3 || j >= nx - 3) continue;
        const size_t k = (size_t)i * nx + j;
        const bool upd = A.upd[k] != 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double dU = 0.0;
            if (upd) dU = -(xs[q][v] + dtdy * (s_f[v][(r + 1) * TX + c] - s_f[v][r * TX + c]));
            const double un = A.Un[v * np + k];
            A.Uout[v * np + k] = (A.stage == 0) ? un + dU : A.ca * un + A.cb * (s_u[v][r + 3][c + 3] + dU);
        }
    }
}


}  // namespace zk
