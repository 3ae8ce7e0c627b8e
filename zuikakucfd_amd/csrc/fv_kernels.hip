// fv_kernels.hip -- WENO5 finite-volume stage kernels for gfx950.
//
// One launch = one SSP-RK3 stage over the whole grid:
//   * fv_stage_kernel: 2-D tiles of the update range [3,ny-3)x[3,nx-3).  The
//     tile plus a 3-cell halo is staged in LDS as SoA FP64 (the ghost fill is
//     applied while staging, through the precomputed source map), every x and
//     y face of the tile is reconstructed and its HLLC flux computed exactly
//     once, then each cell forms dU = -(dt/dx (Fr-Fl) + dt/dy (Fd-Fu)) and the
//     Shu-Osher combination.
//   * fv_ring_kernel: the cells outside the update range (ghost ring, fixed
//     inflow cells) get the same combination with dU = 0.
// The per-face arithmetic follows weno5_reconstruct and hllcFlux of
// src_WENO/fluid.cpp:230-317; the stage combination follows :468-493.
#include "zk_internal.h"
#include "zk_device.h"
#include "fv_face.h"

namespace zk {

namespace {

constexpr int FV_TX = 32;     // tile width (cells along x)
constexpr int FV_TY = 16;     // tile height (cells along y)
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
};

// Face positivity check in FP64 (an option with no reference counterpart).
__device__ __forceinline__ bool face_positive(const double Q[4], double gamma)
{
    if (!(Q[0] >= 1e-12)) return false;
    const double p = (gamma - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) / Q[0]);
    return p >= 1e-12;
}

template <int TX, int TY, bool F32>
__global__ __launch_bounds__(FV_THREADS) void fv_stage_kernel(FvArgs A)
{
    constexpr int SW = TX + 6, SH = TY + 6;
    __shared__ double s_u[4][SH][SW];
    __shared__ double s_fx[4][TY][TX + 1];
    __shared__ double s_fy[4][TY + 1][TX];

    if (A.S->done) return;
    const int tid = threadIdx.x;
    const int bid = blockIdx.x;
    const int tx = bid % A.tiles_x, ty = bid / A.tiles_x;
    const int j0 = 3 + tx * TX, i0 = 3 + ty * TY;
    const int ny = A.ny, nx = A.nx;
    const size_t np = A.np;

    // ---- stage the tile + 3-cell halo with the ghost fill applied ----
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

    // ---- faces: x faces first, then y faces, each exactly once ----
    constexpr int NXF = (TX + 1) * TY, NYF = TX * (TY + 1);
    for (int f = tid; f < NXF + NYF; f += FV_THREADS) {
        const bool isx = f < NXF;
        int r, c, sr, sc, stride_r, stride_c;
        if (isx) {
            r = f / (TX + 1); c = f % (TX + 1);
            // face between tile cells (r, c-1) and (r, c); stencil cols c-3..c+2
            if (i0 + r >= ny - 3 || j0 + c - 1 >= nx - 3) continue;
            sr = r + 3; sc = c; stride_r = 0; stride_c = 1;
        } else {
            const int g = f - NXF;
            r = g / TX; c = g % TX;
            if (j0 + c >= nx - 3 || i0 + r - 1 >= ny - 3) continue;
            sr = r; sc = c + 3; stride_r = 1; stride_c = 0;
        }
        double QL[4], QR[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double w[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = s_u[v][sr + k * stride_r][sc + k * stride_c];
            if (F32) {
                // FP32 reconstruction of the deltas to the face's left cell
                const double r = w[2];
                float dl, dr;
                weno5_pair_f32((float)(w[0] - r), (float)(w[1] - r), 0.0f, (float)(w[3] - r), (float)(w[4] - r),
                               (float)(w[5] - r), dl, dr);
                QL[v] = r + (double)dl;
                QR[v] = r + (double)dr;
            } else {
                weno5_pair_f64(w[0], w[1], w[2], w[3], w[4], w[5], QL[v], QR[v]);
            }
        }
        if (A.positivity && !(face_positive(QL, A.gamma) && face_positive(QR, A.gamma))) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                QL[v] = s_u[v][sr + 2 * stride_r][sc + 2 * stride_c];
                QR[v] = s_u[v][sr + 3 * stride_r][sc + 3 * stride_c];
            }
        }
        // y faces: normal momentum is rho*v (component 2); the flux is called
        // with components 1 and 2 swapped and swapped back (fluid.cpp:444-451).
        double L[4], R[4], F[4];
        L[0] = QL[0]; R[0] = QR[0]; L[3] = QL[3]; R[3] = QR[3];
        L[1] = isx ? QL[1] : QL[2]; L[2] = isx ? QL[2] : QL[1];
        R[1] = isx ? QR[1] : QR[2]; R[2] = isx ? QR[2] : QR[1];
        if (A.flux == ZK_FLUX_HLL) hll_f64(L, R, A.gamma, F);
        else if (F32) hllc_fv_f32(L, R, A.gamma, F);
        else hllc_fv_fast_f64(L, R, A.gamma, F);
        const double F1 = isx ? F[1] : F[2];
        const double F2 = isx ? F[2] : F[1];
        if (isx) {
            s_fx[0][r][c] = F[0]; s_fx[1][r][c] = F1; s_fx[2][r][c] = F2; s_fx[3][r][c] = F[3];
        } else {
            s_fy[0][r][c] = F[0]; s_fy[1][r][c] = F1; s_fy[2][r][c] = F2; s_fy[3][r][c] = F[3];
        }
    }
    __syncthreads();

    // ---- divergence + Shu-Osher combination ----
    const double dt = A.S->dt;
    const double dtdx = dt / A.da, dtdy = dt / A.da;
    for (int t = tid; t < TX * TY; t += FV_THREADS) {
        const int r = t / TX, c = t % TX;
        const int i = i0 + r, j = j0 + c;
        if (i >= ny - 3 || j >= nx - 3) continue;
        const size_t k = (size_t)i * nx + j;
        const bool upd = A.upd[k] != 0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double dU = 0.0;
            if (upd) dU = -(dtdx * (s_fx[v][r][c + 1] - s_fx[v][r][c]) + dtdy * (s_fy[v][r + 1][c] - s_fy[v][r][c]));
            const double un = A.Un[v * np + k];
            double out;
            if (A.stage == 0) out = un + dU;
            else out = A.ca * un + A.cb * (s_u[v][r + 3][c + 3] + dU);
            A.Uout[v * np + k] = out;
        }
    }
}

// Cells outside [3,ny-3)x[3,nx-3): combination with dU = 0.
__global__ void fv_ring_kernel(const double *__restrict__ Un, const double *__restrict__ Uin,
                               double *__restrict__ Uout, const int32_t *__restrict__ src,
                               const int32_t *__restrict__ ring, int nring, size_t np, int nplanes,
                               double ca, double cb, int stage, int neg_c1, int neg_c2,
                               const StepScalars *__restrict__ S)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nring || S->done) return;
    const size_t k = (size_t)ring[t];
    const int32_t e = src[k];
    const size_t s = (size_t)(e >> kSrcShift);
    const int per_var = nplanes / 4;
    for (int p = 0; p < nplanes; ++p) {
        const int var = p / per_var;
        const double un = Un[p * np + k];
        double uf = Uin[p * np + s];
        if (var == 1 && (e & neg_c1)) uf = -uf;
        if (var == 2 && (e & neg_c2)) uf = -uf;
        Uout[p * np + k] = (stage == 0) ? un : ca * un + cb * (uf + 0.0);
    }
}

// max(|u| + c) over every cell (ghosts included), no floors:
// computeMaxSpeed, src_WENO/fluid.cpp:539-553.
__global__ __launch_bounds__(256) void fv_max_speed_kernel(const double *__restrict__ U, size_t np,
                                                           double gamma,
                                                           unsigned long long *__restrict__ out)
{
    double m = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < np;
         k += (size_t)gridDim.x * blockDim.x) {
        const double rho = U[k];
        const double u = U[np + k] / rho, v = U[2 * np + k] / rho;
        const double p = (gamma - 1.0) * (U[3 * np + k] - 0.5 * rho * (u * u + v * v));
        const double a = sqrt(gamma * p / rho);
        const double spd = sqrt(u * u + v * v) + a;
        if (spd > m) m = spd;
    }
    block_max_to(m, out);
}

}  // namespace

void fv_stage(Handle &H, const double *Un, const double *Uin, double *Uout, int stage)
{
    FvArgs A;
    A.Un = Un; A.Uin = Uin; A.Uout = Uout;
    A.src = H.src; A.upd = H.upd; A.S = H.scal;
    A.ny = H.g.ny; A.nx = H.g.nx; A.np = H.g.np;
    A.da = H.g.da; A.gamma = H.g.gamma;
    A.stage = stage;
    A.ca = stage == 1 ? 0.75 : (1.0 / 3.0);
    A.cb = stage == 1 ? 0.25 : (2.0 / 3.0);
    A.flux = H.cfg.flux;
    A.positivity = H.cfg.positivity;
    const int ux = H.g.nx - 6, uy = H.g.ny - 6;
    A.tiles_x = (ux + FV_TX - 1) / FV_TX;
    const int tiles_y = (uy + FV_TY - 1) / FV_TY;
    if (ux > 0 && uy > 0) {
        ProfScope ps(H, "fv_stage");
        if (H.cfg.precision == ZK_PREC_FP32_BULK)
            hipLaunchKernelGGL((fv_stage_kernel<FV_TX, FV_TY, true>), dim3(A.tiles_x * tiles_y), dim3(FV_THREADS),
                               0, H.stream, A);
        else
            hipLaunchKernelGGL((fv_stage_kernel<FV_TX, FV_TY, false>), dim3(A.tiles_x * tiles_y), dim3(FV_THREADS),
                               0, H.stream, A);
    }
    if (H.nring > 0) {
        ProfScope ps(H, "fv_ring");
        hipLaunchKernelGGL(fv_ring_kernel, dim3((H.nring + 255) / 256), dim3(256), 0, H.stream, Un, Uin, Uout,
                           H.src, H.ring, H.nring, H.g.np, 4, A.ca, A.cb, stage, 1, 2, H.scal);
    }
}

void fv_max_speed(Handle &H, const double *U)
{
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, H.stream, H.lam_bits);
    int blocks = (int)((H.g.np + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(fv_max_speed_kernel, dim3(blocks), dim3(256), 0, H.stream, U, H.g.np, H.g.gamma,
                       H.lam_bits);
}

}  // namespace zk
