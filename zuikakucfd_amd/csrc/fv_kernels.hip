// fv_kernels.hip -- WENO5 finite-volume stage kernels for gfx950.
//
// One launch = one SSP-RK3 stage over the whole grid:
//   * fv_stage_kernel: one workgroup per 2-D tile of the update range
//     [3,ny-3)x[3,nx-3) (fv_tile.h), tiles mapped XCD-contiguously.
//   * fv_ring_kernel: the cells outside the update range (ghost ring, fixed
//     inflow cells) get the same combination with dU = 0.
// The per-face arithmetic follows weno5_reconstruct and hllcFlux of
// src_WENO/fluid.cpp:230-317; the stage combination follows :468-493.
#include "fv_tile.h"
#include "tile_map.h"

namespace zk {

namespace {

template <int TX, int TY, bool F32>
__global__ __launch_bounds__(FV_THREADS) void fv_stage_kernel(FvArgs A)
{
    __shared__ double s_u[4][TY + 6][TX + 6];
    __shared__ double s_f[4][(TX + 1) * (TY + 1)];
    if (A.S->done) return;
    fv_tile_body<TX, TY, F32>(A, xcd_tile(blockIdx.x, gridDim.x), s_u, s_f);
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
    if (ux > 0 && uy > 0) {
        ProfScope ps(H, "fv_stage");
        const bool f32 = H.cfg.precision == ZK_PREC_FP32_BULK;
        // tile shape: 0 = automatic (16x15 for small grids, 32x15 otherwise), 1 = 16x15, 2 = 32x15
        int shape = H.fv_tile;
        if (shape == 0) shape = ((size_t)ux * uy < 400000) ? 1 : 2;
        if (shape == 1) {
            A.tiles_x = (ux + 15) / 16;
            const int nb = A.tiles_x * ((uy + 14) / 15);
            if (f32) hipLaunchKernelGGL((fv_stage_kernel<16, 15, true>), dim3(nb), dim3(FV_THREADS), 0, H.stream, A);
            else hipLaunchKernelGGL((fv_stage_kernel<16, 15, false>), dim3(nb), dim3(FV_THREADS), 0, H.stream, A);
        } else {
            A.tiles_x = (ux + 31) / 32;
            const int nb = A.tiles_x * ((uy + 14) / 15);
            if (f32) hipLaunchKernelGGL((fv_stage_kernel<32, 15, true>), dim3(nb), dim3(FV_THREADS), 0, H.stream, A);
            else hipLaunchKernelGGL((fv_stage_kernel<32, 15, false>), dim3(nb), dim3(FV_THREADS), 0, H.stream, A);
        }
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
