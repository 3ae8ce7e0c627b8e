// fv_kernels.hip -- WENO5 finite-volume stage kernels for gfx950.
//
// One launch = one SSP-RK3 stage over the whole grid:
//   * fv_stage_kernel: one workgroup per 2-D tile of the update range
//     [3,ny-3)x[3,nx-3) (fv_tile.h), tiles mapped XCD-contiguously.
//   * the blocks after the tiles give the cells outside the update range
//     (ghost ring, fixed inflow cells) the same combination with dU = 0.
// The per-face arithmetic follows weno5_reconstruct and hllcFlux of
// src_WENO/fluid.cpp:230-317; the stage combination follows :468-493.
#include "fv_tile.h"
#include "tile_map.h"

namespace zk {

namespace {

// Cells outside [3,ny-3)x[3,nx-3) (ghost ring, fixed inflow cells): the stage
// combination with dU = 0, i.e. the stale-ghost mixture of the reference.
__device__ __forceinline__ void fv_ring_cells(const FvArgs &A, int b)
{
    const int t = b * FV_THREADS + threadIdx.x;
    if (t >= A.nring) return;
    const size_t k = (size_t)A.ring[t];
    const int32_t e = A.src[k];
    const size_t s = (size_t)(e >> kSrcShift);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double un = A.Un[v * A.np + k];
        double uf = A.Uin[v * A.np + s];
        if ((v == 1 && (e & 1)) || (v == 2 && (e & 2))) uf = -uf;
        A.Uout[v * A.np + k] = (A.stage == 0) ? un : A.ca * un + A.cb * (uf + 0.0);
    }
}

// Blocks [0, ntiles) update one tile each (XCD-aware order); the blocks after
// them combine the ghost ring, so a stage is a single launch.
template <int TX, int TY, bool F32>
__global__ __launch_bounds__(FV_THREADS) void fv_stage_kernel(FvArgs A)
{
    __shared__ double s_u[4][TY + 6][TX + 6];
    __shared__ double s_f[4][(TX + 1) * (TY + 1)];
    if (A.S->done) return;
    if ((int)blockIdx.x >= A.ntiles) {
        fv_ring_cells(A, (int)blockIdx.x - A.ntiles);
        return;
    }
    fv_tile_body<TX, TY, F32>(A, xcd_tile(blockIdx.x, A.ntiles), s_u, s_f);
}

// Cell positivity fallback, pass 2 (positivity option on): a grid-strided
// launch over five work items per bad cell pass 1 listed (the cell and the
// good neighbours it owns).  The counters rotate over the three stages: this
// launch clears the next stage's.
template <bool F32>
__global__ __launch_bounds__(256) void fv_fallback_kernel(FvArgs A, unsigned int *any, int slot)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) any[(slot + 1) % 3] = 0u;
    if (A.S->done) return;
    const size_t n = (size_t)any[slot] * 5;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        fv_fallback_item<F32>(A, (size_t)A.blist[t / 5].y, (int)(t % 5));
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
        m = speed_max(m, sqrt(u * u + v * v) + a, finite4(rho, U[np + k], U[2 * np + k], U[3 * np + k]));
    }
    block_max_to(m, out);
}

}  // namespace

void fv_stage(Handle &H, const double *Un, const double *Uin, double *Uout, int stage)
{
    FvArgs A;
    A.Un = Un; A.Uin = Uin; A.Uout = Uout;
    A.src = H.src; A.upd = H.upd; A.S = H.scal;
    A.srcy = H.fv_slip ? H.srcy : nullptr;
    A.ny = H.g.ny; A.nx = H.g.nx; A.np = H.g.np;
    A.da = H.g.da; A.gamma = H.g.gamma;
    A.stage = stage;
    A.ca = stage == 1 ? 0.75 : (1.0 / 3.0);
    A.cb = stage == 1 ? 0.25 : (2.0 / 3.0);
    A.flux = H.cfg.flux;
    A.positivity = H.cfg.positivity;
    A.recon = H.recon;
    const int ux = H.g.nx - 6, uy = H.g.ny - 6;
    A.ring = H.ring;
    A.nring = H.nring;
    A.bad = H.cfg.positivity ? H.bad : nullptr;
    A.any = H.cfg.positivity ? H.any + stage : nullptr;
    A.blist = H.blist;
    A.inst = 0;
    const int nrb = (H.nring + FV_THREADS - 1) / FV_THREADS;
    const bool f32 = H.cfg.precision == ZK_PREC_FP32_BULK;
    // tile shape: 0 = automatic (16x15 for small grids, 32x15 otherwise), 1 = 16x15, 2 = 32x15
    int shape = H.fv_tile;
    if (shape == 0) shape = ((size_t)ux * uy < 400000) ? 1 : 2;
    const bool tiles = ux > 0 && uy > 0;
    if (!tiles && nrb == 0) return;
    ProfScope ps(H, "fv_stage");
    if (shape == 1) {
        A.tiles_x = tiles ? (ux + 15) / 16 : 0;
        A.ntiles = tiles ? A.tiles_x * ((uy + 14) / 15) : 0;
        const dim3 grid(A.ntiles + nrb);
        if (f32) hipLaunchKernelGGL((fv_stage_kernel<16, 15, true>), grid, dim3(FV_THREADS), 0, H.stream, A);
        else hipLaunchKernelGGL((fv_stage_kernel<16, 15, false>), grid, dim3(FV_THREADS), 0, H.stream, A);
    } else {
        A.tiles_x = tiles ? (ux + 31) / 32 : 0;
        A.ntiles = tiles ? A.tiles_x * ((uy + 14) / 15) : 0;
        const dim3 grid(A.ntiles + nrb);
        if (f32) hipLaunchKernelGGL((fv_stage_kernel<32, 15, true>), grid, dim3(FV_THREADS), 0, H.stream, A);
        else hipLaunchKernelGGL((fv_stage_kernel<32, 15, false>), grid, dim3(FV_THREADS), 0, H.stream, A);
    }
    if (A.bad && tiles) {
        ProfScope pf(H, "fv_fallback");
        const dim3 grid(256);  // grid-strided over the (few) bad cells; ends at once when there are none
        if (f32) hipLaunchKernelGGL(fv_fallback_kernel<true>, grid, dim3(256), 0, H.stream, A, H.any, stage);
        else hipLaunchKernelGGL(fv_fallback_kernel<false>, grid, dim3(256), 0, H.stream, A, H.any, stage);
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
