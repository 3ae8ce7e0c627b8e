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
    const double *Un;    // U^n; the first stage (stage 0) is called with Uin == Un
    const double *Uin;
    double *Uout;
    const int32_t *src;
    const int32_t *srcy; // slip walls on: the ghost map of the y faces (column mirrors), else null
    const uint8_t *upd;
    const StepScalars *S;
    int ny, nx;
    size_t np;
    double da, gamma;
    double ca, cb;  // Uout = ca*Un + cb*(Uf + dU) for stages 2 and 3
    int stage;
    int flux;
    int positivity;
    int recon;           // face reconstruction: 0 WENO5, 1..3 MUSCL (minmod, van Leer, superbee)
    int tiles_x;
    int ntiles;          // blocks past ntiles combine the ghost ring instead
    const int32_t *ring; // cells outside the update range
    int nring;
    uint8_t *bad;        // cell positivity fallback (positivity on): flags of the stage, else null
    unsigned int *any;   // count of bad cells of this stage
    int2 *blist;         // the bad cells of the stage: (instance, cell)
    int inst;            // instance index written to blist (0 for a single grid)
};

// A stage result is bad when rho or p is not above 1e-12 (or not finite):
// rho > 1e-12 and 2 rho E - |m|^2 > 2 rho 1e-12 / (gamma - 1), which is
// p > 1e-12 multiplied through by 2 rho / (gamma - 1) > 0.
__device__ __forceinline__ bool fv_bad_state(const double U[4], double gamma)
{
    const double rho = U[0];
    const double q = 2.0 * rho * U[3] - (U[1] * U[1] + U[2] * U[2]);
    return !(rho > 1e-12 && q > div_nz(2.0 * rho * 1e-12, gamma - 1.0));
}

// Face positivity check in FP64 (an option with no reference counterpart).
__device__ __forceinline__ bool face_positive(const double Q[4], double gamma)
{
    if (!(Q[0] >= 1e-12)) return false;
    const double p = (gamma - 1.0) * (Q[3] - div_nz(0.5 * (Q[1] * Q[1] + Q[2] * Q[2]), Q[0]));  // Q[0] >= 1e-12
    return p >= 1e-12;
}

// Numerical flux of one face from the 6 cells across it (get(v, k), k = 0..5,
// the face lies between k = 2 and k = 3): WENO5 reconstruction of both face
// states (FP64, or FP32 on deltas to the face's left cell), the optional face
// positivity check, HLLC/HLL; y faces call the flux with components 1 and 2
// swapped and swap the result back (fluid.cpp:444-451).  F is returned as
// (F0, F_rho_u, F_rho_v, F3).
template <bool F32, bool X, typename Get>
__device__ __forceinline__ void fv_face_flux(const FvArgs &A, Get get, double F[4])
{
    double QL[4], QR[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double w[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) w[k] = get(v, k);
        if (A.recon != 0) {  // uniform over the grid
            muscl_f64(w[1], w[2], w[3], w[4], A.recon - 1, QL[v], QR[v]);
        } else if (F32) {
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
            QL[v] = get(v, 2);
            QR[v] = get(v, 3);
        }
    }
    double L[4], R[4], G[4];
    L[0] = QL[0]; R[0] = QR[0]; L[3] = QL[3]; R[3] = QR[3];
    L[1] = X ? QL[1] : QL[2]; L[2] = X ? QL[2] : QL[1];
    R[1] = X ? QR[1] : QR[2]; R[2] = X ? QR[2] : QR[1];
    if (A.flux == ZK_FLUX_HLL) hll_f64(L, R, A.gamma, G);
    else if (F32) hllc_fv_f32(L, R, A.gamma, G);
    else hllc_fv_fast_f64(L, R, A.gamma, G);
    F[0] = G[0];
    F[1] = X ? G[1] : G[2];
    F[2] = X ? G[2] : G[1];
    F[3] = G[3];
}

// First-order flux of a face (the two cell values, FP64), as fv_face_flux.
template <bool X>
__device__ __forceinline__ void fv_face_flux_fo(const FvArgs &A, const double a[4], const double b[4], double F[4])
{
    double L[4], R[4], G[4];
    L[0] = a[0]; R[0] = b[0]; L[3] = a[3]; R[3] = b[3];
    L[1] = X ? a[1] : a[2]; L[2] = X ? a[2] : a[1];
    R[1] = X ? b[1] : b[2]; R[2] = X ? b[2] : b[1];
    if (A.flux == ZK_FLUX_HLL) hll_f64(L, R, A.gamma, G);
    else hllc_fv_fast_f64(L, R, A.gamma, G);
    F[0] = G[0];
    F[1] = X ? G[1] : G[2];
    F[2] = X ? G[2] : G[1];
    F[3] = G[3];
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
        double F[4];
        fv_face_flux<F32, X>(A, [&](int v, int k) { return s_u[v][sr + k * dr][sc + k * dc]; }, F);
#pragma unroll
        for (int v = 0; v < 4; ++v) s_f[v][f] = F[v];
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
    if (A.srcy) {
        // slip walls: a wall cell near updated cells along both its row and
        // its column holds the row mirror; the y faces take the column mirror
        for (int t = tid; t < SH * SW; t += FV_THREADS) {
            const int r = t / SW, c = t % SW;
            const int i = i0 - 3 + r, j = j0 - 3 + c;
            if (i < ny && j < nx) {
                const size_t kk = (size_t)i * nx + j;
                const int32_t e = A.srcy[kk];
                if (e != A.src[kk]) {
                    const size_t s = (size_t)(e >> kSrcShift);
                    double v0 = A.Uin[s], v1 = A.Uin[np + s], v2 = A.Uin[2 * np + s], v3 = A.Uin[3 * np + s];
                    if (e & 1) v1 = -v1;
                    if (e & 2) v2 = -v2;
                    s_u[0][r][c] = v0; s_u[1][r][c] = v1; s_u[2][r][c] = v2; s_u[3][r][c] = v3;
                }
            }
        }
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
        double out[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double dU = 0.0;
            if (upd) dU = -(xs[q][v] + dtdy * (s_f[v][(r + 1) * TX + c] - s_f[v][r * TX + c]));
            // the first stage runs with Uin == Un, and an updated cell's ghost-map
            // entry is itself: its U^n is the staged value, not read again
            const double un = (A.stage == 0 && upd) ? s_u[v][r + 3][c + 3] : A.Un[v * np + k];
            double self = s_u[v][r + 3][c + 3];
            if (A.srcy && !upd) {
                // a wall cell inside the update range may hold its column
                // mirror now; its stored mixture uses the ghost map's value
                const int32_t e = A.src[k];
                self = A.Uin[v * np + (size_t)(e >> kSrcShift)];
                if ((v == 1 && (e & 1)) || (v == 2 && (e & 2))) self = -self;
            }
            out[v] = (A.stage == 0) ? un + dU : A.ca * un + A.cb * (self + dU);
        }
        if (A.bad) {
            // cell positivity fallback, pass 1: a bad cell is left unwritten
            // (the stage may update in place) and recomputed by fv_fallback_cell
            const bool bad = upd && fv_bad_state(out, A.gamma);
            A.bad[k] = (uint8_t)bad;
            if (bad) {
                const unsigned int slot = atomicAdd(A.any, 1u);
                A.blist[slot] = int2{A.inst, (int)k};
                continue;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) A.Uout[v * np + k] = out[v];
    }
}

// Pass 2 work item d of bad cell k: d = 4 is the bad cell itself, d = 0..3
// its good updated neighbour N, W, E, S if k is that neighbour's
// lowest-index bad neighbour (so that every affected cell is handled by
// exactly one item, in a fixed way).  The items are independent.
template <bool F32>
__device__ __forceinline__ void fv_fallback_item(const FvArgs &A, size_t k, int d);

// Cell positivity fallback, pass 2, for updated cell (i, j): a bad cell is
// recomputed with first-order fluxes on its four faces; a good cell next to
// bad cells replaces the high-order flux of each shared face by the
// first-order one (the high-order flux recomputed exactly as pass 1 did).
// Every face touching a bad cell is then first order on both of its sides.
template <bool F32>
__device__ __forceinline__ void fv_fallback_cell(const FvArgs &A, int i, int j)
{
    const int nx = A.nx;
    const size_t np = A.np, k = (size_t)i * nx + j;
    if (!A.upd[k]) return;
    const bool b0 = A.bad[k], bE = A.bad[k + 1], bW = A.bad[k - 1], bS = A.bad[k + nx], bN = A.bad[k - nx];
    if (!(b0 || bE || bW || bS || bN)) return;
    // the y faces read the column-mirror map where slip walls are on
    const int32_t *mapy = A.srcy ? A.srcy : A.src;
    auto Um = [&](const int32_t *map, size_t kk, int v) {
        const int32_t e = map[kk];
        double x = A.Uin[v * np + (size_t)(e >> kSrcShift)];
        if ((v == 1 && (e & 1)) || (v == 2 && (e & 2))) x = -x;
        return x;
    };
    auto U = [&](size_t kk, int v) { return Um(A.src, kk, v); };
    auto Uy = [&](size_t kk, int v) { return Um(mapy, kk, v); };
    auto cell = [&](size_t kk, double c[4]) {
#pragma unroll
        for (int v = 0; v < 4; ++v) c[v] = U(kk, v);
    };
    auto celly = [&](size_t kk, double c[4]) {
#pragma unroll
        for (int v = 0; v < 4; ++v) c[v] = Uy(kk, v);
    };
    const double dt = A.S->dt, dtdx = dt / A.da, dtdy = dt / A.da;
    // face (a | b) along x at row i, columns jl | jl+1, or along y at column j, rows il | il+1
    auto fo_x = [&](int jl, double F[4]) {
        double a[4], b[4];
        cell((size_t)i * nx + jl, a);
        cell((size_t)i * nx + jl + 1, b);
        fv_face_flux_fo<true>(A, a, b, F);
    };
    auto fo_y = [&](int il, double F[4]) {
        double a[4], b[4];
        celly((size_t)il * nx + j, a);
        celly((size_t)(il + 1) * nx + j, b);
        fv_face_flux_fo<false>(A, a, b, F);
    };
    auto hi_x = [&](int jl, double F[4]) {
        fv_face_flux<F32, true>(A, [&](int v, int t) { return U((size_t)i * nx + (jl - 2 + t), v); }, F);
    };
    auto hi_y = [&](int il, double F[4]) {
        fv_face_flux<F32, false>(A, [&](int v, int t) { return Uy((size_t)(il - 2 + t) * nx + j, v); }, F);
    };
    if (b0) {
        double FE[4], FW[4], GS[4], GN[4];
        fo_x(j, FE);
        fo_x(j - 1, FW);
        fo_y(i, GS);
        fo_y(i - 1, GN);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double dU = -(dtdx * (FE[v] - FW[v]) + dtdy * (GS[v] - GN[v]));
            const double un = A.Un[v * np + k];
            A.Uout[v * np + k] = (A.stage == 0) ? un + dU : A.ca * un + A.cb * (U(k, v) + dU);
        }
        return;
    }
    double d[4] = {0.0, 0.0, 0.0, 0.0}, lo[4], hi[4];
    if (bE) {
        fo_x(j, lo); hi_x(j, hi);
#pragma unroll
        for (int v = 0; v < 4; ++v) d[v] -= dtdx * (lo[v] - hi[v]);
    }
    if (bW) {
        fo_x(j - 1, lo); hi_x(j - 1, hi);
#pragma unroll
        for (int v = 0; v < 4; ++v) d[v] += dtdx * (lo[v] - hi[v]);
    }
    if (bS) {
        fo_y(i, lo); hi_y(i, hi);
#pragma unroll
        for (int v = 0; v < 4; ++v) d[v] -= dtdy * (lo[v] - hi[v]);
    }
    if (bN) {
        fo_y(i - 1, lo); hi_y(i - 1, hi);
#pragma unroll
        for (int v = 0; v < 4; ++v) d[v] += dtdy * (lo[v] - hi[v]);
    }
    const double w = (A.stage == 0) ? 1.0 : A.cb;
#pragma unroll
    for (int v = 0; v < 4; ++v) A.Uout[v * np + k] += w * d[v];
}


}  // namespace zk

namespace zk {
template <bool F32>
__device__ __forceinline__ void fv_fallback_item(const FvArgs &A, size_t k, int d)
{
    const int nx = A.nx;
    if (d == 4) {
        fv_fallback_cell<F32>(A, (int)(k / nx), (int)(k % nx));
        return;
    }
    const long off[4] = {-(long)nx, -1, 1, (long)nx};
    const size_t n = (size_t)((long)k + off[d]);
    const int ni = (int)(n / nx), nj = (int)(n % nx);
    if (ni < 3 || ni >= A.ny - 3 || nj < 3 || nj >= nx - 3) return;  // outside the update range
    if (A.bad[n] || !A.upd[n]) return;
    // lowest-index bad neighbour of n
    size_t first = k;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const size_t m = (size_t)((long)n + off[e]);
        if (A.bad[m] && m < first) first = m;
    }
    if (first == k) fv_fallback_cell<F32>(A, ni, nj);
}
}  // namespace zk
