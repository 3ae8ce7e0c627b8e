// dg_kernels.hip -- modal DG (P = 1..3) stage kernels for gfx950.
//
// A stage combines the volume integral (src_DG/fluid.cpp:698-718), the four
// face terms (:731-796), the invMass(m) * 0.5/h scale (:800-804) and the
// Shu-Osher combination (:1022-1053), then the limiter (:824-992, or the TVB
// limiter of this project).  Two kernel structures compute it:
//   * fused (default without the Persson limiter): dg_fused_kernel
//     marches rows of 16-column runs on FP64 matrix cores; every face flux is
//     computed once, the state is read about once per stage, and the kernel
//     certifies cells whose positivity scaling is a no-op (zsok).  In its
//     workgroup form the matrix-core tables sit in LDS and each row arrives
//     by global-to-LDS loads issued while the previous row is computed;
//   * split: dg_face_kernel<P, X> computes each face's projected flux g[c][p]
//     once (the reference evaluates every face twice, from each cell; both see
//     the same inputs), then dg_cell_kernel<P> (VALU, also the Persson path)
//     or dg_cell_mfma3_kernel (P = 3) forms the update.
// The TVB limiter needs the neighbours' new means, so it runs after the stage:
// a per-thread screen lists the cells that are limited or not certified, and a
// matrix-core kernel limits and applies the positivity scaling to those.
// The learned indicator (P = 2) scores only the TVB candidate cells.
#include <atomic>
#include <mutex>

#include "zk_internal.h"
#include "zk_device.h"
#include "mlp_device.h"
#include "tile_map.h"

namespace zk {
// learned indicator hooks (mlp_kernels.hip)
const __bf16 *mlp_w1b(const zk_mlp *M);
const __bf16 *mlp_w2p(const zk_mlp *M);
const float *mlp_w(const zk_mlp *M);
__bf16 *mlp_X(zk_mlp *M);
float *mlp_y(zk_mlp *M);
int *mlp_count(zk_mlp *M);
int mlp_cap(const zk_mlp *M);

namespace {

constexpr int DG_MAXNQ = 4;
constexpr int DG_MAXNM = 16;

// Row and column of update-range cell t (row-major over the nxu columns) in
// 32-bit arithmetic: a grid holds fewer than 2^29 cells (zk_load checks it),
// and a 64-bit division would cost a software routine per thread.
__device__ __forceinline__ void interior_rc(size_t t, int nxu, int &ii, int &jj)
{
    const uint32_t tt = (uint32_t)t, q = tt / (uint32_t)nxu;
    ii = (int)q;
    jj = (int)(tt - q * (uint32_t)nxu);
}
__device__ __forceinline__ size_t interior_cell(size_t t, int nx, int nxu)
{
    int ii, jj;
    interior_rc(t, nxu, ii, jj);
    return (size_t)(ii + 3) * nx + (jj + 3);
}

// Basis tables, filled on the host in plain double arithmetic with the same
// products the reference forms in phi2/dphi2 (src_DG/fluid.cpp:67-76).
struct DgTables {
    double w[DG_MAXNQ];                              // GL weights
    double wl[DG_MAXNQ][DG_MAXNQ];                   // w_q * L_p(s_q)  [p][q]
    double wdl[DG_MAXNQ][DG_MAXNQ];                  // w_q * L'_p(s_q) [p][q]
    double lq[DG_MAXNQ][DG_MAXNQ];                   // L_p(s_q)        [p][q]
    double sgn[DG_MAXNQ];                            // L_p(-1) = (-1)^p
    double tr_x[2][DG_MAXNQ][DG_MAXNM];              // phi_m(-1 / +1, eta_q)
    double tr_y[2][DG_MAXNQ][DG_MAXNM];              // phi_m(xi_q, -1 / +1)
    double phi[DG_MAXNQ][DG_MAXNQ][DG_MAXNM];        // phi_m(xi_a, eta_b)
    double dxi[DG_MAXNQ][DG_MAXNQ][DG_MAXNM];        // d phi_m / d xi
    double det[DG_MAXNQ][DG_MAXNQ][DG_MAXNM];        // d phi_m / d eta
    double ww[DG_MAXNQ][DG_MAXNQ];                   // w_a * w_b
    double corner[4][DG_MAXNM];                      // phi_m at (+-1, +-1) for the CFL speed
    double zs[4][4][DG_MAXNM];                       // truncated basis() at the 4x4 ZS grid
    double zx[20][DG_MAXNM];                         // positivity option: full basis at the face points and
                                                     // corners (4 NQ + 4 of them; the rest repeat the first)
    double invm[DG_MAXNM];                           // (2px+1)(2py+1)
};

// The extra points of the positivity option (no reference counterpart): the
// face quadrature points (-1, eta_q), (+1, eta_q), (xi_q, -1), (xi_q, +1) that
// the fluxes read and the corners (+-1, +-1) that the CFL rule reads.
inline int zs_ext_points(int P, const double *gp, double xs[20], double ys[20])
{
    const int NQ = P + 1;
    int n = 0;
    for (int q = 0; q < NQ; ++q) { xs[n] = -1.0; ys[n++] = gp[q]; }
    for (int q = 0; q < NQ; ++q) { xs[n] = +1.0; ys[n++] = gp[q]; }
    for (int q = 0; q < NQ; ++q) { xs[n] = gp[q]; ys[n++] = -1.0; }
    for (int q = 0; q < NQ; ++q) { xs[n] = gp[q]; ys[n++] = +1.0; }
    for (int c = 0; c < 4; ++c) { xs[n] = (c & 1) ? 1.0 : -1.0; ys[n++] = (c & 2) ? 1.0 : -1.0; }
    for (int e = n; e < 20; ++e) { xs[e] = xs[0]; ys[e] = ys[0]; }
    return n;
}

__constant__ DgTables c_tab[3];  // index P-1

double legP_h(int n, double x)
{
    switch (n) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 0.5 * (3.0 * x * x - 1.0);
    case 3: return 0.5 * (5.0 * x * x * x - 3.0 * x);
    default: return 0.0;
    }
}
double legD_h(int n, double x)
{
    switch (n) {
    case 0: return 0.0;
    case 1: return 1.0;
    case 2: return 3.0 * x;
    case 3: return 0.5 * (15.0 * x * x - 3.0);
    default: return 0.0;
    }
}
double zsP_h(int p, double x)
{
    if (p == 0) return 1.0;
    if (p == 1) return x;
    if (p == 2) return 0.5 * (3.0 * x * x - 1.0);
    return 0.0;
}

// Legendre values, derivatives and Gauss-Legendre nodes/weights as constant
// expressions, so that device code with unrolled indices folds them into
// instruction operands (the fused kernel's 4x4-block form).  Same literals
// and closed forms as legP_h / legD_h and the tables below.
__host__ __device__ constexpr double leg_c(int n, double x)
{
    return n == 0 ? 1.0 : n == 1 ? x : n == 2 ? 0.5 * (3.0 * x * x - 1.0) : n == 3 ? 0.5 * (5.0 * x * x * x - 3.0 * x) : 0.0;
}
__host__ __device__ constexpr double legd_c(int n, double x)
{
    return n == 0 ? 0.0 : n == 1 ? 1.0 : n == 2 ? 3.0 * x : n == 3 ? 0.5 * (15.0 * x * x - 3.0) : 0.0;
}
__host__ __device__ constexpr double gl_x(int nq, int q)
{
    return nq == 2   ? (q == 0 ? -0.5773502691896257 : 0.5773502691896257)
           : nq == 3 ? (q == 0 ? -0.7745966692414834 : q == 1 ? 0.0 : 0.7745966692414834)
           : nq == 4 ? (q == 0   ? -0.8611363115940526
                        : q == 1 ? -0.3399810435848563
                        : q == 2 ? 0.3399810435848563
                                 : 0.8611363115940526)
                     : 0.0;
}
__host__ __device__ constexpr double gl_w(int nq, int q)
{
    return nq == 2   ? 1.0
           : nq == 3 ? (q == 1 ? 0.8888888888888888 : 0.5555555555555556)
           : nq == 4 ? ((q == 0 || q == 3) ? 0.3478548451374538 : 0.6521451548625461)
                     : 0.0;
}

// Gauss-Legendre literals of src_DG/fluid.cpp:94-111.
const double GLP_H[5][4] = {{0, 0, 0, 0},
                            {0.0, 0, 0, 0},
                            {-0.5773502691896257, 0.5773502691896257, 0, 0},
                            {-0.7745966692414834, 0.0, 0.7745966692414834, 0},
                            {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526}};
const double GLW_H[5][4] = {{0, 0, 0, 0},
                            {2.0, 0, 0, 0},
                            {1.0, 1.0, 0, 0},
                            {0.5555555555555556, 0.8888888888888888, 0.5555555555555556, 0},
                            {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538}};

DgTables make_tables(int P)
{
    DgTables T{};
    const int NQ = P + 1, NM = NQ * NQ;
    const double *gp = GLP_H[NQ], *gw = GLW_H[NQ];
    const double zsp[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
    for (int q = 0; q < NQ; ++q) {
        T.w[q] = gw[q];
        for (int p = 0; p < NQ; ++p) {
            T.wl[p][q] = gw[q] * legP_h(p, gp[q]);
            T.wdl[p][q] = gw[q] * legD_h(p, gp[q]);
            T.lq[p][q] = legP_h(p, gp[q]);
        }
    }
    for (int p = 0; p < NQ; ++p) T.sgn[p] = legP_h(p, -1.0);
    const double XI[4] = {-1.0, +1.0, -1.0, +1.0}, ETA[4] = {-1.0, -1.0, +1.0, +1.0};
    for (int m = 0; m < NM; ++m) {
        const int px = m / NQ, py = m % NQ;
        T.invm[m] = (double)((2 * px + 1) * (2 * py + 1));
        for (int q = 0; q < NQ; ++q) {
            T.tr_x[0][q][m] = legP_h(px, -1.0) * legP_h(py, gp[q]);
            T.tr_x[1][q][m] = legP_h(px, +1.0) * legP_h(py, gp[q]);
            T.tr_y[0][q][m] = legP_h(px, gp[q]) * legP_h(py, -1.0);
            T.tr_y[1][q][m] = legP_h(px, gp[q]) * legP_h(py, +1.0);
        }
        for (int a = 0; a < NQ; ++a)
            for (int b = 0; b < NQ; ++b) {
                T.phi[a][b][m] = legP_h(px, gp[a]) * legP_h(py, gp[b]);
                T.dxi[a][b][m] = legD_h(px, gp[a]) * legP_h(py, gp[b]);
                T.det[a][b][m] = legP_h(px, gp[a]) * legD_h(py, gp[b]);
            }
        for (int k = 0; k < 4; ++k) T.corner[k][m] = legP_h(px, XI[k]) * legP_h(py, ETA[k]);
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) T.zs[a][b][m] = zsP_h(px, zsp[a]) * zsP_h(py, zsp[b]);
        double xs[20], ys[20];
        zs_ext_points(P, gp, xs, ys);
        for (int e = 0; e < 20; ++e) T.zx[e][m] = legP_h(px, xs[e]) * legP_h(py, ys[e]);
    }
    for (int a = 0; a < NQ; ++a)
        for (int b = 0; b < NQ; ++b) T.ww[a][b] = gw[a] * gw[b];
    return T;
}

// Euler flux pair with rho and p floored at 1e-3 (eulerFluxX/Y,
// src_DG/fluid.cpp:123-143); both share the primitive recovery.
__device__ __forceinline__ void euler_fg(const double U[4], double gam, double F[4], double G[4])
{
    constexpr double eps = 1e-3;
    const double rho = smax(U[0], eps), ir = rcp_nz(rho);
    const double u = U[1] * ir, v = U[2] * ir;
    const double p = smax((gam - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v)), eps);
    // rho u and rho v are the momenta (to rounding, also where rho is floored)
    const double ruv = U[1] * v, h = U[3] + p;
    F[0] = U[1]; F[1] = U[1] * u + p; F[2] = ruv; F[3] = h * u;
    G[0] = U[2]; G[1] = ruv; G[2] = U[2] * v + p; G[3] = h * v;
}

template <int P>
__device__ __forceinline__ void load_cell(const double *__restrict__ Q, const int32_t *__restrict__ src,
                                          const StLay &L, size_t k, double A[4][(P + 1) * (P + 1)])
{
    constexpr int NM = (P + 1) * (P + 1);
    const int32_t e = src[k];
    const size_t s = L.cell((size_t)(e >> kSrcShift));
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int m = 0; m < NM; ++m) A[c][m] = Q[(size_t)(c * NM + m) * L.ps + s];
    if (e & 1)
#pragma unroll
        for (int m = 0; m < NM; ++m) A[1][m] = -A[1][m];
    if (e & 2)
#pragma unroll
        for (int m = 0; m < NM; ++m) A[2][m] = -A[2][m];
}

struct DgArgs {
    const double *Qn;
    const double *Qin;
    double *Qout;
    const int32_t *src;
    const uint8_t *upd;
    const int32_t *bct;
    const StepScalars *S;
    double *gx, *gy;
    double *means;    // stage means (4 planes) for the TVB limiter, or null
    // fused kernel under the TVB limiters, or null: bit 0 set where the
    // positivity scaling is certified to be a no-op, bit 1 where the cell is
    // a TVB candidate (a linear mode above Mh2, or NaN; at P >= 3 a variable
    // with sum_q |a(q,0)| above Mh2 or some |a(0,q)| above Mh2 / 4)
    uint8_t *zsok;
    double Mh2;
    int ny, nx, nxu, nyu;
    size_t np, nfx, nfy;
    double da, gamma;
    double ca, cb;
    int stage, flux, limiter;
    int zs_ext;       // positivity option: the scaling also at the face points and corners
    int row16;        // rows may be staged with 16-byte loads (nx even: every row segment 16-byte aligned)
    StLay L;          // layout of Qn, Qin and Qout (np above counts cells: per-cell arrays such as the means)
};

// Plane `pl` of a state array whose planes lie `stride` bytes apart.  With a
// 32-bit stride (arrays below 4 GiB) the offset is one 32-bit scalar multiply.
template <typename Ofs, typename T>
__device__ __forceinline__ T *plane_at(T *base, int pl, Ofs stride)
{
    using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return reinterpret_cast<T *>(reinterpret_cast<B *>(base) + (size_t)((Ofs)pl * stride));
}

template <int P, bool X>
__device__ __forceinline__ void load_trace(const double *__restrict__ Q, const int32_t *__restrict__ src, const StLay &L,
                                           size_t k, bool plus, double C[4][P + 1])
{
    constexpr int NQ = P + 1, NM = NQ * NQ;
    const int32_t e = src[k];
    const size_t s = L.cell((size_t)(e >> kSrcShift));
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int t = 0; t < NQ; ++t) C[c][t] = 0.0;
#pragma unroll
        for (int px = 0; px < NQ; ++px)
#pragma unroll
            for (int py = 0; py < NQ; ++py) {
                const double a = Q[(size_t)(c * NM + px * NQ + py) * L.ps + s];
                const int t = X ? py : px;          // direction along the face
                const int o = X ? px : py;          // direction across the face
                if (plus || (o & 1) == 0) C[c][t] += a;
                else C[c][t] -= a;
            }
    }
    if (e & 1)
#pragma unroll
        for (int t = 0; t < NQ; ++t) C[1][t] = -C[1][t];
    if (e & 2)
#pragma unroll
        for (int t = 0; t < NQ; ++t) C[2][t] = -C[2][t];
}

// One thread per face.  X: faces between (i, j-1) and (i, j), j in [3, nx-3];
// otherwise faces between (i-1, j) and (i, j), i in [3, ny-3].
template <int P, bool X>
__global__ __launch_bounds__(128) void dg_face_kernel(DgArgs A)
{
    constexpr int NQ = P + 1;
    const DgTables &T = c_tab[P - 1];
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nf = X ? A.nfx : A.nfy;
    if (f >= nf || A.S->done) return;
    int i, j;
    size_t kl, kr;
    if (X) {
        const int w = A.nxu + 1;
        i = 3 + (int)(f / w); j = 3 + (int)(f % w);
        kl = (size_t)i * A.nx + (j - 1); kr = (size_t)i * A.nx + j;
    } else {
        const int w = A.nxu;
        i = 3 + (int)(f / w); j = 3 + (int)(f % w);
        kl = (size_t)(i - 1) * A.nx + j; kr = (size_t)i * A.nx + j;
    }
    // Trace coefficients along the face.  On an x face the left cell's trace
    // at xi = +1 is sum_py (sum_px a[px][py]) L_py(eta) and the right cell's
    // at xi = -1 weights a[px][py] by (-1)^px; on a y face px and py swap.
    double CL[4][NQ], CR[4][NQ];
    load_trace<P, X>(A.Qin, A.src, A.L, kl, true, CL);
    load_trace<P, X>(A.Qin, A.src, A.L, kr, false, CR);

    double g[4][NQ];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int p = 0; p < NQ; ++p) g[c][p] = 0.0;

#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double UL[4] = {0.0, 0.0, 0.0, 0.0}, UR[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < NQ; ++t)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                UL[c] += CL[c][t] * T.lq[t][q];
                UR[c] += CR[c][t] * T.lq[t][q];
            }
        double Fs[4];
        if (X) {
            if (A.flux == ZK_FLUX_HLL) hll_f64(UL, UR, A.gamma, Fs);
            else hllc_dg_fast_f64(UL, UR, A.gamma, Fs);
        } else {
            const double a[4] = {UL[0], UL[2], UL[1], UL[3]}, b[4] = {UR[0], UR[2], UR[1], UR[3]};
            double Gs[4];
            if (A.flux == ZK_FLUX_HLL) hll_f64(a, b, A.gamma, Gs);
            else hllc_dg_fast_f64(a, b, A.gamma, Gs);
            Fs[0] = Gs[0]; Fs[1] = Gs[2]; Fs[2] = Gs[1]; Fs[3] = Gs[3];
        }
#pragma unroll
        for (int p = 0; p < NQ; ++p)
#pragma unroll
            for (int c = 0; c < 4; ++c) g[c][p] += T.wl[p][q] * Fs[c];
    }
    double *G = X ? A.gx : A.gy;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int p = 0; p < NQ; ++p) G[(size_t)(c * NQ + p) * nf + f] = g[c][p];
}

// Positivity scaling of one cell (src_DG/fluid.cpp:897-991).
template <int P>
__device__ __forceinline__ bool zs_cell(double Q[4][(P + 1) * (P + 1)], double gamma, bool ext = false)
{
    constexpr int NX = 4 * (P + 1) + 4;  // extra points of the positivity option
    bool changed = false;
    constexpr int NM = (P + 1) * (P + 1);
    const DgTables &T = c_tab[P - 1];
    constexpr double RHO_MIN = 1e-5, P_MIN = 1e-5, EPS = 1e-5;
    if (Q[0][0] < RHO_MIN) changed = true;
    Q[0][0] = smax(Q[0][0], RHO_MIN);
    const double rho0 = Q[0][0], mx0 = Q[1][0], my0 = Q[2][0];
    double E0 = Q[3][0];
    const double p0 = (gamma - 1.0) * (E0 - 0.5 * (mx0 * mx0 + my0 * my0) / rho0);
    if (p0 < P_MIN) {
        E0 = P_MIN / (gamma - 1.0) + 0.5 * (mx0 * mx0 + my0 * my0) / rho0;
        Q[3][0] = E0;
        changed = true;
    }
    double theta_rho = 1.0;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
            double rho = 0.0;
#pragma unroll
            for (int m = 0; m < NM; m++) rho += Q[0][m] * T.zs[a][b][m];
            if (rho < RHO_MIN) {
                const double t = (rho0 - RHO_MIN) / (rho0 - rho + EPS);
                theta_rho = smin(theta_rho, t);
            }
        }
    if (ext)
        for (int e = 0; e < NX; e++) {
            double rho = 0.0;
#pragma unroll
            for (int m = 0; m < NM; m++) rho += Q[0][m] * T.zx[e][m];
            if (rho < RHO_MIN) theta_rho = smin(theta_rho, (rho0 - RHO_MIN) / (rho0 - rho + EPS));
        }
    if (theta_rho < 1.0) {
        changed = true;
#pragma unroll
        for (int m = 1; m < NM; m++)
#pragma unroll
            for (int c = 0; c < 4; ++c) Q[c][m] *= theta_rho;
    }

    double theta_p = 1.0;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
            double rho = 0.0, mx = 0.0, my = 0.0, E = 0.0;
#pragma unroll
            for (int m = 0; m < NM; m++) {
                const double phi = T.zs[a][b][m];
                rho += Q[0][m] * phi;
                mx += Q[1][m] * phi;
                my += Q[2][m] * phi;
                E += Q[3][m] * phi;
            }
            if (rho <= RHO_MIN) continue;
            const double u = mx / rho, v = my / rho;
            const double p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
            if (p < P_MIN) {
                const double p_avg = (gamma - 1.0) * (Q[3][0] - 0.5 * (Q[1][0] * Q[1][0] + Q[2][0] * Q[2][0]) / Q[0][0]);
                const double t = (p_avg - P_MIN) / (p_avg - p + EPS);
                theta_p = smin(theta_p, t);
            }
        }
    if (ext)
        for (int e = 0; e < NX; e++) {
            double rho = 0.0, mx = 0.0, my = 0.0, E = 0.0;
#pragma unroll
            for (int m = 0; m < NM; m++) {
                const double phi = T.zx[e][m];
                rho += Q[0][m] * phi;
                mx += Q[1][m] * phi;
                my += Q[2][m] * phi;
                E += Q[3][m] * phi;
            }
            if (rho <= RHO_MIN) continue;
            const double u = mx / rho, v = my / rho;
            const double p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
            if (p < P_MIN) {
                const double p_avg = (gamma - 1.0) * (Q[3][0] - 0.5 * (Q[1][0] * Q[1][0] + Q[2][0] * Q[2][0]) / Q[0][0]);
                theta_p = smin(theta_p, (p_avg - P_MIN) / (p_avg - p + EPS));
            }
        }
    if (theta_p < 1.0) {
        changed = true;
#pragma unroll
        for (int m = 1; m < NM; m++)
#pragma unroll
            for (int c = 0; c < 4; ++c) Q[c][m] *= theta_p;
    }
    return changed;
}

// Persson modal-decay filter of one cell (src_DG/fluid.cpp:841-887).
template <int P>
__device__ __forceinline__ void persson_cell(double Q[4][(P + 1) * (P + 1)])
{
    constexpr double kappa = 1.0, EPS = 1e-5;
    // s0 = -4 log10(P) for P < 3, -3 log10(P) for P >= 3, evaluated as the
    // reference does, with the device log10.
    const double s0 = (P >= 3) ? -3.0 * log10((double)P) : -4.0 * log10((double)P);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        double Etot = 0.0, Ep = 0.0;
#pragma unroll
        for (int px = 0; px <= P; px++)
#pragma unroll
            for (int py = 0; py <= P; py++) {
                const double a = Q[c][px * (P + 1) + py];
                Etot += a * a;
                if (px == P || py == P) Ep += a * a;
            }
        const double s = log10(Ep / (Etot + EPS));
        if (s > s0) {
            double eps;
            if (s > s0 + kappa) eps = 1.0;
            else {
                const double r = (s - s0) / kappa;
                eps = 0.5 * (1.0 + sin(M_PI * (r - 0.5)));
            }
#pragma unroll
            for (int px = 0; px <= P; px++)
#pragma unroll
                for (int py = 0; py <= P; py++)
                    if (px == P || py == P) Q[c][px * (P + 1) + py] *= (1.0 - eps);
        }
    }
}

template <int P>
__global__ __launch_bounds__(128) void dg_cell_kernel(DgArgs A)
{
    constexpr int NQ = P + 1, NM = NQ * NQ;
    const DgTables &T = c_tab[P - 1];
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)A.nxu * A.nyu || A.S->done) return;
    int ii, jj;
    interior_rc(t, A.nxu, ii, jj);
    const int i = ii + 3, j = jj + 3;
    const size_t k = (size_t)i * A.nx + j;
    const size_t np = A.np;
    const size_t sk = A.L.cell(k);  // the cell in the state arrays
    const bool upd = A.upd[k] != 0;

    double Q[4][NM];   // filled input of the cell, then its output
    load_cell<P>(A.Qin, A.src, A.L, k, Q);
    double rhs[4][NM];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int m = 0; m < NM; ++m) rhs[c][m] = 0.0;

    if (upd) {
        // volume term, sum-factorised: for each xi_a collapse px, evaluate the
        // NQ points of that column, project along eta, then along xi.
        //   rhs[c][px,py] += sum_a w_a L'_px(xi_a) SF_a[c][py] + w_a L_px(xi_a) SG_a[c][py]
        //   SF_a[c][py] = sum_b w_b L_py(eta_b) F_c(a,b),  SG_a = sum_b w_b L'_py(eta_b) G_c(a,b)
#pragma unroll
        for (int a = 0; a < NQ; ++a) {
            double Y[4][NQ];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int py = 0; py < NQ; ++py) {
                    double acc = 0.0;
#pragma unroll
                    for (int px = 0; px < NQ; ++px) acc += Q[c][px * NQ + py] * T.lq[px][a];
                    Y[c][py] = acc;
                }
            double SF[4][NQ], SG[4][NQ];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int py = 0; py < NQ; ++py) SF[c][py] = SG[c][py] = 0.0;
#pragma unroll
            for (int b = 0; b < NQ; ++b) {
                double U[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    double acc = 0.0;
#pragma unroll
                    for (int py = 0; py < NQ; ++py) acc += Y[c][py] * T.lq[py][b];
                    U[c] = acc;
                }
                double F[4], G[4];
                euler_fg(U, A.gamma, F, G);
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int py = 0; py < NQ; ++py) {
                        SF[c][py] += T.wl[py][b] * F[c];
                        SG[c][py] += T.wdl[py][b] * G[c];
                    }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int px = 0; px < NQ; ++px)
#pragma unroll
                    for (int py = 0; py < NQ; ++py)
                        rhs[c][px * NQ + py] += T.wdl[px][a] * SF[c][py] + T.wl[px][a] * SG[c][py];
        }
        // face terms: right (x face j+1), left (x face j), down (y face i+1), up (y face i)
        const size_t fr = (size_t)ii * (A.nxu + 1) + (jj + 1), fl = fr - 1;
        const size_t fd = (size_t)(ii + 1) * A.nxu + jj, fu = (size_t)ii * A.nxu + jj;
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int p = 0; p < NQ; ++p) {
                const double gR = A.gx[(size_t)(c * NQ + p) * A.nfx + fr];
                const double gL = A.gx[(size_t)(c * NQ + p) * A.nfx + fl];
                const double gD = A.gy[(size_t)(c * NQ + p) * A.nfy + fd];
                const double gU = A.gy[(size_t)(c * NQ + p) * A.nfy + fu];
#pragma unroll
                for (int o = 0; o < NQ; ++o) {
                    // x faces project onto py (= p), y faces onto px (= p)
                    rhs[c][o * NQ + p] += T.sgn[o] * gL - gR;
                    rhs[c][p * NQ + o] += T.sgn[o] * gU - gD;
                }
            }
    }
    const double dt = A.S->dt;
    const double inv2h = 0.5 / A.da;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            const double dU = rhs[c][m] * (T.invm[m] * inv2h);
            const double qn = A.Qn[(size_t)(c * NM + m) * A.L.ps + sk];
            Q[c][m] = (A.stage == 0) ? qn + dt * dU : A.ca * qn + A.cb * (Q[c][m] + dt * dU);
        }
    if (A.limiter == ZK_LIMITER_PERSSON_ZS) {
        if (A.bct[k] == 0) persson_cell<P>(Q);
        zs_cell<P>(Q, A.gamma, A.zs_ext != 0);
    } else if (A.means) {
#pragma unroll
        for (int c = 0; c < 4; ++c) A.means[(size_t)c * np + k] = Q[c][0];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int m = 0; m < NM; ++m) A.Qout[(size_t)(c * NM + m) * A.L.ps + sk] = Q[c][m];
}

__device__ __forceinline__ double minmod3(double a, double b, double c)
{
    if (a > 0.0 && b > 0.0 && c > 0.0) return fmin(a, fmin(b, c));
    if (a < 0.0 && b < 0.0 && c < 0.0) return fmax(a, fmax(b, c));
    return 0.0;
}

// Mean of variable c of cell n after the ghost fill: the stage mean of its
// source cell, negated for the normal momentum of slip ghosts.
__device__ __forceinline__ double filled_mean(const double *__restrict__ M, const int32_t *__restrict__ src,
                                              size_t np, size_t n, int c)
{
    const int32_t e = src[n];
    double v = M[(size_t)c * np + (size_t)(e >> kSrcShift)];
    if ((c == 1 && (e & 1)) || (c == 2 && (e & 2))) v = -v;
    return v;
}

// TVB-modified minmod: a when |a| <= M h^2, else minmod(a, b, c) (NaN goes
// to the minmod, which returns 0).
__device__ __forceinline__ double tvb_mod(double a, double b, double c, double Mh2)
{
    return fabs(a) <= Mh2 ? a : minmod3(a, b, c);
}

// Face-average deviations of one variable along one axis from its modes
// q[0..P-1] of degree 1..P on that axis (degree 0 on the other): the + face
// average minus the mean, sum q, and the mean minus the - face average,
// sum (-1)^(deg+1) q, each summed in increasing degree.
template <int P>
__device__ __forceinline__ void face_devs(const double q[P], double &fp, double &fm)
{
    fp = 0.0;
    fm = 0.0;
#pragma unroll
    for (int d = 1; d <= P; ++d) {
        fp += q[d - 1];
        if (d & 1)
            fm += q[d - 1];
        else
            fm -= q[d - 1];
    }
}

// TVB test of variable c of cell k (Cockburn & Shu 2001, modified minmod):
// the linear modes ax (x) and ay (y) against the forward and backward
// differences of the filled neighbour means; with faces (P >= 3) the variable
// is also troubled when the one-dimensional test (Cockburn & Shu 1989)
// changes one of the face-average deviations xp, xm, yp, ym.  Leaves the
// limited linear modes in axl/ayl and returns whether the variable is
// troubled.  Neighbour means are read only where a value is above M h^2.
// tvb_axes with the four neighbour means given (east j+1, west j-1, south
// i+1, north i-1): the same arithmetic, for callers that load the means first.
__device__ __forceinline__ bool tvb_axes_nb(double ub, double ax, double ay, bool faces, double xp, double xm,
                                            double yp, double ym, double mE, double mW, double mS, double mN,
                                            double Mh2, double &axl, double &ayl)
{
    bool trouble = false;
    axl = ax;
    ayl = ay;
    if (!(fabs(ax) <= Mh2) || (faces && (!(fabs(xp) <= Mh2) || !(fabs(xm) <= Mh2)))) {
        const double dp = mE - ub, dm = ub - mW;
        axl = tvb_mod(ax, dp, dm, Mh2);
        if (faces) trouble = tvb_mod(xp, dp, dm, Mh2) != xp || tvb_mod(xm, dp, dm, Mh2) != xm;
    }
    if (!(fabs(ay) <= Mh2) || (faces && (!(fabs(yp) <= Mh2) || !(fabs(ym) <= Mh2)))) {
        const double dp = mS - ub, dm = ub - mN;
        ayl = tvb_mod(ay, dp, dm, Mh2);
        if (faces) trouble = trouble || tvb_mod(yp, dp, dm, Mh2) != yp || tvb_mod(ym, dp, dm, Mh2) != ym;
    }
    return trouble || axl != ax || ayl != ay;
}

__device__ __forceinline__ bool tvb_axes(double ub, double ax, double ay, bool faces, double xp, double xm,
                                         double yp, double ym, const double *__restrict__ M,
                                         const int32_t *__restrict__ src, size_t np, size_t k, int nx, int c,
                                         double Mh2, double &axl, double &ayl)
{
    bool trouble = false;
    axl = ax;
    ayl = ay;
    if (!(fabs(ax) <= Mh2) || (faces && (!(fabs(xp) <= Mh2) || !(fabs(xm) <= Mh2)))) {
        const double dp = filled_mean(M, src, np, k + 1, c) - ub, dm = ub - filled_mean(M, src, np, k - 1, c);
        axl = tvb_mod(ax, dp, dm, Mh2);
        if (faces) trouble = tvb_mod(xp, dp, dm, Mh2) != xp || tvb_mod(xm, dp, dm, Mh2) != xm;
    }
    if (!(fabs(ay) <= Mh2) || (faces && (!(fabs(yp) <= Mh2) || !(fabs(ym) <= Mh2)))) {
        const double dp = filled_mean(M, src, np, k + nx, c) - ub, dm = ub - filled_mean(M, src, np, k - nx, c);
        ayl = tvb_mod(ay, dp, dm, Mh2);
        if (faces) trouble = trouble || tvb_mod(yp, dp, dm, Mh2) != yp || tvb_mod(ym, dp, dm, Mh2) != ym;
    }
    return trouble || axl != ax || ayl != ay;
}

// tvb_axes on variable c of a cell held as C[c][px * (P + 1) + py].
template <int P>
__device__ __forceinline__ bool tvb_var(const double Cc[(P + 1) * (P + 1)], const double *__restrict__ M,
                                        const int32_t *__restrict__ src, size_t np, size_t k, int nx, int c,
                                        double Mh2, double &axl, double &ayl, bool lin = false)
{
    constexpr int NQ = P + 1;
    const double ax = Cc[NQ], ay = Cc[1];
    double xp = ax, xm = ax, yp = ay, ym = ay;
    if constexpr (P >= 3) {
        double qx[P], qy[P];
#pragma unroll
        for (int d = 1; d <= P; ++d) {
            qx[d - 1] = Cc[d * NQ];
            qy[d - 1] = Cc[d];
        }
        face_devs<P>(qx, xp, xm);
        face_devs<P>(qy, yp, ym);
    }
    return tvb_axes(Cc[0], ax, ay, P >= 3 && !lin, xp, xm, yp, ym, M, src, np, k, nx, c, Mh2, axl, ayl);
}

// Component-wise TVB limiter (Cockburn & Shu 2001, modified minmod with
// constant M; at P >= 3 also the face-average test of tvb_var) then the
// positivity scaling, in place on the stage output.
// Neighbour means are the combined stage means after the ghost fill; the cell
// is written back only when the limiter or the scaling changed it.
template <int P>
__global__ __launch_bounds__(128) void dg_tvb_kernel(double *__restrict__ Q, const double *__restrict__ M,
                                                     const int32_t *__restrict__ src,
                                                     const int32_t *__restrict__ bct, uint8_t *__restrict__ flags,
                                                     int ny, int nx, size_t np, StLay L, double Mh2, double gamma,
                                                     const StepScalars *__restrict__ S, bool lin, bool ext)
{
    constexpr int NQ = P + 1, NM = NQ * NQ;
    const int nxu = nx - 6, nyu = ny - 6;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nxu * nyu || S->done) return;
    const size_t k = interior_cell(t, nx, nxu);
    const size_t sk = L.cell(k);
    double C[4][NM];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int m = 0; m < NM; ++m) C[c][m] = Q[(size_t)(c * NM + m) * L.ps + sk];
    int flag = 0;
    if (bct[k] == 0) {
        constexpr int mxm = NQ, mym = 1;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double axl, ayl;
            if (tvb_var<P>(C[c], M, src, np, k, nx, c, Mh2, axl, ayl, lin)) {
#pragma unroll
                for (int m = 1; m < NM; ++m) C[c][m] = 0.0;
                C[c][mxm] = axl;
                C[c][mym] = ayl;
                flag = 1;
            }
        }
    }
    const bool changed = zs_cell<P>(C, gamma, ext);
    flags[k] = (uint8_t)flag;
    if (flag || changed) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int m = 0; m < NM; ++m) Q[(size_t)(c * NM + m) * L.ps + sk] = C[c][m];
    }
}

// max(|u| + c) at the four corners of every bctype-0 cell, floors 1e-12
// (computeMaxSpeedDG_omp, src_DG/solver.cpp:26-64).
template <int P>
__global__ __launch_bounds__(256) void dg_max_speed_kernel(const double *__restrict__ Q,
                                                           const int32_t *__restrict__ bct, size_t np, StLay L,
                                                           double gam, unsigned long long *__restrict__ out)
{
    constexpr int NM = (P + 1) * (P + 1);
    const DgTables &T = c_tab[P - 1];
    constexpr double eps = 1e-12;
    double mx = 0.0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (size_t)gridDim.x * blockDim.x) {
        if (bct[k] != 0) continue;
        double A[4][NM];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int m = 0; m < NM; ++m) A[c][m] = Q[L.at(c * NM + m, k)];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double U[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < NM; ++m)
#pragma unroll
                for (int c = 0; c < 4; ++c) U[c] += A[c][m] * T.corner[q][m];
            const double rho = smax(U[0], eps);
            const double u = U[1] / rho, v = U[2] / rho;
            const double p = smax((gam - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v)), eps);
            const double a = sqrt(gam * p / rho);
            mx = speed_max(mx, sqrt(u * u + v * v) + a, finite4(U[0], U[1], U[2], U[3]));
        }
    }
    block_max_to(mx, out);
}

// Cells outside the update range: the stage combination with dU = 0, one
// thread per (plane, ring cell) so that the launch fills the device.  It also
// clears the counts of the limiter lists filled after it (z0, z1, or null),
// which saves two single-thread launches per stage.
struct RingArgs {
    const double *Qn, *Qin;
    double *Qout;
    const int32_t *src, *ring;
    int nring, nm;
    size_t np;  // cells per plane of M
    StLay L;    // layout of Qn, Qin, Qout
    double ca, cb;
    int stage;
    double *M;  // means of the limiter (mode 0 of every variable), or null
};

// Work item t of the ghost-ring combination: (plane, ring cell) pair t.
__device__ __forceinline__ void ring_item(const RingArgs &R, size_t t)
{
    if (t >= (size_t)R.nring * 4 * R.nm) return;
    const uint32_t tt = (uint32_t)t, p = tt / (uint32_t)R.nring;  // fewer than 2^32 (plane, ring cell) pairs
    const size_t k = (size_t)R.ring[tt - p * (uint32_t)R.nring];
    const int32_t e = R.src[k];
    const size_t s = (size_t)(e >> kSrcShift);
    const int var = (int)p / R.nm;
    const double qn = R.Qn[R.L.at((int)p, k)];
    double qf = R.Qin[R.L.at((int)p, s)];
    if ((var == 1 && (e & 1)) || (var == 2 && (e & 2))) qf = -qf;
    const double out = (R.stage == 0) ? qn + 0.0 : R.ca * qn + R.cb * (qf + 0.0);
    R.Qout[R.L.at((int)p, k)] = out;
    if (R.M && (int)p % R.nm == 0) R.M[(size_t)var * R.np + k] = out;
}

__global__ void dg_ring_kernel(RingArgs R, const StepScalars *__restrict__ S, int32_t *__restrict__ z0,
                               int32_t *__restrict__ z1)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        if (z0) *z0 = 0;
        if (z1) *z1 = 0;
    }
    if (S->done) return;
    ring_item(R, t);
}


// ---------------------------------------------------------------------------
// P = 3 cell kernel on FP64 matrix cores.
//
// One wave updates 16 cells.  With v_mfma_f64_16x16x4f64 (A: lane l holds
// A[l%16][l/16]; B: B[l/16][l%16]; D: lane l holds D[l/16 + 4v][l%16], v=0..3)
// the modal-to-point evaluation U_c(16 points x 16 cells) = Phi(16x16) Q_c(16x16)
// and the projection rhs_c = W_F F_c + W_G G_c are 16x16x16 products.  Lane
// group g = l/16 of column j = l%16 holds, for cell j, the modes (px = v, py = g)
// of Q and rhs and the points (xi_a = v, eta_b = g) of U, F and G, v = 0..3, so
// the flux at a point, the face terms and the Shu-Osher combination need no
// data movement between lanes.  Points are numbered p = 4a + b, modes m = 4px + py.
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));

struct MfmaTabs3 {
    double phiA[4][64];  // K-step s: Phi[p = l%16][m = 4s + l/16]
    double wfA[4][64];   // K-step s: w_a w_b L'_px(xi_a) L_py(eta_b) at m = l%16, p = 4s + l/16
    double wgA[4][64];   // K-step s: w_a w_b L_px(xi_a) L'_py(eta_b)
};
__constant__ MfmaTabs3 c_mf3;

MfmaTabs3 make_mfma_tables3()
{
    MfmaTabs3 M{};
    const double *gp = GLP_H[4], *gw = GLW_H[4];
    for (int s = 0; s < 4; ++s)
        for (int l = 0; l < 64; ++l) {
            const int i = l % 16, k = l / 16;
            {   // evaluation: point i = 4a + b, mode 4s + k = (px = s, py = k)
                const int a = i / 4, b = i % 4, px = s, py = k;
                M.phiA[s][l] = legP_h(px, gp[a]) * legP_h(py, gp[b]);
            }
            {   // projection: mode i = (px = i/4, py = i%4), point 4s + k = (a = s, b = k)
                const int px = i / 4, py = i % 4, a = s, b = k;
                const double ww = gw[a] * gw[b];
                M.wfA[s][l] = ww * legD_h(px, gp[a]) * legP_h(py, gp[b]);
                M.wgA[s][l] = ww * legP_h(px, gp[a]) * legD_h(py, gp[b]);
            }
        }
    return M;
}

template <int MINW>
__global__ __launch_bounds__(256, MINW) void dg_cell_mfma3_kernel(DgArgs A)
{
    constexpr int NQ = 4, NM = 16;
    const DgTables &T = c_tab[2];
    if (A.S->done) return;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, col = lane & 15;
    const size_t ncell = (size_t)A.nxu * A.nyu;
    const size_t t0 = ((size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16;
    if (t0 >= ncell) return;  // whole wave past the end
    const size_t t = t0 + col;
    const bool live = t < ncell;
    const size_t tc = live ? t : ncell - 1;   // padded lanes compute on a valid cell, store nothing
    int ii, jj;
    interior_rc(tc, A.nxu, ii, jj);
    const size_t k = (size_t)(ii + 3) * A.nx + (jj + 3);
    const size_t np = A.np;
    const bool upd = A.upd[k] != 0;

    // this lane's modes (px = v, py = g) of the filled input cell
    const int32_t e = A.src[k];
    const size_t sk = A.L.cell((size_t)(e >> kSrcShift));
    double Qr[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) Qr[c][v] = A.Qin[(size_t)(c * NM + 4 * v + g) * A.L.ps + sk];
    if (e & 1)
#pragma unroll
        for (int v = 0; v < 4; ++v) Qr[1][v] = -Qr[1][v];
    if (e & 2)
#pragma unroll
        for (int v = 0; v < 4; ++v) Qr[2][v] = -Qr[2][v];

    // point values U_c at (xi_a = v, eta_b = g)
    d4 U[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(c_mf3.phiA[s][lane], Qr[c][s], acc, 0, 0, 0);
        U[c] = acc;
    }
    // fluxes point by point, projected as soon as they exist
    d4 R[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) R[c] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double Up[4] = {U[0][v], U[1][v], U[2][v], U[3][v]};
        double F[4], G[4];
        euler_fg(Up, A.gamma, F, G);
        const double wf = c_mf3.wfA[v][lane], wg = c_mf3.wgA[v][lane];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            R[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(wf, F[c], R[c], 0, 0, 0);
            R[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(wg, G[c], R[c], 0, 0, 0);
        }
    }
    // face terms: x faces project onto py (= g), y faces onto px (= v)
    const size_t fr = (size_t)ii * (A.nxu + 1) + (jj + 1), fl = fr - 1;
    const size_t fd = (size_t)(ii + 1) * A.nxu + jj, fu = (size_t)ii * A.nxu + jj;
    const double inv2h = 0.5 / A.da;
    const double dt = A.S->dt;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double gR = A.gx[(size_t)(c * NQ + g) * A.nfx + fr];
        const double gL = A.gx[(size_t)(c * NQ + g) * A.nfx + fl];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double gD = A.gy[(size_t)(c * NQ + v) * A.nfy + fd];
            const double gU = A.gy[(size_t)(c * NQ + v) * A.nfy + fu];
            double rhs = R[c][v] + (T.sgn[v] * gL - gR) + (T.sgn[g] * gU - gD);
            if (!upd) rhs = 0.0;
            const int m = 4 * v + g;
            const double dU = rhs * (T.invm[m] * inv2h);
            const size_t idx = (size_t)(c * NM + m) * A.L.ps + A.L.rc(ii + 3, jj + 3);
            const double qn = A.Qn[idx];
            const double out = (A.stage == 0) ? qn + dt * dU : A.ca * qn + A.cb * (Qr[c][v] + dt * dU);
            if (live) {
                A.Qout[idx] = out;
                if (A.means && m == 0) A.means[(size_t)c * np + k] = out;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fused stage kernel (P = 1..3; TVB, learned or no limiter): faces, volume
// term and Shu-Osher combination in one pass over the state.
//
// A wave owns a run of 16 adjacent columns (lanes col = 0..15; columns 1..14
// are updated, 0 and 15 supply the outer face traces) and marches down a
// segment of R rows, plus one halo row above and one below.  Per row:
//   * the modes of the 16 cells are loaded once (lane (g, col) holds modes
//     px = v, py = g, v = 0..3, in 4x4 slots; P < 3 leaves slots past P zero);
//   * one FP64 MFMA product per variable evaluates all 16 face-trace points of
//     every cell (rows r = q + 4 side, sides W, E, N, S; lane group g receives
//     quadrature point q = g of each side);
//   * the volume integral is the MFMA evaluation / projection of dg_cell_mfma3
//     generalised to P = 1..3 (quadrature slots past P + 1 have zero weight);
//   * x faces pair lane col's E trace with lane col+1's W trace (a shuffle in
//     the 16-lane row); y faces pair the carried S trace of the row above with
//     this row's N trace, so every face flux is computed once;
//   * the flux projections onto the face modes are MFMA products whose
//     A operand repeats the weights so that every lane receives exactly the
//     coefficients its own modes need.
// Row r is finished (combined and stored) when row r+1 supplies its S face.
// ---------------------------------------------------------------------------
struct MfmaTabsF {
    double phiA[4][64];  // evaluation at volume point p = 4a + b of mode (px = s, py = l/16)
    double wfA[4][64];   // projection of F at point (a = s, b = l/16) onto mode l%16 = 4px + py
    double wgA[4][64];   // same for G
    double trA[4][64];   // evaluation at trace point r = l%16 (q = r%4, side = r/4) of mode (s, l/16)
    double axA[64];      // x-face projection: row i -> w_q L_{i%4}(s_q), column q = l/16
    double ayA[64];      // y-face projection: row i -> w_q L_{i/4}(s_q)
    double zsA[4][64];   // positivity points: truncated basis() at (a, b), point l%16 = 4a + b
    double zc[16];       // slot 4px+py: max over the 16 positivity points of |truncated basis()|
    double zcx[16];      // the same over those and the extra points of the positivity option (full basis there)
};
__constant__ MfmaTabsF c_mff[3];

// The positivity option's extra points (zs_ext_points) for the matrix-core
// positivity routine: chain h evaluates points 16 h + l % 16 (padded with the
// first point) of mode (px = s, py = l / 16) with the full basis.
struct MfmaTabsX {
    double zxA[2][4][64];
};
__constant__ MfmaTabsX c_mfx[3];

MfmaTabsF make_mfma_tables_fused(int P);

// Tables of the fused kernel's 4x4-block form (F4), for v_mfma_f64_4x4x4f64:
// four independent 4x4x4 products per instruction, D_b = A_b B_b + C_b for
// block b = (lane >> 2) & 3.  Lane 16 r + 4 b + c holds A_b[c][r], B_b[r][c]
// and D_b[r][c] (probed on gfx950).  With B = Q[c][v], the lane (g, col)
// layout of the fused kernel (modes (px = v, py = g) of cell col), block b
// holds cells 4b..4b+3 and D_b[i][col & 3] lands in lane (i, col): a
// constant A contracts over py (the lane group) for 4 cells at once, so the
// tensor-product (sum-factorised) steps need no data movement.  A operands
// are per-lane values (lane -> A[c][r]), the same for every block.
struct MfmaTabs4 {
    double aL[64];      // A[b][py] = L_py(eta_b): values at the Gauss rows eta_b
    double aNv[4][64];  // A[i][py] = L_v(xi_i) L_py(-1): N trace at xi_i, the px = v term
    double aSv[4][64];  // A[i][py] = L_v(xi_i) L_py(+1): S trace
    double aGN[4][64];  // A[i][q] = w_q L_v(xi_q) for every i: N-face projection onto px = v
    double aWL[64];     // A[py][q] = w_q L_py(eta_q): x-face and volume F projection onto py
    double aWD[64];     // A[py][q] = w_q L'_py(eta_q): volume G projection onto py
    double zc[16];      // as MfmaTabsF::zc
    double zcx[16];     // as MfmaTabsF::zcx
};
__constant__ MfmaTabs4 c_mf4[3];

MfmaTabs4 make_mfma_tables4(int P)
{
    MfmaTabs4 M{};
    const int NQ = P + 1;
    const double *gp = GLP_H[NQ], *gw = GLW_H[NQ];
    auto loc = [&](int q) { return q < NQ ? gp[q] : 0.0; };
    auto wt = [&](int q) { return q < NQ ? gw[q] : 0.0; };
    auto L = [&](int n, double x) { return n <= P ? legP_h(n, x) : 0.0; };
    auto D = [&](int n, double x) { return n <= P ? legD_h(n, x) : 0.0; };
    for (int l = 0; l < 64; ++l) {
        const int r = l >> 4, c = l & 3;  // A[c][r] sits in this lane
        M.aL[l] = c < NQ ? L(r, loc(c)) : 0.0;
        for (int v = 0; v < 4; ++v) {
            M.aNv[v][l] = c < NQ ? L(v, loc(c)) * L(r, -1.0) : 0.0;
            M.aSv[v][l] = c < NQ ? L(v, loc(c)) * L(r, +1.0) : 0.0;
            M.aGN[v][l] = wt(r) * L(v, loc(r));
        }
        M.aWL[l] = wt(r) * L(c, loc(r));
        M.aWD[l] = wt(r) * D(c, loc(r));
    }
    const MfmaTabsF F = make_mfma_tables_fused(P);
    for (int i = 0; i < 16; ++i) {
        M.zc[i] = F.zc[i];
        M.zcx[i] = F.zcx[i];
    }
    return M;
}

MfmaTabsF make_mfma_tables_fused(int P)
{
    MfmaTabsF M{};
    const int NQ = P + 1;
    const double *gp = GLP_H[NQ], *gw = GLW_H[NQ];
    // quadrature slot -> location and weight; slots past NQ sit at 0 with weight 0
    auto loc = [&](int q) { return q < NQ ? gp[q] : 0.0; };
    auto wt = [&](int q) { return q < NQ ? gw[q] : 0.0; };
    auto L = [&](int n, double x) { return n <= P ? legP_h(n, x) : 0.0; };
    auto D = [&](int n, double x) { return n <= P ? legD_h(n, x) : 0.0; };
    for (int s = 0; s < 4; ++s)
        for (int l = 0; l < 64; ++l) {
            const int i = l % 16, k = l / 16;
            {
                const int a = i / 4, b = i % 4, px = s, py = k;
                M.phiA[s][l] = L(px, loc(a)) * L(py, loc(b));
            }
            {
                const int px = i / 4, py = i % 4, a = s, b = k;
                const double ww = wt(a) * wt(b);
                M.wfA[s][l] = ww * D(px, loc(a)) * L(py, loc(b));
                M.wgA[s][l] = ww * L(px, loc(a)) * D(py, loc(b));
            }
            {
                const int q = i % 4, side = i / 4, px = s, py = k;
                double v = 0.0;
                switch (side) {
                case 0: v = L(px, -1.0) * L(py, loc(q)); break;  // W, xi = -1
                case 1: v = L(px, +1.0) * L(py, loc(q)); break;  // E, xi = +1
                case 2: v = L(px, loc(q)) * L(py, -1.0); break;  // N, eta = -1 (row i-1 side)
                default: v = L(px, loc(q)) * L(py, +1.0); break; // S, eta = +1 (row i+1 side)
                }
                M.trA[s][l] = v;
            }
        }
    for (int l = 0; l < 64; ++l) {
        const int i = l % 16, q = l / 16;
        M.axA[l] = wt(q) * L(i % 4, loc(q));
        M.ayA[l] = wt(q) * L(i / 4, loc(q));
    }
    // the 4x4 positivity grid of every degree uses the 4-point Gauss nodes and
    // the reference's basis(), which drops degree 3 (src_DG/fluid.cpp:897-991)
    const double zsp[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
    for (int s = 0; s < 4; ++s)
        for (int l = 0; l < 64; ++l) {
            const int i = l % 16, a = i / 4, b = i % 4, px = s, py = l / 16;
            M.zsA[s][l] = (px <= P && py <= P) ? zsP_h(px, zsp[a]) * zsP_h(py, zsp[b]) : 0.0;
        }
    for (int px = 0; px < 4; ++px)
        for (int py = 0; py < 4; ++py) {
            double mx = 0.0;
            if (px <= P && py <= P)
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b) mx = std::max(mx, std::fabs(zsP_h(px, zsp[a]) * zsP_h(py, zsp[b])));
            M.zc[4 * px + py] = mx;
            double xs[20], ys[20];
            const int ne = zs_ext_points(P, gp, xs, ys);
            for (int e = 0; e < ne && px <= P && py <= P; ++e)
                mx = std::max(mx, std::fabs(legP_h(px, xs[e]) * legP_h(py, ys[e])));
            M.zcx[4 * px + py] = mx;
        }
    return M;
}

// Workgroup barrier for the LDS exchanges of the fused kernel.  It waits for
// this wave's LDS accesses only: __syncthreads would also wait for the
// global-to-LDS row loads in flight (they count as vector memory operations).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// WG = false: every wave is independent and owns 16 columns, 14 of them
// updated.  WG = true: the four waves of a workgroup march the same rows side
// by side over 64 columns, 62 of them updated; the x faces between waves take
// the neighbour wave's boundary trace and face projection from LDS (two
// barriers per row), so halo columns exist only at workgroup edges.
// MODE (with WG): 0 reads the matrix-core tables from constant memory every
// row; 1 copies them to LDS once; 2 also brings the rows to LDS by direct
// global-to-LDS loads issued one row ahead (wave w loads the planes of
// variable w), so the next row is in flight while this one is computed,
// without holding registers for it; 3 is 2 with two row buffers, which
// saves the barrier that frees the single buffer.
//
// F4: the products run as 4x4x4 blocks (v_mfma_f64_4x4x4f64, MfmaTabs4) on
// the sum-factorised form: values are contracted over py (the lane group) by
// the matrix core, four cells per block, and over px (the register index) by
// VALU multiply-adds.  Same lane layout, same results up to summation order.
//
// SK: 1 = the first RK stage, 2 = a later one, fixed at compile time so that
// the stage combination carries no selects (0 reads A.stage).  O32: plane
// offsets in 32 bits (state arrays below 4 GiB), which shortens the scalar
// address arithmetic of the 3 x 4 (P+1)^2 plane accesses per row.
//
// AL (with WG, MODE 3, F4): aligned blocks.  Run r covers the 64 columns
// [64 r, 64 r + 64) and updates all of them that lie in the update range, so
// every 16-column segment a wave reads or writes is one 128-byte line and the
// workgroup's plane rows start on 512-byte boundaries.  With the 62-column
// stride of the default form each such segment straddles two lines and the
// stores leave partial lines, which a probe of the same access pattern
// without arithmetic measured at 475 against 348 us for the three arrays of
// the second stage.  A block has 65 x faces for its 64 columns: the two at
// its edges need the traces of the columns either side (64 r - 1 and
// 64 r + 64), whose modes arrive with each row by one global-to-LDS load per
// wave.  At the top of each row wave 1 evaluates the two edge fluxes from
// LDS alone (side columns and columns 0 and 63, with the same trace products
// as the rows, so both blocks at an edge compute the same flux bit for bit),
// and the edge cells take them into the same sums as their inner faces
// (a uniform state stays exactly uniform).
template <int P, int MINW, bool WG, int MODE, bool F4, int SK = 0, bool O32 = false, bool AL = false>
__global__ __launch_bounds__(256, MINW) void dg_fused_kernel(DgArgs A, int R, int nrun)
{
    using Ofs = typename std::conditional<O32, uint32_t, size_t>::type;
    constexpr int NQ = P + 1, NM = NQ * NQ;
    constexpr bool DMA = MODE >= 2, DB = MODE == 3;
    constexpr int NB = DB ? 2 : 1, NS = DB ? 4 : 2;  // row buffers, ghost-map row slots
    static_assert(WG || MODE == 0, "LDS tables need the workgroup form");
    static_assert(!F4 || MODE >= 1, "the 4x4-block form reads its tables from LDS");
    static_assert(!AL || (WG && MODE == 3 && F4), "aligned blocks are a form of the default pipeline");
    constexpr size_t TAB_BYTES = F4 ? sizeof(MfmaTabs4) : sizeof(MfmaTabsF);
    __shared__ double tabs_lds[MODE >= 1 ? TAB_BYTES / 8 : 1];
    const MfmaTabsF &M = (MODE >= 1) ? *reinterpret_cast<const MfmaTabsF *>(tabs_lds) : c_mff[P - 1];
    const MfmaTabs4 &M4 = *reinterpret_cast<const MfmaTabs4 *>(tabs_lds);
    const DgTables &T = c_tab[P - 1];
    __shared__ double lds_w[WG ? 4 : 1][4][4];   // [wave][g][c]: W trace of the wave's column 0
    __shared__ double lds_ge[WG ? 4 : 1][4][4];  // [wave][g][c]: projected flux of its column-15 face
    __shared__ double qbuf[DMA ? NB : 1][DMA ? 4 * NM : 1][DMA ? 64 : 1];  // rows: plane (c, mode) x 64 columns
    // ghost-map entries of row r: slot r % NS; with AL two more, of the side columns
    __shared__ int32_t srcbuf[DMA ? NS : 1][DMA ? (AL ? 66 : 64) : 1];
    // AL: modes of the two side columns [variable][mode][side] per row buffer;
    // the edge traces of the row [which][g][c] (E of the left side column, W of
    // the right one, W of column 0, E of column 63); and the row's projected
    // edge fluxes [side][py][c]
    __shared__ double hbuf[AL ? NB : 1][AL ? 4 * NM * 2 : 1];
    __shared__ double lds_ht[AL ? 4 : 1][4][4];
    __shared__ double lds_bf[AL ? 2 : 1][4][4];
    if (A.S->done) return;
    const bool first = SK == 1 ? true : SK == 2 ? false : A.stage == 0;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, col = lane & 15;
    const int wv = threadIdx.x >> 6;
    const int nseg = (A.nyu + R - 1) / R;
    int run, seg, j;
    bool store_col;
    if (WG) {
        if ((int)blockIdx.x >= nrun * nseg) return;  // whole workgroup
        // neighbouring runs of a segment on one XCD: the cache lines that
        // straddle two runs' columns are then fetched into one L2, not two
        const int t = xcd_tile((int)blockIdx.x, nrun * nseg);
        run = t % nrun;
        seg = t / nrun;
        const int x = 16 * wv + col;
        if (AL) {
            j = 64 * run + x;
            store_col = j >= 3 && j <= A.nx - 4;
        } else {
            j = 2 + run * 62 + x;
            store_col = x >= 1 && x <= 62 && j <= A.nx - 4;
        }
    } else {
        const int w = blockIdx.x * (blockDim.x >> 6) + wv;
        if (w >= nrun * nseg) return;  // whole wave
        run = w % nrun;
        seg = w / nrun;
        j = 2 + run * 14 + col;
        store_col = col >= 1 && col <= 14 && j <= A.nx - 4;
    }
    if (MODE >= 1) {
        const double *src = F4 ? reinterpret_cast<const double *>(&c_mf4[P - 1])
                               : reinterpret_cast<const double *>(&c_mff[P - 1]);
        for (int i = threadIdx.x; i < (int)(TAB_BYTES / 8); i += 256) tabs_lds[i] = src[i];
        __syncthreads();
    }
    auto mf4 = [](double a, double b) { return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, 0.0, 0, 0, 0); };
    auto mf4c = [](double a, double b, double c) { return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0); };
    const int jc = j < A.nx ? j : A.nx - 1;           // clamped for loads past the right edge
    const int r0 = 3 + seg * R;
    const int r1 = (r0 + R < A.ny - 3) ? r0 + R : A.ny - 3;  // rows [r0, r1), halo rows r0-1 and r1
    const size_t np = A.np;
    const StLay L = A.L;
    // element offset of plane 0 of cell kk, 32-bit when the state arrays are below 4 GiB
    auto cel = [&](size_t kk) { return (Ofs)L.cell(kk); };
    const bool slot = g <= P;                         // this lane group holds real modes
    const double gam = A.gamma, dt = A.S->dt, inv2h = 0.5 / A.da;
    // per-slot factor of dU in the combination: dt (stage 1) or cb*dt, times invMass * 0.5/h
    double fdu[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int m = v * NQ + g;
        const double sc = (v <= P && slot) ? T.invm[m < NM ? m : 0] * inv2h : 0.0;
        fdu[v] = sc;
    }

    // Addressing: plane (c, px = v) is a wave-uniform pointer; a lane adds the
    // element offset g*ps + cell, 32-bit below 4 GiB so that addresses need no
    // 64-bit VGPR pairs.
    const Ofs gofs = (Ofs)((size_t)(slot ? g : 0) * L.ps);
    auto load = [&](int row, double Q[4][4]) {
        const size_t k = (size_t)row * A.nx + jc;
        const int32_t e = A.src[k];
        const Ofs o = gofs + cel((size_t)(e >> kSrcShift));
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                Q[c][v] = (v <= P && slot) ? (A.Qin + (size_t)(c * NM + v * NQ) * L.ps)[o] : 0.0;
        if (e & 1)
#pragma unroll
            for (int v = 0; v < 4; ++v) Q[1][v] = -Q[1][v];
        if (e & 2)
#pragma unroll
            for (int v = 0; v < 4; ++v) Q[2][v] = -Q[2][v];
    };

    // Row staging for DMA.  Lane l of wave w moves dword (l & 1) of columns
    // (l >> 1) and 32 + (l >> 1) of every plane of variable w; the source
    // cells come from the ghost map (slip signs are applied on reading).
    // Wave 0 also moves the ghost-map row two ahead into srcbuf.
    const int jrun0 = AL ? 64 * run : 2 + run * 62;
    const int wvu = __builtin_amdgcn_readfirstlane(wv);
    // ghost-map entry of block column x of a row, the column clamped into the
    // grid (x = -1 and x = 64 are the AL side columns)
    auto col_src = [&](int row, int x) {
        int jj = jrun0 + x < A.nx ? jrun0 + x : A.nx - 1;
        jj = jj < 0 ? 0 : jj;
        return A.src[(size_t)row * A.nx + jj];
    };
    auto dma_row = [&](int row, int32_t e0, int32_t e1, Ofs npb) {
        auto qb = qbuf[DB ? (row & 1) : 0];
        const uint32_t b = (uint32_t)(lane & 1) * 4u;
        const Ofs o0 = cel((size_t)(e0 >> kSrcShift)) * 8u + b, o1 = cel((size_t)(e1 >> kSrcShift)) * 8u + b;
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            const int pl = wvu * NM + m;
            const char *plane = reinterpret_cast<const char *>(plane_at(A.Qin, pl, npb));
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)(plane + o0),
                                             (void __attribute__((address_space(3))) *)&qb[pl][0], 4, 0, 0);
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)(plane + o1),
                                             (void __attribute__((address_space(3))) *)&qb[pl][32], 4, 0, 0);
        }
    };
    // A row whose 64 columns are their own sources (no ghost, mirror or
    // clamped column) is copied as it lies, two planes per 16-byte load: lane
    // l moves columns 2 (l % 32) and 2 (l % 32) + 1 of plane pl + l / 32.
    // Every such segment ends inside its row (column jrun0 + 63 < nx).
    auto dma_row16 = [&](int row, Ofs npb) {
        auto qb = qbuf[DB ? (row & 1) : 0];
        const Ofs o = (Ofs)(L.rc(row, jrun0) * 8 + (size_t)(lane & 31) * 16) + ((lane >> 5) ? npb : (Ofs)0);
#pragma unroll
        for (int m = 0; m < NM; m += 2) {
            const char *plane = reinterpret_cast<const char *>(plane_at(A.Qin, wvu * NM + m, npb));
            if (m + 1 < NM || lane < 32)  // an odd plane count ends with one plane (lanes 0-31)
                __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)(plane + o),
                                                 (void __attribute__((address_space(3))) *)&qb[wvu * NM + m][0],
                                                 16, 0, 0);
        }
    };
    auto dma_src = [&](int row) {
        const int32_t *rowp = A.src + (size_t)row * A.nx;  // wave-uniform
        // the lane index is re-derived here rather than kept live across the row
        int l;
        asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
        const uint32_t js = (uint32_t)(jrun0 + l < A.nx ? jrun0 + l : A.nx - 1);
        __builtin_amdgcn_global_load_lds(
            (const void __attribute__((address_space(1))) *)(rowp + js),
            (void __attribute__((address_space(3))) *)&srcbuf[row % NS][0], 4, 0, 0);
        if (AL && l < 2) {  // entries 64 and 65: the columns left and right of the block
            int jh = l == 0 ? jrun0 - 1 : jrun0 + 64;
            jh = jh < 0 ? 0 : (jh < A.nx ? jh : A.nx - 1);
            __builtin_amdgcn_global_load_lds(
                (const void __attribute__((address_space(1))) *)(rowp + jh),
                (void __attribute__((address_space(3))) *)&srcbuf[row % NS][64], 4, 0, 0);
        }
    };
    // AL: the modes of the two side columns of a row, one instruction per wave:
    // lane l of wave w moves dword (l & 1) of side (l >> 1) & 1 of plane
    // w NM + (l >> 2), from the side columns' ghost-map sources el and er
    auto dma_halo = [&](int row, int32_t el, int32_t er, Ofs npb) {
        int l;
        asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
        if (4 * NM >= 64 || l < 4 * NM) {
            const int32_t e = (l & 2) ? er : el;
            const char *plane = reinterpret_cast<const char *>(plane_at(A.Qin, wvu * NM + (l >> 2), npb));
            const Ofs o = cel((size_t)(e >> kSrcShift)) * 8u + (uint32_t)(l & 1) * 4u;
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)(plane + o),
                                             (void __attribute__((address_space(3))) *)&hbuf[DB ? (row & 1) : 0]
                                                                                           [wvu * NM * 2],
                                             4, 0, 0);
        }
    };
    if (DMA) {
        dma_row(r0, col_src(r0, lane >> 1), col_src(r0, 32 + (lane >> 1)), (Ofs)(L.ps * 8));
        if (AL) dma_halo(r0, col_src(r0, -1), col_src(r0, 64), (Ofs)(L.ps * 8));
        if (wvu == 0) {
            dma_src(r0);
            if (r0 + 1 <= r1) dma_src(r0 + 1);
        }
    }

    // S trace of the halo row above the segment
    double Sup[4];
    {
        double Q[4][4];
        load(r0 - 1, Q);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if constexpr (F4) {
                double t = 0.0;
#pragma unroll
                for (int v = 0; v <= P; ++v) t = mf4c(M4.aSv[v][lane], Q[c][v], t);
                Sup[c] = t;
            } else {
                d4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s <= P; ++s)
                    t = __builtin_amdgcn_mfma_f64_16x16x4f64(M.trA[s][lane], Q[c][s], t, 0, 0, 0);
                Sup[c] = t[3];
            }
        }
    }
    // pending row r-1: out = base - fs * gS, its S-face projection gS still to come
    double base[4][4];
    double fs[4];
    size_t kp = 0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) base[c][v] = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) fs[v] = 0.0;

    for (int r = r0; r <= r1; ++r) {
        // the table operands are re-read from constant memory every row: an
        // opaque copy of the lane index keeps the compiler from holding all of
        // them in registers across the loop
        int ln = lane;
        asm volatile("" : "+v"(ln));
        // likewise the plane stride: the plane addresses are formed per row
        // (scalar arithmetic) instead of being held across the loop
        Ofs npb = (Ofs)(L.ps * 8);
        asm volatile("" : "+s"(npb));
        const bool inner = r < r1;
        const size_t k = (size_t)r * A.nx + jc;
        // With two row buffers, row r stays in LDS for the whole iteration, so
        // its modes are read again where they are needed instead of being held
        // in registers across the row (through a fresh opaque lane index, so
        // the reads are not merged with the first one).  At P = 3 this is what
        // keeps the kernel within 256 VGPRs; at P = 1 it makes room for a third
        // wave per SIMD; at P = 2 holding the modes is faster.
        auto reread = [&](double Qr[4][4]) {
            int l2 = lane;
            asm volatile("" : "+v"(l2));
            const int x = 16 * wvu + (l2 & 15);
            const int32_t eo = srcbuf[r % NS][x];
            auto qb = qbuf[DB ? (r & 1) : 0];
            const int gq = (l2 >> 4) <= P ? (l2 >> 4) : 0;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < 4; ++v) Qr[c][v] = (v <= P && slot) ? qb[c * NM + v * NQ + gq][x] : 0.0;
            if (eo & 1)
#pragma unroll
                for (int v = 0; v < 4; ++v) Qr[1][v] = -Qr[1][v];
            if (eo & 2)
#pragma unroll
                for (int v = 0; v < 4; ++v) Qr[2][v] = -Qr[2][v];
        };
        // the same for variable c alone (the aligned form at P = 3: fewer
        // registers live where the volume terms and the combination read them)
        auto reread_var = [&](int c, double q[4]) {
            int l2 = lane;
            asm volatile("" : "+v"(l2));
            const int x = 16 * wvu + (l2 & 15);
            const int32_t eo = srcbuf[r % NS][x];
            auto qb = qbuf[DB ? (r & 1) : 0];
            const int gq = (l2 >> 4) <= P ? (l2 >> 4) : 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) q[v] = (v <= P && slot) ? qb[c * NM + v * NQ + gq][x] : 0.0;
            if ((c == 1 && (eo & 1)) || (c == 2 && (eo & 2)))
#pragma unroll
                for (int v = 0; v < 4; ++v) q[v] = -q[v];
        };
        constexpr bool PERVAR = AL && DB && P != 2;
        double Q[4][4];
        bool upd;  // this lane's cell of row r is updated
        if (DMA) {
            // row r is complete in qbuf once every wave's loads have landed
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            // LDS addresses come from the opaque lane copy, so they are
            // recomputed per row instead of being held (or spilled) across it
            const int x = 16 * wvu + (ln & 15);
            const int32_t eo = srcbuf[r % NS][x];
            upd = (eo & kSrcUpdated) != 0;  // no global load (vmcnt counts in order)
            auto qb = qbuf[DB ? (r & 1) : 0];
            const int gq = (ln >> 4) <= P ? (ln >> 4) : 0;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < 4; ++v) Q[c][v] = (v <= P && slot) ? qb[c * NM + v * NQ + gq][x] : 0.0;
            if (eo & 1)
#pragma unroll
                for (int v = 0; v < 4; ++v) Q[1][v] = -Q[1][v];
            if (eo & 2)
#pragma unroll
                for (int v = 0; v < 4; ++v) Q[2][v] = -Q[2][v];
            // the single buffer is free for the next row once every wave has
            // read it; with two buffers the next row goes to the other one,
            // last read in the previous iteration (before this row's barrier)
            if (!DB) __syncthreads();
            if (AL && inner && wvu == 1) {
                // The two edge fluxes of row r, from LDS alone, before the next
                // row's loads to LDS go out (the compiler makes an LDS read after
                // those loads wait for all of them to land: 2% of the kernel's
                // time at 1024^2 P = 3).  Traces: lanes (g, col) of block
                // b = col / 4 hold variable col % 4 of the E trace of the left
                // side column (b = 0), the W trace of the right one (b = 1), the
                // W trace of column 0 (b = 2) and the E trace of column 63
                // (b = 3), each at point g, with the rows' trace products.
                int l2 = lane;
                asm volatile("" : "+v"(l2));
                const int g2 = l2 >> 4, bk = (l2 >> 2) & 3, cv = l2 & 3, gq = g2 <= P ? g2 : 0;
                const int32_t eh = bk < 2 ? srcbuf[r % NS][64 + bk] : srcbuf[r % NS][bk == 2 ? 0 : 63];
                const bool neg = (cv == 1 && (eh & 1)) || (cv == 2 && (eh & 2));
                auto qb = qbuf[r & 1];
                double tr = 0.0;
#pragma unroll
                for (int v = 0; v <= P; ++v) {
                    const int pl = cv * NM + v * NQ + gq;
                    double q = g2 > P ? 0.0 : bk < 2 ? hbuf[r & 1][pl * 2 + bk] : qb[pl][bk == 2 ? 0 : 63];
                    q = neg ? -q : q;
                    const double t = mf4(M4.aL[l2], q);
                    // W traces (blocks 1, 2) alternate the sign of the odd modes
                    tr = v == 0 ? t : tr + (((bk == 1 || bk == 2) && (v & 1)) ? -t : t);
                }
                lds_ht[bk][g2][cv] = tr;
                // lanes (g, 0): the left edge face at point g (side column | column 0);
                // lanes (g, 1): the right one (column 63 | side column)
                const int c2 = l2 & 15;
                double Lx[4], Rx[4], Fx[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    Lx[c] = lds_ht[c2 == 0 ? 0 : 3][g2][c];
                    Rx[c] = lds_ht[c2 == 0 ? 2 : 1][g2][c];
                }
                if (A.flux == ZK_FLUX_HLL) hll_f64(Lx, Rx, gam, Fx);
                else hllc_dg_fast_f64(Lx, Rx, gam, Fx);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double gb = mf4(M4.aWL[l2], Fx[c]);  // onto py = g
                    if (c2 < 2) lds_bf[c2][g2][c] = gb;
                }
            }
            if (inner) {
                const int32_t *sn = srcbuf[(r + 1) % NS];
                // one column per lane: is row r + 1 its own source throughout?
                const int xs = ln & 63;
                const int32_t es = sn[xs];
                const bool own = jrun0 + xs < A.nx && (es & 3) == 0 &&
                                 (es >> kSrcShift) == (int32_t)((r + 1) * A.nx + jrun0 + xs);
                if (A.row16 && __ballot(own) == ~0ull)
                    dma_row16(r + 1, npb);
                else
                    dma_row(r + 1, sn[ln >> 1], sn[32 + (ln >> 1)], npb);
                if (AL) dma_halo(r + 1, sn[64], sn[65], npb);
                if (wvu == 0 && r + 2 <= r1) dma_src(r + 2);
            }
        } else {
            load(r, Q);
            upd = A.upd[k] != 0;
        }
        // face traces of this row: element side (W, E, N, S) at quadrature point g
        d4 Tr[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if constexpr (F4) {
                // W, E: the modes at eta_g (contracted over py), then L_v(-+1) over v;
                // N, S: L_v(xi_g) L_py(-+1) accumulated over v on the matrix core
                const double t0 = mf4(M4.aL[ln], Q[c][0]);
                double w = t0, e = t0;
                double n = mf4(M4.aNv[0][ln], Q[c][0]), sv = mf4(M4.aSv[0][ln], Q[c][0]);
#pragma unroll
                for (int v = 1; v <= P; ++v) {
                    const double t = mf4(M4.aL[ln], Q[c][v]);
                    w += (v & 1) ? -t : t;
                    e += t;
                    n = mf4c(M4.aNv[v][ln], Q[c][v], n);
                    sv = mf4c(M4.aSv[v][ln], Q[c][v], sv);
                }
                Tr[c] = d4{w, e, n, sv};
            } else {
                d4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s <= P; ++s)
                    t = __builtin_amdgcn_mfma_f64_16x16x4f64(M.trA[s][ln], Q[c][s], t, 0, 0, 0);
                Tr[c] = t;
            }
        }
        // y face between row r-1 (S trace Sup) and row r (N trace), projected onto px = v
        d4 gN[4];
        {
            const double a[4] = {Sup[0], Sup[2], Sup[1], Sup[3]};
            const double b[4] = {Tr[0][2], Tr[2][2], Tr[1][2], Tr[3][2]};
            double Gs[4];
            if (A.flux == ZK_FLUX_HLL) hll_f64(a, b, gam, Gs);
            else hllc_dg_fast_f64(a, b, gam, Gs);
            const double G[4] = {Gs[0], Gs[2], Gs[1], Gs[3]};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if constexpr (F4) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) gN[c][v] = v <= P ? mf4(M4.aGN[v][ln], G[c]) : 0.0;
                } else {
                    gN[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(M.ayA[ln], G[c], d4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
                }
            }
        }
        // sab[c]: sum over the non-mean modes of |mode| * max |basis| at the positivity points
        double sab[4] = {0.0, 0.0, 0.0, 0.0}, mean0[4] = {0.0, 0.0, 0.0, 0.0};
        // this lane's linear mode (x in group 0, y in group 1) is above M h^2, or
        // NaN.  At P >= 3 the TVB test also reads the face-average deviations,
        // sums of the P = 3 modes of degree 0 on one axis: group 0 marks a
        // variable whose sum_q |a(q,0)| is above M h^2 (a bound of the x
        // deviations, exact under monotone rounding), group q one whose |a(0,q)|
        // is above M h^2 / 4 (the y deviations then stay below 3/4 M h^2).  A
        // cell the TVB test can change is always marked.
        bool lin_big = false;
        if (r > r0 && store_col && slot) {
            // finish row r-1: its S face is this row's N face
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                double t0 = 0.0, sx = 0.0;  // P >= 3: |a(0,g)| and sum_q |a(q,0)| of this lane
#pragma unroll
                for (int v = 0; v <= P; ++v) {
                    const double out = base[c][v] - fs[v] * gN[c][v];
                    plane_at(A.Qout, c * NM + v * NQ, npb)[gofs + (Ofs)L.rc(r - 1, jc)] = out;
                    if (v == 0 && g == 0) {
                        mean0[c] = out;
                        if (A.means) A.means[(size_t)c * np + kp] = out;
                    } else {
                        const double *zc = A.zs_ext ? (F4 ? M4.zcx : M.zcx) : (F4 ? M4.zc : M.zc);
                        sab[c] += fabs(out) * zc[4 * v + g];
                    }
                    if constexpr (P >= 3) {
                        if (v == 0)
                            t0 = out;
                        else
                            sx += fabs(out);
                    } else if ((v == 1 && g == 0) || (v == 0 && g == 1)) {
                        lin_big = lin_big || !(fabs(out) <= A.Mh2);
                    }
                }
                if constexpr (P >= 3) lin_big = lin_big || !((g == 0 ? sx : 4.0 * fabs(t0)) <= A.Mh2);
            }
        }
        if (A.zsok && r > r0) {
            // Sufficient condition for the positivity scaling to leave the cell
            // alone: lower bounds of rho and p over its 16 points (means minus
            // the mode sums) clear the 1e-5 floors with a margin.
            // sums over the four lane groups land in group 0, which holds the means
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                sab[c] += __shfl_down(sab[c], 32);
                sab[c] += __shfl_down(sab[c], 16);
            }
            int cand = lin_big ? 1 : 0;
            if (P >= 3) cand |= __shfl_down(cand, 32);  // groups 2 and 3 onto 0 and 1
            cand |= __shfl_down(cand, 16);  // group 1's y-linear test
            if (g == 0 && store_col) {
                const double rlo = mean0[0] - sab[0];
                const double elo = mean0[3] - sab[3];
                const double mx = fabs(mean0[1]) + sab[1], my = fabs(mean0[2]) + sab[2];
                const double tol = 1e-12 * (fabs(mean0[0]) + sab[0] + fabs(mean0[3]) + sab[3] + mx + my);
                bool ok = rlo > 2e-5 + tol;
                if (ok) {
                    // p_lo = (gam - 1) (elo - |m|^2 / (2 rlo)) > 2e-5 + tol, multiplied
                    // through by rlo > 0 (no division; the margin over the 1e-5
                    // floor dwarfs the rounding of either form)
                    ok = (gam - 1.0) * (elo * rlo - 0.5 * (mx * mx + my * my)) > (2e-5 + tol) * rlo;
                }
                A.zsok[kp] = (uint8_t)((ok ? 1 : 0) | (cand ? 2 : 0));
            }
        }
        if (!inner) break;
        double qn[4][4];
        if (store_col && slot && !(first && upd))
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v <= P; ++v) qn[c][v] = plane_at(A.Qn, c * NM + v * NQ, npb)[gofs + (Ofs)L.rc(r, jc)];
        // rhs of row r, all terms but its S face: N face, x faces, then the volume
        // projection accumulated on top by the MFMA
        d4 acc[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double sN = (g & 1) ? -1.0 : 1.0;  // L_g(-1)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[c][v] = sN * gN[c][v];
        }
        {
            double Lx[4], Rx[4], Fx[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                Lx[c] = Tr[c][1];
                Rx[c] = __shfl_down(Tr[c][0], 1, 16);
                Sup[c] = Tr[c][3];
            }
            if (WG) {
                if (col == 0)
#pragma unroll
                    for (int c = 0; c < 4; ++c) lds_w[wvu][ln >> 4][c] = Tr[c][0];
                lds_barrier();
                if (col == 15 && wv < 3)
#pragma unroll
                    for (int c = 0; c < 4; ++c) Rx[c] = lds_w[wvu + 1][ln >> 4][c];
            }
            if (A.flux == ZK_FLUX_HLL) hll_f64(Lx, Rx, gam, Fx);
            else hllc_dg_fast_f64(Lx, Rx, gam, Fx);
            double gE[4], gW[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if constexpr (F4) {
                    gE[c] = mf4(M4.aWL[ln], Fx[c]);      // py = g projection of face (col | col+1)
                } else {
                    const d4 ge =
                        __builtin_amdgcn_mfma_f64_16x16x4f64(M.axA[ln], Fx[c], d4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
                    gE[c] = ge[0];                         // py = g projection of face (col | col+1)
                }
                gW[c] = __shfl_up(gE[c], 1, 16);           // face (col-1 | col)
            }
            if (WG) {
                if (col == 15)
#pragma unroll
                    for (int c = 0; c < 4; ++c) lds_ge[wvu][ln >> 4][c] = gE[c];
                lds_barrier();
                if (col == 0 && wv > 0)
#pragma unroll
                    for (int c = 0; c < 4; ++c) gW[c] = lds_ge[wvu - 1][ln >> 4][c];
                if (AL && (wvu == 0 ? col == 0 : (wvu == 3 && col == 15))) {
                    // the block's edge faces, from wave 1 (written before the first x-face barrier)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const double gb = lds_bf[wvu == 0 ? 0 : 1][ln >> 4][c];
                        if (wvu == 0) gW[c] = gb;
                        else gE[c] = gb;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[c][v] += ((v & 1) ? -gW[c] : gW[c]) - gE[c];  // L_v(-1) gW - gE
        }
        if constexpr (F4) {
            // the coefficients L_v(xi_a), w_a L'_v(xi_a), w_a L_v(xi_a) fold to
            // constants (instruction operands) in the unrolled loops
            if (DB && P != 2 && !PERVAR) reread(Q);
            // the state at the points (xi_a, eta_g): modes at eta_g (contracted
            // over py on the matrix core), then sum_v L_v(xi_a) over px
            double U[4][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if constexpr (PERVAR) reread_var(c, Q[c]);
                double t[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) t[v] = v <= P ? mf4(M4.aL[ln], Q[c][v]) : 0.0;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    double u = t[0];  // L_0 = 1
#pragma unroll
                    for (int v = 1; v <= P; ++v) {
                        const double k = a < NQ ? leg_c(v, gl_x(NQ, a)) : 0.0;
                        if (k != 0.0) u += k * t[v];
                    }
                    U[c][a] = u;
                }
            }
#pragma unroll
            for (int a = 0; a <= P; ++a) {
                // fluxes at (xi_a, eta_g) and their projections: over eta on the
                // matrix core, then w_a L'_px(xi_a) / w_a L_px(xi_a) per px
                const double Up[4] = {U[0][a], U[1][a], U[2][a], U[3][a]};
                double F[4], G[4];
                euler_fg(Up, gam, F, G);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double zf = mf4(M4.aWL[ln], F[c]), zg = mf4(M4.aWD[ln], G[c]);
#pragma unroll
                    for (int v = 0; v <= P; ++v)
                    {
                        // the coefficients are constants after unrolling; the zero
                        // ones (L'_0, and L'_v / L_v at xi = 0 for even / odd v)
                        // contribute nothing and are left out
                        const double kf = gl_w(NQ, a) * legd_c(v, gl_x(NQ, a));
                        const double kg = gl_w(NQ, a) * leg_c(v, gl_x(NQ, a));
                        if (kf != 0.0) acc[c][v] += kf * zf;
                        if (kg != 0.0) acc[c][v] += kg * zg;
                    }
                }
            }
        } else {
            if (DB && P != 2) reread(Q);
            d4 U[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                d4 t = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s <= P; ++s) t = __builtin_amdgcn_mfma_f64_16x16x4f64(M.phiA[s][ln], Q[c][s], t, 0, 0, 0);
                U[c] = t;
            }
#pragma unroll
            for (int v = 0; v <= P; ++v) {
                const double Up[4] = {U[0][v], U[1][v], U[2][v], U[3][v]};
                double F[4], G[4];
                euler_fg(Up, gam, F, G);
                const double wf = M.wfA[v][ln], wg = M.wgA[v][ln];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(wf, F[c], acc[c], 0, 0, 0);
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(wg, G[c], acc[c], 0, 0, 0);
                }
            }
        }
        // combination of row r up to its S-face term
        const double cdt = first ? dt : A.cb * dt;
#pragma unroll
        for (int v = 0; v < 4; ++v) fs[v] = upd ? cdt * fdu[v] : 0.0;
        if (DB && P != 2 && !PERVAR) reread(Q);
        // ca qn + cb (Q + dt du) with du = fdu acc, as ca qn + cb Q + fs acc
        // (fs = cb dt fdu, 0 for cells not updated); the first stage is
        // Q + fs acc for updated cells and qn (not read otherwise) for the rest.
        // At P = 3 the two stage branches cost registers (spills): one
        // expression there, unless the stage is fixed at compile time.
        if constexpr (P == 3 && SK == 0) {
            if (store_col && slot) {
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int v = 0; v <= P; ++v) {
                        const double fa = fs[v] * acc[c][v];
                        base[c][v] = (A.stage == 0) ? (upd ? fa + Q[c][v] : qn[c][v])
                                                    : fa + (A.cb * Q[c][v] + A.ca * qn[c][v]);
                    }
            }
        } else if (store_col && slot) {
            if (first) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if constexpr (PERVAR) reread_var(c, Q[c]);
#pragma unroll
                    for (int v = 0; v <= P; ++v) base[c][v] = upd ? fs[v] * acc[c][v] + Q[c][v] : qn[c][v];
                }
            } else {
                const double ca = A.ca, cb = A.cb;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if constexpr (PERVAR) reread_var(c, Q[c]);
#pragma unroll
                    for (int v = 0; v <= P; ++v)
                        base[c][v] = fs[v] * acc[c][v] + (cb * Q[c][v] + ca * qn[c][v]);
                }
            }
        }
        kp = k;
    }
}

// The TVB test of one cell (per variable: tvb_var, i.e. the two linear modes
// and, at P >= 3, the face-average deviations, against the filled neighbour
// means); returns whether any variable is troubled and leaves the limited
// slopes in axl/ayl.
template <int P>
__device__ __forceinline__ bool tvb_test(const double C[4][(P + 1) * (P + 1)], const double *__restrict__ M,
                                         const int32_t *__restrict__ src, size_t np, size_t k, int nx, double Mh2,
                                         double axl[4], double ayl[4], bool chg[4])
{
    bool any = false;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        chg[c] = tvb_var<P>(C[c], M, src, np, k, nx, c, Mh2, axl[c], ayl[c]);
        any = any || chg[c];
    }
    return any;
}

// Positivity scaling (zs_cell) of the 16 cells of a wave held in the
// matrix-core layout of tvb_mfma_cells (lane (g, col): modes (px = v, py = g)
// of cell col): rho and then every variable at the 4x4 points (and, with ext,
// at the positivity option's extra points) by MFMA products, theta reduced
// across each column's lanes.  Returns whether this lane's cell changed.  The
// whole wave must call it.
template <int P>
__device__ __forceinline__ bool zs_mfma_apply(double C[4][4], double gamma, bool ext)
{
    constexpr int NQ = P + 1;
    const MfmaTabsF &MT = c_mff[P - 1];
    const MfmaTabsX &MX = c_mfx[P - 1];
    constexpr int NXCH = (4 * NQ + 4 + 15) / 16;  // chains of 16 extra points (positivity option)
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    constexpr double RHO_MIN = 1e-5, P_MIN = 1e-5, EPS = 1e-5;
    bool changed = false;
    double m0[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) m0[c] = __shfl(C[c][0], col);
    if (m0[0] < RHO_MIN) changed = true;
    m0[0] = smax(m0[0], RHO_MIN);
    {
        const double p0 = (gamma - 1.0) * (m0[3] - 0.5 * (m0[1] * m0[1] + m0[2] * m0[2]) / m0[0]);
        if (p0 < P_MIN) {
            m0[3] = P_MIN / (gamma - 1.0) + 0.5 * (m0[1] * m0[1] + m0[2] * m0[2]) / m0[0];
            changed = true;
        }
    }
    if (g == 0) {
        C[0][0] = m0[0];
        C[3][0] = m0[3];
    }
    // theta_rho from rho at the 16 points (lane holds points a = v, b = g)
    {
        d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s <= P; ++s) z = __builtin_amdgcn_mfma_f64_16x16x4f64(MT.zsA[s][lane], C[0][s], z, 0, 0, 0);
        double th = 1.0;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (z[v] < RHO_MIN) th = smin(th, (m0[0] - RHO_MIN) / (m0[0] - z[v] + EPS));
        if (ext)
#pragma unroll
            for (int h = 0; h < NXCH; ++h) {
                d4 zx = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int s = 0; s <= P; ++s) zx = __builtin_amdgcn_mfma_f64_16x16x4f64(MX.zxA[h][s][lane], C[0][s], zx, 0, 0, 0);
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (zx[v] < RHO_MIN) th = smin(th, (m0[0] - RHO_MIN) / (m0[0] - zx[v] + EPS));
            }
        th = smin(th, __shfl_xor(th, 16));
        th = smin(th, __shfl_xor(th, 32));
        if (th < 1.0) {
            changed = true;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (v != 0 || g != 0) C[c][v] *= th;
        }
    }
    // theta_p from all variables at the 16 points
    {
        d4 z[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            d4 a = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s <= P; ++s) a = __builtin_amdgcn_mfma_f64_16x16x4f64(MT.zsA[s][lane], C[c][s], a, 0, 0, 0);
            z[c] = a;
        }
        const double p_avg = (gamma - 1.0) * (m0[3] - 0.5 * (m0[1] * m0[1] + m0[2] * m0[2]) / m0[0]);
        double th = 1.0;
        auto pmin = [&](const d4 zz[4]) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double rho = zz[0][v];
                if (rho <= RHO_MIN) continue;
                const double u = zz[1][v] / rho, w = zz[2][v] / rho;
                const double p = (gamma - 1.0) * (zz[3][v] - 0.5 * rho * (u * u + w * w));
                if (p < P_MIN) th = smin(th, (p_avg - P_MIN) / (p_avg - p + EPS));
            }
        };
        pmin(z);
        if (ext)
#pragma unroll
            for (int h = 0; h < NXCH; ++h) {
                d4 zx[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    d4 a = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int s = 0; s <= P; ++s)
                        a = __builtin_amdgcn_mfma_f64_16x16x4f64(MX.zxA[h][s][lane], C[c][s], a, 0, 0, 0);
                    zx[c] = a;
                }
                pmin(zx);
            }
        th = smin(th, __shfl_xor(th, 16));
        th = smin(th, __shfl_xor(th, 32));
        if (th < 1.0) {
            changed = true;
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (v != 0 || g != 0) C[c][v] *= th;
        }
    }
    return changed;
}

// Positivity scaling alone of 16 cells (cell k of column lane col; live is
// false on padding lanes, which compute on a valid cell and store nothing), in
// the matrix-core layout.  The whole wave must call it.
template <int P>
__device__ __forceinline__ void zs_mfma_cells(double *__restrict__ Q, const StLay &L, double gamma, size_t k, bool live,
                                              bool ext)
{
    constexpr int NQ = P + 1, NM = NQ * NQ;
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const bool slot = g <= P;
    const size_t so = (size_t)(slot ? g : 0) * L.ps + L.cell(k);  // this lane's element of plane (c, v)
    double C[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            C[c][v] = (v <= P && slot) ? (Q + (size_t)(c * NM + v * NQ) * L.ps)[so] : 0.0;
    const bool changed = zs_mfma_apply<P>(C, gamma, ext);
    if (live && changed && slot)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int v = 0; v <= P; ++v) (Q + (size_t)(c * NM + v * NQ) * L.ps)[so] = C[c][v];
}

// TVB limiter then positivity scaling in the matrix-core layout of the fused
// kernel: a wave owns 16 consecutive cells of the update range, lane (g, col)
// holds modes (px = v, py = g) of cell col.  The TVB test of a cell runs in its
// lane of group 0 (mean and x-linear mode there, the y-linear mode shuffled in
// from group 1, and at P >= 3 the (0,q) modes from group q for the face-average
// deviations) and its outcome is broadcast down the column; the positivity
// scaling evaluates rho and then all variables at the 4x4 points with MFMA
// products and takes the minimum theta across the column's lanes.  Same
// results as dg_tvb_kernel up to the summation order of the point values.
//
// With zsok (written by the fused stage kernel) the kernel first loads only
// what the TVB test reads (means and the modes of degree 0 on one axis); a wave whose cells are
// all unlimited and certified positive by zsok ends there, and only the other
// waves load every mode for the limiting and the positivity scaling.
// The TVB + positivity work of the 16 cells of one wave (cell k of column
// lane col; live is false on padding lanes, which compute on a valid cell and
// store nothing).  The whole wave must call it.
template <int P>
__device__ __forceinline__ void tvb_mfma_cells(double *__restrict__ Q, const double *__restrict__ Mn,
                                               const int32_t *__restrict__ src, const int32_t *__restrict__ bct,
                                               uint8_t *__restrict__ flags, int nx, size_t np, const StLay &L,
                                               double Mh2, double gamma, const uint8_t *__restrict__ zsok, size_t k,
                                               bool live,
                                               bool lin, bool ext = false)
{
    constexpr int NQ = P + 1, NM = NQ * NQ;
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const bool slot = g <= P;
    const size_t so = (size_t)(slot ? g : 0) * L.ps + L.cell(k);  // this lane's element of plane (c, v)
    double C[4][4];
    // Modes the TVB test reads: (0,0), (1,0) in group 0 and (0,1) in group 1;
    // at P >= 3 every mode (q,0) in group 0 and every mode (0,q) in group q.
    auto tvb_reads = [g](int v) {
        return P >= 3 ? ((g == 0 && v <= P) || (v == 0 && g <= P)) : ((g == 0 && v <= 1) || (g == 1 && v == 0));
    };
    // phase 1: the modes the TVB test reads
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            C[c][v] = tvb_reads(v) ? (Q + (size_t)(c * NM + v * NQ) * L.ps)[so] : 0.0;

    // ---- TVB (cells with bctype 0) ----
    int chg = 0;  // bit c: variable c limited
    double axl[4], ayl[4];
    const bool tvb_cell = bct[k] == 0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double ay_in = __shfl(C[c][0], 16 + col);   // mode (0, 1) of this column's cell
        const double ub = C[c][0], ax = C[c][1];          // in group 0: modes (0, 0) and (1, 0)
        double xp = ax, xm = ax, yp = ay_in, ym = ay_in;
        if constexpr (P >= 3) {
            // face-average deviations: modes (q, 0) in this lane, (0, q) from group q
            const double qx[3] = {C[c][1], C[c][2], C[c][3]};
            const double qy[3] = {ay_in, __shfl(C[c][0], 32 + col), __shfl(C[c][0], 48 + col)};
            face_devs<3>(qx, xp, xm);
            face_devs<3>(qy, yp, ym);
        }
        double a_x = ax, a_y = ay_in;
        if (g == 0 && tvb_cell) {
            if (tvb_axes(ub, ax, ay_in, P >= 3 && !lin, xp, xm, yp, ym, Mn, src, np, k, nx, c, Mh2, a_x, a_y))
                chg |= 1 << c;
        }
        axl[c] = __shfl(a_x, col);
        ayl[c] = __shfl(a_y, col);
    }
    chg = __shfl(chg, col);
    {
        const bool sure = zsok && (zsok[k] & 1);  // positivity scaling certified to be a no-op
        const bool need = live && (chg != 0 || !sure);
        if (!__any(need)) {
            if (live && g == 0) flags[k] = 0;
            return;
        }
    }
    // phase 2: every mode
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (!tvb_reads(v) && v <= P && slot) C[c][v] = (Q + (size_t)(c * NM + v * NQ) * L.ps)[so];
        }
#pragma unroll
    for (int c = 0; c < 4; ++c)
        if (chg & (1 << c))
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double nv = 0.0;
                if (v == 0 && g == 0) nv = C[c][0];
                else if (v == 1 && g == 0) nv = axl[c];
                else if (v == 0 && g == 1) nv = ayl[c];
                C[c][v] = nv;
            }

    // ---- positivity scaling (every cell of the range), as zs_cell ----
    const bool changed = zs_mfma_apply<P>(C, gamma, ext);
    if (!live) return;
    if (g == 0) flags[k] = (uint8_t)(chg != 0);
    if ((chg != 0 || changed) && slot)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int v = 0; v <= P; ++v) (Q + (size_t)(c * NM + v * NQ) * L.ps)[so] = C[c][v];
}

template <int P>
__global__ __launch_bounds__(256) void dg_tvb_mfma_kernel(double *__restrict__ Q, const double *__restrict__ Mn,
                                                          const int32_t *__restrict__ src,
                                                          const int32_t *__restrict__ bct, uint8_t *__restrict__ flags,
                                                          int ny, int nx, size_t np, StLay L, double Mh2,
                                                          double gamma, const uint8_t *__restrict__ zsok,
                                                          const StepScalars *__restrict__ S, bool lin, bool ext)
{
    if (S->done) return;
    const int nxu = nx - 6, nyu = ny - 6;
    const size_t ncell = (size_t)nxu * nyu;
    const int col = threadIdx.x & 15;
    const size_t t0 = ((size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16;
    if (t0 >= ncell) return;  // whole wave
    const bool live = t0 + col < ncell;
    const size_t t = live ? t0 + col : ncell - 1;
    const size_t k = interior_cell(t, nx, nxu);
    tvb_mfma_cells<P>(Q, Mn, src, bct, flags, nx, np, L, Mh2, gamma, zsok, k, live, lin, ext);
}

// Lists the cells the limiter pass must treat, from the byte the fused stage
// kernel leaves per cell (bit 0: positivity scaling certified to be a no-op,
// bit 1: TVB candidate) and the updated-cell mask.  list gets the updated
// candidates and the uncertified cells; all other cells get flag 0 and are
// done.  cand (if not null) gets the updated candidates alone, counted in
// *count.  One thread per cell, one atomic per list and wavefront; the
// counts (entry 0) must be zero on entry.
__device__ __forceinline__ void wave_append(bool in, int32_t *__restrict__ list, unsigned long long *count,
                                            int32_t k)
{
    const unsigned long long mask = __ballot(in);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    const int n = __popcll(mask);
    int base = 0;
    if (lane == leader) {
        base = atomicAdd(list, n);
        if (count) atomicAdd(count, (unsigned long long)n);
    }
    base = __shfl(base, leader);
    if (in) list[1 + base + __popcll(mask & ((1ull << lane) - 1ull))] = k;
}

__global__ __launch_bounds__(256) void dg_list_kernel(const uint8_t *__restrict__ zsok,
                                                      const uint8_t *__restrict__ upd, uint8_t *__restrict__ flags,
                                                      int nx, int nxu, size_t ncell, int32_t *__restrict__ list,
                                                      int32_t *__restrict__ cand,
                                                      unsigned long long *__restrict__ count,
                                                      const StepScalars *__restrict__ S)
{
    if (S->done) return;  // uniform over the grid
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool need = false, is_cand = false;
    size_t k = 0;
    if (t < ncell) {
        k = interior_cell(t, nx, nxu);
        const unsigned z = zsok[k];
        is_cand = (z & 2) && upd[k];
        need = is_cand || !(z & 1);
        if (!need) flags[k] = 0;
    }
    wave_append(need, list, nullptr, (int32_t)k);
    if (cand) wave_append(is_cand, cand, count, (int32_t)k);
}

// Index of the want-th set bit (from 0) of mask, want < popcount(mask): the
// largest p with at most want set bits below p, by binary search.
__device__ __forceinline__ int nth_set_bit(unsigned long long mask, int want)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1)
        if (__popcll(mask & ((1ull << (pos + w)) - 1ull)) <= want) pos += w;
    return pos;
}

// TVB limiter after the fused stage, without a list: a wave scans 64 cells
// of the update range (one per lane) from the byte the fused kernel left per
// cell, gives flag 0 to those that need nothing (certified and not an updated
// candidate), and treats the others itself, 16 at a time, with the
// matrix-core TVB + positivity routine.  Limiting changes no mean, so the
// cells are independent and the order does not matter.
template <int P, int CW>
__global__ __launch_bounds__(256) void dg_tvb_scan_kernel(double *__restrict__ Q, const double *__restrict__ Mn,
                                                          const int32_t *__restrict__ src,
                                                          const int32_t *__restrict__ bct,
                                                          const uint8_t *__restrict__ upd,
                                                          uint8_t *__restrict__ flags,
                                                          const uint8_t *__restrict__ zsok, int nx, int nxu,
                                                          size_t ncell, size_t np, StLay L, double Mh2,
                                                          double gamma, const StepScalars *__restrict__ S, RingArgs R,
                                                          unsigned nscan, bool lin, bool ext)
{
    if (S->done) return;  // uniform over the grid
    if (blockIdx.x >= nscan) {
        // blocks past the scan combine the ghost ring (when the limiter reads
        // no ring cell, so the two parts are independent)
        ring_item(R, (size_t)(blockIdx.x - nscan) * blockDim.x + threadIdx.x);
        return;
    }
    const int lane = threadIdx.x & 63, col = lane & 15;
    const size_t t0 = ((size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * CW;
    if (t0 >= ncell) return;  // whole wave
    const size_t t = t0 + lane;
    bool need = false;
    size_t k = 0;
    if (lane < CW && t < ncell) {
        k = interior_cell(t, nx, nxu);
        const unsigned z = zsok[k];
        need = ((z & 2) && upd[k]) || !(z & 1);
        if (!need) flags[k] = 0;
    }
    const unsigned long long mask = __ballot(need);
    const int n = __popcll(mask);
    for (int g0 = 0; g0 < n; g0 += 16) {
        // column col of this group takes listed cell g0 + col (padding repeats the last one)
        const int want = g0 + col < n ? g0 + col : n - 1;
        const size_t kk = (size_t)__shfl((int)k, nth_set_bit(mask, want));
        tvb_mfma_cells<P>(Q, Mn, src, bct, flags, nx, np, L, Mh2, gamma, nullptr, kk, g0 + col < n, lin, ext);
    }
}

// TVB limiter after the fused stage, one lane per cell.  A wave takes 64
// consecutive cells of the update range and reads their marks.  An updated
// candidate (bit 1) gets the TVB test in its own lane, from the modes the test
// reads (coalesced across the wave), and a limited variable gets its linear
// modes replaced and its other modes zeroed; every cell gets its flag.  The
// positivity scaling follows (the oracle's order: TVB on every cell, then the
// scaling): a certified cell (bit 0: its positivity bounds clear the floors)
// needs none even when the TVB test limited it, because limiting keeps the
// means and never raises the magnitude of a mode (zeroed modes; a
// TVB-modified minmod of the linear modes), so the certificate's bounds only
// tighten.  The few uncertified cells are listed (dcount[stage], dcells; one
// atomic per wave) for dg_tvb_defer_kernel, which applies the scaling to them
// with the matrix-core routine; keeping that routine out of this kernel keeps
// its registers low, so more waves hide the latency of the neighbour-mean
// loads.  Results equal the 16-cell scan's, and the oracle's.
template <int P>
__global__ __launch_bounds__(256) void dg_tvb_lane_kernel(double *__restrict__ Q, const double *__restrict__ Mn,
                                                          const int32_t *__restrict__ src,
                                                          const uint8_t *__restrict__ upd,
                                                          uint8_t *__restrict__ flags,
                                                          const uint8_t *__restrict__ zsok, int nx, int nxu,
                                                          size_t ncell, size_t np, StLay L, double Mh2,
                                                          const StepScalars *__restrict__ S, RingArgs R,
                                                          unsigned nscan, unsigned *__restrict__ dcount,
                                                          int32_t *__restrict__ dcells, bool lin)
{
    constexpr int NQ = P + 1, NM = NQ * NQ;
    if (S->done) return;  // uniform over the grid
    if (blockIdx.x >= nscan) {
        ring_item(R, (size_t)(blockIdx.x - nscan) * blockDim.x + threadIdx.x);
        return;
    }
    const int lane = threadIdx.x & 63;
    const size_t t0 = ((size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
    if (t0 >= ncell) return;  // whole wave
    const size_t t = t0 + lane;
    bool defer = false;
    size_t k = 0;
    if (t < ncell) {
        int ii, jj;
        interior_rc(t, nxu, ii, jj);
        k = (size_t)(ii + 3) * nx + (jj + 3);
        const size_t sk = L.rc(ii + 3, jj + 3);  // the cell in the state arrays
        const unsigned z = zsok[k];
        defer = !(z & 1);
        {
            bool lim = false;
            if ((z & 2) && upd[k]) {
                // every load first (the stores below would otherwise order them):
                // the modes the test reads (the mean and the modes of degree 0 on
                // one axis, x = (q,0), y = (0,q)) and the four neighbour means
                constexpr int NT = P >= 3 ? P : 1;
                double ub[4], X[4][NT], Y[4][NT], mE[4], mW[4], mS[4], mN[4];
                const int32_t eE = src[k + 1], eW = src[k - 1], eS = src[k + nx], eN = src[k - nx];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double *q = Q + (size_t)(c * NM) * L.ps + sk;
                    ub[c] = q[0];
#pragma unroll
                    for (int d = 1; d <= NT; ++d) {
                        X[c][d - 1] = q[(size_t)(d * NQ) * L.ps];
                        Y[c][d - 1] = q[(size_t)d * L.ps];
                    }
                    auto nbm = [&](int32_t e) {
                        double v = Mn[(size_t)c * np + (size_t)(e >> kSrcShift)];
                        if ((c == 1 && (e & 1)) || (c == 2 && (e & 2))) v = -v;
                        return v;
                    };
                    mE[c] = nbm(eE);
                    mW[c] = nbm(eW);
                    mS[c] = nbm(eS);
                    mN[c] = nbm(eN);
                }
                bool chg[4];
                double axl[4], ayl[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    double xp = X[c][0], xm = X[c][0], yp = Y[c][0], ym = Y[c][0];
                    if constexpr (P >= 3) {
                        face_devs<P>(X[c], xp, xm);
                        face_devs<P>(Y[c], yp, ym);
                    }
                    chg[c] = tvb_axes_nb(ub[c], X[c][0], Y[c][0], P >= 3 && !lin, xp, xm, yp, ym, mE[c], mW[c],
                                         mS[c], mN[c], Mh2, axl[c], ayl[c]);
                    lim = lim || chg[c];
                }
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (chg[c])
#pragma unroll
                        for (int m = 1; m < NM; ++m)
                            Q[(size_t)(c * NM + m) * L.ps + sk] = m == NQ ? axl[c] : (m == 1 ? ayl[c] : 0.0);
            }
            flags[k] = (uint8_t)lim;
        }
    }
    const unsigned long long mask = __ballot(defer);
    if (!mask) return;
    const int leader = __ffsll((long long)mask) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(dcount, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    if (defer) dcells[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)k;
}

// The uncertified cells the lane kernel listed (dcount[slot] of them), already
// limited: the matrix-core positivity scaling, 16 cells per wave per pass, over
// a grid-strided range.  Block 0 clears the next stage's count (the counts
// rotate over the three stages, so a captured step needs no memset).
template <int P>
__global__ __launch_bounds__(256) void dg_tvb_defer_kernel(double *__restrict__ Q, StLay L, double gamma,
                                                           const StepScalars *__restrict__ S,
                                                           unsigned *__restrict__ dcount, int slot,
                                                           const int32_t *__restrict__ dcells, bool ext)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) dcount[(slot + 1) % 3] = 0u;
    if (S->done) return;  // uniform over the grid
    const unsigned n = dcount[slot];
    const int col = threadIdx.x & 15;
    const unsigned waves = gridDim.x * (blockDim.x >> 6);
    for (unsigned g0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16u; g0 < n; g0 += waves * 16u) {
        const unsigned i = g0 + col < n ? g0 + col : n - 1;  // padding lanes repeat the last cell, store nothing
        zs_mfma_cells<P>(Q, L, gamma, (size_t)dcells[i], g0 + col < n, ext);
    }
}

// ---------------------------------------------------------------------------
// Learned troubled-cell indicator (P = 2)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ c * 0xC2B2AE3Du;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// A cell can only be changed by the TVB test when one of its linear modes is
// above M h^2 (the minmod is skipped otherwise); such cells are the
// candidates the network is asked about.  NaN counts as a candidate, as the
// test treats it.  At P >= 3 the face-average deviations sum P = 3 modes of
// degree 0 on one axis, so a candidate has, for some variable, the sum of
// |a(q,0)| above M h^2 or some |a(0,q)| above M h^2 / 4.
template <int P>
__device__ __forceinline__ bool tvb_candidate(const double C[4][(P + 1) * (P + 1)], double Mh2)
{
    bool c = false;
    if constexpr (P >= 3) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double sx = 0.0;
#pragma unroll
            for (int d = 1; d <= P; ++d) {
                sx += fabs(C[v][d * (P + 1)]);
                c = c || !(4.0 * fabs(C[v][d]) <= Mh2);
            }
            c = c || !(sx <= Mh2);
        }
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) c = c || !(fabs(C[v][P + 1]) <= Mh2) || !(fabs(C[v][1]) <= Mh2);
    }
    return c;
}

__global__ void zero_i32_kernel(int32_t *p) { *p = 0; }

// Compacts the candidate cells of the stage into cand[1..cand[0]] (one
// atomic per wavefront) and counts them in stats[6].
template <int P>
__global__ __launch_bounds__(256) void dg_cand_kernel(const double *__restrict__ Q, const int32_t *__restrict__ bct,
                                                      int nx, int nxu, size_t ncell, StLay L, double Mh2,
                                                      int32_t *__restrict__ cand,
                                                      unsigned long long *__restrict__ stats,
                                                      const StepScalars *__restrict__ S)
{
    constexpr int NM = (P + 1) * (P + 1);
    if (S->done) return;  // uniform over the grid
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool is = false;
    size_t k = 0;
    if (t < ncell) {
        k = interior_cell(t, nx, nxu);
        if (bct[k] == 0) {
            double C[4][NM];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int d = 1; d <= (P >= 3 ? P : 1); ++d) {
                    C[c][d] = Q[L.at(c * NM + d, k)];
                    C[c][d * (P + 1)] = Q[L.at(c * NM + d * (P + 1), k)];
                }
            is = tvb_candidate<P>(C, Mh2);
        }
    }
    const unsigned long long mask = __ballot(is);
    if (!mask) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    const int n = __popcll(mask);
    int base = 0;
    if (lane == leader) {
        base = atomicAdd(cand, n);
        atomicAdd(stats + 6, (unsigned long long)n);
    }
    base = __shfl(base, leader);
    if (is) cand[1 + base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)k;
}

// Scores the compacted candidates, 16 per wavefront per pass (MFMA), into
// score[k]; a persistent grid walks the list.
__global__ __launch_bounds__(256) void dg_score_list_kernel(const double *__restrict__ Q,
                                                            const int32_t *__restrict__ src, int nx, StLay L,
                                                            const int32_t *__restrict__ cand,
                                                            const __bf16 *__restrict__ w1b,
                                                            const __bf16 *__restrict__ w2p, const float *__restrict__ w,
                                                            float *__restrict__ score, const StepScalars *__restrict__ S)
{
    if (S->done) return;
    const int n = cand[0];
    const int waves = gridDim.x * (blockDim.x >> 6);
    const int i = threadIdx.x & 15;
    for (int base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16; base < n; base += waves * 16) {
        const int t = base + i < n ? base + i : n - 1;
        const size_t k = (size_t)cand[1 + t];
        const float sc = mlp_score16(Q, src, L, nx, k, w1b, w2p, w);
        if ((threadIdx.x & 63) < 16 && base + i < n) score[k] = sc;
    }
}

// One cell of the learned limiter (see dg_learned_kernel).
template <int P>
__device__ __forceinline__ void learned_cell(double *__restrict__ Q, const double *__restrict__ M,
                                             const int32_t *__restrict__ src, const int32_t *__restrict__ bct,
                                             uint8_t *__restrict__ flags, const float *__restrict__ score, int nx,
                                             size_t np, const StLay &L, double Mh2, double gamma, float lo, float hi,
                                             uint32_t audit_thresh, uint32_t salt,
                                             unsigned long long *__restrict__ stats,
                                             const uint8_t *__restrict__ zsok, size_t k, bool ext)
{
    constexpr int NQ = P + 1, NM = NQ * NQ;
    const size_t sk = L.cell(k);  // the cell in the state arrays
    double C[4][NM];
    // the mean and linear modes first: a non-candidate cell whose positivity
    // scaling is certified a no-op (zsok) is left as it is
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        C[c][0] = Q[(size_t)(c * NM) * L.ps + sk];
        C[c][1] = Q[(size_t)(c * NM + 1) * L.ps + sk];
        C[c][NQ] = Q[(size_t)(c * NM + NQ) * L.ps + sk];
    }
    const bool cand = bct[k] == 0 && tvb_candidate<P>(C, Mh2);
    if (!cand && zsok && (zsok[k] & 1)) {
        flags[k] = 0;
        return;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int m = 0; m < NM; ++m)
            if (m != 0 && m != 1 && m != NQ) C[c][m] = Q[(size_t)(c * NM + m) * L.ps + sk];
    int flag = 0;
    if (cand) {
        const float sc = score[k];
        const bool audit = audit_thresh == 0xFFFFFFFFu || hash3((uint32_t)k, salt, 0x5bd1e995u) < audit_thresh;
        const bool band = sc >= lo && sc <= hi;
        double axl[4], ayl[4];
        bool chg[4];
        if (audit || band) {
            const bool tv = tvb_test<P>(C, M, src, np, k, nx, Mh2, axl, ayl, chg);
            if (audit) {
                const bool net = sc > 0.5f;
                atomicAdd(stats + 0, 1ull);
                if (tv) atomicAdd(stats + 7, 1ull);
                if (net == tv) atomicAdd(stats + 1, 1ull);
                else if (tv) atomicAdd(stats + 2, 1ull);
                else atomicAdd(stats + 3, 1ull);
            }
            if (band) atomicAdd(stats + 5, 1ull);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (chg[c]) {
#pragma unroll
                    for (int m = 1; m < NM; ++m) C[c][m] = 0.0;
                    C[c][NQ] = axl[c];
                    C[c][1] = ayl[c];
                    flag = 1;
                }
        } else if (sc > hi) {
            tvb_test<P>(C, M, src, np, k, nx, Mh2, axl, ayl, chg);
            atomicAdd(stats + 4, 1ull);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int m = 1; m < NM; ++m) C[c][m] = 0.0;
                C[c][NQ] = axl[c];
                C[c][1] = ayl[c];
            }
            flag = 1;
        }
    }
    const bool changed = zs_cell<P>(C, gamma, ext);
    flags[k] = (uint8_t)flag;
    if (flag || changed) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int m = 0; m < NM; ++m) Q[(size_t)(c * NM + m) * L.ps + sk] = C[c][m];
    }
}

// Learned limiter.  Non-candidate cells pass unchanged, as they pass the TVB
// test.  For candidates, outside the uncertainty band [lo, hi] the score
// decides (troubled = score > hi, then every variable is limited); inside the
// band and on a random audit fraction the plain TVB test decides.  Audited
// cells also record agreement of the network's 0.5 decision with the TVB
// test.  The positivity scaling follows, as in the TVB limiter.
// LIST: the threads walk list[1..list[0]] (grid-stride) instead of the grid;
// the cells not listed have been finished by the list pass.
template <int P, bool LIST>
__global__ __launch_bounds__(128) void dg_learned_kernel(double *__restrict__ Q, const double *__restrict__ M,
                                                         const int32_t *__restrict__ src, const int32_t *__restrict__ bct,
                                                         uint8_t *__restrict__ flags, const float *__restrict__ score,
                                                         int ny, int nx, size_t np, StLay L, double Mh2, double gamma,
                                                         float lo, float hi, uint32_t audit_thresh, uint32_t salt,
                                                         unsigned long long *__restrict__ stats,
                                                         const uint8_t *__restrict__ zsok,
                                                         const int32_t *__restrict__ list,
                                                         const StepScalars *__restrict__ S, bool ext)
{
    if (S->done) return;
    const int nxu = nx - 6, nyu = ny - 6;
    const size_t n = LIST ? (size_t)list[0] : (size_t)nxu * nyu;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t k = LIST ? (size_t)list[1 + t] : interior_cell(t, nx, nxu);
        learned_cell<P>(Q, M, src, bct, flags, score, nx, np, L, Mh2, gamma, lo, hi, audit_thresh, salt, stats,
                        LIST ? nullptr : zsok, k, ext);
    }
}

// Harvest (features, TVB flag) pairs of the combined (pre-limit) stage state,
// from candidate cells only (the cells the network is asked about): positives
// are kept with probability keep_pos, negatives with keep_neg.  Slots past the
// capacity are dropped.
template <int P>
__global__ __launch_bounds__(256) void dg_harvest_kernel(const double *__restrict__ Q, const double *__restrict__ M,
                                                         const int32_t *__restrict__ src, const int32_t *__restrict__ bct,
                                                         int ny, int nx, size_t np, StLay L, double Mh2, uint32_t keep_pos,
                                                         uint32_t keep_neg, uint32_t salt, __bf16 *__restrict__ X,
                                                         float *__restrict__ y, int *__restrict__ count, int cap,
                                                         const StepScalars *__restrict__ S)
{
    constexpr int NM = (P + 1) * (P + 1);
    const int nxu = nx - 6, nyu = ny - 6;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nxu * nyu || S->done) return;
    const size_t k = interior_cell(t, nx, nxu);
    if (bct[k] != 0) return;
    double C[4][NM];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int m = 0; m < NM; ++m) C[c][m] = Q[L.at(c * NM + m, k)];
    double axl[4], ayl[4];
    bool chg[4];
    const bool tv = tvb_test<P>(C, M, src, np, k, nx, Mh2, axl, ayl, chg);
    if (!tvb_candidate<P>(C, Mh2)) return;  // (a TVB positive is always a candidate)
    const uint32_t keep = tv ? keep_pos : keep_neg;
    if (keep != 0xFFFFFFFFu && hash3((uint32_t)k, salt, 0x27d4eb2fu) >= keep) return;
    if (__atomic_load_n(count, __ATOMIC_RELAXED) >= cap) return;  // bounds the counter's overshoot
    const int slot = atomicAdd(count, 1);
    if (slot < 0 || slot >= cap) return;
    double mean[4], scale[4];
    mlp_scales(Q, L, k, mean, scale);
    for (int f = 0; f < 192; ++f) X[(size_t)slot * 192 + f] = (__bf16)mlp_feature(Q, src, L, nx, k, mean, scale, f);
    y[slot] = tv ? 1.0f : 0.0f;
}

// Constant tables are uploaded once per device (the first DG handle there).
std::mutex g_tables_mutex;
bool g_tables_ready[64] = {false};

template <int P>
void dg_stage_p(Handle &H, const double *Qn, const double *Qin, double *Qout, int stage)
{
    DgArgs A;
    A.Qn = Qn; A.Qin = Qin; A.Qout = Qout;
    A.src = H.src; A.upd = H.upd; A.bct = H.bct; A.S = H.scal;
    A.gx = H.gx; A.gy = H.gy;
    A.means = (H.cfg.limiter == ZK_LIMITER_TVB || H.cfg.limiter == ZK_LIMITER_LEARNED) ? H.means : nullptr;
    A.zsok = nullptr;  // set below when the fused kernel runs under the TVB limiter
    A.Mh2 = H.cfg.tvb_M * H.g.da * H.g.da;
    A.ny = H.g.ny; A.nx = H.g.nx; A.nxu = H.g.nx - 6; A.nyu = H.g.ny - 6;
    A.np = H.g.np;
    A.L = H.lay;
    A.nfx = (size_t)(A.nxu + 1) * A.nyu;
    A.nfy = (size_t)A.nxu * (A.nyu + 1);
    A.da = H.g.da; A.gamma = H.g.gamma;
    A.stage = stage;
    A.ca = stage == 1 ? 0.75 : (1.0 / 3.0);
    A.cb = stage == 1 ? 0.25 : (2.0 / 3.0);
    A.flux = H.cfg.flux;
    A.limiter = H.cfg.limiter;
    A.zs_ext = H.cfg.positivity ? 1 : 0;
    const size_t ncell = (size_t)A.nxu * A.nyu;
    // variant: 0 auto (fused), -1 split with the VALU cell kernel, 2..4 split with the
    // P=3 MFMA cell kernel at that occupancy, 10+R fused with R-row segments
    // auto picks the fused kernel (workgroup form) at every P; at P = 1 its 4x4 slots
    // are three quarters padding, yet with the 4x4x4 blocks it matches the split
    // kernels at 1024^2 and is 17% faster at 2048^2
    const bool fused = H.cfg.limiter != ZK_LIMITER_PERSSON_ZS && !H.no_mfma &&
                       (H.tune == 0 || H.tune >= 10);
    if (ncell > 0 && fused) {
        // automatic: the workgroup form; 10+R / 100+R force a form and segment length
        const bool wg = H.tune == 0 || H.tune >= 100;
        int R = H.tune >= 100 ? H.tune - 100 : (H.tune >= 10 ? H.tune - 10 : 0);
        const int width = wg ? 62 : 14;                 // updated columns per wave (per workgroup for wg)
        // the kernel instance this stage launches
        // the default form with the stage and 32-bit plane offsets fixed at
        // compile time, when the state arrays are below 4 GiB; at P = 3 the
        // two later stages on aligned 64-column blocks unless ZK_FUSED_AL=0
        // (the same results bit for bit).  They read U^n as well, and the
        // aligned accesses save more there than the extra edge-flux call per
        // row costs; the first stage and P = 2 are faster without (1024^2 P=3
        // mid-run: stage 1 412 against 363 us, stages 2-3 439 against 480;
        // 2048^2 P=2 1.13 against 1.02 ms)
        const bool fast = wg && H.dma == 3 && H.mf4 && !H.fused_generic &&
                          (size_t)4 * (P + 1) * (P + 1) * H.g.np * 8 <= 0xFFFFFFFFull;
        const bool al = fast && H.fused_al && P == 3 && stage > 0;
        const int nrun = al ? (A.nx - 4) / 64 + 1 : (A.nxu + width - 1) / width;
        A.row16 = (A.nx % 2 == 0) ? 1 : 0;
        const void *kern = al                         ? (const void *)dg_fused_kernel<P, 2, true, 3, true, 2, true, true>
                           : (fast && stage == 0)     ? (const void *)dg_fused_kernel<P, 2, true, 3, true, 1, true>
                           : fast                     ? (const void *)dg_fused_kernel<P, 2, true, 3, true, 2, true>
                           : (wg && H.dma == 3 && H.mf4) ? (const void *)dg_fused_kernel<P, 2, true, 3, true>
                           : (wg && H.dma == 3)         ? (const void *)dg_fused_kernel<P, 2, true, 3, false>
                           : (wg && H.dma == 2)         ? (const void *)dg_fused_kernel<P, 2, true, 2, false>
                           : (wg && H.dma == 1)         ? (const void *)dg_fused_kernel<P, 2, true, 1, false>
                           : wg                         ? (const void *)dg_fused_kernel<P, 2, true, 0, false>
                                                        : (const void *)dg_fused_kernel<P, 2, false, 0, false>;
        if (R <= 0) {
            // Whole rounds of the device: as many workgroups (or single waves)
            // as the CUs hold at the kernel's occupancy.  The fewest rounds with
            // R <= 64 keep halo rows few and leave no part-filled last round.
            static std::atomic<int> occ_cache[9];  // per kernel instance of this P (0 = not known yet)
            const int slot_i = al ? 8 : fast ? (stage == 0 ? 6 : 7)
                               : !wg ? 5 : H.dma == 3 ? (H.mf4 ? 0 : 1) : H.dma == 2 ? 2 : H.dma == 1 ? 3 : 4;
            int per_cu = occ_cache[slot_i].load();
            if (per_cu < 1) {
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess || per_cu < 1)
                    per_cu = 2;
                occ_cache[slot_i].store(per_cu);
            }
            const long cap = (long)(H.ncu > 0 ? H.ncu : 256) * per_cu * (wg ? 1 : 4);
            for (long m = 1;; ++m) {
                const long nseg = m * cap / nrun;
                if (nseg < 1) continue;
                R = (int)((A.nyu + nseg - 1) / nseg);
                if (R <= 64 || nseg >= A.nyu) break;
            }
            R = R < 4 ? 4 : R;
        }
        const long units = (long)nrun * ((A.nyu + R - 1) / R);
        if (H.cfg.limiter == ZK_LIMITER_TVB || H.cfg.limiter == ZK_LIMITER_LEARNED) A.zsok = H.zsok;
        ProfScope ps(H, "dg_fused");
        const dim3 grid((unsigned)(wg ? units : (units + 3) / 4));
        void *args[] = {&A, &R, (void *)&nrun};
        (void)hipLaunchKernel(kern, grid, dim3(256), args, 0, H.stream);
    } else if (ncell > 0) {
        {
            ProfScope ps(H, "dg_face_x");
            hipLaunchKernelGGL((dg_face_kernel<P, true>), dim3((A.nfx + 127) / 128), dim3(128), 0, H.stream, A);
        }
        {
            ProfScope ps(H, "dg_face_y");
            hipLaunchKernelGGL((dg_face_kernel<P, false>), dim3((A.nfy + 127) / 128), dim3(128), 0, H.stream, A);
        }
        ProfScope ps(H, "dg_cell");
        if (P == 3 && H.cfg.limiter != ZK_LIMITER_PERSSON_ZS && !H.no_mfma) {
            if (H.tune == 4)
                hipLaunchKernelGGL(dg_cell_mfma3_kernel<4>, dim3((ncell + 63) / 64), dim3(256), 0, H.stream, A);
            else if (H.tune == 3)
                hipLaunchKernelGGL(dg_cell_mfma3_kernel<3>, dim3((ncell + 63) / 64), dim3(256), 0, H.stream, A);
            else
                hipLaunchKernelGGL(dg_cell_mfma3_kernel<2>, dim3((ncell + 63) / 64), dim3(256), 0, H.stream, A);
        }
        else
            hipLaunchKernelGGL((dg_cell_kernel<P>), dim3((ncell + 127) / 128), dim3(128), 0, H.stream, A);
    }
    // list counts the limiter pass below fills: cleared by the ring kernel
    int32_t *z0 = nullptr, *z1 = nullptr;
    if (ncell > 0 && H.cfg.limiter == ZK_LIMITER_LEARNED && P == 2) {
        z0 = H.cand;
        if (A.zsok) z1 = H.tlist;
    }
    const RingArgs RA{Qn, Qin, Qout, H.src, H.ring, H.nring, H.nm, H.g.np, H.lay, A.ca, A.cb, stage, A.means};
    const unsigned ring_blocks = (unsigned)(((size_t)H.nring * 4 * H.nm + 255) / 256);
    // The ghost-ring combination writes only ring cells of Qout and of the
    // means.  The TVB scan below changes updated cells and reads the means of
    // their neighbours through the ghost map; unless that map reaches a ring
    // cell (ring_read: fixed cells), the two are independent and the ring
    // rides in the scan's launch as extra blocks.  With kernel timing on, or
    // list counts to clear, it keeps its own launch.
    const bool scan = ncell > 0 && H.cfg.limiter == ZK_LIMITER_TVB && A.zsok && !H.no_mfma;
    const bool fold = scan && H.nring > 0 && !H.ring_read && !z0 && !z1 && !H.harvest && !H.prof;
    if (H.nring > 0 && !fold) {
        ProfScope ps(H, "dg_ring");
        hipLaunchKernelGGL(dg_ring_kernel, dim3(ring_blocks), dim3(256), 0, H.stream, RA, H.scal, z0, z1);
    } else if (H.nring == 0) {
        if (z0) hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, H.stream, z0);
        if (z1) hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(1), 0, H.stream, z1);
    }
    const double Mh2 = A.Mh2;
    H.stage_counter += 1;
    if (ncell > 0 && H.cfg.limiter == ZK_LIMITER_TVB && H.harvest && P == 2) {
        ProfScope ps(H, "dg_harvest");
        auto frac = [](float f) { return f >= 1.0f ? 0xFFFFFFFFu : (uint32_t)(f * 4294967296.0); };
        hipLaunchKernelGGL((dg_harvest_kernel<P>), dim3((ncell + 255) / 256), dim3(256), 0, H.stream, Qout, H.means,
                           H.src, H.bct, H.g.ny, H.g.nx, H.g.np, H.lay, Mh2, frac(H.keep_pos), frac(H.keep_neg),
                           H.stage_counter, mlp_X(H.harvest), mlp_y(H.harvest), mlp_count(H.harvest),
                           mlp_cap(H.harvest), H.scal);
    }
    if (ncell > 0 && H.cfg.limiter == ZK_LIMITER_TVB) {
        ProfScope ps(H, "dg_tvb");
        // after the fused stage (zsok written): a per-thread screen lists the cells that
        // are limited or not certified, and the matrix-core kernel treats only those;
        // otherwise the matrix-core kernel at P = 3 and the per-thread kernel at P <= 2
        if (A.zsok && !H.no_mfma && H.tvb_lane && H.dcount) {
            // one lane per cell, then the matrix-core routine for the uncertified cells
            const unsigned nscan = (unsigned)((ncell + 255) / 256);
            {
                ProfScope pl(H, "dg_tvb_lane");
                hipLaunchKernelGGL((dg_tvb_lane_kernel<P>), dim3(nscan + (fold ? ring_blocks : 0u)), dim3(256), 0,
                                   H.stream, Qout, H.means, H.src, H.upd, H.flags, A.zsok, H.g.nx, A.nxu, ncell,
                                   H.g.np, H.lay, Mh2, H.scal, RA, nscan, H.dcount + stage, H.dcells, H.tvb_linear);
            }
            const unsigned nblk = (unsigned)(H.ncu > 0 ? H.ncu : 256);
            ProfScope pd(H, "dg_tvb_defer");
            hipLaunchKernelGGL((dg_tvb_defer_kernel<P>), dim3(nblk), dim3(256), 0, H.stream, Qout, H.lay,
                               H.g.gamma, H.scal, H.dcount, stage, H.dcells, A.zs_ext != 0);
        } else if (A.zsok && !H.no_mfma) {  // scan the fused kernel's marks and limit in place
            // cells per wave: at P = 3 at most two passes of the routine per wave
            // (the passes are long); at P <= 2 the wider scan is cheaper
            constexpr int CW = P == 3 ? 32 : 64;
            const unsigned nscan = (unsigned)((ncell + 4 * CW - 1) / (4 * CW));
            hipLaunchKernelGGL((dg_tvb_scan_kernel<P, CW>), dim3(nscan + (fold ? ring_blocks : 0u)), dim3(256),
                               0, H.stream,
                               Qout, H.means, H.src, H.bct, H.upd, H.flags, A.zsok, H.g.nx, A.nxu, ncell, H.g.np, H.lay,
                               Mh2, H.g.gamma, H.scal, RA, nscan, H.tvb_linear, A.zs_ext != 0);
        } else if (H.no_mfma || P < 3)
            hipLaunchKernelGGL((dg_tvb_kernel<P>), dim3((ncell + 127) / 128), dim3(128), 0, H.stream, Qout, H.means,
                               H.src, H.bct, H.flags, H.g.ny, H.g.nx, H.g.np, H.lay, Mh2, H.g.gamma, H.scal, H.tvb_linear,
                               A.zs_ext != 0);
        else
            hipLaunchKernelGGL((dg_tvb_mfma_kernel<P>), dim3((unsigned)((ncell + 63) / 64)), dim3(256), 0, H.stream,
                               Qout, H.means, H.src, H.bct, H.flags, H.g.ny, H.g.nx, H.g.np, H.lay, Mh2, H.g.gamma, A.zsok,
                               H.scal, H.tvb_linear, A.zs_ext != 0);
    }
    if (ncell > 0 && H.cfg.limiter == ZK_LIMITER_LEARNED && P == 2) {
        {
            ProfScope ps(H, "dg_candidates");
            // list counts cleared with the ring
            if (A.zsok) {  // the fused kernel marked the candidates: list them and the uncertified cells
                hipLaunchKernelGGL(dg_list_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, H.stream,
                                   A.zsok, H.upd, H.flags, H.g.nx, A.nxu, ncell, H.tlist, H.cand, H.stats + 6, H.scal);
            } else {
                hipLaunchKernelGGL((dg_cand_kernel<P>), dim3((ncell + 255) / 256), dim3(256), 0, H.stream, Qout,
                                   H.bct, H.g.nx, A.nxu, ncell, H.lay, Mh2, H.cand, H.stats, H.scal);
            }
        }
        {
            ProfScope ps(H, "dg_score");
            const size_t want = (ncell + 63) / 64, cap = (size_t)(H.ncu > 0 ? H.ncu : 256) * 8;
            hipLaunchKernelGGL(dg_score_list_kernel, dim3((unsigned)(want < cap ? want : cap)), dim3(256), 0,
                               H.stream, Qout, H.src, H.g.nx, H.lay, H.cand, mlp_w1b(H.mlp), mlp_w2p(H.mlp),
                               mlp_w(H.mlp), H.score, H.scal);
        }
        ProfScope ps(H, "dg_learned");
        const uint32_t audit = H.mlp_audit >= 1.0f ? 0xFFFFFFFFu : (uint32_t)(H.mlp_audit * 4294967296.0);
        if (A.zsok) {
            const size_t want = (ncell + 127) / 128, cap = (size_t)(H.ncu > 0 ? H.ncu : 256) * 16;
            hipLaunchKernelGGL((dg_learned_kernel<P, true>), dim3((unsigned)(want < cap ? want : cap)), dim3(128), 0,
                               H.stream, Qout, H.means, H.src, H.bct, H.flags, H.score, H.g.ny, H.g.nx, H.g.np, H.lay,
                               Mh2, H.g.gamma, H.mlp_lo, H.mlp_hi, audit, H.stage_counter, H.stats, A.zsok, H.tlist,
                               H.scal, A.zs_ext != 0);
        } else {
            hipLaunchKernelGGL((dg_learned_kernel<P, false>), dim3((ncell + 127) / 128), dim3(128), 0, H.stream, Qout,
                               H.means, H.src, H.bct, H.flags, H.score, H.g.ny, H.g.nx, H.g.np, H.lay, Mh2,
                               H.g.gamma, H.mlp_lo, H.mlp_hi, audit, H.stage_counter, H.stats, A.zsok, H.tlist, H.scal,
                               A.zs_ext != 0);
        }
    }
}

}  // namespace

MfmaTabsX make_mfma_tables_ext(int P)
{
    MfmaTabsX X{};
    double xs[20], ys[20];
    zs_ext_points(P, GLP_H[P + 1], xs, ys);
    for (int h = 0; h < 2; ++h)
        for (int s = 0; s < 4; ++s)
            for (int l = 0; l < 64; ++l) {
                const int i = l % 16, py = l / 16, e = 16 * h + i < 20 ? 16 * h + i : 0;
                X.zxA[h][s][l] = (s <= P && py <= P) ? legP_h(s, xs[e]) * legP_h(py, ys[e]) : 0.0;
            }
    return X;
}

// Uploads the basis and quadrature tables to the current device (once per
// device); returns the first HIP error.
hipError_t dg_upload_tables(int device)
{
    std::lock_guard<std::mutex> lock(g_tables_mutex);
    const bool track = device >= 0 && device < 64;
    if (track && g_tables_ready[device]) return hipSuccess;
    const DgTables T[3] = {make_tables(1), make_tables(2), make_tables(3)};
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), T, sizeof(T), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const MfmaTabs3 M = make_mfma_tables3();
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_mf3), &M, sizeof(M), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const MfmaTabsF F[3] = {make_mfma_tables_fused(1), make_mfma_tables_fused(2), make_mfma_tables_fused(3)};
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_mff), F, sizeof(F), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const MfmaTabs4 F4[3] = {make_mfma_tables4(1), make_mfma_tables4(2), make_mfma_tables4(3)};
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_mf4), F4, sizeof(F4), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    const MfmaTabsX FX[3] = {make_mfma_tables_ext(1), make_mfma_tables_ext(2), make_mfma_tables_ext(3)};
    e = hipMemcpyToSymbol(HIP_SYMBOL(c_mfx), FX, sizeof(FX), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    if (track) g_tables_ready[device] = true;
    return hipSuccess;
}

void dg_stage(Handle &H, const double *Qn, const double *Qin, double *Qout, int stage)
{
    switch (H.cfg.P) {
    case 1: dg_stage_p<1>(H, Qn, Qin, Qout, stage); break;
    case 2: dg_stage_p<2>(H, Qn, Qin, Qout, stage); break;
    default: dg_stage_p<3>(H, Qn, Qin, Qout, stage); break;
    }
}

void dg_max_speed(Handle &H, const double *Q)
{
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, H.stream, H.lam_bits);
    int blocks = (int)((H.g.np + 255) / 256);
    if (blocks > 2048) blocks = 2048;
    switch (H.cfg.P) {
    case 1:
        hipLaunchKernelGGL((dg_max_speed_kernel<1>), dim3(blocks), dim3(256), 0, H.stream, Q, H.bct, H.g.np, H.lay,
                           H.g.gamma, H.lam_bits);
        break;
    case 2:
        hipLaunchKernelGGL((dg_max_speed_kernel<2>), dim3(blocks), dim3(256), 0, H.stream, Q, H.bct, H.g.np, H.lay,
                           H.g.gamma, H.lam_bits);
        break;
    default:
        hipLaunchKernelGGL((dg_max_speed_kernel<3>), dim3(blocks), dim3(256), 0, H.stream, Q, H.bct, H.g.np, H.lay,
                           H.g.gamma, H.lam_bits);
        break;
    }
}

}  // namespace zk
