/*
 * zk_oracle.h -- CPU oracle for the explicit 2D Euler update of ZuikakuCFD.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solvers (src_WENO/fluid.cpp, src_DG/fluid.cpp, src_DG/solver.cpp) used as the
 * checker for the HIP library and as the CPU baseline that bench.py times.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product path (libzk.so) never links or calls it.
 *
 * Layout: every field is a row-major plane of ny*nx doubles (row i = y-down row
 * of the input files, column j = x).  A state holds 4*nm planes ordered
 * (variable c, mode m): plane (c*nm + m).  nm = 1 for the finite-volume path and
 * (P+1)^2 for modal DG, mode m = px*(P+1) + py (px along x/j, py along y/i),
 * as src_DG/fluid.cpp:63-65.  The reference stores Eigen column-major
 * matrices; storage order changes no arithmetic.
 *
 * Parity status of each function is stated in its comment; every function is
 * pinned by the closed-form, invariant and brute-force tests under tests/.
 */
#ifndef ZK_ORACLE_H
#define ZK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* bctype codes: 0 updated cell, -1 copy ghost (src_WENO/fluid.h:53-58),
 * 1 slip ghost (DG only, src_DG/fluid.cpp:600-606), -2 periodic ghost (this
 * project's extension, no reference counterpart), -3 MPI ghost (never used on
 * a single domain). */

/* ---------------- point kernels ---------------- */
void zko_weno5(const double s[6], double *UL, double *UR);
void zko_hllc_fv(const double L[4], const double R[4], double gamma, double F[4]);
void zko_hllc_dg(const double L[4], const double R[4], double gamma, double F[4]);
void zko_hll(const double L[4], const double R[4], double gamma, double F[4]);
void zko_hllc_speeds(int kind, const double L[4], const double R[4], double gamma, double out[3]);

/* ---------------- finite volume (WENO5) ---------------- */
/* flux: 0 = HLLC (src_WENO/fluid.cpp:268-317), 1 = HLL (src_DG/fluid.cpp:525-580);
 * or-ed with ZKO_POSITIVITY it also applies the face positivity fallback, and
 * with r << ZKO_RECON_SHIFT it reconstructs with MUSCL (r = 1 minmod, 2 van
 * Leer, 3 superbee) instead of WENO5 (r = 0). */
#define ZKO_POSITIVITY 16
#define ZKO_RECON_SHIFT 5
/* or-ed into flux: bctype-1 cells become mirror slip walls (no reference
 * counterpart; the reference's finite-volume fill leaves them untouched) */
#define ZKO_FV_SLIP 128
void zko_muscl(const double s[4], int limiter, double *UL, double *UR);
void zko_fv_fill_ghosts(int ny, int nx, const int32_t *bct, double *U /*4 planes*/);
void zko_fv_fill_ghosts_slip(int ny, int nx, const int32_t *bct, double *U /*4 planes*/);
void zko_fv_rhs_masked(int ny, int nx, double da, double gamma, const int32_t *bct,
                       double *U, double dt, int flux, double *dU, const uint8_t *bad);
void zko_fv_rhs(int ny, int nx, double da, double gamma, const int32_t *bct,
                double *U, double dt, int flux, double *dU);
int zko_fv_step(int ny, int nx, double da, double gamma, const int32_t *bct,
                 double *U, double dt, int flux);
double zko_fv_max_speed(int ny, int nx, double gamma, const double *U);
void zko_prim_to_cons(int n, double gamma, const double *rho, const double *u,
                      const double *v, const double *p, double *U);

/* ---------------- modal DG ---------------- */
double zko_legP(int n, double x);
double zko_legD(int n, double x);
const double *zko_gl_points(int nq);
const double *zko_gl_weights(int nq);
void zko_dg_fill_ghosts(int ny, int nx, int P, const int32_t *bct, double *Q);
void zko_dg_rhs(int ny, int nx, int P, double da, double gamma, const int32_t *bct,
                double *Q, int flux, double *dQ);
/* limiter: 0 none, 1 Persson + Zhang-Shu (the reference's applyLimiter),
 * 2 TVB (Cockburn-Shu, modified minmod with constant M; at P >= 3 with the
 * face-average test) + Zhang-Shu positivity, 4 the same with the linear-mode
 * test alone at every P; or-ed with 8 (the positivity option, no reference
 * counterpart) the positivity scaling also covers the face quadrature points
 * and the cell corners, with the full basis (zko_dg_limit_zs_ext).
 * flags (optional, ny*nx bytes) receive 1 where the TVB limiter changed a cell. */
void zko_dg_limit_persson_zs(int ny, int nx, int P, double gamma, const int32_t *bct,
                             double *Q);
void zko_dg_limit_tvb(int ny, int nx, int P, double da, double M, const int32_t *bct,
                      double *Q, uint8_t *flags);
void zko_dg_limit_tvb_linear(int ny, int nx, int P, double da, double M, const int32_t *bct,
                             double *Q, uint8_t *flags);
void zko_dg_limit_zs(int ny, int nx, int P, double gamma, double *Q);
void zko_dg_limit_zs_ext(int ny, int nx, int P, double gamma, double *Q);
void zko_dg_step(int ny, int nx, int P, double da, double gamma, const int32_t *bct,
                 double *Q, double dt, int flux, int limiter, double tvb_M,
                 uint8_t *flags);
double zko_dg_max_speed(int ny, int nx, int P, double gamma, const int32_t *bct,
                        const double *Q);
double zko_dg_cfl_limit(int P);

#ifdef __cplusplus
}
#endif
#endif
