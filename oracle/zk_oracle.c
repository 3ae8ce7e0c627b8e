/*
 * zk_oracle.c -- CPU oracle (test infrastructure only; see zk_oracle.h).
 *
 * A plain-C restatement of the reference solver's arithmetic, expression by
 * expression, so that the HIP library can be checked against it.  Reference
 * citations are file:line in midway-official/ZuikakuCFD.  The reference itself
 * cannot be compiled in this image (it needs Eigen3 and MPI headers that are
 * absent), so this restatement is pinned by closed forms, invariants and
 * brute force in tests/test_oracle_*.py.
 *
 * OpenMP mirrors the reference's own parallel structure where it has one:
 * the WENO RHS loop is `parallel for schedule(dynamic,8)` over rows
 * (src_WENO/fluid.cpp:395-397).  The DG RHS is serial per rank in the
 * reference; here rows are split across threads, which changes no arithmetic
 * (each cell writes only its own dU, as in the MPI column split).
 */
#include "zk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j) ((size_t)(i) * (size_t)nx + (size_t)(j))

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
/* std::max(a,b) returns a unless a < b; std::min(a,b) returns a unless b < a. */
static inline double stdmax(double a, double b) { return (a < b) ? b : a; }
static inline double stdmin(double a, double b) { return (b < a) ? b : a; }

/* ======================================================================
 * Point kernels
 * ====================================================================== */

/* MUSCL reconstruction at the face between s[1] and s[2] from the four cells
 * s[0..3] = UL2, UL1, UR1, UR2, with limiter 0 = minmod, 1 = van Leer,
 * 2 = superbee (muscl_reconstruct, src_WENO/fluid.cpp:203-225; the reference
 * defines it but its solver never calls it).  UL = UL1 + sigma_L / 2 with
 * sigma_L = lim(UL1 - UL2, UR1 - UL1); UR = UR1 - sigma_R / 2 with
 * sigma_R = lim(UR1 - UL1, UR2 - UR1). */
static double zko_minmod2(double a, double b) { return a * b <= 0.0 ? 0.0 : (fabs(a) < fabs(b) ? a : b); }
static double zko_vanleer(double a, double b)
{
    const double ab = a * b;
    return ab <= 0.0 ? 0.0 : 2.0 * ab / (a + b);
}
static double zko_superbee(double a, double b)
{
    const double s1 = zko_minmod2(a, 2.0 * b), s2 = zko_minmod2(2.0 * a, b);
    return s1 * s2 <= 0.0 ? 0.0 : (fabs(s1) > fabs(s2) ? s1 : s2);
}
void zko_muscl(const double s[4], int limiter, double *ULo, double *URo)
{
    const double dL = s[1] - s[0], dC = s[2] - s[1], dR = s[3] - s[2];
    double sl = 0.0, sr = 0.0;
    switch (limiter) {
    case 0: sl = zko_minmod2(dL, dC); sr = zko_minmod2(dC, dR); break;
    case 1: sl = zko_vanleer(dL, dC); sr = zko_vanleer(dC, dR); break;
    case 2: sl = zko_superbee(dL, dC); sr = zko_superbee(dC, dR); break;
    default: break;
    }
    *ULo = s[1] + 0.5 * sl;
    *URo = s[2] - 0.5 * sr;
}

/* Face reconstruction of the finite-volume path from the 6-cell stencil s
 * (the face between s[2] and s[3]): recon 0 = WENO5, 1..3 = MUSCL with
 * limiter recon - 1 on s[1..4]. */
static void zko_reconstruct(int recon, const double s[6], double *UL, double *UR);

/* Jiang-Shu WENO5 at the face between s[2] and s[3]: UL from the left-biased
 * stencil s[0..4], UR from the right-biased stencil s[1..5].
 * Follows src_WENO/fluid.cpp:230-263 term by term. */
void zko_weno5(const double s[6], double *ULo, double *URo)
{
    const double UL2 = s[0], UL1 = s[1], UP = s[2], UR1 = s[3], UR2 = s[4], UR3 = s[5];
    const double eps = 1e-6, C1 = 13.0 / 12.0, C2 = 0.25;

    double qL0 = (1.0 / 3.0) * UL2 - (7.0 / 6.0) * UL1 + (11.0 / 6.0) * UP;
    double qL1 = -(1.0 / 6.0) * UL1 + (5.0 / 6.0) * UP + (1.0 / 3.0) * UR1;
    double qL2 = (1.0 / 3.0) * UP + (5.0 / 6.0) * UR1 - (1.0 / 6.0) * UR2;

    double d0 = UL2 - 2 * UL1 + UP, d1 = UL2 - 4 * UL1 + 3 * UP;
    double d2 = UL1 - 2 * UP + UR1, d3 = UL1 - UR1;
    double d4 = UP - 2 * UR1 + UR2, d5 = 3 * UP - 4 * UR1 + UR2;
    double bL0 = C1 * d0 * d0 + C2 * d1 * d1, bL1 = C1 * d2 * d2 + C2 * d3 * d3,
           bL2 = C1 * d4 * d4 + C2 * d5 * d5;

    double t0 = eps + bL0, t1 = eps + bL1, t2 = eps + bL2;
    double aL0 = 0.1 / (t0 * t0), aL1 = 0.6 / (t1 * t1), aL2 = 0.3 / (t2 * t2);
    double inv = 1.0 / (aL0 + aL1 + aL2);
    *ULo = (aL0 * qL0 + aL1 * qL1 + aL2 * qL2) * inv;

    double qR0 = (11.0 / 6.0) * UR1 - (7.0 / 6.0) * UR2 + (1.0 / 3.0) * UR3;
    double qR1 = (1.0 / 3.0) * UP + (5.0 / 6.0) * UR1 - (1.0 / 6.0) * UR2;
    double qR2 = -(1.0 / 6.0) * UL1 + (5.0 / 6.0) * UP + (1.0 / 3.0) * UR1;

    double r0 = UR1 - 2 * UR2 + UR3, r1 = 3 * UR1 - 4 * UR2 + UR3;
    double r2 = UP - 2 * UR1 + UR2, r3 = UP - UR2;
    double r4 = UL1 - 2 * UP + UR1, r5 = UL1 - 4 * UP + 3 * UR1;
    double bR0 = C1 * r0 * r0 + C2 * r1 * r1, bR1 = C1 * r2 * r2 + C2 * r3 * r3,
           bR2 = C1 * r4 * r4 + C2 * r5 * r5;

    double s0 = eps + bR0, s1 = eps + bR1, s2 = eps + bR2;
    double aR0 = 0.1 / (s0 * s0), aR1 = 0.6 / (s1 * s1), aR2 = 0.3 / (s2 * s2);
    double inv2 = 1.0 / (aR0 + aR1 + aR2);
    *URo = (aR0 * qR0 + aR1 * qR1 + aR2 * qR2) * inv2;
}

static void zko_reconstruct(int recon, const double s[6], double *UL, double *UR)
{
    if (recon == 0) zko_weno5(s, UL, UR);
    else zko_muscl(s + 1, recon - 1, UL, UR);
}

/* HLLC flux of the finite-volume solver, normal velocity = component 1.
 * Follows src_WENO/fluid.cpp:268-317 (eps 1e-12 floors, Roe/Davis speeds
 * with c_roe from u_roe only, early supersonic returns, den2 guard). */
void zko_hllc_fv(const double L[4], const double R[4], double gamma, double F[4])
{
    const double eps = 1e-12;
    double rhoL = stdmax(L[0], eps), uL = L[1] / rhoL, vL = L[2] / rhoL, EL = L[3] / rhoL;
    double pL = stdmax((gamma - 1.0) * (L[3] - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    double aL = sqrt(gamma * pL / rhoL);

    double rhoR = stdmax(R[0], eps), uR = R[1] / rhoR, vR = R[2] / rhoR, ER = R[3] / rhoR;
    double pR = stdmax((gamma - 1.0) * (R[3] - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    double aR = sqrt(gamma * pR / rhoR);

    double FL0 = rhoL * uL, FL1 = rhoL * uL * uL + pL, FL2 = rhoL * uL * vL, FL3 = (L[3] + pL) * uL;
    double FR0 = rhoR * uR, FR1 = rhoR * uR * uR + pR, FR2 = rhoR * uR * vR, FR3 = (R[3] + pR) * uR;

    double sqL = sqrt(rhoL), sqR = sqrt(rhoR), den = sqL + sqR;
    double u_roe = (sqL * uL + sqR * uR) / den;
    double h_roe = (sqL * (EL + pL / rhoL) + sqR * (ER + pR / rhoR)) / den;
    double c_roe = sqrt(stdmax((gamma - 1.0) * (h_roe - 0.5 * u_roe * u_roe), eps));

    double SL = stdmin(uL - aL, u_roe - c_roe);
    double SR = stdmax(uR + aR, u_roe + c_roe);

    if (SL >= 0.0) { F[0] = FL0; F[1] = FL1; F[2] = FL2; F[3] = FL3; return; }
    if (SR <= 0.0) { F[0] = FR0; F[1] = FR1; F[2] = FR2; F[3] = FR3; return; }

    double num = pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR);
    double den2 = rhoL * (SL - uL) - rhoR * (SR - uR);
    if (fabs(den2) < 1e-12) den2 = (den2 >= 0 ? 1e-12 : -1e-12);
    double Sstar = num / den2;

    if (Sstar >= 0.0) {
        double c = rhoL * (SL - uL) / (SL - Sstar);
        double s0 = c, s1 = c * Sstar, s2 = c * vL,
               s3 = c * (EL + (Sstar - uL) * (Sstar + pL / (rhoL * (SL - uL))));
        F[0] = FL0 + SL * (s0 - L[0]); F[1] = FL1 + SL * (s1 - L[1]);
        F[2] = FL2 + SL * (s2 - L[2]); F[3] = FL3 + SL * (s3 - L[3]);
    } else {
        double c = rhoR * (SR - uR) / (SR - Sstar);
        double s0 = c, s1 = c * Sstar, s2 = c * vR,
               s3 = c * (ER + (Sstar - uR) * (Sstar + pR / (rhoR * (SR - uR))));
        F[0] = FR0 + SR * (s0 - R[0]); F[1] = FR1 + SR * (s1 - R[1]);
        F[2] = FR2 + SR * (s2 - R[2]); F[3] = FR3 + SR * (s3 - R[3]);
    }
}

/* Wave-speed estimates (SL, SR, S*) of the two HLLC variants, exposed so the
 * tests can check them against closed forms.  kind 0: finite volume
 * (src_WENO/fluid.cpp:274-300), kind 1: DG (src_DG/fluid.cpp:456-496). */
void zko_hllc_speeds(int kind, const double L[4], const double R[4], double gamma, double out[3])
{
    const double eps = kind == 0 ? 1e-12 : 1e-5;
    double rhoL = stdmax(L[0], eps), uL = L[1] / rhoL, vL = L[2] / rhoL, EL = L[3] / rhoL;
    double pL = stdmax((gamma - 1.0) * (L[3] - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    double rhoR = stdmax(R[0], eps), uR = R[1] / rhoR, vR = R[2] / rhoR, ER = R[3] / rhoR;
    double pR = stdmax((gamma - 1.0) * (R[3] - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    double aL, aR;
    if (kind == 0) { aL = sqrt(gamma * pL / rhoL); aR = sqrt(gamma * pR / rhoR); }
    else { aL = sqrt(stdmax(gamma * pL / rhoL, eps)); aR = sqrt(stdmax(gamma * pR / rhoR, eps)); }
    double sqL = sqrt(rhoL), sqR = sqrt(rhoR), den = sqL + sqR;
    if (kind == 1 && den < eps) den = eps;
    double u_roe = (sqL * uL + sqR * uR) / den;
    double h_roe = (sqL * (EL + pL / rhoL) + sqR * (ER + pR / rhoR)) / den;
    double c_roe = sqrt(stdmax((gamma - 1.0) * (h_roe - 0.5 * u_roe * u_roe), eps));
    double SL = stdmin(uL - aL, u_roe - c_roe);
    double SR = stdmax(uR + aR, u_roe + c_roe);
    double num = pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR);
    double den2 = rhoL * (SL - uL) - rhoR * (SR - uR);
    double g = kind == 0 ? 1e-12 : eps;
    if (fabs(den2) < g) den2 = (den2 >= 0 ? g : -g);
    out[0] = SL;
    out[1] = SR;
    out[2] = num / den2;
}

/* HLLC flux of the DG solver: eps 1e-5 floors, sound-speed floor, Roe
 * denominator guard, S - S* guard and NaN/inf guard on the star term.
 * Follows src_DG/fluid.cpp:450-520. */
void zko_hllc_dg(const double L[4], const double R[4], double gamma, double F[4])
{
    const double eps = 1e-5;
    double rhoL = stdmax(L[0], eps), uL = L[1] / rhoL, vL = L[2] / rhoL, EL = L[3] / rhoL;
    double pL = stdmax((gamma - 1.0) * (L[3] - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    double aL = sqrt(stdmax(gamma * pL / rhoL, eps));

    double rhoR = stdmax(R[0], eps), uR = R[1] / rhoR, vR = R[2] / rhoR, ER = R[3] / rhoR;
    double pR = stdmax((gamma - 1.0) * (R[3] - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    double aR = sqrt(stdmax(gamma * pR / rhoR, eps));

    double FL0 = rhoL * uL, FL1 = rhoL * uL * uL + pL, FL2 = rhoL * uL * vL, FL3 = (L[3] + pL) * uL;
    double FR0 = rhoR * uR, FR1 = rhoR * uR * uR + pR, FR2 = rhoR * uR * vR, FR3 = (R[3] + pR) * uR;

    double sqL = sqrt(rhoL), sqR = sqrt(rhoR);
    double den = sqL + sqR;
    if (den < eps) den = eps;
    double u_roe = (sqL * uL + sqR * uR) / den;
    double h_roe = (sqL * (EL + pL / rhoL) + sqR * (ER + pR / rhoR)) / den;
    double c_roe = sqrt(stdmax((gamma - 1.0) * (h_roe - 0.5 * u_roe * u_roe), eps));

    double SL = stdmin(uL - aL, u_roe - c_roe);
    double SR = stdmax(uR + aR, u_roe + c_roe);

    if (SL >= 0.0) { F[0] = FL0; F[1] = FL1; F[2] = FL2; F[3] = FL3; return; }
    if (SR <= 0.0) { F[0] = FR0; F[1] = FR1; F[2] = FR2; F[3] = FR3; return; }

    double num = pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR);
    double den2 = rhoL * (SL - uL) - rhoR * (SR - uR);
    if (fabs(den2) < eps) den2 = (den2 >= 0 ? eps : -eps);
    double Sstar = num / den2;

    double rho, u, v, E, p, S;
    const double *Us;
    double FF0, FF1, FF2, FF3;
    if (Sstar >= 0.0) {
        rho = rhoL; u = uL; v = vL; E = EL; p = pL; S = SL; Us = L;
        FF0 = FL0; FF1 = FL1; FF2 = FL2; FF3 = FL3;
    } else {
        rho = rhoR; u = uR; v = vR; E = ER; p = pR; S = SR; Us = R;
        FF0 = FR0; FF1 = FR1; FF2 = FR2; FF3 = FR3;
    }
    double denom = S - Sstar;
    if (fabs(denom) < eps) denom = (denom >= 0 ? eps : -eps);
    double c = rho * (S - u) / denom;
    double term = (Sstar - u) * (Sstar + p / (rho * (S - u)));
    if (isnan(term) || isinf(term)) term = 0.0;
    double s0 = c, s1 = c * Sstar, s2 = c * v, s3 = c * (E + term);
    F[0] = FF0 + S * (s0 - Us[0]);
    F[1] = FF1 + S * (s1 - Us[1]);
    F[2] = FF2 + S * (s2 - Us[2]);
    F[3] = FF3 + S * (s3 - Us[3]);
}

/* HLL flux with Roe-averaged (u, v) speed bounds, eps 1e-6.
 * Follows hlleFlux, src_DG/fluid.cpp:525-580 (defined but never called in the
 * reference; offered here as the selectable HLL flux). */
void zko_hll(const double L[4], const double R[4], double gamma, double F[4])
{
    const double eps = 1e-6;
    double rhoL = stdmax(L[0], eps), uL = L[1] / rhoL, vL = L[2] / rhoL, EL = L[3];
    double pL = stdmax((gamma - 1.0) * (EL - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    double aL = sqrt(stdmax(gamma * pL / rhoL, eps));
    double HL = (EL + pL) / rhoL;

    double rhoR = stdmax(R[0], eps), uR = R[1] / rhoR, vR = R[2] / rhoR, ER = R[3];
    double pR = stdmax((gamma - 1.0) * (ER - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    double aR = sqrt(stdmax(gamma * pR / rhoR, eps));
    double HR = (ER + pR) / rhoR;

    double sqL = sqrt(rhoL), sqR = sqrt(rhoR);
    double invDen = 1.0 / stdmax(sqL + sqR, eps);
    double u_roe = (sqL * uL + sqR * uR) * invDen;
    double v_roe = (sqL * vL + sqR * vR) * invDen;
    double h_roe = (sqL * HL + sqR * HR) * invDen;
    double c_roe = sqrt(stdmax((gamma - 1.0) * (h_roe - 0.5 * (u_roe * u_roe + v_roe * v_roe)), eps));

    double SL = stdmin(uL - aL, u_roe - c_roe);
    double SR = stdmax(uR + aR, u_roe + c_roe);

    double FL0 = rhoL * uL, FL1 = rhoL * uL * uL + pL, FL2 = rhoL * uL * vL, FL3 = (EL + pL) * uL;
    double FR0 = rhoR * uR, FR1 = rhoR * uR * uR + pR, FR2 = rhoR * uR * vR, FR3 = (ER + pR) * uR;

    if (SL >= 0.0) {
        F[0] = FL0; F[1] = FL1; F[2] = FL2; F[3] = FL3;
    } else if (SR <= 0.0) {
        F[0] = FR0; F[1] = FR1; F[2] = FR2; F[3] = FR3;
    } else {
        double invDiff = 1.0 / stdmax(SR - SL, eps);
        F[0] = (SR * FL0 - SL * FR0 + SL * SR * (R[0] - L[0])) * invDiff;
        F[1] = (SR * FL1 - SL * FR1 + SL * SR * (R[1] - L[1])) * invDiff;
        F[2] = (SR * FL2 - SL * FR2 + SL * SR * (R[2] - L[2])) * invDiff;
        F[3] = (SR * FL3 - SL * FR3 + SL * SR * (R[3] - L[3])) * invDiff;
    }
}

typedef void (*flux_fn)(const double *, const double *, double, double *);

/* Conservative variables from primitives, src_WENO/fluid.cpp:86-93. */
void zko_prim_to_cons(int n, double gamma, const double *rho, const double *u,
                      const double *v, const double *p, double *U)
{
    for (int k = 0; k < n; ++k) {
        double r = rho[k], ux = u[k], vy = v[k], pr = p[k];
        U[k] = r;
        U[(size_t)n + k] = r * ux;
        U[2 * (size_t)n + k] = r * vy;
        U[3 * (size_t)n + k] = pr / (gamma - 1.0) + 0.5 * r * (ux * ux + vy * vy);
    }
}

/* ======================================================================
 * Ghost fill
 * ====================================================================== */

/* Periodic image of a ghost index inside the interior [3, n-3). */
static inline int per_idx(int i, int n)
{
    if (i < 3) return i + (n - 6);
    if (i >= n - 3) return i - (n - 6);
    return i;
}

/* Mirror slip wall of the finite-volume option, one pass (0: along columns,
 * wall normal y, negate rho*v; 1: along rows, negate rho*u).  A bctype-1 cell
 * whose nearest updated cell along the pass lies k <= 3 cells away (ties go
 * to the larger index) takes the cell mirrored across the wall between them,
 * k-1 cells past that updated cell (2k-1 from the ghost), or the nearest
 * updated cell itself where the mirrored one is not updated. */
static void fill_slip_mirror(int ny, int nx, const int32_t *bct, double *U, int pass)
{
    const size_t NP = (size_t)ny * nx;
    const int n = pass == 0 ? ny : nx;
#define ZKO_UPD(a) (pass == 0 ? ((a) >= 3 && (a) < ny - 3 && j >= 3 && j < nx - 3 && bct[IDX((a), j)] == 0) \
                              : ((a) >= 3 && (a) < nx - 3 && i >= 3 && i < ny - 3 && bct[IDX(i, (a))] == 0))
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            if (bct[IDX(i, j)] != 1) continue;
            const int g = pass == 0 ? i : j;
            int k = 0, s = 0;
            for (int d = 1; d <= 3 && !k; ++d) {
                if (g + d < n && ZKO_UPD(g + d)) { k = d; s = 1; }
                else if (g - d >= 0 && ZKO_UPD(g - d)) { k = d; s = -1; }
            }
            if (!k) continue;
            int m = g + s * (2 * k - 1);
            if (!(m >= 0 && m < n && ZKO_UPD(m))) m = g + s * k;
            const size_t src = pass == 0 ? IDX(m, j) : IDX(i, m), dst = IDX(i, j);
            for (int c = 0; c < 4; ++c) {
                const double x = U[(size_t)c * NP + src];
                U[(size_t)c * NP + dst] = (c == (pass == 0 ? 2 : 1)) ? -x : x;
            }
        }
#undef ZKO_UPD
}

/* Copy ghosts of nplanes planes: the row pass then the column pass of
 * src_WENO/fluid.cpp:338-370 / src_DG/fluid.cpp:581-635 (last writer wins),
 * with slip (bctype 1) = 1 (DG) negating the normal momentum planes
 * [plane_c1, plane_c1+nm) or [plane_c2, ...), or slip = 2 (finite volume,
 * nm = 1) filling mirror slip walls after each pass; then the periodic (-2)
 * pass. */
static void fill_ghosts(int ny, int nx, const int32_t *bct, double *Q, int nm, int slip)
{
    const size_t NP = (size_t)ny * nx;
    const int OFF[6] = {-3, -2, -1, 1, 2, 3};
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 3; i < ny - 3; ++i)
            for (int j = 3; j < nx - 3; ++j) {
                if (bct[IDX(i, j)] != 0) continue;
                for (int d = 0; d < 6; ++d) {
                    int ni = i, nj = j;
                    if (pass == 0) { ni = i + OFF[d]; if ((unsigned)ni >= (unsigned)ny) continue; }
                    else { nj = j + OFF[d]; if ((unsigned)nj >= (unsigned)nx) continue; }
                    int bt = bct[IDX(ni, nj)];
                    size_t src = IDX(i, j), dst = IDX(ni, nj);
                    if (bt == -1) {
                        for (int p = 0; p < 4 * nm; ++p) Q[p * NP + dst] = Q[p * NP + src];
                    } else if (bt == 1 && slip == 1) {
                        int neg_c = (pass == 0) ? 2 : 1; /* y pass: rho*v, x pass: rho*u */
                        for (int c = 0; c < 4; ++c)
                            for (int m = 0; m < nm; ++m) {
                                size_t p = (size_t)(c * nm + m) * NP;
                                Q[p + dst] = (c == neg_c) ? -Q[p + src] : Q[p + src];
                            }
                    }
                }
            }
        if (slip == 2) fill_slip_mirror(ny, nx, bct, Q, pass);
    }
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            if (bct[IDX(i, j)] != -2) continue;
            size_t src = IDX(per_idx(i, ny), per_idx(j, nx)), dst = IDX(i, j);
            for (int p = 0; p < 4 * nm; ++p) Q[p * NP + dst] = Q[p * NP + src];
        }
}

void zko_fv_fill_ghosts(int ny, int nx, const int32_t *bct, double *U)
{
    fill_ghosts(ny, nx, bct, U, 1, 0);
}

void zko_fv_fill_ghosts_slip(int ny, int nx, const int32_t *bct, double *U)
{
    fill_ghosts(ny, nx, bct, U, 1, 2);
}

void zko_dg_fill_ghosts(int ny, int nx, int P, const int32_t *bct, double *Q)
{
    fill_ghosts(ny, nx, bct, Q, (P + 1) * (P + 1), 1);
}

/* ======================================================================
 * Finite volume: WENO5 + HLLC/HLL, SSP-RK3
 * ====================================================================== */

/* dU = dt * L(U) on bctype-0 cells in [3,ny-3)x[3,nx-3), 0 elsewhere; the
 * ghosts of U are filled first.  src_WENO/fluid.cpp:328-463. */
/* Face-level positivity check (an option of this project, no reference
 * counterpart): when either reconstructed face state has rho or p below
 * 1e-12, the face falls back to the first-order states (the two cell values). */
static inline int face_positive(const double Q[4], double gamma)
{
    if (!(Q[0] >= 1e-12)) return 0;
    double p = (gamma - 1.0) * (Q[3] - 0.5 * (Q[1] * Q[1] + Q[2] * Q[2]) / Q[0]);
    return p >= 1e-12;
}

static inline void positivity_fix(double QL[4], double QR[4], const double cl[4], const double cr[4], double gamma)
{
    if (face_positive(QL, gamma) && face_positive(QR, gamma)) return;
    for (int c = 0; c < 4; ++c) {
        QL[c] = cl[c];
        QR[c] = cr[c];
    }
}

/* As zko_fv_rhs; when bad is non-null, every face with a bad cell on either
 * side uses the first-order states (the two cell values).  This is the second
 * pass of the cell positivity fallback of zko_fv_step. */
void zko_fv_rhs_masked(int ny, int nx, double da, double gamma, const int32_t *bct,
                       double *U, double dt, int flux, double *dU, const uint8_t *bad)
{
    const size_t NP = (size_t)ny * nx;
    const double dx = da, dy = da;
    const int pos = (flux & ZKO_POSITIVITY) != 0;
    const int recon = (flux >> ZKO_RECON_SHIFT) & 3;
    const int slip = (flux & ZKO_FV_SLIP) != 0;
    flux &= 15;
    flux_fn fl = flux == 1 ? zko_hll : zko_hllc_fv;
    fill_ghosts(ny, nx, bct, U, 1, slip ? 2 : 0);
    memset(dU, 0, 4 * NP * sizeof(double));
    const double *p0 = U, *p1 = U + NP, *p2 = U + 2 * NP, *p3 = U + 3 * NP;
    /* Slip walls: a wall cell that lies within 3 cells of updated cells along
     * both its row and its column holds the row mirror (the row pass comes
     * last), which is the right ghost for the x faces; the y faces read a copy
     * in which the column pass is applied again. */
    double *Uy = NULL;
    const double *q0 = p0, *q1 = p1, *q2 = p2, *q3 = p3;
    if (slip) {
        Uy = (double *)malloc(4 * NP * sizeof(double));
        memcpy(Uy, U, 4 * NP * sizeof(double));
        fill_slip_mirror(ny, nx, bct, Uy, 0);
        q0 = Uy; q1 = Uy + NP; q2 = Uy + 2 * NP; q3 = Uy + 3 * NP;
    }

#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 3; i < ny - 3; ++i) {
        for (int j = 3; j < nx - 3; ++j) {
            if (bct[IDX(i, j)] != 0) continue;
            /* 13-point cross: x offsets -3..3 along the row, y offsets along the column */
            double X[7][4], Y[7][4];
            for (int o = -3; o <= 3; ++o) {
                size_t kx = IDX(i, j + o), ky = IDX(i + o, j);
                X[o + 3][0] = p0[kx]; X[o + 3][1] = p1[kx]; X[o + 3][2] = p2[kx]; X[o + 3][3] = p3[kx];
                Y[o + 3][0] = q0[ky]; Y[o + 3][1] = q1[ky]; Y[o + 3][2] = q2[ky]; Y[o + 3][3] = q3[ky];
            }
            double QLr[4], QRr[4], QLl[4], QRl[4], QLd[4], QRd[4], QLu[4], QRu[4];
            for (int c = 0; c < 4; ++c) {
                double s[6];
                for (int t = 0; t < 6; ++t) s[t] = X[t + 1][c];
                zko_reconstruct(recon, s, &QLr[c], &QRr[c]);  /* face j+1/2 */
                for (int t = 0; t < 6; ++t) s[t] = X[t][c];
                zko_reconstruct(recon, s, &QLl[c], &QRl[c]);  /* face j-1/2 */
                for (int t = 0; t < 6; ++t) s[t] = Y[t + 1][c];
                zko_reconstruct(recon, s, &QLd[c], &QRd[c]);  /* face i+1/2 */
                for (int t = 0; t < 6; ++t) s[t] = Y[t][c];
                zko_reconstruct(recon, s, &QLu[c], &QRu[c]);  /* face i-1/2 */
            }
            if (pos) {
                positivity_fix(QLr, QRr, X[3], X[4], gamma);
                positivity_fix(QLl, QRl, X[2], X[3], gamma);
                positivity_fix(QLd, QRd, Y[3], Y[4], gamma);
                positivity_fix(QLu, QRu, Y[2], Y[3], gamma);
            }
            if (bad) {
                const int b0 = bad[IDX(i, j)];
                if (b0 || bad[IDX(i, j + 1)])
                    for (int c = 0; c < 4; ++c) { QLr[c] = X[3][c]; QRr[c] = X[4][c]; }
                if (b0 || bad[IDX(i, j - 1)])
                    for (int c = 0; c < 4; ++c) { QLl[c] = X[2][c]; QRl[c] = X[3][c]; }
                if (b0 || bad[IDX(i + 1, j)])
                    for (int c = 0; c < 4; ++c) { QLd[c] = Y[3][c]; QRd[c] = Y[4][c]; }
                if (b0 || bad[IDX(i - 1, j)])
                    for (int c = 0; c < 4; ++c) { QLu[c] = Y[2][c]; QRu[c] = Y[3][c]; }
            }
            double Fr[4], Fl[4], Fd[4], Fu[4];
            fl(QLr, QRr, gamma, Fr);
            fl(QLl, QRl, gamma, Fl);
            /* y faces: call with rho*u and rho*v swapped (fluid.cpp:444-451) */
            double a[4] = {QLd[0], QLd[2], QLd[1], QLd[3]}, b[4] = {QRd[0], QRd[2], QRd[1], QRd[3]};
            fl(a, b, gamma, Fd);
            double a2[4] = {QLu[0], QLu[2], QLu[1], QLu[3]}, b2[4] = {QRu[0], QRu[2], QRu[1], QRu[3]};
            fl(a2, b2, gamma, Fu);

            const double dtdx = dt / dx, dtdy = dt / dy;
            size_t k = IDX(i, j);
            dU[k] = -(dtdx * (Fr[0] - Fl[0]) + dtdy * (Fd[0] - Fu[0]));
            dU[NP + k] = -(dtdx * (Fr[1] - Fl[1]) + dtdy * (Fd[2] - Fu[2]));
            dU[2 * NP + k] = -(dtdx * (Fr[2] - Fl[2]) + dtdy * (Fd[1] - Fu[1]));
            dU[3 * NP + k] = -(dtdx * (Fr[3] - Fl[3]) + dtdy * (Fd[3] - Fu[3]));
        }
    }
    free(Uy);
}

void zko_fv_rhs(int ny, int nx, double da, double gamma, const int32_t *bct,
                double *U, double dt, int flux, double *dU)
{
    zko_fv_rhs_masked(ny, nx, da, gamma, bct, U, dt, flux, dU, NULL);
}

/* Shu-Osher combination of one stage: stage 0 T = Un + dU, else
 * T = ca Un + cb (U + dU), over every cell. */
static void fv_combine(size_t N4, int stage, const double *Un, const double *U, const double *dU, double *T)
{
    if (stage == 0)
        for (size_t k = 0; k < N4; ++k) T[k] = Un[k] + dU[k];
    else if (stage == 1)
        for (size_t k = 0; k < N4; ++k) T[k] = 0.75 * Un[k] + 0.25 * (U[k] + dU[k]);
    else
        for (size_t k = 0; k < N4; ++k) T[k] = (1.0 / 3.0) * Un[k] + (2.0 / 3.0) * (U[k] + dU[k]);
}

/* One SSP-RK3 step in Shu-Osher form over every cell (ghosts included, which
 * leaves the stale ghost mixtures of the reference).  src_WENO/fluid.cpp:468-493.
 *
 * With ZKO_POSITIVITY (an option of this project, no reference counterpart)
 * each stage also has a cell positivity fallback: an updated cell whose new
 * state has rho or p not above 1e-12 (or is not finite) is bad, and the stage
 * is recomputed with first-order fluxes at every face touching a bad cell
 * (once; the cells away from bad cells keep their values).  Returns the
 * number of bad cells met in the three stages. */
int zko_fv_step(int ny, int nx, double da, double gamma, const int32_t *bct,
                double *U, double dt, int flux)
{
    int nbad = 0;
    const size_t NP = (size_t)ny * nx, N4 = 4 * NP;
    const int pos = (flux & ZKO_POSITIVITY) != 0;
    double *Un = (double *)malloc(N4 * sizeof(double));
    double *dU = (double *)malloc(N4 * sizeof(double));
    double *T = (double *)malloc(N4 * sizeof(double));
    uint8_t *bad = pos ? (uint8_t *)calloc(NP, 1) : NULL;
    memcpy(Un, U, N4 * sizeof(double));
    for (int stage = 0; stage < 3; ++stage) {
        zko_fv_rhs(ny, nx, da, gamma, bct, U, dt, flux, dU);
        fv_combine(N4, stage, Un, U, dU, T);
        if (pos) {
            int any = 0;
            for (int i = 3; i < ny - 3; ++i)
                for (int j = 3; j < nx - 3; ++j) {
                    const size_t k = IDX(i, j);
                    int b = 0;
                    if (bct[k] == 0) {
                        /* p > 1e-12 multiplied through by 2 rho / (gamma - 1) > 0 */
                        const double rho = T[k];
                        const double q = 2.0 * rho * T[3 * NP + k] - (T[NP + k] * T[NP + k] + T[2 * NP + k] * T[2 * NP + k]);
                        b = !(rho > 1e-12 && q > 2.0 * rho * 1e-12 / (gamma - 1.0));
                    }
                    bad[k] = (uint8_t)b;
                    any |= b;
                    nbad += b;
                }
            if (any) {
                zko_fv_rhs_masked(ny, nx, da, gamma, bct, U, dt, flux, dU, bad);
                fv_combine(N4, stage, Un, U, dU, T);
            }
        }
        memcpy(U, T, N4 * sizeof(double));
    }
    free(Un);
    free(dU);
    free(T);
    free(bad);
    return nbad;
}

/* max(|u| + c) over every cell, ghosts included, no floors.
 * src_WENO/fluid.cpp:539-553. */
double zko_fv_max_speed(int ny, int nx, double gamma, const double *U)
{
    const size_t NP = (size_t)ny * nx;
    double umax = 0.0;
    for (size_t k = 0; k < NP; ++k) {
        double rho = U[k];
        double u = U[NP + k] / rho, v = U[2 * NP + k] / rho;
        double p = (gamma - 1.0) * (U[3 * NP + k] - 0.5 * rho * (u * u + v * v));
        double a = sqrt(gamma * p / rho);
        double spd = sqrt(u * u + v * v) + a;
        if (spd > umax) umax = spd;
    }
    return umax;
}

/* ======================================================================
 * Modal DG
 * ====================================================================== */

/* Legendre polynomials and derivatives, src_DG/fluid.cpp:34-55. */
double zko_legP(int n, double x)
{
    switch (n) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return 0.5 * (3.0 * x * x - 1.0);
    case 3: return 0.5 * (5.0 * x * x * x - 3.0 * x);
    case 4: return 0.125 * (35.0 * x * x * x * x - 30.0 * x * x + 3.0);
    default: return 0.0;
    }
}

double zko_legD(int n, double x)
{
    switch (n) {
    case 0: return 0.0;
    case 1: return 1.0;
    case 2: return 3.0 * x;
    case 3: return 0.5 * (15.0 * x * x - 3.0);
    case 4: return 0.5 * (35.0 * x * x * x - 15.0 * x);
    default: return 0.0;
    }
}

/* Gauss-Legendre nodes and weights as the decimal literals of
 * src_DG/fluid.cpp:94-111 (NQ = 1..5). */
static const double GLP[6][5] = {
    {0, 0, 0, 0, 0},
    {0.0, 0, 0, 0, 0},
    {-0.5773502691896257, 0.5773502691896257, 0, 0, 0},
    {-0.7745966692414834, 0.0, 0.7745966692414834, 0, 0},
    {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526, 0},
    {-0.9061798459386640, -0.5384693101056831, 0.0, 0.5384693101056831, 0.9061798459386640}};
static const double GLW[6][5] = {
    {0, 0, 0, 0, 0},
    {2.0, 0, 0, 0, 0},
    {1.0, 1.0, 0, 0, 0},
    {0.5555555555555556, 0.8888888888888888, 0.5555555555555556, 0, 0},
    {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538, 0},
    {0.2369268850561891, 0.4786286704993665, 0.5688888888888889, 0.4786286704993665, 0.2369268850561891}};

const double *zko_gl_points(int nq) { return GLP[nq]; }
const double *zko_gl_weights(int nq) { return GLW[nq]; }

static inline double phi2(int P, int m, double xi, double eta)
{
    return zko_legP(m / (P + 1), xi) * zko_legP(m % (P + 1), eta);
}

/* Euler fluxes with rho and p floored at 1e-3, src_DG/fluid.cpp:123-143. */
static inline void euler_fx(const double U[4], double F[4], double gam)
{
    const double eps = 1e-3;
    double rho = stdmax(U[0], eps);
    double u = U[1] / rho, v = U[2] / rho;
    double p = stdmax((gam - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v)), eps);
    F[0] = rho * u;
    F[1] = rho * u * u + p;
    F[2] = rho * u * v;
    F[3] = (U[3] + p) * u;
}

static inline void euler_fy(const double U[4], double G[4], double gam)
{
    const double eps = 1e-3;
    double rho = stdmax(U[0], eps);
    double u = U[1] / rho, v = U[2] / rho;
    double p = stdmax((gam - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v)), eps);
    G[0] = rho * v;
    G[1] = rho * u * v;
    G[2] = rho * v * v + p;
    G[3] = (U[3] + p) * v;
}

/* U(xi, eta) = sum_m dof_m phi_m, summed in mode order from 0.0
 * (evalU, src_DG/fluid.cpp:672-681). */
static inline void evalU(const double *Q, size_t NP, int nm, int P, size_t k,
                         double xi, double eta, double U[4])
{
    U[0] = U[1] = U[2] = U[3] = 0.0;
    for (int m = 0; m < nm; ++m) {
        double b = phi2(P, m, xi, eta);
        U[0] += Q[(size_t)(0 * nm + m) * NP + k] * b;
        U[1] += Q[(size_t)(1 * nm + m) * NP + k] * b;
        U[2] += Q[(size_t)(2 * nm + m) * NP + k] * b;
        U[3] += Q[(size_t)(3 * nm + m) * NP + k] * b;
    }
}

/* dQ = dQ/dt (not multiplied by dt) on bctype-0 cells of [3,ny-3)x[3,nx-3);
 * ghosts are filled first.  Volume term with NQ = P+1 Gauss points per
 * direction, four faces each with their own HLLC evaluation (every face is
 * computed by both of its cells, as in the reference), mass-matrix scale
 * invMass(m) * 0.5/h.  src_DG/fluid.cpp:656-807. */
void zko_dg_rhs(int ny, int nx, int P, double da, double gamma, const int32_t *bct,
                double *Q, int flux, double *dQ)
{
    const int nm = (P + 1) * (P + 1), NQ = P + 1;
    const size_t NP = (size_t)ny * nx;
    const double *GP = GLP[NQ], *GW = GLW[NQ];
    const double h = da, gam = gamma;
    flux_fn fl = flux == 1 ? zko_hll : zko_hllc_dg;

    memset(dQ, 0, 4 * (size_t)nm * NP * sizeof(double));
    zko_dg_fill_ghosts(ny, nx, P, bct, Q);
    const double inv2h = 0.5 / h;

#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 3; i < ny - 3; ++i) {
        for (int j = 3; j < nx - 3; ++j) {
            if (bct[IDX(i, j)] != 0) continue;
            const size_t k = IDX(i, j);
            double rhs[4][16];
            memset(rhs, 0, sizeof(rhs));

            for (int qa = 0; qa < NQ; ++qa) {
                double xi = GP[qa], wa = GW[qa];
                for (int qb = 0; qb < NQ; ++qb) {
                    double eta = GP[qb], wb = GW[qb], ww = wa * wb;
                    double U[4], F[4], G[4];
                    evalU(Q, NP, nm, P, k, xi, eta, U);
                    euler_fx(U, F, gam);
                    euler_fy(U, G, gam);
                    for (int m = 0; m < nm; ++m) {
                        int px = m / (P + 1), py = m % (P + 1);
                        double dxi = zko_legD(px, xi) * zko_legP(py, eta);
                        double det = zko_legP(px, xi) * zko_legD(py, eta);
                        double coeff = ww;
                        for (int c = 0; c < 4; ++c) rhs[c][m] += coeff * (F[c] * dxi + G[c] * det);
                    }
                }
            }
            /* right face (xi = +1), neighbour (i, j+1) */
            for (int q = 0; q < NQ; ++q) {
                double eta = GP[q], w = GW[q], UL[4], UR[4], Fs[4];
                evalU(Q, NP, nm, P, k, +1.0, eta, UL);
                evalU(Q, NP, nm, P, IDX(i, j + 1), -1.0, eta, UR);
                fl(UL, UR, gam, Fs);
                for (int m = 0; m < nm; ++m) {
                    double bv = phi2(P, m, +1.0, eta) * w;
                    for (int c = 0; c < 4; ++c) rhs[c][m] -= bv * Fs[c];
                }
            }
            /* left face (xi = -1), neighbour (i, j-1) */
            for (int q = 0; q < NQ; ++q) {
                double eta = GP[q], w = GW[q], UL[4], UR[4], Fs[4];
                evalU(Q, NP, nm, P, IDX(i, j - 1), +1.0, eta, UL);
                evalU(Q, NP, nm, P, k, -1.0, eta, UR);
                fl(UL, UR, gam, Fs);
                for (int m = 0; m < nm; ++m) {
                    double bv = phi2(P, m, -1.0, eta) * w;
                    for (int c = 0; c < 4; ++c) rhs[c][m] += bv * Fs[c];
                }
            }
            /* down face (eta = +1), neighbour (i+1, j); flux called with rho*u <-> rho*v */
            for (int q = 0; q < NQ; ++q) {
                double xi = GP[q], w = GW[q], UL[4], UR[4], Gs[4];
                evalU(Q, NP, nm, P, k, xi, +1.0, UL);
                evalU(Q, NP, nm, P, IDX(i + 1, j), xi, -1.0, UR);
                double a[4] = {UL[0], UL[2], UL[1], UL[3]}, b[4] = {UR[0], UR[2], UR[1], UR[3]};
                fl(a, b, gam, Gs);
                double Gst[4] = {Gs[0], Gs[2], Gs[1], Gs[3]};
                for (int m = 0; m < nm; ++m) {
                    double bv = phi2(P, m, xi, +1.0) * w;
                    for (int c = 0; c < 4; ++c) rhs[c][m] -= bv * Gst[c];
                }
            }
            /* up face (eta = -1), neighbour (i-1, j) */
            for (int q = 0; q < NQ; ++q) {
                double xi = GP[q], w = GW[q], UL[4], UR[4], Gs[4];
                evalU(Q, NP, nm, P, IDX(i - 1, j), xi, +1.0, UL);
                evalU(Q, NP, nm, P, k, xi, -1.0, UR);
                double a[4] = {UL[0], UL[2], UL[1], UL[3]}, b[4] = {UR[0], UR[2], UR[1], UR[3]};
                fl(a, b, gam, Gs);
                double Gst[4] = {Gs[0], Gs[2], Gs[1], Gs[3]};
                for (int m = 0; m < nm; ++m) {
                    double bv = phi2(P, m, xi, -1.0) * w;
                    for (int c = 0; c < 4; ++c) rhs[c][m] += bv * Gst[c];
                }
            }
            for (int m = 0; m < nm; ++m) {
                int px = m / (P + 1), py = m % (P + 1);
                double scale = (double)((2 * px + 1) * (2 * py + 1)) * inv2h;
                for (int c = 0; c < 4; ++c) dQ[(size_t)(c * nm + m) * NP + k] = rhs[c][m] * scale;
            }
        }
    }
}

/* Legendre P0..P2 only; degree 3 evaluates to 0 (basis(), src_DG/fluid.cpp:808-823). */
static inline double zs_basis(int P, int m, double xi, double eta)
{
    int px = m / (P + 1), py = m % (P + 1);
#define ZSP(p, x) ((p) == 0 ? 1.0 : (p) == 1 ? (x) : (p) == 2 ? 0.5 * (3.0 * (x) * (x) - 1.0) : 0.0)
    return ZSP(px, xi) * ZSP(py, eta);
#undef ZSP
}

static const double ZS_PTS[4] = {-0.8611363115940526, -0.3399810435848563,
                                 0.3399810435848563, 0.8611363115940526};

/* Points of the positivity option (no reference counterpart): the face
 * quadrature points the fluxes read, (-1, eta_q), (+1, eta_q), (xi_q, -1),
 * (xi_q, +1), and the four corners (+-1, +-1) the CFL rule reads, with the
 * full Legendre basis.  Returns their count, 4 (P+1) + 4. */
static int ext_points(int P, double xs[], double ys[])
{
    const int NQ = P + 1;
    const double *gp = GLP[NQ];
    int n = 0;
    for (int q = 0; q < NQ; ++q) { xs[n] = -1.0; ys[n++] = gp[q]; }
    for (int q = 0; q < NQ; ++q) { xs[n] = +1.0; ys[n++] = gp[q]; }
    for (int q = 0; q < NQ; ++q) { xs[n] = gp[q]; ys[n++] = -1.0; }
    for (int q = 0; q < NQ; ++q) { xs[n] = gp[q]; ys[n++] = +1.0; }
    for (int c = 0; c < 4; ++c) { xs[n] = (c & 1) ? 1.0 : -1.0; ys[n++] = (c & 2) ? 1.0 : -1.0; }
    return n;
}

/* Zhang-Shu style positivity scaling on every cell of [3,ny-3)x[3,nx-3)
 * (no bctype test).  src_DG/fluid.cpp:892-991.  ext = 1 (the positivity
 * option) also takes both minima over the points of ext_points, with the
 * full basis, after the reference's 16 points. */
static void limit_zs(int ny, int nx, int P, double gamma, double *Q, int ext)
{
    double xs[20], ys[20];
    const int nx_ext = ext ? ext_points(P, xs, ys) : 0;
    const int nm = (P + 1) * (P + 1);
    const size_t NP = (size_t)ny * nx;
    const double RHO_MIN = 1e-5, P_MIN = 1e-5, EPS = 1e-5;
#define D(c, m) Q[(size_t)((c) * nm + (m)) * NP + k]
#pragma omp parallel for schedule(static)
    for (int i = 3; i < ny - 3; i++)
        for (int j = 3; j < nx - 3; j++) {
            const size_t k = IDX(i, j);
            D(0, 0) = stdmax(D(0, 0), RHO_MIN);
            double rho0 = D(0, 0), mx0 = D(1, 0), my0 = D(2, 0), E0 = D(3, 0);
            double p0 = (gamma - 1.0) * (E0 - 0.5 * (mx0 * mx0 + my0 * my0) / rho0);
            if (p0 < P_MIN) {
                E0 = P_MIN / (gamma - 1.0) + 0.5 * (mx0 * mx0 + my0 * my0) / rho0;
                D(3, 0) = E0;
            }
            double theta_rho = 1.0;
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) {
                    double rho = 0.0;
                    for (int m = 0; m < nm; m++) rho += D(0, m) * zs_basis(P, m, ZS_PTS[a], ZS_PTS[b]);
                    if (rho < RHO_MIN) {
                        double t = (rho0 - RHO_MIN) / (rho0 - rho + EPS);
                        theta_rho = stdmin(theta_rho, t);
                    }
                }
            for (int e = 0; e < nx_ext; e++) {
                double rho = 0.0;
                for (int m = 0; m < nm; m++) rho += D(0, m) * phi2(P, m, xs[e], ys[e]);
                if (rho < RHO_MIN) theta_rho = stdmin(theta_rho, (rho0 - RHO_MIN) / (rho0 - rho + EPS));
            }
            if (theta_rho < 1.0)
                for (int m = 1; m < nm; m++)
                    for (int c = 0; c < 4; ++c) D(c, m) *= theta_rho;

            double theta_p = 1.0;
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) {
                    double rho = 0.0, mx = 0.0, my = 0.0, E = 0.0;
                    for (int m = 0; m < nm; m++) {
                        double phi = zs_basis(P, m, ZS_PTS[a], ZS_PTS[b]);
                        rho += D(0, m) * phi;
                        mx += D(1, m) * phi;
                        my += D(2, m) * phi;
                        E += D(3, m) * phi;
                    }
                    if (rho <= RHO_MIN) continue;
                    double u = mx / rho, v = my / rho;
                    double p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
                    if (p < P_MIN) {
                        double p_avg = (gamma - 1.0) * (D(3, 0) - 0.5 * (D(1, 0) * D(1, 0) + D(2, 0) * D(2, 0)) / D(0, 0));
                        double t = (p_avg - P_MIN) / (p_avg - p + EPS);
                        theta_p = stdmin(theta_p, t);
                    }
                }
            for (int e = 0; e < nx_ext; e++) {
                double rho = 0.0, mx = 0.0, my = 0.0, E = 0.0;
                for (int m = 0; m < nm; m++) {
                    double phi = phi2(P, m, xs[e], ys[e]);
                    rho += D(0, m) * phi;
                    mx += D(1, m) * phi;
                    my += D(2, m) * phi;
                    E += D(3, m) * phi;
                }
                if (rho <= RHO_MIN) continue;
                double u = mx / rho, v = my / rho;
                double p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
                if (p < P_MIN) {
                    double p_avg = (gamma - 1.0) * (D(3, 0) - 0.5 * (D(1, 0) * D(1, 0) + D(2, 0) * D(2, 0)) / D(0, 0));
                    theta_p = stdmin(theta_p, (p_avg - P_MIN) / (p_avg - p + EPS));
                }
            }
            if (theta_p < 1.0)
                for (int m = 1; m < nm; m++)
                    for (int c = 0; c < 4; ++c) D(c, m) *= theta_p;
        }
#undef D
}

void zko_dg_limit_zs(int ny, int nx, int P, double gamma, double *Q)
{
    limit_zs(ny, nx, P, gamma, Q, 0);
}

void zko_dg_limit_zs_ext(int ny, int nx, int P, double gamma, double *Q)
{
    limit_zs(ny, nx, P, gamma, Q, 1);
}

/* Persson modal-decay filter on bctype-0 cells of [3,ny-3)x[3,nx-3), then the
 * positivity scaling.  src_DG/fluid.cpp:824-887 and 892-991. */
static void limit_persson_zs(int ny, int nx, int P, double gamma, const int32_t *bct, double *Q, int ext)
{
    const int nm = (P + 1) * (P + 1);
    const size_t NP = (size_t)ny * nx;
    const double kappa = 1.0, EPS = 1e-5;
    double s0;
    if (P >= 3) s0 = -3.0 * log10((double)P);
    else s0 = -4.0 * log10((double)P);
#pragma omp parallel for schedule(static)
    for (int i = 3; i < ny - 3; i++)
        for (int j = 3; j < nx - 3; j++) {
            const size_t k = IDX(i, j);
            if (bct[k] != 0) continue;
            for (int c = 0; c < 4; c++) {
                double Etot = 0.0, Ep = 0.0;
                for (int px = 0; px <= P; px++)
                    for (int py = 0; py <= P; py++) {
                        double a = Q[(size_t)(c * nm + px * (P + 1) + py) * NP + k];
                        Etot += a * a;
                        if (px == P || py == P) Ep += a * a;
                    }
                double s = log10(Ep / (Etot + EPS));
                if (s > s0) {
                    double eps;
                    if (s > s0 + kappa) eps = 1.0;
                    else {
                        double r = (s - s0) / kappa;
                        eps = 0.5 * (1.0 + sin(M_PI * (r - 0.5)));
                    }
                    for (int px = 0; px <= P; px++)
                        for (int py = 0; py <= P; py++)
                            if (px == P || py == P) Q[(size_t)(c * nm + px * (P + 1) + py) * NP + k] *= (1.0 - eps);
                }
            }
        }
    limit_zs(ny, nx, P, gamma, Q, ext);
}

void zko_dg_limit_persson_zs(int ny, int nx, int P, double gamma, const int32_t *bct,
                             double *Q)
{
    limit_persson_zs(ny, nx, P, gamma, bct, Q, 0);
}

static inline double minmod3(double a, double b, double c)
{
    if (a > 0.0 && b > 0.0 && c > 0.0) return dmin(a, dmin(b, c));
    if (a < 0.0 && b < 0.0 && c < 0.0) return dmax(a, dmax(b, c));
    return 0.0;
}

/* TVB-modified minmod (Cockburn & Shu 2001, J. Sci. Comput. 16(3), sec. 2.4):
 * keep a when |a| <= M h^2, else minmod(a, b, c). */
static inline double tvb_minmod(double a, double b, double c, double Mh2)
{
    if (fabs(a) <= Mh2) return a;
    return minmod3(a, b, c);
}

/* Component-wise TVB limiter on the conserved variables.  No reference
 * implementation exists (the reference documents one at src_DG/fluid.h:140-147
 * and README_old.md:137-208 but codes the Persson filter instead).  Per
 * bctype-0 cell of [3,ny-3)x[3,nx-3) and variable c: the x-linear mode
 * (px=1,py=0) and y-linear mode (px=0,py=1) are limited against forward and
 * backward differences of the neighbour means (neighbour means taken after
 * the ghost fill, y forward = +i); if either changes, every other mode >= 1
 * of that variable is zeroed and the cell is flagged.  At P >= 3 the cell is
 * also troubled when the one-dimensional TVB test (Cockburn & Shu 1989,
 * Math. Comp. 52, sec. 2) changes one of its four face-average deviations:
 * east sum_q q(q,0), west sum_q (-1)^(q+1) q(q,0), and the same over q(0,q)
 * for the +i and -i faces, each against the same forward and backward mean
 * differences as the linear mode of its axis.  The linear-mode test alone
 * leaves a cell with small linear but large quadratic and cubic modes
 * unlimited; at P = 3 that lets the four-quadrant problem blow up near
 * t = 0.0054.  faces = 0 gives the linear-mode test alone at every P (the rule
 * README_old.md:137-208 describes; zko_dg_limit_tvb_linear). */
static void limit_tvb(int ny, int nx, int P, double da, double M, const int32_t *bct,
                      double *Q, uint8_t *flags, int faces)
{
    const int nm = (P + 1) * (P + 1);
    const size_t NP = (size_t)ny * nx;
    const double Mh2 = M * da * da;
    const int mxm = 1 * (P + 1) + 0, mym = 0 * (P + 1) + 1;
    /* mode-0 planes after the ghost fill */
    double *A = (double *)malloc(4 * NP * sizeof(double));
    for (int c = 0; c < 4; ++c) memcpy(A + c * NP, Q + (size_t)(c * nm) * NP, NP * sizeof(double));
    fill_ghosts(ny, nx, bct, A, 1, 1);
    if (flags) memset(flags, 0, NP);
#pragma omp parallel for schedule(static)
    for (int i = 3; i < ny - 3; i++)
        for (int j = 3; j < nx - 3; j++) {
            const size_t k = IDX(i, j);
            if (bct[k] != 0) continue;
            int flag = 0;
            for (int c = 0; c < 4; ++c) {
                const double *a0 = A + c * NP;
                double ub = a0[k];
                double ax = Q[(size_t)(c * nm + mxm) * NP + k];
                double ay = Q[(size_t)(c * nm + mym) * NP + k];
                const double dxp = a0[IDX(i, j + 1)] - ub, dxm = ub - a0[IDX(i, j - 1)];
                const double dyp = a0[IDX(i + 1, j)] - ub, dym = ub - a0[IDX(i - 1, j)];
                double axl = tvb_minmod(ax, dxp, dxm, Mh2);
                double ayl = tvb_minmod(ay, dyp, dym, Mh2);
                int trouble = axl != ax || ayl != ay;
                if (P >= 3 && faces) {
                    double e = 0.0, w = 0.0, n = 0.0, so = 0.0;
                    for (int q = 1; q <= P; ++q) {
                        const double qx = Q[(size_t)(c * nm + q * (P + 1)) * NP + k];
                        const double qy = Q[(size_t)(c * nm + q) * NP + k];
                        e += qx;
                        n += qy;
                        if (q & 1) { w += qx; so += qy; }
                        else { w -= qx; so -= qy; }
                    }
                    trouble = trouble || tvb_minmod(e, dxp, dxm, Mh2) != e || tvb_minmod(w, dxp, dxm, Mh2) != w ||
                              tvb_minmod(n, dyp, dym, Mh2) != n || tvb_minmod(so, dyp, dym, Mh2) != so;
                }
                if (trouble) {
                    for (int m = 1; m < nm; ++m) Q[(size_t)(c * nm + m) * NP + k] = 0.0;
                    Q[(size_t)(c * nm + mxm) * NP + k] = axl;
                    Q[(size_t)(c * nm + mym) * NP + k] = ayl;
                    flag = 1;
                }
            }
            if (flags) flags[k] = (uint8_t)flag;
        }
    free(A);
}

void zko_dg_limit_tvb(int ny, int nx, int P, double da, double M, const int32_t *bct,
                      double *Q, uint8_t *flags)
{
    limit_tvb(ny, nx, P, da, M, bct, Q, flags, 1);
}

void zko_dg_limit_tvb_linear(int ny, int nx, int P, double da, double M, const int32_t *bct,
                             double *Q, uint8_t *flags)
{
    limit_tvb(ny, nx, P, da, M, bct, Q, flags, 0);
}

/* SSP-RK3 on all modes with the limiter after every stage.
 * src_DG/fluid.cpp:1005-1054. */
void zko_dg_step(int ny, int nx, int P, double da, double gamma, const int32_t *bct,
                 double *Q, double dt, int flux, int limiter, double tvb_M, uint8_t *flags)
{
    const int nm = (P + 1) * (P + 1);
    const size_t N = 4 * (size_t)nm * ny * nx;
    double *Qn = (double *)malloc(N * sizeof(double));
    double *dQ = (double *)malloc(N * sizeof(double));
    memcpy(Qn, Q, N * sizeof(double));
    for (int stage = 0; stage < 3; ++stage) {
        zko_dg_rhs(ny, nx, P, da, gamma, bct, Q, flux, dQ);
        if (stage == 0)
            for (size_t k = 0; k < N; ++k) Q[k] = Qn[k] + dt * dQ[k];
        else if (stage == 1)
            for (size_t k = 0; k < N; ++k) Q[k] = 0.75 * Qn[k] + 0.25 * (Q[k] + dt * dQ[k]);
        else
            for (size_t k = 0; k < N; ++k) Q[k] = (1.0 / 3.0) * Qn[k] + (2.0 / 3.0) * (Q[k] + dt * dQ[k]);
        const int lim = limiter & 7, ext = (limiter & 8) != 0;
        if (lim == 1) limit_persson_zs(ny, nx, P, gamma, bct, Q, ext);
        else if (lim == 2 || lim == 4) {
            limit_tvb(ny, nx, P, da, tvb_M, bct, Q, flags, lim == 2);
            limit_zs(ny, nx, P, gamma, Q, ext);
        }
    }
    free(Qn);
    free(dQ);
}

/* max(|u| + c) at the four corners (+-1, +-1) of every bctype-0 cell, with
 * floors 1e-12.  computeMaxSpeedDG_omp, src_DG/solver.cpp:26-64. */
double zko_dg_max_speed(int ny, int nx, int P, double gamma, const int32_t *bct,
                        const double *Q)
{
    const int nm = (P + 1) * (P + 1);
    const size_t NP = (size_t)ny * nx;
    const double eps = 1e-12;
    const double XI[4] = {-1.0, +1.0, -1.0, +1.0};
    const double ETA[4] = {-1.0, -1.0, +1.0, +1.0};
    double umax = 0.0;
    for (size_t k = 0; k < NP; ++k) {
        if (bct[k] != 0) continue;
        for (int c4 = 0; c4 < 4; ++c4) {
            double U[4] = {0, 0, 0, 0};
            for (int m = 0; m < nm; ++m) {
                double b = phi2(P, m, XI[c4], ETA[c4]);
                for (int c = 0; c < 4; ++c) U[c] += Q[(size_t)(c * nm + m) * NP + k] * b;
            }
            double rho = stdmax(U[0], eps);
            double u = U[1] / rho, v = U[2] / rho;
            double p = stdmax((gamma - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v)), eps);
            double a = sqrt(gamma * p / rho);
            umax = stdmax(umax, sqrt(u * u + v * v) + a);
        }
    }
    return umax;
}

/* nu = 0.9 / (2 (2P+1)), src_DG/solver.cpp:13-17. */
double zko_dg_cfl_limit(int P) { return 0.9 / (2.0 * (2 * P + 1)); }
