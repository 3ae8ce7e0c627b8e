// zk_device.h -- device-side point kernels shared by the stage kernels.
//
// Each function restates one reference routine (cited per function) in FP64.
// std::max/std::min semantics are kept exactly (a < b ? b : a), since fmax
// differs from them on NaN inputs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace zk {

__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

// Jiang-Shu WENO5 at the face between UP and UR1: UL from UL2..UR2, UR from
// UL1..UR3 (weno5_reconstruct, src_WENO/fluid.cpp:230-263).
__device__ __forceinline__ void weno5_f64(double UL2, double UL1, double UP, double UR1, double UR2,
                                          double UR3, double &UL, double &UR)
{
    constexpr double eps = 1e-6, C1 = 13.0 / 12.0, C2 = 0.25;
    const double qL0 = (1.0 / 3.0) * UL2 - (7.0 / 6.0) * UL1 + (11.0 / 6.0) * UP;
    const double qL1 = -(1.0 / 6.0) * UL1 + (5.0 / 6.0) * UP + (1.0 / 3.0) * UR1;
    const double qL2 = (1.0 / 3.0) * UP + (5.0 / 6.0) * UR1 - (1.0 / 6.0) * UR2;
    const double d0 = UL2 - 2 * UL1 + UP, d1 = UL2 - 4 * UL1 + 3 * UP;
    const double d2 = UL1 - 2 * UP + UR1, d3 = UL1 - UR1;
    const double d4 = UP - 2 * UR1 + UR2, d5 = 3 * UP - 4 * UR1 + UR2;
    const double bL0 = C1 * d0 * d0 + C2 * d1 * d1, bL1 = C1 * d2 * d2 + C2 * d3 * d3,
                 bL2 = C1 * d4 * d4 + C2 * d5 * d5;
    const double t0 = eps + bL0, t1 = eps + bL1, t2 = eps + bL2;
    const double aL0 = 0.1 / (t0 * t0), aL1 = 0.6 / (t1 * t1), aL2 = 0.3 / (t2 * t2);
    const double inv = 1.0 / (aL0 + aL1 + aL2);
    UL = (aL0 * qL0 + aL1 * qL1 + aL2 * qL2) * inv;

    const double qR0 = (11.0 / 6.0) * UR1 - (7.0 / 6.0) * UR2 + (1.0 / 3.0) * UR3;
    const double qR1 = (1.0 / 3.0) * UP + (5.0 / 6.0) * UR1 - (1.0 / 6.0) * UR2;
    const double qR2 = -(1.0 / 6.0) * UL1 + (5.0 / 6.0) * UP + (1.0 / 3.0) * UR1;
    const double r0 = UR1 - 2 * UR2 + UR3, r1 = 3 * UR1 - 4 * UR2 + UR3;
    const double r2 = UP - 2 * UR1 + UR2, r3 = UP - UR2;
    const double r4 = UL1 - 2 * UP + UR1, r5 = UL1 - 4 * UP + 3 * UR1;
    const double bR0 = C1 * r0 * r0 + C2 * r1 * r1, bR1 = C1 * r2 * r2 + C2 * r3 * r3,
                 bR2 = C1 * r4 * r4 + C2 * r5 * r5;
    const double s0 = eps + bR0, s1 = eps + bR1, s2 = eps + bR2;
    const double aR0 = 0.1 / (s0 * s0), aR1 = 0.6 / (s1 * s1), aR2 = 0.3 / (s2 * s2);
    const double inv2 = 1.0 / (aR0 + aR1 + aR2);
    UR = (aR0 * qR0 + aR1 * qR1 + aR2 * qR2) * inv2;
}

// HLLC of the finite-volume solver (hllcFlux, src_WENO/fluid.cpp:268-317).
__device__ __forceinline__ void hllc_fv_f64(const double L[4], const double R[4], double gamma, double F[4])
{
    constexpr double eps = 1e-12;
    const double rhoL = smax(L[0], eps), uL = L[1] / rhoL, vL = L[2] / rhoL, EL = L[3] / rhoL;
    const double pL = smax((gamma - 1.0) * (L[3] - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    const double aL = sqrt(gamma * pL / rhoL);
    const double rhoR = smax(R[0], eps), uR = R[1] / rhoR, vR = R[2] / rhoR, ER = R[3] / rhoR;
    const double pR = smax((gamma - 1.0) * (R[3] - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    const double aR = sqrt(gamma * pR / rhoR);

    const double sqL = sqrt(rhoL), sqR = sqrt(rhoR), den = sqL + sqR;
    const double u_roe = (sqL * uL + sqR * uR) / den;
    const double h_roe = (sqL * (EL + pL / rhoL) + sqR * (ER + pR / rhoR)) / den;
    const double c_roe = sqrt(smax((gamma - 1.0) * (h_roe - 0.5 * u_roe * u_roe), eps));
    const double SL = smin(uL - aL, u_roe - c_roe);
    const double SR = smax(uR + aR, u_roe + c_roe);

    if (SL >= 0.0) {
        F[0] = rhoL * uL; F[1] = rhoL * uL * uL + pL; F[2] = rhoL * uL * vL; F[3] = (L[3] + pL) * uL;
        return;
    }
    if (SR <= 0.0) {
        F[0] = rhoR * uR; F[1] = rhoR * uR * uR + pR; F[2] = rhoR * uR * vR; F[3] = (R[3] + pR) * uR;
        return;
    }
    const double num = pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR);
    double den2 = rhoL * (SL - uL) - rhoR * (SR - uR);
    if (fabs(den2) < 1e-12) den2 = (den2 >= 0 ? 1e-12 : -1e-12);
    const double Sstar = num / den2;

    const bool left = Sstar >= 0.0;
    const double rho = left ? rhoL : rhoR, u = left ? uL : uR, v = left ? vL : vR;
    const double E = left ? EL : ER, p = left ? pL : pR, S = left ? SL : SR;
    const double U0 = left ? L[0] : R[0], U1 = left ? L[1] : R[1], U2 = left ? L[2] : R[2],
                 U3 = left ? L[3] : R[3];
    const double FF0 = rho * u, FF1 = rho * u * u + p, FF2 = rho * u * v, FF3 = (U3 + p) * u;
    const double c = rho * (S - u) / (S - Sstar);
    const double s0 = c, s1 = c * Sstar, s2 = c * v, s3 = c * (E + (Sstar - u) * (Sstar + p / (rho * (S - u))));
    F[0] = FF0 + S * (s0 - U0);
    F[1] = FF1 + S * (s1 - U1);
    F[2] = FF2 + S * (s2 - U2);
    F[3] = FF3 + S * (s3 - U3);
}

// HLLC of the DG solver (hllcFlux, src_DG/fluid.cpp:450-520): eps 1e-5,
// sound speed sqrt(max(gamma p / rho, eps)), Roe speeds with u only, the
// denominator guards and the non-finite star-term reset.  The divisions by
// rho_L, rho_R and the Roe denominator are one reciprocal each (6 divisions
// instead of 15), which moves results by a few ulps; the special cases
// behave identically because the two divisions they involve are kept.
__device__ __forceinline__ void hllc_dg_fast_f64(const double L[4], const double R[4], double gamma, double F[4])
{
    constexpr double eps = 1e-5;
    const double rhoL = smax(L[0], eps), iL = 1.0 / rhoL;
    const double uL = L[1] * iL, vL = L[2] * iL, EL = L[3] * iL;
    const double pL = smax((gamma - 1.0) * (L[3] - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    const double aL = sqrt(smax(gamma * pL * iL, eps));
    const double rhoR = smax(R[0], eps), iR = 1.0 / rhoR;
    const double uR = R[1] * iR, vR = R[2] * iR, ER = R[3] * iR;
    const double pR = smax((gamma - 1.0) * (R[3] - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    const double aR = sqrt(smax(gamma * pR * iR, eps));

    const double sqL = sqrt(rhoL), sqR = sqrt(rhoR);
    const double iden = 1.0 / smax(sqL + sqR, eps);
    const double u_roe = (sqL * uL + sqR * uR) * iden;
    const double h_roe = (sqL * (EL + pL * iL) + sqR * (ER + pR * iR)) * iden;
    const double c_roe = sqrt(smax((gamma - 1.0) * (h_roe - 0.5 * u_roe * u_roe), eps));
    const double SL = smin(uL - aL, u_roe - c_roe);
    const double SR = smax(uR + aR, u_roe + c_roe);

    if (SL >= 0.0) {
        F[0] = rhoL * uL; F[1] = rhoL * uL * uL + pL; F[2] = rhoL * uL * vL; F[3] = (L[3] + pL) * uL;
        return;
    }
    if (SR <= 0.0) {
        F[0] = rhoR * uR; F[1] = rhoR * uR * uR + pR; F[2] = rhoR * uR * vR; F[3] = (R[3] + pR) * uR;
        return;
    }
    const double num = pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR);
    double den2 = rhoL * (SL - uL) - rhoR * (SR - uR);
    if (fabs(den2) < eps) den2 = (den2 >= 0 ? eps : -eps);
    const double Sstar = num / den2;

    const bool left = Sstar >= 0.0;
    const double rho = left ? rhoL : rhoR, u = left ? uL : uR, v = left ? vL : vR;
    const double E = left ? EL : ER, p = left ? pL : pR, S = left ? SL : SR;
    const double U0 = left ? L[0] : R[0], U1 = left ? L[1] : R[1], U2 = left ? L[2] : R[2],
                 U3 = left ? L[3] : R[3];
    const double FF0 = rho * u, FF1 = rho * u * u + p, FF2 = rho * u * v, FF3 = (U3 + p) * u;
    double denom = S - Sstar;
    if (fabs(denom) < eps) denom = (denom >= 0 ? eps : -eps);
    const double q = rho * (S - u);
    const double c = q / denom;
    double term = (Sstar - u) * (Sstar + p / q);
    if (isnan(term) || isinf(term)) term = 0.0;
    F[0] = FF0 + S * (c - U0);
    F[1] = FF1 + S * (c * Sstar - U1);
    F[2] = FF2 + S * (c * v - U2);
    F[3] = FF3 + S * (c * (E + term) - U3);
}

// HLL with Roe-averaged speed bounds (hlleFlux, src_DG/fluid.cpp:525-580).
__device__ __forceinline__ void hll_f64(const double L[4], const double R[4], double gamma, double F[4])
{
    constexpr double eps = 1e-6;
    const double rhoL = smax(L[0], eps), uL = L[1] / rhoL, vL = L[2] / rhoL, EL = L[3];
    const double pL = smax((gamma - 1.0) * (EL - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    const double aL = sqrt(smax(gamma * pL / rhoL, eps));
    const double HL = (EL + pL) / rhoL;
    const double rhoR = smax(R[0], eps), uR = R[1] / rhoR, vR = R[2] / rhoR, ER = R[3];
    const double pR = smax((gamma - 1.0) * (ER - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    const double aR = sqrt(smax(gamma * pR / rhoR, eps));
    const double HR = (ER + pR) / rhoR;
    const double sqL = sqrt(rhoL), sqR = sqrt(rhoR);
    const double invDen = 1.0 / smax(sqL + sqR, eps);
    const double u_roe = (sqL * uL + sqR * uR) * invDen;
    const double v_roe = (sqL * vL + sqR * vR) * invDen;
    const double h_roe = (sqL * HL + sqR * HR) * invDen;
    const double c_roe = sqrt(smax((gamma - 1.0) * (h_roe - 0.5 * (u_roe * u_roe + v_roe * v_roe)), eps));
    const double SL = smin(uL - aL, u_roe - c_roe);
    const double SR = smax(uR + aR, u_roe + c_roe);
    const double FL0 = rhoL * uL, FL1 = rhoL * uL * uL + pL, FL2 = rhoL * uL * vL, FL3 = (EL + pL) * uL;
    const double FR0 = rhoR * uR, FR1 = rhoR * uR * uR + pR, FR2 = rhoR * uR * vR, FR3 = (ER + pR) * uR;
    if (SL >= 0.0) {
        F[0] = FL0; F[1] = FL1; F[2] = FL2; F[3] = FL3;
    } else if (SR <= 0.0) {
        F[0] = FR0; F[1] = FR1; F[2] = FR2; F[3] = FR3;
    } else {
        const double invDiff = 1.0 / smax(SR - SL, eps);
        F[0] = (SR * FL0 - SL * FR0 + SL * SR * (R[0] - L[0])) * invDiff;
        F[1] = (SR * FL1 - SL * FR1 + SL * SR * (R[1] - L[1])) * invDiff;
        F[2] = (SR * FL2 - SL * FR2 + SL * SR * (R[2] - L[2])) * invDiff;
        F[3] = (SR * FL3 - SL * FR3 + SL * SR * (R[3] - L[3])) * invDiff;
    }
}

// Block-wide max of a non-negative double, merged into *out with one
// atomicMax on the IEEE bit pattern (order-preserving for values >= 0).
// NaN never wins: the caller only raises m on a true comparison.
__device__ __forceinline__ void block_max_to(double m, unsigned long long *out)
{
    __shared__ double red[16];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(m, off, 64);
        if (o > m) m = o;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        double b = red[0];
        for (int w = 1; w < nw; ++w)
            if (red[w] > b) b = red[w];
        if (b > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}

static __global__ void zero_u64_kernel(unsigned long long *p)
{
    if (threadIdx.x == 0) *p = 0ull;
}

}  // namespace zk
