// fv_face.h -- one WENO5 + HLLC/HLL face of the finite-volume stage kernel.
//
// Two arithmetic forms of the same scheme (weno5_reconstruct and hllcFlux of
// src_WENO/fluid.cpp:230-317):
//   * FP64: the reference expressions with the divides merged.  The three
//     nonlinear weights d_k/(eps+beta_k)^2 share one denominator, so a face
//     side needs one divide instead of four; HLLC forms 1/rho and the Roe
//     denominator once.  Divisors and radicands that are floored away from
//     zero use div_nz / rcp_nz / sqrt_pos (zk_device.h).  Results differ from
//     the reference only by rounding.
//   * FP32 bulk: reconstruction on FP32 deltas relative to the face's left
//     cell (WENO5 is translation-equivariant, so the FP32 error scales with
//     the local jump rather than the magnitude), added back in FP64; pressure
//     recovery p = (g-1)(E - |m|^2 / 2 rho) and its positivity floor in FP64;
//     wave speeds, star states and fluxes in FP32 with hardware rcp/sqrt.
#pragma once

#include <hip/hip_runtime.h>

#include "zk_device.h"

namespace zk {

// ---------------------------------------------------------------- FP64
__device__ __forceinline__ void weno5_pair_f64(double UL2, double UL1, double UP, double UR1, double UR2, double UR3,
                                               double &UL, double &UR)
{
    constexpr double eps = 1e-6, C1 = 13.0 / 12.0, C2 = 0.25;
    const double qL0 = (1.0 / 3.0) * UL2 - (7.0 / 6.0) * UL1 + (11.0 / 6.0) * UP;
    const double qL1 = -(1.0 / 6.0) * UL1 + (5.0 / 6.0) * UP + (1.0 / 3.0) * UR1;
    const double qL2 = (1.0 / 3.0) * UP + (5.0 / 6.0) * UR1 - (1.0 / 6.0) * UR2;
    const double d0 = UL2 - 2 * UL1 + UP, d1 = UL2 - 4 * UL1 + 3 * UP;
    const double d2 = UL1 - 2 * UP + UR1, d3 = UL1 - UR1;
    const double d4 = UP - 2 * UR1 + UR2, d5 = 3 * UP - 4 * UR1 + UR2;
    // eps + beta_k, accumulated as two FMAs onto eps (eps enters first: one
    // instruction fewer per indicator than adding it last)
    auto ind = [&](double a, double b) { return fma(C1 * a, a, fma(C2 * b, b, eps)); };
    const double t0 = ind(d0, d1);
    const double t1 = ind(d2, d3);
    const double t2 = ind(d4, d5);
    const double s0 = t0 * t0, s1 = t1 * t1, s2 = t2 * t2;
    // alpha_k proportional to d_k * prod_{l != k} s_l, scaled by 10 (the ratio
    // below is unchanged): d = (1, 6, 3) / 10
    const double a0 = s1 * s2, a1 = 6.0 * (s0 * s2), a2 = 3.0 * (s0 * s1);
    UL = div_nz(a0 * qL0 + a1 * qL1 + a2 * qL2, a0 + a1 + a2);  // a_k >= eps^4

    const double qR0 = (11.0 / 6.0) * UR1 - (7.0 / 6.0) * UR2 + (1.0 / 3.0) * UR3;
    const double qR1 = qL2;
    const double qR2 = qL1;
    const double r0 = UR1 - 2 * UR2 + UR3, r1 = 3 * UR1 - 4 * UR2 + UR3;
    const double r5 = UL1 - 4 * UP + 3 * UR1;
    const double u0 = ind(r0, r1);
    const double u1 = ind(d4, UP - UR2);
    const double u2 = ind(d2, r5);
    const double v0 = u0 * u0, v1 = u1 * u1, v2 = u2 * u2;
    const double b0 = v1 * v2, b1 = 6.0 * (v0 * v2), b2 = 3.0 * (v0 * v1);
    UR = div_nz(b0 * qR0 + b1 * qR1 + b2 * qR2, b0 + b1 + b2);
}

// HLLC (fluid.cpp:268-317), normal velocity = component 1.
__device__ __forceinline__ void hllc_fv_fast_f64(const double L[4], const double R[4], double gamma, double F[4])
{
    constexpr double eps = 1e-12;
    const double gm1 = gamma - 1.0;
    const double rhoL = smax(L[0], eps), irL = rcp_nz(rhoL);
    const double uL = L[1] * irL, vL = L[2] * irL, EL = L[3] * irL;
    const double pL = smax(gm1 * (L[3] - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    const double aL = sqrt_pos(gamma * pL * irL);
    const double rhoR = smax(R[0], eps), irR = rcp_nz(rhoR);
    const double uR = R[1] * irR, vR = R[2] * irR, ER = R[3] * irR;
    const double pR = smax(gm1 * (R[3] - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    const double aR = sqrt_pos(gamma * pR * irR);

    const double sqL = sqrt_pos(rhoL), sqR = sqrt_pos(rhoR), iden = rcp_nz(sqL + sqR);
    const double u_roe = (sqL * uL + sqR * uR) * iden;
    const double h_roe = (sqL * (EL + pL * irL) + sqR * (ER + pR * irR)) * iden;
    const double c_roe = sqrt_pos(smax(gm1 * (h_roe - 0.5 * u_roe * u_roe), eps));
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
    const double mL = rhoL * (SL - uL), mR = rhoR * (SR - uR);
    const double num = pR - pL + uL * mL - uR * mR;
    double den2 = mL - mR;
    if (fabs(den2) < 1e-12) den2 = (den2 >= 0 ? 1e-12 : -1e-12);
    const double Sstar = div_nz(num, den2);

    const bool left = Sstar >= 0.0;
    const double rho = left ? rhoL : rhoR, u = left ? uL : uR, v = left ? vL : vR;
    const double E = left ? EL : ER, p = left ? pL : pR, S = left ? SL : SR, m = left ? mL : mR;
    const double U0 = left ? L[0] : R[0], U1 = left ? L[1] : R[1], U2 = left ? L[2] : R[2],
                 U3 = left ? L[3] : R[3];
    const double c = m / (S - Sstar);  // this and p / m are unguarded: IEEE divisions
    F[0] = rho * u + S * (c - U0);
    F[1] = rho * u * u + p + S * (c * Sstar - U1);
    F[2] = rho * u * v + S * (c * v - U2);
    F[3] = (U3 + p) * u + S * (c * (E + (Sstar - u) * (Sstar + p / m)) - U3);
}

// MUSCL at the face between UL1 and UR1 from UL2, UL1, UR1, UR2; limiter 0 =
// minmod, 1 = van Leer, 2 = superbee (muscl_reconstruct, src_WENO/fluid.cpp:
// 203-225, which the reference defines but its solver never calls; here a
// selectable reconstruction, FP64 in both precisions).
__device__ __forceinline__ double minmod2_f64(double a, double b)
{
    return a * b <= 0.0 ? 0.0 : (fabs(a) < fabs(b) ? a : b);
}
__device__ __forceinline__ double muscl_slope(double a, double b, int limiter)
{
    if (limiter == 0) return minmod2_f64(a, b);
    if (limiter == 1) {
        const double ab = a * b;
        return ab <= 0.0 ? 0.0 : 2.0 * ab / (a + b);
    }
    const double s1 = minmod2_f64(a, 2.0 * b), s2 = minmod2_f64(2.0 * a, b);
    return s1 * s2 <= 0.0 ? 0.0 : (fabs(s1) > fabs(s2) ? s1 : s2);
}
__device__ __forceinline__ void muscl_f64(double UL2, double UL1, double UR1, double UR2, int limiter, double &UL,
                                          double &UR)
{
    const double dL = UL1 - UL2, dC = UR1 - UL1, dR = UR2 - UR1;
    UL = UL1 + 0.5 * muscl_slope(dL, dC, limiter);
    UR = UR1 - 0.5 * muscl_slope(dC, dR, limiter);
}

// ---------------------------------------------------------------- FP32 bulk
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ void weno5_pair_f32(float UL2, float UL1, float UP, float UR1, float UR2, float UR3,
                                               float &UL, float &UR)
{
    constexpr float eps = 1e-6f, C1 = 13.0f / 12.0f, C2 = 0.25f;
    const float qL0 = (1.0f / 3.0f) * UL2 - (7.0f / 6.0f) * UL1 + (11.0f / 6.0f) * UP;
    const float qL1 = -(1.0f / 6.0f) * UL1 + (5.0f / 6.0f) * UP + (1.0f / 3.0f) * UR1;
    const float qL2 = (1.0f / 3.0f) * UP + (5.0f / 6.0f) * UR1 - (1.0f / 6.0f) * UR2;
    const float d0 = UL2 - 2 * UL1 + UP, d1 = UL2 - 4 * UL1 + 3 * UP;
    const float d2 = UL1 - 2 * UP + UR1, d3 = UL1 - UR1;
    const float d4 = UP - 2 * UR1 + UR2, d5 = 3 * UP - 4 * UR1 + UR2;
    // 1/t_k^2 stays inside the FP32 range for t_k < 1e19 (deltas keep t_k small)
    const float t0 = eps + (C1 * d0 * d0 + C2 * d1 * d1);
    const float t1 = eps + (C1 * d2 * d2 + C2 * d3 * d3);
    const float t2 = eps + (C1 * d4 * d4 + C2 * d5 * d5);
    const float it0 = frcp(t0), it1 = frcp(t1), it2 = frcp(t2);
    const float a0 = 0.1f * it0 * it0, a1 = 0.6f * it1 * it1, a2 = 0.3f * it2 * it2;
    UL = (a0 * qL0 + a1 * qL1 + a2 * qL2) * frcp(a0 + a1 + a2);

    const float qR0 = (11.0f / 6.0f) * UR1 - (7.0f / 6.0f) * UR2 + (1.0f / 3.0f) * UR3;
    const float r0 = UR1 - 2 * UR2 + UR3, r1 = 3 * UR1 - 4 * UR2 + UR3;
    const float r5 = UL1 - 4 * UP + 3 * UR1;
    const float u0 = eps + (C1 * r0 * r0 + C2 * r1 * r1);
    const float u1 = eps + (C1 * d4 * d4 + C2 * (UP - UR2) * (UP - UR2));
    const float u2 = eps + (C1 * d2 * d2 + C2 * r5 * r5);
    const float iu0 = frcp(u0), iu1 = frcp(u1), iu2 = frcp(u2);
    const float b0 = 0.1f * iu0 * iu0, b1 = 0.6f * iu1 * iu1, b2 = 0.3f * iu2 * iu2;
    UR = (b0 * qR0 + b1 * qL2 + b2 * qL1) * frcp(b0 + b1 + b2);
}

// HLLC with FP64 pressure recovery and positivity floors, FP32 elsewhere.
__device__ __forceinline__ void hllc_fv_f32(const double L[4], const double R[4], double gamma, double F[4])
{
    constexpr double eps = 1e-12;
    const double rhoLd = smax(L[0], eps), rhoRd = smax(R[0], eps);
    const double pLd = smax((gamma - 1.0) * (L[3] - div_nz(0.5 * (L[1] * L[1] + L[2] * L[2]), rhoLd)), eps);
    const double pRd = smax((gamma - 1.0) * (R[3] - div_nz(0.5 * (R[1] * R[1] + R[2] * R[2]), rhoRd)), eps);
    const float g = (float)gamma, gm1 = (float)(gamma - 1.0);
    const float rhoL = (float)rhoLd, rhoR = (float)rhoRd, pL = (float)pLd, pR = (float)pRd;
    const float irL = frcp(rhoL), irR = frcp(rhoR);
    const float uL = (float)L[1] * irL, vL = (float)L[2] * irL, EL = (float)L[3] * irL;
    const float uR = (float)R[1] * irR, vR = (float)R[2] * irR, ER = (float)R[3] * irR;
    const float aL = fsqrt(g * pL * irL), aR = fsqrt(g * pR * irR);
    const float sqL = fsqrt(rhoL), sqR = fsqrt(rhoR), iden = frcp(sqL + sqR);
    const float u_roe = (sqL * uL + sqR * uR) * iden;
    const float h_roe = (sqL * (EL + pL * irL) + sqR * (ER + pR * irR)) * iden;
    const float c_roe = fsqrt(fmaxf(gm1 * (h_roe - 0.5f * u_roe * u_roe), 1e-12f));
    const float SL = fminf(uL - aL, u_roe - c_roe);
    const float SR = fmaxf(uR + aR, u_roe + c_roe);
    const float EtL = (float)L[3], EtR = (float)R[3];
    if (SL >= 0.0f) {
        F[0] = rhoL * uL; F[1] = rhoL * uL * uL + pL; F[2] = rhoL * uL * vL; F[3] = (EtL + pL) * uL;
        return;
    }
    if (SR <= 0.0f) {
        F[0] = rhoR * uR; F[1] = rhoR * uR * uR + pR; F[2] = rhoR * uR * vR; F[3] = (EtR + pR) * uR;
        return;
    }
    const float mL = rhoL * (SL - uL), mR = rhoR * (SR - uR);
    const float num = pR - pL + uL * mL - uR * mR;
    float den2 = mL - mR;
    if (fabsf(den2) < 1e-12f) den2 = (den2 >= 0 ? 1e-12f : -1e-12f);
    const float Sstar = num * frcp(den2);
    const bool left = Sstar >= 0.0f;
    const float rho = left ? rhoL : rhoR, u = left ? uL : uR, v = left ? vL : vR;
    const float E = left ? EL : ER, p = left ? pL : pR, S = left ? SL : SR, m = left ? mL : mR;
    const float U1 = left ? (float)L[1] : (float)R[1], U2 = left ? (float)L[2] : (float)R[2];
    const float U3 = left ? EtL : EtR;
    const float c = m * frcp(S - Sstar);
    F[0] = rho * u + S * (c - (left ? (float)L[0] : (float)R[0]));
    F[1] = rho * u * u + p + S * (c * Sstar - U1);
    F[2] = rho * u * v + S * (c * v - U2);
    F[3] = (U3 + p) * u + S * (c * (E + (Sstar - u) * (Sstar + p * frcp(m))) - U3);
}

}  // namespace zk
