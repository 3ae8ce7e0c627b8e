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

// Reciprocal, quotient and square root without the special-case handling of
// the IEEE sequences (scaling for subnormal and huge operands, exact zeros and
// infinities): for the flux routines, whose divisors and radicands are floored
// away from zero.  Reciprocal: the hardware estimate refined by two
// Newton-Raphson steps r += r (1 - x r); quotient: one correction
// q += r (a - b q); square root: Goldschmidt's iteration on g ~ sqrt(x),
// h ~ 1 / (2 sqrt(x)) from the reciprocal square root estimate, then two
// corrections g += h (x - g^2).  Over 2.7e8 random operands in [1e-8, 1e8]
// all three equal the correctly rounded results.  NaN propagates; an
// infinite operand gives NaN where IEEE gives 0 or inf.
__device__ __forceinline__ double rcp_nz(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double div_nz(double a, double b)
{
    const double r = rcp_nz(b);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ double sqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    g = fma(fma(-g, g, x), h, g);
    return fma(fma(-g, g, x), h, g);
}

// HLLC of the DG solver (hllcFlux, src_DG/fluid.cpp:450-520): eps 1e-5,
// sound speed sqrt(max(gamma p / rho, eps)), Roe speeds with u only, the
// denominator guards and the non-finite star-term reset.  The divisions by
// rho_L, rho_R and the Roe denominator are one reciprocal each (6 divisions
// instead of 15), which moves results by a few ulps.  The floored divisors
// and radicands use rcp_nz / div_nz / sqrt_pos; p / q keeps the IEEE division,
// whose inf and NaN the star-term reset relies on.
__device__ __forceinline__ void hllc_dg_fast_f64(const double L[4], const double R[4], double gamma, double F[4])
{
    constexpr double eps = 1e-5;
    const double rhoL = smax(L[0], eps), iL = rcp_nz(rhoL);
    const double uL = L[1] * iL, vL = L[2] * iL, EL = L[3] * iL;
    const double pL = smax((gamma - 1.0) * (L[3] - 0.5 * rhoL * (uL * uL + vL * vL)), eps);
    const double aL = sqrt_pos(smax(gamma * pL * iL, eps));
    const double rhoR = smax(R[0], eps), iR = rcp_nz(rhoR);
    const double uR = R[1] * iR, vR = R[2] * iR, ER = R[3] * iR;
    const double pR = smax((gamma - 1.0) * (R[3] - 0.5 * rhoR * (uR * uR + vR * vR)), eps);
    const double aR = sqrt_pos(smax(gamma * pR * iR, eps));

    const double sqL = sqrt_pos(rhoL), sqR = sqrt_pos(rhoR);
    const double iden = rcp_nz(smax(sqL + sqR, eps));
    const double u_roe = (sqL * uL + sqR * uR) * iden;
    const double h_roe = (sqL * (EL + pL * iL) + sqR * (ER + pR * iR)) * iden;
    const double c_roe = sqrt_pos(smax((gamma - 1.0) * (h_roe - 0.5 * u_roe * u_roe), eps));
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
    const double Sstar = div_nz(num, den2);

    const bool left = Sstar >= 0.0;
    const double rho = left ? rhoL : rhoR, u = left ? uL : uR, v = left ? vL : vR;
    const double E = left ? EL : ER, p = left ? pL : pR, S = left ? SL : SR;
    const double U0 = left ? L[0] : R[0], U1 = left ? L[1] : R[1], U2 = left ? L[2] : R[2],
                 U3 = left ? L[3] : R[3];
    const double FF0 = rho * u, FF1 = rho * u * u + p, FF2 = rho * u * v, FF3 = (U3 + p) * u;
    double denom = S - Sstar;
    if (fabs(denom) < eps) denom = (denom >= 0 ? eps : -eps);
    const double q = rho * (S - u);
    const double c = div_nz(q, denom);
    double term = (Sstar - u) * (Sstar + p / q);  // q may be 0: IEEE division, then the inf/NaN guard
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

// Running maximum of wave speeds with the reference's comparison (a NaN
// speed, from a negative pressure of a finite state, never wins; an infinite
// one, from a zero density, does), except that a state whose conserved values
// are not all finite gives +inf, so the wave-speed reduction reports a broken
// state instead of skipping it.
__device__ __forceinline__ bool finite4(double a, double b, double c, double d)
{
    constexpr double M = 1.79769313486231570815e308;
    return fabs(a) <= M && fabs(b) <= M && fabs(c) <= M && fabs(d) <= M;
}
__device__ __forceinline__ double speed_max(double m, double s, bool finite_state)
{
    if (!finite_state) return __builtin_huge_val();
    return s > m ? s : m;
}

// Block-wide max of a non-negative double, merged into *out with one
// atomicMax on the IEEE bit pattern (order-preserving for values >= 0; +inf,
// which speed_max gives a non-finite state, orders above every finite value).
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
