/* sanitize_check.c -- runs every entry point of the oracle (zk_oracle.c) on
 * small grids, for an AddressSanitizer / UndefinedBehaviorSanitizer build of
 * the oracle (test infrastructure; built and run by tests/test_sanitizers.py).
 * Grids have irregular masks, slip and periodic codes, and the minimum size
 * (7 x 7), so the fills, stencils and limiters touch every edge case of
 * their index arithmetic.  Prints one line and exits 0 when it finishes. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "zk_oracle.h"

static double frand(unsigned *s)
{
    *s = *s * 1664525u + 1013904223u;
    return (double)(*s >> 8) / 16777216.0;
}

static void mask(int ny, int nx, int code, int obst, int32_t *b, unsigned *seed)
{
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const int ring = i < 3 || i >= ny - 3 || j < 3 || j >= nx - 3;
            b[i * nx + j] = ring ? code : 0;
            if (!ring && obst && frand(seed) < 0.1) b[i * nx + j] = frand(seed) < 0.5 ? -1 : 1;
        }
}

static void state(int ny, int nx, int nm, double gamma, double *Q, unsigned *seed)
{
    const size_t np = (size_t)ny * nx;
    double *r = malloc(np * sizeof(double)), *u = malloc(np * sizeof(double)), *v = malloc(np * sizeof(double)),
           *p = malloc(np * sizeof(double)), *U = malloc(4 * np * sizeof(double));
    for (size_t k = 0; k < np; ++k) {
        r[k] = 0.2 + frand(seed);
        u[k] = frand(seed) - 0.5;
        v[k] = frand(seed) - 0.5;
        p[k] = 0.2 + frand(seed);
    }
    zko_prim_to_cons((int)np, gamma, r, u, v, p, U);
    memset(Q, 0, 4 * (size_t)nm * np * sizeof(double));
    for (int c = 0; c < 4; ++c) {
        memcpy(Q + (size_t)c * nm * np, U + (size_t)c * np, np * sizeof(double));
        for (int m = 1; m < nm; ++m)
            for (size_t k = 0; k < np; ++k) Q[((size_t)c * nm + m) * np + k] = 0.05 * (frand(seed) - 0.5);
    }
    free(r); free(u); free(v); free(p); free(U);
}

int main(void)
{
    unsigned seed = 7;
    const double g = 1.4;
    int runs = 0;
    double chk = 0.0;
    const int sizes[3][2] = {{7, 7}, {9, 13}, {24, 20}};
    for (int si = 0; si < 3; ++si) {
        const int ny = sizes[si][0], nx = sizes[si][1];
        const size_t np = (size_t)ny * nx;
        int32_t *b = malloc(np * sizeof(int32_t));
        for (int code = -2; code <= 1; code += 3) {      /* -2 periodic ring, 1 slip ring */
            for (int obst = 0; obst <= 1; ++obst) {
                mask(ny, nx, code == 1 ? -1 : code, obst, b, &seed);
                if (code == 1) for (int j = 0; j < nx; ++j) b[j] = 1;  /* a slip top edge */
                /* finite volume: WENO5 and MUSCL, HLLC and HLL, positivity, slip walls */
                double *U = malloc(4 * np * sizeof(double)), *dU = malloc(4 * np * sizeof(double));
                const int fluxes[6] = {0, 1, 16, 1 << 5, (2 << 5) | 16, 128};
                for (int f = 0; f < 6; ++f) {
                    state(ny, nx, 1, g, U, &seed);
                    chk += zko_fv_max_speed(ny, nx, g, U);
                    zko_fv_rhs(ny, nx, 0.1, g, b, U, 1e-3, fluxes[f], dU);
                    for (int s = 0; s < 3; ++s) zko_fv_step(ny, nx, 0.1, g, b, U, 1e-3, fluxes[f]);
                    chk += U[np / 2];
                    ++runs;
                }
                zko_fv_fill_ghosts(ny, nx, b, U);
                zko_fv_fill_ghosts_slip(ny, nx, b, U);
                free(U); free(dU);
                /* modal DG P = 1..3 with every limiter and both fluxes */
                for (int P = 1; P <= 3; ++P) {
                    const int nm = (P + 1) * (P + 1);
                    double *Q = malloc(4 * (size_t)nm * np * sizeof(double)), *dQ = malloc(4 * (size_t)nm * np * sizeof(double));
                    uint8_t *fl = malloc(np);
                    for (int lim = 0; lim <= 2; ++lim)
                        for (int flux = 0; flux <= 1; ++flux) {
                            state(ny, nx, nm, g, Q, &seed);
                            chk += zko_dg_max_speed(ny, nx, P, g, b, Q);
                            zko_dg_rhs(ny, nx, P, 0.1, g, b, Q, flux, dQ);
                            for (int s = 0; s < 2; ++s) zko_dg_step(ny, nx, P, 0.1, g, b, Q, 1e-4, flux, lim, 50.0, fl);
                            chk += Q[np / 2];
                            ++runs;
                        }
                    zko_dg_fill_ghosts(ny, nx, P, b, Q);
                    zko_dg_limit_tvb(ny, nx, P, 0.1, 50.0, b, Q, fl);
                    zko_dg_limit_persson_zs(ny, nx, P, g, b, Q);
                    zko_dg_limit_zs(ny, nx, P, g, Q);
                    chk += zko_dg_cfl_limit(P);
                    free(Q); free(dQ); free(fl);
                }
            }
        }
        free(b);
    }
    /* point kernels on extreme states (floors, vacuum, supersonic) */
    const double st[4][4] = {{1, 0, 0, 2.5}, {1e-9, 1e-9, 0, 1e-9}, {-1, 0, 0, 2.5}, {1, 30, 0, 500}};
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) {
            double F[4], sp[3];
            zko_hllc_fv(st[a], st[c], g, F); chk += F[0];
            zko_hllc_dg(st[a], st[c], g, F); chk += F[0];
            zko_hll(st[a], st[c], g, F); chk += F[0];
            zko_hllc_speeds(a & 1, st[a], st[c], g, sp); chk += sp[0];
        }
    for (int n = 0; n <= 4; ++n) chk += zko_legP(n, 0.3) + zko_legD(n, -0.7);
    for (int q = 1; q <= 5; ++q) chk += zko_gl_points(q)[0] + zko_gl_weights(q)[q - 1];
    const double s6[6] = {1, 2, 4, 3, 1, 0};
    double l, r;
    zko_weno5(s6, &l, &r); chk += l + r;
    for (int lim = 0; lim <= 3; ++lim) { zko_muscl(s6, lim, &l, &r); chk += l + r; }
    printf("sanitize_check: %d runs, checksum %.6e\n", runs, chk);
    return 0;
}
