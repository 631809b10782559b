/*
 * eik_oracle.c -- CPU FP64 restatement of SurfATTPP's eikonal hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in surfattpp_amd/ links, imports or calls it.
 *
 * It restates, operation for operation, the double-precision arithmetic of
 *   src/eikonal_solver.cpp:22-317   FSM_UW_PS_lonlat_2d   (forward, factored FSM)
 *   src/eikonal_solver.cpp:323-512  FSM_O1_JSE_lonlat_2d  (adjoint-state FSM)
 *   src/eikonal_solver.cpp:514-536  mask_uniform_grid
 *   include/utils.h:28-36           bilinear
 *   include/utils.h:358-409         interp2d
 *   src/preproc.cpp:19-34,71-148    accumulate_kernels (isotropic) + forward_for_event
 * with the reference's loop order (x outer, y inner), its storage
 * (column-major (ix, iy) -> ix + iy*nx) and its evaluation order of every
 * expression.  It is compiled with -ffp-contract=off so no FMA is formed.
 *
 * Pinning (tests/test_oracle.py): the reference's own known-answer
 * tests (tests/eikonal_test/test_eikonal.cpp tests 1-8) restated; closed
 * forms (great-circle travel time, index-set mask, interp2d on planes);
 * invariants; and the reference's Fortran twin
 * (tests/eikonal_test/eikon_solver_2d.f90:764-1126 forward, :1276-1446
 * adjoint), compiled from its own source into oracle/_ref by oracle/Makefile.
 * Elementary functions.  The reference calls the platform libm (std::sin,
 * std::cos, std::acos, pow(-3)).  Its quadratic update loses ~(dist/dx)^2 of
 * relative precision, so a one-ulp change in T0 moves kernels by ~1e-11 at
 * 200 cells.  By default this file therefore uses correctly rounded values
 * (computed in __float128 with libquadmath and rounded once), which a GPU can
 * reproduce bit for bit; built with -DSAT_ORACLE_GLIBC it calls glibc like the
 * reference does (glibc 2.35 acos is not correctly rounded on ~0.4 % of
 * inputs near 1).  The per-row / per-column sin and cos are evaluated once per
 * row or column instead of per node: the functions are pure, so the values
 * are the same.
 * Parity unpinned: Eigen's reduction order for chi (preproc.cpp:142 uses
 * Eigen's vectorised .sum(); this file sums sequentially).
 *
 * Planted mistakes.  Built with -DSAT_ORACLE_MUTANT=n (oracle/Makefile,
 * liboracle_mutN.so) the file carries one plausible slip, which the pinning
 * tests must reject: 1 a flipped sign in the two-point causality test (:201),
 * 2 the cross-metric term dropped from the left-face flux a1 (:409-413),
 * 3 dx for dy in the two-point y stencil (:181), 4 the near-source fixed box
 * one cell too small in y (:121).  With the default (0) the expressions are
 * the reference's.
 */
#ifndef SAT_ORACLE_MUTANT
#define SAT_ORACLE_MUTANT 0
#endif
#define MUTANT(n) (SAT_ORACLE_MUTANT == (n))
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "eik_oracle.h"

#ifdef SAT_ORACLE_GLIBC
static inline double f_sin(double x) { return sin(x); }
static inline double f_cos(double x) { return cos(x); }
static inline double f_acos(double x) { return acos(x); }
static inline double f_pow_m3(double x) { return pow(x, -3.0); }
#else
#include <quadmath.h>
static inline double f_sin(double x) { return (double)sinq((__float128)x); }
static inline double f_cos(double x) { return (double)cosq((__float128)x); }
static inline double f_acos(double x) { return (double)acosq((__float128)x); }
static inline double f_pow_m3(double x) {
    const __float128 q = (__float128)x;
    return (double)(1.0Q / (q * q * q));
}
#endif

/* include/config.h:54-55,68 */
static const double PI_ = 3.14159265358979323846264338327950288;
static const double R_EARTH_ = 6371.0;
#define DEG2RAD_ (PI_ / 180.0)

/* src/eikonal_solver.cpp:15-16 and :337-339 */
static const int FWD_MAXITER = 2000;
static const double FWD_TOL = 1e-4;
static const int ADJ_MAXITER = 1000;
static const double ADJ_TOL = 1e-9;
static const double ADJ_EPS = 1e-14;

#define IDX(ix, iy) ((ix) + (size_t)(iy) * (size_t)nx)

static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min(a,b) */
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max(a,b) */
static inline int imin(int a, int b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }

/* include/utils.h:28-36 -- bilinear at fractional index (idx0+r1, idy0+r2). */
static double bilinear(const double* M, int nx, int idx0, int idy0, double r1, double r2) {
    return (1 - r1) * (1 - r2) * M[IDX(idx0, idy0)]
         + (1 - r1) * r2 * M[IDX(idx0, idy0 + 1)]
         + r1 * (1 - r2) * M[IDX(idx0 + 1, idy0)]
         + r1 * r2 * M[IDX(idx0 + 1, idy0 + 1)];
}

int sat_oracle_forward(int nx, int ny, const double* xx_deg, const double* yy_deg,
                       const double* spha, const double* sphb, const double* sphc,
                       const double* fun, double x0_deg, double y0_deg,
                       int max_iter, double* T_out, int* iters_out, int* flags_out) {
    /* src/eikonal_solver.cpp:35-36 */
    if (nx < 3 || ny < 3) return SAT_ORACLE_EINVAL;
    const int maxiter = (max_iter > 0) ? max_iter : FWD_MAXITER;
    const size_t N = (size_t)nx * (size_t)ny;
    int flags = 0;

    double* xx = (double*)malloc(sizeof(double) * (size_t)nx);
    double* yy = (double*)malloc(sizeof(double) * (size_t)ny);
    double* a = (double*)malloc(sizeof(double) * N);
    double* b = (double*)malloc(sizeof(double) * N);
    double* c = (double*)malloc(sizeof(double) * N);
    double* T0 = (double*)malloc(sizeof(double) * N);
    double* T0x = (double*)malloc(sizeof(double) * N);
    double* T0y = (double*)malloc(sizeof(double) * N);
    double* tau = (double*)malloc(sizeof(double) * N);
    double* tau_old = (double*)malloc(sizeof(double) * N);
    unsigned char* ischange = (unsigned char*)malloc(N);
    double* trig = (double*)malloc(sizeof(double) * (size_t)(2 * ny + 2 * nx));
    if (!xx || !yy || !a || !b || !c || !T0 || !T0x || !T0y || !tau || !tau_old || !ischange || !trig) {
        free(xx); free(yy); free(a); free(b); free(c); free(T0); free(T0x); free(T0y);
        free(tau); free(tau_old); free(ischange); free(trig);
        return SAT_ORACLE_ENOMEM;
    }
    double* sin_yy = trig;           /* sin(yy(iy)) */
    double* cos_yy = trig + ny;      /* cos(yy(iy)) */
    double* sin_dl = trig + 2 * ny;  /* sin(xx(ix) - x0) */
    double* cos_dl = sin_dl + nx;    /* cos(xx(ix) - x0) */

    /* :41-47 degrees -> radians */
    for (int i = 0; i < nx; i++) xx[i] = xx_deg[i] * DEG2RAD_;
    for (int j = 0; j < ny; j++) yy[j] = yy_deg[j] * DEG2RAD_;
    const double x0 = x0_deg * DEG2RAD_;
    const double y0 = y0_deg * DEG2RAD_;
    const double dx = xx[1] - xx[0];
    const double dy = yy[1] - yy[0];

    /* :55-65 metric tensor */
    for (int iy = 0; iy < ny; iy++) {
        sin_yy[iy] = f_sin(yy[iy]);
        cos_yy[iy] = f_cos(yy[iy]);
    }
    for (int iy = 0; iy < ny; iy++) {
        const double cos_y = cos_yy[iy];
        const double cos2_y = cos_y * cos_y;
        const double R2 = R_EARTH_ * R_EARTH_;
        for (int ix = 0; ix < nx; ix++) {
            a[IDX(ix, iy)] = spha[IDX(ix, iy)] / (R2 * cos2_y);
            b[IDX(ix, iy)] = sphb[IDX(ix, iy)] / R2;
            c[IDX(ix, iy)] = sphc[IDX(ix, iy)] / (R_EARTH_ * R_EARTH_ * cos_y);
        }
    }

    /* :71-82 source cell and bilinear slowness at the source */
    int idx0 = (int)floor((x0 - xx[0]) / dx);
    int idy0 = (int)floor((y0 - yy[0]) / dy);
    idx0 = imax(0, imin(idx0, nx - 2));
    idy0 = imax(0, imin(idy0, ny - 2));
    const double r1 = dmin(1.0, (x0 - xx[idx0]) / dx);
    const double r2 = dmin(1.0, (y0 - yy[idy0]) / dy);
    const double fun0 = bilinear(fun, nx, idx0, idy0, r1, r2);

    /* :96-131 analytic factor T0, its gradient, tau initialisation */
    const double sin_y0 = f_sin(y0), cos_y0 = f_cos(y0);
    for (int ix = 0; ix < nx; ix++) {
        const double dlon = xx[ix] - x0;
        sin_dl[ix] = f_sin(dlon);
        cos_dl[ix] = f_cos(dlon);
    }
    for (int ix = 0; ix < nx; ix++) {
        for (int iy = 0; iy < ny; iy++) {
            const double tmp = sin_yy[iy] * sin_y0 + cos_yy[iy] * cos_y0 * cos_dl[ix];
            const double tmp_c = dmax(-1.0, dmin(1.0, tmp));
            T0[IDX(ix, iy)] = f_acos(tmp_c) * R_EARTH_ * fun0;
            const double sin_d = sqrt(dmax(0.0, 1.0 - tmp_c * tmp_c));
            if (sin_d < 1e-14) {
                T0x[IDX(ix, iy)] = 0.0;
                T0y[IDX(ix, iy)] = 0.0;
            } else {
                T0x[IDX(ix, iy)] = cos_yy[iy] * cos_y0 * sin_dl[ix] / sin_d * R_EARTH_ * fun0;
                T0y[IDX(ix, iy)] = -(cos_yy[iy] * sin_y0 - sin_yy[iy] * cos_y0 * cos_dl[ix])
                                   / sin_d * R_EARTH_ * fun0;
            }
            if (fabs((xx[ix] - x0) / dx) <= 1.0 && fabs((yy[iy] - y0) / dy) <= (MUTANT(4) ? 0.5 : 1.0)) {
                tau[IDX(ix, iy)] = 1.0;
                ischange[IDX(ix, iy)] = 0;
                if (ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1) flags |= SAT_FLAG_SRC_ON_BOUNDARY;
            } else {
                tau[IDX(ix, iy)] = 10.0;
                ischange[IDX(ix, iy)] = 1;
            }
        }
    }

    /* :137-140 sweep table */
    const int SX0[4] = {0, 0, nx - 1, nx - 1};
    const int SX1[4] = {nx - 1, nx - 1, 0, 0};
    const int SY0[4] = {0, ny - 1, 0, ny - 1};
    const int SY1[4] = {ny - 1, 0, ny - 1, 0};
    const int SXq[4] = {-1, -1, +1, +1};
    const int SYq[4] = {-1, +1, -1, +1};

    int iters = maxiter;
    int converged = 0;
    for (int iter = 0; iter < maxiter; iter++) {
        memcpy(tau_old, tau, sizeof(double) * N); /* :147 */
        for (int s = 0; s < 4; s++) {
            const int xd = (SX1[s] >= SX0[s]) ? 1 : -1;
            const int yd = (SY1[s] >= SY0[s]) ? 1 : -1;
            for (int ix = SX0[s]; ix != SX1[s] + xd; ix += xd) {
                for (int iy = SY0[s]; iy != SY1[s] + yd; iy += yd) {
                    if (ischange[IDX(ix, iy)] == 0) continue;
                    const double t0 = T0[IDX(ix, iy)];
                    const double t0x = T0x[IDX(ix, iy)];
                    const double t0y = T0y[IDX(ix, iy)];
                    const double aij = a[IDX(ix, iy)];
                    const double bij = b[IDX(ix, iy)];
                    const double cij = c[IDX(ix, iy)];
                    const double fij = fun[IDX(ix, iy)];
                    const double det = aij * bij - cij * cij;
                    double tnew = tau[IDX(ix, iy)];

                    /* :176-219 two-point quadrant stencils */
                    for (int k = 0; k < 4; k++) {
                        const int sx = SXq[k], sy = SYq[k];
                        const int nix = ix + sx, niy = iy + sy;
                        if (nix < 0 || nix >= nx || niy < 0 || niy >= ny) continue;
                        const double px = t0x - sx * (t0 / dx);
                        const double qx = sx * (t0 / dx) * tau[IDX(ix + sx, iy)];
                        const double py = t0y - sy * (t0 / (MUTANT(3) ? dx : dy));
                        const double qy = sy * (t0 / dy) * tau[IDX(ix, iy + sy)];
                        const double eq_a = aij * px * px + bij * py * py - 2.0 * cij * px * py;
                        const double eq_b = 2.0 * (aij * px * qx + bij * py * qy - cij * (px * qy + py * qx));
                        const double eq_c = aij * qx * qx + bij * qy * qy - 2.0 * cij * qx * qy - fij * fij;
                        const double Delta = eq_b * eq_b - 4.0 * eq_a * eq_c;
                        if (Delta < 0.0 || fabs(eq_a) < 1e-30) continue;
                        const double sq = sqrt(Delta);
                        const double two_a = 2.0 * eq_a;
                        for (int sgn = 0; sgn < 2; sgn++) {
                            const double tmp_tau = (-eq_b + (sgn == 0 ? sq : -sq)) / two_a;
                            const double Tx = px * tmp_tau + qx;
                            const double Ty = py * tmp_tau + qy;
                            const double char_x = aij * Tx - cij * Ty;
                            const double char_y = bij * Ty - cij * Tx;
                            if ((MUTANT(1) ? sx * char_x : -sx * char_x) >= 0.0 && -sy * char_y >= 0.0)
                                tnew = dmin(tnew, tmp_tau);
                        }
                    }

                    /* :224-279 one-point axis stencils */
                    if (ix > 0) {
                        const double px = t0x + t0 / dx;
                        const double qx = -(t0 / dx) * tau[IDX(ix - 1, iy)];
                        if (det > 0.0) {
                            const double dis = sqrt(fij * fij * bij / det);
                            for (int sgn = 0; sgn < 2; sgn++) {
                                const double tmp_tau = ((sgn == 0 ? dis : -dis) - qx) / px;
                                if (tmp_tau * t0 >= tau[IDX(ix - 1, iy)] * T0[IDX(ix - 1, iy)] &&
                                    tmp_tau > tau[IDX(ix - 1, iy)] * 0.5)
                                    tnew = dmin(tnew, tmp_tau);
                            }
                        }
                    }
                    if (ix < nx - 1) {
                        const double px = t0x - t0 / dx;
                        const double qx = (t0 / dx) * tau[IDX(ix + 1, iy)];
                        if (det > 0.0) {
                            const double dis = sqrt(fij * fij * bij / det);
                            for (int sgn = 0; sgn < 2; sgn++) {
                                const double tmp_tau = ((sgn == 0 ? dis : -dis) - qx) / px;
                                if (tmp_tau * t0 >= tau[IDX(ix + 1, iy)] * T0[IDX(ix + 1, iy)] &&
                                    tmp_tau > tau[IDX(ix + 1, iy)] * 0.5)
                                    tnew = dmin(tnew, tmp_tau);
                            }
                        }
                    }
                    if (iy > 0) {
                        const double py = t0y + t0 / dy;
                        const double qy = -(t0 / dy) * tau[IDX(ix, iy - 1)];
                        if (det > 0.0) {
                            const double dis = sqrt(fij * fij * aij / det);
                            for (int sgn = 0; sgn < 2; sgn++) {
                                const double tmp_tau = ((sgn == 0 ? dis : -dis) - qy) / py;
                                if (tmp_tau * t0 >= tau[IDX(ix, iy - 1)] * T0[IDX(ix, iy - 1)] &&
                                    tmp_tau > tau[IDX(ix, iy - 1)] * 0.5)
                                    tnew = dmin(tnew, tmp_tau);
                            }
                        }
                    }
                    if (iy < ny - 1) {
                        const double py = t0y - t0 / dy;
                        const double qy = (t0 / dy) * tau[IDX(ix, iy + 1)];
                        if (det > 0.0) {
                            const double dis = sqrt(fij * fij * aij / det);
                            for (int sgn = 0; sgn < 2; sgn++) {
                                const double tmp_tau = ((sgn == 0 ? dis : -dis) - qy) / py;
                                if (tmp_tau * t0 >= tau[IDX(ix, iy + 1)] * T0[IDX(ix, iy + 1)] &&
                                    tmp_tau > tau[IDX(ix, iy + 1)] * 0.5)
                                    tnew = dmin(tnew, tmp_tau);
                            }
                        }
                    }
                    /* :282-283 accept the minimum (candidates folded in order) */
                    tau[IDX(ix, iy)] = tnew;
                }
            }
        }

        /* :292-302 convergence: L1 and Linf of |dtau|*T0 */
        double L1 = 0.0, Linf = 0.0;
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                const double d = fabs(tau[IDX(ix, iy)] - tau_old[IDX(ix, iy)]) * T0[IDX(ix, iy)];
                L1 += d;
                Linf = dmax(Linf, d);
            }
        L1 /= (double)(nx * ny);
        if (L1 < FWD_TOL && Linf < FWD_TOL) {
            iters = iter + 1;
            converged = 1;
            break;
        }
    }
    if (!converged) flags |= SAT_FLAG_NOT_CONVERGED;

    /* :317 T = tau .* T0 */
    for (size_t k = 0; k < N; k++) T_out[k] = tau[k] * T0[k];

    if (iters_out) *iters_out = iters;
    if (flags_out) *flags_out = flags;
    free(xx); free(yy); free(a); free(b); free(c); free(T0); free(T0x); free(T0y);
    free(tau); free(tau_old); free(ischange); free(trig);
    return SAT_ORACLE_OK;
}

int sat_oracle_adjoint(int nx, int ny, const double* xx_deg, const double* yy_deg,
                       const double* spha, const double* sphb, const double* sphc,
                       const double* T, int nr, const double* xrec_deg, const double* yrec_deg,
                       const double* sourceAdj, int max_iter, double* Ta_out, int* iters_out) {
    if (nx < 3 || ny < 3) return SAT_ORACLE_EINVAL;
    const int maxiter = (max_iter > 0) ? max_iter : ADJ_MAXITER;
    const size_t N = (size_t)nx * (size_t)ny;

    double* xx = (double*)malloc(sizeof(double) * (size_t)nx);
    double* yy = (double*)malloc(sizeof(double) * (size_t)ny);
    double* delta = (double*)calloc(N, sizeof(double));
    double* a = (double*)malloc(sizeof(double) * N);
    double* b = (double*)malloc(sizeof(double) * N);
    double* c = (double*)malloc(sizeof(double) * N);
    double* f8 = (double*)calloc(8 * N, sizeof(double)); /* a1m a1p a2m a2p b1m b1p b2m b2p */
    double* Ta_old = (double*)malloc(sizeof(double) * N);
    if (!xx || !yy || !delta || !a || !b || !c || !f8 || !Ta_old) {
        free(xx); free(yy); free(delta); free(a); free(b); free(c); free(f8); free(Ta_old);
        return SAT_ORACLE_ENOMEM;
    }
    double* a1m = f8 + 0 * N; double* a1p = f8 + 1 * N;
    double* a2m = f8 + 2 * N; double* a2p = f8 + 3 * N;
    double* b1m = f8 + 4 * N; double* b1p = f8 + 5 * N;
    double* b2m = f8 + 6 * N; double* b2p = f8 + 7 * N;
    double* Ta = Ta_out;

    /* src/eikonal_solver.cpp:344-348 */
    for (int i = 0; i < nx; i++) xx[i] = xx_deg[i] * DEG2RAD_;
    for (int j = 0; j < ny; j++) yy[j] = yy_deg[j] * DEG2RAD_;
    const double dx = xx[1] - xx[0];
    const double dy = yy[1] - yy[0];

    /* :356-375 delta source, bilinear spreading, area-normalised */
    for (int ir = 0; ir < nr; ir++) {
        const double xrec = xrec_deg[ir] * DEG2RAD_;
        const double yrec = yrec_deg[ir] * DEG2RAD_;
        int idx0 = (int)floor((xrec - xx[0]) / dx);
        int idy0 = (int)floor((yrec - yy[0]) / dy);
        idx0 = imax(0, imin(idx0, nx - 2));
        idy0 = imax(0, imin(idy0, ny - 2));
        const double r1 = dmin(1.0, (xrec - xx[idx0]) / dx);
        const double r2 = dmin(1.0, (yrec - yy[idy0]) / dy);
        const double area = dx * R_EARTH_ * f_cos(yrec) * dy * R_EARTH_;
        const double w = sourceAdj[ir] / area;
        delta[IDX(idx0, idy0)] += w * (1 - r1) * (1 - r2);
        delta[IDX(idx0, idy0 + 1)] += w * (1 - r1) * r2;
        delta[IDX(idx0 + 1, idy0)] += w * r1 * (1 - r2);
        delta[IDX(idx0 + 1, idy0 + 1)] += w * r1 * r2;
    }

    /* :380-389 metric (note: R2*cos*cos here, R2*(cos*cos) in the forward) */
    for (int ix = 0; ix < nx; ix++) {
        for (int iy = 0; iy < ny; iy++) {
            const double cos_y = f_cos(yy[iy]);
            const double R2 = R_EARTH_ * R_EARTH_;
            a[IDX(ix, iy)] = spha[IDX(ix, iy)] / (R2 * cos_y * cos_y);
            b[IDX(ix, iy)] = sphb[IDX(ix, iy)] / R2;
            c[IDX(ix, iy)] = sphc[IDX(ix, iy)] / (R2 * cos_y);
        }
    }

    /* :407-441 face fluxes and their upwind splits */
#define TT(i, j) T[IDX(i, j)]
    for (int ix = 1; ix < nx - 1; ix++) {
        for (int iy = 1; iy < ny - 1; iy++) {
            const double a1 = -(TT(ix, iy) - TT(ix - 1, iy)) / dx * (a[IDX(ix, iy)] + a[IDX(ix - 1, iy)]) / 2.0
                + (MUTANT(2) ? 0.0
                             : (TT(ix, iy + 1) - TT(ix, iy - 1) + TT(ix - 1, iy + 1) - TT(ix - 1, iy - 1)) / (4.0 * dy)
                                   * (c[IDX(ix, iy)] + c[IDX(ix - 1, iy)]) / 2.0);
            a1m[IDX(ix, iy)] = (a1 - fabs(a1)) / 2.0;
            a1p[IDX(ix, iy)] = (a1 + fabs(a1)) / 2.0;
            const double a2 = -(TT(ix + 1, iy) - TT(ix, iy)) / dx * (a[IDX(ix + 1, iy)] + a[IDX(ix, iy)]) / 2.0
                + (TT(ix + 1, iy + 1) - TT(ix + 1, iy - 1) + TT(ix, iy + 1) - TT(ix, iy - 1)) / (4.0 * dy)
                  * (c[IDX(ix + 1, iy)] + c[IDX(ix, iy)]) / 2.0;
            a2m[IDX(ix, iy)] = (a2 - fabs(a2)) / 2.0;
            a2p[IDX(ix, iy)] = (a2 + fabs(a2)) / 2.0;
            const double b1 = -(TT(ix, iy) - TT(ix, iy - 1)) / dy * (b[IDX(ix, iy)] + b[IDX(ix, iy - 1)]) / 2.0
                + (TT(ix + 1, iy) - TT(ix - 1, iy) + TT(ix + 1, iy - 1) - TT(ix - 1, iy - 1)) / (4.0 * dx)
                  * (c[IDX(ix, iy)] + c[IDX(ix, iy - 1)]) / 2.0;
            b1m[IDX(ix, iy)] = (b1 - fabs(b1)) / 2.0;
            b1p[IDX(ix, iy)] = (b1 + fabs(b1)) / 2.0;
            const double b2 = -(TT(ix, iy + 1) - TT(ix, iy)) / dy * (b[IDX(ix, iy + 1)] + b[IDX(ix, iy)]) / 2.0
                + (TT(ix + 1, iy + 1) - TT(ix - 1, iy + 1) + TT(ix + 1, iy) - TT(ix - 1, iy)) / (4.0 * dx)
                  * (c[IDX(ix, iy + 1)] + c[IDX(ix, iy)]) / 2.0;
            b2m[IDX(ix, iy)] = (b2 - fabs(b2)) / 2.0;
            b2p[IDX(ix, iy)] = (b2 + fabs(b2)) / 2.0;
        }
    }
#undef TT

    /* :446-449 boundary 0, interior 100 */
    for (size_t k = 0; k < N; k++) Ta[k] = 0.0;
    for (int ix = 1; ix < nx - 1; ix++)
        for (int iy = 1; iy < ny - 1; iy++) Ta[IDX(ix, iy)] = 100.0;

    /* :466-469 interior sweep table */
    const int IX0[4] = {1, 1, nx - 2, nx - 2};
    const int IX1[4] = {nx - 2, nx - 2, 1, 1};
    const int IY0[4] = {1, ny - 2, 1, ny - 2};
    const int IY1[4] = {ny - 2, 1, ny - 2, 1};

    int iters = maxiter;
    for (int iter = 0; iter < maxiter; iter++) {
        memcpy(Ta_old, Ta, sizeof(double) * N); /* :472 */
        for (int s = 0; s < 4; s++) {
            const int xd = (IX1[s] >= IX0[s]) ? 1 : -1;
            const int yd = (IY1[s] >= IY0[s]) ? 1 : -1;
            for (int ix = IX0[s]; ix != IX1[s] + xd; ix += xd) {
                for (int iy = IY0[s]; iy != IY1[s] + yd; iy += yd) {
                    const double d = (a2p[IDX(ix, iy)] - a1m[IDX(ix, iy)]) / dx
                                   + (b2p[IDX(ix, iy)] - b1m[IDX(ix, iy)]) / dy;
                    if (fabs(d) < ADJ_EPS) {
                        Ta[IDX(ix, iy)] = 0.0;
                    } else {
                        const double e = (Ta[IDX(ix - 1, iy)] * a1p[IDX(ix, iy)] - Ta[IDX(ix + 1, iy)] * a2m[IDX(ix, iy)]) / dx
                                       + (Ta[IDX(ix, iy - 1)] * b1p[IDX(ix, iy)] - Ta[IDX(ix, iy + 1)] * b2m[IDX(ix, iy)]) / dy;
                        Ta[IDX(ix, iy)] = (delta[IDX(ix, iy)] + e) / d;
                    }
                }
            }
        }
        /* :498-508 convergence */
        double L1 = 0.0, Linf = 0.0;
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                const double dd = fabs(Ta[IDX(ix, iy)] - Ta_old[IDX(ix, iy)]);
                L1 += dd;
                Linf = dmax(Linf, dd);
            }
        L1 /= (double)(nx * ny);
        if (L1 < ADJ_TOL && Linf < ADJ_TOL) {
            iters = iter + 1;
            break;
        }
    }
    if (iters_out) *iters_out = iters;
    free(xx); free(yy); free(delta); free(a); free(b); free(c); free(f8); free(Ta_old);
    return SAT_ORACLE_OK;
}

/* src/eikonal_solver.cpp:514-536 -- coordinates in the caller's units (degrees at preproc.cpp:125). */
void sat_oracle_mask(int nx, int ny, const double* xx, const double* yy, double* Ta, double x0, double y0) {
    if (nx < 2 || ny < 2) return;
    const double dx = xx[1] - xx[0];
    const double dy = yy[1] - yy[0];
    const int ix_lo = imax(0, (int)floor((x0 - 1.5 * dx - xx[0]) / dx) + 1);
    const int ix_hi = imin(nx - 1, (int)ceil((x0 + 1.5 * dx - xx[0]) / dx) - 1);
    const int iy_lo = imax(0, (int)floor((y0 - 1.5 * dy - yy[0]) / dy) + 1);
    const int iy_hi = imin(ny - 1, (int)ceil((y0 + 1.5 * dy - yy[0]) / dy) - 1);
    if (ix_lo <= ix_hi && iy_lo <= iy_hi)
        for (int iy = iy_lo; iy <= iy_hi; iy++)
            for (int ix = ix_lo; ix <= ix_hi; ix++) Ta[IDX(ix, iy)] = 0.0;
}

/* std::lower_bound on an ascending array: first index with v[i] >= q (n if none). */
static int lower_bound_d(const double* v, int n, double q) {
    int lo = 0, cnt = n;
    while (cnt > 0) {
        const int step = cnt / 2;
        const int it = lo + step;
        if (v[it] < q) { lo = it + 1; cnt -= step + 1; } else cnt = step;
    }
    return lo;
}

/* include/utils.h:358-409 */
void sat_oracle_interp2d(int nx, int ny, const double* xgrid, const double* ygrid, const double* z,
                         int nq, const double* xq, const double* yq, double* zq) {
    for (int k = 0; k < nq; ++k) {
        const double qx = xq[k];
        const double qy = yq[k];
        if (qx < xgrid[0] || qx > xgrid[nx - 1] || qy < ygrid[0] || qy > ygrid[ny - 1]) {
            zq[k] = NAN;
            continue;
        }
        int i = lower_bound_d(xgrid, nx, qx);
        if (i == nx) --i;
        if (i > 0 && xgrid[i] > qx) --i;
        i = imax(0, imin(i, nx - 2));
        int j = lower_bound_d(ygrid, ny, qy);
        if (j == ny) --j;
        if (j > 0 && ygrid[j] > qy) --j;
        j = imax(0, imin(j, ny - 2));
        const double x1 = xgrid[i], x2 = xgrid[i + 1];
        const double y1 = ygrid[j], y2 = ygrid[j + 1];
        const double tx = (qx - x1) / (x2 - x1);
        const double ty = (qy - y1) / (y2 - y1);
        zq[k] = (1 - tx) * (1 - ty) * z[IDX(i, j)]
              + (tx) * (1 - ty) * z[IDX(i + 1, j)]
              + (1 - tx) * (ty) * z[IDX(i, j + 1)]
              + (tx) * (ty) * z[IDX(i + 1, j + 1)];
    }
}

/*
 * src/preproc.cpp:71-148 (isotropic, no kden) for a batch of events on one
 * grid, in the caller's event order (the reference iterates its std::map of
 * "{evtname}_{iper}" keys).  Per-period planes are period-major
 * [iper][iy][ix]; the reference's [ix][iy][iper] layout (surf_grid.h:63-65)
 * is transposed by the caller.  adj_s[iper] accumulates Ta .* svel^-3
 * (preproc.cpp:28-31) after the mask (:125).  Returns chi (:142).
 */
int sat_oracle_forward_for_event(int nx, int ny, const double* xgrids, const double* ygrids,
                                 int nperiod, const double* svel, const double* m11,
                                 const double* m12, const double* m22,
                                 int nevent, const int* ev_iper, const double* ev_lon,
                                 const double* ev_lat, const int* ev_rec_off, const int* ev_nrec,
                                 const double* rec_lon, const double* rec_lat,
                                 const double* tt_obs, const double* weight,
                                 int is_calc_adj, double* syn, double* adj_s, double* chi_out,
                                 int* iters_fwd, int* iters_adj, int* flags) {
    return sat_oracle_forward_for_event_kden(nx, ny, xgrids, ygrids, nperiod, svel, m11, m12, m22, nevent,
                                             ev_iper, ev_lon, ev_lat, ev_rec_off, ev_nrec, rec_lon, rec_lat,
                                             tt_obs, weight, is_calc_adj, syn, adj_s, chi_out, iters_fwd,
                                             iters_adj, flags, NULL);
}

/*
 * As sat_oracle_forward_for_event, and with kden_s != NULL the density kernel
 * of is_kden (src/preproc.cpp:127-136): per event a second adjoint solve on the
 * same T from the receiver sources -weight (:128-130), NOT masked (the
 * reference masks adj_field again at :136, not kden_field), accumulated as
 * kden_s[iper] += kden_field .* svel^-3 (:32-34).
 */
int sat_oracle_forward_for_event_kden(int nx, int ny, const double* xgrids, const double* ygrids,
                                      int nperiod, const double* svel, const double* m11,
                                      const double* m12, const double* m22,
                                      int nevent, const int* ev_iper, const double* ev_lon,
                                      const double* ev_lat, const int* ev_rec_off, const int* ev_nrec,
                                      const double* rec_lon, const double* rec_lat,
                                      const double* tt_obs, const double* weight,
                                      int is_calc_adj, double* syn, double* adj_s, double* chi_out,
                                      int* iters_fwd, int* iters_adj, int* flags, double* kden_s) {
    (void)nperiod;
    const size_t N = (size_t)nx * (size_t)ny;
    double* fun = (double*)malloc(sizeof(double) * N);
    double* msphc = (double*)malloc(sizeof(double) * N);
    double* T = (double*)malloc(sizeof(double) * N);
    double* Ta = (double*)malloc(sizeof(double) * N);
    double* adj = NULL;
    int maxrec = 0;
    for (int e = 0; e < nevent; e++) maxrec = imax(maxrec, ev_nrec[e]);
    adj = (double*)malloc(sizeof(double) * (size_t)(maxrec > 0 ? maxrec : 1));
    if (!fun || !msphc || !T || !Ta || !adj) {
        free(fun); free(msphc); free(T); free(Ta); free(adj);
        return SAT_ORACLE_ENOMEM;
    }
    double chi = 0.0;
    int rc = SAT_ORACLE_OK;
    for (int e = 0; e < nevent; e++) {
        const int iper = ev_iper[e];
        const double* sv = svel + (size_t)iper * N;
        const double* a11 = m11 + (size_t)iper * N;
        const double* a12 = m12 + (size_t)iper * N;
        const double* a22 = m22 + (size_t)iper * N;
        for (size_t k = 0; k < N; k++) {
            fun[k] = (1.0 / sv[k]) * 1.0; /* :97-98 */
            msphc[k] = -a12[k];           /* :101 */
        }
        int itf = 0, fl = 0;
        rc = sat_oracle_forward(nx, ny, xgrids, ygrids, a11, a22, msphc, fun, ev_lon[e], ev_lat[e], 0, T,
                                &itf, &fl);
        if (rc) break;
        if (iters_fwd) iters_fwd[e] = itf;
        if (flags) flags[e] = fl;
        const int off = ev_rec_off[e], nr = ev_nrec[e];
        sat_oracle_interp2d(nx, ny, xgrids, ygrids, T, nr, rec_lon + off, rec_lat + off, syn + off); /* :111 */
        if (is_calc_adj) {
            for (int r = 0; r < nr; r++) adj[r] = weight[off + r] * (syn[off + r] - tt_obs[off + r]); /* :115 */
            int ita = 0;
            rc = sat_oracle_adjoint(nx, ny, xgrids, ygrids, a11, a22, msphc, T, nr, rec_lon + off, rec_lat + off,
                                    adj, 0, Ta, &ita);
            if (rc) break;
            if (iters_adj) iters_adj[e] = ita;
            sat_oracle_mask(nx, ny, xgrids, ygrids, Ta, ev_lon[e], ev_lat[e]); /* :125 */
            double* K = adj_s + (size_t)iper * N;
            for (size_t k = 0; k < N; k++) K[k] += Ta[k] * f_pow_m3(sv[k]); /* :29-31 */
            if (kden_s) {
                for (int r = 0; r < nr; r++) adj[r] = -weight[off + r] * 1.0; /* :128-130 */
                rc = sat_oracle_adjoint(nx, ny, xgrids, ygrids, a11, a22, msphc, T, nr, rec_lon + off,
                                        rec_lat + off, adj, 0, Ta, &ita);
                if (rc) break;
                double* D = kden_s + (size_t)iper * N;
                for (size_t k = 0; k < N; k++) D[k] += Ta[k] * f_pow_m3(sv[k]); /* :32-34 */
            }
            double s = 0.0;
            for (int r = 0; r < nr; r++) {
                const double d = syn[off + r] - tt_obs[off + r];
                s += weight[off + r] * (d * d);
            }
            chi += 0.5 * s; /* :142 */
        }
    }
    if (chi_out) *chi_out = chi;
    free(fun); free(msphc); free(T); free(Ta); free(adj);
    return rc;
}

/* ------------------------------------------------------------------------
 * preproc::combine_kernels, isotropic part (src/preproc.cpp:196-342): the
 * per-period 2-D kernels adj_tt (here [nper][ny][nx], ix fastest) combined with
 * the depth sensitivities sen_* ([nx][ny][nz][nper], the reference's RowMajor
 * Tensor4r) into 3-D kernels ker_* ([nx][ny][nz], RowMajor Tensor3r), periods
 * in order (:239-301).  mode 0: vs only (use_alpha_beta_rho false, :280-301,
 * with dalpha_dbeta / drho_dalpha of include/utils.h:198-212); mode 1: vs, vp,
 * rho (:258-278, scale = 1 without anisotropy); mode 2: vs, vp and rho =
 * RHO_SCALING * vs (:276-278).  Outputs are zeroed first (:206-221).
 * ------------------------------------------------------------------------ */
static double o_drho_dalpha(double vp) {  /* include/utils.h:198-200 */
  return 1.6612 - 2*0.4721*vp + 3*0.0671*vp*vp - 4*0.0043*vp*vp*vp + 5*0.000106*vp*vp*vp*vp;
}
static double o_dalpha_dbeta(double vs) { /* include/utils.h:210-212 */
  return 2.0947 - 2*0.8206*vs + 3*0.2683*vs*vs - 4*0.0251*vs*vs*vs;
}

int sat_oracle_combine_kernels(int nx, int ny, int nz, int nper, int mode, const double* adj_tt,
                               const double* sen_vs, const double* sen_vp, const double* sen_rho,
                               const double* vs3d, const double* vp3d, double* ker_vs, double* ker_vp,
                               double* ker_rho) {
  if (nx < 1 || ny < 1 || nz < 1 || nper < 1 || mode < 0 || mode > 2) return SAT_ORACLE_EINVAL;
  const size_t n3 = (size_t)nx * ny * nz;
  for (size_t i = 0; i < n3; i++) {
    ker_vs[i] = 0.0;
    if (mode > 0) { ker_vp[i] = 0.0; ker_rho[i] = 0.0; }
  }
  for (int iper = 0; iper < nper; iper++) {
    for (int ix = 0; ix < nx; ix++) {
      for (int iy = 0; iy < ny; iy++) {
        const double att = adj_tt[(size_t)iper * nx * ny + ix + (size_t)iy * nx];
        for (int k = 0; k < nz; k++) {
          const size_t o3 = ((size_t)ix * ny + iy) * nz + k;
          const size_t o4 = o3 * nper + iper;
          if (mode == 0) {
            const double dab = o_dalpha_dbeta(vs3d[o3]);
            const double dra = o_drho_dalpha(vp3d[o3]);
            ker_vs[o3] -= att * (sen_vs[o4] + sen_vp[o4] * dab + sen_rho[o4] * dra * dab);
          } else {
            const double scale = 1.0 / sqrt(1.0 + 0.0 * 0.0 + 0.0 * 0.0);
            ker_vs[o3] -= att * scale * sen_vs[o4];
            ker_vp[o3] -= att * scale * sen_vp[o4];
            if (mode == 1) ker_rho[o3] -= att * scale * sen_rho[o4];
          }
        }
      }
    }
    if (mode == 2)  /* :276-278, after every period */
      for (size_t i = 0; i < n3; i++) ker_rho[i] = ker_vs[i] * 0.33;
  }
  return SAT_ORACLE_OK;
}

/* ------------------------------------------------------------------------
 * preproc::combine_kernels with the anisotropic and density branches
 * (src/preproc.cpp:196-342): as sat_oracle_combine_kernels, plus
 *  - is_anisotropy (r1 != NULL): scale = 1/sqrt(1 + r1^2 + r2^2) on the vs, vp,
 *    rho (and density) terms of the use_alpha_beta_rho branch (:253-256), and
 *    the Gc, Gs kernels (:307-338) from adj_tt, adj_xi, adj_eta, the phase
 *    velocity svel and sen_gc / sen_gs;
 *  - is_kden (adj_den != NULL): the density kernel, the vs term's expression
 *    with adj_den for adj_tt (:264-267, :294-300).
 * r1, r2: [nx][ny][nper] (SurfGrid r1_loc, r2_loc: RowMajor Tensor3r, from
 * prepare_aniso_media, src/surf_grid.cpp:362-392); adj_xi, adj_eta, adj_den,
 * svel: [nper][ny][nx]; sen_gc, sen_gs: [nx][ny][nz][nper]; kernels [nx][ny][nz].
 * The statements run per period in the reference's order: the isotropic loop
 * (:246-301), the rho scaling (:276-278), then the Gc/Gs loop (:309-338).
 * ------------------------------------------------------------------------ */
int sat_oracle_combine_kernels_ex(int nx, int ny, int nz, int nper, int mode, const double* adj_tt,
                                  const double* sen_vs, const double* sen_vp, const double* sen_rho,
                                  const double* vs3d, const double* vp3d, const double* r1p, const double* r2p,
                                  const double* adj_xi, const double* adj_eta, const double* svel,
                                  const double* sen_gc, const double* sen_gs, const double* adj_den,
                                  double* ker_vs, double* ker_vp, double* ker_rho, double* ker_gc,
                                  double* ker_gs, double* ker_den) {
  if (nx < 1 || ny < 1 || nz < 1 || nper < 1 || mode < 0 || mode > 2) return SAT_ORACLE_EINVAL;
  const int aniso = r1p != NULL, kden = adj_den != NULL;
  const size_t n2 = (size_t)nx * ny, n3 = n2 * nz;
  for (size_t i = 0; i < n3; i++) {
    ker_vs[i] = 0.0;
    if (mode > 0) { ker_vp[i] = 0.0; ker_rho[i] = 0.0; }
    if (aniso) { ker_gc[i] = 0.0; ker_gs[i] = 0.0; }
    if (kden) ker_den[i] = 0.0;
  }
  for (int iper = 0; iper < nper; iper++) {
    for (int ix = 0; ix < nx; ix++) {
      for (int iy = 0; iy < ny; iy++) {
        const size_t o2 = (size_t)iper * n2 + ix + (size_t)iy * nx;
        const double att = adj_tt[o2];
        const double aden = kden ? adj_den[o2] : 0.0;
        const size_t o_r = ((size_t)ix * ny + iy) * nper + iper;
        const double r1 = aniso ? r1p[o_r] : 0.0;
        const double r2 = aniso ? r2p[o_r] : 0.0;
        const double scale = 1.0 / sqrt(1.0 + r1 * r1 + r2 * r2);
        for (int k = 0; k < nz; k++) {
          const size_t o3 = ((size_t)ix * ny + iy) * nz + k;
          const size_t o4 = o3 * nper + iper;
          if (mode == 0) {
            const double dab = o_dalpha_dbeta(vs3d[o3]);
            const double dra = o_drho_dalpha(vp3d[o3]);
            ker_vs[o3] -= att * (sen_vs[o4] + sen_vp[o4] * dab + sen_rho[o4] * dra * dab);
            if (kden) ker_den[o3] -= aden * (sen_vs[o4] + sen_vp[o4] * dab + sen_rho[o4] * dra * dab);
          } else {
            ker_vs[o3] -= att * scale * sen_vs[o4];
            ker_vp[o3] -= att * scale * sen_vp[o4];
            if (kden) ker_den[o3] -= aden * scale * sen_vs[o4];
            if (mode == 1) ker_rho[o3] -= att * scale * sen_rho[o4];
          }
        }
      }
    }
    if (mode == 2)
      for (size_t i = 0; i < n3; i++) ker_rho[i] = ker_vs[i] * 0.33;
    if (aniso) {
      for (int ix = 0; ix < nx; ix++) {
        for (int iy = 0; iy < ny; iy++) {
          const size_t o2 = (size_t)iper * n2 + ix + (size_t)iy * nx;
          const size_t o_r = ((size_t)ix * ny + iy) * nper + iper;
          const double r1 = r1p[o_r], r2 = r2p[o_r];
          const double D = 1.0 + r1 * r1 + r2 * r2;
          const double sqrtD = sqrt(D);
          const double D15 = D * sqrtD;
          const double D2 = D * D;
          const double sv = svel[o2];
          const double att = adj_tt[o2], axi = adj_xi[o2], aeta = adj_eta[o2];
          for (int k = 0; k < nz; k++) {
            const size_t o3 = ((size_t)ix * ny + iy) * nz + k;
            const size_t o4 = o3 * nper + iper;
            ker_gc[o3] += (-att * r1 / D15 + axi * (1.0 - r1 * r1 + r2 * r2) / (sv * D2) +
                           aeta * (-2.0 * r1 * r2) / (sv * D2)) * sen_gc[o4];
            ker_gs[o3] += (-att * r2 / D15 + axi * (-2.0 * r1 * r2) / (sv * D2) +
                           aeta * (1.0 + r1 * r1 - r2 * r2) / (sv * D2)) * sen_gs[o4];
          }
        }
      }
    }
  }
  return SAT_ORACLE_OK;
}
