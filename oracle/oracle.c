/* Test infrastructure only — the oracle.  See oracle.h for the contract.
 *
 * Built with gcc -O2 -ffp-contract=off (no FMA, no reassociation): every expression
 * below is parenthesised exactly as the Fortran source evaluates it (Fortran evaluates
 * a chain of equal-precedence operators left to right unless parentheses say otherwise,
 * and unary minus binds looser than '*').  That is what makes the FP64 results
 * bit-identical to the flang-built reference; tests/test_oracle_golden.py checks it bit
 * for bit against the fixtures that reference wrote (tests/golden/), and tests/test_oracle_invariants.py pins each function independently
 * (moment identities, permutation property, conservation, closed forms).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

const int ORC_CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
const int ORC_CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

/* src/fvm_bardow.F90:90-95 */
#define W0 (4.0 / 9.0)
#define WS (1.0 / 9.0)
#define WD (1.0 / 36.0)
const double ORC_W[9] = {W0, WS, WS, WS, WS, WD, WD, WD, WD};
static const double CSQR = 1.0 / 3.0;

#define F(q, x, y) f[((size_t)(q) * (size_t)nx + (size_t)(x)) * (size_t)ld + (size_t)(y)]

/* ------------------------------------------------------------------ lattice container */

/* src/fvm_bardow.F90:144-147: leading dimension rounded up to a multiple of 16 */
int orc_padded_ld(int ny) {
    int r = ny % 16;
    return r > 0 ? ny + (16 - r) : ny;
}

/* src/fvm_bardow.F90:99-126.  Pinned by the closed-form moments
 * sum feq = rho, sum c feq = rho u, sum cc feq = rho (uu + I/3) (test_oracle_invariants). */
void orc_equilibrium(double rho, double ux, double uy, double feq[9]) {
    double uxx = ux * ux, uyy = uy * uy;
    double indp = 1.0 - 1.5 * (uxx + uyy);
    feq[0] = (W0 * rho) * indp;
    feq[1] = (WS * rho) * ((indp + 3.0 * ux) + 4.5 * uxx);
    feq[2] = (WS * rho) * ((indp + 3.0 * uy) + 4.5 * uyy);
    feq[3] = (WS * rho) * ((indp - 3.0 * ux) + 4.5 * uxx);
    feq[4] = (WS * rho) * ((indp - 3.0 * uy) + 4.5 * uyy);
    double uxpy = ux + uy;
    feq[5] = (WD * rho) * ((indp + 3.0 * uxpy) + (4.5 * uxpy) * uxpy);
    feq[7] = (WD * rho) * ((indp - 3.0 * uxpy) + (4.5 * uxpy) * uxpy);
    double uxmy = ux - uy;
    feq[6] = (WD * rho) * ((indp - 3.0 * uxmy) + (4.5 * uxmy) * uxmy);
    feq[8] = (WD * rho) * ((indp + 3.0 * uxmy) + (4.5 * uxmy) * uxmy);
}

/* src/fvm_bardow.F90:242-269: tau = nu/csqr, omega = dt/(tau + dt/2), default
 * trt_magic = (tau/dt)**2.  Pinned by the closed form (test_oracle_invariants). */
void orc_set_properties(double nu, double dt, double magic, int has_magic, double out[6]) {
    double tau = (1.0 / CSQR) * nu;
    out[0] = nu;
    out[1] = dt;
    out[2] = tau;
    out[3] = dt / (tau + 0.5 * dt);
    out[4] = has_magic ? magic : (tau / dt) * (tau / dt);
    out[5] = CSQR;
}

/* src/collision_trt.F90:24-34.  Pinned by magic_number(le, lambda_d(le, L)) == L and by
 * L = 1/4 giving lambda_d = 2 - omega. */
double orc_magic_number(double le, double ld) {
    return ((2.0 - le) * (2.0 - ld)) / ((4.0 * le) * ld);
}
double orc_lambda_d(double omega, double x) {
    return (4.0 - 2.0 * omega) / ((((4.0 * x) * omega) + 2.0) - omega);
}

/* ------------------------------------------------------------------ streaming */

/* src/periodic_lbm.f90:45-127: pull streaming with periodic wrap,
 * fdst(y,x,q) = fsrc(y - cy_q, x - cx_q, q).  Pure copy, so bit-exact by construction;
 * pinned as a permutation (sum preserved exactly, push-stream inverts it). */
void orc_stream(int nx, int ny, int ld, const double *fsrc, double *fdst) {
    const double *f = fsrc;
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int q = 0; q < 9; ++q) {
            int xs = (x - ORC_CX[q] + nx) % nx;
            for (int y = 0; y < ny; ++y) {
                int ys = (y - ORC_CY[q] + ny) % ny;
                fdst[((size_t)q * nx + x) * ld + y] = F(q, xs, ys);
            }
        }
    }
}

/* ------------------------------------------------------------------ non-LBM streaming */

/* The reference's finite-volume / finite-difference streaming operators.  They loop
 * x = 1..ny, y = 1..nx (src/fvm_bardow.F90:434,437 and 724-726 ...), i.e. square grids;
 * here x runs over nx and y over ny, identical on square grids.  Rest population copied. */
#define IDX(q, x, y) (((size_t)(q) * (size_t)nx + (size_t)(x)) * (size_t)ld + (size_t)(y))

/* stream_fvm_bardow, src/fvm_bardow.F90:410-509 (bilinear face interpolation). */
void orc_stream_fvm_bardow(int nx, int ny, int ld, const double *fold, double *fnew, double dt) {
    const double p2 = 0.5, p8 = 0.125;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) fnew[IDX(0, x, y)] = fold[IDX(0, x, y)];
    for (int q = 1; q < 9; ++q) {
        const double cxq = dt * ORC_CX[q], cyq = dt * ORC_CY[q];
        #pragma omp parallel for schedule(static)
        for (int x = 0; x < nx; ++x) {
            const int xp = (x + 1) % nx, xm = (x + nx - 1) % nx;
            for (int y = 0; y < ny; ++y) {
                const int yp = (y + 1) % ny, ym = (y + ny - 1) % ny;
                const double fc = fold[IDX(q, x, y)], fe = fold[IDX(q, xp, y)],
                             fn = fold[IDX(q, x, yp)], fw = fold[IDX(q, xm, y)],
                             fs = fold[IDX(q, x, ym)], fne = fold[IDX(q, xp, yp)],
                             fnw = fold[IDX(q, xm, yp)], fsw = fold[IDX(q, xm, ym)],
                             fse = fold[IDX(q, xp, ym)];
                const double cfw = (p2 * (fc + fw) - (p2 * cxq) * (fc - fw)) -
                                   (p8 * cyq) * (((fnw + fn) - fsw) - fs);
                const double cfn = (p2 * (fc + fn) - (p2 * cyq) * (fn - fc)) -
                                   (p8 * cxq) * (((fne + fe) - fnw) - fw);
                const double cfe = (p2 * (fc + fe) - (p2 * cxq) * (fe - fc)) -
                                   (p8 * cyq) * (((fne + fn) - fse) - fs);
                const double cfs = (p2 * (fc + fs) - (p2 * cyq) * (fc - fs)) -
                                   (p8 * cxq) * (((fse + fe) - fsw) - fw);
                fnew[IDX(q, x, y)] = (fc - cxq * (cfe - cfw)) - cyq * (cfn - cfs);
            }
        }
    }
}

/* stream_fdm_bardow, src/fvm_bardow.F90:526-685: first and second differences of the
 * 9-point neighbourhood, Lax-Wendroff update.  variant selects the stencil the reference
 * builds with (src/fvm_bardow.F90:591-669): 0 default (central differences), 1 -DFDM_WLS,
 * 2 -DFDM_WLS_GAUSS_V1, 3 -DFDM_WLS_GAUSS_V2, 4 -DFDM_ISO.  Every sum is grouped left to
 * right, as the Fortran source evaluates it. */
void orc_stream_fdm_bardow_v(int variant, int nx, int ny, int ld, const double *fold,
                             double *fnew, double dt) {
    const double p2 = 0.5;
    const double one_sixth = 1.0 / 6.0, one_third = 1.0 / 3.0, two_thirds = 2.0 / 3.0;
    const double five_sixths = 10.0 / 12.0, one_twelth = 1.0 / 12.0;
    const int v1 = variant == 2;
    const double p1s = v1 ? 0.2880584423829145035434 : 0.3934930210807994210853,
                 p1d = v1 ? 0.1059707788085427065949 : 0.05325348945960039354075,
                 p2c = v1 ? -1.152233769531658458263 : -1.573972084323197018207,
                 p2d1 = v1 ? 0.5761168847658292291314 : 0.7869860421615988421706,
                 p2d2 = v1 ? -0.4238831152341712149578 : -0.2130139578384016019186,
                 p2d = v1 ? 0.2119415576170855242122 : 0.1065069789192007732037;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) fnew[IDX(0, x, y)] = fold[IDX(0, x, y)];
    for (int q = 1; q < 9; ++q) {
        const double cxq = dt * ORC_CX[q], cyq = dt * ORC_CY[q];
        const double cxxq = (0.5 * cxq) * cxq, cyyq = (0.5 * cyq) * cyq, cxyq = cxq * cyq;
        #pragma omp parallel for schedule(static)
        for (int x = 0; x < nx; ++x) {
            const int xp = (x + 1) % nx, xm = (x + nx - 1) % nx;
            for (int y = 0; y < ny; ++y) {
                const int yp = (y + 1) % ny, ym = (y + ny - 1) % ny;
                const double fc = fold[IDX(q, x, y)], fe = fold[IDX(q, xp, y)],
                             fn = fold[IDX(q, x, yp)], fw = fold[IDX(q, xm, y)],
                             fs = fold[IDX(q, x, ym)], fne = fold[IDX(q, xp, yp)],
                             fnw = fold[IDX(q, xm, yp)], fsw = fold[IDX(q, xm, ym)],
                             fse = fold[IDX(q, xp, ym)];
                double dfx, dfy, dfxx, dfyy, dfxy;
                switch (variant) {
                case 1: /* FDM_WLS */
                    dfx = one_sixth * (((fne - fnw) + (fe - fw)) + (fse - fsw));
                    dfy = one_sixth * (((fne - fse) + (fn - fs)) + (fnw - fsw));
                    dfxx = (one_third * ((fne - 2.0 * fn) + fnw) +
                            one_third * ((fe - 2.0 * fc) + fw)) +
                           one_third * ((fse - 2.0 * fs) + fsw);
                    dfyy = (one_third * ((fne - 2.0 * fe) + fse) +
                            one_third * ((fn - 2.0 * fc) + fs)) +
                           one_third * ((fnw - 2.0 * fw) + fsw);
                    dfxy = 0.25 * (((fne - fnw) + fsw) - fse);
                    break;
                case 2: /* FDM_WLS_GAUSS_V1 */
                case 3: /* FDM_WLS_GAUSS_V2 */
                    dfx = (p1s * (fe - fw) + p1d * (fne - fnw)) + p1d * (fse - fsw);
                    dfy = (p1s * (fn - fs) + p1d * (fne - fse)) + p1d * (fnw - fsw);
                    dfxx = ((p2c * fc + p2d1 * (fe + fw)) + p2d2 * (fn + fs)) +
                           p2d * (((fne + fnw) + fsw) + fse);
                    dfyy = ((p2c * fc + p2d2 * (fe + fw)) + p2d1 * (fn + fs)) +
                           p2d * (((fne + fnw) + fsw) + fse);
                    dfxy = 0.25 * (((fne - fnw) + fsw) - fse);
                    break;
                case 4: /* FDM_ISO */
                    dfx = p2 * ((one_sixth * (fne - fnw) + two_thirds * (fe - fw)) +
                                one_sixth * (fse - fsw));
                    dfy = p2 * ((one_sixth * (fne - fse) + two_thirds * (fn - fs)) +
                                one_sixth * (fnw - fsw));
                    dfxx = (one_twelth * ((fne - 2.0 * fn) + fnw) +
                            five_sixths * ((fe - 2.0 * fc) + fw)) +
                           one_twelth * ((fse - 2.0 * fs) + fsw);
                    dfyy = (one_twelth * ((fne - 2.0 * fe) + fse) +
                            five_sixths * ((fn - 2.0 * fc) + fs)) +
                           one_twelth * ((fnw - 2.0 * fw) + fsw);
                    dfxy = 0.25 * (((fne - fse) - fnw) + fsw);
                    break;
                default:
                    dfx = p2 * (fe - fw);
                    dfy = p2 * (fn - fs);
                    dfxx = (fe - 2.0 * fc) + fw;
                    dfyy = (fn - 2.0 * fc) + fs;
                    dfxy = 0.25 * (((fne - fse) - fnw) + fsw);
                }
                fnew[IDX(q, x, y)] =
                    ((fc - cxq * dfx) - cyq * dfy) + ((cxxq * dfxx + cxyq * dfxy) + cyyq * dfyy);
            }
        }
    }
}

void orc_stream_fdm_bardow(int nx, int ny, int ld, const double *fold, double *fnew, double dt) {
    orc_stream_fdm_bardow_v(0, nx, ny, ld, fold, fnew, dt);
}

/* stream_fdm_sofonea, src/fvm_bardow.F90:702-889: per direction, a 1-D Lax-Wendroff
 * update along c_q between the upstream node fu and the downstream node fd. */
void orc_stream_fdm_sofonea(int nx, int ny, int ld, const double *fold, double *fnew, double dt) {
    const double p2 = 0.5 / sqrt(2.0);
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) fnew[IDX(0, x, y)] = fold[IDX(0, x, y)];
    /* (dx, dy) of fu for q = 1..8; fd is the opposite neighbour */
    static const int UX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static const int UY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    for (int q = 1; q < 9; ++q) {
        #pragma omp parallel for schedule(static)
        for (int x = 0; x < nx; ++x) {
            const int xu = (x + UX[q] + nx) % nx, xd = (x - UX[q] + nx) % nx;
            for (int y = 0; y < ny; ++y) {
                const int yu = (y + UY[q] + ny) % ny, yd = (y - UY[q] + ny) % ny;
                const double fc = fold[IDX(q, x, y)], fu = fold[IDX(q, xu, yu)],
                             fd = fold[IDX(q, xd, yd)];
                double du1, du2;
                if (q < 5) {
                    du1 = 0.5 * (fu - fd);
                    du2 = (fu - 2.0 * fc) + fd;
                } else {
                    du1 = p2 * (fu - fd);
                    du2 = 0.5 * ((fu - 2.0 * fc) + fd);
                }
                fnew[IDX(q, x, y)] = fc + dt * ((0.5 * dt) * du2 - du1);
            }
        }
    }
}

/* ------------------------------------------------------------------ DUGKS
 * src/periodic_dugks.F90, default build (no -DDUGKS; src/CMakeLists.txt defines none).
 * Pinned bit for bit against the flang-built reference (ref_dugks_* in ref_driver.f90);
 * invariants: omega = 1 in the collision gives the SPLIT-form equilibrium, the stream
 * conserves mass and momentum on a periodic grid, dt = 0 leaves fp unchanged. */

/* kernel_bgk, src/periodic_dugks.F90:80-165 (identical to bgk_kernel_cache,
 * src/collision_bgk.F90:103-170): BGK in the split-weight algebra on one cell. */
static void dugks_bgk_cell(double f[9], double omega) {
    const double omegabar = 1.0 - omega;
    const double om_w0 = (3.0 * omega) * (4.0 / 9.0), om_ws = (3.0 * omega) * (1.0 / 9.0),
                 om_wd = (3.0 * omega) * (1.0 / 36.0);
    const double rho =
        (((f[5] + f[7]) + (f[6] + f[8])) + ((f[1] + f[3]) + (f[2] + f[4]))) + f[0];
    const double invrho = 1.0 / rho;
    const double ux = invrho * (((f[5] - f[7]) + (f[8] - f[6])) + (f[1] - f[3]));
    const double uy = invrho * (((f[5] - f[7]) + (f[6] - f[8])) + (f[2] - f[4]));
    const double indp = 1.0 / 3.0 - 0.5 * (ux * ux + uy * uy);
    f[0] = omegabar * f[0] + (om_w0 * rho) * indp;
    const double t13 = indp + (1.5 * ux) * ux;
    f[1] = omegabar * f[1] + (om_ws * rho) * (t13 + ux);
    f[3] = omegabar * f[3] + (om_ws * rho) * (t13 - ux);
    const double t24 = indp + (1.5 * uy) * uy;
    f[2] = omegabar * f[2] + (om_ws * rho) * (t24 + uy);
    f[4] = omegabar * f[4] + (om_ws * rho) * (t24 - uy);
    const double xpy = ux + uy, t57 = indp + (1.5 * xpy) * xpy;
    f[5] = omegabar * f[5] + (om_wd * rho) * (t57 + xpy);
    f[7] = omegabar * f[7] + (om_wd * rho) * (t57 - xpy);
    const double xmy = ux - uy, t68 = indp + (1.5 * xmy) * xmy;
    f[6] = omegabar * f[6] + (om_wd * rho) * (t68 - xmy);
    f[8] = omegabar * f[8] + (om_wd * rho) * (t68 + xmy);
}

/* bgk_kernel_cache, src/collision_bgk.F90:84-176 (collide_bgk of the -DSPLIT build): the
 * same split-weight algebra as DUGKS's kernel_bgk.  Pinned bit for bit against the
 * reference compiled with -DSPLIT (oracle/Makefile target ref-split). */
void orc_collide_bgk_split(int nx, int ny, int ld, double *f, double omega) {
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) {
            double c[9];
            for (int q = 0; q < 9; ++q) c[q] = f[IDX(q, x, y)];
            dugks_bgk_cell(c, omega);
            for (int q = 0; q < 9; ++q) f[IDX(q, x, y)] = c[q];
        }
}

/* dugks_collide, src/periodic_dugks.F90:46-77: fnew = BGK(fold, omega_full) (the copy
 * then kernel_bgk on inew with grid%omega), fold = BGK(fold, omega_half) in place (the
 * caller passes omega_half = 1/(tau/dt + 1/2)). */
void orc_dugks_collide(int nx, int ny, int ld, double *fold, double *fnew, double omega_full,
                       double omega_half) {
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double a[9], b[9];
            for (int q = 0; q < 9; ++q) a[q] = b[q] = fold[IDX(q, x, y)];
            dugks_bgk_cell(a, omega_full);
            dugks_bgk_cell(b, omega_half);
            for (int q = 0; q < 9; ++q) {
                fnew[IDX(q, x, y)] = a[q];
                fold[IDX(q, x, y)] = b[q];
            }
        }
    }
}

/* dugks_stream / kernel_stream, src/periodic_dugks.F90:168-299 (default build): face
 * values of ft by the bilinear interpolation of stream_fvm_bardow, then the flux balance
 * subtracted from fp for q = 1..8; fp(q=0) is left as it is. */
void orc_dugks_stream(int nx, int ny, int ld, const double *ft, double *fp, double dt) {
    const double p2 = 0.5, p8 = 0.125;
    for (int q = 1; q < 9; ++q) {
        const double cxq = dt * ORC_CX[q], cyq = dt * ORC_CY[q];
        #pragma omp parallel for schedule(static)
        for (int x = 0; x < nx; ++x) {
            const int xp = (x + 1) % nx, xm = (x + nx - 1) % nx;
            for (int y = 0; y < ny; ++y) {
                const int yp = (y + 1) % ny, ym = (y + ny - 1) % ny;
                const double fc = ft[IDX(q, x, y)], fe = ft[IDX(q, xp, y)],
                             fn = ft[IDX(q, x, yp)], fw = ft[IDX(q, xm, y)],
                             fs = ft[IDX(q, x, ym)], fne = ft[IDX(q, xp, yp)],
                             fnw = ft[IDX(q, xm, yp)], fsw = ft[IDX(q, xm, ym)],
                             fse = ft[IDX(q, xp, ym)];
                const double cfw = (p2 * (fc + fw) - (p2 * cxq) * (fc - fw)) -
                                   (p8 * cyq) * (((fnw + fn) - fsw) - fs);
                const double cfn = (p2 * (fc + fn) - (p2 * cyq) * (fn - fc)) -
                                   (p8 * cxq) * (((fne + fe) - fnw) - fw);
                const double cfe = (p2 * (fc + fe) - (p2 * cxq) * (fe - fc)) -
                                   (p8 * cyq) * (((fne + fn) - fse) - fs);
                const double cfs = (p2 * (fc + fs) - (p2 * cyq) * (fc - fs)) -
                                   (p8 * cxq) * (((fse + fe) - fsw) - fw);
                const size_t i = IDX(q, x, y);
                fp[i] = (fp[i] - cxq * (cfe - cfw)) - cyq * (cfn - cfs);
            }
        }
    }
}

/* update_ew / update_ns of the -DDUGKS build (src/periodic_dugks.F90:308-420): a BGK
 * relaxation of the interpolated face values of one face towards the face's own
 * equilibrium, with rate omega; the east/west form updates q = 1, 3, 5, 7, 6, 8 and the
 * north/south form q = 2, 4, 5, 7, 6, 8 (the others carry no flux across that face). */
static void orc_dugks_face_collide(double f[9], double omega, int ew) {
    const double one_third = 1.0 / 3.0, w0 = 4.0 / 9.0, ws = 1.0 / 9.0, wd = 1.0 / 36.0;
    const double omegabar = 1.0 - omega;
    const double omega_ws = (3.0 * omega) * ws, omega_wd = (3.0 * omega) * wd;
    (void)w0;
    const double rho = (((f[5] + f[7]) + (f[6] + f[8])) + ((f[1] + f[3]) + (f[2] + f[4]))) + f[0];
    const double invrho = 1.0 / rho;
    const double ux = invrho * (((f[5] - f[7]) + (f[8] - f[6])) + (f[1] - f[3]));
    const double uy = invrho * (((f[5] - f[7]) + (f[6] - f[8])) + (f[2] - f[4]));
    const double indp = one_third - 0.5 * (ux * ux + uy * uy);
    if (ew) {
        const double vel_trm_13 = indp + (1.5 * ux) * ux;
        f[1] = omegabar * f[1] + (omega_ws * rho) * (vel_trm_13 + ux);
        f[3] = omegabar * f[3] + (omega_ws * rho) * (vel_trm_13 - ux);
    } else {
        const double vel_trm_24 = indp + (1.5 * uy) * uy;
        f[2] = omegabar * f[2] + (omega_ws * rho) * (vel_trm_24 + uy);
        f[4] = omegabar * f[4] + (omega_ws * rho) * (vel_trm_24 - uy);
    }
    const double velxpy = ux + uy, vel_trm_57 = indp + (1.5 * velxpy) * velxpy;
    f[5] = omegabar * f[5] + (omega_wd * rho) * (vel_trm_57 + velxpy);
    f[7] = omegabar * f[7] + (omega_wd * rho) * (vel_trm_57 - velxpy);
    const double velxmy = ux - uy, vel_trm_68 = indp + (1.5 * velxmy) * velxmy;
    f[6] = omegabar * f[6] + (omega_wd * rho) * (vel_trm_68 - velxmy);
    f[8] = omegabar * f[8] + (omega_wd * rho) * (vel_trm_68 + velxmy);
}

/* kernel_stream of the -DDUGKS build (src/periodic_dugks.F90:190-299): the face values of
 * all nine directions (q = 0 included: the faces' moments need them), a face collision
 * on each of the four faces (update_ew on west and east, update_ns on north and south,
 * rate omega_face = 1/(4 tau/dt + 1)), then the flux balance for q = 1..8. */
void orc_dugks_stream_face(int nx, int ny, int ld, const double *ft, double *fp, double dt,
                           double omega_face) {
    const double p2 = 0.5, p8 = 0.125;
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        const int xp = (x + 1) % nx, xm = (x + nx - 1) % nx;
        for (int y = 0; y < ny; ++y) {
            const int yp = (y + 1) % ny, ym = (y + ny - 1) % ny;
            double cfw[9], cfn[9], cfe[9], cfs[9];
            for (int q = 0; q < 9; ++q) {
                const double cxq = dt * ORC_CX[q], cyq = dt * ORC_CY[q];
                const double fc = ft[IDX(q, x, y)], fe = ft[IDX(q, xp, y)],
                             fn = ft[IDX(q, x, yp)], fw = ft[IDX(q, xm, y)],
                             fs = ft[IDX(q, x, ym)], fne = ft[IDX(q, xp, yp)],
                             fnw = ft[IDX(q, xm, yp)], fsw = ft[IDX(q, xm, ym)],
                             fse = ft[IDX(q, xp, ym)];
                cfw[q] = (p2 * (fc + fw) - (p2 * cxq) * (fc - fw)) -
                         (p8 * cyq) * (((fnw + fn) - fsw) - fs);
                cfn[q] = (p2 * (fc + fn) - (p2 * cyq) * (fn - fc)) -
                         (p8 * cxq) * (((fne + fe) - fnw) - fw);
                cfe[q] = (p2 * (fc + fe) - (p2 * cxq) * (fe - fc)) -
                         (p8 * cyq) * (((fne + fn) - fse) - fs);
                cfs[q] = (p2 * (fc + fs) - (p2 * cyq) * (fc - fs)) -
                         (p8 * cxq) * (((fse + fe) - fsw) - fw);
            }
            orc_dugks_face_collide(cfw, omega_face, 1);
            orc_dugks_face_collide(cfe, omega_face, 1);
            orc_dugks_face_collide(cfn, omega_face, 0);
            orc_dugks_face_collide(cfs, omega_face, 0);
            for (int q = 1; q < 9; ++q) {
                const double cxq = dt * ORC_CX[q], cyq = dt * ORC_CY[q];
                const size_t i = IDX(q, x, y);
                fp[i] = (fp[i] - cxq * (cfe[q] - cfw[q])) - cyq * (cfn[q] - cfs[q]);
            }
        }
    }
}
#undef IDX

/* ------------------------------------------------------------------ collisions */

/* src/collision_bgk.F90:35-82 (bgk_kernel; the default build, no -DSPLIT).
 * Pinned: omega=1 gives f=feq(rho,u); mass and momentum conserved; rest is a fixed point. */
void orc_collide_bgk(int nx, int ny, int ld, double *f, double omega) {
    const double omegabar = 1.0 - omega;
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double fs[9], feq[9];
            for (int q = 0; q < 9; ++q) fs[q] = F(q, x, y);
            double rho = (((fs[5] + fs[7]) + (fs[6] + fs[8])) + ((fs[1] + fs[3]) + (fs[2] + fs[4])))
                         + fs[0];
            double invrho = 1.0 / rho;
            double ux = invrho * (((fs[5] - fs[7]) + (fs[8] - fs[6])) + (fs[1] - fs[3]));
            double uy = invrho * (((fs[5] - fs[7]) + (fs[6] - fs[8])) + (fs[2] - fs[4]));
            orc_equilibrium(rho, ux, uy, feq);
            for (int q = 0; q < 9; ++q) F(q, x, y) = omegabar * fs[q] + omega * feq[q];
        }
    }
}

/* src/collision_trt.F90:13-20 prefactors */
#define T0 (4.0 / 9.0)
#define T1X2 ((1.0 / 9.0) * 2.0)
#define T2X2 ((1.0 / 36.0) * 2.0)
#define INV2CSQ2 (1.0 / ((2.0 * (1.0 / 3.0)) * (1.0 / 3.0)))
#define FAC1 (T1X2 * INV2CSQ2)
#define FAC2 (T2X2 * INV2CSQ2)

/* src/collision_trt.F90:64-160 (trt_naive; incompressible equilibrium, j not divided by
 * rho).  Pinned: lambda_d == lambda_e reduces it to BGK with the incompressible feq;
 * mass and j conserved; lambda_e = 1 gives the incompressible feq directly. */
void orc_collide_trt(int nx, int ny, int ld, double *f, double lambda_e, double lambda_d) {
    const double les = 0.5 * lambda_e, lds = 0.5 * lambda_d;
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double vC = F(0, x, y), vE = F(1, x, y), vN = F(2, x, y), vW = F(3, x, y),
                   vS = F(4, x, y), vNE = F(5, x, y), vNW = F(6, x, y), vSW = F(7, x, y),
                   vSE = F(8, x, y);
            double rho = (((vNE + vSW) + (vNW + vSE)) + ((vE + vW) + (vN + vS))) + vC;
            double velX = (((vNE - vSW) + (vSE - vNW)) + (vE - vW));
            double velY = (((vNE - vSW) + (vNW - vSE)) + (vN - vS));
            double velX2 = velX * velX, velY2 = velY * velY;
            double feq_common = rho - 1.5 * (velX2 + velY2);
            F(0, x, y) = vC * (1.0 - lambda_e) + (lambda_e * T0) * feq_common;

            double velXPY = velX + velY;
            double sym = les * (((vNE + vSW) - (FAC2 * velXPY) * velXPY) - T2X2 * feq_common);
            double asym = lds * ((vNE - vSW) - (3.0 * T2X2) * velXPY);
            F(5, x, y) = (vNE - sym) - asym;
            F(7, x, y) = (vSW - sym) + asym;

            double velXMY = velX - velY;
            sym = les * (((vSE + vNW) - (FAC2 * velXMY) * velXMY) - T2X2 * feq_common);
            asym = lds * ((vSE - vNW) - (3.0 * T2X2) * velXMY);
            F(8, x, y) = (vSE - sym) - asym;
            F(6, x, y) = (vNW - sym) + asym;

            sym = les * (((vN + vS) - FAC1 * velY2) - T1X2 * feq_common);
            asym = lds * ((vN - vS) - (3.0 * T1X2) * velY);
            F(2, x, y) = (vN - sym) - asym;
            F(4, x, y) = (vS - sym) + asym;

            sym = les * (((vE + vW) - FAC1 * velX2) - T1X2 * feq_common);
            asym = lds * ((vE - vW) - (3.0 * T1X2) * velX);
            F(1, x, y) = (vE - sym) - asym;
            F(3, x, y) = (vW - sym) + asym;
        }
    }
}

/* src/collision_regularized.F90:40-202 (rr_kernel_naive).  Pinned: mass and momentum
 * conserved, rest state a fixed point, omega=1 returns the 4th-order Hermite feq,
 * which itself has the closed-form moments rho, rho u, rho(uu+I/3). */
void orc_collide_rr(int nx, int ny, int ld, double *f, double omega) {
    const double ow0 = W0 * (1.0 - omega), ows = WS * (1.0 - omega), owd = WD * (1.0 - omega);
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double vC = F(0, x, y), vE = F(1, x, y), vN = F(2, x, y), vW = F(3, x, y),
                   vS = F(4, x, y), vNE = F(5, x, y), vNW = F(6, x, y), vSW = F(7, x, y),
                   vSE = F(8, x, y);
            double feq[9];
            double rho = (((vNE + vSW) + (vNW + vSE)) + ((vE + vW) + (vN + vS))) + vC;
            double invrho = 1.0 / rho;
            double ux = invrho * (((vNE - vSW) + (vSE - vNW)) + (vE - vW));
            double uy = invrho * (((vNE - vSW) + (vNW - vSE)) + (vN - vS));
            double uxx = ux * ux, uyy = uy * uy;
            double uxxy = uxx * uy, uyyx = uyy * ux, uxxyy = uxx * uyy;
            double indp0 = 1.0 - 1.5 * (uxx + uyy);
            double indps = indp0 - 4.5 * uxxyy;
            double indpd = indp0 + 9.0 * uxxyy;
            indp0 = indp0 + 2.25 * uxxyy;

            feq[0] = (W0 * rho) * indp0;
            feq[1] = (WS * rho) * ((indps + 3.0 * ux) + 4.5 * (uxx - uyyx));
            feq[3] = (WS * rho) * ((indps - 3.0 * ux) + 4.5 * (uxx + uyyx));
            feq[2] = (WS * rho) * ((indps + 3.0 * uy) + 4.5 * (uyy - uxxy));
            feq[4] = (WS * rho) * ((indps - 3.0 * uy) + 4.5 * (uyy + uxxy));

            vC = vC - feq[0];
            vE = vE - feq[1];
            vN = vN - feq[2];
            vW = vW - feq[3];
            vS = vS - feq[4];

            double axx = CSQR * ((2.0 * (vE + vW) - (vN + vS)) - vC);
            double ayy = CSQR * ((2.0 * (vN + vS) - (vE + vW)) - vC);

            double u3p = uxxy + uyyx;
            double uxpy = ux + uy;
            double indp57 = indpd + (4.5 * uxpy) * uxpy;
            feq[5] = (WD * rho) * ((indp57 + 3.0 * uxpy) + 9.0 * u3p);
            feq[7] = (WD * rho) * ((indp57 - 3.0 * uxpy) - 9.0 * u3p);

            double u3m = uxxy - uyyx;
            double uxmy = ux - uy;
            double indp68 = indpd + (4.5 * uxmy) * uxmy;
            feq[6] = (WD * rho) * ((indp68 - 3.0 * uxmy) + 9.0 * u3m);
            feq[8] = (WD * rho) * ((indp68 + 3.0 * uxmy) - 9.0 * u3m);

            vNE = vNE - feq[5];
            vNW = vNW - feq[6];
            vSW = vSW - feq[7];
            vSE = vSE - feq[8];

            double tmp = (2.0 * CSQR) * (((vNE + vNW) + vSW) + vSE);
            axx = axx + tmp;
            ayy = ayy + tmp;
            double axy = ((vNE + vSW) - (vNW + vSE));

            double axxy = (2.0 * ux) * axy + uy * axx;
            double ayyx = (2.0 * uy) * axy + ux * ayy;
            double axxyy = (((2.0 * (ux * ayyx + uy * axxy)) - uxx * ayy) - uyy * axx)
                           - ((4.0 * ux) * uy) * axy;

            indp0 = -1.5 * (axx + ayy);
            indps = indp0 - 4.5 * axxyy;
            indpd = 9.0 * axxyy - 2.0 * indp0;
            indp0 = indp0 + 2.25 * axxyy;

            vC = indp0;
            vE = indps + 4.5 * (axx - ayyx);
            vW = indps + 4.5 * (axx + ayyx);
            vN = indps + 4.5 * (ayy - axxy);
            vS = indps + 4.5 * (ayy + axxy);
            vNE = indpd + 9.0 * ((axxy + ayyx) + axy);
            vSW = indpd - 9.0 * ((axxy + ayyx) - axy);
            vNW = indpd + 9.0 * ((axxy - ayyx) - axy);
            vSE = indpd - 9.0 * ((axxy - ayyx) + axy);

            F(0, x, y) = feq[0] + ow0 * vC;
            F(1, x, y) = feq[1] + ows * vE;
            F(2, x, y) = feq[2] + ows * vN;
            F(3, x, y) = feq[3] + ows * vW;
            F(4, x, y) = feq[4] + ows * vS;
            F(5, x, y) = feq[5] + owd * vNE;
            F(6, x, y) = feq[6] + owd * vNW;
            F(7, x, y) = feq[7] + owd * vSW;
            F(8, x, y) = feq[8] + owd * vSE;
        }
    }
}

void orc_collide(int model, int nx, int ny, int ld, double *f, double omega, double trt_magic) {
    switch (model) {
    case 0: orc_collide_bgk(nx, ny, ld, f, omega); break;
    case 1: orc_collide_trt(nx, ny, ld, f, omega, orc_lambda_d(omega, trt_magic)); break;
    case 2: orc_collide_rr(nx, ny, ld, f, omega); break;
    case 3: orc_collide_bgk_split(nx, ny, ld, f, omega); break;
    default: abort();
    }
}

/* ------------------------------------------------------------------ macros / init */

/* src/fvm_bardow.F90:356-388: its own sum order (f0 first), u = (1/rho) j. */
void orc_update_macros(int nx, int ny, int ld, const double *f, double *grho, double *gux,
                       double *guy) {
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double fs[9];
            for (int q = 0; q < 9; ++q) fs[q] = F(q, x, y);
            double rho = fs[0] + (((fs[5] + fs[7]) + (fs[6] + fs[8])) + ((fs[1] + fs[3]) + (fs[2] + fs[4])));
            double invrho = 1.0 / rho;
            size_t c = (size_t)x * ny + y;
            grho[c] = rho;
            gux[c] = invrho * (((fs[5] - fs[7]) + (fs[8] - fs[6])) + (fs[1] - fs[3]));
            guy[c] = invrho * (((fs[5] - fs[7]) + (fs[6] - fs[8])) + (fs[2] - fs[4]));
        }
    }
}

/* src/fvm_bardow.F90:272-305: f(:,:,:,iold) = equilibrium(rho, u) per cell. */
void orc_set_pdf_to_equilibrium(int nx, int ny, int ld, const double *rho, const double *ux,
                                const double *uy, double *f) {
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double feq[9];
            size_t c = (size_t)x * ny + y;
            orc_equilibrium(rho[c], ux[c], uy[c], feq);
            for (int q = 0; q < 9; ++q) F(q, x, y) = feq[q];
        }
    }
}

/* src/periodic_lbm.f90:15-29 repeated: stream iold -> inew, collide inew, swap.  On
 * return *latest_is_inew says whether the last post-collision state is in the buffer
 * passed as f_inew (1) or as f_iold (0).  With the reference's index swap, that buffer is
 * grid%f(:,:,:,grid%iold) after the final step, and the other one (grid%inew) holds the
 * state one step earlier — which is what update_macros reads. */
void orc_run(int model, int nx, int ny, int ld, double *f_iold, double *f_inew, int nsteps,
             double omega, double trt_magic, int *latest_is_inew) {
    double *a = f_iold, *b = f_inew;
    for (int s = 0; s < nsteps; ++s) {
        orc_stream(nx, ny, ld, a, b);
        orc_collide(model, nx, ny, ld, b, omega, trt_magic);
        double *t = a; a = b; b = t;
    }
    *latest_is_inew = (a == f_inew);
}

/* ------------------------------------------------------------------ shifted forms */

/* Populations stored as g_q = f_q - w_q*rho0, rho0 = 1 (precedent: sim/sim.F90:130-131,
 * 366-384).  Equilibrium in deviation form: geq_q = w_q*(drho + rho*P_q) where
 * P_q = feq_q/(w_q rho) - 1 is the velocity polynomial without its leading 1. */
void orc_equilibrium_shifted(double rho, double ux, double uy, double geq[9]) {
    double drho = rho - 1.0;
    double uxx = ux * ux, uyy = uy * uy;
    double p0 = -1.5 * (uxx + uyy);
    double uxpy = ux + uy, uxmy = ux - uy;
    double P[9];
    P[0] = p0;
    P[1] = p0 + 3.0 * ux + 4.5 * uxx;
    P[2] = p0 + 3.0 * uy + 4.5 * uyy;
    P[3] = p0 - 3.0 * ux + 4.5 * uxx;
    P[4] = p0 - 3.0 * uy + 4.5 * uyy;
    P[5] = p0 + 3.0 * uxpy + 4.5 * uxpy * uxpy;
    P[6] = p0 - 3.0 * uxmy + 4.5 * uxmy * uxmy;
    P[7] = p0 - 3.0 * uxpy + 4.5 * uxpy * uxpy;
    P[8] = p0 + 3.0 * uxmy + 4.5 * uxmy * uxmy;
    for (int q = 0; q < 9; ++q) geq[q] = ORC_W[q] * (drho + rho * P[q]);
}

static void rr_shifted_cell(double g[9], double omega) {
    /* Same algebra as rr_kernel_naive, with the constant w_q*rho0 removed from both the
     * populations and the equilibrium (it cancels in every non-equilibrium moment). */
    double drho = (((g[5] + g[7]) + (g[6] + g[8])) + ((g[1] + g[3]) + (g[2] + g[4]))) + g[0];
    double rho = 1.0 + drho;
    double ux = (((g[5] - g[7]) + (g[8] - g[6])) + (g[1] - g[3])) / rho;
    double uy = (((g[5] - g[7]) + (g[6] - g[8])) + (g[2] - g[4])) / rho;
    double uxx = ux * ux, uyy = uy * uy, uxxy = uxx * uy, uyyx = uyy * ux, uxxyy = uxx * uyy;
    double p0 = -1.5 * (uxx + uyy);
    double ps = p0 - 4.5 * uxxyy, pd = p0 + 9.0 * uxxyy;
    double pc = p0 + 2.25 * uxxyy;
    double uxpy = ux + uy, uxmy = ux - uy, u3p = uxxy + uyyx, u3m = uxxy - uyyx;
    double P[9];
    P[0] = pc;
    P[1] = ps + 3.0 * ux + 4.5 * (uxx - uyyx);
    P[3] = ps - 3.0 * ux + 4.5 * (uxx + uyyx);
    P[2] = ps + 3.0 * uy + 4.5 * (uyy - uxxy);
    P[4] = ps - 3.0 * uy + 4.5 * (uyy + uxxy);
    P[5] = pd + 4.5 * uxpy * uxpy + 3.0 * uxpy + 9.0 * u3p;
    P[7] = pd + 4.5 * uxpy * uxpy - 3.0 * uxpy - 9.0 * u3p;
    P[6] = pd + 4.5 * uxmy * uxmy - 3.0 * uxmy + 9.0 * u3m;
    P[8] = pd + 4.5 * uxmy * uxmy + 3.0 * uxmy - 9.0 * u3m;
    double geq[9], n[9];
    for (int q = 0; q < 9; ++q) {
        geq[q] = ORC_W[q] * (drho + rho * P[q]);
        n[q] = g[q] - geq[q];
    }
    double axx = CSQR * (2.0 * (n[1] + n[3]) - (n[2] + n[4]) - n[0]);
    double ayy = CSQR * (2.0 * (n[2] + n[4]) - (n[1] + n[3]) - n[0]);
    double tmp = 2.0 * CSQR * (n[5] + n[6] + n[7] + n[8]);
    axx += tmp;
    ayy += tmp;
    double axy = (n[5] + n[7]) - (n[6] + n[8]);
    double axxy = 2.0 * ux * axy + uy * axx;
    double ayyx = 2.0 * uy * axy + ux * ayy;
    double axxyy = 2.0 * (ux * ayyx + uy * axxy) - uxx * ayy - uyy * axx - 4.0 * ux * uy * axy;
    double i0 = -1.5 * (axx + ayy);
    double is = i0 - 4.5 * axxyy, id = 9.0 * axxyy - 2.0 * i0;
    i0 = i0 + 2.25 * axxyy;
    double h[9];
    h[0] = i0;
    h[1] = is + 4.5 * (axx - ayyx);
    h[3] = is + 4.5 * (axx + ayyx);
    h[2] = is + 4.5 * (ayy - axxy);
    h[4] = is + 4.5 * (ayy + axxy);
    h[5] = id + 9.0 * (axxy + ayyx + axy);
    h[7] = id - 9.0 * (axxy + ayyx - axy);
    h[6] = id + 9.0 * (axxy - ayyx - axy);
    h[8] = id - 9.0 * (axxy - ayyx + axy);
    for (int q = 0; q < 9; ++q) g[q] = geq[q] + ORC_W[q] * (1.0 - omega) * h[q];
}

static void trt_shifted_cell(double g[9], double le, double lo) {
    /* trt_naive with rho = 1 + drho: the w_q*rho0 parts cancel in every symmetric pair
     * (t*rho0 == 2 w rho0) and in the rest population. */
    double drho = (((g[5] + g[7]) + (g[6] + g[8])) + ((g[1] + g[3]) + (g[2] + g[4]))) + g[0];
    double jx = ((g[5] - g[7]) + (g[8] - g[6])) + (g[1] - g[3]);
    double jy = ((g[5] - g[7]) + (g[6] - g[8])) + (g[2] - g[4]);
    double jx2 = jx * jx, jy2 = jy * jy;
    double dcommon = drho - 1.5 * (jx2 + jy2);
    double les = 0.5 * le, los = 0.5 * lo;
    double o[9];
    o[0] = g[0] * (1.0 - le) + le * T0 * dcommon;
    double jp = jx + jy, jm = jx - jy, s, a;
    s = les * (g[5] + g[7] - FAC2 * jp * jp - T2X2 * dcommon);
    a = los * (g[5] - g[7] - 3.0 * T2X2 * jp);
    o[5] = g[5] - s - a;
    o[7] = g[7] - s + a;
    s = les * (g[8] + g[6] - FAC2 * jm * jm - T2X2 * dcommon);
    a = los * (g[8] - g[6] - 3.0 * T2X2 * jm);
    o[8] = g[8] - s - a;
    o[6] = g[6] - s + a;
    s = les * (g[2] + g[4] - FAC1 * jy2 - T1X2 * dcommon);
    a = los * (g[2] - g[4] - 3.0 * T1X2 * jy);
    o[2] = g[2] - s - a;
    o[4] = g[4] - s + a;
    s = les * (g[1] + g[3] - FAC1 * jx2 - T1X2 * dcommon);
    a = los * (g[1] - g[3] - 3.0 * T1X2 * jx);
    o[1] = g[1] - s - a;
    o[3] = g[3] - s + a;
    for (int q = 0; q < 9; ++q) g[q] = o[q];
}

void orc_collide_shifted(int model, int nx, int ny, int ld, double *f, double omega,
                         double lambda_odd) {
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double g[9];
            for (int q = 0; q < 9; ++q) g[q] = F(q, x, y);
            if (model == 0) {
                double drho = (((g[5] + g[7]) + (g[6] + g[8])) + ((g[1] + g[3]) + (g[2] + g[4]))) + g[0];
                double rho = 1.0 + drho;
                double ux = (((g[5] - g[7]) + (g[8] - g[6])) + (g[1] - g[3])) / rho;
                double uy = (((g[5] - g[7]) + (g[6] - g[8])) + (g[2] - g[4])) / rho;
                double geq[9];
                orc_equilibrium_shifted(rho, ux, uy, geq);
                for (int q = 0; q < 9; ++q) g[q] = (1.0 - omega) * g[q] + omega * geq[q];
            } else if (model == 1) {
                trt_shifted_cell(g, omega, lambda_odd);
            } else {
                rr_shifted_cell(g, omega);
            }
            for (int q = 0; q < 9; ++q) F(q, x, y) = g[q];
        }
    }
}

void orc_update_macros_shifted(int nx, int ny, int ld, const double *f, double *grho,
                               double *gux, double *guy) {
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        for (int y = 0; y < ny; ++y) {
            double g[9];
            for (int q = 0; q < 9; ++q) g[q] = F(q, x, y);
            double rho = 1.0 + ((((g[5] + g[7]) + (g[6] + g[8])) + ((g[1] + g[3]) + (g[2] + g[4]))) + g[0]);
            size_t c = (size_t)x * ny + y;
            grho[c] = rho;
            gux[c] = (((g[5] - g[7]) + (g[8] - g[6])) + (g[1] - g[3])) / rho;
            guy[c] = (((g[5] - g[7]) + (g[6] - g[8])) + (g[2] - g[4])) / rho;
        }
    }
}

/* ------------------------------------------------------------------ flow cases */

/* src/benchmarks/taylor_green.f90:31-84 (the constructor's td and eval).  Cell-centred
 * coordinates (x-1)+1/2.  The reference loops y = 1..nx (:68), i.e. square grids; this
 * restatement loops y over ny, identical for the square grids it is used on. */
void orc_tgv_eval(int nx, int ny, double kx, double ky, double umax, double nu, double t,
                  double *p, double *ux, double *uy) {
    double td = 1.0 / (nu * (kx * kx + ky * ky));
    for (int x = 0; x < nx; ++x) {
        double xx = (double)x + 0.5;
        for (int y = 0; y < ny; ++y) {
            double yy = (double)y + 0.5;
            size_t c = (size_t)x * ny + y;
            ux[c] = -((((umax * sqrt(ky / kx)) * cos(kx * xx)) * sin(ky * yy)) * exp(-(t / td)));
            uy[c] = (((umax * sqrt(kx / ky)) * sin(kx * xx)) * cos(ky * yy)) * exp(-(t / td));
            p[c] = -(((0.25 * (umax * umax)) *
                      ((ky / kx) * cos((2.0 * kx) * xx) + (kx / ky) * cos((2.0 * ky) * yy))) *
                     exp(-((2.0 * t) / td)));
        }
    }
}

/* src/benchmarks/shear_layer.f90:25-52 (eval_shear_layer_grid). */
void orc_shear_eval(int nx, int ny, double length, double u0, double k, double delta,
                    double *rho, double *ux, double *uy) {
    const double pi = 4.0 * atan(1.0);
    for (int x = 0; x < nx; ++x) {
        double xd = ((double)x + 0.5) / length;
        for (int y = 0; y < ny; ++y) {
            double yd = ((double)y + 0.5) / length;
            size_t c = (size_t)x * ny + y;
            uy[c] = (u0 * delta) * sin((2.0 * pi) * (xd + 0.25));
            if (yd <= 0.5)
                ux[c] = u0 * tanh(k * (yd - 0.25));
            else
                ux[c] = u0 * tanh(k * (0.75 - yd));
            rho[c] = 1.0;
        }
    }
}

/* Counter-based uniform [0,1) deviate: the splitmix64 output function applied to
 * seed + (counter+1)*golden-gamma, top 53 bits.  (Not in the reference, whose configs
 * name no generator; the product implements the same function, and the tests compare
 * both against a numpy restatement.) */
double orc_noise(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* app/main_taylor_green.f90:174-212: norm2(hypot(u-ua, v-va)) / norm2(hypot(ua, va)),
 * evaluated as plain sums of squares (agrees with flang's norm2 to roundoff, not bits). */
double orc_l2_norm(int n, const double *ux, const double *uy, const double *uxa,
                   const double *uya) {
    double above = 0.0, below = 0.0;
    for (int i = 0; i < n; ++i) {
        double dx = ux[i] - uxa[i], dy = uy[i] - uya[i];
        above += dx * dx + dy * dy;
        below += uxa[i] * uxa[i] + uya[i] * uya[i];
    }
    return sqrt(above) / sqrt(below);
}
