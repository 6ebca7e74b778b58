// d2q9.hpp — per-cell D2Q9 arithmetic shared by every kernel of the engine.
//
// Velocity ordering, weights and the three collision models follow the reference
// (ivan-pi/periodic-lbm): q = 0 rest, 1 E, 2 N, 3 W, 4 S, 5 NE, 6 NW, 7 SW, 8 SE
// (src/fvm_bardow.F90:87-92).
//
// Two families live here:
//  * *_ref: unshifted FP64, every expression parenthesised as the Fortran evaluates it
//    (left-to-right chains, unary minus below '*'), with FMA contraction switched off.
//    Given IEEE double +,-,*,/ (correctly rounded division on gfx950), a kernel built on
//    them produces the reference's numbers bit for bit.
//  * *_shifted<R>: populations stored as deviations g_q = f_q - w_q*rho0 (rho0 = 1),
//    arithmetic in R (double or float).  The constant w_q*rho0 is removed exactly from
//    populations and equilibria, so the O(1) part never rounds; this is what lets FP32
//    and FP16 storage hold the O(u) signal (precedent: sim/sim.F90:130-131,366-384).
#pragma once

#include <hip/hip_runtime.h>

namespace plbm {

// Compile-time D2Q9 tables (src/fvm_bardow.F90:87-92)
__host__ __device__ constexpr int cx(int q) {
    return q == 1 || q == 5 || q == 8 ? 1 : (q == 3 || q == 6 || q == 7 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int q) {
    return q == 2 || q == 5 || q == 6 ? 1 : (q == 4 || q == 7 || q == 8 ? -1 : 0);
}
constexpr double W0 = 4.0 / 9.0, WS = 1.0 / 9.0, WD = 1.0 / 36.0;
__host__ __device__ constexpr double wq(int q) { return q == 0 ? W0 : (q < 5 ? WS : WD); }
constexpr double CSQR = 1.0 / 3.0;

// TRT prefactors, src/collision_trt.F90:13-20
constexpr double T0 = 4.0 / 9.0;
constexpr double T1X2 = (1.0 / 9.0) * 2.0;
constexpr double T2X2 = (1.0 / 36.0) * 2.0;
constexpr double INV2CSQ2 = 1.0 / ((2.0 * (1.0 / 3.0)) * (1.0 / 3.0));
constexpr double FAC1 = T1X2 * INV2CSQ2;
constexpr double FAC2 = T2X2 * INV2CSQ2;

// ======================================================================= reference FP64

// equilibrium(rho, ux, uy), src/fvm_bardow.F90:99-126
__device__ inline void equilibrium_ref(double rho, double ux, double uy, double feq[9]) {
#pragma clang fp contract(off)
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

// bgk_kernel body, src/collision_bgk.F90:55-74
__device__ inline void collide_bgk_ref(double f[9], double omega) {
#pragma clang fp contract(off)
    double rho = (((f[5] + f[7]) + (f[6] + f[8])) + ((f[1] + f[3]) + (f[2] + f[4]))) + f[0];
    double invrho = 1.0 / rho;
    double ux = invrho * (((f[5] - f[7]) + (f[8] - f[6])) + (f[1] - f[3]));
    double uy = invrho * (((f[5] - f[7]) + (f[6] - f[8])) + (f[2] - f[4]));
    double feq[9];
    equilibrium_ref(rho, ux, uy, feq);
    const double omegabar = 1.0 - omega;
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = omegabar * f[q] + omega * feq[q];
}

// bgk_kernel_cache body, src/collision_bgk.F90:103-170 (the -DSPLIT build of collide_bgk):
// the same BGK model in split-weight algebra, f_q <- (1-omega) f_q + 3 omega w_q rho
// (1/3 - u^2/2 + 1.5 (c_q.u)^2 + c_q.u).  kernel_bgk of src/periodic_dugks.F90:80-165 is
// the same code.
__device__ inline void collide_bgk_split_ref(double f[9], double omega) {
#pragma clang fp contract(off)
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

// trt_naive body, src/collision_trt.F90:97-152 (incompressible: j is not divided by rho)
__device__ inline void collide_trt_ref(double f[9], double lambda_e, double lambda_d) {
#pragma clang fp contract(off)
    const double les = 0.5 * lambda_e, lds = 0.5 * lambda_d;
    double vC = f[0], vE = f[1], vN = f[2], vW = f[3], vS = f[4], vNE = f[5], vNW = f[6],
           vSW = f[7], vSE = f[8];
    double rho = (((vNE + vSW) + (vNW + vSE)) + ((vE + vW) + (vN + vS))) + vC;
    double velX = (((vNE - vSW) + (vSE - vNW)) + (vE - vW));
    double velY = (((vNE - vSW) + (vNW - vSE)) + (vN - vS));
    double velX2 = velX * velX, velY2 = velY * velY;
    double feq_common = rho - 1.5 * (velX2 + velY2);
    f[0] = vC * (1.0 - lambda_e) + (lambda_e * T0) * feq_common;

    double velXPY = velX + velY;
    double sym = les * (((vNE + vSW) - (FAC2 * velXPY) * velXPY) - T2X2 * feq_common);
    double asym = lds * ((vNE - vSW) - (3.0 * T2X2) * velXPY);
    f[5] = (vNE - sym) - asym;
    f[7] = (vSW - sym) + asym;

    double velXMY = velX - velY;
    sym = les * (((vSE + vNW) - (FAC2 * velXMY) * velXMY) - T2X2 * feq_common);
    asym = lds * ((vSE - vNW) - (3.0 * T2X2) * velXMY);
    f[8] = (vSE - sym) - asym;
    f[6] = (vNW - sym) + asym;

    sym = les * (((vN + vS) - FAC1 * velY2) - T1X2 * feq_common);
    asym = lds * ((vN - vS) - (3.0 * T1X2) * velY);
    f[2] = (vN - sym) - asym;
    f[4] = (vS - sym) + asym;

    sym = les * (((vE + vW) - FAC1 * velX2) - T1X2 * feq_common);
    asym = lds * ((vE - vW) - (3.0 * T1X2) * velX);
    f[1] = (vE - sym) - asym;
    f[3] = (vW - sym) + asym;
}

// rr_kernel_naive body, src/collision_regularized.F90:71-194
__device__ inline void collide_rr_ref(double f[9], double omega) {
#pragma clang fp contract(off)
    const double ow0 = W0 * (1.0 - omega), ows = WS * (1.0 - omega), owd = WD * (1.0 - omega);
    double vC = f[0], vE = f[1], vN = f[2], vW = f[3], vS = f[4], vNE = f[5], vNW = f[6],
           vSW = f[7], vSE = f[8];
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
    double axxyy =
        (((2.0 * (ux * ayyx + uy * axxy)) - uxx * ayy) - uyy * axx) - ((4.0 * ux) * uy) * axy;

    indp0 = -1.5 * (axx + ayy);
    indps = indp0 - 4.5 * axxyy;
    indpd = 9.0 * axxyy - 2.0 * indp0;
    indp0 = indp0 + 2.25 * axxyy;

    f[0] = feq[0] + ow0 * indp0;
    f[1] = feq[1] + ows * (indps + 4.5 * (axx - ayyx));
    f[3] = feq[3] + ows * (indps + 4.5 * (axx + ayyx));
    f[2] = feq[2] + ows * (indps + 4.5 * (ayy - axxy));
    f[4] = feq[4] + ows * (indps + 4.5 * (ayy + axxy));
    f[5] = feq[5] + owd * (indpd + 9.0 * ((axxy + ayyx) + axy));
    f[7] = feq[7] + owd * (indpd - 9.0 * ((axxy + ayyx) - axy));
    f[6] = feq[6] + owd * (indpd + 9.0 * ((axxy - ayyx) - axy));
    f[8] = feq[8] + owd * (indpd - 9.0 * ((axxy - ayyx) + axy));
}

// update_macros_kernel body, src/fvm_bardow.F90:368-379 (rest population summed first)
__device__ inline void macros_ref(const double f[9], double &rho, double &ux, double &uy) {
#pragma clang fp contract(off)
    rho = f[0] + (((f[5] + f[7]) + (f[6] + f[8])) + ((f[1] + f[3]) + (f[2] + f[4])));
    double invrho = 1.0 / rho;
    ux = invrho * (((f[5] - f[7]) + (f[8] - f[6])) + (f[1] - f[3]));
    uy = invrho * (((f[5] - f[7]) + (f[6] - f[8])) + (f[2] - f[4]));
}

// ======================================================================= shifted forms

// Velocity polynomial of the 2nd-order equilibrium without its leading 1:
// feq_q = w_q*rho*(1 + P_q)  =>  geq_q = feq_q - w_q*rho0 = w_q*(drho + rho*P_q).
template <class R>
__device__ inline void equilibrium_shifted(R drho, R ux, R uy, R geq[9]) {
#pragma clang fp contract(fast)
    const R rho = R(1) + drho;
    const R uxx = ux * ux, uyy = uy * uy;
    const R p0 = R(-1.5) * (uxx + uyy);
    const R uxpy = ux + uy, uxmy = ux - uy;
    const R a = p0 + R(4.5) * uxx, b = p0 + R(4.5) * uyy;
    const R c = p0 + R(4.5) * uxpy * uxpy, d = p0 + R(4.5) * uxmy * uxmy;
    geq[0] = R(W0) * (drho + rho * p0);
    geq[1] = R(WS) * (drho + rho * (a + R(3) * ux));
    geq[3] = R(WS) * (drho + rho * (a - R(3) * ux));
    geq[2] = R(WS) * (drho + rho * (b + R(3) * uy));
    geq[4] = R(WS) * (drho + rho * (b - R(3) * uy));
    geq[5] = R(WD) * (drho + rho * (c + R(3) * uxpy));
    geq[7] = R(WD) * (drho + rho * (c - R(3) * uxpy));
    geq[6] = R(WD) * (drho + rho * (d - R(3) * uxmy));
    geq[8] = R(WD) * (drho + rho * (d + R(3) * uxmy));
}

template <class R>
__device__ inline void moments_shifted(const R g[9], R &drho, R &jx, R &jy) {
#pragma clang fp contract(fast)
    drho = (((g[5] + g[7]) + (g[6] + g[8])) + ((g[1] + g[3]) + (g[2] + g[4]))) + g[0];
    jx = ((g[5] - g[7]) + (g[8] - g[6])) + (g[1] - g[3]);
    jy = ((g[5] - g[7]) + (g[6] - g[8])) + (g[2] - g[4]);
}

// 1/x for the shifted forms: in FP32 the hardware reciprocal (v_rcp_f32, 1 ulp) instead of
// the correctly rounded division sequence (about ten instructions per cell in an
// instruction-bound pass); the FP32 forms are held to error bounds against FP64, not to
// bit equality.  FP64 divides exactly.
template <class R>
__device__ __forceinline__ R recip(R x) {
    if constexpr (sizeof(R) == 4)
        return __builtin_amdgcn_rcpf(x);
    else
        return R(1) / x;
}

template <class R>
__device__ inline void collide_bgk_shifted(R g[9], R omega) {
#pragma clang fp contract(fast)
    R drho, jx, jy;
    moments_shifted(g, drho, jx, jy);
    const R inv = R(1) / (R(1) + drho);
    R geq[9];
    equilibrium_shifted(drho, jx * inv, jy * inv, geq);
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = g[q] + omega * (geq[q] - g[q]);
}

// trt_naive in deviation form: in every symmetric pair t*rho0 = 2*w*rho0 cancels the
// shift of f_q + f_qbar, and the rest population loses t0*rho0 on both sides.
template <class R>
__device__ inline void collide_trt_shifted(R g[9], R le, R lo) {
#pragma clang fp contract(fast)
    R drho, jx, jy;
    moments_shifted(g, drho, jx, jy);
    const R jx2 = jx * jx, jy2 = jy * jy;
    const R dc = drho - R(1.5) * (jx2 + jy2);
    const R les = R(0.5) * le, los = R(0.5) * lo;
    R o0 = g[0] + le * (R(T0) * dc - g[0]);
    const R jp = jx + jy, jm = jx - jy;
    R s, a;
    s = les * (g[5] + g[7] - R(FAC2) * jp * jp - R(T2X2) * dc);
    a = los * (g[5] - g[7] - R(3.0 * T2X2) * jp);
    const R o5 = g[5] - s - a, o7 = g[7] - s + a;
    s = les * (g[8] + g[6] - R(FAC2) * jm * jm - R(T2X2) * dc);
    a = los * (g[8] - g[6] - R(3.0 * T2X2) * jm);
    const R o8 = g[8] - s - a, o6 = g[6] - s + a;
    s = les * (g[2] + g[4] - R(FAC1) * jy2 - R(T1X2) * dc);
    a = los * (g[2] - g[4] - R(3.0 * T1X2) * jy);
    const R o2 = g[2] - s - a, o4 = g[4] - s + a;
    s = les * (g[1] + g[3] - R(FAC1) * jx2 - R(T1X2) * dc);
    a = los * (g[1] - g[3] - R(3.0 * T1X2) * jx);
    const R o1 = g[1] - s - a, o3 = g[3] - s + a;
    g[0] = o0; g[1] = o1; g[2] = o2; g[3] = o3; g[4] = o4;
    g[5] = o5; g[6] = o6; g[7] = o7; g[8] = o8;
}

// rr_kernel_naive in deviation form: the 4th-order Hermite equilibrium minus w_q*rho0,
// the non-equilibrium moments from g - geq (identical to f - feq), and the recursive
// 3rd/4th-order reconstruction unchanged.
template <class R>
__device__ inline void collide_rr_shifted(R g[9], R omega) {
#pragma clang fp contract(fast)
    R drho, jx, jy;
    moments_shifted(g, drho, jx, jy);
    const R rho = R(1) + drho;
    const R inv = recip(rho);
    const R ux = jx * inv, uy = jy * inv;
    const R uxx = ux * ux, uyy = uy * uy, uxxy = uxx * uy, uyyx = uyy * ux, uxxyy = uxx * uyy;
    const R p0 = R(-1.5) * (uxx + uyy);
    const R ps = p0 - R(4.5) * uxxyy, pd = p0 + R(9) * uxxyy, pc = p0 + R(2.25) * uxxyy;
    const R uxpy = ux + uy, uxmy = ux - uy, u3p = uxxy + uyyx, u3m = uxxy - uyyx;
    R geq[9];
    geq[0] = R(W0) * (drho + rho * pc);
    geq[1] = R(WS) * (drho + rho * (ps + R(3) * ux + R(4.5) * (uxx - uyyx)));
    geq[3] = R(WS) * (drho + rho * (ps - R(3) * ux + R(4.5) * (uxx + uyyx)));
    geq[2] = R(WS) * (drho + rho * (ps + R(3) * uy + R(4.5) * (uyy - uxxy)));
    geq[4] = R(WS) * (drho + rho * (ps - R(3) * uy + R(4.5) * (uyy + uxxy)));
    const R e57 = pd + R(4.5) * uxpy * uxpy, e68 = pd + R(4.5) * uxmy * uxmy;
    geq[5] = R(WD) * (drho + rho * (e57 + R(3) * uxpy + R(9) * u3p));
    geq[7] = R(WD) * (drho + rho * (e57 - R(3) * uxpy - R(9) * u3p));
    geq[6] = R(WD) * (drho + rho * (e68 - R(3) * uxmy + R(9) * u3m));
    geq[8] = R(WD) * (drho + rho * (e68 + R(3) * uxmy - R(9) * u3m));
    R n[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) n[q] = g[q] - geq[q];
    const R tmp = R(2.0 * CSQR) * (n[5] + n[6] + n[7] + n[8]);
    const R axx = R(CSQR) * (R(2) * (n[1] + n[3]) - (n[2] + n[4]) - n[0]) + tmp;
    const R ayy = R(CSQR) * (R(2) * (n[2] + n[4]) - (n[1] + n[3]) - n[0]) + tmp;
    const R axy = (n[5] + n[7]) - (n[6] + n[8]);
    const R axxy = R(2) * ux * axy + uy * axx;
    const R ayyx = R(2) * uy * axy + ux * ayy;
    const R axxyy = R(2) * (ux * ayyx + uy * axxy) - uxx * ayy - uyy * axx - R(4) * ux * uy * axy;
    R i0 = R(-1.5) * (axx + ayy);
    const R is = i0 - R(4.5) * axxyy, id = R(9) * axxyy - R(2) * i0;
    i0 = i0 + R(2.25) * axxyy;
    const R om = R(1) - omega;
    g[0] = geq[0] + R(W0) * om * i0;
    g[1] = geq[1] + R(WS) * om * (is + R(4.5) * (axx - ayyx));
    g[3] = geq[3] + R(WS) * om * (is + R(4.5) * (axx + ayyx));
    g[2] = geq[2] + R(WS) * om * (is + R(4.5) * (ayy - axxy));
    g[4] = geq[4] + R(WS) * om * (is + R(4.5) * (ayy + axxy));
    g[5] = geq[5] + R(WD) * om * (id + R(9) * (axxy + ayyx + axy));
    g[7] = geq[7] + R(WD) * om * (id - R(9) * (axxy + ayyx - axy));
    g[6] = geq[6] + R(WD) * om * (id + R(9) * (axxy - ayyx - axy));
    g[8] = geq[8] + R(WD) * om * (id - R(9) * (axxy - ayyx + axy));
}

// FP32 forms of BGK and TRT that conserve mass and momentum by construction.  In FP32 the
// weights w_q and the TRT prefactors are not representable, so the operators above change
// rho and j by a fixed relative amount (~1e-8) per cell and step, a bias that accumulates
// over thousands of steps.  Both operators conserve rho and j exactly in exact arithmetic,
// so the populations that close the three balances can be computed from the others
// without changing the operator: BGK takes geq_1, geq_2 from the momentum balances and
// geq_0 from the mass balance of the equilibrium; TRT takes the antisymmetric parts of
// pairs (1,3), (2,4) from those of (5,7), (8,6) and the rest population from the
// symmetric parts.  What remains is the unbiased rounding of each update.
template <class R>
__device__ inline void collide_bgk_shifted_conserving(R g[9], R omega) {
#pragma clang fp contract(fast)
    R drho, jx, jy;
    moments_shifted(g, drho, jx, jy);
    const R inv = recip(R(1) + drho);
    R geq[9];
    equilibrium_shifted(drho, jx * inv, jy * inv, geq);
    const R e57 = geq[5] - geq[7];
    geq[1] = geq[3] + (jx - (e57 + (geq[8] - geq[6])));
    geq[2] = geq[4] + (jy - (e57 + (geq[6] - geq[8])));
    geq[0] = drho - (((geq[5] + geq[7]) + (geq[6] + geq[8])) +
                     ((geq[1] + geq[3]) + (geq[2] + geq[4])));
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = g[q] + omega * (geq[q] - g[q]);
}

template <class R>
__device__ inline void collide_trt_shifted_conserving(R g[9], R le, R lo) {
#pragma clang fp contract(fast)
    R drho, jx, jy;
    moments_shifted(g, drho, jx, jy);
    const R jx2 = jx * jx, jy2 = jy * jy;
    const R dc = drho - R(1.5) * (jx2 + jy2);
    const R les = R(0.5) * le, los = R(0.5) * lo;
    const R jp = jx + jy, jm = jx - jy;
    const R s57 = les * (g[5] + g[7] - R(FAC2) * jp * jp - R(T2X2) * dc);
    const R a57 = los * (g[5] - g[7] - R(3.0 * T2X2) * jp);
    const R s86 = les * (g[8] + g[6] - R(FAC2) * jm * jm - R(T2X2) * dc);
    const R a86 = los * (g[8] - g[6] - R(3.0 * T2X2) * jm);
    const R s24 = les * (g[2] + g[4] - R(FAC1) * jy2 - R(T1X2) * dc);
    const R s13 = les * (g[1] + g[3] - R(FAC1) * jx2 - R(T1X2) * dc);
    // j_x changes by -2 (a13 + a57 + a86), j_y by -2 (a24 + a57 - a86), the mass by
    // o_0 - g_0 - 2 (s57 + s86 + s24 + s13)
    const R a13 = -(a57 + a86), a24 = a86 - a57;
    const R o0 = g[0] + R(2) * ((s57 + s86) + (s24 + s13));
    const R o5 = g[5] - s57 - a57, o7 = g[7] - s57 + a57;
    const R o8 = g[8] - s86 - a86, o6 = g[6] - s86 + a86;
    const R o2 = g[2] - s24 - a24, o4 = g[4] - s24 + a24;
    const R o1 = g[1] - s13 - a13, o3 = g[3] - s13 + a13;
    g[0] = o0; g[1] = o1; g[2] = o2; g[3] = o3; g[4] = o4;
    g[5] = o5; g[6] = o6; g[7] = o7; g[8] = o8;
}

// Conservation closure for the RR operator in FP32 (the same idea as above): RR conserves
// rho and j exactly in exact arithmetic, so after the update g_1, g_2 are taken from the
// momentum balances and g_0 from the mass balance against the pre-collision moments
// (d0, jx0, jy0); the operator's own formulas for those three populations are then dead
// code.
template <class R>
__device__ inline void close_balances(R g[9], R d0, R jx0, R jy0) {
#pragma clang fp contract(fast)
    const R e57 = g[5] - g[7];
    g[1] = g[3] + (jx0 - (e57 + (g[8] - g[6])));
    g[2] = g[4] + (jy0 - (e57 + (g[6] - g[8])));
    g[0] = d0 - (((g[5] + g[7]) + (g[6] + g[8])) + ((g[1] + g[3]) + (g[2] + g[4])));
}

}  // namespace plbm
