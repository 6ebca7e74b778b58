/* Test infrastructure only — the oracle.
 *
 * A CPU restatement, in plain C, of the reference D2Q9 hot path of ivan-pi/periodic-lbm
 * (Fortran 2008 + OpenMP), written to reproduce the reference's IEEE operation order so
 * that its FP64 results are bit-identical to the flang-built reference.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load liboracle.so, and
 * only as the checker; the product (libplbm.so) never links or calls it.
 *
 * Array conventions follow the reference exactly (Fortran column-major):
 *   f(ld, nx, 0:8)  ->  f[(q*nx + x)*ld + y]      x in [0,nx), y in [0,ny), ld >= ny
 *   rho/ux/uy(ny,nx) -> a[x*ny + y]
 *
 * How each function is pinned (beyond its citation) is stated next to it in oracle.c.
 */
#ifndef PLBM_ORACLE_H
#define PLBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* D2Q9 constants, src/fvm_bardow.F90:87-95 */
extern const int ORC_CX[9];
extern const int ORC_CY[9];
extern const double ORC_W[9];

void orc_equilibrium(double rho, double ux, double uy, double feq[9]);
void orc_set_properties(double nu, double dt, double magic, int has_magic,
                        double out[6] /* nu, dt, tau, omega, trt_magic, csqr */);
double orc_magic_number(double le, double ld);
double orc_lambda_d(double omega, double x);
int orc_padded_ld(int ny);

void orc_stream(int nx, int ny, int ld, const double *fsrc, double *fdst);
/* Finite-volume / finite-difference streaming operators (src/fvm_bardow.F90:393-893) */
void orc_stream_fvm_bardow(int nx, int ny, int ld, const double *fold, double *fnew, double dt);
void orc_stream_fdm_bardow(int nx, int ny, int ld, const double *fold, double *fnew, double dt);
/* variant: 0 default, 1 FDM_WLS, 2 FDM_WLS_GAUSS_V1, 3 FDM_WLS_GAUSS_V2, 4 FDM_ISO */
void orc_stream_fdm_bardow_v(int variant, int nx, int ny, int ld, const double *fold,
                             double *fnew, double dt);
void orc_stream_fdm_sofonea(int nx, int ny, int ld, const double *fold, double *fnew, double dt);
/* DUGKS, src/periodic_dugks.F90 (default build) */
void orc_dugks_collide(int nx, int ny, int ld, double *fold, double *fnew, double omega_full,
                       double omega_half);
void orc_dugks_stream(int nx, int ny, int ld, const double *ft, double *fp, double dt);
/* the -DDUGKS build's stream: face collisions with rate omega_face before the fluxes */
void orc_dugks_stream_face(int nx, int ny, int ld, const double *ft, double *fp, double dt,
                           double omega_face);
void orc_collide_bgk(int nx, int ny, int ld, double *f, double omega);
/* collide_bgk of the -DSPLIT build (bgk_kernel_cache, src/collision_bgk.F90:84-176) */
void orc_collide_bgk_split(int nx, int ny, int ld, double *f, double omega);
void orc_collide_trt(int nx, int ny, int ld, double *f, double lambda_e, double lambda_d);
void orc_collide_rr(int nx, int ny, int ld, double *f, double omega);
/* model: 0 BGK, 1 TRT, 2 RR, 3 BGK (-DSPLIT algebra).  trt_magic feeds lambda_d() as collide_trt does. */
void orc_collide(int model, int nx, int ny, int ld, double *f, double omega, double trt_magic);
void orc_update_macros(int nx, int ny, int ld, const double *f, double *rho, double *ux,
                       double *uy);
void orc_set_pdf_to_equilibrium(int nx, int ny, int ld, const double *rho, const double *ux,
                                const double *uy, double *f);
/* perform_lbm_step x nsteps on the pair (fa = iold, fb = inew); returns through *swapped
 * whether the latest post-collision state sits in fb (0) or fa (1) — i.e. after the call
 * iold points at fa if *swapped==1 ... see oracle.c. */
void orc_run(int model, int nx, int ny, int ld, double *f_iold, double *f_inew, int nsteps,
             double omega, double trt_magic, int *latest_is_inew);

/* Shifted-storage forms: g_q = f_q - w_q*rho0 (rho0 = 1).  Pinned by equality with the
 * unshifted forms to roundoff in FP64. */
void orc_equilibrium_shifted(double rho, double ux, double uy, double geq[9]);
void orc_collide_shifted(int model, int nx, int ny, int ld, double *g, double omega,
                         double lambda_odd);
void orc_update_macros_shifted(int nx, int ny, int ld, const double *g, double *rho,
                               double *ux, double *uy);

/* Flow cases */
void orc_tgv_eval(int nx, int ny, double kx, double ky, double umax, double nu, double t,
                  double *p, double *ux, double *uy);
void orc_shear_eval(int nx, int ny, double length, double u0, double k, double delta,
                    double *rho, double *ux, double *uy);
double orc_noise(uint64_t seed, uint64_t counter);

/* App-side norms, app/main_taylor_green.f90:174-212 (norm2 of hypot) */
double orc_l2_norm(int n, const double *ux, const double *uy, const double *uxa,
                   const double *uya);

#ifdef __cplusplus
}
#endif
#endif
