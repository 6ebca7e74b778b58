/* plbm_sim.h — the reference's sim/ plugin protocol, exported by libplbm.so.
 *
 * Replaces a plugin library of ivan-pi/periodic-lbm's side tree: the callback set of
 * sim/lbm.h:13-19 (siminit / simstep / simvars / simfree) under the names that
 * sim/cases.py:21-32 looks up for a library lib<name>.so, i.e. c_plbm_* here, plus the
 * norm helper of sim/sim_slbm.F90:197-205.  Semantics of the standard-LBM plugin
 * (sim/sim_slbm.F90:30-205); arrays are Fortran order with x fastest, (nx, ny) and
 * (nx, ny, 2).  Errors are reported on stderr; init returns NULL on failure.
 */
#ifndef PLBM_SIM_H
#define PLBM_SIM_H

#ifdef __cplusplus
extern "C" {
#endif

/* dt must be 1.  p: pressure (rho = 1 + 3 p), u: velocity; sigma must be NULL; params
 * NULL or a pointer to an int storage id (default PLBM_F64S). */
void *c_plbm_init(int nx, int ny, double dt, const double *p, const double *u,
                  const double *sigma, void *params);
void c_plbm_step(void *sim, double omega);
void c_plbm_vars(void *sim, double *rho, double *u);
void c_plbm_free(void *sim);
double c_plbm_norm(int nx, int ny, const double *u, const double *ua);

#ifdef __cplusplus
}
#endif
#endif
