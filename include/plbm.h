/* plbm.h — C ABI of the MI355X (gfx950) D2Q9 periodic lattice Boltzmann engine.
 *
 * The hot path of ivan-pi/periodic-lbm is the Fortran module API
 *   fvm_bardow / periodic_lbm / collision_bgk / collision_trt / collision_regularized
 * (lattice container, equilibrium, stream + collide step, macroscopic moments).  This
 * library is what that API binds to: the Fortran modules under periodic-lbm_amd/fortran/
 * call these functions through ISO_C_BINDING, and the Python harness calls them through
 * ctypes.  Each entry point names the reference routine it replaces (file:line).
 *
 * Conventions
 *  - Every function returns an int status: PLBM_OK (0) or a PLBM_ERR_* code; the text of
 *    the last error on the calling thread is returned by plbm_last_error().
 *  - Host population arrays are the reference's f(ld, nx, 0:8) of one buffer, Fortran
 *    order, unshifted FP64, with the caller's leading dimension ld >= ny:
 *    element (y, x, q) (0-based) at f[(q*nx + x)*ld + y].
 *  - Host macroscopic arrays are (ny, nx) Fortran order: element (y, x) at a[x*ny + y].
 *  - Buffer ids are 0-based (the Fortran shim maps grid%iold/inew/imid 1-based ids).
 *  - Entry points with the _dev suffix take device pointers (hipMalloc'd, or a torch
 *    tensor's data_ptr) with the same layouts; they are asynchronous on the lattice's
 *    stream.  Host-pointer entry points are synchronous.
 *  - A lattice is not thread-safe; calls on one lattice must not overlap.
 */
#ifndef PLBM_H
#define PLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLBM_OK 0
#define PLBM_ERR_ARG 1      /* invalid argument (size, id, null pointer, model) */
#define PLBM_ERR_HIP 2      /* a HIP runtime call failed (message has hipGetErrorString) */
#define PLBM_ERR_STATE 3    /* call not valid in the current state (e.g. no params set) */
#define PLBM_ERR_NOMEM 4    /* device or host allocation failed */
#define PLBM_ERR_NUMERIC 5  /* non-finite populations detected by a checked call */
#define PLBM_ERR_TIMEOUT 6  /* a bounded device-side wait gave up (batched engine) */
#define PLBM_ERR_HOST 7     /* a host-side failure other than allocation (C++ exception) */

/* Collision models: collide_bgk (src/collision_bgk.F90:17), collide_trt
 * (src/collision_trt.F90:41), collide_rr (src/collision_regularized.F90:21). */
#define PLBM_BGK 0
#define PLBM_TRT 1
#define PLBM_RR 2
/* collide_bgk as the reference builds it with -DSPLIT (bgk_kernel_cache,
 * src/collision_bgk.F90:84-176): the same BGK model in split-weight algebra, bit-identical
 * to that build in F64 storage (it differs from PLBM_BGK at roundoff level); the shifted
 * storages run the shifted BGK for both. */
#define PLBM_BGK_SPLIT 3

/* Population storage.  F64 reproduces the reference arithmetic bit for bit.  The
 * shifted forms store g_q = f_q - w_q*rho0 (rho0 = 1): F64S in FP64 with FP64 math,
 * F32S in FP32 and F16S in FP16 (scaled by a power of two) with FP32 math. */
#define PLBM_F64 0
#define PLBM_F64S 1
#define PLBM_F32S 2
#define PLBM_F16S 3

/* Reduction slots written by plbm_reduce / plbm_reduce_dev (all FP64, deterministic). */
#define PLBM_RED_MASS 0   /* sum rho                                  */
#define PLBM_RED_KE 1     /* sum rho*(ux^2+uy^2)/2                     */
#define PLBM_RED_ERR2 2   /* sum (ux-uxa)^2 + (uy-uya)^2 (0 if no ua)  */
#define PLBM_RED_REF2 3   /* sum uxa^2 + uya^2 (0 if no ua)            */
#define PLBM_RED_UMAX 4   /* max hypot(ux, uy)                         */
#define PLBM_RED_NONFINITE 5 /* number of cells with a non-finite rho/u */
#define PLBM_RED_U2 6     /* sum ux^2 + uy^2                           */
#define PLBM_NRED 8

typedef struct plbm_lattice plbm_lattice;

typedef struct plbm_info {
    int nx, ny;        /* grid */
    int ldd;           /* device leading dimension (elements), >= ny */
    int nf;            /* number of population buffers */
    int storage;       /* PLBM_F64 ... PLBM_F16S */
    int device;        /* HIP device ordinal */
    double omega;      /* collision parameters last set */
    double trt_magic;
    double lambda_odd; /* lambda_d(omega, trt_magic), src/collision_trt.F90:30-34 */
    double f16_scale;  /* storage scale for PLBM_F16S (power of two), 1 otherwise */
    size_t bytes_per_buffer;
} plbm_info;

/* ---------------------------------------------------------------- lifecycle */

/* alloc_grid (src/fvm_bardow.F90:129-202): device storage for nf population buffers of
 * an nx x ny periodic grid, on the current HIP device, with its own stream. */
int plbm_create(int nx, int ny, int nf, int storage, plbm_lattice **out);
/* dealloc_grid (src/fvm_bardow.F90:204-240). */
int plbm_destroy(plbm_lattice *h);
int plbm_get_info(const plbm_lattice *h, plbm_info *out);
/* Run this lattice's work on `stream` (a hipStream_t; NULL = the lattice's own). */
int plbm_set_stream(plbm_lattice *h, void *stream);
/* Device pointer of population buffer `buf`.  Device layout (column-interleaved SoA):
 * element (q, x, y) at index (x*9 + q)*ldd + y, in the storage type. */
int plbm_buffer_ptr(plbm_lattice *h, int buf, void **dptr);

/* set_properties (src/fvm_bardow.F90:242-269) computes omega and trt_magic on the host
 * side; this hands them to the device path.  lambda_odd = lambda_d(omega, trt_magic). */
int plbm_set_params(plbm_lattice *h, double omega, double trt_magic);
/* Override the FP16 storage scale (power of two, default 1024); set before loading. */
int plbm_set_f16_scale(plbm_lattice *h, double scale);

/* ---------------------------------------------------------------- populations */

/* Copy one buffer from/to the host array f(ld, nx, 0:8) (unshifted FP64). */
int plbm_upload_f(plbm_lattice *h, int buf, const double *f, int ld);
int plbm_download_f(plbm_lattice *h, int buf, double *f, int ld);

/* set_pdf_to_equilibrium (src/fvm_bardow.F90:272-305) with equilibrium()
 * (src/fvm_bardow.F90:99-126): buffer buf = feq(rho, ux, uy) per cell. */
int plbm_equilibrium_init(plbm_lattice *h, int buf, const double *rho, const double *ux,
                          const double *uy);
int plbm_equilibrium_init_dev(plbm_lattice *h, int buf, const double *rho, const double *ux,
                              const double *uy);

/* ---------------------------------------------------------------- stepping */

/* One fused pull-stream + collide step, src -> dst: lbm_stream
 * (src/periodic_lbm.f90:32-129) followed by the collision on the streamed buffer
 * (src/collision_*.F90), in one pass over HBM. */
int plbm_step(plbm_lattice *h, int model, int src, int dst);
/* perform_lbm_step (src/periodic_lbm.f90:15-29) nsteps times with the reference's
 * buffer bookkeeping: on entry *iold / *inew are the current ids; each step reads *iold,
 * writes *inew, then swaps them.  Runs in passes of several steps where that is faster
 * (plbm_set_steps_per_launch, plbm_set_pass_lag_buffer); the slot contents afterwards are
 * the reference's, but the slots' device storage may have moved (fetch plbm_buffer_ptr
 * again).  Asynchronous. */
int plbm_run(plbm_lattice *h, int model, int nsteps, int *iold, int *inew);
/* Unfused operators for the procedure-pointer slots grid%streaming / grid%collision:
 * lbm_stream (src/periodic_lbm.f90:32-129) src -> dst; collide_* in place on buf. */
int plbm_stream(plbm_lattice *h, int src, int dst);
int plbm_collide(plbm_lattice *h, int model, int buf);
/* Buffer copy dst = src (perform_triple_step's pre-collision store,
 * src/fvm_bardow.F90:328). */
int plbm_copy(plbm_lattice *h, int src, int dst);

/* Streaming operators for the grid%streaming slot, src -> dst with time step dt:
 * PLBM_STREAM_EXACT = lbm_stream (dt unused); the others are the reference's
 * finite-volume / finite-difference schemes stream_fvm_bardow (src/fvm_bardow.F90:393-509),
 * stream_fdm_bardow (:511-685, default stencil) and stream_fdm_sofonea (:688-893). */
#define PLBM_STREAM_EXACT 0
#define PLBM_STREAM_FVM_BARDOW 1
#define PLBM_STREAM_FDM_BARDOW 2
#define PLBM_STREAM_FDM_SOFONEA 3
/* stream_fdm_bardow with the stencils the reference selects at compile time
 * (src/fvm_bardow.F90:591-669): -DFDM_WLS, -DFDM_WLS_GAUSS_V1, -DFDM_WLS_GAUSS_V2,
 * -DFDM_ISO.  Bit-identical in FP64 to the reference built with that option. */
#define PLBM_STREAM_FDM_BARDOW_WLS 4
#define PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1 5
#define PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2 6
#define PLBM_STREAM_FDM_BARDOW_ISO 7
int plbm_stream_op(plbm_lattice *h, int op, int src, int dst, double dt);
/* perform_step (src/fvm_bardow.F90:307-320) with grid%streaming = streaming operator `op`
 * and grid%collision = collide_<model>, fused into one pass: src -> stream op -> collide ->
 * dst.  In FP64 storage bit-identical to plbm_stream_op followed by plbm_collide (the
 * unfused path stores the intermediate exactly); in FP32/FP16 storage it skips that
 * intermediate rounding.  op = PLBM_STREAM_EXACT is plbm_step.  Asynchronous. */
int plbm_step_op(plbm_lattice *h, int op, int model, int src, int dst, double dt);
/* nsteps x perform_step with streaming operator `op` and collide_<model> (the fused pair),
 * with the reference's buffer bookkeeping as plbm_run (each step reads *iold, writes
 * *inew, swaps them).  In FP64 storage the steps run in temporally blocked passes of up
 * to plbm_set_op_steps_per_pass steps, the last pass also writing the state one step
 * back into the lattice's third buffer as plbm_run does (bit-identical to single steps;
 * the slots may exchange their device storage, as in plbm_run).  Asynchronous. */
int plbm_run_op(plbm_lattice *h, int op, int model, int nsteps, int *iold, int *inew,
                double dt);
/* Pass depth of plbm_run_op: 1 = single fused steps, 2..3 = passes (FP64 storage only),
 * 0 = chosen from the grid size, operator and model.  The get variant reports the
 * effective value for (op, model). */
int plbm_set_op_steps_per_pass(plbm_lattice *h, int steps);
int plbm_get_op_steps_per_pass(const plbm_lattice *h, int op, int model, int *steps);

/* DUGKS, src/periodic_dugks.F90 (default build, without -DDUGKS; the -DDUGKS form is
 * plbm_dugks_stream_faces / plbm_dugks_step_faces below), FP64 storage only
 * (PLBM_ERR_STATE otherwise).  omega_full is grid%omega, omega_half the reference's
 * 1/(tau/dt + 1/2) (:56).
 *   plbm_dugks_collide  dugks_collide (:46-77): dst = BGK(src, omega_full),
 *                       src = BGK(src, omega_half) in place.
 *   plbm_dugks_stream   dugks_stream / kernel_stream (:168-299): dst_q -= dt c_q . (face
 *                       flux balance of src), q = 1..8.
 *   plbm_dugks_step     nsteps x perform_dugks_step (:25-38) on slots (iold, inew) as they
 *                       are on entry.  On return every slot holds what the reference's
 *                       buffer of the same index holds after the same steps (the caller
 *                       swaps its iold/inew ids once per step, as the reference does).
 *                       The last step moves 360 bytes per cell; the steps before it run in
 *                       temporally blocked passes (plbm_dugks_set_steps_per_pass).  The
 *                       two slots may exchange their device storage, so addresses from
 *                       plbm_buffer_ptr for them must be fetched again.  Asynchronous. */
int plbm_dugks_collide(plbm_lattice *h, int src, int dst, double omega_full,
                       double omega_half);
int plbm_dugks_stream(plbm_lattice *h, int src, int dst, double dt);
int plbm_dugks_step(plbm_lattice *h, int iold, int inew, int nsteps, double dt,
                    double omega_full, double omega_half);
/* Steps per pass of plbm_dugks_step: all steps but the last run in temporally blocked
 * passes of up to this many steps over device memory (1 = the collide + flux pair every
 * step; 2..3 = passes; 0 = chosen from the grid size).  Bit-identical for every value. */
int plbm_dugks_set_steps_per_pass(plbm_lattice *h, int steps);
/* DUGKS as the reference builds it with -DDUGKS (src/periodic_dugks.F90:68,199,267,308):
 * the stream relaxes the interpolated face values of every face towards the face's own
 * equilibrium (update_ew / update_ns) with rate omega_face = 1/(4 tau/dt + 1) before the
 * flux balance, and dugks_collide's half collision runs at 0.75/(tau/dt + 1/2), which the
 * caller passes as omega_half to plbm_dugks_collide / plbm_dugks_step_faces.
 *   plbm_dugks_stream_faces  dugks_stream of that build, src -> dst (q = 1..8).
 *   plbm_dugks_step_faces    nsteps x perform_dugks_step of that build, with the slot
 *                            rules of plbm_dugks_step (one collide + flux pair per step).
 * FP64 storage only.  Asynchronous. */
int plbm_dugks_stream_faces(plbm_lattice *h, int src, int dst, double dt, double omega_face);
int plbm_dugks_step_faces(plbm_lattice *h, int iold, int inew, int nsteps, double dt,
                          double omega_full, double omega_half, double omega_face);
int plbm_dugks_get_steps_per_pass(const plbm_lattice *h, int *steps);

/* ---------------------------------------------------------------- read-back */

/* update_macros (src/fvm_bardow.F90:343-390) on buffer buf: rho and u = j/rho, in the
 * reference's summation order; (ny, nx) Fortran-order outputs. */
int plbm_macros(plbm_lattice *h, int buf, double *rho, double *ux, double *uy);
int plbm_macros_dev(plbm_lattice *h, int buf, double *rho, double *ux, double *uy);

/* FP64 reductions of buffer buf (slots PLBM_RED_*): mass, kinetic energy, max|u|,
 * non-finite count and, when uxa/uya are non-NULL ((ny,nx) analytic velocity), the
 * sums behind the app's relative L2 norm (app/main_taylor_green.f90:174-212). */
int plbm_reduce(plbm_lattice *h, int buf, const double *uxa, const double *uya,
                double out[PLBM_NRED]);
int plbm_reduce_dev(plbm_lattice *h, int buf, const double *uxa, const double *uya,
                    double out[PLBM_NRED]);
/* plbm_reduce_dev without any allocation (no counterpart in the reference; for callers
 * that cannot allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and the same result bits as
 * plbm_reduce_dev, but out_dev is a DEVICE array of PLBM_NRED doubles, written
 * asynchronously on the lattice's stream (no host synchronisation), and the partial sums
 * live in a static device array of 1024 x PLBM_NRED doubles (64 KiB per GPU, held until
 * the process ends).  Grids of any size fit it (at most 1024 reduction workgroups; a
 * larger count is refused with PLBM_ERR_ARG).  All calls on one GPU share that array:
 * calls that overlap on one GPU (other streams, other threads) race. */
int plbm_reduce_static(plbm_lattice *h, int buf, const double *uxa, const double *uya,
                       double *out_dev);

int plbm_sync(plbm_lattice *h);
const char *plbm_last_error(void);

/* ---------------------------------------------------------------- flow cases (host) */

/* taylor_green_t%eval (src/benchmarks/taylor_green.f90:31-84) on an (ny, nx) grid. */
int plbm_field_taylor_green(int nx, int ny, double kx, double ky, double umax, double nu,
                            double t, double *p, double *ux, double *uy);
/* shear_layer_t's grid evaluation eval_shear_layer_grid
 * (src/benchmarks/shear_layer.f90:25-52). */
int plbm_field_shear_layer(int nx, int ny, double length, double u0, double k, double delta,
                           double *rho, double *ux, double *uy);
/* vortex_case_t%eval (src/benchmarks/barotropic_vortex_case.F90:34-83): convected
 * barotropic vortex (advection u0, centre (xc, yc), radius rc, strength eps), rho0 = 1,
 * csqr = 1/3. */
int plbm_field_barotropic_vortex(int nx, int ny, double u0, double xc, double yc, double rc,
                                 double eps, double rho0, double csqr, double *rho, double *ux,
                                 double *uy);
/* ux += amp*(2U-1), uy += amp*(2U'-1) with U, U' the counter-based deviates of cell c
 * (counters 2c and 2c+1) of `seed`; the generator is splitmix64's output function. */
int plbm_field_add_noise(int n, uint64_t seed, double amp, double *ux, double *uy);
double plbm_noise(uint64_t seed, uint64_t counter);

/* lambda_d / magic_number (src/collision_trt.F90:24-34). */
double plbm_lambda_d(double omega, double x);
double plbm_magic_number(double le, double ld);

/* ---------------------------------------------------------------- launch shape */

/* Launch shape of the fused step for one model: threads per workgroup (multiple of 64 in
 * [64, 1024]) and XCD-contiguous workgroup renumbering (0/1).  Default 256 / 1. */
int plbm_set_launch(plbm_lattice *h, int model, int block, int xcd_remap);
int plbm_get_launch(const plbm_lattice *h, int model, int *block, int *xcd_remap);
/* Temporal blocking of plbm_run: steps advanced per pass over HBM (1 = the single-step
 * kernel; 2..8 = that many fused steps per launch, each wavefront walking a segment of
 * columns with the intermediate states in registers; 0 = chosen from the grid size: 8
 * when the default work split gives at least 512 wavefronts, about 512^2 and larger,
 * else 1).  In F64 storage the results are bit-identical for every setting;
 * the shifted storages agree to roundoff (their collisions allow FMA contraction, which
 * the compiler places per kernel).  plbm_run keeps the reference's slot contents (latest
 * in *iold, the state one step earlier in *inew); with passes of more than one step it
 * may exchange the device storage of the two slots, so addresses from plbm_buffer_ptr
 * must be fetched again after it.  The get variant reports the effective value. */
int plbm_set_steps_per_launch(plbm_lattice *h, int model, int steps);
/* One pass of `steps` (1..8) fused steps from buffer src to
 * buffer dst, without slot bookkeeping and without writing the intermediate states:
 * the building block of plbm_run.  Asynchronous. */
int plbm_pass(plbm_lattice *h, int model, int steps, int src, int dst);
int plbm_get_steps_per_launch(const plbm_lattice *h, int model, int *steps);
/* Work split of a pass (engine extension, no counterpart in the reference): each band of
 * 56 (F64 / F64S) or 48 (F32S / F16S) rows is cut into `segments` column ranges of nearly
 * equal length, one wavefront each; 0 = chosen from the grid (64 segments, each of at
 * least 8 columns).  Results do not depend on it.  The get variant reports the setting. */
int plbm_set_pass_segments(plbm_lattice *h, int segments);
int plbm_get_pass_segments(const plbm_lattice *h, int *segments);
/* Third population buffer of plbm_run / plbm_run_op in passes (engine extension, no
 * counterpart in the reference).  With it (enable = 1, the default) the last pass of a run
 * also writes the state one step before the latest, which the other slot must hold
 * (update_macros reads it), into a third buffer allocated on the first such run; the
 * three buffers' storage rotates among the two slots and the spare.  With enable = 0 the
 * buffer is freed and runs end on a single-step launch instead (one more sweep over
 * HBM per run, two buffers of memory).  Results are identical either way. */
int plbm_set_pass_lag_buffer(plbm_lattice *h, int enable);
/* Launch timing of plbm_run (engine extension, no counterpart in the reference): with
 * enable = 1 every later plbm_run records HIP events on the lattice's stream before its
 * first launch and after each launch.  plbm_run_launch_times then waits for the last run
 * and returns, per launch in order, the steps it advanced (depths[k]; 1 = the single-step
 * kernel) and its duration in ms (ms[k], between consecutive events); *nlaunches = the
 * number of launches (up to max_launches entries are written; either array may be
 * NULL).  PLBM_ERR_STATE when timing is off. */
int plbm_set_run_timing(plbm_lattice *h, int enable);
int plbm_run_launch_times(plbm_lattice *h, int max_launches, int *depths, float *ms,
                          int *nlaunches);
/* Autotuner: time every launch shape of the fused step for `model` (reps launches from
 * buffer buf into a scratch buffer; the lattice's buffers are not modified), then passes
 * of 2..8 steps against it per lattice update; keep the fastest for this lattice and
 * cache it per (device, storage, model, grid) for the process.  With passes, it also
 * times the pass that ends a run (the one that also stores the lagged state, into the
 * lattice's third buffer, allocated here if needed) at every depth from the chosen one
 * up, and keeps for it the depth that makes runs cheapest; plbm_set_steps_per_launch
 * resets that choice to the depth it sets.  *best_ms (may be NULL) receives the winner's
 * time per lattice-update step. */
int plbm_autotune(plbm_lattice *h, int model, int buf, int reps, double *best_ms);

/* ---------------------------------------------------------------- batched engine */

/* Many independent periodic lattices of uneven size advanced together (no counterpart in
 * the reference, which runs one grid per process; the per-case arithmetic is the same
 * fused step as plbm_step).  Case c has an nx[c] x ny[c] grid, the batch one storage
 * form and one collision model; per-case omega / trt_magic. */
typedef struct plbm_batch plbm_batch;
int plbm_batch_create(int ncases, const int *nx, const int *ny, int storage, int model,
                      plbm_batch **out);
int plbm_batch_destroy(plbm_batch *b);
int plbm_batch_info(const plbm_batch *b, int *ncases, long long *cells, size_t *arena_bytes);
/* omega[c]; trt_magic may be NULL (Lambda = 1/4 for every case). */
int plbm_batch_set_params(plbm_batch *b, const double *omega, const double *trt_magic);
/* Case c's current buffer = feq(rho, ux, uy), (ny, nx) Fortran-order host arrays. */
int plbm_batch_equilibrium_init(plbm_batch *b, int c, const double *rho, const double *ux,
                                const double *uy);
/* Case c's current populations from/to the host layout f(ld, nx, 0:8), unshifted FP64. */
int plbm_batch_upload_f(plbm_batch *b, int c, const double *f, int ld);
int plbm_batch_download_f(plbm_batch *b, int c, double *f, int ld);
/* rho and u = j/rho of case c's current state (host arrays (ny, nx)). */
int plbm_batch_macros(plbm_batch *b, int c, double *rho, double *ux, double *uy);
/* Per-case FP64 reductions of the current state: out[c*PLBM_NRED + PLBM_RED_*]. */
int plbm_batch_reduce(plbm_batch *b, double *out);
/* Advance every case nsteps steps: replays hipGraphs of up to 100 steps (tile-table step
 * kernel per step + one persistent kernel for the small cases on a forked stream).
 * Asynchronous; plbm_batch_run_direct does the same with plain launches. */
int plbm_batch_run(plbm_batch *b, int nsteps);
int plbm_batch_run_direct(plbm_batch *b, int nsteps);
int plbm_batch_sync(plbm_batch *b);
/* Launch plan: tiles of ty x tx cells (ty*tx a multiple of 64, <= 1024); cases with
 * nx*ny <= resident_cells run in the persistent kernel with resident_block threads. */
int plbm_batch_set_launch(plbm_batch *b, int ty, int tx, int resident_cells, int resident_block);
int plbm_batch_get_launch(const plbm_batch *b, int *ty, int *tx, int *resident_cells,
                          int *resident_block);
/* Steps per launch of the tiled (non-resident) cases: 1 = one step per launch (tile
 * table); 2..8 = temporally blocked passes, as
 * plbm_set_steps_per_launch for one lattice (bit-identical in F64 storage).  Default 1. */
int plbm_batch_set_steps_per_pass(plbm_batch *b, int steps);
int plbm_batch_get_steps_per_pass(const plbm_batch *b, int *steps);
/* Columns each wavefront of a pass walks (engine extension, no counterpart in the
 * reference): a case is cut into segments of this many columns.  Default 64; the
 * autotuner tries 16 to 256.  Results do not depend on it. */
int plbm_batch_set_pass_columns(plbm_batch *b, int columns);
int plbm_batch_get_pass_columns(const plbm_batch *b, int *columns);
/* Autotuner: time `steps` (even) steps of every plan candidate (tile shapes, steps per
 * pass, resident threshold; a plan with passes of ns steps over `steps` rounded up to a
 * multiple of 2 ns, so that it runs whole passes) on a scratch copy of the arena twice
 * and keep the fastest; *best_ms (may be NULL) = its time per step. */
int plbm_batch_autotune(plbm_batch *b, int steps, double *best_ms);

/* ---------------------------------------------------------------- adjoint */

/* Discrete adjoint of nsteps BGK steps (no counterpart in the reference).  Forward:
 * f^0 = feq(rho0, u0), f^{n+1} = collide_bgk(lbm_stream(f^n)) with the engine's fused
 * kernel; loss L = 1/2 sum_cells |u(f^N) - u*|^2, u = j/rho.  States are checkpointed
 * every checkpoint_every steps; the reverse sweep recomputes each segment and runs the
 * fused reverse kernel.  storage: PLBM_F64, PLBM_F64S or PLBM_F32S (FP32
 * forward steps and checkpoints); the adjoint state is FP64 for every storage.  All field pointers are device pointers to (ny, nx) Fortran-order
 * FP64 arrays; the initial fields must stay valid until the backward call. */
typedef struct plbm_adjoint plbm_adjoint;
int plbm_adjoint_create(int nx, int ny, int storage, int nsteps, int checkpoint_every,
                        plbm_adjoint **out);
int plbm_adjoint_destroy(plbm_adjoint *a);
int plbm_adjoint_set_stream(plbm_adjoint *a, void *stream);
int plbm_adjoint_forward_dev(plbm_adjoint *a, const double *rho0, const double *ux0,
                             const double *uy0, double omega, const double *uxt,
                             const double *uyt, double *loss);
/* Gradients of the last forward's loss: dL/drho0 (may be NULL), dL/dux0, dL/duy0
 * (device, (ny,nx)) and dL/domega (host scalar). */
int plbm_adjoint_backward_dev(plbm_adjoint *a, double *grho0, double *gux0, double *guy0,
                              double *gomega);

/* Divergence-free band-limited velocity field (streamfunction modes 0 < |k| <= kmax with
 * amplitude 1/|k|^2 and random phases from `seed`), scaled to max |u| = amp. */
int plbm_field_band_limited(int nx, int ny, int kmax, double amp, uint64_t seed, double *ux,
                            double *uy);

/* ---------------------------------------------------------------- measurement */

/* HBM probe: time `iters` launches of a streaming copy of `bytes` bytes (read + write)
 * on the current device; returns the mean kernel time in milliseconds via *ms. */
int plbm_copy_bandwidth(size_t bytes, int iters, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* PLBM_H */
