// dugks.hip — the discrete unified gas-kinetic scheme of src/periodic_dugks.F90 (default
// build, no -DDUGKS) on FP64 storage:
//
//   dugks_collide  (:46-77)   fnew = BGK(fold, grid%omega)           (the copy + collision)
//                             fold = BGK(fold, 1/(tau/dt + 1/2))      (in place: \bar f^+)
//   dugks_stream   (:172-299) face values of \bar f^+ by bilinear interpolation, then
//                             fnew_q -= dt c_q . (face flux balance), q = 1..8
//   perform_dugks_step (:25-38) = collide, stream, swap iold/inew
//
// Two kernels carry each form:
//   k_dugks_collide   per cell: \bar f^+ into `half` (in place on the source for the
//                     reference's dugks_collide) and optionally BGK(fold, omega) into `full`
//   k_dugks_flux      per cell: 9-point stencil of \bar f^+ per direction, flux balance
//                     subtracted from fp; fp is either read (dugks_stream) or recomputed
//                     as BGK(fold, omega) from the same cell (the fused step)
// The fused step (plbm_dugks_step) runs k_dugks_collide without the full collision
// (\bar f^+ into the other slot) and k_dugks_flux with the collision recomputed from the
// cell's own old populations, writing f^{n+1} over them in place: 360 bytes per cell per
// step against 432 for the reference's copy + collide + stream sequence.  After an odd
// number of steps the two slots exchange their device storage, so that slot contents
// equal the reference's buffers after the same number of perform_dugks_step calls.
// All FP64 expression trees follow the reference with FMA contraction off (bit-exact).
#include <cmath>
#include <utility>

#include "common.hpp"
#include "d2q9.hpp"

namespace plbm {

// kernel_bgk (src/periodic_dugks.F90:80-165): BGK in the split-weight algebra.
__device__ __forceinline__ void dugks_bgk(double f[9], double omega) {
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

// One thread per cell; blockIdx.y = x, blockIdx.x * blockDim.x + threadIdx.x = y.
// `half` may equal `src` (each thread reads and writes only its own cell).
template <bool FULL>
__global__ void __launch_bounds__(256)
    k_dugks_collide(const double *src, double *half, double *full, int ny, int ldd,
                    double om_full, double om_half) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ny) return;
    const size_t L = (size_t)ldd, c0 = (size_t)blockIdx.y * 9 * L + y;
    double a[9], b[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) a[q] = b[q] = src[c0 + q * L];
    dugks_bgk(b, om_half);
#pragma unroll
    for (int q = 0; q < 9; ++q) __builtin_nontemporal_store(b[q], half + c0 + q * L);
    if constexpr (FULL) {
        dugks_bgk(a, om_full);
#pragma unroll
        for (int q = 0; q < 9; ++q) __builtin_nontemporal_store(a[q], full + c0 + q * L);
    }
}

// kernel_stream (src/periodic_dugks.F90:190-299, default build).  COLLIDE: fp is
// BGK(fo, om_full) of the thread's own cell (fo may alias fp), all nine directions are
// written; otherwise fp is read and written for q = 1..8 and q = 0 is left untouched.
template <bool COLLIDE>
__global__ void __launch_bounds__(256)
    k_dugks_flux(const double *__restrict__ ft, const double *fo, double *fp, int nx, int ny,
                 int ldd, double dt, double om_full) {
#pragma clang fp contract(off)
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= ny) return;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const size_t L = (size_t)ldd, cs = 9 * L;
    const size_t c0 = (size_t)x * cs, cm = (size_t)xm * cs, cp = (size_t)xp * cs;
    double g[9];
    if constexpr (COLLIDE) {
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = fo[c0 + q * L + y];
        dugks_bgk(g, om_full);
        __builtin_nontemporal_store(g[0], fp + c0 + y);
    } else {
#pragma unroll
        for (int q = 1; q < 9; ++q) g[q] = fp[c0 + q * L + y];
    }
    const double p2 = 0.5, p8 = 0.125;
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const size_t o = (size_t)q * L;
        const double fc = ft[c0 + o + y], fe = ft[cp + o + y], fn = ft[c0 + o + yp],
                     fw = ft[cm + o + y], fs = ft[c0 + o + ym], fne = ft[cp + o + yp],
                     fnw = ft[cm + o + yp], fsw = ft[cm + o + ym], fse = ft[cp + o + ym];
        const double cxq = dt * double(cx(q)), cyq = dt * double(cy(q));
        const double cfw =
            (p2 * (fc + fw) - (p2 * cxq) * (fc - fw)) - (p8 * cyq) * (((fnw + fn) - fsw) - fs);
        const double cfn =
            (p2 * (fc + fn) - (p2 * cyq) * (fn - fc)) - (p8 * cxq) * (((fne + fe) - fnw) - fw);
        const double cfe =
            (p2 * (fc + fe) - (p2 * cxq) * (fe - fc)) - (p8 * cyq) * (((fne + fn) - fse) - fs);
        const double cfs =
            (p2 * (fc + fs) - (p2 * cyq) * (fc - fs)) - (p8 * cxq) * (((fse + fe) - fsw) - fw);
        __builtin_nontemporal_store((g[q] - cxq * (cfe - cfw)) - cyq * (cfn - cfs),
                                    fp + c0 + o + y);
    }
}

static int dugks_check(plbm_lattice *h, int a, int b) {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(a >= 0 && a < h->nf && b >= 0 && b < h->nf && a != b, PLBM_ERR_ARG,
                 "need distinct valid buffers");
    PLBM_REQUIRE(h->storage == PLBM_F64, PLBM_ERR_STATE,
                 "the DUGKS operators run on FP64 (PLBM_F64) storage");
    return PLBM_OK;
}

static dim3 dugks_grid(const plbm_lattice *h) {
    return dim3((unsigned)((h->ny + 255) / 256), (unsigned)h->nx);
}

}  // namespace plbm

using namespace plbm;

extern "C" int plbm_dugks_collide(plbm_lattice *h, int src, int dst, double omega_full,
                                  double omega_half) {
    PLBM_TRY(dugks_check(h, src, dst));
    PLBM_REQUIRE(std::isfinite(omega_full) && std::isfinite(omega_half), PLBM_ERR_ARG,
                 "non-finite relaxation rate");
    hipLaunchKernelGGL(k_dugks_collide<true>, dugks_grid(h), dim3(256), 0, h->stream,
                       static_cast<const double *>(h->f[src]), static_cast<double *>(h->f[src]),
                       static_cast<double *>(h->f[dst]), h->ny, h->ldd, omega_full, omega_half);
    PLBM_HIP(hipGetLastError());
    return PLBM_OK;
}

extern "C" int plbm_dugks_stream(plbm_lattice *h, int src, int dst, double dt) {
    PLBM_TRY(dugks_check(h, src, dst));
    PLBM_REQUIRE(std::isfinite(dt), PLBM_ERR_ARG, "non-finite dt");
    hipLaunchKernelGGL(k_dugks_flux<false>, dugks_grid(h), dim3(256), 0, h->stream,
                       static_cast<const double *>(h->f[src]), nullptr,
                       static_cast<double *>(h->f[dst]), h->nx, h->ny, h->ldd, dt, 0.0);
    PLBM_HIP(hipGetLastError());
    return PLBM_OK;
}

extern "C" int plbm_dugks_step(plbm_lattice *h, int iold, int inew, int nsteps, double dt,
                               double omega_full, double omega_half) {
    PLBM_TRY(dugks_check(h, iold, inew));
    PLBM_REQUIRE(nsteps >= 0, PLBM_ERR_ARG, "negative step count");
    PLBM_REQUIRE(std::isfinite(dt) && std::isfinite(omega_full) && std::isfinite(omega_half),
                 PLBM_ERR_ARG, "non-finite dt or relaxation rate");
    double *latest = static_cast<double *>(h->f[iold]);
    double *half = static_cast<double *>(h->f[inew]);
    const dim3 grid = dugks_grid(h);
    for (int s = 0; s < nsteps; ++s) {
        hipLaunchKernelGGL(k_dugks_collide<false>, grid, dim3(256), 0, h->stream, latest, half,
                           nullptr, h->ny, h->ldd, omega_full, omega_half);
        hipLaunchKernelGGL(k_dugks_flux<true>, grid, dim3(256), 0, h->stream, half, latest,
                           latest, h->nx, h->ny, h->ldd, dt, omega_full);
    }
    PLBM_HIP(hipGetLastError());
    if (nsteps % 2 == 1) std::swap(h->f[iold], h->f[inew]);
    return PLBM_OK;
}
