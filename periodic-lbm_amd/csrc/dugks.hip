// dugks.hip — the discrete unified gas-kinetic scheme of src/periodic_dugks.F90 on FP64
// storage; the default build (no -DDUGKS):
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
//   k_dugks_flux      per row and strip of columns: 9-point stencil of \bar f^+ per
//                     direction from a register window, flux balance subtracted from fp;
//                     fp is either read (dugks_stream) or recomputed as BGK(fold, omega)
//                     from the same cell (the fused step)
// The fused step (plbm_dugks_step) runs k_dugks_collide without the full collision
// (\bar f^+ into the other slot) and k_dugks_flux with the collision recomputed from the
// cell's own old populations, writing f^{n+1} over them in place: 360 bytes per cell per
// step against 432 for the reference's copy + collide + stream sequence.  A run of several
// steps does all but the last in temporally blocked passes (k_dugks_multi, up to three
// steps per sweep, 144 bytes per cell per pass), and the two slots exchange their device
// storage where needed, so that slot contents equal the reference's buffers after the
// same number of perform_dugks_step calls.
// The -DDUGKS build's form (half collision at 0.75 of its rate, which the caller passes,
// and face collisions in the stream) is plbm_dugks_stream_faces / plbm_dugks_step_faces
// (k_dugks_flux_faces), one collide + flux pair per step.
// All FP64 expression trees follow the reference with FMA contraction off (bit-exact).
#include <algorithm>
#include <cmath>
#include <utility>

#include "common.hpp"
#include "d2q9.hpp"
#include "multistep.hpp"
#include "stepper.hpp"

namespace plbm {

// kernel_bgk (src/periodic_dugks.F90:80-165) is collide_bgk's split-weight algebra.
__device__ __forceinline__ void dugks_bgk(double f[9], double omega) {
    collide_bgk_split_ref(f, omega);
}

// One thread per cell on a 1-D grid of (ny/256) x nx workgroups, renumbered so that each
// XCD streams a contiguous range of columns.  `half` may equal `src` (each thread reads
// and writes only its own cell).
template <bool FULL>
__global__ void __launch_bounds__(256)
    k_dugks_collide(const double *src, double *half, double *full, int ny, int ldd, int nby,
                    double om_full, double om_half) {
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int x = b / nby, y = (b % nby) * 256 + threadIdx.x;
    if (y >= ny) return;
    const size_t L = (size_t)ldd, c0 = (size_t)x * 9 * L + y;
    double a[9], h[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) a[q] = h[q] = src[c0 + q * L];
    dugks_bgk(h, om_half);
#pragma unroll
    for (int q = 0; q < 9; ++q) __builtin_nontemporal_store(h[q], half + c0 + q * L);
    if constexpr (FULL) {
        dugks_bgk(a, om_full);
#pragma unroll
        for (int q = 0; q < 9; ++q) __builtin_nontemporal_store(a[q], full + c0 + q * L);
    }
}

// Face values and flux balance of one direction (kernel_stream, src/periodic_dugks.F90:
// 221-233 and 258-270): c = centre, e/n/w/s and corners are the 3x3 neighbourhood of ft.
__device__ __forceinline__ double dugks_face_update(double fp, double fc, double fe, double fn,
                                                    double fw, double fs, double fne,
                                                    double fnw, double fsw, double fse,
                                                    double cxq, double cyq) {
#pragma clang fp contract(off)
    const double p2 = 0.5, p8 = 0.125;
    const double cfw =
        (p2 * (fc + fw) - (p2 * cxq) * (fc - fw)) - (p8 * cyq) * (((fnw + fn) - fsw) - fs);
    const double cfn =
        (p2 * (fc + fn) - (p2 * cyq) * (fn - fc)) - (p8 * cxq) * (((fne + fe) - fnw) - fw);
    const double cfe =
        (p2 * (fc + fe) - (p2 * cxq) * (fe - fc)) - (p8 * cyq) * (((fne + fn) - fse) - fs);
    const double cfs =
        (p2 * (fc + fs) - (p2 * cyq) * (fc - fs)) - (p8 * cxq) * (((fse + fe) - fsw) - fw);
    return (fp - cxq * (cfe - cfw)) - cyq * (cfn - cfs);
}

// kernel_stream (src/periodic_dugks.F90:190-299, default build).  Each thread owns one row
// y and walks a strip of XS consecutive columns, keeping for every direction the 3x3
// neighbourhood of ft in registers: a step to the next column loads only the new column's
// three values per direction (24 loads per cell instead of 72).
// COLLIDE: fp is BGK(fo, om_full) of the thread's own cell (fo may alias fp: the cell is
// read before it is written and no other thread reads fo) and all nine directions are
// written; otherwise fp is read and written for q = 1..8 and q = 0 is left untouched.
template <bool COLLIDE, int XS>
__global__ void __launch_bounds__(256)
    k_dugks_flux(const double *__restrict__ ft, const double *fo, double *fp, int nx, int ny,
                 int ldd, double dt, double om_full) {
    const int y = blockIdx.x * 256 + threadIdx.x;
    const int x0 = blockIdx.y * XS;
    if (y >= ny) return;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const size_t L = (size_t)ldd, cs = 9 * L;
    auto col = [&](int xx) {
        xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
        return (size_t)xx * cs;
    };
    double w[9][3][3];  // [q][column x-1, x, x+1][row y-1, y, y+1]; q = 0 unused
    {
        const size_t cm = col(x0 - 1), c0 = col(x0);
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const size_t o = (size_t)q * L;
            w[q][0][0] = ft[cm + o + ym], w[q][0][1] = ft[cm + o + y], w[q][0][2] = ft[cm + o + yp];
            w[q][1][0] = ft[c0 + o + ym], w[q][1][1] = ft[c0 + o + y], w[q][1][2] = ft[c0 + o + yp];
        }
    }
    for (int i = 0; i < XS; ++i) {
        const int x = x0 + i;
        if (x >= nx) break;
        const size_t c0 = (size_t)x * cs, cp = col(x + 1);
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const size_t o = (size_t)q * L;
            w[q][2][0] = ft[cp + o + ym], w[q][2][1] = ft[cp + o + y], w[q][2][2] = ft[cp + o + yp];
        }
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
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const double v = dugks_face_update(
                g[q], w[q][1][1], w[q][2][1], w[q][1][2], w[q][0][1], w[q][1][0], w[q][2][2],
                w[q][0][2], w[q][0][0], w[q][2][0], dt * double(cx(q)), dt * double(cy(q)));
            __builtin_nontemporal_store(v, fp + c0 + (size_t)q * L + y);
#pragma unroll
            for (int j = 0; j < 3; ++j) w[q][0][j] = w[q][1][j], w[q][1][j] = w[q][2][j];
        }
    }
}

// update_ew / update_ns of the -DDUGKS build (src/periodic_dugks.F90:308-420): BGK
// relaxation of one face's interpolated values towards the face's own equilibrium (rate
// omega); east/west faces update q = 1, 3, 5, 7, 6, 8, north/south faces q = 2, 4, 5, 7, 6,
// 8 (the others carry no flux across that face).
__device__ __forceinline__ void dugks_face_collide(double f[9], double omega, bool ew) {
#pragma clang fp contract(off)
    const double omegabar = 1.0 - omega;
    const double omega_ws = (3.0 * omega) * (1.0 / 9.0), omega_wd = (3.0 * omega) * (1.0 / 36.0);
    const double rho = (((f[5] + f[7]) + (f[6] + f[8])) + ((f[1] + f[3]) + (f[2] + f[4]))) + f[0];
    const double invrho = 1.0 / rho;
    const double ux = invrho * (((f[5] - f[7]) + (f[8] - f[6])) + (f[1] - f[3]));
    const double uy = invrho * (((f[5] - f[7]) + (f[6] - f[8])) + (f[2] - f[4]));
    const double indp = (1.0 / 3.0) - 0.5 * (ux * ux + uy * uy);
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

// kernel_stream of the -DDUGKS build (src/periodic_dugks.F90:190-299 with the face
// collisions at :253-262): per cell the four faces' interpolated values of all nine
// directions (the faces' moments need them all), a face collision on each face (rate
// om_face = 1/(4 tau/dt + 1)), then the flux balance for q = 1..8.  One thread per cell
// reading its 3x3 neighbourhood (the face values of a cell are recomputed by each cell,
// as the reference computes them, so that every expression is the reference's).
// COLLIDE: fp is BGK(fo, om_full) of the cell (fo may alias fp), as in k_dugks_flux.
template <bool COLLIDE>
__global__ void __launch_bounds__(256)
    k_dugks_flux_faces(const double *__restrict__ ft, const double *fo, double *fp, int nx,
                       int ny, int ldd, double dt, double om_full, double om_face) {
#pragma clang fp contract(off)
    const int y = blockIdx.x * 256 + threadIdx.x, x = blockIdx.y;
    if (y >= ny) return;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const size_t L = (size_t)ldd, cs = 9 * L;
    const size_t c0 = (size_t)x * cs, cm = (size_t)xm * cs, cp = (size_t)xp * cs;
    const double p2 = 0.5, p8 = 0.125;
    double cfw[9], cfn[9], cfe[9], cfs[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const size_t o = (size_t)q * L;
        const double cxq = dt * double(cx(q)), cyq = dt * double(cy(q));
        const double fc = ft[c0 + o + y], fe = ft[cp + o + y], fn = ft[c0 + o + yp];
        const double fw = ft[cm + o + y], fs = ft[c0 + o + ym], fne = ft[cp + o + yp];
        const double fnw = ft[cm + o + yp], fsw = ft[cm + o + ym], fse = ft[cp + o + ym];
        cfw[q] = (p2 * (fc + fw) - (p2 * cxq) * (fc - fw)) - (p8 * cyq) * (((fnw + fn) - fsw) - fs);
        cfn[q] = (p2 * (fc + fn) - (p2 * cyq) * (fn - fc)) - (p8 * cxq) * (((fne + fe) - fnw) - fw);
        cfe[q] = (p2 * (fc + fe) - (p2 * cxq) * (fe - fc)) - (p8 * cyq) * (((fne + fn) - fse) - fs);
        cfs[q] = (p2 * (fc + fs) - (p2 * cyq) * (fc - fs)) - (p8 * cxq) * (((fse + fe) - fsw) - fw);
    }
    dugks_face_collide(cfw, om_face, true);
    dugks_face_collide(cfe, om_face, true);
    dugks_face_collide(cfn, om_face, false);
    dugks_face_collide(cfs, om_face, false);
    double g[9];
    if constexpr (COLLIDE) {
#pragma unroll
        for (int q = 0; q < 9; ++q) g[q] = fo[c0 + q * L + y];
        dugks_bgk(g, om_full);
        fp[c0 + y] = g[0];
    } else {
#pragma unroll
        for (int q = 1; q < 9; ++q) g[q] = fp[c0 + q * L + y];
    }
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const double cxq = dt * double(cx(q)), cyq = dt * double(cy(q));
        fp[c0 + (size_t)q * L + y] = (g[q] - cxq * (cfe[q] - cfw[q])) - cyq * (cfn[q] - cfs[q]);
    }
}

// NS DUGKS steps per sweep over device memory (temporal blocking, as the pass kernel of
// the lattice Boltzmann step): a wavefront owns 64 consecutive rows (56 with output, NS
// halo rows on each side) and walks a strip of kMultiXS columns.  For every level k < NS
// it keeps the half-collided state of the last three columns and the fully collided
// state of the previous column in registers; when level k receives a column, level k+1
// gets the column before it: the face values of that column (from the three-column
// window, y-neighbours from the adjacent lanes) are subtracted from its full collision,
// exactly as k_dugks_collide + k_dugks_flux do.  Only level NS is stored, so a pass
// moves 144 bytes per cell for NS steps; the half-collided state is not materialised.
template <int NS>
__global__ void __launch_bounds__(256)
    k_dugks_multi(const double *__restrict__ src, double *__restrict__ dst, int nx, int ny,
                  int ldd, int waves_y, int nwaves, double dt, double om_full,
                  double om_half) {
    constexpr int VR = 56;
    static_assert(VR + 2 * NS <= 64, "halo rows do not fit in a wavefront");
    const int lane = threadIdx.x & 63;
    const int gw = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (gw >= nwaves) return;  // the whole wavefront
    const int wy = gw % waves_y, xsi = gw / waves_y;
    const int r = VR * wy - NS + lane;  // unwrapped row of this lane
    int y = r % ny;
    if (y < 0) y += ny;
    const bool out = lane >= NS && lane < VR + NS && r < ny;
    const int x0 = xsi * kMultiXS;
    const int ncols = min(kMultiXS, nx - x0);
    const size_t L = (size_t)ldd, cs = 9 * L;
    auto column = [&](int c) {
        c %= nx;
        if (c < 0) c += nx;
        return (size_t)c * cs;
    };
    double H[NS][3][9];  // half-collided state, [level][newest column, -1, -2][q]
    double F[NS][9];     // fully collided state of the newest column of each level
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            F[k][q] = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) H[k][j][q] = 0.0;
        }
    double nxt[9];
    {
        const size_t cc = column(x0 - NS);
#pragma unroll
        for (int q = 0; q < 9; ++q) nxt[q] = src[cc + q * L + y];
    }
    const int iters = ncols + 2 * NS;
    for (int i = 0; i < iters; ++i) {
        double f[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) f[q] = nxt[q];
        if (i + 1 < iters) {
            const size_t cn = column(x0 - NS + i + 1);
#pragma unroll
            for (int q = 0; q < 9; ++q) nxt[q] = src[cn + q * L + y];
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            if (i < 2 * k) break;  // level k has no column yet
            double h[9], b[9], bprev[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) h[q] = b[q] = f[q];
            dugks_bgk(h, om_half);
            dugks_bgk(b, om_full);
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                H[k][2][q] = H[k][1][q];
                H[k][1][q] = H[k][0][q];
                H[k][0][q] = h[q];
                bprev[q] = F[k][q];
                F[k][q] = b[q];
            }
            if (i < 2 * k + 2) break;  // level k+1 needs three columns of level k
            f[0] = bprev[0];
#pragma unroll
            for (int q = 1; q < 9; ++q) {
                const double fc = H[k][1][q], fe = H[k][0][q], fw = H[k][2][q];
                f[q] = dugks_face_update(bprev[q], fc, fe, row_plus(fc), fw, row_minus(fc),
                                         row_plus(fe), row_plus(fw), row_minus(fw),
                                         row_minus(fe), dt * double(cx(q)), dt * double(cy(q)));
            }
            if (k + 1 == NS && out) {
                const size_t co = (size_t)(x0 - 2 * NS + i) * cs;
#pragma unroll
                for (int q = 0; q < 9; ++q) __builtin_nontemporal_store(f[q], dst + co + q * L + y);
            }
        }
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

static int dugks_nby(const plbm_lattice *h) { return (h->ny + 255) / 256; }

static int dugks_strip(const plbm_lattice *h) { return strip_columns(h->nx, h->ny); }

template <bool FULL>
static void launch_collide(plbm_lattice *h, const double *src, double *half, double *full,
                           double om_full, double om_half) {
    const int nby = dugks_nby(h);
    hipLaunchKernelGGL(k_dugks_collide<FULL>, dim3((unsigned)(nby * h->nx)), dim3(256), 0,
                       h->stream, src, half, full, h->ny, h->ldd, nby, om_full, om_half);
}

template <bool COLLIDE>
static void launch_flux(plbm_lattice *h, const double *ft, const double *fo, double *fp,
                        double dt, double om_full) {
    const int xs = dugks_strip(h);
    const dim3 grid((unsigned)dugks_nby(h), (unsigned)((h->nx + xs - 1) / xs));
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, h->stream, ft, fo, fp, h->nx, h->ny,
                           h->ldd, dt, om_full);
    };
    if (xs == 32) go(k_dugks_flux<COLLIDE, 32>);
    else if (xs == 8) go(k_dugks_flux<COLLIDE, 8>);
    else go(k_dugks_flux<COLLIDE, 1>);
}

template <bool COLLIDE>
static void launch_flux_faces(plbm_lattice *h, const double *ft, const double *fo, double *fp,
                              double dt, double om_full, double om_face) {
    const dim3 grid((unsigned)dugks_nby(h), (unsigned)h->nx);
    hipLaunchKernelGGL(k_dugks_flux_faces<COLLIDE>, grid, dim3(256), 0, h->stream, ft, fo, fp,
                       h->nx, h->ny, h->ldd, dt, om_full, om_face);
}

}  // namespace plbm

using namespace plbm;

extern "C" int plbm_dugks_stream_faces(plbm_lattice *h, int src, int dst, double dt,
                                       double omega_face) try {
    PLBM_TRY(dugks_check(h, src, dst));
    PLBM_REQUIRE(std::isfinite(dt) && std::isfinite(omega_face), PLBM_ERR_ARG,
                 "non-finite dt or face relaxation rate");
    launch_flux_faces<false>(h, static_cast<const double *>(h->f[src]), nullptr,
                             static_cast<double *>(h->f[dst]), dt, 0.0, omega_face);
    PLBM_HIP(hipGetLastError());
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_dugks_step_faces(plbm_lattice *h, int iold, int inew, int nsteps,
                                     double dt, double omega_full, double omega_half,
                                     double omega_face) try {
    PLBM_TRY(dugks_check(h, iold, inew));
    PLBM_REQUIRE(nsteps >= 0, PLBM_ERR_ARG, "negative step count");
    PLBM_REQUIRE(std::isfinite(dt) && std::isfinite(omega_full) && std::isfinite(omega_half) &&
                     std::isfinite(omega_face),
                 PLBM_ERR_ARG, "non-finite dt or relaxation rate");
    // each step: the half collision into the other slot, then the full collision of the
    // cell minus the face-collided flux balance, in place over the latest state
    double *latest = static_cast<double *>(h->f[iold]);
    double *other = static_cast<double *>(h->f[inew]);
    for (int s = 0; s < nsteps; ++s) {
        launch_collide<false>(h, latest, other, nullptr, omega_full, omega_half);
        launch_flux_faces<true>(h, other, latest, latest, dt, omega_full, omega_face);
    }
    PLBM_HIP(hipGetLastError());
    const int want = (nsteps % 2) ? inew : iold;
    if (nsteps > 0 && h->f[want] != latest) std::swap(h->f[iold], h->f[inew]);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_dugks_collide(plbm_lattice *h, int src, int dst, double omega_full,
                                  double omega_half) try {
    PLBM_TRY(dugks_check(h, src, dst));
    PLBM_REQUIRE(std::isfinite(omega_full) && std::isfinite(omega_half), PLBM_ERR_ARG,
                 "non-finite relaxation rate");
    launch_collide<true>(h, static_cast<const double *>(h->f[src]),
                         static_cast<double *>(h->f[src]), static_cast<double *>(h->f[dst]),
                         omega_full, omega_half);
    PLBM_HIP(hipGetLastError());
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_dugks_stream(plbm_lattice *h, int src, int dst, double dt) try {
    PLBM_TRY(dugks_check(h, src, dst));
    PLBM_REQUIRE(std::isfinite(dt), PLBM_ERR_ARG, "non-finite dt");
    launch_flux<false>(h, static_cast<const double *>(h->f[src]), nullptr,
                       static_cast<double *>(h->f[dst]), dt, 0.0);
    PLBM_HIP(hipGetLastError());
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

namespace plbm {

static int launch_dugks_multi(plbm_lattice *h, int ns, const double *src, double *dst,
                              double dt, double om_full, double om_half) {
    const int waves_y = (h->ny + 55) / 56;
    const int nw = waves_y * ((h->nx + kMultiXS - 1) / kMultiXS);
    const dim3 grid((unsigned)((nw + 3) / 4));
    switch (ns) {
    case 2:
        hipLaunchKernelGGL(k_dugks_multi<2>, grid, dim3(256), 0, h->stream, src, dst, h->nx,
                           h->ny, h->ldd, waves_y, nw, dt, om_full, om_half);
        break;
    case 3:
        hipLaunchKernelGGL(k_dugks_multi<3>, grid, dim3(256), 0, h->stream, src, dst, h->nx,
                           h->ny, h->ldd, waves_y, nw, dt, om_full, om_half);
        break;
    default:
        set_error("DUGKS steps per pass must be 1..3");
        return PLBM_ERR_ARG;
    }
    PLBM_HIP(hipGetLastError());
    return PLBM_OK;
}

// Steps per pass of plbm_dugks_step when the caller has not chosen: 3 (measured best at
// 8192^2: 2.06 ms per step against 2.18 for 2 and 4.37 for 1) where the grid gives at
// least 2048 wavefronts, else 1.
static int dugks_steps_per_pass(const plbm_lattice *h) {
    if (h->dugks_ns > 0) return h->dugks_ns;
    const long waves = (long)((h->ny + 55) / 56) * ((h->nx + kMultiXS - 1) / kMultiXS);
    return waves >= 2048 ? 3 : 1;
}

}  // namespace plbm

extern "C" int plbm_dugks_set_steps_per_pass(plbm_lattice *h, int steps) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(steps >= 0 && steps <= 3, PLBM_ERR_ARG, "DUGKS steps per pass must be 0..3");
    h->dugks_ns = steps;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_dugks_get_steps_per_pass(const plbm_lattice *h, int *steps) try {
    PLBM_REQUIRE(h && steps, PLBM_ERR_ARG, "null argument");
    *steps = dugks_steps_per_pass(h);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_dugks_step(plbm_lattice *h, int iold, int inew, int nsteps, double dt,
                               double omega_full, double omega_half) try {
    PLBM_TRY(dugks_check(h, iold, inew));
    PLBM_REQUIRE(nsteps >= 0, PLBM_ERR_ARG, "negative step count");
    PLBM_REQUIRE(std::isfinite(dt) && std::isfinite(omega_full) && std::isfinite(omega_half),
                 PLBM_ERR_ARG, "non-finite dt or relaxation rate");
    if (nsteps == 0) return PLBM_OK;
    // All steps but the last in passes of up to ns steps (f only); the last step with the
    // collide + flux pair, which leaves the half-collided state in the other buffer.
    const int ns = dugks_steps_per_pass(h);
    double *latest = static_cast<double *>(h->f[iold]);
    double *other = static_cast<double *>(h->f[inew]);
    for (int done = 0; done < nsteps - 1;) {
        const int k = std::min(ns, nsteps - 1 - done);
        if (k == 1) {
            launch_collide<false>(h, latest, other, nullptr, omega_full, omega_half);
            launch_flux<true>(h, other, latest, latest, dt, omega_full);
        } else {
            PLBM_TRY(launch_dugks_multi(h, k, latest, other, dt, omega_full, omega_half));
            std::swap(latest, other);
        }
        done += k;
    }
    launch_collide<false>(h, latest, other, nullptr, omega_full, omega_half);
    launch_flux<true>(h, other, latest, latest, dt, omega_full);
    PLBM_HIP(hipGetLastError());
    // the reference swaps its ids once per step: after an odd count the latest state
    // belongs to slot inew
    const int want = (nsteps % 2) ? inew : iold;
    if (h->f[want] != latest) std::swap(h->f[iold], h->f[inew]);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}
