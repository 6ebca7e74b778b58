// fields.cpp — host-side initial and analytic fields (flow cases), FP64.
//
// Built with -ffp-contract=off: each expression follows the reference's evaluation
// order so the fields equal the reference's bit for bit when both call the same libm
// (glibc sqrt/sin/cos/exp/tanh).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/plbm.h"

namespace plbm {
int host_exception() noexcept;  // lattice.hip
}

extern "C" {

// taylor_green_t constructor + eval, src/benchmarks/taylor_green.f90:31-84.
// Cell centres are (x-1)+1/2 (:67,69).  The reference's inner loop runs y = 1..nx (:68),
// which is the same as 1..ny on the square grids it is used with; here y runs over ny.
int plbm_field_taylor_green(int nx, int ny, double kx, double ky, double umax, double nu,
                            double t, double *p, double *ux, double *uy) try {
    if (nx < 1 || ny < 1 || !p || !ux || !uy || !(kx > 0) || !(ky > 0) || !(nu > 0))
        return PLBM_ERR_ARG;
    const double td = 1.0 / (nu * (kx * kx + ky * ky));
    const double sx = std::sqrt(ky / kx), sy = std::sqrt(kx / ky);
    const double ex = std::exp(-(t / td)), ep = std::exp(-((2.0 * t) / td));
    const double pa = 0.25 * (umax * umax), rk = ky / kx, rk2 = kx / ky;
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        const double xx = double(x) + 0.5;
        const double cxx = std::cos(kx * xx), sxx = std::sin(kx * xx);
        const double c2x = std::cos((2.0 * kx) * xx);
        for (int y = 0; y < ny; ++y) {
            const double yy = double(y) + 0.5;
            const size_t c = (size_t)x * ny + y;
            ux[c] = -((((umax * sx) * cxx) * std::sin(ky * yy)) * ex);
            uy[c] = (((umax * sy) * sxx) * std::cos(ky * yy)) * ex;
            p[c] = -((pa * (rk * c2x + rk2 * std::cos((2.0 * ky) * yy))) * ep);
        }
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

// eval_shear_layer_grid, src/benchmarks/shear_layer.f90:25-52.
int plbm_field_shear_layer(int nx, int ny, double length, double u0, double k, double delta,
                           double *rho, double *ux, double *uy) try {
    if (nx < 1 || ny < 1 || !rho || !ux || !uy || !(length > 0)) return PLBM_ERR_ARG;
    const double pi = 4.0 * std::atan(1.0);
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        const double xd = (double(x) + 0.5) / length;
        const double v = (u0 * delta) * std::sin((2.0 * pi) * (xd + 0.25));
        for (int y = 0; y < ny; ++y) {
            const double yd = (double(y) + 0.5) / length;
            const size_t c = (size_t)x * ny + y;
            uy[c] = v;
            ux[c] = yd <= 0.5 ? u0 * std::tanh(k * (yd - 0.25)) : u0 * std::tanh(k * (0.75 - yd));
            rho[c] = 1.0;
        }
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

// eval_vortex_case, src/benchmarks/barotropic_vortex_case.F90:34-83 (Wissocq et al. 2020
// convected vortex with barotropic density, Eqs. (3), (4), (20)).
int plbm_field_barotropic_vortex(int nx, int ny, double u0, double xc, double yc, double rc,
                                 double eps, double rho0, double csqr, double *rho, double *ux,
                                 double *uy) try {
    if (nx < 1 || ny < 1 || !rho || !ux || !uy || !(rc > 0) || !(csqr > 0)) return PLBM_ERR_ARG;
    const double half = 1.0 / 2.0;
    const double rcsqr = rc * rc;
    const double vma_sq = (eps * eps) / csqr;
    #pragma omp parallel for schedule(static)
    for (int x = 0; x < nx; ++x) {
        const double xl = double(x) + 0.5;
        for (int y = 0; y < ny; ++y) {
            const double yl = double(y) + 0.5;
            const double xr = xl - xc, yr = yl - yc;
            const double rsqr = xr * xr + yr * yr;
            const double g = std::exp(-((half * rsqr) / rcsqr));
            const size_t c = (size_t)x * ny + y;
            ux[c] = u0 - (eps * (yr / rc)) * g;
            uy[c] = (eps * (xr / rc)) * g;
            rho[c] = rho0 * std::exp(-((half * vma_sq) * std::exp(-(rsqr / rcsqr))));
        }
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

// Counter-based uniform deviate in [0,1): splitmix64's output function applied to
// seed + (counter+1)*0x9E3779B97F4A7C15, top 53 bits.
double plbm_noise(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}

// Divergence-free band-limited velocity: streamfunction
//   psi(x, y) = sum_{0 < |k| <= kmax, half plane} cos(2 pi (m x + n y)/L + phi_k) / |k|^2
// on cell centres, u = d psi/dy, v = -d psi/dx, scaled so that max |u| = amp.  Phases
// phi_k = 2 pi U with U the generator's deviates of `seed` (one per mode, in the order
// the modes are enumerated: m from -kmax to kmax, n from 0 to kmax).
int plbm_field_band_limited(int nx, int ny, int kmax, double amp, uint64_t seed, double *ux,
                            double *uy) try {
    if (nx < 1 || ny < 1 || kmax < 1 || !ux || !uy) return PLBM_ERR_ARG;
    const double pi = 4.0 * std::atan(1.0);
    struct Mode { double kx, ky, a, phi; };
    std::vector<Mode> modes;
    uint64_t counter = 0;
    for (int m = -kmax; m <= kmax; ++m)
        for (int n = 0; n <= kmax; ++n) {
            if (m * m + n * n == 0 || m * m + n * n > kmax * kmax) continue;
            if (n == 0 && m < 0) continue;  // half plane: (m, n) and (-m, -n) are one mode
            const double k2 = double(m * m + n * n);
            modes.push_back({2.0 * pi * m / nx, 2.0 * pi * n / ny, 1.0 / k2,
                             2.0 * pi * plbm_noise(seed, counter++)});
        }
    double umax = 0.0;
    #pragma omp parallel for schedule(static) reduction(max : umax)
    for (int x = 0; x < nx; ++x) {
        const double xx = double(x) + 0.5;
        for (int y = 0; y < ny; ++y) {
            const double yy = double(y) + 0.5;
            double u = 0.0, v = 0.0;
            for (const Mode &md : modes) {
                const double s = std::sin(md.kx * xx + md.ky * yy + md.phi);
                u += -md.a * md.ky * s;  // d/dy cos(...) = -ky sin(...)
                v += md.a * md.kx * s;   // -d/dx cos(...) = kx sin(...)
            }
            const size_t c = (size_t)x * ny + y;
            ux[c] = u;
            uy[c] = v;
            umax = std::max(umax, std::sqrt(u * u + v * v));
        }
    }
    const double s = umax > 0 ? amp / umax : 0.0;
    const size_t n = (size_t)nx * ny;
    for (size_t c = 0; c < n; ++c) {
        ux[c] *= s;
        uy[c] *= s;
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_field_add_noise(int n, uint64_t seed, double amp, double *ux, double *uy) try {
    if (n < 0 || !ux || !uy) return PLBM_ERR_ARG;
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        ux[c] += amp * (2.0 * plbm_noise(seed, 2 * (uint64_t)c) - 1.0);
        uy[c] += amp * (2.0 * plbm_noise(seed, 2 * (uint64_t)c + 1) - 1.0);
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

}  // extern "C"
