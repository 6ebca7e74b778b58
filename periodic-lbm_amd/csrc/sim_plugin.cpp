// sim_plugin.cpp — the reference's side-tree plugin protocol on the GPU engine.
//
// ivan-pi/periodic-lbm drives its sim/ plugins through ctypes by name: a library
// lib<name>.so exports c_<name>_{init,step,vars,free,norm} with the signatures of
// sim/lbm.h:13-19 (found by SimPlugin, sim/cases.py:16-66).  This library is libplbm.so,
// so the plugin entry points are c_plbm_*.  Semantics follow the standard-LBM plugin
// sim/sim_slbm.F90:135-205:
//   init(nx, ny, dt, p, u, sigma, params): dt must be 1; the second argument is the
//     pressure field p(nx, ny) (rho = 1 + p/csqr, sim/sim.F90:181-199), u(nx, ny, 2);
//     populations start at equilibrium; sigma (non-equilibrium start) is not supported.
//   step(sim, omega): one BGK stream-collide step with the given omega.
//   vars(sim, rho, u): density and velocity of the current state.
//   norm(nx, ny, u, ua) = ||u - ua||_2 / ||ua||_2.
// Arrays are Fortran order with x fastest: (nx, ny) and (nx, ny, 2).  The state is kept
// as shifted FP64 populations (the plugin's DDF shift, sim/sim.F90:130-131); params may
// point to an int storage id (PLBM_F64 / F64S / F32S / F16S) to override it.
//
// The reference plugin collides and then push-streams; this one pull-streams and then
// collides.  From an equilibrium start the two produce the same macroscopic fields
// after every step (collision conserves rho and j, and is the identity on feq).
#include <cmath>
#include <cstdio>
#include <exception>
#include <new>
#include <vector>

#include "../../include/plbm.h"
#include "../../include/plbm_sim.h"

namespace {

struct SimPlugin {
    plbm_lattice *h = nullptr;
    int nx = 0, ny = 0;
    int iold = 0, inew = 1;
    std::vector<double> a, b, c;  // (ny, nx) staging
};

void report(const char *what) {
    std::fprintf(stderr, "[c_plbm] %s: %s\n", what, plbm_last_error());
}

}  // namespace

extern "C" {

void *c_plbm_init(int nx, int ny, double dt, const double *p, const double *u,
                  const double *sigma, void *params) {
    if (dt != 1.0) {
        std::fprintf(stderr, "Standard LBM only supports dt = 1.0!\n");
        return nullptr;
    }
    if (sigma) {
        std::fprintf(stderr, "[c_plbm] NotImplementedError: non-equilibrium initialisation\n");
        return nullptr;
    }
    if (nx < 1 || ny < 1 || !p || !u) {
        std::fprintf(stderr, "[c_plbm] invalid arguments\n");
        return nullptr;
    }
    const int storage = params ? *static_cast<const int *>(params) : PLBM_F64S;
    auto *s = new (std::nothrow) SimPlugin();
    if (!s) {
        std::fprintf(stderr, "[c_plbm] init: host memory allocation failed\n");
        return nullptr;
    }
    s->nx = nx;
    s->ny = ny;
    if (plbm_create(nx, ny, 2, storage, &s->h) != PLBM_OK) {
        report("init");
        delete s;
        return nullptr;
    }
    const size_t n = (size_t)nx * ny;
    try {
        s->a.resize(n);
        s->b.resize(n);
        s->c.resize(n);
    } catch (const std::exception &) {  // no exception leaves the plugin's C interface
        std::fprintf(stderr, "[c_plbm] init: host memory allocation failed\n");
        plbm_destroy(s->h);
        delete s;
        return nullptr;
    }
    const double csqr = 1.0 / 3.0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const size_t in = (size_t)y * nx + x, out = (size_t)x * ny + y;
            s->a[out] = 1.0 + p[in] / csqr;
            s->b[out] = u[in];
            s->c[out] = u[n + in];
        }
    if (plbm_equilibrium_init(s->h, 0, s->a.data(), s->b.data(), s->c.data()) != PLBM_OK) {
        report("init");
        plbm_destroy(s->h);
        delete s;
        return nullptr;
    }
    return s;
}

void c_plbm_step(void *sim, double omega) {
    auto *s = static_cast<SimPlugin *>(sim);
    if (!s) return;
    if (plbm_set_params(s->h, omega, 0.25) != PLBM_OK ||
        plbm_step(s->h, PLBM_BGK, s->iold, s->inew) != PLBM_OK) {
        report("step");
        return;
    }
    std::swap(s->iold, s->inew);
}

void c_plbm_vars(void *sim, double *rho, double *u) {
    auto *s = static_cast<SimPlugin *>(sim);
    if (!s || !rho || !u) return;
    if (plbm_macros(s->h, s->iold, s->a.data(), s->b.data(), s->c.data()) != PLBM_OK) {
        report("vars");
        return;
    }
    const size_t n = (size_t)s->nx * s->ny;
    for (int y = 0; y < s->ny; ++y)
        for (int x = 0; x < s->nx; ++x) {
            const size_t out = (size_t)y * s->nx + x, in = (size_t)x * s->ny + y;
            rho[out] = s->a[in];
            u[out] = s->b[in];
            u[n + out] = s->c[in];
        }
}

void c_plbm_free(void *sim) {
    auto *s = static_cast<SimPlugin *>(sim);
    if (!s) return;
    plbm_destroy(s->h);
    delete s;
}

double c_plbm_norm(int nx, int ny, const double *u, const double *ua) {
    double num = 0.0, den = 0.0;
    const size_t n = (size_t)nx * ny;
    for (size_t i = 0; i < n; ++i) {
        const double d = u[i] - ua[i];
        num += d * d;
        den += ua[i] * ua[i];
    }
    return std::sqrt(num) / std::sqrt(den);
}

}  // extern "C"
