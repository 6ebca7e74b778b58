// streaming_ops.hip — the reference's finite-volume / finite-difference streaming
// operators on the device, for the grid%streaming slot (src/fvm_bardow.F90:393-893):
//   PLBM_STREAM_FVM_BARDOW   stream_fvm_bardow  (bilinear face fluxes, :410-509)
//   PLBM_STREAM_FDM_BARDOW   stream_fdm_bardow  (central-difference Lax-Wendroff, default
//                                                stencil without FDM_* options, :526-685)
//   PLBM_STREAM_FDM_SOFONEA  stream_fdm_sofonea (1-D Lax-Wendroff along c_q, :702-889)
//   PLBM_STREAM_FDM_BARDOW_WLS .. _ISO  stream_fdm_bardow built with -DFDM_WLS,
//                            -DFDM_WLS_GAUSS_V1, -DFDM_WLS_GAUSS_V2, -DFDM_ISO (:591-669;
//                            instantiated in streaming_ops_fdm.hip)
// Each is a 9-point (3-point for Sofonea) stencil on one q-segment of the column-
// interleaved layout; the rest population is copied.  plbm_step_op fuses the operator with
// the collision that perform_step applies next (one pass, 144 bytes per cell in FP64).
// The operators are linear and leave constants unchanged, so they act on shifted and
// scaled storage directly.  FP64 expression trees follow the reference with FMA
// contraction off, which makes them bit-identical to it.  The reference loops x over ny
// and y over nx (square grids); here x runs over nx and y over ny.
#include <string>
#include <vector>

#include "fd_ops.hpp"

using namespace plbm;

namespace {

// model < 0: streaming only
int launch_fd(plbm_lattice *h, int op, int model, const void *src, void *dst, double dt) {
    switch (op) {
    case PLBM_STREAM_FVM_BARDOW:
        return launch_fd_op<PLBM_STREAM_FVM_BARDOW>(h, model, src, dst, dt);
    case PLBM_STREAM_FDM_BARDOW:
        return launch_fd_op<PLBM_STREAM_FDM_BARDOW>(h, model, src, dst, dt);
    case PLBM_STREAM_FDM_SOFONEA:
        return launch_fd_op<PLBM_STREAM_FDM_SOFONEA>(h, model, src, dst, dt);
    default:
        return launch_fd_variant(h, op, model, src, dst, dt);
    }
}

int launch_fd_multi(plbm_lattice *h, int op, int model, int ns, const void *src, void *dst,
                    double dt, void *lag = nullptr) {
    switch (op) {
    case PLBM_STREAM_FVM_BARDOW:
        return launch_fd_multi_op<PLBM_STREAM_FVM_BARDOW>(h, model, ns, src, dst, dt, lag);
    case PLBM_STREAM_FDM_BARDOW:
        return launch_fd_multi_op<PLBM_STREAM_FDM_BARDOW>(h, model, ns, src, dst, dt, lag);
    case PLBM_STREAM_FDM_SOFONEA:
        return launch_fd_multi_op<PLBM_STREAM_FDM_SOFONEA>(h, model, ns, src, dst, dt, lag);
    default:
        return launch_fd_multi_variant(h, op, model, ns, src, dst, dt, lag);
    }
}

int check_op_args(plbm_lattice *h, int op, int src, int dst, double dt) {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(src >= 0 && src < h->nf && dst >= 0 && dst < h->nf && src != dst, PLBM_ERR_ARG,
                 "need distinct valid src and dst buffers");
    PLBM_REQUIRE(std::isfinite(dt), PLBM_ERR_ARG, "non-finite dt");
    PLBM_REQUIRE(op >= PLBM_STREAM_EXACT && op <= PLBM_STREAM_FDM_BARDOW_ISO, PLBM_ERR_ARG,
                 "unknown streaming operator " + std::to_string(op));
    return PLBM_OK;
}

}  // namespace

extern "C" int plbm_stream_op(plbm_lattice *h, int op, int src, int dst, double dt) try {
    PLBM_TRY(check_op_args(h, op, src, dst, dt));
    if (op == PLBM_STREAM_EXACT) return plbm_stream(h, src, dst);
    return launch_fd(h, op, -1, h->f[src], h->f[dst], dt);
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_step_op(plbm_lattice *h, int op, int model, int src, int dst, double dt) try {
    PLBM_TRY(check_op_args(h, op, src, dst, dt));
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    PLBM_REQUIRE(h->params_set, PLBM_ERR_STATE, "plbm_set_params has not been called");
    if (op == PLBM_STREAM_EXACT) return plbm_step(h, model, src, dst);
    return launch_fd(h, op, model, h->f[src], h->f[dst], dt);
} catch (...) {
    return plbm::host_exception();
}

// Steps per pass of plbm_run_op when the caller has not chosen, where the grid gives at
// least 2048 wavefronts (FP64 storage; else single steps): 3, the most a pass holds
// (8192^2, ms per step in passes of 2 / 3: FVM + BGK 1.68 / 1.66, FVM + TRT 1.35 / 1.35,
// FDM + BGK 1.11 / 1.08, FDM + TRT 1.09 / 1.04, Sofonea + BGK 1.00 / 0.69; single steps
// about 2.0).
static int fd_steps_per_pass(const plbm_lattice *h, int op, int model) {
    (void)op;
    (void)model;
    if (h->fd_ns > 0) return h->fd_ns;
    if (h->storage != PLBM_F64) return 1;
    const long waves = (long)((h->ny + 55) / 56) * ((h->nx + kMultiXS - 1) / kMultiXS);
    return waves < 2048 ? 1 : 3;
}

extern "C" int plbm_set_op_steps_per_pass(plbm_lattice *h, int steps) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(steps >= 0 && steps <= 3, PLBM_ERR_ARG, "steps per pass must be 0..3");
    PLBM_REQUIRE(steps <= 1 || h->storage == PLBM_F64, PLBM_ERR_STATE,
                 "passes of the streaming operators need FP64 (PLBM_F64) storage");
    h->fd_ns = steps;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_get_op_steps_per_pass(const plbm_lattice *h, int op, int model,
                                          int *steps) try {
    PLBM_REQUIRE(h && steps, PLBM_ERR_ARG, "null argument");
    *steps = fd_steps_per_pass(h, op, model);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_run_op(plbm_lattice *h, int op, int model, int nsteps, int *iold,
                           int *inew, double dt) try {
    PLBM_REQUIRE(h && iold && inew, PLBM_ERR_ARG, "null argument");
    PLBM_REQUIRE(nsteps >= 0, PLBM_ERR_ARG, "nsteps must be >= 0");
    PLBM_TRY(check_op_args(h, op, *iold, *inew, dt));
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    PLBM_REQUIRE(h->params_set, PLBM_ERR_STATE, "plbm_set_params has not been called");
    if (op == PLBM_STREAM_EXACT) return plbm_run(h, model, nsteps, iold, inew);
    const int ns = fd_steps_per_pass(h, op, model);
    if (ns <= 1 || nsteps < 2) {
        for (int s = 0; s < nsteps; ++s) {
            PLBM_TRY(launch_fd(h, op, model, h->f[*iold], h->f[*inew], dt));
            std::swap(*iold, *inew);
        }
        return PLBM_OK;
    }
    // passes of up to ns steps; the last, as deep as a pass goes, also stores the state one
    // step before the latest into the third buffer, which becomes the other slot's storage
    // (as after perform_step).  Without a third buffer the last step is a single fused step
    // instead.
    const int a0 = *iold, b0 = *inew;
    void *latest = h->f[a0], *other = h->f[b0];
    void *aux = pass_aux(h);
    const std::vector<int> plan = aux ? lag_pass_plan(nsteps, ns) : pass_plan(nsteps - 1, ns);
    for (size_t p = 0; p + 1 < plan.size(); ++p) {
        if (plan[p] == 1) PLBM_TRY(launch_fd(h, op, model, latest, other, dt));
        else PLBM_TRY(launch_fd_multi(h, op, model, plan[p], latest, other, dt));
        std::swap(latest, other);
    }
    if (aux) {
        PLBM_TRY(launch_fd_multi(h, op, model, plan.back(), latest, other, dt, aux));
        finish_run(h, a0, b0, nsteps, other, aux, latest, iold, inew);
        return PLBM_OK;
    }
    if (plan.back() == 1) PLBM_TRY(launch_fd(h, op, model, latest, other, dt));
    else PLBM_TRY(launch_fd_multi(h, op, model, plan.back(), latest, other, dt));
    std::swap(latest, other);
    PLBM_TRY(launch_fd(h, op, model, latest, other, dt));
    finish_run(h, a0, b0, nsteps, other, latest, nullptr, iold, inew);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}
