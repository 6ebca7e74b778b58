// streaming_ops.hip — the reference's finite-volume / finite-difference streaming
// operators on the device, for the grid%streaming slot (src/fvm_bardow.F90:393-893):
//   PLBM_STREAM_FVM_BARDOW   stream_fvm_bardow  (bilinear face fluxes, :410-509)
//   PLBM_STREAM_FDM_BARDOW   stream_fdm_bardow  (central-difference Lax-Wendroff, default
//                                                stencil without FDM_* options, :526-685)
//   PLBM_STREAM_FDM_SOFONEA  stream_fdm_sofonea (1-D Lax-Wendroff along c_q, :702-889)
// Each is a 9-point (3-point for Sofonea) stencil on one q-segment of the column-
// interleaved layout; the rest population is copied.  plbm_step_op fuses the operator with
// the collision that perform_step applies next (one pass, 144 bytes per cell in FP64).
// The operators are linear and leave constants unchanged, so they act on shifted and
// scaled storage directly.  FP64 expression trees follow the reference with FMA
// contraction off, which makes them bit-identical to it.  The reference loops x over ny
// and y over nx (square grids); here x runs over nx and y over ny.
#include <cmath>

#include <algorithm>
#include <utility>

#include "common.hpp"
#include "d2q9.hpp"
#include "multistep.hpp"
#include "stepper.hpp"

namespace plbm {

// One direction's update from its 3x3 neighbourhood (c = centre, e/n/w/s and corners).
template <int OP, class R>
__device__ __forceinline__ R fd_update(int q, R fc, R fe, R fn, R fw, R fs, R fne, R fnw, R fsw,
                                       R fse, R dtr, R p2s) {
#pragma clang fp contract(off)
    if constexpr (OP == PLBM_STREAM_FVM_BARDOW || OP == PLBM_STREAM_FDM_BARDOW) {
        const R cxq = dtr * R(cx(q)), cyq = dtr * R(cy(q));
        if constexpr (OP == PLBM_STREAM_FVM_BARDOW) {
            const R p2 = R(0.5), p8 = R(0.125);
            const R cfw = (p2 * (fc + fw) - (p2 * cxq) * (fc - fw)) -
                          (p8 * cyq) * (((fnw + fn) - fsw) - fs);
            const R cfn = (p2 * (fc + fn) - (p2 * cyq) * (fn - fc)) -
                          (p8 * cxq) * (((fne + fe) - fnw) - fw);
            const R cfe = (p2 * (fc + fe) - (p2 * cxq) * (fe - fc)) -
                          (p8 * cyq) * (((fne + fn) - fse) - fs);
            const R cfs = (p2 * (fc + fs) - (p2 * cyq) * (fc - fs)) -
                          (p8 * cxq) * (((fse + fe) - fsw) - fw);
            return (fc - cxq * (cfe - cfw)) - cyq * (cfn - cfs);
        } else {
            const R p2 = R(0.5);
            const R cxxq = (R(0.5) * cxq) * cxq, cyyq = (R(0.5) * cyq) * cyq, cxyq = cxq * cyq;
            const R dfx = p2 * (fe - fw), dfy = p2 * (fn - fs);
            const R dfxx = (fe - R(2) * fc) + fw, dfyy = (fn - R(2) * fc) + fs;
            const R dfxy = R(0.25) * (((fne - fse) - fnw) + fsw);
            return ((fc - cxq * dfx) - cyq * dfy) + ((cxxq * dfxx + cxyq * dfxy) + cyyq * dfyy);
        }
    } else {
        // fu at x + c_q (upstream of the pull), fd at x - c_q
        const int ux = cx(q), uy = cy(q);
        const R fu = ux > 0 ? (uy > 0 ? fne : (uy < 0 ? fse : fe))
                   : ux < 0 ? (uy > 0 ? fnw : (uy < 0 ? fsw : fw))
                            : (uy > 0 ? fn : fs);
        const R fd = ux > 0 ? (uy > 0 ? fsw : (uy < 0 ? fnw : fw))
                   : ux < 0 ? (uy > 0 ? fse : (uy < 0 ? fne : fe))
                            : (uy > 0 ? fs : fn);
        R du1, du2;
        if (q < 5) {
            du1 = R(0.5) * (fu - fd);
            du2 = (fu - R(2) * fc) + fd;
        } else {
            du1 = p2s * (fu - fd);
            du2 = R(0.5) * ((fu - R(2) * fc) + fd);
        }
        return fc + dtr * ((R(0.5) * dtr) * du2 - du1);
    }
}

// Each thread owns one row y and walks a strip of XS consecutive columns, keeping the
// 3x3 neighbourhood of every moving direction in registers, so that a step to the next
// column loads only that column's three values per direction.  MODEL < 0: streaming only,
// the rest population copied as stored.  MODEL >= 0: the streamed cell is collided before
// it is stored (perform_step's stream + collide pair in one pass; in FP64 storage the
// values are the unfused ones, since the unfused path stores the intermediate exactly).
template <int OP, int ST, int MODEL, int XS>
__global__ void __launch_bounds__(256)
    k_stream_fd(StepArgs a, double dt, double p2s) {
    using T = typename Traits<ST>::T;
    using R = typename Traits<ST>::R;
    const int y = blockIdx.x * 256 + threadIdx.x;
    const int x0 = blockIdx.y * XS;
    const int nx = a.nx, ny = a.ny;
    if (y >= ny) return;
    const T *__restrict__ s = static_cast<const T *>(a.src);
    T *__restrict__ d = static_cast<T *>(a.dst);
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const size_t L = (size_t)a.ldd, cs = a.cs;
    auto col = [&](int xx) {
        xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
        return (size_t)xx * cs;
    };
    auto ld_ = [&](size_t c, int yy) { return load_pop<ST>(s + c + yy, a.inv_scale); };
    const R dtr = R(dt), p2r = R(p2s);
    R w[9][3][3];  // [q][column x-1, x, x+1][row y-1, y, y+1]; q = 0 unused
    {
        const size_t cm = col(x0 - 1), c0 = col(x0);
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const size_t o = (size_t)q * L;
            w[q][0][0] = ld_(cm + o, ym), w[q][0][1] = ld_(cm + o, y), w[q][0][2] = ld_(cm + o, yp);
            w[q][1][0] = ld_(c0 + o, ym), w[q][1][1] = ld_(c0 + o, y), w[q][1][2] = ld_(c0 + o, yp);
        }
    }
    for (int i = 0; i < XS; ++i) {
        const int x = x0 + i;
        if (x >= nx) break;
        const size_t c0 = (size_t)x * cs, cp = col(x + 1);
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            const size_t o = (size_t)q * L;
            w[q][2][0] = ld_(cp + o, ym), w[q][2][1] = ld_(cp + o, y), w[q][2][2] = ld_(cp + o, yp);
        }
        R f[9];
#pragma unroll
        for (int q = 1; q < 9; ++q)
            f[q] = fd_update<OP, R>(q, w[q][1][1], w[q][2][1], w[q][1][2], w[q][0][1], w[q][1][0],
                                    w[q][2][2], w[q][0][2], w[q][0][0], w[q][2][0], dtr, p2r);
        if constexpr (MODEL < 0) {
            d[c0 + y] = s[c0 + y];
        } else {
            f[0] = ld_(c0, y);
            collide_cell<MODEL, ST>(f, a);
            store_pop<ST>(d + c0 + y, f[0], a.scale);
        }
#pragma unroll
        for (int q = 1; q < 9; ++q) {
            store_pop<ST>(d + c0 + (size_t)q * L + y, f[q], a.scale);
#pragma unroll
            for (int j = 0; j < 3; ++j) w[q][0][j] = w[q][1][j], w[q][1][j] = w[q][2][j];
        }
    }
}

// NS steps of (streaming operator OP, collision MODEL) per sweep over device memory, FP64
// storage (temporal blocking as the lattice Boltzmann pass kernel): a wavefront owns 64
// consecutive rows (56 with output, NS halo rows on each side) and walks a strip of
// kMultiXS columns, keeping for every level k < NS the last three columns of its own row
// in registers; y-neighbours come from the adjacent lanes (DPP).  The columns are staged
// in LDS two ahead of the walk.  When level k receives a
// column, level k+1 gets the column before it: the operator's 3x3 stencil per direction,
// then the collision, exactly as k_stream_fd<OP, F64, MODEL> computes them.
template <int OP, int MODEL, int NS>
__global__ void __launch_bounds__(256)
    k_stream_fd_multi(StepArgs a, int waves_y, int nwaves, double dt, double p2s) {
    constexpr int VR = 56;
    static_assert(VR + 2 * NS <= 64, "halo rows do not fit in a wavefront");
    const int lane = threadIdx.x & 63;
    const int gw = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (gw >= nwaves) return;  // the whole wavefront
    const int nx = a.nx, ny = a.ny;
    const int wy = gw % waves_y, xsi = gw / waves_y;
    const int r = VR * wy - NS + lane;
    int y = r % ny;
    if (y < 0) y += ny;
    const bool out = lane >= NS && lane < VR + NS && r < ny;
    const int x0 = xsi * kMultiXS;
    const int ncols = min(kMultiXS, nx - x0);
    const size_t L = (size_t)a.ldd, cs = a.cs;
    const double *__restrict__ s = static_cast<const double *>(a.src);
    double *__restrict__ d = static_cast<double *>(a.dst);
    double *__restrict__ lag = static_cast<double *>(a.lag);  // wave-uniform
    auto column = [&](int c) {
        c %= nx;
        if (c < 0) c += nx;
        return (size_t)c * cs;
    };
    double W[NS][3][9];  // [level][newest column, -1, -2][q], own row
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int q = 0; q < 9; ++q) W[k][j][q] = 0.0;
    const int iters = ncols + 2 * NS;
    // the next columns are staged in LDS ahead of the walk (as in the LBM pass kernel),
    // so their latency overlaps the operator and collisions of the current column
    ColumnStage st{};
    st.init(stage_buffers(), s, L, ny, VR * wy - NS, lane);
#pragma unroll
    for (int j = 0; j < kStageAhead; ++j)
        if (j < iters) st.fetch(s, column(x0 - NS + j), L, j % kStageBufs);
    for (int i = 0; i < iters; ++i) {
        double f[9];
        st.wait(i, iters, 2 * NS, a.lag != nullptr);
        {
            const double *slot = st.column(i);
#pragma unroll
            for (int q = 0; q < 9; ++q) f[q] = slot[q * 64 + lane];
        }
        if (i + kStageAhead < iters)
            st.fetch(s, column(x0 - NS + i + kStageAhead), L, (i + kStageAhead) % kStageBufs);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            if (i < 2 * k) break;  // level k has no column yet
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                W[k][2][q] = W[k][1][q];
                W[k][1][q] = W[k][0][q];
                W[k][0][q] = f[q];
            }
            if (i < 2 * k + 2) break;  // level k+1 needs three columns of level k
            f[0] = W[k][1][0];
#pragma unroll
            for (int q = 1; q < 9; ++q) {
                const double fc = W[k][1][q], fe = W[k][0][q], fw = W[k][2][q];
                f[q] = fd_update<OP, double>(q, fc, fe, row_plus(fc), fw, row_minus(fc),
                                             row_plus(fe), row_plus(fw), row_minus(fw),
                                             row_minus(fe), dt, p2s);
            }
            collide_cell<MODEL, PLBM_F64>(f, a);
            if (k + 1 == NS && out) {
                const size_t co = (size_t)(x0 - 2 * NS + i) * cs;
#pragma unroll
                for (int q = 0; q < 9; ++q) __builtin_nontemporal_store(f[q], d + co + q * L + y);
            } else if (k + 2 == NS && lag != nullptr && i >= 2 * NS - 1 &&
                       i < 2 * NS - 1 + ncols && out) {
                // level NS-1 (the state one step before the pass's last) of column
                // x0 + i - (2 NS - 1)
                const size_t co = (size_t)(x0 - 2 * NS + 1 + i) * cs;
#pragma unroll
                for (int q = 0; q < 9; ++q)
                    __builtin_nontemporal_store(f[q], lag + co + q * L + y);
            }
        }
    }
}

}  // namespace plbm

using namespace plbm;

namespace {

StepArgs fd_args(const plbm_lattice *h, const void *src, void *dst) {
    StepArgs a{};
    a.src = src;
    a.dst = dst;
    a.nx = h->nx;
    a.ny = h->ny;
    a.ldd = h->ldd;
    a.nby = (h->ny + 255) / 256;
    a.cs = (size_t)9 * h->ldd;
    a.omega = h->omega;
    a.lambda_odd = h->lambda_odd;
    a.omega_f = float(h->omega);
    a.lambda_odd_f = float(h->lambda_odd);
    a.scale = float(h->scale);
    a.inv_scale = float(1.0 / h->scale);
    return a;
}

// model < 0: streaming only
int launch_fd(plbm_lattice *h, int op, int model, const void *src, void *dst, double dt) {
    const double p2s = 0.5 / std::sqrt(2.0);  // src/fvm_bardow.F90:712
    const StepArgs a = fd_args(h, src, dst);
    const int xs = strip_columns(h->nx, h->ny);
    const dim3 grid((unsigned)a.nby, (unsigned)((h->nx + xs - 1) / xs));
    return dispatch_storage(h->storage, [&](auto S) {
        constexpr int SS = decltype(S)::value;
        auto go = [&](auto O, auto M) {
            constexpr int OO = decltype(O)::value, MM = decltype(M)::value;
            if (xs == 32)
                hipLaunchKernelGGL((k_stream_fd<OO, SS, MM, 32>), grid, dim3(256), 0, h->stream, a,
                                   dt, p2s);
            else if (xs == 8)
                hipLaunchKernelGGL((k_stream_fd<OO, SS, MM, 8>), grid, dim3(256), 0, h->stream, a,
                                   dt, p2s);
            else
                hipLaunchKernelGGL((k_stream_fd<OO, SS, MM, 1>), grid, dim3(256), 0, h->stream, a,
                                   dt, p2s);
            return 0;
        };
        auto with_model = [&](auto O) {
            switch (model) {
            case PLBM_BGK: return go(O, std::integral_constant<int, PLBM_BGK>{});
            case PLBM_TRT: return go(O, std::integral_constant<int, PLBM_TRT>{});
            case PLBM_RR: return go(O, std::integral_constant<int, PLBM_RR>{});
            case PLBM_BGK_SPLIT: return go(O, std::integral_constant<int, PLBM_BGK_SPLIT>{});
            default: return go(O, std::integral_constant<int, -1>{});
            }
        };
        switch (op) {
        case PLBM_STREAM_FVM_BARDOW:
            with_model(std::integral_constant<int, PLBM_STREAM_FVM_BARDOW>{});
            break;
        case PLBM_STREAM_FDM_BARDOW:
            with_model(std::integral_constant<int, PLBM_STREAM_FDM_BARDOW>{});
            break;
        default:
            with_model(std::integral_constant<int, PLBM_STREAM_FDM_SOFONEA>{});
        }
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
}

int check_op_args(plbm_lattice *h, int op, int src, int dst, double dt) {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(src >= 0 && src < h->nf && dst >= 0 && dst < h->nf && src != dst, PLBM_ERR_ARG,
                 "need distinct valid src and dst buffers");
    PLBM_REQUIRE(std::isfinite(dt), PLBM_ERR_ARG, "non-finite dt");
    PLBM_REQUIRE(op >= PLBM_STREAM_EXACT && op <= PLBM_STREAM_FDM_SOFONEA, PLBM_ERR_ARG,
                 "unknown streaming operator " + std::to_string(op));
    return PLBM_OK;
}

int launch_fd_multi(plbm_lattice *h, int op, int model, int ns, const void *src, void *dst,
                    double dt, void *lag = nullptr) {
    const double p2s = 0.5 / std::sqrt(2.0);  // src/fvm_bardow.F90:712
    StepArgs a = fd_args(h, src, dst);
    a.lag = lag;
    const int waves_y = (h->ny + 55) / 56;
    const int nw = waves_y * ((h->nx + kMultiXS - 1) / kMultiXS);
    const dim3 grid((unsigned)((nw + 3) / 4));
    return dispatch_model(model, [&](auto M) {
        constexpr int MM = decltype(M)::value;
        auto go = [&](auto O) {
            constexpr int OO = decltype(O)::value;
            if (ns == 2)
                hipLaunchKernelGGL((k_stream_fd_multi<OO, MM, 2>), grid, dim3(256), 0, h->stream,
                                   a, waves_y, nw, dt, p2s);
            else
                hipLaunchKernelGGL((k_stream_fd_multi<OO, MM, 3>), grid, dim3(256), 0, h->stream,
                                   a, waves_y, nw, dt, p2s);
        };
        switch (op) {
        case PLBM_STREAM_FVM_BARDOW: go(std::integral_constant<int, PLBM_STREAM_FVM_BARDOW>{}); break;
        case PLBM_STREAM_FDM_BARDOW: go(std::integral_constant<int, PLBM_STREAM_FDM_BARDOW>{}); break;
        default: go(std::integral_constant<int, PLBM_STREAM_FDM_SOFONEA>{});
        }
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
}

}  // namespace

extern "C" int plbm_stream_op(plbm_lattice *h, int op, int src, int dst, double dt) {
    PLBM_TRY(check_op_args(h, op, src, dst, dt));
    if (op == PLBM_STREAM_EXACT) return plbm_stream(h, src, dst);
    return launch_fd(h, op, -1, h->f[src], h->f[dst], dt);
}

extern "C" int plbm_step_op(plbm_lattice *h, int op, int model, int src, int dst, double dt) {
    PLBM_TRY(check_op_args(h, op, src, dst, dt));
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    PLBM_REQUIRE(h->params_set, PLBM_ERR_STATE, "plbm_set_params has not been called");
    if (op == PLBM_STREAM_EXACT) return plbm_step(h, model, src, dst);
    return launch_fd(h, op, model, h->f[src], h->f[dst], dt);
}

// Steps per pass of plbm_run_op when the caller has not chosen, where the grid gives at
// least 2048 wavefronts (FP64 storage; else single steps): 3 for the Sofonea operator, 2
// for the 9-point Bardow operators, whose 3-step pass no longer fits two wavefronts per
// SIMD (measured at 8192^2, ms per step in passes of 1 / 2 / 3: FVM + BGK 2.06 / 1.71 /
// 1.89, FDM + TRT 2.02 / 1.05 / 1.30, Sofonea + BGK 1.94 / 1.00 / 0.69).
static int fd_steps_per_pass(const plbm_lattice *h, int op, int model) {
    (void)model;
    if (h->fd_ns > 0) return h->fd_ns;
    if (h->storage != PLBM_F64) return 1;
    const long waves = (long)((h->ny + 55) / 56) * ((h->nx + kMultiXS - 1) / kMultiXS);
    if (waves < 2048) return 1;
    return op == PLBM_STREAM_FDM_SOFONEA ? 3 : 2;
}

extern "C" int plbm_set_op_steps_per_pass(plbm_lattice *h, int steps) {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(steps >= 0 && steps <= 3, PLBM_ERR_ARG, "steps per pass must be 0..3");
    PLBM_REQUIRE(steps <= 1 || h->storage == PLBM_F64, PLBM_ERR_STATE,
                 "passes of the streaming operators need FP64 (PLBM_F64) storage");
    h->fd_ns = steps;
    return PLBM_OK;
}

extern "C" int plbm_get_op_steps_per_pass(const plbm_lattice *h, int op, int model,
                                          int *steps) {
    PLBM_REQUIRE(h && steps, PLBM_ERR_ARG, "null argument");
    *steps = fd_steps_per_pass(h, op, model);
    return PLBM_OK;
}

extern "C" int plbm_run_op(plbm_lattice *h, int op, int model, int nsteps, int *iold,
                           int *inew, double dt) {
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
    // passes of up to ns steps, evenly split, the deepest last; the last pass also stores
    // the state one step before the latest into the third buffer, which becomes the other
    // slot's storage (as after perform_step).  Without a third buffer the last step is a
    // single fused step instead.
    const int a0 = *iold, b0 = *inew;
    void *latest = h->f[a0], *other = h->f[b0];
    void *aux = pass_aux(h);
    const std::vector<int> plan = pass_plan(aux ? nsteps : nsteps - 1, ns);
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
}
