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

#include "common.hpp"
#include "d2q9.hpp"
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
int launch_fd(plbm_lattice *h, int op, int model, int src, int dst, double dt) {
    const double p2s = 0.5 / std::sqrt(2.0);  // src/fvm_bardow.F90:712
    const StepArgs a = fd_args(h, h->f[src], h->f[dst]);
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

}  // namespace

extern "C" int plbm_stream_op(plbm_lattice *h, int op, int src, int dst, double dt) {
    PLBM_TRY(check_op_args(h, op, src, dst, dt));
    if (op == PLBM_STREAM_EXACT) return plbm_stream(h, src, dst);
    return launch_fd(h, op, -1, src, dst, dt);
}

extern "C" int plbm_step_op(plbm_lattice *h, int op, int model, int src, int dst, double dt) {
    PLBM_TRY(check_op_args(h, op, src, dst, dt));
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    PLBM_REQUIRE(h->params_set, PLBM_ERR_STATE, "plbm_set_params has not been called");
    if (op == PLBM_STREAM_EXACT) return plbm_step(h, model, src, dst);
    return launch_fd(h, op, model, src, dst, dt);
}
