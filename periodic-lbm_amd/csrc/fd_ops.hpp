// fd_ops.hpp — the finite-volume / finite-difference streaming operators' kernels
// (src/fvm_bardow.F90:393-893) and their launchers, templated on the operator id, so that
// the operators can be instantiated in more than one translation unit (streaming_ops.hip:
// the default build's three operators; streaming_ops_fdm.hip: the -DFDM_* stencils).
#pragma once

#include <cmath>

#include <algorithm>
#include <utility>

#include "common.hpp"
#include "d2q9.hpp"
#include "multistep.hpp"
#include "stepper.hpp"

namespace plbm {

// stream_fdm_bardow's stencil family: the 9-point first and second derivatives
// (src/fvm_bardow.F90:591-669).  PLBM_STREAM_FDM_BARDOW is the default build's central
// differences; the others are the variants the reference selects at compile time with
// -DFDM_WLS (weighted least squares, no weight), -DFDM_WLS_GAUSS_V1 / _V2 (Gaussian
// weights scaled to the farthest / closest node) and -DFDM_ISO (isotropic differences,
// Kumar 2004).  Each expression is grouped as Fortran evaluates it (left to right), with
// the reference's constants, so that FP64 results are bit-identical to those builds.
__host__ __device__ constexpr bool is_fdm_bardow(int op) {
    return op == PLBM_STREAM_FDM_BARDOW ||
           (op >= PLBM_STREAM_FDM_BARDOW_WLS && op <= PLBM_STREAM_FDM_BARDOW_ISO);
}

template <int OP, class R>
__device__ __forceinline__ void fdm_derivatives(R fc, R fe, R fn, R fw, R fs, R fne, R fnw,
                                                R fsw, R fse, R &dfx, R &dfy, R &dfxx, R &dfyy,
                                                R &dfxy) {
#pragma clang fp contract(off)
    const R two = R(2);
    if constexpr (OP == PLBM_STREAM_FDM_BARDOW) {
        const R p2 = R(0.5);
        dfx = p2 * (fe - fw);
        dfy = p2 * (fn - fs);
        dfxx = (fe - two * fc) + fw;
        dfyy = (fn - two * fc) + fs;
        dfxy = R(0.25) * (((fne - fse) - fnw) + fsw);
    } else if constexpr (OP == PLBM_STREAM_FDM_BARDOW_WLS) {
        const R one_sixth = R(1.0 / 6.0), one_third = R(1.0 / 3.0);
        dfx = one_sixth * (((fne - fnw) + (fe - fw)) + (fse - fsw));
        dfy = one_sixth * (((fne - fse) + (fn - fs)) + (fnw - fsw));
        dfxx = (one_third * ((fne - two * fn) + fnw) + one_third * ((fe - two * fc) + fw)) +
               one_third * ((fse - two * fs) + fsw);
        dfyy = (one_third * ((fne - two * fe) + fse) + one_third * ((fn - two * fc) + fs)) +
               one_third * ((fnw - two * fw) + fsw);
        dfxy = R(0.25) * (((fne - fnw) + fsw) - fse);
    } else if constexpr (OP == PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1 ||
                         OP == PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2) {
        constexpr bool v1 = OP == PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1;
        const R p1s = v1 ? R(0.2880584423829145035434) : R(0.3934930210807994210853);
        const R p1d = v1 ? R(0.1059707788085427065949) : R(0.05325348945960039354075);
        const R p2c = v1 ? R(-1.152233769531658458263) : R(-1.573972084323197018207);
        const R p2d1 = v1 ? R(0.5761168847658292291314) : R(0.7869860421615988421706);
        const R p2d2 = v1 ? R(-0.4238831152341712149578) : R(-0.2130139578384016019186);
        const R p2d = v1 ? R(0.2119415576170855242122) : R(0.1065069789192007732037);
        dfx = (p1s * (fe - fw) + p1d * (fne - fnw)) + p1d * (fse - fsw);
        dfy = (p1s * (fn - fs) + p1d * (fne - fse)) + p1d * (fnw - fsw);
        const R corners = ((fne + fnw) + fsw) + fse;
        dfxx = ((p2c * fc + p2d1 * (fe + fw)) + p2d2 * (fn + fs)) + p2d * corners;
        dfyy = ((p2c * fc + p2d2 * (fe + fw)) + p2d1 * (fn + fs)) + p2d * corners;
        dfxy = R(0.25) * (((fne - fnw) + fsw) - fse);
    } else {
        static_assert(OP == PLBM_STREAM_FDM_BARDOW_ISO, "not an FDM Bardow stencil");
        const R p2 = R(0.5), one_sixth = R(1.0 / 6.0), two_thirds = R(2.0 / 3.0);
        const R five_sixths = R(10.0 / 12.0), one_twelfth = R(1.0 / 12.0);
        dfx = p2 * ((one_sixth * (fne - fnw) + two_thirds * (fe - fw)) + one_sixth * (fse - fsw));
        dfy = p2 * ((one_sixth * (fne - fse) + two_thirds * (fn - fs)) + one_sixth * (fnw - fsw));
        dfxx = (one_twelfth * ((fne - two * fn) + fnw) + five_sixths * ((fe - two * fc) + fw)) +
               one_twelfth * ((fse - two * fs) + fsw);
        dfyy = (one_twelfth * ((fne - two * fe) + fse) + five_sixths * ((fn - two * fc) + fs)) +
               one_twelfth * ((fnw - two * fw) + fsw);
        dfxy = R(0.25) * (((fne - fse) - fnw) + fsw);
    }
}

// One direction's update from its 3x3 neighbourhood (c = centre, e/n/w/s and corners).
template <int OP, class R>
__device__ __forceinline__ R fd_update(int q, R fc, R fe, R fn, R fw, R fs, R fne, R fnw, R fsw,
                                       R fse, R dtr, R p2s) {
#pragma clang fp contract(off)
    if constexpr (OP == PLBM_STREAM_FVM_BARDOW || is_fdm_bardow(OP)) {
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
            const R cxxq = (R(0.5) * cxq) * cxq, cyyq = (R(0.5) * cyq) * cyq, cxyq = cxq * cyq;
            R dfx, dfy, dfxx, dfyy, dfxy;
            fdm_derivatives<OP, R>(fc, fe, fn, fw, fs, fne, fnw, fsw, fse, dfx, dfy, dfxx, dfyy,
                                   dfxy);
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
// Two wavefronts per SIMD are requested: the 3-step passes of the 9-point Bardow
// operators with TRT / RR / split BGK need 258-264 registers and would otherwise run one
// wavefront per SIMD; held to 256 they spill 2-8 registers and run faster (8192^2, ms per
// step in passes of 3, one / two wavefronts per SIMD: FDM + TRT 1.31 / 1.04, FVM + TRT
// 1.80 / 1.35).
template <int OP, int MODEL, int NS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
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

inline StepArgs fd_args(const plbm_lattice *h, const void *src, void *dst) {
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

// One launch of operator OP (+ collision `model`; model < 0: streaming only), src -> dst.
template <int OP>
int launch_fd_op(plbm_lattice *h, int model, const void *src, void *dst, double dt) {
    const double p2s = 0.5 / std::sqrt(2.0);  // src/fvm_bardow.F90:712
    const StepArgs a = fd_args(h, src, dst);
    const int xs = strip_columns(h->nx, h->ny);
    const dim3 grid((unsigned)a.nby, (unsigned)((h->nx + xs - 1) / xs));
    return dispatch_storage(h->storage, [&](auto S) {
        constexpr int SS = decltype(S)::value;
        auto go = [&](auto M) {
            constexpr int MM = decltype(M)::value;
            if (xs == 32)
                hipLaunchKernelGGL((k_stream_fd<OP, SS, MM, 32>), grid, dim3(256), 0, h->stream, a,
                                   dt, p2s);
            else if (xs == 8)
                hipLaunchKernelGGL((k_stream_fd<OP, SS, MM, 8>), grid, dim3(256), 0, h->stream, a,
                                   dt, p2s);
            else
                hipLaunchKernelGGL((k_stream_fd<OP, SS, MM, 1>), grid, dim3(256), 0, h->stream, a,
                                   dt, p2s);
            return 0;
        };
        switch (model) {
        case PLBM_BGK: go(std::integral_constant<int, PLBM_BGK>{}); break;
        case PLBM_TRT: go(std::integral_constant<int, PLBM_TRT>{}); break;
        case PLBM_RR: go(std::integral_constant<int, PLBM_RR>{}); break;
        case PLBM_BGK_SPLIT: go(std::integral_constant<int, PLBM_BGK_SPLIT>{}); break;
        default: go(std::integral_constant<int, -1>{});
        }
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
}

// One temporally blocked pass of ns (2..3) steps of (OP, model), FP64 storage; with lag
// non-null the pass also stores the state one step before its last there.
template <int OP>
int launch_fd_multi_op(plbm_lattice *h, int model, int ns, const void *src, void *dst,
                       double dt, void *lag) {
    const double p2s = 0.5 / std::sqrt(2.0);  // src/fvm_bardow.F90:712
    StepArgs a = fd_args(h, src, dst);
    a.lag = lag;
    const int waves_y = (h->ny + 55) / 56;
    const int nw = waves_y * ((h->nx + kMultiXS - 1) / kMultiXS);
    const dim3 grid((unsigned)((nw + 3) / 4));
    return dispatch_model(model, [&](auto M) {
        constexpr int MM = decltype(M)::value;
        if (ns == 2)
            hipLaunchKernelGGL((k_stream_fd_multi<OP, MM, 2>), grid, dim3(256), 0, h->stream, a,
                               waves_y, nw, dt, p2s);
        else
            hipLaunchKernelGGL((k_stream_fd_multi<OP, MM, 3>), grid, dim3(256), 0, h->stream, a,
                               waves_y, nw, dt, p2s);
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
}

// The -DFDM_* stencils (streaming_ops_fdm.hip): op in PLBM_STREAM_FDM_BARDOW_WLS ..
// PLBM_STREAM_FDM_BARDOW_ISO.
int launch_fd_variant(plbm_lattice *h, int op, int model, const void *src, void *dst, double dt);
int launch_fd_multi_variant(plbm_lattice *h, int op, int model, int ns, const void *src,
                            void *dst, double dt, void *lag);

}  // namespace plbm
