// streaming_ops.hip — the reference's finite-volume / finite-difference streaming
// operators on the device, for the grid%streaming slot (src/fvm_bardow.F90:393-893):
//   PLBM_STREAM_FVM_BARDOW   stream_fvm_bardow  (bilinear face fluxes, :410-509)
//   PLBM_STREAM_FDM_BARDOW   stream_fdm_bardow  (central-difference Lax-Wendroff, default
//                                                stencil without FDM_* options, :526-685)
//   PLBM_STREAM_FDM_SOFONEA  stream_fdm_sofonea (1-D Lax-Wendroff along c_q, :702-889)
// Each is a 9-point (3-point for Sofonea) stencil on one q-segment of the column-
// interleaved layout, one thread per cell looping over q; the rest population is copied.
// The operators are linear and leave constants unchanged, so they act on shifted and
// scaled storage directly.  FP64 expression trees follow the reference with FMA
// contraction off, which makes them bit-identical to it.  The reference loops x over ny
// and y over nx (square grids); here x runs over nx and y over ny.
#include <cmath>

#include "common.hpp"
#include "d2q9.hpp"
#include "stepper.hpp"

namespace plbm {

template <int OP, int ST>
__global__ void __launch_bounds__(256)
    k_stream_fd(const void *src, void *dst, int nx, int ny, int ldd, double dt, double p2s,
                float inv_scale, float scale) {
#pragma clang fp contract(off)
    using T = typename Traits<ST>::T;
    using R = typename Traits<ST>::R;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= ny) return;
    const T *s = static_cast<const T *>(src);
    T *d = static_cast<T *>(dst);
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const size_t L = (size_t)ldd, cs = 9 * L;
    const size_t c0 = (size_t)x * cs, cm = (size_t)xm * cs, cp = (size_t)xp * cs;
    d[c0 + y] = s[c0 + y];  // rest population (stored values, no conversion)
    const R dtr = R(dt);
#pragma unroll
    for (int q = 1; q < 9; ++q) {
        const size_t o = (size_t)q * L;
        auto ld_ = [&](size_t col, int yy) { return load_pop<ST>(s + col + o + yy, inv_scale); };
        R out;
        if constexpr (OP == PLBM_STREAM_FVM_BARDOW || OP == PLBM_STREAM_FDM_BARDOW) {
            const R fc = ld_(c0, y), fe = ld_(cp, y), fn = ld_(c0, yp), fw = ld_(cm, y),
                    fs = ld_(c0, ym), fne = ld_(cp, yp), fnw = ld_(cm, yp), fsw = ld_(cm, ym),
                    fse = ld_(cp, ym);
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
                out = (fc - cxq * (cfe - cfw)) - cyq * (cfn - cfs);
            } else {
                const R p2 = R(0.5);
                const R cxxq = (R(0.5) * cxq) * cxq, cyyq = (R(0.5) * cyq) * cyq,
                        cxyq = cxq * cyq;
                const R dfx = p2 * (fe - fw), dfy = p2 * (fn - fs);
                const R dfxx = (fe - R(2) * fc) + fw, dfyy = (fn - R(2) * fc) + fs;
                const R dfxy = R(0.25) * (((fne - fse) - fnw) + fsw);
                out = ((fc - cxq * dfx) - cyq * dfy) + ((cxxq * dfxx + cxyq * dfxy) + cyyq * dfyy);
            }
        } else {
            // fu at x + c_q, fd at x - c_q
            const size_t cu = cx(q) > 0 ? cp : (cx(q) < 0 ? cm : c0);
            const size_t cd = cx(q) > 0 ? cm : (cx(q) < 0 ? cp : c0);
            const int yu = cy(q) > 0 ? yp : (cy(q) < 0 ? ym : y);
            const int yd = cy(q) > 0 ? ym : (cy(q) < 0 ? yp : y);
            const R fc = ld_(c0, y), fu = ld_(cu, yu), fd = ld_(cd, yd);
            R du1, du2;
            if (q < 5) {
                du1 = R(0.5) * (fu - fd);
                du2 = (fu - R(2) * fc) + fd;
            } else {
                du1 = R(p2s) * (fu - fd);
                du2 = R(0.5) * ((fu - R(2) * fc) + fd);
            }
            out = fc + dtr * ((R(0.5) * dtr) * du2 - du1);
        }
        store_pop<ST, false>(d + c0 + o + y, out, scale);
    }
}

}  // namespace plbm

using namespace plbm;

extern "C" int plbm_stream_op(plbm_lattice *h, int op, int src, int dst, double dt) {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(src >= 0 && src < h->nf && dst >= 0 && dst < h->nf && src != dst, PLBM_ERR_ARG,
                 "need distinct valid src and dst buffers");
    PLBM_REQUIRE(std::isfinite(dt), PLBM_ERR_ARG, "non-finite dt");
    if (op == PLBM_STREAM_EXACT) return plbm_stream(h, src, dst);
    PLBM_REQUIRE(op >= PLBM_STREAM_FVM_BARDOW && op <= PLBM_STREAM_FDM_SOFONEA, PLBM_ERR_ARG,
                 "unknown streaming operator " + std::to_string(op));
    const double p2s = 0.5 / std::sqrt(2.0);  // src/fvm_bardow.F90:712
    const dim3 grid((unsigned)((h->ny + 255) / 256), (unsigned)h->nx);
    return dispatch_storage(h->storage, [&](auto S) {
        constexpr int SS = decltype(S)::value;
        const float is = float(1.0 / h->scale), sc = float(h->scale);
        switch (op) {
        case PLBM_STREAM_FVM_BARDOW:
            hipLaunchKernelGGL((k_stream_fd<PLBM_STREAM_FVM_BARDOW, SS>), grid, dim3(256), 0,
                               h->stream, h->f[src], h->f[dst], h->nx, h->ny, h->ldd, dt, p2s, is,
                               sc);
            break;
        case PLBM_STREAM_FDM_BARDOW:
            hipLaunchKernelGGL((k_stream_fd<PLBM_STREAM_FDM_BARDOW, SS>), grid, dim3(256), 0,
                               h->stream, h->f[src], h->f[dst], h->nx, h->ny, h->ldd, dt, p2s, is,
                               sc);
            break;
        default:
            hipLaunchKernelGGL((k_stream_fd<PLBM_STREAM_FDM_SOFONEA, SS>), grid, dim3(256), 0,
                               h->stream, h->f[src], h->f[dst], h->nx, h->ny, h->ldd, dt, p2s, is,
                               sc);
        }
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
}
