// stepper.hpp — device helpers shared by the single-lattice and batched engines:
// XCD-aware workgroup renumbering, storage load/store, per-cell collision dispatch.
#pragma once

#include "common.hpp"
#include "d2q9.hpp"

namespace plbm {

// Workgroup b of G is dispatched to XCD b % 8.  Give each XCD a contiguous range of
// logical tiles so neighbouring y-tiles of a column (which share the cache line that a
// y+-1 shifted load straddles) sit in the same L2.
__device__ __forceinline__ int xcd_remap(int b, int G) {
    constexpr int NX = 8;
    const int q = G / NX, r = G % NX;
    const int xcd = b % NX, idx = b / NX;
    const int base = xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
    return base + idx;
}

template <int ST>
__device__ __forceinline__ typename Traits<ST>::R load_pop(const typename Traits<ST>::T *p,
                                                           float inv_scale) {
    using R = typename Traits<ST>::R;
    if constexpr (ST == PLBM_F16S)
        return R(*p) * inv_scale;
    else
        return R(*p);
}

// Streaming kernels write every population exactly once per step and read it back only
// in the next step, after the whole grid has been swept: non-temporal stores keep that
// traffic from evicting the lines the current step still reads.  Kernels that keep a
// small lattice resident in L2 across steps pass NT = false.
template <int ST, bool NT = true>
__device__ __forceinline__ void store_pop(typename Traits<ST>::T *p, typename Traits<ST>::R v,
                                          float scale) {
    using T = typename Traits<ST>::T;
    const T t = ST == PLBM_F16S ? T(v * scale) : T(v);
    if constexpr (NT)
        __builtin_nontemporal_store(t, p);
    else
        *p = t;
}

template <int MODEL, int ST>
__device__ __forceinline__ void collide_cell(typename Traits<ST>::R f[9], const StepArgs &a) {
    using R = typename Traits<ST>::R;
    if constexpr (!Traits<ST>::shifted) {
        if constexpr (MODEL == PLBM_BGK) collide_bgk_ref(f, a.omega);
        else if constexpr (MODEL == PLBM_BGK_SPLIT) collide_bgk_split_ref(f, a.omega);
        else if constexpr (MODEL == PLBM_TRT) collide_trt_ref(f, a.omega, a.lambda_odd);
        else collide_rr_ref(f, a.omega);
    } else {
        const R om = sizeof(R) == 8 ? R(a.omega) : R(a.omega_f);
        const R lo = sizeof(R) == 8 ? R(a.lambda_odd) : R(a.lambda_odd_f);
        if constexpr (sizeof(R) == 4 && (MODEL == PLBM_BGK || MODEL == PLBM_BGK_SPLIT)) {
            collide_bgk_shifted_conserving<R>(f, om);
        } else if constexpr (sizeof(R) == 4 && MODEL == PLBM_TRT) {
            collide_trt_shifted_conserving<R>(f, om, lo);
        } else if constexpr (sizeof(R) == 4) {
            R d0, jx0, jy0;
            moments_shifted(f, d0, jx0, jy0);
            collide_rr_shifted<R>(f, om);
            close_balances(f, d0, jx0, jy0);
        } else if constexpr (MODEL == PLBM_BGK || MODEL == PLBM_BGK_SPLIT) {
            collide_bgk_shifted<R>(f, om);
        } else if constexpr (MODEL == PLBM_TRT) {
            collide_trt_shifted<R>(f, om, lo);
        } else {
            collide_rr_shifted<R>(f, om);
        }
    }
}

}  // namespace plbm
