// multistep.hpp — one wavefront's share of a temporally blocked pass (NS fused stream +
// collide steps per sweep over HBM), shared by the single-lattice pass kernel
// (multistep.hip) and the batched engine (batch.hip).  See multistep.hip for the scheme.
#pragma once

#include "common.hpp"
#include "d2q9.hpp"
#include "stepper.hpp"

namespace plbm {

constexpr int kMultiXS = 64;  // columns per strip

// Rows with output per wave: 56 for 8-byte storage (448-byte aligned runs), 48 for 4- and
// 2-byte storage (192 / 96 bytes, 64- / 32-byte aligned).  The halo of a pass of NS steps
// is NS rows on each side, so a pass can hold (64 - VR) / 2 steps: 4 for FP64 storage,
// 8 for FP32 and FP16 storage (whose registers are half as wide).
template <int ST>
constexpr int multi_vr() {
    return sizeof(typename Traits<ST>::T) == 8 ? 56 : 48;
}
template <int ST>
constexpr int max_steps_per_pass() {
    return (64 - multi_vr<ST>()) / 2;
}

// Row y+1 / y-1 of a value held by every lane of a wavefront whose lane i holds row
// base + i: DPP wave_shl:1 gives lane i the value of lane i+1, wave_shr:1 that of lane
// i-1 (the edge lanes keep their own value; they are halo lanes of a pass).
__device__ __forceinline__ double row_plus(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, 0x130, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), 0x130, 0xF, 0xF,
                                               false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double row_minus(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, 0x138, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), 0x138, 0xF, 0xF,
                                               false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// value at row y - cy_q: cy = +1 comes from lane - 1, cy = -1 from lane + 1
template <class R>
__device__ __forceinline__ R pull_y(int q, R v) {
    if (cy(q) > 0) return __shfl_up(v, 1, 64);
    if (cy(q) < 0) return __shfl_down(v, 1, 64);
    return v;
}

// The calling wavefront advances rows [VR*wy, VR*wy + VR) x columns [64*xsi, 64*xsi + 64)
// of the grid described by `a` (clipped to the grid) NS steps, a.src -> a.dst.  All 64
// lanes must call it together (lane shuffles).
template <int MODEL, int ST, int NS>
__device__ __forceinline__ void multi_wave(const StepArgs &a, int wy, int xsi) {
    using T = typename Traits<ST>::T;
    using R = typename Traits<ST>::R;
    constexpr int VR = multi_vr<ST>();
    static_assert(NS >= 1 && VR + 2 * NS <= 64, "halo rows do not fit in a wavefront");
    const int lane = threadIdx.x & 63;
    const int nx = a.nx, ny = a.ny;
    const int r = VR * wy - NS + lane;  // unwrapped row of this lane
    int y = r % ny;
    if (y < 0) y += ny;
    const bool out = lane >= NS && lane < VR + NS && r < ny;
    const int x0 = xsi * kMultiXS;
    const int ncols = min(kMultiXS, nx - x0);
    const size_t L = (size_t)a.ldd, cs = a.cs;
    const T *__restrict__ s = static_cast<const T *>(a.src);
    T *__restrict__ d = static_cast<T *>(a.dst);
    R W[NS][3][9];  // [level][newest column, -1, -2][q]
#pragma unroll
    for (int k = 0; k < NS; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int q = 0; q < 9; ++q) W[k][j][q] = R(0);
    auto column = [&](int c) {
        c %= nx;
        if (c < 0) c += nx;
        return (size_t)c * cs;
    };
    // the next column's loads are issued one iteration ahead, so that their latency
    // overlaps the collisions of the current column
    T nxt[9];
    {
        const size_t cc = column(x0 - NS);
#pragma unroll
        for (int q = 0; q < 9; ++q) nxt[q] = s[cc + q * L + y];
    }
    const int iters = ncols + 2 * NS;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            W[0][2][q] = W[0][1][q];
            W[0][1][q] = W[0][0][q];
            W[0][0][q] = pull_y(q, load_pop<ST>(&nxt[q], a.inv_scale));
        }
        if (i + 1 < iters) {
            const size_t cn = column(x0 - NS + i + 1);
#pragma unroll
            for (int q = 0; q < 9; ++q) nxt[q] = s[cn + q * L + y];
        }
#pragma unroll
        for (int k = 1; k <= NS; ++k) {
            if (i < 2 * k) break;
            R f[9];
#pragma unroll
            for (int q = 0; q < 9; ++q)
                f[q] = cx(q) > 0 ? W[k - 1][2][q] : (cx(q) == 0 ? W[k - 1][1][q] : W[k - 1][0][q]);
            collide_cell<MODEL, ST>(f, a);
            if (k < NS) {
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    R v = f[q];
                    if constexpr (ST == PLBM_F16S) v = R(T(v * a.scale)) * a.inv_scale;
                    W[k][2][q] = W[k][1][q];
                    W[k][1][q] = W[k][0][q];
                    W[k][0][q] = pull_y(q, v);
                }
            } else if (out) {
                const size_t co = (size_t)(x0 - 2 * NS + i) * cs;
#pragma unroll
                for (int q = 0; q < 9; ++q) store_pop<ST>(d + co + q * L + y, f[q], a.scale);
            }
        }
    }
}

}  // namespace plbm
