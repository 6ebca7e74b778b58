// reduce.hpp — deterministic FP64 reductions of the lattice macros (mass, kinetic energy,
// max |u|, non-finite count, error norms against an analytic field), shared by the
// single-lattice and batched engines.
//
// A thread accumulates its cells in a fixed order; a wavefront combines its 64 lanes by a
// butterfly of lane shuffles (xor 32, 16, .., 1: every lane ends with the same value, since
// each step adds two operands that the partner lane adds in the other order, and IEEE
// addition commutes), the wavefronts of a workgroup through one LDS stage in wavefront
// order, and the workgroups' partials in a final pass of the same form.  The result does
// not depend on timing: the same inputs give the same bits on every call.
//
// The app's norms (app/main_taylor_green.f90:151-158, 174-212) are evaluated from these
// slots: max |u| = PLBM_RED_UMAX, relative L2 = sqrt(PLBM_RED_ERR2 / PLBM_RED_REF2).
#pragma once

#include "common.hpp"

namespace plbm {

constexpr int kRedThreads = 256;  // threads per reduction workgroup (4 wavefronts)

__device__ __forceinline__ double red_combine(int k, double a, double b) {
    return k == PLBM_RED_UMAX ? fmax(a, b) : a + b;
}

// Adds cell (rho, u) and, when uxa is non-null, its analytic velocity uxa[i], uya[i].
__device__ __forceinline__ void red_cell(double acc[PLBM_NRED], double rho, double ux,
                                         double uy, const double *uxa, const double *uya,
                                         size_t i) {
    if (!(__builtin_isfinite(rho) && __builtin_isfinite(ux) && __builtin_isfinite(uy))) {
        acc[PLBM_RED_NONFINITE] += 1.0;
        return;
    }
    const double u2 = ux * ux + uy * uy;
    acc[PLBM_RED_MASS] += rho;
    acc[PLBM_RED_KE] += 0.5 * rho * u2;
    acc[PLBM_RED_U2] += u2;
    acc[PLBM_RED_UMAX] = fmax(acc[PLBM_RED_UMAX], sqrt(u2));
    if (uxa) {
        const double ax = uxa[i], ay = uya[i];
        const double ex = ux - ax, ey = uy - ay;
        acc[PLBM_RED_ERR2] += ex * ex + ey * ey;
        acc[PLBM_RED_REF2] += ax * ax + ay * ay;
    }
}

// Butterfly over the 64 lanes of a wavefront; every lane returns the wavefront's value.
__device__ __forceinline__ void red_wave(double acc[PLBM_NRED]) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int k = 0; k < PLBM_NRED; ++k) acc[k] = red_combine(k, acc[k], __shfl_xor(acc[k], m, 64));
}

// Workgroup of kRedThreads threads: wavefront values through LDS, combined in wavefront
// order by threads 0 .. PLBM_NRED-1; thread k (< PLBM_NRED) returns slot k.
__device__ __forceinline__ double red_block(double acc[PLBM_NRED]) {
    constexpr int kWaves = kRedThreads / 64;
    __shared__ double sh[kWaves][PLBM_NRED];
    red_wave(acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < PLBM_NRED; ++k) sh[wave][k] = acc[k];
    __syncthreads();
    double v = 0.0;
    if (threadIdx.x < PLBM_NRED) {
        const int k = threadIdx.x;
        v = sh[0][k];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) v = red_combine(k, v, sh[w][k]);
    }
    return v;
}

// Walk of the cells (x, y) of an nx x ny grid with linear index i = x*ny + y, starting at
// i0 and advancing by `stride` (= sx*ny + sy, 0 <= sy < ny, split on the host): no
// division per cell.
struct CellWalk {
    int x, y;
    __device__ __forceinline__ void start(long long i0, int ny) {
        x = (int)(i0 / ny);
        y = (int)(i0 - (long long)x * ny);
    }
    __device__ __forceinline__ void next(int sx, int sy, int ny) {
        x += sx;
        y += sy;
        if (y >= ny) {
            y -= ny;
            ++x;
        }
    }
};

}  // namespace plbm
