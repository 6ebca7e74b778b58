// adjoint.hip — discrete adjoint of the BGK stream-collide step with checkpointing.
//
// Forward map (one step n):  g^n = S f^n (pull streaming),  f^{n+1} = C(g^n) =
// (1-omega) g^n + omega feq(rho(g^n), u(g^n)).  Initial state f^0 = feq(rho0, u0); loss
// L = 1/2 sum_cells |u(f^N) - u*|^2 with u = j/rho.
//
// Reverse sweep: lambda^N = dL/df^N; for n = N-1 .. 0
//   mu      = J_C(g^n)^T lambda^{n+1}          (collision Jacobian at the pre-collision state)
//   lambda^n_q(x - c_q) = mu_q(x)              (S^T: push streaming)
//   dL/domega += sum_cells sum_q lambda^{n+1}_q (feq_q(g^n) - g^n_q)
// and dL/du0, dL/drho0 = (d feq / d(rho, u))^T lambda^0.  One fused kernel does one reverse
// step: it re-reads f^n with the forward's pull pattern, evaluates the Jacobian-transpose
// product in registers and pushes mu to the neighbours.
//
// Memory: states are checkpointed every K steps during the forward sweep; the reverse
// sweep recomputes each K-step segment from its checkpoint into a segment buffer, then
// runs the K reverse steps.  Forward cost 2N - N/K steps, reverse N steps.
//
// For BGK with rho = sum g, u = sum c g / rho, feq_i = w_i rho (1 + P_i),
// P_i = 3 c_i.u + 4.5 (c_i.u)^2 - 1.5 u.u:
//   d feq_i / d rho = w_i (1 + P_i),  d feq_i / d u_a = w_i rho (3 c_ia + 9 (c_i.u) c_ia - 3 u_a)
//   d rho / d g_j = 1,  d u_a / d g_j = (c_ja - u_a) / rho
//   mu_j = (1-omega) lambda_j + omega (A + sum_a B_a (c_ja - u_a) / rho),
//   A = sum_i lambda_i w_i (1 + P_i),  B_a = sum_i lambda_i d feq_i / d u_a.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "d2q9.hpp"
#include "stepper.hpp"

namespace plbm {

int launch_step(plbm_lattice *h, int model, const void *src, void *dst);
constexpr int kAdjBlocks = 2048;  // reduction partial slots (fixed => deterministic)

template <class R>
__device__ __forceinline__ void bgk_jacobian_t(const R g_unshifted_rho, const R ux, const R uy,
                                               const R lam[9], R omega, R mu[9]) {
    const R rho = g_unshifted_rho;
    const R u2 = ux * ux + uy * uy;
    R A = 0, Bx = 0, By = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const R cu = R(cx(i)) * ux + R(cy(i)) * uy;
        const R P = R(3) * cu + R(4.5) * cu * cu - R(1.5) * u2;
        const R lw = lam[i] * R(wq(i));
        A += lw * (R(1) + P);
        const R t = R(3) + R(9) * cu;
        Bx += lw * (t * R(cx(i)) - R(3) * ux);
        By += lw * (t * R(cy(i)) - R(3) * uy);
    }
    Bx *= rho;
    By *= rho;
    const R inv = R(1) / rho;
#pragma unroll
    for (int j = 0; j < 9; ++j)
        mu[j] = (R(1) - omega) * lam[j] +
                omega * (A + (Bx * (R(cx(j)) - ux) + By * (R(cy(j)) - uy)) * inv);
}

// Macros of one stored cell (pulled or not) from its populations g (deviations
// g_q = f_q - w_q for the shifted storages), in the adjoint's FP64 arithmetic.
template <bool SHIFTED>
__device__ __forceinline__ void cell_state(const double gs[9], double &rho, double &ux,
                                           double &uy, double &drho) {
    double d, jx, jy;
    moments_shifted<double>(gs, d, jx, jy);  // plain sums: valid for shifted and unshifted
    if constexpr (SHIFTED) {
        drho = d;
        rho = 1.0 + d;
    } else {
        rho = d;
        drho = d - 1.0;
    }
    ux = jx / rho;
    uy = jy / rho;
}

// The adjoint state lambda, mu and the omega sensitivity are FP64 for every storage: the
// forward states of FP32 storage convert to FP64 exactly, so the reverse sweep is the
// exact adjoint of the FP64 step evaluated along the FP32 trajectory, and its own
// rounding (500 reverse steps of FP32 would add about 1e-5 relative to the gradient)
// drops out.

// lambda^N and the loss of the final state.
template <int ST>
__global__ void __launch_bounds__(256)
    k_adj_final(const void *fN, int nx, int ny, int ldd, float inv_scale, const double *uxt,
                const double *uyt, double *lam, double *partials) {
    using T = typename Traits<ST>::T;
    __shared__ double sh[256];
    double acc = 0.0;
    const size_t n = (size_t)nx * ny;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ny), y = (int)(i % ny);
        const T *s = static_cast<const T *>(fN);
        const size_t c = pop_index(0, x, y, ldd);
        double g[9];
#pragma unroll
        for (int q = 0; q < 9; ++q)
            g[q] = double(load_pop<ST>(s + c + (size_t)q * ldd, inv_scale));
        double rho, ux, uy, drho;
        cell_state<Traits<ST>::shifted>(g, rho, ux, uy, drho);
        const double ex = ux - uxt[i], ey = uy - uyt[i];
        acc += 0.5 * (ex * ex + ey * ey);
        const double inv = 1.0 / rho;
#pragma unroll
        for (int q = 0; q < 9; ++q)
            lam[c + (size_t)q * ldd] = (ex * (cx(q) - ux) + ey * (cy(q) - uy)) * inv;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// One reverse step: f^n (storage), lambda^{n+1} -> lambda^n, and the omega sensitivity
// of this step accumulated into partials[blockIdx.x] (same slot every step: launches on
// one stream are ordered, so the sum is race-free and deterministic).
template <int ST>
__global__ void __launch_bounds__(256)
    k_adj_step(const void *fn, const double *__restrict__ lam_in, double *__restrict__ lam_out,
               int nx, int ny, int ldd, double omega, float inv_scale, double *partials) {
    using T = typename Traits<ST>::T;
    __shared__ double sh[256];
    double acc = 0.0;
    const size_t n = (size_t)nx * ny;
    const T *s = static_cast<const T *>(fn);
    const size_t L = (size_t)ldd, cs = 9 * L;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ny), y = (int)(i % ny);
        const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
        const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
        const size_t c0 = (size_t)x * cs, cm = (size_t)xm * cs, cp = (size_t)xp * cs;
        // g^n = S f^n at cell (x, y): the forward's pull pattern
        double g[9];
        g[0] = double(load_pop<ST>(s + c0 + 0 * L + y, inv_scale));
        g[1] = double(load_pop<ST>(s + cm + 1 * L + y, inv_scale));
        g[2] = double(load_pop<ST>(s + c0 + 2 * L + ym, inv_scale));
        g[3] = double(load_pop<ST>(s + cp + 3 * L + y, inv_scale));
        g[4] = double(load_pop<ST>(s + c0 + 4 * L + yp, inv_scale));
        g[5] = double(load_pop<ST>(s + cm + 5 * L + ym, inv_scale));
        g[6] = double(load_pop<ST>(s + cp + 6 * L + ym, inv_scale));
        g[7] = double(load_pop<ST>(s + cp + 7 * L + yp, inv_scale));
        g[8] = double(load_pop<ST>(s + cm + 8 * L + yp, inv_scale));
        double rho, ux, uy, drho;
        cell_state<Traits<ST>::shifted>(g, rho, ux, uy, drho);
        double lam[9], mu[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) lam[q] = lam_in[c0 + (size_t)q * L + y];
        // omega sensitivity: sum_q lambda_q (feq_q - g_q), in deviation form
        double geq[9];
        equilibrium_shifted<double>(drho, ux, uy, geq);
        double d = 0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const double gdev = Traits<ST>::shifted ? g[q] : g[q] - wq(q);
            d += lam[q] * (geq[q] - gdev);
        }
        acc += d;
        bgk_jacobian_t<double>(rho, ux, uy, lam, omega, mu);
        // S^T: lambda^n_q(x - c_q, y - cy_q) = mu_q(x, y)
        lam_out[c0 + 0 * L + y] = mu[0];
        lam_out[cm + 1 * L + y] = mu[1];
        lam_out[c0 + 2 * L + ym] = mu[2];
        lam_out[cp + 3 * L + y] = mu[3];
        lam_out[c0 + 4 * L + yp] = mu[4];
        lam_out[cm + 5 * L + ym] = mu[5];
        lam_out[cp + 6 * L + ym] = mu[6];
        lam_out[cp + 7 * L + yp] = mu[7];
        lam_out[cm + 8 * L + yp] = mu[8];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] += sh[0];
}

// Gradient with respect to the initial macroscopic fields: f^0 = feq(rho0, u0).
__global__ void __launch_bounds__(256)
    k_adj_initial(const double *lam0, int nx, int ny, int ldd, const double *rho0, const double *ux0,
                  const double *uy0, double *grho, double *gux, double *guy) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (y >= ny) return;
    const size_t i = (size_t)x * ny + y, c = pop_index(0, x, y, ldd);
    const double rho = rho0[i], ux = ux0[i], uy = uy0[i], u2 = ux * ux + uy * uy;
    double gr = 0, gx = 0, gy = 0;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const double l = lam0[c + (size_t)q * ldd] * wq(q);
        const double cu = cx(q) * ux + cy(q) * uy;
        gr += l * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u2);
        const double t = 3.0 + 9.0 * cu;
        gx += l * rho * (t * cx(q) - 3.0 * ux);
        gy += l * rho * (t * cy(q) - 3.0 * uy);
    }
    if (grho) grho[i] = gr;
    gux[i] = gx;
    guy[i] = gy;
}

__global__ void __launch_bounds__(256) k_adj_sum(const double *partials, int n, double *out) {
    __shared__ double sh[256];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a += partials[i];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ void __launch_bounds__(256) k_adj_zero(double *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace plbm

using namespace plbm;

struct plbm_adjoint {
    int nx = 0, ny = 0, storage = 0, nsteps = 0, every = 0, ncheck = 0;
    plbm_lattice *lat = nullptr;   // forward buffers (ping-pong) + launch plumbing
    void *checkpoints = nullptr;   // ncheck buffers: f^0, f^K, f^2K, ...
    void *segment = nullptr;       // every buffers: states of one recomputed segment
    void *lam[2] = {nullptr, nullptr};
    size_t buf_bytes = 0, lam_bytes = 0;
    double omega = 0;
    bool forward_done = false;
    int final_buf = 0;             // lattice buffer holding f^N
    double *d_partials = nullptr;  // kAdjBlocks
    double *d_scalar = nullptr;    // 2
    const double *rho0 = nullptr, *ux0 = nullptr, *uy0 = nullptr;  // caller's device fields
    hipStream_t stream = nullptr;
};

namespace plbm {

static int adj_reduce_blocks(const plbm_adjoint *a) {
    const size_t n = (size_t)a->nx * a->ny;
    return (int)std::min<size_t>(kAdjBlocks, (n + 255) / 256);
}

static void *ckpt(plbm_adjoint *a, int k) {
    return static_cast<char *>(a->checkpoints) + (size_t)k * a->buf_bytes;
}
static void *seg(plbm_adjoint *a, int k) {
    return static_cast<char *>(a->segment) + (size_t)k * a->buf_bytes;
}

}  // namespace plbm

extern "C" {

int plbm_adjoint_create(int nx, int ny, int storage, int nsteps, int checkpoint_every,
                        plbm_adjoint **out) try {
    PLBM_REQUIRE(out, PLBM_ERR_ARG, "null output pointer");
    *out = nullptr;
    PLBM_REQUIRE(storage == PLBM_F64 || storage == PLBM_F64S || storage == PLBM_F32S,
                 PLBM_ERR_ARG, "adjoint storage must be F64, F64S or F32S");
    PLBM_REQUIRE(nsteps >= 1 && checkpoint_every >= 1, PLBM_ERR_ARG,
                 "nsteps and checkpoint_every must be >= 1");
    auto *a = new plbm_adjoint();
    a->nx = nx;
    a->ny = ny;
    a->storage = storage;
    a->nsteps = nsteps;
    a->every = std::min(checkpoint_every, nsteps);
    a->ncheck = (nsteps + a->every - 1) / a->every;  // checkpoints at 0, K, 2K, ... < N
    int st = plbm_create(nx, ny, 2, storage, &a->lat);
    if (st != PLBM_OK) {
        delete a;
        return st;
    }
    a->stream = a->lat->stream;
    a->buf_bytes = a->lat->elems * storage_bytes(storage);
    a->lam_bytes = a->lat->elems * sizeof(double);  // FP64 adjoint state for every storage
    auto fail = [&](const char *m) {
        (void)hipGetLastError();
        set_error(m);
        plbm_adjoint_destroy(a);
        return PLBM_ERR_NOMEM;
    };
    if (hipMalloc(&a->checkpoints, a->ncheck * a->buf_bytes) != hipSuccess)
        return fail("hipMalloc of adjoint checkpoints failed");
    if (hipMalloc(&a->segment, a->every * a->buf_bytes) != hipSuccess)
        return fail("hipMalloc of the adjoint segment buffer failed");
    if (hipMalloc(&a->lam[0], a->lam_bytes) != hipSuccess ||
        hipMalloc(&a->lam[1], a->lam_bytes) != hipSuccess)
        return fail("hipMalloc of adjoint buffers failed");
    if (hipMalloc(&a->d_partials, kAdjBlocks * sizeof(double)) != hipSuccess ||
        hipMalloc(&a->d_scalar, 2 * sizeof(double)) != hipSuccess)
        return fail("hipMalloc of adjoint scalars failed");
    *out = a;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_adjoint_destroy(plbm_adjoint *a) try {
    if (!a) return PLBM_OK;
    if (a->stream) (void)hipStreamSynchronize(a->stream);
    if (a->checkpoints) (void)hipFree(a->checkpoints);
    if (a->segment) (void)hipFree(a->segment);
    for (void *p : a->lam)
        if (p) (void)hipFree(p);
    if (a->d_partials) (void)hipFree(a->d_partials);
    if (a->d_scalar) (void)hipFree(a->d_scalar);
    if (a->lat) plbm_destroy(a->lat);
    delete a;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_adjoint_set_stream(plbm_adjoint *a, void *stream) try {
    PLBM_REQUIRE(a, PLBM_ERR_ARG, "null adjoint handle");
    PLBM_TRY(plbm_set_stream(a->lat, stream));
    a->stream = a->lat->stream;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_adjoint_forward_dev(plbm_adjoint *a, const double *rho0, const double *ux0,
                             const double *uy0, double omega, const double *uxt,
                             const double *uyt, double *loss) try {
    PLBM_REQUIRE(a && rho0 && ux0 && uy0 && uxt && uyt && loss, PLBM_ERR_ARG, "null argument");
    PLBM_REQUIRE(std::isfinite(omega), PLBM_ERR_ARG, "non-finite omega");
    plbm_lattice *h = a->lat;
    PLBM_TRY(plbm_set_params(h, omega, 0.25));
    a->omega = omega;
    a->rho0 = rho0;
    a->ux0 = ux0;
    a->uy0 = uy0;
    PLBM_TRY(plbm_equilibrium_init_dev(h, 0, rho0, ux0, uy0));
    int cur = 0;
    for (int n = 0; n < a->nsteps; ++n) {
        if (n % a->every == 0)
            PLBM_HIP(hipMemcpyAsync(ckpt(a, n / a->every), h->f[cur], a->buf_bytes,
                                    hipMemcpyDeviceToDevice, a->stream));
        PLBM_TRY(launch_step(h, PLBM_BGK, h->f[cur], h->f[cur ^ 1]));
        cur ^= 1;
    }
    a->final_buf = cur;
    // loss and lambda^N
    const int nb = adj_reduce_blocks(a);
    PLBM_TRY(dispatch_storage(a->storage, [&](auto S) {
        constexpr int SS = decltype(S)::value;
        if constexpr (SS == PLBM_F16S) {
            set_error("adjoint: FP16 storage is not supported");
            return PLBM_ERR_ARG;
        } else {
            hipLaunchKernelGGL(k_adj_final<SS>, dim3(nb), dim3(256), 0, a->stream, h->f[cur], a->nx,
                               a->ny, h->ldd, float(1.0 / h->scale), uxt, uyt,
                               static_cast<double *>(a->lam[0]), a->d_partials);
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        }
    }));
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, a->stream, a->d_partials, nb, a->d_scalar);
    PLBM_HIP(hipGetLastError());
    PLBM_HIP(hipMemcpyAsync(loss, a->d_scalar, sizeof(double), hipMemcpyDeviceToHost, a->stream));
    PLBM_HIP(hipStreamSynchronize(a->stream));
    a->forward_done = true;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_adjoint_backward_dev(plbm_adjoint *a, double *grho0, double *gux0, double *guy0,
                              double *gomega) try {
    PLBM_REQUIRE(a && gux0 && guy0 && gomega, PLBM_ERR_ARG, "null argument");
    PLBM_REQUIRE(a->forward_done, PLBM_ERR_STATE, "run plbm_adjoint_forward_dev first");
    plbm_lattice *h = a->lat;
    const int nb = adj_reduce_blocks(a);
    hipLaunchKernelGGL(k_adj_zero, dim3((nb + 255) / 256), dim3(256), 0, a->stream, a->d_partials,
                       nb);
    PLBM_HIP(hipGetLastError());
    int cur = 0;  // lam[cur] holds lambda^{n+1}
    for (int k = a->ncheck - 1; k >= 0; --k) {
        const int n0 = k * a->every, n1 = std::min(a->nsteps, n0 + a->every);
        // recompute f^{n0} .. f^{n1-1} into the segment buffer
        PLBM_HIP(hipMemcpyAsync(seg(a, 0), ckpt(a, k), a->buf_bytes, hipMemcpyDeviceToDevice,
                                a->stream));
        for (int n = n0 + 1; n < n1; ++n)
            PLBM_TRY(launch_step(h, PLBM_BGK, seg(a, n - 1 - n0), seg(a, n - n0)));
        for (int n = n1 - 1; n >= n0; --n) {
            PLBM_TRY(dispatch_storage(a->storage, [&](auto S) {
                constexpr int SS = decltype(S)::value;
                if constexpr (SS == PLBM_F16S) {
                    return PLBM_ERR_ARG;
                } else {
                    hipLaunchKernelGGL(k_adj_step<SS>, dim3(nb), dim3(256), 0, a->stream,
                                       seg(a, n - n0), static_cast<const double *>(a->lam[cur]),
                                       static_cast<double *>(a->lam[cur ^ 1]), a->nx, a->ny,
                                       h->ldd, a->omega, float(1.0 / h->scale), a->d_partials);
                    PLBM_HIP(hipGetLastError());
                    return PLBM_OK;
                }
            }));
            cur ^= 1;
        }
    }
    // lam[cur] = lambda^0 -> gradients of the initial fields
    hipLaunchKernelGGL(k_adj_initial, dim3((a->ny + 255) / 256, a->nx), dim3(256), 0, a->stream,
                       static_cast<const double *>(a->lam[cur]), a->nx, a->ny, h->ldd, a->rho0,
                       a->ux0, a->uy0, grho0, gux0, guy0);
    PLBM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, a->stream, a->d_partials, nb,
                       a->d_scalar + 1);
    PLBM_HIP(hipGetLastError());
    PLBM_HIP(hipMemcpyAsync(gomega, a->d_scalar + 1, sizeof(double), hipMemcpyDeviceToHost,
                            a->stream));
    PLBM_HIP(hipStreamSynchronize(a->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

}  // extern "C"
