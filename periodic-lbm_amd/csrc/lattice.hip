// lattice.hip — the single-lattice engine: device kernels and the plbm_* C ABI.
//
// HBM layout of one population buffer: column-interleaved SoA (common.hpp, pop_index):
// column x holds the nine q-segments of the reference's f(ld, nx, 0:8)
// (src/fvm_bardow.F90:153) back to back, each ldd = ny rounded up to 64 elements long,
// so every segment starts 128-byte aligned for FP16 and 512-byte aligned for FP64.
//
// Hot path: k_fused, one pull-stream + collide pass per step (lbm_stream,
// src/periodic_lbm.f90:45-127, then collide_*, src/collision_*.F90).  Pull streaming
// is a permutation, so fusing it with the collision gives the same numbers as the
// reference's two passes while touching HBM once: 9 reads + 9 writes per cell.
// Periodic wrap is a select on the neighbour index; there is no halo.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "d2q9.hpp"
#include "reduce.hpp"
#include "stepper.hpp"

namespace plbm {

// ---------------------------------------------------------------------------- errors
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }
int host_exception() noexcept {
    const char *msg = "host-side C++ exception";
    int code = PLBM_ERR_HOST;
    try {
        throw;
    } catch (const std::bad_alloc &) {
        msg = "host memory allocation failed";
        code = PLBM_ERR_NOMEM;
    } catch (const std::length_error &) {
        msg = "host allocation size out of range";
        code = PLBM_ERR_NOMEM;
    } catch (...) {
    }
    try {
        g_err = msg;
    } catch (...) {
        g_err.clear();
    }
    return code;
}

// ---------------------------------------------------------------------------- helpers

// ---------------------------------------------------------------------------- kernels

// Fused pull-stream + collide.  One thread per cell; a workgroup covers blockDim.x
// consecutive y of one column x.  f_q(y, x) <- src_q(y - cy_q, x - cx_q).
template <int MODEL, int ST>
__global__ void __launch_bounds__(1024) k_fused(StepArgs a) {
    using T = typename Traits<ST>::T;
    using R = typename Traits<ST>::R;
    const int b = a.remap ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int by = b % a.nby, x = b / a.nby;
    const int y = by * blockDim.x + threadIdx.x;
    if (y >= a.ny) return;
    const int nx = a.nx, ny = a.ny;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const size_t L = (size_t)a.ldd;
    const size_t c0 = (size_t)x * a.cs, cm = (size_t)xm * a.cs, cp = (size_t)xp * a.cs;
    const T *__restrict__ s = static_cast<const T *>(a.src);
    T *__restrict__ d = static_cast<T *>(a.dst);
    const float is = a.inv_scale;
    R f[9];
    f[0] = load_pop<ST>(s + c0 + 0 * L + y, is);
    f[1] = load_pop<ST>(s + cm + 1 * L + y, is);
    f[2] = load_pop<ST>(s + c0 + 2 * L + ym, is);
    f[3] = load_pop<ST>(s + cp + 3 * L + y, is);
    f[4] = load_pop<ST>(s + c0 + 4 * L + yp, is);
    f[5] = load_pop<ST>(s + cm + 5 * L + ym, is);
    f[6] = load_pop<ST>(s + cp + 6 * L + ym, is);
    f[7] = load_pop<ST>(s + cp + 7 * L + yp, is);
    f[8] = load_pop<ST>(s + cm + 8 * L + yp, is);
    collide_cell<MODEL, ST>(f, a);
#pragma unroll
    for (int q = 0; q < 9; ++q) store_pop<ST>(d + c0 + q * L + y, f[q], a.scale);
}

// lbm_stream alone (src/periodic_lbm.f90:45-127): raw copy of stored values.
template <class T>
__global__ void __launch_bounds__(256) k_stream(const T *__restrict__ s, T *__restrict__ d,
                                                int nx, int ny, int ldd) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= ny) return;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const int xs = (x - cx(q) + nx) % nx, ys = (y - cy(q) + ny) % ny;
        d[pop_index(q, x, y, ldd)] = s[pop_index(q, xs, ys, ldd)];
    }
}

// collide_* alone, in place on one buffer.
template <int MODEL, int ST>
__global__ void __launch_bounds__(256) k_collide(StepArgs a) {
    using T = typename Traits<ST>::T;
    using R = typename Traits<ST>::R;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= a.ny) return;
    T *d = static_cast<T *>(a.dst);
    const size_t c = (size_t)x * a.cs + y;
    R f[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = load_pop<ST>(d + c + (size_t)q * a.ldd, a.inv_scale);
    collide_cell<MODEL, ST>(f, a);
#pragma unroll
    for (int q = 0; q < 9; ++q) store_pop<ST>(d + c + (size_t)q * a.ldd, f[q], a.scale);
}

// Cell macros in FP64.  F64: update_macros order (src/fvm_bardow.F90:371-379).
// Shifted: rho = 1 + sum g, u = j/rho.
template <int ST>
__device__ __forceinline__ void cell_macros(const typename Traits<ST>::T *s, int ldd, int x, int y,
                                            float inv_scale, double &rho, double &ux,
                                            double &uy) {
    double f[9];
    const size_t c = pop_index(0, x, y, ldd);
#pragma unroll
    for (int q = 0; q < 9; ++q) f[q] = double(load_pop<ST>(s + c + (size_t)q * ldd, inv_scale));
    if constexpr (!Traits<ST>::shifted) {
        macros_ref(f, rho, ux, uy);
    } else {
        double drho, jx, jy;
        moments_shifted<double>(f, drho, jx, jy);
        rho = 1.0 + drho;
        ux = jx / rho;
        uy = jy / rho;
    }
}

template <int ST>
__global__ void __launch_bounds__(256) k_macros(const void *src, int nx, int ny, int ldd,
                                                float inv_scale, double *rho, double *ux,
                                                double *uy) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= ny) return;
    const auto *s = static_cast<const typename Traits<ST>::T *>(src);
    double r, u, v;
    cell_macros<ST>(s, ldd, x, y, inv_scale, r, u, v);
    const size_t o = (size_t)x * ny + y;
    rho[o] = r;
    ux[o] = u;
    uy[o] = v;
}

// Deterministic two-pass FP64 reduction (reduce.hpp): workgroup b walks the cells
// b*256 + t, advancing by the grid's thread count, and writes its partial to slot b; one
// workgroup combines the partials.
template <int ST>
__global__ void __launch_bounds__(kRedThreads) k_reduce(const void *src, int nx, int ny, int ldd,
                                                        float inv_scale, const double *uxa,
                                                        const double *uya, int sx, int sy,
                                                        double *partials) {
    const auto *s = static_cast<const typename Traits<ST>::T *>(src);
    double acc[PLBM_NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
    const long long n = (long long)nx * ny, stride = (long long)gridDim.x * kRedThreads;
    const long long i0 = (long long)blockIdx.x * kRedThreads + threadIdx.x;
    CellWalk w;
    w.start(i0, ny);
    for (long long i = i0; i < n; i += stride) {
        double rho, ux, uy;
        cell_macros<ST>(s, ldd, w.x, w.y, inv_scale, rho, ux, uy);
        red_cell(acc, rho, ux, uy, uxa, uya, (size_t)i);
        w.next(sx, sy, ny);
    }
    const double v = red_block(acc);
    if (threadIdx.x < PLBM_NRED) partials[(size_t)blockIdx.x * PLBM_NRED + threadIdx.x] = v;
}

__global__ void __launch_bounds__(kRedThreads) k_reduce_parts(const double *partials, int nparts,
                                                              double *out) {
    double acc[PLBM_NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += kRedThreads)
#pragma unroll
        for (int k = 0; k < PLBM_NRED; ++k)
            acc[k] = red_combine(k, acc[k], partials[(size_t)b * PLBM_NRED + k]);
    const double v = red_block(acc);
    if (threadIdx.x < PLBM_NRED) out[threadIdx.x] = v;
}

// set_pdf_to_equilibrium: buffer = feq(rho, u).  Shifted forms are evaluated in FP64 and
// rounded once to the storage type.
template <int ST>
__global__ void __launch_bounds__(256) k_equilibrium(const double *rho, const double *ux,
                                                     const double *uy, void *dst, int nx,
                                                     int ny, int ldd, float scale) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= ny) return;
    auto *d = static_cast<typename Traits<ST>::T *>(dst);
    const size_t i = (size_t)x * ny + y, c = pop_index(0, x, y, ldd);
    double feq[9];
    if constexpr (!Traits<ST>::shifted)
        equilibrium_ref(rho[i], ux[i], uy[i], feq);
    else
        equilibrium_shifted<double>(rho[i] - 1.0, ux[i], uy[i], feq);
#pragma unroll
    for (int q = 0; q < 9; ++q) store_pop<ST>(d + c + (size_t)q * ldd, feq[q], scale);
}

// One q-plane between host layout (FP64, unshifted, pitch ld) and storage.
template <int ST>
__global__ void __launch_bounds__(256) k_plane_from_f64(const double *src, int ld, void *dst,
                                                        int q, int nx, int ny, int ldd,
                                                        double wshift, float scale) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= ny) return;
    auto *d = static_cast<typename Traits<ST>::T *>(dst);
    const double v = src[(size_t)x * ld + y];
    using T = typename Traits<ST>::T;
    const size_t o = pop_index(q, x, y, ldd);
    if constexpr (ST == PLBM_F16S)
        d[o] = T(float((v - wshift) * double(scale)));
    else if constexpr (Traits<ST>::shifted)
        d[o] = T(v - wshift);
    else
        d[o] = T(v);
}

template <int ST>
__global__ void __launch_bounds__(256) k_plane_to_f64(const void *src, int q, int nx, int ny,
                                                      int ldd, double *dst, int ld, double wshift,
                                                      float inv_scale) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x = blockIdx.y;
    if (y >= ny) return;
    const auto *s = static_cast<const typename Traits<ST>::T *>(src);
    const double v = double(s[pop_index(q, x, y, ldd)]);
    if constexpr (ST == PLBM_F16S)
        dst[(size_t)x * ld + y] = v * double(inv_scale) + wshift;
    else if constexpr (Traits<ST>::shifted)
        dst[(size_t)x * ld + y] = v + wshift;
    else
        dst[(size_t)x * ld + y] = v;
}

// HBM probe: one 16-byte vector per thread, non-temporal store, one pass over the
// array with many small workgroups (the fastest plain copy shape measured on MI355X).
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy(const v4f *__restrict__ s, v4f *__restrict__ d,
                                              size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(s[i], d + i);
}

// ---------------------------------------------------------------------------- dispatch

static StepArgs make_args(const plbm_lattice *h, const void *src, void *dst,
                          LaunchCfg cfg = LaunchCfg()) {
    StepArgs a;
    a.src = src;
    a.dst = dst;
    a.nx = h->nx;
    a.ny = h->ny;
    a.ldd = h->ldd;
    a.nby = (h->ny + cfg.block - 1) / cfg.block;
    a.cs = (size_t)9 * h->ldd;
    a.remap = cfg.remap;
    a.omega = h->omega;
    a.lambda_odd = h->lambda_odd;
    a.omega_f = float(h->omega);
    a.lambda_odd_f = float(h->lambda_odd);
    a.scale = float(h->scale);
    a.inv_scale = float(1.0 / h->scale);
    return a;
}

static int check_buf(const plbm_lattice *h, int buf) {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(buf >= 0 && buf < h->nf, PLBM_ERR_ARG,
                 "buffer id " + std::to_string(buf) + " out of range [0," +
                     std::to_string(h->nf) + ")");
    return PLBM_OK;
}

static int ensure_tmp(plbm_lattice *h, size_t bytes) {
    if (h->tmp_bytes >= bytes) return PLBM_OK;
    if (h->d_tmp) PLBM_HIP(hipFree(h->d_tmp));
    h->d_tmp = nullptr;
    h->tmp_bytes = 0;
    if (hipMalloc(&h->d_tmp, bytes) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc of " + std::to_string(bytes) + " staging bytes failed");
        return PLBM_ERR_NOMEM;
    }
    h->tmp_bytes = bytes;
    return PLBM_OK;
}

// Workgroups of the reduction of an nx x ny grid: one per 256 cells, at most
// kReduceBlocks (the partial slots); each thread then walks every 256*nb-th cell.
static int reduce_blocks(const plbm_lattice *h) {
    const size_t n = (size_t)h->nx * h->ny;
    return (int)std::min<size_t>(kReduceBlocks, (n + kRedThreads - 1) / kRedThreads);
}

// The two reduction kernels on the lattice's stream: partials of reduce_blocks(h)
// workgroups into parts, their combination into out_dev (PLBM_NRED doubles, device).
static int launch_reduce(plbm_lattice *h, int buf, const double *uxa, const double *uya,
                         double *parts, double *out_dev) {
    const int nb = reduce_blocks(h);
    const long long stride = (long long)nb * kRedThreads;
    const int sx = (int)(stride / h->ny), sy = (int)(stride % h->ny);
    PLBM_TRY(dispatch_storage(h->storage, [&](auto S) {
        hipLaunchKernelGGL(k_reduce<decltype(S)::value>, dim3(nb), dim3(kRedThreads), 0,
                           h->stream, h->f[buf], h->nx, h->ny, h->ldd, float(1.0 / h->scale),
                           uxa, uya, sx, sy, parts);
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    }));
    hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(kRedThreads), 0, h->stream, parts, nb,
                       out_dev);
    PLBM_HIP(hipGetLastError());
    return PLBM_OK;
}

// Scratch of plbm_reduce_static, which allocates nothing: the partial sums of the
// reduction workgroups (kStaticReduceParts x PLBM_NRED doubles, 64 KiB), one copy per
// device for the life of the process.  Host code must not pass `scratch` itself to a
// kernel (on the host it names a stand-in, not device memory): its device address is
// looked up once per device with hipGetSymbolAddress and kept in g_static_parts.  Every
// call on a device shares that device's copy; each call overwrites every slot it reads
// (one per workgroup it launches) before reading it, so nothing needs clearing, and calls
// that overlap on one device race.
constexpr int kStaticReduceParts = kReduceBlocks;
constexpr int kStaticMaxDevices = 64;
static __device__ __align__(256) double scratch[kStaticReduceParts * PLBM_NRED];
static double *g_static_parts[kStaticMaxDevices] = {};
static std::mutex g_static_mu;

static int static_reduce_scratch(int dev, double **parts) {
    PLBM_REQUIRE(dev >= 0 && dev < kStaticMaxDevices, PLBM_ERR_ARG,
                 "device ordinal out of range for the static scratch table");
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_static_parts[dev]) {
        int cur = -1;
        PLBM_HIP(hipGetDevice(&cur));
        if (cur != dev) PLBM_HIP(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || p == nullptr) {
            (void)hipGetLastError();
            set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return PLBM_ERR_HIP;
        }
        g_static_parts[dev] = static_cast<double *>(p);
    }
    *parts = g_static_parts[dev];
    return PLBM_OK;
}

static dim3 cell_grid(const plbm_lattice *h) {
    return dim3((unsigned)((h->ny + 255) / 256), (unsigned)h->nx);
}

int launch_step_cfg(plbm_lattice *h, int model, const void *src, void *dst, LaunchCfg cfg) {
    StepArgs a = make_args(h, src, dst, cfg);
    const unsigned nblocks = (unsigned)((size_t)a.nby * h->nx);
    return dispatch_model(model, [&](auto M) {
        return dispatch_storage(h->storage, [&](auto S) {
            hipLaunchKernelGGL((k_fused<decltype(M)::value, decltype(S)::value>), dim3(nblocks),
                               dim3(cfg.block), 0, h->stream, a);
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        });
    });
}

int run_steps(plbm_lattice *h, int model, int ns, int lag_ns, int nsteps, int *iold,
              int *inew);
int default_steps_per_launch(const plbm_lattice *h);

int launch_step(plbm_lattice *h, int model, const void *src, void *dst) {
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    return launch_step_cfg(h, model, src, dst, h->cfg[model]);
}

static bool valid_cfg(LaunchCfg c) {
    return c.block >= 64 && c.block <= 1024 && c.block % 64 == 0 && (c.remap == 0 || c.remap == 1);
}

// ---------------------------------------------------------------------------- autotuner
//
// Picks the fused kernel's launch shape (workgroup size x XCD renumbering) for one
// (device, storage, model, grid) by timing every candidate with HIP events on the
// lattice's stream: `reps` launches from buffer src into a scratch buffer (so the
// lattice's own state is untouched), best of three trials each.  Results are cached per
// process, keyed by device name, storage, model and grid size.

struct TuneKey {
    std::string dev;
    int storage, model, nx, ny;
    bool operator<(const TuneKey &o) const {
        return std::tie(dev, storage, model, nx, ny) <
               std::tie(o.dev, o.storage, o.model, o.nx, o.ny);
    }
};
static std::map<TuneKey, std::pair<LaunchCfg, double>> g_tune_cache;
static std::mutex g_tune_mu;

static std::string device_name(int dev) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
    return std::string(p.gcnArchName) + "/" + std::to_string(p.multiProcessorCount);
}

int autotune_step(plbm_lattice *h, int model, int src, int reps, LaunchCfg *best_out,
                  double *best_ms) {
    TuneKey key{device_name(h->device), h->storage, model, h->nx, h->ny};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            *best_out = it->second.first;
            *best_ms = it->second.second;
            return PLBM_OK;
        }
    }
    void *scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = on_scope_exit([&] {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (scratch) (void)hipFree(scratch);
    });
    const size_t bytes = h->elems * storage_bytes(h->storage);
    if (hipMalloc(&scratch, bytes) != hipSuccess) {
        (void)hipGetLastError();
        scratch = nullptr;
        set_error("autotune: no memory for a scratch buffer");
        return PLBM_ERR_NOMEM;
    }
    PLBM_HIP(hipEventCreate(&e0));
    PLBM_HIP(hipEventCreate(&e1));
    // ms per launch of launch(): the best of three trials of `reps` launches, each trial
    // after one warm-up launch, timed with events on the lattice's stream
    auto time_launch = [&](auto &&launch, double *ms_out) -> int {
        double tmin = 1e300;
        for (int trial = 0; trial < 3; ++trial) {
            PLBM_TRY(launch());  // warm
            (void)hipEventRecord(e0, h->stream);
            for (int r = 0; r < reps; ++r) PLBM_TRY(launch());
            (void)hipEventRecord(e1, h->stream);
            if (hipEventSynchronize(e1) != hipSuccess) {
                (void)hipGetLastError();
                set_error("autotune: event synchronisation failed");
                return PLBM_ERR_HIP;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tmin = std::min(tmin, double(ms) / reps);
        }
        *ms_out = tmin;
        return PLBM_OK;
    };
    LaunchCfg best;
    double tbest = 1e300;
    for (int block : {64, 128, 256, 512, 1024}) {
        for (int remap : {0, 1}) {
            const LaunchCfg c{block, remap};
            double t;
            PLBM_TRY(time_launch(
                [&] { return launch_step_cfg(h, model, h->f[src], scratch, c); }, &t));
            if (t < tbest) {
                tbest = t;
                best = c;
            }
        }
    }
    // steps per pass of plbm_run: the single-step kernel at the shape just chosen against
    // passes of 2..8 steps, compared per lattice update
    best.ns = 1;
    for (int ns = 2; ns <= max_steps_per_pass(h->storage); ++ns) {
        double t;
        PLBM_TRY(time_launch([&] { return launch_multi(h, model, ns, h->f[src], scratch); }, &t));
        if (t / ns < tbest) {
            tbest = t / ns;
            best.ns = ns;
        }
    }
    // depth of a run's last pass, which also stores the lagged state: a run of n steps
    // costs about (n - L) x tbest for its other passes plus that pass's time t(L), so the
    // depth L >= ns that minimises t(L) - L x tbest.  Where that pass is bound by its
    // writes (TRT FP64: about 3 ms at any depth at 8192^2) the deepest wins, and 20 steps
    // run as 6 + 6 + 8 whether ns came out as 7 or 8.  The lag level goes to the
    // lattice's third buffer, whose contents are not needed between runs.
    best.lag_ns = 0;
    void *aux = best.ns > 1 ? pass_aux(h) : nullptr;
    if (aux != nullptr) {
        double cbest = 1e300;
        for (int L = best.ns; L <= max_steps_per_pass(h->storage); ++L) {
            double t;
            PLBM_TRY(time_launch(
                [&] { return launch_multi(h, model, L, h->f[src], scratch, aux); }, &t));
            if (t - L * tbest < cbest) {
                cbest = t - L * tbest;
                best.lag_ns = L;
            }
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = {best, tbest};
    }
    *best_out = best;
    *best_ms = tbest;
    return PLBM_OK;
}

static int require_params(const plbm_lattice *h) {
    PLBM_REQUIRE(h->params_set, PLBM_ERR_STATE,
                 "collision parameters not set: call plbm_set_params first");
    return PLBM_OK;
}

}  // namespace plbm

using namespace plbm;

// ============================================================================ C ABI

extern "C" {

const char *plbm_last_error(void) { return last_error(); }

double plbm_lambda_d(double omega, double x) {
    // src/collision_trt.F90:30-34
    return (4.0 - 2.0 * omega) / ((((4.0 * x) * omega) + 2.0) - omega);
}

double plbm_magic_number(double le, double ld) {
    // src/collision_trt.F90:24-28
    return ((2.0 - le) * (2.0 - ld)) / ((4.0 * le) * ld);
}

int plbm_create(int nx, int ny, int nf, int storage, plbm_lattice **out) try {
    PLBM_REQUIRE(out, PLBM_ERR_ARG, "null output pointer");
    *out = nullptr;
    PLBM_REQUIRE(nx >= 1 && ny >= 1, PLBM_ERR_ARG, "grid size must be positive");
    PLBM_REQUIRE(nf >= 1 && nf <= kMaxBuffers, PLBM_ERR_ARG,
                 "nf must be in [1," + std::to_string(kMaxBuffers) + "]");
    PLBM_REQUIRE(storage >= PLBM_F64 && storage <= PLBM_F16S, PLBM_ERR_ARG, "unknown storage");
    const int ldd = (ny + 63) / 64 * 64;
    PLBM_REQUIRE((size_t)nx < ((size_t)1 << 31) / 9, PLBM_ERR_ARG, "nx too large");
    auto *h = new plbm_lattice();
    h->nx = nx;
    h->ny = ny;
    h->ldd = ldd;
    h->nf = nf;
    h->storage = storage;
    h->elems = (size_t)9 * nx * ldd;
    h->scale = storage == PLBM_F16S ? 1024.0 : 1.0;
    if (hipGetDevice(&h->device) != hipSuccess ||
        hipStreamCreateWithFlags(&h->own, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        set_error("no usable HIP device / stream");
        delete h;
        return PLBM_ERR_HIP;
    }
    h->stream = h->own;
    const size_t bytes = h->elems * storage_bytes(storage);
    for (int b = 0; b < nf; ++b) {
        if (hipMalloc(&h->f[b], bytes) != hipSuccess) {
            (void)hipGetLastError();
            set_error("hipMalloc of population buffer (" + std::to_string(bytes) +
                      " bytes) failed");
            plbm_destroy(h);
            return PLBM_ERR_NOMEM;
        }
    }
    *out = h;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_destroy(plbm_lattice *h) try {
    if (!h) return PLBM_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int b = 0; b < kMaxBuffers; ++b)
        if (h->f[b]) (void)hipFree(h->f[b]);
    if (h->aux) (void)hipFree(h->aux);
    for (hipEvent_t e : h->run_events) (void)hipEventDestroy(e);
    if (h->d_tmp) (void)hipFree(h->d_tmp);
    if (h->d_partials) (void)hipFree(h->d_partials);
    if (h->d_red) (void)hipFree(h->d_red);
    if (h->own) (void)hipStreamDestroy(h->own);
    delete h;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_get_info(const plbm_lattice *h, plbm_info *out) try {
    PLBM_REQUIRE(h && out, PLBM_ERR_ARG, "null argument");
    out->nx = h->nx;
    out->ny = h->ny;
    out->ldd = h->ldd;
    out->nf = h->nf;
    out->storage = h->storage;
    out->device = h->device;
    out->omega = h->omega;
    out->trt_magic = h->trt_magic;
    out->lambda_odd = h->lambda_odd;
    out->f16_scale = h->scale;
    out->bytes_per_buffer = h->elems * storage_bytes(h->storage);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_set_stream(plbm_lattice *h, void *stream) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    h->stream = stream ? static_cast<hipStream_t>(stream) : h->own;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_buffer_ptr(plbm_lattice *h, int buf, void **dptr) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(dptr, PLBM_ERR_ARG, "null output pointer");
    *dptr = h->f[buf];
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_set_params(plbm_lattice *h, double omega, double trt_magic) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(std::isfinite(omega) && std::isfinite(trt_magic), PLBM_ERR_ARG,
                 "non-finite collision parameter");
    h->omega = omega;
    h->trt_magic = trt_magic;
    h->lambda_odd = plbm_lambda_d(omega, trt_magic);
    h->params_set = true;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_set_f16_scale(plbm_lattice *h, double scale) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    int e = 0;
    PLBM_REQUIRE(scale > 0 && std::frexp(scale, &e) == 0.5, PLBM_ERR_ARG,
                 "FP16 scale must be a positive power of two");
    h->scale = h->storage == PLBM_F16S ? scale : 1.0;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_upload_f(plbm_lattice *h, int buf, const double *f, int ld) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(f && ld >= h->ny, PLBM_ERR_ARG, "null array or ld < ny");
    const size_t plane_host = (size_t)ld * h->nx;
    PLBM_TRY(ensure_tmp(h, (size_t)h->nx * h->ny * sizeof(double)));
    std::vector<double> stage;
    for (int q = 0; q < 9; ++q) {
        // host plane (pitch ld) -> compact staging plane (pitch ny)
        PLBM_TRY(plane_from_host(h->d_tmp, f + q * plane_host, ld, h->ny, h->nx, h->stream,
                                 stage));
        const double w = wq(q);
        PLBM_TRY(dispatch_storage(h->storage, [&](auto S) {
            hipLaunchKernelGGL(k_plane_from_f64<decltype(S)::value>, cell_grid(h), dim3(256), 0,
                               h->stream, h->d_tmp, h->ny, h->f[buf], q, h->nx, h->ny, h->ldd, w,
                               float(h->scale));
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        }));
        PLBM_HIP(hipStreamSynchronize(h->stream));
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_download_f(plbm_lattice *h, int buf, double *f, int ld) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(f && ld >= h->ny, PLBM_ERR_ARG, "null array or ld < ny");
    const size_t plane_host = (size_t)ld * h->nx;
    PLBM_TRY(ensure_tmp(h, (size_t)h->nx * h->ny * sizeof(double)));
    std::vector<double> stage;
    for (int q = 0; q < 9; ++q) {
        const double w = wq(q);
        PLBM_TRY(dispatch_storage(h->storage, [&](auto S) {
            hipLaunchKernelGGL(k_plane_to_f64<decltype(S)::value>, cell_grid(h), dim3(256), 0,
                               h->stream, h->f[buf], q, h->nx, h->ny, h->ldd, h->d_tmp, h->ny, w,
                               float(1.0 / h->scale));
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        }));
        // compact staging plane -> host plane (pitch ld); padding rows ny..ld-1 untouched
        PLBM_TRY(plane_to_host(f + q * plane_host, ld, h->d_tmp, h->ny, h->nx, h->stream, stage));
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_equilibrium_init_dev(plbm_lattice *h, int buf, const double *rho, const double *ux,
                              const double *uy) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(rho && ux && uy, PLBM_ERR_ARG, "null field pointer");
    return dispatch_storage(h->storage, [&](auto S) {
        hipLaunchKernelGGL(k_equilibrium<decltype(S)::value>, cell_grid(h), dim3(256), 0,
                           h->stream, rho, ux, uy, h->f[buf], h->nx, h->ny, h->ldd,
                           float(h->scale));
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
} catch (...) {
    return plbm::host_exception();
}

int plbm_equilibrium_init(plbm_lattice *h, int buf, const double *rho, const double *ux,
                          const double *uy) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(rho && ux && uy, PLBM_ERR_ARG, "null field pointer");
    const size_t n = (size_t)h->nx * h->ny;
    PLBM_TRY(ensure_tmp(h, 3 * n * sizeof(double)));
    PLBM_HIP(hipMemcpyAsync(h->d_tmp, rho, n * 8, hipMemcpyHostToDevice, h->stream));
    PLBM_HIP(hipMemcpyAsync(h->d_tmp + n, ux, n * 8, hipMemcpyHostToDevice, h->stream));
    PLBM_HIP(hipMemcpyAsync(h->d_tmp + 2 * n, uy, n * 8, hipMemcpyHostToDevice, h->stream));
    PLBM_TRY(plbm_equilibrium_init_dev(h, buf, h->d_tmp, h->d_tmp + n, h->d_tmp + 2 * n));
    PLBM_HIP(hipStreamSynchronize(h->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_step(plbm_lattice *h, int model, int src, int dst) try {
    PLBM_TRY(check_buf(h, src));
    PLBM_TRY(check_buf(h, dst));
    PLBM_REQUIRE(src != dst, PLBM_ERR_ARG, "fused step needs distinct src and dst buffers");
    PLBM_TRY(require_params(h));
    return launch_step(h, model, h->f[src], h->f[dst]);
} catch (...) {
    return plbm::host_exception();
}

int plbm_run(plbm_lattice *h, int model, int nsteps, int *iold, int *inew) try {
    PLBM_REQUIRE(h && iold && inew, PLBM_ERR_ARG, "null argument");
    PLBM_REQUIRE(nsteps >= 0, PLBM_ERR_ARG, "nsteps must be >= 0");
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    PLBM_TRY(check_buf(h, *iold));
    PLBM_TRY(check_buf(h, *inew));
    PLBM_REQUIRE(*iold != *inew, PLBM_ERR_ARG, "iold == inew");
    PLBM_TRY(require_params(h));
    const int ns = h->cfg[model].ns > 0 ? h->cfg[model].ns : default_steps_per_launch(h);
    return run_steps(h, model, ns, h->cfg[model].lag_ns, nsteps, iold, inew);
} catch (...) {
    return plbm::host_exception();
}

int plbm_stream(plbm_lattice *h, int src, int dst) try {
    PLBM_TRY(check_buf(h, src));
    PLBM_TRY(check_buf(h, dst));
    PLBM_REQUIRE(src != dst, PLBM_ERR_ARG, "streaming needs distinct src and dst buffers");
    return dispatch_storage(h->storage, [&](auto S) {
        using T = typename Traits<decltype(S)::value>::T;
        hipLaunchKernelGGL(k_stream<T>, cell_grid(h), dim3(256), 0, h->stream,
                           static_cast<const T *>(h->f[src]), static_cast<T *>(h->f[dst]), h->nx,
                           h->ny, h->ldd);
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
} catch (...) {
    return plbm::host_exception();
}

int plbm_collide(plbm_lattice *h, int model, int buf) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_TRY(require_params(h));
    StepArgs a = make_args(h, h->f[buf], h->f[buf]);
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    return dispatch_model(model, [&](auto M) {
        return dispatch_storage(h->storage, [&](auto S) {
            hipLaunchKernelGGL((k_collide<decltype(M)::value, decltype(S)::value>), cell_grid(h),
                               dim3(256), 0, h->stream, a);
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        });
    });
} catch (...) {
    return plbm::host_exception();
}

int plbm_copy(plbm_lattice *h, int src, int dst) try {
    PLBM_TRY(check_buf(h, src));
    PLBM_TRY(check_buf(h, dst));
    if (src == dst) return PLBM_OK;
    PLBM_HIP(hipMemcpyAsync(h->f[dst], h->f[src], h->elems * storage_bytes(h->storage),
                            hipMemcpyDeviceToDevice, h->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_macros_dev(plbm_lattice *h, int buf, double *rho, double *ux, double *uy) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(rho && ux && uy, PLBM_ERR_ARG, "null field pointer");
    return dispatch_storage(h->storage, [&](auto S) {
        hipLaunchKernelGGL(k_macros<decltype(S)::value>, cell_grid(h), dim3(256), 0, h->stream,
                           h->f[buf], h->nx, h->ny, h->ldd, float(1.0 / h->scale), rho,
                           ux, uy);
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    });
} catch (...) {
    return plbm::host_exception();
}

int plbm_macros(plbm_lattice *h, int buf, double *rho, double *ux, double *uy) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(rho && ux && uy, PLBM_ERR_ARG, "null field pointer");
    const size_t n = (size_t)h->nx * h->ny;
    PLBM_TRY(ensure_tmp(h, 3 * n * sizeof(double)));
    PLBM_TRY(plbm_macros_dev(h, buf, h->d_tmp, h->d_tmp + n, h->d_tmp + 2 * n));
    PLBM_HIP(hipMemcpyAsync(rho, h->d_tmp, n * 8, hipMemcpyDeviceToHost, h->stream));
    PLBM_HIP(hipMemcpyAsync(ux, h->d_tmp + n, n * 8, hipMemcpyDeviceToHost, h->stream));
    PLBM_HIP(hipMemcpyAsync(uy, h->d_tmp + 2 * n, n * 8, hipMemcpyDeviceToHost, h->stream));
    PLBM_HIP(hipStreamSynchronize(h->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_reduce_dev(plbm_lattice *h, int buf, const double *uxa, const double *uya,
                    double out[PLBM_NRED]) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(out, PLBM_ERR_ARG, "null output");
    PLBM_REQUIRE((uxa == nullptr) == (uya == nullptr), PLBM_ERR_ARG,
                 "uxa and uya must both be given or both be NULL");
    if (!h->d_partials) PLBM_HIP(hipMalloc(&h->d_partials, kReduceBlocks * PLBM_NRED * 8));
    if (!h->d_red) PLBM_HIP(hipMalloc(&h->d_red, PLBM_NRED * 8));
    PLBM_TRY(launch_reduce(h, buf, uxa, uya, h->d_partials, h->d_red));
    PLBM_HIP(hipMemcpyAsync(out, h->d_red, PLBM_NRED * 8, hipMemcpyDeviceToHost, h->stream));
    PLBM_HIP(hipStreamSynchronize(h->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_reduce_static(plbm_lattice *h, int buf, const double *uxa, const double *uya,
                       double *out_dev) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE(out_dev, PLBM_ERR_ARG, "null output");
    PLBM_REQUIRE((uxa == nullptr) == (uya == nullptr), PLBM_ERR_ARG,
                 "uxa and uya must both be given or both be NULL");
    const int nb = reduce_blocks(h);
    PLBM_REQUIRE(nb <= kStaticReduceParts, PLBM_ERR_ARG,
                 "plbm_reduce_static: " + std::to_string(nb) +
                     " partial sums exceed the static scratch of " +
                     std::to_string(kStaticReduceParts));
    double *parts = nullptr;
    PLBM_TRY(static_reduce_scratch(h->device, &parts));
    return launch_reduce(h, buf, uxa, uya, parts, out_dev);
} catch (...) {
    return plbm::host_exception();
}

int plbm_reduce(plbm_lattice *h, int buf, const double *uxa, const double *uya,
                double out[PLBM_NRED]) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_REQUIRE((uxa == nullptr) == (uya == nullptr), PLBM_ERR_ARG,
                 "uxa and uya must both be given or both be NULL");
    if (!uxa) return plbm_reduce_dev(h, buf, nullptr, nullptr, out);
    const size_t n = (size_t)h->nx * h->ny;
    PLBM_TRY(ensure_tmp(h, 2 * n * sizeof(double)));
    PLBM_HIP(hipMemcpyAsync(h->d_tmp, uxa, n * 8, hipMemcpyHostToDevice, h->stream));
    PLBM_HIP(hipMemcpyAsync(h->d_tmp + n, uya, n * 8, hipMemcpyHostToDevice, h->stream));
    return plbm_reduce_dev(h, buf, h->d_tmp, h->d_tmp + n, out);
} catch (...) {
    return plbm::host_exception();
}

int plbm_sync(plbm_lattice *h) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_HIP(hipStreamSynchronize(h->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_set_launch(plbm_lattice *h, int model, int block, int xcd_remap) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG, "unknown collision model");
    LaunchCfg c{block, xcd_remap, h->cfg[model].ns, h->cfg[model].lag_ns};
    PLBM_REQUIRE(valid_cfg(c), PLBM_ERR_ARG,
                 "block must be a multiple of 64 in [64,1024] and xcd_remap 0 or 1");
    h->cfg[model] = c;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_set_steps_per_launch(plbm_lattice *h, int model, int ns) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG, "unknown collision model");
    PLBM_REQUIRE(ns >= 0 && ns <= max_steps_per_pass(h->storage), PLBM_ERR_ARG,
                 "steps per launch must be 0.." + std::to_string(max_steps_per_pass(h->storage)) +
                     " for this storage");
    h->cfg[model].ns = ns;
    h->cfg[model].lag_ns = 0;  // the chosen depth holds for the last pass as well
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_get_steps_per_launch(const plbm_lattice *h, int model, int *ns) try {
    PLBM_REQUIRE(h && ns, PLBM_ERR_ARG, "null argument");
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG, "unknown collision model");
    *ns = h->cfg[model].ns > 0 ? h->cfg[model].ns : default_steps_per_launch(h);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_get_launch(const plbm_lattice *h, int model, int *block, int *xcd_remap) try {
    PLBM_REQUIRE(h && block && xcd_remap, PLBM_ERR_ARG, "null argument");
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG, "unknown collision model");
    *block = h->cfg[model].block;
    *xcd_remap = h->cfg[model].remap;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_autotune(plbm_lattice *h, int model, int buf, int reps, double *best_ms) try {
    PLBM_TRY(check_buf(h, buf));
    PLBM_TRY(require_params(h));
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG, "unknown collision model");
    PLBM_REQUIRE(reps >= 1, PLBM_ERR_ARG, "reps must be >= 1");
    LaunchCfg best;
    double ms = 0;
    PLBM_TRY(autotune_step(h, model, buf, reps, &best, &ms));
    h->cfg[model] = best;
    if (best_ms) *best_ms = ms;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_copy_bandwidth(size_t bytes, int iters, double *ms) try {
    PLBM_REQUIRE(ms && iters > 0 && bytes >= 16, PLBM_ERR_ARG, "bad argument");
    const size_t n = bytes / 16;
    PLBM_REQUIRE(n / 256 < (size_t)1 << 31, PLBM_ERR_ARG, "copy probe size too large");
    v4f *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = on_scope_exit([&] {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    });
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc failed in copy probe");
        return PLBM_ERR_NOMEM;
    }
    PLBM_HIP(hipEventCreate(&e0));
    PLBM_HIP(hipEventCreate(&e1));
    PLBM_HIP(hipMemset(a, 0, n * 16));
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, a, b, n);  // warm-up
    PLBM_HIP(hipGetLastError());
    PLBM_HIP(hipEventRecord(e0, nullptr));
    for (int i = 0; i < iters; ++i)
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, a, b, n);
    PLBM_HIP(hipEventRecord(e1, nullptr));
    PLBM_HIP(hipEventSynchronize(e1));
    float t = 0;
    PLBM_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = double(t) / iters;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

}  // extern "C"
