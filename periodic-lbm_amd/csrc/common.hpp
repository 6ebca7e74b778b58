// common.hpp — host-side plumbing shared by the engine's translation units:
// error reporting, HIP call checking, storage traits, the lattice object.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/plbm.h"

namespace plbm {

// Last error text of the calling thread (plbm_last_error).
void set_error(const std::string &msg);
const char *last_error();
// The handler of the C ABI's int-returning entry points, which are function-try-blocks:
// no C++ exception (a failed host allocation, say) unwinds into a C, Fortran or Python
// caller.  Call it only inside a catch clause; it returns PLBM_ERR_NOMEM for an
// allocation failure, PLBM_ERR_HOST otherwise, and sets the error text.
int host_exception() noexcept;

#define PLBM_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::plbm::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
            return PLBM_ERR_HIP;                                                          \
        }                                                                                 \
    } while (0)

#define PLBM_REQUIRE(cond, code, msg)                                                     \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            ::plbm::set_error(msg);                                                       \
            return (code);                                                                \
        }                                                                                 \
    } while (0)

#define PLBM_TRY(call)                                                                    \
    do {                                                                                  \
        int s_ = (call);                                                                  \
        if (s_ != PLBM_OK) return s_;                                                     \
    } while (0)

// Runs f when the guard leaves scope: releases device buffers and events on every return
// path of a function (PLBM_HIP / PLBM_TRY return early).
template <class F>
struct ScopeExit {
    F f;
    ~ScopeExit() { f(); }
};
template <class F>
ScopeExit<F> on_scope_exit(F f) {
    return ScopeExit<F>{f};
}

// Storage traits: T is what sits in HBM, R is the arithmetic type of the collision.
template <int ST> struct Traits;
template <> struct Traits<PLBM_F64> {
    using T = double;
    using R = double;
    static constexpr bool shifted = false;
};
template <> struct Traits<PLBM_F64S> {
    using T = double;
    using R = double;
    static constexpr bool shifted = true;
};
template <> struct Traits<PLBM_F32S> {
    using T = float;
    using R = float;
    static constexpr bool shifted = true;
};
template <> struct Traits<PLBM_F16S> {
    using T = _Float16;
    using R = float;
    static constexpr bool shifted = true;
};

inline size_t storage_bytes(int st) {
    return st == PLBM_F16S ? 2 : (st == PLBM_F32S ? 4 : 8);
}

// Device layout of one population buffer: column-interleaved SoA.  Column x holds nine
// contiguous q-segments of ldd elements (y unit-stride), so element (q, x, y) sits at
// (x*9 + q)*ldd + y.  The nine read and nine write streams of a workgroup then lie
// within a few hundred KB of each other instead of one plane (hundreds of MB, a power
// of two at 8192^2) apart, which keeps HBM channels and banks evenly loaded.
__host__ __device__ inline size_t pop_index(int q, int x, int y, int ldd) {
    return ((size_t)x * 9 + (size_t)q) * (size_t)ldd + (size_t)y;
}

// Launch shape of the streaming kernels: threads per workgroup (a multiple of 64, at
// most 1024) and whether workgroups are renumbered so each XCD owns a contiguous range
// of tiles.
struct LaunchCfg {
    int block = 256;
    int remap = 1;
    int ns = 0;  // steps per pass of plbm_run (1 = single-step kernel, 2..8 = temporal
                 // blocking, 0 = chosen from the grid size)
    int lag_ns = 0;  // steps of a run's last pass, which also stores the lagged state
                     // (0 = ns; set by the autotuner)
};

// Per-launch scalars of one step, passed by value.
struct StepArgs {
    const void *src;
    void *dst;
    int nx, ny, ldd, nby;   // nby = y-blocks per column
    size_t cs;              // column stride = 9*ldd elements
    int remap;
    double omega, lambda_odd;
    float omega_f, lambda_odd_f;
    float scale, inv_scale; // FP16 storage scale (1 otherwise)
    // Pass kernels only: when non-null, the pass also stores the state one step before
    // its last (level NS-1, post-collision) here, in the layout of dst.
    void *lag = nullptr;
};

// Strip length (columns per thread) of the stencil kernels that walk along x with a
// register window: 32 (measured best at 8192^2), shorter where that would leave too few
// 256-thread workgroups for the 256 CUs (2 resident per CU at those kernels' register
// count: 512 fill the chip once).
inline int strip_columns(int nx, int ny) {
    const long nby = (ny + 255) / 256;
    if (nby * ((nx + 31) / 32) >= 1024) return 32;
    if (nby * ((nx + 7) / 8) >= 512) return 8;
    return 1;
}

// Compile-time dispatch of a runtime storage / model id to a templated callable.
template <class F>
inline int dispatch_storage(int st, F &&f) {
    switch (st) {
    case PLBM_F64: return f(std::integral_constant<int, PLBM_F64>{});
    case PLBM_F64S: return f(std::integral_constant<int, PLBM_F64S>{});
    case PLBM_F32S: return f(std::integral_constant<int, PLBM_F32S>{});
    case PLBM_F16S: return f(std::integral_constant<int, PLBM_F16S>{});
    }
    set_error("unknown storage id " + std::to_string(st));
    return PLBM_ERR_ARG;
}

template <class F>
inline int dispatch_model(int m, F &&f) {
    switch (m) {
    case PLBM_BGK: return f(std::integral_constant<int, PLBM_BGK>{});
    case PLBM_TRT: return f(std::integral_constant<int, PLBM_TRT>{});
    case PLBM_RR: return f(std::integral_constant<int, PLBM_RR>{});
    case PLBM_BGK_SPLIT: return f(std::integral_constant<int, PLBM_BGK_SPLIT>{});
    }
    set_error("unknown collision model " + std::to_string(m));
    return PLBM_ERR_ARG;
}

constexpr int kMaxBuffers = 4;
constexpr int kModels = 4;
constexpr int kMaxStepsPerLaunch = 8;  // temporal-blocking depth of plbm_run
int max_steps_per_pass(int storage);  // PLBM_BGK, PLBM_TRT, PLBM_RR, PLBM_BGK_SPLIT
constexpr int kReduceBlocks = 1024;

// Host plane of nx columns with pitch ld (ld >= ny, the host arrays of the upload /
// download entry points) to / from a compact device plane (pitch ny), synchronous on
// return.  Only 1-D copies: a pitched host plane goes through `stage` (host memory).  A
// pitched 2-D asynchronous copy between device and pageable host memory delivered no data
// when the process ran under rocprofv3's kernel tracing (ROCm 7.2; the host array kept
// its prior contents), while 1-D copies did.
inline int plane_to_host(double *host, size_t ld, const double *dev, size_t ny, size_t nx,
                         hipStream_t s, std::vector<double> &stage) {
    const bool direct = ld == ny;
    if (!direct) stage.resize(ny * nx);
    PLBM_HIP(hipMemcpyAsync(direct ? host : stage.data(), dev, ny * nx * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    PLBM_HIP(hipStreamSynchronize(s));
    if (!direct)
        for (size_t x = 0; x < nx; ++x)
            std::copy(stage.data() + x * ny, stage.data() + (x + 1) * ny, host + x * ld);
    return PLBM_OK;
}
inline int plane_from_host(double *dev, const double *host, size_t ld, size_t ny, size_t nx,
                           hipStream_t s, std::vector<double> &stage) {
    const bool direct = ld == ny;
    if (!direct) {
        stage.resize(ny * nx);
        for (size_t x = 0; x < nx; ++x)
            std::copy(host + x * ld, host + x * ld + ny, stage.data() + x * ny);
    }
    PLBM_HIP(hipMemcpyAsync(dev, direct ? host : stage.data(), ny * nx * sizeof(double),
                            hipMemcpyHostToDevice, s));
    // the source (caller's array or stage) must outlive the copy
    PLBM_HIP(hipStreamSynchronize(s));
    return PLBM_OK;
}

}  // namespace plbm

struct plbm_lattice;

namespace plbm {
// One temporally blocked pass of ns (2..8) steps src -> dst; with lag non-null it also
// stores the state one step before the pass's last there (multistep.hip).
int launch_multi(plbm_lattice *h, int model, int ns, const void *src, void *dst,
                 void *lag = nullptr);
// Run bookkeeping shared by the pass runs (multistep.hip): pass depths of a run, the
// lattice's third buffer (nullptr when it cannot be allocated), and the final assignment
// of device storage to slots.
std::vector<int> pass_plan(int nsteps, int ns);
std::vector<int> lag_pass_plan(int nsteps, int ns, int lag_ns = 0);
void *pass_aux(plbm_lattice *h);
void finish_run(plbm_lattice *h, int a0, int b0, int nsteps, void *latest, void *lag,
                void *spare, int *iold, int *inew);

}  // namespace plbm

struct plbm_lattice {
    int nx = 0, ny = 0, ldd = 0, nf = 0, storage = 0, device = 0;
    hipStream_t stream = nullptr;  // the stream work is issued on
    hipStream_t own = nullptr;     // the lattice's own stream
    void *f[plbm::kMaxBuffers] = {};
    size_t elems = 0;              // elements per buffer = 9*nx*ldd
    double omega = 0, trt_magic = 0, lambda_odd = 0;
    bool params_set = false;
    double scale = 1.0;            // FP16 storage scale
    plbm::LaunchCfg cfg[plbm::kModels];  // fused-step launch shape per collision model
    bool tuned[plbm::kModels] = {};
    int dugks_ns = 0;                    // DUGKS steps per pass (0 = chosen from the grid)
    int fd_ns = 0;                       // FV/FD operator steps per pass of plbm_run_op
    int pass_segs = 0;                   // column segments per row band of a pass (0 = auto)
    // third population buffer of the pass runs (allocated on the first run in passes):
    // the last pass of a run writes the state one step before the latest into it
    void *aux = nullptr;
    bool aux_failed = false;             // its allocation failed once; runs end on a single step
    bool aux_enabled = true;             // plbm_set_pass_lag_buffer
    // per-launch timing of plbm_run (plbm_set_run_timing): events around each launch of
    // the last run and the steps each launch advanced
    bool run_timing = false;
    std::vector<hipEvent_t> run_events;
    std::vector<int> run_depths;
    // scratch (allocated on first use)
    double *d_tmp = nullptr;       // staging for host transfers
    size_t tmp_bytes = 0;
    double *d_partials = nullptr;  // kReduceBlocks * PLBM_NRED
    double *d_red = nullptr;       // PLBM_NRED
};
