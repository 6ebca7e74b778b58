// common.hpp — host-side plumbing shared by the engine's translation units:
// error reporting, HIP call checking, storage traits, the lattice object.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/plbm.h"

namespace plbm {

// Last error text of the calling thread (plbm_last_error).
void set_error(const std::string &msg);
const char *last_error();

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

// Per-launch scalars of one step, passed by value.
struct StepArgs {
    const void *src;
    void *dst;
    int nx, ny, ldd, nby;   // nby = y-blocks per column
    size_t plane;           // elements per q-plane = nx*ldd
    double omega, lambda_odd;
    float omega_f, lambda_odd_f;
    float scale, inv_scale; // FP16 storage scale (1 otherwise)
};

constexpr int kMaxBuffers = 4;
constexpr int kReduceBlocks = 1024;

}  // namespace plbm

struct plbm_lattice {
    int nx = 0, ny = 0, ldd = 0, nf = 0, storage = 0, device = 0;
    hipStream_t stream = nullptr;  // the stream work is issued on
    hipStream_t own = nullptr;     // the lattice's own stream
    void *f[plbm::kMaxBuffers] = {};
    size_t plane = 0;              // elements per q-plane
    double omega = 0, trt_magic = 0, lambda_odd = 0;
    bool params_set = false;
    double scale = 1.0;            // FP16 storage scale
    int block = 256;               // threads per block of the step kernel
    bool xcd_remap = true;
    // scratch (allocated on first use)
    double *d_tmp = nullptr;       // staging for host transfers
    size_t tmp_bytes = 0;
    double *d_partials = nullptr;  // kReduceBlocks * PLBM_NRED
    double *d_red = nullptr;       // PLBM_NRED
};
