// multistep.hip — temporal blocking: NS fused stream + collide steps per launch.
//
// The single-step kernel reads and writes every population once per step (18 s bytes per
// lattice update), so it is bound by HBM at one step per pass.  Here one pass advances the
// lattice NS steps (NS = 2..8) and reads / writes every population once per pass:
//
//  * A wavefront owns 64 consecutive rows y and walks a segment of columns along x (by
//    default 1/64 of the width, at least 8 columns).  Of its rows, VR produce output; NS
//    rows on each side are halo (VR + 2 NS <= 64).  VR is a multiple of 8 so that each
//    wave's stores start and end on 64-byte boundaries.
//  * Level k (k = 0 .. NS-1) keeps the pull-ready values of f^{n+k} for the last three
//    columns in registers.  Loading column c gives level 0 at c; level k+1 at column
//    c - k - 1 follows from level k at columns c - k, c - k - 1, c - k - 2 (the pull
//    f_q(x, y) <- f_q(x - cx_q, y - cy_q) takes x-neighbours from the window and
//    y-neighbours from the adjacent lanes), and level NS is stored.
//  * The next columns' loads are issued ahead (the pass is latency-bound without it:
//    each wavefront walks its strip sequentially): for 8-byte storage two columns ahead,
//    straight into LDS (16-byte global -> LDS loads, two rows per lane), and the pull in
//    y of level 0 reads the neighbouring row's LDS slot; for 4- and 2-byte storage one
//    column ahead into registers.
//  * Each level applies exactly the single-step kernel's per-cell arithmetic
//    (collide_cell) in the storage's arithmetic type, so for FP64, F64S and F32S storage
//    (arithmetic type = storage type) the result equals NS launches of the single-step
//    kernel bit for bit (F64S / F32S up to the compiler's FMA placement).  FP16 storage
//    computes in FP32 and is rounded to FP16 only when a pass stores it: the intermediate
//    states are never stored, so a pass rounds once where NS single steps round NS times
//    (rounding them as well cost 36 conversions per level: 1.83 against 1.32 ms per
//    8-step pass at 8192^2).
//
// Algorithmic traffic per pass: 18 s bytes per cell (one read, one write of each
// population) for NS lattice updates; the halo re-reads (64/VR in y, 2 NS columns per
// segment in x) are the overhead.  Work per output cell and pass: NS collisions plus the halo ones.
#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "multistep.hpp"

namespace plbm {

int launch_step(plbm_lattice *h, int model, const void *src, void *dst);

// The pass covers waves_y row bands x nx columns; each band is cut into m column segments
// of nearly equal length, and wavefront w walks segment w / waves_y of band w % waves_y,
// so that the wavefronts running at the same time hold adjacent bands of the same columns
// (their loads are adjacent in memory, and the rows two bands share are read while both
// are in L2).
//
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD b % 8); they are
// renumbered so that each XCD takes a contiguous range, and adjacent bands, which read
// each other's edge rows, share an XCD's L2 (same-box A/B at 8192^2 and 4096^2, TRT, 8
// steps per pass: +0.6 to +1.5% for every storage).
__device__ __forceinline__ int xcd_contiguous(int b, int nwg) {
    const int base = nwg / 8, rem = nwg % 8, x = b % 8;
    return x * base + min(x, rem) + b / 8;
}

template <int MODEL, int ST, int NS>
__global__ void __launch_bounds__(256) k_multi(StepArgs a, int waves_y, int m, int sync) {
    // wave-uniform: the column walk's indices and addresses are then scalar
    const int wg = xcd_contiguous(blockIdx.x, gridDim.x);
    const int gw = __builtin_amdgcn_readfirstlane(wg * 4 + (threadIdx.x >> 6));
    if (gw >= waves_y * m) return;  // the whole wavefront
    const int band = gw % waves_y, j = gw / waves_y;
    const int x0 = (int)((long long)a.nx * j / m), x1 = (int)((long long)a.nx * (j + 1) / m);
    // the unrolled steady state where it keeps the wavefronts per SIMD and measured faster
    // (8192^2 TRT, same-box A/B, ms per pass: FP64 6 steps 2.11 -> 2.03, 7 steps 2.40 ->
    // 2.26; FP32 storage 8 steps 1.45 -> 1.43, FP16 1.40 -> 1.37, their 20-step runs
    // 1.5% faster; FP32 storage at 6 steps and BGK measured slower with it, FP64 at 8
    // steps needs more than 256 registers)
    constexpr bool rot = MODEL == PLBM_TRT && ((ST == PLBM_F64 && (NS == 6 || NS == 7)) ||
                                              ((ST == PLBM_F32S || ST == PLBM_F16S) && NS == 8));
    if (x1 > x0) multi_wave<MODEL, ST, NS, rot>(a, band, x0, x1 - x0, sync != 0);
}

int max_steps_per_pass(int) { return kMaxPassSteps; }

// Column segments per row band of a pass when the caller has not chosen: 64, each of at
// least 8 columns.  Measured with 8 steps per pass (TRT, GLUPS for 48 / 64 / 96 segments
// per band; FP64, FP32, FP16 storage): 8192^2 181 / 191 / 184, 351 / 366 / 358, 360 / 378
// / 365; 4096^2 168 / 174 / 168, 305 / 326 / 316, 273 / 322 / 333; FP64 2048^2 (32 / 64
// / 128) 145 / 154 / 136, 1024^2 84 / 102 / 94, 512^2 (16 / 32 / 64) 24 / 38 / 53.  With
// fewer wavefronts than resident slots the short segments still pay: the pass is bound by
// each wavefront's sequential walk.
int default_pass_segments(int nx) { return std::max(1, std::min(64, nx / 8)); }

int launch_multi(plbm_lattice *h, int model, int ns, const void *src, void *dst,
                 void *lag) {
    StepArgs a{};
    a.src = src;
    a.dst = dst;
    a.lag = lag;
    a.nx = h->nx;
    a.ny = h->ny;
    a.ldd = h->ldd;
    a.cs = (size_t)9 * h->ldd;
    a.omega = h->omega;
    a.lambda_odd = h->lambda_odd;
    a.omega_f = float(h->omega);
    a.lambda_odd_f = float(h->lambda_odd);
    a.scale = float(h->scale);
    a.inv_scale = float(1.0 / h->scale);
    return dispatch_model(model, [&](auto M) {
        return dispatch_storage(h->storage, [&](auto S) {
            constexpr int MM = decltype(M)::value, SS = decltype(S)::value;
            auto go = [&](auto K) {
                constexpr int NS = decltype(K)::value;
                if constexpr (NS <= max_steps_per_pass<SS>()) {
                    constexpr int vr = multi_vr<SS, NS>();
                    const int waves_y = (h->ny + vr - 1) / vr;
                    const int m = h->pass_segs > 0 ? std::min(h->pass_segs, h->nx)
                                                   : default_pass_segments(h->nx);
                    const long long nw = (long long)waves_y * m;
                    const dim3 grid((unsigned)((nw + 3) / 4));
                    // A barrier per column keeps a workgroup's four wavefronts (adjacent
                    // bands) in step.  It pays in the pass that also stores the lagged level,
                    // bound by its two store streams, and not elsewhere (8192^2 TRT FP64,
                    // same-box A/B, ms per launch without / with: lag-writing pass of 8 steps
                    // 2.99 / 2.92, pass of 6 steps 1.99 / 1.99, pass of 8 steps 2.62 / 2.69).
                    // It needs every wavefront of every workgroup to walk the same number
                    // of columns: equal segments and no idle wavefronts.
                    const int sync = lag != nullptr && h->nx % m == 0 && nw % 4 == 0;
                    hipLaunchKernelGGL((k_multi<MM, SS, NS>), grid, dim3(256), 0, h->stream, a,
                                       waves_y, m, sync);
                }
                return NS <= max_steps_per_pass<SS>();
            };
            bool ok = false;
            switch (ns) {
            case 2: ok = go(std::integral_constant<int, 2>{}); break;
            case 3: ok = go(std::integral_constant<int, 3>{}); break;
            case 4: ok = go(std::integral_constant<int, 4>{}); break;
            case 5: ok = go(std::integral_constant<int, 5>{}); break;
            case 6: ok = go(std::integral_constant<int, 6>{}); break;
            case 7: ok = go(std::integral_constant<int, 7>{}); break;
            case 8: ok = go(std::integral_constant<int, 8>{}); break;
            default: break;
            }
            if (!ok) {
                set_error("steps per pass " + std::to_string(ns) + " out of range for this storage");
                return PLBM_ERR_ARG;
            }
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        });
    });
}

// Depths of the passes of nsteps >= 1 steps with passes of at most ns >= 1 steps: the
// fewest passes, their depths differing by at most one, the deepest last.  A pass of
// depth d costs about one HBM sweep for every d up to 6 at 8192^2 (TRT FP64, us per
// launch for d = 2..8: 1969 2004 1995 1952 1992 2249 2447), so 12 steps as 6 + 6 cost
// less than 8 + 4 (3984 against 4442 us).
std::vector<int> pass_plan(int nsteps, int ns) {
    const int passes = (nsteps + ns - 1) / ns;
    std::vector<int> d((size_t)passes, nsteps / passes);
    for (int p = 0; p < nsteps % passes; ++p) d[(size_t)(passes - 1 - p)] += 1;
    return d;
}

// Depths of a run of nsteps >= 2 steps whose last pass also stores the state one step
// before the latest: that pass writes the whole state twice and is bound by HBM writes
// whatever its depth (8192^2 TRT FP64, ms per launch for d = 2..8 with / without the
// second store, before that pass had its per-column barrier: 3.02 / 1.97, 3.08 / 2.02,
// 3.03 / 2.01, 2.93 / 1.98, 2.92 / 2.06, 3.00 / 2.40, 2.99 / 2.62), so it takes as many
// steps as a pass holds and the rest is split as
// pass_plan does: 20 steps at depth 8 as 6 + 6 + 8 (7.1 ms) instead of 6 + 7 + 7 (7.5 ms).
// lag_ns (> 0) sets that last pass's depth apart from ns (the autotuner's choice).
std::vector<int> lag_pass_plan(int nsteps, int ns, int lag_ns) {
    const int last = std::min(lag_ns > 0 ? lag_ns : ns, nsteps);
    std::vector<int> d = nsteps > last ? pass_plan(nsteps - last, ns) : std::vector<int>();
    d.push_back(last);
    return d;
}

// The third buffer that the last pass of a run writes the lagged state into.  Allocated
// on the first run in passes; when that fails (a grid that fits two buffers but not
// three), or the caller has switched it off (plbm_set_pass_lag_buffer), runs end on a
// single-step launch instead, as the state-lag rule needs.
void *pass_aux(plbm_lattice *h) {
    if (!h->aux_enabled) return nullptr;
    if (h->aux || h->aux_failed) return h->aux;
    if (hipMalloc(&h->aux, h->elems * storage_bytes(h->storage)) != hipSuccess) {
        (void)hipGetLastError();
        h->aux = nullptr;
        h->aux_failed = true;
    }
    return h->aux;
}

// Ends a run whose latest state is at `latest` and whose state one step earlier is at
// `lag` (`spare` is the third buffer, contents no longer needed): slot ids follow the
// reference's swap rule (after nsteps swaps the latest state is in slot b0 when nsteps is
// odd), and the device storage of slots a0 / b0 and of the spare buffer is permuted to
// match.
void finish_run(plbm_lattice *h, int a0, int b0, int nsteps, void *latest, void *lag,
                void *spare, int *iold, int *inew) {
    const int want = (nsteps % 2) ? b0 : a0;
    const int other = want == a0 ? b0 : a0;
    h->f[want] = latest;
    h->f[other] = lag;
    if (spare != nullptr) h->aux = spare;
    *iold = want;
    *inew = other;
}

// nsteps steps of `model` with the reference's bookkeeping (each step reads *iold, writes
// *inew, swaps them).  With ns > 1 the steps run in passes of up to ns steps
// (lag_pass_plan);
// the last pass also stores the state one step before the latest into the third buffer,
// which becomes the other slot's storage, as update_macros expects of it (the reference's
// inew after perform_lbm_step).  Without a third buffer the last step is a single-step
// launch instead.
// With run timing on, an event is recorded on the lattice's stream before the first
// launch of a run and after each launch (plbm_run_launch_times reads them).
static int timing_mark(plbm_lattice *h, int depth) {
    if (!h->run_timing) return PLBM_OK;
    const size_t k = h->run_depths.size() + (depth > 0 ? 1 : 0);
    while (h->run_events.size() <= k) {
        hipEvent_t e;
        PLBM_HIP(hipEventCreate(&e));
        h->run_events.push_back(e);
    }
    if (depth > 0) h->run_depths.push_back(depth);
    PLBM_HIP(hipEventRecord(h->run_events[k], h->stream));
    return PLBM_OK;
}

int run_steps(plbm_lattice *h, int model, int ns, int lag_ns, int nsteps, int *iold,
              int *inew) {
    h->run_depths.clear();
    PLBM_TRY(timing_mark(h, 0));
    if (ns <= 1 || nsteps < 2) {
        for (int s = 0; s < nsteps; ++s) {
            PLBM_TRY(launch_step(h, model, h->f[*iold], h->f[*inew]));
            PLBM_TRY(timing_mark(h, 1));
            std::swap(*iold, *inew);
        }
        return PLBM_OK;
    }
    const int a0 = *iold, b0 = *inew;
    void *latest = h->f[a0], *other = h->f[b0];
    void *aux = pass_aux(h);
    auto launch = [&](int k, void *lag) {
        PLBM_TRY(k == 1 ? launch_step(h, model, latest, other)
                        : launch_multi(h, model, k, latest, other, lag));
        return timing_mark(h, k);
    };
    if (aux == nullptr) {
        for (int k : pass_plan(nsteps - 1, ns)) {
            PLBM_TRY(launch(k, nullptr));
            std::swap(latest, other);
        }
        PLBM_TRY(launch(1, nullptr));
        finish_run(h, a0, b0, nsteps, other, latest, nullptr, iold, inew);
        return PLBM_OK;
    }
    const std::vector<int> plan = lag_pass_plan(nsteps, ns, lag_ns);
    for (size_t p = 0; p + 1 < plan.size(); ++p) {
        PLBM_TRY(launch(plan[p], nullptr));
        std::swap(latest, other);
    }
    PLBM_TRY(launch(plan.back(), aux));
    finish_run(h, a0, b0, nsteps, other, aux, latest, iold, inew);
    return PLBM_OK;
}

}  // namespace plbm

namespace plbm {

// Steps per pass when the caller has not chosen: 8 where the default work split gives at
// least 512 wavefronts, else single steps.  TRT, 8 steps per pass against single steps,
// GLUPS (FP64): 512^2 52.6 / 31.9, 1024^2 101.5 / 35.4, 2048^2 153.6 / 40.8, 4096^2 187
// / 42.5; 320^2 (35 wavefronts of 64 columns) 36 against 5 us per step.  At 8192^2 8 is
// the fastest depth for every storage.
int default_steps_per_launch(const plbm_lattice *h) {
    const int vr = multi_rows(storage_bytes(h->storage), kMaxPassSteps);
    const long waves = (long)((h->ny + vr - 1) / vr) * default_pass_segments(h->nx);
    return waves >= 512 ? kMaxPassSteps : 1;
}

}  // namespace plbm

using namespace plbm;

extern "C" int plbm_pass(plbm_lattice *h, int model, int steps, int src, int dst) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG,
                 "unknown collision model " + std::to_string(model));
    PLBM_REQUIRE(steps >= 1 && steps <= max_steps_per_pass(h->storage), PLBM_ERR_ARG,
                 "steps per pass must be 1.." + std::to_string(max_steps_per_pass(h->storage)) +
                     " for this storage");
    PLBM_REQUIRE(src >= 0 && src < h->nf && dst >= 0 && dst < h->nf && src != dst, PLBM_ERR_ARG,
                 "need distinct valid src and dst buffers");
    PLBM_REQUIRE(h->params_set, PLBM_ERR_STATE,
                 "collision parameters not set: call plbm_set_params first");
    if (steps == 1) return launch_step(h, model, h->f[src], h->f[dst]);
    return launch_multi(h, model, steps, h->f[src], h->f[dst]);
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_set_pass_segments(plbm_lattice *h, int segments) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(segments >= 0, PLBM_ERR_ARG, "segments must be >= 0 (0 = automatic)");
    h->pass_segs = segments;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_set_pass_lag_buffer(plbm_lattice *h, int enable) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(enable == 0 || enable == 1, PLBM_ERR_ARG, "enable must be 0 or 1");
    h->aux_enabled = enable != 0;
    if (!h->aux_enabled && h->aux) {
        PLBM_HIP(hipStreamSynchronize(h->stream));
        PLBM_HIP(hipFree(h->aux));
        h->aux = nullptr;
    }
    if (h->aux_enabled) h->aux_failed = false;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_set_run_timing(plbm_lattice *h, int enable) try {
    PLBM_REQUIRE(h, PLBM_ERR_ARG, "null lattice handle");
    PLBM_REQUIRE(enable == 0 || enable == 1, PLBM_ERR_ARG, "enable must be 0 or 1");
    h->run_timing = enable != 0;
    h->run_depths.clear();
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_run_launch_times(plbm_lattice *h, int max_launches, int *depths, float *ms,
                                     int *nlaunches) try {
    PLBM_REQUIRE(h && nlaunches && max_launches >= 0, PLBM_ERR_ARG, "bad argument");
    PLBM_REQUIRE(h->run_timing, PLBM_ERR_STATE, "run timing is off (plbm_set_run_timing)");
    const int n = (int)h->run_depths.size();
    *nlaunches = n;
    if (n == 0) return PLBM_OK;
    PLBM_HIP(hipEventSynchronize(h->run_events[(size_t)n]));
    for (int k = 0; k < n && k < max_launches; ++k) {
        if (depths) depths[k] = h->run_depths[(size_t)k];
        if (ms) PLBM_HIP(hipEventElapsedTime(&ms[k], h->run_events[(size_t)k],
                                             h->run_events[(size_t)k + 1]));
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

extern "C" int plbm_get_pass_segments(const plbm_lattice *h, int *segments) try {
    PLBM_REQUIRE(h && segments, PLBM_ERR_ARG, "null argument");
    *segments = h->pass_segs;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}
