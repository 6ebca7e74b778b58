// batch.hip — many independent periodic lattices of uneven size advanced together.
//
// All cases live in one device arena, each with an A/B pair of population buffers in the
// single-lattice layout (column-interleaved, element (q,x,y) at (x*9+q)*ldd + y, ldd =
// ny rounded up to 32).  Two kernels share the work of one step:
//
//  * k_batch_step: one launch per step over a flat tile table (tile = ty consecutive y of
//    tx columns of one case), tiles of the largest cases first.  This is the HBM-bound
//    path for the big cases; consecutive steps are captured into a hipGraph and replayed,
//    so a 2000-step run is a handful of graph launches.
//  * k_batch_resident: persistent, one workgroup per small case (nx*ny <= resident_cells)
//    that runs every step of the launch itself, with only workgroup barriers between
//    steps (no dependence between workgroups).  A small case's A/B pair (< 1 MB) stays in
//    the XCD's L2 for the whole run instead of streaming through HBM every step.  It is
//    captured into the same graph on a forked stream, so it overlaps the big-case steps.
//
//  * k_batch_multi (steps per pass ns > 1): the tiled cases advance ns steps per launch
//    with the temporally blocked pass of the single lattice (multistep.hpp), one
//    wavefront per 56 x 64 (FP64, up to 4 steps) or 48 x 64 block of a case, listed in a
//    wave table in place of the tile table.
//
// Tile shape (ty, tx), steps per pass, columns per pass segment and the resident threshold
// are picked by the project's autotuner (plbm_batch_autotune), which times candidates on a scratch copy of
// the arena.
#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "d2q9.hpp"
#include "multistep.hpp"
#include "reduce.hpp"
#include "stepper.hpp"

namespace plbm {

struct BatchCase {
    int nx, ny, ldd, pad;
    unsigned long long off;    // element offset of buffer A in the arena; B follows
    unsigned long long elems;  // elements per buffer = 9*nx*ldd
    double omega, lambda_odd;
    float omega_f, lambda_odd_f;
};

struct BatchTile {
    int c, x0, y0, n;  // case, first column, first row (row band with passes), columns
};

struct BatchCfg {
    int ty = 128, tx = 1;      // tile = ty x tx cells, one thread each
    int resident_cells = 0;    // cases with nx*ny <= this run in k_batch_resident
    int resident_block = 256;
    int ns = 1;                // steps per pass of the tiled cases (> 1: k_batch_multi)
    int xs = kMultiXS;         // columns per wavefront segment of a pass
};

template <int ST>
__device__ __forceinline__ StepArgs case_args(const BatchCase &c, float scale, float inv_scale) {
    StepArgs a;
    a.src = nullptr;
    a.dst = nullptr;
    a.nx = c.nx;
    a.ny = c.ny;
    a.ldd = c.ldd;
    a.nby = 0;
    a.cs = (size_t)9 * c.ldd;
    a.remap = 0;
    a.omega = c.omega;
    a.lambda_odd = c.lambda_odd;
    a.omega_f = c.omega_f;
    a.lambda_odd_f = c.lambda_odd_f;
    a.scale = scale;
    a.inv_scale = inv_scale;
    return a;
}

// One cell of one case: pull from src, collide, store to dst (periodic in the case).
template <int MODEL, int ST, bool NT>
__device__ __forceinline__ void batch_cell(const typename Traits<ST>::T *__restrict__ s,
                                           typename Traits<ST>::T *__restrict__ d,
                                           const StepArgs &a, int x, int y) {
    using R = typename Traits<ST>::R;
    const int nx = a.nx, ny = a.ny;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const size_t L = (size_t)a.ldd;
    const size_t c0 = (size_t)x * a.cs, cm = (size_t)xm * a.cs, cp = (size_t)xp * a.cs;
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
    for (int q = 0; q < 9; ++q) store_pop<ST, NT>(d + c0 + q * L + y, f[q], a.scale);
}

// One step of every tiled case.  parity 0: A -> B, 1: B -> A.
template <int MODEL, int ST>
__global__ void __launch_bounds__(1024)
    k_batch_step(const BatchCase *__restrict__ cases, const BatchTile *__restrict__ tiles,
                 void *arena, int parity, int ty, float scale, float inv_scale) {
    using T = typename Traits<ST>::T;
    const BatchTile t = tiles[blockIdx.x];
    const BatchCase c = cases[t.c];
    const int y = t.y0 + (int)threadIdx.x % ty;
    const int x = t.x0 + (int)threadIdx.x / ty;
    if (x >= c.nx || y >= c.ny) return;
    T *base = static_cast<T *>(arena) + c.off;
    const T *s = parity ? base + c.elems : base;
    T *d = parity ? base : base + c.elems;
    const StepArgs a = case_args<ST>(c, scale, inv_scale);
    batch_cell<MODEL, ST, true>(s, d, a, x, y);
}

// One pass of NS steps of every tiled case; waves[i] = {case, strip, row block}, four
// wavefronts per workgroup.  parity 0: A -> B, 1: B -> A.
template <int MODEL, int ST, int NS>
__global__ void __launch_bounds__(256)
    k_batch_multi(const BatchCase *__restrict__ cases, const BatchTile *__restrict__ waves,
                  int nwaves, void *arena, int parity, float scale, float inv_scale) {
    using T = typename Traits<ST>::T;
    const int gw = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (gw >= nwaves) return;  // the whole wavefront
    const BatchTile w = waves[gw];
    const BatchCase c = cases[w.c];
    T *base = static_cast<T *>(arena) + c.off;
    StepArgs a = case_args<ST>(c, scale, inv_scale);
    a.src = parity ? base + c.elems : base;
    a.dst = parity ? base : base + c.elems;
    multi_wave<MODEL, ST, NS>(a, w.y0, w.x0, w.n);
}

// nsteps steps of one small case per workgroup; only workgroup barriers between steps.
template <int MODEL, int ST>
__global__ void __launch_bounds__(1024)
    k_batch_resident(const BatchCase *__restrict__ cases, const int *__restrict__ list,
                     void *arena, int parity0, int nsteps, float scale, float inv_scale) {
    using T = typename Traits<ST>::T;
    const BatchCase c = cases[list[blockIdx.x]];
    T *base = static_cast<T *>(arena) + c.off;
    const StepArgs a = case_args<ST>(c, scale, inv_scale);
    const int ncell = c.nx * c.ny;
    int parity = parity0;
    for (int s = 0; s < nsteps; ++s) {
        const T *src = parity ? base + c.elems : base;
        T *dst = parity ? base : base + c.elems;
        for (int i = threadIdx.x; i < ncell; i += blockDim.x)
            batch_cell<MODEL, ST, false>(src, dst, a, i / c.ny, i % c.ny);
        __syncthreads();  // this step's stores are visible to the whole workgroup
        parity ^= 1;
    }
}

// Per-case FP64 reductions of the buffer holding the latest state (one workgroup per
// case, reduce.hpp).
template <int ST>
__global__ void __launch_bounds__(kRedThreads)
    k_batch_reduce(const BatchCase *__restrict__ cases, const void *arena, int parity,
                   float inv_scale, double *out) {
    using T = typename Traits<ST>::T;
    const BatchCase c = cases[blockIdx.x];
    const T *s = static_cast<const T *>(arena) + c.off + (parity ? c.elems : 0);
    double acc[PLBM_NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int n = c.nx * c.ny, sx = kRedThreads / c.ny, sy = kRedThreads % c.ny;
    CellWalk w;
    w.start(threadIdx.x, c.ny);
    for (int i = threadIdx.x; i < n; i += kRedThreads) {
        double f[9];
        const size_t base = pop_index(0, w.x, w.y, c.ldd);
#pragma unroll
        for (int q = 0; q < 9; ++q) f[q] = double(load_pop<ST>(s + base + (size_t)q * c.ldd, inv_scale));
        double rho, ux, uy;
        if constexpr (!Traits<ST>::shifted) {
            macros_ref(f, rho, ux, uy);
        } else {
            double dr, jx, jy;
            moments_shifted<double>(f, dr, jx, jy);
            rho = 1.0 + dr;
            ux = jx / rho;
            uy = jy / rho;
        }
        red_cell(acc, rho, ux, uy, nullptr, nullptr, 0);
        w.next(sx, sy, c.ny);
    }
    const double v = red_block(acc);
    if (threadIdx.x < PLBM_NRED) out[(size_t)blockIdx.x * PLBM_NRED + threadIdx.x] = v;
}

// Host-layout <-> arena conversion of one q-plane of one case (as in the single lattice).
template <int ST>
__global__ void __launch_bounds__(256)
    k_batch_from_f64(const double *src, void *dst, int q, int nx, int ny, int ldd, double w,
                     float scale) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (y >= ny) return;
    using T = typename Traits<ST>::T;
    const double v = src[(size_t)x * ny + y];
    T *d = static_cast<T *>(dst);
    const size_t o = pop_index(q, x, y, ldd);
    if constexpr (ST == PLBM_F16S) d[o] = T(float((v - w) * double(scale)));
    else if constexpr (Traits<ST>::shifted) d[o] = T(v - w);
    else d[o] = T(v);
}

template <int ST>
__global__ void __launch_bounds__(256)
    k_batch_to_f64(const void *src, double *dst, int q, int nx, int ny, int ldd, double w,
                   float inv_scale) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (y >= ny) return;
    using T = typename Traits<ST>::T;
    const double v = double(static_cast<const T *>(src)[pop_index(q, x, y, ldd)]);
    if constexpr (ST == PLBM_F16S) dst[(size_t)x * ny + y] = v * double(inv_scale) + w;
    else if constexpr (Traits<ST>::shifted) dst[(size_t)x * ny + y] = v + w;
    else dst[(size_t)x * ny + y] = v;
}

template <int ST>
__global__ void __launch_bounds__(256)
    k_batch_equilibrium(const double *rho, const double *ux, const double *uy, void *dst,
                        int nx, int ny, int ldd, float scale) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x, x = blockIdx.y;
    if (y >= ny) return;
    using T = typename Traits<ST>::T;
    const size_t i = (size_t)x * ny + y, c = pop_index(0, x, y, ldd);
    double feq[9];
    if constexpr (!Traits<ST>::shifted) equilibrium_ref(rho[i], ux[i], uy[i], feq);
    else equilibrium_shifted<double>(rho[i] - 1.0, ux[i], uy[i], feq);
    T *d = static_cast<T *>(dst);
#pragma unroll
    for (int q = 0; q < 9; ++q) store_pop<ST, false>(d + c + (size_t)q * ldd, feq[q], scale);
}

}  // namespace plbm

using namespace plbm;

struct plbm_batch {
    int ncases = 0, storage = 0, model = 0, device = 0;
    std::vector<BatchCase> cases;          // host copy
    BatchCase *d_cases = nullptr;
    void *arena = nullptr;
    size_t arena_elems = 0;
    int parity = 0;                         // 0: latest state in A
    double scale = 1.0;
    bool params_set = false;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    BatchCfg cfg;
    // launch plan for cfg (rebuilt when cfg changes)
    BatchTile *d_tiles = nullptr;
    int ntiles = 0;
    int *d_resident = nullptr;
    int nresident = 0;
    // cached graphs: key = (nsteps, parity at start)
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;
    double *d_red = nullptr;
    double *d_tmp = nullptr;
    size_t tmp_bytes = 0;
};

namespace plbm {

static void drop_graphs(plbm_batch *b) {
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

// Tile table and resident list for the current configuration.
static int build_plan(plbm_batch *b, const BatchCfg &cfg, BatchTile **d_tiles, int *ntiles,
                      int **d_res, int *nres) {
    std::vector<int> order(b->ncases);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int i, int j) {
        return (size_t)b->cases[i].nx * b->cases[i].ny > (size_t)b->cases[j].nx * b->cases[j].ny;
    });
    std::vector<BatchTile> tiles;
    std::vector<int> res;
    for (int c : order) {
        const BatchCase &k = b->cases[c];
        if ((long long)k.nx * k.ny <= cfg.resident_cells) {
            res.push_back(c);
            continue;
        }
        if (cfg.ns > 1) {  // wave table: {case, first column, row band, columns}
            const int vr = multi_rows(storage_bytes(b->storage), cfg.ns);
            for (int x0 = 0; x0 < k.nx; x0 += cfg.xs)
                for (int wy = 0; wy * vr < k.ny; ++wy)
                    tiles.push_back({c, x0, wy, std::min(cfg.xs, k.nx - x0)});
        } else {
            for (int x0 = 0; x0 < k.nx; x0 += cfg.tx)
                for (int y0 = 0; y0 < k.ny; y0 += cfg.ty) tiles.push_back({c, x0, y0, 0});
        }
    }
    *ntiles = (int)tiles.size();
    *nres = (int)res.size();
    *d_tiles = nullptr;
    *d_res = nullptr;
    auto fail = [&](hipError_t e) {
        if (*d_tiles) (void)hipFree(*d_tiles);
        if (*d_res) (void)hipFree(*d_res);
        *d_tiles = nullptr;
        *d_res = nullptr;
        set_error(std::string("batch launch plan: ") + hipGetErrorString(e));
        return PLBM_ERR_HIP;
    };
    hipError_t e;
    if (!tiles.empty()) {
        if ((e = hipMalloc(d_tiles, tiles.size() * sizeof(BatchTile))) != hipSuccess) return fail(e);
        if ((e = hipMemcpy(*d_tiles, tiles.data(), tiles.size() * sizeof(BatchTile),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e);
    }
    if (!res.empty()) {
        if ((e = hipMalloc(d_res, res.size() * sizeof(int))) != hipSuccess) return fail(e);
        if ((e = hipMemcpy(*d_res, res.data(), res.size() * sizeof(int),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e);
    }
    return PLBM_OK;
}

static int set_cfg(plbm_batch *b, const BatchCfg &cfg) {
    BatchTile *t = nullptr;
    int *r = nullptr;
    int nt = 0, nr = 0;
    PLBM_TRY(build_plan(b, cfg, &t, &nt, &r, &nr));
    drop_graphs(b);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_resident) (void)hipFree(b->d_resident);
    b->d_tiles = t;
    b->ntiles = nt;
    b->d_resident = r;
    b->nresident = nr;
    b->cfg = cfg;
    return PLBM_OK;
}

// Pass depths (each <= ns) summing to nsteps whose count has the parity of nsteps: the
// fewest passes of that parity, the steps spread evenly over them (58 steps at ns = 6:
// 6 passes of 6 and 4 of 5, rather than 9 of 6, one of 3 and one of 1; a short pass costs
// nearly a full sweep).
static std::vector<int> pass_depths(int ns, int nsteps) {
    std::vector<int> d;
    if (nsteps <= 0) return d;
    int p = (nsteps + ns - 1) / ns;
    if (p % 2 != nsteps % 2) ++p;  // still <= nsteps: p = nsteps has the parity of nsteps
    const int q = nsteps / p, r = nsteps % p;
    for (int i = 0; i < p; ++i) d.push_back(q + (i < r ? 1 : 0));
    return d;
}

// Enqueue `nsteps` steps starting at `parity` on b->stream (resident cases on b->side).
static int enqueue_steps(plbm_batch *b, void *arena, const BatchTile *tiles, int ntiles,
                         const int *res, int nres, const BatchCfg &cfg, int parity,
                         int nsteps) {
    const float sc = float(b->scale), isc = float(1.0 / b->scale);
    return dispatch_model(b->model, [&](auto M) {
        return dispatch_storage(b->storage, [&](auto S) {
            constexpr int MM = decltype(M)::value, SS = decltype(S)::value;
            if (nres > 0) {
                PLBM_HIP(hipEventRecord(b->fork, b->stream));
                PLBM_HIP(hipStreamWaitEvent(b->side, b->fork, 0));
                hipLaunchKernelGGL((k_batch_resident<MM, SS>), dim3(nres), dim3(cfg.resident_block),
                                   0, b->side, b->d_cases, res, arena, parity, nsteps, sc, isc);
                PLBM_HIP(hipGetLastError());
            }
            if (ntiles > 0 && cfg.ns <= 1) {
                for (int s = 0; s < nsteps; ++s) {
                    hipLaunchKernelGGL((k_batch_step<MM, SS>), dim3(ntiles), dim3(cfg.ty * cfg.tx), 0,
                                       b->stream, b->d_cases, tiles, arena, (parity + s) & 1,
                                       cfg.ty, sc, isc);
                    PLBM_HIP(hipGetLastError());
                }
            } else if (ntiles > 0) {
                // passes of up to ns steps; every pass moves the state to the other buffer,
                // so the number of passes must have the parity of nsteps (as the resident
                // cases, which flip once per step): split a pass in two when it has not
                for (int d : pass_depths(cfg.ns, nsteps)) {
                    const dim3 grid((unsigned)((ntiles + 3) / 4));
                    auto go = [&](auto K) {
                        constexpr int NS = decltype(K)::value;
                        if constexpr (NS <= max_steps_per_pass<SS>())
                            hipLaunchKernelGGL((k_batch_multi<MM, SS, NS>), grid, dim3(256), 0,
                                               b->stream, b->d_cases, tiles, ntiles, arena,
                                               parity, sc, isc);
                    };
                    switch (d) {
                    case 1: go(std::integral_constant<int, 1>{}); break;
                    case 2: go(std::integral_constant<int, 2>{}); break;
                    case 3: go(std::integral_constant<int, 3>{}); break;
                    case 4: go(std::integral_constant<int, 4>{}); break;
                    case 5: go(std::integral_constant<int, 5>{}); break;
                    case 6: go(std::integral_constant<int, 6>{}); break;
                    case 7: go(std::integral_constant<int, 7>{}); break;
                    default: go(std::integral_constant<int, 8>{}); break;
                    }
                    PLBM_HIP(hipGetLastError());
                    parity ^= 1;
                }
            }
            if (nres > 0) {
                PLBM_HIP(hipEventRecord(b->join, b->side));
                PLBM_HIP(hipStreamWaitEvent(b->stream, b->join, 0));
            }
            return PLBM_OK;
        });
    });
}

static int graph_for(plbm_batch *b, int nsteps, int parity, hipGraphExec_t *out) {
    auto key = std::make_pair(nsteps, parity);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return PLBM_OK;
    }
    hipGraph_t g = nullptr;
    PLBM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int st = enqueue_steps(b, b->arena, b->d_tiles, b->ntiles, b->d_resident, b->nresident,
                           b->cfg, parity, nsteps);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != PLBM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return PLBM_ERR_HIP;
    }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        return PLBM_ERR_HIP;
    }
    b->graphs[key] = ex;
    *out = ex;
    return PLBM_OK;
}

static int check_case(const plbm_batch *b, int c) {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_REQUIRE(c >= 0 && c < b->ncases, PLBM_ERR_ARG,
                 "case " + std::to_string(c) + " out of range [0," + std::to_string(b->ncases) + ")");
    return PLBM_OK;
}

static int ensure_btmp(plbm_batch *b, size_t bytes) {
    if (b->tmp_bytes >= bytes) return PLBM_OK;
    if (b->d_tmp) (void)hipFree(b->d_tmp);
    b->d_tmp = nullptr;
    b->tmp_bytes = 0;
    PLBM_HIP(hipMalloc(&b->d_tmp, bytes));
    b->tmp_bytes = bytes;
    return PLBM_OK;
}

}  // namespace plbm

extern "C" {

int plbm_batch_create(int ncases, const int *nx, const int *ny, int storage, int model,
                      plbm_batch **out) try {
    PLBM_REQUIRE(out, PLBM_ERR_ARG, "null output pointer");
    *out = nullptr;
    PLBM_REQUIRE(ncases >= 1 && nx && ny, PLBM_ERR_ARG, "need ncases >= 1 and size arrays");
    PLBM_REQUIRE(storage >= PLBM_F64 && storage <= PLBM_F16S, PLBM_ERR_ARG, "unknown storage");
    PLBM_REQUIRE(model >= PLBM_BGK && model <= PLBM_BGK_SPLIT, PLBM_ERR_ARG, "unknown collision model");
    auto *b = new plbm_batch();
    b->ncases = ncases;
    b->storage = storage;
    b->model = model;
    b->scale = storage == PLBM_F16S ? 1024.0 : 1.0;
    size_t off = 0;
    for (int c = 0; c < ncases; ++c) {
        if (nx[c] < 1 || ny[c] < 1 || (long long)nx[c] * ny[c] > (1ll << 30)) {
            set_error("case " + std::to_string(c) + ": invalid grid size");
            delete b;
            return PLBM_ERR_ARG;
        }
        BatchCase k{};
        k.nx = nx[c];
        k.ny = ny[c];
        k.ldd = (ny[c] + 31) / 32 * 32;
        k.off = off;
        k.elems = (unsigned long long)9 * k.nx * k.ldd;
        off += 2 * k.elems;
        b->cases.push_back(k);
    }
    b->arena_elems = off;
    auto fail = [&](const std::string &m, int code) {
        set_error(m);
        (void)hipGetLastError();
        plbm_batch_destroy(b);
        return code;
    };
    if (hipGetDevice(&b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
        return fail("no usable HIP device / stream", PLBM_ERR_HIP);
    if (hipMalloc(&b->arena, off * storage_bytes(storage)) != hipSuccess)
        return fail("hipMalloc of the batch arena failed", PLBM_ERR_NOMEM);
    if (hipMalloc(&b->d_cases, ncases * sizeof(BatchCase)) != hipSuccess ||
        hipMalloc(&b->d_red, (size_t)ncases * PLBM_NRED * sizeof(double)) != hipSuccess)
        return fail("hipMalloc of batch tables failed", PLBM_ERR_NOMEM);
    if (hipMemcpy(b->d_cases, b->cases.data(), ncases * sizeof(BatchCase),
                  hipMemcpyHostToDevice) != hipSuccess)
        return fail("copy of the case table failed", PLBM_ERR_HIP);
    int st = set_cfg(b, BatchCfg());
    if (st != PLBM_OK) {
        plbm_batch_destroy(b);
        return st;
    }
    *out = b;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_destroy(plbm_batch *b) try {
    if (!b) return PLBM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    drop_graphs(b);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_resident) (void)hipFree(b->d_resident);
    if (b->arena) (void)hipFree(b->arena);
    if (b->d_cases) (void)hipFree(b->d_cases);
    if (b->d_red) (void)hipFree(b->d_red);
    if (b->d_tmp) (void)hipFree(b->d_tmp);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_info(const plbm_batch *b, int *ncases, long long *cells, size_t *arena_bytes) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    long long n = 0;
    for (const auto &k : b->cases) n += (long long)k.nx * k.ny;
    if (ncases) *ncases = b->ncases;
    if (cells) *cells = n;
    if (arena_bytes) *arena_bytes = b->arena_elems * storage_bytes(b->storage);
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_set_params(plbm_batch *b, const double *omega, const double *trt_magic) try {
    PLBM_REQUIRE(b && omega, PLBM_ERR_ARG, "null argument");
    for (int c = 0; c < b->ncases; ++c) {
        const double m = trt_magic ? trt_magic[c] : 0.25;
        PLBM_REQUIRE(std::isfinite(omega[c]) && std::isfinite(m), PLBM_ERR_ARG,
                     "non-finite parameter for case " + std::to_string(c));
        BatchCase &k = b->cases[c];
        k.omega = omega[c];
        k.lambda_odd = plbm_lambda_d(omega[c], m);
        k.omega_f = float(k.omega);
        k.lambda_odd_f = float(k.lambda_odd);
    }
    PLBM_HIP(hipStreamSynchronize(b->stream));
    PLBM_HIP(hipMemcpy(b->d_cases, b->cases.data(), b->ncases * sizeof(BatchCase),
                       hipMemcpyHostToDevice));
    b->params_set = true;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_equilibrium_init(plbm_batch *b, int c, const double *rho, const double *ux,
                                const double *uy) try {
    PLBM_TRY(check_case(b, c));
    PLBM_REQUIRE(rho && ux && uy, PLBM_ERR_ARG, "null field pointer");
    const BatchCase &k = b->cases[c];
    const size_t n = (size_t)k.nx * k.ny;
    PLBM_TRY(ensure_btmp(b, 3 * n * sizeof(double)));
    PLBM_HIP(hipMemcpyAsync(b->d_tmp, rho, n * 8, hipMemcpyHostToDevice, b->stream));
    PLBM_HIP(hipMemcpyAsync(b->d_tmp + n, ux, n * 8, hipMemcpyHostToDevice, b->stream));
    PLBM_HIP(hipMemcpyAsync(b->d_tmp + 2 * n, uy, n * 8, hipMemcpyHostToDevice, b->stream));
    const size_t eb = storage_bytes(b->storage);
    void *dst = static_cast<char *>(b->arena) + (k.off + (b->parity ? k.elems : 0)) * eb;
    PLBM_TRY(dispatch_storage(b->storage, [&](auto S) {
        hipLaunchKernelGGL(k_batch_equilibrium<decltype(S)::value>,
                           dim3((k.ny + 255) / 256, k.nx), dim3(256), 0, b->stream, b->d_tmp,
                           b->d_tmp + n, b->d_tmp + 2 * n, dst, k.nx, k.ny, k.ldd, float(b->scale));
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    }));
    PLBM_HIP(hipStreamSynchronize(b->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_download_f(plbm_batch *b, int c, double *f, int ld) try {
    PLBM_TRY(check_case(b, c));
    const BatchCase &k = b->cases[c];
    PLBM_REQUIRE(f && ld >= k.ny, PLBM_ERR_ARG, "null array or ld < ny");
    const size_t n = (size_t)k.nx * k.ny;
    PLBM_TRY(ensure_btmp(b, n * sizeof(double)));
    const size_t eb = storage_bytes(b->storage);
    const void *src = static_cast<const char *>(b->arena) + (k.off + (b->parity ? k.elems : 0)) * eb;
    std::vector<double> stage;
    for (int q = 0; q < 9; ++q) {
        PLBM_TRY(dispatch_storage(b->storage, [&](auto S) {
            hipLaunchKernelGGL(k_batch_to_f64<decltype(S)::value>, dim3((k.ny + 255) / 256, k.nx),
                               dim3(256), 0, b->stream, src, b->d_tmp, q, k.nx, k.ny, k.ldd, wq(q),
                               float(1.0 / b->scale));
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        }));
        PLBM_TRY(plane_to_host(f + (size_t)q * ld * k.nx, ld, b->d_tmp, k.ny, k.nx, b->stream,
                               stage));
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_upload_f(plbm_batch *b, int c, const double *f, int ld) try {
    PLBM_TRY(check_case(b, c));
    const BatchCase &k = b->cases[c];
    PLBM_REQUIRE(f && ld >= k.ny, PLBM_ERR_ARG, "null array or ld < ny");
    const size_t n = (size_t)k.nx * k.ny;
    PLBM_TRY(ensure_btmp(b, n * sizeof(double)));
    const size_t eb = storage_bytes(b->storage);
    void *dst = static_cast<char *>(b->arena) + (k.off + (b->parity ? k.elems : 0)) * eb;
    std::vector<double> stage;
    for (int q = 0; q < 9; ++q) {
        PLBM_TRY(plane_from_host(b->d_tmp, f + (size_t)q * ld * k.nx, ld, k.ny, k.nx, b->stream,
                                 stage));
        PLBM_TRY(dispatch_storage(b->storage, [&](auto S) {
            hipLaunchKernelGGL(k_batch_from_f64<decltype(S)::value>, dim3((k.ny + 255) / 256, k.nx),
                               dim3(256), 0, b->stream, b->d_tmp, dst, q, k.nx, k.ny, k.ldd, wq(q),
                               float(b->scale));
            PLBM_HIP(hipGetLastError());
            return PLBM_OK;
        }));
        PLBM_HIP(hipStreamSynchronize(b->stream));
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_macros(plbm_batch *b, int c, double *rho, double *ux, double *uy) try {
    PLBM_TRY(check_case(b, c));
    PLBM_REQUIRE(rho && ux && uy, PLBM_ERR_ARG, "null field pointer");
    const BatchCase &k = b->cases[c];
    std::vector<double> f((size_t)k.ny * k.nx * 9);
    PLBM_TRY(plbm_batch_download_f(b, c, f.data(), k.ny));
    const size_t P = (size_t)k.ny * k.nx;
    for (size_t i = 0; i < P; ++i) {
        double v[9];
        for (int q = 0; q < 9; ++q) v[q] = f[q * P + i];
        const double r = v[0] + (((v[5] + v[7]) + (v[6] + v[8])) + ((v[1] + v[3]) + (v[2] + v[4])));
        rho[i] = r;
        ux[i] = (((v[5] - v[7]) + (v[8] - v[6])) + (v[1] - v[3])) / r;
        uy[i] = (((v[5] - v[7]) + (v[6] - v[8])) + (v[2] - v[4])) / r;
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_reduce(plbm_batch *b, double *out) try {
    PLBM_REQUIRE(b && out, PLBM_ERR_ARG, "null argument");
    PLBM_TRY(dispatch_storage(b->storage, [&](auto S) {
        hipLaunchKernelGGL(k_batch_reduce<decltype(S)::value>, dim3(b->ncases), dim3(kRedThreads), 0,
                           b->stream, b->d_cases, b->arena, b->parity, float(1.0 / b->scale),
                           b->d_red);
        PLBM_HIP(hipGetLastError());
        return PLBM_OK;
    }));
    PLBM_HIP(hipMemcpyAsync(out, b->d_red, (size_t)b->ncases * PLBM_NRED * 8,
                            hipMemcpyDeviceToHost, b->stream));
    PLBM_HIP(hipStreamSynchronize(b->stream));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_set_launch(plbm_batch *b, int ty, int tx, int resident_cells, int resident_block) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_REQUIRE(ty >= 1 && tx >= 1 && ty * tx <= 1024 && (ty * tx) % 64 == 0, PLBM_ERR_ARG,
                 "tile must have ty*tx a multiple of 64 and at most 1024 threads");
    PLBM_REQUIRE(resident_cells >= 0 && resident_block >= 64 && resident_block <= 1024 &&
                     resident_block % 64 == 0,
                 PLBM_ERR_ARG, "resident block must be a multiple of 64 in [64,1024]");
    PLBM_HIP(hipStreamSynchronize(b->stream));
    BatchCfg c = b->cfg;
    c.ty = ty;
    c.tx = tx;
    c.resident_cells = resident_cells;
    c.resident_block = resident_block;
    return set_cfg(b, c);
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_set_steps_per_pass(plbm_batch *b, int steps) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_REQUIRE(steps >= 1 && steps <= max_steps_per_pass(b->storage), PLBM_ERR_ARG,
                 "steps per pass must be 1.." + std::to_string(max_steps_per_pass(b->storage)) +
                     " for this storage");
    PLBM_HIP(hipStreamSynchronize(b->stream));
    BatchCfg c = b->cfg;
    c.ns = steps;
    return set_cfg(b, c);
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_get_steps_per_pass(const plbm_batch *b, int *steps) try {
    PLBM_REQUIRE(b && steps, PLBM_ERR_ARG, "null argument");
    *steps = b->cfg.ns;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_set_pass_columns(plbm_batch *b, int columns) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_REQUIRE(columns >= 1, PLBM_ERR_ARG, "columns per pass segment must be >= 1");
    PLBM_HIP(hipStreamSynchronize(b->stream));
    BatchCfg c = b->cfg;
    c.xs = columns;
    return set_cfg(b, c);
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_get_pass_columns(const plbm_batch *b, int *columns) try {
    PLBM_REQUIRE(b && columns, PLBM_ERR_ARG, "null argument");
    *columns = b->cfg.xs;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_get_launch(const plbm_batch *b, int *ty, int *tx, int *resident_cells,
                          int *resident_block) try {
    PLBM_REQUIRE(b && ty && tx && resident_cells && resident_block, PLBM_ERR_ARG, "null argument");
    *ty = b->cfg.ty;
    *tx = b->cfg.tx;
    *resident_cells = b->cfg.resident_cells;
    *resident_block = b->cfg.resident_block;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_run(plbm_batch *b, int nsteps) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_REQUIRE(nsteps >= 0, PLBM_ERR_ARG, "nsteps must be >= 0");
    PLBM_REQUIRE(b->params_set, PLBM_ERR_STATE, "call plbm_batch_set_params first");
    // Replay graphs of up to about 100 steps.  The chunk is even, so parity is preserved
    // and one graph serves every full chunk; with passes of ns steps it is a multiple of
    // 2 ns (96 at ns = 6), so that a full chunk is whole passes of ns steps.
    const int unit = b->cfg.ns > 1 ? 2 * b->cfg.ns : 2;
    const int chunk = std::max(unit, 100 / unit * unit);
    int done = 0;
    while (done < nsteps) {
        const int n = std::min(chunk, nsteps - done);
        hipGraphExec_t ex;
        PLBM_TRY(graph_for(b, n, b->parity, &ex));
        PLBM_HIP(hipGraphLaunch(ex, b->stream));
        b->parity = (b->parity + n) & 1;
        done += n;
    }
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_run_direct(plbm_batch *b, int nsteps) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_REQUIRE(nsteps >= 0, PLBM_ERR_ARG, "nsteps must be >= 0");
    PLBM_REQUIRE(b->params_set, PLBM_ERR_STATE, "call plbm_batch_set_params first");
    if (nsteps == 0) return PLBM_OK;
    PLBM_TRY(enqueue_steps(b, b->arena, b->d_tiles, b->ntiles, b->d_resident, b->nresident,
                           b->cfg, b->parity, nsteps));
    b->parity = (b->parity + nsteps) & 1;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

int plbm_batch_sync(plbm_batch *b) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_HIP(hipStreamSynchronize(b->stream));
    PLBM_HIP(hipStreamSynchronize(b->side));
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

// Autotuner: time `steps` steps of every candidate (tile shape x resident threshold) on a
// scratch copy of the arena (the batch's state is untouched), keep the fastest.
int plbm_batch_autotune(plbm_batch *b, int steps, double *best_ms) try {
    PLBM_REQUIRE(b, PLBM_ERR_ARG, "null batch handle");
    PLBM_REQUIRE(steps >= 2 && steps % 2 == 0, PLBM_ERR_ARG, "steps must be even and >= 2");
    PLBM_REQUIRE(b->params_set, PLBM_ERR_STATE, "call plbm_batch_set_params first");
    PLBM_HIP(hipStreamSynchronize(b->stream));
    void *scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto release = on_scope_exit([&] {
        (void)hipStreamSynchronize(b->side);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (scratch) (void)hipFree(scratch);
    });
    const size_t bytes = b->arena_elems * storage_bytes(b->storage);
    if (hipMalloc(&scratch, bytes) != hipSuccess) {
        (void)hipGetLastError();
        scratch = nullptr;
        set_error("batch autotune: no memory for a scratch arena");
        return PLBM_ERR_NOMEM;
    }
    PLBM_HIP(hipMemcpy(scratch, b->arena, bytes, hipMemcpyDeviceToDevice));
    PLBM_HIP(hipEventCreate(&e0));
    PLBM_HIP(hipEventCreate(&e1));
    std::vector<BatchCfg> cands;
    for (int res : {0, 4096, 16384, 36864})
        for (auto [ty, tx] : std::vector<std::pair<int, int>>{{64, 1}, {128, 1}, {256, 1}, {64, 4},
                                                               {32, 8}, {128, 2}, {512, 1}})
            for (int rb : {256, 1024}) {
                if (res == 0 && rb != 256) continue;
                BatchCfg c;
                c.ty = ty;
                c.tx = tx;
                c.resident_cells = res;
                c.resident_block = rb;
                cands.push_back(c);
            }
    const int max_ns = max_steps_per_pass(b->storage);
    for (int ns = 2; ns <= max_ns; ns += (ns < 4 ? 1 : 2))
        for (int xs : {16, 32, 64, 128, 256})
            for (int res : {0, 4096, 16384, 36864})
                for (int rb : {256, 1024}) {
                    if (res == 0 && rb != 256) continue;
                    BatchCfg c;
                    c.ns = ns;
                    c.xs = xs;
                    c.resident_cells = res;
                    c.resident_block = rb;
                    cands.push_back(c);
                }
    BatchCfg best = b->cfg;
    double tbest = 1e300;
    int status = PLBM_OK;
    for (const BatchCfg &c : cands) {
        BatchTile *t = nullptr;
        int *r = nullptr;
        int nt = 0, nr = 0;
        status = build_plan(b, c, &t, &nt, &r, &nr);
        if (status != PLBM_OK) break;
        // a pass plan is timed on whole passes of its depth, as plbm_batch_run's chunks
        // run it (at least `steps`, rounded up to a multiple of 2 ns; the even split of a
        // shorter count would time shallower passes, e.g. 10 steps at ns = 8 as two of 5)
        const int n = c.ns > 1 ? 2 * c.ns * ((steps + 2 * c.ns - 1) / (2 * c.ns)) : steps;
        double tmin = 1e300;
        for (int trial = 0; trial < 2 && status == PLBM_OK; ++trial) {
            (void)hipEventRecord(e0, b->stream);
            status = enqueue_steps(b, scratch, t, nt, r, nr, c, 0, n);
            (void)hipEventRecord(e1, b->stream);
            if (status == PLBM_OK && hipEventSynchronize(e1) != hipSuccess) {
                set_error("batch autotune: event synchronisation failed");
                status = PLBM_ERR_HIP;
            }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tmin = std::min(tmin, double(ms) / n);
        }
        if (t) (void)hipFree(t);
        if (r) (void)hipFree(r);
        if (status != PLBM_OK) break;
        if (tmin < tbest) {
            tbest = tmin;
            best = c;
        }
    }
    if (status != PLBM_OK) return status;
    PLBM_TRY(set_cfg(b, best));
    if (best_ms) *best_ms = tbest;
    return PLBM_OK;
} catch (...) {
    return plbm::host_exception();
}

}  // extern "C"
