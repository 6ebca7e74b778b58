// multistep.hpp — one wavefront's share of a temporally blocked pass (NS fused stream +
// collide steps per sweep over HBM), shared by the single-lattice pass kernel
// (multistep.hip) and the batched engine (batch.hip).  See multistep.hip for the scheme.
#pragma once

#include "common.hpp"
#include "d2q9.hpp"
#include "stepper.hpp"

namespace plbm {

constexpr int kMultiXS = 64;  // columns per strip

// Rows with output per wave (VR): a pass of NS steps needs NS halo rows on each side, so
// VR <= 64 - 2 NS.  56 for 8-byte storage at up to 4 steps (448-byte runs, 64-byte
// aligned), else 48 (384 / 192 / 96-byte runs), which holds up to 8 steps.
constexpr int kMaxPassSteps = 8;
__host__ __device__ constexpr int multi_rows(int storage_bytes, int ns) {
    return storage_bytes == 8 && ns <= 4 ? 56 : 48;
}
template <int ST, int NS>
constexpr int multi_vr() {
    return multi_rows((int)sizeof(typename Traits<ST>::T), NS);
}
template <int ST>
constexpr int max_steps_per_pass() {
    return kMaxPassSteps;
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

using lds_void = __attribute__((address_space(3))) void;
using gbl_void = __attribute__((address_space(1))) void;

// 8-byte storage stages the next columns in LDS (direct global -> LDS loads into
// kStageAhead + 1 buffers per wavefront) instead of registers, so that the loads in
// flight cost no VGPRs.  A buffer holds the 64 rows of the 9 populations of one column,
// plus one slot of padding, 512 bytes each.  Measured on the 8192^2 FP64 TRT pass of 4
// steps: register prefetch of one column 2.01 ms (172 VGPRs, 2 waves per SIMD); staged,
// one column ahead 2.00 ms (3 waves per SIMD); two ahead 1.95 ms (LDS limits it to 2
// waves per SIMD, with twice the bytes in flight); three ahead 1.96 ms.  Four wavefronts
// per workgroup (the pass kernels launch 256 threads).
constexpr int kStageWaves = 4;
constexpr int kStageAhead = 2;  // columns in flight ahead of the one being consumed
constexpr int kStageBufs = kStageAhead + 1;
constexpr int kStageSlots = 10;

// Wait until at most n vector memory operations of this wavefront are outstanding, n
// rounded down to one of the counts the staged pass produces (rounding down is a stronger
// wait): vmcnt takes an immediate.
__device__ __forceinline__ void wait_vm_at_most(int n) {
    if (n >= 54) asm volatile("s_waitcnt vmcnt(54)" ::: "memory");
    else if (n >= 45) asm volatile("s_waitcnt vmcnt(45)" ::: "memory");
    else if (n >= 41) asm volatile("s_waitcnt vmcnt(41)" ::: "memory");
    else if (n >= 36) asm volatile("s_waitcnt vmcnt(36)" ::: "memory");
    else if (n >= 32) asm volatile("s_waitcnt vmcnt(32)" ::: "memory");
    else if (n >= 27) asm volatile("s_waitcnt vmcnt(27)" ::: "memory");
    else if (n >= 23) asm volatile("s_waitcnt vmcnt(23)" ::: "memory");
    else if (n >= 18) asm volatile("s_waitcnt vmcnt(18)" ::: "memory");
    else if (n >= 14) asm volatile("s_waitcnt vmcnt(14)" ::: "memory");
    else if (n >= 9) asm volatile("s_waitcnt vmcnt(9)" ::: "memory");
    else if (n >= 5) asm volatile("s_waitcnt vmcnt(5)" ::: "memory");
    else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// The next columns of a wavefront's walk along x, staged in LDS (8-byte values).  The
// wavefront's lane j holds row first_row + j (mod ny) of every population; column c is
// staged in buffer c % kStageBufs, population q in slot q (64 values).  When the first
// row and ny are even (and the planes 16-byte aligned), no pair of rows (2m, 2m+1)
// straddles the periodic wrap and one 16-byte load per lane moves two rows: lanes 0-31
// fill the slot of population q, lanes 32-63 that of q + 1 (5 loads per column; the fifth
// moves population 8 twice, the second copy into the padding slot).  Otherwise lane j
// moves the 4-byte half (j & 1) of row (j >> 1) and then of row 32 + (j >> 1) of each
// population (18 loads per column).
struct ColumnStage {
    double *buf0;      // this wavefront's kStageBufs buffers of kStageSlots x 64 values
    unsigned off_a, off_b;
    bool pairs;
    int loads;         // vector memory operations per staged column

    __device__ __forceinline__ void init(double *mem, const double *s, size_t L, int ny,
                                         int first_row, int lane) {
        buf0 = mem;
        auto row_of = [&](int m) {
            int yy = (first_row + m) % ny;
            return yy < 0 ? yy + ny : yy;
        };
        pairs = (first_row % 2 == 0) && (ny % 2 == 0) && (L % 2 == 0) &&
                (reinterpret_cast<uintptr_t>(s) % 16 == 0);
        if (pairs) {
            const unsigned off = (unsigned)row_of(2 * (lane & 31)) * 8u;
            off_a = off + (lane >> 5) * (unsigned)(L * 8);  // populations q, q + 1
            off_b = off;                                     // population 8 twice
        } else {
            off_a = (unsigned)row_of(lane >> 1) * 8u + (lane & 1) * 4u;
            off_b = (unsigned)row_of(32 + (lane >> 1)) * 8u + (lane & 1) * 4u;
        }
        loads = pairs ? 5 : 18;
    }

    // Issue the loads of the column whose planes start at s + cc (plane stride L) into
    // buffer buf.
    __device__ __forceinline__ void fetch(const double *s, size_t cc, size_t L, int buf) {
        double *slots = buf0 + buf * kStageSlots * 64;
        if (pairs) {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const char *plane = reinterpret_cast<const char *>(s + cc + 2 * m * L);
                __builtin_amdgcn_global_load_lds((gbl_void *)(plane + off_a),
                                                 (lds_void *)(slots + 2 * m * 64), 16, 0, 0);
            }
            const char *plane = reinterpret_cast<const char *>(s + cc + 8 * L);
            __builtin_amdgcn_global_load_lds((gbl_void *)(plane + off_b),
                                             (lds_void *)(slots + 8 * 64), 16, 0, 0);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                const char *plane = reinterpret_cast<const char *>(s + cc + q * L);
                __builtin_amdgcn_global_load_lds((gbl_void *)(plane + off_a),
                                                 (lds_void *)(slots + q * 64), 4, 0, 0);
                __builtin_amdgcn_global_load_lds((gbl_void *)(plane + off_b),
                                                 (lds_void *)(slots + q * 64 + 32), 4, 0, 0);
            }
        }
        // keep the stores of this iteration behind these loads in issue order (wait()
        // counts on it)
        asm volatile("" ::: "memory");
    }

    // Wait until column i of a walk of `iters` columns is in LDS.  The prologue issued
    // the loads of columns 0 .. kStageAhead-1; iteration t issues those of column
    // t + kStageAhead (if it exists) and then its stores: 9 of the last level (if
    // t >= first_store) and, when the pass also writes the level before the last (`lag`),
    // 9 of that level (if first_store - 1 <= t < iters - 1).  Vector memory operations
    // complete in issue order, so it suffices that no more remain outstanding than were
    // issued after the loads of column i.  Also waits for the LDS reads of the previous
    // column, whose buffer the next loads overwrite.
    __device__ __forceinline__ void wait(int i, int iters, int first_store, bool lag) const {
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        auto stores = [&](int t) {
            return (t >= first_store ? 9 : 0) +
                   (lag && t >= first_store - 1 && t < iters - 1 ? 9 : 0);
        };
        int after = 0, t0;
        if (i < kStageAhead) {
            for (int j = i + 1; j < kStageAhead; ++j) after += j < iters ? loads : 0;
            t0 = 0;
        } else {
            after += stores(i - kStageAhead);
            t0 = i - kStageAhead + 1;
        }
        for (int t = t0; t < i; ++t) {
            after += t + kStageAhead < iters ? loads : 0;
            after += stores(t);
        }
        wait_vm_at_most(after);
    }

    __device__ __forceinline__ const double *column(int i) const {
        return buf0 + (i % kStageBufs) * kStageSlots * 64;
    }

};

// Lane of a staged slot that lane `lane` reads for the pull in y of a population with
// velocity cy_: the neighbouring row, clamped at the wavefront's edges (halo lanes).
__device__ __forceinline__ int stage_pull_lane(int cy_, int lane) {
    return cy_ > 0 ? max(lane - 1, 0) : (cy_ < 0 ? min(lane + 1, 63) : lane);
}

// This wavefront's LDS staging buffers (four wavefronts per workgroup).
__device__ __forceinline__ double *stage_buffers() {
    __shared__ double stage_mem[kStageWaves][kStageBufs][kStageSlots][64];
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) & (kStageWaves - 1);
    return &stage_mem[wib][0][0][0];
}

// Value at row y - cy_q: cy = +1 comes from lane - 1, cy = -1 from lane + 1.  For 8-byte
// values DPP lane shifts (VALU moves): the FP64 pass is bound by instruction issue and
// latency, and ds_bpermute shuffles put an LDS round trip between consecutive levels
// (8192^2 TRT FP64, 8 steps per pass: 2.88 ms with ds_bpermute, 2.58 ms with DPP).  For
// 4-byte values the shuffles measured at least as fast (8192^2 TRT, ms per step with
// shuffles / DPP: FP32 storage 0.175 / 0.178, FP16 storage 0.227 / 0.255).
template <class R>
__device__ __forceinline__ R pull_y(int q, R v) {
    if constexpr (sizeof(R) == 8) {
        if (cy(q) > 0) return row_minus(v);
        if (cy(q) < 0) return row_plus(v);
        return v;
    } else {
        if (cy(q) > 0) return __shfl_up(v, 1, 64);
        if (cy(q) < 0) return __shfl_down(v, 1, 64);
        return v;
    }
}

// The calling wavefront advances rows [VR*wy, VR*wy + VR) (clipped to the grid) x columns
// [x0, x0 + ncols) (0 <= x0, x0 + ncols <= nx) of the grid described by `a` NS steps,
// a.src -> a.dst, and, when a.lag is set, stores level NS-1 of the same cells there.
// All 64 lanes must call it together (lane shifts and shuffles).  ROT: see the steady
// state below (the arithmetic and results are the same either way).  sync: a workgroup
// barrier at the start of every column, which keeps the workgroup's wavefronts (adjacent
// bands) on the same column; the caller sets it only where every wavefront of every
// workgroup walks the same number of columns, so that all of them reach every barrier.
template <int MODEL, int ST, int NS, bool ROT = false>
__device__ __forceinline__ void multi_wave(const StepArgs &a, int wy, int x0, int ncols,
                                           bool sync = false) {
    using T = typename Traits<ST>::T;
    using R = typename Traits<ST>::R;
    constexpr int VR = multi_vr<ST, NS>();
    static_assert(NS >= 1 && VR + 2 * NS <= 64, "halo rows do not fit in a wavefront");
    const int lane = threadIdx.x & 63;
    const int nx = a.nx, ny = a.ny;
    const int r = VR * wy - NS + lane;  // unwrapped row of this lane
    int y = r % ny;
    if (y < 0) y += ny;
    const bool out = lane >= NS && lane < VR + NS && r < ny;
    const size_t L = (size_t)a.ldd, cs = a.cs;
    const T *__restrict__ s = static_cast<const T *>(a.src);
    T *__restrict__ d = static_cast<T *>(a.dst);
    T *__restrict__ lag = static_cast<T *>(a.lag);  // wave-uniform
    // W[k][.][q]: level k's pull-ready values of its three newest columns.
    R W[NS][3][9];
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
    // The next columns' loads are issued ahead (two columns into LDS for 8-byte storage,
    // one column into registers otherwise), so that their latency overlaps the
    // collisions of the current column.
    constexpr bool kStage = sizeof(T) == 8;
    static_assert(!kStage || kStageBufs == 3, "staging buffers rotate with the window");
    T nxt[9];
    ColumnStage st{};
    int pull_lane[3];  // staged slot lanes of the level-0 pull for cy = -1, 0, +1
    if constexpr (kStage) {
        st.init(stage_buffers(), reinterpret_cast<const double *>(s), L, ny, VR * wy - NS, lane);
#pragma unroll
        for (int c = -1; c <= 1; ++c) pull_lane[c + 1] = stage_pull_lane(c, lane);
    }
    auto fetch = [&](int c, int buf) {
        const size_t cc = column(c);
        if constexpr (kStage) {
            st.fetch(reinterpret_cast<const double *>(s), cc, L, buf);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) nxt[q] = s[cc + q * L + y];
        }
    };
    const int iters = ncols + 2 * NS;
    constexpr int ahead = kStage ? kStageAhead : 1;
#pragma unroll
    for (int j = 0; j < ahead; ++j)
        if (j < iters) fetch(x0 - NS + j, j);

    // Iteration i loads level 0 at relative column i (column x0 - NS + i) and, once
    // i >= 2k, advances level k (k = 1 .. NS) at column i - k from level k-1's columns
    // i - k + 1 (the newest), i - k and i - k - 1.  Iterations past the walk's end (at
    // most two, see the steady state) load and store nothing and only re-read the last
    // column's staged values.
    auto load_level0 = [&](int i, int buf, int slot0, int fetch_buf)
                           __attribute__((always_inline)) {
        if constexpr (kStage) {
            if (!ROT || i < iters) st.wait(i, iters, 2 * NS, lag != nullptr);
            const double *mem = st.buf0 + buf * kStageSlots * 64;
#pragma unroll
            for (int q = 0; q < 9; ++q)
                // the pull in y reads the neighbouring row's slot (edge lanes are halo)
                W[0][slot0][q] = load_pop<ST>(
                    reinterpret_cast<const T *>(mem + q * 64 + pull_lane[cy(q) + 1]), a.inv_scale);
        } else {
#pragma unroll
            for (int q = 0; q < 9; ++q) W[0][slot0][q] = pull_y(q, load_pop<ST>(&nxt[q], a.inv_scale));
        }
        if (i + ahead < iters) fetch(x0 - NS + i + ahead, fetch_buf);
    };
    // Level k's collision of iteration i from level k-1's slots b0 (its newest column),
    // b1 and b2; stores level NS (and level NS-1 into `lag`).  Returns the result in f.
    auto advance = [&](int i, int k, int b0, int b1, int b2, R (&f)[9])
                       __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < 9; ++q)
            f[q] = cx(q) > 0 ? W[k - 1][b2][q] : (cx(q) == 0 ? W[k - 1][b1][q] : W[k - 1][b0][q]);
        collide_cell<MODEL, ST>(f, a);
        // level NS-1 reaches column x0 + j at iteration 2 NS - 1 + j, level NS at 2 NS + j
        if (k == NS - 1 && lag != nullptr && i >= 2 * NS - 1 && i < 2 * NS - 1 + ncols && out) {
            const size_t co = (size_t)(x0 - 2 * NS + 1 + i) * cs;
#pragma unroll
            for (int q = 0; q < 9; ++q) store_pop<ST>(lag + co + q * L + y, f[q], a.scale);
        }
        if (k == NS && out && (!ROT || i < iters)) {
            const size_t co = (size_t)(x0 - 2 * NS + i) * cs;
#pragma unroll
            for (int q = 0; q < 9; ++q) store_pop<ST>(d + co + q * L + y, f[q], a.scale);
        }
    };

    // Start-up (levels switch on one by one), or the whole walk without ROT: slot 0 holds
    // a level's newest column, 1 and 2 the older ones; a level that advances shifts them.
    const int shift_end = ROT ? 2 * NS : iters;
    for (int i = 0; i < shift_end; ++i) {
        if (sync) __builtin_amdgcn_s_barrier();
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            W[0][2][q] = W[0][1][q];
            W[0][1][q] = W[0][0][q];
        }
        load_level0(i, i % 3, 0, (i + ahead) % 3);
#pragma unroll
        for (int k = 1; k <= NS; ++k) {
            if (i < 2 * k) break;
            R f[9];
            advance(i, k, 0, 1, 2, f);
            if (k < NS) {
#pragma unroll
                for (int q = 0; q < 9; ++q) {
                    W[k][2][q] = W[k][1][q];
                    W[k][1][q] = W[k][0][q];
                    W[k][0][q] = pull_y(q, f[q]);
                }
            }
        }
    }
    // ROT: the steady state (every level advances) is unrolled by three columns so that
    // each slot index is a compile-time constant and the windows rotate by renaming
    // registers instead of moving them (60 FP64 moves per column fewer).  Phase
    // P = (i - 2 NS) % 3; the start-up leaves (newest, -1, -2) in slots (0, 1, 2), so in
    // phase P column -back sits in slot (back - P - 1) mod 3 of every level.  The walk is
    // rounded up to whole groups of three.  It costs registers (FP64 TRT: 206 -> 227 at 6
    // steps, 230 -> 255 at 7, beyond 256 at 8; FP32 / FP16 storage at 8 steps: 153 -> 163,
    // 151 -> 160), so the pass kernel asks for it only where it keeps the wavefronts per
    // SIMD and measured faster.
    if constexpr (ROT) {
        auto steady = [&](int i, auto phase) __attribute__((always_inline)) {
            constexpr int P = decltype(phase)::value;
            constexpr int s0 = (3 - P - 1) % 3, s1 = (4 - P - 1) % 3, s2 = (5 - P - 1) % 3;
            if (sync) __builtin_amdgcn_s_barrier();
            load_level0(i, (2 * NS + P) % 3, s0, (2 * NS + P + ahead) % 3);
#pragma unroll
            for (int k = 1; k <= NS; ++k) {
                R f[9];
                advance(i, k, s0, s1, s2, f);
                if (k < NS) {
#pragma unroll
                    for (int q = 0; q < 9; ++q) W[k][s0][q] = pull_y(q, f[q]);
                }
            }
        };
        for (int i = 2 * NS; i < iters; i += 3) {
            steady(i, std::integral_constant<int, 0>{});
            steady(i + 1, std::integral_constant<int, 1>{});
            steady(i + 2, std::integral_constant<int, 2>{});
        }
    }
}

}  // namespace plbm
