// Device-side building blocks shared by every TeaLeaf kernel on gfx950:
// the padded HBM layout of a chunk, the launch geometry (64-lane column
// strips, XCD-aware tile order), clamped stencil loads, and deterministic
// grid reductions (closed inside the sweep or by a one-workgroup launch) that
// hand their results to the host through mapped pinned memory.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#ifndef TL_DIM
#error "compile with -DTL_DIM=2 or -DTL_DIM=3"
#endif

namespace tl {

constexpr int HALO = 2;
// Interior column kk = 2 of every row starts on a 256-byte boundary: a row
// begins XOFF doubles before the first halo column.
constexpr int XOFF = 30;
constexpr int ROW_ALIGN = 32;  // doubles (256 B)
// Zeroed rows before the first and after the last row of every field: the
// temporal-blocked kernels read and store rows up to 2L + 8 past the mesh
// (L <= 12 inner steps) without clamping the row index.
constexpr int SLACK_ROWS = 32;

constexpr int WAVE = 64;
constexpr int WAVES_PER_BLOCK = 4;
constexpr int BLOCK = WAVE * WAVES_PER_BLOCK;
constexpr int NUM_XCD = 8;

// Geometry of one chunk.  Reference coordinates (kk, jj, ii) run over the
// padded extents X = nx + 4, Y = ny + 4, Z = nz + 4 (Z = 1 in 2D), interior
// cells are [2, n + 2) on each axis (tealeaf_2d/ext_shared.hpp:47-49).
// Device element of (kk, jj, ii) is base + ii * sy + jj * sx + kk.
struct Geom {
    int nx, ny, nz;
    int X, Y, Z;
    long sx, sy;
    long base;
    long alloc;  // doubles per field allocation
    // Launch tiling: 64 columns x (4 waves) per block, each lane walks a
    // strip of `strip` cells along the slowest interior axis.
    int strip;
    int tiles_x, tiles_y, tiles_s;
    long nblocks;
};

__host__ __device__ inline long gidx(const Geom& g, int kk, int jj, int ii)
{
    return g.base + (long)ii * g.sy + (long)jj * g.sx + kk;
}

// Workgroups are dispatched round-robin over the 8 XCDs (w -> XCD w % 8).
// Remap so each XCD owns one contiguous run of tiles, which keeps vertical
// stencil neighbours inside that XCD's L2.
__device__ inline long xcd_tile(long w, long n)
{
    const long x = w % NUM_XCD, m = w / NUM_XCD;
    const long q = n / NUM_XCD, r = n % NUM_XCD;
    return x * q + (x < r ? x : r) + m;
}

// Strip owned by the current lane.  In 2D the strip runs along y (rows);
// in 3D along z (planes) with each wave owning one row of the y tile.
struct Strip {
    long i;      // device index of the first cell
    long step;   // stride between consecutive cells of the strip
    int kk, jj, ii;
    int count;   // cells in the strip (0 when the lane is outside the mesh)
};

__device__ inline Strip make_strip(const Geom& g)
{
    const long t = xcd_tile(blockIdx.x, g.nblocks);
    const int lane = threadIdx.x % WAVE;
    const int wave = threadIdx.x / WAVE;
    const int tx = (int)(t % g.tiles_x);
    const long rest = t / g.tiles_x;
    Strip s;
    s.kk = HALO + tx * WAVE + lane;
#if TL_DIM == 3
    const int ty = (int)(rest % g.tiles_y);
    const int ts = (int)(rest / g.tiles_y);
    s.jj = HALO + ty * WAVES_PER_BLOCK + wave;
    s.ii = HALO + ts * g.strip;
    int c = g.nz + HALO - s.ii;
    s.step = g.sy;
    const bool ok = s.kk < g.nx + HALO && s.jj < g.ny + HALO;
#else
    const int ts = (int)rest;
    s.jj = HALO + (ts * WAVES_PER_BLOCK + wave) * g.strip;
    s.ii = 0;
    int c = g.ny + HALO - s.jj;
    s.step = g.sx;
    const bool ok = s.kk < g.nx + HALO;
#endif
    if (c > g.strip) c = g.strip;
    if (c < 0 || !ok) c = 0;
    s.count = c;
    s.i = gidx(g, s.kk, s.jj, s.ii);
    return s;
}

__device__ inline void strip_advance(Strip& s)
{
    s.i += s.step;
#if TL_DIM == 3
    s.ii++;
#else
    s.jj++;
#endif
}

// ---------------------------------------------------------------------------
// Deterministic grid reductions.
//
// Inside each sweep every lane sums its own cells in strip order, the wave
// folds with a fixed xor butterfly, lane 0 of each wave stores to LDS, the
// wave partials are added in wave order and the block's total goes to
// partials[block].  The block partials are then added in block order by
// one workgroup in the pattern of sum_partials, and the scalar epilogue
// that consumes the total runs on the device.  That last phase runs either
// inside the sweep (single pass, reduce_close: the last workgroup to finish,
// found with device-wide counters, closes the reduction) or as a k_finish
// launch after it (the default, which measures faster on MI355X: the launch
// starts as the sweep ends and costs about 4 us, where the single pass
// lengthens the sweep by 8-10 us, each workgroup's count waiting for its own
// stores to drain).  Both give the same bits: the summation order depends
// only on the mesh and the launch geometry, so results are bitwise
// reproducible run to run.

// Solver scalars kept on the device so that consecutive kernels can use the
// recurrences without a host round trip (same IEEE divisions as the host).
struct DevScalars {
    double rro, pw, alpha, rrn, beta, error;
    double rro_prev;  // rro before the last rrn epilogue (a replaced residual recomputes beta from it)
    double bad;       // flexible CG: 1 when the last alpha had p.w <= 0 or a non-finite p.r
    double zr;        // flexible CG, sweeps only: z . r of the iteration, alpha's numerator
    double stop;      // 1 after an rrn at which the driver stops or leaves the iteration it speculated
};

// The scalar recurrence of tea_solve.f90 that consumes a reduction's total.
enum Epilogue {
    EP_STORE = 0, EP_RRO, EP_PW, EP_RRN, EP_RRN_REPLACE, EP_FCG_BETA, EP_FCG_PW, EP_FCG_ALPHA, EP_FCG_PWPR,
    EP_FCG_RRN, EP_FCG_BETA_RRN, EP_FCG_PW_ZR
};

// Counters of a single-pass reduction: one per group of REDUCE_GROUP
// consecutive workgroups and one over the groups, each on its own 256-byte
// line (atomics to one address are serialised, ~7 ns apiece: one counter for
// 4096 workgroups cost a 4096^2 sweep 30 us).
constexpr int REDUCE_GROUP = 64;
constexpr int COUNTER_STRIDE = 64;  // unsigned per counter line
__host__ __device__ constexpr long reduce_counters(long nblocks)
{
    return (1 + (nblocks + REDUCE_GROUP - 1) / REDUCE_GROUP) * COUNTER_STRIDE;
}
// Slots of a single-pass reduction hold SLOT_EMPTY between reductions (a
// NaN no workgroup stores: a partial with these bits is stored as the
// canonical NaN), so an empty slot is one whose store is still in flight.
// SLOT_WAIT bounds a reducer lane's polls for such slots (about a second in
// all): it is never reached while the grid makes progress, and if it were,
// the result would stay undelivered and the host would report it.
constexpr unsigned long long SLOT_EMPTY = ~0ull;
constexpr unsigned long long NAN_BITS = 0x7ff8000000000000ull;
constexpr int SLOT_WAIT = 1 << 20;

// Where the last phase of a reduction delivers its total.  `counter` set:
// the sweep closes the reduction itself; null: a k_finish launch follows.
struct Finish {
    unsigned* counter;  // reduce_counters(grid) counters, all 0 between reductions
    double* slots;      // [grid * NV] partials, all SLOT_EMPTY between reductions
    DevScalars* ds;
    double* dout;       // [4] device copy of the result
    double* hout;       // [4] host-mapped result
    double* hist_a;     // CG alpha / beta history (EP_RRN with hist_n >= 0)
    double* hist_b;
    unsigned* hseq;     // host-mapped: `seq` is written after the result
    int ep;
    int hist_n;
    unsigned seq;
    unsigned max_groups;  // groups the counters serve; a larger grid leaves the result undelivered
    // A speculative iteration (the driver launched it before reading the
    // previous rrn): its kernels do nothing and its finish delivers only the
    // sequence number when ds->stop is set.  An rrn epilogue sets ds->stop
    // when |rrn| < stop_eps, rrn <= stop_le, or (stop_gt finite) not
    // rrn <= stop_gt: a NaN rrn takes the branch the host's tests take.
    int gate;
    double stop_eps, stop_le, stop_gt;
};

struct Reduce {
    double* partials;  // [nblocks * NV] (a reduction closed by k_finish)
    Finish fin;        // fin.counter set: the sweep closes the reduction itself
};

// A kernel of a speculative iteration whose predecessor stopped the solve:
// it returns at once (every workgroup, before any barrier).
__device__ inline bool gated_off(const Reduce& r) { return r.fin.gate && r.fin.ds->stop != 0.0; }

template <int NV>
__device__ inline void wave_fold(double (&v)[NV])
{
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += __shfl_xor(v[k], off, WAVE);
    }
}

// Folds the block's lanes into thread 0 (v valid on thread 0 only).
template <int NV, int NWAVES = WAVES_PER_BLOCK>
__device__ inline void block_fold(double (&v)[NV])
{
    __shared__ double lds[NV][NWAVES];
    wave_fold<NV>(v);
    if (NWAVES == 1) return;
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[k][wave] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = lds[k][0];
#pragma unroll
            for (int w = 1; w < NWAVES; ++w) s += lds[k][w];
            v[k] = s;
        }
    }
}

// One slot of a single-pass reduction, read again while it is still empty
// (its store, issued before the count that covers it, can still be in
// flight; `polls` counts the lane's polls against SLOT_WAIT) and left empty
// for the next reduction.  An absent partial (b >= n) adds +0.0, which leaves
// a sum's bits unchanged (it starts at +0.0 and never becomes -0.0).
__device__ inline double take_slot(unsigned long long* bits, long i, unsigned long long x, bool present, int& polls)
{
    while (present && x == SLOT_EMPTY && polls < SLOT_WAIT) {
        ++polls;
        __builtin_amdgcn_s_sleep(1);
        x = __hip_atomic_load(bits + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (present) bits[i] = SLOT_EMPTY;
    return __longlong_as_double((long long)x);
}

// Total of n block partials, added as one BLOCK-thread workgroup adds them:
// thread t sums partials t, t + BLOCK, t + 2 BLOCK, ... in turn (a chain),
// each wave folds its lanes with the xor butterfly and the waves' sums are
// added in wave order.  A workgroup of NWAVES waves plays BLOCK / WAVE /
// NWAVES of those waves per wave, so every caller gets the same bits (on
// thread 0).  From the slots of a single-pass reduction (SLOTS) the loads go
// in batches that a lane keeps in flight together, across all of its chains
// (a lone wave has four: 32 loads in flight per lane).
template <int NV, int NWAVES, bool SLOTS = false>
__device__ inline void sum_partials(double* __restrict__ partials, long n, double (&out)[NV], bool* lost = nullptr)
{
    constexpr int VW = WAVES_PER_BLOCK;
    static_assert(VW % NWAVES == 0, "the reference workgroup's waves split evenly");
    constexpr int PER = VW / NWAVES;
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    double acc[PER][NV];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[j][k] = 0.0;
    }
    if (!SLOTS) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            for (long b = (long)(wave + j * NWAVES) * WAVE + lane; b < n; b += BLOCK) {
#pragma unroll
                for (int k = 0; k < NV; ++k) acc[j][k] += partials[b * NV + k];
            }
        }
    } else {
        constexpr int INFLIGHT = NWAVES == 1 ? 32 : 8;  // loads per lane per batch
        constexpr int MB = INFLIGHT / PER / NV > 0 ? INFLIGHT / PER / NV : 1;  // chain elements per batch
        auto* bits = reinterpret_cast<unsigned long long*>(partials);
        int polls = 0;
        const long t0 = (long)wave * WAVE + lane;  // chain j starts at t0 + j * NWAVES * WAVE
        for (long m0 = 0; t0 + m0 * BLOCK < n; m0 += MB) {
            unsigned long long x[PER][MB][NV];
#pragma unroll
            for (int j = 0; j < PER; ++j) {
#pragma unroll
                for (int q = 0; q < MB; ++q) {
                    const long b = t0 + (long)j * NWAVES * WAVE + (m0 + q) * BLOCK;
#pragma unroll
                    for (int k = 0; k < NV; ++k) x[j][q][k] = b < n ? bits[b * NV + k] : 0ull;
                }
            }
#pragma unroll
            for (int j = 0; j < PER; ++j) {
#pragma unroll
                for (int q = 0; q < MB; ++q) {
                    const long b = t0 + (long)j * NWAVES * WAVE + (m0 + q) * BLOCK;
#pragma unroll
                    for (int k = 0; k < NV; ++k) acc[j][k] += take_slot(bits, b * NV + k, x[j][q][k], b < n, polls);
                }
            }
        }
        *lost = polls >= SLOT_WAIT;
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) wave_fold<NV>(acc[j]);
    if (NWAVES == 1) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = acc[0][k];
#pragma unroll
            for (int j = 1; j < PER; ++j) s += acc[j][k];
            out[k] = s;
        }
        return;
    }
    __shared__ double lds[NV][VW];
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
#pragma unroll
            for (int k = 0; k < NV; ++k) lds[k][wave + j * NWAVES] = acc[j][k];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = lds[k][0];
#pragma unroll
            for (int w = 1; w < VW; ++w) s += lds[k][w];
            out[k] = s;
        }
    }
}

// Applies the epilogue to a reduction's total v and delivers the result:
// to the device scalars (cur holds their values before the reduction), to
// dout and to host-mapped memory, then the sequence number the host waits
// for.  One thread.  IN_SWEEP: called at the end of the sweep that computed
// the partials, whose stores fill the XCD's L2.
template <int NV, bool IN_SWEEP>
__device__ inline void finish_scalars(const double (&v)[NV], const Finish& f, const DevScalars& cur)
{
    DevScalars* ds = f.ds;
    if (f.gate && cur.stop != 0.0) {
        // a speculative iteration after the stop: nothing to deliver but the number
        __threadfence_system();
        *reinterpret_cast<volatile unsigned*>(f.hseq) = f.seq;
        return;
    }
    double out[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = v[k];
    switch (f.ep) {
    case EP_RRO:  // rro of ext_cg_solver_init
        ds->rro = v[0];
        break;
    case EP_PW: {  // alpha = rro / pw (tea_solve.f90:322,397)
        const double alpha = cur.rro / v[0];
        ds->pw = v[0];
        ds->alpha = alpha;
        out[1] = alpha;
        break;
    }
    case EP_RRN: {  // beta = rrn / rro; error = rrn; rro = rrn (:358,:373-374,:417,:433-434)
        const double beta = v[0] / cur.rro;
        ds->rrn = v[0];
        ds->beta = beta;
        if (f.hist_n >= 0) {
            f.hist_a[f.hist_n] = cur.alpha;
            f.hist_b[f.hist_n] = beta;
        }
        ds->rro_prev = cur.rro;
        ds->rro = v[0];
        ds->error = v[0];
        ds->stop = (std::fabs(v[0]) < f.stop_eps || v[0] <= f.stop_le || (f.stop_gt < HUGE_VAL && !(v[0] <= f.stop_gt)))
                       ? 1.0
                       : 0.0;
        out[1] = beta;
        break;
    }
    case EP_FCG_BETA:  // flexible CG: beta = -(z . w_prev) / (p . w)_prev
        ds->beta = -v[0] / cur.pw;
        break;
    case EP_FCG_PW:  // flexible CG: p . w
        ds->pw = v[0];
        break;
    case EP_FCG_ALPHA:  // flexible CG: alpha = (p . r) / (p . w); a pw <= 0 or non-finite p.r stops the solve
        ds->alpha = v[0] / cur.pw;
        ds->bad = (cur.pw > 0.0 && isfinite(v[0])) ? 0.0 : 1.0;
        break;
    case EP_FCG_PWPR:  // flexible CG: p . w and p . r in one reduction -> alpha
        ds->pw = v[0];
        ds->alpha = v[NV - 1] / v[0];
        ds->bad = (v[0] > 0.0 && isfinite(v[NV - 1])) ? 0.0 : 1.0;
        break;
    case EP_FCG_RRN:  // flexible CG: rrn, and the stop flag of this iteration's alpha
        out[1] = cur.bad;
        break;
    case EP_FCG_BETA_RRN:  // flexible CG: beta and z . r; the previous iteration's rrn and stop flag
        ds->beta = -v[0] / cur.pw;
        ds->zr = v[NV > 2 ? 2 : 0];
        out[0] = v[NV > 1 ? 1 : 0];
        out[1] = cur.bad;
        break;
    case EP_FCG_PW_ZR:  // flexible CG, sweeps only: alpha = (z . r) / (p . w)
        // z . r > 0 for the SPD sweep preconditioner and r != 0: anything
        // else is a breakdown and stops the solve
        ds->pw = v[0];
        ds->alpha = cur.zr / v[0];
        ds->bad = (v[0] > 0.0 && cur.zr > 0.0 && isfinite(cur.zr)) ? 0.0 : 1.0;
        break;
    case EP_RRN_REPLACE: {  // rrn of a replaced residual: beta from the rro before this iteration's rrn
        const double beta = v[0] / cur.rro_prev;
        ds->rrn = v[0];
        ds->beta = beta;
        ds->rro = v[0];
        ds->error = v[0];
        out[1] = beta;
        break;
    }
    default:
        break;
    }
    // the result reaches host memory before the number the host waits for
    if (IN_SWEEP) {
        // system-scope exchanges whose returns are waited for, then the
        // number (a system fence would also write back the sweep's stores
        // from the XCD's L2)
        unsigned long long dep = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f.dout[k] = out[k];
            dep |= __hip_atomic_exchange(reinterpret_cast<unsigned long long*>(f.hout + k),
                                         (unsigned long long)__double_as_longlong(out[k]), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_SYSTEM);
        }
        asm volatile("" ::"v"(dep) : "memory");
        __hip_atomic_store(f.hseq, f.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    } else {
        // a separate launch: its L2 holds next to nothing to write back
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            f.dout[k] = out[k];
            f.hout[k] = out[k];
        }
        __threadfence_system();
        *reinterpret_cast<volatile unsigned*>(f.hseq) = f.seq;
    }
}

// Single pass: thread 0 stores the workgroup's partial in its slot (an
// agent-scope atomic store, coherent across the XCDs' L2 caches on its own
// where a device-wide release fence would write back the XCD's whole L2 at
// the end of every workgroup), then counts the workgroup done in its group;
// the workgroup that completes a group counts the group done.  The workgroup
// that completes the last group takes an agent-scope acquire (it invalidates
// this XCD's stale lines; nothing is written back), adds the partials of the
// whole grid (a 1D grid, one partial per workgroup) in k_finish's order,
// polling a slot still empty, and applies the epilogue.  Every counter
// is back at 0 and every slot empty when the kernel ends.  No workgroup
// waits for another's progress: a count that never completes leaves the
// result undelivered, which the host reports.
template <int NV, int NWAVES>
__device__ inline void reduce_close(const double (&v)[NV], const Reduce& r)
{
    int last = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            const unsigned long long x = (unsigned long long)__double_as_longlong(v[k]);
            __hip_atomic_store(reinterpret_cast<unsigned long long*>(r.fin.slots + (long)blockIdx.x * NV + k),
                               x == SLOT_EMPTY ? NAN_BITS : x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const unsigned grp = blockIdx.x / REDUCE_GROUP, groups = (gridDim.x + REDUCE_GROUP - 1) / REDUCE_GROUP;
        const unsigned members = grp + 1 < groups ? REDUCE_GROUP : gridDim.x - grp * REDUCE_GROUP;
        unsigned* gc = r.fin.counter + (1 + grp) * COUNTER_STRIDE;
        if (groups <= r.fin.max_groups &&
            __hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == members - 1) {
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = groups == 1 ||
                   __hip_atomic_fetch_add(r.fin.counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                       groups - 1;
        }
    }
    if (NWAVES == 1) {
        last = __shfl(last, 0, WAVE);
    } else {
        __shared__ int flag;
        if (threadIdx.x == 0) flag = last;
        __syncthreads();
        last = flag;
    }
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    DevScalars cur{};
    if (threadIdx.x == 0 && r.fin.ds) cur = *r.fin.ds;  // in flight during the sum
    double s[NV];
    bool lost = false;
    sum_partials<NV, NWAVES, true>(r.fin.slots, gridDim.x, s, &lost);
    // a lane whose polls ran out leaves the result undelivered
    lost = NWAVES == 1 ? __any(lost) : __syncthreads_or(lost);
    if (threadIdx.x == 0) {
        if ((gridDim.x + REDUCE_GROUP - 1) / REDUCE_GROUP > 1)
            __hip_atomic_store(r.fin.counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!lost) finish_scalars<NV, true>(s, r.fin, cur);
    }
}

template <int NV, int NWAVES = WAVES_PER_BLOCK>
__device__ inline void block_partial(double (&v)[NV], const Reduce& r)
{
    block_fold<NV, NWAVES>(v);
    if (r.fin.counter) {
        reduce_close<NV, NWAVES>(v, r);
    } else if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) r.partials[(long)blockIdx.x * NV + k] = v[k];
    }
}

}  // namespace tl
