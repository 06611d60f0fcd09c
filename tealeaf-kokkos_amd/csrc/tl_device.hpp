// Device-side building blocks shared by every TeaLeaf kernel on gfx950:
// the padded HBM layout of a chunk, the launch geometry (64-lane column
// strips, XCD-aware tile order), clamped stencil loads, and a deterministic
// single-pass grid reduction that hands its result to the host through
// mapped pinned memory.
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
// Deterministic two-phase reductions.
//
// Phase 1 (inside each sweep): every lane sums its own cells in strip order,
// the wave folds with a fixed xor butterfly, lane 0 of each wave stores to
// LDS, the wave partials are added in wave order and the block's total goes
// to partials[block].  Phase 2 (k_finish, one workgroup) adds the block
// partials in block order with the same pattern and applies the scalar
// epilogue.  No fences or atomics are needed inside the sweep: the kernel
// boundary orders the two phases.  The summation order depends only on the
// mesh and the launch geometry, so results are bitwise reproducible.
struct Reduce {
    double* partials;  // [nblocks * NV]
};

// Solver scalars kept on the device so that consecutive kernels can use the
// recurrences without a host round trip (same IEEE divisions as the host).
struct DevScalars {
    double rro, pw, alpha, rrn, beta, error;
    double rro_prev;  // rro before the last rrn epilogue (a replaced residual recomputes beta from it)
    double bad;       // flexible CG: 1 when the last alpha had p.w <= 0 or a non-finite p.r
    double zr;        // flexible CG, sweeps only: z . r of the iteration, alpha's numerator
};

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

template <int NV, int NWAVES = WAVES_PER_BLOCK>
__device__ inline void block_partial(double (&v)[NV], const Reduce& r)
{
    block_fold<NV, NWAVES>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) r.partials[(long)blockIdx.x * NV + k] = v[k];
    }
}

}  // namespace tl
