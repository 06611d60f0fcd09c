// Temporal-blocked PPCG outer iteration for the 2D 5-point stencil: the
// device code, instantiated per pass size in tl_ppcg_tb_l*.hip.
//
// One PPCG outer iteration of the reference (tea_solve.f90:307-374) makes,
// after CGCalcW, these full sweeps: CGCalcUr (u += alpha p, r -= alpha w),
// PPCGInitSd (sd = r / theta), L inner steps of PPCGCalcU + PPCGCalcSd
// (r -= A sd; u += sd; sd = a_s sd + b_s r; ext_ppcg_kernel.hpp:37-100) and
// Calculate2Norm (rrn = r.r).  None of them needs a global result, so one
// kernel does all of them in a single pass over HBM: a wavefront streams
// down a strip of rows and keeps, per inner level, a three-row register
// window of sd and a two-row window of r; level s at row y - s is updated as
// soon as level s - 1 has produced row y - s + 1.  x-neighbours come from
// the adjacent lanes through DPP row shifts; each level therefore consumes
// one halo column per side, so a 64-lane wave of two columns per lane
// writes the 128 - 2L' columns in its middle (L' = L rounded up to even).
//
// Every cell value is produced by the reference's expressions in the
// reference's order (calc_ur, then init_sd, then step after step), so u and
// r are bitwise those of the unfused sweeps; only rrn's summation order is
// the kernel's own.  HBM traffic per outer iteration drops from
// 48+16+L*(40+24 + 16)+8 bytes per cell (the reference's sweeps) to the
// kernel's 48 + 16 (read u, p, r, w, kx, ky; write u, r).
#pragma once

#include "tl_chunk.hpp"

#include <algorithm>
#include <string>

namespace tl {

#if TL_DIM == 2
namespace tb {

// s_waitcnt immediate for vmcnt(0) with expcnt and lgkmcnt left unwaited
constexpr int VMCNT0 = 0x0F70;

struct TBArgs {
    double* u;
    const double* p;
    const double* r_in;
    const double* w;
    const double* sd_in;
    double* r_out;
    double* sd_out;
    const double* kx;
    const double* ky;
    const DevScalars* ds;
    double theta;
    const double* coef;  // a[0..L) then b[0..L) of this pass (device)
    // byte offsets of the rows a strip reads (cell row, then face row, both
    // reflected into the mesh) for rows row0 .. row0 + count - 1
    const uint2* rowtab;
    int row0;
    Reduce red;  // the last pass closes the rrn reduction
    int tiles_x;
    int rows;  // output rows per wave
    long nblocks;
};

// A wave-uniform double as scalar registers.
__device__ __forceinline__ double uniform(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Lane i receives lane i-1 (x-neighbour at kk-1); lane 0 receives 0.
// Lanes 0 and 63 own no output cell, so the fill never reaches a result.
__device__ __forceinline__ double from_west(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, 0x138, 0xf, 0xf, true),
                            __builtin_amdgcn_mov_dpp(lo, 0x138, 0xf, 0xf, true));
}

// Lane i receives lane i+1 (x-neighbour at kk+1); lane 63 receives 0.
__device__ __forceinline__ double from_east(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, 0x130, 0xf, 0xf, true),
                            __builtin_amdgcn_mov_dpp(lo, 0x130, 0xf, 0xf, true));
}

// Per-wave state of the streaming pass.  Each lane owns two adjacent
// columns (c0 = X0 - L' + 2 lane, c1 = c0 + 1), so a wave spans 128 columns
// and writes the 128 - 2L' in its middle.  Register windows rotate by round
// number instead of moving data: within an unrolled block of rounds the
// slot of every window is a compile-time constant.  u being accumulated
// passes from level to level in a two-row register window per level, as r
// does; the windows that span L+1 rows (kx, ky and the operator's
// diagonal, formed once per row by level 1) live in LDS rings indexed by
// row.
// One prefetched row: r, w (first pass) or sd, u, p and ky as the aligned
// pair that holds the lane's cells, kx at the lane's two faces.
struct TBRow {
    double2 r, b, u, p, ky;
    double kx0, kx1;
};

// Rows in flight ahead of the row being consumed (divides the unroll).
constexpr int TB_PF = 2;

template <int L>
struct TBState {
    double sdw[L][3][2];  // sd of level s, slot (round - ylo) % 3, column
    double rw[L][2][2];   // r of level s, slot (round - ylo) % 2, column
    double2 uw[L][2];     // u after level s, slot (round - ylo) % 2
    TBRow pf[TB_PF];      // prefetched rows
    double acc;
    double ca[L], cb[L];  // step coefficients (wave-uniform)
};

struct TBCtx {
    Geom g;
    int nx, ny, ylo, Y0, yhi;
    // byte offsets within a row: the aligned column pair that holds the
    // lane's two cells reflected into the mesh (hi0, hi1: a cell is the
    // pair's second element), the lane's kx faces reflected into the mesh,
    // the store target (the lane's own pair for lanes that own a cell, else
    // a pad pair past the halo that nothing reads) and the pad pair alone
    // (rows that own no output)
    unsigned bcp, bf0, bf1, bsp, bpad;
    bool hi0, hi1;
    bool out0, out1;
    double alpha, theta;
    double* kxr;  // LDS rings [ring rows][64][2]
    double* dr;
    double* kyr;
};

__device__ __forceinline__ double ld(const double* row, unsigned boff)
{
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(row) + boff);
}

__device__ __forceinline__ void st(double* row, unsigned boff, double v)
{
    *reinterpret_cast<double*>(reinterpret_cast<char*>(row) + boff) = v;
}

__device__ __forceinline__ double2 ldp(const double* row, unsigned boff)
{
    return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(row) + boff);
}

__device__ __forceinline__ void stp(double* row, unsigned boff, double2 v)
{
    *reinterpret_cast<double2*>(reinterpret_cast<char*>(row) + boff) = v;
}

__device__ __forceinline__ double pick(double2 v, bool hi) { return hi ? v.y : v.x; }

// The reflective boundary as data: a cell index outside [HALO, n + 1]
// reads the cell it mirrors, a face index outside [HALO, n + 2] the face it
// mirrors (reflections at both faces, repeated for meshes narrower than the
// reach).  A wave computing cells past a face then holds there the mirror
// image of the cells inside, level by level: the mirrored cell's stencil
// has the same operands with its east and west (north and south) terms
// exchanged, and the two-term sums it forms are commutative in IEEE
// arithmetic, so the image is bitwise equal.  The cell on the face thus
// reads its own value as its outside neighbour, as after the reference's
// reflective halo update, with no per-lane or per-row selection.
__host__ __device__ inline int fold_cell(int i, int n)
{
    const int P = 2 * n;
    int t = (i - HALO) % P;
    if (t < 0) t += P;
    if (t >= n) t = P - 1 - t;
    return HALO + t;
}

__host__ __device__ inline int fold_face(int f, int n)
{
    const int P = 2 * n;
    int t = (f - HALO) % P;
    if (t < 0) t += P;
    if (t > n) t = P - t;
    return HALO + t;
}

// The 16-byte aligned column pair that holds the reflections of cells k0
// and k0 + 1 (k0 even): reflection maps an aligned pair onto an aligned
// pair, reversed past a face, or onto one cell where an odd n folds
// (n + 1 and n + 2 both read n + 1).  hi0, hi1: the cell is the pair's
// second element.
struct CellPair {
    int start;
    bool hi0, hi1;
};

__host__ __device__ inline CellPair cell_pair(int k0, int n)
{
    const int f0 = fold_cell(k0, n), f1 = fold_cell(k0 + 1, n);
    const int s = (f0 < f1 ? f0 : f1) & ~1;
    return CellPair{s, f0 != s, f1 != s};
}

// Column margin of a wave for L levels, rounded up to even so that every
// lane's first column is even (HALO and the wave's output width are even).
__host__ __device__ constexpr int tb_margin(int L) { return L + (L & 1); }

// Loads of row y into a prefetch slot.  Every load is issued on every
// round, so that the number of memory operations in flight is the same on
// every path and the compiler can wait for exactly the oldest slot instead
// of draining all.  Rows outside the mesh read the rows they mirror (cells
// for r, w|sd, u, p, kx; faces for ky), from a table of reflected row
// offsets (one scalar load per round).  Cell columns come as the lane's
// aligned pair (one 16-byte load per array), the two kx faces one by one
// (reflected faces do not map aligned pairs onto aligned pairs); the pair's
// elements are assigned to the lane's cells where the row is consumed, so
// that no load is waited for early.  An address is the array's base plus a
// per-lane 32-bit byte offset of row and column (every field is under 4 GiB).
// A row's entry of the reflected-row table, read through the constant
// address space: a scalar load.
typedef const __attribute__((address_space(4))) unsigned ConstU;

template <bool FIRST>
__device__ __forceinline__ void tb_load(const TBArgs& A, const TBCtx& c, unsigned oy, unsigned of, TBRow& v)
{
    v.r = ldp(A.r_in, oy + c.bcp);
    v.b = ldp(FIRST ? A.w : A.sd_in, oy + c.bcp);
    v.u = ldp(A.u, oy + c.bcp);
    if (FIRST) v.p = ldp(A.p, oy + c.bcp);
    v.kx0 = ld(A.kx, oy + c.bf0);
    v.kx1 = ld(A.kx, oy + c.bf1);
    v.ky = ldp(A.ky, of + c.bcp);
}

// Rounds per unrolled block: a multiple of the register windows' periods
// (3 rows of sd, 2 of r, TB_PF prefetch slots) and of the LDS ring length,
// so that every window slot and ring row is a compile-time constant.
constexpr int TB_UNROLL = 12;

// LDS ring length for L levels: rows y - L .. y are live in a round; the
// smallest divisor of the unroll that holds them.
template <int L>
struct TBRing {
    static constexpr int N = L + 1 <= 2 ? 2 : L + 1 <= 3 ? 3 : L + 1 <= 4 ? 4 : L + 1 <= 6 ? 6 : 12;
    static_assert(L + 1 <= 12, "pass size beyond the LDS ring");
    // ring row of the round PH rounds into a block, minus s
    static constexpr int row(int PH, int s) { return ((PH - s) % N + N) % N; }
};

// One round: level 0 at row y, then level s at row y - s for s = 1..L.  The
// LDS rings of level s + 1 are read before level s computes, so their
// latency overlaps level s's arithmetic (the rows differ, so no write of
// level s is missed).  `tab` is the row table entry of row y + TB_PF - PH.
// WARM: the first block of a wave (y = ylo + PH = Y0 - L + PH).  The strip's
// output rows Y0.. depend on level s only at rows >= Y0 - (L - s), i.e. from
// round PH = 2 s on, so the arithmetic of the levels s > PH / 2 is left out;
// every load and store stays, so that the memory operations in flight at the
// loop head are those of every other block.
template <int L, bool FIRST, bool LAST, int PH, bool WARM>
__device__ __forceinline__ void tb_round(const TBArgs& A, const TBCtx& c, TBState<L>& st, int y, ConstU* tab)
{
    using R = TBRing<L>;
    constexpr int S0 = PH % 3, S1 = (PH + 2) % 3, S2 = (PH + 1) % 3;  // rounds y, y-1, y-2
    constexpr int R0 = PH % 2, R1 = (PH + 1) % 2;                      // rounds y, y-1
    const int lane = threadIdx.x;
    TBRow& v = st.pf[PH % TB_PF];
    double2* kxr = reinterpret_cast<double2*>(c.kxr) + lane;
    double2* kyr = reinterpret_cast<double2*>(c.kyr) + lane;
    double2* dr = reinterpret_cast<double2*>(c.dr) + lane;
    // kx at the east lane's first face: the next element of the ring row
    const double* kxe = c.kxr + 2 * lane + 2;

    // inputs of level 1 (row y - 1)
    double2 nkx = kxr[R::row(PH, 1) * WAVE], nky = kyr[R::row(PH, 1) * WAVE];
    double nkxe = kxe[2 * WAVE * R::row(PH, 1)];
    double2 nd = make_double2(0.0, 0.0);

    // level 0 at row y: CGCalcUr (r -= alpha w, u += alpha p) and PPCGInitSd,
    // evaluated on the loaded pair (elementwise), then assigned to the cells
    double kyn0, kyn1;
    {
        double2 r0 = v.r, sd0 = v.b;
        if (FIRST) {
            r0.x = v.r.x - c.alpha * v.b.x;
            r0.y = v.r.y - c.alpha * v.b.y;
            sd0.x = r0.x / c.theta;
            sd0.y = r0.y / c.theta;
        }
        st.sdw[0][S0][0] = pick(sd0, c.hi0);
        st.sdw[0][S0][1] = pick(sd0, c.hi1);
        st.rw[0][R0][0] = pick(r0, c.hi0);
        st.rw[0][R0][1] = pick(r0, c.hi1);
        double2 uu = v.u;
        if (FIRST) {
            uu.x = v.u.x + c.alpha * v.p.x;
            uu.y = v.u.y + c.alpha * v.p.y;
        }
        // rows outside the strip accumulate nothing that is stored
        const bool yin = y >= c.Y0 && y <= c.yhi;
        st.uw[0][R0] = make_double2(yin ? pick(uu, c.hi0) : 0.0, yin ? pick(uu, c.hi1) : 0.0);
        kxr[R::row(PH, 0) * WAVE] = make_double2(v.kx0, v.kx1);
        kyn0 = pick(v.ky, c.hi0);
        kyn1 = pick(v.ky, c.hi1);
        kyr[R::row(PH, 0) * WAVE] = make_double2(kyn0, kyn1);
    }
    // prefetch row y + TB_PF into the slot just consumed
    tb_load<FIRST>(A, c, tab[2 * PH], tab[2 * PH + 1], v);

#pragma unroll
    for (int s = 1; s <= L; ++s) {
        const int z = y - s;
        const double2 kx = nkx, ky = nky;
        const double kxe1 = nkxe;
        const double2 dp = nd;
        double2 u = st.uw[s - 1][R1];
        const double a = st.ca[s - 1], b = st.cb[s - 1];
        const bool skip = WARM && 2 * s > PH;  // a constant once the loop is unrolled
        if (s < L && !(WARM && 2 * (s + 1) > PH)) {
            nkx = kxr[R::row(PH, s + 1) * WAVE];
            nky = kyr[R::row(PH, s + 1) * WAVE];
            nkxe = kxe[2 * WAVE * R::row(PH, s + 1)];
            nd = dr[R::row(PH, s + 1) * WAVE];
        }
        double rn0 = 0.0, rn1 = 0.0;
        if (!skip) {
            const double c0 = st.sdw[s - 1][S1][0], c1 = st.sdw[s - 1][S1][1];
            const double s0 = st.sdw[s - 1][S2][0], s1 = st.sdw[s - 1][S2][1];
            const double n0 = st.sdw[s - 1][S0][0], n1 = st.sdw[s - 1][S0][1];
            // x-neighbours: c0's west is the previous lane's c1, c1's east the next lane's c0
            const double w0 = from_west(c1), e1 = from_east(c0);
            const double e0 = c1, w1 = c0;
            // SMVP (ext_shared.hpp:37-41) at (c0, z) and (c1, z); the
            // diagonal is the same at every level, so level 1 forms it for
            // row z and the levels after it read it back from the ring
            double2 d = dp;
            if (s == 1) {
                d = make_double2(1.0 + (kx.y + kx.x) + (kyn0 + ky.x), 1.0 + (kxe1 + kx.y) + (kyn1 + ky.y));
                dr[R::row(PH, 1) * WAVE] = d;
            }
            const double m0 = d.x * c0 - (kx.y * e0 + kx.x * w0) - (kyn0 * n0 + ky.x * s0);
            const double m1 = d.y * c1 - (kxe1 * e1 + kx.y * w1) - (kyn1 * n1 + ky.y * s1);
            kyn0 = ky.x;
            kyn1 = ky.y;
            // PPCGCalcU: r -= A sd; u += sd
            rn0 = st.rw[s - 1][R1][0] - m0;
            rn1 = st.rw[s - 1][R1][1] - m1;
            u.x = u.x + c0;
            u.y = u.y + c1;
            if (s < L) {
                // PPCGCalcSd with the coefficients of inner step s
                st.sdw[s][S0][0] = a * c0 + b * rn0;
                st.sdw[s][S0][1] = a * c1 + b * rn1;
                st.rw[s][R0][0] = rn0;
                st.rw[s][R0][1] = rn1;
                st.uw[s][R0] = u;
            }
        }
        if (s == L) {
            // output row: every lane stores, lanes and rows that own no
            // cell into a pad pair, so the store count never depends on a branch
            const bool zin = z >= c.Y0 && z <= c.yhi;
            const bool o0 = zin && c.out0, o1 = zin && c.out1;
            const unsigned bz = zin ? c.bsp : c.bpad;
            const unsigned oz = 8u * (unsigned)(c.g.base + (long)z * c.g.sx);
            const double c0 = st.sdw[s - 1][S1][0], c1 = st.sdw[s - 1][S1][1];
            stp(A.r_out, oz + bz, make_double2(rn0, rn1));
            stp(A.u, oz + bz, u);
            if (!LAST) stp(A.sd_out, oz + bz, make_double2(a * c0 + b * rn0, a * c1 + b * rn1));
            if (LAST) {
                st.acc += o0 ? rn0 * rn0 : 0.0;
                st.acc += o1 ? rn1 * rn1 : 0.0;
            }
        }
    }
}

template <int L, bool FIRST, bool LAST, int PH, bool WARM>
__device__ __forceinline__ void tb_rounds(const TBArgs& A, const TBCtx& c, TBState<L>& st, int y, ConstU* tab)
{
    if constexpr (PH < TB_UNROLL) {
        tb_round<L, FIRST, LAST, PH, WARM>(A, c, st, y + PH, tab);
        // the scheduler may not move work between rounds (interleaving them
        // raises the register count, and the copies it then makes of
        // registers with loads in flight wait for those loads)
        __builtin_amdgcn_sched_barrier(0);
        tb_rounds<L, FIRST, LAST, PH + 1, WARM>(A, c, st, y, tab);
    }
}

template <int L, bool FIRST, bool LAST>
__device__ __forceinline__ void tb_wave(const TBArgs& A, const TBCtx& c)
{
    TBState<L> st;
#pragma unroll
    for (int s = 0; s < L; ++s) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            st.sdw[s][0][q] = st.sdw[s][1][q] = st.sdw[s][2][q] = 0.0;
            st.rw[s][0][q] = st.rw[s][1][q] = 0.0;
        }
        st.uw[s][0] = st.uw[s][1] = make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int k = 0; k < TB_PF; ++k) {
        const double2 z = make_double2(0.0, 0.0);
        st.pf[k].r = st.pf[k].b = st.pf[k].u = st.pf[k].p = st.pf[k].ky = z;
        st.pf[k].kx0 = st.pf[k].kx1 = 0.0;
    }
    st.acc = 0.0;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        st.ca[s] = uniform(A.coef[s]);
        st.cb[s] = uniform(A.coef[L + s]);
    }
    ConstU* tab = (ConstU*)(const unsigned*)A.rowtab + 2 * (c.ylo - A.row0);
#pragma unroll
    for (int k = 0; k < TB_PF; ++k) tb_load<FIRST>(A, c, tab[2 * k], tab[2 * k + 1], st.pf[k]);
    // wait for the first rows here: with loads still in flight on entry
    // the compiler's wait at the loop head has to cover the entry path too,
    // and it then drains all loads at the head of every block of rounds
    __builtin_amdgcn_s_waitcnt(VMCNT0);
    // rows ylo + 12 k ... start each block, so that the LDS ring rows and
    // register window slots of a round depend only on its place in the block
    const int yend = c.yhi + L;
    tb_rounds<L, FIRST, LAST, 0, true>(A, c, st, c.ylo, tab + 2 * TB_PF);
    for (int y = c.ylo + TB_UNROLL; y <= yend; y += TB_UNROLL) {
        tb_rounds<L, FIRST, LAST, 0, false>(A, c, st, y, tab + 2 * (y - c.ylo + TB_PF));
    }
    if (LAST) {
        double v[1] = {st.acc};
        block_partial<1, 1>(v, A.red);
    }
}

template <int L, bool FIRST, bool LAST>
__global__ __launch_bounds__(WAVE) void k_ppcg_tb(Geom g, TBArgs A)
{
    if (gated_off(A.red)) return;
    using R = TBRing<L>;
    // one more element: the last lane's east face reads past the last row
    // (that lane owns no cell, so the value is never used)
    __shared__ double2 kx_ring[R::N * WAVE + 1];
    __shared__ double2 ky_ring[R::N * WAVE];
    __shared__ double2 d_ring[R::N * WAVE];
    const int lane = threadIdx.x;
    const long t = xcd_tile(blockIdx.x, A.nblocks);
    const int tx = (int)(t % A.tiles_x), ty = (int)(t / A.tiles_x);
    constexpr int LP = tb_margin(L);
    constexpr int OW = 2 * WAVE - 2 * LP;
    TBCtx c;
    c.g = g;
    c.nx = g.nx;
    c.ny = g.ny;
    const int X0 = HALO + tx * OW;
    c.Y0 = HALO + ty * A.rows;
    const int k0 = X0 - LP + 2 * lane, k1 = k0 + 1;
    {
        const CellPair cp = cell_pair(k0, c.nx);
        c.bcp = 8u * (unsigned)cp.start;
        c.hi0 = cp.hi0;
        c.hi1 = cp.hi1;
    }
    c.bf0 = 8u * (unsigned)fold_face(k0, c.nx);
    c.bf1 = 8u * (unsigned)fold_face(k1, c.nx);
    // k0 and the tile's bounds are even, so a lane owns both of its cells
    // or neither, except where an odd nx ends the mesh at k0
    c.out0 = k0 >= X0 && k0 < X0 + OW && k0 <= c.nx + 1;
    c.out1 = k1 >= X0 && k1 < X0 + OW && k1 <= c.nx + 1;
    {
        // pad columns X .. sx-1 of every row are never read (at least 30 of
        // them).  A lane whose first cell is the last of an odd nx also
        // writes the halo column nx + 2 past it, with the mirror image of
        // that cell: the value a reflective halo update puts there, and
        // no kernel reads it (neighbours across a face are read mirrored).
        const int xe = (int)g.X + ((int)g.X & 1);
        const int npair = (int)(g.sx - xe) / 2;
        c.bpad = 8u * (unsigned)(xe + 2 * (lane % npair));
        c.bsp = c.out0 ? 8u * (unsigned)k0 : c.bpad;
    }
    c.ylo = c.Y0 - L;
    c.yhi = min(c.Y0 + A.rows - 1, c.ny + 1);
    c.alpha = FIRST ? A.ds->alpha : 0.0;
    c.theta = A.theta;
    c.kxr = reinterpret_cast<double*>(kx_ring);
    c.kyr = reinterpret_cast<double*>(ky_ring);
    c.dr = reinterpret_cast<double*>(d_ring);
    tb_wave<L, FIRST, LAST>(A, c);
}

template <int L, bool FIRST, bool LAST>
void launch(Chunk& c, TBArgs& a, long nblocks)
{
    hipLaunchKernelGGL((k_ppcg_tb<L, FIRST, LAST>), dim3((unsigned)nblocks), dim3(WAVE), 0, c.stream, c.g, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
}

// Launchers of the compiled pass sizes (one translation unit each, so they
// build in parallel): first = the pass starts from CGCalcUr, last = it ends
// with ||r||^2 instead of handing sd over.
void tb_launch_1(Chunk& c, TBArgs& a, bool first, bool last);
void tb_launch_2(Chunk& c, TBArgs& a, bool first, bool last);
void tb_launch_5(Chunk& c, TBArgs& a, bool first, bool last);
void tb_launch_10(Chunk& c, TBArgs& a, bool first, bool last);

#define TL_TB_DEFINE_LAUNCH(N)                                               \
    void tb_launch_##N(Chunk& c, TBArgs& a, bool first, bool last)           \
    {                                                                        \
        if (first && last) launch<N, true, true>(c, a, a.nblocks);           \
        else if (first) launch<N, true, false>(c, a, a.nblocks);             \
        else if (last) launch<N, false, true>(c, a, a.nblocks);              \
        else launch<N, false, false>(c, a, a.nblocks);                       \
    }

}  // namespace tb
#endif

}  // namespace tl
