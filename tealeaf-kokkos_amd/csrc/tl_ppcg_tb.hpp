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
// one halo column per side, so a 64-lane wave writes 64 - 2L columns.
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
    double* partials;
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
// columns (c0 = X0 - L + 2 lane, c1 = c0 + 1), so a wave spans 128 columns
// and writes the 128 - 2L in its middle.  Register windows rotate by round
// number instead of moving data: within a block of six rounds the slot of
// every window is a compile-time constant.  The windows that span L+1 rows
// (u being accumulated, kx, ky) live in LDS rings indexed by row.
template <int L>
struct TBState {
    double sdw[L][3][2];  // sd of level s, slot (round - ylo) % 3, column
    double rw[L][2][2];   // r of level s, slot (round - ylo) % 2, column
    double pf[2][6][2];   // prefetched rows: r, w|sd, u, p, kx, ky
    double acc;
    double ca[L], cb[L];  // step coefficients (wave-uniform)
};

struct TBCtx {
    Geom g;
    int nx, ny, ylo, Y0, yhi;
    // byte offsets within a row: the lane's cells and kx faces reflected
    // into the mesh, store targets (the cell for lanes that own one, else a
    // pad column past the halo that nothing reads) and the pad column alone
    // (rows that own no output)
    unsigned bc0, bc1, bf0, bf1, bs0, bs1, bpad;
    bool out0, out1, out_any;
    double alpha, theta;
    double* ur;  // LDS rings [L+1][64][2]
    double* kxr;
    double* kyr;
    double* kxer;  // LDS ring [L+1][64]: kx at c1 + 1 (the east lane's c0)
};

__device__ __forceinline__ double ld(const double* row, unsigned boff)
{
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(row) + boff);
}

__device__ __forceinline__ void st(double* row, unsigned boff, double v)
{
    *reinterpret_cast<double*>(reinterpret_cast<char*>(row) + boff) = v;
}

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

// Loads of row y into a prefetch slot.  Every load is issued on every
// round, so that the number of memory operations in flight is the same on
// every path and the compiler can wait for exactly the oldest slot instead
// of draining all.  Rows outside the mesh read the rows they mirror (cells
// for r, w|sd, u, p, kx; faces for ky), from a table of reflected row
// offsets (one scalar load per round).  An address is the array's base plus
// a per-lane 32-bit byte offset of row and column (every field is under 4 GiB).
template <bool FIRST>
__device__ __forceinline__ void tb_load(const TBArgs& A, const TBCtx& c, int y, double (&v)[6][2])
{
    // a uniform read through the constant address space: a scalar load
    typedef const __attribute__((address_space(4))) unsigned ConstU;
    ConstU* t = (ConstU*)(const unsigned*)A.rowtab + 2 * (y - A.row0);
    const unsigned oy = t[0], of = t[1];
    v[0][0] = ld(A.r_in, oy + c.bc0);
    v[0][1] = ld(A.r_in, oy + c.bc1);
    const double* b = FIRST ? A.w : A.sd_in;
    v[1][0] = ld(b, oy + c.bc0);
    v[1][1] = ld(b, oy + c.bc1);
    v[2][0] = ld(A.u, oy + c.bc0);
    v[2][1] = ld(A.u, oy + c.bc1);
    if (FIRST) {
        v[3][0] = ld(A.p, oy + c.bc0);
        v[3][1] = ld(A.p, oy + c.bc1);
    }
    v[4][0] = ld(A.kx, oy + c.bf0);
    v[4][1] = ld(A.kx, oy + c.bf1);
    v[5][0] = ld(A.ky, of + c.bc0);
    v[5][1] = ld(A.ky, of + c.bc1);
}

__device__ __forceinline__ int ring(int row, int n)
{
    const int m = row % n;
    return m < 0 ? m + n : m;
}

// One round: level 0 at row y, then level s at row y - s for s = 1..L.  The LDS rings of
// level s + 1 are read before level s computes, so their latency overlaps
// level s's arithmetic (the rows differ, so no write of level s is missed).
template <int L, bool FIRST, bool LAST, int PH>
__device__ __forceinline__ void tb_round(const TBArgs& A, const TBCtx& c, TBState<L>& st, int y)
{
    constexpr int S0 = PH % 3, S1 = (PH + 2) % 3, S2 = (PH + 1) % 3;  // rounds y, y-1, y-2
    constexpr int R0 = PH % 2, R1 = (PH + 1) % 2;                      // rounds y, y-1
    const int lane = threadIdx.x;
    double(&v)[6][2] = st.pf[PH % 2];
    double2* ur = reinterpret_cast<double2*>(c.ur);
    double2* kxr = reinterpret_cast<double2*>(c.kxr);
    double2* kyr = reinterpret_cast<double2*>(c.kyr);

    // inputs of level 1 (row y - 1)
    int rn = ring(y - 1, L + 1);
    double2 nkx = kxr[rn * WAVE + lane], nky = kyr[rn * WAVE + lane], nu = ur[rn * WAVE + lane];
    double nkxe = c.kxer[rn * WAVE + lane];

    // level 0 at row y: CGCalcUr (r -= alpha w, u += alpha p) and PPCGInitSd
    const int ry = ring(y, L + 1);
    double kyn0, kyn1;
    {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            double r0, sd0;
            if (FIRST) {
                r0 = v[0][q] - c.alpha * v[1][q];
                sd0 = r0 / c.theta;
            } else {
                r0 = v[0][q];
                sd0 = v[1][q];
            }
            st.sdw[0][S0][q] = sd0;
            st.rw[0][R0][q] = r0;
        }
        double2 u0 = make_double2(0.0, 0.0);
        if (y >= c.Y0 && y <= c.yhi) {
            u0.x = FIRST ? v[2][0] + c.alpha * v[3][0] : v[2][0];
            u0.y = FIRST ? v[2][1] + c.alpha * v[3][1] : v[2][1];
        }
        ur[ry * WAVE + lane] = u0;
        kxr[ry * WAVE + lane] = make_double2(v[4][0], v[4][1]);
        c.kxer[ry * WAVE + lane] = from_east(v[4][0]);
        kyr[ry * WAVE + lane] = make_double2(v[5][0], v[5][1]);
        kyn0 = v[5][0];
        kyn1 = v[5][1];
    }
    // prefetch row y + 2 into the slot just consumed
    tb_load<FIRST>(A, c, y + 2, v);

#pragma unroll
    for (int s = 1; s <= L; ++s) {
        const int z = y - s;
        const int rz = rn;
        const double2 kx = nkx, ky = nky;
        const double kxe1 = nkxe;
        double2 u = nu;
        const double a = st.ca[s - 1], b = st.cb[s - 1];
        if (s < L) {
            rn = ring(z - 1, L + 1);
            nkx = kxr[rn * WAVE + lane];
            nky = kyr[rn * WAVE + lane];
            nu = ur[rn * WAVE + lane];
            nkxe = c.kxer[rn * WAVE + lane];
        }
        double rn0, rn1;
        {
            const double c0 = st.sdw[s - 1][S1][0], c1 = st.sdw[s - 1][S1][1];
            const double s0 = st.sdw[s - 1][S2][0], s1 = st.sdw[s - 1][S2][1];
            const double n0 = st.sdw[s - 1][S0][0], n1 = st.sdw[s - 1][S0][1];
            // x-neighbours: c0's west is the previous lane's c1, c1's east the next lane's c0
            const double w0 = from_west(c1), e1 = from_east(c0);
            const double e0 = c1, w1 = c0;
            // SMVP (ext_shared.hpp:37-41) at (c0, z) and (c1, z)
            const double m0 = (1.0 + (kx.y + kx.x) + (kyn0 + ky.x)) * c0 - (kx.y * e0 + kx.x * w0) - (kyn0 * n0 + ky.x * s0);
            const double m1 = (1.0 + (kxe1 + kx.y) + (kyn1 + ky.y)) * c1 - (kxe1 * e1 + kx.y * w1) - (kyn1 * n1 + ky.y * s1);
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
                ur[rz * WAVE + lane] = u;
            }
        }
        if (s == L) {
            // output row: every lane stores, lanes and rows that own no
            // cell into a pad column, so the store count never depends on a branch
            const bool zin = z >= c.Y0 && z <= c.yhi;
            const bool o0 = zin && c.out0, o1 = zin && c.out1;
            const unsigned b0 = zin ? c.bs0 : c.bpad, b1 = zin ? c.bs1 : c.bpad;
            const unsigned oz = 8u * (unsigned)(c.g.base + (long)z * c.g.sx);
            const double c0 = st.sdw[s - 1][S1][0], c1 = st.sdw[s - 1][S1][1];
            tb::st(A.r_out, oz + b0, rn0);
            tb::st(A.u, oz + b0, u.x);
            tb::st(A.r_out, oz + b1, rn1);
            tb::st(A.u, oz + b1, u.y);
            if (!LAST) {
                tb::st(A.sd_out, oz + b0, a * c0 + b * rn0);
                tb::st(A.sd_out, oz + b1, a * c1 + b * rn1);
            }
            if (LAST) {
                st.acc += o0 ? rn0 * rn0 : 0.0;
                st.acc += o1 ? rn1 * rn1 : 0.0;
            }
        }
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
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        st.pf[0][k][0] = st.pf[0][k][1] = st.pf[1][k][0] = st.pf[1][k][1] = 0.0;
    }
    st.acc = 0.0;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        st.ca[s] = uniform(A.coef[s]);
        st.cb[s] = uniform(A.coef[L + s]);
    }
    tb_load<FIRST>(A, c, c.ylo, st.pf[0]);
    tb_load<FIRST>(A, c, c.ylo + 1, st.pf[1]);
    // wait for the first two rows here: with loads still in flight on entry
    // the compiler's wait at the loop head has to cover the entry path too,
    // and it then drains all loads at the head of every block of rounds
    __builtin_amdgcn_s_waitcnt(VMCNT0);
    // rounds run in blocks of six; the scheduler may not move work between
    // rounds (interleaving them raises the register count, and the copies
    // it then makes of registers with loads in flight wait for those loads)
    const int yend = c.yhi + L;
    for (int y = c.ylo; y <= yend; y += 6) {
        tb_round<L, FIRST, LAST, 0>(A, c, st, y);
        __builtin_amdgcn_sched_barrier(0);
        tb_round<L, FIRST, LAST, 1>(A, c, st, y + 1);
        __builtin_amdgcn_sched_barrier(0);
        tb_round<L, FIRST, LAST, 2>(A, c, st, y + 2);
        __builtin_amdgcn_sched_barrier(0);
        tb_round<L, FIRST, LAST, 3>(A, c, st, y + 3);
        __builtin_amdgcn_sched_barrier(0);
        tb_round<L, FIRST, LAST, 4>(A, c, st, y + 4);
        __builtin_amdgcn_sched_barrier(0);
        tb_round<L, FIRST, LAST, 5>(A, c, st, y + 5);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (LAST) {
        double v[1] = {st.acc};
        block_partial<1, 1>(v, Reduce{A.partials});
    }
}

template <int L, bool FIRST, bool LAST>
__global__ __launch_bounds__(WAVE) void k_ppcg_tb(Geom g, TBArgs A)
{
    __shared__ double2 u_ring[(L + 1) * WAVE];
    __shared__ double2 kx_ring[(L + 1) * WAVE];
    __shared__ double2 ky_ring[(L + 1) * WAVE];
    __shared__ double kxe_ring[(L + 1) * WAVE];
    const int lane = threadIdx.x;
    const long t = xcd_tile(blockIdx.x, A.nblocks);
    const int tx = (int)(t % A.tiles_x), ty = (int)(t / A.tiles_x);
    constexpr int OW = 2 * WAVE - 2 * L;
    TBCtx c;
    c.g = g;
    c.nx = g.nx;
    c.ny = g.ny;
    const int X0 = HALO + tx * OW;
    c.Y0 = HALO + ty * A.rows;
    const int k0 = X0 - L + 2 * lane, k1 = k0 + 1;
    c.bc0 = 8u * (unsigned)fold_cell(k0, c.nx);
    c.bc1 = 8u * (unsigned)fold_cell(k1, c.nx);
    c.bf0 = 8u * (unsigned)fold_face(k0, c.nx);
    c.bf1 = 8u * (unsigned)fold_face(k1, c.nx);
    c.out0 = k0 >= X0 && k0 < X0 + OW && k0 <= c.nx + 1;
    c.out1 = k1 >= X0 && k1 < X0 + OW && k1 <= c.nx + 1;
    c.out_any = c.out0 || c.out1;
    {
        // pad columns X .. sx-1 of every row are never read
        const int npad = (int)(g.sx - g.X);
        c.bpad = 8u * (unsigned)(g.X + lane % npad);
        c.bs0 = c.out0 ? c.bc0 : c.bpad;
        c.bs1 = c.out1 ? c.bc1 : c.bpad;
    }
    c.ylo = c.Y0 - L;
    c.yhi = min(c.Y0 + A.rows - 1, c.ny + 1);
    c.alpha = FIRST ? A.ds->alpha : 0.0;
    c.theta = A.theta;
    c.ur = reinterpret_cast<double*>(u_ring);
    c.kxr = reinterpret_cast<double*>(kx_ring);
    c.kyr = reinterpret_cast<double*>(ky_ring);
    c.kxer = kxe_ring;
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
