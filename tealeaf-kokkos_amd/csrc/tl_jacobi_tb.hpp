// Temporal-blocked Jacobi iterations for the 2D 5-point stencil: the device
// code, instantiated per pass size in tl_jacobi_tb_l*.hip.
//
// A Jacobi iteration of the reference (JacobiCopyU + JacobiSolve,
// ext_jacobi_kernel.hpp:65-117) is r = u, u = (u0 + sum_f k_f r_f) /
// (1 + sum_f k_f), error = sum |u - r|.  The driver tests the error after
// every iteration, so a pass of K iterations keeps one error sum per
// iteration: the host finds the first iteration that meets eps and, if it is
// not the last of the pass, runs the pass again from its input with that
// many iterations (the input u is kept).  Otherwise the wavefront is the
// one of the temporal-blocked PPCG kernel (tl_ppcg_tb.hpp): iteration s at
// row y - s from a three-row register window of u per level, u0, kx, ky in
// LDS rings, faces as mirrored loads, so u is bitwise that of K separate
// iterations.  HBM traffic per K iterations is 32 + 16 B/cell (read u, u0,
// kx, ky; write u and the previous u as r) instead of K x 40.
#pragma once

#include "tl_ppcg_tb.hpp"

namespace tl {

#if TL_DIM == 2
namespace jtb {

using tb::fold_cell;
using tb::fold_face;
using tb::from_east;
using tb::from_west;
using tb::ld;
using tb::ring;

struct JTBArgs {
    const double* u_in;
    double* u_out;  // != u_in: neighbouring tiles read u_in across their halo
    double* r;      // the u of the pass's second-to-last iteration (JacobiCopyU's r)
    const double* u0;
    const double* kx;
    const double* ky;
    const uint2* rowtab;  // reflected row byte offsets (cell row, face row)
    int row0;
    double* partials;  // K error sums per wave
    int tiles_x;
    int rows;
    long nblocks;
};

template <int K>
struct JTBState {
    double uw[K][3][2];  // u after iteration s, slot (round) % 3, column
    double pf[2][4][2];  // prefetched rows: u, u0, kx, ky
    double err[K];       // sum |u_s - u_(s-1)| over the wave's cells
};

struct JTBCtx {
    Geom g;
    int ylo, Y0, yhi;
    unsigned bc0, bc1, bf0, bf1, bs0, bs1, bpad;
    bool out0, out1;
    double2* u0r;  // LDS rings [K+1][64]
    double2* kxr;
    double2* kyr;
    double* kxer;  // kx at c1 + 1 (the east lane's c0)
};

__device__ __forceinline__ void stv(double* a, unsigned boff, double v)
{
    *reinterpret_cast<double*>(reinterpret_cast<char*>(a) + boff) = v;
}

__device__ __forceinline__ void jtb_load(const JTBArgs& A, const JTBCtx& c, int y, double (&v)[4][2])
{
    typedef const __attribute__((address_space(4))) unsigned ConstU;
    ConstU* t = (ConstU*)(const unsigned*)A.rowtab + 2 * (y - A.row0);
    const unsigned oy = t[0], of = t[1];
    v[0][0] = ld(A.u_in, oy + c.bc0);
    v[0][1] = ld(A.u_in, oy + c.bc1);
    v[1][0] = ld(A.u0, oy + c.bc0);
    v[1][1] = ld(A.u0, oy + c.bc1);
    v[2][0] = ld(A.kx, oy + c.bf0);
    v[2][1] = ld(A.kx, oy + c.bf1);
    v[3][0] = ld(A.ky, of + c.bc0);
    v[3][1] = ld(A.ky, of + c.bc1);
}

// One round: the loaded row y enters as iteration 0's u, iteration s updates
// row y - s (s = 1..K); row y - K leaves the pass.
template <int K, int PH>
__device__ __forceinline__ void jtb_round(const JTBArgs& A, const JTBCtx& c, JTBState<K>& st, int y)
{
    constexpr int S0 = PH % 3, S1 = (PH + 2) % 3, S2 = (PH + 1) % 3;  // rounds y, y-1, y-2
    const int lane = threadIdx.x;
    double(&v)[4][2] = st.pf[PH % 2];

    int rn = ring(y - 1, K + 1);
    double2 nkx = c.kxr[rn * WAVE + lane], nky = c.kyr[rn * WAVE + lane], nu0 = c.u0r[rn * WAVE + lane];
    double nkxe = c.kxer[rn * WAVE + lane];

    const int ry = ring(y, K + 1);
    st.uw[0][S0][0] = v[0][0];
    st.uw[0][S0][1] = v[0][1];
    c.u0r[ry * WAVE + lane] = make_double2(v[1][0], v[1][1]);
    c.kxr[ry * WAVE + lane] = make_double2(v[2][0], v[2][1]);
    c.kxer[ry * WAVE + lane] = from_east(v[2][0]);
    c.kyr[ry * WAVE + lane] = make_double2(v[3][0], v[3][1]);
    double kyn0 = v[3][0], kyn1 = v[3][1];
    jtb_load(A, c, y + 2, v);

#pragma unroll
    for (int s = 1; s <= K; ++s) {
        const int z = y - s;
        const double2 kx = nkx, ky = nky, u0 = nu0;
        const double kxe1 = nkxe;
        if (s < K) {
            rn = ring(z - 1, K + 1);
            nkx = c.kxr[rn * WAVE + lane];
            nky = c.kyr[rn * WAVE + lane];
            nu0 = c.u0r[rn * WAVE + lane];
            nkxe = c.kxer[rn * WAVE + lane];
        }
        const double c0 = st.uw[s - 1][S1][0], c1 = st.uw[s - 1][S1][1];
        const double s0 = st.uw[s - 1][S2][0], s1 = st.uw[s - 1][S2][1];
        const double n0 = st.uw[s - 1][S0][0], n1 = st.uw[s - 1][S0][1];
        const double w0 = from_west(c1), e1 = from_east(c0);
        const double e0 = c1, w1 = c0;
        // JacobiSolve (ext_jacobi_kernel.hpp:85-100) on r = the previous u
        const double un0 = (u0.x + (kx.y * e0 + kx.x * w0) + (kyn0 * n0 + ky.x * s0)) /
                           (1.0 + (kx.x + kx.y) + (ky.x + kyn0));
        const double un1 = (u0.y + (kxe1 * e1 + kx.y * w1) + (kyn1 * n1 + ky.y * s1)) /
                           (1.0 + (kx.y + kxe1) + (ky.y + kyn1));
        kyn0 = ky.x;
        kyn1 = ky.y;
        const bool zin = z >= c.Y0 && z <= c.yhi;
        st.err[s - 1] += zin && c.out0 ? fabs(un0 - c0) : 0.0;
        st.err[s - 1] += zin && c.out1 ? fabs(un1 - c1) : 0.0;
        if (s < K) {
            st.uw[s][S0][0] = un0;
            st.uw[s][S0][1] = un1;
        } else {
            // every lane stores, lanes and rows that own no cell into a pad column
            const unsigned b0 = zin ? c.bs0 : c.bpad, b1 = zin ? c.bs1 : c.bpad;
            const unsigned oz = 8u * (unsigned)(c.g.base + (long)z * c.g.sx);
            stv(A.u_out, oz + b0, un0);
            stv(A.r, oz + b0, c0);
            stv(A.u_out, oz + b1, un1);
            stv(A.r, oz + b1, c1);
        }
    }
}

template <int K>
__device__ __forceinline__ void jtb_wave(const JTBArgs& A, const JTBCtx& c)
{
    JTBState<K> st;
#pragma unroll
    for (int s = 0; s < K; ++s) {
#pragma unroll
        for (int q = 0; q < 2; ++q) st.uw[s][0][q] = st.uw[s][1][q] = st.uw[s][2][q] = 0.0;
        st.err[s] = 0.0;
    }
    jtb_load(A, c, c.ylo, st.pf[0]);
    jtb_load(A, c, c.ylo + 1, st.pf[1]);
    // the first rows complete before the loop, so the loop head waits only
    // for the oldest prefetch slot (see tl_ppcg_tb.hpp)
    __builtin_amdgcn_s_waitcnt(tb::VMCNT0);
    const int yend = c.yhi + K;
    for (int y = c.ylo; y <= yend; y += 6) {
        jtb_round<K, 0>(A, c, st, y);
        __builtin_amdgcn_sched_barrier(0);
        jtb_round<K, 1>(A, c, st, y + 1);
        __builtin_amdgcn_sched_barrier(0);
        jtb_round<K, 2>(A, c, st, y + 2);
        __builtin_amdgcn_sched_barrier(0);
        jtb_round<K, 3>(A, c, st, y + 3);
        __builtin_amdgcn_sched_barrier(0);
        jtb_round<K, 4>(A, c, st, y + 4);
        __builtin_amdgcn_sched_barrier(0);
        jtb_round<K, 5>(A, c, st, y + 5);
        __builtin_amdgcn_sched_barrier(0);
    }
    block_partial<K, 1>(st.err, Reduce{A.partials});
}

template <int K>
__global__ __launch_bounds__(WAVE) void k_jacobi_tb(Geom g, JTBArgs A)
{
    __shared__ double2 u0_ring[(K + 1) * WAVE];
    __shared__ double2 kx_ring[(K + 1) * WAVE];
    __shared__ double2 ky_ring[(K + 1) * WAVE];
    __shared__ double kxe_ring[(K + 1) * WAVE];
    const int lane = threadIdx.x;
    const long t = xcd_tile(blockIdx.x, A.nblocks);
    const int tx = (int)(t % A.tiles_x), ty = (int)(t / A.tiles_x);
    constexpr int OW = 2 * WAVE - 2 * K;
    JTBCtx c;
    c.g = g;
    const int X0 = HALO + tx * OW;
    c.Y0 = HALO + ty * A.rows;
    const int k0 = X0 - K + 2 * lane, k1 = k0 + 1;
    c.bc0 = 8u * (unsigned)fold_cell(k0, g.nx);
    c.bc1 = 8u * (unsigned)fold_cell(k1, g.nx);
    c.bf0 = 8u * (unsigned)fold_face(k0, g.nx);
    c.bf1 = 8u * (unsigned)fold_face(k1, g.nx);
    c.out0 = k0 >= X0 && k0 < X0 + OW && k0 <= g.nx + 1;
    c.out1 = k1 >= X0 && k1 < X0 + OW && k1 <= g.nx + 1;
    {
        // pad columns X .. sx-1 of every row are never read
        const int npad = (int)(g.sx - g.X);
        c.bpad = 8u * (unsigned)(g.X + lane % npad);
        c.bs0 = c.out0 ? c.bc0 : c.bpad;
        c.bs1 = c.out1 ? c.bc1 : c.bpad;
    }
    c.ylo = c.Y0 - K;
    c.yhi = min(c.Y0 + A.rows - 1, g.ny + 1);
    c.u0r = u0_ring;
    c.kxr = kx_ring;
    c.kyr = ky_ring;
    c.kxer = kxe_ring;
    jtb_wave<K>(A, c);
}

template <int K>
void launch(Chunk& c, JTBArgs& a)
{
    hipLaunchKernelGGL((k_jacobi_tb<K>), dim3((unsigned)a.nblocks), dim3(WAVE), 0, c.stream, c.g, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
}

// Launchers of the compiled pass sizes (one translation unit each).
void jtb_launch_1(Chunk& c, JTBArgs& a);
void jtb_launch_2(Chunk& c, JTBArgs& a);
void jtb_launch_5(Chunk& c, JTBArgs& a);
void jtb_launch_10(Chunk& c, JTBArgs& a);

#define TL_JTB_DEFINE_LAUNCH(N) \
    void jtb_launch_##N(Chunk& c, JTBArgs& a) { launch<N>(c, a); }

}  // namespace jtb
#endif

}  // namespace tl
