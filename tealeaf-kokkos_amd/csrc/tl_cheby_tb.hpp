// Temporal-blocked Chebyshev iterations for the 2D 5-point stencil: the
// device code, instantiated per pass size in tl_cheby_tb_l*.hip.
//
// A Chebyshev iteration of the reference (ChebyIterate + ChebyCalcU,
// ext_cheby_kernel.hpp:75-115) is w = A u, r = u0 - w, p = a_s p + b_s r,
// u += p over the whole mesh; the driver checks ||r||^2 only every tenth
// iteration past its estimate (tea_solve.f90:260), so the K iterations up
// to the next check need no global result and one kernel runs them in a
// single pass over HBM, with the wavefront of the temporal-blocked PPCG
// kernel (tl_ppcg_tb.hpp): level s (iteration s of the pass) at row y - s is
// computed as soon as level s - 1 has row y - s + 1, from a three-row
// register window of u and a two-row window of p per level; u0, kx, ky wait
// in LDS rings for the levels that reach their row.  Faces are mirrored
// loads, as there.  Each cell's value is the reference's expressions in its
// order, so u, p and r are bitwise those of K separate iterations; HBM
// traffic per K iterations is 40 + 24 B/cell (read u, p, u0, kx, ky; write
// u, p, r) instead of K x 64.
#pragma once

#include "tl_ppcg_tb.hpp"

namespace tl {

#if TL_DIM == 2
namespace ctb {

using tb::ConstU;
using tb::fold_face;
using tb::from_east;
using tb::from_west;
using tb::ld;
using tb::ldp;
using tb::pick;
using tb::stp;
using tb::uniform;

struct CTBArgs {
    const double* u_in;
    double* u_out;  // != u_in: neighbouring tiles read u_in across their halo
    const double* p_in;
    double* p_out;  // != p_in, for the same reason
    double* r;
    const double* u0;
    const double* kx;
    const double* ky;
    const double* ca;  // a[0..K) and b[0..K) of this pass (device)
    const double* cb;
    const uint2* rowtab;  // reflected row byte offsets (cell row, face row)
    int row0;
    Reduce red;  // the last pass: ||r||^2 (closed when the driver checks it)
    int tiles_x;
    int rows;
    long nblocks;
};

// One prefetched row: u, p, u0 and ky as the aligned pair that holds the
// lane's cells, kx at the lane's two faces.
struct CTBRow {
    double2 u, p, u0, ky;
    double kx0, kx1;
};

template <int K>
struct CTBState {
    double uw[K][3][2];  // u after iteration s, slot (round) % 3, column
    double pw[K][2][2];  // p after iteration s, slot (round) % 2
    CTBRow pf[2];        // prefetched rows
    double acc;
    double ca[K], cb[K];
};

struct CTBCtx {
    Geom g;
    int ny, ylo, Y0, yhi;
    // as in the PPCG kernel: the lane's aligned cell pair (hi0, hi1: a cell
    // is its second element), kx faces, store pair and pad pair
    unsigned bcp, bf0, bf1, bsp, bpad;
    bool hi0, hi1;
    bool out0, out1;
    double* u0r;  // LDS rings [ring rows][64][2]
    double* kxr;
    double* kyr;
};

__device__ __forceinline__ void ctb_load(const CTBArgs& A, const CTBCtx& c, unsigned oy, unsigned of, CTBRow& v)
{
    v.u = ldp(A.u_in, oy + c.bcp);
    v.p = ldp(A.p_in, oy + c.bcp);
    v.u0 = ldp(A.u0, oy + c.bcp);
    v.kx0 = ld(A.kx, oy + c.bf0);
    v.kx1 = ld(A.kx, oy + c.bf1);
    v.ky = ldp(A.ky, of + c.bcp);
}

// One round: the loaded row y enters as iteration 0's u and p, iteration s
// updates row y - s (s = 1..K); row y - K leaves the pass.  `tab` is the
// row table entry of row y + 2 - PH.
template <int K, bool LAST, int PH>
__device__ __forceinline__ void ctb_round(const CTBArgs& A, const CTBCtx& c, CTBState<K>& st, int y, ConstU* tab)
{
    using R = tb::TBRing<K>;
    constexpr int S0 = PH % 3, S1 = (PH + 2) % 3, S2 = (PH + 1) % 3;  // rounds y, y-1, y-2
    constexpr int R0 = PH % 2, R1 = (PH + 1) % 2;                      // rounds y, y-1
    const int lane = threadIdx.x;
    CTBRow& v = st.pf[PH % 2];
    double2* u0r = reinterpret_cast<double2*>(c.u0r) + lane;
    double2* kxr = reinterpret_cast<double2*>(c.kxr) + lane;
    double2* kyr = reinterpret_cast<double2*>(c.kyr) + lane;
    // kx at the east lane's first face: the next element of the ring row
    const double* kxe = c.kxr + 2 * lane + 2;

    double2 nkx = kxr[R::row(PH, 1) * WAVE], nky = kyr[R::row(PH, 1) * WAVE], nu0 = u0r[R::row(PH, 1) * WAVE];
    double nkxe = kxe[2 * WAVE * R::row(PH, 1)];

    st.uw[0][S0][0] = pick(v.u, c.hi0);
    st.uw[0][S0][1] = pick(v.u, c.hi1);
    st.pw[0][R0][0] = pick(v.p, c.hi0);
    st.pw[0][R0][1] = pick(v.p, c.hi1);
    u0r[R::row(PH, 0) * WAVE] = make_double2(pick(v.u0, c.hi0), pick(v.u0, c.hi1));
    kxr[R::row(PH, 0) * WAVE] = make_double2(v.kx0, v.kx1);
    double kyn0 = pick(v.ky, c.hi0), kyn1 = pick(v.ky, c.hi1);
    kyr[R::row(PH, 0) * WAVE] = make_double2(kyn0, kyn1);
    ctb_load(A, c, tab[2 * PH], tab[2 * PH + 1], v);

#pragma unroll
    for (int s = 1; s <= K; ++s) {
        const int z = y - s;
        const double2 kx = nkx, ky = nky, u0 = nu0;
        const double kxe1 = nkxe;
        const double a = st.ca[s - 1], b = st.cb[s - 1];
        if (s < K) {
            nkx = kxr[R::row(PH, s + 1) * WAVE];
            nky = kyr[R::row(PH, s + 1) * WAVE];
            nu0 = u0r[R::row(PH, s + 1) * WAVE];
            nkxe = kxe[2 * WAVE * R::row(PH, s + 1)];
        }
        const double c0 = st.uw[s - 1][S1][0], c1 = st.uw[s - 1][S1][1];
        const double s0 = st.uw[s - 1][S2][0], s1 = st.uw[s - 1][S2][1];
        const double n0 = st.uw[s - 1][S0][0], n1 = st.uw[s - 1][S0][1];
        const double w0 = from_west(c1), e1 = from_east(c0);
        const double e0 = c1, w1 = c0;
        // ChebyIterate: w = A u (SMVP, ext_shared.hpp:37-41), r = u0 - w,
        // p = a p + b r; ChebyCalcU: u += p
        const double m0 = (1.0 + (kx.y + kx.x) + (kyn0 + ky.x)) * c0 - (kx.y * e0 + kx.x * w0) - (kyn0 * n0 + ky.x * s0);
        const double m1 = (1.0 + (kxe1 + kx.y) + (kyn1 + ky.y)) * c1 - (kxe1 * e1 + kx.y * w1) - (kyn1 * n1 + ky.y * s1);
        kyn0 = ky.x;
        kyn1 = ky.y;
        const double r0 = u0.x - m0, r1 = u0.y - m1;
        const double p0 = a * st.pw[s - 1][R1][0] + b * r0;
        const double p1 = a * st.pw[s - 1][R1][1] + b * r1;
        const double un0 = c0 + p0, un1 = c1 + p1;
        if (s < K) {
            st.uw[s][S0][0] = un0;
            st.uw[s][S0][1] = un1;
            st.pw[s][R0][0] = p0;
            st.pw[s][R0][1] = p1;
        } else {
            // every lane stores, lanes and rows that own no cell into a pad pair
            const bool zin = z >= c.Y0 && z <= c.yhi;
            const unsigned bz = zin ? c.bsp : c.bpad;
            const unsigned oz = 8u * (unsigned)(c.g.base + (long)z * c.g.sx);
            stp(A.u_out, oz + bz, make_double2(un0, un1));
            stp(A.p_out, oz + bz, make_double2(p0, p1));
            if (LAST) {
                stp(A.r, oz + bz, make_double2(r0, r1));
                st.acc += zin && c.out0 ? r0 * r0 : 0.0;
                st.acc += zin && c.out1 ? r1 * r1 : 0.0;
            }
        }
    }
}

template <int K, bool LAST, int PH>
__device__ __forceinline__ void ctb_rounds(const CTBArgs& A, const CTBCtx& c, CTBState<K>& st, int y, ConstU* tab)
{
    if constexpr (PH < tb::TB_UNROLL) {
        ctb_round<K, LAST, PH>(A, c, st, y + PH, tab);
        __builtin_amdgcn_sched_barrier(0);
        ctb_rounds<K, LAST, PH + 1>(A, c, st, y, tab);
    }
}

template <int K, bool LAST>
__device__ __forceinline__ void ctb_wave(const CTBArgs& A, const CTBCtx& c)
{
    CTBState<K> st;
#pragma unroll
    for (int s = 0; s < K; ++s) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            st.uw[s][0][q] = st.uw[s][1][q] = st.uw[s][2][q] = 0.0;
            st.pw[s][0][q] = st.pw[s][1][q] = 0.0;
        }
        st.ca[s] = uniform(A.ca[s]);
        st.cb[s] = uniform(A.cb[s]);
    }
    st.acc = 0.0;
    ConstU* tab = (ConstU*)(const unsigned*)A.rowtab + 2 * (c.ylo - A.row0);
    ctb_load(A, c, tab[0], tab[1], st.pf[0]);
    ctb_load(A, c, tab[2], tab[3], st.pf[1]);
    // the first rows complete before the loop, so the loop head waits only
    // for the oldest prefetch slot (as in the PPCG kernel)
    __builtin_amdgcn_s_waitcnt(tb::VMCNT0);
    const int yend = c.yhi + K;
    for (int y = c.ylo; y <= yend; y += tb::TB_UNROLL) ctb_rounds<K, LAST, 0>(A, c, st, y, tab + 2 * (y - c.ylo + 2));
    if (LAST) {
        double v[1] = {st.acc};
        block_partial<1, 1>(v, A.red);
    }
}

template <int K, bool LAST>
__global__ __launch_bounds__(WAVE) void k_cheby_tb(Geom g, CTBArgs A)
{
    using R = tb::TBRing<K>;
    __shared__ double2 u0_ring[R::N * WAVE];
    // one more element: the last lane's east face reads past the last row
    __shared__ double2 kx_ring[R::N * WAVE + 1];
    __shared__ double2 ky_ring[R::N * WAVE];
    const int lane = threadIdx.x;
    const long t = xcd_tile(blockIdx.x, A.nblocks);
    const int tx = (int)(t % A.tiles_x), ty = (int)(t / A.tiles_x);
    constexpr int KP = tb::tb_margin(K);
    constexpr int OW = 2 * WAVE - 2 * KP;
    CTBCtx c;
    c.g = g;
    c.ny = g.ny;
    const int X0 = HALO + tx * OW;
    c.Y0 = HALO + ty * A.rows;
    const int k0 = X0 - KP + 2 * lane, k1 = k0 + 1;
    {
        const tb::CellPair cp = tb::cell_pair(k0, g.nx);
        c.bcp = 8u * (unsigned)cp.start;
        c.hi0 = cp.hi0;
        c.hi1 = cp.hi1;
    }
    c.bf0 = 8u * (unsigned)fold_face(k0, g.nx);
    c.bf1 = 8u * (unsigned)fold_face(k1, g.nx);
    c.out0 = k0 >= X0 && k0 < X0 + OW && k0 <= g.nx + 1;
    c.out1 = k1 >= X0 && k1 < X0 + OW && k1 <= g.nx + 1;
    {
        // pad pairs past the halo are never read; a lane whose first cell
        // is the last of an odd nx also writes the halo column nx + 2 with
        // that cell's mirror image (the PPCG kernel explains why)
        const int xe = (int)g.X + ((int)g.X & 1);
        const int npair = (int)(g.sx - xe) / 2;
        c.bpad = 8u * (unsigned)(xe + 2 * (lane % npair));
        c.bsp = c.out0 ? 8u * (unsigned)k0 : c.bpad;
    }
    c.ylo = c.Y0 - K;
    c.yhi = min(c.Y0 + A.rows - 1, g.ny + 1);
    c.u0r = reinterpret_cast<double*>(u0_ring);
    c.kxr = reinterpret_cast<double*>(kx_ring);
    c.kyr = reinterpret_cast<double*>(ky_ring);
    ctb_wave<K, LAST>(A, c);
}

template <int K, bool LAST>
void launch(Chunk& c, CTBArgs& a)
{
    hipLaunchKernelGGL((k_cheby_tb<K, LAST>), dim3((unsigned)a.nblocks), dim3(WAVE), 0, c.stream, c.g, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
}

// Launchers of the compiled pass sizes (one translation unit each); last =
// the pass writes r and the first phase of ||r||^2.
void ctb_launch_1(Chunk& c, CTBArgs& a, bool last);
void ctb_launch_2(Chunk& c, CTBArgs& a, bool last);
void ctb_launch_5(Chunk& c, CTBArgs& a, bool last);
void ctb_launch_10(Chunk& c, CTBArgs& a, bool last);

#define TL_CTB_DEFINE_LAUNCH(N)                               \
    void ctb_launch_##N(Chunk& c, CTBArgs& a, bool last)      \
    {                                                         \
        if (last) launch<N, true>(c, a);                      \
        else launch<N, false>(c, a);                          \
    }

}  // namespace ctb
#endif

}  // namespace tl
