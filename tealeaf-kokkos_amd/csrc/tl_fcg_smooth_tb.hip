// The flexible CG's preconditioner sweeps as one temporal-blocked pass (2D).
//
// The preconditioner applies K weighted Jacobi sweeps on A z = r,
// z_(s) = z_(s-1) + (r - A z_(s-1)) ((w_1 / D) f_s), D the diagonal of A (1 plus
// the coefficients of the cell's interior faces; a face on the mesh boundary
// carries no flux), starting from z = 0 or from the network's output; the
// sweep weights are w_s = w_1 f_s (f_1 = 1; all f_s = 1 for a fixed weight,
// whose sweeps are then the plain (r - A z) (omega / D)).  Sweep
// s needs sweep s - 1 on the neighbouring rows only, so the wavefront of the
// temporal-blocked Jacobi kernel (tl_jacobi_tb.hpp) runs all K sweeps in one
// pass: sweep s at row y - s from a three-row register window of sweep s - 1,
// r, kx and ky in LDS rings, 24 (32 with a network start) + 8 bytes per cell
// instead of K x 40 and a clearing pass.
//
// Faces on the mesh boundary enter with a zero coefficient: the coefficient
// rows are masked as they enter the rings, so every wave runs one code path
// and a cell's sums are those of the one-sweep kernel, term for term in its
// order (east, west, north, south), with + 0 for a boundary face where that
// kernel skips the term.  Cells outside the mesh are computed from mirrored
// loads and feed no cell inside it (their faces to it are boundary faces).
#include "tl_ppcg_tb.hpp"

#include <algorithm>
#include <string>

namespace tl {

#if TL_DIM == 2
namespace ftb {

using tb::ConstU;
using tb::fold_face;
using tb::from_east;
using tb::from_west;
using tb::ld;
using tb::ldp;
using tb::pick;
using tb::stp;

struct FArgs {
    const double* z_in;  // start (unread when the pass starts from zero)
    double* z_out;
    const double* w;     // with DOTW: the previous direction's A p, for z . w
    double* partials;
    double* r_out;       // with RM 1: r - alpha w (!= r), the previous step's residual update
    const DevScalars* ds;
    const double* r;
    const double* kx;
    const double* ky;
    const uint2* rowtab;
    int row0;
    int ny;
    double w1;    // the first sweep's weight
    double f[4];  // sweep s's weight over w1 (f[0] = 1)
    int tiles_x;
    int rows;
    long nblocks;
};

struct FRow {
    double2 z, r, ky;
    double2 wo;  // w of the output row two rounds on (DOTW)
    double2 wi;  // w of the row itself (RM 1)
    double kx0, kx1;
    bool ky_in;  // the row's south face is inside the mesh
};

template <int K>
struct FState {
    double zw[K][3][2];  // z after sweep s, slot (round) % 3, column
    FRow pf[2];
    double acc;          // z . w over the wave's own cells (DOTW)
    double acc_rr;       // ||r||^2 over the wave's own cells, r as updated (RM 1) or read (RM 2)
    double acc_zr;       // z . r over the wave's own cells (RM 1, 2)
    double2 wd1;         // omega / D of level 1's row, formed this round
};

struct FCtx {
    Geom g;
    int ylo, Y0, yhi;
    unsigned bcp, bf0, bf1, bsp, bpad;
    bool hi0, hi1;
    bool kw0, kw1;  // the lane's x faces (west of c0, west of c1) are inside the mesh
    bool out0, out1;
    double w1, alpha;
    double f[4];
    double* rr;  // LDS rings [ring rows][64][2]
    double* kxr;
    double* kyr;
    double* wdr;  // omega / D of the row's cells
};

template <int K, bool ZERO, bool DOTW, int RM>
__device__ __forceinline__ void f_load(const FArgs& A, const FCtx& c, ConstU* t, int PHk, int y, FRow& v)
{
    const unsigned oy = t[2 * PHk], of = t[2 * PHk + 1];
    if (!ZERO) v.z = ldp(A.z_in, oy + c.bcp);
    if (RM == 1) v.wi = ldp(A.w, oy + c.bcp);
    // the row that leaves the pass when row y enters it (slack rows hold
    // those past the mesh); lanes without output read a pad pair
    if (DOTW) v.wo = ldp(A.w, 8u * (unsigned)(c.g.base + (long)(y - K) * c.g.sx) + c.bsp);
    v.r = ldp(A.r, oy + c.bcp);
    v.kx0 = ld(A.kx, oy + c.bf0);
    v.kx1 = ld(A.kx, oy + c.bf1);
    v.ky = ldp(A.ky, of + c.bcp);
    // the south face of row y is a mesh face iff 3 <= y <= ny + 1 (rows
    // outside the mesh mirror rows inside; their faces are never used)
    v.ky_in = y > HALO && y <= A.ny + 1;
}

// WARM: the first block of a wave (y = ylo + PH = Y0 - K + PH).  The strip's
// output rows depend on sweep s only at rows >= Y0 - (K - s), i.e. from
// round PH = 2 s on, so the arithmetic of the sweeps s > PH / 2 is left out;
// every load and store stays, so that the memory operations in flight at
// the loop head are those of every other block.
template <int K, bool ZERO, bool DOTW, int RM, int PH, bool WARM>
__device__ __forceinline__ void f_round(const FArgs& A, const FCtx& c, FState<K>& st, int y, ConstU* tab)
{
    using R = tb::TBRing<K>;
    constexpr int S0 = PH % 3, S1 = (PH + 2) % 3, S2 = (PH + 1) % 3;  // rounds y, y-1, y-2
    const int lane = threadIdx.x;
    FRow& v = st.pf[PH % 2];
    double2* rr = reinterpret_cast<double2*>(c.rr) + lane;
    double2* kxr = reinterpret_cast<double2*>(c.kxr) + lane;
    double2* kyr = reinterpret_cast<double2*>(c.kyr) + lane;
    double2* wdr = reinterpret_cast<double2*>(c.wdr) + lane;
    const double* kxe = c.kxr + 2 * lane + 2;

    double2 nkx = kxr[R::row(PH, 1) * WAVE], nky = kyr[R::row(PH, 1) * WAVE], nr = rr[R::row(PH, 1) * WAVE];
    double nkxe = kxe[2 * WAVE * R::row(PH, 1)];
    double2 nwd = wdr[R::row(PH, 2) * WAVE];  // level 2's row (level 1's is formed below)

    // row y enters: sweep 0 (the start), r and the masked coefficients
    st.zw[0][S0][0] = ZERO ? 0.0 : pick(v.z, c.hi0);
    st.zw[0][S0][1] = ZERO ? 0.0 : pick(v.z, c.hi1);
    double2 rin = v.r;
    if (RM == 2) {
        // r as read: the wave's own cells sum its square
        const bool yin = y >= c.Y0 && y <= c.yhi;
        const double r0 = pick(rin, c.hi0), r1 = pick(rin, c.hi1);
        st.acc_rr += yin && c.out0 ? r0 * r0 : 0.0;
        st.acc_rr += yin && c.out1 ? r1 * r1 : 0.0;
    }
    if (RM == 1) {
        // CGCalcUr's r -= alpha w of the previous step, on the pair as loaded;
        // the wave's own cells store it and sum its square
        rin.x = v.r.x - c.alpha * v.wi.x;
        rin.y = v.r.y - c.alpha * v.wi.y;
        const double r0 = pick(rin, c.hi0), r1 = pick(rin, c.hi1);
        const bool yin = y >= c.Y0 && y <= c.yhi;
        const unsigned oy = 8u * (unsigned)(c.g.base + (long)y * c.g.sx);
        stp(A.r_out, oy + (yin ? c.bsp : c.bpad), make_double2(r0, r1));
        st.acc_rr += yin && c.out0 ? r0 * r0 : 0.0;
        st.acc_rr += yin && c.out1 ? r1 * r1 : 0.0;
    }
    rr[R::row(PH, 0) * WAVE] = make_double2(pick(rin, c.hi0), pick(rin, c.hi1));
    kxr[R::row(PH, 0) * WAVE] = make_double2(c.kw0 ? v.kx0 : 0.0, c.kw1 ? v.kx1 : 0.0);
    double kyn0 = v.ky_in ? pick(v.ky, c.hi0) : 0.0, kyn1 = v.ky_in ? pick(v.ky, c.hi1) : 0.0;
    kyr[R::row(PH, 0) * WAVE] = make_double2(kyn0, kyn1);
    {
        // omega / D of row y - 1, whose north faces just arrived with row y:
        // D = 1 + east + west + north + south in that order (0 for a face on
        // the mesh boundary), once per cell instead of once per sweep
        double d0 = 1.0, d1 = 1.0;
        d0 += nkx.y;
        d1 += nkxe;
        d0 += nkx.x;
        d1 += nkx.y;
        d0 += kyn0;
        d1 += kyn1;
        d0 += nky.x;
        d1 += nky.y;
        const double2 wd = make_double2(c.w1 / d0, c.w1 / d1);
        wdr[R::row(PH, 1) * WAVE] = wd;
        st.wd1 = wd;
    }
    const double2 wo = v.wo;  // w of this round's output row
    f_load<K, ZERO, DOTW, RM>(A, c, tab, PH, y + 2, v);

#pragma unroll
    for (int s = 1; s <= K; ++s) {
        const int z = y - s;
        const double2 kx = nkx, ky = nky, r = nr;
        const double kxe1 = nkxe;
        const double2 wd = s == 1 ? st.wd1 : nwd;
        const bool skip = WARM && 2 * s > PH;  // a constant once the loop is unrolled
        if (s < K && !(WARM && 2 * (s + 1) > PH)) {
            nkx = kxr[R::row(PH, s + 1) * WAVE];
            nky = kyr[R::row(PH, s + 1) * WAVE];
            nr = rr[R::row(PH, s + 1) * WAVE];
            nkxe = kxe[2 * WAVE * R::row(PH, s + 1)];
            if (s >= 2) nwd = wdr[R::row(PH, s + 1) * WAVE];
        }
        if (skip && s < K) continue;
        double zn0 = 0.0, zn1 = 0.0;
        if (!skip) {
        const double c0 = st.zw[s - 1][S1][0], c1 = st.zw[s - 1][S1][1];
        const double s0 = st.zw[s - 1][S2][0], s1 = st.zw[s - 1][S2][1];
        const double n0 = st.zw[s - 1][S0][0], n1 = st.zw[s - 1][S0][1];
        const double w0 = from_west(c1), e1 = from_east(c0);
        const double e0 = c1, w1 = c0;
        // the one-sweep kernel's order: east, west, north, south
        double az0 = c0, az1 = c1;
        az0 += kx.y * (c0 - e0);
        az1 += kxe1 * (c1 - e1);
        az0 += kx.x * (c0 - w0);
        az1 += kx.y * (c1 - w1);
        az0 += kyn0 * (c0 - n0);
        az1 += kyn1 * (c1 - n1);
        az0 += ky.x * (c0 - s0);
        az1 += ky.y * (c1 - s1);
        const double fs = c.f[s - 1];
        zn0 = c0 + (r.x - az0) * (wd.x * fs);
        zn1 = c1 + (r.y - az1) * (wd.y * fs);
        kyn0 = ky.x;
        kyn1 = ky.y;
        }
        if (s < K) {
            st.zw[s][S0][0] = zn0;
            st.zw[s][S0][1] = zn1;
        } else {
            // every lane stores, lanes and rows that own no cell into a pad pair
            const bool zin = z >= c.Y0 && z <= c.yhi;
            const unsigned bz = zin ? c.bsp : c.bpad;
            const unsigned oz = 8u * (unsigned)(c.g.base + (long)z * c.g.sx);
            stp(A.z_out, oz + bz, make_double2(zn0, zn1));
            if (DOTW) {
                st.acc += zin && c.out0 ? zn0 * wo.x : 0.0;
                st.acc += zin && c.out1 ? zn1 * wo.y : 0.0;
            }
            if (RM != 0) {  // r of the output row (as updated on entry with RM 1)
                st.acc_zr += zin && c.out0 ? zn0 * r.x : 0.0;
                st.acc_zr += zin && c.out1 ? zn1 * r.y : 0.0;
            }
        }
    }
}

template <int K, bool ZERO, bool DOTW, int RM, int PH, bool WARM>
__device__ __forceinline__ void f_rounds(const FArgs& A, const FCtx& c, FState<K>& st, int y, ConstU* tab)
{
    if constexpr (PH < tb::TB_UNROLL) {
        f_round<K, ZERO, DOTW, RM, PH, WARM>(A, c, st, y + PH, tab);
        __builtin_amdgcn_sched_barrier(0);
        f_rounds<K, ZERO, DOTW, RM, PH + 1, WARM>(A, c, st, y, tab);
    }
}

// RM (r mode): 0 r only read; 1 the previous step's r update folded in
// (from zero, with DOTW), ||r||^2 and z.r summed; 2 ||r||^2 and z.r of r as
// read summed (with DOTW).
template <int K, bool ZERO, bool DOTW, int RM>
__global__ __launch_bounds__(WAVE) void k_fcg_smooth_tb(Geom g, FArgs A)
{
    using R = tb::TBRing<K>;
    __shared__ double2 r_ring[R::N * WAVE];
    // one more element: the last lane's east face reads past the last row
    __shared__ double2 kx_ring[R::N * WAVE + 1];
    __shared__ double2 ky_ring[R::N * WAVE];
    __shared__ double2 wd_ring[R::N * WAVE];
    const int lane = threadIdx.x;
    const long t = xcd_tile(blockIdx.x, A.nblocks);
    const int tx = (int)(t % A.tiles_x), ty = (int)(t / A.tiles_x);
    constexpr int KP = tb::tb_margin(K);
    constexpr int OW = 2 * WAVE - 2 * KP;
    FCtx c;
    c.g = g;
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
    // a face is a mesh face iff it lies strictly between two mesh cells
    c.kw0 = k0 > HALO && k0 <= g.nx + 1;
    c.kw1 = k1 > HALO && k1 <= g.nx + 1;
    const bool out0 = k0 >= X0 && k0 < X0 + OW && k0 <= g.nx + 1;
    c.out0 = out0;
    c.out1 = k1 >= X0 && k1 < X0 + OW && k1 <= g.nx + 1;
    {
        // pad pairs past the halo are never read; a lane whose first cell is
        // the last of an odd nx also writes the halo column nx + 2, which no
        // kernel reads
        const int xe = (int)g.X + ((int)g.X & 1);
        const int npair = (int)(g.sx - xe) / 2;
        c.bpad = 8u * (unsigned)(xe + 2 * (lane % npair));
        c.bsp = out0 ? 8u * (unsigned)k0 : c.bpad;
    }
    c.ylo = c.Y0 - K;
    c.yhi = min(c.Y0 + A.rows - 1, g.ny + 1);
    c.w1 = A.w1;
#pragma unroll
    for (int q = 0; q < 4; ++q) c.f[q] = A.f[q];
    c.alpha = RM == 1 ? A.ds->alpha : 0.0;
    c.rr = reinterpret_cast<double*>(r_ring);
    c.kxr = reinterpret_cast<double*>(kx_ring);
    c.kyr = reinterpret_cast<double*>(ky_ring);
    c.wdr = reinterpret_cast<double*>(wd_ring);

    FState<K> st;
#pragma unroll
    for (int s = 0; s < K; ++s)
#pragma unroll
        for (int q = 0; q < 2; ++q) st.zw[s][0][q] = st.zw[s][1][q] = st.zw[s][2][q] = 0.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double2 z = make_double2(0.0, 0.0);
        st.pf[k].z = st.pf[k].r = st.pf[k].ky = st.pf[k].wo = st.pf[k].wi = z;
        st.pf[k].kx0 = st.pf[k].kx1 = 0.0;
        st.pf[k].ky_in = false;
    }
    st.acc = 0.0;
    st.acc_rr = 0.0;
    st.acc_zr = 0.0;
    st.wd1 = make_double2(0.0, 0.0);
    ConstU* tab = (ConstU*)(const unsigned*)A.rowtab + 2 * (c.ylo - A.row0);
    f_load<K, ZERO, DOTW, RM>(A, c, tab, 0, c.ylo, st.pf[0]);
    f_load<K, ZERO, DOTW, RM>(A, c, tab, 1, c.ylo + 1, st.pf[1]);
    __builtin_amdgcn_s_waitcnt(tb::VMCNT0);
    const int yend = c.yhi + K;
    f_rounds<K, ZERO, DOTW, RM, 0, true>(A, c, st, c.ylo, tab + 4);
    for (int y = c.ylo + tb::TB_UNROLL; y <= yend; y += tb::TB_UNROLL)
        f_rounds<K, ZERO, DOTW, RM, 0, false>(A, c, st, y, tab + 2 * (y - c.ylo + 2));
    if (RM != 0) {
        double v[3] = {st.acc, st.acc_rr, st.acc_zr};
        block_partial<3, 1>(v, Reduce{A.partials});
    } else if (DOTW) {
        double v[1] = {st.acc};
        block_partial<1, 1>(v, Reduce{A.partials});
    }
}

}  // namespace ftb

bool fcg_smooth_tb_supported(const Chunk& c, int sweeps)
{
    // the kernel addresses a field with 32-bit byte offsets
    const bool fits = (double)c.g.alloc * sizeof(double) < 4294967296.0;
    return sweeps >= 4 && sweeps % 4 == 0 && c.g.nx >= 2 && c.g.ny >= 2 && fits;
}

// `sweeps` (a multiple of 4) sweeps as sweeps / 4 passes, each from the
// previous one's z, with weights[0 ..) in order; the result in F_Z.
// r_update (from zero): the first pass completes the previous step's
// r -= alpha w, the updated r in F_R.  The last pass leaves z . w (dot_w, or
// with r_update) and, with r_update, ||r||^2 and z . r in the reduction
// partials and returns their count (else 0).
long op_fcg_sweep_passes(Chunk& c, const double* weights, int sweeps, bool from_zero, bool dot_w, bool r_update)
{
    if (!fcg_smooth_tb_supported(c, sweeps)) throw Error("temporal-blocked smoothing: 2D, a multiple of 4 sweeps");
    if (r_update && !from_zero) throw Error("temporal-blocked smoothing: the r update starts from zero");
    const int passes = sweeps / 4;
    long nb = 0;
    for (int q = 0; q < passes; ++q) {
        const bool first = q == 0, last = q + 1 == passes;
        const double* z_in = first && from_zero ? nullptr : c.f[F_Z];
        const double* w = (last && (dot_w || r_update)) || (first && r_update) ? c.f[F_W] : nullptr;
        double* r_out = first && r_update ? c.r_alt : nullptr;
        const bool rsums = last && r_update && !first;
        const long b = op_fcg_smooth_tb(c, z_in, c.f[F_SD], weights, 4, w, r_out, rsums, 4 * q);
        std::swap(c.f[F_Z], c.f[F_SD]);
        if (r_out) std::swap(c.f[F_R], c.r_alt);
        if (last) nb = b;
    }
    return nb;
}

// z_out = `sweeps` weighted Jacobi sweeps on A z = r from z_in (zero when
// z_in is null), in one pass.  z_out != z_in.  With w, the pass also leaves
// the first phase of z_out . w in the reduction partials (returns their
// count, else 0).  With r_out (and w, from zero) it first completes the
// previous step's r -= alpha w (alpha from the device scalars) into r_out,
// sweeps on that r, and leaves z_out . w, ||r_out||^2 and z_out . r_out as
// three partials per block.
long op_fcg_smooth_tb(Chunk& c, const double* z_in, double* z_out, const double* weights, int sweeps,
                      const double* w, double* r_out, bool rsums, int first_sweep)
{
    if (sweeps != 4 || !fcg_smooth_tb_supported(c, sweeps)) throw Error("temporal-blocked smoothing: 2D, 4 sweeps");
    if (r_out && (!w || z_in)) throw Error("temporal-blocked smoothing: the r update needs w and a zero start");
    if (rsums && (!w || !z_in || r_out)) throw Error("temporal-blocked smoothing: r sums need w and a start z");
    ftb::FArgs a{};
    a.z_in = z_in;
    a.z_out = z_out;
    a.w = w;
    a.partials = c.partials;
    a.r_out = r_out;
    a.ds = c.ds;
    a.r = c.f[F_R];
    a.kx = c.f[F_KX];
    a.ky = c.f[F_KY];
    a.rowtab = tb_rowtab(c);
    a.row0 = TB_ROWTAB0;
    a.ny = c.g.ny;
    // every pass factors its weights over the set's first one, as the
    // one-sweep kernel does, so that any split into passes gives the same bits
    a.w1 = weights[0];
    for (int q = 0; q < 4; ++q) a.f[q] = sweep_ratio(weights, first_sweep + q);
    constexpr int K = 4;
    const int ow = 2 * WAVE - 2 * tb::tb_margin(K);
    a.tiles_x = (c.g.nx + ow - 1) / ow;
    a.rows = tb_rows(c.g.ny, a.tiles_x, K, c.simds, tb::TB_UNROLL);
    a.nblocks = (long)a.tiles_x * ((c.g.ny + a.rows - 1) / a.rows);
    if (w && a.nblocks > c.partials_cap) throw Error("temporal-blocked smoothing: more waves than reduction partials");
    const dim3 grid((unsigned)a.nblocks), block(WAVE);
    TL_TIMED(c, r_out ? K_FCG_SMOOTH_R : K_FCG_SMOOTH, {
        if (r_out)
            hipLaunchKernelGGL((ftb::k_fcg_smooth_tb<K, true, true, 1>), grid, block, 0, c.stream, c.g, a);
        else if (rsums)
            hipLaunchKernelGGL((ftb::k_fcg_smooth_tb<K, false, true, 2>), grid, block, 0, c.stream, c.g, a);
        else if (z_in && w)
            hipLaunchKernelGGL((ftb::k_fcg_smooth_tb<K, false, true, 0>), grid, block, 0, c.stream, c.g, a);
        else if (z_in)
            hipLaunchKernelGGL((ftb::k_fcg_smooth_tb<K, false, false, 0>), grid, block, 0, c.stream, c.g, a);
        else if (w)
            hipLaunchKernelGGL((ftb::k_fcg_smooth_tb<K, true, true, 0>), grid, block, 0, c.stream, c.g, a);
        else
            hipLaunchKernelGGL((ftb::k_fcg_smooth_tb<K, true, false, 0>), grid, block, 0, c.stream, c.g, a);
    });
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
    return w ? a.nblocks : 0;
}

#else  // 3D keeps the one-sweep kernel (the flexible CG is 2D)

bool fcg_smooth_tb_supported(const Chunk&, int) { return false; }
long op_fcg_smooth_tb(Chunk&, const double*, double*, const double*, int, const double*, double*, bool, int)
{
    throw Error("temporal-blocked smoothing is implemented for 2D");
}
long op_fcg_sweep_passes(Chunk&, const double*, int, bool, bool, bool)
{
    throw Error("temporal-blocked smoothing is implemented for 2D");
}

#endif

}  // namespace tl
