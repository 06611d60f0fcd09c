// gfx950 kernels for TeaLeaf's linear-solve path and the chunk operations
// that launch them.  Every per-cell expression keeps the operand order of
// the reference functor it replaces (cited at each kernel) and the library
// is built with -ffp-contract=off, so cell values are bitwise those of the
// reference's IEEE evaluation; only the order of the FP64 sums differs.
//
// Boundary handling: a stencil input's neighbour across an external face is
// read as the mirrored interior value.  That is exactly what the reference
// reads after its depth-1 reflective halo update (ext_update_halo.hpp), so
// the solver loop needs no halo launches and the arithmetic is unchanged.
#include "tl_chunk.hpp"

#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <cstring>
#include <cstdlib>

namespace tl {

namespace {

constexpr int FLAT_BLOCK = 256;

struct KPtr {
    const double* __restrict__ kx;
    const double* __restrict__ ky;
    const double* __restrict__ kz;
};

// SMVP: tealeaf_2d/ext_shared.hpp:37-41, tealeaf_3d/ext_shared.hpp:39-45.
__device__ __forceinline__ double smvp(const Geom& g, const double* __restrict__ a, const KPtr& k,
                                       long i, int kk, int jj, int ii)
{
    const double ac = a[i];
    const double aw = a[kk > HALO ? i - 1 : i];
    const double ae = a[kk < g.nx + 1 ? i + 1 : i];
    const double as = a[jj > HALO ? i - g.sx : i];
    const double an = a[jj < g.ny + 1 ? i + g.sx : i];
    const double kxc = k.kx[i], kxe = k.kx[i + 1];
    const double kyc = k.ky[i], kyn = k.ky[i + g.sx];
#if TL_DIM == 3
    const double ab = a[ii > HALO ? i - g.sy : i];
    const double af = a[ii < g.nz + 1 ? i + g.sy : i];
    const double kzc = k.kz[i], kzf = k.kz[i + g.sy];
    return (1.0 + (kxe + kxc) + (kyn + kyc) + (kzf + kzc)) * ac
         - (kxe * ae + kxc * aw)
         - (kyn * an + kyc * as)
         - (kzf * af + kzc * ab);
#else
    (void)ii;
    return (1.0 + (kxe + kxc) + (kyn + kyc)) * ac
         - (kxe * ae + kxc * aw)
         - (kyn * an + kyc * as);
#endif
}

// Diagonal of A (the Jacobi preconditioner's reciprocal), ext_cg_kernel.hpp:119-121.
__device__ __forceinline__ double diag(const Geom& g, const KPtr& k, long i)
{
#if TL_DIM == 3
    return (1.0 + (k.kx[i + 1] + k.kx[i]) + (k.ky[i + g.sx] + k.ky[i]) + (k.kz[i + g.sy] + k.kz[i]));
#else
    return (1.0 + (k.kx[i + 1] + k.kx[i]) + (k.ky[i + g.sx] + k.ky[i]));
#endif
}

// Flat walk over the padded reference domain (setup kernels only).
struct Flat {
    long i;
    int kk, jj, ii;
};

__device__ __forceinline__ bool flat_cell(const Geom& g, long q, Flat& c)
{
    const long n = (long)g.X * g.Y * g.Z;
    if (q >= n) return false;
    c.kk = (int)(q % g.X);
    const long r = q / g.X;
    c.jj = (int)(r % g.Y);
    c.ii = (int)(r / g.Y);
    c.i = gidx(g, c.kk, c.jj, c.ii);
    return true;
}

__device__ __forceinline__ bool in_one(const Geom& g, const Flat& c)
{
#if TL_DIM == 3
    if (!(c.ii > 0 && c.ii < g.Z - 1)) return false;
#endif
    return c.jj > 0 && c.jj < g.Y - 1 && c.kk > 0 && c.kk < g.X - 1;
}

__device__ __forceinline__ bool in_skew(const Geom& g, const Flat& c)
{
#if TL_DIM == 3
    if (!(c.ii >= HALO && c.ii < g.Z - 1)) return false;
#endif
    return c.jj >= HALO && c.jj < g.Y - 1 && c.kk >= HALO && c.kk < g.X - 1;
}

// ----------------------------------------------------------------------------
// Setup kernels

// ext_initialise_chunk.hpp:80-88 (volume = dx*dy[*dz] over the full domain)
__global__ void k_fill(Geom g, double* __restrict__ a, double v)
{
    Flat c;
    if (flat_cell(g, (long)blockIdx.x * FLAT_BLOCK + threadIdx.x, c)) a[c.i] = v;
}

struct StateArgs {
    int n, rect, circ, point;
    const double *density, *energy, *xmin, *xmax, *ymin, *ymax, *zmin, *zmax, *radius;
    const int* geometry;
};

// ext_generate_chunk.hpp:38-86
__global__ void k_generate(Geom g, StateArgs s, const double* __restrict__ vX, const double* __restrict__ vY,
                           const double* __restrict__ vZ, const double* __restrict__ cX,
                           const double* __restrict__ cY, const double* __restrict__ cZ,
                           double* __restrict__ energy0, double* __restrict__ density, double* __restrict__ u)
{
    Flat c;
    if (!flat_cell(g, (long)blockIdx.x * FLAT_BLOCK + threadIdx.x, c)) return;
    const int kk = c.kk, jj = c.jj;
    double e0 = s.energy[0], d0 = s.density[0];
    for (int ss = 1; ss < s.n; ++ss) {
        bool apply = false;
        if (s.geometry[ss] == s.rect) {
            apply = vX[kk + 1] >= s.xmin[ss] && vX[kk] < s.xmax[ss] && vY[jj + 1] >= s.ymin[ss] && vY[jj] < s.ymax[ss];
#if TL_DIM == 3
            apply = apply && vZ[c.ii + 1] >= s.zmin[ss] && vZ[c.ii] < s.zmax[ss];
#endif
        } else if (s.geometry[ss] == s.circ) {
#if TL_DIM == 3
            const double radius = sqrt((cX[kk] - s.xmin[ss]) * (cX[kk] - s.xmin[ss]) +
                                       (cY[jj] - s.ymin[ss]) * (cY[jj] - s.ymin[ss]) +
                                       (cZ[c.ii] - s.zmin[ss]) * (cZ[c.ii] - s.zmin[ss]));
#else
            const double radius = sqrt((cX[kk] - s.xmin[ss]) * (cX[kk] - s.xmin[ss]) +
                                       (cY[jj] - s.ymin[ss]) * (cY[jj] - s.ymin[ss]));
#endif
            apply = radius <= s.radius[ss];
        } else if (s.geometry[ss] == s.point) {
            apply = vX[kk] == s.xmin[ss] && vY[jj] == s.ymin[ss];
#if TL_DIM == 3
            apply = apply && vZ[c.ii] == s.zmin[ss];
#endif
        }
        if (apply) {
            e0 = s.energy[ss];
            d0 = s.density[ss];
        }
    }
    energy0[c.i] = e0;
    density[c.i] = d0;
    if (in_one(g, c)) u[c.i] = e0 * d0;
    (void)vZ; (void)cZ;
}

// ext_field_summary.hpp:20-31: vol, mass, ie, temp over the interior.
__global__ __launch_bounds__(BLOCK) void k_field_summary(Geom g, const double* __restrict__ volume,
                                                         const double* __restrict__ density,
                                                         const double* __restrict__ energy0,
                                                         const double* __restrict__ u, Reduce red)
{
    Strip s = make_strip(g);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const double cellVol = volume[s.i];
        const double cellMass = cellVol * density[s.i];
        v[0] += cellVol;
        v[1] += cellMass;
        v[2] += cellMass * energy0[s.i];
        v[3] += cellMass * u[s.i];
    }
    block_partial<4>(v, red);
}

// ----------------------------------------------------------------------------
// Reflective halo (ext_core.cpp:305-336, ext_update_halo.hpp:11-76 in 2D;
// tealeaf_3d/ext_update_halo.hpp:11-136 in 3D) as one in-place gather.
//
// 2D: the reference updates left/right over every row (halo rows included)
// and then top/bottom over every column, so a cell in the top/bottom band of
// depth <= d takes the value of (mirrored column if its column is in the
// left/right band of depth <= d, else its own column; mirrored row); a cell
// only in the left/right band takes (mirrored column, own row).  3D updates
// face interiors only.  No source cell is written by the same launch, so all
// threads read pre-launch values.
struct HaloFields {
    double* f[NUM_FIELDS];
    int count;
};

__device__ __forceinline__ int hdepth(int k, int n)  // 0 inside, 1 or 2 in the halo
{
    return k < HALO ? HALO - k : (k >= n + HALO ? k - (n + HALO - 1) : 0);
}

__device__ __forceinline__ int mirror(int k, int n)
{
    return k < HALO ? 2 * HALO - 1 - k : 2 * (n + HALO) - 1 - k;
}

__global__ void k_update_halo(Geom g, HaloFields hf, int depth, int mask)
{
    const long t = (long)blockIdx.x * FLAT_BLOCK + threadIdx.x;
    int kk, jj;
#if TL_DIM == 2
    // Ring of width 2: rows 0,1,Y-2,Y-1 over the full width, then columns
    // 0,1,X-2,X-1 over the interior rows.
    const long nrow = 4L * g.X, ncol = 4L * g.ny;
    if (t < nrow) {
        const int r = (int)(t / g.X);
        kk = (int)(t % g.X);
        jj = r < 2 ? r : g.Y - 4 + r;
    } else if (t < nrow + ncol) {
        const long q = t - nrow;
        const int cidx = (int)(q / g.ny);
        jj = HALO + (int)(q % g.ny);
        kk = cidx < 2 ? cidx : g.X - 4 + cidx;
    } else {
        return;
    }
    const int kd = hdepth(kk, g.nx), jd = hdepth(jj, g.ny);
    const bool kin = kd >= 1 && kd <= depth && (mask & (kk < HALO ? 1 : 2));
    const bool jin = jd >= 1 && jd <= depth && (mask & (jj < HALO ? 4 : 8));
    int sk, sj;
    if (jin) {
        sk = kin ? mirror(kk, g.nx) : kk;
        sj = mirror(jj, g.ny);
    } else if (kin) {
        sk = mirror(kk, g.nx);
        sj = jj;
    } else {
        return;
    }
    const long dst = gidx(g, kk, jj, 0), src = gidx(g, sk, sj, 0);
#else
    // Six face slabs of two layers over the face interiors.
    int ii;
    const long nlr = 4L * g.ny * g.nz, ntb = 4L * g.nx * g.nz, nfb = 4L * g.nx * g.ny;
    int sk, sj, si, axis, layer;
    if (t < nlr) {
        axis = 0;
        layer = (int)(t / ((long)g.ny * g.nz));
        const long q = t % ((long)g.ny * g.nz);
        jj = HALO + (int)(q % g.ny);
        ii = HALO + (int)(q / g.ny);
        kk = layer < 2 ? layer : g.X - 4 + layer;
    } else if (t < nlr + ntb) {
        axis = 1;
        const long q0 = t - nlr;
        layer = (int)(q0 / ((long)g.nx * g.nz));
        const long q = q0 % ((long)g.nx * g.nz);
        kk = HALO + (int)(q % g.nx);
        ii = HALO + (int)(q / g.nx);
        jj = layer < 2 ? layer : g.Y - 4 + layer;
    } else if (t < nlr + ntb + nfb) {
        axis = 2;
        const long q0 = t - nlr - ntb;
        layer = (int)(q0 / ((long)g.nx * g.ny));
        const long q = q0 % ((long)g.nx * g.ny);
        kk = HALO + (int)(q % g.nx);
        jj = HALO + (int)(q / g.nx);
        ii = layer < 2 ? layer : g.Z - 4 + layer;
    } else {
        return;
    }
    sk = kk; sj = jj; si = ii;
    if (axis == 0) {
        if (hdepth(kk, g.nx) > depth || !(mask & (kk < HALO ? 1 : 2))) return;
        sk = mirror(kk, g.nx);
    } else if (axis == 1) {
        if (hdepth(jj, g.ny) > depth || !(mask & (jj < HALO ? 4 : 8))) return;
        sj = mirror(jj, g.ny);
    } else {
        if (hdepth(ii, g.nz) > depth || !(mask & (ii < HALO ? 16 : 32))) return;
        si = mirror(ii, g.nz);
    }
    const long dst = gidx(g, kk, jj, ii), src = gidx(g, sk, sj, si);
#endif
    for (int q = 0; q < hf.count; ++q) hf.f[q][dst] = hf.f[q][src];
}

// ----------------------------------------------------------------------------
// CG (ext_cg_kernel.hpp)

// CGInitU (:22-27) over the full domain, CGInitW (:47-55) on the one-halo
// domain and CGInitK (:74-83) on the skew domain, in one pass.  kx/ky/kz
// read w = rho or 1/rho of the left/lower neighbour, which the skew domain
// keeps inside the one-halo domain, so w is recomputed from density.
__global__ void k_cg_init_k(Geom g, int coef, double rx, double ry, double rz, const double* __restrict__ density,
                            const double* __restrict__ energy, double* __restrict__ u, double* __restrict__ p,
                            double* __restrict__ r, double* __restrict__ w, double* __restrict__ kx,
                            double* __restrict__ ky, double* __restrict__ kz)
{
    Flat c;
    if (!flat_cell(g, (long)blockIdx.x * FLAT_BLOCK + threadIdx.x, c)) return;
    const long i = c.i;
    p[i] = 0.0;
    r[i] = 0.0;
    u[i] = energy[i] * density[i];
    if (in_one(g, c)) w[i] = (coef == CONDUCTIVITY) ? density[i] : 1.0 / density[i];
    if (in_skew(g, c)) {
        const double wc = (coef == CONDUCTIVITY) ? density[i] : 1.0 / density[i];
        const double wl = (coef == CONDUCTIVITY) ? density[i - 1] : 1.0 / density[i - 1];
        const double wd = (coef == CONDUCTIVITY) ? density[i - g.sx] : 1.0 / density[i - g.sx];
        kx[i] = rx * (wl + wc) / (2.0 * wl * wc);
        ky[i] = ry * (wd + wc) / (2.0 * wd * wc);
#if TL_DIM == 3
        const double wb = (coef == CONDUCTIVITY) ? density[i - g.sy] : 1.0 / density[i - g.sy];
        kz[i] = rz * (wb + wc) / (2.0 * wb * wc);
#endif
    }
    (void)rz; (void)kz;
}

// CGCalcRRO (:107-132)
__global__ __launch_bounds__(BLOCK) void k_cg_calc_rro(Geom g, KPtr k, int pre, const double* __restrict__ u,
                                                       double* __restrict__ w, double* __restrict__ r,
                                                       double* __restrict__ p, double* __restrict__ mi,
                                                       double* __restrict__ z, Reduce red)
{
    Strip s = make_strip(g);
    double v[1] = {0.0};
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        const double sm = smvp(g, u, k, i, s.kk, s.jj, s.ii);
        w[i] = sm;
        const double rv = u[i] - sm;
        r[i] = rv;
        double pv;
        if (pre) {
            const double m = 1.0 / diag(g, k, i);
            mi[i] = m;
            pv = m * rv;
            z[i] = pv;
        } else {
            pv = rv;
        }
        p[i] = pv;
        v[0] += rv * pv;
    }
    block_partial<1>(v, red);
}

// CGCalcW (:178-189): w = A p, partial p.w.  With FOLD the pending
// CGCalcP of the previous iteration (:301-310, p = beta p + r or z) is
// applied on the fly: every stencil point of p is rebuilt from the old p and
// r (z) with the same expression, the new p goes to p_out (a second buffer,
// since neighbours still read the old one) and A is applied to it.
// With U (and FOLD) the previous iteration's CGCalcU, u += alpha p with that
// iteration's alpha and p (both still in place), rides along: CGCalcUr's
// sweep then updates r only (ext_cg_kernel.hpp:244-264, same expression).
template <bool FOLD, bool U = false>
__global__ __launch_bounds__(BLOCK) void k_cg_calc_w(Geom g, KPtr k, const double* __restrict__ p,
                                                     const double* __restrict__ q, const DevScalars* __restrict__ ds,
                                                     double* __restrict__ p_out, double* __restrict__ w, Reduce red,
                                                     double* __restrict__ u)
{
    if (gated_off(red)) return;
    Strip s = make_strip(g);
    double v[1] = {0.0};
    const double beta = FOLD ? ds->beta : 0.0;
    const double alpha = U ? ds->alpha : 0.0;
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        double sm, pc;
        if (FOLD) {
            const long iw = s.kk > HALO ? i - 1 : i, ie = s.kk < g.nx + 1 ? i + 1 : i;
            const long is = s.jj > HALO ? i - g.sx : i, in = s.jj < g.ny + 1 ? i + g.sx : i;
            pc = beta * p[i] + q[i];
            const double pw_ = beta * p[iw] + q[iw], pe = beta * p[ie] + q[ie];
            const double ps = beta * p[is] + q[is], pn = beta * p[in] + q[in];
            const double kxc = k.kx[i], kxe = k.kx[i + 1], kyc = k.ky[i], kyn = k.ky[i + g.sx];
#if TL_DIM == 3
            const long ib = s.ii > HALO ? i - g.sy : i, iff = s.ii < g.nz + 1 ? i + g.sy : i;
            const double pb = beta * p[ib] + q[ib], pf = beta * p[iff] + q[iff];
            const double kzc = k.kz[i], kzf = k.kz[i + g.sy];
            sm = (1.0 + (kxe + kxc) + (kyn + kyc) + (kzf + kzc)) * pc - (kxe * pe + kxc * pw_) - (kyn * pn + kyc * ps)
               - (kzf * pf + kzc * pb);
#else
            sm = (1.0 + (kxe + kxc) + (kyn + kyc)) * pc - (kxe * pe + kxc * pw_) - (kyn * pn + kyc * ps);
#endif
            p_out[i] = pc;
            if (U) u[i] = u[i] + alpha * p[i];
        } else {
            pc = p[i];
            sm = smvp(g, p, k, i, s.kk, s.jj, s.ii);
        }
        w[i] = sm;
        v[0] += sm * pc;
    }
    block_partial<1>(v, red);
}

// Lane i receives lane i-1 / i+1 through DPP row shifts (wave_shr:1 /
// wave_shl:1); the lane at the wave's edge keeps its own value.
__device__ __forceinline__ double dpp_from_west(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false),
                            __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ double dpp_from_east(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false),
                            __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false));
}

#if TL_DIM == 3
// 3D sweeps along z with a register window.  A lane owns column kk, a wave
// row jj, and the lane walks the block's strip of planes (the tiling of
// make_strip) keeping the stencil operand of planes ii-1, ii, ii+1 in
// registers: x-neighbours come from the adjacent lanes (DPP; at the wave's
// edges one extra load), y-neighbours are loaded, and kz of a plane's back
// face is the previous plane's front face.  Raw inputs of the next plane
// are loaded while the current plane is computed.  The loop count is the
// same for every lane of a wave (lanes past the last column run on the
// clamped column and store nothing), so the DPP sources are always live.
// Same expressions, summation order and partials as the per-cell kernels
// they replace.
struct ZWalk {
    long i;        // device index of the lane's cell in the current plane
    long dys, dyn; // offsets of the y-neighbours (0 on the mesh faces)
    long de;       // offset of the x-neighbour beyond the wave's edge
    int i0, count; // first plane and planes in the strip (wave-uniform)
    int lane;
    bool xok;      // the lane's column is in the mesh (stores, sums)
    bool edge;     // lane 0 or lane 63
    bool east;     // the column is the last one (east neighbour = itself)
};

__device__ inline ZWalk make_zwalk(const Geom& g)
{
    const long t = xcd_tile(blockIdx.x, g.nblocks);
    ZWalk z;
    z.lane = threadIdx.x % WAVE;
    const int wave = threadIdx.x / WAVE;
    const int tx = (int)(t % g.tiles_x);
    const long rest = t / g.tiles_x;
    const int ty = (int)(rest % g.tiles_y);
    const int ts = (int)(rest / g.tiles_y);
    const int kk = HALO + tx * WAVE + z.lane;
    const int jj = HALO + ty * WAVES_PER_BLOCK + wave;
    z.i0 = HALO + ts * g.strip;
    z.count = min(g.strip, g.nz + HALO - z.i0);
    if (jj > g.ny + 1 || z.count < 0) z.count = 0;
    z.xok = kk <= g.nx + 1;
    const int kc = min(kk, g.nx + 1);
    z.east = kc == g.nx + 1;
    z.edge = z.lane == 0 || z.lane == WAVE - 1;
    const int ce = z.lane == 0 ? max(kc - 1, HALO) : min(kc + 1, g.nx + 1);
    z.de = z.edge ? (long)(ce - kc) : 0;  // other lanes load their own cell (no branch around the load)
    z.dys = jj > HALO ? -g.sx : 0;
    z.dyn = jj < g.ny + 1 ? g.sx : 0;
    z.i = gidx(g, kc, jj, z.i0);
    return z;
}

// Index of plane ii + 1 (the lane's column), clamped to the last plane.
__device__ inline long z_next(const Geom& g, long i, int ii)
{
    return ii < g.nz + 1 ? i + g.sy : i;
}

// x-neighbours of the centre value c: adjacent lanes, the extra load ec at
// the wave's edges, the cell itself on the mesh faces (ec is the cell's own
// value at the west face).
__device__ inline void z_xnb(const ZWalk& z, double c, double ec, double& w, double& e)
{
    const double dw = dpp_from_west(c), de = dpp_from_east(c);
    w = z.lane == 0 ? ec : dw;
    e = z.east ? c : (z.lane == WAVE - 1 ? ec : de);
}

// Face coefficients of the current plane, and of the next one in flight.
struct ZCoef {
    double kxc, kxe, kyc, kyn, kzc, kzf;
    double kxc1, kxe1, kyc1, kyn1, kzf1;
    __device__ void load(const Geom& g, const KPtr& k, long i)
    {
        kxc = k.kx[i];
        kxe = k.kx[i + 1];
        kyc = k.ky[i];
        kyn = k.ky[i + g.sx];
        kzc = k.kz[i];
        kzf = k.kz[i + g.sy];
    }
    __device__ void prefetch(const Geom& g, const KPtr& k, long i1)
    {
        kxc1 = k.kx[i1];
        kxe1 = k.kx[i1 + 1];
        kyc1 = k.ky[i1];
        kyn1 = k.ky[i1 + g.sx];
        kzf1 = k.kz[i1 + g.sy];
    }
    __device__ void advance()
    {
        kxc = kxc1;
        kxe = kxe1;
        kyc = kyc1;
        kyn = kyn1;
        kzc = kzf;
        kzf = kzf1;
    }
    // the 7-point SMVP of smvp() from the neighbour values
    __device__ double apply(double c, double w, double e, double s, double n, double b, double f) const
    {
        return (1.0 + (kxe + kxc) + (kyn + kyc) + (kzf + kzc)) * c - (kxe * e + kxc * w) - (kyn * n + kyc * s)
             - (kzf * f + kzc * b);
    }
};

// CGCalcP folded into CGCalcW (k_cg_calc_w<true, U>) along z: the operand
// is the new p, beta p + r (z), rebuilt from the raw p and r (z) of every
// point it is needed at; the old p of the centre feeds the folded u update.
template <bool U>
__global__ __launch_bounds__(BLOCK) void k_cg_pw_z(Geom g, KPtr k, const double* __restrict__ p,
                                                   const double* __restrict__ q, const DevScalars* __restrict__ ds,
                                                   double* __restrict__ p_out, double* __restrict__ w,
                                                   double* __restrict__ u, Reduce red)
{
    if (gated_off(red)) return;
    ZWalk z = make_zwalk(g);
    const double beta = ds->beta;
    const double alpha = U ? ds->alpha : 0.0;
    double acc = 0.0;
    if (z.count > 0) {
        long i = z.i;
        // window: new p of planes ii-1 (b), ii (c), ii+1 (f); old p of ii, ii+1
        const long ib = z.i0 > HALO ? i - g.sy : i;
        double pb = beta * p[ib] + q[ib];
        double poc = p[i];
        double pc = beta * poc + q[i];
        const long if0 = z_next(g, i, z.i0);
        double pof = p[if0];
        double pf = beta * pof + q[if0];
        if (z.i0 == HALO) pb = pc;
        double ps = beta * p[i + z.dys] + q[i + z.dys];
        double pn = beta * p[i + z.dyn] + q[i + z.dyn];
        double ec = beta * p[i + z.de] + q[i + z.de];
        ZCoef kc;
        kc.load(g, k, i);
        double uc = U ? u[i] : 0.0;
        for (int n = 0; n < z.count; ++n) {
            const int ii = z.i0 + n;
            // raw inputs of plane ii + 1 (whose front neighbour is plane ii + 2)
            const long i1 = z_next(g, i, ii);
            const long i2 = z_next(g, i1, ii + 1);
            const double pp2 = p[i2], qq2 = q[i2];
            const double ps1 = beta * p[i1 + z.dys] + q[i1 + z.dys];
            const double pn1 = beta * p[i1 + z.dyn] + q[i1 + z.dyn];
            const double pe1 = p[i1 + z.de], qe1 = q[i1 + z.de];
            kc.prefetch(g, k, i1);
            const double u1 = U ? u[i1] : 0.0;

            double pw_, pe;
            z_xnb(z, pc, ec, pw_, pe);
            const double sm = kc.apply(pc, pw_, pe, ps, pn, pb, ii < g.nz + 1 ? pf : pc);
            if (z.xok) {
                p_out[i] = pc;
                w[i] = sm;
                if (U) u[i] = uc + alpha * poc;
                acc += sm * pc;
            }
            pb = pc;
            pc = pf;
            poc = pof;
            pof = pp2;
            pf = beta * pp2 + qq2;
            ps = ps1;
            pn = pn1;
            ec = beta * pe1 + qe1;
            kc.advance();
            uc = u1;
            i = i1;
        }
    }
    double v[1] = {acc};
    block_partial<1>(v, red);
}

// k_ppcg_inner_fused along z: r -= A sd, u += sd, sd' = a sd + b M r.
__global__ __launch_bounds__(BLOCK) void k_ppcg_inner_z(Geom g, KPtr k, int pre, double a, double b,
                                                        const double* __restrict__ sd, double* __restrict__ r,
                                                        double* __restrict__ u, const double* __restrict__ mi,
                                                        double* __restrict__ sd_out)
{
    ZWalk z = make_zwalk(g);
    if (z.count == 0) return;
    long i = z.i;
    double sb = sd[z.i0 > HALO ? i - g.sy : i];
    double sc = sd[i];
    double sf = sd[z_next(g, i, z.i0)];
    if (z.i0 == HALO) sb = sc;
    double ss = sd[i + z.dys], sn = sd[i + z.dyn];
    double ec = sd[i + z.de];
    ZCoef kc;
    kc.load(g, k, i);
    double rc = r[i], uc = u[i], mc = pre ? mi[i] : 0.0;
    for (int n = 0; n < z.count; ++n) {
        const int ii = z.i0 + n;
        const long i1 = z_next(g, i, ii);
        const long i2 = z_next(g, i1, ii + 1);
        const double sf2 = sd[i2];
        const double ss1 = sd[i1 + z.dys], sn1 = sd[i1 + z.dyn];
        const double ec1 = sd[i1 + z.de];
        kc.prefetch(g, k, i1);
        const double r1 = r[i1], u1 = u[i1], m1 = pre ? mi[i1] : 0.0;

        double sw, se;
        z_xnb(z, sc, ec, sw, se);
        const double sm = kc.apply(sc, sw, se, ss, sn, sb, ii < g.nz + 1 ? sf : sc);
        if (z.xok) {
            const double rv = rc - sm;
            r[i] = rv;
            u[i] = uc + sc;
            sd_out[i] = a * sc + b * (pre ? mc * rv : rv);
        }
        sb = sc;
        sc = sf;
        sf = sf2;
        ss = ss1;
        sn = sn1;
        ec = ec1;
        kc.advance();
        rc = r1;
        uc = u1;
        mc = m1;
        i = i1;
    }
}

// Two PPCG inner steps (PPCGCalcU + PPCGCalcSd twice) in one 3D sweep
// along z.  A workgroup of 16 waves owns a tile of 64 columns (lane l:
// column kk0 - 2 + l) by 16 rows (wave w: row jj0 - 2 + w) and writes the
// 60 x 12 cells in its middle: the first step's sd is needed one cell
// beyond them, the input sd two.  At step t the input sd of plane
// a = i0 - 2 + t arrives, the first step runs at plane a - 1 and the second
// at a - 2; row neighbours go through LDS (two buffers per level, one
// barrier per step), column neighbours through DPP, plane neighbours stay in
// registers, and the next step's inputs are loaded while this one computes.
// Cells outside the mesh run on the clamped cell and are never used: every
// neighbour across a mesh face is replaced by the cell itself, as after the
// reflective halo update.  Neighbouring tiles read the input r and sd at
// their halo, so r goes to r_out and sd to sd_out; u is read only at the
// cells this tile writes and is updated in place.  Same expressions as
// k_ppcg_inner_z: bitwise two of its launches.
constexpr int PZ_WAVES = 16;
constexpr int PZ_TX = WAVE - 4, PZ_TY = PZ_WAVES - 4;

struct PZGeom {
    int tiles_x, tiles_y, strip;
    long nblocks;
};

__global__ __launch_bounds__(WAVE * PZ_WAVES) void k_ppcg_inner2_z(
    Geom g, PZGeom pz, KPtr k, int pre, double a1, double b1, double a2, double b2, const double* __restrict__ sd,
    const double* __restrict__ r, double* __restrict__ u, const double* __restrict__ mi, double* __restrict__ sd_out,
    double* __restrict__ r_out)
{
    __shared__ double l0[2][PZ_WAVES][WAVE];  // input sd of one plane, per row
    __shared__ double l1[2][PZ_WAVES][WAVE];  // first-step sd of one plane, per row
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const long blk = xcd_tile(blockIdx.x, pz.nblocks);
    const int tx = (int)(blk % pz.tiles_x);
    const long rest = blk / pz.tiles_x;
    const int ty = (int)(rest % pz.tiles_y);
    const int ts = (int)(rest / pz.tiles_y);
    const int kk = HALO + tx * PZ_TX - 2 + lane;
    const int jj = HALO + ty * PZ_TY - 2 + w;
    const int i0 = HALO + ts * pz.strip;
    const int i1 = min(i0 + pz.strip, g.nz + HALO);  // output planes [i0, i1)
    const int nz1 = g.nz + 1;
    const int kc = min(max(kk, HALO), g.nx + 1), jc = min(max(jj, HALO), g.ny + 1);
    const bool out = lane >= 2 && lane < WAVE - 2 && w >= 2 && w < PZ_WAVES - 2 && kk <= g.nx + 1 && jj <= g.ny + 1;
    const bool fw = kk > HALO, fe = kk < g.nx + 1, fs = jj > HALO, fn = jj < g.ny + 1;
    const int ws = max(w - 1, 0), wn = min(w + 1, PZ_WAVES - 1);
    const long col = gidx(g, kc, jc, 0);
    auto at = [&](int ii) { return col + (long)min(max(ii, HALO), nz1) * g.sy; };

    // inputs of the coming step: sd of plane a, r, u, mi and faces of plane a - 1
    double nsd = sd[at(i0 - 2)];
    long ib = at(i0 - 3);
    double nr = r[ib], nu = u[ib], nm = pre ? mi[ib] : 0.0;
    ZCoef nk;
    nk.load(g, k, ib);
    double s0b = 0.0, s0c = 0.0, s0f = 0.0;  // input sd of planes b - 1, b, b + 1 (b = a - 1)
    double s1b = 0.0, s1c = 0.0, s1f = 0.0;  // first-step sd of planes c - 1, c, c + 1 (c = a - 2)
    double r1c = 0.0, u1c = 0.0, m1c = 0.0;  // first-step r, u and mi of plane c
    ZCoef kcp = nk;                          // faces of plane c
    const int steps = i1 - i0 + 4;
    for (int t = 0; t < steps; ++t) {
        const int a = i0 - 2 + t, b = a - 1, c = a - 2;
        const double sda = nsd, r0b = nr, u0b = nu, mib = nm;
        const ZCoef kb = nk;
        // the next step's inputs
        nsd = sd[at(a + 1)];
        ib = at(a);
        nr = r[ib];
        nu = u[ib];
        nm = pre ? mi[ib] : 0.0;
        nk.load(g, k, ib);

        s0b = s0c;
        s0c = s0f;
        s0f = sda;
        const int rb = (t + 1) & 1;  // buffers written at step t - 1
        const double s0s = l0[rb][ws][lane], s0n = l0[rb][wn][lane];
        const double s1s = l1[rb][ws][lane], s1n = l1[rb][wn][lane];
        // first step at plane b
        double sd1b = 0.0, r1b = 0.0, u1b = 0.0;
        if (t >= 2) {
            const double dw = dpp_from_west(s0c), de = dpp_from_east(s0c);
            const double sm = kb.apply(s0c, fw ? dw : s0c, fe ? de : s0c, fs ? s0s : s0c, fn ? s0n : s0c,
                                       b > HALO ? s0b : s0c, b < nz1 ? s0f : s0c);
            r1b = r0b - sm;
            u1b = u0b + s0c;
            sd1b = a1 * s0c + b1 * (pre ? mib * r1b : r1b);
        }
        // second step at plane c
        s1b = s1c;
        s1c = s1f;
        s1f = sd1b;
        if (t >= 4) {
            const double dw = dpp_from_west(s1c), de = dpp_from_east(s1c);
            const double sm = kcp.apply(s1c, fw ? dw : s1c, fe ? de : s1c, fs ? s1s : s1c, fn ? s1n : s1c,
                                        c > HALO ? s1b : s1c, c < nz1 ? s1f : s1c);
            const double rv = r1c - sm;
            if (out) {
                const long i = gidx(g, kc, jc, c);
                r_out[i] = rv;
                u[i] = u1c + s1c;
                sd_out[i] = a2 * s1c + b2 * (pre ? m1c * rv : rv);
            }
        }
        r1c = r1b;
        u1c = u1b;
        m1c = mib;
        kcp = kb;
        l0[t & 1][w][lane] = sda;
        l1[t & 1][w][lane] = sd1b;
        __syncthreads();
    }
}

// k_cheby<false, true> along z: r = u0 - A u, p = a p + b M r, u' = u + p.
__global__ __launch_bounds__(BLOCK) void k_cheby_z(Geom g, KPtr k, int pre, double a, double b,
                                                   const double* __restrict__ u, const double* __restrict__ u0,
                                                   const double* __restrict__ mi, double* __restrict__ r,
                                                   double* __restrict__ p, double* __restrict__ u_out)
{
    ZWalk z = make_zwalk(g);
    if (z.count == 0) return;
    long i = z.i;
    double vb = u[z.i0 > HALO ? i - g.sy : i];
    double vc = u[i];
    double vf = u[z_next(g, i, z.i0)];
    if (z.i0 == HALO) vb = vc;
    double vs = u[i + z.dys], vn = u[i + z.dyn];
    double ec = u[i + z.de];
    ZCoef kc;
    kc.load(g, k, i);
    double u0c = u0[i], pc = p[i], mc = pre ? mi[i] : 0.0;
    for (int n = 0; n < z.count; ++n) {
        const int ii = z.i0 + n;
        const long i1 = z_next(g, i, ii);
        const long i2 = z_next(g, i1, ii + 1);
        const double vf2 = u[i2];
        const double vs1 = u[i1 + z.dys], vn1 = u[i1 + z.dyn];
        const double ec1 = u[i1 + z.de];
        kc.prefetch(g, k, i1);
        const double u01 = u0[i1], p1 = p[i1], m1 = pre ? mi[i1] : 0.0;

        double vw, ve;
        z_xnb(z, vc, ec, vw, ve);
        const double sm = kc.apply(vc, vw, ve, vs, vn, vb, ii < g.nz + 1 ? vf : vc);
        if (z.xok) {
            const double rv = u0c - sm;
            r[i] = rv;
            const double zr = pre ? mc * rv : rv;
            const double pv = a * pc + b * zr;
            p[i] = pv;
            u_out[i] = vc + pv;
        }
        vb = vc;
        vc = vf;
        vf = vf2;
        vs = vs1;
        vn = vn1;
        ec = ec1;
        kc.advance();
        u0c = u01;
        pc = p1;
        mc = m1;
        i = i1;
    }
}
#endif

#if TL_DIM == 2

// CGCalcP folded into CGCalcW as a streaming 2D sweep: each lane owns two
// adjacent columns, a wave 128 columns and a strip of `rows` rows; the new p
// of rows j-1, j, j+1 is kept in registers, x-neighbours come from adjacent
// lanes (DPP) and, at the wave's edges, from one extra load.  Every input
// is read from HBM once: 48 bytes per cell.  Same expressions as k_cg_calc_w.
// With U, the previous iteration's CGCalcU (u += alpha p, with that
// iteration's alpha and p, both still in place) is folded in as well, so
// CGCalcUr's second sweep only updates r: 16 more bytes here, 24 fewer there.
struct PWStrip {
    int tiles_x, rows;
    long nwaves;
};

// Strip geometry of the p, w sweeps: strips of 64 rows (fewer strip starts,
// each re-reading two rows of p and r), shorter only while there would be
// under two waves per SIMD, 128 where that still leaves four or more.
// Heights of 1 to 256 rows move the sweep by 1-3% at 4096^2 and 8192^2.
inline PWStrip pw_strip(const Chunk& c)
{
    PWStrip t;
    t.tiles_x = (c.g.nx + 2 * WAVE - 1) / (2 * WAVE);
    t.rows = 64;
    while (t.rows > 4 && (long)t.tiles_x * ((c.g.ny + t.rows - 1) / t.rows) < 2L * c.simds) t.rows /= 2;
    if ((long)t.tiles_x * ((c.g.ny + 127) / 128) >= 4L * c.simds) t.rows = 128;
    t.nwaves = (long)t.tiles_x * ((c.g.ny + t.rows - 1) / t.rows);
    return t;
}

// With PR (the flexible CG), p' . r is summed as well (a second partial),
// r read beside the other inputs.
template <bool U, bool PR = false>
__global__ __launch_bounds__(BLOCK) void k_cg_pw_strip(Geom g, PWStrip t, KPtr k, const double* __restrict__ p,
                                                       const double* __restrict__ q,
                                                       const DevScalars* __restrict__ ds, double* __restrict__ p_out,
                                                       double* __restrict__ w, double* __restrict__ u, Reduce red,
                                                       const double* __restrict__ rin = nullptr)
{
    if (gated_off(red)) return;
    const int lane = threadIdx.x % WAVE;
    // the XCD-aware order is taken per workgroup (workgroups, not waves, go
    // round-robin over the XCDs); a workgroup's waves take consecutive
    // tiles, adjacent column ranges of one strip row, so the column beyond
    // each wave's edge is its neighbour's and stays in that XCD's L2
    const long nblk = (t.nwaves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    const long tw = xcd_tile(blockIdx.x, nblk) * WAVES_PER_BLOCK + threadIdx.x / WAVE;
    double acc = 0.0, acc_pr = 0.0;
    if (tw < t.nwaves) {
        const int tx = (int)(tw % t.tiles_x), ty = (int)(tw / t.tiles_x);
        const int nx = g.nx, ny = g.ny;
        const int c0 = HALO + tx * 2 * WAVE + 2 * lane, c1 = c0 + 1;
        const bool v0 = c0 <= nx + 1, v1 = c1 <= nx + 1;
        // a lane's two columns are one 16-byte load (c0 is even): cells from
        // the pair clamped to start at or before nx + 1, kx faces at or
        // before nx + 2; for an odd nx the last pair holds the halo column
        // nx + 2, which the face select below never reads
        const int k0 = min(c0, (nx + 1) & ~1);
        const int f0 = min(c0, (nx + 2) & ~1);
        // the column beyond the wave's edge (west of lane 0, east of lane 63)
        const int ce = lane == 0 ? max(c0 - 1, HALO) : min(c1 + 1, nx + 1);
        const bool edge = lane == 0 || lane == WAVE - 1;
        const int j0 = HALO + ty * t.rows, j1 = min(j0 + t.rows, ny + HALO);  // rows [j0, j1)
        const double beta = ds->beta;
        const double alpha = U ? ds->alpha : 0.0;
        auto pnew = [&](int col, int row) {
            const long i = gidx(g, col, row, 0);
            return beta * p[i] + q[i];
        };
        auto ld2 = [&](const double* a, long i) { return *reinterpret_cast<const double2*>(a + i); };
        // the new p of a row's column pair; `raw` keeps the old p (for u)
        double2 raw;
        auto pnew2 = [&](int row) {
            const long i = gidx(g, k0, row, 0);
            const double2 pp = ld2(p, i), qq = ld2(q, i);
            raw = pp;
            return make_double2(beta * pp.x + qq.x, beta * pp.y + qq.y);
        };
        // window: rows j-1 (s), j (c), j+1 (n) of the new p, two columns
        const int js = max(j0 - 1, HALO);
        double2 t2 = pnew2(js);
        double s0 = t2.x, s1 = t2.y;
        t2 = pnew2(j0);
        double c0v = t2.x, c1v = t2.y;
        double2 pold_c = raw;  // old p of rows j and j + 1, u of row j
        double2 uc = U ? ld2(u, gidx(g, k0, j0, 0)) : make_double2(0.0, 0.0);
        double2 rc = PR ? ld2(rin, gidx(g, k0, j0, 0)) : make_double2(0.0, 0.0);
        double ec = edge ? pnew(ce, j0) : 0.0;
        if (j0 == HALO) {  // bottom face: the reflective mirror
            s0 = c0v;
            s1 = c1v;
        }
        t2 = ld2(k.ky, gidx(g, k0, j0, 0));
        double kyc0 = t2.x, kyc1 = t2.y;
        // Raw inputs are loaded one row ahead of their use, so each row's
        // loads are in flight while the previous row is computed: for row
        // j + 1 the new p (n, en) and ky; for row j the x-face coefficients.
        const int jn = min(j0 + 1, ny + 1);
        t2 = pnew2(jn);
        double n0 = t2.x, n1 = t2.y;
        double2 pold_n = raw;
        double en = edge ? pnew(ce, jn) : 0.0;
        t2 = ld2(k.ky, gidx(g, k0, j0 + 1, 0));  // rows <= ny + 2
        double kyn0 = t2.x, kyn1 = t2.y;
        t2 = ld2(k.kx, gidx(g, f0, j0, 0));
        double kx0 = t2.x, kx1 = t2.y;
        double kxl = lane == WAVE - 1 ? k.kx[gidx(g, min(c1 + 1, nx + 2), j0, 0)] : 0.0;
        for (int j = j0; j < j1; ++j) {
            // next row's raw inputs (row j + 2 for p, q, ky; row j + 1 for kx)
            const int j2 = min(j + 2, ny + 1), j1k = min(j + 1, ny + 1);
            const long a0 = gidx(g, k0, j2, 0), ae = gidx(g, ce, j2, 0);
            const double2 pp = ld2(p, a0), qq = ld2(q, a0);
            const double pp0 = pp.x, pp1 = pp.y, qq0 = qq.x, qq1 = qq.y;
            const double ppe = edge ? p[ae] : 0.0, qqe = edge ? q[ae] : 0.0;
            // ky of row j + 2's south face = the north face of row j + 1
            const double2 kyq = ld2(k.ky, gidx(g, k0, min(j + 2, ny + 2), 0));
            const double kyq0 = kyq.x, kyq1 = kyq.y;
            const double2 kxq = ld2(k.kx, gidx(g, f0, j1k, 0));
            const double kxq0 = kxq.x, kxq1 = kxq.y;
            const double kxql = lane == WAVE - 1 ? k.kx[gidx(g, min(c1 + 1, nx + 2), j1k, 0)] : 0.0;
            const double2 un = U ? ld2(u, gidx(g, k0, j1k, 0)) : make_double2(0.0, 0.0);
            const double2 rn = PR ? ld2(rin, gidx(g, k0, j1k, 0)) : make_double2(0.0, 0.0);

            double nn0 = n0, nn1 = n1;
            if (j == ny + 1) {  // top face
                nn0 = c0v;
                nn1 = c1v;
            }
            const long i0 = gidx(g, c0, j, 0);
            // kx lives on faces up to nx + 2: the lane past the last column
            // must hold kx of the right face for its west neighbour's DPP
            double kxe1 = dpp_from_east(kx0);
            if (lane == WAVE - 1) kxe1 = kxl;
            double w0 = dpp_from_west(c1v), e1 = dpp_from_east(c0v);
            if (lane == 0) w0 = ec;
            if (lane == WAVE - 1) e1 = ec;
            double e0 = c1v, w1 = c0v;
            if (c0 == HALO) w0 = c0v;
            if (c0 == nx + 1) e0 = c0v;
            if (c1 == nx + 1) e1 = c1v;
            const double m0 = (1.0 + (kx1 + kx0) + (kyn0 + kyc0)) * c0v - (kx1 * e0 + kx0 * w0) - (kyn0 * nn0 + kyc0 * s0);
            const double m1 = (1.0 + (kxe1 + kx1) + (kyn1 + kyc1)) * c1v - (kxe1 * e1 + kx1 * w1) - (kyn1 * nn1 + kyc1 * s1);
            // CGCalcUr's u + alpha p of the previous iteration (c0 = k0 where a cell is stored)
            const double2 uo = make_double2(uc.x + alpha * pold_c.x, uc.y + alpha * pold_c.y);
            if (v1) {
                *reinterpret_cast<double2*>(p_out + i0) = make_double2(c0v, c1v);
                *reinterpret_cast<double2*>(w + i0) = make_double2(m0, m1);
                if (U) *reinterpret_cast<double2*>(u + i0) = uo;
                acc += m0 * c0v;
                acc += m1 * c1v;
                if (PR) {
                    acc_pr += c0v * rc.x;
                    acc_pr += c1v * rc.y;
                }
            } else if (v0) {
                p_out[i0] = c0v;
                w[i0] = m0;
                if (U) u[i0] = uo.x;
                acc += m0 * c0v;
                if (PR) acc_pr += c0v * rc.x;
            }
            pold_c = pold_n;
            pold_n = pp;
            uc = un;
            rc = rn;
            s0 = c0v;
            s1 = c1v;
            c0v = n0;
            c1v = n1;
            ec = en;
            kyc0 = kyn0;
            kyc1 = kyn1;
            kx0 = kxq0;
            kx1 = kxq1;
            kxl = kxql;
            // the new p of row j + 2 from the prefetched values (CGCalcP's beta p + r)
            n0 = beta * pp0 + qq0;
            n1 = beta * pp1 + qq1;
            en = beta * ppe + qqe;
            kyn0 = kyq0;
            kyn1 = kyq1;
        }
    }
    if (PR) {
        double v[2] = {acc, acc_pr};
        block_partial<2>(v, red);
    } else {
        double v[1] = {acc};
        block_partial<1>(v, red);
    }
}
#endif

// CGCalcUr (:244-264); alpha from the device scalars when ds is given.
// Without U the u update is left to the next k_cg_pw_strip<true> (or
// k_cg_calc_u), which applies it with the same alpha and p.
template <bool U>
__global__ __launch_bounds__(BLOCK) void k_cg_calc_ur(Geom g, double alpha_v, const DevScalars* __restrict__ ds,
                                                      int pre, double* __restrict__ u, double* __restrict__ r,
                                                      const double* __restrict__ p, const double* __restrict__ w,
                                                      const double* __restrict__ mi, double* __restrict__ z, Reduce red)
{
    if (gated_off(red)) return;
    Strip s = make_strip(g);
    double v[1] = {0.0};
    const double alpha = ds ? ds->alpha : alpha_v;
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        if (U) u[i] = u[i] + alpha * p[i];
        const double rv = r[i] - alpha * w[i];
        r[i] = rv;
        if (pre) {
            const double zv = mi[i] * rv;
            z[i] = zv;
            v[0] += rv * zv;
        } else {
            v[0] += rv * rv;
        }
    }
    block_partial<1>(v, red);
}

// The u half of CGCalcUr (u += alpha p, alpha from the device scalars): the
// update a k_cg_calc_ur<false> left pending, where no pw strip follows.
__global__ __launch_bounds__(BLOCK) void k_cg_calc_u(Geom g, const DevScalars* __restrict__ ds,
                                                     double* __restrict__ u, const double* __restrict__ p)
{
    Strip s = make_strip(g);
    const double alpha = ds->alpha;
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        u[i] = u[i] + alpha * p[i];
    }
}

// CGCalcP (:301-310); beta from the device scalars when ds is given.
__global__ __launch_bounds__(BLOCK) void k_cg_calc_p(Geom g, double beta_v, const DevScalars* __restrict__ ds,
                                                     double* __restrict__ p, const double* __restrict__ q)
{
    Strip s = make_strip(g);
    const double beta = ds ? ds->beta : beta_v;
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        p[i] = beta * p[i] + q[i];
    }
}

// Boundary cells' p for a CG iteration that follows Chebyshev steps.  One
// value per boundary cell and face, faces in the order west, east
// (ny [x nz] cells), south, north (nx [x nz]) and, in 3D, back, front
// (nx x ny); within a face the slower of its two axes is the outer one.
struct FaceOffsets {
    long w, e, s, n, b, f, count;
};

__host__ __device__ inline FaceOffsets face_offsets(const Geom& g)
{
    FaceOffsets o;
#if TL_DIM == 3
    const long yz = (long)g.ny * g.nz, xz = (long)g.nx * g.nz, xy = (long)g.nx * g.ny;
#else
    const long yz = g.ny, xz = g.nx, xy = 0;
#endif
    o.w = 0;
    o.e = yz;
    o.s = 2 * yz;
    o.n = 2 * yz + xz;
    o.b = 2 * yz + 2 * xz;
    o.f = o.b + xy;
    o.count = o.f + xy;
    return o;
}

// faces[t] = p (or beta p + q with the pending CGCalcP) of the boundary cell
// that face entry t belongs to.
__global__ void k_capture_faces(Geom g, FaceOffsets o, int pending, const DevScalars* __restrict__ ds,
                                const double* __restrict__ p, const double* __restrict__ q,
                                double* __restrict__ faces)
{
    const long t = (long)blockIdx.x * FLAT_BLOCK + threadIdx.x;
    if (t >= o.count) return;
    int kk, jj, ii = 0;
    if (t < o.s) {  // west, east: t' = (ii - 2) ny + (jj - 2)
        const long u = t < o.e ? t : t - o.e;
        kk = t < o.e ? HALO : g.nx + 1;
        jj = HALO + (int)(u % g.ny);
#if TL_DIM == 3
        ii = HALO + (int)(u / g.ny);
#endif
    } else if (t < o.b) {  // south, north: t' = (ii - 2) nx + (kk - 2)
        const long u = t < o.n ? t - o.s : t - o.n;
        jj = t < o.n ? HALO : g.ny + 1;
        kk = HALO + (int)(u % g.nx);
#if TL_DIM == 3
        ii = HALO + (int)(u / g.nx);
#endif
    } else {  // back, front (3D): t' = (jj - 2) nx + (kk - 2)
        const long u = t < o.f ? t - o.b : t - o.f;
        ii = t < o.f ? HALO : g.nz + 1;
        kk = HALO + (int)(u % g.nx);
        jj = HALO + (int)(u / g.nx);
    }
    const long i = gidx(g, kk, jj, ii);
    faces[t] = pending ? ds->beta * p[i] + q[i] : p[i];
}

// CGCalcW (:178-189) whose neighbours outside the mesh are the captured
// faces instead of the reflected cell (the reference's SMVP reads a depth-1
// halo of p that is still the one of the last CG iteration before the
// Chebyshev switch).  Same expression, tiling and partials as k_cg_calc_w.
__global__ __launch_bounds__(BLOCK) void k_cg_calc_w_faces(Geom g, FaceOffsets o, KPtr k,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ faces,
                                                           double* __restrict__ w, Reduce red)
{
    Strip s = make_strip(g);
    double v[1] = {0.0};
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        const int kk = s.kk, jj = s.jj;
#if TL_DIM == 3
        const int ii = s.ii;
        const long yz = (long)(ii - HALO) * g.ny + (jj - HALO), xz = (long)(ii - HALO) * g.nx + (kk - HALO);
#else
        const long yz = jj - HALO, xz = kk - HALO;
#endif
        const double ac = p[i];
        const double aw = kk > HALO ? p[i - 1] : faces[o.w + yz];
        const double ae = kk < g.nx + 1 ? p[i + 1] : faces[o.e + yz];
        const double as = jj > HALO ? p[i - g.sx] : faces[o.s + xz];
        const double an = jj < g.ny + 1 ? p[i + g.sx] : faces[o.n + xz];
        const double kxc = k.kx[i], kxe = k.kx[i + 1];
        const double kyc = k.ky[i], kyn = k.ky[i + g.sx];
#if TL_DIM == 3
        const long xy = (long)(jj - HALO) * g.nx + (kk - HALO);
        const double ab = ii > HALO ? p[i - g.sy] : faces[o.b + xy];
        const double af = ii < g.nz + 1 ? p[i + g.sy] : faces[o.f + xy];
        const double kzc = k.kz[i], kzf = k.kz[i + g.sy];
        const double sm = (1.0 + (kxe + kxc) + (kyn + kyc) + (kzf + kzc)) * ac - (kxe * ae + kxc * aw)
                        - (kyn * an + kyc * as) - (kzf * af + kzc * ab);
#else
        const double sm = (1.0 + (kxe + kxc) + (kyn + kyc)) * ac - (kxe * ae + kxc * aw) - (kyn * an + kyc * as);
#endif
        w[i] = sm;
        v[0] += sm * ac;
    }
    block_partial<1>(v, red);
}

// Last phase of a reduction whose partials a previous launch left: one
// workgroup adds them in block order and applies the epilogue (the sweeps
// that close their own reduction run the same code in their last
// workgroup).
template <int NV>
__global__ __launch_bounds__(BLOCK) void k_finish(double* __restrict__ partials, long nblocks, Finish f)
{
    DevScalars cur{};
    if (threadIdx.x == 0 && f.ds) cur = *f.ds;
    double v[NV];
    sum_partials<NV, WAVES_PER_BLOCK>(partials, nblocks, v);
    if (threadIdx.x == 0) finish_scalars<NV, false>(v, f, cur);
}

// ----------------------------------------------------------------------------
// Chebyshev (ext_cheby_kernel.hpp)

// ChebyInit (:44-54) when first, ChebyIterate (:78-90) otherwise.  The plain
// form writes w and updates p in place; ChebyCalcU (:50-58) follows as its
// own launch.  The fused form also does u += p, writing u_out (!= u).
template <bool FIRST, bool FUSED>
__global__ __launch_bounds__(BLOCK) void k_cheby(Geom g, KPtr k, int pre, double theta, double a, double b,
                                                 const double* __restrict__ u, const double* __restrict__ u0,
                                                 const double* __restrict__ mi, double* __restrict__ w,
                                                 double* __restrict__ r, double* __restrict__ p,
                                                 double* __restrict__ u_out)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        const double sm = smvp(g, u, k, i, s.kk, s.jj, s.ii);
        if (!FUSED) w[i] = sm;
        const double rv = u0[i] - sm;
        r[i] = rv;
        const double zr = pre ? mi[i] * rv : rv;
        const double pv = FIRST ? zr / theta : a * p[i] + b * zr;
        p[i] = pv;
        if (FUSED) u_out[i] = u[i] + pv;
    }
}

// ChebyCalcU (:50-58)
__global__ __launch_bounds__(BLOCK) void k_add_p(Geom g, double* __restrict__ u, const double* __restrict__ p)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) u[s.i] = u[s.i] + p[s.i];
}

// ----------------------------------------------------------------------------
// PPCG (ext_ppcg_kernel.hpp)

// PPCGInitSd (:18-26)
__global__ __launch_bounds__(BLOCK) void k_ppcg_init_sd(Geom g, int pre, double theta, const double* __restrict__ r,
                                                        const double* __restrict__ mi, double* __restrict__ sd)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        sd[i] = (pre ? mi[i] * r[i] : r[i]) / theta;
    }
}

// PPCGCalcU (:47-57): r -= A sd; u += sd
__global__ __launch_bounds__(BLOCK) void k_ppcg_calc_u(Geom g, KPtr k, const double* __restrict__ sd,
                                                       double* __restrict__ r, double* __restrict__ u)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        const double sm = smvp(g, sd, k, i, s.kk, s.jj, s.ii);
        r[i] = r[i] - sm;
        u[i] = u[i] + sd[i];
    }
}

// PPCGCalcSd (:80-89)
__global__ __launch_bounds__(BLOCK) void k_ppcg_calc_sd(Geom g, int pre, double a, double b,
                                                        const double* __restrict__ r, const double* __restrict__ mi,
                                                        double* __restrict__ sd)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        sd[i] = a * sd[i] + b * (pre ? mi[i] * r[i] : r[i]);
    }
}

// PPCGCalcU followed by PPCGCalcSd in one sweep; sd_out != sd.
__global__ __launch_bounds__(BLOCK) void k_ppcg_inner_fused(Geom g, KPtr k, int pre, double a, double b,
                                                            const double* __restrict__ sd,
                                                            double* __restrict__ r, double* __restrict__ u,
                                                            const double* __restrict__ mi,
                                                            double* __restrict__ sd_out)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        const double sdc = sd[i];
        const double sm = smvp(g, sd, k, i, s.kk, s.jj, s.ii);
        const double rv = r[i] - sm;
        r[i] = rv;
        u[i] = u[i] + sdc;
        sd_out[i] = a * sdc + b * (pre ? mi[i] * rv : rv);
    }
}

// ----------------------------------------------------------------------------
// Jacobi (ext_jacobi_kernel.hpp)

// JacobiInit (:126-150)
__global__ void k_jacobi_init(Geom g, int coef, double rx, double ry, double rz, const double* __restrict__ density,
                              const double* __restrict__ energy, double* __restrict__ u0, double* __restrict__ u,
                              double* __restrict__ kx, double* __restrict__ ky, double* __restrict__ kz)
{
    Flat c;
    if (!flat_cell(g, (long)blockIdx.x * FLAT_BLOCK + threadIdx.x, c)) return;
    const long i = c.i;
    if (in_one(g, c)) {
        const double v = energy[i] * density[i];
        u0[i] = v;
        u[i] = v;
    }
    if (in_skew(g, c)) {
        const double dC = (coef == CONDUCTIVITY) ? density[i] : 1.0 / density[i];
        const double dL = (coef == CONDUCTIVITY) ? density[i - 1] : 1.0 / density[i - 1];
        const double dD = (coef == CONDUCTIVITY) ? density[i - g.sx] : 1.0 / density[i - g.sx];
        kx[i] = rx * (dL + dC) / (2.0 * dL * dC);
        ky[i] = ry * (dD + dC) / (2.0 * dD * dC);
#if TL_DIM == 3
        const double dB = (coef == CONDUCTIVITY) ? density[i - g.sy] : 1.0 / density[i - g.sy];
        kz[i] = rz * (dB + dC) / (2.0 * dB * dC);
#endif
    }
    (void)rz; (void)kz;
}

// JacobiSolve (:200-209): un = (u0 + sum k*r_nbr) / (1 + sum k); error += |un - r|.
// `prev` is the previous iterate (the reference's copy r = u); the result
// goes to `next`, which may be the buffer u lived in only when prev != next.
__global__ __launch_bounds__(BLOCK) void k_jacobi_solve(Geom g, KPtr k, const double* __restrict__ u0,
                                                        const double* __restrict__ prev, double* __restrict__ next,
                                                        Reduce red)
{
    Strip s = make_strip(g);
    double v[1] = {0.0};
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const long i = s.i;
        const double rc = prev[i];
        const double rw = prev[s.kk > HALO ? i - 1 : i];
        const double re = prev[s.kk < g.nx + 1 ? i + 1 : i];
        const double rs = prev[s.jj > HALO ? i - g.sx : i];
        const double rn = prev[s.jj < g.ny + 1 ? i + g.sx : i];
        const double kxc = k.kx[i], kxe = k.kx[i + 1], kyc = k.ky[i], kyn = k.ky[i + g.sx];
#if TL_DIM == 3
        const double rb = prev[s.ii > HALO ? i - g.sy : i];
        const double rf = prev[s.ii < g.nz + 1 ? i + g.sy : i];
        const double kzc = k.kz[i], kzf = k.kz[i + g.sy];
        const double un = (u0[i] + (kxe * re + kxc * rw) + (kyn * rn + kyc * rs) + (kzf * rf + kzc * rb))
                        / (1.0 + (kxc + kxe) + (kyc + kyn) + (kzc + kzf));
#else
        const double un = (u0[i] + (kxe * re + kxc * rw) + (kyn * rn + kyc * rs))
                        / (1.0 + (kxc + kxe) + (kyc + kyn));
#endif
        next[i] = un;
        v[0] += fabs(un - rc);
    }
    block_partial<1>(v, red);
}

// ----------------------------------------------------------------------------
// Shared solver methods (ext_solver_methods.hpp)

// CopyU (:19-25)
__global__ __launch_bounds__(BLOCK) void k_copy_u(Geom g, const double* __restrict__ u, double* __restrict__ u0)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) u0[s.i] = u[s.i];
}

// CalculateResidual (:48-58)
__global__ __launch_bounds__(BLOCK) void k_residual(Geom g, KPtr k, const double* __restrict__ u,
                                                    const double* __restrict__ u0, double* __restrict__ r)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        const double sm = smvp(g, u, k, s.i, s.kk, s.jj, s.ii);
        r[s.i] = u0[s.i] - sm;
    }
}

// Calculate2Norm (:80-89)
__global__ __launch_bounds__(BLOCK) void k_2norm(Geom g, const double* __restrict__ a, Reduce red)
{
    Strip s = make_strip(g);
    double v[1] = {0.0};
    for (int n = 0; n < s.count; ++n, strip_advance(s)) v[0] += a[s.i] * a[s.i];
    block_partial<1>(v, red);
}

// Sum over the interior faces (each once; a face on the mesh boundary
// carries no flux) of their coefficients: sum of diag(A) = cells + 2 x this.
__global__ __launch_bounds__(BLOCK) void k_face_sum(Geom g, const double* __restrict__ kx,
                                                    const double* __restrict__ ky, const double* __restrict__ kz,
                                                    Reduce red)
{
    Strip s = make_strip(g);
    double v[1] = {0.0};
    for (int n = 0; n < s.count; ++n, strip_advance(s)) {
        if (s.kk > HALO) v[0] += kx[s.i];
        if (s.jj > HALO) v[0] += ky[s.i];
#if TL_DIM == 3
        if (s.ii > HALO) v[0] += kz[s.i];
#endif
    }
    block_partial<1>(v, red);
}

// a . b over the interior (flexible CG)
__global__ __launch_bounds__(BLOCK) void k_dot(Geom g, const double* __restrict__ a, const double* __restrict__ b,
                                               Reduce red)
{
    Strip s = make_strip(g);
    double v[1] = {0.0};
    for (int n = 0; n < s.count; ++n, strip_advance(s)) v[0] += a[s.i] * b[s.i];
    block_partial<1>(v, red);
}

// p = z + beta p over the interior (flexible CG direction); beta from the
// device scalars when ds is given (0 on the first iteration)
__global__ __launch_bounds__(BLOCK) void k_fcg_p(Geom g, double beta_v, const DevScalars* __restrict__ ds,
                                                 int first, const double* __restrict__ z, double* __restrict__ p)
{
    Strip s = make_strip(g);
    const double beta = ds ? (first ? 0.0 : ds->beta) : beta_v;
    for (int n = 0; n < s.count; ++n, strip_advance(s)) p[s.i] = z[s.i] + beta * p[s.i];
}

// Finalise (:104-111)
__global__ __launch_bounds__(BLOCK) void k_finalise(Geom g, const double* __restrict__ u,
                                                    const double* __restrict__ density, double* __restrict__ energy)
{
    Strip s = make_strip(g);
    for (int n = 0; n < s.count; ++n, strip_advance(s)) energy[s.i] = u[s.i] / density[s.i];
}

inline unsigned flat_blocks(const Geom& g)
{
    const long n = (long)g.X * g.Y * g.Z;
    return (unsigned)((n + FLAT_BLOCK - 1) / FLAT_BLOCK);
}

inline KPtr kptr(const Chunk& c) { return KPtr{c.f[F_KX], c.f[F_KY], c.f[F_KZ]}; }

inline void check_launch()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
}

#define STRIP_LAUNCH(kernel, c, ...)                                                          \
    do {                                                                                      \
        hipLaunchKernelGGL(kernel, dim3((unsigned)(c).g.nblocks), dim3(BLOCK), 0, (c).stream, \
                           (c).g, __VA_ARGS__);                                               \
        check_launch();                                                                       \
    } while (0)

#define TIMED TL_TIMED

#define FLAT_LAUNCH(kernel, c, ...)                                                               \
    do {                                                                                          \
        hipLaunchKernelGGL(kernel, dim3(flat_blocks((c).g)), dim3(FLAT_BLOCK), 0, (c).stream, (c).g, \
                           __VA_ARGS__);                                                          \
        check_launch();                                                                           \
    } while (0)

// Plain device-to-device copy of n doubles, 16 bytes per lane per step: the
// calibration kernel for HBM byte counters (it moves exactly 8n bytes each way).
constexpr int COPY_BLOCK = 512;
constexpr unsigned COPY_GRID = 65536;  // the fastest shape of a launch sweep on MI355X

__global__ __launch_bounds__(COPY_BLOCK) void k_stream_copy(const double2* __restrict__ a, double2* __restrict__ b,
                                                            long n2)
{
    for (long i = (long)blockIdx.x * COPY_BLOCK + threadIdx.x; i < n2; i += (long)gridDim.x * COPY_BLOCK) b[i] = a[i];
}

// Streaming probes of the attainable HBM rate: one 16-byte element per
// thread over the whole array (the launch shape that streams fastest), plain
// or with nontemporal loads and stores; read-only, write-only; and the stream
// mixes of the solver's sweeps: 4 reads + 2 writes (p, w strip), 6 reads +
// 2 writes (PPCG outer iteration).
typedef double d2v __attribute__((ext_vector_type(2)));
template <bool NT>
__global__ __launch_bounds__(256) void k_probe_copy(const d2v* __restrict__ a, d2v* __restrict__ b, long n2)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n2) return;
    if (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
    else b[i] = a[i];
}
__global__ __launch_bounds__(256) void k_probe_read(const d2v* __restrict__ a, long n2, double* __restrict__ out)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n2) return;
    const d2v v = a[i];
    if (v.x + v.y == -1.0) out[0] = v.x;  // keeps the load; the arrays hold zeros
}
__global__ __launch_bounds__(256) void k_probe_write(d2v* __restrict__ b, long n2)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n2) b[i] = d2v{1.0, 2.0};
}
struct ProbeStreams {
    const d2v* r[6];
    d2v* w[2];
};
template <int NR>
__global__ __launch_bounds__(256) void k_probe_mix(ProbeStreams s, long n2)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n2) return;
    d2v v[NR];
    for (int k = 0; k < NR; ++k) v[k] = s.r[k][i];
    s.w[0][i] = v[0] + v[1] * v[2];
    s.w[1][i] = NR > 4 ? v[3] + v[4] * v[5] : v[3] + v[1];
}

// Probe of the grid reduction: workgroup b contributes x[b] as its partial
// (thread 0 holds it, the other lanes 0), then the reduction closes in the
// last workgroup (single pass) or waits for k_finish.
template <int NWAVES>
__global__ __launch_bounds__(NWAVES * WAVE) void k_reduce_probe(const double* __restrict__ x, Reduce red)
{
    double v[1] = {threadIdx.x == 0 ? x[blockIdx.x] : 0.0};
    block_partial<1, NWAVES>(v, red);
}

// Last phase of a reduction whose first phase wrote `nblocks` partials.
template <int NV>
void finish(Chunk& c, long nblocks, Slot slot, int ep, int hist_n = -1)
{
    Finish f = c.finish_args(slot, ep, hist_n);
    f.counter = nullptr;
    hipLaunchKernelGGL(k_finish<NV>, dim3(1), dim3(BLOCK), 0, c.stream, c.partials, nblocks, f);
    check_launch();
}

}  // namespace

// ============================================================================
// Kernel timer

static bool g_single_pass = false;
void set_single_pass(bool on) { g_single_pass = on; }
bool single_pass_default() { return g_single_pass; }

static int g_timing_from = 0;
static int g_timing_only = -1;
void set_kernel_timing_from(int step) { g_timing_from = step < 0 ? 0 : step; }
int kernel_timing_from() { return g_timing_from; }
void set_kernel_timing_only(int kid) { g_timing_only = kid; }
int kernel_timing_only() { return g_timing_only; }

int kernel_id(const char* name)
{
    for (int k = 0; name && k < K_COUNT; ++k)
        if (std::strcmp(kernel_name(k), name) == 0) return k;
    return -1;
}

const char* kernel_name(int k)
{
    static const char* names[K_COUNT] = {"cg_init", "cg_calc_w", "cg_calc_p_w", "cg_calc_ur", "cg_calc_p",
                                         "cheby_iterate", "ppcg_init_sd", "ppcg_inner", "ppcg_outer",
                                         "jacobi_solve", "calc_residual", "calc_2norm", "copy_u", "finalise",
                                         "field_summary", "update_halo", "ppcg_outer_mixed", "cheby_tb",
                                         "jacobi_tb", "cg_calc_p_w_u", "cg_calc_r", "cg_calc_u", "fcg_smooth",
                                         "fcg_p_w", "fcg_smooth_r", "fcg_p_w_zr", "ppcg_inner2"};
    return (k >= 0 && k < K_COUNT) ? names[k] : "?";
}

double kernel_bytes_per_cell(int k)
{
    const double kb = 8.0 * TL_DIM;  // the kx, ky[, kz] coefficient arrays
    switch (k) {
    case K_CG_W: return 16.0 + kb;         // p, coefficients -> w
    case K_CG_WP: return 32.0 + kb;        // p, r, coefficients -> p', w
    case K_PPCG_OUTER: return 48.0 + kb;   // u, p, r, w, coefficients -> u, r
    case K_CG_UR: return 48.0;             // u, p, r, w -> u, r
    case K_CG_P: return 24.0;              // p, r -> p
    case K_CHEBY: return 48.0 + kb;        // u, u0, p, coefficients -> r, p, u'
    case K_PPCG_SD: return 16.0;           // r -> sd
    case K_PPCG_INNER: return 48.0 + kb;   // sd, r, u, coefficients -> r, u, sd'
    case K_PPCG_INNER2: return 48.0 + kb;  // the same for two inner steps
    case K_JACOBI: return 24.0 + kb;       // u, u0, coefficients -> u'
    case K_RESIDUAL: return 24.0 + kb;     // u, u0, coefficients -> r
    case K_2NORM: return 8.0;
    case K_COPY_U: return 16.0;
    case K_FINALISE: return 24.0;
    case K_SUMMARY: return 32.0;
    case K_PPCG_MIXED: return 48.0 + kb / 2;  // u, p, r, w, FP32 coefficients -> u, r
    case K_CHEBY_TB: return 48.0 + kb;        // u, p, u0, coefficients -> u', p', r (per pass)
    case K_JACOBI_TB: return 32.0 + kb;       // u, u0, coefficients -> u', r (per pass)
    case K_CG_WPU: return 48.0 + kb;          // u, p, r, coefficients -> u', p', w
    case K_CG_R: return 24.0;                 // r, w -> r
    case K_CG_U: return 24.0;                 // u, p -> u
    case K_FCG_SMOOTH: return 24.0 + kb;      // r, w, coefficients -> z (4 sweeps, z.w)
    case K_FCG_STRIP: return 56.0 + kb;       // u, p, z, r, coefficients -> u', p', w (p.w, p.r)
    case K_FCG_SMOOTH_R: return 32.0 + kb;    // r, w, coefficients -> r', z (4 sweeps, z.w, r'.r', z.r')
    case K_FCG_STRIP_ZR: return 48.0 + kb;    // u, p, z, coefficients -> u', p', w (p.w)
    default: return 0.0;
    }
}

hipEvent_t KernelTimer::take()
{
    if (pool.empty()) {
        hipEvent_t e;
        TL_HIP(hipEventCreate(&e));
        return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
}

void KernelTimer::begin(int k, hipStream_t s, Pending& p)
{
    if (pending.size() >= 512) flush();
    p.k = k;
    p.a = take();
    p.b = take();
    TL_HIP(hipEventRecord(p.a, s));
}

void KernelTimer::end(Pending& p, hipStream_t s)
{
    TL_HIP(hipEventRecord(p.b, s));
    pending.push_back(p);
}

void KernelTimer::flush()
{
    if (pending.empty()) return;
    TL_HIP(hipEventSynchronize(pending.back().b));
    for (Pending& p : pending) {
        float t = 0.0f;
        TL_HIP(hipEventElapsedTime(&t, p.a, p.b));
        ms[p.k] += t;
        count[p.k] += 1;
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
    ++flushes;
}

void KernelTimer::discard_since(const Mark& m)
{
    if (m.flushes != flushes) return;
    for (size_t i = m.n; i < pending.size(); ++i) {
        pool.push_back(pending[i].a);
        pool.push_back(pending[i].b);
    }
    if (m.n < pending.size()) pending.resize(m.n);
}

KernelTimer::~KernelTimer()
{
    for (Pending& p : pending) {
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

// ============================================================================
// Chunk

static Geom make_geom(int nx, int ny, int nz)
{
    Geom g{};
    g.nx = nx;
    g.ny = ny;
    g.X = nx + 2 * HALO;
    g.Y = ny + 2 * HALO;
#if TL_DIM == 3
    g.nz = nz;
    g.Z = nz + 2 * HALO;
#else
    (void)nz;
    g.nz = 1;
    g.Z = 1;
#endif
    g.sx = ((XOFF + g.X + ROW_ALIGN - 1) / ROW_ALIGN) * ROW_ALIGN;
    g.sy = g.sx * g.Y;
    g.base = XOFF + (long)SLACK_ROWS * g.sx;
    g.alloc = g.base + g.sy * g.Z + (long)SLACK_ROWS * g.sx + ROW_ALIGN;
    g.tiles_x = (nx + WAVE - 1) / WAVE;
#if TL_DIM == 3
    g.tiles_y = (ny + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
    const long cols = (long)g.tiles_x * g.tiles_y;
    const int len = nz;
    const int per = 1;
#else
    g.tiles_y = 1;
    const long cols = g.tiles_x;
    const int len = ny;
    const int per = WAVES_PER_BLOCK;
#endif
    // Longest strip (<= 16) that still gives >= 2048 workgroups.
    int strip = 16;
    while (strip > 1 && cols * ((len + (long)per * strip - 1) / ((long)per * strip)) < 2048) strip /= 2;
    g.strip = strip;
    g.tiles_s = (int)((len + (long)per * strip - 1) / ((long)per * strip));
    g.nblocks = cols * g.tiles_s;
    return g;
}

Chunk::Chunk(int nx, int ny, int nz, int rank_) : rank(rank_)
{
#if TL_DIM == 3
    if (nx < 2 || ny < 2 || nz < 2) throw Error("a chunk needs at least 2 interior cells per axis");
#else
    if (nx < 2 || ny < 2) throw Error("a chunk needs at least 2 interior cells per axis");
#endif
    g = make_geom(nx, ny, nz);
    TL_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    const size_t bytes = sizeof(double) * (size_t)g.alloc;
    for (int k = 0; k < F_COUNT; ++k) {
#if TL_DIM == 2
        if (k == F_KZ) continue;
#endif
        TL_HIP(hipMalloc(&f[k], bytes));
        TL_HIP(hipMemsetAsync(f[k], 0, bytes, stream));
    }
    for (double** q : {&u_alt, &sd_alt, &p_alt, &r_alt}) {
        TL_HIP(hipMalloc(q, bytes));
        TL_HIP(hipMemsetAsync(*q, 0, bytes, stream));
    }
    TL_HIP(hipMalloc(&vertexX, sizeof(double) * (g.X + 1)));
    TL_HIP(hipMalloc(&vertexY, sizeof(double) * (g.Y + 1)));
    TL_HIP(hipMalloc(&vertexZ, sizeof(double) * (g.Z + 1)));
    TL_HIP(hipMalloc(&cellX, sizeof(double) * g.X));
    TL_HIP(hipMalloc(&cellY, sizeof(double) * g.Y));
    TL_HIP(hipMalloc(&cellZ, sizeof(double) * g.Z));
    partials_cap = std::max<long>(g.nblocks, 65536);
    TL_HIP(hipMalloc(&partials, sizeof(double) * 4 * (size_t)partials_cap));
    int dev = 0, cus = 0;
    TL_HIP(hipGetDevice(&dev));
    TL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus > 0) simds = 4 * cus;
    TL_HIP(hipMalloc(&ds, sizeof(DevScalars)));
    TL_HIP(hipMemsetAsync(ds, 0, sizeof(DevScalars), stream));
    single_pass = single_pass_default();
    TL_HIP(hipMalloc(&d_counter, sizeof(unsigned) * reduce_counters(partials_cap)));
    TL_HIP(hipMemsetAsync(d_counter, 0, sizeof(unsigned) * reduce_counters(partials_cap), stream));
    TL_HIP(hipMalloc(&d_slots, sizeof(double) * 4 * (size_t)partials_cap));
    TL_HIP(hipMemsetAsync(d_slots, 0xff, sizeof(double) * 4 * (size_t)partials_cap, stream));  // SLOT_EMPTY
    TL_HIP(hipMalloc(&d_result, sizeof(double) * 4 * S_COUNT));
    double* h = nullptr;
    TL_HIP(hipHostMalloc(&h, sizeof(double) * 4 * S_COUNT, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(h, 0, sizeof(double) * 4 * S_COUNT);
    h_result = h;
    TL_HIP(hipHostGetDevicePointer((void**)&h_result_dev, h, 0));
    unsigned* hs = nullptr;
    TL_HIP(hipHostMalloc(&hs, sizeof(unsigned) * S_COUNT, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(hs, 0, sizeof(unsigned) * S_COUNT);
    h_seq = hs;
    TL_HIP(hipHostGetDevicePointer((void**)&h_seq_dev, hs, 0));
    TL_HIP(hipStreamSynchronize(stream));
}

Chunk::~Chunk()
{
    if (stream) (void)hipStreamSynchronize(stream);
    for (double* p : f) (void)hipFree(p);
    (void)hipFree(u_alt);
    (void)hipFree(sd_alt);
    (void)hipFree(p_alt);
    (void)hipFree(r_alt);
    (void)hipFree(p_faces);
    (void)hipFree(kxy32);
    (void)hipFree(d_rowtab);
    (void)hipFree(d_pack);
    (void)hipFree(d_cheby);
    (void)hipFree(d_levels);
    (void)hipFree(d_coef32);
    (void)hipFree(d_coef);
    (void)hipFree(ds);
    (void)hipFree(d_counter);
    (void)hipFree(d_slots);
    (void)hipFree(hist_a);
    (void)hipFree(hist_b);
    (void)hipFree(vertexX);
    (void)hipFree(vertexY);
    (void)hipFree(vertexZ);
    (void)hipFree(cellX);
    (void)hipFree(cellY);
    (void)hipFree(cellZ);
    (void)hipFree(partials);
    (void)hipFree(d_result);
    if (h_result) (void)hipHostFree((void*)h_result);
    if (h_seq) (void)hipHostFree((void*)h_seq);
    if (stream) (void)hipStreamDestroy(stream);
}

void Chunk::ensure_history(int n)
{
    if (n <= hist_cap) return;
    sync();
    (void)hipFree(hist_a);
    (void)hipFree(hist_b);
    hist_a = hist_b = nullptr;
    TL_HIP(hipMalloc(&hist_a, sizeof(double) * (size_t)(n + 1)));
    TL_HIP(hipMalloc(&hist_b, sizeof(double) * (size_t)(n + 1)));
    TL_HIP(hipMemsetAsync(hist_a, 0, sizeof(double) * (size_t)(n + 1), stream));
    TL_HIP(hipMemsetAsync(hist_b, 0, sizeof(double) * (size_t)(n + 1), stream));
    hist_cap = n;
}

void Chunk::sync() const
{
    TL_HIP(hipStreamSynchronize(stream));
    if (timer.on) timer.flush();
}

// A slot's result once its reduction's second phase has written it: the
// host spins on the number that phase writes after the result (the device
// fences the result first), so the next launches need not wait for a
// stream synchronisation.  A wait longer than 50 ms, or more than 1024
// pending timer events, falls back to synchronising the stream (which also
// reports a failed launch).
Finish Chunk::finish_args(Slot slot, int ep, int hist_n)
{
    const unsigned seq = ++fin_seq[slot];
    return Finish{d_counter, d_slots, ds, d_result + 4 * slot, h_result_dev + 4 * slot, hist_a, hist_b,
                  h_seq_dev + slot, ep, hist_n, seq, (unsigned)((partials_cap + REDUCE_GROUP - 1) / REDUCE_GROUP),
                  gate ? 1 : 0, stop_eps, stop_le, stop_gt};
}

Reduce Chunk::reducer(Slot slot, int ep, int hist_n)
{
    Finish f = finish_args(slot, ep, hist_n);
    pending_open = !single_pass;
    if (pending_open) {
        pending = f;
        f.counter = nullptr;
    }
    return Reduce{partials, f};
}

void Chunk::close(long nblocks, int nv)
{
    if (!pending_open) return;
    pending_open = false;
    Finish f = pending;
    f.counter = nullptr;
    switch (nv) {
    case 1: hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(BLOCK), 0, stream, partials, nblocks, f); break;
    case 2: hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(BLOCK), 0, stream, partials, nblocks, f); break;
    case 3: hipLaunchKernelGGL(k_finish<3>, dim3(1), dim3(BLOCK), 0, stream, partials, nblocks, f); break;
    case 4: hipLaunchKernelGGL(k_finish<4>, dim3(1), dim3(BLOCK), 0, stream, partials, nblocks, f); break;
    default: throw Error("reductions carry 1 to 4 sums");
    }
    check_launch();
}

double Chunk::result(Slot s, int k) const
{
    const unsigned want = fin_seq[s];
    if (timer.pending.size() < 1024) {
        const auto t0 = std::chrono::steady_clock::now();
        for (long i = 1; h_seq[s] != want; ++i)
            if ((i & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) break;
    }
    if (h_seq[s] != want) sync();
    if (h_seq[s] != want) throw Error("reduction result of slot " + std::to_string((int)s) + " not delivered");
    std::atomic_thread_fence(std::memory_order_acquire);
    return h_result[4 * s + k];
}

void Chunk::get_field(int id, double* host) const
{
    if (id < 0 || id >= F_COUNT || !f[id]) throw Error("invalid field id " + std::to_string(id));
    TL_HIP(hipMemcpy2DAsync(host, sizeof(double) * g.X, f[id] + g.base, sizeof(double) * g.sx,
                            sizeof(double) * g.X, (size_t)g.Y * g.Z, hipMemcpyDeviceToHost, stream));
    sync();
}

void Chunk::set_field(int id, const double* host)
{
    if (id < 0 || id >= F_COUNT || !f[id]) throw Error("invalid field id " + std::to_string(id));
    TL_HIP(hipMemcpy2DAsync(f[id] + g.base, sizeof(double) * g.sx, host, sizeof(double) * g.X,
                            sizeof(double) * g.X, (size_t)g.Y * g.Z, hipMemcpyHostToDevice, stream));
    sync();
}

// ============================================================================
// Operations

// ext_core.cpp:33-52 (vertices and cell centres are a few thousand values,
// computed on the host with the reference's expressions and copied once).
void op_initialise_chunk(Chunk& c, double xmin, double ymin, double zmin, double dx, double dy, double dz)
{
    c.dx = dx;
    c.dy = dy;
    const Geom& g = c.g;
    std::vector<double> vx(g.X + 1), vy(g.Y + 1), vz(g.Z + 1), cx(g.X), cy(g.Y), cz(g.Z);
    for (int i = 0; i < g.X + 1; ++i) vx[i] = xmin + dx * (i - HALO);
    for (int i = 0; i < g.Y + 1; ++i) vy[i] = ymin + dy * (i - HALO);
    for (int i = 0; i < g.X; ++i) cx[i] = 0.5 * (vx[i] + vx[i + 1]);
    for (int i = 0; i < g.Y; ++i) cy[i] = 0.5 * (vy[i] + vy[i + 1]);
#if TL_DIM == 3
    c.dz = dz;
    for (int i = 0; i < g.Z + 1; ++i) vz[i] = zmin + dz * (i - HALO);
    for (int i = 0; i < g.Z; ++i) cz[i] = 0.5 * (vz[i] + vz[i + 1]);
    const double vol = dx * dy * dz;
#else
    (void)zmin;
    (void)dz;
    const double vol = dx * dy;
#endif
    TL_HIP(hipMemcpyAsync(c.vertexX, vx.data(), sizeof(double) * vx.size(), hipMemcpyHostToDevice, c.stream));
    TL_HIP(hipMemcpyAsync(c.vertexY, vy.data(), sizeof(double) * vy.size(), hipMemcpyHostToDevice, c.stream));
    TL_HIP(hipMemcpyAsync(c.vertexZ, vz.data(), sizeof(double) * vz.size(), hipMemcpyHostToDevice, c.stream));
    TL_HIP(hipMemcpyAsync(c.cellX, cx.data(), sizeof(double) * cx.size(), hipMemcpyHostToDevice, c.stream));
    TL_HIP(hipMemcpyAsync(c.cellY, cy.data(), sizeof(double) * cy.size(), hipMemcpyHostToDevice, c.stream));
    TL_HIP(hipMemcpyAsync(c.cellZ, cz.data(), sizeof(double) * cz.size(), hipMemcpyHostToDevice, c.stream));
    FLAT_LAUNCH(k_fill, c, c.f[F_VOLUME], vol);
    c.sync();  // the host vectors die here
}

// ext_core.cpp:57-126
void op_generate_chunk(Chunk& c, int n, const double* density, const double* energy, const double* xmin,
                       const double* xmax, const double* ymin, const double* ymax, const double* zmin,
                       const double* zmax, const double* radius, const int* geometry, int rect, int circ,
                       int point)
{
    if (n < 1) throw Error("generate_chunk needs at least one state");
    std::vector<double> hs(9 * (size_t)n, 0.0);
    const double* src[9] = {density, energy, xmin, xmax, ymin, ymax, zmin, zmax, radius};
    for (int a = 0; a < 9; ++a)
        if (src[a]) std::memcpy(&hs[(size_t)a * n], src[a], sizeof(double) * n);
    double* ds = nullptr;
    int* dg = nullptr;
    TL_HIP(hipMalloc(&ds, sizeof(double) * hs.size()));
    TL_HIP(hipMalloc(&dg, sizeof(int) * n));
    TL_HIP(hipMemcpyAsync(ds, hs.data(), sizeof(double) * hs.size(), hipMemcpyHostToDevice, c.stream));
    TL_HIP(hipMemcpyAsync(dg, geometry, sizeof(int) * n, hipMemcpyHostToDevice, c.stream));
    StateArgs s{n, rect, circ, point, ds, ds + n, ds + 2 * n, ds + 3 * n, ds + 4 * n, ds + 5 * n,
                ds + 6 * n, ds + 7 * n, ds + 8 * n, dg};
    FLAT_LAUNCH(k_generate, c, s, c.vertexX, c.vertexY, c.vertexZ, c.cellX, c.cellY, c.cellZ, c.f[F_ENERGY0],
                c.f[F_DENSITY], c.f[F_U]);
    c.sync();
    (void)hipFree(ds);
    (void)hipFree(dg);
}

// ext_core.cpp:249-256 (energy1 = energy0 over the full domain)
void op_set_field(Chunk& c)
{
    TL_HIP(hipMemcpyAsync(c.f[F_ENERGY1], c.f[F_ENERGY0], sizeof(double) * c.g.alloc, hipMemcpyDeviceToDevice,
                          c.stream));
}

void op_field_summary(Chunk& c, double* vol, double* mass, double* ie, double* temp)
{
    TIMED(c, K_SUMMARY, STRIP_LAUNCH(k_field_summary, c, c.f[F_VOLUME], c.f[F_DENSITY], c.f[F_ENERGY0], c.f[F_U],
                                     c.reducer(S_SUMMARY, EP_STORE)));
    c.close(c.g.nblocks, 4);
    c.sync();
    *vol = c.h_result[4 * S_SUMMARY + 0];
    *mass = c.h_result[4 * S_SUMMARY + 1];
    *ie = c.h_result[4 * S_SUMMARY + 2];
    *temp = c.h_result[4 * S_SUMMARY + 3];
}

// ---------------------------------------------------------------------------
// Halo messages between chunks (ext_pack_message_, ext_core.cpp:151-244 2D,
// :168-270 3D; index order of ext_pack_kernel.hpp).  Buffer element b of a
// face's message maps to one cell: the strip `depth` deep along the face,
// inside the mesh to pack, in the halo to unpack.  The order is the
// reference's: across the strip fastest for left/right, along the face
// fastest for the others.  In 2D the reference's strips start at row 0
// (left/right) and column 0 (top/bottom), not at the first mesh row or
// column; that offset is kept, so that a message means the same to both.

struct PackMap {
    int face, depth, pack;
    int lead;  // first column (left/right), row (top/bottom) or plane (back/front) of the strip
};

__device__ __forceinline__ long pack_cell(const Geom& g, const PackMap& m, long b)
{
#if TL_DIM == 2
    if (m.face == CHUNK_LEFT || m.face == CHUNK_RIGHT)
        return gidx(g, m.lead + (int)(b % m.depth), (int)(b / m.depth), 0);
    return gidx(g, (int)(b % g.nx), m.lead + (int)(b / g.nx), 0);
#else
    if (m.face == CHUNK_LEFT || m.face == CHUNK_RIGHT) {
        const long line = b / m.depth;
        return gidx(g, m.lead + (int)(b % m.depth), HALO + (int)(line % g.ny), HALO + (int)(line / g.ny));
    }
    const long line = b / g.nx;
    if (m.face == CHUNK_BOTTOM || m.face == CHUNK_TOP)
        return gidx(g, HALO + (int)(b % g.nx), m.lead + (int)(line % m.depth), HALO + (int)(line / m.depth));
    return gidx(g, HALO + (int)(b % g.nx), HALO + (int)(line % g.ny), m.lead + (int)(line / g.ny));
#endif
}

__global__ void k_pack(Geom g, PackMap m, double* __restrict__ field, double* __restrict__ buf, long n)
{
    for (long b = (long)blockIdx.x * blockDim.x + threadIdx.x; b < n; b += (long)gridDim.x * blockDim.x) {
        const long i = pack_cell(g, m, b);
        if (m.pack) buf[b] = field[i];
        else field[i] = buf[b];
    }
}

void op_pack_message(Chunk& c, const int* fields, const int* offsets, int depth, int face, double* buffer, bool pack)
{
    const Geom& g = c.g;
    if (depth < 1 || depth > HALO) throw Error("halo message depth must be 1 or 2");
    c.internal_faces = true;
    PackMap m{face, depth, pack ? 1 : 0, 0};
    long n = 0;
    const bool lo = face == CHUNK_LEFT || face == CHUNK_BOTTOM || face == CHUNK_BACK;
    const int ext = face == CHUNK_LEFT || face == CHUNK_RIGHT ? g.X : face == CHUNK_BOTTOM || face == CHUNK_TOP ? g.Y : g.Z;
    switch (face) {
    case CHUNK_LEFT:
    case CHUNK_RIGHT: n = (long)g.ny * g.nz * depth; break;
    case CHUNK_BOTTOM:
    case CHUNK_TOP: n = (long)g.nx * g.nz * depth; break;
    case CHUNK_BACK:
    case CHUNK_FRONT:
        if (TL_DIM == 3) {
            n = (long)g.nx * g.ny * depth;
            break;
        }
        [[fallthrough]];
    default: throw Error("halo message: incorrect face " + std::to_string(face));
    }
    // the mesh strip next to the face (pack) or the halo strip beyond it (unpack)
    m.lead = pack ? (lo ? HALO : ext - HALO - depth) : (lo ? HALO - depth : ext - HALO);
    static const int field_of[NUM_FIELDS] = {F_DENSITY, F_ENERGY0, F_ENERGY1, F_U, F_P, F_SD};
    if (n > c.pack_cap) {
        c.sync();
        (void)hipFree(c.d_pack);
        c.d_pack = nullptr;
        TL_HIP(hipMalloc(&c.d_pack, sizeof(double) * n));
        c.pack_cap = n;
    }
    for (int q = 0; q < NUM_FIELDS; ++q) {
        if (!fields[q]) continue;
        double* host = buffer + offsets[q];
        if (!pack) TL_HIP(hipMemcpyAsync(c.d_pack, host, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        hipLaunchKernelGGL(k_pack, dim3((unsigned)std::min<long>((n + 255) / 256, 4096)), dim3(256), 0, c.stream, g, m,
                           c.f[field_of[q]], c.d_pack, n);
        check_launch();
        if (pack) TL_HIP(hipMemcpyAsync(host, c.d_pack, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    }
}

void op_update_halo(Chunk& c, const int* neighbours, const int* fields, int depth)
{
    if (depth < 1 || depth > HALO) throw Error("halo depth must be 1 or 2");
    int mask = 0;
    for (int f = 0; f < NUM_FACES; ++f) {
        if (neighbours[f] != EXTERNAL_FACE) {
            c.internal_faces = true;
            continue;
        }
        switch (f + 1) {
        case CHUNK_LEFT: mask |= 1; break;
        case CHUNK_RIGHT: mask |= 2; break;
        case CHUNK_BOTTOM: mask |= 4; break;
        case CHUNK_TOP: mask |= 8; break;
        case CHUNK_BACK: mask |= 16; break;
        case CHUNK_FRONT: mask |= 32; break;
        }
    }
    HaloFields hf{};
    static const int field_of[NUM_FIELDS] = {F_DENSITY, F_ENERGY0, F_ENERGY1, F_U, F_P, F_SD};
    for (int q = 0; q < NUM_FIELDS; ++q)
        if (fields[q]) hf.f[hf.count++] = c.f[field_of[q]];
    if (hf.count == 0 || mask == 0) return;
    const Geom& g = c.g;
#if TL_DIM == 3
    const long n = 4L * g.ny * g.nz + 4L * g.nx * g.nz + 4L * g.nx * g.ny;
#else
    const long n = 4L * g.X + 4L * g.ny;
#endif
    TIMED(c, K_HALO, hipLaunchKernelGGL(k_update_halo, dim3((unsigned)((n + FLAT_BLOCK - 1) / FLAT_BLOCK)),
                                        dim3(FLAT_BLOCK), 0, c.stream, g, hf, depth, mask));
    check_launch();
}

// ext_core.cpp:377-413
double op_cg_init(Chunk& c, int coefficient, bool preconditioner, double dt)
{
    if (coefficient < CONDUCTIVITY && coefficient < RECIP_CONDUCTIVITY)
        throw Error("Coefficient " + std::to_string(coefficient) + " is not valid.");
    c.preconditioner = preconditioner;
    const double rx = dt / (c.dx * c.dx);
    const double ry = dt / (c.dy * c.dy);
    const double rz = TL_DIM == 3 ? dt / (c.dz * c.dz) : 0.0;
    TIMED(c, K_CG_INIT, {
        FLAT_LAUNCH(k_cg_init_k, c, coefficient, rx, ry, rz, c.f[F_DENSITY], c.f[F_ENERGY1], c.f[F_U], c.f[F_P],
                    c.f[F_R], c.f[F_W], c.f[F_KX], c.f[F_KY], c.f[F_KZ]);
        STRIP_LAUNCH(k_cg_calc_rro, c, kptr(c), (int)preconditioner, c.f[F_U], c.f[F_W], c.f[F_R], c.f[F_P],
                     c.f[F_MI], c.f[F_Z], c.reducer(S_RRO, EP_RRO));
    });
    c.close(c.g.nblocks, 1);
    return c.result(S_RRO);
}

// ext_core.cpp:417-429
double op_cg_calc_w(Chunk& c)
{
    TIMED(c, K_CG_W, STRIP_LAUNCH(k_cg_calc_w<false>, c, kptr(c), c.f[F_P], nullptr, nullptr, nullptr, c.f[F_W],
                                  c.reducer(S_PW, EP_STORE), nullptr));
    c.close(c.g.nblocks, 1);
    return c.result(S_PW);
}

// ext_core.cpp:433-447
double op_cg_calc_ur(Chunk& c, double alpha)
{
    TIMED(c, K_CG_UR, STRIP_LAUNCH(k_cg_calc_ur<true>, c, alpha, nullptr, (int)c.preconditioner, c.f[F_U], c.f[F_R],
                                   c.f[F_P], c.f[F_W], c.f[F_MI], c.f[F_Z], c.reducer(S_RRN, EP_STORE)));
    c.close(c.g.nblocks, 1);
    return c.result(S_RRN);
}

// ext_core.cpp:451-464
void op_cg_calc_p(Chunk& c, double beta)
{
    TIMED(c, K_CG_P, STRIP_LAUNCH(k_cg_calc_p, c, beta, nullptr, c.f[F_P], c.preconditioner ? c.f[F_Z] : c.f[F_R]));
}

void op_cg_calc_w_dev(Chunk& c, bool fold, bool fold_u)
{
    const double* q = c.preconditioner ? c.f[F_Z] : c.f[F_R];
    if (fold_u && !fold) throw Error("the pending u update folds only into the p, w sweep");
    if (fold) {
#if TL_DIM == 2
        const PWStrip t = pw_strip(c);
        const long nb = (t.nwaves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
        if (fold_u) {
            TIMED(c, K_CG_WPU, {
                hipLaunchKernelGGL(k_cg_pw_strip<true>, dim3((unsigned)nb), dim3(BLOCK), 0, c.stream, c.g, t, kptr(c),
                                   c.f[F_P], q, c.ds, c.p_alt, c.f[F_W], c.f[F_U], c.reducer(S_PW, EP_PW));
                check_launch();
            });
        } else {
            TIMED(c, K_CG_WP, {
                hipLaunchKernelGGL(k_cg_pw_strip<false>, dim3((unsigned)nb), dim3(BLOCK), 0, c.stream, c.g, t,
                                   kptr(c), c.f[F_P], q, c.ds, c.p_alt, c.f[F_W], nullptr, c.reducer(S_PW, EP_PW));
                check_launch();
            });
        }
        std::swap(c.f[F_P], c.p_alt);
        c.close(nb, 1);
        return;
#else
        if (fold_u) {
            TIMED(c, K_CG_WPU, STRIP_LAUNCH(k_cg_pw_z<true>, c, kptr(c), c.f[F_P], q, c.ds, c.p_alt, c.f[F_W],
                                            c.f[F_U], c.reducer(S_PW, EP_PW)));
        } else {
            TIMED(c, K_CG_WP, STRIP_LAUNCH(k_cg_pw_z<false>, c, kptr(c), c.f[F_P], q, c.ds, c.p_alt, c.f[F_W],
                                          nullptr, c.reducer(S_PW, EP_PW)));
        }
        std::swap(c.f[F_P], c.p_alt);
#endif
    } else {
        TIMED(c, K_CG_W, STRIP_LAUNCH(k_cg_calc_w<false>, c, kptr(c), c.f[F_P], nullptr, nullptr, nullptr, c.f[F_W],
                                      c.reducer(S_PW, EP_PW), nullptr));
    }
    c.close(c.g.nblocks, 1);
}

void op_cg_calc_ur_dev(Chunk& c, int n)
{
    if (n > 0) c.ensure_history(n);
    TIMED(c, K_CG_UR, STRIP_LAUNCH(k_cg_calc_ur<true>, c, 0.0, c.ds, (int)c.preconditioner, c.f[F_U], c.f[F_R], c.f[F_P],
                                   c.f[F_W], c.f[F_MI], c.f[F_Z], c.reducer(S_RRN, EP_RRN, n > 0 ? n : -1)));
    c.close(c.g.nblocks, 1);
}

void op_cg_calc_r_dev(Chunk& c, int n, Slot slot)
{
    if (n > 0) c.ensure_history(n);
    TIMED(c, K_CG_R, STRIP_LAUNCH(k_cg_calc_ur<false>, c, 0.0, c.ds, (int)c.preconditioner, c.f[F_U], c.f[F_R],
                                  c.f[F_P], c.f[F_W], c.f[F_MI], c.f[F_Z], c.reducer(slot, EP_RRN, n > 0 ? n : -1)));
    c.close(c.g.nblocks, 1);
}

void op_cg_calc_u_dev(Chunk& c)
{
    TIMED(c, K_CG_U, STRIP_LAUNCH(k_cg_calc_u, c, c.ds, c.f[F_U], c.f[F_P]));
}

void op_ppcg_calc_ur_dev(Chunk& c)
{
    TIMED(c, K_CG_UR, STRIP_LAUNCH(k_cg_calc_ur<true>, c, 0.0, c.ds, (int)c.preconditioner, c.f[F_U], c.f[F_R], c.f[F_P],
                                   c.f[F_W], c.f[F_MI], c.f[F_Z], c.reducer()));
}

void op_cg_calc_p_dev(Chunk& c)
{
    TIMED(c, K_CG_P, STRIP_LAUNCH(k_cg_calc_p, c, 0.0, c.ds, c.f[F_P], c.preconditioner ? c.f[F_Z] : c.f[F_R]));
}

void op_capture_p_faces(Chunk& c, bool pending)
{
    const FaceOffsets o = face_offsets(c.g);
    if (!c.p_faces) TL_HIP(hipMalloc(&c.p_faces, sizeof(double) * (size_t)o.count));
    const unsigned nb = (unsigned)((o.count + FLAT_BLOCK - 1) / FLAT_BLOCK);
    hipLaunchKernelGGL(k_capture_faces, dim3(nb), dim3(FLAT_BLOCK), 0, c.stream, c.g, o, (int)pending, c.ds,
                       c.f[F_P], c.preconditioner ? c.f[F_Z] : c.f[F_R], c.p_faces);
    check_launch();
}

void op_cg_calc_w_faces_dev(Chunk& c)
{
    if (!c.p_faces) throw Error("no boundary faces of p were captured");
    TIMED(c, K_CG_W, STRIP_LAUNCH(k_cg_calc_w_faces, c, face_offsets(c.g), kptr(c), c.f[F_P], c.p_faces, c.f[F_W],
                                  c.reducer(S_PW, EP_PW)));
    c.close(c.g.nblocks, 1);
}

void op_2norm_rrn_dev(Chunk& c)
{
    TIMED(c, K_2NORM, STRIP_LAUNCH(k_2norm, c, c.f[F_R], c.reducer(S_RRN, EP_RRN)));
    c.close(c.g.nblocks, 1);
}

double op_error(Chunk& c) { return c.result(S_RRN); }

void op_finish_sums3(Chunk& c, long nblocks, double* out)
{
    finish<3>(c, nblocks, S_NORM, EP_STORE);
    for (int k = 0; k < 3; ++k) out[k] = c.result(S_NORM, k);
}

void op_2norm_replace_dev(Chunk& c)
{
    TIMED(c, K_2NORM, STRIP_LAUNCH(k_2norm, c, c.f[F_R], c.reducer(S_RRN, EP_RRN_REPLACE)));
    c.close(c.g.nblocks, 1);
}

void reduce_probe(long nblocks, int nwaves, const double* x, double* out)
{
    if (nblocks < 1 || nblocks > (1L << 22)) throw Error("reduce probe: 1 to 2^22 workgroups");
    if (nwaves != 1 && nwaves != WAVES_PER_BLOCK) throw Error("reduce probe: workgroups of 1 or 4 waves");
    hipStream_t s = nullptr;
    double *dx = nullptr, *part = nullptr, *slots = nullptr, *res = nullptr;
    unsigned *counter = nullptr, *seq = nullptr;
    const long ncount = reduce_counters(nblocks);
    auto release = [&] {
        (void)hipFree(dx);
        (void)hipFree(part);
        (void)hipFree(slots);
        (void)hipFree(res);
        (void)hipFree(counter);
        (void)hipFree(seq);
    };
    try {
        TL_HIP(hipMalloc(&dx, sizeof(double) * nblocks));
        TL_HIP(hipMalloc(&part, sizeof(double) * nblocks));
        TL_HIP(hipMalloc(&slots, sizeof(double) * nblocks));
        TL_HIP(hipMalloc(&res, sizeof(double) * 12));
        TL_HIP(hipMalloc(&counter, sizeof(unsigned) * ncount));
        TL_HIP(hipMalloc(&seq, sizeof(unsigned)));
        TL_HIP(hipMemcpy(dx, x, sizeof(double) * nblocks, hipMemcpyHostToDevice));
        TL_HIP(hipMemset(slots, 0xff, sizeof(double) * nblocks));  // SLOT_EMPTY
        TL_HIP(hipMemset(counter, 0, sizeof(unsigned) * ncount));
        TL_HIP(hipMemset(res, 0, sizeof(double) * 12));
        // three reductions in a row: single pass twice (counters and slots
        // must be back at their empty state after the first), then the
        // separate k_finish launch
        for (int k = 0; k < 3; ++k) {
            Reduce red{part, Finish{k < 2 ? counter : nullptr, slots, nullptr, res + 4 * k, res + 4 * k, nullptr,
                                    nullptr, seq, EP_STORE, -1, (unsigned)k + 1,
                                    (unsigned)((nblocks + REDUCE_GROUP - 1) / REDUCE_GROUP)}};
            if (nwaves == 1)
                hipLaunchKernelGGL(k_reduce_probe<1>, dim3((unsigned)nblocks), dim3(WAVE), 0, s, dx, red);
            else
                hipLaunchKernelGGL(k_reduce_probe<WAVES_PER_BLOCK>, dim3((unsigned)nblocks), dim3(BLOCK), 0, s, dx,
                                   red);
            check_launch();
            if (k == 2) {
                hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(BLOCK), 0, s, part, nblocks, red.fin);
                check_launch();
            }
        }
        double h[12];
        TL_HIP(hipMemcpy(h, res, sizeof h, hipMemcpyDeviceToHost));
        std::vector<unsigned> left((size_t)ncount, 1u);
        TL_HIP(hipMemcpy(left.data(), counter, sizeof(unsigned) * left.size(), hipMemcpyDeviceToHost));
        for (unsigned v : left)
            if (v != 0) throw Error("reduce probe: a counter was not reset");
        std::vector<unsigned long long> empty((size_t)nblocks, 0ull);
        TL_HIP(hipMemcpy(empty.data(), slots, sizeof(double) * nblocks, hipMemcpyDeviceToHost));
        for (unsigned long long v : empty)
            if (v != SLOT_EMPTY) throw Error("reduce probe: a partial slot was not left empty");
        out[0] = h[0];
        out[1] = h[4];
        out[2] = h[8];
    } catch (...) {
        release();
        throw;
    }
    release();
}

double stream_copy_probe(long n, int reps)
{
    if (n < 2 || reps < 1) throw Error("copy probe needs n >= 2 and reps >= 1");
    n &= ~1L;
    double *a = nullptr, *b = nullptr;
    TL_HIP(hipMalloc(&a, sizeof(double) * n));
    TL_HIP(hipMalloc(&b, sizeof(double) * n));
    TL_HIP(hipMemset(a, 0, sizeof(double) * n));
    hipEvent_t e0, e1;
    TL_HIP(hipEventCreate(&e0));
    TL_HIP(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_stream_copy, dim3(COPY_GRID), dim3(COPY_BLOCK), 0, 0, (const double2*)a, (double2*)b, n / 2);
    TL_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_stream_copy, dim3(COPY_GRID), dim3(COPY_BLOCK), 0, 0, (const double2*)a, (double2*)b, n / 2);
    TL_HIP(hipEventRecord(e1, 0));
    TL_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    TL_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return 16.0 * (double)n * reps / (ms * 1e-3) / 1e9;  // GB/s, read + write
}

double stream_probe(int kind, long n, int reps)
{
    if (kind == 0) return stream_copy_probe(n, reps);
    if (kind < 0 || kind > 6) throw Error("stream probe kinds are 0-6");
    if (n < 2 || reps < 1) throw Error("stream probe needs n >= 2 and reps >= 1");
    n &= ~1L;
    const long n2 = n / 2;
    const int nbuf = kind >= 5 ? 8 : 2;
    std::vector<double*> buf(nbuf, nullptr);
    for (double*& q : buf) {
        TL_HIP(hipMalloc(&q, sizeof(double) * n));
        TL_HIP(hipMemset(q, 0, sizeof(double) * n));
    }
    ProbeStreams st{};
    for (int k = 0; k < 6 && k < nbuf; ++k) st.r[k] = reinterpret_cast<const d2v*>(buf[k]);
    if (kind >= 5) {
        st.w[0] = reinterpret_cast<d2v*>(buf[6]);
        st.w[1] = reinterpret_cast<d2v*>(buf[7]);
    }
    const d2v* a = reinterpret_cast<const d2v*>(buf[0]);
    d2v* b = reinterpret_cast<d2v*>(buf[1]);
    const unsigned grid = (unsigned)((n2 + 255) / 256);
    double per = 16.0;  // bytes per element and launch
    auto launch = [&] {
        switch (kind) {
        case 1: hipLaunchKernelGGL(k_probe_copy<false>, dim3(grid), dim3(256), 0, 0, a, b, n2); break;
        case 2: hipLaunchKernelGGL(k_probe_copy<true>, dim3(grid), dim3(256), 0, 0, a, b, n2); break;
        case 3: hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(256), 0, 0, a, n2, buf[1]); break;
        case 4: hipLaunchKernelGGL(k_probe_write, dim3(grid), dim3(256), 0, 0, b, n2); break;
        case 5: hipLaunchKernelGGL(k_probe_mix<4>, dim3(grid), dim3(256), 0, 0, st, n2); break;
        default: hipLaunchKernelGGL(k_probe_mix<6>, dim3(grid), dim3(256), 0, 0, st, n2); break;
        }
    };
    if (kind == 1 || kind == 2) per = 32.0;
    if (kind == 5) per = 96.0;
    if (kind == 6) per = 128.0;
    hipEvent_t e0, e1;
    TL_HIP(hipEventCreate(&e0));
    TL_HIP(hipEventCreate(&e1));
    launch();
    check_launch();
    TL_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r) launch();
    TL_HIP(hipEventRecord(e1, 0));
    TL_HIP(hipEventSynchronize(e1));
    check_launch();
    float ms = 0.0f;
    TL_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (double* q : buf) (void)hipFree(q);
    return per * (double)n2 * reps / (ms * 1e-3) / 1e9;
}

void op_read_history(Chunk& c, int n, std::vector<double>& a, std::vector<double>& b)
{
    a.assign(n + 2, 0.0);
    b.assign(n + 2, 0.0);
    if (n <= 0) return;
    if (n > c.hist_cap) throw Error("CG history shorter than requested");
    TL_HIP(hipMemcpyAsync(a.data() + 1, c.hist_a + 1, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    TL_HIP(hipMemcpyAsync(b.data() + 1, c.hist_b + 1, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
}

static void store_coefs(Chunk& c, const double* alphas, const double* betas, int n)
{
    if (n < 0) throw Error("negative coefficient count");
    c.alphas.assign(alphas, alphas + n);
    c.betas.assign(betas, betas + n);
}

// Device copy of the stored Chebyshev coefficients for the temporal-blocked
// iterations: a[0..n) then b[0..n).
void op_cheby_upload(Chunk& c)
{
    const int n = (int)c.alphas.size();
    if (n == 0) throw Error("no Chebyshev coefficients stored");
    if (2 * n > c.d_cheby_cap) {
        c.sync();
        (void)hipFree(c.d_cheby);
        c.d_cheby = nullptr;
        TL_HIP(hipMalloc(&c.d_cheby, sizeof(double) * 2 * (size_t)n));
        c.d_cheby_cap = 2 * n;
    }
    std::vector<double> h(c.alphas);
    h.insert(h.end(), c.betas.begin(), c.betas.end());
    c.sync();
    TL_HIP(hipMemcpy(c.d_cheby, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    c.d_cheby_n = n;
}

// ext_core.cpp:468-497
void op_cheby_init(Chunk& c, const double* alphas, const double* betas, int n, double theta, bool preconditioner,
                   bool fused)
{
    c.preconditioner = preconditioner;
    store_coefs(c, alphas, betas, n);
    const int pre = preconditioner;
    if (fused && TL_DIM == 2 && n > 0) op_cheby_upload(c);
    if (fused) {
        STRIP_LAUNCH((k_cheby<true, true>), c, kptr(c), pre, theta, 0.0, 0.0, c.f[F_U], c.f[F_U0], c.f[F_MI],
                     c.f[F_W], c.f[F_R], c.f[F_P], c.u_alt);
        std::swap(c.f[F_U], c.u_alt);
    } else {
        STRIP_LAUNCH((k_cheby<true, false>), c, kptr(c), pre, theta, 0.0, 0.0, c.f[F_U], c.f[F_U0], c.f[F_MI],
                     c.f[F_W], c.f[F_R], c.f[F_P], c.u_alt);
        STRIP_LAUNCH(k_add_p, c, c.f[F_U], c.f[F_P]);
    }
}

// ext_core.cpp:501-515
void op_cheby_iterate(Chunk& c, int step1, bool fused)
{
    const int s = step1 - 1;
    if (s < 0 || s >= (int)c.alphas.size())
        throw Error("Chebyshev step " + std::to_string(step1) + " outside the stored coefficients");
    const double a = c.alphas[s], b = c.betas[s];
    const int pre = c.preconditioner;
    if (fused) {
#if TL_DIM == 3
        TIMED(c, K_CHEBY, STRIP_LAUNCH(k_cheby_z, c, kptr(c), pre, a, b, c.f[F_U], c.f[F_U0], c.f[F_MI], c.f[F_R],
                                       c.f[F_P], c.u_alt));
#else
        TIMED(c, K_CHEBY, STRIP_LAUNCH((k_cheby<false, true>), c, kptr(c), pre, 0.0, a, b, c.f[F_U], c.f[F_U0],
                                       c.f[F_MI], c.f[F_W], c.f[F_R], c.f[F_P], c.u_alt));
#endif
        std::swap(c.f[F_U], c.u_alt);
    } else {
        STRIP_LAUNCH((k_cheby<false, false>), c, kptr(c), pre, 0.0, a, b, c.f[F_U], c.f[F_U0], c.f[F_MI], c.f[F_W],
                     c.f[F_R], c.f[F_P], c.u_alt);
        STRIP_LAUNCH(k_add_p, c, c.f[F_U], c.f[F_P]);
    }
}

// ext_core.cpp:519-539
void op_ppcg_init(Chunk& c, const double* alphas, const double* betas, int n, double theta)
{
    c.theta = theta;
    store_coefs(c, alphas, betas, n);
    c.d_coef_steps = 0;  // the temporal-blocked passes lay their coefficients out again
}

// ext_core.cpp:543-550
void op_ppcg_init_sd(Chunk& c)
{
    TIMED(c, K_PPCG_SD, STRIP_LAUNCH(k_ppcg_init_sd, c, (int)c.preconditioner, c.theta, c.f[F_R], c.f[F_MI],
                                     c.f[F_SD]));
}

// ext_core.cpp:554-567
void op_ppcg_inner(Chunk& c, int step1, bool fused)
{
    const int s = step1 - 1;
    if (s < 0 || s >= (int)c.alphas.size())
        throw Error("PPCG step " + std::to_string(step1) + " outside the stored coefficients");
    const double a = c.alphas[s], b = c.betas[s];
    const int pre = c.preconditioner;
    if (fused) {
#if TL_DIM == 3
        TIMED(c, K_PPCG_INNER, STRIP_LAUNCH(k_ppcg_inner_z, c, kptr(c), pre, a, b, c.f[F_SD], c.f[F_R], c.f[F_U],
                                            c.f[F_MI], c.sd_alt));
#else
        TIMED(c, K_PPCG_INNER, STRIP_LAUNCH(k_ppcg_inner_fused, c, kptr(c), pre, a, b, c.f[F_SD], c.f[F_R],
                                            c.f[F_U], c.f[F_MI], c.sd_alt));
#endif
        std::swap(c.f[F_SD], c.sd_alt);
    } else {
        STRIP_LAUNCH(k_ppcg_calc_u, c, kptr(c), c.f[F_SD], c.f[F_R], c.f[F_U]);
        STRIP_LAUNCH(k_ppcg_calc_sd, c, pre, a, b, c.f[F_R], c.f[F_MI], c.f[F_SD]);
    }
}

void op_ppcg_inner2(Chunk& c, int step1)
{
#if TL_DIM == 3
    const int s = step1 - 1;
    if (s < 0 || s + 1 >= (int)c.alphas.size())
        throw Error("PPCG steps " + std::to_string(step1) + ", " + std::to_string(step1 + 1) +
                    " outside the stored coefficients");
    PZGeom pz;
    pz.tiles_x = (c.g.nx + PZ_TX - 1) / PZ_TX;
    pz.tiles_y = (c.g.ny + PZ_TY - 1) / PZ_TY;
    // strip length: fewest rounds of resident workgroups (one per CU) times
    // the steps each walks (its planes + 4)
    const long cus = std::max(1, c.simds / 4), cols = (long)pz.tiles_x * pz.tiles_y;
    long best = -1;
    pz.strip = c.g.nz;
    for (int S : {64, 48, 32, 24, 16, 12, 8}) {
        const int len = std::min(S, c.g.nz);
        const long blocks = cols * ((c.g.nz + len - 1) / len);
        const long cost = ((blocks + cus - 1) / cus) * (len + 4);
        if (best < 0 || cost < best) {
            best = cost;
            pz.strip = len;
        }
    }
    pz.nblocks = cols * ((c.g.nz + pz.strip - 1) / pz.strip);
    const int pre = c.preconditioner;
    TIMED(c, K_PPCG_INNER2, {
        hipLaunchKernelGGL(k_ppcg_inner2_z, dim3((unsigned)pz.nblocks), dim3(WAVE * PZ_WAVES), 0, c.stream, c.g, pz,
                           kptr(c), pre, c.alphas[s], c.betas[s], c.alphas[s + 1], c.betas[s + 1], c.f[F_SD],
                           c.f[F_R], c.f[F_U], c.f[F_MI], c.sd_alt, c.r_alt);
        check_launch();
    });
    std::swap(c.f[F_SD], c.sd_alt);
    std::swap(c.f[F_R], c.r_alt);
#else
    (void)c;
    (void)step1;
    throw Error("the two-step PPCG sweep is 3D only");
#endif
}

// ext_core.cpp:340-358
void op_jacobi_init(Chunk& c, int coefficient, double dt)
{
    if (coefficient < CONDUCTIVITY && coefficient < RECIP_CONDUCTIVITY)
        throw Error("Coefficient " + std::to_string(coefficient) + " is not valid.");
    const double rx = dt / (c.dx * c.dx);
    const double ry = dt / (c.dy * c.dy);
    const double rz = TL_DIM == 3 ? dt / (c.dz * c.dz) : 0.0;
    FLAT_LAUNCH(k_jacobi_init, c, coefficient, rx, ry, rz, c.f[F_DENSITY], c.f[F_ENERGY1], c.f[F_U0], c.f[F_U],
                c.f[F_KX], c.f[F_KY], c.f[F_KZ]);
}

// ext_core.cpp:362-373.  Plain form: r = u over the whole allocation, then
// u from r in place.  Fused form: the new iterate goes into r's buffer and
// the two swap, leaving r holding the previous iterate as the reference's
// copy does.
double op_jacobi_solve(Chunk& c, bool fused)
{
    if (fused) {
        TIMED(c, K_JACOBI, STRIP_LAUNCH(k_jacobi_solve, c, kptr(c), c.f[F_U0], c.f[F_U], c.f[F_R], c.reducer(S_ERR, EP_STORE)));
        std::swap(c.f[F_U], c.f[F_R]);
    } else {
        TL_HIP(hipMemcpyAsync(c.f[F_R], c.f[F_U], sizeof(double) * c.g.alloc, hipMemcpyDeviceToDevice, c.stream));
        STRIP_LAUNCH(k_jacobi_solve, c, kptr(c), c.f[F_U0], c.f[F_R], c.f[F_U], c.reducer(S_ERR, EP_STORE));
    }
    c.close(c.g.nblocks, 1);
    return c.result(S_ERR);
}

void op_copy_u(Chunk& c) { TIMED(c, K_COPY_U, STRIP_LAUNCH(k_copy_u, c, c.f[F_U], c.f[F_U0])); }

void op_calc_residual(Chunk& c)
{
    TIMED(c, K_RESIDUAL, STRIP_LAUNCH(k_residual, c, kptr(c), c.f[F_U], c.f[F_U0], c.f[F_R]));
}

double op_dot(Chunk& c, const double* a, const double* b)
{
    STRIP_LAUNCH(k_dot, c, a, b, c.reducer(S_NORM, EP_STORE));
    c.close(c.g.nblocks, 1);
    return c.result(S_NORM);
}

void op_fcg_update_p(Chunk& c, double beta) { STRIP_LAUNCH(k_fcg_p, c, beta, nullptr, 0, c.f[F_Z], c.f[F_P]); }

// The flexible CG's scalars on the device: one host read per iteration.
// beta_rrn: the preconditioner pass left z . w and the previous iteration's
// ||r||^2 (it applied that iteration's r update), read from S_ERR.  defer_r:
// leave this iteration's r -= alpha w to the next preconditioner pass (2D);
// no rrn reaches S_RRN then.
void op_fcg_iteration_dev(Chunk& c, bool first, long beta_blocks, bool& u_pending, bool beta_rrn, bool defer_r,
                          bool alpha_zr)
{
    if (beta_rrn) {
        if (first || beta_blocks <= 0) throw Error("flexible CG: no preconditioner partials for beta and rrn");
        finish<3>(c, beta_blocks, S_ERR, EP_FCG_BETA_RRN);
    } else if (!first && beta_blocks > 0) {
        finish<1>(c, beta_blocks, S_ERR, EP_FCG_BETA);
    } else if (!first) {  // beta = -(z . w) / pw_prev, w still A p_prev
        STRIP_LAUNCH(k_dot, c, c.f[F_Z], c.f[F_W], c.reducer(S_ERR, EP_FCG_BETA));
        c.close(c.g.nblocks, 1);
    }
#if TL_DIM == 2
    if (!first) {
        // p = z + beta p, w = A p, p.w and p.r in one streaming sweep
        const PWStrip t = pw_strip(c);
        const long nb = (t.nwaves + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
        // with u_pending the previous iteration's u += alpha p rides along
        // (alpha and the old p are still in place).  With beta_rrn the
        // preconditioner pass left z . r, alpha's numerator for the fixed
        // (symmetric) sweep preconditioner, and the strip reads no r (unless
        // alpha_zr is off: after a residual replacement p_prev . r = 0, on
        // which z . r = p . r rests, no longer holds).
        if (beta_rrn && alpha_zr) {
            TIMED(c, K_FCG_STRIP_ZR, {
                if (u_pending)
                    hipLaunchKernelGGL((k_cg_pw_strip<true, false>), dim3((unsigned)nb), dim3(BLOCK), 0, c.stream,
                                       c.g, t, kptr(c), c.f[F_P], c.f[F_Z], c.ds, c.p_alt, c.f[F_W], c.f[F_U],
                                       c.reducer(S_PW, EP_FCG_PW_ZR));
                else
                    hipLaunchKernelGGL((k_cg_pw_strip<false, false>), dim3((unsigned)nb), dim3(BLOCK), 0, c.stream,
                                       c.g, t, kptr(c), c.f[F_P], c.f[F_Z], c.ds, c.p_alt, c.f[F_W], nullptr,
                                       c.reducer(S_PW, EP_FCG_PW_ZR));
                check_launch();
            });
            u_pending = false;
            std::swap(c.f[F_P], c.p_alt);
            c.close(nb, 1);
        } else {
        TIMED(c, K_FCG_STRIP, {
            if (u_pending)
                hipLaunchKernelGGL((k_cg_pw_strip<true, true>), dim3((unsigned)nb), dim3(BLOCK), 0, c.stream, c.g, t,
                                   kptr(c), c.f[F_P], c.f[F_Z], c.ds, c.p_alt, c.f[F_W], c.f[F_U], c.reducer(S_PW, EP_FCG_PWPR),
                                   c.f[F_R]);
            else
                hipLaunchKernelGGL((k_cg_pw_strip<false, true>), dim3((unsigned)nb), dim3(BLOCK), 0, c.stream, c.g,
                                   t, kptr(c), c.f[F_P], c.f[F_Z], c.ds, c.p_alt, c.f[F_W], nullptr, c.reducer(S_PW, EP_FCG_PWPR),
                                   c.f[F_R]);
            check_launch();
        });
        u_pending = false;
        std::swap(c.f[F_P], c.p_alt);
        c.close(nb, 2);
        }
    } else
#endif
    {
        if (u_pending) {
            op_cg_calc_u_dev(c);
            u_pending = false;
        }
        STRIP_LAUNCH(k_fcg_p, c, 0.0, c.ds, first ? 1 : 0, c.f[F_Z], c.f[F_P]);
        TIMED(c, K_CG_W, STRIP_LAUNCH(k_cg_calc_w<false>, c, kptr(c), c.f[F_P], nullptr, nullptr, nullptr, c.f[F_W],
                                      c.reducer(S_PW, EP_FCG_PW), nullptr));
        c.close(c.g.nblocks, 1);
        STRIP_LAUNCH(k_dot, c, c.f[F_P], c.f[F_R], c.reducer(S_ERR, EP_FCG_ALPHA));
        c.close(c.g.nblocks, 1);
    }
    if (TL_DIM == 2 && defer_r) {
        u_pending = true;
        return;
    }
    if (TL_DIM == 2) {
        // r only: u += alpha p is left to the next strip (or op_cg_calc_u_dev)
        TIMED(c, K_CG_R, STRIP_LAUNCH(k_cg_calc_ur<false>, c, 0.0, c.ds, (int)c.preconditioner, c.f[F_U],
                                      c.f[F_R], c.f[F_P], c.f[F_W], c.f[F_MI], c.f[F_Z], c.reducer(S_RRN, EP_FCG_RRN)));
        u_pending = true;
    } else {
        TIMED(c, K_CG_UR, STRIP_LAUNCH(k_cg_calc_ur<true>, c, 0.0, c.ds, (int)c.preconditioner, c.f[F_U],
                                       c.f[F_R], c.f[F_P], c.f[F_W], c.f[F_MI], c.f[F_Z], c.reducer(S_RRN, EP_FCG_RRN)));
    }
    c.close(c.g.nblocks, 1);
}

double op_mean_diag(Chunk& c)
{
    STRIP_LAUNCH(k_face_sum, c, c.f[F_KX], c.f[F_KY], TL_DIM == 3 ? c.f[F_KZ] : nullptr,
                 c.reducer(S_NORM, EP_STORE));
    c.close(c.g.nblocks, 1);
    return 1.0 + 2.0 * c.result(S_NORM) / ((double)c.g.nx * c.g.ny * c.g.nz);
}

double op_calc_2norm(Chunk& c, bool norm_r)
{
    TIMED(c, K_2NORM, STRIP_LAUNCH(k_2norm, c, norm_r ? c.f[F_R] : c.f[F_U0], c.reducer(S_NORM, EP_STORE)));
    c.close(c.g.nblocks, 1);
    return c.result(S_NORM);
}

void op_finalise(Chunk& c)
{
    TIMED(c, K_FINALISE, STRIP_LAUNCH(k_finalise, c, c.f[F_U], c.f[F_DENSITY], c.f[F_ENERGY1]));
}

}  // namespace tl
