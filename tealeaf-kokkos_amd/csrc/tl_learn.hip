// Learned preconditioner for TeaLeaf's implicit solve (no counterpart in
// the reference): training pairs from the FP64 PPCG solver, training of the
// BF16 U-Net (tl_unet.hip) on the GPU, and flexible CG with the network as
// its preconditioner, with FP64 PPCG as the fallback when it stalls.
//
// Training pairs.  While the driver's FP64 PPCG solves a deck's first
// timestep, the solve hooks copy u_k and r_k = u0 - A u_k at a set of outer
// iterations k, and u* at convergence.  The PPCG correction of snapshot k is
// e_k = u* - u_k (A e_k = r_k up to the solver tolerance).  Random crop x crop
// windows of (r_k, density, e_k) become pairs: the network input is r / s
// (s = rms of r_k over the whole mesh, as the flexible CG scales its input)
// with the density feature (log10(rho) - 0.5)/1.5, the target e / (s G), G
// one global scale over the data set.  Residual and correction share the
// factor s, so the map stays linear in r as A^-1 is.  Training minimises the
// A-norm of the error over each batch, sum (z - e).A(z - e) / sum e.A e.
//
// Flexible CG (Notay's FCG with one stored direction): z = s G N(r / s),
// p = z - (z.w_prev / p_prev.w_prev) p_prev, w = A p, alpha = p.r / p.w,
// u += alpha p, r -= alpha w, all vectors and dot products in FP64.  It is
// accepted only on the true residual ||u0 - A u||^2 < eps; when the
// residual has not halved over a window of iterations the driver solves the
// timestep again with FP64 PPCG.
#include "tl_chunk.hpp"
#include "tl_unet.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

namespace tl {

struct Learner {
    UNet* net = nullptr;
    uint64_t rng = 0;
    int crop = 0;
    long pairs = 0;
    bool consistent = true;                             // input residual A_c e of the crop (else r_k)
    int phases = 3;                                     // snapshots from: 1 CG presteps, 2 PPCG outer iterations
    float *dr = nullptr, *df = nullptr, *de = nullptr;  // [pairs][crop][crop]
    float *dkx = nullptr, *dky = nullptr;               // the pair's face coefficients
    long dr_cap = 0, df_cap = 0, de_cap = 0, dkx_cap = 0, dky_cap = 0;
    double escale = 0.0;                                // G; 0 until the first training call
    float *br = nullptr, *bf = nullptr, *be = nullptr;  // one training batch
    float *bkx = nullptr, *bky = nullptr;
    long br_cap = 0, bf_cap = 0, be_cap = 0, bkx_cap = 0, bky_cap = 0;
    float *pr = nullptr, *pf = nullptr, *pz = nullptr;  // FCG planes
    long pr_cap = 0, pf_cap = 0, pz_cap = 0;
    // coarse pairs: whole coarse grids of crop x crop cells (the network is
    // then the coarse-grid solve of the two-grid preconditioner)
    bool coarse = false;
    double* crc = nullptr;  // the coarse right-hand side (FP64) and its rms
    long crc_cap = 0;
    float *ckx = nullptr, *cky = nullptr;  // FCG: the coarse operator (unused at inference) 
    long ckx_cap = 0, cky_cap = 0;
    double* scratch = nullptr;                          // per-crop sums
    long scratch_cap = 0;
};

namespace {

Learner* g_active = nullptr;
SolveHooks g_hooks;

uint64_t next(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <class T>
void regrow(T*& p, long& cap, long need, long keep)
{
    if (need <= cap) return;
    T* q = nullptr;
    TL_HIP(hipMalloc(&q, sizeof(T) * (size_t)need));
    if (p && keep > 0) TL_HIP(hipMemcpy(q, p, sizeof(T) * (size_t)keep, hipMemcpyDeviceToDevice));
    (void)hipFree(p);
    p = q;
    cap = need;
}

// One workgroup per crop: copy the window at (x0, y0) (interior cell
// coordinates) of r_k, the density feature, e = u* - u_k and the face
// coefficients kx, ky, and add the crop's sum of r^2 into ss[crop].
__global__ void k_extract(Geom g, const double* __restrict__ uk, const double* __restrict__ ustar,
                          const double* __restrict__ rk, const double* __restrict__ density,
                          const double* __restrict__ kx, const double* __restrict__ ky, const int* __restrict__ xy,
                          int crop, float* __restrict__ out_r, float* __restrict__ out_f, float* __restrict__ out_e,
                          float* __restrict__ out_kx, float* __restrict__ out_ky, double* __restrict__ ss)
{
    const int q = blockIdx.x;
    const int x0 = xy[2 * q], y0 = xy[2 * q + 1];
    const long base = (long)q * crop * crop;
    double s = 0.0;
    for (int i = threadIdx.x; i < crop * crop; i += blockDim.x) {
        const int x = i % crop, y = i / crop;
        const long gi = gidx(g, HALO + x0 + x, HALO + y0 + y, 0);
        const double r = rk[gi];
        out_r[base + i] = (float)r;
        out_f[base + i] = (float)((std::log10(density[gi]) - 0.5) / 1.5);
        out_e[base + i] = (float)(ustar[gi] - uk[gi]);
        out_kx[base + i] = (float)kx[gi];
        out_ky[base + i] = (float)ky[gi];
        s += r * r;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (threadIdx.x % WAVE == 0) atomicAdd(ss + q, s);
}

// The residual a crop's error leaves under the crop's own operator:
// r = A_c e, A_c = I + sum over the faces inside the crop of k (e_i - e_j)
// (no flux across the crop's edge; kx the west and ky the south face of each
// cell), FP64 arithmetic on the stored FP32 values.  The pair (A_c e, e) is
// then consistent: the target is exactly A_c^-1 of the input, a function of
// what the network sees, whereas the solve's own r_k inside a crop also
// carries the part of e driven by residual outside it.
__global__ void k_crop_residual(int crop, const float* __restrict__ e, const float* __restrict__ kx,
                                const float* __restrict__ ky, float* __restrict__ r)
{
    const int q = blockIdx.x;
    const long base = (long)q * crop * crop;
    const float* eq = e + base;
    const float* kxq = kx + base;
    const float* kyq = ky + base;
    for (int i = threadIdx.x; i < crop * crop; i += blockDim.x) {
        const int x = i % crop, y = i / crop;
        const double c = eq[i];
        double v = c;
        if (x + 1 < crop) v += (double)kxq[i + 1] * (c - eq[i + 1]);
        if (x > 0) v += (double)kxq[i] * (c - eq[i - 1]);
        if (y + 1 < crop) v += (double)kyq[i + crop] * (c - eq[i + crop]);
        if (y > 0) v += (double)kyq[i] * (c - eq[i - crop]);
        r[base + i] = (float)v;
    }
}

// r /= s, e /= s G over every crop of one snapshot (s the snapshot's rms residual)
__global__ void k_normalise(int crop, double inv, double inv_g, float* __restrict__ r, float* __restrict__ e)
{
    const int q = blockIdx.x;
    const long base = (long)q * crop * crop;
    for (int i = threadIdx.x; i < crop * crop; i += blockDim.x) {
        r[base + i] = (float)(r[base + i] * inv);
        e[base + i] = (float)(e[base + i] * inv * inv_g);
    }
}

// sum of a^2 over the interior of a field
__global__ void k_field_sumsq(Geom g, const double* __restrict__ a, double* __restrict__ out)
{
    const long n = (long)g.nx * g.ny;
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = a[gidx(g, HALO + (int)(i % g.nx), HALO + (int)(i / g.nx), 0)];
        s += v * v;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (threadIdx.x % WAVE == 0) atomicAdd(out, s);
}

__global__ void k_sumsq(const float* __restrict__ a, long n, double* __restrict__ out)
{
    double s = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += (double)a[i] * a[i];
    for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (threadIdx.x % WAVE == 0) atomicAdd(out, s);
}

__global__ void k_scale(float* __restrict__ a, long n, float f)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] *= f;
}

// FCG planes: r / s and the density feature over the interior (row-major ny x nx)
__global__ void k_plane_in(Geom g, const double* __restrict__ r, const double* __restrict__ density, double inv_s,
                           int with_feature, float* __restrict__ pr, float* __restrict__ pf)
{
    const long n = (long)g.nx * g.ny;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int x = (int)(i % g.nx), y = (int)(i / g.nx);
        const long gi = gidx(g, HALO + x, HALO + y, 0);
        pr[i] = (float)(r[gi] * inv_s);
        if (with_feature) pf[i] = (float)((std::log10(density[gi]) - 0.5) / 1.5);
    }
}

// z = s * network output, FP64, interior
__global__ void k_plane_out(Geom g, const float* __restrict__ pz, double s, double* __restrict__ z)
{
    const long n = (long)g.nx * g.ny;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int x = (int)(i % g.nx), y = (int)(i / g.nx);
        z[gidx(g, HALO + x, HALO + y, 0)] = s * (double)pz[i];
    }
}

// One weighted Jacobi sweep on A z = r: z_out = z + (r - A z) ((w1 / D) f),
// the sweep's weight w1 f, D the diagonal of A (1 + the coefficients of the
// cell's interior faces; a face on the mesh boundary carries no flux).
// Interior cells.
__global__ void k_fcg_smooth(Geom g, const double* __restrict__ kx, const double* __restrict__ ky,
                             const double* __restrict__ r, const double* __restrict__ z, double w1, double f,
                             double* __restrict__ zo)
{
    const long n = (long)g.nx * g.ny;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int x = (int)(i % g.nx) + HALO, y = (int)(i / g.nx) + HALO;
        const long c = gidx(g, x, y, 0);
        const double zc = z[c];
        double az = zc, d = 1.0;
        if (x < g.nx + 1) { az += kx[c + 1] * (zc - z[c + 1]); d += kx[c + 1]; }
        if (x > HALO) { az += kx[c] * (zc - z[c - 1]); d += kx[c]; }
        if (y < g.ny + 1) { az += ky[c + g.sx] * (zc - z[c + g.sx]); d += ky[c + g.sx]; }
        if (y > HALO) { az += ky[c] * (zc - z[c - g.sx]); d += ky[c]; }
        zo[c] = zc + (r[c] - az) * ((w1 / d) * f);
    }
}

unsigned blocks_for(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 16384)); }

}  // namespace

// z = `sweeps` weighted Jacobi sweeps on A z = r from z (from zero with
// from_zero): one temporal-blocked pass where it applies, else one launch
// per sweep.  The result is left in F_Z.
void fcg_sweep_weights(const FcgOptions& opt, double* w)
{
    const int m = opt.smooth;
    if (opt.cheby_ratio > 0.0) {
        const double hi = 2.0, lo = 2.0 / opt.cheby_ratio, pi = 3.14159265358979323846;
        // roots in decreasing order: the iterate never grows (every partial
        // product of (1 - x / theta) stays within 1 on [0, 2]), but a
        // rounding error of an early sweep is multiplied by the large
        // late weights (up to 5.5e7 for 16 sweeps on [2/3000, 2])
        std::vector<double> th(m);
        for (int k = 0; k < m; ++k) th[k] = 0.5 * (hi + lo) + 0.5 * (hi - lo) * std::cos((2 * k + 1) * pi / (2 * m));
        // for m a power of two, the Lebedev-Finogenov order (root i paired
        // with root n + 1 - i at each doubling) keeps both the partial
        // products and those rounding factors small (90 and 353 for 16)
        std::vector<int> idx(1, 1);
        if ((m & (m - 1)) == 0) {
            while ((int)idx.size() < m) {
                const int n2 = 2 * (int)idx.size();
                std::vector<int> next;
                for (int v : idx) {
                    next.push_back(v);
                    next.push_back(n2 + 1 - v);
                }
                idx.swap(next);
            }
        } else {
            idx.resize(m);
            for (int k = 0; k < m; ++k) idx[k] = k + 1;
        }
        for (int k = 0; k < m; ++k) w[k] = 1.0 / th[idx[k] - 1];
    } else {
        for (int k = 0; k < m; ++k) w[k] = opt.omega;
    }
}

double fcg_auto_ratio(Chunk& c, int sweeps)
{
    const double m = sweeps / 16.0;
    const double a = 30.0 * m * m * std::sqrt(op_mean_diag(c));
    return std::min(30000.0, std::max(30.0, a));
}

long fcg_smooth(Chunk& c, int sweeps, const double* weights, bool from_zero, bool allow_fused, bool dot_w)
{
    // the sweep kernels are the 5-point operator on one plane
    if (TL_DIM != 2) throw Error("the flexible CG's preconditioner sweeps are 2D only");
    if (sweeps < 1) {
        if (from_zero) TL_HIP(hipMemsetAsync(c.f[F_Z], 0, sizeof(double) * (size_t)c.g.alloc, c.stream));
        return 0;
    }
    if (allow_fused && fcg_smooth_tb_supported(c, sweeps))
        return op_fcg_sweep_passes(c, weights, sweeps, from_zero, dot_w, false);
    if (from_zero) TL_HIP(hipMemsetAsync(c.f[F_Z], 0, sizeof(double) * (size_t)c.g.alloc, c.stream));
    const long n = (long)c.g.nx * c.g.ny * c.g.nz;
    for (int k = 0; k < sweeps; ++k) {
        hipLaunchKernelGGL(k_fcg_smooth, dim3(blocks_for(n)), dim3(256), 0, c.stream, c.g, c.f[F_KX], c.f[F_KY],
                           c.f[F_R], c.f[F_Z], weights[0], sweep_ratio(weights, k), c.f[F_SD]);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) throw Error(std::string("FCG smoothing: ") + hipGetErrorString(e));
        std::swap(c.f[F_Z], c.f[F_SD]);
    }
    return 0;
}

namespace {

void launched(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// Outer iterations at which a solve is sampled: spread over the whole
// convergence history (early, large residuals to late, small ones).
const int kSchedule[] = {1,   2,   3,   5,   8,   13,   21,   34,   55,   89,   144,  233,
                         377, 610, 987, 1597, 2584, 4181, 6765, 10946};

}  // namespace

SolveHooks& solve_hooks() { return g_hooks; }
Learner* active_learner() { return g_active; }
void learn_set_active(Learner* L) { g_active = L; }
UNet& learn_net(Learner& L) { return *L.net; }
long learn_pairs(const Learner& L) { return L.pairs; }
void learn_set_consistent(Learner& L, bool on) { L.consistent = on; }
void learn_set_phases(Learner& L, int mask)
{
    if (mask < 1 || mask > 3) throw Error("snapshot phases: 1 (CG presteps), 2 (PPCG outer) or 3 (both)");
    L.phases = mask;
}

Learner* learn_create(unsigned long long seed)
{
    auto L = std::make_unique<Learner>();
    L->net = unet_create(seed);
    L->rng = seed ^ 0x5DEECE66Dull;
    return L.release();
}

void learn_destroy(Learner* L)
{
    if (!L) return;
    if (g_active == L) g_active = nullptr;
    unet_destroy(L->net);
    for (float* p : {L->dr, L->df, L->de, L->dkx, L->dky, L->br, L->bf, L->be, L->bkx, L->bky, L->pr, L->pf, L->pz,
                     L->ckx, L->cky})
        (void)hipFree(p);
    (void)hipFree(L->crc);
    (void)hipFree(L->scratch);
    delete L;
}

namespace {

struct Snap {
    double *u = nullptr, *r = nullptr;
};

// The iterates of one FP64 PPCG solve (the deck's first timestep) that
// training pairs are made from: u_k, r_k at the schedule's iterations of the
// phases the learner samples, u* at convergence, and the density and face
// coefficients of the chunk.
struct SnapSet {
    std::vector<Snap> snaps;
    double *ustar = nullptr, *dens = nullptr, *kx = nullptr, *ky = nullptr;
    Geom geom{};
    ~SnapSet()
    {
        for (Snap& s : snaps) {
            (void)hipFree(s.u);
            (void)hipFree(s.r);
        }
        (void)hipFree(ustar);
        (void)hipFree(dens);
        (void)hipFree(kx);
        (void)hipFree(ky);
    }
};

void collect_snapshots(Learner& L, const std::string& deck, int snapshots, SnapSet& set)
{
    snapshots = std::max(1, std::min<int>(snapshots, (int)(sizeof kSchedule / sizeof kSchedule[0])));
    bool seen = false;
    // `snapshots` iterates of each phase, at the schedule's iterations: the
    // CG presteps (early, sharp residuals, like the flexible CG's own) and
    // the PPCG outer iterations (late, smooth errors)
    int taken[2] = {0, 0};
    auto snap = [&](Chunk& c, int k, int phase) {
        if (TL_DIM != 2 || !(L.phases & (1 << phase)) || taken[phase] >= snapshots || k != kSchedule[taken[phase]])
            return;
        ++taken[phase];
        const size_t bytes = sizeof(double) * (size_t)c.g.alloc;
        Snap s;
        TL_HIP(hipMalloc(&s.u, bytes));
        set.snaps.push_back(s);
        TL_HIP(hipMalloc(&set.snaps.back().r, bytes));
        TL_HIP(hipMemcpyAsync(set.snaps.back().u, c.f[F_U], bytes, hipMemcpyDeviceToDevice, c.stream));
        TL_HIP(hipMemcpyAsync(set.snaps.back().r, c.f[F_R], bytes, hipMemcpyDeviceToDevice, c.stream));
    };
    g_hooks.on_cg = [&](Chunk& c, int n) { snap(c, n, 0); };
    g_hooks.on_outer = [&](Chunk& c, int k) { snap(c, k, 1); };
    g_hooks.on_end = [&](Chunk& c) {
        if (seen) return;  // the first timestep only
        seen = true;
        set.geom = c.g;
        const size_t bytes = sizeof(double) * (size_t)c.g.alloc;
        for (double** q : {&set.ustar, &set.dens, &set.kx, &set.ky}) TL_HIP(hipMalloc(q, bytes));
        TL_HIP(hipMemcpyAsync(set.ustar, c.f[F_U], bytes, hipMemcpyDeviceToDevice, c.stream));
        TL_HIP(hipMemcpyAsync(set.dens, c.f[F_DENSITY], bytes, hipMemcpyDeviceToDevice, c.stream));
        TL_HIP(hipMemcpyAsync(set.kx, c.f[F_KX], bytes, hipMemcpyDeviceToDevice, c.stream));
        TL_HIP(hipMemcpyAsync(set.ky, c.f[F_KY], bytes, hipMemcpyDeviceToDevice, c.stream));
        TL_HIP(hipStreamSynchronize(c.stream));
    };
    try {
        (void)driver_run_deck(deck, 1);
    } catch (...) {
        g_hooks = SolveHooks{};
        throw;
    }
    g_hooks = SolveHooks{};
    if (!seen || set.snaps.empty()) throw Error("the deck's solve gave no training snapshots (a 2D PPCG deck is needed)");
}

}  // namespace

long learn_add_deck(Learner& L, const std::string& deck, int snapshots, int crops, int crop, unsigned long long seed)
{
    if (crop < 8 || crop % 8) throw Error("training crops must be a multiple of 8 pixels");
    if (L.crop && L.crop != crop) throw Error("all training crops must have the same size");
    if (L.coarse) throw Error("the learner holds coarse-grid pairs; crops cannot be added");
    SnapSet set;
    collect_snapshots(L, deck, snapshots, set);
    const Geom geom = set.geom;
    if (geom.nx < crop || geom.ny < crop) throw Error("the deck's mesh is smaller than the training crop");
    L.crop = crop;
    std::vector<Snap>& snaps = set.snaps;
    double *ustar = set.ustar, *dens = set.dens, *kx = set.kx, *ky = set.ky;
    const long n_new = (long)snaps.size() * crops;
    const long px = (long)crop * crop;
    const long need = (L.pairs + n_new) * px, keep = L.pairs * px;
    regrow(L.dr, L.dr_cap, need, keep);
    regrow(L.df, L.df_cap, need, keep);
    regrow(L.de, L.de_cap, need, keep);
    regrow(L.dkx, L.dkx_cap, need, keep);
    regrow(L.dky, L.dky_cap, need, keep);
    regrow(L.scratch, L.scratch_cap, crops + 1, 0);
    int* dxy = nullptr;
    TL_HIP(hipMalloc(&dxy, sizeof(int) * 2 * crops));
    uint64_t rs = seed;
    std::vector<int> xy(2 * crops);
    for (size_t k = 0; k < snaps.size(); ++k) {
        for (int q = 0; q < crops; ++q) {
            xy[2 * q] = (int)(next(rs) % (uint64_t)(geom.nx - crop + 1));
            xy[2 * q + 1] = (int)(next(rs) % (uint64_t)(geom.ny - crop + 1));
        }
        TL_HIP(hipMemcpy(dxy, xy.data(), sizeof(int) * 2 * crops, hipMemcpyHostToDevice));
        TL_HIP(hipMemset(L.scratch, 0, sizeof(double) * crops));
        const long off = L.pairs * px;
        hipLaunchKernelGGL(k_extract, dim3(crops), dim3(256), 0, nullptr, geom, snaps[k].u, ustar, snaps[k].r, dens, kx,
                           ky, dxy, crop, L.dr + off, L.df + off, L.de + off, L.dkx + off, L.dky + off, L.scratch);
        launched("training pair extraction");
        if (L.consistent) {
            hipLaunchKernelGGL(k_crop_residual, dim3(crops), dim3(256), 0, nullptr, crop, L.de + off, L.dkx + off,
                               L.dky + off, L.dr + off);
            launched("training pair residual");
        }
        // the snapshot's residual rms over the whole mesh scales r and e of
        // all its crops, as the flexible CG scales the network's input
        double* fs = L.scratch + crops;
        TL_HIP(hipMemset(fs, 0, sizeof(double)));
        hipLaunchKernelGGL(k_field_sumsq, dim3(blocks_for((long)geom.nx * geom.ny)), dim3(256), 0, nullptr, geom,
                           snaps[k].r, fs);
        launched("snapshot residual norm");
        double ss = 0.0;
        TL_HIP(hipMemcpy(&ss, fs, sizeof(double), hipMemcpyDeviceToHost));
        const double rms = std::sqrt(ss / ((double)geom.nx * geom.ny));
        const double inv_g = L.escale > 0.0 ? 1.0 / L.escale : 1.0;
        hipLaunchKernelGGL(k_normalise, dim3(crops), dim3(256), 0, nullptr, crop, rms > 0.0 ? 1.0 / rms : 0.0, inv_g,
                           L.dr + off, L.de + off);
        launched("training pair normalisation");
        TL_HIP(hipDeviceSynchronize());
        L.pairs += crops;
    }
    (void)hipFree(dxy);
    return n_new;
}

long learn_add_deck_coarse(Learner& L, const std::string& deck, int snapshots, int nc)
{
    if (L.crop && (!L.coarse || L.crop != nc)) throw Error("all training pairs must be coarse grids of one size");
    SnapSet set;
    collect_snapshots(L, deck, snapshots, set);
    const CoarseGrid cg = coarse_grid(set.geom, nc);
    if (cg.B == 0 || cg.ncy != nc)
        throw Error("the deck's mesh does not divide into " + std::to_string(nc) + " x " + std::to_string(nc) +
                    " square blocks");
    L.crop = nc;
    L.coarse = true;
    const long px = (long)nc * nc;
    const long n_new = (long)set.snaps.size();
    const long need = (L.pairs + n_new) * px, keep = L.pairs * px;
    regrow(L.dr, L.dr_cap, need, keep);
    regrow(L.df, L.df_cap, need, keep);
    regrow(L.de, L.de_cap, need, keep);
    regrow(L.dkx, L.dkx_cap, need, keep);
    regrow(L.dky, L.dky_cap, need, keep);
    regrow(L.scratch, L.scratch_cap, 1, 0);
    for (size_t k = 0; k < set.snaps.size(); ++k) {
        const long off = L.pairs * px;
        // the restricted error e_c of the snapshot and the coarse operator;
        // the input is the consistent coarse residual (I + K_c) e_c
        coarse_fields(set.geom, cg, set.kx, set.ky, set.dens, L.dkx + off, L.dky + off, L.df + off, nullptr);
        coarse_mean_diff(set.geom, cg, set.ustar, set.snaps[k].u, L.de + off, nullptr);
        hipLaunchKernelGGL(k_crop_residual, dim3(1), dim3(256), 0, nullptr, nc, L.de + off, L.dkx + off,
                           L.dky + off, L.dr + off);
        launched("coarse pair residual");
        // scaled by the rms of its own input, as the preconditioner scales it
        TL_HIP(hipMemset(L.scratch, 0, sizeof(double)));
        hipLaunchKernelGGL(k_sumsq, dim3(blocks_for(px)), dim3(256), 0, nullptr, L.dr + off, px, L.scratch);
        launched("coarse pair norm");
        double ss = 0.0;
        TL_HIP(hipMemcpy(&ss, L.scratch, sizeof(double), hipMemcpyDeviceToHost));
        const double rms = std::sqrt(ss / (double)px);
        const double inv_g = L.escale > 0.0 ? 1.0 / L.escale : 1.0;
        hipLaunchKernelGGL(k_normalise, dim3(1), dim3(256), 0, nullptr, nc, rms > 0.0 ? 1.0 / rms : 0.0, inv_g,
                           L.dr + off, L.de + off);
        launched("coarse pair normalisation");
        TL_HIP(hipDeviceSynchronize());
        L.pairs += 1;
    }
    return n_new;
}

void learn_train(Learner& L, int steps, int batch, double lr, double* losses)
{
    if (L.pairs < 1) throw Error("no training pairs");
    if (batch < 1 || steps < 1) throw Error("training needs steps >= 1 and batch >= 1");
    const long px = (long)L.crop * L.crop;
    hipStream_t s = unet_stream(*L.net);
    if (L.escale <= 0.0) {
        // one global target scale: unit rms over the data set
        regrow(L.scratch, L.scratch_cap, 1, 0);
        TL_HIP(hipMemset(L.scratch, 0, sizeof(double)));
        hipLaunchKernelGGL(k_sumsq, dim3(blocks_for(L.pairs * px)), dim3(256), 0, nullptr, L.de, L.pairs * px,
                           L.scratch);
        launched("target scale");
        double ss = 0.0;
        TL_HIP(hipMemcpy(&ss, L.scratch, sizeof(double), hipMemcpyDeviceToHost));
        L.escale = ss > 0.0 ? std::sqrt(ss / (double)(L.pairs * px)) : 1.0;
        hipLaunchKernelGGL(k_scale, dim3(blocks_for(L.pairs * px)), dim3(256), 0, nullptr, L.de, L.pairs * px,
                           (float)(1.0 / L.escale));
        launched("target scale");
        TL_HIP(hipDeviceSynchronize());
    }
    regrow(L.br, L.br_cap, batch * px, 0);
    regrow(L.bf, L.bf_cap, batch * px, 0);
    regrow(L.be, L.be_cap, batch * px, 0);
    regrow(L.bkx, L.bkx_cap, batch * px, 0);
    regrow(L.bky, L.bky_cap, batch * px, 0);
    for (int it = 0; it < steps; ++it) {
        for (int b = 0; b < batch; ++b) {
            const long q = (long)(next(L.rng) % (uint64_t)L.pairs);
            TL_HIP(hipMemcpyAsync(L.br + b * px, L.dr + q * px, sizeof(float) * px, hipMemcpyDeviceToDevice, s));
            TL_HIP(hipMemcpyAsync(L.bf + b * px, L.df + q * px, sizeof(float) * px, hipMemcpyDeviceToDevice, s));
            TL_HIP(hipMemcpyAsync(L.be + b * px, L.de + q * px, sizeof(float) * px, hipMemcpyDeviceToDevice, s));
            TL_HIP(hipMemcpyAsync(L.bkx + b * px, L.dkx + q * px, sizeof(float) * px, hipMemcpyDeviceToDevice, s));
            TL_HIP(hipMemcpyAsync(L.bky + b * px, L.dky + q * px, sizeof(float) * px, hipMemcpyDeviceToDevice, s));
        }
        // the A-norm of the error: the norm flexible CG minimises
        const float loss = unet_loss_grad(*L.net, L.br, L.bf, L.be, L.bkx, L.bky, batch, L.crop, L.crop);
        // cosine learning-rate decay to 5% of lr
        const double frac = (double)it / (double)steps;
        const double lr_t = lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac)));
        unet_adam(*L.net, (float)lr_t);
        if (losses) losses[it] = loss;
    }
    TL_HIP(hipStreamSynchronize(s));
}

bool fcg_learned_solve(Chunk& c, Learner* L, const FcgOptions& opt, double eps, int max_iters, int& iters,
                       double& error)
{
    iters = 0;
    error = 0.0;
    if (TL_DIM != 2 || c.g.nx % 8 || c.g.ny % 8 || c.preconditioner || (!L && opt.network)) return false;
    if (opt.smooth > 32) throw Error("tl_fcg_smooth: at most 32 sweeps");
    double sw[32] = {};
    if (opt.smooth > 0) {
        FcgOptions o = opt;
        if (o.cheby_ratio < 0.0) o.cheby_ratio = fcg_auto_ratio(c, o.smooth);
        fcg_sweep_weights(o, sw);
    }
    const long n = (long)c.g.nx * c.g.ny;
    float *pf = nullptr, *pz = nullptr;
    hipStream_t keep = nullptr;
    CoarseGrid cg{};
    if (opt.coarse) {
        if (!opt.network) throw Error("tl_fcg_coarse needs the network (tl_fcg_no_network is set)");
        if (!L->coarse) throw Error("tl_fcg_coarse: the loaded learned preconditioner was not trained on coarse grids");
        cg = coarse_grid(c.g, L->crop);
        if (cg.B == 0)
            throw Error("tl_fcg_coarse: the mesh does not divide into the network's " + std::to_string(L->crop) +
                        " coarse cells a side (with a multiple of 8 the other way)");
        const long nc = (long)cg.ncx * cg.ncy;
        regrow(L->pr, L->pr_cap, nc, 0);
        regrow(L->pf, L->pf_cap, nc, 0);
        regrow(L->pz, L->pz_cap, nc, 0);
        regrow(L->crc, L->crc_cap, nc + 1, 0);
        regrow(L->ckx, L->ckx_cap, nc, 0);
        regrow(L->cky, L->cky_cap, nc, 0);
        // the coarse density feature (the coarse operator is not needed here)
        coarse_fields(c.g, cg, c.f[F_KX], c.f[F_KY], c.f[F_DENSITY], L->ckx, L->cky, L->pf, c.stream);
        pf = L->pf;
        pz = L->pz;
        keep = unet_stream(*L->net);
        unet_set_stream(*L->net, c.stream);
    } else if (opt.network) {
        regrow(L->pr, L->pr_cap, n, 0);
        regrow(L->pf, L->pf_cap, n, 0);
        regrow(L->pz, L->pz_cap, n, 0);
        pf = L->pf;
        pz = L->pz;
        keep = unet_stream(*L->net);
        unet_set_stream(*L->net, c.stream);
    }
    struct Restore {
        UNet* net;
        hipStream_t s;
        ~Restore()
        {
            if (net) unet_set_stream(*net, s);
        }
    } restore{opt.network ? L->net : nullptr, keep};

    const double cells = (double)n;
    double rrn = op_calc_2norm(c, true);
    if (rrn < eps) {
        error = rrn;
        return true;
    }
    double best = rrn;
    bool u_pending = false;  // u += alpha p of the last step not yet applied
    // (a solve that returns false is re-run from its start by the caller,
    // so only the accepting path needs u current)
    auto flush_u = [&] {
        if (u_pending) op_cg_calc_u_dev(c);
        u_pending = false;
    };
    int since_best = 0;
    const int window = 200;
    bool replaced = false;  // check() put the true residual in r
    bool pr_next = false;   // so the next strip takes alpha from p . r
    // returns 1 to accept (u current, true residual below eps), -1 stalled
    auto check = [&](double r2, int it) -> int {
        iters = it;
        if (!std::isfinite(r2)) return -1;
        if (r2 < eps) {
            // accept on the true residual only (u brought up to date first)
            flush_u();
            op_calc_residual(c);
            r2 = op_calc_2norm(c, true);
            if (r2 < eps) {
                error = r2;
                return 1;
            }
            replaced = true;
        }
        rrn = r2;
        if (r2 < 0.5 * best) {
            best = r2;
            since_best = 0;
        } else if (++since_best >= window) {
            return -1;
        }
        return 0;
    };
    // Before an iteration without the network (2D): iteration k's r -= alpha w
    // and its ||r||^2 ride in iteration k + 1's preconditioner pass, and the
    // host reads them while that iteration's strip runs (which also brings u
    // up to iteration k).  The network reads r first, so the iteration before
    // one that applies it keeps r current.
    const bool sweeps_tb = TL_DIM == 2 && fcg_smooth_tb_supported(c, opt.smooth);
    auto uses_net = [&](int j) {
        if (opt.coarse) return true;  // the two-grid preconditioner in every iteration
        return opt.network && (j - 1) % std::max(1, opt.net_every) == 0 && (opt.net_iters <= 0 || j <= opt.net_iters);
    };
    bool r_pending = false;  // r -= alpha w of the last step not yet applied
    bool first = true;
    for (int it = 1; it <= max_iters; ++it) {
        // z = s N(r / s), then weighted Jacobi sweeps on A z = r from it
        const double s = std::sqrt(rrn / cells);
        long beta_blocks = 0;  // z . w left by the preconditioner pass
        const bool beta_rrn = r_pending;
        if (r_pending) {
            beta_blocks = op_fcg_sweep_passes(c, sw, opt.smooth, true, true, true);
            r_pending = false;
        } else if (opt.coarse) {
            // the two-grid preconditioner: z = S r from zero, the coarse
            // correction z += s G P N(rc / s) with rc = P^T (r - A z) / B^2,
            // then S again from that z (S: the weighted Jacobi sweeps)
            if (opt.coarse_pre) (void)fcg_smooth(c, opt.smooth, sw, true, true, false);
            const long nc = (long)cg.ncx * cg.ncy;
            op_coarse_restrict(c, cg, opt.coarse_pre, L->crc, L->pr, L->crc + nc);
            unet_apply(*L->net, L->pr, pf, pz, 1, cg.ncy, cg.ncx);
            op_coarse_prolong(c, cg, pz, L->crc + nc, L->escale > 0.0 ? L->escale : 1.0, opt.coarse_pre);
            beta_blocks = fcg_smooth(c, opt.smooth, sw, false, true, !first);
        } else if (uses_net(it)) {
            hipLaunchKernelGGL(k_plane_in, dim3(blocks_for(n)), dim3(256), 0, c.stream, c.g, c.f[F_R],
                               c.f[F_DENSITY], s > 0.0 ? 1.0 / s : 0.0, it == 1 ? 1 : 0, L->pr, pf);
            launched("FCG input");
            unet_apply(*L->net, L->pr, pf, pz, 1, c.g.ny, c.g.nx);
            // the network predicts e / (s G): undo both scales
            const double g_scale = L->escale > 0.0 ? L->escale : 1.0;
            hipLaunchKernelGGL(k_plane_out, dim3(blocks_for(n)), dim3(256), 0, c.stream, c.g, pz, s * g_scale,
                               c.f[F_Z]);
            launched("FCG output");
            beta_blocks = fcg_smooth(c, opt.smooth, sw, false, true, !first);
        } else {
            beta_blocks = fcg_smooth(c, opt.smooth, sw, true, true, !first);
        }
        // beta, p, w, alpha, u[, r] with the scalars on the device; one read
        // after the last network iteration a restart begins a new recurrence
        // from the improved iterate (p = z), which needs r current
        const bool restart_next = !opt.coarse && opt.net_restart && uses_net(it) && !uses_net(it + 1);
        const bool defer_r = sweeps_tb && !uses_net(it + 1) && !restart_next && it < max_iters;
        op_fcg_iteration_dev(c, first, beta_blocks, u_pending, beta_rrn, defer_r, !pr_next);
        pr_next = false;
        first = restart_next;
        if (beta_rrn) {
            // iteration it - 1 is complete: r updated by the pass, u by the
            // strip; what is pending now belongs to iteration it
            const double r2 = c.result(S_ERR, 0);
            if (c.result(S_ERR, 1) != 0.0) return false;
            const bool pend = u_pending;
            u_pending = false;
            const int v = check(r2, it - 1);
            if (v != 0) return v > 0;
            u_pending = pend;
            pr_next = replaced;
            replaced = false;
        }
        if (defer_r) {
            r_pending = true;
            continue;
        }
        const double r2 = c.result(S_RRN, 0);
        if (c.result(S_RRN, 1) != 0.0) return false;
        const int v = check(r2, it);
        if (v != 0) return v > 0;
    }
    return false;
}

}  // namespace tl
