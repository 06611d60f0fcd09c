// Coarse grid of the learned two-grid preconditioner (no counterpart in the
// reference): the flexible CG's preconditioner smooths A z = r with weighted
// Jacobi sweeps, restricts the remaining residual to a grid of B x B cell
// blocks, solves the coarse problem with the BF16 U-Net, adds the
// prolongated correction and smooths again.
//
// With piecewise-constant prolongation P (every fine cell takes its block's
// value) the Galerkin coarse operator P^T A P, divided by B^2, has the form
// of the fine one: unit diagonal plus, per coarse face, the sum of the fine
// face coefficients across it over B^2, and no flux across the mesh
// boundary (A's reflective boundary carries none).  So the coarse equation
// is  (I + K_c) e_c = P^T (r - A z) / B^2, and the network, trained on whole
// coarse grids with that operator, maps the scaled right-hand side to e_c.
#include "tl_chunk.hpp"

#include <cmath>

namespace tl {

#if TL_DIM == 2

namespace {

constexpr int CBLOCK = 256;

// rc[J * ncx + I] = sum over block (I, J) of (r - A z) / B^2, A with no flux
// across the mesh boundary (z null: of r alone).  One workgroup per coarse cell; thread t takes
// the block's cells t, t + 256, ... and the workgroup folds in a fixed order.
__global__ __launch_bounds__(CBLOCK) void k_coarse_restrict(Geom g, int B, int ncx, const double* __restrict__ r,
                                                             const double* __restrict__ z,
                                                             const double* __restrict__ kx,
                                                             const double* __restrict__ ky, double inv_b2,
                                                             double* __restrict__ rc)
{
    const int I = blockIdx.x % ncx, J = blockIdx.x / ncx;
    double v[1] = {0.0};
    for (int t = threadIdx.x; t < B * B; t += CBLOCK) {
        const int x = HALO + I * B + t % B, y = HALO + J * B + t / B;
        const long c = gidx(g, x, y, 0);
        double az = 0.0;
        if (z) {
            const double zc = z[c];
            az = zc;
            if (x < g.nx + 1) az += kx[c + 1] * (zc - z[c + 1]);
            if (x > HALO) az += kx[c] * (zc - z[c - 1]);
            if (y < g.ny + 1) az += ky[c + g.sx] * (zc - z[c + g.sx]);
            if (y > HALO) az += ky[c] * (zc - z[c - g.sx]);
        }
        v[0] += r[c] - az;
    }
    block_fold<1, CBLOCK / WAVE>(v);
    if (threadIdx.x == 0) rc[blockIdx.x] = v[0] * inv_b2;
}

// The network's input plane: rc / s (FP32) with s the rms of rc, kept in
// *s_out for the prolongation.  One workgroup.
__global__ __launch_bounds__(CBLOCK) void k_coarse_plane(const double* __restrict__ rc, long n,
                                                          float* __restrict__ pr, double* __restrict__ s_out)
{
    __shared__ double inv;
    double v[1] = {0.0};
    for (long i = threadIdx.x; i < n; i += CBLOCK) v[0] += rc[i] * rc[i];
    block_fold<1, CBLOCK / WAVE>(v);
    if (threadIdx.x == 0) {
        const double s = std::sqrt(v[0] / (double)n);
        *s_out = s;
        inv = s > 0.0 ? 1.0 / s : 0.0;
    }
    __syncthreads();
    for (long i = threadIdx.x; i < n; i += CBLOCK) pr[i] = (float)(rc[i] * inv);
}

// z += scale * s * ez[block] on every interior cell (P e_c, e_c = s G N(rc / s));
// z = that without `add`.
__global__ void k_coarse_prolong(Geom g, int B, int ncx, const float* __restrict__ ez,
                                 const double* __restrict__ s_dev, double scale, int add, double* __restrict__ z)
{
    const double f = scale * *s_dev;
    const long n = (long)g.nx * g.ny;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int x = (int)(i % g.nx), y = (int)(i / g.nx);
        const long c = gidx(g, HALO + x, HALO + y, 0);
        const double e = f * (double)ez[(long)(y / B) * ncx + x / B];
        z[c] = add ? z[c] + e : e;
    }
}

// The coarse operator and feature, per coarse cell (I, J): the west face
// kxc = sum of the fine kx across it / B^2 (0 on the mesh boundary), the
// south face kyc likewise, and the mean of the density feature
// (log10(rho) - 0.5) / 1.5 over the block.  One workgroup per coarse cell.
__global__ __launch_bounds__(CBLOCK) void k_coarse_fields(Geom g, int B, int ncx, const double* __restrict__ kx,
                                                           const double* __restrict__ ky,
                                                           const double* __restrict__ density, double inv_b2,
                                                           float* __restrict__ kxc, float* __restrict__ kyc,
                                                           float* __restrict__ fc)
{
    const int I = blockIdx.x % ncx, J = blockIdx.x / ncx;
    double v[3] = {0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < B * B; t += CBLOCK) {
        const int x = HALO + I * B + t % B, y = HALO + J * B + t / B;
        const long c = gidx(g, x, y, 0);
        v[2] += (std::log10(density[c]) - 0.5) / 1.5;
        if (I > 0 && t % B == 0) v[0] += kx[c];
        if (J > 0 && t / B == 0) v[1] += ky[c];
    }
    block_fold<3, CBLOCK / WAVE>(v);
    if (threadIdx.x == 0) {
        kxc[blockIdx.x] = (float)(v[0] * inv_b2);
        kyc[blockIdx.x] = (float)(v[1] * inv_b2);
        fc[blockIdx.x] = (float)(v[2] * inv_b2);
    }
}

// Block mean of a - b (the restricted error of a snapshot: u* - u_k).
__global__ __launch_bounds__(CBLOCK) void k_coarse_mean_diff(Geom g, int B, int ncx, const double* __restrict__ a,
                                                              const double* __restrict__ b, double inv_b2,
                                                              float* __restrict__ out)
{
    const int I = blockIdx.x % ncx, J = blockIdx.x / ncx;
    double v[1] = {0.0};
    for (int t = threadIdx.x; t < B * B; t += CBLOCK) {
        const long c = gidx(g, HALO + I * B + t % B, HALO + J * B + t / B, 0);
        v[0] += a[c] - b[c];
    }
    block_fold<1, CBLOCK / WAVE>(v);
    if (threadIdx.x == 0) out[blockIdx.x] = (float)(v[0] * inv_b2);
}

void launched(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

CoarseGrid coarse_grid(const Geom& g, int nc)
{
    CoarseGrid cg{};
    if (nc < 8 || nc % 8 || g.nx % nc) return cg;
    const int B = g.nx / nc;
    if (B < 2 || g.ny % B || (g.ny / B) % 8) return cg;
    cg.B = B;
    cg.ncx = nc;
    cg.ncy = g.ny / B;
    return cg;
}

void op_coarse_restrict(Chunk& c, const CoarseGrid& cg, bool with_z, double* rc, float* plane, double* s_dev)
{
    const double inv = 1.0 / ((double)cg.B * cg.B);
    const long n = (long)cg.ncx * cg.ncy;
    hipLaunchKernelGGL(k_coarse_restrict, dim3((unsigned)n), dim3(CBLOCK), 0, c.stream, c.g, cg.B, cg.ncx, c.f[F_R],
                       with_z ? c.f[F_Z] : nullptr, c.f[F_KX], c.f[F_KY], inv, rc);
    launched("coarse restriction");
    hipLaunchKernelGGL(k_coarse_plane, dim3(1), dim3(CBLOCK), 0, c.stream, rc, n, plane, s_dev);
    launched("coarse input plane");
}

void op_coarse_prolong(Chunk& c, const CoarseGrid& cg, const float* ez, const double* s_dev, double scale, bool add)
{
    const long n = (long)c.g.nx * c.g.ny;
    const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_coarse_prolong, dim3(blocks), dim3(256), 0, c.stream, c.g, cg.B, cg.ncx, ez, s_dev, scale,
                       add ? 1 : 0, c.f[F_Z]);
    launched("coarse prolongation");
}

void coarse_fields(const Geom& g, const CoarseGrid& cg, const double* kx, const double* ky, const double* density,
                   float* kxc, float* kyc, float* fc, hipStream_t s)
{
    const double inv = 1.0 / ((double)cg.B * cg.B);
    hipLaunchKernelGGL(k_coarse_fields, dim3((unsigned)((long)cg.ncx * cg.ncy)), dim3(CBLOCK), 0, s, g, cg.B, cg.ncx,
                       kx, ky, density, inv, kxc, kyc, fc);
    launched("coarse operator");
}

void coarse_mean_diff(const Geom& g, const CoarseGrid& cg, const double* a, const double* b, float* out,
                      hipStream_t s)
{
    const double inv = 1.0 / ((double)cg.B * cg.B);
    hipLaunchKernelGGL(k_coarse_mean_diff, dim3((unsigned)((long)cg.ncx * cg.ncy)), dim3(CBLOCK), 0, s, g, cg.B,
                       cg.ncx, a, b, inv, out);
    launched("coarse restriction of the error");
}

#else  // the learned preconditioner is 2D

CoarseGrid coarse_grid(const Geom&, int) { return CoarseGrid{}; }
void op_coarse_restrict(Chunk&, const CoarseGrid&, bool, double*, float*, double*)
{
    throw Error("the coarse-grid preconditioner is 2D");
}
void op_coarse_prolong(Chunk&, const CoarseGrid&, const float*, const double*, double, bool)
{
    throw Error("the coarse-grid preconditioner is 2D");
}
void coarse_fields(const Geom&, const CoarseGrid&, const double*, const double*, const double*, float*, float*, float*,
                   hipStream_t)
{
    throw Error("the coarse-grid preconditioner is 2D");
}
void coarse_mean_diff(const Geom&, const CoarseGrid&, const double*, const double*, float*, hipStream_t)
{
    throw Error("the coarse-grid preconditioner is 2D");
}

#endif

}  // namespace tl
