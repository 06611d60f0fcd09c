// Allocation-free form of the CG matvec-and-dot (ext_cg_calc_w_) for callers
// that cannot allocate device memory: code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc.  It has no counterpart in
// the reference.
//
// The caller passes device arrays in the reference's flat layout
// (tealeaf_2d/ext_chunk.cpp:50-54: X = nx + 4 fastest, halo 2; 3D adds
// Y = ny + 4 and Z = nz + 4) and one device double for the result.  As in
// ext_cg_calc_w_, p's halo must hold the reflective copies the reference's
// update_halo writes (tealeaf_2d/ext_update_halo.hpp); w is written on the
// interior cells only.  w = A p is bitwise the reference's SMVP
// (tealeaf_2d/ext_shared.hpp:37-41, tealeaf_3d/ext_shared.hpp:39-45, same
// operand order, built with -ffp-contract=off) and *pw = sum w.p
// (ext_cg_kernel.hpp:148-197) is summed in a fixed order, so repeated calls
// give the same bits.
//
// Work memory: one FP64 partial sum per mesh row (ny rows in 2D, ny * nz in
// 3D), in the file-scope device array `scratch` of kMaxRows doubles
// (512 KiB, held by every GPU the library runs on until the process ends).
// Meshes with more rows are refused.  Each row's partial is written before
// it is read and exactly `rows` partials are read, so nothing needs
// clearing between calls.  Calls that overlap on one GPU (on different
// streams, or from several host threads) share that GPU's array and race.
#include "tl_chunk.hpp"

#include <mutex>

namespace tl {

namespace {

constexpr int kMaxRows = 65536;
constexpr int kBlock = 256;

static __device__ __align__(256) double scratch[kMaxRows];

struct FlatMesh {
    int nx, ny, nz;
    long X, XY;
};

// One workgroup per row (jj, ii): w = A p on the row's interior cells and
// the row's sum of w.p (lanes in column order, then a fixed tree).
__global__ void __launch_bounds__(kBlock) k_cg_calc_w_rows(FlatMesh m, const double* __restrict__ p,
                                                           const double* __restrict__ kx,
                                                           const double* __restrict__ ky,
                                                           const double* __restrict__ kz, double* __restrict__ w,
                                                           double* __restrict__ part)
{
    const int q = blockIdx.x;
    const int jj = HALO + q % m.ny;
    const int ii = TL_DIM == 3 ? HALO + q / m.ny : 0;
    const long row = (long)ii * m.XY + (long)jj * m.X;
    double s = 0.0;
    for (int kk = HALO + (int)threadIdx.x; kk < m.nx + HALO; kk += kBlock) {
        const long i = row + kk;
#if TL_DIM == 3
        const double v = (1.0 + (kx[i + 1] + kx[i]) + (ky[i + m.X] + ky[i]) + (kz[i + m.XY] + kz[i])) * p[i]
                       - (kx[i + 1] * p[i + 1] + kx[i] * p[i - 1])
                       - (ky[i + m.X] * p[i + m.X] + ky[i] * p[i - m.X])
                       - (kz[i + m.XY] * p[i + m.XY] + kz[i] * p[i - m.XY]);
#else
        (void)kz;
        const double v = (1.0 + (kx[i + 1] + kx[i]) + (ky[i + m.X] + ky[i])) * p[i]
                       - (kx[i + 1] * p[i + 1] + kx[i] * p[i - 1])
                       - (ky[i + m.X] * p[i + m.X] + ky[i] * p[i - m.X]);
#endif
        w[i] = v;
        s += v * p[i];
    }
    __shared__ double ws[kBlock / WAVE];
    for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (threadIdx.x % WAVE == 0) ws[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = ws[0];
        for (int k = 1; k < kBlock / WAVE; ++k) t += ws[k];
        part[q] = t;
    }
}

// Adds the `rows` partials in a fixed order into *out.
__global__ void __launch_bounds__(kBlock) k_sum_partials(const double* __restrict__ part, int rows,
                                                         double* __restrict__ out)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < rows; k += kBlock) s += part[k];
    __shared__ double ws[kBlock / WAVE];
    for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (threadIdx.x % WAVE == 0) ws[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = ws[0];
        for (int k = 1; k < kBlock / WAVE; ++k) t += ws[k];
        *out = t;
    }
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (in host code the array's name is only a stand-in for it).
double* scratch_ptr()
{
    int dev = 0;
    TL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error("tl_cg_calc_w_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(std::string("tl_cg_calc_w_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                        hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    return g_scratch[dev];
}

}  // namespace

void cg_calc_w_noalloc(int nx, int ny, int nz, const double* p, const double* kx, const double* ky,
                       const double* kz, double* w, double* pw, hipStream_t stream)
{
    if (nx < 1 || ny < 1 || nz < 1) throw Error("tl_cg_calc_w_noalloc: mesh extents must be at least 1");
    if (!p || !kx || !ky || !w || !pw || (TL_DIM == 3 && !kz))
        throw Error("tl_cg_calc_w_noalloc: null device pointer");
    const long rows = TL_DIM == 3 ? (long)ny * nz : (long)ny;
    if (rows > kMaxRows)
        throw Error("tl_cg_calc_w_noalloc: " + std::to_string(rows) + " mesh rows exceed the " +
                    std::to_string(kMaxRows) + " partial sums its scratch array holds");
    FlatMesh m{nx, ny, nz, (long)nx + 2 * HALO, ((long)nx + 2 * HALO) * ((long)ny + 2 * HALO)};
    double* part = scratch_ptr();
    hipLaunchKernelGGL(k_cg_calc_w_rows, dim3((unsigned)rows), dim3(kBlock), 0, stream, m, p, kx, ky, kz, w, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("tl_cg_calc_w_noalloc: ") + hipGetErrorString(e));
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, stream, part, (int)rows, pw);
    e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("tl_cg_calc_w_noalloc: ") + hipGetErrorString(e));
}

}  // namespace tl
