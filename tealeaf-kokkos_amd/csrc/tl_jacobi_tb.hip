// Host side of the temporal-blocked Jacobi iterations (2D): one pass per
// call, its per-iteration error sums, and the undo that lets the driver
// stop inside a pass; the device code is tl_jacobi_tb.hpp.
#include "tl_jacobi_tb.hpp"

#include <vector>

namespace tl {

#if TL_DIM == 2
namespace {

// Sums of the per-wave error partials of a K-iteration pass.
template <int K>
__global__ __launch_bounds__(BLOCK) void k_finish_levels(const double* __restrict__ partials, long nblocks,
                                                         double* __restrict__ out)
{
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    for (long b = threadIdx.x; b < nblocks; b += BLOCK) {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += partials[b * K + k];
    }
    block_fold<K>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = v[k];
    }
}

}  // namespace

bool jacobi_tb_supported(const Chunk& c)
{
    // the kernel addresses a field with 32-bit byte offsets
    const bool fits = (double)c.g.alloc * sizeof(double) < 4294967296.0;
    return c.g.nx >= 2 && c.g.ny >= 2 && fits;
}

int jacobi_tb_pass(int remaining)
{
    for (int k : {10, 5, 2, 1})
        if (k <= remaining) return k;
    return 1;
}

// K iterations of ext_jacobi_kernel_solve_ (ext_core.cpp:362-373) in one
// sweep, K from jacobi_tb_pass: u advances by K (its input is kept as
// Chunk::u_alt until the next call), r holds the u before the last of them.
// Returns the error sum of each iteration.
std::vector<double> op_jacobi_tb(Chunk& c, int K)
{
    if (!jacobi_tb_supported(c)) throw Error("temporal-blocked Jacobi is 2D");
    if (K != 1 && K != 2 && K != 5 && K != 10) throw Error("temporal-blocked Jacobi passes are 1, 2, 5 or 10 iterations");
    if (!c.d_levels) TL_HIP(hipMalloc(&c.d_levels, sizeof(double) * 16));
    jtb::JTBArgs a{};
    a.u_in = c.f[F_U];
    a.u_out = c.u_alt;
    a.r = c.f[F_R];
    a.u0 = c.f[F_U0];
    a.kx = c.f[F_KX];
    a.ky = c.f[F_KY];
    a.rowtab = tb_rowtab(c);
    a.row0 = TB_ROWTAB0;
    a.partials = c.partials;
    const int ow = 2 * WAVE - 2 * tb::tb_margin(K);
    a.tiles_x = (c.g.nx + ow - 1) / ow;
    a.rows = tb_rows(c.g.ny, a.tiles_x, K, c.simds, tb::TB_UNROLL);
    a.nblocks = (long)a.tiles_x * ((c.g.ny + a.rows - 1) / a.rows);
    if (a.nblocks * K > 4 * c.partials_cap) throw Error("temporal-blocked Jacobi: more waves than reduction partials");
    TL_TIMED(c, K_JACOBI_TB, {
        switch (K) {
        case 10: jtb::jtb_launch_10(c, a); break;
        case 5: jtb::jtb_launch_5(c, a); break;
        case 2: jtb::jtb_launch_2(c, a); break;
        default: jtb::jtb_launch_1(c, a); break;
        }
    });
    switch (K) {
    case 10: hipLaunchKernelGGL(k_finish_levels<10>, dim3(1), dim3(BLOCK), 0, c.stream, c.partials, a.nblocks, c.d_levels); break;
    case 5: hipLaunchKernelGGL(k_finish_levels<5>, dim3(1), dim3(BLOCK), 0, c.stream, c.partials, a.nblocks, c.d_levels); break;
    case 2: hipLaunchKernelGGL(k_finish_levels<2>, dim3(1), dim3(BLOCK), 0, c.stream, c.partials, a.nblocks, c.d_levels); break;
    default: hipLaunchKernelGGL(k_finish_levels<1>, dim3(1), dim3(BLOCK), 0, c.stream, c.partials, a.nblocks, c.d_levels); break;
    }
    TL_HIP(hipGetLastError());
    std::vector<double> err(K);
    TL_HIP(hipMemcpyAsync(err.data(), c.d_levels, sizeof(double) * K, hipMemcpyDeviceToHost, c.stream));
    TL_HIP(hipStreamSynchronize(c.stream));
    std::swap(c.f[F_U], c.u_alt);
    return err;
}

// Back to the input of the last op_jacobi_tb (r is left as that pass wrote it).
void op_jacobi_tb_undo(Chunk& c) { std::swap(c.f[F_U], c.u_alt); }

#else  // the 7-point stencil runs one iteration per launch

bool jacobi_tb_supported(const Chunk&) { return false; }
int jacobi_tb_pass(int) { return 1; }
std::vector<double> op_jacobi_tb(Chunk&, int) { throw Error("temporal-blocked Jacobi is implemented for 2D"); }
void op_jacobi_tb_undo(Chunk&) { throw Error("temporal-blocked Jacobi is implemented for 2D"); }

#endif

}  // namespace tl
