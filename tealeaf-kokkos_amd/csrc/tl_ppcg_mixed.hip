// Mixed-precision PPCG outer iteration for the 2D 5-point stencil.
//
// The FP64 outer iteration of PPCG (tea_solve.f90:307-374) is kept: CGCalcW,
// alpha = rro / pw, u += alpha p, r -= alpha w, rrn = ||r||^2 and the beta /
// p recurrence all stay in FP64, as do the eigenvalue estimate and the
// Chebyshev coefficients.  Only the polynomial correction of the inner steps
// (PPCGInitSd, then L x [r -= A sd; u += sd; sd = a sd + b r],
// ext_ppcg_kernel.hpp:8-100) runs in FP32, on the FP32 rounding of the FP64
// residual:
//
//   r32 = (float) r',  sd = r32 / theta,  du = sum of the sd's,
//   r32 <- r32 - A32 sd  per step,
//   u_out = u' + du,   r_out = r' + (r32_L - r32_0)          (FP64 adds)
//
// so the FP32 error enters the new residual only through r32_L - r32_0, a
// quantity the outer iteration measures again in FP64 (rrn), and the driver
// replaces the recurrence residual by the true FP64 residual u0 - A u
// whenever ||r||^2 has fallen by a set factor and before it accepts
// convergence.
//
// A32 sd is evaluated in difference form, (A sd)_c = sd_c + sum_f k_f
// (sd_c - sd_f), with FMAs: the reference's (1 + sum k) sd_c - sum k sd_f
// cancels catastrophically in FP32 when the face coefficients are large
// (k ~ 1e4 on fine meshes), the difference form does not.
//
// The sweep is the temporal-blocked wavefront of the FP64 kernel: a wave
// owns 128 columns (two per lane), streams down a strip of rows and keeps
// per level a three-row window of sd, two-row windows of r and du; the FP32
// kx/ky rows and the FP64 r' = r - alpha w of level 0 live in LDS rings
// until the output row, where u' = u + alpha p is formed from u, p loaded
// there, so every input is read from memory once.
#include "tl_chunk.hpp"

#include <algorithm>

namespace tl {

#if TL_DIM == 2
namespace {

constexpr int MX_MAX_LEVELS = 12;

// Inner-step counts with a compiled kernel (each costs build time); other
// counts run the FP64 kernels.
__host__ __device__ constexpr bool mx_levels_built(int L) { return L == 1 || L == 3 || L == 10; }
#ifndef TL_MX_WAVES_PER_SIMD
#define TL_MX_WAVES_PER_SIMD 1
#endif
constexpr int MX_WAVES_PER_SIMD = TL_MX_WAVES_PER_SIMD;  // register budget: 512 / this per lane

struct MXArgs {
    double* u;
    const double* p;
    const double* r_in;
    const double* w;
    double* r_out;
    const float* kx;  // FP32 copies, same element index as the FP64 fields
    const float* ky;
    const DevScalars* ds;
    double inv_theta;
    const float* coef;  // a[0..L) then b[0..L)
    double* partials;
    int tiles_x;
    int rows;
    long nblocks;
};

__device__ __forceinline__ float f_from_west(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

__device__ __forceinline__ float f_from_east(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

template <int L>
struct MXState {
    float sdw[L][3][2];   // sd of level s, slot (round) % 3, column
    float rw[L][2][2];    // r32 of level s, slot (round) % 2
    float duw[L][2][2];   // du after level s (s >= 1), slot (round) % 2
    double pin[2][2][2];  // prefetched level-0 rows: r, w
    float pk[2][2][2];    // prefetched level-0 rows: kx, ky
    double po[2][2][2];   // prefetched output rows: u, p
    double acc;
};

struct MXCtx {
    Geom g;
    int nx, ny, ylo, Y0, yhi;
    unsigned bc0, bc1, bf0, bf1, bs0, bs1, bpad;  // byte offsets of doubles in a row
    unsigned hc0, hc1, hf0, hf1;                  // byte offsets of floats in a row
    bool wcl0, wcl1, ecl0, ecl1;
    bool out0, out1;
    double alpha, inv_theta;
    float4* kxr;   // LDS ring [L+1][64]: kx at c0, c1, c1 + 1
    float2* kyr;   // LDS ring [L+1][64]: ky at c0, c1
    double2* rpr;  // LDS ring [L+1][64]: r' at c0, c1
    const float* coef;
};

template <class T>
__device__ __forceinline__ T ldb(const T* row, unsigned boff)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(row) + boff);
}

__device__ __forceinline__ void stb(double* row, unsigned boff, double v)
{
    *reinterpret_cast<double*>(reinterpret_cast<char*>(row) + boff) = v;
}

__device__ __forceinline__ int ring(int row, int n)
{
    const int m = row % n;
    return m < 0 ? m + n : m;
}

// Prefetch for round y + 2: level-0 inputs of row y + 2 and the output row
// y + 2 - L.  Every load is issued on every round, so the count of loads in
// flight never depends on a branch.  Waves that reach a y face (YE) clamp
// the rows; the others read rows y + 2 and y + 2 - L as they are (the wave
// placement keeps them inside the mesh), one row offset for every array.
template <int L, bool YE>
__device__ __forceinline__ void mx_load(const MXArgs& A, const MXCtx& c, int y, int slot, MXState<L>& st)
{
    // byte offsets of the rows (the host keeps every field under 4 GiB);
    // each load is the array's base plus a per-lane 32-bit offset
    const unsigned oy = (unsigned)(c.g.base + (long)y * c.g.sx);
    const unsigned oyr = YE ? (unsigned)(c.g.base + (long)min(y, c.ny + 1) * c.g.sx) : oy;
    const unsigned oyk = YE ? (unsigned)(c.g.base + (long)min(y, c.ny + 2) * c.g.sx) : oy;  // ky also on the top face
    const unsigned oyo = YE ? (unsigned)(c.g.base + (long)min(max(y - L, c.Y0), c.yhi) * c.g.sx)
                            : oy - (unsigned)(L * c.g.sx);
    st.pin[slot][0][0] = ldb(A.r_in, 8u * oyr + c.bc0);
    st.pin[slot][0][1] = ldb(A.r_in, 8u * oyr + c.bc1);
    st.pin[slot][1][0] = ldb(A.w, 8u * oyr + c.bc0);
    st.pin[slot][1][1] = ldb(A.w, 8u * oyr + c.bc1);
    st.pk[slot][0][0] = ldb(A.kx, 4u * oyr + c.hf0);
    st.pk[slot][0][1] = ldb(A.kx, 4u * oyr + c.hf1);
    st.pk[slot][1][0] = ldb(A.ky, 4u * oyk + c.hc0);
    st.pk[slot][1][1] = ldb(A.ky, 4u * oyk + c.hc1);
    st.po[slot][0][0] = ldb(A.u, 8u * oyo + c.bc0);
    st.po[slot][0][1] = ldb(A.u, 8u * oyo + c.bc1);
    st.po[slot][1][0] = ldb(A.p, 8u * oyo + c.bc0);
    st.po[slot][1][1] = ldb(A.p, 8u * oyo + c.bc1);
}

// One round: level 0 at row y, level s at row y - s (s = 1..L); level L's
// row is the output row.
template <int L, bool XE, bool YE, int PH>
__device__ __forceinline__ void mx_round(const MXArgs& A, const MXCtx& c, MXState<L>& st, int y)
{
    constexpr int S0 = PH % 3, S1 = (PH + 2) % 3, S2 = (PH + 1) % 3;  // rounds y, y-1, y-2
    constexpr int R0 = PH % 2, R1 = (PH + 1) % 2;                      // rounds y, y-1
    constexpr int P = PH % 2;                                          // prefetch slot of this round
    const int lane = threadIdx.x;
    const int ny = c.ny;

    int rn = ring(y - 1, L + 1);
    float4 nkx = c.kxr[rn * WAVE + lane];
    float2 nky = c.kyr[rn * WAVE + lane];

    // level 0 at row y: r' = r - alpha w in FP64, then its FP32 rounding
    const int ry = ring(y, L + 1);
    float kyn0, kyn1;
    if (!YE || y <= ny + 1) {
        double rp[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            rp[q] = st.pin[P][0][q] - c.alpha * st.pin[P][1][q];
            st.rw[0][R0][q] = (float)rp[q];
            st.sdw[0][S0][q] = (float)(rp[q] * c.inv_theta);
        }
        c.rpr[ry * WAVE + lane] = make_double2(rp[0], rp[1]);
        const float kx0 = st.pk[P][0][0], kx1 = st.pk[P][0][1];
        c.kxr[ry * WAVE + lane] = make_float4(kx0, kx1, f_from_east(kx0), 0.0f);
        c.kyr[ry * WAVE + lane] = make_float2(st.pk[P][1][0], st.pk[P][1][1]);
        kyn0 = st.pk[P][1][0];
        kyn1 = st.pk[P][1][1];
    } else {
        // beyond the top face: the reflective mirror repeats the last row
        const bool face = y == ny + 2;
        kyn0 = face ? st.pk[P][1][0] : 0.0f;
        kyn1 = face ? st.pk[P][1][1] : 0.0f;
        c.kyr[ry * WAVE + lane] = make_float2(kyn0, kyn1);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            st.sdw[0][S0][q] = st.sdw[0][S1][q];
            st.rw[0][R0][q] = st.rw[0][R1][q];
        }
    }
    // the output row of this round: u, p loaded two rounds ago, r' from
    // the ring (row y - L shares no slot with row y)
    double po[2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        po[k][0] = st.po[P][k][0];
        po[k][1] = st.po[P][k][1];
    }
    const double2 rpo = c.rpr[ring(y - L, L + 1) * WAVE + lane];
    mx_load<L, YE>(A, c, y + 2, P, st);

#pragma unroll
    for (int s = 1; s <= L; ++s) {
        const int z = y - s;
        const float4 kx = nkx;
        const float2 ky = nky;
        const float a = c.coef[s - 1], b = c.coef[L + s - 1];
        if (s < L) {
            rn = ring(z - 1, L + 1);
            nkx = c.kxr[rn * WAVE + lane];
            nky = c.kyr[rn * WAVE + lane];
        }
        float t0 = 0.0f, t1 = 0.0f, d0 = 0.0f, d1 = 0.0f;
        if (!YE || z <= ny + 1) {
            const float c0 = st.sdw[s - 1][S1][0], c1 = st.sdw[s - 1][S1][1];
            float s0 = st.sdw[s - 1][S2][0], s1 = st.sdw[s - 1][S2][1];
            float n0 = st.sdw[s - 1][S0][0], n1 = st.sdw[s - 1][S0][1];
            if (YE) {
                if (z == HALO) { s0 = c0; s1 = c1; }
                if (z == ny + 1) { n0 = c0; n1 = c1; }
            }
            float w0 = f_from_west(c1), e1 = f_from_east(c0);
            float e0 = c1, w1 = c0;
            if (XE) {
                if (c.wcl0) w0 = c0;
                if (c.wcl1) w1 = c1;
                if (c.ecl0) e0 = c0;
                if (c.ecl1) e1 = c1;
            }
            // r - A sd = (r - sd) + sum_f k_f (sd_f - sd)
            t0 = st.rw[s - 1][R1][0] - c0;
            t0 = __builtin_fmaf(kx.y, e0 - c0, t0);
            t0 = __builtin_fmaf(kx.x, w0 - c0, t0);
            t0 = __builtin_fmaf(kyn0, n0 - c0, t0);
            t0 = __builtin_fmaf(ky.x, s0 - c0, t0);
            t1 = st.rw[s - 1][R1][1] - c1;
            t1 = __builtin_fmaf(kx.z, e1 - c1, t1);
            t1 = __builtin_fmaf(kx.y, w1 - c1, t1);
            t1 = __builtin_fmaf(kyn1, n1 - c1, t1);
            t1 = __builtin_fmaf(ky.y, s1 - c1, t1);
            kyn0 = ky.x;
            kyn1 = ky.y;
            d0 = s == 1 ? c0 : st.duw[s - 1][R1][0] + c0;
            d1 = s == 1 ? c1 : st.duw[s - 1][R1][1] + c1;
            if (s < L) {
                st.sdw[s][S0][0] = __builtin_fmaf(a, c0, b * t0);
                st.sdw[s][S0][1] = __builtin_fmaf(a, c1, b * t1);
                st.rw[s][R0][0] = t0;
                st.rw[s][R0][1] = t1;
                st.duw[s][R0][0] = d0;
                st.duw[s][R0][1] = d1;
            }
        } else {
            kyn0 = ky.x;
            kyn1 = ky.y;
            if (s < L) {
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    st.sdw[s][S0][q] = st.sdw[s][S1][q];
                    st.rw[s][R0][q] = st.rw[s][R1][q];
                    st.duw[s][R0][q] = st.duw[s][R1][q];
                }
            }
        }
        if (s == L) {
            // u_out = u' + du, r_out = r' + (r32_L - r32_0), in FP64; every
            // lane stores, lanes and rows that own no cell into a pad column
            const bool zin = (!YE || z <= ny + 1) && z >= c.Y0 && z <= c.yhi;
            const bool o0 = zin && c.out0, o1 = zin && c.out1;
            const unsigned b0 = zin ? c.bs0 : c.bpad, b1 = zin ? c.bs1 : c.bpad;
            const unsigned oz = 8u * (unsigned)(c.g.base + (long)(YE ? min(max(z, 0), ny + 2 * HALO - 1) : z) * c.g.sx);
            const double rp0 = rpo.x, rp1 = rpo.y;
            const double up0 = po[0][0] + c.alpha * po[1][0], up1 = po[0][1] + c.alpha * po[1][1];
            const double rr0 = rp0 + (double)(t0 - (float)rp0);
            const double rr1 = rp1 + (double)(t1 - (float)rp1);
            stb(A.r_out, oz + b0, rr0);
            stb(A.u, oz + b0, up0 + (double)d0);
            stb(A.r_out, oz + b1, rr1);
            stb(A.u, oz + b1, up1 + (double)d1);
            st.acc += o0 ? rr0 * rr0 : 0.0;
            st.acc += o1 ? rr1 * rr1 : 0.0;
        }
    }
}

template <int L, bool XE, bool YE>
__device__ __forceinline__ void mx_wave(const MXArgs& A, const MXCtx& c)
{
    MXState<L> st;
#pragma unroll
    for (int s = 0; s < L; ++s) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            st.sdw[s][0][q] = st.sdw[s][1][q] = st.sdw[s][2][q] = 0.0f;
            st.rw[s][0][q] = st.rw[s][1][q] = 0.0f;
            st.duw[s][0][q] = st.duw[s][1][q] = 0.0f;
        }
    }
    st.acc = 0.0;
    mx_load<L, YE>(A, c, c.ylo, 0, st);
    mx_load<L, YE>(A, c, c.ylo + 1, 1, st);
    const int yend = c.yhi + L;
    for (int y = c.ylo; y <= yend; y += 6) {
        mx_round<L, XE, YE, 0>(A, c, st, y);
        mx_round<L, XE, YE, 1>(A, c, st, y + 1);
        mx_round<L, XE, YE, 2>(A, c, st, y + 2);
        mx_round<L, XE, YE, 3>(A, c, st, y + 3);
        mx_round<L, XE, YE, 4>(A, c, st, y + 4);
        mx_round<L, XE, YE, 5>(A, c, st, y + 5);
    }
    double v[1] = {st.acc};
    block_partial<1, 1>(v, Reduce{A.partials});
}

template <int L>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(TL_MX_WAVES_PER_SIMD))) void k_ppcg_mixed(
    Geom g, MXArgs A)
{
    __shared__ float4 kx_ring[(L + 1) * WAVE];
    __shared__ float2 ky_ring[(L + 1) * WAVE];
    __shared__ double2 rp_ring[(L + 1) * WAVE];
    __shared__ float coef[2 * L];
    const int lane = threadIdx.x;
    if (lane < 2 * L) coef[lane] = A.coef[lane];
    __syncthreads();
    const long t = xcd_tile(blockIdx.x, A.nblocks);
    const int tx = (int)(t % A.tiles_x), ty = (int)(t / A.tiles_x);
    constexpr int OW = 2 * WAVE - 2 * L;
    MXCtx c;
    c.g = g;
    c.nx = g.nx;
    c.ny = g.ny;
    const int X0 = HALO + tx * OW;
    c.Y0 = HALO + ty * A.rows;
    const int k0 = X0 - L + 2 * lane, k1 = k0 + 1;
    const int kc0 = min(max(k0, HALO), c.nx + 1), kc1 = min(max(k1, HALO), c.nx + 1);
    const int kf0 = min(max(k0, HALO), c.nx + 2), kf1 = min(max(k1, HALO), c.nx + 2);
    c.bc0 = 8u * (unsigned)kc0;
    c.bc1 = 8u * (unsigned)kc1;
    c.bf0 = 8u * (unsigned)kf0;
    c.bf1 = 8u * (unsigned)kf1;
    c.hc0 = 4u * (unsigned)kc0;
    c.hc1 = 4u * (unsigned)kc1;
    c.hf0 = 4u * (unsigned)kf0;
    c.hf1 = 4u * (unsigned)kf1;
    c.wcl0 = k0 == HALO;
    c.wcl1 = k1 == HALO;
    c.ecl0 = k0 == c.nx + 1;
    c.ecl1 = k1 == c.nx + 1;
    c.out0 = k0 >= X0 && k0 < X0 + OW && k0 <= c.nx + 1;
    c.out1 = k1 >= X0 && k1 < X0 + OW && k1 <= c.nx + 1;
    {
        // pad columns X .. sx-1 of every row are never read
        const int npad = (int)(g.sx - g.X);
        c.bpad = 8u * (unsigned)(g.X + lane % npad);
        c.bs0 = c.out0 ? c.bc0 : c.bpad;
        c.bs1 = c.out1 ? c.bc1 : c.bpad;
    }
    c.ylo = max(c.Y0 - L, HALO);
    c.yhi = min(c.Y0 + A.rows - 1, c.ny + 1);
    c.alpha = A.ds->alpha;
    c.inv_theta = A.inv_theta;
    c.kxr = kx_ring;
    c.kyr = ky_ring;
    c.rpr = rp_ring;
    c.coef = coef;
    const bool xedge = X0 - L <= HALO || X0 - L + 2 * WAVE - 1 >= c.nx + 1;
    // waves near a y face clamp their rows; the others read rows ylo + 2 - L
    // .. yhi + L + 7 (prefetch and the six-round unroll) unclamped
    const bool yedge = c.ylo == HALO || c.Y0 < 2 * L + HALO || c.yhi + L + 8 >= c.ny + 1;
    if (xedge) {
        if (yedge) mx_wave<L, true, true>(A, c);
        else mx_wave<L, true, false>(A, c);
    } else {
        if (yedge) mx_wave<L, false, true>(A, c);
        else mx_wave<L, false, false>(A, c);
    }
}

template <int L>
void launch(Chunk& c, MXArgs& a)
{
    hipLaunchKernelGGL((k_ppcg_mixed<L>), dim3((unsigned)a.nblocks), dim3(WAVE), 0, c.stream, c.g, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
}

__global__ void k_to_f32(const double* __restrict__ a, float* __restrict__ b, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

}  // namespace

bool ppcg_mixed_supported(const Chunk& c, int inner_steps)
{
    // the kernel addresses a field with 32-bit byte offsets
    const bool fits = (double)c.g.alloc * sizeof(double) < 4294967296.0;
    return !c.preconditioner && mx_levels_built(inner_steps) && c.g.nx >= 2 && c.g.ny >= 2 && fits;
}

void op_mixed_prepare(Chunk& c, const double* alphas, const double* betas, int inner_steps, double theta)
{
    if (!ppcg_mixed_supported(c, inner_steps))
        throw Error("mixed-precision PPCG needs no preconditioner and 1, 3 or 10 inner steps");
    const size_t n = (size_t)c.g.alloc;
    if (!c.kx32) TL_HIP(hipMalloc(&c.kx32, sizeof(float) * n));
    if (!c.ky32) TL_HIP(hipMalloc(&c.ky32, sizeof(float) * n));
    if (!c.d_coef32) TL_HIP(hipMalloc(&c.d_coef32, sizeof(float) * 2 * MX_MAX_LEVELS));
    const unsigned blocks = (unsigned)std::min<long>(((long)n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_f32, dim3(blocks), dim3(256), 0, c.stream, c.f[F_KX], c.kx32, (long)n);
    hipLaunchKernelGGL(k_to_f32, dim3(blocks), dim3(256), 0, c.stream, c.f[F_KY], c.ky32, (long)n);
    TL_HIP(hipGetLastError());
    float h[2 * MX_MAX_LEVELS] = {};
    for (int k = 0; k < inner_steps; ++k) {
        h[k] = (float)alphas[k];
        h[inner_steps + k] = (float)betas[k];
    }
    TL_HIP(hipMemcpyAsync(c.d_coef32, h, sizeof(float) * 2 * inner_steps, hipMemcpyHostToDevice, c.stream));
    TL_HIP(hipStreamSynchronize(c.stream));  // h is a stack array
    c.mixed_steps = inner_steps;
    c.inv_theta = 1.0 / theta;
}

// One mixed-precision PPCG outer iteration after CGCalcW (alpha on the
// device): u and r updated in place of CGCalcUr .. Calculate2Norm, the rrn
// reduction's first phase left in Chunk::partials (returns its block count).
long op_ppcg_outer_mixed(Chunk& c, int inner_steps)
{
    if (c.mixed_steps != inner_steps || !c.kx32) throw Error("mixed-precision PPCG used before op_mixed_prepare");
    MXArgs a{};
    a.u = c.f[F_U];
    a.p = c.f[F_P];
    a.r_in = c.f[F_R];
    a.w = c.f[F_W];
    a.r_out = c.r_alt;
    a.kx = c.kx32;
    a.ky = c.ky32;
    a.ds = c.ds;
    a.inv_theta = c.inv_theta;
    a.coef = c.d_coef32;
    a.partials = c.partials;
    const int L = inner_steps;
    const int ow = 2 * WAVE - 2 * L;
    a.tiles_x = (c.g.nx + ow - 1) / ow;
    a.rows = tb_rows(c.g.ny, a.tiles_x, L, MX_WAVES_PER_SIMD * c.simds);
    a.nblocks = (long)a.tiles_x * ((c.g.ny + a.rows - 1) / a.rows);
    if (a.nblocks > c.partials_cap) throw Error("mixed-precision PPCG: more waves than reduction partials");
    switch (L) {
#define TL_MX_CASE(n) \
    case n: launch<n>(c, a); break;
        TL_MX_CASE(1) TL_MX_CASE(3) TL_MX_CASE(10)
#undef TL_MX_CASE
    default: throw Error("mixed-precision PPCG pass needs 1, 3 or 10 levels");
    }
    std::swap(c.f[F_R], c.r_alt);
    return a.nblocks;
}

#else  // the 7-point stencil has no mixed-precision form

bool ppcg_mixed_supported(const Chunk&, int) { return false; }
void op_mixed_prepare(Chunk&, const double*, const double*, int, double)
{
    throw Error("mixed-precision PPCG is implemented for 2D");
}
long op_ppcg_outer_mixed(Chunk&, int) { throw Error("mixed-precision PPCG is implemented for 2D"); }

#endif

}  // namespace tl
