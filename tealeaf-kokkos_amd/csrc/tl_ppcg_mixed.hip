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
constexpr int VMCNT0 = 0x0F70;  // s_waitcnt immediate: vmcnt(0), expcnt and lgkmcnt unwaited
#ifndef TL_MX_PF
#define TL_MX_PF 3
#endif
constexpr int MX_PF = TL_MX_PF;  // rounds between a row's loads and their use (divides 6)

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
    const float* kxy;  // FP32 kx, ky interleaved: element i of the FP64 fields at [2i], [2i + 1]
    const DevScalars* ds;
    double inv_theta;
    const float* coef;  // a[0..L) then b[0..L)
    Reduce red;  // closes the rrn reduction
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
    float sdw[L][2][2];   // sd of level s, slot (round) % 2, column
    float dnw[L + 1][2];  // level s: sd(z) - sd(z - 1) of level s - 1 at its last row z
    float rw[L][2][2];    // r32 of level s, slot (round) % 2
    float duw[L][2][2];   // du after level s (s >= 1), slot (round) % 2
    double pin[MX_PF][2][2];  // prefetched level-0 rows: r, w
    float pk[MX_PF][2][2];    // prefetched level-0 rows: kx, ky
    double po[MX_PF][2][2];   // prefetched output rows: u, p
    double acc;
    float ca[L], cb[L];   // step coefficients (wave-uniform)
};

struct MXCtx {
    Geom g;
    int nx, ny, ylo, Y0, yhi;
    unsigned bs;    // store target of the lane's column pair (a pad pair for lanes without output)
    unsigned bpad;  // a pad pair past the halo that nothing reads (rows without output)
    unsigned bp;    // the lane's column pair, clamped into the mesh
    unsigned bop;   // u, p: lanes without output read an output pair
    bool kw0, kw1;  // kx at c0, c1 is not a boundary face
    bool out0, out1;
    double alpha, inv_theta;
    float4* kr;    // LDS ring [16][64]: kx at c0, c1, ky at c0, c1
    double2* rpr;  // LDS ring [16][64]: r' at c0, c1
};

template <class T>
__device__ __forceinline__ T ldb(const T* row, unsigned boff)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(row) + boff);
}

__device__ __forceinline__ void stb(double* row, unsigned boff, double2 v)
{
    *reinterpret_cast<double2*>(reinterpret_cast<char*>(row) + boff) = v;
}

// Ring slot of a row: 16 rows, more than the L + 1 a wave keeps.
constexpr int NRING = 16;
__device__ __forceinline__ int ring(int row) { return row & (NRING - 1); }

// Prefetch for round y: level-0 inputs of row y and the output row y - L,
// issued MX_PF rounds ahead.  Every load is issued on every round, so the count of loads in
// flight never depends on a branch.  Rows are not clamped: a strip reaches
// at most 2L rows below and L + 5 + MX_PF rows above the mesh, which fall in the
// halo and the zeroed slack rows of the allocation.  Each load is the
// array's base plus a per-lane 32-bit byte offset (the host keeps every
// field under 4 GiB).
template <int L>
__device__ __forceinline__ void mx_load(const MXArgs& A, const MXCtx& c, int y, int slot, MXState<L>& st)
{
    const unsigned oyr = (unsigned)(c.g.base + (long)y * c.g.sx);
    const unsigned oyo = oyr - (unsigned)(L * c.g.sx);
    // a lane's two columns are one 16-byte (8-byte for floats) load
    const double2 r = ldb(reinterpret_cast<const double2*>(A.r_in), 8u * oyr + c.bp);
    const double2 w = ldb(reinterpret_cast<const double2*>(A.w), 8u * oyr + c.bp);
    const float4 k = ldb(reinterpret_cast<const float4*>(A.kxy), 8u * oyr + c.bp);  // kx0 ky0 kx1 ky1
    const double2 u = ldb(reinterpret_cast<const double2*>(A.u), 8u * oyo + c.bop);
    const double2 p = ldb(reinterpret_cast<const double2*>(A.p), 8u * oyo + c.bop);
    st.pin[slot][0][0] = r.x;
    st.pin[slot][0][1] = r.y;
    st.pin[slot][1][0] = w.x;
    st.pin[slot][1][1] = w.y;
    st.pk[slot][0][0] = k.x;
    st.pk[slot][0][1] = k.z;
    st.pk[slot][1][0] = k.y;
    st.pk[slot][1][1] = k.w;
    st.po[slot][0][0] = u.x;
    st.po[slot][0][1] = u.y;
    st.po[slot][1][0] = p.x;
    st.po[slot][1][1] = p.y;
}

// One round: level 0 at row y, level s at row y - s (s = 1..L); level L's
// row is the output row.
//
// The reflective faces are folded into the coefficients: a mirrored
// neighbour equals the cell, so its term k_f (sd_f - sd) is zero, and the
// rings hold k_f = 0 for the faces on the mesh boundary (west of column 2,
// east of column nx + 1, south of row 2, north of row ny + 1).  Every level
// then runs the same arithmetic on every row and column, with the same
// result as substituting the mirrored value; rows and columns outside the
// mesh compute finite values (halo and slack rows hold finite values,
// columns are clamped, the rings start zeroed) that reach no cell of the
// mesh.
template <int L, int PH>
__device__ __forceinline__ void mx_round(const MXArgs& A, const MXCtx& c, MXState<L>& st, int y)
{
    constexpr int R0 = PH % 2, R1 = (PH + 1) % 2;  // rounds y, y-1
    constexpr int P = PH % MX_PF;                  // prefetch slot of this round
    const int lane = threadIdx.x;

    float4 nk = c.kr[ring(y - 1) * WAVE + lane];

    // level 0 at row y: r' = r - alpha w in FP64, then its FP32 rounding
    const int ry = ring(y);
    float kyn0, kyn1;
    {
        double rp[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            rp[q] = st.pin[P][0][q] - c.alpha * st.pin[P][1][q];
            st.rw[0][R0][q] = (float)rp[q];
            st.sdw[0][R0][q] = (float)(rp[q] * c.inv_theta);
        }
        c.rpr[ry * WAVE + lane] = make_double2(rp[0], rp[1]);
        kyn0 = st.pk[P][1][0];
        kyn1 = st.pk[P][1][1];
        if (y == HALO || y > c.ny + 1) kyn0 = kyn1 = 0.0f;
        c.kr[ry * WAVE + lane] = make_float4(c.kw0 ? st.pk[P][0][0] : 0.0f, c.kw1 ? st.pk[P][0][1] : 0.0f, kyn0, kyn1);
    }
    // the output row of this round: u, p loaded MX_PF rounds ago, r' from
    // the ring
    double po[2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        po[k][0] = st.po[P][k][0];
        po[k][1] = st.po[P][k][1];
    }
    const double2 rpo = c.rpr[ring(y - L) * WAVE + lane];
    mx_load<L>(A, c, y + MX_PF, P, st);

#pragma unroll
    for (int s = 1; s <= L; ++s) {
        const int z = y - s;
        const float4 k = nk;  // kx at c0, c1 and ky at c0, c1 of row z
        const float a = st.ca[s - 1], b = st.cb[s - 1];
        if (s < L) nk = c.kr[ring(z - 1) * WAVE + lane];
        // level s - 1 at rows z (c) and z + 1 (n); its row z - 1 enters
        // through the difference carried from the previous round.  The
        // differences are the ones the direct form takes (sd_f - sd), so
        // the sums below round exactly as they would:
        //   e0 - c0 = c1 - c0,     w1 - c1 = -(c1 - c0),
        //   e1 - c1 = c0(+1) - c1, w0 - c0 = -(e1 - c1 of the lane to the west),
        //   s - c = -(c - s), the north difference of row z - 1.
        const float c0 = st.sdw[s - 1][R1][0], c1 = st.sdw[s - 1][R1][1];
        const float n0 = st.sdw[s - 1][R0][0], n1 = st.sdw[s - 1][R0][1];
        const float d01 = c1 - c0;
        const float de1 = f_from_east(c0) - c1;
        const float dw0 = f_from_west(de1);
        const float dn0 = n0 - c0, dn1 = n1 - c1;
        const float kx2 = f_from_east(k.x);  // kx at c1 + 1
        // r - A sd = (r - sd) + sum_f k_f (sd_f - sd), the north face
        // last: only that term needs the row the level below produced in
        // this round, so the rest of every level is off the chain that
        // runs through the L levels of a round
        float t0 = st.rw[s - 1][R1][0] - c0;
        t0 = __builtin_fmaf(k.y, d01, t0);
        t0 = __builtin_fmaf(k.x, -dw0, t0);
        t0 = __builtin_fmaf(k.z, -st.dnw[s][0], t0);
        t0 = __builtin_fmaf(kyn0, dn0, t0);
        float t1 = st.rw[s - 1][R1][1] - c1;
        t1 = __builtin_fmaf(kx2, de1, t1);
        t1 = __builtin_fmaf(k.y, -d01, t1);
        t1 = __builtin_fmaf(k.w, -st.dnw[s][1], t1);
        t1 = __builtin_fmaf(kyn1, dn1, t1);
        st.dnw[s][0] = dn0;
        st.dnw[s][1] = dn1;
        kyn0 = k.z;
        kyn1 = k.w;
        const float d0 = s == 1 ? c0 : st.duw[s - 1][R1][0] + c0;
        const float d1 = s == 1 ? c1 : st.duw[s - 1][R1][1] + c1;
        if (s < L) {
            st.sdw[s][R0][0] = __builtin_fmaf(a, c0, b * t0);
            st.sdw[s][R0][1] = __builtin_fmaf(a, c1, b * t1);
            st.rw[s][R0][0] = t0;
            st.rw[s][R0][1] = t1;
            st.duw[s][R0][0] = d0;
            st.duw[s][R0][1] = d1;
        } else {
            // u_out = u' + du, r_out = r' + (r32_L - r32_0), in FP64; every
            // lane stores, lanes and rows that own no cell into a pad column
            const bool zin = z >= c.Y0 && z <= c.yhi;
            const bool o0 = zin && c.out0, o1 = zin && c.out1;
            const unsigned bz = zin ? c.bs : c.bpad;
            const unsigned oz = 8u * (unsigned)(c.g.base + (long)z * c.g.sx);
            const double rp0 = rpo.x, rp1 = rpo.y;
            const double up0 = po[0][0] + c.alpha * po[1][0], up1 = po[0][1] + c.alpha * po[1][1];
            const double rr0 = rp0 + (double)(t0 - (float)rp0);
            const double rr1 = rp1 + (double)(t1 - (float)rp1);
            stb(A.r_out, oz + bz, make_double2(rr0, rr1));
            stb(A.u, oz + bz, make_double2(up0 + (double)d0, up1 + (double)d1));
            st.acc += o0 ? rr0 * rr0 : 0.0;
            st.acc += o1 ? rr1 * rr1 : 0.0;
        }
    }
}

template <int L>
__device__ __forceinline__ void mx_wave(const MXArgs& A, const MXCtx& c)
{
    MXState<L> st;
#pragma unroll
    for (int s = 0; s < L; ++s) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            st.sdw[s][0][q] = st.sdw[s][1][q] = 0.0f;
            st.dnw[s + 1][q] = 0.0f;
            st.rw[s][0][q] = st.rw[s][1][q] = 0.0f;
            st.duw[s][0][q] = st.duw[s][1][q] = 0.0f;
        }
    }
    st.acc = 0.0;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        st.ca[s] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(A.coef[s])));
        st.cb[s] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(A.coef[L + s])));
    }
#pragma unroll
    for (int k = 0; k < MX_PF; ++k) mx_load<L>(A, c, c.ylo + k, k, st);
    // wait for the first rows here: with loads still in flight on entry
    // the compiler's wait at the loop head has to cover the entry path too,
    // and it then drains all loads at the head of every block of rounds
    __builtin_amdgcn_s_waitcnt(VMCNT0);
    const int yend = c.yhi + L;
    // the scheduler may not move work between rounds: interleaving them
    // raises the register count past 256, and the copies it then makes of
    // registers with loads in flight wait for those loads
    for (int y = c.ylo; y <= yend; y += 6) {
        mx_round<L, 0>(A, c, st, y);
        __builtin_amdgcn_sched_barrier(0);
        mx_round<L, 1>(A, c, st, y + 1);
        __builtin_amdgcn_sched_barrier(0);
        mx_round<L, 2>(A, c, st, y + 2);
        __builtin_amdgcn_sched_barrier(0);
        mx_round<L, 3>(A, c, st, y + 3);
        __builtin_amdgcn_sched_barrier(0);
        mx_round<L, 4>(A, c, st, y + 4);
        __builtin_amdgcn_sched_barrier(0);
        mx_round<L, 5>(A, c, st, y + 5);
        __builtin_amdgcn_sched_barrier(0);
    }
    double v[1] = {st.acc};
    block_partial<1, 1>(v, A.red);
}

template <int L>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(TL_MX_WAVES_PER_SIMD))) void k_ppcg_mixed(
    Geom g, MXArgs A)
{
    static_assert(L + 1 <= NRING, "ring too short");
    __shared__ float4 k_ring[NRING * WAVE];
    __shared__ double2 rp_ring[NRING * WAVE];
    const int lane = threadIdx.x;
    // rows below the strip read ring slots before any row is written there
    for (int k = 0; k < NRING; ++k) k_ring[k * WAVE + lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    __syncthreads();
    const long t = xcd_tile(blockIdx.x, A.nblocks);
    const int tx = (int)(t % A.tiles_x), ty = (int)(t / A.tiles_x);
    // the wave's first column is even, so a lane's column pair is 16-byte
    // aligned: LP = L rounded up to even columns of margin on each side
    constexpr int LP = L + (L & 1);
    constexpr int OW = 2 * WAVE - 2 * LP;
    MXCtx c;
    c.g = g;
    c.nx = g.nx;
    c.ny = g.ny;
    const int X0 = HALO + tx * OW;
    c.Y0 = HALO + ty * A.rows;
    const int k0 = X0 - LP + 2 * lane, k1 = k0 + 1;
    {
        // pairs start on even columns: the last one may hold the halo
        // column nx + 2, which no cell of the mesh reads (its face is zero)
        const int kp = min(max(k0, HALO), (c.nx + 1) & ~1);
        c.bp = 8u * (unsigned)kp;
    }
    // faces HALO (west of column 2) and nx + 2 (east of column nx + 1)
    c.kw0 = k0 != HALO && k0 != c.nx + 2;
    c.kw1 = k1 != HALO && k1 != c.nx + 2;
    c.out0 = k0 >= X0 && k0 < X0 + OW && k0 <= c.nx + 1;
    c.out1 = k1 >= X0 && k1 < X0 + OW && k1 <= c.nx + 1;
    {
        // u and p of lanes without output come from the same cache lines
        const int xlast = min(X0 + OW - 1, c.nx + 1);
        c.bop = 8u * (unsigned)min(max(k0, X0), xlast & ~1);
    }
    {
        // pad columns X .. sx-1 of every row are never read (at least 30
        // of them); a lane whose first column is the last cell of an odd
        // nx also writes the halo column nx + 2 past it, which no kernel
        // reads either (faces are read as mirrored values)
        const int xe = (int)g.X + ((int)g.X & 1);
        const int npair = (int)(g.sx - xe) / 2;
        c.bpad = 8u * (unsigned)(xe + 2 * (lane % npair));
        c.bs = (c.out0 || c.out1) ? 8u * (unsigned)k0 : c.bpad;
    }
    c.ylo = c.Y0 - L;
    c.yhi = min(c.Y0 + A.rows - 1, c.ny + 1);
    c.alpha = A.ds->alpha;
    c.inv_theta = A.inv_theta;
    c.kr = k_ring;
    c.rpr = rp_ring;
    mx_wave<L>(A, c);
}

template <int L>
void launch(Chunk& c, MXArgs& a)
{
    hipLaunchKernelGGL((k_ppcg_mixed<L>), dim3((unsigned)a.nblocks), dim3(WAVE), 0, c.stream, c.g, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string("kernel launch: ") + hipGetErrorString(e));
}

__global__ void k_to_f32(const double* __restrict__ kx, const double* __restrict__ ky, float* __restrict__ kxy,
                         long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        reinterpret_cast<float2*>(kxy)[i] = make_float2((float)kx[i], (float)ky[i]);
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
    if (!c.kxy32) TL_HIP(hipMalloc(&c.kxy32, sizeof(float) * 2 * n));
    if (!c.d_coef32) TL_HIP(hipMalloc(&c.d_coef32, sizeof(float) * 2 * MX_MAX_LEVELS));
    const unsigned blocks = (unsigned)std::min<long>(((long)n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_to_f32, dim3(blocks), dim3(256), 0, c.stream, c.f[F_KX], c.f[F_KY], c.kxy32, (long)n);
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
// device): u and r updated in place of CGCalcUr .. Calculate2Norm, and the
// rrn reduction closed (beta, error and rro on the device, rrn in S_RRN).
void op_ppcg_outer_mixed(Chunk& c, int inner_steps)
{
    if (c.mixed_steps != inner_steps || !c.kxy32) throw Error("mixed-precision PPCG used before op_mixed_prepare");
    MXArgs a{};
    a.u = c.f[F_U];
    a.p = c.f[F_P];
    a.r_in = c.f[F_R];
    a.w = c.f[F_W];
    a.r_out = c.r_alt;
    a.kxy = c.kxy32;
    a.ds = c.ds;
    a.inv_theta = c.inv_theta;
    a.coef = c.d_coef32;
    a.red = c.reducer(S_RRN, EP_RRN);
    const int L = inner_steps;
    const int ow = 2 * WAVE - 2 * (L + (L & 1));
    a.tiles_x = (c.g.nx + ow - 1) / ow;
    a.rows = tb_rows(c.g.ny, a.tiles_x, L, MX_WAVES_PER_SIMD * c.simds);
    a.nblocks = (long)a.tiles_x * ((c.g.ny + a.rows - 1) / a.rows);
    if (a.nblocks > c.partials_cap) throw Error("mixed-precision PPCG: more waves than reduction partials");
    if (L != 1 && L != 3 && L != 10) throw Error("mixed-precision PPCG pass needs 1, 3 or 10 levels");
    // the timer brackets the pass alone, not the finish of its reduction
    TL_TIMED(c, K_PPCG_MIXED, {
        switch (L) {
        case 1: launch<1>(c, a); break;
        case 3: launch<3>(c, a); break;
        default: launch<10>(c, a); break;
        }
    });
    c.close(a.nblocks, 1);
    std::swap(c.f[F_R], c.r_alt);
}

#else  // the 7-point stencil has no mixed-precision form

bool ppcg_mixed_supported(const Chunk&, int) { return false; }
void op_mixed_prepare(Chunk&, const double*, const double*, int, double)
{
    throw Error("mixed-precision PPCG is implemented for 2D");
}
void op_ppcg_outer_mixed(Chunk&, int) { throw Error("mixed-precision PPCG is implemented for 2D"); }

#endif

}  // namespace tl
