// The learned preconditioner: a 4-level convolutional U-Net (3x3 convs,
// 16/32/64/128 channels) in BF16 with FP32 accumulation, trained on the GPU
// by the kernels below, and applied by flexible CG (tl_driver.hip).
//
// Layout: activations are NHWC bf16 (channels fastest, padded to a multiple
// of 8), weights FP32 masters [Cout][9][Cin] with bf16 copies in the two
// GEMM layouts the kernels read:
//   wf[co][t*Cin + ci]          forward: out = sum over (tap, ci)
//   wb[ci][t*Cout + co]         backward-data, taps mirrored (t -> 8 - t)
// each row zero-padded to a multiple of 32 (the MFMA K).
//
// Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16
// (gfx950): A = weights (16 output channels x 32 k), B = input (32 k x 16
// pixels), D = 16 channels x 16 pixels in FP32.  A lane loads its 8
// consecutive k of one tap as one 16-byte read (channels are contiguous),
// and stores 4 consecutive output channels of one pixel.  The weight
// gradient is the same instruction with the pixels as K.
#include "tl_chunk.hpp"
#include "tl_unet.hpp"

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace tl {

typedef unsigned short u16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

__host__ __device__ inline u16 f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (u16)(u >> 16);
}

__host__ __device__ inline float bf2f(u16 h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// 16-channel column tiles per wave: 4 when the channel count allows it
inline int col_tiles(int channels) { return channels % 64 == 0 ? 4 : channels % 32 == 0 ? 2 : 1; }

__device__ inline uint4 ld16(const u16* p) { return *reinterpret_cast<const uint4*>(p); }

__device__ inline bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// zero the bf16 lanes of v whose mask value is not positive (ReLU')
__device__ inline uint4 relu_mask(uint4 v, uint4 m)
{
    uint32_t* pv = reinterpret_cast<uint32_t*>(&v);
    const uint32_t* pm = reinterpret_cast<const uint32_t*>(&m);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t lo = pm[q] & 0xffffu, hi = pm[q] >> 16;
        const bool plo = lo != 0 && !(lo & 0x8000u), phi = hi != 0 && !(hi & 0x8000u);
        pv[q] &= (plo ? 0x0000ffffu : 0u) | (phi ? 0xffff0000u : 0u);
    }
    return v;
}

struct ConvArgs {
    const u16* x;     // [n][h][w][cin]
    const u16* m;     // ReLU mask source, same shape as x (or null)
    const u16* wt;    // [cout][kp]
    const float* bias;
    u16* y;           // [n][h][w][cout]
    int n, h, w, cin, cout, kp;
    int xtiles, cotiles;
    long waves;
};

// One wave: 64 pixels of one row x 16*NT output channels.
template <int NT, bool RELU, bool MASK>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a)
{
    const long wid = (long)blockIdx.x * 4 + threadIdx.x / WAVE;
    if (wid >= a.waves) return;
    const int lane = threadIdx.x % WAVE;
    const int j = lane % 16, kq = lane / 16;
    const int cot = (int)(wid % a.cotiles);
    long rest = wid / a.cotiles;
    const int xt = (int)(rest % a.xtiles);
    const long row = rest / a.xtiles;
    const int n = (int)(row / a.h), y = (int)(row % a.h);
    const int px0 = xt * 64, co0 = cot * 16 * NT;

    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int kreal = 9 * a.cin;
    for (int kc = 0; kc < a.kp; kc += 32) {
        const int k = kc + kq * 8;
        const int tap = k / a.cin, ci = k - tap * a.cin;
        const bool kin = k < kreal;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        bf16x8 av[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) av[nt] = as_bf8(ld16(a.wt + (long)(co0 + 16 * nt + j) * a.kp + k));
        const int yy = y + dy;
        const bool yin = kin && yy >= 0 && yy < a.h;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int xx = px0 + 16 * mt + j + dx;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (yin && xx >= 0 && xx < a.w) {
                const long off = (((long)n * a.h + yy) * a.w + xx) * a.cin + ci;
                v = ld16(a.x + off);
                if (MASK) v = relu_mask(v, ld16(a.m + off));
            }
            const bf16x8 bv = as_bf8(v);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[nt], bv, acc[mt][nt], 0, 0, 0);
        }
    }
    // D[co = 4 kq + v][px = j] of every (mt, nt) tile
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int px = px0 + 16 * mt + j;
        if (px >= a.w) continue;
        u16* out = a.y + (((long)n * a.h + y) * a.w + px) * a.cout;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = co0 + 16 * nt + 4 * kq;
            u16 h4[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float s = acc[mt][nt][v] + (a.bias ? a.bias[co + v] : 0.0f);
                if (RELU) s = s > 0.0f ? s : 0.0f;
                h4[v] = f2bf(s);
            }
            uint2 pk;
            pk.x = (uint32_t)h4[0] | ((uint32_t)h4[1] << 16);
            pk.y = (uint32_t)h4[2] | ((uint32_t)h4[3] << 16);
            *reinterpret_cast<uint2*>(out + co) = pk;
        }
    }
}

// The same convolution with the input staged in LDS: a workgroup copies the
// (TR + 2) input rows x 66 pixels x cin channels that its TR output rows x
// 64 pixels read (zero outside the image, ReLU' applied for MASK) into LDS
// once.  Its 4 waves split the TR rows and, 4 / TR ways, the cotiles: a
// wave computes a whole 64-pixel row for its cotiles, so each weight it
// loads serves 64 pixels, with B read from LDS and the next K chunk's
// weights loaded while the current one's MFMAs run.  Every accumulator sees
// the same MFMA sequence as in k_conv3x3, so the output is bitwise the same;
// the input is read from L2 (TR + 2) / TR times instead of 9 times per
// cotile.  A staged pixel holds cin + 8 channels, so that the 16-byte reads
// of 8 consecutive lanes (8 consecutive pixels) fall in distinct banks.
template <int NT, int TR, bool RELU, bool MASK>
__global__ __launch_bounds__(256) void k_conv3x3_lds(ConvArgs a)
{
    extern __shared__ uint4 stage[];
    const int pu = a.cin / 8 + 1;  // 16-byte units per staged pixel
    const int xt = (int)(blockIdx.x % a.xtiles);
    const long rt = blockIdx.x / a.xtiles;
    const int rtiles = (a.h + TR - 1) / TR;
    const int n = (int)(rt / rtiles), y0 = (int)(rt % rtiles) * TR;
    const int px0 = xt * 64;
    const int c8 = a.cin / 8;
    const int units = (TR + 2) * 66 * c8;
    // SB loads of 16 bytes in flight per thread before any is stored
    constexpr int SB = 8;
    for (int u0 = 0; u0 < units; u0 += 256 * SB) {
        uint4 v[SB], mv[SB];
        int at[SB];
#pragma unroll
        for (int q = 0; q < SB; ++q) {
            const int u = u0 + (int)threadIdx.x + 256 * q;
            v[q] = mv[q] = make_uint4(0u, 0u, 0u, 0u);
            at[q] = -1;
            if (u < units) {
                const int c = u % c8, pq = u / c8;
                const int sx = pq % 66, sr = pq / 66;
                const int yy = y0 - 1 + sr, xx = px0 - 1 + sx;
                at[q] = (sr * 66 + sx) * pu + c;
                if (yy >= 0 && yy < a.h && xx >= 0 && xx < a.w) {
                    const long off = (((long)n * a.h + yy) * a.w + xx) * a.cin + 8 * c;
                    v[q] = ld16(a.x + off);
                    if (MASK) mv[q] = ld16(a.m + off);
                }
            }
        }
#pragma unroll
        for (int q = 0; q < SB; ++q)
            if (at[q] >= 0) stage[at[q]] = MASK ? relu_mask(v[q], mv[q]) : v[q];
    }
    __syncthreads();
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int j = lane % 16, kq = lane / 16;
    constexpr int G = 4 / TR;  // waves per output row, each taking every G-th cotile
    const int r = wave / G;
    const int y = y0 + r;
    if (y >= a.h) return;  // no barrier follows
    for (int cot = wave % G; cot < a.cotiles; cot += G) {
        const int co0 = cot * 16 * NT;
        const u16* wrow = a.wt + (long)(co0 + j) * a.kp + kq * 8;
        f32x4 acc[4][NT];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
        uint4 wv[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) wv[nt] = ld16(wrow + (long)16 * nt * a.kp);
        // this lane's k = kc + 8 kq as (tap, ci), advanced by 32 per chunk
        int ci = kq * 8, tap = 0;
        while (ci >= a.cin) {
            ci -= a.cin;
            ++tap;
        }
        int tdy = tap / 3 - 1, tdx = tap % 3 - 1;
        for (int kc = 0; kc < a.kp; kc += 32) {
            bf16x8 av[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) av[nt] = as_bf8(wv[nt]);
            if (kc + 32 < a.kp) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) wv[nt] = ld16(wrow + (long)16 * nt * a.kp + kc + 32);
            }
            const bool kin = tap < 9;  // k < 9 cin: the padded k carry zero weights and inputs
            const uint4* srow = stage + ((r + 1 + (kin ? tdy : 0)) * 66 + 1 + tdx + j) * pu + (ci >> 3);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                const uint4 v = kin ? srow[16 * mt * pu] : make_uint4(0u, 0u, 0u, 0u);
                const bf16x8 bv = as_bf8(v);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[nt], bv, acc[mt][nt], 0, 0, 0);
            }
            ci += 32;
            while (ci >= a.cin) {
                ci -= a.cin;
                ++tap;
                if (++tdx > 1) {
                    tdx = -1;
                    ++tdy;
                }
            }
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int px = px0 + 16 * mt + j;
            if (px >= a.w) continue;
            u16* out = a.y + (((long)n * a.h + y) * a.w + px) * a.cout;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = co0 + 16 * nt + 4 * kq;
                u16 h4[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float sv = acc[mt][nt][v] + (a.bias ? a.bias[co + v] : 0.0f);
                    if (RELU) sv = sv > 0.0f ? sv : 0.0f;
                    h4[v] = f2bf(sv);
                }
                uint2 pk;
                pk.x = (uint32_t)h4[0] | ((uint32_t)h4[1] << 16);
                pk.y = (uint32_t)h4[2] | ((uint32_t)h4[3] << 16);
                *reinterpret_cast<uint2*>(out + co) = pk;
            }
        }
    }
}

struct WGradArgs {
    const u16* x;   // layer input [n][h][w][cin]
    const u16* dy;  // gradient at the layer output [n][h][w][cout]
    const u16* ym;  // layer output (ReLU mask) or null
    float* dw;      // [cout][9][cin], accumulated
    int n, h, w, cin, cout;
    int cotiles, citiles;
    long pixels;    // n*h*w
    int chunk;      // pixels per wave (multiple of 32)
    long waves;
};

// dW[co][t][ci] += sum over a pixel chunk of dY'[p][co] x[p + off(t)][ci],
// dY' = dY masked by ReLU'.  One wave: 16 co x 16 ci of one tap.
template <bool MASK>
__global__ __launch_bounds__(256) void k_conv_wgrad(WGradArgs a)
{
    const long wid = (long)blockIdx.x * 4 + threadIdx.x / WAVE;
    if (wid >= a.waves) return;
    const int lane = threadIdx.x % WAVE;
    const int j = lane % 16, kq = lane / 16;
    long r = wid;
    const int cit = (int)(r % a.citiles);
    r /= a.citiles;
    const int cot = (int)(r % a.cotiles);
    r /= a.cotiles;
    const int t = (int)(r % 9);
    const long chunk_id = r / 9;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int co = cot * 16 + j;
    const int ci = cit * 16 + j;
    const bool ciok = ci < a.cin;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const long p_begin = chunk_id * a.chunk;
    const long p_end = p_begin + a.chunk < a.pixels ? p_begin + a.chunk : a.pixels;
    for (long pb = p_begin; pb < p_end; pb += 32) {
        uint32_t ap[4], bp[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            u16 av[2], bv[2];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const long p = pb + kq * 8 + 2 * q + e;
                u16 gval = 0, xval = 0;
                if (p < p_end) {
                    const long o = p * a.cout + co;
                    gval = a.dy[o];
                    if (MASK) {
                        const u16 mv = a.ym[o];
                        if (mv == 0 || (mv & 0x8000u)) gval = 0;
                    }
                    const int x = (int)(p % a.w);
                    const long ny = p / a.w;
                    const int y = (int)(ny % a.h);
                    const int yy = y + dy, xx = x + dx;
                    if (ciok && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w) xval = a.x[(p + (long)dy * a.w + dx) * a.cin + ci];
                }
                av[e] = gval;
                bv[e] = xval;
            }
            ap[q] = (uint32_t)av[0] | ((uint32_t)av[1] << 16);
            bp[q] = (uint32_t)bv[0] | ((uint32_t)bv[1] << 16);
        }
        const bf16x8 A = as_bf8(make_uint4(ap[0], ap[1], ap[2], ap[3]));
        const bf16x8 B = as_bf8(make_uint4(bp[0], bp[1], bp[2], bp[3]));
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
    }
    // D[co' = 4 kq + v][ci' = j]
    if (!ciok) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int cw = cot * 16 + 4 * kq + v;
        atomicAdd(a.dw + ((long)cw * 9 + t) * a.cin + ci, acc[v]);
    }
}

// db[co] += sum_p dY'[p][co]
template <bool MASK>
__global__ void k_bias_grad(const u16* dy, const u16* ym, float* db, long pixels, int cout)
{
    const int co = threadIdx.x % cout;
    const int lanes = blockDim.x / cout;
    float s = 0.0f;
    for (long p = (long)blockIdx.x * lanes + threadIdx.x / cout; p < pixels; p += (long)gridDim.x * lanes) {
        const u16 g = dy[p * cout + co];
        if (MASK) {
            const u16 mv = ym[p * cout + co];
            if (mv == 0 || (mv & 0x8000u)) continue;
        }
        s += bf2f(g);
    }
    atomicAdd(db + co, s);
}

__global__ void k_pack_weights(const float* w, u16* wf, u16* wb, int cin, int cout, int kp, int kpb)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)cout * 9 * cin;
    if (i >= total) return;
    const int ci = (int)(i % cin);
    const int t = (int)((i / cin) % 9);
    const int co = (int)(i / (9L * cin));
    const u16 h = f2bf(w[i]);
    wf[(long)co * kp + t * cin + ci] = h;
    wb[(long)ci * kpb + (8 - t) * cout + co] = h;
}

__global__ void k_avgpool(const u16* x, u16* y, int n, int h, int w, int c)
{
    const int ho = h / 2, wo = w / 2;
    const long total = (long)n * ho * wo * c;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const long q = i / c;
        const int xo = (int)(q % wo);
        const long r = q / wo;
        const int yo = (int)(r % ho), b = (int)(r / ho);
        const long base = (((long)b * h + 2 * yo) * w + 2 * xo) * c + ch;
        const float s = bf2f(x[base]) + bf2f(x[base + c]) + bf2f(x[base + (long)w * c]) + bf2f(x[base + (long)w * c + c]);
        y[i] = f2bf(0.25f * s);
    }
}

// dX (+)= the 2x2 spread of dY / 4
__global__ void k_avgpool_bwd(const u16* dy, u16* dx, int n, int h, int w, int c, int accumulate)
{
    const long total = (long)n * h * w * c;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const long q = i / c;
        const int x = (int)(q % w);
        const long r = q / w;
        const int y = (int)(r % h), b = (int)(r / h);
        const float g = 0.25f * bf2f(dy[(((long)b * (h / 2) + y / 2) * (w / 2) + x / 2) * c + ch]);
        dx[i] = f2bf(accumulate ? bf2f(dx[i]) + g : g);
    }
}

// cat[..][0:ca] = nearest 2x upsample of a (h/2 x w/2), cat[..][ca:ca+cb] = skip
// 8 channels (16 bytes) per thread: ca and cb are multiples of 8
__global__ void k_up_concat(const u16* a, const u16* skip, u16* cat, int n, int h, int w, int ca, int cb)
{
    const int c = ca + cb, c8 = c / 8, ca8 = ca / 8;
    const long total = (long)n * h * w * c8;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int j = (int)(i % c8);
        const long q = i / c8;  // pixel
        const int x = (int)(q % w);
        const long r = q / w;
        const int y = (int)(r % h), b = (int)(r / h);
        const uint4 v = j < ca8 ? ld16(a + (((long)b * (h / 2) + y / 2) * (w / 2) + x / 2) * ca + 8 * j)
                                : ld16(skip + q * cb + 8 * (j - ca8));
        *reinterpret_cast<uint4*>(cat + q * c + 8 * j) = v;
    }
}

// backward of k_up_concat: da = 2x2 sum of the first ca channels, dskip = the rest
__global__ void k_up_concat_bwd(const u16* dcat, u16* da, u16* dskip, int n, int h, int w, int ca, int cb)
{
    const int c = ca + cb, ca8 = ca / 8, cb8 = cb / 8;
    const int ho = h / 2, wo = w / 2;
    const long ta = (long)n * ho * wo * ca8;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ta; i += (long)gridDim.x * blockDim.x) {
        const int j = (int)(i % ca8);
        const long q = i / ca8;  // coarse pixel
        const int xo = (int)(q % wo);
        const long r = q / wo;
        const int yo = (int)(r % ho), b = (int)(r / ho);
        const long base = (((long)b * h + 2 * yo) * w + 2 * xo) * c + 8 * j;
        const uint4 v0 = ld16(dcat + base), v1 = ld16(dcat + base + c);
        const uint4 v2 = ld16(dcat + base + (long)w * c), v3 = ld16(dcat + base + (long)w * c + c);
        const uint32_t* p0 = reinterpret_cast<const uint32_t*>(&v0);
        const uint32_t* p1 = reinterpret_cast<const uint32_t*>(&v1);
        const uint32_t* p2 = reinterpret_cast<const uint32_t*>(&v2);
        const uint32_t* p3 = reinterpret_cast<const uint32_t*>(&v3);
        uint32_t o[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float lo = bf2f((u16)p0[k]) + bf2f((u16)p1[k]) + bf2f((u16)p2[k]) + bf2f((u16)p3[k]);
            const float hi = bf2f((u16)(p0[k] >> 16)) + bf2f((u16)(p1[k] >> 16)) + bf2f((u16)(p2[k] >> 16)) +
                             bf2f((u16)(p3[k] >> 16));
            o[k] = (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
        }
        *reinterpret_cast<uint4*>(da + q * ca + 8 * j) = make_uint4(o[0], o[1], o[2], o[3]);
    }
    const long ts = (long)n * h * w * cb8;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ts; i += (long)gridDim.x * blockDim.x) {
        const int j = (int)(i % cb8);
        const long q = i / cb8;
        *reinterpret_cast<uint4*>(dskip + q * cb + 8 * j) = ld16(dcat + q * c + ca + 8 * j);
    }
}

// loss = mean_p (out[p][0] - target[p])^2 over valid pixels; dout = its gradient
__global__ void k_mse(const u16* out, int cout, const float* target, u16* dout, long pixels, float* loss)
{
    float s = 0.0f;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += (long)gridDim.x * blockDim.x) {
        const float d = bf2f(out[p * cout]) - target[p];
        s += d * d;
        for (int c = 0; c < cout; ++c) dout[p * cout + c] = c == 0 ? f2bf(2.0f * d / (float)pixels) : (u16)0;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, WAVE);
    if (threadIdx.x % WAVE == 0) atomicAdd(loss, s / (float)pixels);
}

// A-norm loss over a batch of h x w samples: with d = out - t and A the
// 5-point operator of each sample's face coefficients (kx: west face of each
// cell, ky: south face; faces on the sample's edge carry no flux),
// loss = sum(d.A d) / sum(t.A t) over the batch and dout = 2 A d / sum(t.A t).
// First pass, one workgroup per sample: d, and the two sums (into sums[0..1]).
__device__ inline float a_apply(const float* v, const float* kx, const float* ky, int x, int y, int h, int w)
{
    const int i = y * w + x;
    const float c = v[i];
    float r = c;
    if (x + 1 < w) r += kx[i + 1] * (c - v[i + 1]);
    if (x > 0) r += kx[i] * (c - v[i - 1]);
    if (y + 1 < h) r += ky[i + w] * (c - v[i + w]);
    if (y > 0) r += ky[i] * (c - v[i - w]);
    return r;
}

__global__ __launch_bounds__(256) void k_anorm_sums(const u16* out, int cout, const float* target, const float* kx,
                                                    const float* ky, float* dbuf, int h, int w, float* sums)
{
    __shared__ float red[2][256 / WAVE];
    const int q = blockIdx.x;
    const long px = (long)h * w, base = (long)q * px;
    const float* t = target + base;
    const float* kxq = kx + base;
    const float* kyq = ky + base;
    float* d = dbuf + base;
    for (long i = threadIdx.x; i < px; i += blockDim.x) d[i] = bf2f(out[(base + i) * cout]) - t[i];
    __syncthreads();
    float num = 0.0f, den = 0.0f;
    for (long i = threadIdx.x; i < px; i += blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        num += d[i] * a_apply(d, kxq, kyq, x, y, h, w);
        den += t[i] * a_apply(t, kxq, kyq, x, y, h, w);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        num += __shfl_xor(num, off, WAVE);
        den += __shfl_xor(den, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0) {
        red[0][threadIdx.x / WAVE] = num;
        red[1][threadIdx.x / WAVE] = den;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        num = den = 0.0f;
        for (int k = 0; k < (int)blockDim.x / WAVE; ++k) {
            num += red[0][k];
            den += red[1][k];
        }
        atomicAdd(sums, num);
        atomicAdd(sums + 1, den);
    }
}

// Second pass: dout = 2 A d / sum(t.A t) (channel 0), loss = sum(d.A d) / sum(t.A t).
__global__ __launch_bounds__(256) void k_anorm_grad(const float* kx, const float* ky, const float* dbuf, u16* dout,
                                                    int cout, int h, int w, const float* sums, float* loss)
{
    const int q = blockIdx.x;
    const long px = (long)h * w, base = (long)q * px;
    const float inv = sums[1] > 0.0f ? 1.0f / sums[1] : 0.0f;
    for (long i = threadIdx.x; i < px; i += blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        const float g = 2.0f * a_apply(dbuf + base, kx + base, ky + base, x, y, h, w) * inv;
        for (int c = 0; c < cout; ++c) dout[(base + i) * cout + c] = c == 0 ? f2bf(g) : (u16)0;
    }
    if (q == 0 && threadIdx.x == 0) *loss = sums[0] * inv;
}

// Diagonal of a sample's operator: 1 + the coefficients of the cell's faces
// inside the sample.
__device__ inline float a_diag(const float* kx, const float* ky, int x, int y, int h, int w)
{
    const int i = y * w + x;
    float d = 1.0f;
    if (x + 1 < w) d += kx[i + 1];
    if (x > 0) d += kx[i];
    if (y + 1 < h) d += ky[i + w];
    if (y > 0) d += ky[i];
    return d;
}

// A-norm loss after the preconditioner's smoothing: the flexible CG applies
// m weighted Jacobi sweeps (weight omega) on A z = r starting from the
// network's output, which multiply the output's error d = out - t by
// M = (I - omega D^-1 A)^m.  Per sample, sum(Md.A Md) / den with
// den = max(sum(Mt.A Mt), its batch mean); the loss is the mean over the n
// samples (at most 1 for a zero output: the sweeps alone), so a sample of
// large error does not drown the others and one the sweeps already remove
// does not divide by ~0; dout = 2 M^T A M d / (n den),
// M^T = (I - omega A D^-1)^m.  buf holds four planes of the
// batch: d and its ping-pong partner, t and its partner.  First pass, one
// workgroup per sample: M d (left in plane m % 2) and the sample's two sums
// (sums[2q], sums[2q + 1]).
__global__ __launch_bounds__(256) void k_anorm_smooth_sums(const u16* out, int cout, const float* target,
                                                           const float* kx, const float* ky, float* buf, long plane,
                                                           int h, int w, int m, float omega, float* sums)
{
    __shared__ float red[2][256 / WAVE];
    const int q = blockIdx.x;
    const long px = (long)h * w, base = (long)q * px;
    const float* kxq = kx + base;
    const float* kyq = ky + base;
    float* d[2] = {buf + base, buf + plane + base};
    float* t[2] = {buf + 2 * plane + base, buf + 3 * plane + base};
    for (long i = threadIdx.x; i < px; i += blockDim.x) {
        d[0][i] = bf2f(out[(base + i) * cout]) - target[base + i];
        t[0][i] = target[base + i];
    }
    __syncthreads();
    for (int j = 0; j < m; ++j) {
        const int a = j & 1, b = a ^ 1;
        for (long i = threadIdx.x; i < px; i += blockDim.x) {
            const int x = (int)(i % w), y = (int)(i / w);
            const float s = omega / a_diag(kxq, kyq, x, y, h, w);
            d[b][i] = d[a][i] - s * a_apply(d[a], kxq, kyq, x, y, h, w);
            t[b][i] = t[a][i] - s * a_apply(t[a], kxq, kyq, x, y, h, w);
        }
        __syncthreads();
    }
    const float* dm = d[m & 1];
    const float* tm = t[m & 1];
    float num = 0.0f, den = 0.0f;
    for (long i = threadIdx.x; i < px; i += blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        num += dm[i] * a_apply(dm, kxq, kyq, x, y, h, w);
        den += tm[i] * a_apply(tm, kxq, kyq, x, y, h, w);
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        num += __shfl_xor(num, off, WAVE);
        den += __shfl_xor(den, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0) {
        red[0][threadIdx.x / WAVE] = num;
        red[1][threadIdx.x / WAVE] = den;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        num = den = 0.0f;
        for (int k = 0; k < (int)blockDim.x / WAVE; ++k) {
            num += red[0][k];
            den += red[1][k];
        }
        sums[2 * q] = num;
        sums[2 * q + 1] = den;
    }
}

// Second pass: g = 2 A (M d) / (n sum(Mt.A Mt)) per sample, then M^T g by m steps
// g <- g - omega A (D^-1 g) (planes 2 and 3 hold g and D^-1 g), into dout.
__global__ __launch_bounds__(256) void k_anorm_smooth_grad(const float* kx, const float* ky, float* buf, long plane,
                                                           u16* dout, int cout, int h, int w, int m, float omega,
                                                           const float* sums, float* loss)
{
    const int q = blockIdx.x;
    const long px = (long)h * w, base = (long)q * px;
    const float* kxq = kx + base;
    const float* kyq = ky + base;
    const float* dm = buf + (m & 1) * plane + base;
    float* g = buf + 2 * plane + base;
    float* v = buf + 3 * plane + base;
    // each sample's denominator floored at the batch mean, so a sample whose
    // target the sweeps already remove does not divide by ~0
    float mean = 0.0f;
    for (int k = 0; k < (int)gridDim.x; ++k) mean += sums[2 * k + 1];
    mean /= (float)gridDim.x;
    const float den = fmaxf(sums[2 * q + 1], mean);
    const float inv = den > 0.0f ? 1.0f / (den * (float)gridDim.x) : 0.0f;
    for (long i = threadIdx.x; i < px; i += blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        g[i] = 2.0f * a_apply(dm, kxq, kyq, x, y, h, w) * inv;
    }
    __syncthreads();
    for (int j = 0; j < m; ++j) {
        for (long i = threadIdx.x; i < px; i += blockDim.x) {
            const int x = (int)(i % w), y = (int)(i / w);
            v[i] = g[i] / a_diag(kxq, kyq, x, y, h, w);
        }
        __syncthreads();
        for (long i = threadIdx.x; i < px; i += blockDim.x) {
            const int x = (int)(i % w), y = (int)(i / w);
            g[i] -= omega * a_apply(v, kxq, kyq, x, y, h, w);
        }
        __syncthreads();
    }
    for (long i = threadIdx.x; i < px; i += blockDim.x)
        for (int c = 0; c < cout; ++c) dout[(base + i) * cout + c] = c == 0 ? f2bf(g[i]) : (u16)0;
    if (threadIdx.x == 0) atomicAdd(loss, sums[2 * q] * inv);
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Network input from FP32 host-side planes: channel 0 = residual, 1 = the
// density feature, 2..7 = 0.
__global__ void k_pack_input(const float* r, const float* f, u16* x, long pixels)
{
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += (long)gridDim.x * blockDim.x) {
        uint4 v = make_uint4((uint32_t)f2bf(r[p]) | ((uint32_t)f2bf(f[p]) << 16), 0u, 0u, 0u);
        *reinterpret_cast<uint4*>(x + p * 8) = v;
    }
}

__global__ void k_unpack_output(const u16* out, int cout, float* z, long pixels)
{
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += (long)gridDim.x * blockDim.x)
        z[p] = bf2f(out[p * cout]);
}

unsigned grid_for(long n, int block = 256)
{
    const long b = (n + block - 1) / block;
    return (unsigned)(b < 65536 ? (b > 0 ? b : 1) : 65536);
}

void check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

bool g_conv_staging = true;

// LDS bytes of a staged tile of TR output rows
long staged_bytes(int tr, int cin) { return (long)(tr + 2) * 66 * (cin + 8) * 2; }

// LDS a workgroup may use on the current device
long lds_limit()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
    return v;
}

template <int NT, int TR, bool RELU, bool MASK>
void launch_staged(const ConvArgs& a, hipStream_t s)
{
    const long bytes = staged_bytes(TR, a.cin);
    static bool raised = false;  // per instantiation: allow more than the default dynamic LDS
    if (!raised && bytes > 65536) {
        TL_HIP(hipFuncSetAttribute((const void*)k_conv3x3_lds<NT, TR, RELU, MASK>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_limit()));
        raised = true;
    }
    const long blocks = (long)a.n * ((a.h + TR - 1) / TR) * a.xtiles;
    hipLaunchKernelGGL((k_conv3x3_lds<NT, TR, RELU, MASK>), dim3((unsigned)blocks), dim3(256), (size_t)bytes, s, a);
}

template <int NT, bool RELU, bool MASK>
void launch_conv_nt(ConvArgs& a, int tr, hipStream_t s)
{
    switch (tr) {
    case 4: launch_staged<NT, 4, RELU, MASK>(a, s); return;
    case 2: launch_staged<NT, 2, RELU, MASK>(a, s); return;
    case 1: launch_staged<NT, 1, RELU, MASK>(a, s); return;
    default: break;
    }
    a.waves = (long)a.n * a.h * a.xtiles * a.cotiles;
    hipLaunchKernelGGL((k_conv3x3<NT, RELU, MASK>), dim3((unsigned)((a.waves + 3) / 4)), dim3(256), 0, s, a);
}

// One 3x3 convolution of ConvArgs a (xtiles set; cout a multiple of 16 * nt):
// the staged kernel where its tile fits the LDS, with its column tiles
// narrowed until the cotiles split evenly over the waves of a row, else the
// one that reads through L2.
template <bool RELU, bool MASK>
void launch_conv(ConvArgs& a, int nt, hipStream_t s)
{
    static const long limit = lds_limit();
    int tr = 0;
    if (g_conv_staging && (long)a.n * a.h * a.xtiles < (1L << 31) && a.cin % 8 == 0) {
        if (staged_bytes(4, a.cin) <= std::min(limit, 65536L)) tr = 4;
        else if (staged_bytes(2, a.cin) <= std::min(limit, 81920L)) tr = 2;
        else if (staged_bytes(1, a.cin) <= limit) tr = 1;
    }
    if (tr) {
        const int g = 4 / tr;
        while (nt > 1 && (a.cout / (16 * nt)) % g) nt /= 2;
        if ((a.cout / (16 * nt)) % g) tr = 0;
    }
    a.cotiles = a.cout / (16 * nt);
    if (nt == 4) launch_conv_nt<4, RELU, MASK>(a, tr, s);
    else if (nt == 2) launch_conv_nt<2, RELU, MASK>(a, tr, s);
    else launch_conv_nt<1, RELU, MASK>(a, tr, s);
}

}  // namespace

void set_conv_staging(bool on) { g_conv_staging = on; }

// ---------------------------------------------------------------------------
// Network

struct UConv {
    int cin = 0, cout = 0, kp = 0, kpb = 0;
    bool relu = true;
    long nw = 0;               // cout * 9 * cin
    float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    u16 *wf = nullptr, *wb = nullptr;
};

struct UBuf {
    u16* d = nullptr;
    long cap = 0;  // elements
};

// Activations of one forward pass at a given batch shape, and their gradients.
struct UNet {
    static constexpr int NL = 15;
    UConv c[NL];
    hipStream_t stream = nullptr;  // where the kernels run
    hipStream_t own = nullptr;     // the stream created with the network
    int step = 0;
    // activation and gradient buffers, indexed by the slot enum below
    enum {
        X0, E0A, E0B, P0, E1A, E1B, P1, E2A, E2B, P2, E3A, E3B, CAT2, D2A, D2B, CAT1, D1A, D1B, CAT0, D0A, D0B, OUT,
        NSLOT
    };
    UBuf act[NSLOT], grad[NSLOT];
    float* target = nullptr;
    long target_cap = 0;
    float* dbuf = nullptr;  // A-norm loss: out - target (and the smoothed loss's work planes)
    long dbuf_cap = 0;
    int smooth = 4;         // A-norm loss measured after this many weighted Jacobi sweeps
    float omega = 0.8f;     // their weight
    float* sums = nullptr;  // A-norm loss: sum(d.A d), sum(t.A t) (per sample for the smoothed loss)
    long sums_cap = 2;
    float* loss = nullptr;
    float* stage_r = nullptr;  // input staging (FP32 planes)
    float* stage_f = nullptr;
    long stage_cap = 0;

    explicit UNet(uint64_t seed);
    ~UNet();
    void pack_weights();
    void ensure(int n, int h, int w);
    void forward(int n, int h, int w);
    float loss_and_backward(int n, int h, int w, const float* kx, const float* ky);
    void adam(float lr);
    long param_count() const;
    void get_params(float* host) const;
    void set_params(const float* host);
    void get_grads(float* host) const;

  private:
    void conv_fwd(int l, const u16* x, u16* y, int n, int h, int w);
    void conv_bwd(int l, const u16* x, const u16* y, const u16* dy, u16* dx, int n, int h, int w);
};

namespace {

// layer shapes: (cin, cout, relu)
const int kShapes[UNet::NL][3] = {
    {8, 16, 1},    {16, 16, 1},  {16, 32, 1},  {32, 32, 1},  {32, 64, 1}, {64, 64, 1}, {64, 128, 1}, {128, 128, 1},
    {192, 64, 1},  {64, 64, 1},  {96, 32, 1},  {32, 32, 1},  {48, 16, 1}, {16, 16, 1}, {16, 16, 0}};

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double normal(uint64_t& s)
{
    const double u1 = ((splitmix(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = (splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

template <class T>
T* dalloc(long n)
{
    T* p = nullptr;
    TL_HIP(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    TL_HIP(hipMemset(p, 0, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    // the fill runs on the null stream, which is not ordered with the
    // network's non-blocking stream: finish it before a kernel writes p
    TL_HIP(hipDeviceSynchronize());
    return p;
}

// Frees a buffer that kernels still queued on the network's stream may use.
void dfree(void* p)
{
    if (!p) return;
    TL_HIP(hipDeviceSynchronize());
    (void)hipFree(p);
}

void grow(UBuf& b, long n)
{
    if (n <= b.cap) return;
    dfree(b.d);
    b.d = dalloc<u16>(n);
    b.cap = n;
}

}  // namespace

UNet::UNet(uint64_t seed)
{
    TL_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    stream = own;
    uint64_t s = seed;
    for (int l = 0; l < NL; ++l) {
        UConv& k = c[l];
        k.cin = kShapes[l][0];
        k.cout = kShapes[l][1];
        k.relu = kShapes[l][2] != 0;
        k.kp = (9 * k.cin + 31) / 32 * 32;
        k.kpb = (9 * k.cout + 31) / 32 * 32;
        k.nw = (long)k.cout * 9 * k.cin;
        // He-normal over the real fan-in (the first layer has 2 live inputs)
        const int fan_in = 9 * (l == 0 ? 2 : k.cin);
        const double sd = std::sqrt(2.0 / fan_in) * (k.relu ? 1.0 : 0.5);
        std::vector<float> hw(k.nw);
        for (long i = 0; i < k.nw; ++i) {
            const int ci = (int)(i % k.cin);
            hw[i] = (l == 0 && ci >= 2) ? 0.0f : (float)(sd * normal(s));
        }
        k.w = dalloc<float>(k.nw);
        k.gw = dalloc<float>(k.nw);
        k.mw = dalloc<float>(k.nw);
        k.vw = dalloc<float>(k.nw);
        k.b = dalloc<float>(k.cout);
        k.gb = dalloc<float>(k.cout);
        k.mb = dalloc<float>(k.cout);
        k.vb = dalloc<float>(k.cout);
        k.wf = dalloc<u16>((long)k.cout * k.kp);
        k.wb = dalloc<u16>((long)k.cin * k.kpb);
        TL_HIP(hipMemcpy(k.w, hw.data(), sizeof(float) * k.nw, hipMemcpyHostToDevice));
    }
    loss = dalloc<float>(1);
    sums = dalloc<float>(2);
    pack_weights();
    TL_HIP(hipStreamSynchronize(stream));
}

UNet::~UNet()
{
    if (stream) (void)hipStreamSynchronize(stream);
    if (own) (void)hipStreamSynchronize(own);
    for (UConv& k : c) {
        for (float* p : {k.w, k.b, k.gw, k.gb, k.mw, k.vw, k.mb, k.vb}) (void)hipFree(p);
        (void)hipFree(k.wf);
        (void)hipFree(k.wb);
    }
    for (int i = 0; i < NSLOT; ++i) {
        (void)hipFree(act[i].d);
        (void)hipFree(grad[i].d);
    }
    (void)hipFree(target);
    (void)hipFree(dbuf);
    (void)hipFree(sums);
    (void)hipFree(loss);
    (void)hipFree(stage_r);
    (void)hipFree(stage_f);
    if (own) (void)hipStreamDestroy(own);
}

void UNet::pack_weights()
{
    for (UConv& k : c) {
        hipLaunchKernelGGL(k_pack_weights, dim3(grid_for(k.nw)), dim3(256), 0, stream, k.w, k.wf, k.wb, k.cin, k.cout,
                           k.kp, k.kpb);
        check_launch("pack weights");
    }
}

long UNet::param_count() const
{
    long n = 0;
    for (const UConv& k : c) n += k.nw + k.cout;
    return n;
}

void UNet::get_params(float* host) const
{
    TL_HIP(hipStreamSynchronize(stream));
    for (const UConv& k : c) {
        TL_HIP(hipMemcpy(host, k.w, sizeof(float) * k.nw, hipMemcpyDeviceToHost));
        host += k.nw;
        TL_HIP(hipMemcpy(host, k.b, sizeof(float) * k.cout, hipMemcpyDeviceToHost));
        host += k.cout;
    }
}

void UNet::set_params(const float* host)
{
    TL_HIP(hipStreamSynchronize(stream));
    for (UConv& k : c) {
        TL_HIP(hipMemcpy(k.w, host, sizeof(float) * k.nw, hipMemcpyHostToDevice));
        host += k.nw;
        TL_HIP(hipMemcpy(k.b, host, sizeof(float) * k.cout, hipMemcpyHostToDevice));
        host += k.cout;
    }
    pack_weights();
    TL_HIP(hipStreamSynchronize(stream));
}

void UNet::get_grads(float* host) const
{
    TL_HIP(hipStreamSynchronize(stream));
    for (const UConv& k : c) {
        TL_HIP(hipMemcpy(host, k.gw, sizeof(float) * k.nw, hipMemcpyDeviceToHost));
        host += k.nw;
        TL_HIP(hipMemcpy(host, k.gb, sizeof(float) * k.cout, hipMemcpyDeviceToHost));
        host += k.cout;
    }
}

// Activation sizes of an (n, h, w) batch; h and w divisible by 8.
void UNet::ensure(int n, int h, int w)
{
    if (n < 1 || h < 8 || w < 8 || h % 8 || w % 8) throw Error("U-Net input must be a batch of h x w, both multiples of 8");
    const long p0 = (long)n * h * w, p1 = p0 / 4, p2 = p0 / 16, p3 = p0 / 64;
    const long sz[NSLOT] = {p0 * 8,  p0 * 16, p0 * 16, p1 * 16, p1 * 32, p1 * 32, p2 * 32, p2 * 64,
                            p2 * 64, p3 * 64, p3 * 128, p3 * 128, p2 * 192, p2 * 64, p2 * 64, p1 * 96,
                            p1 * 32, p1 * 32, p0 * 48, p0 * 16, p0 * 16, p0 * 16};
    for (int i = 0; i < NSLOT; ++i) grow(act[i], sz[i]);
}

void UNet::conv_fwd(int l, const u16* x, u16* y, int n, int h, int w)
{
    const UConv& k = c[l];
    ConvArgs a{};
    a.x = x;
    a.m = nullptr;
    a.wt = k.wf;
    a.bias = k.b;
    a.y = y;
    a.n = n; a.h = h; a.w = w;
    a.cin = k.cin; a.cout = k.cout; a.kp = k.kp;
    const int nt = col_tiles(k.cout);
    a.cotiles = k.cout / (16 * nt);
    a.xtiles = (w + 63) / 64;
    if (k.relu) launch_conv<true, false>(a, nt, stream);
    else launch_conv<false, false>(a, nt, stream);
    check_launch("conv forward");
}

// Gradients of layer l: dw, db accumulated from (x, dY masked by ReLU'(y)),
// and dx = the backward-data convolution (if dx is not null).
void UNet::conv_bwd(int l, const u16* x, const u16* y, const u16* dy, u16* dx, int n, int h, int w)
{
    const UConv& k = c[l];
    const long pixels = (long)n * h * w;
    WGradArgs g{};
    g.x = x; g.dy = dy; g.ym = y; g.dw = k.gw;
    g.n = n; g.h = h; g.w = w; g.cin = k.cin; g.cout = k.cout;
    g.cotiles = k.cout / 16;
    g.citiles = (k.cin + 15) / 16;
    g.pixels = pixels;
    g.chunk = 2048;
    g.waves = 9L * g.cotiles * g.citiles * ((pixels + g.chunk - 1) / g.chunk);
    const dim3 grid((unsigned)((g.waves + 3) / 4)), block(256);
    if (k.relu) hipLaunchKernelGGL((k_conv_wgrad<true>), grid, block, 0, stream, g);
    else hipLaunchKernelGGL((k_conv_wgrad<false>), grid, block, 0, stream, g);
    check_launch("conv weight gradient");
    const int bthreads = 256 / k.cout * k.cout;
    if (k.relu) hipLaunchKernelGGL((k_bias_grad<true>), dim3(256), dim3(bthreads), 0, stream, dy, y, k.gb, pixels, k.cout);
    else hipLaunchKernelGGL((k_bias_grad<false>), dim3(256), dim3(bthreads), 0, stream, dy, y, k.gb, pixels, k.cout);
    check_launch("conv bias gradient");
    if (!dx) return;
    ConvArgs a{};
    a.x = dy;
    a.m = y;
    a.wt = k.wb;
    a.bias = nullptr;
    a.y = dx;
    a.n = n; a.h = h; a.w = w;
    a.cin = k.cout; a.cout = k.cin; a.kp = k.kpb;
    if (k.cin % 16) throw Error("backward-data convolution needs a multiple of 16 input channels");
    const int nt = col_tiles(k.cin);
    a.cotiles = k.cin / (16 * nt);
    a.xtiles = (w + 63) / 64;
    if (k.relu) launch_conv<false, true>(a, nt, stream);
    else launch_conv<false, false>(a, nt, stream);
    check_launch("conv backward data");
}

void UNet::forward(int n, int h, int w)
{
    const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
    const long p0 = (long)n * h * w, p1 = p0 / 4, p2 = p0 / 16;
    conv_fwd(0, act[X0].d, act[E0A].d, n, h, w);
    conv_fwd(1, act[E0A].d, act[E0B].d, n, h, w);
    hipLaunchKernelGGL(k_avgpool, dim3(grid_for(p1 * 16)), dim3(256), 0, stream, act[E0B].d, act[P0].d, n, h, w, 16);
    conv_fwd(2, act[P0].d, act[E1A].d, n, h1, w1);
    conv_fwd(3, act[E1A].d, act[E1B].d, n, h1, w1);
    hipLaunchKernelGGL(k_avgpool, dim3(grid_for(p2 * 32)), dim3(256), 0, stream, act[E1B].d, act[P1].d, n, h1, w1, 32);
    conv_fwd(4, act[P1].d, act[E2A].d, n, h2, w2);
    conv_fwd(5, act[E2A].d, act[E2B].d, n, h2, w2);
    hipLaunchKernelGGL(k_avgpool, dim3(grid_for(p2 / 4 * 64)), dim3(256), 0, stream, act[E2B].d, act[P2].d, n, h2, w2,
                       64);
    conv_fwd(6, act[P2].d, act[E3A].d, n, h3, w3);
    conv_fwd(7, act[E3A].d, act[E3B].d, n, h3, w3);
    hipLaunchKernelGGL(k_up_concat, dim3(grid_for(p2 * 192 / 8)), dim3(256), 0, stream, act[E3B].d, act[E2B].d,
                       act[CAT2].d, n, h2, w2, 128, 64);
    conv_fwd(8, act[CAT2].d, act[D2A].d, n, h2, w2);
    conv_fwd(9, act[D2A].d, act[D2B].d, n, h2, w2);
    hipLaunchKernelGGL(k_up_concat, dim3(grid_for(p1 * 96 / 8)), dim3(256), 0, stream, act[D2B].d, act[E1B].d,
                       act[CAT1].d, n, h1, w1, 64, 32);
    conv_fwd(10, act[CAT1].d, act[D1A].d, n, h1, w1);
    conv_fwd(11, act[D1A].d, act[D1B].d, n, h1, w1);
    hipLaunchKernelGGL(k_up_concat, dim3(grid_for(p0 * 48 / 8)), dim3(256), 0, stream, act[D1B].d, act[E0B].d,
                       act[CAT0].d, n, h, w, 32, 16);
    conv_fwd(12, act[CAT0].d, act[D0A].d, n, h, w);
    conv_fwd(13, act[D0A].d, act[D0B].d, n, h, w);
    conv_fwd(14, act[D0B].d, act[OUT].d, n, h, w);
    check_launch("U-Net forward");
}

float UNet::loss_and_backward(int n, int h, int w, const float* kx, const float* ky)
{
    const int h1 = h / 2, w1 = w / 2, h2 = h / 4, w2 = w / 4, h3 = h / 8, w3 = w / 8;
    const long p0 = (long)n * h * w;
    for (int i = 0; i < NSLOT; ++i) grow(grad[i], act[i].cap);
    for (UConv& k : c) {
        TL_HIP(hipMemsetAsync(k.gw, 0, sizeof(float) * k.nw, stream));
        TL_HIP(hipMemsetAsync(k.gb, 0, sizeof(float) * k.cout, stream));
    }
    TL_HIP(hipMemsetAsync(loss, 0, sizeof(float), stream));
    if (kx && ky) {
        if (4 * p0 > dbuf_cap) {
            dfree(dbuf);
            dbuf = dalloc<float>(4 * p0);
            dbuf_cap = 4 * p0;
        }
        if (2L * n > sums_cap) {
            dfree(sums);
            sums = dalloc<float>(2L * n);
            sums_cap = 2L * n;
        }
        TL_HIP(hipMemsetAsync(sums, 0, sizeof(float) * 2, stream));
        if (smooth > 0) {
            hipLaunchKernelGGL(k_anorm_smooth_sums, dim3(n), dim3(256), 0, stream, act[OUT].d, 16, target, kx, ky,
                               dbuf, p0, h, w, smooth, omega, sums);
            hipLaunchKernelGGL(k_anorm_smooth_grad, dim3(n), dim3(256), 0, stream, kx, ky, dbuf, p0, grad[OUT].d, 16,
                               h, w, smooth, omega, sums, loss);
        } else {
            hipLaunchKernelGGL(k_anorm_sums, dim3(n), dim3(256), 0, stream, act[OUT].d, 16, target, kx, ky, dbuf, h,
                               w, sums);
            hipLaunchKernelGGL(k_anorm_grad, dim3(n), dim3(256), 0, stream, kx, ky, dbuf, grad[OUT].d, 16, h, w, sums,
                               loss);
        }
    } else {
        hipLaunchKernelGGL(k_mse, dim3(grid_for(p0)), dim3(256), 0, stream, act[OUT].d, 16, target, grad[OUT].d, p0,
                           loss);
    }
    check_launch("loss");
    conv_bwd(14, act[D0B].d, act[OUT].d, grad[OUT].d, grad[D0B].d, n, h, w);
    conv_bwd(13, act[D0A].d, act[D0B].d, grad[D0B].d, grad[D0A].d, n, h, w);
    conv_bwd(12, act[CAT0].d, act[D0A].d, grad[D0A].d, grad[CAT0].d, n, h, w);
    // dCAT0 -> dD1B (upsampled part), dE0B (skip part; the pool path adds later)
    hipLaunchKernelGGL(k_up_concat_bwd, dim3(grid_for(p0 * 16)), dim3(256), 0, stream, grad[CAT0].d, grad[D1B].d,
                       grad[E0B].d, n, h, w, 32, 16);
    conv_bwd(11, act[D1A].d, act[D1B].d, grad[D1B].d, grad[D1A].d, n, h1, w1);
    conv_bwd(10, act[CAT1].d, act[D1A].d, grad[D1A].d, grad[CAT1].d, n, h1, w1);
    hipLaunchKernelGGL(k_up_concat_bwd, dim3(grid_for(p0 / 4 * 32)), dim3(256), 0, stream, grad[CAT1].d, grad[D2B].d,
                       grad[E1B].d, n, h1, w1, 64, 32);
    conv_bwd(9, act[D2A].d, act[D2B].d, grad[D2B].d, grad[D2A].d, n, h2, w2);
    conv_bwd(8, act[CAT2].d, act[D2A].d, grad[D2A].d, grad[CAT2].d, n, h2, w2);
    hipLaunchKernelGGL(k_up_concat_bwd, dim3(grid_for(p0 / 16 * 64)), dim3(256), 0, stream, grad[CAT2].d, grad[E3B].d,
                       grad[E2B].d, n, h2, w2, 128, 64);
    conv_bwd(7, act[E3A].d, act[E3B].d, grad[E3B].d, grad[E3A].d, n, h3, w3);
    conv_bwd(6, act[P2].d, act[E3A].d, grad[E3A].d, grad[P2].d, n, h3, w3);
    hipLaunchKernelGGL(k_avgpool_bwd, dim3(grid_for(p0 / 16 * 64)), dim3(256), 0, stream, grad[P2].d, grad[E2B].d, n,
                       h2, w2, 64, 1);
    conv_bwd(5, act[E2A].d, act[E2B].d, grad[E2B].d, grad[E2A].d, n, h2, w2);
    conv_bwd(4, act[P1].d, act[E2A].d, grad[E2A].d, grad[P1].d, n, h2, w2);
    hipLaunchKernelGGL(k_avgpool_bwd, dim3(grid_for(p0 / 4 * 32)), dim3(256), 0, stream, grad[P1].d, grad[E1B].d, n,
                       h1, w1, 32, 1);
    conv_bwd(3, act[E1A].d, act[E1B].d, grad[E1B].d, grad[E1A].d, n, h1, w1);
    conv_bwd(2, act[P0].d, act[E1A].d, grad[E1A].d, grad[P0].d, n, h1, w1);
    hipLaunchKernelGGL(k_avgpool_bwd, dim3(grid_for(p0 * 16)), dim3(256), 0, stream, grad[P0].d, grad[E0B].d, n, h, w,
                       16, 1);
    conv_bwd(1, act[E0A].d, act[E0B].d, grad[E0B].d, grad[E0A].d, n, h, w);
    conv_bwd(0, act[X0].d, act[E0A].d, grad[E0A].d, nullptr, n, h, w);
    check_launch("U-Net backward");
    float l = 0.0f;
    TL_HIP(hipMemcpyAsync(&l, loss, sizeof(float), hipMemcpyDeviceToHost, stream));
    TL_HIP(hipStreamSynchronize(stream));
    return l;
}

void UNet::adam(float lr)
{
    ++step;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - std::pow(b1, (float)step), bc2 = 1.0f - std::pow(b2, (float)step);
    for (UConv& k : c) {
        hipLaunchKernelGGL(k_adam, dim3(grid_for(k.nw)), dim3(256), 0, stream, k.w, k.gw, k.mw, k.vw, k.nw, lr, b1,
                           b2, eps, bc1, bc2);
        hipLaunchKernelGGL(k_adam, dim3(grid_for(k.cout)), dim3(256), 0, stream, k.b, k.gb, k.mb, k.vb, (long)k.cout,
                           lr, b1, b2, eps, bc1, bc2);
    }
    check_launch("adam");
    pack_weights();
}

// ---------------------------------------------------------------------------
// Host entry points used by the learned-preconditioner code and the tests.

UNet* unet_create(uint64_t seed) { return new UNet(seed); }
void unet_destroy(UNet* u) { delete u; }

// Network input from device planes r (residual, already scaled) and f (the
// density feature), n x h x w pixels; the output's channel 0 into z.
void unet_apply(UNet& u, const float* r, const float* f, float* z, int n, int h, int w)
{
    u.ensure(n, h, w);
    const long p = (long)n * h * w;
    hipLaunchKernelGGL(k_pack_input, dim3(grid_for(p)), dim3(256), 0, u.stream, r, f, u.act[UNet::X0].d, p);
    check_launch("pack input");
    u.forward(n, h, w);
    hipLaunchKernelGGL(k_unpack_output, dim3(grid_for(p)), dim3(256), 0, u.stream, u.act[UNet::OUT].d, 16, z, p);
    check_launch("unpack output");
}

// One loss + gradient evaluation on device planes (r, f, target; kx, ky for
// the A-norm loss, null for the mean squared error); returns the loss.
float unet_loss_grad(UNet& u, const float* r, const float* f, const float* target, const float* kx, const float* ky,
                     int n, int h, int w)
{
    u.ensure(n, h, w);
    const long p = (long)n * h * w;
    if (p > u.target_cap) {
        dfree(u.target);
        u.target = dalloc<float>(p);
        u.target_cap = p;
    }
    TL_HIP(hipMemcpyAsync(u.target, target, sizeof(float) * p, hipMemcpyDeviceToDevice, u.stream));
    hipLaunchKernelGGL(k_pack_input, dim3(grid_for(p)), dim3(256), 0, u.stream, r, f, u.act[UNet::X0].d, p);
    check_launch("pack input");
    u.forward(n, h, w);
    return u.loss_and_backward(n, h, w, kx, ky);
}

void unet_adam(UNet& u, float lr) { u.adam(lr); }

void unet_set_loss_smoothing(UNet& u, int sweeps, float omega)
{
    if (sweeps < 0 || sweeps > 16) throw Error("loss smoothing: sweeps must be in [0, 16]");
    if (!(omega > 0.0f && omega <= 1.0f)) throw Error("loss smoothing: omega must be in (0, 1]");
    u.smooth = sweeps;
    u.omega = omega;
}

double unet_forward_macs(int n, int h, int w)
{
    // layer -> pyramid level of its output
    static const int level[UNet::NL] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    double macs = 0.0;
    for (int l = 0; l < UNet::NL; ++l) {
        const double px = (double)n * h * w / (double)(1 << (2 * level[l]));
        macs += px * 9.0 * kShapes[l][0] * kShapes[l][1];
    }
    return macs;
}
double unet_forward_bytes(int n, int h, int w)
{
    // every tensor of the forward pass read or written once, as stored
    // (BF16 activations with their padded channel counts, FP32 planes in and out)
    static const int level[UNet::NL] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    const double p0 = (double)n * h * w;
    auto px = [&](int lev) { return p0 / (double)(1 << (2 * lev)); };
    double b = p0 * (8.0 + 2.0 * kShapes[0][0]);  // r, f (FP32) -> input channels
    for (int l = 0; l < UNet::NL; ++l)
        b += 2.0 * px(level[l]) * (kShapes[l][0] + kShapes[l][1]) + 2.0 * 9.0 * kShapes[l][0] * kShapes[l][1];
    const int pool_c[3] = {16, 32, 64};  // 2x2 average pooling after levels 0, 1, 2
    for (int lev = 0; lev < 3; ++lev) b += 2.0 * pool_c[lev] * (px(lev) + px(lev + 1));
    // up-sampling with concatenation into levels 2, 1, 0: (up, skip) channels
    const int up_c[3][2] = {{128, 64}, {64, 32}, {32, 16}};
    for (int k = 0; k < 3; ++k) {
        const int lev = 2 - k;
        b += 2.0 * (px(lev + 1) * up_c[k][0] + px(lev) * up_c[k][1] + px(lev) * (up_c[k][0] + up_c[k][1]));
    }
    b += p0 * (2.0 + 4.0);  // output channel -> FP32 plane
    return b;
}

hipStream_t unet_stream(UNet& u) { return u.stream; }
void unet_set_stream(UNet& u, hipStream_t s) { u.stream = s; }
long unet_param_count(const UNet& u) { return u.param_count(); }
void unet_get_params(const UNet& u, float* host) { u.get_params(host); }
void unet_set_params(UNet& u, const float* host) { u.set_params(host); }
void unet_get_grads(const UNet& u, float* host) { u.get_grads(host); }

}  // namespace tl
