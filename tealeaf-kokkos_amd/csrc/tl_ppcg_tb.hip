// Temporal-blocked PPCG outer iteration for the 2D 5-point stencil: pass
// planning and the host entry (device code in tl_ppcg_tb.hpp).
#include "tl_ppcg_tb.hpp"

#include <algorithm>
#include <vector>

namespace tl {

#if TL_DIM == 2

// Output rows per wave.  Every wave holds a whole SIMD (one resident wave
// each), so the pass takes ceil(waves / simds) turns of the slowest SIMD, each
// of (rows + 2L) rounds rounded up to the unrolled block; minimise that,
// and among equal costs take the fewest waves (least recomputed overlap).
int tb_rows(int ny, int tiles_x, int L, int simds, int unroll)
{
    int best_rows = 16;
    long best = -1;
    for (int rows = 16; rows <= std::max(16, ny); ++rows) {
        const long tiles_y = (ny + rows - 1) / rows;
        if (rows > 16 && (ny + rows - 2) / (rows - 1) == tiles_y) continue;  // same tiling as rows - 1
        const long waves = (long)tiles_x * tiles_y;
        const long rounds = ((long)std::min(rows, ny) + 2 * L + unroll - 1) / unroll * unroll;
        const long cost = (waves + simds - 1) / simds * rounds;
        if (best < 0 || cost < best || (cost == best && tiles_y < (ny + best_rows - 1) / best_rows)) {
            best = cost;
            best_rows = rows;
        }
    }
    return best_rows;
}

bool ppcg_tb_supported(const Chunk& c, int inner_steps)
{
    // the kernel addresses a field with 32-bit byte offsets
    const bool fits = (double)c.g.alloc * sizeof(double) < 4294967296.0;
    return TL_DIM == 2 && !c.preconditioner && inner_steps >= 1 && c.g.nx >= 2 && c.g.ny >= 2 && fits;
}

std::vector<int> tb_passes(int inner_steps)
{
    static const int sizes[] = {10, 5, 2, 1};
    std::vector<int> passes;
    int left = inner_steps;
    for (int n : sizes)
        while (left >= n) {
            passes.push_back(n);
            left -= n;
        }
    return passes;
}

namespace {

// The device coefficients in the layout of the pass plan (once per solve).
void tb_coefs(Chunk& c, int inner_steps, const std::vector<int>& passes)
{
    if (c.d_coef_steps == inner_steps) return;
    const int need = 2 * inner_steps;
    if (need > c.d_coef_cap) {
        c.sync();
        (void)hipFree(c.d_coef);
        c.d_coef = nullptr;
        TL_HIP(hipMalloc(&c.d_coef, sizeof(double) * need));
        c.d_coef_cap = need;
    }
    std::vector<double> h(need, 0.0);
    int k = 0;
    for (int m : passes) {
        for (int q = 0; q < m; ++q) {
            h[2 * k + q] = c.alphas[k + q];
            h[2 * k + m + q] = c.betas[k + q];
        }
        k += m;
    }
    c.sync();
    TL_HIP(hipMemcpy(c.d_coef, h.data(), sizeof(double) * need, hipMemcpyHostToDevice));
    c.d_coef_steps = inner_steps;
}

}  // namespace

// Rows a strip reads: HALO - 12 .. ny + 20 (12 levels at most, the
// prefetch and the six-round unroll); one uint2 (cell row, face row byte
// offsets, reflected into the mesh) per row, built once per chunk.
const uint2* tb_rowtab(Chunk& c)
{
    if (!c.d_rowtab) {
        const int count = c.g.ny + 48;
        std::vector<unsigned> h(2 * (size_t)count);
        for (int k = 0; k < count; ++k) {
            const int y = TB_ROWTAB0 + k;
            h[2 * k] = 8u * (unsigned)(c.g.base + (long)tb::fold_cell(y, c.g.ny) * c.g.sx);
            h[2 * k + 1] = 8u * (unsigned)(c.g.base + (long)tb::fold_face(y, c.g.ny) * c.g.sx);
        }
        TL_HIP(hipMalloc(&c.d_rowtab, sizeof(unsigned) * h.size()));
        TL_HIP(hipMemcpy(c.d_rowtab, h.data(), sizeof(unsigned) * h.size(), hipMemcpyHostToDevice));
    }
    return reinterpret_cast<const uint2*>(c.d_rowtab);
}

// One PPCG outer iteration after CGCalcW: alpha is read from the device
// scalars, u and r are updated, and the last pass closes the rrn reduction
// (beta, error and rro on the device, rrn in S_RRN).  The inner steps run as
// passes of 10, 5, 2 and 1 levels (largest first), handing sd over through
// memory between passes; every cell's arithmetic is the same in any split.
void op_ppcg_outer_tb(Chunk& c, int inner_steps, Slot slot)
{
    if (!ppcg_tb_supported(c, inner_steps)) throw Error("temporal-blocked PPCG is 2D, no preconditioner");
    if ((int)c.alphas.size() < inner_steps) throw Error("PPCG coefficients shorter than the inner steps");
    const std::vector<int> passes = tb_passes(inner_steps);
    tb_coefs(c, inner_steps, passes);
    int done = 0;
    for (size_t q = 0; q < passes.size(); ++q) {
        const int L = passes[q];
        const bool first = q == 0, last = q + 1 == passes.size();
        tb::TBArgs a{};
        a.u = c.f[F_U];
        a.p = c.f[F_P];
        a.r_in = c.f[F_R];
        a.w = c.f[F_W];
        a.sd_in = c.f[F_SD];
        a.r_out = c.r_alt;
        a.sd_out = c.sd_alt;
        a.kx = c.f[F_KX];
        a.ky = c.f[F_KY];
        a.ds = c.ds;
        a.theta = c.theta;
        a.coef = c.d_coef + 2 * done;
        a.rowtab = tb_rowtab(c);
        a.row0 = TB_ROWTAB0;
        a.red = last ? c.reducer(slot, EP_RRN) : c.reducer();
        const int ow = 2 * WAVE - 2 * tb::tb_margin(L);
        a.tiles_x = (c.g.nx + ow - 1) / ow;
        a.rows = tb_rows(c.g.ny, a.tiles_x, L, c.simds, tb::TB_UNROLL);
        const long tiles_y = (c.g.ny + a.rows - 1) / a.rows;
        a.nblocks = (long)a.tiles_x * tiles_y;
        if (a.nblocks > c.partials_cap) throw Error("temporal-blocked PPCG: more waves than reduction partials");
        // the timer brackets each pass alone, not the finish of its reduction
        TL_TIMED(c, K_PPCG_OUTER, {
            switch (L) {
            case 10: tb::tb_launch_10(c, a, first, last); break;
            case 5: tb::tb_launch_5(c, a, first, last); break;
            case 2: tb::tb_launch_2(c, a, first, last); break;
            default: tb::tb_launch_1(c, a, first, last); break;
            }
        });
        if (last) c.close(a.nblocks, 1);
        std::swap(c.f[F_R], c.r_alt);
        if (!last) std::swap(c.f[F_SD], c.sd_alt);
        done += L;
    }
}

#else  // the 7-point stencil runs the PPCG steps as separate sweeps

bool ppcg_tb_supported(const Chunk&, int) { return false; }
std::vector<int> tb_passes(int) { return {}; }
void op_ppcg_outer_tb(Chunk&, int, Slot) { throw Error("temporal-blocked PPCG is implemented for 2D"); }

#endif

}  // namespace tl
