// Host side of the temporal-blocked Chebyshev iterations (2D): pass plan
// and launches; the device code is tl_cheby_tb.hpp.
#include "tl_cheby_tb.hpp"

#include <vector>

namespace tl {

#if TL_DIM == 2

bool cheby_tb_supported(const Chunk& c)
{
    // the kernel addresses a field with 32-bit byte offsets
    const bool fits = (double)c.g.alloc * sizeof(double) < 4294967296.0;
    return !c.preconditioner && c.g.nx >= 2 && c.g.ny >= 2 && fits && c.d_cheby;
}

// Iterations step1 .. step1 + count - 1 (ext_cheby_solver_iterate_,
// ext_core.cpp:501-515, each), as passes of 10, 5, 2 and 1 iterations
// that hand u and p over through memory.
void op_cheby_tb(Chunk& c, int step1, int count, bool norm)
{
    if (!cheby_tb_supported(c)) throw Error("temporal-blocked Chebyshev is 2D, no preconditioner");
    if (count < 1) throw Error("temporal-blocked Chebyshev needs at least one iteration");
    if (step1 < 1 || step1 + count - 1 > c.d_cheby_n)
        throw Error("Chebyshev steps " + std::to_string(step1) + ".." + std::to_string(step1 + count - 1) +
                    " outside the stored coefficients");
    const std::vector<int> passes = tb_passes(count);
    int step = step1;
    for (size_t q = 0; q < passes.size(); ++q) {
        const int K = passes[q];
        const bool last = q + 1 == passes.size();
        ctb::CTBArgs a{};
        a.u_in = c.f[F_U];
        a.u_out = c.u_alt;
        a.p_in = c.f[F_P];
        a.p_out = c.p_alt;
        a.r = c.f[F_R];
        a.u0 = c.f[F_U0];
        a.kx = c.f[F_KX];
        a.ky = c.f[F_KY];
        a.ca = c.d_cheby + (step - 1);
        a.cb = c.d_cheby + c.d_cheby_n + (step - 1);
        a.rowtab = tb_rowtab(c);
        a.row0 = TB_ROWTAB0;
        a.red = last && norm ? c.reducer(S_NORM, EP_STORE) : c.reducer();
        const int ow = 2 * WAVE - 2 * tb::tb_margin(K);
        a.tiles_x = (c.g.nx + ow - 1) / ow;
        a.rows = tb_rows(c.g.ny, a.tiles_x, K, c.simds, tb::TB_UNROLL);
        a.nblocks = (long)a.tiles_x * ((c.g.ny + a.rows - 1) / a.rows);
        if (a.nblocks > c.partials_cap) throw Error("temporal-blocked Chebyshev: more waves than reduction partials");
        TL_TIMED(c, K_CHEBY_TB, {
            switch (K) {
            case 10: ctb::ctb_launch_10(c, a, last); break;
            case 5: ctb::ctb_launch_5(c, a, last); break;
            case 2: ctb::ctb_launch_2(c, a, last); break;
            default: ctb::ctb_launch_1(c, a, last); break;
            }
        });
        if (last && norm) c.close(a.nblocks, 1);
        std::swap(c.f[F_U], c.u_alt);
        std::swap(c.f[F_P], c.p_alt);
        step += K;
    }
}

#else  // the 7-point stencil runs one iteration per launch

bool cheby_tb_supported(const Chunk&) { return false; }
void op_cheby_tb(Chunk&, int, int, bool) { throw Error("temporal-blocked Chebyshev is implemented for 2D"); }

#endif

}  // namespace tl
