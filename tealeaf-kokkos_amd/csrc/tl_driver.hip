// Host driver: TeaLeaf's Fortran caller of the ext_* layer, restated in C++
// so that a deck runs end to end on the GPU with no Fortran.  It follows
//   read_input.f90 / parse.f90       deck keywords, defaults, state shrink
//   start.f90, diffuse.f90           set-up order and the timestep loop
//   tea_solve.f90                    CG, Chebyshev, PPCG and Jacobi control
//   tea_leaf_cheby.f90:233-360       tqli, eigenvalues, Chebyshev coefficients
//   field_summary.f90                per-step totals
// and calls the chunk operations directly, choosing the fused kernels.
// Per-iteration halo updates of u, p and sd are not issued: every stencil
// kernel already reads the reflective mirror across external faces, which is
// what those updates would have put in memory.
#include "tl_chunk.hpp"

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace tl {
namespace {

struct State {
    bool defined = false;
    double density = 0, energy = 0, xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0, radius = 0;
    int geometry = 0;
};

struct Deck {
    // Defaults of read_input.f90:39-80.  Unsuffixed Fortran literals are
    // single precision, widened to double exactly as the reference does.
    double xmin = 0.0, ymin = 0.0, zmin = 0.0, xmax = 100.0, ymax = 100.0, zmax = 100.0;
    int x_cells = 10, y_cells = 10, z_cells = 10;
    double dtinit = (double)0.1f;
    int max_iters = 1000;
    double eps = (double)1.0e-10f;
    double end_time = 10.0;
    int end_step = 640000;  // g_ibig (data.f90:28)
    int summary_frequency = 10;
    int visit_frequency = 0;
    int coefficient = CONDUCTIVITY;
    bool errswitch = false;
    int presteps = 30;
    double epslim = (double)1e-5f;
    bool check_result = false;
    int ppcg_inner_steps = 10;
    bool preconditioner = false;
    bool use_cheby = false, use_cg = false, use_ppcg = false, use_jacobi = true;
    // Mixed-precision PPCG (no counterpart in the reference): FP32 inner
    // steps, the FP64 residual replaced once ||r||^2 fell by this factor.
    bool ppcg_mixed = false;
    double mixed_replace = 1.0e-2;
    // FP64 PPCG under the same residual replacement, accepting only the true
    // residual (no counterpart in the reference): the FP64 solve at the
    // accuracy the mixed and learned forms are held to.
    bool ppcg_true_residual = false;
    // tl_no_lookahead: launch each CG or PPCG iteration only after the
    // previous one's rrn is read (no counterpart in the reference; the
    // results are the same either way)
    bool lookahead = true;
    // Flexible CG with the learned preconditioner (no counterpart in the
    // reference), FP64 PPCG as its fallback solver.
    bool use_fcg = false;
    FcgOptions fcg;  // tl_fcg_smooth=<sweeps>, tl_fcg_omega=<w>, tl_fcg_cheby, tl_fcg_cheby_ratio=<a>,
                     // tl_fcg_net_every=<m>, tl_fcg_net_iters=<m>, tl_fcg_no_network, tl_fcg_net_restart,
                     // tl_fcg_coarse
    int test_problem = 0;
    std::vector<State> states;
};

// ---- deck reader (parse.f90:130-300, read_input.f90:83-297) --------------

// One content line of the *tea block, already cleaned.
std::vector<std::string> deck_lines(const std::string& text)
{
    std::vector<std::string> out;
    std::string here;
    std::istringstream in(text);
    std::string raw;
    while (std::getline(in, raw)) {
        std::string l = raw.substr(0, std::min<size_t>(raw.size(), 100));  // READ (a100)
        for (char& ch : l)
            if ((unsigned char)ch < 32 || (unsigned char)ch > 128) ch = ' ';
        auto trim = [](std::string& s) {
            size_t a = s.find_first_not_of(' ');
            if (a == std::string::npos) { s.clear(); return; }
            size_t b = s.find_last_not_of(' ');
            s = s.substr(a, b - a + 1);
        };
        trim(l);
        for (char cut : {'!', ';'}) {
            size_t k = l.find(cut);
            if (k != std::string::npos) { l.resize(k); trim(l); }
        }
        for (char& ch : l) {
            if (ch >= 'A' && ch <= 'Z') ch = (char)(ch + 32);
            if (ch == '=' || ch == ',') ch = ' ';
        }
        if (!l.empty() && l[0] == '*') {
            std::string nugget = l.substr(0, l.find(' '));
            if (nugget.compare(1, 3, "end") == 0) {
                std::string want = "*" + nugget.substr(4);
                size_t last = here.rfind('*');
                if (last == std::string::npos || here.substr(last) != want) throw Error("Unmatched */*end pair.");
                if (here == "*tea") return out;
                here.resize(last);
            } else {
                here += nugget;
            }
            continue;
        }
        if (here == "*tea" && !l.empty()) out.push_back(l);
    }
    return out;
}

struct Words {
    const std::vector<std::string>& lines;
    size_t line = 0, pos = 0;
    bool started = false;

    bool next_line()
    {
        if (started) ++line;
        started = true;
        pos = 0;
        return line < lines.size();
    }
    // parse_getword: next blank-separated word; with `wrap` an empty line
    // end continues on the next line.
    std::string word(bool wrap)
    {
        for (;;) {
            if (line >= lines.size()) return "";
            const std::string& l = lines[line];
            while (pos < l.size() && l[pos] == ' ') ++pos;
            if (pos < l.size()) {
                size_t e = l.find(' ', pos);
                if (e == std::string::npos) e = l.size();
                std::string w = l.substr(pos, e - pos);
                pos = e;
                return w;
            }
            if (!wrap) return "";
            ++line;
            pos = 0;
        }
    }
};

int parse_int(const std::string& w)  // READ (I7)
{
    std::string s = w.substr(0, std::min<size_t>(w.size(), 7));
    char* end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (end == s.c_str() || *end) throw Error("Error attempting to convert to integer:" + w);
    return (int)v;
}

double parse_real(const std::string& w)  // '.0' appended when no '.', then READ (E27.20)
{
    std::string s = w.find('.') == std::string::npos ? w + ".0" : w;
    char* end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end) throw Error("Error attempting to convert to real:" + w);
    return v;
}

Deck parse_deck(const std::string& text)
{
    Deck d;
    const std::vector<std::string> lines = deck_lines(text);
    int state_max = 0;
    {
        Words w{lines};
        while (w.next_line()) {
            for (std::string word = w.word(false); !word.empty(); word = w.word(false)) {
                if (word == "state") {
                    state_max = std::max(state_max, parse_int(w.word(true)));
                    break;
                }
            }
        }
    }
    if (state_max < 1) throw Error("No states defined.");
    d.states.resize(state_max);
    Words w{lines};
    bool smooth_set = false;  // tl_fcg_smooth given
    while (w.next_line()) {
        for (std::string k = w.word(false); !k.empty(); k = w.word(false)) {
            auto R = [&] { return parse_real(w.word(true)); };
            auto I = [&] { return parse_int(w.word(true)); };
            if (k == "initial_timestep") d.dtinit = R();
            else if (k == "end_time") d.end_time = R();
            else if (k == "end_step") d.end_step = I();
            else if (k == "xmin") d.xmin = R();
            else if (k == "xmax") d.xmax = R();
            else if (k == "ymin") d.ymin = R();
            else if (k == "ymax") d.ymax = R();
            else if (k == "x_cells") d.x_cells = I();
            else if (k == "y_cells") d.y_cells = I();
#if TL_DIM == 3
            else if (k == "zmin") d.zmin = R();
            else if (k == "zmax") d.zmax = R();
            else if (k == "z_cells") d.z_cells = I();
#endif
            else if (k == "visit_frequency") d.visit_frequency = I();
            else if (k == "summary_frequency") d.summary_frequency = I();
            else if (k == "tl_ch_cg_presteps") d.presteps = I();
            else if (k == "tl_ppcg_inner_steps") d.ppcg_inner_steps = I();
            else if (k == "tl_ch_cg_epslim") d.epslim = R();
            else if (k == "tl_check_result") d.check_result = true;
            else if (k == "tl_ch_cg_errswitch") d.errswitch = true;
            else if (k == "tl_preconditioner_on") d.preconditioner = true;
            else if (k == "tl_use_jacobi") { d.use_cheby = d.use_cg = d.use_ppcg = d.ppcg_mixed = d.use_fcg = false; d.use_jacobi = true; }
            else if (k == "tl_use_cg") { d.use_cheby = d.use_ppcg = d.use_jacobi = d.ppcg_mixed = d.use_fcg = false; d.use_cg = true; }
            else if (k == "tl_use_ppcg") { d.use_cheby = d.use_cg = d.use_jacobi = d.ppcg_mixed = d.use_fcg = false; d.use_ppcg = true; }
            else if (k == "tl_use_fcg_learned") { d.use_cheby = d.use_cg = d.use_jacobi = d.ppcg_mixed = false; d.use_ppcg = d.use_fcg = true; }
            else if (k == "tl_use_ppcg_mixed") { d.use_cheby = d.use_cg = d.use_jacobi = d.use_fcg = false; d.use_ppcg = d.ppcg_mixed = true; }
            else if (k == "tl_ppcg_mixed_replace") d.mixed_replace = R();
            else if (k == "tl_ppcg_true_residual") d.ppcg_true_residual = true;
            else if (k == "tl_no_lookahead") d.lookahead = false;
            else if (k == "tl_fcg_smooth") {
                d.fcg.smooth = I();
                smooth_set = true;
            }
            else if (k == "tl_fcg_omega") {
                // a fixed weight instead of the Chebyshev weights, with the
                // 4 sweeps the fixed weight was tuned for unless set
                d.fcg.omega = R();
                d.fcg.cheby_ratio = 0.0;
                if (!smooth_set) d.fcg.smooth = 4;
                // beyond 1 the sweeps stop being an SPD preconditioner (and can diverge)
                if (!(d.fcg.omega > 0.0 && d.fcg.omega <= 1.0)) throw Error("tl_fcg_omega must lie in (0, 1]");
            }
            else if (k == "tl_fcg_cheby") d.fcg.cheby_ratio = -1.0;  // the ratio from the operator (default)
            else if (k == "tl_fcg_cheby_ratio") {
                d.fcg.cheby_ratio = R();
                if (!(d.fcg.cheby_ratio > 1.0)) throw Error("tl_fcg_cheby_ratio must exceed 1");
            }
            else if (k == "tl_fcg_no_network") d.fcg.network = false;
            else if (k == "tl_fcg_net_every") d.fcg.net_every = I();
            else if (k == "tl_fcg_net_iters") d.fcg.net_iters = I();
            else if (k == "tl_fcg_net_restart") d.fcg.net_restart = true;
            else if (k == "tl_fcg_coarse") d.fcg.coarse = true;
            else if (k == "tl_fcg_coarse_post") d.fcg.coarse = true, d.fcg.coarse_pre = false;
            else if (k == "tl_use_chebyshev") { d.use_cg = d.use_ppcg = d.use_jacobi = d.ppcg_mixed = d.use_fcg = false; d.use_cheby = true; }
            else if (k == "tl_max_iters") d.max_iters = I();
            else if (k == "tl_eps") d.eps = R();
            else if (k == "tl_coefficient_density") d.coefficient = CONDUCTIVITY;
            else if (k == "tl_coefficient_inverrse_density") d.coefficient = RECIP_CONDUCTIVITY;
            else if (k == "test_problem") d.test_problem = I();
            else if (k == "state") {
                const int s = I();
                if (s < 1 || s > state_max) throw Error("state " + std::to_string(s) + " out of range");
                State& st = d.states[s - 1];
                if (st.defined) throw Error("State defined twice.");
                st.defined = true;
                for (std::string q = w.word(false); !q.empty(); q = w.word(false)) {
                    if (q == "xmin") st.xmin = R();
                    else if (q == "ymin") st.ymin = R();
                    else if (q == "xmax") st.xmax = R();
                    else if (q == "ymax") st.ymax = R();
#if TL_DIM == 3
                    else if (q == "zmin") st.zmin = R();
                    else if (q == "zmax") st.zmax = R();
#endif
                    else if (q == "radius") st.radius = R();
                    else if (q == "density") st.density = R();
                    else if (q == "energy") st.energy = R();
                    else if (q == "geometry") {
                        const std::string gname = w.word(true);
#if TL_DIM == 3
                        if (gname == "cuboid") st.geometry = G_RECT;
#else
                        if (gname == "rectangle") st.geometry = G_RECT;
#endif
                        else if (gname == "circle") st.geometry = G_CIRC;
                        else if (gname == "point") st.geometry = G_POINT;
                    }
                }
            }
        }
    }
    // State extents pulled a hundredth of a cell inwards (read_input.f90:284-297).
    const double dx = (d.xmax - d.xmin) / (double)(float)d.x_cells;
    const double dy = (d.ymax - d.ymin) / (double)(float)d.y_cells;
    const double dz = (d.zmax - d.zmin) / (double)(float)d.z_cells;
    for (size_t n = 1; n < d.states.size(); ++n) {
        State& s = d.states[n];
        s.xmin = s.xmin + (dx / 100.0);
        s.ymin = s.ymin + (dy / 100.0);
        s.xmax = s.xmax - (dx / 100.0);
        s.ymax = s.ymax - (dy / 100.0);
        if (TL_DIM == 3) {
            s.zmin = s.zmin + (dz / 100.0);
            s.zmax = s.zmax - (dz / 100.0);
        }
    }
    return d;
}

// ---- eigenvalue estimate and Chebyshev coefficients -----------------------

// tqli (tea_leaf_cheby.f90:233-283): implicit-shift QL on the symmetric
// tridiagonal (d, e), 1-based, e(2..n) the sub-diagonal before the shift.
int tqli(std::vector<double>& d, std::vector<double>& e, int n)
{
    for (int i = 1; i < n; ++i) e[i] = e[i + 1];
    if (n >= 1) e[n] = 0.0;
    for (int l = 1; l <= n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m <= n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (iter == 30) return 1;
            ++iter;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::sqrt(g * g + 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(std::fabs(r), g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool restart = false;
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i];
                const double b = c * e[i];
                r = std::sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = d[i + 1] - p;
                    e[m] = 0.0;
                    restart = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
            }
            if (restart) continue;
            d[l] = d[l] - p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    return 0;
}

// tea_calc_eigenvalues (tea_leaf_cheby.f90:285-332), CG coefficients 1..m.
int calc_eigenvalues(const std::vector<double>& a, const std::vector<double>& b, int m, double& emin, double& emax)
{
    std::vector<double> diag(m + 2, 0.0), off(m + 2, 0.0);
    for (int n = 1; n <= m; ++n) {
        diag[n] = 1.0 / a[n];
        if (n > 1) diag[n] = diag[n] + b[n - 1] / a[n - 1];
        if (n < m) off[n + 1] = std::sqrt(b[n]) / a[n];
    }
    if (tqli(diag, off, m)) return 1;
    std::vector<double> v(diag.begin() + 1, diag.begin() + 1 + m);
    std::sort(v.begin(), v.end());
    emin = v.front();
    emax = v.back();
    return (emin < 0.0 || emax < 0.0) ? 1 : 0;
}

// tea_calc_ch_coefs (tea_leaf_cheby.f90:334-360), coefficients 1..n.
void calc_ch_coefs(std::vector<double>& ca, std::vector<double>& cb, double emin, double emax, double& theta, int n)
{
    theta = (emax + emin) / 2;
    const double delta = (emax - emin) / 2;
    const double sigma = theta / delta;
    double rho_old = 1.0 / sigma;
    for (int k = 1; k <= n; ++k) {
        const double rho_new = 1.0 / (2.0 * sigma - rho_old);
        ca[k] = rho_new * rho_old;
        cb[k] = 2.0 * rho_new / delta;
        rho_old = rho_new;
    }
}

// ---- tea.out lines ----------------------------------------------------------
//
// The reference writes its per-step diagnostics and field summaries to
// tea.out with fixed Fortran edit descriptors (field_summary.f90:44,81,148;
// tea_solve.f90:216-218,535-543,744-745).  These produce the same text as a
// Fortran runtime for Aw, Iw, Fw.d and Ew.d: right-justified in w columns,
// w asterisks when the value does not fit, Ew.d as 0.d1..dd E+ee with the
// E dropped for a three-digit exponent, NaN and Inf spelled as they are.

namespace fortran {

std::string edit_A(const std::string& t, int w)
{
    if ((int)t.size() >= w) return t.substr(0, (size_t)w);
    return std::string((size_t)w - t.size(), ' ') + t;
}

std::string stars(int w) { return std::string((size_t)std::max(w, 0), '*'); }

std::string edit_I(long v, int w)
{
    const std::string t = std::to_string(v);
    return (int)t.size() > w ? stars(w) : edit_A(t, w);
}

std::string non_finite(double v, int w)
{
    const std::string t = std::isnan(v) ? "NaN" : (v < 0 ? "-Inf" : "Inf");
    return (int)t.size() > w ? stars(w) : edit_A(t, w);
}

std::string edit_F(double v, int w, int d)
{
    if (!std::isfinite(v)) return non_finite(v, w);
    char buf[400];
    std::snprintf(buf, sizeof buf, "%.*f", d, v);
    const std::string t = buf;
    return (int)t.size() > w ? stars(w) : edit_A(t, w);
}

std::string edit_E(double v, int w, int d)
{
    if (!std::isfinite(v)) return non_finite(v, w);
    // d significant digits, correctly rounded, as d1.d2..dd e+XX
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(v));
    std::string digits;
    const char* q = buf;
    for (; *q && *q != 'e'; ++q)
        if (*q != '.') digits += *q;
    int ex = v == 0.0 ? 0 : std::atoi(q + 1) + 1;  // 0.d1d2.. x 10^ex
    char eb[16];
    if (std::abs(ex) <= 99)
        std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    else
        std::snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
    const std::string sign = std::signbit(v) ? "-" : "";
    std::string t = sign + "0." + digits + eb;
    if ((int)t.size() > w) t = sign + "." + digits + eb;  // the leading zero is optional
    return (int)t.size() > w ? stars(w) : edit_A(t, w);
}

}  // namespace fortran

// ---- VTK dumps (visit.f90:20-112) -----------------------------------------
//
// At every visit_frequency-th step, and again when the run completes, the
// reference appends the chunk's file name to tea.visit (started with an
// "!NBLOCKS" line on the first call) and writes tea.<chunk>.<step>.vtk: a
// rectilinear-grid header, the vertex coordinates in E12.4 and the headers
// of two fields whose values it never writes.  The 2D reference leaves the
// third DIMENSIONS entry uninitialised (tealeaf_2d/visit.f90:95); 1 is
// written there.  Files go to the working directory, as the reference's do.
std::string visit_tag(int n)
{
    std::string t = fortran::edit_I(n + 100000, 6);  // i6, the leading digit replaced by '.'
    t[0] = '.';
    return t;
}

void write_visit(const Deck& d, double dx, double dy, double dz, int step, bool& first_call)
{
    if (first_call) {
        FILE* f = std::fopen("tea.visit", "w");
        if (!f) throw Error("cannot write tea.visit");
        std::fprintf(f, "!NBLOCKS %s\n", fortran::edit_I(1, 5).c_str());
        std::fclose(f);
        first_call = false;
    }
    const std::string name = "tea" + visit_tag(1) + visit_tag(step) + ".vtk";
    FILE* f = std::fopen("tea.visit", "a");
    if (!f) throw Error("cannot append to tea.visit");
    std::fprintf(f, "%s\n", name.c_str());
    std::fclose(f);
    const long nxv = d.x_cells + 1, nyv = d.y_cells + 1, nzv = TL_DIM == 3 ? d.z_cells + 1 : 1;
    std::string t = "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET RECTILINEAR_GRID\n";
    t += "DIMENSIONS" + fortran::edit_I(nxv, 12) + fortran::edit_I(nyv, 12) + fortran::edit_I(nzv, 12) + "\n";
    auto axis = [&](const char* label, long n, double lo, double h) {
        t += std::string(label) + fortran::edit_I(n, 5) + " double\n";
        for (long j = 0; j < n; ++j) t += fortran::edit_E(lo + h * (double)j, 12, 4) + "\n";  // vertexx(j)
    };
    axis("X_COORDINATES ", nxv, d.xmin, dx);
    axis("Y_COORDINATES ", nyv, d.ymin, dy);
#if TL_DIM == 3
    axis("Z_COORDINATES ", nzv, d.zmin, dz);
#else
    (void)dz;
#endif
    const long nv = nxv * nyv * nzv;
    t += "FIELD FieldData 3\n";
    t += "x_vel 1 " + fortran::edit_I(nv, 20) + " double\n";
    t += "y_vel 1 " + fortran::edit_I(nv, 20) + " double\n";
    FILE* v = std::fopen(name.c_str(), "w");
    if (!v) throw Error("cannot write " + name);
    std::fwrite(t.data(), 1, t.size(), v);
    std::fclose(v);
}

// One summary in tea.out (field_summary.f90:44-45,81): the column heads and
// the ' step:' line; the list-directed 'Time' line is left out.
void tea_out_summary(std::string& out, int step, double vol, double mass, double ie, double temp)
{
    using namespace fortran;
    out += edit_A("           ", 13) + edit_A("Volume", 16) + edit_A("Mass", 16) + edit_A("Density", 16) +
           edit_A("Energy", 16) + edit_A("U", 16) + "\n";
    out += edit_A(" step:", 6) + edit_I(step, 7) + edit_E(vol, 16, 7) + edit_E(mass, 16, 7) +
           edit_E(mass / vol, 16, 7) + edit_E(ie, 16, 7) + edit_E(temp, 16, 7) + "\n";
}

// The reference's QA line for its test problems 6-23 (field_summary.f90:86-155).
void tea_out_qa(std::string& out, int test_problem, double temp)
{
    // expected final temperatures of the shipped test problems
    static const double expected[24] = {0, 0, 0, 0, 0, 0,
                                        103.88639125996923, 177.00175203427580, 139.97102840522848,
                                        118.33195538063963, 108.12687625327908, 103.46970921076068,
                                        101.21009327612924, 100.16633923123922, 99.695059768963887,
                                        99.756584330021866, 106.22536751207716, 103.75236286020100,
                                        102.95322957172701, 102.61257379535277, 101.99792927438858,
                                        101.93825682423312, 101.54636465546844, 101.51302189154072};
    if (test_problem < 6 || test_problem > 23) return;
    const double qa_diff = std::fabs((100.0 * (temp / expected[test_problem])) - 100.0);
    out += "Test problem" + fortran::edit_I(test_problem, 4) + " is within" + fortran::edit_E(qa_diff, 16, 7) +
           "% of the expected solution\n";
    out += qa_diff < 0.001 ? " This test is considered PASSED\n" : " This test is considered NOT PASSED\n";
}

// ---- solve ----------------------------------------------------------------

struct StepInfo {
    int iters = 0, switch_step = 0, est_itc = 0, cg_steps = 0, cheby_steps = 0;
    double error = 0, exact_error = 0, eigmin = 0, eigmax = 0, solve_s = 0;
    // mixed-precision PPCG: FP32 outer iterations, residual replacements,
    // whether the FP64 inner steps took over
    int mixed_steps = 0, replacements = 0;
    bool mixed_fallback = false;
    // learned flexible CG: its iterations, and whether FP64 PPCG took over
    int fcg_iters = 0;
    bool fcg = false, fcg_fallback = false;
    double fcg_ratio = 0.0;  // the sweeps' Chebyshev interval ratio (0: a fixed weight)
    // iterations launched before the previous rrn was read, and how many of
    // them the stop flag turned into no-ops
    int ahead = 0, ahead_undone = 0;
    std::string tea_out;  // the step's tea.out lines
};

const int kAllExternal[6] = {EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE};

void halo(Chunk& c, std::initializer_list<int> which, int depth)
{
    int fields[NUM_FIELDS] = {0, 0, 0, 0, 0, 0};
    for (int f : which) fields[f - 1] = 1;
    op_update_halo(c, kAllExternal, fields, depth);
}

// tea_leaf (tea_solve.f90:36-605) for one chunk.  The CG and PPCG scalars
// (rro, pw, alpha, rrn, beta, error) live on the device and follow the
// Fortran recurrences there; the host reads `error` once per iteration for
// the convergence and switch tests.  CGCalcP is deferred into the next
// CGCalcW (`p_pending`) and applied on its own only where r changes first
// (the residual recomputed at the PPCG switch); in 2D so is the u half of
// CGCalcUr (`u_pending`), applied on its own at the switch and at the end.
StepInfo tea_leaf(Chunk& c, const Deck& d, double dt)
{
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    StepInfo si;
    const int max_iters = d.max_iters;
    double error = 1e10, exact_error = 0.0;
    int cheby_calc_steps = 0, cg_calc_steps = 0;
    double eigmin = 0.0, eigmax = 0.0, theta = 0.0, cn = 0.0;
    int est_itc = 0, max_cheby_iters = 0, switch_step = 0;
    bool p_pending = false;
    // the next CG iteration is the first since the Chebyshev switch: the
    // reference has refreshed only u's halo since (tea_solve.f90:203-204),
    // so its CGCalcW reads p's depth-1 halo of the last CG iteration
    bool stale_p_halo = false;
    // CG iterations leave u += alpha p to the next iteration's p, w sweep:
    // u_pending marks an update not yet applied
    const bool defer_u = true;
    bool u_pending = false;
    // mixed-precision PPCG state: rrn at the last residual replacement, and
    // true-residual checks that failed without halving the true rrn
    // mixed_form: the solve uses residual replacement and accepts only the
    // true residual; mixed: the FP32 inner steps are (still) in use
    const bool mixed_ok = d.use_ppcg && d.ppcg_mixed && ppcg_mixed_supported(c, d.ppcg_inner_steps);
    const bool mixed_form = mixed_ok || (d.use_ppcg && !d.use_fcg && d.ppcg_true_residual);
    bool mixed = mixed_ok;
    double rrn_replaced = 0.0, last_true = HUGE_VAL;
    int stalls = 0;
    std::vector<double> cg_alphas, cg_betas;
    std::vector<double> ch_alphas(max_iters + 3, 0.0), ch_betas(max_iters + 3, 0.0);
    const bool want_history = d.use_cheby || d.use_ppcg;

    // One iteration ahead.  A CG iteration before any switch, or an FP64
    // PPCG outer iteration of the temporal-blocked form, launches the next
    // iteration of its kind before the host reads its rrn, so the GPU never
    // waits for the host between them.  The rrn's epilogue sets the device
    // stop flag exactly when the host's tests (convergence, the switch rule
    // of tea_solve.f90:173-179) make the next iteration something else; the
    // launched iteration then does nothing, and the host undoes its swaps of
    // the buffers.  Iteration counts and results are those of the
    // one-at-a-time loop.
    const bool pipe = d.lookahead && !c.single_pass && !solve_hooks().on_cg && !solve_hooks().on_outer;
    bool ahead = false;  // iteration n's kernels were launched during iteration n - 1
    struct Bufs {
        double *p, *p_alt, *r, *r_alt, *sd, *sd_alt;
        KernelTimer::Mark mark;
    };
    auto save_bufs = [&] { return Bufs{c.f[F_P], c.p_alt, c.f[F_R], c.r_alt, c.f[F_SD], c.sd_alt, c.timer.mark()}; };
    // a speculative iteration that did nothing: its swaps undone, its timer pairs dropped
    auto restore_bufs = [&](const Bufs& b) {
        c.f[F_P] = b.p;
        c.p_alt = b.p_alt;
        c.f[F_R] = b.r;
        c.r_alt = b.r_alt;
        c.f[F_SD] = b.sd;
        c.sd_alt = b.sd_alt;
        c.timer.discard_since(b.mark);
    };
    auto rrn_slot = [](int k) { return (k & 1) ? S_RRN2 : S_RRN; };
    // the error at or below which iteration k is a switch to Chebyshev or PPCG
    // rather than a CG iteration (while no switch has happened)
    auto switch_le = [&](int k) {
        if (!(d.use_cheby || d.use_ppcg)) return -HUGE_VAL;
        if (d.errswitch) return k >= 20 ? d.epslim : -HUGE_VAL;
        return k >= d.presteps ? 1.0 : -HUGE_VAL;
    };
    // after a PPCG outer iteration the next is a CG iteration when not
    // error <= 1 (the switch test of tea_solve.f90:176-177 fails again);
    // under errswitch it is always PPCG
    const double ppcg_gt = d.errswitch ? HUGE_VAL : 1.0;
    auto ppcg_stops = [&](double e) { return std::fabs(e) < d.eps || (ppcg_gt < HUGE_VAL && !(e <= ppcg_gt)); };

    halo(c, {FIELD_ENERGY1, FIELD_DENSITY}, 2);
    if (d.use_cg || d.use_cheby || d.use_ppcg) {
        (void)op_cg_init(c, d.coefficient, d.preconditioner, dt);
        if (want_history) c.ensure_history(max_iters + 1);
    } else if (d.use_jacobi) {
        op_jacobi_init(c, d.coefficient, dt);
    }
    op_copy_u(c);

    int n = 1;
    Learner* learner = d.use_fcg ? active_learner() : nullptr;
    if (d.use_fcg && d.fcg.network && !learner) throw Error("tl_use_fcg_learned: no learned preconditioner is loaded");
    bool fcg_done = false;
    if (d.use_fcg) {
        int it = 0;
        double err = 0.0;
        si.fcg = true;
        FcgOptions o = d.fcg;
        if (TL_DIM == 2 && o.smooth > 0 && o.cheby_ratio < 0.0) o.cheby_ratio = fcg_auto_ratio(c, o.smooth);
        si.fcg_ratio = o.cheby_ratio;
        fcg_done = fcg_learned_solve(c, learner, o, d.eps, max_iters, it, err);
        si.fcg_iters = it;
        if (fcg_done) {
            error = err;
            n = it + 1;
        } else {
            // the flexible CG stalled: solve this timestep with FP64 PPCG from the start
            si.fcg_fallback = true;
            halo(c, {FIELD_ENERGY1, FIELD_DENSITY}, 2);
            (void)op_cg_init(c, d.coefficient, d.preconditioner, dt);
            op_copy_u(c);
        }
    }

    for (n = fcg_done ? n : 1; !fcg_done && n <= max_iters; ++n) {
        bool ch_switch_check;
        if (d.errswitch)
            ch_switch_check = (cheby_calc_steps > 0) || ((error <= d.epslim) && (n >= 20));
        else
            ch_switch_check = (n >= d.presteps) && (error <= 1.0);

        if ((d.use_cheby || d.use_ppcg) && ch_switch_check) {
            if (u_pending) {
                op_cg_calc_u_dev(c);
                u_pending = false;
            }
            if (cheby_calc_steps == 0) {
                max_cheby_iters = max_iters - n + 2;
                op_read_history(c, n - 1, cg_alphas, cg_betas);
                if (calc_eigenvalues(cg_alphas, cg_betas, n - 1, eigmin, eigmax))
                    throw Error("tea_leaf: Error in calculating eigenvalues");
                eigmin = eigmin * (double)0.95f;
                eigmax = eigmax * (double)1.05f;
                calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, theta,
                              d.use_cheby ? max_cheby_iters : d.ppcg_inner_steps);
                cn = eigmax / eigmin;
                // tea_solve.f90:216-218 (rro = error at the switch)
                using namespace fortran;
                si.tea_out += "Switching after " + edit_I(n, 3) + " steps, error " + edit_E(error, 15, 7) + "\n";
                si.tea_out += edit_A("eigmin", 11) + edit_A("eigmax", 11) + edit_A("cn", 11) + edit_A("error", 11) + "\n";
                si.tea_out += edit_F(eigmin, 11, 5) + edit_F(eigmax, 11, 5) + edit_E(cn, 11, 4) + edit_E(error, 11, 4) +
                              "\n";
            }
            if (d.use_cheby) {
                if (cheby_calc_steps > 0 && p_pending) {
                    // a CG iteration ran since the last Chebyshev step (the
                    // error was above 1): ChebyIterate reads its new p
                    op_cg_calc_p_dev(c);
                    p_pending = false;
                }
                if (cheby_calc_steps == 0) {
                    // tea_leaf_cheby_first_step (tea_solve.f90:658-750); the
                    // pending p update is dead in the mesh (ChebyInit
                    // overwrites p) but lives on in the halo of p, which a
                    // CG iteration after this step would read
                    op_capture_p_faces(c, p_pending);
                    stale_p_halo = true;
                    p_pending = false;
                    const double bb = op_calc_2norm(c, false);
                    op_cheby_init(c, &ch_alphas[1], &ch_betas[1], max_cheby_iters, theta, d.preconditioner, true);
                    op_cheby_iterate(c, max_cheby_iters, true);
                    error = op_calc_2norm(c, true);
                    const double it_alpha = DBL_EPSILON * bb / (4.0 * error);
                    const double gamm = (std::sqrt(cn) - 1.0) / (std::sqrt(cn) + 1.0);
                    est_itc = (int)std::lround(std::log(it_alpha) / (2.0 * std::log(gamm)));
                    si.tea_out += fortran::edit_A("est itc", 11) + "\n" + fortran::edit_I(est_itc, 11) + "\n";
                    cheby_calc_steps = 1;
                    switch_step = n;
                } else if (cheby_tb_supported(c)) {
                    // iterations n .. m in one temporal-blocked sweep, m the
                    // next that checks ||r||^2 (or the last allowed): nothing
                    // between them reads a global result
                    int m = n;
                    while (m < max_iters && !((m >= est_itc) && (m % 10 == 0))) ++m;
                    const bool check = (m >= est_itc) && (m % 10 == 0);
                    op_cheby_tb(c, cheby_calc_steps, m - n + 1, check);
                    if (check) error = c.result(S_NORM);
                    cheby_calc_steps += m - n;
                    n = m;
                } else {
                    op_cheby_iterate(c, cheby_calc_steps, true);
                    if ((n >= est_itc) && (n % 10 == 0)) error = op_calc_2norm(c, true);
                }
            } else {
                if (cheby_calc_steps == 0) {
                    cheby_calc_steps = 1;
                    op_ppcg_init(c, &ch_alphas[1], &ch_betas[1], max_cheby_iters, theta);
                    if (p_pending) {
                        op_cg_calc_p_dev(c);
                        p_pending = false;
                    }
                    op_calc_residual(c);
                    switch_step = n;
                    if (mixed) op_mixed_prepare(c, &ch_alphas[1], &ch_betas[1], d.ppcg_inner_steps, theta);
                    if (mixed_form) rrn_replaced = error;
                }
                // rro on the device is the last rrn (= error), as at tea_solve.f90:187,304
                const bool tb = !mixed && ppcg_tb_supported(c, d.ppcg_inner_steps);
                const bool spec = pipe && tb && !mixed_form;
                if (spec && ahead) {
                    // launched during the previous iteration
                } else if (mixed) {
                    op_cg_calc_w_dev(c, p_pending);
                    op_ppcg_outer_mixed(c, d.ppcg_inner_steps);
                    ++si.mixed_steps;
                } else if (tb) {
                    // calc_ur, init_sd, the inner steps and ||r||^2 in one sweep
                    op_cg_calc_w_dev(c, p_pending);
                    if (spec) {
                        c.stop_eps = d.eps;
                        c.stop_gt = ppcg_gt;
                    }
                    op_ppcg_outer_tb(c, d.ppcg_inner_steps, rrn_slot(n));
                } else {
                    op_cg_calc_w_dev(c, p_pending);
                    op_ppcg_calc_ur_dev(c);
                    // tea_leaf_run_ppcg_inner_steps (tea_solve.f90:607-656)
                    op_ppcg_init_sd(c);
                    for (int step = 1; step <= d.ppcg_inner_steps;) {
                        if (TL_DIM == 3 && step < d.ppcg_inner_steps) {
                            op_ppcg_inner2(c, step);  // two inner steps per sweep
                            step += 2;
                        } else {
                            op_ppcg_inner(c, step++, true);
                        }
                    }
                    op_2norm_rrn_dev(c);
                }
                p_pending = true;
                if (spec) {
                    ahead = false;
                    Bufs bufs{};
                    const bool next = n + 1 <= max_iters;
                    if (next) {
                        bufs = save_bufs();
                        c.gate = true;
                        c.stop_eps = d.eps;
                        c.stop_gt = ppcg_gt;
                        op_cg_calc_w_dev(c, true);
                        op_ppcg_outer_tb(c, d.ppcg_inner_steps, rrn_slot(n + 1));
                        c.gate = false;
                    }
                    error = c.result(rrn_slot(n));
                    if (next) {
                        ++si.ahead;
                        if (ppcg_stops(error)) {
                            restore_bufs(bufs);
                            ++si.ahead_undone;
                        } else {
                            ahead = true;
                        }
                    }
                    c.stop_eps = -1.0;
                    c.stop_gt = HUGE_VAL;
                } else if (tb) {
                    error = c.result(rrn_slot(n));
                } else {
                    error = op_error(c);
                }
                if (solve_hooks().on_outer) solve_hooks().on_outer(c, cheby_calc_steps);
                if (mixed_form) {
                    const bool converged = std::fabs(error) < d.eps;
                    if (converged || error < rrn_replaced * d.mixed_replace) {
                        // FP64 residual replacement r = u0 - A u; beta is
                        // recomputed from it.  Convergence is accepted only
                        // on the true residual.
                        op_calc_residual(c);
                        op_2norm_replace_dev(c);
                        error = op_error(c);
                        rrn_replaced = error;
                        ++si.replacements;
                        if (converged && !(std::fabs(error) < d.eps)) {
                            // checks that keep failing without 10% progress:
                            // the FP32 steps have stalled, finish in FP64
                            stalls = error > 0.9 * last_true ? stalls + 1 : 0;
                            last_true = error;
                            if (mixed && stalls >= 5) {
                                mixed = false;
                                stalls = 0;
                                si.mixed_fallback = true;
                            }
                        }
                    }
                }
            }
            ++cheby_calc_steps;
        } else if (d.use_cg || d.use_cheby || d.use_ppcg) {
            ++cg_calc_steps;
            bool via_strip = false;  // the p, w strip and calc_r path (rrn in rrn_slot(n))
            if (stale_p_halo) {
                // p and u are up to date here: Chebyshev steps ran since the
                // last CG iteration
                op_cg_calc_w_faces_dev(c);
                stale_p_halo = false;
                op_cg_calc_ur_dev(c, want_history ? n : 0);
            } else if (defer_u) {
                // CGCalcUr's u += alpha p rides on the next iteration's p, w
                // strip, which reads p anyway (bitwise the same update)
                via_strip = true;
                if (!ahead) {
                    c.stop_eps = d.eps;
                    c.stop_le = switch_le(n + 1);
                    op_cg_calc_w_dev(c, p_pending, u_pending);
                    op_cg_calc_r_dev(c, want_history ? n : 0, rrn_slot(n));
                }
                u_pending = true;
            } else {
                op_cg_calc_w_dev(c, p_pending);
                op_cg_calc_ur_dev(c, want_history ? n : 0);
            }
            p_pending = true;
            if (via_strip) {
                ahead = false;
                Bufs bufs{};
                const bool next = pipe && cheby_calc_steps == 0 && n + 1 <= max_iters;
                if (next) {
                    bufs = save_bufs();
                    c.gate = true;
                    c.stop_eps = d.eps;
                    c.stop_le = switch_le(n + 2);
                    op_cg_calc_w_dev(c, true, true);
                    op_cg_calc_r_dev(c, want_history ? n + 1 : 0, rrn_slot(n + 1));
                    c.gate = false;
                }
                error = c.result(rrn_slot(n));
                if (next) {
                    // the test the rrn's epilogue applied to set the stop flag
                    ++si.ahead;
                    if (std::fabs(error) < d.eps || error <= switch_le(n + 1)) {
                        restore_bufs(bufs);
                        ++si.ahead_undone;
                    } else {
                        ahead = true;
                    }
                }
                c.stop_eps = -1.0;
                c.stop_le = -HUGE_VAL;
            } else {
                error = op_error(c);
            }
            if (solve_hooks().on_cg) solve_hooks().on_cg(c, n);
        } else if (d.use_jacobi && jacobi_tb_supported(c)) {
            // iterations n .. n + K - 1 in one temporal-blocked pass, ending
            // at the next multiple of 50 (the residual check) at the latest;
            // if an iteration inside it meets eps, the pass is run again from
            // its input up to that iteration
            const int next50 = (n + 49) / 50 * 50;
            int K = jacobi_tb_pass(std::min(max_iters, next50) - n + 1);
            std::vector<double> err = op_jacobi_tb(c, K);
            int k = 0;
            while (k < K && !(std::fabs(err[k]) < d.eps)) ++k;
            if (k < K - 1) {
                op_jacobi_tb_undo(c);
                for (int left = k + 1; left > 0;) {
                    const int m = jacobi_tb_pass(left);
                    (void)op_jacobi_tb(c, m);
                    left -= m;
                }
                K = k + 1;
            }
            error = err[K - 1];
            n += K - 1;
            if (n % 50 == 0) {
                op_calc_residual(c);
                error = op_calc_2norm(c, true);
            }
        } else if (d.use_jacobi) {
            error = op_jacobi_solve(c, true);
            if (n % 50 == 0) {
                op_calc_residual(c);
                error = op_calc_2norm(c, true);
            }
        }
        if (std::fabs(error) < d.eps) break;
    }
    if (u_pending) {
        op_cg_calc_u_dev(c);
        u_pending = false;
    }
    if (solve_hooks().on_end) solve_hooks().on_end(c);
    if (d.check_result) {
        op_calc_residual(c);
        exact_error = op_calc_2norm(c, true);
    }
    // tea_solve.f90:535-543
    si.tea_out += "Conduction error " + fortran::edit_E(error, 14, 7) + "\n";
    if (d.check_result) si.tea_out += "EXACT error calculated as" + fortran::edit_E(exact_error, 14, 7) + "\n";
    si.tea_out += "Iteration count " + fortran::edit_I(n - 1, 8) + "\n";
    si.iters = n - 1;
    si.error = error;
    si.exact_error = exact_error;
    si.switch_step = switch_step;
    si.est_itc = est_itc;
    si.eigmin = eigmin;
    si.eigmax = eigmax;
    si.cg_steps = cg_calc_steps;
    si.cheby_steps = cheby_calc_steps;
    op_finalise(c);
    halo(c, {FIELD_ENERGY1}, 1);
    c.sync();
    si.solve_s = std::chrono::duration<double>(clock::now() - t0).count();
    return si;
}

struct Json {
    std::string s;
    void f(const char* fmt, ...)
    {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        s += buf;
    }
};

// field_summary.f90: the totals into the report and their tea.out lines;
// returns temp
double summary(Json& o, std::string& tea_out, Chunk& c, int step, double time)
{
    double vol, mass, ie, temp;
    op_field_summary(c, &vol, &mass, &ie, &temp);
    o.f("\"step\": %d, \"time\": %.17g, \"vol\": %.17g, \"mass\": %.17g, \"ie\": %.17g, \"temp\": %.17g", step, time,
        vol, mass, ie, temp);
    tea_out_summary(tea_out, step, vol, mass, ie, temp);
    return temp;
}

// initialise -> start -> diffuse (start.f90:24-162, diffuse.f90:23-160).
bool g_progress = false;  // a line on stderr after every timestep
std::string g_tea_out;    // the last run's tea.out lines

std::string run_deck(const std::string& text, int max_steps)
{
    const Deck d = parse_deck(text);
    const int nz = TL_DIM == 3 ? d.z_cells : 1;
    auto c = std::make_unique<Chunk>(d.x_cells, d.y_cells, nz, 0);
    const double dx = (d.xmax - d.xmin) / (double)(float)d.x_cells;
    const double dy = (d.ymax - d.ymin) / (double)(float)d.y_cells;
    const double dz = (d.zmax - d.zmin) / (double)(float)d.z_cells;
    op_initialise_chunk(*c, d.xmin, d.ymin, d.zmin, dx, dy, dz);
    const int ns = (int)d.states.size();
    std::vector<double> sd(ns), se(ns), x0(ns), x1(ns), y0(ns), y1(ns), z0(ns), z1(ns), sr(ns);
    std::vector<int> sg(ns);
    for (int k = 0; k < ns; ++k) {
        const State& s = d.states[k];
        sd[k] = s.density; se[k] = s.energy;
        x0[k] = s.xmin; x1[k] = s.xmax; y0[k] = s.ymin; y1[k] = s.ymax; z0[k] = s.zmin; z1[k] = s.zmax;
        sr[k] = s.radius; sg[k] = s.geometry;
    }
    op_generate_chunk(*c, ns, sd.data(), se.data(), x0.data(), x1.data(), y0.data(), y1.data(), z0.data(), z1.data(),
                      sr.data(), sg.data(), G_RECT, G_CIRC, G_POINT);
    halo(*c, {FIELD_DENSITY, FIELD_ENERGY0, FIELD_ENERGY1}, 2);

    Json o;
    o.f("{\"dim\": %d, \"x_cells\": %d, \"y_cells\": %d, \"z_cells\": %d, \"device\": \"hip\", ", TL_DIM, d.x_cells,
        d.y_cells, nz);
    o.f("\"reductions\": \"%s\", ", c->single_pass ? "single-pass" : "separate finish");
    o.f("\"solver\": \"%s\", \"eps\": %.17g, \"dtinit\": %.17g, \"max_iters\": %d, ",
        d.use_cg ? "cg" : d.use_cheby ? "chebyshev" : d.use_ppcg ? (d.use_fcg ? "fcg_learned" : d.ppcg_mixed ? "ppcg_mixed" : "ppcg") : "jacobi",
        d.eps, d.dtinit, d.max_iters);
    std::string& tea_out = g_tea_out;
    tea_out.clear();
    o.f("\"initial\": {");
    summary(o, tea_out, *c, 0, 0.0);
    o.f("}, \"steps\": [");

    op_set_field(*c);
    double time = 0.0, dt = d.dtinit, total = 0.0;
    const double g_small = (double)1.0e-16f;
    int step = 0;
    bool visit_first = true;
    for (;;) {
        ++step;
        c->timer.on = kernel_timing_from() > 0 && step >= kernel_timing_from();
        c->timer.only = kernel_timing_only();
        if (d.dtinit <= dt) dt = d.dtinit;  // timestep.f90 / calc_dt.f90:35
        const StepInfo si = tea_leaf(*c, d, dt);
        total += si.solve_s;
        time = time + dt;
        if (g_progress) {
            std::fprintf(stderr, "timestep %d: %d iterations, %.3f s\n", step, si.iters, si.solve_s);
            std::fflush(stderr);
        }
        tea_out += si.tea_out;
        o.f("%s{", step > 1 ? ", " : "");
        o.f("\"iters\": %d, \"error\": %.17g, \"exact_error\": %.17g, \"switch_step\": %d, \"est_itc\": %d, "
            "\"eigmin\": %.17g, \"eigmax\": %.17g, \"cg_steps\": %d, \"cheby_steps\": %d, \"solve_s\": %.9g, "
            "\"dt\": %.17g, \"mixed_steps\": %d, \"replacements\": %d, \"mixed_fallback\": %d, "
            "\"fcg\": %d, \"fcg_iters\": %d, \"fcg_fallback\": %d, \"fcg_ratio\": %.17g, \"ahead\": %d, "
            "\"ahead_undone\": %d, ",
            si.iters, si.error, si.exact_error, si.switch_step, si.est_itc, si.eigmin, si.eigmax, si.cg_steps,
            si.cheby_steps, si.solve_s, dt, si.mixed_steps, si.replacements, si.mixed_fallback ? 1 : 0,
            si.fcg ? 1 : 0, si.fcg_iters, si.fcg_fallback ? 1 : 0, si.fcg_ratio, si.ahead, si.ahead_undone);
        if (d.summary_frequency != 0 && step % d.summary_frequency == 0)
            summary(o, tea_out, *c, step, time);
        else
            o.f("\"step\": %d, \"time\": %.17g", step, time);
        o.f("}");
        if (d.visit_frequency != 0 && step % d.visit_frequency == 0) write_visit(d, dx, dy, dz, step, visit_first);
        const bool complete = time + g_small > d.end_time || step >= d.end_step;
        if (complete && d.visit_frequency != 0) write_visit(d, dx, dy, dz, step, visit_first);  // diffuse.f90:93
        if (complete || (max_steps > 0 && step >= max_steps)) break;
    }
    o.f("], \"final\": {");
    tea_out_qa(tea_out, d.test_problem, summary(o, tea_out, *c, step, time));
    o.f("}, \"test_problem\": %d, \"total_solve_s\": %.9g, \"cells\": %ld", d.test_problem, total,
        (long)d.x_cells * d.y_cells * nz);
    if (kernel_timing_from() > 0) {
        c->sync();
        o.f(", \"kernels\": {");
        bool first = true;
        for (int k = 0; k < K_COUNT; ++k) {
            if (!c->timer.count[k]) continue;
            o.f("%s\"%s\": {\"count\": %ld, \"ms\": %.6f, \"bytes_per_cell\": %.1f}", first ? "" : ", ",
                kernel_name(k), c->timer.count[k], c->timer.ms[k], kernel_bytes_per_cell(k));
            first = false;
        }
        o.f("}");
    }
    o.f("}");
    return o.s;
}

std::string g_driver_error;

}  // namespace

std::string driver_run_deck(const std::string& text, int max_steps) { return run_deck(text, max_steps); }

}  // namespace tl

extern "C" {

// Runs a tea.in deck end to end on the GPU and writes a JSON report (per-step
// iteration counts, solver diagnostics, field summaries, solve seconds) into
// json[0..cap).  max_steps > 0 stops after that many timesteps.  Returns 0,
// -1 on an error (message in tl_driver_error()), -2 if cap is too small.
int tl_run_deck(const char* deck, char* json, int cap, int max_steps)
{
    try {
        tl::g_driver_error.clear();
        const std::string out = tl::run_deck(deck ? deck : "", max_steps);
        if ((int)out.size() + 1 > cap) {
            tl::g_driver_error = "JSON buffer too small: need " + std::to_string(out.size() + 1);
            return -2;
        }
        std::memcpy(json, out.c_str(), out.size() + 1);
        return 0;
    } catch (const std::exception& e) {
        tl::g_driver_error = e.what();
        return -1;
    }
}

const char* tl_driver_error(void) { return tl::g_driver_error.c_str(); }

// The last tl_run_deck's tea.out lines with fixed formats (see the tea.out
// section above), NUL-terminated; empty before the first run.
const char* tl_tea_out(void) { return tl::g_tea_out.c_str(); }

// One value through a Fortran edit descriptor "aW", "iW", "fW.D" or "eW.D"
// (the first of v_int, v_real or text that the letter reads) into out[0..cap).
// W is 1..200 and D at most 40.  Returns the length, -1 for a descriptor it
// does not handle, -2 if cap is too small.
int tl_fortran_edit(const char* desc, long v_int, double v_real, const char* text, char* out, int cap)
{
    if (!desc || !out || !desc[0]) return -1;
    const char k = (char)std::tolower((unsigned char)desc[0]);
    int w = 0, dd = -1;
    if (std::sscanf(desc + 1, "%d.%d", &w, &dd) < 1 || w <= 0 || w > 200 || dd > 40) return -1;
    std::string r;
    if (k == 'a') r = tl::fortran::edit_A(text ? text : "", w);
    else if (k == 'i') r = tl::fortran::edit_I(v_int, w);
    else if (k == 'f' && dd >= 0) r = tl::fortran::edit_F(v_real, w, dd);
    else if (k == 'e' && dd >= 1) r = tl::fortran::edit_E(v_real, w, dd);
    else return -1;
    if ((int)r.size() + 1 > cap) return -2;
    std::memcpy(out, r.c_str(), r.size() + 1);
    return (int)r.size();
}

// Per-launch kernel timing with HIP events on the chunk's stream, for the
// timesteps from `from_step` on (0 turns it off); tl_run_deck then adds a
// "kernels" table (launch count, summed milliseconds, algorithmic bytes per
// cell) to its report.
void tl_set_kernel_timing(int from_step) { tl::set_kernel_timing_from(from_step); }

// Restricts that timing to the kernel of the given report name (NULL or ""
// times every kernel again): two events per launch of one kernel perturb
// the timestep less than two around every launch.  Returns -1 for an
// unknown name.
int tl_set_kernel_timing_only(const char* name)
{
    if (!name || !*name) {
        tl::set_kernel_timing_only(-1);
        return 0;
    }
    const int k = tl::kernel_id(name);
    if (k < 0) return -1;
    tl::set_kernel_timing_only(k);
    return 0;
}

// A line on stderr after every timestep of tl_run_deck (0 turns it off).
void tl_set_progress(int on) { tl::g_progress = on != 0; }

// Chunks created after this call close every reduction inside the sweep
// that computes it (on != 0) or with a separate one-workgroup launch (0, the
// default); both give the same bits.
void tl_set_single_pass(int on) { tl::set_single_pass(on != 0); }

}  // extern "C"
