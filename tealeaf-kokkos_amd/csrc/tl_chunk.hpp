// Host-side view of one TeaLeaf chunk resident in HBM, and the operations
// the ext_* C-ABI and the solver driver run on it.  Each operation is one
// (or, for setup, a few) hand-written gfx950 kernel launches on the chunk's
// stream; operations that reduce return the FP64 total to the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <functional>
#include <vector>

#include "tl_device.hpp"

namespace tl {

// Field ids of the tl_get_field / tl_set_field hooks.
enum FieldId {
    F_DENSITY = 0, F_ENERGY0, F_ENERGY1, F_U, F_U0, F_P, F_R, F_MI, F_W, F_Z,
    F_KX, F_KY, F_KZ, F_SD, F_VOLUME, F_COUNT
};

// Fortran-side codes (tealeaf_2d/ext_shared.hpp:10-29, data.f90:39-64).
enum { CHUNK_LEFT = 1, CHUNK_RIGHT = 2, CHUNK_BOTTOM = 3, CHUNK_TOP = 4, CHUNK_BACK = 5, CHUNK_FRONT = 6 };
enum { FIELD_DENSITY = 1, FIELD_ENERGY0 = 2, FIELD_ENERGY1 = 3, FIELD_U = 4, FIELD_P = 5, FIELD_SD = 6, NUM_FIELDS = 6 };
enum { CONDUCTIVITY = 1, RECIP_CONDUCTIVITY = 2 };
enum { G_RECT = 1, G_CIRC = 2, G_POINT = 3 };
constexpr int EXTERNAL_FACE = -1;
#if TL_DIM == 3
constexpr int NUM_FACES = 6;
#else
constexpr int NUM_FACES = 4;
#endif

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define TL_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw ::tl::Error(std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Result slots of the host-mapped reduction buffer (4 doubles each).
// S_RRN2: the rrn of odd iterations when the driver runs one iteration
// ahead (the host reads iteration n's while n + 1's is being computed).
enum Slot { S_RRO = 0, S_PW, S_RRN, S_NORM, S_ERR, S_SUMMARY, S_RRN2, S_COUNT };

// Kernels timed by the optional per-launch event timer.
enum KernelId {
    K_CG_INIT = 0, K_CG_W, K_CG_WP, K_CG_UR, K_CG_P, K_CHEBY, K_PPCG_SD, K_PPCG_INNER, K_PPCG_OUTER,
    K_JACOBI, K_RESIDUAL, K_2NORM, K_COPY_U, K_FINALISE, K_SUMMARY, K_HALO, K_PPCG_MIXED, K_CHEBY_TB, K_JACOBI_TB,
    K_CG_WPU, K_CG_R, K_CG_U, K_FCG_SMOOTH, K_FCG_STRIP, K_FCG_SMOOTH_R, K_FCG_STRIP_ZR, K_PPCG_INNER2,
    K_COUNT
};
const char* kernel_name(int k);
// Algorithmic HBM bytes per interior cell of one launch (each array read or
// written once), 0 where the kernel is not a full sweep.
double kernel_bytes_per_cell(int k);

// Per-launch timing with HIP events recorded on the chunk's stream; pending
// pairs are resolved at the next host synchronisation.
struct KernelTimer {
    bool on = false;
    int only = -1;  // time only this KernelId (-1: all)
    double ms[K_COUNT] = {};
    long count[K_COUNT] = {};
    struct Pending {
        int k;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    long flushes = 0;
    // Pairs recorded after a mark can be dropped (launches that turned out to
    // do nothing) unless a flush has resolved them since.
    struct Mark {
        long flushes;
        size_t n;
    };
    Mark mark() const { return Mark{flushes, pending.size()}; }
    void discard_since(const Mark& m);
    hipEvent_t take();
    void begin(int k, hipStream_t s, Pending& p);
    void end(Pending& p, hipStream_t s);
    void flush();
    ~KernelTimer();
};
// Timesteps from which the driver turns each chunk's timer on (0 = never),
// and the one kernel it then times (-1: all).
void set_kernel_timing_from(int step);
int kernel_timing_from();
void set_kernel_timing_only(int kid);
int kernel_timing_only();
int kernel_id(const char* name);  // -1 for an unknown name
// Whether chunks created from now on close their reductions inside the
// sweeps (single pass) or with a k_finish launch (the default: faster here).
void set_single_pass(bool on);
bool single_pass_default();

// Runs STMT between two timer events when the chunk's timer is on.
#define TL_TIMED(c, kid, STMT)                               \
    do {                                                     \
        if ((c).timer.on && ((c).timer.only < 0 || (c).timer.only == (kid))) { \
            ::tl::KernelTimer::Pending p_;                   \
            (c).timer.begin((kid), (c).stream, p_);          \
            STMT;                                            \
            (c).timer.end(p_, (c).stream);                   \
        } else {                                             \
            STMT;                                            \
        }                                                    \
    } while (0)

struct Chunk {
    Geom g{};
    int rank = 0;
    double dx = 0, dy = 0, dz = 0;
    double theta = 0;
    bool preconditioner = false;

    double* f[F_COUNT] = {};
    double* u_alt = nullptr;   // ping-pong partner of u (Chebyshev)
    double* sd_alt = nullptr;  // ping-pong partner of sd (PPCG inner steps)
    double* p_alt = nullptr;   // ping-pong partner of p (CGCalcP folded into CGCalcW)
    double* r_alt = nullptr;   // ping-pong partner of r (temporal-blocked PPCG)
    // p of the mesh's boundary cells as the depth-1 halo of p holds it
    // after the last CG iteration before the Chebyshev switch (per face:
    // west, east, south, north[, back, front]); allocated at first use
    double* p_faces = nullptr;

    // Mixed-precision PPCG: FP32 copies of kx, ky (same layout, float
    // elements) and of the inner-step coefficients, made at the switch.
    float* kxy32 = nullptr;  // FP32 kx, ky interleaved: element i at [2i], [2i + 1]
    bool internal_faces = false;  // a face has a neighbour chunk (halo update or message)
    double* d_pack = nullptr;     // staging of halo messages (ext_pack_message_)
    long pack_cap = 0;
    unsigned* d_rowtab = nullptr;  // reflected row byte offsets of the temporal-blocked kernels (cell, face)
    double* d_cheby = nullptr;     // Chebyshev a[0..n) then b[0..n) on the device
    double* d_levels = nullptr;    // per-iteration error sums of a temporal-blocked Jacobi pass
    int d_cheby_n = 0, d_cheby_cap = 0;
    float* d_coef32 = nullptr;  // a[0..n) then b[0..n)
    int mixed_steps = 0;
    double inv_theta = 0.0;

    // Mesh coordinates (device, reference indexing).
    double *vertexX = nullptr, *vertexY = nullptr, *vertexZ = nullptr;
    double *cellX = nullptr, *cellY = nullptr, *cellZ = nullptr;

    // Chebyshev / PPCG coefficients, copied from the caller at init.  The
    // device copy serves the temporal-blocked PPCG passes: for the pass that
    // starts at inner step k with m levels, a[k..k+m) then b[k..k+m) at
    // offset 2k; laid out for d_coef_steps inner steps (0: not yet).
    std::vector<double> alphas, betas;
    double* d_coef = nullptr;
    int d_coef_cap = 0;
    int d_coef_steps = 0;

    // Reduction scratch and device-resident solver scalars.
    double* partials = nullptr;
    DevScalars* ds = nullptr;
    double* hist_a = nullptr;  // CG alphas / betas by iteration (device)
    double* hist_b = nullptr;
    int hist_cap = 0;
    double* d_result = nullptr;           // [S_COUNT * 4] device
    volatile double* h_result = nullptr;  // [S_COUNT * 4] host-mapped
    double* h_result_dev = nullptr;       // device alias of h_result
    // per slot, the number of the last reduction launched and the number
    // its second phase writes (host-mapped) after its result: the host
    // reads a result as soon as the number arrives, without a stream sync
    unsigned fin_seq[S_COUNT] = {};
    volatile unsigned* h_seq = nullptr;
    unsigned* h_seq_dev = nullptr;

    hipStream_t stream = nullptr;
    int simds = 1024;            // SIMDs of the device (CUs x 4): one resident wave each
    long partials_cap = 0;       // block partials the reduction scratch holds
    mutable KernelTimer timer;

    Chunk(int nx, int ny, int nz, int rank);
    ~Chunk();
    Chunk(const Chunk&) = delete;
    Chunk& operator=(const Chunk&) = delete;

    // Single-pass reductions (single_pass): reduce_counters(partials_cap)
    // counters, 0 between reductions, and 4 * partials_cap slots, SLOT_EMPTY
    // between them.  Otherwise each reduction is closed by a k_finish launch.
    unsigned* d_counter = nullptr;
    double* d_slots = nullptr;
    bool single_pass = false;
    Finish pending{};  // the reduction close() launches k_finish for
    bool pending_open = false;

    // Speculation: while `gate` is set, the reductions (and through them the
    // kernels that take them) launched belong to an iteration launched
    // before the previous rrn was read, and do nothing if that rrn stopped
    // the solve; rrn epilogues set ds->stop by stop_eps, stop_le, stop_gt.
    bool gate = false;
    double stop_eps = -1.0, stop_le = -HUGE_VAL, stop_gt = HUGE_VAL;

    // Partials only: an explicit finish closes the reduction.
    Reduce reducer() const
    {
        Finish f{};
        f.ds = ds;
        f.gate = gate ? 1 : 0;
        return Reduce{partials, f};
    }
    // The next reduction of `slot`, with epilogue `ep`: closed inside the
    // sweep that takes it (single_pass), else by close() after its launch.
    Reduce reducer(Slot slot, int ep, int hist_n = -1);
    Finish finish_args(Slot slot, int ep, int hist_n = -1);
    void close(long nblocks, int nv);  // launches k_finish unless the sweep closed the reduction
    double result(Slot s, int k = 0) const;  // waits for the slot's last reduction
    void ensure_history(int n);
    void sync() const;

    // Field access in the reference's flat layout (X*Y*Z doubles).
    void get_field(int id, double* host) const;
    void set_field(int id, const double* host);
};

// ---- operations (tl_kernels.hip) ----------------------------------------
void op_initialise_chunk(Chunk& c, double xmin, double ymin, double zmin, double dx, double dy, double dz);
void op_generate_chunk(Chunk& c, int nstates, const double* density, const double* energy,
                       const double* xmin, const double* xmax, const double* ymin, const double* ymax,
                       const double* zmin, const double* zmax, const double* radius, const int* geometry,
                       int rect, int circ, int point);
void op_set_field(Chunk& c);
void op_field_summary(Chunk& c, double* vol, double* mass, double* ie, double* temp);
// fields[6] flags in FIELD_* order, neighbours[NUM_FACES] (EXTERNAL_FACE marks a reflective face)
void op_update_halo(Chunk& c, const int* neighbours, const int* fields, int depth);
// Pack (pack) or unpack one face's halo message of the flagged fields to or
// from buffer + offsets[field]: ext_pack_message_.
void op_pack_message(Chunk& c, const int* fields, const int* offsets, int depth, int face, double* buffer, bool pack);

double op_cg_init(Chunk& c, int coefficient, bool preconditioner, double dt);
double op_cg_calc_w(Chunk& c);
double op_cg_calc_ur(Chunk& c, double alpha);
void op_cg_calc_p(Chunk& c, double beta);

// Device-resident forms for the driver: the scalars (rro, pw, alpha, rrn,
// beta, error) stay in Chunk::ds and are updated by the reductions' second
// phase exactly as tea_solve.f90 updates them; nothing waits for the host.
// calc_w with `fold` first applies the pending p = beta p + r (or z), with
// `fold_u` (2D) also the pending u += alpha p that calc_r left.
// calc_ur / calc_r record alpha/beta of CG iteration `n` (1-based) when
// n > 0; calc_r updates r only and leaves u += alpha p pending, for the
// next calc_w with fold_u or for calc_u.
void op_cg_calc_w_dev(Chunk& c, bool fold, bool fold_u = false);
void op_cg_calc_ur_dev(Chunk& c, int n);
void op_cg_calc_r_dev(Chunk& c, int n, Slot slot = S_RRN);
void op_cg_calc_u_dev(Chunk& c);
void op_cg_calc_p_dev(Chunk& c);
// Chebyshev switch: keep the boundary cells' p (with the pending CGCalcP
// applied when `pending`) for op_cg_calc_w_faces_dev.
void op_capture_p_faces(Chunk& c, bool pending);
// CGCalcW whose out-of-mesh neighbours come from the captured faces (a CG
// iteration after Chebyshev steps reads that stale halo of p).
void op_cg_calc_w_faces_dev(Chunk& c);
void op_ppcg_calc_ur_dev(Chunk& c);  // u += alpha p, r -= alpha w; its r.r is not used
void op_2norm_rrn_dev(Chunk& c);  // rrn = ||r||^2, then beta, error, rro as in the PPCG loop
double op_error(Chunk& c);        // waits and returns ds->error
double stream_copy_probe(long n, int reps);   // GB/s of a plain copy of n doubles
// Sum of x[0..nblocks) through the grid reduction with workgroups of
// nwaves (1 or 4) waves: out[0], out[1] single pass twice in a row, out[2]
// with the separate k_finish launch.
void reduce_probe(long nblocks, int nwaves, const double* x, double* out);
// GB/s of a streaming probe over arrays of n doubles (kinds in tl_stream_probe)
double stream_probe(int kind, long n, int reps);
// w = A p and *pw = sum w.p on caller device arrays in the reference's flat
// layout, with no device allocation (tl_noalloc.hip)
void cg_calc_w_noalloc(int nx, int ny, int nz, const double* p, const double* kx, const double* ky,
                       const double* kz, double* w, double* pw, hipStream_t stream);

// Temporal-blocked PPCG outer iteration (tl_ppcg_tb.hip): after CGCalcW,
// CGCalcUr + PPCGInitSd + inner_steps inner steps + ||r||^2 in one sweep,
// which also closes the rrn reduction (PPCG recurrences, rrn in S_RRN).
// Output rows per wave of a streaming pass with L levels over tiles_x
// column tiles, for `slots` resident waves and blocks of `unroll` rounds.
int tb_rows(int ny, int tiles_x, int L, int slots, int unroll = 6);
// Level counts of the passes of a temporal-blocked PPCG outer iteration
// with the given inner steps (each a compiled pass size).
std::vector<int> tb_passes(int inner_steps);
bool ppcg_tb_supported(const Chunk& c, int inner_steps);
void op_ppcg_outer_tb(Chunk& c, int inner_steps, Slot slot = S_RRN);
// First row of the reflected-row table of the temporal-blocked kernels
// (2D): entry y - TB_ROWTAB0 holds the byte offsets of the cell row and the
// face row that row y reads.
constexpr int TB_ROWTAB0 = HALO - 14;
#if TL_DIM == 2
const uint2* tb_rowtab(Chunk& c);
#endif
// Temporal-blocked Chebyshev (2D, no preconditioner): `count` iterations
// from the 1-based step `step1` in passes of at most 10; the last writes r
// and, with `norm`, closes ||r||^2 into S_NORM.
bool cheby_tb_supported(const Chunk& c);
void op_cheby_tb(Chunk& c, int step1, int count, bool norm);
// Closes a reduction of three sums per block that an earlier pass left in
// the partials (out[0..3)).
void op_finish_sums3(Chunk& c, long nblocks, double* out);
// Temporal-blocked Jacobi (2D): one pass of K (1, 2, 5 or 10) iterations,
// K = jacobi_tb_pass(iterations left); returns each iteration's error sum;
// op_jacobi_tb_undo returns u to the pass's input.
bool jacobi_tb_supported(const Chunk& c);
int jacobi_tb_pass(int remaining);
std::vector<double> op_jacobi_tb(Chunk& c, int K);
void op_jacobi_tb_undo(Chunk& c);

// Flexible CG (tl_learn.hip): a . b over the interior; p = z + beta p.
double op_dot(Chunk& c, const double* a, const double* b);
void op_fcg_update_p(Chunk& c, double beta);
// One flexible-CG step after the preconditioner, its scalars on the device:
// beta = -(z.w)/pw_prev (not on the first), p = z + beta p, w = A p, alpha =
// (p.r)/(p.w), u += alpha p, r -= alpha w; result(S_RRN, 0) is ||r||^2 and
// result(S_RRN, 1) is 1 when alpha could not be formed (p.w <= 0 or a
// non-finite p.r).
// beta_blocks > 0: z . w is already in the reduction partials (from the
// preconditioner pass), with that many blocks.  In 2D the step leaves
// u += alpha p pending (u_pending) for the next step's strip; apply it with
// op_cg_calc_u_dev before u is read.  With beta_rrn the pass also applied
// the previous iteration's deferred r update and left its ||r||^2 (read from
// S_ERR); with defer_r this iteration's r update is deferred the same way and
// S_RRN gets nothing.
void op_fcg_iteration_dev(Chunk& c, bool first, long beta_blocks, bool& u_pending, bool beta_rrn = false,
                          bool defer_r = false, bool alpha_zr = true);

// Observers of the driver's FP64 PPCG solve (no counterpart in the
// reference): after every outer iteration (u, r current) and at the end of
// the solve before the energy is finalised.  Used to record training pairs.
struct SolveHooks {
    std::function<void(Chunk&, int)> on_outer;
    std::function<void(Chunk&, int)> on_cg;  // after CG iteration n (u may lag r by the pending alpha p)
    std::function<void(Chunk&)> on_end;
};
SolveHooks& solve_hooks();

// The learned preconditioner the driver's flexible CG applies (tl_learn.hip).
struct Learner;
Learner* active_learner();
// Flexible CG on the chunk set up by the driver (u, u0, r = u0 - A u, rro in
// the device scalars): true when it reached eps on the true residual, false
// when it stalled (the caller then solves with FP64 PPCG).
// z = s N(r / s) (network on, every net_every-th iteration) or 0, then
// `smooth` weighted-Jacobi sweeps z += w_s D^-1 (r - A z): the Chebyshev
// step weights of [2 / cheby_ratio, 2] (fcg_sweep_weights), or w_s = omega
// with cheby_ratio = 0; a negative cheby_ratio is set per solve by
// fcg_auto_ratio.
struct FcgOptions {
    bool network = true;
    int net_every = 1;
    int net_iters = 0;  // > 0: the network only in the first net_iters iterations, sweeps alone after
    bool net_restart = false;  // start a new recurrence (p = z) after the last network iteration
    int smooth = 32;  // eight temporal-blocked passes of 4
    double omega = 0.8;
    double cheby_ratio = -1.0;
    // the network as the coarse-grid solve of a two-grid preconditioner
    // (tl_coarse.hip), in every iteration: sweeps, restriction, network,
    // prolongation, sweeps
    bool coarse = false;
    bool coarse_pre = true;  // smooth before the coarse correction too (else only after it)
};

// Coarse grid of the two-grid preconditioner: blocks of B x B cells, ncx x
// ncy of them (B = 0: the mesh does not divide into nc blocks a side with
// ncy a multiple of 8).
struct CoarseGrid {
    int B, ncx, ncy;
};
CoarseGrid coarse_grid(const Geom& g, int nc);
// rc = P^T (r - A z) / B^2 (FP64, ncx * ncy; without with_z, of r alone) and
// the network's input plane rc / s, s = rms(rc) left in *s_dev; r, z: the
// chunk's F_R, F_Z.
void op_coarse_restrict(Chunk& c, const CoarseGrid& cg, bool with_z, double* rc, float* plane, double* s_dev);
// F_Z += scale * s * P ez (ez the network's output plane); F_Z = that without `add`.
void op_coarse_prolong(Chunk& c, const CoarseGrid& cg, const float* ez, const double* s_dev, double scale, bool add);
// The coarse operator's face coefficients (west kxc, south kyc; 0 on the
// mesh boundary) and the block mean of the density feature.
void coarse_fields(const Geom& g, const CoarseGrid& cg, const double* kx, const double* ky, const double* density,
                   float* kxc, float* kyc, float* fc, hipStream_t s);
// out = block mean of a - b.
void coarse_mean_diff(const Geom& g, const CoarseGrid& cg, const double* a, const double* b, float* out,
                      hipStream_t s);
bool fcg_learned_solve(Chunk& c, Learner* L, const FcgOptions& opt, double eps, int max_iters, int& iters,
                       double& error);
// The sweep weights of the options, w[0 .. opt.smooth - 1].  Chebyshev:
// w_k = 1 / theta_k, theta_k the roots of the degree-`smooth` Chebyshev
// polynomial on [2 / cheby_ratio, 2]
// ((hi + lo) / 2 + (hi - lo) / 2 cos((2 k + 1) pi / (2 smooth))), in the
// Lebedev-Finogenov order for a power-of-two count, else in decreasing
// order.  The residual polynomial prod_k (1 - x / theta_k) then stays below
// 1 in magnitude on that interval, and D^-1 A's spectrum lies in (0, 2).
void fcg_sweep_weights(const FcgOptions& opt, double* w);
// The Chebyshev interval ratio for the chunk's operator and m sweeps:
// 30 (m / 16)^2 sqrt(mean diag A), within [30, 30000].  D^-1 A's bulk
// spectrum widens with the face coefficients (dt / dx^2 times the
// conductivity), and a higher degree covers a wider interval; this fits the
// best ratios measured from 192^2 to 8192^2 (for 16 sweeps about 100 at
// 192^2, 300 at 512^2, 3000 at 4096^2 and 8192^2, a flat optimum from 1000
// to 10000 at the largest; for 32 sweeps 3x that a little better than 1x).
double fcg_auto_ratio(Chunk& c, int sweeps);
// Sweep q's weight over the first one, the factor the sweep kernels apply to
// w_1 / D (exactly 1 for q = 0 and for equal weights).
inline double sweep_ratio(const double* w, int q) { return w[q] / w[0]; }
// The flexible CG's preconditioner sweeps (weighted Jacobi on A z = r with
// weights[0 .. sweeps - 1], into F_Z): one temporal-blocked pass where
// supported and allowed.  With dot_w the pass also leaves the first phase
// of z . w in the reduction partials and returns their count (else 0).
long fcg_smooth(Chunk& c, int sweeps, const double* weights, bool from_zero, bool allow_fused = true,
                bool dot_w = false);
bool fcg_smooth_tb_supported(const Chunk& c, int sweeps);  // 2D, a multiple of 4 sweeps
// One temporal-blocked pass of 4 sweeps (tl_fcg_smooth_tb.hip), with the
// weights weights[first_sweep ..] of a set that starts at weights[0].
long op_fcg_smooth_tb(Chunk& c, const double* z_in, double* z_out, const double* weights, int sweeps,
                      const double* w = nullptr, double* r_out = nullptr, bool rsums = false, int first_sweep = 0);
// `sweeps` (a multiple of 4) sweeps as chained passes, into F_Z.
long op_fcg_sweep_passes(Chunk& c, const double* weights, int sweeps, bool from_zero, bool dot_w, bool r_update);

// The C++ driver (tl_driver.hip): a deck run to its end, or to max_steps
// timesteps when that is positive; returns the JSON report.
std::string driver_run_deck(const std::string& text, int max_steps);

// Learned preconditioner (tl_learn.hip): training pairs recorded from the
// FP64 PPCG solves of decks, training of the network on them, and the
// network the driver's flexible CG uses.
Learner* learn_create(unsigned long long seed);
void learn_destroy(Learner* L);
// Runs the first timestep of `deck` with FP64 PPCG, records (residual,
// correction) at up to `snapshots` outer iterations and cuts `crops` random
// crop x crop pairs from each; returns the number of pairs added.
long learn_add_deck(Learner& L, const std::string& deck, int snapshots, int crops, int crop,
                    unsigned long long seed);
long learn_pairs(const Learner& L);
// Runs the first timestep of `deck` with FP64 PPCG and adds one pair per
// snapshot for the network as the coarse-grid solve of the two-grid
// preconditioner: the whole nc x nc coarse grid (blocks of n / nc cells a
// side) with the snapshot's restricted error e_c, the consistent input
// (I + K_c) e_c, the coarse operator and density feature.  Returns the
// number of pairs added.
long learn_add_deck_coarse(Learner& L, const std::string& deck, int snapshots, int nc);
void learn_set_consistent(Learner& L, bool on);
void learn_set_phases(Learner& L, int mask);
void learn_train(Learner& L, int steps, int batch, double lr, double* losses);
void learn_set_active(Learner* L);
struct UNet;
UNet& learn_net(Learner& L);

// Mixed-precision PPCG (tl_ppcg_mixed.hip): FP32 inner steps inside the
// FP64 outer iteration.  prepare makes the FP32 coefficient copies after
// op_ppcg_init; the outer op mirrors op_ppcg_outer_tb.
bool ppcg_mixed_supported(const Chunk& c, int inner_steps);
void op_mixed_prepare(Chunk& c, const double* alphas, const double* betas, int inner_steps, double theta);
void op_ppcg_outer_mixed(Chunk& c, int inner_steps);
// rrn = ||r||^2 of a replaced residual; beta from the rro before this iteration
void op_2norm_replace_dev(Chunk& c);
void op_read_history(Chunk& c, int n, std::vector<double>& a, std::vector<double>& b);  // 1..n

// Chebyshev, PPCG and Jacobi come in two forms.  The plain form keeps the
// reference's memory effects exactly (scratch w written, updates in place).
// The fused form computes the same values in one sweep, skips the write of
// the never-read scratch w, and ping-pongs u (Chebyshev, Jacobi) or sd
// (PPCG) between two buffers, so a cell's stencil inputs are never
// overwritten while its neighbours still need them.
void op_cheby_init(Chunk& c, const double* alphas, const double* betas, int n, double theta,
                   bool preconditioner, bool fused);
void op_cheby_upload(Chunk& c);  // stored coefficients to the device (temporal-blocked passes)
void op_cheby_iterate(Chunk& c, int step1, bool fused);  // 1-based step

void op_ppcg_init(Chunk& c, const double* alphas, const double* betas, int n, double theta);
void op_ppcg_init_sd(Chunk& c);
void op_ppcg_inner(Chunk& c, int step1, bool fused);  // 1-based step
// 3D: inner steps step1 and step1 + 1 in one sweep (r and sd change
// buffers); throws in 2D.
void op_ppcg_inner2(Chunk& c, int step1);

void op_jacobi_init(Chunk& c, int coefficient, double dt);
double op_jacobi_solve(Chunk& c, bool fused);

void op_copy_u(Chunk& c);
void op_calc_residual(Chunk& c);
double op_calc_2norm(Chunk& c, bool norm_r);
// Mean of A's diagonal over the mesh cells (1 + the cell's interior face
// coefficients), deterministic.
double op_mean_diag(Chunk& c);
void op_finalise(Chunk& c);

}  // namespace tl
