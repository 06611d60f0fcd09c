/*
 * C ABI of libtealeaf2d_hip.so: TeaLeaf's 2D kernel layer (5-point stencil)
 * on an MI355X (gfx950).
 *
 * The ext_*_ entry points keep the names, argument order and conventions of
 * the reference's extension interface (tealeaf_2d/ext_core.cpp,
 * tealeaf_2d/ext_chunk.cpp), which the Fortran driver binds through implicit
 * interfaces:
 *   - every argument is passed by pointer;
 *   - a chunk is a 1-based index into a process-global list filled by
 *     ext_init_ (Chunks[*chunk-1], ext_chunk.cpp:15);
 *   - double* outputs are overwritten with the chunk's reduction; the value
 *     is in host memory when the call returns;
 *   - Fortran LOGICALs arrive as const int* or bool* (low byte read);
 *   - errors print a message and exit(1) (ext_core.cpp:571-583) unless
 *     tl_set_abort_on_error(0) was called, in which case the call returns
 *     and tl_last_error() holds the message.
 * Device state is resident in HBM; the layout is internal (fields are read
 * and written in the reference's flat layout through tl_get_field /
 * tl_set_field).  A stencil input's depth-1 halo across an external face is
 * taken to be the reflective mirror of the interior, which is what
 * ext_update_halo_kernel_ establishes and what the reference's driver keeps
 * in place before every stencil call.
 */
#ifndef TEALEAF_EXT2D_H
#define TEALEAF_EXT2D_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ext_chunk.cpp:18-25.  Creates chunk n+1 of xMax*yMax interior cells
 * (the Fortran caller passes 13 further arrays, start.f90:77-90, which are
 * not read). */
void ext_init_(int* xMax, int* yMax, int* rank);
/* ext_chunk.cpp:28-33 */
void ext_finalise_(void);

/* ext_core.cpp:33-52 (vertices, cell centres, volume, areas) */
void ext_initialise_chunk_(const int* chunk, const double* xMin, const double* yMin, const double* dx,
                           const double* dy);
/* ext_core.cpp:57-126 (paints states 2..n over state 1) */
void ext_generate_chunk_(const int* chunk, const int* numStates, const double* stateDensity,
                         const double* stateEnergy, const double* stateXMin, const double* stateXMax,
                         const double* stateYMin, const double* stateYMax, const double* stateRadius,
                         const int* stateGeometry, const int* rectParam, const int* circParam,
                         const int* pointParam);
/* ext_core.cpp:249-256 (energy1 = energy0) */
void ext_set_field_kernel_(const int* chunk);
/* ext_core.cpp:130-147 (interior sums of V, V*rho, V*rho*e0, V*rho*u) */
void ext_field_summary_kernel_(const int* chunk, double* vol, double* mass, double* ie, double* temp);
/* ext_core.cpp:305-336 (reflective copy on faces whose neighbour is -1;
 * fields[6] flags density, energy0, energy1, u, p, sd; depth 1 or 2) */
void ext_update_halo_kernel_(const int* chunk, const int* chunkNeighbours, const int* fields, const int* depth);
/* ext_core.cpp:151-244, ext_pack_kernel.hpp:7-97.  Halo message of one face
 * (face 1-4 = left, right, bottom, top; depth 1 or 2) for a neighbouring
 * chunk: *pack != 0 copies each flagged field's mesh strip next to the face
 * into buffer + offsets[field], *pack == 0 copies buffer + offsets[field]
 * into the halo strip beyond the face.  Message lengths: ny*depth (left,
 * right), nx*depth (top, bottom), in the reference's element order (its
 * strips start at row 0 / column 0, and so do these). */
void ext_pack_message_(const int* chunk, const int* fields, const int* offsets, const int* depth, const int* face,
                       double* buffer, bool* pack);

/* ext_core.cpp:377-413 */
void ext_cg_solver_init_(const int* chunk, const int* coefficient, const int* preconditioner, double* dt,
                         double* rro);
/* ext_core.cpp:417-429 (w = A p, *pw = p.w) */
void ext_cg_calc_w_(const int* chunk, double* pw);
/* ext_core.cpp:433-447 (u += alpha p, r -= alpha w, *rrn = r.r or r.z) */
void ext_cg_calc_ur_(const int* chunk, const double* alpha, double* rrn);
/* ext_core.cpp:451-464 (p = beta p + r or z) */
void ext_cg_calc_p_(const int* chunk, const double* beta);

/* ext_core.cpp:468-497 */
void ext_cheby_solver_init_(const int* chunk, double* alphas, double* betas, const double* theta,
                            const int* preconditioner, const int* maxChebyIters);
/* ext_core.cpp:501-515 (chebyCalcStep is 1-based) */
void ext_cheby_solver_iterate_(const int* chunk, const int* chebyCalcStep);

/* ext_core.cpp:519-539 */
void ext_ppcg_init_(const int* chunk, double* alphas, double* betas, const double* theta,
                    const int* maxChebyIters);
/* ext_core.cpp:543-550 */
void ext_ppcg_init_sd_(const int* chunk);
/* ext_core.cpp:554-567 (currentStep is 1-based) */
void ext_ppcg_inner_(const int* chunk, int* currentStep);

/* ext_core.cpp:340-358 */
void ext_jacobi_kernel_init_(const int* chunk, const int* coefficient, const double* dt);
/* ext_core.cpp:362-373 (*error = sum |u_new - u_old|) */
void ext_jacobi_kernel_solve_(const int* chunk, double* error);

/* ext_core.cpp:260-267 (u0 = u) */
void ext_solver_copy_u_(const int* chunk);
/* ext_core.cpp:271-278 (r = u0 - A u) */
void ext_calculate_residual_(const int* chunk);
/* ext_core.cpp:282-291 (*norm = sum r^2 if *normArray else sum u0^2) */
void ext_calculate_2norm_(const int* chunk, double* norm, bool* normArray);
/* ext_core.cpp:295-302 (energy1 = u / rho) */
void ext_solver_finalise_(const int* chunk);

/* timer_c.cpp:30 (wall-clock seconds) */
void timer_c_(double* elapsed);

/* ---- Beside the reference interface --------------------------------- */

/* The solver driver (tea_solve.f90 and its callers restated in C++): runs a
 * tea.in deck end to end on the GPU and writes a JSON report (per step:
 * iterations, error, solver diagnostics, field summary, solve seconds) into
 * json[0..cap).  max_steps > 0 stops after that many timesteps.  Returns 0,
 * -1 on error (tl_driver_error()), -2 when cap is too small. */
int tl_run_deck(const char* deck, char* json, int cap, int max_steps);
const char* tl_driver_error(void);
/* The last tl_run_deck's tea.out lines that the reference writes with fixed
 * formats: the field summaries (field_summary.f90:44-45,81, from start.f90:154
 * and diffuse.f90:60,92), the switch to Chebyshev/PPCG and est itc
 * (tea_solve.f90:216-218,744-745), Conduction error / EXACT error / Iteration
 * count (tea_solve.f90:535-543) and the QA verdict (field_summary.f90:148-155).
 * The list-directed lines (times, wall clocks, banners) are not written. */
const char* tl_tea_out(void);
/* One value through the Fortran edit descriptor desc ("aW", "iW", "fW.D",
 * "eW.D"; W 1..200, D <= 40), as those lines use them; v_int, v_real or
 * text per the letter.
 * Returns the length, -1 for an unhandled descriptor, -2 if cap is too small.
 * No counterpart in the reference (its Fortran runtime formats). */
int tl_fortran_edit(const char* desc, long v_int, double v_real, const char* text, char* out, int cap);
/* Per-launch kernel timing with HIP events on the chunk's stream for the
 * timesteps from `from_step` on (0 = off); tl_run_deck's report then has a
 * "kernels" table: launches, summed milliseconds, algorithmic bytes/cell. */
void tl_set_kernel_timing(int from_step);
/* Restricts that timing to one kernel, by its name in the report's
 * "kernels" table (NULL or "" = every kernel): fewer events in the stream,
 * less perturbed timesteps.  No counterpart in the reference.  Returns -1
 * for an unknown name. */
int tl_set_kernel_timing_only(const char* name);
/* A progress line on stderr after every timestep of tl_run_deck (0 = off). */
void tl_set_progress(int on);
/* How chunks created after the call close their grid reductions: on != 0,
 * inside the sweep that computes the sum (single pass: the last workgroup
 * to finish adds the partials and applies the solver recurrence); 0 (the
 * default), with a separate one-workgroup launch after it, which measures
 * faster on MI355X.  Both add the partials in the same order and give the
 * same bits.  No counterpart in the reference (its Kokkos CUDA backend
 * closes reductions in the last block,
 * kokkos_src/core/src/Cuda/Kokkos_Cuda_ReduceScan.hpp:331-403). */
void tl_set_single_pass(int on);

/* The fused PPCG outer iteration the driver uses (2D, no preconditioner),
 * on a chunk made by ext_init_: with the given alpha, CGCalcUr + PPCGInitSd
 * + nsteps inner steps (alphas/betas[0..nsteps), theta) + ||r||^2 in one
 * temporal-blocked sweep.  Returns 0, or -1 with tl_last_error() set. */
int tl_ppcg_outer_fused(const int* chunk, double alpha, double theta, const double* alphas, const double* betas,
                        int nsteps, double* rrn);

/* PPCG inner steps step1 and step1 + 1 (1-based) in one sweep, as the 3D
 * driver runs them (no counterpart in the reference; the two
 * ext_ppcg_inner_ calls, ext_core.cpp:554-567, fused), on a chunk after
 * ext_ppcg_init_: r, u and sd as after those calls.  3D only; otherwise
 * returns -1 with tl_last_error() set. */
int tl_ppcg_inner_pair(const int* chunk, int step1);

/* Chebyshev iterations step1 .. step1 + count - 1 (1-based) as the driver
 * runs them (no counterpart in the reference; the ext_cheby_solver_iterate_
 * calls of those steps, ext_core.cpp:501-515, in temporal-blocked passes of
 * up to 10), on a chunk after ext_cheby_solver_init_: u, p and r as after
 * those calls, ||r||^2 in *rrn.  2D, no preconditioner; otherwise returns -1
 * with tl_last_error() set. */
int tl_cheby_iterate_fused(const int* chunk, int step1, int count, double* rrn);

/* count Jacobi iterations as the driver runs them (no counterpart in the
 * reference; count ext_jacobi_kernel_solve_ calls, ext_core.cpp:362-373, in
 * temporal-blocked passes of up to 10): u and r as after those calls, the
 * error of each iteration in errors[0..count).  2D; otherwise returns -1
 * with tl_last_error() set. */
int tl_jacobi_iterate_fused(const int* chunk, int count, double* errors);

/* The flexible CG's preconditioner (no counterpart in the reference): z =
 * `sweeps` weighted Jacobi sweeps z += omega (r - A z) / D on the chunk's z
 * and r fields (from z = 0 with from_zero), as one temporal-blocked pass
 * (fused = 1; 2D, 4 sweeps) or one launch per sweep (fused = 0).  Returns 0,
 * or -1 with tl_last_error() set. */
int tl_fcg_smooth(const int* chunk, int sweeps, double omega, int from_zero, int fused);
/* The same with a weight per sweep, weights[0 .. sweeps) (1 to 32 sweeps,
 * each weight positive): sweep s applies z += (r - A z) ((w_0 / D) f_s),
 * f_s = w_s / w_0, the form the Chebyshev-weighted sweeps
 * (tl_fcg_cheby) take.  No counterpart in the reference. */
int tl_fcg_smooth_weights(const int* chunk, int sweeps, const double* weights, int from_zero, int fused);
/* The weights the flexible CG's sweeps use, weights[0 .. sweeps): omega
 * each, or with cheby_ratio > 1 (tl_fcg_cheby_ratio; tl_fcg_cheby is 30)
 * 1 / theta_k, theta_k the roots of the degree-`sweeps` Chebyshev polynomial
 * on [2 / cheby_ratio, 2] in decreasing order.  No GPU work.  No
 * counterpart in the reference.  Returns 0, or -1 (tl_last_error()). */
int tl_fcg_sweep_weights(int sweeps, double omega, double cheby_ratio, double* weights);
/* The sweeps-only flexible CG's preconditioner pass with the previous
 * step's r update folded in (2D, 4 sweeps): r -= alpha w, then z = the four
 * weighted Jacobi sweeps on that r from zero, in one pass; sums[0..3) =
 * z.w (w as given), ||r||^2, z.r.  No counterpart in the reference (a test
 * hook of the flexible CG form).  Returns 0, or -1 (tl_last_error()). */
int tl_fcg_smooth_update(const int* chunk, double omega, double alpha, double* sums);

/* Mixed-precision form of tl_ppcg_outer_fused (no counterpart in the
 * reference): the nsteps inner steps in FP32 on the FP32 rounding of the
 * residual, u and r updated in FP64 (u += alpha p + sum sd, r = r' + the FP32
 * change of the residual), ||r||^2 in FP64.  2D, no preconditioner, nsteps
 * 1, 3 or 10; otherwise returns -1 with tl_last_error() set. */
int tl_ppcg_outer_mixed(const int* chunk, double alpha, double theta, const double* alphas, const double* betas,
                        int nsteps, double* rrn);

/* GB/s (read + write) of a plain device copy of n doubles repeated reps
 * times: the attainable streaming rate, and the known-size kernel that
 * calibrates HBM byte counters.  Negative on error. */
double tl_copy_probe(long n, int reps);

/* The grid reduction of every sweep, on given values: workgroup b of a
 * 1D grid of nblocks workgroups of nwaves waves (1 or 4) contributes x[b]
 * as its partial.  out[0] and out[1]: two single-pass reductions in a row
 * (the last workgroup to finish adds the partials, a device-wide counter
 * finds it and is left at 0 for the next); out[2]: the same partials added
 * by a separate one-workgroup launch.  The three are bitwise equal by
 * construction.  No counterpart in the reference (its Kokkos CUDA backend
 * closes reductions in the last block,
 * kokkos_src/core/src/Cuda/Kokkos_Cuda_ReduceScan.hpp:331-403).  0, or -1
 * with tl_last_error() set. */
int tl_reduce_probe(long nblocks, int nwaves, const double* x, double* out);

/* GB/s of a streaming probe over arrays of n doubles, repeated reps times;
 * one 16-byte element per thread unless stated.  kind 0: the plain copy of
 * tl_copy_probe (grid-stride); 1: copy; 2: copy with nontemporal loads and
 * stores; 3: read only; 4: write only; 5: 4 reads + 2 writes (the p, w
 * strip's stream mix); 6: 6 reads + 2 writes (the PPCG outer iteration's).
 * The attainable rate the kernels' HBM figures are compared with; no
 * counterpart in the reference.  Negative on error. */
double tl_stream_probe(int kind, long n, int reps);

/* No counterpart in the reference: ext_cg_calc_w_ for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  p, kx, ky, w: caller-owned device arrays
 * in the reference's flat layout ((nx + 4) x (ny + 4), x fastest, halo 2; p's
 * halo holding the reflective copies update_halo writes); w is written on the
 * interior; *pw (one device double) receives sum w.p.  Launches on `stream`
 * (a hipStream_t, NULL for the default stream) without synchronising.  Work
 * memory is a static 512 KiB device array of one partial per mesh row:
 * meshes of more than 65536 rows are refused; calls that overlap on one GPU
 * race on it.  0, or -1 with tl_last_error() set. */
int tl_cg_calc_w_noalloc(int nx, int ny, const double* p, const double* kx, const double* ky, double* w,
                         double* pw, void* stream);

/* ---- Learned preconditioner (no counterpart in the reference) ------------
 * A 4-level BF16 U-Net (3x3 convs, 16/32/64/128 channels, FP32 accumulate)
 * trained on the GPU on (residual, PPCG correction) pairs, applied by the
 * driver's flexible CG for decks that select tl_use_fcg_learned (2D, x_cells
 * and y_cells multiples of 8; FP64 PPCG solves the timestep when the flexible
 * CG stalls).  Handles are opaque; calls return -1 (or NULL) on error with
 * tl_last_error() set. */
void* tl_learn_create(unsigned long long seed);
void tl_learn_destroy(void* learner);
/* FP64 PPCG solve of the deck's first timestep; (residual, correction) pairs
 * of crop x crop cells from up to `snapshots` outer iterations, `crops` each.
 * Returns the number of pairs added. */
long tl_learn_add_deck(void* learner, const char* deck, int snapshots, int crops, int crop,
                       unsigned long long seed);
/* Test hook of the two-grid preconditioner's coarse pieces on a chunk (2D;
 * nc coarse cells a side, blocks of B = x_cells / nc cells): kxc, kyc the
 * west and south faces of the coarse operator P^T A P / B^2 = I + K_c (the
 * sums of the fine faces across each coarse face over B^2, 0 on the mesh
 * boundary), fc the block mean of the density feature, rc the restricted
 * residual P^T (r - A z) / B^2 of the chunk's r and z fields (of r alone
 * without with_z); row-major (nc x y_cells / B), any output may be NULL.  No
 * counterpart in the reference.  0, or -1 with tl_last_error() set. */
int tl_coarse_probe(const int* chunk, int nc, int with_z, float* kxc, float* kyc, float* fc, double* rc);
/* The network as the coarse-grid solve of a two-grid preconditioner (deck
 * keyword tl_fcg_coarse with tl_use_fcg_learned): FP64 PPCG solve of the
 * deck's first timestep, one pair per snapshot, each the whole nc x nc
 * coarse grid of the mesh's (n / nc)^2-cell blocks (nc a multiple of 8
 * dividing x_cells and y_cells, y_cells / (x_cells / nc) a multiple of 8
 * too): target the snapshot's restricted error e_c, input the coarse
 * residual (I + K_c) e_c of the Galerkin coarse operator scaled by 1 / B^2.
 * A learner holds crops or coarse pairs, not both.  Returns the number of
 * pairs added. */
long tl_learn_add_deck_coarse(void* learner, const char* deck, int snapshots, int nc);
long tl_learn_pairs(void* learner);
/* on != 0 (the default): each pair's input residual is A_c e, the crop's
 * operator (no flux across the crop edge) applied to its correction e, so the
 * target is exactly A_c^-1 of the input; 0: the solve's residual r_k over the
 * crop.  Applies to pairs added afterwards. */
int tl_learn_consistent_pairs(void* learner, int on);
/* Solve phases later pairs are sampled from: 1 the CG presteps of the FP64
 * PPCG solve, 2 its outer iterations, 3 both (the default); 0, or -1. */
int tl_learn_sample_phases(void* learner, int mask);
/* The A-norm training loss measured after `sweeps` weighted Jacobi sweeps of
 * weight omega, as the flexible CG applies them after the network:
 * per sample sum(Md.A Md) / max(sum(Mt.A Mt), batch mean), averaged,
 * M = (I - omega D^-1 A)^sweeps (default 4 and
 * 0.8, the flexible CG's defaults; 0 sweeps: the plain A-norm).  sweeps in
 * [0, 16], omega in (0, 1]; 0, or -1 on error. */
int tl_learn_loss_smoothing(void* learner, int sweeps, double omega);
/* Adam on random batches of the pairs; losses[steps] receives each step's loss. */
int tl_learn_train(void* learner, int steps, int batch, double lr, double* losses);
/* The network tl_run_deck's flexible CG uses (NULL: none). */
int tl_learn_use(void* learner);
/* Parameters in layer order, each layer W[cout][9][cin] then b[cout]. */
long tl_learn_param_count(void* learner);
int tl_learn_get_params(void* learner, float* out);
int tl_learn_set_params(void* learner, const float* in);
/* Loss of the output against target and its parameter gradients (host
 * planes of n x h x w; h, w multiples of 8).  With kx, ky (each pixel's west
 * and south face coefficients) the A-norm sum(d.A d) / sum(t.A t) over the n
 * samples, d = out - target, as in training; with NULL the mean squared
 * error. */
double tl_learn_loss_grad(void* learner, const float* r, const float* f, const float* target, const float* kx,
                          const float* ky, int n, int h, int w, float* grads);
/* Network output (channel 0) for host planes r (scaled residual) and f
 * (density feature). */
int tl_learn_apply(void* learner, const float* r, const float* f, float* z, int n, int h, int w);
/* Mean milliseconds of one forward pass on device-resident n x h x w planes
 * over `reps` timed passes (HIP events); *macs = its multiply-accumulates. */
double tl_learn_time_forward(void* learner, int n, int h, int w, int reps, double* macs);
/* Bytes one forward pass on n x h x w planes moves, each tensor read or
 * written once as stored (BF16 activations with padded channels, FP32 planes
 * in and out): the HBM side of the network's roofline. */
double tl_learn_forward_bytes(int n, int h, int w);
/* The network's 3x3 convolutions: 1 (default) stages each workgroup's input
 * tile in LDS, 0 reads it through L2 once per tap; outputs and gradients are
 * bitwise the same either way.  No counterpart in the reference. */
void tl_learn_conv_staging(int on);

/* Error handling of the ext_*_ calls. */
void tl_set_abort_on_error(int on);
const char* tl_last_error(void);
void tl_clear_error(void);

/* Test access to chunk state in the reference's flat layout (X = nx+4
 * fastest, Y = ny+4 rows).  Field ids: 0 density, 1 energy0, 2 energy1, 3 u,
 * 4 u0, 5 p, 6 r, 7 mi, 8 w, 9 z, 10 kx, 11 ky, 12 kz (3D only), 13 sd,
 * 14 volume.  dims receives X, Y, Z, nx, ny, nz. */
int tl_dim(void);
int tl_num_fields(void);
int tl_chunk_dims(const int* chunk, int* dims);
int tl_get_field(const int* chunk, int field, double* out);
int tl_set_field(const int* chunk, int field, const double* in);

#ifdef __cplusplus
}
#endif

#endif
