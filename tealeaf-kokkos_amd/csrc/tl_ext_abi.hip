// The ext_* C-ABI of TeaLeaf's kernel layer, re-exported on top of the
// gfx950 chunk operations.  Names, argument order and meaning follow
// tealeaf_2d/ext_core.cpp and ext_chunk.cpp (tealeaf_3d/ for TL_DIM == 3):
// every argument is a pointer (Fortran implicit interfaces), a chunk is a
// 1-based index into a process-global list, double* outputs are overwritten
// with the chunk's reduction, and a LOGICAL passed as bool* is read through
// its low byte.  Errors print a message and exit(1) as the reference's
// abort() does (ext_core.cpp:571-583), unless a caller has switched that off
// with tl_set_abort_on_error(0), in which case the call returns and the
// message is kept for tl_last_error().
#include "tl_chunk.hpp"
#include "tl_unet.hpp"

#include <sys/time.h>

#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

using namespace tl;

namespace {

std::vector<std::unique_ptr<Chunk>> g_chunks;
std::string g_error;
bool g_abort_on_error = true;

void fail(const std::string& msg)
{
    g_error = msg;
    if (g_abort_on_error) {
        std::fprintf(stderr, "\nError in the TeaLeaf HIP library: %s\n", msg.c_str());
        std::exit(1);
    }
}

Chunk* chunk_at(const int* id)
{
    if (!id || *id < 1 || *id > (int)g_chunks.size() || !g_chunks[*id - 1])
        throw Error("invalid chunk " + std::to_string(id ? *id : 0));
    return g_chunks[*id - 1].get();
}

// A chunk for a stencil entry point.  The stencil kernels read a face's
// outside neighbours as the mirrored interior (the reflective halo), which
// is right only where every face is external: a chunk that has exchanged
// halo messages with neighbour chunks is refused rather than solved wrongly.
Chunk* stencil_chunk(const int* id)
{
    Chunk* c = chunk_at(id);
    if (c->internal_faces)
        throw Error("chunk " + std::to_string(*id) +
                    " has neighbour chunks: this library's stencil kernels support external faces only");
    return c;
}

template <class F>
void guarded(F&& f)
{
    try {
        f();
    } catch (const std::exception& e) {
        fail(e.what());
    }
}

}  // namespace

extern "C" {

// ---- test and driver hooks (no counterpart in the reference) --------------

const char* tl_last_error(void) { return g_error.c_str(); }
void tl_clear_error(void) { g_error.clear(); }
void tl_set_abort_on_error(int on) { g_abort_on_error = on != 0; }
int tl_dim(void) { return TL_DIM; }
int tl_num_fields(void) { return F_COUNT; }

int tl_chunk_dims(const int* chunk, int* dims)
{
    int rc = -1;
    guarded([&] {
        Chunk* c = chunk_at(chunk);
        dims[0] = c->g.X; dims[1] = c->g.Y; dims[2] = c->g.Z;
        dims[3] = c->g.nx; dims[4] = c->g.ny; dims[5] = TL_DIM == 3 ? c->g.nz : 1;
        rc = 0;
    });
    return rc;
}

int tl_get_field(const int* chunk, int field, double* out)
{
    int rc = -1;
    guarded([&] { chunk_at(chunk)->get_field(field, out); rc = 0; });
    return rc;
}

int tl_set_field(const int* chunk, int field, const double* in)
{
    int rc = -1;
    guarded([&] { chunk_at(chunk)->set_field(field, in); rc = 0; });
    return rc;
}

// Fused forms used by the C++ driver, exposed for parity tests on a chunk
// created with ext_init_.  One PPCG outer iteration after CGCalcW with the
// given alpha: CGCalcUr + PPCGInitSd + nsteps inner steps (coefficients
// alphas/betas[0..nsteps), theta) + ||r||^2, as one temporal-blocked sweep.
// Returns 0, or -1 (tl_last_error) when the fused form does not apply.
int tl_ppcg_outer_fused(const int* chunk, double alpha, double theta, const double* alphas, const double* betas,
                        int nsteps, double* rrn)
{
    int rc = -1;
    guarded([&] {
        Chunk* c = stencil_chunk(chunk);
        DevScalars s{};
        s.rro = 1.0;
        s.alpha = alpha;
        TL_HIP(hipMemcpyAsync(c->ds, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
        op_ppcg_init(*c, alphas, betas, nsteps, theta);
        op_ppcg_outer_tb(*c, nsteps);
        *rrn = op_error(*c);
        rc = 0;
    });
    return rc;
}

// PPCG inner steps step1 and step1 + 1 in one 3D sweep, as the driver runs
// them, on a chunk after ext_ppcg_init_: r, u and sd as after the two
// ext_ppcg_inner_ calls.
int tl_ppcg_inner_pair(const int* chunk, int step1)
{
    int rc = -1;
    guarded([&] {
        op_ppcg_inner2(*stencil_chunk(chunk), step1);
        rc = 0;
    });
    return rc;
}

// Chebyshev iterations step1 .. step1 + count - 1 on a chunk after
// ext_cheby_solver_init_, as the driver runs them: temporal-blocked passes
// of up to 10, the last writing r and the first phase of ||r||^2.  u, p and
// r end as after the same ext_cheby_solver_iterate_ calls.
int tl_cheby_iterate_fused(const int* chunk, int step1, int count, double* rrn)
{
    int rc = -1;
    guarded([&] {
        Chunk* c = stencil_chunk(chunk);
        op_cheby_upload(*c);
        op_cheby_tb(*c, step1, count, true);
        *rrn = c->result(S_NORM);
        rc = 0;
    });
    return rc;
}

// count Jacobi iterations (JacobiCopyU + JacobiSolve each) as the driver
// runs them: temporal-blocked passes of up to 10 with one error sum per
// iteration, written to errors[0..count).  u and r end as after the same
// ext_jacobi_kernel_solve_ calls.
int tl_jacobi_iterate_fused(const int* chunk, int count, double* errors)
{
    int rc = -1;
    guarded([&] {
        Chunk* c = stencil_chunk(chunk);
        if (count < 1) throw Error("Jacobi needs at least one iteration");
        int done = 0;
        while (done < count) {
            const std::vector<double> e = op_jacobi_tb(*c, jacobi_tb_pass(count - done));
            for (double v : e) errors[done++] = v;
        }
        rc = 0;
    });
    return rc;
}

// The flexible CG's preconditioner sweeps on a chunk built through the ABI:
// z (the u0 field's neighbour in the chunk, F_Z) = `sweeps` weighted Jacobi
// sweeps on A z = r from z, or from zero with from_zero; fused = 1 runs them
// as one temporal-blocked pass, 0 as one launch per sweep.
int tl_fcg_smooth_weights(const int* chunk, int sweeps, const double* weights, int from_zero, int fused)
{
    int rc = -1;
    guarded([&] {
        Chunk* c = stencil_chunk(chunk);
        if (sweeps < 1 || sweeps > 32 || !weights) throw Error("tl_fcg_smooth_weights: 1 to 32 sweeps and their weights");
        for (int k = 0; k < sweeps; ++k)
            if (!(weights[k] > 0.0 && std::isfinite(weights[k]))) throw Error("sweep weights must be positive");
        if (fused && !fcg_smooth_tb_supported(*c, sweeps)) throw Error("no temporal-blocked pass for these sweeps");
        fcg_smooth(*c, sweeps, weights, from_zero != 0, fused != 0);
        c->sync();
        rc = 0;
    });
    return rc;
}

// The sweep weights the flexible CG uses for `sweeps` sweeps: omega each, or
// with cheby_ratio > 1 the Chebyshev step weights of [2 / cheby_ratio, 2]
// (tl_fcg_cheby, tl_fcg_cheby_ratio).  No GPU work.
int tl_fcg_sweep_weights(int sweeps, double omega, double cheby_ratio, double* weights)
{
    int rc = -1;
    guarded([&] {
        if (sweeps < 1 || sweeps > 32 || !weights) throw Error("tl_fcg_sweep_weights: 1 to 32 sweeps");
        if (cheby_ratio != 0.0 && !(cheby_ratio > 1.0))
            throw Error("the Chebyshev ratio must exceed 1 (the ratio from the operator needs a chunk)");
        FcgOptions o;
        o.smooth = sweeps;
        o.omega = omega;
        o.cheby_ratio = cheby_ratio;
        fcg_sweep_weights(o, weights);
        rc = 0;
    });
    return rc;
}

// The same with one weight for every sweep.
int tl_fcg_smooth(const int* chunk, int sweeps, double omega, int from_zero, int fused)
{
    double w[32];
    for (int k = 0; k < 32; ++k) w[k] = omega;
    return tl_fcg_smooth_weights(chunk, sweeps, w, from_zero, fused);
}

// The two-grid preconditioner's coarse pieces on a chunk (test hook):
// nc coarse cells a side; kxc, kyc, fc (ncx * ncy, row-major) the coarse
// operator's west and south faces and the density feature; rc the
// restricted residual P^T (r - A z) / B^2 of the chunk's r and z (of r
// alone without with_z).  Any output may be null.  0, or -1 on error.
int tl_coarse_probe(const int* chunk, int nc, int with_z, float* kxc, float* kyc, float* fc, double* rc)
{
    int rc_ = -1;
    guarded([&] {
        Chunk* c = stencil_chunk(chunk);
        const CoarseGrid cg = coarse_grid(c->g, nc);
        if (cg.B == 0) throw Error("the mesh does not divide into " + std::to_string(nc) + " coarse cells a side");
        const long n = (long)cg.ncx * cg.ncy;
        float* d32 = nullptr;
        double* d64 = nullptr;
        float* plane = nullptr;
        TL_HIP(hipMalloc(&d32, sizeof(float) * 3 * n));
        TL_HIP(hipMalloc(&d64, sizeof(double) * (n + 1)));
        TL_HIP(hipMalloc(&plane, sizeof(float) * n));
        try {
            coarse_fields(c->g, cg, c->f[F_KX], c->f[F_KY], c->f[F_DENSITY], d32, d32 + n, d32 + 2 * n, c->stream);
            op_coarse_restrict(*c, cg, with_z != 0, d64, plane, d64 + n);
            c->sync();
            if (kxc) TL_HIP(hipMemcpy(kxc, d32, sizeof(float) * n, hipMemcpyDeviceToHost));
            if (kyc) TL_HIP(hipMemcpy(kyc, d32 + n, sizeof(float) * n, hipMemcpyDeviceToHost));
            if (fc) TL_HIP(hipMemcpy(fc, d32 + 2 * n, sizeof(float) * n, hipMemcpyDeviceToHost));
            if (rc) TL_HIP(hipMemcpy(rc, d64, sizeof(double) * n, hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(d32);
            (void)hipFree(d64);
            (void)hipFree(plane);
            throw;
        }
        (void)hipFree(d32);
        (void)hipFree(d64);
        (void)hipFree(plane);
        rc_ = 0;
    });
    return rc_;
}

// The sweeps-only flexible CG's preconditioner pass with the previous
// step's r update folded in (2D, 4 sweeps): r -= alpha w, then z = the four
// sweeps on that r from zero; sums[0..3) = z.w (w as given), ||r||^2, z.r.
int tl_fcg_smooth_update(const int* chunk, double omega, double alpha, double* sums)
{
    int rc = -1;
    guarded([&] {
        Chunk* c = stencil_chunk(chunk);
        if (!fcg_smooth_tb_supported(*c, 4)) throw Error("no temporal-blocked pass for this chunk");
        DevScalars s{};
        s.alpha = alpha;
        TL_HIP(hipMemcpyAsync(c->ds, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
        const double w[4] = {omega, omega, omega, omega};
        const long nb = op_fcg_smooth_tb(*c, nullptr, c->f[F_SD], w, 4, c->f[F_W], c->r_alt);
        std::swap(c->f[F_Z], c->f[F_SD]);
        std::swap(c->f[F_R], c->r_alt);
        op_finish_sums3(*c, nb, sums);
        rc = 0;
    });
    return rc;
}

// The mixed-precision form of the same outer iteration: FP32 inner steps,
// u and r updated in FP64 (2D only; nsteps with a compiled kernel).
int tl_ppcg_outer_mixed(const int* chunk, double alpha, double theta, const double* alphas, const double* betas,
                        int nsteps, double* rrn)
{
    int rc = -1;
    guarded([&] {
        Chunk* c = stencil_chunk(chunk);
        if (!ppcg_mixed_supported(*c, nsteps)) throw Error("mixed-precision PPCG does not apply to this chunk");
        DevScalars s{};
        s.rro = 1.0;
        s.alpha = alpha;
        TL_HIP(hipMemcpyAsync(c->ds, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
        op_ppcg_init(*c, alphas, betas, nsteps, theta);
        op_mixed_prepare(*c, alphas, betas, nsteps, theta);
        op_ppcg_outer_mixed(*c, nsteps);
        *rrn = op_error(*c);
        rc = 0;
    });
    return rc;
}

// ---- learned preconditioner (tl_learn.hip, tl_unet.hip) -------------------

void* tl_learn_create(unsigned long long seed)
{
    void* h = nullptr;
    guarded([&] { h = learn_create(seed); });
    return h;
}

void tl_learn_destroy(void* h) { guarded([&] { learn_destroy(static_cast<Learner*>(h)); }); }

long tl_learn_add_deck(void* h, const char* deck, int snapshots, int crops, int crop, unsigned long long seed)
{
    long n = -1;
    guarded([&] {
        if (!h || !deck) throw Error("tl_learn_add_deck: null argument");
        n = learn_add_deck(*static_cast<Learner*>(h), deck, snapshots, crops, crop, seed);
    });
    return n;
}

// One whole-coarse-grid pair per snapshot of the deck's FP64 PPCG solve (the
// network as the two-grid preconditioner's coarse solve); -1 on error.
long tl_learn_add_deck_coarse(void* h, const char* deck, int snapshots, int nc)
{
    long n = -1;
    guarded([&] {
        if (!h || !deck) throw Error("tl_learn_add_deck_coarse: null argument");
        n = learn_add_deck_coarse(*static_cast<Learner*>(h), deck, snapshots, nc);
    });
    return n;
}

long tl_learn_pairs(void* h) { return h ? learn_pairs(*static_cast<Learner*>(h)) : -1; }

// Solve phases later pairs are sampled from: 1 the CG presteps, 2 the PPCG
// outer iterations, 3 both (the default).  0, or -1 on error.
int tl_learn_sample_phases(void* h, int mask)
{
    int rc = -1;
    guarded([&] {
        if (!h) throw Error("tl_learn_sample_phases: null learner");
        learn_set_phases(*static_cast<Learner*>(h), mask);
        rc = 0;
    });
    return rc;
}

// The A-norm training loss after `sweeps` weighted Jacobi sweeps of weight
// omega (the flexible CG's smoothing after the network); 0 sweeps: the plain
// A-norm.  0, or -1 on error.
int tl_learn_loss_smoothing(void* h, int sweeps, double omega)
{
    int rc = -1;
    guarded([&] {
        if (!h) throw Error("tl_learn_loss_smoothing: null learner");
        unet_set_loss_smoothing(learn_net(*static_cast<Learner*>(h)), sweeps, (float)omega);
        rc = 0;
    });
    return rc;
}

// on != 0 (the default): a pair's input residual is A_c e, the crop
// operator applied to the crop's correction; 0: the solve's own residual
// r_k over the crop.  Applies to pairs added afterwards.
int tl_learn_consistent_pairs(void* h, int on)
{
    if (!h) return -1;
    learn_set_consistent(*static_cast<Learner*>(h), on != 0);
    return 0;
}

int tl_learn_train(void* h, int steps, int batch, double lr, double* losses)
{
    int rc = -1;
    guarded([&] {
        if (!h) throw Error("tl_learn_train: null learner");
        learn_train(*static_cast<Learner*>(h), steps, batch, lr, losses);
        rc = 0;
    });
    return rc;
}

int tl_learn_use(void* h)
{
    learn_set_active(static_cast<Learner*>(h));
    return 0;
}

long tl_learn_param_count(void* h) { return h ? unet_param_count(learn_net(*static_cast<Learner*>(h))) : -1; }

int tl_learn_get_params(void* h, float* out)
{
    int rc = -1;
    guarded([&] {
        unet_get_params(learn_net(*static_cast<Learner*>(h)), out);
        rc = 0;
    });
    return rc;
}

int tl_learn_set_params(void* h, const float* in)
{
    int rc = -1;
    guarded([&] {
        unet_set_params(learn_net(*static_cast<Learner*>(h)), in);
        rc = 0;
    });
    return rc;
}

namespace {

struct DevPlanes {
    float* d[4] = {nullptr, nullptr, nullptr, nullptr};
    ~DevPlanes()
    {
        for (float* p : d) (void)hipFree(p);
    }
};

}  // namespace

// Loss and parameter gradients of the network on host planes r, f, target
// (n x h x w each), with kx, ky the A-norm loss, else the mean squared
// error; grads receives tl_learn_param_count floats.
double tl_learn_loss_grad(void* h, const float* r, const float* f, const float* target, const float* kx,
                          const float* ky, int n, int hh, int ww, float* grads)
{
    double loss = -1.0;
    guarded([&] {
        UNet& u = learn_net(*static_cast<Learner*>(h));
        const size_t bytes = sizeof(float) * (size_t)n * hh * ww;
        DevPlanes p, q;
        for (int k = 0; k < 3; ++k) TL_HIP(hipMalloc(&p.d[k], bytes));
        TL_HIP(hipMemcpy(p.d[0], r, bytes, hipMemcpyHostToDevice));
        TL_HIP(hipMemcpy(p.d[1], f, bytes, hipMemcpyHostToDevice));
        TL_HIP(hipMemcpy(p.d[2], target, bytes, hipMemcpyHostToDevice));
        const bool anorm = kx && ky;
        if (anorm) {
            for (int k = 0; k < 2; ++k) TL_HIP(hipMalloc(&q.d[k], bytes));
            TL_HIP(hipMemcpy(q.d[0], kx, bytes, hipMemcpyHostToDevice));
            TL_HIP(hipMemcpy(q.d[1], ky, bytes, hipMemcpyHostToDevice));
        }
        loss = unet_loss_grad(u, p.d[0], p.d[1], p.d[2], anorm ? q.d[0] : nullptr, anorm ? q.d[1] : nullptr, n, hh,
                              ww);
        unet_get_grads(u, grads);
    });
    return loss;
}

// The network's output channel 0 for host planes r, f (n x h x w) into z.
int tl_learn_apply(void* h, const float* r, const float* f, float* z, int n, int hh, int ww)
{
    int rc = -1;
    guarded([&] {
        UNet& u = learn_net(*static_cast<Learner*>(h));
        const size_t bytes = sizeof(float) * (size_t)n * hh * ww;
        DevPlanes p;
        for (int k = 0; k < 3; ++k) TL_HIP(hipMalloc(&p.d[k], bytes));
        TL_HIP(hipMemcpy(p.d[0], r, bytes, hipMemcpyHostToDevice));
        TL_HIP(hipMemcpy(p.d[1], f, bytes, hipMemcpyHostToDevice));
        unet_apply(u, p.d[0], p.d[1], p.d[2], n, hh, ww);
        TL_HIP(hipStreamSynchronize(unet_stream(u)));
        TL_HIP(hipMemcpy(z, p.d[2], bytes, hipMemcpyDeviceToHost));
        rc = 0;
    });
    return rc;
}

// Mean milliseconds of one network forward pass on device-resident n x h x w
// planes (HIP events over `reps` passes after one untimed pass), and the
// multiply-accumulates of one pass in *macs.
double tl_learn_time_forward(void* h, int n, int hh, int ww, int reps, double* macs)
{
    double ms = -1.0;
    guarded([&] {
        UNet& u = learn_net(*static_cast<Learner*>(h));
        const size_t px = (size_t)n * hh * ww;
        DevPlanes p;
        for (int k = 0; k < 3; ++k) {
            TL_HIP(hipMalloc(&p.d[k], sizeof(float) * px));
            TL_HIP(hipMemset(p.d[k], 0, sizeof(float) * px));
        }
        TL_HIP(hipDeviceSynchronize());  // the fills, before the network's own stream reads the planes
        unet_apply(u, p.d[0], p.d[1], p.d[2], n, hh, ww);
        hipEvent_t e0, e1;
        TL_HIP(hipEventCreate(&e0));
        TL_HIP(hipEventCreate(&e1));
        TL_HIP(hipEventRecord(e0, unet_stream(u)));
        for (int k = 0; k < reps; ++k) unet_apply(u, p.d[0], p.d[1], p.d[2], n, hh, ww);
        TL_HIP(hipEventRecord(e1, unet_stream(u)));
        TL_HIP(hipEventSynchronize(e1));
        float t = 0.0f;
        TL_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ms = t / reps;
        if (macs) *macs = unet_forward_macs(n, hh, ww);
    });
    return ms;
}

// The network's 3x3 convolutions: 1 stages each workgroup's input tile in LDS
// (the default), 0 reads it through L2 once per tap; the outputs and
// gradients are bitwise the same.  Applies to every network of the library.
void tl_learn_conv_staging(int on) { set_conv_staging(on != 0); }

// Bytes one network forward pass on n x h x w planes moves, each tensor read
// or written once as stored (BF16 activations, FP32 planes in and out).
double tl_learn_forward_bytes(int n, int hh, int ww) { return unet_forward_bytes(n, hh, ww); }

// Bandwidth of a plain device copy of n doubles repeated reps times (GB/s,
// read + write): the machine's attainable streaming rate, and the known-size
// kernel that calibrates HBM byte counters.  Returns a negative value on error.
double tl_copy_probe(long n, int reps)
{
    double gbs = -1.0;
    guarded([&] { gbs = stream_copy_probe(n, reps); });
    return gbs;
}

// The grid reduction on given partials (see the header); 0, or -1 on error.
int tl_reduce_probe(long nblocks, int nwaves, const double* x, double* out)
{
    int rc = -1;
    guarded([&] {
        if (!x || !out) throw Error("reduce probe: null pointer");
        reduce_probe(nblocks, nwaves, x, out);
        rc = 0;
    });
    return rc;
}

// GB/s of a streaming probe over arrays of n doubles (see the header).
double tl_stream_probe(int kind, long n, int reps)
{
    double gbs = -1.0;
    guarded([&] { gbs = stream_probe(kind, n, reps); });
    return gbs;
}

// ext_cg_calc_w_ on caller-owned device arrays with no device allocation
// (the work partials live in a static device array); 0, or -1 on error.
#if TL_DIM == 3
int tl_cg_calc_w_noalloc(int nx, int ny, int nz, const double* p, const double* kx, const double* ky,
                         const double* kz, double* w, double* pw, void* stream)
{
    int rc = -1;
    guarded([&] {
        cg_calc_w_noalloc(nx, ny, nz, p, kx, ky, kz, w, pw, static_cast<hipStream_t>(stream));
        rc = 0;
    });
    return rc;
}
#else
int tl_cg_calc_w_noalloc(int nx, int ny, const double* p, const double* kx, const double* ky, double* w,
                         double* pw, void* stream)
{
    int rc = -1;
    guarded([&] {
        cg_calc_w_noalloc(nx, ny, 1, p, kx, ky, nullptr, w, pw, static_cast<hipStream_t>(stream));
        rc = 0;
    });
    return rc;
}
#endif

// ---- ext_chunk.cpp:18-33 --------------------------------------------------

#if TL_DIM == 3
void ext_init_(int* xMax, int* yMax, int* zMax, int* rank)
{
    guarded([&] { g_chunks.emplace_back(new Chunk(*xMax, *yMax, *zMax, *rank)); });
}
#else
void ext_init_(int* xMax, int* yMax, int* rank)
{
    guarded([&] { g_chunks.emplace_back(new Chunk(*xMax, *yMax, 1, *rank)); });
}
#endif

// The reference deletes its chunks but keeps the list; here the list is
// emptied too, so an ext_init_ after ext_finalise_ starts again at chunk 1.
void ext_finalise_(void) { g_chunks.clear(); }

// ---- ext_core.cpp ----------------------------------------------------------

#if TL_DIM == 3
void ext_initialise_chunk_(const int* chunk, const double* xMin, const double* yMin, const double* zMin,
                           const double* dx, const double* dy, const double* dz)
{
    guarded([&] { op_initialise_chunk(*chunk_at(chunk), *xMin, *yMin, *zMin, *dx, *dy, *dz); });
}

void ext_generate_chunk_(const int* chunk, const int* numStates, const double* density, const double* energy,
                         const double* xMin, const double* xMax, const double* yMin, const double* yMax,
                         const double* zMin, const double* zMax, const double* radius, const int* geometry,
                         const int* rectParam, const int* circParam, const int* pointParam)
{
    guarded([&] {
        op_generate_chunk(*chunk_at(chunk), *numStates, density, energy, xMin, xMax, yMin, yMax, zMin, zMax,
                          radius, geometry, *rectParam, *circParam, *pointParam);
    });
}
#else
void ext_initialise_chunk_(const int* chunk, const double* xMin, const double* yMin, const double* dx,
                           const double* dy)
{
    guarded([&] { op_initialise_chunk(*chunk_at(chunk), *xMin, *yMin, 0.0, *dx, *dy, 0.0); });
}

void ext_generate_chunk_(const int* chunk, const int* numStates, const double* density, const double* energy,
                         const double* xMin, const double* xMax, const double* yMin, const double* yMax,
                         const double* radius, const int* geometry, const int* rectParam, const int* circParam,
                         const int* pointParam)
{
    guarded([&] {
        op_generate_chunk(*chunk_at(chunk), *numStates, density, energy, xMin, xMax, yMin, yMax, nullptr, nullptr,
                          radius, geometry, *rectParam, *circParam, *pointParam);
    });
}
#endif

void ext_field_summary_kernel_(const int* chunk, double* vol, double* mass, double* ie, double* temp)
{
    guarded([&] { op_field_summary(*chunk_at(chunk), vol, mass, ie, temp); });
}

// Halo message of one face for a neighbouring chunk (ext_core.cpp:151-244;
// 3D :168-270, whose fieldType is not read): pack copies the flagged fields'
// mesh strip next to the face into buffer + offsets[field], unpack copies
// buffer + offsets[field] into the halo strip beyond it.  With one chunk per
// process and rank, the Fortran driver's MPI exchange (tea.f90:246-407)
// moves these buffers between processes, one GPU each.
#if TL_DIM == 3
void ext_pack_message_(const int* chunk, const int* fields, const int* offsets, const int* depth, const int* face,
                       const int* /*fieldType*/, double* buffer, bool* pack)
#else
void ext_pack_message_(const int* chunk, const int* fields, const int* offsets, const int* depth, const int* face,
                       double* buffer, bool* pack)
#endif
{
    guarded([&] {
        const bool p = *reinterpret_cast<const unsigned char*>(pack) != 0;
        op_pack_message(*chunk_at(chunk), fields, offsets, *depth, *face, buffer, p);
    });
}

void ext_set_field_kernel_(const int* chunk)
{
    guarded([&] { op_set_field(*chunk_at(chunk)); });
}

void ext_solver_copy_u_(const int* chunk)
{
    guarded([&] { op_copy_u(*chunk_at(chunk)); });
}

void ext_calculate_residual_(const int* chunk)
{
    guarded([&] { op_calc_residual(*stencil_chunk(chunk)); });
}

void ext_calculate_2norm_(const int* chunk, double* norm, bool* normArray)
{
    guarded([&] {
        const bool use_r = *reinterpret_cast<const unsigned char*>(normArray) != 0;
        *norm = op_calc_2norm(*chunk_at(chunk), use_r);
    });
}

void ext_solver_finalise_(const int* chunk)
{
    guarded([&] { op_finalise(*chunk_at(chunk)); });
}

void ext_update_halo_kernel_(const int* chunk, const int* chunkNeighbours, const int* fields, const int* depth)
{
    guarded([&] { op_update_halo(*chunk_at(chunk), chunkNeighbours, fields, *depth); });
}

void ext_jacobi_kernel_init_(const int* chunk, const int* coefficient, const double* dt)
{
    guarded([&] { op_jacobi_init(*stencil_chunk(chunk), *coefficient, *dt); });
}

void ext_jacobi_kernel_solve_(const int* chunk, double* error)
{
    guarded([&] { *error = op_jacobi_solve(*stencil_chunk(chunk), false); });
}

void ext_cg_solver_init_(const int* chunk, const int* coefficient, const int* preconditioner, double* dt,
                         double* rro)
{
    guarded([&] { *rro = op_cg_init(*stencil_chunk(chunk), *coefficient, *preconditioner != 0, *dt); });
}

void ext_cg_calc_w_(const int* chunk, double* pw)
{
    guarded([&] { *pw = op_cg_calc_w(*stencil_chunk(chunk)); });
}

void ext_cg_calc_ur_(const int* chunk, const double* alpha, double* rrn)
{
    guarded([&] { *rrn = op_cg_calc_ur(*chunk_at(chunk), *alpha); });
}

void ext_cg_calc_p_(const int* chunk, const double* beta)
{
    guarded([&] { op_cg_calc_p(*chunk_at(chunk), *beta); });
}

void ext_cheby_solver_init_(const int* chunk, double* alphas, double* betas, const double* theta,
                            const int* preconditioner, const int* maxChebyIters)
{
    guarded([&] {
        op_cheby_init(*stencil_chunk(chunk), alphas, betas, *maxChebyIters, *theta, *preconditioner != 0, false);
    });
}

void ext_cheby_solver_iterate_(const int* chunk, const int* chebyCalcStep)
{
    guarded([&] { op_cheby_iterate(*stencil_chunk(chunk), *chebyCalcStep, false); });
}

void ext_ppcg_init_(const int* chunk, double* alphas, double* betas, const double* theta, const int* maxChebyIters)
{
    guarded([&] { op_ppcg_init(*chunk_at(chunk), alphas, betas, *maxChebyIters, *theta); });
}

void ext_ppcg_init_sd_(const int* chunk)
{
    guarded([&] { op_ppcg_init_sd(*chunk_at(chunk)); });
}

void ext_ppcg_inner_(const int* chunk, int* currentStep)
{
    guarded([&] { op_ppcg_inner(*stencil_chunk(chunk), *currentStep, false); });
}

// timer_c.cpp:30 (wall-clock seconds for the Fortran timer)
void timer_c_(double* elapsed)
{
    struct timeval t;
    gettimeofday(&t, nullptr);
    *elapsed = (double)t.tv_sec + 1.0e-6 * (double)t.tv_usec;
}

}  // extern "C"
