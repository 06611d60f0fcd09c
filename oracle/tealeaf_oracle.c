/*
 * CPU oracle for TeaLeaf's linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against; nothing in the product links, loads or calls it.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load the
 * shared objects built from it.
 *
 * It restates, in plain C with OpenMP, the reference's behaviour for:
 *   - the ext_* C-ABI kernels     tealeaf_2d/ext_core.cpp, ext_*.hpp
 *                                 tealeaf_3d/ext_core.cpp, ext_*.hpp
 *   - the Fortran caller          tealeaf_2d/tea_solve.f90, tea_leaf_cheby.f90,
 *                                 diffuse.f90, start.f90, field_summary.f90,
 *                                 read_input.f90, parse.f90
 *   - the summation order of Kokkos' OpenMP parallel_reduce
 *                                 kokkos_src/core/src/Kokkos_ExecPolicy.hpp:200-219
 *                                 kokkos_src/core/src/OpenMP/Kokkos_OpenMP_Parallel.hpp:161-193
 *
 * Built twice: -DTL_DIM=2 (5-point stencil) and -DTL_DIM=3 (7-point stencil),
 * with -ffp-contract=off so every per-cell expression is evaluated exactly as
 * written (no fused multiply-add), the IEEE semantics the reference's QA
 * values were produced with (tealeaf_2d/field_summary.f90:24-25).
 *
 * Exported symbols use the reference's ext_*_ names and signatures so that a
 * test can drive the oracle and the HIP library through the same calls.
 * tl_* symbols are test hooks (field access, the driver entry point).
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#ifndef TL_DIM
#error "compile with -DTL_DIM=2 or -DTL_DIM=3"
#endif

#define HALO 2

/* Face, field and coefficient codes: tealeaf_2d/ext_shared.hpp:10-29,
 * tealeaf_3d/ext_shared.hpp:10-31 */
#define CHUNK_LEFT 1
#define CHUNK_RIGHT 2
#define CHUNK_BOTTOM 3
#define CHUNK_TOP 4
#define CHUNK_BACK 5
#define CHUNK_FRONT 6
#define EXTERNAL_FACE (-1)
#define FIELD_DENSITY 1
#define FIELD_ENERGY0 2
#define FIELD_ENERGY1 3
#define FIELD_U 4
#define FIELD_P 5
#define FIELD_SD 6
#define NUM_FIELDS 6
#define CONDUCTIVITY 1
#define RECIP_CONDUCTIVITY 2

/* Geometry codes: tealeaf_2d/data.f90:62-64 */
#define G_RECT 1
#define G_CIRC 2
#define G_POINT 3

/* Field ids of the tl_get_field / tl_set_field test hooks (shared with the
 * HIP library). */
enum {
    TF_DENSITY = 0, TF_ENERGY0, TF_ENERGY1, TF_U, TF_U0, TF_P, TF_R, TF_MI,
    TF_W, TF_Z, TF_KX, TF_KY, TF_KZ, TF_SD, TF_VOLUME, TF_COUNT
};

typedef struct {
    int nx, ny, nz;   /* interior cells */
    int X, Y, Z;      /* padded extents (Z == 1 in 2D) */
    long full;
    int rank;
    double dx, dy, dz;
    double theta;
    int preconditioner;
    double *f[TF_COUNT];
    double *vertexX, *vertexY, *vertexZ, *cellX, *cellY, *cellZ;
    double *alphas, *betas;
    int ncoefs;
} Chunk;

static Chunk **g_chunks = NULL;
static int g_nchunks = 0;
/* Number of partitions of Kokkos' OpenMP reduce (the pool size T). */
static int g_parts = 1;
static int g_threads = 0;
static char g_errbuf[512];
static int g_err = 0;

static void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_errbuf, sizeof g_errbuf, fmt, ap);
    va_end(ap);
    g_err = 1;
}

/* ext_core.cpp:571-583 prints and exits; the oracle runs inside the test
 * process, so it records the message and aborts the current call instead. */
#define CHECK_CHUNK(c, id)                                               \
    Chunk *c = ((id) >= 1 && (id) <= g_nchunks) ? g_chunks[(id)-1] : NULL; \
    if (!c) { set_error("invalid chunk %d", (id)); return; }

void tl_oracle_set_reduce_parts(int parts) { g_parts = parts < 1 ? 1 : parts; }
void tl_oracle_set_threads(int threads) { g_threads = threads; if (threads > 0) omp_set_num_threads(threads); }
/* keep != 0: tl_run_deck leaves its chunk allocated after the run, so the
 * tests can read the final fields through tl_get_field; the next
 * tl_run_deck or ext_finalise_ releases it. */
static int g_keep_chunk = 0;
void tl_oracle_keep_chunk(int keep) { g_keep_chunk = keep; }
int tl_oracle_get_reduce_parts(void) { return g_parts; }
const char *tl_last_error(void) { return g_err ? g_errbuf : ""; }
void tl_clear_error(void) { g_err = 0; g_errbuf[0] = 0; }

/* ------------------------------------------------------------------------ */
/* Index domains: tealeaf_2d/ext_shared.hpp:43-57, tealeaf_3d/ext_shared.hpp */
#if TL_DIM == 3
#define INNER(c) (kk >= HALO && kk < (c)->X - HALO && jj >= HALO && jj < (c)->Y - HALO && ii >= HALO && ii < (c)->Z - HALO)
#define ONE(c) (ii > 0 && ii < (c)->Z - 1 && jj > 0 && jj < (c)->Y - 1 && kk > 0 && kk < (c)->X - 1)
#define SKEW(c) (ii >= HALO && ii < (c)->Z - 1 && jj >= HALO && jj < (c)->Y - 1 && kk >= HALO && kk < (c)->X - 1)
#else
#define INNER(c) (kk >= HALO && kk < (c)->X - HALO && jj >= HALO && jj < (c)->Y - HALO)
#define ONE(c) (jj > 0 && jj < (c)->Y - 1 && kk > 0 && kk < (c)->X - 1)
#define SKEW(c) (jj >= HALO && jj < (c)->Y - 1 && kk >= HALO && kk < (c)->X - 1)
#endif

/* Kokkos::parallel_for over [0, fullDomain): order-free, so rows are split
 * over threads. */
#define FOR_FULL(c, BODY)                                                     \
    do {                                                                      \
        const long nrows_ = (long)(c)->Y * (c)->Z;                            \
        _Pragma("omp parallel for schedule(static)")                          \
        for (long row_ = 0; row_ < nrows_; row_++) {                          \
            const int jj = (int)(row_ % (c)->Y);                              \
            const int ii = (int)(row_ / (c)->Y);                              \
            (void)ii; (void)jj;                                               \
            for (int kk = 0; kk < (c)->X; kk++) {                             \
                const long index = row_ * (c)->X + kk;                        \
                (void)index;                                                  \
                BODY                                                          \
            }                                                                 \
        }                                                                     \
    } while (0)

/* Kokkos::parallel_reduce over [0, fullDomain) on the OpenMP backend:
 * partition p of T takes the contiguous range [p*wp, (p+1)*wp) with
 * wp = ((N + T - 1)/T + 7) & ~7 (Kokkos_ExecPolicy.hpp:200-219), sums its
 * terms in index order from 0.0, and the partials are joined serially from
 * the highest rank down: ((s[T-1] + s[T-2]) + ...) + s[0]
 * (Kokkos_OpenMP_Parallel.hpp:175-181 with pool_rev, Kokkos_OpenMPexec.hpp:116-117,
 * 133-134).  NV is the number of values reduced together. */
#define REDUCE_FULL(c, NV, out, BODY)                                         \
    do {                                                                      \
        const int T_ = g_parts;                                               \
        const long N_ = (c)->full;                                            \
        const long wp_ = ((N_ + T_ - 1) / T_ + 7) & ~7L;                      \
        double *part_ = calloc((size_t)T_ * (NV), sizeof(double));            \
        _Pragma("omp parallel for schedule(dynamic,1)")                       \
        for (int t_ = 0; t_ < T_; t_++) {                                     \
            long b_ = wp_ * t_, e_ = b_ + wp_;                                \
            if (b_ > N_) b_ = N_;                                             \
            if (e_ > N_) e_ = N_;                                             \
            double acc[(NV)];                                                 \
            for (int v_ = 0; v_ < (NV); v_++) acc[v_] = 0.0;                  \
            int kk = (int)(b_ % (c)->X);                                      \
            long q_ = b_ / (c)->X;                                            \
            int jj = (int)(q_ % (c)->Y);                                      \
            int ii = (int)(q_ / (c)->Y);                                      \
            (void)ii;                                                         \
            for (long index = b_; index < e_; index++) {                      \
                BODY                                                          \
                if (++kk == (c)->X) { kk = 0; if (++jj == (c)->Y) { jj = 0; ii++; } } \
            }                                                                 \
            for (int v_ = 0; v_ < (NV); v_++) part_[(size_t)t_ * (NV) + v_] = acc[v_]; \
        }                                                                     \
        for (int v_ = 0; v_ < (NV); v_++) {                                   \
            double s_ = part_[(size_t)(T_ - 1) * (NV) + v_];                  \
            for (int t_ = T_ - 2; t_ >= 0; t_--) s_ += part_[(size_t)t_ * (NV) + v_]; \
            (out)[v_] = s_;                                                   \
        }                                                                     \
        free(part_);                                                          \
    } while (0)

/* SMVP: tealeaf_2d/ext_shared.hpp:37-41, tealeaf_3d/ext_shared.hpp:39-45.
 * kx/ky/kz already carry rx/ry/rz. */
static inline double smvp(const Chunk *c, const double *a, long i)
{
    const double *kx = c->f[TF_KX], *ky = c->f[TF_KY];
    const long X = c->X;
#if TL_DIM == 3
    const double *kz = c->f[TF_KZ];
    const long P = (long)c->X * c->Y;
    return (1.0 + (kx[i + 1] + kx[i]) + (ky[i + X] + ky[i]) + (kz[i + P] + kz[i])) * a[i]
         - (kx[i + 1] * a[i + 1] + kx[i] * a[i - 1])
         - (ky[i + X] * a[i + X] + ky[i] * a[i - X])
         - (kz[i + P] * a[i + P] + kz[i] * a[i - P]);
#else
    return (1.0 + (kx[i + 1] + kx[i]) + (ky[i + X] + ky[i])) * a[i]
         - (kx[i + 1] * a[i + 1] + kx[i] * a[i - 1])
         - (ky[i + X] * a[i + X] + ky[i] * a[i - X]);
#endif
}

/* ------------------------------------------------------------------------ */
/* Chunk lifetime: tealeaf_2d/ext_chunk.cpp:18-82, tealeaf_3d/ext_chunk.cpp */

static Chunk *chunk_new(int nx, int ny, int nz, int rank)
{
    Chunk *c = calloc(1, sizeof(Chunk));
    c->nx = nx; c->ny = ny; c->nz = nz; c->rank = rank;
    c->X = nx + 2 * HALO;
    c->Y = ny + 2 * HALO;
#if TL_DIM == 3
    c->Z = nz + 2 * HALO;
#else
    c->Z = 1;
#endif
    c->full = (long)c->X * c->Y * c->Z;
    for (int k = 0; k < TF_COUNT; k++) {
#if TL_DIM == 2
        if (k == TF_KZ) continue;
#endif
        c->f[k] = calloc((size_t)c->full, sizeof(double));
    }
    c->vertexX = calloc((size_t)c->X + 1, sizeof(double));
    c->vertexY = calloc((size_t)c->Y + 1, sizeof(double));
    c->vertexZ = calloc((size_t)c->Z + 1, sizeof(double));
    c->cellX = calloc((size_t)c->X, sizeof(double));
    c->cellY = calloc((size_t)c->Y, sizeof(double));
    c->cellZ = calloc((size_t)c->Z, sizeof(double));
    return c;
}

static void chunk_free(Chunk *c)
{
    if (!c) return;
    for (int k = 0; k < TF_COUNT; k++) free(c->f[k]);
    free(c->vertexX); free(c->vertexY); free(c->vertexZ);
    free(c->cellX); free(c->cellY); free(c->cellZ);
    free(c->alphas); free(c->betas);
    free(c);
}

#if TL_DIM == 3
void ext_init_(int *xMax, int *yMax, int *zMax, int *rank)
{
    int nz = *zMax;
#else
void ext_init_(int *xMax, int *yMax, int *rank)
{
    int nz = 1;
#endif
    g_chunks = realloc(g_chunks, sizeof(Chunk *) * (size_t)(g_nchunks + 1));
    g_chunks[g_nchunks++] = chunk_new(*xMax, *yMax, nz, *rank);
}

/* The reference deletes the chunks but keeps the (dangling) list; here the
 * list is emptied so a later ext_init_ in the same process starts at 1. */
void ext_finalise_(void)
{
    for (int k = 0; k < g_nchunks; k++) chunk_free(g_chunks[k]);
    free(g_chunks);
    g_chunks = NULL;
    g_nchunks = 0;
}

/* ------------------------------------------------------------------------ */
/* Test hooks */

int tl_num_fields(void) { return TF_COUNT; }

int tl_chunk_dims(const int *chunk, int *dims /* X, Y, Z, nx, ny, nz */)
{
    int id = *chunk;
    if (id < 1 || id > g_nchunks) return -1;
    Chunk *c = g_chunks[id - 1];
    dims[0] = c->X; dims[1] = c->Y; dims[2] = c->Z;
    dims[3] = c->nx; dims[4] = c->ny; dims[5] = c->nz;
    return 0;
}

int tl_get_field(const int *chunk, int field, double *out)
{
    int id = *chunk;
    if (id < 1 || id > g_nchunks || field < 0 || field >= TF_COUNT) return -1;
    Chunk *c = g_chunks[id - 1];
    if (!c->f[field]) return -1;
    memcpy(out, c->f[field], sizeof(double) * (size_t)c->full);
    return 0;
}

int tl_set_field(const int *chunk, int field, const double *in)
{
    int id = *chunk;
    if (id < 1 || id > g_nchunks || field < 0 || field >= TF_COUNT) return -1;
    Chunk *c = g_chunks[id - 1];
    if (!c->f[field]) return -1;
    memcpy(c->f[field], in, sizeof(double) * (size_t)c->full);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Setup kernels */

/* ext_core.cpp:33-52; ext_initialise_chunk.hpp:12-88 */
#if TL_DIM == 3
void ext_initialise_chunk_(const int *chunk, const double *xMin, const double *yMin,
                           const double *zMin, const double *dx, const double *dy,
                           const double *dz)
#else
void ext_initialise_chunk_(const int *chunk, const double *xMin, const double *yMin,
                           const double *dx, const double *dy)
#endif
{
    CHECK_CHUNK(c, *chunk);
    c->dx = *dx;
    c->dy = *dy;
    for (int i = 0; i < c->X + 1; i++) c->vertexX[i] = *xMin + *dx * (i - HALO);
    for (int i = 0; i < c->Y + 1; i++) c->vertexY[i] = *yMin + *dy * (i - HALO);
    for (int i = 0; i < c->X; i++) c->cellX[i] = 0.5 * (c->vertexX[i] + c->vertexX[i + 1]);
    for (int i = 0; i < c->Y; i++) c->cellY[i] = 0.5 * (c->vertexY[i] + c->vertexY[i + 1]);
#if TL_DIM == 3
    c->dz = *dz;
    for (int i = 0; i < c->Z + 1; i++) c->vertexZ[i] = *zMin + *dz * (i - HALO);
    for (int i = 0; i < c->Z; i++) c->cellZ[i] = 0.5 * (c->vertexZ[i] + c->vertexZ[i + 1]);
    const double vol = *dx * *dy * *dz;
#else
    const double vol = *dx * *dy;
#endif
    double *volume = c->f[TF_VOLUME];
    FOR_FULL(c, { volume[index] = vol; });
}

/* ext_core.cpp:57-126; ext_generate_chunk.hpp:12-101 */
#if TL_DIM == 3
void ext_generate_chunk_(const int *chunk, const int *numStates, const double *sDensity,
                         const double *sEnergy, const double *sXMin, const double *sXMax,
                         const double *sYMin, const double *sYMax, const double *sZMin,
                         const double *sZMax, const double *sRadius, const int *sGeometry,
                         const int *rectParam, const int *circParam, const int *pointParam)
#else
void ext_generate_chunk_(const int *chunk, const int *numStates, const double *sDensity,
                         const double *sEnergy, const double *sXMin, const double *sXMax,
                         const double *sYMin, const double *sYMax, const double *sRadius,
                         const int *sGeometry, const int *rectParam, const int *circParam,
                         const int *pointParam)
#endif
{
    CHECK_CHUNK(c, *chunk);
    const int ns = *numStates;
    double *energy0 = c->f[TF_ENERGY0], *density = c->f[TF_DENSITY], *u = c->f[TF_U];
    FOR_FULL(c, {
        energy0[index] = sEnergy[0];
        density[index] = sDensity[0];
        for (int ss = 1; ss < ns; ss++) {
            int apply = 0;
            if (sGeometry[ss] == *rectParam) {
                apply = c->vertexX[kk + 1] >= sXMin[ss] && c->vertexX[kk] < sXMax[ss] &&
                        c->vertexY[jj + 1] >= sYMin[ss] && c->vertexY[jj] < sYMax[ss];
#if TL_DIM == 3
                apply = apply && c->vertexZ[ii + 1] >= sZMin[ss] && c->vertexZ[ii] < sZMax[ss];
#endif
            } else if (sGeometry[ss] == *circParam) {
#if TL_DIM == 3
                double radius = sqrt((c->cellX[kk] - sXMin[ss]) * (c->cellX[kk] - sXMin[ss]) +
                                     (c->cellY[jj] - sYMin[ss]) * (c->cellY[jj] - sYMin[ss]) +
                                     (c->cellZ[ii] - sZMin[ss]) * (c->cellZ[ii] - sZMin[ss]));
#else
                double radius = sqrt((c->cellX[kk] - sXMin[ss]) * (c->cellX[kk] - sXMin[ss]) +
                                     (c->cellY[jj] - sYMin[ss]) * (c->cellY[jj] - sYMin[ss]));
#endif
                apply = radius <= sRadius[ss];
            } else if (sGeometry[ss] == *pointParam) {
                apply = c->vertexX[kk] == sXMin[ss] && c->vertexY[jj] == sYMin[ss];
#if TL_DIM == 3
                apply = apply && c->vertexZ[ii] == sZMin[ss];
#endif
            }
            if (apply) {
                energy0[index] = sEnergy[ss];
                density[index] = sDensity[ss];
            }
        }
        if (ONE(c)) u[index] = energy0[index] * density[index];
    });
}

/* ext_core.cpp:249-256; ext_set_field.hpp:12-29 */
void ext_set_field_kernel_(const int *chunk)
{
    CHECK_CHUNK(c, *chunk);
    double *e = c->f[TF_ENERGY1], *e0 = c->f[TF_ENERGY0];
    FOR_FULL(c, { e[index] = e0[index]; });
}

/* ext_core.cpp:130-147; ext_field_summary.hpp:12-67 */
void ext_field_summary_kernel_(const int *chunk, double *vol, double *mass, double *ie, double *temp)
{
    CHECK_CHUNK(c, *chunk);
    const double *volume = c->f[TF_VOLUME], *density = c->f[TF_DENSITY];
    const double *energy0 = c->f[TF_ENERGY0], *u = c->f[TF_U];
    double res[4];
    REDUCE_FULL(c, 4, res, {
        if (INNER(c)) {
            double cellVol = volume[index];
            double cellMass = cellVol * density[index];
            acc[0] += cellVol;
            acc[1] += cellMass;
            acc[2] += cellMass * energy0[index];
            acc[3] += cellMass * u[index];
        }
    });
    *vol = res[0]; *mass = res[1]; *ie = res[2]; *temp = res[3];
}

/* ------------------------------------------------------------------------ */
/* Reflective halo: ext_core.cpp:305-336 and ext_update_halo.hpp:11-76 (2D);
 * tealeaf_3d/ext_core.cpp:332-367 and ext_update_halo.hpp:11-136 (3D). */

static void halo_face(Chunk *c, double *buf, int face, int depth)
{
    const int X = c->X, Y = c->Y;
#if TL_DIM == 2
    switch (face) {
    case CHUNK_LEFT:
        for (int idx = 0; idx < Y * depth; idx++) {
            int flip = idx % depth, lines = idx / depth;
            int from = HALO + lines * (X - depth) + idx;
            buf[from - (1 + flip * 2)] = buf[from];
        }
        break;
    case CHUNK_RIGHT:
        for (int idx = 0; idx < Y * depth; idx++) {
            int flip = idx % depth, lines = idx / depth;
            int to = X - HALO + lines * (X - depth) + idx;
            buf[to] = buf[to - (1 + flip * 2)];
        }
        break;
    case CHUNK_TOP:
        for (int idx = 0; idx < X * depth; idx++) {
            int lines = idx / X;
            int to = X * (Y - HALO) + idx;
            buf[to] = buf[to - (1 + lines * 2) * X];
        }
        break;
    case CHUNK_BOTTOM:
        for (int idx = 0; idx < X * depth; idx++) {
            int lines = idx / X;
            int from = X * HALO + idx;
            buf[from - (1 + lines * 2) * X] = buf[from];
        }
        break;
    }
#else
    const int Z = c->Z;
    const long page = (long)X * Y;
    const int HP2 = 2 * HALO;
    switch (face) {
    case CHUNK_LEFT:
    case CHUNK_RIGHT: {
        long n = (long)(Y - HP2) * (Z - HP2) * depth;
        int height = Y - HP2;
        for (long idx = 0; idx < n; idx++) {
            long lines = idx / depth;
            int flip = (int)(idx % depth);
            long base = page * HALO + X * HALO + lines * (X - depth) + (lines / height) * X * HP2;
            if (face == CHUNK_LEFT) {
                long from = base + HALO + idx;
                buf[from - (1 + flip * 2)] = buf[from];
            } else {
                long to = base + (X - HALO) + idx;
                buf[to] = buf[to - (1 + flip * 2)];
            }
        }
        break;
    }
    case CHUNK_TOP:
    case CHUNK_BOTTOM: {
        long n = (long)(X - HP2) * (Z - HP2) * depth;
        int width = X - HP2;
        for (long idx = 0; idx < n; idx++) {
            long lines = idx / width;
            int flip = (int)(lines % depth);
            long tail = HALO + lines * HP2 + (long)X * (Y - depth) * (lines / depth);
            if (face == CHUNK_TOP) {
                long to = page * HALO + (long)X * (Y - HALO) + tail + idx;
                buf[to] = buf[to - (1 + flip * 2) * X];
            } else {
                long from = page * HALO + (long)X * HALO + tail + idx;
                buf[from - (1 + flip * 2) * X] = buf[from];
            }
        }
        break;
    }
    case CHUNK_FRONT:
    case CHUNK_BACK: {
        long n = (long)(X - HP2) * (Y - HP2) * depth;
        int width = X - HP2, height = Y - HP2;
        for (long idx = 0; idx < n; idx++) {
            long lines = idx / width;
            int flip = (int)(idx / ((long)width * height));
            long tail = (long)X * HALO + HALO + lines * HP2 + (lines / height) * X * HP2;
            if (face == CHUNK_FRONT) {
                long to = page * (Z - HALO) + tail + idx;
                buf[to] = buf[to - (1 + flip * 2) * page];
            } else {
                long from = page * HALO + tail + idx;
                buf[from - (1 + flip * 2) * page] = buf[from];
            }
        }
        break;
    }
    }
#endif
}

void ext_update_halo_kernel_(const int *chunk, const int *chunkNeighbours, const int *fields,
                             const int *depth)
{
    CHECK_CHUNK(c, *chunk);
    static const int order[NUM_FIELDS] = {FIELD_DENSITY, FIELD_P, FIELD_ENERGY0,
                                          FIELD_ENERGY1, FIELD_U, FIELD_SD};
#if TL_DIM == 3
    static const int faces[] = {CHUNK_LEFT, CHUNK_RIGHT, CHUNK_TOP, CHUNK_BOTTOM, CHUNK_FRONT, CHUNK_BACK};
    const int nfaces = 6;
#else
    static const int faces[] = {CHUNK_LEFT, CHUNK_RIGHT, CHUNK_TOP, CHUNK_BOTTOM};
    const int nfaces = 4;
#endif
    for (int k = 0; k < NUM_FIELDS; k++) {
        int fid = order[k];
        if (!fields[fid - 1]) continue;
        double *buf = NULL;
        switch (fid) {
        case FIELD_DENSITY: buf = c->f[TF_DENSITY]; break;
        case FIELD_P: buf = c->f[TF_P]; break;
        case FIELD_ENERGY0: buf = c->f[TF_ENERGY0]; break;
        case FIELD_ENERGY1: buf = c->f[TF_ENERGY1]; break;
        case FIELD_U: buf = c->f[TF_U]; break;
        case FIELD_SD: buf = c->f[TF_SD]; break;
        }
        for (int q = 0; q < nfaces; q++)
            if (chunkNeighbours[faces[q] - 1] == EXTERNAL_FACE) halo_face(c, buf, faces[q], *depth);
    }
}

/* Halo messages for a neighbouring chunk: ext_core.cpp:151-244 (2D),
 * :168-270 (3D), element b of a face's message at the flat index the
 * packers of ext_pack_kernel.hpp give (2D :10-39 top/bottom, :58-85
 * left/right; 3D :21-61 top/bottom, :83-122 left/right, :146-185
 * back/front), restated index for index. */
static long pack_index(const Chunk *c, int face, int depth, int pack, long b)
{
    const long X = c->X, Y = c->Y;
#if TL_DIM == 2
    if (face == CHUNK_TOP || face == CHUNK_BOTTOM) {
        const long width = X - 2 * HALO, lines = b / width;
        long off;
        if (face == CHUNK_TOP && pack) off = X * (Y - HALO - depth) + lines * 2 * HALO;
        else if (face == CHUNK_BOTTOM && pack) off = X * HALO + lines * 2 * HALO;
        else if (face == CHUNK_TOP) off = X * (Y - HALO) + lines * 2 * HALO;
        else off = X * (HALO - depth) + lines * 2 * HALO;
        return off + b;
    } else {
        const long lines = b / depth;
        long off;
        if (face == CHUNK_LEFT && pack) off = HALO + lines * (X - depth);
        else if (face == CHUNK_RIGHT && pack) off = X - HALO - depth + lines * (X - depth);
        else if (face == CHUNK_LEFT) off = HALO - depth + lines * (X - depth);
        else off = X - HALO + lines * (X - depth);
        return off + b;
    }
#else
    const long Z = c->Z, XY = X * Y, HP2 = 2 * HALO;
    if (face == CHUNK_TOP || face == CHUNK_BOTTOM) {
        const long lines = b / (X - HP2);
        long row;
        if (face == CHUNK_TOP && pack) row = Y - HALO - depth;
        else if (face == CHUNK_BOTTOM && pack) row = HALO;
        else if (face == CHUNK_TOP) row = Y - HALO;
        else row = HALO - depth;
        return XY * HALO + X * row + HALO + lines * HP2 + X * (Y - depth) * (lines / depth) + b;
    } else if (face == CHUNK_LEFT || face == CHUNK_RIGHT) {
        const long height = Y - HP2, lines = b / depth;
        long col;
        if (face == CHUNK_LEFT && pack) col = HALO;
        else if (face == CHUNK_RIGHT && pack) col = X - HALO - depth;
        else if (face == CHUNK_LEFT) col = HALO - depth;
        else col = X - HALO;
        return XY * HALO + X * HALO + col + lines * (X - depth) + (lines / height) * X * HP2 + b;
    } else {
        const long height = Y - HP2, lines = b / (X - HP2);
        long plane;
        if (face == CHUNK_FRONT && pack) plane = Z - HALO - depth;
        else if (face == CHUNK_BACK && pack) plane = HALO;
        else if (face == CHUNK_FRONT) plane = Z - HALO;
        else plane = HALO - depth;
        return XY * plane + X * HALO + HALO + lines * HP2 + (lines / height) * X * HP2 + b;
    }
#endif
}

#if TL_DIM == 3
void ext_pack_message_(const int *chunk, const int *fields, const int *offsets, const int *depth, const int *face,
                       const int *fieldType, double *buffer, bool *pack)
#else
void ext_pack_message_(const int *chunk, const int *fields, const int *offsets, const int *depth, const int *face,
                       double *buffer, bool *pack)
#endif
{
#if TL_DIM == 3
    (void)fieldType;
#endif
    CHECK_CHUNK(c, *chunk);
    const int pk = *(const unsigned char *)pack != 0;
    long len;
    /* lrLineLen = yMax (*zMax), tbLineLen = xMax (*zMax), fbLineLen = xMax*yMax
       (ext_chunk.cpp:78-79; 3D :97-99) */
    switch (*face) {
    case CHUNK_LEFT: case CHUNK_RIGHT: len = (long)c->ny * c->nz * *depth; break;
    case CHUNK_TOP: case CHUNK_BOTTOM: len = (long)c->nx * c->nz * *depth; break;
#if TL_DIM == 3
    case CHUNK_BACK: case CHUNK_FRONT: len = (long)c->nx * c->ny * *depth; break;
#endif
    default: set_error("Incorrect face provided: %d.", *face); return;
    }
    for (int q = 0; q < NUM_FIELDS; q++) {
        if (!fields[q]) continue;
        double *f = NULL;
        switch (q + 1) {
        case FIELD_DENSITY: f = c->f[TF_DENSITY]; break;
        case FIELD_ENERGY0: f = c->f[TF_ENERGY0]; break;
        case FIELD_ENERGY1: f = c->f[TF_ENERGY1]; break;
        case FIELD_U: f = c->f[TF_U]; break;
        case FIELD_P: f = c->f[TF_P]; break;
        case FIELD_SD: f = c->f[TF_SD]; break;
        }
        double *ext = buffer + offsets[q];
        for (long b = 0; b < len; b++) {
            const long i = pack_index(c, *face, *depth, pk, b);
            if (pk) ext[b] = f[i];
            else f[i] = ext[b];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Shared solver methods: ext_core.cpp:260-302; ext_solver_methods.hpp */

void ext_solver_copy_u_(const int *chunk)
{
    CHECK_CHUNK(c, *chunk);
    double *u = c->f[TF_U], *u0 = c->f[TF_U0];
    FOR_FULL(c, { if (INNER(c)) u0[index] = u[index]; });
}

void ext_calculate_residual_(const int *chunk)
{
    CHECK_CHUNK(c, *chunk);
    double *u = c->f[TF_U], *u0 = c->f[TF_U0], *r = c->f[TF_R];
    FOR_FULL(c, { if (INNER(c)) { const double s = smvp(c, u, index); r[index] = u0[index] - s; } });
}

void ext_calculate_2norm_(const int *chunk, double *norm, bool *normArray)
{
    CHECK_CHUNK(c, *chunk);
    const double *buf = (*(const unsigned char *)normArray) ? c->f[TF_R] : c->f[TF_U0];
    double res[1];
    REDUCE_FULL(c, 1, res, { if (INNER(c)) acc[0] += buf[index] * buf[index]; });
    *norm = res[0];
}

void ext_solver_finalise_(const int *chunk)
{
    CHECK_CHUNK(c, *chunk);
    double *u = c->f[TF_U], *density = c->f[TF_DENSITY], *energy = c->f[TF_ENERGY1];
    FOR_FULL(c, { if (INNER(c)) energy[index] = u[index] / density[index]; });
}

/* ------------------------------------------------------------------------ */
/* Jacobi: ext_core.cpp:340-373; ext_jacobi_kernel.hpp:11-117 */

void ext_jacobi_kernel_init_(const int *chunk, const int *coefficient, const double *dt)
{
    CHECK_CHUNK(c, *chunk);
    if (*coefficient < CONDUCTIVITY && *coefficient < RECIP_CONDUCTIVITY) {
        set_error("Coefficient %d is not valid.", *coefficient);
        return;
    }
    const double rx = *dt / (c->dx * c->dx);
    const double ry = *dt / (c->dy * c->dy);
    const int coef = *coefficient;
    const long X = c->X;
    double *u = c->f[TF_U], *u0 = c->f[TF_U0], *density = c->f[TF_DENSITY];
    double *energy = c->f[TF_ENERGY1], *kx = c->f[TF_KX], *ky = c->f[TF_KY];
#if TL_DIM == 3
    const double rz = *dt / (c->dz * c->dz);
    double *kz = c->f[TF_KZ];
    const long P = X * c->Y;
#endif
    FOR_FULL(c, {
        if (ONE(c)) {
            u0[index] = energy[index] * density[index];
            u[index] = u0[index];
        }
        if (SKEW(c)) {
            double dC = (coef == CONDUCTIVITY) ? density[index] : 1.0 / density[index];
            double dL = (coef == CONDUCTIVITY) ? density[index - 1] : 1.0 / density[index - 1];
            double dD = (coef == CONDUCTIVITY) ? density[index - X] : 1.0 / density[index - X];
            kx[index] = rx * (dL + dC) / (2.0 * dL * dC);
            ky[index] = ry * (dD + dC) / (2.0 * dD * dC);
#if TL_DIM == 3
            double dB = (coef == CONDUCTIVITY) ? density[index - P] : 1.0 / density[index - P];
            kz[index] = rz * (dB + dC) / (2.0 * dB * dC);
#endif
        }
    });
}

void ext_jacobi_kernel_solve_(const int *chunk, double *error)
{
    CHECK_CHUNK(c, *chunk);
    double *u = c->f[TF_U], *u0 = c->f[TF_U0], *r = c->f[TF_R];
    const double *kx = c->f[TF_KX], *ky = c->f[TF_KY];
    const long X = c->X;
#if TL_DIM == 3
    const double *kz = c->f[TF_KZ];
    const long P = X * c->Y;
#endif
    FOR_FULL(c, { r[index] = u[index]; });
    double res[1];
    REDUCE_FULL(c, 1, res, {
        if (INNER(c)) {
#if TL_DIM == 3
            u[index] = (u0[index] + (kx[index + 1] * r[index + 1] + kx[index] * r[index - 1])
                        + (ky[index + X] * r[index + X] + ky[index] * r[index - X])
                        + (kz[index + P] * r[index + P] + kz[index] * r[index - P]))
                     / (1.0 + (kx[index] + kx[index + 1]) + (ky[index] + ky[index + X])
                        + (kz[index] + kz[index + P]));
#else
            u[index] = (u0[index] + (kx[index + 1] * r[index + 1] + kx[index] * r[index - 1])
                        + (ky[index + X] * r[index + X] + ky[index] * r[index - X]))
                     / (1.0 + (kx[index] + kx[index + 1]) + (ky[index] + ky[index + X]));
#endif
            acc[0] += fabs(u[index] - r[index]);
        }
    });
    *error = res[0];
}

/* ------------------------------------------------------------------------ */
/* CG: ext_core.cpp:377-464; ext_cg_kernel.hpp:12-319 */

void ext_cg_solver_init_(const int *chunk, const int *coefficient, const int *preconditioner,
                         double *dt, double *rro)
{
    CHECK_CHUNK(c, *chunk);
    if (*coefficient < CONDUCTIVITY && *coefficient < RECIP_CONDUCTIVITY) {
        set_error("Coefficient %d is not valid.", *coefficient);
        return;
    }
    c->preconditioner = *preconditioner;
    const int coef = *coefficient;
    const long X = c->X;
    double *p = c->f[TF_P], *r = c->f[TF_R], *u = c->f[TF_U], *density = c->f[TF_DENSITY];
    double *energy = c->f[TF_ENERGY1], *w = c->f[TF_W], *kx = c->f[TF_KX], *ky = c->f[TF_KY];
    double *mi = c->f[TF_MI], *z = c->f[TF_Z];
    /* CGInitU */
    FOR_FULL(c, { p[index] = 0.0; r[index] = 0.0; u[index] = energy[index] * density[index]; });
    const double rx = *dt / (c->dx * c->dx);
    const double ry = *dt / (c->dy * c->dy);
    /* CGInitW */
    FOR_FULL(c, { if (ONE(c)) w[index] = (coef == CONDUCTIVITY) ? density[index] : 1.0 / density[index]; });
    /* CGInitK */
#if TL_DIM == 3
    const double rz = *dt / (c->dz * c->dz);
    double *kz = c->f[TF_KZ];
    const long P = X * c->Y;
#endif
    FOR_FULL(c, {
        if (SKEW(c)) {
            kx[index] = rx * (w[index - 1] + w[index]) / (2.0 * w[index - 1] * w[index]);
            ky[index] = ry * (w[index - X] + w[index]) / (2.0 * w[index - X] * w[index]);
#if TL_DIM == 3
            kz[index] = rz * (w[index - P] + w[index]) / (2.0 * w[index - P] * w[index]);
#endif
        }
    });
    /* CGCalcRRO */
    const int pre = c->preconditioner;
    double res[1];
    REDUCE_FULL(c, 1, res, {
        if (INNER(c)) {
            const double s = smvp(c, u, index);
            w[index] = s;
            r[index] = u[index] - w[index];
            if (pre) {
#if TL_DIM == 3
                mi[index] = (1.0 + (kx[index + 1] + kx[index]) + (ky[index + X] + ky[index])
                             + (kz[index + P] + kz[index]));
#else
                mi[index] = (1.0 + (kx[index + 1] + kx[index]) + (ky[index + X] + ky[index]));
#endif
                mi[index] = 1.0 / mi[index];
                z[index] = mi[index] * r[index];
                p[index] = z[index];
            } else {
                p[index] = r[index];
            }
            acc[0] += r[index] * p[index];
        }
    });
    *rro = res[0];
}

void ext_cg_calc_w_(const int *chunk, double *pw)
{
    CHECK_CHUNK(c, *chunk);
    double *w = c->f[TF_W];
    const double *p = c->f[TF_P];
    double res[1];
    REDUCE_FULL(c, 1, res, {
        if (INNER(c)) {
            const double s = smvp(c, p, index);
            w[index] = s;
            acc[0] += w[index] * p[index];
        }
    });
    *pw = res[0];
}

void ext_cg_calc_ur_(const int *chunk, const double *alpha, double *rrn)
{
    CHECK_CHUNK(c, *chunk);
    double *u = c->f[TF_U], *r = c->f[TF_R], *z = c->f[TF_Z];
    const double *p = c->f[TF_P], *w = c->f[TF_W], *mi = c->f[TF_MI];
    const double a = *alpha;
    const int pre = c->preconditioner;
    double res[1];
    REDUCE_FULL(c, 1, res, {
        if (INNER(c)) {
            u[index] += a * p[index];
            r[index] -= a * w[index];
            if (pre) {
                z[index] = mi[index] * r[index];
                acc[0] += r[index] * z[index];
            } else {
                acc[0] += r[index] * r[index];
            }
        }
    });
    *rrn = res[0];
}

void ext_cg_calc_p_(const int *chunk, const double *beta)
{
    CHECK_CHUNK(c, *chunk);
    double *p = c->f[TF_P];
    const double *r = c->f[TF_R], *z = c->f[TF_Z];
    const double b = *beta;
    const int pre = c->preconditioner;
    FOR_FULL(c, { if (INNER(c)) p[index] = b * p[index] + (pre ? z[index] : r[index]); });
}

/* ------------------------------------------------------------------------ */
/* Chebyshev: ext_core.cpp:468-515; ext_cheby_kernel.hpp:11-115 */

static void store_coefs(Chunk *c, const double *alphas, const double *betas, int n)
{
    free(c->alphas); free(c->betas);
    c->ncoefs = n;
    c->alphas = calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    c->betas = calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    if (n > 0) {
        memcpy(c->alphas, alphas, sizeof(double) * (size_t)n);
        memcpy(c->betas, betas, sizeof(double) * (size_t)n);
    }
}

static void cheby_calc_u(Chunk *c)
{
    double *u = c->f[TF_U];
    const double *p = c->f[TF_P];
    FOR_FULL(c, { if (INNER(c)) u[index] += p[index]; });
}

void ext_cheby_solver_init_(const int *chunk, double *alphas, double *betas, const double *theta,
                            const int *preconditioner, const int *maxChebyIters)
{
    CHECK_CHUNK(c, *chunk);
    c->preconditioner = *preconditioner;
    store_coefs(c, alphas, betas, *maxChebyIters);
    double *p = c->f[TF_P], *r = c->f[TF_R], *w = c->f[TF_W];
    const double *u = c->f[TF_U], *u0 = c->f[TF_U0], *mi = c->f[TF_MI];
    const double th = *theta;
    const int pre = *preconditioner;
    FOR_FULL(c, {
        if (INNER(c)) {
            const double s = smvp(c, u, index);
            w[index] = s;
            r[index] = u0[index] - w[index];
            p[index] = (pre ? mi[index] * r[index] : r[index]) / th;
        }
    });
    cheby_calc_u(c);
}

void ext_cheby_solver_iterate_(const int *chunk, const int *chebyCalcStep)
{
    CHECK_CHUNK(c, *chunk);
    const int step = *chebyCalcStep - 1;
    if (step < 0 || step >= c->ncoefs) {
        set_error("Chebyshev step %d outside the %d stored coefficients", step + 1, c->ncoefs);
        return;
    }
    double *p = c->f[TF_P], *r = c->f[TF_R], *w = c->f[TF_W];
    const double *u = c->f[TF_U], *u0 = c->f[TF_U0], *mi = c->f[TF_MI];
    const double a = c->alphas[step], b = c->betas[step];
    const int pre = c->preconditioner;
    FOR_FULL(c, {
        if (INNER(c)) {
            const double s = smvp(c, u, index);
            w[index] = s;
            r[index] = u0[index] - w[index];
            p[index] = a * p[index] + b * (pre ? mi[index] * r[index] : r[index]);
        }
    });
    cheby_calc_u(c);
}

/* ------------------------------------------------------------------------ */
/* PPCG: ext_core.cpp:519-567; ext_ppcg_kernel.hpp:8-100 */

void ext_ppcg_init_(const int *chunk, double *alphas, double *betas, const double *theta,
                    const int *maxChebyIters)
{
    CHECK_CHUNK(c, *chunk);
    c->theta = *theta;
    store_coefs(c, alphas, betas, *maxChebyIters);
}

void ext_ppcg_init_sd_(const int *chunk)
{
    CHECK_CHUNK(c, *chunk);
    double *sd = c->f[TF_SD];
    const double *r = c->f[TF_R], *mi = c->f[TF_MI];
    const double th = c->theta;
    const int pre = c->preconditioner;
    FOR_FULL(c, { if (INNER(c)) sd[index] = (pre ? mi[index] * r[index] : r[index]) / th; });
}

void ext_ppcg_inner_(const int *chunk, int *currentStep)
{
    CHECK_CHUNK(c, *chunk);
    const int step = *currentStep - 1;
    if (step < 0 || step >= c->ncoefs) {
        set_error("PPCG step %d outside the %d stored coefficients", step + 1, c->ncoefs);
        return;
    }
    double *sd = c->f[TF_SD], *r = c->f[TF_R], *u = c->f[TF_U];
    const double *mi = c->f[TF_MI];
    /* PPCGCalcU */
    FOR_FULL(c, {
        if (INNER(c)) {
            const double s = smvp(c, sd, index);
            r[index] -= s;
            u[index] += sd[index];
        }
    });
    /* PPCGCalcSd */
    const double a = c->alphas[step], b = c->betas[step];
    const int pre = c->preconditioner;
    FOR_FULL(c, { if (INNER(c)) sd[index] = a * sd[index] + b * (pre ? mi[index] * r[index] : r[index]); });
}

/* ------------------------------------------------------------------------ */
/* Timer for the Fortran side: timer_c.cpp:30 */
void timer_c_(double *t)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    *t = (double)tv.tv_sec + 1.0e-6 * (double)tv.tv_usec;
}

static double wall(void)
{
    double t;
    timer_c_(&t);
    return t;
}

/* ======================================================================== */
/* Host driver: the Fortran caller restated.                                 */
/* ======================================================================== */

typedef struct {
    int defined;
    double density, energy, xmin, xmax, ymin, ymax, zmin, zmax, radius;
    int geometry;
} State;

typedef struct {
    double xmin, ymin, zmin, xmax, ymax, zmax;
    int x_cells, y_cells, z_cells;
    double dtinit, end_time, eps, epslim;
    int end_step, max_iters, summary_frequency, presteps, ppcg_inner_steps;
    int errswitch, check_result, preconditioner, use_cheby, use_cg, use_ppcg, use_jacobi;
    int coefficient, test_problem;
    int nstates;
    State *states;
} Deck;

/* Defaults: tealeaf_2d/read_input.f90:39-80.  Literals the Fortran writes
 * without a kind suffix are single precision and are widened here the same
 * way (0.1 -> 0.100000001490116...). */
static void deck_defaults(Deck *d)
{
    memset(d, 0, sizeof *d);
    d->xmin = 0.0; d->ymin = 0.0; d->zmin = 0.0;
    d->xmax = 100.0; d->ymax = 100.0; d->zmax = 100.0;
    d->x_cells = d->y_cells = d->z_cells = 10;
    d->dtinit = (double)0.1f;
    d->max_iters = 1000;
    d->eps = (double)1.0e-10f;
    d->end_time = 10.0;
    d->end_step = 640000; /* g_ibig, data.f90:28 */
    d->summary_frequency = 10;
    d->coefficient = CONDUCTIVITY;
    d->presteps = 30;
    d->epslim = (double)1e-5f;
    d->ppcg_inner_steps = 10;
    d->use_jacobi = 1;
}

/* Line reader of parse.f90:130-255 (parse_getline): lines are cut at 100
 * characters, control characters blanked, '!' and ';' start comments, text
 * is lower-cased, '=' and ',' separate words, and only lines inside the
 * '*tea' ... '*endtea' block count. */
typedef struct {
    char **lines;
    int nlines;
} DeckLines;

static void trim_inplace(char *s)
{
    size_t n = strlen(s);
    while (n > 0 && s[n - 1] == ' ') s[--n] = 0;
    size_t k = 0;
    while (s[k] == ' ') k++;
    if (k) memmove(s, s + k, strlen(s + k) + 1);
}

static int deck_lines(const char *text, DeckLines *out)
{
    out->lines = NULL;
    out->nlines = 0;
    char here[512] = "";
    const char *p = text;
    while (*p) {
        const char *e = strchr(p, '\n');
        size_t len = e ? (size_t)(e - p) : strlen(p);
        char l[128];
        size_t take = len < 100 ? len : 100;
        memcpy(l, p, take);
        l[take] = 0;
        p = e ? e + 1 : p + len;
        for (size_t i = 0; i < take; i++)
            if ((unsigned char)l[i] < 32 || (unsigned char)l[i] > 128) l[i] = ' ';
        trim_inplace(l);
        char *s = strchr(l, '!');
        if (s) { *s = 0; trim_inplace(l); }
        s = strchr(l, ';');
        if (s) { *s = 0; trim_inplace(l); }
        for (char *q = l; *q; q++) {
            if (*q >= 'A' && *q <= 'Z') *q = (char)(*q + 32);
            if (*q == '=' || *q == ',') *q = ' ';
        }
        if (l[0] == '*') {
            char nugget[128];
            size_t k = strcspn(l, " ");
            memcpy(nugget, l, k);
            nugget[k] = 0;
            if (strncmp(nugget + 1, "end", 3) == 0) {
                char want[128];
                snprintf(want, sizeof want, "*%s", nugget + 4);
                char *last = strrchr(here, '*');
                if (!last || strcmp(last, want) != 0) {
                    set_error("Unmatched */*end pair.");
                    return -1;
                }
                if (strcmp(here, "*tea") == 0) return 0; /* end of the deck */
                *last = 0;
            } else {
                strncat(here, nugget, sizeof here - strlen(here) - 1);
            }
            continue;
        }
        if (strcmp(here, "*tea") != 0 || l[0] == 0) continue;
        out->lines = realloc(out->lines, sizeof(char *) * (size_t)(out->nlines + 1));
        out->lines[out->nlines++] = strdup(l);
    }
    return 0;
}

static void deck_lines_free(DeckLines *dl)
{
    for (int i = 0; i < dl->nlines; i++) free(dl->lines[i]);
    free(dl->lines);
}

/* Word cursor over the deck lines (parse_getword with wrap). */
typedef struct {
    DeckLines *dl;
    int line;
    const char *pos;
} Cursor;

static int next_word(Cursor *cur, char *word, size_t cap, int wrap)
{
    for (;;) {
        if (cur->line >= cur->dl->nlines) { word[0] = 0; return 0; }
        const char *p = cur->pos;
        while (*p == ' ') p++;
        if (*p) {
            size_t k = strcspn(p, " ");
            if (k >= cap) k = cap - 1;
            memcpy(word, p, k);
            word[k] = 0;
            cur->pos = p + k;
            return 1;
        }
        if (!wrap) { word[0] = 0; return 0; }
        cur->line++;
        if (cur->line < cur->dl->nlines) cur->pos = cur->dl->lines[cur->line];
    }
}

static int next_line(Cursor *cur)
{
    if (cur->pos == NULL) {
        cur->line = 0;
    } else {
        cur->line++;
    }
    if (cur->line >= cur->dl->nlines) return 0;
    cur->pos = cur->dl->lines[cur->line];
    return 1;
}

/* parse_getival: READ with (I7) */
static int parse_ival(const char *w, int *out)
{
    char buf[8];
    size_t n = strlen(w);
    if (n > 7) n = 7;
    memcpy(buf, w, n);
    buf[n] = 0;
    char *end;
    long v = strtol(buf, &end, 10);
    while (*end == ' ') end++;
    if (end == buf || *end) { set_error("Error attempting to convert to integer:%s", w); return -1; }
    *out = (int)v;
    return 0;
}

/* parse_getrval: a word without '.' gets '.0' appended, then READ (E27.20)
 * (parse.f90:332-354). */
static int parse_rval(const char *w, double *out)
{
    char buf[160];
    snprintf(buf, sizeof buf, "%s%s", w, strchr(w, '.') ? "" : ".0");
    char *end;
    double v = strtod(buf, &end);
    if (end == buf || *end) { set_error("Error attempting to convert to real:%s", w); return -1; }
    *out = v;
    return 0;
}

/* read_input.f90:83-297 */
static int deck_parse(const char *text, Deck *d)
{
    deck_defaults(d);
    DeckLines dl;
    if (deck_lines(text, &dl)) return -1;
    char w[128], v[128];
    int state_max = 0;
    Cursor cur = {&dl, 0, NULL};
    while (next_line(&cur)) {
        while (next_word(&cur, w, sizeof w, 0)) {
            if (strcmp(w, "state") == 0) {
                int s;
                next_word(&cur, v, sizeof v, 1);
                if (parse_ival(v, &s)) { deck_lines_free(&dl); return -1; }
                if (s > state_max) state_max = s;
                break;
            }
        }
    }
    if (state_max < 1) { set_error("No states defined."); deck_lines_free(&dl); return -1; }
    d->nstates = state_max;
    d->states = calloc((size_t)state_max, sizeof(State));
    cur.line = 0; cur.pos = NULL;
    int rc = 0;
#define RVAL(dst) do { next_word(&cur, v, sizeof v, 1); if (parse_rval(v, &(dst))) { rc = -1; goto done; } } while (0)
#define IVAL(dst) do { next_word(&cur, v, sizeof v, 1); if (parse_ival(v, &(dst))) { rc = -1; goto done; } } while (0)
    while (next_line(&cur)) {
        while (next_word(&cur, w, sizeof w, 0)) {
            if (!strcmp(w, "initial_timestep")) RVAL(d->dtinit);
            else if (!strcmp(w, "end_time")) RVAL(d->end_time);
            else if (!strcmp(w, "end_step")) IVAL(d->end_step);
            else if (!strcmp(w, "xmin")) RVAL(d->xmin);
            else if (!strcmp(w, "xmax")) RVAL(d->xmax);
            else if (!strcmp(w, "ymin")) RVAL(d->ymin);
            else if (!strcmp(w, "ymax")) RVAL(d->ymax);
            else if (!strcmp(w, "x_cells")) IVAL(d->x_cells);
            else if (!strcmp(w, "y_cells")) IVAL(d->y_cells);
#if TL_DIM == 3
            else if (!strcmp(w, "zmin")) RVAL(d->zmin);
            else if (!strcmp(w, "zmax")) RVAL(d->zmax);
            else if (!strcmp(w, "z_cells")) IVAL(d->z_cells);
#endif
            else if (!strcmp(w, "visit_frequency")) { int tmp; IVAL(tmp); }
            else if (!strcmp(w, "summary_frequency")) IVAL(d->summary_frequency);
            else if (!strcmp(w, "tl_ch_cg_presteps")) IVAL(d->presteps);
            else if (!strcmp(w, "tl_ppcg_inner_steps")) IVAL(d->ppcg_inner_steps);
            else if (!strcmp(w, "tl_ch_cg_epslim")) RVAL(d->epslim);
            else if (!strcmp(w, "tl_check_result")) d->check_result = 1;
            else if (!strcmp(w, "tl_ch_cg_errswitch")) d->errswitch = 1;
            else if (!strcmp(w, "tl_preconditioner_on")) d->preconditioner = 1;
            else if (!strcmp(w, "tl_use_jacobi")) { d->use_cheby = d->use_cg = d->use_ppcg = 0; d->use_jacobi = 1; }
            else if (!strcmp(w, "tl_use_cg")) { d->use_cheby = d->use_ppcg = d->use_jacobi = 0; d->use_cg = 1; }
            else if (!strcmp(w, "tl_use_ppcg")) { d->use_cheby = d->use_cg = d->use_jacobi = 0; d->use_ppcg = 1; }
            else if (!strcmp(w, "tl_use_chebyshev")) { d->use_cg = d->use_ppcg = d->use_jacobi = 0; d->use_cheby = 1; }
            else if (!strcmp(w, "tl_max_iters")) IVAL(d->max_iters);
            else if (!strcmp(w, "tl_eps")) RVAL(d->eps);
            else if (!strcmp(w, "tl_coefficient_density")) d->coefficient = CONDUCTIVITY;
            else if (!strcmp(w, "tl_coefficient_inverrse_density")) d->coefficient = RECIP_CONDUCTIVITY;
            else if (!strcmp(w, "test_problem")) IVAL(d->test_problem);
            else if (!strcmp(w, "state")) {
                int s;
                IVAL(s);
                if (s < 1 || s > d->nstates) { set_error("state %d out of range", s); rc = -1; goto done; }
                State *st = &d->states[s - 1];
                if (st->defined) { set_error("State defined twice."); rc = -1; goto done; }
                st->defined = 1;
                while (next_word(&cur, w, sizeof w, 0)) {
                    if (!strcmp(w, "xmin")) RVAL(st->xmin);
                    else if (!strcmp(w, "ymin")) RVAL(st->ymin);
                    else if (!strcmp(w, "xmax")) RVAL(st->xmax);
                    else if (!strcmp(w, "ymax")) RVAL(st->ymax);
#if TL_DIM == 3
                    else if (!strcmp(w, "zmin")) RVAL(st->zmin);
                    else if (!strcmp(w, "zmax")) RVAL(st->zmax);
#endif
                    else if (!strcmp(w, "radius")) RVAL(st->radius);
                    else if (!strcmp(w, "density")) RVAL(st->density);
                    else if (!strcmp(w, "energy")) RVAL(st->energy);
                    else if (!strcmp(w, "geometry")) {
                        next_word(&cur, v, sizeof v, 1);
#if TL_DIM == 3
                        if (!strcmp(v, "cuboid")) st->geometry = G_RECT;
#else
                        if (!strcmp(v, "rectangle")) st->geometry = G_RECT;
#endif
                        else if (!strcmp(v, "circle")) st->geometry = G_CIRC;
                        else if (!strcmp(v, "point")) st->geometry = G_POINT;
                    }
                }
            }
        }
    }
    /* State shrink: read_input.f90:284-297 */
    {
        const double dx = (d->xmax - d->xmin) / (double)(float)d->x_cells;
        const double dy = (d->ymax - d->ymin) / (double)(float)d->y_cells;
#if TL_DIM == 3
        const double dz = (d->zmax - d->zmin) / (double)(float)d->z_cells;
#endif
        for (int n = 1; n < d->nstates; n++) {
            d->states[n].xmin = d->states[n].xmin + (dx / 100.0);
            d->states[n].ymin = d->states[n].ymin + (dy / 100.0);
            d->states[n].xmax = d->states[n].xmax - (dx / 100.0);
            d->states[n].ymax = d->states[n].ymax - (dy / 100.0);
#if TL_DIM == 3
            d->states[n].zmin = d->states[n].zmin + (dz / 100.0);
            d->states[n].zmax = d->states[n].zmax - (dz / 100.0);
#endif
        }
    }
done:
#undef RVAL
#undef IVAL
    deck_lines_free(&dl);
    if (rc) { free(d->states); d->states = NULL; }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* tqli (tea_leaf_cheby.f90:233-283), 1-based arrays d[1..n], e[1..n]. */
static void tqli(double *d, double *e, int n, int *info)
{
    for (int i = 1; i < n; i++) e[i] = e[i + 1]; /* e = eoshift(e, 1) */
    if (n >= 1) e[n] = 0.0;
    *info = 0;
    for (int l = 1; l <= n; l++) {
        int iter = 0;
        for (;;) {
            int m;
            for (m = l; m <= n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) + dd == dd) break;
            }
            if (m == l) break;
            if (iter == 30) { *info = 1; return; }
            iter++;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = sqrt(g * g + 1.0);
            g = d[m] - d[l] + e[l] / (g + copysign(fabs(r), g));
            double s = 1.0, c = 1.0, p = 0.0;
            int restart = 0;
            for (int i = m - 1; i >= l; i--) {
                double f = s * e[i];
                double b = c * e[i];
                r = sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] = d[i + 1] - p;
                    e[m] = 0.0;
                    restart = 1;
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
}

/* tea_calc_eigenvalues (tea_leaf_cheby.f90:285-332); alphas/betas 1-based. */
static void calc_eigenvalues(const double *cg_alphas, const double *cg_betas, double *eigmin,
                             double *eigmax, int m, int *info)
{
    double *diag = calloc((size_t)m + 2, sizeof(double));
    double *offdiag = calloc((size_t)m + 2, sizeof(double));
    for (int n = 1; n <= m; n++) {
        diag[n] = 1.0 / cg_alphas[n];
        if (n > 1) diag[n] = diag[n] + cg_betas[n - 1] / cg_alphas[n - 1];
        if (n < m) offdiag[n + 1] = sqrt(cg_betas[n]) / cg_alphas[n];
    }
    tqli(diag, offdiag, m, info);
    if (*info == 0) {
        /* The reference bubble-sorts ascending; only the extremes are used. */
        for (int pass = 1; pass < m; pass++)
            for (int n = 1; n <= m - pass; n++)
                if (diag[n] > diag[n + 1]) { double t = diag[n]; diag[n] = diag[n + 1]; diag[n + 1] = t; }
        *eigmin = diag[1];
        *eigmax = diag[m];
        if (*eigmin < 0.0 || *eigmax < 0.0) *info = 1;
    }
    free(diag);
    free(offdiag);
}

/* tea_calc_ch_coefs (tea_leaf_cheby.f90:334-360); arrays 1-based. */
static void calc_ch_coefs(double *ch_alphas, double *ch_betas, double eigmin, double eigmax,
                          double *theta, int max_cheby_iters)
{
    *theta = (eigmax + eigmin) / 2;
    double delta = (eigmax - eigmin) / 2;
    double sigma = *theta / delta;
    double rho_old = 1.0 / sigma;
    for (int n = 1; n <= max_cheby_iters; n++) {
        double rho_new = 1.0 / (2.0 * sigma - rho_old);
        double cur_alpha = rho_new * rho_old;
        double cur_beta = 2.0 * rho_new / delta;
        ch_alphas[n] = cur_alpha;
        ch_betas[n] = cur_beta;
        rho_old = rho_new;
    }
}

/* Exported for tests that pin the eigenvalue and coefficient routines. */
int tl_calc_eigenvalues(const double *alphas0, const double *betas0, int m, double *eig)
{
    double *a = calloc((size_t)m + 2, sizeof(double)), *b = calloc((size_t)m + 2, sizeof(double));
    for (int i = 0; i < m; i++) { a[i + 1] = alphas0[i]; b[i + 1] = betas0[i]; }
    int info = 0;
    calc_eigenvalues(a, b, &eig[0], &eig[1], m, &info);
    free(a); free(b);
    return info;
}

void tl_calc_ch_coefs(double eigmin, double eigmax, int n, double *alphas0, double *betas0, double *theta)
{
    double *a = calloc((size_t)n + 2, sizeof(double)), *b = calloc((size_t)n + 2, sizeof(double));
    calc_ch_coefs(a, b, eigmin, eigmax, theta, n);
    for (int i = 0; i < n; i++) { alphas0[i] = a[i + 1]; betas0[i] = b[i + 1]; }
    free(a); free(b);
}

/* ------------------------------------------------------------------------ */

typedef struct {
    char *buf;
    size_t cap, len;
} Out;

static void outf(Out *o, const char *fmt, ...)
{
    if (!o->buf || o->len >= o->cap) return;
    va_list ap;
    va_start(ap, fmt);
    int k = vsnprintf(o->buf + o->len, o->cap - o->len, fmt, ap);
    va_end(ap);
    if (k > 0) o->len += (size_t)k;
    if (o->len >= o->cap) o->len = o->cap;
}

typedef struct {
    int iters;
    double error, exact_error;
    int switch_step, est_itc, cheby_steps, cg_steps;
    double eigmin, eigmax;
    double solve_s;
} StepInfo;

#if TL_DIM == 3
static const int g_neigh[6] = {EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE};
#else
static const int g_neigh[4] = {EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE, EXTERNAL_FACE};
#endif

static void update_halo(int chunk, const int *fields, int depth)
{
    ext_update_halo_kernel_(&chunk, g_neigh, fields, &depth);
}

static void set_fields(int *fields, int a, int b)
{
    for (int k = 0; k < NUM_FIELDS; k++) fields[k] = 0;
    if (a) fields[a - 1] = 1;
    if (b) fields[b - 1] = 1;
}

/* tea_leaf (tea_solve.f90:36-605), one chunk, ext kernels. */
static int tea_leaf(const Deck *d, int chunk, double dt, StepInfo *si)
{
    int fields[NUM_FIELDS];
    const int max_iters = d->max_iters;
    double error = 1e10, exact_error = 0.0;
    int cheby_calc_steps = 0, cg_calc_steps = 0;
    double rro = 0.0, pw = 0.0, rrn = 0.0, alpha, beta;
    double eigmin = 0.0, eigmax = 0.0, theta = 0.0, cn = 0.0;
    int est_itc = 0, max_cheby_iters = 0, info = 0, switch_step = 0;
    double *cg_alphas = calloc((size_t)max_iters + 3, sizeof(double));
    double *cg_betas = calloc((size_t)max_iters + 3, sizeof(double));
    double *ch_alphas = calloc((size_t)max_iters + 3, sizeof(double));
    double *ch_betas = calloc((size_t)max_iters + 3, sizeof(double));
    const int pre = d->preconditioner;
    const int one = 1, zero = 0;
    double t0 = wall();

    set_fields(fields, FIELD_ENERGY1, FIELD_DENSITY);
    update_halo(chunk, fields, 2);

    if (d->use_cg || d->use_cheby || d->use_ppcg) {
        ext_cg_solver_init_(&chunk, &d->coefficient, &pre, &dt, &rro);
        set_fields(fields, FIELD_U, FIELD_P);
        update_halo(chunk, fields, 1);
    } else if (d->use_jacobi) {
        ext_jacobi_kernel_init_(&chunk, &d->coefficient, &dt);
        set_fields(fields, FIELD_U, 0);
    }
    ext_solver_copy_u_(&chunk);
    if (g_err) goto fail;

    int n;
    for (n = 1; n <= max_iters; n++) {
        int ch_switch_check;
        if (d->errswitch)
            ch_switch_check = (cheby_calc_steps > 0) || ((error <= d->epslim) && (n >= 20));
        else
            ch_switch_check = (n >= d->presteps) && (error <= 1.0);

        if ((d->use_cheby || d->use_ppcg) && ch_switch_check) {
            if (cheby_calc_steps == 0) {
                max_cheby_iters = max_iters - n + 2;
                rro = error;
                calc_eigenvalues(cg_alphas, cg_betas, &eigmin, &eigmax, n - 1, &info);
                if (info != 0) { set_error("tea_leaf: Error in calculating eigenvalues"); goto fail; }
                eigmin = eigmin * (double)0.95f;
                eigmax = eigmax * (double)1.05f;
                if (d->use_cheby) {
                    calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, &theta, max_cheby_iters);
                    set_fields(fields, FIELD_U, 0);
                } else {
                    calc_ch_coefs(ch_alphas, ch_betas, eigmin, eigmax, &theta, d->ppcg_inner_steps);
                }
                cn = eigmax / eigmin;
            }
            if (d->use_cheby) {
                if (cheby_calc_steps == 0) {
                    /* tea_leaf_cheby_first_step, tea_solve.f90:658-750 */
                    double bb;
                    bool f0 = false, f1 = true;
                    ext_calculate_2norm_(&chunk, &bb, &f0);
                    ext_cheby_solver_init_(&chunk, ch_alphas + 1, ch_betas + 1, &theta, &pre, &max_cheby_iters);
                    update_halo(chunk, fields, 1);
                    ext_cheby_solver_iterate_(&chunk, &max_cheby_iters);
                    ext_calculate_2norm_(&chunk, &error, &f1);
                    double it_alpha = DBL_EPSILON * bb / (4.0 * error);
                    double gamm = (sqrt(cn) - 1.0) / (sqrt(cn) + 1.0);
                    est_itc = (int)lround(log(it_alpha) / (2.0 * log(gamm)));
                    cheby_calc_steps = 1;
                    switch_step = n;
                } else {
                    ext_cheby_solver_iterate_(&chunk, &cheby_calc_steps);
                    if ((n >= est_itc) && (n % 10 == 0)) {
                        bool f1 = true;
                        ext_calculate_2norm_(&chunk, &error, &f1);
                    }
                }
            } else if (d->use_ppcg) {
                if (cheby_calc_steps == 0) {
                    cheby_calc_steps = 1;
                    ext_ppcg_init_(&chunk, ch_alphas + 1, ch_betas + 1, &theta, &max_cheby_iters);
                    ext_calculate_residual_(&chunk);
                    update_halo(chunk, fields, 1);
                    switch_step = n;
                }
                ext_cg_calc_w_(&chunk, &pw);
                alpha = rro / pw;
                ext_cg_calc_ur_(&chunk, &alpha, &rrn);
                /* tea_leaf_run_ppcg_inner_steps, tea_solve.f90:607-656 */
                ext_ppcg_init_sd_(&chunk);
                int sdf[NUM_FIELDS];
                set_fields(sdf, FIELD_SD, 0);
                for (int step = 1; step <= d->ppcg_inner_steps; step++) {
                    update_halo(chunk, sdf, 1);
                    ext_ppcg_inner_(&chunk, &step);
                }
                bool f1 = true;
                ext_calculate_2norm_(&chunk, &rrn, &f1);
                beta = rrn / rro;
                ext_cg_calc_p_(&chunk, &beta);
                error = rrn;
                rro = rrn;
            }
            cheby_calc_steps = cheby_calc_steps + 1;
        } else if (d->use_cg || d->use_cheby || d->use_ppcg) {
            fields[FIELD_P - 1] = 1;
            cg_calc_steps = cg_calc_steps + 1;
            ext_cg_calc_w_(&chunk, &pw);
            alpha = rro / pw;
            if (d->use_cheby || d->use_ppcg) cg_alphas[n] = alpha;
            ext_cg_calc_ur_(&chunk, &alpha, &rrn);
            beta = rrn / rro;
            if (d->use_cheby || d->use_ppcg) cg_betas[n] = beta;
            ext_cg_calc_p_(&chunk, &beta);
            error = rrn;
            rro = rrn;
        } else if (d->use_jacobi) {
            ext_jacobi_kernel_solve_(&chunk, &error);
            if (n % 50 == 0) {
                bool f1 = true;
                update_halo(chunk, fields, 1);
                ext_calculate_residual_(&chunk);
                ext_calculate_2norm_(&chunk, &error, &f1);
            }
        }
        update_halo(chunk, fields, 1);
        if (g_err) goto fail;
        if (fabs(error) < d->eps) break;
    }
    if (d->check_result) {
        bool f1 = true;
        ext_calculate_residual_(&chunk);
        ext_calculate_2norm_(&chunk, &exact_error, &f1);
    }
    si->iters = n - 1;
    si->error = error;
    si->exact_error = exact_error;
    si->switch_step = switch_step;
    si->est_itc = est_itc;
    si->eigmin = eigmin;
    si->eigmax = eigmax;
    si->cheby_steps = cheby_calc_steps;
    si->cg_steps = cg_calc_steps;
    (void)one; (void)zero;

    ext_solver_finalise_(&chunk);
    set_fields(fields, FIELD_ENERGY1, 0);
    update_halo(chunk, fields, 1);
    si->solve_s = wall() - t0;
    free(cg_alphas); free(cg_betas); free(ch_alphas); free(ch_betas);
    return g_err ? -1 : 0;
fail:
    free(cg_alphas); free(cg_betas); free(ch_alphas); free(ch_betas);
    return -1;
}

static void summary_json(Out *o, int chunk, int step, double time)
{
    double vol, mass, ie, temp;
    ext_field_summary_kernel_(&chunk, &vol, &mass, &ie, &temp);
    outf(o, "\"step\": %d, \"time\": %.17g, \"vol\": %.17g, \"mass\": %.17g, \"ie\": %.17g, \"temp\": %.17g",
         step, time, vol, mass, ie, temp);
}

/* Runs a deck the way tea_leaf.f90 -> initialise -> start -> diffuse does
 * (start.f90:24-162, diffuse.f90:23-160), one chunk covering the mesh.
 * Writes a JSON report into json[0..cap).  max_steps > 0 caps the number of
 * timesteps (a bounded sample); 0 runs the deck as written. */
int tl_run_deck(const char *deck_text, char *json, int cap, int max_steps)
{
    tl_clear_error();
    Out o = {json, (size_t)cap, 0};
    if (json && cap > 0) json[0] = 0;
    Deck d;
    if (deck_parse(deck_text, &d)) return -1;

    ext_finalise_();
    int rank = 0, chunk = 1;
    int nx = d.x_cells, ny = d.y_cells;
#if TL_DIM == 3
    int nz = d.z_cells;
    if (nx < 2 || ny < 2 || nz < 2) { set_error("meshes need at least 2 cells per axis"); free(d.states); return -1; }
    ext_init_(&nx, &ny, &nz, &rank);
    const double dz = (d.zmax - d.zmin) / (double)(float)d.z_cells;
#else
    if (nx < 2 || ny < 2) { set_error("meshes need at least 2 cells per axis"); free(d.states); return -1; }
    ext_init_(&nx, &ny, &rank);
#endif
    const double dx = (d.xmax - d.xmin) / (double)(float)d.x_cells;
    const double dy = (d.ymax - d.ymin) / (double)(float)d.y_cells;
#if TL_DIM == 3
    ext_initialise_chunk_(&chunk, &d.xmin, &d.ymin, &d.zmin, &dx, &dy, &dz);
#else
    ext_initialise_chunk_(&chunk, &d.xmin, &d.ymin, &dx, &dy);
#endif
    const int ns = d.nstates;
    double *sd = calloc((size_t)ns, sizeof(double)), *se = calloc((size_t)ns, sizeof(double));
    double *sx0 = calloc((size_t)ns, sizeof(double)), *sx1 = calloc((size_t)ns, sizeof(double));
    double *sy0 = calloc((size_t)ns, sizeof(double)), *sy1 = calloc((size_t)ns, sizeof(double));
    double *sz0 = calloc((size_t)ns, sizeof(double)), *sz1 = calloc((size_t)ns, sizeof(double));
    double *sr = calloc((size_t)ns, sizeof(double));
    int *sg = calloc((size_t)ns, sizeof(int));
    for (int k = 0; k < ns; k++) {
        sd[k] = d.states[k].density; se[k] = d.states[k].energy;
        sx0[k] = d.states[k].xmin; sx1[k] = d.states[k].xmax;
        sy0[k] = d.states[k].ymin; sy1[k] = d.states[k].ymax;
        sz0[k] = d.states[k].zmin; sz1[k] = d.states[k].zmax;
        sr[k] = d.states[k].radius; sg[k] = d.states[k].geometry;
    }
    const int gr = G_RECT, gc = G_CIRC, gp = G_POINT;
#if TL_DIM == 3
    ext_generate_chunk_(&chunk, &ns, sd, se, sx0, sx1, sy0, sy1, sz0, sz1, sr, sg, &gr, &gc, &gp);
#else
    ext_generate_chunk_(&chunk, &ns, sd, se, sx0, sx1, sy0, sy1, sr, sg, &gr, &gc, &gp);
    (void)sz0; (void)sz1;
#endif
    free(sd); free(se); free(sx0); free(sx1); free(sy0); free(sy1); free(sz0); free(sz1); free(sr); free(sg);

    int fields[NUM_FIELDS];
    for (int k = 0; k < NUM_FIELDS; k++) fields[k] = 0;
    fields[FIELD_DENSITY - 1] = fields[FIELD_ENERGY0 - 1] = fields[FIELD_ENERGY1 - 1] = 1;
    update_halo(chunk, fields, 2);

    outf(&o, "{\"dim\": %d, \"x_cells\": %d, \"y_cells\": %d, \"z_cells\": %d, \"reduce_parts\": %d, \"threads\": %d, ",
         TL_DIM, d.x_cells, d.y_cells, TL_DIM == 3 ? d.z_cells : 1, g_parts, omp_get_max_threads());
    outf(&o, "\"solver\": \"%s\", \"eps\": %.17g, \"dtinit\": %.17g, \"max_iters\": %d, ",
         d.use_cg ? "cg" : d.use_cheby ? "chebyshev" : d.use_ppcg ? "ppcg" : "jacobi", d.eps, d.dtinit, d.max_iters);
    outf(&o, "\"initial\": {");
    summary_json(&o, chunk, 0, 0.0);
    outf(&o, "}, \"steps\": [");

    ext_set_field_kernel_(&chunk);
    double time = 0.0, dt = d.dtinit;
    int step = 0, rc = 0;
    double total_solve = 0.0;
    const double g_small = (double)1.0e-16f;
    for (;;) {
        step = step + 1;
        /* timestep / calc_dt (timestep.f90, calc_dt.f90:35) */
        if (d.dtinit <= dt) dt = d.dtinit;
        StepInfo si;
        memset(&si, 0, sizeof si);
        if (tea_leaf(&d, chunk, dt, &si)) { rc = -1; break; }
        total_solve += si.solve_s;
        time = time + dt;
        outf(&o, "%s{", step > 1 ? ", " : "");
        outf(&o, "\"iters\": %d, \"error\": %.17g, \"exact_error\": %.17g, \"switch_step\": %d, "
                 "\"est_itc\": %d, \"eigmin\": %.17g, \"eigmax\": %.17g, \"cg_steps\": %d, "
                 "\"cheby_steps\": %d, \"solve_s\": %.9g, \"dt\": %.17g, ",
             si.iters, si.error, si.exact_error, si.switch_step, si.est_itc, si.eigmin, si.eigmax,
             si.cg_steps, si.cheby_steps, si.solve_s, dt);
        if (d.summary_frequency != 0 && step % d.summary_frequency == 0) {
            summary_json(&o, chunk, step, time);
        } else {
            outf(&o, "\"step\": %d, \"time\": %.17g", step, time);
        }
        outf(&o, "}");
        if (time + g_small > d.end_time || step >= d.end_step || (max_steps > 0 && step >= max_steps)) break;
    }
    if (rc == 0) {
        outf(&o, "], \"final\": {");
        summary_json(&o, chunk, step, time);
        outf(&o, "}, \"test_problem\": %d, \"total_solve_s\": %.9g}", d.test_problem, total_solve);
        if (o.buf && o.len >= o.cap) { set_error("JSON buffer too small"); rc = -2; }
    }
    free(d.states);
    if (!g_keep_chunk) ext_finalise_();
    return rc;
}
