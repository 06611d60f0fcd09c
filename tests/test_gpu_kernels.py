"""Per-entry-point parity of the HIP library against the oracle.

Both libraries export the reference's ext_*_ C ABI; each test builds the same
chunk state in both (through the ABI and the tl_set_field hook), calls the
same entry point on both and compares every array the entry point writes.
Per-cell values must be bitwise equal: both sides evaluate the reference's
expressions in the same operand order without FMA contraction.  Reductions
are FP64 sums in a different order (tree on the GPU, Kokkos-OpenMP blocks in
the oracle), so they are compared to a relative 1e-12 of the sum of
magnitudes.

Stencil inputs: the reference reads the depth-1 halo from memory after
update_halo; the HIP kernels read the mirrored interior value.  The tests
therefore refresh the halos of stencil inputs exactly where the reference's
driver does (tea_solve.f90) before comparing.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELD_DENSITY, FIELD_ENERGY0, FIELD_ENERGY1, FIELD_U, FIELD_P, FIELD_SD = 1, 2, 3, 4, 5, 6

SHAPES = {2: (37, 23, 1), 3: (13, 11, 9)}


def pair(hip, oracle, dim):
    nx, ny, nz = SHAPES[dim]
    g = hip.ExtChunk(hip.lib(dim), dim, nx, ny, nz)
    o = hip.ExtChunk(oracle.lib(dim), dim, nx, ny, nz)
    return g, o


def mesh(ch, dim, with_circle=True):
    """initialise_chunk + generate_chunk + the start-up halo and set_field."""
    n = [ch.nx, ch.ny, ch.nz]
    dx = 10.0 / n[0]
    dy = 10.0 / n[1]
    dz = 10.0 / n[2] if dim == 3 else 0.0
    ch.initialise_chunk(0.0, 0.0, dx, dy, 0.0, dz)
    geo = [0, 1, 2, 3] if with_circle else [0, 1, 1, 1]
    # shrink by a hundredth of a cell as read_input.f90 does for states 2..n
    s = 0.01
    ch.generate_chunk(
        density=[100.0, 0.1, 3.0, 0.5], energy=[1e-4, 25.0, 0.7, 2.0],
        xmin=[0, 0 + s * dx, 6.0, 4 * dx], xmax=[0, 1.0 - s * dx, 0, 0],
        ymin=[0, 0 + s * dy, 6.5, 3 * dy], ymax=[0, 2.0 - s * dy, 0, 0],
        radius=[0, 0, 2.1, 0], geometry=geo,
        zmin=[0, 0, 5.0, 2 * dz] if dim == 3 else None, zmax=[0, 4.0, 0, 0] if dim == 3 else None)
    ch.update_halo([FIELD_DENSITY, FIELD_ENERGY0, FIELD_ENERGY1], 2)
    ch.set_field_kernel()


def interior(ch, a):
    if ch.dim == 3:
        return a[2:-2, 2:-2, 2:-2]
    return a[:, 2:-2, 2:-2]


def assert_same(g, o, names, region="full"):
    for name in names:
        a, b = g.get(name), o.get(name)
        if region == "interior":
            a, b = interior(g, a), interior(o, b)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            k = tuple(bad[0])
            raise AssertionError(f"{name}: {len(bad)} cells differ, first at {k}: gpu {a[k]!r} oracle {b[k]!r}")


def assert_sum(x, y, scale, rtol=1e-12):
    assert abs(x - y) <= rtol * max(abs(scale), 1e-300), (x, y)


def random_fields(g, o, names, rng, lo=0.1, hi=2.0):
    for name in names:
        a = rng.uniform(lo, hi, size=g.shape)
        g.set(name, a)
        o.set(name, a)


@pytest.mark.parametrize("dim", [2, 3])
def test_setup_and_field_summary(hip, oracle, dim):
    g, o = pair(hip, oracle, dim)
    for ch in (g, o):
        mesh(ch, dim)
    assert_same(g, o, ["density", "energy0", "energy1", "volume", "u"])
    sg, so = g.field_summary(), o.field_summary()
    for x, y in zip(sg, so):
        assert_sum(x, y, y)
    g.close()
    o.close()


# one tile; x, y and z ragged; three tiles each way with a 5-column and
# 2-row remainder; the smallest mesh
PAIR_SHAPES = [(13, 11, 9), (70, 9, 37), (125, 26, 40), (2, 2, 2)]


@pytest.mark.parametrize("shape", PAIR_SHAPES)
@pytest.mark.parametrize("pre", [False, True])
def test_ppcg_inner_pair_3d(hip, oracle, shape, pre):
    """The 3D driver's two-step PPCG sweep (tl_ppcg_inner_pair) against two
    ext_ppcg_inner_ calls each (with the depth-1 halo update of sd before
    each, as tea_solve.f90 makes it) on the oracle, for 10 inner steps:
    sd, u and r bitwise on the interior."""
    nx, ny, nz = shape
    g = hip.ExtChunk(hip.lib(3), 3, nx, ny, nz)
    o = hip.ExtChunk(oracle.lib(3), 3, nx, ny, nz)
    cg_prelude(g, o, 3, pre)
    a, b, theta = oracle.calc_ch_coefs(0.9, 40.0, 10)
    for ch in (g, o):
        ch.ppcg_init(a, b, theta)
        ch.calculate_residual()
        ch.update_halo([FIELD_U, FIELD_P], 1)
    pw = o.cg_calc_w()
    g.cg_calc_w()
    alpha = 0.5 / pw
    for ch in (g, o):
        ch.cg_calc_ur(alpha)
        ch.ppcg_init_sd()
    for step in range(1, 11, 2):
        rc = g.L.tl_ppcg_inner_pair(g.c, step)
        assert rc == 0, g.L.tl_last_error()
        for s in (step, step + 1):
            o.update_halo([FIELD_SD], 1)
            o.ppcg_inner(s)
        assert_same(g, o, ["sd", "u", "r"], region="interior")
    g.close()
    o.close()


def test_ppcg_inner_pair_refused_in_2d(hip):
    """The two-step sweep is 3D only: the 2D library refuses it."""
    g = hip.ExtChunk(hip.lib(2), 2, 16, 16)
    a = [0.5] * 4
    g.ppcg_init(a, a, 1.0)
    assert g.L.tl_ppcg_inner_pair(g.c, 1) == -1
    assert b"3D only" in g.L.tl_last_error()
    g.L.tl_clear_error()
    g.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("coef", [1, 2])
@pytest.mark.parametrize("pre", [False, True])
def test_cg_init_and_iteration(hip, oracle, dim, coef, pre):
    g, o = pair(hip, oracle, dim)
    dt = 0.004
    res = []
    for ch in (g, o):
        mesh(ch, dim)
        ch.update_halo([FIELD_ENERGY1, FIELD_DENSITY], 2)
        res.append(ch.cg_solver_init(coef, pre, dt))
        ch.update_halo([FIELD_U, FIELD_P], 1)
    ks = ["kx", "ky"] + (["kz"] if dim == 3 else [])
    assert_same(g, o, ["u", "p", "r", "w"] + ks + (["mi", "z"] if pre else []))
    assert_sum(res[0], res[1], res[1])
    rro = res[1]
    for it in range(3):
        pw = [g.cg_calc_w(), o.cg_calc_w()]
        assert_same(g, o, ["w"])
        assert_sum(pw[0], pw[1], abs(pw[1]))
        alpha = rro / pw[1]
        rrn = [g.cg_calc_ur(alpha), o.cg_calc_ur(alpha)]
        assert_same(g, o, ["u", "r"] + (["z"] if pre else []))
        assert_sum(rrn[0], rrn[1], rrn[1])
        beta = rrn[1] / rro
        g.cg_calc_p(beta)
        o.cg_calc_p(beta)
        assert_same(g, o, ["p"])
        for ch in (g, o):
            ch.update_halo([FIELD_U, FIELD_P], 1)
        rro = rrn[1]
    g.close()
    o.close()


def cg_prelude(g, o, dim, pre=False, iters=4):
    """Run CG init plus a few CG iterations on both sides with the oracle's
    scalars, so later solver tests start from a realistic state."""
    dt = 0.004
    rro = None
    for ch in (g, o):
        mesh(ch, dim)
        ch.update_halo([FIELD_ENERGY1, FIELD_DENSITY], 2)
        r = ch.cg_solver_init(1, pre, dt)
        ch.update_halo([FIELD_U, FIELD_P], 1)
        ch.copy_u()
        rro = r
    for _ in range(iters):
        g.cg_calc_w()
        pw = o.cg_calc_w()
        alpha = rro / pw
        g.cg_calc_ur(alpha)
        rrn = o.cg_calc_ur(alpha)
        g.cg_calc_p(rrn / rro)
        o.cg_calc_p(rrn / rro)
        for ch in (g, o):
            ch.update_halo([FIELD_U, FIELD_P], 1)
        rro = rrn
    assert_same(g, o, ["u", "u0", "p", "r", "w"])
    return rro


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("pre", [False, True])
def test_chebyshev(hip, oracle, dim, pre):
    g, o = pair(hip, oracle, dim)
    cg_prelude(g, o, dim, pre)
    a, b, theta = oracle.calc_ch_coefs(0.9, 40.0, 12)
    for ch in (g, o):
        ch.cheby_solver_init(a, b, theta, pre)
        ch.update_halo([FIELD_U], 1)
    assert_same(g, o, ["u", "p", "r", "w"])
    for step in (12, 2, 3, 4):
        for ch in (g, o):
            ch.cheby_solver_iterate(step)
            ch.update_halo([FIELD_U], 1)
        assert_same(g, o, ["u", "p", "r", "w"])
    n = [g.calculate_2norm(True), o.calculate_2norm(True)]
    assert_sum(n[0], n[1], n[1])
    g.close()
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("pre", [False, True])
def test_ppcg(hip, oracle, dim, pre):
    g, o = pair(hip, oracle, dim)
    cg_prelude(g, o, dim, pre)
    a, b, theta = oracle.calc_ch_coefs(0.9, 40.0, 10)
    for ch in (g, o):
        ch.ppcg_init(a, b, theta)
        ch.calculate_residual()
        ch.update_halo([FIELD_U, FIELD_P], 1)
    assert_same(g, o, ["r"])
    pw = [g.cg_calc_w(), o.cg_calc_w()]
    alpha = 0.5 / pw[1]
    g.cg_calc_ur(alpha)
    o.cg_calc_ur(alpha)
    for ch in (g, o):
        ch.ppcg_init_sd()
    assert_same(g, o, ["sd", "u", "r"])
    for step in range(1, 11):
        for ch in (g, o):
            ch.update_halo([FIELD_SD], 1)
            ch.ppcg_inner(step)
        assert_same(g, o, ["sd", "u", "r"])
    n = [g.calculate_2norm(True), o.calculate_2norm(True)]
    assert_sum(n[0], n[1], n[1])
    g.close()
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("coef", [1, 2])
def test_jacobi(hip, oracle, dim, coef):
    g, o = pair(hip, oracle, dim)
    for ch in (g, o):
        mesh(ch, dim)
        ch.update_halo([FIELD_ENERGY1, FIELD_DENSITY], 2)
        ch.jacobi_init(coef, 0.004)
        ch.copy_u()
    ks = ["kx", "ky"] + (["kz"] if dim == 3 else [])
    assert_same(g, o, ["u", "u0"] + ks)
    for _ in range(4):
        e = [g.jacobi_solve(), o.jacobi_solve()]
        assert_same(g, o, ["u", "r"], region="interior")
        assert_sum(e[0], e[1], e[1])
        for ch in (g, o):
            ch.update_halo([FIELD_U], 1)
    for ch in (g, o):
        ch.calculate_residual()
    assert_same(g, o, ["r"], region="interior")
    g.close()
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_residual_norm_finalise(hip, oracle, dim):
    g, o = pair(hip, oracle, dim)
    cg_prelude(g, o, dim)
    for ch in (g, o):
        ch.calculate_residual()
    assert_same(g, o, ["r"])
    for flag in (True, False):
        n = [g.calculate_2norm(flag), o.calculate_2norm(flag)]
        assert_sum(n[0], n[1], n[1])
    for ch in (g, o):
        ch.finalise()
    assert_same(g, o, ["energy1"])
    g.close()
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("depth", [1, 2])
def test_update_halo_random(hip, oracle, dim, depth):
    """Random values everywhere (halo cells included), then one halo update
    of several fields: every cell of every field must match, which pins the
    order-dependent corner values of the 2D update and the face-only 3D one."""
    rng = np.random.default_rng(7 + dim + depth)
    g, o = pair(hip, oracle, dim)
    names = ["density", "energy0", "energy1", "u", "p", "sd"]
    random_fields(g, o, names, rng)
    for ch in (g, o):
        ch.update_halo([FIELD_DENSITY, FIELD_P, FIELD_ENERGY1, FIELD_SD], depth)
    assert_same(g, o, names)
    # a chunk with only some external faces (the others are left untouched)
    nf = 6 if dim == 3 else 4
    nb = [-1 if k % 2 == 0 else 7 for k in range(nf)]
    random_fields(g, o, ["u"], rng)
    for ch in (g, o):
        ch.update_halo([FIELD_U], depth, neighbours=nb)
    assert_same(g, o, ["u"])
    g.close()
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_random_stencil_inputs(hip, oracle, dim):
    """cg_calc_w on random p and random positive kx/ky[/kz]: the SMVP
    arithmetic itself, away from any physical state."""
    rng = np.random.default_rng(11 + dim)
    g, o = pair(hip, oracle, dim)
    ks = ["kx", "ky"] + (["kz"] if dim == 3 else [])
    random_fields(g, o, ["p"] + ks, rng, lo=-1.0, hi=3.0)
    for ch in (g, o):
        ch.update_halo([FIELD_P], 1)
    pw = [g.cg_calc_w(), o.cg_calc_w()]
    assert_same(g, o, ["w"], region="interior")
    w = o.get("w")
    assert_sum(pw[0], pw[1], float(np.sum(np.abs(interior(o, w) * interior(o, o.get("p"))))))
    g.close()
    o.close()


@pytest.mark.parametrize("shape,nsteps", [((37, 23), 10), ((150, 140), 10), ((150, 140), 3), ((64, 200), 1),
                                          ((130, 70), 12), ((90, 100), 15), ((45, 66), 25),
                                          # meshes narrower than the levels' reach: reflections repeat
                                          ((5, 4), 10), ((9, 3), 12), ((2, 17), 5), ((131, 6), 10)])
def test_ppcg_outer_fused(hip, oracle, shape, nsteps):
    """The driver's temporal-blocked PPCG outer iteration (calc_ur, init_sd,
    all inner steps and ||r||^2 in one sweep) after CGCalcW against the
    reference's separate sweeps on the oracle: u and r bitwise, rrn to
    1e-12."""
    import ctypes
    nx, ny = shape
    g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
    o = hip.ExtChunk(oracle.lib(2), 2, nx, ny)
    rng = np.random.default_rng(nx * 1000 + ny + nsteps)
    random_fields(g, o, ["u", "p", "r"], rng, lo=-1.0, hi=1.0)
    random_fields(g, o, ["kx", "ky"], rng, lo=0.1, hi=40.0)
    # the state after CGCalcW (tea_solve.f90:379-383): w = A p from p with
    # its reflective halo
    o.update_halo([FIELD_P], 1)
    g.cg_calc_w()
    o.cg_calc_w()
    a, b, theta = oracle.calc_ch_coefs(0.7, 150.0, nsteps)
    alpha = 0.37
    ca = (ctypes.c_double * nsteps)(*a)
    cb = (ctypes.c_double * nsteps)(*b)
    rrn_g = ctypes.c_double()
    rc = g.L.tl_ppcg_outer_fused(g.c, ctypes.c_double(alpha), ctypes.c_double(theta), ca, cb, nsteps,
                                 ctypes.byref(rrn_g))
    assert rc == 0, g.L.tl_last_error()
    o.ppcg_init(a, b, theta)
    o.cg_calc_ur(alpha)
    o.ppcg_init_sd()
    for step in range(1, nsteps + 1):
        o.update_halo([FIELD_SD], 1)
        o.ppcg_inner(step)
    rrn_o = o.calculate_2norm(True)
    assert_same(g, o, ["u", "r"], region="interior")
    assert_sum(rrn_g.value, rrn_o, rrn_o)
    g.close()
    o.close()


# meshes narrower than the passes' reach, odd and even widths, and iteration
# counts that need passes of 10, 5, 2 and 1
FUSED_SHAPES = [((37, 23), 10), ((5, 4), 13), ((9, 7), 17), ((131, 6), 10), ((67, 33), 3), ((2, 17), 6),
                ((130, 41), 8)]


@pytest.mark.parametrize("shape,count", FUSED_SHAPES)
def test_cheby_iterate_fused(hip, oracle, shape, count):
    """The driver's temporal-blocked Chebyshev passes against the reference's
    one-iteration ChebyIterate + ChebyCalcU calls (with the depth-1 halo
    update of u before each, as tea_solve.f90 makes it) on the oracle: u, p
    and r bitwise, ||r||^2 to 1e-12."""
    import ctypes
    nx, ny = shape
    g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
    o = hip.ExtChunk(oracle.lib(2), 2, nx, ny)
    rng = np.random.default_rng(nx * 1000 + ny + count)
    random_fields(g, o, ["u", "u0", "p"], rng, lo=-1.0, hi=1.0)
    random_fields(g, o, ["kx", "ky"], rng, lo=0.1, hi=40.0)
    step1 = 2
    a, b, theta = oracle.calc_ch_coefs(0.7, 150.0, step1 + count + 1)
    for ch in (g, o):
        ch.update_halo([FIELD_U], 1)
        ch.cheby_solver_init(a, b, theta, False)
        ch.update_halo([FIELD_U], 1)
    rrn_g = ctypes.c_double()
    rc = g.L.tl_cheby_iterate_fused(g.c, ctypes.c_int(step1), ctypes.c_int(count), ctypes.byref(rrn_g))
    assert rc == 0, g.L.tl_last_error()
    for step in range(step1, step1 + count):
        o.cheby_solver_iterate(step)
        o.update_halo([FIELD_U], 1)
    rrn_o = o.calculate_2norm(True)
    assert_same(g, o, ["u", "p", "r"], region="interior")
    assert_sum(rrn_g.value, rrn_o, rrn_o)
    g.close()
    o.close()


@pytest.mark.parametrize("shape,count", FUSED_SHAPES)
def test_jacobi_iterate_fused(hip, oracle, shape, count):
    """The driver's temporal-blocked Jacobi passes against the reference's
    JacobiCopyU + JacobiSolve calls (depth-1 halo update of u before each):
    u and r bitwise, each iteration's error sum to 1e-12."""
    import ctypes
    nx, ny = shape
    g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
    o = hip.ExtChunk(oracle.lib(2), 2, nx, ny)
    rng = np.random.default_rng(nx * 1000 + ny + 7 * count)
    random_fields(g, o, ["u", "u0"], rng, lo=0.1, hi=2.0)
    random_fields(g, o, ["kx", "ky"], rng, lo=0.1, hi=40.0)
    errs = (ctypes.c_double * count)()
    rc = g.L.tl_jacobi_iterate_fused(g.c, ctypes.c_int(count), errs)
    assert rc == 0, g.L.tl_last_error()
    for k in range(count):
        o.update_halo([FIELD_U], 1)
        e = o.jacobi_solve()
        assert_sum(errs[k], e, e)
    assert_same(g, o, ["u", "r"], region="interior")
    g.close()
    o.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("depth", [1, 2])
def test_pack_message(hip, oracle, dim, depth):
    """ext_pack_message_: every face's message packed from and unpacked into
    the flagged fields, against the oracle's restatement of the reference's
    packers: buffers and fields bitwise."""
    from test_oracle import pack_message
    g, o = pair(hip, oracle, dim)
    rng = np.random.default_rng(100 * dim + depth)
    names = ["density", "energy0", "energy1", "u", "p", "sd"]
    for face in range(1, 2 * dim + 1):
        fields = [int(v) for v in rng.integers(0, 2, size=6)]
        fields[face % 6] = 1
        random_fields(g, o, names, rng, lo=-1.0, hi=1.0)
        n = {1: g.ny * g.nz, 2: g.ny * g.nz, 3: g.nx * g.nz, 4: g.nx * g.nz, 5: g.nx * g.ny,
             6: g.nx * g.ny}[face] * depth
        offsets = [q * n + 3 for q in range(6)]
        bufs = [np.full(6 * n + 3, -7.0) for _ in range(2)]
        for ch, b in zip((g, o), bufs):
            pack_message(ch, fields, offsets, depth, face, b, True)
        np.testing.assert_array_equal(bufs[0], bufs[1])
        msg = rng.uniform(-1, 1, size=6 * n + 3)
        for ch in (g, o):
            pack_message(ch, fields, offsets, depth, face, msg.copy(), False)
        assert_same(g, o, names)
    g.close()
    o.close()


def test_stencil_refused_with_neighbour_chunks(hip, oracle):
    """The stencil kernels read reflective boundaries: once a chunk has a
    neighbour chunk (a halo update naming one), stencil calls are refused
    with an error instead of computing with the wrong boundary."""
    g = hip.ExtChunk(hip.lib(2), 2, 37, 23)
    mesh(g, 2)
    g.update_halo([FIELD_ENERGY1, FIELD_DENSITY], 2)
    g.cg_solver_init(1, False, 0.004)
    g.update_halo([FIELD_P], 1, neighbours=[-1, 2, -1, -1])
    with pytest.raises(RuntimeError, match="neighbour chunks"):
        g.cg_calc_w()
    g.close()


def jacobi_sweeps_numpy(r, z, kx, ky, omega, sweeps):
    """Weighted Jacobi sweeps z <- z + omega (r - A z) / D on the mesh cells
    (omega a number, or one weight per sweep),
    A = I + sum over interior faces of k (e_i - e_j)(e_i - e_j)^T (a face on
    the mesh boundary carries no flux) and D = diag(A); arrays in the
    reference's flat layout, (1, ny + 4, nx + 4), kx the west and ky the south
    face of each cell.  Written from the operator, independently of the
    kernels."""
    R = r[0, 2:-2, 2:-2]
    Z = z[0, 2:-2, 2:-2].copy()
    kw, ke = kx[0, 2:-2, 2:-2].copy(), kx[0, 2:-2, 3:-1].copy()
    ks, kn = ky[0, 2:-2, 2:-2].copy(), ky[0, 3:-1, 2:-2].copy()
    kw[:, 0] = 0.0
    ke[:, -1] = 0.0
    ks[0, :] = 0.0
    kn[-1, :] = 0.0
    D = 1.0 + kw + ke + ks + kn
    ws = np.broadcast_to(np.asarray(omega, dtype=float), (sweeps,))
    for s in range(sweeps):
        P = np.pad(Z, 1, mode="edge")
        AZ = Z + kw * (Z - P[1:-1, :-2]) + ke * (Z - P[1:-1, 2:]) + ks * (Z - P[:-2, 1:-1]) + kn * (Z - P[2:, 1:-1])
        Z = Z + ws[s] * (R - AZ) / D
    return Z


@pytest.mark.parametrize("shape", [(37, 23), (5, 4), (9, 7), (131, 6), (130, 41), (2, 17), (64, 64)])
@pytest.mark.parametrize("from_zero", [True, False])
def test_fcg_smooth_fused(hip, shape, from_zero):
    """The flexible CG's four weighted Jacobi sweeps as one temporal-blocked
    pass against one launch per sweep: z equal on every mesh cell (the pass
    adds + 0 for a boundary face where the one-sweep kernel skips the term,
    which can only change the sign of a zero)."""
    import ctypes
    nx, ny = shape
    rng = np.random.default_rng(nx * 131 + ny + from_zero)
    vals = {"r": rng.uniform(-1, 1, size=(1, ny + 4, nx + 4)), "z": rng.uniform(-1, 1, size=(1, ny + 4, nx + 4)),
            "kx": rng.uniform(0.1, 40.0, size=(1, ny + 4, nx + 4)), "ky": rng.uniform(0.1, 40.0, size=(1, ny + 4, nx + 4))}
    out = []
    for fused in (0, 1):
        g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
        for k, v in vals.items():
            g.set(k, v)
        rc = g.L.tl_fcg_smooth(g.c, ctypes.c_int(4), ctypes.c_double(0.8), ctypes.c_int(int(from_zero)),
                               ctypes.c_int(fused))
        assert rc == 0, g.L.tl_last_error()
        out.append(interior(g, g.get("z")))
        g.close()
    np.testing.assert_array_equal(out[0], out[1])
    z0 = np.zeros_like(vals["z"]) if from_zero else vals["z"]
    want = jacobi_sweeps_numpy(vals["r"], z0, vals["kx"], vals["ky"], 0.8, 4)
    got = out[1].reshape(want.shape)
    assert np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("shape", [(37, 23), (5, 4), (130, 41), (64, 64)])
@pytest.mark.parametrize("from_zero", [True, False])
@pytest.mark.parametrize("m,ratio", [(4, 30.0), (8, 300.0), (12, 1000.0), (16, 3000.0), (32, 13000.0)])
def test_fcg_smooth_chebyshev_weights(hip, shape, from_zero, m, ratio):
    """The sweeps with a weight per sweep (the Chebyshev step weights of
    [2 / ratio, 2] from tl_fcg_sweep_weights, up to 182 for 16 sweeps): the
    temporal-blocked passes (m / 4 of them, each from the previous one's z)
    equal to one launch per sweep on every mesh cell, and both against
    numpy's sweeps with those weights."""
    import ctypes
    nx, ny = shape
    w = (ctypes.c_double * m)()
    L = hip.lib(2)
    assert L.tl_fcg_sweep_weights(ctypes.c_int(m), ctypes.c_double(0.8), ctypes.c_double(ratio), w) == 0
    assert max(w[:]) > 7.0
    rng = np.random.default_rng(nx * 53 + ny + from_zero)
    full = (1, ny + 4, nx + 4)
    vals = {"r": rng.uniform(-1, 1, size=full), "z": rng.uniform(-1, 1, size=full),
            "kx": rng.uniform(0.1, 40.0, size=full), "ky": rng.uniform(0.1, 40.0, size=full)}
    print(f"fcg chebyshev sweeps {shape} from_zero={from_zero}", flush=True)
    out = []
    for fused in (0, 1):
        g = hip.ExtChunk(L, 2, nx, ny)
        for k, v in vals.items():
            g.set(k, v)
        rc = g.L.tl_fcg_smooth_weights(g.c, ctypes.c_int(m), w, ctypes.c_int(int(from_zero)), ctypes.c_int(fused))
        assert rc == 0, g.L.tl_last_error()
        out.append(interior(g, g.get("z")))
        g.close()
    np.testing.assert_array_equal(out[0], out[1])
    z0 = np.zeros_like(vals["z"]) if from_zero else vals["z"]
    want = jacobi_sweeps_numpy(vals["r"], z0, vals["kx"], vals["ky"], list(w), m)
    # large weights amplify the rounding of intermediate sweeps before the
    # polynomial brings them back down
    tol = 1e-12 if m == 4 else 1e-9 if m < 12 else 1e-7
    assert np.max(np.abs(out[1].reshape(want.shape) - want)) <= tol * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("shape", [(37, 23), (5, 4), (131, 6), (130, 41), (2, 17), (64, 64)])
def test_fcg_smooth_update(hip, shape):
    """The preconditioner pass with the previous step's r update folded in:
    r - alpha w bitwise on every mesh cell, z equal to one launch per sweep
    on that r (the fused pass adds + 0 where the one-sweep kernel skips a
    boundary term), and z.w, ||r||^2, z.r against numpy over the mesh."""
    import ctypes
    nx, ny = shape
    rng = np.random.default_rng(nx * 977 + ny)
    full = (1, ny + 4, nx + 4)
    r, w = rng.uniform(-1, 1, size=full), rng.uniform(-1, 1, size=full)
    kx, ky = rng.uniform(0.1, 40.0, size=full), rng.uniform(0.1, 40.0, size=full)
    alpha, omega = 0.37, 0.8
    print(f"fcg smooth update {shape}", flush=True)
    g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
    for k, v in (("r", r), ("w", w), ("kx", kx), ("ky", ky)):
        g.set(k, v)
    sums = (ctypes.c_double * 3)()
    rc = g.L.tl_fcg_smooth_update(g.c, ctypes.c_double(omega), ctypes.c_double(alpha), sums)
    assert rc == 0, g.L.tl_last_error()
    r_got, z_got = interior(g, g.get("r")), interior(g, g.get("z"))
    g.close()
    r_new = r - alpha * w
    np.testing.assert_array_equal(r_got, interior(g, r_new))

    g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
    for k, v in (("r", r_new), ("kx", kx), ("ky", ky)):
        g.set(k, v)
    rc = g.L.tl_fcg_smooth(g.c, ctypes.c_int(4), ctypes.c_double(omega), ctypes.c_int(1), ctypes.c_int(0))
    assert rc == 0, g.L.tl_last_error()
    z_ref = interior(g, g.get("z"))
    g.close()
    np.testing.assert_array_equal(z_got, z_ref)
    want = jacobi_sweeps_numpy(r_new, np.zeros_like(r_new), kx, ky, omega, 4)
    assert np.max(np.abs(z_got.reshape(want.shape) - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))

    wi, ri = interior(g, w), interior(g, r_new)
    want = [float(np.sum(z_ref * wi)), float(np.sum(ri * ri)), float(np.sum(z_ref * ri))]
    for k in range(3):
        assert abs(sums[k] - want[k]) <= 1e-12 * max(1.0, float(np.sum(np.abs(z_ref * wi))), float(np.sum(ri * ri))), \
            (k, sums[k], want[k])


def test_fcg_smooth_refused_in_3d(hip):
    """The flexible CG's sweeps are 2D only: the 3D library refuses them."""
    import ctypes
    g = hip.ExtChunk(hip.lib(3), 3, 6, 5, 4)
    rc = g.L.tl_fcg_smooth(g.c, ctypes.c_int(4), ctypes.c_double(0.8), ctypes.c_int(1), ctypes.c_int(0))
    assert rc == -1
    assert b"2D only" in g.L.tl_last_error()
    g.L.tl_clear_error()
    g.close()


def reduction_order(x):
    """The grid reduction's summation order: thread t of a 256-lane
    workgroup adds partials t, t + 256, ... in turn; each 64-lane wave folds
    with the xor butterfly (offsets 32 .. 1); the four waves' sums are added
    in wave order."""
    lanes = np.zeros(256)
    for t in range(256):
        s = 0.0
        for v in x[t::256]:
            s = s + v
        lanes[t] = s
    idx = np.arange(64)
    waves = []
    for w in range(4):
        v = lanes[64 * w:64 * (w + 1)].copy()
        for off in (32, 16, 8, 4, 2, 1):
            v = v + v[idx ^ off]
        waves.append(v[0])
    s = waves[0]
    for w in waves[1:]:
        s = s + w
    return s


@pytest.mark.parametrize("nwaves", [1, 4])
@pytest.mark.parametrize("nblocks", [1, 3, 64, 255, 256, 257, 1000, 4097, 70000])
def test_single_pass_reduction(hip, nwaves, nblocks):
    """The last workgroup's sum (single pass, counter reset for the next
    reduction) equals the separate k_finish launch and the stated order
    bitwise, for workgroups of one wave (temporal-blocked passes) and four
    (streaming sweeps), on values spread over 16 decades so that any other
    order would round differently."""
    rng = np.random.default_rng(nblocks * 7 + nwaves)
    x = rng.standard_normal(nblocks) * 10.0 ** rng.uniform(-8, 8, nblocks)
    out = np.zeros(3)
    L = hip.lib(2)
    ptr = ctypes.POINTER(ctypes.c_double)
    rc = L.tl_reduce_probe(nblocks, nwaves, x.ctypes.data_as(ptr), out.ctypes.data_as(ptr))
    assert rc == 0, L.tl_last_error().decode()
    want = reduction_order(x)
    assert out[0] == want and out[1] == want and out[2] == want, (out, want)
    assert abs(want - np.sum(np.sort(x))) <= 1e-12 * np.sum(np.abs(x))
