"""Pins of the oracle's Chebyshev, PPCG and Jacobi driver branches (no GPU).

The oracle's restated driver decides iteration counts, so its branch logic
is checked here against independent numpy models of the reference's solve
loop (tealeaf_2d/tea_solve.f90:169-503, 607-750) on tiny dense systems.  The
models share no code with the oracle and are written in other forms:

  - Chebyshev in the three-term "omega" form
        e_{k+1} = w_{k+1} (I - A/theta) e_k + (1 - w_{k+1}) e_{k-1},
        w_{k+1} = 2 sigma T_k(sigma) / T_{k+1}(sigma)
    with T_k(sigma) = cosh(k arccosh sigma) in closed form (the reference's
    alpha_k = w_{k+1} - 1 and beta_k = w_{k+1}/theta);
  - PPCG's L inner steps as the residual polynomial they apply,
        r <- R_L(A) r,  u <- u + A^-1 (I - R_L(A)) r,
        R_L(lambda) = T_L((theta - lambda)/delta) / T_L(sigma),
    evaluated on the eigendecomposition of A;
  - Jacobi as the matrix splitting A = D' - N' the reference's kernel
    iterates (D' includes the boundary faces' coefficients, whose mirrored
    neighbour is the cell itself, ext_jacobi_kernel.hpp:85-117);
  - preconditioned Chebyshev in explicit p form with the closed-form
    coefficients, including the CG iterations the switch test lets run
    between Chebyshev steps and the stale halo of p the first of them
    reads.

Each test runs the oracle's whole driver on a deck (capped or to
convergence), reads its final fields, and compares u, the iteration count,
the switch step and est_itc with the model.  Each also runs the model with
one plausible mistake planted (the first Chebyshev step's coefficient,
the inner-step coefficient index, the norm-check cadence, the Jacobi
residual replacement) and requires that the oracle disagrees with it, so
the comparison is shown to detect that class of mistake.
"""
import ctypes
import math

import numpy as np
import pytest

EPS64 = 2.0 ** -52
F95, F105 = float(np.float32(0.95)), float(np.float32(1.05))  # default-real literals, tea_solve.f90:194-195


def kept_run(oracle, deck, parts=4):
    """Run a 1-step deck on the oracle and return (report, fields) with the
    final fields of its chunk, in the reference's flat layout."""
    L = oracle.lib(2)
    L.tl_oracle_keep_chunk.argtypes = [ctypes.c_int]
    L.tl_oracle_keep_chunk(1)
    try:
        rep = oracle.run_deck(deck, dim=2, reduce_parts=parts, threads=4)
        chunk = ctypes.c_int(1)
        dims = (ctypes.c_int * 6)()
        assert L.tl_chunk_dims(ctypes.byref(chunk), dims) == 0
        shape = (dims[1], dims[0])
        out = {}
        for name in ("u", "u0", "kx", "ky"):
            a = np.empty(shape)
            assert L.tl_get_field(ctypes.byref(chunk), oracle.FIELD_ID[name],
                                  a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
            out[name] = a
    finally:
        L.tl_oracle_keep_chunk(0)
        L.ext_finalise_()
    return rep, out


class System:
    """A u = b on the interior cells, from the chunk's face coefficients:
    A = I + sum over interior faces of k (e_i - e_j)(e_i - e_j)^T."""

    def __init__(self, f):
        kx, ky = f["kx"], f["ky"]
        Y, X = kx.shape
        ny, nx = Y - 4, X - 4
        n = nx * ny
        idx = lambda x, y: (y - 2) * nx + (x - 2)
        A = np.eye(n)
        bsum = np.zeros(n)  # boundary-face coefficients of each cell
        for y in range(2, ny + 2):
            for x in range(2, nx + 2):
                i = idx(x, y)
                for (xn, yn, k) in ((x - 1, y, kx[y, x]), (x, y - 1, ky[y, x])):
                    if 2 <= xn < nx + 2 and 2 <= yn < ny + 2:
                        j = idx(xn, yn)
                        A[i, i] += k
                        A[j, j] += k
                        A[i, j] -= k
                        A[j, i] -= k
                    else:
                        bsum[i] += k
                if x == nx + 1:
                    bsum[i] += kx[y, x + 1]
                if y == ny + 1:
                    bsum[i] += ky[y + 1, x]
        self.A = A
        self.b = f["u0"][2:-2, 2:-2].ravel().copy()
        self.bsum = bsum
        self.lam, self.V = np.linalg.eigh(A)
        self.shape = (ny, nx)


def interior(f, name):
    return f[name][2:-2, 2:-2].ravel()


def lanczos_eigs(alphas, betas):
    """Extreme eigenvalues of the CG Lanczos tridiagonal
    (tea_leaf_cheby.f90:298-302), by LAPACK."""
    m = len(alphas)
    T = np.zeros((m, m))
    for k in range(m):
        T[k, k] = 1.0 / alphas[k] + (betas[k - 1] / alphas[k - 1] if k else 0.0)
        if k + 1 < m:
            T[k, k + 1] = T[k + 1, k] = math.sqrt(betas[k]) / alphas[k]
    ev = np.linalg.eigvalsh(T)
    return ev[0], ev[-1]


def omega(k, sigma):
    """w_{k+1} = 2 sigma T_k(sigma) / T_{k+1}(sigma), sigma > 1, in the
    overflow-free closed form of cosh ratios."""
    phi = math.acosh(sigma)
    return 2.0 * sigma * math.exp(-phi) * (1.0 + math.exp(-2.0 * k * phi)) / (1.0 + math.exp(-2.0 * (k + 1) * phi))


def cheb_T(L, x):
    c = np.zeros(L + 1)
    c[L] = 1.0
    return np.polynomial.chebyshev.chebval(x, c)


def model(sys_, solver, presteps, max_iters, eps, inner=10, first_coef="last", inner_shift=0, cadence=10):
    """The reference's solve loop for one timestep on the dense system.
    Returns u and the loop's bookkeeping.  The keyword arguments plant
    mistakes: first_coef="first" uses coefficient 1 for the first Chebyshev
    step, inner_shift moves PPCG's coefficient index, cadence changes the
    Chebyshev norm-check period."""
    A, b = sys_.A, sys_.b
    u = b.copy()  # cg_solver_init: u = energy * density, u0 = u
    r = b - A @ u
    p = r.copy()
    rro = r @ r
    error = 1e10
    cg_a, cg_b = [], []
    steps = 0
    out = {"switch_step": 0, "est_itc": 0}
    n_done = max_iters
    for n in range(1, max_iters + 1):
        switch = n >= presteps and error <= 1.0
        if solver in ("chebyshev", "ppcg") and switch:
            if steps == 0:
                M = max_iters - n + 2
                rro = error
                lo, hi = lanczos_eigs(cg_a, cg_b)
                eigmin, eigmax = lo * F95, hi * F105
                theta, delta = (eigmax + eigmin) / 2, (eigmax - eigmin) / 2
                sigma = theta / delta
                cn = eigmax / eigmin
                out.update(switch_step=n, eigmin=eigmin, eigmax=eigmax)
            if solver == "chebyshev":
                if steps == 0:
                    bb = b @ b
                    u_prev = u.copy()
                    u = u + (b - A @ u) / theta
                    r = b - A @ u
                    w = omega(M if first_coef == "last" else 1, sigma)
                    u_prev, u = u, w * (u + r / theta) + (1.0 - w) * u_prev
                    error = r @ r
                    gamm = (math.sqrt(cn) - 1.0) / (math.sqrt(cn) + 1.0)
                    est = math.log(EPS64 * bb / (4.0 * error)) / (2.0 * math.log(gamm))
                    out["est_itc"] = int(math.floor(est + 0.5))
                    steps = 1
                else:
                    r = b - A @ u
                    w = omega(steps, sigma)
                    u_prev, u = u, w * (u + r / theta) + (1.0 - w) * u_prev
                    if n >= out["est_itc"] and n % cadence == 0:
                        error = r @ r
            else:
                if steps == 0:
                    steps = 1
                    r = b - A @ u  # the true residual replaces the recurrence one
                w = A @ p
                alpha = rro / (p @ w)
                u = u + alpha * p
                r = r - alpha * w
                c = sys_.V.T @ r
                x = (theta - sys_.lam) / delta
                if inner_shift == 0:
                    R = cheb_T(inner, x) / cheb_T(inner, sigma)
                else:
                    R = cheb_T(inner + inner_shift, x) / cheb_T(inner + inner_shift, sigma)
                u = u + sys_.V @ ((1.0 - R) / sys_.lam * c)
                r = sys_.V @ (R * c)
                rrn = r @ r
                p = r + (rrn / rro) * p
                error = rro = rrn
            steps += 1
        elif solver in ("cg", "chebyshev", "ppcg"):
            w = A @ p
            alpha = rro / (p @ w)
            u = u + alpha * p
            r = r - alpha * w
            rrn = r @ r
            beta = rrn / rro
            cg_a.append(alpha)
            cg_b.append(beta)
            p = r + beta * p
            error = rro = rrn
        if abs(error) < eps:
            n_done = n - 1
            break
    out["iters"] = n_done
    out["u"] = u
    return out


def jacobi_model(sys_, max_iters, eps, period=50):
    """Jacobi as the reference's kernel iterates it: u' = D'^-1 (b + N' u)
    with D' = diag(A) + boundary-face coefficients and N' = D' - A; the
    error is sum |u' - u|, replaced by ||b - A u'||^2 every `period`
    iterations (tea_solve.f90:437-485)."""
    A, b = sys_.A, sys_.b
    D = np.diag(A) + sys_.bsum
    N = np.diag(D) - A
    u = b.copy()
    hist = []
    n_done = max_iters
    for n in range(1, max_iters + 1):
        un = (b + N @ u) / D
        error = np.abs(un - u).sum()
        u = un
        l1 = error
        rr = None
        if period and n % period == 0:
            rr = (b - A @ u) @ (b - A @ u)
            error = rr
        hist.append((l1, rr))
        if abs(error) < eps:
            n_done = n - 1
            break
    return {"u": u, "iters": n_done, "hist": hist}


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def deck(solver, max_iters, eps, presteps=12, inner=10, dt=2.0, n=14, ny=11):
    import decks
    return decks.two_rect_2d(n, solver, steps=1, eps=eps, dt=dt, max_iters=max_iters, presteps=presteps,
                             inner=inner, ny=ny)


TOL = 1e-10   # model and oracle differ only by rounding
MISS = 1e-6   # a planted mistake moves u by more than this


def test_system_model_is_the_oracle_operator(oracle):
    """The dense model applies the same operator as the oracle: one CG
    iteration from u = u0 gives the same u."""
    rep, f = kept_run(oracle, deck("cg", 1, 0.0))
    s = System(f)
    m = model(s, "cg", 12, 1, 0.0)
    assert rel(interior(f, "u"), m["u"]) < 1e-13
    assert s.lam[0] >= 1.0 - 1e-12  # A = I + positive semidefinite flux terms


@pytest.mark.parametrize("extra", [0, 1, 2, 5, 13])
def test_chebyshev_trajectory(oracle, extra):
    """Capped Chebyshev solves: u after the switch step (whose second update
    uses the last coefficient, tea_solve.f90:707) and `extra` further steps
    equals the omega-form model; the first step with coefficient 1 does not."""
    full, _ = kept_run(oracle, deck("chebyshev", 10000, 1e-30))
    ns = full["steps"][0]["switch_step"]
    assert ns >= 12
    cap = ns + extra
    rep, f = kept_run(oracle, deck("chebyshev", cap, 0.0))
    s = System(f)
    m = model(s, "chebyshev", 12, cap, 0.0)
    st = rep["steps"][0]
    assert st["switch_step"] == m["switch_step"] == ns
    assert st["iters"] == m["iters"] == cap
    assert st["eigmin"] == pytest.approx(m["eigmin"], rel=1e-9)
    assert st["eigmax"] == pytest.approx(m["eigmax"], rel=1e-9)
    assert rel(interior(f, "u"), m["u"]) < TOL
    wrong = model(s, "chebyshev", 12, cap, 0.0, first_coef="first")
    assert rel(interior(f, "u"), wrong["u"]) > MISS


def test_chebyshev_cadence_and_est_itc(oracle):
    """A converged Chebyshev solve: est_itc is the reference's estimate from
    ||u0||^2 and the first step's residual, norms are checked only at
    iterations n >= est_itc with n % 10 == 0, so the loop ends at such an n
    (reported as n - 1), and no earlier check point met eps."""
    eps = 1e-16
    rep, f = kept_run(oracle, deck("chebyshev", 10000, eps))
    s = System(f)
    m = model(s, "chebyshev", 12, 10000, eps)
    st = rep["steps"][0]
    assert st["est_itc"] == m["est_itc"] > st["switch_step"]
    n_end = st["iters"] + 1
    assert n_end % 10 == 0 and n_end >= st["est_itc"]
    assert st["iters"] == m["iters"]
    assert rel(interior(f, "u"), m["u"]) < TOL
    for cadence in (1, 9, 11):
        assert model(s, "chebyshev", 12, 10000, eps, cadence=cadence)["iters"] != st["iters"]


@pytest.mark.parametrize("inner", [1, 4, 10])
@pytest.mark.parametrize("extra", [0, 1, 3])
def test_ppcg_trajectory(oracle, inner, extra):
    """Capped PPCG solves: at the switch r is recomputed and p kept, each
    outer iteration applies the degree-L Chebyshev residual polynomial
    (coefficients 1..L, tea_solve.f90:629-637), and rrn is ||r||^2.  The
    cap keeps at least L coefficients stored (max_iters - n + 2 >= L; the
    reference reads past its coefficient array otherwise)."""
    full, _ = kept_run(oracle, deck("ppcg", 10000, 1e-30, inner=inner))
    ns = full["steps"][0]["switch_step"]
    cap = ns + max(0, inner - 2) + extra
    rep, f = kept_run(oracle, deck("ppcg", cap, 0.0, inner=inner))
    s = System(f)
    m = model(s, "ppcg", 12, cap, 0.0, inner=inner)
    st = rep["steps"][0]
    assert st["switch_step"] == m["switch_step"] == ns
    assert rel(interior(f, "u"), m["u"]) < TOL
    wrong = model(s, "ppcg", 12, cap, 0.0, inner=inner, inner_shift=1)
    assert rel(interior(f, "u"), wrong["u"]) > MISS


def test_ppcg_converged(oracle):
    eps = 1e-20
    rep, f = kept_run(oracle, deck("ppcg", 10000, eps, inner=6))
    s = System(f)
    m = model(s, "ppcg", 12, 10000, eps, inner=6)
    assert rep["steps"][0]["iters"] == m["iters"]
    assert rel(interior(f, "u"), m["u"]) < TOL


def test_jacobi_trajectory(oracle):
    """130 capped Jacobi iterations (two residual replacements on the way)
    give the splitting model's u."""
    rep, f = kept_run(oracle, deck("jacobi", 130, 0.0, dt=0.5))
    s = System(f)
    m = jacobi_model(s, 130, 0.0)
    assert rep["steps"][0]["iters"] == 130
    assert rel(interior(f, "u"), m["u"]) < TOL


def test_jacobi_residual_replacement_every_50(oracle):
    """An eps that the L1 error does not meet in iterations 1-49 but the true
    ||r||^2 meets at iteration 50 stops the loop at n = 50 exactly (reported
    49): the error is replaced by ||r||^2 at n % 50 == 0 only."""
    _, f = kept_run(oracle, deck("jacobi", 1, 0.0, dt=0.1))
    s = System(f)
    hist = jacobi_model(s, 100, 0.0)["hist"]
    l1_min = min(h[0] for h in hist[:49])
    rr50 = hist[49][1]
    assert rr50 is not None and rr50 < 1e-3 * l1_min, (rr50, l1_min)
    eps = math.sqrt(rr50 * l1_min)
    rep, f2 = kept_run(oracle, deck("jacobi", 1000, eps, dt=0.1))
    m = jacobi_model(s, 1000, eps)
    assert m["iters"] == 49
    assert rep["steps"][0]["iters"] == 49
    assert rel(interior(f2, "u"), m["u"]) < TOL
    for period in (49, 51, 0):
        assert jacobi_model(s, 1000, eps, period=period)["iters"] != 49


def test_jacobi_l1_stop(oracle):
    """An eps between two consecutive L1 errors (no residual check in
    between) stops the loop at the first iteration below it."""
    _, f = kept_run(oracle, deck("jacobi", 1, 0.0, dt=0.5))
    s = System(f)
    hist = jacobi_model(s, 40, 0.0)["hist"]
    n_stop = 30
    eps = math.sqrt(hist[n_stop - 1][0] * hist[n_stop - 2][0])
    rep, _ = kept_run(oracle, deck("jacobi", 1000, eps, dt=0.5))
    assert rep["steps"][0]["iters"] == n_stop - 1


def pre_chebyshev_model(sys_, presteps, max_iters, eps, stale=True, p_after_cg=True):
    """Preconditioned Chebyshev (tl_preconditioner_on, M = 1 / diagonal of
    the SMVP including the boundary faces, ext_cg_kernel.hpp:119-121) in
    explicit p form: p = a_k p + b_k M r, u += p, with a_k = w_{k+1} - 1 and
    b_k = w_{k+1} / theta from the closed-form omega.  While ||r||^2 after a
    Chebyshev step is above 1 the reference's switch test fails and a CG
    iteration runs on the Chebyshev state (tea_solve.f90:169-180, 378-433).
    The first such iteration's SMVP reads the depth-1 halo of p that the
    last CG iteration before the switch left (only u's halo is refreshed in
    Chebyshev mode, tea_solve.f90:203-204): A p + bsum (p - p_stale).
    Planted mistakes: stale=False uses the mirrored halo there;
    p_after_cg=False lets Chebyshev resume on the p from before that CG
    iteration's p update."""
    A, b = sys_.A, sys_.b
    Mi = 1.0 / (np.diag(A) + sys_.bsum)
    u = b.copy()
    r = b - A @ u
    p = Mi * r
    rro = r @ p
    error = 1e10
    cg_a, cg_b = [], []
    steps = 0
    out = {"switch_step": 0, "est_itc": 0, "hybrid": 0}
    p_stale = None
    n_done = max_iters
    for n in range(1, max_iters + 1):
        if n >= presteps and error <= 1.0:
            if steps == 0:
                M = max_iters - n + 2
                rro = error
                lo, hi = lanczos_eigs(cg_a, cg_b)
                eigmin, eigmax = lo * F95, hi * F105
                theta, delta = (eigmax + eigmin) / 2, (eigmax - eigmin) / 2
                sigma = theta / delta
                cn = eigmax / eigmin
                out.update(switch_step=n, eigmin=eigmin, eigmax=eigmax)
                p_stale = p.copy()
                bb = b @ b
                r = b - A @ u
                p = Mi * r / theta
                u = u + p
                r = b - A @ u
                w = omega(M, sigma)
                p = (w - 1.0) * p + (w / theta) * (Mi * r)
                u = u + p
                error = r @ r
                gamm = (math.sqrt(cn) - 1.0) / (math.sqrt(cn) + 1.0)
                out["est_itc"] = int(math.floor(math.log(EPS64 * bb / (4.0 * error)) / (2.0 * math.log(gamm)) + 0.5))
                steps = 1
            else:
                r = b - A @ u
                w = omega(steps, sigma)
                p = (w - 1.0) * p + (w / theta) * (Mi * r)
                u = u + p
                if n >= out["est_itc"] and n % 10 == 0:
                    error = r @ r
            steps += 1
        else:
            wv = A @ p
            if p_stale is not None and stale:
                wv = wv + sys_.bsum * (p - p_stale)
            if steps > 0:
                out["hybrid"] += 1
                p_stale = None  # the CG iteration's halo update refreshes p's halo from now on
            alpha = rro / (p @ wv)
            u = u + alpha * p
            r = r - alpha * wv
            z = Mi * r
            rrn = r @ z
            beta = rrn / rro
            if steps == 0:
                cg_a.append(alpha)
                cg_b.append(beta)
            p_new = z + beta * p
            if steps == 0 or p_after_cg:
                p = p_new
            error = rro = rrn
        if abs(error) < eps:
            n_done = n - 1
            break
    out["iters"] = n_done
    out["u"] = u
    return out


@pytest.mark.parametrize("extra", [1, 3, 6, 12])
def test_preconditioned_chebyshev_with_cg_iterations(oracle, extra):
    """Preconditioned Chebyshev on a deck whose ||r||^2 stays above 1 after
    the first Chebyshev step, so CG iterations run between Chebyshev steps:
    capped solves give the model's u; a mirrored halo in the first of those
    CG iterations, or Chebyshev resuming on the p from before the CG
    iteration's update, does not."""
    import decks

    def pdeck(cap, eps):
        return decks.two_rect_2d(14, "chebyshev", steps=1, eps=eps, dt=2.0, max_iters=cap, presteps=8, ny=11,
                                 extra=("tl_preconditioner_on",))

    full, _ = kept_run(oracle, pdeck(10000, 1e-16))
    st = full["steps"][0]
    ns = st["switch_step"]
    assert st["cg_steps"] > ns  # CG iterations after the switch
    cap = ns + extra
    rep, f = kept_run(oracle, pdeck(cap, 0.0))
    s = System(f)
    m = pre_chebyshev_model(s, 8, cap, 0.0)
    assert m["switch_step"] == rep["steps"][0]["switch_step"] == ns
    assert m["hybrid"] >= 1
    assert rel(interior(f, "u"), m["u"]) < TOL
    # a planted mistake may also diverge (nan): anything but agreement counts
    with np.errstate(all="ignore"):
        assert not rel(interior(f, "u"), pre_chebyshev_model(s, 8, cap, 0.0, stale=False)["u"]) <= MISS
        if extra >= 3:
            assert not rel(interior(f, "u"), pre_chebyshev_model(s, 8, cap, 0.0, p_after_cg=False)["u"]) <= MISS
