"""Mixed-precision PPCG (FP32 inner steps, FP64 outer iteration) on the GPU.

The mixed form has no counterpart in the reference, so its bar is stated
here rather than inherited:

* one outer iteration keeps u and r consistent in FP64: with u' = u + alpha p
  and r' = r - alpha w, the new residual equals r' - A (u_new - u') up to
  the FP32 recurrence error of the inner steps, checked against an FP64 numpy
  evaluation of A (the 5-point operator of ext_shared.hpp:37-41 with
  reflective faces);
* u and r stay within an FP32-sized distance of the oracle's FP64 outer
  iteration, and rrn is the FP64 sum of the new r squared;
* a whole solve converges on the true FP64 residual (||u0 - A u||^2 < eps,
  the deck's tl_check_result value), and its per-step field summaries agree
  with the oracle's FP64 PPCG run of the same deck to MIXED_RTOL.
"""
import ctypes

import numpy as np
import pytest

from test_gpu_kernels import FIELD_SD, assert_sum, random_fields

pytestmark = pytest.mark.gpu

# FP32 inner steps: relative bound on the recurrence/true-residual gap of one
# outer iteration, and on u, r against the FP64 iteration (relative to the
# size of the inner-step correction).
GAP_RTOL = 1e-5
STEP_RTOL = 1e-4
# per-step field summaries of a full mixed solve against FP64 PPCG, both
# converged to ||r||^2 < 1e-10
MIXED_RTOL = 1e-7


def apply_a(x, kx, ky):
    """(A x) on the interior of full 2D arrays [Y][X], reflective faces."""
    ny, nx = x.shape[0] - 4, x.shape[1] - 4
    j = np.arange(2, ny + 2)[:, None]
    k = np.arange(2, nx + 2)[None, :]
    c = x[j, k]
    e = x[j, np.minimum(k + 1, nx + 1)]
    w = x[j, np.maximum(k - 1, 2)]
    n = x[np.minimum(j + 1, ny + 1), k]
    s = x[np.maximum(j - 1, 2), k]
    return c + kx[j, k + 1] * (c - e) + kx[j, k] * (c - w) + ky[j + 1, k] * (c - n) + ky[j, k] * (c - s)


@pytest.mark.parametrize("shape,nsteps", [((37, 23), 10), ((150, 140), 10), ((150, 140), 3), ((64, 200), 1),
                                          ((300, 257), 10), ((5, 4), 10), ((3, 9), 3)])
def test_ppcg_outer_mixed(hip, oracle, shape, nsteps):
    nx, ny = shape
    print(f"mixed outer {shape} nsteps={nsteps}", flush=True)
    g = hip.ExtChunk(hip.lib(2), 2, nx, ny)
    o = hip.ExtChunk(oracle.lib(2), 2, nx, ny)
    rng = np.random.default_rng(nx * 1000 + ny + nsteps + 7)
    random_fields(g, o, ["u", "p", "r", "w"], rng, lo=-1.0, hi=1.0)
    random_fields(g, o, ["kx", "ky"], rng, lo=0.1, hi=40.0)
    u0, p0, r0, w0 = (o.get(f)[0] for f in ("u", "p", "r", "w"))
    kx, ky = o.get("kx")[0], o.get("ky")[0]
    # Chebyshev coefficients for an interval that holds the spectrum of A
    # (1 <= lambda <= 1 + 8 max k): outside it the polynomial grows, and so
    # would the FP32 error of the inner steps
    a, b, theta = oracle.calc_ch_coefs(1.0, 1.05 * (1.0 + 8 * 40.0), nsteps)
    alpha = 0.37
    ca = (ctypes.c_double * nsteps)(*a)
    cb = (ctypes.c_double * nsteps)(*b)
    rrn_g = ctypes.c_double()
    rc = g.L.tl_ppcg_outer_mixed(g.c, ctypes.c_double(alpha), ctypes.c_double(theta), ca, cb, nsteps,
                                 ctypes.byref(rrn_g))
    assert rc == 0, g.L.tl_last_error()
    ug, rg = g.get("u")[0], g.get("r")[0]

    # FP64 reference of the same outer iteration (separate sweeps on the oracle)
    o.ppcg_init(a, b, theta)
    o.cg_calc_ur(alpha)
    o.ppcg_init_sd()
    for step in range(1, nsteps + 1):
        o.update_halo([FIELD_SD], 1)
        o.ppcg_inner(step)
    uo, ro = o.get("u")[0], o.get("r")[0]

    inner = (slice(2, ny + 2), slice(2, nx + 2))
    up = (u0 + alpha * p0)[inner]
    rp = (r0 - alpha * w0)[inner]
    du_g = ug[inner] - up
    full = np.zeros_like(ug)
    full[inner] = du_g
    gap = rg[inner] - (rp - apply_a(full, kx, ky))
    scale_r = np.max(np.abs(rp))
    print(f"  gap {np.max(np.abs(gap)) / scale_r:.3e}  |du| {np.max(np.abs(du_g)):.3e}", flush=True)
    assert np.max(np.abs(gap)) <= GAP_RTOL * scale_r

    du_o = uo[inner] - up
    dr_o = ro[inner] - rp
    eu = np.max(np.abs(ug[inner] - uo[inner])) / np.max(np.abs(du_o))
    er = np.max(np.abs(rg[inner] - ro[inner])) / np.max(np.abs(dr_o))
    print(f"  u {eu:.3e}  r {er:.3e}", flush=True)
    assert eu <= STEP_RTOL and er <= STEP_RTOL
    # the inner steps did run in reduced precision
    assert eu > 0.0
    assert_sum(rrn_g.value, float(np.sum(rg[inner] ** 2)), float(np.sum(rg[inner] ** 2)))
    g.close()
    o.close()


def _fp64_deck(deck):
    return deck.replace("tl_use_ppcg_mixed", "tl_use_ppcg")


def _compare_mixed(gpu, ref, eps):
    assert gpu["solver"] == "ppcg_mixed"
    assert len(gpu["steps"]) == len(ref["steps"])
    worst = 0.0
    for sg, so in zip(gpu["steps"], ref["steps"]):
        assert sg["mixed_steps"] > 0
        assert sg["exact_error"] < eps, (sg["step"], sg["exact_error"])
        for key in ("vol", "mass", "ie", "temp"):
            rel = abs(sg[key] - so[key]) / abs(so[key])
            worst = max(worst, rel)
            assert rel <= MIXED_RTOL, (sg["step"], key, sg[key], so[key])
        # the same number of outer iterations to within a few
        assert abs(sg["iters"] - so["iters"]) <= max(5, so["iters"] // 20), (sg["step"], sg["iters"], so["iters"])
    print(f"  worst field-summary difference {worst:.3e}", flush=True)


@pytest.mark.parametrize("n,steps", [(256, 3), (200, 2)])
def test_mixed_two_rect(hip, oracle, n, steps):
    import decks
    deck = decks.two_rect_2d(n, "ppcg_mixed", steps=steps, extra=("tl_check_result",))
    print(f"mixed deck two_rect n={n}", flush=True)
    gpu = hip.run_deck(deck, dim=2)
    ref = oracle.run_deck(_fp64_deck(deck), dim=2, reduce_parts=8, threads=8)
    _compare_mixed(gpu, ref, 1e-10)


@pytest.mark.parametrize("seed", [0, 3])
def test_mixed_random_rects(hip, oracle, seed):
    import decks
    deck = decks.random_rects_2d(192, seed, steps=2, extra=("tl_check_result",)).replace(
        "tl_use_ppcg", "tl_use_ppcg_mixed")
    print(f"mixed deck random rects seed={seed}", flush=True)
    gpu = hip.run_deck(deck, dim=2)
    ref = oracle.run_deck(_fp64_deck(deck), dim=2, reduce_parts=8, threads=8)
    _compare_mixed(gpu, ref, 1e-10)


def test_mixed_unsupported_inner_steps_runs_fp64(hip, oracle):
    """Inner-step counts without a compiled FP32 kernel run the FP64 path:
    then the results are the FP64 PPCG's (1e-11, iterations +-1)."""
    import decks
    from test_gpu_decks import compare
    deck = decks.two_rect_2d(128, "ppcg_mixed", steps=2, inner=4)
    gpu = hip.run_deck(deck, dim=2)
    ref = oracle.run_deck(_fp64_deck(deck), dim=2, reduce_parts=8, threads=8)
    assert all(s["mixed_steps"] == 0 for s in gpu["steps"])
    compare(gpu, ref)


def test_mixed_hook_rejects_3d(hip):
    L = hip.lib(3)
    ch = hip.ExtChunk(L, 3, 8, 8, 8)
    out = ctypes.c_double()
    a = (ctypes.c_double * 10)(*([0.5] * 10))
    rc = L.tl_ppcg_outer_mixed(ch.c, ctypes.c_double(0.1), ctypes.c_double(2.0), a, a, 10, ctypes.byref(out))
    assert rc == -1
    assert b"mixed" in L.tl_last_error()
    L.tl_clear_error()
    ch.close()
