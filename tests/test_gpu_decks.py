"""End-to-end parity: decks run by the C++ driver on the GPU against the
oracle's restated Fortran driver on the CPU.

Bar: every per-step field-summary total
(vol, mass, ie, temp) within 1e-11 relative and the iteration count of every
timestep within +-1.  The QA decks additionally meet the reference's own
test: final temperature within 0.001 % of the value the reference ships
(tealeaf_2d/field_summary.f90:107, tealeaf_3d/field_summary.f90:115).
"""
import math

import pytest

from conftest import read_deck

pytestmark = pytest.mark.gpu

RTOL = 1e-11


def close(a, b):
    if math.isnan(b):
        return math.isnan(a)
    return abs(a - b) <= RTOL * abs(b)


def compare(gpu: dict, ref: dict):
    assert len(gpu["steps"]) == len(ref["steps"])
    for key in ("vol", "mass", "ie", "temp"):
        a, b = gpu["initial"][key], ref["initial"][key]
        assert close(a, b), ("initial", key, a, b)
    for sg, so in zip(gpu["steps"], ref["steps"]):
        assert abs(sg["iters"] - so["iters"]) <= 1, (sg["step"], sg["iters"], so["iters"])
        for key in ("vol", "mass", "ie", "temp"):
            if key in so:
                a, b = sg[key], so[key]
                assert close(a, b), (sg["step"], key, a, b)
    for key in ("vol", "mass", "ie", "temp"):
        a, b = gpu["final"][key], ref["final"][key]
        assert close(a, b), ("final", key, a, b)


def run_both(hip, oracle, deck, dim):
    print(f"deck dim={dim} cells={hip_cells(deck)}", flush=True)
    gpu = hip.run_deck(deck, dim=dim)
    ref = oracle.run_deck(deck, dim=dim, reduce_parts=8, threads=8)
    return gpu, ref


def hip_cells(deck):
    import decks
    return decks.cells(deck)


CASES_2D = [
    ("cg", 128, 3, {}), ("ppcg", 256, 2, {}), ("chebyshev", 256, 2, {}), ("jacobi", 128, 2, {"max_iters": 1200}),
    ("cg", 96, 2, {"extra": ("tl_preconditioner_on",)}),
    ("chebyshev", 180, 2, {"extra": ("tl_preconditioner_on",)}),
    ("chebyshev", 160, 2, {"extra": ("tl_ch_cg_errswitch", "tl_ch_cg_epslim=0.001")}),
    # temporal-blocked passes: a non-square mesh, and a cap that falls between norm checks
    ("chebyshev", 130, 2, {"ny": 47}), ("chebyshev", 128, 1, {"max_iters": 157}),
    # Jacobi that converges inside a pass (the pass is run again up to that iteration), and a cap
    # that is not a multiple of the pass
    ("jacobi", 40, 2, {"max_iters": 20000}), ("jacobi", 72, 1, {"max_iters": 1237, "ny": 51}),
    ("cg", 130, 2, {"extra": ("tl_coefficient_inverrse_density", "tl_check_result")}),
    # meshes a few cells wide: every strip and tile is an edge one, reflections repeat
    ("chebyshev", 3, 2, {"ny": 40, "presteps": 5}), ("ppcg", 2, 2, {"ny": 17, "presteps": 5}),
    ("chebyshev", 17, 2, {"ny": 2, "presteps": 3}),
]


@pytest.mark.parametrize("solver,n,steps,kw", CASES_2D)
def test_decks_2d(hip, oracle, solver, n, steps, kw):
    import decks
    deck = decks.two_rect_2d(n, solver, steps=steps, **kw)
    gpu, ref = run_both(hip, oracle, deck, 2)
    compare(gpu, ref)


@pytest.mark.parametrize("solver,n,steps", [("cg", 32, 2), ("ppcg", 48, 2), ("chebyshev", 40, 2), ("jacobi", 24, 2)])
def test_decks_3d(hip, oracle, solver, n, steps):
    import decks
    deck = decks.cube_3d(n, solver, steps=steps, max_iters=3000)
    gpu, ref = run_both(hip, oracle, deck, 3)
    compare(gpu, ref)


def test_random_rect_deck(hip, oracle):
    import decks
    deck = decks.random_rects_2d(192, seed=3, nrect=32, steps=2)
    gpu, ref = run_both(hip, oracle, deck, 2)
    compare(gpu, ref)


def test_random_shapes_deck(hip, oracle):
    import decks
    deck = decks.random_shapes_2d(160, seed=1001, steps=2)
    gpu, ref = run_both(hip, oracle, deck, 2)
    compare(gpu, ref)


def test_c1_512_cg(hip, oracle):
    """512^2 CG, 10 steps, eps 1e-10: the reference's CPU-runnable case."""
    import decks
    deck = decks.two_rect_2d(512, "cg", steps=10)
    gpu, ref = run_both(hip, oracle, deck, 2)
    compare(gpu, ref)


def test_qa_tp11_2d(hip, oracle):
    gpu, ref = run_both(hip, oracle, read_deck("tp11_2d.in"), 2)
    compare(gpu, ref)
    t = gpu["final"]["temp"]
    assert abs(100.0 * (t / 103.46970921076068) - 100.0) < 0.001


def test_qa_tp11_3d(hip, oracle):
    gpu, ref = run_both(hip, oracle, read_deck("tp11_3d.in"), 3)
    compare(gpu, ref)
    t = gpu["final"]["temp"]
    assert abs(100.0 * (t / 58.442896702081427) - 100.0) < 0.001
