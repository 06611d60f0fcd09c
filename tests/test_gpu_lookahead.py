"""Speculative lookahead: the driver launches CG iteration n + 1 (before any
switch) and FP64 PPCG outer iteration n + 1 (the temporal-blocked form)
before it reads iteration n's rrn; an rrn that converges or leaves the
method sets a device flag that turns the launched iteration into no-ops.

Bar: bitwise equal to the one-at-a-time loop (``tl_no_lookahead``) in every
per-step iteration count, error, eigenvalue estimate, switch step and
field-summary total, and the speculative path must actually have run.
"""
import pytest

import decks

pytestmark = pytest.mark.gpu

STEP_KEYS = ("iters", "error", "switch_step", "est_itc", "eigmin", "eigmax", "cg_steps", "cheby_steps",
             "vol", "mass", "ie", "temp")

CASES = [
    # CG: converges with an iteration in flight; a cap that ends it early
    ("cg", 2, 128, dict(steps=3)),
    ("cg", 2, 64, dict(steps=1, max_iters=37)),
    ("cg", 3, 24, dict(steps=2)),
    # CG up to the switch (the iteration in flight is undone), then Chebyshev or PPCG
    ("chebyshev", 2, 256, dict(steps=2)),
    ("chebyshev", 2, 160, dict(steps=2, extra=("tl_ch_cg_errswitch", "tl_ch_cg_epslim=0.001"))),
    ("ppcg", 2, 256, dict(steps=2)),
    ("ppcg", 2, 130, dict(steps=2, ny=47, inner=7)),
    ("ppcg", 2, 160, dict(steps=2, extra=("tl_ch_cg_errswitch", "tl_ch_cg_epslim=0.001"))),
    # a PPCG outer iteration whose rrn is above 1: CG iterations follow it (no
    # switch test passes again before the cap)
    ("ppcg", 2, 2, dict(steps=2, ny=17, dt=0.4, presteps=2, max_iters=300)),
    # a cap inside the PPCG iterations
    ("ppcg", 2, 128, dict(steps=1, max_iters=45)),
    ("ppcg", 3, 24, dict(steps=2)),
]


def deck_for(solver, dim, n, kw, extra=()):
    kw = dict(kw)
    kw["extra"] = tuple(kw.get("extra", ())) + tuple(extra)
    if dim == 2:
        return decks.two_rect_2d(n, solver, **kw)
    return decks.cube_3d(n, solver, **kw)


@pytest.mark.parametrize("solver,dim,n,kw", CASES)
def test_lookahead_matches_one_at_a_time(hip, solver, dim, n, kw):
    print(f"lookahead {solver} dim={dim} n={n} {kw}", flush=True)
    fast = hip.run_deck(deck_for(solver, dim, n, kw), dim=dim)
    slow = hip.run_deck(deck_for(solver, dim, n, kw, ("tl_no_lookahead",)), dim=dim)
    assert len(fast["steps"]) == len(slow["steps"])
    for a, b in zip(fast["steps"], slow["steps"]):
        for key in STEP_KEYS:
            if key in b:
                assert a[key] == b[key], (a["step"], key, a[key], b[key])
        assert b["ahead"] == 0
    for key in ("vol", "mass", "ie", "temp"):
        assert fast["final"][key] == slow["final"][key], key
    ahead = sum(s["ahead"] for s in fast["steps"])
    undone = sum(s["ahead_undone"] for s in fast["steps"])
    assert ahead > 0
    # at most one launched iteration per method change or end of solve is undone
    assert undone <= 3 * len(fast["steps"])
