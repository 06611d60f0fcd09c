"""The two ways a grid reduction is closed give the same bits.

Every sweep that computes a dot product or norm leaves one partial per
workgroup; the partials are added in a fixed order (thread t of a 256-lane
workgroup adds partials t, t + 256, ... in turn, then the xor butterfly per
wave, then the waves in order) either by a separate one-workgroup launch
(the default) or, with tl_set_single_pass(1), inside the sweep by the last
workgroup to finish.  Whole solves of every solver run both ways here and
must agree exactly: iteration counts, and every field-summary total
bitwise.  The meshes give grids of several 64-workgroup counter groups.
"""
import pytest

import decks

pytestmark = pytest.mark.gpu

CASES = [
    ("2d-cg", 2, lambda: decks.two_rect_2d(384, "cg", steps=2)),
    ("2d-ppcg", 2, lambda: decks.two_rect_2d(384, "ppcg", steps=2)),
    ("2d-chebyshev", 2, lambda: decks.two_rect_2d(320, "chebyshev", steps=2)),
    ("2d-jacobi", 2, lambda: decks.two_rect_2d(256, "jacobi", steps=2, max_iters=600)),
    ("2d-ppcg-mixed", 2, lambda: decks.two_rect_2d(256, "ppcg_mixed", steps=2, extra=("tl_check_result",))),
    ("2d-fcg-sweeps", 2, lambda: decks.two_rect_2d(256, "ppcg", steps=2, extra=("tl_check_result",)).replace(
        "tl_use_ppcg", "tl_use_fcg_learned\ntl_fcg_no_network")),
    ("3d-cg", 3, lambda: decks.cube_3d(48, "cg", steps=2)),
    ("3d-ppcg", 3, lambda: decks.cube_3d(48, "ppcg", steps=2)),
]


def run(hip, deck, dim, single):
    L = hip.lib(dim)
    L.tl_set_single_pass(1 if single else 0)
    try:
        return hip.run_deck(deck, dim=dim)
    finally:
        L.tl_set_single_pass(0)


@pytest.mark.parametrize("name,dim,make", CASES, ids=[c[0] for c in CASES])
def test_single_pass_matches_separate_finish(hip, name, dim, make):
    deck = make()
    print(f"reduction forms: {name}", flush=True)
    two = run(hip, deck, dim, False)
    one = run(hip, deck, dim, True)
    assert two["reductions"] == "separate finish" and one["reductions"] == "single-pass"
    assert len(one["steps"]) == len(two["steps"]) > 0
    for a, b in zip(one["steps"], two["steps"]):
        assert a["iters"] == b["iters"], (name, a["step"], a["iters"], b["iters"])
        for key in ("vol", "mass", "ie", "temp"):
            assert a[key] == b[key], (name, a["step"], key, a[key], b[key])
    for key in ("vol", "mass", "ie", "temp"):
        assert one["final"][key] == two["final"][key], (name, key)
