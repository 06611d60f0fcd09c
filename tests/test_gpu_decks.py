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
    # preconditioned Chebyshev whose ||r||^2 is still above 1 after the first step: a CG
    # iteration follows it (reading the halo of p the reference left from before the switch),
    # then Chebyshev resumes on that iteration's p
    ("chebyshev", 64, 2, {"dt": 0.4, "presteps": 20, "extra": ("tl_preconditioner_on",)}),
    ("chebyshev", 96, 2, {"ny": 40, "dt": 0.4, "presteps": 20, "extra": ("tl_preconditioner_on",)}),
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


# Ragged 3D meshes for the sweeps along z: a second, partial 64-column tile
# (wave-edge neighbours across tiles, lanes past the last column), a partial
# 4-row tile and a partial strip of planes, with and without the diagonal
# preconditioner.
CASES_3D_RAGGED = [
    ("cg", 70, 9, 37, 0.04, 20, ()), ("cg", 70, 9, 37, 0.04, 20, ("tl_preconditioner_on",)),
    ("ppcg", 67, 6, 21, 0.4, 5, ()), ("ppcg", 67, 6, 21, 0.04, 20, ("tl_preconditioner_on",)),
    ("chebyshev", 65, 7, 19, 0.4, 5, ()), ("chebyshev", 65, 7, 19, 0.4, 20, ("tl_preconditioner_on",)),
    # a CG iteration after the first Chebyshev step (see the 2D cases)
    ("chebyshev", 24, 24, 24, 0.4, 20, ("tl_preconditioner_on",)),
]


@pytest.mark.parametrize("solver,nx,ny,nz,dt,presteps,extra", CASES_3D_RAGGED)
def test_decks_3d_ragged(hip, oracle, solver, nx, ny, nz, dt, presteps, extra):
    import decks
    deck = decks.cube_3d(nx, solver, steps=2, max_iters=3000, ny=ny, nz=nz, dt=dt, presteps=presteps, extra=extra)
    gpu, ref = run_both(hip, oracle, deck, 3)
    compare(gpu, ref)


@pytest.mark.parametrize("inner", [1, 3, 4])
def test_decks_3d_ppcg_inner_counts(hip, oracle, inner):
    """3D PPCG runs its inner steps two per sweep: odd counts end with a
    single step, and one step is a single sweep."""
    import decks
    deck = decks.cube_3d(67, "ppcg", steps=2, max_iters=3000, ny=6, nz=21, dt=0.4, presteps=5, inner=inner)
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


# ---- tea.out lines -----------------------------------------------------------

def _num(text: str) -> float:
    """A Fortran-written number; Ew.d drops the E of a three-digit exponent."""
    t = text.strip()
    if "E" not in t and ("+" in t[1:] or "-" in t[1:]):
        k = max(t.rfind("+"), t.rfind("-"))
        t = t[:k] + "E" + t[k:]
    return float(t)


def _expected_tea_out(hip, ref: dict, check_result: bool, cheby: bool, dim: int):
    """The tea.out lines for the oracle's report, as fixed-width pieces:
    text, or (descriptor, value, kind) with kind "num" (within the printed
    precision), "iters" (within +-1) or "any" (not in the oracle's report)."""
    lines = []

    def summary(s):
        lines.append([hip.fortran_edit("a13", "           ", dim)] +
                     [hip.fortran_edit("a16", h, dim) for h in ("Volume", "Mass", "Density", "Energy", "U")])
        lines.append([" step:", ("i7", s["step"], "exact")] +
                     [("e16.7", v, "num") for v in (s["vol"], s["mass"], s["mass"] / s["vol"], s["ie"], s["temp"])])

    summary(ref["initial"])
    for s in ref["steps"]:
        if s["switch_step"] > 0:
            lines.append(["Switching after ", ("i3", s["switch_step"], "iters"), " steps, error ",
                          ("e15.7", 0.0, "any")])
            lines.append([hip.fortran_edit("a11", h, dim) for h in ("eigmin", "eigmax", "cn", "error")])
            lines.append([("f11.5", s["eigmin"], "f5"), ("f11.5", s["eigmax"], "f5"),
                          ("e11.4", s["eigmax"] / s["eigmin"], "num4"), ("e11.4", 0.0, "any")])
            if cheby:
                lines.append([hip.fortran_edit("a11", "est itc", dim)])
                lines.append([("i11", s["est_itc"], "iters")])
        lines.append(["Conduction error ", ("e14.7", s["error"], "any")])
        if check_result:
            lines.append(["EXACT error calculated as", ("e14.7", s["exact_error"], "any")])
        lines.append(["Iteration count ", ("i8", s["iters"], "iters")])
        if "vol" in s:
            summary(s)
    summary(ref["final"])
    return lines


def _check_line(hip, got: str, pieces, dim: int):
    pos = 0
    for p in pieces:
        if isinstance(p, str):
            assert got[pos:pos + len(p)] == p, (got, p)
            pos += len(p)
            continue
        desc, value, kind = p
        w = int(desc[1:].split(".")[0])
        field = got[pos:pos + w]
        pos += w
        assert len(field) == w and field == field.rjust(w) and "*" not in field, (got, desc)
        if kind == "exact":
            assert field == hip.fortran_edit(desc, value, dim), (got, desc, value)
        elif kind == "iters":
            assert abs(int(field) - value) <= 1, (got, value)
        elif kind == "any":
            _num(field)
        else:
            tol = {"num": 1.1e-6, "num4": 1.1e-3}.get(kind)  # one unit of the last printed digit
            if tol is None:  # f11.5: five decimals
                assert abs(float(field) - value) <= 1.1e-5 + 1e-9 * abs(value), (got, value)
            else:
                assert abs(_num(field) - value) <= tol * abs(value) + 1e-300, (got, value)
    assert pos == len(got), (got, pos)


TEA_OUT_CASES = [("chebyshev", 2, 128), ("ppcg", 2, 128), ("cg", 2, 48), ("chebyshev", 3, 24)]


@pytest.mark.parametrize("solver,dim,n", TEA_OUT_CASES)
def test_tea_out_lines(hip, oracle, solver, dim, n):
    """The driver's tea.out lines have the reference's fixed layout and the
    oracle's values: field summaries within the 7 printed digits, switch
    step, eigenvalue line and est itc, conduction error, iteration counts
    within +-1."""
    import decks
    if dim == 2:
        deck = decks.two_rect_2d(n, solver, steps=3, extra=("tl_check_result",))
    else:
        deck = decks.cube_3d(n, solver, steps=2, max_iters=3000, extra=("tl_check_result",))
    print(f"tea.out {solver} dim={dim} n={n}", flush=True)
    gpu = hip.run_deck(deck, dim=dim)
    ref = oracle.run_deck(deck, dim=dim, reduce_parts=8, threads=8)
    got = gpu["tea_out"].splitlines()
    exp = _expected_tea_out(hip, ref, True, solver == "chebyshev", dim)
    assert len(got) == len(exp), (len(got), len(exp), got[:12])
    for g, e in zip(got, exp):
        _check_line(hip, g, e, dim)
    if dim == 2 and solver in ("chebyshev", "ppcg"):  # these meshes switch after 30 CG steps
        assert any(line.startswith("Switching after") for line in got)


def test_tea_out_qa_verdict(hip):
    """The reference's own QA deck (test problem 11, 512^2): tea.out ends
    with its verdict line."""
    rep = hip.run_deck(read_deck("tp11_2d.in"), dim=2)
    lines = rep["tea_out"].splitlines()
    assert lines[-2].startswith("Test problem  11 is within")
    assert lines[-2].endswith("% of the expected solution")
    qa = _num(lines[-2][len("Test problem  11 is within"):len("Test problem  11 is within") + 16])
    assert qa < 0.001
    assert lines[-1] == " This test is considered PASSED"


@pytest.mark.parametrize("dim", [2, 3])
def test_visit_files(hip, tmp_path, monkeypatch, dim):
    """visit_frequency (diffuse.f90:62-63,93; visit.f90): tea.visit starts
    with '!NBLOCKS     1' and lists tea.00001.<step>.vtk for every
    visit_frequency-th step and again for the last one; each file is the
    reference's header-only rectilinear grid with the vertex coordinates in
    E12.4 (no field values: the reference writes none)."""
    import decks
    monkeypatch.chdir(tmp_path)
    if dim == 2:
        n, nz = 16, 1
        deck = decks.two_rect_2d(n, "cg", steps=5, extra=("visit_frequency=2",))
        visits = [2, 4, 5]
    else:
        n, nz = 8, 8
        deck = decks.cube_3d(n, "cg", steps=2, extra=("visit_frequency=1",))
        visits = [1, 2, 2]
    hip.run_deck(deck, dim=dim)
    listed = (tmp_path / "tea.visit").read_text().splitlines()
    assert listed == ["!NBLOCKS     1"] + [f"tea.00001.{s:05d}.vtk" for s in visits]
    lines = (tmp_path / f"tea.00001.{visits[-1]:05d}.vtk").read_text().splitlines()
    nv = n + 1
    exp = ["# vtk DataFile Version 3.0", "vtk output", "ASCII", "DATASET RECTILINEAR_GRID",
           f"DIMENSIONS{nv:12d}{nv:12d}{(nz + 1) if dim == 3 else 1:12d}"]
    dx = 10.0 / n
    for axis in ("X", "Y") + (("Z",) if dim == 3 else ()):
        exp.append(f"{axis}_COORDINATES {nv:5d} double")
        exp += [hip.fortran_edit("e12.4", dx * j, dim) for j in range(nv)]
    total = nv * nv * (nz + 1 if dim == 3 else 1)
    exp += ["FIELD FieldData 3", f"x_vel 1 {total:20d} double", f"y_vel 1 {total:20d} double"]
    assert lines == exp
