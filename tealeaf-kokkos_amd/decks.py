"""Generators of ``tea.in`` decks for the benchmark and test configurations.

Every deck is plain text in the reference's input format (``*tea`` ...
``*endtea``, see tealeaf_2d/read_input.f90), so the same file drives the HIP
library's C++ driver and any other TeaLeaf implementation.  All decks set
the common keywords: ext kernels, explicit ``tl_eps`` and
``initial_timestep`` (the reference's defaults are single-precision
literals), ``summary_frequency=1`` and a 10x10[x10] domain.

Random layouts use splitmix64 so a seed gives the same deck everywhere.
"""
from __future__ import annotations

SOLVER_KEYWORD = {"cg": "tl_use_cg", "chebyshev": "tl_use_chebyshev", "ppcg": "tl_use_ppcg",
                  "jacobi": "tl_use_jacobi",
                  # mixed-precision PPCG of the HIP driver (no counterpart in the reference)
                  "ppcg_mixed": "tl_use_ppcg_mixed"}


def fnum(x: float) -> str:
    """Shortest round-trip decimal that always contains a '.' (the reference
    appends '.0' to a number without one, parse.f90:342)."""
    s = repr(float(x))
    if "." in s:
        return s
    if "e" in s:
        m, e = s.split("e")
        return f"{m}.0e{e}"
    return s + ".0"


def _finish(lines: list[str]) -> str:
    for ln in lines:
        if len(ln) > 100:
            raise ValueError(f"deck line longer than 100 characters: {ln!r}")
    return "\n".join(["*tea"] + lines + ["*endtea", ""])


def _common(solver: str, steps: int, eps: float, dt: float, max_iters: int, presteps: int,
            inner: int, extra: list[str]) -> list[str]:
    out = [
        f"initial_timestep={fnum(dt)}",
        f"end_step={steps}",
        "summary_frequency=1",
        f"tl_max_iters={max_iters}",
        f"tl_eps={fnum(eps)}",
        "tl_coefficient_density",
        "use_ext_kernels",
        SOLVER_KEYWORD[solver],
    ]
    if solver in ("chebyshev", "ppcg", "ppcg_mixed"):
        out.append(f"tl_ch_cg_presteps={presteps}")
    if solver in ("ppcg", "ppcg_mixed"):
        out.append(f"tl_ppcg_inner_steps={inner}")
    return out + list(extra)


def two_rect_2d(n: int, solver: str = "cg", steps: int = 10, eps: float = 1e-10, dt: float = 0.004,
                max_iters: int = 10000, presteps: int = 30, inner: int = 10, ny: int | None = None,
                extra: tuple[str, ...] = ()) -> str:
    """Ambient density 100 / energy 1e-4 with a hot rectangle (0,0)-(1,2)
    (density 0.1, energy 25) and a rectangle (1,1)-(6,6) (density 0.1,
    energy 0.1) on 10x10: the 512^2 CG and 4096^2 PPCG/Chebyshev/Jacobi
    benchmark layouts."""
    lines = [
        "state 1 density=100.0 energy=0.0001",
        "state 2 density=0.1 energy=25.0 geometry=rectangle xmin=0.0 xmax=1.0 ymin=0.0 ymax=2.0",
        "state 3 density=0.1 energy=0.1 geometry=rectangle xmin=1.0 xmax=6.0 ymin=1.0 ymax=6.0",
        f"x_cells={n}",
        f"y_cells={ny or n}",
        "xmin=0.0", "ymin=0.0", "xmax=10.0", "ymax=10.0",
    ]
    return _finish(lines + _common(solver, steps, eps, dt, max_iters, presteps, inner, list(extra)))


def cube_3d(n: int, solver: str = "cg", steps: int = 5, eps: float = 1e-10, dt: float = 0.004,
            max_iters: int = 10000, presteps: int = 30, inner: int = 10, ny: int | None = None,
            nz: int | None = None, extra: tuple[str, ...] = ()) -> str:
    """3D layout: ambient density 100 / energy 1e-4, a hot box over the first
    10% of each axis (density 0.1, energy 25) and a cuboid over 10-60% of
    each axis (density 0.1, energy 0.1) on 10x10x10; n x ny x nz cells
    (ny, nz default to n)."""
    lines = [
        "state 1 density=100.0 energy=0.0001",
        "state 2 density=0.1 energy=25. geometry=cuboid xmin=0. xmax=1. ymin=0. ymax=1. zmin=0. zmax=1.",
        "state 3 density=0.1 energy=0.1 geometry=cuboid xmin=1. xmax=6. ymin=1. ymax=6. zmin=1. zmax=6.",
        f"x_cells={n}", f"y_cells={ny or n}", f"z_cells={nz or n}",
        "xmin=0.0", "ymin=0.0", "zmin=0.0", "xmax=10.0", "ymax=10.0", "zmax=10.0",
    ]
    return _finish(lines + _common(solver, steps, eps, dt, max_iters, presteps, inner, list(extra)))


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): x += 0x9E3779B97F4A7C15, then
    two xor-shift-multiply rounds."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.x = seed & self.MASK

    def next(self) -> int:
        self.x = (self.x + 0x9E3779B97F4A7C15) & self.MASK
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def randint(self, a: int, b: int) -> int:  # inclusive
        return a + self.next() % (b - a + 1)


def _sig(x: float, digits: int) -> float:
    return float(f"{x:.{digits}g}")


def _rect_line(k: int, dens: float, ener: float, box: tuple[float, float, float, float]) -> str:
    """A rectangle state on one line of at most 100 characters (the reader
    truncates longer lines, parse.f90:138): coordinates keep 3 decimals, or
    2 when that would not fit."""
    for nd in (3, 2, 1):
        x0, x1, y0, y1 = (round(v, nd) for v in box)
        ln = (f"state {k} density={fnum(dens)} energy={fnum(ener)} geometry=rectangle "
              f"xmin={fnum(x0)} xmax={fnum(x1)} ymin={fnum(y0)} ymax={fnum(y1)}")
        if len(ln) <= 100:
            return ln
    raise ValueError("rectangle state does not fit on one deck line")


def random_rects_2d(n: int, seed: int, nrect: int = 32, solver: str = "ppcg", steps: int = 10,
                    eps: float = 1e-10, dt: float = 0.004, max_iters: int = 10000, presteps: int = 30,
                    inner: int = 10, extra: tuple[str, ...] = ()) -> str:
    """Ambient density 100 / energy 1e-4 plus `nrect` axis-aligned rectangles
    with corners uniform in [0,10]^2 (sorted per axis), density
    10^U(-1, 2) and energy U(1e-4, 25), three significant digits."""
    rng = SplitMix64(seed)
    lines = ["state 1 density=100.0 energy=0.0001"]
    for k in range(nrect):
        x0, x1 = sorted((rng.uniform(0, 10), rng.uniform(0, 10)))
        y0, y1 = sorted((rng.uniform(0, 10), rng.uniform(0, 10)))
        dens = _sig(10.0 ** rng.uniform(-1.0, 2.0), 3)
        ener = _sig(rng.uniform(1e-4, 25.0), 3)
        lines.append(_rect_line(k + 2, dens, ener, (x0, x1, y0, y1)))
    lines += [f"x_cells={n}", f"y_cells={n}", "xmin=0.0", "ymin=0.0", "xmax=10.0", "ymax=10.0"]
    return _finish(lines + _common(solver, steps, eps, dt, max_iters, presteps, inner, list(extra)))


def random_shapes_2d(n: int, seed: int, nmin: int = 8, nmax: int = 64, solver: str = "ppcg", steps: int = 1,
                     eps: float = 1e-10, dt: float = 0.004, max_iters: int = 10000, presteps: int = 30,
                     inner: int = 10, extra: tuple[str, ...] = ()) -> str:
    """Ambient density 100 / energy 1e-4 plus between nmin and nmax
    rectangles and circles (centre uniform in [0,10]^2, radius U(0.2, 2)),
    density 10^U(-1, 2), energy U(1e-4, 25): the learned-preconditioner
    training and evaluation layouts."""
    rng = SplitMix64(seed)
    count = rng.randint(nmin, nmax)
    lines = ["state 1 density=100.0 energy=0.0001"]
    for k in range(count):
        dens = _sig(10.0 ** rng.uniform(-1.0, 2.0), 3)
        ener = _sig(rng.uniform(1e-4, 25.0), 3)
        if rng.next() & 1:
            x0, x1 = sorted((rng.uniform(0, 10), rng.uniform(0, 10)))
            y0, y1 = sorted((rng.uniform(0, 10), rng.uniform(0, 10)))
            lines.append(_rect_line(k + 2, dens, ener, (x0, x1, y0, y1)))
        else:
            cx, cy = round(rng.uniform(0, 10), 3), round(rng.uniform(0, 10), 3)
            rad = round(rng.uniform(0.2, 2.0), 3)
            lines.append(f"state {k + 2} density={fnum(dens)} energy={fnum(ener)} geometry=circle "
                         f"xmin={fnum(cx)} ymin={fnum(cy)} radius={fnum(rad)}")
    lines += [f"x_cells={n}", f"y_cells={n}", "xmin=0.0", "ymin=0.0", "xmax=10.0", "ymax=10.0"]
    return _finish(lines + _common(solver, steps, eps, dt, max_iters, presteps, inner, list(extra)))


GEOMETRY = {"rectangle": 1, "circle": 2, "point": 3}


def states(deck: str) -> dict:
    """The 2D states of a deck as ext_generate_chunk_ takes them (lists
    indexed by state; geometry 0 for the ambient state 1), with the extents
    of states 2..n moved inwards by a hundredth of a cell
    (tealeaf_2d/read_input.f90:284-297; dx from the single-precision cell
    count, read_input.f90:290)."""
    import numpy as np

    grid = {"xmin": 0.0, "xmax": 100.0, "ymin": 0.0, "ymax": 100.0, "x_cells": 10, "y_cells": 10}
    st = {}
    for ln in deck.splitlines():
        words = ln.replace("=", " ").split()
        if not words:
            continue
        if words[0] == "state":
            k = int(words[1])
            kv = dict(zip(words[2::2], words[3::2]))
            st[k] = {"density": float(kv["density"]), "energy": float(kv["energy"]),
                     "geometry": GEOMETRY.get(kv.get("geometry", ""), 0),
                     **{key: float(kv.get(key, 0.0)) for key in ("xmin", "xmax", "ymin", "ymax", "radius")}}
        elif words[0] in grid and len(words) > 1:
            grid[words[0]] = type(grid[words[0]])(float(words[1]) if words[0].endswith("cells") else words[1])
    dx = (grid["xmax"] - grid["xmin"]) / float(np.float32(grid["x_cells"]))
    dy = (grid["ymax"] - grid["ymin"]) / float(np.float32(grid["y_cells"]))
    out = {key: [] for key in ("density", "energy", "xmin", "xmax", "ymin", "ymax", "radius", "geometry")}
    for k in sorted(st):
        s = dict(st[k])
        if k >= 2:
            s["xmin"] += dx / 100.0
            s["ymin"] += dy / 100.0
            s["xmax"] -= dx / 100.0
            s["ymax"] -= dy / 100.0
        if k == 1:
            s["geometry"] = 0
        for key in out:
            out[key].append(s[key])
    out.update(dx=dx, dy=dy, grid_xmin=grid["xmin"], grid_ymin=grid["ymin"])
    return out


def cells(deck: str) -> int:
    """Interior cell count of a deck (x_cells * y_cells [* z_cells])."""
    n = 1
    for key in ("x_cells", "y_cells", "z_cells"):
        for ln in deck.splitlines():
            ln = ln.replace("=", " ").split()
            if len(ln) >= 2 and ln[0] == key:
                n *= int(ln[1])
    return n

