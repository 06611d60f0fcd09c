"""Writes the full-size oracle fixtures under tests/golden/fullsize/.

Each fixture is the oracle's FP64 run of one benchmark deck at its real
size, repeated for two Kokkos-OpenMP pool sizes (``reduce_parts`` 4 and 8):
the pool size fixes the order of every FP64 sum in the reference's CPU build
(Kokkos_ExecPolicy.hpp:200-219, Kokkos_OpenMP_Parallel.hpp:175-193), so the
two runs show how far the reference's own results move with nothing but the
summation order changed.  Per timestep the fixture keeps the iteration
count, the solver switch step, est_itc, eigmin/eigmax, the recurrence and
true residuals, and the field-summary totals (vol, mass, ie, temp) as JSON
doubles (round-trip exact).

The decks are those of the benchmark configurations:
  c2_<solver>_4096  4096^2 two rectangles, PPCG (30 presteps, 10 inner
                    steps), Chebyshev (30 presteps) and Jacobi (cap 20000),
                    the first two timesteps;
  c3_<solver>_256   256^3 hot box and cuboid, CG and PPCG, five timesteps.
Every deck also sets tl_check_result, so the true residual ||u0 - A u||^2 of
each solve is recorded.

Run from the repository root on a CPU host (the 4096^2 PPCG and Jacobi
cases take tens of minutes each on 8 cores):
    python3 tests/golden/make_fullsize_fixtures.py [case ...]
"""
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tealeaf-kokkos_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import decks  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(HERE, "fullsize")
PARTS = (4, 8)
STEP_KEYS = ("step", "iters", "switch_step", "est_itc", "eigmin", "eigmax", "cg_steps", "cheby_steps", "error",
             "exact_error", "vol", "mass", "ie", "temp", "solve_s")


def deck_for(case: str) -> tuple[str, int]:
    """(deck text, dimension) of a fixture case."""
    cfg, solver, n = case.split("_")
    n = int(n)
    extra = ("tl_check_result",)
    if cfg == "c2":
        max_iters = 20000 if solver == "jacobi" else 10000
        return decks.two_rect_2d(n, solver, steps=2, presteps=30, inner=10, max_iters=max_iters, extra=extra), 2
    if cfg == "c3":
        return decks.cube_3d(n, solver, steps=5, presteps=30, inner=10, max_iters=10000, extra=extra), 3
    raise ValueError(case)


CASES = ["c2_chebyshev_4096", "c3_cg_256", "c3_ppcg_256", "c2_ppcg_4096", "c2_jacobi_4096"]


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_case(case: str, threads: int) -> dict:
    deck, dim = deck_for(case)
    runs = []
    for parts in PARTS:
        t0 = time.time()
        rep = oracle.run_deck(deck, dim=dim, reduce_parts=parts, threads=threads)
        wall = time.time() - t0
        runs.append({
            "reduce_parts": parts,
            "wall_s": round(wall, 1),
            "initial": {k: rep["initial"][k] for k in ("vol", "mass", "ie", "temp")},
            "steps": [{k: s[k] for k in STEP_KEYS if k in s} for s in rep["steps"]],
        })
        print(f"{case} parts={parts}: {wall:.0f} s, iters "
              f"{[s['iters'] for s in rep['steps']]}", flush=True)
    return {"case": case, "dim": dim, "deck": deck, "threads": threads, "cpu": cpu_model(), "runs": runs}


def main(argv: list[str]) -> int:
    oracle.build()
    os.makedirs(OUT, exist_ok=True)
    threads = os.cpu_count() or 8
    for case in argv or CASES:
        fix = run_case(case, threads)
        path = os.path.join(OUT, f"{case}.json")
        with open(path, "w") as f:
            json.dump(fix, f, indent=1)
            f.write("\n")
        print(f"wrote {path}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
