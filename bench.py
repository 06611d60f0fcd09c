#!/usr/bin/env python3
"""Benchmark: TeaLeaf's linear solve on one MI355X.

A step is one timestep, i.e. one ``tea_leaf()`` solve from its first halo
update to the finalised energy (the reference's "Tea Solve" time per step,
tealeaf_2d/diffuse.f90:149).  --warmup W timesteps run untimed first, then
--steps K are timed; the driver times each solve with a device
synchronisation at its end.

Default workload c2: the 2D 4096^2 two-rectangle deck (ambient density 100,
energy 1e-4; rectangles (0,0)-(1,2) and (1,1)-(6,6) of density 0.1) on a
10x10 domain, eps 1e-10, dt 0.004.  `value` is its FP64 PPCG solve (30 CG
presteps, 10 inner Chebyshev steps: the reference's solver, rebuilt) in
seconds per timestep over the K timed steps; `forms` holds the same deck
and timesteps solved by mixed-precision PPCG and by the flexible CG with
Jacobi sweeps.  The line also carries the roofline of the FP64 solve's
dominant kernel (algorithmic bytes per launch over its mean HIP-event
launch time) and a CPU baseline from the oracle (see cpu_baseline()).

Attached to it, over a fixed --c4-warmup + --c4-steps timesteps (1 + 2) so
that the run's length does not grow with K: the comparison the target is
stated on, c4, 8192^2 with 32 seeded random rectangles (splitmix64 seed 0,
density log-uniform in [0.1, 100], energy uniform in [1e-4, 25]).  FP64
PPCG against mixed-precision PPCG (FP32 inner steps), learned flexible CG
(BF16 U-Net + 4 weighted Jacobi sweeps, trained first), the same flexible
CG without the network (the ablation), and FP64 PPCG accepted on the true
residual ||u0 - A u||^2 (the equal-accuracy baseline: the reference's PPCG
stops on its recurrence residual), under `forms_c4`; `north_star` sums up
the speedups against the 1.5x target.  --workload c4 runs that comparison
over W + K timesteps as a line of its own (about 15 s per FP64 timestep),
c3 the 3D 256^3 cube, c5 the 4096^2 learned-FCG configuration; --solver
picks CG, Chebyshev or Jacobi instead of PPCG for c2/c3.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tealeaf-kokkos_amd"))

import decks  # noqa: E402
import tealeaf_hip  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def cpu_info() -> dict:
    """The host CPU as the baseline uses it: model name, physical cores in
    this process's CPU set, and the OpenMP thread cap of the environment."""
    model = ""
    cores = set()
    allowed = os.sched_getaffinity(0)
    try:
        with open("/proc/cpuinfo") as f:
            cpu = phys = core = None
            for ln in f.read().splitlines() + [""]:
                if ln.startswith("processor"):
                    cpu = int(ln.split(":")[1])
                elif ln.startswith("model name") and not model:
                    model = ln.split(":", 1)[1].strip()
                elif ln.startswith("physical id"):
                    phys = ln.split(":")[1].strip()
                elif ln.startswith("core id"):
                    core = ln.split(":")[1].strip()
                elif not ln.strip() and cpu is not None:
                    if cpu in allowed:
                        cores.add((phys, core))
                    cpu = phys = core = None
    except OSError:
        pass
    physical = len(cores) or len(allowed)
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"model": model, "physical_cores": physical, "logical_cpus": len(allowed),
            "threads": min(physical, cap) if cap > 0 else physical}


def cpu_baseline(deck: str, n: int, threads: int, steps: list[dict], budget_s: float = 20.0) -> dict:
    """The oracle (C/OpenMP restatement of the reference) timed on a bounded
    sample of the same workload: the deck's n^2 chunk in its first-timestep
    state (layout rebuilt through ext_generate_chunk_ from the deck's states)
    runs k CG iterations and m PPCG outer iterations (each with the halo
    updates the reference's driver issues), sized to about `budget_s`
    seconds.  Seconds per timestep = init + cg_steps * t_cg + outer * t_outer
    with the iteration counts of the timed GPU steps (the per-step counts the
    parity tests show the oracle reproduces)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    info = cpu_info()
    threads = threads or info["threads"]
    L = oracle.lib(2)
    L.tl_oracle_set_reduce_parts(threads)
    L.tl_oracle_set_threads(threads)
    ch = tealeaf_hip.ExtChunk(L, 2, n, n)
    st = decks.states(deck)
    ch.initialise_chunk(st["grid_xmin"], st["grid_ymin"], st["dx"], st["dy"])
    ch.generate_chunk(density=st["density"], energy=st["energy"], xmin=st["xmin"], xmax=st["xmax"],
                      ymin=st["ymin"], ymax=st["ymax"], radius=st["radius"], geometry=st["geometry"])
    ch.update_halo([1, 2, 3], 2)
    ch.set_field_kernel()
    t0 = time.perf_counter()
    ch.update_halo([3, 1], 2)
    rro = ch.cg_solver_init(1, False, 0.004)
    ch.update_halo([4, 5], 1)
    ch.copy_u()
    t_init = time.perf_counter() - t0

    def cg_iteration(rro):
        pw = ch.cg_calc_w()
        rrn = ch.cg_calc_ur(rro / pw)
        ch.cg_calc_p(rrn / rro)
        ch.update_halo([4, 5], 1)
        return rrn

    t0 = time.perf_counter()
    rro = cg_iteration(rro)
    t_one = time.perf_counter() - t0
    k = max(2, min(60, int(0.4 * budget_s / max(t_one, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(k):
        rro = cg_iteration(rro)
    t_cg = (time.perf_counter() - t0) / k

    a, b, theta = oracle.calc_ch_coefs(0.5, 4000.0, 10)
    ch.ppcg_init(a, b, theta)
    ch.calculate_residual()
    ch.update_halo([4, 5], 1)

    def outer(rro):
        pw = ch.cg_calc_w()
        ch.cg_calc_ur(rro / pw)
        ch.ppcg_init_sd()
        for step in range(1, 11):
            ch.update_halo([6], 1)
            ch.ppcg_inner(step)
        rrn = ch.calculate_2norm(True)
        ch.cg_calc_p(rrn / rro)
        ch.update_halo([4, 5], 1)
        return rrn

    m = max(1, min(20, int(0.5 * budget_s / max(6.0 * t_cg, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(m):
        rro = outer(rro)
    t_outer = (time.perf_counter() - t0) / m
    ch.close()

    per_step = []
    for s in steps:
        n_outer = max(s["cheby_steps"] - 1, 0)
        per_step.append(t_init + s["cg_steps"] * t_cg + n_outer * t_outer)
    return {
        "value": statistics.mean(per_step),
        "unit": "s/timestep",
        "cores": threads,
        "kind": "port",
        "cpu_model": info["model"],
        "physical_cores_available": info["physical_cores"],
        "sample": (f"oracle (FP64 PPCG) on the deck's {n}^2 chunk, {threads} OpenMP threads: {k} CG iterations "
                   f"({t_cg * 1e3:.1f} ms each) and {m} PPCG outer iterations ({t_outer * 1e3:.1f} ms each), "
                   f"scaled by the timed steps' iteration counts"),
        "per_iteration_s": {"init": t_init, "cg": t_cg, "ppcg_outer": t_outer},
    }


def inner_probe(n: int, launches: int) -> dict:
    """A short program for hardware-counter runs of the 3D PPCG inner steps:
    an n^3 chunk through the ext ABI with seeded random sd, r, u and face
    coefficients, then `launches` two-step sweeps (tl_ppcg_inner_pair, as the
    3D driver runs them); the kernel trace of the run times them."""
    import numpy as np

    L = tealeaf_hip.lib(3)
    ch = tealeaf_hip.ExtChunk(L, 3, n, n, n)
    rng = np.random.default_rng(1)
    for name in ("sd", "r", "u"):
        ch.set(name, rng.uniform(-1.0, 1.0, size=ch.shape))
    for name in ("kx", "ky", "kz"):
        ch.set(name, rng.uniform(0.1, 4.0, size=ch.shape))
    ch.ppcg_init([0.3] * 10, [0.3] * 10, 1.0)
    for _ in range(launches):
        if L.tl_ppcg_inner_pair(ch.c, 1) != 0:
            raise RuntimeError(L.tl_last_error().decode())
    ch.get("u")  # waits for the sweeps
    ch.close()
    return {"probe": "ppcg_inner2 (3D two-step sweep)", "n": n, "launches": launches}


def kernel_probe(n: int, launches: int) -> dict:
    """A short program for hardware-counter runs: the 2D n^2 two-rectangle
    chunk set up through the ext ABI (generate, CG init, 40 CG iterations),
    then `launches` + 1 fused PPCG outer iterations (tl_ppcg_outer_fused, 10
    inner steps) and as many mixed-precision ones (tl_ppcg_outer_mixed).
    Prints the mean host time per FP64 launch."""
    import ctypes

    L = tealeaf_hip.lib(2)
    ch = tealeaf_hip.ExtChunk(L, 2, n, n)
    dx = 10.0 / n
    s = dx / 100.0
    ch.initialise_chunk(0.0, 0.0, dx, dx)
    ch.generate_chunk(density=[100.0, 0.1, 0.1], energy=[1e-4, 25.0, 0.1], xmin=[0, s, 1 + s],
                      xmax=[0, 1 - s, 6 - s], ymin=[0, s, 1 + s], ymax=[0, 2 - s, 6 - s], radius=[0, 0, 0],
                      geometry=[0, 1, 1])
    ch.update_halo([1, 2, 3], 2)
    ch.set_field_kernel()
    rro = ch.cg_solver_init(1, False, 0.004)
    ch.copy_u()
    for _ in range(40):
        pw = ch.cg_calc_w()
        rrn = ch.cg_calc_ur(rro / pw)
        ch.cg_calc_p(rrn / rro)
        rro = rrn
    ch.cg_calc_w()
    a = [0.9 - 0.01 * k for k in range(10)]
    b = [1e-3 * (k + 1) for k in range(10)]
    ca = (ctypes.c_double * 10)(*a)
    cb = (ctypes.c_double * 10)(*b)
    out = ctypes.c_double()
    L.tl_ppcg_outer_fused(ch.c, ctypes.c_double(1e-3), ctypes.c_double(2000.0), ca, cb, 10, ctypes.byref(out))
    t0 = time.perf_counter()
    for _ in range(launches):
        rc = L.tl_ppcg_outer_fused(ch.c, ctypes.c_double(1e-3), ctypes.c_double(2000.0), ca, cb, 10,
                                   ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(L.tl_last_error().decode())
    dt = (time.perf_counter() - t0) / launches
    # the same outer iteration with FP32 inner steps
    for _ in range(launches + 1):
        rc = L.tl_ppcg_outer_mixed(ch.c, ctypes.c_double(1e-3), ctypes.c_double(2000.0), ca, cb, 10,
                                   ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(L.tl_last_error().decode())
    ch.close()
    return {"probe": "ppcg_outer, ppcg_outer_mixed", "n": n, "launches": launches, "host_s_per_launch": dt}


# keywords added to every deck run_solve runs (--no-lookahead)
DECK_OPTS: list[str] = []


def run_solve(deck: str, dim: int, warmup: int, timing: bool = True, only: str | None = None,
              max_steps: int = 0) -> dict:
    """Run a deck through the C++ driver; the timed steps follow `warmup`.
    With `timing`, HIP events bracket every launch in the timed steps, or
    only the launches of kernel `only` (two events per launch of every
    kernel add about 3% to a 4096^2 PPCG timestep)."""
    if DECK_OPTS:
        deck = deck.replace("*endtea", "\n".join(DECK_OPTS) + "\n*endtea")
    L = tealeaf_hip.lib(dim)
    L.tl_set_progress(1)
    L.tl_set_kernel_timing(warmup + 1 if timing else 0)
    if L.tl_set_kernel_timing_only(only.encode() if only else None) != 0:
        raise ValueError(f"no kernel named {only}")
    t0 = time.perf_counter()
    try:
        rep = tealeaf_hip.run_deck(deck, dim=dim, max_steps=max_steps)
    finally:
        L.tl_set_kernel_timing(0)
        L.tl_set_kernel_timing_only(None)
    rep["wall_s"] = time.perf_counter() - t0
    rep["timed"] = rep["steps"][warmup:]
    rep["s_per_step"] = statistics.mean(s["solve_s"] for s in rep["timed"])
    return rep


def kernel_table(rep: dict) -> dict:
    cells = rep["cells"]
    return {k: {"launches": v["count"], "avg_us": round(1e3 * v["ms"] / v["count"], 2),
                "gbs": round(v["bytes_per_cell"] * cells / (v["ms"] / v["count"] * 1e-3) / 1e9, 1)}
            for k, v in rep.get("kernels", {}).items()}


def roofline_of(rep: dict, n: int) -> dict | None:
    """Roofline of the kernel with the most time: algorithmic bytes per
    launch over its mean HIP-event launch time."""
    kern = rep.get("kernels", {})
    if not kern:
        return None
    name, k = max(kern.items(), key=lambda kv: kv[1]["ms"])
    avg_s = k["ms"] / k["count"] * 1e-3
    bytes_per_launch = k["bytes_per_cell"] * rep["cells"]
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get(f"{name}@{n}")
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": name,
            "avg_launch_us": round(avg_s * 1e6, 2), "bytes_per_launch": bytes_per_launch}


def fcg_form(ref: dict, fcg: dict) -> dict:
    """The flexible CG with Jacobi-sweep preconditioning (the learned form's
    solver with the network off) against the FP64 PPCG solve of the same
    deck: time per timestep, ratio, per-step field-summary difference from
    the FP64 solve, iterations, true residual, fallbacks."""
    per_step = []
    for a, b in zip(fcg["timed"], ref["timed"]):
        per_step.append({
            "step": a["step"], "iters": a["fcg_iters"], "fp64_iters": b["iters"],
            "fallback": a["fcg_fallback"], "true_rrn": a["exact_error"], "fp64_true_rrn": b["exact_error"],
            "max_summary_rel_err": max(abs(a[k] - b[k]) / abs(b[k]) for k in ("vol", "mass", "ie", "temp")),
        })
    return {
        "s_per_step": round(fcg["s_per_step"], 6),
        "speedup_vs_fp64": round(ref["s_per_step"] / fcg["s_per_step"], 3),
        "max_summary_rel_err": max(p["max_summary_rel_err"] for p in per_step),
        "per_step": per_step,
    }


def mixed_form(ref: dict, mix: dict) -> dict:
    """The mixed-precision PPCG solve of the same deck against the FP64 one:
    time per timestep and its ratio, and per timestep the relative
    difference of every field-summary total from the FP64 solve, the true
    residual ||u0 - A u||^2 the mixed solve ended with, iterations and
    residual replacements."""
    per_step = []
    for a, b in zip(mix["timed"], ref["timed"]):
        per_step.append({
            "step": a["step"],
            "temp_rel_err": abs(a["temp"] - b["temp"]) / abs(b["temp"]),
            "max_summary_rel_err": max(abs(a[k] - b[k]) / abs(b[k]) for k in ("vol", "mass", "ie", "temp")),
            "true_rrn": a["exact_error"], "iters": a["iters"], "fp64_iters": b["iters"],
            "replacements": a["replacements"], "fallback": a["mixed_fallback"],
        })
    return {
        "s_per_step": round(mix["s_per_step"], 6),
        "speedup_vs_fp64": round(ref["s_per_step"] / mix["s_per_step"], 3),
        "max_temp_rel_err": max(p["temp_rel_err"] for p in per_step),
        "per_step": per_step,
        "kernels": kernel_table({"cells": mix["cells"], "kernels": mix.get("kernels_first_step", mix["kernels"])}),
        "roofline": roofline_of(mix, mix["x_cells"]),
    }


def train_learner(args, n_train: int, dt_train: float):
    """The BF16 U-Net trained on (residual, PPCG correction) pairs from FP64
    PPCG solves of seeded random layouts of 8-64 rectangles and circles
    (training seeds 1000..): crops of the iterates of the CG presteps and
    the PPCG outer iterations, each pair's input residual the crop
    operator applied to its correction, the A-norm loss measured after the
    flexible CG's 4 Jacobi sweeps.  n_train/dt_train set the training
    meshes; the network sees the operator per cell, whose face coefficients
    scale with dt/dx^2, so a smaller mesh with dt scaled by the square of
    the size ratio trains it for a larger one."""
    learner = tealeaf_hip.Learner(seed=args.seed + 17)
    t0 = time.perf_counter()
    for k in range(args.train_decks):
        s = 1000 + k
        print(f"bench: training pairs from layout seed {s} ({n_train}^2, dt {dt_train:g})", file=sys.stderr,
              flush=True)
        learner.add_deck(decks.random_shapes_2d(n_train, s, steps=1, dt=dt_train), snapshots=args.snapshots,
                         crops=args.crops, crop=min(args.crop, n_train), seed=s)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = []
    chunks = max(1, args.train_steps // 100)
    crop = min(args.crop, n_train)
    for k in range(chunks):
        lr = 2e-3 * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * k / chunks)))
        losses += list(learner.train(args.train_steps // chunks, batch=max(1, 32 * 64 * 64 // (crop * crop)), lr=lr))
        if k % 4 == 3 or k == chunks - 1:
            print(f"bench: training step {len(losses)}, loss {statistics.mean(losses[-20:]):.4f}", file=sys.stderr,
                  flush=True)
    t_train = time.perf_counter() - t0
    stats = {"pairs": int(learner.pairs), "decks": args.train_decks, "mesh": n_train, "dt": dt_train,
             "crop": crop, "steps": len(losses), "loss_first": float(statistics.mean(losses[:20])),
             "loss_last": float(statistics.mean(losses[-20:])), "data_s": round(t_data, 2),
             "train_s": round(t_train, 2),
             "loss": "A-norm of the error after 4 weighted Jacobi sweeps, relative to the sweeps alone"}
    return learner, stats


def unet_roofline(learner, n: int) -> dict:
    fwd_ms, macs = learner.time_forward(1, n, n, reps=3)
    tflops = 2.0 * macs / (fwd_ms * 1e-3) / 1e12
    fwd_bytes = learner.forward_bytes(1, n, n)
    fwd_gbs = fwd_bytes / (fwd_ms * 1e-3) / 1e9
    return {"bound": "mfma", "achieved": round(tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
            "frac": round(tflops / 2500.0, 5), "forward_ms": round(fwd_ms, 3), "macs_per_forward": macs,
            "dtype": "bf16 (fp32 accumulate)", "bytes_per_forward": fwd_bytes, "hbm_achieved_gbs": round(fwd_gbs, 1),
            "hbm_frac": round(fwd_gbs / HBM_PEAK_GBS, 4)}


def learned_workload(args, n: int) -> dict:
    """The learned-FCG configuration: train the network (train_learner), then
    solve held-out layouts (seeds 0..) with flexible CG (network + 4
    weighted Jacobi sweeps, FP64 PPCG as fallback), with the same flexible
    CG without the network, and with FP64 PPCG.  Returns the bench line's
    fields: `value` is the FP64 PPCG time, `forms.learned_fcg` the learned
    solve against it."""
    learner, training = train_learner(args, args.train_n or n, 0.004 * ((n / (args.train_n or n)) ** 2))
    roof = unet_roofline(learner, n)
    steps = args.warmup + args.steps
    per_seed, fcg_s, ref_s, sweep_s = [], [], [], []
    ref0 = None
    for s in range(args.eval_decks):
        deck = decks.random_shapes_2d(n, s, steps=steps, extra=("tl_check_result",))
        print(f"bench: held-out layout seed {s}: FP64 PPCG", file=sys.stderr, flush=True)
        ref = run_solve(deck, 2, args.warmup, timing=False)
        if ref0 is None:
            ref0 = (deck, ref)
        learner.use()
        print(f"bench: held-out layout seed {s}: learned FCG", file=sys.stderr, flush=True)
        got = run_solve(deck.replace("tl_use_ppcg", f"tl_use_fcg_learned\ntl_fcg_net_every={args.net_every}\ntl_fcg_net_iters={args.net_iters}"
                                        + ("\ntl_fcg_net_restart" if args.net_restart else "")), 2,
                        args.warmup, timing=False)
        learner.use(False)
        print(f"bench: held-out layout seed {s}: flexible CG, sweeps only", file=sys.stderr, flush=True)
        sw = run_solve(deck.replace("tl_use_ppcg", "tl_use_fcg_learned\ntl_fcg_no_network"), 2, args.warmup,
                       timing=False)
        ref_s.append(ref["s_per_step"])
        fcg_s.append(got["s_per_step"])
        sweep_s.append(sw["s_per_step"])
        for a, b, c in zip(got["timed"], ref["timed"], sw["timed"]):
            per_seed.append({
                "seed": s, "step": a["step"],
                "fcg_s": a["solve_s"], "ppcg_s": b["solve_s"], "sweeps_only_s": c["solve_s"],
                "fcg_iters": a["fcg_iters"], "fcg_fallback": a["fcg_fallback"],
                "sweeps_only_iters": c["fcg_iters"], "sweeps_only_fallback": c["fcg_fallback"],
                "ppcg_iters": b["iters"], "true_rrn": a["exact_error"],
                "max_summary_rel_err": max(abs(a[k] - b[k]) / abs(b[k]) for k in ("vol", "mass", "ie", "temp")),
            })
    learner.close()
    return {
        "ref_s": statistics.mean(ref_s), "cells": n * n, "ref0": ref0,
        "learned_fcg": {
            "s_per_step": round(statistics.mean(fcg_s), 6),
            "speedup_vs_fp64": round(statistics.mean(ref_s) / statistics.mean(fcg_s), 3),
            "sweeps_only_s_per_step": round(statistics.mean(sweep_s), 6),
            "gain_vs_sweeps_only": round(statistics.mean(sweep_s) / statistics.mean(fcg_s), 3),
            "net_every": args.net_every,
            "per_step": per_seed,
            "training": training,
            "unet_roofline": roof,
        },
    }


def true_residual_form(ref: dict, tr: dict) -> dict:
    """FP64 PPCG accepted on the true residual (tl_ppcg_true_residual)
    against the reference-faithful FP64 PPCG, which stops on its recurrence
    residual."""
    return {
        "s_per_step": round(tr["s_per_step"], 6),
        "vs_fp64_ppcg": round(ref["s_per_step"] / tr["s_per_step"], 3),
        "per_step": [{"step": a["step"], "iters": a["iters"], "fp64_iters": b["iters"], "true_rrn": a["exact_error"],
                      "fp64_true_rrn": b["exact_error"], "replacements": a["replacements"],
                      "max_summary_rel_err": max(abs(a[k] - b[k]) / abs(b[k]) for k in ("vol", "mass", "ie", "temp"))}
                     for a, b in zip(tr["timed"], ref["timed"])],
    }


def learned_form(args, deck: str, n: int, ref: dict, sweeps: dict | None, warmup: int) -> dict:
    """Learned flexible CG on the benchmark deck: the network is trained first
    (train_learner, meshes of --train-n cells per side with dt scaled to
    the benchmark's face coefficients), then applied every --net-every
    iterations."""
    n_train = args.train_n or min(n, 2048)
    learner, training = train_learner(args, n_train, 0.004 * (n / n_train) ** 2)
    learner.use()
    print("bench: learned flexible CG on the same deck", file=sys.stderr, flush=True)
    got = run_solve(deck.replace("tl_use_ppcg", f"tl_use_fcg_learned\ntl_fcg_net_every={args.net_every}\ntl_fcg_net_iters={args.net_iters}"
                                        + ("\ntl_fcg_net_restart" if args.net_restart else "")), 2,
                    warmup, timing=False)
    learner.use(False)
    roof = unet_roofline(learner, n)
    learner.close()
    per_step = []
    for k, (a, b) in enumerate(zip(got["timed"], ref["timed"])):
        row = {"step": a["step"], "iters": a["fcg_iters"], "fallback": a["fcg_fallback"], "fp64_iters": b["iters"],
               "true_rrn": a["exact_error"], "fp64_true_rrn": b["exact_error"],
               "max_summary_rel_err": max(abs(a[q] - b[q]) / abs(b[q]) for q in ("vol", "mass", "ie", "temp"))}
        if sweeps:
            row["sweeps_only_iters"] = sweeps["per_step"][k]["iters"]
        per_step.append(row)
    out = {
        "s_per_step": round(got["s_per_step"], 6),
        "speedup_vs_fp64": round(ref["s_per_step"] / got["s_per_step"], 3),
        "net_every": args.net_every,
        "net_iters": args.net_iters,
        "net_restart": bool(args.net_restart),
        "per_step": per_step,
        "training": training,
        "unet_roofline": roof,
    }
    if sweeps:
        out["gain_vs_sweeps_only"] = round(sweeps["s_per_step"] / got["s_per_step"], 3)
    return out


METRIC = "solve s/timestep @eps 1e-10 (learned-FCG, mixed PPCG vs FP64 PPCG); HBM GB/s"


def line(value_s: float, args, workload: str, cells: int, **extra) -> dict:
    out = {
        "metric": METRIC, "value": round(value_s, 6), "unit": "s/timestep", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * value_s, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "cells": cells, "parallelism": "single chunk, one GPU",
                   "reductions": "single pass" if args.single_pass else "separate finish launch",
                   "lookahead": not args.no_lookahead},
    }
    out.update(extra)
    return out


def copy_rate() -> float:
    import ctypes
    L = tealeaf_hip.lib(2)
    L.tl_copy_probe.restype = ctypes.c_double
    return round(L.tl_copy_probe(1 << 27, 10), 1)


def c4_comparison(args, n: int, warmup: int, steps: int, kernel_table_run: bool) -> dict:
    """C4: FP64 PPCG against mixed PPCG, learned FCG, the flexible CG without
    the network and FP64 PPCG accepted on the true residual, on the same
    deck and timesteps (`warmup` untimed, then `steps` timed)."""
    deck = decks.random_rects_2d(n, args.seed, steps=warmup + steps, solver="ppcg", presteps=30, inner=10,
                                 max_iters=10000, extra=("tl_check_result",))
    what = (f"2D {n}x{n}, 32 random rectangles (splitmix64 seed {args.seed}), eps 1e-10, dt 0.004; "
            f"FP64 PPCG (30 presteps, 10 inner steps) vs mixed PPCG vs learned FCG")
    print(f"bench: {what}, {warmup} + {steps} timesteps", file=sys.stderr, flush=True)
    first, top = None, "ppcg_outer"
    if kernel_table_run:
        first = run_solve(deck, 2, 0, max_steps=1)
        top = max(first["kernels"].items(), key=lambda kv: kv[1]["ms"])[0]
        print(f"bench: FP64 PPCG, kernel with the most time: {top}", file=sys.stderr, flush=True)
    ref = run_solve(deck, 2, warmup, only=top)
    forms = {}
    print("bench: mixed-precision PPCG", file=sys.stderr, flush=True)
    mix = run_solve(deck.replace("tl_use_ppcg", "tl_use_ppcg_mixed"), 2, warmup, only="ppcg_outer_mixed")
    forms["mixed_ppcg"] = mixed_form(ref, mix)
    print("bench: FP64 PPCG accepted on the true residual", file=sys.stderr, flush=True)
    tr = run_solve(deck.replace("tl_use_ppcg", "tl_use_ppcg\ntl_ppcg_true_residual"), 2, warmup, timing=False)
    forms["fp64_ppcg_true_residual"] = true_residual_form(ref, tr)
    print("bench: flexible CG, Jacobi sweeps only", file=sys.stderr, flush=True)
    sw = run_solve(deck.replace("tl_use_ppcg", "tl_use_fcg_learned\ntl_fcg_no_network"), 2, warmup, timing=False)
    forms["fcg_sweeps_only"] = fcg_form(ref, sw)
    if not args.no_learned:
        forms["learned_fcg"] = learned_form(args, deck, n, ref, forms["fcg_sweeps_only"], warmup)
    t_ref, t_true = ref["s_per_step"], tr["s_per_step"]
    ns = {"target_speedup": 1.5, "cells": n * n, "warmup": warmup, "steps": steps, "fp64_ppcg_s": round(t_ref, 6),
          "fp64_ppcg_true_residual_s": round(t_true, 6)}
    for key in ("mixed_ppcg", "learned_fcg", "fcg_sweeps_only"):
        if key in forms:
            t = forms[key]["s_per_step"]
            ns[key] = {"s": t, "vs_fp64_ppcg": round(t_ref / t, 3), "vs_fp64_ppcg_true_residual": round(t_true / t, 3)}
    ns["met"] = any(ns[k]["vs_fp64_ppcg"] >= 1.5 for k in ("mixed_ppcg", "learned_fcg") if k in ns)
    if "learned_fcg" in forms:
        # what the trained network itself adds: the learned solve against the
        # same flexible CG with the network switched off
        ns["learned_fcg"]["network_gain_vs_sweeps_only"] = forms["learned_fcg"].get("gain_vs_sweeps_only")
    return {"deck": deck, "what": what, "first": first, "ref": ref, "forms": forms, "north_star": ns}


def north_star(args, n: int) -> dict:
    """The c4 workload's line: the C4 comparison over --warmup + --steps
    timesteps, FP64 PPCG as `value`."""
    r = c4_comparison(args, n, args.warmup, args.steps, True)
    deck, what, first, ref, forms, ns = r["deck"], r["what"], r["first"], r["ref"], r["forms"], r["north_star"]
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle sample)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(deck, n, args.cpu_threads, ref["timed"])
    timed = ref["timed"]
    return line(t_ref, args, what, ref["cells"],
                median_s_per_step=round(statistics.median(s["solve_s"] for s in timed), 6),
                iterations=[s["iters"] for s in timed], temp=[s["temp"] for s in timed],
                true_rrn=[s["exact_error"] for s in timed], step1_s=ref["steps"][0]["solve_s"] if args.warmup else None,
                wall_s=round(ref["wall_s"], 3), kernels=kernel_table(first),
                kernels_note="every kernel, HIP events over the first timestep (a run of its own); roofline: "
                             "events around the roofline kernel only, over the timed steps",
                copy_gbs=copy_rate(), roofline=roofline_of(ref, n), north_star=ns, forms=forms, cpu_baseline=cpu)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed timesteps (default 3 for c4, 9 otherwise)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed timesteps before them")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2 (default): 2D 4096^2 two rectangles, with the c4 comparison attached; c4: 2D 8192^2, "
                         "32 random rectangles (the target's configuration) over --warmup + --steps timesteps; "
                         "c3: 3D 256^3 cube; c5: 2D 4096^2 learned FCG on random rectangles and circles")
    ap.add_argument("--c4-warmup", type=int, default=1, help="c2: untimed 8192^2 timesteps of the attached comparison")
    ap.add_argument("--c4-steps", type=int, default=2, help="c2: timed 8192^2 timesteps of the attached comparison")
    ap.add_argument("--c4-n", type=int, default=8192, help="c2: cells per side of the attached comparison")
    ap.add_argument("--no-c4", action="store_true", help="c2: skip the attached 8192^2 comparison")
    ap.add_argument("--train-decks", type=int, default=8, help="learned form: layouts that give training pairs")
    ap.add_argument("--train-n", type=int, default=None, help="learned form: training mesh (default min(n, 2048))")
    ap.add_argument("--snapshots", type=int, default=16, help="learned form: iterates sampled per solve phase")
    ap.add_argument("--crops", type=int, default=160, help="learned form: pairs cut from each iterate")
    ap.add_argument("--train-steps", type=int, default=1600, help="learned form: Adam steps")
    ap.add_argument("--eval-decks", type=int, default=2, help="c5: held-out layouts")
    ap.add_argument("--net-every", type=int, default=1, help="learned form: network every m-th FCG iteration")
    ap.add_argument("--net-restart", type=int, default=1,
                    help="learned form: 1 starts a new flexible-CG recurrence after the network iterations")
    ap.add_argument("--net-iters", type=int, default=1,
                    help="learned form: network only in the first m FCG iterations (0: throughout)")
    ap.add_argument("--crop", type=int, default=64, help="learned form: training pair size (cells per side)")
    ap.add_argument("--no-learned", action="store_true", help="c4: skip the learned form")
    ap.add_argument("--n", type=int, default=None, help="cells per side (default 8192 for c4, 256 for c3, else 4096)")
    ap.add_argument("--solver", default="ppcg", choices=["ppcg", "cg", "chebyshev", "jacobi", "fcg"],
                    help="c2/c3; fcg: flexible CG with Jacobi-sweep preconditioning on the PPCG deck")
    ap.add_argument("--seed", type=int, default=0, help="c4 layout seed")
    ap.add_argument("--no-mixed", action="store_true", help="c2: skip the mixed-precision PPCG solve")
    ap.add_argument("--no-fcg", action="store_true", help="c2: skip the flexible CG (Jacobi sweeps) solve")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (default: physical cores, "
                                                                 "capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-lookahead", action="store_true",
                    help="launch each CG / PPCG iteration only after reading the previous rrn")
    ap.add_argument("--single-pass", action="store_true",
                    help="close every grid reduction inside its sweep (default: a separate one-workgroup launch)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain copy kernel (counter calibration)")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="only run fused PPCG outer iterations on an n^2 chunk (counter runs)")
    ap.add_argument("--inner-probe", action="store_true",
                    help="only run 3D two-step PPCG inner sweeps on an n^3 chunk (counter runs)")
    ap.add_argument("--unet-probe", action="store_true",
                    help="only time the U-Net forward pass (untrained weights) at 4096^2 and 8192^2, with its 3x3 "
                         "convolutions staged in LDS and reading through L2")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only run the streaming probes of the attainable HBM rate")
    ap.add_argument("--cpu-check", action="store_true",
                    help="only check the CPU baseline's sampled estimate against one fully timed oracle timestep "
                         "(2D n^2 two-rectangle PPCG deck; takes minutes)")
    ap.add_argument("--counter-deck", action="store_true",
                    help="only run the c4 deck capped at 100 iterations (counter runs)")
    args = ap.parse_args()
    n = args.n or {"c2": 4096, "c3": 256, "c4": 8192, "c5": 4096}[args.workload]
    if args.no_lookahead:
        DECK_OPTS.append("tl_no_lookahead")
    if args.single_pass:
        for d in (2, 3):
            tealeaf_hip.lib(d).tl_set_single_pass(1)
    if args.steps is None:
        args.steps = 3 if args.workload == "c4" else 9

    if args.copy_probe:
        L = tealeaf_hip.lib(2)
        L.tl_copy_probe.restype = __import__("ctypes").c_double
        m = 1 << 28  # 2 GiB per buffer
        print(json.dumps({"probe": "copy", "doubles": m, "bytes_each_way": 8 * m,
                          "gbs": L.tl_copy_probe(m, 10)}), flush=True)
        return 0
    if args.kernel_probe:
        print(json.dumps(kernel_probe(n, 20)), flush=True)
        return 0
    if args.inner_probe:
        print(json.dumps(inner_probe(args.n or 256, 9)), flush=True)
        return 0
    if args.unet_probe:
        learner = tealeaf_hip.Learner(seed=3)
        for staged in (1, 0):
            tealeaf_hip.lib(2).tl_learn_conv_staging(staged)
            for m in (4096, 8192):
                roof = unet_roofline(learner, m)
                print(json.dumps({"probe": "unet forward", "n": m, "conv_staging": bool(staged), **roof}), flush=True)
        tealeaf_hip.lib(2).tl_learn_conv_staging(1)
        learner.close()
        return 0
    if args.stream_probe:
        import ctypes
        L = tealeaf_hip.lib(2)
        L.tl_stream_probe.restype = ctypes.c_double
        L.tl_stream_probe.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_int]
        names = {0: "copy, grid-stride (tl_copy_probe)", 1: "copy, one element per thread",
                 2: "copy, nontemporal", 3: "read only", 4: "write only",
                 5: "4 reads + 2 writes (p, w strip mix)", 6: "6 reads + 2 writes (PPCG outer mix)"}
        for kind, name in names.items():
            for m2 in (1 << 28, 1 << 26):
                if kind >= 5 and m2 == 1 << 28:
                    continue  # eight 2 GiB arrays: the 512 MiB streams of the 8192^2 fields suffice
                gbs = L.tl_stream_probe(kind, m2, 10)
                print(json.dumps({"probe": name, "kind": kind, "doubles_per_array": m2, "gbs": round(gbs, 1),
                                  "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 3)}), flush=True)
        return 0
    if args.cpu_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        info = cpu_info()
        threads = args.cpu_threads or info["threads"]
        deck = decks.two_rect_2d(n, "ppcg", steps=1, presteps=30, inner=10, max_iters=10000)
        t0 = time.perf_counter()
        full = oracle.run_deck(deck, dim=2, reduce_parts=threads, threads=threads, max_steps=1)
        wall = time.perf_counter() - t0
        step = full["steps"][0]
        est = cpu_baseline(deck, n, threads, [step])
        print(json.dumps({"probe": "cpu baseline check", "n": n, "threads": threads, "cpu_model": info["model"],
                          "full_step_solve_s": step["solve_s"], "full_run_wall_s": round(wall, 1),
                          "iters": step["iters"], "cg_steps": step["cg_steps"], "cheby_steps": step["cheby_steps"],
                          "sampled_estimate_s": round(est["value"], 3),
                          "ratio_estimate_to_full": round(est["value"] / step["solve_s"], 4),
                          "sample": est["sample"]}), flush=True)
        return 0
    if args.counter_deck:
        # the 8192^2 deck, one timestep capped at 100 iterations, switching to
        # PPCG after 20: the CG strips and the fused PPCG outer kernel, short
        # enough for hardware-counter runs
        deck = decks.random_rects_2d(n, args.seed, steps=1, solver="ppcg", max_iters=100,
                                     extra=("tl_ch_cg_errswitch", "tl_ch_cg_epslim=1.0e30"))
        rep = tealeaf_hip.run_deck(deck, dim=2)
        print(json.dumps({"probe": "counter deck", "n": n, "iters": rep["steps"][0]["iters"]}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solve runs as a single chunk on one MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0:
        print("need --steps >= 1 and --warmup >= 0", file=sys.stderr)
        return 2

    if args.workload == "c4":
        print(json.dumps(north_star(args, n)), flush=True)
        return 0

    if args.workload == "c5":
        r = learned_workload(args, n)
        deck0, ref0 = r["ref0"]
        cpu = None if args.no_cpu_baseline else cpu_baseline(deck0, n, args.cpu_threads, ref0["timed"])
        out = line(r["ref_s"], args,
                   f"2D {n}x{n} random rectangles and circles (held-out seeds 0-{args.eval_decks - 1}), FP64 PPCG "
                   f"vs learned FCG (BF16 U-Net), eps 1e-10, dt 0.004", r["cells"],
                   forms={"learned_fcg": r["learned_fcg"]}, roofline=r["learned_fcg"]["unet_roofline"],
                   cpu_baseline=cpu)
        print(json.dumps(out), flush=True)
        return 0

    total = args.warmup + args.steps
    max_iters = 10000 if args.solver != "jacobi" else 20000
    extra = ("tl_check_result",)
    dim = 3 if args.workload == "c3" else 2
    # fcg: the flexible CG with Jacobi-sweep preconditioning (the learned
    # form with its network off) on the PPCG deck
    solver = "ppcg" if args.solver == "fcg" else args.solver
    if args.workload == "c2":
        deck = decks.two_rect_2d(n, solver, steps=total, presteps=30, inner=10, max_iters=max_iters,
                                 extra=extra)
        what = f"2D {n}x{n} {args.solver.upper()} two-rectangle deck"
    else:
        deck = decks.cube_3d(n, solver, steps=total, presteps=30, inner=10, max_iters=max_iters, extra=extra)
        what = f"3D {n}^3 {args.solver.upper()} hot-box/cuboid deck"
    if args.solver == "fcg":
        if dim != 2:
            print("bench: the flexible CG is 2D", file=sys.stderr)
            return 2
        deck = deck.replace("tl_use_ppcg", "tl_use_fcg_learned\ntl_fcg_no_network")

    print(f"bench: {what}, {args.warmup} + {args.steps} timesteps", file=sys.stderr, flush=True)
    # every kernel timed over the first timestep: the kernel table and the
    # kernel with the most time; the timed steps then bracket only that
    # kernel's launches with events (its roofline), so the timestep itself
    # is measured almost unperturbed
    first = run_solve(deck, dim, 0, max_steps=1)
    top = max(first["kernels"].items(), key=lambda kv: kv[1]["ms"])[0]
    print(f"bench: kernel with the most time: {top}", file=sys.stderr, flush=True)
    ref = run_solve(deck, dim, args.warmup, only=top)
    forms = {}
    if dim == 2 and args.solver == "ppcg" and not args.no_mixed:
        print("bench: mixed-precision PPCG on the same deck", file=sys.stderr, flush=True)
        mdeck = deck.replace("tl_use_ppcg", "tl_use_ppcg_mixed")
        mfirst = run_solve(mdeck, dim, 0, max_steps=1)
        mix = run_solve(mdeck, dim, args.warmup, only="ppcg_outer_mixed")
        mix["kernels_first_step"] = mfirst["kernels"]
        forms["mixed_ppcg"] = mixed_form(ref, mix)
    if dim == 2 and args.solver == "ppcg" and not args.no_fcg:
        # the flexible CG of the learned form with its network switched off:
        # z = 4 weighted Jacobi sweeps on r, convergence on the true residual
        print("bench: flexible CG with Jacobi sweeps on the same deck", file=sys.stderr, flush=True)
        fdeck = deck.replace("tl_use_ppcg", "tl_use_fcg_learned\ntl_fcg_no_network")
        fcg = run_solve(fdeck, dim, args.warmup, timing=False)
        forms["fcg_sweeps"] = fcg_form(ref, fcg)

    cpu = None
    if not args.no_cpu_baseline and args.workload == "c2" and args.solver == "ppcg":
        cpu = cpu_baseline(deck, n, args.cpu_threads, ref["timed"])
    c4 = {}
    if args.workload == "c2" and args.solver == "ppcg" and not args.no_c4 and args.c4_steps > 0:
        # the target's comparison at 8192^2 over a fixed number of timesteps,
        # so that the run's length does not grow with --steps
        r = c4_comparison(args, args.c4_n, args.c4_warmup, args.c4_steps, False)
        c4 = {"north_star": r["north_star"], "forms_c4": r["forms"],
              "c4_workload": r["what"], "c4_iterations": [st["iters"] for st in r["ref"]["timed"]],
              "c4_true_rrn": [st["exact_error"] for st in r["ref"]["timed"]],
              "c4_roofline": roofline_of(r["ref"], args.c4_n)}

    timed = ref["timed"]
    out = line(ref["s_per_step"], args, f"{what}, FP64, eps 1e-10, dt 0.004, presteps 30, inner steps 10",
               ref["cells"], median_s_per_step=round(statistics.median(s["solve_s"] for s in timed), 6),
               iterations=[s["iters"] for s in timed], temp=[s["temp"] for s in timed],
               true_rrn=[s["exact_error"] for s in timed],
               step1_s=ref["steps"][0]["solve_s"] if args.warmup else None, wall_s=round(ref["wall_s"], 3),
               kernels=kernel_table(first),
               kernels_note="every kernel, HIP events over the first timestep (a run of its own); roofline: "
                            "events around the roofline kernel only, over the timed steps",
               copy_gbs=copy_rate(), roofline=roofline_of(ref, n), forms=forms, cpu_baseline=cpu, **c4)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
