"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates TeaLeaf's linear-solve path in plain C (see
``tealeaf_oracle.c``); this module loads the two shared objects built from it
(``liboracle2d.so`` / ``liboracle3d.so``) and exposes the same calls the HIP
library exports, so tests can drive both through one interface.

Only the tests, ``__graft_entry__.smoke`` and the CPU-baseline leg of
``bench.py`` use this module.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import json
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))

# Field ids shared by tl_get_field / tl_set_field in both libraries.
FIELDS = ["density", "energy0", "energy1", "u", "u0", "p", "r", "mi", "w", "z",
          "kx", "ky", "kz", "sd", "volume"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}


def build() -> None:
    """Compile the oracle shared objects with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_LIBS: dict[int, ctypes.CDLL] = {}


def parse_report(text: str) -> dict:
    """Load a driver report; C's printf spells non-finite values nan/inf,
    which become JSON's NaN/Infinity."""
    text = re.sub(r"(:\s*)-?nan\b", r"\1NaN", text)
    text = re.sub(r"(:\s*)(-?)inf\b", r"\1\2Infinity", text)
    return json.loads(text)


def lib(dim: int) -> ctypes.CDLL:
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    if dim not in _LIBS:
        path = os.path.join(HERE, f"liboracle{dim}d.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run oracle.build()")
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        L.tl_run_deck.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.tl_run_deck.restype = ctypes.c_int
        L.tl_last_error.restype = ctypes.c_char_p
        L.tl_oracle_set_reduce_parts.argtypes = [ctypes.c_int]
        L.tl_oracle_set_threads.argtypes = [ctypes.c_int]
        L.tl_calc_eigenvalues.argtypes = [ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double)]
        L.tl_calc_eigenvalues.restype = ctypes.c_int
        L.tl_calc_ch_coefs.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
        _LIBS[dim] = L
    return _LIBS[dim]


def run_deck(deck: str, dim: int = 2, reduce_parts: int = 1, threads: int = 0,
             max_steps: int = 0, json_cap: int = 1 << 22) -> dict:
    """Run a ``tea.in`` deck through the oracle's restated driver.

    ``reduce_parts`` is the Kokkos-OpenMP pool size whose partition defines
    the order of every FP64 sum; ``threads`` only sets how many OpenMP
    threads compute (0 keeps the current setting).  ``max_steps`` > 0 stops
    after that many timesteps.
    """
    L = lib(dim)
    L.tl_oracle_set_reduce_parts(int(reduce_parts))
    if threads:
        L.tl_oracle_set_threads(int(threads))
    buf = ctypes.create_string_buffer(json_cap)
    rc = L.tl_run_deck(deck.encode(), buf, json_cap, int(max_steps))
    if rc != 0:
        raise RuntimeError(f"oracle run failed ({rc}): {L.tl_last_error().decode()}")
    return parse_report(buf.value.decode())


def calc_eigenvalues(alphas, betas):
    """tea_calc_eigenvalues on CG alphas/betas (0-based lists of equal length)."""
    L = lib(2)
    m = len(alphas)
    a = (ctypes.c_double * m)(*alphas)
    b = (ctypes.c_double * m)(*betas)
    out = (ctypes.c_double * 2)()
    info = L.tl_calc_eigenvalues(a, b, m, out)
    return info, out[0], out[1]


def calc_ch_coefs(eigmin: float, eigmax: float, n: int):
    L = lib(2)
    a = (ctypes.c_double * n)()
    b = (ctypes.c_double * n)()
    th = ctypes.c_double()
    L.tl_calc_ch_coefs(eigmin, eigmax, n, a, b, ctypes.byref(th))
    return list(a), list(b), th.value
