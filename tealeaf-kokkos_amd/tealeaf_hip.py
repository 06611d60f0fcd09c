"""Python access to the gfx950 TeaLeaf libraries through their C ABI.

The libraries (``libtealeaf2d_hip.so`` and ``libtealeaf3d_hip.so``, built
beside this file by its Makefile) export the reference's ``ext_*_`` kernel
entry points plus a C++ driver entry point ``tl_run_deck``.  This module
loads them with ctypes; nothing here computes anything on the CPU, and every
call fails loudly when a library is missing.

``ExtChunk`` wraps one chunk of any library exporting the ``ext_*_`` ABI
with the reference's calling conventions (every argument by pointer), so the
same test code can drive the HIP library and, for comparison, the oracle.
"""
from __future__ import annotations

import ctypes
import json
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FIELDS = ["density", "energy0", "energy1", "u", "u0", "p", "r", "mi", "w", "z",
          "kx", "ky", "kz", "sd", "volume"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}

# Fortran-side codes (tealeaf_2d/ext_shared.hpp:10-29).
FIELD_DENSITY, FIELD_ENERGY0, FIELD_ENERGY1, FIELD_U, FIELD_P, FIELD_SD = 1, 2, 3, 4, 5, 6
CONDUCTIVITY, RECIP_CONDUCTIVITY = 1, 2
G_RECT, G_CIRC, G_POINT = 1, 2, 3
EXTERNAL_FACE = -1

_P_INT = ctypes.POINTER(ctypes.c_int)
_P_DBL = ctypes.POINTER(ctypes.c_double)


def build(jobs: int = 8) -> None:
    """Compile both libraries for gfx950 with the package Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}"], check=True)


def library_path(dim: int) -> str:
    return os.path.join(HERE, f"libtealeaf{dim}d_hip.so")


_LIBS: dict[int, ctypes.CDLL] = {}


def parse_report(text: str) -> dict:
    """Load a driver report; C's printf spells non-finite values nan/inf,
    which become JSON's NaN/Infinity."""
    text = re.sub(r"(:\s*)-?nan\b", r"\1NaN", text)
    text = re.sub(r"(:\s*)(-?)inf\b", r"\1\2Infinity", text)
    return json.loads(text)


def lib(dim: int) -> ctypes.CDLL:
    """Load the HIP library for ``dim`` (2 or 3); raises when it is not built."""
    if dim not in (2, 3):
        raise ValueError("dim must be 2 or 3")
    if dim not in _LIBS:
        path = library_path(dim)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run tealeaf_hip.build()")
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        L.tl_run_deck.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
        L.tl_run_deck.restype = ctypes.c_int
        L.tl_driver_error.restype = ctypes.c_char_p
        L.tl_tea_out.restype = ctypes.c_char_p
        L.tl_fortran_edit.argtypes = [ctypes.c_char_p, ctypes.c_long, ctypes.c_double, ctypes.c_char_p,
                                      ctypes.c_char_p, ctypes.c_int]
        L.tl_fortran_edit.restype = ctypes.c_int
        L.tl_last_error.restype = ctypes.c_char_p
        L.tl_copy_probe.argtypes = [ctypes.c_long, ctypes.c_int]
        L.tl_copy_probe.restype = ctypes.c_double
        vp, fp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)
        sigs = {
            "tl_learn_create": (vp, [ctypes.c_ulonglong]),
            "tl_learn_destroy": (None, [vp]),
            "tl_learn_add_deck": (ctypes.c_long, [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_ulonglong]),
            "tl_learn_pairs": (ctypes.c_long, [vp]),
            "tl_learn_add_deck_coarse": (ctypes.c_long, [vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
            "tl_learn_consistent_pairs": (ctypes.c_int, [vp, ctypes.c_int]),
            "tl_learn_loss_smoothing": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double]),
            "tl_learn_sample_phases": (ctypes.c_int, [vp, ctypes.c_int]),
            "tl_learn_train": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp]),
            "tl_learn_use": (ctypes.c_int, [vp]),
            "tl_learn_param_count": (ctypes.c_long, [vp]),
            "tl_learn_get_params": (ctypes.c_int, [vp, fp]),
            "tl_learn_set_params": (ctypes.c_int, [vp, fp]),
            "tl_learn_loss_grad": (ctypes.c_double, [vp, fp, fp, fp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                     fp]),
            "tl_learn_apply": (ctypes.c_int, [vp, fp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "tl_learn_forward_bytes": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "tl_learn_conv_staging": (None, [ctypes.c_int]),
            "tl_reduce_probe": (ctypes.c_int, [ctypes.c_long, ctypes.c_int, dp, dp]),
            "tl_set_single_pass": (None, [ctypes.c_int]),
            "tl_coarse_probe": (ctypes.c_int, [_P_INT, ctypes.c_int, ctypes.c_int, fp, fp, fp, dp]),
            "tl_learn_time_forward": (ctypes.c_double, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                        dp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.tl_set_abort_on_error(0)
        _LIBS[dim] = L
    return _LIBS[dim]


def run_deck(deck: str, dim: int = 2, max_steps: int = 0, json_cap: int = 1 << 22) -> dict:
    """Run a ``tea.in`` deck on the GPU through the C++ driver; returns its JSON report."""
    L = lib(dim)
    buf = ctypes.create_string_buffer(json_cap)
    rc = L.tl_run_deck(deck.encode(), buf, json_cap, int(max_steps))
    if rc != 0:
        raise RuntimeError(f"tl_run_deck failed ({rc}): {L.tl_driver_error().decode()}")
    rep = parse_report(buf.value.decode())
    rep["tea_out"] = L.tl_tea_out().decode()
    return rep


def fortran_edit(desc: str, value, dim: int = 2) -> str:
    """``value`` written through the Fortran edit descriptor ``desc``
    (``aW``, ``iW``, ``fW.D``, ``eW.D``) as the driver writes tea.out."""
    L = lib(dim)
    buf = ctypes.create_string_buffer(256)
    k = desc[0].lower()
    n = L.tl_fortran_edit(desc.encode(), int(value) if k == "i" else 0,
                          float(value) if k in "ef" else 0.0, str(value).encode() if k == "a" else None, buf, 256)
    if n < 0:
        raise ValueError(f"tl_fortran_edit({desc!r}) failed ({n})")
    return buf.value.decode()


def _i(v: int):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(v: float):
    return ctypes.byref(ctypes.c_double(float(v)))


class ExtChunk:
    """One chunk driven through the ``ext_*_`` C ABI of ``libhandle``.

    Arrays are exchanged in the reference's flat layout: shape (Z, Y, X) with
    X = nx + 4 fastest (Z = 1 in 2D).
    """

    def __init__(self, libhandle: ctypes.CDLL, dim: int, nx: int, ny: int, nz: int = 1):
        self.L = libhandle
        self.dim = dim
        self.L.ext_finalise_()
        if dim == 3:
            self.L.ext_init_(_i(nx), _i(ny), _i(nz), _i(0))
        else:
            self.L.ext_init_(_i(nx), _i(ny), _i(0))
        self._check()
        self.chunk = ctypes.c_int(1)
        dims = (ctypes.c_int * 6)()
        if self.L.tl_chunk_dims(ctypes.byref(self.chunk), dims) != 0:
            raise RuntimeError("chunk was not created")
        self.X, self.Y, self.Z = dims[0], dims[1], dims[2]
        self.nx, self.ny, self.nz = dims[3], dims[4], dims[5]
        self.shape = (self.Z, self.Y, self.X)

    def _check(self):
        msg = self.L.tl_last_error()
        if msg:
            self.L.tl_clear_error()
            raise RuntimeError(msg.decode())

    @property
    def c(self):
        return ctypes.byref(self.chunk)

    def close(self):
        self.L.ext_finalise_()

    # ---- fields --------------------------------------------------------
    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        rc = self.L.tl_get_field(self.c, FIELD_ID[name], out.ctypes.data_as(_P_DBL))
        if rc != 0:
            self._check()
            raise RuntimeError(f"tl_get_field({name}) failed")
        return out

    def set(self, name: str, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        rc = self.L.tl_set_field(self.c, FIELD_ID[name], a.ctypes.data_as(_P_DBL))
        if rc != 0:
            self._check()
            raise RuntimeError(f"tl_set_field({name}) failed")

    # ---- ext_* entry points --------------------------------------------
    def initialise_chunk(self, xmin, ymin, dx, dy, zmin=0.0, dz=0.0):
        if self.dim == 3:
            self.L.ext_initialise_chunk_(self.c, _d(xmin), _d(ymin), _d(zmin), _d(dx), _d(dy), _d(dz))
        else:
            self.L.ext_initialise_chunk_(self.c, _d(xmin), _d(ymin), _d(dx), _d(dy))
        self._check()

    def generate_chunk(self, density, energy, xmin, xmax, ymin, ymax, radius, geometry,
                       zmin=None, zmax=None):
        n = len(density)
        arr = lambda v: (ctypes.c_double * n)(*[float(x) for x in v])
        geo = (ctypes.c_int * n)(*[int(x) for x in geometry])
        if self.dim == 3:
            self.L.ext_generate_chunk_(self.c, _i(n), arr(density), arr(energy), arr(xmin), arr(xmax),
                                       arr(ymin), arr(ymax), arr(zmin), arr(zmax), arr(radius), geo,
                                       _i(G_RECT), _i(G_CIRC), _i(G_POINT))
        else:
            self.L.ext_generate_chunk_(self.c, _i(n), arr(density), arr(energy), arr(xmin), arr(xmax),
                                       arr(ymin), arr(ymax), arr(radius), geo,
                                       _i(G_RECT), _i(G_CIRC), _i(G_POINT))
        self._check()

    def set_field_kernel(self):
        self.L.ext_set_field_kernel_(self.c)
        self._check()

    def field_summary(self):
        v = [ctypes.c_double() for _ in range(4)]
        self.L.ext_field_summary_kernel_(self.c, *[ctypes.byref(x) for x in v])
        self._check()
        return tuple(x.value for x in v)

    def update_halo(self, fields, depth, neighbours=None):
        nf = 6 if self.dim == 3 else 4
        nb = (ctypes.c_int * nf)(*(neighbours or [EXTERNAL_FACE] * nf))
        fl = (ctypes.c_int * 6)(*[1 if (k + 1) in fields else 0 for k in range(6)])
        self.L.ext_update_halo_kernel_(self.c, nb, fl, _i(depth))
        self._check()

    def cg_solver_init(self, coefficient, preconditioner, dt):
        rro = ctypes.c_double()
        self.L.ext_cg_solver_init_(self.c, _i(coefficient), _i(1 if preconditioner else 0), _d(dt),
                                   ctypes.byref(rro))
        self._check()
        return rro.value

    def cg_calc_w(self):
        pw = ctypes.c_double()
        self.L.ext_cg_calc_w_(self.c, ctypes.byref(pw))
        self._check()
        return pw.value

    def cg_calc_ur(self, alpha):
        rrn = ctypes.c_double()
        self.L.ext_cg_calc_ur_(self.c, _d(alpha), ctypes.byref(rrn))
        self._check()
        return rrn.value

    def cg_calc_p(self, beta):
        self.L.ext_cg_calc_p_(self.c, _d(beta))
        self._check()

    def cheby_solver_init(self, alphas, betas, theta, preconditioner):
        n = len(alphas)
        a = (ctypes.c_double * n)(*alphas)
        b = (ctypes.c_double * n)(*betas)
        self.L.ext_cheby_solver_init_(self.c, a, b, _d(theta), _i(1 if preconditioner else 0), _i(n))
        self._check()

    def cheby_solver_iterate(self, step):
        self.L.ext_cheby_solver_iterate_(self.c, _i(step))
        self._check()

    def ppcg_init(self, alphas, betas, theta):
        n = len(alphas)
        a = (ctypes.c_double * n)(*alphas)
        b = (ctypes.c_double * n)(*betas)
        self.L.ext_ppcg_init_(self.c, a, b, _d(theta), _i(n))
        self._check()

    def ppcg_init_sd(self):
        self.L.ext_ppcg_init_sd_(self.c)
        self._check()

    def ppcg_inner(self, step):
        self.L.ext_ppcg_inner_(self.c, _i(step))
        self._check()

    def jacobi_init(self, coefficient, dt):
        self.L.ext_jacobi_kernel_init_(self.c, _i(coefficient), _d(dt))
        self._check()

    def jacobi_solve(self):
        err = ctypes.c_double()
        self.L.ext_jacobi_kernel_solve_(self.c, ctypes.byref(err))
        self._check()
        return err.value

    def copy_u(self):
        self.L.ext_solver_copy_u_(self.c)
        self._check()

    def calculate_residual(self):
        self.L.ext_calculate_residual_(self.c)
        self._check()

    def calculate_2norm(self, norm_r: bool):
        v = ctypes.c_double()
        flag = ctypes.c_bool(bool(norm_r))
        self.L.ext_calculate_2norm_(self.c, ctypes.byref(v), ctypes.byref(flag))
        self._check()
        return v.value

    def finalise(self):
        self.L.ext_solver_finalise_(self.c)
        self._check()


class Learner:
    """The learned preconditioner of the 2D library: training pairs from FP64
    PPCG solves, GPU training of the BF16 U-Net, and the network the driver's
    flexible CG (tl_use_fcg_learned) applies once ``use()`` is called."""

    def __init__(self, seed: int = 1):
        self.L = lib(2)
        self.h = self.L.tl_learn_create(seed)
        if not self.h:
            raise RuntimeError(self.L.tl_last_error().decode())

    def _check(self, rc):
        if rc is None or rc < 0:
            raise RuntimeError(self.L.tl_last_error().decode())
        return rc

    def add_deck(self, deck: str, snapshots: int = 12, crops: int = 64, crop: int = 64, seed: int = 0) -> int:
        return self._check(self.L.tl_learn_add_deck(self.h, deck.encode(), snapshots, crops, crop, seed))

    def add_deck_coarse(self, deck: str, snapshots: int = 12, nc: int = 256) -> int:
        """One pair per snapshot of the deck's FP64 PPCG solve, each the whole
        nc x nc coarse grid (the network as the two-grid preconditioner's
        coarse solve, deck keyword tl_fcg_coarse)."""
        return self._check(self.L.tl_learn_add_deck_coarse(self.h, deck.encode(), snapshots, nc))

    def consistent_pairs(self, on: bool = True) -> None:
        """Input residual of later pairs: A_c e of the crop (default) or the
        solve's r_k."""
        self._check(self.L.tl_learn_consistent_pairs(self.h, 1 if on else 0))

    def sample_phases(self, cg: bool = True, outer: bool = True) -> None:
        """Solve phases later pairs are sampled from."""
        self._check(self.L.tl_learn_sample_phases(self.h, (1 if cg else 0) | (2 if outer else 0)))

    def loss_smoothing(self, sweeps: int = 4, omega: float = 0.8) -> None:
        """Measure the A-norm loss after `sweeps` weighted Jacobi sweeps."""
        self._check(self.L.tl_learn_loss_smoothing(self.h, sweeps, omega))

    @property
    def pairs(self) -> int:
        return self.L.tl_learn_pairs(self.h)

    def train(self, steps: int, batch: int = 8, lr: float = 1e-3):
        import numpy as np
        losses = np.zeros(steps)
        self._check(self.L.tl_learn_train(self.h, steps, batch, lr,
                                          losses.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return losses

    def use(self, on: bool = True):
        self.L.tl_learn_use(self.h if on else None)

    def params(self):
        import numpy as np
        out = np.zeros(self.L.tl_learn_param_count(self.h), dtype=np.float32)
        self._check(self.L.tl_learn_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.L.tl_learn_set_params(self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def loss_grad(self, r, f, target, kx=None, ky=None):
        """Loss and parameter gradients: the A-norm loss with kx, ky, else MSE."""
        import numpy as np
        n, h, w = r.shape
        fp = ctypes.POINTER(ctypes.c_float)
        planes = [np.ascontiguousarray(a, dtype=np.float32) for a in (r, f, target)]
        ks = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (kx, ky)]
        kp = [None if a is None else a.ctypes.data_as(fp) for a in ks]
        grads = np.zeros(self.L.tl_learn_param_count(self.h), dtype=np.float32)
        loss = self.L.tl_learn_loss_grad(self.h, *[a.ctypes.data_as(fp) for a in planes], kp[0], kp[1], n, h, w,
                                         grads.ctypes.data_as(fp))
        if loss < 0:
            raise RuntimeError(self.L.tl_last_error().decode())
        return loss, grads

    def apply(self, r, f):
        import numpy as np
        n, h, w = r.shape
        fp = ctypes.POINTER(ctypes.c_float)
        planes = [np.ascontiguousarray(a, dtype=np.float32) for a in (r, f)]
        z = np.zeros((n, h, w), dtype=np.float32)
        self._check(self.L.tl_learn_apply(self.h, planes[0].ctypes.data_as(fp), planes[1].ctypes.data_as(fp),
                                          z.ctypes.data_as(fp), n, h, w))
        return z

    def time_forward(self, n: int, h: int, w: int, reps: int = 5):
        """(ms per forward pass, multiply-accumulates per pass)."""
        macs = ctypes.c_double()
        ms = self.L.tl_learn_time_forward(self.h, n, h, w, reps, ctypes.byref(macs))
        if ms < 0:
            raise RuntimeError(self.L.tl_last_error().decode())
        return ms, macs.value

    def forward_bytes(self, n: int, h: int, w: int) -> float:
        """HBM bytes of one forward pass, each tensor read or written once."""
        return self.L.tl_learn_forward_bytes(n, h, w)

    def close(self):
        if self.h:
            self.L.tl_learn_destroy(self.h)
            self.h = None
