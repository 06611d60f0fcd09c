"""The allocation-free entry point tl_cg_calc_w_noalloc against the existing
ext_cg_calc_w_ (tealeaf_2d/ext_core.cpp:417-429) on the same data.

The existing entry point runs on a chunk the library owns; the new one on
caller-owned device arrays (allocated by the test with hipMalloc, in the
reference's flat layout).  Bar: w bitwise equal on the interior (same per-cell operand
order); pw within 1e-13 relative (the two sum the same products in
different orders).  Two calls in a row, the second on a smaller mesh, so the
second runs on a scratch array that still holds the first call's partials.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELD_P = 5
VP = ctypes.c_void_p


def reference_call(hip, dim, n, rng):
    """Random kx, ky[, kz], p on a chunk; p's halo updated; the existing
    entry point's w and pw; the arrays as the reference lays them out."""
    ch = hip.ExtChunk(hip.lib(dim), dim, *n)
    names = ["kx", "ky"] + (["kz"] if dim == 3 else [])
    for name in names:
        ch.set(name, rng.uniform(0.1, 5.0, size=ch.shape))
    ch.set("p", rng.uniform(-1.0, 1.0, size=ch.shape))
    ch.update_halo([FIELD_P], 1)
    pw = ch.cg_calc_w()
    arrays = {name: ch.get(name) for name in names + ["p", "w"]}
    ch.close()
    return arrays, pw


class Hip:
    """hipMalloc / hipMemcpy through the HIP runtime the library itself
    links (its RUNPATH), so the test's buffers live in the same runtime."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.rt = ctypes.CDLL("/opt/rocm-7.2.0/lib/libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(VP), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [VP, VP, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [VP]
        self.bufs = []

    def alloc(self, nbytes):
        ptr = VP()
        assert self.rt.hipMalloc(ctypes.byref(ptr), nbytes) == 0
        self.bufs.append(ptr)
        return ptr

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        ptr = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(ptr, VP(a.ctypes.data), a.nbytes, self.H2D) == 0
        return ptr

    def download(self, ptr, shape):
        out = np.empty(shape, dtype=np.float64)
        assert self.rt.hipMemcpy(VP(out.ctypes.data), ptr, out.nbytes, self.D2H) == 0
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for ptr in self.bufs:
            self.rt.hipFree(ptr)
        self.bufs = []


def noalloc_call(hip, dim, n, arrays):
    L = hip.lib(dim)
    rt = Hip()
    try:
        dev = {k: rt.upload(v) for k, v in arrays.items() if k != "w"}
        w = rt.upload(np.zeros_like(arrays["p"]))
        pw = rt.upload(np.zeros(1))
        if dim == 3:
            L.tl_cg_calc_w_noalloc.argtypes = [ctypes.c_int] * 3 + [VP] * 7
            rc = L.tl_cg_calc_w_noalloc(n[0], n[1], n[2], dev["p"], dev["kx"], dev["ky"], dev["kz"], w, pw, None)
        else:
            L.tl_cg_calc_w_noalloc.argtypes = [ctypes.c_int] * 2 + [VP] * 6
            rc = L.tl_cg_calc_w_noalloc(n[0], n[1], dev["p"], dev["kx"], dev["ky"], w, pw, None)
        assert rc == 0, L.tl_last_error().decode()
        rt.sync()
        return rt.download(w, arrays["p"].shape), float(rt.download(pw, (1,))[0])
    finally:
        rt.free()


@pytest.mark.parametrize("dim,sizes", [(2, [(300, 211, 1), (64, 37, 1)]), (3, [(40, 33, 29), (17, 9, 6)])])
def test_noalloc_matches_cg_calc_w(hip, dim, sizes):
    rng = np.random.default_rng(11 + dim)
    for n in sizes:
        print(f"noalloc dim={dim} mesh={n}", flush=True)
        arrays, pw_ref = reference_call(hip, dim, n, rng)
        w, pw = noalloc_call(hip, dim, n, arrays)
        sl = (slice(2, -2), slice(2, -2), slice(2, -2)) if dim == 3 else (slice(None), slice(2, -2), slice(2, -2))
        np.testing.assert_array_equal(w[sl], arrays["w"][sl])
        assert abs(pw - pw_ref) <= 1e-13 * abs(pw_ref), (pw, pw_ref)


def test_noalloc_refuses_oversize_and_null(hip):
    L = hip.lib(2)
    L.tl_cg_calc_w_noalloc.argtypes = [ctypes.c_int] * 2 + [VP] * 6
    one = VP(1)  # never dereferenced: the call is refused before any launch
    assert L.tl_cg_calc_w_noalloc(8, 70000, one, one, one, one, one, None) == -1
    assert b"65536" in L.tl_last_error()
    L.tl_clear_error()
    assert L.tl_cg_calc_w_noalloc(8, 8, None, one, one, one, one, None) == -1
    assert b"null" in L.tl_last_error()
    L.tl_clear_error()
