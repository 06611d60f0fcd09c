"""The C-ABI libraries load without a GPU and export every symbol their
headers declare (no compute calls here)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADERS = {2: "tealeaf_ext2d.h", 3: "tealeaf_ext3d.h"}


def declared(header: str) -> list[str]:
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(\w+)\s*\(", text)
    keep = [n for n in names if n.startswith(("ext_", "tl_", "timer_c_"))]
    return sorted(set(keep))


@pytest.mark.parametrize("dim", [2, 3])
def test_library_exports_header(hip, dim):
    path = hip.library_path(dim)
    assert os.path.exists(path), f"{path} not built"
    L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    names = declared(HEADERS[dim])
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    L.tl_dim.restype = ctypes.c_int
    assert L.tl_dim() == dim


@pytest.mark.parametrize("dim", [2, 3])
def test_reference_entry_points_present(dim):
    """Every ext_*_ entry point of the reference's extension interface
    (ext_core.cpp, ext_chunk.cpp) is declared in the header."""
    expected = {"ext_init_", "ext_finalise_", "ext_initialise_chunk_", "ext_generate_chunk_",
                "ext_field_summary_kernel_", "ext_pack_message_", "ext_set_field_kernel_",
                "ext_solver_copy_u_", "ext_calculate_residual_", "ext_calculate_2norm_",
                "ext_solver_finalise_", "ext_update_halo_kernel_", "ext_jacobi_kernel_init_",
                "ext_jacobi_kernel_solve_", "ext_cg_solver_init_", "ext_cg_calc_w_", "ext_cg_calc_ur_",
                "ext_cg_calc_p_", "ext_cheby_solver_init_", "ext_cheby_solver_iterate_", "ext_ppcg_init_",
                "ext_ppcg_init_sd_", "ext_ppcg_inner_", "timer_c_"}
    assert expected <= set(declared(HEADERS[dim]))


def test_missing_library_fails_loudly(hip, monkeypatch):
    monkeypatch.setattr(hip, "library_path", lambda dim: "/nonexistent/libtealeaf.so")
    monkeypatch.setattr(hip, "_LIBS", {})
    with pytest.raises(FileNotFoundError):
        hip.lib(2)


@pytest.mark.parametrize("dim", [2, 3])
def test_fortran_edit_golden(hip, dim):
    """The driver's tea.out formatter against a Fortran runtime: every value
    of tests/golden/fortran_format.json (written by flang from exact bit
    patterns through the descriptors tea.out uses) comes out the same.  Host
    code only: no GPU call."""
    import json
    import os
    import struct
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fortran_format.json")))
    bad = []
    for c in golden["cases"]:
        if c["kind"] == "real":
            v = struct.unpack("<d", struct.pack("<Q", int(c["bits"], 16)))[0]
        else:
            v = c["value"]
        got = hip.fortran_edit(c["desc"], v, dim=dim)
        if got != c["out"]:
            bad.append((c["desc"], c["value"], got, c["out"]))
    assert not bad, bad[:10]
    assert len(golden["cases"]) > 250


def test_fortran_edit_rejects_bad_descriptors(hip):
    """Unhandled or out-of-range descriptors are refused, not read past."""
    import ctypes
    L = hip.lib(2)
    buf = ctypes.create_string_buffer(64)
    for desc in (b"", b"x5", b"e16", b"e0.7", b"f300.2", b"e16.41", b"q3"):
        assert L.tl_fortran_edit(desc, 0, 1.0, None, buf, 64) == -1, desc
    assert L.tl_fortran_edit(b"e16.7", 0, 1.0, None, buf, 8) == -2
    assert L.tl_fortran_edit(b"i4", 42, 0.0, None, buf, 64) == 4 and buf.value == b"  42"
