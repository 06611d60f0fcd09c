"""Writes fortran_format.json: values written through the fixed edit
descriptors of TeaLeaf's tea.out lines (e16.7, e15.7, e14.7, e11.4, f11.5,
i3/i4/i7/i8/i11, a11/a13/a16; field_summary.f90:44,81,148,
tea_solve.f90:216-218,535-543,744-745) by a Fortran runtime.

A small Fortran program of our own is generated, compiled with flang (LLVM
Flang under /opt/rocm/lib/llvm/bin), run, and its output parsed; the values
enter the program as exact IEEE-754 bit patterns.  Run from the repository
root: python3 tests/golden/make_fortran_format.py
"""
import json
import math
import os
import struct
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

REALS = [0.0, -0.0, 1.0, -1.0, 100.0, 103.46970921076068, 58.442896702081427, 9.96e-11, -9.96e-11, 1e-10,
         0.99999999, 0.999999949, 0.9999999500000001, 9.9999995e-5, 1.2345675e-3, 1.5e-300, 1.234567890123e300,
         -2.5e-120, 1e100, 9.99999999e99, 0.5, 0.125, 123456.789, 2.0 ** 60, 5e-324, 1.7976931348623157e308,
         3.1415926535897931, -337500.827244273, 16777216.0, 0.00049999999, 12345.678901, 99999.999999,
         -99999.999999, 1e-5, -1e-6, float("nan"), float("inf"), float("-inf")]
EDESC = ["e16.7", "e15.7", "e14.7", "e11.4"]
FDESC = ["f11.5"]
INTS = [0, 1, -1, 7, 99, 100, 999, 1000, -99, -100, 813, 9503, 12345678, 123456789, -1234567, 2147483647]
IDESC = ["i3", "i4", "i7", "i8", "i11"]
STRS = [("a13", "           "), ("a16", "Volume"), ("a16", "Density"), ("a16", "U"), ("a11", "eigmin"),
        ("a11", "est itc"), ("a11", "cn"), ("a6", " step:"), ("a3", "abcdef")]


def bits(x: float) -> int:
    return struct.unpack("<q", struct.pack("<d", x))[0]


def program() -> tuple[str, list]:
    lines = ["program fmt", "  implicit none", "  real(kind=8) :: x", "  integer :: i", "  character(len=32) :: s"]
    cases = []
    for desc in EDESC + FDESC:
        for v in REALS:
            lines.append(f"  x = transfer({bits(v)}_8, 1.0_8)")
            lines.append(f"  write(*,'(a,{desc},a)') '[', x, ']'")
            cases.append({"desc": desc, "kind": "real", "bits": f"{bits(v) & 0xFFFFFFFFFFFFFFFF:016x}",
                          "value": repr(v)})
    for desc in IDESC:
        for v in INTS:
            lines.append(f"  i = {v}")
            lines.append(f"  write(*,'(a,{desc},a)') '[', i, ']'")
            cases.append({"desc": desc, "kind": "int", "value": v})
    for desc, t in STRS:
        lines.append(f"  write(*,'(a,{desc},a)') '[', '{t}', ']'")
        cases.append({"desc": desc, "kind": "str", "value": t})
    lines.append("end program fmt")
    return "\n".join(lines) + "\n", cases


def main() -> int:
    src, cases = program()
    with tempfile.TemporaryDirectory() as tmp:
        f90 = os.path.join(tmp, "fmt.f90")
        exe = os.path.join(tmp, "fmt")
        with open(f90, "w") as fh:
            fh.write(src)
        subprocess.run([FLANG, "-O0", "-o", exe, f90], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    assert len(out) == len(cases), (len(out), len(cases))
    for c, line in zip(cases, out):
        assert line.startswith("[") and line.endswith("]"), line
        c["out"] = line[1:-1]
    ver = subprocess.run([FLANG, "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    with open(os.path.join(HERE, "fortran_format.json"), "w") as fh:
        json.dump({"compiler": ver, "cases": cases}, fh, indent=0)
    print(f"{len(cases)} cases ({ver})")
    return 0


if __name__ == "__main__":
    sys.exit(main())
