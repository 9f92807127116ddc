"""The text formats of WriteTecplot (src/SELF_Model2D.f90:1101-1207) on the CPU: the Python
writer's ES16.7E3 numbers and solution.<stamp>.tec time stamps against what the Fortran layer
itself prints for the same values (a small amdflang program over SELF_Model2D's TimeStamp and
the ES16.7E3 edit descriptor).  Skipped where amdflang or the Fortran build is absent."""
import os
import subprocess

import pytest

from self_amd import _es16, time_stamp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/bin/amdflang"
BUILD = os.path.join(ROOT, "_build", "fortran")
FLIB = os.path.join(ROOT, "self_amd", "fortran", "libself_amd_fortran.a")
VALUES = [0.0, -0.0, 1.0, -1.0, 0.5, 1.0 / 3.0, -2.0 / 3.0, 123456.789, -9.87654321e-5, 1.5e-300, 9.99999995e99,
          2.5e-8, 1.23456785, 1.0e100, -7.0e-101]
TIMES = [0.0, 0.25, 1.5, 59.9996, 61.0, 3725.0005, 86399.9999, 90061.25, 1.0e-4]

PROGRAM = """
PROGRAM fmt_check
  USE SELF_Constants
  USE SELF_Model2D, ONLY: TimeStamp
  IMPLICIT NONE
  REAL(prec) :: v(%(nv)d) = [%(values)s]
  REAL(prec) :: t(%(nt)d) = [%(times)s]
  INTEGER :: k
  DO k = 1, SIZE(v)
    WRITE (*, '(ES16.7E3,1X)') v(k)
  END DO
  DO k = 1, SIZE(t)
    WRITE (*, '(A)') TimeStamp(t(k))
  END DO
END PROGRAM fmt_check
"""


def test_es16_and_time_stamp_match_fortran(tmp_path):
    if not (os.path.exists(FC) and os.path.exists(FLIB) and os.path.isdir(BUILD)):
        pytest.skip("amdflang or the Fortran build (make -C self_amd/fortran) is absent")
    src = tmp_path / "fmt_check.f90"
    src.write_text(PROGRAM % {"nv": len(VALUES), "values": ", ".join(f"{v!r}_prec" for v in VALUES),
                              "nt": len(TIMES), "times": ", ".join(f"{t!r}_prec" for t in TIMES)})
    exe = str(tmp_path / "fmt_check")
    r = subprocess.run([FC, "-O0", "-I" + BUILD, str(src), "-o", exe, FLIB, "-L" + os.path.join(ROOT, "self_amd"),
                        "-lself_amd", "-Wl,-rpath," + os.path.join(ROOT, "self_amd")], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip(f"amdflang could not build the check program: {r.stderr[-300:]}")
    out = subprocess.run([exe], capture_output=True, text=True, check=True, timeout=60).stdout.splitlines()
    for v, line in zip(VALUES, out[:len(VALUES)]):
        assert _es16(v) == line, (v, line)  # a trailing 1X writes nothing
    for t, line in zip(TIMES, out[len(VALUES):]):
        assert time_stamp(t) == line, (t, line)
