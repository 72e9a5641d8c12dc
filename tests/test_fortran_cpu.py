"""The ISO_C_BINDING module compiles with the image's Fortran compiler and
a Fortran program links against the HIP library and calls through it (an
argument-check path that needs no GPU: izem_status and the surrogate entry
points on a null context)."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or shutil.which("flang")

PROGRAM = """
program shim_check
  use iso_c_binding
  use izem_hip_mod
  implicit none
  integer(c_int) :: rc, bits, r1, r2, r3, r4
  type(izem_grid) :: g
  type(c_ptr) :: wt(4), bb(4)
  rc = izem_status(c_null_ptr, bits)
  g = izem_grid(16, 16, 1, 3, 2)
  wt = c_null_ptr
  bb = c_null_ptr
  r1 = izem_surr_features(c_null_ptr, 0, c_null_ptr, c_null_ptr, c_null_ptr, 0.0d0, &
                          c_null_ptr, c_null_ptr, c_null_ptr)
  r2 = izem_mlp_infer(c_null_ptr, 0, c_null_ptr, wt, bb, c_null_ptr)
  r3 = izem_surr_apply(c_null_ptr, 0, c_null_ptr, c_null_ptr, c_null_ptr, 64_c_long, &
                       c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                       6.0, 1.0d0, 1.0d0, c_null_ptr, c_null_ptr)
  r4 = izem_surr_fallback(c_null_ptr, 0, c_null_ptr, c_null_ptr, c_null_ptr, 0.0d0, &
                          c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                          c_null_ptr, c_null_ptr)
  write (*, '(i0,1x,i0,4(1x,i0))') rc, g%nx + g%ng, r1, r2, r3, r4
end program
"""


@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_module_links_and_calls():
    lib = os.path.join(ROOT, "phd_amd", "libizem_hip.so")
    if not os.path.exists(lib):
        pytest.skip("HIP library not built")
    with tempfile.TemporaryDirectory() as d:
        mod = os.path.join(ROOT, "phd_amd", "fortran", "izem_hip_mod.f90")
        prog = os.path.join(d, "shim_check.f90")
        with open(prog, "w") as f:
            f.write(PROGRAM)
        exe = os.path.join(d, "shim_check")
        r = subprocess.run([FC, "-J", d, mod, prog, "-o", exe, "-L" + os.path.dirname(lib),
                            "-lizem_hip", "-Wl,-rpath," + os.path.dirname(lib)],
                           capture_output=True, text=True, cwd=d)
        assert r.returncode == 0, r.stdout + r.stderr
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        # every entry point rejects a null context before touching a device
        assert out.stdout.split() == ["-1", "19", "-1", "-1", "-1", "-1"]
