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


THD_PROGRAM = """
program thd_image_check
  use iso_c_binding
  use izem_hip_mod
  implicit none
  integer, parameter :: ld = 29
  real(c_double) :: W(ld), Wc(ld), Wc_i(ld), HeatCap(7, 2, ld), Tmin(ld), Tlim(ld), Tmax(ld)
  real(c_double) :: CofD(4, ld, ld), CofLam(4, ld), CofEta(4, ld), image(IZEM_THD_IMAGE_LEN)
  integer :: c, r, k, l
  integer(c_long) :: n
  do k = 1, ld
    W(k) = 1.0d0 + k
    Wc(k) = 2.0d0 * k
    Wc_i(k) = 3.0d0 * k
    Tmin(k) = 4.0d0 * k
    Tlim(k) = 5.0d0 * k
    Tmax(k) = 6.0d0 * k
    do r = 1, 2
      do c = 1, 7
        HeatCap(c, r, k) = c + 10 * r + 100 * k
      end do
    end do
    do c = 1, 4
      CofLam(c, k) = c + 10 * k
      CofEta(c, k) = -(c + 10 * k)
      do l = 1, ld
        CofD(c, l, k) = c + 10 * l + 1000 * k
      end do
    end do
  end do
  n = izem_thd_image(9_c_int, int(ld, c_int), W, Wc, Wc_i, HeatCap, Tmin, Tlim, Tmax, CofD, &
                     CofLam, CofEta, 0.25d0, 1.0d0, 2.0d0, 3.0d0, 4.0d0, 5.0d0, 6.0d0, image, &
                     IZEM_THD_IMAGE_LEN)
  write (*, '(i0)') n
  write (*, '(es25.17)') image
end program
"""


@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_thd_image_from_declared_arrays():
    """izem_thd_image through the module's interface on arrays declared as
    thd_type declares them (HeatCap(7,2,29), CofD(4,29,29), ...): every
    slot lands where the image layout puts it."""
    import numpy as np
    lib = os.path.join(ROOT, "phd_amd", "libizem_hip.so")
    if not os.path.exists(lib):
        pytest.skip("HIP library not built")
    with tempfile.TemporaryDirectory() as d:
        mod = os.path.join(ROOT, "phd_amd", "fortran", "izem_hip_mod.f90")
        prog = os.path.join(d, "thd_image_check.f90")
        with open(prog, "w") as f:
            f.write(THD_PROGRAM)
        exe = os.path.join(d, "thd_image_check")
        r = subprocess.run([FC, "-J", d, mod, prog, "-o", exe, "-L" + os.path.dirname(lib),
                            "-lizem_hip", "-Wl,-rpath," + os.path.dirname(lib)],
                           capture_output=True, text=True, cwd=d)
        assert r.returncode == 0, r.stdout + r.stderr
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
    vals = out.stdout.split()
    assert int(vals[0]) == 745
    img = np.array([float(x) for x in vals[1:]])
    k1 = np.arange(1, 10)
    W = 1.0 + k1
    assert np.array_equal(img[0:9], W) and np.array_equal(img[18:27], 3.0 * k1)
    heat = img[27:153].reshape(9, 2, 7)
    c, r_, k = np.meshgrid(np.arange(1, 8), np.arange(1, 3), k1, indexing="ij")
    assert np.array_equal(heat, np.transpose(c + 10 * r_ + 100 * k, (2, 1, 0)))
    assert np.array_equal(img[153:162], 4.0 * k1) and np.array_equal(img[171:180], 6.0 * k1)
    cofd = img[180:504].reshape(9, 9, 4)
    l, m, c = np.meshgrid(k1, k1, np.arange(1, 5), indexing="ij")
    assert np.array_equal(cofd, c + 10 * l + 1000 * m)        # [l][m][c] = CofD(c, l, m)
    coflam = img[504:540].reshape(9, 4)
    assert np.array_equal(coflam, np.arange(1, 5)[None, :] + 10 * k1[:, None])
    c1 = img[576:657].reshape(9, 9)
    assert np.allclose(c1, (W[None, :] / W[:, None]) ** 0.25, rtol=1e-15, atol=0)
    assert np.array_equal(img[738:745], [0.25, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0])


def test_fortran_host_library_exports():
    """libizem_fortran.so (built with the library when a Fortran compiler
    exists) exports the host driver's bind(C) entry points."""
    so = os.path.join(ROOT, "phd_amd", "libizem_fortran.so")
    if not os.path.exists(so):
        pytest.skip("libizem_fortran.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    for sym in ("izem_f_create", "izem_f_chem_init", "izem_f_main_iteration"):
        assert f" T {sym}\n" in out, sym
