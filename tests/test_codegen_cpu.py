"""The generated straight-line kinetics (phd_amd/csrc/mech_gen.h): it is
what the generator writes for the bundled mechanism, and its fused
right-hand side rounds exactly like the split form (gen_kT + gen_rhs) that
the finite-difference Jacobian columns use, so that a column difference
f(y + r e_j) - f(y) carries no rounding gap between the two."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "phd_amd", "csrc")

HARNESS = r"""
#include <math.h>
#define IZ_HD static
#include "mech_gen.h"
extern "C" int nkt() { return GEN_NKT; }
extern "C" void both(double rho, const double *y, double *f_full, double *f_split) {
  double kt[GEN_NKT];
  gen_rhs_full(rho, y, f_full);
  gen_kT(y[0], kt);
  gen_rhs(rho, y, kt, f_split);
}
"""


def test_generated_file_is_current():
    from phd_amd import chemdata as cd
    from phd_amd.codegen_mech import OUT, generate
    with open(OUT) as f:
        assert f.read() == generate(cd.read_mechanism())


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("gen")
    (d / "gen.cpp").write_text(HARNESS)
    so = d / "libgen.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-I", CSRC,
                    "-o", str(so), str(d / "gen.cpp")], check=True)
    L = C.CDLL(str(so))
    L.both.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def test_fused_rhs_rounds_like_split(lib):
    rng = np.random.default_rng(5)
    for _ in range(2000):
        T = rng.uniform(300.0, 3000.0)
        Y = rng.dirichlet(np.ones(9) * 0.3)
        Y[rng.random(9) < 0.2] = 0.0   # absent species
        y = np.concatenate([[T], Y])
        rho = 10.0 ** rng.uniform(-5, -2)
        f1, f2 = np.zeros(10), np.zeros(10)
        lib.both(rho, y.ctypes.data, f1.ctypes.data, f2.ctypes.data)
        assert np.array_equal(f1, f2), (T, rho)
