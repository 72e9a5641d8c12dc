"""Host build of the chemistry integrator's dense LU (chem_device.h: d_lu,
d_lusolve), checked against the mathematics and against LAPACK.

The kernels compile the same header for the GPU; here g++ compiles it for
the host, one cell at a time (the wave-wide pivot test IZ_ANY reduces to the
lane's own predicate).  The factorisation follows LINPACK's dgefa/dgesl
(chemkin/vode.f:7033, 7630): partial pivoting by the largest magnitude, the
row interchange of step k applied to columns k..N-1 only, and the solve
applying the interchanges interleaved with the forward elimination.  The
diagonal of the factors holds the reciprocal pivots, which the back
substitution multiplies by.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "phd_amd", "csrc")
N = 10  # CK_NEQ: T + 9 species

HARNESS = r"""
#include <math.h>
#define IZ_HD static
#define IZ_MHD
#define IZ_FORCEINLINE inline
#define IZ_ANY(p) (p)
#define IZ_CMECH h_mech
#define IZ_CBDF h_bdf
#include "chem_tables.h"
static IzMech h_mech;
static IzBdf h_bdf;
#include "chem_device.h"
extern "C" int lu_solve(const double *A, const double *b, double *lu, int *piv_out, double *x) {
  double M[N_][N_];
  int piv[N_];
  double xx[N_];
  for (int i = 0; i < N_; i++) {
    for (int j = 0; j < N_; j++) M[i][j] = A[i * N_ + j];
    xx[i] = b[i];
  }
  const bool ok = d_lu(M, piv);
  if (ok) d_lusolve(M, piv, xx);
  for (int i = 0; i < N_; i++) {
    for (int j = 0; j < N_; j++) lu[i * N_ + j] = M[i][j];
    piv_out[i] = piv[i];
    x[i] = xx[i];
  }
  return ok ? 1 : 0;
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("lu")
    src = d / "lu.cpp"
    src.write_text(HARNESS)
    so = d / "liblu.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-I", CSRC,
                    "-o", str(so), str(src)], check=True)
    L = C.CDLL(str(so))
    L.lu_solve.argtypes = [C.c_void_p] * 5
    return L


def run(lib, A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lu = np.zeros((N, N))
    piv = np.zeros(N, dtype=np.int32)
    x = np.zeros(N)
    ok = lib.lu_solve(A.ctypes.data, b.ctypes.data, lu.ctypes.data, piv.ctypes.data,
                      x.ctypes.data)
    return ok, lu, piv, x


def matrices(rng):
    """Random dense matrices, ones whose diagonal is small (every step
    pivots), and iteration matrices I - h l0 J of a stiff Jacobian shape."""
    out = []
    for _ in range(50):
        out.append(rng.standard_normal((N, N)))
    for _ in range(20):
        A = rng.standard_normal((N, N))
        A[np.arange(N), np.arange(N)] *= 1e-6
        out.append(A)
    for _ in range(20):
        J = rng.standard_normal((N, N)) * 10.0 ** rng.uniform(-2, 8, size=(N, 1))
        out.append(np.eye(N) - 10.0 ** rng.uniform(-9, -5) * J)
    out.append(np.eye(N)[::-1].copy())  # anti-diagonal permutation: a swap at every step
    return out


def test_lu_solve_residual_and_lapack(lib):
    rng = np.random.default_rng(7)
    for A in matrices(rng):
        b = rng.standard_normal(N)
        ok, lu, piv, x = run(lib, A, b)
        assert ok == 1
        x_ref = np.linalg.solve(A, b)
        cond = np.linalg.cond(A)
        # backward error of a pivoted LU solve
        r = np.abs(A @ x - b).max() / (np.abs(A).max() * np.abs(x).max() + np.abs(b).max())
        assert r <= 1e-13, (r, cond)
        assert np.abs(x - x_ref).max() <= 1e-12 * cond * np.abs(x_ref).max()
        # same pivot rows and the same U as LAPACK's getrf (which also picks
        # the first largest magnitude); L is stored LINPACK-style, so only
        # U is compared entry by entry, its diagonal from the reciprocals
        lu_ref, piv_ref = sla.lu_factor(A)
        assert np.array_equal(piv, piv_ref)
        U, U_ref = np.triu(lu, 1), np.triu(lu_ref)
        U[np.arange(N), np.arange(N)] = 1.0 / np.diag(lu)
        assert np.abs(U - U_ref).max() <= 1e-12 * np.abs(U_ref).max()


def test_lu_singular_column(lib):
    """A zero column has no pivot: the factorisation reports it (dgefa's
    INFO), which the integrator treats as a corrector failure."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((N, N))
    A[:, 4] = 0.0
    ok, lu, piv, x = run(lib, A, rng.standard_normal(N))
    assert ok == 0
