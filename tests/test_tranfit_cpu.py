"""The transport-fit front end (phd_amd/tranfit.py: kinetic-theory fits of
the kind the reference's TRANFIT, devel-izem/src/tranfit/tranfit.f, writes
to tranfit.out) against

* the reference's own fixture devel-izem/testcases/shockedshearlayer/
  tranfit.out (phd_amd/data/tranfit.out), made by a working build of that
  program for the same nine species: viscosities and binary diffusion
  coefficients do not depend on the thermodynamic data, so every one of
  them must agree; conductivities depend on c_v, and that fixture was made
  with other thermodynamic data, so only the atoms' (c_v = 3/2 R) must;
* the same program built here from its sources (oracle/_ref/tranfit) on the
  bundled mechanism, therm.dat and tran.dat: every entry it computes.  That
  build returns NaN wherever the reduced dipole moment is nonzero (H2O's
  viscosity and conductivity and the H2O-H2O diffusion coefficient, the
  two-dimensional interpolation of the collision integrals), which the
  fixture above pins instead.

Comparisons are of ln(property) at the 50 fit temperatures (300..3500 K):
the cubic coefficients themselves are ill-conditioned in ln T."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_OUT = os.path.join(ROOT, "oracle", "_ref", "tranfit", "tranfit.out")
LNT = np.log(np.linspace(300.0, 3500.0, 50))
V = np.vander(LNT, 4, increasing=True)


@pytest.fixture(scope="module")
def fits():
    from phd_amd import chemdata as cd
    from phd_amd.tranfit import fit_transport
    return fit_transport(cd.read_mechanism())


def _ln(c):
    return c @ V.T


def test_fits_vs_reference_fixture(fits):
    from phd_amd import chemdata as cd
    ref = cd.read_tranfit()
    assert fits.names == ref.names
    assert np.array_equal(fits.kncf, ref.kncf)
    assert np.max(np.abs(fits.W - ref.W) / ref.W) <= 1e-15
    eta = np.max(np.abs(_ln(fits.cofeta) - _ln(ref.cofeta)))
    dif = np.max(np.abs(np.einsum("jkn,tn->jkt", fits.cofd - ref.cofd, V)))
    atoms = [k for k, s in enumerate(fits.names) if s in ("H", "O")]
    lam = np.max(np.abs(_ln(fits.coflam[atoms]) - _ln(ref.coflam[atoms])))
    print(f"ln eta {eta:.1e}, ln D {dif:.1e}, ln lambda (atoms) {lam:.1e}")
    assert eta <= 1e-12 and dif <= 1e-12 and lam <= 1e-12


@pytest.mark.skipif(not os.path.exists(REF_OUT), reason="oracle/_ref/tranfit not built")
def test_fits_vs_reference_program(fits):
    from phd_amd import chemdata as cd
    ref = cd.read_tranfit(REF_OUT)
    iw = fits.names.index("H2O")
    for name, a, b in (("lambda", fits.coflam, ref.coflam), ("eta", fits.cofeta, ref.cofeta)):
        bad = [k for k in range(len(fits.names)) if not np.all(np.isfinite(b[k]))]
        assert bad in ([], [iw]), (name, bad)
        ok = [k for k in range(len(fits.names)) if k not in bad]
        err = np.max(np.abs(_ln(a[ok]) - _ln(b[ok])))
        print(f"ln {name}: {err:.1e}")
        assert err <= 1e-12
    d = np.einsum("jkn,tn->jkt", fits.cofd - ref.cofd, V)
    fin = np.isfinite(d)
    assert np.all(fin | ((np.arange(9)[:, None, None] == iw) & (np.arange(9)[None, :, None] == iw)))
    print(f"ln D: {np.max(np.abs(d[fin])):.1e}")
    assert np.max(np.abs(d[fin])) <= 1e-12


def test_collision_integrals_at_table_nodes():
    """At a tabulated reduced temperature and dipole moment the interpolation
    returns the Monchick-Mason table entry; above T* = 75 (non-polar only)
    the cubics take over continuously, and a polar pair beyond it is
    refused."""
    from phd_amd import tranfit as tf
    Om = tf.collision_integral
    assert abs(Om(1, tf.T_STAR[9], 0.0) - tf.TABLES[1][9, 0]) <= 1e-15
    assert abs(Om(2, tf.T_STAR[19], 0.0) - tf.TABLES[2][19, 0]) <= 1e-15
    assert abs(Om(2, tf.T_STAR[14], tf.DELTA_STAR[4]) - tf.TABLES[2][14, 4]) <= 1e-14
    # between nodes the value lies between its neighbours (monotone columns)
    for i in (3, 12, 25):
        x = 0.5 * (tf.T_STAR[i] + tf.T_STAR[i + 1])
        lo, hi = sorted((tf.TABLES[1][i, 0], tf.TABLES[1][i + 1, 0]))
        assert lo <= Om(1, x, 0.0) <= hi
    # Omega(2,2)* / Omega(1,1)* = A* lies near 1.1 for the LJ potential
    for tr in (0.5, 1.0, 3.0, 10.0, 50.0):
        a = Om(2, tr, 0.0) / Om(1, tr, 0.0)
        assert 1.0 < a < 1.2
    assert abs(Om(1, 80.0, 0.0) - Om(1, 75.0, 0.0)) < 0.02
    with pytest.raises(ValueError):
        Om(1, 100.0, 0.5)


def test_fits_round_trip_and_thd(fits, tmp_path):
    """write_tranfit writes what read_tranfit and build_thd read."""
    from phd_amd import chemdata as cd
    from phd_amd.tranfit import write_tranfit
    p = tmp_path / "tranfit.out"
    write_tranfit(fits, str(p), nreac=21)
    back = cd.read_tranfit(str(p))
    for name in ("cofd", "coflam", "cofeta", "W", "awt", "kncf"):
        assert np.array_equal(getattr(back, name), getattr(fits, name)), name
    adi = cd.Adim()
    thd = cd.build_thd(adi, fits=fits)
    thd_file = cd.build_thd(adi, tranfit_path=str(p))
    assert np.array_equal(thd.flat(), thd_file.flat())


def test_physical_values(fits):
    """N2 at 300 K and 1 atm: viscosity 1.78e-4 g/(cm s) and conductivity
    2.6e-4 W/(cm K) within 3 % (tabulated data); the N2-O2 diffusion
    coefficient 0.20 cm^2/s within 5 %."""
    k = fits.names.index("N2")
    o = fits.names.index("O2")
    t = np.array([1.0, np.log(300.0), np.log(300.0) ** 2, np.log(300.0) ** 3])
    mu = np.exp(fits.cofeta[k] @ t)
    lam = np.exp(fits.coflam[k] @ t) * 1e-7  # erg/(cm s K) -> W/(cm K)
    d = np.exp(fits.cofd[k, o] @ t)
    assert abs(mu / 1.78e-4 - 1) < 0.03
    assert abs(lam / 2.6e-4 - 1) < 0.03
    assert abs(d / 0.20 - 1) < 0.05
