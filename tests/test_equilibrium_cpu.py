"""The equilibrium solver behind the batched-chemistry cells' product
blend (phd_amd/equilibrium.py), pinned by what thermodynamics fixes
independently of the solver: the law of mass action of two dissociation
reactions, element conservation, complete combustion at low temperature,
the first law at constant pressure, and the adiabatic flame temperature of
stoichiometric H2/air (2380-2400 K with dissociation, e.g. Turns,
An Introduction to Combustion, table of adiabatic flame temperatures)."""
from __future__ import annotations

import numpy as np
import pytest


@pytest.fixture(scope="module")
def m():
    from phd_amd import chemdata as cd
    return cd.read_mechanism()


def _idx(m):
    return {s.upper(): k for k, s in enumerate(m.species)}


@pytest.mark.parametrize("T,p", [(2500.0, 101325.0), (3000.0, 1013250.0)])
def test_mass_action_and_elements(m, T, p):
    from phd_amd import equilibrium as eq
    i = _idx(m)
    n0 = eq.premix_moles(m, 0.8)
    A = eq.element_matrix(m)
    n = eq.equilibrate_TP(m, A @ n0, T, p)
    assert np.all(n >= 0)
    assert np.allclose(A @ n, A @ n0, rtol=1e-11, atol=0)
    x = n / n.sum()
    g = eq.g_RT(m, T)
    P = p / eq.P0
    # H2O = H2 + 1/2 O2 and OH = 1/2 O2 + H (K_p in partial pressures)
    K1 = np.exp(-(g[i["H2"]] + 0.5 * g[i["O2"]] - g[i["H2O"]]))
    assert x[i["H2"]] * x[i["O2"]] ** 0.5 / x[i["H2O"]] * P ** 0.5 == pytest.approx(K1, rel=1e-9)
    K2 = np.exp(-(0.5 * g[i["O2"]] + g[i["H"]] - g[i["OH"]]))
    assert x[i["H"]] * x[i["O2"]] ** 0.5 / x[i["OH"]] * P ** 0.5 == pytest.approx(K2, rel=1e-9)


def test_complete_combustion_at_low_temperature(m):
    from phd_amd import equilibrium as eq
    i = _idx(m)
    n0 = eq.premix_moles(m, 1.0)
    n = eq.equilibrate_TP(m, eq.element_matrix(m) @ n0, 600.0)
    x = n / n.sum()
    assert x[i["H2O"]] == pytest.approx(2.0 / (2.0 + 0.79 / 0.21), rel=1e-8)
    assert x[i["N2"]] == pytest.approx((0.79 / 0.21) / (2.0 + 0.79 / 0.21), rel=1e-8)


def test_adiabatic_flame_temperature(m):
    from phd_amd import equilibrium as eq
    n0 = eq.premix_moles(m, 1.0)
    T, n = eq.equilibrate_HP(m, n0, 300.0)
    assert 2380.0 <= T <= 2400.0
    # the first law to the bisection's 1e-9 relative temperature (c_p T of
    # the products is about 2.6e6 J against an enthalpy of 369 J)
    assert abs(eq.mixture_h(m, n, T) - eq.mixture_h(m, n0, 300.0)) <= 1e-9 * 2.6e6
    # lean: less dissociation, cooler
    T_lean, _ = eq.equilibrate_HP(m, eq.premix_moles(m, 0.5), 300.0)
    assert T_lean < T - 500.0


def test_batched_cells_carry_equilibrium_products(thd):
    from phd_amd import ics
    T, p, Y = ics.config3_cells(thd, 4096, seed=3)
    assert np.allclose(Y.sum(axis=1), 1.0, atol=1e-12)
    names = [n.upper() for n in thd.names]
    # radicals of the dissociated products are present in the blend
    for sp in ("OH", "H", "O", "H2O"):
        assert np.median(Y[:, names.index(sp)]) > 0
    assert np.all(Y >= 0)
