"""Chemical equilibrium of an ideal-gas mixture of the mechanism's species,
for the batched-chemistry workload's product blend.

Equilibrium at (T, p) minimises the Gibbs energy subject to the element
amounts.  With element potentials lambda_e (the Lagrange multipliers), the
minimiser is

    n_k = N exp(-g_k/RT - ln(p/p0) + sum_e lambda_e a_ek),     sum_k n_k = N,

so the unknowns are the element potentials and ln N, found by Newton's
method on the element balances and the mole sum (the element-potential
method of STANJAN, W. C. Reynolds 1986).  The adiabatic state at constant
p and h (the burnt gas of a premix) adds the temperature, found by bisection
on the mixture enthalpy.  Thermodynamic data: the mechanism's NASA-7
polynomials (CHEMKIN form, as CKSMH / CKHMS use them).
"""
from __future__ import annotations

import numpy as np

RU = 8.314510  # J/(mol K), as the CHEMKIN tables
P0 = 101325.0


def _coeffs(mech, T):
    a = np.asarray(mech.a)
    hi = np.asarray(T) > np.asarray(mech.tmid)
    return np.where(hi[:, None], a[:, 1, :], a[:, 0, :])


def h_RT(mech, T):
    c = _coeffs(mech, T)
    return (c[:, 0] + c[:, 1] * T / 2 + c[:, 2] * T ** 2 / 3 + c[:, 3] * T ** 3 / 4
            + c[:, 4] * T ** 4 / 5 + c[:, 5] / T)


def s_R(mech, T):
    c = _coeffs(mech, T)
    return (c[:, 0] * np.log(T) + c[:, 1] * T + c[:, 2] * T ** 2 / 2 + c[:, 3] * T ** 3 / 3
            + c[:, 4] * T ** 4 / 4 + c[:, 6])


def g_RT(mech, T):
    return h_RT(mech, T) - s_R(mech, T)


def element_matrix(mech) -> np.ndarray:
    """a[e][k]: atoms of element e in species k."""
    return np.asarray(mech.ncf, dtype=np.float64).T


def equilibrate_TP(mech, b, T, p=P0, tol=1e-12, itmax=200):
    """Equilibrium mole numbers n_k at (T, p) for element amounts b[e]
    (moles of atoms).  Elements with b = 0 are dropped with their species."""
    A = element_matrix(mech)
    b = np.asarray(b, dtype=np.float64)
    keep_e = b > 0
    keep_k = ~np.any(A[~keep_e] > 0, axis=0)
    A, b = A[keep_e][:, keep_k], b[keep_e]
    mu0 = g_RT(mech, T)[keep_k] + np.log(p / P0)
    ne = len(b)
    # start: the element potentials fitted (least squares) to a guess that
    # spreads the atoms over the species of lowest Gibbs energy per atom
    # (complete combustion, the excess reactant left over)
    n = _guess(A, b, mu0)
    N = n.sum()
    major = n > 1e-3 * N
    lam = np.linalg.lstsq(A[:, major].T, mu0[major] + np.log(n[major] / N), rcond=None)[0]
    lnN = np.log(N)
    for it in range(itmax):
        ln_x = np.clip(-mu0 + lam @ A, -700.0, 700.0)
        x = np.exp(ln_x)
        N = np.exp(lnN)
        n = N * x
        r = np.concatenate([A @ n - b, [x.sum() - 1.0]])
        if np.max(np.abs(r[:-1]) / b) < tol and abs(r[-1]) < tol:
            break
        # Jacobian of (A n - b, sum x - 1) in (lambda, ln N)
        Jm = np.zeros((ne + 1, ne + 1))
        Jm[:ne, :ne] = (A * n) @ A.T
        Jm[:ne, ne] = A @ n
        Jm[ne, :ne] = A @ x
        d = np.linalg.solve(Jm, -r)
        step = 1.0
        big = np.max(np.abs(d))
        if big > 2.0:
            step = 2.0 / big
        lam = lam + step * d[:ne]
        lnN = lnN + step * d[ne]
    else:
        raise RuntimeError("equilibrium did not converge")
    out = np.zeros(len(mech.species))
    out[keep_k] = n
    return out


def _guess(A, b, mu0):
    """Greedy placement of the atoms: repeatedly the species with the lowest
    Gibbs energy per atom among those whose atoms are all still free takes
    as many of them as it can; every species keeps a trace."""
    free = b.astype(np.float64).copy()
    n = np.zeros(A.shape[1])
    order = np.argsort(mu0 / np.maximum(A.sum(axis=0), 1.0))
    for k in order:
        need = A[:, k]
        if not np.any(need > 0):
            continue
        amount = np.min(free[need > 0] / need[need > 0])
        if amount > 0:
            n[k] += amount
            free -= amount * need
    return n + 1e-10 * max(b.sum(), 1.0)


def mixture_h(mech, n, T):
    """Enthalpy of n moles [J]."""
    return float(np.dot(n, h_RT(mech, T)) * RU * T)


def equilibrate_HP(mech, n0, T0, p=P0, Tlo=300.0, Thi=4000.0):
    """Adiabatic equilibrium at constant p from the mixture n0 at T0:
    returns (T, n)."""
    A = element_matrix(mech)
    b = A @ np.asarray(n0, dtype=np.float64)
    h0 = mixture_h(mech, n0, T0)
    f = lambda T: mixture_h(mech, equilibrate_TP(mech, b, T, p), T) - h0
    lo, hi = Tlo, Thi
    flo = f(lo)
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        fm = f(mid)
        if (fm > 0) == (flo > 0):
            lo, flo = mid, fm
        else:
            hi = mid
        if hi - lo < 1e-9 * mid:
            break
    T = 0.5 * (lo + hi)
    return T, equilibrate_TP(mech, b, T, p)


def premix_moles(mech, phi: float) -> np.ndarray:
    """Moles of H2/air at equivalence ratio phi per mole of O2 (air = 0.21 O2
    + 0.79 N2 by mole)."""
    idx = {s.upper(): k for k, s in enumerate(mech.species)}
    n = np.zeros(len(mech.species))
    n[idx["H2"]] = 2.0 * phi
    n[idx["O2"]] = 1.0
    n[idx["N2"]] = 0.79 / 0.21
    return n


def burnt_mass_fractions(mech, phi: float, T0: float = 300.0, p: float = P0) -> np.ndarray:
    """Mass fractions of the adiabatic equilibrium products of the premix at
    phi, burnt at constant p from T0."""
    _, n = equilibrate_HP(mech, premix_moles(mech, phi), T0, p)
    y = n * np.asarray(mech.wt)
    return y / y.sum()
