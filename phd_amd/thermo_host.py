"""Vectorised host evaluation of the flow solver's species enthalpies, used
only to build initial states (the device kernels do the per-cell thermo).

Follows ha_scalar (devel-izem/src/module_thermodynamics.f90:873-978): NASA-7
enthalpy in the low range below Tpolylim, high range up to Tpolymax, linear
extrapolation outside [Tpolymin, Tpolymax]; nondimensionalised by
Wc_i / Tinf.
"""
from __future__ import annotations

import numpy as np


def _h(c, t):
    return c[0] * t + c[1] * t * t * 0.5 + c[2] * t ** 3 / 3.0 + c[3] * t ** 4 * 0.25 + \
        c[4] * t ** 5 * 0.2 + c[5]


def _dh(c, t):
    return c[0] + c[1] * t + c[2] * t ** 2 + c[3] * t ** 3 + c[4] * t ** 4


def ha_scalar_vec(thd, Tn: np.ndarray) -> np.ndarray:
    TK = Tn * thd.Tinf
    out = np.empty((len(thd.names),) + np.shape(Tn))
    for s in range(len(thd.names)):
        lo, hi = thd.heatcap[s, 0], thd.heatcap[s, 1]
        tmin, tlim, tmax = thd.tpolymin[s], thd.tpolylim[s], thd.tpolymax[s]
        h = np.where(TK < tlim, _h(lo, TK), _h(hi, TK))
        hlo = _dh(lo, tmin) * (TK - tmin) + _h(lo, tmin)
        hhi = _dh(hi, tmax) * (TK - tmax) + _h(hi, tmax)
        h = np.where(TK < tmin, hlo, np.where(TK >= tmax, hhi, h))
        out[s] = thd.Wc_i[s] / thd.Tinf * h
    return out


def x_to_y(thd, X) -> np.ndarray:
    """X_to_Y (module_thermodynamics.f90:45-68): mole to mass fractions, in
    the reference's summation order."""
    wrk = 0.0
    for l in range(len(thd.W)):
        wrk = wrk + float(X[l]) * float(thd.Wc[l])
    wrk = 1.0 / wrk
    return np.array([float(X[l]) * float(thd.Wc[l]) * wrk for l in range(len(thd.W))])


def wmix_i_scalar(thd, Y) -> float:
    """Wmix_i_scalar (module_thermodynamics.f90:739-756): sum_l Y_l Wc_i_l."""
    w = 0.0
    for l in range(len(thd.W)):
        w = w + float(Y[l]) * float(thd.Wc_i[l])
    return w


def ha_scalar(thd, temp: float) -> np.ndarray:
    """ha_scalar (module_thermodynamics.f90:873-978) at one nondimensional
    temperature, in the reference's operation order (powers by repeated
    products, 1/3 as a product), so that a state built from it equals the
    reference's to the last bit."""
    Tinf_i, onethird = 1.0 / thd.Tinf, 1.0 / 3.0

    def h(c, t1):
        t2 = t1 * t1
        t3 = t2 * t1
        t4 = t3 * t1
        t5 = t4 * t1
        return (c[0] * t1 + c[1] * t2 * 0.50 + c[2] * t3 * onethird + c[3] * t4 * 0.25
                + c[4] * t5 * 0.20 + c[5])

    def dh(c, t1):
        t2 = t1 * t1
        t3 = t2 * t1
        t4 = t3 * t1
        return c[0] + c[1] * t1 + c[2] * t2 + c[3] * t3 + c[4] * t4

    out = np.empty(len(thd.W))
    t1 = temp * thd.Tinf
    for s in range(len(thd.W)):
        lo = [float(x) for x in thd.heatcap[s, 0]]
        hi = [float(x) for x in thd.heatcap[s, 1]]
        tmin, tlim, tmax = float(thd.tpolymin[s]), float(thd.tpolylim[s]), float(thd.tpolymax[s])
        fac = float(thd.Wc_i[s]) * Tinf_i
        if t1 < tmin:
            val = fac * (dh(lo, tmin) * (temp * thd.Tinf - tmin) + h(lo, tmin))
        elif t1 < tlim:
            val = fac * h(lo, t1)
        elif t1 < tmax:
            val = fac * h(hi, t1)
        else:
            val = fac * (dh(hi, tmax) * (temp * thd.Tinf - tmax) + h(hi, tmax))
        out[s] = 0.0 + val
    return out


def cp_scalar(thd, temp: float, Y) -> float:
    """cp_scalar (module_thermodynamics.f90:767-862) at one nondimensional
    temperature, in the reference's operation order; cp is linear outside
    [Tpolymin, Tpolymax]."""
    cp = 0.0
    for s in range(len(thd.W)):
        lo = [float(x) for x in thd.heatcap[s, 0]]
        hi = [float(x) for x in thd.heatcap[s, 1]]
        tmin, tlim, tmax = float(thd.tpolymin[s]), float(thd.tpolylim[s]), float(thd.tpolymax[s])
        fac = float(Y[s]) * thd.gam2 * float(thd.Wc_i[s])
        t1 = temp * thd.Tinf

        def poly(c, t1):
            t2 = t1 * t1
            t3 = t2 * t1
            t4 = t3 * t1
            return c[0] + c[1] * t1 + c[2] * t2 + c[3] * t3 + c[4] * t4

        def dpoly(c, t1):
            t2 = t1 * t1
            t3 = t2 * t1
            return c[1] + c[2] * t1 * 2.0 + c[3] * t2 * 3.0 + c[4] * t3 * 4.0

        if t1 < tmin:
            cp = cp + fac * (dpoly(lo, tmin) * (temp * thd.Tinf - tmin) + poly(lo, tmin))
        elif t1 < tlim:
            cp = cp + fac * poly(lo, t1)
        elif t1 < tmax:
            cp = cp + fac * poly(hi, t1)
        else:
            cp = cp + fac * (dpoly(hi, tmax) * (temp * thd.Tinf - tmax) + poly(hi, tmax))
    return cp
