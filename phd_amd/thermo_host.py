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
