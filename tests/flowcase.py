"""Shared set-up of small flow states for the parity tests: the same numpy
arrays are handed to the oracle and to the HIP library."""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle.oracle import Grid as OGrid
from oracle.oracle import Opts as OOpts
from oracle.oracle import make_adi, make_thd
from phd_amd import ics


def P(a):
    return C.c_void_p(a.ctypes.data)


def small_state(adi, thd, kind: str, n: int, seed: int = 7):
    """kind '2d': config-1 recipe on n x n; '3d': config-2 recipe on n^3,
    both with strong (hot-spot / turbulence) structure; '2ds': the 2-D state
    on a stretched grid, its metric arrays computed from the coordinates
    by phd_amd.grid as the reference computes them from grid.dat."""
    if kind == "2ds":
        import dataclasses
        from phd_amd.grid import metric
        box, v, T = ics.config1_state(adi, thd, n=n, seed=seed)
        i = np.arange(n, dtype=np.float64)
        mets = []
        for d in range(2):
            h = box.L[d] / n
            x = h * (i + 0.8 * n / (2 * np.pi) * np.sin(2 * np.pi * i / n) * (0.3 + 0.1 * d))
            mets.append(metric(x) * adi.L_ref)
        mets.append(np.ones(box.nz + 2 * box.ng) * adi.L_ref)
        return dataclasses.replace(box, metrics=tuple(mets)), v, T
    if kind == "2d":
        box, v, T = ics.config1_state(adi, thd, n=n, seed=seed)
    else:
        box, v, T = ics.config2_state(adi, thd, n=n, seed=seed, L=0.001 * n / 16)
    return box, v, T


class OracleRun:
    """Oracle state + helpers in the reference layout."""

    def __init__(self, oracle, adi, thd, box, v, T):
        self.o = oracle
        self.box = box
        self.g = OGrid(box.nx, box.ny, box.nz, box.ng, box.ndim)
        self.t = make_thd(thd)
        self.a = make_adi(adi)
        self.v = np.ascontiguousarray(v.copy())
        self.T = np.ascontiguousarray(T.copy())
        shp = box.shape
        self.W_i = np.zeros(shp)
        self.cp = np.zeros(shp)
        self.ha = np.zeros((9,) + shp)
        self.sensor = np.zeros(shp, np.uint8)
        self.d_i = [np.ascontiguousarray(x) for x in box.inv_spacing(adi)]

    def domains(self):
        b = self.box
        return [0] + [s * q for q in range(1, b.ndim + 1) for s in (-1, 1)]

    def prim_all(self):
        it = C.c_int()
        for d in self.domains():
            s = self.o.lib.ora_prim_inv_var(d, C.byref(self.g), C.byref(self.t), P(self.v),
                                            P(self.W_i), P(self.T), P(self.cp), P(self.ha),
                                            C.byref(it))
            assert s == 0
        return it.value

    def opts(self, vis=1, reaction=0, weno_avg=1, mrw=-1.0, forcesum=1):
        return OOpts(weno_avg, mrw, vis, 1, reaction, forcesum, 0.5, 0.1, 0.1, 20.0)
