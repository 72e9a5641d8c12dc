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

    def opts(self, vis=1, reaction=0, weno_avg=1, mrw=-1.0, forcesum=1, CPR=20.0,
             sensor_type=0, ducros_treduc=0.5):
        return OOpts(weno_avg, mrw, vis, 1, reaction, forcesum, 0.5, 0.1, 0.1, CPR, sensor_type,
                     ducros_treduc)


def oracle_reaction_dvode(run: OracleRun, rc, adi, dt_nondim: float, rtol: float, atol: float,
                          trim: bool = True) -> float:
    """reaction_dvode (devel-izem/src/module_reaction.f90:102-304) on the
    oracle state with the reference's own compiled PSR_CHEMKIN/DVODE per
    cell: the reacting box trimmed by ng+1 planes at the upper faces
    (:138-150), CGS conversion, T updated from the reactor but not rho e
    (:278), Y clipped with the last species closing the sum (:285-293);
    then upd_prim_var_woT, comm_cons and the ghost Newton (:73-83).
    Returns DTmax [K]."""
    box = run.box
    ng, nx, ny, nz = box.ng, box.nx, box.ny, box.nz
    dt_phys = dt_nondim * adi.time_ref
    rho_cgs = adi.rho_ref * 1.0e-3
    DTmax = 0.0
    jr = range(ng, ny - ng - 1) if trim else range(0, ny)
    ir = range(ng, nx - ng - 1) if trim else range(0, nx)
    if box.ndim == 3:
        kr = range(ng, min(ny - 1, nz - ng - 2) + 1) if trim else range(0, nz)
    else:
        kr = range(0, 1)
    for k in kr:
        for j in jr:
            for i in ir:
                c = (k + ng, j + ng, i + ng)
                rho0 = run.v[(0,) + c]
                Ya = run.v[(slice(5, 14),) + c] / rho0
                T0 = run.T[c] * adi.T_ref
                Tn, Yn, ist, *_ = rc.psr(T0, rho0 * rho_cgs, Ya, dt_phys, rtol, atol)
                assert ist == 2
                run.T[c] = Tn * (1.0 / adi.T_ref)
                DTmax = max(DTmax, abs(Tn - T0))
                Yc = np.clip(Yn[:8], 0.0, 1.0)
                run.v[(slice(5, 13),) + c] = rho0 * Yc
                run.v[(13,) + c] = rho0 * (1.0 - Yc.sum())
    run.o.lib.ora_upd_prim_var_woT(0, C.byref(run.g), C.byref(run.t), P(run.v), P(run.T),
                                   P(run.W_i), P(run.cp), P(run.ha))
    run.o.lib.ora_comm_cons(C.byref(run.g), P(run.v))
    st = run.o.lib.ora_upd_prim_var_ghost(C.byref(run.g), C.byref(run.t), 1, P(run.v),
                                          P(run.W_i), P(run.T), P(run.cp), P(run.ha))
    assert st == 0
    return DTmax


def negative_T_case(adi, thd, cells, drop: float = 100.0, n: int = 16, twin: bool = False):
    """The 2-D hot-spot state on n^2 with the listed interior cells (i, j)
    spoiled so that their Newton iterate turns negative: N2 and H2 swapped
    (so that their mixture molar mass differs from their neighbours') and
    the energy lowered by drop * rho.  twin: the right x neighbour of each
    cell is made a copy of the left one first (in both states), so that the
    repaired cell equals its neighbours.  Returns box, the clean v, the
    spoiled v, T and the array cells."""
    box, v, T = small_state(adi, thd, "2d", n)
    g = box.ng
    if twin:
        for (i, j) in cells:
            v[:, g, j + g, i + g + 1] = v[:, g, j + g, i + g - 1]
            T[g, j + g, i + g + 1] = T[g, j + g, i + g - 1]
    vb = v.copy()
    names = [nm.upper() for nm in thd.names]
    h2, n2 = 5 + names.index("H2"), 5 + names.index("N2")
    out = []
    for (i, j) in cells:
        c = (g, j + g, i + g)
        vb[(h2,) + c], vb[(n2,) + c] = v[(n2,) + c], v[(h2,) + c]
        vb[(4,) + c] -= vb[(0,) + c] * drop
        out.append(c)
    return box, v, vb, T, out
