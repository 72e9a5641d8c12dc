"""The reference's own 2-D Sod shock-tube run (devel-izem/testcases/sod:
sod.inp, output.cfg, grid.dat, tranfit.out, restarts/restart_0000k.h5,
times_stat.out; the fixtures under tests/golden are copies of those files).

Case: O2/N2 air (2 species), 400 x 11 cells over 1 m x ~0.025 m, viscous,
WENO7 (ng = 4) with JS weights where the Adams-Shariff sensor (PERCENT_WEIGHT
0.15) has fired, linear weights elsewhere, extrapolation in x and periodic
in y, CFL 0.5 without the viscous limits, no chemistry, run on 2 MPI ranks
along x.  Initial state ini_sod (devel-izem/src/module_ICs.f90:2715-2881).
Statistics (and restart files) at TIME_OFFSET + (k-1) DTIME, the step before
each shortened to land on it (izem.f90:254-263, module_input.f90:718-727).

SodRun drives the oracle's restatement of that run (built for 2 species)
from the initial state: step() is one main-loop iteration; run_to_stat(k)
iterates until statistics time k has been written.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
GRID = os.path.join(GOLDEN, "sod_grid.dat")
TRANFIT = os.path.join(GOLDEN, "sod_tranfit.out")
NX, NY, NZ, NG, NDIM = 400, 11, 1, 4, 2
# sod.inp / output.cfg
CFL, CFLMIN, FO, CPR, PERCENT_WEIGHT = 0.5, 0.1, 0.1, 20.0, 0.15
TIME_OFFSET = 7.0968962141250196e-05
DTIME = 7.0968962141250196e-06
NSTAT = 30


def restart_path(k: int) -> str:
    return os.path.join(GOLDEN, f"sod_restart_{k:05d}.h5")


def sod_thd(adi):
    from phd_amd import chemdata as cd
    return cd.build_thd(adi, tranfit_path=TRANFIT)


def sod_box(adi):
    from phd_amd import grid
    return grid.box_from_grid(GRID, NX, NY, NZ, NDIM, adi, ng=NG)


def stat_times(adi) -> np.ndarray:
    """Nondimensional statistics times: TIME_OFFSET, then DTIME added one at
    a time (module_input.f90:722-725), divided by time_ref
    (module_tools.f90:180)."""
    t = [TIME_OFFSET]
    for _ in range(1, NSTAT):
        t.append(t[-1] + DTIME)
    return np.array(t) / adi.time_ref


def initial_state(oracle, adi, thd, box):
    """ini_sod: the left half (x < 0.5 m) at p_ref, rho_ref, the right half at
    0.1 p_ref, 0.125 rho_ref, both air (mole fractions 0.233 O2 / 0.767 N2
    turned into mass fractions by X_to_Y, module_thermodynamics.f90:45-68),
    at rest.  Returns v and the initial temperature guess T (interior)."""
    from phd_amd import grid
    t = oracle.thd_struct(thd)
    xt, _, _ = grid.read_grid(GRID, NX, NY, NZ)
    x = xt * (1.0 / adi.L_ref)
    nrv = len(thd.W)
    Xa = np.array([0.233, 0.767])
    wrk = 0.0
    for l in range(nrv):
        wrk = wrk + Xa[l] * thd.Wc[l]
    wrk = 1.0 / wrk
    Ya = np.array([Xa[l] * thd.Wc[l] * wrk for l in range(nrv)])
    Ws_i = oracle.lib.ora_wmix_i_scalar(C.byref(t), Ya.ctypes.data_as(C.POINTER(C.c_double)))
    states = []
    for P, rho in ((adi.p_ref / adi.p_ref, adi.rho_ref / adi.rho_ref),
                   (0.1 * adi.p_ref / adi.p_ref, 0.125 * adi.rho_ref / adi.rho_ref)):
        Temp = P / (rho * Ws_i)
        has = np.zeros(nrv)
        oracle.lib.ora_ha_scalar(C.byref(t), C.c_double(Temp),
                                 has.ctypes.data_as(C.POINTER(C.c_double)))
        hm = 0.0
        for l in range(nrv):
            hm = hm + has[l] * Ya[l]
        if len(states) == 1:
            rho = P / (Temp * Ws_i)
        e = rho * hm - P + 0.5 * rho * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0)
        states.append((rho, e, Temp))
    v = np.zeros((5 + nrv,) + box.shape)
    T = np.zeros(box.shape)
    I = box.interior()
    for i in range(NX):
        rho, e, Temp = states[0] if x[i] * adi.L_ref < 0.5 else states[1]
        sl = (I[0], I[1], NG + i)
        v[(0,) + sl] = rho
        v[(4,) + sl] = e
        for l in range(nrv):
            v[(5 + l,) + sl] = rho * Ya[l]
        T[sl] = Temp
    return v, T


class SodRun:
    """The oracle restatement of the run (liboracle_nrv2.so)."""

    def __init__(self, adi, oracle=None, restart=None):
        from oracle.oracle import BC_EXTRAPOLATION, Grid, Opts, Oracle, make_adi
        self.o = oracle or Oracle(nrv=2)
        self.adi = adi
        self.thd = sod_thd(adi)
        self.box = box = sod_box(adi)
        self.g = Grid(NX, NY, NZ, NG, NDIM)
        self.g.bc[0] = self.g.bc[1] = BC_EXTRAPOLATION
        self.g.nproc[0] = 2
        self.t = self.o.thd_struct(self.thd)
        self.a = make_adi(adi)
        self.opts = Opts(1, PERCENT_WEIGHT, 1, 0, 0, 1, CFL, CFLMIN, FO, CPR, 0, 0.1)
        self.d_i = [np.ascontiguousarray(m) for m in box.metrics]
        shp = box.shape
        self.W_i = np.zeros(shp)
        self.cp = np.zeros(shp)
        self.ha = np.zeros((2,) + shp)
        self.sensor = np.zeros(shp, np.uint8)
        self.ite = 0
        self.dtmin, self.DTmax = 1.0, 1.0
        self.dt_hist = []
        if restart is None:
            self.v, self.T = initial_state(self.o, adi, self.thd, box)
            self._init_primitives()
            self.times = stat_times(adi)
            self.time, self.nb_stat = 0.0, 0
        else:
            # read_restart_hdf5 (module_io.f90:107-192): the interior from the
            # file, T from zero, the same closing steps as ini_sod (the
            # reader's upd_boundaries repeats the extrapolation), an empty
            # sensor; statistics times from NUMBER_RESTART + 1 copies of
            # TIME_OFFSET (module_input.f90:718-721)
            from phd_amd import restart as rs
            path, nb = restart
            self.v, self.time = rs.read(path, box, [n.strip().upper() for n in self.thd.names])
            self.v = np.ascontiguousarray(self.v)
            self.T = np.zeros(shp)
            self._init_primitives()
            t = [TIME_OFFSET] * (nb + 1)
            while len(t) < NSTAT:
                t.append(t[-1] + DTIME)
            self.times = np.array(t) / adi.time_ref
            self.nb_stat = nb

    def _p(self, a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def _prim_mach(self, dom):
        it = C.c_int()
        st = self.o.lib.ora_prim_inv_var_mach(dom, C.byref(self.g), C.byref(self.t),
                                              self._p(self.v), self._p(self.W_i),
                                              self._p(self.T), self._p(self.cp),
                                              self._p(self.ha), C.byref(it))
        assert st >= 0, st

    def _init_primitives(self):
        """comm_cons, prim_inv_var_err_mach on the interior, upd_extrapolation
        (bc_extrapolation on every face, periodic ones included), then the
        ghost faces (module_ICs.f90:2872-2877, module_BCs.f90:154-174)."""
        L = self.o.lib
        L.ora_comm_cons(C.byref(self.g), self._p(self.v))
        self._prim_mach(0)
        from oracle.oracle import BC_EXTRAPOLATION, Grid
        gx = Grid(NX, NY, NZ, NG, NDIM)
        for f in range(2 * NDIM):
            gx.bc[f] = BC_EXTRAPOLATION
        L.ora_upd_boundaries(C.byref(gx), self._p(self.v))
        for dom in (-1, 1, -2, 2):
            self._prim_mach(dom)

    def step(self):
        """One main-loop iteration (izem.f90:242-282 without chemistry)."""
        L = self.o.lib
        self.ite += 1
        dt_4 = np.zeros(4)
        d = C.c_double(self.dtmin)
        L.ora_timestep(C.byref(self.g), C.byref(self.t), C.byref(self.a), C.byref(self.opts),
                       *[self._p(x) for x in self.d_i], self._p(self.cp), self._p(self.W_i),
                       self._p(self.T), self._p(self.v), C.c_double(self.DTmax),
                       self._p(dt_4), C.byref(d))
        dtmin = d.value
        write = False
        nxt = self.times[self.nb_stat] if self.nb_stat < len(self.times) else np.inf
        if self.time + dtmin > nxt and self.time - dtmin < nxt:
            dtmin = nxt - self.time
            write = True
        st = L.ora_rk3(C.byref(self.g), C.byref(self.t), C.byref(self.a), C.byref(self.opts),
                       C.c_double(dtmin), *[self._p(x) for x in self.d_i], self._p(self.T),
                       self._p(self.W_i), self._p(self.cp), self._p(self.ha),
                       self.sensor.ctypes.data_as(C.POINTER(C.c_uint8)), self._p(self.v), None)
        assert st == 0, st
        self.time += dtmin
        self.dtmin = dtmin
        self.dt_hist.append((dtmin, dt_4.copy()))
        if write:
            self.nb_stat += 1
        return write

    def run_to_stat(self, k: int, progress=None):
        while self.nb_stat < k:
            self.step()
            if progress and self.ite % 100 == 0:
                progress(self)


def read_reference(adi, k: int, thd=None):
    """Interior conserved variables and time of the reference's restart k."""
    from phd_amd import restart
    thd = thd or sod_thd(adi)
    box = sod_box(adi)
    return restart.read(restart_path(k), box, [n.strip().upper() for n in thd.names])


def reference_stats(path=None):
    """Rows of the reference's times_stat.out: statistics number, time [s],
    dtmin [s] and dt_4(1..4) / dtmin (izem.f90:286-300)."""
    path = path or os.path.join(GOLDEN, "sod_times_stat.out")
    rows = [ln.split() for ln in open(path) if ln.strip()]
    return {int(r[0]): [float(x) for x in r[1:]] for r in rows}
