"""The reference's double Mach reflection deck (devel-izem/testcases/dmr:
dmr.inp, grid.dat, therm.dat, chem.inp, tran.dat; copies under tests/golden
with the prefix dmr_).

Case: a Mach 4.5 shock in N2 meeting a 36-degree wedge, O2/N2 (2 species),
900 x 511 cells, WENO7 with JS weights where the Adams-Shariff sensor
(PERCENT_WEIGHT 0.15) has fired, inviscid, extrapolated x faces and top
face, the post-shock slip wall below, CFL 0.3, 4 MPI ranks along x; ini_DMR
(module_ICs.f90:265-412).  The reference ships no output for it.

DmrRun drives the oracle's restatement of that run (built for 2 species):
step() is one main-loop iteration; run_to_stat(k) iterates until statistics
time k has been written.  write_deck() lays the case out in a directory for
the run driver, optionally on a coarser copy of the grid.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
NX, NY, NZ, NG, NDIM = 900, 511, 1, 4, 2
NPROC = (4, 1, 1)
CFL, CFLMIN, FO, CPR, PERCENT_WEIGHT = 0.3, 0.1, 0.1, 20.0, 0.15
TIME_OFFSET, DTIME = 1.0e-10, 2.0e-08


def coordinates(step_x: int = 1, step_y: int = 1, ny: int | None = None):
    from phd_amd import grid
    xt, yt, zt = grid.read_grid(os.path.join(GOLDEN, "dmr_grid.dat"), NX, NY, NZ)
    yt = yt[::step_y]
    return xt[::step_x], (yt[:ny] if ny else yt), zt


def write_grid(path: str, xt, yt, zt) -> None:
    with open(path, "w") as f:
        for word, t in (("X-DIRECTION", xt), ("Y-DIRECTION", yt), ("Z-DIRECTION", zt)):
            f.write(f"{word}\n\n" + "".join(f"{x:.17f}\n" for x in t) + "\n")
        f.write("END\n")


def write_deck(d: str, nstat: int, coarse=None) -> None:
    """The case in directory d (dmr.inp with NSTAT = nstat); coarse =
    (step_x, step_y, ny) takes every step_x-th x and step_y-th y point."""
    deck = open(os.path.join(GOLDEN, "dmr.inp")).read()
    assert deck.count("NSTAT = 56") == 1
    with open(os.path.join(d, "dmr.inp"), "w") as f:
        f.write(deck.replace("NSTAT = 56", f"NSTAT = {nstat}"))
    for name in ("therm.dat", "chem.inp", "tran.dat"):
        shutil.copy(os.path.join(GOLDEN, f"dmr_{name}"), os.path.join(d, name))
    if coarse:
        write_grid(os.path.join(d, "grid.dat"), *coordinates(*coarse))
    else:
        shutil.copy(os.path.join(GOLDEN, "dmr_grid.dat"), os.path.join(d, "grid.dat"))


def species(adi, d: str = GOLDEN, prefix: str = "dmr_"):
    from phd_amd import chemdata as cd
    from phd_amd import tranfit as tf
    therm = os.path.join(d, prefix + "therm.dat")
    mech = cd.read_mechanism(os.path.join(d, prefix + "chem.inp"), therm)
    return cd.build_thd(adi, therm_path=therm,
                        fits=tf.fit_transport(mech, os.path.join(d, prefix + "tran.dat"), therm))


def stat_times(adi, nstat: int) -> np.ndarray:
    t = [TIME_OFFSET]
    for _ in range(1, nstat):
        t.append(t[-1] + DTIME)
    return np.array(t) / adi.time_ref


class DmrRun:
    """The oracle restatement of the run (liboracle_nrv2.so) on the grid of
    grid_path."""

    def __init__(self, adi, grid_path: str, nstat: int, oracle=None, restart=None,
                 time_offset: float = TIME_OFFSET):
        from oracle.oracle import (BC_EXTRAPOLATION, BC_POSTSHOCK_SLIPWALL, Grid, Opts, Oracle,
                                   make_adi)
        from phd_amd import grid, shock
        self.o = oracle or Oracle(nrv=2)
        self.adi = adi
        self.thd = species(adi)
        nx, ny, nz = grid.count_points(grid_path)
        self.box = box = grid.box_from_grid(grid_path, nx, ny, nz, NDIM, adi, ng=NG)
        xt, yt, _ = grid.read_grid(grid_path, nx, ny, nz)
        self.g = Grid(nx, ny, nz, NG, NDIM)
        for f, b in enumerate((BC_EXTRAPOLATION, BC_EXTRAPOLATION, BC_POSTSHOCK_SLIPWALL,
                               BC_EXTRAPOLATION)):
            self.g.bc[f] = b
        self.g.nproc[0] = NPROC[0]
        self.wall = np.ascontiguousarray(shock.dmr_wall_state(adi, self.thd))
        self.g.nfix = shock.dmr_wall_cells(adi, xt)
        self.g.wall = self.wall.ctypes.data
        self.t = self.o.thd_struct(self.thd)
        self.a = make_adi(adi)
        self.opts = Opts(1, PERCENT_WEIGHT, 0, 0, 0, 1, CFL, CFLMIN, FO, CPR, 0, 0.1)
        self.d_i = [np.ascontiguousarray(m) for m in box.metrics]
        shp = box.shape
        self.W_i = np.zeros(shp)
        self.cp = np.zeros(shp)
        self.ha = np.zeros((2,) + shp)
        self.sensor = np.zeros(shp, np.uint8)
        self.nstat = nstat
        self.ite = 0
        self.dtmin, self.DTmax = 1.0, 1.0
        if restart is None:
            self.v, self.T = shock.dmr_initial_state(adi, self.thd, box, xt, yt)
            self._init_primitives()
            self.times = stat_times(adi, nstat)
            self.time, self.nb_stat = 0.0, 0
        else:
            path, nb = restart
            self._read_restart(path)
            # statistics times of a resumed run (module_input.f90:718-721)
            t = [time_offset] * (nb + 1)
            while len(t) < nstat:
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
        """ini_DMR's end: comm_cons, prim_inv_var_err_mach on the interior,
        upd_extrapolation on every face, then the ghost slabs
        (module_ICs.f90:404-409)."""
        from oracle.oracle import BC_EXTRAPOLATION, Grid
        L = self.o.lib
        L.ora_comm_cons(C.byref(self.g), self._p(self.v))
        self._prim_mach(0)
        gx = Grid(self.g.nx, self.g.ny, self.g.nz, NG, NDIM)
        for f in range(2 * NDIM):
            gx.bc[f] = BC_EXTRAPOLATION
        L.ora_upd_boundaries(C.byref(gx), self._p(self.v))
        for dom in (-1, 1, -2, 2):
            self._prim_mach(dom)

    def _read_restart(self, path: str):
        """read_restart_hdf5 (module_io.f90:107-192): the interior from the
        file, comm_cons, prim_inv_var_err_mach on the interior from T = 0,
        upd_extrapolation, upd_boundaries (the wall's condition and its
        row's primitives), then the ghost slabs."""
        from oracle.oracle import BC_EXTRAPOLATION, Grid
        from phd_amd import restart
        L = self.o.lib
        self.v, self.time = restart.read(path, self.box, self.thd.names)
        self.v = np.ascontiguousarray(self.v)
        self.T = np.zeros(self.box.shape)
        L.ora_comm_cons(C.byref(self.g), self._p(self.v))
        self._prim_mach(0)
        gx = Grid(self.g.nx, self.g.ny, self.g.nz, NG, NDIM)
        for f in range(2 * NDIM):
            gx.bc[f] = BC_EXTRAPOLATION
        L.ora_upd_boundaries(C.byref(gx), self._p(self.v))
        L.ora_upd_boundaries(C.byref(self.g), self._p(self.v))
        st = L.ora_upd_dirichlet(C.byref(self.g), C.byref(self.t), 1, self._p(self.v),
                                 self._p(self.W_i), self._p(self.T), self._p(self.cp),
                                 self._p(self.ha))
        assert st in (0, 1), st
        for dom in (-1, 1, -2, 2):
            self._prim_mach(dom)

    def step(self) -> bool:
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
        nxt = self.times[self.nb_stat] if self.nb_stat < self.nstat else np.inf
        if self.time + dtmin > nxt and self.time - dtmin < nxt:
            dtmin = nxt - self.time
            write = True
        st = L.ora_rk3(C.byref(self.g), C.byref(self.t), C.byref(self.a), C.byref(self.opts),
                       C.c_double(dtmin), *[self._p(x) for x in self.d_i], self._p(self.T),
                       self._p(self.W_i), self._p(self.cp), self._p(self.ha),
                       self.sensor.ctypes.data_as(C.POINTER(C.c_uint8)), self._p(self.v), None)
        # 1: a negative temperature was repaired (the reference warns)
        assert st in (0, 1), st
        self.status = st
        self.time += dtmin
        self.dtmin = dtmin
        if write:
            self.nb_stat += 1
        return write

    def run_to_stat(self, k: int):
        while self.nb_stat < k:
            self.step()
