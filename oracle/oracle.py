"""ctypes front end of the oracle.  TEST INFRASTRUCTURE ONLY: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
the product package.

Two independent checkers live here:
  * ``Oracle``  -- liboracle.so, the C restatement of the reference flow path
                   and kinetics (izem_oracle.c);
  * ``RefChem`` -- oracle/_ref/libpsrref.so, the reference's own CHEMKIN-II +
                   DVODE Fortran compiled from its sources, driven through the
                   harness in oracle/ref/psr_driver.f.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NRV, NIV, NV = 9, 5, 14
KMAX, IMAX, TBMAX = 16, 48, 12

_dp = C.POINTER(C.c_double)
_u8 = C.POINTER(C.c_uint8)


BC_PERIODIC, BC_EXTRAPOLATION, BC_POSTSHOCK_SLIPWALL, BC_NSCBC_OUT = 0, 1, 2, 3


class Grid(C.Structure):
    """ora_grid: bc per face W, E, S, N, B, F (zero: periodic), the MPI
    ranks per direction of the run being restated (zero: one), and for a
    post-shock slip wall the cells of the face that take the post-shock
    state and that state (a pointer to NV doubles, kept alive by the
    caller)."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("ndim", C.c_int), ("bc", C.c_int * 6), ("nproc", C.c_int * 3),
                ("nfix", C.c_int), ("wall", C.c_void_p), ("nscbc_Lx", C.c_double),
                ("nscbc_pinf", C.c_double)]


def _thd_struct(nrv: int):
    class Thd(C.Structure):
        _fields_ = [("W", C.c_double * nrv), ("Wc", C.c_double * nrv),
                    ("Wc_i", C.c_double * nrv), ("HeatCap", C.c_double * (nrv * 2 * 7)),
                    ("Tpolymin", C.c_double * nrv), ("Tpolylim", C.c_double * nrv),
                    ("Tpolymax", C.c_double * nrv), ("CofD", C.c_double * (nrv * nrv * 4)),
                    ("CofLam", C.c_double * (nrv * 4)), ("CofEta", C.c_double * (nrv * 4)),
                    ("wilke_c1", C.c_double * (nrv * nrv)),
                    ("wilke_c2", C.c_double * (nrv * nrv)), ("gam2", C.c_double),
                    ("Tinf", C.c_double), ("p0", C.c_double), ("pinf", C.c_double),
                    ("Dinf", C.c_double), ("laminf", C.c_double), ("muinf", C.c_double)]
    return Thd


Thd = _thd_struct(NRV)
_THD = {NRV: Thd, 2: _thd_struct(2)}


class Adi(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("sqgmr", "Sc", "Pr", "ggmopr", "time_ref", "T_ref", "rho_ref", "L_ref")]


class Opts(C.Structure):
    _fields_ = [("weno_avg", C.c_int), ("max_rel_weight", C.c_double), ("vis", C.c_int),
                ("vis_dt", C.c_int), ("reaction", C.c_int), ("forcesum", C.c_int),
                ("CFL", C.c_double), ("CFLmin", C.c_double), ("Fo", C.c_double),
                ("CPR", C.c_double), ("sensor_type", C.c_int), ("ducros_treduc", C.c_double)]


class Mech(C.Structure):
    _fields_ = [("kk", C.c_int), ("ii", C.c_int), ("wt", C.c_double * KMAX),
                ("a", C.c_double * (KMAX * 14)), ("tmid", C.c_double * KMAX),
                ("A", C.c_double * IMAX), ("b", C.c_double * IMAX), ("EoR", C.c_double * IMAX),
                ("nu", C.c_int * (IMAX * 6)), ("nunk", C.c_int * (IMAX * 6)),
                ("reversible", C.c_int * IMAX), ("ntb", C.c_int * IMAX),
                ("tb_k", C.c_int * (IMAX * TBMAX)), ("tb_eff", C.c_double * (IMAX * TBMAX)),
                ("fal", C.c_int * IMAX), ("fpar", C.c_double * (IMAX * 8)),
                ("kfal", C.c_int * IMAX), ("rpar", C.c_double * (IMAX * 3)),
                ("ru", C.c_double), ("patm", C.c_double), ("small_", C.c_double),
                ("big_", C.c_double), ("exparg", C.c_double)]


def _fill(arr, values):
    flat = np.ascontiguousarray(values, dtype=np.float64 if isinstance(arr[0], float) else None).ravel()
    for i, x in enumerate(flat):
        arr[i] = x


def make_thd(thd):
    """The ora_thd struct of a chemdata.Thd (9 or 2 species)."""
    t = _THD[len(thd.W)]()
    for name in ("W", "Wc", "Wc_i"):
        _fill(getattr(t, name), getattr(thd, name))
    _fill(t.HeatCap, thd.heatcap)
    _fill(t.Tpolymin, thd.tpolymin)
    _fill(t.Tpolylim, thd.tpolylim)
    _fill(t.Tpolymax, thd.tpolymax)
    _fill(t.CofD, thd.cofd)
    _fill(t.CofLam, thd.coflam)
    _fill(t.CofEta, thd.cofeta)
    _fill(t.wilke_c1, thd.wilke_c1)
    _fill(t.wilke_c2, thd.wilke_c2)
    for name in ("gam2", "Tinf", "p0", "pinf", "Dinf", "laminf", "muinf"):
        setattr(t, name, float(getattr(thd, name)))
    return t


def make_adi(adi) -> Adi:
    a = Adi()
    for name in ("sqgmr", "Sc", "time_ref", "T_ref", "rho_ref", "L_ref"):
        setattr(a, name, float(getattr(adi, name)))
    a.Pr = adi.pr
    a.ggmopr = adi.ggmopr
    return a


def make_mech(mech) -> Mech:
    t = mech.tables()
    m = Mech()
    m.kk, m.ii = t["kk"], t["ii"]
    for name in ("wt", "a", "tmid", "A", "b", "EoR", "tb_eff", "fpar", "rpar"):
        arr = getattr(m, name)
        for i, x in enumerate(np.asarray(t[name], dtype=np.float64).ravel()):
            arr[i] = float(x)
    for name in ("nu", "nunk", "reversible", "ntb", "tb_k", "fal", "kfal"):
        arr = getattr(m, name)
        for i, x in enumerate(np.asarray(t[name]).ravel()):
            arr[i] = int(x)
    m.ru, m.patm = 8.314510e7, 1.01325e6
    m.small_ = 10.0 ** (-300)
    m.big_ = 10.0 ** (+300)
    m.exparg = float(np.log(m.big_))
    return m


def _p(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    """The C restatement.  Arrays use the Fortran-order layout with ghosts."""

    def __init__(self, lib_path: str | None = None, nrv: int = NRV):
        name = "liboracle.so" if nrv == NRV else f"liboracle_nrv{nrv}.so"
        lib_path = lib_path or os.path.join(HERE, name)
        if not os.path.exists(lib_path):
            subprocess.run(["make", "-C", HERE, name], check=True, stdout=subprocess.DEVNULL)
        self.nrv = nrv
        self.lib = C.CDLL(lib_path)
        L = self.lib
        L.ora_euler_llf.restype = C.c_int
        L.ora_prim_inv_var.restype = C.c_int
        L.ora_rk3.restype = C.c_int
        L.ora_upd_prim_var_ghost.restype = C.c_int
        L.ora_upd_dirichlet.restype = C.c_int
        L.ora_boundaries.restype = C.c_int
        L.ora_wmix_i_scalar.restype = C.c_double
        L.ora_cp_scalar.restype = C.c_double
        L.ora_ckcvbs.restype = C.c_double
        L.ora_cp_scalar.argtypes = [C.c_void_p, C.c_double, _dp]
        L.ora_ha_scalar.argtypes = [C.c_void_p, C.c_double, _dp]
        L.ora_ckcvbs.argtypes = [C.c_void_p, C.c_double, _dp]
        L.ora_ckwyr.argtypes = [C.c_void_p, C.c_double, C.c_double, _dp, _dp]
        L.ora_ckhms.argtypes = [C.c_void_p, C.c_double, _dp]
        L.ora_reactor_rhs.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.ora_transport_cell.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, _dp,
                                         _dp, _dp, _dp]

    def thd_struct(self, thd):
        if len(thd.W) != self.nrv:
            raise ValueError(f"{len(thd.W)} species for an oracle built for {self.nrv}")
        return make_thd(thd)


class RefChem:
    """The reference CHEMKIN-II/DVODE built from its own sources (oracle/_ref).

    The Fortran keeps its state in COMMON blocks, so calls are serialised,
    and each linking file has its own copy of the library (one per directory
    of oracle/_ref: the bundled mechanism in _ref, the front-end test
    mechanism in _ref/sri)."""

    _lock = threading.Lock()
    _libs: dict = {}

    def __init__(self, ref_dir: str | None = None):
        self.dir = os.path.abspath(ref_dir or os.path.join(HERE, "_ref"))
        so = os.path.join(self.dir, "libpsrref.so")
        if not os.path.exists(so):
            raise FileNotFoundError(f"{so} missing: run `make -C oracle ref`")
        with RefChem._lock:
            if self.dir not in RefChem._libs:
                lib = C.CDLL(so)
                cwd = os.getcwd()
                try:
                    os.chdir(self.dir)   # INITCHEMKIN opens 'chem.bin' in the cwd
                    lib.refinit_()
                finally:
                    os.chdir(cwd)
                RefChem._libs[self.dir] = lib
        self.lib = RefChem._libs[self.dir]
        kk, neq = C.c_int(), C.c_int()
        self.lib.refkk_(C.byref(kk), C.byref(neq))
        self.kk = kk.value

    def wdot(self, rho, T, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        w = np.zeros(self.kk)
        with RefChem._lock:
            self.lib.refwdot_(C.byref(C.c_double(rho)), C.byref(C.c_double(T)), _p(Y), _p(w))
        return w

    def hms(self, T):
        h = np.zeros(self.kk)
        with RefChem._lock:
            self.lib.refhms_(C.byref(C.c_double(T)), _p(h))
        return h

    def cvbs(self, T, Y):
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        c = C.c_double()
        with RefChem._lock:
            self.lib.refcvb_(C.byref(C.c_double(T)), _p(Y), C.byref(c))
        return c.value

    def rhs(self, rho, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        zp = np.zeros(self.kk + 1)
        with RefChem._lock:
            self.lib.refrhs_(C.byref(C.c_double(rho)), _p(z), _p(zp))
        return zp

    def psr(self, T, rho, Y, dt, rtol=None, atol=None):
        """One constant-volume reactor substep.  With rtol/atol None the
        reference PSR_CHEMKIN runs with its compiled tolerances (1e-6/1e-10);
        otherwise the REACTION_V loop runs with the given ones.  Returns
        (T, Y, istate, nst, nfe, nje)."""
        Yc = np.ascontiguousarray(Y, dtype=np.float64).copy()
        Tc = C.c_double(T)
        ist = C.c_int(0)
        with RefChem._lock:
            if rtol is None:
                self.lib.refpsr_(C.byref(Tc), C.byref(C.c_double(rho)), C.byref(C.c_double(0.0)),
                                 _p(Yc), C.byref(C.c_double(dt)), C.byref(ist))
                return Tc.value, Yc, ist.value, -1, -1, -1
            nst, nfe, nje = C.c_int(), C.c_int(), C.c_int()
            self.lib.refpsrt_(C.byref(Tc), C.byref(C.c_double(rho)), _p(Yc),
                              C.byref(C.c_double(dt)), C.byref(C.c_double(rtol)),
                              C.byref(C.c_double(atol)), C.byref(ist), C.byref(nst),
                              C.byref(nfe), C.byref(nje))
        return Tc.value, Yc, ist.value, nst.value, nfe.value, nje.value
