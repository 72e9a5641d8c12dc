"""Helpers for the tests of a Fortran host: the thd_type arrays in the
memory layout thd_init leaves them in (Fortran order, species dimension
Nspcmax_thd = 29, devel-izem/src/thd/type_thd.f90:58-73 and
param_thd.f90:23), and CHEMKIN-II's work arrays as INITCHEMKIN leaves them
in the reference library's COMMON blocks (chemkin/chemkin.common:57,
ckstrt.common:5)."""
from __future__ import annotations

import ctypes as C

import numpy as np

NSPCMAX_THD = 29
LENIWK = 4000           # chemkin.common: IWORK(LENIWK), RWORK(LENRWK)
CKSTRT_LEN = 85         # integers in /CKSTRT/


def thd_arrays(thd, ld: int = NSPCMAX_THD, fill: float = np.nan) -> dict:
    """Fortran-ordered arrays of thd_type filled from a chemdata.Thd; the
    slots past the species count hold `fill`."""
    n = len(thd.W)
    F = dict(order="F")

    def vec(x):
        a = np.full(ld, fill)
        a[:n] = x
        return a

    heat = np.full((7, 2, ld), fill, **F)
    heat[:, :, :n] = np.transpose(thd.heatcap, (2, 1, 0))     # heatcap[k][r][c]
    cofd = np.full((4, ld, ld), fill, **F)
    cofd[:, :n, :n] = np.transpose(thd.cofd, (2, 0, 1))       # cofd[l][m][c] = CofD(c,l,m)
    coflam = np.full((4, ld), fill, **F)
    coflam[:, :n] = thd.coflam.T
    cofeta = np.full((4, ld), fill, **F)
    cofeta[:, :n] = thd.cofeta.T
    return dict(W=vec(thd.W), Wc=vec(thd.Wc), Wc_i=vec(thd.Wc_i), HeatCap=heat,
                Tpolymin=vec(thd.tpolymin), Tpolylim=vec(thd.tpolylim),
                Tpolymax=vec(thd.tpolymax), CofD=cofd, CofLam=coflam, CofEta=cofeta)


THD_ORDER = ("W", "Wc", "Wc_i", "HeatCap", "Tpolymin", "Tpolylim", "Tpolymax", "CofD",
             "CofLam", "CofEta")


def thd_ref_quantities(thd) -> list:
    """gam2, T_ref, p_ref, D_ref, lbda_ref, mu_ref as izem_f_create takes
    them."""
    return [thd.gam2, thd.Tinf, thd.pinf, thd.Dinf, thd.laminf, thd.muinf]


def thd_image(lib, thd, arrays: dict, nspc: int = 9, cap: int = 745,
              ld: int | None = None) -> tuple[int, np.ndarray]:
    img = np.zeros(max(cap, 1))
    ld = arrays["W"].shape[0] if ld is None else ld
    ptrs = [arrays[k].ctypes.data for k in THD_ORDER]
    n = lib.izem_thd_image(nspc, ld, *ptrs, thd.gam2, thd.Tinf, thd.p0, thd.pinf, thd.Dinf,
                           thd.laminf, thd.muinf, img.ctypes.data, cap)
    return n, img


def ckwrk_addresses(refchem) -> tuple[int, int, int]:
    """Host addresses of IWORK, RWORK (/CKPASR/) and
    /CKSTRT/ in a loaded oracle.oracle.RefChem library, after INITCHEMKIN."""
    base = C.addressof(C.c_int.in_dll(refchem.lib, "ckpasr_"))
    ck = C.addressof(C.c_int.in_dll(refchem.lib, "ckstrt_"))
    return base, base + 4 * LENIWK, ck
