"""The grid front end: the reference's grid.dat format and the metric
arrays 1/dx that the flow kernels take (metrics, module_tools.f90:223-442;
dscalar, module_deriv.f90:80-201; adimensionalize, module_tools.f90:160-212).

grid.dat holds the keyword X-DIRECTION, ntx coordinates one per record, then
Y-DIRECTION and nty, Z-DIRECTION and ntz, and END (blank records are skipped
by the list-directed reads).  The metric of a direction is the derivative
of its coordinates with respect to the index, taken with the reference's
stencils: sixth-order central in the interior (ng = 3; eighth order with
sixth order at the fourth point from each end for ng = 4), fourth-order central
at the third point from each end and fourth-order one-sided and upwind
closures at the first two; it is inverted on 1..nt, and the ghost entries
keep the value 1 that the reference initialises them to.  Nondimensional
arrays multiply every entry by L_ref and divide the coordinates by it.
"""
from __future__ import annotations

import numpy as np

from . import chemdata as cd
from .ics import Box

# stencil coefficients (module_deriv.f90:23-66), index offset -> weight
_A6 = {-3: -1.0 / 60.0, -2: 9.0 / 60.0, -1: -45.0 / 60.0, 0: 0.0, 1: 45.0 / 60.0,
       2: -9.0 / 60.0, 3: 1.0 / 60.0}
_A8 = {-4: 3.0 / 840.0, -3: -32.0 / 840.0, -2: 168.0 / 840.0, -1: -672.0 / 840.0, 0: 0.0,
       1: 672.0 / 840.0, 2: -168.0 / 840.0, 3: 32.0 / 840.0, 4: -3.0 / 840.0}
_A4 = {-2: 1.0 / 12.0, -1: -8.0 / 12.0, 0: 0.0, 1: 8.0 / 12.0, 2: -1.0 / 12.0}
_A4OL = {0: -25.0 / 12.0, 1: 48.0 / 12.0, 2: -36.0 / 12.0, 3: 16.0 / 12.0, 4: -3.0 / 12.0}
_A4UL = {-1: -3.0 / 12.0, 0: -10.0 / 12.0, 1: 18.0 / 12.0, 2: -6.0 / 12.0, 3: 1.0 / 12.0}
_A4UR = {-3: 1.0 / 12.0, -2: -6.0 / 12.0, -1: 18.0 / 12.0, 0: -10.0 / 12.0, 1: -3.0 / 12.0}
_A4OR = {-4: -3.0 / 12.0, -3: 16.0 / 12.0, -2: -36.0 / 12.0, -1: 48.0 / 12.0, 0: -25.0 / 12.0}


def count_points(path: str) -> tuple[int, int, int]:
    """Points per direction of a grid.dat, counted as the reference counts
    them (module_parallel.f90:271-306): the records between X-DIRECTION and
    Y-DIRECTION, Y-DIRECTION and Z-DIRECTION, Z-DIRECTION and END (a
    list-directed read skips blank records)."""
    words = [ln.split()[0] for ln in open(path).read().splitlines() if ln.split()]
    try:
        p = [words.index(w) for w in ("X-DIRECTION", "Y-DIRECTION", "Z-DIRECTION", "END")]
    except ValueError:
        raise ValueError(f"{path}: missing direction in grid.dat") from None
    if not p[0] < p[1] < p[2] < p[3]:
        raise ValueError(f"{path}: directions out of order")
    return p[1] - p[0] - 1, p[2] - p[1] - 1, p[3] - p[2] - 1


def read_grid(path: str, ntx: int, nty: int, ntz: int):
    """Coordinates (xt, yt, zt) of a grid.dat (module_tools.f90:276-297)."""
    recs = [ln.split() for ln in open(path).read().splitlines()]
    recs = [r for r in recs if r]
    p = 0
    out = []
    for word, n in (("X-DIRECTION", ntx), ("Y-DIRECTION", nty), ("Z-DIRECTION", ntz)):
        if recs[p][0] != word:
            raise ValueError(f"grid.dat: key word {word} is missing")
        p += 1
        out.append(np.array([float(recs[p + i][0].replace("D", "E").replace("d", "e"))
                             for i in range(n)]))
        p += n
    if p >= len(recs) or recs[p][0] != "END":
        raise ValueError("grid.dat: key word END is missing")
    return tuple(out)


def _deriv(v: np.ndarray, m: int, coef: dict) -> float:
    # dd = 0; dd = dd + a(l) v(m+l) in increasing l (dscalar's loops)
    dd = 0.0
    for l in sorted(coef):
        dd = dd + coef[l] * v[m + l]
    return dd


def metric(t: np.ndarray, ng: int = 3) -> np.ndarray:
    """1/(dt/di) of one direction's nt coordinates, as n + 2 ng entries with
    the ghost entries at 1 (dscalar with d_i = 1, then the inversion).  ng = 3
    (WENO5): sixth order inside; ng = 4 (WENO7): eighth order inside and
    sixth order at the fourth point from each end."""
    if ng not in (3, 4):
        raise ValueError("the metric is defined for ng = 3 (WENO5) and ng = 4 (WENO7)")
    nt = t.size
    if nt < 2 * ng + 1:
        raise ValueError("not enough points for the metric stencils")
    dv = np.ones(nt)
    for m in range(ng, nt - ng):
        dv[m] = _deriv(t, m, _A6 if ng == 3 else _A8) * 1.0
    if ng == 4:
        dv[3] = _deriv(t, 3, _A6) * 1.0
        dv[nt - 4] = _deriv(t, nt - 4, _A6) * 1.0
    dv[0] = _deriv(t, 0, _A4OL) * 1.0
    dv[1] = _deriv(t, 1, _A4UL) * 1.0
    dv[2] = _deriv(t, 2, _A4) * 1.0
    dv[nt - 3] = _deriv(t, nt - 3, _A4) * 1.0
    dv[nt - 2] = -_deriv(t, nt - 2, _A4UR) * 1.0
    dv[nt - 1] = -_deriv(t, nt - 1, _A4OR) * 1.0
    out = np.ones(nt + 2 * ng)
    out[ng:ng + nt] = 1.0 / dv
    return out


def box_from_grid(path: str, ntx: int, nty: int, ntz: int, ndim: int,
                  adi: "cd.Adim | None" = None, ng: int = 3) -> Box:
    """A one-rank Box whose metric arrays are the reference's for the grid
    in path, nondimensionalised with L_ref (the directions beyond ndim keep
    1 * L_ref, as the reference leaves them)."""
    adi = adi or cd.Adim()
    xt, yt, zt = read_grid(path, ntx, nty, ntz)
    mets = []
    for d, t in enumerate((xt, yt, zt)):
        a = metric(t, ng) if d < ndim else np.ones(t.size + 2 * ng)
        mets.append(a * adi.L_ref)
    # physical sides of the periodic box: extent plus one spacing
    L = tuple(float(t[-1] - t[0] + (t[1] - t[0])) if t.size > 1 else 1.0 for t in (xt, yt, zt))
    return Box(ntx, nty, ntz, ng, ndim, L, metrics=tuple(mets))
