"""Shock states of the double Mach reflection case: the Rankine-Hugoniot
jump of a mixture with temperature-dependent heat capacities, the initial
state of ini_DMR and the post-shock state of its bottom boundary.

The jump (RH_variables, devel-izem/src/module_Rankine_Hugoniot.f90:66-308)
is solved for the temperature ratio phi = T2 / T1.  With the mass and
momentum balances the pressure ratio is a function of phi,

    psi(phi) = (1 + m + sqrt((1 + m)^2 - 4 m phi)) / 2,   m = rho1 u1^2 / p1

(the strong root; :377-406), and the energy balance leaves one equation,

    F(phi) = h1 - h(T1 phi) + u1^2 / 2 (1 - phi^2 / psi^2) = 0      (:451-465)

solved by Newton's method with a forward difference of step 1e-5 in phi,
started from the calorically perfect ratio (:323-341) and stopped when the
next correction is below 1e-6 (the correction is not applied; the mixture
enthalpy returned is the one of the differencing point, as the reference
returns it); bisection between phi = 1 and the largest ratio with a real
psi (:352-366) when Newton's method fails.  All arithmetic follows the
reference's order, so the states equal its own to the last bit.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .thermo_host import cp_scalar, ha_scalar, wmix_i_scalar

ITMAX = 50          # module_Rankine_Hugoniot.f90:47-50
EPS_PHI = 1.0e-6
DPHI = 1.0e-5
NEG_TOL = -1.0e-2


class ShockError(RuntimeError):
    pass


@dataclass
class Jump:
    rho1: float
    h1: float
    p2: float
    T2: float
    rho2: float
    w2: float       # post-shock speed relative to the shock
    h2: float
    theta: float    # flow deflection


def _pressure_ratio(rho1: float, u1: float, p1: float, phi: float) -> tuple[float, bool]:
    """psi(phi) and whether the discriminant was too negative
    (module_Rankine_Hugoniot.f90:377-406)."""
    m = rho1 * u1 * u1 / p1
    disc = (1.0 + m) * (1.0 + m)
    disc = disc - 4.0 * rho1 * u1 * u1 * phi / p1
    bad = disc < NEG_TOL
    if NEG_TOL < disc < 0.0:
        disc = 0.0
    root = math.sqrt(disc) if disc >= 0.0 else math.nan
    return 0.5 * (root + 1.0 + m), bad


def _mixture_h(thd, Y, T: float) -> float:
    h = 0.0
    ha = ha_scalar(thd, T)
    for l in range(len(thd.W)):
        h = h + ha[l] * float(Y[l])
    return h


def jump(thd, Y, p1: float, T1: float, w1: float, beta: float,
         newton_itmax: int = ITMAX) -> Jump:
    """The oblique-shock jump of a gas at p1, T1 (nondimensional) crossing
    the shock at speed w1 and angle beta (module_Rankine_Hugoniot.f90:
    66-308).  newton_itmax: Newton iterations before the bisection takes
    over (the reference's 50)."""
    Wi1 = wmix_i_scalar(thd, Y)
    cp1 = cp_scalar(thd, T1, Y)
    rho1 = p1 / (T1 * Wi1)
    u1 = w1 * math.sin(beta)
    g1 = thd.gam2 * Wi1 / cp1
    g1 = 1.0 / (1.0 - g1)
    cs1 = math.sqrt(g1 * p1 / rho1)
    M1n = u1 / cs1
    # calorically perfect ratio (RH_phi0)
    a = 0.5 * (g1 + 1.0) * M1n
    phi = g1 * M1n * M1n - 0.5 * (g1 - 1.0)
    phi = phi * (1.0 + 0.5 * (g1 - 1.0) * M1n * M1n)
    phi = phi / (a * a)
    h1 = _mixture_h(thd, Y, T1)
    m = rho1 * u1 * u1 / p1

    def F(phi, psi, h2):
        return h1 - h2 + 0.5 * u1 * u1 * (1.0 - phi * phi / (psi * psi))

    bisect, it = False, 0
    psi = h2 = math.nan
    while True:
        it += 1
        if it > newton_itmax:
            bisect = True
            break
        h2 = _mixture_h(thd, Y, T1 * phi)
        psi, bad = _pressure_ratio(rho1, u1, p1, phi)
        if bad:
            bisect = True
            break
        f = F(phi, psi, h2)
        h2 = _mixture_h(thd, Y, T1 * (phi + DPHI))
        psi_d, bad = _pressure_ratio(rho1, u1, p1, phi + DPHI)
        if bad:
            bisect = True
            break
        df = (F(phi + DPHI, psi_d, h2) - f) / DPHI
        if abs(f / df) < EPS_PHI:
            break
        phi = phi - f / df
    if bisect:
        lo = 1.0
        hi = (1.0 + m) * (1.0 + m) * p1 / (4.0 * rho1 * u1 * u1)      # RH_max_phi
        ends = []
        for ph in (lo, hi):
            h2 = _mixture_h(thd, Y, T1 * ph)
            ps, bad = _pressure_ratio(rho1, u1, p1, ph)
            if bad:
                raise ShockError("Rankine-Hugoniot bisection: no real pressure ratio at an end")
            ends.append(F(ph, ps, h2))
        f_lo, f_hi = ends
        if f_lo * f_hi > 0.0:
            raise ShockError("Rankine-Hugoniot bisection: no root between 1 and phi_max")
        it = 0
        while True:
            it += 1
            if it > ITMAX:
                raise ShockError("Rankine-Hugoniot bisection did not converge")
            phi = 0.5 * (lo + hi)
            h2 = _mixture_h(thd, Y, T1 * phi)
            psi, bad = _pressure_ratio(rho1, u1, p1, phi)
            if bad:
                raise ShockError("Rankine-Hugoniot bisection: no real pressure ratio")
            f = F(phi, psi, h2)
            if abs(f) < EPS_PHI:
                break
            if f_lo * f < 0.0:
                hi, f_hi = phi, f
            else:
                lo, f_lo = phi, f
    T2 = T1 * phi
    p2 = p1 * psi
    rho2 = p2 / (T2 * Wi1)
    u2 = rho1 * u1 / rho2
    theta = beta - math.atan(u2 * math.tan(beta) / u1)
    w2 = u2 / math.sin(beta - theta)
    return Jump(rho1, h1, p2, T2, rho2, w2, h2, theta)


# ----------------------------------------------------------------------------
# double Mach reflection (ini_DMR, module_ICs.f90:265-412; bc_postshock_slipwall,
# module_BCs.f90:1065-1172): a Mach 4.5 shock in N2 at 1 atm and 300 K meeting
# a 36-degree wedge at x0 = 0
MACH = 4.5
WEDGE_DEG = 36.0


@dataclass
class DmrStates:
    Y: np.ndarray
    p1: float
    T1: float
    rho1: float
    h1: float
    Us: float           # shock speed
    jump: Jump
    w2: float           # post-shock speed in the laboratory frame


def dmr_states(adi, thd) -> DmrStates:
    """Pre- and post-shock states of the case (the same computation opens
    ini_DMR and bc_postshock_slipwall).  The species set must be O2, N2:
    the reference sets Y = (0, 1) by position."""
    if len(thd.W) != 2:
        raise ShockError(f"the double Mach reflection is set for 2 species (O2, N2), "
                         f"not {len(thd.W)}")
    Y = np.array([0.00, 1.00])
    p1 = 101325.0 / adi.p_ref
    T1 = 300.0 / adi.T_ref
    Wi = wmix_i_scalar(thd, Y)
    rho1 = p1 / (T1 * Wi)
    cp1 = cp_scalar(thd, T1, Y)
    g1 = thd.gam2 * Wi / cp1
    g1 = 1.0 / (1.0 - g1)
    cs1 = math.sqrt(g1 * p1 / rho1)
    Us = MACH * cs1
    pi = math.acos(-1.0)
    j = jump(thd, Y, p1, T1, Us, pi / 2.0)
    return DmrStates(Y, p1, T1, j.rho1, j.h1, Us, j, Us - j.w2)


def dmr_wall_state(adi, thd) -> np.ndarray:
    """The conserved post-shock state bc_postshock_slipwall writes on
    x <= x0 (module_BCs.f90:1131-1141): velocity along the wedge-normal
    shock's direction of motion, (cos 36, -sin 36)."""
    s = dmr_states(adi, thd)
    pi = math.acos(-1.0)
    tw = WEDGE_DEG * pi / 180.0
    r2, w2 = s.jump.rho2, s.w2
    out = [r2, r2 * w2 * math.cos(tw), -r2 * w2 * math.sin(tw), 0.0,
           r2 * s.jump.h2 - s.jump.p2 + 0.5 * r2 * w2 * w2]
    out += [r2 * float(y) for y in s.Y]
    return np.array(out)


def dmr_initial_state(adi, thd, box, xt: np.ndarray, yt: np.ndarray):
    """ini_DMR (module_ICs.f90:265-412): the post-shock state left of the
    shock line through x0 at 90 - 36 degrees to the wall (x <= x0 + y /
    tan(beta) and y >= tan(beta) (x - x0)), the gas at rest right of it;
    xt, yt: the cell coordinates [m] of the grid file.  Returns v (interior)
    and the temperature guess T."""
    s = dmr_states(adi, thd)
    pi = math.acos(-1.0)
    x0 = 0.0 / adi.L_ref
    tw = WEDGE_DEG * pi / 180.0
    beta = pi / 2.0 - tw
    slope = math.tan(beta)
    slope_i = 1.0 / slope
    u2 = s.w2 * math.sin(beta)
    v2 = -s.w2 * math.cos(beta)
    w2 = 0.0
    r1, r2 = s.rho1, s.jump.rho2
    post = [r2, r2 * u2, r2 * v2, r2 * w2,
            r2 * s.jump.h2 - s.jump.p2 + 0.5 * r2 * (u2 * u2 + v2 * v2 + w2 * w2)]
    post += [r2 * float(y) for y in s.Y]
    pre = [r1, r1 * 0.0, r1 * 0.0, r1 * 0.0,
           r1 * s.h1 - s.p1 + 0.5 * r1 * (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0)]
    pre += [r1 * float(y) for y in s.Y]
    x = np.asarray(xt, dtype=np.float64)[:box.nx] * (1.0 / adi.L_ref)
    y = np.asarray(yt, dtype=np.float64)[:box.ny] * (1.0 / adi.L_ref)
    # elementwise IEEE arithmetic, as the reference's per-cell test
    left = (x[None, :] <= (x0 + y[:, None] * slope_i)) & (y[:, None] >= (slope * (x[None, :] - x0)))
    nv = 5 + len(thd.W)
    v = np.zeros((nv,) + box.shape)
    T = np.zeros(box.shape)
    I = box.interior()
    for q in range(nv):
        v[(q,) + I] = np.where(left, post[q], pre[q])[None]
    T[I] = np.where(left, s.jump.T2, s.T1)[None]
    return v, T


def dmr_wall_cells(adi, xt: np.ndarray) -> int:
    """Cells of the bottom face that take the post-shock state: x(i) <= x0,
    a leading run of the grid's ascending x (module_BCs.f90:1126)."""
    x = np.asarray(xt, dtype=np.float64) * (1.0 / adi.L_ref)
    x0 = 0.0 / adi.L_ref
    n = int(np.sum(x <= x0))
    if not np.all(x[:n] <= x0) or np.any(x[n:] <= x0):
        raise ShockError("grid x is not ascending")
    return n
