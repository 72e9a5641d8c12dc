"""The oracle's characteristic outflow (bc_nscbc_out) pinned by the LODI
relations of a subsonic outflow (Poinsot and Lele, J. Comput. Phys. 101,
1992) with the partially reflecting relaxation of Rudy and Strikwerda
(J. Comput. Phys. 36, 1980):

    d(rho)/dt = -d1,  d(p)/dt = -d2,  d(u)/dt = -d3,  d(v)/dt = -d4, ...
    d2 = (L5 + L1) / 2,  d1 = (L2 + d2) / c^2,  d3 = (L5 - L1) / (2 rho c),

and on a uniform state every derivative vanishes, so L2 = L3 = L4 = L5 = 0
and L1 = K (p - p_inf), K = sigma (1 - M^2) c / Lx, sigma = 0.28: the last
cell's density, velocity and pressure move, to first order in dt, by
-dt K (p - p_inf) / (2 c^2), +dt K (p - p_inf) / (2 rho c) and
-dt K (p - p_inf) / 2, the pressure toward p_inf, the tangential velocity
not at all; the ghosts take the last cell's new state.  The energy update
is the calorically perfect form (d(rho E) = d(p) / (gamma - 1) + ...), so
the gas of the check has a constant c_p = 3.5 R (e = c_v T); for the
NASA-7 mixture the pressure change differs by the (e - c_v T) d1 term the
reference leaves out.  A uniform supersonic outflow is left as it is.  The
relaxation taken with the wrong sign fails the check.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINF = 94232.25


def _perfect_gas(thd):
    import copy
    t = copy.deepcopy(thd)
    t.heatcap = np.zeros_like(thd.heatcap)
    t.heatcap[:, :, 0] = 3.5
    return t


def _uniform(oracle, adi, thd, mach_x, p_pa):
    from oracle.oracle import BC_EXTRAPOLATION, BC_NSCBC_OUT
    from phd_amd import ics
    from phd_amd.ics import Box
    from tests.flowcase import OracleRun
    thd = _perfect_gas(thd)
    nx, ny = 16, 8
    box = Box(nx, ny, 1, 3, 2, (0.004, 0.002, 1.0))
    Y = np.zeros(9)
    Y[1], Y[8] = 0.233, 0.767
    Tk = np.full((1, ny, nx), 900.0)
    # the frozen sound speed of the mixture at 900 K (gamma from c_p)
    from phd_amd.thermo_host import cp_scalar, wmix_i_scalar
    Wi = wmix_i_scalar(thd, Y)
    g = thd.gam2 * Wi / cp_scalar(thd, 900.0 / adi.T_ref, Y)
    g = 1.0 / (1.0 - g)
    c = np.sqrt(g * 900.0 / adi.T_ref * Wi) * adi.u_ref
    u = np.zeros((3, 1, ny, nx))
    u[0] = mach_x * c
    u[1] = 0.05 * c
    Yb = np.broadcast_to(Y[:, None, None, None], (9, 1, ny, nx))
    v, T = ics.conserved_from_primitive(box, adi, thd, Tk, p_pa, Yb, u)
    run = OracleRun(oracle, adi, thd, box, v, T)
    run.g.bc[0], run.g.bc[1] = BC_EXTRAPOLATION, BC_NSCBC_OUT
    run.g.nscbc_Lx = box.L[0] * (nx - 1) / nx / adi.L_ref
    run.g.nscbc_pinf = PINF / adi.p_ref
    run.prim_all()
    return box, run


def _apply(oracle, run, dt):
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    oracle.lib.ora_bc_nscbc_out(C.byref(run.g), C.byref(run.t), C.c_double(dt), P(run.d_i[0]),
                                P(run.T), P(run.W_i), P(run.cp), P(run.v))


def _lodi_errors(oracle, adi, thd):
    """Relative deviations of the last cell's changes from the LODI ones."""
    box, run = _uniform(oracle, adi, thd, 0.3, 1.2 * PINF)
    g, ex = box.ng, box.ng + box.nx - 1
    c0 = (g, g + 3, ex)
    rho, T, Wi, cp = run.v[(0,) + c0], run.T[c0], run.W_i[c0], run.cp[c0]
    u, w = run.v[(1,) + c0] / rho, run.v[(2,) + c0] / rho
    p = rho * T * Wi
    gam = 1.0 / (1.0 - run.t.gam2 * Wi / cp)
    assert abs(gam - 1.4) < 1e-12
    c = np.sqrt(gam * p / rho)
    M = np.sqrt(u * u + w * w) / c
    pinf = PINF / adi.p_ref
    K = 0.28 * (1.0 - M * M) * c / run.g.nscbc_Lx
    dt = 1e-4 / K
    _apply(oracle, run, dt)
    r2 = run.v[(0,) + c0]
    u2, w2 = run.v[(1,) + c0] / r2, run.v[(2,) + c0] / r2
    run.prim_all()
    p2 = r2 * run.T[c0] * run.W_i[c0]
    L1 = K * (p - pinf)
    errs = {"rho": (r2 - rho) / (-dt * L1 / (2 * c * c)) - 1.0,
            "u": (u2 - u) / (dt * L1 / (2 * rho * c)) - 1.0,
            "p": (p2 - p) / (-dt * L1 / 2) - 1.0,
            "w": (w2 - w) / w}
    ghosts = [run.v[(slice(None), g, g + 3, ex + q)] for q in range(1, g + 1)]
    same = all(np.array_equal(gq, run.v[(slice(None),) + c0]) for gq in ghosts)
    return errs, same


def test_subsonic_outflow_follows_lodi(oracle, adi, thd):
    errs, same = _lodi_errors(oracle, adi, thd)
    print({k: f"{v:.1e}" for k, v in errs.items()})
    assert abs(errs["rho"]) < 1e-3 and abs(errs["u"]) < 1e-3 and abs(errs["p"]) < 1e-3
    assert abs(errs["w"]) < 1e-12
    assert same


def test_supersonic_uniform_outflow_is_left_alone(oracle, adi, thd):
    box, run = _uniform(oracle, adi, thd, 1.5, 1.2 * PINF)
    v0 = run.v.copy()
    _apply(oracle, run, 1e-2)
    ex = box.ng + box.nx - 1
    I = (slice(None), slice(box.ng, box.ng + 1), slice(box.ng, box.ng + box.ny), ex)
    assert np.allclose(run.v[I], v0[I], rtol=1e-13, atol=0)


def test_outflow_pin_sees_planted_mistake(adi, thd, tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    from oracle.oracle import Oracle
    src = open(os.path.join(ROOT, "oracle", "izem_oracle.c")).read()
    old = "cs / Lx * (Pres - Pinf);"
    assert src.count(old) == 1
    shutil.copy(os.path.join(ROOT, "oracle", "izem_oracle.h"), tmp_path / "izem_oracle.h")
    (tmp_path / "izem_oracle.c").write_text(src.replace(old, "cs / Lx * (Pinf - Pres);"))
    so = tmp_path / "liboracle.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-std=c11", "-shared", "-o",
                    str(so), str(tmp_path / "izem_oracle.c"), "-lm"], check=True)
    errs, _ = _lodi_errors(Oracle(str(so)), adi, thd)
    assert abs(errs["p"]) > 1.0
