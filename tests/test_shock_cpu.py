"""The double Mach reflection's host side (phd_amd.shock) and the post-shock
slip wall of the oracle, pinned by what the physics fixes.

* The Rankine-Hugoniot jump (RH_variables): for a calorically perfect gas
  (NASA-7 tables with a constant c_p = 3.5 R) the textbook normal-shock
  ratios at Mach 4.5 and the theta-beta-Mach relation of an oblique shock;
  for N2 with its real c_p(T) the mass and momentum balances to rounding
  and the energy balance to the reference's Newton tolerance; the bisection
  path gives the same jump; the returned post-shock enthalpy is the one at
  the differencing point T1 (phi + 1e-5), the reference's own.
* The case's states: the laboratory post-shock speed from the mass balance,
  the initial shock line at 90 - 36 degrees to the wall through x0, the wall
  state moving along (cos 36, -sin 36).
* The deck: the reference's dmr.inp, grid.dat, therm.dat (a duplicated N2
  entry: the first is read, as the reference's reader reads it).
* The slip wall in the oracle: on a 2-D state with the wall on the whole
  bottom face, the Euler right-hand side equals, bit for bit, the one of the
  state extended by its mirror image below the wall, on every row whose
  stencils stay clear of the top face; a planted sign error in the mirrored
  rho v, or a mirror about the face instead of the first row, breaks it.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def dmr_thd(adi):
    from phd_amd import chemdata as cd
    from phd_amd import tranfit as tf
    therm = os.path.join(GOLD, "dmr_therm.dat")
    mech = cd.read_mechanism(os.path.join(GOLD, "dmr_chem.inp"), therm)
    fits = tf.fit_transport(mech, os.path.join(GOLD, "dmr_tran.dat"), therm)
    return cd.build_thd(adi, therm_path=therm, fits=fits)


def _perfect_gas(thd, gamma=1.4):
    """The same species with c_p = gamma / (gamma - 1) R in both ranges."""
    import copy
    t = copy.deepcopy(thd)
    t.heatcap = np.zeros_like(thd.heatcap)
    t.heatcap[:, :, 0] = gamma / (gamma - 1.0)
    return t


def _state1(adi, thd):
    from phd_amd.thermo_host import cp_scalar, wmix_i_scalar
    Y = np.array([0.0, 1.0])
    p1, T1 = 101325.0 / adi.p_ref, 300.0 / adi.T_ref
    Wi = wmix_i_scalar(thd, Y)
    rho1 = p1 / (T1 * Wi)
    g = thd.gam2 * Wi / cp_scalar(thd, T1, Y)
    g = 1.0 / (1.0 - g)
    return Y, p1, T1, rho1, g, math.sqrt(g * p1 / rho1)


def test_perfect_gas_normal_shock(adi, dmr_thd):
    from phd_amd import shock
    thd = _perfect_gas(dmr_thd)
    Y, p1, T1, rho1, g, cs1 = _state1(adi, thd)
    assert abs(g - 1.4) < 1e-12
    M = 4.5
    j = shock.jump(thd, Y, p1, T1, M * cs1, math.pi / 2.0)
    pr = 1.0 + 2.0 * g / (g + 1.0) * (M * M - 1.0)
    rr = (g + 1.0) * M * M / ((g - 1.0) * M * M + 2.0)
    assert abs(j.p2 / p1 / pr - 1.0) < 2e-6
    assert abs(j.rho2 / rho1 / rr - 1.0) < 2e-6
    assert abs(j.T2 / T1 / (pr / rr) - 1.0) < 2e-6


def test_perfect_gas_oblique_shock(adi, dmr_thd):
    """theta-beta-Mach: tan theta = 2 cot beta (M^2 sin^2 beta - 1) /
    (M^2 (gamma + cos 2 beta) + 2)."""
    from phd_amd import shock
    thd = _perfect_gas(dmr_thd)
    Y, p1, T1, rho1, g, cs1 = _state1(adi, thd)
    M, beta = 3.0, math.radians(40.0)
    j = shock.jump(thd, Y, p1, T1, M * cs1, beta)
    tt = 2.0 / math.tan(beta) * (M * M * math.sin(beta) ** 2 - 1.0) / \
        (M * M * (g + math.cos(2.0 * beta)) + 2.0)
    assert abs(math.tan(j.theta) / tt - 1.0) < 2e-6
    # the tangential speed is conserved across the shock
    w1 = M * cs1
    assert abs(j.w2 * math.cos(beta - j.theta) / (w1 * math.cos(beta)) - 1.0) < 1e-12


@pytest.mark.parametrize("newton", [True, False])
def test_real_gas_jump_balances(adi, dmr_thd, newton):
    from phd_amd import shock
    from phd_amd.thermo_host import ha_scalar
    Y, p1, T1, rho1, g, cs1 = _state1(adi, dmr_thd)
    w1 = 4.5 * cs1
    j = shock.jump(dmr_thd, Y, p1, T1, w1, math.pi / 2.0,
                   newton_itmax=shock.ITMAX if newton else 0)
    u2 = j.w2
    assert abs(rho1 * w1 - j.rho2 * u2) <= 1e-13 * rho1 * w1
    assert abs(p1 + rho1 * w1 * w1 - (j.p2 + j.rho2 * u2 * u2)) <= 1e-12 * (p1 + rho1 * w1 * w1)
    h2 = float(np.dot(ha_scalar(dmr_thd, j.T2), Y))
    e1 = j.h1 + 0.5 * w1 * w1
    assert abs(e1 - (h2 + 0.5 * u2 * u2)) <= 1e-6 * e1
    # the real gas's c_p(T) rises with T: a colder, denser jump than gamma 1.4
    M = 4.5
    assert j.T2 / T1 < (1.0 + 2.8 / 2.4 * (M * M - 1.0)) / (2.4 * M * M / (0.4 * M * M + 2.0))
    if newton:
        # the enthalpy returned is the one at the differencing point
        hd = float(np.dot(ha_scalar(dmr_thd, T1 * (j.T2 / T1 + shock.DPHI)), Y))
        assert abs(j.h2 - hd) <= 1e-15 * abs(hd) + 1e-300
        jb = shock.jump(dmr_thd, Y, p1, T1, w1, math.pi / 2.0, newton_itmax=0)
        assert abs(jb.T2 / j.T2 - 1.0) < 1e-5


def test_dmr_states(adi, dmr_thd):
    from phd_amd import grid, shock
    s = shock.dmr_states(adi, dmr_thd)
    assert abs(s.w2 - s.Us * (1.0 - s.rho1 / s.jump.rho2)) <= 1e-13 * s.Us
    assert 1500.0 < s.Us * adi.u_ref < 1700.0            # Mach 4.5 in N2 at 300 K
    w = shock.dmr_wall_state(adi, dmr_thd)
    assert abs(w[2] / w[1] + math.tan(math.radians(36.0))) < 1e-12
    assert abs(math.hypot(w[1], w[2]) / (s.jump.rho2 * s.w2) - 1.0) < 1e-14
    assert w[3] == 0.0 and w[5] == 0.0 and w[6] == s.jump.rho2
    # the initial shock line on a coarse copy of the case's grid
    box = grid.box_from_grid(os.path.join(GOLD, "dmr_grid.dat"), 900, 511, 1, 2, adi, ng=4)
    xt, yt, _ = grid.read_grid(os.path.join(GOLD, "dmr_grid.dat"), 900, 511, 1)
    assert shock.dmr_wall_cells(adi, xt) == int(np.sum(xt <= 0.0)) > 0
    sub = type(box)(60, 40, 1, 4, 2, box.L)
    xs, ys = xt[::15], yt[::12][:40]
    v, T = shock.dmr_initial_state(adi, dmr_thd, sub, xs, ys)
    I = sub.interior()
    post = v[(0,) + I][0] == s.jump.rho2
    X, Yg = np.meshgrid(xs, ys)
    line = X <= Yg / math.tan(math.radians(90.0 - 36.0))
    assert np.array_equal(post, line) and 0 < post.sum() < post.size


def test_dmr_deck(adi):
    from phd_amd.deck import Deck, DeckError, decomposition, face_conditions, library_key, options
    d = Deck.read(os.path.join(GOLD, "dmr.inp"))
    o = options(d)
    assert (o.vis, o.reaction, o.CFL) == (0, 0, 0.3)
    assert face_conditions(d) == (1, 1, 2, 1, 0, 0)
    assert library_key(d) == (2, 4) and decomposition(d) == (4, 1, 1)
    with pytest.raises(DeckError, match="BCXMAX"):
        options(Deck.parse("[GENERAL]\n  NRV = 9\n  INIT = 'DMR'\n[WENO]\n  ORDER_WENO = 5\n"
                           "[GRID]\n  BCXMAX = 'POSTSHOCK_SLIPWALL'\n"))
    with pytest.raises(DeckError, match="INIT = 'DMR'"):
        options(Deck.parse("[GENERAL]\n  NRV = 9\n[WENO]\n  ORDER_WENO = 5\n"
                           "[GRID]\n  BCYMIN = 'POSTSHOCK_SLIPWALL'\n"))


def test_therm_first_entry_is_read():
    from phd_amd import chemdata as cd
    nasa, _ = cd.read_therm(os.path.join(GOLD, "dmr_therm.dat"))
    assert nasa["N2"].low[1] == 0.01408240e-01      # the first N2 entry (0.14082404E-02 later)


# ---------------------------------------------------------------- slip wall pin
def _mirror_mismatch(oracle, adi, thd) -> float:
    """Largest difference, relative to the largest value, between the wall
    box's y-flux divergence and the mirrored box's on the rows whose stencils
    stay clear of the top face."""
    from phd_amd.ics import Box
    from tests.flowcase import OracleRun
    from oracle.oracle import BC_EXTRAPOLATION, BC_POSTSHOCK_SLIPWALL
    nx, ny, ng = 8, 20, 3
    M = ny - 1

    def state(rows):
        y = np.asarray(rows, dtype=np.float64)
        T = 5.0 + 0.6 * np.cos(0.31 * y)           # even about the wall row
        u = 0.3 + 0.1 * np.cos(0.2 * y)
        w = 0.25 * np.sin(0.45 * y)                 # odd about the wall row
        return T, u, w

    def build(ny_box, rows):
        # the same spacing in both boxes
        box = Box(nx, ny_box, 1, ng, 2, (1.0, ny_box / ny, 1.0))
        I = box.interior()
        T, u, w = state(rows)
        Y = np.zeros(9)
        Y[8], Y[1] = 0.77, 0.23
        from phd_amd import ics
        Tk = np.broadcast_to(T[None, :, None], (1, ny_box, nx)) * adi.T_ref
        uu = np.zeros((3, 1, ny_box, nx))
        uu[0] = u[None, :, None] * adi.u_ref
        uu[1] = w[None, :, None] * adi.u_ref
        Yb = np.broadcast_to(Y[:, None, None, None], (9, 1, ny_box, nx))
        v, Tn = ics.conserved_from_primitive(box, adi, thd, Tk, 101325.0, Yb, uu)
        return box, v, Tn

    results = {}
    for kind in ("wall", "mirror"):
        if kind == "wall":
            box, v, T = build(ny, np.arange(ny))
        else:
            box, v, T = build(2 * M + 1, np.arange(-M, M + 1))
        run = OracleRun(oracle, adi, thd, box, v, T)
        run.g.bc[2], run.g.bc[3] = (BC_POSTSHOCK_SLIPWALL if kind == "wall" else BC_EXTRAPOLATION,
                                    BC_EXTRAPOLATION)
        run.g.bc[0] = run.g.bc[1] = 0
        run.g.nfix = 0
        oracle.lib.ora_upd_boundaries(C.byref(run.g), run.v.ctypes.data_as(C.POINTER(C.c_double)))
        run.prim_all()
        f = np.zeros_like(run.v)
        P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        sensor = np.ones(box.shape, np.uint8)
        st = oracle.lib.ora_euler_llf(1, C.byref(run.g), C.byref(run.t), P(run.d_i[1]),
                                      P(run.T), P(run.W_i), P(run.cp), P(run.ha),
                                      sensor.ctypes.data_as(C.POINTER(C.c_uint8)), 1, P(run.v),
                                      P(f))
        assert st == 0
        off = 0 if kind == "wall" else M
        rows = slice(ng + off, ng + off + ny - ng - 1)
        results[kind] = f[:, ng, rows, ng:ng + nx]
    a, b = results["wall"], results["mirror"]
    assert np.max(np.abs(a[2])) > 1e-7 * np.max(np.abs(a))     # rho v has a real divergence
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_oracle_slip_wall_equals_mirror_image(oracle, adi, thd):
    # the ghost-slab Newton exits on its own slab's maximum, the mirrored
    # box's on the interior's: temperatures differ in the last bit
    assert _mirror_mismatch(oracle, adi, thd) <= 1e-12


@pytest.mark.parametrize("edit", ["sign", "offset"])
def test_slip_wall_pin_sees_planted_mistakes(adi, thd, tmp_path, edit):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    from oracle.oracle import Oracle
    src = open(os.path.join(ROOT, "oracle", "izem_oracle.c")).read()
    if edit == "sign":
        old, new = "(l == 2) ? -a : a;", "a;"
    else:
        old, new = "cidx(g, i, s, k)];", "cidx(g, i, s - 1, k)];"
    assert src.count(old) == 1
    shutil.copy(os.path.join(ROOT, "oracle", "izem_oracle.h"), tmp_path / "izem_oracle.h")
    (tmp_path / "izem_oracle.c").write_text(src.replace(old, new))
    so = tmp_path / "liboracle.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-std=c11", "-shared", "-o",
                    str(so), str(tmp_path / "izem_oracle.c"), "-lm"], check=True)
    assert _mirror_mismatch(Oracle(str(so)), adi, thd) > 1e-4
