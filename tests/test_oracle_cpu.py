"""CPU tests that pin the oracle (the checker of the HIP path).

Each test fixes a property that a plausible mistake in the restatement
would break: the reference's own compiled CHEMKIN (oracle/_ref), closed
forms, invariants (free stream, discrete conservation, e(T(e)) = e,
el.er = I, R Lambda L = dF/dU), polynomial exactness of the stencils and
the documented quirks of the reference (the reversed WENO5 linear
coefficients, the dtc = 1 floor of the convective time step).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from tests.flowcase import OracleRun, P, small_state

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libpsrref.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ------------------------------------------------------------ kinetics
@needs_ref
def test_oracle_ckwyr_matches_reference_chemkin(oracle, mech):
    """The restated CKWYR (with this project's mechanism parser) against the
    reference CHEMKIN-II compiled from its sources with its own interpreter
    (ckinterp39 -> chem.bin): pins unit conversions, falloff, third bodies,
    reverse rates through equilibrium constants."""
    from oracle.oracle import RefChem, make_mech
    rc = RefChem()
    m = make_mech(mech)
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(300):
        T = rng.uniform(300.0, 3200.0)
        rho = 10 ** rng.uniform(-6.0, -2.0)
        Y = rng.random(9) ** 3
        Y /= Y.sum()
        w = np.zeros(9)
        oracle.lib.ora_ckwyr(C.byref(m), rho, T, _dp(Y), _dp(w))
        wr = rc.wdot(rho, T, Y)
        worst = max(worst, np.max(np.abs(w - wr)) / np.max(np.abs(wr)))
    assert worst <= 1e-13


SRI_DIR = os.path.join(os.path.dirname(REF), "sri")
SRI_MECH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h2_air_sri_rev.inp")


@pytest.mark.skipif(not os.path.exists(os.path.join(SRI_DIR, "chem.bin")),
                    reason="oracle/_ref/sri not built")
def test_front_end_mechanism_matches_reference_chemkin(oracle, mech):
    """The mechanism front end and the restated CKWYR on a mechanism that
    uses SRI (5 values), 4-value Troe, a named falloff collider (+N2), REV
    and an irreversible reaction (tests/golden/h2_air_sri_rev.inp), against
    the reference's interpreter and CKWYR on the same file.  The bundled
    mechanism's rates differ from it at these states, so the comparison is
    not with the wrong linking file."""
    from oracle.oracle import RefChem, make_mech
    from phd_amd import chemdata as cd
    m_sri = cd.read_mechanism(SRI_MECH)
    assert [r["fal"] for r in m_sri.reactions if r["fal"]] == [1, 2, 4]
    assert sum(1 for r in m_sri.reactions if r["rpar"]) == 1
    assert sum(1 for r in m_sri.reactions if not r["rev"]) == 1
    rc = RefChem(SRI_DIR)
    rc0 = RefChem()
    m = make_mech(m_sri)
    rng = np.random.default_rng(4)
    worst, differs = 0.0, 0.0
    for _ in range(300):
        T = rng.uniform(300.0, 3200.0)
        rho = 10 ** rng.uniform(-6.0, -2.0)
        Y = rng.random(9) ** 3
        Y /= Y.sum()
        w = np.zeros(9)
        oracle.lib.ora_ckwyr(C.byref(m), rho, T, _dp(Y), _dp(w))
        wr = rc.wdot(rho, T, Y)
        scale = np.max(np.abs(wr))
        worst = max(worst, np.max(np.abs(w - wr)) / scale)
        differs = max(differs, np.max(np.abs(rc0.wdot(rho, T, Y) - wr)) / scale)
    assert worst <= 1e-13
    assert differs > 1e-3


@needs_ref
def test_oracle_reactor_rhs_matches_reference_fun(oracle, mech):
    from oracle.oracle import RefChem, make_mech
    rc = RefChem()
    m = make_mech(mech)
    rng = np.random.default_rng(1)
    for _ in range(100):
        z = np.concatenate([[rng.uniform(500, 3000)], rng.random(9)])
        z[1:] /= z[1:].sum()
        rho = 10 ** rng.uniform(-5, -2)
        zp = np.zeros(10)
        oracle.lib.ora_reactor_rhs(C.byref(m), rho, _dp(z), _dp(zp))
        ref = rc.rhs(rho, z)
        assert np.max(np.abs(zp - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_detailed_balance(oracle, mech):
    """Reverse rates come from equilibrium constants (CKRATT, vode.f:983-1036):
    with only reaction 1 (H + O2 = O + OH) in the table, concentrations set
    in chemical equilibrium -- c_O c_OH / (c_H c_O2) = K_c(T) computed here
    from the NASA-7 Gibbs energies -- give a vanishing net rate."""
    import copy
    from oracle.oracle import make_mech
    one = copy.deepcopy(mech)
    one.reactions = one.reactions[:1]
    m = make_mech(one)
    sp = {s: i for i, s in enumerate(mech.species)}

    def g_RT(k, T):
        a = mech.a[k][1 if T > mech.tmid[k] else 0]
        h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + \
            a[5] / T
        s = a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + \
            a[4] * T ** 4 / 4 + a[6]
        return h - s

    for T in (900.0, 1500.0, 2400.0):
        dG = g_RT(sp["O"], T) + g_RT(sp["OH"], T) - g_RT(sp["H"], T) - g_RT(sp["O2"], T)
        Kc = np.exp(-dG)  # no change in mole number
        c = np.zeros(9)
        c[sp["H"]], c[sp["O2"]], c[sp["O"]] = 1e-8, 2e-6, 3e-8
        c[sp["OH"]] = Kc * c[sp["H"]] * c[sp["O2"]] / c[sp["O"]]
        c[sp["N2"]] = 1e-5
        rho = float(np.dot(c, mech.wt))
        Y = c * mech.wt / rho
        w = np.zeros(9)
        oracle.lib.ora_ckwyr(C.byref(m), rho, T, _dp(Y), _dp(w))
        r = one.reactions[0]
        kf = r["A"] * np.exp(r["b"] * np.log(T) - r["EoR"] / T)
        fwd = kf * c[sp["H"]] * c[sp["O2"]]
        assert abs(w[sp["OH"]]) <= 1e-10 * fwd, (T, w[sp["OH"]], fwd)
        # and away from equilibrium the net rate has the sign of the affinity
        c2 = c.copy()
        c2[sp["OH"]] *= 0.5
        Y2 = c2 * mech.wt / float(np.dot(c2, mech.wt))
        oracle.lib.ora_ckwyr(C.byref(m), float(np.dot(c2, mech.wt)), T, _dp(Y2), _dp(w))
        assert w[sp["OH"]] > 0


def test_mechanism_tables(mech, thd):
    assert mech.kk == 9 and mech.ii == 21
    assert mech.species == [s.upper() for s in thd.names]
    # molecular weights: CHEMKIN float32 atomic weights vs the tranfit values
    assert np.allclose(mech.wt, thd.W * 1000.0, rtol=1e-6)
    fal = [r for r in mech.reactions if r["fal"]]
    assert len(fal) == 2 and all(r["fal"] == 3 for r in fal)
    dup = [r for r in mech.reactions if r["dup"]]
    assert len(dup) == 4
    # E/R uses the interpreter's float32 4.184 / 8.314510 conversion
    assert mech.reactions[0]["EoR"] == pytest.approx(1.6599e4 * float(np.float32(4.184)) /
                                                     8.314510, rel=0, abs=0)
    # elements balance in every reaction
    for r in mech.reactions:
        lhs = sum(mech.ncf[k] * n for k, n in r["reac"])
        rhs = sum(mech.ncf[k] * n for k, n in r["prod"])
        assert np.array_equal(lhs, rhs)


# ---------------------------------------------------------------- thermo
def test_nondimensional_constants(adi):
    """constants() closed forms (module_tools.f90:453-562)."""
    assert adi.sqgmr == pytest.approx(np.sqrt(1.4) * 2.25 / 65000.0, rel=1e-15)
    assert adi.W_ref == pytest.approx(8.31451 / 287.15, rel=1e-15)
    assert adi.c_inf == pytest.approx(np.sqrt(1.4 * 287.15 * 169.44), rel=1e-14)
    assert adi.u_ref == pytest.approx(np.sqrt(287.15 * 169.44), rel=1e-14)
    mu = 1.458e-6 * 169.44 ** 1.5 / (169.44 + 110.4)
    assert adi.mu_ref == pytest.approx(mu, rel=1e-15)
    assert adi.p_ref == pytest.approx(adi.rho_ref * adi.u_ref ** 2, rel=1e-15)
    assert adi.time_ref == pytest.approx(adi.L_ref / adi.u_ref, rel=1e-15)


def test_nasa_thermo_branches(oracle, thd):
    """cp_scalar / ha_scalar: continuity at Tpolylim (GRI data are fitted
    to join), C1 linear extrapolation below Tmin and above Tmax
    (module_thermodynamics.f90:797-857), and h' = cp."""
    lib = oracle.lib
    from oracle.oracle import make_thd
    t = make_thd(thd)
    Tinf = thd.Tinf
    for s in range(9):
        Y = np.zeros(9)
        Y[s] = 1.0
        ha = np.zeros(9)

        def h(TK):
            lib.ora_ha_scalar(C.byref(t), TK / Tinf, _dp(ha))
            return ha[s]

        def cp(TK):
            return lib.ora_cp_scalar(C.byref(t), TK / Tinf, _dp(Y))

        for Tk in (thd.tpolylim[s], thd.tpolymin[s], thd.tpolymax[s]):
            eps = 1e-6 * Tk
            # no jump: the change across [T-eps, T+eps] is the slope times 2 eps
            slope = max(cp(Tk + eps), cp(Tk - eps)) / thd.gam2 * (2 * eps / Tinf)
            assert abs(h(Tk + eps) - h(Tk - eps)) <= 1.5 * slope + 1e-12
            assert abs(cp(Tk + eps) - cp(Tk - eps)) <= 1e-3 * cp(Tk)
        # dh/dT = cp / gam2 inside [Tmin, Tmax] (h in units of R T_ref / W_ref,
        # cp in cp_ref = R / (W_ref gam2)); outside, h continues with the slope
        # it has at the end of the range while cp is extrapolated linearly
        d = 1e-3
        for Tk in (900.0, 1500.0, 2500.0):
            dh = (h(Tk + d) - h(Tk - d)) / (2 * d / Tinf)
            assert dh == pytest.approx(cp(Tk) / thd.gam2, rel=1e-6)
        for Tk, Te in ((thd.tpolymin[s] - 60.0, thd.tpolymin[s]),
                       (thd.tpolymax[s] + 500.0, thd.tpolymax[s])):
            dh = (h(Tk + d) - h(Tk - d)) / (2 * d / Tinf)
            assert dh == pytest.approx(cp(Te) / thd.gam2, rel=1e-6)


def test_newton_recovers_temperature(oracle, adi, thd):
    """prim_inv_var from a 5 % perturbed start returns the temperature the
    state was built from (e(T(e)) = e) and stops after the global exit."""
    box, v, T = small_state(adi, thd, "2d", 16)
    rng = np.random.default_rng(2)
    T0 = T.copy()
    run = OracleRun(oracle, adi, thd, box, v, T * (1 + 0.05 * rng.standard_normal(T.shape)))
    it = run.prim_all()
    I = box.interior()
    assert 2 <= it <= 10
    assert np.max(np.abs(run.T[I] - T0[I]) / T0[I]) <= 1e-9


# ------------------------------------------------------------------ WENO
def _weno(oracle, vp, vm, js=True):
    a, b = C.c_double(), C.c_double()
    fn = oracle.lib.ora_wenorec if js else oracle.lib.ora_wenorec_nw
    fn(_dp(np.ascontiguousarray(vp, dtype=float)), _dp(np.ascontiguousarray(vm, dtype=float)),
       C.byref(a), C.byref(b))
    return a.value, b.value


def test_weno_js_linear_data_uses_ideal_weights(oracle):
    """For linear data all smoothness indicators are equal, so the JS
    weights are the ideal ones and both sides reconstruct the face value."""
    f = np.arange(1, 7, dtype=float)  # points i_l-2 .. i_l+3, face between 3 and 4
    vm_, vp_ = _weno(oracle, f, f)
    assert vm_ == pytest.approx(3.5, abs=1e-14)
    assert vp_ == pytest.approx(3.5, abs=1e-14)


def test_weno5_linear_formula_reference_quirk(oracle):
    """The reference's ng = 3 linear reconstruction (module_weno.f90:262-266)
    has its coefficients reversed: on linear data it returns the value at
    the face i-1/2 (2.5) instead of i+1/2 (3.5).  The restatement keeps it."""
    f = np.arange(1, 7, dtype=float)
    vm_, vp_ = _weno(oracle, f, f, js=False)
    assert vm_ == pytest.approx(2.5, abs=1e-14)


def test_weno_js_is_essentially_non_oscillatory(oracle):
    step = np.array([0, 0, 0, 1, 1, 1], dtype=float)
    vm_, vp_ = _weno(oracle, step, step)
    assert abs(vm_) < 1e-12          # the smooth left stencil wins
    assert abs(vp_ - 1.0) < 1e-12    # the smooth right stencil wins


def test_weno_js_fifth_order_on_smooth_data(oracle):
    """Error of the reconstructed face value of a smooth function drops
    like h^5 (JS weights approach the ideal ones)."""
    errs = []
    for h in (0.1, 0.05, 0.025):
        x = (np.arange(6) - 2.5) * h          # face at x = 0
        # point values of the flux whose cell averages are sin(x+1):
        # the FD-WENO reconstructs H(x_f) with f = (1/h) int H
        H = lambda s: np.sin(s + 1.0)
        f = np.array([(np.cos(xi - h / 2 + 1.0) - np.cos(xi + h / 2 + 1.0)) / h for xi in x])
        vm_, _ = _weno(oracle, f, f)
        errs.append(abs(vm_ - H(0.0)))
    r1, r2 = np.log2(errs[0] / errs[1]), np.log2(errs[1] / errs[2])
    assert r1 > 4.5 and r2 > 4.5, errs


# ---------------------------------------------------- eigenvectors / flux
def _roe_state(thd, T_K=1300.0, seed=0, dir=0):
    from phd_amd.chemdata import Adim
    rng = np.random.default_rng(seed)
    Y = rng.random(9)
    Y /= Y.sum()
    u = rng.standard_normal(3) * 0.3
    return Y, u, T_K


def test_eigenvectors_left_right_inverse(oracle, thd):
    """el . er = I at random states in every direction (module_weno.f90:
    775-951): a dropped term or a wrong sign breaks the identity."""
    from oracle.oracle import make_thd
    t = make_thd(thd)
    rng = np.random.default_rng(4)
    for trial in range(20):
        Y = rng.random(9)
        Y /= Y.sum()
        uc = rng.standard_normal(3)
        Tn = rng.uniform(2.0, 15.0)
        Wi = float(np.dot(Y, thd.Wc_i))
        cp = oracle.lib.ora_cp_scalar(C.byref(t), Tn, _dp(Y))
        gam = 1.0 / (1.0 - thd.gam2 * Wi / cp)
        cs = np.sqrt(gam * Tn * Wi)
        ha = np.zeros(9)
        oracle.lib.ora_ha_scalar(C.byref(t), Tn, _dp(ha))
        psi = -ha + (gam / (gam - 1.0)) * Tn * thd.Wc_i
        ht = float(np.dot(ha, Y)) + 0.5 * float(np.dot(uc, uc))
        for d in range(3):
            el = np.zeros((14, 14))
            er = np.zeros((14, 14))
            oracle.lib.ora_eigvecs(d, C.byref(t), _dp(np.ascontiguousarray(uc)),
                                   _dp(np.ascontiguousarray(Y)), _dp(np.ascontiguousarray(psi)),
                                   C.c_double(ht), C.c_double(cs), C.c_double(gam), _dp(el),
                                   _dp(er))
            prod = el @ er
            assert np.max(np.abs(prod - np.eye(14))) <= 1e-9 * np.max(np.abs(el)) * \
                np.max(np.abs(er)), (trial, d)


def _euler_flux_and_jacobian(thd, oracle, d, U):
    """Exact Euler flux F(U) of the reference's EOS and its finite-difference
    Jacobian, for a state vector U (14)."""
    from oracle.oracle import make_thd
    t = make_thd(thd)

    def temperature(U):
        # solve e(T) = e with Newton on the restated thermo
        rho = U[0]
        Y = U[5:] / rho
        e = U[4] / rho - 0.5 * np.dot(U[1:4], U[1:4]) / rho ** 2
        Wi = float(np.dot(Y, thd.Wc_i))
        T = 6.0
        ha = np.zeros(9)
        for _ in range(60):
            oracle.lib.ora_ha_scalar(C.byref(t), T, _dp(ha))
            F = np.dot(Y, ha) - T * Wi - e
            cp = oracle.lib.ora_cp_scalar(C.byref(t), T, _dp(np.ascontiguousarray(Y)))
            dT = F / (cp / thd.gam2 - Wi)
            T -= dT
            if abs(dT) < 1e-14 * T:
                break
        return T, Wi

    def flux(U):
        T, Wi = temperature(U)
        rho = U[0]
        un = U[1 + d] / rho
        P = rho * T * Wi
        F = U * un
        F[0] = U[1 + d]
        F[1 + d] += P
        F[4] = un * (U[4] + P)
        return F

    F0 = flux(U)
    J = np.zeros((14, 14))
    for j in range(14):
        h = 1e-6 * max(abs(U[j]), 1e-3 * abs(U[0]))
        Up, Um = U.copy(), U.copy()
        Up[j] += h
        Um[j] -= h
        J[:, j] = (flux(Up) - flux(Um)) / (2 * h)
    return F0, J, temperature(U)


def test_eigensystem_reconstructs_flux_jacobian(oracle, thd):
    """R Lambda L = dF/dU at a state (the Roe average of two equal states is
    the state itself).  The species equations are closed with the last
    species eliminated, so the check acts on the 13 independent directions
    that keep sum(rho Y) = rho (dU with sum of species increments = d rho)."""
    from oracle.oracle import make_thd
    t = make_thd(thd)
    rng = np.random.default_rng(9)
    for trial in range(3):
        Y = rng.random(9) + 0.05
        Y /= Y.sum()
        rho = 1.3
        uc = rng.standard_normal(3) * 0.5
        Tn = 7.0
        ha = np.zeros(9)
        oracle.lib.ora_ha_scalar(C.byref(t), Tn, _dp(ha))
        Wi = float(np.dot(Y, thd.Wc_i))
        e = float(np.dot(Y, ha)) - Tn * Wi
        U = np.concatenate([[rho], rho * uc, [rho * (e + 0.5 * np.dot(uc, uc))], rho * Y])
        for d in range(3):
            F0, J, (T, Wi2) = _euler_flux_and_jacobian(thd, oracle, d, U)
            cp = oracle.lib.ora_cp_scalar(C.byref(t), T, _dp(np.ascontiguousarray(Y)))
            gam = 1.0 / (1.0 - thd.gam2 * Wi2 / cp)
            cs = np.sqrt(gam * T * Wi2)
            oracle.lib.ora_ha_scalar(C.byref(t), T, _dp(ha))
            psi = -ha + (gam / (gam - 1.0)) * T * thd.Wc_i
            ht = float(np.dot(ha, Y)) + 0.5 * float(np.dot(uc, uc))
            el = np.zeros((14, 14))
            er = np.zeros((14, 14))
            oracle.lib.ora_eigvecs(d, C.byref(t), _dp(np.ascontiguousarray(uc)),
                                   _dp(np.ascontiguousarray(Y)), _dp(np.ascontiguousarray(psi)),
                                   C.c_double(ht), C.c_double(cs), C.c_double(gam), _dp(el),
                                   _dp(er))
            un = uc[d]
            lam = np.full(14, un)
            lam[0], lam[2] = un - cs, un + cs
            A = er @ np.diag(lam) @ el
            # tangent directions of the manifold rho = sum rho Y_k: the four
            # momentum/energy unit vectors and "add mass of species k"
            B = np.zeros((14, 13))
            for q in range(4):
                B[1 + q, q] = 1.0
            for k in range(9):
                B[0, 4 + k] = 1.0
                B[5 + k, 4 + k] = 1.0
            err = np.max(np.abs((A - J) @ B)) / np.max(np.abs(J @ B))
            assert err <= 1e-6, (trial, d, err)


def test_free_stream_preservation(oracle, adi, thd):
    """A uniform state gives f = 0 exactly (Euler and viscous)."""
    from phd_amd import ics
    box = ics.Box(12, 12, 1, 3, 2, (1e-3, 1e-3, 1e-3))
    T = np.full((1, 12, 12), 1100.0)
    Y = np.broadcast_to(np.array([0.01, 0.02, 0.0, 0.2, 0.01, 0.05, 0.0, 0.0, 0.71])[:, None, None,
                                  None], (9, 1, 12, 12)).copy()
    Y /= Y.sum(axis=0)
    u = np.zeros((3, 1, 12, 12))
    u[0] += 30.0
    u[1] -= 12.0
    v, Tn = ics.conserved_from_primitive(box, adi, thd, T, 101325.0, Y, u)
    run = OracleRun(oracle, adi, thd, box, v, Tn)
    run.prim_all()
    run.sensor[box.interior()] = 1
    f = np.zeros((14,) + box.shape)
    for d in range(2):
        oracle.lib.ora_euler_llf(d, C.byref(run.g), C.byref(run.t), P(run.d_i[d]), P(run.T),
                                 P(run.W_i), P(run.cp), P(run.ha),
                                 run.sensor.ctypes.data_as(C.POINTER(C.c_uint8)), 1, P(run.v),
                                 P(f))
    oracle.lib.ora_viscflux(C.byref(run.g), C.byref(run.t), C.byref(run.a), P(run.d_i[0]),
                            P(run.d_i[1]), P(run.d_i[2]), P(run.T), P(run.W_i), P(run.ha),
                            P(run.v), P(f))
    I = (slice(None),) + box.interior()
    # a uniform state has exactly equal stencil values: every flux
    # difference is F - F = 0 bit for bit
    assert np.max(np.abs(f[I])) <= 1e-30


@pytest.mark.parametrize("kind,n", [("2d", 20), ("3d", 10)])
def test_discrete_conservation(oracle, adi, thd, kind, n):
    """On a periodic box the flux differences telescope: sum over cells of
    f / d_i vanishes to round-off for every equation (Euler + viscous)."""
    box, v, T = small_state(adi, thd, kind, n)
    run = OracleRun(oracle, adi, thd, box, v, T)
    run.prim_all()
    run.sensor[box.interior()] = 1
    f = np.zeros((14,) + box.shape)
    for d in range(box.ndim):
        oracle.lib.ora_euler_llf(d, C.byref(run.g), C.byref(run.t), P(run.d_i[d]), P(run.T),
                                 P(run.W_i), P(run.cp), P(run.ha),
                                 run.sensor.ctypes.data_as(C.POINTER(C.c_uint8)), 1, P(run.v),
                                 P(f))
    oracle.lib.ora_viscflux(C.byref(run.g), C.byref(run.t), C.byref(run.a), P(run.d_i[0]),
                            P(run.d_i[1]), P(run.d_i[2]), P(run.T), P(run.W_i), P(run.ha),
                            P(run.v), P(f))
    I = (slice(None),) + box.interior()
    tot = f[I].sum(axis=(1, 2, 3))
    mag = np.abs(f[I]).sum(axis=(1, 2, 3)) + 1e-300
    assert np.max(np.abs(tot) / mag) <= 1e-12


def test_derivative_stencil_exact_on_sextics(oracle):
    """alpha6 (module_deriv.f90:37-40) differentiates polynomials of degree
    <= 6 exactly and has 6th-order error on sin."""
    n = 40
    x = np.linspace(-1.0, 1.0, n)
    h = x[1] - x[0]
    for p in range(7):
        a = x ** p
        out = np.zeros(n)
        oracle.lib.ora_deriv7_line(n, _dp(a), C.c_double(1.0 / h), _dp(out))
        ref = p * x ** (p - 1) if p else np.zeros(n)
        assert np.max(np.abs(out[3:-3] - ref[3:-3])) <= 1e-10
    errs = []
    for n in (20, 40):
        x = np.linspace(0, 1, n)
        h = x[1] - x[0]
        out = np.zeros(n)
        oracle.lib.ora_deriv7_line(n, _dp(np.sin(3 * x)), C.c_double(1.0 / h), _dp(out))
        errs.append(np.max(np.abs(out[3:-3] - 3 * np.cos(3 * x[3:-3]))))
    assert np.log2(errs[0] / errs[1]) > 5.5


def test_transport_closed_forms(oracle, thd):
    """Mixture-averaged transport (prim_vis_vars): a single-species gas has
    mu = Ku exp(CofEta . lnT) and lambda = Kl (1/lam + lam)."""
    lib = oracle.lib
    from oracle.oracle import make_thd
    t = make_thd(thd)
    TK = 1200.0
    Tn = TK / thd.Tinf
    X = np.zeros(9)
    X[8] = 1.0  # pure N2
    dm = np.zeros(9)
    mu, ct = C.c_double(), C.c_double()
    lib.ora_transport_cell(C.byref(t), 1.0, Tn, thd.Wc_i[8], _dp(X), _dp(dm), C.byref(mu),
                           C.byref(ct))
    lt = np.log(TK)
    eta = np.exp(sum(thd.cofeta[8][q] * lt ** q for q in range(4)))
    lam = np.exp(sum(thd.coflam[8][q] * lt ** q for q in range(4)))
    assert mu.value == pytest.approx(0.1 / thd.muinf * eta, rel=1e-13)
    assert ct.value == pytest.approx(0.5e-5 / thd.laminf * (lam + lam), rel=1e-13)


def test_timestep_convective_floor_quirk(oracle, adi, thd):
    """timestep starts the convective maximum at dtc = 1 (module_solver.f90:
    250-292): on a coarse grid the convective step is capped at CFL."""
    box, v, T = small_state(adi, thd, "2d", 24)
    run = OracleRun(oracle, adi, thd, box, v, T)
    run.prim_all()
    opts = run.opts(vis=0)
    ts = np.zeros(4)
    dtmin = C.c_double(1.0)
    oracle.lib.ora_timestep(C.byref(run.g), C.byref(run.t), C.byref(run.a), C.byref(opts),
                            P(run.d_i[0]), P(run.d_i[1]), P(run.d_i[2]), P(run.cp), P(run.W_i),
                            P(run.T), P(run.v), C.c_double(0.0), P(ts), C.byref(dtmin))
    assert ts[0] == 0.5
