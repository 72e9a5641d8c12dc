"""GPU chemistry against the oracle and the reference's own compiled
CHEMKIN-II / DVODE (oracle/_ref), through the C ABI.

Bars:
  * rates (CKWYR): per cell max_k |dw_k| <= 1e-12 * max_k |w_k|; the
    kernel evaluates the CHEMKIN formulas with a different operation order
    (and device exp/log), so this is a tolerance, not a bit identity;
  * reactor substeps: the GPU BDF and DVODE are two solutions of the same
    ODE at rtol = 1e-6, atol = 1e-12.  Both are measured against DVODE at
    rtol 1e-11 / atol 1e-17 (the exact solution) in DVODE's own error norm;
    the GPU's error must be within twice DVODE's at the same tolerances
    (largest and mean over the cells), and each component's (T, Y_k) mean
    error within three times (tests/chembars.py).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from tests import chembars

pytestmark = pytest.mark.gpu

RU = 8.314510e7


def _solver(mech, n=8, rtol=1e-6, atol=1e-12):
    from phd_amd.ics import Box
    from phd_amd.solver import Options, Solver
    return Solver(Box(n, n, 1, 3, 2, (1, 1, 1)), Options(rtol=rtol, atol=atol), mech=mech)


def _cells(thd, n, seed=11):
    from phd_amd import ics
    T, p, Y = ics.config3_cells(thd, n, seed=seed)
    Wmix = 1.0 / (Y / (np.array(thd.W) * 1000.0)).sum(axis=1)
    rho = (p * 10.0) * Wmix / (RU * T)  # g/cm^3 (p in dyn/cm^2)
    return T, rho, Y


def _refchem():
    from oracle.oracle import RefChem
    return RefChem()


def test_ckwyr_gpu_vs_reference(mech, thd):
    import torch
    rc = _refchem()
    s = _solver(mech)
    rng = np.random.default_rng(0)
    n = 512
    T = rng.uniform(300.0, 3200.0, n)
    rho = 10 ** rng.uniform(-5.0, -2.5, n)
    Y = rng.random((n, 9)) ** 3
    Y /= Y.sum(axis=1, keepdims=True)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    w = torch.zeros((n, 9), dtype=torch.float64, device="cuda:0")
    assert s.lib.izem_ckwyr(s.ctx, n, C.c_void_p(drho.data_ptr()), C.c_void_p(dT.data_ptr()),
                            C.c_void_p(dY.data_ptr()), C.c_void_p(w.data_ptr())) == 0
    s.sync()
    wg = w.cpu().numpy()
    worst, worst_ref = 0.0, 0.0
    for c in range(n):
        wr = rc.wdot(rho[c], T[c], Y[c])
        e = np.max(np.abs(wg[c] - wr)) / np.max(np.abs(wr))
        worst = max(worst, e)
    print("ckwyr worst per-cell rel err", worst)
    assert worst <= 1e-12


SRI_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                       "_ref", "sri")
SRI_MECH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "h2_air_sri_rev.inp")
needs_sri = pytest.mark.skipif(not os.path.exists(os.path.join(SRI_DIR, "chem.bin")),
                               reason="oracle/_ref/sri not built")


@needs_sri
def test_front_end_mechanism_on_the_gpu(thd):
    """A mechanism with SRI, 4-value Troe, a (+N2) collider, REV and an
    irreversible reaction (tests/golden/h2_air_sri_rev.inp) loaded through
    izem_chem_setup: the context falls back to the table-driven kinetics (the
    straight-line code is for the bundled mechanism), izem_ckwyr matches the
    reference CKWYR on the same file, and izem_psr_batch matches the
    reference DVODE on igniting cells."""
    import torch
    from oracle.oracle import RefChem
    from phd_amd import chemdata as cd, ics
    m = cd.read_mechanism(SRI_MECH)
    rc = RefChem(SRI_DIR)
    s = _solver(m)
    assert s.lib.izem_chem_tune(s.ctx, 1, 4, 50, 1) == -1  # no straight-line code for it
    rng = np.random.default_rng(6)
    n = 512
    T = rng.uniform(300.0, 3200.0, n)
    rho = 10 ** rng.uniform(-5.0, -2.5, n)
    Y = rng.random((n, 9)) ** 3
    Y /= Y.sum(axis=1, keepdims=True)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    w = torch.zeros((n, 9), dtype=torch.float64, device="cuda:0")
    assert s.lib.izem_ckwyr(s.ctx, n, C.c_void_p(drho.data_ptr()), C.c_void_p(dT.data_ptr()),
                            C.c_void_p(dY.data_ptr()), C.c_void_p(w.data_ptr())) == 0
    s.sync()
    wg = w.cpu().numpy()
    worst = max(np.max(np.abs(wg[c] - rc.wdot(rho[c], T[c], Y[c]))) /
                np.max(np.abs(rc.wdot(rho[c], T[c], Y[c]))) for c in range(n))
    print(f"front-end mechanism: ckwyr worst per-cell rel err {worst:.2e}")
    assert worst <= 1e-12
    nc = 64
    Tc, pc, Yc = ics.config3_cells(thd, nc, seed=12)
    Wmix = 1.0 / (Yc / (np.array(thd.W) * 1000.0)).sum(axis=1)
    rhoc = (pc * 10.0) * Wmix / (8.314510e7 * Tc)
    dTc, drc, dYc = d(Tc), d(rhoc), d(Yc)
    nst = torch.zeros(nc, dtype=torch.int32, device="cuda:0")
    assert s.lib.izem_psr_batch(s.ctx, nc, C.c_void_p(dTc.data_ptr()), C.c_void_p(drc.data_ptr()),
                                C.c_void_p(dYc.data_ptr()), 1e-6, C.c_void_p(nst.data_ptr())) == 0
    s.sync()
    assert s.status() == 0
    Tg, Yg = dTc.cpu().numpy(), dYc.cpu().numpy()
    e = chembars.dvode_errors(rc, Tc, rhoc, Yc, 1e-6, Tg, Yg)
    print(f"front-end mechanism: psr {chembars.summary(e)}")
    chembars.assert_like_dvode(e)
    assert np.max(np.abs(Tg - Tc)) > 1.0


def test_element_conservation_of_rates(mech):
    """sum_k ncf[k,e] W_k^-1... : elements are conserved by every reaction, so
    sum_k ncf[k, e] * wdot_k = 0 to round-off relative to the largest term."""
    import torch
    s = _solver(mech)
    rng = np.random.default_rng(1)
    n = 256
    T = rng.uniform(800.0, 3000.0, n)
    rho = 10 ** rng.uniform(-5.0, -3.0, n)
    Y = rng.random((n, 9))
    Y /= Y.sum(axis=1, keepdims=True)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    w = torch.zeros((n, 9), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    s.lib.izem_ckwyr(s.ctx, n, C.c_void_p(drho.data_ptr()), C.c_void_p(dT.data_ptr()),
                     C.c_void_p(dY.data_ptr()), C.c_void_p(w.data_ptr()))
    s.sync()
    wg = w.cpu().numpy()
    ncf = mech.ncf  # [k][element]
    for e in range(ncf.shape[1]):
        tot = wg @ ncf[:, e]
        scale = np.abs(wg) @ np.abs(ncf[:, e]) + 1e-300
        assert np.max(np.abs(tot) / scale) <= 1e-12


@pytest.mark.parametrize("kinetics,waves", [(0, 1), (0, 2), (0, 4), (1, 1), (2, 1)])
@pytest.mark.parametrize("rtol,atol", [(1e-6, 1e-12)])
def test_psr_batch_vs_reference_dvode(mech, thd, rtol, atol, kinetics, waves):
    """Both kinetics forms (table-driven, straight-line generated) and two
    register budgets of the persistent kernel against DVODE."""
    import torch
    rc = _refchem()
    s = _solver(mech, rtol=rtol, atol=atol)
    assert s.lib.izem_chem_tune(s.ctx, waves, 4 * waves, 50, kinetics) == 0
    n = 96
    T, rho, Y = _cells(thd, n)
    dt = 1.0e-6
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    nst = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    assert s.lib.izem_psr_batch(s.ctx, n, C.c_void_p(dT.data_ptr()), C.c_void_p(drho.data_ptr()),
                                C.c_void_p(dY.data_ptr()), dt, C.c_void_p(nst.data_ptr())) == 0
    s.sync()
    assert s.status() == 0
    Tg, Yg = dT.cpu().numpy(), dY.cpu().numpy()
    e = chembars.dvode_errors(rc, T, rho, Y, dt, Tg, Yg, rtol, atol)
    print(f"psr kinetics {kinetics} waves {waves}: {chembars.summary(e)}, "
          f"steps gpu {nst.cpu().numpy().mean():.1f}")
    chembars.assert_like_dvode(e)


def test_psr_conserves_mass_and_energy(mech, thd):
    """Constant-volume reactor invariants: sum Y = 1 (linear, kept to
    round-off by the BDF) and the mixture internal energy sum_k Y_k u_k(T),
    conserved by the ODE and kept by the integration to the tolerance level:
    |du| <= 10 * rtol * cv T (the reference DVODE's drift is printed beside)."""
    import torch
    rc = _refchem()
    s = _solver(mech)
    n = 64
    T, rho, Y = _cells(thd, n, seed=5)
    dt = 1e-5
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    s.lib.izem_psr_batch(s.ctx, n, C.c_void_p(dT.data_ptr()), C.c_void_p(drho.data_ptr()),
                         C.c_void_p(dY.data_ptr()), dt, None)
    s.sync()
    Tg, Yg = dT.cpu().numpy(), dY.cpu().numpy()

    def u_mix(TT, YY):
        h = rc.hms(TT)
        return float(np.dot(YY, h - RU * TT / (mech.wt)))

    worst, worst_ref = 0.0, 0.0
    for c in range(n):
        assert abs(Yg[c].sum() - 1.0) <= 1e-12
        Tr, Yr, ist, *_ = rc.psr(T[c], rho[c], Y[c], dt, 1e-6, 1e-12)
        scale = abs(rc.cvbs(T[c], Y[c]) * T[c])
        u0 = u_mix(T[c], Y[c])
        du_gpu = abs(u_mix(Tg[c], Yg[c]) - u0) / scale
        du_ref = abs(u_mix(Tr, Yr) - u0) / scale
        worst = max(worst, du_gpu)
        worst_ref = max(worst_ref, du_ref)
        assert du_gpu <= 10.0 * 1e-6, (c, du_gpu, du_ref)
    print(f"max relative internal-energy drift: gpu {worst:.2e}, dvode {worst_ref:.2e}")


def test_reaction_dvode_vs_reference(oracle, adi, thd, mech):
    """reaction_selec / reaction_dvode (module_reaction.f90:47-90, 102-304)
    on a 3-D 12^3 state of the decaying-HIT recipe (T = 1000 K +- 50 K, so
    the 125 cells of the reacting box all differ): the GPU path against the
    reference's own PSR_CHEMKIN loop (compiled DVODE) applied cell by cell
    on the same reacting box, followed by the oracle's upd_prim_var_woT,
    comm_cons and ghost Newton.  The half step is 2e-4 s, over which the
    hottest cells ignite (DTmax ~ 1780 K); the bars are DVODE's own errors
    against DVODE at rtol 1e-11 (chembars)."""
    import torch
    from phd_amd.solver import Options, Solver
    from tests.flowcase import OracleRun, oracle_reaction_dvode, small_state
    rc = _refchem()
    box, v, T = small_state(adi, thd, "3d", 12)
    run = OracleRun(oracle, adi, thd, box, v, T)
    run.prim_all()
    rtol, atol = 1e-6, 1e-12
    s = Solver(box, Options(rtol=rtol, atol=atol), adi=adi, thd=thd, mech=mech)
    s.v.copy_(torch.from_numpy(run.v)); s.T.copy_(torch.from_numpy(run.T))
    s.W_i.copy_(torch.from_numpy(run.W_i)); s.cp.copy_(torch.from_numpy(run.cp))
    s.ha.copy_(torch.from_numpy(run.ha))
    dtmin = 4.0e-4 / adi.time_ref  # nondimensional step; the half step is 2e-4 s
    dt = torch.tensor([dtmin], dtype=torch.float64, device="cuda:0")
    s.reaction_selec(dt.data_ptr(), 0.5)
    s.sync()
    assert s.status() == 0
    DTmax_gpu = s.time_state()[5]
    runs = {}
    for name, tol in (("exact", chembars.TIGHT), ("dvode", (rtol, atol))):
        r = OracleRun(oracle, adi, thd, box, run.v, run.T)
        r.W_i[:], r.cp[:], r.ha[:] = run.W_i, run.cp, run.ha
        runs[name] = (r, oracle_reaction_dvode(r, rc, adi, 0.5 * dtmin, *tol))
    vg, Tg = s.v.cpu().numpy(), s.T.cpu().numpy()
    I = box.interior()
    (rt, DTt), (rd, DTd) = runs["exact"], runs["dvode"]

    def z(vv, TT):
        return np.concatenate([TT[I].reshape(-1, 1) * adi.T_ref,
                               (vv[(slice(5, 14),) + I] / vv[(0,) + I]).reshape(9, -1).T], axis=1)

    e = chembars.compare(z(vg, Tg), z(rd.v, rd.T), z(rt.v, rt.T), rtol, atol)
    print(f"reaction_dvode: {chembars.summary(e)}; DTmax gpu {DTmax_gpu:.6e} K, DVODE "
          f"{DTd:.6e} K, exact {DTt:.6e} K")
    assert DTt > 1000.0
    chembars.assert_like_dvode(e)
    # DTmax is one cell's value: twice DVODE's distance plus ten tolerance
    # units of that cell's end temperature
    Tend = np.max(rt.T[I]) * adi.T_ref
    assert abs(DTmax_gpu - DTt) <= 2.0 * abs(DTd - DTt) + 10.0 * rtol * Tend


def test_straight_line_kinetics_one_wave_only(mech):
    """The straight-line kinetics exist at one wave per SIMD only: asking for
    them at two or four is refused, and 'straight-line when available' at
    those budgets runs the table-driven kinetics."""
    s = _solver(mech)
    for w in (2, 4):
        assert s.lib.izem_chem_tune(s.ctx, w, 4 * w, 50, 1) == -1
        assert s.lib.izem_chem_tune(s.ctx, w, 4 * w, 50, 2) == -1
        assert s.lib.izem_chem_tune(s.ctx, w, 4 * w, 50, -1) == 0
    assert s.lib.izem_chem_tune(s.ctx, 1, 4, 50, 1) == 0
    assert s.lib.izem_chem_tune(s.ctx, 1, 4, 50, 2) == 0
    assert s.lib.izem_chem_tune(s.ctx, 1, 4, 50, 3) == -1


def test_setup_self_check_is_clean(mech):
    """izem_chem_setup integrates 64 ignition cells with both kinetics forms
    to check the straight-line build against the table-driven one: on this
    build they agree (the straight-line kinetics stay selectable), and the
    check leaves no trace in the work counters or the status word."""
    s = _solver(mech)
    out = np.zeros(4, dtype=np.uint64)
    assert s.lib.izem_chem_stats(s.ctx, out.ctypes.data) == 0
    assert np.all(out == 0)
    assert s.status() == 0
    assert s.lib.izem_chem_tune(s.ctx, 1, 4, 50, 1) == 0


@pytest.mark.parametrize("kinetics,waves", [(0, 1), (0, 2), (0, 4), (1, 1), (2, 1)])
def test_variants_follow_host_build(mech, thd, host_bdf, kinetics, waves):
    """Every kernel variant integrates the same source as a g++ build of it;
    the two differ only in the last bits of exp/log, so most cells take the
    same steps and every end state agrees far inside the tolerance.  Bars:
    |dT|/T <= 1e-6 per cell, at least 80 % of the cells with the same step
    count, mean step count within 3 %.  A variant whose compiled code
    departs from its source (seen as steps doubled on some cells, or
    failures) fails here even when its end states stay within DVODE's."""
    import torch
    L = host_bdf
    n = 96
    T, rho, Y = _cells(thd, n)
    Th, Yh = T.copy(), np.ascontiguousarray(Y.copy())
    nst_h, ok_h = np.zeros(n, np.int32), np.zeros(n, np.int32)
    L.host_psr(kinetics, n, Th.ctypes.data, rho.ctypes.data, Yh.ctypes.data, 1e-6, 1e-6, 1e-12,
               50, nst_h.ctypes.data, ok_h.ctypes.data)
    assert ok_h.all()
    s = _solver(mech)
    assert s.lib.izem_chem_tune(s.ctx, waves, 4 * waves, 50, kinetics) == 0
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    nst = torch.zeros(n, dtype=torch.int32, device="cuda:0")
    assert s.lib.izem_psr_batch(s.ctx, n, C.c_void_p(dT.data_ptr()), C.c_void_p(drho.data_ptr()),
                                C.c_void_p(dY.data_ptr()), 1e-6, C.c_void_p(nst.data_ptr())) == 0
    s.sync()
    st = s.status()
    ng = nst.cpu().numpy()
    rel = np.abs(dT.cpu().numpy() - Th) / Th
    same = int((ng == nst_h).sum())
    print(f"kinetics {kinetics} waves {waves}: status {st}, max |dT|/T {np.nanmax(rel):.2e}, "
          f"same steps {same}/{n}, mean steps {ng.mean():.1f} (host {nst_h.mean():.1f})")
    assert st == 0
    assert np.all(rel <= 1e-6)
    assert same >= 0.8 * n
    assert abs(ng.mean() - nst_h.mean()) <= 0.03 * nst_h.mean()


@pytest.mark.parametrize("kinetics,waves", [(0, 1), (0, 2), (0, 4), (1, 1), (2, 1)])
def test_reaction_queue_variants_follow_host_build(oracle, adi, thd, mech, host_bdf, kinetics,
                                                   waves):
    """The reacting-step kernel (k_reaction_queue, behind reaction_selec) in
    every variant against the g++ build of the integrator on the same cells:
    a 2-D box whose reacting cells hold the C3 ignition recipe (T, p, Y),
    advanced over a 1e-6 s half step.  End temperatures within 1e-6
    relative, and the total BDF step count (the library's work counters)
    within 3 % of the host build's."""
    import torch
    from phd_amd import ics
    from phd_amd.solver import Options, Solver
    from tests.flowcase import OracleRun, small_state
    box, v, T = small_state(adi, thd, "2d", 32)
    ng, nx, ny = box.ng, box.nx, box.ny
    cells = [(ng, j + ng, i + ng) for j in range(ng, ny - ng - 1) for i in range(ng, nx - ng - 1)]
    Tc, pc, Yc = ics.config3_cells(thd, len(cells), seed=17)
    Wmix = 1.0 / (Yc / (np.array(thd.W) * 1000.0)).sum(axis=1)
    rho_c = (pc * 10.0) * Wmix / (RU * Tc) / (adi.rho_ref * 1.0e-3)  # nondimensional
    for q, c in enumerate(cells):
        v[(0,) + c] = rho_c[q]
        v[(slice(5, 14),) + c] = rho_c[q] * Yc[q]
        T[c] = Tc[q] / adi.T_ref
    run = OracleRun(oracle, adi, thd, box, v, T)
    run.prim_all()
    s = Solver(box, Options(), adi=adi, thd=thd, mech=mech)
    s.v.copy_(torch.from_numpy(run.v)); s.T.copy_(torch.from_numpy(run.T))
    s.W_i.copy_(torch.from_numpy(run.W_i)); s.cp.copy_(torch.from_numpy(run.cp))
    s.ha.copy_(torch.from_numpy(run.ha))
    assert s.lib.izem_chem_tune(s.ctx, waves, 4 * waves, 50, kinetics) == 0
    stats = np.zeros(4, dtype=np.uint64)
    s.lib.izem_chem_stats(s.ctx, stats.ctypes.data)  # clears the counters
    dtmin = 2.0e-6 / adi.time_ref
    dt = torch.tensor([dtmin], dtype=torch.float64, device="cuda:0")
    s.reaction_selec(dt.data_ptr(), 0.5)
    s.sync()
    st = s.status()
    assert st & 16 == 0  # no integration failure (the C3 cells need not be in EOS balance)
    s.lib.izem_chem_stats(s.ctx, stats.ctypes.data)
    # the same cells (the trimmed reacting box) through the host build
    n = len(cells)
    rho_cgs = adi.rho_ref * 1.0e-3
    T0 = np.array([run.T[c] * adi.T_ref for c in cells])
    rho = np.array([run.v[(0,) + c] * rho_cgs for c in cells])
    Y = np.ascontiguousarray([run.v[(slice(5, 14),) + c] / run.v[(0,) + c] for c in cells])
    Th = T0.copy()
    nst, ok = np.zeros(n, np.int32), np.zeros(n, np.int32)
    host_bdf.host_psr(kinetics, n, Th.ctypes.data, rho.ctypes.data, Y.ctypes.data,
                      0.5 * dtmin * adi.time_ref, 1e-6, 1e-12, 50, nst.ctypes.data,
                      ok.ctypes.data)
    assert ok.all()
    Tg = s.T.cpu().numpy()
    Tgc = np.array([Tg[c] * adi.T_ref for c in cells])
    rel = np.abs(Tgc - Th) / Th
    print(f"kinetics {kinetics} waves {waves}: {n} cells, max |dT|/T {rel.max():.2e}, "
          f"steps {int(stats[0])} (host {int(nst.sum())}), max |T - T0| {np.max(np.abs(Th - T0)):.1f} K")
    assert rel.max() <= 1e-6
    assert abs(float(stats[0]) - float(nst.sum())) <= 0.03 * float(nst.sum())


@pytest.mark.parametrize("n", [0, 1, 63, 65, 257])
def test_psr_batch_sizes(mech, thd, host_bdf, n):
    """Empty, single and ragged batches (not multiples of the 64-lane wave,
    fewer cells than lanes): every cell integrated once, end states as the
    host build's; n = 0 is a no-op that succeeds."""
    import torch
    s = _solver(mech)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    if n == 0:
        z = torch.zeros(1, dtype=torch.float64, device="cuda:0")
        assert s.lib.izem_psr_batch(s.ctx, 0, C.c_void_p(z.data_ptr()), C.c_void_p(z.data_ptr()),
                                    C.c_void_p(z.data_ptr()), 1e-6, None) == 0
        s.sync()
        assert s.status() == 0
        return
    T, rho, Y = _cells(thd, n, seed=23)
    dT, drho, dY = d(T), d(rho), d(Y)
    nst = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
    assert s.lib.izem_psr_batch(s.ctx, n, C.c_void_p(dT.data_ptr()), C.c_void_p(drho.data_ptr()),
                                C.c_void_p(dY.data_ptr()), 1e-6, C.c_void_p(nst.data_ptr())) == 0
    s.sync()
    assert s.status() == 0
    ng = nst.cpu().numpy()
    assert np.all(ng > 0)  # every cell taken from the queue exactly once and integrated
    Th, Yh = T.copy(), np.ascontiguousarray(Y.copy())
    nh, ok = np.zeros(n, np.int32), np.zeros(n, np.int32)
    host_bdf.host_psr(1, n, Th.ctypes.data, rho.ctypes.data, Yh.ctypes.data, 1e-6, 1e-6, 1e-12,
                      50, nh.ctypes.data, ok.ctypes.data)
    assert np.all(np.abs(dT.cpu().numpy() - Th) / Th <= 1e-6)


@pytest.mark.parametrize("n", [1, 65, 4096 + 37, (1 << 17) + 37])
def test_stiffness_order_changes_no_result(mech, thd, n):
    """With izem_chem_order on, izem_psr_batch hands cells out stiffest
    first once the cells outnumber the resident lanes (65536 at one wave
    per SIMD, four blocks per CU: the last size); the end states and per-
    cell step counts are bit-identical to index order, since every lane
    integrates its cell on its own.  Cells of an ignition ensemble (step
    counts from a few tens to several hundred) at single, ragged and larger
    counts; an out-of-range switch is refused."""
    import torch
    s = _solver(mech)
    assert s.lib.izem_chem_tune(s.ctx, 1, 4, 50, -1) == 0
    T, rho, Y = _cells(thd, n, seed=5)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    out = []
    for on in (1, 0):
        assert s.lib.izem_chem_order(s.ctx, on) == 0
        dT, drho, dY = d(T), d(rho), d(Y)
        nst = torch.full((n,), -1, dtype=torch.int32, device="cuda:0")
        assert s.lib.izem_psr_batch(s.ctx, n, C.c_void_p(dT.data_ptr()),
                                    C.c_void_p(drho.data_ptr()), C.c_void_p(dY.data_ptr()), 1e-6,
                                    C.c_void_p(nst.data_ptr())) == 0
        s.sync()
        assert s.status() == 0
        out.append((dT.cpu().numpy(), dY.cpu().numpy(), nst.cpu().numpy()))
        if on:
            _check_order_ran(s, n, n > 65536, nst=out[-1][2])
    (T1, Y1, n1), (T0, Y0, n0) = out
    assert np.all(n1 > 0)
    if n > 1000:
        assert n1.max() > 4 * n1.min()
    assert np.array_equal(n1, n0)
    assert np.array_equal(T1.view(np.uint64), T0.view(np.uint64))
    assert np.array_equal(Y1.view(np.uint64), Y0.view(np.uint64))
    assert s.lib.izem_chem_order(s.ctx, 2) == -1


def _check_order_ran(s, n, sorted_, nst=None):
    """The context's last stiffness sort (izem_chem_order_read): with more
    cells than resident lanes it covers all n cells, its permutation is one
    of 0..n-1 and the cells' slots are non-decreasing along it (stiffest
    first); otherwise no sort ran."""
    cap = max(n, 1)
    nn = C.c_long(-1)
    perm = np.zeros(cap, np.int32)
    bins = np.zeros(cap, np.int32)
    assert s.lib.izem_chem_order_read(s.ctx, cap, C.byref(nn), perm.ctypes.data,
                                      bins.ctypes.data) == 0
    if not sorted_:
        assert nn.value == 0
        return
    assert nn.value == n
    assert np.array_equal(np.sort(perm), np.arange(n))
    b = bins[perm]
    assert np.all(np.diff(b) >= 0) and b[0] < b[-1]
    if nst is not None:
        # the stiffest tenth of the queue takes more steps than the least stiff
        k = max(n // 10, 1)
        assert nst[perm[:k]].mean() > nst[perm[-k:]].mean()


def test_resetup_keeps_work_counters(mech, thd):
    """A second izem_chem_setup on a live context (new tolerances) runs its
    self-check on counters of its own: the work counters of the integrations
    so far survive it."""
    import torch
    s = _solver(mech)
    T, rho, Y = _cells(thd, 256, seed=9)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    dT, drho, dY = d(T), d(rho), d(Y)
    assert s.lib.izem_psr_batch(s.ctx, 256, C.c_void_p(dT.data_ptr()), C.c_void_p(drho.data_ptr()),
                                C.c_void_p(dY.data_ptr()), 1e-6, None) == 0
    s.sync()
    from phd_amd.solver import mech_flat
    mf = mech_flat(mech)
    assert s.lib.izem_chem_setup(s.ctx, mf.ctypes.data, mf.size, 1e-7, 1e-13, 1) == 0
    st = np.zeros(4, dtype=np.uint64)
    assert s.lib.izem_chem_stats(s.ctx, st.ctypes.data) == 0
    assert st[0] > 256 and st[1] > st[0] and st[2] >= 256, st
    assert s.status() == 0


@pytest.mark.parametrize("kinetics", [0, 1, 2])
def test_stiffness_order_flow_state(adi, thd, mech, kinetics):
    """izem_reaction_dvode with izem_chem_order on takes the cells of the
    reacting box stiffest first; the conserved state, temperature, DTmax and
    work counters are bit-identical to index order (2-D hot-spot state of
    288^2 cells, more reacting cells than the 65536 resident lanes)."""
    import torch
    from phd_amd.solver import Options, Solver
    from tests.flowcase import small_state
    box, v, T = small_state(adi, thd, "2d", 288)
    s = Solver(box, Options(rtol=1e-6, atol=1e-12), adi=adi, thd=thd, mech=mech)
    assert s.lib.izem_chem_tune(s.ctx, 1, 4, 50, kinetics) == 0
    dt = torch.tensor([2.0e-7 / adi.time_ref], dtype=torch.float64, device="cuda:0")
    out = []
    for on in (1, 0):
        s.load_state(v, T)
        assert s.lib.izem_chem_order(s.ctx, on) == 0
        st = np.zeros(4, dtype=np.uint64)
        s.lib.izem_chem_stats(s.ctx, st.ctypes.data)
        s.reaction_selec(dt.data_ptr(), 0.5)
        s.sync()
        assert s.status() == 0
        s.lib.izem_chem_stats(s.ctx, st.ctypes.data)
        out.append((s.v.cpu().numpy().copy(), s.T.cpu().numpy().copy(), s.time_state()[5],
                    st.copy()))
        if on:
            nr = (box.nx - 7) * (box.ny - 7)
            _check_order_ran(s, nr, True)
    (v1, T1, d1, s1), (v0, T0, d0, s0) = out
    assert d1 > 0.0 and d1 == d0
    assert np.array_equal(s1, s0)
    assert np.array_equal(v1.view(np.uint64), v0.view(np.uint64))
    assert np.array_equal(T1.view(np.uint64), T0.view(np.uint64))
