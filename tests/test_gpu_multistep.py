"""Multi-step parity of the main loop (devel-izem/src/izem.f90:242-389)
against the oracle, through the C ABI.

Flow only (config-1 recipe at 32^2, config-2 recipe at 16^3): 100 steps of
timestep -> rk3 on the GPU (Solver.step, dt kept on the device) and in the
oracle; the dt history and the state after steps 1 and 100 agree to 1e-12.
This runs the Newton warm start, the sticky shock sensor and the time-step
control across steps.

Reacting (config-1 recipe at 32^2 with the hot spot's peak at 1800 K, so
that it ignites within the run): 100 steps of timestep -> reaction(dt/2)
-> rk3 -> reaction(dt/2) with the reference's own DVODE in the oracle loop.
The GPU integrator and DVODE are different solutions of the chemistry at
rtol 1e-6, so the oracle loop is run twice, with DVODE at rtol 1e-6 and at
rtol 1e-11; the GPU trajectory must stay within twice DVODE's own distance
from the tight one (dt history, T, species).  CPR = 1 K makes the chemical
limit dt_4(4) = dt CPR / DTmax (module_solver.f90:378-389) set dt once the
hot spot ignites, so the DTmax -> dt feedback is run; so is the
chemistry's update of T without rho e (module_reaction.f90:278).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.flowcase import OracleRun, P, oracle_reaction_dvode, small_state

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _relerr(a, b, box):
    I = (slice(None),) + box.interior()
    a, b = a[I], b[I]
    s = np.max(np.abs(b), axis=(1, 2, 3))
    d = np.max(np.abs(a - b), axis=(1, 2, 3))
    return np.where(s > 0, d / np.where(s > 0, s, 1.0), d)


def _upload(s, run):
    import torch
    for name in ("v", "T", "W_i", "cp", "ha", "sensor"):
        getattr(s, name).copy_(torch.from_numpy(getattr(run, name)))
    torch.cuda.synchronize()


def _oracle_timestep(run, opts, dtmin, DTmax):
    ts = np.zeros(4)
    d = C.c_double(dtmin)
    run.o.lib.ora_timestep(C.byref(run.g), C.byref(run.t), C.byref(run.a), C.byref(opts),
                           P(run.d_i[0]), P(run.d_i[1]), P(run.d_i[2]), P(run.cp), P(run.W_i),
                           P(run.T), P(run.v), C.c_double(DTmax), P(ts), C.byref(d))
    return d.value


def _oracle_rk3(run, opts, dt):
    st = run.o.lib.ora_rk3(C.byref(run.g), C.byref(run.t), C.byref(run.a), C.byref(opts),
                           C.c_double(dt), P(run.d_i[0]), P(run.d_i[1]), P(run.d_i[2]),
                           P(run.T), P(run.W_i), P(run.cp), P(run.ha),
                           run.sensor.ctypes.data_as(C.POINTER(C.c_uint8)), P(run.v), None)
    assert st == 0


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("kind,n", [("2d", 32), ("3d", 16)])
def test_flow_multistep_parity(oracle, adi, thd, kind, n, mode):
    from phd_amd.solver import Options, Solver
    steps = 100
    box, v, T = small_state(adi, thd, kind, n)
    run = OracleRun(oracle, adi, thd, box, v, T)
    run.prim_all()
    s = Solver(box, Options(reaction=0), adi=adi, thd=thd)
    assert s.lib.izem_flux_mode(s.ctx, mode) == 0
    _upload(s, run)
    s.set_time_state([1.0, 1.0, 1.0, 1.0, 1.0, 0.0])
    opts = run.opts(vis=1)
    dtmin = 1.0
    worst_dt = 0.0
    for it in range(1, steps + 1):
        dtmin = _oracle_timestep(run, opts, dtmin, 0.0)
        _oracle_rk3(run, opts, dtmin)
        s.step()
        dtg = s.time_state()[4]
        worst_dt = max(worst_dt, abs(dtg - dtmin) / dtmin)
        if it in (1, steps):
            assert s.status() == 0
            e = _relerr(s.v.cpu().numpy(), run.v, box)
            I = box.interior()
            eT = np.max(np.abs(s.T.cpu().numpy()[I] - run.T[I]) / run.T[I])
            print(f"{kind} {n} mode {mode}: step {it}: state max rel err {e.max():.2e}, T {eT:.2e}, "
                  f"dt {worst_dt:.2e}")
            assert e.max() <= TOL and eT <= TOL
            assert np.array_equal(s.sensor.cpu().numpy(), run.sensor)
    assert worst_dt <= TOL


def test_reacting_multistep_dvode_bars(oracle, adi, thd, mech):
    import torch  # noqa: F401
    from oracle.oracle import RefChem
    from phd_amd.solver import Options, Solver
    from tests import chembars
    rc = RefChem()
    steps, n = 100, 32
    rtol, atol = 1e-6, 1e-12
    from phd_amd import ics
    box, v, T = ics.config1_state(adi, thd, n=n, seed=7, T_peak=1800.0)
    run0 = OracleRun(oracle, adi, thd, box, v, T)
    run0.prim_all()
    s = Solver(box, Options(rtol=rtol, atol=atol, CPR=1.0), adi=adi, thd=thd, mech=mech)
    _upload(s, run0)
    s.set_time_state([1.0, 1.0, 1.0, 1.0, 1.0, 0.0])
    gpu_dt = []
    for it in range(steps):
        s.step()
        gpu_dt.append(s.time_state()[4])
    assert s.status() == 0
    traj = {}
    for name, tol in (("exact", chembars.TIGHT), ("dvode", (rtol, atol))):
        r = OracleRun(oracle, adi, thd, box, run0.v, run0.T)
        r.W_i[:], r.cp[:], r.ha[:] = run0.W_i, run0.cp, run0.ha
        opts = r.opts(vis=1, reaction=1, CPR=1.0)
        dtmin, DTmax, dts = 1.0, 0.0, []
        for it in range(steps):
            dtmin = _oracle_timestep(r, opts, dtmin, DTmax)
            dts.append(dtmin)
            oracle_reaction_dvode(r, rc, adi, 0.5 * dtmin, *tol)
            _oracle_rk3(r, opts, dtmin)
            DTmax = oracle_reaction_dvode(r, rc, adi, 0.5 * dtmin, *tol)
        traj[name] = (r, np.array(dts))
    (rt, dt_t), (rd, dt_d) = traj["exact"], traj["dvode"]
    gpu_dt = np.array(gpu_dt)
    I = box.interior()
    Tg = s.T.cpu().numpy()
    vg = s.v.cpu().numpy()
    e_dt = (np.max(np.abs(gpu_dt - dt_t) / dt_t), np.max(np.abs(dt_d - dt_t) / dt_t))

    def z(vv, TT):
        return np.concatenate([TT[I].reshape(-1, 1) * adi.T_ref,
                               (vv[(slice(5, 14),) + I] / vv[(0,) + I]).reshape(9, -1).T], axis=1)

    e = chembars.compare(z(vg, Tg), z(rd.v, rd.T), z(rt.v, rt.T), rtol, atol)
    ign = np.max(rt.T[I]) * adi.T_ref
    assert np.min(dt_t) < 0.5 * np.max(dt_t)       # the chemical limit set dt
    print(f"reacting 32^2 x {steps}: peak T {ign:.1f} K, time {dt_t.sum() * adi.time_ref:.3e} s; "
          f"dt history {e_dt[0]:.2e} (DVODE {e_dt[1]:.2e}); {chembars.summary(e)}")
    assert e_dt[0] <= 2.0 * e_dt[1] + 1e-12
    chembars.assert_like_dvode(e)
