"""The chemistry integrator's source (phd_amd/csrc/chem_device.h), built for
the host with g++, against the reference's own compiled DVODE
(oracle/_ref): both kinetics forms, the C3 cell recipe, the reactor
substep of izem_psr_batch.  Bars as in the GPU test: |dT| <= 2e-4 T and
|dY_k| <= 2e-4 |Y_k| + 1e-8 per cell, every cell integrated (ok)."""
import os

import numpy as np
import pytest

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libpsrref.so")


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")
@pytest.mark.parametrize("kinetics", [0, 1])
def test_host_integrator_vs_reference_dvode(host_bdf, thd, kinetics):
    from oracle.oracle import RefChem
    from phd_amd import ics
    rc = RefChem()
    n = 48
    T, p, Y = ics.config3_cells(thd, n, seed=11)
    Wmix = 1.0 / (Y / (np.array(thd.W) * 1000.0)).sum(axis=1)
    rho = (p * 10.0) * Wmix / (8.314510e7 * T)
    dt, rtol, atol = 1e-6, 1e-6, 1e-12
    Th, Yh = T.copy(), np.ascontiguousarray(Y.copy())
    nst, ok = np.zeros(n, np.int32), np.zeros(n, np.int32)
    host_bdf.host_psr(kinetics, n, Th.ctypes.data, rho.ctypes.data, Yh.ctypes.data, dt, rtol,
                      atol, 50, nst.ctypes.data, ok.ctypes.data)
    assert ok.all()
    errT, errY, steps = 0.0, 0.0, []
    for c in range(n):
        Tr, Yr, ist, ns, nfe, nje = rc.psr(T[c], rho[c], Y[c], dt, rtol, atol)
        assert ist == 2
        steps.append(ns)
        errT = max(errT, abs(Th[c] - Tr) / Tr)
        errY = max(errY, np.max(np.abs(Yh[c] - Yr) / (2e-4 * np.abs(Yr) + 1e-8)))
    print(f"host BDF vs DVODE: max |dT|/T {errT:.2e}, Y err / bound {errY:.2f}, "
          f"steps {nst.mean():.1f} (DVODE {np.mean(steps):.1f})")
    assert errT <= 2e-4
    assert errY <= 1.0
    assert np.all(np.abs(Yh.sum(axis=1) - 1.0) <= 1e-12)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")
def test_host_integrator_global_error_like_dvode(host_bdf, thd):
    """At the same rtol/atol the integrator's global error is like DVODE's:
    both measured against DVODE run at rtol 1e-11, atol 1e-17, in DVODE's
    error norm (tests/chembars.py), on C3 cells at two substeps and on the
    cells of a 3-D decaying-HIT state over an igniting 2e-4 s substep."""
    from oracle.oracle import Oracle, RefChem
    from phd_amd import chemdata as cd, ics
    from tests import chembars
    from tests.flowcase import OracleRun, small_state
    rc = RefChem()
    cases = []
    for seed, n, dt in ((11, 64, 1e-6), (7, 48, 1e-7), (5, 48, 1e-5)):
        T, p, Y = ics.config3_cells(thd, n, seed=seed)
        Wmix = 1.0 / (Y / (np.array(thd.W) * 1000.0)).sum(axis=1)
        cases.append((T, (p * 10.0) * Wmix / (8.314510e7 * T), Y, dt))
    adi = cd.Adim()
    box, v, T = small_state(adi, thd, "3d", 10)
    run = OracleRun(Oracle(), adi, thd, box, v, T)
    run.prim_all()
    I = box.interior()
    rho = run.v[(0,) + I].ravel()
    cases.append((run.T[I].ravel() * adi.T_ref, rho * adi.rho_ref * 1e-3,
                  np.ascontiguousarray((run.v[(slice(5, 14),) + I] / run.v[(0,) + I]).reshape(9, -1).T),
                  2e-4))
    for T, rho, Y, dt in cases:
        n = len(T)
        Th, Yh = T.copy(), np.ascontiguousarray(Y.copy())
        nst, ok = np.zeros(n, np.int32), np.zeros(n, np.int32)
        host_bdf.host_psr(1, n, Th.ctypes.data, np.ascontiguousarray(rho).ctypes.data,
                          Yh.ctypes.data, dt, 1e-6, 1e-12, 50, nst.ctypes.data, ok.ctypes.data)
        assert ok.all()
        e = chembars.dvode_errors(rc, T, rho, Y, dt, Th, Yh)
        print(f"{n} cells, dt {dt:g}: {chembars.summary(e)}")
        chembars.assert_like_dvode(e)
