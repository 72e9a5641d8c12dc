"""The oracle against the reference's own trajectory of its 2-D Sod shock
tube (devel-izem/testcases/sod; the restart files, times_stat.out, grid.dat
and tranfit.out are copied under tests/golden).

From ini_sod's initial state the oracle built for 2 species runs the
reference's main loop (WENO7, Adams-Shariff sensor at 0.15 with JS weights
where it fired, extrapolated x faces with the one-sided derivative closures,
periodic y, viscous fluxes, CFL 0.5) until the first and second statistics
times (647 and 712 iterations) and is compared with the reference's
restart_00001.h5 and restart_00002.h5 and with its times_stat.out.  The run
being restated used 2 MPI ranks along x; two consequences for the sensor are
restated too: each rank's sweep stops one cell short of its own last cell,
and a rank sees only its own interior cells' flags (the sensor is neither
set on ghosts nor exchanged).  Without that the fields differ by 2e-2; with
it they agree to 5e-8 (the remainder: each rank's Newton exits on its own
maximum, and compiler rounding).  A planted mistake in the rank views makes
the comparison fail.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

from tests import sodcase as sc

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
# per-variable bars, relative to the variable's largest magnitude
BARS = {0: 2e-7, 1: 5e-7, 4: 2e-7, 5: 2e-7, 6: 2e-7}
ABS_TRANSVERSE = 1e-12     # rho v, rho w: the reference's own noise is 3e-13


def _io_or_skip():
    from phd_amd import restart
    if not os.path.exists(restart.IO_PATH):
        pytest.skip("libizem_io.so not built (needs the HDF5 C library)")


def _errors(run, adi, k):
    ref, t = sc.read_reference(adi, k, run.thd)
    I = (slice(None),) + run.box.interior()
    a, b = run.v[I], ref[I]
    rel = {}
    for l in range(a.shape[0]):
        d = np.max(np.abs(a[l] - b[l]))
        rel[l] = d / np.max(np.abs(b[l])) if l in BARS else d
    return t, rel


def test_sod_oracle_reproduces_reference_trajectory(adi):
    _io_or_skip()
    run = sc.SodRun(adi)
    stats = sc.reference_stats()
    for k in (1, 2):
        run.run_to_stat(k)
        t, rel = _errors(run, adi, k)
        print(f"statistics {k}: iteration {run.ite}, "
              + ", ".join(f"var {l} {e:.1e}" for l, e in rel.items()))
        assert run.time == t                       # the same statistics time
        for l, bar in BARS.items():
            assert rel[l] <= bar, (k, l, rel[l])
        assert rel[2] <= ABS_TRANSVERSE and rel[3] <= ABS_TRANSVERSE
        # the step that lands on the statistics time and the four limits
        # (times_stat.out prints 9 significant digits)
        dtmin, dt_4 = run.dt_hist[-1]
        row = stats[k]
        assert abs(dtmin * adi.time_ref - row[1]) <= 1e-8 * row[1]
        for q in range(4):
            assert abs(dt_4[q] / dtmin - row[2 + q]) <= 1e-8 * row[2 + q]
    assert run.ite == 712


def test_sod_trajectory_sees_rank_views(adi, tmp_path):
    """Planted mistake: every rank sees every cell's flag."""
    _io_or_skip()
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    from oracle.oracle import Oracle
    src = open(os.path.join(ORACLE, "izem_oracle.c")).read()
    old = "for (int p = p0 > lo ? p0 : lo; p <= p1 && p <= hi; p++)"
    assert src.count(old) == 1
    shutil.copy(os.path.join(ORACLE, "izem_oracle.h"), tmp_path / "izem_oracle.h")
    (tmp_path / "izem_oracle.c").write_text(src.replace(old, "for (int p = p0; p <= p1; p++)"))
    so = tmp_path / "liboracle_nrv2.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-std=c11", "-DORA_NRV=2",
                    "-shared", "-o", str(so), str(tmp_path / "izem_oracle.c"), "-lm"], check=True)
    run = sc.SodRun(adi, oracle=Oracle(str(so), nrv=2))
    run.run_to_stat(1)
    _, rel = _errors(run, adi, 1)
    assert rel[0] > 100 * BARS[0], rel


def test_product_sod_state_equals_oracle_state(adi):
    """phd_amd.ics.sod_state (the run driver's ini_sod) builds the oracle's
    initial state to the last bit: the same X_to_Y, Wmix_i and ha_scalar
    operation order, the right state's density recomputed from its
    temperature."""
    from oracle.oracle import Oracle
    from phd_amd import grid, ics
    o = Oracle(nrv=2)
    thd, box = sc.sod_thd(adi), sc.sod_box(adi)
    v0, T0 = sc.initial_state(o, adi, thd, box)
    assert grid.count_points(sc.GRID) == (sc.NX, sc.NY, sc.NZ)
    xt, _, _ = grid.read_grid(sc.GRID, sc.NX, sc.NY, sc.NZ)
    v, T = ics.sod_state(adi, thd, box, xt)
    assert np.array_equal(v, v0) and np.array_equal(T, T0)
    I = box.interior()
    rho = v[(0,) + I][0, 0]
    assert rho[0] == 1.0 and abs(rho[-1] - 0.125) < 1e-15 and (rho == 1.0).sum() == sc.NX // 2


def test_host_ha_scalar_equals_oracle(adi):
    """phd_amd.thermo_host.ha_scalar against the oracle's ha_scalar, bit for
    bit, below Tpolymin, in both NASA ranges and above Tpolymax."""
    import ctypes as C
    from oracle.oracle import Oracle
    from phd_amd.thermo_host import ha_scalar
    o = Oracle(nrv=2)
    thd = sc.sod_thd(adi)
    t = o.thd_struct(thd)
    for TK in (150.0, 299.0, 300.0, 800.0, 999.9, 1000.0, 2500.0, 4999.0, 5000.0, 7000.0):
        Tn = TK / adi.T_ref
        ref = np.zeros(2)
        o.lib.ora_ha_scalar(C.byref(t), C.c_double(Tn), ref.ctypes.data_as(C.POINTER(C.c_double)))
        assert np.array_equal(ha_scalar(thd, Tn), ref), TK

