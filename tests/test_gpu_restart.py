"""GPU side of restarts: the reference's restart initialisation
(read_restart_hdf5, module_io.f90:165-180: periodic ghost fill, the
machine-precision inversion from T = 0, the edge extrapolation of every
face, the ghost slabs' inversion) against the oracle's, and a run resumed
from a restart against the same run without the break."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.flowcase import OracleRun, P, small_state

pytestmark = pytest.mark.gpu


def _oracle_init(oracle, adi, thd, box, v, ghosts):
    from phd_amd import restart
    run = OracleRun(oracle, adi, thd, box, v, np.zeros(box.shape))
    oracle.lib.ora_comm_cons(C.byref(run.g), P(run.v))
    it = C.c_int()
    doms = [0] + [s * q for q in range(1, box.ndim + 1) for s in (-1, 1)]
    for d in doms:
        if d == doms[1] and ghosts == "reference":
            restart.extrapolate_faces(box, run.v)
        st = oracle.lib.ora_prim_inv_var_mach(d, C.byref(run.g), C.byref(run.t), P(run.v),
                                              P(run.W_i), P(run.T), P(run.cp), P(run.ha),
                                              C.byref(it))
        assert st == 0
    return run


@pytest.mark.parametrize("ghosts", ["reference", "periodic"])
@pytest.mark.parametrize("kind,n", [("2d", 16), ("3d", 8)])
def test_restart_initialisation_parity(tmp_path, oracle, adi, thd, kind, n, ghosts):
    from phd_amd import restart
    from phd_amd.solver import Options, Solver
    box, v, T = small_state(adi, thd, kind, n)
    path = str(tmp_path / "restart_00001.h5")
    restart.write(path, box, thd.names, v, 3.25)
    vf, _ = restart.read(path, box, thd.names)
    run = _oracle_init(oracle, adi, thd, box, vf, ghosts)
    s = Solver(box, Options(reaction=0))
    assert s.load_restart(path, ghosts=ghosts) == 3.25
    assert s.status() == 0
    assert np.array_equal(s.v.cpu().numpy(), run.v)
    for name, ref in (("T", run.T), ("cp", run.cp), ("W_i", run.W_i)):
        g = getattr(s, name).cpu().numpy()
        sel = ref != 0
        assert np.array_equal(g[~sel], ref[~sel]), name
        assert np.max(np.abs(g[sel] - ref[sel]) / np.abs(ref[sel])) <= 1e-14, name
    h = s.ha.cpu().numpy()
    assert np.max(np.abs(h - run.ha)) <= 1e-14 * np.max(np.abs(run.ha))
    # the interior temperature the state was built from, at machine precision
    I = box.interior()
    assert np.max(np.abs(run.T[I] - T[I]) / T[I]) <= 1e-12


def test_resumed_run_matches_the_unbroken_run(tmp_path, adi, thd, mech):
    """Four reacting steps against two, a restart written and read back with
    periodic ghosts into a new context, and two more: the states agree to
    1e-9 (the restart's inversion converges to 1e-10 where the running one
    stops at 1e-5, so the temperatures differ in the last digits)."""
    import torch
    from phd_amd import ics
    from phd_amd.solver import Options, Solver
    box, v, T = ics.config1_state(adi, thd, n=24, T_peak=1800.0)
    a = Solver(box, Options(flux_mode=1), adi=adi, thd=thd, mech=mech)
    a.load_state(v, T)
    for _ in range(4):
        a.step()
    a.sync()
    va = a.v.cpu().numpy()
    b = Solver(box, Options(flux_mode=1), adi=adi, thd=thd, mech=mech)
    b.load_state(v, T)
    for _ in range(2):
        b.step()
    b.sync()
    path = str(tmp_path / "restart_00002.h5")
    b.save_restart(path, 1.0)
    sensor = b.sensor.clone()
    c = Solver(box, Options(flux_mode=1), adi=adi, thd=thd, mech=mech)
    assert c.load_restart(path, ghosts="periodic") == 1.0
    c.sensor.copy_(sensor)  # the sticky sensor is not part of a restart file
    c.set_time_state(b.time_state())
    for _ in range(2):
        c.step()
    c.sync()
    assert a.status() == 0 and c.status() == 0
    vc = c.v.cpu().numpy()
    I = (slice(None),) + box.interior()
    scale = np.max(np.abs(va[I]), axis=(1, 2, 3))
    err = np.max(np.abs(vc[I] - va[I]), axis=(1, 2, 3)) / np.where(scale > 0, scale, 1)
    print("resumed run, max rel err per field", err.max())
    assert err.max() <= 1e-9
