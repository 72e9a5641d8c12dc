"""The Adams-Shariff sensor at a real threshold on the GPU, and RK3 steps in
which it leaves faces unflagged (the linear wenorec_nw path of
module_weno.f90:236-275 on those faces), against the oracle."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests import pins
from tests.flowcase import OracleRun, P, small_state

pytestmark = pytest.mark.gpu
MRW = 0.15


def _solver(box):
    from phd_amd.solver import Options, Solver
    return Solver(box, Options(reaction=0, max_rel_weight=MRW), device="cuda:0")


def _upload(s, run):
    import torch
    for name in ("v", "T", "W_i", "cp", "ha", "sensor"):
        getattr(s, name).copy_(torch.from_numpy(getattr(run, name)))
    torch.cuda.synchronize()


def test_shock_sensor_threshold(oracle, adi, thd):
    """Flags exactly the cells of the reference's definition: |d rho| /
    max(rho) and |d P| / min(P) both above 0.15 towards the next cell."""
    box, v, Tn, want = pins.sensor_case(adi, thd)
    run = OracleRun(oracle, adi, thd, box, v, Tn)
    pins.exact_prims(run)
    s = _solver(box)
    _upload(s, run)
    s.shock_sensor()
    s.sync()
    n, g = box.nx, box.ng
    got = s.sensor.cpu().numpy()[g, g:g + n, g:g + n].astype(bool)
    assert np.array_equal(got, want), np.argwhere(got != want)
    # sticky: a second call on a uniform state keeps the flags
    import torch
    s.v.copy_(torch.from_numpy(np.broadcast_to(run.v[:, g:g + 1, g:g + 1, g:g + 1], run.v.shape).copy()))
    s.shock_sensor()
    s.sync()
    assert np.array_equal(s.sensor.cpu().numpy()[g, g:g + n, g:g + n].astype(bool), want)


@pytest.mark.parametrize("case", ["steps", "hotspot", "hit3d"])
def test_rk3_with_sensor_threshold(oracle, adi, thd, case):
    import torch
    if case == "steps":
        box, v, Tn, _ = pins.sensor_case(adi, thd)
    elif case == "hotspot":
        box, v, Tn = small_state(adi, thd, "2d", 24)
    else:
        box, v, Tn = small_state(adi, thd, "3d", 12)
    run = OracleRun(oracle, adi, thd, box, v, Tn)
    run.prim_all()
    s = _solver(box)
    _upload(s, run)
    opts = run.opts(vis=1, mrw=MRW)
    dt = 0.2
    st = oracle.lib.ora_rk3(C.byref(run.g), C.byref(run.t), C.byref(run.a), C.byref(opts),
                            C.c_double(dt), P(run.d_i[0]), P(run.d_i[1]), P(run.d_i[2]),
                            P(run.T), P(run.W_i), P(run.cp), P(run.ha),
                            run.sensor.ctypes.data_as(C.POINTER(C.c_uint8)), P(run.v), None)
    assert st == 0
    s.rk3(torch.tensor([dt], dtype=torch.float64, device="cuda:0"))
    s.sync()
    assert s.status() == 0
    I = (slice(None),) + box.interior()
    flagged = int(run.sensor[box.interior()].sum())
    assert np.array_equal(s.sensor.cpu().numpy(), run.sensor)
    vg = s.v.cpu().numpy()
    sc = np.max(np.abs(run.v[I]), axis=(1, 2, 3))
    err = np.max(np.abs(vg[I] - run.v[I]), axis=(1, 2, 3)) / np.where(sc > 0, sc, 1.0)
    print(f"{case}: {flagged} of {box.nx * box.ny * box.nz} cells flagged, rk3 max rel err "
          f"{err.max():.2e}")
    assert err.max() <= 1e-12


@pytest.mark.parametrize("case", ["waves2d", "waves3d", "hotspot", "hit3d"])
def test_ducros_sensor_and_rk3(oracle, adi, thd, case):
    """shock_sensor_ducros (WENO%SHOCK_SENSOR_TYPE = DUCROS, threshold 0.5)
    on the GPU against the oracle: the same flags, and an RK3 step that
    takes the JS or the linear reconstruction by them."""
    import torch
    from phd_amd.solver import Options, Solver
    if case.startswith("waves"):
        nd = 2 if case == "waves2d" else 3
        box, v, Tn, _, _ = pins.ducros_case(adi, thd, nd, n=24 if nd == 2 else 12)
    elif case == "hotspot":
        box, v, Tn = small_state(adi, thd, "2d", 24)
    else:
        box, v, Tn = small_state(adi, thd, "3d", 12)
    run = OracleRun(oracle, adi, thd, box, v, Tn)
    run.prim_all()
    s = Solver(box, Options(reaction=0, sensor_type=1, ducros_treduc=0.5), device="cuda:0")
    _upload(s, run)
    s.shock_sensor_ducros()
    s.sync()
    ref = np.zeros(box.shape, np.uint8)
    oracle.lib.ora_sensor_ducros(C.byref(run.g), C.c_double(0.5), P(run.d_i[0]), P(run.d_i[1]),
                                 P(run.d_i[2]), P(run.v),
                                 ref.ctypes.data_as(C.POINTER(C.c_uint8)))
    got = s.sensor.cpu().numpy()
    assert np.array_equal(got, ref)
    flagged = int(ref[box.interior()].sum())
    s.sensor.zero_()
    opts = run.opts(vis=1, sensor_type=1, ducros_treduc=0.5)
    dt = 0.2
    st = oracle.lib.ora_rk3(C.byref(run.g), C.byref(run.t), C.byref(run.a), C.byref(opts),
                            C.c_double(dt), P(run.d_i[0]), P(run.d_i[1]), P(run.d_i[2]),
                            P(run.T), P(run.W_i), P(run.cp), P(run.ha),
                            run.sensor.ctypes.data_as(C.POINTER(C.c_uint8)), P(run.v), None)
    assert st == 0
    s.rk3(torch.tensor([dt], dtype=torch.float64, device="cuda:0"))
    s.sync()
    assert s.status() == 0
    assert np.array_equal(s.sensor.cpu().numpy(), run.sensor)
    I = (slice(None),) + box.interior()
    vg = s.v.cpu().numpy()
    sc = np.max(np.abs(run.v[I]), axis=(1, 2, 3))
    err = np.max(np.abs(vg[I] - run.v[I]), axis=(1, 2, 3)) / np.where(sc > 0, sc, 1.0)
    print(f"ducros {case}: {flagged} cells flagged, rk3 max rel err {err.max():.2e}")
    assert err.max() <= 1e-12
