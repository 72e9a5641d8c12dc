"""The characteristic outflow (GRID%BCXMAX = 'NSCBC_OUT', bc_nscbc_out,
module_BCs.f90:2331-2516) on the GPU against the oracle: the 2-D hot spot
and the 3-D turbulence recipe with x extrapolated on W and the outflow on
E (their flow leaves and enters through it, so both branches of the wave
amplitudes act; the gas is at 1 atm against the reference's far-field
94232.25 Pa, so the relaxation acts), viscous: one SSP-RK3 step and 20
main-loop iterations (the outflow takes the step's dt from the device) to
1e-12.  The oracle's outflow is pinned by the LODI relations of a subsonic
outflow (tests/test_nscbc_cpu.py)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.flowcase import OracleRun, P, small_state
from tests.test_gpu_bc import _oracle_rk3, _oracle_timestep, _relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12
PINF = 94232.25
CASES = [("2d", 24, (1, 3, 0, 0, 0, 0)), ("3d", 12, (1, 3, 0, 0, 1, 1))]
IDS = ["2d", "3d"]


def _setup(oracle, adi, thd, kind, n, bc):
    import torch
    from phd_amd.solver import Options, Solver
    box, v, T = small_state(adi, thd, kind, n)
    run = OracleRun(oracle, adi, thd, box, v, T)
    for f in range(6):
        run.g.bc[f] = bc[f]
    Lx = box.L[0] * (box.nx - 1) / box.nx / adi.L_ref
    run.g.nscbc_Lx, run.g.nscbc_pinf = Lx, PINF / adi.p_ref
    oracle.lib.ora_comm_cons(C.byref(run.g), P(run.v))
    oracle.lib.ora_upd_boundaries(C.byref(run.g), P(run.v))
    run.prim_all()
    s = Solver(box, Options(reaction=0), device="cuda:0")
    s.set_bc(bc)
    s.set_nscbc(Lx, PINF / adi.p_ref, run.d_i[0][box.nx - 1 + box.ng])
    for name in ("v", "T", "W_i", "cp", "ha", "sensor"):
        getattr(s, name).copy_(torch.from_numpy(getattr(run, name)))
    torch.cuda.synchronize()
    print(f"case {kind} n={n} bc={bc}", flush=True)
    return box, run, s


@pytest.mark.parametrize("kind,n,bc", CASES, ids=IDS)
def test_outflow_rk3_and_steps(oracle, adi, thd, kind, n, bc):
    import torch
    box, run, s = _setup(oracle, adi, thd, kind, n, bc)
    I = box.interior()
    ux = run.v[1][I][..., -1] / run.v[0][I][..., -1]
    assert (ux > 0).any() and (ux < 0).any()          # out- and inflow cells on E
    # the outflow needs the primitives and dt: the plain ghost fill refuses
    assert s.lib.izem_comm_cons(s.ctx, C.c_void_p(s.v.data_ptr())) == -4   # IZEM_E_STATE
    opts = run.opts(vis=1)
    dt = _oracle_timestep(run, opts)
    s.timestep()
    assert abs(s.time_state()[4] - dt) <= 1e-14 * dt
    _oracle_rk3(run, opts, dt)
    s.rk3(torch.tensor([dt], dtype=torch.float64, device="cuda:0"))
    s.sync()
    assert s.status() == 0
    e_v = _relerr(s.v.cpu().numpy(), run.v, box).max()
    print(f"  rk3 state {e_v:.2e}")
    assert e_v <= TOL
    s.set_time_state([1.0, 1.0, 1.0, 1.0, dt, 0.0])
    dtmin, worst = dt, 0.0
    for it in range(20):
        dtmin = _oracle_timestep(run, opts, dtmin)
        _oracle_rk3(run, opts, dtmin)
        s.step()
        worst = max(worst, abs(s.time_state()[4] - dtmin) / dtmin)
    assert s.status() == 0
    e = _relerr(s.v.cpu().numpy(), run.v, box).max()
    print(f"  21 steps: state {e:.2e}, dt {worst:.2e}")
    assert e <= TOL and worst <= TOL
    s.close()


def test_outflow_refused_elsewhere(adi, thd):
    from phd_amd import _lib
    from phd_amd.solver import Options, Solver
    box, v, T = small_state(adi, thd, "2d", 16)
    s = Solver(box, Options(reaction=0), device="cuda:0")
    for bc in ((3, 1, 0, 0, 0, 0), (0, 0, 1, 3, 0, 0)):
        with pytest.raises(_lib.IzemError):
            s.set_bc(bc)
    with pytest.raises(_lib.IzemError):
        s.set_nscbc(0.0, 1.0, 1.0)
    s.close()
