"""A decomposed run reproduced on one GPU (izem_set_decomposition) against
the oracle run with the same ranks.

The reference's results depend on its MPI decomposition through the shock
sensor: each rank's Adams-Shariff sweep stops one cell short of its own last
cell (module_tools.f90:2733-2814), a rank sees only its own cells' flags
(the sensor is neither set on ghosts nor exchanged), and the face between two
ranks is formed by both, each with its own view (module_weno.f90:955-1020).
The oracle restates that (tests/test_sod_reference_cpu.py pins it on the
reference's own 2-rank Sod trajectory).  Here the 2-D hot spot and the 3-D
turbulence recipe, with 2 or 3 ranks per direction, on the WENO5 library and
the 9-species WENO7 variant: the sensor bit for bit, the right-hand side and
one SSP-RK3 step to 1e-12 (reference operation order; 1e-11 with the
structured projections).  The first stage's right-hand side is formed before
any flag is set; the step's later stages see the flags, and the undecomposed
oracle's step differs from the decomposed one by 3e-3 to 2e-1, so each case
exercises the seams.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests.flowcase import OracleRun, P, small_state
from tests.test_gpu_weno7 import _ork3, _otimestep, _relerr, _repad, _upload

pytestmark = pytest.mark.gpu
TOL = {0: 1e-12, 1: 1e-11}

# thresholds at which the flags of one step are sparse (a fifth to a third
# of the cells), so that which rank sees which flag changes the faces
CASES = [  # kind, n, sensor threshold, ranks, ghost layers, flux mode
    ("3d", 16, 0.01, (2, 3, 2), 3, 0),
    ("2d", 24, 0.001, (3, 2, 1), 3, 0),
    ("2d", 24, 0.001, (3, 2, 1), 3, 1),
    ("3d", 16, 0.005, (2, 1, 3), 4, 0),
    ("2d", 24, 0.0001, (2, 3, 1), 4, 1),
]


@pytest.mark.parametrize("kind,n,mrw,nproc,ng,mode", CASES)
def test_decomposed_rhs_and_step(oracle, adi, thd, kind, n, mrw, nproc, ng, mode):
    import torch
    from phd_amd import _lib
    from phd_amd.solver import Options, Solver
    print(f"case {kind} n={n} ng={ng} ranks={nproc} flux mode {mode}", flush=True)
    box, v, T = small_state(adi, thd, kind, n)
    if ng == 4:
        box, v, T = _repad(box, v, T)
    run = OracleRun(oracle, adi, thd, box, v, T)
    for q in range(3):
        run.g.nproc[q] = nproc[q]
    run.prim_all()
    whole = OracleRun(oracle, adi, thd, box, run.v, run.T)
    whole.W_i, whole.cp, whole.ha = run.W_i.copy(), run.cp.copy(), run.ha.copy()
    opts = run.opts(vis=1, mrw=mrw)
    s = Solver(box, Options(reaction=0, max_rel_weight=mrw, vis=1, flux_mode=mode), adi=adi,
               thd=thd, lib=None if ng == 3 else _lib.load_variant(9, 4))
    s.set_decomposition(nproc)
    _upload(s, run)
    f_gpu = torch.zeros((14,) + box.shape, dtype=torch.float64, device="cuda:0")
    s.rhs(f_gpu)
    s.sync()
    dt = _otimestep(run, opts)
    s.timestep()
    assert abs(s.time_state()[4] - dt) <= 1e-14 * dt
    f1 = np.zeros((14,) + box.shape)
    _ork3(run, opts, dt, f1)
    f_whole = np.zeros((14,) + box.shape)
    _ork3(whole, opts, dt, f_whole)
    s.rk3(torch.tensor([dt], dtype=torch.float64, device="cuda:0"))
    s.sync()
    assert s.status() == 0
    e_rhs = _relerr(f_gpu.cpu().numpy(), f1, box).max()
    e_v = _relerr(s.v.cpu().numpy(), run.v, box).max()
    d_whole = _relerr(whole.v, run.v, box).max()
    ncell = box.nx * box.ny * box.nz
    print(f"  rhs {e_rhs:.2e}, rk3 state {e_v:.2e}; the undecomposed step differs by "
          f"{d_whole:.2e}; {int(run.sensor.sum())} of {ncell} cells flagged")
    assert np.array_equal(s.sensor.cpu().numpy(), run.sensor)
    assert 0 < run.sensor.sum() < ncell
    assert d_whole > 1e-4
    assert e_rhs <= TOL[mode] and e_v <= TOL[mode]
    s.close()


def test_set_decomposition_refuses(adi, thd):
    from phd_amd import _lib
    from phd_amd.solver import Options, Solver
    box, v, T = small_state(adi, thd, "2d", 16)
    s = Solver(box, Options(reaction=0), adi=adi, thd=thd)
    for bad in ((0, 1, 1), (1, 1, 2), (17, 1, 1), (1, -1, 1)):
        with pytest.raises(_lib.IzemError):
            s.set_decomposition(bad)
    s.set_decomposition((16, 2, 1))
    s.set_decomposition((1, 1, 1))
    s.close()
