"""The reference's double Mach reflection deck on the GPU, through the run
driver (python -m phd_amd.run), against the oracle's restatement of the run
(tests/dmrcase.py): ini_DMR with its Rankine-Hugoniot states
(phd_amd.shock, pinned in tests/test_shock_cpu.py), the post-shock slip
wall below (izem_set_bc, izem_set_wall_state: the post-shock state on the
first cells and the first interior row, its primitives recomputed as
upd_prim_var_dirichlet does, the mirror beyond), extrapolated faces
elsewhere, WENO7 on the 2-species library, the deck's 4 ranks along x.

Both runs take the reference's operation order in the flux kernels
(--flux-mode 0): the structured projections of the default mode differ by
rounding (1e-13 after one step), which the Mach 4.5 shock and WENO7's
weights lift to 2e-10 over a few steps.

* A coarse copy of the case (every 15th x and 12th y point, 60 x 40 cells)
  to 8 statistics times: every restart file equal to the oracle's state
  within 1e-12.  The first stage of a run sees no sensor flags yet, so the
  incident shock is crossed by linear WENO7 faces and undershoots; the
  negative-temperature repair (a warning in the reference) acts in both.
* The deck as shipped (900 x 511, NSTAT cut to 2: one iteration to 1e-10 s,
  then to 2.01e-8 s): restart file 2 against the oracle within 1e-12.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from tests import dmrcase as dc

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _io_or_skip():
    from phd_amd import restart
    if not os.path.exists(restart.IO_PATH):
        pytest.skip("libizem_io.so not built (needs the HDF5 C library)")


def _relerr(a, b, box):
    I = (slice(None),) + box.interior()
    a, b = a[I], b[I]
    out = []
    for m in range(a.shape[0]):
        s = np.max(np.abs(b[m]))
        d = np.max(np.abs(a[m] - b[m]))
        out.append(0.0 if s == 0 and d == 0 else d / (s if s > 0 else 1.0))
    return np.array(out)


def _compare(tmp_path, adi, run, k):
    from phd_amd import restart
    ours, t = restart.read(str(tmp_path / f"restart_{k:05d}.h5"), run.box, run.thd.names)
    assert t == run.time, (k, t, run.time)
    e = _relerr(ours, run.v, run.box)
    print(f"  statistics {k}: iteration {run.ite}, state {e.max():.2e}", flush=True)
    return e.max()


def test_dmr_coarse_deck_matches_oracle(tmp_path, adi, capsys):
    _io_or_skip()
    from phd_amd.run import main
    nstat = 8
    dc.write_deck(str(tmp_path), nstat, coarse=(15, 12, 40))
    print("coarse double Mach reflection deck through the driver", flush=True)
    assert main([str(tmp_path / "dmr.inp"), "--flux-mode", "0"]) == 0
    out = capsys.readouterr().out
    assert "entering the main loop (60 x 40 x 1, inert)" in out
    run = dc.DmrRun(adi, str(tmp_path / "grid.dat"), nstat)
    worst, repaired = 0.0, 0
    for k in range(1, nstat + 1):
        run.run_to_stat(k)
        repaired += run.status
        worst = max(worst, _compare(tmp_path, adi, run, k))
    assert repaired > 0                      # the repair was exercised
    assert worst <= TOL


def test_dmr_coarse_resumed_from_restart(tmp_path, adi, capsys):
    """A run resumed from restart file 3 (READ_RESTART): the reader's
    upd_boundaries puts the wall's state back on its first cells and first
    row before the ghost slabs, as read_restart_hdf5 does; 3 more
    statistics times against the oracle resumed from the same file."""
    _io_or_skip()
    from phd_amd.run import main
    nstat = 6
    dc.write_deck(str(tmp_path), nstat, coarse=(15, 12, 40))
    assert main([str(tmp_path / "dmr.inp"), "--flux-mode", "0"]) == 0
    deck = (tmp_path / "dmr.inp").read_text()
    # statistics times restart from TIME_OFFSET (module_input.f90:718-721):
    # put it at the restart time, so that the first is met by a step of
    # (about) zero length
    assert deck.count("TIME_OFFSET = 1.0e-10") == 1
    deck = deck.replace("TIME_OFFSET = 1.0e-10", "TIME_OFFSET = 4.01e-08").replace(
        "[MONITORING]", "[MONITORING]\n        READ_RESTART = T\n"
        "        NAME_RESTART = 'restart'\n        NUMBER_RESTART = 3")
    os.makedirs(tmp_path / "resume")
    for f in ("grid.dat", "therm.dat", "chem.inp", "tran.dat", "restart_00003.h5"):
        os.link(tmp_path / f, tmp_path / "resume" / f)
    (tmp_path / "resume" / "dmr.inp").write_text(deck)
    assert main([str(tmp_path / "resume" / "dmr.inp"), "--flux-mode", "0"]) == 0
    capsys.readouterr()
    run = dc.DmrRun(adi, str(tmp_path / "grid.dat"), nstat,
                    restart=(str(tmp_path / "restart_00003.h5"), 3), time_offset=4.01e-08)
    worst = 0.0
    for k in range(4, nstat + 1):
        run.run_to_stat(k)
        worst = max(worst, _compare(tmp_path / "resume", adi, run, k))
    assert worst <= TOL


def test_dmr_reference_deck(tmp_path, adi, capsys):
    _io_or_skip()
    from phd_amd.run import main
    dc.write_deck(str(tmp_path), 2)
    print("double Mach reflection deck as shipped (NSTAT 2)", flush=True)
    assert main([str(tmp_path / "dmr.inp"), "--flux-mode", "0"]) == 0
    out = capsys.readouterr().out
    assert "entering the main loop (900 x 511 x 1, inert)" in out
    run = dc.DmrRun(adi, os.path.join(dc.GOLDEN, "dmr_grid.dat"), 2)
    run.run_to_stat(2)
    assert _compare(tmp_path, adi, run, 2) <= TOL
    # the gas ahead of the shock is untouched: the top right corner
    from phd_amd import shock
    I = run.box.interior()
    rho = run.v[(0,) + I][0]
    assert np.all(rho[-5:, -5:] == shock.dmr_states(adi, run.thd).rho1)
