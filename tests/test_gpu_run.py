"""The deck driver (python -m phd_amd.run): a reacting 2-D run from a deck
that lands on its statistics times and writes a restart at each, then a
second run read from one of those restarts."""
from __future__ import annotations

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DECK = """# reacting hot spot, WENO5, every statistics time a restart
[GENERAL]
  DIM = 2
  NRV = 9
  INIT = 'HOTSPOT'
[TIMESTEPPING]
  CFL = 0.5
[WENO]
  ORDER_WENO = 5
  PERCENT_WEIGHT = -1.0
[REACTING]
  REACTION = T
[MONITORING]
  ITSHOW = 5
  ITLIM = T
  ITMAX = 60
  STAT = T
  NSTAT = 3
  TIME_OFFSET = 2.0e-7
  DTIME = 2.0e-7
{extra}
"""


def test_deck_run_and_restart(tmp_path, capsys):
    from phd_amd import restart
    from phd_amd.run import main
    from phd_amd import chemdata as cd
    deck = tmp_path / "hotspot.ini"
    deck.write_text(DECK.format(extra=""))
    assert main([str(deck), "--n", "16"]) == 0
    lines = (tmp_path / "times_stat.out").read_text().splitlines()
    assert len(lines) == 3
    t = np.array([float(l.split()[1]) for l in lines])
    assert np.allclose(t, [2e-7, 4e-7, 6e-7], rtol=1e-9, atol=0)
    adi = cd.Adim()
    for k in (1, 2, 3):
        (nx, ny, nz), tk = restart.dims(str(tmp_path / f"restart_{k:05d}.h5"))
        assert (nx, ny, nz) == (16, 16, 1)
        assert tk * adi.time_ref == pytest.approx(2e-7 * k, rel=1e-9)
    out = capsys.readouterr().out
    assert "entering the main loop" in out
    # resume from the second restart.  As in the reference, a restarted run's
    # statistics times begin with NUMBER_RESTART + 1 copies of TIME_OFFSET
    # (module_input.f90:718-722): with TIME_OFFSET at the restart time the
    # first is met at once by a step of zero length (the chemistry's DVODE
    # returns untouched for tout = t), the next one DTIME later
    deck2 = tmp_path / "resume.ini"
    deck2.write_text(DECK.format(extra="  READ_RESTART = T\n  NAME_RESTART = 'restart'\n"
                                       "  NUMBER_RESTART = 2\n").replace(
        "NSTAT = 3", "NSTAT = 4").replace("TIME_OFFSET = 2.0e-7", "TIME_OFFSET = 4.0e-7"))
    os.remove(tmp_path / "times_stat.out")
    assert main([str(deck2), "--n", "16"]) == 0
    lines = (tmp_path / "times_stat.out").read_text().splitlines()
    assert [int(l.split()[0]) for l in lines] == [3, 4]
    t = np.array([float(l.split()[1]) for l in lines])
    assert np.allclose(t, [4e-7, 6e-7], rtol=1e-9, atol=0)


def test_reference_sod_deck_reproduces_reference_restarts(tmp_path, capsys):
    """The reference's own Sod deck (devel-izem/testcases/sod: sod.inp,
    grid.dat, tranfit.out; NSTAT cut from 30 to 2) through the run driver:
    ini_sod on the grid file, the 2-species WENO7 library, extrapolated x
    faces, the deck's 2 ranks along x reproduced on the GPU.  Its restart
    files 1 and 2 (iterations 647 and 712) against the reference's own, to
    the bars the oracle meets (tests/test_sod_reference_cpu.py), and its
    times_stat.out against the reference's."""
    import shutil
    from phd_amd import chemdata as cd
    from phd_amd import restart
    from phd_amd.run import main
    from tests import sodcase as sc
    from tests.test_sod_reference_cpu import ABS_TRANSVERSE, BARS
    if not os.path.exists(restart.IO_PATH):
        pytest.skip("libizem_io.so not built (needs the HDF5 C library)")
    deck = open(os.path.join(sc.GOLDEN, "sod.inp")).read()
    assert deck.count("NSTAT = 30") == 1
    (tmp_path / "sod.inp").write_text(deck.replace("NSTAT = 30", "NSTAT = 2"))
    shutil.copy(sc.GRID, tmp_path / "grid.dat")
    shutil.copy(sc.TRANFIT, tmp_path / "tranfit.out")
    assert main([str(tmp_path / "sod.inp")]) == 0
    out = capsys.readouterr().out
    assert "entering the main loop (400 x 11 x 1, inert)" in out
    assert " iteration      712 " in out and " iteration      713 " not in out
    adi = cd.Adim()
    thd = sc.sod_thd(adi)
    box = sc.sod_box(adi)
    stats = sc.reference_stats()
    rows = [[float(x) for x in ln.split()] for ln in
            (tmp_path / "times_stat.out").read_text().splitlines()]
    assert [int(r[0]) for r in rows] == [1, 2]
    I = (slice(None),) + box.interior()
    for k in (1, 2):
        ours, t = restart.read(str(tmp_path / f"restart_{k:05d}.h5"), box, thd.names)
        ref, t_ref = sc.read_reference(adi, k, thd)
        assert t == t_ref
        rel = {}
        for l in range(ours.shape[0]):
            d = np.max(np.abs(ours[I][l] - ref[I][l]))
            rel[l] = d / np.max(np.abs(ref[I][l])) if l in BARS else d
        print(f"statistics {k}: " + ", ".join(f"var {l} {e:.1e}" for l, e in rel.items()))
        for l, bar in BARS.items():
            assert rel[l] <= bar, (k, l, rel[l])
        assert rel[2] <= ABS_TRANSVERSE and rel[3] <= ABS_TRANSVERSE
        # time, dtmin and the four limits / dtmin (9 significant digits)
        for q in range(1, 7):
            assert abs(rows[k - 1][q] - stats[k][q - 1]) <= 1e-8 * abs(stats[k][q - 1]), (k, q)


def test_reference_sod_deck_resumed_from_reference_restart(tmp_path, capsys):
    """The reference's Sod deck resumed (READ_RESTART) from the reference's
    own restart_00001.h5, against the oracle resumed from the same file.  A
    resumed run's statistics times start with NUMBER_RESTART + 1 copies of
    TIME_OFFSET (module_input.f90:718-721): the first is met at once by a
    step of zero length and written as statistics 2, statistics 3 falls at
    the original statistics 2's time.  Against the oracle in the reference's
    operation order within 1e-10: the temperature re-derived from the file
    (prim_inv_var_err_mach from T = 0) agrees to 1e-14, not bit for bit
    (tests/test_gpu_restart.py), and 66 WENO7 iterations lift that to 2e-11.
    It is not the writing run's trajectory:
    the sticky shock sensor restarts empty and T is re-derived from the
    file, so at the original statistics 2's time it lies 1e-3 from the
    reference's restart_00002.h5 (the oracle's uninterrupted run: 5e-8)."""
    import shutil
    from phd_amd import chemdata as cd
    from phd_amd import restart
    from phd_amd.run import main
    from tests import sodcase as sc
    if not os.path.exists(restart.IO_PATH):
        pytest.skip("libizem_io.so not built (needs the HDF5 C library)")
    deck = open(os.path.join(sc.GOLDEN, "sod.inp")).read()
    deck = deck.replace("NSTAT = 30", "NSTAT = 3").replace(
        "[MONITORING]", "[MONITORING]\n        READ_RESTART = T\n"
        "        NAME_RESTART = 'restart'\n        NUMBER_RESTART = 1")
    (tmp_path / "sod.inp").write_text(deck)
    shutil.copy(sc.GRID, tmp_path / "grid.dat")
    shutil.copy(sc.TRANFIT, tmp_path / "tranfit.out")
    shutil.copy(sc.restart_path(1), tmp_path / "restart_00001.h5")
    assert main([str(tmp_path / "sod.inp"), "--flux-mode", "0"]) == 0
    capsys.readouterr()
    rows = [[float(x) for x in ln.split()] for ln in
            (tmp_path / "times_stat.out").read_text().splitlines()]
    assert [int(r[0]) for r in rows] == [2, 3]
    adi = cd.Adim()
    run = sc.SodRun(adi, restart=(sc.restart_path(1), 1))
    run.run_to_stat(3)
    box = run.box
    ours, t = restart.read(str(tmp_path / "restart_00003.h5"), box, run.thd.names)
    assert t == run.time
    I = (slice(None),) + box.interior()
    e = max(np.max(np.abs(ours[I][l] - run.v[I][l])) / np.max(np.abs(run.v[I][l]))
            for l in range(ours.shape[0]) if np.max(np.abs(run.v[I][l])) > 0)
    ref, _ = sc.read_reference(adi, 2, run.thd)
    d = np.max(np.abs(ours[I][0] - ref[I][0])) / np.max(np.abs(ref[I][0]))
    print(f"resumed: {run.ite} iterations, against the oracle {e:.2e}, "
          f"against the reference's uninterrupted run {d:.1e}")
    assert e <= 1e-10
    assert 1e-5 < d < 1e-2


def test_reference_sod_deck_to_late_statistics(tmp_path, capsys):
    """The reference's Sod deck through the driver to statistics time 29
    (iteration 2500 or so: shock, contact and expansion well apart), its
    restart files 10, 20 and 29 against the reference's own, to the bars
    of statistics 1 and 2 (tests/test_sod_reference_cpu.py): the
    trajectory does not drift from the reference's."""
    import shutil
    from phd_amd import chemdata as cd
    from phd_amd import restart
    from phd_amd.run import main
    from tests import sodcase as sc
    from tests.test_sod_reference_cpu import ABS_TRANSVERSE, BARS
    if not os.path.exists(restart.IO_PATH):
        pytest.skip("libizem_io.so not built (needs the HDF5 C library)")
    deck = open(os.path.join(sc.GOLDEN, "sod.inp")).read()
    assert deck.count("NSTAT = 30") == 1 and deck.count("ITSHOW = 1\n") == 1
    (tmp_path / "sod.inp").write_text(deck.replace("NSTAT = 30", "NSTAT = 29")
                                      .replace("ITSHOW = 1\n", "ITSHOW = 500\n"))
    shutil.copy(sc.GRID, tmp_path / "grid.dat")
    shutil.copy(sc.TRANFIT, tmp_path / "tranfit.out")
    assert main([str(tmp_path / "sod.inp")]) == 0
    capsys.readouterr()
    adi = cd.Adim()
    thd = sc.sod_thd(adi)
    box = sc.sod_box(adi)
    I = (slice(None),) + box.interior()
    for k in (10, 20, 29):
        ours, t = restart.read(str(tmp_path / f"restart_{k:05d}.h5"), box, thd.names)
        ref, t_ref = sc.read_reference(adi, k, thd)
        assert abs(t / t_ref - 1.0) < 1e-13
        rel = {}
        for l in range(ours.shape[0]):
            d = np.max(np.abs(ours[I][l] - ref[I][l]))
            rel[l] = d / np.max(np.abs(ref[I][l])) if l in BARS else d
        print(f"statistics {k}: " + ", ".join(f"var {l} {e:.1e}" for l, e in rel.items()),
              flush=True)
        for l, bar in BARS.items():
            assert rel[l] <= bar, (k, l, rel[l])
        # rho v and rho w are rounding noise in both codes; the reference's
        # grows to 7.3e-13 by statistics 29
        for l in (2, 3):
            noise = np.max(np.abs(ref[I][l]))
            assert rel[l] <= max(ABS_TRANSVERSE, 3.0 * noise), (k, l, rel[l], noise)
