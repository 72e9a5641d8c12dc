#!/usr/bin/env python3
"""Benchmark of the IZEM reacting-flow step on one MI355X.

Workload (the headline configuration): 3-D periodic 256^3 decaying reacting
HIT, premixed H2/air at phi = 0.8, T0 = 1000 K with 50 K Gaussian-random
fluctuations, u_rms = 20 m/s, FP64, 9 species / 21-reaction mechanism.  A
step is one iteration of the reference main loop (devel-izem/src/izem.f90:
254-277): timestep -> chemistry(dt/2) -> SSP-RK3(dt) -> chemistry(dt/2),
WENO5-JS characteristic LLF fluxes, viscous + mixture-averaged diffusion,
Newton T inversion, batched BDF chemistry (rtol 1e-6, atol 1e-12).

Prints one JSON line (metric, value in cell-steps/s, roofline of the
dominant kernel measured with HIP events, CPU baseline timed on the host).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s 256^3 reacting H2 (FP64); chem cells/s surrogate vs VODE @|dT|<=2K"
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0  # dense MFMA
PEAK_HBM_GBS = 8000.0


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(kernel: str, n: int):
    """HBM bytes per launch of a kernel from its committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel factor)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{n}cube_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def _cpu_step_worker(arg):
    """One full reacting step (timestep, chemistry(dt/2), SSP-RK3,
    chemistry(dt/2)) on an n^3 box of the workload recipe: the oracle's C
    restatement of the flow path and the reference's compiled
    PSR_CHEMKIN/DVODE per cell.  Returns the step's seconds."""
    import ctypes as C

    import numpy as np

    from oracle.oracle import Oracle, RefChem
    from phd_amd import chemdata as cd
    from phd_amd import ics
    from tests.flowcase import OracleRun, P

    n, seed = arg
    adi = cd.Adim()
    thd = cd.build_thd(adi)
    box, v, T = ics.config2_state(adi, thd, n=n, seed=seed, L=0.01 * n / 256)
    run = OracleRun(Oracle(), adi, thd, box, v, T)
    run.prim_all()
    rc = RefChem()
    opts = run.opts(vis=1, reaction=1)
    t0 = time.perf_counter()
    ts = np.zeros(4)
    dtmin = C.c_double(1.0)
    run.o.lib.ora_timestep(C.byref(run.g), C.byref(run.t), C.byref(run.a), C.byref(opts),
                           P(run.d_i[0]), P(run.d_i[1]), P(run.d_i[2]), P(run.cp), P(run.W_i),
                           P(run.T), P(run.v), C.c_double(0.0), P(ts), C.byref(dtmin))
    dt = dtmin.value
    rho_cgs = adi.rho_ref * 1e-3

    def chem():
        g = box.ng
        for k in range(g, box.nz - g - 1):
            for j in range(g, box.ny - g - 1):
                for i in range(g, box.nx - g - 1):
                    c = (k + g, j + g, i + g)
                    rho0 = run.v[(0,) + c]
                    Ya = run.v[(slice(5, 14),) + c] / rho0
                    Tn, Yn, ist, *_ = rc.psr(run.T[c] * adi.T_ref, rho0 * rho_cgs, Ya,
                                             0.5 * dt * adi.time_ref, 1e-6, 1e-12)
                    run.T[c] = Tn / adi.T_ref
                    Yc = np.clip(Yn[:8], 0, 1)
                    run.v[(slice(5, 13),) + c] = rho0 * Yc
                    run.v[(13,) + c] = rho0 * (1 - Yc.sum())
        run.o.lib.ora_upd_prim_var_woT(0, C.byref(run.g), C.byref(run.t), P(run.v), P(run.T),
                                       P(run.W_i), P(run.cp), P(run.ha))
        run.o.lib.ora_comm_cons(C.byref(run.g), P(run.v))
        run.o.lib.ora_upd_prim_var_ghost(C.byref(run.g), C.byref(run.t), 1, P(run.v),
                                         P(run.W_i), P(run.T), P(run.cp), P(run.ha))

    chem()
    sens = np.zeros(box.shape, np.uint8)
    run.o.lib.ora_rk3(C.byref(run.g), C.byref(run.t), C.byref(run.a), C.byref(opts),
                      C.c_double(dt), P(run.d_i[0]), P(run.d_i[1]), P(run.d_i[2]), P(run.T),
                      P(run.W_i), P(run.cp), P(run.ha), sens.ctypes.data_as(C.POINTER(C.c_uint8)),
                      P(run.v), None)
    chem()
    return time.perf_counter() - t0


def _cpu_c3_worker(arg):
    """m independent cells of the batched-chemistry workload through the
    reference's compiled PSR_CHEMKIN/DVODE over 1e-6 s (rtol 1e-6, atol
    1e-12, the GPU path's tolerances).  Returns (seconds, BDF steps)."""
    import numpy as np

    from oracle.oracle import RefChem
    from phd_amd import chemdata as cd
    from phd_amd import ics

    m, seed = arg
    thd = cd.build_thd(cd.Adim())
    T, p, Y = ics.config3_cells(thd, m, seed=seed)
    Wmix = 1.0 / (Y / (np.array(thd.W) * 1000.0)).sum(axis=1)
    rho = (p * 10.0) * Wmix / (8.314510e7 * T)
    rc = RefChem()
    nst = 0
    t0 = time.perf_counter()
    for c in range(m):
        nst += rc.psr(T[c], rho[c], Y[c], 1.0e-6, 1.0e-6, 1.0e-12)[3]
    return time.perf_counter() - t0, nst


def cpu_baseline(n: int = 48, m: int = 4096):
    """The CPU path on the host's cores (at most 16, the GPU's share of the
    box): one process per core, each stepping its own n^3 box of the
    workload recipe (processes run together; no halo exchange, so the sum
    of their rates bounds an MPI run of the same cores from above), and
    the same processes each integrating m batched-chemistry cells.  Must
    run before the parent process touches the GPU (the workers are
    forked)."""
    import multiprocessing as mp

    procs = max(1, min(16, len(os.sched_getaffinity(0))))
    with mp.get_context("fork").Pool(procs) as pool:
        t0 = time.perf_counter()
        step_s = pool.map(_cpu_step_worker, [(n, 2 + r) for r in range(procs)])
        wall_step = time.perf_counter() - t0
        t0 = time.perf_counter()
        c3 = pool.map(_cpu_c3_worker, [(m, 300 + r) for r in range(procs)])
        wall_c3 = time.perf_counter() - t0
    rate = sum(n ** 3 / el for el in step_s)
    c3_rate = sum(m / el for el, _ in c3)
    return {"value": rate, "unit": "cell-steps/s", "cores": procs, "kind": "port",
            "sample": f"{procs} processes, each one full reacting step (timestep, 2 x chemistry, "
                      f"SSP-RK3) on its own {n}^3 box of the same recipe: the oracle's C "
                      f"restatement of the flow path + the reference's compiled "
                      f"PSR_CHEMKIN/DVODE per cell; step {min(step_s):.1f}-{max(step_s):.1f} s "
                      f"per process, {wall_step:.1f} s wall with set-up",
            "c3_chemistry": {"cells_per_s": c3_rate, "cells": procs * m, "dt_s": 1.0e-6,
                             "rtol": 1.0e-6, "atol": 1.0e-12, "cores": procs,
                             "kind": "reference",
                             "bdf_steps_mean": sum(k for _, k in c3) / (procs * m),
                             "sample": f"{procs} processes x {m} independent cells of the "
                                       f"batched-chemistry recipe through the reference's "
                                       f"compiled PSR_CHEMKIN/DVODE; {wall_c3:.1f} s wall"}}


def chem_vode_rate(s, thd, ncell: int):
    import ctypes as C

    import numpy as np
    import torch

    from phd_amd import ics

    T, p, Y = ics.config3_cells(thd, ncell, seed=3)
    Wmix = 1.0 / (Y / (np.array(thd.W) * 1000.0)).sum(axis=1)
    rho = (p * 10.0) * Wmix / (8.314510e7 * T)
    d = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    drho = d(rho)
    nst = torch.zeros(ncell, dtype=torch.int32, device="cuda:0")
    el, rc = {}, 0
    # index order, then stiffest first (izem_chem_order): same end states,
    # the waves of the second hold cells of like step counts
    for order in (0, 1):
        dT, dY = d(T), d(Y)
        rc |= s.lib.izem_chem_order(s.ctx, order)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc |= s.lib.izem_psr_batch(s.ctx, ncell, C.c_void_p(dT.data_ptr()),
                                   C.c_void_p(drho.data_ptr()), C.c_void_p(dY.data_ptr()), 1.0e-6,
                                   C.c_void_p(nst.data_ptr()))
        s.sync()
        torch.cuda.synchronize()
        el[order] = time.perf_counter() - t0
    rc |= s.lib.izem_chem_order(s.ctx, 0)
    st = nst.cpu().numpy()
    return {"cells": ncell, "dt_s": 1e-6, "queue_order": "stiffness",
            "vode_cells_per_s": ncell / el[1], "seconds": el[1],
            "vode_cells_per_s_index_order": ncell / el[0],
            "bdf_steps_mean": float(st.mean()), "bdf_steps_max": int(st.max()), "rc": rc}


def surrogate_rate(s, adi, n: int, dt_half: float):
    """Train the chemistry surrogate on chemistry trajectories of cells of
    the benchmark state (labelled by the BDF path), calibrate its indicator
    on other cells, and time surrogate + fallback against the BDF path on a
    third sample at the step's half dt."""
    from phd_amd import autotune
    from phd_amd import surrogate as su

    t0 = time.perf_counter()
    dts = [0.5 * dt_half, dt_half, 2.0 * dt_half] * 2
    train = []
    for seed in (11, 12):
        Ts, rs, Ys = autotune.sample_reacting_cells(s, n, seed=seed)
        train += su.trajectory(s, Ts.cpu().numpy(), rs.cpu().numpy(), Ys.cpu().numpy(), dts)
    sur = su.Surrogate(s, seed=0)
    sur.set_normalisation(train)
    hist = sur.fit(train, epochs=8, batch=65536, lr=1e-3)
    Tc, rc, Yc = autotune.sample_reacting_cells(s, n // 2, seed=13)
    cal = sur.calibrate(su.concat(su.trajectory(s, Tc.cpu().numpy(), rc.cpu().numpy(),
                                                Yc.cpu().numpy(), [dt_half] * 3)))
    t_train = time.perf_counter() - t0
    Tt, rt, Yt = autotune.sample_reacting_cells(s, 2 * n, seed=14)
    L = su.label_cells(s, Tt.cpu().numpy(), rt.cpu().numpy(), Yt.cpu().numpy(), dt_half)
    res = su.evaluate(sur, L)
    # the fused inference kernel alone, event-timed on the library stream
    from phd_amd.solver import profile_read, profile_start
    X = sur.features(L.T, L.rho, L.Y, L.dt)
    sur.infer(X)
    profile_start(s)
    for _ in range(5):
        sur.infer(X)
    c_ms, c_cnt = profile_read(s)["chemistry"]
    avg_s = c_ms / c_cnt / 1e3
    flop = 2.0 * X.shape[0] * sum(k * n for k, n in su.TRUE_DIMS)
    res["roofline_mlp"] = {"kernel": "k_mlp_fused", "bound": "mfma",
                           "achieved": flop / avg_s / 1e12, "peak": PEAK_BF16_TFLOPS,
                           "unit": "TFLOP/s", "frac": flop / avg_s / 1e12 / PEAK_BF16_TFLOPS,
                           "avg_launch_ms": avg_s * 1e3, "rows": int(X.shape[0]),
                           "flop_per_row": flop / X.shape[0]}
    res.update({"train_samples": sum(p.T.numel() for p in train), "train_epochs": 8,
                "final_loss": hist[-1], "train_and_calibrate_s": t_train,
                "tol_T_K": cal["tol_T"], "dT_max_K": cal["dT_max"],
                "cells": "cells of the 256^3 reacting state, Strang half step"})
    return res


def surrogate_c3(s, thd, n_train_cells: int = 1 << 22, n_eval: int = 1 << 22,
                 dt: float = 1.0e-6):
    """The chemistry surrogate on the batched-chemistry workload (independent
    cells, T in [900, 2000] K, p in [1, 10] atm, dt = 1e-6 s), where a substep
    moves T by up to ~1500 K: trained on 16 M states labelled by the BDF
    path (four chained 1e-6 s substeps of 2^22 cells, so that burning and
    burnt states are in the data), Adam lr 3e-4, batch 65536; indicator
    calibrated on 2^20 held-out cells for max |T - T_BDF| <= 2 K; timed
    against the BDF path on 2^22 fresh cells.  The error of a predictor that
    returns no change is given beside the surrogate's."""
    import torch

    from phd_amd import surrogate as su

    t0 = time.perf_counter()
    T, r, Y = su.reactor_cells(thd, n_train_cells, seed=31)
    train = su.trajectory(s, T, r, Y, [dt] * 4)
    t_label = time.perf_counter() - t0
    sur = su.Surrogate(s, seed=0)
    sur.set_normalisation(train)
    hist = sur.fit(train, epochs=8, batch=65536, lr=3e-4)
    del train
    torch.cuda.empty_cache()
    Tc, rc, Yc = su.reactor_cells(thd, 1 << 20, seed=32)
    cal = sur.calibrate(su.label_cells(s, Tc, rc, Yc, dt))
    t_train = time.perf_counter() - t0
    Te, re_, Ye = su.reactor_cells(thd, n_eval, seed=33)
    L = su.label_cells(s, Te, re_, Ye, dt)
    res = su.evaluate(sur, L)
    res["no_change_max_abs_dT_K"] = float((L.T1 - L.T).abs().max())
    res["no_change_frac_cells_over_2K"] = float(((L.T1 - L.T).abs() > 2.0).double().mean())
    res.update({"train_samples": 4 * n_train_cells, "train_epochs": 8, "lr": 3e-4,
                "batch": 65536, "final_loss": hist[-1], "label_s": t_label,
                "train_and_calibrate_s": t_train, "tol_T_K": cal["tol_T"],
                "dT_max_K": cal["dT_max"], "calibration_accept_fraction": cal["accept_fraction"],
                "cells": "independent reactor cells of the batched-chemistry workload"})
    return res


def coupled_shear_layer(adi, thd, mech, n: int, steps: int, tune, flux_mode: int):
    """The coupled learned-chemistry workload: the periodic double shear
    layer on n^3 cells (fuel 50 % H2 / 50 % N2 at 300 K and +100 m/s, air
    at 1200 K and -100 m/s, vorticity thickness 0.5 mm, 8 mm box), the
    surrogate trained on the run's own cells (2^20 cells aged by 0, 10, 100
    and 1000 half steps, each labelled over 0.5, 1 and 2 half steps: 12.6 M
    states; Adam lr 3e-4, batch 65536, 8 epochs) and calibrated for 2 K;
    then the same steps from the same state with reaction_selec on the BDF
    path and on the surrogate, each timed over `steps` graph-replayed steps
    after one eager and one graph warm-up step, with one event-profiled
    eager step for the chemistry time.  Peak T and heat-release rate of the
    two end states are compared."""
    import numpy as np
    import torch

    from phd_amd import _lib, coupled, ics
    from phd_amd.solver import Options, Solver, profile_read, profile_start

    t0 = time.perf_counter()
    box, v, T = ics.config5_state(adi, thd, n=n)
    s = Solver(box, Options(chem_tune=tune, flux_mode=flux_mode), adi=adi, thd=thd, mech=mech)
    s.load_state(v, T)
    del v, T
    t_state = time.perf_counter() - t0
    ts0 = s.time_state()
    s.timestep()
    dt_half = 0.5 * float(s.time_state()[4]) * adi.time_ref
    s.set_time_state(ts0)
    log(f"shear layer {n}^3 resident ({t_state:.0f} s); training the surrogate on its cells")
    sur, info = coupled.train_flow_surrogate(s, dt_half, n_cells=1 << 20,
                                             ages=(0, 10, 100, 1000), epochs=8, batch=65536,
                                             lr=3e-4, n_cal=1 << 18)
    s.attach_surrogate(sur)
    snap = [x.clone() for x in (s.T, s.v, s.W_i, s.cp, s.ha)]
    fields = {}
    out = {"workload": f"3-D periodic double shear layer {n}^3, H2/N2 (300 K, +100 m/s) "
                       f"against air (1200 K, -100 m/s), vorticity thickness 0.5 mm, 8 mm box, "
                       f"2 m/s broadband perturbation in the layers",
           "cells": n ** 3, "dt_s": 2.0 * dt_half, "surrogate_training": info}
    for which in ("dvode", "surrogate"):
        for dst, src in zip((s.T, s.v, s.W_i, s.cp, s.ha), snap):
            dst.copy_(src)
        s.sensor.zero_()
        s.set_time_state(ts0)
        s.reac_selec(_lib.REAC_SURROGATE if which == "surrogate" else _lib.REAC_DVODE)
        s.step(graph=False)
        s.step(graph=True)
        s.sync()
        if which == "surrogate":
            s.surrogate_counts()  # clear
        log(f"shear layer: timing {steps} steps with reaction_selec = {which}")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for q in range(steps):
            s.step(graph=True)
            if steps > 50 and (q + 1) % 50 == 0:
                s.sync()  # a progress line every 50 steps of long runs
                log(f"shear layer ({which}): {q + 1} steps, "
                    f"{time.perf_counter() - t1:.1f} s")
        s.sync()
        torch.cuda.synchronize()
        el = time.perf_counter() - t1
        fall = None
        if which == "surrogate":
            cells, nf = s.surrogate_counts()
            fall = nf / max(cells, 1)
        profile_start(s)
        s.step(graph=False)
        s.sync()
        prof = profile_read(s)
        if which == "surrogate":
            s.surrogate_counts()
        st = coupled.flame_stats(s)
        Tc, rc, Yc = coupled.reacting_cells(s)
        fields[which] = (Tc, Yc, coupled.heat_release(s, Tc, rc, Yc))
        out[which] = {"ms_per_step": 1e3 * el / steps, "cell_steps_per_s": n ** 3 * steps / el,
                      "chemistry_ms_per_step": prof["chemistry"][0],
                      "kernel_ms_per_step": {k: v[0] for k, v in prof.items()},
                      "fallback_rate": fall, "status_bits": s.status(),
                      "steps_from_initial_state": steps + 3, **st}
    d, g = out["dvode"], out["surrogate"]
    out["chemistry_speedup"] = d["chemistry_ms_per_step"] / g["chemistry_ms_per_step"]
    out["T_peak_diff_K"] = g["T_peak_K"] - d["T_peak_K"]
    (Td, Yd, qd), (Tg, Yg, qg) = fields["dvode"], fields["surrogate"]
    major = [thd.names.index(nm) for nm in ("H2", "O2", "H2O", "N2")]
    out["max_abs_dT_K"] = float((Tg - Td).abs().max())
    out["max_abs_dY_major"] = float((Yg[:, major] - Yd[:, major]).abs().max())
    out["hrr_peak_diff_W_m3"] = g["hrr_peak_W_m3"] - d["hrr_peak_W_m3"]
    out["hrr_mean_rel_diff"] = (g["hrr_mean_W_m3"] - d["hrr_mean_W_m3"]) / abs(d["hrr_mean_W_m3"])
    out["hrr_l2_rel_diff"] = float((qg - qd).norm() / qd.norm())
    del fields, snap
    s.attach_surrogate(None)
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--chem-cells", type=int, default=1 << 22)
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-surrogate-c3", action="store_true")
    ap.add_argument("--shear-n", type=int, default=384,
                    help="cells per side of the coupled learned-chemistry shear layer (0: skip)")
    ap.add_argument("--shear-steps", type=int, default=5)
    ap.add_argument("--flux-mode", type=int, default=1,
                    help="Euler face kernels: 0 reference operation order (bit parity), "
                         "1 structured projections (within 1e-12 of the reference)")
    ap.add_argument("--chem-tune", type=str, default=None,
                    help="waves,blocks_per_cu,jac_age,kinetics for the chemistry kernels "
                         "(default: autotuned on a sample of the reacting cells)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one MI355X: --gpus {args.gpus} is not supported "
              f"(one process drives one GPU; a multi-GPU decomposition is not built)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from phd_amd import chemdata as cd
    from phd_amd import ics
    from phd_amd.flops import EULER_FLOP_PER_FACE, EULER_FLOP_PER_FACE_MODE
    from phd_amd.solver import Options, Solver, profile_read, profile_start

    cpu = None
    if not args.no_cpu_baseline:
        # before anything touches the GPU: the workers are forked
        log("CPU baseline (oracle flow + reference DVODE, one process per core)")
        cpu = cpu_baseline()
        log(f"CPU baseline: {cpu['value']:.3e} cell-steps/s on {cpu['cores']} cores, "
            f"batched chemistry {cpu['c3_chemistry']['cells_per_s']:.3e} cells/s")
    n = args.n
    adi = cd.Adim()
    thd = cd.build_thd(adi)
    mech = cd.read_mechanism()
    log(f"building the {n}^3 reacting HIT state")
    box, v, T = ics.config2_state(adi, thd, n=n, seed=2)
    tune = tuple(int(x) for x in args.chem_tune.split(",")) if args.chem_tune else None
    s = Solver(box, Options(chem_tune=tune, flux_mode=args.flux_mode), adi=adi, thd=thd,
               mech=mech)
    s.load_state(v, T)
    del v, T
    tuned = None
    if tune is None:
        from phd_amd import autotune
        ts0 = s.time_state()
        s.timestep()  # the step's dt, to sample the chemistry at its half step
        dt_half = 0.5 * float(s.time_state()[4]) * adi.time_ref
        s.set_time_state(ts0)
        Ts, rs, Ys = autotune.sample_reacting_cells(s, 1 << 20)
        tuned = autotune.tune(s, Ts, rs, Ys, dt_half, key=f"reacting-hit-{n}", use_cache=False)
        log(f"chemistry kernels tuned: waves {tuned[0]}, blocks/CU {tuned[1]}, "
            f"kinetics {autotune.KINETICS[tuned[3]]}, "
            f"{'stiffness' if tuned[5] else 'index'} order ({tuned[4]:.3e} cells/s)")
        del Ts, rs, Ys
    log("state resident on the GPU; warm-up (eager, then the captured graph)")
    for _ in range(args.warmup):
        s.step(graph=False)
        s.step(graph=True)
    s.sync()
    torch.cuda.synchronize()
    # timed: the main-loop iteration replayed from its HIP graph (izem_step)
    log(f"timing {args.steps} steps (graph)")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for q in range(args.steps):
        s.step(graph=True)
    s.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the same number of eager steps with every kernel bracketed by events:
    # the per-kernel breakdown and the rooflines
    log(f"timing {args.steps} steps (eager, event-profiled)")
    chem_stats = np.zeros(4, dtype=np.uint64)
    if s.opts.reaction:
        s.lib.izem_chem_stats(s.ctx, chem_stats.ctypes.data)  # clear the counters
    profile_start(s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for q in range(args.steps):
        s.step(graph=False)
    s.sync()
    torch.cuda.synchronize()
    el_eager = time.perf_counter() - t1
    prof = profile_read(s)
    if s.opts.reaction:
        s.lib.izem_chem_stats(s.ctx, chem_stats.ctypes.data)
    status = s.status()
    ts = s.time_state()
    ncell = n ** 3
    value = ncell * args.steps / el
    ms_step = 1e3 * el / args.steps
    ms_eager = 1e3 * el_eager / args.steps
    ms_kernels = sum(v[0] for v in prof.values()) / args.steps
    e_ms, e_cnt = prof["euler_faces"]
    nfaces = (n + 1) * n * n
    e_avg_s = e_ms / e_cnt / 1e3
    flop_face = EULER_FLOP_PER_FACE_MODE[args.flux_mode]
    achieved = flop_face * nfaces / e_avg_s / 1e12
    euler_roof = {"kernel": "k_euler_faces (x, y, z)", "bound": "fp64-valu",
                  "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                  "frac": achieved / PEAK_FP64_TFLOPS,
                  "traffic": read_traffic("euler_faces", n),
                  "avg_launch_ms": e_avg_s * 1e3, "launches": e_cnt,
                  "flop_per_face": flop_face, "faces_per_launch": nfaces,
                  "flux_mode": args.flux_mode,
                  "flop_per_face_reference_order": EULER_FLOP_PER_FACE}
    c_ms, c_cnt = prof["chemistry"]
    chem_roof = None
    if c_cnt:
        from phd_amd.flops import chem_flops
        c_avg_s = c_ms / c_cnt / 1e3
        c_flop = chem_flops(chem_stats) / c_cnt
        c_ach = c_flop / c_avg_s / 1e12
        chem_roof = {"kernel": "k_reaction_queue", "bound": "fp64-valu", "achieved": c_ach,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": c_ach / PEAK_FP64_TFLOPS,
                     "traffic": read_traffic("reaction_queue", n),
                     "avg_launch_ms": c_avg_s * 1e3, "launches": c_cnt,
                     "flop_per_launch": c_flop,
                     "work_per_launch": {"bdf_steps": float(chem_stats[0]) / c_cnt,
                                         "rhs_evals": float(chem_stats[1]) / c_cnt,
                                         "jacobians": float(chem_stats[2]) / c_cnt,
                                         "lu_solves": float(chem_stats[3]) / c_cnt}}
    # the memory-bound flow classes against the HBM peak: algorithmic bytes
    # per cell (phd_amd.flops) x cells per launch / event-timed launch time
    from phd_amd.flops import BYTES_PER_CELL, STEP_BYTES_PER_CELL
    hbm_roofs = {}
    for cls, cells in (("rhs_update", ncell), ("viscous", ncell + 6 * 3 * n * n)):
        ms, cnt = prof[cls]
        if not cnt:
            continue
        avg_s = ms / cnt / 1e3  # one record per RK stage (viscous: prim + fused)
        by = BYTES_PER_CELL[cls] * cells
        hbm_roofs[cls] = {"kernel": "k_rhs_update" if cls == "rhs_update"
                          else "k_visc_prim + k_visc_fused", "bound": "hbm",
                          "achieved": by / avg_s / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": by / avg_s / 1e9 / PEAK_HBM_GBS,
                          "traffic": read_traffic(cls, n), "avg_stage_ms": avg_s * 1e3,
                          "bytes_per_cell": BYTES_PER_CELL[cls], "cells": cells}
    step_hbm = {"bytes_per_cell_step": STEP_BYTES_PER_CELL,
                "achieved": value * STEP_BYTES_PER_CELL / 1e9, "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": value * STEP_BYTES_PER_CELL / 1e9 / PEAK_HBM_GBS}
    # the roofline object describes the kernel with the largest share of
    # the step; the other one is reported beside it
    dominant = chem_roof if (chem_roof and c_ms > e_ms) else euler_roof
    log(f"{ms_step:.1f} ms/step with the graph, {ms_eager:.1f} eager; kernel time per eager "
        f"step: " + ", ".join(f"{k} {v[0] / args.steps:.1f} ms" for k, v in prof.items()))
    surr = None
    if s.opts.reaction and not args.no_surrogate:
        log("chemistry surrogate: train, calibrate, time against the BDF path")
        surr = surrogate_rate(s, adi, 1 << 19, 0.5 * float(ts[4]) * adi.time_ref)
        log(f"surrogate {surr['surrogate_cells_per_s']:.3e} cells/s vs BDF "
            f"{surr['vode_cells_per_s']:.3e} (x{surr['speedup']:.2f}), fallback "
            f"{100 * surr['fallback_rate']:.2f} %, max |dT| {surr['max_abs_dT_K']:.2e} K")
    surr_c3 = None
    if s.opts.reaction and not args.no_surrogate_c3:
        log("chemistry surrogate on independent cells at dt = 1e-6 s (16 M training states)")
        surr_c3 = surrogate_c3(s, thd)
        log(f"surrogate {surr_c3['surrogate_cells_per_s']:.3e} cells/s vs BDF "
            f"{surr_c3['vode_cells_per_s']:.3e} (x{surr_c3['speedup']:.2f}), fallback "
            f"{100 * surr_c3['fallback_rate']:.1f} %, max |dT| {surr_c3['max_abs_dT_K']:.2f} K "
            f"(no-change predictor {surr_c3['no_change_max_abs_dT_K']:.1f} K)")
    chem = None
    if args.chem_cells > 0:
        log(f"batched VODE-path chemistry on {args.chem_cells} independent cells")
        chem = chem_vode_rate(s, thd, args.chem_cells)
    s.close()
    del s
    torch.cuda.empty_cache()
    shear = None
    if args.shear_n > 0:
        shear_tune = tune if tune is not None else (tuned[:4] if tuned else None)
        shear = coupled_shear_layer(adi, thd, mech, args.shear_n, args.shear_steps, shear_tune,
                                    args.flux_mode)
        log(f"shear layer {args.shear_n}^3: {shear['dvode']['ms_per_step']:.0f} ms/step with "
            f"the BDF path, {shear['surrogate']['ms_per_step']:.0f} with the surrogate "
            f"(chemistry x{shear['chemistry_speedup']:.2f}, fallback "
            f"{100 * shear['surrogate']['fallback_rate']:.1f} %), peak T difference "
            f"{shear['T_peak_diff_K']:.2e} K, max |dT| {shear['max_abs_dT_K']:.2e} K, "
            f"heat-release field {100 * shear['hrr_l2_rel_diff']:.2f} % (L2)")
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "ms_per_step_eager": ms_eager,
        "ms_per_step_kernels": ms_kernels,
        "ms_per_step_unattributed_eager": ms_eager - ms_kernels,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D periodic {n}^3 decaying reacting HIT, H2/air phi=0.8, "
                               f"T0=1000 K +- 50 K, u_rms=20 m/s, 14 conserved variables, "
                               f"WENO5-JS LLF + viscous + mixture-averaged diffusion, SSP-RK3 "
                               f"with Strang BDF chemistry (rtol 1e-6, atol 1e-12)",
                   "cells": ncell, "dt_s": float(ts[4] * adi.time_ref),
                   "flux_mode": args.flux_mode,
                   "status_bits": status},
        "roofline": dominant,
        "roofline_kernels": {"k_reaction_queue": chem_roof, "k_euler_faces": euler_roof,
                             **{v["kernel"]: v for v in hbm_roofs.values()}},
        "step_hbm_roofline": step_hbm,
        "kernel_ms_per_step": {k: v[0] / args.steps for k, v in prof.items()},
        "chem_tuning": (None if tuned is None else
                        {"waves_per_simd": tuned[0], "blocks_per_cu": tuned[1],
                         "jac_age": tuned[2], "kinetics": autotune.KINETICS[tuned[3]],
                         "queue_order": "stiffness" if tuned[5] else "index",
                         "sample_cells_per_s": tuned[4]}),
        "chemistry": chem,
        "chem_surrogate": surr,
        "chem_surrogate_c3": surr_c3,
        "coupled_shear_layer": shear,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
