"""The coupled reacting run with learned chemistry: the flow step with the
chemistry surrogate inside reaction_selec, beside the same run with the BDF
path, and the flame diagnostics the two are compared by.

The surrogate is trained on cell states the solver itself holds: a random
sample of the reacting box, aged along chemistry trajectories (chained BDF
substeps) so that the radical pools that build up over a run are in the
data, each aged state labelled by the BDF path over the step's half dt (and
half and twice that).  Nothing here has a counterpart in the reference; the
step it runs is the reference's main loop (izem.f90:254-277) with
reaction_selec choosing the surrogate (module_reaction.f90:47-90).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib
from . import autotune
from . import surrogate as su

RU_CGS = 8.314510e7  # erg/(mol K), as the mechanism tables


def reacting_cells(solver):
    """Every cell of the reacting box (CHEMKIN units): T [K], rho [g/cm^3],
    Y (n, 9), contiguous device tensors."""
    box, adi = solver.box, solver.adi
    g = box.ng
    if box.ndim == 3:
        sl = (slice(2 * g, box.nz - 1), slice(2 * g, box.ny - 1), slice(2 * g, box.nx - 1))
    else:
        sl = (slice(g, g + 1), slice(2 * g, box.ny - 1), slice(2 * g, box.nx - 1))
    rho = solver.v[0][sl].reshape(-1)
    T = solver.T[sl].reshape(-1) * adi.T_ref
    rho_i = 1.0 / rho
    Y = torch.stack([solver.v[5 + l][sl].reshape(-1) * rho_i for l in range(9)], dim=1)
    return T.contiguous(), (rho * (adi.rho_ref * 1.0e-3)).contiguous(), Y.contiguous()


def heat_release(solver, T, rho, Y) -> torch.Tensor:
    """Heat-release rate -sum_k h_k wdot_k of each cell [W/m^3]: wdot from
    the library's CKWYR (izem_ckwyr), molar enthalpies from the
    mechanism's NASA-7 tables (CKHMS form, chemkin/vode.f:228-285)."""
    n = T.numel()
    wdot = torch.empty((n, 9), dtype=torch.float64, device=T.device)
    _lib.check(solver.lib.izem_ckwyr(solver.ctx, n, C.c_void_p(rho.data_ptr()),
                                     C.c_void_p(T.data_ptr()), C.c_void_p(Y.data_ptr()),
                                     C.c_void_p(wdot.data_ptr())), "ckwyr")
    t = solver.mech.tables()
    a = torch.tensor(np.asarray(t["a"])[:9], dtype=torch.float64, device=T.device)  # (9, 2, 7)
    tmid = torch.tensor(np.asarray(t["tmid"])[:9], dtype=torch.float64, device=T.device)
    hi = (T[:, None] > tmid[None, :])[..., None]
    c = torch.where(hi, a[None, :, 1, :], a[None, :, 0, :])  # (n, 9, 7)
    Tc = T[:, None]
    h_RT = (c[..., 0] + c[..., 1] * Tc / 2.0 + c[..., 2] * Tc ** 2 / 3.0 +
            c[..., 3] * Tc ** 3 / 4.0 + c[..., 4] * Tc ** 4 / 5.0 + c[..., 5] / Tc)
    q = -(RU_CGS * Tc * h_RT * wdot).sum(dim=1)  # erg/(cm^3 s)
    return 0.1 * q


def flame_stats(solver) -> dict:
    """Peak temperature, peak and mean heat-release rate over the reacting
    box of the solver's current state."""
    T, rho, Y = reacting_cells(solver)
    q = heat_release(solver, T, rho, Y)
    solver.sync()
    return {"T_peak_K": float(T.max()), "hrr_peak_W_m3": float(q.max()),
            "hrr_mean_W_m3": float(q.mean())}


def train_flow_surrogate(solver, dt_half: float, n_cells: int = 1 << 18,
                         ages=(0, 10, 100, 1000), epochs: int = 8, batch: int = 65536,
                         lr: float = 3e-4, n_cal: int = 1 << 17, seed: int = 0,
                         verbose: bool = False):
    """Train and calibrate a surrogate for the solver's reaction step.

    n_cells random cells of the reacting box are aged by chained BDF
    substeps to each of `ages` (in units of dt_half); every aged state is
    labelled over 0.5, 1 and 2 dt_half.  The indicator is calibrated (max
    |T - T_BDF| within 0.75 x 2 K) on n_cal other cells aged the same way
    and labelled over dt_half."""
    t0 = time.perf_counter()

    def aged(n, sd):
        Ts, rs, Ys = autotune.sample_reacting_cells(solver, n, seed=sd)
        T, r, Y = Ts.cpu().numpy(), rs.cpu().numpy(), Ys.cpu().numpy()
        out, done = [], 0
        for age in sorted(ages):
            if age > done:
                L = su.label_cells(solver, T, r, Y, (age - done) * dt_half)
                T, Y = L.T1.cpu().numpy(), L.Y1.cpu().numpy()
                done = age
            out.append((T.copy(), r, Y.copy()))
        return out

    train = []
    for T, r, Y in aged(n_cells, seed + 1):
        for f in (0.5, 1.0, 2.0):
            train.append(su.label_cells(solver, T, r, Y, f * dt_half))
    sur = su.Surrogate(solver, seed=seed)
    sur.set_normalisation(train)
    hist = sur.fit(train, epochs=epochs, batch=min(batch, sum(p.T.numel() for p in train)),
                   lr=lr, verbose=verbose)
    ntrain = sum(p.T.numel() for p in train)
    del train
    cal_pairs = [su.label_cells(solver, T, r, Y, dt_half) for T, r, Y in aged(n_cal, seed + 2)]
    cal = sur.calibrate(su.concat(cal_pairs))
    info = {"train_samples": ntrain, "epochs": epochs, "lr": lr, "final_loss": hist[-1],
            "ages_in_half_steps": list(ages), "calibration_accept_fraction":
            cal["accept_fraction"], "tol_T_K": cal["tol_T"], "dT_max_K": cal["dT_max"],
            "train_and_calibrate_s": time.perf_counter() - t0}
    return sur, info


def indicator_breakdown(sur, T, rho, Y, dt: float) -> dict:
    """Which of the surrogate's indicator tests send cells (CHEMKIN units,
    device tensors) to the BDF path: the fractions failing the feature
    range, the mass-fraction floor (with positive T) and, among the others,
    the energy-temperature gap and the dT limit."""
    n = T.numel()
    X = sur.features(T, rho, Y, dt)
    bits = X[:, :su.NF].cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    feat_bad = torch.from_numpy((np.abs(bits.view(np.float32)) > sur.feat_lim).any(axis=1))
    T1, Y1 = T.clone(), Y.clone()
    Te = torch.empty_like(T)
    flag = sur.apply(T1, rho, Y1, dt, tol_T=1e30, dT_max=1e30, T_energy=Te).cpu().bool()
    ok = ~flag
    gap = (Te - T1).abs().cpu()
    dT = (T1 - T).abs().cpu()
    return {"cells": n, "feature_range": float(feat_bad.double().mean()),
            "mass_fraction_floor": float((flag & ~feat_bad).double().mean()),
            "energy_gap": float((ok & (gap > sur.tol_T)).double().mean()),
            "dT_limit": float((ok & (gap <= sur.tol_T) & (dT > sur.dT_max)).double().mean()),
            "accepted": float((ok & (gap <= sur.tol_T) & (dT <= sur.dT_max)).double().mean()),
            "gap_q50_q99_K": [float(x) for x in np.quantile(gap[ok].numpy(), [0.5, 0.99])]
            if ok.any() else None,
            "dT_q50_q99_K": [float(x) for x in np.quantile(dT[ok].numpy(), [0.5, 0.99])]
            if ok.any() else None}
