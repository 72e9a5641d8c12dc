"""Autotuner of the batched chemistry kernels.

The persistent work-queue kernels (izem_psr_batch, izem_reaction_dvode) are
compiled in several variants: table-driven kinetics at a register budget of
1, 2 or 4 waves per SIMD, straight-line kinetics at 1 wave per SIMD with a
finite-difference Jacobian or with its mass-fraction columns in closed form,
and
launched with a chosen number of resident 64-lane blocks per CU.  Which combination is fastest depends on
the cells' stiffness (steps per cell), the device and the mechanism.  The
tuner times each candidate with izem_psr_batch on a sample of the cells the
caller will integrate, keeps the fastest and applies it with
izem_chem_tune; with that variant it then times izem_psr_batch's stiffness
order (izem_chem_order) against index order and keeps the faster.  Every
candidate integrates the same BDF method with the same tolerances and
Jacobian reuse ages (DVODE's MSBP = 20, MSBJ = 50); the kinetics forms
differ at round-off level and the Jacobian forms in the iteration matrix
only, so the choice changes speed, and results within the tolerance.

Results are cached per (device, mechanism hash, workload key) in a JSON
file, by default ~/.cache/phd_amd/chem_tune.json.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np
import torch

from . import _lib
from .codegen_mech import mechanism_hash

CANDIDATES = ([(1, b, k) for k in (2, 1) for b in (4, 2)] +
              [(w, b, 0) for w in (4, 2, 1) for b in (4 * w, 2 * w)])
KINETICS = {0: "tables", 1: "straight-line", 2: "straight-line, closed-form Jacobian"}
JAC_AGE = 50  # DVODE MSBJ
ORDER_MARGIN = 1.05  # the stiffness order must be at least 5 % faster to be kept


def _cache_path() -> str:
    root = os.environ.get("PHD_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache",
                                                        "phd_amd"))
    return os.path.join(root, "chem_tune.json")


def _load_cache() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(c: dict) -> None:
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def sample_reacting_cells(solver, n: int, seed: int = 0):
    """A random sample of the reacting box of the solver's state in CHEMKIN
    units: T [K], rho [g/cm^3], Y (n, 9) as contiguous device tensors."""
    box, adi = solver.box, solver.adi
    g = box.ng
    if box.ndim == 3:
        sl = (slice(2 * g, box.nz - 1), slice(2 * g, box.ny - 1), slice(2 * g, box.nx - 1))
    else:
        sl = (slice(g, g + 1), slice(2 * g, box.ny - 1), slice(2 * g, box.nx - 1))
    rho = solver.v[0][sl].reshape(-1)
    T = solver.T[sl].reshape(-1)
    Y = torch.stack([solver.v[5 + l][sl].reshape(-1) for l in range(9)], dim=1)
    gen = torch.Generator(device="cpu").manual_seed(seed)
    idx = torch.randint(0, rho.numel(), (min(n, rho.numel()),), generator=gen).to(rho.device)
    r = rho[idx]
    return ((T[idx] * adi.T_ref).contiguous(), (r * (adi.rho_ref * 1.0e-3)).contiguous(),
            (Y[idx] / r[:, None]).contiguous())


def tune(solver, T_K: torch.Tensor, rho: torch.Tensor, Y: torch.Tensor, dt: float,
         key: str = "", candidates=None, use_cache: bool = True, verbose: bool = False):
    """Time the candidates on (T_K, rho, Y) advanced by dt seconds, apply the
    fastest to the solver's context and return (waves, blocks_per_cu,
    jac_age, kinetics, cells_per_s, order)."""
    lib, ctx = solver.lib, solver.ctx
    dev = torch.cuda.get_device_properties(T_K.device).name
    mech_hash = mechanism_hash(solver.mech)
    ck = f"{dev}|{mech_hash}|{key}|n={T_K.numel()}"
    cache = _load_cache() if use_cache else {}
    if ck in cache:
        w, b, j, k, rate, o = (cache[ck] + [0])[:6]
        if lib.izem_chem_tune(ctx, w, b, j, k) == 0 and lib.izem_chem_order(ctx, o) == 0:
            return (w, b, j, k, rate, o)
    n = T_K.numel()
    best = None

    def rate_of(label):
        el = []
        for _ in range(2):
            t, r, y = T_K.clone(), rho.clone(), Y.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc = lib.izem_psr_batch(ctx, n, C.c_void_p(t.data_ptr()), C.c_void_p(r.data_ptr()),
                                    C.c_void_p(y.data_ptr()), float(dt), None)
            solver.sync()
            el.append(time.perf_counter() - t0)
            if rc != 0:
                return None
        if solver.status() != 0:
            return None
        if verbose:
            print(f"[autotune] {label}: {n / min(el):.3e} cells/s", flush=True)
        return n / min(el)

    _lib.check(lib.izem_chem_order(ctx, 0), "izem_chem_order")
    for (w, b, k) in (candidates or CANDIDATES):
        if lib.izem_chem_tune(ctx, w, b, JAC_AGE, k) != 0:
            continue  # e.g. straight-line kinetics for another mechanism
        rate = rate_of(f"waves {w} blocks/CU {b} kinetics {k}")
        if rate is not None and (best is None or rate > best[4]):
            best = (w, b, JAC_AGE, k, rate, 0)
    if best is None:
        raise _lib.IzemError("no chemistry kernel variant ran on the sample")
    _lib.check(lib.izem_chem_tune(ctx, *best[:4]), "izem_chem_tune")
    # The library sorts only batches with more cells than resident lanes
    # (64 x CUs x blocks per CU); on a smaller sample both timings would run
    # the same code and the choice would be noise, so index order is kept.
    # The order must also win by a clear margin: the key pass it adds is
    # about as expensive as a short integration.
    ncu = torch.cuda.get_device_properties(T_K.device).multi_processor_count
    if n > 64 * ncu * best[1]:
        _lib.check(lib.izem_chem_order(ctx, 1), "izem_chem_order")
        rate = rate_of("stiffness order")
        if rate is not None and rate > ORDER_MARGIN * best[4]:
            best = best[:4] + (rate, 1)
    _lib.check(lib.izem_chem_order(ctx, best[5]), "izem_chem_order")
    if use_cache:
        cache[ck] = list(best)
        _store_cache(cache)
    return best
