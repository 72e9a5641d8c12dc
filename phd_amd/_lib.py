"""ctypes binding of the C ABI in include/izem_hip.h (libizem_hip.so).

The product path always goes through this library; when it is missing the
import fails loudly (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libizem_hip.so")

EXPORTS = [
    "izem_create", "izem_destroy", "izem_synchronize", "izem_set_stream", "izem_status",
    "izem_time_state", "izem_get_time_state", "izem_set_time_state", "izem_profile",
    "izem_profile_read",
    "izem_wenorec", "izem_euler_llf_x", "izem_euler_llf_y", "izem_euler_llf_z",
    "izem_viscflux", "izem_force_sumY", "izem_prim_inv_var", "izem_upd_prim_var_woT",
    "izem_comm_cons", "izem_upd_prim_var_ghost", "izem_prim_inv_var_mach", "izem_shock_sensor", "izem_shock_sensor_ducros", "izem_rk3", "izem_rhs",
    "izem_timestep", "izem_chem_setup", "izem_chem_tune", "izem_chem_stats", "izem_ckwyr", "izem_psr_batch", "izem_psr_batch_noalloc", "izem_chem_order", "izem_chem_order_read", "izem_reaction_dvode", "izem_step", "izem_flux_mode",
    "izem_mlp_gemm", "izem_mlp_gemm_splitk", "izem_mlp_colsum", "izem_mlp_adam", "izem_mlp_to_bf16",
    "izem_mlp_mse", "izem_mlp_infer", "izem_surr_features", "izem_surr_apply", "izem_surr_fallback",
    "izem_reac_selec", "izem_reaction_selec", "izem_surr_attach", "izem_reaction_surrogate",
    "izem_surr_counts", "izem_thd_image", "izem_mech_image_ckwrk", "izem_set_bc",
    "izem_set_decomposition", "izem_set_wall_state", "izem_comm_cons_exchange",
    "izem_set_nscbc", "izem_upd_boundaries",
]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ng", C.c_int),
                ("ndim", C.c_int)]


class Opts(C.Structure):
    _fields_ = [("weno_avg", C.c_int), ("max_rel_weight", C.c_double), ("vis", C.c_int),
                ("vis_dt", C.c_int), ("reaction", C.c_int), ("forcesum", C.c_int),
                ("CFL", C.c_double), ("CFLmin", C.c_double), ("Fo", C.c_double),
                ("CPR", C.c_double), ("sqgmr", C.c_double), ("Sc", C.c_double),
                ("Pr", C.c_double), ("ggmopr", C.c_double), ("time_ref", C.c_double),
                ("T_ref", C.c_double), ("rho_ref", C.c_double), ("L_ref", C.c_double),
                ("sensor_type", C.c_int), ("ducros_treduc", C.c_double)]


class SurrogateDesc(C.Structure):
    """izem_surrogate: the device arrays of a trained surrogate."""
    _fields_ = [("Wt", C.c_void_p * 4), ("b", C.c_void_p * 4), ("in_mean", C.c_void_p),
                ("in_istd", C.c_void_p), ("out_mean", C.c_void_p), ("out_std", C.c_void_p),
                ("out_c", C.c_void_p), ("proj", C.c_void_p), ("feat_lim", C.c_float),
                ("tol_T", C.c_double), ("dT_max", C.c_double)]


REAC_DVODE, REAC_SURROGATE = 0, 1


class IzemError(RuntimeError):
    pass


_lib = None


def load(path: str | None = None):
    """Load the HIP library (raises when it has not been built).  The
    environment variable PHD_AMD_LIB names another build of it."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("PHD_AMD_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise IzemError(f"{path} not built: run `python -m phd_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(path)
    _declare(lib)
    _lib = lib
    return lib


class _Lenient:
    """Attribute access to a library build that lacks some entry points (a
    flow-only variant): declarations of absent ones go nowhere."""

    class _Absent:
        def __setattr__(self, name, value):
            pass

    def __init__(self, lib):
        object.__setattr__(self, "_l", lib)

    def __getattr__(self, name):
        try:
            return getattr(self._l, name)
        except AttributeError:
            return _Lenient._Absent()


_variants: dict = {}


def load_variant(nrv: int, ng: int):
    """A flow-only build for nrv species and ng ghost layers (ng = 4: WENO7),
    built by `python -m phd_amd.build` (phd_amd/build.py FLOW_VARIANTS)."""
    from .build import flow_variant_path
    path = flow_variant_path(nrv, ng)
    if path not in _variants:
        if not os.path.exists(path):
            raise IzemError(f"{path} not built: run `python -m phd_amd.build` (hipcc, gfx950)")
        lib = C.CDLL(path)
        _declare(_Lenient(lib))
        _variants[path] = lib
    return _variants[path]


def _declare(lib):
    vp = C.c_void_p
    for name in EXPORTS:
        getattr(lib, name).restype = C.c_int
    lib.izem_time_state.restype = C.c_void_p
    lib.izem_create.argtypes = [C.POINTER(vp), C.POINTER(Grid), vp, C.c_long, C.POINTER(Opts)]
    lib.izem_destroy.argtypes = [vp]
    lib.izem_synchronize.argtypes = [vp]
    lib.izem_status.argtypes = [vp, C.POINTER(C.c_int)]
    lib.izem_time_state.argtypes = [vp]
    lib.izem_set_stream.argtypes = [vp, vp]
    lib.izem_get_time_state.argtypes = [vp, vp]
    lib.izem_set_time_state.argtypes = [vp, vp]
    lib.izem_profile.argtypes = [vp, C.c_int]
    lib.izem_profile_read.argtypes = [vp, vp, vp]
    lib.izem_wenorec.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    for n in ("x", "y", "z"):
        getattr(lib, f"izem_euler_llf_{n}").argtypes = [vp] * 9
    lib.izem_viscflux.argtypes = [vp] * 9
    lib.izem_force_sumY.argtypes = [vp, C.c_int, vp]
    lib.izem_prim_inv_var.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, C.POINTER(C.c_int)]
    lib.izem_upd_prim_var_woT.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    lib.izem_prim_inv_var_mach.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, C.POINTER(C.c_int)]
    lib.izem_comm_cons.argtypes = [vp, vp]
    lib.izem_upd_prim_var_ghost.argtypes = [vp] * 6
    lib.izem_shock_sensor.argtypes = [vp] * 5
    lib.izem_shock_sensor_ducros.argtypes = [vp] * 6
    lib.izem_rk3.argtypes = [vp] * 11
    lib.izem_rhs.argtypes = [vp] * 11
    lib.izem_timestep.argtypes = [vp] * 8
    lib.izem_chem_setup.argtypes = [vp, vp, C.c_long, C.c_double, C.c_double, C.c_int]
    lib.izem_chem_tune.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.izem_chem_stats.argtypes = [vp, vp]
    L, I, D, F = C.c_long, C.c_int, C.c_double, C.c_float
    lib.izem_mlp_gemm.argtypes = [vp, I, I, I, I, I, I, vp, L, vp, L, vp, vp, L, vp, L, vp, L]
    lib.izem_mlp_gemm_splitk.argtypes = [vp, I, I, I, I, I, vp, L, vp, L, I, vp, vp]
    lib.izem_mlp_colsum.argtypes = [vp, I, I, vp, L, I, vp, vp]
    lib.izem_mlp_adam.argtypes = [vp, I, I, vp, vp, vp, vp, F, F, F, F, I, vp, L, vp, L]
    lib.izem_mlp_to_bf16.argtypes = [vp, I, I, vp, vp, L, vp, L]
    lib.izem_mlp_mse.argtypes = [vp, I, I, I, vp, L, vp, L, vp, L, vp]
    lib.izem_mlp_infer.argtypes = [vp, I, vp, vp, vp, vp]
    lib.izem_surr_features.argtypes = [vp, I, vp, vp, vp, D, vp, vp, vp]
    lib.izem_surr_apply.argtypes = [vp, I, vp, vp, vp, L, vp, vp, vp, vp, vp, F, D, D, vp, vp]
    lib.izem_surr_fallback.argtypes = [vp, I, vp, vp, vp, D, vp, vp, vp, vp, vp, vp, vp]
    lib.izem_ckwyr.argtypes = [vp, C.c_int, vp, vp, vp, vp]
    lib.izem_psr_batch.argtypes = [vp, C.c_int, vp, vp, vp, C.c_double, vp]
    lib.izem_psr_batch_noalloc.argtypes = [vp, C.c_int, vp, vp, vp, C.c_double, vp]
    lib.izem_chem_order.argtypes = [vp, C.c_int]
    lib.izem_chem_order_read.argtypes = [vp, C.c_long, C.POINTER(C.c_long), vp, vp]
    lib.izem_reaction_dvode.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp, vp]
    lib.izem_step.argtypes = [vp] + [vp] * 9 + [C.c_int]
    lib.izem_flux_mode.argtypes = [vp, C.c_int]
    lib.izem_reac_selec.argtypes = [vp, C.c_int]
    lib.izem_reaction_selec.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp, vp]
    lib.izem_reaction_surrogate.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp, vp]
    lib.izem_surr_attach.argtypes = [vp, C.POINTER(SurrogateDesc), C.c_long]
    lib.izem_surr_counts.argtypes = [vp, vp]
    lib.izem_thd_image.restype = C.c_long
    lib.izem_thd_image.argtypes = [C.c_int, C.c_int] + [vp] * 10 + [C.c_double] * 7 + [vp,
                                                                                      C.c_long]
    lib.izem_set_bc.argtypes = [vp, vp]
    lib.izem_set_decomposition.argtypes = [vp, vp]
    lib.izem_set_wall_state.argtypes = [vp, vp, C.c_int]
    lib.izem_comm_cons_exchange.argtypes = [vp, vp]
    lib.izem_set_nscbc.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    lib.izem_upd_boundaries.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.izem_mech_image_ckwrk.restype = C.c_long
    lib.izem_mech_image_ckwrk.argtypes = [vp, vp, vp, vp, C.c_long]


def check(rc: int, what: str = ""):
    if rc != 0:
        raise IzemError(f"{what} failed with code {rc}")
