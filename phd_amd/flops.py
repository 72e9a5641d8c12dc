"""Algorithmic operation counts of the hot kernels (for the roofline).

Counts are of the arithmetic the algorithm performs, not of the machine
instructions: +, -, *, / are one FP64 operation each; a transcendental
(exp, log, log10, pow, sqrt) counts as one operation too (its polynomial
evaluation inside the math library is implementation).  Comparisons,
fmin/fmax/fabs and selects are not counted.

Chemistry: the work of a batch follows from the integrator's counters
(izem_chem_stats: BDF steps, RHS evaluations, Jacobians, LU solves) and the
per-call counts of the straight-line kinetics read from the generated
source (phd_amd/csrc/mech_gen.h); both kinetics forms do the same
arithmetic in a different order.
"""
from __future__ import annotations

import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
NEQ = 10  # T + 9 species

_TOKEN = re.compile(r"\d+\.\d*(?:[eE][+-]?\d+)?|\d+(?:[eE][+-]?\d+)|\d+|[A-Za-z_]\w*|"
                    r"\+=|-=|\*=|/=|[-+*/()\[\],;=?:<>!&|{}]")
_TRANS = {"exp", "log", "log10", "pow", "sqrt"}


def count_ops(src: str) -> tuple[int, int]:
    """(arithmetic operations, transcendental calls) in a C function body."""
    arith = trans = 0
    prev = None
    for tok in _TOKEN.findall(src):
        if tok in ("+=", "-=", "*=", "/="):
            arith += 1
        elif tok in ("*", "/"):
            arith += 1
        elif tok in ("+", "-"):
            # binary when it follows an operand
            if prev is not None and (prev[0].isalnum() or prev[0] == "_" or prev in (")", "]")):
                arith += 1
        elif tok in _TRANS:
            trans += 1
        prev = tok
    return arith, trans


def _function_body(src: str, name: str) -> str:
    i = src.index(name + "(")
    i = src.index("{", i)
    depth, j = 0, i
    while True:
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
            if depth == 0:
                return src[i:j + 1]
        j += 1


def kinetics_counts(path: str | None = None) -> dict:
    """Per-call operation counts of the temperature part (gen_kT) and the
    concentration part plus reactor RHS (gen_rhs) of the kinetics."""
    path = path or os.path.join(HERE, "csrc", "mech_gen.h")
    src = open(path).read()
    kt = count_ops(_function_body(src, "gen_kT"))
    rhs = count_ops(_function_body(src, "gen_rhs"))
    return {"kT": kt[0] + kt[1], "conc": rhs[0] + rhs[1], "kT_trans": kt[1],
            "conc_trans": rhs[1]}


# Operations of one Euler face flux (euler_LLF_x/_y/_z, module_weno.f90:
# 537-1044) with the JS reconstruction, counted from the reference's
# expressions per section (the per-stencil presolve is counted per face,
# as the reference evaluates it):
EULER_FACE_SECTIONS = {
    # rho^-1, u_n, P, gamma, c and the |u_n -+ c| of 6 stencil points
    "stencil presolve (6 x 13)": 78,
    # Roe weights, u, Y_k, h_t, e_t averages of i_l and i_r
    "Roe average": 124,
    # W^-1, T, c_p (NASA-7, 9 species), gamma, c, h_k(T_c), psi_k
    "EOS-consistent average state": 408,
    # 1/c, c^2, |u|^2/2, b1, b2, xi
    "eigenvector scalars": 14,
    # rows of el: 20 + 22 + 20 + 14 + 9 x 19
    "left eigenvector rows": 247,
    # w = el.U and g = el.F with the flux at 6 points for 12 dense rows
    # (73 per point) and 2 two-entry rows (9.5 per point)
    "characteristic projections": 12 * 438 + 2 * 57,
    # two JS reconstructions (betas 34, weights 14, candidates 15,
    # combination 5) and g^ = g- + g+ for 14 fields
    "WENO5-JS reconstructions": 14 * 137,
    # F^ = er . g^ with er's nonzero pattern
    "back projection": 97,
}
EULER_FLOP_PER_FACE = sum(EULER_FACE_SECTIONS.values())

# Flux mode 1 (izem_flux_mode): the projections through the structure of el
# (k_euler_faces<DIR, true>) replace the dense rows and their assembly.  Per
# stencil point: the shared scalars S of U and of F (-b1 U_0 + b2 u.m - b2 E
# - b2 sum psi_l U_l: 27 and 32), sum psi_l U_l (18) and its flux (1), and
# u_n U_0 - U_n with its flux (6): 84 minus the 25 the two S share = 59;
# then per point the 14 fields' w and g from them: rows 0 and 2 (6 each),
# row 1 (2), row 5 (4), the 9 species rows (5 each) and the two tangential
# rows (5.5 each) = 74.
EULER_FACE_SECTIONS_MODE1 = dict(EULER_FACE_SECTIONS)
del EULER_FACE_SECTIONS_MODE1["left eigenvector rows"]
EULER_FACE_SECTIONS_MODE1["characteristic projections"] = 6 * 59 + 6 * 74
EULER_FLOP_PER_FACE_MODE = {0: EULER_FLOP_PER_FACE,
                            1: sum(EULER_FACE_SECTIONS_MODE1.values())}


# Algorithmic HBM bytes per cell (3-D, FP64, 9 species) of the memory-bound
# flow kernels: every field a kernel must read or write, once (stencil
# neighbours are the same arrays and count once).
# k_rhs_update (flux difference + viscous divergence + RK stage update +
# force_sumY + Newton probe): face fluxes 3 x 14, stresses and heat flux 9,
# species diffusion fluxes 27, U^0 14, U 14 read and 14 written, T 1.
RHS_UPDATE_FIELDS = 3 * 14 + 9 + 27 + 14 + 14 + 14 + 1
# viscous pass: k_visc_prim (reads rho, rho u_i, rho Y_k, W_i; writes
# u_i, X_k) then k_visc_fused (reads u_i, X_k, rho, rho u_i, rho Y_k, T, W_i,
# h_k; writes the stresses and heat flux 9 and the diffusion fluxes 27)
VISC_PRIM_FIELDS = (1 + 3 + 9 + 1) + (3 + 9)
VISC_FUSED_FIELDS = (3 + 9) + (1 + 3 + 9) + 1 + 1 + 9 + 9 + 27
BYTES_PER_CELL = {"rhs_update": 8 * RHS_UPDATE_FIELDS,
                  "viscous": 8 * (VISC_PRIM_FIELDS + VISC_FUSED_FIELDS)}
# the whole reacting step's compulsory traffic: RHS + SSP-RK3 944 B (stage
# 1 reads U and writes U^(1) with T; stages 2 and 3 also read U^0) plus the
# two Strang chemistry substeps 2 x 168 B
STEP_BYTES_PER_CELL = 944 + 336


def lu_ops(n: int = NEQ) -> int:
    """LU with partial pivoting of an n x n matrix: per column one reciprocal,
    per row below one multiplier and 2 operations per updated entry."""
    return sum(1 + m * (1 + 2 * m) for m in range(n))


def solve_ops(n: int = NEQ) -> int:
    """Forward and back substitution."""
    return n * (n - 1) + n * (n - 1) + n


def chem_flops(stats, counts: dict | None = None) -> float:
    """Operations of a batch of integrations from its counters
    (steps, RHS evaluations, Jacobians, LU solves).  A Jacobian is one
    concentration part per perturbed column and the temperature part of the
    temperature column, plus the difference quotients (3 per entry, 6 per
    column for the increment) and the diagonal: the species columns reuse
    the temperature part of the evaluation of f(y) that precedes every
    Jacobian (counted among the plain evaluations).  One factorisation is
    counted per Jacobian, DVODE's minimum."""
    nst, nfe, nje, nsol = [float(x) for x in stats]
    c = counts or kinetics_counts()
    rhs = c["kT"] + c["conc"]
    jac = c["kT"] + NEQ * c["conc"] + NEQ * (3 * NEQ + 6) + NEQ
    plain_rhs = nfe - NEQ * nje  # evaluations outside the Jacobians
    return plain_rhs * rhs + nje * (jac + lu_ops()) + nsol * solve_ops()


if __name__ == "__main__":
    c = kinetics_counts()
    print(c, "LU", lu_ops(), "solve", solve_ops(), "Euler face", EULER_FLOP_PER_FACE)
