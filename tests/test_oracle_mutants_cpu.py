"""Each pin of tests/pins.py fails on an oracle with a planted mistake.

A scratch copy of oracle/izem_oracle.c gets one edit (a flipped sign, a
swapped index, a wrong coefficient), is built with the oracle's own flags
into a temporary directory, and the checks that should see the mistake are
run against it: each must raise AssertionError.  The oracle itself is not
touched.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from tests import pins

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

P = pins
MUTANTS = [
    ("llf_split_sign", "gp[st] = 0.5 * (gc + eigenmax * wc);",
     "gp[st] = 0.5 * (gc - eigenmax * wc);",
     [P.check_euler_linear_upwind, P.check_euler_riemann_face]),
    ("eigenvalue_per_field", "int lam = (m == 0) ? 0 : (m == 2 ? 2 : 1);",
     "int lam = (m == 0) ? 2 : (m == 2 ? 0 : 1);",
     [P.check_euler_linear_upwind, P.check_euler_riemann_face]),
    ("roe_weight_inverted", "double r = sqrt(V(0, ir) * rho_i[ql]);",
     "double r = 1.0 / sqrt(V(0, ir) * rho_i[ql]);", [P.check_euler_riemann_face]),
    ("energy_flux_no_pressure_work", "fs[st][4] = un * (V(4, is) + P[q]);",
     "fs[st][4] = un * V(4, is);", [P.check_euler_convergence]),
    ("weno_substencil_coefficient", "c[1][0] = 1.0 / 3.0; c[1][1] = 5.0 / 6.0;",
     "c[1][0] = 1.0 / 3.0; c[1][1] = 4.0 / 6.0;",
     [P.check_euler_convergence, P.check_euler_linear_upwind]),
    ("z_rotation", "const int t2 = (dir == 2) ? 2 : 3;\n  double wrk;",
     "const int t2 = (dir == 2) ? 3 : 3;\n  double wrk;", [P.check_euler_axis_permutation]),
    ("y_rotation_tangent", "el[3][0] = -uc[t1 - 1]; el[3][t1] = 1.0;",
     "el[3][0] = -uc[t1 - 1]; el[3][(dir == 1) ? 3 : t1] = 1.0;",
     [P.check_euler_axis_permutation]),
    ("stress_two_thirds", "const double twothirds = 2.0 / 3.0;",
     "const double twothirds = 1.0 / 3.0;", [P.check_visc_stress]),
    ("wilke_index", "double c1 = thd->wilke_c1[m][l];", "double c1 = thd->wilke_c1[l][m];",
     [P.check_visc_stress]),
    ("wilke_mole_fraction", "c4 = c4 + c2 * Xa[l];", "c4 = c4 + c2 * Xa[m];",
     [P.check_visc_stress, P.check_species_split_invariance]),
    ("mathur_harmonic_term", "*ct = Kl * (1.0 / c + w1);", "*ct = Kl * (c + w1);",
     [P.check_visc_heat]),
    ("conductivity_scale", "const double rlam0 = adi->sqgmr * adi->ggmopr;",
     "const double rlam0 = adi->sqgmr / adi->Pr;", [P.check_visc_heat]),
    ("diffusion_scale", "double Kd = 1.0e-4 * thd->p0 / (thd->pinf * thd->Dinf);",
     "double Kd = 1.0e-4 * thd->p0 / (thd->pinf * thd->Dinf) * 1.001;",
     [P.check_visc_binary_fick]),
    ("correction_velocity_2d", "double vdtx = vdcx * Ya, vdty = vdcy * Ya;",
     "double vdtx = 0.0 * vdcx * Ya, vdty = 0.0 * vdcy * Ya;", [P.check_visc_species_sum]),
    ("correction_velocity_3d", "double vdtx = vdcx * Ya, vdty = vdcy * Ya, vdtz = vdcz * Ya;",
     "double vdtx = vdcx * Ya, vdty = 0.0 * vdcy * Ya, vdtz = vdcz * Ya;",
     [P.check_visc_species_sum]),
    ("hc_sum_skips_self", "      if (m == l) continue;\n      double dkj",
     "      if (m == l + 1) continue;\n      double dkj",
     [P.check_visc_binary_fick, P.check_timestep_closed_form]),
    ("rk3_coefficients", "static const double bk3[3] = {0.00, 0.75, 1.00 / 3.00};\n"
     "  static const double ck3[3] = {1.00, 0.25, 2.00 / 3.00};",
     "static const double bk3[3] = {0.00, 0.75, 0.30};\n"
     "  static const double ck3[3] = {1.00, 0.25, 0.70};", [P.check_rk3_third_order]),
    ("rk3_stage2", "static const double bk3[3] = {0.00, 0.75, 1.00 / 3.00};\n"
     "  static const double ck3[3] = {1.00, 0.25, 2.00 / 3.00};",
     "static const double bk3[3] = {0.00, 0.70, 1.00 / 3.00};\n"
     "  static const double ck3[3] = {1.00, 0.30, 2.00 / 3.00};", [P.check_rk3_third_order]),
    ("timestep_species_pick", "if (diffmax < dmall[l * ni + p]) { ldmax = l;",
     "if (diffmax > dmall[l * ni + p]) { ldmax = l;", [P.check_timestep_closed_form]),
    ("timestep_thermal_density", "double b = d2 * v[c] / (ctall[q] / cp[c] + adi->Pr * 0.0);",
     "double b = d2 / (ctall[q] / cp[c] + adi->Pr * 0.0);", [P.check_timestep_closed_form]),
    ("timestep_sound_speed", "double csound = sqrt(gamma * P * rho_i);\n        double dtcx_i",
     "double csound = sqrt(P * rho_i);\n        double dtcx_i", [P.check_timestep_closed_form]),
    ("ducros_vorticity_sign", "          wrk = du[1][0] - du[0][1];\n          vort = wrk * wrk;\n        } else {",
     "          wrk = du[1][0] + du[0][1];\n          vort = wrk * wrk;\n        } else {",
     [P.check_ducros_sensor]),
    ("ducros_divergence_3d", "          div = du[0][0] + du[1][1] + du[2][2];",
     "          div = du[0][0] + du[1][1] + du[2][1];", [P.check_ducros_sensor]),
    ("sensor_density_measure", "double drho = fabs(v[c2] - v[c]) * fmin(ri, ri2);",
     "double drho = fabs(v[c2] - v[c]) * fmax(ri, ri2);", [P.check_sensor_threshold]),
    ("negT_average_along_y", "v[l * nc + c] = 0.5 * (v[l * nc + c - 1] + v[l * nc + c + 1]);",
     "v[l * nc + c] = 0.5 * (v[l * nc + c - NX] + v[l * nc + c + NX]);",
     [P.check_negT_repair_converges, P.check_negT_repair_adjacent_pair]),
    ("negT_no_restart", "            it = 0;\n            correct_T = 1;",
     "            correct_T = 1;", [P.check_negT_repair_converges]),
    ("negT_no_T_reset", "for (int l = 0; l < ipt; l++) T[rec[l]] = 0.5 * (T[rec[l] - 1] + T[rec[l] + 1]);",
     "for (int l = 0; l < ipt; l++) (void)rec[l];", [P.check_negT_repair_runs_to_itmax]),
    ("weno7_indicator_coefficient", "{547.0, -2522.0, 1922.0, -494.0,",
     "{547.0, -2522.0, 1922.0, -495.0,", [P.check_weno7_betas]),
    ("weno7_mirror_indicator", "{267.0, -1642.0, 1602.0, -494.0, 2843.0,",
     "{267.0, -1642.0, 1602.0, -494.0, 2834.0,", [P.check_weno7_betas]),
    ("weno7_linear_coefficient", "319.0 * vp[4]", "318.0 * vp[4]", [P.check_weno7_linear_exact]),
    ("weno7_ideal_weights", "{4.0 / 35.0, 18.0 / 35.0, 12.0 / 35.0, 1.0 / 35.0}",
     "{4.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 1.0 / 35.0}", [P.check_weno7_js_ideal]),
    ("weno7_candidate_row", "{-1.0 / 12.0, 7.0 / 12.0, 7.0 / 12.0, -1.0 / 12.0},",
     "{-1.0 / 12.0, 7.0 / 12.0, 7.0 / 12.0, 1.0 / 12.0},", [P.check_weno7_js_ideal]),
    ("closure_upwind_weight", "{-3.0 / 12.0, -10.0 / 12.0, 18.0 / 12.0, -6.0 / 12.0,",
     "{-3.0 / 12.0, -10.0 / 12.0, 18.0 / 12.0, -5.0 / 12.0,", [P.check_deriv_closures]),
    ("closure_right_sign", "if (p == n - 2) return -(wsum(D4UR, -3, 1, a, c, s) * dinv);",
     "if (p == n - 2) return wsum(D4UR, -3, 1, a, c, s) * dinv;", [P.check_deriv_closures]),
    ("closure_sixth_order_point", "    return wsum(D6, -3, 3, a, c, s) * dinv;           /* p == 3, ng = 4 */",
     "    return wsum(D8, -4, 4, a, c, s) * dinv;           /* p == 3, ng = 4 */",
     [P.check_deriv_closures]),
    ("negT_wmix_after_repair", "  if (iters) *iters = it;\n  free(DF);",
     "  ora_wmix_i(domain, g, thd, v, W_i);\n  if (iters) *iters = it;\n  free(DF);",
     [P.check_negT_repair_converges]),
]


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    return tmp_path_factory.mktemp("oracle_mutants")


def _build(dirpath, name, old, new):
    src = open(os.path.join(ORACLE, "izem_oracle.c")).read()
    assert src.count(old) == 1, f"{name}: the edited text must occur exactly once"
    d = dirpath / name
    d.mkdir()
    shutil.copy(os.path.join(ORACLE, "izem_oracle.h"), d / "izem_oracle.h")
    (d / "izem_oracle.c").write_text(src.replace(old, new))
    so = d / "liboracle.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-std=c11", "-shared", "-o",
                    str(so), str(d / "izem_oracle.c"), "-lm"], check=True)
    return str(so)


@pytest.mark.parametrize("name,old,new,checks", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_fails_on_mutant(mutant_dir, adi, thd, name, old, new, checks):
    from oracle.oracle import Oracle
    o = Oracle(_build(mutant_dir, name, old, new))
    for check in checks:
        with pytest.raises(AssertionError):
            check(o, adi, thd)
