/*
 * izem_oracle.c -- CPU restatement of the IZEM reacting-flow path.
 *
 * TEST INFRASTRUCTURE ONLY (see izem_oracle.h).  Build with
 * -ffp-contract=off: each expression below keeps the reference's
 * left-to-right evaluation order so that a HIP build with the same order
 * agrees bit-for-bit on the flow kernels.
 *
 * Reference: saras13015/Phd, devel-izem/src/ (cited as file:line).
 */
#include "izem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NRV ORA_NRV
#define NIV ORA_NIV
#define NV ORA_NV

static inline long NXg(const ora_grid *g) { return g->nx + 2 * g->ng; }
static inline long NYg(const ora_grid *g) { return g->ny + 2 * g->ng; }
static inline long NZg(const ora_grid *g) { return g->nz + 2 * g->ng; }
static inline long ncell(const ora_grid *g) { return NXg(g) * NYg(g) * NZg(g); }
static inline long cidx(const ora_grid *g, int i, int j, int k) {
  return (long)(i + g->ng) + NXg(g) * ((long)(j + g->ng) + NYg(g) * (long)(k + g->ng));
}

/* domain_select: src/module_parallel.f90:1936-2009.  0 = interior,
 * -1/+1 x ghost slabs, -2/+2 y, -3/+3 z (faces only). */
static inline int is_periodic(const ora_grid *g, int d) {
  return g->bc[2 * d] == ORA_BC_PERIODIC && g->bc[2 * d + 1] == ORA_BC_PERIODIC;
}

static void domain_select(const ora_grid *g, int d, int *i0, int *i1, int *j0, int *j1,
                          int *k0, int *k1) {
  int nx = g->nx, ny = g->ny, nz = g->nz, ng = g->ng;
  *i0 = 0; *i1 = nx - 1; *j0 = 0; *j1 = ny - 1; *k0 = 0; *k1 = nz - 1;
  switch (d) {
    case -1: *i0 = -ng; *i1 = -1; break;
    case 1: *i0 = nx; *i1 = nx + ng - 1; break;
    case -2: *j0 = -ng; *j1 = -1; break;
    case 2: *j0 = ny; *j1 = ny + ng - 1; break;
    case -3: *k0 = -ng; *k1 = -1; break;
    case 3: *k0 = nz; *k1 = nz + ng - 1; break;
    /* the Dirichlet rows: the first interior layer of a face
     * (module_parallel.f90:1946-1957, 1983-1994) */
    case -10: *i1 = 0; break;
    case 10: *i0 = nx - 1; break;
    case -20: *j1 = 0; break;
    case 20: *j0 = ny - 1; break;
    case -30: *k1 = 0; break;
    case 30: *k0 = nz - 1; break;
    default: break;
  }
}

/* ---------------------------------------------------------------- thermo */

/* Wmix_i_scalar: src/module_thermodynamics.f90:739-756 */
double ora_wmix_i_scalar(const ora_thd *thd, const double *Y) {
  double w = 0.0;
  for (int l = 0; l < NRV; l++) w = w + Y[l] * thd->Wc_i[l];
  return w;
}

/* cp_scalar: src/module_thermodynamics.f90:767-862 (linear extrapolation of
 * cp outside [Tpolymin, Tpolymax]) */
double ora_cp_scalar(const ora_thd *thd, double temp, const double *Y) {
  double cp = 0.0;
  for (int is = 0; is < NRV; is++) {
    double tk1 = temp * thd->Tinf, tk2 = tk1 * tk1, tk3 = tk2 * tk1, tk4 = tk3 * tk1;
    const double *hl = thd->HeatCap[is][0], *hh = thd->HeatCap[is][1];
    double fac = Y[is] * thd->gam2 * thd->Wc_i[is];
    if (tk1 < thd->Tpolymin[is]) {
      tk1 = thd->Tpolymin[is]; tk2 = tk1 * tk1; tk3 = tk2 * tk1; tk4 = tk3 * tk1;
      double cp0 = hl[0] + hl[1] * tk1 + hl[2] * tk2 + hl[3] * tk3 + hl[4] * tk4;
      double dcp0 = hl[1] + hl[2] * tk1 * 2.0 + hl[3] * tk2 * 3.0 + hl[4] * tk3 * 4.0;
      cp = cp + fac * (dcp0 * (temp * thd->Tinf - thd->Tpolymin[is]) + cp0);
    } else if (tk1 < thd->Tpolylim[is]) {
      cp = cp + fac * (hl[0] + hl[1] * tk1 + hl[2] * tk2 + hl[3] * tk3 + hl[4] * tk4);
    } else if (tk1 < thd->Tpolymax[is]) {
      cp = cp + fac * (hh[0] + hh[1] * tk1 + hh[2] * tk2 + hh[3] * tk3 + hh[4] * tk4);
    } else {
      tk1 = thd->Tpolymax[is]; tk2 = tk1 * tk1; tk3 = tk2 * tk1; tk4 = tk3 * tk1;
      double cp0 = hh[0] + hh[1] * tk1 + hh[2] * tk2 + hh[3] * tk3 + hh[4] * tk4;
      double dcp0 = hh[1] + hh[2] * tk1 * 2.0 + hh[3] * tk2 * 3.0 + hh[4] * tk3 * 4.0;
      cp = cp + fac * (dcp0 * (temp * thd->Tinf - thd->Tpolymax[is]) + cp0);
    }
  }
  return cp;
}

static inline double h_poly(const double *h, double t1, double t2, double t3, double t4,
                            double t5, double onethird) {
  return h[0] * t1 + h[1] * t2 * 0.50 + h[2] * t3 * onethird + h[3] * t4 * 0.25 +
         h[4] * t5 * 0.20 + h[5];
}
static inline double dh_poly(const double *h, double t1, double t2, double t3, double t4) {
  return h[0] + h[1] * t1 + h[2] * t2 + h[3] * t3 + h[4] * t4;
}

/* ha_scalar: src/module_thermodynamics.f90:873-978 */
void ora_ha_scalar(const ora_thd *thd, double temp, double *ha) {
  double Tinf_i = 1.0 / thd->Tinf, onethird = 1.0 / 3.0;
  for (int is = 0; is < NRV; is++) {
    double t1 = temp * thd->Tinf, t2 = t1 * t1, t3 = t2 * t1, t4 = t3 * t1, t5 = t4 * t1;
    const double *hl = thd->HeatCap[is][0], *hh = thd->HeatCap[is][1];
    double fac = thd->Wc_i[is] * Tinf_i;
    double h = 0.0;
    if (t1 < thd->Tpolymin[is]) {
      t1 = thd->Tpolymin[is]; t2 = t1 * t1; t3 = t2 * t1; t4 = t3 * t1; t5 = t4 * t1;
      double ha0 = h_poly(hl, t1, t2, t3, t4, t5, onethird);
      double dha0 = dh_poly(hl, t1, t2, t3, t4);
      h = h + fac * (dha0 * (temp * thd->Tinf - thd->Tpolymin[is]) + ha0);
    } else if (t1 < thd->Tpolylim[is]) {
      h = h + fac * h_poly(hl, t1, t2, t3, t4, t5, onethird);
    } else if (t1 < thd->Tpolymax[is]) {
      h = h + fac * h_poly(hh, t1, t2, t3, t4, t5, onethird);
    } else {
      t1 = thd->Tpolymax[is]; t2 = t1 * t1; t3 = t2 * t1; t4 = t3 * t1; t5 = t4 * t1;
      double ha0 = h_poly(hh, t1, t2, t3, t4, t5, onethird);
      double dha0 = dh_poly(hh, t1, t2, t3, t4);
      h = h + fac * (dha0 * (temp * thd->Tinf - thd->Tpolymax[is]) + ha0);
    }
    ha[is] = h;
  }
}

/* ------------------------------------------------------------------ WENO */

/* wenopar (ng=3): src/module_weno.f90:114-140, 176-190 */
static const double Z1 = 13.0 / 12.0, Z2 = 1.0 / 4.0, ZA = -2.0, ZB = -4.0, ZC = 3.0, ZD = -1.0;
static const double EPS_WENO = 1.0e-10; /* src/module_weno.f90:31 */

static void cweno_tab(double c[4][3]) {
  c[0][0] = 11.0 / 6.0; c[0][1] = -7.0 / 6.0; c[0][2] = 1.0 / 3.0;   /* r = -1 */
  c[1][0] = 1.0 / 3.0; c[1][1] = 5.0 / 6.0; c[1][2] = -1.0 / 6.0;    /* r = 0 */
  c[2][0] = -1.0 / 6.0; c[2][1] = 5.0 / 6.0; c[2][2] = 1.0 / 3.0;    /* r = 1 */
  c[3][0] = 1.0 / 3.0; c[3][1] = -7.0 / 6.0; c[3][2] = 11.0 / 6.0;   /* r = 2 */
}

static inline void betas(const double *q, int i, double *beta) {
  /* q is 1-based as in the reference (q[1..6]) */
  double a0 = q[i] + ZA * q[i + 1] + q[i + 2], b0 = ZC * q[i] + ZB * q[i + 1] + q[i + 2];
  beta[0] = Z1 * a0 * a0 + Z2 * b0 * b0;
  double a1 = q[i - 1] + ZA * q[i] + q[i + 1], b1 = q[i - 1] + ZD * q[i + 1];
  beta[1] = Z1 * a1 * a1 + Z2 * b1 * b1;
  double a2 = q[i - 2] + ZA * q[i - 1] + q[i], b2 = q[i - 2] + ZB * q[i - 1] + ZC * q[i];
  beta[2] = Z1 * a2 * a2 + Z2 * b2 * b2;
}

/* wenorec (JS weights): src/module_weno.f90:286-527, ng = 3 branch.
 * vp, vm hold the 6 stencil values (0-based here, reference is 1-based). */
void ora_wenorec(const double *vp0, const double *vm0, double *vminus, double *vplus) {
  double c[4][3];
  cweno_tab(c);
  const double d[3] = {3.0 / 10.0, 3.0 / 5.0, 1.0 / 10.0};
  const double dt[3] = {d[2], d[1], d[0]};
  const double *vp = vp0 - 1, *vm = vm0 - 1; /* 1-based views */
  double vpr[3], beta[3], alfa[3], omega[3], sum;
  int i = 3;
  for (int r = 0; r < 3; r++) {
    vpr[r] = 0.0;
    for (int j = 0; j < 3; j++) vpr[r] = vpr[r] + c[r + 1][j] * vp[i - r + j];
  }
  betas(vp, i, beta);
  sum = 0.0;
  for (int r = 0; r < 3; r++) {
    alfa[r] = d[r] / ((EPS_WENO + beta[r]) * (EPS_WENO + beta[r]));
    sum = sum + alfa[r];
  }
  sum = 1.0 / sum;
  for (int r = 0; r < 3; r++) omega[r] = alfa[r] * sum;
  double vmn = 0.0;
  for (int r = 0; r < 3; r++) vmn = vmn + omega[r] * vpr[r];
  *vminus = vmn;

  i = 4;
  double vmr[3];
  for (int r = 0; r < 3; r++) {
    vmr[r] = 0.0;
    for (int j = 0; j < 3; j++) vmr[r] = vmr[r] + c[r][j] * vm[i - r + j];
  }
  betas(vm, i, beta);
  sum = 0.0;
  for (int r = 0; r < 3; r++) {
    alfa[r] = dt[r] / ((EPS_WENO + beta[r]) * (EPS_WENO + beta[r]));
    sum = sum + alfa[r];
  }
  sum = 1.0 / sum;
  for (int r = 0; r < 3; r++) omega[r] = alfa[r] * sum;
  double vpl = 0.0;
  for (int r = 0; r < 3; r++) vpl = vpl + omega[r] * vmr[r];
  *vplus = vpl;
}

/* wenorec_nw, ng = 3: src/module_weno.f90:258-266.  Kept exactly as the
 * reference writes it (its coefficient order reconstructs at i-1/2, not
 * i+1/2; the JS path above is the one parity runs use). */
void ora_wenorec_nw(const double *vp0, const double *vm0, double *vminus, double *vplus) {
  const double *vp = vp0 - 1, *vm = vm0 - 1;
  double denom = 1.0 / 60.0;
  *vminus = (-3.0 * vp[1] + 27.0 * vp[2] + 47.0 * vp[3] - 13.0 * vp[4] + 2.0 * vp[5]) * denom;
  *vplus = (-3.0 * vm[6] + 27.0 * vm[5] + 47.0 * vm[4] - 13.0 * vm[3] + 2.0 * vm[2]) * denom;
}

/* WENO7 (ng = 4): candidate-stencil coefficients and ideal weights
 * (4, 18, 12, 1)/35 with their mirror for the '+' state
 * (src/module_weno.f90:71-111), smoothness indicators (:154-172, 330-353,
 * 451-471) and the JS weights (:369-413, 476-519).  The indicators are
 * Balsara and Shu's r = 4 forms (J. Comput. Phys. 160 (2000) 405-452): for
 * the four values a, b, c, d of a candidate stencil
 *   beta = a (k0 a + k1 b + k2 c + k3 d) + b (k4 b + k5 c + k6 d)
 *        + c (k7 c + k8 d) + d (k9 d),
 * with one coefficient row per stencil position (the last two rows mirror
 * the first two). */
static const double CW7[5][4] = {
    {25.0 / 12.0, -23.0 / 12.0, 13.0 / 12.0, -1.0 / 4.0},   /* r = -1 */
    {1.0 / 4.0, 13.0 / 12.0, -5.0 / 12.0, 1.0 / 12.0},      /* r = 0 */
    {-1.0 / 12.0, 7.0 / 12.0, 7.0 / 12.0, -1.0 / 12.0},     /* r = 1 */
    {1.0 / 12.0, -5.0 / 12.0, 13.0 / 12.0, 1.0 / 4.0},      /* r = 2 */
    {-1.0 / 4.0, 13.0 / 12.0, -23.0 / 12.0, 25.0 / 12.0}};  /* r = 3 */
static const double B7[4][10] = {
    {2107.0, -9402.0, 7042.0, -1854.0, 11003.0, -17246.0, 4642.0, 7043.0, -3882.0, 547.0},
    {547.0, -2522.0, 1922.0, -494.0, 3443.0, -5966.0, 1602.0, 2843.0, -1642.0, 267.0},
    {267.0, -1642.0, 1602.0, -494.0, 2843.0, -5966.0, 1922.0, 3443.0, -2522.0, 547.0},
    {547.0, -3882.0, 4642.0, -1854.0, 7043.0, -17246.0, 7042.0, 11003.0, -9402.0, 2107.0}};

static inline double beta7(const double *k, double a, double b, double c, double d) {
  return a * (k[0] * a + k[1] * b + k[2] * c + k[3] * d) + b * (k[4] * b + k[5] * c + k[6] * d) +
         c * (k[7] * c + k[8] * d) + d * (k[9] * d);
}

/* smoothness indicator of WENO7 candidate r on its four values q[0..3] */
double ora_weno7_beta(int r, const double *q) { return beta7(B7[r], q[0], q[1], q[2], q[3]); }

/* one side of the WENO7 reconstruction: q is 1-based, i the centre,
 * row0 the coefficient row of candidate r = 0 (1 for '-', 0 for '+') */
static double weno7_side(const double *q, int i, int row0, const double *dw) {
  double vr[4], beta[4], alfa[4], sum = 0.0, out = 0.0;
  for (int r = 0; r < 4; r++) {
    vr[r] = 0.0;
    for (int j = 0; j < 4; j++) vr[r] = vr[r] + CW7[r + row0][j] * q[i - r + j];
  }
  for (int r = 0; r < 4; r++)
    beta[r] = beta7(B7[r], q[i - r], q[i - r + 1], q[i - r + 2], q[i - r + 3]);
  for (int r = 0; r < 4; r++) {
    alfa[r] = dw[r] / ((EPS_WENO + beta[r]) * (EPS_WENO + beta[r]));
    sum = sum + alfa[r];
  }
  sum = 1.0 / sum;
  for (int r = 0; r < 4; r++) out = out + alfa[r] * sum * vr[r];
  return out;
}

/* wenorec, ng = 4 branch: vp, vm hold the 8 stencil values (0-based) */
void ora_wenorec7(const double *vp0, const double *vm0, double *vminus, double *vplus) {
  static const double d[4] = {4.0 / 35.0, 18.0 / 35.0, 12.0 / 35.0, 1.0 / 35.0};
  static const double dt[4] = {1.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 4.0 / 35.0};
  *vminus = weno7_side(vp0 - 1, 4, 1, d);
  *vplus = weno7_side(vm0 - 1, 5, 0, dt);
}

/* wenorec_nw, ng = 4: src/module_weno.f90:246-256 (the 7th-order linear
 * upwind face value; the ng = 4 formula reconstructs at i+1/2) */
void ora_wenorec7_nw(const double *vp0, const double *vm0, double *vminus, double *vplus) {
  const double *vp = vp0 - 1, *vm = vm0 - 1;
  double denom = 1.0 / 420.0;
  *vminus = (-3.0 * vp[1] + 25.0 * vp[2] - 101.0 * vp[3] + 319.0 * vp[4] + 214.0 * vp[5] -
             38.0 * vp[6] + 4.0 * vp[7]) * denom;
  *vplus = (-3.0 * vm[8] + 25.0 * vm[7] - 101.0 * vm[6] + 319.0 * vm[5] + 214.0 * vm[4] -
            38.0 * vm[3] + 4.0 * vm[2]) * denom;
}

/* --------------------------------------------------------- Euler fluxes */

/* Left and right eigenvector matrices at an averaged state
 * (module_weno.f90:763-951 and the y/z rotations).  Shared by the flux
 * routine and the eigenvector tests. */
void ora_eigvecs(int dir, const ora_thd *thd, const double *uc, const double *Yc,
                 const double *psi_c, double ht_c, double cs_c, double gam_c,
                 double el[ORA_NV][ORA_NV], double er[ORA_NV][ORA_NV]) {
  const int cn = 1 + dir;
  const int t1 = (dir == 0) ? 2 : 1;
  const int t2 = (dir == 2) ? 2 : 3;
  double wrk;
  (void)thd;
  double cs_c_i = 1.0 / cs_c, cs_c_2 = cs_c * cs_c;
  double ux_c = uc[0], vy_c = uc[1], wz_c = uc[2];
  double q = 0.5 * (ux_c * ux_c + vy_c * vy_c + wz_c * wz_c);
  double b2 = (gam_c - 1.0) / cs_c_2;
  double b1 = b2 * q;
  double xi = b1 - b2 * ht_c;
  double un_c = uc[cn - 1];
  /* left eigenvectors: src/module_weno.f90:775-874 (and y/z rotations) */
  memset(el, 0, sizeof(double) * NV * NV);
  memset(er, 0, sizeof(double) * NV * NV);
  el[0][0] = 0.5 * (b1 + un_c * cs_c_i);
  el[2][0] = 0.5 * (b1 - un_c * cs_c_i);
  for (int c = 1; c <= 3; c++) {
    if (c == cn) {
      el[0][c] = -0.5 * (b2 * uc[c - 1] + cs_c_i);
      el[2][c] = -0.5 * (b2 * uc[c - 1] - cs_c_i);
    } else {
      el[0][c] = -0.5 * (b2 * uc[c - 1]);
      el[2][c] = -0.5 * (b2 * uc[c - 1]);
    }
  }
  el[0][4] = 0.5 * b2;
  el[2][4] = 0.5 * b2;
  el[1][0] = 1.0 - b1;
  el[1][1] = b2 * ux_c; el[1][2] = b2 * vy_c; el[1][3] = b2 * wz_c;
  el[1][4] = -b2;
  el[3][0] = -uc[t1 - 1]; el[3][t1] = 1.0;
  el[4][0] = -uc[t2 - 1]; el[4][t2] = 1.0;
  wrk = 0.5 * b2;
  for (int l = NIV; l < NV; l++) el[0][l] = wrk * psi_c[l - NIV];
  for (int l = NIV; l < NV; l++) el[1][l] = -el[0][l] - el[0][l];
  for (int l = NIV; l < NV; l++) el[2][l] = el[0][l];
  el[5][0] = (1 + xi) * q;
  el[5][1] = -(1 + xi) * ux_c;
  el[5][2] = -(1 + xi) * vy_c;
  el[5][3] = -(1 + xi) * wz_c;
  el[5][4] = (1 + xi);
  for (int l = NIV; l < NV; l++) el[5][l] = xi * psi_c[l - NIV];
  for (int l = 6; l < NV; l++) {
    double y = Yc[l - 6];
    el[l][0] = -b1 * y;
    el[l][1] = b2 * ux_c * y;
    el[l][2] = b2 * vy_c * y;
    el[l][3] = b2 * wz_c * y;
    el[l][4] = -b2 * y;
  }
  for (int s1 = 0; s1 < NRV - 1; s1++)
    for (int s2 = 0; s2 < NRV; s2++) el[s1 + 6][s2 + 5] = -b2 * Yc[s1] * psi_c[s2];
  for (int l = 0; l < NRV - 1; l++) el[l + 6][l + 5] = el[l + 6][l + 5] + 1.0;
  /* right eigenvectors: src/module_weno.f90:881-951 */
  er[0][0] = 1.0; er[0][1] = 1.0; er[0][2] = 1.0;
  for (int c = 1; c <= 3; c++) {
    if (c == cn) {
      er[c][0] = uc[c - 1] - cs_c;
      er[c][1] = uc[c - 1];
      er[c][2] = uc[c - 1] + cs_c;
    } else {
      er[c][0] = uc[c - 1]; er[c][1] = uc[c - 1]; er[c][2] = uc[c - 1];
    }
  }
  er[t1][3] = 1.0;
  er[t2][4] = 1.0;
  er[4][0] = ht_c - un_c * cs_c;
  er[4][1] = q;
  er[4][2] = ht_c + un_c * cs_c;
  er[4][3] = uc[t1 - 1];
  er[4][4] = uc[t2 - 1];
  for (int l = NIV; l < NV; l++) { er[l][0] = Yc[l - NIV]; er[l][2] = Yc[l - NIV]; }
  for (int l = NIV; l < NV; l++) er[l - 1][l] = 1.0;
  er[NV - 1][5] = -1.0 / psi_c[NRV - 1];
  wrk = -1.0 / psi_c[NRV - 1];
  for (int l = 0; l < NRV - 1; l++) er[NV - 1][l + NIV + 1] = wrk * psi_c[l];
}

/* ranks along a direction own (r n)/np .. ((r+1) n)/np - 1
 * (src/module_parallel.f90:547-548); ghost positions belong to the nearest
 * end rank */
static inline int rank_lo(int r, int n, int np) { return (int)(((long)r * n) / np); }
static inline int rank_hi(int r, int n, int np) { return (int)(((long)(r + 1) * n) / np) - 1; }
static inline int rank_of(int p, int n, int np) {
  int r = 0;
  while (r + 1 < np && rank_lo(r + 1, n, np) <= p) r++;
  return r;
}

/* any sensor flag among line positions p0..p1 that lie in [lo, hi] */
static inline int flags_in(const uint8_t *sensor, long base, long stride, int p0, int p1, int lo,
                           int hi) {
  for (int p = p0 > lo ? p0 : lo; p <= p1 && p <= hi; p++)
    if (sensor[base + (long)p * stride]) return 1;
  return 0;
}

static inline double weno_face(int ng, int js, const double *gp, const double *gm) {
  double gl, gr;
  if (ng == 4) {
    if (js) ora_wenorec7(gp, gm, &gl, &gr);
    else ora_wenorec7_nw(gp, gm, &gl, &gr);
  } else {
    if (js) ora_wenorec(gp, gm, &gl, &gr);
    else ora_wenorec_nw(gp, gm, &gl, &gr);
  }
  return gl + gr;
}

/* euler_LLF_x / _y / _z: src/module_weno.f90:537-1044, 1055-1584, 1595-2118.
 * dir 0/1/2.  x assigns fl, y and z accumulate (src/module_weno.f90:1002-1018
 * vs the += form of the y/z routines). */
int ora_euler_llf(int dir, const ora_grid *g, const ora_thd *thd, const double *d_i,
                  const double *T, const double *W_i, const double *cp, const double *ha,
                  const uint8_t *sensor, int weno_avg, const double *v, double *fl) {
  const int ng = g->ng;
  const long nc = ncell(g);
  int n, na, nb;
  long stride;
  if (dir == 0) { n = g->nx; na = g->ny; nb = g->nz; stride = 1; }
  else if (dir == 1) { n = g->ny; na = g->nx; nb = g->nz; stride = NXg(g); }
  else { n = g->nz; na = g->nx; nb = g->ny; stride = NXg(g) * NYg(g); }
  const int cn = 1 + dir;                       /* normal momentum column */
  const int t1 = (dir == 0) ? 2 : 1;            /* first tangential column */
  const int t2 = (dir == 2) ? 2 : 3;            /* second tangential column */
  const int L = n + 2 * ng;
  const int S = 2 * ng; /* stencil of a face: 6 (WENO5) or 8 (WENO7) points */
  if (ng != 3 && ng != 4) return -2;
  double *rho_i = malloc(sizeof(double) * L), *P = malloc(sizeof(double) * L),
         *gam = malloc(sizeof(double) * L), *cs = malloc(sizeof(double) * L),
         *vel = malloc(sizeof(double) * 3 * L), *ht = malloc(sizeof(double) * L),
         *et = malloc(sizeof(double) * L), *Ya = malloc(sizeof(double) * NRV * L),
         *fhat = malloc(sizeof(double) * NV * (n + 1)),
         *fhat_alt = malloc(sizeof(double) * NV * (n + 1));
  const int np_dir = g->nproc[dir] > 1 ? g->nproc[dir] : 1;
  int status = 0;
  for (int b = 0; b < nb; b++)
    for (int a = 0; a < na; a++) {
      long base;
      if (dir == 0) base = cidx(g, 0, a, b);
      else if (dir == 1) base = cidx(g, a, 0, b);
      else base = cidx(g, a, b, 0);
#define C(p) (base + (long)(p) * stride)
#define V(l, p) v[(long)(l) * nc + C(p)]
      /* presolve 1/2: src/module_weno.f90:634-650 */
      for (int p = -ng; p < n + ng; p++) {
        int q = p + ng;
        rho_i[q] = 1.0 / V(0, p);
        vel[cn - 1 + 3 * q] = V(cn, p) * rho_i[q];
        P[q] = V(0, p) * T[C(p)] * W_i[C(p)];
        if (P[q] <= 0.0 && status == 0) status = -1;
        gam[q] = thd->gam2 * W_i[C(p)] / cp[C(p)];
        gam[q] = 1.0 / (1.0 - gam[q]);
        cs[q] = sqrt(gam[q] * P[q] * rho_i[q]);
      }
      /* presolve 2/2: src/module_weno.f90:654-667 */
      for (int p = -1; p <= n; p++) {
        int q = p + ng;
        vel[t1 - 1 + 3 * q] = V(t1, p) * rho_i[q];
        vel[t2 - 1 + 3 * q] = V(t2, p) * rho_i[q];
        for (int l = 0; l < NRV; l++) Ya[l + NRV * q] = V(NIV + l, p) * rho_i[q];
        double h = 0.0;
        for (int l = 0; l < NRV; l++) h = h + ha[(long)l * nc + C(p)] * Ya[l + NRV * q];
        double ux = vel[3 * q], vy = vel[3 * q + 1], wz = vel[3 * q + 2];
        ht[q] = h + 0.5 * (ux * ux + vy * vy + wz * wz);
        et[q] = V(4, p) * rho_i[q];
      }
      for (int il = -1; il < n; il++) {
        int ir = il + 1, ql = il + ng, qr = ir + ng;
        double fs[8][NV], Ls[8][3];
        for (int st = 0; st < S; st++) {
          int is = il + st + 1 - ng, q = is + ng;
          double un = vel[cn - 1 + 3 * q];
          fs[st][0] = V(cn, is);
          for (int c = 1; c <= 3; c++) {
            if (c == cn) fs[st][c] = V(c, is) * un + P[q];
            else fs[st][c] = V(c, is) * un;
          }
          fs[st][4] = un * (V(4, is) + P[q]);
          for (int l = NIV; l < NV; l++) fs[st][l] = V(l, is) * un;
          Ls[st][0] = fabs(un - cs[q]);
          Ls[st][1] = fabs(un);
          Ls[st][2] = fabs(un + cs[q]);
        }
        /* Roe average: src/module_weno.f90:726-735 */
        double r = sqrt(V(0, ir) * rho_i[ql]);
        double r1 = 1.0 / (r + 1.0);
        double uc[3];
        for (int c = 0; c < 3; c++) uc[c] = (r * vel[c + 3 * qr] + vel[c + 3 * ql]) * r1;
        double Yc[NRV];
        for (int l = 0; l < NRV; l++) Yc[l] = (r * (Ya[l + NRV * qr]) + Ya[l + NRV * ql]) * r1;
        double ht_c = (r * ht[qr] + ht[ql]) * r1;
        double et_c = (r * et[qr] + et[ql]) * r1;
        /* EOS-compatible average state: src/module_weno.f90:750-761 */
        double W_i_c = ora_wmix_i_scalar(thd, Yc);
        double T_c = (ht_c - et_c) / W_i_c;
        double cp_c = ora_cp_scalar(thd, T_c, Yc);
        double gam_c = thd->gam2 * W_i_c / cp_c;
        gam_c = 1.0 / (1.0 - gam_c);
        double cs_c = sqrt(gam_c * T_c * W_i_c);
        double ha_c[NRV], psi_c[NRV];
        ora_ha_scalar(thd, T_c, ha_c);
        double wrk = (gam_c / (gam_c - 1.0)) * T_c;
        for (int l = 0; l < NRV; l++) {
          psi_c[l] = -(r * ha_c[l] + ha_c[l]) * r1;
          psi_c[l] = psi_c[l] + wrk * thd->Wc_i[l];
        }
        double el[NV][NV], er[NV][NV];
        ora_eigvecs(dir, thd, uc, Yc, psi_c, ht_c, cs_c, gam_c, el, er);
        /* characteristic LLF + WENO per field: src/module_weno.f90:955-989.
         * The flags a rank sees are its own interior cells' (the sensor is
         * neither set on ghosts nor exchanged): the face takes the view of
         * the rank of the cell to its right; a face between two ranks is
         * also formed with the left rank's view (shk_alt) for that rank's
         * last cell. */
        const int rr = rank_of(il + 1, n, np_dir), rl = rank_of(il, n, np_dir);
        const int shk = flags_in(sensor, base, stride, il - ng + 1, il + ng,
                                 rank_lo(rr, n, np_dir), rank_hi(rr, n, np_dir));
        const int split = il >= 0 && il + 1 < n && rl != rr;
        const int shk_alt = split ? flags_in(sensor, base, stride, il - ng + 1, il + ng,
                                             rank_lo(rl, n, np_dir), rank_hi(rl, n, np_dir))
                                  : shk;
        double ghat[NV], ghat_alt[NV];
        for (int m = 0; m < NV; m++) {
          int lam = (m == 0) ? 0 : (m == 2 ? 2 : 1);
          double eigenmax = -1.0;
          for (int st = 0; st < S; st++) eigenmax = fmax(Ls[st][lam], eigenmax);
          double gp[8], gm[8];
          for (int st = 0; st < S; st++) {
            int is = il + st + 1 - ng;
            double wc = 0.0, gc = 0.0;
            for (int mm = 0; mm < NV; mm++) {
              wc = wc + el[m][mm] * V(mm, is);
              gc = gc + el[m][mm] * fs[st][mm];
            }
            gp[st] = 0.5 * (gc + eigenmax * wc);
            gm[st] = gc - gp[st];
          }
          ghat[m] = weno_face(ng, shk && weno_avg, gp, gm);
          ghat_alt[m] = (shk_alt == shk) ? ghat[m] : weno_face(ng, shk_alt && weno_avg, gp, gm);
        }
        for (int m = 0; m < NV; m++) {
          double s = 0.0, sa = 0.0;
          for (int mm = 0; mm < NV; mm++) s = s + er[m][mm] * ghat[mm];
          for (int mm = 0; mm < NV; mm++) sa = sa + er[m][mm] * ghat_alt[mm];
          fhat[m * (n + 1) + (il + 1)] = s;
          fhat_alt[m * (n + 1) + (il + 1)] = sa;
        }
      }
      /* flux difference: src/module_weno.f90:998-1020 (a rank's last cell
       * takes its right face from its own view) */
      for (int m = 0; m < NV; m++)
        for (int ic = 0; ic < n; ic++) {
          const double *fr = (ic + 1 < n && rank_of(ic + 1, n, np_dir) != rank_of(ic, n, np_dir))
                                 ? fhat_alt : fhat;
          double df = (fr[m * (n + 1) + ic + 1] - fhat[m * (n + 1) + ic]) * d_i[ic + ng];
          long cc = (long)m * nc + C(ic);
          if (dir == 0) fl[cc] = df * 1.0;
          else fl[cc] = fl[cc] + df * 1.0;
        }
#undef V
#undef C
    }
  free(rho_i); free(P); free(gam); free(cs); free(vel); free(ht); free(et); free(Ya);
  free(fhat); free(fhat_alt);
  return status;
}

/* ------------------------------------------------- primitive variables */

/* Wmix_i: src/module_thermodynamics.f90:700-728 */
void ora_wmix_i(int domain, const ora_grid *g, const ora_thd *thd, const double *v, double *W_i) {
  int i0, i1, j0, j1, k0, k1;
  long nc = ncell(g);
  domain_select(g, domain, &i0, &i1, &j0, &j1, &k0, &k1);
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        long c = cidx(g, i, j, k);
        double w = 0.0;
        for (int l = 0; l < NRV; l++) w = w + (v[(NIV + l) * nc + c] / v[c]) * thd->Wc_i[l];
        W_i[c] = w;
      }
}

/* force_sumY (npv = 0 branch): src/module_thermodynamics.f90:1709-1760 */
void ora_force_sumY(int domain, const ora_grid *g, double *v) {
  int i0, i1, j0, j1, k0, k1;
  long nc = ncell(g);
  domain_select(g, domain, &i0, &i1, &j0, &j1, &k0, &k1);
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        long c = cidx(g, i, j, k);
        double rho = v[c], rho_i = 1.0 / rho, sum = 0.0, Y;
        for (int l = 0; l < NRV - 1; l++) {
          Y = v[(NIV + l) * nc + c] * rho_i;
          Y = fmin(fmax(Y, 0.0), 1.0);
          v[(NIV + l) * nc + c] = rho * Y;
          sum = sum + Y;
        }
        Y = 1.0 - sum;
        Y = fmin(fmax(Y, 0.0), 1.0);
        v[(NIV + NRV - 1) * nc + c] = rho * Y;
      }
}

/* per-cell body of newton: src/module_thermodynamics.f90:271-508.  Writes
 * cp and ha at the current T and returns DF = F / (sum Y Wc_i - cp/gam2). */
static double newton_cell(const ora_thd *thd, const double *vc, double T, double *cp_out,
                          double *ha_out) {
  double Tinf_i = 1.0 / thd->Tinf, onethird = 1.0 / 3.0, gam2_i = 1.0 / thd->gam2;
  double rho_i = 1.0 / vc[0];
  double F = -vc[4] * rho_i + 0.5 * (vc[1] * vc[1] + vc[2] * vc[2] + vc[3] * vc[3]) *
                                  (rho_i * rho_i);
  double Ya[NRV];
  for (int is = 0; is < NRV; is++) Ya[is] = vc[NIV + is] * rho_i;
  double cp = 0.0;
  for (int is = 0; is < NRV; is++) {
    double t1 = T * thd->Tinf, t2 = t1 * t1, t3 = t2 * t1, t4 = t3 * t1, t5 = t4 * t1;
    double wrk1 = -T;
    const double *hl = thd->HeatCap[is][0], *hh = thd->HeatCap[is][1];
    double fh = thd->Wc_i[is] * Tinf_i;
    double fc = Ya[is] * thd->gam2 * thd->Wc_i[is];
    double h = 0.0;
    if (t1 < thd->Tpolymin[is]) {
      t1 = thd->Tpolymin[is]; t2 = t1 * t1; t3 = t2 * t1; t4 = t3 * t1; t5 = t4 * t1;
      double ha0 = h_poly(hl, t1, t2, t3, t4, t5, onethird);
      double dha0 = dh_poly(hl, t1, t2, t3, t4);
      h = h + fh * (dha0 * (T * thd->Tinf - thd->Tpolymin[is]) + ha0);
      double cp0 = dha0;
      double dcp0 = hl[1] + hl[2] * t1 * 2.0 + hl[3] * t2 * 3.0 + hl[4] * t3 * 4.0;
      cp = cp + fc * (dcp0 * (T * thd->Tinf - thd->Tpolymin[is]) + cp0);
    } else if (t1 < thd->Tpolylim[is]) {
      h = h + fh * h_poly(hl, t1, t2, t3, t4, t5, onethird);
      cp = cp + fc * dh_poly(hl, t1, t2, t3, t4);
    } else if (t1 < thd->Tpolymax[is]) {
      h = h + fh * h_poly(hh, t1, t2, t3, t4, t5, onethird);
      cp = cp + fc * dh_poly(hh, t1, t2, t3, t4);
    } else {
      t1 = thd->Tpolymax[is]; t2 = t1 * t1; t3 = t2 * t1; t4 = t3 * t1; t5 = t4 * t1;
      double ha0 = h_poly(hh, t1, t2, t3, t4, t5, onethird);
      double dha0 = dh_poly(hh, t1, t2, t3, t4);
      h = h + fh * (dha0 * (T * thd->Tinf - thd->Tpolymax[is]) + ha0);
      double cp0 = dha0;
      double dcp0 = hh[1] + hh[2] * t1 * 2.0 + hh[3] * t2 * 3.0 + hh[4] * t3 * 4.0;
      cp = cp + fc * (dcp0 * (T * thd->Tinf - thd->Tpolymax[is]) + cp0);
    }
    ha_out[is] = h;
    F = F + Ya[is] * (h + wrk1 * thd->Wc_i[is]);
  }
  double wrk = 0.0;
  for (int is = 0; is < NRV; is++) wrk = wrk + Ya[is] * thd->Wc_i[is];
  *cp_out = cp;
  return F / (wrk - cp * gam2_i);
}

/* prim_inv_var: src/module_thermodynamics.f90:79-186.  All cells of the
 * domain iterate together until sqrt(max DF^2) < 1e-5 or 200 iterations;
 * cp and ha keep the values of the last evaluation (pre-update T).  W_i is
 * computed once, before the iterations (:102).
 *
 * Negative-temperature repair (:120-157): each iteration first evaluates
 * the increments of every cell of the domain (newton, :116), then sweeps
 * the cells in storage order adding them; a cell whose T turns negative is
 * recorded, its conserved variables are replaced by the mean of its x
 * neighbours (so a negative neighbour met earlier in the same sweep enters
 * with its replaced values), and the iteration count restarts.  After the
 * exit test, every cell recorded so far gets T = mean of its x neighbours'
 * T, at every later iteration.  The report file of :141-153 is not
 * written.  Returns 0, 1 (a repair was made: the reference's warning),
 * -2 (more than 40000 repaired cells, where the reference aborts, or a
 * repaired cell whose x neighbour lies outside the allocated array, where
 * it reads out of bounds) or -3 (NaN). */
int ora_prim_inv_var(int domain, const ora_grid *g, const ora_thd *thd, double *v, double *W_i,
                     double *T, double *cp, double *ha, int *iters) {
  const int itmax = 200, nptnegT_max = 40000;
  const double eps_temp = 1.0e-5;
  int i0, i1, j0, j1, k0, k1;
  long nc = ncell(g);
  const int NX = g->nx + 2 * g->ng;
  domain_select(g, domain, &i0, &i1, &j0, &j1, &k0, &k1);
  ora_wmix_i(domain, g, thd, v, W_i);
  const long nd = (long)(i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1);
  double *DF = malloc(sizeof(double) * (nd > 0 ? nd : 1));
  long *rec = NULL;
  int it = 0, status = 0, ipt = 0, correct_T = 0;
  double error = 0.0;
  for (;;) {
    it = it + 1;
    if (it > itmax) break;
    long q = 0;
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++, q++) {
          long c = cidx(g, i, j, k);
          double vc[NV], hc[NRV], cpc;
          for (int l = 0; l < NV; l++) vc[l] = v[l * nc + c];
          DF[q] = newton_cell(thd, vc, T[c], &cpc, hc);
          cp[c] = cpc;
          for (int l = 0; l < NRV; l++) ha[l * nc + c] = hc[l];
        }
    double maxdf = -INFINITY;
    int nan_seen = 0;
    q = 0;
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++, q++) {
          long c = cidx(g, i, j, k);
          T[c] = T[c] + DF[q];
          double d2 = DF[q] * DF[q];
          if (T[c] < 0.0) {
            ipt = ipt + 1;
            const int ia = i + g->ng;  /* array index along x */
            if (ipt > nptnegT_max || ia < 1 || ia > NX - 2) {
              status = -2;
              goto done;
            }
            if (!rec) rec = malloc(sizeof(long) * nptnegT_max);
            rec[ipt - 1] = c;
            it = 0;
            correct_T = 1;
            status = 1;
            for (int l = 0; l < NV; l++)
              v[l * nc + c] = 0.5 * (v[l * nc + c - 1] + v[l * nc + c + 1]);
          }
          if (d2 != d2) nan_seen = 1;
          maxdf = fmax(maxdf, d2);
        }
    error = nan_seen ? NAN : sqrt(maxdf);
    if (error < eps_temp) break;
    if (correct_T)
      for (int l = 0; l < ipt; l++) T[rec[l]] = 0.5 * (T[rec[l] - 1] + T[rec[l] + 1]);
  }
  if (error != error) status = -3;
done:
  if (iters) *iters = it;
  free(DF);
  free(rec);
  return status;
}

/* prim_inv_var_err_mach: src/module_thermodynamics.f90:201-260, the
 * inversion of initialisation and restart: the global loop of prim_inv_var
 * to 1e-10 and at most 100 iterations, without the repair branch.  Returns
 * 0, or the warnings it prints as bits: 2 a negative temperature left,
 * 8 no convergence. */
int ora_prim_inv_var_mach(int domain, const ora_grid *g, const ora_thd *thd, double *v,
                          double *W_i, double *T, double *cp, double *ha, int *iters) {
  const int itmax_mach = 100;
  const double eps_temp_mach = 1.0e-10;
  int i0, i1, j0, j1, k0, k1;
  long nc = ncell(g);
  domain_select(g, domain, &i0, &i1, &j0, &j1, &k0, &k1);
  ora_wmix_i(domain, g, thd, v, W_i);
  int it = 0;
  double error = 0.0;
  for (;;) {
    it = it + 1;
    if (it > itmax_mach) break;
    double maxdf = -INFINITY;
    int nan_seen = 0;
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
          long c = cidx(g, i, j, k);
          double vc[NV], hc[NRV], cpc;
          for (int l = 0; l < NV; l++) vc[l] = v[l * nc + c];
          double DF = newton_cell(thd, vc, T[c], &cpc, hc);
          cp[c] = cpc;
          for (int l = 0; l < NRV; l++) ha[l * nc + c] = hc[l];
          T[c] = T[c] + DF;
          DF = DF * DF;
          if (DF != DF) nan_seen = 1;
          maxdf = fmax(maxdf, DF);
        }
    error = nan_seen ? NAN : sqrt(maxdf);
    if (error < eps_temp_mach) break;
  }
  int warn = 0;
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++)
        if (T[cidx(g, i, j, k)] < 0.0) warn |= 2;
  if (it > itmax_mach) warn |= 8;
  if (iters) *iters = it;
  return warn;
}

/* upd_prim_var_woT: src/module_thermodynamics.f90:522-689 */
void ora_upd_prim_var_woT(int domain, const ora_grid *g, const ora_thd *thd, const double *v,
                          const double *T, double *W_i, double *cp, double *ha) {
  int i0, i1, j0, j1, k0, k1;
  long nc = ncell(g);
  domain_select(g, domain, &i0, &i1, &j0, &j1, &k0, &k1);
  ora_wmix_i(domain, g, thd, v, W_i);
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        long c = cidx(g, i, j, k);
        double vc[NV], hc[NRV], cpc;
        for (int l = 0; l < NV; l++) vc[l] = v[l * nc + c];
        /* the same per-species branches as newton; the F part is unused */
        (void)newton_cell(thd, vc, T[c], &cpc, hc);
        cp[c] = cpc;
        for (int l = 0; l < NRV; l++) ha[l * nc + c] = hc[l];
      }
}

/* comm_cons on one rank: src/module_parallel.f90:1715-1777.  Periodic
 * directions exchange with themselves through the periodic Cartesian
 * communicator (faces only, edges and corners are not written); a physical
 * (non-periodic) face has no neighbour and is left to upd_boundaries. */
void ora_comm_cons(const ora_grid *g, double *v) {
  int nx = g->nx, ny = g->ny, nz = g->nz, ng = g->ng;
  long nc = ncell(g);
  for (int l = 0; l < NV; l++) {
    double *a = v + l * nc;
    if (is_periodic(g, 0))
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int s = 1; s <= ng; s++) {
            a[cidx(g, -s, j, k)] = a[cidx(g, nx - s, j, k)];
            a[cidx(g, nx - 1 + s, j, k)] = a[cidx(g, s - 1, j, k)];
          }
    if (g->ndim >= 2 && is_periodic(g, 1))
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++)
          for (int s = 1; s <= ng; s++) {
            a[cidx(g, i, -s, k)] = a[cidx(g, i, ny - s, k)];
            a[cidx(g, i, ny - 1 + s, k)] = a[cidx(g, i, s - 1, k)];
          }
    if (g->ndim == 3 && is_periodic(g, 2))
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++)
          for (int s = 1; s <= ng; s++) {
            a[cidx(g, i, j, -s)] = a[cidx(g, i, j, nz - s)];
            a[cidx(g, i, j, nz - 1 + s)] = a[cidx(g, i, j, s - 1)];
          }
  }
}

/* upd_boundaries with bc_extrapolation (src/module_BCs.f90:41-141,
 * 327-412): the ng ghost layers of an extrapolated face take the edge
 * value of the interior, face slab only, faces in the order W, E, S, N, B,
 * F. */
void ora_upd_boundaries(const ora_grid *g, double *v) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, ng = g->ng;
  const long nc = ncell(g);
  for (int f = 0; f < 2 * g->ndim; f++) {
    if (g->bc[f] == ORA_BC_POSTSHOCK_SLIPWALL && f == 2) {
      /* bc_postshock_slipwall (src/module_BCs.f90:1065-1172), face S only:
       * on x(i) <= x0 (the first nfix cells) the post-shock state on the
       * ghost rows and on the first interior row; beyond, the ghost rows
       * mirror the rows above the first one (rho v odd) */
      for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
          if (i < g->nfix) {
            for (int j = -ng; j <= 0; j++)
              for (int l = 0; l < NV; l++) v[l * nc + cidx(g, i, j, k)] = g->wall[l];
          } else {
            for (int s = 1; s <= ng; s++)
              for (int l = 0; l < NV; l++) {
                const double a = v[l * nc + cidx(g, i, s, k)];
                v[l * nc + cidx(g, i, -s, k)] = (l == 2) ? -a : a;
              }
          }
        }
      continue;
    }
    if (g->bc[f] != ORA_BC_EXTRAPOLATION) continue;
    const int d = f / 2, hi = f % 2;
    for (int l = 0; l < NV; l++) {
      double *a = v + l * nc;
      if (d == 0)
        for (int k = 0; k < nz; k++)
          for (int j = 0; j < ny; j++)
            for (int s = 1; s <= ng; s++) {
              if (hi) a[cidx(g, nx - 1 + s, j, k)] = a[cidx(g, nx - 1, j, k)];
              else a[cidx(g, -s, j, k)] = a[cidx(g, 0, j, k)];
            }
      else if (d == 1)
        for (int k = 0; k < nz; k++)
          for (int s = 1; s <= ng; s++)
            for (int i = 0; i < nx; i++) {
              if (hi) a[cidx(g, i, ny - 1 + s, k)] = a[cidx(g, i, ny - 1, k)];
              else a[cidx(g, i, -s, k)] = a[cidx(g, i, 0, k)];
            }
      else
        for (int s = 1; s <= ng; s++)
          for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
              if (hi) a[cidx(g, i, j, nz - 1 + s)] = a[cidx(g, i, j, nz - 1)];
              else a[cidx(g, i, j, -s)] = a[cidx(g, i, j, 0)];
            }
    }
  }
}

/* bc_nscbc_out: src/module_BCs.f90:2331-2516 (face E).  Per row: the
 * one-sided fourth-order derivatives (alph4or) of rho, u, v, w, P and Y
 * over cells ex-4..ex, the LODI amplitudes of a subsonic outflow with L1
 * relaxed toward pinf, and the state of cell ex advanced by -dt d(.)
 * written to ex..ex+ng. */
void ora_bc_nscbc_out(const ora_grid *g, const ora_thd *thd, double dt, const double *dx_i,
                      const double *T, const double *W_i, const double *cp, double *v) {
  const long nc = ncell(g);
  const int ex = g->nx - 1;
  static const double alph4or[5] = {-3.0 / 12.0, 16.0 / 12.0, -36.0 / 12.0, 48.0 / 12.0,
                                    -25.0 / 12.0};
  const double sigma1 = 0.28, Pinf = g->nscbc_pinf, Lx = g->nscbc_Lx, dtime = dt;
  const double dxi = dx_i[ex + g->ng];
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++) {
      double Machm = 0.0;
      long c = cidx(g, ex, j, k);
      double rho = v[c], rho_i = 1.0 / rho;
      double ux = v[nc + c] * rho_i, vy = v[2 * nc + c] * rho_i, wz = v[3 * nc + c] * rho_i;
      double Pres = v[c] * T[c] * W_i[c];
      double gam = thd->gam2 * W_i[c] / cp[c];
      gam = 1.0 / (1.0 - gam);
      double cs = sqrt(gam * Pres * rho_i);
      double Mach = sqrt(ux * ux + vy * vy + wz * wz) / cs;
      Machm = fmax(Machm, Mach);
      double drho = 0.0, dux = 0.0, dvy = 0.0, dwz = 0.0, dPres = 0.0, dYa[NRV], Ya[NRV];
      for (int l = 0; l < NRV; l++) dYa[l] = 0.0;
      for (int m = -4; m <= 0; m++) {
        long cm = cidx(g, ex + m, j, k);
        rho = v[cm];
        rho_i = 1.0 / rho;
        ux = v[nc + cm] * rho_i;
        vy = v[2 * nc + cm] * rho_i;
        wz = v[3 * nc + cm] * rho_i;
        Pres = v[cm] * T[cm] * W_i[cm];
        for (int l = 0; l < NRV; l++) Ya[l] = v[(NIV + l) * nc + cm] * rho_i;
        drho = drho + alph4or[m + 4] * rho;
        dux = dux + alph4or[m + 4] * ux;
        dvy = dvy + alph4or[m + 4] * vy;
        dwz = dwz + alph4or[m + 4] * wz;
        dPres = dPres + alph4or[m + 4] * Pres;
        for (int l = 0; l < NRV; l++) dYa[l] = dYa[l] + alph4or[m + 4] * Ya[l];
      }
      drho = -drho * dxi;
      dux = -dux * dxi;
      dvy = -dvy * dxi;
      dwz = -dwz * dxi;
      dPres = -dPres * dxi;
      for (int l = 0; l < NRV; l++) dYa[l] = -dYa[l] * dxi;
      double L1 = 0.0, L2 = 0.0, L3 = 0.0, L4 = 0.0, L5 = 0.0, L5_Ya[NRV];
      for (int l = 0; l < NRV; l++) L5_Ya[l] = 0.0;
      gam = thd->gam2 * W_i[c] / cp[c];
      gam = 1.0 / (1.0 - gam);
      cs = sqrt(gam * Pres * rho_i);
      if (ux - cs < 0.0) L1 = sigma1 * (1.0 - Machm * Machm) * cs / Lx * (Pres - Pinf);
      else L1 = (ux - cs) * (dPres - rho * cs * dux);
      if (ux + cs > 0.0) L5 = (ux + cs) * (dPres + rho * cs * dux);
      else L5 = (ux + cs) * (2.0 * dPres - L1 / (ux - cs));
      if (ux > 0.0) {
        L2 = ux * (cs * cs * drho - dPres);
        L3 = ux * dvy;
        L4 = ux * dwz;
        for (int l = 0; l < NRV; l++) L5_Ya[l] = ux * dYa[l];
      } else {
        L2 = ux * cs * cs * drho - 0.5 * ux * (L5 / (ux + cs) + L1 / (ux - cs));
      }
      double d2 = 0.5 * (L5 + L1);
      double d1 = (L2 + d2) / (cs * cs);
      double d3 = 0.5 * (L5 - L1) / (cs * rho);
      double d4 = L3, d5 = L4, d5_Ya[NRV];
      for (int l = 0; l < NRV; l++) d5_Ya[l] = L5_Ya[l];
      double has[NRV];
      ora_ha_scalar(thd, T[c], has);
      double hm = 0.0;
      for (int l = 0; l < NRV; l++) hm = hm + has[l] * Ya[l];
      for (int i = ex; i <= ex + g->ng; i++) {
        long ci = cidx(g, i, j, k);
        v[ci] = rho - d1 * dtime;
        v[nc + ci] = (rho * ux) - (ux * d1 + rho * d3) * dtime;
        v[2 * nc + ci] = (rho * vy) - (vy * d1 + rho * d4) * dtime;
        v[3 * nc + ci] = (rho * wz) - (wz * d1 + rho * d5) * dtime;
        v[4 * nc + ci] = (rho * hm - Pres + 0.5 * rho * (ux * ux + vy * vy + wz * wz)) -
                         (0.5 * (ux * ux + vy * vy + wz * wz) * d1 +
                          rho * (ux * d3 + vy * d4 + wz * d5) + d2 / (gam - 1.0)) * dtime;
        for (int l = 0; l < NRV; l++)
          v[(NIV + l) * nc + ci] = (rho * Ya[l]) - (Ya[l] * d1 + rho * d5_Ya[l]) * dtime;
      }
    }
}

/* upd_boundaries in full: src/module_BCs.f90:41-141, the faces in the order
 * W, E, S, N, B, F, each followed, when its condition writes the first
 * interior row (the post-shock wall, the characteristic outflow), by
 * upd_prim_var_dirichlet on that row (force_sumY and prim_inv_var,
 * :131, 297-314).  dt: the step (for the outflow). */
int ora_boundaries(const ora_grid *g, const ora_thd *thd, int forcesum, double dt,
                   const double *dx_i, double *v, double *W_i, double *T, double *cp,
                   double *ha) {
  static const int row[6] = {-10, 10, -20, 20, -30, 30};
  int st = 0;
  for (int f = 0; f < 2 * g->ndim; f++) {
    if (g->bc[f] == ORA_BC_PERIODIC) continue;
    int dirichlet = 0;
    if (g->bc[f] == ORA_BC_NSCBC_OUT && f == 1) {
      ora_bc_nscbc_out(g, thd, dt, dx_i, T, W_i, cp, v);
      dirichlet = 1;
    } else {
      ora_grid one = *g;
      for (int q = 0; q < 6; q++)
        if (q != f) one.bc[q] = ORA_BC_PERIODIC;
      ora_upd_boundaries(&one, v);
      dirichlet = g->bc[f] == ORA_BC_POSTSHOCK_SLIPWALL;
    }
    if (dirichlet) {
      if (forcesum) ora_force_sumY(row[f], g, v);
      int s = ora_prim_inv_var(row[f], g, thd, v, W_i, T, cp, ha, NULL);
      if (s && !st) st = s;
    }
  }
  return st;
}

/* the primitives of the Dirichlet row of each face whose condition writes
 * the first interior row (upd_prim_var_dirichlet after the face's
 * condition, src/module_BCs.f90:93-96, 131, 297-314): force_sumY and
 * prim_inv_var on that row */
int ora_upd_dirichlet(const ora_grid *g, const ora_thd *thd, int forcesum, double *v,
                      double *W_i, double *T, double *cp, double *ha) {
  static const int row[6] = {-10, 10, -20, 20, -30, 30};
  int st = 0;
  for (int f = 0; f < 2 * g->ndim; f++) {
    if (g->bc[f] != ORA_BC_POSTSHOCK_SLIPWALL) continue;
    if (forcesum) ora_force_sumY(row[f], g, v);
    int s = ora_prim_inv_var(row[f], g, thd, v, W_i, T, cp, ha, NULL);
    if (s && !st) st = s;
  }
  return st;
}

/* upd_prim_var_ghost: src/module_BCs.f90:254-284 (faces -1,1,-2,2,-3,3) */
int ora_upd_prim_var_ghost(const ora_grid *g, const ora_thd *thd, int forcesum, double *v,
                           double *W_i, double *T, double *cp, double *ha) {
  static const int face[6] = {-1, 1, -2, 2, -3, 3};
  int st = 0;
  for (int l = 0; l < 2 * g->ndim; l++) {
    if (forcesum) ora_force_sumY(face[l], g, v);
    int s = ora_prim_inv_var(face[l], g, thd, v, W_i, T, cp, ha, NULL);
    if (s && !st) st = s;
  }
  return st;
}

/* shock_sensor_modified_adams_shariff: src/module_tools.f90:2733-2814.
 * Sticky: flags are only ever set. */
void ora_sensor_adams_shariff(const ora_grid *g, double max_rel_weight, const double *v,
                              const double *T, const double *W_i, uint8_t *sensor) {
  int nx = g->nx, ny = g->ny, nz = g->nz;
  long nc = ncell(g);
  (void)nc;
  for (int d = 0; d < g->ndim; d++) {
    int n = (d == 0) ? nx : (d == 1 ? ny : nz);
    int np = g->nproc[d] > 1 ? g->nproc[d] : 1;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          int p = (d == 0) ? i : (d == 1 ? j : k);
          /* the sweep of each rank stops one short of its last cell
           * (ranks own (r n)/np .. ((r+1) n)/np - 1, module_parallel.f90:
           * 547-548) */
          int r = 0;
          while (r + 1 < np && ((long)(r + 1) * n) / np <= p) r++;
          if (p >= ((long)(r + 1) * n) / np - 1) continue;
          int i2 = i + (d == 0), j2 = j + (d == 1), k2 = k + (d == 2);
          long c = cidx(g, i, j, k), c2 = cidx(g, i2, j2, k2);
          double ri = 1.0 / v[c], ri2 = 1.0 / v[c2];
          double P = v[c] * T[c] * W_i[c], P2 = v[c2] * T[c2] * W_i[c2];
          double drho = fabs(v[c2] - v[c]) * fmin(ri, ri2);
          double dpres = fabs(P2 - P) / fmin(P, P2);
          if (drho > max_rel_weight && dpres > max_rel_weight) sensor[c] = 1;
        }
  }
}

/* shock_sensor_ducros: src/module_tools.f90:2665-2719, with the velocity
 * gradients of compute_duixj (:2521-2651: velocities on the interior and
 * the face ghost slabs, then d?_fixed1, derivatives as dgrid below).
 * Sticky: flags are only ever set. */
static inline double dgrid(const ora_grid *g, const double *a, long c, int d, int i, int j, int k,
                           double dinv);

void ora_sensor_ducros(const ora_grid *g, double treduc, const double *dx_i, const double *dy_i,
                       const double *dz_i, const double *v, uint8_t *sensor) {
  const long nc = ncell(g);
  const int ng = g->ng, nd = g->ndim;
  double *vel = calloc((size_t)3 * nc, sizeof(double));
  int i0, i1, j0, j1, k0, k1;
  for (int d = -nd; d <= nd; d++) {
    domain_select(g, d, &i0, &i1, &j0, &j1, &k0, &k1);
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
          long c = cidx(g, i, j, k);
          double rho_i = 1.0 / v[c];
          for (int q = 0; q < nd; q++) vel[q * nc + c] = v[(1 + q) * nc + c] * rho_i;
        }
  }
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++) {
        long c = cidx(g, i, j, k);
        const double dinv[3] = {dx_i[i + ng], dy_i[j + ng], dz_i[k + ng]};
        double du[3][3] = {{0}}; /* du[a][b] = d u_a / d x_b: duixj(3 a + b + 1) */
        for (int a = 0; a < nd; a++)
          for (int b = 0; b < nd; b++) du[a][b] = dgrid(g, vel + a * nc, c, b, i, j, k, dinv[b]);
        double div, vort, wrk;
        if (nd == 2) {
          div = du[0][0] + du[1][1];
          div = div * div;
          wrk = du[1][0] - du[0][1];
          vort = wrk * wrk;
        } else {
          div = du[0][0] + du[1][1] + du[2][2];
          div = div * div;
          wrk = du[2][1] - du[1][2];
          vort = wrk * wrk;
          wrk = du[0][2] - du[2][0];
          wrk = wrk * wrk;
          vort = vort + wrk;
          wrk = du[1][0] - du[0][1];
          wrk = wrk * wrk;
          vort = vort + wrk;
        }
        double ducloc = div / (div + vort + 1.0e-30);
        if (ducloc > treduc) sensor[c] = 1;
      }
  free(vel);
}

/* --------------------------------------------------------- transport */

/* prim_vis_vars for one cell: src/module_thermodynamics.f90:989-1230 */
void ora_transport_cell(const ora_thd *thd, double rho, double T, double W_i, const double *Xa,
                        double *dm, double *mu, double *ct) {
  const double eps = 1.0e-12; /* epsm12_asp, src/thd/param_thd.f90:19 */
  double Kl = 0.5 * 1.0e-5 / thd->laminf;
  double Ku = 1.0e-1 / thd->muinf;
  double lt[4];
  lt[0] = 1.0;
  lt[1] = log(T * thd->Tinf);
  lt[2] = lt[1] * lt[1];
  lt[3] = lt[2] * lt[1];
  double P = rho * T * W_i;
  for (int m = 0; m < NRV; m++) {
    double Kd = 1.0e-4 * thd->p0 / (thd->pinf * thd->Dinf);
    double wrk1 = 0.0, d = 0.0;
    for (int l = 0; l < NRV; l++) {
      if (m == l) continue;
      double dkj = 0.0;
      for (int n = 0; n < 4; n++) dkj = dkj + thd->CofD[m][l][n] * lt[n];
      wrk1 = wrk1 + (Xa[l] + eps) * thd->Wc[l] * W_i;
      d = d + (Xa[l] + eps) * (1.0 / exp(dkj));
    }
    dm[m] = wrk1 * Kd / (d * P);
  }
  double w1 = 0.0, c = 0.0;
  for (int l = 0; l < NRV; l++) {
    double s = 0.0;
    for (int n = 0; n < 4; n++) s = s + thd->CofLam[l][n] * lt[n];
    double c1 = exp(s), c2 = Xa[l];
    c = c + c2 / c1;
    w1 = w1 + c2 * c1;
  }
  *ct = Kl * (1.0 / c + w1);
  double eta[NRV];
  for (int l = 0; l < NRV; l++) {
    double s = 0.0;
    for (int n = 0; n < 4; n++) s = s + thd->CofEta[l][n] * lt[n];
    eta[l] = exp(s);
  }
  double mus = 0.0;
  for (int m = 0; m < NRV; m++) {
    double c4 = 0.0;
    for (int l = 0; l < NRV; l++) {
      double c1 = thd->wilke_c1[m][l];
      double c2 = thd->wilke_c2[m][l];
      double c3 = 1.0 + c1 * sqrt(eta[m] / eta[l]);
      c2 = c3 * c3 * c2;
      c4 = c4 + c2 * Xa[l];
    }
    mus = mus + Ku * (eta[m] * Xa[m] / c4);
  }
  *mu = mus;
}

/* First derivatives: src/module_deriv.f90:20-78 (weights), 212-469 (dx;
 * d?_fixed1 at :926-1171 and d?_fixed2 at :1644-1720 have the same
 * structure).  Central 6th order (ng = 3) or 8th order (ng = 4), summed over
 * the whole stencil including its zero weight; next to a physical
 * (non-periodic) boundary the ng points nearest to it take the closures:
 * 4th-order one-sided, 4th-order upwind, 4th-order central and, for ng = 4,
 * 6th-order central (the right-hand ones negated).  The right boundary's
 * block is applied after the left one's.  The reference's z operators use
 * the 8-point stencil whatever ng is, which reads k+-4 of an ng = 3 halo;
 * with ng = 3 this restatement uses the 6th-order stencil in z as in x and
 * y. */
static const double D6[7] = {-1.0 / 60.0, 9.0 / 60.0, -45.0 / 60.0, 0.0,
                             45.0 / 60.0, -9.0 / 60.0, 1.0 / 60.0};
static const double D8[9] = {3.0 / 840.0, -32.0 / 840.0, 168.0 / 840.0, -672.0 / 840.0, 0.0,
                             672.0 / 840.0, -168.0 / 840.0, 32.0 / 840.0, -3.0 / 840.0};
static const double D4[5] = {1.0 / 12.0, -8.0 / 12.0, 0.0, 8.0 / 12.0, -1.0 / 12.0};
static const double D4OL[5] = {-25.0 / 12.0, 48.0 / 12.0, -36.0 / 12.0, 16.0 / 12.0,
                               -3.0 / 12.0};                      /* l = 0..4 */
static const double D4UL[5] = {-3.0 / 12.0, -10.0 / 12.0, 18.0 / 12.0, -6.0 / 12.0,
                               1.0 / 12.0};                       /* l = -1..3 */
static const double D4UR[5] = {1.0 / 12.0, -6.0 / 12.0, 18.0 / 12.0, -10.0 / 12.0,
                               -3.0 / 12.0};                      /* l = -3..1 */
static const double D4OR[5] = {-3.0 / 12.0, 16.0 / 12.0, -36.0 / 12.0, 48.0 / 12.0,
                               -25.0 / 12.0};                     /* l = -4..0 */

static inline double wsum(const double *w, int lo, int hi, const double *a, long c, long s) {
  double dd = 0.0;
  for (int l = lo; l <= hi; l++) dd = dd + w[l - lo] * a[c + l * s];
  return dd;
}

static double deriv_at(const double *a, long c, long s, double dinv, int p, int n, int ng,
                       int lphys, int rphys) {
  if (rphys && p >= n - ng) {
    if (p == n - 1) return -(wsum(D4OR, -4, 0, a, c, s) * dinv);
    if (p == n - 2) return -(wsum(D4UR, -3, 1, a, c, s) * dinv);
    if (p == n - 3) return wsum(D4, -2, 2, a, c, s) * dinv;
    return wsum(D6, -3, 3, a, c, s) * dinv;           /* p == n - 4, ng = 4 */
  }
  if (lphys && p < ng) {
    if (p == 0) return wsum(D4OL, 0, 4, a, c, s) * dinv;
    if (p == 1) return wsum(D4UL, -1, 3, a, c, s) * dinv;
    if (p == 2) return wsum(D4, -2, 2, a, c, s) * dinv;
    return wsum(D6, -3, 3, a, c, s) * dinv;           /* p == 3, ng = 4 */
  }
  if (ng == 4) return wsum(D8, -4, 4, a, c, s) * dinv;
  return wsum(D6, -3, 3, a, c, s) * dinv;
}

double ora_deriv_line(const double *a, long stride, double dinv, int p, int n, int ng,
                      int lphys, int rphys) {
  return deriv_at(a, (long)p * stride, stride, dinv, p, n, ng, lphys, rphys);
}

/* derivative along direction d of field a at interior cell (i, j, k) = c */
static inline double dgrid(const ora_grid *g, const double *a, long c, int d, int i, int j, int k,
                           double dinv) {
  const int n = (d == 0) ? g->nx : (d == 1 ? g->ny : g->nz);
  const int p = (d == 0) ? i : (d == 1 ? j : k);
  const long s = (d == 0) ? 1 : (d == 1 ? NXg(g) : NXg(g) * NYg(g));
  return deriv_at(a, c, s, dinv, p, n, g->ng, g->bc[2 * d] != ORA_BC_PERIODIC,
                  g->bc[2 * d + 1] != ORA_BC_PERIODIC);
}

/* viscflux_xy / viscflux_xyz: src/module_viscflux.f90:721-1135, 1660-2189.
 * Periodic or extrapolated faces (no wall, symmetry or injection branches);
 * comm_deriv fills the derivative ghosts of periodic directions only. */
int ora_viscflux(const ora_grid *g, const ora_thd *thd, const ora_adi *adi, const double *dx_i,
                 const double *dy_i, const double *dz_i, const double *T, const double *W_i,
                 const double *ha, const double *v, double *fl) {
  const int ndim = g->ndim, ng = g->ng;
  const long nc = ncell(g);
  const int nfluid = ndim * (ndim + 1);
  const int nderiv = nfluid + ndim * NRV;
  const double twothirds = 2.0 / 3.0;
  const double rvd0 = adi->sqgmr / adi->Sc;
  const double rlam0 = adi->sqgmr * adi->ggmopr;
  double *vel = calloc((size_t)3 * nc, sizeof(double));
  double *Xa = calloc((size_t)NRV * nc, sizeof(double));
  double *fd = calloc((size_t)nderiv * nc, sizeof(double));
  int nflux = (ndim == 3) ? 9 : 5; /* sx_x..sz_z / q_x.. */
  double *sf = calloc((size_t)nflux * nc, sizeof(double));
  int i0, i1, j0, j1, k0, k1;
  /* molefrac + velocities on interior and face ghosts (:1738-1760) */
  for (int d = -ndim; d <= ndim; d++) {
    domain_select(g, d, &i0, &i1, &j0, &j1, &k0, &k1);
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
          long c = cidx(g, i, j, k);
          for (int l = 0; l < NRV; l++)
            Xa[l * nc + c] = v[(NIV + l) * nc + c] / (v[c] * W_i[c] * thd->Wc[l]);
          double rho_i = 1.0 / v[c];
          for (int q = 0; q < ndim; q++) vel[q * nc + c] = v[(1 + q) * nc + c] * rho_i;
        }
  }
  /* first derivatives on the interior (d?_fixed1) */
  const double *dinv[3] = {dx_i, dy_i, dz_i};
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++) {
        long c = cidx(g, i, j, k);
        int idx3[3] = {i + ng, j + ng, k + ng};
        int slot = 0;
        for (int q = 0; q < ndim; q++)          /* du/dx_j, dv/dx_j, dw/dx_j */
          for (int dd = 0; dd < ndim; dd++)
            fd[(slot++) * nc + c] = dgrid(g, vel + q * nc, c, dd, i, j, k, dinv[dd][idx3[dd]]);
        for (int dd = 0; dd < ndim; dd++)       /* dT/dx_j */
          fd[(slot++) * nc + c] = dgrid(g, T, c, dd, i, j, k, dinv[dd][idx3[dd]]);
        for (int l = 0; l < NRV; l++)           /* dX_l/dx_j */
          for (int dd = 0; dd < ndim; dd++)
            fd[(slot++) * nc + c] = dgrid(g, Xa + l * nc, c, dd, i, j, k, dinv[dd][idx3[dd]]);
      }
  /* comm_deriv: periodic self-exchange of fd on the faces (:1923) */
  {
    ora_grid gd = *g;
    for (int f = 0; f < nderiv; f += NV) {
      /* reuse comm_cons on blocks of NV fields */
      int nb = (nderiv - f < NV) ? nderiv - f : NV;
      if (nb == NV) ora_comm_cons(&gd, fd + (long)f * nc);
      else {
        double *tmp = calloc((size_t)NV * nc, sizeof(double));
        memcpy(tmp, fd + (long)f * nc, sizeof(double) * nb * nc);
        ora_comm_cons(&gd, tmp);
        memcpy(fd + (long)f * nc, tmp, sizeof(double) * nb * nc);
        free(tmp);
      }
    }
  }
  /* fluxes on interior and face ghosts (:1930-2041) */
  for (int d = -ndim; d <= ndim; d++) {
    domain_select(g, d, &i0, &i1, &j0, &j1, &k0, &k1);
    for (int k = k0; k <= k1; k++)
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
          long c = cidx(g, i, j, k);
          double X[NRV], dm[NRV], mu, ct;
          for (int l = 0; l < NRV; l++) X[l] = Xa[l * nc + c];
          ora_transport_cell(thd, v[c], T[c], W_i[c], X, dm, &mu, &ct);
          double rho_i = 1.0 / v[c];
          double uxs = v[nc + c] * rho_i, vys = v[2 * nc + c] * rho_i, wzs = v[3 * nc + c] * rho_i;
          double rmu = -adi->sqgmr * mu;
          double rvd = rvd0 * v[c];
          double rlam = -rlam0 * ct;
#define FD(s) fd[(long)((s) - 1) * nc + c]
          if (ndim == 3) {
            double sxx = twothirds * rmu * (FD(1) + FD(1) - (FD(5) + FD(9)));
            double sxy = rmu * (FD(4) + FD(2));
            double sxz = rmu * (FD(7) + FD(3));
            double syy = twothirds * rmu * (FD(5) + FD(5) - (FD(1) + FD(9)));
            double syz = rmu * (FD(8) + FD(6));
            double szz = twothirds * rmu * (FD(9) + FD(9) - (FD(1) + FD(5)));
            double vdcx = 0.0, vdcy = 0.0, vdcz = 0.0;
            for (int l = 1; l <= NRV; l++) {
              double wrk = dm[l - 1] * thd->Wc[l - 1] * W_i[c];
              vdcx = vdcx + wrk * FD(nfluid + ndim * l - 2);
              vdcy = vdcy + wrk * FD(nfluid + ndim * l - 1);
              vdcz = vdcz + wrk * FD(nfluid + ndim * l);
            }
            double vdhx = 0.0, vdhy = 0.0, vdhz = 0.0;
            for (int l = 1; l <= NRV; l++) {
              double Ya = v[(NIV + l - 1) * nc + c] * rho_i;
              double vdtx = vdcx * Ya, vdty = vdcy * Ya, vdtz = vdcz * Ya;
              double wrk = -dm[l - 1] * thd->Wc[l - 1] * W_i[c];
              vdtx = vdtx + wrk * FD(nfluid + ndim * l - 2);
              vdty = vdty + wrk * FD(nfluid + ndim * l - 1);
              vdtz = vdtz + wrk * FD(nfluid + ndim * l);
              FD(nfluid + ndim * l - 2) = rvd * vdtx;
              FD(nfluid + ndim * l - 1) = rvd * vdty;
              FD(nfluid + ndim * l) = rvd * vdtz;
              double h = ha[(long)(l - 1) * nc + c];
              vdhx = vdhx + h * FD(nfluid + ndim * l - 2);
              vdhy = vdhy + h * FD(nfluid + ndim * l - 1);
              vdhz = vdhz + h * FD(nfluid + ndim * l);
            }
            sf[0 * nc + c] = sxx; sf[1 * nc + c] = sxy; sf[2 * nc + c] = sxz;
            sf[3 * nc + c] = syy; sf[4 * nc + c] = syz; sf[5 * nc + c] = szz;
            sf[6 * nc + c] = uxs * sxx + vys * sxy + wzs * sxz + rlam * FD(10) + vdhx;
            sf[7 * nc + c] = uxs * sxy + vys * syy + wzs * syz + rlam * FD(11) + vdhy;
            sf[8 * nc + c] = uxs * sxz + vys * syz + wzs * szz + rlam * FD(12) + vdhz;
          } else {
            double sxx = twothirds * rmu * (FD(1) + FD(1) - FD(4));
            double sxy = rmu * (FD(3) + FD(2));
            double syy = twothirds * rmu * (FD(4) + FD(4) - FD(1));
            double vdcx = 0.0, vdcy = 0.0;
            for (int l = 1; l <= NRV; l++) {
              double wrk = dm[l - 1] * thd->Wc[l - 1] * W_i[c];
              vdcx = vdcx + wrk * FD(nfluid + ndim * l - 1);
              vdcy = vdcy + wrk * FD(nfluid + ndim * l);
            }
            double vdhx = 0.0, vdhy = 0.0;
            for (int l = 1; l <= NRV; l++) {
              double Ya = v[(NIV + l - 1) * nc + c] * rho_i;
              double vdtx = vdcx * Ya, vdty = vdcy * Ya;
              double wrk = -dm[l - 1] * thd->Wc[l - 1] * W_i[c];
              vdtx = vdtx + wrk * FD(nfluid + ndim * l - 1);
              vdty = vdty + wrk * FD(nfluid + ndim * l);
              FD(nfluid + ndim * l - 1) = rvd * vdtx;
              FD(nfluid + ndim * l) = rvd * vdty;
              double h = ha[(long)(l - 1) * nc + c];
              vdhx = vdhx + h * FD(nfluid + ndim * l - 1);
              vdhy = vdhy + h * FD(nfluid + ndim * l);
            }
            sf[0 * nc + c] = sxx; sf[1 * nc + c] = sxy; sf[2 * nc + c] = syy;
            sf[3 * nc + c] = uxs * sxx + vys * sxy + rlam * FD(5) + vdhx;
            sf[4 * nc + c] = uxs * sxy + vys * syy + rlam * FD(6) + vdhy;
          }
#undef FD
        }
  }
  /* divergence (dx/dy/dz, d?_fixed2) accumulated into fl (:2087-2183) */
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++) {
        long c = cidx(g, i, j, k);
        double dxi = dx_i[i + ng], dyi = dy_i[j + ng], dzi = dz_i[k + ng];
        if (ndim == 3) {
          double wx, wy, wz;
          wx = dgrid(g, sf + 0 * nc, c, 0, i, j, k, dxi); wy = dgrid(g, sf + 1 * nc, c, 1, i, j, k, dyi);
          wz = dgrid(g, sf + 2 * nc, c, 2, i, j, k, dzi);
          fl[1 * nc + c] = fl[1 * nc + c] + (wx + wy + wz);
          wx = dgrid(g, sf + 1 * nc, c, 0, i, j, k, dxi); wy = dgrid(g, sf + 3 * nc, c, 1, i, j, k, dyi);
          wz = dgrid(g, sf + 4 * nc, c, 2, i, j, k, dzi);
          fl[2 * nc + c] = fl[2 * nc + c] + (wx + wy + wz);
          wx = dgrid(g, sf + 2 * nc, c, 0, i, j, k, dxi); wy = dgrid(g, sf + 4 * nc, c, 1, i, j, k, dyi);
          wz = dgrid(g, sf + 5 * nc, c, 2, i, j, k, dzi);
          fl[3 * nc + c] = fl[3 * nc + c] + (wx + wy + wz);
          wx = dgrid(g, sf + 6 * nc, c, 0, i, j, k, dxi); wy = dgrid(g, sf + 7 * nc, c, 1, i, j, k, dyi);
          wz = dgrid(g, sf + 8 * nc, c, 2, i, j, k, dzi);
          fl[4 * nc + c] = fl[4 * nc + c] + (wx + wy + wz);
          for (int l = 1; l <= NRV; l++) {
            wx = dgrid(g, fd + (long)(nfluid + ndim * l - 3) * nc, c, 0, i, j, k, dxi);
            wy = dgrid(g, fd + (long)(nfluid + ndim * l - 2) * nc, c, 1, i, j, k, dyi);
            wz = dgrid(g, fd + (long)(nfluid + ndim * l - 1) * nc, c, 2, i, j, k, dzi);
            fl[(NIV + l - 1) * nc + c] = fl[(NIV + l - 1) * nc + c] + (wx + wy + wz);
          }
        } else {
          double wx, wy;
          wx = dgrid(g, sf + 0 * nc, c, 0, i, j, k, dxi); wy = dgrid(g, sf + 1 * nc, c, 1, i, j, k, dyi);
          fl[1 * nc + c] = fl[1 * nc + c] + (wx + wy);
          wx = dgrid(g, sf + 1 * nc, c, 0, i, j, k, dxi); wy = dgrid(g, sf + 2 * nc, c, 1, i, j, k, dyi);
          fl[2 * nc + c] = fl[2 * nc + c] + (wx + wy);
          wx = dgrid(g, sf + 3 * nc, c, 0, i, j, k, dxi); wy = dgrid(g, sf + 4 * nc, c, 1, i, j, k, dyi);
          fl[4 * nc + c] = fl[4 * nc + c] + (wx + wy);
          for (int l = 1; l <= NRV; l++) {
            wx = dgrid(g, fd + (long)(nfluid + ndim * l - 2) * nc, c, 0, i, j, k, dxi);
            wy = dgrid(g, fd + (long)(nfluid + ndim * l - 1) * nc, c, 1, i, j, k, dyi);
            fl[(NIV + l - 1) * nc + c] = fl[(NIV + l - 1) * nc + c] + (wx + wy);
          }
        }
      }
  free(vel); free(Xa); free(fd); free(sf);
  return 0;
}

/* ------------------------------------------------------------ SSP-RK3 */

/* rk3: src/module_solver.f90:694-856 (periodic or extrapolated faces, no
 * LES/EGLIB/NSCBC).  f_out (optional, ncell*NV) receives the RHS of
 * the first stage. */
int ora_rk3(const ora_grid *g, const ora_thd *thd, const ora_adi *adi, const ora_opts *op,
            double dt, const double *dx_i, const double *dy_i, const double *dz_i, double *T,
            double *W_i, double *cp, double *ha, uint8_t *sensor, double *v, double *f_out) {
  static const double bk3[3] = {0.00, 0.75, 1.00 / 3.00};
  static const double ck3[3] = {1.00, 0.25, 2.00 / 3.00};
  const long nc = ncell(g);
  double *v0 = malloc(sizeof(double) * NV * nc);
  double *f = calloc((size_t)NV * nc, sizeof(double));
  memcpy(v0, v, sizeof(double) * NV * nc);
  int status = 0;
  for (int ik = 0; ik < 3; ik++) {
    /* module_solver.f90:758-762 */
    if (op->sensor_type == 1) ora_sensor_ducros(g, op->ducros_treduc, dx_i, dy_i, dz_i, v, sensor);
    else ora_sensor_adams_shariff(g, op->max_rel_weight, v, T, W_i, sensor);
    int s;
    const double *dd[3] = {dx_i, dy_i, dz_i};
    for (int d = 0; d < g->ndim; d++) {
      s = ora_euler_llf(d, g, thd, dd[d], T, W_i, cp, ha, sensor, op->weno_avg, v, f);
      if (s && !status) status = s;
    }
    if (op->vis) ora_viscflux(g, thd, adi, dx_i, dy_i, dz_i, T, W_i, ha, v, f);
    if (ik == 0 && f_out) memcpy(f_out, f, sizeof(double) * NV * nc);
    for (int l = 0; l < NV; l++)
      for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ny; j++)
          for (int i = 0; i < g->nx; i++) {
            long c = (long)l * nc + cidx(g, i, j, k);
            v[c] = v0[c] * bk3[ik] + ck3[ik] * (v[c] - dt * f[c]);
          }
    if (op->forcesum) ora_force_sumY(0, g, v);
    s = ora_prim_inv_var(0, g, thd, v, W_i, T, cp, ha, NULL);
    if (s && !status) status = s;
    ora_comm_cons(g, v);
    s = ora_boundaries(g, thd, op->forcesum, dt, dx_i, v, W_i, T, cp, ha);
    if (s && !status) status = s;
    s = ora_upd_prim_var_ghost(g, thd, op->forcesum, v, W_i, T, cp, ha);
    if (s && !status) status = s;
  }
  free(v0); free(f);
  return status;
}

/* timestep: src/module_solver.f90:202-408 (no LES: mu_SGS = 0; FIXEDDT off) */
void ora_timestep(const ora_grid *g, const ora_thd *thd, const ora_adi *adi, const ora_opts *op,
                  const double *dx_i, const double *dy_i, const double *dz_i, const double *cp,
                  const double *W_i, const double *T, const double *v, double DTmax,
                  double *dt_4, double *dtmin) {
  const long nc = ncell(g);
  const int ng = g->ng;
  double rvd0_i = adi->Sc / adi->sqgmr, rlamcp0_i = adi->Pr / adi->sqgmr;
  double dtc = 1.0;
  for (int q = 0; q < 4; q++) dt_4[q] = 1.0;
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++) {
        long c = cidx(g, i, j, k);
        double rho_i = 1.0 / v[c];
        double ux = v[nc + c] * rho_i, vy = v[2 * nc + c] * rho_i, wz = v[3 * nc + c] * rho_i;
        double gamma = thd->gam2 * W_i[c] / cp[c];
        gamma = 1.0 / (1.0 - gamma);
        double P = v[c] * T[c] * W_i[c];
        double csound = sqrt(gamma * P * rho_i);
        double dtcx_i = (fabs(ux) + csound) * dx_i[i + ng];
        double dtcy_i = (fabs(vy) + csound) * dy_i[j + ng];
        double dtcz_i = (fabs(wz) + csound) * dz_i[k + ng];
        dtc = fmax(fmax(fmax(dtc, dtcx_i), dtcy_i), dtcz_i);
      }
  dt_4[0] = op->CFL / dtc;
  if (op->vis && op->vis_dt) {
    /* molefrac + prim_vis_vars_wo_mu + filter_width (:297-374) */
    int ldmax = 0;
    double diffmax = -INFINITY;
    double *dmall = malloc(sizeof(double) * NRV * g->nx * g->ny * g->nz);
    double *ctall = malloc(sizeof(double) * g->nx * g->ny * g->nz);
    long q = 0;
    for (int k = 0; k < g->nz; k++)
      for (int j = 0; j < g->ny; j++)
        for (int i = 0; i < g->nx; i++, q++) {
          long c = cidx(g, i, j, k);
          double X[NRV], dm[NRV], mu, ct;
          for (int l = 0; l < NRV; l++) X[l] = v[(NIV + l) * nc + c] / (v[c] * W_i[c] * thd->Wc[l]);
          ora_transport_cell(thd, v[c], T[c], W_i[c], X, dm, &mu, &ct);
          for (int l = 0; l < NRV; l++) dmall[l * g->nx * g->ny * g->nz + q] = dm[l];
          ctall[q] = ct;
        }
    long ni = (long)g->nx * g->ny * g->nz;
    diffmax = dmall[0];
    for (int l = 0; l < NRV; l++)
      for (long p = 0; p < ni; p++)
        if (diffmax < dmall[l * ni + p]) { ldmax = l; diffmax = dmall[l * ni + p]; }
    double m2 = INFINITY, m3 = INFINITY;
    q = 0;
    for (int k = 0; k < g->nz; k++)
      for (int j = 0; j < g->ny; j++)
        for (int i = 0; i < g->nx; i++, q++) {
          long c = cidx(g, i, j, k);
          double delta;
          if (g->ndim == 3)
            delta = 1.0 / pow(dx_i[i + ng] * dy_i[j + ng] * dz_i[k + ng], 1.0 / 3.0);
          else if (g->ndim == 2)
            delta = 1.0 / sqrt(dx_i[i + ng] * dy_i[j + ng]);
          else
            delta = 1.0 / dx_i[i + ng];
          double d2 = delta * delta;
          double a = d2 / (dmall[ldmax * ni + q] + adi->Sc * 0.0 / v[c]);
          double b = d2 * v[c] / (ctall[q] / cp[c] + adi->Pr * 0.0);
          m2 = fmin(m2, a);
          m3 = fmin(m3, b);
        }
    dt_4[1] = rvd0_i * op->Fo * m2;
    dt_4[2] = rlamcp0_i * op->Fo * m3;
    free(dmall); free(ctall);
  }
  if (op->reaction) {
    dt_4[3] = *dtmin * op->CPR / DTmax;
    dt_4[3] = fmin(dt_4[3], 1.0);
    dt_4[3] = fmax(dt_4[3], dt_4[0] * op->CFLmin / op->CFL);
  }
  double mn = dt_4[0];
  for (int q2 = 1; q2 < 4; q2++) mn = fmin(mn, dt_4[q2]);
  *dtmin = mn;
}

/* ------------------------------------------------------------ kinetics */

/* CKSMH: src/chemkin/vode.f:1474-1535 */
static void cksmh(const ora_mech *m, double T, double *smh) {
  double tn[5];
  tn[0] = log(T) - 1.0;
  for (int n = 2; n <= 5; n++) tn[n - 1] = pow(T, n - 1) / ((n - 1) * n);
  for (int k = 0; k < m->kk; k++) {
    int L = (T > m->tmid[k]) ? 1 : 0;
    const double *a = m->a[k][L];
    double sum = 0.0;
    for (int n = 0; n < 5; n++) sum = sum + tn[n] * a[n];
    smh[k] = sum + a[6] - a[5] / T;
  }
}

static inline double ipow(double x, int n) {
  double r = 1.0;
  for (int q = 0; q < n; q++) r *= x;
  return r;
}

/* CKWYR = CKRATT + CKYTCR + CKRATX + stoichiometric sum:
 * src/chemkin/vode.f:1797-1904, 935-1126, 3171-3221, 1130-1314 */
void ora_ckwyr(const ora_mech *m, double rho, double T, const double *Y, double *wdot) {
  int ii = m->ii, kk = m->kk;
  double rkft[ORA_IMAX], rkrt[ORA_IMAX], eqk[ORA_IMAX], smh[ORA_KMAX], c[ORA_KMAX];
  double alogt = log(T);
  for (int i = 0; i < ii; i++) rkft[i] = m->A[i] * exp(m->b[i] * alogt - m->EoR[i] / T);
  cksmh(m, T, smh);
  for (int i = 0; i < ii; i++) {
    double sumsmh = 0.0;
    for (int n = 0; n < 6; n++)
      if (m->nunk[i][n] != 0) sumsmh = sumsmh + (double)m->nu[i][n] * smh[m->nunk[i][n] - 1];
    if (sumsmh != 0.0) eqk[i] = exp(fmin(sumsmh, m->exparg));
    else eqk[i] = 0.0; /* left unset in the reference; unused when the sum is 0 */
  }
  double pfac = m->patm / (m->ru * T);
  for (int i = 0; i < ii; i++) {
    int nusumk = 0;
    for (int n = 0; n < 6; n++) nusumk += m->nu[i][n];
    eqk[i] = eqk[i] * (nusumk >= 0 ? ipow(pfac, nusumk) : 1.0 / ipow(pfac, -nusumk));
  }
  for (int i = 0; i < ii; i++) {
    rkrt[i] = 0.0;
    if (m->reversible[i]) rkrt[i] = rkft[i] / fmax(eqk[i], m->small_);
  }
  /* reverse parameters given (REV): vode.f:1028-1032 */
  for (int i = 0; i < ii; i++)
    if (m->reversible[i] == 2)
      rkrt[i] = m->rpar[i][0] * exp(m->rpar[i][1] * alogt - m->rpar[i][2] / T);
  for (int k = 0; k < kk; k++) c[k] = rho * Y[k] / m->wt[k];
  double ctb[ORA_IMAX];
  for (int i = 0; i < ii; i++) ctb[i] = 1.0;
  double ctot = 0.0;
  for (int k = 0; k < kk; k++) ctot = ctot + c[k];
  for (int i = 0; i < ii; i++) {
    if (m->ntb[i] < 0) continue;
    ctb[i] = ctot;
    for (int l = 0; l < m->ntb[i]; l++)
      ctb[i] = ctb[i] + (m->tb_eff[i][l] - 1.0) * c[m->tb_k[i][l] - 1];
  }
  for (int i = 0; i < ii; i++) {
    if (!m->fal[i]) continue;
    const double *fp = m->fpar[i];
    double rklow = fp[0] * exp(fp[1] * alogt - fp[2] / T);
    double pr;
    if (m->kfal[i] == 0) { /* vode.f:1193-1199 */
      pr = rklow * ctb[i] / rkft[i];
      ctb[i] = 1.0;
    } else {
      pr = rklow * c[m->kfal[i] - 1] / rkft[i];
    }
    double pcor = pr / (1.0 + pr);
    if (m->fal[i] == 2) { /* 8-parameter SRI form, vode.f:1207-1216 */
      double prlog = log10(fmax(pr, m->small_));
      double xp = 1.0 / (1.0 + prlog * prlog);
      double fc = pow(fp[3] * exp(-fp[4] / T) + exp(-T / fp[5]), xp) * fp[6] * pow(T, fp[7]);
      pcor = fc * pcor;
    } else if (m->fal[i] > 2) {
      double prlog = log10(fmax(pr, m->small_));
      double fcent = (1.0 - fp[3]) * exp(-T / fp[4]) + fp[3] * exp(-T / fp[5]);
      if (m->fal[i] == 4) fcent = fcent + exp(-fp[6] / T);
      double fclog = log10(fmax(fcent, m->small_));
      double xn = 0.75 - 1.27 * fclog;
      double cprlog = prlog - (0.4 + 0.67 * fclog);
      double flog = fclog / (1.0 + (cprlog / (xn - 0.14 * cprlog)) * (cprlog / (xn - 0.14 * cprlog)));
      double fc = pow(10.0, flog);
      pcor = fc * pcor;
    }
    rkft[i] = rkft[i] * pcor;
    rkrt[i] = rkrt[i] * pcor;
  }
  double rkf[ORA_IMAX], rkr[ORA_IMAX];
  for (int i = 0; i < ii; i++) {
    rkft[i] = rkft[i] * ctb[i] * 1.0;
    rkrt[i] = rkrt[i] * ctb[i] * 1.0;
    rkf[i] = 0.0; rkr[i] = 0.0;
    if (m->nu[i][0] != 0) {
      rkf[i] = rkft[i] * ipow(c[m->nunk[i][0] - 1], abs(m->nu[i][0]));
      rkr[i] = rkrt[i] * ipow(c[m->nunk[i][3] - 1], m->nu[i][3]);
      if (m->nunk[i][1] != 0) {
        rkf[i] = rkf[i] * ipow(c[m->nunk[i][1] - 1], abs(m->nu[i][1]));
        if (m->nunk[i][2] != 0) rkf[i] = rkf[i] * ipow(c[m->nunk[i][2] - 1], abs(m->nu[i][2]));
      }
      if (m->nunk[i][4] != 0) {
        rkr[i] = rkr[i] * ipow(c[m->nunk[i][4] - 1], m->nu[i][4]);
        if (m->nunk[i][5] != 0) rkr[i] = rkr[i] * ipow(c[m->nunk[i][5] - 1], m->nu[i][5]);
      }
    }
  }
  for (int k = 0; k < kk; k++) wdot[k] = 0.0;
  for (int n = 0; n < 6; n++)
    for (int i = 0; i < ii; i++) {
      int k = m->nunk[i][n];
      if (k != 0) wdot[k - 1] = wdot[k - 1] + (rkf[i] - rkr[i]) * (double)m->nu[i][n];
    }
}

/* CKCVBS via CKCVMS/CKCPMS: src/chemkin/vode.f:64-226 */
double ora_ckcvbs(const ora_mech *m, double T, const double *Y) {
  double tn[5];
  tn[0] = 1.0;
  for (int n = 2; n <= 5; n++) tn[n - 1] = pow(T, n - 1);
  double cvb = 0.0;
  for (int k = 0; k < m->kk; k++) {
    int L = (T > m->tmid[k]) ? 1 : 0;
    const double *a = m->a[k][L];
    double sum = 0.0;
    for (int n = 0; n < 5; n++) sum = sum + tn[n] * a[n];
    double cpms = m->ru * sum / m->wt[k];
    double cvms = cpms - m->ru / m->wt[k];
    cvb = cvb + Y[k] * cvms;
  }
  return cvb;
}

/* CKHMS: src/chemkin/vode.f:228-285 */
void ora_ckhms(const ora_mech *m, double T, double *hms) {
  double rut = T * m->ru, tn[5];
  tn[0] = 1.0;
  for (int n = 2; n <= 5; n++) tn[n - 1] = pow(T, n - 1) / n;
  for (int k = 0; k < m->kk; k++) {
    int L = (T > m->tmid[k]) ? 1 : 0;
    const double *a = m->a[k][L];
    double sum = 0.0;
    for (int n = 0; n < 5; n++) sum = sum + tn[n] * a[n];
    hms[k] = rut * (sum + a[5] / T) / m->wt[k];
  }
}

/* FUN, constant-volume adiabatic reactor (OMS = 1): src/chemkin/reaction_v.f:69-114.
 * z = (T, Y_1..Y_KK), zp = dz/dt. */
void ora_reactor_rhs(const ora_mech *m, double rho, const double *z, double *zp) {
  double wdot[ORA_KMAX], h[ORA_KMAX];
  double cvb = ora_ckcvbs(m, z[0], z + 1);
  ora_ckwyr(m, rho, z[0], z + 1, wdot);
  ora_ckhms(m, z[0], h);
  double rho_i = 1.0 / rho * 1.0;
  double mru = -m->ru;
  double sum = 0.0;
  for (int k = 0; k < m->kk; k++) {
    zp[k + 1] = wdot[k] * m->wt[k] * rho_i;
    sum = sum + h[k] * wdot[k] * m->wt[k] + mru * z[0] * wdot[k];
  }
  zp[0] = -sum / (rho * cvb) * 1.0;
}

/* one line of the 6th-order central derivative (alpha6), for tests:
 * out[i] = d/dx a at a[i], i = 3..n-4 */
void ora_deriv7_line(int n, const double *a, double dinv, double *out) {
  for (int i = 3; i < n - 3; i++) out[i] = wsum(D6, -3, 3, a, i, 1) * dinv;
}
