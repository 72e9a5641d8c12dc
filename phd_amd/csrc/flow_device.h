// flow_device.h -- per-cell FP64 device functions of the flow solver.
//
// Every function keeps the operation order of the reference Fortran (cited
// per function) and is compiled with -ffp-contract=off, so that the results
// are the reference's own roundings, not a re-associated variant.
#pragma once

#include "izem_common.h"

extern __constant__ IzThd c_thd;

// Wmix_i_scalar: devel-izem/src/module_thermodynamics.f90:739-756
__device__ inline double d_wmix_scalar(const double *Y) {
  double w = 0.0;
#pragma unroll
  for (int l = 0; l < IZ_NRV; l++) w = w + Y[l] * c_thd.Wc_i[l];
  return w;
}

// cp_scalar: module_thermodynamics.f90:767-862
__device__ inline double d_cp_scalar(double temp, const double *Y) {
  double cp = 0.0;
  for (int is = 0; is < IZ_NRV; is++) {
    double tk1 = temp * c_thd.Tinf, tk2 = tk1 * tk1, tk3 = tk2 * tk1, tk4 = tk3 * tk1;
    const double *hl = c_thd.HeatCap[is][0], *hh = c_thd.HeatCap[is][1];
    double fac = Y[is] * c_thd.gam2 * c_thd.Wc_i[is];
    if (tk1 < c_thd.Tpolymin[is]) {
      tk1 = c_thd.Tpolymin[is]; tk2 = tk1 * tk1; tk3 = tk2 * tk1; tk4 = tk3 * tk1;
      double cp0 = hl[0] + hl[1] * tk1 + hl[2] * tk2 + hl[3] * tk3 + hl[4] * tk4;
      double dcp0 = hl[1] + hl[2] * tk1 * 2.0 + hl[3] * tk2 * 3.0 + hl[4] * tk3 * 4.0;
      cp = cp + fac * (dcp0 * (temp * c_thd.Tinf - c_thd.Tpolymin[is]) + cp0);
    } else if (tk1 < c_thd.Tpolylim[is]) {
      cp = cp + fac * (hl[0] + hl[1] * tk1 + hl[2] * tk2 + hl[3] * tk3 + hl[4] * tk4);
    } else if (tk1 < c_thd.Tpolymax[is]) {
      cp = cp + fac * (hh[0] + hh[1] * tk1 + hh[2] * tk2 + hh[3] * tk3 + hh[4] * tk4);
    } else {
      tk1 = c_thd.Tpolymax[is]; tk2 = tk1 * tk1; tk3 = tk2 * tk1; tk4 = tk3 * tk1;
      double cp0 = hh[0] + hh[1] * tk1 + hh[2] * tk2 + hh[3] * tk3 + hh[4] * tk4;
      double dcp0 = hh[1] + hh[2] * tk1 * 2.0 + hh[3] * tk2 * 3.0 + hh[4] * tk3 * 4.0;
      cp = cp + fac * (dcp0 * (temp * c_thd.Tinf - c_thd.Tpolymax[is]) + cp0);
    }
  }
  return cp;
}

__device__ inline double d_hpoly(const double *h, double t1, double t2, double t3, double t4,
                                 double t5, double onethird) {
  return h[0] * t1 + h[1] * t2 * 0.50 + h[2] * t3 * onethird + h[3] * t4 * 0.25 +
         h[4] * t5 * 0.20 + h[5];
}
__device__ inline double d_dhpoly(const double *h, double t1, double t2, double t3, double t4) {
  return h[0] + h[1] * t1 + h[2] * t2 + h[3] * t3 + h[4] * t4;
}

// Per-species enthalpy and cp contribution at nondimensional T, with the
// four-branch NASA-7 evaluation of newton / upd_prim_var_woT / ha_scalar
// (module_thermodynamics.f90:327-498, 522-689, 873-978).
__device__ inline void d_species_h_cp(int is, double T, double Ya, double &h, double &cpc) {
  const double Tinf_i = 1.0 / c_thd.Tinf, onethird = 1.0 / 3.0;
  double t1 = T * c_thd.Tinf, t2 = t1 * t1, t3 = t2 * t1, t4 = t3 * t1, t5 = t4 * t1;
  const double *hl = c_thd.HeatCap[is][0], *hh = c_thd.HeatCap[is][1];
  double fh = c_thd.Wc_i[is] * Tinf_i;
  double fc = Ya * c_thd.gam2 * c_thd.Wc_i[is];
  double hv = 0.0;
  if (t1 < c_thd.Tpolymin[is]) {
    t1 = c_thd.Tpolymin[is]; t2 = t1 * t1; t3 = t2 * t1; t4 = t3 * t1; t5 = t4 * t1;
    double ha0 = d_hpoly(hl, t1, t2, t3, t4, t5, onethird);
    double dha0 = d_dhpoly(hl, t1, t2, t3, t4);
    hv = hv + fh * (dha0 * (T * c_thd.Tinf - c_thd.Tpolymin[is]) + ha0);
    double dcp0 = hl[1] + hl[2] * t1 * 2.0 + hl[3] * t2 * 3.0 + hl[4] * t3 * 4.0;
    cpc = fc * (dcp0 * (T * c_thd.Tinf - c_thd.Tpolymin[is]) + dha0);
  } else if (t1 < c_thd.Tpolylim[is]) {
    hv = hv + fh * d_hpoly(hl, t1, t2, t3, t4, t5, onethird);
    cpc = fc * d_dhpoly(hl, t1, t2, t3, t4);
  } else if (t1 < c_thd.Tpolymax[is]) {
    hv = hv + fh * d_hpoly(hh, t1, t2, t3, t4, t5, onethird);
    cpc = fc * d_dhpoly(hh, t1, t2, t3, t4);
  } else {
    t1 = c_thd.Tpolymax[is]; t2 = t1 * t1; t3 = t2 * t1; t4 = t3 * t1; t5 = t4 * t1;
    double ha0 = d_hpoly(hh, t1, t2, t3, t4, t5, onethird);
    double dha0 = d_dhpoly(hh, t1, t2, t3, t4);
    hv = hv + fh * (dha0 * (T * c_thd.Tinf - c_thd.Tpolymax[is]) + ha0);
    double dcp0 = hh[1] + hh[2] * t1 * 2.0 + hh[3] * t2 * 3.0 + hh[4] * t3 * 4.0;
    cpc = fc * (dcp0 * (T * c_thd.Tinf - c_thd.Tpolymax[is]) + dha0);
  }
  h = hv;
}

// ha_scalar: module_thermodynamics.f90:873-978
__device__ inline void d_ha_scalar(double temp, double *ha) {
#pragma unroll
  for (int is = 0; is < IZ_NRV; is++) {
    double h, cpc;
    d_species_h_cp(is, temp, 0.0, h, cpc);
    ha[is] = h;
  }
}

// newton (one cell): module_thermodynamics.f90:271-508.  Returns DF and the
// cp / ha of this evaluation.
__device__ inline double d_newton_cell(const double *vc, double T, double &cp_out,
                                       double *ha_out) {
  const double gam2_i = 1.0 / c_thd.gam2;
  double rho_i = 1.0 / vc[0];
  double F = -vc[4] * rho_i + 0.5 * (vc[1] * vc[1] + vc[2] * vc[2] + vc[3] * vc[3]) *
                                  (rho_i * rho_i);
  double Ya[IZ_NRV];
#pragma unroll
  for (int is = 0; is < IZ_NRV; is++) Ya[is] = vc[IZ_NIV + is] * rho_i;
  double cp = 0.0;
#pragma unroll
  for (int is = 0; is < IZ_NRV; is++) {
    double wrk1 = -T;
    double h, cpc;
    d_species_h_cp(is, T, Ya[is], h, cpc);
    cp = cp + cpc;
    ha_out[is] = h;
    F = F + Ya[is] * (h + wrk1 * c_thd.Wc_i[is]);
  }
  double wrk = 0.0;
#pragma unroll
  for (int is = 0; is < IZ_NRV; is++) wrk = wrk + Ya[is] * c_thd.Wc_i[is];
  cp_out = cp;
  return F / (wrk - cp * gam2_i);
}

// ------------------------------------------------------------------ WENO
// wenopar / wenorec / wenorec_nw, ng = 3: module_weno.f90:114-190, 236-527
#define IZ_Z1 (13.0 / 12.0)
#define IZ_Z2 (1.0 / 4.0)
#define IZ_EPS_WENO 1.0e-10

__device__ inline void d_betas(const double *q, int i, double *beta) {
  // q is 1-based
  double a0 = q[i] + (-2.0) * q[i + 1] + q[i + 2], b0 = 3.0 * q[i] + (-4.0) * q[i + 1] + q[i + 2];
  beta[0] = IZ_Z1 * a0 * a0 + IZ_Z2 * b0 * b0;
  double a1 = q[i - 1] + (-2.0) * q[i] + q[i + 1], b1 = q[i - 1] + (-1.0) * q[i + 1];
  beta[1] = IZ_Z1 * a1 * a1 + IZ_Z2 * b1 * b1;
  double a2 = q[i - 2] + (-2.0) * q[i - 1] + q[i],
         b2 = q[i - 2] + (-4.0) * q[i - 1] + 3.0 * q[i];
  beta[2] = IZ_Z1 * a2 * a2 + IZ_Z2 * b2 * b2;
}

__device__ inline void d_wenorec(const double *vp0, const double *vm0, double &vminus,
                                 double &vplus) {
  const double c[4][3] = {{11.0 / 6.0, -7.0 / 6.0, 1.0 / 3.0},
                          {1.0 / 3.0, 5.0 / 6.0, -1.0 / 6.0},
                          {-1.0 / 6.0, 5.0 / 6.0, 1.0 / 3.0},
                          {1.0 / 3.0, -7.0 / 6.0, 11.0 / 6.0}};
  const double d[3] = {3.0 / 10.0, 3.0 / 5.0, 1.0 / 10.0};
  const double *vp = vp0 - 1, *vm = vm0 - 1;
  double vpr[3], beta[3], alfa[3], sum;
  int i = 3;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    vpr[r] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) vpr[r] = vpr[r] + c[r + 1][j] * vp[i - r + j];
  }
  d_betas(vp, i, beta);
  sum = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    alfa[r] = d[r] / ((IZ_EPS_WENO + beta[r]) * (IZ_EPS_WENO + beta[r]));
    sum = sum + alfa[r];
  }
  sum = 1.0 / sum;
  double out = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++) out = out + (alfa[r] * sum) * vpr[r];
  vminus = out;
  i = 4;
  double vmr[3];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    vmr[r] = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) vmr[r] = vmr[r] + c[r][j] * vm[i - r + j];
  }
  d_betas(vm, i, beta);
  sum = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    alfa[r] = d[2 - r] / ((IZ_EPS_WENO + beta[r]) * (IZ_EPS_WENO + beta[r]));
    sum = sum + alfa[r];
  }
  sum = 1.0 / sum;
  out = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++) out = out + (alfa[r] * sum) * vmr[r];
  vplus = out;
}

__device__ inline void d_wenorec_nw(const double *vp0, const double *vm0, double &vminus,
                                    double &vplus) {
  const double *vp = vp0 - 1, *vm = vm0 - 1;
  const double denom = 1.0 / 60.0;
  vminus = (-3.0 * vp[1] + 27.0 * vp[2] + 47.0 * vp[3] - 13.0 * vp[4] + 2.0 * vp[5]) * denom;
  vplus = (-3.0 * vm[6] + 27.0 * vm[5] + 47.0 * vm[4] - 13.0 * vm[3] + 2.0 * vm[2]) * denom;
}

// WENO7 (ng = 4): the candidate stencils' coefficients and the ideal weights
// (4, 18, 12, 1)/35 mirrored for the '+' state (module_weno.f90:71-111), and
// Balsara and Shu's r = 4 smoothness indicators (:154-172, 330-353): for the
// four values a..d of a candidate,
//   beta = a (k0 a + k1 b + k2 c + k3 d) + b (k4 b + k5 c + k6 d)
//        + c (k7 c + k8 d) + d (k9 d).
__device__ inline double d_beta7(int r, double a, double b, double c, double d) {
  const double K[4][10] = {
      {2107.0, -9402.0, 7042.0, -1854.0, 11003.0, -17246.0, 4642.0, 7043.0, -3882.0, 547.0},
      {547.0, -2522.0, 1922.0, -494.0, 3443.0, -5966.0, 1602.0, 2843.0, -1642.0, 267.0},
      {267.0, -1642.0, 1602.0, -494.0, 2843.0, -5966.0, 1922.0, 3443.0, -2522.0, 547.0},
      {547.0, -3882.0, 4642.0, -1854.0, 7043.0, -17246.0, 7042.0, 11003.0, -9402.0, 2107.0}};
  const double *k = K[r];
  return a * (k[0] * a + k[1] * b + k[2] * c + k[3] * d) + b * (k[4] * b + k[5] * c + k[6] * d) +
         c * (k[7] * c + k[8] * d) + d * (k[9] * d);
}

// one side of the WENO7 reconstruction: q 1-based, i the centre, row0 the
// coefficient row of candidate 0 (1 for the '-' state, 0 for '+')
__device__ inline double d_weno7_side(const double *q, int i, int row0, const double *dw) {
  const double c[5][4] = {{25.0 / 12.0, -23.0 / 12.0, 13.0 / 12.0, -1.0 / 4.0},
                          {1.0 / 4.0, 13.0 / 12.0, -5.0 / 12.0, 1.0 / 12.0},
                          {-1.0 / 12.0, 7.0 / 12.0, 7.0 / 12.0, -1.0 / 12.0},
                          {1.0 / 12.0, -5.0 / 12.0, 13.0 / 12.0, 1.0 / 4.0},
                          {-1.0 / 4.0, 13.0 / 12.0, -23.0 / 12.0, 25.0 / 12.0}};
  double vr[4], alfa[4], sum = 0.0, out = 0.0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    vr[r] = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) vr[r] = vr[r] + c[r + row0][j] * q[i - r + j];
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const double beta = d_beta7(r, q[i - r], q[i - r + 1], q[i - r + 2], q[i - r + 3]);
    alfa[r] = dw[r] / ((IZ_EPS_WENO + beta) * (IZ_EPS_WENO + beta));
    sum = sum + alfa[r];
  }
  sum = 1.0 / sum;
#pragma unroll
  for (int r = 0; r < 4; r++) out = out + alfa[r] * sum * vr[r];
  return out;
}

__device__ inline void d_wenorec7(const double *vp0, const double *vm0, double &vminus,
                                  double &vplus) {
  const double d[4] = {4.0 / 35.0, 18.0 / 35.0, 12.0 / 35.0, 1.0 / 35.0};
  const double dt[4] = {1.0 / 35.0, 12.0 / 35.0, 18.0 / 35.0, 4.0 / 35.0};
  vminus = d_weno7_side(vp0 - 1, 4, 1, d);
  vplus = d_weno7_side(vm0 - 1, 5, 0, dt);
}

// wenorec_nw, ng = 4 (module_weno.f90:246-256): the 7th-order linear face value
__device__ inline void d_wenorec7_nw(const double *vp0, const double *vm0, double &vminus,
                                     double &vplus) {
  const double *vp = vp0 - 1, *vm = vm0 - 1;
  const double denom = 1.0 / 420.0;
  vminus = (-3.0 * vp[1] + 25.0 * vp[2] - 101.0 * vp[3] + 319.0 * vp[4] + 214.0 * vp[5] -
            38.0 * vp[6] + 4.0 * vp[7]) * denom;
  vplus = (-3.0 * vm[8] + 25.0 * vm[7] - 101.0 * vm[6] + 319.0 * vm[5] + 214.0 * vm[4] -
           38.0 * vm[3] + 4.0 * vm[2]) * denom;
}

// the reconstruction of the build's order on IZ_S stencil values
__device__ inline void d_weno_face(bool js, const double *gp, const double *gm, double &gl,
                                   double &gr) {
#if IZ_NG == 4
  if (js) d_wenorec7(gp, gm, gl, gr);
  else d_wenorec7_nw(gp, gm, gl, gr);
#else
  if (js) d_wenorec(gp, gm, gl, gr);
  else d_wenorec_nw(gp, gm, gl, gr);
#endif
}

// 6th-order central first derivative (alpha6), periodic branch of
// dx/dy/dz, d?_fixed1, d?_fixed2: module_deriv.f90:37-40
__device__ inline double d_deriv7(const double *__restrict__ a, long c, long s, double dinv) {
  double dd = 0.0;
  dd = dd + (-1.0 / 60.0) * a[c - 3 * s];
  dd = dd + (9.0 / 60.0) * a[c - 2 * s];
  dd = dd + (-45.0 / 60.0) * a[c - s];
  dd = dd + (45.0 / 60.0) * a[c + s];
  dd = dd + (-9.0 / 60.0) * a[c + 2 * s];
  dd = dd + (1.0 / 60.0) * a[c + 3 * s];
  return dd * dinv;
}

// First derivative at position p (0 .. n-1) of a line whose values at
// offsets l from p are f(l), next to physical faces (lp: the face before
// position 0, rp: the one after n-1): the three points nearest a physical
// face take 4th-order closures, one-sided at the face point, upwind at the
// next, central at the third, the right-hand ones negated, and for ng = 4
// 6th-order central at the fourth (module_deriv.f90:20-78, 268-469);
// elsewhere the central stencil of the build's order.  Weights summed in
// increasing offset.
// central first derivative of the build's order at offsets f(l): 6th order
// (alpha6) for ng = 3, 8th order (alph8) for ng = 4 (module_deriv.f90:37-78)
template <typename Fn>
__device__ inline double d_deriv_central(Fn f, double dinv) {
  double dd = 0.0;
#if IZ_NG == 4
  dd = dd + (3.0 / 840.0) * f(-4);
  dd = dd + (-32.0 / 840.0) * f(-3);
  dd = dd + (168.0 / 840.0) * f(-2);
  dd = dd + (-672.0 / 840.0) * f(-1);
  dd = dd + (672.0 / 840.0) * f(1);
  dd = dd + (-168.0 / 840.0) * f(2);
  dd = dd + (32.0 / 840.0) * f(3);
  dd = dd + (-3.0 / 840.0) * f(4);
#else
  dd = dd + (-1.0 / 60.0) * f(-3);
  dd = dd + (9.0 / 60.0) * f(-2);
  dd = dd + (-45.0 / 60.0) * f(-1);
  dd = dd + (45.0 / 60.0) * f(1);
  dd = dd + (-9.0 / 60.0) * f(2);
  dd = dd + (1.0 / 60.0) * f(3);
#endif
  return dd * dinv;
}

template <typename Fn>
__device__ inline double d_deriv6(Fn f, double dinv) {
  double dd = 0.0;
  dd = dd + (-1.0 / 60.0) * f(-3);
  dd = dd + (9.0 / 60.0) * f(-2);
  dd = dd + (-45.0 / 60.0) * f(-1);
  dd = dd + (45.0 / 60.0) * f(1);
  dd = dd + (-9.0 / 60.0) * f(2);
  dd = dd + (1.0 / 60.0) * f(3);
  return dd * dinv;
}

template <typename Fn>
__device__ inline double d_deriv_line(Fn f, double dinv, int p, int n, bool lp, bool rp) {
  double dd = 0.0;
  if (rp && p >= n - IZ_NG) {
    if (IZ_NG == 4 && p == n - 4) return d_deriv6(f, dinv);
    if (p == n - 1) {
      dd = dd + (-3.0 / 12.0) * f(-4);
      dd = dd + (16.0 / 12.0) * f(-3);
      dd = dd + (-36.0 / 12.0) * f(-2);
      dd = dd + (48.0 / 12.0) * f(-1);
      dd = dd + (-25.0 / 12.0) * f(0);
      return -(dd * dinv);
    }
    if (p == n - 2) {
      dd = dd + (1.0 / 12.0) * f(-3);
      dd = dd + (-6.0 / 12.0) * f(-2);
      dd = dd + (18.0 / 12.0) * f(-1);
      dd = dd + (-10.0 / 12.0) * f(0);
      dd = dd + (-3.0 / 12.0) * f(1);
      return -(dd * dinv);
    }
    dd = dd + (1.0 / 12.0) * f(-2);
    dd = dd + (-8.0 / 12.0) * f(-1);
    dd = dd + (8.0 / 12.0) * f(1);
    dd = dd + (-1.0 / 12.0) * f(2);
    return dd * dinv;
  }
  if (lp && p < IZ_NG) {
    if (IZ_NG == 4 && p == 3) return d_deriv6(f, dinv);
    if (p == 0) {
      dd = dd + (-25.0 / 12.0) * f(0);
      dd = dd + (48.0 / 12.0) * f(1);
      dd = dd + (-36.0 / 12.0) * f(2);
      dd = dd + (16.0 / 12.0) * f(3);
      dd = dd + (-3.0 / 12.0) * f(4);
    } else if (p == 1) {
      dd = dd + (-3.0 / 12.0) * f(-1);
      dd = dd + (-10.0 / 12.0) * f(0);
      dd = dd + (18.0 / 12.0) * f(1);
      dd = dd + (-6.0 / 12.0) * f(2);
      dd = dd + (1.0 / 12.0) * f(3);
    } else {
      dd = dd + (1.0 / 12.0) * f(-2);
      dd = dd + (-8.0 / 12.0) * f(-1);
      dd = dd + (8.0 / 12.0) * f(1);
      dd = dd + (-1.0 / 12.0) * f(2);
    }
    return dd * dinv;
  }
  return d_deriv_central(f, dinv);
}

// derivative along direction dir of the array a at cell c, interior
// position p along dir; PHYS = false: every face periodic (d_deriv7)
template <bool PHYS>
__device__ inline double d_deriv_dir(const IzDims &d, const double *__restrict__ a, long c, int dir,
                                     int p, double dinv) {
  const long s = dir == 0 ? 1 : (dir == 1 ? d.NX : d.NX * d.NY);
#if IZ_NG == 3
  if (!PHYS) return d_deriv7(a, c, s, dinv);
#else
  if (!PHYS) return d_deriv_central([&](int l) { return a[c + l * s]; }, dinv);
#endif
  const int n = dir == 0 ? d.nx : (dir == 1 ? d.ny : d.nz);
  const bool lp = (d.phys >> (2 * dir)) & 1, rp = (d.phys >> (2 * dir + 1)) & 1;
  return d_deriv_line([&](int l) { return a[c + l * s]; }, dinv, p, n, lp, rp);
}
