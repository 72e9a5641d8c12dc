// chem_device.h -- CHEMKIN-form kinetics and the BDF integrator of one
// reactor cell.  Device code; the functions are also host-compilable (IZ_HD)
// so that the integrator can be exercised on a CPU build.
#pragma once

#include <math.h>

#ifndef IZ_HD
#define IZ_HD __device__
#endif
#ifndef IZ_CMECH
#define IZ_CMECH c_mech
#define IZ_CBDF c_bdf
#endif

#include "chem_tables.h"

// --------------------------------------------------------------- kinetics
// CKSMH (vode.f:1474-1535): S/R - H/RT per species
IZ_HD inline void d_smh(double T, double *smh) {
  double lt = log(T);
  double t2 = T * T, t3 = t2 * T, t4 = t3 * T;
  double tn0 = lt - 1.0, tn1 = T / 2.0, tn2 = t2 / 6.0, tn3 = t3 / 12.0, tn4 = t4 / 20.0;
  double Ti = 1.0 / T;
  for (int k = 0; k < IZ_CMECH.kk; k++) {
    const double *a = IZ_CMECH.a[k][(T > IZ_CMECH.tmid[k]) ? 1 : 0];
    smh[k] = tn0 * a[0] + tn1 * a[1] + tn2 * a[2] + tn3 * a[3] + tn4 * a[4] + a[6] - a[5] * Ti;
  }
}

IZ_HD inline double d_ipow(double x, int n) {
  double r = 1.0;
  for (int q = 0; q < n; q++) r *= x;
  return r;
}

// CKWYR = CKRATT + CKYTCR + CKRATX + stoichiometric sum (vode.f:935-1314,
// 1797-1904, 3171-3221); mol/cm^3/s
IZ_HD inline void d_ckwyr(double rho, double T, const double *Y, double *wdot) {
  const int ii = IZ_CMECH.ii, kk = IZ_CMECH.kk;
  double smh[CK_KMAX], c[CK_KMAX];
  d_smh(T, smh);
  const double alogt = log(T), Ti = 1.0 / T;
  const double pfac = 1.01325e6 / (8.314510e7 * T);
  double ctot = 0.0;
  for (int k = 0; k < kk; k++) {
    c[k] = rho * Y[k] / IZ_CMECH.wt[k];
    ctot += c[k];
  }
  for (int k = 0; k < kk; k++) wdot[k] = 0.0;
  for (int i = 0; i < ii; i++) {
    double kf = IZ_CMECH.A[i] * exp(IZ_CMECH.b[i] * alogt - IZ_CMECH.EoR[i] * Ti);
    double kr = 0.0;
    if (IZ_CMECH.reversible[i] == 2) {
      // explicit reverse parameters (CKRATT, vode.f:1028-1032)
      const double *rp = IZ_CMECH.rpar[i];
      kr = rp[0] * exp(rp[1] * alogt - rp[2] * Ti);
    } else if (IZ_CMECH.reversible[i]) {
      double s = 0.0;
      int nus = 0;
      for (int n = 0; n < 6; n++) {
        int k = IZ_CMECH.nunk[i][n];
        if (k) { s += (double)IZ_CMECH.nu[i][n] * smh[k - 1]; nus += IZ_CMECH.nu[i][n]; }
      }
      double eqk = exp(fmin(s, 690.77552789821368));  // EXPARG = ln(1e300)
      eqk *= (nus >= 0) ? d_ipow(pfac, nus) : 1.0 / d_ipow(pfac, -nus);
      kr = kf / fmax(eqk, 1e-300);
    }
    double ctb = 1.0;
    if (IZ_CMECH.ntb[i] >= 0) {
      ctb = ctot;
      for (int l = 0; l < IZ_CMECH.ntb[i]; l++)
        ctb += (IZ_CMECH.tb_eff[i][l] - 1.0) * c[IZ_CMECH.tb_k[i][l] - 1];
    }
    if (IZ_CMECH.fal[i]) {
      const double *fp = IZ_CMECH.fpar[i];
      double rklow = fp[0] * exp(fp[1] * alogt - fp[2] * Ti);
      // [M] is the third-body sum, or one species' concentration (KFAL)
      const int kf_m = IZ_CMECH.kfal[i];
      double pr = rklow * (kf_m ? c[kf_m - 1] : ctb) / kf;
      ctb = 1.0;
      double pcor = pr / (1.0 + pr);
      if (IZ_CMECH.fal[i] == 2) {
        // SRI form (CKRATX, vode.f:1211-1216)
        double prlog = log10(fmax(pr, 1e-300));
        double xp = 1.0 / (1.0 + prlog * prlog);
        double fc = pow(fp[3] * exp(-fp[4] / T) + exp(-T / fp[5]), xp) * fp[6] * pow(T, fp[7]);
        pcor *= fc;
      } else if (IZ_CMECH.fal[i] > 2) {
        double prlog = log10(fmax(pr, 1e-300));
        double fcent = (1.0 - fp[3]) * exp(-T / fp[4]) + fp[3] * exp(-T / fp[5]);
        if (IZ_CMECH.fal[i] == 4) fcent += exp(-fp[6] * Ti);
        double fclog = log10(fmax(fcent, 1e-300));
        double xn = 0.75 - 1.27 * fclog;
        double cprlog = prlog - (0.4 + 0.67 * fclog);
        double q = cprlog / (xn - 0.14 * cprlog);
        double flog = fclog / (1.0 + q * q);
        pcor *= exp(flog * 2.302585092994046);  // 10**flog
      }
      kf *= pcor;
      kr *= pcor;
    }
    kf *= ctb;
    kr *= ctb;
    double rf = kf, rr = kr;
    for (int n = 0; n < 3; n++) {
      int k = IZ_CMECH.nunk[i][n];
      if (k) rf *= d_ipow(c[k - 1], -IZ_CMECH.nu[i][n]);
    }
    for (int n = 3; n < 6; n++) {
      int k = IZ_CMECH.nunk[i][n];
      if (k) rr *= d_ipow(c[k - 1], IZ_CMECH.nu[i][n]);
    }
    double rop = rf - rr;
    for (int n = 0; n < 6; n++) {
      int k = IZ_CMECH.nunk[i][n];
      if (k) wdot[k - 1] += rop * (double)IZ_CMECH.nu[i][n];
    }
  }
}

// FUN (reaction_v.f:69-114) with CKCVBS (vode.f:64) and CKHMS (vode.f:228):
// constant-volume adiabatic reactor, z = (T, Y_1..Y_kk)
IZ_HD inline void d_reactor_rhs(double rho, const double *z, double *zp) {
  const int kk = IZ_CMECH.kk;
  const double T = z[0], ru = 8.314510e7;
  double wdot[CK_KMAX];
  d_ckwyr(rho, T, z + 1, wdot);
  double t2 = T * T, t3 = t2 * T, t4 = t3 * T;
  double cvb = 0.0, sum = 0.0;
  const double rho_i = 1.0 / rho;
  for (int k = 0; k < kk; k++) {
    const double *a = IZ_CMECH.a[k][(T > IZ_CMECH.tmid[k]) ? 1 : 0];
    const double w = IZ_CMECH.wt[k];
    double cp = ru * (a[0] + a[1] * T + a[2] * t2 + a[3] * t3 + a[4] * t4) / w;
    cvb += z[1 + k] * (cp - ru / w);
    double h = ru * T * (a[0] + a[1] * T / 2.0 + a[2] * t2 / 3.0 + a[3] * t3 / 4.0 +
                         a[4] * t4 / 5.0 + a[5] / T) / w;
    zp[1 + k] = wdot[k] * w * rho_i;
    sum += h * wdot[k] * w - ru * T * wdot[k];
  }
  zp[0] = -sum / (rho * cvb);
}

// ------------------------------------------------------------ BDF solver
#define N_ CK_NEQ

// weighted RMS norm with the reciprocal error weights w = 1 / (rtol |y| + atol)
// that DVODE keeps in EWT (vode.f:5624; DVNORM, vode.f:4260-4282)
IZ_HD inline double d_wrms(const double *v, const double *w) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < N_; i++) { double x = v[i] * w[i]; s += x * x; }
  return sqrt(s / N_);
}

// x^e for x >= 0: the step-ratio powers of the step and order selection
// (DVSTEP's (BIAS*DSM)**(1/L), vode.f:6516, 6693-6708)
IZ_HD inline double d_powpos(double x, double e) { return exp(e * log(x)); }

// LU with partial pivoting in LINPACK's form (dgefa / dgesl, vode.f:7033,
// 7630): step k swaps rows k and p over columns k..N-1 only, so the
// multipliers of the earlier steps stay in the rows where the solve
// applies them, interleaved with the same swaps.  Every index is a
// compile-time constant, so that the factors can stay in registers: the
// row swap is a run of selects, skipped when no lane of the wave pivots.
// The diagonal keeps the reciprocal of each pivot (0 for a missing one), so
// the back substitution multiplies where dgesl divides.
IZ_HD inline bool d_lu(double (&A)[N_][N_], int (&piv)[N_]) {
  bool ok = true;
#pragma unroll
  for (int k = 0; k < N_; k++) {
    int p = k;
    double amax = fabs(A[k][k]);
#pragma unroll
    for (int i = k + 1; i < N_; i++) {
      const double a = fabs(A[i][k]);
      p = (a > amax) ? i : p;
      amax = (a > amax) ? a : amax;
    }
    piv[k] = p;
    if (IZ_ANY(p != k)) {
#pragma unroll
      for (int j = k; j < N_; j++) {
        const double akj = A[k][j];
        double apj = akj;
#pragma unroll
        for (int i = k + 1; i < N_; i++) {
          apj = (p == i) ? A[i][j] : apj;
          A[i][j] = (p == i) ? akj : A[i][j];
        }
        A[k][j] = apj;
      }
    }
    if (amax == 0.0) ok = false;  // dgefa's INFO = k: column k has no pivot
    const double inv = (amax == 0.0) ? 0.0 : 1.0 / A[k][k];
    A[k][k] = inv;
#pragma unroll
    for (int i = k + 1; i < N_; i++) {
      const double m = A[i][k] * inv;
      A[i][k] = m;
#pragma unroll
      for (int j = k + 1; j < N_; j++) A[i][j] -= m * A[k][j];
    }
  }
  return ok;
}
IZ_HD inline void d_lusolve(const double (&A)[N_][N_], const int (&piv)[N_], double (&x)[N_]) {
#pragma unroll
  for (int k = 0; k < N_; k++) {
    const int p = piv[k];
    if (IZ_ANY(p != k)) {
      const double xk = x[k];
      double xp = xk;
#pragma unroll
      for (int i = k + 1; i < N_; i++) {
        xp = (p == i) ? x[i] : xp;
        x[i] = (p == i) ? xk : x[i];
      }
      x[k] = xp;
    }
    const double t = x[k];
#pragma unroll
    for (int i = k + 1; i < N_; i++) x[i] -= A[i][k] * t;
  }
#pragma unroll
  for (int k = N_ - 1; k >= 0; k--) {
    double s = x[k];
#pragma unroll
    for (int j = k + 1; j < N_; j++) s -= A[k][j] * x[j];
    x[k] = s * A[k][k];
  }
}

// Nordsieck history of one cell, z(j, n) for column j = 0..5 and component
// n.  ZArr holds it in the lane's own (private) memory; ZLds holds it in
// LDS with one 64-lane wave's worth of stride between elements, so that
// element (j, n) of lane l sits at base[(j * N_ + n) * 64 + l] and a wave
// reads a column component as one contiguous 512-byte row.
struct ZArr {
  double a[6][N_];
  IZ_MHD inline double &operator()(int j, int n) { return a[j][n]; }
};
struct ZLds {
  double *p;  // base + lane
  IZ_MHD inline double &operator()(int j, int n) const { return p[(j * N_ + n) * 64]; }
};

template <class Z>
IZ_HD inline double d_wrms_col(Z &z, int j, const double *w) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < N_; i++) { double x = z(j, i) * w[i]; s += x * x; }
  return sqrt(s / N_);
}

// Pascal-triangle update of the Nordsieck history (sign +1: predict to t+h,
// sign -1: retract); sweeps jb = 1..q over columns q-jb..q-1 in ascending
// order, which makes the -1 sweep the exact inverse of the +1 sweep.
template <class Z>
IZ_HD inline void d_nordsieck_shift(Z &z, int q, double sign) {
  for (int jb = 1; jb <= q; jb++)
    for (int i = q - jb; i <= q - 1; i++)
#pragma unroll
      for (int n = 0; n < N_; n++) z(i, n) += sign * z(i + 1, n);
}

// rescale the Nordsieck history for a step change h -> rh*h
template <class Z>
IZ_HD inline void d_nordsieck_rescale(Z &z, int q, double rh) {
  double f = rh;
  for (int j = 1; j <= q; j++) {
#pragma unroll
    for (int n = 0; n < N_; n++) z(j, n) *= f;
    f *= rh;
  }
}

struct BdfResult {
  int nst, nfe, nje, ok;
  int nsol;  // LU back-substitutions (one per corrector iteration)
};

// finite-difference increment of Jacobian column j: DVJAC's
// max(srur |y_j|, r0 / w_j) (vode.f:3739, w_j = 1 / (rtol |y_j| + atol)),
// with a floor of srur * 1e-6 on the mass-fraction columns.  For a species
// that is absent (y_j = 0) DVJAC's increment is r0 * atol, of order
// 1e-20 here, and the temperature row of that column, a difference of two
// evaluations of dT/dt ~ 1e8 K/s, is then a few ulps over 1e-20: noise of
// 1e8..1e10 in an entry whose true value is of that size.  The Newton
// iteration converges or fails with that noise from one build to the next;
// with the floor the entries are accurate to about 1e-3.  The species
// rates are polynomial in the concentrations, so the larger increment
// costs no truncation accuracy that matters to a chord Newton iteration.
IZ_HD inline double d_fd_increment(int j, double yj, double wj, double r0) {
  const double srur = 1.4901161193847656e-08;
  double r = fmax(srur * fabs(yj), r0 / wj);
  return j == 0 ? r : fmax(r, srur * 1.0e-6);
}

// Integrate the reactor of one cell from 0 to tout (ITASK = 1: step past tout
// and interpolate the Nordsieck history back to it).
// Kinetics policies: rhs(rho, y, f) evaluates the reactor right-hand side;
// rhs_c(rho, y, f, c) does the same and keeps in c what the Jacobian at the
// same y can reuse; jac_c(rho, y, f0, c, hl0, w, r0, P), given f0 = f(y) and
// the c of that evaluation, writes P = -hl0 * J with a finite-difference
// Jacobian (DVJAC with MF = 22, vode.f:3728-3770).  None of them writes y.
struct KinTables {
  struct Cache {};
  IZ_MHD inline void rhs(double rho, const double *y, double *f) const {
    d_reactor_rhs(rho, y, f);
  }
  IZ_MHD inline void rhs_c(double rho, const double *y, double *f, Cache &) const {
    d_reactor_rhs(rho, y, f);
  }
  IZ_MHD inline void jac_c(double rho, const double *y, const double *f0, const Cache &,
                           double hl0, const double *w, double r0, double (*P)[N_]) const {
    for (int j = 0; j < N_; j++) {
      double yl[N_], ftmp[N_];
#pragma unroll
      for (int n = 0; n < N_; n++) yl[n] = y[n];
      const double yj = yl[j];
      const double r = d_fd_increment(j, yj, w[j], r0);
      yl[j] = yj + r;
      d_reactor_rhs(rho, yl, ftmp);
      const double ri = 1.0 / r;
#pragma unroll
      for (int i = 0; i < N_; i++) P[i][j] = -hl0 * (ftmp[i] - f0[i]) * ri;
    }
  }
};

// Jacobian and iteration-matrix reuse follow DVODE (DVNLSD / DVJAC,
// vode.f:4032-4259, 3596-3780): the finite-difference Jacobian J is kept,
// and the iteration matrix P = I - h l0 J is refactored from it whenever
// h l0 has changed by more than CCMAX = 0.3 or MSBP = 20 steps have
// passed; J itself is re-evaluated only at the first step, after jac_age
// (DVODE's MSBJ, default 50) steps, or after a corrector failure that a
// new h l0 cannot explain (DRC < CCMXJ = 0.2) or that a current J did not
// prevent.
// J is the caller's storage for the saved Jacobian.  The first step of
// every integration evaluates a new one before any read, so the caller
// initialises J once (the kernels once per lane, not once per cell) and
// may pass the storage of a previous integration.
template <class Kin, class Z>
IZ_HD IZ_FORCEINLINE BdfResult d_bdf_integrate(const Kin &kin, Z &z, double (&J)[N_][N_],
                                               double rho, double *y, double tout, double rtol,
                                               double atol, int jac_age) {
  const double uround = 2.220446049250313e-16;
  const int maxstep = 10000, maxcor = 3, msbp = 20;
  const double ccmax = 0.3, ccmxj = 0.2;
  BdfResult res{0, 0, 0, 1, 0};
  // a first call with tout = t returns at once, y untouched (vode.f:5475)
  if (!(tout > 0.0)) return res;
  double acor[N_], ewt[N_], savf[N_], yy[N_];  // ewt: reciprocal weights
  double P[N_][N_];
  int piv[N_];
#pragma unroll
  for (int i = 0; i < N_; i++) { z(0, i) = y[i]; ewt[i] = 1.0 / (rtol * fabs(y[i]) + atol); }
  kin.rhs(rho, y, savf);
  res.nfe++;
  // initial step by DVHIN's rule (vode.f:3338-3460): between 100 uround tout
  // and a tenth of tout, and no component moving by more than a tenth of
  // itself plus atol in one step; within that, h ~ sqrt(2 / ||y''||) with
  // y'' from a difference of derivatives, refined at most four times, then
  // halved
  double h;
  {
    const double hlb = 100.0 * uround * tout;
    double hub = 0.1 * tout;
#pragma unroll
    for (int i = 0; i < N_; i++) {
      const double dely = 0.1 * fabs(y[i]) + atol, af = fabs(savf[i]);
      if (af * hub > dely) hub = dely / af;
    }
    double hg = sqrt(hlb * hub);
    if (hub < hlb) {
      h = hg;
    } else {
      double hnew = hg;
      for (int it = 1; it <= 4; it++) {
        double fy[N_];
#pragma unroll
        for (int i = 0; i < N_; i++) yy[i] = y[i] + hg * savf[i];
        kin.rhs(rho, yy, fy);
        res.nfe++;
        const double rhg = 1.0 / hg;
#pragma unroll
        for (int i = 0; i < N_; i++) fy[i] = (fy[i] - savf[i]) * rhg;
        const double ydd = d_wrms(fy, ewt);
        hnew = (ydd * hub * hub > 2.0) ? sqrt(2.0 / ydd) : sqrt(hg * hub);
        if (it == 4) break;
        const double hrat = hnew / hg;
        if (hrat > 0.5 && hrat < 2.0) break;
        if (it >= 2 && hnew > 2.0 * hg) { hnew = hg; break; }
        hg = hnew;
      }
      h = fmin(fmax(0.5 * hnew, hlb), hub);
    }
  }
#pragma unroll
  for (int i = 0; i < N_; i++) z(1, i) = h * savf[i];
  int q = 1, nqwait = 2, ncf = 0, kflag = 0;
  // etamax: largest step growth (DVSTEP's ETAMX1 = 1e4 after the first step,
  // then 10; 1 right after a failure)
  double t = 0.0, rc = 0.0, crate = 1.0, hl0 = 0.0, etamax = 1.0e4, hsave = 1.0;
  bool need_p = true;   // DVODE's IPUP
  bool p_ok = true;     // the last factorisation found a pivot in every column
  int nslp = 0, nslj = 0;
  int icf = 0;          // 1: corrector failed with a saved J, 2: with a current J
  while (t < tout) {
    if (res.nst >= maxstep) { res.ok = 0; break; }
    // error weights at the start of the step
#pragma unroll
    for (int i = 0; i < N_; i++) ewt[i] = 1.0 / (rtol * fabs(z(0, i)) + atol);
    // predictor: multiply the Nordsieck array by the Pascal triangle
    d_nordsieck_shift(z, q, 1.0);
    const double l0 = IZ_CBDF.el[q][0];
    hl0 = h * l0;
    const double drc = (rc != 0.0) ? fabs(hl0 / rc - 1.0) : 0.0;
    if (drc > ccmax || res.nst >= nslp + msbp) need_p = true;
    // corrector (chord Newton on G(e) = h f(z0 + l0 e) - z1 - e)
#pragma unroll
    for (int n = 0; n < N_; n++) { yy[n] = z(0, n); acor[n] = 0.0; }
    typename Kin::Cache kc;
    kin.rhs_c(rho, yy, savf, kc);
    res.nfe++;
    bool jcur = false;
    // The iteration matrix is factored at every step: a lane that asked
    // for a new one builds it from h l0 (after a new J where one is due),
    // every other lane rebuilds its current one from the same J and the
    // h l0 it was built with (rc), which reproduces it bit for bit.  So the
    // factors are never live across the step's kinetics (the predictor
    // right-hand side, the Jacobian, the error test and order selection):
    // the register budget goes to the kinetics instead of spills, which
    // costs less than the extra factorisations (8.7 % less time on the
    // 256^3 reacting half step, 4 % on the ignition ensemble).
    {
      const bool new_j = need_p && (res.nst == 0 || res.nst > nslj + jac_age ||
                                    (icf == 1 && drc < ccmxj) || icf == 2);
      if (new_j) {
        double fac = d_wrms(savf, ewt);
        double r0 = 1000.0 * fabs(h) * uround * N_ * fac;
        if (r0 == 0.0) r0 = 1.0;
        kin.jac_c(rho, yy, savf, kc, -1.0, ewt, r0, J);  // J = (f(y + r e_j) - f(y)) / r
        res.nfe += N_;
        res.nje++;
        nslj = res.nst;
        jcur = true;
      }
      const double con = need_p ? -hl0 : -rc;
#pragma unroll
      for (int i = 0; i < N_; i++)
#pragma unroll
        for (int j = 0; j < N_; j++) P[i][j] = con * J[i][j];
#pragma unroll
      for (int i = 0; i < N_; i++) P[i][i] += 1.0;
      p_ok = d_lu(P, piv);
      if (need_p) {
        rc = hl0;
        nslp = res.nst;
        need_p = false;
        crate = 1.0;
      }
    }
    bool conv = false;
    double del = 0.0, delp = 0.0;
    // a singular iteration matrix counts as a corrector failure (DVNLSD,
    // IERPJ != 0)
    for (int m = 0; m < maxcor && p_ok; m++) {
      double d[N_];
#pragma unroll
      for (int n = 0; n < N_; n++) d[n] = h * savf[n] - z(1, n) - acor[n];
      d_lusolve(P, piv, d);
      res.nsol++;
      if (rc != hl0) {  // chord correction for a stale h*l0
        double cs = 2.0 / (1.0 + hl0 / rc);
#pragma unroll
        for (int n = 0; n < N_; n++) d[n] *= cs;
      }
      del = d_wrms(d, ewt);
#pragma unroll
      for (int n = 0; n < N_; n++) { acor[n] += d[n]; yy[n] = z(0, n) + l0 * acor[n]; }
      // DVNLSD's convergence test: the rate estimate decays by CRDOWN = 0.3,
      // and the correction must be below CORTES = 0.1 of the error-test
      // tolerance
      if (m > 0) crate = fmax(0.3 * crate, del / delp);
      double dcon = del * fmin(1.0, crate) / (0.1 * IZ_CBDF.tesco[q][1]);
      if (dcon <= 1.0) { conv = true; break; }
      if (m >= 1 && del > 2.0 * delp) break;
      delp = del;
      kin.rhs(rho, yy, savf);
      res.nfe++;
    }
    if (!conv) {
      // undo the prediction; retry with a new iteration matrix, and with a
      // quarter of the step (DVSTEP's ETACF) when the Jacobian was already
      // current; the step after a failure keeps its size (ETAMAX = 1)
      d_nordsieck_shift(z, q, -1.0);
      ncf++;
      need_p = true;
      etamax = 1.0;
      const bool hard = jcur || !p_ok;
      icf = hard ? 2 : 1;
      if (hard || ncf > 10) {
        if (ncf > 10) { res.ok = 0; break; }
        d_nordsieck_rescale(z, q, 0.25);
        h *= 0.25;
      }
      continue;
    }
    icf = 0;
    // local error test
    double dsm = d_wrms(acor, ewt) / IZ_CBDF.tesco[q][1];
#ifdef IZ_BDF_TRACE
    printf("t %.4e h %.3e q %d dsm %.3e conv %d jcur %d T %.6f\n", t, h, q, dsm, (int)conv,
           (int)jcur, z(0, 0));
#endif
    if (dsm > 1.0) {
      // DVSTEP's policy after an error-test failure: eta = 1 / ((6 dsm)^(1/L)
      // + 1e-6), at least 0.1 and at most 0.2 after the second failure in a
      // row; after the third, one order lower at a tenth of the step, or at
      // order 1 a restart from a fresh derivative; the seventh gives up
      d_nordsieck_shift(z, q, -1.0);
      kflag--;
      etamax = 1.0;
      if (kflag <= -7) { res.ok = 0; break; }
      if (kflag > -3) {
        double eta = fmax(1.0 / (d_powpos(6.0 * dsm, IZ_CBDF.rexp[q + 1]) + 1.0e-6), 0.1);
        if (kflag <= -2) eta = fmin(eta, 0.2);
        d_nordsieck_rescale(z, q, eta);
        h *= eta;
        continue;
      }
      if (q > 1) {
        q--;
        d_nordsieck_rescale(z, q, 0.1);
        h *= 0.1;
        nqwait = q + 1;
        continue;
      }
      h *= 0.1;
#pragma unroll
      for (int n = 0; n < N_; n++) yy[n] = z(0, n);
      kin.rhs(rho, yy, savf);
      res.nfe++;
#pragma unroll
      for (int n = 0; n < N_; n++) z(1, n) = h * savf[n];
      nqwait = 10;
      need_p = true;
      continue;
    }
    // accept
    kflag = 0;
    ncf = 0;
    t += h;
    res.nst++;
    for (int j = 0; j <= q; j++) {
      double e = IZ_CBDF.el[q][j];
#pragma unroll
      for (int n = 0; n < N_; n++) z(j, n) += e * acor[n];
    }
    nqwait--;
    if (nqwait == 1 && q < 5) {
#pragma unroll
      for (int n = 0; n < N_; n++) z(5, n) = acor[n];  // kept for the order-up estimate
      hsave = h;
    }
    if (etamax == 1.0) {
      // the step after a failure keeps h and q
      if (nqwait < 2) nqwait = 2;
      etamax = 10.0;
      continue;
    }
    // step size every step, order every q+1 steps (DVSTEP: ETAQ, and
    // ETAQM1 / ETAQP1 when NQWAIT reaches 0; biases 6, 6 and 10)
    double eta = 1.0 / (d_powpos(6.0 * dsm, IZ_CBDF.rexp[q + 1]) + 1.0e-6);
    int newq = q;
    if (nqwait == 0) {
      nqwait = 2;
      double etam = 0.0, etap = 0.0;
      if (q > 1) {
        double ddn = d_wrms_col(z, q, ewt) / IZ_CBDF.tesco[q][0];
        etam = 1.0 / (d_powpos(6.0 * ddn, IZ_CBDF.rexp[q]) + 1.0e-6);
      }
      if (q < 5) {
        // the saved correction, scaled to the current step (acor ~ h^(q+1))
        const double cq = d_powpos(h / hsave, (double)(q + 1));
        double dd[N_];
#pragma unroll
        for (int n = 0; n < N_; n++) dd[n] = acor[n] - cq * z(5, n);
        double dup = d_wrms(dd, ewt) / IZ_CBDF.tesco[q][2];
        etap = 1.0 / (d_powpos(10.0 * dup, IZ_CBDF.rexp[q + 2]) + 1.0e-6);
      }
      if (eta >= etap) {
        if (eta < etam) { eta = etam; newq = q - 1; }
      } else if (etap > etam) {
        eta = etap; newq = q + 1;
      } else {
        eta = etam; newq = q - 1;
      }
    }
    if (eta < 1.5) { etamax = 10.0; continue; }  // THRESH: keep h and q
    eta = fmin(eta, etamax);
    etamax = 10.0;
    if (newq > q) {
      // new highest-order column from the last correction
      double e = IZ_CBDF.el[q][q] / (q + 1);
#pragma unroll
      for (int n = 0; n < N_; n++) z(newq, n) = acor[n] * e;
    }
    if (newq != q) nqwait = newq + 1;
    q = newq;
    d_nordsieck_rescale(z, q, eta);
    h *= eta;
  }
  // interpolate back to tout: y = sum_j z(j) s^j
  double s = (tout - t) / h;
#pragma unroll
  for (int n = 0; n < N_; n++) {
    double acc = z(q, n);
    for (int j = q - 1; j >= 0; j--) acc = acc * s + z(j, n);
    y[n] = acc;
  }
  return res;
}


#ifdef GEN_KK
// straight-line kinetics of the compiled-in mechanism; the T-only part kt is
// shared by the Jacobian columns that perturb mass fractions
struct KinGen {
  struct Cache { double kt[GEN_NKT]; };  // rate constants at the evaluated T
  IZ_MHD inline void rhs(double rho, const double *y, double *f) const {
    double kt[GEN_NKT];
    gen_kT(y[0], kt);
    gen_rhs(rho, y, kt, f);
  }
  IZ_MHD inline void rhs_c(double rho, const double *y, double *f, Cache &c) const {
    gen_kT(y[0], c.kt);
    gen_rhs(rho, y, c.kt, f);
  }
  // The species columns reuse the rate constants of the evaluation of f0 (T
  // is not perturbed) and come first; the temperature column, with the
  // constants of the perturbed temperature, comes last, so that only one
  // set of constants is live at a time.  Each column perturbs a local copy
  // of the state: perturbing the caller's vector in place and restoring it
  // gave wrong Jacobians in a GPU build of the unrolled loop (correct on
  // the CPU), so the state is never written here.
  IZ_MHD inline void jac_c(double rho, const double *y, const double *f0, const Cache &c,
                           double hl0, const double *w, double r0, double (*P)[N_]) const {
    const double *kt = c.kt;
    for (int j = 1; j < N_; j++) {
      double yl[N_], ftmp[N_];
#pragma unroll
      for (int n = 0; n < N_; n++) yl[n] = y[n];
      const double yj = yl[j];
      const double r = d_fd_increment(j, yj, w[j], r0);
      yl[j] = yj + r;
      gen_rhs(rho, yl, kt, ftmp);
      const double ri = 1.0 / r;
#pragma unroll
      for (int i = 0; i < N_; i++) P[i][j] = -hl0 * (ftmp[i] - f0[i]) * ri;
    }
    double yl[N_], ftmp[N_], kt0[GEN_NKT];
#pragma unroll
    for (int n = 0; n < N_; n++) yl[n] = y[n];
    const double r = d_fd_increment(0, yl[0], w[0], r0);
    yl[0] = yl[0] + r;
    gen_kT(yl[0], kt0);
    gen_rhs(rho, yl, kt0, ftmp);
    const double ri = 1.0 / r;
#pragma unroll
    for (int i = 0; i < N_; i++) P[i][0] = -hl0 * (ftmp[i] - f0[i]) * ri;
  }
};

// The straight-line kinetics with the mass-fraction columns of the
// Jacobian in closed form (gen_jac, generated with the rates: mass action,
// third-body sums, falloff factors differentiated) and the temperature
// column by a finite difference as above: one evaluation of the kinetics
// per Jacobian instead of GEN_KK + 1.  DVODE's user-Jacobian form (MF = 21,
// vode.f:3596-3727) in place of the reference's MF = 22; the integration
// keeps the same error control, so results stay within its tolerance.
struct KinGenAJ : KinGen {
  IZ_MHD inline void jac_c(double rho, const double *y, const double *f0, const Cache &c,
                           double hl0, const double *w, double r0, double (*P)[N_]) const {
    gen_jac(rho, y, c.kt, f0[0], P);
    if (hl0 != -1.0) {
#pragma unroll
      for (int i = 0; i < N_; i++)
#pragma unroll
        for (int j = 1; j < N_; j++) P[i][j] = -hl0 * P[i][j];
    }
    double yl[N_], ftmp[N_], kt0[GEN_NKT];
#pragma unroll
    for (int n = 0; n < N_; n++) yl[n] = y[n];
    const double r = d_fd_increment(0, yl[0], w[0], r0);
    yl[0] = yl[0] + r;
    gen_kT(yl[0], kt0);
    gen_rhs(rho, yl, kt0, ftmp);
    const double ri = 1.0 / r;
#pragma unroll
    for (int i = 0; i < N_; i++) P[i][0] = -hl0 * (ftmp[i] - f0[i]) * ri;
  }
};
#endif
