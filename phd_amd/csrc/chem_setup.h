// chem_setup.h -- host-side construction of the chemistry tables: the
// mechanism image (IzMech) from its flat double array, and the BDF
// coefficients (IzBdf).  Plain C++, shared by the library and by host
// builds of the integrator.
#pragma once

#include <string.h>

#include "chem_tables.h"

// The flat image holds, in order: kk, ii; wt[KMAX]; a[KMAX][2][7];
// tmid[KMAX]; A, b, EoR [IMAX each]; nu[IMAX][6]; nunk[IMAX][6];
// reversible[IMAX]; ntb[IMAX]; tb_k[IMAX][TBMAX]; tb_eff[IMAX][TBMAX];
// fal[IMAX]; fpar[IMAX][8]; kfal[IMAX]; rpar[IMAX][3] (the content of the CHEMKIN linking file,
// chemkin/vode.f:391).  Returns false when the length or kk does not fit.
#define IZ_MECH_IMAGE_LEN                                                                      \
  (2L + CK_KMAX + CK_KMAX * 14L + CK_KMAX + 3L * CK_IMAX + 12L * CK_IMAX + 2L * CK_IMAX +        \
   2L * CK_IMAX * CK_TBMAX + CK_IMAX + 8L * CK_IMAX + CK_IMAX + 3L * CK_IMAX)

static inline bool iz_mech_from_flat(const double *mf, long nmech, int kk_expected, IzMech &m) {
  memset(&m, 0, sizeof m);
  if (!mf || nmech != IZ_MECH_IMAGE_LEN) return false;
  long p = 0;
  auto take = [&](long n) { long s = p; p += n; return s; };
  long o;
  o = take(2); m.kk = (int)mf[o]; m.ii = (int)mf[o + 1];
  if (m.kk != kk_expected || m.ii < 1 || m.ii > CK_IMAX) return false;
  o = take(CK_KMAX); for (int k = 0; k < CK_KMAX; k++) m.wt[k] = mf[o + k];
  o = take(CK_KMAX * 14); memcpy(m.a, mf + o, sizeof(double) * CK_KMAX * 14);
  o = take(CK_KMAX); for (int k = 0; k < CK_KMAX; k++) m.tmid[k] = mf[o + k];
  o = take(CK_IMAX); for (int i = 0; i < CK_IMAX; i++) m.A[i] = mf[o + i];
  o = take(CK_IMAX); for (int i = 0; i < CK_IMAX; i++) m.b[i] = mf[o + i];
  o = take(CK_IMAX); for (int i = 0; i < CK_IMAX; i++) m.EoR[i] = mf[o + i];
  o = take(6 * CK_IMAX); for (int i = 0; i < 6 * CK_IMAX; i++) (&m.nu[0][0])[i] = (int)mf[o + i];
  o = take(6 * CK_IMAX); for (int i = 0; i < 6 * CK_IMAX; i++) (&m.nunk[0][0])[i] = (int)mf[o + i];
  o = take(CK_IMAX); for (int i = 0; i < CK_IMAX; i++) m.reversible[i] = (int)mf[o + i];
  o = take(CK_IMAX); for (int i = 0; i < CK_IMAX; i++) m.ntb[i] = (int)mf[o + i];
  o = take(CK_IMAX * CK_TBMAX);
  for (int i = 0; i < CK_IMAX * CK_TBMAX; i++) (&m.tb_k[0][0])[i] = (int)mf[o + i];
  o = take(CK_IMAX * CK_TBMAX);
  for (int i = 0; i < CK_IMAX * CK_TBMAX; i++) (&m.tb_eff[0][0])[i] = mf[o + i];
  o = take(CK_IMAX); for (int i = 0; i < CK_IMAX; i++) m.fal[i] = (int)mf[o + i];
  o = take(8 * CK_IMAX); memcpy(m.fpar, mf + o, sizeof(double) * 8 * CK_IMAX);
  o = take(CK_IMAX); for (int i = 0; i < CK_IMAX; i++) m.kfal[i] = (int)mf[o + i];
  o = take(3 * CK_IMAX); memcpy(m.rpar, mf + o, sizeof(double) * 3 * CK_IMAX);
  for (int i = 0; i < m.ii; i++)
    if (m.kfal[i] < 0 || m.kfal[i] > m.kk || m.fal[i] < 0 || m.fal[i] > 4 ||
        m.reversible[i] < 0 || m.reversible[i] > 2)
      return false;
  return true;
}

// Nordsieck l-vectors and error constants of the BDF orders 1..5
static inline void iz_bdf_tables(IzBdf &t) {
  memset(&t, 0, sizeof t);
  for (int q = 1; q <= 5; q++) {
    // coefficients of prod_{i=1..q} (1 + x/i), normalised so that l1 = 1
    double c[7] = {1, 0, 0, 0, 0, 0, 0};
    for (int i = 1; i <= q; i++)
      for (int j = i; j >= 1; j--) c[j] += c[j - 1] / i;
    for (int j = 0; j <= q; j++) t.el[q][j] = c[j] / c[1];
    double fact = 1.0;
    for (int i = 2; i <= q; i++) fact *= i;  // q!
    t.tesco[q][0] = 1.0 / fact;
    t.tesco[q][1] = (q + 1) / t.el[q][0];
    t.tesco[q][2] = (q + 2) / t.el[q][0];
  }
  for (int m = 1; m < 8; m++) t.rexp[m] = 1.0 / m;
}
