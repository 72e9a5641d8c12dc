// images.hip -- the flat table images izem_create and izem_chem_setup take,
// built from what a reference host already holds (host code, no kernels):
//
//  * izem_thd_image: from the arrays of its thd_type after thd_init
//    (devel-izem/src/thd/thd_tools_perfectgas.f90:12-248) and the reference
//    quantities of adi_type;
//  * izem_mech_image_ckwrk: from CHEMKIN-II's work arrays after CKINIT
//    (devel-izem/src/chemkin/vode.f:391), read through the pointers of its
//    /CKSTRT/ common block (chemkin/ckstrt.common) in their declared order.
//
// So a Fortran host hands the library its own tables and needs no other
// front end.
#include <math.h>
#include <string.h>

#include "izem_ctx.h"
#include "chem_setup.h"

static_assert(IZEM_THD_IMAGE_LEN == sizeof(IzThd) / sizeof(double), "thd image length");
static_assert(IZEM_MECH_IMAGE_LEN == IZ_MECH_IMAGE_LEN, "mechanism image length");

// Wilke's mixing factors of the viscosity (module_thermodynamics.f90:
// 1166-1224 evaluates (W_l / W_m)^(1/4) and 1 / sqrt(8 (1 + W_m / W_l)))
extern "C" long izem_thd_image(int nspc, int ld, const double *W, const double *Wc,
                               const double *Wc_i, const double *HeatCap,
                               const double *Tpolymin, const double *Tpolylim,
                               const double *Tpolymax, const double *CofD, const double *CofLam,
                               const double *CofEta, double gam2, double Tinf, double p0,
                               double pinf, double Dinf, double laminf, double muinf,
                               double *image, long cap) {
  const int n = IZ_NRV;
  const long len = IZEM_THD_IMAGE_LEN;
  if (nspc != n || ld < n || !W || !Wc || !Wc_i || !HeatCap || !Tpolymin || !Tpolylim ||
      !Tpolymax || !CofD || !CofLam || !CofEta || !image)
    return IZEM_E_ARG;
  if (cap < len) return IZEM_E_ARG;
  long p = 0;
  for (int k = 0; k < n; k++) image[p++] = W[k];
  for (int k = 0; k < n; k++) image[p++] = Wc[k];
  for (int k = 0; k < n; k++) image[p++] = Wc_i[k];
  // HeatCap(7, 2, ld): coefficient fastest, then the range, then species
  for (int k = 0; k < n; k++)
    for (int r = 0; r < 2; r++)
      for (int c = 0; c < 7; c++) image[p++] = HeatCap[c + 7 * (r + 2 * (long)k)];
  for (int k = 0; k < n; k++) image[p++] = Tpolymin[k];
  for (int k = 0; k < n; k++) image[p++] = Tpolylim[k];
  for (int k = 0; k < n; k++) image[p++] = Tpolymax[k];
  // CofD(4, ld, ld) as image [l][m][4] = CofD(:, l, m); CofLam(4, ld),
  // CofEta(4, ld)
  for (int l = 0; l < n; l++)
    for (int m = 0; m < n; m++)
      for (int c = 0; c < 4; c++) image[p++] = CofD[c + 4 * (l + (long)ld * m)];
  for (int k = 0; k < n; k++)
    for (int c = 0; c < 4; c++) image[p++] = CofLam[c + 4L * k];
  for (int k = 0; k < n; k++)
    for (int c = 0; c < 4; c++) image[p++] = CofEta[c + 4L * k];
  for (int m = 0; m < n; m++)
    for (int l = 0; l < n; l++) image[p++] = pow(W[l] / W[m], 0.25);
  for (int m = 0; m < n; m++)
    for (int l = 0; l < n; l++) image[p++] = 1.0 / sqrt((1.0 + W[m] / W[l]) * 8.0);
  const double tail[7] = {gam2, Tinf, p0, pinf, Dinf, laminf, muinf};
  for (int q = 0; q < 7; q++) image[p++] = tail[q];
  return p;
}

// /CKSTRT/ in declaration order (chemkin/ckstrt.common)
enum CkStrt {
  NMM, NKK, NII, MXSP, MXTB, MXTP, NCP, NCP1, NCP2, NCP2T, NPAR, NLAR, NFAR, NLAN, NFAL, NREV,
  NTHB, NRLT, NWL, NEIM, NJAN, NJAR, NFT1, NF1R, NEXC, NRNU, NORD, MXORD,
  IcMM, IcKK, IcNC, IcPH, IcCH, IcNT, IcNU, IcNK, IcNS, IcNR, IcLT, IcRL, IcRV, IcWL, IcFL,
  IcFO, IcKF, IcTB, IcKN, IcKT, IcEI, IcTD, IcJN, IcF1, IcEX, IcRNU, IcORD, IcKOR,
  NcAW, NcWT, NcTT, NcAA, NcCO, NcRV, NcLT, NcRL, NcFL, NcKT, NcWL, NcJN, NcF1, NcEX, NcRU,
  NcRC, NcPA, NcKF, NcKR, NcRNU, NcKOR
};

extern "C" long izem_mech_image_ckwrk(const int *ickwrk, const double *rckwrk,
                                      const int *ckstrt, double *image, long cap) {
  if (!ickwrk || !rckwrk || !ckstrt || !image) return IZEM_E_ARG;
  const int *s = ckstrt;
  const int kk = s[NKK], ii = s[NII], mxsp = s[MXSP], mxtb = s[MXTB], mxtp = s[MXTP];
  const int ncp2 = s[NCP2], ncp2t = s[NCP2T], npar = s[NPAR], nfar = s[NFAR];
  // forms the kernels do not evaluate: Landau-Teller, real stoichiometry,
  // orders, photon / electron / Janev / fit-1 / excitation reactions
  if (kk < 1 || kk > CK_KMAX || ii < 1 || ii > CK_IMAX || mxsp != 6 || mxtp != 3 ||
      ncp2 != 7 || npar != 3 || (s[NFAL] > 0 && nfar > 8) || mxtb > CK_TBMAX || s[NLAN] ||
      s[NRLT] || s[NWL] || s[NEIM] || s[NJAN] || s[NFT1] || s[NEXC] || s[NRNU] || s[NORD])
    return IZEM_E_ARG;
  const long len = IZEM_MECH_IMAGE_LEN;
  if (cap < len) return IZEM_E_ARG;
  memset(image, 0, sizeof(double) * len);
  // Fortran 1-based pointers into the work arrays
  auto I = [&](int ptr, long off) { return ickwrk[ptr - 1 + off]; };
  auto R = [&](int ptr, long off) { return rckwrk[ptr - 1 + off]; };
  long p = 0;
  auto take = [&](long n) { long o = p; p += n; return o; };
  long o = take(2);
  image[o] = kk;
  image[o + 1] = ii;
  const long owt = take(CK_KMAX), oa = take(CK_KMAX * 14), otm = take(CK_KMAX);
  for (int k = 0; k < kk; k++) {
    image[owt + k] = R(s[NcWT], k);
    // AA(NcAA + (L-1) NCP2 + (K-1) NCP2T + N-1): L = 1 the range up to the
    // middle temperature TT(2)
    for (int r = 0; r < 2; r++)
      for (int c = 0; c < 7; c++)
        image[oa + 14L * k + 7 * r + c] = R(s[NcAA], (long)r * ncp2 + (long)k * ncp2t + c);
    image[otm + k] = R(s[NcTT], (long)k * mxtp + 1);
  }
  const long oA = take(CK_IMAX), ob = take(CK_IMAX), oE = take(CK_IMAX);
  const long onu = take(6L * CK_IMAX), onk = take(6L * CK_IMAX), orev = take(CK_IMAX);
  const long ontb = take(CK_IMAX), otbk = take(CK_IMAX * CK_TBMAX);
  const long otbe = take(CK_IMAX * CK_TBMAX), ofal = take(CK_IMAX);
  const long ofp = take(8L * CK_IMAX), okf = take(CK_IMAX), orp = take(3L * CK_IMAX);
  // no third body: -1, also in the unused slots
  for (int i = 0; i < CK_IMAX; i++) image[ontb + i] = -1;
  for (int i = 0; i < ii; i++) {
    image[oA + i] = R(s[NcCO], (long)i * (npar + 1));
    image[ob + i] = R(s[NcCO], (long)i * (npar + 1) + 1);
    image[oE + i] = R(s[NcCO], (long)i * (npar + 1) + 2);
    for (int n = 0; n < 6; n++) {
      image[onu + 6L * i + n] = I(s[IcNU], (long)i * mxsp + n);
      image[onk + 6L * i + n] = I(s[IcNK], (long)i * mxsp + n);
    }
    // NS < 0: irreversible
    image[orev + i] = I(s[IcNS], i) < 0 ? 0 : 1;
  }
  for (int q = 0; q < s[NREV]; q++) {
    const int i = I(s[IcRV], q) - 1;
    image[orev + i] = 2;
    for (int c = 0; c < 3; c++) image[orp + 3L * i + c] = R(s[NcRV], (long)q * (npar + 1) + c);
  }
  for (int q = 0; q < s[NTHB]; q++) {
    const int i = I(s[IcTB], q) - 1, nk = I(s[IcKN], q);
    image[ontb + i] = nk;
    for (int l = 0; l < nk; l++) {
      image[otbk + (long)CK_TBMAX * i + l] = I(s[IcKT], (long)q * mxtb + l);
      image[otbe + (long)CK_TBMAX * i + l] = R(s[NcKT], (long)q * mxtb + l);
    }
  }
  for (int q = 0; q < s[NFAL]; q++) {
    const int i = I(s[IcFL], q) - 1;
    image[ofal + i] = I(s[IcFO], q);
    image[okf + i] = I(s[IcKF], q);
    for (int c = 0; c < nfar; c++) image[ofp + 8L * i + c] = R(s[NcFL], (long)q * nfar + c);
  }
  return p;
}
