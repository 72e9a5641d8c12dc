// chem_tables.h -- mechanism and BDF coefficient tables of the chemistry
// kernels (the content of the CHEMKIN linking file, chemkin/vode.f:391).
#pragma once

// qualifier of device free functions and of policy member functions; a CPU
// build of these headers (for debugging the integrator) defines them away
#ifndef IZ_HD
#define IZ_HD __device__
#endif
#ifndef IZ_MHD
#define IZ_MHD __device__
#endif
// the integrator is inlined into each kernel variant, so that every
// register budget gets its own allocation of the per-cell state
#ifndef IZ_FORCEINLINE
#define IZ_FORCEINLINE __attribute__((always_inline)) inline
#endif
// true when the predicate holds on any active lane of the wave (a CPU
// build, one cell at a time, defines it as the predicate itself)
#ifndef IZ_ANY
#define IZ_ANY(pred) (__any((int)(pred)) != 0)
#endif

#define CK_KMAX 16
#define CK_IMAX 48
#define CK_TBMAX 12
#define CK_NEQ 10  // T + 9 species (the H2/air mechanism); checked at setup

struct IzMech {
  int kk, ii;
  double wt[CK_KMAX];
  double a[CK_KMAX][2][7];
  double tmid[CK_KMAX];
  double A[CK_IMAX], b[CK_IMAX], EoR[CK_IMAX];
  int nu[CK_IMAX][6], nunk[CK_IMAX][6];
  int reversible[CK_IMAX];
  int ntb[CK_IMAX];
  int tb_k[CK_IMAX][CK_TBMAX];
  double tb_eff[CK_IMAX][CK_TBMAX];
  int fal[CK_IMAX];            // 0 none, 1 Lindemann, 2 SRI, 3/4 Troe (CKINTP's IFOP)
  double fpar[CK_IMAX][8];     // low-pressure A, b, E/R, then the SRI or Troe values
  int kfal[CK_IMAX];           // falloff [M]: 0 the third-body sum, k + 1 species k
  double rpar[CK_IMAX][3];     // REV: reverse A, b, E/R (reversible[i] == 2)
};


// BDF coefficients (Nordsieck l-vectors and error constants) for orders 1..5
struct IzBdf {
  double el[6][6];     // el[q][j], j = 0..q, el[q][1] = 1
  double tesco[6][3];  // [q][0] order-down, [1] same order, [2] order-up constants
  double rexp[8];      // rexp[m] = 1 / m, the exponents of the step ratios
};


