// Host build of the chemistry integrator (phd_amd/csrc/chem_device.h) for
// the tests: the GPU kernels and this build compile the same source, so a
// GPU variant that follows a different trajectory from this one (more
// steps, another end state) points at its compilation, not at the method.
// One cell at a time: the wave-wide predicate IZ_ANY is the lane's own.
#include <math.h>
#include <string.h>
#define IZ_HD static
#define IZ_MHD
#define IZ_FORCEINLINE inline
#define IZ_ANY(p) (p)
#define IZ_CMECH h_mech
#define IZ_CBDF h_bdf
#include "chem_tables.h"
#include "chem_setup.h"
static IzMech h_mech;
static IzBdf h_bdf;
#include "mech_gen.h"
#include "chem_device.h"

extern "C" int host_setup(const double *mf, long n) {
  iz_bdf_tables(h_bdf);
  return iz_mech_from_flat(mf, n, CK_NEQ - 1, h_mech) ? 0 : -1;
}

// izem_psr_batch's integration of each cell
extern "C" void host_psr(int kinetics, int n, double *T, const double *rho, double *Y, double dt,
                         double rtol, double atol, int jac_age, int *nst, int *ok) {
  for (int c = 0; c < n; c++) {
    double y[N_];
    y[0] = T[c];
    for (int k = 0; k < N_ - 1; k++) y[1 + k] = Y[c * (N_ - 1) + k];
    ZArr z;
    double J[N_][N_] = {};
    BdfResult r = kinetics == 2
        ? d_bdf_integrate(KinGenAJ(), z, J, rho[c], y, dt, rtol, atol, jac_age)
        : kinetics
        ? d_bdf_integrate(KinGen(), z, J, rho[c], y, dt, rtol, atol, jac_age)
        : d_bdf_integrate(KinTables(), z, J, rho[c], y, dt, rtol, atol, jac_age);
    T[c] = y[0];
    for (int k = 0; k < N_ - 1; k++) Y[c * (N_ - 1) + k] = y[1 + k];
    nst[c] = r.nst;
    ok[c] = r.ok;
  }
}

// the reactor right-hand side (FUN) of either kinetics form
extern "C" void host_rhs(int kinetics, double rho, const double *z, double *zp) {
  if (kinetics) KinGen().rhs(rho, z, zp);
  else KinTables().rhs(rho, z, zp);
}

// the mass-fraction columns of the reactor Jacobian in closed form
// (gen_jac) at z; J row-major 10 x 10, column 0 untouched
extern "C" void host_jac(double rho, const double *z, double *J) {
  double kt[GEN_NKT], f[N_];
  gen_kT(z[0], kt);
  gen_rhs(rho, z, kt, f);
  gen_jac(rho, z, kt, f[0], (double (*)[N_])J);
}
