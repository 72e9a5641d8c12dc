// chem.hip -- batched stiff chemistry (entry points return IZEM_E_STATE
// until the kinetics tables are set up)
#include "izem_ctx.h"

void iz_chem_destroy(izem_ctx *ctx) { (void)ctx; }

extern "C" int izem_chem_setup(izem_ctx *, const double *, long, double, double, int) {
  return IZEM_E_STATE;
}
extern "C" int izem_ckwyr(izem_ctx *, int, const double *, const double *, const double *,
                          double *) {
  return IZEM_E_STATE;
}
extern "C" int izem_psr_batch(izem_ctx *, int, double *, const double *, double *, double,
                              int *) {
  return IZEM_E_STATE;
}
extern "C" int izem_reaction_dvode(izem_ctx *, const double *, double, double *, double *,
                                   double *, double *, double *) {
  return IZEM_E_STATE;
}
