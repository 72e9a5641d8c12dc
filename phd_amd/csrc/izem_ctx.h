// izem_ctx.h -- host-side context shared by the translation units of the
// IZEM HIP library (flow.hip, chem.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/izem_hip.h"
#include "izem_common.h"

#define IZ_CHECK(x)                                  \
  do {                                               \
    hipError_t e_ = (x);                             \
    if (e_ != hipSuccess) return IZEM_E_HIP;         \
  } while (0)

struct IzChem;  // chem.hip

struct izem_ctx {
  IzDims d;
  izem_opts o;
  hipStream_t stream;
  hipStream_t own_stream;
  // Euler face arrays (14 x faces per direction)
  double *F[3];
  long nfaces[3];
  // viscous scratch: velocities + mole fractions, gradients / diffusion
  // fluxes, stresses + heat flux
  double *prim, *fd, *sf;
  int nderiv, nsf;
  double *v0;
  // device words: Newton masks [7], status, iteration count
  unsigned long long *masks;
  int *status;
  int *iters;
  double *red;  // timestep reductions [12]
  double *ts;   // dt_4[4], dtmin, DTmax
  IzChem *chem;
};

static inline unsigned iz_blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// internal entry points across translation units
int iz_ghost_newton(izem_ctx *ctx, double *T, double *W_i, double *cp, double *ha, double *v);
int iz_comm_cons(izem_ctx *ctx, double *a, int nf);
int iz_upd_prim_woT(izem_ctx *ctx, int domain, const double *v, const double *T, double *W_i,
                    double *cp, double *ha);
void iz_chem_destroy(izem_ctx *ctx);
