// izem_common.h -- device-side types shared by the IZEM HIP kernels.
//
// Layout of every field array is the reference's Fortran allocation
// v(sx-ng:ex+ng, sy-ng:ey+ng, sz-ng:ez+ng, 1:nv): i fastest, ghosts included,
// field-major for multi-field arrays (devel-izem/src/module_ICs.f90:2364-2369).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// species count (9 for the H2/air mechanism) and ghost layers (3 for WENO5,
// 4 for WENO7: ng = (ORDER_WENO + 1) / 2, module_input.f90:712) are fixed
// when a library is built; the default build is 9 species and WENO5, and
// flow-only variants are built for other pairs (phd_amd/build.py)
#ifndef IZ_NRV
#define IZ_NRV 9
#endif
#ifndef IZ_NG
#define IZ_NG 3
#endif
#define IZ_S (2 * IZ_NG)  // points of a face's WENO stencil
#define IZ_NIV 5
#define IZ_NV (IZ_NIV + IZ_NRV)

// thd_type subset (devel-izem/src/thd/type_thd.f90:10-91); flat image
// produced by phd_amd.chemdata.Thd.flat().
struct IzThd {
  double W[IZ_NRV];
  double Wc[IZ_NRV];
  double Wc_i[IZ_NRV];
  double HeatCap[IZ_NRV][2][7];
  double Tpolymin[IZ_NRV], Tpolylim[IZ_NRV], Tpolymax[IZ_NRV];
  double CofD[IZ_NRV][IZ_NRV][4];
  double CofLam[IZ_NRV][4];
  double CofEta[IZ_NRV][4];
  double wilke_c1[IZ_NRV][IZ_NRV];
  double wilke_c2[IZ_NRV][IZ_NRV];
  double gam2, Tinf, p0, pinf, Dinf, laminf, muinf;
};

// phys: bit f set when face f (W, E, S, N, B, F) is a physical boundary
// with extrapolated ghosts (bc_extrapolation) instead of a periodic one
// np: ranks of the decomposed run whose shock-sensor views are reproduced,
// per direction (1: the whole line is one rank)
// wall: face S is a post-shock slip wall (its first nfix cells take a
// post-shock state, the others mirror); nscbc: face E is a subsonic
// characteristic outflow
struct IzDims {
  int nx, ny, nz, ng, ndim;
  long NX, NY, NZ, nc;
  int phys;
  int np[3];
  int wall, nfix;
  int nscbc;
};

// rank r of np along a line of n cells owns cells (r n)/np .. ((r+1) n)/np - 1
// (sx = coords*ntx/dims + 1, ex = (coords+1)*ntx/dims: module_parallel.f90:
// 547-548, zero-based here); positions past either end go to the end ranks
__host__ __device__ inline int iz_rank_lo(int r, int n, int np) {
  return (int)(((long)r * n) / np);
}
__host__ __device__ inline int iz_rank_of(int p, int n, int np) {
  int r = 0;
  while (r + 1 < np && iz_rank_lo(r + 1, n, np) <= p) r++;
  return r;
}
// index of the seam face to the right of cell p when p is the last cell of a
// rank that is not the last one ((r) + (np - 1) line), else -1
__host__ __device__ inline long iz_seam(int p, int n, int np, long line) {
  if (np <= 1) return -1;
  const int r = iz_rank_of(p, n, np);
  if (r + 1 < np && p == iz_rank_lo(r + 1, n, np) - 1) return r + (long)(np - 1) * line;
  return -1;
}

__host__ __device__ inline long iz_cidx(const IzDims &d, int i, int j, int k) {
  return (long)(i + d.ng) + d.NX * ((long)(j + d.ng) + d.NY * (long)(k + d.ng));
}

// status word bits reported back through the C ABI
#define IZ_ST_NEG_PRESSURE 1
#define IZ_ST_NEG_TEMPERATURE 2
#define IZ_ST_NAN 4
#define IZ_ST_NEWTON_ITMAX 8
#define IZ_ST_VODE_FAIL 16
#define IZ_ST_NEG_T_ABORT 32
