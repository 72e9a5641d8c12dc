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

// kernel classes timed by the event profiler (izem_profile_*)
enum IzProfClass {
  IZP_EULER = 0,   // Euler face kernels (x, y, z)
  IZP_VISC = 1,    // viscous prim / gradient / flux kernels
  IZP_RHS = 2,     // fused flux difference + RK stage update + Newton probe
  IZP_NEWTON = 3,  // Newton apply / ghost Newton
  IZP_CHEM = 4,    // chemistry queue kernels
  IZP_OTHER = 5,   // ghost fills, sensor, time step
  IZP_NCLASS = 6
};

struct IzProf {
  int on, n, cap;
  hipEvent_t *ev;  // 2 per record
  int *cls;
};

// a captured main-loop iteration (izem_step) and what it was captured with
#define IZ_GRAPH_PTRS 9
#define IZ_CHEM_SIG 8
struct IzGraph {
  hipGraph_t g;
  hipGraphExec_t exec;
  const void *ptr[IZ_GRAPH_PTRS];
  int react, flux_mode;
  double chem[IZ_CHEM_SIG];
};

struct izem_ctx {
  IzDims d;
  izem_opts o;
  hipStream_t stream;
  hipStream_t own_stream;
  // Euler face arrays (14 x faces per direction)
  double *F[3];
  long nfaces[3];
  // faces between two ranks formed with the left rank's view (d.np > 1):
  // 14 x nseam per direction
  double *Fs[3];
  long nseam[3];
  // the post-shock state of a slip-wall face (IZ_NV doubles) and whether
  // it was given
  double *wall_state;
  int wall_set;
  // the characteristic outflow's relaxation: domain length Lx, far-field
  // pressure and 1/dx of the last cell (nondimensional), and whether given
  double nscbc_Lx, nscbc_pinf, nscbc_dxi;
  int nscbc_set;
  // viscous scratch: velocities + mole fractions, gradients / diffusion
  // fluxes, stresses + heat flux
  double *prim, *fd, *sf;
  int nderiv, nsf;
  double *v0;
  // device words: Newton probe masks (blocking iterations [0..6] and
  // negative temperatures [8..14] per domain), status, iteration counts
  unsigned long long *masks;
  int *negT_rec;  // cells recorded by the negative-temperature repair
  int *status;
  int *iters;
  double *red;  // timestep reductions [24]
  double *ts;   // dt_4[4], dtmin, DTmax
  IzChem *chem;
  IzProf prof;
  IzGraph graph;
  int flux_mode;  // 0: reference operation order (bit parity), 1: structured projections
  int ncu;  // compute units of the device (grid caps of reducing kernels)
};

// bracket a kernel launch with events on the context stream
static inline int iz_prof_begin(izem_ctx *c, int cls) {
  IzProf &p = c->prof;
  if (!p.on || p.n >= p.cap) return -1;
  int k = p.n++;
  p.cls[k] = cls;
  (void)hipEventRecord(p.ev[2 * k], c->stream);
  return k;
}
static inline void iz_prof_end(izem_ctx *c, int k) {
  if (k >= 0) (void)hipEventRecord(c->prof.ev[2 * k + 1], c->stream);
}

static inline unsigned iz_blocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
// grid of a grid-stride kernel: enough blocks for n items, at most per_cu
// blocks per compute unit
static inline unsigned iz_blocks_capped(const izem_ctx *c, long n, int bs, int per_cu) {
  long nb = (n + bs - 1) / bs, cap = (long)c->ncu * per_cu;
  if (nb > cap) nb = cap;
  return (unsigned)(nb < 1 ? 1 : nb);
}

// internal entry points across translation units
int iz_ghost_newton(izem_ctx *ctx, double *T, double *W_i, double *cp, double *ha, double *v);
// bcs = 1: the periodic exchange and the non-periodic faces' conditions
// (comm_cons + upd_boundaries); 0: the exchange alone.  The characteristic
// outflow needs the primitives T, W_i, cp and the step dt (scale x *dtp);
// without them (nullptr) a context with such a face returns IZEM_E_STATE.
int iz_comm_cons(izem_ctx *ctx, double *a, int nf, int bcs, const double *T = nullptr,
                 const double *W_i = nullptr, const double *cp = nullptr,
                 const double *dtp = nullptr, double scale = 1.0);
int iz_upd_prim_woT(izem_ctx *ctx, int domain, const double *v, const double *T, double *W_i,
                    double *cp, double *ha);
void iz_chem_destroy(izem_ctx *ctx);
void iz_graph_destroy(izem_ctx *ctx);
