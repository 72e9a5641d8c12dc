/*
 * izem_oracle.h -- CPU restatement of the IZEM reacting-flow right-hand side,
 * time step and CHEMKIN-form kinetics.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * Every function follows the reference Fortran operation by operation (same
 * loop order, same left-to-right expression trees, no FMA contraction: build
 * with -ffp-contract=off) so that results are comparable bit-for-bit with a
 * HIP build that keeps the same order.  Citations are to
 * devel-izem/src/<file>:<line> of saras13015/Phd.
 *
 * Array layout (identical to the reference allocations
 * v(sx-ng:ex+ng, sy-ng:ey+ng, sz-ng:ez+ng, 1:nv), Fortran column-major):
 *   cell (i,j,k), i in [-ng, nx+ng), ... -> (i+ng) + NX*((j+ng) + NY*(k+ng))
 *   field l of a multi-field array at  l*ncell + cell
 * Interior index i = 0..nx-1 corresponds to Fortran sx..ex.
 * Ghost z-planes exist even in 2-D (the reference allocates them too).
 */
#ifndef IZEM_ORACLE_H
#define IZEM_ORACLE_H

#include <stdint.h>

/* species count: 9 (the H2/air mechanism) unless the build sets it (the
 * 2-species O2/N2 shock-tube case is built with -DORA_NRV=2) */
#ifndef ORA_NRV
#define ORA_NRV 9
#endif
#define ORA_NIV 5
#define ORA_NV (ORA_NIV + ORA_NRV)

/* ng = 3 (WENO5) or 4 (WENO7): ng = (ORDER_WENO + 1) / 2,
 * src/module_input.f90:712.  bc[face] for the faces W, E, S, N, B, F:
 * ORA_BC_PERIODIC, ORA_BC_EXTRAPOLATION (src/module_BCs.f90:41-141,
 * 327-412) or, on face S, ORA_BC_POSTSHOCK_SLIPWALL (:1065-1172).  nproc[d]: MPI ranks along direction d in the run being
 * restated (0 or 1: one); only the shock sensor depends on it, because its
 * sweep stops one cell short of each rank's end (src/module_tools.f90:
 * 2758-2770). */
#define ORA_BC_PERIODIC 0
#define ORA_BC_EXTRAPOLATION 1
#define ORA_BC_POSTSHOCK_SLIPWALL 2
#define ORA_BC_NSCBC_OUT 3
typedef struct {
  int nx, ny, nz, ng, ndim;
  int bc[6];
  int nproc[3];
  /* ORA_BC_POSTSHOCK_SLIPWALL: the conserved post-shock state (NV values)
   * and the cells of the face that take it */
  int nfix;
  const double *wall;
  /* ORA_BC_NSCBC_OUT: the domain length and the far-field pressure */
  double nscbc_Lx, nscbc_pinf;
} ora_grid;

/* thd_type subset used by the flow kernels (src/thd/type_thd.f90:10-91) */
typedef struct {
  double W[ORA_NRV];        /* kg/mol */
  double Wc[ORA_NRV];       /* W / Winf */
  double Wc_i[ORA_NRV];     /* 1 / Wc */
  double HeatCap[ORA_NRV][2][7]; /* [species][range 0=low 1=high][coef] */
  double Tpolymin[ORA_NRV], Tpolylim[ORA_NRV], Tpolymax[ORA_NRV];
  double CofD[ORA_NRV][ORA_NRV][4]; /* CofD(n,l,m) stored [l][m][n] */
  double CofLam[ORA_NRV][4];        /* CofLam(n,l) stored [l][n] */
  double CofEta[ORA_NRV][4];
  double wilke_c1[ORA_NRV][ORA_NRV]; /* (W(l)/W(m))**0.25 stored [m][l] */
  double wilke_c2[ORA_NRV][ORA_NRV]; /* 1/sqrt((1+W(m)/W(l))*8) stored [m][l] */
  double gam2, Tinf, p0, pinf, Dinf, laminf, muinf;
} ora_thd;

/* adi_type subset (src/module_adim.f90, values from src/module_tools.f90:453-562) */
typedef struct {
  double sqgmr, Sc, Pr, ggmopr, time_ref, T_ref, rho_ref, L_ref;
} ora_adi;

/* run-time switches of the path (src/module_parameters.f90:53-54) */
typedef struct {
  int weno_avg;       /* WENO%WEIGHTS */
  double max_rel_weight; /* WENO%PERCENT_WEIGHT */
  int vis, vis_dt, reaction, forcesum;
  double CFL, CFLmin, Fo, CPR;
  int sensor_type;      /* WENO%SHOCK_SENSOR_TYPE: 0 ADAMSSHARIFF, 1 DUCROS */
  double ducros_treduc; /* WENO%DUCROS_TREDUC */
} ora_opts;

/* CHEMKIN-II mechanism tables (the content of the linking file, see
 * src/chemkin/vode.f:391 CKINIT for the layout it replaces). */
#define ORA_KMAX 16
#define ORA_IMAX 48
#define ORA_TBMAX 12
typedef struct {
  int kk, ii;
  double wt[ORA_KMAX];                 /* g/mol */
  double a[ORA_KMAX][2][7];            /* NASA-7, range 0 = T <= Tmid */
  double tmid[ORA_KMAX];
  double A[ORA_IMAX], b[ORA_IMAX], EoR[ORA_IMAX];
  int nu[ORA_IMAX][6], nunk[ORA_IMAX][6]; /* 1-based species; slots 0..2 reactants (nu<0), 3..5 products */
  int reversible[ORA_IMAX];
  int ntb[ORA_IMAX];                   /* -1: no third body; else number of efficiencies */
  int tb_k[ORA_IMAX][ORA_TBMAX];       /* 1-based species */
  double tb_eff[ORA_IMAX][ORA_TBMAX];
  int fal[ORA_IMAX];                   /* 0 none, 1 Lindemann, 2 SRI, 3 Troe-3, 4 Troe-4 */
  double fpar[ORA_IMAX][8];            /* A0, b0, E0/R, then alpha, T***, T*, T** or a..e */
  int kfal[ORA_IMAX];                  /* falloff [M]: 0 third-body sum, k+1 species k */
  double rpar[ORA_IMAX][3];            /* REV parameters (reversible == 2) */
  double ru, patm;                     /* erg/(mol K), dyn/cm^2 */
  double small_, big_, exparg;
} ora_mech;

#ifdef __cplusplus
extern "C" {
#endif

/* thermo scalars */
double ora_wmix_i_scalar(const ora_thd *thd, const double *Y);
double ora_cp_scalar(const ora_thd *thd, double temp, const double *Y);
void ora_ha_scalar(const ora_thd *thd, double temp, double *ha);

/* WENO: 5th order on 6 values (ng = 3), 7th order on 8 values (ng = 4) */
void ora_wenorec(const double *vp, const double *vm, double *vminus, double *vplus);
void ora_wenorec_nw(const double *vp, const double *vm, double *vminus, double *vplus);
void ora_wenorec7(const double *vp, const double *vm, double *vminus, double *vplus);
void ora_wenorec7_nw(const double *vp, const double *vm, double *vminus, double *vplus);
double ora_weno7_beta(int r, const double *q);
/* first derivative of a line of n interior values (a[0] is interior point
 * 0, ghosts at negative indices and past n) at interior point p: central
 * 6th (ng = 3) or 8th (ng = 4) order, with the one-sided closures of a
 * physical boundary where lphys / rphys is set (src/module_deriv.f90:
 * 212-469) */
double ora_deriv_line(const double *a, long stride, double dinv, int p, int n, int ng,
                      int lphys, int rphys);
/* the conditions of the non-periodic faces, after ora_comm_cons
 * (upd_boundaries, src/module_BCs.f90:41-141); ora_upd_dirichlet then
 * updates the primitives of the rows a condition wrote (:131) */
void ora_upd_boundaries(const ora_grid *g, double *v);
int ora_upd_dirichlet(const ora_grid *g, const ora_thd *thd, int forcesum, double *v,
                      double *W_i, double *T, double *cp, double *ha);
void ora_bc_nscbc_out(const ora_grid *g, const ora_thd *thd, double dt, const double *dx_i,
                      const double *T, const double *W_i, const double *cp, double *v);
int ora_boundaries(const ora_grid *g, const ora_thd *thd, int forcesum, double dt,
                   const double *dx_i, double *v, double *W_i, double *T, double *cp,
                   double *ha);

void ora_eigvecs(int dir, const ora_thd *thd, const double *uc, const double *Yc,
                 const double *psi_c, double ht_c, double cs_c, double gam_c,
                 double el[ORA_NV][ORA_NV], double er[ORA_NV][ORA_NV]);
void ora_deriv7_line(int n, const double *a, double dinv, double *out);

/* flow-solver kernels; all return 0 or a negative status */
int ora_euler_llf(int dir, const ora_grid *g, const ora_thd *thd, const double *d_i,
                  const double *T, const double *W_i, const double *cp, const double *ha,
                  const uint8_t *sensor, int weno_avg, const double *v, double *fl);
void ora_wmix_i(int domain, const ora_grid *g, const ora_thd *thd, const double *v, double *W_i);
void ora_force_sumY(int domain, const ora_grid *g, double *v);
int ora_prim_inv_var(int domain, const ora_grid *g, const ora_thd *thd, double *v,
                     double *W_i, double *T, double *cp, double *ha, int *iters);
int ora_prim_inv_var_mach(int domain, const ora_grid *g, const ora_thd *thd, double *v,
                          double *W_i, double *T, double *cp, double *ha, int *iters);
void ora_upd_prim_var_woT(int domain, const ora_grid *g, const ora_thd *thd, const double *v,
                          const double *T, double *W_i, double *cp, double *ha);
void ora_comm_cons(const ora_grid *g, double *v);
int ora_upd_prim_var_ghost(const ora_grid *g, const ora_thd *thd, int forcesum, double *v,
                           double *W_i, double *T, double *cp, double *ha);
void ora_sensor_adams_shariff(const ora_grid *g, double max_rel_weight, const double *v,
                              const double *T, const double *W_i, uint8_t *sensor);
void ora_sensor_ducros(const ora_grid *g, double treduc, const double *dx_i, const double *dy_i,
                       const double *dz_i, const double *v, uint8_t *sensor);
void ora_transport_cell(const ora_thd *thd, double rho, double T, double W_i, const double *Xa,
                        double *dm, double *mu, double *ct);
int ora_viscflux(const ora_grid *g, const ora_thd *thd, const ora_adi *adi, const double *dx_i,
                 const double *dy_i, const double *dz_i, const double *T, const double *W_i,
                 const double *ha, const double *v, double *fl);
int ora_rk3(const ora_grid *g, const ora_thd *thd, const ora_adi *adi, const ora_opts *op,
            double dt, const double *dx_i, const double *dy_i, const double *dz_i,
            double *T, double *W_i, double *cp, double *ha, uint8_t *sensor, double *v,
            double *f_out);
void ora_timestep(const ora_grid *g, const ora_thd *thd, const ora_adi *adi, const ora_opts *op,
                  const double *dx_i, const double *dy_i, const double *dz_i, const double *cp,
                  const double *W_i, const double *T, const double *v, double DTmax,
                  double *dt_4, double *dtmin);

/* kinetics (cgs units, as CHEMKIN) */
void ora_ckwyr(const ora_mech *m, double rho, double T, const double *Y, double *wdot);
double ora_ckcvbs(const ora_mech *m, double T, const double *Y);
void ora_ckhms(const ora_mech *m, double T, double *hms);
void ora_reactor_rhs(const ora_mech *m, double rho, const double *z, double *zp);

#ifdef __cplusplus
}
#endif
#endif
