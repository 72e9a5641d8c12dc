/*
 * izem_hip.h -- C ABI of the MI355X (gfx950) implementation of the IZEM
 * reacting compressible Navier-Stokes right-hand side, SSP-RK3 time step
 * and operator-split stiff chemistry.
 *
 * Every array argument is a DEVICE pointer to FP64 data in the layout of
 * the reference's Fortran allocations (devel-izem/src/module_ICs.f90:2364):
 *     a(sx-ng:ex+ng, sy-ng:ey+ng, sz-ng:ez+ng[, 1:nf])
 * i.e. column-major, i fastest, ghost cells included, one field after the
 * other for multi-field arrays (v: nf = 14 = 5 + 9 species; ha: nf = 9).
 * Ghost z-planes exist in 2-D too.  The 1/dx arrays have n+2*ng entries.
 * A Fortran host passes c_loc(v(sx-ng,sy-ng,sz-ng,1)) of device-resident
 * arrays (see the ISO_C_BINDING module phd_amd/fortran/izem_hip_mod.f90).
 *
 * All functions enqueue on the context's HIP stream and return 0 on
 * success or a negative IZEM_E_* code; the physical events that the
 * reference aborts on (negative pressure, NaN, unrepairable negative
 * temperature, VODE failure) or warns about (a repaired negative
 * temperature, Newton non-convergence) are accumulated in a status word
 * read by izem_status().  A context is not thread-safe; one context per host
 * thread.  Citations are to saras13015/Phd devel-izem/src/<file>:<line>.
 */
#ifndef IZEM_HIP_H
#define IZEM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IZEM_OK 0
#define IZEM_E_ARG -1
#define IZEM_E_HIP -2
#define IZEM_E_NOMEM -3
#define IZEM_E_STATE -4

/* status bits (izem_status) */
#define IZEM_ST_NEG_PRESSURE 1  /* module_weno.f90:1027-1036 */
#define IZEM_ST_NEG_TEMPERATURE 2 /* repaired (a warning): module_thermodynamics.f90:120-157 */
#define IZEM_ST_NAN 4           /* module_thermodynamics.f90:180 */
#define IZEM_ST_NEWTON_ITMAX 8  /* module_thermodynamics.f90:177 */
#define IZEM_ST_VODE_FAIL 16    /* chemkin/reaction_v.f:59-62 */
#define IZEM_ST_NEG_T_ABORT 32  /* repair impossible (>40000 cells or at the array edge): :131 */

typedef struct izem_ctx izem_ctx;

/* local box (sx..ex etc. of module_parameters); ng = 3 (WENO5) for the default
 * library, 4 (WENO7) for the ng4 flow-only builds */
typedef struct {
  int nx, ny, nz, ng, ndim;
} izem_grid;

/* run-time switches of the path (module_parameters.f90:53-54,
 * module_input.f90:266-399) */
typedef struct {
  int weno_avg;          /* WENO%WEIGHTS */
  double max_rel_weight; /* WENO%PERCENT_WEIGHT */
  int vis, vis_dt, reaction, forcesum;
  double CFL, CFLmin, Fo, CPR;
  /* nondimensional constants (adi_type, module_tools.f90:453-562) */
  double sqgmr, Sc, Pr, ggmopr, time_ref, T_ref, rho_ref, L_ref;
  /* WENO%SHOCK_SENSOR_TYPE: 0 ADAMSSHARIFF, 1 DUCROS (module_solver.f90:758-762),
   * and WENO%DUCROS_TREDUC */
  int sensor_type;
  double ducros_treduc;
} izem_opts;

/* ---- context ---------------------------------------------------------- */
/* lengths (doubles) of the flat table images izem_create and
 * izem_chem_setup take */
#define IZEM_THD_IMAGE_LEN 745
#define IZEM_MECH_IMAGE_LEN 2850

/* face conditions (GRID%BCXMIN, BCXMAX, BCYMIN, BCYMAX, BCZMIN, BCZMAX;
 * module_BCs.f90:41-141): IZEM_BC_PERIODIC, the default, or
 * IZEM_BC_EXTRAPOLATION (bc_extrapolation, :327-412: the ng ghost layers
 * take the edge value, and the first derivatives of the viscous fluxes and
 * of the Ducros sensor take the one-sided closures of a face without a
 * neighbour, module_deriv.f90:268-469).  A direction is periodic on both
 * faces or on neither.  izem_comm_cons then fills extrapolated faces as
 * comm_cons followed by upd_boundaries does. */
#define IZEM_BC_PERIODIC 0
#define IZEM_BC_EXTRAPOLATION 1
/* bc_postshock_slipwall (module_BCs.f90:1065-1172), face S only: the first
 * nfix cells of the face (x(i) <= x0) take a post-shock state on the ng
 * ghost rows and on the first interior row; the others mirror the rows
 * above the first one, with rho v odd.  The state is given by
 * izem_set_wall_state; the first interior row's primitives are then
 * recomputed as the reference's upd_prim_var_dirichlet does (force_sumY and
 * prim_inv_var on that row, :93-96, 131, 297-314) before the ghost slabs.
 * Derivatives next to the face take the one-sided closures. */
#define IZEM_BC_POSTSHOCK_SLIPWALL 2
/* bc_nscbc_out (module_BCs.f90:2331-2516), face E only: a characteristic
 * subsonic outflow.  Per row, the fourth-order one-sided x derivatives of
 * rho, u, v, w, p and Y over the last five cells give the wave amplitudes
 * L1..L5; the incoming one is relaxed toward the far-field pressure,
 * L1 = 0.28 (1 - M^2) c / Lx (p - pinf); the last cell's state advanced by
 * -dt times the resulting changes is written to it and to its ghosts, and
 * its primitives recomputed (force_sumY, prim_inv_var).  Lx, pinf and 1/dx
 * of the last cell come from izem_set_nscbc; the step dt from the call
 * (izem_rk3's, the chemistry's scale x dt, izem_upd_boundaries'). */
#define IZEM_BC_NSCBC_OUT 3

/* thd: flat image of the thermodynamic/transport tables (thd_init,
 * thd/thd_tools_perfectgas.f90:12-248), see phd_amd.chemdata.Thd.flat(),
 * or izem_thd_image below */
int izem_create(izem_ctx **out, const izem_grid *grid, const double *thd, long nthd,
                const izem_opts *opts);
int izem_destroy(izem_ctx *ctx);
/* bc[6] in the order W, E, S, N, B, F (faces beyond ndim must be periodic) */
int izem_set_bc(izem_ctx *ctx, const int bc[6]);
/* the post-shock state of a IZEM_BC_POSTSHOCK_SLIPWALL face: 5 + nspc
 * conserved values (host memory) and the number nfix of its cells that take
 * it (phd_amd.shock.dmr_wall_state, dmr_wall_cells) */
int izem_set_wall_state(izem_ctx *ctx, const double *state, int nfix);
/* the characteristic outflow's constants (nondimensional): the domain
 * length Lx = x(ntx) - x(1) (module_tools.f90:319, 489), the far-field
 * pressure (94232.25 Pa in the reference, module_BCs.f90:2389) and 1/dx of
 * the last cell */
int izem_set_nscbc(izem_ctx *ctx, double Lx, double pinf, double dxi_last);
/* Reproduce on one GPU a run decomposed into nproc[0] x nproc[1] x nproc[2]
 * ranks (GRID%NUMPROCX/Y/Z, module_input.f90:361-363, 498-500; rank r of
 * np along a line of n cells owns cells r n / np .. (r + 1) n / np - 1,
 * module_parallel.f90:547-548).  The results of the reference's decomposed run depend on the
 * decomposition through the shock sensor alone: each rank's
 * shock_sensor_modified_adams_shariff sweep stops one short of its last
 * cell (module_tools.f90:2733-2814), flags are neither set on ghosts nor
 * exchanged, so a face sees the flags of the rank that forms it, and the
 * face between two ranks is formed by both, each with its own view
 * (module_weno.f90:955-1020 per rank).  {1, 1, 1} (the default) is the
 * undecomposed run; directions beyond ndim take 1, and a direction takes at
 * most one rank per cell. */
int izem_set_decomposition(izem_ctx *ctx, const int nproc[3]);
int izem_synchronize(izem_ctx *ctx);
/* run on a caller-owned stream (hipStream_t); NULL restores the context's */
int izem_set_stream(izem_ctx *ctx, void *stream);
/* OR of all status bits raised since the last call (synchronises) */
int izem_status(izem_ctx *ctx, int *bits);
/* device pointer to 6 doubles: dt_4(1:4), dtmin, DTmax [K] of the last
 * izem_timestep / izem_reaction_dvode (keeps dt resident on the device) */
double *izem_time_state(izem_ctx *ctx);
/* copy the 6 time-state doubles to host memory (synchronises) */
int izem_get_time_state(izem_ctx *ctx, double *host6);
/* overwrite the time state from host memory (dtmin seeds the chemical
 * limit of the next izem_timestep) */
int izem_set_time_state(izem_ctx *ctx, const double *host6);

/* event profiler: capacity > 0 starts recording up to capacity kernel
 * launches of the classes 0 Euler faces, 1 viscous, 2 RHS + RK update,
 * 3 Newton, 4 chemistry, 5 other; 0 stops.  izem_profile_read sums the
 * recorded milliseconds and launch counts per class (6 entries each) and
 * clears the record. */
int izem_profile(izem_ctx *ctx, int capacity);
int izem_profile_read(izem_ctx *ctx, double *ms, int *count);

/* ---- module_weno ------------------------------------------------------ */
/* wenorec (js = 1, module_weno.f90:286) or wenorec_nw (js = 0, :236) on n
 * stencils of 6 values; out = (vminus, vplus) pairs */
int izem_wenorec(izem_ctx *ctx, int n, int js, const double *vp, const double *vm,
                 double *out);
/* euler_LLF_x / _y / _z (module_weno.f90:537, 1055, 1595): x assigns fl,
 * y and z add to it; sensor: one byte per cell (logical shock_sensor) */
int izem_euler_llf_x(izem_ctx *ctx, const double *dx_i, const double *T, const double *W_i,
                     const double *cp, const double *ha, const uint8_t *sensor,
                     const double *v, double *fl);
int izem_euler_llf_y(izem_ctx *ctx, const double *dy_i, const double *T, const double *W_i,
                     const double *cp, const double *ha, const uint8_t *sensor,
                     const double *v, double *fl);
int izem_euler_llf_z(izem_ctx *ctx, const double *dz_i, const double *T, const double *W_i,
                     const double *cp, const double *ha, const uint8_t *sensor,
                     const double *v, double *fl);

/* ---- module_viscflux -------------------------------------------------- */
/* viscflux_xy (ndim 2) / viscflux_xyz (ndim 3) (module_viscflux.f90:721,
 * 1660), periodic faces; adds to fl */
int izem_viscflux(izem_ctx *ctx, const double *dx_i, const double *dy_i, const double *dz_i,
                  const double *T, const double *W_i, const double *ha, const double *v,
                  double *fl);

/* ---- module_thermodynamics / module_BCs / module_parallel ------------- */
/* domain: 0 interior, -1/1 x ghost slabs, -2/2 y, -3/3 z (domain_select,
 * module_parallel.f90:1936) */
int izem_force_sumY(izem_ctx *ctx, int domain, double *v);
/* prim_inv_var (module_thermodynamics.f90:79): Newton with the reference's
 * global exit test; iters (host, may be NULL) receives the iteration count */
int izem_prim_inv_var(izem_ctx *ctx, int domain, double *v, double *W_i, double *T,
                      double *cp, double *ha, int *iters);
/* upd_prim_var_woT (module_thermodynamics.f90:522) */
/* prim_inv_var_err_mach (module_thermodynamics.f90:201-260), the
 * inversion of initialisation and restart: as izem_prim_inv_var to 1e-10
 * and at most 100 iterations, with no repair (a negative temperature left
 * at the end raises IZEM_ST_NEG_TEMPERATURE, the reference's warning) */
int izem_prim_inv_var_mach(izem_ctx *ctx, int domain, double *v, double *W_i, double *T,
                           double *cp, double *ha, int *iters);
int izem_upd_prim_var_woT(izem_ctx *ctx, int domain, const double *v, const double *T,
                          double *W_i, double *cp, double *ha);
/* comm_cons on one rank (module_parallel.f90:1715-1777: the periodic wrap)
 * followed by upd_boundaries' conditions of the non-periodic faces
 * (module_BCs.f90:41-141), as the reference's time step calls them; a
 * context with a characteristic outflow face returns IZEM_E_STATE (its
 * condition needs the primitives and dt: izem_upd_boundaries) */
int izem_comm_cons(izem_ctx *ctx, double *v);
/* the same with the primitives and the step dt (device pointer) that the
 * characteristic outflow reads: comm_cons + upd_boundaries */
int izem_upd_boundaries(izem_ctx *ctx, const double *dt, const double *T, const double *W_i,
                        const double *cp, double *v);
/* comm_cons alone: the periodic wrap, non-periodic faces untouched (the
 * reference's initialisations call it before their own face treatment,
 * module_ICs.f90:404, 2872) */
int izem_comm_cons_exchange(izem_ctx *ctx, double *v);
/* upd_prim_var_ghost (module_BCs.f90:254), after the primitives of the
 * first interior rows that a post-shock slip wall or a characteristic
 * outflow wrote (upd_prim_var_dirichlet, module_BCs.f90:297-314; their
 * force_sumY came with the face condition) */
int izem_upd_prim_var_ghost(izem_ctx *ctx, double *T, double *W_i, double *cp, double *ha,
                            double *v);
/* shock_sensor_modified_adams_shariff (module_tools.f90:2733), sticky */
int izem_shock_sensor(izem_ctx *ctx, const double *v, const double *T, const double *W_i,
                      uint8_t *sensor);
/* shock_sensor_ducros with compute_duixj (module_tools.f90:2665-2719,
 * 2521-2651): flags an interior cell (sticky) where (div u)^2 / ((div u)^2 +
 * |curl u|^2 + 1e-30) > ducros_treduc; velocity gradients by the sixth-order
 * central stencil (also in z, where the reference's eighth-order operator
 * reads beyond an ng = 3 halo).  Selected in rk3 by sensor_type = 1. */
int izem_shock_sensor_ducros(izem_ctx *ctx, const double *dx_i, const double *dy_i,
                             const double *dz_i, const double *v, uint8_t *sensor);

/* ---- module_solver ---------------------------------------------------- */
/* rk3 (module_solver.f90:694): one SSP-RK3 step with the time step read
 * from the device pointer dt (nondimensional) */
int izem_rk3(izem_ctx *ctx, const double *dt, const double *dx_i, const double *dy_i,
             const double *dz_i, double *T, double *W_i, double *cp, double *ha,
             uint8_t *sensor, double *v);
/* right-hand side f of one stage (sensor, Euler x/y/z, viscous) without
 * the update; f has the layout of v */
int izem_rhs(izem_ctx *ctx, const double *dx_i, const double *dy_i, const double *dz_i,
             const double *T, const double *W_i, const double *cp, const double *ha,
             uint8_t *sensor, const double *v, double *f);
/* timestep (module_solver.f90:202): results in izem_time_state(ctx);
 * DTmax is taken from the time state (set by izem_reaction_dvode) and the
 * previous dtmin from slot 4 */
int izem_timestep(izem_ctx *ctx, const double *dx_i, const double *dy_i, const double *dz_i,
                  const double *cp, const double *W_i, const double *T, const double *v);

/* ---- table images from a reference host's own arrays (host code) ------ */
/* The thd image from the arrays of thd_type after thd_init
 * (thd/type_thd.f90:58-73 declares them; thd_tools_perfectgas.f90:12-248
 * fills them) and the reference quantities of the adi_type: nspc species
 * (the library's 9), ld the declared species dimension (Nspcmax_thd, 29),
 * HeatCap(7, 2, ld), CofD(4, ld, ld), CofLam(4, ld), CofEta(4, ld) in
 * Fortran order; gam2 = (gamma - 1) / gamma, Tinf = T_ref, p0 = 101325 Pa,
 * pinf = p_ref, Dinf = D_ref, laminf = lbda_ref, muinf = mu_ref
 * (module_tools.f90:453-562).  Wilke's factors are computed here
 * (module_thermodynamics.f90:1166-1224).  Returns the image length
 * (IZEM_THD_IMAGE_LEN) or IZEM_E_ARG. */
long izem_thd_image(int nspc, int ld, const double *W, const double *Wc, const double *Wc_i,
                    const double *HeatCap, const double *Tpolymin, const double *Tpolylim,
                    const double *Tpolymax, const double *CofD, const double *CofLam,
                    const double *CofEta, double gam2, double Tinf, double p0, double pinf,
                    double Dinf, double laminf, double muinf, double *image, long cap);
/* The mechanism image from CHEMKIN-II's work arrays after CKINIT
 * (chemkin/vode.f:391-790): ickwrk / rckwrk as CKINIT filled them and
 * ckstrt the /CKSTRT/ common block (chemkin/ckstrt.common), whose pointers
 * locate the tables.  IZEM_E_ARG for reaction forms the kernels do not
 * evaluate (Landau-Teller, real stoichiometric coefficients, explicit
 * orders, radiation wavelength, electron-impact, Janev, fit-1, excitation)
 * or tables beyond 16 species / 48 reactions / 12 third bodies.  Returns
 * the image length (IZEM_MECH_IMAGE_LEN). */
long izem_mech_image_ckwrk(const int *ickwrk, const double *rckwrk, const int *ckstrt,
                           double *image, long cap);

/* ---- module_reaction / chemkin ---------------------------------------- */
/* mechanism tables (the linking-file content read by CKINIT,
 * chemkin/vode.f:391), flat image of phd_amd.chemdata.Mechanism or of
 * izem_mech_image_ckwrk */
int izem_chem_setup(izem_ctx *ctx, const double *mech, long nmech, double rtol, double atol,
                    int trim_box);
/* parameters of the persistent chemistry kernels, as chosen by the
 * autotuner (phd_amd/autotune.py): register budget as waves per SIMD the
 * kernel variant is compiled for (1, 2 or 4), resident 64-lane blocks per
 * CU, Jacobian reuse age in BDF steps, kinetics form (0 table-driven,
 * 1 straight-line code generated for the bundled mechanism -- IZEM_E_ARG if
 * the loaded mechanism is another one or waves is not 1 --, 2 the same with
 * the closed-form mass-fraction columns of the Jacobian (same conditions),
 * -1 straight-line when it matches and waves is 1, table-driven otherwise).  No counterpart
 * in the reference (DVODE runs on the CPU). */
int izem_chem_tune(izem_ctx *ctx, int waves, int blocks_per_cu, int jac_age, int kinetics);
/* work counters of the chemistry integrations since the previous call:
 * out[0] BDF steps, out[1] RHS evaluations, out[2] Jacobian evaluations,
 * out[3] LU solves; synchronises and clears them.
 * Instrumentation without a reference counterpart (DVODE keeps the same
 * counts per call in IWORK(11..13), vode.f:4286). */
int izem_chem_stats(izem_ctx *ctx, unsigned long long out[4]);
/* ckwyr (chemkin/vode.f:1797) on n cells: rho [g/cm3], T [K], Y[n][kk] ->
 * wdot[n][kk] [mol/cm3/s] */
int izem_ckwyr(izem_ctx *ctx, int n, const double *rho, const double *T, const double *Y,
               double *wdot);
/* PSR_CHEMKIN constant-volume reactor (chemkin/psr_chemkin.f:1, reaction_v.f:1)
 * on n independent cells: T [K] in/out, rho [g/cm3], Y[n][kk] in/out, dt [s];
 * nsteps (may be NULL) receives the BDF step count of each cell */
int izem_psr_batch(izem_ctx *ctx, int n, double *T, const double *rho, double *Y, double dt,
                   int *nsteps);
/* izem_psr_batch that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The work
 * queue head and the stiffness order's buffers live in a static device
 * array of the library, one copy per GPU held until the process ends (about
 * 8 MB), located by izem_chem_setup; n may be at most IZEM_NOALLOC_MAX_CELLS
 * (IZEM_E_ARG above it).  Every call on a GPU shares that GPU's array, so
 * calls that overlap on one GPU (two streams, threads or contexts) race.
 * Results, step counts and work counters are those of izem_psr_batch. */
#define IZEM_NOALLOC_MAX_CELLS (1 << 20)
int izem_psr_batch_noalloc(izem_ctx *ctx, int n, double *T, const double *rho, double *Y,
                           double dt, int *nsteps);
/* order in which izem_psr_batch and izem_reaction_dvode hand cells to
 * lanes: 0 (the default) in index order; 1 stiffest first -- a key pass
 * evaluates max_k |df_k/dY_k|
 * of each cell (ten RHS evaluations) and counting-sorts the cells by it, so
 * the cells sharing a 64-lane wave have like step counts.  Pays where step
 * counts differ widely (ignition ensembles); where every cell takes a few
 * steps the key pass costs more than it saves.  Applied only to calls with
 * more cells than resident lanes (64 x CUs x blocks per CU): with fewer,
 * each lane takes one cell in any order.  Results are the same either
 * way.  The context keeps two int buffers of the largest cell count
 * ordered so far; a call that grows them synchronises the context stream
 * and allocates.  No counterpart in the reference (PSR_CHEMKIN integrates
 * one cell per call). */
int izem_chem_order(izem_ctx *ctx, int on);
/* the last stiffness sort of the context: *n = its cell count (0 when the
 * last izem_psr_batch / izem_reaction_dvode call ran unsorted), perm[q] =
 * cell at queue position q, bin[c] = stiffness slot of cell c (smaller =
 * stiffer) for the first *n entries (host arrays of at least cap entries;
 * IZEM_E_ARG if *n > cap).  Synchronises the context stream.  A diagnostic
 * with no counterpart in the reference. */
int izem_chem_order_read(izem_ctx *ctx, long cap, long *n, int *perm, int *bin);
/* reaction_selec with reac_selec = DVODE (module_reaction.f90:47, 102):
 * integrates every cell over scale*dt (Strang half step: scale = 0.5),
 * refreshes W_i/cp/ha, fills ghosts, ghost Newton; DTmax into the time state */
int izem_reaction_dvode(izem_ctx *ctx, const double *dt, double scale, double *T, double *W_i,
                        double *cp, double *ha, double *v);
/* reaction_selec (module_reaction.f90:47-90): the chemistry substep of the
 * solver chosen by izem_reac_selec (the reference's inp%reac_selec), then
 * the W_i/cp/ha refresh, ghost fill and ghost Newton of :79-83 */
#define IZEM_REAC_DVODE 0     /* reaction_dvode */
#define IZEM_REAC_SURROGATE 1 /* the attached surrogate, BDF on its rejected cells */
int izem_reac_selec(izem_ctx *ctx, int solver);
int izem_reaction_selec(izem_ctx *ctx, const double *dt, double scale, double *T, double *W_i,
                        double *cp, double *ha, double *v);
/* operation order of the Euler face kernels: 0 (the default) the
 * reference's dense characteristic projections in its summation order (bit
 * parity with it); 1 the projections through the structure of the left
 * eigenvectors (a shared scalar per stencil point; a tenth of the work,
 * results within rounding of mode 0).  No counterpart in the reference. */
int izem_flux_mode(izem_ctx *ctx, int mode);
/* one iteration of the main loop (izem.f90:242-389): timestep, then with
 * chemistry reaction_selec(dt/2) -> rk3(dt) -> reaction_selec(dt/2)
 * (:254-277), dt kept in the time state.  use_graph = 1 replays the
 * iteration from a HIP graph captured at the first call (and again whenever
 * the arrays or the chemistry settings change): same results, no launch
 * gaps.  IZEM_E_STATE with use_graph = 1 and the stiffness order on. */
int izem_step(izem_ctx *ctx, const double *dx_i, const double *dy_i, const double *dz_i,
              double *T, double *W_i, double *cp, double *ha, uint8_t *sensor, double *v,
              int use_graph);

/* ---- chemistry surrogate (no reference counterpart) -------------------
 * A BF16 MLP (12 -> 256 -> 256 -> 256 -> 10, GELU) stands in for the
 * constant-volume reactor substep of izem_psr_batch on the cells its error
 * indicator accepts; the others fall back to izem_psr_batch.  BF16 operands
 * are raw 16-bit arrays.  Matrices are row-major with leading dimensions in
 * elements. */
/* C[M][N] = A[M][K] . B[K][N] + epilogue on the matrix cores:
 * ta = 0: A row-major, 1: A stored [K][M]; tb = 0: B stored [N][K],
 * 1: B stored [K][N]; epi 0: C = AB (fp32 Cf), 1: Z = AB + bias (fp32 Cf)
 * and gelu(Z) (bf16 Cb), 2: Cf = AB + bias, 3: Cb = AB * gelu'(aux),
 * 5: Cb = gelu(AB + bias) only (inference) */
int izem_mlp_gemm(izem_ctx *ctx, int ta, int tb, int epi, int M, int N, int K, const void *A,
                  long lda, const void *B, long ldb, const float *bias, const float *aux,
                  long ldaux, float *Cf, long ldcf, void *Cb, long ldcb);
/* C[M][N] = A . B with the K range split over nslices blocks (weight
 * gradients, K = batch); work holds nslices * M * N floats */
int izem_mlp_gemm_splitk(izem_ctx *ctx, int ta, int tb, int M, int N, int K, const void *A,
                         long lda, const void *B, long ldb, int nslices, float *work, float *C);
/* out[n] = sum over rows of the bf16 matrix G[M][N] (bias gradients);
 * work holds nslices * N floats */
int izem_mlp_colsum(izem_ctx *ctx, int M, int N, const void *G, long ld, int nslices,
                    float *work, float *out);
/* Adam step `step` (>= 1) on the fp32 master matrix P[rows][cols]; writes
 * bf16 copies P^T (ldt) and P (ldw) where given (NULL skips) */
int izem_mlp_adam(izem_ctx *ctx, int rows, int cols, float *P, const float *G, float *m,
                  float *v, float lr, float b1, float b2, float eps, int step, void *Pt,
                  long ldt, void *Pw, long ldw);
int izem_mlp_to_bf16(izem_ctx *ctx, int rows, int cols, const float *P, void *Pt, long ldt,
                     void *Pw, long ldw);
/* mean-square loss gradient dY = 2 (pred - tgt) / (Bn nout), bf16, padded
 * columns zero; *loss (device) accumulates the sum of squares */
int izem_mlp_mse(izem_ctx *ctx, int Bn, int nout, int npad, const float *pred, long ldp,
                 const float *tgt, long ldt, void *dY, long ldd, float *loss);
/* the whole network on M cells in one kernel (inference): X[M][32] bf16
 * features, Wt[l] the bf16 weights of layer l stored [N][K] (K = 32, 256,
 * 256, 256; N = 256, 256, 256, 64), b[l] fp32 biases, pred[M][64] fp32
 * outputs; activations stay on chip */
int izem_mlp_infer(izem_ctx *ctx, int M, const void *X, const void *const *Wt,
                   const float *const *b, float *pred);
/* normalised bf16 features (ln p, T, Y_k, ln dt) of n cells, rows of 32 */
int izem_surr_features(izem_ctx *ctx, int n, const double *T, const double *rho, const double *Y,
                       double dt, const float *mean, const float *istd, void *X);
/* apply the network output pred[n][ldp] (dT, dY_k as standardised
 * asinh(change / oc): change = oc sinh(pred ostd + omean)): project dY onto
 * element conservation (proj: 9 x 9), T += dT; flag[c] = 1 (cell
 * untouched) when the indicator rejects it:
 * a feature beyond feat_lim, a mass fraction below -1e-7, |T(energy of the
 * projected Y) - (T0 + dT)| > tol_T or |dT| > dT_max; T_energy (may be
 * NULL) receives that energy temperature */
int izem_surr_apply(izem_ctx *ctx, int n, double *T, double *Y, const float *pred, long ldp,
                    const float *omean, const float *ostd, const float *oc, const double *proj,
                    const void *X,
                    float feat_lim, double tol_T, double dT_max, int *flag, double *T_energy);
/* the flagged cells through izem_psr_batch (device work arrays: counter[1],
 * idx[n], Tw[n], rw[n], Yw[9 n]); *nfall (host) receives their number */
int izem_surr_fallback(izem_ctx *ctx, int n, double *T, const double *rho, double *Y, double dt,
                       const int *flag, int *counter, int *idx, double *Tw, double *rw,
                       double *Yw, int *nfall);

/* ---- the surrogate inside the flow's reaction step (no reference
 * counterpart; the dispatch it joins is reaction_selec) ----------------
 * The caller's trained network, all device arrays, read at every reaction
 * call (they must outlive the attachment): */
typedef struct {
  const void *Wt[4];     /* bf16 weights of layer l stored [N][K], 16-byte aligned */
  const float *b[4];     /* fp32 biases (N = 256, 256, 256, 64) */
  const float *in_mean, *in_istd;            /* feature normalisation, 12 each */
  const float *out_mean, *out_std, *out_c;   /* output scaling, 10 each */
  const double *proj;    /* 9 x 9 element-conservation projector */
  float feat_lim;        /* indicator thresholds as izem_surr_apply */
  double tol_T, dT_max;  /* dT_max < 0 sends every cell to the BDF path */
} izem_surrogate;
/* attach s (copied; NULL detaches and selects DVODE) with a workspace for
 * chunk cells per network pass (128 <= chunk <= 2^26; 320 bytes each)
 * plus a 4-byte fallback-list entry per reacting cell */
int izem_surr_attach(izem_ctx *ctx, const izem_surrogate *s, long chunk);
/* reaction_selec's substep with the surrogate: on every cell of the
 * reacting box the network's substep over scale*dt where the indicator
 * accepts it, the reaction_dvode reactor kernel on the others; then the
 * tail of reaction_selec.  DTmax over both sets.  IZEM_E_STATE when no
 * surrogate is attached. */
int izem_reaction_surrogate(izem_ctx *ctx, const double *dt, double scale, double *T,
                            double *W_i, double *cp, double *ha, double *v);
/* cells through izem_reaction_surrogate and how many fell back to the BDF
 * since the previous call (synchronises the stream, then clears them) */
int izem_surr_counts(izem_ctx *ctx, unsigned long long out[2]);

#ifdef __cplusplus
}
#endif
#endif
