! izem_hip_mod.f90 -- ISO_C_BINDING interface of the MI355X IZEM library
! (include/izem_hip.h) for a Fortran host.
!
! The reference keeps its state in allocatable arrays with ghost-inclusive
! lower bounds, e.g. v(sx-ng:ex+ng, sy-ng:ey+ng, sz-ng:ez+ng, 1:nv)
! (devel-izem/src/module_ICs.f90:2364-2369).  The library takes the same
! layout as DEVICE memory: a host built with OpenMP target / hipfort device
! arrays passes c_loc of the first element; a host whose arrays live in host
! memory copies them to device buffers once (hipMemcpy) and keeps them
! resident between calls, as the library does.
!
! Part 1 is the bind(C) interface block, one entry per C function.  Part 2
! are thin wrappers with the reference's subroutine names and argument
! meaning (module_solver.f90:202, 694; module_weno.f90:537, 1055, 1595;
! module_viscflux.f90:721, 1660; module_reaction.f90:47) that take the
! device addresses as type(c_ptr) and turn a nonzero status into the
! reference's abort_mpi (module_parallel.f90:200-211) message.
module izem_hip_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: IZEM_OK = 0, IZEM_E_ARG = -1, IZEM_E_HIP = -2, &
                                       IZEM_E_NOMEM = -3, IZEM_E_STATE = -4
  integer(c_int), parameter, public :: IZEM_ST_NEG_PRESSURE = 1, IZEM_ST_NEG_TEMPERATURE = 2, &
                                       IZEM_ST_NAN = 4, IZEM_ST_NEWTON_ITMAX = 8, &
                                       IZEM_ST_VODE_FAIL = 16, IZEM_ST_NEG_T_ABORT = 32

  type, bind(C), public :: izem_grid
    integer(c_int) :: nx, ny, nz, ng, ndim
  end type izem_grid

  type, bind(C), public :: izem_opts
    integer(c_int) :: weno_avg
    real(c_double) :: max_rel_weight
    integer(c_int) :: vis, vis_dt, reaction, forcesum
    real(c_double) :: CFL, CFLmin, Fo, CPR
    real(c_double) :: sqgmr, Sc, Pr, ggmopr, time_ref, T_ref, rho_ref, L_ref
    ! WENO%SHOCK_SENSOR_TYPE (0 ADAMSSHARIFF, 1 DUCROS), WENO%DUCROS_TREDUC
    integer(c_int) :: sensor_type
    real(c_double) :: ducros_treduc
  end type izem_opts

  public :: izem_create, izem_destroy, izem_synchronize, izem_set_stream, izem_status, &
            izem_time_state, izem_get_time_state, izem_set_time_state, izem_profile, &
            izem_profile_read, izem_wenorec, izem_euler_llf_x, izem_euler_llf_y, &
            izem_euler_llf_z, izem_viscflux, izem_force_sumY, izem_prim_inv_var, &
            izem_upd_prim_var_woT, izem_comm_cons, izem_comm_cons_exchange, &
            izem_upd_prim_var_ghost, &
            izem_shock_sensor, izem_rk3, izem_rhs, izem_timestep, izem_chem_setup, &
            izem_chem_tune, izem_chem_stats, izem_ckwyr, izem_psr_batch, izem_reaction_dvode, &
            izem_chem_order
  public :: izem_surr_features, izem_mlp_infer, izem_surr_apply, izem_surr_fallback
  public :: izem_thd_image, izem_mech_image_ckwrk, izem_reaction_selec, izem_reac_selec
  public :: izem_check, hip_rk3, hip_timestep, hip_euler_llf, hip_viscflux, hip_reaction_selec
  public :: hip_create, hip_chem_init, hip_dtmin

  ! face conditions of izem_set_bc (GRID%BCXMIN .. BCZMAX)
  integer(c_int), parameter, public :: IZEM_BC_PERIODIC = 0, IZEM_BC_EXTRAPOLATION = 1, &
                                       IZEM_BC_POSTSHOCK_SLIPWALL = 2, IZEM_BC_NSCBC_OUT = 3
  public :: izem_set_bc, izem_set_wall_state, izem_set_nscbc, izem_upd_boundaries
  ! ranks of a decomposed run to reproduce (GRID%NUMPROCX/Y/Z)
  public :: izem_set_decomposition

  ! lengths of the flat table images (doubles)
  integer(c_long), parameter, public :: IZEM_THD_IMAGE_LEN = 745, IZEM_MECH_IMAGE_LEN = 2850

  ! ------------------------------------------------------------ part 1
  interface
    integer(c_int) function izem_create(ctx, grid, thd, nthd, opts) bind(C, name="izem_create")
      import :: c_int, c_ptr, c_long, izem_grid, izem_opts
      type(c_ptr), intent(out) :: ctx
      type(izem_grid), intent(in) :: grid
      type(c_ptr), value :: thd
      integer(c_long), value :: nthd
      type(izem_opts), intent(in) :: opts
    end function
    integer(c_int) function izem_set_bc(ctx, bc) bind(C, name="izem_set_bc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: bc(6)
    end function
    integer(c_int) function izem_set_wall_state(ctx, state, nfix) &
        bind(C, name="izem_set_wall_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: state(*)
      integer(c_int), value :: nfix
    end function
    integer(c_int) function izem_set_nscbc(ctx, Lx, pinf, dxi_last) &
        bind(C, name="izem_set_nscbc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: Lx, pinf, dxi_last
    end function
    integer(c_int) function izem_upd_boundaries(ctx, dt, T, W_i, cp, v) &
        bind(C, name="izem_upd_boundaries")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dt, T, W_i, cp, v
    end function
    integer(c_int) function izem_set_decomposition(ctx, nproc) &
        bind(C, name="izem_set_decomposition")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(in) :: nproc(3)
    end function
    integer(c_int) function izem_destroy(ctx) bind(C, name="izem_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function izem_synchronize(ctx) bind(C, name="izem_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function izem_set_stream(ctx, stream) bind(C, name="izem_set_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream
    end function
    integer(c_int) function izem_status(ctx, bits) bind(C, name="izem_status")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: bits
    end function
    type(c_ptr) function izem_time_state(ctx) bind(C, name="izem_time_state")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function izem_get_time_state(ctx, host6) bind(C, name="izem_get_time_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: host6(6)
    end function
    integer(c_int) function izem_set_time_state(ctx, host6) bind(C, name="izem_set_time_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: host6(6)
    end function
    integer(c_int) function izem_profile(ctx, capacity) bind(C, name="izem_profile")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: capacity
    end function
    integer(c_int) function izem_profile_read(ctx, ms, count) bind(C, name="izem_profile_read")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(6)
      integer(c_int), intent(out) :: count(6)
    end function
    integer(c_int) function izem_wenorec(ctx, n, js, vp, vm, out) bind(C, name="izem_wenorec")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, vp, vm, out
      integer(c_int), value :: n, js
    end function
    integer(c_int) function izem_euler_llf_x(ctx, d_i, T, W_i, cp, ha, sensor, v, fl) &
        bind(C, name="izem_euler_llf_x")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_i, T, W_i, cp, ha, sensor, v, fl
    end function
    integer(c_int) function izem_euler_llf_y(ctx, d_i, T, W_i, cp, ha, sensor, v, fl) &
        bind(C, name="izem_euler_llf_y")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_i, T, W_i, cp, ha, sensor, v, fl
    end function
    integer(c_int) function izem_euler_llf_z(ctx, d_i, T, W_i, cp, ha, sensor, v, fl) &
        bind(C, name="izem_euler_llf_z")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_i, T, W_i, cp, ha, sensor, v, fl
    end function
    integer(c_int) function izem_viscflux(ctx, dx_i, dy_i, dz_i, T, W_i, ha, v, fl) &
        bind(C, name="izem_viscflux")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dx_i, dy_i, dz_i, T, W_i, ha, v, fl
    end function
    integer(c_int) function izem_force_sumY(ctx, domain, v) bind(C, name="izem_force_sumY")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, v
      integer(c_int), value :: domain
    end function
    integer(c_int) function izem_prim_inv_var(ctx, domain, v, W_i, T, cp, ha, iters) &
        bind(C, name="izem_prim_inv_var")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, v, W_i, T, cp, ha, iters
      integer(c_int), value :: domain
    end function
    integer(c_int) function izem_upd_prim_var_woT(ctx, domain, v, T, W_i, cp, ha) &
        bind(C, name="izem_upd_prim_var_woT")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, v, T, W_i, cp, ha
      integer(c_int), value :: domain
    end function
    integer(c_int) function izem_comm_cons(ctx, v) bind(C, name="izem_comm_cons")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, v
    end function
    integer(c_int) function izem_comm_cons_exchange(ctx, v) &
        bind(C, name="izem_comm_cons_exchange")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, v
    end function
    integer(c_int) function izem_upd_prim_var_ghost(ctx, T, W_i, cp, ha, v) &
        bind(C, name="izem_upd_prim_var_ghost")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, T, W_i, cp, ha, v
    end function
    integer(c_int) function izem_shock_sensor(ctx, v, T, W_i, sensor) &
        bind(C, name="izem_shock_sensor")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, v, T, W_i, sensor
    end function
    integer(c_int) function izem_rk3(ctx, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v) &
        bind(C, name="izem_rk3")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v
    end function
    integer(c_int) function izem_rhs(ctx, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v, f) &
        bind(C, name="izem_rhs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v, f
    end function
    integer(c_int) function izem_timestep(ctx, dx_i, dy_i, dz_i, cp, W_i, T, v) &
        bind(C, name="izem_timestep")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dx_i, dy_i, dz_i, cp, W_i, T, v
    end function
    integer(c_int) function izem_chem_setup(ctx, mech, nmech, rtol, atol, trim_box) &
        bind(C, name="izem_chem_setup")
      import :: c_int, c_ptr, c_long, c_double
      type(c_ptr), value :: ctx, mech
      integer(c_long), value :: nmech
      real(c_double), value :: rtol, atol
      integer(c_int), value :: trim_box
    end function
    ! table images from the host's own arrays: thd_type after thd_init,
    ! CHEMKIN-II's work arrays after CKINIT and its /CKSTRT/ pointers
    integer(c_long) function izem_thd_image(nspc, ld, W, Wc, Wc_i, HeatCap, Tpolymin, &
                                            Tpolylim, Tpolymax, CofD, CofLam, CofEta, gam2, &
                                            Tinf, p0, pinf, Dinf, laminf, muinf, image, cap) &
        bind(C, name="izem_thd_image")
      import :: c_int, c_long, c_double
      integer(c_int), value :: nspc, ld
      real(c_double), intent(in) :: W(*), Wc(*), Wc_i(*), HeatCap(*), Tpolymin(*), &
                                    Tpolylim(*), Tpolymax(*), CofD(*), CofLam(*), CofEta(*)
      real(c_double), value :: gam2, Tinf, p0, pinf, Dinf, laminf, muinf
      real(c_double), intent(out) :: image(*)
      integer(c_long), value :: cap
    end function
    integer(c_long) function izem_mech_image_ckwrk(ickwrk, rckwrk, ckstrt, image, cap) &
        bind(C, name="izem_mech_image_ckwrk")
      import :: c_int, c_long, c_double
      integer(c_int), intent(in) :: ickwrk(*), ckstrt(*)
      real(c_double), intent(in) :: rckwrk(*)
      real(c_double), intent(out) :: image(*)
      integer(c_long), value :: cap
    end function
    integer(c_int) function izem_chem_tune(ctx, waves, blocks_per_cu, jac_age, kinetics) &
        bind(C, name="izem_chem_tune")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: waves, blocks_per_cu, jac_age, kinetics
    end function
    integer(c_int) function izem_chem_order(ctx, on) bind(C, name="izem_chem_order")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function izem_chem_stats(ctx, out) bind(C, name="izem_chem_stats")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: out(4)
    end function
    integer(c_int) function izem_ckwyr(ctx, n, rho, T, Y, wdot) bind(C, name="izem_ckwyr")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rho, T, Y, wdot
      integer(c_int), value :: n
    end function
    integer(c_int) function izem_psr_batch(ctx, n, T, rho, Y, dt, nsteps) &
        bind(C, name="izem_psr_batch")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, T, rho, Y, nsteps
      integer(c_int), value :: n
      real(c_double), value :: dt
    end function
    integer(c_int) function izem_reaction_dvode(ctx, dt, scale, T, W_i, cp, ha, v) &
        bind(C, name="izem_reaction_dvode")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dt, T, W_i, cp, ha, v
      real(c_double), value :: scale
    end function
    ! reaction_selec (module_reaction.f90:47): the solver izem_reac_selec
    ! chose (0 DVODE, 1 the attached surrogate)
    integer(c_int) function izem_reaction_selec(ctx, dt, scale, T, W_i, cp, ha, v) &
        bind(C, name="izem_reaction_selec")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, dt, T, W_i, cp, ha, v
      real(c_double), value :: scale
    end function
    integer(c_int) function izem_reac_selec(ctx, solver) bind(C, name="izem_reac_selec")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: solver
    end function
    ! chemistry surrogate (no reference counterpart): apply a trained network
    ! to n cells, with the BDF fallback for the flagged ones
    integer(c_int) function izem_surr_features(ctx, n, T, rho, Y, dt, mean, istd, X) &
        bind(C, name="izem_surr_features")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, T, rho, Y, mean, istd, X
      integer(c_int), value :: n
      real(c_double), value :: dt
    end function
    integer(c_int) function izem_mlp_infer(ctx, M, X, Wt, b, pred) bind(C, name="izem_mlp_infer")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, X, pred
      integer(c_int), value :: M
      type(c_ptr), intent(in) :: Wt(4), b(4)
    end function
    integer(c_int) function izem_surr_apply(ctx, n, T, Y, pred, ldp, omean, ostd, oc, proj, X, &
                                            feat_lim, tol_T, dT_max, flag, T_energy) &
        bind(C, name="izem_surr_apply")
      import :: c_int, c_ptr, c_long, c_float, c_double
      type(c_ptr), value :: ctx, T, Y, pred, omean, ostd, oc, proj, X, flag, T_energy
      integer(c_int), value :: n
      integer(c_long), value :: ldp
      real(c_float), value :: feat_lim
      real(c_double), value :: tol_T, dT_max
    end function
    integer(c_int) function izem_surr_fallback(ctx, n, T, rho, Y, dt, flag, counter, idx, Tw, rw, &
                                               Yw, nfall) bind(C, name="izem_surr_fallback")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, T, rho, Y, flag, counter, idx, Tw, rw, Yw, nfall
      integer(c_int), value :: n
      real(c_double), value :: dt
    end function
  end interface

contains

  ! ------------------------------------------------------------ part 2
  ! A nonzero return or an aborting status bit ends the run with the
  ! reference's messages (abort_mpi in the reference; error stop here); a
  ! repaired negative temperature and Newton non-convergence are warnings
  ! there (module_thermodynamics.f90:126, 177) and here.
  subroutine izem_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    integer(c_int) :: bits, r2
    if (rc /= IZEM_OK) then
      write (*, '(a,a,a,i0)') 'izem_hip: ', where, ' failed with code ', rc
      error stop 1
    end if
    r2 = izem_status(ctx, bits)
    if (r2 /= IZEM_OK .or. bits == 0) return
    if (iand(bits, IZEM_ST_NEG_TEMPERATURE) /= 0) &
      write (*, '(a,a)') 'WARNING: negative temperature repaired after ', where
    if (iand(bits, IZEM_ST_NEWTON_ITMAX) /= 0) &
      write (*, '(a,a)') 'WARNING: failure in temperatures convergence after ', where
    if (iand(bits, ior(ior(IZEM_ST_NEG_PRESSURE, IZEM_ST_NAN), &
                       ior(IZEM_ST_VODE_FAIL, IZEM_ST_NEG_T_ABORT))) == 0) return
    if (iand(bits, IZEM_ST_NEG_PRESSURE) /= 0) write (*, '(a)') 'negative pressure'
    if (iand(bits, IZEM_ST_NEG_T_ABORT) /= 0) write (*, '(a)') 'too many points with negative temperature'
    if (iand(bits, IZEM_ST_NAN) /= 0) write (*, '(a)') 'error is a NaN'
    if (iand(bits, IZEM_ST_VODE_FAIL) /= 0) write (*, '(a)') 'DVODE failure (ISTATE <= -2)'
    write (*, '(a,a)') 'izem_hip: abort after ', where
    error stop 2
  end subroutine

  ! rk3 (module_solver.f90:694): dt is the device address of the
  ! nondimensional step (slot 5 of izem_time_state after hip_timestep)
  subroutine hip_rk3(ctx, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, shock_sensor, v)
    type(c_ptr), intent(in) :: ctx, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, shock_sensor, v
    call izem_check(ctx, izem_rk3(ctx, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, shock_sensor, &
                                  v), 'rk3')
  end subroutine

  ! timestep (module_solver.f90:202): dt_4 and dtmin returned in host memory
  subroutine hip_timestep(ctx, dx_i, dy_i, dz_i, cp, W_i, T, v, dt_4, dtmin)
    type(c_ptr), intent(in) :: ctx, dx_i, dy_i, dz_i, cp, W_i, T, v
    real(c_double), intent(out) :: dt_4(4), dtmin
    real(c_double) :: ts(6)
    call izem_check(ctx, izem_timestep(ctx, dx_i, dy_i, dz_i, cp, W_i, T, v), 'timestep')
    call izem_check(ctx, izem_get_time_state(ctx, ts), 'timestep')
    dt_4 = ts(1:4)
    dtmin = ts(5)
  end subroutine

  ! euler_LLF_x/_y/_z (module_weno.f90:537, 1055, 1595), direction 1..3
  subroutine hip_euler_llf(ctx, direction, d_i, T, W_i, cp, ha, shock_sensor, v, fl)
    type(c_ptr), intent(in) :: ctx, d_i, T, W_i, cp, ha, shock_sensor, v, fl
    integer, intent(in) :: direction
    select case (direction)
    case (1)
      call izem_check(ctx, izem_euler_llf_x(ctx, d_i, T, W_i, cp, ha, shock_sensor, v, fl), &
                      'euler_LLF_x')
    case (2)
      call izem_check(ctx, izem_euler_llf_y(ctx, d_i, T, W_i, cp, ha, shock_sensor, v, fl), &
                      'euler_LLF_y')
    case default
      call izem_check(ctx, izem_euler_llf_z(ctx, d_i, T, W_i, cp, ha, shock_sensor, v, fl), &
                      'euler_LLF_z')
    end select
  end subroutine

  ! viscflux_xy / viscflux_xyz (module_viscflux.f90:721, 1660)
  subroutine hip_viscflux(ctx, dx_i, dy_i, dz_i, T, W_i, ha, v, fl)
    type(c_ptr), intent(in) :: ctx, dx_i, dy_i, dz_i, T, W_i, ha, v, fl
    call izem_check(ctx, izem_viscflux(ctx, dx_i, dy_i, dz_i, T, W_i, ha, v, fl), 'viscflux')
  end subroutine

  ! reaction_selec (module_reaction.f90:47), DVODE unless izem_reac_selec
  ! chose the surrogate: one Strang half step of dt (device address) when
  ! scale = 0.5; DTmax [K] returned in host memory
  subroutine hip_reaction_selec(ctx, dt, scale, T, W_i, cp, ha, v, DTmax)
    type(c_ptr), intent(in) :: ctx, dt, T, W_i, cp, ha, v
    real(c_double), intent(in) :: scale
    real(c_double), intent(out) :: DTmax
    real(c_double) :: ts(6)
    call izem_check(ctx, izem_reaction_selec(ctx, dt, scale, T, W_i, cp, ha, v), &
                    'reaction_selec')
    call izem_check(ctx, izem_get_time_state(ctx, ts), 'reaction_selec')
    DTmax = ts(6)
  end subroutine

  ! device address of dtmin (slot 5 of izem_time_state): the dt that
  ! hip_rk3 and hip_reaction_selec take, kept on the device
  type(c_ptr) function hip_dtmin(ctx)
    type(c_ptr), intent(in) :: ctx
    integer(c_intptr_t) :: a
    a = transfer(izem_time_state(ctx), a)
    hip_dtmin = transfer(a + 4_c_intptr_t * c_sizeof(0.0_c_double), hip_dtmin)
  end function

  ! izem_create from the tables thd_init filled (thd/type_thd.f90:58-73,
  ! declared with the species dimension Nspcmax_thd) and the reference
  ! quantities of adi_type (module_tools.f90:453-562): gam2 = (gamma-1)/gamma
  subroutine hip_create(ctx, grid, opts, nspc, W, Wc, Wc_i, HeatCap, Tpolymin, Tpolylim, &
                        Tpolymax, CofD, CofLam, CofEta, gam2, T_ref, p_ref, D_ref, lbda_ref, &
                        mu_ref)
    type(c_ptr), intent(out) :: ctx
    type(izem_grid), intent(in) :: grid
    type(izem_opts), intent(in) :: opts
    integer(c_int), intent(in) :: nspc
    real(c_double), intent(in) :: W(:), Wc(:), Wc_i(:), HeatCap(:, :, :), Tpolymin(:), &
                                  Tpolylim(:), Tpolymax(:), CofD(:, :, :), CofLam(:, :), &
                                  CofEta(:, :)
    real(c_double), intent(in) :: gam2, T_ref, p_ref, D_ref, lbda_ref, mu_ref
    real(c_double), target :: image(IZEM_THD_IMAGE_LEN)
    integer(c_long) :: n
    if (size(HeatCap, 1) /= 7 .or. size(HeatCap, 2) /= 2 .or. size(CofD, 1) /= 4 .or. &
        size(CofD, 2) /= size(W) .or. size(CofLam, 1) /= 4 .or. size(CofEta, 1) /= 4) then
      write (*, '(a)') 'izem_hip: thd arrays are not shaped as thd_type declares them'
      error stop 1
    end if
    n = izem_thd_image(nspc, int(size(W), c_int), W, Wc, Wc_i, HeatCap, Tpolymin, Tpolylim, &
                       Tpolymax, CofD, CofLam, CofEta, gam2, T_ref, 101325.0_c_double, p_ref, &
                       D_ref, lbda_ref, mu_ref, image, IZEM_THD_IMAGE_LEN)
    if (n /= IZEM_THD_IMAGE_LEN) call izem_check(c_null_ptr, int(n, c_int), 'thd_image')
    ctx = c_null_ptr
    call izem_check(ctx, izem_create(ctx, grid, c_loc(image), n, opts), 'izem_create')
  end subroutine

  ! izem_chem_setup from CKINIT's work arrays (chemkin/vode.f:391) and the
  ! /CKSTRT/ pointers (ckstrt: its 85 integers, e.g. an array equivalenced
  ! to NMM)
  subroutine hip_chem_init(ctx, ickwrk, rckwrk, ckstrt, rtol, atol)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: ickwrk(*), ckstrt(*)
    real(c_double), intent(in) :: rckwrk(*)
    real(c_double), intent(in) :: rtol, atol
    real(c_double), target :: image(IZEM_MECH_IMAGE_LEN)
    integer(c_long) :: n
    n = izem_mech_image_ckwrk(ickwrk, rckwrk, ckstrt, image, IZEM_MECH_IMAGE_LEN)
    if (n /= IZEM_MECH_IMAGE_LEN) call izem_check(ctx, int(n, c_int), 'mech_image_ckwrk')
    call izem_check(ctx, izem_chem_setup(ctx, c_loc(image), n, rtol, atol, 1_c_int), &
                    'izem_chem_setup')
  end subroutine

end module izem_hip_mod
