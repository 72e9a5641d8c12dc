! izem_fdriver.f90 -- a Fortran host of the library: the reference's set-up
! and one iteration of its main loop (devel-izem/src/izem.f90:180-277)
! written against izem_hip_mod only, as a reference host would call it.
! The tables come from the host's own arrays (thd_type after thd_init,
! CHEMKIN-II's work arrays after CKINIT), so no other front end is needed.
! Built into libizem_fortran.so; the bind(C) names let a test harness drive
! it.
module izem_fdriver
  use iso_c_binding
  use izem_hip_mod
  implicit none
  private
  public :: izem_f_create, izem_f_chem_init, izem_f_main_iteration

contains

  ! thd_init's arrays, declared (7, 2, ld), (4, ld, ld), ... as in
  ! thd/type_thd.f90:58-73; ref = gam2, T_ref, p_ref, D_ref, lbda_ref, mu_ref
  subroutine izem_f_create(ctx, grid, opts, nspc, ld, W, Wc, Wc_i, HeatCap, Tpolymin, &
                           Tpolylim, Tpolymax, CofD, CofLam, CofEta, ref) &
      bind(C, name="izem_f_create")
    type(c_ptr), intent(out) :: ctx
    type(izem_grid), intent(in) :: grid
    type(izem_opts), intent(in) :: opts
    integer(c_int), value :: nspc, ld
    real(c_double), intent(in) :: W(ld), Wc(ld), Wc_i(ld), HeatCap(7, 2, ld), Tpolymin(ld), &
                                  Tpolylim(ld), Tpolymax(ld), CofD(4, ld, ld), CofLam(4, ld), &
                                  CofEta(4, ld), ref(6)
    call hip_create(ctx, grid, opts, nspc, W, Wc, Wc_i, HeatCap, Tpolymin, Tpolylim, Tpolymax, &
                    CofD, CofLam, CofEta, ref(1), ref(2), ref(3), ref(4), ref(5), ref(6))
  end subroutine

  ! INITCHEMKIN's IWORK / RWORK after CKINIT (chemkin/initchemkin.f:25) and
  ! the /CKSTRT/ common block
  subroutine izem_f_chem_init(ctx, ickwrk, rckwrk, ckstrt, rtol, atol) &
      bind(C, name="izem_f_chem_init")
    type(c_ptr), value :: ctx
    integer(c_int), intent(in) :: ickwrk(*), ckstrt(*)
    real(c_double), intent(in) :: rckwrk(*)
    real(c_double), value :: rtol, atol
    call hip_chem_init(ctx, ickwrk, rckwrk, ckstrt, rtol, atol)
  end subroutine

  ! One iteration of the main loop (izem.f90:254-277): timestep, then the
  ! Strang split reaction(dt/2) -> rk3(dt) -> reaction(dt/2) when reacting,
  ! rk3(dt) alone otherwise; dt stays on the device.  Returns dtmin and
  ! DTmax (0 when inert) in host memory.
  subroutine izem_f_main_iteration(ctx, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v, reaction, &
                                   dtmin, DTmax) bind(C, name="izem_f_main_iteration")
    type(c_ptr), value :: ctx, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v
    integer(c_int), value :: reaction
    real(c_double), intent(out) :: dtmin, DTmax
    real(c_double) :: dt_4(4)
    type(c_ptr) :: dt
    call hip_timestep(ctx, dx_i, dy_i, dz_i, cp, W_i, T, v, dt_4, dtmin)
    dt = hip_dtmin(ctx)
    DTmax = 0.0_c_double
    if (reaction /= 0) then
      call hip_reaction_selec(ctx, dt, 0.5_c_double, T, W_i, cp, ha, v, DTmax)
      call hip_rk3(ctx, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v)
      call hip_reaction_selec(ctx, dt, 0.5_c_double, T, W_i, cp, ha, v, DTmax)
    else
      call hip_rk3(ctx, dt, dx_i, dy_i, dz_i, T, W_i, cp, ha, sensor, v)
    end if
  end subroutine

end module izem_fdriver
