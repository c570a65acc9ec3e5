! hotcross_dropin.f90 - the reference's timestep subroutines, re-implemented as
! thin Fortran wrappers over the HIP library (include/hotcross.h).
!
! Each subroutine keeps the reference's name and argument list
! (hotcross80-revised2014.f:277-350, mpdata-hotcross80.f:16), so a driver
! written against the reference calls them unchanged; the state they act on
! lives in HBM inside the context held by module hotcross_dropin_state.
! Host module arrays are synchronised only at explicit points
! (hc_upload/hc_download), as the reference's output code would need.
! Errors stop the program with the reference's own messages.
module hotcross_dropin_state
  use iso_c_binding
  use hotcross_capi
  implicit none
  type(c_ptr), save :: hc = c_null_ptr
  type(hc_params_t), save :: prm
  integer, save :: mtime_ = 0, leap_ = 1, npatch_ = 0
  real(8), save :: time_ = 0
  real(8), allocatable, save :: t_bkg_(:), s_bkg_(:)
  ! host addresses of the tracer module arrays, to map TRACER31's dummy c
  type(c_ptr), save :: addr_t = c_null_ptr, addr_s = c_null_ptr, addr_c = c_null_ptr
  ! random draws of the last SOURCE_terms call (3 per patch cell)
  real(8), allocatable, target, save :: noise_(:)
contains
  subroutine hc_check(r, where)
    integer(c_int), intent(in) :: r
    character(len=*), intent(in) :: where
    if (r /= HC_OK) then
      write (6, *) 'hotcross: ', where, ': ', trim(hc_message(hc))
      if (r == HC_ECFL) stop 'stop on impending cfl violation'
      if (r == HC_ENAN) stop 'stop on NaN cfl violation'
      stop 'hotcross device error'
    endif
  end subroutine
end module hotcross_dropin_state

! time-loop control, hotcross80-revised2014.f:264-275
subroutine hc_set_step(m, time)
  use hotcross_dropin_state
  integer, intent(in) :: m
  real(8), intent(in) :: time
  mtime_ = m
  time_ = time
  call hc_check(hc_begin_step(hc, int(m, c_int32_t), time), 'begin_step')
end subroutine

subroutine body_force
  use hotcross_dropin_state
  call hc_check(hc_body_force(hc), 'BODY_FORCE')
end subroutine

subroutine forces
  use hotcross_dropin_state
  call hc_check(hc_forces(hc), 'FORCES')
end subroutine

subroutine pressure
  use hotcross_dropin_state
  call hc_check(hc_pressure(hc), 'PRESSURE')
end subroutine

subroutine momentum_u
  use hotcross_dropin_state
  call hc_check(hc_momentum_u(hc), 'MOMENTUM_U')
end subroutine

subroutine momentum_v
  use hotcross_dropin_state
  call hc_check(hc_momentum_v(hc), 'MOMENTUM_V')
end subroutine

subroutine momentum_w
  use hotcross_dropin_state
  call hc_check(hc_momentum_w(hc), 'MOMENTUM_W')
end subroutine

subroutine viscosity3d
  use hotcross_dropin_state
  call hc_check(hc_viscosity3d(hc), 'VISCOSITY3D')
end subroutine

! SOURCE_terms, source80.f:8-32: the random_number draws stay on the host
! (source80.inc:40-42, three per patch cell, j outer, i inner) and only the
! draws travel to the device
subroutine source_terms
  use hotcross_dropin_state
  integer :: q
  real(8) :: r
  do q = 1, 3 * npatch_
    call random_number(r)
    noise_(q) = r
  enddo
  call hc_check(hc_source_terms(hc, c_loc(noise_)), 'SOURCE_terms')
end subroutine

subroutine mixing
  use hotcross_dropin_state
  call hc_check(hc_mixing(hc), 'MIXING')
end subroutine

! work velocities 0.5*(up+u), hotcross80-revised2014.f:303-305
subroutine work_velocities(tau3)
  use hotcross_dropin_state
  real(8), intent(in) :: tau3
  call hc_check(hc_work_velocity(hc, tau3), 'work velocities')
end subroutine

! TRACER31, mpdata-hotcross80.f:16.  The dummy c is one of the module arrays
! t, s, c; its address selects the device field.  ucc/vcc/wcc are the work
! velocities already formed on the device by work_velocities.
subroutine tracer31(c, q_c, c_bkg, cmin, cmax, c_total0, c_end, ucc, vcc, wcc, tau3, output_unit_number)
  use hotcross_dropin_state
  real(8), target :: c(*)
  real(8) :: q_c(*), ucc(*), vcc(*), wcc(*)
  real(8), intent(in) :: c_bkg(*), cmin, cmax, tau3
  real(8), intent(inout) :: c_total0
  integer :: c_end(*), output_unit_number
  integer(c_int) :: field
  real(c_double) :: budgets(3)
  if (c_associated(c_loc(c(1)), addr_t)) then
    field = HC_T
  else if (c_associated(c_loc(c(1)), addr_s)) then
    field = HC_S
  else if (c_associated(c_loc(c(1)), addr_c)) then
    field = HC_C
  else
    stop 'TRACER31: c is not a registered tracer array'
  endif
  budgets = 0
  call hc_check(hc_tracer31(hc, field, c_bkg, cmin, cmax, tau3, budgets), 'TRACER31')
  if (mod(mtime_, 100) == 0) c_total0 = budgets(1)
end subroutine

! vent ghost cells after TRACER31, hotcross80-revised2014.f:311-315
subroutine vent_ghost(field, value)
  use hotcross_dropin_state
  integer(c_int), intent(in) :: field
  real(8), intent(in) :: value
  call hc_check(hc_vent_ghost(hc, field, value), 'vent ghost')
end subroutine

subroutine density
  use hotcross_dropin_state
  call hc_check(hc_density(hc), 'DENSITY')
end subroutine

subroutine refresh
  use hotcross_dropin_state
  call hc_check(hc_refresh(hc), 'REFRESH')
end subroutine

! CFL_VALUES, hotcross80-revised2014.f:374-438 (writes unit 81 as the
! reference does; the bailout stops with the reference's message)
subroutine cfl_values
  use hotcross_dropin_state
  real(c_double) :: cfl(6)
  integer(c_int) :: r
  r = hc_cfl_values(hc, cfl)
  write (81, '(6e15.7,i10)') cfl(1), cfl(2), cfl(3), cfl(4), cfl(5), cfl(6), mtime_
  call hc_check(r, 'CFL_VALUES')
end subroutine
