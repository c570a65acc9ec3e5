! hotcross_driver.f90 - a Fortran driver with the reference's time-loop call
! sequence (hotcross80-revised2014.f:264-358), running the timestep on the
! MI355X through the drop-in subroutines of hotcross_dropin.f90.
!
! usage: hotcross_driver case.nml out.bin
!   case.nml holds namelist /case/ (the variables of the reference's case file
!   and the workload's profile); out.bin receives, after the run, the state
!   fields in the order u v w up vp wp um vm wm p t s c ss ss_v rhop (Fortran
!   order over each field's bounds, FP64) followed by the source draws of every
!   step.
program hotcross_driver
  use iso_c_binding
  use hotcross_capi
  use hotcross_dropin_state
  implicit none
  integer :: nx = 16, ny = 12, nz = 10, nsteps = 10, ndata, m, i, j, k, u_out, nnx, nny
  real(8) :: dx = 1, dy = 1, dz = 1, dt = 0.25d0, t_noise = 1.0d-4
  real(8) :: zstart = 1750, tmin, smax, time
  real(8), allocatable :: zd(:), td(:), sd(:), buf(:), allnoise(:,:)
  real(8), allocatable, target :: t(:,:,:), s(:,:,:), c(:,:,:), q_t(:,:,:)
  integer, allocatable :: p_end(:,:)
  real(8) :: t_total = 0, s_total = 0, c_total = 0, c_bkg0(0:4095)
  integer(c_int32_t) :: lo(3), hi(3)
  integer :: fld
  character(len=256) :: case_file, out_file
  namelist /case/ nx, ny, nz, nsteps, dx, dy, dz, dt, zstart, t_noise

  call get_command_argument(1, case_file)
  call get_command_argument(2, out_file)
  open (unit=10, file=trim(case_file), status='old')
  read (10, nml=case)
  close (10)

  ! case parameters: the reference's constants and a Dante-like set of
  ! switches (initialize_constants, modules80-revised2014.f:51-60)
  prm%nx = nx; prm%ny = ny; prm%nz = nz; prm%nt = nsteps
  prm%dx = dx; prm%dy = dy; prm%dz = dz; prm%dt = dt
  prm%zstart = zstart; prm%zend = zstart + nz * dz
  prm%bit2 = 1; prm%bit3 = 1; prm%bit4 = 1; prm%bit5 = 1; prm%bit7 = 1; prm%bit8 = 0; prm%bit9 = 0
  prm%bit10 = 1; prm%bit2k = 1; prm%bit5k = 1; prm%ismol = 1; prm%ismag_h = 1; prm%ismag_v = 1
  prm%bcondx = 1; prm%bcondy = 1; prm%bottom_slip = 1; prm%side_slip = 1; prm%open_box = 1
  prm%bdbit = 1; prm%modulo_x = 0; prm%modulo_y = 0
  prm%toggle_t = 1; prm%toggle_s = 1; prm%toggle_c = 1
  prm%cs_h = 0.2d0; prm%cs_v = 0.2d0; prm%prandtl = 1; prm%axmin = 1.0d-4; prm%aymin = 1.0d-4
  prm%azmin = 1.0d-5; prm%azmax = 1.0d-3; prm%mixing_height_scale = 3; prm%sponge_factor = 0.05d0
  prm%lat = 47.95d0; prm%g = 9.81d0; prm%rho0 = 1027.65d0; prm%assel = 0.01d0; prm%eps = 1.0d-10
  prm%omega = 2.0d0 * (4.0d0 * atan2(1.0d0, 1.0d0)) / 24.0d0 / 3600.0d0
  prm%source_length = 4 * dx; prm%source_width = 4.5d0 * dy
  prm%dqtdt = 50.0d6 / (1027.65d0 * 4200.0d0) / 20.0d0; prm%dqsdt = -1.0d-3; prm%dqcdt = 1
  prm%tsource = 0.5d0; prm%ssource = -0.01d0; prm%wsource = 0.01d0
  prm%source_ramp_hrs = 2; prm%noise_amp = 0.025d0; prm%wramp_hrs = 6
  prm%flow_u0 = 0.032d0; prm%flow_u1 = 0.036d0; prm%flow_v1 = 0.0024d0
  prm%flow_omega = 2.0d0 * (4.0d0 * atan2(1.0d0, 1.0d0)) / 12.42d0 / 3600.0d0; prm%flow_ramp_hrs = 2
  prm%eos = 0; prm%eos_rho_ref = 1027.65d0

  call hc_check(hc_create(prm, HC_FP64, hc), 'hc_create')

  ! INITIALIZE_PROFILES input: the linear table of the reference's
  ! site_stratification_linear.inc (header lines 2-3), on its 2 m grid
  ndata = 220
  allocate (zd(ndata), td(ndata), sd(ndata))
  do i = 1, ndata
    zd(i) = 1690.0d0 + 2.0d0 * (i - 1)
    td(i) = 1.673d0 + 6.666d-4 * (2101.0d0 - zd(i))
    sd(i) = 34.612d0 - 1.135d-4 * (2101.0d0 - zd(i))
  enddo
  call hc_check(hc_init_case(hc, zd, td, sd, int(ndata, c_int32_t), t_noise, 7_c_int64_t), 'init')

  ! host mirrors of the tracer module arrays (registry for TRACER31) and of
  ! the background profiles
  allocate (t(0:nx+1,0:ny+1,0:nz+1), s(0:nx+1,0:ny+1,0:nz+1), c(0:nx+1,0:ny+1,0:nz+1))
  allocate (q_t(0:nx+1,0:ny+1,0:nz+1), p_end(0:nx+1,0:ny+1))
  addr_t = c_loc(t); addr_s = c_loc(s); addr_c = c_loc(c)
  allocate (t_bkg_(0:nz+1), s_bkg_(0:nz+1))
  lo = [0, 0, 0]; hi = [nx+1, ny+1, nz+1]
  call hc_check(hc_download(hc, HC_T, c_loc(t), lo, hi), 'download')
  call hc_check(hc_download(hc, HC_S, c_loc(s), lo, hi), 'download')
  ! background = the initial fields at the first column (tracers start at
  ! background, initialize_profiles80.f:141-153); T noise excluded by using S
  ! and the profile table directly
  call profile_from_table(nz, zstart, dz, zd, td, sd, ndata, t_bkg_, s_bkg_)
  nnx = nint(prm%source_length / dx); nny = nint(prm%source_width / dy)
  npatch_ = nnx * nny
  allocate (noise_(3 * npatch_), allnoise(3 * npatch_, nsteps))
  open (unit=81, file='cfl.out')

  time = -dt
  p_end = nz
  do m = 0, nsteps - 1
    time = time + dt
    call hc_set_step(m, time)
    call body_force
    call forces
    call pressure
    call momentum_u
    call momentum_v
    call momentum_w
    call viscosity3d
    call source_terms
    allnoise(:, m + 1) = noise_
    call mixing
    call work_velocities(dt)
    tmin = minval(t_bkg_)
    call tracer31(t, q_t, t_bkg_, tmin, 1.0d10, t_total, p_end, t, t, t, dt, 401)
    call vent_ghost(HC_T, prm%tsource + t_bkg_(nz))
    smax = maxval(s_bkg_)
    call tracer31(s, q_t, s_bkg_, 0.0d0, smax, s_total, p_end, t, t, t, dt, 402)
    call vent_ghost(HC_S, prm%ssource + s_bkg_(nz))
    call density
    c_bkg0 = 0
    call tracer31(c, q_t, c_bkg0, 0.0d0, 1.0d10, c_total, p_end, t, t, t, dt, 403)
    call refresh
    if (mod(m, 10) == 0) call cfl_values
  enddo

  ! state out
  open (newunit=u_out, file=trim(out_file), access='stream', form='unformatted', status='replace')
  do fld = 0, 15
    lo = [0, 0, 0]; hi = [nx+1, ny+1, nz+1]
    if (fld == HC_U .or. fld == HC_UP .or. fld == HC_UM) lo(1) = -1
    if (fld == HC_V .or. fld == HC_VP .or. fld == HC_VM) lo(2) = -1
    if (fld == HC_W .or. fld == HC_WP .or. fld == HC_WM) lo(3) = -1
    allocate (buf(product(hi - lo + 1)))
    call dl(fld, buf, lo, hi)
    write (u_out) buf
    deallocate (buf)
  enddo
  write (u_out) allnoise
  close (u_out)
  call hc_check(hc_destroy(hc), 'destroy')
  write (6, *) 'hotcross_driver: ', nsteps, ' steps on ', nx, ny, nz

contains

  subroutine dl(fld, b, lo, hi)
    integer, intent(in) :: fld
    real(8), target :: b(:)
    integer(c_int32_t), intent(in) :: lo(3), hi(3)
    call hc_check(hc_download(hc, int(fld, c_int), c_loc(b), lo, hi), 'download')
  end subroutine

  ! t_bkg, s_bkg by linear interpolation (initialize_profiles80.f:46-69)
  subroutine profile_from_table(nz, zstart, dz, zd, td, sd, nd, tb, sb)
    integer, intent(in) :: nz, nd
    real(8), intent(in) :: zstart, dz, zd(nd), td(nd), sd(nd)
    real(8), intent(out) :: tb(0:nz+1), sb(0:nz+1)
    integer :: k, kup, ilo, ihi, imid
    real(8) :: zz
    do k = 0, nz
      zz = (zstart - dz / 2 + k * dz)
      ilo = 1; ihi = nd
      do while (ihi - ilo > 1)
        imid = (ihi + ilo) / 2
        if (zz >= zd(imid)) then
          ilo = imid
        else
          ihi = imid
        endif
      enddo
      kup = ilo
      if (zz == zd(1)) kup = 1
      if (zz == zd(nd)) kup = nd - 1
      tb(k) = td(kup) + (td(kup+1) - td(kup)) / (zd(kup+1) - zd(kup)) * (zz - zd(kup))
      sb(k) = sd(kup) + (sd(kup+1) - sd(kup)) / (zd(kup+1) - zd(kup)) * (zz - zd(kup))
    enddo
    tb(0) = tb(1); tb(nz+1) = tb(nz)
    sb(0) = sb(1); sb(nz+1) = sb(nz)
  end subroutine

end program hotcross_driver
