! ref_harness.f90 -- TEST INFRASTRUCTURE, not product code.
!
! A C-callable shim over the reference's own subroutines, compiled in place
! from /root/reference by oracle/build_ref.py into oracle/_ref/.  It sets the
! module state the reference's case file would set, calls the reference
! subroutines one at a time or as the driver's step sequence
! (hotcross80-revised2014.f:264-350), and copies module arrays in and out.
!
! Every reference subroutine runs as compiled from its own source, including
! VISCOSITY3D (smagorinsky_split_mixing-revised2014.f), which oracle/build_ref.py
! compiles with one preprocessor definition for the csqrt of a REAL at :112.
!
! The reference's parallel loops in VISCOSITY3D keep the strain components
! s11..s23 shared between threads (declared at
! smagorinsky_split_mixing-revised2014.f:16, loops at :26-65 and :72-122), so its results depend on thread timing whenever more than one thread
! runs; href_set_threads(1) makes the reference deterministic for checks.

module href
  use iso_c_binding
  implicit none
  private :: nx, ny, nz
  include 'array_size.inc'
  real(8), save :: worku(-1:nx+1,0:ny+1,0:nz+1), workv(0:nx+1,-1:ny+1,0:nz+1), &
                   workw(0:nx+1,0:ny+1,-1:nz+1)
  real(8), save :: h_lat = 0
  real(8), save :: h_noise(3*4096)
  integer, save :: h_nnoise = 0
contains


subroutine href_sizes(out) bind(C, name="href_sizes")
  use iso_c_binding
  implicit none
  integer(c_int), intent(out) :: out(3)
  out(1) = nx; out(2) = ny; out(3) = nz
end subroutine

! ip: bit2 bit3 bit4 bit5 bit7 bit8 bit9 bit10 bit2k bit5k ismol ismag_h ismag_v
!     bcondx bcondy bottom_slip side_slip open_box bdbit(1:4) modulo_x modulo_y
!     toggle6 toggle7 toggle8 nt                                    (28 values)
! rp: dx dy dz dt zstart zend cs_h cs_v prandtl axmin aymin azmin azmax
!     mixing_height_scale sponge_factor lat source_length source_width dqtdt
!     dqsdt dqcdt tsource ssource wsource                           (24 values)
subroutine href_setup(ip, rp) bind(C, name="href_setup")
  use iso_c_binding
  use velocity_module
  use tracer_module
  use viscosity_module
  use constants_module
  use signals_module
  use grid_module
  use force_module
  use source_module
  use geometry_module
  use tracer_smol_module
  implicit none
  integer(c_int), intent(in) :: ip(28)
  real(c_double), intent(in) :: rp(24)

  call initialize_constants
  bit_z_hp = 1; bit_z_hx = 1; bit_z_hy = 1; bit_z_hz = 1
  mp_ = 1.0; np_ = 1.0; mp_k = 1.0; np_k = 1.0; mpx_ = 1.0; npy_ = 1.0
  hzmn = 1.0; hzmnx = 1.0d0; hzmny = 1.0d0; hznx = 1.0d0; hzmy = 1.0d0
  hzp = 1.0d0; hzw = 1.0d0

  bit2 = ip(1); bit3 = ip(2); bit4 = ip(3); bit5 = ip(4); bit7 = ip(5); bit8 = ip(6)
  bit9 = ip(7); bit10 = ip(8); bit2k = ip(9); bit5k = ip(10); ismol = ip(11)
  ismag_h = ip(12); ismag_v = ip(13); bcondx = ip(14); bcondy = ip(15)
  bottom_slip = ip(16); side_slip = ip(17); open_box = ip(18)
  bdbit(1) = ip(19); bdbit(2) = ip(20); bdbit(3) = ip(21); bdbit(4) = ip(22)
  modulo_x = ip(23); modulo_y = ip(24)
  toggle = 0; toggle(6) = ip(25); toggle(7) = ip(26); toggle(8) = ip(27)
  nt = ip(28)

  dx = rp(1); dy = rp(2); dz = rp(3); dt = rp(4); zstart = rp(5); zend = rp(6)
  cs_h = rp(7); cs_v = rp(8); prandtl = rp(9); axmin = rp(10); aymin = rp(11)
  azmin = rp(12); azmax = rp(13); mixing_height_scale = rp(14)
  sponge_factor = rp(15); h_lat = rp(16); source_length = rp(17)
  source_width = rp(18); dQtdt = rp(19); dQSdt = rp(20); dQcdt = rp(21)
  tsource = rp(22); ssource = rp(23); wsource = rp(24)

  allocate (p(0:nx+1,0:ny+1,0:nz+1))
  allocate (u(-1:nx+1,0:ny+1,0:nz+1), up(-1:nx+1,0:ny+1,0:nz+1), um(-1:nx+1,0:ny+1,0:nz+1))
  allocate (v(0:nx+1,-1:ny+1,0:nz+1), vp(0:nx+1,-1:ny+1,0:nz+1), vm(0:nx+1,-1:ny+1,0:nz+1))
  allocate (w(0:nx+1,0:ny+1,-1:nz+1), wm(0:nx+1,0:ny+1,-1:nz+1), wp(0:nx+1,0:ny+1,-1:nz+1))
  p=0.;w=0.;up=0.;u=0.;um=0.;vp=0.;v=0.;vm=0.;wp=0.;w=0;wm=0.
  allocate (kxt(-1:nx+1,0:ny+1,0:nz+1), kyt(0:nx+1,-1:ny+1,0:nz+1), kzt(0:nx+1,0:ny+1,-1:nz+1))
  allocate (s(0:nx+1,0:ny+1,0:nz+1), Q_s(0:nx+1,0:ny+1,0:nz+1))
  allocate (t(0:nx+1,0:ny+1,0:nz+1), Q_t(0:nx+1,0:ny+1,0:nz+1))
  allocate (c(0:nx+1,0:ny+1,0:nz+1), Q_c(0:nx+1,0:ny+1,0:nz+1))
  kxt=0.;kyt=0.;kzt=0.;s=0.;q_s=0.;t=0.;q_t=0.;c=0.;Q_c=0.
  allocate ( uc(-1:nx+1,0:ny+1,0:nz+1), vc(0:nx+1,-1:ny+1,0:nz+1), wc(0:nx+1,0:ny+1,-1:nz+1))
  allocate ( cp(0:nx+1,0:ny+1,0:nz+1) , cc(0:nx+1,0:ny+1,0:nz+1))
  allocate ( ud(-1:nx+1,0:ny+1,0:nz+1), vd(0:nx+1,-1:ny+1,0:nz+1), wd(0:nx+1,0:ny+1,-1:nz+1))
  uc = 0.;vc=0.;wc=0.;cp=0.;cc=0.;ud=0.;vd=0.;wd=0.

  ds = dz
  fc = 2.0 * omega * sin (pi * h_lat / 180.0)
  fc_y = real(bit9) * 2.0 * omega * cos (pi * h_lat / 180.0)
  pgrady_con = 0.0
  pgradx_con = 0.0
  time = -dt
end subroutine

! INITIALIZE_PROFILES reads unit 1010 (initialize_profiles80.f:29-37)
subroutine href_init_profiles(path, n) bind(C, name="href_init_profiles")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  character(kind=c_char), intent(in) :: path(n)
  character(len=n) :: fname
  integer :: q
  do q = 1, n
    fname(q:q) = path(q)
  enddo
  open (unit=1010, file=fname, status='old')
  call initialize_profiles
  close (1010)
end subroutine

subroutine href_density() bind(C, name="href_density")
  call density
end subroutine

subroutine href_initialize() bind(C, name="href_initialize")
  call initialize
end subroutine

! SPONGE_MAKER + alphae_complement, hotcross80-revised2014.f:249-255
subroutine href_sponge_maker() bind(C, name="href_sponge_maker")
  use velocity_module
  call sponge_maker
  alphae_complement = 1.0d0
  where (alphae .gt. 0.0) alphae_complement = 0.0d0
end subroutine

! time-loop control, hotcross80-revised2014.f:264-275
subroutine href_begin_step(m) bind(C, name="href_begin_step")
  use iso_c_binding
  use grid_module
  use signals_module
  implicit none
  integer(c_int), value :: m
  time = time + dt
  mtime = m
  leap = 2
  delta = 1
  if ( m .eq. 0 ) then
    leap = 1
    delta = 0
  endif
end subroutine

subroutine href_body_force() bind(C, name="href_body_force")
  call body_force
end subroutine
subroutine href_forces() bind(C, name="href_forces")
  call forces
end subroutine
subroutine href_pressure() bind(C, name="href_pressure")
  call pressure
end subroutine
subroutine href_momentum_u() bind(C, name="href_momentum_u")
  call momentum_u
end subroutine
subroutine href_momentum_v() bind(C, name="href_momentum_v")
  call momentum_v
end subroutine
subroutine href_momentum_w() bind(C, name="href_momentum_w")
  call momentum_w
end subroutine
subroutine href_mixing() bind(C, name="href_mixing")
  call mixing
end subroutine
subroutine href_refresh() bind(C, name="href_refresh")
  call refresh
end subroutine

! SOURCE_terms, recording the random_number draws it makes so that another
! implementation can be fed the same noise (3 draws per patch cell)
subroutine href_source_terms() bind(C, name="href_source_terms")
  implicit none
  integer :: nseed, q
  integer, allocatable :: seed(:)
  real(8) :: r
  call random_seed(size=nseed)
  allocate (seed(nseed))
  call random_seed(get=seed)
  call source_terms
  call random_seed(put=seed)
  call href_patch_count(h_nnoise)
  do q = 1, 3 * h_nnoise
    call random_number(r)
    h_noise(q) = r
  enddo
  deallocate (seed)
end subroutine

subroutine href_patch_count(n)
  use tracer_module, only: isource_map
  implicit none
  integer, intent(out) :: n
  n = count(isource_map .eq. 1)
end subroutine

subroutine href_get_noise(buf, n) bind(C, name="href_get_noise")
  use iso_c_binding
  implicit none
  real(c_double), intent(out) :: buf(*)
  integer(c_int), intent(out) :: n
  n = h_nnoise
  buf(1:3*h_nnoise) = h_noise(1:3*h_nnoise)
end subroutine

! driver lines 303-305
subroutine href_work_velocity() bind(C, name="href_work_velocity")
  use velocity_module
  implicit none
  worku = 0.5 * ( up + u )
  workv = 0.5 * ( vp + v )
  workw = 0.5 * ( wp + w )
end subroutine

! TRACER31 as the driver calls it for T (which=1, with the vent ghost of
! :311-315), S (which=2, :321-332) and the passive tracer (which=3, :338-345)
subroutine href_tracer(which) bind(C, name="href_tracer")
  use iso_c_binding
  use tracer_module
  use grid_module
  implicit none
  integer(c_int), value :: which
  integer :: i, j
  real(8) :: tmin2, tmax2, smin2, smax2
  integer :: pend(0:nx+1,0:ny+1)
  pend = nz
  if (which .eq. 1) then
    tmin2 = minval(t_bkg); tmax2 = 1.0e+10
    call tracer31 (t, Q_t, t_bkg, tmin2, tmax2, t_total, pend, worku, workv, workw, dt, 401)
    do j = 1,ny; do i=1,nx
      t(i,j,nz+1) = (tsource + t_bkg(nz)) * isource_map(i,j) + real ( 1 - isource_map(i,j))  * t(i,j,nz+1)
    enddo; enddo
  else if (which .eq. 2) then
    smin2 = 0
    smax2 = maxval(s_bkg)
    call tracer31 (s, Q_s, s_bkg, smin2, smax2, s_total, pend, worku, workv, workw, dt, 402)
    do j = 1, ny; do i=1,nx
      s(i,j,nz+1) = (ssource + s_bkg(nz)) * isource_map(i,j) + real ( 1 - isource_map(i,j)) * s(i,j,nz+1)
    enddo; enddo
  else
    cmin = 0.0; cmax = 1.0e+10; c_bkg = 0.0
    call tracer31 (c, Q_c, c_bkg, cmin, cmax, c_total, pend, worku, workv, workw, dt, 403)
  endif
end subroutine

subroutine href_viscosity3d() bind(C, name="href_viscosity3d")
  call viscosity3d
end subroutine

! OpenMP thread count of the reference's parallel regions (the reference's
! main program fixes 8, hotcross80-revised2014.f:88-90, which never runs here)
subroutine href_set_threads(n) bind(C, name="href_set_threads")
  use iso_c_binding
  use omp_lib
  implicit none
  integer(c_int), value :: n
  call omp_set_num_threads(n)
end subroutine

! one pass of the time loop body, hotcross80-revised2014.f:264-350, without
! CFL_VALUES (its bailout path calls the netCDF writer) and output
subroutine href_step(m) bind(C, name="href_step")
  use iso_c_binding
  use signals_module
  implicit none
  integer(c_int), value :: m
  call href_begin_step(m)
  call body_force
  call forces
  call pressure
  call momentum_u
  call momentum_v
  call momentum_w
  call viscosity3d
  call href_source_terms
  call mixing
  if (toggle(6) .eq. 1) then
    call href_work_velocity
    call href_tracer(1)
  endif
  if (toggle(7) .eq. 1) call href_tracer(2)
  if (bit3 .eq. 1) call density
  if (toggle(8) .eq. 1) then
    call href_tracer(3)
  else
    call href_clear_cbkg
  endif
  call refresh
end subroutine

subroutine href_clear_cbkg()
  use tracer_module
  cmin = 0.0; cmax = 1.0e+10; c_bkg = 0.0
end subroutine

! the reference's CFL_VALUES (hotcross80-revised2014.f:374-438) on the current
! state: it writes maxu, minu, maxv, minv, maxw, minw and mtime to unit 81
! (format 6e15.7,i10); flushed here so the caller can read the line back.
! Only for states under the 0.75 bailout, whose path calls the netCDF writer.
subroutine href_cfl_values() bind(C, name="href_cfl_values")
  implicit none
  call cfl_values
  flush(81)
end subroutine

subroutine href_scalars(out) bind(C, name="href_scalars")
  use iso_c_binding
  use grid_module
  use velocity_module
  use force_module
  use constants_module
  implicit none
  real(c_double), intent(out) :: out(16)
  out(1) = time; out(2) = mtime; out(3) = leap; out(4) = 0
  out(5) = ubkg_m(0,1); out(6) = ubkg(0,1); out(7) = ubkg_p(0,1)
  out(8) = vbkg_m(1,0); out(9) = vbkg(1,0); out(10) = vbkg_p(1,0)
  out(11) = fbx(0,1,1); out(12) = fby(1,0,1); out(13) = fc; out(14) = fc_y
  out(15) = 0; out(16) = 0
end subroutine

! ---- array copies ---------------------------------------------------------
subroutine cp3(a, buf, dir)
  real(8) :: a(:,:,:)
  real(8) :: buf(*)
  integer :: dir, n
  n = size(a)
  if (dir .eq. 0) then
    buf(1:n) = reshape(a, [n])
  else
    a = reshape(buf(1:n), shape(a))
  endif
end subroutine

subroutine cp2(a, buf, dir)
  real(8) :: a(:,:)
  real(8) :: buf(*)
  integer :: dir, n
  n = size(a)
  if (dir .eq. 0) then
    buf(1:n) = reshape(a, [n])
  else
    a = reshape(buf(1:n), shape(a))
  endif
end subroutine

subroutine cp1(a, buf, dir)
  real(8) :: a(:)
  real(8) :: buf(*)
  integer :: dir, n
  n = size(a)
  if (dir .eq. 0) then
    buf(1:n) = a
  else
    a = buf(1:n)
  endif
end subroutine

! dir = 0: module array -> buf; dir = 1: buf -> module array.  Returns the
! number of elements, or -1 for an unknown name.
function href_copy(cname, n, buf, dir) bind(C, name="href_copy") result(cnt)
  use iso_c_binding
  use velocity_module
  use tracer_module
  use viscosity_module
  use force_module
  use tracer_smol_module
  implicit none
  integer(c_int), value :: n, dir
  character(kind=c_char), intent(in) :: cname(n)
  real(c_double) :: buf(*)
  integer(c_int) :: cnt
  character(len=32) :: nm
  real(8), allocatable :: tmp(:,:)
  integer :: q
  nm = ' '
  do q = 1, n
    nm(q:q) = cname(q)
  enddo
  cnt = 0
  select case (trim(nm))
  case ('u'); call cp3(u, buf, dir); cnt = size(u)
  case ('up'); call cp3(up, buf, dir); cnt = size(up)
  case ('um'); call cp3(um, buf, dir); cnt = size(um)
  case ('v'); call cp3(v, buf, dir); cnt = size(v)
  case ('vp'); call cp3(vp, buf, dir); cnt = size(vp)
  case ('vm'); call cp3(vm, buf, dir); cnt = size(vm)
  case ('w'); call cp3(w, buf, dir); cnt = size(w)
  case ('wp'); call cp3(wp, buf, dir); cnt = size(wp)
  case ('wm'); call cp3(wm, buf, dir); cnt = size(wm)
  case ('p'); call cp3(p, buf, dir); cnt = size(p)
  case ('fx'); call cp3(fx, buf, dir); cnt = size(fx)
  case ('fy'); call cp3(fy, buf, dir); cnt = size(fy)
  case ('fz'); call cp3(fz, buf, dir); cnt = size(fz)
  case ('t'); call cp3(t, buf, dir); cnt = size(t)
  case ('s'); call cp3(s, buf, dir); cnt = size(s)
  case ('c'); call cp3(c, buf, dir); cnt = size(c)
  case ('q_t'); call cp3(q_t, buf, dir); cnt = size(q_t)
  case ('q_s'); call cp3(q_s, buf, dir); cnt = size(q_s)
  case ('q_c'); call cp3(q_c, buf, dir); cnt = size(q_c)
  case ('kxt'); call cp3(kxt, buf, dir); cnt = size(kxt)
  case ('kyt'); call cp3(kyt, buf, dir); cnt = size(kyt)
  case ('kzt'); call cp3(kzt, buf, dir); cnt = size(kzt)
  case ('rhop'); call cp3(rhop, buf, dir); cnt = size(rhop)
  case ('rrho'); call cp3(rrho, buf, dir); cnt = size(rrho)
  case ('ss'); call cp3(ss, buf, dir); cnt = size(ss)
  case ('ss_v'); call cp3(ss_v, buf, dir); cnt = size(ss_v)
  case ('ri'); call cp3(ri, buf, dir); cnt = size(ri)
  case ('epsh'); call cp3(epsh, buf, dir); cnt = size(epsh)
  case ('epsv'); call cp3(epsv, buf, dir); cnt = size(epsv)
  case ('chi_v'); call cp3(chi_v, buf, dir); cnt = size(chi_v)
  case ('chi_h'); call cp3(chi_h, buf, dir); cnt = size(chi_h)
  case ('axu'); call cp3(axu, buf, dir); cnt = size(axu)
  case ('ayu'); call cp3(ayu, buf, dir); cnt = size(ayu)
  case ('axv'); call cp3(axv, buf, dir); cnt = size(axv)
  case ('ayv'); call cp3(ayv, buf, dir); cnt = size(ayv)
  case ('azu'); call cp3(azu, buf, dir); cnt = size(azu)
  case ('azv'); call cp3(azv, buf, dir); cnt = size(azv)
  case ('uc'); call cp3(uc, buf, dir); cnt = size(uc)
  case ('vc'); call cp3(vc, buf, dir); cnt = size(vc)
  case ('wc'); call cp3(wc, buf, dir); cnt = size(wc)
  case ('ud'); call cp3(ud, buf, dir); cnt = size(ud)
  case ('vd'); call cp3(vd, buf, dir); cnt = size(vd)
  case ('wd'); call cp3(wd, buf, dir); cnt = size(wd)
  case ('cp'); call cp3(cp, buf, dir); cnt = size(cp)
  case ('cc'); call cp3(cc, buf, dir); cnt = size(cc)
  case ('worku'); call cp3(worku, buf, dir); cnt = size(worku)
  case ('workv'); call cp3(workv, buf, dir); cnt = size(workv)
  case ('workw'); call cp3(workw, buf, dir); cnt = size(workw)
  case ('alphaee'); call cp3(alphaee, buf, dir); cnt = size(alphaee)
  case ('alphau'); call cp2(alphau, buf, dir); cnt = size(alphau)
  case ('alphav'); call cp2(alphav, buf, dir); cnt = size(alphav)
  case ('alphae'); call cp2(alphae, buf, dir); cnt = size(alphae)
  case ('isource_map')
    allocate (tmp(lbound(isource_map,1):ubound(isource_map,1), lbound(isource_map,2):ubound(isource_map,2)))
    if (dir .eq. 0) then
      tmp = isource_map
      call cp2(tmp, buf, 0)
    else
      call cp2(tmp, buf, 1)
      isource_map = nint(tmp)
    endif
    cnt = size(tmp)
    deallocate (tmp)
  case ('alphaw'); call cp1(alphaw, buf, dir); cnt = size(alphaw)
  case ('t_bkg'); call cp1(t_bkg, buf, dir); cnt = size(t_bkg)
  case ('s_bkg'); call cp1(s_bkg, buf, dir); cnt = size(s_bkg)
  case ('c_bkg'); call cp1(c_bkg, buf, dir); cnt = size(c_bkg)
  case ('rho_bkg'); call cp1(rho_bkg, buf, dir); cnt = size(rho_bkg)
  case ('azzmin'); call cp1(azzmin, buf, dir); cnt = size(azzmin)
  case ('pgradx'); call cp1(pgradx, buf, dir); cnt = size(pgradx)
  case ('pgrady'); call cp1(pgrady, buf, dir); cnt = size(pgrady)
  case default
    cnt = -1
  end select
end function

end module href
