! hotcross_capi.f90 - ISO_C_BINDING interfaces of include/hotcross.h.
!
! The C-ABI exports one entry per reference timestep subroutine; this module
! lets Fortran host code (the reference's own driver, or the drop-in driver in
! hotcross_driver.f90) call them.  hc_params_t mirrors struct hc_params
! member for member.
module hotcross_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: HC_OK = 0, HC_EINVAL = -1, HC_EDEVICE = -2, HC_ENOMEM = -3, HC_ECFL = -4, &
                               HC_ENAN = -5
  integer(c_int), parameter :: HC_FP64 = 0, HC_FP32 = 1
  integer(c_int), parameter :: HC_U = 0, HC_V = 1, HC_W = 2, HC_UP = 3, HC_VP = 4, HC_WP = 5, HC_UM = 6, &
                               HC_VM = 7, HC_WM = 8, HC_P = 9, HC_T = 10, HC_S = 11, HC_C = 12, HC_SS = 13, &
                               HC_SSV = 14, HC_RHOP = 15, HC_FX = 16, HC_FY = 17, HC_FZ = 18
  ! adjoint-only fields: the work velocities' Courant numbers
  integer(c_int), parameter :: HC_ADJ_UC = 32, HC_ADJ_VC = 33, HC_ADJ_WC = 34
  ! running sums per cell of a statistics slab (hc_stats_get)
  integer(c_int), parameter :: HC_NSTAT = 10

  type, bind(C) :: hc_params_t
    integer(c_int32_t) :: nx, ny, nz, nt
    real(c_double) :: dx, dy, dz, dt, zstart, zend
    integer(c_int32_t) :: bit2, bit3, bit4, bit5, bit7, bit8, bit9, bit10, bit2k, bit5k
    integer(c_int32_t) :: ismol, ismag_h, ismag_v
    integer(c_int32_t) :: bcondx, bcondy, bottom_slip, side_slip, open_box
    integer(c_int32_t) :: bdbit(4)
    integer(c_int32_t) :: modulo_x, modulo_y
    integer(c_int32_t) :: toggle_t, toggle_s, toggle_c, pad0
    real(c_double) :: cs_h, cs_v, prandtl, axmin, aymin, azmin, azmax, mixing_height_scale
    real(c_double) :: sponge_factor, lat
    real(c_double) :: g, rho0, assel, eps, omega
    real(c_double) :: source_length, source_width
    real(c_double) :: dqtdt, dqsdt, dqcdt
    real(c_double) :: tsource, ssource, wsource
    real(c_double) :: source_ramp_hrs, noise_amp, wramp_hrs
    real(c_double) :: flow_u0, flow_u1, flow_v1, flow_omega, flow_ramp_hrs
    integer(c_int32_t) :: eos, pad1
    real(c_double) :: eos_rho_ref, eos_alpha, eos_beta, eos_t0, eos_s0
  end type

  interface
    integer(c_int) function hc_create(prm, precision, ctx) bind(C, name="hc_create")
      import :: c_int, c_ptr, hc_params_t
      type(hc_params_t), intent(in) :: prm
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function hc_destroy(ctx) bind(C, name="hc_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function hc_last_error(ctx) bind(C, name="hc_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_synchronize(ctx) bind(C, name="hc_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_upload(ctx, field, host, lo, hi) bind(C, name="hc_upload")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: field
      integer(c_int32_t), intent(in) :: lo(3), hi(3)
    end function
    integer(c_int) function hc_download(ctx, field, host, lo, hi) bind(C, name="hc_download")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: field
      integer(c_int32_t), intent(in) :: lo(3), hi(3)
    end function
    integer(c_int) function hc_download_f32(ctx, field, host, lo, hi, scale) bind(C, name="hc_download_f32")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: field
      integer(c_int32_t), intent(in) :: lo(3), hi(3)
      real(c_double), value :: scale
    end function
    integer(c_int) function hc_clock(ctx, time, mtime, next_step) bind(C, name="hc_clock")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: time
      integer(c_int32_t), intent(out) :: mtime, next_step
    end function
    integer(c_int) function hc_set_profiles(ctx, t_bkg, s_bkg, rho_bkg, azzmin) bind(C, name="hc_set_profiles")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: t_bkg(*), s_bkg(*), rho_bkg(*), azzmin(*)
    end function
    integer(c_int) function hc_set_sponge(ctx, alphau, alphav, alphae, alphaw) bind(C, name="hc_set_sponge")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: alphau(*), alphav(*), alphae(*), alphaw(*)
    end function
    integer(c_int) function hc_init_case(ctx, zd, td, sd, nd, t_noise, seed) bind(C, name="hc_init_case")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: zd(*), td(*), sd(*)
      integer(c_int32_t), value :: nd
      real(c_double), value :: t_noise
      integer(c_int64_t), value :: seed
    end function
    integer(c_int) function hc_begin_step(ctx, mtime, time) bind(C, name="hc_begin_step")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mtime
      real(c_double), value :: time
    end function
    integer(c_int) function hc_body_force(ctx) bind(C, name="hc_body_force")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_forces(ctx) bind(C, name="hc_forces")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_pressure(ctx) bind(C, name="hc_pressure")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_momentum_u(ctx) bind(C, name="hc_momentum_u")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_momentum_v(ctx) bind(C, name="hc_momentum_v")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_momentum_w(ctx) bind(C, name="hc_momentum_w")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_viscosity3d(ctx) bind(C, name="hc_viscosity3d")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_source_terms(ctx, noise) bind(C, name="hc_source_terms")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, noise
    end function
    integer(c_int) function hc_mixing(ctx) bind(C, name="hc_mixing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_work_velocity(ctx, tau3) bind(C, name="hc_work_velocity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tau3
    end function
    integer(c_int) function hc_tracer31(ctx, field, c_bkg, cmin, cmax, tau3, budgets) bind(C, name="hc_tracer31")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(in) :: c_bkg(*)
      real(c_double), value :: cmin, cmax, tau3
      real(c_double), intent(out) :: budgets(3)
    end function
    integer(c_int) function hc_vent_ghost(ctx, field, value) bind(C, name="hc_vent_ghost")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), value :: value
    end function
    integer(c_int) function hc_density(ctx) bind(C, name="hc_density")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_refresh(ctx) bind(C, name="hc_refresh")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_cfl_values(ctx, out) bind(C, name="hc_cfl_values")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(6)
    end function
    integer(c_int) function hc_step(ctx, nsteps) bind(C, name="hc_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    ! ---- plane statistics: time sums over planes k0..k1 ---------------------
    integer(c_int) function hc_stats_reset(ctx, k0, k1, every) bind(C, name="hc_stats_reset")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: k0, k1, every
    end function
    integer(c_int) function hc_stats_accumulate(ctx) bind(C, name="hc_stats_accumulate")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_stats_get(ctx, sums, count) bind(C, name="hc_stats_get")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, sums
      integer(c_int32_t), intent(out) :: count
    end function
    ! ---- controls and discrete adjoint (no reference counterpart) ----------
    integer(c_int) function hc_set_controls(ctx, qmap, flow_scale, cs) bind(C, name="hc_set_controls")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, qmap
      real(c_double), value :: flow_scale, cs
    end function
    integer(c_int) function hc_adj_reset(ctx) bind(C, name="hc_adj_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_upload(ctx, field, host, lo, hi) bind(C, name="hc_adj_upload")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: field
      integer(c_int32_t), intent(in) :: lo(3), hi(3)
    end function
    integer(c_int) function hc_adj_download(ctx, field, host, lo, hi) bind(C, name="hc_adj_download")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: field
      integer(c_int32_t), intent(in) :: lo(3), hi(3)
    end function
    integer(c_int) function hc_adj_forces(ctx) bind(C, name="hc_adj_forces")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_pressure(ctx) bind(C, name="hc_adj_pressure")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_momentum_u(ctx) bind(C, name="hc_adj_momentum_u")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_momentum_v(ctx) bind(C, name="hc_adj_momentum_v")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_momentum_w(ctx) bind(C, name="hc_adj_momentum_w")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_viscosity3d(ctx) bind(C, name="hc_adj_viscosity3d")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_work_velocity(ctx, tau3) bind(C, name="hc_adj_work_velocity")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tau3
    end function
    integer(c_int) function hc_adj_tracer31(ctx, field, c_bkg, cmin, cmax, tau3) bind(C, name="hc_adj_tracer31")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
      real(c_double), intent(in) :: c_bkg(*)
      real(c_double), value :: cmin, cmax, tau3
    end function
    integer(c_int) function hc_adj_vent_ghost(ctx, field) bind(C, name="hc_adj_vent_ghost")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field
    end function
    integer(c_int) function hc_adj_density(ctx) bind(C, name="hc_adj_density")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_refresh(ctx) bind(C, name="hc_adj_refresh")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hc_adj_gradients(ctx, gq, out) bind(C, name="hc_adj_gradients")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, gq
      real(c_double), intent(out) :: out(3)
    end function
    integer(c_int) function hc_adj_set_probes(ctx, nprobe, ijk, every, wgt_w, wgt_t) bind(C, name="hc_adj_set_probes")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nprobe, every
      integer(c_int32_t), intent(in) :: ijk(3, *)
      real(c_double), value :: wgt_w, wgt_t
    end function
    integer(c_int) function hc_adj_window_bytes(ctx, nsteps, seg, bytes) bind(C, name="hc_adj_window_bytes")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps, seg
      integer(c_int64_t), intent(out) :: bytes
    end function
    integer(c_int) function hc_adj_window(ctx, nsteps, seg, obs, samples, loss, gq, gs, end_state) &
        bind(C, name="hc_adj_window")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx, samples, gq
      integer(c_int32_t), value :: nsteps, seg, end_state
      real(c_double), intent(in) :: obs(*)
      real(c_double), intent(out) :: loss, gs(3)
    end function
    ! the same window under binomial checkpointing with nsnap snapshot slots
    integer(c_int) function hc_adj_window_binomial_bytes(ctx, nsteps, nsnap, bytes) &
        bind(C, name="hc_adj_window_binomial_bytes")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps, nsnap
      integer(c_int64_t), intent(out) :: bytes
    end function
    integer(c_int) function hc_adj_window_binomial(ctx, nsteps, nsnap, obs, samples, loss, gq, gs, end_state) &
        bind(C, name="hc_adj_window_binomial")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx, samples, gq
      integer(c_int32_t), value :: nsteps, nsnap, end_state
      real(c_double), intent(in) :: obs(*)
      real(c_double), intent(out) :: loss, gs(3)
    end function
    integer(c_int) function hc_binomial_schedule(nsteps, nsnap, acts, cap, nacts, nfwd) &
        bind(C, name="hc_binomial_schedule")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      integer(c_int32_t), value :: nsteps, nsnap
      type(c_ptr), value :: acts
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: nacts, nfwd
    end function
    ! output diagnostics and the case arrays the context holds
    integer(c_int) function hc_set_diagnostics(ctx, on) bind(C, name="hc_set_diagnostics")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function hc_get_diagnostic(ctx, which, host) bind(C, name="hc_get_diagnostic")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function hc_get_profiles(ctx, t_bkg, s_bkg, rho_bkg, azzmin) bind(C, name="hc_get_profiles")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, t_bkg, s_bkg, rho_bkg, azzmin
    end function
    integer(c_int) function hc_get_sponge(ctx, alphau, alphav, alphae, alphaw) bind(C, name="hc_get_sponge")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, alphau, alphav, alphae, alphaw
    end function
  end interface

contains

  ! message of the last error of a context, as a Fortran string
  function hc_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: q
    msg = ' '
    call c_f_pointer(hc_last_error(ctx), p, [256])
    do q = 1, 256
      if (p(q) == c_null_char) exit
      msg(q:q) = p(q)
    enddo
  end function

end module hotcross_capi
