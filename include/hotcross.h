/*
 * hotcross.h - C-ABI of the MI355X-native HotCross LES timestep.
 *
 * The reference (diioriod/PlumeModel-Hotcross) is a Fortran program whose
 * timestep is a fixed sequence of argument-less subroutines acting on module
 * state (hotcross80-revised2014.f:264-358).  This library keeps that state in
 * HBM and exports one entry point per reference subroutine, so that Fortran
 * host code (through ISO_C_BINDING, see plumemodel-hotcross_amd/fortran/) or
 * any other host language can call them in the reference's own order.
 *
 * Conventions
 *  - Every entry returns 0 on success and a negative HC_E* code on failure;
 *    hc_last_error() gives the message.  The reference's equivalent is a
 *    Fortran `stop 'message'` (hotcross80-revised2014.f:187-192,413,1314).
 *  - Host arrays are column-major (x fastest) with explicit Fortran bounds
 *    lo[3]..hi[3], exactly the module arrays of modules80-revised2014.f.
 *  - One context = one device stream.  Entries that return host-visible
 *    scalars (hc_cfl_values, hc_tracer31 budgets, hc_download) synchronise;
 *    the rest are asynchronous on the context stream.
 *  - Single caller thread per context (the reference calls these from the
 *    master thread between OpenMP regions).
 */
#ifndef HOTCROSS_H
#define HOTCROSS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HC_OK 0
#define HC_EINVAL (-1)   /* bad argument / unsupported configuration      */
#define HC_EDEVICE (-2)  /* HIP runtime error                             */
#define HC_ENOMEM (-3)   /* device allocation failed                      */
#define HC_ECFL (-4)     /* "stop on impending cfl violation"             */
#define HC_ENAN (-5)     /* non-finite velocity found by the CFL check    */

#define HC_FP64 0        /* plain form: every field and kernel in FP64    */
#define HC_FP32 1        /* chosen form: FP32 fields, FP64 islands        */

/* Case parameters: the variables the reference reads from its case file
 * (case.inc -> DanteExp1.case) and module constants
 * (modules80-revised2014.f:49-62,77-109,256-275). */
typedef struct hc_params {
  int32_t nx, ny, nz, nt;                 /* array_size.inc:6, grid_module nt */
  double dx, dy, dz, dt;                  /* grid_module                    */
  double zstart, zend;                    /* grid_module                    */
  /* signals_module switches */
  int32_t bit2, bit3, bit4, bit5, bit7, bit8, bit9, bit10, bit2k, bit5k;
  int32_t ismol, ismag_h, ismag_v;
  int32_t bcondx, bcondy, bottom_slip, side_slip, open_box;
  int32_t bdbit[4];
  int32_t modulo_x, modulo_y;
  int32_t toggle_t, toggle_s, toggle_c;   /* toggle(6), toggle(7), toggle(8) */
  int32_t pad0;
  /* viscosity_module */
  double cs_h, cs_v, prandtl, axmin, aymin, azmin, azmax, mixing_height_scale;
  double sponge_factor, lat;
  /* constants_module */
  double g, rho0, assel, eps, omega;
  /* vent source: source_module/tracer_module + source80.inc constants */
  double source_length, source_width;
  double dqtdt, dqsdt, dqcdt;
  double tsource, ssource, wsource;
  double source_ramp_hrs;   /* source80.inc:8  (reference: 2 h)            */
  double noise_amp;         /* source80.inc:32 wee (reference: 0.025)      */
  double wramp_hrs;         /* hotcross80-revised2014.f:835 (6 h)          */
  /* crossflow series, site_forcing.inc:51-72:
   * ut(l) = ramp(l) * (flow_u0 + flow_u1*sin(flow_omega*l*dt)),
   * vt(l) = ramp(l) * flow_v1*cos(flow_omega*l*dt),
   * ramp(l) = min(1, l*dt/(flow_ramp_hrs*3600)), or 1 if flow_ramp_hrs<=0 */
  double flow_u0, flow_u1, flow_v1, flow_omega, flow_ramp_hrs;
  /* equation of state: 0 = UNESCO one-atmosphere polynomial of
   * potential_density80.f:94-101; 1 = linear
   * rho = eos_rho_ref*(1 - eos_alpha*(T-eos_t0) + eos_beta*(S-eos_s0)) */
  int32_t eos, pad1;
  double eos_rho_ref, eos_alpha, eos_beta, eos_t0, eos_s0;
} hc_params;

/* Device-resident fields (module arrays of modules80-revised2014.f).
 * Fortran bounds: u-class (-1:nx+1,0:ny+1,0:nz+1), v-class
 * (0:nx+1,-1:ny+1,0:nz+1), w-class (0:nx+1,0:ny+1,-1:nz+1), scalar-class
 * (0:nx+1,0:ny+1,0:nz+1). */
enum hc_field {
  HC_U = 0, HC_V, HC_W, HC_UP, HC_VP, HC_WP, HC_UM, HC_VM, HC_WM,
  HC_P, HC_T, HC_S, HC_C, HC_SS, HC_SSV, HC_RHOP, HC_FX, HC_FY, HC_FZ,
  HC_NFIELDS
};

typedef struct hc_ctx hc_ctx;

/* Create a context and allocate the state in HBM (all zero).
 * precision: HC_FP64 or HC_FP32.  Replaces the module allocations of
 * hotcross80-revised2014.f:120-135 and the HS3CRI set-up call (:1309). */
int hc_create(const hc_params *prm, int precision, hc_ctx **out);
int hc_destroy(hc_ctx *ctx);
const char *hc_last_error(const hc_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that record events. */
void *hc_stream(hc_ctx *ctx);
int hc_synchronize(hc_ctx *ctx);

/* Host <-> device copies of one field over the Fortran box lo..hi, which must
 * lie inside the field's bounds; host is column-major over that box. */
int hc_upload(hc_ctx *ctx, int field, const double *host, const int32_t lo[3],
              const int32_t hi[3]);
int hc_download(hc_ctx *ctx, int field, double *host, const int32_t lo[3],
                const int32_t hi[3]);
/* FP32 copy of scale*field over lo..hi, converted on the device: what the
 * reference's netCDF writers store (real(field, kind=4),
 * write_snapshota.f:269-374; w is written as -w, scale = -1,
 * put_block1.f:265).  Half the PCIe bytes of hc_download. */
int hc_download_f32(hc_ctx *ctx, int field, float *host, const int32_t lo[3],
                    const int32_t hi[3], double scale);

/* Background profiles, each of length nz+2 (index 0..nz+1): t_bkg, s_bkg,
 * rho_bkg (tracer_module), azzmin (viscosity_module).  Set by the
 * reference's INITIALIZE_PROFILES (initialize_profiles80.f:46-159). */
int hc_set_profiles(hc_ctx *ctx, const double *t_bkg, const double *s_bkg,
                    const double *rho_bkg, const double *azzmin);
/* Sponge masks from SPONGE_MAKER (sponge80.f:3-142): alphau(-1:nx+1,0:ny+1),
 * alphav(0:nx+1,-1:ny+1), alphae(0:nx+1,0:ny+1), alphaw(-1:nz+1). */
int hc_set_sponge(hc_ctx *ctx, const double *alphau, const double *alphav,
                  const double *alphae, const double *alphaw);
/* The profiles and sponge masks the context holds (set by hc_set_profiles /
 * hc_set_sponge or hc_init_case), in the layouts of the setters; any pointer
 * may be null.  No reference counterpart (the reference reads its module
 * arrays directly); for output writers and drivers that need them. */
int hc_get_profiles(hc_ctx *ctx, double *t_bkg, double *s_bkg, double *rho_bkg, double *azzmin);
int hc_get_sponge(hc_ctx *ctx, double *alphau, double *alphav, double *alphae, double *alphaw);
/* Output diagnostics of VISCOSITY3D and MIXING: with on = 1, every later
 * hc_viscosity3d (also inside hc_step) also forms the Richardson number Ri
 * (smagorinsky_split_mixing-revised2014.f:101-102), epsh = S_H^2 (A_H +
 * axmin) (:54-59), epsv = S_V^2 (A_V + azmin) (:108-114), and chi_v, chi_h
 * (mixing80-revised2014.f:64-94, from T and ss, ss_v as they are then, which
 * is what MIXING sees in the reference's call order), over the interior; the
 * ghost cells are zero.  Off by default (one more pass over the state per
 * step).  hc_get_diagnostic copies diagnostic `which` (0 Ri, 1 epsh, 2 epsv,
 * 3 chi_v, 4 chi_h) of the last step as FP64 over (0:nx+1, 0:ny+1, 0:nz+1),
 * x fastest. */
int hc_set_diagnostics(hc_ctx *ctx, int32_t on);
int hc_get_diagnostic(hc_ctx *ctx, int32_t which, double *host);
/* Host-side case set-up in the library: profiles by LOCATOR interpolation of
 * a (z, T, S) table (initialize_profiles80.f:46-64,164-197), sponge masks
 * (sponge80.f), and the initial state of INITIALIZE
 * (hotcross80-revised2014.f:444-560): zero velocities with the -1e34 edge
 * fill, tracers at background, plus optional uniform T noise of amplitude
 * t_noise drawn from a counter-based generator seeded with `seed`. */
int hc_init_case(hc_ctx *ctx, const double *zdata, const double *tdata,
                 const double *sdata, int32_t ndata, double t_noise,
                 uint64_t seed);

/* ---- one entry per reference subroutine of the timestep ---------------- */
/* Time-loop control, hotcross80-revised2014.f:266-275: time, mtime, leap,
 * delta for step m. */
int hc_begin_step(hc_ctx *ctx, int32_t mtime, double time);
/* BODY_FORCE, body_force80.f:4-22 + site_forcing.inc:81-171. */
int hc_body_force(hc_ctx *ctx);
/* FORCES, hotcross80-revised2014.f:890-1139. */
int hc_forces(hc_ctx *ctx);
/* PRESSURE (+ HS3CRT solve), hotcross80-revised2014.f:1146-1405,
 * hs3crt_composite80.f:4-143. */
int hc_pressure(hc_ctx *ctx);
/* MOMENTUM_U/V/W, hotcross80-revised2014.f:671-887. */
int hc_momentum_u(hc_ctx *ctx);
int hc_momentum_v(hc_ctx *ctx);
int hc_momentum_w(hc_ctx *ctx);
/* VISCOSITY3D, smagorinsky_split_mixing-revised2014.f:5-150. */
int hc_viscosity3d(hc_ctx *ctx);
/* SOURCE_terms, source80.f:8-32 + source80.inc.  `noise` holds the three
 * uniform [0,1) draws per patch cell in the reference's order (j outer, i
 * inner; r1 r2 r3), or NULL to use the library's counter-based generator. */
int hc_source_terms(hc_ctx *ctx, const double *noise);
/* MIXING, mixing80-revised2014.f:1-99.  The diffusivities kxt/kyt/kzt are
 * formed inside the TRACER31 kernels from ss/ss_v with the same expressions,
 * so this entry only validates state (the chi diagnostics are output-only). */
int hc_mixing(hc_ctx *ctx);
/* Work velocities 0.5*(up+u) of hotcross80-revised2014.f:303-305, turned
 * into the Courant numbers of mpdata-hotcross80.f:94-104 with tau3. */
int hc_work_velocity(hc_ctx *ctx, double tau3);
/* TRACER31(c, q_c, c_bkg, cmin, cmax, c_total0, ...), mpdata-hotcross80.f:16
 * for field HC_T, HC_S or HC_C.  c_bkg has nz+2 entries.  If budgets is not
 * NULL and mod(mtime,100)==0 it receives c_total, sum_sponge, sum_source
 * (mpdata-hotcross80.f:384-423) and the call synchronises. */
int hc_tracer31(hc_ctx *ctx, int field, const double *c_bkg, double cmin,
                double cmax, double tau3, double *budgets);
/* Vent ghost cells after TRACER31, hotcross80-revised2014.f:311-315:
 * c(i,j,nz+1) = value on source cells. */
int hc_vent_ghost(hc_ctx *ctx, int field, double value);
/* DENSITY, potential_density80.f:3-71. */
int hc_density(hc_ctx *ctx);
/* REFRESH, hotcross80-revised2014.f:1411-1493. */
int hc_refresh(hc_ctx *ctx);
/* CFL_VALUES, hotcross80-revised2014.f:374-438: out = maxu, minu, maxv,
 * minv, maxw, minw (Courant numbers).  Returns HC_ECFL when the largest
 * magnitude exceeds 0.75, and HC_ENAN (out all NaN) when any velocity it
 * scans is NaN or infinite (the NaN bailout the reference keeps commented
 * out at :417-431; min/max alone pass over NaN).  The state is left as it
 * is, for the caller's emergency snapshot (the reference's :407-411). */
int hc_cfl_values(hc_ctx *ctx, double out[6]);

/* CFL_VALUES of hc_cfl_values over caller-owned device velocity fields,
 * allocating nothing.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc), which also cannot create a
 * context.  u, v, w (FP64 or FP32 by precision) are device buffers laid out
 * as the library's fields: one padded box per field over i=-1..nx+1,
 * j=-1..ny+1, k=-1..nz+1 with element (i,j,k) at
 *   (i+16) + px*((j+1) + (ny+3)*(k+1)),  px = 16*ceil((nx+18)/16),
 * whatever the field's staggering (only the ranges CFL_VALUES scans are
 * read).  The reduction's partial results live in one static device array
 * of 57,400 bytes per GPU, whatever the grid; calls that overlap on one GPU
 * (other host threads, other streams) share it and race.  Grids whose padded
 * box holds 2^31 elements or more are rejected (HC_EINVAL).  Runs on
 * `stream` (a hipStream_t, NULL for the null stream) and synchronises it;
 * out and the return code as hc_cfl_values (HC_ECFL, HC_ENAN), with no
 * message (there is no context). */
int hc_cfl_values_noalloc(int32_t nx, int32_t ny, int32_t nz, int precision, const void *u, const void *v,
                          const void *w, double dx, double dy, double dz, double dt, void *stream,
                          double out[6]);

/* The whole reference time loop body (hotcross80-revised2014.f:264-350) for
 * nsteps steps starting at the context's next step, with the source noise
 * from the counter-based generator.  CFL every 10 steps as in the
 * reference; returns HC_ECFL or HC_ENAN on the bailout conditions, after
 * the step whose check failed (the state stays on the device). */
int hc_step(hc_ctx *ctx, int32_t nsteps);

/* The three uniform draws per source-patch cell that hc_step's generator
 * uses at step mtime (patch order as hc_source_terms); returns the number of
 * patch cells.  For checkers that replay a native run. */
int hc_source_noise(hc_ctx *ctx, int32_t mtime, double *out);

/* Model time and step counters: time and mtime of the last step begun
 * (hotcross80-revised2014.f:266-268), and the step hc_step runs next.  Any
 * pointer may be NULL. */
int hc_clock(hc_ctx *ctx, double *time, int32_t *mtime, int32_t *next_step);

/* ---- plane statistics ---------------------------------------------------
 * Running FP64 time sums over the planes k0..k1 (1 <= k0 <= k1 <= nz), for
 * the time-mean w, T, TKE and heat flux of the plume runs (the reference
 * writes 1 Hz planes, put_block1.f:242-330, and averages them afterwards).
 * Per cell (i,j,k), x fastest then y then k, HC_NSTAT sums laid out
 * stat-major (sums[s*ncell + cell]):
 *   0 w at the w point (as stored: positive toward increasing k, i.e.
 *     downward; the writers output -w, put_block1.f:265), 1 T,
 *   2-4 the cell-centred u, v, w, 5-7 their squares, 8 T^2, 9 (cell w)*T.
 * hc_stats_reset zeroes the sums (k0 = k1 = 0 frees them); hc_step adds one
 * sample after every `every`-th step (m+1 divisible by every; 0 = only
 * hc_stats_accumulate adds).  hc_stats_get returns the sums and the sample
 * count (synchronises). */
#define HC_NSTAT 10
int hc_stats_reset(hc_ctx *ctx, int32_t k0, int32_t k1, int32_t every);
int hc_stats_accumulate(hc_ctx *ctx);
int hc_stats_get(hc_ctx *ctx, double *sums, int32_t *count);

/* ---- controls and discrete adjoint --------------------------------------
 * The reference has no adjoint; these entries differentiate the timestep
 * above.  Each hc_adj_<routine> is the transpose of the linearisation of the
 * forward entry of the same name at the context's current state (the state
 * the forward entry would read): it takes the adjoint of what the forward
 * entry writes from the adjoint fields, sets that adjoint to zero (the
 * forward overwrites it) and adds to the adjoints of what the forward reads.
 * Calling them in the reverse order of a forward sequence, each at the state
 * its forward call saw, gives the adjoint of the sequence.  Non-smooth
 * operations differentiate along the branch the forward took (|x| -> sign x,
 * the MPDATA limiter's zeroed fluxes held fixed, sqrt(max(0,x)) -> 0 below
 * zero). */

/* Controls the adjoint differentiates: qmap (npatch values, NULL = 1) scales
 * the vent heat flux of each source cell (q_t of source80.inc:46), flow_scale
 * multiplies the crossflow series ut, vt (site_forcing.inc:51-72), cs > 0
 * sets cs_h = cs_v (the Smagorinsky constants; cs <= 0 leaves them). */
int hc_set_controls(hc_ctx *ctx, const double *qmap, double flow_scale, double cs);

/* Adjoint fields use the forward field ids, plus the adjoints of the work
 * velocities' Courant numbers (mpdata-hotcross80.f:96-102), which have the
 * bounds of u, v and w. */
#define HC_ADJ_UC 32
#define HC_ADJ_VC 33
#define HC_ADJ_WC 34
/* Allocate (first call) and zero every adjoint field and gradient. */
int hc_adj_reset(hc_ctx *ctx);
int hc_adj_upload(hc_ctx *ctx, int field, const double *host, const int32_t lo[3], const int32_t hi[3]);
int hc_adj_download(hc_ctx *ctx, int field, double *host, const int32_t lo[3], const int32_t hi[3]);

int hc_adj_forces(hc_ctx *ctx);
int hc_adj_pressure(hc_ctx *ctx);
int hc_adj_momentum_u(hc_ctx *ctx);
int hc_adj_momentum_v(hc_ctx *ctx);
int hc_adj_momentum_w(hc_ctx *ctx);
int hc_adj_viscosity3d(hc_ctx *ctx);
int hc_adj_work_velocity(hc_ctx *ctx, double tau3);
/* TRACER31 of field (T, S or C) with the arguments of hc_tracer31; the
 * field's adjoint becomes the adjoint of its value before the call. */
int hc_adj_tracer31(hc_ctx *ctx, int field, const double *c_bkg, double cmin, double cmax, double tau3);
int hc_adj_vent_ghost(hc_ctx *ctx, int field);
int hc_adj_density(hc_ctx *ctx);
int hc_adj_refresh(hc_ctx *ctx);
/* Gradients accumulated since hc_adj_reset: gq (npatch values, may be NULL)
 * = dL/dqmap; out = dL/dflow_scale, dL/dcs_h, dL/dcs_v. */
int hc_adj_gradients(hc_ctx *ctx, double *gq, double out[3]);

/* Probe-misfit window.  Probes at interior cells (i,j,k) sample w and T
 * every `every` steps; the loss is
 *   L = sum over samples and probes of 0.5*(wgt_w (w - w_obs)^2 + wgt_t (T - T_obs)^2).
 * hc_adj_window runs nsteps steps of the native loop from the current state
 * (without the passive tracer c), keeping a checkpoint every `seg` steps in
 * HBM, then sweeps backwards, recomputing each segment into a tape; obs holds
 * (w_obs, T_obs) per probe per sample (nsteps/every samples), samples
 * (optional) receives the model's values in the same order.  On return the
 * context holds the window's end state (end_state = 1, to continue the run)
 * or its initial state and step counter (end_state = 0, to evaluate the
 * window again); gq and gs as hc_adj_gradients. */
int hc_adj_set_probes(hc_ctx *ctx, int32_t nprobe, const int32_t *ijk, int32_t every, double wgt_w,
                      double wgt_t);
/* HBM (bytes) the checkpoints and tapes of hc_adj_window(nsteps, seg) take
 * beyond the context's state and adjoint fields; for choosing seg under a
 * memory budget (the recompute is one forward step per step for any seg,
 * so the smallest footprint, near seg = sqrt(nsteps * 14 / 30), is the
 * usual choice). */
int hc_adj_window_bytes(hc_ctx *ctx, int32_t nsteps, int32_t seg, uint64_t *bytes);
int hc_adj_window(hc_ctx *ctx, int32_t nsteps, int32_t seg, const double *obs, double *samples, double *loss,
                  double *gq, double gs[3], int32_t end_state);

/* The same window under binomial checkpointing (Griewank & Walther's
 * revolve schedule): nsnap snapshot slots (slot 0 holds the window's start)
 * and the fewest forward steps for them; for budgets the fixed-interval
 * window does not fit (it needs about 2 sqrt(nsteps) states, this nsnap + 2
 * states and one step of tracer intermediates).  Every reversed step is run
 * once more just before its adjoint, with its tracers' intermediates taped.
 * Arguments and results as hc_adj_window.  No reference counterpart (the
 * reference has no adjoint); it serves the forward + adjoint configuration
 * at 256^3 over 1800 steps under a fixed HBM budget. */
int hc_adj_window_binomial_bytes(hc_ctx *ctx, int32_t nsteps, int32_t nsnap, uint64_t *bytes);
int hc_adj_window_binomial(hc_ctx *ctx, int32_t nsteps, int32_t nsnap, const double *obs, double *samples,
                           double *loss, double *gq, double gs[3], int32_t end_state);
/* The binomial schedule itself, computed on the host (no context, no device):
 * up to cap actions of four int32 (kind, step, slot, 0) into acts (may be
 * null; kind 0 advance the live state by step `step`, 1 snapshot of the live
 * state X_step into slot, 2 restore slot (X_step), 3 turn step `step` from
 * X_step held in slot), their number into nacts and the forward steps run
 * (advances and turns) into nfwd. */
int hc_binomial_schedule(int32_t nsteps, int32_t nsnap, int32_t *acts, int64_t cap, int64_t *nacts, int64_t *nfwd);

/* Per-phase device time accumulated since the last reset (ms), measured with
 * HIP events on the context stream when enabled.  names/ms have room for
 * `cap` entries; returns the number of phases. */
int hc_phase_timing(hc_ctx *ctx, int enable);
int hc_phase_times(hc_ctx *ctx, const char **names, double *ms, int32_t *counts,
                   int32_t cap);

#ifdef __cplusplus
}
#endif
#endif
