// hc_ctx.h - the library context (HBM state, case arrays, step scalars) and
// the helpers shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/hotcross.h"
#include "hc_common.h"
#include "hc_kernels.h"

using namespace hc;

struct AdjCtx;  // adjoint state (hc_adjoint.hip)

struct Phase {
  const char *name;
  double ms;
  int count;
  hipEvent_t a, b;
  bool pending;
};

struct hc_ctx {
  hc_params P;
  int prec;
  size_t esz;
  Layout L;
  hipStream_t st = nullptr;
  std::string err;
  void *f[HC_NFIELDS] = {};
  // case arrays (device + host copies)
  Aux aux{};
  double *d_tbkg = nullptr, *d_sbkg = nullptr, *d_zero = nullptr, *d_cbkg = nullptr;
  std::vector<double> h_tbkg, h_sbkg, h_rho_bkg, h_azzmin, h_alphae, h_alphae0, h_alphau, h_alphav, h_alphaw;
  TracerWork tw{};
  PoissonWork pw{};
  double *d_tw = nullptr;
  double *d_q[3] = {};       // q_t, q_s, q_c on the patch
  double *d_noise = nullptr; // 3 draws per patch cell
  double *d_qbase = nullptr; // q_t of the patch before the heat-flux map factor
  double *d_stage = nullptr; // host<->device staging
  size_t stage_n = 0;
  double *d_cflpart = nullptr, *d_cflout = nullptr;
  double *h_pin = nullptr;   // pinned host scratch for small results
  // step state
  int mtime = -1, leap = 1, delta = 0;
  double time = 0, tau = 0;
  double ubkg_m = 0, ubkg = 0, ubkg_p = 0, vbkg_m = 0, vbkg = 0, vbkg_p = 0, fbx = 0, fby = 0;
  double fc = 0, fc_y = 0;
  double acc[8] = {};
  int nxs = 0, nxe = 0, nys = 0, nye = 0, npatch = 0;
  bool map_valid = false;
  // after REFRESH the reference's up/vp/wp equal u/v/w (u = up copies,
  // hotcross80-revised2014.f:1421-1488); here the buffers are swapped instead and
  // up_alias marks that the logical up is the u buffer until MOMENTUM writes it
  bool up_alias[3] = {false, false, false};
  double sumup = 0, sumdown = 0;
  uint64_t seed = 12345;
  int next_step = 0;
  double last_budget[3] = {};
  // timing
  bool timing = false;
  std::vector<Phase> phases;
  int mark_q = -1;
  // multiply by the exact reciprocal of power-of-two spacings (HC_NO_P2=1
  // at creation forces true divisions, for comparing the two forms)
  bool allow_p2 = true;
  // native loop: MOMENTUM also forms the Courant numbers and REFRESH's
  // filtered fields (into spare[c]); HC_NO_FUSE=1 at creation turns it off
  bool fuse = true;
  // output diagnostics (hc_set_diagnostics): Ri, epsh, epsv, chi_v, chi_h
  bool diag = false;
  void *d_diag[5] = {};
  void *spare[3] = {nullptr, nullptr, nullptr};
  void *c1x[2] = {nullptr, nullptr};  // pass-1 results of S and C in the native loop
  // the native loop's S and C tracers run on streams of their own beside T
  // (each with its own corrective-pass scratch), so that the blocks of one
  // tracer's passes fill the tail of the other's; HC_SERIAL_TRACERS=1 at
  // creation keeps them on the context stream.  Created on first use.
  bool par_tracers = true;
  hipStream_t tst[2] = {nullptr, nullptr};
  hipEvent_t tev[3] = {nullptr, nullptr, nullptr};
  TracerWork tw_par[2] = {};
  // controls that the adjoint differentiates: a per-cell multiplier of the
  // vent heat flux (null: 1 everywhere), the crossflow amplitude (a factor on
  // the series of site_forcing.inc) and the Smagorinsky constants in P
  double *d_qmap = nullptr;
  double flow_scale = 1.0;
  // d/d(flow_scale) of ubkg_m, ubkg, vbkg_m, vbkg, fbx, fby at this step
  double dflow[6] = {};
  AdjCtx *adj = nullptr;
  // running time sums over the planes k0..k1 (hc_stats.hip); the native loop
  // adds to them after every stats_every-th step (0: only on request)
  double *d_stats = nullptr;
  int stats_k0 = 0, stats_k1 = 0, stats_every = 0, stats_count = 0;
  // non-finite velocities found by the last CFL check
  int64_t nonfinite = 0;
};

void adj_free(hc_ctx *ctx);
void stats_free(hc_ctx *ctx);
int stats_accumulate(hc_ctx *ctx);
// shared by the C-ABI translation units (hc_api.hip)
int box_copy(hc_ctx *ctx, void *dev, int bounds_field, double *host, const int32_t lo[3], const int32_t hi[3],
             bool up);
void update_sums(hc_ctx *ctx);
void native_bounds(hc_ctx *ctx, double *tmin, double *smax);
int native_tracer(hc_ctx *ctx, int field, void *c1 = nullptr);
// the same on stream st with the corrective-pass scratch of tw (the
// context's own when null)
int native_tracer_on(hc_ctx *ctx, int field, void *c1, hipStream_t st, const TracerWork *tw);
// creates ctx->tst[0..1] and ctx->tev if missing (HC_OK or an error code)
int tracer_streams(hc_ctx *ctx);
int native_pass1(hc_ctx *ctx, int n, const int *fields, void *const *dst);
void *native_c1(hc_ctx *ctx, int q, int *rc);
// MOMENTUM with the work velocity and REFRESH's filter fused in, and the
// buffer rotation that completes REFRESH (the native loop's fused path)
int momentum_fused(hc_ctx *ctx, int comp, bool cour, double tau3);
void refresh_rotate(hc_ctx *ctx);

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);                     \
      return HC_EDEVICE;                                                             \
    }                                                                                \
  } while (0)

static inline int fail(hc_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

// ---------------------------------------------------------------------------
// phase timing with HIP events on the context stream
static inline int phase_begin(hc_ctx *ctx, const char *name) {
  if (!ctx->timing) return -1;
  for (size_t q = 0; q < ctx->phases.size(); ++q)
    if (!strcmp(ctx->phases[q].name, name)) {
      Phase &p = ctx->phases[q];
      if (p.pending) {
        float ms = 0;
        hipEventSynchronize(p.b);
        hipEventElapsedTime(&ms, p.a, p.b);
        p.ms += ms;
        p.count++;
        p.pending = false;
      }
      hipEventRecord(p.a, ctx->st);
      return (int)q;
    }
  Phase p{name, 0.0, 0, nullptr, nullptr, false};
  hipEventCreate(&p.a);
  hipEventCreate(&p.b);
  hipEventRecord(p.a, ctx->st);
  ctx->phases.push_back(p);
  return (int)ctx->phases.size() - 1;
}
static inline void phase_end(hc_ctx *ctx, int q) {
  if (q < 0) return;
  hipEventRecord(ctx->phases[q].b, ctx->st);
  ctx->phases[q].pending = true;
}
struct PhaseScope {
  hc_ctx *c;
  int q;
  PhaseScope(hc_ctx *c_, const char *n) : c(c_), q(phase_begin(c_, n)) {}
  ~PhaseScope() { phase_end(c, q); }
};

// ---------------------------------------------------------------------------
template <typename T>
inline FieldPtrs<T> ptrs(hc_ctx *ctx) {
  FieldPtrs<T> p;
  T **a = (T **)&p;
  for (int q = 0; q < HC_NFIELDS; ++q) a[q] = (T *)ctx->f[q];
  return p;
}

static inline StepArgs step_args(hc_ctx *ctx) {
  StepArgs s{};
  s.dx = ctx->P.dx; s.dy = ctx->P.dy; s.dz = ctx->P.dz; s.dt = ctx->P.dt; s.tau3 = ctx->P.dt;
  s.idx = 1.0 / s.dx; s.idy = 1.0 / s.dy; s.idz = 1.0 / s.dz;
  s.rpr = 1.0 / ctx->P.prandtl; s.rrho0 = 1.0 / ctx->P.rho0;
  s.rprf = 1.0f / (float)ctx->P.prandtl; s.rrho0f = 1.0f / (float)ctx->P.rho0;
  s.rdt = 1.0 / s.dt; s.rdtf = 1.0f / (float)s.dt;
  s.idxf = 1.0f / (float)s.dx; s.idyf = 1.0f / (float)s.dy; s.idzf = 1.0f / (float)s.dz;
  s.p2 = ctx->prec == HC_FP64
             ? (pow2_spacing<double>(s.dx) && pow2_spacing<double>(s.dy) && pow2_spacing<double>(s.dz))
             : (pow2_spacing<float>(s.dx) && pow2_spacing<float>(s.dy) && pow2_spacing<float>(s.dz));
  if (!ctx->allow_p2) s.p2 = 0;
  s.leap = ctx->leap; s.delta = ctx->delta; s.mtime = ctx->mtime;
  s.ubkg_m = ctx->ubkg_m; s.ubkg = ctx->ubkg; s.vbkg_m = ctx->vbkg_m; s.vbkg = ctx->vbkg;
  s.fbx = ctx->fbx; s.fby = ctx->fby; s.fc = ctx->fc; s.fc_y = ctx->fc_y;
  for (int q = 0; q < 8; ++q) s.acc[q] = ctx->acc[q];
  return s;
}

static inline VentArgs vent_args(hc_ctx *ctx) {
  VentArgs v{};
  v.map_valid = ctx->map_valid;
  v.nxs = ctx->nxs; v.nxe = ctx->nxe; v.nys = ctx->nys; v.nye = ctx->nye;
  double wup = ctx->P.wsource * fmin(1.0, (ctx->mtime * ctx->P.dt) / (ctx->P.wramp_hrs * 3600.0));
  v.wup = wup;
  v.sumup = ctx->sumup;
  v.sumdown = ctx->sumdown;
  return v;
}

// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_fill(size_t n, T *__restrict__ p, double v) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = (T)v;
}

static inline void field_bounds(int field, const hc_params &P, int lo[3], int hi[3]) {
  lo[0] = 0; lo[1] = 0; lo[2] = 0;
  hi[0] = P.nx + 1; hi[1] = P.ny + 1; hi[2] = P.nz + 1;
  switch (field) {
    case HC_U: case HC_UP: case HC_UM: case HC_FX: lo[0] = -1; break;
    case HC_V: case HC_VP: case HC_VM: case HC_FY: lo[1] = -1; break;
    case HC_W: case HC_WP: case HC_WM: case HC_FZ: lo[2] = -1; break;
    default: break;
  }
}

template <typename T>
inline void launch_fill(hc_ctx *ctx, void *p, double v) {
  size_t n = ctx->L.n;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->st, n, (T *)p, v);
}
static inline void fill(hc_ctx *ctx, void *p, double v) {
  if (ctx->prec == HC_FP64) launch_fill<double>(ctx, p, v);
  else launch_fill<float>(ctx, p, v);
}

// host -> device copy ordered on the context stream, complete on return.
// dalloc clears new buffers with an asynchronous memset on that stream, and
// the stream does not wait for the null stream (nor the null stream for
// it), so a synchronous hipMemcpy could land before a pending clear of the
// same buffer and be wiped.
static inline hipError_t h2d(hc_ctx *ctx, void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->st);
}

static inline void *dalloc(hc_ctx *ctx, size_t bytes, int *rc) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    *rc = HC_ENOMEM;
    ctx->err = "hipMalloc failed for " + std::to_string(bytes) + " bytes";
    return nullptr;
  }
  hipMemsetAsync(p, 0, bytes, ctx->st);
  return p;
}

