// hc_api.hip - the C-ABI of include/hotcross.h: context, HBM state, host-side
// case set-up and the per-subroutine entry points in the reference's order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/hotcross.h"
#include "hc_common.h"
#include "hc_ctx.h"
#include "hc_kernels.h"
#include "hc_rng.h"

using namespace hc;

// sumup/sumdown of MOMENTUM_W (hotcross80-revised2014.f:823-833)
void update_sums(hc_ctx *ctx) {
  int nx = ctx->P.nx, ny = ctx->P.ny;
  double su = 0, sd = 0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      if (ctx->h_alphae[i + (nx + 2) * j] == 0) {
        bool src = ctx->map_valid && i >= ctx->nxs && i <= ctx->nxe && j >= ctx->nys && j <= ctx->nye;
        if (src) su = su + 1.0; else sd = sd + 1.0;
      }
  ctx->sumup = su;
  ctx->sumdown = sd;
}

// ---------------------------------------------------------------------------
// pack/unpack between a Fortran box (column-major, FP64) and the padded field
template <typename T>
__global__ void k_scatter(Layout L, int l0, int l1, int l2, int n0, int n1, int n2, const double *__restrict__ src,
                          T *__restrict__ dst) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)n0 * n1 * n2;
  if (q >= n) return;
  int i = (int)(q % n0), j = (int)((q / n0) % n1), k = (int)(q / ((size_t)n0 * n1));
  dst[IX(L, l0 + i, l1 + j, l2 + k)] = (T)src[q];
}
template <typename T>
__global__ void k_gather(Layout L, int l0, int l1, int l2, int n0, int n1, int n2, const T *__restrict__ src,
                         double *__restrict__ dst) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)n0 * n1 * n2;
  if (q >= n) return;
  int i = (int)(q % n0), j = (int)((q / n0) % n1), k = (int)(q / ((size_t)n0 * n1));
  dst[q] = (double)src[IX(L, l0 + i, l1 + j, l2 + k)];
}

// FP32 copy of a Fortran box for the output writers (the reference writes
// real(..., kind=4) of each field, write_snapshota.f:269-374), times `scale`
// (-1 for w, which the writers output as -w, put_block1.f:265)
template <typename T>
__global__ void k_gather32(Layout L, int l0, int l1, int l2, int n0, int n1, int n2, const T *__restrict__ src,
                           double scale, float *__restrict__ dst) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)n0 * n1 * n2;
  if (q >= n) return;
  int i = (int)(q % n0), j = (int)((q / n0) % n1), k = (int)(q / ((size_t)n0 * n1));
  dst[q] = (float)(scale * (double)src[IX(L, l0 + i, l1 + j, l2 + k)]);
}

// q patch of SOURCE_terms (source80.inc:40-49) from the noise draws
__global__ void k_source(int npatch, double fq_t, double fq_s, double wee, const double *__restrict__ noise,
                         int use_hash, uint64_t seed, int mtime, const double *__restrict__ qmap,
                         double *__restrict__ qt, double *__restrict__ qs, double *__restrict__ qc,
                         double *__restrict__ qbase) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= npatch) return;
  double r1, r2;
  if (use_hash) {
    r1 = hash_uniform(seed, (uint64_t)mtime, (uint64_t)n, 0);
    r2 = hash_uniform(seed, (uint64_t)mtime, (uint64_t)n, 1);
  } else {
    r1 = noise[3 * n];
    r2 = noise[3 * n + 1];
  }
  qt[n] = fq_t * (1.0 + wee * (2.0 * (r1 - .5)));
  qbase[n] = qt[n];
  if (qmap) qt[n] = qt[n] * qmap[n];
  qs[n] = fq_s * (1.0 + wee * (2.0 * (r2 - .5)));
  qc[n] = 1.0;
}

// ---------------------------------------------------------------------------
extern "C" int hc_create(const hc_params *prm, int precision, hc_ctx **out) {
  if (!prm || !out) return HC_EINVAL;
  *out = nullptr;
  hc_ctx *ctx = new hc_ctx();
  ctx->P = *prm;
  const hc_params &P = ctx->P;
  ctx->prec = precision;
  if (precision != HC_FP64 && precision != HC_FP32) {
    int r = fail(ctx, HC_EINVAL, "precision must be HC_FP64 or HC_FP32");
    *out = ctx;
    return r;
  }
  auto fft_ok = [](int n) {
    if (n % 2) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
  };
  if (P.nz > 1024) {
    // the pressure solve keeps a z column in LDS
    int r = fail(ctx, HC_EINVAL, "grid must have nz <= 1024");
    *out = ctx;
    return r;
  }
  if (P.nx < 8 || P.ny < 4 || P.nz < 3 || !fft_ok(P.nx) || !fft_ok(P.ny)) {
    // nx >= 8: the entrance ramp of FORCES touches i = 0..7
    // (hotcross80-revised2014.f:953-956); nx, ny = 2^a 3^b, even, for the
    // LDS transforms
    int r = fail(ctx, HC_EINVAL, "grid must have nx>=8, ny>=4, nz>=3 and nx, ny even products of 2 and 3");
    *out = ctx;
    return r;
  }
  ctx->esz = precision == HC_FP64 ? 8 : 4;
  {
    const char *e = getenv("HC_XCD_REMAP");
    box_remap_flag() = (e && e[0] == '1') ? 1 : 0;
    const char *np2 = getenv("HC_NO_P2");
    ctx->allow_p2 = !(np2 && np2[0] == '1');
    const char *nf = getenv("HC_NO_FUSE");
    ctx->fuse = !(nf && nf[0] == '1');
  }
  ctx->L = make_layout(P.nx, P.ny, P.nz);
  if (ctx->L.n >= ((size_t)1 << 31)) {
    *out = ctx;
    return fail(ctx, HC_EINVAL, "grid too large: the padded box must hold fewer than 2^31 cells");
  }
  {
    const char *e = getenv("HC_SERIAL_TRACERS");
    ctx->par_tracers = !(e && e[0] == '1');
  }
  *out = ctx;
  HIPCHK(hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking));
  int rc = 0;
  size_t fb = ctx->L.n * ctx->esz;
  for (int q = 0; q < HC_NFIELDS; ++q)
    if (!(ctx->f[q] = dalloc(ctx, fb, &rc))) return rc;
  fill(ctx, ctx->f[HC_FX], MISSING);
  fill(ctx, ctx->f[HC_FY], MISSING);
  // tracer scratch
  void **tws[] = {&ctx->tw.uc, &ctx->tw.vc, &ctx->tw.wc, &ctx->tw.ud2, &ctx->tw.vd2, &ctx->tw.wd2,
                  &ctx->tw.c1, &ctx->tw.c2};
  for (auto p : tws)
    if (!(*p = dalloc(ctx, fb, &rc))) return rc;
  if (!(ctx->tw.budget_part = (double *)dalloc(ctx, (size_t)3 * P.ny * P.nz * 8, &rc))) return rc;
  if (!(ctx->tw.budget = (double *)dalloc(ctx, 64, &rc))) return rc;
  // case arrays
  int nz2 = P.nz + 2;
  double **a1[] = {&ctx->aux.azzmin, &ctx->aux.rho_bkg, &ctx->aux.pgradx, &ctx->aux.pgrady, &ctx->d_tbkg,
                   &ctx->d_sbkg, &ctx->d_zero, &ctx->d_cbkg};
  for (auto p : a1)
    if (!(*p = (double *)dalloc(ctx, (size_t)nz2 * 8, &rc))) return rc;
  if (!(ctx->aux.alphaw = (double *)dalloc(ctx, (size_t)(P.nz + 3) * 8, &rc))) return rc;
  size_t n2 = (size_t)(P.nx + 3) * (P.ny + 3) * 8;
  double **a2[] = {&ctx->aux.alphau, &ctx->aux.alphav, &ctx->aux.alphae, &ctx->aux.alphae0};
  for (auto p : a2)
    if (!(*p = (double *)dalloc(ctx, n2, &rc))) return rc;
  ctx->h_alphae.assign((size_t)(P.nx + 2) * (P.ny + 2), 0.0);
  // source patch
  {
    int nx = P.nx, ny = P.ny;
    int nnx = (int)lround(P.source_length / P.dx), nny = (int)lround(P.source_width / P.dy);
    if (nnx % 2 == 0) { ctx->nxs = nx / 3 - (nnx / 2 - 1); ctx->nxe = nx / 3 + (nnx / 2 - 1) + 1; }
    else { ctx->nxs = nx / 3 - (nnx / 2 - 1) - 1; ctx->nxe = nx / 3 + (nnx / 2 - 1) + 1; }
    if (nny % 2 == 0) { ctx->nys = ny / 2 - (nny / 2 - 1); ctx->nye = ny / 2 + (nny / 2 - 1) + 1; }
    else { ctx->nys = ny / 2 - (nny / 2 - 1) - 1; ctx->nye = ny / 2 + (nny / 2 - 1) + 1; }
    if (ctx->nxs < 1 || ctx->nxe > nx || ctx->nys < 1 || ctx->nye > ny)
      return fail(ctx, HC_EINVAL, "source patch does not fit the grid");
    ctx->npatch = (ctx->nxe - ctx->nxs + 1) * (ctx->nye - ctx->nys + 1);
  }
  for (int q = 0; q < 3; ++q)
    if (!(ctx->d_q[q] = (double *)dalloc(ctx, (size_t)ctx->npatch * 8, &rc))) return rc;
  if (!(ctx->d_noise = (double *)dalloc(ctx, (size_t)ctx->npatch * 3 * 8, &rc))) return rc;
  if (!(ctx->d_qbase = (double *)dalloc(ctx, (size_t)ctx->npatch * 8, &rc))) return rc;
  ctx->stage_n = (size_t)(P.nx + 3) * (P.ny + 3) * (P.nz + 3);
  if (!(ctx->d_stage = (double *)dalloc(ctx, ctx->stage_n * 8, &rc))) return rc;
  if (!(ctx->d_cflpart = (double *)dalloc(ctx, (size_t)CFL_NPART * CFL_BLOCKS * 8, &rc))) return rc;
  if (!(ctx->d_cflout = (double *)dalloc(ctx, 64, &rc))) return rc;
  HIPCHK(hipHostMalloc((void **)&ctx->h_pin, 4096, hipHostMallocDefault));
  // pressure solve
  size_t nc = (size_t)P.nx * P.ny * P.nz;
  if (!(ctx->pw.r = dalloc(ctx, nc * ctx->esz, &rc))) return rc;
  if (!(ctx->pw.piv = (double *)dalloc(ctx, nc * 8, &rc))) return rc;
  if (!(ctx->pw.part = (double *)dalloc(ctx, (size_t)P.ny * P.nz * 8, &rc))) return rc;
  if (!(ctx->pw.delta = (double *)dalloc(ctx, 64, &rc))) return rc;
  if (!(ctx->d_tw = (double *)dalloc(ctx, (size_t)4 * (P.nx + P.ny) * 8, &rc))) return rc;
  {
    // twiddles exp(-2 pi i m/N) and DCT rotations (cos, sin)(pi k/(2N))
    std::vector<double> tw((size_t)4 * (P.nx + P.ny));
    double pi = 4.0 * atan(1.0);
    size_t o = 0;
    for (int N : {P.nx, P.ny}) {
      for (int m = 0; m < N; ++m) { tw[o++] = cos(2 * pi * m / N); tw[o++] = -sin(2 * pi * m / N); }
      for (int k = 0; k < N; ++k) { tw[o++] = cos(pi * k / (2.0 * N)); tw[o++] = sin(pi * k / (2.0 * N)); }
    }
    HIPCHK(h2d(ctx, ctx->d_tw, tw.data(), tw.size() * 8));
    ctx->pw.tw_x = ctx->d_tw;
    ctx->pw.tw_y = ctx->d_tw + 4 * P.nx;
    // pivots of a*x(k-1) + (b_k + lambda)*x(k) + a*x(k+1) = g(k), b = -2a, b_1 = b_nz = -a,
    // lambda = -4/dx^2 sin^2(pi kx/(2nx)) - 4/dy^2 sin^2(pi ky/(2ny)), the DCT-II
    // eigenvalues of the x and y operators (hs3crt_composite80.f:200-225, 604-611)
    std::vector<double> piv(nc);
    double a = 1.0 / (P.dz * P.dz);
    double cx = 1.0 / (P.dx * P.dx), cy = 1.0 / (P.dy * P.dy);
    for (int ky = 0; ky < P.ny; ++ky)
      for (int kx = 0; kx < P.nx; ++kx) {
        double sx = sin(kx * pi / (2.0 * P.nx)), sy = sin(ky * pi / (2.0 * P.ny));
        double lam = -4.0 * cx * (sx * sx) + -4.0 * cy * (sy * sy);
        double d = 0;
        for (int k = 0; k < P.nz; ++k) {
          double b = (k == 0 || k == P.nz - 1) ? -a : -2.0 * a;
          double den = (k == 0) ? (b + lam) : (b + lam - a * a * d);
          if (kx == 0 && ky == 0 && k == P.nz - 1) den = 1.0;  // singular mode, hs3crt :627-629
          d = 1.0 / den;
          piv[((size_t)k * P.ny + ky) * P.nx + kx] = d;
        }
      }
    HIPCHK(h2d(ctx, ctx->pw.piv, piv.data(), nc * 8));
    if (ctx->prec == HC_FP64) {
      ctx->pw.pivt = ctx->pw.piv;
    } else {
      // the pivots the FP32 sweeps use, each rounded once (the value a
      // conversion at the point of use gives), half the bytes per sweep
      std::vector<float> pf(nc);
      for (size_t q = 0; q < nc; ++q) pf[q] = (float)piv[q];
      if (!(ctx->pw.pivt = dalloc(ctx, nc * 4, &rc))) return rc;
      HIPCHK(h2d(ctx, ctx->pw.pivt, pf.data(), nc * 4));
    }
  }
  double pi = 4.0 * atan2(1.0, 1.0);
  ctx->fc = 2.0 * P.omega * sin(pi * P.lat / 180.0);
  ctx->fc_y = (double)P.bit9 * 2.0 * P.omega * cos(pi * P.lat / 180.0);
  ctx->time = -P.dt;
  HIPCHK(hipStreamSynchronize(ctx->st));
  return HC_OK;
}

extern "C" int hc_destroy(hc_ctx *ctx) {
  if (!ctx) return HC_OK;
  if (ctx->st) hipStreamSynchronize(ctx->st);
  adj_free(ctx);  // before the side streams it uses are destroyed
  for (int q = 0; q < HC_NFIELDS; ++q) hipFree(ctx->f[q]);
  void *ps[] = {ctx->tw.uc, ctx->tw.vc, ctx->tw.wc, ctx->tw.ud2, ctx->tw.vd2, ctx->tw.wd2, ctx->tw.c1, ctx->tw.c2, ctx->tw.budget_part,
                ctx->tw.budget, ctx->aux.azzmin, ctx->aux.rho_bkg, ctx->aux.pgradx, ctx->aux.pgrady,
                ctx->aux.alphaw, ctx->aux.alphau, ctx->aux.alphav, ctx->aux.alphae, ctx->aux.alphae0,
                ctx->d_tbkg, ctx->d_sbkg, ctx->d_zero, ctx->d_cbkg, ctx->d_q[0], ctx->d_q[1], ctx->d_q[2],
                ctx->d_noise, ctx->d_qbase, ctx->d_stage, ctx->d_cflpart, ctx->d_cflout, ctx->pw.r, ctx->pw.piv, ctx->pw.part,
                ctx->pw.delta, ctx->d_tw};
  for (void *p : ps)
    if (p) hipFree(p);
  for (void *p : ctx->c1x)
    if (p) hipFree(p);
  for (int q = 0; q < 2; ++q) {
    if (ctx->tst[q]) {
      hipStreamSynchronize(ctx->tst[q]);
      hipStreamDestroy(ctx->tst[q]);
    }
    void *pq[] = {ctx->tw_par[q].ud2, ctx->tw_par[q].vd2, ctx->tw_par[q].wd2, ctx->tw_par[q].c2,
                  ctx->tw_par[q].budget_part, ctx->tw_par[q].budget};
    for (void *p : pq)
      if (p) hipFree(p);
  }
  for (hipEvent_t e : ctx->tev)
    if (e) hipEventDestroy(e);
  for (void *p : ctx->d_diag)
    if (p) hipFree(p);
  if (ctx->pw.pivt && ctx->pw.pivt != (void *)ctx->pw.piv) hipFree(ctx->pw.pivt);
  for (void *p : ctx->spare)
    if (p) hipFree(p);
  if (ctx->h_pin) hipHostFree(ctx->h_pin);
  if (ctx->d_qmap) hipFree(ctx->d_qmap);
  stats_free(ctx);
  for (auto &p : ctx->phases) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
  if (ctx->st) hipStreamDestroy(ctx->st);
  delete ctx;
  return HC_OK;
}

extern "C" const char *hc_last_error(const hc_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }
extern "C" void *hc_stream(hc_ctx *ctx) { return ctx ? (void *)ctx->st : nullptr; }
extern "C" int hc_synchronize(hc_ctx *ctx) {
  HIPCHK(hipStreamSynchronize(ctx->st));
  return HC_OK;
}

// copy the u buffer into the up buffer where the logical up still aliases u
static int materialise_alias(hc_ctx *ctx, int comp) {
  if (!ctx->up_alias[comp]) return HC_OK;
  int uf = HC_U + comp, upf = HC_UP + comp;
  HIPCHK(hipMemcpyAsync(ctx->f[upf], ctx->f[uf], ctx->L.n * ctx->esz, hipMemcpyDeviceToDevice, ctx->st));
  ctx->up_alias[comp] = false;
  return HC_OK;
}

static int check_box(hc_ctx *ctx, int field, const int32_t lo[3], const int32_t hi[3], int n[3]) {
  if (field < 0 || field >= HC_NFIELDS) return fail(ctx, HC_EINVAL, "bad field id");
  int fl[3], fh[3];
  field_bounds(field, ctx->P, fl, fh);
  for (int q = 0; q < 3; ++q) {
    if (lo[q] < fl[q] || hi[q] > fh[q] || hi[q] < lo[q]) return fail(ctx, HC_EINVAL, "box outside field bounds");
    n[q] = hi[q] - lo[q] + 1;
  }
  return HC_OK;
}

// copy between a host Fortran box and a padded device field (dev) with the
// bounds of field id bounds_field; up = host to device
int box_copy(hc_ctx *ctx, void *dev, int bounds_field, double *host, const int32_t lo[3], const int32_t hi[3],
             bool up) {
  int n[3];
  int r = check_box(ctx, bounds_field, lo, hi, n);
  if (r) return r;
  size_t cnt = (size_t)n[0] * n[1] * n[2];
  dim3 g((unsigned)((cnt + 255) / 256));
  if (up) {
    HIPCHK(hipMemcpyAsync(ctx->d_stage, host, cnt * 8, hipMemcpyHostToDevice, ctx->st));
    if (ctx->prec == HC_FP64)
      hipLaunchKernelGGL(k_scatter<double>, g, dim3(256), 0, ctx->st, ctx->L, lo[0], lo[1], lo[2], n[0], n[1], n[2],
                         ctx->d_stage, (double *)dev);
    else
      hipLaunchKernelGGL(k_scatter<float>, g, dim3(256), 0, ctx->st, ctx->L, lo[0], lo[1], lo[2], n[0], n[1], n[2],
                         ctx->d_stage, (float *)dev);
  } else {
    if (ctx->prec == HC_FP64)
      hipLaunchKernelGGL(k_gather<double>, g, dim3(256), 0, ctx->st, ctx->L, lo[0], lo[1], lo[2], n[0], n[1], n[2],
                         (const double *)dev, ctx->d_stage);
    else
      hipLaunchKernelGGL(k_gather<float>, g, dim3(256), 0, ctx->st, ctx->L, lo[0], lo[1], lo[2], n[0], n[1], n[2],
                         (const float *)dev, ctx->d_stage);
    HIPCHK(hipMemcpyAsync(host, ctx->d_stage, cnt * 8, hipMemcpyDeviceToHost, ctx->st));
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return HC_OK;
}

extern "C" int hc_upload(hc_ctx *ctx, int field, const double *host, const int32_t lo[3], const int32_t hi[3]) {
  if (field < 0 || field >= HC_NFIELDS) return fail(ctx, HC_EINVAL, "bad field id");
  if (field >= HC_U && field <= HC_WP) {
    int r = materialise_alias(ctx, field % 3);
    if (r) return r;
  }
  return box_copy(ctx, ctx->f[field], field, (double *)host, lo, hi, true);
}

extern "C" int hc_download_f32(hc_ctx *ctx, int field, float *host, const int32_t lo[3], const int32_t hi[3],
                               double scale) {
  if (field < 0 || field >= HC_NFIELDS) return fail(ctx, HC_EINVAL, "bad field id");
  int n[3];
  int r = check_box(ctx, field, lo, hi, n);
  if (r) return r;
  const int src = (field >= HC_UP && field <= HC_WP && ctx->up_alias[field - HC_UP]) ? field - 3 : field;
  const size_t cnt = (size_t)n[0] * n[1] * n[2];
  dim3 g((unsigned)((cnt + 255) / 256));
  float *stage = (float *)ctx->d_stage;
  if (ctx->prec == HC_FP64)
    hipLaunchKernelGGL(k_gather32<double>, g, dim3(256), 0, ctx->st, ctx->L, lo[0], lo[1], lo[2], n[0], n[1], n[2],
                       (const double *)ctx->f[src], scale, stage);
  else
    hipLaunchKernelGGL(k_gather32<float>, g, dim3(256), 0, ctx->st, ctx->L, lo[0], lo[1], lo[2], n[0], n[1], n[2],
                       (const float *)ctx->f[src], scale, stage);
  HIPCHK(hipMemcpyAsync(host, stage, cnt * 4, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return HC_OK;
}

extern "C" int hc_download(hc_ctx *ctx, int field, double *host, const int32_t lo[3], const int32_t hi[3]) {
  if (field < 0 || field >= HC_NFIELDS) return fail(ctx, HC_EINVAL, "bad field id");
  int src = (field >= HC_UP && field <= HC_WP && ctx->up_alias[field - HC_UP]) ? field - 3 : field;
  return box_copy(ctx, ctx->f[src], field, host, lo, hi, false);
}

extern "C" int hc_set_profiles(hc_ctx *ctx, const double *t_bkg, const double *s_bkg, const double *rho_bkg,
                               const double *azzmin) {
  int nz2 = ctx->P.nz + 2;
  ctx->h_tbkg.assign(t_bkg, t_bkg + nz2);
  ctx->h_sbkg.assign(s_bkg, s_bkg + nz2);
  ctx->h_rho_bkg.assign(rho_bkg, rho_bkg + nz2);
  ctx->h_azzmin.assign(azzmin, azzmin + nz2);
  HIPCHK(h2d(ctx, ctx->d_tbkg, t_bkg, nz2 * 8));
  HIPCHK(h2d(ctx, ctx->d_sbkg, s_bkg, nz2 * 8));
  HIPCHK(h2d(ctx, ctx->aux.rho_bkg, rho_bkg, nz2 * 8));
  HIPCHK(h2d(ctx, ctx->aux.azzmin, azzmin, nz2 * 8));
  return HC_OK;
}

extern "C" int hc_get_profiles(hc_ctx *ctx, double *t_bkg, double *s_bkg, double *rho_bkg, double *azzmin) {
  if (!ctx) return HC_EINVAL;
  if (ctx->h_tbkg.empty()) return fail(ctx, HC_EINVAL, "profiles not set (hc_init_case or hc_set_profiles)");
  const size_t n = ctx->h_tbkg.size() * 8;
  if (t_bkg) memcpy(t_bkg, ctx->h_tbkg.data(), n);
  if (s_bkg) memcpy(s_bkg, ctx->h_sbkg.data(), n);
  if (rho_bkg) memcpy(rho_bkg, ctx->h_rho_bkg.data(), n);
  if (azzmin) memcpy(azzmin, ctx->h_azzmin.data(), n);
  return HC_OK;
}

extern "C" int hc_get_sponge(hc_ctx *ctx, double *alphau, double *alphav, double *alphae, double *alphaw) {
  if (!ctx) return HC_EINVAL;
  if (ctx->h_alphau.empty()) return fail(ctx, HC_EINVAL, "sponge not set (hc_init_case or hc_set_sponge)");
  if (alphau) memcpy(alphau, ctx->h_alphau.data(), ctx->h_alphau.size() * 8);
  if (alphav) memcpy(alphav, ctx->h_alphav.data(), ctx->h_alphav.size() * 8);
  if (alphae) memcpy(alphae, ctx->h_alphae.data(), ctx->h_alphae.size() * 8);
  if (alphaw) memcpy(alphaw, ctx->h_alphaw.data(), ctx->h_alphaw.size() * 8);
  return HC_OK;
}

extern "C" int hc_set_sponge(hc_ctx *ctx, const double *alphau, const double *alphav, const double *alphae,
                             const double *alphaw) {
  const hc_params &P = ctx->P;
  int nx = P.nx, ny = P.ny, nz = P.nz;
  size_t nu = (size_t)(nx + 3) * (ny + 2), nv = (size_t)(nx + 2) * (ny + 3), ne = (size_t)(nx + 2) * (ny + 2);
  ctx->h_alphau.assign(alphau, alphau + nu);
  ctx->h_alphav.assign(alphav, alphav + nv);
  ctx->h_alphae.assign(alphae, alphae + ne);
  ctx->h_alphaw.assign(alphaw, alphaw + nz + 3);
  // alphaee (sponge80.f:115-119) is formed from alphae before its periodic
  // copies (:128-138); rebuild that mask
  ctx->h_alphae0.assign(ne, 0.0);
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 0; i <= nx + 1; ++i)
      ctx->h_alphae0[i + (nx + 2) * j] =
          fmax((alphau[(i + 1) + (nx + 3) * j] + alphau[i + (nx + 3) * j]),
               (alphav[i + (nx + 2) * (j + 1)] + alphav[i + (nx + 2) * j])) / 2.0;
  HIPCHK(h2d(ctx, ctx->aux.alphau, alphau, nu * 8));
  HIPCHK(h2d(ctx, ctx->aux.alphav, alphav, nv * 8));
  HIPCHK(h2d(ctx, ctx->aux.alphae, alphae, ne * 8));
  HIPCHK(h2d(ctx, ctx->aux.alphae0, ctx->h_alphae0.data(), ne * 8));
  HIPCHK(h2d(ctx, ctx->aux.alphaw, alphaw, (nz + 3) * 8));
  update_sums(ctx);
  return HC_OK;
}

// ---------------------------------------------------------------------------
// host-side set-up of a case (not on the timed path)
static double rho_rho_host(const hc_params &P, double s, double t) {
  if (P.eos == 1) return P.eos_rho_ref * (1.0 - P.eos_alpha * (t - P.eos_t0) + P.eos_beta * (s - P.eos_s0));
  double t2 = t * t, t3 = t2 * t, t4 = t2 * t2, t5 = t4 * t;
  double r0 = 999.842594 + 6.793952e-2 * t - 9.095290e-3 * t * t + 1.001685e-4 * t3 - 1.120083e-6 * t4 +
              6.536332e-9 * t5;
  return r0 + s * (0.824493 - 4.0899e-3 * t + 7.6438e-5 * t * t - 8.2467e-7 * t3 + 5.3875e-9 * t4) +
         pow(s, 1.5) * (-5.72466e-3 + 1.0227e-4 * t - 1.6546e-6 * t * t) + 4.8314e-4 * s * s;
}

extern "C" int hc_init_case(hc_ctx *ctx, const double *zd, const double *td, const double *sd, int32_t nd,
                            double t_noise, uint64_t seed) {
  const hc_params &P = ctx->P;
  int nx = P.nx, ny = P.ny, nz = P.nz;
  if (nd < 2) return fail(ctx, HC_EINVAL, "profile table needs two rows");
  ctx->seed = seed;
  // INITIALIZE_PROFILES, initialize_profiles80.f:46-64 with LOCATOR :164-197
  std::vector<double> tb(nz + 2), sb(nz + 2), az(nz + 2), rb(nz + 2);
  auto locate = [&](double x) {
    int ilo = 1, ihi = nd;
    while (ihi - ilo > 1) {
      int im = (ihi + ilo) / 2;
      if (x >= zd[im - 1]) ilo = im; else ihi = im;
    }
    if (x == zd[0]) return 1;
    if (x == zd[nd - 1]) return nd - 1;
    return ilo;
  };
  for (int k = 0; k <= nz; ++k) {
    double zz = (P.zstart - P.dz / 2 + k * P.dz);
    int kup = locate(zz);
    double z0 = zd[kup - 1], z1 = zd[kup];
    tb[k] = td[kup - 1] + (td[kup] - td[kup - 1]) / (z1 - z0) * (zz - z0);
    sb[k] = sd[kup - 1] + (sd[kup] - sd[kup - 1]) / (z1 - z0) * (zz - z0);
    az[k] = (P.azmin + (P.azmax - P.azmin) * exp(-((P.zend - P.zstart) - k * P.dz) / P.mixing_height_scale));
  }
  tb[0] = tb[1]; tb[nz + 1] = tb[nz];
  sb[0] = sb[1]; sb[nz + 1] = sb[nz];
  for (int k = 0; k <= nz + 1; ++k) rb[k] = rho_rho_host(P, sb[k], tb[k]);
  int r = hc_set_profiles(ctx, tb.data(), sb.data(), rb.data(), az.data());
  if (r) return r;
  // SPONGE_MAKER, sponge80.f:20-138
  std::vector<double> au((size_t)(nx + 3) * (ny + 2), 0.0), av((size_t)(nx + 2) * (ny + 3), 0.0),
      ae((size_t)(nx + 2) * (ny + 2), 0.0), aw(nz + 3, 0.0);
  double beta = 0.95;
#define AU(i, j) au[(size_t)((i) + 1) + (size_t)(nx + 3) * (j)]
#define AV(i, j) av[(size_t)(i) + (size_t)(nx + 2) * ((j) + 1)]
#define AE(i, j) ae[(size_t)(i) + (size_t)(nx + 2) * (j)]
  int ncut = (int)fmin(8.0, (double)lround(.10 * (double)nx));
  for (int j = 0; j <= ny + 1; ++j) {
    for (int i = -1; i <= ncut; ++i) AU(i, j) = beta * (double)(ncut - i) / (double)(ncut + 1);
    for (int i = nx + 1 - ncut - 1; i <= nx + 1; ++i) AU(i, j) = beta * (double)(i - (nx + 1 - ncut - 1)) / (double)(ncut + 1);
  }
  if (P.bcondx == 0)
    for (int j = 0; j <= ny + 1; ++j) { AU(-1, j) = AU(nx - 1, j); AU(nx + 1, j) = AU(1, j); }
  for (auto &x : au) x = x * x;
  ncut = (int)fmin(8.0, (double)lround(.10 * (double)ny));
  for (int i = 0; i <= nx + 1; ++i) {
    for (int j = -1; j <= ncut; ++j) AV(i, j) = beta * (double)(ncut - j) / (double)(ncut + 1);
    for (int j = ny + 1 - ncut - 1; j <= ny + 1; ++j) AV(i, j) = beta * (double)(j - (ny + 1 - ncut - 1)) / (double)(ncut + 1);
  }
  if (P.bcondy == 0)
    for (int i = 0; i <= nx + 1; ++i) { AV(i, -1) = AV(i, ny - 1); AV(i, ny + 1) = AV(i, 1); }
  for (auto &x : av) x = x * x;
  ncut = (int)fmin(10.0, (double)lround(.20 * (double)nz));
  for (int k = -1; k <= ncut; ++k) aw[k + 1] = beta * (double)(ncut - k) / (double)ncut;
  for (auto &x : aw) x = 1.0 * (x * x);
  double sf = P.sponge_factor / P.dt;
  for (auto &x : au) x = sf * x;
  for (auto &x : av) x = sf * x;
  for (auto &x : aw) x = sf * x;
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 0; i <= nx + 1; ++i) AE(i, j) = fmax((AU(i, j) + AU(i - 1, j)), (AV(i, j) + AV(i, j - 1))) / 2.0;
  if (P.bcondx == 0)
    for (int j = 0; j <= ny + 1; ++j) { AE(0, j) = AE(nx, j); AE(nx + 1, j) = AE(1, j); }
  if (P.bcondy == 0)
    for (int i = 0; i <= nx + 1; ++i) { AE(i, 0) = AE(i, ny); AE(i, ny + 1) = AE(i, 1); }
#undef AU
#undef AV
#undef AE
  r = hc_set_sponge(ctx, au.data(), av.data(), ae.data(), aw.data());
  if (r) return r;
  // INITIALIZE (hotcross80-revised2014.f:444-560): velocities and pressure
  // zero with the -1e34 fill on the 12 box edges of u, v, w, up, vp, wp, p;
  // tracers at background (initialize_profiles80.f:141-153), T noise
  for (int fld = 0; fld < HC_NFIELDS; ++fld) {
    if (fld == HC_FX || fld == HC_FY || fld == HC_FZ || fld == HC_SS || fld == HC_SSV || fld == HC_RHOP) continue;
    int lo[3], hi[3];
    field_bounds(fld, P, lo, hi);
    int n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1, n2 = hi[2] - lo[2] + 1;
    std::vector<double> a((size_t)n0 * n1 * n2, 0.0);
    for (int k = 0; k < n2; ++k)
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i) {
          size_t q = (size_t)i + (size_t)n0 * (j + (size_t)n1 * k);
          int kk = lo[2] + k;
          if (fld == HC_T) {
            a[q] = tb[kk];
            if (t_noise != 0.0)
              a[q] += t_noise * (2.0 * hash_uniform(seed ^ 0x7E57ull, (uint64_t)i, (uint64_t)j, (uint64_t)k) - 1.0);
          } else if (fld == HC_S) {
            a[q] = sb[kk];
          }
          bool edge = ((i == 0 || i == n0 - 1) + (j == 0 || j == n1 - 1) + (k == 0 || k == n2 - 1)) >= 2;
          bool filled = fld == HC_U || fld == HC_V || fld == HC_W || fld == HC_UP || fld == HC_VP ||
                        fld == HC_WP || fld == HC_P;
          if (edge && filled) a[q] = MISSING;
        }
    r = hc_upload(ctx, fld, a.data(), lo, hi);
    if (r) return r;
  }
  // DENSITY of the initial state (hotcross80-revised2014.f:213)
  return hc_density(ctx);
}

// ---------------------------------------------------------------------------
extern "C" int hc_begin_step(hc_ctx *ctx, int32_t mtime, double time) {
  ctx->time = time;
  ctx->mtime = mtime;
  ctx->leap = 2;
  ctx->delta = 1;
  if (mtime == 0) { ctx->leap = 1; ctx->delta = 0; }
  return HC_OK;
}

// crossflow series, site_forcing.inc:47-72 (ut(-1) = vt(-1) = 0)
static void flow_series(const hc_params &P, int l, double *ut, double *vt) {
  if (l < 0) { *ut = 0; *vt = 0; return; }
  double time_ = P.dt * l + 0.0;
  double u = P.flow_u0 + P.flow_u1 * sin(P.flow_omega * time_);
  double v = P.flow_v1 * cos(P.flow_omega * time_);
  double fac = 1.0;
  if (P.flow_ramp_hrs > 0) fac = fmin(1.0, (l * P.dt) / (P.flow_ramp_hrs * 3600.0));
  *ut = fac * u;
  *vt = fac * v;
}

extern "C" int hc_body_force(hc_ctx *ctx) {
  const hc_params &P = ctx->P;
  ctx->tau = ctx->leap * P.dt;
  int l = (int)lround(ctx->time / P.dt);
  // the series at unit amplitude, then scaled by the crossflow amplitude
  // (1 unless a caller sets it: x * 1.0 is exact)
  double d, um1, u0, up1, vm1, v0, vp1;
  flow_series(P, l - 1, &um1, &d); flow_series(P, l - 1, &d, &vm1);
  flow_series(P, l, &u0, &d); flow_series(P, l, &d, &v0);
  flow_series(P, l + 1, &up1, &d); flow_series(P, l + 1, &d, &vp1);
  const double A = ctx->flow_scale;
  ctx->ubkg_m = A * um1; ctx->ubkg = A * u0; ctx->ubkg_p = A * up1;
  ctx->vbkg_m = A * vm1; ctx->vbkg = A * v0; ctx->vbkg_p = A * vp1;
  auto fbx_of = [&](double upp, double uc, double umm, double vc) {
    double v4 = ((vc + vc) + vc) + vc;
    return +(upp - (1 - ctx->delta) * uc - ctx->delta * umm) / ctx->tau - P.bit7 * ctx->fc * v4 / 4.0;
  };
  auto fby_of = [&](double vpp, double vc, double vmm, double uc) {
    double u4 = ((uc + uc) + uc) + uc;
    return +(vpp - (1 - ctx->delta) * vc - ctx->delta * vmm) / ctx->tau + P.bit7 * ctx->fc * u4 / 4.0;
  };
  ctx->fbx = fbx_of(ctx->ubkg_p, ctx->ubkg, ctx->ubkg_m, ctx->vbkg);
  ctx->fby = fby_of(ctx->vbkg_p, ctx->vbkg, ctx->vbkg_m, ctx->ubkg);
  // every one of these scalars is linear in the amplitude
  ctx->dflow[0] = um1; ctx->dflow[1] = u0; ctx->dflow[2] = vm1; ctx->dflow[3] = v0;
  ctx->dflow[4] = fbx_of(up1, u0, um1, v0);
  ctx->dflow[5] = fby_of(vp1, v0, vm1, u0);
  return HC_OK;
}

extern "C" int hc_forces(hc_ctx *ctx) {
  PhaseScope ps(ctx, "forces");
  // the entrance ramp (hotcross80-revised2014.f:953-956) accumulates on the
  // boundary entries that :937-944 never recompute
  for (int i = 0; i <= 7; ++i) ctx->acc[i] = ctx->acc[i] + (double)(7 - i) * .01 / 7.0;
  StepArgs s = step_args(ctx);
  if (ctx->prec == HC_FP64) launch_forces<double>(ctx->st, ctx->L, s, ctx->P, ptrs<double>(ctx), ctx->aux);
  else launch_forces<float>(ctx->st, ctx->L, s, ctx->P, ptrs<float>(ctx), ctx->aux);
  return HC_OK;
}

extern "C" int hc_pressure(hc_ctx *ctx) {
  PhaseScope ps(ctx, "pressure");
  StepArgs s = step_args(ctx);
  if (ctx->prec == HC_FP64) launch_pressure<double>(ctx->st, ctx->L, s, ctx->P, ptrs<double>(ctx), ctx->aux, ctx->pw);
  else launch_pressure<float>(ctx->st, ctx->L, s, ctx->P, ptrs<float>(ctx), ctx->aux, ctx->pw);
  return HC_OK;
}

static int momentum(hc_ctx *ctx, int comp) {
  PhaseScope ps(ctx, "momentum");
  ctx->up_alias[comp] = false;  // MOMENTUM writes the up buffer
  StepArgs s = step_args(ctx);
  VentArgs va = vent_args(ctx);
  if (ctx->prec == HC_FP64) launch_momentum<double>(ctx->st, comp, ctx->L, s, ctx->P, va, ptrs<double>(ctx), ctx->aux);
  else launch_momentum<float>(ctx->st, comp, ctx->L, s, ctx->P, va, ptrs<float>(ctx), ctx->aux);
  return HC_OK;
}
extern "C" int hc_momentum_u(hc_ctx *ctx) { return momentum(ctx, 0); }
extern "C" int hc_momentum_v(hc_ctx *ctx) { return momentum(ctx, 1); }
extern "C" int hc_momentum_w(hc_ctx *ctx) { return momentum(ctx, 2); }

extern "C" int hc_viscosity3d(hc_ctx *ctx) {
  PhaseScope ps(ctx, "viscosity3d");
  StepArgs s = step_args(ctx);
  if (ctx->prec == HC_FP64) launch_viscosity<double>(ctx->st, ctx->L, s, ctx->P, ptrs<double>(ctx), ctx->aux);
  else launch_viscosity<float>(ctx->st, ctx->L, s, ctx->P, ptrs<float>(ctx), ctx->aux);
  if (ctx->diag) {
    if (ctx->prec == HC_FP64)
      launch_diag<double>(ctx->st, ctx->L, s, ctx->P, ptrs<double>(ctx), ctx->aux, ctx->d_tbkg, ctx->d_diag);
    else
      launch_diag<float>(ctx->st, ctx->L, s, ctx->P, ptrs<float>(ctx), ctx->aux, ctx->d_tbkg, ctx->d_diag);
  }
  return HC_OK;
}

extern "C" int hc_set_diagnostics(hc_ctx *ctx, int32_t on) {
  if (!ctx) return HC_EINVAL;
  if (on) {
    if (ctx->h_tbkg.empty()) return fail(ctx, HC_EINVAL, "profiles not set (hc_init_case or hc_set_profiles)");
    for (void *&p : ctx->d_diag) {
      if (!p) {
        int rc = 0;
        if (!(p = dalloc(ctx, ctx->L.n * ctx->esz, &rc))) return rc;
      }
      HIPCHK(hipMemsetAsync(p, 0, ctx->L.n * ctx->esz, ctx->st));
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
  }
  ctx->diag = on != 0;
  return HC_OK;
}

extern "C" int hc_get_diagnostic(hc_ctx *ctx, int32_t which, double *host) {
  if (!ctx || !host) return HC_EINVAL;
  if (which < 0 || which > 4) return fail(ctx, HC_EINVAL, "diagnostic 0..4 (ri, epsh, epsv, chi_v, chi_h)");
  if (!ctx->d_diag[which]) return fail(ctx, HC_EINVAL, "diagnostics are off (hc_set_diagnostics)");
  const int32_t lo[3] = {0, 0, 0}, hi[3] = {ctx->P.nx + 1, ctx->P.ny + 1, ctx->P.nz + 1};
  return box_copy(ctx, ctx->d_diag[which], HC_T, host, lo, hi, false);
}

extern "C" int hc_source_terms(hc_ctx *ctx, const double *noise) {
  const hc_params &P = ctx->P;
  double f = fmin(1.0, ctx->mtime * P.dt / (P.source_ramp_hrs * 3600.0));
  if (noise) HIPCHK(hipMemcpyAsync(ctx->d_noise, noise, (size_t)ctx->npatch * 3 * 8, hipMemcpyHostToDevice, ctx->st));
  hipLaunchKernelGGL(k_source, dim3((ctx->npatch + 255) / 256), dim3(256), 0, ctx->st, ctx->npatch, P.dqtdt * f,
                     P.dqsdt * f, P.noise_amp, ctx->d_noise, noise ? 0 : 1, ctx->seed, ctx->mtime, ctx->d_qmap, ctx->d_q[0],
                     ctx->d_q[1], ctx->d_q[2], ctx->d_qbase);
  if (noise) HIPCHK(hipStreamSynchronize(ctx->st));
  if (!ctx->map_valid) {
    ctx->map_valid = true;
    update_sums(ctx);
  }
  return HC_OK;
}

extern "C" int hc_mixing(hc_ctx *ctx) {
  (void)ctx;
  return HC_OK;
}

extern "C" int hc_work_velocity(hc_ctx *ctx, double tau3) {
  PhaseScope ps(ctx, "work_velocity");
  StepArgs s = step_args(ctx);
  s.tau3 = tau3;
  if (ctx->prec == HC_FP64) launch_courant<double>(ctx->st, ctx->L, s, ptrs<double>(ctx), ctx->tw);
  else launch_courant<float>(ctx->st, ctx->L, s, ptrs<float>(ctx), ctx->tw);
  return HC_OK;
}

// kernel timing of the corrective passes, as the phase "mpdata_pass"
static void mark_pass(void *arg, int end) {
  hc_ctx *ctx = (hc_ctx *)arg;
  if (!end) ctx->mark_q = phase_begin(ctx, "mpdata_pass");
  else phase_end(ctx, ctx->mark_q);
}

// c1: if set, holds the tracer's pass-1 result already (launch_pass1_multi)
// TRACER31 of `field` on stream st with the corrective-pass scratch of
// `work` (the context's own when null)
static int tracer_on(hc_ctx *ctx, hipStream_t st, const TracerWork *work, int field, const double *d_cbkg,
                     double cmin, double cmax, double tau3, int vent, double vent_value, int want_budget,
                     void *c1) {
  PhaseScope ps(ctx, "tracer31");
  int which = field == HC_T ? 0 : field == HC_S ? 1 : 2;
  StepArgs s = step_args(ctx);
  VentArgs va = vent_args(ctx);
  TracerArgs ta{cmin, cmax, tau3, vent, vent_value, want_budget, c1 != nullptr};
  TracerWork tw = work ? *work : ctx->tw;
  if (c1) tw.c1 = c1;
  MarkFn mk = ctx->timing ? mark_pass : nullptr;
  if (ctx->prec == HC_FP64)
    launch_tracer<double>(st, ctx->L, s, ctx->P, va, ptrs<double>(ctx), ctx->aux, tw,
                          (double *)ctx->f[field], d_cbkg, ctx->d_q[which], ta, mk, ctx);
  else
    launch_tracer<float>(st, ctx->L, s, ctx->P, va, ptrs<float>(ctx), ctx->aux, tw,
                         (float *)ctx->f[field], d_cbkg, ctx->d_q[which], ta, mk, ctx);
  return HC_OK;
}

static int tracer(hc_ctx *ctx, int field, const double *d_cbkg, double cmin, double cmax, double tau3, int vent,
                  double vent_value, int want_budget, void *c1 = nullptr) {
  return tracer_on(ctx, ctx->st, nullptr, field, d_cbkg, cmin, cmax, tau3, vent, vent_value, want_budget, c1);
}

// the two streams beside the context stream and their events: the S and C
// tracers of the native loop run on them, an adjoint window's S tracer and
// tape copies too (one set per context: with few hardware queues per
// process, every further stream would share a queue with one of these)
int tracer_streams(hc_ctx *ctx) {
  if (ctx->tst[0] && ctx->tst[1]) return HC_OK;
  for (int q = 0; q < 2; ++q)
    if (!ctx->tst[q]) HIPCHK(hipStreamCreateWithFlags(&ctx->tst[q], hipStreamNonBlocking));
  for (hipEvent_t &e : ctx->tev)
    if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return HC_OK;
}

// the streams, events and scratch of the concurrent S and C tracers
static int par_tracers_init(hc_ctx *ctx) {
  if (ctx->tw_par[1].budget) return tracer_streams(ctx);
  int rc = 0;
  const size_t fb = ctx->L.n * ctx->esz;
  for (int q = 0; q < 2; ++q) {
    TracerWork &w = ctx->tw_par[q];
    w.uc = ctx->tw.uc; w.vc = ctx->tw.vc; w.wc = ctx->tw.wc;  // read only
    void **ps[] = {&w.ud2, &w.vd2, &w.wd2, &w.c2};
    for (auto p : ps)
      if (!(*p = dalloc(ctx, fb, &rc))) return rc;
    if (!(w.budget_part = (double *)dalloc(ctx, (size_t)3 * ctx->P.ny * ctx->P.nz * 8, &rc))) return rc;
    if (!(w.budget = (double *)dalloc(ctx, 64, &rc))) return rc;
  }
  return tracer_streams(ctx);
}

// pass 1 of the native loop's tracers fields[0..n) (T, S or C) in one
// launch, results (boundary images included) into dst[q]
int native_pass1(hc_ctx *ctx, int n, const int *fields, void *const *dst) {
  if (n < 1) return HC_OK;
  if (n > 3) return fail(ctx, HC_EINVAL, "at most three tracers per pass-1 launch");
  void *c[3];
  const double *b[3];
  for (int q = 0; q < n; ++q) {
    c[q] = ctx->f[fields[q]];
    b[q] = fields[q] == HC_T ? ctx->d_tbkg : fields[q] == HC_S ? ctx->d_sbkg : ctx->d_zero;
  }
  PhaseScope ps(ctx, "tracer31");
  StepArgs s = step_args(ctx);
  VentArgs va = vent_args(ctx);
  if (ctx->prec == HC_FP64)
    launch_pass1_multi<double>(ctx->st, ctx->L, s, ctx->P, va, ptrs<double>(ctx), ctx->aux, ctx->tw, n,
                               (double *const *)c, b, (double *const *)dst, ctx->P.dt);
  else
    launch_pass1_multi<float>(ctx->st, ctx->L, s, ctx->P, va, ptrs<float>(ctx), ctx->aux, ctx->tw, n,
                              (float *const *)c, b, (float *const *)dst, ctx->P.dt);
  return HC_OK;
}

// the pass-1 buffer of S (q = 1) or C (q = 2) in the native loop; T uses
// the tracer scratch's own
void *native_c1(hc_ctx *ctx, int q, int *rc) {
  if (q == 0) return ctx->tw.c1;
  if (!ctx->c1x[q - 1]) ctx->c1x[q - 1] = dalloc(ctx, ctx->L.n * ctx->esz, rc);
  return ctx->c1x[q - 1];
}

// pass 1 of the native loop's tracers (T, S, C as switched on) in one launch;
// c1[q] (q = 0 T, 1 S, 2 C) receives where each result went, nullptr for the
// tracers that are off
static int tracers_pass1(hc_ctx *ctx, void *c1[3]) {
  const hc_params &P = ctx->P;
  const int on[3] = {P.toggle_t == 1, P.toggle_s == 1, P.toggle_c == 1};
  const int fid[3] = {HC_T, HC_S, HC_C};
  int fields[3], n = 0;
  void *dst[3];
  for (int q = 0; q < 3; ++q) {
    c1[q] = nullptr;
    if (!on[q]) continue;
    int rc = 0;
    if (!(c1[q] = native_c1(ctx, q, &rc))) return rc;
    fields[n] = fid[q];
    dst[n++] = c1[q];
  }
  return native_pass1(ctx, n, fields, dst);
}

// cmin of T and cmax of S in the native loop (hotcross80-revised2014.f:307,
// 321-324): the background profiles' extremes
void native_bounds(hc_ctx *ctx, double *tmin, double *smax) {
  *tmin = 1e300;
  *smax = -1e300;
  for (int k = 0; k <= ctx->P.nz + 1; ++k) {
    *tmin = fmin(*tmin, ctx->h_tbkg[k]);
    *smax = fmax(*smax, ctx->h_sbkg[k]);
  }
}

// TRACER31 for T, S or C as the native loop calls it (bounds, background,
// vent ghost cells of T and S), without budgets; c1, if set, holds the
// call's pass-1 result already (native_pass1)
int native_tracer_on(hc_ctx *ctx, int field, void *c1, hipStream_t st, const TracerWork *tw) {
  const hc_params &P = ctx->P;
  double tmin, smax;
  native_bounds(ctx, &tmin, &smax);
  const int nz = P.nz;
  if (field == HC_T)
    return tracer_on(ctx, st, tw, HC_T, ctx->d_tbkg, tmin, 1.0e+10, P.dt, 1, P.tsource + ctx->h_tbkg[nz], 0, c1);
  if (field == HC_S)
    return tracer_on(ctx, st, tw, HC_S, ctx->d_sbkg, 0.0, smax, P.dt, 1, P.ssource + ctx->h_sbkg[nz], 0, c1);
  return tracer_on(ctx, st, tw, HC_C, ctx->d_zero, 0.0, 1.0e+10, P.dt, 0, 0.0, 0, c1);
}

int native_tracer(hc_ctx *ctx, int field, void *c1) { return native_tracer_on(ctx, field, c1, ctx->st, nullptr); }

extern "C" int hc_tracer31(hc_ctx *ctx, int field, const double *c_bkg, double cmin, double cmax, double tau3,
                           double *budgets) {
  if (field != HC_T && field != HC_S && field != HC_C) return fail(ctx, HC_EINVAL, "tracer field must be T, S or C");
  if (ctx->P.ismol != 1) return fail(ctx, HC_EINVAL, "only ismol = 1 (two corrective passes) is supported");
  HIPCHK(hipMemcpyAsync(ctx->d_cbkg, c_bkg, (size_t)(ctx->P.nz + 2) * 8, hipMemcpyHostToDevice, ctx->st));
  int want = (budgets != nullptr) && (ctx->mtime % 100 == 0);
  int r = tracer(ctx, field, ctx->d_cbkg, cmin, cmax, tau3, 0, 0.0, want);
  if (r) return r;
  if (want) {
    HIPCHK(hipMemcpyAsync(ctx->h_pin, ctx->tw.budget, 3 * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int q = 0; q < 3; ++q) budgets[q] = ctx->h_pin[q];
  } else {
    HIPCHK(hipStreamSynchronize(ctx->st));
  }
  return HC_OK;
}

extern "C" int hc_vent_ghost(hc_ctx *ctx, int field, double value) {
  VentArgs va = vent_args(ctx);
  if (ctx->prec == HC_FP64) launch_vent_ghost<double>(ctx->st, ctx->L, va, (double *)ctx->f[field], value);
  else launch_vent_ghost<float>(ctx->st, ctx->L, va, (float *)ctx->f[field], value);
  return HC_OK;
}

extern "C" int hc_density(hc_ctx *ctx) {
  PhaseScope ps(ctx, "density");
  if (ctx->prec == HC_FP64) launch_density<double>(ctx->st, ctx->L, ctx->P, ptrs<double>(ctx), ctx->aux);
  else launch_density<float>(ctx->st, ctx->L, ctx->P, ptrs<float>(ctx), ctx->aux);
  return HC_OK;
}

extern "C" int hc_refresh(hc_ctx *ctx) {
  PhaseScope ps(ctx, "refresh");
  int tri[3][3] = {{HC_U, HC_UP, HC_UM}, {HC_V, HC_VP, HC_VM}, {HC_W, HC_WP, HC_WM}};
  for (auto &t : tri) {
    if (ctx->prec == HC_FP64)
      launch_refresh<double>(ctx->st, ctx->L, ctx->leap, ctx->P.assel, (double *)ctx->f[t[0]], (double *)ctx->f[t[1]],
                             (double *)ctx->f[t[2]]);
    else
      launch_refresh<float>(ctx->st, ctx->L, ctx->leap, ctx->P.assel, (float *)ctx->f[t[0]], (float *)ctx->f[t[1]],
                            (float *)ctx->f[t[2]]);
    // u = up: the up buffer becomes u; the old u buffer is overwritten by the
    // next MOMENTUM call on every cell MOMENTUM writes
    void *tmp = ctx->f[t[0]];
    ctx->f[t[0]] = ctx->f[t[1]];
    ctx->f[t[1]] = tmp;
  }
  for (int q = 0; q < 3; ++q) ctx->up_alias[q] = true;
  return HC_OK;
}

extern "C" int hc_cfl_values(hc_ctx *ctx, double out[6]) {
  if (ctx->prec == HC_FP64) launch_cfl<double>(ctx->st, ctx->L, ptrs<double>(ctx), ctx->d_cflpart, ctx->d_cflout);
  else launch_cfl<float>(ctx->st, ctx->L, ptrs<float>(ctx), ctx->d_cflpart, ctx->d_cflout);
  HIPCHK(hipMemcpyAsync(ctx->h_pin, ctx->d_cflout, CFL_NPART * 8, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  const hc_params &P = ctx->P;
  for (int q = 0; q < 6; ++q) out[q] = ctx->h_pin[q] * P.dt / (q < 2 ? P.dx : q < 4 ? P.dy : P.dz);
  ctx->nonfinite = (int64_t)ctx->h_pin[6];
  if (ctx->nonfinite > 0) {
    for (int q = 0; q < 6; ++q) out[q] = NAN;
    return fail(ctx, HC_ENAN, "stop on NaN cfl violation (" + std::to_string(ctx->nonfinite) +
                                  " non-finite velocities at step " + std::to_string(ctx->mtime) + ")");
  }
  double m = fmax(fmax(fmax(out[0], out[2]), fmax(-out[1], -out[3])), fmax(out[4], -out[5]));
  if (m > 0.75) return fail(ctx, HC_ECFL, "stop on impending cfl violation");
  return HC_OK;
}

// the reference's time loop body, hotcross80-revised2014.f:264-350
// MOMENTUM of the native loop with the work velocity (tau3) and REFRESH's
// filter fused in; the filtered field waits in spare[comp] until the loop's
// REFRESH point, where the buffers rotate
int momentum_fused(hc_ctx *ctx, int comp, bool cour, double tau3) {
  PhaseScope ps(ctx, "momentum");
  ctx->up_alias[comp] = false;
  if (!ctx->spare[comp]) {
    int rc = 0;
    if (!(ctx->spare[comp] = dalloc(ctx, ctx->L.n * ctx->esz, &rc))) return rc;
  }
  StepArgs s = step_args(ctx);
  VentArgs va = vent_args(ctx);
  void *c = comp == 0 ? ctx->tw.uc : comp == 1 ? ctx->tw.vc : ctx->tw.wc;
  if (ctx->prec == HC_FP64) {
    MomFuse<double> mf{cour ? (double *)c : nullptr, (double *)ctx->spare[comp], tau3, ctx->P.assel, ctx->leap};
    launch_momentum<double>(ctx->st, comp, ctx->L, s, ctx->P, va, ptrs<double>(ctx), ctx->aux, &mf);
  } else {
    MomFuse<float> mf{cour ? (float *)c : nullptr, (float *)ctx->spare[comp], tau3, ctx->P.assel, ctx->leap};
    launch_momentum<float>(ctx->st, comp, ctx->L, s, ctx->P, va, ptrs<float>(ctx), ctx->aux, &mf);
  }
  return HC_OK;
}

// REFRESH of the native loop after momentum_fused: u <- up (buffer swap),
// um <- the filtered field, the old um buffer becomes the spare
void refresh_rotate(hc_ctx *ctx) {
  int tri[3][3] = {{HC_U, HC_UP, HC_UM}, {HC_V, HC_VP, HC_VM}, {HC_W, HC_WP, HC_WM}};
  for (int c = 0; c < 3; ++c) {
    void *u = ctx->f[tri[c][0]], *up = ctx->f[tri[c][1]], *um = ctx->f[tri[c][2]];
    ctx->f[tri[c][0]] = up;
    ctx->f[tri[c][1]] = u;
    ctx->f[tri[c][2]] = ctx->spare[c];
    ctx->spare[c] = um;
    ctx->up_alias[c] = true;
  }
}

extern "C" int hc_step(hc_ctx *ctx, int32_t nsteps) {
  const hc_params &P = ctx->P;
  int nz = P.nz;
  if (ctx->h_tbkg.empty()) return fail(ctx, HC_EINVAL, "profiles not set (hc_init_case or hc_set_profiles)");
  double tmin = 1e300, smax = -1e300;
  for (int k = 0; k <= nz + 1; ++k) { tmin = fmin(tmin, ctx->h_tbkg[k]); smax = fmax(smax, ctx->h_sbkg[k]); }
  for (int n = 0; n < nsteps; ++n) {
    int m = ctx->next_step++;
    hc_begin_step(ctx, m, ctx->time + P.dt);
    hc_body_force(ctx);
    hc_forces(ctx);
    hc_pressure(ctx);
    const bool fuse = ctx->fuse;
    for (int c = 0; c < 3; ++c) {
      int r = fuse ? momentum_fused(ctx, c, P.toggle_t == 1, P.dt) : momentum(ctx, c);
      if (r) return r;
    }
    hc_viscosity3d(ctx);
    {
      PhaseScope ps(ctx, "source");
      hc_source_terms(ctx, nullptr);
    }
    int budgets = (m % 100 == 0);
    // with fusion, pass 1 of T, S and C runs first, in one launch: it reads
    // the tracers at the step's start and the shared Courant numbers and
    // viscosities, none of which the T and S calls or DENSITY change
    void *c1[3] = {nullptr, nullptr, nullptr};
    if (fuse) {
      int r = tracers_pass1(ctx, c1);
      if (r) return r;
    }
    if (fuse && ctx->par_tracers && !ctx->timing && (P.toggle_s == 1 || P.toggle_c == 1)) {
      // S and C on streams of their own, T on the context stream; DENSITY
      // waits for S, the step's end for C.  Pass 1 of all three ran above,
      // and the passes read nothing another tracer writes.
      int r = par_tracers_init(ctx);
      if (r) return r;
      HIPCHK(hipEventRecord(ctx->tev[0], ctx->st));
      if (P.toggle_s == 1) {
        HIPCHK(hipStreamWaitEvent(ctx->tst[0], ctx->tev[0], 0));
        tracer_on(ctx, ctx->tst[0], &ctx->tw_par[0], HC_S, ctx->d_sbkg, 0.0, smax, P.dt, 1,
                  P.ssource + ctx->h_sbkg[nz], budgets, c1[1]);
        HIPCHK(hipEventRecord(ctx->tev[1], ctx->tst[0]));
      }
      if (P.toggle_c == 1) {
        HIPCHK(hipStreamWaitEvent(ctx->tst[1], ctx->tev[0], 0));
        tracer_on(ctx, ctx->tst[1], &ctx->tw_par[1], HC_C, ctx->d_zero, 0.0, 1.0e+10, P.dt, 0, 0.0, budgets, c1[2]);
        HIPCHK(hipEventRecord(ctx->tev[2], ctx->tst[1]));
      }
      if (P.toggle_t == 1)
        tracer(ctx, HC_T, ctx->d_tbkg, tmin, 1.0e+10, P.dt, 1, P.tsource + ctx->h_tbkg[nz], budgets, c1[0]);
      if (P.toggle_s == 1) HIPCHK(hipStreamWaitEvent(ctx->st, ctx->tev[1], 0));
      if (P.bit3 == 1) hc_density(ctx);
      if (P.toggle_c == 1) HIPCHK(hipStreamWaitEvent(ctx->st, ctx->tev[2], 0));
    } else {
    if (P.toggle_t == 1) {
      if (!fuse) hc_work_velocity(ctx, P.dt);
      tracer(ctx, HC_T, ctx->d_tbkg, tmin, 1.0e+10, P.dt, 1, P.tsource + ctx->h_tbkg[nz], budgets, c1[0]);
    }
    if (P.toggle_s == 1)
      tracer(ctx, HC_S, ctx->d_sbkg, 0.0, smax, P.dt, 1, P.ssource + ctx->h_sbkg[nz], budgets, c1[1]);
    if (P.bit3 == 1) hc_density(ctx);
    if (P.toggle_c == 1) tracer(ctx, HC_C, ctx->d_zero, 0.0, 1.0e+10, P.dt, 0, 0.0, budgets, c1[2]);
    }
    if (fuse) {
      PhaseScope ps(ctx, "refresh");
      refresh_rotate(ctx);
    } else {
      hc_refresh(ctx);
    }
    if (ctx->stats_every > 0 && (m + 1) % ctx->stats_every == 0) stats_accumulate(ctx);
    if (m % 10 == 0) {
      double cfl[6];
      int r = hc_cfl_values(ctx, cfl);
      if (r) return r;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, HC_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  }
  return HC_OK;
}

extern "C" int hc_clock(hc_ctx *ctx, double *time, int32_t *mtime, int32_t *next_step) {
  if (!ctx) return HC_EINVAL;
  if (time) *time = ctx->time;
  if (mtime) *mtime = ctx->mtime;
  if (next_step) *next_step = ctx->next_step;
  return HC_OK;
}

// noise draws of the native loop's generator (for checkers replaying it)
extern "C" int hc_source_noise(hc_ctx *ctx, int32_t mtime, double *out) {
  for (int n = 0; n < ctx->npatch; ++n) {
    out[3 * n] = hash_uniform(ctx->seed, (uint64_t)mtime, (uint64_t)n, 0);
    out[3 * n + 1] = hash_uniform(ctx->seed, (uint64_t)mtime, (uint64_t)n, 1);
    out[3 * n + 2] = hash_uniform(ctx->seed, (uint64_t)mtime, (uint64_t)n, 2);
  }
  return ctx->npatch;
}

extern "C" int hc_phase_timing(hc_ctx *ctx, int enable) {
  hipStreamSynchronize(ctx->st);
  for (auto &p : ctx->phases) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
  ctx->phases.clear();
  ctx->timing = enable != 0;
  return HC_OK;
}

extern "C" int hc_phase_times(hc_ctx *ctx, const char **names, double *ms, int32_t *counts, int32_t cap) {
  hipStreamSynchronize(ctx->st);
  int n = 0;
  for (auto &p : ctx->phases) {
    if (p.pending) {
      float t = 0;
      hipEventElapsedTime(&t, p.a, p.b);
      p.ms += t;
      p.count++;
      p.pending = false;
    }
    if (n < cap) {
      names[n] = p.name;
      ms[n] = p.ms;
      counts[n] = p.count;
    }
    n++;
  }
  return n;
}
