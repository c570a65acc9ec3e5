// hc_stats.hip - running time sums over a slab of horizontal planes, for the
// plume statistics: time-mean w and T on the plane 20 m above the vent, the
// turbulent kinetic energy and the heat flux w'T'.  The reference writes the
// fields of such planes at 1 Hz (the y-slab writer put_block1.f:242-330, and
// the cone file of README.md:26) and leaves the averaging to post-processing;
// here the sums stay in HBM, are added to after every `every`-th step of the
// native loop, and are read once at the end of a period.
#include <hip/hip_runtime.h>

#include "hc_ctx.h"

namespace hc {

#define F3(a, i, j, k) (a)[IX(L, (i), (j), (k))]

// per cell (i,j,k) of the slab, FP64 sums of: w at the w point (i,j,k), T,
// the cell-centred u, v, w, their squares, T^2 and (cell-centred w)*T.
// w is kept as the model stores it (positive toward increasing k, which is
// downward); the writers flip its sign (put_block1.f:265).
template <typename T>
__global__ void __launch_bounds__(256) k_stats(Layout L, int k0, size_t plane, const T *__restrict__ u,
                                               const T *__restrict__ v, const T *__restrict__ w,
                                               const T *__restrict__ t, double *__restrict__ sums, size_t n) {
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x), j = 1 + (int)blockIdx.y, k = k0 + (int)blockIdx.z;
  if (i > L.nx) return;
  const double uc = 0.5 * ((double)F3(u, i - 1, j, k) + (double)F3(u, i, j, k));
  const double vc = 0.5 * ((double)F3(v, i, j - 1, k) + (double)F3(v, i, j, k));
  const double wc = 0.5 * ((double)F3(w, i, j, k - 1) + (double)F3(w, i, j, k));
  const double tt = (double)F3(t, i, j, k);
  const size_t q = (size_t)(i - 1) + (size_t)L.nx * (j - 1) + plane * blockIdx.z;
  const double vals[HC_NSTAT] = {(double)F3(w, i, j, k), tt, uc, vc, wc, uc * uc, vc * vc, wc * wc, tt * tt, wc * tt};
#pragma unroll
  for (int s = 0; s < HC_NSTAT; ++s) sums[s * n + q] += vals[s];
}

}  // namespace hc

static size_t stats_cells(const hc_ctx *ctx) {
  return (size_t)ctx->P.nx * ctx->P.ny * (size_t)(ctx->stats_k1 - ctx->stats_k0 + 1);
}

int stats_accumulate(hc_ctx *ctx) {
  const Layout &L = ctx->L;
  const size_t plane = (size_t)L.nx * L.ny, n = stats_cells(ctx);
  dim3 g((L.nx + 255) / 256, L.ny, ctx->stats_k1 - ctx->stats_k0 + 1);
  if (ctx->prec == HC_FP64)
    hipLaunchKernelGGL(k_stats<double>, g, dim3(256), 0, ctx->st, L, ctx->stats_k0, plane,
                       (const double *)ctx->f[HC_U], (const double *)ctx->f[HC_V], (const double *)ctx->f[HC_W],
                       (const double *)ctx->f[HC_T], ctx->d_stats, n);
  else
    hipLaunchKernelGGL(k_stats<float>, g, dim3(256), 0, ctx->st, L, ctx->stats_k0, plane,
                       (const float *)ctx->f[HC_U], (const float *)ctx->f[HC_V], (const float *)ctx->f[HC_W],
                       (const float *)ctx->f[HC_T], ctx->d_stats, n);
  ctx->stats_count++;
  return HC_OK;
}

void stats_free(hc_ctx *ctx) {
  if (ctx->d_stats) hipFree(ctx->d_stats);
  ctx->d_stats = nullptr;
  ctx->stats_k0 = ctx->stats_k1 = 0;
  ctx->stats_every = 0;
}

extern "C" int hc_stats_reset(hc_ctx *ctx, int32_t k0, int32_t k1, int32_t every) {
  if (!ctx) return HC_EINVAL;
  if (k0 == 0 && k1 == 0) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    stats_free(ctx);
    return HC_OK;
  }
  if (k0 < 1 || k1 < k0 || k1 > ctx->P.nz) return fail(ctx, HC_EINVAL, "statistics slab must satisfy 1 <= k0 <= k1 <= nz");
  if (every < 0) return fail(ctx, HC_EINVAL, "statistics interval must be >= 0");
  const size_t bytes = (size_t)HC_NSTAT * ctx->P.nx * ctx->P.ny * (size_t)(k1 - k0 + 1) * 8;
  if (ctx->d_stats && (k1 - k0 != ctx->stats_k1 - ctx->stats_k0)) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    stats_free(ctx);
  }
  if (!ctx->d_stats) {
    int rc = 0;
    if (!(ctx->d_stats = (double *)dalloc(ctx, bytes, &rc))) return rc;
  }
  HIPCHK(hipMemsetAsync(ctx->d_stats, 0, bytes, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  ctx->stats_k0 = k0;
  ctx->stats_k1 = k1;
  ctx->stats_every = every;
  ctx->stats_count = 0;
  return HC_OK;
}

extern "C" int hc_stats_accumulate(hc_ctx *ctx) {
  if (!ctx->d_stats) return fail(ctx, HC_EINVAL, "no statistics slab (hc_stats_reset)");
  int r = stats_accumulate(ctx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, HC_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return r;
}

extern "C" int hc_stats_get(hc_ctx *ctx, double *sums, int32_t *count) {
  if (!ctx->d_stats) return fail(ctx, HC_EINVAL, "no statistics slab (hc_stats_reset)");
  HIPCHK(hipMemcpyAsync(sums, ctx->d_stats, (size_t)HC_NSTAT * stats_cells(ctx) * 8, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  *count = ctx->stats_count;
  return HC_OK;
}
