// hc_cfl_static.hip - CFL_VALUES over caller-owned device velocity fields,
// with no allocation: the reduction's partial results live in a static
// device array (`scratch`, below) instead of context memory.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc); they hold
// their velocities in device buffers of their own, laid out as the library's
// fields (hotcross.h, hc_cfl_values_noalloc).  The kernels are the ones
// hc_cfl_values launches (k_cfl, k_cfl_final), so both entries give the same
// values bit for bit.
#include <mutex>

#include "hc_ctx.h"

namespace hc {
namespace {

// partials of k_cfl (CFL_NPART values for each of at most CFL_BLOCKS blocks)
// followed by the CFL_NPART results of k_cfl_final: 57,400 bytes, held by
// every GPU the library runs on until the process ends
constexpr size_t SCRATCH_DOUBLES = (size_t)CFL_NPART * CFL_BLOCKS + CFL_NPART;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[MAX_DEVICES];  // device address of `scratch` on each device

// The device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU holds its own copy at its own
// address; the array's name is a host-side stand-in, never passed to a
// kernel or a copy).
hipError_t scratch_on_device(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_ptr[dev] = (double *)p;
  }
  *out = scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace
}  // namespace hc

using namespace hc;

extern "C" int hc_cfl_values_noalloc(int32_t nx, int32_t ny, int32_t nz, int precision, const void *u, const void *v,
                                     const void *w, double dx, double dy, double dz, double dt, void *stream,
                                     double out[6]) {
  if (!u || !v || !w || !out) return HC_EINVAL;
  if (precision != HC_FP64 && precision != HC_FP32) return HC_EINVAL;
  if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || !(dy > 0) || !(dz > 0) || !(dt > 0)) return HC_EINVAL;
  const Layout L = make_layout(nx, ny, nz);
  // the kernels index in signed 32-bit arithmetic (hc_create's limit)
  if (L.n >= ((size_t)1 << 31)) return HC_EINVAL;
  // the reduction never uses more than CFL_BLOCKS blocks, so the scratch
  // size does not depend on the grid
  const int nrows = (ny + 1) * (nz + 1);
  if ((size_t)CFL_NPART * std::min(CFL_BLOCKS, (nrows + 3) / 4) + CFL_NPART > SCRATCH_DOUBLES) return HC_EINVAL;
  double *sp = nullptr;
  if (scratch_on_device(&sp) != hipSuccess) return HC_EDEVICE;
  double *part = sp, *res = sp + (size_t)CFL_NPART * CFL_BLOCKS;
  hipStream_t st = (hipStream_t)stream;
  FieldPtrs<double> f64{};
  FieldPtrs<float> f32{};
  if (precision == HC_FP64) {
    f64.u = (double *)u; f64.v = (double *)v; f64.w = (double *)w;
    launch_cfl<double>(st, L, f64, part, res);
  } else {
    f32.u = (float *)u; f32.v = (float *)v; f32.w = (float *)w;
    launch_cfl<float>(st, L, f32, part, res);
  }
  if (hipGetLastError() != hipSuccess) return HC_EDEVICE;
  double h[CFL_NPART];
  if (hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st) != hipSuccess) return HC_EDEVICE;
  if (hipStreamSynchronize(st) != hipSuccess) return HC_EDEVICE;
  for (int q = 0; q < 6; ++q) out[q] = h[q] * dt / (q < 2 ? dx : q < 4 ? dy : dz);
  if (h[6] > 0) {
    for (int q = 0; q < 6; ++q) out[q] = NAN;
    return HC_ENAN;
  }
  const double m = fmax(fmax(fmax(out[0], out[2]), fmax(-out[1], -out[3])), fmax(out[4], -out[5]));
  return m > 0.75 ? HC_ECFL : HC_OK;
}
