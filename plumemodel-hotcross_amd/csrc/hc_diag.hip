// hc_diag.hip - the diagnostics the reference forms inside VISCOSITY3D and
// MIXING for its output files only: the Richardson number Ri and the
// dissipations epsh = S_H^2 (A_H + axmin), epsv = S_V^2 (A_V + azmin)
// (smagorinsky_split_mixing-revised2014.f:54-59, 99-114), and the
// temperature-variance dissipations chi_v, chi_h
// (mixing80-revised2014.f:64-94).  One cell-parallel pass over the interior,
// run after VISCOSITY3D when the context's diagnostics are on (off by
// default: the timestep does not need them); the ghost cells stay zero, as
// the reference leaves them.
//
// Ri, epsh and epsv need the step's um, vm, wm and rhop as VISCOSITY3D read
// them, which REFRESH and DENSITY overwrite later in the step; chi needs T
// before the tracers and ss, ss_v after VISCOSITY3D.  Right after VISCOSITY3D
// all of them are current (SOURCE_terms and MIXING, which the reference runs
// in between, change none of them).  The formulas follow the reference's
// evaluation order term by term, with true divisions.
#include "hc_common.h"
#include "hc_kernels.h"

namespace hc {

#define F3(a, i, j, k) (a)[IX(L, (i), (j), (k))]

template <typename T>
__global__ void __launch_bounds__(256) k_diag(Layout L, StepArgs s, Phys ph, const T *__restrict__ um,
                                              const T *__restrict__ vm, const T *__restrict__ wm,
                                              const T *__restrict__ rhop, const T *__restrict__ ss,
                                              const T *__restrict__ ssv, const T *__restrict__ t,
                                              const double *__restrict__ rho_bkg, const double *__restrict__ azzmin,
                                              const double *__restrict__ t_bkg, T *__restrict__ ri,
                                              T *__restrict__ epsh, T *__restrict__ epsv, T *__restrict__ chiv,
                                              T *__restrict__ chih) {
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = 1 + (int)blockIdx.y, k = 1 + (int)blockIdx.z;
  if (i > L.nx) return;
  const T dx = (T)s.dx, dy = (T)s.dy, dz = (T)s.dz;
  // VISCOSITY3D, horizontal: S_H^2 and A_H (:31-59)
  const double ls = sqrt(s.dx * s.dy);
  const T coefh = (T)(ph.ismag_h * ((ph.cs_h * ls) * (ph.cs_h * ls)));
  const T s11 = (F3(um, i, j, k) - F3(um, i - 1, j, k)) / dx;
  const T s22 = (F3(vm, i, j, k) - F3(vm, i, j - 1, k)) / dy;
  const T s33 = (F3(wm, i, j, k) - F3(wm, i, j, k - 1)) / dz;
  const T s12 = (T)0.25 *
                ((F3(um, i, j + 1, k) + F3(um, i - 1, j + 1, k) - F3(um, i, j - 1, k) - F3(um, i - 1, j - 1, k)) / dy +
                 (F3(vm, i + 1, j, k) + F3(vm, i + 1, j - 1, k) - F3(vm, i - 1, j, k) - F3(vm, i - 1, j - 1, k)) / dx) /
                (T)2.0;
  const T q = (T)2.0 * (s11 * s11) + (T)2.0 * (s22 * s22) + (T)4.0 * (s12 * s12) + (T)2.0 * (s33 * s33);
  const T ah = coefh * sqrt(q + (T)1.0e-30);
  F3(epsh, i, j, k) = q * (ah + (T)ph.axmin);
  // vertical: S_V^2, Ri (in FP64) and A_V (:72-114)
  const T coefv = (T)(ph.ismag_v * ((ph.cs_v * s.dz) * (ph.cs_v * s.dz)));
  const T s13 = (T)0.25 *
                ((F3(um, i, j, k - 1) + F3(um, i - 1, j, k - 1) - F3(um, i, j, k + 1) - F3(um, i - 1, j, k + 1)) / dz +
                 (F3(wm, i + 1, j, k) + F3(wm, i + 1, j, k - 1) - F3(wm, i - 1, j, k) - F3(wm, i - 1, j, k - 1)) / dx) /
                (T)2.0;
  const T s23 = (T)0.25 *
                ((F3(vm, i, j, k - 1) + F3(vm, i, j - 1, k - 1) - F3(vm, i, j, k + 1) - F3(vm, i, j - 1, k + 1)) / dz +
                 (F3(wm, i, j + 1, k) + F3(wm, i, j + 1, k - 1) - F3(wm, i, j - 1, k) - F3(wm, i, j - 1, k - 1)) / dy) /
                (T)2.0;
  const T sv = (T)4.0 * (s13 * s13) + (T)4.0 * (s23 * s23);
  const double rkm = (double)F3(rhop, i, j, k - 1) + rho_bkg[k - 1];
  const double rkp = (double)F3(rhop, i, j, k + 1) + rho_bkg[k + 1];
  const double r = -(ph.g * (rkm - rkp) / 2.0 / s.dz / ph.rho0 / ((double)sv + 1.0e-30));
  F3(ri, i, j, k) = (T)r;
  const T av = coefv * sqrt(sv + (T)1.0e-30) * (T)sqrt(fmax(0.0, 1.0 + ph.eps - r / 3.0));
  F3(epsv, i, j, k) = sv * (av + (T)ph.azmin);
  // MIXING: the diffusivities at the cell (:22-59) from ss, ss_v after
  // VISCOSITY3D, and chi from T - t_bkg (:64-94)
  const T pr = (T)ph.prandtl;
  const T kxt = (T)ph.bit5k * (((T)0.5 * (F3(ss, i, j, k) + F3(ss, i + 1, j, k))) / pr + (T)ph.axmin / pr);
  const T kyt = (T)ph.bit5k * (((T)0.5 * (F3(ss, i, j, k) + F3(ss, i, j + 1, k))) / pr + (T)ph.aymin / pr);
  const T kzt = k <= L.nz - 1
                    ? (T)ph.bit2k * (((T)0.5 * (F3(ssv, i, j, k) + F3(ssv, i, j, k + 1))) / pr + (T)azzmin[k] / pr)
                    : (T)0;
  auto tp = [&](int a, int b, int c) { return F3(t, a, b, c) - (T)t_bkg[c]; };
  const T d = (tp(i, j, k - 1) - tp(i, j, k + 1)) / ((T)2.0 * dz);
  F3(chiv, i, j, k) = kzt * (d * d);
  const T a = (tp(i + 1, j, k) - tp(i - 1, j, k)) / ((T)2.0 * dx);
  const T b = (tp(i, j + 1, k) - tp(i, j - 1, k)) / ((T)2.0 * dy);
  F3(chih, i, j, k) = kxt * (a * a) + kyt * (b * b);
}

template <typename T>
void launch_diag(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                 const Aux &aux, const double *t_bkg, void *const out[5]) {
  dim3 g((unsigned)((L.nx + 255) / 256), (unsigned)L.ny, (unsigned)L.nz);
  hipLaunchKernelGGL(k_diag<T>, g, dim3(256), 0, st, L, s, ph, (const T *)f.um, (const T *)f.vm, (const T *)f.wm,
                     (const T *)f.rhop, (const T *)f.ss, (const T *)f.ssv, (const T *)f.t, (const double *)aux.rho_bkg,
                     (const double *)aux.azzmin, t_bkg, (T *)out[0], (T *)out[1], (T *)out[2], (T *)out[3],
                     (T *)out[4]);
}
template void launch_diag<double>(hipStream_t, const Layout &, const StepArgs &, const Phys &,
                                  const FieldPtrs<double> &, const Aux &, const double *, void *const[5]);
template void launch_diag<float>(hipStream_t, const Layout &, const StepArgs &, const Phys &,
                                 const FieldPtrs<float> &, const Aux &, const double *, void *const[5]);

}  // namespace hc
