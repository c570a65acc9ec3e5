// hc_cell.h - per-cell formulas of the timestep (FORCES, VISCOSITY3D,
// DENSITY, TRACER31 pass 1, BOUNDARY_CONDITIONS) shared by the forward
// kernels, the adjoint kernels and host builds of both (every function here
// compiles for the host and the device).
#pragma once
#include "hc_common.h"
#include "hc_kernels.h"
#include "hc_mpdata.h"

namespace hc {

#define F3(a, i, j, k) (a)[IX(L, (i), (j), (k))]

// FORCES, hotcross80-revised2014.f:890-1139.  azu/azv are formed on the fly;
// their never-recomputed boundary entries are the persistent accumulators
// acc[i] (the entrance ramp of :953-956 adds (7-i)*0.01/7 to them per call).
template <typename T, class V>
HC_HD inline T azu_at(const Layout &L, const V &ssv, const double *azzmin,
                           const StepArgs &s, int i, int j, int k) {
  // azu(i,j,k) for i=0..nx, j=1..ny, k=0..nz
  if (i == 0) return k >= 1 ? (T)s.acc[0] : (T)0;
  if (i == L.nx) return (T)0;
  T v = (T).25 * (ssv(i, j, k) + ssv(i + 1, j, k) + ssv(i, j, k + 1) + ssv(i + 1, j, k + 1)) +
        (T)azzmin[k];
  if (i <= 7 && k >= 1) v = v + (T)((double)(7 - i) * .01 / 7.0);
  return v;
}

template <typename T, class V>
HC_HD inline T azv_at(const Layout &L, const V &ssv, const double *azzmin,
                           const StepArgs &s, int i, int j, int k) {
  // azv(i,j,k) for i=1..nx, j=0..ny, k=0..nz
  if (j == 0 || j == L.ny) {
    if (k == 0 || i > 7) return (T)0;
    return (T)s.acc[i];
  }
  T v = (T).25 * (ssv(i, j, k) + ssv(i, j + 1, k) + ssv(i, j, k + 1) + ssv(i, j + 1, k + 1)) +
        (T)azzmin[k];
  if (i <= 7 && k >= 1) v = v + (T)((double)(7 - i) * .01 / 7.0);
  return v;
}

template <typename T, bool P2, class V>
HC_HD inline void forces_x_at(const Layout &L, int i, int j, int k, const StepArgs &s, const Phys &ph, const V &u,
                                                  const V &v, const V &w,
                                                  const V &um, const V &vm,
                                                  const V &wm, const V &ss,
                                                  const V &ssv, const double *__restrict__ azzmin,
                                                  const double *__restrict__ alphau, T *__restrict__ fx) {
  const Sp<T, P2> dx((T)s.dx, (T)s.idx), dy((T)s.dy, (T)s.idy), dz((T)s.dz, (T)s.idz);
  const T bit2 = (T)ph.bit2, bit5 = (T)ph.bit5;
  const T axmin = (T)ph.axmin, aymin = (T)ph.aymin;
  const T coe_4x = (T)((double)ph.bit4 / 4.0 / s.dx), coe_4y = (T)((double)ph.bit4 / 4.0 / s.dy),
          coe_4z = (T)((double)ph.bit4 / 4.0 / s.dz);
  const T coe_7 = (T)(ph.bit7 * s.fc / 4.0), coe_9 = (T)(ph.bit9 * s.fc_y / 4.0), coe_10 = (T)ph.bit10;
  T axu_p = ss(i + 1, j, k) + axmin, axu_c = ss(i, j, k) + axmin;
  T ayu_c = (T).25 * (ss(i, j, k) + ss(i + 1, j, k) + ss(i, j + 1, k) + ss(i + 1, j + 1, k)) + aymin;
  T ayu_m = (T).25 * (ss(i, j - 1, k) + ss(i + 1, j - 1, k) + ss(i, j, k) + ss(i + 1, j, k)) + aymin;
  T azu_c = azu_at<T>(L, ssv, azzmin, s, i, j, k), azu_m = azu_at<T>(L, ssv, azzmin, s, i, j, k - 1);
  T umc = um(i, j, k);
  T dis = (T)2.0 * bit5 * (axu_p * (um(i + 1, j, k) - umc) - axu_c * (umc - um(i - 1, j, k))) / dx / dx +
          bit5 * (ayu_c * (um(i, j + 1, k) - umc) - ayu_m * (umc - um(i, j - 1, k))) / dy / dy +
          bit5 * (ayu_c * (vm(i + 1, j, k) - vm(i, j, k)) - ayu_m * (vm(i + 1, j - 1, k) - vm(i, j - 1, k))) / dx / dy +
          bit2 * (azu_c * (um(i, j, k + 1) - umc) - azu_m * (umc - um(i, j, k - 1))) / dz / dz +
          bit2 * (azu_c * (wm(i + 1, j, k) - wm(i, j, k)) - azu_m * (wm(i + 1, j, k - 1) - wm(i, j, k - 1))) / dx / dz;
  T uc_ = u(i, j, k);
  T a1 = u(i + 1, j, k) + uc_, a2 = uc_ + u(i - 1, j, k);
  T r = dis + (T)ph.bit3 * (T)0.0 - (T)0.5 * (T)alphau[(i + 1) + (L.nx + 3) * j] * (umc - (T)s.ubkg_m) -
        coe_4x * (a1 * a1 - a2 * a2) -
        coe_4z * ((w(i, j, k) + w(i + 1, j, k)) * (u(i, j, k + 1) + uc_) -
                  (w(i, j, k - 1) + w(i + 1, j, k - 1)) * (uc_ + u(i, j, k - 1))) -
        coe_4y * ((v(i, j, k) + v(i + 1, j, k)) * (u(i, j + 1, k) + uc_) -
                  (v(i, j - 1, k) + v(i + 1, j - 1, k)) * (uc_ + u(i, j - 1, k))) +
        coe_7 * (v(i + 1, j, k) + v(i, j, k) + v(i + 1, j - 1, k) + v(i, j - 1, k)) -
        coe_9 * (w(i + 1, j, k) + w(i + 1, j, k - 1) + w(i, j, k) + w(i, j, k - 1)) +
        coe_10 * (T)s.fbx;
  F3(fx, i, j, k) = r;
}

template <typename T, bool P2, class V>
HC_HD inline void forces_y_at(const Layout &L, int i, int j, int k, const StepArgs &s, const Phys &ph, const V &u,
                                                  const V &v, const V &w,
                                                  const V &um, const V &vm,
                                                  const V &wm, const V &ss,
                                                  const V &ssv, const double *__restrict__ azzmin,
                                                  const double *__restrict__ alphav, T *__restrict__ fy) {
  const Sp<T, P2> dx((T)s.dx, (T)s.idx), dy((T)s.dy, (T)s.idy), dz((T)s.dz, (T)s.idz);
  const T bit2 = (T)ph.bit2, bit5 = (T)ph.bit5;
  const T axmin = (T)ph.axmin, aymin = (T)ph.aymin;
  const T coe_4x = (T)((double)ph.bit4 / 4.0 / s.dx), coe_4y = (T)((double)ph.bit4 / 4.0 / s.dy),
          coe_4z = (T)((double)ph.bit4 / 4.0 / s.dz);
  const T coe_7 = (T)(ph.bit7 * s.fc / 4.0), coe_10 = (T)ph.bit10;
  T axv_c = (T).25 * (ss(i, j, k) + ss(i + 1, j, k) + ss(i, j + 1, k) + ss(i + 1, j + 1, k)) + axmin;
  T axv_m = (T).25 * (ss(i - 1, j, k) + ss(i, j, k) + ss(i - 1, j + 1, k) + ss(i, j + 1, k)) + axmin;
  T ayv_p = ss(i, j + 1, k) + aymin, ayv_c = ss(i, j, k) + aymin;
  T azv_c = azv_at<T>(L, ssv, azzmin, s, i, j, k), azv_m = azv_at<T>(L, ssv, azzmin, s, i, j, k - 1);
  T vmc = vm(i, j, k);
  T dis = +bit5 * (axv_c * (vm(i + 1, j, k) - vmc) - axv_m * (vmc - vm(i - 1, j, k))) / dx / dx +
          bit5 * (axv_c * (um(i, j + 1, k) - um(i, j, k)) - axv_m * (um(i - 1, j + 1, k) - um(i - 1, j, k))) / dx / dy +
          (T)2.0 * bit5 * (ayv_p * (vm(i, j + 1, k) - vmc) - ayv_c * (vmc - vm(i, j - 1, k))) / dy / dy +
          bit2 * (azv_c * (vm(i, j, k + 1) - vmc) - azv_m * (vmc - vm(i, j, k - 1))) / dz / dz +
          bit2 * (azv_c * (wm(i, j + 1, k) - wm(i, j, k)) - azv_m * (wm(i, j + 1, k - 1) - wm(i, j, k - 1))) / dy / dz;
  T vc_ = v(i, j, k);
  T b1 = v(i, j + 1, k) + vc_, b2 = vc_ + v(i, j - 1, k);
  T r = dis + (T)ph.bit3 * (T)0.0 - (T)0.5 * (T)alphav[i + (L.nx + 2) * (j + 1)] * (vmc - (T)s.vbkg_m) -
        coe_4y * (b1 * b1 - b2 * b2) -
        coe_4z * ((v(i, j, k + 1) + vc_) * (w(i, j + 1, k) + w(i, j, k)) -
                  (vc_ + v(i, j, k - 1)) * (w(i, j + 1, k - 1) + w(i, j, k - 1))) -
        coe_4x * ((v(i + 1, j, k) + vc_) * (u(i, j, k) + u(i, j + 1, k)) -
                  (vc_ + v(i - 1, j, k)) * (u(i - 1, j, k) + u(i - 1, j + 1, k))) -
        coe_7 * (u(i, j, k) + u(i - 1, j, k) + u(i, j + 1, k) + u(i - 1, j + 1, k)) +
        coe_10 * (T)s.fby;
  F3(fy, i, j, k) = r;
}

template <typename T, bool P2, class V, class VR>
HC_HD inline void forces_z_at(const Layout &L, int i, int j, int k, const StepArgs &s, const Phys &ph, const V &u,
                                                  const V &v, const V &w,
                                                  const V &um, const V &vm,
                                                  const V &wm, const V &ss,
                                                  const V &ssv, const VR &rhop,
                                                  const double *__restrict__ azzmin, const double *__restrict__ alphae,
                                                  const double *__restrict__ alphaw, T *__restrict__ fz) {
  const Sp<T, P2> dx((T)s.dx, (T)s.idx), dy((T)s.dy, (T)s.idy), dz((T)s.dz, (T)s.idz);
  const T bit2 = (T)ph.bit2, bit5 = (T)ph.bit5;
  const T axmin = (T)ph.axmin;
  const T coe_4x = (T)((double)ph.bit4 / 4.0 / s.dx), coe_4y = (T)((double)ph.bit4 / 4.0 / s.dy),
          coe_4z = (T)((double)ph.bit4 / 4.0 / s.dz);
  const T coe_9 = (T)(ph.bit9 * s.fc_y / 4.0), coe_10 = (T)ph.bit10;
  T ahp = ss(i, j, k + 1) + axmin, ahm = ss(i, j, k) + axmin;
  T zk = (T)azzmin[k];
  T s0 = ssv(i, j, k + 1), s1 = ssv(i, j, k);
  T avxp = (T)0.25 * (s0 + s1 + ssv(i + 1, j, k + 1) + ssv(i + 1, j, k)) + zk;
  T avxm = (T)0.25 * (s0 + s1 + ssv(i - 1, j, k + 1) + ssv(i - 1, j, k)) + zk;
  T avyp = (T)0.25 * (s0 + s1 + ssv(i, j + 1, k + 1) + ssv(i, j + 1, k)) + zk;
  T avym = (T)0.25 * (s0 + s1 + ssv(i, j - 1, k + 1) + ssv(i, j - 1, k)) + zk;
  T wmc = wm(i, j, k);
  T dis = (T)2.0 * bit5 * (ahp * (wm(i, j, k + 1) - wmc) - ahm * (wmc - wm(i, j, k - 1))) / dz / dz +
          bit2 * (avxp * (wm(i + 1, j, k) - wmc) - avxm * (wmc - wm(i - 1, j, k))) / dx / dx +
          bit2 * (avxp * (um(i, j, k) - um(i, j, k + 1)) - avxm * (um(i - 1, j, k) - um(i - 1, j, k + 1))) / dz / dx +
          bit2 * (avyp * (wm(i, j + 1, k) - wmc) - avym * (wmc - wm(i, j - 1, k))) / dy / dy +
          bit2 * (avyp * (vm(i, j, k) - vm(i, j, k + 1)) - avym * (vm(i, j - 1, k) - vm(i, j - 1, k + 1))) / dz / dy;
  T wc_ = w(i, j, k);
  T c1 = w(i, j, k + 1) + wc_, c2 = wc_ + w(i, j, k - 1);
  // alphaee(i,j,k) = max(alphae(i,j), (alphaw(k)+alphaw(k-1))/2), sponge80.f:115-119
  // (alphae here is the mask before the periodic copies of :128-138)
  double ae = fmax(alphae[i + (L.nx + 2) * j], (alphaw[k + 1] + alphaw[k]) / 2.0);
  const UDiv<T> rho0 = udiv_t<T>(ph.rho0, s.rrho0, s.rrho0f);
  T r = dis + rho0((T)ph.bit3 * (T)ph.g * (rhop(i, j, k) + rhop(i, j, k + 1))) / (T)2.0 -
        (T)0.0 * (T)alphaw[k + 1] * wmc - (T)0.5 * (T)ae * wmc - coe_4z * (c1 * c1 - c2 * c2) -
        coe_4x * ((w(i + 1, j, k) + wc_) * (u(i, j, k) + u(i, j, k + 1)) -
                  (wc_ + w(i - 1, j, k)) * (u(i - 1, j, k) + u(i - 1, j, k + 1))) -
        coe_4y * ((w(i, j + 1, k) + wc_) * (v(i, j, k) + v(i, j, k + 1)) -
                  (wc_ + w(i, j - 1, k)) * (v(i, j - 1, k) + v(i, j - 1, k + 1))) +
        coe_9 * (u(i, j, k + 1) + u(i - 1, j, k + 1) + u(i, j, k) + u(i - 1, j, k)) +
        coe_10 * (T)0.0;
  F3(fz, i, j, k) = r;
}

// ---------------------------------------------------------------------------
// VISCOSITY3D, smagorinsky_split_mixing-revised2014.f:5-150.  ss is rewritten
// over its whole box (the reference zeroes it first, :24); ss_v over its
// interior and z ghosts.  The complex root at :112 keeps its real part,
// sqrt(max(0,x)).
template <typename T, bool P2, class V>
HC_HD inline T ss_val(const Layout &L, const Phys &ph, const StepArgs &s, const V &um, const V &vm,
                           const V &wm, int i, int j, int k) {
  const Sp<T, P2> dx((T)s.dx, (T)s.idx), dy((T)s.dy, (T)s.idy), dz((T)s.dz, (T)s.idz);
  double ls = sqrt(s.dx * s.dy);
  T coef = (T)(ph.ismag_h * ((ph.cs_h * ls) * (ph.cs_h * ls)));
  T s11 = (um(i, j, k) - um(i - 1, j, k)) / dx;
  T s22 = (vm(i, j, k) - vm(i, j - 1, k)) / dy;
  T s33 = (wm(i, j, k) - wm(i, j, k - 1)) / dz;
  T s12 = (T)0.25 *
          ((um(i, j + 1, k) + um(i - 1, j + 1, k) - um(i, j - 1, k) - um(i - 1, j - 1, k)) / dy +
           (vm(i + 1, j, k) + vm(i + 1, j - 1, k) - vm(i - 1, j, k) - vm(i - 1, j - 1, k)) / dx) /
          (T)2.0;
  T q = (T)2.0 * (s11 * s11) + (T)2.0 * (s22 * s22) + (T)4.0 * (s12 * s12) + (T)2.0 * (s33 * s33);
  return coef * sqrt(q + (T)1.0e-30);
}

template <typename T, bool P2, class V, class VR>
HC_HD inline T ssv_val(const Layout &L, const Phys &ph, const StepArgs &s, const V &um, const V &vm,
                            const V &wm, const VR &rhop, const double *rho_bkg, int i, int j, int k) {
  const Sp<T, P2> dx((T)s.dx, (T)s.idx), dy((T)s.dy, (T)s.idy), dz((T)s.dz, (T)s.idz);
  T coef = (T)(ph.ismag_v * ((ph.cs_v * s.dz) * (ph.cs_v * s.dz)));
  T s13 = (T)0.25 *
          ((um(i, j, k - 1) + um(i - 1, j, k - 1) - um(i, j, k + 1) - um(i - 1, j, k + 1)) / dz +
           (wm(i + 1, j, k) + wm(i + 1, j, k - 1) - wm(i - 1, j, k) - wm(i - 1, j, k - 1)) / dx) /
          (T)2.0;
  T s23 = (T)0.25 *
          ((vm(i, j, k - 1) + vm(i, j - 1, k - 1) - vm(i, j, k + 1) - vm(i, j - 1, k + 1)) / dz +
           (wm(i, j + 1, k) + wm(i, j + 1, k - 1) - wm(i, j - 1, k) - wm(i, j - 1, k - 1)) / dy) /
          (T)2.0;
  T sv = (T)4.0 * (s13 * s13) + (T)4.0 * (s23 * s23);
  // Richardson number in FP64: rho' + rho_bkg is ~1027 and its 1-m change is
  // ~1e-4 of that
  double rkm = (double)rhop(i, j, k - 1) + rho_bkg[k - 1];
  double rkp = (double)rhop(i, j, k + 1) + rho_bkg[k + 1];
  // the uniform divisors through their reciprocals (UDiv) and the
  // data-dependent one through ddiv: the same correctly rounded quotients
  const UDiv<double> rho0(ph.rho0, s.rrho0), three(3.0, 1.0 / 3.0);
  double ri = -ddiv(rho0(ph.g * (rkm - rkp) / 2.0 / Sp<double, P2>(s.dz, s.idz)), (double)sv + 1.0e-30);
  double arg = 1.0 + ph.eps - three(ri);
  return coef * sqrt(sv + (T)1.0e-30) * (T)sqrt(fmax(0.0, arg));
}

// one cell of the viscosity box (interior formula, ghost-cell rules)
template <typename T, bool P2, class V, class VR>
HC_HD inline void visc_at(const Layout &L, const StepArgs &s, const Phys &ph, const V &um, const V &vm,
                               const V &wm, const VR &rhop, const double *rho_bkg, T *ss, T *ssv, int i, int j,
                               int k) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  bool gi = (i == 0 || i == nx + 1), gj = (j == 0 || j == ny + 1), gk = (k == 0 || k == nz + 1);
  int ng = (int)gi + (int)gj + (int)gk;
  // ss
  T sval = (T)0;
  if (ng == 0) {
    sval = ss_val<T, P2>(L, ph, s, um, vm, wm, i, j, k);
  } else if (ng == 1 && gi) {
    const T bx = (T)ph.bcondx;
    T a = ss_val<T, P2>(L, ph, s, um, vm, wm, i == 0 ? 1 : nx, j, k);
    T c = ss_val<T, P2>(L, ph, s, um, vm, wm, i == 0 ? nx : 1, j, k);
    sval = a * bx + (T)(1 - ph.bcondx) * c;
  } else if (ng == 1 && gj) {
    const T by = (T)ph.bcondy;
    T a = ss_val<T, P2>(L, ph, s, um, vm, wm, i, j == 0 ? 1 : ny, k);
    T c = ss_val<T, P2>(L, ph, s, um, vm, wm, i, j == 0 ? ny : 1, k);
    sval = a * by + (T)(1 - ph.bcondy) * c;
  }
  F3(ss, i, j, k) = sval;
  // ss_v
  if (ng == 0) {
    F3(ssv, i, j, k) = ssv_val<T, P2>(L, ph, s, um, vm, wm, rhop, rho_bkg, i, j, k);
  } else if (ng == 1 && gk) {
    F3(ssv, i, j, k) = ssv_val<T, P2>(L, ph, s, um, vm, wm, rhop, rho_bkg, i, j, k == 0 ? 1 : nz);
  }
}

// ---------------------------------------------------------------------------
// DENSITY, potential_density80.f:3-71: interior and the six faces.
//
// s**1.5 rounded correctly (but for exact values within ~2^-104 of a
// rounding midpoint): r = sqrt(s) is correctly rounded, so s - r*r is exact
// as one fma; sqrt(s) = r + (s - r*r)/(2r) to ~2^-106 relative, and the
// product s*sqrt(s) is carried as p + (pe + s*c) with pe the exact error of
// s*r.  A host libm pow is accurate to about half an ulp, so both round to
// the same double except for rare near-midpoint values.
HC_HD inline double pow15(double s) {
  if (!(s > 0.0) || isinf(s)) return pow(s, 1.5);
  double r = sqrt(s);
  double e = fma(-r, r, s);
  double c = e / (2.0 * r);
  double p = s * r;
  double pe = fma(s, r, -p);
  return p + fma(s, c, pe);
}

template <typename T>
HC_HD inline double rho_rho(const Phys &ph, double s, double t) {
  if (ph.eos == 1) return ph.eos_rho_ref * (1.0 - ph.eos_alpha * (t - ph.eos_t0) + ph.eos_beta * (s - ph.eos_s0));
  double t2 = t * t, t3 = t2 * t, t4 = t2 * t2, t5 = t4 * t;
  double r0 = 999.842594 + 6.793952e-2 * t - 9.095290e-3 * t * t + 1.001685e-4 * t3 - 1.120083e-6 * t4 +
              6.536332e-9 * t5;
  return r0 + s * (0.824493 - 4.0899e-3 * t + 7.6438e-5 * t * t - 8.2467e-7 * t3 + 5.3875e-9 * t4) +
         pow15(s) * (-5.72466e-3 + 1.0227e-4 * t - 1.6546e-6 * t * t) + 4.8314e-4 * s * s;
}

// pass 1: diffusion (mpdata-hotcross80.f:141-174) + upwind step (:176-184)
template <typename T, class V>
HC_HD inline T pass1_at(const Layout &L, const StepArgs &s, const Phys &ph, const V &c, const V &uc,
                             const V &vc, const V &wc, const V &ss, const V &ssv, const double *azzmin,
                             const double *cbkg, int i, int j, int k) {
  // the divisions by the Prandtl number and by dx, dy, dz through their
  // reciprocals (UDiv: the same correctly rounded quotients)
  const UDiv<double> prd(ph.prandtl, s.rpr);
  const UDiv<T> tpr = udiv_t<T>(ph.prandtl, s.rpr, s.rprf);
  const UDiv<double> ux(s.dx, s.idx), uy(s.dy, s.idy), uz(s.dz, s.idz);
  const T xfac = (T)(ux(ux((double)(1 * ph.bit5k))) * s.tau3 * 1);
  const T yfac = (T)(uy(uy((double)(1 * ph.bit5k))) * s.tau3 * 1);
  const T zfac = (T)(uz(uz((double)ph.bit2k)) * s.tau3 * 1);
  const T b5 = (T)ph.bit5k, b2 = (T)ph.bit2k;
  const T axp = (T)prd(ph.axmin), ayp = (T)prd(ph.aymin);
  T cp = c(i, j, k);
  T kx1 = b5 * (tpr((T).5 * (ss(i, j, k) + ss(i + 1, j, k))) + axp);
  T kx0 = b5 * (tpr((T).5 * (ss(i - 1, j, k) + ss(i, j, k))) + axp);
  T ky1 = b5 * (tpr((T).5 * (ss(i, j, k) + ss(i, j + 1, k))) + ayp);
  T ky0 = b5 * (tpr((T).5 * (ss(i, j - 1, k) + ss(i, j, k))) + ayp);
  // kzt(k) = 0 at k = 0 and k = nz (mixing80-revised2014.f:49-59)
  T kz1 = (k == L.nz) ? (T)0 : b2 * (tpr((T)0.5 * (ssv(i, j, k) + ssv(i, j, k + 1))) + (T)prd(azzmin[k]));
  T kz0 = (k == 1) ? (T)0 : b2 * (tpr((T)0.5 * (ssv(i, j, k - 1) + ssv(i, j, k))) + (T)prd(azzmin[k - 1]));
  T cx = xfac * (kx1 * (c(i + 1, j, k) - cp) - kx0 * (cp - c(i - 1, j, k)));
  T cy = yfac * (ky1 * (c(i, j + 1, k) - cp) - ky0 * (cp - c(i, j - 1, k)));
  T bk = (T)cbkg[k], bkm = (T)cbkg[k - 1], bkp = (T)cbkg[k + 1];
  T cz = zfac * (kz0 * ((c(i, j, k - 1) + bk) - (bkm + cp)) - kz1 * ((cp + bkp) - (bk + c(i, j, k + 1))));
  T cd = cp + (+cx + cy + cz);
  T ce = mp_donor<T>(cp, c(i - 1, j, k), c(i + 1, j, k), c(i, j - 1, k), c(i, j + 1, k),
                     c(i, j, k - 1), c(i, j, k + 1), uc(i - 1, j, k), uc(i, j, k),
                     vc(i, j - 1, k), vc(i, j, k), wc(i, j, k - 1), wc(i, j, k), false, 0, 0, 0, 0,
                     0, 0, 0);
  return cd + ce;
}


// Interior image of a ghost cell of a scalar field under
// BOUNDARY_CONDITIONS (mpdata-hotcross80.f:435-494): z ghosts copy the
// adjacent plane (:444-453); x and y ghosts copy the adjacent cell (bcond = 1)
// or the periodic image (bcond = 0, :455-478); for modulo_x xor modulo_y the
// corner columns take the periodic image in that direction (:485-492).  Every
// ghost's value ends up equal to that of one interior cell, the one returned.
HC_HD inline void bc_source(const Layout &L, const Phys &ph, int i, int j, int k, int &si, int &sj, int &sk) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const bool gi = (i == 0 || i == nx + 1), gj = (j == 0 || j == ny + 1);
  si = i; sj = j; sk = k;
  if (k == 0) sk = 1;
  if (k == nz + 1) sk = nz;
  if (i == 0) si = ph.bcondx ? 1 : nx;
  if (i == nx + 1) si = ph.bcondx ? nx : 1;
  if (j == 0) sj = ph.bcondy ? 1 : ny;
  if (j == ny + 1) sj = ph.bcondy ? ny : 1;
  const int mx = ph.modulo_x, my = ph.modulo_y;
  if (gi && gj && k >= 1 && k <= nz && ((mx == 1 && my == 0) || (mx == 0 && my == 1))) {
    if (mx == 1) si = (i == 0) ? nx : 1;
    else sj = (j == 0) ? ny : 1;
  }
}

}  // namespace hc
