// hc_adj.h - hand-written reverse (adjoint) formulas of the HotCross
// timestep, one function per forward formula of a cell or face.
//
// Each function takes the forward inputs through read views r(i,j,k) and the
// adjoint of the formula's output (lam), recomputes the forward
// intermediates it needs, and adds lam * d(output)/d(input) to the adjoint
// of every input through accumulators a(i,j,k,value): atomic adds in the
// device kernels, plain += in host builds.  The derivative of every
// non-smooth operation follows the branch the forward took: |x| -> sign(x)
// (0 at x = 0), max(0,x) -> 0 below the kink, the limiter's zeroed fluxes
// -> 0 (the flags are recorded from the forward values, not differentiated).
//
// Scalar parameters: gflow collects d/d(flow_scale) through the crossflow
// scalars of the step (their derivatives in FlowSens).
#pragma once
#include "hc_cell.h"

namespace hc {

template <typename T>
struct RView {
  const T *p;
  Layout L;
  HC_HD T operator()(int i, int j, int k) const { return p[IX(L, i, j, k)]; }
};

template <typename T>
HC_HD inline T adj_sign(T x) {
  return x > (T)0 ? (T)1 : (x < (T)0 ? (T)-1 : (T)0);
}

// ---------------------------------------------------------------------------
// FORCES (hotcross80-revised2014.f:961-1134; forward: forces_x/y/z_at)

// azu(i,j,k) = .25*(ssv(i,j,k)+ssv(i+1,j,k)+ssv(i,j,k+1)+ssv(i+1,j,k+1)) + const
// for 0 < i < nx; the entries i = 0 and i = nx are constants
template <class A, typename T>
HC_HD inline void adj_azu(const Layout &L, int i, int j, int k, T lam, const A &assv) {
  if (i == 0 || i == L.nx) return;
  const T q = (T).25 * lam;
  assv(i, j, k, q); assv(i + 1, j, k, q); assv(i, j, k + 1, q); assv(i + 1, j, k + 1, q);
}
template <class A, typename T>
HC_HD inline void adj_azv(const Layout &L, int i, int j, int k, T lam, const A &assv) {
  if (j == 0 || j == L.ny) return;
  const T q = (T).25 * lam;
  assv(i, j, k, q); assv(i, j + 1, k, q); assv(i, j, k + 1, q); assv(i, j + 1, k + 1, q);
}

struct AdjConsts {
  double idx, idy, idz;
  double bit2, bit5, axmin, aymin;
  double coe_4x, coe_4y, coe_4z, coe_7, coe_9, coe_10;
};
HC_HD inline AdjConsts adj_consts(const StepArgs &s, const Phys &ph) {
  AdjConsts c;
  c.idx = 1.0 / s.dx; c.idy = 1.0 / s.dy; c.idz = 1.0 / s.dz;
  c.bit2 = ph.bit2; c.bit5 = ph.bit5; c.axmin = ph.axmin; c.aymin = ph.aymin;
  c.coe_4x = (double)ph.bit4 / 4.0 / s.dx; c.coe_4y = (double)ph.bit4 / 4.0 / s.dy;
  c.coe_4z = (double)ph.bit4 / 4.0 / s.dz;
  c.coe_7 = ph.bit7 * s.fc / 4.0; c.coe_9 = ph.bit9 * s.fc_y / 4.0; c.coe_10 = ph.bit10;
  return c;
}

// PART selects the terms of a component: bit 1 advection and the Coriolis
// and crossflow terms (the adjoints of u, v, w and d/d flow through fbx,
// fby), bit 2 or 4 diffusion, buoyancy and sponge, whose adjoints go to um,
// vm, wm (accumulators of type AM, with d/d flow through the sponge's
// background under bit 2) and to ss, ssv, rhop (type AS); a caller that
// takes one group passes accumulators of the other that discard their
// values, and the terms that only feed those drop out
template <typename T, int PART = 7, class R, class A, class AM, class AS>
HC_HD inline void adj_forces_x(const Layout &L, int i, int j, int k, const StepArgs &s, const Phys &ph,
                               const FlowSens &fs, const R &u, const R &v, const R &w, const R &um, const R &vm,
                               const R &wm, const R &ss, const R &ssv, const double *azzmin, const double *alphau,
                               T lam, const A &au, const A &av, const A &aw, const AM &aum, const AM &avm,
                               const AM &awm, const AS &ass, const AS &assv, double &gflow) {
  (void)ssv; (void)azzmin;
  const AdjConsts c = adj_consts(s, ph);
  if constexpr ((PART & 6) != 0) {
    const T axu_p = ss(i + 1, j, k) + (T)c.axmin, axu_c = ss(i, j, k) + (T)c.axmin;
    const T ayu_c = (T).25 * (ss(i, j, k) + ss(i + 1, j, k) + ss(i, j + 1, k) + ss(i + 1, j + 1, k)) + (T)c.aymin;
    const T ayu_m = (T).25 * (ss(i, j - 1, k) + ss(i + 1, j - 1, k) + ss(i, j, k) + ss(i + 1, j, k)) + (T)c.aymin;
    const T azu_c = azu_at<T>(L, ssv, azzmin, s, i, j, k), azu_m = azu_at<T>(L, ssv, azzmin, s, i, j, k - 1);
    const T umc = um(i, j, k);
    T lumc = 0;
    // 2 bit5 (axu_p (um(i+1)-umc) - axu_c (umc-um(i-1))) / dx^2
    const T g1 = lam * (T)(2.0 * c.bit5 * c.idx * c.idx);
    ass(i + 1, j, k, g1 * (um(i + 1, j, k) - umc));
    ass(i, j, k, -g1 * (umc - um(i - 1, j, k)));
    aum(i + 1, j, k, g1 * axu_p); aum(i - 1, j, k, g1 * axu_c);
    lumc -= g1 * (axu_p + axu_c);
    // bit5 (ayu_c (um(j+1)-umc) - ayu_m (umc-um(j-1))) / dy^2
    const T g2 = lam * (T)(c.bit5 * c.idy * c.idy);
    T lay_c = g2 * (um(i, j + 1, k) - umc), lay_m = -g2 * (umc - um(i, j - 1, k));
    aum(i, j + 1, k, g2 * ayu_c); aum(i, j - 1, k, g2 * ayu_m);
    lumc -= g2 * (ayu_c + ayu_m);
    // bit5 (ayu_c (vm(i+1,j)-vm(i,j)) - ayu_m (vm(i+1,j-1)-vm(i,j-1))) / (dx dy)
    const T g3 = lam * (T)(c.bit5 * c.idx * c.idy);
    lay_c += g3 * (vm(i + 1, j, k) - vm(i, j, k));
    lay_m -= g3 * (vm(i + 1, j - 1, k) - vm(i, j - 1, k));
    avm(i + 1, j, k, g3 * ayu_c); avm(i, j, k, -g3 * ayu_c);
    avm(i + 1, j - 1, k, -g3 * ayu_m); avm(i, j - 1, k, g3 * ayu_m);
    {
      const T qc = (T).25 * lay_c, qm = (T).25 * lay_m;
      ass(i, j, k, qc + qm); ass(i + 1, j, k, qc + qm);
      ass(i, j + 1, k, qc); ass(i + 1, j + 1, k, qc);
      ass(i, j - 1, k, qm); ass(i + 1, j - 1, k, qm);
    }
    // bit2 (azu_c (um(k+1)-umc) - azu_m (umc-um(k-1))) / dz^2
    const T g4 = lam * (T)(c.bit2 * c.idz * c.idz);
    T laz_c = g4 * (um(i, j, k + 1) - umc), laz_m = -g4 * (umc - um(i, j, k - 1));
    aum(i, j, k + 1, g4 * azu_c); aum(i, j, k - 1, g4 * azu_m);
    lumc -= g4 * (azu_c + azu_m);
    // bit2 (azu_c (wm(i+1,k)-wm(i,k)) - azu_m (wm(i+1,k-1)-wm(i,k-1))) / (dx dz)
    const T g5 = lam * (T)(c.bit2 * c.idx * c.idz);
    laz_c += g5 * (wm(i + 1, j, k) - wm(i, j, k));
    laz_m -= g5 * (wm(i + 1, j, k - 1) - wm(i, j, k - 1));
    awm(i + 1, j, k, g5 * azu_c); awm(i, j, k, -g5 * azu_c);
    awm(i + 1, j, k - 1, -g5 * azu_m); awm(i, j, k - 1, g5 * azu_m);
    adj_azu(L, i, j, k, laz_c, assv);
    adj_azu(L, i, j, k - 1, laz_m, assv);
    // sponge -0.5 alphau (um - ubkg_m)
    const double au_ = alphau[(i + 1) + (L.nx + 3) * j];
    lumc -= (T)(0.5 * au_) * lam;
    if constexpr ((PART & 2) != 0) gflow += 0.5 * au_ * (double)lam * fs.dubkg_m;
    aum(i, j, k, lumc);
  }
  if constexpr ((PART & 1) != 0) {
    // advection
    const T uc = u(i, j, k);
    const T a1 = u(i + 1, j, k) + uc, a2 = uc + u(i - 1, j, k);
    const T la1 = -(T)c.coe_4x * (T)2 * a1 * lam, la2 = (T)c.coe_4x * (T)2 * a2 * lam;
    T luc = la1 + la2;
    au(i + 1, j, k, la1); au(i - 1, j, k, la2);
    {  // -coe_4z ((w(i,k)+w(i+1,k)) (u(k+1)+uc) - (w(i,k-1)+w(i+1,k-1)) (uc+u(k-1)))
      const T P1 = w(i, j, k) + w(i + 1, j, k), Q1 = u(i, j, k + 1) + uc;
      const T P2 = w(i, j, k - 1) + w(i + 1, j, k - 1), Q2 = uc + u(i, j, k - 1);
      const T z = (T)c.coe_4z * lam, c9 = (T)c.coe_9 * lam;
      aw(i, j, k, -z * Q1 - c9); aw(i + 1, j, k, -z * Q1 - c9);
      aw(i, j, k - 1, z * Q2 - c9); aw(i + 1, j, k - 1, z * Q2 - c9);
      au(i, j, k + 1, -z * P1); au(i, j, k - 1, z * P2);
      luc += -z * P1 + z * P2;
    }
    {  // -coe_4y ((v(i,j)+v(i+1,j)) (u(j+1)+uc) - (v(i,j-1)+v(i+1,j-1)) (uc+u(j-1))) + coe_7 (...)
      const T P1 = v(i, j, k) + v(i + 1, j, k), Q1 = u(i, j + 1, k) + uc;
      const T P2 = v(i, j - 1, k) + v(i + 1, j - 1, k), Q2 = uc + u(i, j - 1, k);
      const T y = (T)c.coe_4y * lam, c7 = (T)c.coe_7 * lam;
      av(i, j, k, -y * Q1 + c7); av(i + 1, j, k, -y * Q1 + c7);
      av(i, j - 1, k, y * Q2 + c7); av(i + 1, j - 1, k, y * Q2 + c7);
      au(i, j + 1, k, -y * P1); au(i, j - 1, k, y * P2);
      luc += -y * P1 + y * P2;
    }
    au(i, j, k, luc);
    gflow += c.coe_10 * (double)lam * fs.dfbx;
  }
}

template <typename T, int PART = 7, class R, class A, class AM, class AS>
HC_HD inline void adj_forces_y(const Layout &L, int i, int j, int k, const StepArgs &s, const Phys &ph,
                               const FlowSens &fs, const R &u, const R &v, const R &w, const R &um, const R &vm,
                               const R &wm, const R &ss, const R &ssv, const double *azzmin, const double *alphav,
                               T lam, const A &au, const A &av, const A &aw, const AM &aum, const AM &avm,
                               const AM &awm, const AS &ass, const AS &assv, double &gflow) {
  const AdjConsts c = adj_consts(s, ph);
  if constexpr ((PART & 6) != 0) {
    const T axv_c = (T).25 * (ss(i, j, k) + ss(i + 1, j, k) + ss(i, j + 1, k) + ss(i + 1, j + 1, k)) + (T)c.axmin;
    const T axv_m = (T).25 * (ss(i - 1, j, k) + ss(i, j, k) + ss(i - 1, j + 1, k) + ss(i, j + 1, k)) + (T)c.axmin;
    const T ayv_p = ss(i, j + 1, k) + (T)c.aymin, ayv_c = ss(i, j, k) + (T)c.aymin;
    const T azv_c = azv_at<T>(L, ssv, azzmin, s, i, j, k), azv_m = azv_at<T>(L, ssv, azzmin, s, i, j, k - 1);
    const T vmc = vm(i, j, k);
    T lvmc = 0;
    // bit5 (axv_c (vm(i+1)-vmc) - axv_m (vmc-vm(i-1))) / dx^2
    const T g1 = lam * (T)(c.bit5 * c.idx * c.idx);
    T lax_c = g1 * (vm(i + 1, j, k) - vmc), lax_m = -g1 * (vmc - vm(i - 1, j, k));
    avm(i + 1, j, k, g1 * axv_c); avm(i - 1, j, k, g1 * axv_m);
    lvmc -= g1 * (axv_c + axv_m);
    // bit5 (axv_c (um(i,j+1)-um(i,j)) - axv_m (um(i-1,j+1)-um(i-1,j))) / (dx dy)
    const T g2 = lam * (T)(c.bit5 * c.idx * c.idy);
    lax_c += g2 * (um(i, j + 1, k) - um(i, j, k));
    lax_m -= g2 * (um(i - 1, j + 1, k) - um(i - 1, j, k));
    aum(i, j + 1, k, g2 * axv_c); aum(i, j, k, -g2 * axv_c);
    aum(i - 1, j + 1, k, -g2 * axv_m); aum(i - 1, j, k, g2 * axv_m);
    {
      const T qc = (T).25 * lax_c, qm = (T).25 * lax_m;
      ass(i, j, k, qc + qm); ass(i, j + 1, k, qc + qm);
      ass(i + 1, j, k, qc); ass(i + 1, j + 1, k, qc);
      ass(i - 1, j, k, qm); ass(i - 1, j + 1, k, qm);
    }
    // 2 bit5 (ayv_p (vm(j+1)-vmc) - ayv_c (vmc-vm(j-1))) / dy^2
    const T g3 = lam * (T)(2.0 * c.bit5 * c.idy * c.idy);
    ass(i, j + 1, k, g3 * (vm(i, j + 1, k) - vmc));
    ass(i, j, k, -g3 * (vmc - vm(i, j - 1, k)));
    avm(i, j + 1, k, g3 * ayv_p); avm(i, j - 1, k, g3 * ayv_c);
    lvmc -= g3 * (ayv_p + ayv_c);
    // bit2 (azv_c (vm(k+1)-vmc) - azv_m (vmc-vm(k-1))) / dz^2
    const T g4 = lam * (T)(c.bit2 * c.idz * c.idz);
    T laz_c = g4 * (vm(i, j, k + 1) - vmc), laz_m = -g4 * (vmc - vm(i, j, k - 1));
    avm(i, j, k + 1, g4 * azv_c); avm(i, j, k - 1, g4 * azv_m);
    lvmc -= g4 * (azv_c + azv_m);
    // bit2 (azv_c (wm(j+1,k)-wm(j,k)) - azv_m (wm(j+1,k-1)-wm(j,k-1))) / (dy dz)
    const T g5 = lam * (T)(c.bit2 * c.idy * c.idz);
    laz_c += g5 * (wm(i, j + 1, k) - wm(i, j, k));
    laz_m -= g5 * (wm(i, j + 1, k - 1) - wm(i, j, k - 1));
    awm(i, j + 1, k, g5 * azv_c); awm(i, j, k, -g5 * azv_c);
    awm(i, j + 1, k - 1, -g5 * azv_m); awm(i, j, k - 1, g5 * azv_m);
    adj_azv(L, i, j, k, laz_c, assv);
    adj_azv(L, i, j, k - 1, laz_m, assv);
    // sponge -0.5 alphav (vm - vbkg_m)
    const double av_ = alphav[i + (L.nx + 2) * (j + 1)];
    lvmc -= (T)(0.5 * av_) * lam;
    if constexpr ((PART & 2) != 0) gflow += 0.5 * av_ * (double)lam * fs.dvbkg_m;
    avm(i, j, k, lvmc);
  }
  if constexpr ((PART & 1) != 0) {
    // advection
    const T vc = v(i, j, k);
    const T b1 = v(i, j + 1, k) + vc, b2 = vc + v(i, j - 1, k);
    const T lb1 = -(T)c.coe_4y * (T)2 * b1 * lam, lb2 = (T)c.coe_4y * (T)2 * b2 * lam;
    T lvc = lb1 + lb2;
    av(i, j + 1, k, lb1); av(i, j - 1, k, lb2);
    {  // -coe_4z ((v(k+1)+vc) (w(j+1,k)+w(j,k)) - (vc+v(k-1)) (w(j+1,k-1)+w(j,k-1)))
      const T P1 = v(i, j, k + 1) + vc, Q1 = w(i, j + 1, k) + w(i, j, k);
      const T P2 = vc + v(i, j, k - 1), Q2 = w(i, j + 1, k - 1) + w(i, j, k - 1);
      const T z = (T)c.coe_4z * lam;
      av(i, j, k + 1, -z * Q1); av(i, j, k - 1, z * Q2);
      lvc += -z * Q1 + z * Q2;
      aw(i, j + 1, k, -z * P1); aw(i, j, k, -z * P1);
      aw(i, j + 1, k - 1, z * P2); aw(i, j, k - 1, z * P2);
    }
    {  // -coe_4x ((v(i+1)+vc) (u(i,j)+u(i,j+1)) - (vc+v(i-1)) (u(i-1,j)+u(i-1,j+1))) - coe_7 (u ...)
      const T P1 = v(i + 1, j, k) + vc, Q1 = u(i, j, k) + u(i, j + 1, k);
      const T P2 = vc + v(i - 1, j, k), Q2 = u(i - 1, j, k) + u(i - 1, j + 1, k);
      const T x = (T)c.coe_4x * lam, c7 = (T)c.coe_7 * lam;
      av(i + 1, j, k, -x * Q1); av(i - 1, j, k, x * Q2);
      lvc += -x * Q1 + x * Q2;
      au(i, j, k, -x * P1 - c7); au(i, j + 1, k, -x * P1 - c7);
      au(i - 1, j, k, x * P2 - c7); au(i - 1, j + 1, k, x * P2 - c7);
    }
    av(i, j, k, lvc);
    gflow += c.coe_10 * (double)lam * fs.dfby;
  }
}

template <typename T, int PART = 7, class R, class RR, class A, class AM, class AS>
HC_HD inline void adj_forces_z(const Layout &L, int i, int j, int k, const StepArgs &s, const Phys &ph,
                               const R &u, const R &v, const R &w, const R &um, const R &vm, const R &wm,
                               const R &ss, const R &ssv, const RR &rhop, const double *azzmin,
                               const double *alphae, const double *alphaw, T lam, const A &au, const A &av,
                               const A &aw, const AM &aum, const AM &avm, const AM &awm, const AS &ass,
                               const AS &assv, const AS &arhop) {
  (void)rhop;
  const AdjConsts c = adj_consts(s, ph);
  if constexpr ((PART & 6) != 0) {
    const T ahp = ss(i, j, k + 1) + (T)c.axmin, ahm = ss(i, j, k) + (T)c.axmin;
    const T zk = (T)azzmin[k];
    const T s0 = ssv(i, j, k + 1), s1 = ssv(i, j, k);
    const T avxp = (T)0.25 * (s0 + s1 + ssv(i + 1, j, k + 1) + ssv(i + 1, j, k)) + zk;
    const T avxm = (T)0.25 * (s0 + s1 + ssv(i - 1, j, k + 1) + ssv(i - 1, j, k)) + zk;
    const T avyp = (T)0.25 * (s0 + s1 + ssv(i, j + 1, k + 1) + ssv(i, j + 1, k)) + zk;
    const T avym = (T)0.25 * (s0 + s1 + ssv(i, j - 1, k + 1) + ssv(i, j - 1, k)) + zk;
    const T wmc = wm(i, j, k);
    T lwmc = 0;
    // 2 bit5 (ahp (wm(k+1)-wmc) - ahm (wmc-wm(k-1))) / dz^2
    const T g1 = lam * (T)(2.0 * c.bit5 * c.idz * c.idz);
    ass(i, j, k + 1, g1 * (wm(i, j, k + 1) - wmc));
    ass(i, j, k, -g1 * (wmc - wm(i, j, k - 1)));
    awm(i, j, k + 1, g1 * ahp); awm(i, j, k - 1, g1 * ahm);
    lwmc -= g1 * (ahp + ahm);
    // bit2 (avxp (wm(i+1)-wmc) - avxm (wmc-wm(i-1))) / dx^2
    const T g2 = lam * (T)(c.bit2 * c.idx * c.idx);
    T lxp = g2 * (wm(i + 1, j, k) - wmc), lxm = -g2 * (wmc - wm(i - 1, j, k));
    awm(i + 1, j, k, g2 * avxp); awm(i - 1, j, k, g2 * avxm);
    lwmc -= g2 * (avxp + avxm);
    // bit2 (avxp (um(i,k)-um(i,k+1)) - avxm (um(i-1,k)-um(i-1,k+1))) / (dz dx)
    const T g3 = lam * (T)(c.bit2 * c.idz * c.idx);
    lxp += g3 * (um(i, j, k) - um(i, j, k + 1));
    lxm -= g3 * (um(i - 1, j, k) - um(i - 1, j, k + 1));
    aum(i, j, k, g3 * avxp); aum(i, j, k + 1, -g3 * avxp);
    aum(i - 1, j, k, -g3 * avxm); aum(i - 1, j, k + 1, g3 * avxm);
    // bit2 (avyp (wm(j+1)-wmc) - avym (wmc-wm(j-1))) / dy^2
    const T g4 = lam * (T)(c.bit2 * c.idy * c.idy);
    T lyp = g4 * (wm(i, j + 1, k) - wmc), lym = -g4 * (wmc - wm(i, j - 1, k));
    awm(i, j + 1, k, g4 * avyp); awm(i, j - 1, k, g4 * avym);
    lwmc -= g4 * (avyp + avym);
    // bit2 (avyp (vm(j,k)-vm(j,k+1)) - avym (vm(j-1,k)-vm(j-1,k+1))) / (dz dy)
    const T g5 = lam * (T)(c.bit2 * c.idz * c.idy);
    lyp += g5 * (vm(i, j, k) - vm(i, j, k + 1));
    lym -= g5 * (vm(i, j - 1, k) - vm(i, j - 1, k + 1));
    avm(i, j, k, g5 * avyp); avm(i, j, k + 1, -g5 * avyp);
    avm(i, j - 1, k, -g5 * avym); avm(i, j - 1, k + 1, g5 * avym);
    {
      // the four averaged vertical viscosities share s0, s1
      const T qxp = (T).25 * lxp, qxm = (T).25 * lxm, qyp = (T).25 * lyp, qym = (T).25 * lym;
      const T q01 = qxp + qxm + qyp + qym;
      assv(i, j, k + 1, q01); assv(i, j, k, q01);
      assv(i + 1, j, k + 1, qxp); assv(i + 1, j, k, qxp);
      assv(i - 1, j, k + 1, qxm); assv(i - 1, j, k, qxm);
      assv(i, j + 1, k + 1, qyp); assv(i, j + 1, k, qyp);
      assv(i, j - 1, k + 1, qym); assv(i, j - 1, k, qym);
    }
    // buoyancy bit3 g (rhop(k) + rhop(k+1)) / rho0 / 2
    {
      const T b = (T)((double)ph.bit3 * ph.g / ph.rho0 / 2.0) * lam;
      arhop(i, j, k, b); arhop(i, j, k + 1, b);
    }
    // sponge -0.5 alphaee wm
    const double ae = fmax(alphae[i + (L.nx + 2) * j], (alphaw[k + 1] + alphaw[k]) / 2.0);
    lwmc -= (T)(0.5 * ae) * lam;
    awm(i, j, k, lwmc);
  }
  if constexpr ((PART & 1) != 0) {
    // advection
    const T wc = w(i, j, k);
    const T c1 = w(i, j, k + 1) + wc, c2 = wc + w(i, j, k - 1);
    const T lc1 = -(T)c.coe_4z * (T)2 * c1 * lam, lc2 = (T)c.coe_4z * (T)2 * c2 * lam;
    T lwc = lc1 + lc2;
    aw(i, j, k + 1, lc1); aw(i, j, k - 1, lc2);
    {  // -coe_4x ((w(i+1)+wc) (u(i,k)+u(i,k+1)) - (wc+w(i-1)) (u(i-1,k)+u(i-1,k+1))) + coe_9 (u ...)
      const T P1 = w(i + 1, j, k) + wc, Q1 = u(i, j, k) + u(i, j, k + 1);
      const T P2 = wc + w(i - 1, j, k), Q2 = u(i - 1, j, k) + u(i - 1, j, k + 1);
      const T x = (T)c.coe_4x * lam, c9 = (T)c.coe_9 * lam;
      aw(i + 1, j, k, -x * Q1); aw(i - 1, j, k, x * Q2);
      lwc += -x * Q1 + x * Q2;
      au(i, j, k, -x * P1 + c9); au(i, j, k + 1, -x * P1 + c9);
      au(i - 1, j, k, x * P2 + c9); au(i - 1, j, k + 1, x * P2 + c9);
    }
    {  // -coe_4y ((w(j+1)+wc) (v(j,k)+v(j,k+1)) - (wc+w(j-1)) (v(j-1,k)+v(j-1,k+1)))
      const T P1 = w(i, j + 1, k) + wc, Q1 = v(i, j, k) + v(i, j, k + 1);
      const T P2 = wc + w(i, j - 1, k), Q2 = v(i, j - 1, k) + v(i, j - 1, k + 1);
      const T y = (T)c.coe_4y * lam;
      aw(i, j + 1, k, -y * Q1); aw(i, j - 1, k, y * Q2);
      lwc += -y * Q1 + y * Q2;
      av(i, j, k, -y * P1); av(i, j, k + 1, -y * P1);
      av(i, j - 1, k, y * P2); av(i, j - 1, k + 1, y * P2);
    }
    aw(i, j, k, lwc);
  }
}

// ---------------------------------------------------------------------------
// VISCOSITY3D (smagorinsky_split_mixing-revised2014.f:26-122; forward:
// ss_val / ssv_val)

template <typename T, class R, class A>
HC_HD inline void adj_ss_val(const Layout &L, const Phys &ph, const StepArgs &s, const R &um, const R &vm,
                             const R &wm, int i, int j, int k, T lam, const A &aum, const A &avm, const A &awm) {
  (void)L;
  if (lam == (T)0) return;
  const double idx = 1.0 / s.dx, idy = 1.0 / s.dy, idz = 1.0 / s.dz;
  const double ls = sqrt(s.dx * s.dy);
  const T coef = (T)(ph.ismag_h * ((ph.cs_h * ls) * (ph.cs_h * ls)));
  const T s11 = (um(i, j, k) - um(i - 1, j, k)) * (T)idx;
  const T s22 = (vm(i, j, k) - vm(i, j - 1, k)) * (T)idy;
  const T s33 = (wm(i, j, k) - wm(i, j, k - 1)) * (T)idz;
  const T s12 = (T)0.25 *
                ((um(i, j + 1, k) + um(i - 1, j + 1, k) - um(i, j - 1, k) - um(i - 1, j - 1, k)) * (T)idy +
                 (vm(i + 1, j, k) + vm(i + 1, j - 1, k) - vm(i - 1, j, k) - vm(i - 1, j - 1, k)) * (T)idx) /
                (T)2.0;
  const T q = (T)2.0 * (s11 * s11) + (T)2.0 * (s22 * s22) + (T)4.0 * (s12 * s12) + (T)2.0 * (s33 * s33);
  const T lq = lam * coef * (T)0.5 / sqrt(q + (T)1.0e-30);
  const T l11 = (T)4 * s11 * lq * (T)idx, l22 = (T)4 * s22 * lq * (T)idy, l33 = (T)4 * s33 * lq * (T)idz;
  aum(i, j, k, l11); aum(i - 1, j, k, -l11);
  avm(i, j, k, l22); avm(i, j - 1, k, -l22);
  awm(i, j, k, l33); awm(i, j, k - 1, -l33);
  const T la = (T)8 * s12 * lq * (T)0.125;
  const T ly = la * (T)idy, lx = la * (T)idx;
  aum(i, j + 1, k, ly); aum(i - 1, j + 1, k, ly); aum(i, j - 1, k, -ly); aum(i - 1, j - 1, k, -ly);
  avm(i + 1, j, k, lx); avm(i + 1, j - 1, k, lx); avm(i - 1, j, k, -lx); avm(i - 1, j - 1, k, -lx);
}

template <typename T, class R, class RR, class A>
HC_HD inline void adj_ssv_val(const Layout &L, const Phys &ph, const StepArgs &s, const R &um, const R &vm,
                              const R &wm, const RR &rhop, const double *rho_bkg, int i, int j, int k, T lam,
                              const A &aum, const A &avm, const A &awm, const A &arhop) {
  (void)L;
  if (lam == (T)0) return;
  const double idx = 1.0 / s.dx, idy = 1.0 / s.dy, idz = 1.0 / s.dz;
  const T coef = (T)(ph.ismag_v * ((ph.cs_v * s.dz) * (ph.cs_v * s.dz)));
  const T s13 = (T)0.25 *
                ((um(i, j, k - 1) + um(i - 1, j, k - 1) - um(i, j, k + 1) - um(i - 1, j, k + 1)) * (T)idz +
                 (wm(i + 1, j, k) + wm(i + 1, j, k - 1) - wm(i - 1, j, k) - wm(i - 1, j, k - 1)) * (T)idx) /
                (T)2.0;
  const T s23 = (T)0.25 *
                ((vm(i, j, k - 1) + vm(i, j - 1, k - 1) - vm(i, j, k + 1) - vm(i, j - 1, k + 1)) * (T)idz +
                 (wm(i, j + 1, k) + wm(i, j + 1, k - 1) - wm(i, j - 1, k) - wm(i, j - 1, k - 1)) * (T)idy) /
                (T)2.0;
  const double sv = (double)((T)4.0 * (s13 * s13) + (T)4.0 * (s23 * s23));
  const double den = sv + 1.0e-30;
  const double Aq = sqrt(den);
  const double rkm = (double)rhop(i, j, k - 1) + rho_bkg[k - 1];
  const double rkp = (double)rhop(i, j, k + 1) + rho_bkg[k + 1];
  const double G = ph.g / 2.0 * idz / ph.rho0;
  const double ri = -(G * (rkm - rkp) / den);
  const double arg = 1.0 + ph.eps - ri / 3.0;
  const double B = sqrt(fmax(0.0, arg));
  const double l = (double)lam * (double)coef;
  double lsv = l * B * 0.5 / Aq;
  if (arg > 0.0) {
    const double lri = l * Aq * 0.5 / B * (-1.0 / 3.0);
    lsv += lri * (-ri / den);
    arhop(i, j, k - 1, (T)(lri * (-G / den)));
    arhop(i, j, k + 1, (T)(lri * (G / den)));
  }
  const T la = (T)(8.0 * (double)s13 * lsv * 0.125), lb = (T)(8.0 * (double)s23 * lsv * 0.125);
  const T laz = la * (T)idz, lax = la * (T)idx, lbz = lb * (T)idz, lby = lb * (T)idy;
  aum(i, j, k - 1, laz); aum(i - 1, j, k - 1, laz); aum(i, j, k + 1, -laz); aum(i - 1, j, k + 1, -laz);
  awm(i + 1, j, k, lax); awm(i + 1, j, k - 1, lax); awm(i - 1, j, k, -lax); awm(i - 1, j, k - 1, -lax);
  avm(i, j, k - 1, lbz); avm(i, j - 1, k - 1, lbz); avm(i, j, k + 1, -lbz); avm(i, j - 1, k + 1, -lbz);
  awm(i, j + 1, k, lby); awm(i, j + 1, k - 1, lby); awm(i, j - 1, k, -lby); awm(i, j - 1, k - 1, -lby);
}

// one cell of the viscosity box: the adjoints of the ss and ss_v values the
// forward wrote there (visc_at), routed to the interior formula cells they
// were computed from
template <typename T, class R, class RR, class A>
HC_HD inline void adj_visc_at(const Layout &L, const StepArgs &s, const Phys &ph, const R &um, const R &vm,
                              const R &wm, const RR &rhop, const double *rho_bkg, T lss, T lssv, int i, int j,
                              int k, const A &aum, const A &avm, const A &awm, const A &arhop) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const bool gi = (i == 0 || i == nx + 1), gj = (j == 0 || j == ny + 1), gk = (k == 0 || k == nz + 1);
  const int ng = (int)gi + (int)gj + (int)gk;
  if (ng == 0) {
    adj_ss_val<T>(L, ph, s, um, vm, wm, i, j, k, lss, aum, avm, awm);
    adj_ssv_val<T>(L, ph, s, um, vm, wm, rhop, rho_bkg, i, j, k, lssv, aum, avm, awm, arhop);
  } else if (ng == 1 && gi) {
    const T bx = (T)ph.bcondx;
    adj_ss_val<T>(L, ph, s, um, vm, wm, i == 0 ? 1 : nx, j, k, lss * bx, aum, avm, awm);
    adj_ss_val<T>(L, ph, s, um, vm, wm, i == 0 ? nx : 1, j, k, lss * ((T)1 - bx), aum, avm, awm);
  } else if (ng == 1 && gj) {
    const T by = (T)ph.bcondy;
    adj_ss_val<T>(L, ph, s, um, vm, wm, i, j == 0 ? 1 : ny, k, lss * by, aum, avm, awm);
    adj_ss_val<T>(L, ph, s, um, vm, wm, i, j == 0 ? ny : 1, k, lss * ((T)1 - by), aum, avm, awm);
  } else if (ng == 1 && gk) {
    adj_ssv_val<T>(L, ph, s, um, vm, wm, rhop, rho_bkg, i, j, k == 0 ? 1 : nz, lssv, aum, avm, awm, arhop);
  }
}

// the cells whose ss / ss_v values VISCOSITY3D writes from a formula (the
// others hold constants)
HC_HD inline bool visc_ss_written(const Layout &L, int i, int j, int k) {
  const bool gi = (i == 0 || i == L.nx + 1), gj = (j == 0 || j == L.ny + 1), gk = (k == 0 || k == L.nz + 1);
  return ((int)gi + (int)gj + (int)gk) == 0 || ((int)gi + (int)gj + (int)gk == 1 && !gk);
}
HC_HD inline bool visc_ssv_written(const Layout &L, int i, int j, int k) {
  const bool gi = (i == 0 || i == L.nx + 1), gj = (j == 0 || j == L.ny + 1), gk = (k == 0 || k == L.nz + 1);
  return !gi && !gj;
}

// ---------------------------------------------------------------------------
// DENSITY (potential_density80.f:20-67, RHO_RHO :94-101): d rho / dT, dS
HC_HD inline void drho(const Phys &ph, double s, double t, double &drdt, double &drds) {
  if (ph.eos == 1) {
    drdt = -ph.eos_rho_ref * ph.eos_alpha;
    drds = ph.eos_rho_ref * ph.eos_beta;
    return;
  }
  const double t2 = t * t, t3 = t2 * t, t4 = t2 * t2;
  const double dr0 = 6.793952e-2 - 2.0 * 9.095290e-3 * t + 3.0 * 1.001685e-4 * t2 - 4.0 * 1.120083e-6 * t3 +
                     5.0 * 6.536332e-9 * t4;
  const double a = 0.824493 - 4.0899e-3 * t + 7.6438e-5 * t2 - 8.2467e-7 * t3 + 5.3875e-9 * t4;
  const double da = -4.0899e-3 + 2.0 * 7.6438e-5 * t - 3.0 * 8.2467e-7 * t2 + 4.0 * 5.3875e-9 * t3;
  const double b = -5.72466e-3 + 1.0227e-4 * t - 1.6546e-6 * t2;
  const double db = 1.0227e-4 - 2.0 * 1.6546e-6 * t;
  const double rs = s > 0.0 ? sqrt(s) : 0.0;
  drdt = dr0 + s * da + s * rs * db;
  drds = a + 1.5 * rs * b + 2.0 * 4.8314e-4 * s;
}

// ---------------------------------------------------------------------------
// MPDATA (mpdata-hotcross80.f:218-326, 527-608, 660-720)

// reverse of mp_donor: F(cc, nbrs, face Courant numbers) with the limiter's
// zeroed split velocities; outputs the adjoints of the 7 concentrations
// (cc, x-, x+, y-, y+, z-, z+) and the 6 face velocities (u0, u1, v0, v1,
// w0, w1)
template <typename T>
HC_HD inline void adj_mp_donor(T cc, T cxm, T cxp, T cym, T cyp, T czm, T czp, T u0, T u1, T v0, T v1, T w0, T w1,
                               bool lim, int fc, int fxm, int fxp, int fym, int fyp, int fzm, int fzp, T lam,
                               T lc[7], T lv[6]) {
  // split velocities p = .5 (x + |x|), m = .5 (x - |x|) with their masks
  T mp[6] = {1, 1, 1, 1, 1, 1}, mm[6] = {1, 1, 1, 1, 1, 1};
  if (lim) {
    if (fxm == 1 || fc == 2) mp[0] = 0;
    if (fc == 1 || fxm == 2) mm[0] = 0;
    if (fc == 1 || fxp == 2) mp[1] = 0;
    if (fxp == 1 || fc == 2) mm[1] = 0;
    if (fym == 1 || fc == 2) mp[2] = 0;
    if (fc == 1 || fym == 2) mm[2] = 0;
    if (fc == 1 || fyp == 2) mp[3] = 0;
    if (fyp == 1 || fc == 2) mm[3] = 0;
    if (fc == 1 || fzm == 2) mp[4] = 0;
    if (fzm == 1 || fc == 2) mm[4] = 0;
    if (fzp == 1 || fc == 2) mp[5] = 0;
    if (fc == 1 || fzp == 2) mm[5] = 0;
  }
  const T vel[6] = {u0, u1, v0, v1, w0, w1};
  T p[6], m[6], dp[6], dm[6];
  for (int q = 0; q < 6; ++q) {
    const T a = vel[q], sg = adj_sign(a);
    p[q] = (T)0.5 * (a + mp_fabs(a)) * mp[q];
    m[q] = (T)0.5 * (a - mp_fabs(a)) * mm[q];
    dp[q] = (T)0.5 * ((T)1 + sg) * mp[q];
    dm[q] = (T)0.5 * ((T)1 - sg) * mm[q];
  }
  // F = (p0 cxm + m0 cc) + (p2 cym + m2 cc) + (m5 cc + p5 czp)
  //   - ((p1 cc + m1 cxp) + (p3 cc + m3 cyp) + (m4 czm + p4 cc))
  lc[0] = lam * (m[0] + m[2] + m[5] - p[1] - p[3] - p[4]);
  lc[1] = lam * p[0];
  lc[2] = -lam * m[1];
  lc[3] = lam * p[2];
  lc[4] = -lam * m[3];
  lc[5] = -lam * m[4];
  lc[6] = lam * p[5];
  lv[0] = lam * (dp[0] * cxm + dm[0] * cc);
  lv[1] = -lam * (dp[1] * cc + dm[1] * cxp);
  lv[2] = lam * (dp[2] * cym + dm[2] * cc);
  lv[3] = -lam * (dp[3] * cc + dm[3] * cyp);
  lv[4] = -lam * (dm[4] * czm + dp[4] * cc);
  lv[5] = lam * (dm[5] * cc + dp[5] * czp);
}

// 1 / x for the antidiffusive ratios' reverse (their values and
// derivatives to an ulp or two, which the adjoint's coefficients need no more
// precisely): the hardware reciprocal refined by two Newton steps in FP64,
// the hardware reciprocal in FP32; one reciprocal per ratio replaces the two
// divisions of r and of its derivative
template <typename T>
HC_HD inline T adj_recip(T x);
template <>
HC_HD inline double adj_recip<double>(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;
#endif
}
template <>
HC_HD inline float adj_recip<float>(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// d r / d ca and d r / d cb of r = (ca - cb) / (ca + cb + eps), with
// rs = 1 / (ca + cb + eps)
template <typename T>
HC_HD inline void adj_ratio(T ca, T cb, T rs, T lr, T &lca, T &lcb) {
  const T q = lr * rs * rs;
  lca = q * ((T)2 * cb + (T)1.0e-30);
  lcb = -q * ((T)2 * ca + (T)1.0e-30);
}

// reverse of mp_anti_u at face (i,j,k): C(a,b,c) reads cc(i+a, j+b, k+c),
// accumulators likewise relative to the face
template <typename T, class FC, class FU, class FV, class FW, class AC, class AU, class AV, class AW>
HC_HD inline void adj_anti_u(const FC &C, const FU &U, const FV &V, const FW &W, T lam, const AC &ac, const AU &aU,
                             const AV &aV, const AW &aW) {
  if (lam == (T)0) return;
  const T Uc = U(0, 0, 0);
  const T ca0 = C(1, 0, 0), cb0 = C(0, 0, 0);
  const T r0 = adj_recip<T>(ca0 + cb0 + (T)1.0e-30), A0 = (ca0 - cb0) * r0;
  const T ca1 = C(1, 1, 0) + C(0, 1, 0), cb1 = C(1, -1, 0) + C(0, -1, 0);
  const T r1 = adj_recip<T>(ca1 + cb1 + (T)1.0e-30), A1 = (ca1 - cb1) * r1;
  const T ca2 = C(1, 0, -1) + C(0, 0, -1), cb2 = C(1, 0, 1) + C(0, 0, 1);
  const T r2 = adj_recip<T>(ca2 + cb2 + (T)1.0e-30), A2 = (ca2 - cb2) * r2;
  const T Vb = (T)0.25 * (V(1, 0, 0) + V(0, 0, 0) + V(1, -1, 0) + V(0, -1, 0));
  const T Wb = (T)0.25 * (W(1, 0, 0) + W(0, 0, 0) + W(1, 0, -1) + W(0, 0, -1));
  aU(0, 0, 0, lam * ((adj_sign(Uc) - (T)2 * Uc) * A0 - (T)0.5 * A1 * Vb - (T)0.5 * A2 * Wb));
  const T lvb = (T)0.25 * (-(T)0.5 * lam * Uc * A1), lwb = (T)0.25 * (-(T)0.5 * lam * Uc * A2);
  aV(1, 0, 0, lvb); aV(0, 0, 0, lvb); aV(1, -1, 0, lvb); aV(0, -1, 0, lvb);
  aW(1, 0, 0, lwb); aW(0, 0, 0, lwb); aW(1, 0, -1, lwb); aW(0, 0, -1, lwb);
  T la, lb;
  adj_ratio<T>(ca0, cb0, r0, lam * (mp_fabs(Uc) - Uc * Uc), la, lb);
  ac(1, 0, 0, la); ac(0, 0, 0, lb);
  adj_ratio<T>(ca1, cb1, r1, -(T)0.5 * lam * Uc * Vb, la, lb);
  ac(1, 1, 0, la); ac(0, 1, 0, la); ac(1, -1, 0, lb); ac(0, -1, 0, lb);
  adj_ratio<T>(ca2, cb2, r2, -(T)0.5 * lam * Uc * Wb, la, lb);
  ac(1, 0, -1, la); ac(0, 0, -1, la); ac(1, 0, 1, lb); ac(0, 0, 1, lb);
}

// reverse of mp_anti_v at face (i,j,k)
template <typename T, class FC, class FU, class FV, class FW, class AC, class AU, class AV, class AW>
HC_HD inline void adj_anti_v(const FC &C, const FU &U, const FV &V, const FW &W, T lam, const AC &ac, const AU &aU,
                             const AV &aV, const AW &aW) {
  if (lam == (T)0) return;
  const T Vc = V(0, 0, 0);
  const T ca0 = C(0, 1, 0), cb0 = C(0, 0, 0);
  const T r0 = adj_recip<T>(ca0 + cb0 + (T)1.0e-30), A0 = (ca0 - cb0) * r0;
  const T ca1 = C(1, 1, 0) + C(1, 0, 0), cb1 = C(-1, 1, 0) + C(-1, 0, 0);
  const T r1 = adj_recip<T>(ca1 + cb1 + (T)1.0e-30), A1 = (ca1 - cb1) * r1;
  const T ca2 = C(0, 1, -1) + C(0, 0, -1), cb2 = C(0, 1, 1) + C(0, 0, 1);
  const T r2 = adj_recip<T>(ca2 + cb2 + (T)1.0e-30), A2 = (ca2 - cb2) * r2;
  const T Ub = (T)0.25 * (U(0, 1, 0) + U(0, 0, 0) + U(-1, 1, 0) + U(-1, 0, 0));
  const T Wb = (T)0.25 * (W(0, 0, 0) + W(0, 1, 0) + W(0, 0, -1) + W(0, 1, -1));
  aV(0, 0, 0, lam * ((adj_sign(Vc) - (T)2 * Vc) * A0 - (T)0.5 * A1 * Ub - (T)0.5 * A2 * Wb));
  const T lub = (T)0.25 * (-(T)0.5 * lam * Vc * A1), lwb = (T)0.25 * (-(T)0.5 * lam * Vc * A2);
  aU(0, 1, 0, lub); aU(0, 0, 0, lub); aU(-1, 1, 0, lub); aU(-1, 0, 0, lub);
  aW(0, 0, 0, lwb); aW(0, 1, 0, lwb); aW(0, 0, -1, lwb); aW(0, 1, -1, lwb);
  T la, lb;
  adj_ratio<T>(ca0, cb0, r0, lam * (mp_fabs(Vc) - Vc * Vc), la, lb);
  ac(0, 1, 0, la); ac(0, 0, 0, lb);
  adj_ratio<T>(ca1, cb1, r1, -(T)0.5 * lam * Vc * Ub, la, lb);
  ac(1, 1, 0, la); ac(1, 0, 0, la); ac(-1, 1, 0, lb); ac(-1, 0, 0, lb);
  adj_ratio<T>(ca2, cb2, r2, -(T)0.5 * lam * Vc * Wb, la, lb);
  ac(0, 1, -1, la); ac(0, 0, -1, la); ac(0, 1, 1, lb); ac(0, 0, 1, lb);
}

// reverse of mp_anti_w at face (i,j,k) (between planes k and k+1)
template <typename T, class FC, class FU, class FV, class FW, class AC, class AU, class AV, class AW>
HC_HD inline void adj_anti_w(const FC &C, const FU &U, const FV &V, const FW &W, T lam, const AC &ac, const AU &aU,
                             const AV &aV, const AW &aW) {
  if (lam == (T)0) return;
  const T Wc = W(0, 0, 0);
  const T ca0 = C(0, 0, 0), cb0 = C(0, 0, 1);
  const T r0 = adj_recip<T>(ca0 + cb0 + (T)1.0e-30), A0 = (ca0 - cb0) * r0;
  const T ca1 = C(1, 0, 0) + C(1, 0, 1), cb1 = C(-1, 0, 0) + C(-1, 0, 1);
  const T r1 = adj_recip<T>(ca1 + cb1 + (T)1.0e-30), A1 = (ca1 - cb1) * r1;
  const T ca2 = C(0, 1, 1) + C(0, 1, 0), cb2 = C(0, -1, 1) + C(0, -1, 0);
  const T r2 = adj_recip<T>(ca2 + cb2 + (T)1.0e-30), A2 = (ca2 - cb2) * r2;
  const T Ub = (T)0.25 * (U(0, 0, 0) + U(0, 0, 1) + U(-1, 0, 1) + U(-1, 0, 0));
  const T Vb = (T)0.25 * (V(0, 0, 0) + V(0, 0, 1) + V(0, -1, 1) + V(0, -1, 0));
  aW(0, 0, 0, lam * ((adj_sign(Wc) - (T)2 * Wc) * A0 - (T)0.5 * A1 * Ub - (T)0.5 * A2 * Vb));
  const T lub = (T)0.25 * (-(T)0.5 * lam * Wc * A1), lvb = (T)0.25 * (-(T)0.5 * lam * Wc * A2);
  aU(0, 0, 0, lub); aU(0, 0, 1, lub); aU(-1, 0, 1, lub); aU(-1, 0, 0, lub);
  aV(0, 0, 0, lvb); aV(0, 0, 1, lvb); aV(0, -1, 1, lvb); aV(0, -1, 0, lvb);
  T la, lb;
  adj_ratio<T>(ca0, cb0, r0, lam * (mp_fabs(Wc) - Wc * Wc), la, lb);
  ac(0, 0, 0, la); ac(0, 0, 1, lb);
  adj_ratio<T>(ca1, cb1, r1, -(T)0.5 * lam * Wc * Ub, la, lb);
  ac(1, 0, 0, la); ac(1, 0, 1, la); ac(-1, 0, 0, lb); ac(-1, 0, 1, lb);
  adj_ratio<T>(ca2, cb2, r2, -(T)0.5 * lam * Wc * Vb, la, lb);
  ac(0, 1, 1, la); ac(0, 1, 0, la); ac(0, -1, 1, lb); ac(0, -1, 0, lb);
}

// reverse of pass1_at at interior cell (i,j,k): explicit diffusion with the
// diffusivities of MIXING and the unlimited donor-cell step
template <typename T, class R, class A>
HC_HD inline void adj_pass1(const Layout &L, const StepArgs &s, const Phys &ph, const R &c, const R &uc, const R &vc,
                            const R &wc, const R &ss, const R &ssv, const double *azzmin, const double *cbkg, int i,
                            int j, int k, T lam, const A &acc, const A &auc, const A &avc, const A &awc,
                            const A &ass, const A &assv) {
  if (lam == (T)0) return;
  // the divisions by dx, dy, dz and by the Prandtl number through host
  // reciprocals (UDiv: the same correctly rounded quotients as pass1_at's)
  const UDiv<double> ux(s.dx, s.idx), uy(s.dy, s.idy), uz(s.dz, s.idz), prd(ph.prandtl, s.rpr);
  const UDiv<T> tpr = udiv_t<T>(ph.prandtl, s.rpr, s.rprf);
  const T xfac = (T)(ux(ux((double)(1 * ph.bit5k))) * s.tau3 * 1);
  const T yfac = (T)(uy(uy((double)(1 * ph.bit5k))) * s.tau3 * 1);
  const T zfac = (T)(uz(uz((double)ph.bit2k)) * s.tau3 * 1);
  const T b5 = (T)ph.bit5k, b2 = (T)ph.bit2k;
  const T axp = (T)prd(ph.axmin), ayp = (T)prd(ph.aymin);
  const T h5 = tpr(b5 * (T)0.5), h2 = tpr(b2 * (T)0.5);
  const T cp = c(i, j, k);
  const T kx1 = b5 * (tpr((T).5 * (ss(i, j, k) + ss(i + 1, j, k))) + axp);
  const T kx0 = b5 * (tpr((T).5 * (ss(i - 1, j, k) + ss(i, j, k))) + axp);
  const T ky1 = b5 * (tpr((T).5 * (ss(i, j, k) + ss(i, j + 1, k))) + ayp);
  const T ky0 = b5 * (tpr((T).5 * (ss(i, j - 1, k) + ss(i, j, k))) + ayp);
  const T kz1 = (k == L.nz) ? (T)0 : b2 * (tpr((T)0.5 * (ssv(i, j, k) + ssv(i, j, k + 1))) + (T)prd(azzmin[k]));
  const T kz0 = (k == 1) ? (T)0 : b2 * (tpr((T)0.5 * (ssv(i, j, k - 1) + ssv(i, j, k))) + (T)prd(azzmin[k - 1]));
  const T bk = (T)cbkg[k], bkm = (T)cbkg[k - 1], bkp = (T)cbkg[k + 1];
  T lcp = lam;  // cd = cp + ...
  {  // cx = xfac (kx1 (c(i+1)-cp) - kx0 (cp-c(i-1)))
    const T g = xfac * lam;
    const T lk1 = g * (c(i + 1, j, k) - cp), lk0 = -g * (cp - c(i - 1, j, k));
    acc(i + 1, j, k, g * kx1); acc(i - 1, j, k, g * kx0);
    lcp -= g * (kx1 + kx0);
    ass(i, j, k, h5 * (lk1 + lk0)); ass(i + 1, j, k, h5 * lk1); ass(i - 1, j, k, h5 * lk0);
  }
  {
    const T g = yfac * lam;
    const T lk1 = g * (c(i, j + 1, k) - cp), lk0 = -g * (cp - c(i, j - 1, k));
    acc(i, j + 1, k, g * ky1); acc(i, j - 1, k, g * ky0);
    lcp -= g * (ky1 + ky0);
    ass(i, j, k, h5 * (lk1 + lk0)); ass(i, j + 1, k, h5 * lk1); ass(i, j - 1, k, h5 * lk0);
  }
  {  // cz = zfac (kz0 ((c(k-1)+bk) - (bkm+cp)) - kz1 ((cp+bkp) - (bk+c(k+1))))
    const T g = zfac * lam;
    const T lk0 = g * ((c(i, j, k - 1) + bk) - (bkm + cp)), lk1 = -g * ((cp + bkp) - (bk + c(i, j, k + 1)));
    acc(i, j, k - 1, g * kz0); acc(i, j, k + 1, g * kz1);
    lcp -= g * (kz0 + kz1);
    if (k != 1) { assv(i, j, k - 1, h2 * lk0); assv(i, j, k, h2 * lk0); }
    if (k != L.nz) { assv(i, j, k, h2 * lk1); assv(i, j, k + 1, h2 * lk1); }
  }
  T lc[7], lv[6];
  adj_mp_donor<T>(cp, c(i - 1, j, k), c(i + 1, j, k), c(i, j - 1, k), c(i, j + 1, k), c(i, j, k - 1),
                  c(i, j, k + 1), uc(i - 1, j, k), uc(i, j, k), vc(i, j - 1, k), vc(i, j, k), wc(i, j, k - 1),
                  wc(i, j, k), false, 0, 0, 0, 0, 0, 0, 0, lam, lc, lv);
  acc(i, j, k, lcp + lc[0]);
  acc(i - 1, j, k, lc[1]); acc(i + 1, j, k, lc[2]); acc(i, j - 1, k, lc[3]); acc(i, j + 1, k, lc[4]);
  acc(i, j, k - 1, lc[5]); acc(i, j, k + 1, lc[6]);
  auc(i - 1, j, k, lv[0]); auc(i, j, k, lv[1]); avc(i, j - 1, k, lv[2]); avc(i, j, k, lv[3]);
  awc(i, j, k - 1, lv[4]); awc(i, j, k, lv[5]);
}

}  // namespace hc
