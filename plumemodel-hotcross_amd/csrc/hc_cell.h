// hc_cell.h - per-cell formulas of TRACER31 shared by the forward kernels,
// the adjoint kernels and host builds of both (every function here compiles
// for the host and the device).
#pragma once
#include "hc_common.h"
#include "hc_kernels.h"
#include "hc_mpdata.h"

namespace hc {

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

// pass 1: diffusion (mpdata-hotcross80.f:141-174) + upwind step (:176-184)
template <typename T, class V>
HC_HD inline T pass1_at(const Layout &L, const StepArgs &s, const Phys &ph, const V &c, const V &uc,
                             const V &vc, const V &wc, const V &ss, const V &ssv, const double *azzmin,
                             const double *cbkg, int i, int j, int k) {
  const double pr = ph.prandtl;
  const T xfac = (T)((double)(1 * ph.bit5k) / s.dx / s.dx * s.tau3 * 1);
  const T yfac = (T)((double)(1 * ph.bit5k) / s.dy / s.dy * s.tau3 * 1);
  const T zfac = (T)((double)ph.bit2k / s.dz / s.dz * s.tau3 * 1);
  const T b5 = (T)ph.bit5k, b2 = (T)ph.bit2k;
  const T axp = (T)(ph.axmin / pr), ayp = (T)(ph.aymin / pr), tpr = (T)pr;
  T cp = c(i, j, k);
  T kx1 = b5 * (((T).5 * (ss(i, j, k) + ss(i + 1, j, k))) / tpr + axp);
  T kx0 = b5 * (((T).5 * (ss(i - 1, j, k) + ss(i, j, k))) / tpr + axp);
  T ky1 = b5 * (((T).5 * (ss(i, j, k) + ss(i, j + 1, k))) / tpr + ayp);
  T ky0 = b5 * (((T).5 * (ss(i, j - 1, k) + ss(i, j, k))) / tpr + ayp);
  // kzt(k) = 0 at k = 0 and k = nz (mixing80-revised2014.f:49-59)
  T kz1 = (k == L.nz) ? (T)0 : b2 * (((T)0.5 * (ssv(i, j, k) + ssv(i, j, k + 1))) / tpr + (T)(azzmin[k] / pr));
  T kz0 = (k == 1) ? (T)0
                   : b2 * (((T)0.5 * (ssv(i, j, k - 1) + ssv(i, j, k))) / tpr + (T)(azzmin[k - 1] / pr));
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
