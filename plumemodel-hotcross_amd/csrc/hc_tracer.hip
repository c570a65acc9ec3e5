// hc_tracer.hip - TRACER31 / MPDATA / BOUNDARY_CONDITIONS on gfx950.
//
// mpdata-hotcross80.f:16-494.  One tracer call is
//   pass 1: explicit diffusion + donor-cell upwind step      (k_pass1)
//   pass 2/3: antidiffusive velocities (k_antidiff), then the limiter flags
//            and the limited donor-cell step (k_limited)
//   sponge + source on the last pass, then the boundary images (k_bc).
// The diffusivities kxt/kyt/kzt of MIXING (mixing80-revised2014.f:22-59) are
// formed inside k_pass1 from ss/ss_v with the reference's expressions.
//
// The limiter (mpdata-hotcross80.f:660-690) zeroes face entries from both of
// a face's cells; the reference does this by scattering zeros from every cell
// (a benign write-write race under OpenMP).  Here each cell has a flag
// (1: c+dc <= cmin, 2: c+dc >= cmax), recomputed by each cell for itself and
// its six neighbours, and every face reads the flags of its two cells, which
// is race-free and gives the same zero set.
//
// Ghost cells: BOUNDARY_CONDITIONS leaves every ghost equal to an interior
// cell (z: 0<-1, nz+1<-nz; x/y: clamped for bcond=1, periodic for bcond=0),
// so k_bc copies from the mapped interior cell instead of sweeping faces in
// the reference's order.
#include "hc_common.h"
#include "hc_kernels.h"

namespace hc {

#define F3(a, i, j, k) (a)[IX(L, (i), (j), (k))]

template <typename T>
__global__ void __launch_bounds__(256) k_courant(Layout L, Box b, double tau3, double dx, double dy, double dz,
                                                 const T *__restrict__ u, const T *__restrict__ up,
                                                 const T *__restrict__ v, const T *__restrict__ vp,
                                                 const T *__restrict__ w, const T *__restrict__ wp,
                                                 T *__restrict__ uc, T *__restrict__ vc, T *__restrict__ wc) {
  HC_BOX_IJK(b);  // 0..nx+1, 0..ny+1, 0..nz+1
  size_t q = IX(L, i, j, k);
  // worku = 0.5*(up+u) (hotcross80-revised2014.f:303), uc = worku*tau3/dx
  // (mpdata-hotcross80.f:96-102)
  if (i <= L.nx) uc[q] = (T)0.5 * (up[q] + u[q]) * (T)tau3 / (T)dx;
  if (j <= L.ny) vc[q] = (T)0.5 * (vp[q] + v[q]) * (T)tau3 / (T)dy;
  if (k <= L.nz) wc[q] = -((T)0.5 * (wp[q] + w[q])) * (T)tau3 / (T)dz;
}

// donor-cell flux divergence of MPDATA (mpdata-hotcross80.f:590-608 with the
// split velocities of :527-565), optionally with limited face entries
template <typename T>
__device__ inline T donor(const Layout &L, const T *ca, const T *U, const T *V, const T *W, int i, int j, int k,
                          const int *fl7) {
  T u0 = F3(U, i - 1, j, k), u1 = F3(U, i, j, k);
  T v0 = F3(V, i, j - 1, k), v1 = F3(V, i, j, k);
  T w0 = F3(W, i, j, k - 1), w1 = F3(W, i, j, k);
  T upi0 = (T)0.5 * (u0 + fabs(u0)), umi0 = (T)0.5 * (u0 - fabs(u0));
  T upi1 = (T)0.5 * (u1 + fabs(u1)), umi1 = (T)0.5 * (u1 - fabs(u1));
  T vpj0 = (T)0.5 * (v0 + fabs(v0)), vmj0 = (T)0.5 * (v0 - fabs(v0));
  T vpj1 = (T)0.5 * (v1 + fabs(v1)), vmj1 = (T)0.5 * (v1 - fabs(v1));
  T wpk0 = (T)0.5 * (w0 + fabs(w0)), wmk0 = (T)0.5 * (w0 - fabs(w0));
  T wpk1 = (T)0.5 * (w1 + fabs(w1)), wmk1 = (T)0.5 * (w1 - fabs(w1));
  if (fl7) {
    // flags of the cell and its six neighbours: centre, x-, x+, y-, y+, z-, z+
    int fc = fl7[0], fxm = fl7[1], fxp = fl7[2], fym = fl7[3], fyp = fl7[4], fzm = fl7[5], fzp = fl7[6];
    // x face i-1 (cells i-1, i) and face i (cells i, i+1)
    if (fxm == 1 || fc == 2) upi0 = (T)0.0;
    if (fc == 1 || fxm == 2) umi0 = (T)0.0;
    if (fc == 1 || fxp == 2) upi1 = (T)0.0;
    if (fxp == 1 || fc == 2) umi1 = (T)0.0;
    if (fym == 1 || fc == 2) vpj0 = (T)0.0;
    if (fc == 1 || fym == 2) vmj0 = (T)0.0;
    if (fc == 1 || fyp == 2) vpj1 = (T)0.0;
    if (fyp == 1 || fc == 2) vmj1 = (T)0.0;
    // z face k-1 (cells k-1, k): wpk zeroed by cell k <= cmin or cell k-1 >= cmax
    if (fc == 1 || fzm == 2) wpk0 = (T)0.0;
    if (fzm == 1 || fc == 2) wmk0 = (T)0.0;
    if (fzp == 1 || fc == 2) wpk1 = (T)0.0;
    if (fc == 1 || fzp == 2) wmk1 = (T)0.0;
  }
  T cc = F3(ca, i, j, k);
  T F = +(upi0 * F3(ca, i - 1, j, k) + umi0 * cc) + (vpj0 * F3(ca, i, j - 1, k) + vmj0 * cc) +
        (wmk1 * cc + wpk1 * F3(ca, i, j, k + 1)) -
        ((upi1 * cc + umi1 * F3(ca, i + 1, j, k)) + (vpj1 * cc + vmj1 * F3(ca, i, j + 1, k)) +
         (wmk0 * F3(ca, i, j, k - 1) + wpk0 * cc));
  return (T)0 * cc + F / (T)1.0;
}

// pass 1: diffusion (mpdata-hotcross80.f:141-174) + upwind step (:176-184)
template <typename T>
__global__ void __launch_bounds__(256) k_pass1(Layout L, Box b, StepArgs s, Phys ph, const T *__restrict__ c,
                                               const T *__restrict__ uc, const T *__restrict__ vc,
                                               const T *__restrict__ wc, const T *__restrict__ ss,
                                               const T *__restrict__ ssv, const double *__restrict__ azzmin,
                                               const double *__restrict__ cbkg, T *__restrict__ c1) {
  HC_BOX_IJK(b);  // interior
  const double pr = ph.prandtl;
  const T xfac = (T)((double)(1 * ph.bit5k) / s.dx / s.dx * s.tau3 * 1);
  const T yfac = (T)((double)(1 * ph.bit5k) / s.dy / s.dy * s.tau3 * 1);
  const T zfac = (T)((double)ph.bit2k / s.dz / s.dz * s.tau3 * 1);
  const T b5 = (T)ph.bit5k, b2 = (T)ph.bit2k;
  const T axp = (T)(ph.axmin / pr), ayp = (T)(ph.aymin / pr), tpr = (T)pr;
  T cp = F3(c, i, j, k);
  T kx1 = b5 * (((T).5 * (F3(ss, i, j, k) + F3(ss, i + 1, j, k))) / tpr + axp);
  T kx0 = b5 * (((T).5 * (F3(ss, i - 1, j, k) + F3(ss, i, j, k))) / tpr + axp);
  T ky1 = b5 * (((T).5 * (F3(ss, i, j, k) + F3(ss, i, j + 1, k))) / tpr + ayp);
  T ky0 = b5 * (((T).5 * (F3(ss, i, j - 1, k) + F3(ss, i, j, k))) / tpr + ayp);
  // kzt(k) = 0 at k = 0 and k = nz (mixing80-revised2014.f:49-59)
  T kz1 = (k == L.nz) ? (T)0 : b2 * (((T)0.5 * (F3(ssv, i, j, k) + F3(ssv, i, j, k + 1))) / tpr + (T)(azzmin[k] / pr));
  T kz0 = (k == 1) ? (T)0
                   : b2 * (((T)0.5 * (F3(ssv, i, j, k - 1) + F3(ssv, i, j, k))) / tpr + (T)(azzmin[k - 1] / pr));
  T cx = xfac * (kx1 * (F3(c, i + 1, j, k) - cp) - kx0 * (cp - F3(c, i - 1, j, k)));
  T cy = yfac * (ky1 * (F3(c, i, j + 1, k) - cp) - ky0 * (cp - F3(c, i, j - 1, k)));
  T bk = (T)cbkg[k], bkm = (T)cbkg[k - 1], bkp = (T)cbkg[k + 1];
  T cz = zfac * (kz0 * ((F3(c, i, j, k - 1) + bk) - (bkm + cp)) - kz1 * ((cp + bkp) - (bk + F3(c, i, j, k + 1))));
  T cd = cp + (+cx + cy + cz);
  T ce = donor<T>(L, c, uc, vc, wc, i, j, k, nullptr);
  F3(c1, i, j, k) = cd + ce;
}

// antidiffusive velocities, mpdata-hotcross80.f:218-326
template <typename T>
__global__ void __launch_bounds__(256) k_antidiff(Layout L, Box b, const T *__restrict__ cc, const T *__restrict__ uc,
                                                  const T *__restrict__ vc, const T *__restrict__ wc,
                                                  T *__restrict__ ud, T *__restrict__ vd, T *__restrict__ wd) {
  HC_BOX_IJK(b);  // 0..nx, 0..ny, 0..nz
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const T epsa = (T)1.0e-30;
  if (j >= 1 && k >= 1) {  // ud: i=0..nx, j=1..ny, k=1..nz
    T ca = F3(cc, i + 1, j, k), cb = F3(cc, i, j, k);
    T U = F3(uc, i, j, k);
    T r = (fabs(U) - U * U / (T)1.0) * (ca - cb) / (ca + cb + epsa);
    ca = (F3(cc, i + 1, j + 1, k) + F3(cc, i, j + 1, k));
    cb = (F3(cc, i + 1, j - 1, k) + F3(cc, i, j - 1, k));
    r = r - (T)1 * (T)0.5 * U / (T)1.0 * (ca - cb) / (ca + cb + epsa) * (T)0.25 *
                (F3(vc, i + 1, j, k) + F3(vc, i, j, k) + F3(vc, i + 1, j - 1, k) + F3(vc, i, j - 1, k)) * (T)1;
    ca = (F3(cc, i + 1, j, k - 1) + F3(cc, i, j, k - 1));
    cb = (F3(cc, i + 1, j, k + 1) + F3(cc, i, j, k + 1));
    r = r - (T)0.5 * U / (T)1.0 * (ca - cb) / (ca + cb + epsa) * (T)0.25 *
                (F3(wc, i + 1, j, k) + F3(wc, i, j, k) + F3(wc, i + 1, j, k - 1) + F3(wc, i, j, k - 1)) * (T)1;
    F3(ud, i, j, k) = r;
  }
  if (i >= 1 && k >= 1) {  // vd: i=1..nx, j=0..ny, k=1..nz
    T ca = F3(cc, i, j + 1, k), cb = F3(cc, i, j, k);
    T V = F3(vc, i, j, k);
    T r = (fabs(V) - V * V / (T)1.0) * (T)1 * (ca - cb) / (ca + cb + epsa);
    ca = (F3(cc, i + 1, j + 1, k) + F3(cc, i + 1, j, k));
    cb = (F3(cc, i - 1, j + 1, k) + F3(cc, i - 1, j, k));
    r = r - (T)0.5 * V / (T)1.0 * (ca - cb) / (ca + cb + epsa) * (T)0.25 *
                (F3(uc, i, j + 1, k) + F3(uc, i, j, k) + F3(uc, i - 1, j + 1, k) + F3(uc, i - 1, j, k)) * (T)1;
    ca = (F3(cc, i, j + 1, k - 1) + F3(cc, i, j, k - 1));
    cb = (F3(cc, i, j + 1, k + 1) + F3(cc, i, j, k + 1));
    r = r - (T)0.5 * V / (T)1.0 * (ca - cb) / (ca + cb + epsa) * (T)0.25 *
                (F3(wc, i, j, k) + F3(wc, i, j + 1, k) + F3(wc, i, j, k - 1) + F3(wc, i, j + 1, k - 1)) * (T)1;
    F3(vd, i, j, k) = r;
  }
  if (i >= 1 && j >= 1) {  // wd: i=1..nx, j=1..ny, k=0..nz
    T ca = F3(cc, i, j, k), cb = F3(cc, i, j, k + 1);
    T W = F3(wc, i, j, k);
    T r = (fabs(W) - W * W / (T)1.0) * (ca - cb) / (ca + cb + epsa);
    ca = (F3(cc, i + 1, j, k) + F3(cc, i + 1, j, k + 1));
    cb = (F3(cc, i - 1, j, k) + F3(cc, i - 1, j, k + 1));
    r = r - (T)1 * (T)0.5 * W / (T)1.0 * (ca - cb) / (ca + cb + epsa) * (T)0.25 *
                (F3(uc, i, j, k) + F3(uc, i, j, k + 1) + F3(uc, i - 1, j, k + 1) + F3(uc, i - 1, j, k)) * (T)1;
    ca = (F3(cc, i, j + 1, k + 1) + F3(cc, i, j + 1, k));
    cb = (F3(cc, i, j - 1, k + 1) + F3(cc, i, j - 1, k));
    r = r - (T)1 * (T)0.5 * W / (T)1.0 * (ca - cb) / (ca + cb + epsa) * (T)0.25 *
                (F3(vc, i, j, k) + F3(vc, i, j, k + 1) + F3(vc, i, j - 1, k + 1) + F3(vc, i, j - 1, k)) * (T)1;
    F3(wd, i, j, k) = r;
  }
  (void)nx; (void)ny; (void)nz;
}

// limiter flag of one cell from its unlimited corrective step,
// mpdata-hotcross80.f:660-690 (1: c+dc <= cmin, 2: c+dc >= cmax); cells
// outside the interior carry no flag
template <typename T>
__device__ inline int cell_flag(const Layout &L, const T *cc, const T *ud, const T *vd, const T *wd, int i, int j,
                                int k, double cmin, double cmax) {
  if (i < 1 || i > L.nx || j < 1 || j > L.ny || k < 1 || k > L.nz) return 0;
  T x = F3(cc, i, j, k) + donor<T>(L, cc, ud, vd, wd, i, j, k, nullptr);
  if (x <= (T)cmin) return 1;
  if (x >= (T)cmax) return 2;
  return 0;
}

// limited corrective step (mpdata-hotcross80.f:705-720, then c = c + ce).
// The flags of the cell and of its six neighbours are recomputed here from
// the antidiffusive velocities (read through the caches) instead of being
// stored.  On the last pass also the sponge and source of :356-381, with the
// pre-sponge value kept in cpre for the boundary images.
template <typename T>
__global__ void __launch_bounds__(256) k_limited(Layout L, Box b, StepArgs s, Phys ph, VentArgs va, int last,
                                                 double cmin, double cmax, const T *__restrict__ cc,
                                                 const T *__restrict__ ud, const T *__restrict__ vd,
                                                 const T *__restrict__ wd, T *__restrict__ cout,
                                                 T *__restrict__ cpre, const double *__restrict__ cbkg,
                                                 const double *__restrict__ alphae0, const double *__restrict__ alphaw,
                                                 const double *__restrict__ qpatch) {
  HC_BOX_IJK(b);  // interior
  int f7[7];
  f7[0] = cell_flag<T>(L, cc, ud, vd, wd, i, j, k, cmin, cmax);
  f7[1] = cell_flag<T>(L, cc, ud, vd, wd, i - 1, j, k, cmin, cmax);
  f7[2] = cell_flag<T>(L, cc, ud, vd, wd, i + 1, j, k, cmin, cmax);
  f7[3] = cell_flag<T>(L, cc, ud, vd, wd, i, j - 1, k, cmin, cmax);
  f7[4] = cell_flag<T>(L, cc, ud, vd, wd, i, j + 1, k, cmin, cmax);
  f7[5] = cell_flag<T>(L, cc, ud, vd, wd, i, j, k - 1, cmin, cmax);
  f7[6] = cell_flag<T>(L, cc, ud, vd, wd, i, j, k + 1, cmin, cmax);
  T ce = donor<T>(L, cc, ud, vd, wd, i, j, k, f7);
  T cn = F3(cc, i, j, k) + ce;
  if (!last) {
    F3(cout, i, j, k) = cn;
    return;
  }
  F3(cpre, i, j, k) = cn;
  const double tau3 = s.tau3;
  // sponge: c = c - 0.5*alphaee*tau3*(cp - c_bkg(k)), cp = the field at entry
  double ae = fmax(alphae0[i + (L.nx + 2) * j], (alphaw[k + 1] + alphaw[k]) / 2.0);
  T cp = F3(cout, i, j, k);
  T r = cn - (T)1 * (T)0.5 * (T)ae * (T)tau3 * (cp - (T)cbkg[k]);
  double q = 0.0;
  if (k == L.nz && va.map_valid && i >= va.nxs && i <= va.nxe && j >= va.nys && j <= va.nye)
    q = qpatch[(i - va.nxs) + (va.nxe - va.nxs + 1) * (j - va.nys)];
  r = r + (T)tau3 * (T)q / (T)(s.dx * s.dy * s.dz) / (T)1.0 * (T)1;
  F3(cout, i, j, k) = r;
}

// ghost cells of a scalar field from their interior images.  One thread per
// ghost cell: the x faces (with edges), then the y faces, then the z faces.
template <typename T>
__global__ void __launch_bounds__(256) k_bc(Layout L, Phys ph, VentArgs va, T *__restrict__ dst,
                                            const T *__restrict__ src, int vent, double vent_value) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long na = 2L * (ny + 2) * (nz + 2), nb = 2L * nx * (nz + 2), nc = 2L * nx * ny;
  int i, j, k;
  if (t < na) {
    j = (int)(t % (ny + 2)); long r = t / (ny + 2);
    k = (int)(r % (nz + 2)); i = (r / (nz + 2)) ? nx + 1 : 0;
  } else if ((t -= na) < nb) {
    i = 1 + (int)(t % nx); long r = t / nx;
    k = (int)(r % (nz + 2)); j = (r / (nz + 2)) ? ny + 1 : 0;
  } else if ((t -= nb) < nc) {
    i = 1 + (int)(t % nx); long r = t / nx;
    j = 1 + (int)(r % ny); k = (r / ny) ? nz + 1 : 0;
  } else {
    return;
  }
  bool gi = (i == 0 || i == nx + 1), gj = (j == 0 || j == ny + 1);
  int si = i, sj = j, sk = k;
  if (k == 0) sk = 1;
  if (k == nz + 1) sk = nz;
  if (i == 0) si = ph.bcondx ? 1 : nx;
  if (i == nx + 1) si = ph.bcondx ? nx : 1;
  if (j == 0) sj = ph.bcondy ? 1 : ny;
  if (j == ny + 1) sj = ph.bcondy ? ny : 1;
  // corner columns of mpdata-hotcross80.f:485-492
  int mx = ph.modulo_x, my = ph.modulo_y;
  if (gi && gj && k >= 1 && k <= nz && ((mx == 1 && my == 0) || (mx == 0 && my == 1))) {
    if (mx == 1) si = (i == 0) ? nx : 1;
    else sj = (j == 0) ? ny : 1;
  }
  T val = src[IX(L, si, sj, sk)];
  if (vent && k == nz + 1 && !gi && !gj && va.map_valid && i >= va.nxs && i <= va.nxe && j >= va.nys &&
      j <= va.nye)
    val = (T)vent_value * (T)1 + (T)0.0 * val;
  dst[IX(L, i, j, k)] = val;
}

template <typename T>
__global__ void __launch_bounds__(256) k_vent(Layout L, VentArgs va, T *__restrict__ c, double value) {
  int i = va.nxs + (int)threadIdx.x, j = va.nys + (int)blockIdx.x;
  if (!va.map_valid || i > va.nxe || j > va.nye) return;
  size_t q = IX(L, i, j, L.nz + 1);
  c[q] = (T)value * (T)1 + (T)0.0 * c[q];
}

// budgets (mpdata-hotcross80.f:384-423): FP64 partial sums per block
template <typename T>
__global__ void __launch_bounds__(256) k_budget(Layout L, double tau3, VentArgs va, const T *__restrict__ c,
                                                const double *__restrict__ cbkg, const double *__restrict__ alphae,
                                                const double *__restrict__ qpatch, double *__restrict__ part) {
  // block per (j,k) row, 256 threads over i
  int j = 1 + (int)blockIdx.x % L.ny, k = 1 + (int)blockIdx.x / L.ny;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int i = 1 + (int)threadIdx.x; i <= L.nx; i += blockDim.x) {
    double d = (double)F3(c, i, j, k) - cbkg[k];
    s0 += d;
    s1 -= alphae[i + (L.nx + 2) * j] * d;
    if (k == L.nz && va.map_valid && i >= va.nxs && i <= va.nxe && j >= va.nys && j <= va.nye)
      s2 += tau3 * qpatch[(i - va.nxs) + (va.nxe - va.nxs + 1) * (j - va.nys)];
  }
  __shared__ double sh[3][256];
  sh[0][threadIdx.x] = s0; sh[1][threadIdx.x] = s1; sh[2][threadIdx.x] = s2;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 3; ++q) part[(size_t)q * gridDim.x + blockIdx.x] = sh[q][0];
}

__global__ void __launch_bounds__(256) k_sum3(int n, const double *__restrict__ part, double *__restrict__ out,
                                              double scale0, double scale1) {
  __shared__ double sh[3][256];
  for (int q = 0; q < 3; ++q) {
    double a = 0;
    for (int r = threadIdx.x; r < n; r += 256) a += part[(size_t)q * n + r];
    sh[q][threadIdx.x] = a;
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0][0] * scale0;
    out[1] = sh[1][0] * scale1;
    out[2] = sh[2][0];
  }
}

// ---------------------------------------------------------------------------
template <typename T>
void launch_courant(hipStream_t st, const Layout &L, const StepArgs &s, const FieldPtrs<T> &f, TracerWork &tw) {
  Box b{0, L.nx + 1, 0, L.ny + 1, 0, L.nz + 1};
  hipLaunchKernelGGL(k_courant<T>, box_grid(b), box_block(), 0, st, L, b, s.tau3, s.dx, s.dy, s.dz, f.u, f.up, f.v,
                     f.vp, f.w, f.wp, (T *)tw.uc, (T *)tw.vc, (T *)tw.wc);
}

template <typename T>
void launch_tracer(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const VentArgs &va,
                   const FieldPtrs<T> &f, const Aux &aux, TracerWork &tw, T *c, const double *c_bkg,
                   const double *qpatch, const TracerArgs &ta) {
  StepArgs s2 = s;
  s2.tau3 = ta.tau3;
  Box in{1, L.nx, 1, L.ny, 1, L.nz};
  const long nghost = 2L * (L.ny + 2) * (L.nz + 2) + 2L * L.nx * (L.nz + 2) + 2L * L.nx * L.ny;
  const unsigned nbc = (unsigned)((nghost + 255) / 256);
  Box ad{0, L.nx, 0, L.ny, 0, L.nz};
  T *c1 = (T *)tw.c1, *c2 = (T *)tw.c2;
  hipLaunchKernelGGL(k_pass1<T>, box_grid(in), box_block(), 0, st, L, in, s2, ph, c, (T *)tw.uc, (T *)tw.vc,
                     (T *)tw.wc, f.ss, f.ssv, aux.azzmin, c_bkg, c1);
  hipLaunchKernelGGL(k_bc<T>, dim3(nbc), dim3(256), 0, st, L, ph, va, c1, c1, 0, 0.0);
  if (ph.ismol != 1) {
    // no corrective passes: sponge/source straight after pass 1 through the
    // limited kernel with zero antidiffusive velocities would change nothing;
    // this configuration is not used by the reference runs
    return;
  }
  // pass 2
  hipLaunchKernelGGL(k_antidiff<T>, box_grid(ad), box_block(), 0, st, L, ad, c1, (T *)tw.uc, (T *)tw.vc,
                     (T *)tw.wc, (T *)tw.ud2, (T *)tw.vd2, (T *)tw.wd2);
  hipLaunchKernelGGL(k_limited<T>, box_grid(in), box_block(), 0, st, L, in, s2, ph, va, 0, ta.cmin, ta.cmax, c1,
                     (T *)tw.ud2, (T *)tw.vd2, (T *)tw.wd2, c2, (T *)nullptr, c_bkg, aux.alphae0, aux.alphaw, qpatch);
  hipLaunchKernelGGL(k_bc<T>, dim3(nbc), dim3(256), 0, st, L, ph, va, c2, c2, 0, 0.0);
  // pass 3 (the antidiffusive velocities of pass 2 play the Courant numbers)
  hipLaunchKernelGGL(k_antidiff<T>, box_grid(ad), box_block(), 0, st, L, ad, c2, (T *)tw.ud2, (T *)tw.vd2,
                     (T *)tw.wd2, (T *)tw.ud, (T *)tw.vd, (T *)tw.wd);
  hipLaunchKernelGGL(k_limited<T>, box_grid(in), box_block(), 0, st, L, in, s2, ph, va, 1, ta.cmin, ta.cmax, c2,
                     (T *)tw.ud, (T *)tw.vd, (T *)tw.wd, c, c1, c_bkg, aux.alphae0, aux.alphaw, qpatch);
  hipLaunchKernelGGL(k_bc<T>, dim3(nbc), dim3(256), 0, st, L, ph, va, c, c1, ta.vent, ta.vent_value);
  if (ta.budgets) {
    int nb = L.ny * L.nz;
    hipLaunchKernelGGL(k_budget<T>, dim3(nb), dim3(256), 0, st, L, ta.tau3, va, c, c_bkg, aux.alphae, qpatch,
                       tw.budget_part);
    hipLaunchKernelGGL(k_sum3, dim3(1), dim3(256), 0, st, nb, tw.budget_part, tw.budget,
                       s.dx * s.dy * s.dz, s.dx * s.dy * s.dz * ta.tau3);
  }
}

template <typename T>
void launch_vent_ghost(hipStream_t st, const Layout &L, const VentArgs &va, T *c, double value) {
  if (!va.map_valid) return;
  hipLaunchKernelGGL(k_vent<T>, dim3(va.nye - va.nys + 1), dim3(va.nxe - va.nxs + 1), 0, st, L, va, c, value);
}

// CFL_VALUES reductions, hotcross80-revised2014.f:395-402
template <typename T>
__global__ void __launch_bounds__(256) k_cfl(Layout L, const T *__restrict__ u, const T *__restrict__ v,
                                             const T *__restrict__ w, double *__restrict__ part) {
  int j = (int)blockIdx.x % (L.ny + 1), k = (int)blockIdx.x / (L.ny + 1);  // j,k = 0..ny, 0..nz
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int i = (int)threadIdx.x; i <= L.nx; i += blockDim.x) {
    if (j >= 1 && k >= 1) { double a = F3(u, i, j, k); mn[0] = fmin(mn[0], a); mx[0] = fmax(mx[0], a); }
    if (i >= 1 && k >= 1) { double a = F3(v, i, j, k); mn[1] = fmin(mn[1], a); mx[1] = fmax(mx[1], a); }
    if (i >= 1 && j >= 1) { double a = F3(w, i, j, k); mn[2] = fmin(mn[2], a); mx[2] = fmax(mx[2], a); }
  }
  __shared__ double sh[6][256];
  for (int q = 0; q < 3; ++q) { sh[2 * q][threadIdx.x] = mx[q]; sh[2 * q + 1][threadIdx.x] = mn[q]; }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 3; ++q) {
        sh[2 * q][threadIdx.x] = fmax(sh[2 * q][threadIdx.x], sh[2 * q][threadIdx.x + off]);
        sh[2 * q + 1][threadIdx.x] = fmin(sh[2 * q + 1][threadIdx.x], sh[2 * q + 1][threadIdx.x + off]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) part[(size_t)q * gridDim.x + blockIdx.x] = sh[q][0];
}

__global__ void __launch_bounds__(256) k_cfl_final(int n, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[6][256];
  for (int q = 0; q < 6; ++q) {
    double a = (q % 2 == 0) ? -1e300 : 1e300;
    for (int r = threadIdx.x; r < n; r += 256)
      a = (q % 2 == 0) ? fmax(a, part[(size_t)q * n + r]) : fmin(a, part[(size_t)q * n + r]);
    sh[q][threadIdx.x] = a;
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 6; ++q)
        sh[q][threadIdx.x] = (q % 2 == 0) ? fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + off])
                                          : fmin(sh[q][threadIdx.x], sh[q][threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) out[q] = sh[q][0];
}

template <typename T>
void launch_cfl(hipStream_t st, const Layout &L, const FieldPtrs<T> &f, double *part, double *out) {
  int nb = (L.ny + 1) * (L.nz + 1);
  hipLaunchKernelGGL(k_cfl<T>, dim3(nb), dim3(256), 0, st, L, f.u, f.v, f.w, part);
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, st, nb, part, out);
}

#define HC_INST_T(T)                                                                                          \
  template void launch_courant<T>(hipStream_t, const Layout &, const StepArgs &, const FieldPtrs<T> &,        \
                                  TracerWork &);                                                              \
  template void launch_tracer<T>(hipStream_t, const Layout &, const StepArgs &, const Phys &,                 \
                                 const VentArgs &, const FieldPtrs<T> &, const Aux &, TracerWork &, T *,      \
                                 const double *, const double *, const TracerArgs &);                         \
  template void launch_vent_ghost<T>(hipStream_t, const Layout &, const VentArgs &, T *, double);             \
  template void launch_cfl<T>(hipStream_t, const Layout &, const FieldPtrs<T> &, double *, double *);
HC_INST_T(double)
HC_INST_T(float)

}  // namespace hc
