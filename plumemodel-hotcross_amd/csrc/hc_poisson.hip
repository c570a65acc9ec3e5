// hc_poisson.hip - PRESSURE and its HS3CRT solve on gfx950.
//
// The reference (hotcross80-revised2014.f:1146-1405, hs3crt_composite80.f)
// solves the 7-point staggered Laplacian with Neumann data on all six faces
// by cosine transforms in y and z and a pre-factored Thomas solve in x whose
// singular pivot (zero wavenumber in y and z) is replaced by 1 (:627-629).
// That replacement solves the compatible system obtained by removing the
// right-hand side's total delta from the plane i = nx, delta/(ny*nz) per cell;
// the null-space constant is then removed by PRESSURE's mean subtraction.
//
// The solve here keeps the operator and that compatibility rule but transforms
// in the two directions that map onto LDS best and solves along z:
//   k_rhs       RHS + boundary data (HS3CRT :38-106), FP64 block sums of it
//   k_sum1      delta = total of the RHS (FP64)
//   k_dct_x<0>  rows along x: delta correction at i=nx, DCT-II   (LDS FFT)
//   k_dct_y<0>  16-column tiles along y: DCT-II                  (LDS FFT)
//   k_tridiag   Thomas solve per (kx,ky) column in z with pre-factored
//               pivots; the singular (0,0) column in FP64 with its last pivot
//               set to 1, then its k-mean removed (PRESSURE :1325-1345)
//   k_dct_y<1>  inverse
//   k_dct_x<1>  inverse, writes p and its x and z ghost faces (:1348-1372)
//   k_pyghost   y ghost faces (:1374-1381)
// The transforms are orthonormal DCT-II/III computed as one complex FFT per
// pair of real sequences (Makhoul's reordering), mixed radix 4/2/3 Stockham
// in LDS with an FP64 twiddle table.
#include <stdlib.h>

#include "hc_common.h"
#include "hc_kernels.h"

namespace hc {

#define F3(a, i, j, k) (a)[IX(L, (i), (j), (k))]

// Sum over a block of one double per thread (blockDim.x a multiple of 64, at
// most 1024): shuffles within each wavefront, then the wavefronts' sums
// through LDS and shuffles in the first wavefront.  The total is valid in
// thread 0; every thread must call it, once per kernel.
__device__ inline double block_sum(double v) {
  __shared__ double ws[16];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
  if (lane == 0) ws[w] = v;
  __syncthreads();
  double t = 0.0;
  if (w == 0) {
    t = lane < nw ? ws[lane] : 0.0;
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
  }
  return t;
}

template <typename T>
struct cpx {
  T x, y;
};
template <typename T>
__device__ inline cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ inline cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ inline cpx<T> cmul(cpx<T> a, cpx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

struct FftPlan {
  int N;
  int nst;
  int radix[16];
};

inline FftPlan make_plan(int N) {
  FftPlan p{};
  p.N = N;
  int n = N;
  while (n % 4 == 0) { p.radix[p.nst++] = 4; n /= 4; }
  while (n % 2 == 0) { p.radix[p.nst++] = 2; n /= 2; }
  while (n % 3 == 0) { p.radix[p.nst++] = 3; n /= 3; }
  if (n != 1) p.nst = -1;
  return p;
}

// One complex FFT of length N on each of nseq sequences held in LDS (stride
// N); sign -1 forward, +1 inverse (unnormalised).  Returns the buffer that
// holds the result.
template <typename T>
__device__ cpx<T> *lds_fft(cpx<T> *a, cpx<T> *b, const FftPlan &pl, int nseq, const double *tw, int sign) {
  const int N = pl.N;
  int Ns = 1;
  const int nthr = blockDim.x;
  for (int st = 0; st < pl.nst; ++st) {
    const int R = pl.radix[st];
    const int nb = N / R;
    const int step = N / (Ns * R);
    for (int t = threadIdx.x; t < nb * nseq; t += nthr) {
      int s = t / nb, j = t - s * nb;
      const cpx<T> *in = a + (size_t)s * N;
      cpx<T> *out = b + (size_t)s * N;
      int jm = j % Ns;
      cpx<T> v[4];
      for (int r = 0; r < R; ++r) v[r] = in[j + r * nb];
      for (int r = 1; r < R; ++r) {
        int m = (r * jm * step) % N;
        cpx<T> w{(T)tw[2 * m], (T)(sign > 0 ? -tw[2 * m + 1] : tw[2 * m + 1])};
        v[r] = cmul(v[r], w);
      }
      cpx<T> y[4];
      if (R == 2) {
        y[0] = cadd(v[0], v[1]);
        y[1] = csub(v[0], v[1]);
      } else if (R == 4) {
        cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // d * (sign * i)
        cpx<T> t3 = sign < 0 ? cpx<T>{d.y, -d.x} : cpx<T>{-d.y, d.x};
        y[0] = cadd(t0, t2);
        y[2] = csub(t0, t2);
        y[1] = cadd(t1, t3);
        y[3] = csub(t1, t3);
      } else {  // R == 3
        const T c = (T)-0.5, sn = (T)(sign * 0.86602540378443864676);
        cpx<T> s12 = cadd(v[1], v[2]), d12 = csub(v[1], v[2]);
        cpx<T> t1{v[0].x + c * s12.x, v[0].y + c * s12.y};
        cpx<T> rot{-sn * d12.y, sn * d12.x};  // i * sn * d12
        y[0] = cadd(v[0], s12);
        y[1] = cadd(t1, rot);
        y[2] = csub(t1, rot);
      }
      int idxD = (j / Ns) * Ns * R + jm;
      for (int r = 0; r < R; ++r) out[idxD + r * Ns] = y[r];
    }
    __syncthreads();
    cpx<T> *tmp = a; a = b; b = tmp;
    Ns *= R;
  }
  return a;
}

// x index of the Makhoul reordering: v[n] = x[2n] (n < N/2), v[N-1-n] = x[2n+1]
__device__ inline int mperm(int n, int N) { return n < N / 2 ? 2 * n : 2 * (N - 1 - n) + 1; }

// DCT post-processing of a packed pair: Z = FFT(a + i b)  ->  orthonormal
// DCT-II of a and b at index k
template <typename T>
__device__ inline void dct_post(const cpx<T> *Z, int N, int k, const double *ct, T &xa, T &xb) {
  cpx<T> zk = Z[k], zc = Z[(N - k) % N];
  zc.y = -zc.y;
  cpx<T> A{(zk.x + zc.x) * (T)0.5, (zk.y + zc.y) * (T)0.5};
  cpx<T> D{zk.x - zc.x, zk.y - zc.y};
  cpx<T> B{D.y * (T)0.5, -D.x * (T)0.5};  // D * (-i/2)
  T c = (T)ct[2 * k], s = (T)ct[2 * k + 1];
  T sk = (T)(k == 0 ? sqrt(1.0 / N) : sqrt(2.0 / N));
  xa = sk * (A.x * c + A.y * s);
  xb = sk * (B.x * c + B.y * s);
}

// inverse pre-processing: orthonormal coefficients X -> V[k] of one sequence
template <typename T>
__device__ inline cpx<T> dct_pre_inv(T xk, T xnk, int k, int N, const double *ct) {
  T sk = (T)(k == 0 ? sqrt((double)N) : sqrt(N / 2.0));
  T snk = (T)sqrt(N / 2.0);
  T p = xk * sk, q = (k == 0) ? (T)0 : xnk * snk;
  T c = (T)ct[2 * k], s = (T)ct[2 * k + 1];
  return {p * c + q * s, p * s - q * c};
}

// ---------------------------------------------------------------------------
// right-hand side (hotcross80-revised2014.f:1191-1257 + hs3crt :38-106)
template <typename T>
__global__ void __launch_bounds__(256) k_rhs(Layout L, StepArgs s, Phys ph, const T *__restrict__ um,
                                             const T *__restrict__ vm, const T *__restrict__ wm,
                                             const T *__restrict__ w, const T *__restrict__ fx,
                                             const T *__restrict__ fy, const T *__restrict__ fz,
                                             const double *__restrict__ pgradx, const double *__restrict__ pgrady,
                                             T *__restrict__ r, double *__restrict__ part) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const int j = 1 + (int)blockIdx.x % ny, k = 1 + (int)blockIdx.x / ny;
  const double dx = s.dx, dy = s.dy, dz = s.dz, dt = s.dt;
  const double dlx = 1.0 / ((nx * dx - 0.0) / nx), dly = 1.0 / ((ny * dy - 0.0) / ny), dlz = 1.0 / ((nz * dz - 0.0) / nz);
  // the divisions by dx, dy, dz and dt through host reciprocals (UDiv: the
  // same correctly rounded quotients); leap is 1 or 2, so x / leap equals
  // x * (1 / leap) exactly
  const UDiv<T> ux = udiv_t<T>(dx, s.idx, s.idxf), uy = udiv_t<T>(dy, s.idy, s.idyf),
                uz = udiv_t<T>(dz, s.idz, s.idzf), ut = udiv_t<T>(dt, s.rdt, s.rdtf);
  const T rleap = (T)(1.0 / s.leap);
  double acc = 0.0;
  for (int i = 1 + (int)threadIdx.x; i <= nx; i += blockDim.x) {
    T v = ut(ux(F3(um, i, j, k) - F3(um, i - 1, j, k)) + uy(F3(vm, i, j, k) - F3(vm, i, j - 1, k)) +
             uz(F3(wm, i, j, k) - F3(wm, i, j, k - 1))) * rleap +
          (ux(F3(fx, i, j, k) - F3(fx, i - 1, j, k)) + uy(F3(fy, i, j, k) - F3(fy, i, j - 1, k)) +
           uz(F3(fz, i, j, k) - F3(fz, i, j, k - 1)));
    if (i == 1) {
      T bd = (T)(-(s.ubkg - s.ubkg_m) / dt - pgradx[k]) + F3(fx, 0, j, k);
      v = v + bd * (T)dlx;
    }
    if (i == nx) {
      T bd = (T)(-(s.ubkg - s.ubkg_m) / dt - pgradx[k]) + F3(fx, nx, j, k);
      v = v - bd * (T)dlx;
    }
    if (j == 1) {
      T bd = (T)(-(s.vbkg - s.vbkg_m) / dt - pgrady[k]) + F3(fy, i, 0, k);
      v = v + bd * (T)dly;
    }
    if (j == ny) {
      T bd = (T)(-(s.vbkg - s.vbkg_m) / dt - pgrady[k]) + F3(fy, i, ny, k);
      v = v - bd * (T)dly;
    }
    if (k == 1) {
      T bd = -(F3(w, i, j, 0) - F3(wm, i, j, 0)) / (T)dt + F3(fz, i, j, 0);
      v = v + bd * (T)dlz;
    }
    if (k == nz) {
      T bd = -(F3(w, i, j, nz) - F3(wm, i, j, nz)) / (T)dt + F3(fz, i, j, nz);
      v = v - bd * (T)dlz;
    }
    r[((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1)] = v;
    acc += (double)v;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// total of n block partials (one block of SUM1_THREADS threads).  Each thread
// keeps SUM1_ACC strided accumulators so that SUM1_ACC loads are in flight at
// once (a single dependent chain waited on every load: 88 us for 65,536
// partials at 256^3); the order of the additions is fixed for a given n.
constexpr int SUM1_THREADS = 1024, SUM1_ACC = 8;
__global__ void __launch_bounds__(SUM1_THREADS) k_sum1(int n, const double *__restrict__ part,
                                                       double *__restrict__ out) {
  double a[SUM1_ACC] = {};
  const int stride = SUM1_THREADS * SUM1_ACC;
  for (int q0 = (int)threadIdx.x; q0 < n; q0 += stride) {
#pragma unroll
    for (int t = 0; t < SUM1_ACC; ++t) {
      const int q = q0 + t * SUM1_THREADS;
      if (q < n) a[t] += part[q];
    }
  }
  double v = 0.0;
#pragma unroll
  for (int t = 0; t < SUM1_ACC; ++t) v += a[t];
  v = block_sum(v);
  if (threadIdx.x == 0) out[0] = v;
}

// ---------------------------------------------------------------------------
// DCT along x over row pairs.  INV=0: r -> rh (delta correction first);
// INV=1: rh -> p (padded field) with the x and z ghost faces.
template <typename T, int INV>
__global__ void __launch_bounds__(256) k_dct_x(Layout L, FftPlan pl, int npair, const double *__restrict__ tw,
                                               const double *__restrict__ ct, const T *__restrict__ in,
                                               T *__restrict__ out, const double *__restrict__ delta, StepArgs s,
                                               Phys ph, T *__restrict__ p, const T *__restrict__ fx,
                                               const T *__restrict__ w, const T *__restrict__ wm,
                                               const T *__restrict__ fz, const double *__restrict__ pgradx) {
  extern __shared__ char smem_raw[];
  const int N = L.nx, ny = L.ny, nz = L.nz;
  const int nrows = ny * nz;
  cpx<T> *A = (cpx<T> *)smem_raw, *B = A + (size_t)npair * N;
  const int row0 = blockIdx.x * 2 * npair;
  const double corr = INV ? 0.0 : delta[0] / ((double)ny * nz);
  // load
  for (int t = threadIdx.x; t < npair * N; t += blockDim.x) {
    int pq = t / N, n = t - pq * N;
    int ra = row0 + 2 * pq, rb = ra + 1;
    if (!INV) {
      int x = mperm(n, N);
      T a = ra < nrows ? in[(size_t)ra * N + x] : (T)0;
      T b = rb < nrows ? in[(size_t)rb * N + x] : (T)0;
      if (x == N - 1) { a = a - (T)corr; b = b - (T)corr; }
      A[(size_t)pq * N + n] = {a, b};
    } else {
      int k = n, nk = (N - k) % N;
      T xa = ra < nrows ? in[(size_t)ra * N + k] : (T)0, xb = rb < nrows ? in[(size_t)rb * N + k] : (T)0;
      T xna = ra < nrows ? in[(size_t)ra * N + nk] : (T)0, xnb = rb < nrows ? in[(size_t)rb * N + nk] : (T)0;
      cpx<T> va = dct_pre_inv(xa, xna, k, N, ct), vb = dct_pre_inv(xb, xnb, k, N, ct);
      A[(size_t)pq * N + n] = {va.x - vb.y, va.y + vb.x};  // va + i vb
    }
  }
  __syncthreads();
  cpx<T> *Z = lds_fft<T>(A, B, pl, npair, tw, INV ? +1 : -1);
  if (!INV) {
    for (int t = threadIdx.x; t < npair * N; t += blockDim.x) {
      int pq = t / N, k = t - pq * N;
      int ra = row0 + 2 * pq, rb = ra + 1;
      T xa, xb;
      dct_post(Z + (size_t)pq * N, N, k, ct, xa, xb);
      if (ra < nrows) out[(size_t)ra * N + k] = xa;
      if (rb < nrows) out[(size_t)rb * N + k] = xb;
    }
  } else {
    const T invN = (T)(1.0 / N);
    for (int t = threadIdx.x; t < npair * N; t += blockDim.x) {
      int pq = t / N, n = t - pq * N;
      int x = mperm(n, N);
      cpx<T> z = Z[(size_t)pq * N + n];
      for (int h = 0; h < 2; ++h) {
        int rr = row0 + 2 * pq + h;
        if (rr >= nrows) continue;
        int j = 1 + rr % ny, k = 1 + rr / ny;
        T val = (h == 0 ? z.x : z.y) * invN;
        int i = x + 1;
        F3(p, i, j, k) = val;
        // z ghost faces, :1348-1350
        if (k == 1) {
          T bd = -(F3(w, i, j, 0) - F3(wm, i, j, 0)) / (T)s.dt + F3(fz, i, j, 0);
          F3(p, i, j, 0) = val - bd * (T)s.dz;
        }
        if (k == nz) {
          T bd = -(F3(w, i, j, nz) - F3(wm, i, j, nz)) / (T)s.dt + F3(fz, i, j, nz);
          F3(p, i, j, nz + 1) = val + bd * (T)s.dz;
        }
      }
    }
    __syncthreads();
    // x ghost faces, :1352-1372 (need the values at i = 1 and i = nx)
    for (int t = threadIdx.x; t < 2 * npair; t += blockDim.x) {
      int rr = row0 + t;
      if (rr >= nrows) continue;
      int j = 1 + rr % ny, k = 1 + rr / ny;
      T p1 = F3(p, 1, j, k), pn = F3(p, N, j, k);
      T bxs = (T)(-(s.ubkg - s.ubkg_m) / s.dt - pgradx[k]) + F3(fx, 0, j, k);
      T bxf = (T)(-(s.ubkg - s.ubkg_m) / s.dt - pgradx[k]) + F3(fx, N, j, k);
      if (ph.open_box == 1) {
        F3(p, 0, j, k) = (p1 - bxs * (T)s.dx);
        F3(p, N + 1, j, k) = (pn + bxf * (T)s.dx);
      } else {
        T bx = (T)ph.bcondx;
        F3(p, 0, j, k) = (p1 - bxs * (T)s.dx) * bx + (T)(1 - ph.bcondx) * pn;
        F3(p, N + 1, j, k) = (pn + bxf * (T)s.dx) * bx + (T)(1 - ph.bcondx) * p1;
      }
    }
  }
}

// DCT along y on tiles of 2*npair x-columns of one z-plane, in place on buf
template <typename T, int INV>
__global__ void __launch_bounds__(256) k_dct_y(Layout L, FftPlan pl, int npair, const double *__restrict__ tw,
                                               const double *__restrict__ ct, T *__restrict__ buf) {
  extern __shared__ char smem_raw[];
  const int nx = L.nx, N = L.ny;
  cpx<T> *A = (cpx<T> *)smem_raw, *B = A + (size_t)npair * N;
  const int ntile = (nx + 2 * npair - 1) / (2 * npair);
  const int k = blockIdx.x / ntile, x0 = (blockIdx.x % ntile) * 2 * npair;
  T *plane = buf + (size_t)k * N * nx;
  const int TX = 2 * npair;
  for (int t = threadIdx.x; t < TX * N; t += blockDim.x) {
    int c = t % TX, n = t / TX;  // coalesced over c
    int x = x0 + c;
    int pq = c >> 1, h = c & 1;
    if (!INV) {
      // element y = n goes to position m with mperm(m) = n
      int m = (n % 2 == 0) ? n / 2 : N - 1 - (n - 1) / 2;
      T v = x < nx ? plane[(size_t)n * nx + x] : (T)0;
      T *dst = (T *)&A[(size_t)pq * N + m];
      dst[h] = v;
    } else {
      T *dst = (T *)&B[(size_t)pq * N + n];  // stage raw coefficients in B
      dst[h] = x < nx ? plane[(size_t)n * nx + x] : (T)0;
    }
  }
  __syncthreads();
  if (INV) {
    for (int t = threadIdx.x; t < npair * N; t += blockDim.x) {
      int pq = t / N, k2 = t - pq * N, nk = (N - k2) % N;
      cpx<T> xr = B[(size_t)pq * N + k2], xn = B[(size_t)pq * N + nk];
      cpx<T> va = dct_pre_inv(xr.x, xn.x, k2, N, ct), vb = dct_pre_inv(xr.y, xn.y, k2, N, ct);
      A[(size_t)pq * N + k2] = {va.x - vb.y, va.y + vb.x};
    }
    __syncthreads();
  }
  cpx<T> *Z = lds_fft<T>(A, B, pl, npair, tw, INV ? +1 : -1);
  if (!INV) {
    for (int t = threadIdx.x; t < TX * N; t += blockDim.x) {
      int c = t % TX, k2 = t / TX, x = x0 + c;
      int pq = c >> 1, h = c & 1;
      T xa, xb;
      dct_post(Z + (size_t)pq * N, N, k2, ct, xa, xb);
      if (x < nx) plane[(size_t)k2 * nx + x] = h ? xb : xa;
    }
  } else {
    const T invN = (T)(1.0 / N);
    for (int t = threadIdx.x; t < TX * N; t += blockDim.x) {
      int c = t % TX, y = t / TX, x = x0 + c;
      int pq = c >> 1, h = c & 1;
      int m = (y % 2 == 0) ? y / 2 : N - 1 - (y - 1) / 2;
      cpx<T> z = Z[(size_t)pq * N + m];
      if (x < nx) plane[(size_t)y * nx + x] = (h ? z.y : z.x) * invN;
    }
  }
}

// Thomas solve along z per (kx,ky), hs3crt_composite80.f:408-431 in the z
// direction; the (0,0) column in FP64 with the k-mean removed
// the column of the (0,0) mode lives in LDS while one thread runs its
// recurrences (nz <= TRI_MAXZ); the block that holds it loads and stores it
// with all its threads
constexpr int TRI_MAXZ = 1024;

template <typename T, int KB>
__global__ void __launch_bounds__(256) k_tridiag(Layout L, double dz, const double *__restrict__ piv,
                                                 const T *__restrict__ pivt, T *__restrict__ buf,
                                                 double *__restrict__ y00) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double a = 1.0 / (dz * dz);
  if (blockIdx.x == 0) {
    // the singular (0,0) column: FP64 throughout, last pivot 1 (set with the
    // table), then the k-mean removed
    __shared__ double g[TRI_MAXZ], d[TRI_MAXZ];
    for (int k = (int)threadIdx.x; k < nz; k += blockDim.x) {
      g[k] = (double)buf[k * plane];
      d[k] = piv[k * plane];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double y = g[0] * d[0];
      g[0] = y;
      for (int k = 1; k < nz; ++k) {
        y = (g[k] - a * y) * d[k];
        g[k] = y;
      }
      double x = y, sum = x;
      for (int k = nz - 2; k >= 0; --k) {
        x = g[k] - a * d[k] * x;
        g[k] = x;
        sum += x;
      }
      g[nz - 1] = y;
      y00[0] = sum / nz;
    }
    __syncthreads();
    const double mean = y00[0];
    for (int k = (int)threadIdx.x; k < nz; k += blockDim.x) buf[k * plane] = (T)(g[k] - mean);
  }
  if (q == 0 || q >= plane) return;
  // the sweeps are sequential in k; loads are issued in batches of KB levels
  // ahead of the recurrence so their latency overlaps, whole batches without
  // a per-level test.  The pivots come in the context's precision (pivt, the
  // FP64 values rounded once on the host), so no conversion waits on a load.
  auto load = [&](int k, int dir, int kn, T *g, T *d) {
    if (kn == KB) {
#pragma unroll
      for (int t = 0; t < KB; ++t) {
        g[t] = buf[(size_t)(k + dir * t) * plane + q];
        d[t] = pivt[(size_t)(k + dir * t) * plane + q];
      }
    } else {
#pragma unroll
      for (int t = 0; t < KB; ++t)
        if (t < kn) {
          g[t] = buf[(size_t)(k + dir * t) * plane + q];
          d[t] = pivt[(size_t)(k + dir * t) * plane + q];
        }
    }
  };
  T y = (T)0;
  for (int k0 = 0; k0 < nz; k0 += KB) {
    T g[KB], d[KB];
    const int kn = min(KB, nz - k0);
    load(k0, 1, kn, g, d);
#pragma unroll
    for (int t = 0; t < KB; ++t)
      if (t < kn) {
        y = (k0 + t == 0) ? g[t] * d[t] : (g[t] - (T)a * y) * d[t];
        buf[(size_t)(k0 + t) * plane + q] = y;
      }
  }
  T x = y;
  for (int k1 = nz - 2; k1 >= 0; k1 -= KB) {
    T g[KB], d[KB];
    const int kn = min(KB, k1 + 1);
    load(k1, -1, kn, g, d);
#pragma unroll
    for (int t = 0; t < KB; ++t)
      if (t < kn) {
        x = g[t] - (T)a * d[t] * x;
        buf[(size_t)(k1 - t) * plane + q] = x;
      }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_pyghost(Layout L, StepArgs s, Phys ph, const T *__restrict__ fy,
                                                 const double *__restrict__ pgrady, T *__restrict__ p) {
  int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x), k = 1 + (int)blockIdx.y;
  if (i > L.nx) return;
  const int ny = L.ny;
  T by = (T)ph.bcondy, oby = (T)(1 - ph.bcondy);
  T bys = (T)(-(s.vbkg - s.vbkg_m) / s.dt - pgrady[k]) + F3(fy, i, 0, k);
  T byf = (T)(-(s.vbkg - s.vbkg_m) / s.dt - pgrady[k]) + F3(fy, i, ny, k);
  T p1 = F3(p, i, 1, k), pn = F3(p, i, ny, k);
  F3(p, i, 0, k) = (p1 - bys * (T)s.dy) * by + oby * pn;
  F3(p, i, ny + 1, k) = (pn + byf * (T)s.dy) * by + oby * p1;
}

// ---------------------------------------------------------------------------
// Transforms of compile-time length N (the sizes listed in FAST_SIZES).
// Same mathematics as the kernels above, written for the hardware: the radix
// plan, every index split and every divisor are compile-time constants; the
// twiddle and DCT rotation tables sit in LDS; each Stockham stage runs in
// place (all of a stage's reads into registers, a barrier, then its writes),
// so a block holds one sequence buffer; global reads and writes run in
// memory order with the Makhoul permutation applied on the LDS side.
constexpr int fft_nst(int n) {
  int s = 0;
  while (n % 4 == 0) { n /= 4; ++s; }
  while (n % 2 == 0) { n /= 2; ++s; }
  while (n % 3 == 0) { n /= 3; ++s; }
  return n == 1 ? s : -1;
}
constexpr int fft_radix(int n, int st) {
  int s = 0;
  while (n % 4 == 0) { if (s == st) return 4; n /= 4; ++s; }
  while (n % 2 == 0) { if (s == st) return 2; n /= 2; ++s; }
  while (n % 3 == 0) { if (s == st) return 3; n /= 3; ++s; }
  return 1;
}

template <typename T, int R>
__device__ inline void butterfly(cpx<T> *v, cpx<T> *y, int sign) {
  if constexpr (R == 2) {
    y[0] = cadd(v[0], v[1]);
    y[1] = csub(v[0], v[1]);
  } else if constexpr (R == 4) {
    cpx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    cpx<T> t3 = sign < 0 ? cpx<T>{d.y, -d.x} : cpx<T>{-d.y, d.x};  // d * (sign * i)
    y[0] = cadd(t0, t2);
    y[2] = csub(t0, t2);
    y[1] = cadd(t1, t3);
    y[3] = csub(t1, t3);
  } else {  // R == 3
    const T c = (T)-0.5, sn = (T)(sign * 0.86602540378443864676);
    cpx<T> s12 = cadd(v[1], v[2]), d12 = csub(v[1], v[2]);
    cpx<T> t1{v[0].x + c * s12.x, v[0].y + c * s12.y};
    cpx<T> rot{-sn * d12.y, sn * d12.x};  // i * sn * d12
    y[0] = cadd(v[0], s12);
    y[1] = cadd(t1, rot);
    y[2] = csub(t1, rot);
  }
}

// stages ST.. of an in-place Stockham FFT of SEQ sequences of length N in
// buf, NT threads; NS = product of the radices already applied
template <typename T, int N, int SEQ, int NT, int ST, int NS>
__device__ inline void fft_stages(cpx<T> *buf, const cpx<T> *tw, int sign) {
  if constexpr (ST < fft_nst(N)) {
    constexpr int R = fft_radix(N, ST);
    constexpr int NB = N / R, STEP = N / (NS * R);
    constexpr int TOT = SEQ * NB, NBT = (TOT + NT - 1) / NT;
    cpx<T> v[NBT][R];
#pragma unroll
    for (int u = 0; u < NBT; ++u) {
      const int b = (int)threadIdx.x + u * NT;
      if (TOT % NT == 0 || b < TOT) {
        const cpx<T> *in = buf + (b / NB) * N + b % NB;
#pragma unroll
        for (int r = 0; r < R; ++r) v[u][r] = in[r * NB];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NBT; ++u) {
      const int b = (int)threadIdx.x + u * NT;
      if (TOT % NT == 0 || b < TOT) {
        const int sq = b / NB, j = b % NB, jm = j % NS;
        if constexpr (NS > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            cpx<T> w = tw[r * jm * STEP];
            if (sign > 0) w.y = -w.y;
            v[u][r] = cmul(v[u][r], w);
          }
        }
        cpx<T> y[R];
        butterfly<T, R>(v[u], y, sign);
        cpx<T> *out = buf + sq * N + (j / NS) * NS * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) out[r * NS] = y[r];
      }
    }
    __syncthreads();
    fft_stages<T, N, SEQ, NT, ST + 1, NS * R>(buf, tw, sign);
  }
}

// LDS copies of the twiddles exp(-2 pi i m/N) and rotations (cos, sin)(pi k/2N)
template <typename T, int N, int NT>
__device__ inline void load_tables(cpx<T> *tw, cpx<T> *ct, const double *g_tw, const double *g_ct) {
  for (int m = (int)threadIdx.x; m < N; m += NT) {
    tw[m] = {(T)g_tw[2 * m], (T)g_tw[2 * m + 1]};
    ct[m] = {(T)g_ct[2 * m], (T)g_ct[2 * m + 1]};
  }
}

// position of x in the Makhoul order: v[n] = x[2n] (n < N/2), v[N-1-n] = x[2n+1]
template <int N>
__device__ inline int minv(int x) { return (x % 2 == 0) ? x / 2 : N - 1 - (x - 1) / 2; }

// orthonormal DCT-II of the pair packed in Z at index k
template <typename T, int N>
__device__ inline void dct_post_t(const cpx<T> *Z, const cpx<T> *ct, int k, T &xa, T &xb) {
  cpx<T> zk = Z[k], zc = Z[(N - k) % N];
  zc.y = -zc.y;
  cpx<T> A{(zk.x + zc.x) * (T)0.5, (zk.y + zc.y) * (T)0.5};
  cpx<T> D{zk.x - zc.x, zk.y - zc.y};
  cpx<T> B{D.y * (T)0.5, -D.x * (T)0.5};  // D * (-i/2)
  const T c = ct[k].x, sn = ct[k].y;
  const T sk = k == 0 ? (T)sqrt(1.0 / N) : (T)sqrt(2.0 / N);
  xa = sk * (A.x * c + A.y * sn);
  xb = sk * (B.x * c + B.y * sn);
}

// orthonormal coefficients -> V[k] of one sequence (inverse pre-processing)
template <typename T, int N>
__device__ inline cpx<T> dct_pre_t(T xk, T xnk, int k, const cpx<T> *ct) {
  const T sk = k == 0 ? (T)sqrt((double)N) : (T)sqrt(N / 2.0);
  const T snk = (T)sqrt(N / 2.0);
  T pp = xk * sk, qq = (k == 0) ? (T)0 : xnk * snk;
  const T c = ct[k].x, sn = ct[k].y;
  return {pp * c + qq * sn, pp * sn - qq * c};
}

template <int N>
struct DctCfg {
  static constexpr int SEQ = (2048 / N) < 1 ? 1 : 2048 / N;  // x: row pairs per block
  static constexpr int SEQY = (2048 / N) < 1 ? 1 : ((2048 / N) > 8 ? 8 : 2048 / N);  // y: column pairs
};

// DCT along x, INV=0: r -> r (delta correction at x = N-1 first);
// INV=1: r -> p with the x and z ghost faces
template <typename T, int N, int INV>
__global__ void __launch_bounds__(256) k_dctx_fast(Layout L, const double *__restrict__ g_tw,
                                                   const double *__restrict__ g_ct, const T *__restrict__ in,
                                                   T *__restrict__ out, const double *__restrict__ delta,
                                                   StepArgs s, Phys ph, T *__restrict__ p, const T *__restrict__ fx,
                                                   const T *__restrict__ w, const T *__restrict__ wm,
                                                   const T *__restrict__ fz, const double *__restrict__ pgradx) {
  constexpr int SEQ = DctCfg<N>::SEQ, NT = 256;
  __shared__ cpx<T> buf[SEQ * N];
  __shared__ cpx<T> tw[N], ct[N];
  const int ny = L.ny, nz = L.nz;
  const int nrows = ny * nz;
  const int row0 = (int)blockIdx.x * 2 * SEQ;
  load_tables<T, N, NT>(tw, ct, g_tw, g_ct);
  if (!INV) {
    const T corr = (T)(delta[0] / ((double)ny * nz));
    for (int t = (int)threadIdx.x; t < 2 * SEQ * N; t += NT) {
      const int rr = t / N, x = t % N;  // memory order
      const int row = row0 + rr;
      T v = row < nrows ? in[(size_t)row * N + x] : (T)0;
      if (x == N - 1) v = v - corr;
      T *dst = (T *)&buf[(rr >> 1) * N + minv<N>(x)];
      dst[rr & 1] = v;
    }
  } else {
    __syncthreads();  // ct
    for (int t = (int)threadIdx.x; t < SEQ * N; t += NT) {
      const int pq = t / N, k = t % N, nk = (N - k) % N;
      const int ra = row0 + 2 * pq, rb = ra + 1;
      T xa = ra < nrows ? in[(size_t)ra * N + k] : (T)0, xb = rb < nrows ? in[(size_t)rb * N + k] : (T)0;
      T xna = ra < nrows ? in[(size_t)ra * N + nk] : (T)0, xnb = rb < nrows ? in[(size_t)rb * N + nk] : (T)0;
      cpx<T> va = dct_pre_t<T, N>(xa, xna, k, ct), vb = dct_pre_t<T, N>(xb, xnb, k, ct);
      buf[pq * N + k] = {va.x - vb.y, va.y + vb.x};  // va + i vb
    }
  }
  __syncthreads();
  fft_stages<T, N, SEQ, NT, 0, 1>(buf, tw, INV ? +1 : -1);
  if (!INV) {
    for (int t = (int)threadIdx.x; t < 2 * SEQ * N; t += NT) {
      const int rr = t / N, k = t % N;
      const int row = row0 + rr;
      T xa, xb;
      dct_post_t<T, N>(buf + (rr >> 1) * N, ct, k, xa, xb);
      if (row < nrows) out[(size_t)row * N + k] = (rr & 1) ? xb : xa;
    }
  } else {
    const T invN = (T)(1.0 / N);
    for (int t = (int)threadIdx.x; t < 2 * SEQ * N; t += NT) {
      const int rr = t / N, x = t % N;
      const int row = row0 + rr;
      if (row >= nrows) continue;
      const cpx<T> z = buf[(rr >> 1) * N + minv<N>(x)];
      const int j = 1 + row % ny, k = 1 + row / ny, i = x + 1;
      const T val = ((rr & 1) ? z.y : z.x) * invN;
      F3(p, i, j, k) = val;
      // z ghost faces, hotcross80-revised2014.f:1348-1350
      if (k == 1) {
        T bd = -(F3(w, i, j, 0) - F3(wm, i, j, 0)) / (T)s.dt + F3(fz, i, j, 0);
        F3(p, i, j, 0) = val - bd * (T)s.dz;
      }
      if (k == nz) {
        T bd = -(F3(w, i, j, nz) - F3(wm, i, j, nz)) / (T)s.dt + F3(fz, i, j, nz);
        F3(p, i, j, nz + 1) = val + bd * (T)s.dz;
      }
    }
    // x ghost faces, :1352-1372, from the values at i = 1 and i = N (in LDS)
    for (int t = (int)threadIdx.x; t < 2 * SEQ; t += NT) {
      const int row = row0 + t;
      if (row >= nrows) continue;
      const int j = 1 + row % ny, k = 1 + row / ny;
      const cpx<T> z1 = buf[(t >> 1) * N + minv<N>(0)], zn = buf[(t >> 1) * N + minv<N>(N - 1)];
      const T p1 = ((t & 1) ? z1.y : z1.x) * invN, pn = ((t & 1) ? zn.y : zn.x) * invN;
      T bxs = (T)(-(s.ubkg - s.ubkg_m) / s.dt - pgradx[k]) + F3(fx, 0, j, k);
      T bxf = (T)(-(s.ubkg - s.ubkg_m) / s.dt - pgradx[k]) + F3(fx, N, j, k);
      if (ph.open_box == 1) {
        F3(p, 0, j, k) = (p1 - bxs * (T)s.dx);
        F3(p, N + 1, j, k) = (pn + bxf * (T)s.dx);
      } else {
        T bx = (T)ph.bcondx;
        F3(p, 0, j, k) = (p1 - bxs * (T)s.dx) * bx + (T)(1 - ph.bcondx) * pn;
        F3(p, N + 1, j, k) = (pn + bxf * (T)s.dx) * bx + (T)(1 - ph.bcondx) * p1;
      }
    }
  }
}

// DCT along y on tiles of 2*SEQY x-columns of one z-plane, in place on buf
template <typename T, int N, int INV>
__global__ void __launch_bounds__(256) k_dcty_fast(Layout L, const double *__restrict__ g_tw,
                                                   const double *__restrict__ g_ct, T *__restrict__ data) {
  constexpr int SEQ = DctCfg<N>::SEQY, TX = 2 * SEQ, NT = 256;
  __shared__ cpx<T> buf[SEQ * N];
  __shared__ cpx<T> tw[N], ct[N];
  const int nx = L.nx;
  const int ntile = (nx + TX - 1) / TX;
  const int k = (int)blockIdx.x / ntile, x0 = ((int)blockIdx.x % ntile) * TX;
  T *plane = data + (size_t)k * N * nx;
  load_tables<T, N, NT>(tw, ct, g_tw, g_ct);
  if (!INV) {
    for (int t = (int)threadIdx.x; t < TX * N; t += NT) {
      const int c = t % TX, n = t / TX, x = x0 + c;  // coalesced over c
      T v = x < nx ? plane[(size_t)n * nx + x] : (T)0;
      T *dst = (T *)&buf[(c >> 1) * N + minv<N>(n)];
      dst[c & 1] = v;
    }
    __syncthreads();
  } else {
    // coefficients straight from global memory into the pre-processing
    __syncthreads();  // ct
    for (int t = (int)threadIdx.x; t < SEQ * N; t += NT) {
      const int pq = t % SEQ, k2 = t / SEQ, nk = (N - k2) % N;
      const int xa = x0 + 2 * pq, xb = xa + 1;
      T ra = xa < nx ? plane[(size_t)k2 * nx + xa] : (T)0, rb = xb < nx ? plane[(size_t)k2 * nx + xb] : (T)0;
      T rna = xa < nx ? plane[(size_t)nk * nx + xa] : (T)0, rnb = xb < nx ? plane[(size_t)nk * nx + xb] : (T)0;
      cpx<T> va = dct_pre_t<T, N>(ra, rna, k2, ct), vb = dct_pre_t<T, N>(rb, rnb, k2, ct);
      buf[pq * N + k2] = {va.x - vb.y, va.y + vb.x};
    }
    __syncthreads();
  }
  fft_stages<T, N, SEQ, NT, 0, 1>(buf, tw, INV ? +1 : -1);
  if (!INV) {
    for (int t = (int)threadIdx.x; t < TX * N; t += NT) {
      const int c = t % TX, k2 = t / TX, x = x0 + c;
      T xa, xb;
      dct_post_t<T, N>(buf + (c >> 1) * N, ct, k2, xa, xb);
      if (x < nx) plane[(size_t)k2 * nx + x] = (c & 1) ? xb : xa;
    }
  } else {
    const T invN = (T)(1.0 / N);
    for (int t = (int)threadIdx.x; t < TX * N; t += NT) {
      const int c = t % TX, y = t / TX, x = x0 + c;
      const cpx<T> z = buf[(c >> 1) * N + minv<N>(y)];
      if (x < nx) plane[(size_t)y * nx + x] = ((c & 1) ? z.y : z.x) * invN;
    }
  }
}

// lengths with compile-time transforms; others use the generic kernels
#define FAST_SIZES(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

template <typename T, int INV>
static bool launch_dctx_fast(hipStream_t st, const Layout &L, const double *twx, const double *ctx, const T *in,
                             T *out, const double *delta, const StepArgs &s, const Phys &ph, T *p, const T *fx,
                             const T *w, const T *wm, const T *fz, const double *pgradx) {
  const int nrows = L.ny * L.nz;
  switch (L.nx) {
#define HC_X(NN)                                                                                         \
  case NN: {                                                                                             \
    const int g = (nrows + 2 * DctCfg<NN>::SEQ - 1) / (2 * DctCfg<NN>::SEQ);                             \
    hipLaunchKernelGGL((k_dctx_fast<T, NN, INV>), dim3(g), dim3(256), 0, st, L, twx, ctx, in, out, delta, s, \
                       ph, p, fx, w, wm, fz, pgradx);                                                    \
    return true;                                                                                         \
  }
    FAST_SIZES(HC_X)
#undef HC_X
    default:
      return false;
  }
}

template <typename T, int INV>
static bool launch_dcty_fast(hipStream_t st, const Layout &L, const double *twy, const double *cty, T *data) {
  switch (L.ny) {
#define HC_Y(NN)                                                                        \
  case NN: {                                                                            \
    const int g = L.nz * ((L.nx + 2 * DctCfg<NN>::SEQY - 1) / (2 * DctCfg<NN>::SEQY));  \
    hipLaunchKernelGGL((k_dcty_fast<T, NN, INV>), dim3(g), dim3(256), 0, st, L, twy, cty, data); \
    return true;                                                                        \
  }
    FAST_SIZES(HC_Y)
#undef HC_Y
    default:
      return false;
  }
}

// ---------------------------------------------------------------------------
static int pick_npair(int N, size_t elem) {
  // LDS: 2 buffers x npair x N complex elements, at most 64 KiB
  int np = 8;
  while (np > 1 && (size_t)2 * np * N * 2 * elem > 65536) np >>= 1;
  return np;
}

// compile-time transforms on (HC_GENERIC_DCT=1 selects the generic kernels,
// for comparing the two)
static bool fast_transforms() {
  static bool f = [] {
    const char *e = getenv("HC_GENERIC_DCT");
    return !(e && e[0] == '1');
  }();
  return f;
}

template <typename T>
void launch_pressure(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                     const Aux &aux, PoissonWork &pw) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  T *r = (T *)pw.r;
  hipLaunchKernelGGL(k_rhs<T>, dim3(ny * nz), dim3(256), 0, st, L, s, ph, f.um, f.vm, f.wm, f.w, f.fx, f.fy, f.fz,
                     aux.pgradx, aux.pgrady, r, pw.part);
  hipLaunchKernelGGL(k_sum1, dim3(1), dim3(SUM1_THREADS), 0, st, ny * nz, pw.part, pw.delta);
  FftPlan px = make_plan(nx), py = make_plan(ny);
  int npx = pick_npair(nx, sizeof(T)), npy = pick_npair(ny, sizeof(T));
  size_t shx = (size_t)2 * npx * nx * 2 * sizeof(T), shy = (size_t)2 * npy * ny * 2 * sizeof(T);
  int nrows = ny * nz;
  int gx = (nrows + 2 * npx - 1) / (2 * npx);
  const double *twx = pw.tw_x, *ctx = pw.tw_x + 2 * nx;
  const double *twy = pw.tw_y, *cty = pw.tw_y + 2 * ny;
  const bool fast = fast_transforms();
  if (!(fast && launch_dctx_fast<T, 0>(st, L, twx, ctx, r, r, pw.delta, s, ph, (T *)nullptr, (const T *)nullptr,
                                       (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                                       (const double *)nullptr)))
    hipLaunchKernelGGL((k_dct_x<T, 0>), dim3(gx), dim3(256), shx, st, L, px, npx, twx, ctx, r, r, pw.delta, s, ph,
                       (T *)nullptr, (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                       (const T *)nullptr, (const double *)nullptr);
  int gy = nz * ((nx + 2 * npy - 1) / (2 * npy));
  if (!(fast && launch_dcty_fast<T, 0>(st, L, twy, cty, r)))
    hipLaunchKernelGGL((k_dct_y<T, 0>), dim3(gy), dim3(256), shy, st, L, py, npy, twy, cty, r);
  {
    // levels per batch of loads in the z sweeps: 32 in FP64 (pressure 0.77
    // -> 0.69 ms per step at 256^3 against 8), 16 in FP32 (no difference
    // measured); HC_TRI_KB = 8, 16 or 32 overrides
    static const int kb_env = [] {
      const char *e = getenv("HC_TRI_KB");
      const int v = e ? atoi(e) : 0;
      return (v == 8 || v == 16 || v == 32) ? v : 0;
    }();
    const int kb = kb_env ? kb_env : (sizeof(T) == 8 ? 32 : 16);
    const dim3 tg((unsigned)(((size_t)nx * ny + 255) / 256));
    const T *pt = (const T *)pw.pivt;
    if (kb == 8) hipLaunchKernelGGL((k_tridiag<T, 8>), tg, dim3(256), 0, st, L, s.dz, pw.piv, pt, r, pw.part);
    else if (kb == 32) hipLaunchKernelGGL((k_tridiag<T, 32>), tg, dim3(256), 0, st, L, s.dz, pw.piv, pt, r, pw.part);
    else hipLaunchKernelGGL((k_tridiag<T, 16>), tg, dim3(256), 0, st, L, s.dz, pw.piv, pt, r, pw.part);
  }
  if (!(fast && launch_dcty_fast<T, 1>(st, L, twy, cty, r)))
    hipLaunchKernelGGL((k_dct_y<T, 1>), dim3(gy), dim3(256), shy, st, L, py, npy, twy, cty, r);
  if (!(fast && launch_dctx_fast<T, 1>(st, L, twx, ctx, r, (T *)nullptr, pw.delta, s, ph, f.p, f.fx, f.w, f.wm,
                                       f.fz, aux.pgradx)))
    hipLaunchKernelGGL((k_dct_x<T, 1>), dim3(gx), dim3(256), shx, st, L, px, npx, twx, ctx, r, (T *)nullptr,
                       pw.delta, s, ph, f.p, f.fx, f.w, f.wm, f.fz, aux.pgradx);
  hipLaunchKernelGGL(k_pyghost<T>, dim3((nx + 255) / 256, nz), dim3(256), 0, st, L, s, ph, f.fy, aux.pgrady, f.p);
}


// ---------------------------------------------------------------------------
// Adjoint of PRESSURE.  The forward map is affine in its inputs:
//   r  = RHS(um, vm, wm, w, fx, fy, fz) + boundary constants    (k_rhs)
//   r' = r - e_nx * sum(r) / (ny nz)                               (delta rule)
//   X  = Dy Dx r'        (orthonormal DCT-II, x then y)
//   Y  = S X             (z tridiagonal per (kx,ky); the (0,0) column with
//                         its last pivot 1, then its k-mean removed)
//   p  = Dx^-1 Dy^-1 Y,  ghosts of p from p and the boundary data.
// Its transpose runs the steps backwards: the ghost rules' adjoint, Dy Dx
// (the inverse of an orthonormal transform is its transpose), the transposed
// Thomas recurrences, Dx^-1 Dy^-1, the delta rule's transpose
// (subtract the i = nx plane's mean from every cell), and the divergence's
// transpose (a difference in the other direction).

// interior adjoint of p in the compact layout, plus the boundary data's
// share of the ghost adjoints (one thread per interior cell; the ghost cells
// next to it are handled by it alone)
template <typename T>
__global__ void __launch_bounds__(256) k_padj_ghost(Layout L, StepArgs s, Phys ph, FlowSens fs,
                                                    const T *__restrict__ lp, T *__restrict__ lfx,
                                                    T *__restrict__ lfy, T *__restrict__ lfz, T *__restrict__ lw,
                                                    T *__restrict__ lwm, T *__restrict__ g, double *gflow) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x), j = 1 + (int)blockIdx.y, k = 1 + (int)blockIdx.z;
  double gf = 0.0;
  if (i <= nx) {
    T v = F3(lp, i, j, k);
    // z ghosts (:1348-1350): p(i,j,0) = p(i,j,1) - bd0 dz, p(i,j,nz+1) = p(i,j,nz) + bdn dz,
    // bd = -(w - wm)/dt + fz at k = 0 / nz
    if (k == 1) {
      const T a = F3(lp, i, j, 0);
      v += a;
      F3(lw, i, j, 0) += a * (T)(s.dz / s.dt);
      F3(lwm, i, j, 0) -= a * (T)(s.dz / s.dt);
      F3(lfz, i, j, 0) -= a * (T)s.dz;
    }
    if (k == nz) {
      const T a = F3(lp, i, j, nz + 1);
      v += a;
      F3(lw, i, j, nz) -= a * (T)(s.dz / s.dt);
      F3(lwm, i, j, nz) += a * (T)(s.dz / s.dt);
      F3(lfz, i, j, nz) += a * (T)s.dz;
    }
    // x ghosts (:1352-1372), bxs/bxf = -(ubkg - ubkg_m)/dt - pgradx + fx(0|nx)
    const T bx = ph.open_box == 1 ? (T)1 : (T)ph.bcondx, obx = ph.open_box == 1 ? (T)0 : (T)(1 - ph.bcondx);
    const double dbx = -(fs.dubkg - fs.dubkg_m) / s.dt;
    if (i == 1) {
      const T a0 = F3(lp, 0, j, k), an = F3(lp, nx + 1, j, k);
      v += bx * a0 + obx * an;
      F3(lfx, 0, j, k) -= a0 * bx * (T)s.dx;
      gf += -(double)(a0 * bx) * s.dx * dbx;
    }
    if (i == nx) {
      const T a0 = F3(lp, 0, j, k), an = F3(lp, nx + 1, j, k);
      v += bx * an + obx * a0;
      F3(lfx, nx, j, k) += an * bx * (T)s.dx;
      gf += (double)(an * bx) * s.dx * dbx;
    }
    // y ghosts (:1374-1381)
    const T by = (T)ph.bcondy, oby = (T)(1 - ph.bcondy);
    const double dby = -(fs.dvbkg - fs.dvbkg_m) / s.dt;
    if (j == 1) {
      const T a0 = F3(lp, i, 0, k), an = F3(lp, i, ny + 1, k);
      v += by * a0 + oby * an;
      F3(lfy, i, 0, k) -= a0 * by * (T)s.dy;
      gf += -(double)(a0 * by) * s.dy * dby;
    }
    if (j == ny) {
      const T a0 = F3(lp, i, 0, k), an = F3(lp, i, ny + 1, k);
      v += by * an + oby * a0;
      F3(lfy, i, ny, k) += an * by * (T)s.dy;
      gf += (double)(an * by) * s.dy * dby;
    }
    g[((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1)] = v;
  }
  __shared__ double sh[256];
  const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
  sh[tid] = gf;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  // one of ADJ_SLOTS partial sums (blocks spread over them)
  const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (tid == 0 && sh[0] != 0.0) atomicAdd(gflow + (bid & (ADJ_SLOTS - 1)), sh[0]);
}

// transposed Thomas recurrences of k_tridiag, per (kx,ky) column
template <typename T>
__global__ void __launch_bounds__(256) k_tridiag_adj(Layout L, double dz, const double *__restrict__ piv,
                                                     const T *__restrict__ pivt, T *__restrict__ buf) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const size_t plane = (size_t)nx * ny;
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double a = 1.0 / (dz * dz);
  if (blockIdx.x == 0) {
    // the (0,0) column in FP64 and in LDS: mean removal (a symmetric
    // projection) first, then the recurrences transposed
    __shared__ double g[TRI_MAXZ], d[TRI_MAXZ];
    for (int k = (int)threadIdx.x; k < nz; k += blockDim.x) {
      g[k] = (double)buf[k * plane];
      d[k] = piv[k * plane];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double mean = 0.0;
      for (int k = 0; k < nz; ++k) mean += g[k];
      mean /= nz;
      // back substitution x_k = y_k - a d_k x_{k+1}, transposed (k ascending)
      double lx_next = g[0] - mean;
      for (int k = 0; k < nz - 1; ++k) {
        const double ly = lx_next;
        lx_next = (g[k + 1] - mean) - a * d[k] * ly;
        g[k] = ly;
      }
      g[nz - 1] = lx_next;
      // forward elimination y_k = (g_k - a y_{k-1}) d_k, transposed (k descending)
      double ly = g[nz - 1];
      for (int k = nz - 1; k >= 1; --k) {
        const double lg = d[k] * ly;
        ly = g[k - 1] - a * lg;
        g[k] = lg;
      }
      g[0] = d[0] * ly;
    }
    __syncthreads();
    for (int k = (int)threadIdx.x; k < nz; k += blockDim.x) buf[k * plane] = (T)g[k];
  }
  if (q == 0 || q >= plane) return;
  // as k_tridiag: loads in batches of KB levels ahead of the recurrence,
  // pivots in the context's precision.  Each batch loads the levels it reads
  // before it stores any, and every level is read before the sweep stores to
  // it, as in the level-by-level form.
  constexpr int KB = sizeof(T) == 8 ? 32 : 16;
  // levels kg + dir*t of buf and kd + dir*t of the pivots, t < kn
  auto load = [&](int kg, int kd, int dir, int kn, T *g, T *d) {
    if (kn == KB) {
#pragma unroll
      for (int t = 0; t < KB; ++t) {
        g[t] = buf[(size_t)(kg + dir * t) * plane + q];
        d[t] = pivt[(size_t)(kd + dir * t) * plane + q];
      }
    } else {
#pragma unroll
      for (int t = 0; t < KB; ++t)
        if (t < kn) {
          g[t] = buf[(size_t)(kg + dir * t) * plane + q];
          d[t] = pivt[(size_t)(kd + dir * t) * plane + q];
        }
    }
  };
  T lx = buf[q];
  for (int k0 = 0; k0 < nz - 1; k0 += KB) {
    const int kn = min(KB, nz - 1 - k0);
    T g[KB], d[KB];
    load(k0 + 1, k0, 1, kn, g, d);
#pragma unroll
    for (int t = 0; t < KB; ++t)
      if (t < kn) {
        const T ly = lx;
        buf[(size_t)(k0 + t) * plane + q] = ly;
        lx = g[t] - (T)a * d[t] * ly;
      }
  }
  buf[(size_t)(nz - 1) * plane + q] = lx;
  T ly = lx;
  for (int k1 = nz - 1; k1 >= 1; k1 -= KB) {
    const int kn = min(KB, k1);
    T g[KB], d[KB];
    load(k1 - 1, k1, -1, kn, g, d);
#pragma unroll
    for (int t = 0; t < KB; ++t)
      if (t < kn) {
        const T lg = d[t] * ly;
        buf[(size_t)(k1 - t) * plane + q] = lg;
        ly = g[t] - (T)a * lg;
      }
  }
  buf[q] = pivt[q] * ly;
}

// sum of the adjoint over the plane i = nx (the delta rule's transpose)
template <typename T>
__global__ void __launch_bounds__(256) k_padj_plane(Layout L, const T *__restrict__ pad, double *__restrict__ part) {
  const int k = 1 + (int)blockIdx.x;
  double a = 0;
  for (int j = 1 + (int)threadIdx.x; j <= L.ny; j += blockDim.x) a += (double)F3(pad, L.nx, j, k);
  __shared__ double sh[256];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// the quotients k_padj_rhs uses, formed once on the host (the same
// correctly rounded FP64 values each thread would form)
struct PadjConsts {
  double rdx, rdy, rdz, rdt;  // 1/dx, 1/dy, 1/dz, 1/dt
  double dlx, dly, dlz;       // 1 / ((n d - 0) / n) per direction
  double cm;                  // 1 / dt / leap
  double dbdu, dbdv;          // -(d ubkg - d ubkg_m) / dt, the same for v
};
static PadjConsts padj_consts(const Layout &L, const StepArgs &s, const FlowSens &fs) {
  PadjConsts c;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const double dx = s.dx, dy = s.dy, dz = s.dz, dt = s.dt;
  c.rdx = 1.0 / dx; c.rdy = 1.0 / dy; c.rdz = 1.0 / dz; c.rdt = 1.0 / dt;
  c.dlx = 1.0 / ((nx * dx - 0.0) / nx); c.dly = 1.0 / ((ny * dy - 0.0) / ny); c.dlz = 1.0 / ((nz * dz - 0.0) / nz);
  c.cm = 1.0 / dt / s.leap;
  c.dbdu = -(fs.dubkg - fs.dubkg_m) / dt;
  c.dbdv = -(fs.dvbkg - fs.dvbkg_m) / dt;
  return c;
}

// transpose of k_rhs: one thread per face (i,j,k), i,j,k = 0..n; the 256
// threads of a block take 256 consecutive (i,j) of a plane (i fastest, rows
// of nx+1 wrapping), blockIdx.z = k
template <typename T>
__global__ void __launch_bounds__(256) k_padj_rhs(Layout L, PadjConsts c, const T *__restrict__ pad,
                                                  const double *__restrict__ dsum, T *__restrict__ lum,
                                                  T *__restrict__ lvm, T *__restrict__ lwm, T *__restrict__ lw,
                                                  T *__restrict__ lfx, T *__restrict__ lfy, T *__restrict__ lfz,
                                                  double *gflow) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const int t = (int)(blockIdx.x * 256u + threadIdx.y * 64u + threadIdx.x);
  const int j = t / (nx + 1), i = t - j * (nx + 1), k = (int)blockIdx.z;
  double gf = 0.0;
  if (j <= ny) {
    const T corr = (T)(dsum[0] / ((double)ny * nz));
    // every value the face (i,j,k) needs is loaded first, unconditionally
    // (all at box indices 0..n+1, inside the padded allocation), and the
    // terms that do not exist are masked afterwards, so that the loads are
    // in flight together.  The adjoint of p is taken at the interior cells
    // only (zero at the ghosts); the boundary data's terms read the same
    // four values: at i = 0 p(1) is p(i+1), at i = nx p(nx) is p(i), and
    // the same in y and z.
    const T p0 = F3(pad, i, j, k), px = F3(pad, i + 1, j, k), py = F3(pad, i, j + 1, k), pz = F3(pad, i, j, k + 1);
    const T oum = F3(lum, i, j, k), ofx = F3(lfx, i, j, k), ovm = F3(lvm, i, j, k), ofy = F3(lfy, i, j, k);
    const T owm = F3(lwm, i, j, k), ofz = F3(lfz, i, j, k), ow = F3(lw, i, j, k);
    const T L0 = (i >= 1 && j >= 1 && k >= 1) ? p0 - corr : (T)0;
    const T Lx = (i + 1 <= nx && j >= 1 && k >= 1) ? px - corr : (T)0;
    const T Ly = (i >= 1 && j + 1 <= ny && k >= 1) ? py - corr : (T)0;
    const T Lz = (i >= 1 && j >= 1 && k + 1 <= nz) ? pz - corr : (T)0;
    const double dlx = c.dlx, dly = c.dly, dlz = c.dlz;
    const T cm = (T)c.cm;
    if (j >= 1 && k >= 1) {
      const T d = L0 - Lx;
      F3(lum, i, j, k) = oum + d * (T)c.rdx * cm;
      T f = d * (T)c.rdx;
      const double dbd = c.dbdu;
      if (i == 0) { f += Lx * (T)dlx; gf += (double)Lx * dlx * dbd; }
      if (i == nx) { f -= L0 * (T)dlx; gf -= (double)L0 * dlx * dbd; }
      F3(lfx, i, j, k) = ofx + f;
    }
    if (i >= 1 && k >= 1) {
      const T d = L0 - Ly;
      F3(lvm, i, j, k) = ovm + d * (T)c.rdy * cm;
      T f = d * (T)c.rdy;
      const double dbd = c.dbdv;
      if (j == 0) { f += Ly * (T)dly; gf += (double)Ly * dly * dbd; }
      if (j == ny) { f -= L0 * (T)dly; gf -= (double)L0 * dly * dbd; }
      F3(lfy, i, j, k) = ofy + f;
    }
    if (i >= 1 && j >= 1) {
      const T d = L0 - Lz;
      T wmv = d * (T)c.rdz * cm;
      T f = d * (T)c.rdz;
      if (k == 0) {
        const T a = Lz * (T)dlz;
        f += a;
        F3(lw, i, j, 0) = ow - a * (T)c.rdt;
        wmv += a * (T)c.rdt;
      }
      if (k == nz) {
        const T a = L0 * (T)dlz;
        f -= a;
        F3(lw, i, j, nz) = ow + a * (T)c.rdt;
        wmv -= a * (T)c.rdt;
      }
      F3(lwm, i, j, k) = owm + wmv;
      F3(lfz, i, j, k) = ofz + f;
    }
  }
  __shared__ double sh[256];
  const int tid = (int)(threadIdx.x + blockDim.x * threadIdx.y);
  sh[tid] = gf;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  // one of ADJ_SLOTS partial sums (blocks spread over them)
  const unsigned bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (tid == 0 && sh[0] != 0.0) atomicAdd(gflow + (bid & (ADJ_SLOTS - 1)), sh[0]);
}

template <typename T>
void launch_pressure_adj(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                         const FieldPtrs<T> &lam, const Aux &aux, PoissonWork &pw, const FlowSens &fs,
                         double *gflow, T *g, const double *zero, double *part) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  hipLaunchKernelGGL(k_padj_ghost<T>, dim3((nx + 255) / 256, ny, nz), dim3(256), 0, st, L, s, ph, fs,
                     (const T *)lam.p, lam.fx, lam.fy, lam.fz, lam.w, lam.wm, g, gflow);
  FftPlan px = make_plan(nx), py = make_plan(ny);
  int npx = pick_npair(nx, sizeof(T)), npy = pick_npair(ny, sizeof(T));
  size_t shx = (size_t)2 * npx * nx * 2 * sizeof(T), shy = (size_t)2 * npy * ny * 2 * sizeof(T);
  int gx = (ny * nz + 2 * npx - 1) / (2 * npx);
  int gy = nz * ((nx + 2 * npy - 1) / (2 * npy));
  const double *twx = pw.tw_x, *ctx = pw.tw_x + 2 * nx;
  const double *twy = pw.tw_y, *cty = pw.tw_y + 2 * ny;
  const bool fast = fast_transforms();
  // Dy Dx (forward transforms, no delta rule)
  if (!(fast && launch_dctx_fast<T, 0>(st, L, twx, ctx, g, g, zero, s, ph, (T *)nullptr, (const T *)nullptr,
                                       (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                                       (const double *)nullptr)))
    hipLaunchKernelGGL((k_dct_x<T, 0>), dim3(gx), dim3(256), shx, st, L, px, npx, twx, ctx, g, g, zero, s, ph,
                       (T *)nullptr, (const T *)nullptr, (const T *)nullptr, (const T *)nullptr,
                       (const T *)nullptr, (const double *)nullptr);
  if (!(fast && launch_dcty_fast<T, 0>(st, L, twy, cty, g)))
    hipLaunchKernelGGL((k_dct_y<T, 0>), dim3(gy), dim3(256), shy, st, L, py, npy, twy, cty, g);
  hipLaunchKernelGGL(k_tridiag_adj<T>, dim3((unsigned)(((size_t)nx * ny + 255) / 256)), dim3(256), 0, st, L, s.dz,
                     pw.piv, (const T *)pw.pivt, g);
  // Dx^-1 Dy^-1 into the interior of lam.p (its ghosts get values that are
  // not used and cleared with the rest of lam.p afterwards)
  if (!(fast && launch_dcty_fast<T, 1>(st, L, twy, cty, g)))
    hipLaunchKernelGGL((k_dct_y<T, 1>), dim3(gy), dim3(256), shy, st, L, py, npy, twy, cty, g);
  if (!(fast && launch_dctx_fast<T, 1>(st, L, twx, ctx, g, (T *)nullptr, zero, s, ph, lam.p, f.fx, f.w, f.wm,
                                       f.fz, aux.pgradx)))
    hipLaunchKernelGGL((k_dct_x<T, 1>), dim3(gx), dim3(256), shx, st, L, px, npx, twx, ctx, g, (T *)nullptr,
                       zero, s, ph, lam.p, f.fx, f.w, f.wm, f.fz, aux.pgradx);
  hipLaunchKernelGGL(k_padj_plane<T>, dim3(nz), dim3(256), 0, st, L, (const T *)lam.p, part);
  hipLaunchKernelGGL(k_sum1, dim3(1), dim3(SUM1_THREADS), 0, st, nz, part, part + nz);
  const unsigned nf = (unsigned)(((long)(nx + 1) * (ny + 1) + 255) / 256);
  hipLaunchKernelGGL(k_padj_rhs<T>, dim3(nf, 1, nz + 1), dim3(64, 4), 0, st, L, padj_consts(L, s, fs),
                     (const T *)lam.p, (const double *)(part + nz), lam.um, lam.vm, lam.wm, lam.w, lam.fx, lam.fy,
                     lam.fz, gflow);
  hipMemsetAsync(lam.p, 0, L.n * sizeof(T), st);
}

#define HC_INST_P(T)                                                                                          \
  template void launch_pressure<T>(hipStream_t, const Layout &, const StepArgs &, const Phys &,               \
                                   const FieldPtrs<T> &, const Aux &, PoissonWork &);                        \
  template void launch_pressure_adj<T>(hipStream_t, const Layout &, const StepArgs &, const Phys &,           \
                                       const FieldPtrs<T> &, const FieldPtrs<T> &, const Aux &, PoissonWork &,  \
                                       const FlowSens &, double *, T *, const double *, double *);
HC_INST_P(double)
HC_INST_P(float)

}  // namespace hc
