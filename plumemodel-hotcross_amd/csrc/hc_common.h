// hc_common.h - device layout, launch helpers and per-step scalars shared by
// the HotCross HIP kernels and the C-ABI implementation.
//
// HBM layout: every 3-D field (whatever its staggering) is stored in one
// padded box that covers i=-1..nx+1, j=-1..ny+1, k=-1..nz+1, x fastest.  The
// storage offset puts i=0 on a 16-element boundary (128 B for FP64) so that
// interior rows start cache-line aligned; the row pitch px is a multiple of 16
// elements.  A Fortran module array of any class (u: -1:nx+1, v: -1:ny+1,
// w: -1:nz+1, scalar) maps onto the same index function, so kernels that mix
// staggered fields share one address computation.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <math.h>
#include <stdint.h>

namespace hc {

constexpr int OXP = 16;  // storage index of i = 0 is OXP

struct Layout {
  int nx, ny, nz;
  int px, py, pz;  // padded extents
  size_t n;        // elements per field
};

// Element index of (i,j,k).  Computed in signed 32-bit arithmetic: a
// context refuses grids whose padded box holds 2^31 elements or more, so the
// index never overflows, and the compiler may fold a stencil's constant
// neighbour offsets into the loads.
__host__ __device__ inline int IX(const Layout &L, int i, int j, int k) {
  return (i + OXP) + L.px * ((j + 1) + L.py * (k + 1));
}

inline Layout make_layout(int nx, int ny, int nz) {
  Layout L;
  L.nx = nx; L.ny = ny; L.nz = nz;
  L.px = ((OXP + nx + 2) + 15) / 16 * 16;
  L.py = ny + 3;
  L.pz = nz + 3;
  L.n = (size_t)L.px * L.py * L.pz;
  return L;
}

// Scalars of one timestep that kernels take by value.
struct StepArgs {
  double dx, dy, dz, dt, tau3;
  double idx, idy, idz;  // 1/dx, 1/dy, 1/dz
  // correctly rounded reciprocals of the Prandtl number and rho0, in FP64
  // and FP32, for the divisions by them (UDiv)
  double rpr, rrho0;
  float rprf, rrho0f;
  // the same for dt (FP64, FP32) and for dx, dy, dz in FP32
  double rdt;
  float rdtf, idxf, idyf, idzf;
  int p2;                // dx, dy, dz are powers of two (exact reciprocals)
  int leap, delta, mtime;
  double ubkg_m, ubkg, vbkg_m, vbkg, fbx, fby;
  double fc, fc_y;
  double acc[8];  // persistent azu/azv boundary accumulators
};

// XCD-aware block order on/off for the box launches (off unless the
// environment sets HC_XCD_REMAP=1 when a context is created; on MI355X at
// 256^3 it measured 1.7% slower than the dispatcher's own order)
inline int &box_remap_flag() {
  static int flag = 0;
  return flag;
}

// Box of cells for the cell-parallel launches (box_grid, HC_BOX_IJK).
struct Box {
  int i0, i1, j0, j1, k0, k1;            // inclusive Fortran ranges
  int remap = box_remap_flag();          // XCD-aware block order (box_block_coords)
};

// XCD-aware block order.  The dispatcher deals workgroups to the 8 XCDs
// round-robin by linear id; remapping id -> logical block so that XCD x gets
// the x-th contiguous slice of the (i,j,k) block space keeps the j/k halo
// rows a block re-reads inside its own XCD's L2.  The map is a bijection on
// [0, nblocks), so every cell is still covered exactly once.
__device__ inline void box_block_coords(int remap, unsigned &bx, unsigned &by, unsigned &bz) {
  bx = blockIdx.x; by = blockIdx.y; bz = blockIdx.z;
  if (!remap) return;
  const unsigned gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
  const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned xcd = id % 8u, slot = id / 8u, per = n / 8u, rem = n % 8u;
  const unsigned lid = (xcd < rem) ? xcd * (per + 1u) + slot : rem * (per + 1u) + (xcd - rem) * per + slot;
  bx = lid % gx;
  by = (lid / gx) % gy;
  bz = lid / (gx * gy);
}

// Cell-parallel kernels over a box: 256 threads per block take 256
// consecutive cells of one (i,j) plane in memory order (i fastest, rows
// wrapping), so a padded row of nx+2 or nx+3 cells does not leave most of a
// fifth 64-wide block idle; blockIdx.z is the plane.
inline dim3 box_grid(const Box &b) {
  long ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
  int nk = b.k1 - b.k0 + 1;
  return dim3((unsigned)((ni * nj + 255) / 256), 1, nk);
}
inline dim3 box_block() { return dim3(64, 4, 1); }

#define HC_BOX_IJK(b)                                                 \
  unsigned bx_, by_, bz_;                                             \
  box_block_coords((b).remap, bx_, by_, bz_);                         \
  const int ni_ = (b).i1 - (b).i0 + 1;                                \
  const int t_ = (int)(bx_ * 256u + threadIdx.y * 64u + threadIdx.x); \
  const int jr_ = t_ / ni_;                                           \
  const int i = (b).i0 + (t_ - jr_ * ni_);                            \
  const int j = (b).j0 + jr_;                                         \
  const int k = (b).k0 + (int)bz_;                                    \
  if (j > (b).j1) return;

// Planes per block of a k-marching tile kernel: kmax, halved (not below 2)
// until the launch has at least min_blocks blocks, so that a small grid
// still fills the chip's 256 CUs (at 64^3 the corrective pass has 8 tiles a
// plane: 2 planes per block gave it 256 blocks and took a launch from 30 to
// 23 us, the step from 0.35 to 0.31 ms, though each block stages two
// planes beyond its own).
inline int tile_chunk(int nz, long tiles, int kmax, long min_blocks) {
  int kc = kmax;
  while (kc > 2 && tiles * ((nz + kc - 1) / kc) < min_blocks) kc /= 2;
  return kc < nz ? kc : nz;
}

constexpr double MISSING = -1.0e+34;

// A grid spacing as a divisor.  x / d for a power of two d equals x * (1/d)
// exactly (both are one correctly rounded operation on the same real
// number), so kernels instantiated with P2 = true multiply instead of
// dividing; the results are bit-identical to the division.
template <typename T, bool P2>
struct Sp {
  T d, inv;
  __host__ __device__ Sp(T d_, T inv_) : d(d_), inv(inv_) {}
};
template <typename T, bool P2>
__host__ __device__ inline T operator/(T x, const Sp<T, P2> &s) {
  return P2 ? x * s.inv : x / s.d;
}

// x / b for a divisor b that is the same over a launch: with r the correctly
// rounded 1/b (computed once on the host), q = x*r corrected once by the exact
// remainder x - b*q (Markstein) is the correctly rounded quotient, in three
// operations instead of a division's eleven.  Checked against IEEE division
// on gfx950: bit for bit on 5e8 random operand pairs in FP64 and in FP32,
// divisors over 2^-30..2^30 and the all-ones significand among them.
template <typename T>
struct UDiv {
  T b, r;
  __host__ __device__ UDiv(T b_, T r_) : b(b_), r(r_) {}
  __host__ __device__ T operator()(T x) const {
#ifdef __HIP_DEVICE_COMPILE__
    const T q = x * r;
    return fma(fma(-b, q, x), r, q);
#else
    return x / b;
#endif
  }
};
// the divisor b as type T with its reciprocal of that type
template <typename T>
__host__ __device__ inline UDiv<T> udiv_t(double b, double r64, float r32);
template <>
__host__ __device__ inline UDiv<double> udiv_t<double>(double b, double r64, float) { return UDiv<double>(b, r64); }
template <>
__host__ __device__ inline UDiv<float> udiv_t<float>(double b, double, float r32) { return UDiv<float>((float)b, r32); }

// a / b in FP64 for a data-dependent b: the hardware reciprocal, two Newton
// steps and one remainder correction, the arithmetic of the compiler's IEEE
// division without its operand scaling and special-case fix-up (8
// instructions instead of 11).  For a normal b away from the overflow range
// and a finite quotient it is the IEEE quotient (an in-range sweep of 8e8
// operand pairs on gfx950 matched bit for bit); b = 0, infinities and NaNs
// give NaN, where IEEE gives an infinity.
__host__ __device__ inline double ddiv(double a, double b) {
#ifdef __HIP_DEVICE_COMPILE__
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
#else
  return a / b;
#endif
}

// d > 0 is a power of two whose reciprocal is a normal number of type T
template <typename T>
inline bool pow2_spacing(double d) {
  T t = (T)d;
  if (!(t > (T)0)) return false;
  int e;
  double m = frexp((double)t, &e);
  T inv = (T)1 / t;
  return m == 0.5 && inv * t == (T)1 && inv > (T)0 && (double)inv < 1e30;
}

}  // namespace hc
