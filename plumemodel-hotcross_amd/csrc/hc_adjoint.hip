// hc_adjoint.hip - discrete adjoint of the HotCross timestep on gfx950.
//
// The adjoint of each reference subroutine is a set of kernels over the
// same boxes as the forward kernels.  Every adjoint entry "consumes" the
// adjoint of what its forward counterpart overwrites (sets it to zero) and
// adds to the adjoints of what it reads, so the adjoint of a sequence of
// subroutine calls is the sequence of adjoint calls in reverse order.
//
// Stencils scatter: a thread takes the adjoint of one output cell (or face)
// and adds lam * d(output)/d(input) to each input's adjoint with hardware
// FP64 atomic adds (per-cell reverse formulas in hc_adj.h).  The linear
// maps with regular structure (MOMENTUM, the pressure solve, REFRESH, the
// work velocities, DENSITY) are gathers without atomics.
//
// MPDATA's non-smooth parts follow the forward branch: the limiter flags
// of both corrective passes are the ones the forward kernels formed (taped
// with the passes' intermediates, or written by a forward recompute of the
// step) and are held fixed; |U| differentiates to sign(U).
//
// The window driver (hc_adj_window) runs the forward with a checkpoint
// every `seg` steps, evaluates a probe-misfit loss, then sweeps the segments
// backwards: each segment is recomputed from its checkpoint into a tape of
// its states in HBM (the tracers' pass intermediates taped beside them), and
// its steps are reversed one by one.  The S tracer's forward passes and its
// adjoint run on a side stream beside T's, and the tape copies of each state
// on the other side stream beside the next step's FORCES.
#include <stdlib.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "hc_adj.h"
#include "hc_ctx.h"

namespace hc {

#define F3(a, i, j, k) (a)[IX(L, (i), (j), (k))]

// Atomic adds of adjoint contributions into the adjoint fields (global
// memory).  They are issued as global (not flat) atomics: a flat atomic also
// counts against the LDS wait counter, so every block barrier (which waits
// for outstanding LDS operations) would wait for the device-scope atomics to
// complete; as global atomics they stay in flight across the barriers of the
// k-marching tiles.  In the FP32 form, contributions of magnitude below the
// smallest normal float (zeros and denormals) are dropped, below the form's
// resolution; FP64 adds every contribution.
__device__ inline bool adj_keep(double) { return true; }
__device__ inline bool adj_keep(float v) { return !(fabsf(v) < 1.17549435e-38f); }
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) float gfloat;
__device__ inline void adj_add(double *p, double v) {
  if (adj_keep(v)) __builtin_amdgcn_global_atomic_fadd_f64((gdouble *)p, v);
}
__device__ inline void adj_add(float *p, float v) {
  if (adj_keep(v)) __builtin_amdgcn_global_atomic_fadd_f32((gfloat *)p, v);
}
// LDS ring accumulation in FP64: every contribution is added, without a
// test (a branch per contribution cost more than the adds)
template <typename T>
__device__ inline void adj_ladd(double *p, T v) {
  atomicAdd(p, (double)v);
}

// accumulator: hardware FP64 / FP32 atomic add into a padded field
template <typename T>
struct AAcc {
  T *p;
  Layout L;
  __device__ void operator()(int i, int j, int k, T v) const { adj_add(p + IX(L, i, j, k), v); }
};
// views and accumulators relative to a face, for the MPDATA face formulas
template <typename T>
struct RelR {
  const T *p;
  Layout L;
  int i, j, k;
  __device__ T operator()(int a, int b, int c) const { return p[IX(L, i + a, j + b, k + c)]; }
};
template <typename T>
struct RelA {
  T *p;
  Layout L;
  int i, j, k;
  __device__ void operator()(int a, int b, int c, T v) const { adj_add(p + IX(L, i + a, j + b, k + c), v); }
};

// sum of one double per thread into one of the 256 slots at `slots` (one
// atomic per block, spread over the slots so that blocks do not queue on one
// address; the slots are summed on the host); every thread of the block must
// call it.  The closing barrier keeps a following call (which reuses `sh`)
// from overwriting the partial sums before thread 0 has read them.
__device__ inline void block_add(double v, double *slots) {
  const int tid = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
  const int nt = (int)(blockDim.x * blockDim.y * blockDim.z);
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __shared__ double sh[16];
  if (tid % 64 == 0) sh[tid / 64] = v;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < (nt + 63) / 64; ++w) s += sh[w];
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (s != 0.0) atomicAdd(slots + (b & (ADJ_SLOTS - 1)), s);
  }
  __syncthreads();
}

struct Box3 {
  int i0, i1, j0, j1, k0, k1;
};
// cell-parallel launches over a box (blocks of 64 x 4 threads): the 256
// threads of a block take 256 consecutive cells of one plane of the box, i
// fastest and rows wrapping, so rows of n+2 cells leave no block mostly
// idle; blockIdx.z is the plane
inline dim3 grid3(const Box3 &b) {
  const long ni = b.i1 - b.i0 + 1, nj = b.j1 - b.j0 + 1;
  return dim3((unsigned)((ni * nj + 255) / 256), 1, (unsigned)(b.k1 - b.k0 + 1));
}
#define ADJ_IJK(b)                                                         \
  const int ni_ = (b).i1 - (b).i0 + 1;                                     \
  const int t_ = (int)(blockIdx.x * 256u + threadIdx.y * 64u + threadIdx.x); \
  const int jr_ = t_ / ni_;                                                \
  const int i = (b).i0 + (t_ - jr_ * ni_);                                 \
  const int j = (b).j0 + jr_;                                              \
  const int k = (b).k0 + (int)blockIdx.z;                                  \
  const bool in = j <= (b).j1;

// ---------------------------------------------------------------------------
// k-marching LDS accumulation.  A block owns a TI x TJ column of output
// cells and marches down the planes ka..kb; the contributions of the
// outputs of plane k land in three LDS planes (k-1, k, k+1) of every adjoint
// field the formula feeds, over the tile plus a one-cell halo.  After plane
// k, plane k-1 receives nothing more from this block: it is added to HBM
// (one atomic per touched cell and field, instead of one per contribution)
// and its ring slot cleared for plane k+2.  Kernels whose formulas add to
// one field at one offset several times per cell sum those contributions
// in registers first (RegRel, RegAbs) and add each sum to the ring once.
// The rings accumulate in FP64 in both forms: the LDS FP64 atomic add is
// fast on gfx950, while the FP32 one halved the FP32 window's speed at 256^3
// (measured: the same window with the ring adds made non-atomic ran in
// 23.6 ms per step instead of 43.7; FP64 was unchanged).
template <typename T, int TI, int TJ, int NF>
struct AdjRing {
  static constexpr int SI = TI + 2, SJ = TJ + 2, S = SI * SJ, NT = TI * TJ;
  double a[NF][3][S];
};

template <typename T, int SI, int S>
struct LAcc {
  double *ring;  // [3][S] of one field
  int ib, jb;
  __device__ void operator()(int i, int j, int k, T v) const {
    adj_ladd(&ring[((unsigned)(k + 3) % 3u) * S + (i - ib) + SI * (j - jb)], v);  // k >= -3
  }
};
// the same, relative to a face (MPDATA face formulas)
template <typename T, int SI, int S>
struct LRel {
  double *ring;
  int ib, jb, i, j, k;
  __device__ void operator()(int a, int b, int c, T v) const {
    adj_ladd(&ring[((unsigned)(k + c + 3) % 3u) * S + (i + a - ib) + SI * (j + b - jb)], v);  // k + c >= -3
  }
};

// values of a field at offsets (-1..1)^3 around a cell (i,j,k), loaded
// into registers up front and read by absolute indices; the formulas call
// it with the cell's indices plus constants, which fold after inlining
// (entries never loaded are never read)
template <typename T>
struct NbAt {
  T v[27];
  int i, j, k;
  __device__ NbAt(int i_, int j_, int k_) : i(i_), j(j_), k(k_) {}
  __device__ T operator()(int a, int b, int c) const { return v[(a - i + 1) + 3 * (b - j + 1) + 9 * (c - k + 1)]; }
  template <int N>
  __device__ void load(const T *p, const Layout &L, const int (&o)[N][3]) {
#pragma unroll
    for (int q = 0; q < N; ++q)
      v[(o[q][0] + 1) + 3 * (o[q][1] + 1) + 9 * (o[q][2] + 1)] = p[IX(L, i + o[q][0], j + o[q][1], k + o[q][2])];
  }
};
// the cell and its six face neighbours, and with two edge neighbours more
#define HC_NB7 {0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}

// Accumulators in a thread's registers, 27 entries at the offsets (-1..1)^3
// around a cell, indexed by constants after inlining: the contributions a
// cell's formulas make to one field at one offset add up here and reach the
// LDS ring as one atomic add per offset (RegRel: offsets relative to the
// cell; RegAbs: absolute indices).  reg_flush adds the entries at the listed
// offsets to a ring accumulator.
template <typename T>
struct RegRel {
  T *v;
  __device__ void operator()(int a, int b, int c, T x) const { v[(a + 1) + 3 * (b + 1) + 9 * (c + 1)] += x; }
};
template <typename T>
struct RegAbs {
  T *v;
  int i, j, k;
  __device__ void operator()(int a, int b, int c, T x) const {
    v[(a - i + 1) + 3 * (b - j + 1) + 9 * (c - k + 1)] += x;
  }
};
template <typename T, int N, class A>
__device__ inline void reg_flush(const T *v, const int (&o)[N][3], const A &ring) {
#pragma unroll
  for (int q = 0; q < N; ++q) ring(o[q][0], o[q][1], o[q][2], v[(o[q][0] + 1) + 3 * (o[q][1] + 1) + 9 * (o[q][2] + 1)]);
}
template <typename T, int N, class A>
__device__ inline void reg_flush_at(const T *v, const int (&o)[N][3], const A &ring, int i, int j, int k) {
#pragma unroll
  for (int q = 0; q < N; ++q)
    ring(i + o[q][0], j + o[q][1], k + o[q][2], v[(o[q][0] + 1) + 3 * (o[q][1] + 1) + 9 * (o[q][2] + 1)]);
}

template <typename T, int TI, int TJ, int NF, class Body>
__device__ inline void adj_march(AdjRing<T, TI, TJ, NF> &sh, const Layout &L, T *const *dst, int i0, int j0, int ka,
                                 int kb, Body &&body) {
  typedef AdjRing<T, TI, TJ, NF> R;
  constexpr int SI = R::SI, S = R::S, NT = R::NT;
  const int tid = (int)threadIdx.x;
  double *flat = &sh.a[0][0][0];
  for (int e = tid; e < NF * 3 * S; e += NT) flat[e] = 0.0;
  __syncthreads();
  const int i = i0 + tid % TI, j = j0 + tid / TI;
  const bool own = i <= L.nx && j <= L.ny;
  auto flush = [&](int k) {
    const int slot = k % 3;
    for (int e = tid; e < NF * S; e += NT) {
      const int f = e / S, r = e - f * S;
      double *cell = &sh.a[f][slot][r];
      const T v = (T)*cell;
      const int ii = i0 - 1 + r % SI, jj = j0 - 1 + r / SI;
      if (v != (T)0 && ii <= L.nx + 1 && jj <= L.ny + 1) adj_add(dst[f] + IX(L, ii, jj, k), v);
      *cell = 0.0;
    }
  };
  for (int k = ka; k <= kb; ++k) {
    if (own) body(i, j, k);
    __syncthreads();
    flush(k - 1);
    __syncthreads();
  }
  flush(kb);
  flush(kb + 1);
}

// The same march with a ring depth per field: field f takes contributions
// at planes k + lo .. k + hi of the plane k being processed (lo = -1 or 0,
// hi = 0 or 1; NP::lo[f], NP::hi[f]), so its ring holds hi - lo + 1 planes.
// After plane k its plane k + lo is complete: it is added to HBM in that
// plane's flush and its slot cleared; after the last plane kb the planes
// kb + lo + 1 .. kb + hi follow.  Shallower rings mean less LDS per block and
// more blocks per CU.
template <typename T, int TI, int TJ, int NF, class NP>
struct AdjRingV {
  static constexpr int SI = TI + 2, SJ = TJ + 2, S = SI * SJ, NT = TI * TJ;
  static constexpr int depth(int f) { return NP::hi[f] - NP::lo[f] + 1; }
  static constexpr int base(int f) {
    int o = 0;
    for (int q = 0; q < f; ++q) o += depth(q);
    return o * S;
  }
  static constexpr int TOT = base(NF);
  // the same for a field index known only at run time: a chain of selects
  // between constants (no loop, no table in memory)
  struct Of {
    int np, lo, base;
  };
  static __device__ Of of(int f) {
    Of o{depth(0), NP::lo[0], 0};
#pragma unroll
    for (int q = 1; q < NF; ++q)
      if (f == q) o = Of{depth(q), NP::lo[q], base(q)};
    return o;
  }
  double a[TOT];
};
template <typename T, int SI, int S, int NPL>
struct LAccV {
  double *ring;  // [NPL][S] of one field, plane p in slot (p + 3) mod NPL
  int ib, jb;
  __device__ void operator()(int i, int j, int k, T v) const {
    adj_ladd(&ring[((unsigned)(k + 3) % (unsigned)NPL) * S + (i - ib) + SI * (j - jb)], v);  // k >= -3
  }
};
template <typename T, int TI, int TJ, int NF, class NP, class Body>
__device__ inline void adj_march_v(AdjRingV<T, TI, TJ, NF, NP> &sh, const Layout &L, T *const *dst, int i0, int j0,
                                   int ka, int kb, Body &&body) {
  typedef AdjRingV<T, TI, TJ, NF, NP> R;
  constexpr int SI = R::SI, S = R::S, NT = R::NT;
  const int tid = (int)threadIdx.x;
  for (int e = tid; e < R::TOT; e += NT) sh.a[e] = 0.0;
  __syncthreads();
  const int i = i0 + tid % TI, j = j0 + tid / TI;
  const bool own = i <= L.nx && j <= L.ny;
  // plane k + lo of every field (d = 0), or after the last plane the
  // plane kb + lo + d of the fields with hi - lo >= d (d = 1, 2)
  auto flush = [&](int k, int d) {
    for (int e = tid; e < NF * S; e += NT) {
      const int f = e / S, r = e - f * S;
      const typename R::Of o = R::of(f);
      if (d >= o.np) continue;
      const int p = k + o.lo + d;
      double *cell = &sh.a[o.base + (int)((unsigned)(p + 3) % (unsigned)o.np) * S + r];
      const T v = (T)*cell;
      const int ii = i0 - 1 + r % SI, jj = j0 - 1 + r / SI;
      if (v != (T)0 && ii <= L.nx + 1 && jj <= L.ny + 1) adj_add(dst[f] + IX(L, ii, jj, p), v);
      *cell = 0.0;
    }
  };
  for (int k = ka; k <= kb; ++k) {
    if (own) body(i, j, k);
    __syncthreads();
    flush(k, 0);
    __syncthreads();
  }
  flush(kb, 1);
  flush(kb, 2);
}

template <int TI, int TJ>
inline dim3 adj_tile_grid(const Layout &L, int kc) {
  return dim3((unsigned)((L.nx + TI - 1) / TI), (unsigned)((L.ny + TJ - 1) / TJ), (unsigned)((L.nz + kc - 1) / kc));
}
constexpr int ATI = 32, ATJ = 8, AKC = 32;
// planes per block of the adjoint tiles (HC_ADJ_MINB overrides the least
// number of blocks a launch should have)
static int adj_chunk(const Layout &L) {
  static long minb = [] {
    const char *e = getenv("HC_ADJ_MINB");
    return e ? atol(e) : 1024L;
  }();
  return tile_chunk(L.nz, (long)((L.nx + ATI - 1) / ATI) * ((L.ny + ATJ - 1) / ATJ), AKC, minb);
}

// ---------------------------------------------------------------------------
// FORCES
template <typename T>
__global__ void __launch_bounds__(256) k_adj_forces(Layout L, Box3 b, StepArgs s, Phys ph, FlowSens fs,
                                                    FieldPtrs<T> f, const T *__restrict__ ss0,
                                                    const T *__restrict__ ssv0, FieldPtrs<T> lam,
                                                    const double *__restrict__ azzmin,
                                                    const double *__restrict__ alphau,
                                                    const double *__restrict__ alphav,
                                                    const double *__restrict__ alphae0,
                                                    const double *__restrict__ alphaw, double *gflow) {
  ADJ_IJK(b);
  double gf = 0.0;
  if (in) {
    const RView<T> U{f.u, L}, V{f.v, L}, W{f.w, L}, UM{f.um, L}, VM{f.vm, L}, WM{f.wm, L}, SS{ss0, L}, SV{ssv0, L},
        RH{f.rhop, L};
    const AAcc<T> au{lam.u, L}, av{lam.v, L}, aw{lam.w, L}, aum{lam.um, L}, avm{lam.vm, L}, awm{lam.wm, L},
        ass{lam.ss, L}, assv{lam.ssv, L}, arh{lam.rhop, L};
    // each cell's output adjoints are consumed here (read, then cleared:
    // FORCES overwrites fx, fy, fz)
    if (j >= 1 && k >= 1) {
      const T l = F3(lam.fx, i, j, k);
      F3(lam.fx, i, j, k) = (T)0;
      if (l != (T)0)
        adj_forces_x<T>(L, i, j, k, s, ph, fs, U, V, W, UM, VM, WM, SS, SV, azzmin, alphau, l, au, av, aw, aum, avm,
                        awm, ass, assv, gf);
    }
    if (i >= 1 && k >= 1) {
      const T l = F3(lam.fy, i, j, k);
      F3(lam.fy, i, j, k) = (T)0;
      if (l != (T)0)
        adj_forces_y<T>(L, i, j, k, s, ph, fs, U, V, W, UM, VM, WM, SS, SV, azzmin, alphav, l, au, av, aw, aum, avm,
                        awm, ass, assv, gf);
    }
    if (i >= 1 && j >= 1) {
      const T l = F3(lam.fz, i, j, k);
      F3(lam.fz, i, j, k) = (T)0;
      if (l != (T)0)
        adj_forces_z<T>(L, i, j, k, s, ph, U, V, W, UM, VM, WM, SS, SV, RH, azzmin, alphae0, alphaw, l, au, av, aw,
                        aum, avm, awm, ass, assv, arh);
    }
  }
  block_add(gf, gflow);
}


// an accumulator that discards its values (the terms that only feed it
// drop out of the formulas)
struct NullAcc {
  template <typename T>
  __device__ void operator()(int, int, int, T) const {}
};

// FORCES, interior cells i, j, k >= 1 of all three components as a tile,
// in three launches of three LDS rings each (24.6 KB per block instead of
// 73.6 KB for all nine): PART 1 the advection terms (rings of u, v, w), 2
// and 4 the diffusion, buoyancy and sponge terms into um, vm, wm and into
// ss, ssv, rhop; the last launch clears the output adjoints all three read.
template <typename T, int PART>
__global__ void __launch_bounds__(ATI *ATJ) k_adj_forces_tile(Layout L, StepArgs s, Phys ph, FlowSens fs,
                                                              FieldPtrs<T> f, const T *__restrict__ ss0,
                                                              const T *__restrict__ ssv0, FieldPtrs<T> lam,
                                                              const double *__restrict__ azzmin,
                                                              const double *__restrict__ alphau,
                                                              const double *__restrict__ alphav,
                                                              const double *__restrict__ alphae0,
                                                              const double *__restrict__ alphaw, double *gflow, int kc) {
  static_assert(PART == 1 || PART == 2 || PART == 4, "one part per launch");
  typedef AdjRing<T, ATI, ATJ, 3> R;
  __shared__ R sh;
  const int i0 = 1 + (int)blockIdx.x * ATI, j0 = 1 + (int)blockIdx.y * ATJ;
  const int ka = 1 + (int)blockIdx.z * kc, kb = min(L.nz, ka + kc - 1);
  T *const dst1[3] = {lam.u, lam.v, lam.w};
  T *const dst2[3] = {lam.um, lam.vm, lam.wm};
  T *const dst4[3] = {lam.ss, lam.ssv, lam.rhop};
  T *const *dst = PART == 1 ? dst1 : PART == 2 ? dst2 : dst4;
  const RView<T> RH{f.rhop, L};
  typedef LAcc<T, R::SI, R::S> A;
  const A a0{sh.a[0][0], i0 - 1, j0 - 1}, a1{sh.a[1][0], i0 - 1, j0 - 1}, a2{sh.a[2][0], i0 - 1, j0 - 1};
  const NullAcc nul;
  double gf = 0.0;
  adj_march(sh, L, dst, i0, j0, ka, kb, [&](int i, int j, int k) {
    const T lx = F3(lam.fx, i, j, k), ly = F3(lam.fy, i, j, k), lz = F3(lam.fz, i, j, k);
    if (PART == 4) {
      F3(lam.fx, i, j, k) = (T)0;  // consumed (FORCES overwrites them)
      F3(lam.fy, i, j, k) = (T)0;
      F3(lam.fz, i, j, k) = (T)0;
    }
    // the inputs the part reads, loaded before the zero tests of the three
    // components so that a cell's loads are in flight together
    NbAt<T> u(i, j, k), v(i, j, k), w(i, j, k), um(i, j, k), vm(i, j, k), wm(i, j, k), ss(i, j, k), sv(i, j, k);
    if constexpr (PART == 1) {
      static constexpr int ou[9][3] = {HC_NB7, {-1, 1, 0}, {-1, 0, 1}};
      static constexpr int ov[9][3] = {HC_NB7, {1, -1, 0}, {0, -1, 1}};
      static constexpr int ow[9][3] = {HC_NB7, {1, 0, -1}, {0, 1, -1}};
      u.load(f.u, L, ou);
      v.load(f.v, L, ov);
      w.load(f.w, L, ow);
      if (lx != (T)0)
        adj_forces_x<T, 1>(L, i, j, k, s, ph, fs, u, v, w, um, vm, wm, ss, sv, azzmin, alphau, lx, a0, a1, a2, nul,
                           nul, nul, nul, nul, gf);
      if (ly != (T)0)
        adj_forces_y<T, 1>(L, i, j, k, s, ph, fs, u, v, w, um, vm, wm, ss, sv, azzmin, alphav, ly, a0, a1, a2, nul,
                           nul, nul, nul, nul, gf);
      if (lz != (T)0)
        adj_forces_z<T, 1>(L, i, j, k, s, ph, u, v, w, um, vm, wm, ss, sv, RH, azzmin, alphae0, alphaw, lz, a0, a1, a2,
                           nul, nul, nul, nul, nul, nul);
    } else if constexpr (PART == 2) {  // the coefficients: ss and ssv
      static constexpr int oss[9][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                        {1, 1, 0}, {1, -1, 0}, {-1, 1, 0}, {0, 0, 1}};
      static constexpr int osv[13][3] = {HC_NB7,     {1, 0, 1},  {1, 0, -1}, {0, 1, 1},
                                         {0, 1, -1}, {-1, 0, 1}, {0, -1, 1}};
      ss.load(ss0, L, oss);
      sv.load(ssv0, L, osv);
      if (lx != (T)0)
        adj_forces_x<T, 2>(L, i, j, k, s, ph, fs, u, v, w, um, vm, wm, ss, sv, azzmin, alphau, lx, nul, nul, nul, a0,
                           a1, a2, nul, nul, gf);
      if (ly != (T)0)
        adj_forces_y<T, 2>(L, i, j, k, s, ph, fs, u, v, w, um, vm, wm, ss, sv, azzmin, alphav, ly, nul, nul, nul, a0,
                           a1, a2, nul, nul, gf);
      if (lz != (T)0)
        adj_forces_z<T, 2>(L, i, j, k, s, ph, u, v, w, um, vm, wm, ss, sv, RH, azzmin, alphae0, alphaw, lz, nul, nul,
                           nul, a0, a1, a2, nul, nul, nul);
    } else {  // the differences of um, vm, wm
      static constexpr int oum[9][3] = {HC_NB7, {-1, 1, 0}, {-1, 0, 1}};
      static constexpr int ovm[9][3] = {HC_NB7, {1, -1, 0}, {0, -1, 1}};
      static constexpr int owm[9][3] = {HC_NB7, {1, 0, -1}, {0, 1, -1}};
      um.load(f.um, L, oum);
      vm.load(f.vm, L, ovm);
      wm.load(f.wm, L, owm);
      if (lx != (T)0)
        adj_forces_x<T, 4>(L, i, j, k, s, ph, fs, u, v, w, um, vm, wm, ss, sv, azzmin, alphau, lx, nul, nul, nul, nul,
                           nul, nul, a0, a1, gf);
      if (ly != (T)0)
        adj_forces_y<T, 4>(L, i, j, k, s, ph, fs, u, v, w, um, vm, wm, ss, sv, azzmin, alphav, ly, nul, nul, nul, nul,
                           nul, nul, a0, a1, gf);
      if (lz != (T)0)
        adj_forces_z<T, 4>(L, i, j, k, s, ph, u, v, w, um, vm, wm, ss, sv, RH, azzmin, alphae0, alphaw, lz, nul, nul,
                           nul, nul, nul, nul, a0, a1, a2);
    }
  });
  block_add(gf, gflow);
}

// ---------------------------------------------------------------------------
// MOMENTUM_U/V/W: Lambda = the adjoint of the interior formula at a cell,
// collected from the cell and from the ghost cells that copy it; then the
// formula's (linear) adjoint
template <typename T>
__global__ void __launch_bounds__(256) k_adj_mom_lambda(Layout L, Box3 b, int comp, StepArgs s, Phys ph,
                                                        const T *__restrict__ lup, T *__restrict__ lamb,
                                                        T *__restrict__ lu, T *__restrict__ lum, T *__restrict__ lf) {
  ADJ_IJK(b);
  if (!in) return;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  T a = F3(lup, i, j, k);
  if (comp == 0) {
    if (k == 1) a += F3(lup, i, j, 0);
    if (k == nz) a += (T)ph.bottom_slip * F3(lup, i, j, nz + 1);
    if (j == 1) a += F3(lup, i, 0, k);
    if (j == ny) a += F3(lup, i, ny + 1, k);
    if (i == 0) a += F3(lup, -1, j, k);
    if (i == nx) a += F3(lup, nx + 1, j, k);
  } else if (comp == 1) {
    if (k == 1) a += F3(lup, i, j, 0);
    if (k == nz) a += (T)ph.bottom_slip * F3(lup, i, j, nz + 1);
    if (i == 1) a += F3(lup, 0, j, k);
    if (i == nx) a += F3(lup, nx + 1, j, k);
    if (j == 0) a += F3(lup, i, -1, k);
    if (j == ny) a += F3(lup, i, ny + 1, k);
  }
  F3(lamb, i, j, k) = a;
  F3(lu, i, j, k) += (T)(1 - s.delta) * a;
  F3(lum, i, j, k) += (T)s.delta * a;
  F3(lf, i, j, k) += (T)(s.dt * s.leap) * a;
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_mom_p(Layout L, Box3 b, Box3 fb, int comp, T c,
                                                   const T *__restrict__ lamb, T *__restrict__ lp) {
  ADJ_IJK(b);
  if (!in) return;
  // Lambda exists on the formula's box fb (k_adj_mom_lambda writes it
  // there); it is zero elsewhere
  auto la = [&](int a, int d, int e) -> T {
    return (a >= fb.i0 && a <= fb.i1 && d >= fb.j0 && d <= fb.j1 && e >= fb.k0 && e <= fb.k1) ? F3(lamb, a, d, e) : (T)0;
  };
  const T prev = comp == 0 ? la(i - 1, j, k) : (comp == 1 ? la(i, j - 1, k) : la(i, j, k - 1));
  F3(lp, i, j, k) += c * (la(i, j, k) - prev);
}

// ---------------------------------------------------------------------------
// VISCOSITY3D
// VISCOSITY3D: interior cells as a tile, the ghost shell cell by cell
template <typename T>
__global__ void __launch_bounds__(ATI *ATJ) k_adj_visc_tile(Layout L, StepArgs s, Phys ph, const T *__restrict__ um,
                                                            const T *__restrict__ vm, const T *__restrict__ wm,
                                                            const T *__restrict__ rhop,
                                                            const double *__restrict__ rho_bkg,
                                                            const T *__restrict__ ss1, const T *__restrict__ ssv1,
                                                            T *__restrict__ lss, T *__restrict__ lssv,
                                                            T *lum, T *lvm, T *lwm, T *lrhop, double *gcs, int kc) {
  typedef AdjRing<T, ATI, ATJ, 4> R;
  __shared__ R sh;
  const int i0 = 1 + (int)blockIdx.x * ATI, j0 = 1 + (int)blockIdx.y * ATJ;
  const int ka = 1 + (int)blockIdx.z * kc, kb = min(L.nz, ka + kc - 1);
  T *const dst[4] = {lum, lvm, lwm, lrhop};
  const RView<T> UM{um, L}, VM{vm, L}, WM{wm, L}, RH{rhop, L};
  typedef LAcc<T, R::SI, R::S> A;
  const A aum{sh.a[0][0], i0 - 1, j0 - 1}, avm{sh.a[1][0], i0 - 1, j0 - 1}, awm{sh.a[2][0], i0 - 1, j0 - 1},
      arh{sh.a[3][0], i0 - 1, j0 - 1};
  double sh_ = 0.0, sv_ = 0.0;
  adj_march(sh, L, dst, i0, j0, ka, kb, [&](int i, int j, int k) {
    const T a = F3(lss, i, j, k), c = F3(lssv, i, j, k);
    F3(lss, i, j, k) = (T)0;  // consumed: VISCOSITY3D overwrites both
    F3(lssv, i, j, k) = (T)0;
    adj_ss_val<T>(L, ph, s, UM, VM, WM, i, j, k, a, aum, avm, awm);
    adj_ssv_val<T>(L, ph, s, UM, VM, WM, RH, rho_bkg, i, j, k, c, aum, avm, awm, arh);
    sh_ += (double)a * (double)F3(ss1, i, j, k);
    sv_ += (double)c * (double)F3(ssv1, i, j, k);
  });
  block_add(sh_, gcs);
  block_add(sv_, gcs + ADJ_SLOTS);
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_visc_shell(Layout L, StepArgs s, Phys ph, const T *__restrict__ um,
                                                        const T *__restrict__ vm, const T *__restrict__ wm,
                                                        const T *__restrict__ rhop, const double *__restrict__ rho_bkg,
                                                        const T *__restrict__ ss1, const T *__restrict__ ssv1,
                                                        T *__restrict__ lss, T *__restrict__ lssv,
                                                        T *__restrict__ lum, T *__restrict__ lvm, T *__restrict__ lwm,
                                                        T *__restrict__ lrhop, double *gcs) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long na = 2L * (ny + 2) * (nz + 2), nb = 2L * nx * (nz + 2), nc = 2L * nx * ny;
  int i = 0, j = 0, k = 0;
  bool in = true;
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
    in = false;
  }
  double sh = 0.0, sv = 0.0;
  if (in) {
    const bool wc = visc_ssv_written(L, i, j, k);
    const T a = visc_ss_written(L, i, j, k) ? F3(lss, i, j, k) : (T)0;
    const T c = wc ? F3(lssv, i, j, k) : (T)0;
    // consumed: VISCOSITY3D overwrites ss on the whole shell (constants on
    // the edges and the k ghosts) and ss_v on the interior columns
    F3(lss, i, j, k) = (T)0;
    if (wc) F3(lssv, i, j, k) = (T)0;
    if (a != (T)0 || c != (T)0) {
      const RView<T> UM{um, L}, VM{vm, L}, WM{wm, L}, RH{rhop, L};
      const AAcc<T> aum{lum, L}, avm{lvm, L}, awm{lwm, L}, arh{lrhop, L};
      adj_visc_at<T>(L, s, ph, UM, VM, WM, RH, rho_bkg, a, c, i, j, k, aum, avm, awm, arh);
      sh = (double)a * (double)F3(ss1, i, j, k);
      sv = (double)c * (double)F3(ssv1, i, j, k);
    }
  }
  block_add(sh, gcs);
  block_add(sv, gcs + ADJ_SLOTS);
}

// ---------------------------------------------------------------------------
// up to eight whole fields set to zero in one launch (16-byte stores; a
// padded field is a multiple of 64 bytes); a kernel, not a memset, so that
// it is also safe inside a captured graph
struct ZeroList {
  uint4 *p[8];
};
__global__ void __launch_bounds__(256) k_zero_fields(ZeroList z, size_t n16) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n16) z.p[blockIdx.y][e] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------------------
// DENSITY, REFRESH, work velocities (elementwise)
template <typename T>
__global__ void __launch_bounds__(256) k_adj_density(Layout L, Box3 b, Phys ph, const T *__restrict__ t,
                                                     const T *__restrict__ sal, const T *__restrict__ lrhop,
                                                     T *__restrict__ lt, T *__restrict__ ls) {
  ADJ_IJK(b);
  if (!in) return;
  const int ng = (int)(i == 0 || i == L.nx + 1) + (int)(j == 0 || j == L.ny + 1) + (int)(k == 0 || k == L.nz + 1);
  if (ng > 1) return;
  const T lr = F3(lrhop, i, j, k);
  if (lr == (T)0) return;
  double dt_, ds_;
  drho(ph, (double)F3(sal, i, j, k), (double)F3(t, i, j, k), dt_, ds_);
  F3(lt, i, j, k) += (T)((double)lr * dt_);
  F3(ls, i, j, k) += (T)((double)lr * ds_);
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_refresh(size_t n, int leap, double assel, T *__restrict__ lu,
                                                     T *__restrict__ lup, T *__restrict__ lum) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const T a = lu[q], m = lum[q];
  if (leap == 1) {
    lup[q] += a + m;
    lu[q] = (T)0;
    lum[q] = (T)0;
  } else {
    lup[q] += a + (T)assel * m;
    lu[q] = ((T)1 - (T)2 * (T)assel) * m;
    lum[q] = (T)assel * m;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_adj_courant(Layout L, Box3 b, double tau3, double dx, double dy, double dz,
                                                     T *__restrict__ luc, T *__restrict__ lvc, T *__restrict__ lwc,
                                                     T *__restrict__ lu, T *__restrict__ lup, T *__restrict__ lv,
                                                     T *__restrict__ lvp, T *__restrict__ lw, T *__restrict__ lwp) {
  ADJ_IJK(b);
  if (!in) return;
  const size_t q = IX(L, i, j, k);
  if (i <= L.nx) {
    const T c = (T)(0.5 * tau3 / dx) * luc[q];
    lup[q] += c; lu[q] += c; luc[q] = (T)0;
  }
  if (j <= L.ny) {
    const T c = (T)(0.5 * tau3 / dy) * lvc[q];
    lvp[q] += c; lv[q] += c; lvc[q] = (T)0;
  }
  if (k <= L.nz) {
    const T c = -(T)(0.5 * tau3 / dz) * lwc[q];
    lwp[q] += c; lw[q] += c; lwc[q] = (T)0;
  }
}

// ---------------------------------------------------------------------------
// TRACER31 / MPDATA

// BOUNDARY_CONDITIONS transposed: a ghost's adjoint moves to its interior
// image.  mode 0: in place (the ghost's adjoint is consumed); mode 1: from
// the tracer's output into the pre-sponge accumulator, skipping the vent
// ghost cells whose value the driver sets to a constant.
template <typename T>
__global__ void __launch_bounds__(256) k_adj_bc(Layout L, Phys ph, VentArgs va, int mode, int vent,
                                                T *__restrict__ lsrc, T *__restrict__ ldst) {
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
  const bool gi = (i == 0 || i == nx + 1), gj = (j == 0 || j == ny + 1);
  if (mode == 1 && vent && k == nz + 1 && !gi && !gj && va.map_valid && i >= va.nxs && i <= va.nxe &&
      j >= va.nys && j <= va.nye)
    return;
  int si, sj, sk;
  bc_source(L, ph, i, j, k, si, sj, sk);
  const T v = F3(lsrc, i, j, k);
  if (mode == 0) F3(lsrc, i, j, k) = (T)0;
  if (v != (T)0) adj_add(&F3(ldst, si, sj, sk), v);
}

struct DonorLast {
  int last;
  // k_adj_donor_march: the outputs it is the first to write, which it
  // assigns (over its whole box) instead of adding to: 1 the entry
  // concentration's adjoint of the pass (lcc), 2 the face velocities' (lud,
  // lvd, lwd), 4 the tracer's entry value (lentry, sponge term)
  int assign;
  double tau3, dxdydz;
  const double *alphae0, *alphaw, *qbase;
  double *gq;
  int map_valid, nxs, nxe, nys, nye;
};

// the antidiffusive velocities reversed on the domain's first faces, in one
// launch: ud at i = 0 (j, k >= 1), vd at j = 0 (i, k >= 1), wd at k = 0 (i,
// j >= 1), one thread per face (inside the tile kernel, these faces raised
// its registers and took longer than this launch)
template <typename T>
__global__ void __launch_bounds__(256) k_adj_anti_faces(Layout L, const T *__restrict__ cc, const T *__restrict__ U,
                                                        const T *__restrict__ V, const T *__restrict__ W,
                                                        const T *__restrict__ lud, const T *__restrict__ lvd,
                                                        const T *__restrict__ lwd, T *__restrict__ lcc,
                                                        T *__restrict__ lU, T *__restrict__ lV, T *__restrict__ lW) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nu = (long)L.ny * L.nz, nv = (long)L.nx * L.nz, nw = (long)L.nx * L.ny;
  int i, j, k;
  if (t < nu) {
    i = 0; j = 1 + (int)(t % L.ny); k = 1 + (int)(t / L.ny);
  } else if ((t -= nu) < nv) {
    i = 1 + (int)(t % L.nx); j = 0; k = 1 + (int)(t / L.nx);
  } else if ((t -= nv) < nw) {
    i = 1 + (int)(t % L.nx); j = 1 + (int)(t / L.nx); k = 0;
  } else {
    return;
  }
  const RelR<T> C{cc, L, i, j, k}, Uf{U, L, i, j, k}, Vf{V, L, i, j, k}, Wf{W, L, i, j, k};
  const RelA<T> ac{lcc, L, i, j, k}, aU{lU, L, i, j, k}, aV{lV, L, i, j, k}, aW{lW, L, i, j, k};
  if (i == 0) adj_anti_u<T>(C, Uf, Vf, Wf, F3(lud, i, j, k), ac, aU, aV, aW);
  else if (j == 0) adj_anti_v<T>(C, Uf, Vf, Wf, F3(lvd, i, j, k), ac, aU, aV, aW);
  else adj_anti_w<T>(C, Uf, Vf, Wf, F3(lwd, i, j, k), ac, aU, aV, aW);
}

// The limited donor-cell step of a pass reversed as a gather over the box
// 0..n+1 (no atomics).  In flux form the step is cn = cc + sum of signed
// face fluxes: an x (or y) face between L = i and R = i+1 carries
// f = p cc(L) + m cc(R) with p = (u+|u|)/2, m = (u-|u|)/2 times the
// limiter's masks, taken from L and given to R; a z face carries
// g = wm cc(L) + wp cc(R), given to L and taken from R (mp_donor).  So each
// cell collects lam_f * (d flux / d cc) from its six faces, with
// lam_f = lam_n(R) - lam_n(L) (x, y) or lam_n(L) - lam_n(R) (z), and each
// face's velocity adjoint is lam_f * d flux / d u, written by the face's
// L cell.  lam_n is the adjoint of the pass's result, zero at ghost cells.
//
// A block takes 256 consecutive columns (i,j) of the box (i fastest, rows
// of nx+2 wrapping) and marches down kc of its planes; each thread carries
// its column's values at k-1, k and k+1 (lam_n, the flags, cc, the z face
// velocity below) in registers from one plane to the next, so the z
// neighbours are loaded once, not three times over (one plane per block
// fetched 3.6x the algorithmic bytes into L2 at 256^3 in FP32).
template <typename T>
__global__ void __launch_bounds__(256) k_adj_donor_march(Layout L, DonorLast dl, int kc, const T *__restrict__ lout,
                                                         const T *__restrict__ lpre, const T *__restrict__ cc,
                                                         const T *__restrict__ ud, const T *__restrict__ vd,
                                                         const T *__restrict__ wd, const uint8_t *__restrict__ fl,
                                                         T *__restrict__ lcc, T *__restrict__ lud,
                                                         T *__restrict__ lvd, T *__restrict__ lwd,
                                                         T *__restrict__ lentry) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const int t = (int)(blockIdx.x * 256u + threadIdx.x);
  const int j = t / (nx + 2), i = t - j * (nx + 2);
  if (j > ny + 1) return;
  const int ka = (int)blockIdx.y * kc, kb = min(nz + 1, ka + kc - 1);
  auto inner = [&](int a, int c, int d) { return a >= 1 && a <= nx && c >= 1 && c <= ny && d >= 1 && d <= nz; };
  auto lnv = [&](int a, int c, int d) -> T {
    if (!inner(a, c, d)) return (T)0;
    return F3(lout, a, c, d) + (lpre ? F3(lpre, a, c, d) : (T)0);
  };
  // split velocity and its derivative with a mask
  auto pos = [](T u, bool on, T &v, T &dv) {
    v = on ? (T)0.5 * (u + mp_fabs(u)) : (T)0;
    dv = on ? (T)0.5 * ((T)1 + adj_sign(u)) : (T)0;
  };
  auto neg = [](T u, bool on, T &v, T &dv) {
    v = on ? (T)0.5 * (u - mp_fabs(u)) : (T)0;
    dv = on ? (T)0.5 * ((T)1 - adj_sign(u)) : (T)0;
  };
  // masks: A = !(fL == 1 || fR == 2), B = !(fR == 1 || fL == 2)
  auto mA = [](int fL, int fR) { return !(fL == 1 || fR == 2); };
  auto mB = [](int fL, int fR) { return !(fR == 1 || fL == 2); };
  const bool col = i >= 1 && i <= nx && j >= 1 && j <= ny;  // an interior column: it has z faces
  const bool set_c = dl.assign & 1, set_u = dl.assign & 2;
  // the column at k-1 and k (the index k = -1 lies inside the padded box)
  T lm = lnv(i, j, ka - 1), lc = lnv(i, j, ka);
  int fm = ka >= 1 ? F3(fl, i, j, ka - 1) : 0, f0 = F3(fl, i, j, ka);
  T c0 = F3(cc, i, j, ka);
  T wm1 = (col && ka >= 1) ? F3(wd, i, j, ka - 1) : (T)0;
  // lam_n unmasked (the ghost cells of lout and lpre hold finite values)
  auto raw = [&](int a, int c, int d) -> T { return F3(lout, a, c, d) + (lpre ? F3(lpre, a, c, d) : (T)0); };
  for (int k = ka; k <= kb; ++k) {
    // every value the cell may use is loaded first, at neighbour indices
    // clamped into the box (0..n+1, so inside the padded allocation), and
    // the faces that do not exist are masked afterwards: the loads of a
    // plane are all in flight together instead of behind the face tests
    const bool up = k <= nz;  // the column at k+1 and the z face k are in the box
    const int kp = up ? k + 1 : k;
    const int ip = min(i + 1, nx + 1), im = max(i - 1, 0), jp = min(j + 1, ny + 1), jm = max(j - 1, 0);
    const T rxp = raw(ip, j, k), rxm = raw(im, j, k), ryp = raw(i, jp, k), rym = raw(i, jm, k), rzp = raw(i, j, kp);
    const int fxp = F3(fl, ip, j, k), fxm = F3(fl, im, j, k), fyp = F3(fl, i, jp, k), fym = F3(fl, i, jm, k);
    const int fzp = F3(fl, i, j, kp);
    const T u_i = F3(ud, i, j, k), u_im = F3(ud, im, j, k), v_j = F3(vd, i, j, k), v_jm = F3(vd, i, jm, k);
    const T wz = F3(wd, i, j, k);
    const T cxp = F3(cc, ip, j, k), cyp = F3(cc, i, jp, k), czp = F3(cc, i, j, kp);
    const T lp = inner(i, j, k + 1) ? rzp : (T)0;
    const int fp = up ? fzp : 0;
    const T cp = up ? czp : (T)0;
    const T w0 = (col && up) ? wz : (T)0;
    const bool cin = inner(i, j, k);
    T acc = lc;
    if ((dl.assign & 4) && !(dl.last && cin)) F3(lentry, i, j, k) = (T)0;
    if (dl.last && cin) {
      const T lo = F3(lout, i, j, k);
      const double ae = fmax(dl.alphae0[i + (nx + 2) * j], (dl.alphaw[k + 1] + dl.alphaw[k]) / 2.0);
      const T le = -(T)(0.5 * ae * dl.tau3) * lo;
      if (dl.assign & 4) F3(lentry, i, j, k) = le;
      else F3(lentry, i, j, k) += le;
      if (k == nz && dl.map_valid && i >= dl.nxs && i <= dl.nxe && j >= dl.nys && j <= dl.nye) {
        const int n = (i - dl.nxs) + (dl.nxe - dl.nxs + 1) * (j - dl.nys);
        dl.gq[n] += (double)lo * dl.tau3 / dl.dxdydz * dl.qbase[n];
      }
    }
    const bool rowx = j >= 1 && j <= ny && k >= 1 && k <= nz, rowy = i >= 1 && i <= nx && k >= 1 && k <= nz;
    const bool xf = rowx && i <= nx, xb = rowx && i >= 1, yf = rowy && j <= ny, yb = rowy && j >= 1;
    const bool zf = col && up, zb = col && k >= 1;
    T v, dv, v2, dv2;
    // x face i: L = (i,j,k), R = (i+1,j,k)
    pos(u_i, mA(f0, fxp), v, dv);
    neg(u_i, mB(f0, fxp), v2, dv2);
    const T lfx = (inner(i + 1, j, k) ? rxp : (T)0) - lc;
    if (xf) acc += lfx * v;
    const T gx = lfx * (dv * c0 + dv2 * cxp);
    // x face i-1: L = (i-1,j,k), R = (i,j,k)
    neg(u_im, mB(fxm, f0), v2, dv2);
    if (xb) acc += (lc - (inner(i - 1, j, k) ? rxm : (T)0)) * v2;
    // y face j
    pos(v_j, mA(f0, fyp), v, dv);
    neg(v_j, mB(f0, fyp), v2, dv2);
    const T lfy = (inner(i, j + 1, k) ? ryp : (T)0) - lc;
    if (yf) acc += lfy * v;
    const T gy = lfy * (dv * c0 + dv2 * cyp);
    // y face j-1
    neg(v_jm, mB(fym, f0), v2, dv2);
    if (yb) acc += (lc - (inner(i, j - 1, k) ? rym : (T)0)) * v2;
    // z face k: g = wm cc(L) + wp cc(R), L = k, R = k+1
    neg(w0, mA(f0, fp), v2, dv2);  // wm
    pos(w0, mB(f0, fp), v, dv);    // wp
    const T lg = lc - lp;
    if (zf) acc += lg * v2;
    const T gz = lg * (dv2 * c0 + dv * cp);
    // z face k-1: L = k-1, R = k
    pos(wm1, mB(fm, f0), v, dv);
    if (zb) acc += (lm - lc) * v;
    if (xf) {
      if (set_u) F3(lud, i, j, k) = gx;
      else F3(lud, i, j, k) += gx;
    }
    if (yf) {
      if (set_u) F3(lvd, i, j, k) = gy;
      else F3(lvd, i, j, k) += gy;
    }
    if (zf) {
      if (set_u) F3(lwd, i, j, k) = gz;
      else F3(lwd, i, j, k) += gz;
    }
    if (set_c) F3(lcc, i, j, k) = acc;
    else F3(lcc, i, j, k) += acc;
    lm = lc; lc = lp;
    fm = f0; f0 = fp;
    c0 = cp;
    wm1 = w0;
  }
}

// planes per block of k_adj_donor_march: at most 16, fewer until the launch
// has 2048 blocks (8 waves on each SIMD)
static dim3 donor_grid(const Layout &L, int &kc) {
  const long nbp = ((long)(L.nx + 2) * (L.ny + 2) + 255) / 256;
  kc = tile_chunk(L.nz + 2, nbp, 16, 2048);
  return dim3((unsigned)nbp, (unsigned)((L.nz + 2 + kc - 1) / kc));
}

// values of a field at offsets (-1..1)^3 around a cell, loaded into
// registers up front; a reader called with constant offsets after inlining
// (entries never loaded are never read)
template <typename T>
struct Nb27 {
  T v[27];
  __device__ T operator()(int a, int b, int c) const { return v[(a + 1) + 3 * (b + 1) + 9 * (c + 1)]; }
  __device__ void load(const T *p, const Layout &L, int i, int j, int k, int a, int b, int c) {
    v[(a + 1) + 3 * (b + 1) + 9 * (c + 1)] = p[IX(L, i + a, j + b, k + c)];
  }
};

// the antidiffusive velocities reversed on the faces of the interior cells
// (ud, vd, wd at (i, j, k) with i, j, k >= 1) as a tile.  The values the
// three face formulas read (cc at 16 offsets, U, V, W at 6 each) are loaded
// before their zero tests, so that a cell's loads are in flight together
// rather than in three rounds, one per face.
template <typename T>
__global__ void __launch_bounds__(ATI *ATJ) k_adj_anti_tile(Layout L, const T *__restrict__ cc,
                                                            const T *__restrict__ U, const T *__restrict__ V,
                                                            const T *__restrict__ W, const T *__restrict__ lud,
                                                            const T *__restrict__ lvd, const T *__restrict__ lwd,
                                                            T *lcc, T *lU, T *lV, T *lW, int kc) {
  typedef AdjRing<T, ATI, ATJ, 4> R;
  __shared__ R sh;
  const int i0 = 1 + (int)blockIdx.x * ATI, j0 = 1 + (int)blockIdx.y * ATJ;
  const int ka = 1 + (int)blockIdx.z * kc, kb = min(L.nz, ka + kc - 1);
  T *const dst[4] = {lcc, lU, lV, lW};
  typedef LRel<T, R::SI, R::S> A;
  adj_march(sh, L, dst, i0, j0, ka, kb, [&](int i, int j, int k) {
    const T lu = F3(lud, i, j, k), lv = F3(lvd, i, j, k), lw = F3(lwd, i, j, k);
    Nb27<T> C, Uf, Vf, Wf;
    // cc: the cell, its six face neighbours and nine edge neighbours
    static constexpr int oc[16][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
                                      {0, 0, -1}, {1, 1, 0},  {1, -1, 0}, {-1, 1, 0}, {1, 0, -1}, {1, 0, 1},
                                      {-1, 0, 1}, {0, 1, -1}, {0, 1, 1},  {0, -1, 1}};
    static constexpr int ou[6][3] = {{0, 0, 0}, {0, 1, 0}, {-1, 1, 0}, {-1, 0, 0}, {0, 0, 1}, {-1, 0, 1}};
    static constexpr int ov[6][3] = {{0, 0, 0}, {1, 0, 0}, {1, -1, 0}, {0, -1, 0}, {0, 0, 1}, {0, -1, 1}};
    static constexpr int ow[6][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, -1}, {0, 0, -1}, {0, 1, 0}, {0, 1, -1}};
#pragma unroll
    for (int q = 0; q < 16; ++q) C.load(cc, L, i, j, k, oc[q][0], oc[q][1], oc[q][2]);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      Uf.load(U, L, i, j, k, ou[q][0], ou[q][1], ou[q][2]);
      Vf.load(V, L, i, j, k, ov[q][0], ov[q][1], ov[q][2]);
      Wf.load(W, L, i, j, k, ow[q][0], ow[q][1], ow[q][2]);
    }
    if (lu == (T)0 && lv == (T)0 && lw == (T)0) return;
    // the three faces' contributions summed per field and offset in
    // registers, then one LDS atomic per offset: 34 instead of 57 a cell
    T rc[27] = {}, rU[27] = {}, rV[27] = {}, rW[27] = {};
    const RegRel<T> gc{rc}, gU{rU}, gV{rV}, gW{rW};
    adj_anti_u<T>(C, Uf, Vf, Wf, lu, gc, gU, gV, gW);
    adj_anti_v<T>(C, Uf, Vf, Wf, lv, gc, gU, gV, gW);
    adj_anti_w<T>(C, Uf, Vf, Wf, lw, gc, gU, gV, gW);
    const A ac{sh.a[0][0], i0 - 1, j0 - 1, i, j, k}, aU{sh.a[1][0], i0 - 1, j0 - 1, i, j, k},
        aV{sh.a[2][0], i0 - 1, j0 - 1, i, j, k}, aW{sh.a[3][0], i0 - 1, j0 - 1, i, j, k};
    reg_flush(rc, oc, ac);
    reg_flush(rU, ou, aU);
    reg_flush(rV, ov, aV);
    reg_flush(rW, ow, aW);
  });
}

// ring depths of pass 1's adjoint: the entry value and ss_v take
// contributions at k-1, k and k+1, the z Courant number at k-1 and k, the x
// and y ones and ss at k only
struct Pass1Planes {
  static constexpr int lo[6] = {-1, 0, 0, -1, 0, -1}, hi[6] = {1, 0, 0, 0, 0, 1};
};

// pass 1 reversed, as a tile
template <typename T>
__global__ void __launch_bounds__(ATI *ATJ) k_adj_pass1_tile(Layout L, StepArgs s, Phys ph, FieldPtrs<T> f,
                                                             const T *__restrict__ c, const T *__restrict__ uc,
                                                             const T *__restrict__ vc, const T *__restrict__ wc,
                                                             const double *__restrict__ azzmin,
                                                             const double *__restrict__ cbkg, const T *__restrict__ lc1,
                                                             T *lentry, T *luc, T *lvc, T *lwc, T *lss, T *lssv, int kc) {
  typedef AdjRingV<T, ATI, ATJ, 6, Pass1Planes> R;
  __shared__ R sh;
  const int i0 = 1 + (int)blockIdx.x * ATI, j0 = 1 + (int)blockIdx.y * ATJ;
  const int ka = 1 + (int)blockIdx.z * kc, kb = min(L.nz, ka + kc - 1);
  T *const dst[6] = {lentry, luc, lvc, lwc, lss, lssv};
  const RView<T> C{c, L}, UC{uc, L}, VC{vc, L}, WC{wc, L}, SS{f.ss, L}, SV{f.ssv, L};
  const LAccV<T, R::SI, R::S, 3> ac{&sh.a[R::base(0)], i0 - 1, j0 - 1}, assv{&sh.a[R::base(5)], i0 - 1, j0 - 1};
  const LAccV<T, R::SI, R::S, 1> auc{&sh.a[R::base(1)], i0 - 1, j0 - 1}, avc{&sh.a[R::base(2)], i0 - 1, j0 - 1},
      ass{&sh.a[R::base(4)], i0 - 1, j0 - 1};
  const LAccV<T, R::SI, R::S, 2> awc{&sh.a[R::base(3)], i0 - 1, j0 - 1};
  adj_march_v(sh, L, dst, i0, j0, ka, kb, [&](int i, int j, int k) {
    const T l = F3(lc1, i, j, k);
    if (l == (T)0) return;
    // the contributions summed per field and offset in registers, then one
    // LDS atomic per offset: 21 instead of 29 a cell
    T rc[27] = {}, ru[27] = {}, rv[27] = {}, rw[27] = {}, rs[27] = {}, rsv[27] = {};
    const RegAbs<T> gc{rc, i, j, k}, gu{ru, i, j, k}, gv{rv, i, j, k}, gw{rw, i, j, k}, gs{rs, i, j, k},
        gsv{rsv, i, j, k};
    adj_pass1<T>(L, s, ph, C, UC, VC, WC, SS, SV, azzmin, cbkg, i, j, k, l, gc, gu, gv, gw, gs, gsv);
    static constexpr int o7[7][3] = {HC_NB7};
    static constexpr int o5[5][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
    static constexpr int oz[3][3] = {{0, 0, -1}, {0, 0, 0}, {0, 0, 1}};
    static constexpr int ox[2][3] = {{-1, 0, 0}, {0, 0, 0}}, oy[2][3] = {{0, -1, 0}, {0, 0, 0}},
                         ozw[2][3] = {{0, 0, -1}, {0, 0, 0}};
    reg_flush_at(rc, o7, ac, i, j, k);
    reg_flush_at(ru, ox, auc, i, j, k);
    reg_flush_at(rv, oy, avc, i, j, k);
    reg_flush_at(rw, ozw, awc, i, j, k);
    reg_flush_at(rs, o5, ass, i, j, k);
    // (at k = 1 the entry below and at k = nz the one above stay zero: pass
    // 1 has no z face there, and the flush skips zero entries)
    reg_flush_at(rsv, oz, assv, i, j, k);
  });
}

// ---------------------------------------------------------------------------
// probes
__global__ void k_probe_sample(int n, const int *__restrict__ iw, const int *__restrict__ it, const double *w,
                               const double *t, double *out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  out[2 * q] = w[iw[q]];
  out[2 * q + 1] = t[it[q]];
}
__global__ void k_probe_sample32(int n, const int *__restrict__ iw, const int *__restrict__ it, const float *w,
                                 const float *t, double *out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  out[2 * q] = w[iw[q]];
  out[2 * q + 1] = t[it[q]];
}
template <typename T>
__global__ void k_probe_seed(int n, const int *__restrict__ iw, const int *__restrict__ it,
                             const double *__restrict__ samp, const double *__restrict__ obs, double ww, double wt,
                             T *lw, T *lt) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  adj_add(lw + iw[q], (T)(ww * (samp[2 * q] - obs[2 * q])));
  adj_add(lt + it[q], (T)(wt * (samp[2 * q + 1] - obs[2 * q + 1])));
}

}  // namespace hc

// ===========================================================================
// adjoint state of a context
// a state of the window: the 11 fields the step reads, and the up, vp, wp
// buffers, whose cells MOMENTUM does not write become u, v, w at REFRESH
static const int NTAPE = 14;
static const int TAPE_FIELDS[NTAPE] = {HC_U,    HC_V,   HC_W, HC_UM, HC_VM, HC_WM, HC_SS,
                                       HC_SSV,  HC_RHOP, HC_T, HC_S,  HC_UP, HC_VP, HC_WP};

// a step's taped tracer intermediates: the results of passes 1 and 2 and
// the antidiffusive velocities of passes 2 and 3
static const int TTAPE = 8;
// bytes of one tracer's taped step: TTAPE fields and the limiter flags of
// passes 2 and 3 (one byte per cell each, 256-byte aligned)
static size_t ttape_block(const hc_ctx *ctx) {
  const size_t fl = (ctx->L.n + 255) / 256 * 256;
  return (size_t)TTAPE * ctx->L.n * ctx->esz + 2 * fl;
}
struct TracerSlot {
  void *c1, *c2, *ud2[3], *ud3[3];
  uint8_t *fl2, *fl3;  // limiter flags of passes 2 and 3
};

struct HostState {
  double time;
  double acc[8];
  int next_step;
  bool map_valid;
  bool up_alias[3];
};

// scratch of one TRACER31 adjoint: the recompute's intermediates (when no
// tape holds them) and the adjoints of the passes' intermediates
struct AdjTracerScratch {
  void *lcnew = nullptr;
  void *cw = nullptr, *c1 = nullptr, *c2 = nullptr, *cpre = nullptr, *ud2[3] = {}, *ud3[3] = {};
  void *lc1 = nullptr, *lc2 = nullptr, *lpre = nullptr, *lud2[3] = {}, *lud3[3] = {};
  uint8_t *fl2 = nullptr, *fl3 = nullptr;
};

struct AdjCtx {
  void *lam[HC_NFIELDS] = {};
  void *luc[3] = {};
  void *lamb = nullptr;  // MOMENTUM's Lambda
  // the T adjoint's scratch (and that of the per-subroutine entry), and the
  // S adjoint's when a window's step runs S on a stream of its own beside T
  AdjTracerScratch sc[2];
  // the context's two side streams (ctx->tst): S's tracer work on st2, the
  // tape's copies of a window's states on stc while the next step's FORCES
  // runs (copy_pending until the context stream has waited for them)
  hipStream_t st2 = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipStream_t stc = nullptr;
  hipEvent_t evc[2] = {nullptr, nullptr};
  bool copy_pending = false;
  void *g = nullptr;          // compact pressure work
  // ADJ_SLOTS partial sums each of d/d flow_scale, sum lam_ss*ss and
  // sum lam_ssv*ssv, then a device 0.0
  double *d_sc = nullptr;
  double *d_part = nullptr;   // nz+2 partial sums
  double *d_gq = nullptr;     // npatch
  // window
  void *tape = nullptr;       // (nck + seg + 1) states of NTAPE fields
  size_t tape_slots = 0;
  // the tracers' intermediates of the recomputed steps of a segment (steps
  // a .. b-2), TTAPE fields per tracer (T, S) per step, so that their
  // adjoints need not run the tracer passes once more
  void *ttape = nullptr;
  size_t ttape_steps = 0;
  int nprobe = 0, every = 1;
  double wgt_w = 1.0, wgt_t = 1.0;
  int *d_iw = nullptr, *d_it = nullptr;
  double *d_samp = nullptr, *d_obs = nullptr;
  size_t nsamp_cap = 0;
};

void adj_free(hc_ctx *ctx) {
  AdjCtx *A = ctx->adj;
  if (!A) return;
  hipStreamSynchronize(ctx->st);
  std::vector<void *> ps;
  if (A->st2) hipStreamSynchronize(A->st2);  // the context's streams: not destroyed here
  if (A->stc) hipStreamSynchronize(A->stc);
  for (hipEvent_t e : A->ev)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : A->evc)
    if (e) hipEventDestroy(e);
  for (void *p : A->lam) ps.push_back(p);
  for (void *p : A->luc) ps.push_back(p);
  for (AdjTracerScratch &X : A->sc) {
    for (int q = 0; q < 3; ++q) { ps.push_back(X.ud2[q]); ps.push_back(X.ud3[q]); ps.push_back(X.lud2[q]); ps.push_back(X.lud3[q]); }
    void *xs[] = {X.lcnew, X.cw, X.c1, X.c2, X.cpre, X.lc1, X.lc2, X.lpre, X.fl2, X.fl3};
    for (void *p : xs) ps.push_back(p);
  }
  void *rest[] = {A->lamb, A->g, A->d_sc, A->d_part, A->d_gq, A->tape, A->ttape, A->d_iw, A->d_it, A->d_samp,
                  A->d_obs};
  for (void *p : rest) ps.push_back(p);
  for (void *p : ps)
    if (p) hipFree(p);
  delete A;
  ctx->adj = nullptr;
}

static int scratch_alloc(hc_ctx *ctx, AdjTracerScratch &X) {
  const size_t fb = ctx->L.n * ctx->esz;
  int rc = 0;
  void **fields[] = {&X.lcnew, &X.cw, &X.c1, &X.c2, &X.cpre, &X.ud2[0], &X.ud2[1], &X.ud2[2], &X.ud3[0], &X.ud3[1],
                     &X.ud3[2], &X.lc1, &X.lc2, &X.lpre, &X.lud2[0], &X.lud2[1], &X.lud2[2], &X.lud3[0], &X.lud3[1],
                     &X.lud3[2]};
  for (auto p : fields)
    if (!(*p = dalloc(ctx, fb, &rc))) return rc;
  if (!(X.fl2 = (uint8_t *)dalloc(ctx, ctx->L.n, &rc))) return rc;
  if (!(X.fl3 = (uint8_t *)dalloc(ctx, ctx->L.n, &rc))) return rc;
  return HC_OK;
}

// the S adjoint's stream and scratch, for running it beside T's
static int adj_par_init(hc_ctx *ctx) {
  AdjCtx *A = ctx->adj;
  if (A->st2) return HC_OK;
  int r = scratch_alloc(ctx, A->sc[1]);
  if (r) return r;
  if ((r = tracer_streams(ctx))) return r;
  for (hipEvent_t &e : A->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  A->st2 = ctx->tst[0];
  return HC_OK;
}

static int adj_alloc(hc_ctx *ctx) {
  if (ctx->adj) return HC_OK;
  AdjCtx *A = new AdjCtx();
  ctx->adj = A;
  const size_t fb = ctx->L.n * ctx->esz;
  int rc = 0;
  for (int q = 0; q < HC_NFIELDS; ++q)
    if (!(A->lam[q] = dalloc(ctx, fb, &rc))) return rc;
  void **fields[] = {&A->luc[0], &A->luc[1], &A->luc[2], &A->lamb};
  for (auto p : fields)
    if (!(*p = dalloc(ctx, fb, &rc))) return rc;
  if ((rc = scratch_alloc(ctx, A->sc[0]))) return rc;
  if (!(A->g = dalloc(ctx, (size_t)ctx->P.nx * ctx->P.ny * ctx->P.nz * ctx->esz, &rc))) return rc;
  if (!(A->d_sc = (double *)dalloc(ctx, (3 * ADJ_SLOTS + 1) * 8, &rc))) return rc;
  if (!(A->d_part = (double *)dalloc(ctx, (size_t)(ctx->P.nz + 2) * 8, &rc))) return rc;
  if (!(A->d_gq = (double *)dalloc(ctx, (size_t)ctx->npatch * 8, &rc))) return rc;
  return HC_OK;
}

static int adj_need(hc_ctx *ctx) {
  int r = adj_alloc(ctx);
  if (r) return r;
  return HC_OK;
}

template <typename T>
static FieldPtrs<T> lptrs(hc_ctx *ctx) {
  FieldPtrs<T> p;
  T **a = (T **)&p;
  for (int q = 0; q < HC_NFIELDS; ++q) a[q] = (T *)ctx->adj->lam[q];
  return p;
}

static FlowSens flow_sens(hc_ctx *ctx) {
  FlowSens f;
  f.dubkg_m = ctx->dflow[0]; f.dubkg = ctx->dflow[1]; f.dvbkg_m = ctx->dflow[2]; f.dvbkg = ctx->dflow[3];
  f.dfbx = ctx->dflow[4]; f.dfby = ctx->dflow[5];
  return f;
}

static void zero(hc_ctx *ctx, void *p) { hipMemsetAsync(p, 0, ctx->L.n * ctx->esz, ctx->st); }
// n <= 8 whole fields to zero, one launch on stream st
static void zero_fields(hc_ctx *ctx, hipStream_t st, void *const *ps, int n) {
  ZeroList z{};
  for (int q = 0; q < n; ++q) z.p[q] = (uint4 *)ps[q];
  const size_t n16 = ctx->L.n * ctx->esz / 16;
  hipLaunchKernelGGL(k_zero_fields, dim3((unsigned)((n16 + 255) / 256), (unsigned)n), dim3(256), 0, st, z, n16);
}

static Box3 interior(const Layout &L) { return Box3{1, L.nx, 1, L.ny, 1, L.nz}; }

// ---------------------------------------------------------------------------
// per-routine adjoints (templated on the precision)
template <typename T>
static void adj_forces_t(hc_ctx *ctx, const T *ss0, const T *ssv0) {
  AdjCtx *A = ctx->adj;
  const Layout &L = ctx->L;
  FieldPtrs<T> lam = lptrs<T>(ctx);
  const StepArgs s = step_args(ctx);
  const FlowSens fs = flow_sens(ctx);
  // the faces i = 0 (fx), j = 0 (fy), k = 0 (fz) cell by cell, the interior
  // as LDS-accumulating tiles
  const Box3 faces[3] = {{0, 0, 0, L.ny, 0, L.nz}, {1, L.nx, 0, 0, 0, L.nz}, {1, L.nx, 1, L.ny, 0, 0}};
  for (const Box3 &b : faces)
    hipLaunchKernelGGL(k_adj_forces<T>, grid3(b), dim3(64, 4), 0, ctx->st, L, b, s, ctx->P, fs, ptrs<T>(ctx), ss0,
                       ssv0, lam, ctx->aux.azzmin, ctx->aux.alphau, ctx->aux.alphav, ctx->aux.alphae0,
                       ctx->aux.alphaw, A->d_sc);
  const int kc = adj_chunk(L);
  const dim3 tg = adj_tile_grid<ATI, ATJ>(L, kc);
  hipLaunchKernelGGL((k_adj_forces_tile<T, 1>), tg, dim3(ATI * ATJ), 0, ctx->st, L, s, ctx->P,
                     fs, ptrs<T>(ctx), ss0, ssv0, lam, ctx->aux.azzmin, ctx->aux.alphau, ctx->aux.alphav,
                     ctx->aux.alphae0, ctx->aux.alphaw, A->d_sc, kc);
  hipLaunchKernelGGL((k_adj_forces_tile<T, 2>), tg, dim3(ATI * ATJ), 0, ctx->st, L, s, ctx->P,
                     fs, ptrs<T>(ctx), ss0, ssv0, lam, ctx->aux.azzmin, ctx->aux.alphau, ctx->aux.alphav,
                     ctx->aux.alphae0, ctx->aux.alphaw, A->d_sc, kc);
  hipLaunchKernelGGL((k_adj_forces_tile<T, 4>), tg, dim3(ATI * ATJ), 0, ctx->st, L, s, ctx->P,
                     fs, ptrs<T>(ctx), ss0, ssv0, lam, ctx->aux.azzmin, ctx->aux.alphau, ctx->aux.alphav,
                     ctx->aux.alphae0, ctx->aux.alphaw, A->d_sc, kc);
  // lam.fx, fy, fz are nonzero only where FORCES writes (MOMENTUM's and
  // PRESSURE's adjoints add to them there), and the kernels above cleared
  // every such cell after reading it
}

template <typename T>
static void adj_momentum_t(hc_ctx *ctx, int comp, bool clear = true) {
  AdjCtx *A = ctx->adj;
  const Layout &L = ctx->L;
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  FieldPtrs<T> lam = lptrs<T>(ctx);
  StepArgs s = step_args(ctx);
  T *lup = comp == 0 ? lam.up : comp == 1 ? lam.vp : lam.wp;
  T *lu = comp == 0 ? lam.u : comp == 1 ? lam.v : lam.w;
  T *lum = comp == 0 ? lam.um : comp == 1 ? lam.vm : lam.wm;
  T *lf = comp == 0 ? lam.fx : comp == 1 ? lam.fy : lam.fz;
  Box3 fb = comp == 0 ? Box3{0, nx, 1, ny, 1, nz} : comp == 1 ? Box3{1, nx, 0, ny, 1, nz} : Box3{1, nx, 1, ny, 1, nz - 1};
  Box3 pb = comp == 0 ? Box3{0, nx + 1, 1, ny, 1, nz} : comp == 1 ? Box3{1, nx, 0, ny + 1, 1, nz} : Box3{1, nx, 1, ny, 1, nz};
  hipLaunchKernelGGL(k_adj_mom_lambda<T>, grid3(fb), dim3(64, 4), 0, ctx->st, L, fb, comp, s, ctx->P, (const T *)lup,
                     (T *)A->lamb, lu, lum, lf);
  const double d = comp == 0 ? s.dx : comp == 1 ? s.dy : s.dz;
  hipLaunchKernelGGL(k_adj_mom_p<T>, grid3(pb), dim3(64, 4), 0, ctx->st, L, pb, fb, comp, (T)(s.dt * s.leap / d),
                     (const T *)A->lamb, lam.p);
  if (clear) zero(ctx, lup);  // consumed (else the caller clears it)
}

template <typename T>
static void adj_pressure_t(hc_ctx *ctx) {
  AdjCtx *A = ctx->adj;
  launch_pressure_adj<T>(ctx->st, ctx->L, step_args(ctx), ctx->P, ptrs<T>(ctx), lptrs<T>(ctx), ctx->aux, ctx->pw,
                         flow_sens(ctx), A->d_sc, (T *)A->g, A->d_sc + 3 * ADJ_SLOTS, A->d_part);
}

template <typename T>
static void adj_viscosity_t(hc_ctx *ctx, const T *ss1, const T *ssv1) {
  AdjCtx *A = ctx->adj;
  const Layout &L = ctx->L;
  FieldPtrs<T> f = ptrs<T>(ctx), lam = lptrs<T>(ctx);
  const StepArgs s = step_args(ctx);
  const long nshell = 2L * (L.ny + 2) * (L.nz + 2) + 2L * L.nx * (L.nz + 2) + 2L * L.nx * L.ny;
  hipLaunchKernelGGL(k_adj_visc_shell<T>, dim3((unsigned)((nshell + 255) / 256)), dim3(256), 0, ctx->st, L, s, ctx->P,
                     (const T *)f.um, (const T *)f.vm, (const T *)f.wm, (const T *)f.rhop,
                     (const double *)ctx->aux.rho_bkg, ss1, ssv1, lam.ss, lam.ssv, lam.um,
                     lam.vm, lam.wm, lam.rhop, A->d_sc + ADJ_SLOTS);
  const int kc = adj_chunk(L);
  const dim3 tg = adj_tile_grid<ATI, ATJ>(L, kc);
  hipLaunchKernelGGL(k_adj_visc_tile<T>, tg, dim3(ATI * ATJ), 0, ctx->st, L, s, ctx->P,
                     (const T *)f.um, (const T *)f.vm, (const T *)f.wm, (const T *)f.rhop,
                     (const double *)ctx->aux.rho_bkg, ss1, ssv1, lam.ss, lam.ssv, lam.um,
                     lam.vm, lam.wm, lam.rhop, A->d_sc + ADJ_SLOTS, kc);
  // VISCOSITY3D writes every cell of ss (constants on the edges and the k
  // ghosts) but leaves ss_v on the i and j ghost columns as it was: the two
  // kernels clear lam.ss everywhere and lam.ssv on the interior columns
  // after reading them
}

template <typename T>
static void adj_density_t(hc_ctx *ctx, const T *t1, const T *s1) {
  const Layout &L = ctx->L;
  Box3 b{0, L.nx + 1, 0, L.ny + 1, 0, L.nz + 1};
  FieldPtrs<T> lam = lptrs<T>(ctx);
  hipLaunchKernelGGL(k_adj_density<T>, grid3(b), dim3(64, 4), 0, ctx->st, L, b, ctx->P, t1, s1, (const T *)lam.rhop,
                     lam.t, lam.s);
  zero(ctx, lam.rhop);
}

template <typename T>
static void adj_refresh_t(hc_ctx *ctx) {
  FieldPtrs<T> lam = lptrs<T>(ctx);
  size_t n = ctx->L.n;
  T *tri[3][3] = {{lam.u, lam.up, lam.um}, {lam.v, lam.vp, lam.vm}, {lam.w, lam.wp, lam.wm}};
  for (auto &t : tri)
    hipLaunchKernelGGL(k_adj_refresh<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->st, n, ctx->leap,
                       ctx->P.assel, t[0], t[1], t[2]);
}

template <typename T>
static void adj_courant_t(hc_ctx *ctx, double tau3) {
  AdjCtx *A = ctx->adj;
  const Layout &L = ctx->L;
  Box3 b{0, L.nx + 1, 0, L.ny + 1, 0, L.nz + 1};
  FieldPtrs<T> lam = lptrs<T>(ctx);
  hipLaunchKernelGGL(k_adj_courant<T>, grid3(b), dim3(64, 4), 0, ctx->st, L, b, tau3, ctx->P.dx, ctx->P.dy, ctx->P.dz,
                     (T *)A->luc[0], (T *)A->luc[1], (T *)A->luc[2], lam.u, lam.up, lam.v, lam.vp, lam.w, lam.wp);
}

// TRACER31 for field `field` at the current state: lam[field] holds the
// adjoint of the call's output and receives the adjoint of its input.  The
// passes' intermediates come from `taped` (a window's segment recompute) or
// are recomputed here on a copy of the entry value.
template <typename T>
static void adj_tracer_t(hc_ctx *ctx, int field, const double *d_cbkg, double cmin, double cmax, double tau3,
                         int vent, const double *qpatch, const TracerSlot *taped = nullptr, int which = 0) {
  AdjCtx *A = ctx->adj;
  AdjTracerScratch &X = A->sc[which];
  const Layout &L = ctx->L;
  hipStream_t st = which ? A->st2 : ctx->st;
  const size_t fb = L.n * sizeof(T);
  StepArgs s = step_args(ctx);
  s.tau3 = tau3;
  VentArgs va = vent_args(ctx);
  FieldPtrs<T> f = ptrs<T>(ctx);
  const T *c = (const T *)ctx->f[field];
  TracerSlot is{X.c1, X.c2, {X.ud2[0], X.ud2[1], X.ud2[2]}, {X.ud3[0], X.ud3[1], X.ud3[2]}, X.fl2, X.fl3};
  if (taped) {
    is = *taped;
  } else {
    hipMemcpyAsync(X.cw, c, fb, hipMemcpyDeviceToDevice, st);
    TracerWork tw{};
    tw.uc = ctx->tw.uc; tw.vc = ctx->tw.vc; tw.wc = ctx->tw.wc;
    tw.ud2 = is.ud2[0]; tw.vd2 = is.ud2[1]; tw.wd2 = is.ud2[2];
    tw.c1 = is.c1; tw.c2 = is.c2; tw.cpre = X.cpre;
    tw.ud3 = is.ud3[0]; tw.vd3 = is.ud3[1]; tw.wd3 = is.ud3[2];
    tw.fl2 = is.fl2; tw.fl3 = is.fl3;
    tw.budget_part = ctx->tw.budget_part; tw.budget = ctx->tw.budget;
    TracerArgs ta{cmin, cmax, tau3, 0, 0.0, 0};
    launch_tracer<T>(st, L, s, ctx->P, va, f, ctx->aux, tw, (T *)X.cw, d_cbkg, qpatch, ta);
  }
  const T *c1 = (const T *)is.c1, *c2 = (const T *)is.c2;
  const T *ud2 = (const T *)is.ud2[0], *vd2 = (const T *)is.ud2[1], *wd2 = (const T *)is.ud2[2];
  const T *ud3 = (const T *)is.ud3[0], *vd3 = (const T *)is.ud3[1], *wd3 = (const T *)is.ud3[2];
  // the limiter flags of both passes, as the forward passes formed them
  // (taped, or written by the recompute above)
  const uint8_t *fl2 = is.fl2, *fl3 = is.fl3;
  // adjoint buffers: lpre and the pass-2 face velocities' adjoints are
  // first added to by atomics; the donor gathers assign the others on their
  // first write (lc2, lud3 and lcnew in pass 3, lc1 in pass 2)
  void *const zs[] = {X.lpre, X.lud2[0], X.lud2[1], X.lud2[2]};
  zero_fields(ctx, st, zs, 4);
  T *lout = (T *)A->lam[field];
  const long nghost = 2L * (L.ny + 2) * (L.nz + 2) + 2L * L.nx * (L.nz + 2) + 2L * L.nx * L.ny;
  const unsigned nbc = (unsigned)((nghost + 255) / 256);
  // final BOUNDARY_CONDITIONS from the pre-sponge values
  hipLaunchKernelGGL(k_adj_bc<T>, dim3(nbc), dim3(256), 0, st, L, ctx->P, va, 1, vent, lout, (T *)X.lpre);
  // pass 3
  DonorLast dl{};
  dl.last = 1; dl.tau3 = tau3;
  dl.assign = 1 | 2 | 4; dl.dxdydz = ctx->P.dx * ctx->P.dy * ctx->P.dz;
  dl.alphae0 = ctx->aux.alphae0; dl.alphaw = ctx->aux.alphaw; dl.qbase = ctx->d_qbase;
  dl.gq = (field == HC_T) ? A->d_gq : nullptr;
  dl.map_valid = (field == HC_T && ctx->d_qbase) ? va.map_valid : 0;
  dl.nxs = va.nxs; dl.nxe = va.nxe; dl.nys = va.nys; dl.nye = va.nye;
  const int kc = adj_chunk(L);
  const dim3 tg = adj_tile_grid<ATI, ATJ>(L, kc), tb(ATI * ATJ);
  // the antidiffusive velocities' adjoint: the domain's first faces in one
  // launch, the interior faces as tiles
  const long nfaces = (long)L.ny * L.nz + (long)L.nx * L.nz + (long)L.nx * L.ny;
  auto anti = [&](const T *cc, const T *U, const T *V, const T *W, const T *lu, const T *lv, const T *lw, T *lcc,
                  T *lU, T *lV, T *lW) {
    hipLaunchKernelGGL(k_adj_anti_faces<T>, dim3((unsigned)((nfaces + 255) / 256)), dim3(256), 0, st, L, cc, U, V, W,
                       lu, lv, lw, lcc, lU, lV, lW);
    hipLaunchKernelGGL(k_adj_anti_tile<T>, tg, tb, 0, st, L, cc, U, V, W, lu, lv, lw, lcc, lU, lV, lW, kc);
  };
  int dkc = 0;
  const dim3 dg = donor_grid(L, dkc);
  hipLaunchKernelGGL(k_adj_donor_march<T>, dg, dim3(256), 0, st, L, dl, dkc, (const T *)lout,
                     (const T *)X.lpre, c2, ud3, vd3, wd3, fl3, (T *)X.lc2, (T *)X.lud3[0],
                     (T *)X.lud3[1], (T *)X.lud3[2], (T *)X.lcnew);
  anti(c2, ud2, vd2, wd2, (const T *)X.lud3[0], (const T *)X.lud3[1], (const T *)X.lud3[2], (T *)X.lc2,
       (T *)X.lud2[0], (T *)X.lud2[1], (T *)X.lud2[2]);
  hipLaunchKernelGGL(k_adj_bc<T>, dim3(nbc), dim3(256), 0, st, L, ctx->P, va, 0, 0, (T *)X.lc2, (T *)X.lc2);
  // pass 2
  DonorLast d2{};
  d2.assign = 1;
  hipLaunchKernelGGL(k_adj_donor_march<T>, dg, dim3(256), 0, st, L, d2, dkc, (const T *)X.lc2,
                     (const T *)nullptr, c1, ud2, vd2, wd2, fl2, (T *)X.lc1,
                     (T *)X.lud2[0], (T *)X.lud2[1], (T *)X.lud2[2], (T *)nullptr);
  anti(c1, (const T *)ctx->tw.uc, (const T *)ctx->tw.vc, (const T *)ctx->tw.wc, (const T *)X.lud2[0],
       (const T *)X.lud2[1], (const T *)X.lud2[2], (T *)X.lc1, (T *)A->luc[0], (T *)A->luc[1], (T *)A->luc[2]);
  hipLaunchKernelGGL(k_adj_bc<T>, dim3(nbc), dim3(256), 0, st, L, ctx->P, va, 0, 0, (T *)X.lc1, (T *)X.lc1);
  // pass 1
  FieldPtrs<T> lam = lptrs<T>(ctx);
  hipLaunchKernelGGL(k_adj_pass1_tile<T>, tg, tb, 0, st, L, s, ctx->P, f, c, (const T *)ctx->tw.uc,
                     (const T *)ctx->tw.vc, (const T *)ctx->tw.wc, (const double *)ctx->aux.azzmin, d_cbkg,
                     (const T *)X.lc1, (T *)X.lcnew, (T *)A->luc[0], (T *)A->luc[1], (T *)A->luc[2], lam.ss,
                     lam.ssv, kc);
  // the entry value's adjoint becomes the field's
  void *tmp = A->lam[field];
  A->lam[field] = X.lcnew;
  X.lcnew = tmp;
}

// ---------------------------------------------------------------------------
// C-ABI
#define ADJ_DISPATCH(fn, ...)                        \
  do {                                               \
    if (ctx->prec == HC_FP64) fn<double>(__VA_ARGS__); \
    else fn<float>(__VA_ARGS__);                     \
  } while (0)

static int adj_field_ptr(hc_ctx *ctx, int field, void **p, int *bounds_field) {
  if (field >= 0 && field < HC_NFIELDS) {
    *p = ctx->adj->lam[field];
    *bounds_field = field;
    return HC_OK;
  }
  if (field >= HC_ADJ_UC && field <= HC_ADJ_WC) {
    *p = ctx->adj->luc[field - HC_ADJ_UC];
    *bounds_field = field == HC_ADJ_UC ? HC_U : field == HC_ADJ_VC ? HC_V : HC_W;
    return HC_OK;
  }
  return fail(ctx, HC_EINVAL, "bad adjoint field id");
}

extern "C" int hc_adj_reset(hc_ctx *ctx) {
  int r = adj_need(ctx);
  if (r) return r;
  AdjCtx *A = ctx->adj;
  for (void *p : A->lam) zero(ctx, p);
  for (void *p : A->luc) zero(ctx, p);
  HIPCHK(hipMemsetAsync(A->d_sc, 0, (3 * ADJ_SLOTS + 1) * 8, ctx->st));
  HIPCHK(hipMemsetAsync(A->d_gq, 0, (size_t)ctx->npatch * 8, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  return HC_OK;
}

extern "C" int hc_adj_upload(hc_ctx *ctx, int field, const double *host, const int32_t lo[3], const int32_t hi[3]) {
  int r = adj_need(ctx);
  if (r) return r;
  void *p;
  int bf;
  if ((r = adj_field_ptr(ctx, field, &p, &bf))) return r;
  return box_copy(ctx, p, bf, (double *)host, lo, hi, true);
}

extern "C" int hc_adj_download(hc_ctx *ctx, int field, double *host, const int32_t lo[3], const int32_t hi[3]) {
  int r = adj_need(ctx);
  if (r) return r;
  void *p;
  int bf;
  if ((r = adj_field_ptr(ctx, field, &p, &bf))) return r;
  return box_copy(ctx, p, bf, host, lo, hi, false);
}

extern "C" int hc_adj_forces(hc_ctx *ctx) {
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_forces");
  if (ctx->prec == HC_FP64) adj_forces_t<double>(ctx, (const double *)ctx->f[HC_SS], (const double *)ctx->f[HC_SSV]);
  else adj_forces_t<float>(ctx, (const float *)ctx->f[HC_SS], (const float *)ctx->f[HC_SSV]);
  return HC_OK;
}

extern "C" int hc_adj_pressure(hc_ctx *ctx) {
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_pressure");
  ADJ_DISPATCH(adj_pressure_t, ctx);
  return HC_OK;
}

static int adj_mom(hc_ctx *ctx, int comp) {
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_momentum");
  ADJ_DISPATCH(adj_momentum_t, ctx, comp);
  return HC_OK;
}
extern "C" int hc_adj_momentum_u(hc_ctx *ctx) { return adj_mom(ctx, 0); }
extern "C" int hc_adj_momentum_v(hc_ctx *ctx) { return adj_mom(ctx, 1); }
extern "C" int hc_adj_momentum_w(hc_ctx *ctx) { return adj_mom(ctx, 2); }

extern "C" int hc_adj_viscosity3d(hc_ctx *ctx) {
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_viscosity3d");
  if (ctx->prec == HC_FP64)
    adj_viscosity_t<double>(ctx, (const double *)ctx->f[HC_SS], (const double *)ctx->f[HC_SSV]);
  else
    adj_viscosity_t<float>(ctx, (const float *)ctx->f[HC_SS], (const float *)ctx->f[HC_SSV]);
  return HC_OK;
}

extern "C" int hc_adj_work_velocity(hc_ctx *ctx, double tau3) {
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_work_velocity");
  ADJ_DISPATCH(adj_courant_t, ctx, tau3);
  return HC_OK;
}

extern "C" int hc_adj_tracer31(hc_ctx *ctx, int field, const double *c_bkg, double cmin, double cmax, double tau3) {
  if (field != HC_T && field != HC_S && field != HC_C) return fail(ctx, HC_EINVAL, "tracer field must be T, S or C");
  if (ctx->P.ismol != 1) return fail(ctx, HC_EINVAL, "only ismol = 1 (two corrective passes) is supported");
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_tracer31");
  HIPCHK(hipMemcpyAsync(ctx->d_cbkg, c_bkg, (size_t)(ctx->P.nz + 2) * 8, hipMemcpyHostToDevice, ctx->st));
  const int which = field == HC_T ? 0 : field == HC_S ? 1 : 2;
  ADJ_DISPATCH(adj_tracer_t, ctx, field, ctx->d_cbkg, cmin, cmax, tau3, 0, ctx->d_q[which]);
  HIPCHK(hipStreamSynchronize(ctx->st));
  return HC_OK;
}

template <typename T>
__global__ void k_adj_vent(Layout L, VentArgs va, T *lc) {
  int i = va.nxs + (int)threadIdx.x, j = va.nys + (int)blockIdx.x;
  if (!va.map_valid || i > va.nxe || j > va.nye) return;
  lc[IX(L, i, j, L.nz + 1)] = (T)0;
}

extern "C" int hc_adj_vent_ghost(hc_ctx *ctx, int field) {
  int r = adj_need(ctx);
  if (r) return r;
  VentArgs va = vent_args(ctx);
  if (!va.map_valid) return HC_OK;
  dim3 g(va.nye - va.nys + 1), b(va.nxe - va.nxs + 1);
  if (ctx->prec == HC_FP64) hipLaunchKernelGGL(k_adj_vent<double>, g, b, 0, ctx->st, ctx->L, va, (double *)ctx->adj->lam[field]);
  else hipLaunchKernelGGL(k_adj_vent<float>, g, b, 0, ctx->st, ctx->L, va, (float *)ctx->adj->lam[field]);
  return HC_OK;
}

extern "C" int hc_adj_density(hc_ctx *ctx) {
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_density");
  if (ctx->prec == HC_FP64) adj_density_t<double>(ctx, (const double *)ctx->f[HC_T], (const double *)ctx->f[HC_S]);
  else adj_density_t<float>(ctx, (const float *)ctx->f[HC_T], (const float *)ctx->f[HC_S]);
  return HC_OK;
}

extern "C" int hc_adj_refresh(hc_ctx *ctx) {
  int r = adj_need(ctx);
  if (r) return r;
  PhaseScope ps(ctx, "adj_refresh");
  ADJ_DISPATCH(adj_refresh_t, ctx);
  return HC_OK;
}

extern "C" int hc_adj_gradients(hc_ctx *ctx, double *gq, double out[3]) {
  int r = adj_need(ctx);
  if (r) return r;
  AdjCtx *A = ctx->adj;
  std::vector<double> sc(3 * ADJ_SLOTS);
  HIPCHK(hipMemcpyAsync(sc.data(), A->d_sc, 3 * ADJ_SLOTS * 8, hipMemcpyDeviceToHost, ctx->st));
  if (gq) HIPCHK(hipMemcpyAsync(gq, A->d_gq, (size_t)ctx->npatch * 8, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  double sum[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < 3; ++q)
    for (int e = 0; e < ADJ_SLOTS; ++e) sum[q] += sc[q * ADJ_SLOTS + e];
  out[0] = sum[0];
  out[1] = ctx->P.cs_h != 0.0 ? 2.0 * sum[1] / ctx->P.cs_h : 0.0;
  out[2] = ctx->P.cs_v != 0.0 ? 2.0 * sum[2] / ctx->P.cs_v : 0.0;
  return HC_OK;
}

extern "C" int hc_set_controls(hc_ctx *ctx, const double *qmap, double flow_scale, double cs) {
  if (!(flow_scale == flow_scale)) return fail(ctx, HC_EINVAL, "flow_scale is NaN");
  ctx->flow_scale = flow_scale;
  if (cs > 0.0) {
    ctx->P.cs_h = cs;
    ctx->P.cs_v = cs;
  }
  if (qmap) {
    if (!ctx->d_qmap) {
      int rc = 0;
      if (!(ctx->d_qmap = (double *)dalloc(ctx, (size_t)ctx->npatch * 8, &rc))) return rc;
    }
    HIPCHK(hipMemcpyAsync(ctx->d_qmap, qmap, (size_t)ctx->npatch * 8, hipMemcpyHostToDevice, ctx->st));
  } else if (ctx->d_qmap) {
    HIPCHK(hipStreamSynchronize(ctx->st));
    hipFree(ctx->d_qmap);
    ctx->d_qmap = nullptr;
  }
  HIPCHK(hipStreamSynchronize(ctx->st));
  return HC_OK;
}

// ---------------------------------------------------------------------------
// window: forward with checkpoints, probe-misfit loss, reverse sweep
extern "C" int hc_adj_set_probes(hc_ctx *ctx, int32_t nprobe, const int32_t *ijk, int32_t every, double wgt_w,
                                 double wgt_t) {
  int r = adj_need(ctx);
  if (r) return r;
  AdjCtx *A = ctx->adj;
  if (nprobe < 1 || every < 1) return fail(ctx, HC_EINVAL, "need at least one probe and every >= 1");
  const hc_params &P = ctx->P;
  std::vector<int> iw(nprobe), it(nprobe);
  for (int q = 0; q < nprobe; ++q) {
    int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
    if (i < 1 || i > P.nx || j < 1 || j > P.ny || k < 1 || k > P.nz - 1)
      return fail(ctx, HC_EINVAL, "probe outside the interior (1..nx, 1..ny, 1..nz-1)");
    iw[q] = IX(ctx->L, i, j, k);
    it[q] = IX(ctx->L, i, j, k);
  }
  if (A->d_iw) { hipFree(A->d_iw); hipFree(A->d_it); }
  int rc = 0;
  if (!(A->d_iw = (int *)dalloc(ctx, (size_t)nprobe * 4, &rc))) return rc;
  if (!(A->d_it = (int *)dalloc(ctx, (size_t)nprobe * 4, &rc))) return rc;
  HIPCHK(h2d(ctx, A->d_iw, iw.data(), (size_t)nprobe * 4));
  HIPCHK(h2d(ctx, A->d_it, it.data(), (size_t)nprobe * 4));
  A->nprobe = nprobe;
  A->every = every;
  A->wgt_w = wgt_w;
  A->wgt_t = wgt_t;
  return HC_OK;
}

static HostState host_state(hc_ctx *ctx) {
  HostState h;
  h.time = ctx->time;
  for (int q = 0; q < 8; ++q) h.acc[q] = ctx->acc[q];
  h.next_step = ctx->next_step;
  h.map_valid = ctx->map_valid;
  for (int c = 0; c < 3; ++c) h.up_alias[c] = ctx->up_alias[c];
  return h;
}

static void set_host_state(hc_ctx *ctx, const HostState &h) {
  ctx->time = h.time;
  for (int q = 0; q < 8; ++q) ctx->acc[q] = h.acc[q];
  ctx->next_step = h.next_step;
  for (int c = 0; c < 3; ++c) ctx->up_alias[c] = h.up_alias[c];
  if (ctx->map_valid != h.map_valid) {
    ctx->map_valid = h.map_valid;
    update_sums(ctx);
  }
}

static void *slot_ptr(hc_ctx *ctx, size_t slot, int f) {
  return (char *)ctx->adj->tape + (slot * NTAPE + f) * ctx->L.n * ctx->esz;
}
// a reverse-sweep tape slot holds the 11 fields the step reads: its state
// is read by the adjoint (adj_head, adj_step) and never restored to run
// forward from, so the up, vp, wp buffers are not kept
static const int NTAPE_REV = 11;
static void save_slot(hc_ctx *ctx, size_t slot, int nf = NTAPE) {
  for (int f = 0; f < nf; ++f)
    hipMemcpyAsync(slot_ptr(ctx, slot, f), ctx->f[TAPE_FIELDS[f]], ctx->L.n * ctx->esz, hipMemcpyDeviceToDevice, ctx->st);
}
// the context stream waits for the tape copies in flight
static void copy_join(hc_ctx *ctx) {
  AdjCtx *A = ctx->adj;
  if (A && A->copy_pending) {
    hipStreamWaitEvent(ctx->st, A->evc[1], 0);
    A->copy_pending = false;
  }
}
// save_slot on the copy stream, after the work queued so far on the context
// stream; the fields it reads must not be written on the context stream
// before copy_join (fwd_step joins before PRESSURE, which is the first of a
// step's kernels to write a taped field).  Without the stream it copies in
// order on the context stream.
static void save_slot_async(hc_ctx *ctx, size_t slot, int nf = NTAPE) {
  AdjCtx *A = ctx->adj;
  if (!A->stc) {
    bool ok = tracer_streams(ctx) == HC_OK;
    for (hipEvent_t &e : A->evc) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (ok) {
      A->stc = ctx->tst[1];
    } else {
      (void)hipGetLastError();
      for (hipEvent_t &e : A->evc)
        if (e) hipEventDestroy(e), e = nullptr;
    }
  }
  if (!A->stc) {
    save_slot(ctx, slot, nf);
    return;
  }
  hipEventRecord(A->evc[0], ctx->st);
  hipStreamWaitEvent(A->stc, A->evc[0], 0);
  for (int f = 0; f < nf; ++f)
    hipMemcpyAsync(slot_ptr(ctx, slot, f), ctx->f[TAPE_FIELDS[f]], ctx->L.n * ctx->esz, hipMemcpyDeviceToDevice, A->stc);
  hipEventRecord(A->evc[1], A->stc);
  A->copy_pending = true;
}
static void load_slot(hc_ctx *ctx, size_t slot) {
  copy_join(ctx);
  for (int f = 0; f < NTAPE; ++f)
    hipMemcpyAsync(ctx->f[TAPE_FIELDS[f]], slot_ptr(ctx, slot, f), ctx->L.n * ctx->esz, hipMemcpyDeviceToDevice, ctx->st);
}

// the steps of the native loop up to the work velocities (no tracers)
static void step_head(hc_ctx *ctx, int m) {
  hc_begin_step(ctx, m, ctx->time + ctx->P.dt);
  hc_body_force(ctx);
  hc_forces(ctx);
  hc_pressure(ctx);
  hc_momentum_u(ctx);
  hc_momentum_v(ctx);
  hc_momentum_w(ctx);
  hc_viscosity3d(ctx);
  hc_source_terms(ctx, nullptr);
  hc_work_velocity(ctx, ctx->P.dt);
}

// The head of reverse step n without running it again.  Its inputs are the
// state X_n (tape slot s0), and the adjoint reads of its outputs only those
// that X_n+1 (slot s1) holds: VISCOSITY3D's ss and ss_v, which nothing later
// in the step changes, and MOMENTUM's up, vp, wp, which REFRESH made X_n+1's
// u, v, w (the buffers themselves).  The work velocities are formed from the
// two states.  The fields point into the tape for the step's adjoint (which
// writes no forward field); the caller restores the pointers afterwards.
// Host state advances as the head would advance it: step counters, body
// force, FORCES' entrance-ramp accumulators and the source terms.
// x0, x1: the NTAPE fields (TAPE_FIELDS order) of X_n and X_n+1
static void adj_head_p(hc_ctx *ctx, int m, void *const x0[NTAPE], void *const x1[NTAPE]) {
  hc_begin_step(ctx, m, ctx->time + ctx->P.dt);
  hc_body_force(ctx);
  for (int i = 0; i <= 7; ++i) ctx->acc[i] = ctx->acc[i] + (double)(7 - i) * .01 / 7.0;
  for (int f = 0; f < NTAPE; ++f) ctx->f[TAPE_FIELDS[f]] = x0[f];
  ctx->f[HC_SS] = x1[6];
  ctx->f[HC_SSV] = x1[7];
  for (int c = 0; c < 3; ++c) {
    ctx->f[HC_UP + c] = x1[c];
    ctx->up_alias[c] = false;
  }
  hc_source_terms(ctx, nullptr);
  hc_work_velocity(ctx, ctx->P.dt);
}
static void adj_head(hc_ctx *ctx, int m, size_t s0, size_t s1) {
  void *x0[NTAPE], *x1[NTAPE];
  for (int f = 0; f < NTAPE; ++f) {
    x0[f] = slot_ptr(ctx, s0, f);
    x1[f] = slot_ptr(ctx, s1, f);
  }
  adj_head_p(ctx, m, x0, x1);
}

// the taped intermediates of tracer `tr` (0: T, 1: S) of the segment's step
// index `q`
static TracerSlot tracer_slot(hc_ctx *ctx, int q, int tr) {
  const size_t fb = ctx->L.n * ctx->esz;
  char *b = (char *)ctx->adj->ttape + ((size_t)q * 2 + tr) * ttape_block(ctx);
  TracerSlot t;
  t.c1 = b;
  t.c2 = b + fb;
  for (int d = 0; d < 3; ++d) {
    t.ud2[d] = b + (2 + d) * fb;
    t.ud3[d] = b + (5 + d) * fb;
  }
  t.fl2 = (uint8_t *)(b + TTAPE * fb);
  t.fl3 = t.fl2 + (ctx->L.n + 255) / 256 * 256;
  return t;
}

// a tracer of the native loop whose passes leave their intermediates in
// slot `t` (and the pre-sponge boundary values in the adjoint's scratch);
// pass1_done: t.c1 holds pass 1's result already
static void taped_tracer(hc_ctx *ctx, int field, const TracerSlot &t, bool pass1_done = false) {
  const TracerWork keep = ctx->tw;
  ctx->tw.c1 = t.c1; ctx->tw.c2 = t.c2; ctx->tw.cpre = ctx->adj->sc[0].cpre;
  ctx->tw.ud2 = t.ud2[0]; ctx->tw.vd2 = t.ud2[1]; ctx->tw.wd2 = t.ud2[2];
  ctx->tw.ud3 = t.ud3[0]; ctx->tw.vd3 = t.ud3[1]; ctx->tw.wd3 = t.ud3[2];
  ctx->tw.fl2 = t.fl2; ctx->tw.fl3 = t.fl3;
  native_tracer(ctx, field, pass1_done ? t.c1 : nullptr);
  ctx->tw = keep;
}

// one forward step of the window (the native loop without the passive
// tracer c, which no loss here depends on, and without the CFL check);
// tq >= 0 tapes the tracers' intermediates at segment step index tq.
// Returns HC_ENOMEM when the pass-1 buffer of S cannot be allocated (as
// hc_step does), else 0.
static int fwd_step(hc_ctx *ctx, int m, int tq = -1) {
  const hc_params &P = ctx->P;
  const bool tape = tq >= 0 && ctx->adj && ctx->adj->ttape;
  if (!ctx->fuse) {
    copy_join(ctx);
    step_head(ctx, m);
    if (P.toggle_t == 1) {
      if (tape) taped_tracer(ctx, HC_T, tracer_slot(ctx, tq, 0));
      else native_tracer(ctx, HC_T);
    }
    if (P.toggle_s == 1) {
      if (tape) taped_tracer(ctx, HC_S, tracer_slot(ctx, tq, 1));
      else native_tracer(ctx, HC_S);
    }
    if (P.bit3 == 1) hc_density(ctx);
    hc_refresh(ctx);
    ctx->next_step = m + 1;
    return HC_OK;
  }
  // the native loop's fused path: MOMENTUM with the work velocities and
  // REFRESH's filter, pass 1 of T and S in one launch, REFRESH as a buffer
  // rotation (the same values as the path above)
  hc_begin_step(ctx, m, ctx->time + P.dt);
  hc_body_force(ctx);
  hc_forces(ctx);  // reads the taped fields only: the previous step's tape copies may still run
  copy_join(ctx);
  hc_pressure(ctx);
  for (int c = 0; c < 3; ++c) momentum_fused(ctx, c, true, P.dt);
  hc_viscosity3d(ctx);
  hc_source_terms(ctx, nullptr);
  const int on[2] = {P.toggle_t == 1, P.toggle_s == 1}, fid[2] = {HC_T, HC_S};
  TracerSlot ts[2];
  void *c1[2] = {nullptr, nullptr}, *dst[2];
  int fields[2], n = 0;
  for (int q = 0; q < 2; ++q) {
    if (!on[q]) continue;
    int rc = 0;
    if (tape) {
      ts[q] = tracer_slot(ctx, tq, q);
      c1[q] = ts[q].c1;
    } else {
      c1[q] = native_c1(ctx, q, &rc);
    }
    if (!c1[q]) return rc ? rc : fail(ctx, HC_ENOMEM, "pass-1 buffer: allocation failed");
    fields[n] = fid[q];
    dst[n++] = c1[q];
  }
  native_pass1(ctx, n, fields, dst);
  // S on the adjoint's second stream beside T (its own corrective-pass
  // scratch, and the taped slot's buffers when taping), DENSITY after both
  bool par = ctx->adj && ctx->par_tracers && !ctx->timing && on[0] && on[1];
  if (par && adj_par_init(ctx) != HC_OK) {
    par = false;
    (void)hipGetLastError();
  }
  if (par) {
    AdjCtx *A = ctx->adj;
    AdjTracerScratch &X = A->sc[1];
    TracerWork tw{};
    tw.uc = ctx->tw.uc; tw.vc = ctx->tw.vc; tw.wc = ctx->tw.wc;
    tw.budget_part = ctx->tw.budget_part; tw.budget = ctx->tw.budget;  // no budgets in a window's step
    tw.cpre = X.cpre;
    if (tape) {
      const TracerSlot &t = ts[1];
      tw.c2 = t.c2;
      tw.ud2 = t.ud2[0]; tw.vd2 = t.ud2[1]; tw.wd2 = t.ud2[2];
      tw.ud3 = t.ud3[0]; tw.vd3 = t.ud3[1]; tw.wd3 = t.ud3[2];
      tw.fl2 = t.fl2; tw.fl3 = t.fl3;
    } else {
      tw.c2 = X.c2;
      tw.ud2 = X.ud2[0]; tw.vd2 = X.ud2[1]; tw.wd2 = X.ud2[2];
    }
    hipEventRecord(A->ev[0], ctx->st);
    hipStreamWaitEvent(A->st2, A->ev[0], 0);
    native_tracer_on(ctx, HC_S, c1[1], A->st2, &tw);
    hipEventRecord(A->ev[1], A->st2);
    if (tape) taped_tracer(ctx, HC_T, ts[0], true);
    else native_tracer(ctx, HC_T, c1[0]);
    hipStreamWaitEvent(ctx->st, A->ev[1], 0);
  } else {
    for (int q = 0; q < 2; ++q) {
      if (!on[q]) continue;
      if (tape) taped_tracer(ctx, fid[q], ts[q], c1[q] != nullptr);
      else native_tracer(ctx, fid[q], c1[q]);
    }
  }
  if (P.bit3 == 1) hc_density(ctx);
  refresh_rotate(ctx);
  ctx->next_step = m + 1;
  return HC_OK;
}

// the adjoint of step m, with the live state holding the step's
// recomputed head (step_head from its taped state); t1/s1 are T and S after
// the step, ss0/ssv0 the viscosities at its start
template <typename T>
static void adj_step(hc_ctx *ctx, int m, const T *t1, const T *s1, const T *ss0, const T *ssv0, int tq = -1) {
  (void)m;
  const hc_params &P = ctx->P;
  adj_refresh_t<T>(ctx);
  if (P.bit3 == 1) adj_density_t<T>(ctx, t1, s1);
  double tmin, smax;
  native_bounds(ctx, &tmin, &smax);
  const int nz = P.nz;
  const bool taped = tq >= 0 && ctx->adj->ttape;
  TracerSlot ts[2];
  if (taped)
    for (int tr = 0; tr < 2; ++tr) ts[tr] = tracer_slot(ctx, tq, tr);
  // the two tracers' adjoints read and consume disjoint fields and add to
  // the shared adjoints (Courant numbers, ss, ss_v) by atomics only, so S
  // runs on a stream of its own beside T (with its own scratch)
  bool par = ctx->par_tracers && P.toggle_s == 1 && P.toggle_t == 1;
  if (par && adj_par_init(ctx) != HC_OK) {
    par = false;  // no memory for the second scratch: one stream
    (void)hipGetLastError();
  }
  AdjCtx *A = ctx->adj;
  if (par) {
    hipEventRecord(A->ev[0], ctx->st);
    hipStreamWaitEvent(A->st2, A->ev[0], 0);
  }
  if (P.toggle_s == 1)
    adj_tracer_t<T>(ctx, HC_S, ctx->d_sbkg, 0.0, smax, P.dt, 1, ctx->d_q[1], taped ? &ts[1] : nullptr, par ? 1 : 0);
  if (par) hipEventRecord(A->ev[1], A->st2);
  if (P.toggle_t == 1)
    adj_tracer_t<T>(ctx, HC_T, ctx->d_tbkg, tmin, 1.0e+10, P.dt, 1, ctx->d_q[0], taped ? &ts[0] : nullptr);
  if (par) hipStreamWaitEvent(ctx->st, A->ev[1], 0);
  (void)nz;
  adj_courant_t<T>(ctx, P.dt);
  adj_viscosity_t<T>(ctx, (const T *)ctx->f[HC_SS], (const T *)ctx->f[HC_SSV]);
  adj_momentum_t<T>(ctx, 2, false);
  adj_momentum_t<T>(ctx, 1, false);
  adj_momentum_t<T>(ctx, 0, false);
  {
    FieldPtrs<T> lam = lptrs<T>(ctx);
    void *const ups[3] = {lam.up, lam.vp, lam.wp};  // consumed by the three
    zero_fields(ctx, ctx->st, ups, 3);
  }
  adj_pressure_t<T>(ctx);
  adj_forces_t<T>(ctx, ss0, ssv0);
}

template <typename T>
static void probe_sample(hc_ctx *ctx, size_t q) {
  AdjCtx *A = ctx->adj;
  dim3 g((A->nprobe + 255) / 256);
  if (sizeof(T) == 8)
    hipLaunchKernelGGL(k_probe_sample, g, dim3(256), 0, ctx->st, A->nprobe, A->d_iw, A->d_it,
                       (const double *)ctx->f[HC_W], (const double *)ctx->f[HC_T], A->d_samp + 2 * A->nprobe * q);
  else
    hipLaunchKernelGGL(k_probe_sample32, g, dim3(256), 0, ctx->st, A->nprobe, A->d_iw, A->d_it,
                       (const float *)ctx->f[HC_W], (const float *)ctx->f[HC_T], A->d_samp + 2 * A->nprobe * q);
}

template <typename T>
static int window_t(hc_ctx *ctx, int nsteps, int seg, const double *obs, double *samples, double *loss, double *gq,
                    double gs[3], int end_state) {
  AdjCtx *A = ctx->adj;
  const int m0 = ctx->next_step;
  const int nck = (nsteps + seg - 1) / seg;  // checkpoints at n = 0, seg, ...
  const size_t slots = (size_t)nck + 1 + (size_t)seg;
  const size_t slot_bytes = (size_t)NTAPE * ctx->L.n * ctx->esz;
  if (A->tape_slots < slots) {
    if (A->tape) { hipStreamSynchronize(ctx->st); hipFree(A->tape); A->tape = nullptr; A->tape_slots = 0; }
    if (hipMalloc(&A->tape, slots * slot_bytes) != hipSuccess) {
      A->tape = nullptr;
      return fail(ctx, HC_ENOMEM, "checkpoint tape: hipMalloc of " + std::to_string(slots * slot_bytes) +
                                      " bytes failed (use a smaller segment length)");
    }
    A->tape_slots = slots;
  }
  const int nsamp = nsteps / A->every;
  const size_t ns = (size_t)nsamp * 2 * A->nprobe;
  if (A->nsamp_cap < ns) {
    if (A->d_samp) { hipFree(A->d_samp); hipFree(A->d_obs); }
    int rc = 0;
    if (!(A->d_samp = (double *)dalloc(ctx, ns * 8, &rc))) return rc;
    if (!(A->d_obs = (double *)dalloc(ctx, ns * 8, &rc))) return rc;
    A->nsamp_cap = ns;
  }
  if (ns) HIPCHK(hipMemcpyAsync(A->d_obs, obs, ns * 8, hipMemcpyHostToDevice, ctx->st));
  // the tracers' intermediates of a segment's recomputed steps; without room
  // for them the adjoint recomputes the tracer passes step by step
  const size_t tsteps = (size_t)std::max(seg - 1, 0);
  if (A->ttape_steps < tsteps) {
    if (A->ttape) { hipStreamSynchronize(ctx->st); hipFree(A->ttape); A->ttape = nullptr; A->ttape_steps = 0; }
    // zeroed once: the passes leave the ghost faces of the antidiffusive
    // velocities unwritten, and the adjoint kernels read them (with zero
    // weight), so they must hold finite values
    const size_t tb = tsteps * 2 * ttape_block(ctx);
    if (hipMalloc(&A->ttape, tb) == hipSuccess) {
      A->ttape_steps = tsteps;
      HIPCHK(hipMemsetAsync(A->ttape, 0, tb, ctx->st));
    } else {
      A->ttape = nullptr;
      (void)hipGetLastError();
    }
  }
  const size_t tape0 = (size_t)nck + 1;  // first slot of the segment tape
  std::vector<HostState> ck(nck + 1), tp(seg + 1);
  // forward with checkpoints; the last segment, which the reverse sweep
  // takes first, is taped here as its recompute would tape it (its steps
  // but the last, with the tracers' intermediates), so it is not recomputed
  const int alast = (nck - 1) * seg;
  for (int n = 0; n < nsteps; ++n) {
    if (n % seg == 0) {
      save_slot_async(ctx, n / seg);
      ck[n / seg] = host_state(ctx);
    }
    int rs;
    if (n >= alast && n < nsteps - 1) {
      rs = fwd_step(ctx, m0 + n, n - alast);
      save_slot_async(ctx, tape0 + (n + 1 - alast), NTAPE_REV);
      tp[n + 1 - alast] = host_state(ctx);
    } else {
      rs = fwd_step(ctx, m0 + n);
    }
    if (rs) {
      copy_join(ctx);
      return rs;
    }
    if ((n + 1) % A->every == 0) probe_sample<T>(ctx, (size_t)((n + 1) / A->every - 1));
  }
  copy_join(ctx);
  save_slot(ctx, nck);
  ck[nck] = host_state(ctx);
  // the last segment's first state, also when that segment is a single step
  // (nsteps % seg == 1), whose forward step the taping branch above skips
  tp[0] = ck[nck - 1];
  std::vector<double> samp(ns);
  if (ns) HIPCHK(hipMemcpyAsync(samp.data(), A->d_samp, ns * 8, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  double L2 = 0.0;
  for (size_t q = 0; q < ns; q += 2) {
    const double dw = samp[q] - obs[q], dtt = samp[q + 1] - obs[q + 1];
    L2 += 0.5 * (A->wgt_w * dw * dw + A->wgt_t * dtt * dtt);
  }
  *loss = L2;
  if (samples && ns) memcpy(samples, samp.data(), ns * 8);
  // reverse sweep
  int r = hc_adj_reset(ctx);
  if (r) return r;

  for (int sg = nck - 1; sg >= 0; --sg) {
    const int a = sg * seg, b = std::min(nsteps, a + seg);
    // the segment's states: X_a is checkpoint sg, X_b checkpoint sg + 1,
    // X_a+1 .. X_b-1 recomputed into the tape
    auto slot_of = [&](int n) -> size_t {
      return n == a ? (size_t)sg : (n == b ? (size_t)sg + 1 : tape0 + (size_t)(n - a));
    };
    if (sg != nck - 1) {  // the last segment was taped by the forward sweep
      load_slot(ctx, sg);
      set_host_state(ctx, ck[sg]);
      tp[0] = ck[sg];
      for (int n = a; n < b - 1; ++n) {
        const int rs = fwd_step(ctx, m0 + n, n - a);
        if (rs) {
          copy_join(ctx);
          return rs;
        }
        save_slot_async(ctx, tape0 + (n + 1 - a), NTAPE_REV);
        tp[n + 1 - a] = host_state(ctx);
      }
      copy_join(ctx);  // the adjoint steps read the tape
    }
    for (int n = b - 1; n >= a; --n) {
      FieldPtrs<T> lam = lptrs<T>(ctx);
      if ((n + 1) % A->every == 0) {
        const size_t q = (size_t)((n + 1) / A->every - 1);
        hipLaunchKernelGGL(k_probe_seed<T>, dim3((A->nprobe + 255) / 256), dim3(256), 0, ctx->st, A->nprobe, A->d_iw,
                           A->d_it, A->d_samp + 2 * A->nprobe * q, A->d_obs + 2 * A->nprobe * q, A->wgt_w, A->wgt_t,
                           lam.w, lam.t);
      }
      set_host_state(ctx, tp[n - a]);
      const size_t s1 = slot_of(n + 1), s0 = slot_of(n);
      void *own[HC_NFIELDS];  // the field buffers (the fused forward rotates them)
      for (int f = 0; f < HC_NFIELDS; ++f) own[f] = ctx->f[f];
      adj_head(ctx, m0 + n, s0, s1);
      // step b-1 is not recomputed (its end state is X_b): its tracers'
      // intermediates are recomputed in the adjoint
      adj_step<T>(ctx, m0 + n, (const T *)slot_ptr(ctx, s1, 9), (const T *)slot_ptr(ctx, s1, 10),
                  (const T *)slot_ptr(ctx, s0, 6), (const T *)slot_ptr(ctx, s0, 7), n < b - 1 ? n - a : -1);
      for (int f = 0; f < HC_NFIELDS; ++f) ctx->f[f] = own[f];
    }
  }
  // the window's end state (to continue the run) or its initial state (to
  // evaluate the window again)
  const int keep = end_state ? nck : 0;
  load_slot(ctx, keep);
  set_host_state(ctx, ck[keep]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, HC_EDEVICE, std::string("adjoint window: ") + hipGetErrorString(e));
  return hc_adj_gradients(ctx, gq, gs);
}

// ---------------------------------------------------------------------------
// binomial checkpointing (Griewank 1992; Griewank & Walther 2000, "Algorithm
// 799: revolve"): the window's reversal under a fixed number of snapshot
// slots with the fewest forward steps.  reverse(a, b, s) reverses steps
// a .. b-1 from X_a, which a slot holds and the live state equals, with s
// free slots: one step is a turn (the step run again with its tracers'
// intermediates taped, then its adjoint); more steps advance to the split
// point m, take a snapshot of X_m, reverse(m, b, s-1), restore X_a and
// reverse(a, m, s).  With C(l, s) the advances this takes for l steps,
//   C(1, s) = 0,  C(l, 0) = l(l-1)/2,
//   C(l, s) = min over 1 <= m < l of m + C(l-m, s-1) + C(m, s)
//           = t l - beta(s+2, t-1),  t the least with beta(s+1, t) >= l,
// beta(s, t) = (s+t)! / (s! t!); the split is a minimiser of the bracket.
// Without free slots the steps are turned last to first, each from X_a.

// beta(s, t) capped at cap (the counts only need comparing with l)
static long long binom_beta(long long s, long long t, long long cap) {
  long long b = 1;
  for (long long i = 1; i <= t; ++i) {
    b = b * (s + i) / i;  // exact: b is C(s+i, i) after each step
    if (b > cap) return cap + 1;
  }
  return b;
}

static long long binom_cost(long long l, long long s) {
  if (l <= 1) return 0;
  if (s <= 0) return l * (l - 1) / 2;
  long long t = 1;
  while (binom_beta(s + 1, t, l) < l) ++t;
  return t * l - binom_beta(s + 2, t - 1, (long long)4e18);
}

static long long binom_split(long long l, long long s) {
  long long best = 1, bc = -1;
  for (long long m = 1; m < l; ++m) {
    const long long c = m + binom_cost(l - m, s - 1) + binom_cost(m, s);
    if (bc < 0 || c < bc) { bc = c; best = m; }
  }
  return best;
}

// actions: {kind, step, slot, 0}; kind 0: advance the live state by one step
// (step = the step run), 1: snapshot of the live state (X_step) into slot,
// 2: restore slot (X_step), 3: turn step `step` from X_step held in slot
enum { BA_ADVANCE = 0, BA_SNAP = 1, BA_RESTORE = 2, BA_TURN = 3 };
struct BinomAct {
  int kind, step, slot;
};

// slots: 0 .. nsnap-1 snapshots (slot 0 holds X_0), nsnap the temporary of
// the turns without a free slot
static void binom_reverse(std::vector<BinomAct> &out, int a, int b, int s, int slot_a, int next, int temp) {
  if (b - a == 1) {
    out.push_back({BA_TURN, a, slot_a});
    return;
  }
  if (s == 0) {
    for (int n = b - 1; n >= a; --n) {
      if (n != b - 1) out.push_back({BA_RESTORE, a, slot_a});
      for (int q = a; q < n; ++q) out.push_back({BA_ADVANCE, q, -1});
      if (n == a) {
        out.push_back({BA_TURN, a, slot_a});
      } else {
        out.push_back({BA_SNAP, n, temp});
        out.push_back({BA_TURN, n, temp});
      }
    }
    return;
  }
  const int m = a + (int)binom_split(b - a, s);
  for (int q = a; q < m; ++q) out.push_back({BA_ADVANCE, q, -1});
  out.push_back({BA_SNAP, m, next});
  binom_reverse(out, m, b, s - 1, next, next + 1, temp);
  out.push_back({BA_RESTORE, a, slot_a});
  binom_reverse(out, a, m, s, slot_a, next, temp);
}

static std::vector<BinomAct> binom_schedule(int nsteps, int nsnap) {
  std::vector<BinomAct> out;
  out.push_back({BA_SNAP, 0, 0});
  binom_reverse(out, 0, nsteps, nsnap - 1, 0, 1, nsnap);
  return out;
}

// The schedule for nsteps steps and nsnap snapshot slots (host only, no
// device): up to cap actions of 4 int32 (kind, step, slot, 0) into acts
// (may be null), their number into nacts and the forward steps it runs
// (advances and turns) into nfwd.
extern "C" int hc_binomial_schedule(int32_t nsteps, int32_t nsnap, int32_t *acts, int64_t cap, int64_t *nacts,
                                    int64_t *nfwd) {
  if (nsteps < 1 || nsnap < 1) return HC_EINVAL;
  const std::vector<BinomAct> v = binom_schedule(nsteps, nsnap);
  int64_t nf = 0;
  for (size_t q = 0; q < v.size(); ++q) {
    if (v[q].kind == BA_ADVANCE || v[q].kind == BA_TURN) ++nf;
    if (acts && (int64_t)q < cap) {
      acts[4 * q] = v[q].kind;
      acts[4 * q + 1] = v[q].step;
      acts[4 * q + 2] = v[q].slot;
      acts[4 * q + 3] = 0;
    }
  }
  if (nacts) *nacts = (int64_t)v.size();
  if (nfwd) *nfwd = nf;
  return HC_OK;
}

template <typename T>
static int window_binomial_t(hc_ctx *ctx, int nsteps, int nsnap, const double *obs, double *samples, double *loss,
                             double *gq, double gs[3], int end_state) {
  AdjCtx *A = ctx->adj;
  const int m0 = ctx->next_step;
  // slots: nsnap snapshots, the turns' temporary, the window's end state
  const size_t slots = (size_t)nsnap + 2, temp = (size_t)nsnap, endslot = (size_t)nsnap + 1;
  const size_t slot_bytes = (size_t)NTAPE * ctx->L.n * ctx->esz;
  if (A->tape_slots < slots) {
    if (A->tape) { hipStreamSynchronize(ctx->st); hipFree(A->tape); A->tape = nullptr; A->tape_slots = 0; }
    if (hipMalloc(&A->tape, slots * slot_bytes) != hipSuccess) {
      A->tape = nullptr;
      return fail(ctx, HC_ENOMEM, "snapshot slots: hipMalloc of " + std::to_string(slots * slot_bytes) +
                                      " bytes failed (use fewer snapshots)");
    }
    A->tape_slots = slots;
  }
  const int nsamp = nsteps / A->every;
  const size_t ns = (size_t)nsamp * 2 * A->nprobe;
  if (A->nsamp_cap < ns) {
    if (A->d_samp) { hipFree(A->d_samp); hipFree(A->d_obs); }
    int rc = 0;
    if (!(A->d_samp = (double *)dalloc(ctx, ns * 8, &rc))) return rc;
    if (!(A->d_obs = (double *)dalloc(ctx, ns * 8, &rc))) return rc;
    A->nsamp_cap = ns;
  }
  if (ns) HIPCHK(hipMemcpyAsync(A->d_obs, obs, ns * 8, hipMemcpyHostToDevice, ctx->st));
  if (A->ttape_steps < 1) {  // one step of the tracers' intermediates, for the turns
    if (A->ttape) { hipStreamSynchronize(ctx->st); hipFree(A->ttape); A->ttape = nullptr; A->ttape_steps = 0; }
    const size_t tb = 2 * ttape_block(ctx);
    if (hipMalloc(&A->ttape, tb) == hipSuccess) {
      A->ttape_steps = 1;
      HIPCHK(hipMemsetAsync(A->ttape, 0, tb, ctx->st));  // finite ghost faces, as in the fixed-interval window
    } else {
      A->ttape = nullptr;
      (void)hipGetLastError();
    }
  }
  const std::vector<BinomAct> plan = binom_schedule(nsteps, nsnap);
  std::vector<HostState> hs(slots);
  HostState h_end{};
  int done = 0;        // states X_1 .. X_done computed (and sampled) so far
  int live = 0;        // the step of the live state
  bool reset = false;  // the adjoint fields zeroed (at the first turn, after every sample exists)
  auto sample = [&](int n1) {  // X_n1 was computed just now
    if (n1 > done) {
      if (n1 % A->every == 0) probe_sample<T>(ctx, (size_t)(n1 / A->every - 1));
      done = n1;
    }
  };
  for (const BinomAct &ac : plan) {
    if (ac.kind == BA_ADVANCE) {
      const int rs = fwd_step(ctx, m0 + ac.step);
      if (rs) return rs;
      live = ac.step + 1;
      sample(live);
      if (live == nsteps) {
        save_slot(ctx, endslot);
        h_end = host_state(ctx);
      }
    } else if (ac.kind == BA_SNAP) {
      save_slot(ctx, (size_t)ac.slot);
      hs[ac.slot] = host_state(ctx);
    } else if (ac.kind == BA_RESTORE) {
      load_slot(ctx, (size_t)ac.slot);
      set_host_state(ctx, hs[ac.slot]);
      live = ac.step;
    } else {  // turn: step n from X_n (slot), the adjoint reads X_n+1 from the live buffers
      const int n = ac.step;
      const HostState h0 = hs[ac.slot];
      const int rs = fwd_step(ctx, m0 + n, 0);
      if (rs) return rs;
      sample(n + 1);
      if (n + 1 == nsteps) {
        save_slot(ctx, endslot);
        h_end = host_state(ctx);
      }
      if (!reset) {  // the last step is turned first: every sample exists now
        const int r = hc_adj_reset(ctx);
        if (r) return r;
        reset = true;
      }
      FieldPtrs<T> lam = lptrs<T>(ctx);
      if ((n + 1) % A->every == 0) {
        const size_t q = (size_t)((n + 1) / A->every - 1);
        hipLaunchKernelGGL(k_probe_seed<T>, dim3((A->nprobe + 255) / 256), dim3(256), 0, ctx->st, A->nprobe, A->d_iw,
                           A->d_it, A->d_samp + 2 * A->nprobe * q, A->d_obs + 2 * A->nprobe * q, A->wgt_w, A->wgt_t,
                           lam.w, lam.t);
      }
      void *own[HC_NFIELDS];  // X_n+1 in the field buffers (the fused forward rotates them)
      for (int f = 0; f < HC_NFIELDS; ++f) own[f] = ctx->f[f];
      void *x0[NTAPE], *x1[NTAPE];
      for (int f = 0; f < NTAPE; ++f) {
        x0[f] = slot_ptr(ctx, (size_t)ac.slot, f);
        x1[f] = own[TAPE_FIELDS[f]];
      }
      set_host_state(ctx, h0);
      adj_head_p(ctx, m0 + n, x0, x1);
      adj_step<T>(ctx, m0 + n, (const T *)x1[9], (const T *)x1[10], (const T *)x0[6], (const T *)x0[7], 0);
      for (int f = 0; f < HC_NFIELDS; ++f) ctx->f[f] = own[f];
      live = -1;  // the live buffers are X_n+1, but the host state is X_n's: restore before use
    }
  }
  std::vector<double> samp(ns);
  if (ns) HIPCHK(hipMemcpyAsync(samp.data(), A->d_samp, ns * 8, hipMemcpyDeviceToHost, ctx->st));
  HIPCHK(hipStreamSynchronize(ctx->st));
  double L2 = 0.0;
  for (size_t q = 0; q < ns; q += 2) {
    const double dw = samp[q] - obs[q], dtt = samp[q + 1] - obs[q + 1];
    L2 += 0.5 * (A->wgt_w * dw * dw + A->wgt_t * dtt * dtt);
  }
  *loss = L2;
  if (samples && ns) memcpy(samples, samp.data(), ns * 8);
  (void)live;
  if (end_state) {
    load_slot(ctx, endslot);
    set_host_state(ctx, h_end);
  } else {
    load_slot(ctx, 0);
    set_host_state(ctx, hs[0]);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, HC_EDEVICE, std::string("adjoint window: ") + hipGetErrorString(e));
  return hc_adj_gradients(ctx, gq, gs);
}

// HBM the window's checkpoints and tapes take (beyond the context's own
// state and the adjoint fields): (nck + 1 + seg) states of NTAPE fields and
// (seg - 1) steps of TTAPE fields and two flag planes for each of the two
// tracers
extern "C" int hc_adj_window_bytes(hc_ctx *ctx, int32_t nsteps, int32_t seg, uint64_t *bytes) {
  if (!ctx || !bytes || nsteps < 1 || seg < 1) return fail(ctx, HC_EINVAL, "nsteps and seg must be positive");
  const uint64_t fb = (uint64_t)ctx->L.n * ctx->esz;
  const uint64_t nck = (uint64_t)((nsteps + seg - 1) / seg);
  *bytes = (nck + 1 + (uint64_t)seg) * NTAPE * fb + (uint64_t)(seg - 1) * 2 * ttape_block(ctx);
  return HC_OK;
}

// HBM the binomial window takes: nsnap + 2 states of NTAPE fields (the
// snapshots, the turns' temporary, the end state) and one step of TTAPE
// fields and two flag planes for each of the two tracers
extern "C" int hc_adj_window_binomial_bytes(hc_ctx *ctx, int32_t nsteps, int32_t nsnap, uint64_t *bytes) {
  if (!ctx || !bytes || nsteps < 1 || nsnap < 1) return fail(ctx, HC_EINVAL, "nsteps and nsnap must be positive");
  const uint64_t fb = (uint64_t)ctx->L.n * ctx->esz;
  *bytes = ((uint64_t)nsnap + 2) * NTAPE * fb + 2 * ttape_block(ctx);
  return HC_OK;
}

extern "C" int hc_adj_window_binomial(hc_ctx *ctx, int32_t nsteps, int32_t nsnap, const double *obs,
                                      double *samples, double *loss, double *gq, double gs[3], int32_t end_state) {
  int r = adj_need(ctx);
  if (r) return r;
  if (ctx->h_tbkg.empty()) return fail(ctx, HC_EINVAL, "profiles not set (hc_init_case or hc_set_profiles)");
  if (!ctx->adj->d_iw) return fail(ctx, HC_EINVAL, "no probes (hc_adj_set_probes)");
  if (nsteps < 1 || nsnap < 1) return fail(ctx, HC_EINVAL, "nsteps and nsnap must be positive");
  if (ctx->P.ismol != 1) return fail(ctx, HC_EINVAL, "only ismol = 1 (two corrective passes) is supported");
  PhaseScope ps(ctx, "adj_window");
  if (ctx->prec == HC_FP64) return window_binomial_t<double>(ctx, nsteps, nsnap, obs, samples, loss, gq, gs, end_state);
  return window_binomial_t<float>(ctx, nsteps, nsnap, obs, samples, loss, gq, gs, end_state);
}

extern "C" int hc_adj_window(hc_ctx *ctx, int32_t nsteps, int32_t seg, const double *obs, double *samples,
                             double *loss, double *gq, double gs[3], int32_t end_state) {
  int r = adj_need(ctx);
  if (r) return r;
  if (ctx->h_tbkg.empty()) return fail(ctx, HC_EINVAL, "profiles not set (hc_init_case or hc_set_profiles)");
  if (!ctx->adj->d_iw) return fail(ctx, HC_EINVAL, "no probes (hc_adj_set_probes)");
  if (nsteps < 1 || seg < 1) return fail(ctx, HC_EINVAL, "nsteps and seg must be positive");
  if (ctx->P.ismol != 1) return fail(ctx, HC_EINVAL, "only ismol = 1 (two corrective passes) is supported");
  PhaseScope ps(ctx, "adj_window");
  if (ctx->prec == HC_FP64) return window_t<double>(ctx, nsteps, seg, obs, samples, loss, gq, gs, end_state);
  return window_t<float>(ctx, nsteps, seg, obs, samples, loss, gq, gs, end_state);
}
