// hc_tracer.hip - TRACER31 / MPDATA / BOUNDARY_CONDITIONS on gfx950.
//
// mpdata-hotcross80.f:16-494.  One tracer call is
//   pass 1: explicit diffusion + donor-cell upwind step      (k_pass1_tile)
//   pass 2/3: antidiffusive velocities, limiter flags and the limited
//            donor-cell step in one k-marching LDS tile (k_mpdata_pass,
//            hc_mpdata.h)
//   sponge + source on the last pass, then the boundary images (k_bc).
// The diffusivities kxt/kyt/kzt of MIXING (mixing80-revised2014.f:22-59) are
// formed inside pass 1 from ss/ss_v with the reference's expressions.
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
#include <stdlib.h>

#include "hc_common.h"
#include "hc_kernels.h"
#include "hc_cell.h"
#include "hc_mpdata.h"
#include "hc_mpdata2.h"
#include "hc_mpdata3.h"
#include "hc_tile.h"

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

// pass 1 as a k-marching LDS tile (planes k-1, k, k+1 of the six inputs)
template <typename T, int TI, int TJ>
__global__ void __launch_bounds__(TI *TJ) k_pass1_tile(Layout L, StepArgs s, Phys ph, const T *__restrict__ c,
                                                       const T *__restrict__ uc, const T *__restrict__ vc,
                                                       const T *__restrict__ wc, const T *__restrict__ ss,
                                                       const T *__restrict__ ssv, const double *__restrict__ azzmin,
                                                       const double *__restrict__ cbkg, T *__restrict__ c1, int kc,
                                                       int remap) {
  __shared__ Ring3<T, TI, TJ, 6> sh;
  unsigned bx, by, bz;
  box_block_coords(remap, bx, by, bz);
  const int i0 = 1 + (int)bx * TI, j0 = 1 + (int)by * TJ;
  const int ka = 1 + (int)bz * kc, kb = min(L.nz, ka + kc - 1);
  const T *src[6] = {c, uc, vc, wc, ss, ssv};
  ring3_march(sh, L, src, i0, j0, ka, kb, [&](int i, int j, int k, auto &view) {
    F3(c1, i, j, k) = pass1_at<T>(L, s, ph, view(0), view(1), view(2), view(3), view(4), view(5), azzmin, cbkg, i,
                                  j, k);
  });
}

// Pass 1 of NC tracers in one k-march.  The tracers of a native step read the
// same Courant numbers and viscosities, so uc, vc, wc, ss and ss_v are staged
// once for all of them: 5 + NC field moves in and NC out per cell instead of
// 7 per tracer.  LDS holds planes k and k+1 of every field with a one-cell
// halo in i and j; pass 1 reads plane k-1 only in the cell's own column
// (c, wc, ss_v), so that value rolls in a register.
template <typename T, int TI, int TJ, int NF>
struct Ring2 {
  static constexpr int SI = TI + 2, S = SI * (TJ + 2), NT = TI * TJ, RS = (S + NT - 1) / NT;
  T f[NF][2][S];
};

// planes k (p0) and k+1 (pp) in LDS; m = the own column's value at plane k-1
// (a read at plane k-1 must be at the own column)
template <typename T, int SI>
struct R2View {
  const T *p0, *pp;
  T m;
  int k, ib, jb;
  __device__ T operator()(int i, int j, int kk) const {
    if (kk < k) return m;
    return (kk > k ? pp : p0)[(i - ib) + SI * (j - jb)];
  }
};

template <typename T>
struct Pass1Set {
  const T *src[8];       // uc, vc, wc, ss, ss_v, then the tracers
  const double *cbkg[3];
  T *out[3];
};

template <typename T, int TI, int TJ, int NC>
__global__ void __launch_bounds__(TI *TJ) k_pass1_multi(Layout L, StepArgs s, Phys ph, Pass1Set<T> set,
                                                        const double *__restrict__ azzmin, int kc, int remap) {
  constexpr int NF = 5 + NC;
  typedef Ring2<T, TI, TJ, NF> R;
  constexpr int SI = R::SI, S = R::S, NT = R::NT, RS = R::RS;
  __shared__ R sh;
  unsigned bx, by, bz;
  box_block_coords(remap, bx, by, bz);
  const int i0 = 1 + (int)bx * TI, j0 = 1 + (int)by * TJ;
  const int ka = 1 + (int)bz * kc, kb = min(L.nz, ka + kc - 1);
  const int tid = (int)threadIdx.x;
  int so[RS];
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int e = tid + r * NT;
    const int i = i0 - 1 + e % SI, j = j0 - 1 + e / SI;
    so[r] = (e < S && i <= L.nx + 1 && j <= L.ny + 1) ? (int)IX(L, i, j, -1) : -1;
  }
  const long pstride = (long)L.px * L.py;
  T pf[NF][RS];
  auto fetch = [&](int k) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int r = 0; r < RS; ++r) pf[f][r] = so[r] >= 0 ? set.src[f][so[r] + (k + 1) * pstride] : (T)0;
  };
  auto commit = [&](int k) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        const int e = tid + r * NT;
        if (e < S) sh.f[f][k & 1][e] = pf[f][r];
      }
  };
  const int i = i0 + tid % TI, j = j0 + tid / TI;
  const bool own = i <= L.nx && j <= L.ny;
  const int cp = (tid % TI + 1) + SI * (tid / TI + 1);  // own cell in the LDS plane
  T pm[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) pm[f] = own ? set.src[f][IX(L, i, j, ka - 1)] : (T)0;
  fetch(ka);
  commit(ka);
  fetch(ka + 1);  // ka + 1 <= nz + 1
  commit(ka + 1);
  if (ka + 2 <= kb + 1) fetch(ka + 2);
  __syncthreads();
  for (int k = ka; k <= kb; ++k) {
    if (own) {
      auto view = [&](int f) { return R2View<T, SI>{sh.f[f][k & 1], sh.f[f][(k + 1) & 1], pm[f], k, i0 - 1, j0 - 1}; };
#pragma unroll
      for (int q = 0; q < NC; ++q)
        F3(set.out[q], i, j, k) = pass1_at<T>(L, s, ph, view(5 + q), view(0), view(1), view(2), view(3), view(4),
                                              azzmin, set.cbkg[q], i, j, k);
#pragma unroll
      for (int f = 0; f < NF; ++f) pm[f] = sh.f[f][k & 1][cp];
    }
    if (k < kb) {
      __syncthreads();  // every read of plane k's slot is done
      commit(k + 2);    // k + 2 <= kb + 1
      __syncthreads();
      if (k + 3 <= kb + 1) fetch(k + 3);  // in flight during plane k + 1
    }
  }
}

// one corrective pass (antidiffusive velocities, limiter flags, limited
// donor-cell step; sponge and source on the last pass) as a k-marching tile
template <typename T, int TI, int TJ>
__global__ void __launch_bounds__(TI *TJ) k_mpdata_pass(Layout L, MpPass<T> a, int kc, int remap) {
  typedef MpTile<T, TI, TJ> Tl;
  __shared__ MpShared<T, TI, TJ> sh;
  unsigned bx, by, bz;
  box_block_coords(remap, bx, by, bz);  // XCD x gets the x-th slice of tiles
  MpGeom g;
  g.nx = L.nx; g.ny = L.ny; g.nz = L.nz; g.px = L.px; g.py = L.py;
  g.i0 = 1 + (int)bx * TI;
  g.j0 = 1 + (int)by * TJ;
  g.ka = 1 + (int)bz * kc;
  g.kb = min(L.nz, g.ka + kc - 1);
  const int tid = (int)threadIdx.x;
  const int qs = max(g.ka - 2, 0);
  MpRegs<T> r{};
  typename Tl::Plan pl;
  typename Tl::Prefetch pf;
  Tl::plan(g, tid, (int)(bx + by + bz), pl);
  Tl::stage_initial(g, tid, pl.so, sh, a, qs);
  Tl::fetch(g, pl.so, a, qs, pf);
  for (int q = qs; q <= g.kb + 1; ++q) {
    // no barrier before the staging: it overwrites the slots of c(q-3),
    // U(q-1), V(q-1), W(q-2), which the previous plane's finishing step does
    // not read and its faces and flags phases finished reading before the
    // barrier ahead of that step
    Tl::commit(tid, sh, q, pf);
    __syncthreads();
    if (q <= g.kb) Tl::fetch(g, pl.so, a, q + 1, pf);  // in flight during this plane
    Tl::faces(g, pl, sh, a, q, q >= g.ka - 1);
    __syncthreads();
    if (q >= g.ka - 1) Tl::flags(g, pl, sh, a, q);
    __syncthreads();
    Tl::finish(g, pl, sh, a, q, r);
  }
}

// the block executor of mp2_block on the device: a phase is each thread's
// own call, the barrier is __syncthreads
template <typename T, int TI, int TJ>
struct DevExec {
  Mp2State<T, TI, TJ> st;
  template <class F>
  __device__ void each(F f) { f((int)threadIdx.x, st); }
  __device__ void sync() { __syncthreads(); }
};

// the same pass with the fixed per-thread assignment of MpTile2; at least
// four waves per SIMD (two 512-thread blocks per CU, what the LDS of a
// 32 x 16 FP64 tile allows), which caps the registers at 128
template <typename T, int TI, int TJ>
__global__ void __launch_bounds__(TI *TJ, 4) k_mpdata_pass2(Layout L, MpPass<T> a, int kc, int remap) {
  __shared__ MpShared<T, TI, TJ> sh;
  unsigned bx, by, bz;
  box_block_coords(remap, bx, by, bz);  // XCD x gets the x-th slice of tiles
  MpGeom g;
  g.nx = L.nx; g.ny = L.ny; g.nz = L.nz; g.px = L.px; g.py = L.py;
  g.i0 = 1 + (int)bx * TI;
  g.j0 = 1 + (int)by * TJ;
  g.ka = 1 + (int)bz * kc;
  g.kb = min(L.nz, g.ka + kc - 1);
  DevExec<T, TI, TJ> x;
  mp2_block<T, TI, TJ>(x, g, sh, a);
}

// the pass with each face's donor-cell flux formed once (MpTile3): the
// fluxes of the flags and the limited fluxes of the finishing step come from
// products formed right after the face velocities
template <typename T, int TI, int TJ>
struct DevExec3 {
  Mp3State<T, TI, TJ> st;
  template <class F>
  __device__ void each(F f) { f((int)threadIdx.x, st); }
  __device__ void sync() { __syncthreads(); }
};

template <typename T, int TI, int TJ>
__global__ void __launch_bounds__(TI *TJ, 4) k_mpdata_pass3(Layout L, MpPass<T> a, int kc, int remap) {
  __shared__ Mp3Shared<T, TI, TJ> sh;
  unsigned bx, by, bz;
  box_block_coords(remap, bx, by, bz);  // XCD x gets the x-th slice of tiles
  MpGeom g;
  g.nx = L.nx; g.ny = L.ny; g.nz = L.nz; g.px = L.px; g.py = L.py;
  g.i0 = 1 + (int)bx * TI;
  g.j0 = 1 + (int)by * TJ;
  g.ka = 1 + (int)bz * kc;
  g.kb = min(L.nz, g.ka + kc - 1);
  DevExec3<T, TI, TJ> x;
  mp3_block<T, TI, TJ>(x, g, sh, a);
}

// ghost cells of a scalar field from their interior images.  One thread per
// ghost cell: the x faces (with edges), then the y faces, then the z faces.
// ghost cell number t: the x faces (with edges), then the y faces, then the
// z faces; false past the last
__device__ inline bool bc_ghost(const Layout &L, long t, int &i, int &j, int &k) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const long na = 2L * (ny + 2) * (nz + 2), nb = 2L * nx * (nz + 2), nc = 2L * nx * ny;
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
    return false;
  }
  return true;
}

// the ghost cells of up to three fields in place, field blockIdx.y
template <typename T>
struct BcSet {
  T *f[3];
};
template <typename T>
__global__ void __launch_bounds__(256) k_bc_set(Layout L, Phys ph, BcSet<T> set) {
  int i, j, k, si, sj, sk;
  if (!bc_ghost(L, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  bc_source(L, ph, i, j, k, si, sj, sk);
  T *c = set.f[blockIdx.y];
  c[IX(L, i, j, k)] = c[IX(L, si, sj, sk)];
}

template <typename T>
__global__ void __launch_bounds__(256) k_bc(Layout L, Phys ph, VentArgs va, T *__restrict__ dst,
                                            const T *__restrict__ src, int vent, double vent_value) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  int i, j, k;
  if (!bc_ghost(L, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  bool gi = (i == 0 || i == nx + 1), gj = (j == 0 || j == ny + 1);
  int si, sj, sk;
  bc_source(L, ph, i, j, k, si, sj, sk);
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

// planes per block of the corrective-pass kernel (HC_MP_KC overrides): 64
// (1-2% faster than 32 at 256^3 in both forms: fewer chunk overlaps), fewer
// where the grid has too few tiles to give two 512-thread blocks per CU
static int mp_chunk(const Layout &L, int TI, int TJ) {
  static int kc_env = [] {
    const char *e = getenv("HC_MP_KC");
    return e ? atoi(e) : 0;
  }();
  if (kc_env > 0) return kc_env < L.nz ? kc_env : L.nz;
  const long tiles = (long)((L.nx + TI - 1) / TI) * ((L.ny + TJ - 1) / TJ);
  return tile_chunk(L.nz, tiles, 64, 512);
}

// XCD-aware tile order for the LDS-tiled kernels (HC_TILE_XCD=0 turns it off)
int tile_remap() {
  static int r = [] {
    const char *e = getenv("HC_TILE_XCD");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  return r;
}

// tile of the corrective-pass kernel (HC_MP_TILE = 0, 1, 2: MpTile at
// 32x8, 16x16, 32x16; 3 and 4: the fixed-assignment MpTile2 at 32x16 and
// 16x16; 5, the default, 6 and 7: MpTile3, each face's donor-cell flux formed
// once, at 32x16, 32x12 and 16x16).  All give the same values bit for bit in
// FP64 (the FP32 form of MpTile3 forms the antidiffusive velocities with fused
// multiply-adds).  At 256^3 the default takes 0.379 ms (FP64) and 0.219 ms
// (FP32) per launch against 0.401 and 0.241 for MpTile2 32x16 and 0.461 and
// 0.304 for MpTile 32x16.
static int mp_tile() {
  static int t = [] {
    const char *e = getenv("HC_MP_TILE");
    return e ? atoi(e) : 5;
  }();
  return t;
}

template <typename T, int TI, int TJ>
static void launch_mpdata_tile(hipStream_t st, const Layout &L, const MpPass<T> &a) {
  const int kc = mp_chunk(L, TI, TJ);
  dim3 grid((L.nx + TI - 1) / TI, (L.ny + TJ - 1) / TJ, (L.nz + kc - 1) / kc);
  hipLaunchKernelGGL((k_mpdata_pass<T, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, a, kc, tile_remap());
}

template <typename T, int TI, int TJ>
static void launch_mpdata_tile2(hipStream_t st, const Layout &L, const MpPass<T> &a) {
  const int kc = mp_chunk(L, TI, TJ);
  dim3 grid((L.nx + TI - 1) / TI, (L.ny + TJ - 1) / TJ, (L.nz + kc - 1) / kc);
  hipLaunchKernelGGL((k_mpdata_pass2<T, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, a, kc, tile_remap());
}

template <typename T, int TI, int TJ>
static void launch_mpdata_tile3(hipStream_t st, const Layout &L, const MpPass<T> &a) {
  const int kc = mp_chunk(L, TI, TJ);
  dim3 grid((L.nx + TI - 1) / TI, (L.ny + TJ - 1) / TJ, (L.nz + kc - 1) / kc);
  hipLaunchKernelGGL((k_mpdata_pass3<T, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, a, kc, tile_remap());
}

template <typename T>
static void launch_mpdata_pass(hipStream_t st, const Layout &L, const MpPass<T> &a) {
  // the first form (MpTile) does not write the flags: a pass that keeps them
  // runs MpTile2
  const int t = (a.fl && mp_tile() < 3) ? 3 : mp_tile();
  switch (t) {
    case 5: launch_mpdata_tile3<T, 32, 16>(st, L, a); break;
    case 6: launch_mpdata_tile3<T, 32, 12>(st, L, a); break;
    case 7: launch_mpdata_tile3<T, 16, 16>(st, L, a); break;
    case 1: launch_mpdata_tile<T, 16, 16>(st, L, a); break;
    case 2: launch_mpdata_tile<T, 32, 16>(st, L, a); break;
    case 3: launch_mpdata_tile2<T, 32, 16>(st, L, a); break;
    case 4: launch_mpdata_tile2<T, 16, 16>(st, L, a); break;
    default: launch_mpdata_tile<T, 32, 8>(st, L, a); break;
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
                   const double *qpatch, const TracerArgs &ta, MarkFn mark, void *mark_arg) {
  StepArgs s2 = s;
  s2.tau3 = ta.tau3;
  const long nghost = 2L * (L.ny + 2) * (L.nz + 2) + 2L * L.nx * (L.nz + 2) + 2L * L.nx * L.ny;
  const unsigned nbc = (unsigned)((nghost + 255) / 256);
  T *c1 = (T *)tw.c1, *c2 = (T *)tw.c2;
  if (!ta.pass1_done) {
    constexpr int TI = 32, TJ = 8;
    const int kc = tile_chunk(L.nz, (long)((L.nx + TI - 1) / TI) * ((L.ny + TJ - 1) / TJ), 32, 768);
    dim3 grid((L.nx + TI - 1) / TI, (L.ny + TJ - 1) / TJ, (L.nz + kc - 1) / kc);
    hipLaunchKernelGGL((k_pass1_tile<T, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, s2, ph, c, (T *)tw.uc,
                       (T *)tw.vc, (T *)tw.wc, f.ss, f.ssv, aux.azzmin, c_bkg, c1, kc, tile_remap());
  }
  if (!ta.pass1_done) hipLaunchKernelGGL(k_bc<T>, dim3(nbc), dim3(256), 0, st, L, ph, va, c1, c1, 0, 0.0);
  if (ph.ismol != 1) {
    // no corrective passes: sponge/source straight after pass 1 through the
    // limited kernel with zero antidiffusive velocities would change nothing;
    // this configuration is not used by the reference runs
    return;
  }
  // pass 2: the antidiffusive velocities are kept for pass 3
  MpPass<T> a{};
  a.cmin = ta.cmin; a.cmax = ta.cmax;
  a.tau3 = ta.tau3; a.dxdydz = s.dx * s.dy * s.dz;
  a.cbkg = c_bkg; a.alphae0 = aux.alphae0; a.alphaw = aux.alphaw; a.qpatch = qpatch;
  a.map_valid = va.map_valid; a.nxs = va.nxs; a.nxe = va.nxe; a.nys = va.nys; a.nye = va.nye;
  a.cc = c1; a.U = (T *)tw.uc; a.V = (T *)tw.vc; a.W = (T *)tw.wc;
  a.ud = (T *)tw.ud2; a.vd = (T *)tw.vd2; a.wd = (T *)tw.wd2;
  a.cout = c2; a.cpre = nullptr; a.last = 0;
  a.fl = tw.fl2;
  if (mark) mark(mark_arg, 0);
  launch_mpdata_pass<T>(st, L, a);
  if (mark) mark(mark_arg, 1);
  hipLaunchKernelGGL(k_bc<T>, dim3(nbc), dim3(256), 0, st, L, ph, va, c2, c2, 0, 0.0);
  // pass 3 (the antidiffusive velocities of pass 2 play the Courant numbers);
  // c1 receives the pre-sponge values of the cells next to a boundary
  a.cc = c2; a.U = (T *)tw.ud2; a.V = (T *)tw.vd2; a.W = (T *)tw.wd2;
  a.ud = (T *)tw.ud3; a.vd = (T *)tw.vd3; a.wd = (T *)tw.wd3;
  T *cpre = tw.cpre ? (T *)tw.cpre : c1;
  a.cout = c; a.cpre = cpre; a.last = 1;
  a.fl = tw.fl3;
  if (mark) mark(mark_arg, 0);
  launch_mpdata_pass<T>(st, L, a);
  if (mark) mark(mark_arg, 1);
  hipLaunchKernelGGL(k_bc<T>, dim3(nbc), dim3(256), 0, st, L, ph, va, c, cpre, ta.vent, ta.vent_value);
  if (ta.budgets) {
    int nb = L.ny * L.nz;
    hipLaunchKernelGGL(k_budget<T>, dim3(nb), dim3(256), 0, st, L, ta.tau3, va, c, c_bkg, aux.alphae, qpatch,
                       tw.budget_part);
    hipLaunchKernelGGL(k_sum3, dim3(1), dim3(256), 0, st, nb, tw.budget_part, tw.budget,
                       s.dx * s.dy * s.dz, s.dx * s.dy * s.dz * ta.tau3);
  }
}

template <typename T, int NC>
static void launch_pass1_set(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph,
                             const Pass1Set<T> &set, const double *azzmin) {
  constexpr int TI = 32, TJ = 8;
  const int kc = tile_chunk(L.nz, (long)((L.nx + TI - 1) / TI) * ((L.ny + TJ - 1) / TJ), 32, 768);
  dim3 grid((L.nx + TI - 1) / TI, (L.ny + TJ - 1) / TJ, (L.nz + kc - 1) / kc);
  hipLaunchKernelGGL((k_pass1_multi<T, TI, TJ, NC>), grid, dim3(TI * TJ), 0, st, L, s, ph, set, azzmin, kc,
                     tile_remap());
}

template <typename T>
void launch_pass1_multi(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const VentArgs &va,
                        const FieldPtrs<T> &f, const Aux &aux, const TracerWork &tw, int n, T *const *c,
                        const double *const *c_bkg, T *const *c1, double tau3) {
  if (n < 1 || n > 3) return;
  StepArgs s2 = s;
  s2.tau3 = tau3;
  Pass1Set<T> set{};
  set.src[0] = (const T *)tw.uc; set.src[1] = (const T *)tw.vc; set.src[2] = (const T *)tw.wc;
  set.src[3] = f.ss; set.src[4] = f.ssv;
  for (int q = 0; q < n; ++q) {
    set.src[5 + q] = c[q];
    set.cbkg[q] = c_bkg[q];
    set.out[q] = c1[q];
  }
  if (n == 1) launch_pass1_set<T, 1>(st, L, s2, ph, set, aux.azzmin);
  else if (n == 2) launch_pass1_set<T, 2>(st, L, s2, ph, set, aux.azzmin);
  else launch_pass1_set<T, 3>(st, L, s2, ph, set, aux.azzmin);
  const long nghost = 2L * (L.ny + 2) * (L.nz + 2) + 2L * L.nx * (L.nz + 2) + 2L * L.nx * L.ny;
  const unsigned nbc = (unsigned)((nghost + 255) / 256);
  (void)va;
  BcSet<T> bs{};
  for (int q = 0; q < n; ++q) bs.f[q] = c1[q];
  hipLaunchKernelGGL(k_bc_set<T>, dim3(nbc, (unsigned)n), dim3(256), 0, st, L, ph, bs);
}

template <typename T>
void launch_vent_ghost(hipStream_t st, const Layout &L, const VentArgs &va, T *c, double value) {
  if (!va.map_valid) return;
  hipLaunchKernelGGL(k_vent<T>, dim3(va.nye - va.nys + 1), dim3(va.nxe - va.nxs + 1), 0, st, L, va, c, value);
}

// CFL_VALUES reductions, hotcross80-revised2014.f:395-402, plus a count of
// non-finite velocities (the NaN bailout the reference keeps commented out at
// :417-431).  One wave per (j,k) row at a time, rows strided over the grid's
// waves; fmin/fmax ignore NaN, so the count is what catches a NaN state.
// part holds CFL_NPART values per block: max u, min u, max v, min v, max w,
// min w, non-finite count.
template <typename T>
__global__ void __launch_bounds__(256) k_cfl(Layout L, const T *__restrict__ u, const T *__restrict__ v,
                                             const T *__restrict__ w, double *__restrict__ part) {
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  const int nrows = (L.ny + 1) * (L.nz + 1), nwaves = (int)gridDim.x * 4;
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300}, bad = 0.0;
  for (int row = (int)blockIdx.x * 4 + wv; row < nrows; row += nwaves) {
    const int j = row % (L.ny + 1), k = row / (L.ny + 1);  // j,k = 0..ny, 0..nz
    for (int i = lane; i <= L.nx; i += 64) {
      if (j >= 1 && k >= 1) {
        double a = F3(u, i, j, k);
        mn[0] = fmin(mn[0], a); mx[0] = fmax(mx[0], a); bad += !(fabs(a) <= 1.7976931348623157e308);
      }
      if (i >= 1 && k >= 1) {
        double a = F3(v, i, j, k);
        mn[1] = fmin(mn[1], a); mx[1] = fmax(mx[1], a); bad += !(fabs(a) <= 1.7976931348623157e308);
      }
      if (i >= 1 && j >= 1) {
        double a = F3(w, i, j, k);
        mn[2] = fmin(mn[2], a); mx[2] = fmax(mx[2], a); bad += !(fabs(a) <= 1.7976931348623157e308);
      }
    }
  }
  double r[CFL_NPART] = {mx[0], mn[0], mx[1], mn[1], mx[2], mn[2], bad};
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < CFL_NPART; ++q) {
      double o = __shfl_xor(r[q], off, 64);
      r[q] = q == 6 ? r[q] + o : (q % 2 == 0) ? fmax(r[q], o) : fmin(r[q], o);
    }
  __shared__ double sh[4][CFL_NPART];
  if (lane == 0)
    for (int q = 0; q < CFL_NPART; ++q) sh[wv][q] = r[q];
  __syncthreads();
  if (threadIdx.x < CFL_NPART) {
    const int q = (int)threadIdx.x;
    double a = sh[0][q];
    for (int x = 1; x < 4; ++x) a = q == 6 ? a + sh[x][q] : (q % 2 == 0) ? fmax(a, sh[x][q]) : fmin(a, sh[x][q]);
    part[(size_t)q * gridDim.x + blockIdx.x] = a;
  }
}

__global__ void __launch_bounds__(256) k_cfl_final(int n, const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[CFL_NPART][256];
  for (int q = 0; q < CFL_NPART; ++q) {
    double a = q == 6 ? 0.0 : (q % 2 == 0) ? -1e300 : 1e300;
    for (int r = threadIdx.x; r < n; r += 256) {
      double b = part[(size_t)q * n + r];
      a = q == 6 ? a + b : (q % 2 == 0) ? fmax(a, b) : fmin(a, b);
    }
    sh[q][threadIdx.x] = a;
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < CFL_NPART; ++q) {
        double a = sh[q][threadIdx.x], b = sh[q][threadIdx.x + off];
        sh[q][threadIdx.x] = q == 6 ? a + b : (q % 2 == 0) ? fmax(a, b) : fmin(a, b);
      }
    __syncthreads();
  }
  if (threadIdx.x < CFL_NPART) out[threadIdx.x] = sh[threadIdx.x][0];
}

template <typename T>
void launch_cfl(hipStream_t st, const Layout &L, const FieldPtrs<T> &f, double *part, double *out) {
  const int nrows = (L.ny + 1) * (L.nz + 1);
  const int nb = std::min(CFL_BLOCKS, (nrows + 3) / 4);
  hipLaunchKernelGGL(k_cfl<T>, dim3(nb), dim3(256), 0, st, L, f.u, f.v, f.w, part);
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, st, nb, part, out);
}

#define HC_INST_T(T)                                                                                          \
  template void launch_courant<T>(hipStream_t, const Layout &, const StepArgs &, const FieldPtrs<T> &,        \
                                  TracerWork &);                                                              \
  template void launch_tracer<T>(hipStream_t, const Layout &, const StepArgs &, const Phys &,                 \
                                 const VentArgs &, const FieldPtrs<T> &, const Aux &, TracerWork &, T *,      \
                                 const double *, const double *, const TracerArgs &, MarkFn, void *);         \
  template void launch_pass1_multi<T>(hipStream_t, const Layout &, const StepArgs &, const Phys &,            \
                                      const VentArgs &, const FieldPtrs<T> &, const Aux &, const TracerWork &,   \
                                      int, T *const *, const double *const *, T *const *, double);             \
  template void launch_vent_ghost<T>(hipStream_t, const Layout &, const VentArgs &, T *, double);             \
  template void launch_cfl<T>(hipStream_t, const Layout &, const FieldPtrs<T> &, double *, double *);
HC_INST_T(double)
HC_INST_T(float)

}  // namespace hc
