// hc_tile.h - field views and the three-plane LDS ring shared by the
// k-marching stencil kernels.
//
// A stencil formula is written once against "views" of its input fields,
// v(i, j, k): GView reads the padded box in HBM, LView reads a plane ring in
// LDS.  ring3_march runs a block over a TI x TJ column of cells and the planes
// ka..kb: it keeps planes k-1, k, k+1 of NF fields in LDS with a one-cell
// halo in i and j, fetches plane k+2 into registers while plane k is being
// computed, and calls the cell body with LDS views.  Each field plane is read
// from HBM once per block (plus the halo), instead of once per stencil point.
#pragma once
#include "hc_common.h"

namespace hc {

// a field read through the padded-box index
template <typename T>
struct GView {
  const T *p;
  Layout L;
  __device__ T operator()(int i, int j, int k) const { return p[IX(L, i, j, k)]; }
};

// planes k-1, k, k+1 of one field in LDS; (ib, jb) = (i0-1, j0-1) is the
// LDS origin of the tile with its halo
template <typename T, int SI>
struct LView {
  const T *pm, *p0, *pp;
  int k, ib, jb;
  __device__ T operator()(int i, int j, int kk) const {
    const T *b = kk < k ? pm : (kk > k ? pp : p0);
    return b[(i - ib) + SI * (j - jb)];
  }
};

template <typename T, int TI, int TJ, int NF>
struct Ring3 {
  static constexpr int SI = TI + 2, S = SI * (TJ + 2), NT = TI * TJ, RS = (S + NT - 1) / NT;
  T f[NF][3][S];
};

// body(i, j, k, view) for every cell of the tile inside the grid (i <= nx,
// j <= ny), planes ka..kb in order (0 <= ka, kb <= nz); view(f) is the
// LView of field f.  All threads of the block must call this (it
// synchronizes the block).
template <typename T, int TI, int TJ, int NF, class Body>
__device__ inline void ring3_march(Ring3<T, TI, TJ, NF> &sh, const Layout &L, const T *const *src, int i0, int j0,
                                   int ka, int kb, Body &&body) {
  typedef Ring3<T, TI, TJ, NF> R;
  constexpr int SI = R::SI, S = R::S, NT = R::NT, RS = R::RS;
  const int tid = (int)threadIdx.x;
  int so[RS];
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int e = tid + r * NT;
    const int i = i0 - 1 + e % SI, j = j0 - 1 + e / SI;
    so[r] = (e < S && i <= L.nx + 1 && j <= L.ny + 1) ? IX(L, i, j, -1) : -1;
  }
  const int pstride = L.px * L.py;
  T pf[NF][RS];
  auto fetch = [&](int k) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int r = 0; r < RS; ++r) pf[f][r] = so[r] >= 0 ? src[f][so[r] + (k + 1) * pstride] : (T)0;
  };
  auto commit = [&](int k) {
    const int slot = (k + 3) % 3;  // k >= -1
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        const int e = tid + r * NT;
        if (e < S) sh.f[f][slot][e] = pf[f][r];
      }
  };
  fetch(ka - 1);
  commit(ka - 1);
  fetch(ka);
  commit(ka);
  fetch(ka + 1);
  const int i = i0 + tid % TI, j = j0 + tid / TI;
  const bool own = i <= L.nx && j <= L.ny;
  for (int k = ka; k <= kb; ++k) {
    __syncthreads();  // the previous plane's reads of slot (k+1)%3 are done
    commit(k + 1);
    __syncthreads();
    if (k < kb) fetch(k + 2);  // in flight during this plane
    if (own) {
      const int sm = (k + 2) % 3, s0 = k % 3, sp = (k + 1) % 3;
      auto view = [&](int f) { return LView<T, SI>{sh.f[f][sm], sh.f[f][s0], sh.f[f][sp], k, i0 - 1, j0 - 1}; };
      body(i, j, k, view);
    }
  }
}

}  // namespace hc
