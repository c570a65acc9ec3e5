// hc_stencil.hip - stencil kernels of the HotCross timestep for gfx950.
//
// One thread per cell of a box, block (64,4): a wavefront sweeps 64
// consecutive i of one (j,k) row, so every load and store of a field is a
// coalesced x-fastest sweep.  All kernels are HBM-bandwidth bound (about
// 1.5 flop per byte, far below the FP64 ridge), so they are written for the
// fewest passes over HBM, not for arithmetic.  Expressions keep the
// reference's evaluation order term by term (and the library is built with
// -ffp-contract=off) so the FP64 form reproduces the reference's rounding.
#include <stdlib.h>

#include <algorithm>

#include "hc_common.h"
#include "hc_cell.h"
#include "hc_kernels.h"
#include "hc_tile.h"

namespace hc {

#define F3(a, i, j, k) (a)[IX(L, (i), (j), (k))]


// ---------------------------------------------------------------------------
// FORCES, hotcross80-revised2014.f:890-1139 (the per-cell formulas are in hc_cell.h).
// one pass over the union of the three output boxes: the inputs shared by
// fx, fy and fz are read from HBM once per block
template <typename T, bool P2>
__global__ void __launch_bounds__(256) k_forces(Layout L, Box b, StepArgs s, Phys ph, const T *__restrict__ u,
                                                const T *__restrict__ v, const T *__restrict__ w,
                                                const T *__restrict__ um, const T *__restrict__ vm,
                                                const T *__restrict__ wm, const T *__restrict__ ss,
                                                const T *__restrict__ ssv, const T *__restrict__ rhop,
                                                const double *__restrict__ azzmin, const double *__restrict__ alphau,
                                                const double *__restrict__ alphav, const double *__restrict__ alphae0,
                                                const double *__restrict__ alphaw, T *__restrict__ fx,
                                                T *__restrict__ fy, T *__restrict__ fz) {
  HC_BOX_IJK(b);  // i = 0..nx, j = 0..ny, k = 0..nz
  const GView<T> U{u, L}, Vv{v, L}, W{w, L}, UM{um, L}, VM{vm, L}, WM{wm, L}, SS{ss, L}, SV{ssv, L}, RH{rhop, L};
  if (j >= 1 && k >= 1)
    forces_x_at<T, P2>(L, i, j, k, s, ph, U, Vv, W, UM, VM, WM, SS, SV, azzmin, alphau, fx);
  if (i >= 1 && k >= 1)
    forces_y_at<T, P2>(L, i, j, k, s, ph, U, Vv, W, UM, VM, WM, SS, SV, azzmin, alphav, fy);
  if (i >= 1 && j >= 1)
    forces_z_at<T, P2>(L, i, j, k, s, ph, U, Vv, W, UM, VM, WM, SS, SV, RH, azzmin, alphae0, alphaw, fz);
}

// The interior box i = 1..nx, j = 1..ny, k = 1..nz as a k-marching LDS
// tile: a block owns a TI x TJ column of cells and keeps the planes k-1, k,
// k+1 of the eight stencil inputs in LDS rings with a one-cell halo in i and
// j.  Each step reads one new plane per field from HBM (prefetched into
// registers a plane ahead), and the three force components of the cell are
// formed from LDS with the same expressions as k_forces.  The blocks of the
// first chunk start one plane higher, at k = 0, where only fz is formed (the
// k = 0 face over i, j >= 1); the faces i = 0 (fx) and j = 0 (fy) are left to
// k_forces.
template <typename T, bool P2, int TI, int TJ>
__global__ void __launch_bounds__(TI *TJ) k_forces_tile(Layout L, StepArgs s, Phys ph, FieldPtrs<T> fp, int kc,
                                                        int remap, const double *__restrict__ azzmin,
                                                        const double *__restrict__ alphau,
                                                        const double *__restrict__ alphav,
                                                        const double *__restrict__ alphae0,
                                                        const double *__restrict__ alphaw) {
  __shared__ Ring3<T, TI, TJ, 8> sh;
  unsigned bx, by, bz;
  box_block_coords(remap, bx, by, bz);  // XCD x gets the x-th slice of tiles
  const int i0 = 1 + (int)bx * TI, j0 = 1 + (int)by * TJ;
  const int ka = 1 + (int)bz * kc, kb = min(L.nz, ka + kc - 1);
  const T *src[8] = {fp.u, fp.v, fp.w, fp.um, fp.vm, fp.wm, fp.ss, fp.ssv};
  ring3_march(sh, L, src, i0, j0, ka == 1 ? 0 : ka, kb, [&](int i, int j, int k, auto &view) {
    const auto U = view(0), Vv = view(1), W = view(2), UM = view(3), VM = view(4), WM = view(5), SS = view(6),
               SV = view(7);
    const GView<T> RH{fp.rhop, L};
    if (k >= 1) {
      forces_x_at<T, P2>(L, i, j, k, s, ph, U, Vv, W, UM, VM, WM, SS, SV, azzmin, alphau, fp.fx);
      forces_y_at<T, P2>(L, i, j, k, s, ph, U, Vv, W, UM, VM, WM, SS, SV, azzmin, alphav, fp.fy);
    }
    forces_z_at<T, P2>(L, i, j, k, s, ph, U, Vv, W, UM, VM, WM, SS, SV, RH, azzmin, alphae0, alphaw, fp.fz);
  });
}

// ---------------------------------------------------------------------------
// MOMENTUM_U/V/W, hotcross80-revised2014.f:671-887.  Each thread computes the
// value of one cell of the written set, re-evaluating the interior formula at
// the source cell of a boundary copy (a pure function of the inputs), so one
// pass writes interior and ghosts with the reference's values.
// Fused epilogue of the native loop: with the new up of a cell at hand, also
// the work velocity's Courant number (hotcross80-revised2014.f:303-305,
// mpdata-hotcross80.f:96-102, the expressions of k_courant) and REFRESH's
// filtered value (:1421-1488, the expression of k_refresh) into a spare
// buffer that becomes um when the loop reaches REFRESH.  Every cell of the
// component's box is visited, written by MOMENTUM or not.
template <typename T, bool P2>
__device__ inline void mom_fuse(const Layout &L, const StepArgs &s, const MomFuse<T> &mf, bool wr, T val,
                                const T *u, const T *um, const T *up, int i, int j, int k, int comp, bool cour_box) {
  const T upv = wr ? val : F3(up, i, j, k);
  const T a = F3(u, i, j, k);
  if (mf.cour && cour_box) {
    if (comp == 0) F3(mf.cour, i, j, k) = (T)0.5 * (upv + a) * (T)mf.tau3 / Sp<T, P2>((T)s.dx, (T)s.idx);
    else if (comp == 1) F3(mf.cour, i, j, k) = (T)0.5 * (upv + a) * (T)mf.tau3 / Sp<T, P2>((T)s.dy, (T)s.idy);
    else F3(mf.cour, i, j, k) = -((T)0.5 * (upv + a)) * (T)mf.tau3 / Sp<T, P2>((T)s.dz, (T)s.idz);
  }
  F3(mf.umn, i, j, k) = mf.leap == 1 ? upv : a + (T)mf.assel * (upv - (T)2 * a + F3(um, i, j, k));
}

template <typename T, bool P2>
__device__ inline T mom_u_val(const Layout &L, const StepArgs &s, const T *u, const T *um, const T *p,
                              const T *fx, const double *pgradx, int i, int j, int k) {
  return (T)(1 - s.delta) * F3(u, i, j, k) + (T)s.delta * F3(um, i, j, k) +
         (T)s.dt * (T)s.leap * (-(F3(p, i + 1, j, k) - F3(p, i, j, k)) / Sp<T, P2>((T)s.dx, (T)s.idx) - (T)pgradx[k] + F3(fx, i, j, k));
}

template <typename T, bool P2, bool FUSE>
__global__ void __launch_bounds__(256) k_momentum_u(Layout L, Box b, StepArgs s, Phys ph, const T *__restrict__ u,
                                                    const T *__restrict__ um, const T *__restrict__ p,
                                                    const T *__restrict__ fx, const double *__restrict__ pgradx,
                                                    T *__restrict__ up, MomFuse<T> mf) {
  HC_BOX_IJK(b);  // box: i=-1..nx+1, j=0..ny+1, k=0..nz+1
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  bool gi = (i == -1 || i == nx + 1), gj = (j == 0 || j == ny + 1), gk = (k == 0 || k == nz + 1);
  const bool wr = (int)gi + (int)gj + (int)gk <= 1;  // edges keep -1e34
  if (!FUSE && !wr) return;
  T val = (T)0;
  if (!wr) {
  } else if (!gi && !gj && !gk) {
    val = mom_u_val<T, P2>(L, s, u, um, p, fx, pgradx, i, j, k);
  } else if (gk) {
    if (k == 0) val = mom_u_val<T, P2>(L, s, u, um, p, fx, pgradx, i, j, 1);
    else val = (T)ph.bottom_slip * mom_u_val<T, P2>(L, s, u, um, p, fx, pgradx, i, j, nz);
  } else if (gj) {
    val = mom_u_val<T, P2>(L, s, u, um, p, fx, pgradx, i, j == 0 ? 1 : ny, k);
  } else {
    val = mom_u_val<T, P2>(L, s, u, um, p, fx, pgradx, i == -1 ? 0 : nx, j, k);
  }
  if (wr) F3(up, i, j, k) = val;
  if (FUSE) mom_fuse<T, P2>(L, s, mf, wr, val, u, um, up, i, j, k, 0, i >= 0 && i <= nx);
}

template <typename T, bool P2>
__device__ inline T mom_v_val(const Layout &L, const StepArgs &s, const T *v, const T *vm, const T *p,
                              const T *fy, const double *pgrady, int i, int j, int k) {
  return (T)(1 - s.delta) * F3(v, i, j, k) + (T)s.delta * F3(vm, i, j, k) +
         (T)s.dt * (T)s.leap * (-(F3(p, i, j + 1, k) - F3(p, i, j, k)) / Sp<T, P2>((T)s.dy, (T)s.idy) - (T)pgrady[k] + F3(fy, i, j, k));
}

template <typename T, bool P2, bool FUSE>
__global__ void __launch_bounds__(256) k_momentum_v(Layout L, Box b, StepArgs s, Phys ph, const T *__restrict__ v,
                                                    const T *__restrict__ vm, const T *__restrict__ p,
                                                    const T *__restrict__ fy, const double *__restrict__ pgrady,
                                                    T *__restrict__ vp, MomFuse<T> mf) {
  HC_BOX_IJK(b);  // box: i=0..nx+1, j=-1..ny+1, k=0..nz+1
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  bool gi = (i == 0 || i == nx + 1), gj = (j == -1 || j == ny + 1), gk = (k == 0 || k == nz + 1);
  const bool wr = (int)gi + (int)gj + (int)gk <= 1;
  if (!FUSE && !wr) return;
  T val = (T)0;
  if (!wr) {
  } else if (!gi && !gj && !gk) {
    val = mom_v_val<T, P2>(L, s, v, vm, p, fy, pgrady, i, j, k);
  } else if (gk) {
    if (k == 0) val = mom_v_val<T, P2>(L, s, v, vm, p, fy, pgrady, i, j, 1);
    else val = (T)ph.bottom_slip * mom_v_val<T, P2>(L, s, v, vm, p, fy, pgrady, i, j, nz);
  } else if (gi) {
    val = mom_v_val<T, P2>(L, s, v, vm, p, fy, pgrady, i == 0 ? 1 : nx, j, k);
  } else {
    val = mom_v_val<T, P2>(L, s, v, vm, p, fy, pgrady, i, j == -1 ? 0 : ny, k);
  }
  if (wr) F3(vp, i, j, k) = val;
  if (FUSE) mom_fuse<T, P2>(L, s, mf, wr, val, v, vm, vp, i, j, k, 1, j >= 0 && j <= ny);
}

template <typename T, bool P2>
__device__ inline T mom_w_val(const Layout &L, const StepArgs &s, const T *w, const T *wm, const T *p,
                              const T *fz, int i, int j, int k) {
  return (T)(1 - s.delta) * F3(w, i, j, k) + (T)s.delta * F3(wm, i, j, k) +
         (T)s.leap * (T)s.dt * (-(F3(p, i, j, k + 1) - F3(p, i, j, k)) / Sp<T, P2>((T)s.dz, (T)s.idz) + F3(fz, i, j, k));
}

// vent value of wp at the bottom (k = nz, nz+1), :821-857
template <typename T>
__device__ inline T w_vent(const VentArgs &va, const double *alphae, const Layout &L, int i, int j) {
  if (alphae[i + (L.nx + 2) * j] == 0) {
    bool src = va.map_valid && i >= va.nxs && i <= va.nxe && j >= va.nys && j <= va.nye;
    if (src) return (T)(-va.wup);
    return (T)(va.wup * va.sumup / va.sumdown);
  }
  return (T)0.0;
}

template <typename T, bool P2, bool FUSE>
__global__ void __launch_bounds__(256) k_momentum_w(Layout L, Box b, StepArgs s, Phys ph, VentArgs va,
                                                    const T *__restrict__ w, const T *__restrict__ wm,
                                                    const T *__restrict__ p, const T *__restrict__ fz,
                                                    const double *__restrict__ alphae, T *__restrict__ wp,
                                                    MomFuse<T> mf) {
  HC_BOX_IJK(b);  // box: i=0..nx+1, j=0..ny+1, k=-1..nz+1
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  bool ii = (i >= 1 && i <= nx), jj = (j >= 1 && j <= ny);
  auto inner = [&](int i2, int j2, int k2) -> T {
    // wp(i2,j2,k2) for i2,j2 interior after the interior update and bottom/top rows
    if (k2 >= 1 && k2 <= nz - 1) return mom_w_val<T, P2>(L, s, w, wm, p, fz, i2, j2, k2);
    if (k2 == nz) return w_vent<T>(va, alphae, L, i2, j2);
    return (T)0.0;  // k2 = 0
  };
  T val = (T)0;
  bool wr = true;
  if (ii && jj) {
    if (k >= 1 && k <= nz - 1) val = mom_w_val<T, P2>(L, s, w, wm, p, fz, i, j, k);
    else if (k == nz || k == nz + 1) val = w_vent<T>(va, alphae, L, i, j);
    else val = (T)0.0;  // k = -1, 0
  } else if (ii && (j == 0 || j == ny + 1) && k >= 0 && k <= nz) {
    // y ends, :863-875 (0 * wp(i,1,k) / 0 * wp(i,ny,k))
    val = (T)0 * inner(i, j == 0 ? 1 : ny, k);
  } else if ((i == 0 || i == nx + 1) && j >= 0 && j <= ny && k >= 1 && k <= nz) {
    // x ends, :877-884 (0 * wp(1,j,k) / 0 * wp(nx,j,k)); wp(1,0,k) is the y-end value
    int i2 = i == 0 ? 1 : nx;
    T src = (j >= 1) ? inner(i2, j, k) : (T)0 * inner(i2, 1, k);
    val = (T)0 * src;
  } else {
    wr = false;  // not written by MOMENTUM_W
    if (!FUSE) return;
  }
  if (wr) F3(wp, i, j, k) = val;
  if (FUSE) mom_fuse<T, P2>(L, s, mf, wr, val, w, wm, wp, i, j, k, 2, k >= 0 && k <= nz);
}

template <typename T, bool P2>
__global__ void __launch_bounds__(256) k_viscosity(Layout L, Box b, StepArgs s, Phys ph, const T *__restrict__ um,
                                                   const T *__restrict__ vm, const T *__restrict__ wm,
                                                   const T *__restrict__ rhop, const double *__restrict__ rho_bkg,
                                                   T *__restrict__ ss, T *__restrict__ ssv) {
  HC_BOX_IJK(b);  // box: i=0..nx+1, j=0..ny+1, k=0..nz+1
  const GView<T> UM{um, L}, VM{vm, L}, WM{wm, L}, RH{rhop, L};
  visc_at<T, P2>(L, s, ph, UM, VM, WM, RH, rho_bkg, ss, ssv, i, j, k);
}

// the ghost shell of the viscosity box, one thread per shell cell: the x
// faces (with edges), the y faces, then the z faces
template <typename T, bool P2>
__global__ void __launch_bounds__(256) k_visc_shell(Layout L, StepArgs s, Phys ph, const T *__restrict__ um,
                                                    const T *__restrict__ vm, const T *__restrict__ wm,
                                                    const T *__restrict__ rhop, const double *__restrict__ rho_bkg,
                                                    T *__restrict__ ss, T *__restrict__ ssv) {
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
  const GView<T> UM{um, L}, VM{vm, L}, WM{wm, L}, RH{rhop, L};
  visc_at<T, P2>(L, s, ph, UM, VM, WM, RH, rho_bkg, ss, ssv, i, j, k);
}

// the interior i = 1..nx, j = 1..ny, k = 1..nz as a k-marching LDS tile with
// planes k-1, k, k+1 of um, vm, wm (one-cell halo), as k_forces_tile
template <typename T, bool P2, int TI, int TJ>
__global__ void __launch_bounds__(TI *TJ) k_visc_tile(Layout L, StepArgs s, Phys ph, const T *__restrict__ um,
                                                      const T *__restrict__ vm, const T *__restrict__ wm,
                                                      const T *__restrict__ rhop,
                                                      const double *__restrict__ rho_bkg, T *__restrict__ ss,
                                                      T *__restrict__ ssv, int kc, int remap) {
  __shared__ Ring3<T, TI, TJ, 3> sh;
  unsigned bx, by, bz;
  box_block_coords(remap, bx, by, bz);
  const int i0 = 1 + (int)bx * TI, j0 = 1 + (int)by * TJ;
  const int ka = 1 + (int)bz * kc, kb = min(L.nz, ka + kc - 1);
  const T *src[3] = {um, vm, wm};
  ring3_march(sh, L, src, i0, j0, ka, kb, [&](int i, int j, int k, auto &view) {
    const auto UM = view(0), VM = view(1), WM = view(2);
    const GView<T> RH{rhop, L};
    F3(ss, i, j, k) = ss_val<T, P2>(L, ph, s, UM, VM, WM, i, j, k);
    F3(ssv, i, j, k) = ssv_val<T, P2>(L, ph, s, UM, VM, WM, RH, rho_bkg, i, j, k);
  });
}

template <typename T>
__global__ void __launch_bounds__(256) k_density(Layout L, Box b, Phys ph, const T *__restrict__ t,
                                                 const T *__restrict__ sal, const double *__restrict__ rho_bkg,
                                                 T *__restrict__ rhop) {
  HC_BOX_IJK(b);  // box 0..nx+1, 0..ny+1, 0..nz+1
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  int ng = (int)(i == 0 || i == nx + 1) + (int)(j == 0 || j == ny + 1) + (int)(k == 0 || k == nz + 1);
  if (ng > 1) return;
  // equation of state evaluated in FP64 from the stored T and S
  double r = rho_rho<T>(ph, (double)F3(sal, i, j, k), (double)F3(t, i, j, k)) - rho_bkg[k];
  F3(rhop, i, j, k) = (T)r;
}

// ---------------------------------------------------------------------------
// REFRESH, hotcross80-revised2014.f:1411-1493, as "um <- filtered u" over the
// whole box; the caller then swaps the u and up buffers (u = up).
template <typename T>
__global__ void __launch_bounds__(256) k_refresh(size_t n, int leap, double assel, const T *__restrict__ u,
                                                 const T *__restrict__ up, T *__restrict__ um) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (leap == 1) {
    um[q] = up[q];
  } else {
    T a = u[q];
    um[q] = a + (T)assel * (up[q] - (T)2 * a + um[q]);
  }
}

// ---------------------------------------------------------------------------
// host launchers
// launch the power-of-two-spacing instantiation when s.p2 is set
// the LDS-tiled FORCES and VISCOSITY3D (HC_FORCES_CELL=1 selects the
// cell-parallel kernels for the whole boxes, for comparing the two)
static bool forces_tiled() {
  static bool t = [] {
    const char *e = getenv("HC_FORCES_CELL");
    return !(e && e[0] == '1');
  }();
  return t;
}

#define HC_P2_LAUNCH(K, ...)                                   \
  do {                                                         \
    if (s.p2) hipLaunchKernelGGL((K<T, true>), __VA_ARGS__);   \
    else hipLaunchKernelGGL((K<T, false>), __VA_ARGS__);       \
  } while (0)

template <typename T>
void launch_forces(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                   const Aux &aux) {
  if (!forces_tiled()) {
    Box b{0, L.nx, 0, L.ny, 0, L.nz};
    HC_P2_LAUNCH(k_forces, box_grid(b), box_block(), 0, st, L, b, s, ph, f.u, f.v, f.w, f.um, f.vm, f.wm, f.ss,
                 f.ssv, f.rhop, aux.azzmin, aux.alphau, aux.alphav, aux.alphae0, aux.alphaw, f.fx, f.fy, f.fz);
    return;
  }
  // faces i = 0 and j = 0 with the cell-parallel kernel (the tile forms the
  // k = 0 face)
  const Box faces[2] = {{0, 0, 0, L.ny, 0, L.nz}, {1, L.nx, 0, 0, 0, L.nz}};
  for (const Box &b : faces)
    HC_P2_LAUNCH(k_forces, box_grid(b), box_block(), 0, st, L, b, s, ph, f.u, f.v, f.w, f.um, f.vm, f.wm, f.ss,
                 f.ssv, f.rhop, aux.azzmin, aux.alphau, aux.alphav, aux.alphae0, aux.alphaw, f.fx, f.fy, f.fz);
  constexpr int TI = 32, TJ = 8;
  const int kc = tile_chunk(L.nz, (long)((L.nx + TI - 1) / TI) * ((L.ny + TJ - 1) / TJ), 32, 768);
  dim3 grid((L.nx + TI - 1) / TI, (L.ny + TJ - 1) / TJ, (L.nz + kc - 1) / kc);
  if (s.p2)
    hipLaunchKernelGGL((k_forces_tile<T, true, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, s, ph, f, kc, tile_remap(), aux.azzmin,
                       aux.alphau, aux.alphav, aux.alphae0, aux.alphaw);
  else
    hipLaunchKernelGGL((k_forces_tile<T, false, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, s, ph, f, kc, tile_remap(), aux.azzmin,
                       aux.alphau, aux.alphav, aux.alphae0, aux.alphaw);
}

template <typename T, class... A>
static void mom_launch(void (*k00)(A...), void (*k01)(A...), void (*k10)(A...), void (*k11)(A...), bool p2, bool fuse,
                       dim3 g, hipStream_t st, A... args) {
  auto k = p2 ? (fuse ? k11 : k10) : (fuse ? k01 : k00);
  hipLaunchKernelGGL(k, g, box_block(), 0, st, args...);
}

template <typename T>
void launch_momentum(hipStream_t st, int comp, const Layout &L, const StepArgs &s, const Phys &ph,
                     const VentArgs &va, const FieldPtrs<T> &f, const Aux &aux, const MomFuse<T> *mf) {
  const MomFuse<T> m = mf ? *mf : MomFuse<T>{};
  const bool fu = mf != nullptr;
  if (comp == 0) {
    Box b{-1, L.nx + 1, 0, L.ny + 1, 0, L.nz + 1};
    mom_launch<T>(k_momentum_u<T, false, false>, k_momentum_u<T, false, true>, k_momentum_u<T, true, false>,
                  k_momentum_u<T, true, true>, s.p2 != 0, fu, box_grid(b), st, L, b, s, ph, (const T *)f.u,
                  (const T *)f.um, (const T *)f.p, (const T *)f.fx, (const double *)aux.pgradx, f.up, m);
  } else if (comp == 1) {
    Box b{0, L.nx + 1, -1, L.ny + 1, 0, L.nz + 1};
    mom_launch<T>(k_momentum_v<T, false, false>, k_momentum_v<T, false, true>, k_momentum_v<T, true, false>,
                  k_momentum_v<T, true, true>, s.p2 != 0, fu, box_grid(b), st, L, b, s, ph, (const T *)f.v,
                  (const T *)f.vm, (const T *)f.p, (const T *)f.fy, (const double *)aux.pgrady, f.vp, m);
  } else {
    Box b{0, L.nx + 1, 0, L.ny + 1, -1, L.nz + 1};
    mom_launch<T>(k_momentum_w<T, false, false>, k_momentum_w<T, false, true>, k_momentum_w<T, true, false>,
                  k_momentum_w<T, true, true>, s.p2 != 0, fu, box_grid(b), st, L, b, s, ph, va, (const T *)f.w,
                  (const T *)f.wm, (const T *)f.p, (const T *)f.fz, (const double *)aux.alphae, f.wp, m);
  }
}

template <typename T>
void launch_viscosity(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                      const Aux &aux) {
  if (!forces_tiled()) {
    Box b{0, L.nx + 1, 0, L.ny + 1, 0, L.nz + 1};
    HC_P2_LAUNCH(k_viscosity, box_grid(b), box_block(), 0, st, L, b, s, ph, f.um, f.vm, f.wm, f.rhop, aux.rho_bkg,
                 f.ss, f.ssv);
    return;
  }
  const long nshell = 2L * (L.ny + 2) * (L.nz + 2) + 2L * L.nx * (L.nz + 2) + 2L * L.nx * L.ny;
  HC_P2_LAUNCH(k_visc_shell, dim3((unsigned)((nshell + 255) / 256)), dim3(256), 0, st, L, s, ph, f.um, f.vm, f.wm,
               f.rhop, aux.rho_bkg, f.ss, f.ssv);
  constexpr int TI = 32, TJ = 8;
  const int kc = tile_chunk(L.nz, (long)((L.nx + TI - 1) / TI) * ((L.ny + TJ - 1) / TJ), 32, 768);
  dim3 grid((L.nx + TI - 1) / TI, (L.ny + TJ - 1) / TJ, (L.nz + kc - 1) / kc);
  if (s.p2)
    hipLaunchKernelGGL((k_visc_tile<T, true, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, s, ph, f.um, f.vm, f.wm,
                       f.rhop, aux.rho_bkg, f.ss, f.ssv, kc, tile_remap());
  else
    hipLaunchKernelGGL((k_visc_tile<T, false, TI, TJ>), grid, dim3(TI * TJ), 0, st, L, s, ph, f.um, f.vm, f.wm,
                       f.rhop, aux.rho_bkg, f.ss, f.ssv, kc, tile_remap());
}

template <typename T>
void launch_density(hipStream_t st, const Layout &L, const Phys &ph, const FieldPtrs<T> &f, const Aux &aux) {
  Box b{0, L.nx + 1, 0, L.ny + 1, 0, L.nz + 1};
  hipLaunchKernelGGL(k_density<T>, box_grid(b), box_block(), 0, st, L, b, ph, f.t, f.s, aux.rho_bkg, f.rhop);
}

template <typename T>
void launch_refresh(hipStream_t st, const Layout &L, int leap, double assel, const T *u, const T *up, T *um) {
  size_t n = L.n;
  hipLaunchKernelGGL(k_refresh<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, leap, assel, u, up, um);
}

#define HC_INST(T)                                                                                            \
  template void launch_forces<T>(hipStream_t, const Layout &, const StepArgs &, const Phys &,                 \
                                 const FieldPtrs<T> &, const Aux &);                                          \
  template void launch_momentum<T>(hipStream_t, int, const Layout &, const StepArgs &, const Phys &,          \
                                   const VentArgs &, const FieldPtrs<T> &, const Aux &, const MomFuse<T> *);  \
  template void launch_viscosity<T>(hipStream_t, const Layout &, const StepArgs &, const Phys &,              \
                                    const FieldPtrs<T> &, const Aux &);                                       \
  template void launch_density<T>(hipStream_t, const Layout &, const Phys &, const FieldPtrs<T> &,            \
                                  const Aux &);                                                               \
  template void launch_refresh<T>(hipStream_t, const Layout &, int, double, const T *, const T *, T *);
HC_INST(double)
HC_INST(float)

}  // namespace hc
