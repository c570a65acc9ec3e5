// hc_mpdata.h - one MPDATA corrective pass as a k-marching LDS tile.
//
// A corrective pass of TRACER31 (mpdata-hotcross80.f:218-326 antidiffusive
// velocities, :660-690 limiter flags, :705-720 limited donor-cell step, and
// on the last pass the sponge and source of :356-381) is one kernel here.
// A block owns a TI x TJ column of cells and marches down a chunk of planes
// [ka, kb].  For each new plane q it stages the planes of the concentration
// and the velocities it needs into LDS (with a two-cell halo in i and j),
// forms the antidiffusive velocities ud/vd/wd of plane q and the limiter
// flags of plane q over the tile plus a one-cell halo, and then finishes
// the limited step of plane q-1.  Nothing between the staged inputs and the
// new concentration goes through HBM; per plane the block reads four planes
// (c, U, V, W) and writes one (plus ud/vd/wd when the next pass needs them).
//
// Rings in LDS, indexed by plane number:
//   c  : q-2 .. q+1  (4)     U, V : q, q+1 (2)     W : q-1, q (2)
//   ud, vd : q (1)           wd, flags : q-1, q (2)
// The values the finishing step of plane q-1 takes from the cell's own
// column (its six face velocities, the flags above and below, and c below)
// roll in registers from one plane to the next.  Because the finishing step
// reads only planes q-1 and q of the c ring, the next plane's staging (which
// overwrites plane q-2's slot) needs no barrier after it.
//
// The face and donor formulas keep the reference's evaluation order term by
// term, so with -ffp-contract=off the FP64 form rounds as the reference.
// The tile logic is written as phases separated by barriers and compiles for
// the host as well, where a serial loop over the threads of a block checks
// the index and ring arithmetic against the plain per-face formulation.
#pragma once
#include <math.h>
#include <stdint.h>

#include "hc_common.h"

#ifndef HC_HD
#define HC_HD __host__ __device__
#endif

// a host build may define HC_MP_CHECK(x) / HC_MP_CHECKC(x) to range-check
// every LDS index of a staged / computed plane
#ifndef HC_MP_CHECK
#define HC_MP_CHECK(x) (x)
#endif
#ifndef HC_MP_CHECKC
#define HC_MP_CHECKC(x) (x)
#endif
// and HC_MP_GIDX(x) to range-check the element index of every staged load
#ifndef HC_MP_GIDX
#define HC_MP_GIDX(x) (x)
#endif

namespace hc {

// Shared-memory tile of one block.
// Staged planes cover rows j0-2 .. j0+TJ+1 (S entries); the computed ones
// (ud, vd, wd, flags) rows j0-2 .. j0+TJ only (SC entries), same pitch.
template <typename T, int TI, int TJ>
struct MpShared {
  static constexpr int SI = TI + 4, SJ = TJ + 4, S = SI * SJ, SC = SI * (TJ + 3);
  T c[4][S];
  T U[2][S], V[2][S], W[2][S];
  T ud[SC], vd[SC];
  T wd[2][SC];
  unsigned char fl[2][SC];
};

// Per-thread state that rolls from plane to plane.
template <typename T>
struct MpRegs {
  T u0, u1, v0, v1;  // ud(i-1), ud(i), vd(j-1), vd(j) at the plane being finished
  T wA, wB;          // wd at planes q-2, q-1 (the finished plane's bottom/top faces)
  int fA, fB;        // flags of the cell at planes q-2, q-1
  T cB;              // the cell's own c at plane q-2 (below the finished plane)
};

// Arguments of one pass.
template <typename T>
struct MpPass {
  const T *cc;              // input concentration (ghosts set)
  const T *U, *V, *W;       // Courant numbers / previous antidiffusive velocities
  T *ud, *vd, *wd;          // antidiffusive velocities written for the next pass (or null)
  T *cout;                  // result; on the last pass also read as the entry value
  T *cpre;                  // last pass: pre-sponge value for the boundary images
  unsigned char *fl;        // limiter flags of the interior cells (the padded index), or null
  double cmin, cmax;
  int last;
  // last pass only
  double tau3, dxdydz;
  const double *cbkg, *alphae0, *alphaw, *qpatch;
  int map_valid, nxs, nxe, nys, nye;
};

HC_HD inline double mp_fabs(double x) { return fabs(x); }
HC_HD inline float mp_fabs(float x) { return fabsf(x); }

// the quotients of the antidiffusive velocities, nine per face triple: each
// term is (coefficient * (a - b)) / (a + b + eps), multiplied before it is
// divided, as the reference evaluates it left to right
// (mpdata-hotcross80.f:228-229, 236-237, 245-246 and the vd and wd lines).
// FP64: ddiv, the correctly rounded quotient (the reference's rounding)
// without the division's scaling and fix-up, which the ratios never need:
// b = a' + b' + 1e-30 of bounded concentrations is a normal number far from
// the overflow range.  FP32 form: the quotient through the hardware
// reciprocal (v_rcp_f32, 1 ulp; the product within 2 ulp), which the form's
// stated error against FP64 absorbs.
HC_HD inline double mp_ratio(double a, double b) { return ddiv(a, b); }
HC_HD inline float mp_ratio(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
  return a * __builtin_amdgcn_rcpf(b);
#else
  return a / b;
#endif
}

struct MpGeom {
  int nx, ny, nz;
  int px, py;   // padded pitches (Layout)
  int i0, j0;   // first cell of the tile
  int ka, kb;   // planes finished by this block
};

HC_HD inline long mp_ix(const MpGeom &g, int i, int j, int k) {
  return (long)(i + 16) + (long)g.px * ((long)(j + 1) + (long)g.py * (long)(k + 1));
}

// ---------------------------------------------------------------------------
// formulas on accessors: C(a,b,c) = cc(i+a, j+b, q+c), likewise U, V, W.

// ud at face (i,j,q), mpdata-hotcross80.f:218-252
template <typename T, class FC, class FU, class FV, class FW>
HC_HD inline T mp_anti_u(const FC &C, const FU &U, const FV &V, const FW &W) {
  const T epsa = (T)1.0e-30;
  T ca = C(1, 0, 0), cb = C(0, 0, 0);
  T Uc = U(0, 0, 0);
  T r = mp_ratio((mp_fabs(Uc) - Uc * Uc / (T)1.0) * (ca - cb), ca + cb + epsa);
  ca = (C(1, 1, 0) + C(0, 1, 0));
  cb = (C(1, -1, 0) + C(0, -1, 0));
  r = r - mp_ratio((T)1 * (T)0.5 * Uc / (T)1.0 * (ca - cb), ca + cb + epsa) * (T)0.25 *
              (V(1, 0, 0) + V(0, 0, 0) + V(1, -1, 0) + V(0, -1, 0)) * (T)1;
  ca = (C(1, 0, -1) + C(0, 0, -1));
  cb = (C(1, 0, 1) + C(0, 0, 1));
  r = r - mp_ratio((T)0.5 * Uc / (T)1.0 * (ca - cb), ca + cb + epsa) * (T)0.25 *
              (W(1, 0, 0) + W(0, 0, 0) + W(1, 0, -1) + W(0, 0, -1)) * (T)1;
  return r;
}

// vd at face (i,j,q), mpdata-hotcross80.f:254-289
template <typename T, class FC, class FU, class FV, class FW>
HC_HD inline T mp_anti_v(const FC &C, const FU &U, const FV &V, const FW &W) {
  const T epsa = (T)1.0e-30;
  T ca = C(0, 1, 0), cb = C(0, 0, 0);
  T Vc = V(0, 0, 0);
  T r = mp_ratio((mp_fabs(Vc) - Vc * Vc / (T)1.0) * (T)1 * (ca - cb), ca + cb + epsa);
  ca = (C(1, 1, 0) + C(1, 0, 0));
  cb = (C(-1, 1, 0) + C(-1, 0, 0));
  r = r - mp_ratio((T)0.5 * Vc / (T)1.0 * (ca - cb), ca + cb + epsa) * (T)0.25 *
              (U(0, 1, 0) + U(0, 0, 0) + U(-1, 1, 0) + U(-1, 0, 0)) * (T)1;
  ca = (C(0, 1, -1) + C(0, 0, -1));
  cb = (C(0, 1, 1) + C(0, 0, 1));
  r = r - mp_ratio((T)0.5 * Vc / (T)1.0 * (ca - cb), ca + cb + epsa) * (T)0.25 *
              (W(0, 0, 0) + W(0, 1, 0) + W(0, 0, -1) + W(0, 1, -1)) * (T)1;
  return r;
}

// wd at face (i,j,q) (between planes q and q+1), mpdata-hotcross80.f:291-326
template <typename T, class FC, class FU, class FV, class FW>
HC_HD inline T mp_anti_w(const FC &C, const FU &U, const FV &V, const FW &W) {
  const T epsa = (T)1.0e-30;
  T ca = C(0, 0, 0), cb = C(0, 0, 1);
  T Wc = W(0, 0, 0);
  T r = mp_ratio((mp_fabs(Wc) - Wc * Wc / (T)1.0) * (ca - cb), ca + cb + epsa);
  ca = (C(1, 0, 0) + C(1, 0, 1));
  cb = (C(-1, 0, 0) + C(-1, 0, 1));
  r = r - mp_ratio((T)1 * (T)0.5 * Wc / (T)1.0 * (ca - cb), ca + cb + epsa) * (T)0.25 *
              (U(0, 0, 0) + U(0, 0, 1) + U(-1, 0, 1) + U(-1, 0, 0)) * (T)1;
  ca = (C(0, 1, 1) + C(0, 1, 0));
  cb = (C(0, -1, 1) + C(0, -1, 0));
  r = r - mp_ratio((T)1 * (T)0.5 * Wc / (T)1.0 * (ca - cb), ca + cb + epsa) * (T)0.25 *
              (V(0, 0, 0) + V(0, 0, 1) + V(0, -1, 1) + V(0, -1, 0)) * (T)1;
  return r;
}

// donor-cell flux divergence (mpdata-hotcross80.f:590-608 with the split
// velocities of :527-565); with flags, the limited form of :705-720.
// Flags: 1 = c+dc <= cmin, 2 = c+dc >= cmax, for the cell (fc) and its six
// neighbours x-, x+, y-, y+, z-, z+.
template <typename T>
HC_HD inline T mp_donor(T cc, T cxm, T cxp, T cym, T cyp, T czm, T czp, T u0, T u1, T v0, T v1, T w0, T w1,
                        bool lim, int fc, int fxm, int fxp, int fym, int fyp, int fzm, int fzp) {
  T upi0 = (T)0.5 * (u0 + mp_fabs(u0)), umi0 = (T)0.5 * (u0 - mp_fabs(u0));
  T upi1 = (T)0.5 * (u1 + mp_fabs(u1)), umi1 = (T)0.5 * (u1 - mp_fabs(u1));
  T vpj0 = (T)0.5 * (v0 + mp_fabs(v0)), vmj0 = (T)0.5 * (v0 - mp_fabs(v0));
  T vpj1 = (T)0.5 * (v1 + mp_fabs(v1)), vmj1 = (T)0.5 * (v1 - mp_fabs(v1));
  T wpk0 = (T)0.5 * (w0 + mp_fabs(w0)), wmk0 = (T)0.5 * (w0 - mp_fabs(w0));
  T wpk1 = (T)0.5 * (w1 + mp_fabs(w1)), wmk1 = (T)0.5 * (w1 - mp_fabs(w1));
  if (lim) {
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
  T F = +(upi0 * cxm + umi0 * cc) + (vpj0 * cym + vmj0 * cc) + (wmk1 * cc + wpk1 * czp) -
        ((upi1 * cc + umi1 * cxp) + (vpj1 * cc + vmj1 * cyp) + (wmk0 * czm + wpk0 * cc));
  return (T)0 * cc + F / (T)1.0;
}

// ---------------------------------------------------------------------------
// phases of one plane step
//
// Work of a block per plane, as flat lists over the block's threads:
//   faces: ud on i in [i0-2, i0+TI] x j in [j0-1, j0+TJ], vd on
//          [i0-1, i0+TI] x [j0-2, j0+TJ], wd on [i0-1, i0+TI] x [j0-1, j0+TJ],
//          packed back to back;
//   flags: cells [i0-1, i0+TI] x [j0-1, j0+TJ];
//   finish: one tile cell per thread.
// A thread's items are the same for every plane, so their LDS indexes and
// global (i,j) offsets are worked out once per block (Plan); per plane only
// the plane's base address changes, and that is uniform across the block.
template <typename T, int TI, int TJ>
struct MpTile {
  typedef MpShared<T, TI, TJ> Sh;
  static constexpr int SI = Sh::SI, SJ = Sh::SJ, S = Sh::S, SC = Sh::SC, NT = TI * TJ;
  static constexpr int pad64(int n) { return (n + 63) / 64 * 64; }
  // The face lists are packed back to back without padding: ud over
  // (TI+3) x (TJ+2) faces, vd over (TI+2) x (TJ+3), wd over (TI+2) x (TJ+2);
  // only the two wavefronts that straddle a list boundary compute two kinds.
  // The flag list runs over whole LDS rows.
  static constexpr int NIU = TI + 3, NJU = TJ + 2, NIV = TI + 2, NJV = TJ + 3, NIW = TI + 2, NJW = TJ + 2;
  static constexpr int LU = NIU * NJU, LV = NIV * NJV, LW = NIW * NJW;
  static constexpr int NF = LU + LV + LW, RF = (NF + NT - 1) / NT;
  static constexpr int NJF = TJ + 2, NFL = SI * NJF, RFL = (NFL + NT - 1) / NT;
  static constexpr int RS = (S + NT - 1) / NT;
  static constexpr int INTERIOR = 1 << 30;

  struct Plan {
    int fp[RF];   // face item: LDS index | kind << 28 (0 ud, 1 vd, 2 wd), or -1
    int fo[RF];   // face item: global (i,j) offset of an owned face, or -1
    int lp[RFL];  // flag cell: LDS index | INTERIOR, or -1
    int so[RS];   // staged element: global (i,j) offset, or -1 outside the box
    int cp, co;   // finishing cell: LDS index, global (i,j) offset or -1
    int ci, cj;   // finishing cell indexes
  };

  HC_HD static int at(const MpGeom &g, int i, int j) { return (i - (g.i0 - 2)) + SI * (j - (g.j0 - 2)); }
  HC_HD static int ij(const MpGeom &g, int i, int j) { return i + 16 + g.px * (j + 1); }
  HC_HD static long plane(const MpGeom &g, int k) { return (long)g.px * g.py * (long)(k + 1); }

  // rot rotates the assignment of face and flag items to wavefronts, so
  // that the waves that carry an extra round differ from block to block; the
  // flag items are rotated by a further half block so that the waves with an
  // extra face round are not the ones with an extra flag round
  HC_HD static void plan(const MpGeom &g, int tid, int rot, Plan &pl) {
    const int nw = NT / 64;
    const int vt = (tid + NT - 64 * (rot % nw)) % NT;
    const int vf = (vt + 64 * (nw / 2)) % NT;
    for (int r = 0; r < RF; ++r) {
      int e = vt + r * NT, p = -1, o = -1, kind = 0;
      if (e < LU) {  // ud: faces i in [i0-2, i0+TI], rows [j0-1, j0+TJ]
        const int i = g.i0 - 2 + e % NIU, j = g.j0 - 1 + e / NIU;
        if (i >= 0 && i <= g.nx && j >= 1 && j <= g.ny) {
          p = at(g, i, j);
          if (j < g.j0 + TJ && ((i >= g.i0 && i < g.i0 + TI) || (i == 0 && g.i0 == 1))) o = ij(g, i, j);
        }
      } else if ((e -= LU) < LV) {  // vd: columns [i0-1, i0+TI], faces j in [j0-2, j0+TJ]
        kind = 1;
        const int i = g.i0 - 1 + e % NIV, j = g.j0 - 2 + e / NIV;
        if (i >= 1 && i <= g.nx && j >= 0 && j <= g.ny) {
          p = at(g, i, j);
          if (i < g.i0 + TI && i >= g.i0 && ((j >= g.j0 && j < g.j0 + TJ) || (j == 0 && g.j0 == 1))) o = ij(g, i, j);
        }
      } else if ((e -= LV) < LW) {  // wd: columns [i0-1, i0+TI], rows [j0-1, j0+TJ]
        kind = 2;
        const int i = g.i0 - 1 + e % NIW, j = g.j0 - 1 + e / NIW;
        if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) {
          p = at(g, i, j);
          if (i >= g.i0 && i < g.i0 + TI && j >= g.j0 && j < g.j0 + TJ) o = ij(g, i, j);
        }
      }
      pl.fp[r] = p < 0 ? -1 : (p | (kind << 28));
      pl.fo[r] = o;
    }
    for (int r = 0; r < RFL; ++r) {  // flags: cells [i0-1, i0+TI] x [j0-1, j0+TJ]
      const int e = vf + r * NT;
      const int i = g.i0 - 2 + e % SI, j = g.j0 - 1 + e / SI;
      const bool need = e < NFL && i >= g.i0 - 1 && i <= g.i0 + TI;
      pl.lp[r] = need ? (at(g, i, j) | ((i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) ? INTERIOR : 0)) : -1;
    }
    for (int r = 0; r < RS; ++r) {
      const int e = tid + r * NT;
      const int i = g.i0 - 2 + e % SI, j = g.j0 - 2 + e / SI;
      pl.so[r] = (e < S && i <= g.nx + 1 && j <= g.ny + 1) ? ij(g, i, j) : -1;
    }
    pl.ci = g.i0 + tid % TI;
    pl.cj = g.j0 + tid / TI;
    pl.cp = at(g, pl.ci, pl.cj);
    pl.co = (pl.ci <= g.nx && pl.cj <= g.ny) ? ij(g, pl.ci, pl.cj) : -1;
  }

  // The planes a step needs next (c(q+1), U(q+1), V(q+1), W(q)) are fetched
  // into registers one step ahead, so their HBM latency overlaps the
  // arithmetic of the current plane, and written to LDS at the next step.
  struct Prefetch {
    T c[RS], u[RS], v[RS], w[RS];
  };

  // elements outside the box (so < 0, or planes beyond -1 .. nz+1: the
  // prefetch of the last plane step asks for nz+2) stage as zero without a
  // load (loading a clamped in-box address instead measured 10% slower:
  // a third of the second staging round has no element)
  HC_HD static T fetch1(const MpGeom &g, const T *src, int so, int k) {
    return (so >= 0 && k >= -1 && k <= g.nz + 1) ? src[HC_MP_GIDX(plane(g, k) + so)] : (T)0;
  }

  HC_HD static void fetch(const MpGeom &g, const int (&so)[RS], const MpPass<T> &a, int q, Prefetch &pf) {
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      pf.c[r] = fetch1(g, a.cc, so[r], q + 1);
      pf.u[r] = fetch1(g, a.U, so[r], q + 1);
      pf.v[r] = fetch1(g, a.V, so[r], q + 1);
      pf.w[r] = fetch1(g, a.W, so[r], q);
    }
  }

  HC_HD static void commit(int tid, Sh &sh, int q, const Prefetch &pf) {
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int e = tid + r * NT;
      if (e < S) {
        sh.c[(q + 1) & 3][e] = pf.c[r];
        sh.U[(q + 1) & 1][e] = pf.u[r];
        sh.V[(q + 1) & 1][e] = pf.v[r];
        sh.W[q & 1][e] = pf.w[r];
      }
    }
  }

  // the planes the first step needs besides its own prefetch:
  // c(qs-1), c(qs), U(qs), V(qs), W(qs-1); all loads are issued before the
  // first LDS store
  HC_HD static void stage_initial(const MpGeom &g, int tid, const int (&so)[RS], Sh &sh, const MpPass<T> &a,
                                  int qs) {
    T v[5][RS];
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      v[0][r] = fetch1(g, a.cc, so[r], qs - 1);
      v[1][r] = fetch1(g, a.cc, so[r], qs);
      v[2][r] = fetch1(g, a.U, so[r], qs);
      v[3][r] = fetch1(g, a.V, so[r], qs);
      v[4][r] = fetch1(g, a.W, so[r], qs - 1);
    }
#pragma unroll
    for (int r = 0; r < RS; ++r) {
      const int e = tid + r * NT;
      if (e < S) {
        sh.c[(qs - 1) & 3][e] = v[0][r];
        sh.c[qs & 3][e] = v[1][r];
        sh.U[qs & 1][e] = v[2][r];
        sh.V[qs & 1][e] = v[3][r];
        sh.W[(qs - 1) & 1][e] = v[4][r];
      }
    }
  }

  // phase A: ud, vd, wd of plane q on the face ranges the flags need
  HC_HD static void faces(const MpGeom &g, const Plan &pl, Sh &sh, const MpPass<T> &a, int q, bool full) {
    const bool quv = full && q >= 1 && q <= g.nz;
    const bool qw = q >= 0 && q <= g.nz;
    const bool wuv = q >= g.ka && q <= g.kb;
    const bool ww = wuv || (q == 0 && g.ka == 1);
    const T *c0 = sh.c[q & 3], *cm = sh.c[(q - 1) & 3], *cq = sh.c[(q + 1) & 3];
    const T *u0 = sh.U[q & 1], *u1 = sh.U[(q + 1) & 1];
    const T *v0 = sh.V[q & 1], *v1 = sh.V[(q + 1) & 1];
    const T *w0 = sh.W[q & 1], *wm = sh.W[(q - 1) & 1];
    const long pq = plane(g, q);
#pragma unroll
    for (int r = 0; r < RF; ++r) {
      const int fp = pl.fp[r];
      if (fp < 0) continue;
      const int p = fp & ((1 << 28) - 1), kind = fp >> 28;
      auto C = [&](int di, int dj, int dk) {
        const T *b = dk < 0 ? cm : (dk > 0 ? cq : c0);
        return b[HC_MP_CHECK(p + di + SI * dj)];
      };
      auto Uf = [&](int di, int dj, int dk) { return (dk > 0 ? u1 : u0)[HC_MP_CHECK(p + di + SI * dj)]; };
      auto Vf = [&](int di, int dj, int dk) { return (dk > 0 ? v1 : v0)[HC_MP_CHECK(p + di + SI * dj)]; };
      auto Wf = [&](int di, int dj, int dk) { return (dk < 0 ? wm : w0)[HC_MP_CHECK(p + di + SI * dj)]; };
      if (kind == 0) {
        if (quv) {
          T v = mp_anti_u<T>(C, Uf, Vf, Wf);
          sh.ud[HC_MP_CHECKC(p)] = v;
          if (a.ud && wuv && pl.fo[r] >= 0) (a.ud + pq)[pl.fo[r]] = v;
        }
      } else if (kind == 1) {
        if (quv) {
          T v = mp_anti_v<T>(C, Uf, Vf, Wf);
          sh.vd[HC_MP_CHECKC(p)] = v;
          if (a.vd && wuv && pl.fo[r] >= 0) (a.vd + pq)[pl.fo[r]] = v;
        }
      } else {
        if (qw) {
          T v = mp_anti_w<T>(C, Uf, Vf, Wf);
          sh.wd[q & 1][HC_MP_CHECKC(p)] = v;
          if (a.wd && ww && pl.fo[r] >= 0) (a.wd + pq)[pl.fo[r]] = v;
        }
      }
    }
  }

  // phase B: limiter flags of plane q over the tile and a one-cell halo
  HC_HD static void flags(const MpGeom &g, const Plan &pl, Sh &sh, const MpPass<T> &a, int q) {
    const bool qin = q >= 1 && q <= g.nz;
    const T *c0 = sh.c[q & 3], *cm = sh.c[(q - 1) & 3], *cq = sh.c[(q + 1) & 3];
    const T *wm = sh.wd[(q - 1) & 1], *w0 = sh.wd[q & 1];
    unsigned char *fl = sh.fl[q & 1];
#pragma unroll
    for (int r = 0; r < RFL; ++r) {
      const int lp = pl.lp[r];
      if (lp < 0) continue;
      const int p = lp & (INTERIOR - 1);
      int f = 0;
      if (qin && (lp & INTERIOR)) {
        T cc = c0[p];
        T x = cc + mp_donor<T>(cc, c0[HC_MP_CHECK(p - 1)], c0[HC_MP_CHECK(p + 1)], c0[HC_MP_CHECK(p - SI)],
                               c0[HC_MP_CHECK(p + SI)], cm[p], cq[p], sh.ud[HC_MP_CHECK(p - 1)], sh.ud[p],
                               sh.vd[HC_MP_CHECK(p - SI)], sh.vd[p], wm[p], w0[p], false, 0, 0, 0, 0, 0, 0, 0);
        if (x <= (T)a.cmin) f = 1;
        else if (x >= (T)a.cmax) f = 2;
      }
      fl[HC_MP_CHECKC(p)] = (unsigned char)f;
    }
  }

  // phase C: after the flags of plane q are in LDS, take the own column's
  // values of plane q into registers and finish plane kk = q-1
  HC_HD static void finish(const MpGeom &g, const Plan &pl, Sh &sh, const MpPass<T> &a, int q, MpRegs<T> &r) {
    finish_cell(g, pl.cp, pl.co, pl.ci, pl.cj, sh, a, q, r);
  }

  // the finishing step of the cell (ci, cj) at LDS index p and global (i,j)
  // offset co (-1: outside the box)
  HC_HD static void finish_cell(const MpGeom &g, int p, int co, int ci, int cj, Sh &sh, const MpPass<T> &a, int q,
                                MpRegs<T> &r) {
    const T nu0 = sh.ud[HC_MP_CHECKC(p - 1)], nu1 = sh.ud[HC_MP_CHECKC(p)], nv0 = sh.vd[HC_MP_CHECKC(p - SI)], nv1 = sh.vd[p];
    const T nw = sh.wd[q & 1][p];
    const int nf = sh.fl[q & 1][p];
    const int kk = q - 1;
    if (kk >= g.ka && kk <= g.kb && co >= 0) {
      const T *c0 = sh.c[kk & 3];
      const unsigned char *f0 = sh.fl[kk & 1];
      T cc = c0[p];
      T ce = mp_donor<T>(cc, c0[HC_MP_CHECK(p - 1)], c0[HC_MP_CHECK(p + 1)], c0[HC_MP_CHECK(p - SI)],
                         c0[HC_MP_CHECK(p + SI)], r.cB, sh.c[(kk + 1) & 3][p], r.u0, r.u1, r.v0,
                         r.v1, r.wA, r.wB, true, r.fB, f0[HC_MP_CHECKC(p - 1)], f0[HC_MP_CHECKC(p + 1)],
                         f0[HC_MP_CHECKC(p - SI)], f0[HC_MP_CHECKC(p + SI)], r.fA, nf);
      T cn = cc + ce;
      const long pk = plane(g, kk);
      if (!a.last) {
        (a.cout + pk)[co] = cn;
      } else {
        const int i = ci, j = cj;
        // the boundary images take the pre-sponge value of the cells next
        // to a boundary; only those are stored
        if (i == 1 || i == g.nx || j == 1 || j == g.ny || kk == 1 || kk == g.nz) (a.cpre + pk)[co] = cn;
        // sponge: c = c - 0.5*alphaee*tau3*(cp - c_bkg(k)), cp = the field at entry
        double ae = fmax(a.alphae0[i + (g.nx + 2) * j], (a.alphaw[kk + 1] + a.alphaw[kk]) / 2.0);
        T cp = (a.cout + pk)[co];
        T rr = cn - (T)1 * (T)0.5 * (T)ae * (T)a.tau3 * (cp - (T)a.cbkg[kk]);
        double qv = 0.0;
        if (kk == g.nz && a.map_valid && i >= a.nxs && i <= a.nxe && j >= a.nys && j <= a.nye)
          qv = a.qpatch[(i - a.nxs) + (a.nxe - a.nxs + 1) * (j - a.nys)];
        rr = rr + (T)a.tau3 * (T)qv / (T)a.dxdydz / (T)1.0 * (T)1;
        (a.cout + pk)[co] = rr;
      }
    }
    r.u0 = nu0; r.u1 = nu1; r.v0 = nv0; r.v1 = nv1;
    r.wA = r.wB; r.wB = nw;
    r.fA = r.fB; r.fB = nf;
    r.cB = sh.c[kk & 3][p];
  }
};

}  // namespace hc
