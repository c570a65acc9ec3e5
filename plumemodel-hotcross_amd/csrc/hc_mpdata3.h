// hc_mpdata3.h - the MPDATA corrective pass with each face's donor-cell flux
// formed once.
//
// The pass of MpTile2 (hc_mpdata2.h: antidiffusive velocities of
// mpdata-hotcross80.f:218-326, limiter flags of :660-690, the limited
// donor-cell step of :705-720, on the last pass the sponge and source of
// :356-381), with the same thread assignment and the same staged planes, but
// with the donor-cell arithmetic organised by face instead of by cell.
//
// The donor-cell step of a cell sums six face fluxes,
//   F = ((x0 + y0) + z1) - ((x1 + y1) + z0),
// where the flux through a face is one expression of the face's velocity u
// and the two cells it separates, u+ * c(P) + u- * c(N) (u+ = (u + |u|)/2 =
// max(u, 0), u- = (u - |u|)/2 = min(u, 0); P is the cell whose value u+
// carries: the left cell of an x or y face, the upper cell of a z face).  The
// two cells that share a face evaluate that same expression, and the limited
// step zeroes u+ or u- on conditions that are the same from either side:
// u+ when P is at its lower bound or N at its upper one, u- when N is at its
// lower bound or P at its upper one.  So each face's products u+ c(P) and
// u- c(N) are formed once, right after the face velocity:
//   - their sum (the unlimited flux) goes to LDS for the flags of plane q;
//   - the products of the own faces stay in registers and, once the flags of
//     their plane are known, give the limited flux (a product is replaced by
//     zero where its condition holds), which the next finishing step sums.
// The per-cell flags and finishing steps then only add six fluxes.  The
// values are the reference's: the same products of the same operands, added
// in the same order (max(u, 0) and min(u, 0) equal (u + |u|)/2 and
// (u - |u|)/2 exactly; zeroing a velocity half zeroes its product).
//
// Per plane step q:
//   A (after the staging barrier): the own cell's three faces and one halo
//     face of plane q (ud, vd, wd as in MpTile2, written to HBM where the next
//     pass needs them) with their products; the unlimited fluxes of the x and
//     y faces into fx, fy, of the halo z faces into fz; the limited fluxes of
//     plane q-1 through the own x and y faces (and, on the threads whose halo
//     face is a tile's lower x or y boundary face, through that face) into
//     lx, ly;
//   B: the flags of plane q (own cell and one halo cell);
//   C: the finishing step of plane q-1: the limited flux through the own
//     column's upper z face (its products in registers, both flags known
//     now), the lower one rolled from the step before, the four others from
//     lx, ly.
// The tile logic compiles for the host too, where a serial loop over the
// threads of a block runs each phase (tests/mpcheck).
#pragma once
#include "hc_mpdata2.h"

// a host build may define HC_MP_CHECKZ(x) to range-check the halo z-flux index
#ifndef HC_MP_CHECKZ
#define HC_MP_CHECKZ(x) (x)
#endif

namespace hc {

// The concentration ring holds planes q-1, q, q+1 (slot q mod 3): the
// finishing step takes the own cell's value from a register, so the staging
// of plane q+2 may overwrite plane q-1 while other threads still finish.
// fz holds the z fluxes of the halo columns only (the own column's stay in
// registers), one entry per halo flag cell, in the order of MpTile2's halo
// flag and halo wd face lists (the same cells).
template <typename T, int TI, int TJ>
struct Mp3Shared {
  static constexpr int SI = TI + 4, SJ = TJ + 4, S = SI * SJ, SC = SI * (TJ + 3);
  static constexpr int NHL = 2 * (TJ + 2) + 2 * TI;  // MpTile2's halo flag cells
  T c[3][S];
  T U[2][S], V[2][S], W[2][S];
  T fx[SC], fy[SC];   // unlimited flux through the ud / vd faces of plane q
  T fz[2][NHL];       // through the halo columns' wd faces of planes q-1, q
  T lx[SC], ly[SC];   // limited flux through the ud / vd faces of plane q-1
  unsigned char fl[2][SC];
};

HC_HD inline int mp_slot3(int q) { return (q + 3) % 3; }

// per-thread state that rolls from plane to plane
template <typename T>
struct Mp3Regs {
  T xP, xN, yP, yN;  // products of the own ud and vd faces of the latest plane
  T zP, zN;          // products of the own wd face of plane q-1 (the finished cell's upper face)
  T zPn, zNn;        // the same of plane q
  T fzb, fzt;        // unlimited flux through the own wd faces of planes q-1 and q
  T zlo;             // limited flux through the own column's wd face below the plane being finished
  T hP, hN;          // products of the halo face whose limited flux a tile cell needs
  T cc, ccn;         // own concentration of planes q-1 and q
  int fB, fC;        // own flags of planes q-1 and q
};

HC_HD inline double mp_pos(double u) { return fmax(u, 0.0); }
HC_HD inline float mp_pos(float u) { return fmaxf(u, 0.0f); }
HC_HD inline double mp_neg(double u) { return fmin(u, 0.0); }
HC_HD inline float mp_neg(float u) { return fminf(u, 0.0f); }

// the limited flux of a face from its two products and the flags of P and N
// (1: at or below the lower bound, 2: at or above the upper one)
template <typename T>
HC_HD inline T mp_lim(T pP, T pN, int fP, int fN) {
  const T a = ((fP & 1) | (fN & 2)) ? (T)0 : pP;
  const T b = ((fN & 1) | (fP & 2)) ? (T)0 : pN;
  return a + b;
}

// The FP32 form's antidiffusive velocities: the formulas of mp_anti_u/v/w
// (mpdata-hotcross80.f:218-326) with the products and sums grouped for fused
// multiply-adds.  (|U| - U^2) is one fused multiply-add; the second and third
// terms scale the face velocity once, by 0.5 * 0.25 (a power of two, so
// folding it changes no rounding), and subtract through fused multiply-adds;
// the four-point velocity averages add in pairs.  The quotients go through
// the hardware reciprocal, as in the FP32 form of mp_ratio.
HC_HD inline float mp_rcp(float b) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(b);
#else
  return 1.0f / b;
#endif
}
template <class FC, class FU, class FV, class FW>
HC_HD inline float mp_anti_u32(const FC &C, const FU &U, const FV &V, const FW &W) {
  const float eps = 1.0e-30f;
  const float Uc = U(0, 0, 0), h = 0.125f * Uc;
  const float c1 = C(1, 0, 0), c0 = C(0, 0, 0);
  float r = fmaf(-Uc, Uc, fabsf(Uc)) * (c1 - c0) * mp_rcp(c1 + c0 + eps);
  float ca = C(1, 1, 0) + C(0, 1, 0), cb = C(1, -1, 0) + C(0, -1, 0);
  r = fmaf(-(h * (ca - cb) * mp_rcp(ca + cb + eps)), (V(1, 0, 0) + V(0, 0, 0)) + (V(1, -1, 0) + V(0, -1, 0)), r);
  ca = C(1, 0, -1) + C(0, 0, -1);
  cb = C(1, 0, 1) + C(0, 0, 1);
  return fmaf(-(h * (ca - cb) * mp_rcp(ca + cb + eps)), (W(1, 0, 0) + W(0, 0, 0)) + (W(1, 0, -1) + W(0, 0, -1)), r);
}
template <class FC, class FU, class FV, class FW>
HC_HD inline float mp_anti_v32(const FC &C, const FU &U, const FV &V, const FW &W) {
  const float eps = 1.0e-30f;
  const float Vc = V(0, 0, 0), h = 0.125f * Vc;
  const float c1 = C(0, 1, 0), c0 = C(0, 0, 0);
  float r = fmaf(-Vc, Vc, fabsf(Vc)) * (c1 - c0) * mp_rcp(c1 + c0 + eps);
  float ca = C(1, 1, 0) + C(1, 0, 0), cb = C(-1, 1, 0) + C(-1, 0, 0);
  r = fmaf(-(h * (ca - cb) * mp_rcp(ca + cb + eps)), (U(0, 1, 0) + U(0, 0, 0)) + (U(-1, 1, 0) + U(-1, 0, 0)), r);
  ca = C(0, 1, -1) + C(0, 0, -1);
  cb = C(0, 1, 1) + C(0, 0, 1);
  return fmaf(-(h * (ca - cb) * mp_rcp(ca + cb + eps)), (W(0, 0, 0) + W(0, 1, 0)) + (W(0, 0, -1) + W(0, 1, -1)), r);
}
template <class FC, class FU, class FV, class FW>
HC_HD inline float mp_anti_w32(const FC &C, const FU &U, const FV &V, const FW &W) {
  const float eps = 1.0e-30f;
  const float Wc = W(0, 0, 0), h = 0.125f * Wc;
  const float c0 = C(0, 0, 0), c1 = C(0, 0, 1);
  float r = fmaf(-Wc, Wc, fabsf(Wc)) * (c0 - c1) * mp_rcp(c0 + c1 + eps);
  float ca = C(1, 0, 0) + C(1, 0, 1), cb = C(-1, 0, 0) + C(-1, 0, 1);
  r = fmaf(-(h * (ca - cb) * mp_rcp(ca + cb + eps)), (U(0, 0, 0) + U(0, 0, 1)) + (U(-1, 0, 1) + U(-1, 0, 0)), r);
  ca = C(0, 1, 1) + C(0, 1, 0);
  cb = C(0, -1, 1) + C(0, -1, 0);
  return fmaf(-(h * (ca - cb) * mp_rcp(ca + cb + eps)), (V(0, 0, 0) + V(0, 0, 1)) + (V(0, -1, 1) + V(0, -1, 0)), r);
}

template <typename T, int TI, int TJ>
struct MpTile3 {
  typedef Mp3Shared<T, TI, TJ> Sh;
  typedef MpTile2<T, TI, TJ> T2;  // thread assignment and index functions
  typedef typename T2::Plan Plan;
  static constexpr int SI = Sh::SI, S = Sh::S, SC = Sh::SC, NT = TI * TJ;
  static constexpr int RS = T2::RS;
  static constexpr int INTERIOR = T2::INTERIOR;
  static_assert(RS == (S + NT - 1) / NT, "staging rounds");
  static_assert(Sh::NHL == T2::NHL, "halo flag cells");

  struct Prefetch {
    T c[RS], u[RS], v[RS], w[RS];
  };

  // which halo face's limited flux a tile cell reads: 1 for ud at the
  // column i0-1 (rows of the tile), 2 for vd at the row j0-1 (columns of the
  // tile), else 0 (also where the face is outside the grid: no cell it
  // bounds is finished)
  HC_HD static int halo_limited(const MpGeom &g, int tid, const Plan &pl) {
    if (tid >= T2::NHF || pl.hp < 0) return 0;
    int kind, i, j;
    T2::halo_face(g, tid, kind, i, j);
    if (kind == 0 && i == g.i0 - 1 && j >= g.j0 && j < g.j0 + TJ) return 1;
    if (kind == 1 && j == g.j0 - 1 && i >= g.i0 && i < g.i0 + TI) return 2;
    return 0;
  }

  // staging of c, U, V, W: the planes MpTile fetches and commits, into this
  // tile's shared layout
  HC_HD static T fetch1(const MpGeom &g, const T *src, int so, int k) {
    return (so >= 0 && k >= -1 && k <= g.nz + 1) ? src[HC_MP_GIDX(T2::plane(g, k) + so)] : (T)0;
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
        sh.c[mp_slot3(q + 1)][e] = pf.c[r];
        sh.U[(q + 1) & 1][e] = pf.u[r];
        sh.V[(q + 1) & 1][e] = pf.v[r];
        sh.W[q & 1][e] = pf.w[r];
      }
    }
  }
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
        sh.c[mp_slot3(qs - 1)][e] = v[0][r];
        sh.c[mp_slot3(qs)][e] = v[1][r];
        sh.U[qs & 1][e] = v[2][r];
        sh.V[qs & 1][e] = v[3][r];
        sh.W[(qs - 1) & 1][e] = v[4][r];
      }
    }
  }

  // the three face formulas at LDS index p of plane q
  template <int KIND>
  HC_HD static T face(const Sh &sh, int p, int q) {
    const T *c0 = sh.c[mp_slot3(q)], *cm = sh.c[mp_slot3(q - 1)], *cq = sh.c[mp_slot3(q + 1)];
    const T *u0 = sh.U[q & 1], *u1 = sh.U[(q + 1) & 1];
    const T *v0 = sh.V[q & 1], *v1 = sh.V[(q + 1) & 1];
    const T *w0 = sh.W[q & 1], *wm = sh.W[(q - 1) & 1];
    auto C = [&](int di, int dj, int dk) {
      const T *b = dk < 0 ? cm : (dk > 0 ? cq : c0);
      return b[HC_MP_CHECK(p + di + SI * dj)];
    };
    auto Uf = [&](int di, int dj, int dk) { return (dk > 0 ? u1 : u0)[HC_MP_CHECK(p + di + SI * dj)]; };
    auto Vf = [&](int di, int dj, int dk) { return (dk > 0 ? v1 : v0)[HC_MP_CHECK(p + di + SI * dj)]; };
    auto Wf = [&](int di, int dj, int dk) { return (dk < 0 ? wm : w0)[HC_MP_CHECK(p + di + SI * dj)]; };
    if constexpr (sizeof(T) == 4) {  // the FP32 form
      if (KIND == 0) return mp_anti_u32(C, Uf, Vf, Wf);
      if (KIND == 1) return mp_anti_v32(C, Uf, Vf, Wf);
      return mp_anti_w32(C, Uf, Vf, Wf);
    } else {
      if (KIND == 0) return mp_anti_u<T>(C, Uf, Vf, Wf);
      if (KIND == 1) return mp_anti_v<T>(C, Uf, Vf, Wf);
      return mp_anti_w<T>(C, Uf, Vf, Wf);
    }
  }

  // phase A
  HC_HD static void faces(const MpGeom &g, const Plan &pl, int hl, int hz, Sh &sh, const MpPass<T> &a, int q,
                          bool full, Mp3Regs<T> &r) {
    const bool quv = full && q >= 1 && q <= g.nz;
    const bool qw = q >= 0 && q <= g.nz;
    const bool wuv = q >= g.ka && q <= g.kb;
    const bool ww = wuv || (q == 0 && g.ka == 1);
    const bool lim = q - 1 >= g.ka && q - 1 <= g.kb;  // plane q-1 is finished at this step
    const long pq = T2::plane(g, q);
    const int p = pl.po;
    const bool own = pl.go >= 0;
    const T *c0 = sh.c[mp_slot3(q)], *cq = sh.c[mp_slot3(q + 1)];
    const int hp = pl.hp & ((1 << 28) - 1), kind = pl.hp >> 28;
    r.ccn = c0[HC_MP_CHECK(p)];
    if (lim) {  // limited fluxes of plane q-1 (the products are still those of plane q-1)
      const unsigned char *fm = sh.fl[(q - 1) & 1];
      sh.lx[HC_MP_CHECKC(p)] = mp_lim(r.xP, r.xN, r.fB, (int)fm[HC_MP_CHECKC(p + 1)]);
      sh.ly[HC_MP_CHECKC(p)] = mp_lim(r.yP, r.yN, r.fB, (int)fm[HC_MP_CHECKC(p + SI)]);
      if (hl == 1) sh.lx[HC_MP_CHECKC(hp)] = mp_lim(r.hP, r.hN, (int)fm[HC_MP_CHECKC(hp)], (int)fm[HC_MP_CHECKC(hp + 1)]);
      if (hl == 2) sh.ly[HC_MP_CHECKC(hp)] = mp_lim(r.hP, r.hN, (int)fm[HC_MP_CHECKC(hp)], (int)fm[HC_MP_CHECKC(hp + SI)]);
    }
    if (quv) {
      const T ud = face<0>(sh, p, q);
      const T vd = face<1>(sh, p, q);
      if (own && wuv && a.ud) {
        (a.ud + pq)[pl.go] = ud;
        (a.vd + pq)[pl.go] = vd;
      }
      const T cc = r.ccn;
      r.xP = mp_pos(ud) * cc;
      r.xN = mp_neg(ud) * c0[HC_MP_CHECK(p + 1)];
      r.yP = mp_pos(vd) * cc;
      r.yN = mp_neg(vd) * c0[HC_MP_CHECK(p + SI)];
      sh.fx[HC_MP_CHECKC(p)] = r.xP + r.xN;
      sh.fy[HC_MP_CHECKC(p)] = r.yP + r.yN;
    }
    r.fzb = r.fzt;
    if (qw) {
      const T wd = face<2>(sh, p, q);
      if (own && ww && a.wd) (a.wd + pq)[pl.go] = wd;
      r.zPn = mp_pos(wd) * cq[HC_MP_CHECK(p)];
      r.zNn = mp_neg(wd) * c0[HC_MP_CHECK(p)];
      r.fzt = r.zNn + r.zPn;
    }
    if (pl.hp >= 0) {
      if (kind == 0) {
        if (quv) {
          const T v = face<0>(sh, hp, q);
          if (a.ud && wuv && pl.ho >= 0) (a.ud + pq)[pl.ho] = v;
          const T pP = mp_pos(v) * c0[HC_MP_CHECK(hp)], pN = mp_neg(v) * c0[HC_MP_CHECK(hp + 1)];
          sh.fx[HC_MP_CHECKC(hp)] = pP + pN;
          r.hP = pP;
          r.hN = pN;
        }
      } else if (kind == 1) {
        if (quv) {
          const T v = face<1>(sh, hp, q);
          if (a.vd && wuv && pl.ho >= 0) (a.vd + pq)[pl.ho] = v;
          const T pP = mp_pos(v) * c0[HC_MP_CHECK(hp)], pN = mp_neg(v) * c0[HC_MP_CHECK(hp + SI)];
          sh.fy[HC_MP_CHECKC(hp)] = pP + pN;
          r.hP = pP;
          r.hN = pN;
        }
      } else if (qw) {
        const T v = face<2>(sh, hp, q);
        sh.fz[q & 1][HC_MP_CHECKZ(hz)] = mp_neg(v) * c0[HC_MP_CHECK(hp)] + mp_pos(v) * cq[HC_MP_CHECK(hp)];
      }
    }
  }

  // phase B: the flags of plane q, the own cell's and one halo cell's
  HC_HD static void flags(const MpGeom &g, const Plan &pl, int lz, Sh &sh, const MpPass<T> &a, int q,
                          Mp3Regs<T> &r) {
    const bool qin = q >= 1 && q <= g.nz;
    const T *c0 = sh.c[mp_slot3(q)];
    unsigned char *fl = sh.fl[q & 1];
    {
      const int p = pl.po;
      const T F = ((sh.fx[HC_MP_CHECKC(p - 1)] + sh.fy[HC_MP_CHECKC(p - SI)]) + r.fzt) -
                  ((sh.fx[HC_MP_CHECKC(p)] + sh.fy[HC_MP_CHECKC(p)]) + r.fzb);
      const T x = r.ccn + F;
      const int f = x <= (T)a.cmin ? 1 : (x >= (T)a.cmax ? 2 : 0);  // formed everywhere, kept where the cell exists
      r.fC = (qin && pl.go >= 0) ? f : 0;
      fl[HC_MP_CHECKC(p)] = (unsigned char)r.fC;
      // the flags of the planes this block finishes, for a caller that keeps
      // them (the adjoint's tape)
      if (a.fl && q >= g.ka && q <= g.kb && pl.go >= 0) (a.fl + T2::plane(g, q))[HC_MP_GIDX(pl.go)] = (unsigned char)f;
    }
    if (pl.lp >= 0) {
      const int p = pl.lp & (INTERIOR - 1);
      int f = 0;
      if (qin && (pl.lp & INTERIOR)) {
        const T F = ((sh.fx[HC_MP_CHECKC(p - 1)] + sh.fy[HC_MP_CHECKC(p - SI)]) + sh.fz[q & 1][HC_MP_CHECKZ(lz)]) -
                    ((sh.fx[HC_MP_CHECKC(p)] + sh.fy[HC_MP_CHECKC(p)]) + sh.fz[(q - 1) & 1][HC_MP_CHECKZ(lz)]);
        const T x = c0[HC_MP_CHECK(p)] + F;
        f = x <= (T)a.cmin ? 1 : (x >= (T)a.cmax ? 2 : 0);
      }
      fl[HC_MP_CHECKC(p)] = (unsigned char)f;
    }
  }

  // phase C: the finishing step of plane kk = q-1
  HC_HD static void finish(const MpGeom &g, const Plan &pl, Sh &sh, const MpPass<T> &a, int q, Mp3Regs<T> &r) {
    const int kk = q - 1;
    const int p = pl.po;
    // the own column's face between kk and kk+1: P = the cell above (flag of
    // plane q), N = the cell itself
    const T z1 = mp_lim(r.zP, r.zN, r.fC, r.fB);
    if (kk >= g.ka && kk <= g.kb && pl.go >= 0) {
      const T cc = r.cc;
      const T F = ((sh.lx[HC_MP_CHECKC(p - 1)] + sh.ly[HC_MP_CHECKC(p - SI)]) + z1) -
                  ((sh.lx[HC_MP_CHECKC(p)] + sh.ly[HC_MP_CHECKC(p)]) + r.zlo);
      const T cn = cc + F;
      const long pk = T2::plane(g, kk);
      const int co = pl.go;
      if (!a.last) {
        (a.cout + pk)[co] = cn;
      } else {
        const int i = pl.ci, j = pl.cj;
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
    r.zlo = z1;
    r.cc = r.ccn;
    r.zP = r.zPn;
    r.zN = r.zNn;
    r.fB = r.fC;
  }
};

template <typename T, int TI, int TJ>
struct Mp3State {
  typename MpTile3<T, TI, TJ>::Plan pl;
  typename MpTile3<T, TI, TJ>::Prefetch pf;
  Mp3Regs<T> r;
  int hl;  // halo_limited
  int hz;  // a halo wd face's entry in fz
  int lz;  // a halo flag cell's entry in fz
};

// One block of the pass, for a block executor X (as mp2_block)
template <typename T, int TI, int TJ, class X>
HC_HD inline void mp3_block(X &x, const MpGeom &g, typename MpTile3<T, TI, TJ>::Sh &sh, const MpPass<T> &a) {
  typedef MpTile3<T, TI, TJ> Tl;
  const int qs = g.ka - 2 > 0 ? g.ka - 2 : 0;
  x.each([&](int tid, auto &ps) {
    Tl::T2::plan(g, tid, ps.pl);
    ps.hl = Tl::halo_limited(g, tid, ps.pl);
    ps.hz = tid - (Tl::T2::HU + Tl::T2::HV);
    ps.lz = tid - (Tl::NT - Tl::T2::NHL);
    ps.r = Mp3Regs<T>{};
    Tl::stage_initial(g, tid, ps.pl.so, sh, a, qs);
    Tl::fetch(g, ps.pl.so, a, qs, ps.pf);
  });
  for (int q = qs; q <= g.kb + 1; ++q) {
    // the staging overwrites the slots of c(q-2), U(q-1), V(q-1), W(q-2),
    // which only the face phase of the previous step read, before its
    // second barrier
    x.each([&](int tid, auto &ps) { Tl::commit(tid, sh, q, ps.pf); });
    x.sync();
    x.each([&](int tid, auto &ps) {
      if (q <= g.kb) Tl::fetch(g, ps.pl.so, a, q + 1, ps.pf);  // in flight during this plane
      Tl::faces(g, ps.pl, ps.hl, ps.hz, sh, a, q, q >= g.ka - 1, ps.r);
    });
    x.sync();
    x.each([&](int tid, auto &ps) {
      if (q >= g.ka - 1) Tl::flags(g, ps.pl, ps.lz, sh, a, q, ps.r);
    });
    x.sync();
    x.each([&](int tid, auto &ps) { Tl::finish(g, ps.pl, sh, a, q, ps.r); });
  }
}

}  // namespace hc
