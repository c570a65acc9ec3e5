// hc_mpdata2.h - the MPDATA corrective pass as a k-marching LDS tile with a
// fixed per-thread work assignment.
//
// The same pass as MpTile (hc_mpdata.h: antidiffusive velocities of
// mpdata-hotcross80.f:218-326, limiter flags of :660-690, the limited
// donor-cell step of :705-720, on the last pass the sponge and source of
// :356-381), the same LDS rings and the same formulas in the same order (so
// FP64 rounds as the reference), with the work of a plane dealt out
// differently:
//   - every thread forms the three faces of its own cell (ud at i+1/2, vd at
//     j+1/2, wd at k+1/2) in straight-line code, one formula per round;
//   - the faces of the one- and two-cell halo that the flags need (352 for a
//     32 x 16 tile) form one more round over the first threads, sorted by
//     kind so that at most two wavefronts mix two formulas;
//   - every thread forms its own cell's flag; the 100 halo flags go to the
//     last threads, which the halo face round left idle;
//   - the finishing step is the own cell's, as in MpTile.
// The per-thread LDS indexes and global offsets are fixed for the block and
// worked out once (Plan2).  The tile logic compiles for the host too, where a
// serial loop over the threads of a block runs each phase (tests/mpcheck).
#pragma once
#include "hc_mpdata.h"

namespace hc {

template <typename T, int TI, int TJ>
struct MpTile2 {
  typedef MpShared<T, TI, TJ> Sh;
  typedef MpTile<T, TI, TJ> T1;  // staging (fetch, commit, stage_initial) is shared
  static constexpr int SI = Sh::SI, S = Sh::S, SC = Sh::SC, NT = TI * TJ;
  static constexpr int RS = T1::RS;
  // halo faces: ud on columns i0-2, i0-1, i0+TI (rows j0-1 .. j0+TJ) and on
  // rows j0-1, j0+TJ (columns i0 .. i0+TI-1); vd on columns i0-1, i0+TI
  // (rows j0-2 .. j0+TJ) and rows j0-2, j0-1, j0+TJ; wd on columns i0-1,
  // i0+TI (rows j0-1 .. j0+TJ) and rows j0-1, j0+TJ
  static constexpr int HU = 3 * (TJ + 2) + 2 * TI, HV = 2 * (TJ + 3) + 3 * TI, HW = 2 * (TJ + 2) + 2 * TI;
  static constexpr int NHF = HU + HV + HW;
  // halo flags: columns i0-1, i0+TI (rows j0-1 .. j0+TJ), rows j0-1, j0+TJ
  static constexpr int NHL = 2 * (TJ + 2) + 2 * TI;
  static_assert(NHF <= NT, "one round of halo faces");
  static_assert(NHL <= NT, "one round of halo flags");
  static constexpr int INTERIOR = 1 << 30;

  struct Plan {
    int po;       // own cell's LDS index
    int go;       // own cell's global (i,j) offset, or -1 outside the box
    int hp;       // halo face: LDS index | kind << 28 (0 ud, 1 vd, 2 wd), or -1
    int ho;       // halo face: global (i,j) offset of a face this tile writes (i = 0, j = 0), or -1
    int lp;       // halo flag: LDS index | INTERIOR, or -1
    int so[RS];   // staged element: global (i,j) offset, or -1 outside the box
    int ci, cj;   // own cell
  };

  HC_HD static int at(const MpGeom &g, int i, int j) { return (i - (g.i0 - 2)) + SI * (j - (g.j0 - 2)); }
  HC_HD static int ij(const MpGeom &g, int i, int j) { return i + 16 + g.px * (j + 1); }
  HC_HD static long plane(const MpGeom &g, int k) { return (long)g.px * g.py * (long)(k + 1); }

  // halo face number e (0 .. NHF-1) -> (kind, i, j)
  HC_HD static void halo_face(const MpGeom &g, int e, int &kind, int &i, int &j) {
    if (e < HU) {
      kind = 0;
      if (e < 3 * (TJ + 2)) {
        const int c = e / (TJ + 2);
        i = c == 0 ? g.i0 - 2 : c == 1 ? g.i0 - 1 : g.i0 + TI;
        j = g.j0 - 1 + e % (TJ + 2);
      } else {
        e -= 3 * (TJ + 2);
        i = g.i0 + e % TI;
        j = e / TI == 0 ? g.j0 - 1 : g.j0 + TJ;
      }
    } else if ((e -= HU) < HV) {
      kind = 1;
      if (e < 2 * (TJ + 3)) {
        i = e / (TJ + 3) == 0 ? g.i0 - 1 : g.i0 + TI;
        j = g.j0 - 2 + e % (TJ + 3);
      } else {
        e -= 2 * (TJ + 3);
        i = g.i0 + e % TI;
        const int r = e / TI;
        j = r == 0 ? g.j0 - 2 : r == 1 ? g.j0 - 1 : g.j0 + TJ;
      }
    } else {
      e -= HV;
      kind = 2;
      if (e < 2 * (TJ + 2)) {
        i = e / (TJ + 2) == 0 ? g.i0 - 1 : g.i0 + TI;
        j = g.j0 - 1 + e % (TJ + 2);
      } else {
        e -= 2 * (TJ + 2);
        i = g.i0 + e % TI;
        j = e / TI == 0 ? g.j0 - 1 : g.j0 + TJ;
      }
    }
  }

  HC_HD static void plan(const MpGeom &g, int tid, Plan &pl) {
    for (int r = 0; r < RS; ++r) {  // staged planes: rows j0-2 .. j0+TJ+1, columns i0-2 .. i0+TI+1
      const int e = tid + r * NT;
      const int i = g.i0 - 2 + e % SI, j = g.j0 - 2 + e / SI;
      pl.so[r] = (e < S && i <= g.nx + 1 && j <= g.ny + 1) ? ij(g, i, j) : -1;
    }
    pl.ci = g.i0 + tid % TI;
    pl.cj = g.j0 + tid / TI;
    pl.po = at(g, pl.ci, pl.cj);
    pl.go = (pl.ci <= g.nx && pl.cj <= g.ny) ? ij(g, pl.ci, pl.cj) : -1;
    pl.hp = pl.ho = -1;
    if (tid < NHF) {
      int kind, i, j;
      halo_face(g, tid, kind, i, j);
      bool ok;
      if (kind == 0) ok = i >= 0 && i <= g.nx && j >= 1 && j <= g.ny;
      else if (kind == 1) ok = i >= 1 && i <= g.nx && j >= 0 && j <= g.ny;
      else ok = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny;
      if (ok) {
        pl.hp = at(g, i, j) | (kind << 28);
        // the domain's first face in x (ud, i = 0) and in y (vd, j = 0)
        // belongs to the first tile of its row or column
        if ((kind == 0 && i == 0 && j >= g.j0 && j < g.j0 + TJ) || (kind == 1 && j == 0 && i >= g.i0 && i < g.i0 + TI))
          pl.ho = ij(g, i, j);
      }
    }
    pl.lp = -1;
    const int e = tid - (NT - NHL);
    if (e >= 0) {
      int i, j;
      if (e < 2 * (TJ + 2)) {
        i = e / (TJ + 2) == 0 ? g.i0 - 1 : g.i0 + TI;
        j = g.j0 - 1 + e % (TJ + 2);
      } else {
        const int f = e - 2 * (TJ + 2);
        i = g.i0 + f % TI;
        j = f / TI == 0 ? g.j0 - 1 : g.j0 + TJ;
      }
      pl.lp = at(g, i, j) | ((i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) ? INTERIOR : 0);
    }
  }

  // the three face formulas at LDS index p of plane q
  template <int KIND>
  HC_HD static T face(const Sh &sh, int p, int q) {
    const T *c0 = sh.c[q & 3], *cm = sh.c[(q - 1) & 3], *cq = sh.c[(q + 1) & 3];
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
    if (KIND == 0) return mp_anti_u<T>(C, Uf, Vf, Wf);
    if (KIND == 1) return mp_anti_v<T>(C, Uf, Vf, Wf);
    return mp_anti_w<T>(C, Uf, Vf, Wf);
  }

  // phase A: the own cell's faces of plane q, then one halo face.  The own
  // faces are formed for every thread of the block, also for cells past nx
  // or ny, where no face exists: every staged value is finite (elements
  // outside the box stage as zero), nothing reads those slots for a cell it
  // updates, and the per-thread guard cost more than the work (only the
  // uniform guards of the plane number remain, and the stores to HBM are
  // guarded per thread).
  HC_HD static void faces(const MpGeom &g, const Plan &pl, Sh &sh, const MpPass<T> &a, int q, bool full) {
    const bool quv = full && q >= 1 && q <= g.nz;
    const bool qw = q >= 0 && q <= g.nz;
    const bool wuv = q >= g.ka && q <= g.kb;
    const bool ww = wuv || (q == 0 && g.ka == 1);
    const long pq = plane(g, q);
    const int p = pl.po;
    const bool own = pl.go >= 0;
    if (quv) {
      const T ud = face<0>(sh, p, q);
      const T vd = face<1>(sh, p, q);
      sh.ud[HC_MP_CHECKC(p)] = ud;
      sh.vd[HC_MP_CHECKC(p)] = vd;
      if (own && wuv && a.ud) {
        (a.ud + pq)[pl.go] = ud;
        (a.vd + pq)[pl.go] = vd;
      }
    }
    if (qw) {
      const T wd = face<2>(sh, p, q);
      sh.wd[q & 1][HC_MP_CHECKC(p)] = wd;
      if (own && ww && a.wd) (a.wd + pq)[pl.go] = wd;
    }
    if (pl.hp >= 0) {
      const int hp = pl.hp & ((1 << 28) - 1), kind = pl.hp >> 28;
      if (kind == 0) {
        if (quv) {
          const T v = face<0>(sh, hp, q);
          sh.ud[HC_MP_CHECKC(hp)] = v;
          if (a.ud && wuv && pl.ho >= 0) (a.ud + pq)[pl.ho] = v;
        }
      } else if (kind == 1) {
        if (quv) {
          const T v = face<1>(sh, hp, q);
          sh.vd[HC_MP_CHECKC(hp)] = v;
          if (a.vd && wuv && pl.ho >= 0) (a.vd + pq)[pl.ho] = v;
        }
      } else if (qw) {
        sh.wd[q & 1][HC_MP_CHECKC(hp)] = face<2>(sh, hp, q);
      }
    }
  }

  // the unlimited donor-cell flag of the cell at LDS index p of plane q
  HC_HD static int flag_at(const Sh &sh, const MpPass<T> &a, int p, int q) {
    const T *c0 = sh.c[q & 3], *cm = sh.c[(q - 1) & 3], *cq = sh.c[(q + 1) & 3];
    const T *wm = sh.wd[(q - 1) & 1], *w0 = sh.wd[q & 1];
    const T cc = c0[p];
    const T x = cc + mp_donor<T>(cc, c0[HC_MP_CHECK(p - 1)], c0[HC_MP_CHECK(p + 1)], c0[HC_MP_CHECK(p - SI)],
                                 c0[HC_MP_CHECK(p + SI)], cm[p], cq[p], sh.ud[HC_MP_CHECKC(p - 1)], sh.ud[p],
                                 sh.vd[HC_MP_CHECKC(p - SI)], sh.vd[p], wm[p], w0[p], false, 0, 0, 0, 0, 0, 0, 0);
    return x <= (T)a.cmin ? 1 : (x >= (T)a.cmax ? 2 : 0);
  }

  // phase B: the own cell's flag of plane q, and one halo flag on the last
  // threads
  HC_HD static void flags(const MpGeom &g, const Plan &pl, Sh &sh, const MpPass<T> &a, int q) {
    const bool qin = q >= 1 && q <= g.nz;
    unsigned char *fl = sh.fl[q & 1];
    const int f = flag_at(sh, a, pl.po, q);  // formed everywhere, kept where the cell exists
    fl[HC_MP_CHECKC(pl.po)] = (unsigned char)((qin && pl.go >= 0) ? f : 0);
    if (a.fl && q >= g.ka && q <= g.kb && pl.go >= 0) (a.fl + plane(g, q))[HC_MP_GIDX(pl.go)] = (unsigned char)f;
    if (pl.lp >= 0) {
      const int p = pl.lp & (INTERIOR - 1);
      fl[HC_MP_CHECKC(p)] = (unsigned char)((qin && (pl.lp & INTERIOR)) ? flag_at(sh, a, p, q) : 0);
    }
  }

  HC_HD static void finish(const MpGeom &g, const Plan &pl, Sh &sh, const MpPass<T> &a, int q, MpRegs<T> &r) {
    T1::finish_cell(g, pl.po, pl.go, pl.ci, pl.cj, sh, a, q, r);
  }
};

// One block of the pass: the phase sequence of the device kernel, for a
// block executor X that runs a phase for every thread of the block and
// separates phases as a barrier does:
//   X::template each(f) calls f(tid, ps) for each thread with its state ps;
//   X::sync() is the barrier.
template <typename T, int TI, int TJ, class X>
HC_HD inline void mp2_block(X &x, const MpGeom &g, typename MpTile2<T, TI, TJ>::Sh &sh, const MpPass<T> &a) {
  typedef MpTile2<T, TI, TJ> Tl;
  const int qs = g.ka - 2 > 0 ? g.ka - 2 : 0;
  x.each([&](int tid, auto &ps) {
    Tl::plan(g, tid, ps.pl);
    ps.r = MpRegs<T>{};
    Tl::T1::stage_initial(g, tid, ps.pl.so, sh, a, qs);
    Tl::T1::fetch(g, ps.pl.so, a, qs, ps.pf);
  });
  for (int q = qs; q <= g.kb + 1; ++q) {
    // no barrier before the staging: it overwrites the slots of c(q-3),
    // U(q-1), V(q-1), W(q-2), which the previous plane's finishing step does
    // not read and its faces and flags phases finished reading before the
    // barrier ahead of that step
    x.each([&](int tid, auto &ps) { Tl::T1::commit(tid, sh, q, ps.pf); });
    x.sync();
    x.each([&](int tid, auto &ps) {
      if (q <= g.kb) Tl::T1::fetch(g, ps.pl.so, a, q + 1, ps.pf);  // in flight during this plane
      Tl::faces(g, ps.pl, sh, a, q, q >= g.ka - 1);
    });
    x.sync();
    x.each([&](int tid, auto &ps) {
      if (q >= g.ka - 1) Tl::flags(g, ps.pl, sh, a, q);
    });
    x.sync();
    x.each([&](int tid, auto &ps) { Tl::finish(g, ps.pl, sh, a, q, ps.r); });
  }
}

template <typename T, int TI, int TJ>
struct Mp2State {
  typename MpTile2<T, TI, TJ>::Plan pl;
  typename MpTile<T, TI, TJ>::Prefetch pf;
  MpRegs<T> r;
};

}  // namespace hc
