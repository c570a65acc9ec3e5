// mpcheck.hip - host build of the MPDATA corrective-pass tiles (hc_mpdata.h:
// MpTile, the kernel k_mpdata_pass runs; hc_mpdata2.h: MpTile2; hc_mpdata3.h:
// MpTile3), for CPU
// tests: every block of a launch runs serially, each phase over all the
// block's threads before the next (what the barriers guarantee on the
// device), with every LDS index range-checked and LDS filled with NaN at the
// start of each block, so that a read of a slot no phase wrote shows up in
// the results; every staged load's element index is range-checked against
// the padded field too.  Host code only; nothing here launches a kernel.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <memory>
#include <vector>

static int g_lim_s = 0, g_lim_c = 0;  // staged / computed LDS plane sizes
static int g_lim_z = 0;                // halo z-flux entries (MpTile3)
static long g_lim_n = 0;               // elements of a padded field
static long g_bad = 0;                 // LDS or field indexes out of range
static inline int chk(int x, int lim) {
  if (x < 0 || x >= lim) {
    ++g_bad;
    return 0;
  }
  return x;
}
#define HC_MP_CHECK(x) chk((x), g_lim_s)
#define HC_MP_CHECKC(x) chk((x), g_lim_c)
#define HC_MP_CHECKZ(x) chk((x), g_lim_z)
static inline long chkg(long x) {
  if (x < 0 || x >= g_lim_n) {
    ++g_bad;
    return 0;
  }
  return x;
}
#define HC_MP_GIDX(x) chkg(x)

#include "../../plumemodel-hotcross_amd/csrc/hc_mpdata3.h"

using namespace hc;

namespace {

template <class State>
struct HostExec {
  std::vector<State> st;
  explicit HostExec(int nt) : st(nt) {}
  template <class F>
  void each(F f) {
    for (int t = 0; t < (int)st.size(); ++t) f(t, st[t]);
  }
  void sync() {}
};

template <class Sh>
void fill_nan(Sh &sh) {
  memset(&sh, 0xff, sizeof(sh));  // NaN in every T slot, 255 (no flag) in the flag planes
}

// the loop of k_mpdata_pass (hc_tracer.hip) for one block, on the host
template <typename T, int TI, int TJ>
void v1_block(const MpGeom &g, MpShared<T, TI, TJ> &sh, const MpPass<T> &a, int rot) {
  typedef MpTile<T, TI, TJ> Tl;
  struct St {
    typename Tl::Plan pl;
    typename Tl::Prefetch pf;
    MpRegs<T> r;
  };
  HostExec<St> x(TI * TJ);
  const int qs = g.ka - 2 > 0 ? g.ka - 2 : 0;
  x.each([&](int tid, St &s) {
    Tl::plan(g, tid, rot, s.pl);
    s.r = MpRegs<T>{};
    Tl::stage_initial(g, tid, s.pl.so, sh, a, qs);
    Tl::fetch(g, s.pl.so, a, qs, s.pf);
  });
  for (int q = qs; q <= g.kb + 1; ++q) {
    x.each([&](int tid, St &s) { Tl::commit(tid, sh, q, s.pf); });
    x.each([&](int tid, St &s) {
      if (q <= g.kb) Tl::fetch(g, s.pl.so, a, q + 1, s.pf);
      Tl::faces(g, s.pl, sh, a, q, q >= g.ka - 1);
    });
    x.each([&](int tid, St &s) {
      if (q >= g.ka - 1) Tl::flags(g, s.pl, sh, a, q);
    });
    x.each([&](int tid, St &s) { Tl::finish(g, s.pl, sh, a, q, s.r); });
  }
}

template <typename T, int TI, int TJ>
long run(int variant, int nx, int ny, int nz, int kc, const MpPass<T> &a) {
  Layout L = make_layout(nx, ny, nz);
  typedef MpShared<T, TI, TJ> Sh;
  g_lim_s = Sh::S;
  g_lim_c = Sh::SC;
  g_lim_n = (long)L.n;
  g_bad = 0;
  std::unique_ptr<Sh> sh(new Sh);
  typedef Mp3Shared<T, TI, TJ> Sh3;
  static_assert(Sh3::S == Sh::S && Sh3::SC == Sh::SC, "the same plane sizes");
  std::unique_ptr<Sh3> sh3(new Sh3);
  g_lim_z = Sh3::NHL;
  const int bxn = (nx + TI - 1) / TI, byn = (ny + TJ - 1) / TJ, bzn = (nz + kc - 1) / kc;
  for (int bz = 0; bz < bzn; ++bz)
    for (int by = 0; by < byn; ++by)
      for (int bx = 0; bx < bxn; ++bx) {
        MpGeom g;
        g.nx = nx; g.ny = ny; g.nz = nz; g.px = L.px; g.py = L.py;
        g.i0 = 1 + bx * TI;
        g.j0 = 1 + by * TJ;
        g.ka = 1 + bz * kc;
        g.kb = g.ka + kc - 1 < nz ? g.ka + kc - 1 : nz;
        fill_nan(*sh);
        fill_nan(*sh3);
        if (variant == 1) {
          v1_block<T, TI, TJ>(g, *sh, a, bx + by + bz);
        } else if constexpr (TJ >= 16) {  // MpTile2 and MpTile3 need one round of halo faces
          if (variant == 2) {
            HostExec<Mp2State<T, TI, TJ>> x(TI * TJ);
            mp2_block<T, TI, TJ>(x, g, *sh, a);
          } else {
            HostExec<Mp3State<T, TI, TJ>> x(TI * TJ);
            mp3_block<T, TI, TJ>(x, g, *sh3, a);
          }
        }
      }
  return g_bad;
}

}  // namespace

extern "C" {

int64_t mpc_layout(int nx, int ny, int nz) { return (int64_t)make_layout(nx, ny, nz).n; }

// One corrective pass over padded FP64 fields (the library's layout).
// variant 1: the MpTile loop of k_mpdata_pass; 2: mp2_block; 3: mp3_block.  tile 0: 32x16,
// 1: 16x16, 2: 32x8 (variant 1 only).  ud/vd/wd/cpre may be null.  dbl
// holds cmin, cmax, tau3, dxdydz; ints holds last, map_valid, nxs, nxe, nys,
// nye.  Returns the number of LDS indexes out of range.
int64_t mpc_pass(int variant, int tile, int nx, int ny, int nz, int kc, const double *cc, const double *U,
                 const double *V, const double *W, double *ud, double *vd, double *wd, double *cout, double *cpre,
                 const double *dbl, const int *ints, const double *cbkg, const double *alphae0,
                 const double *alphaw, const double *qpatch) {
  MpPass<double> a{};
  a.cc = cc; a.U = U; a.V = V; a.W = W;
  a.ud = ud; a.vd = vd; a.wd = wd;
  a.cout = cout; a.cpre = cpre;
  a.cmin = dbl[0]; a.cmax = dbl[1]; a.tau3 = dbl[2]; a.dxdydz = dbl[3];
  a.last = ints[0]; a.map_valid = ints[1];
  a.nxs = ints[2]; a.nxe = ints[3]; a.nys = ints[4]; a.nye = ints[5];
  a.cbkg = cbkg; a.alphae0 = alphae0; a.alphaw = alphaw; a.qpatch = qpatch;
  if (tile == 0) return run<double, 32, 16>(variant, nx, ny, nz, kc, a);
  if (tile == 1) return run<double, 16, 16>(variant, nx, ny, nz, kc, a);
  if (variant == 1) return run<double, 32, 8>(variant, nx, ny, nz, kc, a);
  return -1;
}

}  // extern "C"
