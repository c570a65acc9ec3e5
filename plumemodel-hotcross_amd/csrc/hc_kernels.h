// hc_kernels.h - kernel launcher declarations of the HotCross timestep.
#pragma once
#include "../../include/hotcross.h"
#include "hc_common.h"

namespace hc {

typedef hc_params Phys;

// XCD-aware tile order for the LDS-tiled kernels, on unless HC_TILE_XCD=0
int tile_remap();

// vent / source patch state used by MOMENTUM_W and the tracer kernels
struct VentArgs {
  int map_valid;            // isource_map set by a SOURCE_terms call
  int nxs, nxe, nys, nye;   // patch bounds, source80.inc:10-27
  double wup, sumup, sumdown;
};

template <typename T>
struct FieldPtrs {
  T *u, *v, *w, *up, *vp, *wp, *um, *vm, *wm, *p, *t, *s, *c, *ss, *ssv, *rhop, *fx, *fy, *fz;
};

// 1-D / 2-D case arrays, FP64, Fortran order with the reference's bounds
struct Aux {
  double *azzmin, *rho_bkg, *pgradx, *pgrady;  // (0:nz+1)
  double *alphau;   // (-1:nx+1, 0:ny+1)
  double *alphav;   // (0:nx+1, -1:ny+1)
  double *alphae;   // (0:nx+1, 0:ny+1) after the periodic copies
  double *alphae0;  // (0:nx+1, 0:ny+1) as used to form alphaee
  double *alphaw;   // (-1:nz+1)
};

// partial-sum slots per scalar sensitivity of the adjoint (a power of two)
constexpr int ADJ_SLOTS = 256;
// CFL_VALUES: partial results per block (six extremes and a non-finite count)
// and the number of blocks of its reduction
constexpr int CFL_NPART = 7;
constexpr int CFL_BLOCKS = 1024;

// d(ubkg_m, ubkg, vbkg_m, vbkg, fbx, fby)/d(flow_scale) at one step
struct FlowSens {
  double dubkg_m, dubkg, dvbkg_m, dvbkg, dfbx, dfby;
};

template <typename T>
void launch_forces(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                   const Aux &aux);
// Fused epilogue of MOMENTUM in the native loop: the work velocity's Courant
// number and REFRESH's filtered value, formed from the new up of each cell
template <typename T>
struct MomFuse {
  T *cour;      // uc / vc / wc, or null when the work velocities are not formed
  T *umn;       // filtered field, the next um
  double tau3, assel;
  int leap;
};
template <typename T>
void launch_momentum(hipStream_t st, int comp, const Layout &L, const StepArgs &s, const Phys &ph,
                     const VentArgs &va, const FieldPtrs<T> &f, const Aux &aux, const MomFuse<T> *mf = nullptr);
template <typename T>
void launch_viscosity(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                      const Aux &aux);
template <typename T>
void launch_density(hipStream_t st, const Layout &L, const Phys &ph, const FieldPtrs<T> &f, const Aux &aux);
// output diagnostics after VISCOSITY3D (hc_diag.hip): out = Ri, epsh, epsv,
// chi_v, chi_h, scalar-shaped fields of type T
template <typename T>
void launch_diag(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                 const Aux &aux, const double *t_bkg, void *const out[5]);
template <typename T>
void launch_refresh(hipStream_t st, const Layout &L, int leap, double assel, const T *u, const T *up, T *um);

// tracer transport (TRACER31 / MPDATA)
struct TracerWork {
  void *uc, *vc, *wc;      // Courant numbers of the work velocities
  void *ud2, *vd2, *wd2;   // antidiffusive velocities of pass 2 (pass 3's Courant numbers)
  void *c1, *c2;           // pass results (c1: pass 1, then the pre-sponge boundary values)
  void *cpre;              // if set, receives the pre-sponge boundary values instead of c1
  void *ud3, *vd3, *wd3;   // if set, receive pass 3's antidiffusive velocities
  uint8_t *fl2, *fl3;      // if set, receive the limiter flags of passes 2 and 3 (interior cells)
  double *budget_part;     // per-block partial sums (FP64)
  double *budget;          // 3 sums
};
struct TracerArgs {
  double cmin, cmax, tau3;
  int vent;          // apply the driver's vent ghost override after the call
  double vent_value;
  int budgets;       // mod(mtime,100) == 0
  int pass1_done;    // tw.c1 already holds pass 1's result with its boundary images
};
template <typename T>
void launch_courant(hipStream_t st, const Layout &L, const StepArgs &s, const FieldPtrs<T> &f, TracerWork &tw);
// optional hook around each launch of the corrective-pass kernel (for kernel
// timing with events on the stream): called with 0 before and 1 after
typedef void (*MarkFn)(void *arg, int end);
template <typename T>
void launch_tracer(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const VentArgs &va,
                   const FieldPtrs<T> &f, const Aux &aux, TracerWork &tw, T *c, const double *c_bkg,
                   const double *qpatch, const TracerArgs &ta, MarkFn mark = nullptr, void *mark_arg = nullptr);
// pass 1 of n <= 3 tracers c[q] (backgrounds c_bkg[q]) into c1[q], boundary
// images included, in one launch that stages the shared inputs once
template <typename T>
void launch_pass1_multi(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const VentArgs &va,
                        const FieldPtrs<T> &f, const Aux &aux, const TracerWork &tw, int n, T *const *c,
                        const double *const *c_bkg, T *const *c1, double tau3);
template <typename T>
void launch_vent_ghost(hipStream_t st, const Layout &L, const VentArgs &va, T *c, double value);
template <typename T>
void launch_cfl(hipStream_t st, const Layout &L, const FieldPtrs<T> &f, double *part, double *out);

// pressure solve (PRESSURE + HS3CRT)
struct PoissonWork {
  void *r;         // compact (nx,ny,nz) right-hand side / transform work
  void *rh;        // second compact buffer
  double *piv;     // pivots of the z tridiagonal systems, (nx,ny,nz)
  void *pivt;      // the same in the context's precision (piv itself in FP64); the
                   // singular (0,0) column reads piv in either form
  double *lam;     // lambda_x(kx) + lambda_y(ky), (nx,ny)
  const double *tw_x, *tw_y;  // FFT twiddles exp(-2 pi i m/N), complex (2N)
  double *part;    // per-block partial sums of the right-hand side
  double *delta;   // its total
  int nparts;
};
template <typename T>
void launch_pressure(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                     const Aux &aux, PoissonWork &pw);
void poisson_setup(const Layout &L, const Phys &ph, PoissonWork &pw, hipStream_t st);
// adjoint of PRESSURE: consumes lam.p (the adjoint of the pressure field,
// ghosts included) and adds to lam.um, vm, wm, w, fx, fy, fz; gflow (one
// FP64 device slots, ADJ_SLOTS of them) collects d/d(flow_scale) through the
// boundary data.
// g is a compact (nx,ny,nz) work array; zero points to a device 0.0.
template <typename T>
void launch_pressure_adj(hipStream_t st, const Layout &L, const StepArgs &s, const Phys &ph, const FieldPtrs<T> &f,
                         const FieldPtrs<T> &lam, const Aux &aux, PoissonWork &pw, const FlowSens &fs,
                         double *gflow, T *g, const double *zero, double *part);

}  // namespace hc
