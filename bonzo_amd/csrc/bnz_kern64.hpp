// bnz_kern64.hpp -- per-face, per-edge and per-cell bodies of the FP64 VL
// step, shared by the single-grid kernels (bnz_gpu.hip) and the batched
// kernels (bnz_batch.hip), so that a box of a batch follows exactly the
// arithmetic of a single grid.
#pragma once
#include <type_traits>

#include "bnz_common.hpp"

namespace bnz {

// Godunov fluxes of one direction (godunov.pyx:20-210 with the double-buffered
// left state of D1).  Face index c is the lower face of cell c, between cells
// c-1 and c.  The flux is stored in global variable order, flux[n][c], except
// the normal-field slot, which no consumer reads.  flux_face_to hands the
// flux (local slots) of a face with a matched region to sink(f, c);
// flux_face stores it.
// CP (PLM only): the slopes are limited in characteristic variables of each
// cell's own state (charact_proj, reconstr_c.c:64-80), which reads the
// cell-centred normal field too.
template <int DIR, int ORDER, int RS, class SINK, bool CP = false>
__device__ __forceinline__ void flux_face_to(const Geo &g, const double *__restrict__ w,
                                             const double *__restrict__ bf, int i, int j, int k,
                                             const SINK &sink) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sd = (DIR == 0) ? 1 : (DIR == 1) ? g.sy : g.sz;
  const double bn = bf[(long)DIR * g.sv + c];

  double wl[NMODE], wr[NMODE];
  if (ORDER == 1 && CP) {
    double wm2[NMODE], wm1[NMODE], w0[NMODE], wp1[NMODE];
#pragma unroll
    for (int n = 0; n < NMODE; ++n) {
      const double *p = w + (long)perm<DIR>(n) * g.sv + c;
      wm2[n] = p[-2 * sd];
      wm1[n] = p[-sd];
      w0[n] = p[0];
      wp1[n] = p[sd];
    }
    double d[NMODE];
    plm_half_slope<true>(d, wm2, wm1, w0, g.gam);  // cell c-1, upper state
#pragma unroll
    for (int n = 0; n < NMODE; ++n) wl[n] = wm1[n] + d[n];
    plm_half_slope<true>(d, wm1, w0, wp1, g.gam);  // cell c, lower state
#pragma unroll
    for (int n = 0; n < NMODE; ++n) wr[n] = w0[n] - d[n];
  } else {
#pragma unroll
    for (int n = 0; n < NMODE; ++n) {
      if (n == BX) continue;
      const double *p = w + (long)perm<DIR>(n) * g.sv + c;
      if (ORDER == 0) {
        wl[n] = p[-sd];
        wr[n] = p[0];
      } else {
        const double wm2 = p[-2 * sd], wm1 = p[-sd], w0 = p[0], wp1 = p[sd];
        // cell c-1, upper state; cell c, lower state (reconstr_c.c:54-89)
        const double dl = vl_half(wm1 - wm2, w0 - wm1);
        const double dr = vl_half(w0 - wm1, wp1 - w0);
        wl[n] = wm1 + dl;
        wr[n] = w0 - dr;
      }
    }
  }
  wl[BX] = bn;
  wr[BX] = bn;

  double f[NMODE];
  int reg;
  if (RS == BNZ_RS_HLLD)
    reg = flux_hlld(f, wl, wr, g.gam, g.gamm1i);
  else if (RS == BNZ_RS_HLLT)
    reg = flux_hllt(f, wl, wr, g.gam, g.gamm1i);
  else if (RS == BNZ_RS_HLLC)
    reg = flux_hllc(f, wl, wr, g.gam, g.gamm1i);
  else if (RS == RS_HLL_HYDRO)
    reg = flux_hll<double, true>(f, wl, wr, g.gam, g.gamm1i);
  else
    reg = flux_hll(f, wl, wr, g.gam, g.gamm1i);
  if (reg == 0) return;  // reference leaves the flux row untouched
  sink(f, c);
}

template <int DIR, int ORDER, int RS, bool CP = false>
__device__ __forceinline__ void flux_face(const Geo &g, const double *__restrict__ w,
                                          const double *__restrict__ bf,
                                          double *__restrict__ fl, int i, int j, int k) {
  const long sv = g.sv;
  const auto store = [fl, sv](const double *f, long c) {
    // streaming stores: the fluxes are read back by the next kernels, long
    // after any L2 could hold them, and bypassing it keeps the stencil rows
    // of the prim there (-4 % on the flux launches, -3 % on the EMFs)
#pragma unroll
    for (int n = 0; n < NMODE; ++n) {
      if (n == BX) continue;
      __builtin_nontemporal_store(f[n], &fl[(long)perm<DIR>(n) * sv + c]);
    }
  };
  flux_face_to<DIR, ORDER, RS, decltype(store), CP>(g, w, bf, i, j, k, store);
}

// Every face flux of one stage whose lower cell is (i, j, k): the x, y and
// z faces of godunov_fluxes' three sweeps (godunov.pyx:20-210), each only
// where that sweep's range includes it.  Over flux3_box(L) this covers
// exactly the faces of the three per-direction launches.  DIMS > 0 fixes
// the dimensionality at compile time; 0 reads it from g.
template <int ORDER, int RS, int DIMS = 0, bool CP = false>
__device__ __forceinline__ void flux_cell(const Geo &g, const double *__restrict__ w,
                                          const double *__restrict__ bf,
                                          double *__restrict__ fx, double *__restrict__ fy,
                                          double *__restrict__ fz, const Lims &L, int i, int j,
                                          int k) {
  const int dims = DIMS > 0 ? DIMS : g.dims;
  if (i >= L.lo[0]) flux_face<0, ORDER, RS, CP>(g, w, bf, fx, i, j, k);
  if (dims >= 2 && j >= L.lo[1]) flux_face<1, ORDER, RS, CP>(g, w, bf, fy, i, j, k);
  if (dims >= 3 && k >= L.lo[2]) flux_face<2, ORDER, RS, CP>(g, w, bf, fz, i, j, k);
}

__device__ __forceinline__ double ec_comp(const double *w, long sv, long c, int comp) {
  const double vx = w[VX * sv + c], vy = w[VY * sv + c], vz = w[VZ * sv + c];
  const double bx = w[BX * sv + c], by = w[BY * sv + c], bz = w[BZ * sv + c];
  if (comp == 0) return -(vy * bz - vz * by);
  if (comp == 1) return (vx * bz - vz * bx);
  return -(vx * by - vy * bx);
}

// Edge EMFs, interp_ee1 (src/integration/ct.pyx:153-215), for a grid of
// DIMS dimensions: the three edges whose lower corner is cell (i, j, k).
template <int DIMS>
__device__ __forceinline__ void emf_edge_d(const Geo &g, const double *__restrict__ w,
                                           const double *__restrict__ fx,
                                           const double *__restrict__ fy,
                                           const double *__restrict__ fz,
                                           double *__restrict__ ee, int i, int j, int k) {
  const long sv = g.sv;
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sy = g.sy, sz = g.sz;
  if (DIMS == 3) {
    ee[0 * sv + c] = (0.5 * (-fy[BZ * sv + c - sz] - fy[BZ * sv + c] + fz[BY * sv + c - sy] +
                             fz[BY * sv + c]) -
                      0.25 * (ec_comp(w, sv, c - sy, 0) + ec_comp(w, sv, c, 0) +
                              ec_comp(w, sv, c - sz - sy, 0) + ec_comp(w, sv, c - sz, 0)));
    ee[1 * sv + c] = (0.5 * (-fz[BX * sv + c - 1] - fz[BX * sv + c] + fx[BZ * sv + c - sz] +
                             fx[BZ * sv + c]) -
                      0.25 * (ec_comp(w, sv, c - 1, 1) + ec_comp(w, sv, c, 1) +
                              ec_comp(w, sv, c - sz - 1, 1) + ec_comp(w, sv, c - sz, 1)));
    ee[2 * sv + c] = (0.5 * (-fx[BY * sv + c - sy] - fx[BY * sv + c] + fy[BX * sv + c - 1] +
                             fy[BX * sv + c]) -
                      0.25 * (ec_comp(w, sv, c - 1, 2) + ec_comp(w, sv, c, 2) +
                              ec_comp(w, sv, c - sy - 1, 2) + ec_comp(w, sv, c - sy, 2)));
  } else if (DIMS == 2) {
    ee[0 * sv + c] = -fy[BZ * sv + c];
    ee[1 * sv + c] = fx[BZ * sv + c];
    ee[2 * sv + c] = (0.5 * (-fx[BY * sv + c - sy] - fx[BY * sv + c] + fy[BX * sv + c - 1] +
                             fy[BX * sv + c]) -
                      0.25 * (ec_comp(w, sv, c - 1, 2) + ec_comp(w, sv, c, 2) +
                              ec_comp(w, sv, c - sy - 1, 2) + ec_comp(w, sv, c - sy, 2)));
  } else {
    ee[0 * sv + c] = -ec_comp(w, sv, c, 0);
    ee[1 * sv + c] = fx[BZ * sv + c];
    ee[2 * sv + c] = -fx[BY * sv + c];
  }
}

// Weight of the lower cell's EMF slope in the upwind corner EMF,
// get_ct_weight of src/integration/ct.pyx:221-226.
__device__ __forceinline__ double ct_weight(double frho, double rho_l, double rho_r, double ds,
                                            double dt) {
  const double v_c = 1e3 * frho * dt / (fmin(rho_l, rho_r) * ds);
  return 0.5 + fmax(-0.5, fmin(0.5, v_c));
}

// Upwind corner EMFs of cell corner (i, j, k), interp_ee2 of
// src/integration/ct.pyx:229-358: each EMF slope at an edge quarter is taken
// from the side the mass flux through the neighbouring face comes from
// (smoothed by ct_weight).  3D: the reference's three components with the
// two cell-EMF indices of its Ex slopes corrected to the cyclic image of its
// Ez and Ey (ec[0,k-1,j-1,i]); 2D: its Ez, with Ex = -Fy[Bz] and
// Ey = Fx[Bz].  dt is the step's dt.  Same operation order as the oracle.
template <int DIMS>
__device__ __forceinline__ void emf_edge_upwind(const Geo &g, const double *__restrict__ w,
                                                const double *__restrict__ fx,
                                                const double *__restrict__ fy,
                                                const double *__restrict__ fz,
                                                double *__restrict__ ee, double dt, int i, int j,
                                                int k) {
  const long sv = g.sv;
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  if (DIMS == 1) {  // interp_ee2's 1D branch (ct.pyx:303-304, 341-342, 282-283)
    ee[0 * sv + c] = ec_comp(w, sv, c, 0);
    ee[1 * sv + c] = fx[BZ * sv + c];
    ee[2 * sv + c] = -fx[BY * sv + c];
    return;
  }
  const long sy = g.sy, sz = DIMS == 3 ? g.sz : 0;
  const double *rho = w;  // slot RHO
  {  // Ez at (i-1/2, j-1/2)
    const double wgtx = ct_weight(fx[RHO * sv + c], rho[c - 1], rho[c], g.dlf[0][i], dt);
    const double dedy_14 = (wgtx * (ec_comp(w, sv, c - 1, 2) - fy[BX * sv + c - 1]) +
                            (1. - wgtx) * (ec_comp(w, sv, c, 2) - fy[BX * sv + c]));
    const double dedy_34 = (wgtx * (fy[BX * sv + c - 1] - ec_comp(w, sv, c - sy - 1, 2)) +
                            (1. - wgtx) * (fy[BX * sv + c] - ec_comp(w, sv, c - sy, 2)));
    const double wgty = ct_weight(fy[RHO * sv + c], rho[c - sy], rho[c], g.dlv[1][j], dt);
    const double dedx_14 = (wgty * (ec_comp(w, sv, c - sy, 2) + fx[BY * sv + c - sy]) +
                            (1. - wgty) * (ec_comp(w, sv, c, 2) + fx[BY * sv + c]));
    const double dedx_34 = (wgty * (-fx[BY * sv + c - sy] - ec_comp(w, sv, c - sy - 1, 2)) +
                            (1. - wgty) * (-fx[BY * sv + c] - ec_comp(w, sv, c - 1, 2)));
    ee[2 * sv + c] = 0.25 * (-fx[BY * sv + c - sy] - fx[BY * sv + c] + fy[BX * sv + c - 1] +
                             fy[BX * sv + c] + dedy_34 - dedy_14 + dedx_34 - dedx_14);
  }
  if (DIMS == 2) {
    ee[0 * sv + c] = -fy[BZ * sv + c];
    ee[1 * sv + c] = fx[BZ * sv + c];
    return;
  }
  {  // Ex at (j-1/2, k-1/2)
    const double wgty = ct_weight(fy[RHO * sv + c], rho[c - sy], rho[c], g.dlf[1][j], dt);
    const double dedz_14 = (wgty * (ec_comp(w, sv, c - sy, 0) - fz[BY * sv + c - sy]) +
                            (1. - wgty) * (ec_comp(w, sv, c, 0) - fz[BY * sv + c]));
    const double dedz_34 = (wgty * (fz[BY * sv + c - sy] - ec_comp(w, sv, c - sz - sy, 0)) +
                            (1. - wgty) * (fz[BY * sv + c] - ec_comp(w, sv, c - sz, 0)));
    const double wgtz = ct_weight(fz[RHO * sv + c], rho[c - sz], rho[c], g.dlv[2][k], dt);
    const double dedy_14 = (wgtz * (ec_comp(w, sv, c - sz, 0) + fy[BZ * sv + c - sz]) +
                            (1. - wgtz) * (ec_comp(w, sv, c, 0) + fy[BZ * sv + c]));
    const double dedy_34 = (wgtz * (-fy[BZ * sv + c - sz] - ec_comp(w, sv, c - sz - sy, 0)) +
                            (1. - wgtz) * (-fy[BZ * sv + c] - ec_comp(w, sv, c - sy, 0)));
    ee[0 * sv + c] = 0.25 * (-fy[BZ * sv + c - sz] - fy[BZ * sv + c] + fz[BY * sv + c - sy] +
                             fz[BY * sv + c] + dedz_34 - dedz_14 + dedy_34 - dedy_14);
  }
  {  // Ey at (i-1/2, k-1/2)
    const double wgtz = ct_weight(fz[RHO * sv + c], rho[c - sz], rho[c], g.dlf[2][k], dt);
    const double dedx_14 = (wgtz * (ec_comp(w, sv, c - sz, 1) - fx[BZ * sv + c - sz]) +
                            (1. - wgtz) * (ec_comp(w, sv, c, 1) - fx[BZ * sv + c]));
    const double dedx_34 = (wgtz * (fx[BZ * sv + c - sz] - ec_comp(w, sv, c - sz - 1, 1)) +
                            (1. - wgtz) * (fx[BZ * sv + c] - ec_comp(w, sv, c - 1, 1)));
    const double wgtx = ct_weight(fx[RHO * sv + c], rho[c - 1], rho[c], g.dlv[0][i], dt);
    const double dedz_14 = (wgtx * (ec_comp(w, sv, c - 1, 1) + fz[BX * sv + c - 1]) +
                            (1. - wgtx) * (ec_comp(w, sv, c, 1) + fz[BX * sv + c]));
    const double dedz_34 = (wgtx * (-fz[BX * sv + c - 1] - ec_comp(w, sv, c - sz - 1, 1)) +
                            (1. - wgtx) * (-fz[BX * sv + c] - ec_comp(w, sv, c - sz, 1)));
    ee[1 * sv + c] = 0.25 * (-fz[BX * sv + c - 1] - fz[BX * sv + c] + fx[BZ * sv + c - sz] +
                             fx[BZ * sv + c] + dedx_34 - dedx_14 + dedz_34 - dedz_14);
  }
}

// --------------------------------------------------------------------------
// Stage update: advance_hydro (godunov.pyx:215-252), advance_b (ct.pyx:15-91),
// interp_bc (ct.pyx:96-123) and cons2prim (eos.pyx:13-41) fused, one thread
// per cell of the whole grid.  u0 = prim2cons(w0) is the reference's cons
// array.  PRED: writes ws (a separate buffer); cells outside the limits take
// w0, as the reference's ws aliases prim (vanleer.pyx:72).  Corrector: w1 may
// alias w0 (each thread reads only its own cell of w0).
//
// Every operand is loaded up front and unconditionally: neighbour offsets
// are zero where the cell needs no neighbour on that side, which keeps all
// loads in bounds, and only the stores depend on the cell's role.  The loads
// then form one block the compiler can issue together instead of a chain of
// dependent round trips.  Face values by Stokes (advance_b):
//   Bx(i)   = b - dt/(dy dz) (dz Ez(j+1) - dz Ez(j)) + dt/(dy dz) (dy Ey(k+1) - dy Ey(k))
//   By(j)   = b + dt/(dx dz) (dz Ez(i+1) - dz Ez(i)) - dt/(dx dz) (dx Ex(k+1) - dx Ex(k))
//   Bz(k)   = b - dt/(dx dy) (dy Ey(i+1) - dy Ey(i)) + dt/(dx dy) (dx Ex(j+1) - dx Ex(j))
//
// The edge EMFs come from em(comp, off): component comp at the edge whose
// lower corner is cell c + off.  EmfGlobal reads the ee array of the EMF
// kernel; on 1D grids EmfLds1D (below) forms them from the x fluxes as that
// kernel does (ey = F_x[Bz], ez = -F_x[By]; a 1D cell reads no Ex), so 1D
// steps need no EMF kernel.
struct EmfGlobal {
  const double *ee;
  long sv, c;
  __device__ __forceinline__ double operator()(int comp, long off) const {
    return ee[comp * sv + c + off];
  }
};

// The 2D edge EMFs formed where the update needs them (small 2D grids, no
// EMF launch): emf_edge_d<2>'s expressions at the edge whose lower corner is
// cell c + off, from the stage's prim we and flux arrays.  Each edge is
// formed by the (up to) four cells around it, the same value each time.
// Threads with no role in the update read at off = 0; an edge on the grid's
// first row or column has no cells below it and gives 0 (no role reads it).
struct EmfFly2D {
  const double *we, *fx, *fy;
  long sv, sy, c;
  int i, j;
  __device__ __forceinline__ double operator()(int comp, long off) const {
    const long e = c + off;
    if (comp == 0) return -fy[BZ * sv + e];
    if (comp == 1) return fx[BZ * sv + e];
    const int je = j + (off >= sy ? 1 : 0);
    const int ie = i + (int)(off >= sy ? off - sy : off);
    if (ie < 1 || je < 1) return 0.;
    const double *w = we;
    return (0.5 * (-fx[BY * sv + e - sy] - fx[BY * sv + e] + fy[BX * sv + e - 1] +
                   fy[BX * sv + e]) -
            0.25 * (ec_comp(w, sv, e - 1, 2) + ec_comp(w, sv, e, 2) +
                    ec_comp(w, sv, e - sy - 1, 2) + ec_comp(w, sv, e - sy, 2)));
  }
};

// Face fluxes as the update reads them: variable n of the face whose lower
// cell is c + off.  FluxGlobal reads the arrays of the flux kernels.
struct FluxGlobal {
  const double *fx, *fy, *fz;
  long sv, c;
  __device__ __forceinline__ double x(int n, long off) const { return fx[n * sv + c + off]; }
  __device__ __forceinline__ double y(int n, long off) const { return fy[n * sv + c + off]; }
  __device__ __forceinline__ double z(int n, long off) const { return fz[n * sv + c + off]; }
};

// Fluxes of a run of 1D faces held in LDS, s[n * BLK + slot]: slot t + off
// is the face whose lower cell is the updating thread's cell + off.
struct FluxLds1D {
  const double *s;
  int t;
  __device__ __forceinline__ double x(int n, long off) const { return s[n * BLK + t + off]; }
  __device__ __forceinline__ double y(int, long) const { return 0.; }
  __device__ __forceinline__ double z(int, long) const { return 0.; }
};
struct EmfLds1D {  // the 1D EMFs from the LDS fluxes
  const double *s;
  int t;
  __device__ __forceinline__ double operator()(int comp, long off) const {
    return comp == 2 ? -s[BY * BLK + t + off] : comp == 1 ? s[BZ * BLK + t + off] : 0.;
  }
};

// LDS_OUT (predictor of a fused 1D step): the new prim goes to the LDS
// array w1 as w1[n * BLK + lslot] and no face is stored (b1 unused).
template <int DIMS, bool PRED, bool DTRED, class EMF, class GM, class FL, bool LDS_OUT = false>
__device__ __forceinline__ double update_cell_e(const Geo &g, const GM &gm, const double *w0,
                                                const FL &fl, const EMF &em,
                                                const double *__restrict__ b0,
                                                double *__restrict__ b1, double *w1,
                                                const Lims &L, const double *__restrict__ dts,
                                                const Box &act, int fms_ref, int i, int j, int k,
                                                int lslot = 0) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sv = g.sv;
  const double dt = PRED ? 0.5 * dts[0] : dts[0];

  const bool inI = i >= L.lo[0] && i <= L.hi[0];
  const bool inJ = j >= L.lo[1] && j <= L.hi[1];
  const bool inK = k >= L.lo[2] && k <= L.hi[2];
  const bool cell = inI && inJ && inK;
  const bool fX = (i >= L.lo[0] && i <= L.hi[0] + 1) && inJ && inK;
  const bool fY = inI && (j >= L.lo[1] && j <= L.hi[1] + (DIMS >= 2 ? 1 : 0)) && inK;
  const bool fZ = inI && inJ && (k >= L.lo[2] && k <= L.hi[2] + (DIMS >= 3 ? 1 : 0));
  // neighbour offsets: each is used only by roles that keep that index
  // inside the limits, so c + ox + oy + oz is a cell of the grid
  const long ox = (cell || fY || fZ) ? 1 : 0;
  const long oy = (DIMS >= 2 && (cell || fX || fZ)) ? g.sy : 0;
  const long oz = (DIMS >= 3 && (cell || fX || fY)) ? g.sz : 0;

  // ---- loads
  double wv[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) wv[n] = w0[n * sv + c];
  const double bxc = b0[c], byc = b0[sv + c], bzc = b0[2 * sv + c];
  const double bxp = b0[c + ox], byp = b0[sv + c + oy], bzp = b0[2 * sv + c + oz];
  const double ez00 = em(2, 0), ez10 = em(2, ox), ez01 = em(2, oy), ez11 = em(2, ox + oy);
  const double ey00 = em(1, 0), ey10 = em(1, ox), ey01 = em(1, oz), ey11 = em(1, ox + oz);
  const double ex00 = em(0, 0), ex10 = em(0, oy), ex01 = em(0, oz), ex11 = em(0, oy + oz);
  double fxl[NMODE - 3], fxu[NMODE - 3], fyl[NMODE - 3], fyu[NMODE - 3], fzl[NMODE - 3],
      fzu[NMODE - 3];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    fxl[n] = fl.x(n, 0);
    fxu[n] = fl.x(n, ox);
    if (DIMS >= 2) {
      fyl[n] = fl.y(n, 0);
      fyu[n] = fl.y(n, oy);
    }
    if (DIMS >= 3) {
      fzl[n] = fl.z(n, 0);
      fzu[n] = fl.z(n, oz);
    }
  }
  const double dlx = gm.dlf(0, i), dly = gm.dlf(1, j), dlz = gm.dlf(2, k);

  // ---- faces (advance_b); the lower faces of the cell and, for the cell
  // interpolation, its upper faces
  double bxl, bxu, byl, byu = 0., bzl, bzu = 0.;
  {
    const double dtda = dt / (dly * dlz);
    double v = bxc, vu = bxp;
    if (DIMS >= 2) {
      const double ds = dlz;
      v -= dtda * (ds * ez01 - ds * ez00);
      vu -= dtda * (ds * ez11 - ds * ez10);
    }
    if (DIMS >= 3) {
      const double ds = dly;
      v += dtda * (ds * ey01 - ds * ey00);
      vu += dtda * (ds * ey11 - ds * ey10);
    }
    bxl = v;
    bxu = vu;
  }
  {
    const double dtda = dt / (dlx * dlz);
    double ds = dlz;
    double v = byc + dtda * (ds * ez10 - ds * ez00);
    double vu = byp + dtda * (ds * ez11 - ds * ez01);
    if (DIMS >= 3) {
      ds = dlx;
      v -= dtda * (ds * ex01 - ds * ex00);
      vu -= dtda * (ds * ex11 - ds * ex10);
    }
    byl = v;
    if (DIMS >= 2) byu = vu;
  }
  {
    const double dtda = dt / (dlx * dly);
    double ds = dly;
    double v = bzc - dtda * (ds * ey10 - ds * ey00);
    double vu = bzp - dtda * (ds * ey11 - ds * ey01);
    if (DIMS >= 2) {
      ds = dlx;
      v += dtda * (ds * ex10 - ds * ex00);
      vu += dtda * (ds * ex11 - ds * ex01);
    }
    bzl = v;
    if (DIMS >= 3) bzu = vu;
  }

  // ---- conserved update (advance_hydro) and cell-centred field (interp_bc)
  double u[NMODE];
  prim2cons(u, wv, g.gamm1i);
  const double dtdv = dt / (dlx * dly * dlz);
  const double ax = dly * dlz, ay = dlx * dlz, az = dlx * dly;
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    double v = u[n] - dtdv * (ax * fxu[n] - ax * fxl[n]);
    if (DIMS >= 2) v -= dtdv * (ay * fyu[n] - ay * fyl[n]);
    if (DIMS >= 3) v -= dtdv * (az * fzu[n] - az * fzl[n]);
    u[n] = v;
  }
  u[BX] = (gm.icm(0, i) * bxl + gm.icp(0, i) * bxu) * gm.dlfi(0, i);
  if (DIMS >= 2)
    u[BY] = (gm.icm(1, j) * byl + gm.icp(1, j) * byu) * gm.dlfi(1, j);
  else
    u[BY] = byl;
  if (DIMS >= 3)
    u[BZ] = (gm.icm(2, k) * bzl + gm.icp(2, k) * bzu) * gm.dlfi(2, k);
  else
    u[BZ] = bzl;
  double wn[NMODE];
  cons2prim(wn, u, g.gam);

  // ---- stores
  if (LDS_OUT) {
    if (cell) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n) w1[n * BLK + lslot] = wn[n];
    } else if (PRED) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n) w1[n * BLK + lslot] = wv[n];
    }
    return 0.;
  }
  if (fX) b1[c] = bxl;
  if (fY) b1[sv + c] = byl;
  if (fZ) b1[2 * sv + c] = bzl;
  if (cell) {
#pragma unroll
    for (int n = 0; n < NMODE; ++n) w1[n * sv + c] = wn[n];
  } else if (PRED) {
#pragma unroll
    for (int n = 0; n < NMODE; ++n) w1[n * sv + c] = wv[n];
  }
  if (DTRED && cell && i >= act.lo[0] && i < act.lo[0] + act.n[0] && j >= act.lo[1] &&
      j < act.lo[1] + act.n[1] && k >= act.lo[2] && k < act.lo[2] + act.n[2])
    return cds_cell_g(gm, g, wn, fms_ref, i, j, k);
  return 0.;
}

// FLY (2D): the edge EMFs formed from the stage's prim we (EmfFly2D) instead
// of read from ee.
template <int DIMS, bool PRED, bool DTRED, bool UG = false, bool FLY = false>
__device__ __forceinline__ double update_cell_d(const Geo &g, const double *w0,
                                                const double *__restrict__ fx,
                                                const double *__restrict__ fy,
                                                const double *__restrict__ fz,
                                                const double *__restrict__ ee,
                                                const double *__restrict__ b0,
                                                double *__restrict__ b1, double *w1,
                                                const Lims &L, const double *__restrict__ dts,
                                                const Box &act, int fms_ref, int i, int j, int k,
                                                const double *we = nullptr) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  using GM = typename std::conditional<UG, GeoUni, GeoArr>::type;
  const GM gm{g};
  static_assert(DIMS >= 2, "1D grids run stage1_cell");
  static_assert(!FLY || DIMS == 2, "EMFs in the update on 2D grids only");
  const FluxGlobal fl{fx, fy, fz, g.sv, c};
  if (FLY)
    return update_cell_e<DIMS, PRED, DTRED>(g, gm, w0, fl,
                                            EmfFly2D{we, fx, fy, g.sv, (long)g.sy, c, i, j}, b0,
                                            b1, w1, L, dts, act, fms_ref, i, j, k);
  return update_cell_e<DIMS, PRED, DTRED>(g, gm, w0, fl, EmfGlobal{ee, g.sv, c}, b0, b1, w1, L,
                                          dts, act, fms_ref, i, j, k);
}

// One stage of a 1D grid for thread threadIdx.x of a workgroup of BLK that
// owns faces i0 .. i0 + BLK - 1 and cells i0 .. i0 + BLK - 2 (i = i0 +
// threadIdx.x): the flux of face i (flux_face_to on the stage's prim wf)
// goes to sf (NMODE x BLK doubles of LDS), then after a barrier cell i is
// updated from faces i and i + 1 (update_cell_e), so the fluxes never go
// through HBM.  Every thread must call it.  The normal field of the faces
// is read from the step's faces b0: in 1D the predictor copies x faces
// unchanged, so they equal the predictor's, and the corrector's writes to
// b1 cannot race these reads.  A face whose solver matches no region
// (non-finite speeds) takes the old row of fx_old, which a fused pass never
// writes.  Returns the CFL speed of the updated cell (corrector).
template <bool PRED, int ORDER, int RS>
__device__ __forceinline__ double stage1_cell(const Geo &g, const double *__restrict__ wf,
                                              const double *b0, double *b1, const double *w0,
                                              double *w1, const double *__restrict__ fx_old,
                                              const Lims &L, const double *__restrict__ dts,
                                              const Box &act, int fms_ref, int i, double *sf) {
  const int tid = threadIdx.x;
  if (i >= L.lo[0] && i <= L.hi[0] + 1) {  // the stage's faces (flux3_box)
    bool got = false;
    flux_face_to<0, ORDER, RS>(g, wf, b0, i, 0, 0, [&](const double *f, long) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n)
        if (n != BX) sf[n * BLK + tid] = f[n];
      got = true;
    });
    if (!got)
      for (int n = 0; n < NMODE; ++n)
        if (n != BX) sf[n * BLK + tid] = fx_old[n * g.sv + i];
  }
  __syncthreads();
  double cds = 0.;
  if (tid < BLK - 1 && i < g.nt[0]) {
    const double r = update_cell_e<1, PRED, !PRED>(g, GeoArr{g}, w0, FluxLds1D{sf, tid},
                                                   EmfLds1D{sf, tid}, b0, b1, w1, L, dts, act,
                                                   fms_ref, i, 0, 0);
    if (!PRED) cds = r;
  }
  return cds;
}

// A whole VL step of a 1D grid for a workgroup of BLK threads that owns the
// STEP1_CELLS cells i0 .. i0 + STEP1_CELLS - 1.  Thread t computes the
// predictor face and cell i0 - 2 + t (donor cell, dt/2; fluxes in sf, the
// new prims in sw; the 2 cells below and 3 above the owned range are
// computed by the neighbouring workgroups too), then the corrector face
// i0 + t from the prims in sw and, for t < STEP1_CELLS, the corrector cell
// i0 + t.  The arithmetic is stage1_cell's.  It reads the step's prim w and
// faces b only and writes the new prim and faces to wout and bout, so no
// read races a write; the predictor's faces are not kept (in 1D the
// corrector needs only the x faces, which the predictor copies).  Every
// thread must call it; returns the CFL speed of the thread's corrector cell.
template <int ORDER, int RS, bool CP = false>
__device__ __forceinline__ double step1_block(const Geo &g, const double *w,
                                              const double *bf, double *wout, double *bout,
                                              const double *__restrict__ fx_old,
                                              const Lims &L0, const Lims &L1,
                                              const double *__restrict__ dts, const Box &act,
                                              int fms_ref, int i0, double *sf, double *sw) {
  constexpr int C = STEP1_CELLS;
  const int t = threadIdx.x;
  const long sv = g.sv;
  const auto sink = [&](const double *f, long) {
#pragma unroll
    for (int n = 0; n < NMODE; ++n)
      if (n != BX) sf[n * BLK + t] = f[n];
  };
  // predictor (donor cell, dt/2)
  const int ip = i0 - 2 + t;
  if (ip >= L0.lo[0] && ip <= L0.hi[0] + 1) {
    bool got = false;
    flux_face_to<0, 0, RS>(g, w, bf, ip, 0, 0, [&](const double *f, long q) {
      sink(f, q);
      got = true;
    });
    if (!got)  // no region matched: the flux row keeps what it held (flux_face)
      for (int n = 0; n < NMODE; ++n)
        if (n != BX) sf[n * BLK + t] = fx_old[n * sv + ip];
  }
  __syncthreads();
  if (t < BLK - 1 && ip >= 0 && ip < g.nt[0])
    update_cell_e<1, true, false, EmfLds1D, GeoArr, FluxLds1D, true>(
        g, GeoArr{g}, w, FluxLds1D{sf, t}, EmfLds1D{sf, t}, bf, nullptr, sw, L0, dts, act,
        fms_ref, ip, 0, 0, t);
  __syncthreads();
  // corrector (reconstruction order, dt): fluxes from the prims in sw, read
  // through a geometry whose variable stride is the LDS row (BLK) and whose
  // cell ip sits at slot ip - (i0 - 2)
  const int ic = i0 + t;
  if (t <= C && ic >= L1.lo[0] && ic <= L1.hi[0] + 1) {
    Geo gl = g;
    gl.sv = BLK;
    bool got = false;
    const auto sink_c = [&](const double *f, long q) {
      sink(f, q);
      got = true;
    };
    flux_face_to<0, ORDER, RS, decltype(sink_c), CP>(gl, sw - (i0 - 2), bf, ic, 0, 0, sink_c);
    if (!got)
      for (int n = 0; n < NMODE; ++n)
        if (n != BX) sf[n * BLK + t] = fx_old[n * sv + ic];
  }
  __syncthreads();
  double cds = 0.;
  if (t < C && ic < g.nt[0])
    cds = update_cell_e<1, false, true>(g, GeoArr{g}, w, FluxLds1D{sf, t}, EmfLds1D{sf, t}, bf,
                                        bout, wout, L1, dts, act, fms_ref, ic, 0, 0);
  return cds;
}

}  // namespace bnz
