// bnz_rk3.hpp -- the RK3 integrator (src/integration/_rk3.pyx:20-175) on one
// FP64 grid, included by bnz_step64.hpp (both FP64 units).
//
// Shu & Osher's three-stage SSP Runge-Kutta, every stage a full Godunov +
// CT stage at the reconstruction order with the step's dt on the same
// limits:
//   U*  = U + dt L(U)
//   U** = 3/4 U + 1/4 (U* + dt L(U*))
//   U'  = 1/3 U + (1 - 1/3) (U** + dt L(U**))
// (and the faces alike).  The conserved stage states live in arrays of
// their own (U: uc, U*: ws, U**: uss; faces B*: bs, B**: bss); the prim
// array is rewritten on the limits after each stage, as the reference's
// Ws and Wss alias W, with no ghost fill between the stages (ng = 9 keeps
// the stale values the stages read near the grid edges out of the active
// cells).  The update of a stage is the VL update's arithmetic (advance_hydro,
// advance_b, interp_bc, cons2prim) with the stage's conserved input in place
// of prim2cons(w), and the combination with U and B on the limits.
#pragma once
#include "bnz_kern64.hpp"

namespace bnz {

// Stage weights of the combination c1 * U + c2 * (stage result); stage 1
// has none (_rk3.pyx:118-121, 152-154).
template <int STAGE> struct RkWeights;
template <> struct RkWeights<1> {
  static constexpr double c1 = 0., c2 = 1.;
};
template <> struct RkWeights<2> {
  static constexpr double c1 = 0.75, c2 = 0.25;
};
template <> struct RkWeights<3> {
  static constexpr double c1 = 1. / 3, c2 = 1. - 1. / 3;
};

// One RK3 stage of cell (i, j, k): u0 the stage's conserved input, b0 its
// faces; ua, ba the step's U and B (stages 2, 3; they may alias u1, b1, as
// stage 3 writes U and B in place: each cell reads and writes only its own
// slots of them).  Writes the stage result (combined) to u1 and its lower
// faces to b1 (faces outside the limits uncombined, as the reference's
// combine covers the cells of the limits only), and cons2prim of it to w1.
// Returns the CFL speed of an active cell (DTRED, stage 3).
template <int DIMS, int STAGE, bool DTRED>
__device__ __forceinline__ double update_cell_rk(const Geo &g, const double *__restrict__ u0,
                                                 const FluxGlobal &fl, const EmfGlobal &em,
                                                 const double *__restrict__ b0, const double *ua,
                                                 const double *ba, double *u1, double *b1,
                                                 double *w1, const Lims &L, double dt,
                                                 const Box &act, int fms_ref, int i, int j,
                                                 int k) {
  const GeoArr gm{g};
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sv = g.sv;
  const bool inI = i >= L.lo[0] && i <= L.hi[0];
  const bool inJ = j >= L.lo[1] && j <= L.hi[1];
  const bool inK = k >= L.lo[2] && k <= L.hi[2];
  const bool cell = inI && inJ && inK;
  const bool fX = (i >= L.lo[0] && i <= L.hi[0] + 1) && inJ && inK;
  const bool fY = inI && (j >= L.lo[1] && j <= L.hi[1] + (DIMS >= 2 ? 1 : 0)) && inK;
  const bool fZ = inI && inJ && (k >= L.lo[2] && k <= L.hi[2] + (DIMS >= 3 ? 1 : 0));
  if (!(cell || fX || fY || fZ)) return 0.;
  // neighbour offsets, each used only by roles that keep it in the grid
  const long ox = (cell || fY || fZ) ? 1 : 0;
  const long oy = (DIMS >= 2 && (cell || fX || fZ)) ? g.sy : 0;
  const long oz = (DIMS >= 3 && (cell || fX || fY)) ? g.sz : 0;

  const double bxc = b0[c], byc = b0[sv + c], bzc = b0[2 * sv + c];
  const double bxp = b0[c + ox], byp = b0[sv + c + oy], bzp = b0[2 * sv + c + oz];
  const double ez00 = em(2, 0), ez10 = em(2, ox), ez01 = em(2, oy), ez11 = em(2, ox + oy);
  const double ey00 = em(1, 0), ey10 = em(1, ox), ey01 = em(1, oz), ey11 = em(1, ox + oz);
  const double ex00 = em(0, 0), ex10 = em(0, oy), ex01 = em(0, oz), ex11 = em(0, oy + oz);
  const double dlx = gm.dlf(0, i), dly = gm.dlf(1, j), dlz = gm.dlf(2, k);

  // ---- faces (advance_b), as in update_cell_e
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
  constexpr double c1 = RkWeights<STAGE>::c1, c2 = RkWeights<STAGE>::c2;
  // the lower faces of this cell: combined on the limits (stages 2, 3)
  if (fX) b1[c] = (STAGE > 1 && cell) ? c1 * ba[c] + c2 * bxl : bxl;
  if (fY) b1[sv + c] = (STAGE > 1 && cell) ? c1 * ba[sv + c] + c2 * byl : byl;
  if (fZ) b1[2 * sv + c] = (STAGE > 1 && cell) ? c1 * ba[2 * sv + c] + c2 * bzl : bzl;
  if (!cell) return 0.;

  // ---- conserved update (advance_hydro) and cell-centred field (interp_bc)
  double u[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) u[n] = u0[n * sv + c];
  const double dtdv = dt / (dlx * dly * dlz);
  const double ax = dly * dlz, ay = dlx * dlz, az = dlx * dly;
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    double v = u[n] - dtdv * (ax * fl.x(n, ox) - ax * fl.x(n, 0));
    if (DIMS >= 2) v -= dtdv * (ay * fl.y(n, oy) - ay * fl.y(n, 0));
    if (DIMS >= 3) v -= dtdv * (az * fl.z(n, oz) - az * fl.z(n, 0));
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
  if (STAGE > 1) {
#pragma unroll
    for (int n = 0; n < NMODE; ++n) u[n] = c1 * ua[n * sv + c] + c2 * u[n];
  }
  double wn[NMODE];
  cons2prim(wn, u, g.gam);
#pragma unroll
  for (int n = 0; n < NMODE; ++n) {
    u1[n * sv + c] = u[n];
    w1[n * sv + c] = wn[n];
  }
  if (DTRED && i >= act.lo[0] && i < act.lo[0] + act.n[0] && j >= act.lo[1] &&
      j < act.lo[1] + act.n[1] && k >= act.lo[2] && k < act.lo[2] + act.n[2])
    return cds_cell_g(gm, g, wn, fms_ref, i, j, k);
  return 0.;
}

}  // namespace bnz
