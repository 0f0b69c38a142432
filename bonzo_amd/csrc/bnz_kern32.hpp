// bnz_kern32.hpp -- per-face, per-edge and per-cell bodies of the
// FP32-compensated VL step, shared by the single-grid kernels
// (bnz_gpu32.hip) and the batched kernels (bnz_batch32.hip), so that a box
// of a batch follows exactly the arithmetic of a single grid.  Compiled with
// FMA contraction of single expressions only (-ffp-contract=on).
#pragma once
#include "bnz_common.hpp"

namespace bnz {

__device__ __forceinline__ void split(double v, float &hi, float &lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}
__device__ __forceinline__ double join(float hi, float lo) { return (double)hi + (double)lo; }


// Godunov flux of one face, FP32 (same stencil as the FP64 flux_face); ng
// counts guarded HLLD star denominators.
// flux32_face_to hands the flux (local slots) of a face with a matched region
// to sink(f, c); flux32_face stores it.
template <int DIR, int ORDER, int RS, typename BF, class SINK>
__device__ __forceinline__ void flux32_face_to(const Geo &g, const float *__restrict__ w,
                                               const BF *__restrict__ bf, int i, int j, int k,
                                               int &ng, const SINK &sink) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sd = (DIR == 0) ? 1 : (DIR == 1) ? g.sy : g.sz;
  const float bn = (float)bf[(long)DIR * g.sv + c];
  float wl[NMODE], wr[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) {
    if (n == BX) continue;
    const float *p = w + (long)perm<DIR>(n) * g.sv + c;
    if (ORDER == 0) {
      wl[n] = p[-sd];
      wr[n] = p[0];
    } else {
      const float wm2 = p[-2 * sd], wm1 = p[-sd], w0 = p[0], wp1 = p[sd];
      wl[n] = wm1 + vl_half(wm1 - wm2, w0 - wm1);
      wr[n] = w0 - vl_half(w0 - wm1, wp1 - w0);
    }
  }
  wl[BX] = bn;
  wr[BX] = bn;
  float f[NMODE];
  const float gam = (float)g.gam, gm1i = (float)g.gamm1i;
  const int reg = (RS == BNZ_RS_HLLD) ? flux_hlld(f, wl, wr, gam, gm1i, &ng)
                                      : flux_hll(f, wl, wr, gam, gm1i);
  if (reg == 0) return;
  sink(f, c);
}

template <int DIR, int ORDER, int RS, typename BF>
__device__ __forceinline__ void flux32_face(const Geo &g, const float *__restrict__ w,
                                            const BF *__restrict__ bf, float *__restrict__ fl,
                                            int i, int j, int k, int &ng) {
  const long sv = g.sv;
  flux32_face_to<DIR, ORDER, RS>(g, w, bf, i, j, k, ng, [fl, sv](const float *f, long c) {
#pragma unroll
    for (int n = 0; n < NMODE; ++n) {
      if (n == BX) continue;
      fl[(long)perm<DIR>(n) * sv + c] = f[n];
    }
  });
}


// All face fluxes of one stage whose lower cell is (i, j, k) (as the FP64
// flux_cell, DIMS included); ng counts guarded HLLD star denominators.
template <int ORDER, int RS, typename BF, int DIMS = 0>
__device__ __forceinline__ void flux32_cell(const Geo &g, const float *__restrict__ w,
                                            const BF *__restrict__ bf, float *__restrict__ fx,
                                            float *__restrict__ fy, float *__restrict__ fz,
                                            const Lims &L, int i, int j, int k, int &ng) {
  const int dims = DIMS > 0 ? DIMS : g.dims;
  if (i >= L.lo[0]) flux32_face<0, ORDER, RS>(g, w, bf, fx, i, j, k, ng);
  if (dims >= 2 && j >= L.lo[1]) flux32_face<1, ORDER, RS>(g, w, bf, fy, i, j, k, ng);
  if (dims >= 3 && k >= L.lo[2]) flux32_face<2, ORDER, RS>(g, w, bf, fz, i, j, k, ng);
}


__device__ __forceinline__ float ec32(const float *w, long sv, long c, int comp) {
  const float vx = w[VX * sv + c], vy = w[VY * sv + c], vz = w[VZ * sv + c];
  const float bx = w[BX * sv + c], by = w[BY * sv + c], bz = w[BZ * sv + c];
  if (comp == 0) return -(vy * bz - vz * by);
  if (comp == 1) return (vx * bz - vz * bx);
  return -(vx * by - vy * bx);
}

// Index of the reflecting wall plane of each side (x1, x2, y1, y2, z1, z2),
// or -1.
struct Walls {
  int at[6];
};


// Edge EMFs, FP32 (interp_ee1, src/integration/ct.pyx:153-215), for a grid
// of DIMS dimensions.  On a reflecting wall the two EMF components lying in
// the wall are set to zero (v_n = 0 and B_n = 0 there, so the tangential E
// vanishes): the wall face of the normal field, which the boundary fill
// zeroes after every update, then receives no FP32 round-off from them and
// div B stays at FP64 round-off (D16; the FP64 path keeps the reference's
// computed EMFs, which cancel there to FP64 round-off).
template <int DIMS>
__device__ __forceinline__ void emf32_edge(const Geo &g, const float *__restrict__ w,
                                           const float *__restrict__ fx,
                                           const float *__restrict__ fy,
                                           const float *__restrict__ fz,
                                           float *__restrict__ ee, const Walls &wl, int i, int j,
                                           int k) {
  const long sv = g.sv, sy = g.sy, sz = g.sz;
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const bool wx = i == wl.at[0] || i == wl.at[1];
  const bool wy = DIMS >= 2 && (j == wl.at[2] || j == wl.at[3]);
  const bool wz = DIMS >= 3 && (k == wl.at[4] || k == wl.at[5]);
  float ex, ey, ez;
  if (DIMS == 3) {
    ex = 0.5f * (-fy[BZ * sv + c - sz] - fy[BZ * sv + c] + fz[BY * sv + c - sy] + fz[BY * sv + c]) -
         0.25f * (ec32(w, sv, c - sy, 0) + ec32(w, sv, c, 0) + ec32(w, sv, c - sz - sy, 0) +
                  ec32(w, sv, c - sz, 0));
    ey = 0.5f * (-fz[BX * sv + c - 1] - fz[BX * sv + c] + fx[BZ * sv + c - sz] + fx[BZ * sv + c]) -
         0.25f * (ec32(w, sv, c - 1, 1) + ec32(w, sv, c, 1) + ec32(w, sv, c - sz - 1, 1) +
                  ec32(w, sv, c - sz, 1));
  } else if (DIMS == 2) {
    ex = -fy[BZ * sv + c];
    ey = fx[BZ * sv + c];
  } else {
    ex = -ec32(w, sv, c, 0);
    ey = fx[BZ * sv + c];
  }
  if (DIMS >= 2)
    ez = 0.5f * (-fx[BY * sv + c - sy] - fx[BY * sv + c] + fy[BX * sv + c - 1] + fy[BX * sv + c]) -
         0.25f * (ec32(w, sv, c - 1, 2) + ec32(w, sv, c, 2) + ec32(w, sv, c - sy - 1, 2) +
                  ec32(w, sv, c - sy, 2));
  else
    ez = -fx[BY * sv + c];
  // components lying in a reflecting wall plane
  if (wy || wz) ex = 0.f;
  if (wx || wz) ey = 0.f;
  if (wx || wy) ez = 0.f;
  ee[c] = ex;
  ee[sv + c] = ey;
  ee[2 * sv + c] = ez;
}


// FP32 edge EMFs of the stage update: em(comp, off) is component comp at
// the edge whose lower corner is cell c + off.  Emf32Global reads the ee
// array of the EMF kernel; on 1D grids EmfLds32x1D (below) forms them from
// the x fluxes as emf32_edge<1> does (ey = F_x[Bz], ez = -F_x[By], both
// zero on a reflecting wall plane; a 1D cell reads no Ex), so 1D steps need
// no EMF kernel.
struct Emf32Global {
  const float *ee;
  long sv, c;
  __device__ __forceinline__ float operator()(int comp, long off) const {
    return ee[comp * sv + c + off];
  }
};

// The FP32 2D edge EMFs formed where the update needs them (small 2D grids,
// no EMF launch): emf32_edge<2>'s expressions and wall zeroing at the edge
// whose lower corner is cell c + off, from the stage's FP32 prim we (as the
// FP64 EmfFly2D; an edge on the grid's first row or column gives 0).
struct Emf32Fly2D {
  const float *we, *fx, *fy;
  long sv, sy, c;
  int i, j;
  Walls wl;
  __device__ __forceinline__ float operator()(int comp, long off) const {
    const long e = c + off;
    const int je = j + (off >= sy ? 1 : 0);
    const int ie = i + (int)(off >= sy ? off - sy : off);
    const bool wx = ie == wl.at[0] || ie == wl.at[1];
    const bool wy = je == wl.at[2] || je == wl.at[3];
    if (comp == 0) return wy ? 0.f : -fy[BZ * sv + e];
    if (comp == 1) return wx ? 0.f : fx[BZ * sv + e];
    if (wx || wy || ie < 1 || je < 1) return 0.f;
    const float *w = we;
    return 0.5f * (-fx[BY * sv + e - sy] - fx[BY * sv + e] + fy[BX * sv + e - 1] +
                   fy[BX * sv + e]) -
           0.25f * (ec32(w, sv, e - 1, 2) + ec32(w, sv, e, 2) + ec32(w, sv, e - sy - 1, 2) +
                    ec32(w, sv, e - sy, 2));
  }
};

// Face-field update by Stokes (ct.pyx:15-91) in working type R from face
// values b and FP32 EMFs em: the lower faces of cell c (lo) and its upper
// faces (up).  Neighbour offsets ox, oy, oz are zero where the cell needs no
// neighbour (all loads stay in bounds; see the FP64 update_cell_d).
template <int DIMS, typename R, typename BF, class EM, class GM>
__device__ __forceinline__ void faces_d(const GM &gm, BF b, const EM &em, double dt, int i, int j,
                                        int k, long c, long ox, long oy, long oz, R *lo, R *up) {
  const R ez00 = em(2, 0), ez10 = em(2, ox), ez01 = em(2, oy), ez11 = em(2, ox + oy);
  const R ey00 = em(1, 0), ey10 = em(1, ox), ey01 = em(1, oz), ey11 = em(1, ox + oz);
  const R ex00 = em(0, 0), ex10 = em(0, oy), ex01 = em(0, oz), ex11 = em(0, oy + oz);
  const R bxc = b(0, c), byc = b(1, c), bzc = b(2, c);
  const R bxp = b(0, c + ox), byp = b(1, c + oy), bzp = b(2, c + oz);
  const double dlx = gm.dlf(0, i), dly = gm.dlf(1, j), dlz = gm.dlf(2, k);
  {
    const R dtda = (R)(dt / (dly * dlz));
    R v = bxc, vu = bxp;
    if (DIMS >= 2) {
      const R ds = (R)dlz;
      v -= dtda * (ds * ez01 - ds * ez00);
      vu -= dtda * (ds * ez11 - ds * ez10);
    }
    if (DIMS >= 3) {
      const R ds = (R)dly;
      v += dtda * (ds * ey01 - ds * ey00);
      vu += dtda * (ds * ey11 - ds * ey10);
    }
    lo[0] = v;
    up[0] = vu;
  }
  {
    const R dtda = (R)(dt / (dlx * dlz));
    R ds = (R)dlz;
    R v = byc + dtda * (ds * ez10 - ds * ez00);
    R vu = byp + dtda * (ds * ez11 - ds * ez01);
    if (DIMS >= 3) {
      ds = (R)dlx;
      v -= dtda * (ds * ex01 - ds * ex00);
      vu -= dtda * (ds * ex11 - ds * ex10);
    }
    lo[1] = v;
    up[1] = vu;
  }
  {
    const R dtda = (R)(dt / (dlx * dly));
    R ds = (R)dly;
    R v = bzc - dtda * (ds * ey10 - ds * ey00);
    R vu = bzp - dtda * (ds * ey11 - ds * ey01);
    if (DIMS >= 2) {
      ds = (R)dlx;
      v += dtda * (ds * ex10 - ds * ex00);
      vu += dtda * (ds * ex11 - ds * ex01);
    }
    lo[2] = v;
    up[2] = vu;
  }
}

// Cell roles of the stage update (as update_cell_d) and neighbour offsets.
template <int DIMS>
struct Roles {
  bool cell, fX, fY, fZ;
  long ox, oy, oz;
  __device__ Roles(const Geo &g, const Lims &L, int i, int j, int k) {
    const bool inI = i >= L.lo[0] && i <= L.hi[0], inJ = j >= L.lo[1] && j <= L.hi[1];
    const bool inK = k >= L.lo[2] && k <= L.hi[2];
    cell = inI && inJ && inK;
    fX = (i >= L.lo[0] && i <= L.hi[0] + 1) && inJ && inK;
    fY = inI && (j >= L.lo[1] && j <= L.hi[1] + (DIMS >= 2 ? 1 : 0)) && inK;
    fZ = inI && inJ && (k >= L.lo[2] && k <= L.hi[2] + (DIMS >= 3 ? 1 : 0));
    ox = (cell || fY || fZ) ? 1 : 0;
    oy = (DIMS >= 2 && (cell || fX || fZ)) ? g.sy : 0;
    oz = (DIMS >= 3 && (cell || fX || fY)) ? g.sz : 0;
  }
};

struct Faces32 {
  const double *b;
  long sv;
  __device__ float operator()(int a, long c) const { return (float)b[a * sv + c]; }
};
struct Faces64 {
  const double *b;
  long sv;
  __device__ double operator()(int a, long c) const { return b[a * sv + c]; }
};

// FP32 face fluxes as the updates read them: variable n of the face whose
// lower cell is c + off (the arrays of the flux kernels).
struct FluxArr32 {
  const float *fx, *fy, *fz;
  long sv, c;
  __device__ __forceinline__ float x(int n, long off) const { return fx[n * sv + c + off]; }
  __device__ __forceinline__ float y(int n, long off) const { return fy[n * sv + c + off]; }
  __device__ __forceinline__ float z(int n, long off) const { return fz[n * sv + c + off]; }
};

struct Pred32 {
  const float *w32, *uh;
  const double *b;
  const float *fx, *fy, *fz, *ee;
  float *bs, *ws;
};


// Predictor (dt/2): FP32 update of the hi words, FP64 pressure recovery.
// All operands are loaded up front; only the stores (and the copy of w into
// ws outside the limits) depend on the roles.
// Fluxes come from fl (x, y, z accessors), EMFs from em.  LDS_OUT (the
// predictor of a fused 1D step): the new prim goes to the LDS array a.ws as
// a.ws[n * BLK + lslot] and no face is stored.
template <int DIMS, class GM, class FL, class EM, bool LDS_OUT = false>
__device__ __forceinline__ void update32_pred_e(const Geo &g, const Pred32 &a, const Lims &L,
                                                const double *__restrict__ dts, const FL &fl,
                                                const EM &em, int i, int j, int k,
                                                int lslot = 0) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i, sv = g.sv;
  const double dt = 0.5 * dts[0];
  const Roles<DIMS> ro(g, L, i, j, k);
  float uh[NMODE - 3], fxl[NMODE - 3], fxu[NMODE - 3], fyl[NMODE - 3], fyu[NMODE - 3],
      fzl[NMODE - 3], fzu[NMODE - 3];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    uh[n] = a.uh[n * sv + c];
    fxl[n] = fl.x(n, 0);
    fxu[n] = fl.x(n, ro.ox);
    if (DIMS >= 2) {
      fyl[n] = fl.y(n, 0);
      fyu[n] = fl.y(n, ro.oy);
    }
    if (DIMS >= 3) {
      fzl[n] = fl.z(n, 0);
      fzu[n] = fl.z(n, ro.oz);
    }
  }
  float lo[3], up[3];
  const GM gm{g};
  faces_d<DIMS, float>(gm, Faces32{a.b, sv}, em, dt, i, j, k, c, ro.ox, ro.oy, ro.oz, lo, up);
  const float dlx = (float)gm.dlf(0, i), dly = (float)gm.dlf(1, j), dlz = (float)gm.dlf(2, k);
  const float dtdv = (float)(dt / (gm.dlf(0, i) * gm.dlf(1, j) * gm.dlf(2, k)));
  const float ax = dly * dlz, ay = dlx * dlz, az = dlx * dly;
  double u[NMODE], wv[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    float v = uh[n] - dtdv * (ax * fxu[n] - ax * fxl[n]);
    if (DIMS >= 2) v -= dtdv * (ay * fyu[n] - ay * fyl[n]);
    if (DIMS >= 3) v -= dtdv * (az * fzu[n] - az * fzl[n]);
    u[n] = v;
  }
  u[BX] = ((float)gm.icm(0, i) * lo[0] + (float)gm.icp(0, i) * up[0]) * (float)gm.dlfi(0, i);
  if (DIMS >= 2)
    u[BY] = ((float)gm.icm(1, j) * lo[1] + (float)gm.icp(1, j) * up[1]) * (float)gm.dlfi(1, j);
  else
    u[BY] = lo[1];
  if (DIMS >= 3)
    u[BZ] = ((float)gm.icm(2, k) * lo[2] + (float)gm.icp(2, k) * up[2]) * (float)gm.dlfi(2, k);
  else
    u[BZ] = lo[2];
  cons2prim(wv, u, g.gam);
  if (LDS_OUT) {
    if (ro.cell) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n) a.ws[n * BLK + lslot] = (float)wv[n];
    } else {
#pragma unroll
      for (int n = 0; n < NMODE; ++n) a.ws[n * BLK + lslot] = a.w32[n * sv + c];
    }
    return;
  }
  if (ro.fX) a.bs[c] = lo[0];
  if (ro.fY) a.bs[sv + c] = lo[1];
  if (ro.fZ) a.bs[2 * sv + c] = lo[2];
  if (ro.cell) {
#pragma unroll
    for (int n = 0; n < NMODE; ++n) a.ws[n * sv + c] = (float)wv[n];
  } else {  // outside the limits ws keeps w (read only here: the thin outer shell)
#pragma unroll
    for (int n = 0; n < NMODE; ++n) a.ws[n * sv + c] = a.w32[n * sv + c];
  }
}

// 2D and 3D updates read the flux and EMF arrays (1D grids run stage32_cell);
// FLY (2D) forms the EMFs from the stage's prim we instead (Emf32Fly2D).
template <int DIMS, class GM = GeoArr, bool FLY = false>
__device__ __forceinline__ void update32_pred(const Geo &g, const Pred32 &a, const Lims &L,
                                              const double *__restrict__ dts, int i, int j,
                                              int k, const float *we = nullptr,
                                              const Walls &wl = Walls{}) {
  static_assert(DIMS >= 2, "1D grids run stage32_cell");
  static_assert(!FLY || DIMS == 2, "EMFs in the update on 2D grids only");
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i, sv = g.sv;
  if (FLY)
    update32_pred_e<DIMS, GM>(g, a, L, dts, FluxArr32{a.fx, a.fy, a.fz, sv, c},
                              Emf32Fly2D{we, a.fx, a.fy, sv, (long)g.sy, c, i, j, wl}, i, j, k);
  else
    update32_pred_e<DIMS, GM>(g, a, L, dts, FluxArr32{a.fx, a.fy, a.fz, sv, c},
                              Emf32Global{a.ee, sv, c}, i, j, k);
}


struct Corr32 {
  float *uh, *ul;
  const double *b;
  const float *fx, *fy, *fz, *ee;
  double *b1;
  float *w32;
};


// Corrector (dt): FP64 update of the hi + lo state, FP64 pressure recovery,
// returns the CFL speed of the new state (0 outside the active cells).  Operands loaded up front as in the
// predictor.
// SPLIT (the corrector of a fused 1D step): the new hi and lo words go to
// uh1 and ul1 instead of back to a.uh and a.ul.
template <int DIMS, class GM, class FL, class EM, bool SPLIT = false>
__device__ __forceinline__ double update32_corr_e(const Geo &g, const Corr32 &a, const Lims &L,
                                                  const double *__restrict__ dts, const Box &act,
                                                  int fms_ref, const FL &fl, const EM &em, int i,
                                                  int j, int k, float *uh1 = nullptr,
                                                  float *ul1 = nullptr) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i, sv = g.sv;
  double cdsmax = 0.;
  const double dt = dts[0];
  const Roles<DIMS> ro(g, L, i, j, k);
  float uh[NMODE - 3], ul[NMODE - 3], fxl[NMODE - 3], fxu[NMODE - 3], fyl[NMODE - 3],
      fyu[NMODE - 3], fzl[NMODE - 3], fzu[NMODE - 3];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    uh[n] = a.uh[n * sv + c];
    ul[n] = a.ul[n * sv + c];
    fxl[n] = fl.x(n, 0);
    fxu[n] = fl.x(n, ro.ox);
    if (DIMS >= 2) {
      fyl[n] = fl.y(n, 0);
      fyu[n] = fl.y(n, ro.oy);
    }
    if (DIMS >= 3) {
      fzl[n] = fl.z(n, 0);
      fzu[n] = fl.z(n, ro.oz);
    }
  }
  double lo[3], up[3];
  const GM gm{g};
  faces_d<DIMS, double>(gm, Faces64{a.b, sv}, em, dt, i, j, k, c, ro.ox, ro.oy, ro.oz, lo, up);
  const double dlx = gm.dlf(0, i), dly = gm.dlf(1, j), dlz = gm.dlf(2, k);
  const double dtdv = dt / (dlx * dly * dlz);
  const double ax = dly * dlz, ay = dlx * dlz, az = dlx * dly;
  double u[NMODE], wv[NMODE];
  float nh[NMODE - 3], nl[NMODE - 3];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    double v = join(uh[n], ul[n]) - dtdv * (ax * (double)fxu[n] - ax * (double)fxl[n]);
    if (DIMS >= 2) v -= dtdv * (ay * (double)fyu[n] - ay * (double)fyl[n]);
    if (DIMS >= 3) v -= dtdv * (az * (double)fzu[n] - az * (double)fzl[n]);
    u[n] = v;
    split(v, nh[n], nl[n]);
  }
  u[BX] = (gm.icm(0, i) * lo[0] + gm.icp(0, i) * up[0]) * gm.dlfi(0, i);
  if (DIMS >= 2)
    u[BY] = (gm.icm(1, j) * lo[1] + gm.icp(1, j) * up[1]) * gm.dlfi(1, j);
  else
    u[BY] = lo[1];
  if (DIMS >= 3)
    u[BZ] = (gm.icm(2, k) * lo[2] + gm.icp(2, k) * up[2]) * gm.dlfi(2, k);
  else
    u[BZ] = lo[2];
  cons2prim(wv, u, g.gam);
  if (ro.fX) a.b1[c] = lo[0];
  if (ro.fY) a.b1[sv + c] = lo[1];
  if (ro.fZ) a.b1[2 * sv + c] = lo[2];
  if (ro.cell) {
#pragma unroll
    for (int n = 0; n < NMODE - 3; ++n) {
      (SPLIT ? uh1 : a.uh)[n * sv + c] = nh[n];
      (SPLIT ? ul1 : a.ul)[n * sv + c] = nl[n];
    }
#pragma unroll
    for (int n = 0; n < NMODE; ++n) a.w32[n * sv + c] = (float)wv[n];
    if (i >= act.lo[0] && i < act.lo[0] + act.n[0] && j >= act.lo[1] &&
        j < act.lo[1] + act.n[1] && k >= act.lo[2] && k < act.lo[2] + act.n[2])
      cdsmax = cds_cell_g(gm, g, wv, fms_ref, i, j, k);
  }
  return cdsmax;
}

template <int DIMS, class GM = GeoArr, bool FLY = false>
__device__ __forceinline__ double update32_corr(const Geo &g, const Corr32 &a, const Lims &L,
                                                const double *__restrict__ dts, const Box &act,
                                                int fms_ref, int i, int j, int k,
                                                const float *we = nullptr,
                                                const Walls &wl = Walls{}) {
  static_assert(DIMS >= 2, "1D grids run stage32_cell");
  static_assert(!FLY || DIMS == 2, "EMFs in the update on 2D grids only");
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i, sv = g.sv;
  if (FLY)
    return update32_corr_e<DIMS, GM>(g, a, L, dts, act, fms_ref,
                                     FluxArr32{a.fx, a.fy, a.fz, sv, c},
                                     Emf32Fly2D{we, a.fx, a.fy, sv, (long)g.sy, c, i, j, wl}, i,
                                     j, k);
  return update32_corr_e<DIMS, GM>(g, a, L, dts, act, fms_ref,
                                   FluxArr32{a.fx, a.fy, a.fz, sv, c}, Emf32Global{a.ee, sv, c},
                                   i, j, k);
}


// Fluxes of a run of 1D faces in LDS (as the FP64 FluxLds1D) and the 1D
// EMFs formed from them, zero on the reflecting wall planes wlo / whi.
struct FluxLds32 {
  const float *s;
  int t;
  __device__ __forceinline__ float x(int n, long off) const { return s[n * BLK + t + off]; }
  __device__ __forceinline__ float y(int, long) const { return 0.f; }
  __device__ __forceinline__ float z(int, long) const { return 0.f; }
};
struct EmfLds32x1D {
  const float *s;
  int t, i, wlo, whi;
  __device__ __forceinline__ float operator()(int comp, long off) const {
    const int e = i + (int)off;
    if (comp == 0 || e == wlo || e == whi) return 0.f;
    return comp == 2 ? -s[BY * BLK + t + off] : s[BZ * BLK + t + off];
  }
};

// One FP32C stage of a 1D grid for thread threadIdx.x of a workgroup of BLK,
// as the FP64 stage1_cell: the flux of face i into sf (NMODE x BLK floats),
// a barrier, then the update of cell i (update32_pred_e / update32_corr_e on
// the structs ap / ac) from faces i and i + 1.  The predictor's fluxes read
// the FP32 prim and the FP64 faces of the step (ap.w32, ap.b); the
// corrector's read the predictor's FP32 prim and faces (ws32, bs32), which
// the corrector does not write.  wlo / whi are the x wall planes (-1: none).
// Returns the CFL speed of the updated cell (corrector).
template <bool PRED, int ORDER, int RS>
__device__ __forceinline__ double stage32_cell(const Geo &g, const float *__restrict__ ws32,
                                               const float *__restrict__ bs32, const Pred32 &ap,
                                               const Corr32 &ac, const Lims &L,
                                               const double *__restrict__ dts, const Box &act,
                                               int fms_ref, int wlo, int whi, int i, float *sf,
                                               int &ng) {
  const int tid = threadIdx.x;
  if (i >= L.lo[0] && i <= L.hi[0] + 1) {  // the stage's faces (flux3_box)
    bool got = false;
    const auto sink = [&](const float *f, long) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n)
        if (n != BX) sf[n * BLK + tid] = f[n];
      got = true;
    };
    if (PRED)
      flux32_face_to<0, ORDER, RS>(g, ap.w32, ap.b, i, 0, 0, ng, sink);
    else
      flux32_face_to<0, ORDER, RS>(g, ws32, bs32, i, 0, 0, ng, sink);
    if (!got)  // no region matched: the flux row keeps what it held
      for (int n = 0; n < NMODE; ++n)
        if (n != BX) sf[n * BLK + tid] = ap.fx[n * g.sv + i];
  }
  __syncthreads();
  double cds = 0.;
  if (tid < BLK - 1 && i < g.nt[0]) {
    const FluxLds32 fl{sf, tid};
    const EmfLds32x1D em{sf, tid, i, wlo, whi};
    if (PRED)
      update32_pred_e<1, GeoArr>(g, ap, L, dts, fl, em, i, 0, 0);
    else
      cds = update32_corr_e<1, GeoArr>(g, ac, L, dts, act, fms_ref, fl, em, i, 0, 0);
  }
  return cds;
}

// CFL speed of one active cell from the FP32 primitives.
__device__ __forceinline__ double cfl32_cell(const Geo &g, const float *__restrict__ w,
                                             int fms_ref, int i, int j, int k) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  double wv[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) wv[n] = (double)w[n * g.sv + c];
  return cds_cell(g, wv, fms_ref, i, j, k);
}

// The ghost fill of the FP32C state: conserved hi words (uh) and FP32 prim
// (w32) of every variable, and the face fields.  The low words (ul) get no
// ghosts: only the corrector reads them, and of the cells it updates only the
// active ones keep their result (the fill overwrites the others' hi words and
// prim after the step).
constexpr int NGHOST32 = 6 + NMODE;
struct Ghost32 {
  float *a[NGHOST32];  // uh (6), w32 (9)
  double *d[3];        // face fields
  const float *fix32;  // user fixed values: NGHOST32 state words per row
  const double *fix64; // and NMODE + 3 FP64 values (faces at NMODE..)
};

// One (ghost, source) pair: sources are active cells the fill never writes,
// so every value is loaded before any is stored.
__device__ __forceinline__ void ghost32_pair(const Ghost32 &p, int dst, int src) {
  float v[NGHOST32];
  double f[3];
#pragma unroll
  for (int n = 0; n < NGHOST32; ++n) v[n] = p.a[n][src];
#pragma unroll
  for (int n = 0; n < 3; ++n) f[n] = p.d[n][src];
#pragma unroll
  for (int n = 0; n < NGHOST32; ++n) p.a[n][dst] = v[n];
#pragma unroll
  for (int n = 0; n < 3; ++n) p.d[n][dst] = f[n];
}

// A ghost cell a reflecting side touches (GhostOp); the sign bit of slot v
// applies to conserved slot v of uh (momentum flips with velocity) and to
// prim slot v of w32.
__device__ __forceinline__ void ghost32_op(const Ghost32 &p, const GhostOp &o) {
  const bool cells = !(o.bits & (1u << 15));
  float v[NGHOST32];
  double f[3];
  if (o.bits & GHOST_FIXED) {
#pragma unroll
    for (int n = 0; n < NGHOST32; ++n) v[n] = p.fix32[(long)o.src * NGHOST32 + n];
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = p.fix64[(long)o.src * (NMODE + 3) + NMODE + a];
  } else {
#pragma unroll
    for (int n = 0; n < NGHOST32; ++n) v[n] = cells ? p.a[n][o.src] : 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = ((o.bits >> (12 + a)) & 1u) ? 0. : p.d[a][o.srcf[a]];
  }
  if (cells) {
#pragma unroll
    for (int n = 0; n < NGHOST32; ++n) {
      const int slot = n < 6 ? n : n - 6;
      p.a[n][o.dst] = ((o.bits >> slot) & 1u) ? -v[n] : v[n];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) p.d[a][o.dst] = ((o.bits >> (9 + a)) & 1u) ? -f[a] : f[a];
}

// FP64 prim of cell c -> FP32-compensated state (conserved hi + lo, FP32 prim).
__device__ __forceinline__ void from64_cell(const Geo &g, const double *__restrict__ w,
                                            float *uh, float *ul, float *w32, long c) {
  const long sv = g.sv;
  double wv[NMODE], u[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) {
    wv[n] = w[n * sv + c];
    w32[n * sv + c] = (float)wv[n];
  }
  prim2cons(u, wv, g.gamm1i);
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) split(u[n], uh[n * sv + c], ul[n * sv + c]);
}

// FP32-compensated state of active cell (i, j, k) -> FP64 prim.
__device__ __forceinline__ void to64_cell(const Geo &g, const float *uh, const float *ul,
                                          const double *b, double *w, int i, int j, int k) {
  const long sv = g.sv;
  const long t = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  double u[NMODE], wv[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) u[n] = join(uh[n * sv + t], ul[n * sv + t]);
  u[BX] = (g.icm[0][i] * b[t] + g.icp[0][i] * b[t + 1]) * g.dlfi[0][i];
  if (g.dims >= 2)
    u[BY] = (g.icm[1][j] * b[sv + t] + g.icp[1][j] * b[sv + t + g.sy]) * g.dlfi[1][j];
  else
    u[BY] = b[sv + t];
  if (g.dims >= 3)
    u[BZ] = (g.icm[2][k] * b[2 * sv + t] + g.icp[2][k] * b[2 * sv + t + g.sz]) * g.dlfi[2][k];
  else
    u[BZ] = b[2 * sv + t];
  cons2prim(wv, u, g.gam);
#pragma unroll
  for (int n = 0; n < NMODE; ++n) w[n * sv + t] = wv[n];
}

}  // namespace bnz
