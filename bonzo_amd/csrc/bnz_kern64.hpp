// bnz_kern64.hpp -- per-face, per-edge and per-cell bodies of the FP64 VL
// step, shared by the single-grid kernels (bnz_gpu.hip) and the batched
// kernels (bnz_batch.hip), so that a box of a batch follows exactly the
// arithmetic of a single grid.
#pragma once
#include "bnz_common.hpp"

namespace bnz {

// Godunov fluxes of one direction (godunov.pyx:20-210 with the double-buffered
// left state of D1).  Face index c is the lower face of cell c, between cells
// c-1 and c.  The flux is stored in global variable order, flux[n][c], except
// the normal-field slot, which no consumer reads.
template <int DIR, int ORDER, int RS>
__device__ __forceinline__ void flux_face(const Geo &g, const double *__restrict__ w,
                                          const double *__restrict__ bf,
                                          double *__restrict__ fl, int i, int j, int k) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sd = (DIR == 0) ? 1 : (DIR == 1) ? g.sy : g.sz;
  const double bn = bf[(long)DIR * g.sv + c];

  double wl[NMODE], wr[NMODE];
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
      const double dl = 0.5 * vl_lim(wm1 - wm2, w0 - wm1);
      const double dr = 0.5 * vl_lim(w0 - wm1, wp1 - w0);
      wl[n] = wm1 + dl;
      wr[n] = w0 - dr;
    }
  }
  wl[BX] = bn;
  wr[BX] = bn;

  double f[NMODE];
  int reg;
  if (RS == BNZ_RS_HLLD)
    reg = flux_hlld(f, wl, wr, g.gam, g.gamm1i);
  else
    reg = flux_hll(f, wl, wr, g.gam, g.gamm1i);
  if (reg == 0) return;  // reference leaves the flux row untouched
#pragma unroll
  for (int n = 0; n < NMODE; ++n) {
    if (n == BX) continue;
    fl[(long)perm<DIR>(n) * g.sv + c] = f[n];
  }
}

__device__ __forceinline__ double ec_comp(const double *w, long sv, long c, int comp) {
  const double vx = w[VX * sv + c], vy = w[VY * sv + c], vz = w[VZ * sv + c];
  const double bx = w[BX * sv + c], by = w[BY * sv + c], bz = w[BZ * sv + c];
  if (comp == 0) return -(vy * bz - vz * by);
  if (comp == 1) return (vx * bz - vz * bx);
  return -(vx * by - vy * bx);
}

// Edge EMFs, interp_ee1 (src/integration/ct.pyx:153-215).
__device__ __forceinline__ void emf_edge(const Geo &g, const double *__restrict__ w,
                                         const double *__restrict__ fx,
                                         const double *__restrict__ fy,
                                         const double *__restrict__ fz,
                                         double *__restrict__ ee, int i, int j, int k) {
  const long sv = g.sv;
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sy = g.sy, sz = g.sz;
  if (g.dims == 3) {
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
  } else if (g.dims == 2) {
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

// --------------------------------------------------------------------------
// Face-field update by Stokes (advance_b, ct.pyx:15-91) for one face.
__device__ __forceinline__ double bface_x(const Geo &g, const double *b0, const double *ee,
                                          double dt, int i, int j, int k, long c) {
  const long sv = g.sv;
  const double dtda = dt / (g.dlf[1][j] * g.dlf[2][k]);
  double v = b0[c];
  if (g.dims >= 2) {
    const double ds = g.dlf[2][k];
    v -= dtda * (ds * ee[2 * sv + c + g.sy] - ds * ee[2 * sv + c]);
  }
  if (g.dims >= 3) {
    const double ds = g.dlf[1][j];
    v += dtda * (ds * ee[1 * sv + c + g.sz] - ds * ee[1 * sv + c]);
  }
  return v;
}

__device__ __forceinline__ double bface_y(const Geo &g, const double *b0, const double *ee,
                                          double dt, int i, int j, int k, long c) {
  const long sv = g.sv;
  const double dtda = dt / (g.dlf[0][i] * g.dlf[2][k]);
  double ds = g.dlf[2][k];
  double v = b0[1 * sv + c] + dtda * (ds * ee[2 * sv + c + 1] - ds * ee[2 * sv + c]);
  if (g.dims >= 3) {
    ds = g.dlf[0][i];
    v -= dtda * (ds * ee[0 * sv + c + g.sz] - ds * ee[0 * sv + c]);
  }
  return v;
}

__device__ __forceinline__ double bface_z(const Geo &g, const double *b0, const double *ee,
                                          double dt, int i, int j, int k, long c) {
  const long sv = g.sv;
  const double dtda = dt / (g.dlf[0][i] * g.dlf[1][j]);
  double ds = g.dlf[1][j];
  double v = b0[2 * sv + c] - dtda * (ds * ee[1 * sv + c + 1] - ds * ee[1 * sv + c]);
  if (g.dims >= 2) {
    ds = g.dlf[0][i];
    v += dtda * (ds * ee[0 * sv + c + g.sy] - ds * ee[0 * sv + c]);
  }
  return v;
}


// Stage update: advance_hydro (godunov.pyx:215-252), advance_b, interp_bc
// (ct.pyx:96-123) and cons2prim (eos.pyx:13-41) fused, one thread per cell
// of the whole grid.  u0 = prim2cons(w0) is the reference's cons array.
// PRED: writes ws (a separate buffer); cells outside the limits take w0,
// as the reference's ws aliases prim (vanleer.pyx:72).  Corrector: w1 may
// alias w0 (each thread reads only its own cell of w0).
template <bool PRED, bool DTRED>
__device__ __forceinline__ double update_cell(const Geo &g, const double *w0,
                                              const double *__restrict__ fx,
                                              const double *__restrict__ fy,
                                              const double *__restrict__ fz,
                                              const double *__restrict__ ee,
                                              const double *__restrict__ b0,
                                              double *__restrict__ b1, double *w1, const Lims &L,
                                              const double *__restrict__ dts, const Box &act,
                                              int fms_ref, int i, int j, int k) {
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sv = g.sv;
  const double dt = PRED ? 0.5 * dts[0] : dts[0];

  const bool inI = i >= L.lo[0] && i <= L.hi[0];
  const bool inJ = j >= L.lo[1] && j <= L.hi[1];
  const bool inK = k >= L.lo[2] && k <= L.hi[2];
  const bool cell = inI && inJ && inK;
  const bool fX = (i >= L.lo[0] && i <= L.hi[0] + 1) && inJ && inK;
  const bool fY = inI && (j >= L.lo[1] && j <= L.hi[1] + (g.dims >= 2 ? 1 : 0)) && inK;
  const bool fZ = inI && inJ && (k >= L.lo[2] && k <= L.hi[2] + (g.dims >= 3 ? 1 : 0));

  double bxl = 0., byl = 0., bzl = 0.;
  if (fX) {
    bxl = bface_x(g, b0, ee, dt, i, j, k, c);
    b1[c] = bxl;
  }
  if (fY) {
    byl = bface_y(g, b0, ee, dt, i, j, k, c);
    b1[sv + c] = byl;
  }
  if (fZ) {
    bzl = bface_z(g, b0, ee, dt, i, j, k, c);
    b1[2 * sv + c] = bzl;
  }

  if (!cell) {
    if (PRED) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n) w1[n * sv + c] = w0[n * sv + c];
    }
    return 0.;
  }

  double wv[NMODE], u[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) wv[n] = w0[n * sv + c];
  prim2cons(u, wv, g.gamm1i);

  const double dlx = g.dlf[0][i], dly = g.dlf[1][j], dlz = g.dlf[2][k];
  const double dtdv = dt / (dlx * dly * dlz);
  const double ax = dly * dlz, ay = dlx * dlz, az = dlx * dly;
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    double v = u[n] - dtdv * (ax * fx[n * sv + c + 1] - ax * fx[n * sv + c]);
    if (g.dims >= 2) v -= dtdv * (ay * fy[n * sv + c + g.sy] - ay * fy[n * sv + c]);
    if (g.dims >= 3) v -= dtdv * (az * fz[n * sv + c + g.sz] - az * fz[n * sv + c]);
    u[n] = v;
  }

  // interp_bc with the upper faces recomputed in-thread
  const double bxu = bface_x(g, b0, ee, dt, i + 1, j, k, c + 1);
  u[BX] = (g.icm[0][i] * bxl + g.icp[0][i] * bxu) * g.dlfi[0][i];
  if (g.dims >= 2) {
    const double byu = bface_y(g, b0, ee, dt, i, j + 1, k, c + g.sy);
    u[BY] = (g.icm[1][j] * byl + g.icp[1][j] * byu) * g.dlfi[1][j];
  } else {
    u[BY] = byl;
  }
  if (g.dims >= 3) {
    const double bzu = bface_z(g, b0, ee, dt, i, j, k + 1, c + g.sz);
    u[BZ] = (g.icm[2][k] * bzl + g.icp[2][k] * bzu) * g.dlfi[2][k];
  } else {
    u[BZ] = bzl;
  }
  cons2prim(wv, u, g.gam);
#pragma unroll
  for (int n = 0; n < NMODE; ++n) w1[n * sv + c] = wv[n];
  if (DTRED && i >= act.lo[0] && i < act.lo[0] + act.n[0] && j >= act.lo[1] &&
      j < act.lo[1] + act.n[1] && k >= act.lo[2] && k < act.lo[2] + act.n[2])
    return cds_cell(g, wv, fms_ref, i, j, k);
  return 0.;
}


}  // namespace bnz
