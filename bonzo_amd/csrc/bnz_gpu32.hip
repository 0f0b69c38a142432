// bnz_gpu32.hip -- FP32-compensated Van-Leer step (BNZ_PREC_FP32C).
//
// Same sequence as the FP64 step (src/integration/vanleer.pyx:100-163) with
// the arithmetic split by need:
//   FP32   reconstruction, Riemann solvers, EMFs (hardware rcp/sqrt, FMA)
//          and the predictor's half-step update;
//   FP64   the corrector's conserved and face-field updates (applied to the
//          hi + lo pairs of the state, whose low words carry the rounding
//          residual from step to step), the pressure recovery
//          p = (gam-1)(E - (B^2 + rho v^2)/2), which cancels at low beta,
//          and the CFL maximum.
// Face fields are stored in FP64 (the same 8 bytes as a hi + lo pair, without
// the pair's 48-bit rounding of every face each step); their increments are
// FP64 functions of FP32 edge EMFs that each edge shares with the faces
// around it, so div B stays at FP64 round-off.  The conserved state is the
// hi + lo pair (not primitive variables): the low words are only meaningful
// on the quantities that the update accumulates, and the predictor reads the
// hi words alone.
#include "bnz_common.hpp"

using namespace bnz;

namespace {

__device__ __forceinline__ void split(double v, float &hi, float &lo) {
  hi = (float)v;
  lo = (float)(v - (double)hi);
}
__device__ __forceinline__ double join(float hi, float lo) { return (double)hi + (double)lo; }

// Godunov flux of one face, FP32 (same stencil as the FP64 flux_face); ng
// counts guarded HLLD star denominators.
template <int DIR, int ORDER, int RS, typename BF>
__device__ __forceinline__ void flux32_face(const Geo &g, const float *__restrict__ w,
                                            const BF *__restrict__ bf, float *__restrict__ fl,
                                            int i, int j, int k, int &ng) {
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
      wl[n] = wm1 + 0.5f * vl_lim(wm1 - wm2, w0 - wm1);
      wr[n] = w0 - 0.5f * vl_lim(w0 - wm1, wp1 - w0);
    }
  }
  wl[BX] = bn;
  wr[BX] = bn;
  float f[NMODE];
  const float gam = (float)g.gam, gm1i = (float)g.gamm1i;
  const int reg = (RS == BNZ_RS_HLLD) ? flux_hlld(f, wl, wr, gam, gm1i, &ng)
                                      : flux_hll(f, wl, wr, gam, gm1i);
  if (reg == 0) return;
#pragma unroll
  for (int n = 0; n < NMODE; ++n) {
    if (n == BX) continue;
    fl[(long)perm<DIR>(n) * g.sv + c] = f[n];
  }
}

// All face directions of a stage in one launch, one thread per cell of the
// union of the face ranges (as the FP64 k_flux3).
template <int ORDER, int RS, typename BF>
__global__ __launch_bounds__(BLK) void k32_flux3(Geo g, const float *__restrict__ w,
                                                 const BF *__restrict__ bf,
                                                 float *__restrict__ fx, float *__restrict__ fy,
                                                 float *__restrict__ fz, Box box, Lims L,
                                                 int *nguard) {
  int i, j, k;
  if (!cell_index<false, 1, 2>(box, 0, i, j, k)) return;
  int ng = 0;
  if (i >= L.lo[0]) flux32_face<0, ORDER, RS>(g, w, bf, fx, i, j, k, ng);
  if (g.dims >= 2 && j >= L.lo[1]) flux32_face<1, ORDER, RS>(g, w, bf, fy, i, j, k, ng);
  if (g.dims >= 3 && k >= L.lo[2]) flux32_face<2, ORDER, RS>(g, w, bf, fz, i, j, k, ng);
  if (ng) atomicAdd(nguard, ng);
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
__global__ __launch_bounds__(BLK) void k32_emf(Geo g, const float *__restrict__ w,
                                               const float *__restrict__ fx,
                                               const float *__restrict__ fy,
                                               const float *__restrict__ fz,
                                               float *__restrict__ ee, Box box, Walls wl) {
  int i, j, k;
  if (!cell_index<false, 1, 2>(box, 0, i, j, k)) return;
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

// Face-field update by Stokes (ct.pyx:15-91) in working type R from face
// values b and FP32 EMFs: the lower faces of cell c (lo) and its upper faces
// (up).  Neighbour offsets ox, oy, oz are zero where the cell needs no
// neighbour (all loads stay in bounds; see the FP64 update_cell_d).
template <int DIMS, typename R, typename BF>
__device__ __forceinline__ void faces_d(const Geo &g, BF b, const float *__restrict__ ee,
                                        double dt, int i, int j, int k, long c, long ox, long oy,
                                        long oz, R *lo, R *up) {
  const long sv = g.sv;
  const float *ex = ee, *ey = ee + sv, *ez = ee + 2 * sv;
  const R ez00 = ez[c], ez10 = ez[c + ox], ez01 = ez[c + oy], ez11 = ez[c + ox + oy];
  const R ey00 = ey[c], ey10 = ey[c + ox], ey01 = ey[c + oz], ey11 = ey[c + ox + oz];
  const R ex00 = ex[c], ex10 = ex[c + oy], ex01 = ex[c + oz], ex11 = ex[c + oy + oz];
  const R bxc = b(0, c), byc = b(1, c), bzc = b(2, c);
  const R bxp = b(0, c + ox), byp = b(1, c + oy), bzp = b(2, c + oz);
  const double dlx = g.dlf[0][i], dly = g.dlf[1][j], dlz = g.dlf[2][k];
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

struct Pred32 {
  const float *w32, *uh;
  const double *b;
  const float *fx, *fy, *fz, *ee;
  float *bs, *ws;
};

// Predictor (dt/2): FP32 update of the hi words, FP64 pressure recovery.
// All operands are loaded up front; only the stores (and the copy of w into
// ws outside the limits) depend on the roles.
template <int DIMS>
__global__ __launch_bounds__(BLK) void k32_update_pred(Geo g, Pred32 a, Lims L,
                                                       const double *__restrict__ dts) {
  const long t = xcd_block() * BLK + threadIdx.x;
  const long ncell = (long)g.nt[0] * g.nt[1] * g.nt[2];
  if (t >= ncell) return;
  const int i = (int)(t % g.nt[0]);
  const long r = t / g.nt[0];
  const int j = (int)(r % g.nt[1]), k = (int)(r / g.nt[1]);
  const long c = t, sv = g.sv;
  const double dt = 0.5 * dts[0];
  const Roles<DIMS> ro(g, L, i, j, k);
  float uh[NMODE - 3], fxl[NMODE - 3], fxu[NMODE - 3], fyl[NMODE - 3], fyu[NMODE - 3],
      fzl[NMODE - 3], fzu[NMODE - 3];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    uh[n] = a.uh[n * sv + c];
    fxl[n] = a.fx[n * sv + c];
    fxu[n] = a.fx[n * sv + c + ro.ox];
    if (DIMS >= 2) {
      fyl[n] = a.fy[n * sv + c];
      fyu[n] = a.fy[n * sv + c + ro.oy];
    }
    if (DIMS >= 3) {
      fzl[n] = a.fz[n * sv + c];
      fzu[n] = a.fz[n * sv + c + ro.oz];
    }
  }
  float lo[3], up[3];
  faces_d<DIMS, float>(g, Faces32{a.b, sv}, a.ee, dt, i, j, k, c, ro.ox, ro.oy, ro.oz, lo, up);
  const float dlx = (float)g.dlf[0][i], dly = (float)g.dlf[1][j], dlz = (float)g.dlf[2][k];
  const float dtdv = (float)(dt / (g.dlf[0][i] * g.dlf[1][j] * g.dlf[2][k]));
  const float ax = dly * dlz, ay = dlx * dlz, az = dlx * dly;
  double u[NMODE], wv[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) {
    float v = uh[n] - dtdv * (ax * fxu[n] - ax * fxl[n]);
    if (DIMS >= 2) v -= dtdv * (ay * fyu[n] - ay * fyl[n]);
    if (DIMS >= 3) v -= dtdv * (az * fzu[n] - az * fzl[n]);
    u[n] = v;
  }
  u[BX] = ((float)g.icm[0][i] * lo[0] + (float)g.icp[0][i] * up[0]) * (float)g.dlfi[0][i];
  if (DIMS >= 2)
    u[BY] = ((float)g.icm[1][j] * lo[1] + (float)g.icp[1][j] * up[1]) * (float)g.dlfi[1][j];
  else
    u[BY] = lo[1];
  if (DIMS >= 3)
    u[BZ] = ((float)g.icm[2][k] * lo[2] + (float)g.icp[2][k] * up[2]) * (float)g.dlfi[2][k];
  else
    u[BZ] = lo[2];
  cons2prim(wv, u, g.gam);
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

struct Corr32 {
  float *uh, *ul;
  const double *b;
  const float *fx, *fy, *fz, *ee;
  double *b1;
  float *w32;
};

// Corrector (dt): FP64 update of the hi + lo state, FP64 pressure recovery,
// fused CFL maximum of the new state.  Operands loaded up front as in the
// predictor.
template <int DIMS>
__global__ __launch_bounds__(BLK) void k32_update_corr(Geo g, Corr32 a, Lims L,
                                                       const double *__restrict__ dts, Box act,
                                                       int fms_ref, double *__restrict__ part) {
  const long t = xcd_block() * BLK + threadIdx.x;
  const long ncell = (long)g.nt[0] * g.nt[1] * g.nt[2];
  double cdsmax = 0.;
  if (t < ncell) {
    const int i = (int)(t % g.nt[0]);
    const long r = t / g.nt[0];
    const int j = (int)(r % g.nt[1]), k = (int)(r / g.nt[1]);
    const long c = t, sv = g.sv;
    const double dt = dts[0];
    const Roles<DIMS> ro(g, L, i, j, k);
    float uh[NMODE - 3], ul[NMODE - 3], fxl[NMODE - 3], fxu[NMODE - 3], fyl[NMODE - 3],
        fyu[NMODE - 3], fzl[NMODE - 3], fzu[NMODE - 3];
#pragma unroll
    for (int n = 0; n < NMODE - 3; ++n) {
      uh[n] = a.uh[n * sv + c];
      ul[n] = a.ul[n * sv + c];
      fxl[n] = a.fx[n * sv + c];
      fxu[n] = a.fx[n * sv + c + ro.ox];
      if (DIMS >= 2) {
        fyl[n] = a.fy[n * sv + c];
        fyu[n] = a.fy[n * sv + c + ro.oy];
      }
      if (DIMS >= 3) {
        fzl[n] = a.fz[n * sv + c];
        fzu[n] = a.fz[n * sv + c + ro.oz];
      }
    }
    double lo[3], up[3];
    faces_d<DIMS, double>(g, Faces64{a.b, sv}, a.ee, dt, i, j, k, c, ro.ox, ro.oy, ro.oz, lo,
                          up);
    const double dlx = g.dlf[0][i], dly = g.dlf[1][j], dlz = g.dlf[2][k];
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
    u[BX] = (g.icm[0][i] * lo[0] + g.icp[0][i] * up[0]) * g.dlfi[0][i];
    if (DIMS >= 2)
      u[BY] = (g.icm[1][j] * lo[1] + g.icp[1][j] * up[1]) * g.dlfi[1][j];
    else
      u[BY] = lo[1];
    if (DIMS >= 3)
      u[BZ] = (g.icm[2][k] * lo[2] + g.icp[2][k] * up[2]) * g.dlfi[2][k];
    else
      u[BZ] = lo[2];
    cons2prim(wv, u, g.gam);
    if (ro.fX) a.b1[c] = lo[0];
    if (ro.fY) a.b1[sv + c] = lo[1];
    if (ro.fZ) a.b1[2 * sv + c] = lo[2];
    if (ro.cell) {
#pragma unroll
      for (int n = 0; n < NMODE - 3; ++n) {
        a.uh[n * sv + c] = nh[n];
        a.ul[n * sv + c] = nl[n];
      }
#pragma unroll
      for (int n = 0; n < NMODE; ++n) a.w32[n * sv + c] = (float)wv[n];
      if (i >= act.lo[0] && i < act.lo[0] + act.n[0] && j >= act.lo[1] &&
          j < act.lo[1] + act.n[1] && k >= act.lo[2] && k < act.lo[2] + act.n[2])
        cdsmax = cds_cell(g, wv, fms_ref, i, j, k);
    }
  }
  block_max_to(cdsmax, part);
}

// CFL maximum from the FP32 primitives (first step after an upload).
__global__ __launch_bounds__(BLK) void k32_dt_partial(Geo g, const float *__restrict__ w, Box box,
                                                      int fms_ref, double *__restrict__ part) {
  int i, j, k;
  double cds = 0.;
  if (cell_index<false, 1, 2>(box, 0, i, j, k)) {
    const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
    double wv[NMODE];
#pragma unroll
    for (int n = 0; n < NMODE; ++n) wv[n] = (double)w[n * g.sv + c];
    cds = cds_cell(g, wv, fms_ref, i, j, k);
  }
  block_max_to(cds, part);
}

struct Ghost32 {
  float *a[21];   // uh, ul, w32
  double *d[3];   // face fields
};

__global__ __launch_bounds__(BLK) void k32_ghost(Geo g, Ghost32 p, const int2 *__restrict__ pairs,
                                                 int npairs) {
  const long t = xcd_block() * BLK + threadIdx.x;
  if (t >= npairs) return;
  const int2 ds = pairs[t];
  // sources are active cells the fill never writes: load all, then store
  float v[21];
  double f[3];
#pragma unroll
  for (int n = 0; n < 21; ++n) v[n] = p.a[n][ds.y];
#pragma unroll
  for (int n = 0; n < 3; ++n) f[n] = p.d[n][ds.y];
#pragma unroll
  for (int n = 0; n < 21; ++n) p.a[n][ds.x] = v[n];
#pragma unroll
  for (int n = 0; n < 3; ++n) p.d[n][ds.x] = f[n];
}

// Ghost cells a reflecting side touches (GhostOp); the sign bit of slot v
// applies to conserved slot v of uh/ul (momentum flips with velocity) and to
// prim slot v of w32.
__global__ __launch_bounds__(BLK) void k32_ghost_ops(Geo g, Ghost32 p,
                                                     const GhostOp *__restrict__ ops, int nops) {
  const long t = xcd_block() * BLK + threadIdx.x;
  if (t >= nops) return;
  const GhostOp o = ops[t];
  const bool cells = !(o.bits & (1u << 15));
  float v[21];
  double f[3];
#pragma unroll
  for (int n = 0; n < 21; ++n) v[n] = cells ? p.a[n][o.src] : 0.f;
#pragma unroll
  for (int a = 0; a < 3; ++a) f[a] = ((o.bits >> (12 + a)) & 1u) ? 0. : p.d[a][o.srcf[a]];
  if (cells) {
#pragma unroll
    for (int n = 0; n < 21; ++n) {
      const int slot = n < 12 ? n % 6 : n - 12;
      p.a[n][o.dst] = ((o.bits >> slot) & 1u) ? -v[n] : v[n];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) p.d[a][o.dst] = ((o.bits >> (9 + a)) & 1u) ? -f[a] : f[a];
}

// FP64 prim (all cells) -> FP32-compensated state.
__global__ __launch_bounds__(BLK) void k32_from64(Geo g, const double *__restrict__ w, float *uh,
                                                  float *ul, float *w32) {
  const long t = xcd_block() * BLK + threadIdx.x;
  const long ncell = (long)g.nt[0] * g.nt[1] * g.nt[2];
  if (t >= ncell) return;
  const long sv = g.sv;
  double wv[NMODE], u[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) {
    wv[n] = w[n * sv + t];
    w32[n * sv + t] = (float)wv[n];
  }
  prim2cons(u, wv, g.gamm1i);
#pragma unroll
  for (int n = 0; n < NMODE - 3; ++n) split(u[n], uh[n * sv + t], ul[n * sv + t]);
}

// FP32-compensated state -> FP64 prim of the active cells (the caller fills
// the ghosts).
__global__ __launch_bounds__(BLK) void k32_to64(Geo g, const float *uh, const float *ul,
                                                const double *b, double *w, Box act) {
  const long t = xcd_block() * BLK + threadIdx.x;
  const long ncell = (long)g.nt[0] * g.nt[1] * g.nt[2];
  if (t >= ncell) return;
  const long sv = g.sv;
  const int i = (int)(t % g.nt[0]);
  const long r = t / g.nt[0];
  const int j = (int)(r % g.nt[1]), k = (int)(r / g.nt[1]);
  if (i < act.lo[0] || i >= act.lo[0] + act.n[0] || j < act.lo[1] || j >= act.lo[1] + act.n[1] ||
      k < act.lo[2] || k >= act.lo[2] + act.n[2])
    return;
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

template <int ORDER, int RS, typename BF>
void launch_fluxes32(bnz_ctx *c, const float *w, const BF *bf, const Lims &L) {
  int *ng = (int *)(c->dev + 4);
  const Box u = flux3_box(c->dims, L);
  Launch t(c, ORDER ? "flux3_plm" : "flux3_dc");
  hipLaunchKernelGGL((k32_flux3<ORDER, RS, BF>), dim3(blocks_for(u.total())), dim3(BLK), 0,
                     c->stream, c->g, w, bf, c->fx32, c->fy32, c->fz32, u, L, ng);
}

template <typename BF>
void fluxes32(bnz_ctx *c, int order, const float *w, const BF *bf, const Lims &L) {
  const bool hlld = c->p.rsolver == BNZ_RS_HLLD;
  if (order == 0) {
    if (hlld) launch_fluxes32<0, BNZ_RS_HLLD>(c, w, bf, L);
    else launch_fluxes32<0, BNZ_RS_HLL>(c, w, bf, L);
  } else {
    if (hlld) launch_fluxes32<1, BNZ_RS_HLLD>(c, w, bf, L);
    else launch_fluxes32<1, BNZ_RS_HLL>(c, w, bf, L);
  }
}

void emf32(bnz_ctx *c, const float *w, const Lims &L, const char *name) {
  Box e;
  e.lo[0] = L.lo[0]; e.n[0] = L.hi[0] + 1 - L.lo[0] + 1;
  e.lo[1] = (c->dims >= 2) ? L.lo[1] : 0; e.n[1] = (c->dims >= 2) ? L.hi[1] + 1 - L.lo[1] + 1 : 1;
  e.lo[2] = (c->dims >= 3) ? L.lo[2] : 0; e.n[2] = (c->dims >= 3) ? L.hi[2] + 1 - L.lo[2] + 1 : 1;
  Launch t(c, name);
  Walls wl;
  const int lo[3] = {c->i1, c->j1, c->k1}, hi[3] = {c->i2, c->j2, c->k2};
  for (int a = 0; a < 3; ++a) {
    wl.at[2 * a] = (a < c->dims && c->p.bc[2 * a] == BNZ_BC_REFLECT) ? lo[a] : -1;
    wl.at[2 * a + 1] = (a < c->dims && c->p.bc[2 * a + 1] == BNZ_BC_REFLECT) ? hi[a] + 1 : -1;
  }
#define E32_ARGS \
  dim3(blocks_for(e.total())), dim3(BLK), 0, c->stream, c->g, w, c->fx32, c->fy32, c->fz32, c->ee32, e, wl
  if (c->dims == 3) hipLaunchKernelGGL(k32_emf<3>, E32_ARGS);
  else if (c->dims == 2) hipLaunchKernelGGL(k32_emf<2>, E32_ARGS);
  else hipLaunchKernelGGL(k32_emf<1>, E32_ARGS);
#undef E32_ARGS
}

void ghost32(bnz_ctx *c) {
  Ghost32 p;
  int q = 0;
  for (int n = 0; n < 6; ++n) p.a[q++] = c->uh + n * c->ncell;
  for (int n = 0; n < 6; ++n) p.a[q++] = c->ul + n * c->ncell;
  for (int n = 0; n < 9; ++n) p.a[q++] = c->w32 + n * c->ncell;
  for (int n = 0; n < 3; ++n) p.d[n] = c->b + n * c->ncell;
  Launch t(c, "ghost");
  if (c->nghost > 0)
    hipLaunchKernelGGL(k32_ghost, dim3(blocks_for(c->nghost)), dim3(BLK), 0, c->stream, c->g, p,
                       c->ghost, c->nghost);
  if (c->ngops > 0)
    hipLaunchKernelGGL(k32_ghost_ops, dim3(blocks_for(c->ngops)), dim3(BLK), 0, c->stream, c->g,
                       p, c->gops, c->ngops);
}

}  // namespace

namespace bnz {

int f32_alloc(bnz_ctx *c) {
  const size_t n = (size_t)c->ncell * sizeof(float);
  struct {
    float **p;
    int rows;
  } bufs[] = {{&c->uh, 6},   {&c->ul, 6},   {&c->w32, 9},  {&c->ws32, 9}, {&c->bs32, 3},
              {&c->fx32, 9}, {&c->fy32, 9}, {&c->fz32, 9}, {&c->ee32, 3}};
  for (auto &q : bufs) {
    if (hipMalloc(q.p, q.rows * n) != hipSuccess)
      return fail(c, BNZ_ENOMEM, "device allocation of FP32 arrays failed");
    HIPCHK(c, hipMemsetAsync(*q.p, 0, q.rows * n, c->stream));
  }
  return BNZ_OK;
}

void f32_free(bnz_ctx *c) {
  float *bufs[] = {c->uh, c->ul, c->w32, c->ws32, c->bs32, c->fx32, c->fy32, c->fz32, c->ee32};
  for (float *q : bufs)
    if (q) (void)hipFree(q);
}

int f32_from64(bnz_ctx *c) {
  hipLaunchKernelGGL(k32_from64, dim3(blocks_for(c->ncell)), dim3(BLK), 0, c->stream, c->g, c->w,
                     c->uh, c->ul, c->w32);
  c->dt_parts = 0;
  return check_launch(c);
}

int f32_to64(bnz_ctx *c) {
  hipLaunchKernelGGL(k32_to64, dim3(blocks_for(c->ncell)), dim3(BLK), 0, c->stream, c->g, c->uh,
                     c->ul, c->b, c->w, active_box(c));
  launch_ghost64(c);
  return check_launch(c);
}

void f32_ghost(bnz_ctx *c) { ghost32(c); }

void f32_dt(bnz_ctx *c, int use_fixed, double fixed_dt) {
  int nb = c->dt_parts;
  if (nb <= 0) {
    Box a = active_box(c);
    nb = blocks_for(a.total());
    Launch t(c, "dt_partial");
    hipLaunchKernelGGL(k32_dt_partial, dim3(nb), dim3(BLK), 0, c->stream, c->g, c->w32, a,
                       c->p.dt_fms_ref, c->part);
  }
  launch_dt_final(c, nb, use_fixed, fixed_dt);
}

void f32_step_body(bnz_ctx *c) {
  const int nb = blocks_for(c->ncell);
  Lims L0 = integr_lims(c, 0);
  fluxes32(c, 0, c->w32, c->b, L0);
  emf32(c, c->w32, L0, "emf_pred");
  {
    Pred32 a{c->w32, c->uh, c->b, c->fx32, c->fy32, c->fz32, c->ee32, c->bs32, c->ws32};
    Launch t(c, "update_pred");
#define P32_ARGS dim3(nb), dim3(BLK), 0, c->stream, c->g, a, L0, c->dev
    if (c->dims == 3) hipLaunchKernelGGL(k32_update_pred<3>, P32_ARGS);
    else if (c->dims == 2) hipLaunchKernelGGL(k32_update_pred<2>, P32_ARGS);
    else hipLaunchKernelGGL(k32_update_pred<1>, P32_ARGS);
#undef P32_ARGS
  }
  Lims L1 = integr_lims(c, c->order);
  fluxes32(c, c->order, c->ws32, c->bs32, L1);
  emf32(c, c->ws32, L1, "emf_corr");
  {
    Corr32 a{c->uh, c->ul, c->b, c->fx32, c->fy32, c->fz32, c->ee32, c->bs, c->w32};
    Launch t(c, "update_corr");
#define C32_ARGS \
  dim3(nb), dim3(BLK), 0, c->stream, c->g, a, L1, c->dev, active_box(c), c->p.dt_fms_ref, c->part
    if (c->dims == 3) hipLaunchKernelGGL(k32_update_corr<3>, C32_ARGS);
    else if (c->dims == 2) hipLaunchKernelGGL(k32_update_corr<2>, C32_ARGS);
    else hipLaunchKernelGGL(k32_update_corr<1>, C32_ARGS);
#undef C32_ARGS
  }
  c->dt_parts = nb;
  std::swap(c->b, c->bs);
  ghost32(c);
  launch_time_advance(c);
}

}  // namespace bnz
