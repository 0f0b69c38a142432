// bnz_phys.hpp -- device-side ideal-MHD point physics for the VL step (FP64).
//
// Operation order follows the reference expression by expression
// (src/mhd/eos_c.h, src/mhd/fluxes_c.c, src/coordinates/reconstr_c.h) so that
// an IEEE build (-ffp-contract=off) reproduces the CPU oracle bit for bit.
// States are 9-slot register arrays in the local frame of the sweep:
//   0 rho, 1 vx|mx, 2 vy|my, 3 vz|mz, 4 p|E, 5 passive scalar, 6 Bx, 7 By, 8 Bz
// (src/defs_c.h:68-82).
#pragma once
#include <hip/hip_runtime.h>

namespace bnz {

enum { RHO = 0, MX = 1, MY = 2, MZ = 3, EN = 4, PSC = 5, BX = 6, BY = 7, BZ = 8 };
enum { VX = 1, VY = 2, VZ = 3, PR = 4 };
constexpr int NMODE = 9;

__device__ __forceinline__ double sqr(double x) { return x * x; }
// SIGN(0)=0 (src/defs_c.h:150)
__device__ __forceinline__ double sign0(double a) { return (double)((0.0 < a) - (a < 0.0)); }

// src/mhd/eos_c.h:60-95; gamm1i = 1/(gam-1) is passed in (same value the
// reference recomputes on every call).
__device__ __forceinline__ void prim2cons(double *u, const double *w, double gamm1i) {
  u[RHO] = w[RHO];
  u[MX] = w[RHO] * w[VX];
  u[MY] = w[RHO] * w[VY];
  u[MZ] = w[RHO] * w[VZ];
  double v2 = sqr(w[VX]) + sqr(w[VY]) + sqr(w[VZ]);
  double b2 = sqr(w[BX]) + sqr(w[BY]) + sqr(w[BZ]);
  u[EN] = 0.5 * (w[RHO] * v2 + b2) + w[PR] * gamm1i;
  u[PSC] = w[PSC];
  u[BX] = w[BX];
  u[BY] = w[BY];
  u[BZ] = w[BZ];
}

// src/mhd/eos_c.h:20-55
__device__ __forceinline__ void cons2prim(double *w, const double *u, double gam) {
  double rhoi = 1. / u[RHO];
  w[RHO] = u[RHO];
  w[VX] = u[MX] * rhoi;
  w[VY] = u[MY] * rhoi;
  w[VZ] = u[MZ] * rhoi;
  double v2 = sqr(w[VX]) + sqr(w[VY]) + sqr(w[VZ]);
  double b2 = sqr(u[BX]) + sqr(u[BY]) + sqr(u[BZ]);
  w[PR] = (gam - 1.) * (u[EN] - 0.5 * (b2 + v2 * u[RHO]));
  w[PSC] = u[PSC];
  w[BX] = u[BX];
  w[BY] = u[BY];
  w[BZ] = u[BZ];
}

// src/mhd/eos_c.h:416-461 with a=1 (no CGL/TWOTEMP)
__device__ __forceinline__ void primcons2flux(double *f, const double *w, const double *u) {
  double bx2 = sqr(w[BX]);
  double b2 = bx2 + sqr(w[BY]) + sqr(w[BZ]);
  double bxby = w[BX] * w[BY];
  double bxbz = w[BX] * w[BZ];
  double ptot = 0.;
  ptot += 0.5 * b2;
  ptot += w[PR];
#pragma unroll
  for (int n = 0; n < BX; ++n) f[n] = u[n] * w[VX];
  f[MX] += ptot;
  f[EN] += ptot * w[VX];
  f[MX] -= 1. * bx2;
  f[MY] -= 1. * bxby;
  f[MZ] -= 1. * bxbz;
  f[EN] -= 1. * (w[VX] * bx2 + w[VY] * bxby + w[VZ] * bxbz);
  f[BX] = 0.;
  f[BY] = w[VX] * w[BY] - w[VY] * w[BX];
  f[BZ] = w[VX] * w[BZ] - w[VZ] * w[BX];
}

// Fast magnetosonic speed, src/mhd/eos_c.h:468-530.  Reads w[BX] (the
// reference ignores its bx argument).
__device__ __forceinline__ double fms(const double *w, double gam) {
  double bx2 = sqr(w[BX]);
  double b2 = bx2 + sqr(w[BY]) + sqr(w[BZ]);
  double gamp = gam * w[PR];
  double c1 = gamp + b2;
  double c2 = sqrt(fabs(c1 * c1 - 4. * gamp * bx2));
  return sqrt(0.5 / w[RHO] * (c1 + c2));
}

// van Leer limiter, src/coordinates/reconstr_c.h:39-44
__device__ __forceinline__ double vl_lim(double a, double b) {
  double ab = a * b;
  return (ab > 0.) ? (2. * ab / (a + b)) : 0.;
}

// sLsR_simple, src/mhd/fluxes_c.c:25-42 (FMAX/FMIN as fmax/fmin)
__device__ __forceinline__ void speeds_simple(double &sL, double &sR, const double *wL,
                                              const double *wR, double gam) {
  double cfL = fms(wL, gam);
  double cfR = fms(wR, gam);
  double cf_max = fmax(cfL, cfR);
  sL = fmin(wL[VX], wR[VX]) - cf_max;
  sR = fmax(wL[VX], wR[VX]) + cf_max;
}

// HLL flux, src/mhd/fluxes_c.c:175-223.  Always writes (returns 1).
__device__ __forceinline__ int flux_hll(double *flux, const double *wL, const double *wR,
                                       double gam, double gamm1i) {
  double uL[NMODE], uR[NMODE], fL[NMODE], fR[NMODE];
  prim2cons(uL, wL, gamm1i);
  prim2cons(uR, wR, gamm1i);
  primcons2flux(fL, wL, uL);
  primcons2flux(fR, wR, uR);
  double sL, sR;
  speeds_simple(sL, sR, wL, wR, gam);
  double aL = fmin(sL, 0.);
  double aR = fmax(sR, 0.);
  double aLaR = aL * aR;
  double aRmaLi = 1. / (aR - aL);
#pragma unroll
  for (int n = 0; n < NMODE; ++n)
    flux[n] = (aR * fL[n] - aL * fR[n] + aLaR * (uR[n] - uL[n])) * aRmaLi;
  flux[BX] = 0.;
  return 1;
}

// HLLD flux, src/mhd/fluxes_c.c:616-834.  Returns the region (1..6) chosen
// by the reference's chain of independent ifs (last true wins), or 0 when
// no condition holds (only with NaN speeds): the caller then keeps the old
// flux, as the reference leaves the flux row untouched.
__device__ __forceinline__ int flux_hlld(double *flux, const double *wL, const double *wR,
                                        double gam, double gamm1i) {
  const double bx = wL[BX];
  double uL[NMODE], uR[NMODE], fL[NMODE], fR[NMODE];
  prim2cons(uL, wL, gamm1i);
  prim2cons(uR, wR, gamm1i);
  primcons2flux(fL, wL, uL);
  primcons2flux(fR, wR, uR);
  double sL, sR;
  speeds_simple(sL, sR, wL, wR, gam);

  double b2hL = 0.5 * (sqr(bx) + sqr(wL[BY]) + sqr(wL[BZ]));
  double b2hR = 0.5 * (sqr(bx) + sqr(wR[BY]) + sqr(wR[BZ]));
  double ptL = wL[PR] + b2hL;
  double ptR = wR[PR] + b2hR;

  double deni = 1. / (sR * uR[RHO] - sL * uL[RHO] - (fR[RHO] - fL[RHO]));
  double sM = (sR * uR[MX] - sL * uL[MX] - (fR[MX] - fL[MX])) * deni;

  double sLvxL = sL - wL[VX];
  double sRvxR = sR - wR[VX];
  double sLsMi = 1. / (sL - sM);
  double sRsMi = 1. / (sR - sM);
  double xL = sLvxL * sLsMi;
  double xR = sRvxR * sRsMi;

  double uLs[NMODE], uRs[NMODE];
  uLs[RHO] = uL[RHO] * xL;
  uRs[RHO] = uR[RHO] * xR;
  uLs[PSC] = uL[PSC] * xL;
  uRs[PSC] = uR[PSC] * xR;

  double pts = wL[RHO] * sLvxL * (sM - wL[VX]) + ptL;

  double rhosLvxL = wL[RHO] * sLvxL;
  double rhosRvxR = wR[RHO] * sRvxR;
  double bx2 = sqr(bx);
  double wvbL = 1. / (rhosLvxL * (sL - sM) - bx2);
  double wvbR = 1. / (rhosRvxR * (sR - sM) - bx2);
  double wvL = (bx * (sM - wL[VX])) * wvbL;
  double wvR = (bx * (sM - wR[VX])) * wvbR;
  double vyLs = wL[VY] - wL[BY] * wvL;
  double vyRs = wR[VY] - wR[BY] * wvR;
  double vzLs = wL[VZ] - wL[BZ] * wvL;
  double vzRs = wR[VZ] - wR[BZ] * wvR;

  uLs[MX] = uLs[RHO] * sM;
  uRs[MX] = uRs[RHO] * sM;
  uLs[MY] = uLs[RHO] * vyLs;
  uRs[MY] = uRs[RHO] * vyRs;
  uLs[MZ] = uLs[RHO] * vzLs;
  uRs[MZ] = uRs[RHO] * vzRs;

  double wbL = (rhosLvxL * sLvxL - bx2) * wvbL;
  double wbR = (rhosRvxR * sRvxR - bx2) * wvbR;
  uLs[BY] = wL[BY] * wbL;
  uRs[BY] = wR[BY] * wbR;
  uLs[BZ] = wL[BZ] * wbL;
  uRs[BZ] = wR[BZ] * wbR;

  double vsbsL = sM * bx + vyLs * uLs[BY] + vzLs * uLs[BZ];
  double vsbsR = sM * bx + vyRs * uRs[BY] + vzRs * uRs[BZ];
  uLs[EN] = (sLvxL * uL[EN] - ptL * wL[VX] + pts * sM +
             bx * (wL[VX] * bx + wL[VY] * wL[BY] + wL[VZ] * wL[BZ] - vsbsL)) * sLsMi;
  uRs[EN] = (sRvxR * uR[EN] - ptR * wR[VX] + pts * sM +
             bx * (wR[VX] * bx + wR[VY] * wR[BY] + wR[VZ] * wR[BZ] - vsbsR)) * sRsMi;

  double sqrt_rhoLs = sqrt(uLs[RHO]);
  double sqrt_rhoRs = sqrt(uRs[RHO]);
  double sLs = sM - fabs(bx) / sqrt_rhoLs;
  double sRs = sM + fabs(bx) / sqrt_rhoRs;
  double rhoLRi = 1. / (sqrt_rhoLs + sqrt_rhoRs);
  double sgnb = sign0(bx);

  double vyss = (sqrt_rhoLs * vyLs + sqrt_rhoRs * vyRs + (uRs[BY] - uLs[BY]) * sgnb) * rhoLRi;
  double vzss = (sqrt_rhoLs * vzLs + sqrt_rhoRs * vzRs + (uRs[BZ] - uLs[BZ]) * sgnb) * rhoLRi;
  double byss = (sqrt_rhoLs * uRs[BY] + sqrt_rhoRs * uLs[BY] +
                 sqrt_rhoLs * sqrt_rhoRs * (vyRs - vyLs) * sgnb) * rhoLRi;
  double bzss = (sqrt_rhoLs * uRs[BZ] + sqrt_rhoRs * uLs[BZ] +
                 sqrt_rhoLs * sqrt_rhoRs * (vzRs - vzLs) * sgnb) * rhoLRi;
  double vssbss = sM * bx + vyss * byss + vzss * bzss;

  int reg = 0;
  if (sL > 0.) reg = 1;
  if (sL <= 0. && 0. < sLs) reg = 2;
  if (sLs <= 0. && 0. < sM) reg = 3;
  if (sM <= 0. && 0. < sRs) reg = 4;
  if (sRs <= 0. && 0. < sR) reg = 5;
  if (sR <= 0.) reg = 6;
  if (reg == 0) return 0;

  // Evaluate only the chosen side; each region's expression keeps the
  // reference's association: ((f + s*(u* - u)) + s*(u** - u*)).
  const bool left = reg <= 3;
  const double *f = left ? fL : fR;
  const double *u = left ? uL : uR;
  const double *us = left ? uLs : uRs;
  const double s1 = left ? sL : sR;
  const double s2 = left ? sLs : sRs;
  const double sq = left ? sqrt_rhoLs : sqrt_rhoRs;
  const double vyS = left ? vyLs : vyRs, vzS = left ? vzLs : vzRs;
  const double vsbs = left ? vsbsL : vsbsR;
  const double rhoS = us[RHO];
  double uss[NMODE];
  uss[RHO] = rhoS;
  uss[MX] = us[MX];
  uss[PSC] = us[PSC];
  uss[MY] = rhoS * vyss;
  uss[MZ] = rhoS * vzss;
  uss[BY] = byss;
  uss[BZ] = bzss;
  uss[EN] = left ? (us[EN] - sq * (vsbs - vssbss) * sgnb) : (us[EN] + sq * (vsbs - vssbss) * sgnb);
  (void)vyS;
  (void)vzS;
  const int lvl = left ? (reg == 1 ? 0 : reg == 2 ? 1 : 2) : (reg == 6 ? 0 : reg == 5 ? 1 : 2);
#pragma unroll
  for (int n = 0; n < NMODE; ++n) {
    if (n == BX) continue;
    double v = f[n];
    if (lvl >= 1) v = f[n] + s1 * (us[n] - u[n]);
    if (lvl >= 2) v = v + s2 * (uss[n] - us[n]);
    flux[n] = v;
  }
  flux[BX] = 0.;
  return reg;
}

}  // namespace bnz
