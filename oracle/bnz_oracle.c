/*
 * bnz_oracle.c -- CPU oracle for the Van-Leer MHD step (TEST INFRASTRUCTURE).
 *
 * The checker for the HIP path, never the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Restates sergekomarov/bonzo's algorithm with the reference's operation
 * order, so that an IEEE build (no contraction, no fast-math) is reproducible
 * bit for bit by any other implementation that keeps the same order.
 * Decisions where the shipped reference cannot be followed literally:
 *   D1  swap_array_ptrs is a no-op (src/util.pyx:204-208); the y/z sweeps
 *       here use the double-buffered left state the code evidently intends
 *       (src/integration/godunov.pyx:141-160,185-204).
 *   D2  fms() ignores its bx argument (src/mhd/eos_c.h:474-476), so new_dt
 *       uses the x fast speed for all axes (new_dt.pyx:76-91).  Kept
 *       (dt_fms_ref=1, default); dt_fms_ref=0 gives per-axis fast speeds.
 *   D11 advance_b's By loop lacks the +1 face (src/integration/ct.pyx:47-57).
 *       The never-written face feeds interp_bc and, through cons2prim's
 *       magnetic pressure, the PLM state of the last active y face, so the
 *       face is updated here like Bx and Bz (ct.pyx:26,75).
 *   D14 z1 outflow passes k1 instead of k1-1 (grid_bc_funcs.pyx:179 vs
 *       :121,:150) and overwrites the first active plane; the intent (k1-1)
 *       is followed.
 *   FMAX/FMIN are one-argument macros (src/defs_c.h:142-143) called with two
 *       arguments (src/mhd/fluxes_c.c:38-40,208-209); fmax/fmin are used.
 *   2D: advance_b's Bz loop runs k in [0,1] with Ntot_z=1 (ct.pyx:75), an
 *       out-of-bounds write; only k=0 is updated here.
 *   D15 reflect boundaries follow the evident intent of grid_bc_funcs.pyx
 *       (see bc_axis).
 *
 * Parity unpinned: the energy flux of or_hllt_flux in the Alfven-fan regions
 * 2-4.  hllt_flux never sets the star-state energy (and normal-field) slots
 * (fluxes_c.c:504-573), so no closed form, invariant or golden vector fixes
 * what they should be; this restatement uses 0 there, as the device does.
 * Every other function is pinned by tests/test_oracle_pins.py.
 */
#include "bnz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Variable slots, src/defs_c.h:68-82 with MFIELD, NMODE=9. */
enum { RHO = 0, MX = 1, MY = 2, MZ = 3, EN = 4, PSC = 5, BX = 6, BY = 7, BZ = 8 };
enum { VX = 1, VY = 2, VZ = 3, PR = 4 };
#define NMODE OR_NMODE

static inline double sqr(double x) { return x * x; }
/* SIGN(0)=0, src/defs_c.h:150 */
static inline double sign0(double a) { return (double)((0.0 < a) - (a < 0.0)); }

/* ------------------------------------------------------------------------ */
/* EOS point functions                                                       */
/* ------------------------------------------------------------------------ */

/* src/mhd/eos_c.h:20-55 */
void or_cons2prim(double *w, const double *u, double gam)
{
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

/* src/mhd/eos_c.h:60-95 */
void or_prim2cons(double *u, const double *w, double gam)
{
  double gamm1i = 1. / (gam - 1.);
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

/* src/mhd/eos_c.h:416-461 (MFIELD, no CGL/TWOTEMP: a=1) */
void or_primcons2flux(double *f, const double *w, const double *u, double gam)
{
  (void)gam;
  double bx2 = sqr(w[BX]);
  double b2 = bx2 + sqr(w[BY]) + sqr(w[BZ]);
  double bxby = w[BX] * w[BY];
  double bxbz = w[BX] * w[BZ];
  double ptot = 0.;
  ptot += 0.5 * b2;
  double a = 1.;
  ptot += w[PR];
  for (int n = 0; n < BX; ++n) f[n] = u[n] * w[VX];
  f[MX] += ptot;
  f[EN] += ptot * w[VX];
  f[MX] -= a * bx2;
  f[MY] -= a * bxby;
  f[MZ] -= a * bxbz;
  f[EN] -= a * (w[VX] * bx2 + w[VY] * bxby + w[VZ] * bxbz);
  f[BX] = 0.;
  f[BY] = w[VX] * w[BY] - w[VY] * w[BX];
  f[BZ] = w[VX] * w[BZ] - w[VZ] * w[BX];
}

/* src/mhd/eos_c.h:468-530 (MFIELD branch).  bx is ignored as in the
   reference (D2): the x component w[BX] is always used. */
double or_fms(const double *w, double bx, double gam)
{
  (void)bx;
  double bx2 = sqr(w[BX]);
  double b2 = bx2 + sqr(w[BY]) + sqr(w[BZ]);
  double p = w[PR];
  double gamp = gam * p;
  double c1 = gamp + b2;
  double c2 = sqrt(fabs(c1 * c1 - 4. * gamp * bx2));
  return sqrt(0.5 / w[RHO] * (c1 + c2));
}

/* Sound speed of the hydro build (MFIELD=0): src/mhd/eos_c.h:468-530, the
   branch without field, SQRT(gamp / w[RHO]). */
double or_fms_hydro(const double *w, double gam)
{
  double p = w[PR];
  double gamp = gam * p;
  return sqrt(gamp / w[RHO]);
}

/* src/coordinates/reconstr_c.h:39-44 */
double or_vl_lim(double a, double b)
{
  double ab = a * b;
  return (ab > 0.) ? (2. * ab / (a + b)) : 0.;
}

/* ------------------------------------------------------------------------ */
/* Riemann solvers, pencil ABI                                               */
/* ------------------------------------------------------------------------ */

/* src/mhd/fluxes_c.c:25-42 (FMAX/FMIN read as fmax/fmin) */
static inline void sLsR_simple(double *sL, double *sR, const double *wL, const double *wR,
                               double gam)
{
  double bx = wL[BX];
  double cfL = or_fms(wL, bx, gam);
  double cfR = or_fms(wR, bx, gam);
  double cf_max = fmax(cfL, cfR);
  *sL = fmin(wL[VX], wR[VX]) - cf_max;
  *sR = fmax(wL[VX], wR[VX]) + cf_max;
}

/* sLsR_simple of the hydro build: the same estimate with the sound speed
   (src/mhd/fluxes_c.c:25-42, bx = 0 without MFIELD) */
static inline void sLsR_simple_hydro(double *sL, double *sR, const double *wL, const double *wR,
                                     double gam)
{
  double cfL = or_fms_hydro(wL, gam);
  double cfR = or_fms_hydro(wR, gam);
  double cf_max = fmax(cfL, cfR);
  *sL = fmin(wL[VX], wR[VX]) - cf_max;
  *sR = fmax(wL[VX], wR[VX]) + cf_max;
}

/* Toro's pressure-based wave speeds (PVRS), src/mhd/fluxes_c.c:138-168,
   hydro build (fms is the sound speed; no TWOTEMP) */
static inline void sLsR_pressure(double *sL, double *sR, const double *wL, const double *wR,
                                 double gam)
{
  double rhoL, rhoR, vxL, vxR;
  double aL, aR, ptL, ptR, qL, qR, pvrs;
  double g1gh = 0.5 * (gam + 1.) / gam;
  rhoL = wL[RHO];
  rhoR = wR[RHO];
  vxL = wL[VX];
  vxR = wR[VX];
  ptL = wL[PR];
  ptR = wR[PR];
  aL = or_fms_hydro(wL, gam);
  aR = or_fms_hydro(wR, gam);
  pvrs = fmax(0., 0.5 * (ptL + ptR) - 0.125 * (vxR - vxL) * (rhoL + rhoR) * (aL + aR));
  if (pvrs < ptL) qL = 1.;
  else qL = sqrt(1. + g1gh * (pvrs / ptL - 1.));
  if (pvrs < ptR) qR = 1.;
  else qR = sqrt(1. + g1gh * (pvrs / ptR - 1.));
  *sL = vxL - aL * qL;
  *sR = vxR + aR * qR;
}

static inline void load_face(double *wL, double *wR, double **_wL, double **_wR,
                             const double *_bx, int i)
{
  for (int n = 0; n < NMODE; ++n) {
    wL[n] = _wL[n][i];
    wR[n] = _wR[n][i];
  }
  wL[BX] = _bx[i];
  wR[BX] = _bx[i];
}

/* src/mhd/fluxes_c.c:175-223 */
void or_hll_flux(double **_flux, double **_wL, double **_wR, double *_bx, int i1, int i2,
                 double gam)
{
  double uL[NMODE], uR[NMODE], wL[NMODE], wR[NMODE], fL[NMODE], fR[NMODE];
  for (int i = i1; i <= i2; ++i) {
    load_face(wL, wR, _wL, _wR, _bx, i);
    or_prim2cons(uL, wL, gam);
    or_prim2cons(uR, wR, gam);
    or_primcons2flux(fL, wL, uL, gam);
    or_primcons2flux(fR, wR, uR, gam);
    double sL = 0., sR = 0.;
    sLsR_simple(&sL, &sR, wL, wR, gam);
    double aL = fmin(sL, 0.);
    double aR = fmax(sR, 0.);
    double aLaR = aL * aR;
    double aRmaLi = 1. / (aR - aL);
    for (int n = 0; n < NMODE; ++n)
      _flux[n][i] = (aR * fL[n] - aL * fR[n] + aLaR * (uR[n] - uL[n])) * aRmaLi;
    _flux[BX][i] = 0.;
  }
}

/* hll_flux of the hydro build (src/mhd/fluxes_c.c:175-223 without MFIELD):
   the sound speed in the wave estimate.  Rows are the 9-slot MHD layout
   with zero field (prim2cons, primcons2flux and their B terms then reduce
   to the hydro forms exactly); the B slots of the flux come out zero. */
void or_hll_flux_hydro(double **_flux, double **_wL, double **_wR, double *_bx, int i1,
                       int i2, double gam)
{
  double uL[NMODE], uR[NMODE], wL[NMODE], wR[NMODE], fL[NMODE], fR[NMODE];
  for (int i = i1; i <= i2; ++i) {
    load_face(wL, wR, _wL, _wR, _bx, i);
    or_prim2cons(uL, wL, gam);
    or_prim2cons(uR, wR, gam);
    or_primcons2flux(fL, wL, uL, gam);
    or_primcons2flux(fR, wR, uR, gam);
    double sL = 0., sR = 0.;
    sLsR_simple_hydro(&sL, &sR, wL, wR, gam);
    double aL = fmin(sL, 0.);
    double aR = fmax(sR, 0.);
    double aLaR = aL * aR;
    double aRmaLi = 1. / (aR - aL);
    for (int n = 0; n < NMODE; ++n)
      _flux[n][i] = (aR * fL[n] - aL * fR[n] + aLaR * (uR[n] - uL[n])) * aRmaLi;
    _flux[BX][i] = 0.;
  }
}

/* hllc_flux, src/mhd/fluxes_c.c:230-327 (hydro build, no TWOTEMP): Toro's
   HLLC in the form F* = z (F - s U) + s/(s - sM) p* (0, 1, 0, 0, sM), with
   z = -sM/(s - sM) on the side sM points away from and the outer speeds
   clipped at 0 so the supersonic cases take F_L or F_R.  9-slot rows with
   zero field, as for or_hll_flux_hydro. */
void or_hllc_flux(double **_flux, double **_wL, double **_wR, double *_bx, int i1, int i2,
                  double gam)
{
  double wL[NMODE], wR[NMODE], uL[NMODE], uR[NMODE], fL[NMODE], fR[NMODE];
  for (int i = i1; i <= i2; ++i) {
    load_face(wL, wR, _wL, _wR, _bx, i);
    double sL = 0., sR = 0.;
    or_prim2cons(uL, wL, gam);
    or_prim2cons(uR, wR, gam);
    sLsR_pressure(&sL, &sR, wL, wR, gam);
    double ptL = wL[PR];
    double ptR = wR[PR];
    double vLsL = wL[VX] - sL;
    double vRsR = wR[VX] - sR;
    double xL = uL[RHO] * vLsL;
    double xR = uR[RHO] * vRsR;
    double xRxLi = 1. / (xR - xL);
    double yL = ptL + uL[MX] * vLsL;
    double yR = ptR + uR[MX] * vRsR;
    double sM = xRxLi * (yR - yL);
    double pts = xRxLi * (xR * yL - xL * yR);
    sL = fmin(sL, 0.);
    sR = fmax(sR, 0.);
    vLsL = wL[VX] - sL;
    vRsR = wR[VX] - sR;
    for (int n = 0; n < NMODE; ++n) {
      fL[n] = uL[n] * vLsL;
      fR[n] = uR[n] * vRsR;
    }
    fL[MX] += ptL;
    fR[MX] += ptR;
    fL[EN] += ptL * wL[VX];
    fR[EN] += ptR * wR[VX];
    double sLsMi = 1. / (sL - sM);
    double sRsMi = 1. / (sR - sM);
    double zL, zR, sLR;
    if (sM > 0.) {
      zL = -sM * sLsMi;
      zR = 0.;
      sLR = sL * sLsMi;
    } else {
      zL = 0.;
      zR = -sM * sRsMi;
      sLR = sR * sRsMi;
    }
    for (int n = 0; n < NMODE; n++) _flux[n][i] = zL * fL[n] + zR * fR[n];
    _flux[MX][i] += sLR * pts;
    _flux[EN][i] += sLR * pts * sM;
    _flux[BX][i] = 0.;
  }
}

/* src/mhd/fluxes_c.c:447-607 (MFIELD, no TWOTEMP) */
void or_hllt_flux(double **_flux, double **_wL, double **_wR, double *_bx, int i1, int i2,
                  double gam)
{
  double uL[NMODE], uR[NMODE], uLs[NMODE], uRs[NMODE], uCs[NMODE];
  double wL[NMODE], wR[NMODE], fL[NMODE], fR[NMODE];
  for (int i = i1; i <= i2; ++i) {
    load_face(wL, wR, _wL, _wR, _bx, i);
    double bx = _bx[i];
    or_prim2cons(uL, wL, gam);
    or_prim2cons(uR, wR, gam);
    or_primcons2flux(fL, wL, uL, gam);
    or_primcons2flux(fR, wR, uR, gam);
    double sL = 0., sR = 0.;
    sLsR_simple(&sL, &sR, wL, wR, gam);
    double sRsLi = 1. / (sR - sL);
    double rhos = (sR * uR[RHO] - sL * uL[RHO] - (fR[RHO] - fL[RHO])) * sRsLi;
    double mxs = (sR * uR[MX] - sL * uL[MX] - (fR[MX] - fL[MX])) * sRsLi;
    double frhos = (sR * fL[RHO] - sL * fR[RHO] + sR * sL * (uR[RHO] - uL[RHO])) * sRsLi;
    double us = frhos / rhos;
    double pscs = (sR * uR[PSC] - sL * uL[PSC] - (fR[PSC] - fL[PSC])) * sRsLi;
    uLs[RHO] = rhos;  uRs[RHO] = rhos;
    uLs[MX] = mxs;    uRs[MX] = mxs;
    uLs[PSC] = pscs;  uRs[PSC] = pscs;
    uLs[MY] = rhos * wL[VY];
    uRs[MY] = rhos * wR[VY];
    uLs[MZ] = rhos * wL[VZ];
    uRs[MZ] = rhos * wR[VZ];
    double sLs = us, sRs = us;
    double rhosi = 1. / rhos;
    double sqrt_rhosi = sqrt(rhosi);
    sLs -= fabs(bx) * sqrt_rhosi;
    sRs += fabs(bx) * sqrt_rhosi;
    double xL = 1. / ((sL - sLs) * (sL - sRs));
    double xR = 1. / ((sR - sLs) * (sR - sRs));
    double yL = bx * xL * (us - wL[VX]);
    double yR = bx * xR * (us - wR[VX]);
    uLs[MY] -= yL * wL[BY];
    uRs[MY] -= yR * wR[BY];
    uLs[MZ] -= yL * wL[BZ];
    uRs[MZ] -= yR * wR[BZ];
    yL = (wL[RHO] * sqr(sL - wL[VX]) - sqr(bx)) * rhosi * xL;
    yR = (wR[RHO] * sqr(sR - wR[VX]) - sqr(bx)) * rhosi * xR;
    uLs[BY] = yL * wL[BY];
    uRs[BY] = yR * wR[BY];
    uLs[BZ] = yL * wL[BZ];
    uRs[BZ] = yR * wR[BZ];
    uCs[RHO] = rhos;
    uCs[MX] = mxs;
    uCs[PSC] = pscs;
    double sgnb = sign0(bx);
    double x = 0.5 * sgnb / sqrt_rhosi;
    uCs[MY] = 0.5 * (uLs[MY] + uRs[MY]) + x * (uRs[BY] - uLs[BY]);
    uCs[MZ] = 0.5 * (uLs[MZ] + uRs[MZ]) + x * (uRs[BZ] - uLs[BZ]);
    uCs[BY] = 0.5 * (uLs[BY] + uRs[BY]) + x * rhosi * (uRs[MY] - uLs[MY]);
    uCs[BZ] = 0.5 * (uLs[BZ] + uRs[BZ]) + x * rhosi * (uRs[MZ] - uLs[MZ]);
    /* The reference never sets the energy and Bx slots of the star states
       (fluxes_c.c:504-573), so its EN flux in regions 2-4 reads
       uninitialised stack memory.  Zero is used for those slots here; the
       energy flux of HLLT is therefore only defined in regions 1 and 5. */
    uLs[EN] = uRs[EN] = uCs[EN] = 0.;
    uLs[BX] = uRs[BX] = uCs[BX] = 0.;
    int reg = 0;
    if (sL > 0.) reg = 1;
    if (sL <= 0. && 0. < sLs) reg = 2;
    if (sLs <= 0. && 0. < sRs) reg = 3;
    if (sRs <= 0. && 0. < sR) reg = 4;
    if (sR <= 0.) reg = 5;
    for (int n = 0; n < NMODE; ++n) {
      if (reg == 1) _flux[n][i] = fL[n];
      else if (reg == 2) _flux[n][i] = fL[n] + sL * (uLs[n] - uL[n]);
      else if (reg == 3)
        _flux[n][i] = fL[n] + sL * (uLs[n] - uL[n]) + sLs * (uCs[n] - uLs[n]);
      else if (reg == 4) _flux[n][i] = fR[n] + sR * (uRs[n] - uR[n]);
      else if (reg == 5) _flux[n][i] = fR[n];
    }
    _flux[BX][i] = 0.;
  }
}

/* src/mhd/fluxes_c.c:616-834 (no TWOTEMP).  The region chain is a
   sequence of independent ifs, so the last true condition wins and, if none
   holds (only with NaN speeds), the flux row keeps its previous content. */
void or_hlld_flux(double **_flux, double **_wL, double **_wR, double *_bx, int i1, int i2,
                  double gam)
{
  double uL[NMODE], uR[NMODE], uLs[NMODE], uRs[NMODE], uLss[NMODE], uRss[NMODE];
  double wL[NMODE], wR[NMODE], fL[NMODE], fR[NMODE];
  for (int i = i1; i <= i2; ++i) {
    load_face(wL, wR, _wL, _wR, _bx, i);
    double bx = _bx[i];
    or_prim2cons(uL, wL, gam);
    or_prim2cons(uR, wR, gam);
    or_primcons2flux(fL, wL, uL, gam);
    or_primcons2flux(fR, wR, uR, gam);
    double sL = 0., sR = 0.;
    sLsR_simple(&sL, &sR, wL, wR, gam);

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
    double bxsMvxL = bx * (sM - wL[VX]);
    double bxsMvxR = bx * (sM - wR[VX]);
    double wvL = bxsMvxL * wvbL;
    double wvR = bxsMvxR * wvbR;
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

    double rhosLvxL2_bx2 = rhosLvxL * sLvxL - bx2;
    double rhosRvxR2_bx2 = rhosRvxR * sRvxR - bx2;
    double wbL = rhosLvxL2_bx2 * wvbL;
    double wbR = rhosRvxR2_bx2 * wvbR;
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

    uLss[RHO] = uLs[RHO];
    uRss[RHO] = uRs[RHO];
    uLss[MX] = uLs[MX];
    uRss[MX] = uRs[MX];
    uLss[PSC] = uLs[PSC];
    uRss[PSC] = uRs[PSC];

    double sqrt_rhoLs = sqrt(uLs[RHO]);
    double sqrt_rhoRs = sqrt(uRs[RHO]);
    double sLs = sM - fabs(bx) / sqrt_rhoLs;
    double sRs = sM + fabs(bx) / sqrt_rhoRs;
    double rhoLRi = 1. / (sqrt_rhoLs + sqrt_rhoRs);
    double sgnb = sign0(bx);

    double vyss = (sqrt_rhoLs * vyLs + sqrt_rhoRs * vyRs + (uRs[BY] - uLs[BY]) * sgnb) * rhoLRi;
    double vzss = (sqrt_rhoLs * vzLs + sqrt_rhoRs * vzRs + (uRs[BZ] - uLs[BZ]) * sgnb) * rhoLRi;
    uLss[MY] = uLs[RHO] * vyss;
    uRss[MY] = uRs[RHO] * vyss;
    uLss[MZ] = uLs[RHO] * vzss;
    uRss[MZ] = uRs[RHO] * vzss;

    double byss = (sqrt_rhoLs * uRs[BY] + sqrt_rhoRs * uLs[BY] +
                   sqrt_rhoLs * sqrt_rhoRs * (vyRs - vyLs) * sgnb) * rhoLRi;
    double bzss = (sqrt_rhoLs * uRs[BZ] + sqrt_rhoRs * uLs[BZ] +
                   sqrt_rhoLs * sqrt_rhoRs * (vzRs - vzLs) * sgnb) * rhoLRi;
    uLss[BY] = byss;
    uRss[BY] = byss;
    uLss[BZ] = bzss;
    uRss[BZ] = bzss;

    double vssbss = sM * bx + vyss * byss + vzss * bzss;
    uLss[EN] = uLs[EN] - sqrt_rhoLs * (vsbsL - vssbss) * sgnb;
    uRss[EN] = uRs[EN] + sqrt_rhoRs * (vsbsR - vssbss) * sgnb;

    /* Bx slot of the star states is never set in the reference; its flux is
       overwritten with 0 below, so any value will do. */
    uLs[BX] = uRs[BX] = uLss[BX] = uRss[BX] = bx;

    int reg = 0;
    if (sL > 0.) reg = 1;
    if (sL <= 0. && 0. < sLs) reg = 2;
    if (sLs <= 0. && 0. < sM) reg = 3;
    if (sM <= 0. && 0. < sRs) reg = 4;
    if (sRs <= 0. && 0. < sR) reg = 5;
    if (sR <= 0.) reg = 6;
    for (int n = 0; n < NMODE; ++n) {
      if (reg == 1) _flux[n][i] = fL[n];
      else if (reg == 2) _flux[n][i] = fL[n] + sL * (uLs[n] - uL[n]);
      else if (reg == 3)
        _flux[n][i] = fL[n] + sL * (uLs[n] - uL[n]) + sLs * (uLss[n] - uLs[n]);
      else if (reg == 4)
        _flux[n][i] = fR[n] + sR * (uRs[n] - uR[n]) + sRs * (uRss[n] - uRs[n]);
      else if (reg == 5) _flux[n][i] = fR[n] + sR * (uRs[n] - uR[n]);
      else if (reg == 6) _flux[n][i] = fR[n];
    }
    _flux[BX][i] = 0.;
  }
}

/* ------------------------------------------------------------------------ */
/* Characteristic projection (MFIELD branch of src/mhd/eos_c.h:100-280)      */
/* ------------------------------------------------------------------------ */

/* Wave speeds and eigenvector weights of the state w, shared by prim2char
   (eos_c.h:100-161) and char2prim (eos_c.h:222-262). */
typedef struct {
  double rhoi, s, a2rho, a, cf, cs, alphaf, alphas, betay, betaz;
} or_charw;

static void char_weights(or_charw *c, const double *w, double gam)
{
  double w_rho = w[RHO], w_p = w[PR], w_bx = w[BX], w_by = w[BY], w_bz = w[BZ];
  c->rhoi = 1. / w_rho;
  c->s = sign0(w_bx);
  c->a2rho = gam * w_p;
  c->a = sqrt(c->a2rho * c->rhoi);
  double bx2 = sqr(w_bx);
  double bperp2 = sqr(w_by) + sqr(w_bz);
  double b2 = bx2 + bperp2;
  double x = c->a2rho + b2;
  double y = sqrt(fabs(x * x - 4. * c->a2rho * bx2));
  double cf2rho = 0.5 * (x + y);
  double cs2rho = 0.5 * (x - y);
  c->cf = sqrt(c->rhoi * cf2rho);
  c->cs = sqrt(c->rhoi * fabs(cs2rho));
  double z = 1. / (cf2rho - cs2rho);
  if (c->cf <= c->cs) {
    c->alphaf = 1.;
    c->alphas = 0.;
  } else if (c->a <= c->cs) {
    c->alphaf = 0.;
    c->alphas = 1.;
  } else if (c->cf <= c->a) {
    c->alphaf = 1.;
    c->alphas = 0.;
  } else {
    c->alphaf = sqrt(z * fabs(c->a2rho - cs2rho));
    c->alphas = sqrt(z * fabs(cf2rho - c->a2rho));
  }
  c->betay = 0.;
  c->betaz = 0.;
  if (bperp2 != 0.) {
    double bperpi = 1. / sqrt(bperp2);
    c->betay = bperpi * w_by;
    c->betaz = bperpi * w_bz;
  }
}

/* prim2char: primitive vector vc -> characteristic amplitudes in the
   eigensystem of state w (slots RHO fast-, VX Alfven-, VY slow-, VZ entropy,
   PR slow+, BY Alfven+, BZ fast+; PSC and BX pass through). */
void or_prim2char(double *vc, const double *w, double gam)
{
  double vc_rho = vc[RHO], vc_vx = vc[VX], vc_vy = vc[VY], vc_vz = vc[VZ], vc_p = vc[PR];
  double vc_by = vc[BY], vc_bz = vc[BZ];
  or_charw c;
  char_weights(&c, w, gam);
  double rhoi = c.rhoi, cf = c.cf, cs = c.cs, betay = c.betay, betaz = c.betaz;
  double nf = 0.5 / (c.a2rho * rhoi);
  double qf_ = nf * cf * c.alphaf * c.s;
  double qs_ = nf * cs * c.alphas * c.s;
  double alphaf_ = nf * c.alphaf;
  double alphas_ = nf * c.alphas;
  double rhoisr = sqrt(rhoi);
  double arhoisr = c.a * rhoisr;
  double Af_ = alphaf_ * arhoisr;
  double As_ = alphas_ * arhoisr;
  double x = c.s * rhoisr;
  vc[RHO] = (alphaf_ * (rhoi * vc_p - cf * vc_vx) + qs_ * (betay * vc_vy + betaz * vc_vz) +
             As_ * (betay * vc_by + betaz * vc_bz));
  vc[VX] = 0.5 * ((-betaz * vc_vy + betay * vc_vz) + x * (-betaz * vc_by + betay * vc_bz));
  vc[VY] = (alphas_ * (rhoi * vc_p - cs * vc_vx) - qf_ * (betay * vc_vy + betaz * vc_vz) -
            Af_ * (betay * vc_by + betaz * vc_bz));
  vc[VZ] = vc_rho - 2. * nf * vc_p;
  vc[PR] = (alphas_ * (rhoi * vc_p + cs * vc_vx) + qf_ * (betay * vc_vy + betaz * vc_vz) -
            Af_ * (betay * vc_by + betaz * vc_bz));
  vc[BY] = 0.5 * ((betaz * vc_vy - betay * vc_vz) + x * (-betaz * vc_by + betay * vc_bz));
  vc[BZ] = (alphaf_ * (rhoi * vc_p + cf * vc_vx) - qs_ * (betay * vc_vy + betaz * vc_vz) +
            As_ * (betay * vc_by + betaz * vc_bz));
}

/* char2prim: the inverse map (eos_c.h:191-280). */
void or_char2prim(double *vc, const double *w, double gam)
{
  double vc_rho = vc[RHO], vc_vx = vc[VX], vc_vy = vc[VY], vc_vz = vc[VZ], vc_p = vc[PR];
  double vc_by = vc[BY], vc_bz = vc[BZ];
  double w_rho = w[RHO];
  or_charw c;
  char_weights(&c, w, gam);
  double cf = c.cf, cs = c.cs, alphaf = c.alphaf, alphas = c.alphas, s = c.s;
  double betay = c.betay, betaz = c.betaz;
  double qf = cf * alphaf * s;
  double qs = cs * alphas * s;
  double rhosr = sqrt(w_rho);
  double arhosr = c.a * rhosr;
  double Af = alphaf * arhosr;
  double As = alphas * arhosr;
  vc[RHO] = w_rho * (alphaf * (vc_rho + vc_bz) + alphas * (vc_vy + vc_p)) + vc_vz;
  vc[VX] = cf * alphaf * (vc_bz - vc_rho) + cs * alphas * (vc_p - vc_vy);
  vc[VY] = betay * (qs * (vc_rho - vc_bz) - qf * (vc_vy - vc_p)) - betaz * (vc_vx - vc_by);
  vc[VZ] = betaz * (qs * (vc_rho - vc_bz) - qf * (vc_vy - vc_p)) + betay * (vc_vx - vc_by);
  vc[PR] = c.a2rho * (alphaf * (vc_rho + vc_bz) + alphas * (vc_vy + vc_p));
  vc[BY] = betay * (As * (vc_rho + vc_bz) - Af * (vc_vy + vc_p)) - betaz * s * rhosr * (vc_vx + vc_by);
  vc[BZ] = betaz * (As * (vc_rho + vc_bz) - Af * (vc_vy + vc_p)) + betay * s * rhosr * (vc_vx + vc_by);
}

/* prim2char_1 / char2prim_1 (src/mhd/eos_c.c:53-94) as their loops evidently
   intend (D4: the shipped bodies put only the w[n] copy inside the n loop,
   leaving vc[n] with an undeclared n): every slot of rows vc1 and w1 at
   each i in [i1, i2] is gathered, projected, and written back. */
static void prim2char_rows(double **vc1, double **w1, int i1, int i2, double gam, int inverse)
{
  double w[NMODE], vc[NMODE];
  for (int i = i1; i <= i2; ++i) {
    for (int n = 0; n < NMODE; ++n) {
      w[n] = w1[n][i];
      vc[n] = vc1[n][i];
    }
    if (inverse) or_char2prim(vc, w, gam);
    else or_prim2char(vc, w, gam);
    for (int n = 0; n < NMODE; ++n) vc1[n][i] = vc[n];
  }
}

/* ------------------------------------------------------------------------ */
/* Reconstruction                                                            */
/* ------------------------------------------------------------------------ */

/* src/coordinates/reconstr_c.c:8-22 */
void or_reconstr_const(double **wl, double **wr, double ***w, double ***scr, int ax, int i1,
                       int i2)
{
  (void)scr;
  int ip = (ax == 0) ? 1 : 0;
  for (int n = 0; n < NMODE; ++n)
    for (int i = i1; i <= i2; ++i) {
      wr[n][i] = w[0][n][i];
      wl[n][i + ip] = w[0][n][i];
    }
}

/* src/coordinates/reconstr_c.c:29-91, uniform flat branch; char_proj
   limits the slopes in characteristic variables of the cell's own state
   (reconstr_c.c:64-67,79-80). */
void or_reconstr_linear_cp(double **wl, double **wr, double ***w, double ***scr, int ax, int i1,
                           int i2, int char_proj, double gam)
{
  double **dm = scr[0], **dp = scr[1], **d = scr[0], **wc = scr[2];
  int ip = (ax == 0) ? 1 : 0;
  for (int n = 0; n < NMODE; n++)
    for (int i = i1; i <= i2; i++) {
      wc[n][i] = w[1][n][i];
      dm[n][i] = w[1][n][i] - w[0][n][i];
      dp[n][i] = w[2][n][i] - w[1][n][i];
    }
  if (char_proj) {
    prim2char_rows(dm, wc, i1, i2, gam, 0);
    prim2char_rows(dp, wc, i1, i2, gam, 0);
  }
  for (int n = 0; n < NMODE; n++)
    for (int i = i1; i <= i2; i++) d[n][i] = 0.5 * or_vl_lim(dm[n][i], dp[n][i]);
  if (char_proj) prim2char_rows(d, wc, i1, i2, gam, 1);
  for (int n = 0; n < NMODE; n++)
    for (int i = i1; i <= i2; i++) {
      wr[n][i] = wc[n][i] - d[n][i];
      wl[n][i + ip] = wc[n][i] + d[n][i];
    }
}

void or_reconstr_linear(double **wl, double **wr, double ***w, double ***scr, int ax, int i1,
                        int i2)
{
  or_reconstr_linear_cp(wl, wr, w, scr, ax, i1, i2, 0, 0.);
}

/* ------------------------------------------------------------------------ */
/* Grid, coordinates and simulation state                                    */
/* ------------------------------------------------------------------------ */

typedef void (*rsolver_fn)(double **, double **, double **, double *, int, int, double);
typedef void (*reconstr_fn)(double **, double **, double ***, double ***, int, int, int);

struct or_sim {
  int dims, ng, reconstr_order, nthreads, dt_fms_ref, char_proj, hydro, ct_upwind, tintegr;
  int Nact[3], Ntot[3];
  int i1, i2, j1, j2, k1, k2;
  double lmin[3], lmax[3];
  double *lf[3], *lv[3], *dlf[3], *dlf_inv[3], *dlv[3];
  double gam, cour;
  int rsolver;
  int bc[3][2];
  long ncell;
  double *prim, *cons, *bfld, *cons_s, *bfld_s, *flx[3], *ec, *ee;
  double *cons_ss, *bfld_ss;  /* RK3's second stage (Uss, Bss) */
  double time, dt;
  int step;
  /* user sides (flag 3): per slab cell a source cell, -1 (fixed values,
     NMODE + 3 per cell) or -2 (left alone) */
  int *usr_src[3][2];
  double *usr_fix[3][2];
};

#define IDX(s, n, k, j, i) \
  ((((long)(n) * (s)->Ntot[2] + (k)) * (s)->Ntot[1] + (j)) * (s)->Ntot[0] + (i))

int or_ntot(const or_sim *s, int axis) { return s->Ntot[axis]; }
int or_ng(const or_sim *s) { return s->ng; }
double *or_prim(or_sim *s) { return s->prim; }
double *or_cons(or_sim *s) { return s->cons; }
double *or_bfld(or_sim *s) { return s->bfld; }
double or_time(const or_sim *s) { return s->time; }
int or_step(const or_sim *s) { return s->step; }
double or_last_dt(const or_sim *s) { return s->dt; }

/* get_hst_vars_grid (src/io/output.pyx:20-138) with MFIELD: sums over the
   active cells of the conserved state, in loop order, divided by the cell
   count; the last three columns become sqrt(|<B^2> - <B>^2|). */
void or_history(or_sim *s, double *out)
{
  double sum[14];
  for (int q = 0; q < 14; ++q) sum[q] = 0.;
  double w[NMODE], u[NMODE];
  for (int k = s->k1; k <= s->k2; ++k)
    for (int j = s->j1; j <= s->j2; ++j)
      for (int i = s->i1; i <= s->i2; ++i) {
        for (int n = 0; n < NMODE; ++n) w[n] = s->prim[IDX(s, n, k, j, i)];
        or_prim2cons(u, w, s->gam);
        double rhoi = 1. / u[RHO];
        double engk = 0.5 * rhoi * (sqr(u[MX]) + sqr(u[MY]) + sqr(u[MZ]));
        double bx2 = sqr(u[BX]), by2 = sqr(u[BY]), bz2 = sqr(u[BZ]);
        double engm = 0.5 * (bx2 + by2 + bz2);
        double engt = u[EN] - engk - engm;
        double v[14] = {u[RHO], engk, engt, engm, u[EN], u[MX], u[MY], u[MZ],
                        u[BX], u[BY], u[BZ], bx2, by2, bz2};
        for (int q = 0; q < 14; ++q) sum[q] += v[q];
      }
  double ncells = (double)s->Nact[0] * s->Nact[1] * s->Nact[2];
  out[0] = s->step;
  out[1] = s->time;
  for (int q = 0; q < 14; ++q) out[2 + q] = sum[q] / ncells;
  for (int q = 13; q < 16; ++q) out[q] = sqrt(fabs(out[q] - sqr(out[q - 3])));
}
void or_set_dt_fms_ref(or_sim *s, int on) { s->dt_fms_ref = on; }
void or_set_char_proj(or_sim *s, int on) { s->char_proj = on; }
/* hydro build (MFIELD=0): the sound speed in the wave estimates and the CFL
   dt, HLL or HLLC; the field arrays stay in place and must hold zeros (CT
   then keeps them exactly zero) */
void or_set_hydro(or_sim *s, int on) { s->hydro = on; }
/* corner EMF: 0 interp_ee1 (the reference's), 1 the upwind interp_ee2 */
void or_set_ct_upwind(or_sim *s, int on) { s->ct_upwind = on; }

/* User side (flag 3): src and fixed over the side's slab (ng layers along
   the axis, full extent of the others, x fastest); copied.  0 or -1. */
int or_set_user_bc(or_sim *s, int axis, int side, const int *src, const double *fixed)
{
  if (axis < 0 || axis >= s->dims || side < 0 || side > 1 || s->bc[axis][side] != 3) return -1;
  long n = (long)s->ng;
  for (int a = 0; a < 3; ++a)
    if (a != axis) n *= s->Ntot[a];
  int *sc = malloc(n * sizeof(int));
  double *fc = calloc(n * (NMODE + 3), sizeof(double));
  for (long e = 0; e < n; ++e) {
    if (src[e] < -2 || src[e] >= s->ncell) {
      free(sc);
      free(fc);
      return -1;
    }
    sc[e] = src[e];
  }
  if (fixed) memcpy(fc, fixed, n * (NMODE + 3) * sizeof(double));
  free(s->usr_src[axis][side]);
  free(s->usr_fix[axis][side]);
  s->usr_src[axis][side] = sc;
  s->usr_fix[axis][side] = fc;
  return 0;
}
/* For the oracle's own tests: the face-flux arrays [9][cells] and the edge
   EMFs [3][cells], and one corner-EMF evaluation (ec_from_prim, then
   interp_ee1 or interp_ee2) over the corrector's limits from the current
   prim and the flux arrays as they stand. */
static void ec_from_prim(or_sim *s, const double *w, const int *lims);
static void interp_ee1(or_sim *s, const int *lims);
static void interp_ee2(or_sim *s, const int *lims, const double *w, double dt);
static void integr_lims(const or_sim *s, int order, int *l);
double *or_flux(or_sim *s, int axis) { return s->flx[axis]; }
double *or_edge_emf(or_sim *s) { return s->ee; }
void or_corner_emf(or_sim *s, int upwind, double dt)
{
  int lims[6];
  integr_lims(s, s->reconstr_order, lims);
  ec_from_prim(s, s->prim, lims);
  if (upwind) interp_ee2(s, lims, s->prim, dt);
  else interp_ee1(s, lims);
}

double *or_coord(or_sim *s, int which, int axis)
{
  switch (which) {
  case 0: return s->lf[axis];
  case 1: return s->lv[axis];
  case 2: return s->dlf[axis];
  case 3: return s->dlf_inv[axis];
  default: return s->dlv[axis];
  }
}

/* Uniform Cartesian spacings: src/coordinates/coord.pyx:201-257 (uniform
   branch, one block, ghost faces re-spaced for periodic axes) and
   src/coordinates/cartesian.pyx:18-43. */
static void set_coordinates(or_sim *s)
{
  for (int n = 0; n < 3; ++n) {
    int N = s->Ntot[n];
    s->lf[n] = calloc(N + 2, sizeof(double));
    s->lv[n] = calloc(N + 2, sizeof(double));
    s->dlf[n] = calloc(N + 2, sizeof(double));
    s->dlf_inv[n] = calloc(N + 2, sizeof(double));
    s->dlv[n] = calloc(N + 2, sizeof(double));
    double *lf = s->lf[n];
    if (N == 1) {
      lf[0] = s->lmin[n];
      lf[1] = s->lmax[n];
    } else {
      int ind1 = (n == 0) ? s->i1 : (n == 1) ? s->j1 : s->k1;
      int ind2 = (n == 0) ? s->i2 : (n == 1) ? s->j2 : s->k2;
      double dx = (s->lmax[n] - s->lmin[n]) / s->Nact[n];
      double l1 = 0 * s->Nact[n] * dx; /* pos[n]=0 for a single block */
      for (int i = 0; i < N + 1; ++i) lf[i] = l1 + (i - ind1) * dx;
      if (s->bc[n][0] == 0)
        for (int i = ind1 - 1; i > -1; --i) {
          int i_ = ind2 - i + (ind1 - 1);
          double dlf_ = lf[i_ + 1] - lf[i_];
          lf[i] = lf[i + 1] - dlf_;
        }
      if (s->bc[n][1] == 0)
        for (int i = ind2 + 1; i < N; ++i) {
          int i_ = ind1 + i - (ind2 + 1);
          double dlf_ = lf[i_ + 1] - lf[i_];
          lf[i + 1] = lf[i] + dlf_;
        }
    }
    for (int i = 0; i < N; ++i) {
      s->dlf[n][i] = lf[i + 1] - lf[i];
      s->dlf_inv[n][i] = 1. / s->dlf[n][i];
    }
    for (int i = 0; i < N; ++i) s->lv[n][i] = 0.5 * (lf[i] + lf[i + 1]);
    for (int i = 1; i < N; ++i) s->dlv[n][i] = s->lv[n][i] - s->lv[n][i - 1];
  }
}

or_sim *or_create(int nx, int ny, int nz, int reconstr_order, const double *lmin,
                  const double *lmax, double gam, double cour, int rsolver, const int *bc_flags,
                  int nthreads)
{
  return or_create_ex(nx, ny, nz, reconstr_order, lmin, lmax, gam, cour, rsolver, bc_flags,
                      nthreads, 0);
}

/* tintegr 1: the RK3 integrator (src/integration/_rk3.pyx), whose three
   stages run without a ghost fill between them and so need ng = 9
   (src/coordinates/grid.pyx:142-143) */
or_sim *or_create_ex(int nx, int ny, int nz, int reconstr_order, const double *lmin,
                     const double *lmax, double gam, double cour, int rsolver,
                     const int *bc_flags, int nthreads, int tintegr)
{
  or_sim *s = calloc(1, sizeof(or_sim));
  s->dims = (nz > 1) ? 3 : (ny > 1) ? 2 : 1;
  /* ng: src/coordinates/grid.pyx:129-143 */
  s->reconstr_order = reconstr_order;
  s->tintegr = tintegr;
  s->ng = tintegr ? 9 : (reconstr_order == 0) ? 2 : 3;
  s->Nact[0] = nx;
  s->Nact[1] = ny;
  s->Nact[2] = nz;
  s->Ntot[0] = nx + 2 * s->ng;
  s->Ntot[1] = (s->dims >= 2) ? ny + 2 * s->ng : 1;
  s->Ntot[2] = (s->dims >= 3) ? nz + 2 * s->ng : 1;
  /* index ranges: src/coordinates/grid.pyx:155-161 */
  s->i1 = s->ng;
  s->i2 = nx + s->ng - 1;
  s->j1 = (s->dims >= 2) ? s->ng : 0;
  s->j2 = (s->dims >= 2) ? ny + s->ng - 1 : 0;
  s->k1 = (s->dims >= 3) ? s->ng : 0;
  s->k2 = (s->dims >= 3) ? nz + s->ng - 1 : 0;
  for (int n = 0; n < 3; ++n) {
    s->lmin[n] = lmin[n];
    s->lmax[n] = lmax[n];
    s->bc[n][0] = bc_flags[2 * n];
    s->bc[n][1] = bc_flags[2 * n + 1];
  }
  s->gam = gam;
  s->cour = cour;
  s->rsolver = rsolver;
  s->nthreads = nthreads > 0 ? nthreads : 1;
  s->dt_fms_ref = 1;
  set_coordinates(s);
  s->ncell = (long)s->Ntot[0] * s->Ntot[1] * s->Ntot[2];
  s->prim = calloc(NMODE * s->ncell, sizeof(double));
  s->cons = calloc(NMODE * s->ncell, sizeof(double));
  s->cons_s = calloc(NMODE * s->ncell, sizeof(double));
  s->bfld = calloc(3 * s->ncell, sizeof(double));
  s->bfld_s = calloc(3 * s->ncell, sizeof(double));
  for (int d = 0; d < 3; ++d) s->flx[d] = calloc(NMODE * s->ncell, sizeof(double));
  s->ec = calloc(3 * s->ncell, sizeof(double));
  s->ee = calloc(3 * s->ncell, sizeof(double));
  if (tintegr) {
    s->cons_ss = calloc(NMODE * s->ncell, sizeof(double));
    s->bfld_ss = calloc(3 * s->ncell, sizeof(double));
  }
  return s;
}

void or_destroy(or_sim *s)
{
  if (!s) return;
  for (int n = 0; n < 3; ++n) {
    free(s->lf[n]);
    free(s->lv[n]);
    free(s->dlf[n]);
    free(s->dlf_inv[n]);
    free(s->dlv[n]);
    free(s->flx[n]);
  }
  free(s->prim);
  free(s->cons);
  free(s->cons_s);
  free(s->cons_ss);
  free(s->bfld_ss);
  free(s->bfld);
  free(s->bfld_s);
  free(s->ec);
  free(s->ee);
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 2; ++d) {
      free(s->usr_src[a][d]);
      free(s->usr_fix[a][d]);
    }
  free(s);
}

/* get_integr_lims: src/integration/integrator.pyx:375-383 */
static void integr_lims(const or_sim *s, int order, int *l)
{
  l[0] = order + 1;
  l[1] = s->Ntot[0] - 2 - order;
  l[2] = order + 1;
  l[3] = s->Ntot[1] - 2 - order;
  l[4] = order + 1;
  l[5] = s->Ntot[2] - 2 - order;
  if (s->dims < 2) l[2] = l[3] = 0;
  if (s->dims < 3) l[4] = l[5] = 0;
}

/* swap_mhd_vec: src/integration/godunov.pyx:297-324 */
static void swap_mhd_vec(int ax, int *vars)
{
  vars[RHO] = RHO;
  for (int n = 4; n < NMODE - 3; ++n) vars[n] = n;
  if (ax == 0) {
    vars[VX] = VX; vars[VY] = VY; vars[VZ] = VZ;
    vars[BX] = BX; vars[BY] = BY; vars[BZ] = BZ;
  } else if (ax == 1) {
    vars[VX] = VY; vars[VY] = VZ; vars[VZ] = VX;
    vars[BX] = BY; vars[BY] = BZ; vars[BZ] = BX;
  } else {
    vars[VX] = VZ; vars[VY] = VX; vars[VZ] = VY;
    vars[BX] = BZ; vars[BY] = BX; vars[BZ] = BY;
  }
}

/* integrator.pyx:253-268: the solver the build binds to the selection */
static rsolver_fn pick_rsolver(int id, int hydro)
{
  if (hydro) return id == OR_RS_HLLC ? or_hllc_flux : or_hll_flux_hydro;
  switch (id) {
  case OR_RS_HLLD: return or_hlld_flux;
  case OR_RS_HLLT: return or_hllt_flux;
  default: return or_hll_flux;
  }
}

static double **alloc_rows(int nrow, int len)
{
  double **r = malloc(nrow * sizeof(double *));
  double *buf = calloc((size_t)nrow * len, sizeof(double));
  for (int n = 0; n < nrow; ++n) r[n] = buf + (size_t)n * len;
  return r;
}
static void free_rows(double **r)
{
  free(r[0]);
  free(r);
}

/* godunov_fluxes: src/integration/godunov.pyx:20-210 (D1 applied) */
/* Reconstruction of the stage (reconstr_func of godunov.pyx:103-197): donor
   cell for order 0, PLM otherwise, in characteristic variables when
   char_proj is set (integrator.pyx:245, D4). */
#define RECONSTR(wl_, wr_, w_, scr_, ax_, a_, b_)                                    \
  ((order) == 0 ? or_reconstr_const(wl_, wr_, w_, scr_, ax_, a_, b_)                \
                : or_reconstr_linear_cp(wl_, wr_, w_, scr_, ax_, a_, b_, s->char_proj, \
                                        s->gam))

static void godunov_fluxes(or_sim *s, const double *prim, const double *bfld, const int *lims,
                           int order)
{
  int il = lims[0], iu = lims[1], jl = lims[2], ju = lims[3], kl = lims[4], ku = lims[5];
  int varsx[NMODE], varsy[NMODE], varsz[NMODE];
  swap_mhd_vec(0, varsx);
  swap_mhd_vec(1, varsy);
  swap_mhd_vec(2, varsz);
  rsolver_fn rs = pick_rsolver(s->rsolver, s->hydro);
  int nx = s->Ntot[0];
  int len = nx + 4;
  int nst = 2 * order + 1;
  double *pbase = (double *)prim;

#pragma omp parallel num_threads(s->nthreads)
  {
    double **wl = alloc_rows(NMODE, len), **wl_ = alloc_rows(NMODE, len);
    double **wr = alloc_rows(NMODE, len);
    double **scrrows[3] = {alloc_rows(NMODE, len), alloc_rows(NMODE, len),
                           alloc_rows(NMODE, len)};
    double ***scr = scrrows;
    double *prim1_rows[3][NMODE];
    double **prim1[3] = {prim1_rows[0], prim1_rows[1], prim1_rows[2]};
    double *flx1[NMODE];

    /* x-direction, godunov.pyx:84-112 */
    int kb = kl - 1, ke = ku + 1, jb = jl - 1, je = ju + 1;
    if (s->dims < 2) jb = je = 0;
    if (s->dims < 3) kb = ke = 0;
#pragma omp for schedule(dynamic)
    for (int k = kb; k <= ke; ++k)
      for (int j = jb; j <= je; ++j) {
        for (int d = 0; d < nst; ++d)
          for (int n = 0; n < NMODE; ++n)
            prim1[d][n] = pbase + IDX(s, varsx[n], k, j, 0) + d - order;
        RECONSTR(wl, wr, prim1, scr, 0, il - 1, iu + 1);
        for (int n = 0; n < NMODE; ++n) flx1[n] = s->flx[0] + IDX(s, varsx[n], k, j, 0);
        rs(flx1, wl, wr, (double *)bfld + IDX(s, 0, k, j, 0), il, iu + 1, s->gam);
      }

    if (s->dims >= 2) {
      /* y-direction, godunov.pyx:118-160 */
      int ib = il - 1, ie = iu + 1;
      kb = kl - 1;
      ke = ku + 1;
      if (s->dims < 3) kb = ke = 0;
#pragma omp for schedule(dynamic)
      for (int k = kb; k <= ke; ++k) {
        double **wlc = wl, **wln = wl_;
        for (int d = 0; d < nst; ++d)
          for (int n = 0; n < NMODE; ++n)
            prim1[d][n] = pbase + IDX(s, varsy[n], k, jl - 1 + d - order, 0);
        RECONSTR(wlc, wr, prim1, scr, 1, ib, ie);
        for (int j = jl; j <= ju + 1; ++j) {
          for (int d = 0; d < nst; ++d)
            for (int n = 0; n < NMODE; ++n)
              prim1[d][n] = pbase + IDX(s, varsy[n], k, j + d - order, 0);
          RECONSTR(wln, wr, prim1, scr, 1, ib, ie);
          for (int n = 0; n < NMODE; ++n) flx1[n] = s->flx[1] + IDX(s, varsy[n], k, j, 0);
          rs(flx1, wlc, wr, (double *)bfld + IDX(s, 1, k, j, 0), ib, ie, s->gam);
          double **t = wlc; /* D1: the intended swap */
          wlc = wln;
          wln = t;
        }
      }
    }

    if (s->dims >= 3) {
      /* z-direction, godunov.pyx:166-204 */
      int ib = il - 1, ie = iu + 1;
      jb = jl - 1;
      je = ju + 1;
#pragma omp for schedule(dynamic)
      for (int j = jb; j <= je; ++j) {
        double **wlc = wl, **wln = wl_;
        for (int d = 0; d < nst; ++d)
          for (int n = 0; n < NMODE; ++n)
            prim1[d][n] = pbase + IDX(s, varsz[n], kl - 1 + d - order, j, 0);
        RECONSTR(wlc, wr, prim1, scr, 2, ib, ie);
        for (int k = kl; k <= ku + 1; ++k) {
          for (int d = 0; d < nst; ++d)
            for (int n = 0; n < NMODE; ++n)
              prim1[d][n] = pbase + IDX(s, varsz[n], k + d - order, j, 0);
          RECONSTR(wln, wr, prim1, scr, 2, ib, ie);
          for (int n = 0; n < NMODE; ++n) flx1[n] = s->flx[2] + IDX(s, varsz[n], k, j, 0);
          rs(flx1, wlc, wr, (double *)bfld + IDX(s, 2, k, j, 0), ib, ie, s->gam);
          double **t = wlc;
          wlc = wln;
          wln = t;
        }
      }
    }
    free_rows(wl);
    free_rows(wl_);
    free_rows(wr);
    for (int d = 0; d < 3; ++d) free_rows(scrrows[d]);
  }
}

/* advance_hydro: src/integration/godunov.pyx:215-252 (Cartesian metric,
   src/coordinates/cartesian.pyx:99-115) */
static void advance_hydro(or_sim *s, double *u1, const double *u0, const int *lims, double dt)
{
  double *fx = s->flx[0], *fy = s->flx[1], *fz = s->flx[2];
  double **dlf = s->dlf;
  for (int n = 0; n < NMODE - 3; ++n)
    for (int k = lims[4]; k <= lims[5]; ++k)
      for (int j = lims[2]; j <= lims[3]; ++j)
        for (int i = lims[0]; i <= lims[1]; ++i) {
          double dtdv = dt / (dlf[0][i] * dlf[1][j] * dlf[2][k]);
          double dam = dlf[1][j] * dlf[2][k];
          double dap = dlf[1][j] * dlf[2][k];
          long c = IDX(s, n, k, j, i);
          u1[c] = u0[c] - dtdv * (dap * fx[IDX(s, n, k, j, i + 1)] - dam * fx[c]);
          if (s->dims >= 2) {
            dam = dlf[0][i] * dlf[2][k];
            dap = dlf[0][i] * dlf[2][k];
            u1[c] -= dtdv * (dap * fy[IDX(s, n, k, j + 1, i)] - dam * fy[c]);
          }
          if (s->dims >= 3) {
            dam = dlf[0][i] * dlf[1][j];
            dap = dlf[0][i] * dlf[1][j];
            u1[c] -= dtdv * (dap * fz[IDX(s, n, k + 1, j, i)] - dam * fz[c]);
          }
        }
}

/* ec_from_prim: src/integration/ct.pyx:128-148 */
static void ec_from_prim(or_sim *s, const double *w, const int *lims)
{
  int il = lims[0] - 1, iu = lims[1] + 1, jl = lims[2] - 1, ju = lims[3] + 1;
  int kl = lims[4] - 1, ku = lims[5] + 1;
  if (s->dims < 2) jl = ju = 0;
  if (s->dims < 3) kl = ku = 0;
  double *ec = s->ec;
  for (int k = kl; k <= ku; ++k)
    for (int j = jl; j <= ju; ++j)
      for (int i = il; i <= iu; ++i) {
        double vx = w[IDX(s, VX, k, j, i)], vy = w[IDX(s, VY, k, j, i)];
        double vz = w[IDX(s, VZ, k, j, i)], bx = w[IDX(s, BX, k, j, i)];
        double by = w[IDX(s, BY, k, j, i)], bz = w[IDX(s, BZ, k, j, i)];
        ec[IDX(s, 0, k, j, i)] = -(vy * bz - vz * by);
        ec[IDX(s, 1, k, j, i)] = (vx * bz - vz * bx);
        ec[IDX(s, 2, k, j, i)] = -(vx * by - vy * bx);
      }
}

/* Weight of the cell-centred EMFs in interp_ee1's corner average
   (ct.pyx:181-205 subtracts 0.25 * the sum of four).  Only a mutant build
   for the oracle's own sensitivity test (oracle/Makefile `mutants`, built
   with -DOR_MUTANT_EC_SIGN) flips its sign; every other build uses 0.25. */
#ifdef OR_MUTANT_EC_SIGN
#define OR_EC_W (-0.25)
#else
#define OR_EC_W 0.25
#endif

/* interp_ee1: src/integration/ct.pyx:153-215 */
static void interp_ee1(or_sim *s, const int *lims)
{
  int il = lims[0], iu = lims[1] + 1, jl = lims[2], ju = lims[3] + 1;
  int kl = lims[4], ku = lims[5] + 1;
  if (s->dims < 2) jl = ju = 0;
  if (s->dims < 3) kl = ku = 0;
  double *ee = s->ee, *ec = s->ec, *fx = s->flx[0], *fy = s->flx[1], *fz = s->flx[2];
#define A(arr, n, k, j, i) arr[IDX(s, n, k, j, i)]
  if (s->dims == 3) {
    for (int k = kl; k <= ku; ++k)
      for (int j = jl; j <= ju; ++j)
        for (int i = il; i <= iu; ++i) {
          A(ee, 0, k, j, i) = (0.5 * (-A(fy, BZ, k - 1, j, i) - A(fy, BZ, k, j, i) +
                                      A(fz, BY, k, j - 1, i) + A(fz, BY, k, j, i)) -
                               OR_EC_W * (A(ec, 0, k, j - 1, i) + A(ec, 0, k, j, i) +
                                       A(ec, 0, k - 1, j - 1, i) + A(ec, 0, k - 1, j, i)));
          A(ee, 1, k, j, i) = (0.5 * (-A(fz, BX, k, j, i - 1) - A(fz, BX, k, j, i) +
                                      A(fx, BZ, k - 1, j, i) + A(fx, BZ, k, j, i)) -
                               OR_EC_W * (A(ec, 1, k, j, i - 1) + A(ec, 1, k, j, i) +
                                       A(ec, 1, k - 1, j, i - 1) + A(ec, 1, k - 1, j, i)));
          A(ee, 2, k, j, i) = (0.5 * (-A(fx, BY, k, j - 1, i) - A(fx, BY, k, j, i) +
                                      A(fy, BX, k, j, i - 1) + A(fy, BX, k, j, i)) -
                               OR_EC_W * (A(ec, 2, k, j, i - 1) + A(ec, 2, k, j, i) +
                                       A(ec, 2, k, j - 1, i - 1) + A(ec, 2, k, j - 1, i)));
        }
  } else if (s->dims == 2) {
    for (int j = jl; j <= ju; ++j)
      for (int i = il; i <= iu; ++i) {
        A(ee, 0, 0, j, i) = -A(fy, BZ, 0, j, i);
        A(ee, 1, 0, j, i) = A(fx, BZ, 0, j, i);
        A(ee, 2, 0, j, i) = (0.5 * (-A(fx, BY, 0, j - 1, i) - A(fx, BY, 0, j, i) +
                                    A(fy, BX, 0, j, i - 1) + A(fy, BX, 0, j, i)) -
                             OR_EC_W * (A(ec, 2, 0, j, i - 1) + A(ec, 2, 0, j, i) +
                                     A(ec, 2, 0, j - 1, i - 1) + A(ec, 2, 0, j - 1, i)));
      }
  } else {
    for (int i = il; i <= iu; ++i) {
      A(ee, 0, 0, 0, i) = -A(ec, 0, 0, 0, i);
      A(ee, 1, 0, 0, i) = A(fx, BZ, 0, 0, i);
      A(ee, 2, 0, 0, i) = -A(fx, BY, 0, 0, i);
    }
  }
}

/* get_ct_weight: src/integration/ct.pyx:221-226.  Weight of the lower
   (left) cell's EMF slope: 1/2 + clip(1e3 frho dt / (min(rho) ds), +-1/2),
   a smoothed sign of the mass flux through the face. */
static inline double ct_weight(double frho, double rho_l, double rho_r, double ds, double dt)
{
  double v_c = 1e3 * frho * dt / (fmin(rho_l, rho_r) * ds);
  return 0.5 + fmax(-0.5, fmin(0.5, v_c));
}

/* interp_ee2, the upwind corner EMF (Gardiner & Stone 2005 with a smoothed
   upwind weight): src/integration/ct.pyx:229-358, which the reference never
   calls (integrator.pyx:106-108 comments it out).  Followed in 3D with two
   cell-EMF indices corrected: in the Ex slopes dedz_34 and dedy_34 the
   reference reads ec[0,k,j-1,i] where the cyclic image of its own Ez and Ey
   formulas reads ec[0,k-1,j-1,i]; with the correction the three components
   are cyclic images of each other and the scheme reduces to interp_ee1 when
   every weight is 1/2 (no mass flux), which fails with the indices as
   written.  In 2D and 1D the reference's loops are empty (their k range is
   range(0, 0)); the 2D Ez formula and the Ex/Ey branches it intends are
   used.  rho is the stage's density, dt the full step (self.dt). */
static void interp_ee2(or_sim *s, const int *lims, const double *w, double dt)
{
  int il = lims[0], iu = lims[1] + 1, jl = lims[2], ju = lims[3] + 1;
  int kl = lims[4], ku = lims[5] + 1;
  if (s->dims < 2) jl = ju = 0;
  if (s->dims < 3) kl = ku = 0;
  double *ee = s->ee, *ec = s->ec, *fx = s->flx[0], *fy = s->flx[1], *fz = s->flx[2];
  double **dlf = s->dlf, **dlv = s->dlv;
  for (int k = kl; k <= ku; ++k)
    for (int j = jl; j <= ju; ++j)
      for (int i = il; i <= iu; ++i) {
        if (s->dims == 1) {
          A(ee, 0, 0, 0, i) = A(ec, 0, 0, 0, i);
          A(ee, 1, 0, 0, i) = A(fx, BZ, 0, 0, i);
          A(ee, 2, 0, 0, i) = -A(fx, BY, 0, 0, i);
          continue;
        }
        /* ---- Ez at (i-1/2, j-1/2) */
        {
          double dsx = dlf[0][i];
          double wgtx = ct_weight(A(fx, RHO, k, j, i), A(w, RHO, k, j, i - 1),
                                  A(w, RHO, k, j, i), dsx, dt);
          double dedy_14 = (wgtx * (A(ec, 2, k, j, i - 1) - A(fy, BX, k, j, i - 1)) +
                            (1. - wgtx) * (A(ec, 2, k, j, i) - A(fy, BX, k, j, i)));
          double dedy_34 = (wgtx * (A(fy, BX, k, j, i - 1) - A(ec, 2, k, j - 1, i - 1)) +
                            (1. - wgtx) * (A(fy, BX, k, j, i) - A(ec, 2, k, j - 1, i)));
          double dsy = dlv[1][j];
          double wgty = ct_weight(A(fy, RHO, k, j, i), A(w, RHO, k, j - 1, i),
                                  A(w, RHO, k, j, i), dsy, dt);
          double dedx_14 = (wgty * (A(ec, 2, k, j - 1, i) + A(fx, BY, k, j - 1, i)) +
                            (1. - wgty) * (A(ec, 2, k, j, i) + A(fx, BY, k, j, i)));
          double dedx_34 = (wgty * (-A(fx, BY, k, j - 1, i) - A(ec, 2, k, j - 1, i - 1)) +
                            (1. - wgty) * (-A(fx, BY, k, j, i) - A(ec, 2, k, j, i - 1)));
          A(ee, 2, k, j, i) = 0.25 * (-A(fx, BY, k, j - 1, i) - A(fx, BY, k, j, i) +
                                      A(fy, BX, k, j, i - 1) + A(fy, BX, k, j, i) + dedy_34 -
                                      dedy_14 + dedx_34 - dedx_14);
        }
        if (s->dims == 2) {
          A(ee, 0, k, j, i) = -A(fy, BZ, k, j, i);
          A(ee, 1, k, j, i) = A(fx, BZ, k, j, i);
          continue;
        }
        /* ---- Ex at (j-1/2, k-1/2) */
        {
          double dsy = dlf[1][j];
          double wgty = ct_weight(A(fy, RHO, k, j, i), A(w, RHO, k, j - 1, i),
                                  A(w, RHO, k, j, i), dsy, dt);
          double dedz_14 = (wgty * (A(ec, 0, k, j - 1, i) - A(fz, BY, k, j - 1, i)) +
                            (1. - wgty) * (A(ec, 0, k, j, i) - A(fz, BY, k, j, i)));
          double dedz_34 = (wgty * (A(fz, BY, k, j - 1, i) - A(ec, 0, k - 1, j - 1, i)) +
                            (1. - wgty) * (A(fz, BY, k, j, i) - A(ec, 0, k - 1, j, i)));
          double dsz = dlv[2][k];
          double wgtz = ct_weight(A(fz, RHO, k, j, i), A(w, RHO, k - 1, j, i),
                                  A(w, RHO, k, j, i), dsz, dt);
          double dedy_14 = (wgtz * (A(ec, 0, k - 1, j, i) + A(fy, BZ, k - 1, j, i)) +
                            (1. - wgtz) * (A(ec, 0, k, j, i) + A(fy, BZ, k, j, i)));
          double dedy_34 = (wgtz * (-A(fy, BZ, k - 1, j, i) - A(ec, 0, k - 1, j - 1, i)) +
                            (1. - wgtz) * (-A(fy, BZ, k, j, i) - A(ec, 0, k, j - 1, i)));
          A(ee, 0, k, j, i) = 0.25 * (-A(fy, BZ, k - 1, j, i) - A(fy, BZ, k, j, i) +
                                      A(fz, BY, k, j - 1, i) + A(fz, BY, k, j, i) + dedz_34 -
                                      dedz_14 + dedy_34 - dedy_14);
        }
        /* ---- Ey at (i-1/2, k-1/2) */
        {
          double dsz = dlf[2][k];
          double wgtz = ct_weight(A(fz, RHO, k, j, i), A(w, RHO, k - 1, j, i),
                                  A(w, RHO, k, j, i), dsz, dt);
          double dedx_14 = (wgtz * (A(ec, 1, k - 1, j, i) - A(fx, BZ, k - 1, j, i)) +
                            (1. - wgtz) * (A(ec, 1, k, j, i) - A(fx, BZ, k, j, i)));
          double dedx_34 = (wgtz * (A(fx, BZ, k - 1, j, i) - A(ec, 1, k - 1, j, i - 1)) +
                            (1. - wgtz) * (A(fx, BZ, k, j, i) - A(ec, 1, k, j, i - 1)));
          double dsx = dlv[0][i];
          double wgtx = ct_weight(A(fx, RHO, k, j, i), A(w, RHO, k, j, i - 1),
                                  A(w, RHO, k, j, i), dsx, dt);
          double dedz_14 = (wgtx * (A(ec, 1, k, j, i - 1) + A(fz, BX, k, j, i - 1)) +
                            (1. - wgtx) * (A(ec, 1, k, j, i) + A(fz, BX, k, j, i)));
          double dedz_34 = (wgtx * (-A(fz, BX, k, j, i - 1) - A(ec, 1, k - 1, j, i - 1)) +
                            (1. - wgtx) * (-A(fz, BX, k, j, i) - A(ec, 1, k - 1, j, i)));
          A(ee, 1, k, j, i) = 0.25 * (-A(fz, BX, k, j, i - 1) - A(fz, BX, k, j, i) +
                                      A(fx, BZ, k - 1, j, i) + A(fx, BZ, k, j, i) + dedx_34 -
                                      dedx_14 + dedz_34 - dedz_14);
        }
      }
}

/* advance_b: src/integration/ct.pyx:15-91 (D11: By over j in [jl,ju+1];
   2D Bz only at k=0).  Edge lengths and face areas per
   src/coordinates/cartesian.pyx:57-109. */
static void advance_b(or_sim *s, double *b1, const double *b0, const int *lims, double dt)
{
  double **dlf = s->dlf;
  double *ee = s->ee;
  int kz_hi = (s->dims >= 3) ? lims[5] + 1 : lims[5];
  int jy_hi = (s->dims >= 2) ? lims[3] + 1 : lims[3];
  for (int k = lims[4]; k <= lims[5]; ++k)
    for (int j = lims[2]; j <= lims[3]; ++j)
      for (int i = lims[0]; i <= lims[1] + 1; ++i) {
        double dtda = dt / (dlf[1][j] * dlf[2][k]);
        long c = IDX(s, 0, k, j, i);
        b1[c] = b0[c];
        if (s->dims >= 2) {
          double dsm = dlf[2][k], dsp = dlf[2][k];
          b1[c] -= dtda * (dsp * A(ee, 2, k, j + 1, i) - dsm * A(ee, 2, k, j, i));
        }
        if (s->dims >= 3) {
          double dsm = dlf[1][j], dsp = dlf[1][j];
          b1[c] += dtda * (dsp * A(ee, 1, k + 1, j, i) - dsm * A(ee, 1, k, j, i));
        }
      }
  for (int k = lims[4]; k <= lims[5]; ++k)
    for (int j = lims[2]; j <= jy_hi; ++j)
      for (int i = lims[0]; i <= lims[1]; ++i) {
        double dtda = dt / (dlf[0][i] * dlf[2][k]);
        double dsm = dlf[2][k], dsp = dlf[2][k];
        long c = IDX(s, 1, k, j, i);
        b1[c] = b0[c] + dtda * (dsp * A(ee, 2, k, j, i + 1) - dsm * A(ee, 2, k, j, i));
        if (s->dims >= 3) {
          dsm = dlf[0][i];
          dsp = dlf[0][i];
          b1[c] -= dtda * (dsp * A(ee, 0, k + 1, j, i) - dsm * A(ee, 0, k, j, i));
        }
      }
  for (int k = lims[4]; k <= kz_hi; ++k)
    for (int j = lims[2]; j <= lims[3]; ++j)
      for (int i = lims[0]; i <= lims[1]; ++i) {
        double dtda = dt / (dlf[0][i] * dlf[1][j]);
        double dsm = dlf[1][j], dsp = dlf[1][j];
        long c = IDX(s, 2, k, j, i);
        b1[c] = b0[c] - dtda * (dsp * A(ee, 1, k, j, i + 1) - dsm * A(ee, 1, k, j, i));
        if (s->dims >= 2) {
          dsm = dlf[0][i];
          dsp = dlf[0][i];
          b1[c] += dtda * (dsp * A(ee, 0, k, j + 1, i) - dsm * A(ee, 0, k, j, i));
        }
      }
}

/* interp_bc: src/integration/ct.pyx:96-123 */
static void interp_bc(or_sim *s, double *u, const double *b, const int *lims)
{
  double **lf = s->lf, **lv = s->lv, **dlfi = s->dlf_inv;
  for (int k = lims[4]; k <= lims[5]; ++k)
    for (int j = lims[2]; j <= lims[3]; ++j)
      for (int i = lims[0]; i <= lims[1]; ++i) {
        A(u, BX, k, j, i) = ((lv[0][i] - lf[0][i]) * A(b, 0, k, j, i) +
                             (lf[0][i + 1] - lv[0][i]) * A(b, 0, k, j, i + 1)) * dlfi[0][i];
        if (s->dims >= 2)
          A(u, BY, k, j, i) = ((lv[1][j] - lf[1][j]) * A(b, 1, k, j, i) +
                               (lf[1][j + 1] - lv[1][j]) * A(b, 1, k, j + 1, i)) * dlfi[1][j];
        else
          A(u, BY, k, j, i) = A(b, 1, k, j, i);
        if (s->dims >= 3)
          A(u, BZ, k, j, i) = ((lv[2][k] - lf[2][k]) * A(b, 2, k, j, i) +
                               (lf[2][k + 1] - lv[2][k]) * A(b, 2, k + 1, j, i)) * dlfi[2][k];
        else
          A(u, BZ, k, j, i) = A(b, 2, k, j, i);
      }
}

/* src/mhd/eos.pyx:13-78 with src/mhd/eos_c.c:9-48 */
static void cons2prim_lims(or_sim *s, double *w, const double *u, const int *lims)
{
  double uu[NMODE], ww[NMODE];
  for (int k = lims[4]; k <= lims[5]; ++k)
    for (int j = lims[2]; j <= lims[3]; ++j)
      for (int i = lims[0]; i <= lims[1]; ++i) {
        for (int n = 0; n < NMODE; ++n) uu[n] = A(u, n, k, j, i);
        or_cons2prim(ww, uu, s->gam);
        for (int n = 0; n < NMODE; ++n) A(w, n, k, j, i) = ww[n];
      }
}

static void prim2cons_lims(or_sim *s, double *u, const double *w, const int *lims)
{
  double uu[NMODE], ww[NMODE];
  for (int k = lims[4]; k <= lims[5]; ++k)
    for (int j = lims[2]; j <= lims[3]; ++j)
      for (int i = lims[0]; i <= lims[1]; ++i) {
        for (int n = 0; n < NMODE; ++n) ww[n] = A(w, n, k, j, i);
        or_prim2cons(uu, ww, s->gam);
        for (int n = 0; n < NMODE; ++n) A(u, n, k, j, i) = uu[n];
      }
}

/* ------------------------------------------------------------------------ */
/* Ghost fill: src/bc/grid_bc.pyx:612-623 (z, y, x; left then right),        */
/* periodic copy_layer_* and outflow prolong_* from src/bc/util_bc.pyx,      */
/* reflect from grid_bc_funcs.pyx:203-330 with util_bc.pyx:48-90,172-205.    */
/* ------------------------------------------------------------------------ */

/* One side of one axis for array A3, which is prim slot n (face = 0) or face
   component n (face = 1).
   Reflect (D15): the shipped reflect functions test an undefined name
   (bvar[n], grid_bc_funcs.pyx:213) and, through an if / if-else chain
   (:213-221), overwrite the sign-flipped copy of the normal velocity and
   normal cell field with an unflipped one.  The evident intent is followed:
   normal velocity and normal cell-centred field are mirrored with a sign
   flip, every other cell variable and the transverse face fields mirrored
   as they are (copy_reflect_layer_*, util_bc.pyx:48-90); the face-normal
   field is zeroed on the wall face (set_layer_*, util_bc.pyx:172-205) and
   mirrored about it with a sign flip, over ng faces on the left and ng-1 on
   the right (:218, :239). */
static void bc_axis(or_sim *s, double *A3, int ax, int side, int face, int n)
{
  int ng = s->ng;
  int nx = s->Ntot[0], ny = s->Ntot[1], nz = s->Ntot[2];
  int a1 = (ax == 0) ? s->i1 : (ax == 1) ? s->j1 : s->k1;
  int a2 = (ax == 0) ? s->i2 : (ax == 1) ? s->j2 : s->k2;
  int flag = s->bc[ax][side];
#define AT(k, j, i) A3[((long)(k) * ny + (j)) * nx + (i)]
#define LAYER(d, expr)                                                      \
  do {                                                                      \
    if (ax == 0) {                                                          \
      for (int k = 0; k < nz; ++k)                                          \
        for (int j = 0; j < ny; ++j) AT(k, j, d) = (expr(AT(k, j, src)));   \
    } else if (ax == 1) {                                                   \
      for (int k = 0; k < nz; ++k)                                          \
        for (int i = 0; i < nx; ++i) AT(k, d, i) = (expr(AT(k, src, i)));   \
    } else {                                                                \
      for (int j = 0; j < ny; ++j)                                          \
        for (int i = 0; i < nx; ++i) AT(d, j, i) = (expr(AT(src, j, i)));   \
    }                                                                       \
  } while (0)
#define PLUS(x) (x)
#define MINUS(x) (-1 * (x))
#define ZERO(x) (0.0)
  if (flag == 3) {
    /* user side: the fill the problem file's hook would do, given as data
       (set_bc_grid_ptrs_user, src/problem/problem.pxd:13; the kinds of fill
       of st2d.pyx:198-372: fixed values, copies from chosen cells) */
    const int *src = s->usr_src[ax][side];
    const double *fix = s->usr_fix[ax][side];
    if (!src) return;
    int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
    lo[ax] = side ? s->Ntot[ax] - ng : 0;
    hi[ax] = lo[ax] + ng;
    long e = 0;
    for (int k = lo[2]; k < hi[2]; ++k)
      for (int j = lo[1]; j < hi[1]; ++j)
        for (int i = lo[0]; i < hi[0]; ++i, ++e) {
          if (src[e] >= 0) AT(k, j, i) = A3[src[e]];
          else if (src[e] == -1) AT(k, j, i) = fix[e * (NMODE + 3) + (face ? NMODE + n : n)];
        }
    return;
  }
  if (flag == 2) {
    if (face && n == ax) {
      int src = 0;
      int wall = side == 0 ? a1 : a2 + 1;
      LAYER(wall, ZERO);
      int cnt = side == 0 ? ng : ng - 1;
      for (int g = 0; g < cnt; ++g) {
        int d = side == 0 ? a1 - 1 - g : a2 + 2 + g;
        src = side == 0 ? a1 + 1 + g : a2 - g;
        LAYER(d, MINUS);
      }
    } else {
      int odd = !face && (n == VX + ax || n == BX + ax);
      for (int g = 0; g < ng; ++g) {
        int d = side == 0 ? a1 - 1 - g : a2 + 1 + g;
        int src = side == 0 ? a1 + g : a2 - g;
        if (odd) LAYER(d, MINUS);
        else LAYER(d, PLUS);
      }
    }
  } else {
    for (int g = 0; g < ng; ++g) {
      int d, src;
      if (flag == 0) { /* periodic, grid_bc_funcs.pyx:36-105 */
        d = side == 0 ? a1 - ng + g : a2 + 1 + g;
        src = side == 0 ? a2 - ng + 1 + g : a1 + g;
      } else { /* outflow, grid_bc_funcs.pyx:112-196, util_bc.pyx:208-249 (D14) */
        d = side == 0 ? a1 - ng + g : a2 + 1 + g;
        src = side == 0 ? a1 : a2;
      }
      LAYER(d, PLUS);
    }
  }
#undef ZERO
#undef MINUS
#undef PLUS
#undef LAYER
#undef AT
}

/* narr arrays from base; the first is prim slot (or face component) n0 */
static void apply_bc_arrays(or_sim *s, double *base, int narr, int face, int n0)
{
  for (int q = 0; q < narr; ++q) {
    double *A3 = base + (long)q * s->ncell;
    int n = n0 + q;
    if (s->dims >= 3) { bc_axis(s, A3, 2, 0, face, n); bc_axis(s, A3, 2, 1, face, n); }
    if (s->dims >= 2) { bc_axis(s, A3, 1, 0, face, n); bc_axis(s, A3, 1, 1, face, n); }
    bc_axis(s, A3, 0, 0, face, n);
    bc_axis(s, A3, 0, 1, face, n);
  }
}

/* apply_grid_bc over all NVAR: prim[0..8] and bfld[0..2]
   (src/bc/grid_bc_funcs.pyx:10-29).  The per-variable ordering of the
   reference (all variables per side) gives the same result because the
   copies of different variables are independent. */
void or_apply_grid_bc(or_sim *s)
{
  apply_bc_arrays(s, s->prim, NMODE, 0, 0);
  apply_bc_arrays(s, s->bfld, 3, 1, 0);
}

/* init_sim after set_problem: src/integration/vanleer.pyx:38-48 */
void or_init(or_sim *s)
{
  int lims_act[6] = {s->i1, s->i2, s->j1, s->j2, s->k1, s->k2};
  int lims_tot[6] = {0, s->Ntot[0] - 1, 0, s->Ntot[1] - 1, 0, s->Ntot[2] - 1};
  or_apply_grid_bc(s);
  interp_bc(s, s->prim, s->bfld, lims_act);
  apply_bc_arrays(s, s->prim + (long)BX * s->ncell, 3, 0, BX);
  prim2cons_lims(s, s->cons, s->prim, lims_tot);
  s->time = 0.;
  s->step = 0;
}

/* new_dt: src/integration/new_dt.pyx:18-135 (Cartesian widths) */
double or_new_dt(or_sim *s)
{
  double cdsmax = 0.;
  double w[NMODE];
  for (int k = s->k1; k <= s->k2; ++k) {
    double cx = 0., cy = 0., cz = 0.;
    for (int j = s->j1; j <= s->j2; ++j)
      for (int i = s->i1; i <= s->i2; ++i) {
        for (int n = 0; n < NMODE; ++n) w[n] = A(s->prim, n, k, j, i);
        double bx = w[BX], by = w[BY], bz = w[BZ];
        double wy[NMODE], wz[NMODE];
        const double *py = w, *pz = w;
        if (!s->dt_fms_ref) {
          /* corrected mode: fast speed along each axis */
          memcpy(wy, w, sizeof(w));
          memcpy(wz, w, sizeof(w));
          wy[BX] = by; wy[BY] = bx;
          wz[BX] = bz; wz[BZ] = bx;
          py = wy;
          pz = wz;
        }
        /* the hydro build's fms is the sound speed (eos_c.h:523-527) */
        double vfx = s->hydro ? or_fms_hydro(w, s->gam) : or_fms(w, bx, s->gam);
        cx = fabs(w[VX]) + vfx;
        if (s->dims >= 2) {
          double vfy = s->hydro ? vfx : or_fms(py, by, s->gam);
          cy = fabs(w[VY]) + vfy;
        }
        if (s->dims >= 3) {
          double vfz = s->hydro ? vfx : or_fms(pz, bz, s->gam);
          cz = fabs(w[VZ]) + vfz;
        }
        double dsx = s->dlf[0][i], dsy = s->dlf[1][j], dsz = s->dlf[2][k];
        double cds = fmax(fmax(cx / dsx, cy / dsy), cz / dsz);
        if (cds > cdsmax) cdsmax = cds;
      }
  }
  return s->cour / cdsmax;
}

/* One VL step with a given dt: src/integration/vanleer.pyx:100-163 */
void or_vl_step(or_sim *s, double dt_full)
{
  int lims[6];
  int lims_tot[6] = {0, s->Ntot[0] - 1, 0, s->Ntot[1] - 1, 0, s->Ntot[2] - 1};
  double *w = s->prim, *u = s->cons, *b = s->bfld;
  double *ws = s->prim, *us = s->cons_s, *bs = s->bfld_s;

  /* predictor */
  int order = 0;
  double dt = 0.5 * dt_full;
  integr_lims(s, order, lims);
  godunov_fluxes(s, w, b, lims, order);
  advance_hydro(s, us, u, lims, dt);
  ec_from_prim(s, w, lims);
  if (s->ct_upwind) interp_ee2(s, lims, w, dt_full);
  else interp_ee1(s, lims);
  advance_b(s, bs, b, lims, dt);
  interp_bc(s, us, bs, lims);
  cons2prim_lims(s, ws, us, lims);

  /* corrector */
  order = s->reconstr_order;
  dt = dt_full;
  integr_lims(s, order, lims);
  godunov_fluxes(s, ws, bs, lims, order);
  advance_hydro(s, u, u, lims, dt);
  ec_from_prim(s, ws, lims);
  if (s->ct_upwind) interp_ee2(s, lims, ws, dt_full);
  else interp_ee1(s, lims);
  advance_b(s, b, b, lims, dt);
  interp_bc(s, u, b, lims);
  cons2prim_lims(s, w, u, lims);

  or_apply_grid_bc(s);
  prim2cons_lims(s, u, w, lims_tot);
}

/* combine of src/integration/_rk3.pyx:196-204: a3 = c1 a1 + c2 a2 over the
   cells of lims, nvar variables (a3 may alias a1 or a2) */
static void combine(or_sim *s, double *a3, const double *a1, const double *a2, double c1,
                    double c2, const int *lims, int nvar)
{
  for (int n = 0; n < nvar; ++n)
    for (int k = lims[4]; k <= lims[5]; ++k)
      for (int j = lims[2]; j <= lims[3]; ++j)
        for (int i = lims[0]; i <= lims[1]; ++i) {
          long c = IDX(s, n, k, j, i);
          a3[c] = c1 * a1[c] + c2 * a2[c];
        }
}

/* One stage of _rk3.pyx: integrate_hydro (u1 = u0 + dt L) and
   integrate_field (b1 = b0 + dt L, cell-centred field of u1) from the
   primitive state w and the faces b0 */
static void rk3_stage(or_sim *s, double *u1, const double *u0, double *b1, const double *b0,
                      const int *lims, double dt)
{
  double *w = s->prim;
  godunov_fluxes(s, w, b0, lims, s->reconstr_order);
  advance_hydro(s, u1, u0, lims, dt);
  ec_from_prim(s, w, lims);
  if (s->ct_upwind) interp_ee2(s, lims, w, dt);
  else interp_ee1(s, lims);
  advance_b(s, b1, b0, lims, dt);
  interp_bc(s, u1, b1, lims);
}

/* One RK3 step with a given dt: src/integration/_rk3.pyx:20-175, the
   three-stage strong-stability-preserving Runge-Kutta of Shu & Osher
   (U* = U + dt L(U); U** = 3/4 U + 1/4 (U* + dt L(U*)); U_{n+1} = 1/3 U +
   2/3 (U** + dt L(U**))), every stage at the reconstruction order and dt on
   the same limits.  The primitive array is updated in place on the limits
   after each stage (Ws and Wss alias W), with no ghost fill between the
   stages: the cells next to the limits read the step's old values, and ng
   = 9 keeps that three-stage contamination out of the active cells.  No
   sources; the user hook (do_user_work_cons) is not a part of the path. */
void or_rk3_step(or_sim *s, double dt)
{
  int lims[6];
  int lims_tot[6] = {0, s->Ntot[0] - 1, 0, s->Ntot[1] - 1, 0, s->Ntot[2] - 1};
  double *w = s->prim, *u = s->cons, *b = s->bfld;
  double *us = s->cons_s, *bs = s->bfld_s, *uss = s->cons_ss, *bss = s->bfld_ss;
  integr_lims(s, s->reconstr_order, lims);
  /* I: Us = U + dt L(W, B) */
  rk3_stage(s, us, u, bs, b, lims, dt);
  cons2prim_lims(s, w, us, lims);
  /* II: Uss = 3/4 U + 1/4 (Us + dt L(Ws, Bs)) */
  rk3_stage(s, uss, us, bss, bs, lims, dt);
  combine(s, uss, u, uss, 0.75, 0.25, lims, NMODE);
  combine(s, bss, b, bss, 0.75, 0.25, lims, 3);
  cons2prim_lims(s, w, uss, lims);
  /* III: U = 1/3 U + 2/3 (Uss + dt L(Wss, Bss)), Us holding the bracket */
  rk3_stage(s, us, uss, bs, bss, lims, dt);
  combine(s, u, u, us, 1. / 3, 1. - 1. / 3, lims, NMODE);
  combine(s, b, b, bs, 1. / 3, 1. - 1. / 3, lims, 3);
  cons2prim_lims(s, w, u, lims);
  or_apply_grid_bc(s);
  prim2cons_lims(s, u, w, lims_tot);
}

static void step_any(or_sim *s, double dt)
{
  if (s->tintegr) or_rk3_step(s, dt);
  else or_vl_step(s, dt);
}

/* vanleer.advance loop: src/integration/vanleer.pyx:87-167 (the RK3 loop of
   _rk3.pyx:56-70 takes its dt the same way) */
int or_advance(or_sim *s, double tmax, int max_steps)
{
  int n = 0;
  while (s->time < tmax && (max_steps <= 0 || n < max_steps)) {
    double dt = or_new_dt(s);
    if (s->time + dt > tmax) dt = tmax - s->time;
    s->dt = dt;
    step_any(s, dt);
    s->time += dt;
    s->step += 1;
    ++n;
  }
  return n;
}

int or_advance_steps(or_sim *s, int nsteps)
{
  for (int n = 0; n < nsteps; ++n) {
    double dt = or_new_dt(s);
    s->dt = dt;
    step_any(s, dt);
    s->time += dt;
    s->step += 1;
  }
  return nsteps;
}
#undef A
