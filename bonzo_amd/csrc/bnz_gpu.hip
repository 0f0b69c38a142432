// bnz_gpu.hip -- MI355X (gfx950) kernels and C ABI of the Van-Leer MHD step.
//
// One VL step (src/integration/vanleer.pyx:87-167) is a chain of kernels on
// one stream:
//   dt_final (dt_partial on the first step)   CFL dt, device resident
//   predictor: flux3, emf, update<pred>        (donor cell, dt/2)
//   corrector: flux3, emf, update<corr>        (PLM, dt; CFL maximum fused)
//   ghost (+ ghost_ops for reflecting sides), time_advance
// Grid arrays keep the reference's SoA layout [var][k][j][i] in HBM so every
// wavefront reads 64 consecutive doubles of one x-row.  The persistent state
// is prim (9) + face B (3); conserved variables are recomputed from prim
// where the reference reads its cons array (u_n = prim2cons(w_n) on all
// cells, vanleer.pyx:163).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bnz_step64.hpp"

using namespace bnz;

namespace {

// --------------------------------------------------------------------------
// Ghost fill (src/bc/grid_bc.pyx:612-623, util_bc.pyx).  The z, y, x copy
// sequence composes per-axis index maps (periodic wrap, outflow clamp), so
// each ghost cell gathers directly from the active cell its maps name.
__global__ __launch_bounds__(BLK) void k_ghost(Geo g, double *w, double *b,
                                               const int2 *__restrict__ pairs, int npairs,
                                               double *adv) {
  time_advance_once(adv);
  const long t = xcd_block() * BLK + threadIdx.x;
  if (t >= npairs) return;
  const int2 ds = pairs[t];  // (ghost cell, active source cell)
  const long sv = g.sv;
  // sources are active cells the fill never writes: load everything, then
  // store, so the copies do not wait on one another
  double v[NMODE + 3];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) v[n] = w[n * sv + ds.y];
#pragma unroll
  for (int n = 0; n < 3; ++n) v[NMODE + n] = b[n * sv + ds.y];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) w[n * sv + ds.x] = v[n];
#pragma unroll
  for (int n = 0; n < 3; ++n) b[n * sv + ds.x] = v[NMODE + n];
}

// Ghost cells a reflecting side touches (see GhostOp): per-class sources,
// sign flips and zeroed wall faces; with user boundaries, every entry of a
// one-side pass, fixed values (rows of fix, NMODE + 3 each) included.
// bonly stores only the cell-centred field (the second fill of the init).
__global__ __launch_bounds__(BLK) void k_ghost_ops(Geo g, double *w, double *b,
                                                   const GhostOp *__restrict__ ops, int nops,
                                                   const double *__restrict__ fix, int bonly,
                                                   double *adv) {
  time_advance_once(adv);
  const long t = xcd_block() * BLK + threadIdx.x;
  if (t >= nops) return;
  const GhostOp o = ops[t];
  const long sv = g.sv;
  double v[NMODE], f[3];
  bool cells = !(o.bits & (1u << 15));
  if (o.bits & GHOST_FIXED) {
    const double *r = fix + (long)o.src * (NMODE + 3);
#pragma unroll
    for (int n = 0; n < NMODE; ++n) v[n] = r[n];
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = r[NMODE + a];
  } else {
#pragma unroll
    for (int n = 0; n < NMODE; ++n) v[n] = cells ? w[n * sv + o.src] : 0.;
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = ((o.bits >> (12 + a)) & 1u) ? 0. : b[a * sv + o.srcf[a]];
  }
  if (cells) {
#pragma unroll
    for (int n = 0; n < NMODE; ++n)
      if (!bonly || n >= BX) w[n * sv + o.dst] = ((o.bits >> n) & 1u) ? -v[n] : v[n];
  }
  if (bonly) return;
#pragma unroll
  for (int a = 0; a < 3; ++a) b[a * sv + o.dst] = ((o.bits >> (9 + a)) & 1u) ? -f[a] : f[a];
}

// interp_bc over the active cells into prim (init only, vanleer.pyx:44)
__global__ __launch_bounds__(BLK) void k_interp_active(Geo g, double *w, const double *b, Box box) {
  long t = xcd_block() * BLK + threadIdx.x;
  if (t >= box.total()) return;
  int i, j, k;
  box_idx(box, t, i, j, k);
  const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
  const long sv = g.sv;
  w[BX * sv + c] = (g.icm[0][i] * b[c] + g.icp[0][i] * b[c + 1]) * g.dlfi[0][i];
  if (g.dims >= 2)
    w[BY * sv + c] = (g.icm[1][j] * b[sv + c] + g.icp[1][j] * b[sv + c + g.sy]) * g.dlfi[1][j];
  else
    w[BY * sv + c] = b[sv + c];
  if (g.dims >= 3)
    w[BZ * sv + c] = (g.icm[2][k] * b[2 * sv + c] + g.icp[2][k] * b[2 * sv + c + g.sz]) * g.dlfi[2][k];
  else
    w[BZ * sv + c] = b[2 * sv + c];
}

// cons = prim2cons(prim) over all cells (vanleer.pyx:163)
__global__ __launch_bounds__(BLK) void k_prim2cons(Geo g, const double *w, double *u) {
  long t = xcd_block() * BLK + threadIdx.x;
  const long ncell = (long)g.nt[0] * g.nt[1] * g.nt[2];
  if (t >= ncell) return;
  double wv[NMODE], uv[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) wv[n] = w[n * g.sv + t];
  prim2cons(uv, wv, g.gamm1i);
#pragma unroll
  for (int n = 0; n < NMODE; ++n) u[n * g.sv + t] = uv[n];
}

// --------------------------------------------------------------------------
// CFL dt (new_dt.pyx:55-117): per-block max of max(cx/dx, cy/dy, cz/dz) over
// active cells, wavefront shuffles then LDS across the block's 4 waves.

__global__ __launch_bounds__(BLK) void k_dt_partial(Geo g, const double *__restrict__ w, Box box,
                                                    int fms_ref, double *__restrict__ part) {
  __shared__ double red[BLK / 64];
  long t = xcd_block() * BLK + threadIdx.x;
  double cds = 0.;
  if (t < box.total()) {
    int i, j, k;
    box_idx(box, t, i, j, k);
    const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
    double wv[NMODE];
#pragma unroll
    for (int n = 0; n < NMODE; ++n) wv[n] = w[n * g.sv + c];
    const double cx = fabs(wv[VX]) + fms(wv, g.gam);
    double cy = 0., cz = 0.;
    if (g.dims >= 2) {
      double wy[NMODE];
#pragma unroll
      for (int n = 0; n < NMODE; ++n) wy[n] = wv[n];
      if (!fms_ref) { wy[BX] = wv[BY]; wy[BY] = wv[BX]; }
      cy = fabs(wv[VY]) + fms(wy, g.gam);
    }
    if (g.dims >= 3) {
      double wz[NMODE];
#pragma unroll
      for (int n = 0; n < NMODE; ++n) wz[n] = wv[n];
      if (!fms_ref) { wz[BX] = wv[BZ]; wz[BZ] = wv[BX]; }
      cz = fabs(wv[VZ]) + fms(wz, g.gam);
    }
    cds = fmax(fmax(cx / g.dlf[0][i], cy / g.dlf[1][j]), cz / g.dlf[2][k]);
    // NaN never wins the reference's `cds > cdsmaxloc` test
    if (!(cds > 0.)) cds = 0.;
  }
  cds = wave_max(cds);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cds;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
#pragma unroll
    for (int q = 1; q < BLK / 64; ++q) m = fmax(m, red[q]);
    part[blockIdx.x] = m;
  }
}

// The dt reduction as a kernel of its own (dt_reduce_block)
__global__ __launch_bounds__(BLK) void k_dt_final(DtFold f, double *dev) { dt_reduce_block(f, dev); }

__global__ void k_time_advance(double *dev) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && dev[0] != 0.) {
    dev[1] += dev[0];
    dev[3] += 1.0;
  }
}

// --------------------------------------------------------------------------
// Diagnostics: sums and extrema over active cells, one partial per block.
// History sums over the active cells (get_hst_vars_grid, src/io/output.pyx:
// 20-138) from u = prim2cons(prim): rho, Ek, Et, Em, E, mx, my, mz, Bx, By,
// Bz, Bx^2, By^2, Bz^2; one partial per block.
constexpr int NHIST = 14;
__global__ __launch_bounds__(BLK) void k_hist(Geo g, const double *__restrict__ w, Box box,
                                              double *__restrict__ part) {
  __shared__ double red[NHIST][BLK / 64];
  const long t = xcd_block() * BLK + threadIdx.x;
  double v[NHIST];
#pragma unroll
  for (int q = 0; q < NHIST; ++q) v[q] = 0.;
  if (t < box.total()) {
    int i, j, k;
    box_idx(box, t, i, j, k);
    const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
    double wv[NMODE], u[NMODE];
#pragma unroll
    for (int n = 0; n < NMODE; ++n) wv[n] = w[n * g.sv + c];
    prim2cons(u, wv, g.gamm1i);
    const double rhoi = 1. / u[RHO];
    const double engk = 0.5 * rhoi * (sqr(u[MX]) + sqr(u[MY]) + sqr(u[MZ]));
    const double bx2 = sqr(u[BX]), by2 = sqr(u[BY]), bz2 = sqr(u[BZ]);
    const double engm = 0.5 * (bx2 + by2 + bz2);
    v[0] = u[RHO];
    v[1] = engk;
    v[2] = u[EN] - engk - engm;
    v[3] = engm;
    v[4] = u[EN];
    v[5] = u[MX];
    v[6] = u[MY];
    v[7] = u[MZ];
    v[8] = u[BX];
    v[9] = u[BY];
    v[10] = u[BZ];
    v[11] = bx2;
    v[12] = by2;
    v[13] = bz2;
  }
#pragma unroll
  for (int q = 0; q < NHIST; ++q) {
    double x = v[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < NHIST) {
    double x = 0.;
#pragma unroll
    for (int r = 0; r < BLK / 64; ++r) x += red[threadIdx.x][r];
    part[(long)blockIdx.x * NHIST + threadIdx.x] = x;
  }
}

// Per-block partials of the diagnostics: sums of mass, mx, my, mz, E, psc
// (0-5), maxima of |div B|, |B|, -p, -rho (6-9), counts of non-finite cells,
// cells with p <= 0 and with rho <= 0 (10-12), and the largest -index of a
// failing cell (13: -1e300 if none), so the host finds the lowest index.
constexpr int NDIAG = 14;
__device__ __forceinline__ bool diag_is_max(int q) { return (q >= 6 && q <= 9) || q == 13; }
__global__ __launch_bounds__(BLK) void k_diag(Geo g, const double *__restrict__ w,
                                              const double *__restrict__ b, Box box,
                                              double *__restrict__ part) {
  __shared__ double red[NDIAG][BLK / 64];
  long t = xcd_block() * BLK + threadIdx.x;
  double v[NDIAG];
#pragma unroll
  for (int q = 0; q < NDIAG; ++q) v[q] = 0.;
  v[8] = -1e300;
  v[9] = -1e300;
  v[13] = -1e300;
  if (t < box.total()) {
    int i, j, k;
    box_idx(box, t, i, j, k);
    const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
    double wv[NMODE], u[NMODE];
    bool finite = true;
#pragma unroll
    for (int n = 0; n < NMODE; ++n) {
      wv[n] = w[n * g.sv + c];
      finite = finite && isfinite(wv[n]);
    }
    prim2cons(u, wv, g.gamm1i);
    const double vol = g.dlf[0][i] * g.dlf[1][j] * g.dlf[2][k];
#pragma unroll
    for (int q = 0; q < 6; ++q) v[q] = u[q] * vol;
    double div = (b[c + 1] - b[c]) * g.dlfi[0][i];
    if (g.dims >= 2) div += (b[g.sv + c + g.sy] - b[g.sv + c]) * g.dlfi[1][j];
    if (g.dims >= 3) div += (b[2 * g.sv + c + g.sz] - b[2 * g.sv + c]) * g.dlfi[2][k];
    v[6] = fabs(div);
    v[7] = sqrt(sqr(wv[BX]) + sqr(wv[BY]) + sqr(wv[BZ]));
    v[8] = -wv[PR];
    v[9] = -wv[RHO];
    const bool pneg = wv[PR] <= 0., rneg = wv[RHO] <= 0.;
    v[10] = finite ? 0. : 1.;
    v[11] = pneg ? 1. : 0.;
    v[12] = rneg ? 1. : 0.;
    if (!finite || pneg || rneg) v[13] = -(double)c;
  }
#pragma unroll
  for (int q = 0; q < NDIAG; ++q) {
    double x = v[q];
    const bool is_max = diag_is_max(q);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double y = __shfl_xor(x, off, 64);
      x = is_max ? fmax(x, y) : x + y;
    }
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < NDIAG) {
    const int q = threadIdx.x;
    const bool is_max = diag_is_max(q);
    double x = red[q][0];
    for (int p = 1; p < BLK / 64; ++p) x = is_max ? fmax(x, red[q][p]) : x + red[q][p];
    part[(long)blockIdx.x * NDIAG + q] = x;
  }
}

// Pencil solver entry for unit tests: [9][n] local-frame states.
template <int RS>
__global__ __launch_bounds__(BLK) void k_pencil(const double *wl_, const double *wr_,
                                                const double *bx, double *flux, long n,
                                                double gam, double gamm1i) {
  long t = xcd_block() * BLK + threadIdx.x;
  if (t >= n) return;
  double wl[NMODE], wr[NMODE], f[NMODE];
#pragma unroll
  for (int q = 0; q < NMODE; ++q) {
    wl[q] = wl_[q * n + t];
    wr[q] = wr_[q * n + t];
  }
  wl[BX] = bx[t];
  wr[BX] = bx[t];
  int reg = (RS == BNZ_RS_HLLD)   ? flux_hlld(f, wl, wr, gam, gamm1i)
            : (RS == BNZ_RS_HLLT) ? flux_hllt(f, wl, wr, gam, gamm1i)
            : (RS == BNZ_RS_HLLC) ? flux_hllc(f, wl, wr, gam, gamm1i)
                                  : flux_hll(f, wl, wr, gam, gamm1i);
  if (reg == 0) return;
#pragma unroll
  for (int q = 0; q < NMODE; ++q) flux[q * n + t] = f[q];
}

}  // namespace

// ==========================================================================
// Host side
// ==========================================================================

namespace {

// Uniform Cartesian coordinates as src/coordinates/coord.pyx:201-257 and
// src/coordinates/cartesian.pyx:18-43 compute them (one MPI block).
void set_coordinates(bnz_ctx *c) {
  for (int n = 0; n < 3; ++n) {
    const int N = c->nt[n];
    std::vector<double> lf(N + 1), lv(N);
    if (N == 1) {
      lf[0] = c->p.lmin[n];
      lf[1] = c->p.lmax[n];
    } else {
      const int ind1 = (n == 0) ? c->i1 : (n == 1) ? c->j1 : c->k1;
      const int ind2 = (n == 0) ? c->i2 : (n == 1) ? c->j2 : c->k2;
      const int nact = (n == 0) ? c->p.nx : (n == 1) ? c->p.ny : c->p.nz;
      const double dx = (c->p.lmax[n] - c->p.lmin[n]) / nact;
      const double l1 = 0 * nact * dx;
      for (int i = 0; i < N + 1; ++i) lf[i] = l1 + (i - ind1) * dx;
      if (c->p.bc[2 * n] == 0)
        for (int i = ind1 - 1; i > -1; --i) {
          const int i_ = ind2 - i + (ind1 - 1);
          const double d = lf[i_ + 1] - lf[i_];
          lf[i] = lf[i + 1] - d;
        }
      if (c->p.bc[2 * n + 1] == 0)
        for (int i = ind2 + 1; i < N; ++i) {
          const int i_ = ind1 + i - (ind2 + 1);
          const double d = lf[i_ + 1] - lf[i_];
          lf[i + 1] = lf[i] + d;
        }
    }
    c->h_dlf[n].resize(N);
    c->h_dlfi[n].resize(N);
    c->h_icm[n].resize(N);
    c->h_icp[n].resize(N);
    c->h_dlv[n].assign(N, 0.);
    for (int i = 0; i < N; ++i) {
      c->h_dlf[n][i] = lf[i + 1] - lf[i];
      c->h_dlfi[n][i] = 1. / c->h_dlf[n][i];
      lv[i] = 0.5 * (lf[i] + lf[i + 1]);
      c->h_icm[n][i] = lv[i] - lf[i];
      c->h_icp[n][i] = lf[i + 1] - lv[i];
    }
    for (int i = 1; i < N; ++i) c->h_dlv[n][i] = lv[i] - lv[i - 1];
  }
}

// The ghost fill; adv (a step's fill) is the time and step counter that its
// first launch advances.  Returns whether a launch took adv.
bool launch_ghost(bnz_ctx *c, int bonly = 0, double *adv = nullptr) {
  Launch t(c, "ghost");
  if (c->npass > 0) {  // user boundaries: one pass per side, in order
    for (int q = 0; q < 6; ++q) {
      const int n = c->pass_off[q + 1] - c->pass_off[q];
      if (n > 0) {
        hipLaunchKernelGGL(k_ghost_ops, dim3(blocks_for(n)), dim3(BLK), 0, c->stream, c->g, c->w,
                           c->b, c->gops + c->pass_off[q], n, c->ufix, bonly, adv);
        adv = nullptr;
      }
    }
    return adv == nullptr;
  }
  if (c->nghost > 0) {
    hipLaunchKernelGGL(k_ghost, dim3(blocks_for(c->nghost)), dim3(BLK), 0, c->stream, c->g, c->w,
                       c->b, c->ghost, c->nghost, adv);
    adv = nullptr;
  }
  if (c->ngops > 0) {
    hipLaunchKernelGGL(k_ghost_ops, dim3(blocks_for(c->ngops)), dim3(BLK), 0, c->stream, c->g,
                       c->w, c->b, c->gops, c->ngops, (const double *)nullptr, 0, adv);
    adv = nullptr;
  }
  return adv == nullptr;
}

// The CFL partials of the step (the last corrector's, or k_dt_partial's),
// and their reduction left to the step's first flux launch (take_dt).
void launch_dt(bnz_ctx *c, int use_fixed, double fixed_dt) {
  int nb = c->dt_parts;
  if (nb <= 0) {
    Box a = active_box(c);
    nb = blocks_for(a.total());
    Launch t(c, "dt_partial");
    hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(BLK), 0, c->stream, c->g, c->w, a,
                       cfl_mode(c), c->part);
  }
  defer_dt_final(c, nb, use_fixed, fixed_dt);
}

}  // namespace

namespace bnz {

// Leaves the reduction of npart partials pending: the step's first flux
// launch runs it (take_dt), or flush_dt as a kernel of its own.
void defer_dt_final(bnz_ctx *c, int npart, int use_fixed, double fixed_dt) {
  c->pdt.part = c->part;
  c->pdt.npart = npart;
  c->pdt.use_fixed = use_fixed;
  c->pdt.cour = c->p.cour;
  c->pdt.fixed_dt = fixed_dt;
  c->pdt_pending = true;
}

void flush_dt(bnz_ctx *c) {
  const DtFold f = take_dt(c);
  if (f.npart == 0) return;
  Launch t(c, "dt_final");
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(BLK), 0, c->stream, f, c->dev);
}

void launch_time_advance(bnz_ctx *c) {
  hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(64), 0, c->stream, c->dev);
}

void launch_ghost64(bnz_ctx *c) { launch_ghost(c); }

// A step's ghost fill, which also advances its time and step counter.
void launch_ghost_step64(bnz_ctx *c) {
  if (!launch_ghost(c, 0, c->dev)) launch_time_advance(c);
}

}  // namespace bnz

namespace {

// Precision dispatch of the step-level calls.
void any_dt(bnz_ctx *c, int use_fixed, double fixed_dt) {
  if (c->f32) f32_dt(c, use_fixed, fixed_dt);
  else launch_dt(c, use_fixed, fixed_dt);
}
void any_step(bnz_ctx *c) {
  // 1D VL steps and RK3 steps read dt in their first launch; 2D and 3D VL
  // steps reduce it in their first flux launch
  if (c->dims == 1 || c->p.tintegr) flush_dt(c);
  if (c->f32) f32_step_body(c);
  else if (c->p.precision == BNZ_PREC_FP64_FMA) fma_step_body(c);
  else launch_step_body(c);
}

// Validation, shape, index ranges and coordinates of a grid (the host part
// of BnzGrid/BnzIntegr construction); no device work.
int configure_grid(bnz_ctx *c, const bnz_params *p) {
  c->p = *p;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1) return fail(c, BNZ_EINVAL, "nx, ny, nz must be >= 1");
  if (p->nx > (1 << 30) || p->ny > (1 << 30) || p->nz > (1 << 30))
    return fail(c, BNZ_EINVAL, "grid too large: cell indices are 32-bit");
  if (p->nz > 1 && p->ny == 1) return fail(c, BNZ_EINVAL, "nz > 1 needs ny > 1 (grid.pyx:114-122)");
  if (!(p->gam > 1.0)) return fail(c, BNZ_EINVAL, "gam must be > 1");
  if (!(p->cour > 0.0)) return fail(c, BNZ_EINVAL, "cour must be > 0");
  if (p->precision != BNZ_PREC_FP64 && p->precision != BNZ_PREC_FP32C &&
      p->precision != BNZ_PREC_FP64_FMA)
    return fail(c, BNZ_EINVAL, "precision must be BNZ_PREC_FP64, BNZ_PREC_FP32C or BNZ_PREC_FP64_FMA");
  if (p->hydro != 0 && p->hydro != 1) return fail(c, BNZ_EINVAL, "hydro must be 0 or 1");
  if (p->hydro) {
    // the hydro build binds hll and hllc here (integrator.pyx:253-262)
    if (p->rsolver != BNZ_RS_HLL && p->rsolver != BNZ_RS_HLLC)
      return fail(c, BNZ_EUNSUPPORTED, "the hydro build runs hll or hllc on the device");
    if (p->precision == BNZ_PREC_FP32C)
      return fail(c, BNZ_EUNSUPPORTED, "the hydro build runs in FP64 (BNZ_PREC_FP64 or _FMA)");
    if (p->char_proj)
      return fail(c, BNZ_EUNSUPPORTED, "characteristic limiting of the hydro build is not built");
  } else if (p->rsolver == BNZ_RS_HLLC) {
    return fail(c, BNZ_EUNSUPPORTED,
                "hllc is a hydro-only solver (fluxes_c.c:230); set hydro = 1, or use hll or hlld");
  }
  if (p->rsolver != BNZ_RS_HLLD && p->rsolver != BNZ_RS_HLL && p->rsolver != BNZ_RS_HLLT &&
      p->rsolver != BNZ_RS_HLLC)
    return fail(c, BNZ_EUNSUPPORTED, "rsolver not implemented on the device (hll, hlld, hllt, hllc are)");
  if (p->rsolver == BNZ_RS_HLLT && p->precision == BNZ_PREC_FP32C)
    return fail(c, BNZ_EUNSUPPORTED, "hllt runs in FP64 only");
  if (p->reconstr != BNZ_RCN_CONST && p->reconstr != BNZ_RCN_LINEAR)
    return fail(c, BNZ_EUNSUPPORTED, "reconstr must be const or linear (weno/parab are commented out in reconstr_c.c)");
  if (p->char_proj != 0 && p->char_proj != 1) return fail(c, BNZ_EINVAL, "char_proj must be 0 or 1");
  if (p->ct_upwind != 0 && p->ct_upwind != 1) return fail(c, BNZ_EINVAL, "ct_upwind must be 0 or 1");
  if (p->tintegr != 0 && p->tintegr != 1) return fail(c, BNZ_EINVAL, "tintegr must be 0 (vl) or 1 (rk3)");
  if (p->tintegr && p->precision == BNZ_PREC_FP32C)
    return fail(c, BNZ_EUNSUPPORTED, "the RK3 integrator runs in FP64 only");
  if (p->ct_upwind && p->precision == BNZ_PREC_FP32C)
    return fail(c, BNZ_EUNSUPPORTED, "the upwind corner EMF runs in FP64 only");
  if (p->char_proj && p->precision == BNZ_PREC_FP32C)
    return fail(c, BNZ_EUNSUPPORTED, "characteristic limiting runs in FP64 only");
  for (int q = 0; q < 6; ++q)
    if (p->bc[q] != BNZ_BC_PERIODIC && p->bc[q] != BNZ_BC_OUTFLOW &&
        p->bc[q] != BNZ_BC_REFLECT && p->bc[q] != BNZ_BC_USER)
      return fail(c, BNZ_EUNSUPPORTED,
                 "boundary flags must be 0 (periodic), 1 (outflow), 2 (reflect) or 3 (user)");
  c->dims = (p->nz > 1) ? 3 : (p->ny > 1) ? 2 : 1;
  c->order = (p->reconstr == BNZ_RCN_CONST) ? 0 : 1;
  // grid.pyx:129-143: RK3's stages run without a ghost fill between them
  c->ng = p->tintegr ? 9 : (c->order == 0) ? 2 : 3;
  c->nt[0] = p->nx + 2 * c->ng;
  c->nt[1] = (c->dims >= 2) ? p->ny + 2 * c->ng : 1;
  c->nt[2] = (c->dims >= 3) ? p->nz + 2 * c->ng : 1;
  c->i1 = c->ng; c->i2 = p->nx + c->ng - 1;
  c->j1 = (c->dims >= 2) ? c->ng : 0; c->j2 = (c->dims >= 2) ? p->ny + c->ng - 1 : 0;
  c->k1 = (c->dims >= 3) ? c->ng : 0; c->k2 = (c->dims >= 3) ? p->nz + c->ng - 1 : 0;
  c->ncell = (long)c->nt[0] * c->nt[1] * c->nt[2];
  if (c->ncell >= (1L << 31)) return fail(c, BNZ_EINVAL, "grid too large: cell indices are 32-bit");
  {
    const int nact[3] = {p->nx, p->ny, p->nz};
    auto copies = [](int f) { return f == BNZ_BC_PERIODIC || f == BNZ_BC_REFLECT; };
    for (int a = 0; a < c->dims; ++a)
      if ((copies(p->bc[2 * a]) || copies(p->bc[2 * a + 1])) && nact[a] < c->ng)
        // with fewer active cells than ghost layers the reference's periodic
        // and reflect copies read ghost cells they have not filled yet
        return fail(c, BNZ_EINVAL,
                   "a periodic or reflecting axis needs at least ng active cells");
  }
  set_coordinates(c);

  return BNZ_OK;
}

// Per-axis geometry arrays (dlf, dlfi, icm, icp, dlv) as one host table,
// with the offset of each array in it.
size_t geo_table(const bnz_ctx *c, std::vector<double> &hg, size_t off[3][5]) {
  hg.clear();
  for (int a = 0; a < 3; ++a) {
    off[a][0] = hg.size(); hg.insert(hg.end(), c->h_dlf[a].begin(), c->h_dlf[a].end());
    off[a][1] = hg.size(); hg.insert(hg.end(), c->h_dlfi[a].begin(), c->h_dlfi[a].end());
    off[a][2] = hg.size(); hg.insert(hg.end(), c->h_icm[a].begin(), c->h_icm[a].end());
    off[a][3] = hg.size(); hg.insert(hg.end(), c->h_icp[a].begin(), c->h_icp[a].end());
    off[a][4] = hg.size(); hg.insert(hg.end(), c->h_dlv[a].begin(), c->h_dlv[a].end());
  }
  return hg.size();
}

// The kernels' Geo of a configured grid whose geometry table sits at c->geo.
void set_geo(bnz_ctx *c, const size_t off[3][5]) {
  Geo &g = c->g;
  for (int q = 0; q < 3; ++q) g.nt[q] = c->nt[q];
  g.dims = c->dims;
  g.sv = c->ncell;
  g.sy = c->nt[0];
  g.sz = c->nt[0] * c->nt[1];
  for (int q = 0; q < 3; ++q) {
    g.dlf[q] = c->geo + off[q][0];
    g.dlfi[q] = c->geo + off[q][1];
    g.icm[q] = c->geo + off[q][2];
    g.icp[q] = c->geo + off[q][3];
    g.dlv[q] = c->geo + off[q][4];
  }
  g.gam = c->p.gam;
  g.gamm1i = 1. / (c->p.gam - 1.);
  g.unif = 1;
  for (int q = 0; q < 3; ++q) {
    const std::vector<double> *arr[4] = {&c->h_dlf[q], &c->h_dlfi[q], &c->h_icm[q], &c->h_icp[q]};
    double *u[4] = {&g.udlf[q], &g.udlfi[q], &g.uicm[q], &g.uicp[q]};
    for (int m = 0; m < 4; ++m) {
      *u[m] = (*arr[m])[0];
      for (double v : *arr[m])
        if (v != *u[m]) g.unif = 0;
    }
  }
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
// Exact state for checkpoint/resume: a header, then the arrays the next step
// reads (FP64: prim and face B; FP32C: conserved hi, lo, FP32 prim, face B),
// all cells.  The header carries time, step, last dt and the CFL maximum the
// last corrector left for the next step, so a resumed run takes the same dt.
struct StateHeader {
  char magic[8];
  int precision, nt[3];
  double time, step, dt, cflmax;
};

namespace {
struct StatePart {
  void *dev;
  size_t bytes;
};
// bnz_advance's captured chunk: its key and its release
bnz_ctx::ChunkKey chunk_key(const bnz_ctx *c) {
  return bnz_ctx::ChunkKey{c->w,          c->ws,         c->b,
                           c->bs,         c->part,       c->dt_parts,
                           c->flux_chunk, c->emf_chunk,  c->update_chunk};
}
void drop_chunk_graph(bnz_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

// User boundaries.  Slab of side q = 2 * axis + side: the ng ghost layers
// along the axis times the full extent of the other axes.
Box user_slab(const bnz_ctx *c, int q) {
  Box s;
  for (int d = 0; d < 3; ++d) {
    s.lo[d] = 0;
    s.n[d] = c->nt[d];
  }
  const int a = q / 2;
  s.lo[a] = (q & 1) ? c->nt[a] - c->ng : 0;
  s.n[a] = c->ng;
  return s;
}

// One gather pass per side, in the order of GridBC.apply_grid_bc (z, then
// y, then x, left before right: src/bc/grid_bc.pyx:612-623): a built-in side
// from build_ghosts with every other side left alone (its copies over the
// full extent of the other axes, as the per-side functions of
// grid_bc_funcs.pyx run), a user side from its slab entries.  No entry of a
// pass reads a cell the pass writes (built-in sources are active along the
// axis; bnz_set_user_bc checks user sources), so each pass runs its entries
// in any order.
//   A pass may read a ghost cell that a later pass writes: it then gets what
// the cell held between fills, which depends on how a step's stages use the
// ghost cells (the reference's predictor writes them in place), not on the
// boundaries.  Such a value may be overwritten later in the fill, but it may
// not end up in a ghost cell: every ghost cell's final value must trace back
// to active cells or fixed values through earlier passes ("clean").
int check_clean(bnz_ctx *c, const std::vector<GhostOp> &all, const int off[7]) {
  std::vector<signed char> st((size_t)c->ncell, -1);  // -1 unwritten, 0 stale, 1 clean
  const int lo[3] = {c->i1, c->j1, c->k1}, hi[3] = {c->i2, c->j2, c->k2};
  auto active = [&](long q) {
    const int i = (int)(q % c->nt[0]), j = (int)((q / c->nt[0]) % c->nt[1]),
              k = (int)(q / ((long)c->nt[0] * c->nt[1]));
    return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
  };
  auto clean_src = [&](int q) { return active(q) || st[q] == 1; };
  std::vector<signed char> res;
  for (int t = 0; t < 6; ++t) {
    res.assign(off[t + 1] - off[t], 0);
    for (int e = off[t]; e < off[t + 1]; ++e) {
      const GhostOp &o = all[e];
      bool ok = (o.bits & GHOST_FIXED) != 0;
      if (!ok) {
        ok = (o.bits & (1u << 15)) || clean_src(o.src);
        for (int a = 0; a < 3; ++a) ok = ok && (((o.bits >> (12 + a)) & 1u) || clean_src(o.srcf[a]));
      }
      res[e - off[t]] = ok ? 1 : 0;
    }
    for (int e = off[t]; e < off[t + 1]; ++e) st[all[e].dst] = res[e - off[t]];
  }
  for (long q = 0; q < c->ncell; ++q)
    if (!active(q) && st[q] != 1) {
      const int i = (int)(q % c->nt[0]), j = (int)((q / c->nt[0]) % c->nt[1]),
                k = (int)(q / ((long)c->nt[0] * c->nt[1]));
      return fail(c, BNZ_EINVAL,
                  std::string("user boundaries: ghost cell (") + std::to_string(i) + ", " +
                      std::to_string(j) + ", " + std::to_string(k) + ") " +
                      (st[q] < 0 ? "is written by no side"
                                 : "would take a value read from a ghost cell that a later side "
                                   "fills (sides fill z1, z2, y1, y2, x1, x2 in turn)"));
    }
  return BNZ_OK;
}

// A new device buffer *dst holding v (at least one element's room); on a
// failure *dst may hold the allocation, which the caller frees.
template <typename T>
int dev_upload(bnz_ctx *c, T **dst, const std::vector<T> &v) {
  HIPCHK(c, hipMalloc(dst, std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!v.empty())
    HIPCHK(c, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return BNZ_OK;
}

int build_passes(bnz_ctx *c) {
  static const int order[6] = {4, 5, 2, 3, 0, 1};
  std::vector<GhostOp> all;
  std::vector<double> fix;
  int off[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int t = 0; t < 6; ++t) {
    const int q = order[t];
    off[t] = (int)all.size();
    if (q / 2 >= c->dims) continue;
    const int flag = c->p.bc[q];
    if (flag != BNZ_BC_USER) {
      int one[6] = {-1, -1, -1, -1, -1, -1};
      one[q] = flag;
      std::vector<int2> pairs;
      std::vector<GhostOp> ops;
      build_ghosts(c->nt, c->dims, c->ng, one, 0, pairs, ops);
      for (const int2 &d : pairs) {
        GhostOp o;
        o.dst = d.x;
        o.src = o.srcf[0] = o.srcf[1] = o.srcf[2] = d.y;
        o.bits = 0;
        all.push_back(o);
      }
      all.insert(all.end(), ops.begin(), ops.end());
      continue;
    }
    const Box sl = user_slab(c, q);
    const std::vector<int> &src = c->usr_src[q];
    for (long e = 0; e < sl.total(); ++e) {
      if (src[e] == -2) continue;
      const int i = sl.lo[0] + (int)(e % sl.n[0]);
      const int j = sl.lo[1] + (int)((e / sl.n[0]) % sl.n[1]);
      const int k = sl.lo[2] + (int)(e / ((long)sl.n[0] * sl.n[1]));
      GhostOp o;
      o.dst = (int)(((long)k * c->nt[1] + j) * c->nt[0] + i);
      o.bits = 0;
      if (src[e] >= 0) {
        o.src = o.srcf[0] = o.srcf[1] = o.srcf[2] = src[e];
      } else {
        o.src = (int)(fix.size() / (NMODE + 3));
        o.srcf[0] = o.srcf[1] = o.srcf[2] = 0;
        o.bits = GHOST_FIXED;
        fix.insert(fix.end(), c->usr_fix[q].begin() + e * (NMODE + 3),
                   c->usr_fix[q].begin() + (e + 1) * (NMODE + 3));
      }
      all.push_back(o);
    }
  }
  off[6] = (int)all.size();
  if (const int rc = check_clean(c, all, off)) return rc;
  // the new device tables are complete before the old ones go, so that a
  // failed allocation, copy or launch leaves the boundaries as they were
  GhostOp *gops = nullptr;
  double *ufix = nullptr;
  float *ufix32 = nullptr;
  int rc = dev_upload(c, &gops, all);
  if (!rc) rc = dev_upload(c, &ufix, fix);
  if (!rc && c->f32) rc = f32_fixed_rows(c, ufix, (int)(fix.size() / (NMODE + 3)), &ufix32);
  // the old tables may be in use by queued fills
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess)
    rc = fail(c, BNZ_EHIP, "user boundary: hipStreamSynchronize failed");
  if (rc) {
    if (gops) (void)hipFree(gops);
    if (ufix) (void)hipFree(ufix);
    if (ufix32) (void)hipFree(ufix32);
    return rc;
  }
  drop_chunk_graph(c);
  c->has_steady = false;
  if (c->gops) (void)hipFree(c->gops);
  if (c->ufix) (void)hipFree(c->ufix);
  if (c->ufix32) (void)hipFree(c->ufix32);
  c->gops = gops;
  c->ufix = ufix;
  c->ufix32 = ufix32;
  c->ngops = (int)all.size();
  c->nghost = 0;
  for (int t = 0; t < 7; ++t) c->pass_off[t] = off[t];
  c->npass = 6;
  return BNZ_OK;
}

int user_sides_ready(bnz_ctx *c) {
  static const char *names[6] = {"x1", "x2", "y1", "y2", "z1", "z2"};
  for (int q = 0; q < 2 * c->dims; ++q)
    if (c->p.bc[q] == BNZ_BC_USER && !c->usr_set[q])
      return fail(c, BNZ_ESTATE,
                  std::string("user boundary of side ") + names[q] +
                      " not set (bnz_set_user_bc)");
  return BNZ_OK;
}

std::vector<StatePart> state_parts(bnz_ctx *c) {
  const size_t n = (size_t)c->ncell;
  if (c->f32)
    return {{c->uh, 6 * n * sizeof(float)}, {c->ul, 6 * n * sizeof(float)},
            {c->w32, 9 * n * sizeof(float)}, {c->b, 3 * n * sizeof(double)}};
  return {{c->w, 9 * n * sizeof(double)}, {c->b, 3 * n * sizeof(double)}};
}
}  // namespace

extern "C" {

const char *bnz_last_error(const bnz_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bnz_create(const bnz_params *p, bnz_ctx **out) {
  if (!p || !out) return BNZ_EINVAL;
  *out = nullptr;
  bnz_ctx *c = new bnz_ctx();
  // on a configuration error the context is kept so that the caller can
  // read the message, then destroy it
  *out = c;
  {
    const int rc = configure_grid(c, p);
    if (rc) return rc;
  }
  if (hipSetDevice(p->device) != hipSuccess) return fail(c, BNZ_EHIP, "hipSetDevice failed");
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(c, hipHostMalloc(&c->hbuf, 8 * sizeof(double), hipHostMallocDefault));
  for (int q = 0; q < 2; ++q) HIPCHK(c, hipEventCreateWithFlags(&c->hev[q], hipEventDisableTiming));
  const size_t cells = (size_t)c->ncell;
  c->f32 = (p->precision == BNZ_PREC_FP32C);
  // FP64 prim and faces always exist (upload, init, download); the FP64
  // work arrays only for the FP64 step
  std::vector<double **> bufs9 = {&c->w}, bufs3 = {&c->b, &c->bs};
  if (!c->f32) {
    bufs9.insert(bufs9.end(), {&c->ws, &c->fx, &c->fy, &c->fz});
    bufs3.insert(bufs3.end(), {&c->ee});
  }
  if (p->tintegr) {
    bufs9.insert(bufs9.end(), {&c->uc, &c->uss});
    bufs3.insert(bufs3.end(), {&c->bss});
  }
  for (double **q : bufs9)
    if (hipMalloc(q, NVAR_STATE * cells * sizeof(double)) != hipSuccess)
      return fail(c, BNZ_ENOMEM, "device allocation of grid arrays failed");
  for (double **q : bufs3)
    if (hipMalloc(q, 3 * cells * sizeof(double)) != hipSuccess)
      return fail(c, BNZ_ENOMEM, "device allocation of field arrays failed");
  // zero everything once so stale reads of never-written ghost faces are defined
  for (double **q : bufs9) HIPCHK(c, hipMemsetAsync(*q, 0, NVAR_STATE * cells * sizeof(double), c->stream));
  for (double **q : bufs3) HIPCHK(c, hipMemsetAsync(*q, 0, 3 * cells * sizeof(double), c->stream));
  if (c->f32) {
    int rc = f32_alloc(c);
    if (rc) return rc;
  }

  std::vector<double> hg;
  size_t off[3][5];
  const size_t geo_n = geo_table(c, hg, off);
  HIPCHK(c, hipMalloc(&c->geo, geo_n * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->geo, hg.data(), geo_n * sizeof(double), hipMemcpyHostToDevice));

  {
    // (ghost, source) pairs of every ghost cell, from the per-axis maps
    // (user sides are left alone here: bnz_set_user_bc builds the passes)
    std::vector<int2> pairs;
    std::vector<GhostOp> ops;
    int bc0[6];
    for (int q = 0; q < 6; ++q) bc0[q] = p->bc[q] == BNZ_BC_USER ? -1 : p->bc[q];
    build_ghosts(c->nt, c->dims, c->ng, bc0, 0, pairs, ops);
    c->nghost = (int)pairs.size();
    c->ngops = (int)ops.size();
    HIPCHK(c, hipMalloc(&c->ghost, std::max<size_t>(1, pairs.size()) * sizeof(int2)));
    if (!pairs.empty())
      HIPCHK(c, hipMemcpy(c->ghost, pairs.data(), pairs.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHK(c, hipMalloc(&c->gops, std::max<size_t>(1, ops.size()) * sizeof(GhostOp)));
    if (!ops.empty())
      HIPCHK(c, hipMemcpy(c->gops, ops.data(), ops.size() * sizeof(GhostOp), hipMemcpyHostToDevice));
  }

  HIPCHK(c, hipMalloc(&c->dev, 8 * sizeof(double)));
  const double dev0[8] = {0., 0., -1., 0., 0., 0., 0., 0.};
  HIPCHK(c, hipMemcpy(c->dev, dev0, sizeof(dev0), hipMemcpyHostToDevice));
  Box a = active_box(c);
  {
    Box grid;
    for (int q = 0; q < 3; ++q) {
      grid.lo[q] = 0;
      grid.n[q] = c->nt[q];
    }
    c->npart_cap = std::max({(long)blocks_for(a.total()) * std::max(NDIAG, NHIST) + 64,
                             (long)blocks_for(c->ncell), (long)tiled_blocks(grid, 1, 2),
                             (long)(c->nt[0] + STEP1_CELLS - 1) / STEP1_CELLS});
  }
  HIPCHK(c, hipMalloc(&c->part, c->npart_cap * sizeof(double)));
  // diag and history reduce into their own buffer, so the CFL partials the
  // last corrector left stay valid and the next dt does not depend on them
  HIPCHK(c, hipMalloc(&c->rpart, (size_t)blocks_for(a.total()) * std::max(NDIAG, NHIST) *
                                     sizeof(double)));

  set_geo(c, off);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BNZ_OK;
}

int bnz_destroy(bnz_ctx *c) {
  if (!c) return BNZ_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  collect_timing(c);
  double *bufs[] = {c->w,   c->ws,  c->b,    c->bs,  c->fx,  c->fy, c->fz, c->ee,
                    c->geo, c->dev, c->part, c->rpart, c->uc, c->uss, c->bss};
  for (double *q : bufs)
    if (q) (void)hipFree(q);
  drop_chunk_graph(c);
  if (c->ghost) (void)hipFree(c->ghost);
  if (c->gops) (void)hipFree(c->gops);
  if (c->ufix) (void)hipFree(c->ufix);
  f32_free(c);
  for (hipEvent_t e : c->hev)
    if (e) (void)hipEventDestroy(e);
  if (c->hbuf) (void)hipHostFree(c->hbuf);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return BNZ_OK;
}

int bnz_get_shape(const bnz_ctx *c, int *ntot3, int *ng) {
  if (!c) return BNZ_EINVAL;
  if (ntot3)
    for (int q = 0; q < 3; ++q) ntot3[q] = c->nt[q];
  if (ng) *ng = c->ng;
  return BNZ_OK;
}

int bnz_upload(bnz_ctx *c, const double *prim, const double *bfld) {
  if (!c || !c->stream) return BNZ_EINVAL;
  c->dt_parts = 0;
  c->pdt_pending = false;
  // ordered after the steps still running on the context's stream (a plain
  // hipMemcpy would not wait for a non-blocking stream)
  if (c->p.hydro && bfld)  // the hydro build has no field
    for (long q = 0; q < 3 * c->ncell; ++q)
      if (bfld[q] != 0.)
        return fail(c, BNZ_EINVAL, "hydro build: every face field value must be zero");
  if (prim)
    HIPCHK(c, hipMemcpyAsync(c->w, prim, NMODE * c->ncell * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
  if (bfld)
    HIPCHK(c, hipMemcpyAsync(c->b, bfld, 3 * c->ncell * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
  else if (c->p.hydro)
    HIPCHK(c, hipMemsetAsync(c->b, 0, 3 * c->ncell * sizeof(double), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the uploaded state has no ghosts, cell-centred field or FP32C words yet
  c->ready = false;
  return BNZ_OK;
}

int bnz_download(bnz_ctx *c, double *prim, double *bfld, double *cons) {
  if (!c || !c->stream) return BNZ_EINVAL;
  if (c->f32 && c->ready) {
    int rc = f32_to64(c);
    if (rc) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (prim) HIPCHK(c, hipMemcpy(prim, c->w, NMODE * c->ncell * sizeof(double), hipMemcpyDeviceToHost));
  if (bfld) HIPCHK(c, hipMemcpy(bfld, c->b, 3 * c->ncell * sizeof(double), hipMemcpyDeviceToHost));
  if (cons) {
    // ws is dead between steps; it holds prim2cons(prim) for the copy (the
    // FP32 step has no FP64 ws, so a temporary is used)
    double *tmp = c->ws;
    if (!tmp) HIPCHK(c, hipMalloc(&tmp, NMODE * c->ncell * sizeof(double)));
    hipLaunchKernelGGL(k_prim2cons, dim3(blocks_for(c->ncell)), dim3(BLK), 0, c->stream, c->g,
                       c->w, tmp);
    int rc = check_launch(c);
    if (!rc) {
      hipError_t e1 = hipStreamSynchronize(c->stream);
      hipError_t e2 = hipMemcpy(cons, tmp, NMODE * c->ncell * sizeof(double), hipMemcpyDeviceToHost);
      if (e1 != hipSuccess || e2 != hipSuccess) rc = fail(c, BNZ_EHIP, "copy of cons failed");
    }
    if (tmp != c->ws) (void)hipFree(tmp);
    if (rc) return rc;
  }
  return BNZ_OK;
}

int bnz_fill_ghosts(bnz_ctx *c) {
  if (!c || !c->stream) return BNZ_EINVAL;
  if (const int rc = user_sides_ready(c)) return rc;
  if (c->f32 && c->ready) f32_ghost(c);
  else launch_ghost(c);
  return check_launch(c);
}

int bnz_init_state(bnz_ctx *c) {
  if (!c || !c->stream) return BNZ_EINVAL;
  if (const int rc = user_sides_ready(c)) return rc;
  c->dt_parts = 0;
  c->pdt_pending = false;
  launch_ghost(c);
  Box a = active_box(c);
  hipLaunchKernelGGL(k_interp_active, dim3(blocks_for(a.total())), dim3(BLK), 0, c->stream,
                     c->g, c->w, c->b, a);
  // the cell-centred field again (vanleer.pyx:44-46); without user sides
  // every source is active and refilling all variables gives the same
  launch_ghost(c, 1);
  if (c->f32) {
    int rc = f32_from64(c);
    if (rc) return rc;
  }
  const double dev0[4] = {0., 0., -1., 0.};
  HIPCHK(c, hipMemcpyAsync(c->dev, dev0, sizeof(dev0), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ready = true;
  return check_launch(c);
}

int bnz_new_dt(bnz_ctx *c, double *dt) {
  if (!c || !dt) return BNZ_EINVAL;
  if (!c->ready) return fail(c, BNZ_ESTATE, "bnz_init_state has not run");
  const double nomax = -1.;
  HIPCHK(c, hipMemcpyAsync(c->dev + 2, &nomax, sizeof(double), hipMemcpyHostToDevice, c->stream));
  any_dt(c, 0, 0.);
  flush_dt(c);
  int rc = check_launch(c);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(dt, c->dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BNZ_OK;
}

int bnz_vl_step(bnz_ctx *c, double dt) {
  if (!c) return BNZ_EINVAL;
  if (!c->ready) return fail(c, BNZ_ESTATE, "bnz_init_state has not run");
  if (!(dt > 0.)) return fail(c, BNZ_EINVAL, "dt must be > 0");
  any_dt(c, 1, dt);
  any_step(c);
  return check_launch(c);
}

int bnz_run_steps(bnz_ctx *c, int nsteps) {
  if (!c) return BNZ_EINVAL;
  if (!c->ready) return fail(c, BNZ_ESTATE, "bnz_init_state has not run");
  const double nomax = -1.;
  HIPCHK(c, hipMemcpyAsync(c->dev + 2, &nomax, sizeof(double), hipMemcpyHostToDevice, c->stream));
  for (int s = 0; s < nsteps; ++s) {
    any_dt(c, 0, 0.);
    any_step(c);
  }
  return check_launch(c);
}

int bnz_advance(bnz_ctx *c, double tmax, int max_steps, int *nsteps) {
  if (!c) return BNZ_EINVAL;
  if (!c->ready) return fail(c, BNZ_ESTATE, "bnz_init_state has not run");
  HIPCHK(c, hipMemcpyAsync(c->dev + 2, &tmax, sizeof(double), hipMemcpyHostToDevice, c->stream));
  double d0[4];
  HIPCHK(c, hipMemcpyAsync(d0, c->dev, sizeof(d0), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int n = 0;
  if (d0[1] < tmax) {
    // Steps go out in chunks; the host reads the time after a chunk while
    // the next one runs.  A step that starts at t >= t_max gets dt = 0 and
    // changes nothing (k_dt_final, k_update), so a chunk may overrun t_max.
    // With step graphs on, a full chunk in the steady state (kernel timing
    // off, and buffers, CFL partial count and work shapes the same before
    // and after a full eager chunk) is captured once in a hipGraph and
    // replayed: its launches carry only device pointers and sizes, and the
    // buffers it swaps are back in place after its even number of steps.
    // Right after an upload, a state load or a change of work shapes the
    // first chunk runs eagerly.
    constexpr int CHUNK = 8;
    int launched = 0;
    auto chunk = [&](int slot) -> int {
      const bool full = max_steps <= 0 || launched + CHUNK <= max_steps;
      const bnz_ctx::ChunkKey key = chunk_key(c);
      const bool replay = c->gexec && key == c->gkey;
      if (c->step_graph && full && !c->timing &&
          (replay || (c->has_steady && key == c->steady))) {
        if (!replay) drop_chunk_graph(c);
        if (!c->gexec) {
          HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
          for (int r = 0; r < CHUNK; ++r) {
            any_dt(c, 0, 0.);
            any_step(c);
          }
          hipGraph_t g = nullptr;
          const hipError_t e = hipStreamEndCapture(c->stream, &g);
          if (e != hipSuccess)
            return fail(c, BNZ_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
          c->graph = g;
          HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
          c->gkey = key;
        }
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        launched += CHUNK;
        if (!(chunk_key(c) == key)) {
          // the steps moved the context off the key: this launch holds the
          // state the host now points at, but the graph cannot be replayed
          HIPCHK(c, hipStreamSynchronize(c->stream));
          drop_chunk_graph(c);
          c->has_steady = false;
        }
      } else {
        int r = 0;
        for (; r < CHUNK && (max_steps <= 0 || launched < max_steps); ++r, ++launched) {
          any_dt(c, 0, 0.);
          any_step(c);
        }
        if (r == CHUNK && chunk_key(c) == key) {
          c->steady = key;
          c->has_steady = true;
        }
      }
      int rc = check_launch(c);
      if (rc) return rc;
      HIPCHK(c, hipMemcpyAsync(c->hbuf + 4 * slot, c->dev, 4 * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipEventRecord(c->hev[slot], c->stream));
      return BNZ_OK;
    };
    int slot = 0;
    int rc = chunk(slot);
    if (rc) return rc;
    for (;;) {
      const bool more = max_steps <= 0 || launched < max_steps;
      if (more && (rc = chunk(slot ^ 1)) != 0) return rc;
      HIPCHK(c, hipEventSynchronize(c->hev[slot]));
      if (!(c->hbuf[4 * slot + 1] < tmax) || !more) break;
      slot ^= 1;
    }
    double d1[4];
    HIPCHK(c, hipMemcpyAsync(d1, c->dev, sizeof(d1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    n = (int)(d1[3] - d0[3]);
  }
  if (nsteps) *nsteps = n;
  return BNZ_OK;
}

int bnz_get_time(bnz_ctx *c, double *time, int *step, double *last_dt) {
  if (!c) return BNZ_EINVAL;
  double d[8];
  HIPCHK(c, hipMemcpyAsync(d, c->dev, sizeof(d), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (time) *time = d[1];
  if (step) *step = (int)d[3];
  if (last_dt) *last_dt = d[7];
  return BNZ_OK;
}

int bnz_set_time(bnz_ctx *c, double time, int step) {
  if (!c) return BNZ_EINVAL;
  double d[2] = {time, 0.};
  HIPCHK(c, hipMemcpyAsync(c->dev + 1, &d[0], sizeof(double), hipMemcpyHostToDevice, c->stream));
  d[1] = (double)step;
  HIPCHK(c, hipMemcpyAsync(c->dev + 3, &d[1], sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BNZ_OK;
}

int bnz_state_size(const bnz_ctx *c, long *bytes) {
  if (!c || !bytes) return BNZ_EINVAL;
  size_t t = sizeof(StateHeader);
  for (auto &q : state_parts(const_cast<bnz_ctx *>(c))) t += q.bytes;
  *bytes = (long)t;
  return BNZ_OK;
}

int bnz_save_state(bnz_ctx *c, void *host, long bytes) {
  if (!c || !host) return BNZ_EINVAL;
  if (!c->ready) return fail(c, BNZ_ESTATE, "bnz_init_state has not run");
  long need = 0;
  bnz_state_size(c, &need);
  if (bytes < need) return fail(c, BNZ_EINVAL, "state buffer too small");
  StateHeader h;
  std::memcpy(h.magic, "BNZSTAT1", 8);
  h.precision = c->p.precision;
  for (int a = 0; a < 3; ++a) h.nt[a] = c->nt[a];
  double d[8];
  HIPCHK(c, hipMemcpyAsync(d, c->dev, sizeof(d), hipMemcpyDeviceToHost, c->stream));
  h.cflmax = -1.;
  std::vector<double> parts((size_t)std::max(c->dt_parts, 0));
  if (c->dt_parts > 0)
    HIPCHK(c, hipMemcpyAsync(parts.data(), c->part, parts.size() * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (double x : parts) h.cflmax = std::fmax(h.cflmax, x);
  h.dt = d[7];
  h.time = d[1];
  h.step = d[3];
  char *out = (char *)host;
  std::memcpy(out, &h, sizeof(h));
  out += sizeof(h);
  for (auto &q : state_parts(c)) {
    HIPCHK(c, hipMemcpy(out, q.dev, q.bytes, hipMemcpyDeviceToHost));
    out += q.bytes;
  }
  return BNZ_OK;
}

int bnz_load_state(bnz_ctx *c, const void *host, long bytes) {
  if (!c || !host) return BNZ_EINVAL;
  long need = 0;
  bnz_state_size(c, &need);
  if (bytes < need) return fail(c, BNZ_EINVAL, "state buffer too small for this context");
  StateHeader h;
  std::memcpy(&h, host, sizeof(h));
  if (std::memcmp(h.magic, "BNZSTAT1", 8) != 0) return fail(c, BNZ_EINVAL, "not a state buffer");
  if (h.precision != c->p.precision || h.nt[0] != c->nt[0] || h.nt[1] != c->nt[1] ||
      h.nt[2] != c->nt[2])
    return fail(c, BNZ_EINVAL, "state was saved by a context of another shape or precision");
  const char *in = (const char *)host + sizeof(h);
  for (auto &q : state_parts(c)) {
    HIPCHK(c, hipMemcpyAsync(q.dev, in, q.bytes, hipMemcpyHostToDevice, c->stream));
    in += q.bytes;
  }
  const double d[4] = {h.dt, h.time, -1., h.step};
  HIPCHK(c, hipMemcpyAsync(c->dev, d, sizeof(d), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->dev + 7, &h.dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (h.cflmax >= 0.) {
    HIPCHK(c, hipMemcpyAsync(c->part, &h.cflmax, sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    c->dt_parts = 1;
    c->pdt_pending = false;
  } else {
    c->dt_parts = 0;
    c->pdt_pending = false;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ready = true;
  return BNZ_OK;
}

int bnz_diag(bnz_ctx *c, bnz_diag_out *out) {
  if (!c || !out) return BNZ_EINVAL;
  if (c->f32 && c->ready) {
    int rc0 = f32_to64(c);
    if (rc0) return rc0;
  }
  Box a = active_box(c);
  const int nb = blocks_for(a.total());
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(BLK), 0, c->stream, c->g, c->w, c->b, a, c->rpart);
  int rc = check_launch(c);
  if (rc) return rc;
  std::vector<double> h((size_t)nb * NDIAG);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->rpart, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double s[NDIAG] = {0};
  s[6] = s[7] = 0.;
  s[8] = s[9] = s[13] = -1e300;
  for (int q = 0; q < nb; ++q)
    for (int v = 0; v < NDIAG; ++v) {
      const double x = h[(size_t)q * NDIAG + v];
      if ((v >= 6 && v <= 9) || v == 13) s[v] = std::fmax(s[v], x);
      else s[v] += x;
    }
  out->mass = s[0];
  out->mom[0] = s[1];
  out->mom[1] = s[2];
  out->mom[2] = s[3];
  out->energy = s[4];
  out->scalar = s[5];
  out->divb_max = s[6];
  out->bmax = s[7];
  out->pmin = -s[8];
  out->rhomin = -s[9];
  out->nonfinite = (long)s[10];
  out->npneg = (long)s[11];
  out->nrhoneg = (long)s[12];
  out->bad_cell = -1;
  out->bad_kind = 0;
  for (int a = 0; a < 3; ++a) out->bad_ijk[a] = -1;
  if (s[13] > -1e299) {  // the failing cell of lowest index, and what fails there
    const long cell = (long)(-s[13]);
    out->bad_cell = cell;
    out->bad_ijk[0] = (int)(cell % c->nt[0]);
    out->bad_ijk[1] = (int)((cell / c->nt[0]) % c->nt[1]);
    out->bad_ijk[2] = (int)(cell / ((long)c->nt[0] * c->nt[1]));
    double wv[NMODE];
    for (int n = 0; n < NMODE; ++n)
      HIPCHK(c, hipMemcpy(&wv[n], c->w + n * c->ncell + cell, sizeof(double),
                          hipMemcpyDeviceToHost));
    for (int n = 0; n < NMODE; ++n)
      if (!std::isfinite(wv[n])) out->bad_kind |= 1;
    if (wv[PR] <= 0.) out->bad_kind |= 2;
    if (wv[RHO] <= 0.) out->bad_kind |= 4;
  }
  int ng = 0;
  HIPCHK(c, hipMemcpy(&ng, (int *)(c->dev + 4), sizeof(int), hipMemcpyDeviceToHost));
  out->hlld_guards = ng;
  double d[4];
  HIPCHK(c, hipMemcpy(d, c->dev, sizeof(d), hipMemcpyDeviceToHost));
  out->time = d[1];
  out->step = (int)d[3];
  return BNZ_OK;
}

int bnz_history(bnz_ctx *c, double *out16) {
  if (!c || !out16) return BNZ_EINVAL;
  if (!c->ready) return fail(c, BNZ_ESTATE, "bnz_init_state has not run");
  if (c->f32) {
    int rc0 = f32_to64(c);
    if (rc0) return rc0;
  }
  Box a = active_box(c);
  const int nb = blocks_for(a.total());
  hipLaunchKernelGGL(k_hist, dim3(nb), dim3(BLK), 0, c->stream, c->g, c->w, a, c->rpart);
  int rc = check_launch(c);
  if (rc) return rc;
  std::vector<double> h((size_t)nb * NHIST);
  double d[4];
  HIPCHK(c, hipMemcpyAsync(h.data(), c->rpart, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(d, c->dev, sizeof(d), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double sum[NHIST] = {0};
  for (int q = 0; q < nb; ++q)
    for (int v = 0; v < NHIST; ++v) sum[v] += h[(size_t)q * NHIST + v];
  const double ncells = (double)c->p.nx * c->p.ny * c->p.nz;
  out16[0] = d[3];
  out16[1] = d[1];
  for (int v = 0; v < NHIST; ++v) out16[2 + v] = sum[v] / ncells;
  for (int q = 13; q < 16; ++q) out16[q] = std::sqrt(std::fabs(out16[q] - out16[q - 3] * out16[q - 3]));
  return BNZ_OK;
}

int bnz_synchronize(bnz_ctx *c) {
  if (!c) return BNZ_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return BNZ_OK;
}

int bnz_set_tuning(bnz_ctx *c, int flux_chunk, int emf_chunk, int update_chunk) {
  if (!c || flux_chunk < 0 || emf_chunk < 0 || update_chunk < 0) return BNZ_EINVAL;
  c->flux_chunk = flux_chunk;
  c->emf_chunk = emf_chunk;
  c->update_chunk = update_chunk;
  Box grid;
  for (int q = 0; q < 3; ++q) {
    grid.lo[q] = 0;
    grid.n[q] = c->nt[q];
  }
  const long need = c->f32 ? work_shape32(grid, c->dims, update_chunk).blocks
                           : work_shape(grid, c->dims, update_chunk).blocks;
  if (need > c->npart_cap) {
    double *np = nullptr;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMalloc(&np, need * sizeof(double)));
    // the last corrector's CFL partials move along: the next dt is the same
    if (c->dt_parts > 0)
      HIPCHK(c, hipMemcpy(np, c->part, c->dt_parts * sizeof(double), hipMemcpyDeviceToDevice));
    (void)hipFree(c->part);
    c->part = np;
    c->pdt_pending = false;
    c->npart_cap = need;
  }
  return BNZ_OK;
}

int bnz_set_user_bc(bnz_ctx *c, int axis, int side, const int *src, const double *fixed) {
  if (!c || !c->stream) return BNZ_EINVAL;
  if (axis < 0 || axis >= c->dims || side < 0 || side > 1 || !src)
    return fail(c, BNZ_EINVAL, "user boundary: axis must be below the grid's dimension, side 0 or 1");
  const int q = 2 * axis + side;
  if (c->p.bc[q] != BNZ_BC_USER)
    return fail(c, BNZ_EINVAL, "user boundary: the side's flag is not 3 (BNZ_BC_USER)");
  const Box sl = user_slab(c, q);
  const long n = sl.total();
  // a source may not be a cell this side writes: the pass gathers in any order
  std::vector<char> written(n, 0);
  bool any_fixed = false;
  for (long e = 0; e < n; ++e) {
    if (src[e] < -2 || src[e] >= c->ncell)
      return fail(c, BNZ_EINVAL, "user boundary: source index out of range");
    written[e] = src[e] != -2;
    any_fixed |= src[e] == -1;
  }
  if (any_fixed && !fixed)
    return fail(c, BNZ_EINVAL, "user boundary: fixed values missing for entries marked -1");
  for (long e = 0; e < n; ++e) {
    if (src[e] < 0) continue;
    const long s = src[e];
    const int i = (int)(s % c->nt[0]), j = (int)((s / c->nt[0]) % c->nt[1]),
              k = (int)(s / ((long)c->nt[0] * c->nt[1]));
    const int id[3] = {i - sl.lo[0], j - sl.lo[1], k - sl.lo[2]};
    if (id[0] >= 0 && id[0] < sl.n[0] && id[1] >= 0 && id[1] < sl.n[1] && id[2] >= 0 &&
        id[2] < sl.n[2] && written[((long)id[2] * sl.n[1] + id[1]) * sl.n[0] + id[0]])
      return fail(c, BNZ_EINVAL, "user boundary: a source is a cell the same side writes");
  }
  // a refused side leaves the boundaries as they were
  std::vector<int> old_src = std::move(c->usr_src[q]);
  std::vector<double> old_fix = std::move(c->usr_fix[q]);
  const bool old_set = c->usr_set[q];
  c->usr_src[q].assign(src, src + n);
  if (fixed) c->usr_fix[q].assign(fixed, fixed + n * (NMODE + 3));
  else c->usr_fix[q].assign(n * (NMODE + 3), 0.);
  c->usr_set[q] = true;
  for (int r = 0; r < 2 * c->dims; ++r)
    if (c->p.bc[r] == BNZ_BC_USER && !c->usr_set[r]) return BNZ_OK;
  const int rc = build_passes(c);
  if (rc) {
    c->usr_src[q] = std::move(old_src);
    c->usr_fix[q] = std::move(old_fix);
    c->usr_set[q] = old_set;
  }
  return rc;
}

int bnz_set_step_graph(bnz_ctx *c, int on) {
  if (!c || on < 0 || on > 1) return BNZ_EINVAL;
  c->step_graph = on != 0;
  if (!c->step_graph) drop_chunk_graph(c);
  return BNZ_OK;
}

int bnz_set_kernel_timing(bnz_ctx *c, int on) {
  if (!c) return BNZ_EINVAL;
  collect_timing(c);
  if (on) c->ktimes.clear();
  c->timing = on != 0;
  return BNZ_OK;
}

int bnz_kernel_timing(bnz_ctx *c, int idx, char *name, int name_len, double *total_ms,
                      long *launches) {
  if (!c) return BNZ_EINVAL;
  collect_timing(c);
  if (idx < 0 || idx >= (int)c->ktimes.size()) return BNZ_EINVAL;
  if (name && name_len > 0) {
    std::strncpy(name, c->ktimes[idx].name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (total_ms) *total_ms = c->ktimes[idx].total_ms;
  if (launches) *launches = c->ktimes[idx].launches;
  return BNZ_OK;
}

int bnz_gpu_flux_pencils(int rsolver, const double *wl, const double *wr, const double *bx,
                         double *flux, long n, double gam) {
  if (!wl || !wr || !bx || !flux || n < 0 || !(gam > 1.)) return BNZ_EINVAL;
  if (n == 0) return BNZ_OK;
  const double gm1i = 1. / (gam - 1.);
  if (rsolver == BNZ_RS_HLLD)
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLLD>, dim3(blocks_for(n)), dim3(BLK), 0, 0, wl, wr, bx,
                       flux, n, gam, gm1i);
  else if (rsolver == BNZ_RS_HLLC)
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLLC>, dim3(blocks_for(n)), dim3(BLK), 0, 0, wl, wr, bx,
                       flux, n, gam, gm1i);
  else if (rsolver == BNZ_RS_HLL)
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLL>, dim3(blocks_for(n)), dim3(BLK), 0, 0, wl, wr, bx,
                       flux, n, gam, gm1i);
  else if (rsolver == BNZ_RS_HLLT)
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLLT>, dim3(blocks_for(n)), dim3(BLK), 0, 0, wl, wr, bx,
                       flux, n, gam, gm1i);
  else
    return BNZ_EUNSUPPORTED;
  if (hipGetLastError() != hipSuccess) return BNZ_EHIP;
  if (hipDeviceSynchronize() != hipSuccess) return BNZ_EHIP;
  return BNZ_OK;
}


// The reference's Riemann-solver seam, RSolverFunc: host row pointers flux[n],
// wl[n], wr[n] (n < rows) and bx, faces i1..i2.  Each call packs the rows into
// one host block (wl 9n | wr 9n | bx n | flux 9n), copies it to the device,
// runs k_pencil there and unpacks the flux rows.  The flux rows travel to
// the device too, so a face whose solver matches no region keeps its old
// row, as in the reference.  rows is 9 (the MHD build: hll, hlld, hllt) or
// HYDRO_ROWS (the hydro build's hllc, fluxes_c.c:230-327: rho, v, p and the
// scalar, no field rows and no bx, which it never reads); the hydro rows are
// staged with a zero field, for which the device HLLC is the hydro one bit
// for bit.
constexpr int HYDRO_ROWS = 6;
static int seam_check(int rs, double **flux, double **wl, double **wr, double *bx, int i1,
                      int i2, double gam, int rows = NMODE) {
  if (!flux || !wl || !wr || i2 < i1 || !(gam > 1.)) return BNZ_EINVAL;
  if (rows == NMODE) {
    if (!bx || (rs != BNZ_RS_HLL && rs != BNZ_RS_HLLD && rs != BNZ_RS_HLLT)) return BNZ_EINVAL;
  } else if (rows != HYDRO_ROWS || rs != BNZ_RS_HLLC) {
    return BNZ_EINVAL;
  }
  for (int q = 0; q < rows; ++q)
    if (!flux[q] || !wl[q] || !wr[q]) return BNZ_EINVAL;
  return BNZ_OK;
}

constexpr int SEAM_WORDS = 28;  // doubles staged per face

// Pack, solve on the device through dev (SEAM_WORDS * n doubles) and unpack.
static int seam_solve(int rs, double **flux, double **wl, double **wr, double *bx, int i1,
                      long n, double gam, double *h, double *dev, int rows = NMODE) {
  for (int q = 0; q < NMODE; ++q)
    for (long i = 0; i < n; ++i) {
      h[q * n + i] = q < rows ? wl[q][i1 + i] : 0.;
      h[(9 + q) * n + i] = q < rows ? wr[q][i1 + i] : 0.;
      h[(19 + q) * n + i] = q < rows ? flux[q][i1 + i] : 0.;
    }
  for (long i = 0; i < n; ++i) h[18 * n + i] = bx ? bx[i1 + i] : 0.;
  if (hipMemcpy(dev, h, SEAM_WORDS * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return BNZ_EHIP;
  const double gm1i = 1. / (gam - 1.);
  const double *dwl = dev, *dwr = dev + 9 * n, *dbx = dev + 18 * n;
  double *dfl = dev + 19 * n;
  if (rs == BNZ_RS_HLLD)
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLLD>, dim3(blocks_for(n)), dim3(BLK), 0, 0, dwl, dwr, dbx,
                       dfl, n, gam, gm1i);
  else if (rs == BNZ_RS_HLLT)
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLLT>, dim3(blocks_for(n)), dim3(BLK), 0, 0, dwl, dwr, dbx,
                       dfl, n, gam, gm1i);
  else if (rs == BNZ_RS_HLLC)
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLLC>, dim3(blocks_for(n)), dim3(BLK), 0, 0, dwl, dwr, dbx,
                       dfl, n, gam, gm1i);
  else
    hipLaunchKernelGGL(k_pencil<BNZ_RS_HLL>, dim3(blocks_for(n)), dim3(BLK), 0, 0, dwl, dwr, dbx,
                       dfl, n, gam, gm1i);
  if (hipGetLastError() != hipSuccess) return BNZ_EHIP;
  if (hipMemcpy(h + 19 * n, dfl, 9 * n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return BNZ_EHIP;
  for (int q = 0; q < rows; ++q)
    for (long i = 0; i < n; ++i) flux[q][i1 + i] = h[(19 + q) * n + i];
  return BNZ_OK;
}

// Fixed staging of bnz_rsolver_rows_noalloc: one device array per GPU (each
// GPU that runs the entry point holds its own copy until the process ends)
// and one host array, sized for BNZ_NOALLOC_MAX_FACES faces.  Host code
// never passes the name `scratch` to a kernel or a copy: it stands for no
// device address on the host.  The device address of each GPU's copy comes
// from hipGetSymbolAddress, once per device, on the entry point's first call
// with that device current, and only that address is used.
static __device__ __align__(256) double scratch[(size_t)SEAM_WORDS * BNZ_NOALLOC_MAX_FACES];
static double g_noalloc_host[(size_t)SEAM_WORDS * BNZ_NOALLOC_MAX_FACES];
constexpr int NOALLOC_MAX_DEVICES = 64;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_for_current_device(double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return BNZ_EHIP;
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return BNZ_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BNZ_EHIP;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return BNZ_OK;
}

// Staging memory of the seam is kept per host thread (the reference calls
// the solver from OpenMP threads) and grows with the longest pencil.
static int pencil_rows(int rs, double **flux, double **wl, double **wr, double *bx, int i1,
                       int i2, double gam, int rows) {
  const int rc = seam_check(rs, flux, wl, wr, bx, i1, i2, gam, rows);
  if (rc != BNZ_OK) return rc;
  const long n = (long)i2 - i1 + 1;
  thread_local std::vector<double> host;
  thread_local double *dev = nullptr;
  thread_local size_t cap = 0;
  const size_t need = (size_t)SEAM_WORDS * n;
  if (need > cap) {
    if (dev) (void)hipFree(dev);
    dev = nullptr;
    cap = 0;
    if (hipMalloc(&dev, need * sizeof(double)) != hipSuccess) return BNZ_ENOMEM;
    cap = need;
  }
  host.resize(need);
  return seam_solve(rs, flux, wl, wr, bx, i1, n, gam, host.data(), dev, rows);
}

static void pencil_seam(const char *name, int rs, double **flux, double **wl, double **wr,
                        double *bx, int i1, int i2, double gam, int rows = NMODE) {
  const int rc = pencil_rows(rs, flux, wl, wr, bx, i1, i2, gam, rows);
  if (rc != BNZ_OK)  // the seam returns nothing: report, leave the flux rows as they were
    std::fprintf(stderr, "%s: faces %d..%d not computed (error %d)\n", name, i1, i2, rc);
}

void bnz_hll_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                  double gam) {
  pencil_seam("bnz_hll_flux", BNZ_RS_HLL, flux, wl, wr, bx, i1, i2, gam);
}
void bnz_hlld_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                   double gam) {
  pencil_seam("bnz_hlld_flux", BNZ_RS_HLLD, flux, wl, wr, bx, i1, i2, gam);
}
void bnz_hllt_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                   double gam) {
  pencil_seam("bnz_hllt_flux", BNZ_RS_HLLT, flux, wl, wr, bx, i1, i2, gam);
}
void bnz_hllc_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                   double gam) {
  pencil_seam("bnz_hllc_flux", BNZ_RS_HLLC, flux, wl, wr, bx, i1, i2, gam, HYDRO_ROWS);
}

int bnz_rsolver_rows_noalloc(int rsolver, double **flux, double **wl, double **wr, double *bx,
                             int i1, int i2, double gam) {
  int rc = seam_check(rsolver, flux, wl, wr, bx, i1, i2, gam);
  if (rc != BNZ_OK) return rc;
  const long n = (long)i2 - i1 + 1;
  if (n > BNZ_NOALLOC_MAX_FACES) {
    std::fprintf(stderr,
                 "bnz_rsolver_rows_noalloc: %ld faces exceed the fixed staging of %d faces; "
                 "split the pencil or call bnz_hll_flux / bnz_hlld_flux / bnz_hllt_flux\n",
                 n, BNZ_NOALLOC_MAX_FACES);
    return BNZ_ERANGE;
  }
  double *dev = nullptr;
  rc = scratch_for_current_device(&dev);
  if (rc != BNZ_OK) return rc;
  return seam_solve(rsolver, flux, wl, wr, bx, i1, n, gam, g_noalloc_host, dev);
}

}  // extern "C"
