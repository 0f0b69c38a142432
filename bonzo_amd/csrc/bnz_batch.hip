// bnz_batch.hip -- many independent boxes of uneven size advanced together,
// one launch per stage over all boxes (FP64).
//
// Every box is a complete grid of its own (sizes, ghosts, boundary flags,
// t_max) stored at an offset in shared pools.  The work of each stage
// (faces, edges or cells of every box) is cut into work blocks of BLK items
// that never straddle two boxes, so a workgroup finds its box with one
// uniform load and skips a finished box as a whole.  Launches are persistent:
// a fixed grid walks the work blocks.  The per-item arithmetic is the
// single-grid arithmetic (bnz_kern64.hpp), so a box reproduces its
// single-grid run bit for bit.  Each box keeps its own CFL dt, time and
// "running" flag on the device, so a VL step -- and, in pairs, the
// face-buffer ping-pong -- is captured once in a hipGraph and replayed until
// every box has reached its t_max (vanleer.pyx:87-96 per box).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "bnz_batch.hpp"
#include "bnz_kern64.hpp"

using namespace bnz;
using namespace bnzb;

namespace {

// The hot kernels (flux, EMF, update) are compiled per dimensionality and
// launched over the work blocks of the boxes of that dimensionality: a 1D
// box then runs 1D code with the registers 1D code needs.  Their workgroups
// take the work blocks in run_block order (runs of 16 dealt round-robin over
// the XCDs, so that an XCD walks neighbouring rows of a box): 2D blasts
// +3 % (FP64) and +5 % (FP32C) over launch order; the 1D stage kernels keep
// launch order, which is faster for them.
template <int ORDER, int RS, int DIMS>
__global__ __launch_bounds__(BLK) void kb_flux(const BoxDev *__restrict__ boxes, WList wl,
                                               BoxState st, const double *__restrict__ w,
                                               const double *__restrict__ bf,
                                               double *__restrict__ fx, double *__restrict__ fy,
                                               double *__restrict__ fz) {
  for (int wb = wl.lo + (int)run_block(); wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    if (!st.run[b]) continue;
    const BoxDev &d = boxes[b];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    const Box &it = d.items[wl.kt];
    if (t >= it.total()) continue;
    int i, j, k;
    box_idx(it, t, i, j, k);
    const long o = d.off;
    flux_cell<ORDER, RS, DIMS>(d.g, w + o, bf + o, fx + o, fy + o, fz + o,
                               d.L[wl.kt == KF0 ? 0 : 1], i, j, k);
  }
}

template <int DIMS>
__global__ __launch_bounds__(BLK) void kb_emf(const BoxDev *__restrict__ boxes, WList wl,
                                              BoxState st, const double *__restrict__ w,
                                              const double *__restrict__ fx,
                                              const double *__restrict__ fy,
                                              const double *__restrict__ fz,
                                              double *__restrict__ ee) {
  for (int wb = wl.lo + (int)run_block(); wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    if (!st.run[b]) continue;
    const BoxDev &d = boxes[b];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    const Box &it = d.items[wl.kt];
    if (t >= it.total()) continue;
    int i, j, k;
    box_idx(it, t, i, j, k);
    const long o = d.off;
    emf_edge_d<DIMS>(d.g, w + o, fx + o, fy + o, fz + o, ee + o, i, j, k);
  }
}

// Stage update of every cell of the running boxes.  The corrector also
// folds the CFL maximum of the new state into the box's cmax (one atomic per
// work block).  A finished box is left alone: its faces stay in the buffer
// its last step wrote, which kb_time_dt tracks (BoxState::fpar) and
// kb_face_fix makes current when an advance ends.
template <bool PRED, int DIMS>
__global__ __launch_bounds__(BLK) void kb_update(const BoxDev *__restrict__ boxes, WList wl,
                                                 BoxState st, Pools p, int fms_ref) {
  for (int wb = wl.lo + (int)run_block(); wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    if (!st.run[b]) continue;
    const BoxDev &d = boxes[b];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    const Box &it = d.items[wl.kt];
    const long o = d.off;
    double cds = 0.;
    if (t < it.total()) {
      int i, j, k;
      box_idx(it, t, i, j, k);
      if (PRED)
        update_cell_d<DIMS, true, false>(d.g, p.w + o, p.fx + o, p.fy + o, p.fz + o, p.ee + o,
                                         p.b + o, p.bs + o, p.ws + o, d.L[0], st.dt + b, d.act,
                                         fms_ref, i, j, k);
      else
        cds = update_cell_d<DIMS, false, true>(d.g, p.w + o, p.fx + o, p.fy + o, p.fz + o,
                                               p.ee + o, p.b + o, p.bs + o, p.w + o, d.L[1],
                                               st.dt + b, d.act, fms_ref, i, j, k);
    }
    if (!PRED) block_max_atomic(cds, st.cmax + b);
  }
}

// One stage of the 1D boxes in one pass (stage1_cell): the fluxes of BLK
// faces go to LDS, the BLK - 1 cells between them are updated from there,
// so the fluxes never go through HBM.
template <bool PRED, int ORDER, int RS>
__global__ __launch_bounds__(BLK) void kb_stage1(const BoxDev *__restrict__ boxes, WList wl,
                                                 BoxState st, Pools p, int fms_ref) {
  __shared__ double sf[NMODE * BLK];
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    if (!st.run[b]) continue;
    const BoxDev &d = boxes[b];
    const long o = d.off;
    const int i = (wb - d.blk0[wl.kt]) * (BLK - 1) + threadIdx.x;
    const double cds = stage1_cell<PRED, ORDER, RS>(
        d.g, PRED ? p.w + o : p.ws + o, p.b + o, p.bs + o, p.w + o, PRED ? p.ws + o : p.w + o,
        p.fx + o, d.L[PRED ? 0 : 1], st.dt + b, d.act, fms_ref, i, sf);
    if (!PRED)
      block_max_atomic(cds, st.cmax + b);  // ends with a barrier: sf is free again
    else
      __syncthreads();
  }
}

// A whole VL step of a 1D box in one pass (all-1D FP64 batches), work
// block by work block (step1_block): the pass reads the step's prim w and
// faces b only and writes the new ones to ws and bs, which the caller swaps
// with w and b.
template <int ORDER, int RS>
__global__ __launch_bounds__(BLK) void kb_step1(const BoxDev *__restrict__ boxes, WList wl,
                                                BoxState st, Pools p, int fms_ref) {
  __shared__ double sf[NMODE * BLK];
  __shared__ double sw[NMODE * BLK];
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    if (!st.run[b]) continue;
    const BoxDev &d = boxes[b];
    const long o = d.off;
    const double cds = step1_block<ORDER, RS>(d.g, p.w + o, p.b + o, p.ws + o, p.bs + o,
                                              p.fx + o, d.L[0], d.L[1], st.dt + b, d.act,
                                              fms_ref, (wb - d.blk0[wl.kt]) * STEP1_CELLS, sf,
                                              sw);
    block_max_atomic(cds, st.cmax + b);  // ends with a barrier: sf, sw are free again
  }
}

// CFL maximum of the current prim over the active cells (after init).
__global__ __launch_bounds__(BLK) void kb_cfl(const BoxDev *__restrict__ boxes, WList wl,
                                              BoxState st, const double *__restrict__ w,
                                              int fms_ref) {
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    const BoxDev &d = boxes[b];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    const Box &it = d.items[wl.kt];
    double cds = 0.;
    if (t < it.total()) {
      int i, j, k;
      box_idx(it, t, i, j, k);
      const long c = d.off + ((long)k * d.g.nt[1] + j) * d.g.nt[0] + i;
      double wv[NMODE];
#pragma unroll
      for (int n = 0; n < NMODE; ++n) wv[n] = w[n * d.g.sv + c];
      cds = cds_cell(d.g, wv, fms_ref, i, j, k);
    }
    block_max_atomic(cds, st.cmax + b);
  }
}

// interp_bc over the active cells (init only, vanleer.pyx:44)
__global__ __launch_bounds__(BLK) void kb_interp(const BoxDev *__restrict__ boxes, WList wl,
                                                 double *w, const double *b) {
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const BoxDev &d = boxes[wl.bmap[wb]];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    const Box &it = d.items[wl.kt];
    if (t >= it.total()) continue;
    int i, j, k;
    box_idx(it, t, i, j, k);
    const Geo &g = d.g;
    const long c = d.off + ((long)k * g.nt[1] + j) * g.nt[0] + i, sv = g.sv;
    w[BX * sv + c] = (g.icm[0][i] * b[c] + g.icp[0][i] * b[c + 1]) * g.dlfi[0][i];
    if (g.dims >= 2)
      w[BY * sv + c] = (g.icm[1][j] * b[sv + c] + g.icp[1][j] * b[sv + c + g.sy]) * g.dlfi[1][j];
    else
      w[BY * sv + c] = b[sv + c];
    if (g.dims >= 3)
      w[BZ * sv + c] =
          (g.icm[2][k] * b[2 * sv + c] + g.icp[2][k] * b[2 * sv + c + g.sz]) * g.dlfi[2][k];
    else
      w[BZ * sv + c] = b[2 * sv + c];
  }
}

// Per-box dt from the CFL maximum, clipped to t_max; a box at t_max stops
// (vanleer.pyx:87-96).  Clears the maxima for the next accumulation.
__global__ __launch_bounds__(BLK) void kb_dt(BoxState st, int nbox, double cour) {
  for (int b = threadIdx.x; b < nbox; b += BLK) {
    const double m = __longlong_as_double((long long)st.cmax[b]);
    st.cmax[b] = 0ull;
    const double t = st.t[b], tmax = st.tmax[b];
    int run = 0;
    if (t < tmax) {
      double dt = cour / m;
      if (t + dt > tmax) dt = tmax - t;
      st.dt[b] = dt;
      run = 1;
    }
    st.run[b] = run;
  }
}

// The end of one step and the dt of the next (time_dt_block); alone when a
// batch has no ghost cells, else in the last workgroup of the step's last
// ghost launch.
__global__ __launch_bounds__(BLK) void kb_time_dt(BoxState st, int nbox, double cour) {
  time_dt_block(st, nbox, cour);
}

// Ghost pairs (ghost cell, source cell), pool indices.  With with_time the
// last workgroup runs time_dt_block instead (it touches no cell).
__global__ __launch_bounds__(BLK) void kb_ghost(long sv, double *w, double *b,
                                                const int2 *__restrict__ pairs, long npairs,
                                                BoxState st, int nbox, double cour,
                                                int with_time) {
  if (with_time && blockIdx.x == gridDim.x - 1) {
    time_dt_block(st, nbox, cour);
    return;
  }
  const long nb = gridDim.x - with_time;
  for (long t = (long)blockIdx.x * BLK + threadIdx.x; t < npairs; t += nb * BLK) {
    const int2 ds = pairs[t];
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
}

// Ghost cells a reflecting side touches (GhostOp), pool indices.
__global__ __launch_bounds__(BLK) void kb_ghost_ops(long sv, double *w, double *b,
                                                    const GhostOp *__restrict__ ops, long nops,
                                                    BoxState st, int nbox, double cour,
                                                    int with_time) {
  if (with_time && blockIdx.x == gridDim.x - 1) {
    time_dt_block(st, nbox, cour);
    return;
  }
  const long nb = gridDim.x - with_time;
  for (long t = (long)blockIdx.x * BLK + threadIdx.x; t < nops; t += nb * BLK) {
    const GhostOp o = ops[t];
    if (!(o.bits & (1u << 15))) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n) {
        const double v = w[n * sv + o.src];
        w[n * sv + o.dst] = ((o.bits >> n) & 1u) ? -v : v;
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double v = ((o.bits >> (12 + a)) & 1u) ? 0. : b[a * sv + o.srcf[a]];
      b[a * sv + o.dst] = ((o.bits >> (9 + a)) & 1u) ? -v : v;
    }
  }
}

// Faces (and, with fused 1D steps, prims) of the boxes that sat out an odd
// number of steps, from the other buffer into the current one (all cells of
// the box), at the end of an advance; kb_fpar_clear then resets the
// parities.
__global__ __launch_bounds__(BLK) void kb_face_fix(const BoxDev *__restrict__ boxes, WList wl,
                                                   BoxState st, double *__restrict__ b,
                                                   const double *__restrict__ other,
                                                   double *__restrict__ w,
                                                   const double *__restrict__ wother) {
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const int bx = wl.bmap[wb];
    if (!st.fpar[bx]) continue;
    const BoxDev &d = boxes[bx];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    if (t >= d.items[wl.kt].total()) continue;
    const long sv = d.g.sv, c = d.off + t;
#pragma unroll
    for (int n = 0; n < 3; ++n) b[n * sv + c] = other[n * sv + c];
    if (w)  // fused 1D steps: the prim buffers swap as well
      for (int n = 0; n < NMODE; ++n) w[n * sv + c] = wother[n * sv + c];
  }
}

__global__ void kb_fpar_clear(int *fpar, int n) {
  for (int q = threadIdx.x + blockIdx.x * blockDim.x; q < n; q += blockDim.x * gridDim.x)
    fpar[q] = 0;
}

__global__ void kb_zero(unsigned long long *a, int n) {
  for (int q = threadIdx.x + blockIdx.x * blockDim.x; q < n; q += blockDim.x * gridDim.x)
    a[q] = 0ull;
}

}  // namespace

namespace bnzb {
void b_time_dt(bnz_batch *c) {
  hipLaunchKernelGGL(kb_time_dt, dim3(1), dim3(BLK), 0, c->stream, c->S, c->nbox, c->p.cour);
}
}  // namespace bnzb

// ==========================================================================
// Host side
// ==========================================================================


namespace {

int bfail(bnz_batch *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

#define BCHK(ctx, call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bfail(ctx, BNZ_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

// Coordinates of one axis of one box (coord.pyx:201-257, cartesian.pyx:18-43),
// the same construction as the single grid's.
void axis_coords(int N, int ind1, int ind2, int nact, double lmin, double lmax, int bclo,
                 int bchi, std::vector<double> *out) {
  std::vector<double> lf(N + 1);
  if (N == 1) {
    lf[0] = lmin;
    lf[1] = lmax;
  } else {
    const double dx = (lmax - lmin) / nact;
    const double l1 = 0 * nact * dx;
    for (int i = 0; i < N + 1; ++i) lf[i] = l1 + (i - ind1) * dx;
    if (bclo == 0)
      for (int i = ind1 - 1; i > -1; --i) {
        const int i_ = ind2 - i + (ind1 - 1);
        const double d = lf[i_ + 1] - lf[i_];
        lf[i] = lf[i + 1] - d;
      }
    if (bchi == 0)
      for (int i = ind2 + 1; i < N; ++i) {
        const int i_ = ind1 + i - (ind2 + 1);
        const double d = lf[i_ + 1] - lf[i_];
        lf[i + 1] = lf[i] + d;
      }
  }
  for (int q = 0; q < 4; ++q) out[q].resize(N);
  for (int i = 0; i < N; ++i) {
    const double dlf = lf[i + 1] - lf[i];
    const double lv = 0.5 * (lf[i] + lf[i + 1]);
    out[0][i] = dlf;
    out[1][i] = 1. / dlf;
    out[2][i] = lv - lf[i];
    out[3][i] = lf[i + 1] - lv;
  }
}

Box mkbox(int i0, int ni, int j0, int nj, int k0, int nk) {
  Box b;
  b.lo[0] = i0; b.n[0] = ni;
  b.lo[1] = j0; b.n[1] = nj;
  b.lo[2] = k0; b.n[2] = nk;
  return b;
}



// fluxes of the 2D and 3D boxes (1D boxes: kb_stage1)
template <int ORDER, int RS>
void b_fluxes(bnz_batch *c, const double *w, const double *bf, int kt) {
  for (int d = 2; d <= 3; ++d) {
    const WList wl = wlist_dims(c, kt, d);
    if (wl.hi == wl.lo) continue;
    const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TF]));
    if (d == 2)
      hipLaunchKernelGGL((kb_flux<ORDER, RS, 2>), grid, dim3(BLK), 0, c->stream, c->boxes, wl,
                         c->S, w, bf, c->P.fx, c->P.fy, c->P.fz);
    else
      hipLaunchKernelGGL((kb_flux<ORDER, RS, 3>), grid, dim3(BLK), 0, c->stream, c->boxes, wl,
                         c->S, w, bf, c->P.fx, c->P.fy, c->P.fz);
  }
}

// EMFs of the 2D and 3D boxes (1D boxes form theirs in the update)
void b_emf(bnz_batch *c, const double *w, int kt) {
  for (int d = 2; d <= 3; ++d) {
    const WList wl = wlist_dims(c, kt, d);
    if (wl.hi == wl.lo) continue;
    const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TE]));
    if (d == 2)
      hipLaunchKernelGGL(kb_emf<2>, grid, dim3(BLK), 0, c->stream, c->boxes, wl, c->S, w,
                         c->P.fx, c->P.fy, c->P.fz, c->P.ee);
    else
      hipLaunchKernelGGL(kb_emf<3>, grid, dim3(BLK), 0, c->stream, c->boxes, wl, c->S, w,
                         c->P.fx, c->P.fy, c->P.fz, c->P.ee);
  }
}

// stage updates of the 2D and 3D boxes, then the fused stage of the 1D boxes
template <bool PRED, int ORDER, int RS>
void b_stage1(bnz_batch *c) {
  const WList wl = wlist(c, KS1);
  if (wl.hi == wl.lo) return;
  const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TU]));
  hipLaunchKernelGGL((kb_stage1<PRED, ORDER, RS>), grid, dim3(BLK), 0,
                     c->stream, c->boxes, wl, c->S, c->P, c->p.dt_fms_ref);
}

template <bool PRED>
void b_update(bnz_batch *c) {
  const int fr = c->p.dt_fms_ref;
  for (int d = 2; d <= 3; ++d) {
    const WList wl = wlist_dims(c, KUPD, d);
    if (wl.hi == wl.lo) continue;
    const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TU]));
    if (d == 2)
      hipLaunchKernelGGL((kb_update<PRED, 2>), grid, dim3(BLK), 0, c->stream, c->boxes, wl, c->S,
                         c->P, fr);
    else
      hipLaunchKernelGGL((kb_update<PRED, 3>), grid, dim3(BLK), 0, c->stream, c->boxes, wl, c->S,
                         c->P, fr);
  }
  const int order = PRED ? 0 : c->order, rs = c->p.rsolver;
  if (order == 0) {
    if (rs == BNZ_RS_HLLD) b_stage1<PRED, 0, BNZ_RS_HLLD>(c);
    else if (rs == BNZ_RS_HLLT) b_stage1<PRED, 0, BNZ_RS_HLLT>(c);
    else b_stage1<PRED, 0, BNZ_RS_HLL>(c);
  } else {
    if (rs == BNZ_RS_HLLD) b_stage1<PRED, 1, BNZ_RS_HLLD>(c);
    else if (rs == BNZ_RS_HLLT) b_stage1<PRED, 1, BNZ_RS_HLLT>(c);
    else b_stage1<PRED, 1, BNZ_RS_HLL>(c);
  }
}

void b_flux_dispatch(bnz_batch *c, int order, const double *w, const double *bf, int k0) {
  const int rs = c->p.rsolver;
  if (order == 0) {
    if (rs == BNZ_RS_HLLD) b_fluxes<0, BNZ_RS_HLLD>(c, w, bf, k0);
    else if (rs == BNZ_RS_HLLT) b_fluxes<0, BNZ_RS_HLLT>(c, w, bf, k0);
    else b_fluxes<0, BNZ_RS_HLL>(c, w, bf, k0);
  } else {
    if (rs == BNZ_RS_HLLD) b_fluxes<1, BNZ_RS_HLLD>(c, w, bf, k0);
    else if (rs == BNZ_RS_HLLT) b_fluxes<1, BNZ_RS_HLLT>(c, w, bf, k0);
    else b_fluxes<1, BNZ_RS_HLL>(c, w, bf, k0);
  }
}

// Ghost fill of every box; with_time: the end-of-step time and dt kernel
// rides in the last ghost launch (or runs alone if there is none).
void b_ghost(bnz_batch *c, bool with_time) {
  const int tp = with_time && c->ngops == 0 ? 1 : 0, to = with_time ? 1 : 0;
  if (c->nghost > 0)
    hipLaunchKernelGGL(kb_ghost, dim3(grid_for((c->nghost + BLK - 1) / BLK) + tp), dim3(BLK), 0,
                       c->stream, c->ncell, c->P.w, c->P.b, c->ghost, c->nghost, c->S, c->nbox,
                       c->p.cour, tp);
  if (c->ngops > 0)
    hipLaunchKernelGGL(kb_ghost_ops, dim3(grid_for((c->ngops + BLK - 1) / BLK) + to), dim3(BLK),
                       0, c->stream, c->ncell, c->P.w, c->P.b, c->gops, c->ngops, c->S, c->nbox,
                       c->p.cour, to);
  else if (with_time && c->nghost == 0)
    b_time_dt(c);
}

// One VL step of every running box (vanleer.pyx:94-167): the same launch
// sequence as the single grid's step.
// One VL step of every running box; the box dt's were set by the previous
// step's kb_time_dt (or by init_state's kb_dt).
template <int ORDER, int RS>
void b_step1(bnz_batch *c) {
  const WList wl = wlist(c, KT1);
  hipLaunchKernelGGL((kb_step1<ORDER, RS>), dim3(grid_for(wl.hi - wl.lo, c->cap[TU])), dim3(BLK),
                     0, c->stream, c->boxes, wl, c->S, c->P, c->p.dt_fms_ref);
}

void b_step(bnz_batch *c) {
  if (c->f32) {
    b32_step(c);  // ends with the ghost fill and the time kernel
    return;
  }
  if (c->step1) {  // every box 1D: one fused pass per step
    const int rs = c->p.rsolver;
    if (c->order == 0) {
      if (rs == BNZ_RS_HLLD) b_step1<0, BNZ_RS_HLLD>(c);
      else if (rs == BNZ_RS_HLLT) b_step1<0, BNZ_RS_HLLT>(c);
      else b_step1<0, BNZ_RS_HLL>(c);
    } else {
      if (rs == BNZ_RS_HLLD) b_step1<1, BNZ_RS_HLLD>(c);
      else if (rs == BNZ_RS_HLLT) b_step1<1, BNZ_RS_HLLT>(c);
      else b_step1<1, BNZ_RS_HLL>(c);
    }
    std::swap(c->P.b, c->P.bs);
    std::swap(c->P.w, c->P.ws);
    b_ghost(c, true);
    return;
  }
  // predictor: donor cell, dt/2, limits of order 0
  b_flux_dispatch(c, 0, c->P.w, c->P.b, KF0);
  b_emf(c, c->P.w, KEM0);
  b_update<true>(c);
  // corrector: reconstruction order, dt; b_{n+1} goes to the predictor's
  // face buffer, which then becomes the current one
  b_flux_dispatch(c, c->order, c->P.ws, c->P.bs, KF1);
  b_emf(c, c->P.ws, KEM1);
  b_update<false>(c);
  std::swap(c->P.b, c->P.bs);
  b_ghost(c, true);
}

int blaunch_check(bnz_batch *c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return bfail(c, BNZ_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return BNZ_OK;
}

void drop_graph(bnz_batch *c) {
  if (c->exec) (void)hipGraphExecDestroy(c->exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->exec = nullptr;
  c->graph = nullptr;
  c->graph_b = nullptr;
}

}  // namespace

extern "C" {

const char *bnz_batch_last_error(const bnz_batch *c) { return c ? c->err.c_str() : "null batch"; }

int bnz_batch_create(const bnz_params *p, const bnz_box *boxes, int nbox, bnz_batch **out) {
  if (!out) return BNZ_EINVAL;
  *out = nullptr;
  if (!p || !boxes || nbox < 1) return BNZ_EINVAL;
  bnz_batch *c = new bnz_batch();
  *out = c;
  c->p = *p;
  c->nbox = nbox;
  if (p->precision == BNZ_PREC_FP64_FMA)
    return bfail(c, BNZ_EUNSUPPORTED, "batches run BNZ_PREC_FP64 or BNZ_PREC_FP32C");
  if (p->precision != BNZ_PREC_FP64 && p->precision != BNZ_PREC_FP32C)
    return bfail(c, BNZ_EINVAL, "precision must be BNZ_PREC_FP64 or BNZ_PREC_FP32C");
  c->f32 = p->precision == BNZ_PREC_FP32C;
  if (p->rsolver != BNZ_RS_HLLD && p->rsolver != BNZ_RS_HLL && p->rsolver != BNZ_RS_HLLT)
    return bfail(c, BNZ_EUNSUPPORTED, "rsolver must be hll, hlld or hllt");
  if (p->rsolver == BNZ_RS_HLLT && c->f32)
    return bfail(c, BNZ_EUNSUPPORTED, "hllt runs in FP64 only");
  if (p->reconstr != BNZ_RCN_CONST && p->reconstr != BNZ_RCN_LINEAR)
    return bfail(c, BNZ_EUNSUPPORTED, "reconstr must be const or linear");
  if (p->hydro)
    return bfail(c, BNZ_EUNSUPPORTED, "the hydro build runs on single grids only");
  if (p->ct_upwind)
    return bfail(c, BNZ_EUNSUPPORTED, "the upwind corner EMF runs on single grids only");
  if (p->tintegr)
    return bfail(c, BNZ_EUNSUPPORTED, "the RK3 integrator runs on single grids only");
  if (p->char_proj)
    return bfail(c, BNZ_EUNSUPPORTED, "characteristic limiting runs on single grids only");
  if (!(p->gam > 1.) || !(p->cour > 0.)) return bfail(c, BNZ_EINVAL, "need gam > 1 and cour > 0");
  c->order = (p->reconstr == BNZ_RCN_CONST) ? 0 : 1;
  c->ng = c->order ? 3 : 2;  // grid.pyx:132-150
  const int ng = c->ng;
  std::vector<double> hgeo;
  std::vector<size_t> geo_off;  // 12 per box: (dlf, dlfi, icm, icp) x 3 axes
  std::vector<int2> pairs;
  std::vector<GhostOp> gops;
  c->hbox.resize(nbox);
  long off = 0;
  for (int b = 0; b < nbox; ++b) {
    const bnz_box &s = boxes[b];
    if (s.nx < 1 || s.ny < 1 || s.nz < 1 || (s.nz > 1 && s.ny == 1))
      return bfail(c, BNZ_EINVAL, "box " + std::to_string(b) + ": bad size");
    {
      // before any per-box host array is sized from it
      const long mx = 1L << 30, ext = 2 * ng;
      const long cells = (long)std::min<long>(s.nx, mx) + ext;
      const long cy = s.ny > 1 ? (long)std::min<long>(s.ny, mx) + ext : 1;
      const long cz = s.nz > 1 ? (long)std::min<long>(s.nz, mx) + ext : 1;
      if (s.nx > mx || s.ny > mx || s.nz > mx || off + cells * cy * cz >= (1L << 31))
        return bfail(c, BNZ_EINVAL, "batch too large: pool cell indices are 32-bit");
    }
    for (int q = 0; q < 6; ++q)
      if (s.bc[q] != BNZ_BC_PERIODIC && s.bc[q] != BNZ_BC_OUTFLOW && s.bc[q] != BNZ_BC_REFLECT)
        return bfail(c, BNZ_EUNSUPPORTED,
                     "box " + std::to_string(b) +
                         ": boundary flags must be periodic, outflow or reflect");
    for (int a = 0; a < 3; ++a)
      if (!(s.lmax[a] > s.lmin[a]))
        return bfail(c, BNZ_EINVAL, "box " + std::to_string(b) + ": need lmax > lmin");
    BoxDev &d = c->hbox[b];
    std::memset(&d, 0, sizeof(d));
    const int dims = (s.nz > 1) ? 3 : (s.ny > 1) ? 2 : 1;
    const int nt[3] = {s.nx + 2 * ng, dims >= 2 ? s.ny + 2 * ng : 1, dims >= 3 ? s.nz + 2 * ng : 1};
    const int a1[3] = {ng, dims >= 2 ? ng : 0, dims >= 3 ? ng : 0};
    const int a2[3] = {s.nx + ng - 1, dims >= 2 ? s.ny + ng - 1 : 0, dims >= 3 ? s.nz + ng - 1 : 0};
    const int nact[3] = {s.nx, s.ny, s.nz};
    for (int a = 0; a < dims; ++a)
      if ((s.bc[2 * a] != BNZ_BC_OUTFLOW || s.bc[2 * a + 1] != BNZ_BC_OUTFLOW) && nact[a] < ng)
        return bfail(c, BNZ_EINVAL, "box " + std::to_string(b) +
                                        ": a periodic or reflecting axis needs at least ng cells");
    for (int q = 0; q < 3; ++q) d.g.nt[q] = nt[q];
    d.g.dims = dims;
    d.g.sy = nt[0];
    d.g.sz = nt[0] * nt[1];
    d.g.gam = p->gam;
    d.g.gamm1i = 1. / (p->gam - 1.);
    d.g.unif = 0;  // batch kernels read the per-cell geometry arrays
    d.off = off;
    for (int a = 0; a < 3; ++a) {
      std::vector<double> v[4];
      axis_coords(nt[a], a1[a], a2[a], nact[a], s.lmin[a], s.lmax[a], s.bc[2 * a],
                  s.bc[2 * a + 1], v);
      for (int q = 0; q < 4; ++q) {
        geo_off.push_back(hgeo.size());
        hgeo.insert(hgeo.end(), v[q].begin(), v[q].end());
      }
    }
    // work items of each list: the single grid's launch boxes
    for (int o = 0; o < 2; ++o) {
      const int ord = o == 0 ? 0 : c->order;
      Lims &L = d.L[o];
      for (int a = 0; a < 3; ++a) {
        L.lo[a] = ord + 1;
        L.hi[a] = nt[a] - 2 - ord;
      }
      if (dims < 2) L.lo[1] = L.hi[1] = 0;
      if (dims < 3) L.lo[2] = L.hi[2] = 0;
      const int il = L.lo[0], iu = L.hi[0], jl = L.lo[1], ju = L.hi[1], kl = L.lo[2], ku = L.hi[2];
      d.items[o == 0 ? KF0 : KF1] = flux3_box(dims, L);
      // 1D boxes form their EMFs from the LDS fluxes (EmfLds1D, EmfLds32x1D): no EMF items
      d.items[o == 0 ? KEM0 : KEM1] =
          dims == 1
              ? mkbox(0, 0, 0, 0, 0, 0)
              : mkbox(il, iu + 2 - il, dims >= 2 ? jl : 0, dims >= 2 ? ju + 2 - jl : 1,
                      dims >= 3 ? kl : 0, dims >= 3 ? ku + 2 - kl : 1);
    }
    d.items[KUPD] = mkbox(0, nt[0], 0, nt[1], 0, nt[2]);
    d.items[KS1] = dims == 1 ? mkbox(0, nt[0], 0, 1, 0, 1) : mkbox(0, 0, 0, 0, 0, 0);
    d.items[KT1] = d.items[KS1];
    d.act = mkbox(a1[0], s.nx, a1[1], dims >= 2 ? s.ny : 1, a1[2], dims >= 3 ? s.nz : 1);
    d.items[KACT] = d.act;
    for (int a = 0; a < 3; ++a) {  // reflecting wall planes (FP32C EMFs)
      d.walls[2 * a] = (a < dims && s.bc[2 * a] == BNZ_BC_REFLECT) ? a1[a] : -1;
      d.walls[2 * a + 1] = (a < dims && s.bc[2 * a + 1] == BNZ_BC_REFLECT) ? a2[a] + 1 : -1;
    }
    build_ghosts(nt, dims, ng, s.bc, off, pairs, gops);
    off += (long)nt[0] * nt[1] * nt[2];
    if (off >= (1L << 31))
      return bfail(c, BNZ_EINVAL, "batch too large: pool cell indices are 32-bit");
  }
  c->ncell = off;
  c->step1 = true;
  for (int b = 0; b < nbox; ++b) c->step1 = c->step1 && c->hbox[b].g.dims == 1;
  c->nghost = (long)pairs.size();
  c->ngops = (long)gops.size();
  // work blocks: box b owns ceil(items / BLK) consecutive blocks of each
  // list; the 1D boxes come first, then the 2D and the 3D boxes
  std::vector<int> hmap;
  for (int kt = 0; kt < NK; ++kt) {
    c->bm_off[kt] = (int)hmap.size();
    long nb = 0;
    for (int dims = 1; dims <= 3; ++dims) {
      c->cls[kt][dims - 1] = (int)nb;
      for (int b = 0; b < nbox; ++b) {
        BoxDev &d = c->hbox[b];
        if (d.g.dims != dims) continue;
        d.blk0[kt] = (int)nb;
        const long blocks = (d.items[kt].total() + blk_items(kt) - 1) / blk_items(kt);
        for (long q = 0; q < blocks; ++q) hmap.push_back(b);
        nb += blocks;
      }
    }
    c->cls[kt][3] = (int)nb;
    c->nblk[kt] = (int)nb;
  }

  if (hipSetDevice(p->device) != hipSuccess) return bfail(c, BNZ_EHIP, "hipSetDevice failed");
  BCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  BCHK(c, hipHostMalloc(&c->hrun, 2 * sizeof(int), hipHostMallocDefault));
  for (int q = 0; q < 2; ++q) BCHK(c, hipEventCreateWithFlags(&c->ev[q], hipEventDisableTiming));
  const size_t cells = (size_t)c->ncell;
  struct {
    double **q;
    int rows;
  } bufs[] = {{&c->P.w, NMODE}, {&c->P.b, 3},          {&c->P.bs, 3},
              {&c->P.ws, NMODE}, {&c->P.fx, NMODE}, {&c->P.fy, NMODE},
              {&c->P.fz, NMODE}, {&c->P.ee, 3}};
  // FP32C keeps prim (upload, download) and the faces in FP64; its work
  // arrays are the FP32 pools below
  const int nb64 = c->f32 ? 3 : 8;
  for (int q = 0; q < nb64; ++q) {
    if (hipMalloc(bufs[q].q, bufs[q].rows * cells * sizeof(double)) != hipSuccess)
      return bfail(c, BNZ_ENOMEM, "device allocation of the box pools failed");
    // zeroed once so reads of never-written ghost faces are defined
    BCHK(c, hipMemsetAsync(*bufs[q].q, 0, bufs[q].rows * cells * sizeof(double), c->stream));
  }
  if (c->f32) {
    struct {
      float **q;
      int rows;
    } b32[] = {{&c->Q.uh, 6},   {&c->Q.ul, 6},  {&c->Q.w32, 9}, {&c->Q.ws32, 9},
               {&c->Q.bs32, 3}, {&c->Q.fx, 9},  {&c->Q.fy, 9},  {&c->Q.fz, 9},
               {&c->Q.ee, 3}};
    for (auto &q : b32) {
      if (hipMalloc(q.q, q.rows * cells * sizeof(float)) != hipSuccess)
        return bfail(c, BNZ_ENOMEM, "device allocation of the FP32 box pools failed");
      BCHK(c, hipMemsetAsync(*q.q, 0, q.rows * cells * sizeof(float), c->stream));
    }
    if (c->step1)  // fused 1D steps write the new hi and lo words to a second buffer
      for (float **q : {&c->Q.uh2, &c->Q.ul2}) {
        if (hipMalloc(q, 6 * cells * sizeof(float)) != hipSuccess)
          return bfail(c, BNZ_ENOMEM, "device allocation of the FP32 box pools failed");
        BCHK(c, hipMemsetAsync(*q, 0, 6 * cells * sizeof(float), c->stream));
      }
    BCHK(c, hipMalloc(&c->nguard, sizeof(int)));
    BCHK(c, hipMemsetAsync(c->nguard, 0, sizeof(int), c->stream));
  }
  BCHK(c, hipMalloc(&c->geo, hgeo.size() * sizeof(double)));
  BCHK(c, hipMemcpy(c->geo, hgeo.data(), hgeo.size() * sizeof(double), hipMemcpyHostToDevice));
  for (int b = 0; b < nbox; ++b) {
    Geo &g = c->hbox[b].g;
    g.sv = c->ncell;
    for (int a = 0; a < 3; ++a) {
      g.dlf[a] = c->geo + geo_off[b * 12 + a * 4 + 0];
      g.dlfi[a] = c->geo + geo_off[b * 12 + a * 4 + 1];
      g.icm[a] = c->geo + geo_off[b * 12 + a * 4 + 2];
      g.icp[a] = c->geo + geo_off[b * 12 + a * 4 + 3];
      g.dlv[a] = nullptr;  // only the upwind corner EMF reads it (single grids)
    }
  }
  BCHK(c, hipMalloc(&c->boxes, nbox * sizeof(BoxDev)));
  BCHK(c, hipMemcpy(c->boxes, c->hbox.data(), nbox * sizeof(BoxDev), hipMemcpyHostToDevice));
  BCHK(c, hipMalloc(&c->bmap, std::max<size_t>(1, hmap.size()) * sizeof(int)));
  BCHK(c, hipMemcpy(c->bmap, hmap.data(), hmap.size() * sizeof(int), hipMemcpyHostToDevice));
  BCHK(c, hipMalloc(&c->ghost, std::max<size_t>(1, pairs.size()) * sizeof(int2)));
  if (!pairs.empty())
    BCHK(c, hipMemcpy(c->ghost, pairs.data(), pairs.size() * sizeof(int2), hipMemcpyHostToDevice));
  BCHK(c, hipMalloc(&c->gops, std::max<size_t>(1, gops.size()) * sizeof(GhostOp)));
  if (!gops.empty())
    BCHK(c, hipMemcpy(c->gops, gops.data(), gops.size() * sizeof(GhostOp), hipMemcpyHostToDevice));
  // box scalars: dt, t, tmax, cmax (8 bytes each), run, steps, fpar (4 bytes),
  // nrun
  const size_t nb = nbox;
  const size_t words = 4 * nb + (3 * nb + 1 + 1) / 2;
  BCHK(c, hipMalloc(&c->sbuf, words * sizeof(double)));
  c->S.dt = c->sbuf;
  c->S.t = c->sbuf + nb;
  c->S.tmax = c->sbuf + 2 * nb;
  c->S.cmax = (unsigned long long *)(c->sbuf + 3 * nb);
  c->S.run = (int *)(c->sbuf + 4 * nb);
  c->S.steps = c->S.run + nb;
  c->S.fpar = c->S.steps + nb;
  c->S.nrun = c->S.fpar + nb;
  std::vector<double> h(words, 0.);
  for (int b = 0; b < nbox; ++b) h[2 * nb + b] = boxes[b].tmax;
  BCHK(c, hipMemcpy(c->sbuf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  BCHK(c, hipStreamSynchronize(c->stream));
  return BNZ_OK;
}

int bnz_batch_destroy(bnz_batch *c) {
  if (!c) return BNZ_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  double *bufs[] = {c->P.w, c->P.ws, c->P.b,  c->P.bs, c->P.fx,
                    c->P.fy, c->P.fz, c->P.ee, c->geo,  c->sbuf};
  for (double *q : bufs)
    if (q) (void)hipFree(q);
  float *b32[] = {c->Q.uh, c->Q.ul, c->Q.w32, c->Q.ws32, c->Q.bs32, c->Q.fx,
                  c->Q.fy, c->Q.fz, c->Q.ee,  c->Q.uh2,  c->Q.ul2};
  for (float *q : b32)
    if (q) (void)hipFree(q);
  if (c->nguard) (void)hipFree(c->nguard);
  if (c->boxes) (void)hipFree(c->boxes);
  if (c->bmap) (void)hipFree(c->bmap);
  if (c->ghost) (void)hipFree(c->ghost);
  if (c->gops) (void)hipFree(c->gops);
  for (hipEvent_t e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->hrun) (void)hipHostFree(c->hrun);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return BNZ_OK;
}

int bnz_batch_box_shape(const bnz_batch *c, int box, int *ntot3) {
  if (!c || box < 0 || box >= c->nbox || !ntot3 || c->hbox.size() != (size_t)c->nbox)
    return BNZ_EINVAL;
  for (int q = 0; q < 3; ++q) ntot3[q] = c->hbox[box].g.nt[q];
  return BNZ_OK;
}

int bnz_batch_upload(bnz_batch *c, int box, const double *prim, const double *bfld) {
  if (!c || box < 0 || box >= c->nbox) return BNZ_EINVAL;
  if (!c->stream) return bfail(c, BNZ_ESTATE, "batch was not created");
  const BoxDev &d = c->hbox[box];
  const size_t n = (size_t)d.g.nt[0] * d.g.nt[1] * d.g.nt[2];
  for (int v = 0; v < NMODE && prim; ++v)
    BCHK(c, hipMemcpyAsync(c->P.w + v * c->ncell + d.off, prim + v * n, n * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  for (int v = 0; v < 3 && bfld; ++v)
    BCHK(c, hipMemcpyAsync(c->P.b + v * c->ncell + d.off, bfld + v * n, n * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  BCHK(c, hipStreamSynchronize(c->stream));
  c->ready = false;
  c->w_stale = false;
  return BNZ_OK;
}

int bnz_batch_download(bnz_batch *c, int box, double *prim, double *bfld) {
  if (!c || box < 0 || box >= c->nbox) return BNZ_EINVAL;
  if (!c->stream) return bfail(c, BNZ_ESTATE, "batch was not created");
  if (c->f32 && c->w_stale) {  // FP32C state -> FP64 prim of every box, then its ghosts
    b32_to64(c);
    b_ghost(c, false);
    int rc = blaunch_check(c);
    if (rc) return rc;
    c->w_stale = false;
  }
  BCHK(c, hipStreamSynchronize(c->stream));
  const BoxDev &d = c->hbox[box];
  const size_t n = (size_t)d.g.nt[0] * d.g.nt[1] * d.g.nt[2];
  for (int v = 0; v < NMODE && prim; ++v)
    BCHK(c, hipMemcpy(prim + v * n, c->P.w + v * c->ncell + d.off, n * sizeof(double),
                      hipMemcpyDeviceToHost));
  for (int v = 0; v < 3 && bfld; ++v)
    BCHK(c, hipMemcpy(bfld + v * n, c->P.b + v * c->ncell + d.off, n * sizeof(double),
                      hipMemcpyDeviceToHost));
  return BNZ_OK;
}

int bnz_batch_init_state(bnz_batch *c) {
  if (!c) return BNZ_EINVAL;
  if (!c->stream) return bfail(c, BNZ_ESTATE, "batch was not created");
  const int nb = c->nbox;
  b_ghost(c, false);
  hipLaunchKernelGGL(kb_interp, dim3(grid_for(c->nblk[KACT])), dim3(BLK), 0, c->stream, c->boxes,
                     wlist(c, KACT), c->P.w, c->P.b);
  b_ghost(c, false);
  std::vector<double> z(nb, 0.);
  std::vector<int> zi(2 * nb + 1, 0);  // steps, fpar, nrun
  BCHK(c, hipMemcpyAsync(c->S.t, z.data(), nb * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BCHK(c, hipMemcpyAsync(c->S.steps, zi.data(), (2 * nb + 1) * sizeof(int), hipMemcpyHostToDevice,
                         c->stream));
  hipLaunchKernelGGL(kb_zero, dim3(1), dim3(BLK), 0, c->stream, c->S.cmax, nb);
  if (c->f32)  // FP32C state from the FP64 prim, first CFL maxima from its FP32 prim
    b32_init(c);
  else
    hipLaunchKernelGGL(kb_cfl, dim3(grid_for(c->nblk[KACT])), dim3(BLK), 0, c->stream,
                       c->boxes, wlist(c, KACT), c->S, c->P.w, c->p.dt_fms_ref);
  // dt and run flags of the first step (later steps get theirs from kb_time_dt)
  hipLaunchKernelGGL(kb_dt, dim3(1), dim3(BLK), 0, c->stream, c->S, c->nbox, c->p.cour);
  c->w_stale = false;
  int rc = blaunch_check(c);
  if (rc) return rc;
  BCHK(c, hipStreamSynchronize(c->stream));
  c->ready = true;
  return BNZ_OK;
}

int bnz_batch_advance(bnz_batch *c, int max_steps, int use_graph, int *nsteps) {
  if (!c) return BNZ_EINVAL;
  if (!c->ready) return bfail(c, BNZ_ESTATE, "bnz_batch_init_state has not run");
  if (use_graph && c->exec && c->graph_b != c->P.b) drop_graph(c);
  if (use_graph && !c->exec) {
    c->graph_b = c->P.b;
    BCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    b_step(c);
    b_step(c);  // two swaps: the face buffers are back where capture began
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) {
      c->graph_b = nullptr;
      return bfail(c, BNZ_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    c->graph = g;
    BCHK(c, hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0));
  }
  // Steps go out in chunks (REPLAYS graph replays, or one eager step), each
  // followed by a copy of the running-box count into a pinned slot.  The
  // host reads a chunk's count while the next chunk runs, so the device does
  // not drain between chunks; a chunk issued after the last box finished only
  // launches idle kernels.
  constexpr int REPLAYS = 4;
  int n = 0;
  auto chunk = [&](int slot) -> int {
    if (use_graph) {
      for (int r = 0; r < REPLAYS && (max_steps <= 0 || n < max_steps); ++r) {
        BCHK(c, hipGraphLaunch(c->exec, c->stream));
        n += 2;
      }
    } else {
      b_step(c);
      n += 1;
    }
    int rc = blaunch_check(c);
    if (rc) return rc;
    BCHK(c, hipMemcpyAsync(c->hrun + slot, c->S.nrun, sizeof(int), hipMemcpyDeviceToHost,
                           c->stream));
    BCHK(c, hipEventRecord(c->ev[slot], c->stream));
    return BNZ_OK;
  };
  int slot = 0;
  int rc = chunk(slot);
  if (rc) return rc;
  for (;;) {
    const bool more = max_steps <= 0 || n < max_steps;
    if (more && (rc = chunk(slot ^ 1)) != 0) return rc;
    BCHK(c, hipEventSynchronize(c->ev[slot]));
    if (c->hrun[slot] == 0 || !more) break;
    slot ^= 1;
  }
  // faces of the boxes that finished an odd number of steps ago
  hipLaunchKernelGGL(kb_face_fix, dim3(grid_for(c->nblk[KUPD])), dim3(BLK), 0, c->stream,
                     c->boxes, wlist(c, KUPD), c->S, c->P.b, c->P.bs,
                     c->step1 && !c->f32 ? c->P.w : nullptr, c->P.ws);
  if (c->step1 && c->f32) b32_fix(c);
  hipLaunchKernelGGL(kb_fpar_clear, dim3(1), dim3(BLK), 0, c->stream, c->S.fpar, c->nbox);
  rc = blaunch_check(c);
  if (rc) return rc;
  BCHK(c, hipStreamSynchronize(c->stream));
  if (nsteps) *nsteps = n;
  c->w_stale = c->f32;
  return BNZ_OK;
}

int bnz_batch_set_tuning(bnz_batch *c, int flux_blocks, int emf_blocks, int update_blocks) {
  if (!c) return BNZ_EINVAL;
  const int v[NTUNE] = {flux_blocks, emf_blocks, update_blocks};
  for (int q = 0; q < NTUNE; ++q)
    if (v[q] < 0 || v[q] > (1 << 24)) return bfail(c, BNZ_EINVAL, "workgroup cap out of range");
  if (c->stream) BCHK(c, hipStreamSynchronize(c->stream));
  drop_graph(c);  // the captured launches carry the old grid sizes
  for (int q = 0; q < NTUNE; ++q) c->cap[q] = v[q] ? v[q] : PERSIST;
  return BNZ_OK;
}

int bnz_batch_get_times(bnz_batch *c, double *t, int *steps) {
  if (!c) return BNZ_EINVAL;
  if (!c->stream) return bfail(c, BNZ_ESTATE, "batch was not created");
  BCHK(c, hipStreamSynchronize(c->stream));
  if (t) BCHK(c, hipMemcpy(t, c->S.t, c->nbox * sizeof(double), hipMemcpyDeviceToHost));
  if (steps) BCHK(c, hipMemcpy(steps, c->S.steps, c->nbox * sizeof(int), hipMemcpyDeviceToHost));
  return BNZ_OK;
}

int bnz_batch_synchronize(bnz_batch *c) {
  if (!c) return BNZ_EINVAL;
  if (!c->stream) return bfail(c, BNZ_ESTATE, "batch was not created");
  BCHK(c, hipStreamSynchronize(c->stream));
  return BNZ_OK;
}

}  // extern "C"
