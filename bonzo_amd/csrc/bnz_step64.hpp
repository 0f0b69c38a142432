// bnz_step64.hpp -- the FP64 step's kernels and their launches, included by
// two translation units that compile the same code with different
// arithmetic: bnz_gpu.hip (no contraction: bit for bit the CPU oracle) and
// bnz_gpu64f.hip (a*b+c contracted into FMAs: the fast FP64 mode).  Every
// definition is in an anonymous namespace, so each unit gets its own copy of
// the kernels, compiled with its own flags.
//
// One VL step (src/integration/vanleer.pyx:87-167), after the dt kernels:
//   predictor: flux3, emf, update<pred>        (donor cell, dt/2)
//   corrector: flux3, emf, update<corr>        (PLM, dt; CFL maximum fused)
//   ghost (+ ghost_ops for reflecting sides), time_advance
// 1D grids run the whole step as one fused pass (k_step1).
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "bnz_kern64.hpp"
#include "bnz_rk3.hpp"

namespace {

using namespace bnz;

// --------------------------------------------------------------------------
// Work shapes (3D): ty == 0 one thread per item in x-fastest order; TILED
// 64 x 4 tiles ordered along z; ty > 1 slab columns of ty rows walked plane
// by plane (see cell_index).
//
// All three face directions of a stage in one launch: one thread per cell of
// the union of the face ranges computes the x, y and z faces whose lower
// cell it is, so the prim stencil of a cell is fetched once per stage.
// A step's first flux launch also reduces its dt in workgroup 0 (DtFold).
// SPLIT > 0 (small grids, flat order): workgroup b computes the faces of
// direction b % SPLIT only, so a thread solves one face instead of SPLIT and
// the launch, bound there by one thread's chain of solves, ends sooner.
template <int ORDER, int RS, bool TILED = false, bool CP = false, int SPLIT = 0>
__global__ __launch_bounds__(BLK) void k_flux3(Geo g, const double *__restrict__ w,
                                               const double *__restrict__ bf,
                                               double *__restrict__ fx, double *__restrict__ fy,
                                               double *__restrict__ fz, Box box, Lims L,
                                               int ty, double *dev, DtFold dtf) {
  if (dtf.npart > 0 && blockIdx.x == 0) dt_reduce_block(dtf, dev);
  if (step_past_tmax(dtf, dev)) return;  // a step past t_max (see k_update)
  int i, j, k;
  if (SPLIT > 0) {
    const long lb = xcd_block();
    const int dir = (int)(lb % SPLIT);
    const long t = (lb / SPLIT) * BLK + threadIdx.x;
    if (t >= box.total()) return;
    box_idx(box, t, i, j, k);
    // the faces flux_cell computes for this cell, one direction
    if (dir == 0) {
      if (i >= L.lo[0]) flux_face<0, ORDER, RS, CP>(g, w, bf, fx, i, j, k);
    } else if (dir == 1) {
      if (j >= L.lo[1]) flux_face<1, ORDER, RS, CP>(g, w, bf, fy, i, j, k);
    } else {
      if (k >= L.lo[2]) flux_face<2, ORDER, RS, CP>(g, w, bf, fz, i, j, k);
    }
    return;
  }
  if (!cell_index<TILED, 1, 2>(box, ty, i, j, k)) return;
  flux_cell<ORDER, RS, 0, CP>(g, w, bf, fx, fy, fz, L, i, j, k);
}

// A whole VL step of a 1D grid in one pass (step1_block): workgroup b owns
// the cells from b * STEP1_CELLS; the new prim and faces go to wout and
// bout, which the host then swaps with w and b; the workgroup's CFL maximum
// goes to part.
template <int ORDER, int RS, bool CP = false>
__global__ __launch_bounds__(BLK) void k_step1(Geo g, const double *__restrict__ w,
                                               const double *__restrict__ bf,
                                               double *__restrict__ wout,
                                               double *__restrict__ bout,
                                               const double *__restrict__ fx_old, Lims L0,
                                               Lims L1, const double *__restrict__ dts, Box act,
                                               int fms_ref, double *__restrict__ part) {
  __shared__ double sf[NMODE * BLK];
  __shared__ double sw[NMODE * BLK];
  const int i0 = blockIdx.x * STEP1_CELLS;
  if (dts[0] == 0.) {  // a step past t_max: the state is carried into the new buffers
    const int i = i0 + threadIdx.x;
    if (threadIdx.x < STEP1_CELLS && i < g.nt[0]) {
      for (int n = 0; n < NMODE; ++n) wout[n * g.sv + i] = w[n * g.sv + i];
      for (int a = 0; a < 3; ++a) bout[a * g.sv + i] = bf[a * g.sv + i];
    }
    return;
  }
  const double cds = step1_block<ORDER, RS, CP>(g, w, bf, wout, bout, fx_old, L0, L1, dts, act,
                                                fms_ref, i0, sf, sw);
  block_max_to(cds, part);
}

// --------------------------------------------------------------------------
// Edge EMFs, interp_ee1 (src/integration/ct.pyx:153-215) with the cell-centred
// E = -v x B of ec_from_prim (ct.pyx:128-148) computed on the fly.
template <int DIMS, bool TILED = false>
__global__ __launch_bounds__(BLK) void k_emf(Geo g, const double *__restrict__ w,
                                             const double *__restrict__ fx,
                                             const double *__restrict__ fy,
                                             const double *__restrict__ fz,
                                             double *__restrict__ ee, Box box, int ty) {
  int i, j, k;
  if (!cell_index<TILED, 1, 2>(box, ty, i, j, k)) return;
  emf_edge_d<DIMS>(g, w, fx, fy, fz, ee, i, j, k);
}

// The upwind corner EMFs (interp_ee2, emf_edge_upwind), weights from the
// step's dt.
template <int DIMS, bool TILED = false>
__global__ __launch_bounds__(BLK) void k_emf_upw(Geo g, const double *__restrict__ w,
                                                 const double *__restrict__ fx,
                                                 const double *__restrict__ fy,
                                                 const double *__restrict__ fz,
                                                 double *__restrict__ ee, Box box, int ty,
                                                 const double *__restrict__ dts) {
  int i, j, k;
  if (!cell_index<TILED, 1, 2>(box, ty, i, j, k)) return;
  emf_edge_upwind<DIMS>(g, w, fx, fy, fz, ee, dts[0], i, j, k);
}

// One RK3 stage update (update_cell_rk) over the whole grid; stage 3 leaves
// the next step's CFL partial maxima in part.
template <int DIMS, int STAGE>
__global__ __launch_bounds__(BLK) void k_update_rk(
    Geo g, const double *__restrict__ u0, const double *__restrict__ fx,
    const double *__restrict__ fy, const double *__restrict__ fz, const double *__restrict__ ee,
    const double *__restrict__ b0, const double *ua, const double *ba, double *u1, double *b1,
    double *w1, Lims L, const double *__restrict__ dts, Box grid, Box act, int fms_ref,
    double *__restrict__ part) {
  int i, j, k;
  double cdsmax = 0.;
  const bool have = cell_index<false, 1, 2>(grid, 0, i, j, k);
  if (dts[0] == 0.) return;  // a step past t_max changes nothing (uniform)
  if (have) {
    const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
    cdsmax = update_cell_rk<DIMS, STAGE, STAGE == 3>(
        g, u0, FluxGlobal{fx, fy, fz, g.sv, c}, EmfGlobal{ee, g.sv, c}, b0, ua, ba, u1, b1, w1,
        L, dts[0], act, fms_ref, i, j, k);
  }
  if (STAGE == 3) block_max_to(cdsmax, part);
}

// U = prim2cons(W) over every cell (the RK3 step's U, _rk3.pyx:167).
__global__ __launch_bounds__(BLK) void k_cons_of_prim(Geo g, const double *__restrict__ w,
                                                      double *__restrict__ u,
                                                      const double *__restrict__ dts) {
  const long t = xcd_block() * BLK + threadIdx.x;
  if (dts[0] == 0. || t >= g.sv) return;
  double wv[NMODE], uv[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) wv[n] = w[n * g.sv + t];
  prim2cons(uv, wv, g.gamm1i);
#pragma unroll
  for (int n = 0; n < NMODE; ++n) u[n * g.sv + t] = uv[n];
}

template <int DIMS, bool PRED, bool DTRED = false, bool TILED = false, bool UG = false,
          bool FLY = false>
__global__ __launch_bounds__(BLK) void k_update(Geo g, const double *w0, const double *__restrict__ fx,
                                                const double *__restrict__ fy,
                                                const double *__restrict__ fz,
                                                const double *__restrict__ ee,
                                                const double *__restrict__ b0,
                                                double *__restrict__ b1, double *w1, Lims L,
                                                const double *__restrict__ dts, Box grid, Box act,
                                                int fms_ref, double *__restrict__ part,
                                                int ty, const double *we = nullptr) {
  int i, j, k;
  double cdsmax = 0.;
  const bool have = cell_index<TILED, 1, 2>(grid, ty, i, j, k);
  if (dts[0] == 0.) {  // a step past t_max: state unchanged, faces carried over
    if (!PRED && have) {
      const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
      for (int a = 0; a < 3; ++a) b1[a * g.sv + c] = b0[a * g.sv + c];
    }
    return;  // uniform over the workgroup; the CFL partials stay those of the state
  }
  if (have)
    cdsmax = update_cell_d<DIMS, PRED, DTRED, UG, FLY>(g, w0, fx, fy, fz, ee, b0, b1, w1, L, dts,
                                                       act, fms_ref, i, j, k, we);
  if (DTRED) block_max_to(cdsmax, part);
}

// ==========================================================================
// Launches
// ==========================================================================

template <int ORDER, int RS, bool CP>
void launch_fluxes_cp(bnz_ctx *c, const double *w, const double *bf, const Lims &L) {
  // all face directions in one launch over the union of the face ranges
  const Box u = flux3_box(c->dims, L);
  const Shape sh = flat_order(c, work_shape(u, c->dims, c->flux_chunk));
  const DtFold dtf = take_dt(c);
  Launch t(c, ORDER ? (CP ? "flux3_plm_char" : "flux3_plm") : "flux3_dc");
  if (flux_split(c, sh)) {
    const int nb = blocks_for(u.total());
    if (c->dims == 3)
      hipLaunchKernelGGL((k_flux3<ORDER, RS, false, CP, 3>), dim3(3 * nb), dim3(BLK), 0,
                         c->stream, c->g, w, bf, c->fx, c->fy, c->fz, u, L, 0, c->dev, dtf);
    else
      hipLaunchKernelGGL((k_flux3<ORDER, RS, false, CP, 2>), dim3(2 * nb), dim3(BLK), 0,
                         c->stream, c->g, w, bf, c->fx, c->fy, c->fz, u, L, 0, c->dev, dtf);
  } else if (sh.tiled)
    hipLaunchKernelGGL((k_flux3<ORDER, RS, true, CP>), dim3(sh.blocks), dim3(BLK), 0, c->stream,
                       c->g, w, bf, c->fx, c->fy, c->fz, u, L, 0, c->dev, dtf);
  else
    hipLaunchKernelGGL((k_flux3<ORDER, RS, false, CP>), dim3(sh.blocks), dim3(BLK), 0, c->stream,
                       c->g, w, bf, c->fx, c->fy, c->fz, u, L, sh.ty, c->dev, dtf);
}

// PLM limits in characteristic variables when the context asks for it
template <int ORDER, int RS>
void launch_fluxes(bnz_ctx *c, const double *w, const double *bf, const Lims &L) {
  if (ORDER == 1 && c->p.char_proj) launch_fluxes_cp<ORDER, RS, true>(c, w, bf, L);
  else launch_fluxes_cp<ORDER, RS, false>(c, w, bf, L);
}

void launch_flux_dispatch(bnz_ctx *c, int order, const double *w, const double *bf,
                          const Lims &L) {
  const int rs = rs_kern(c);
  if (order == 0) {
    if (rs == BNZ_RS_HLLD) launch_fluxes<0, BNZ_RS_HLLD>(c, w, bf, L);
    else if (rs == BNZ_RS_HLLT) launch_fluxes<0, BNZ_RS_HLLT>(c, w, bf, L);
    else if (rs == BNZ_RS_HLLC) launch_fluxes<0, BNZ_RS_HLLC>(c, w, bf, L);
    else if (rs == RS_HLL_HYDRO) launch_fluxes<0, RS_HLL_HYDRO>(c, w, bf, L);
    else launch_fluxes<0, BNZ_RS_HLL>(c, w, bf, L);
  } else {
    if (rs == BNZ_RS_HLLD) launch_fluxes<1, BNZ_RS_HLLD>(c, w, bf, L);
    else if (rs == BNZ_RS_HLLT) launch_fluxes<1, BNZ_RS_HLLT>(c, w, bf, L);
    else if (rs == BNZ_RS_HLLC) launch_fluxes<1, BNZ_RS_HLLC>(c, w, bf, L);
    else if (rs == RS_HLL_HYDRO) launch_fluxes<1, RS_HLL_HYDRO>(c, w, bf, L);
    else launch_fluxes<1, BNZ_RS_HLL>(c, w, bf, L);
  }
}

void launch_emf(bnz_ctx *c, const double *w, const Lims &L, const char *name) {
  Box e;
  e.lo[0] = L.lo[0]; e.n[0] = L.hi[0] + 1 - L.lo[0] + 1;
  e.lo[1] = (c->dims >= 2) ? L.lo[1] : 0; e.n[1] = (c->dims >= 2) ? L.hi[1] + 1 - L.lo[1] + 1 : 1;
  e.lo[2] = (c->dims >= 3) ? L.lo[2] : 0; e.n[2] = (c->dims >= 3) ? L.hi[2] + 1 - L.lo[2] + 1 : 1;
  const Shape sh = flat_order(c, work_shape(e, c->dims, c->emf_chunk));
  Launch t(c, name);
#define EMF_ARGS                                                                              \
  dim3(sh.blocks), dim3(BLK), 0, c->stream, c->g, w, c->fx, c->fy, c->fz, c->ee, e, sh.ty
  // (1D grids reach here only with RK3: the VL step of a 1D grid is k_step1)
  if (c->p.ct_upwind) {
    if (sh.tiled) hipLaunchKernelGGL((k_emf_upw<3, true>), EMF_ARGS, c->dev);
    else if (c->dims == 3) hipLaunchKernelGGL((k_emf_upw<3, false>), EMF_ARGS, c->dev);
    else if (c->dims == 2) hipLaunchKernelGGL((k_emf_upw<2, false>), EMF_ARGS, c->dev);
    else hipLaunchKernelGGL((k_emf_upw<1, false>), EMF_ARGS, c->dev);
  } else if (sh.tiled) {
    hipLaunchKernelGGL((k_emf<3, true>), EMF_ARGS);
  } else if (c->dims == 3) {
    hipLaunchKernelGGL((k_emf<3, false>), EMF_ARGS);
  } else if (c->dims == 2) {
    hipLaunchKernelGGL((k_emf<2, false>), EMF_ARGS);
  } else {
    hipLaunchKernelGGL((k_emf<1, false>), EMF_ARGS);
  }
#undef EMF_ARGS
}

// One VL step after the dt kernels (vanleer.pyx:100-163).
// we: with emf_fly, the prim the stage's EMFs come from
void launch_update(bnz_ctx *c, bool pred, const Lims &L, const double *we = nullptr) {
  double *w1 = pred ? c->ws : c->w;
  Box grid;
  for (int q = 0; q < 3; ++q) {
    grid.lo[q] = 0;
    grid.n[q] = c->nt[q];
  }
  const Box act = active_box(c);
  const int fr = cfl_mode(c);
  const Shape sh = flat_order(c, work_shape(grid, c->dims, c->update_chunk));
  Launch t(c, pred ? "update_pred" : "update_corr");
#define UPD_ARGS                                                                              \
  dim3(sh.blocks), dim3(BLK), 0, c->stream, c->g, c->w, c->fx, c->fy, c->fz, c->ee, c->b, c->bs, \
      w1, L, c->dev, grid, act, fr, c->part, sh.ty
  // a uniform grid (every cell width bit for bit the same) takes its
  // geometry from kernel arguments instead of per-cell loads
  const bool ug = c->g.unif != 0;
  if (sh.tiled) {  // 3D only
    if (pred) hipLaunchKernelGGL((k_update<3, true, false, true>), UPD_ARGS);
    else hipLaunchKernelGGL((k_update<3, false, true, true>), UPD_ARGS);
  } else if (c->dims == 3 && ug) {
    if (pred) hipLaunchKernelGGL((k_update<3, true, false, false, true>), UPD_ARGS);
    else hipLaunchKernelGGL((k_update<3, false, true, false, true>), UPD_ARGS);
  } else if (c->dims == 3) {
    if (pred) hipLaunchKernelGGL((k_update<3, true, false, false>), UPD_ARGS);
    else hipLaunchKernelGGL((k_update<3, false, true, false>), UPD_ARGS);
  } else if (we) {  // 2D, EMFs formed in the update
    if (pred) hipLaunchKernelGGL((k_update<2, true, false, false, false, true>), UPD_ARGS, we);
    else hipLaunchKernelGGL((k_update<2, false, true, false, false, true>), UPD_ARGS, we);
  } else {  // 2D (1D grids run k_step1)
    if (pred) hipLaunchKernelGGL((k_update<2, true, false, false>), UPD_ARGS);
    else hipLaunchKernelGGL((k_update<2, false, true, false>), UPD_ARGS);
  }
#undef UPD_ARGS
  // the corrector leaves the next step's CFL partials in part[0..blocks)
  c->dt_parts = pred ? 0 : sh.blocks;
}

// A 1D step as one fused launch (k_step1): w, b -> ws, bs, then swapped.
template <int ORDER, int RS>
void step1_launch(bnz_ctx *c, const Lims &L0, const Lims &L1) {
  const int nb = (c->nt[0] + STEP1_CELLS - 1) / STEP1_CELLS;
  Launch t(c, "step1");
  if (ORDER == 1 && c->p.char_proj)
    hipLaunchKernelGGL((k_step1<ORDER, RS, true>), dim3(nb), dim3(BLK), 0, c->stream, c->g, c->w,
                       c->b, c->ws, c->bs, c->fx, L0, L1, c->dev, active_box(c),
                       cfl_mode(c), c->part);
  else
    hipLaunchKernelGGL((k_step1<ORDER, RS, false>), dim3(nb), dim3(BLK), 0, c->stream, c->g, c->w,
                       c->b, c->ws, c->bs, c->fx, L0, L1, c->dev, active_box(c),
                       cfl_mode(c), c->part);
  c->dt_parts = nb;
  std::swap(c->w, c->ws);  // the faces swap with the other stages' below
}

void launch_step1(bnz_ctx *c, const Lims &L0, const Lims &L1) {
  const int rs = rs_kern(c);
  if (c->order == 0) {
    if (rs == BNZ_RS_HLLD) step1_launch<0, BNZ_RS_HLLD>(c, L0, L1);
    else if (rs == BNZ_RS_HLLT) step1_launch<0, BNZ_RS_HLLT>(c, L0, L1);
    else if (rs == BNZ_RS_HLLC) step1_launch<0, BNZ_RS_HLLC>(c, L0, L1);
    else if (rs == RS_HLL_HYDRO) step1_launch<0, RS_HLL_HYDRO>(c, L0, L1);
    else step1_launch<0, BNZ_RS_HLL>(c, L0, L1);
  } else {
    if (rs == BNZ_RS_HLLD) step1_launch<1, BNZ_RS_HLLD>(c, L0, L1);
    else if (rs == BNZ_RS_HLLT) step1_launch<1, BNZ_RS_HLLT>(c, L0, L1);
    else if (rs == BNZ_RS_HLLC) step1_launch<1, BNZ_RS_HLLC>(c, L0, L1);
    else if (rs == RS_HLL_HYDRO) step1_launch<1, RS_HLL_HYDRO>(c, L0, L1);
    else step1_launch<1, BNZ_RS_HLL>(c, L0, L1);
  }
}

// One RK3 stage update launch: conserved input u0 and faces b0, results to
// u1 and b1 (and prim w), combined with U (uc) and B (b) in stages 2, 3.
template <int STAGE>
void launch_update_rk(bnz_ctx *c, const double *u0, const double *b0, double *u1, double *b1,
                      const Lims &L) {
  Box grid;
  for (int q = 0; q < 3; ++q) {
    grid.lo[q] = 0;
    grid.n[q] = c->nt[q];
  }
  const Box act = active_box(c);
  const int nb = blocks_for(grid.total());
  Launch t(c, STAGE == 1 ? "update_rk1" : STAGE == 2 ? "update_rk2" : "update_rk3");
#define RK_ARGS                                                                                 \
  dim3(nb), dim3(BLK), 0, c->stream, c->g, u0, c->fx, c->fy, c->fz, c->ee, b0, c->uc, c->b, u1, \
      b1, c->w, L, c->dev, grid, act, cfl_mode(c), c->part
  if (c->dims == 3) hipLaunchKernelGGL((k_update_rk<3, STAGE>), RK_ARGS);
  else if (c->dims == 2) hipLaunchKernelGGL((k_update_rk<2, STAGE>), RK_ARGS);
  else hipLaunchKernelGGL((k_update_rk<1, STAGE>), RK_ARGS);
#undef RK_ARGS
  if (STAGE == 3) c->dt_parts = nb;
}

// One RK3 step after the dt kernels (_rk3.pyx:56-167): U from the prim, the
// three stages on the limits of the reconstruction order, the ghost fill.
void launch_rk3_body(bnz_ctx *c) {
  const Lims L = integr_lims(c, c->order);
  {
    Launch t(c, "cons_of_prim");
    hipLaunchKernelGGL(k_cons_of_prim, dim3(blocks_for(c->ncell)), dim3(BLK), 0, c->stream,
                       c->g, c->w, c->uc, c->dev);
  }
  // I: U* = U + dt L(W, B)
  launch_flux_dispatch(c, c->order, c->w, c->b, L);
  launch_emf(c, c->w, L, "emf_rk");
  launch_update_rk<1>(c, c->uc, c->b, c->ws, c->bs, L);
  // II: U** = 3/4 U + 1/4 (U* + dt L(W*, B*))
  launch_flux_dispatch(c, c->order, c->w, c->bs, L);
  launch_emf(c, c->w, L, "emf_rk");
  launch_update_rk<2>(c, c->ws, c->bs, c->uss, c->bss, L);
  // III: U = 1/3 U + 2/3 (U** + dt L(W**, B**)), in place on U and B
  launch_flux_dispatch(c, c->order, c->w, c->bss, L);
  launch_emf(c, c->w, L, "emf_rk");
  launch_update_rk<3>(c, c->uss, c->bss, c->uc, c->b, L);
  launch_ghost_step64(c);  // and the time advance
}

// One VL step after the dt kernels (vanleer.pyx:100-163).
void launch_step_body(bnz_ctx *c) {
  if (c->p.tintegr) {
    launch_rk3_body(c);
    return;
  }
  Lims L0 = integr_lims(c, 0);
  Lims L1 = integr_lims(c, c->order);
  if (c->dims == 1) {  // the whole step through LDS
    launch_step1(c, L0, L1);
  } else {
    // predictor: donor cell, dt/2, lims of order 0; ws and bs are new buffers
    launch_flux_dispatch(c, 0, c->w, c->b, L0);
    const bool fly = emf_fly(c);
    if (!fly) launch_emf(c, c->w, L0, "emf_pred");
    launch_update(c, true, L0, fly ? c->w : nullptr);
    // corrector: reconstruction order, dt; prim is updated in place (each
    // cell reads only its own w_n), b_{n+1} goes to the predictor's face
    // buffer, which is dead once the corrector fluxes and EMFs exist
    launch_flux_dispatch(c, c->order, c->ws, c->bs, L1);
    if (!fly) launch_emf(c, c->ws, L1, "emf_corr");
    launch_update(c, false, L1, fly ? c->ws : nullptr);
  }
  std::swap(c->b, c->bs);
  launch_ghost_step64(c);  // and the time advance
}

}  // namespace
