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
#include <type_traits>

#include "bnz_kern32.hpp"

using namespace bnz;

namespace {

// All face directions of a stage in one launch, one thread per cell of the
// union of the face ranges (as the FP64 k_flux3); the step's first one also
// reduces its dt in workgroup 0 (DtFold).
// SPLIT > 0 (small grids, flat order): workgroup b computes the faces of
// direction b % SPLIT only (as the FP64 k_flux3).
template <int ORDER, int RS, typename BF, int SPLIT = 0>
__global__ __launch_bounds__(BLK) void k32_flux3(Geo g, const float *__restrict__ w,
                                                 const BF *__restrict__ bf,
                                                 float *__restrict__ fx, float *__restrict__ fy,
                                                 float *__restrict__ fz, Box box, Lims L,
                                                 int *nguard, int ty, double *dev, DtFold dtf) {
  if (dtf.npart > 0 && blockIdx.x == 0) dt_reduce_block(dtf, dev);
  if (step_past_tmax(dtf, dev)) return;  // a step past t_max (see k_update)
  int i, j, k;
  int ng = 0;
  if (SPLIT > 0) {
    const long lb = xcd_block();
    const int dir = (int)(lb % SPLIT);
    const long t = (lb / SPLIT) * BLK + threadIdx.x;
    if (t >= box.total()) return;
    box_idx(box, t, i, j, k);
    // the faces flux32_cell computes for this cell, one direction
    if (dir == 0) {
      if (i >= L.lo[0]) flux32_face<0, ORDER, RS>(g, w, bf, fx, i, j, k, ng);
    } else if (dir == 1) {
      if (j >= L.lo[1]) flux32_face<1, ORDER, RS>(g, w, bf, fy, i, j, k, ng);
    } else {
      if (k >= L.lo[2]) flux32_face<2, ORDER, RS>(g, w, bf, fz, i, j, k, ng);
    }
    if (ng) atomicAdd(nguard, ng);
    return;
  }
  if (!cell_index<false, 1, 2>(box, ty, i, j, k)) return;
  flux32_cell<ORDER, RS>(g, w, bf, fx, fy, fz, L, i, j, k, ng);
  if (ng) atomicAdd(nguard, ng);
}

// Edge EMFs, FP32 (emf32_edge).
template <int DIMS>
__global__ __launch_bounds__(BLK) void k32_emf(Geo g, const float *__restrict__ w,
                                               const float *__restrict__ fx,
                                               const float *__restrict__ fy,
                                               const float *__restrict__ fz,
                                               float *__restrict__ ee, Box box, Walls wl,
                                               int ty) {
  int i, j, k;
  if (!cell_index<false, 1, 2>(box, ty, i, j, k)) return;
  emf32_edge<DIMS>(g, w, fx, fy, fz, ee, wl, i, j, k);
}

// Predictor (update32_pred), one thread per cell of the grid.
// FLY (2D): the edge EMFs formed from the stage's prim we (Emf32Fly2D).
template <int DIMS, bool UG = false, bool FLY = false>
__global__ __launch_bounds__(BLK) void k32_update_pred(Geo g, Pred32 a, Lims L,
                                                       const double *__restrict__ dts,
                                                       Box grid, int ty, const float *we = nullptr,
                                                       Walls wl = Walls{}) {
  int i, j, k;
  if (dts[0] == 0.) return;  // a step past t_max (see k_update)
  if (!cell_index<false, 1, 2>(grid, ty, i, j, k)) return;
  update32_pred<DIMS, typename std::conditional<UG, GeoUni, GeoArr>::type, FLY>(g, a, L, dts, i,
                                                                               j, k, we, wl);
}

// Corrector (update32_corr) with the CFL maximum reduced per block.
template <int DIMS, bool UG = false, bool FLY = false>
__global__ __launch_bounds__(BLK) void k32_update_corr(Geo g, Corr32 a, Lims L,
                                                       const double *__restrict__ dts, Box act,
                                                       int fms_ref, double *__restrict__ part,
                                                       Box grid, int ty, const float *we = nullptr,
                                                       Walls wl = Walls{}) {
  int i, j, k;
  double cdsmax = 0.;
  const bool have = cell_index<false, 1, 2>(grid, ty, i, j, k);
  if (dts[0] == 0.) {  // a step past t_max: state unchanged, faces carried over
    if (have) {
      const long c = ((long)k * g.nt[1] + j) * g.nt[0] + i;
      for (int q = 0; q < 3; ++q) a.b1[q * g.sv + c] = a.b[q * g.sv + c];
    }
    return;
  }
  if (have)
    cdsmax = update32_corr<DIMS, typename std::conditional<UG, GeoUni, GeoArr>::type, FLY>(
        g, a, L, dts, act, fms_ref, i, j, k, we, wl);
  block_max_to(cdsmax, part);
}

// One FP32C stage of a 1D grid in one pass (stage32_cell): workgroup b owns
// faces b * (BLK - 1) .. + BLK - 1 and the BLK - 1 cells between them; the
// corrector leaves its workgroup's CFL maximum in part.
template <bool PRED, int ORDER, int RS>
__global__ __launch_bounds__(BLK) void k32_stage1(Geo g, Pred32 ap, Corr32 ac, Lims L,
                                                  const double *__restrict__ dts, Box act,
                                                  int fms_ref, double *__restrict__ part,
                                                  Walls wl, int *nguard) {
  __shared__ float sf[NMODE * BLK];
  const int i = blockIdx.x * (BLK - 1) + threadIdx.x;
  if (dts[0] == 0.) {  // a step past t_max: state unchanged, faces carried over
    if (!PRED && threadIdx.x < BLK - 1 && i < g.nt[0])
      for (int q = 0; q < 3; ++q) ac.b1[q * g.sv + i] = ac.b[q * g.sv + i];
    return;
  }
  int ng = 0;
  const double cds = stage32_cell<PRED, ORDER, RS>(g, ap.ws, ap.bs, ap, ac, L, dts, act,
                                                   fms_ref, wl.at[0], wl.at[1], i, sf, ng);
  if (!PRED) block_max_to(cds, part);
  if (ng) atomicAdd(nguard, ng);
}

// CFL maximum from the FP32 primitives (first step after an upload).
__global__ __launch_bounds__(BLK) void k32_dt_partial(Geo g, const float *__restrict__ w, Box box,
                                                      int fms_ref, double *__restrict__ part) {
  int i, j, k;
  double cds = 0.;
  if (cell_index<false, 1, 2>(box, 0, i, j, k)) cds = cfl32_cell(g, w, fms_ref, i, j, k);
  block_max_to(cds, part);
}

// adv: the time and step counter a step's fill advances (one launch of it)
__global__ __launch_bounds__(BLK) void k32_ghost(Geo g, Ghost32 p, const int2 *__restrict__ pairs,
                                                 int npairs, double *adv) {
  time_advance_once(adv);
  const long t = xcd_block() * BLK + threadIdx.x;
  if (t >= npairs) return;
  const int2 ds = pairs[t];
  ghost32_pair(p, ds.x, ds.y);
}

__global__ __launch_bounds__(BLK) void k32_ghost_ops(Geo g, Ghost32 p,
                                                     const GhostOp *__restrict__ ops, int nops,
                                                     double *adv) {
  time_advance_once(adv);
  const long t = xcd_block() * BLK + threadIdx.x;
  if (t >= nops) return;
  ghost32_op(p, ops[t]);
}

// FP32C state of the user fixed-value rows (NMODE prim + 3 faces each): the
// conserved hi words and FP32 prim a cell of that prim gets from k32_from64.
__global__ __launch_bounds__(BLK) void k32_fixed_rows(const double *__restrict__ fix,
                                                      float *fix32, int rows, double gamm1i) {
  const long t = (long)blockIdx.x * BLK + threadIdx.x;
  if (t >= rows) return;
  double wv[NMODE], u[NMODE];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) wv[n] = fix[t * (NMODE + 3) + n];
  prim2cons(u, wv, gamm1i);
#pragma unroll
  for (int n = 0; n < 6; ++n) fix32[t * NGHOST32 + n] = (float)u[n];
#pragma unroll
  for (int n = 0; n < NMODE; ++n) fix32[t * NGHOST32 + 6 + n] = (float)wv[n];
}

// FP64 prim (all cells) -> FP32-compensated state.
__global__ __launch_bounds__(BLK) void k32_from64(Geo g, const double *__restrict__ w, float *uh,
                                                  float *ul, float *w32) {
  const long t = xcd_block() * BLK + threadIdx.x;
  const long ncell = (long)g.nt[0] * g.nt[1] * g.nt[2];
  if (t >= ncell) return;
  from64_cell(g, w, uh, ul, w32, t);
}

// FP32-compensated state -> FP64 prim of the active cells (the caller fills
// the ghosts).
__global__ __launch_bounds__(BLK) void k32_to64(Geo g, const float *uh, const float *ul,
                                                const double *b, double *w, Box act) {
  int i, j, k;
  if (!cell_index<false, 1, 2>(act, 0, i, j, k)) return;
  to64_cell(g, uh, ul, b, w, i, j, k);
}

template <int ORDER, int RS, typename BF>
void launch_fluxes32(bnz_ctx *c, const float *w, const BF *bf, const Lims &L) {
  int *ng = (int *)(c->dev + 4);
  const Box u = flux3_box(c->dims, L);
  const Shape sh = flat_order(c, work_shape32(u, c->dims, c->flux_chunk));
  const DtFold dtf = take_dt(c);
  Launch t(c, ORDER ? "flux3_plm" : "flux3_dc");
  if (flux_split(c, sh)) {
    const int nb = blocks_for(u.total());
    if (c->dims == 3)
      hipLaunchKernelGGL((k32_flux3<ORDER, RS, BF, 3>), dim3(3 * nb), dim3(BLK), 0, c->stream,
                         c->g, w, bf, c->fx32, c->fy32, c->fz32, u, L, ng, 0, c->dev, dtf);
    else
      hipLaunchKernelGGL((k32_flux3<ORDER, RS, BF, 2>), dim3(2 * nb), dim3(BLK), 0, c->stream,
                         c->g, w, bf, c->fx32, c->fy32, c->fz32, u, L, ng, 0, c->dev, dtf);
  } else {
    hipLaunchKernelGGL((k32_flux3<ORDER, RS, BF>), dim3(sh.blocks), dim3(BLK), 0, c->stream,
                       c->g, w, bf, c->fx32, c->fy32, c->fz32, u, L, ng, sh.ty, c->dev, dtf);
  }
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

// Reflecting wall planes of the grid (their in-plane EMF components are zero).
Walls walls32(const bnz_ctx *c) {
  Walls wl;
  const int lo[3] = {c->i1, c->j1, c->k1}, hi[3] = {c->i2, c->j2, c->k2};
  for (int a = 0; a < 3; ++a) {
    wl.at[2 * a] = (a < c->dims && c->p.bc[2 * a] == BNZ_BC_REFLECT) ? lo[a] : -1;
    wl.at[2 * a + 1] = (a < c->dims && c->p.bc[2 * a + 1] == BNZ_BC_REFLECT) ? hi[a] + 1 : -1;
  }
  return wl;
}

void emf32(bnz_ctx *c, const float *w, const Lims &L, const char *name) {
  Box e;
  e.lo[0] = L.lo[0]; e.n[0] = L.hi[0] + 1 - L.lo[0] + 1;
  e.lo[1] = (c->dims >= 2) ? L.lo[1] : 0; e.n[1] = (c->dims >= 2) ? L.hi[1] + 1 - L.lo[1] + 1 : 1;
  e.lo[2] = (c->dims >= 3) ? L.lo[2] : 0; e.n[2] = (c->dims >= 3) ? L.hi[2] + 1 - L.lo[2] + 1 : 1;
  Launch t(c, name);
  const Walls wl = walls32(c);
  const Shape sh = flat_order(c, work_shape32(e, c->dims, c->emf_chunk));
#define E32_ARGS \
  dim3(sh.blocks), dim3(BLK), 0, c->stream, c->g, w, c->fx32, c->fy32, c->fz32, c->ee32, e, wl, sh.ty
  if (c->dims == 3) hipLaunchKernelGGL(k32_emf<3>, E32_ARGS);
  else hipLaunchKernelGGL(k32_emf<2>, E32_ARGS);  // 1D grids run k32_stage1
#undef E32_ARGS
}

// The FP32C ghost fill; a step's fill (step) also advances the time and
// step counter.
void ghost32(bnz_ctx *c, bool step = false) {
  Ghost32 p;
  int q = 0;
  for (int n = 0; n < 6; ++n) p.a[q++] = c->uh + n * c->ncell;
  for (int n = 0; n < 9; ++n) p.a[q++] = c->w32 + n * c->ncell;
  for (int n = 0; n < 3; ++n) p.d[n] = c->b + n * c->ncell;
  p.fix32 = c->ufix32;
  p.fix64 = c->ufix;
  double *adv = step ? c->dev : nullptr;
  {
    Launch t(c, "ghost");
    if (c->npass > 0) {  // user boundaries: one pass per side, in order
      for (int s = 0; s < 6; ++s) {
        const int n = c->pass_off[s + 1] - c->pass_off[s];
        if (n > 0) {
          hipLaunchKernelGGL(k32_ghost_ops, dim3(blocks_for(n)), dim3(BLK), 0, c->stream, c->g,
                             p, c->gops + c->pass_off[s], n, adv);
          adv = nullptr;
        }
      }
    } else {
      if (c->nghost > 0) {
        hipLaunchKernelGGL(k32_ghost, dim3(blocks_for(c->nghost)), dim3(BLK), 0, c->stream, c->g,
                           p, c->ghost, c->nghost, adv);
        adv = nullptr;
      }
      if (c->ngops > 0) {
        hipLaunchKernelGGL(k32_ghost_ops, dim3(blocks_for(c->ngops)), dim3(BLK), 0, c->stream,
                           c->g, p, c->gops, c->ngops, adv);
        adv = nullptr;
      }
    }
  }
  if (adv) launch_time_advance(c);
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
  float *bufs[] = {c->uh,   c->ul,   c->w32,  c->ws32,  c->bs32,
                   c->fx32, c->fy32, c->fz32, c->ee32, c->ufix32};
  for (float *q : bufs)
    if (q) (void)hipFree(q);
}

int f32_from64(bnz_ctx *c) {
  hipLaunchKernelGGL(k32_from64, dim3(blocks_for(c->ncell)), dim3(BLK), 0, c->stream, c->g, c->w,
                     c->uh, c->ul, c->w32);
  c->dt_parts = 0;
  c->pdt_pending = false;
  return check_launch(c);
}

int f32_to64(bnz_ctx *c) {
  const Box act = active_box(c);
  hipLaunchKernelGGL(k32_to64, dim3(blocks_for(act.total())), dim3(BLK), 0, c->stream, c->g,
                     c->uh, c->ul, c->b, c->w, act);
  launch_ghost64(c);
  return check_launch(c);
}

void f32_ghost(bnz_ctx *c) { ghost32(c); }

int f32_fixed_rows(bnz_ctx *c, const double *ufix, int rows, float **out) {
  HIPCHK(c, hipMalloc(out, (size_t)std::max(1, rows) * NGHOST32 * sizeof(float)));
  if (rows > 0)
    hipLaunchKernelGGL(k32_fixed_rows, dim3(blocks_for(rows)), dim3(BLK), 0, c->stream, ufix,
                       *out, rows, c->g.gamm1i);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

void f32_dt(bnz_ctx *c, int use_fixed, double fixed_dt) {
  int nb = c->dt_parts;
  if (nb <= 0) {
    Box a = active_box(c);
    nb = blocks_for(a.total());
    Launch t(c, "dt_partial");
    hipLaunchKernelGGL(k32_dt_partial, dim3(nb), dim3(BLK), 0, c->stream, c->g, c->w32, a,
                       c->p.dt_fms_ref, c->part);
  }
  defer_dt_final(c, nb, use_fixed, fixed_dt);
}

template <bool PRED, int ORDER, int RS>
void stage32_launch(bnz_ctx *c, const Lims &L) {
  const int nb = (c->nt[0] + BLK - 2) / (BLK - 1);
  const Pred32 ap{c->w32, c->uh, c->b, c->fx32, c->fy32, c->fz32, c->ee32, c->bs32, c->ws32};
  const Corr32 ac{c->uh, c->ul, c->b, c->fx32, c->fy32, c->fz32, c->ee32, c->bs, c->w32};
  Launch t(c, PRED ? "stage1_pred" : "stage1_corr");
  hipLaunchKernelGGL((k32_stage1<PRED, ORDER, RS>), dim3(nb), dim3(BLK), 0, c->stream, c->g, ap,
                     ac, L, c->dev, active_box(c), c->p.dt_fms_ref, c->part, walls32(c),
                     (int *)(c->dev + 4));  // guarded-face counter, as launch_fluxes32
  c->dt_parts = PRED ? 0 : nb;
}

template <bool PRED>
void stage32(bnz_ctx *c, int order, const Lims &L) {
  const bool d = c->p.rsolver == BNZ_RS_HLLD;
  if (order == 0) {
    if (d) stage32_launch<PRED, 0, BNZ_RS_HLLD>(c, L);
    else stage32_launch<PRED, 0, BNZ_RS_HLL>(c, L);
  } else {
    if (d) stage32_launch<PRED, 1, BNZ_RS_HLLD>(c, L);
    else stage32_launch<PRED, 1, BNZ_RS_HLL>(c, L);
  }
}

void f32_step_body(bnz_ctx *c) {
  if (c->dims == 1) {  // fluxes through LDS
    stage32<true>(c, 0, integr_lims(c, 0));
    stage32<false>(c, c->order, integr_lims(c, c->order));
    std::swap(c->b, c->bs);
    ghost32(c, true);
    return;
  }
  Box grid;
  for (int q = 0; q < 3; ++q) {
    grid.lo[q] = 0;
    grid.n[q] = c->nt[q];
  }
  const Shape us = flat_order(c, work_shape32(grid, c->dims, c->update_chunk));
  const int nb = us.blocks;
  Lims L0 = integr_lims(c, 0);
  const bool fly = emf_fly(c);
  const Walls wl = walls32(c);
  fluxes32(c, 0, c->w32, c->b, L0);
  if (!fly) emf32(c, c->w32, L0, "emf_pred");
  {
    Pred32 a{c->w32, c->uh, c->b, c->fx32, c->fy32, c->fz32, c->ee32, c->bs32, c->ws32};
    Launch t(c, "update_pred");
#define P32_ARGS dim3(nb), dim3(BLK), 0, c->stream, c->g, a, L0, c->dev, grid, us.ty
    if (c->dims == 3 && c->g.unif) hipLaunchKernelGGL((k32_update_pred<3, true>), P32_ARGS);
    else if (c->dims == 3) hipLaunchKernelGGL(k32_update_pred<3>, P32_ARGS);
    else if (fly)
      hipLaunchKernelGGL((k32_update_pred<2, false, true>), P32_ARGS, (const float *)c->w32, wl);
    else hipLaunchKernelGGL(k32_update_pred<2>, P32_ARGS);
#undef P32_ARGS
  }
  Lims L1 = integr_lims(c, c->order);
  fluxes32(c, c->order, c->ws32, c->bs32, L1);
  if (!fly) emf32(c, c->ws32, L1, "emf_corr");
  {
    Corr32 a{c->uh, c->ul, c->b, c->fx32, c->fy32, c->fz32, c->ee32, c->bs, c->w32};
    Launch t(c, "update_corr");
#define C32_ARGS \
  dim3(nb), dim3(BLK), 0, c->stream, c->g, a, L1, c->dev, active_box(c), c->p.dt_fms_ref, c->part, \
      grid, us.ty
    if (c->dims == 3 && c->g.unif) hipLaunchKernelGGL((k32_update_corr<3, true>), C32_ARGS);
    else if (c->dims == 3) hipLaunchKernelGGL(k32_update_corr<3>, C32_ARGS);
    else if (fly)
      hipLaunchKernelGGL((k32_update_corr<2, false, true>), C32_ARGS, (const float *)c->ws32, wl);
    else hipLaunchKernelGGL(k32_update_corr<2>, C32_ARGS);
#undef C32_ARGS
  }
  c->dt_parts = nb;
  std::swap(c->b, c->bs);
  ghost32(c, true);
}

}  // namespace bnz
