// bnz_batch32.hip -- FP32-compensated kernels of the batched-box step.
//
// The same work lists, box descriptors and per-box dt / t / run flags as the
// FP64 batch (bnz_batch.hip), with the per-item bodies of the single-grid
// FP32C step (bnz_kern32.hpp), so that a box of an FP32C batch follows
// exactly the arithmetic of an FP32C single grid.  Built, like
// bnz_gpu32.hip, with FMA contraction of single expressions (-ffp-contract=on).
#include <hip/hip_runtime.h>

#include "bnz_batch.hpp"
#include "bnz_kern32.hpp"

using namespace bnz;
using namespace bnzb;

namespace {

__device__ __forceinline__ Walls box_walls(const BoxDev &d) {
  Walls wl;
#pragma unroll
  for (int q = 0; q < 6; ++q) wl.at[q] = d.walls[q];
  return wl;
}

// The flux, EMF and update kernels are compiled per dimensionality and
// launched over the work blocks of the boxes of that dimensionality.
template <int ORDER, int RS, typename BF, int DIMS>
__global__ __launch_bounds__(BLK) void kb32_flux(const BoxDev *__restrict__ boxes, WList wl,
                                                 BoxState st, const float *__restrict__ w,
                                                 const BF *__restrict__ bf,
                                                 float *__restrict__ fx, float *__restrict__ fy,
                                                 float *__restrict__ fz, int *nguard) {
  int ng = 0;
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
    flux32_cell<ORDER, RS, BF, DIMS>(d.g, w + o, bf + o, fx + o, fy + o, fz + o,
                                     d.L[wl.kt == KF0 ? 0 : 1], i, j, k, ng);
  }
  if (ng) atomicAdd(nguard, ng);
}

template <int DIMS>
__global__ __launch_bounds__(BLK) void kb32_emf(const BoxDev *__restrict__ boxes, WList wl,
                                                BoxState st, const float *__restrict__ w,
                                                const float *__restrict__ fx,
                                                const float *__restrict__ fy,
                                                const float *__restrict__ fz,
                                                float *__restrict__ ee) {
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
    emf32_edge<DIMS>(d.g, w + o, fx + o, fy + o, fz + o, ee + o, box_walls(d), i, j, k);
  }
}

// Stage update of every cell of the running boxes (update32_pred /
// update32_corr); the corrector folds the CFL maximum into the box's cmax.
// A finished box is left alone (its faces: BoxState::fpar, kb_face_fix).
template <bool PRED, int DIMS>
__global__ __launch_bounds__(BLK) void kb32_update(const BoxDev *__restrict__ boxes, WList wl,
                                                   BoxState st, Pools P, Pools32 Q,
                                                   int fms_ref) {
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
      if (PRED) {
        const Pred32 a{Q.w32 + o, Q.uh + o, P.b + o, Q.fx + o, Q.fy + o,
                       Q.fz + o,  Q.ee + o, Q.bs32 + o, Q.ws32 + o};
        update32_pred<DIMS>(d.g, a, d.L[0], st.dt + b, i, j, k);
      } else {
        const Corr32 a{Q.uh + o, Q.ul + o, P.b + o, Q.fx + o, Q.fy + o,
                       Q.fz + o, Q.ee + o, P.bs + o, Q.w32 + o};
        cds = update32_corr<DIMS>(d.g, a, d.L[1], st.dt + b, d.act, fms_ref, i, j, k);
      }
    }
    if (!PRED) block_max_atomic(cds, st.cmax + b);
  }
}

// One FP32C stage of the 1D boxes in one pass (stage32_cell), as the FP64
// kb_stage1.
template <bool PRED, int ORDER, int RS>
__global__ __launch_bounds__(BLK) void kb32_stage1(const BoxDev *__restrict__ boxes, WList wl,
                                                   BoxState st, Pools P, Pools32 Q, int fms_ref,
                                                   int *nguard) {
  __shared__ float sf[NMODE * BLK];
  int ng = 0;
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    if (!st.run[b]) continue;
    const BoxDev &d = boxes[b];
    const long o = d.off;
    const int i = (wb - d.blk0[wl.kt]) * (BLK - 1) + threadIdx.x;
    const Pred32 ap{Q.w32 + o, Q.uh + o, P.b + o, Q.fx + o, Q.fy + o,
                    Q.fz + o,  Q.ee + o, Q.bs32 + o, Q.ws32 + o};
    const Corr32 ac{Q.uh + o, Q.ul + o, P.b + o, Q.fx + o, Q.fy + o,
                    Q.fz + o, Q.ee + o, P.bs + o, Q.w32 + o};
    const double cds = stage32_cell<PRED, ORDER, RS>(
        d.g, Q.ws32 + o, Q.bs32 + o, ap, ac, d.L[PRED ? 0 : 1], st.dt + b, d.act, fms_ref,
        d.walls[0], d.walls[1], i, sf, ng);
    if (!PRED)
      block_max_atomic(cds, st.cmax + b);  // ends with a barrier: sf is free again
    else
      __syncthreads();
  }
  if (ng) atomicAdd(nguard, ng);
}

// A whole FP32C VL step of a 1D box in one pass (all-1D batches), as the
// FP64 kb_step1: predictor fluxes and prims, then corrector fluxes, in LDS;
// the corrector reads the step's hi / lo words, FP32 prim and faces and
// writes the new ones to the other buffers (uh2, ul2, ws32, P.bs), which
// the caller swaps in.  The corrector's fluxes take the normal field from
// the FP64 faces rounded to FP32, the value the predictor's FP32 faces hold.
template <int ORDER, int RS>
__global__ __launch_bounds__(BLK) void kb32_step1(const BoxDev *__restrict__ boxes, WList wl,
                                                  BoxState st, Pools P, Pools32 Q, int fms_ref,
                                                  int *nguard) {
  constexpr int C = STEP1_CELLS;
  __shared__ float sf[NMODE * BLK];  // face fluxes: predictor, then corrector
  __shared__ float sw[NMODE * BLK];  // predictor FP32 prim of cell i0 - 2 + t
  const int t = threadIdx.x;
  int ng = 0;
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const int b = wl.bmap[wb];
    if (!st.run[b]) continue;
    const BoxDev &d = boxes[b];
    const Geo &g = d.g;
    const long o = d.off, sv = g.sv;
    const Lims &L0 = d.L[0], &L1 = d.L[1];
    const int i0 = (wb - d.blk0[wl.kt]) * C;
    const double *bf = P.b + o;
    const auto sink = [&](const float *f, long) {
#pragma unroll
      for (int n = 0; n < NMODE; ++n)
        if (n != BX) sf[n * BLK + t] = f[n];
    };
    const int ip = i0 - 2 + t;
    if (ip >= L0.lo[0] && ip <= L0.hi[0] + 1) {
      bool got = false;
      flux32_face_to<0, 0, RS>(g, Q.w32 + o, bf, ip, 0, 0, ng, [&](const float *f, long q) {
        sink(f, q);
        got = true;
      });
      if (!got)  // no region matched: the flux row keeps what it held
        for (int n = 0; n < NMODE; ++n)
          if (n != BX) sf[n * BLK + t] = Q.fx[n * sv + o + ip];
    }
    __syncthreads();
    if (t < BLK - 1 && ip >= 0 && ip < g.nt[0]) {
      const Pred32 ap{Q.w32 + o, Q.uh + o, bf, Q.fx + o, Q.fy + o, Q.fz + o, Q.ee + o,
                      nullptr, sw};
      update32_pred_e<1, GeoArr, FluxLds32, EmfLds32x1D, true>(
          g, ap, L0, st.dt + b, FluxLds32{sf, t}, EmfLds32x1D{sf, t, ip, d.walls[0], d.walls[1]},
          ip, 0, 0, t);
    }
    __syncthreads();
    const int ic = i0 + t;
    if (t <= C && ic >= L1.lo[0] && ic <= L1.hi[0] + 1) {
      Geo gl = g;  // the LDS prims: variable stride BLK, cell ip at slot ip - (i0 - 2)
      gl.sv = BLK;
      bool got = false;
      flux32_face_to<0, ORDER, RS>(gl, sw - (i0 - 2), bf, ic, 0, 0, ng,
                                   [&](const float *f, long q) {
                                     sink(f, q);
                                     got = true;
                                   });
      if (!got)
        for (int n = 0; n < NMODE; ++n)
          if (n != BX) sf[n * BLK + t] = Q.fx[n * sv + o + ic];
    }
    __syncthreads();
    double cds = 0.;
    if (t < C && ic < g.nt[0]) {
      const Corr32 ac{Q.uh + o, Q.ul + o, bf, Q.fx + o, Q.fy + o, Q.fz + o, Q.ee + o, P.bs + o,
                      Q.ws32 + o};
      cds = update32_corr_e<1, GeoArr, FluxLds32, EmfLds32x1D, true>(
          g, ac, L1, st.dt + b, d.act, fms_ref, FluxLds32{sf, t},
          EmfLds32x1D{sf, t, ic, d.walls[0], d.walls[1]}, ic, 0, 0, Q.uh2 + o, Q.ul2 + o);
    }
    block_max_atomic(cds, st.cmax + b);  // ends with a barrier: sf, sw are free again
  }
  if (ng) atomicAdd(nguard, ng);
}

// FP32C state (hi, lo, FP32 prim) of the boxes that sat out an odd number
// of fused steps, from the other buffers into the current ones (the faces:
// kb_face_fix).
__global__ __launch_bounds__(BLK) void kb32_fix(const BoxDev *__restrict__ boxes, WList wl,
                                                BoxState st, Pools32 Q) {
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const int bx = wl.bmap[wb];
    if (!st.fpar[bx]) continue;
    const BoxDev &d = boxes[bx];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    if (t >= d.items[wl.kt].total()) continue;
    const long sv = d.g.sv, c = d.off + t;
    for (int n = 0; n < NMODE - 3; ++n) {
      Q.uh[n * sv + c] = Q.uh2[n * sv + c];
      Q.ul[n * sv + c] = Q.ul2[n * sv + c];
    }
    for (int n = 0; n < NMODE; ++n) Q.w32[n * sv + c] = Q.ws32[n * sv + c];
  }
}

// FP64 prim -> FP32C state over every cell of every box (init).
__global__ __launch_bounds__(BLK) void kb32_from64(const BoxDev *__restrict__ boxes, WList wl,
                                                   Pools P, Pools32 Q) {
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const BoxDev &d = boxes[wl.bmap[wb]];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    if (t >= d.items[wl.kt].total()) continue;
    const long o = d.off;
    from64_cell(d.g, P.w + o, Q.uh + o, Q.ul + o, Q.w32 + o, t);
  }
}

// CFL maximum of the FP32 prims over the active cells (first step).
__global__ __launch_bounds__(BLK) void kb32_cfl(const BoxDev *__restrict__ boxes, WList wl,
                                                BoxState st, const float *__restrict__ w32,
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
      cds = cfl32_cell(d.g, w32 + d.off, fms_ref, i, j, k);
    }
    block_max_atomic(cds, st.cmax + b);
  }
}

// FP32C state -> FP64 prim of the active cells of every box (download).
__global__ __launch_bounds__(BLK) void kb32_to64(const BoxDev *__restrict__ boxes, WList wl,
                                                 Pools P, Pools32 Q) {
  for (int wb = wl.lo + blockIdx.x; wb < wl.hi; wb += gridDim.x) {
    const BoxDev &d = boxes[wl.bmap[wb]];
    const long t = (long)(wb - d.blk0[wl.kt]) * BLK + threadIdx.x;
    const Box &it = d.items[wl.kt];
    if (t >= it.total()) continue;
    int i, j, k;
    box_idx(it, t, i, j, k);
    const long o = d.off;
    to64_cell(d.g, Q.uh + o, Q.ul + o, P.b + o, P.w + o, i, j, k);
  }
}

// Ghost fill; with with_time the last workgroup ends the step instead
// (time_dt_block), as in the FP64 kb_ghost.
__global__ __launch_bounds__(BLK) void kb32_ghost(Ghost32 p, const int2 *__restrict__ pairs,
                                                  long npairs, BoxState st, int nbox,
                                                  double cour, int with_time) {
  if (with_time && blockIdx.x == gridDim.x - 1) {
    time_dt_block(st, nbox, cour);
    return;
  }
  const long nb = gridDim.x - with_time;
  for (long t = (long)blockIdx.x * BLK + threadIdx.x; t < npairs; t += nb * BLK) {
    const int2 ds = pairs[t];
    ghost32_pair(p, ds.x, ds.y);
  }
}

__global__ __launch_bounds__(BLK) void kb32_ghost_ops(Ghost32 p, const GhostOp *__restrict__ ops,
                                                      long nops, BoxState st, int nbox,
                                                      double cour, int with_time) {
  if (with_time && blockIdx.x == gridDim.x - 1) {
    time_dt_block(st, nbox, cour);
    return;
  }
  const long nb = gridDim.x - with_time;
  for (long t = (long)blockIdx.x * BLK + threadIdx.x; t < nops; t += nb * BLK)
    ghost32_op(p, ops[t]);
}

// fluxes of the 2D and 3D boxes (1D boxes: kb32_stage1)
template <int ORDER, int RS, typename BF>
void flux_launch(bnz_batch *c, const float *w, const BF *bf, int kt) {
  for (int d = 2; d <= 3; ++d) {
    const WList wl = wlist_dims(c, kt, d);
    if (wl.hi == wl.lo) continue;
    const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TF]));
    if (d == 2)
      hipLaunchKernelGGL((kb32_flux<ORDER, RS, BF, 2>), grid, dim3(BLK), 0, c->stream, c->boxes,
                         wl, c->S, w, bf, c->Q.fx, c->Q.fy, c->Q.fz, c->nguard);
    else
      hipLaunchKernelGGL((kb32_flux<ORDER, RS, BF, 3>), grid, dim3(BLK), 0, c->stream, c->boxes,
                         wl, c->S, w, bf, c->Q.fx, c->Q.fy, c->Q.fz, c->nguard);
  }
}

template <typename BF>
void fluxes(bnz_batch *c, int order, const float *w, const BF *bf, int kt) {
  const int rs = c->p.rsolver;
  if (order == 0) {
    if (rs == BNZ_RS_HLLD) flux_launch<0, BNZ_RS_HLLD>(c, w, bf, kt);
    else flux_launch<0, BNZ_RS_HLL>(c, w, bf, kt);
  } else {
    if (rs == BNZ_RS_HLLD) flux_launch<1, BNZ_RS_HLLD>(c, w, bf, kt);
    else flux_launch<1, BNZ_RS_HLL>(c, w, bf, kt);
  }
}

// EMFs of the 2D and 3D boxes (1D boxes form theirs in the update)
void emfs(bnz_batch *c, const float *w, int kt) {
  for (int d = 2; d <= 3; ++d) {
    const WList wl = wlist_dims(c, kt, d);
    if (wl.hi == wl.lo) continue;
    const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TE]));
    if (d == 2)
      hipLaunchKernelGGL(kb32_emf<2>, grid, dim3(BLK), 0, c->stream, c->boxes, wl, c->S, w,
                         c->Q.fx, c->Q.fy, c->Q.fz, c->Q.ee);
    else
      hipLaunchKernelGGL(kb32_emf<3>, grid, dim3(BLK), 0, c->stream, c->boxes, wl, c->S, w,
                         c->Q.fx, c->Q.fy, c->Q.fz, c->Q.ee);
  }
}

template <bool PRED, int ORDER, int RS>
void stage1(bnz_batch *c) {
  const WList wl = wlist(c, KS1);
  if (wl.hi == wl.lo) return;
  const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TU]));
  hipLaunchKernelGGL((kb32_stage1<PRED, ORDER, RS>), grid, dim3(BLK), 0,
                     c->stream, c->boxes, wl, c->S, c->P, c->Q, c->p.dt_fms_ref, c->nguard);
}

// stage updates of the 2D and 3D boxes, then the fused stage of the 1D boxes
template <bool PRED>
void updates(bnz_batch *c) {
  const int fr = c->p.dt_fms_ref;
  for (int d = 2; d <= 3; ++d) {
    const WList wl = wlist_dims(c, KUPD, d);
    if (wl.hi == wl.lo) continue;
    const dim3 grid(grid_for(wl.hi - wl.lo, c->cap[TU]));
    if (d == 2)
      hipLaunchKernelGGL((kb32_update<PRED, 2>), grid, dim3(BLK), 0, c->stream, c->boxes, wl,
                         c->S, c->P, c->Q, fr);
    else
      hipLaunchKernelGGL((kb32_update<PRED, 3>), grid, dim3(BLK), 0, c->stream, c->boxes, wl,
                         c->S, c->P, c->Q, fr);
  }
  const int order = PRED ? 0 : c->order;
  if (order == 0) {
    if (c->p.rsolver == BNZ_RS_HLLD) stage1<PRED, 0, BNZ_RS_HLLD>(c);
    else stage1<PRED, 0, BNZ_RS_HLL>(c);
  } else {
    if (c->p.rsolver == BNZ_RS_HLLD) stage1<PRED, 1, BNZ_RS_HLLD>(c);
    else stage1<PRED, 1, BNZ_RS_HLL>(c);
  }
}

Ghost32 ghost_arrays(bnz_batch *c) {
  Ghost32 p;
  int q = 0;
  const long n = c->ncell;
  for (int v = 0; v < 6; ++v) p.a[q++] = c->Q.uh + v * n;
  for (int v = 0; v < 9; ++v) p.a[q++] = c->Q.w32 + v * n;
  for (int v = 0; v < 3; ++v) p.d[v] = c->P.b + v * n;
  return p;
}

}  // namespace

namespace bnzb {

// The FP32C stages of one VL step of every running box, between the dt
// and time kernels of the caller (the sequence of the single-grid
// f32_step_body).
template <int ORDER, int RS>
void step1_launch(bnz_batch *c) {
  const WList wl = wlist(c, KT1);
  hipLaunchKernelGGL((kb32_step1<ORDER, RS>), dim3(grid_for(wl.hi - wl.lo, c->cap[TU])),
                     dim3(BLK), 0, c->stream, c->boxes, wl, c->S, c->P, c->Q, c->p.dt_fms_ref,
                     c->nguard);
}

void b32_step(bnz_batch *c) {
  if (c->step1) {  // every box 1D: one fused pass per step
    const bool d = c->p.rsolver == BNZ_RS_HLLD;
    if (c->order == 0) {
      if (d) step1_launch<0, BNZ_RS_HLLD>(c);
      else step1_launch<0, BNZ_RS_HLL>(c);
    } else {
      if (d) step1_launch<1, BNZ_RS_HLLD>(c);
      else step1_launch<1, BNZ_RS_HLL>(c);
    }
    std::swap(c->Q.uh, c->Q.uh2);
    std::swap(c->Q.ul, c->Q.ul2);
    std::swap(c->Q.w32, c->Q.ws32);
  } else {
    fluxes(c, 0, c->Q.w32, (const double *)c->P.b, KF0);
    emfs(c, c->Q.w32, KEM0);
    updates<true>(c);
    fluxes(c, c->order, c->Q.ws32, (const float *)c->Q.bs32, KF1);
    emfs(c, c->Q.ws32, KEM1);
    updates<false>(c);
  }
  std::swap(c->P.b, c->P.bs);
  const Ghost32 p = ghost_arrays(c);
  // the end-of-step time and dt kernel rides in the last ghost launch
  const int tp = c->ngops == 0 ? 1 : 0;
  if (c->nghost > 0)
    hipLaunchKernelGGL(kb32_ghost, dim3(grid_for((c->nghost + BLK - 1) / BLK) + tp), dim3(BLK),
                       0, c->stream, p, c->ghost, c->nghost, c->S, c->nbox, c->p.cour, tp);
  if (c->ngops > 0)
    hipLaunchKernelGGL(kb32_ghost_ops, dim3(grid_for((c->ngops + BLK - 1) / BLK) + 1), dim3(BLK),
                       0, c->stream, p, c->gops, c->ngops, c->S, c->nbox, c->p.cour, 1);
  else if (c->nghost == 0)
    b_time_dt(c);
}

void b32_init(bnz_batch *c) {
  hipLaunchKernelGGL(kb32_from64, dim3(grid_for(c->nblk[KUPD])), dim3(BLK), 0, c->stream,
                     c->boxes, wlist(c, KUPD), c->P, c->Q);
  hipLaunchKernelGGL(kb32_cfl, dim3(grid_for(c->nblk[KACT])), dim3(BLK), 0, c->stream, c->boxes,
                     wlist(c, KACT), c->S, c->Q.w32, c->p.dt_fms_ref);
}

void b32_fix(bnz_batch *c) {
  hipLaunchKernelGGL(kb32_fix, dim3(grid_for(c->nblk[KUPD])), dim3(BLK), 0, c->stream, c->boxes,
                     wlist(c, KUPD), c->S, c->Q);
}

void b32_to64(bnz_batch *c) {
  hipLaunchKernelGGL(kb32_to64, dim3(grid_for(c->nblk[KACT])), dim3(BLK), 0, c->stream,
                     c->boxes, wlist(c, KACT), c->P, c->Q);
}

}  // namespace bnzb
