// bnz_batch.hpp -- types shared by the FP64 (bnz_batch.hip) and
// FP32-compensated (bnz_batch32.hip) translation units of the batched-box
// step: box descriptors, work lists, pools and the batch object.
#pragma once
#include <string>
#include <vector>

#include "bnz_common.hpp"

namespace bnzb {
using namespace bnz;

// work lists: fluxes (all directions, one item per cell of the union of the
// face ranges) and EMFs of the predictor and of the corrector, all cells
// (stage updates), active cells (init and the first CFL maximum), the cells
// of the 1D boxes for the fused 1D stage (BLK - 1 per work block) and for
// the fused 1D step of all-1D FP64 batches (BLK - 5 per work block)
enum { KF0 = 0, KEM0, KF1, KEM1, KUPD, KACT, KS1, KT1, NK };

// cells per work block of list kt
__host__ __device__ constexpr int blk_items(int kt) {
  return kt == KS1 ? BLK - 1 : kt == KT1 ? STEP1_CELLS : BLK;
}

struct BoxDev {
  Geo g;         // strides and geometry of the box; g.sv is the pool stride
  long off;      // first cell of the box in the pools
  Lims L[2];     // predictor / corrector limits (integrator.pyx:375-383)
  Box act;       // active cells
  Box items[NK];
  int blk0[NK];  // first work block of the box in each list
  int walls[6];  // reflecting wall plane of each side, or -1 (FP32C EMFs)
};

struct Pools {
  double *w, *ws, *b, *bs, *fx, *fy, *fz, *ee;
};

// FP32-compensated pools (face fields stay in Pools::b / bs, FP64)
struct Pools32 {
  float *uh, *ul, *w32, *ws32, *bs32, *fx, *fy, *fz, *ee;
  float *uh2, *ul2;  // the other hi / lo words of fused 1D steps (all-1D batches only)
};

struct BoxState {
  double *dt, *t, *tmax;
  unsigned long long *cmax;  // CFL maximum, as the bits of a non-negative double
  int *run;                  // box advances in the current step
  int *steps;
  int *nrun;                 // boxes still short of t_max after the last step
  int *fpar;                 // steps a finished box sat out, mod 2: its faces are
                             // in the other face buffer when 1
};

// Maximum of v over the workgroup folded into *addr with one atomic (the
// bits of a non-negative double order like its value).  Every thread of the
// workgroup calls it (a work block belongs to one box); the closing barrier
// lets the next work block of a persistent loop reuse red.
__device__ __forceinline__ void block_max_atomic(double v, unsigned long long *addr) {
  __shared__ double red[BLK / 64];
  v = bnz::wave_max(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
#pragma unroll
    for (int q = 1; q < BLK / 64; ++q) m = fmax(m, red[q]);
    if (m > 0.) atomicMax(addr, (unsigned long long)__double_as_longlong(m));
  }
  __syncthreads();
}

// The end of one step and the dt of the next, by one workgroup of BLK
// threads: t += dt of the running boxes and the count of boxes still short
// of t_max, then the dt and run flag of every box from the CFL maximum the
// corrector just folded (what kb_dt at the start of the next step would
// compute from the same maximum and time).
__device__ __forceinline__ void time_dt_block(const BoxState &st, int nbox, double cour) {
  __shared__ int cnt[BLK / 64];
  int n = 0;
  for (int b = threadIdx.x; b < nbox; b += BLK) {
    double t = st.t[b];
    if (st.run[b]) {
      t += st.dt[b];
      st.t[b] = t;
      st.steps[b] += 1;
    } else {
      st.fpar[b] ^= 1;  // the face buffers swapped without this box
    }
    const double tmax = st.tmax[b];
    n += (t < tmax) ? 1 : 0;
    const double m = __longlong_as_double((long long)st.cmax[b]);
    st.cmax[b] = 0ull;
    int run = 0;
    if (t < tmax) {
      double dt = cour / m;
      if (t + dt > tmax) dt = tmax - t;
      st.dt[b] = dt;
      run = 1;
    }
    st.run[b] = run;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if ((threadIdx.x & 63) == 0) cnt[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
#pragma unroll
    for (int q = 0; q < BLK / 64; ++q) tot += cnt[q];
    *st.nrun = tot;
  }
}

// Work list kt as seen by a kernel: work blocks [lo, hi).  Each list holds
// the blocks of the 1D boxes, then those of the 2D and of the 3D boxes, so
// the hot kernels run one launch per dimensionality present, compiled for it.
struct WList {
  const int *bmap;  // work block -> box
  int lo, hi;
  int kt;
};

// Workgroups of a launch over nblk work blocks: at most cap (each walks the
// blocks cap apart), at least one.  PERSIST is the cap of the setup kernels
// and the default of the stage kernels, whose caps bnz_batch_set_tuning sets
// per kernel class.
constexpr int PERSIST = 16384;
enum { TF = 0, TE, TU, NTUNE };  // tuning classes: fluxes, EMFs, updates (+ 1D stages)

inline int grid_for(long nblk, long cap = PERSIST) {
  return (int)std::max<long>(1, std::min<long>(nblk, cap));
}

}  // namespace bnzb

struct bnz_batch {
  bnz_params p{};
  int nbox = 0, order = 1, ng = 3;
  std::vector<bnzb::BoxDev> hbox;
  int nblk[bnzb::NK] = {};
  int bm_off[bnzb::NK] = {};
  int cls[bnzb::NK][4] = {};  // first work block of the 1D, 2D, 3D boxes; nblk
  int cap[bnzb::NTUNE] = {bnzb::PERSIST, bnzb::PERSIST, bnzb::PERSIST};
  long ncell = 0, nghost = 0, ngops = 0;
  hipStream_t stream = nullptr;
  bnzb::BoxDev *boxes = nullptr;
  int *bmap = nullptr;
  double *geo = nullptr;
  int2 *ghost = nullptr;
  bnz::GhostOp *gops = nullptr;
  bnzb::Pools P{};
  bool f32 = false;
  bool step1 = false;  // all boxes 1D: fused steps, state buffers swap like the faces
  bnzb::Pools32 Q{};
  int *nguard = nullptr;  // guarded FP32 HLLD faces
  bool w_stale = false;   // FP32C: the FP64 prim pool lags the state
  bnzb::BoxState S{};
  double *sbuf = nullptr;      // backing store of BoxState
  hipGraph_t graph = nullptr;  // two captured steps
  hipGraphExec_t exec = nullptr;
  double *graph_b = nullptr;   // current face buffer when the graph was captured
  int *hrun = nullptr;         // pinned host copies of the running-box count (2 slots)
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ready = false;
  std::string err;
};

namespace bnzb {

inline WList wlist(const bnz_batch *c, int kt) {
  return WList{c->bmap + c->bm_off[kt], 0, c->nblk[kt], kt};
}

// the work blocks of list kt that belong to boxes of dimensionality dims
inline WList wlist_dims(const bnz_batch *c, int kt, int dims) {
  return WList{c->bmap + c->bm_off[kt], c->cls[kt][dims - 1], c->cls[kt][dims], kt};
}

// FP32-compensated step (bnz_batch32.hip)
void b32_step(bnz_batch *c);
// the end-of-step time and dt kernel alone (bnz_batch.hip)
void b_time_dt(bnz_batch *c);
void b32_init(bnz_batch *c);   // FP64 prim pool -> FP32C state; first CFL maxima
void b32_fix(bnz_batch *c);    // FP32C state of finished boxes into the current buffers
void b32_to64(bnz_batch *c);   // FP32C state -> FP64 prim pool (active cells)

}  // namespace bnzb
