// bnz_common.hpp -- types and helpers shared by every translation unit of
// libbnzgpu.so: single grids in FP64 (bnz_gpu.hip; with FMA contraction
// bnz_gpu64f.hip) and FP32C (bnz_gpu32.hip), batched boxes (bnz_batch.hip,
// bnz_batch32.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bnzgpu.h"
#include "bnz_phys.hpp"

namespace bnz {

constexpr int NVAR_STATE = 9;
constexpr int BLK = 256;
// cells a workgroup owns in a fused 1D step (step1_block): the predictor
// also computes 2 cells below and 3 above them
constexpr int STEP1_CELLS = BLK - 5;

// --------------------------------------------------------------------------
// Geometry passed by value to every kernel.
// Kernel-side solver id of the hydro build's HLL (sound speed in the wave
// estimate); the public ids are BNZ_RS_*.  CFL_HYDRO: the fms_ref value of
// the hydro build's dt.
constexpr int RS_HLL_HYDRO = 5;
constexpr int CFL_HYDRO = 2;

struct Geo {
  int nt[3];       // Ntot per axis
  int dims;
  long sv;         // stride between variables (cells per array)
  int sy, sz;      // strides of j and k
  const double *dlf[3];   // cell widths (coord.pyx:254-257)
  const double *dlfi[3];  // 1/width
  const double *icm[3];   // lv - lf       (interp_bc weights, ct.pyx:110-123)
  const double *icp[3];   // lf[i+1] - lv
  const double *dlv[3];   // lv[i] - lv[i-1] (0 at i = 0), the upwind EMF's ds
  double gam, gamm1i;
  // unif = 1 when each of the four arrays above is constant along every
  // axis (bit for bit, e.g. 2^-8 wide cells); the constants are then here
  int unif;
  double udlf[3], udlfi[3], uicm[3], uicp[3];
};

// Geometry lookups of a kernel: GeoArr reads the per-cell arrays, GeoUni the
// constants of a uniform grid (the same values, with no per-cell loads).
struct GeoArr {
  const Geo &g;
  __device__ __forceinline__ double dlf(int a, int q) const { return g.dlf[a][q]; }
  __device__ __forceinline__ double dlfi(int a, int q) const { return g.dlfi[a][q]; }
  __device__ __forceinline__ double icm(int a, int q) const { return g.icm[a][q]; }
  __device__ __forceinline__ double icp(int a, int q) const { return g.icp[a][q]; }
};
struct GeoUni {
  const Geo &g;
  __device__ __forceinline__ double dlf(int a, int) const { return g.udlf[a]; }
  __device__ __forceinline__ double dlfi(int a, int) const { return g.udlfi[a]; }
  __device__ __forceinline__ double icm(int a, int) const { return g.uicm[a]; }
  __device__ __forceinline__ double icp(int a, int) const { return g.uicp[a]; }
};

struct GhostOp {
  int dst;       // cell written
  int src;       // source of the cell variables (with bit 16: fixed-value row)
  int srcf[3];   // source of each face component
  unsigned bits; // bit v < 9: negate slot v; 9 + a: negate face a;
                 // 12 + a: face a = 0; 15: cell variables untouched;
                 // 16: every value from row src of the fixed-value table
};
constexpr unsigned GHOST_FIXED = 1u << 16;

struct Lims {
  int lo[3], hi[3];  // inclusive cell limits of get_integr_lims (integrator.pyx:375-383)
};

struct Box {
  int lo[3];
  int n[3];
  __host__ __device__ long total() const { return (long)n[0] * n[1] * n[2]; }
};

// Workgroups are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Renumber them so that each XCD owns one contiguous run of logical blocks:
// neighbouring rows and planes of a stencil then share that XCD's L2.
constexpr int NXCD = 8;
__device__ __forceinline__ long xcd_block() {
  const long b = blockIdx.x, nb = gridDim.x;
  const long q = nb / NXCD, r = nb % NXCD;
  const long x = b % NXCD, idx = b / NXCD;
  return x * q + (x < r ? x : r) + idx;
}

// Workgroups in runs of XCD_RUN consecutive logical blocks, the runs dealt
// round-robin over the XCDs (logical run r on XCD r % NXCD): all XCDs work
// on the same few planes at once and share what those planes bring into the
// MALL, while each XCD still walks whole rows.  Blocks past the last whole
// round keep their launch order.  Runs of 4 to 32 measured alike; 1 (launch
// order) and 64 or more were slower.
constexpr long XCD_RUN = 16;
__device__ __forceinline__ long run_block() {
  const long b = blockIdx.x, nb = gridDim.x;
  const long nfull = nb / (NXCD * XCD_RUN) * (NXCD * XCD_RUN);
  if (b >= nfull) return b;
  const long x = b % NXCD, idx = b / NXCD;
  return ((idx / XCD_RUN) * NXCD + x) * XCD_RUN + idx % XCD_RUN;
}

__device__ __forceinline__ void box_idx(const Box &b, long t, int &i, int &j, int &k) {
  int nx = b.n[0], ny = b.n[1];
  i = b.lo[0] + (int)(t % nx);
  long r = t / nx;
  j = b.lo[1] + (int)(r % ny);
  k = b.lo[2] + (int)(r / ny);
}

// Thread -> cell of box b.
//   flat (ty <= 0): x fastest over the whole box; ty == 0 in xcd_block
//     order, ty < 0 in run_block order (see flat_order);
//   TILED (TA, MA): a workgroup is 64 x-lanes by 4 rows of axis TA at one
//     index of axis MA, and consecutive workgroups step along MA, so stencil
//     neighbours along MA are read by workgroups that run back to back on
//     the same XCD;
//   slab (ty > 1): the box is cut into columns of ty rows (y); a column is
//     walked plane by plane (z), x fastest within a plane, so the rows of
//     plane k+1 that plane k reads as neighbours are still in the XCD's L2
//     when plane k+1 is processed.
template <bool TILED, int TA, int MA>
__device__ __forceinline__ bool cell_index(const Box &b, int ty, int &i, int &j, int &k) {
  if (TILED) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nbi = (b.n[0] + 63) / 64, nm = b.n[MA];
    long L = xcd_block();
    const int m = (int)(L % nm);
    L /= nm;
    const int bi = (int)(L % nbi), bt = (int)(L / nbi);
    const int ii = bi * 64 + lane, tt = bt * 4 + wv;
    if (ii >= b.n[0] || tt >= b.n[TA]) return false;
    i = b.lo[0] + ii;
    const int vt = b.lo[TA] + tt, vm = b.lo[MA] + m;
    j = (TA == 1) ? vt : vm;
    k = (TA == 2) ? vt : vm;
    return true;
  }
  if (ty <= 0) {
    const long t = (ty < 0 ? run_block() : xcd_block()) * BLK + threadIdx.x;
    if (t >= b.total()) return false;
    box_idx(b, t, i, j, k);
    return true;
  }
  const int per = ty * b.n[0];                // cells of one column plane
  const int bpp = (per + BLK - 1) / BLK;      // workgroups per column plane
  const long L = xcd_block();
  const long bcol = (long)bpp * b.n[2];       // workgroups per column
  const int col = (int)(L / bcol);
  const long rem = L - col * bcol;
  const int kk = (int)(rem / bpp);
  const int t = (int)(rem - (long)kk * bpp) * BLK + threadIdx.x;
  if (t >= per) return false;
  const int jj = col * ty + t / b.n[0];
  if (jj >= b.n[1]) return false;
  i = b.lo[0] + t % b.n[0];
  j = b.lo[1] + jj;
  k = b.lo[2] + kk;
  return true;
}

// Global variable read by local slot n of a sweep along DIR
// (swap_mhd_vec, src/integration/godunov.pyx:297-324).
template <int DIR>
__device__ __forceinline__ constexpr int perm(int n) {
  if (DIR == 0) return n;
  if (DIR == 1) {
    switch (n) {
      case VX: return VY;
      case VY: return VZ;
      case VZ: return VX;
      case BX: return BY;
      case BY: return BZ;
      case BZ: return BX;
      default: return n;
    }
  }
  switch (n) {
    case VX: return VZ;
    case VY: return VX;
    case VZ: return VY;
    case BX: return BZ;
    case BY: return BX;
    case BZ: return BY;
    default: return n;
  }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Maximum over the workgroup into part[blockIdx.x] (wave shuffles, then LDS).
__device__ __forceinline__ void block_max_to(double v, double *part) {
  __shared__ double red[BLK / 64];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
#pragma unroll
    for (int q = 1; q < BLK / 64; ++q) m = fmax(m, red[q]);
    part[blockIdx.x] = m;
  }
}

// The dt of a step from its CFL partials (new_dt.pyx:18-135 after the
// per-cell maxima): dt = cour / max over npart partials, or a fixed dt,
// clipped to t_max.  dev[0]=dt, dev[1]=t, dev[2]=t_max (<0: no clip),
// dev[3]=step counter, dev[7]=the dt of the last step taken.  k_dt_final
// runs it as a kernel of its own; a step's first flux launch runs it in its
// workgroup 0, where nothing reads dt before the predictor's update (DtFold).
struct DtFold {
  const double *part = nullptr;
  int npart = 0;  // 0: dev[0] already holds the step's dt
  int use_fixed = 0;
  double cour = 0., fixed_dt = 0.;
};

// Called by every thread of one workgroup.
__device__ __forceinline__ void dt_reduce_block(const DtFold &f, double *dev) {
  __shared__ double red[BLK / 64];
  double m[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
  int q = threadIdx.x;
  for (; q + 7 * BLK < f.npart; q += 8 * BLK) {
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = fmax(m[u], f.part[q + u * BLK]);
  }
  for (; q < f.npart; q += BLK) m[0] = fmax(m[0], f.part[q]);
  double mm = m[0];
#pragma unroll
  for (int u = 1; u < 8; ++u) mm = fmax(mm, m[u]);
  mm = wave_max(mm);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mm;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = red[0];
    for (int r = 1; r < BLK / 64; ++r) x = fmax(x, red[r]);
    double dt = f.use_fixed ? f.fixed_dt : f.cour / x;
    const double tnow = dev[1], tmax = dev[2];
    if (!f.use_fixed && tmax >= 0.) {
      if (!(tnow < tmax))
        dt = 0.;  // already at t_max: the step is a no-op (see k_update)
      else if (tnow + dt > tmax)
        dt = tmax - tnow;
    }
    dev[0] = dt;
    if (dt != 0.) dev[7] = dt;  // the dt of the last step taken
  }
}

// Whether a step is one past t_max, for the workgroups of a launch whose
// workgroup 0 may still be computing dt: with a fold pending, from the
// fixed dt or from t and t_max (dt = 0 exactly then, apart from an infinite
// CFL speed, after which the update still sees dt = 0); else from dev[0].
__device__ __forceinline__ bool step_past_tmax(const DtFold &f, const double *dev) {
  if (f.npart == 0) return dev[0] == 0.;
  if (f.use_fixed) return f.fixed_dt == 0.;
  return dev[2] >= 0. && !(dev[1] < dev[2]);
}

// The time advance of a finished step, run by one thread of the step's
// last launch (the ghost fill): t += dt, step += 1, unless the step was one
// past t_max.
__device__ __forceinline__ void time_advance_once(double *adv) {
  if (adv && blockIdx.x == 0 && threadIdx.x == 0 && adv[0] != 0.) {
    adv[1] += adv[0];
    adv[3] += 1.0;
  }
}

// CFL speed of one active cell (new_dt.pyx:74-91).  fms_ref is 1 (the
// reference's x fast speed on every axis, D2), 0 (per-axis fast speeds) or
// CFL_HYDRO (the hydro build: the sound speed on every axis).
template <class GM>
__device__ __forceinline__ double cds_cell_g(const GM &gm, const Geo &g, const double *wv,
                                             int fms_ref, int i, int j, int k) {
  if (fms_ref == CFL_HYDRO) {
    const double a = fms_hydro(wv, g.gam);
    const double cx = fabs(wv[VX]) + a;
    const double cy = g.dims >= 2 ? fabs(wv[VY]) + a : 0.;
    const double cz = g.dims >= 3 ? fabs(wv[VZ]) + a : 0.;
    double cds = fmax(fmax(cx / gm.dlf(0, i), cy / gm.dlf(1, j)), cz / gm.dlf(2, k));
    return (cds > 0.) ? cds : 0.;
  }
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
  double cds = fmax(fmax(cx / gm.dlf(0, i), cy / gm.dlf(1, j)), cz / gm.dlf(2, k));
  return (cds > 0.) ? cds : 0.;  // NaN never wins the reference's `>` test
}
__device__ __forceinline__ double cds_cell(const Geo &g, const double *wv, int fms_ref, int i,
                                           int j, int k) {
  return cds_cell_g(GeoArr{g}, g, wv, fms_ref, i, j, k);
}

}  // namespace bnz

// ==========================================================================
// Host side
// ==========================================================================

struct KTime {
  std::string name;
  double total_ms = 0.;
  long launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct bnz_ctx {
  bnz_params p{};
  int dims = 1, ng = 3, order = 1;
  int nt[3] = {1, 1, 1};
  int i1, i2, j1, j2, k1, k2;
  long ncell = 0;
  hipStream_t stream = nullptr;
  double *w = nullptr, *ws = nullptr, *b = nullptr, *bs = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr, *ee = nullptr;
  // RK3 (tintegr = 1): U, U** (conserved) and B** (faces); U* and B* use ws
  // and bs
  double *uc = nullptr, *uss = nullptr, *bss = nullptr;
  // bnz_advance: a captured chunk of steps, replayed while the buffers,
  // CFL partial count and work shapes it was captured with are current
  struct ChunkKey {
    const void *w, *ws, *b, *bs, *part;
    int dt_parts, flux_chunk, emf_chunk, update_chunk;
    bool operator==(const ChunkKey &o) const {
      return w == o.w && ws == o.ws && b == o.b && bs == o.bs && part == o.part &&
             dt_parts == o.dt_parts && flux_chunk == o.flux_chunk &&
             emf_chunk == o.emf_chunk && update_chunk == o.update_chunk;
    }
  };
  bool step_graph = false;  // bnz_set_step_graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  ChunkKey gkey{};
  // a key that a full eager chunk started and ended on: capturing from it
  // gives a graph that leaves the context as it found it
  ChunkKey steady{};
  bool has_steady = false;
  double *geo = nullptr;
  int2 *ghost = nullptr;      // (ghost cell, source cell) pairs, see build_ghosts
  int nghost = 0;
  bnz::GhostOp *gops = nullptr;  // ghost cells touched by a reflecting side
  int ngops = 0;
  // User boundaries (flag 3, bnz_set_user_bc): with one on any side the fill
  // runs as one gather pass per side in the order z1, z2, y1, y2, x1, x2
  // (gops[pass_off[p] .. pass_off[p + 1])), each pass reading what the
  // earlier ones wrote.  Per side: the caller's slab entries (source cell,
  // -1 fixed values, -2 untouched) and NMODE + 3 fixed values per entry.
  std::vector<int> usr_src[6];
  std::vector<double> usr_fix[6];
  bool usr_set[6] = {false, false, false, false, false, false};
  int npass = 0;
  int pass_off[7] = {0, 0, 0, 0, 0, 0, 0};
  double *ufix = nullptr;  // fixed values in use: NMODE prim + 3 faces each
  float *ufix32 = nullptr; // their FP32C state: 6 conserved hi words + NMODE prim
  double *dev = nullptr;      // dt, t, tmax, step
  double *part = nullptr;     // CFL partials (the last corrector's, kept for the next dt)
  long npart_cap = 0;
  double *rpart = nullptr;    // partials of the diag and history reductions
  std::vector<double> h_dlf[3], h_dlfi[3], h_icm[3], h_icp[3], h_dlv[3];
  bnz::Geo g{};
  // FP32-compensated state (BNZ_PREC_FP32C): conserved hydro variables as
  // FP32 hi + lo pairs, face fields in b/bs (FP64), FP32 primitives of the
  // current and predicted state, FP32 predicted faces, fluxes and EMFs; w
  // above is the FP64 prim mirror used for upload, init and download.
  bool f32 = false;
  float *uh = nullptr, *ul = nullptr;     // [6][cells]
  float *w32 = nullptr, *ws32 = nullptr;  // [9][cells]
  float *bs32 = nullptr;                  // [3][cells] predicted faces
  float *fx32 = nullptr, *fy32 = nullptr, *fz32 = nullptr;  // [9][cells]
  float *ee32 = nullptr;                  // [3][cells]
  bool ready = false;
  bool timing = false;
  // tuning: work shapes of the flux, EMF and update kernels (3D), see
  // work_shape()
  int flux_chunk = 0, emf_chunk = 0, update_chunk = 0;
  // CFL partial maxima left by the last corrector for the next step
  int dt_parts = 0;
  // the next step's dt reduction, left for its first flux launch (take_dt)
  bnz::DtFold pdt{};
  bool pdt_pending = false;
  // bnz_advance: pinned copies of (t, tmax, step) per chunk of steps, and the
  // events that mark them
  double *hbuf = nullptr;
  hipEvent_t hev[2] = {nullptr, nullptr};
  std::vector<KTime> ktimes;
  std::string err;
};


namespace bnz {

inline int fail(bnz_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, BNZ_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

inline int blocks_for(long n) { return (int)((n + BLK - 1) / BLK); }

// Kernel-side solver id and CFL mode of a context: the hydro build's HLL
// takes the sound speed in its estimate and the hydro dt the sound speed.
inline int rs_kern(const bnz_ctx *c) {
  return (c->p.hydro && c->p.rsolver == BNZ_RS_HLL) ? RS_HLL_HYDRO : c->p.rsolver;
}
inline int cfl_mode(const bnz_ctx *c) { return c->p.hydro ? CFL_HYDRO : c->p.dt_fms_ref; }

// Workgroups of a tiled per-cell kernel over box b (see cell_index).
inline int tiled_blocks(const Box &b, int ta, int ma) {
  return ((b.n[0] + 63) / 64) * ((b.n[ta] + 3) / 4) * b.n[ma];
}

// Workgroups of a slab-column launch over box b with columns of ty rows.
inline int slab_blocks(const Box &b, int ty) {
  return ((b.n[1] + ty - 1) / ty) * b.n[2] * ((ty * b.n[0] + BLK - 1) / BLK);
}

// Work shape of a kernel class from its tuning value (3D only): 0 flat,
// 1 tiles of 64 x 4 along z, > 1 slab columns of that many rows.  ty is
// what cell_index takes: 0 flat, -1 flat in run_block order (flat_order),
// > 1 the rows of a slab column.
struct Shape {
  bool tiled;
  int ty;
  int blocks;
};
inline Shape work_shape(const Box &b, int dims, int chunk) {
  if (dims < 3 || chunk <= 0) return Shape{false, 0, blocks_for(b.total())};
  if (chunk == 1) return Shape{true, 0, tiled_blocks(b, 1, 2)};
  return Shape{false, chunk, slab_blocks(b, chunk)};
}

// FP32C kernels take flat or slab-column shapes; the tile shape runs flat.
inline Shape work_shape32(const Box &b, int dims, int chunk) {
  return work_shape(b, dims, chunk == 1 ? 0 : chunk);
}

// The flat shape of the flux, EMF and update launches of large 3D grids
// takes its workgroups in runs dealt round-robin over the XCDs (ty -1,
// run_block): there all XCDs then share the planes the MALL holds, and the
// launches run faster than with one contiguous run per XCD (256^3: step
// 7.42 -> 6.97 ms; 128^3 -4 %), while 2D and small 3D grids run slower so
// (single-block runs: OT 128^2 +6 %, turbulence 32^3 +5 %).  The EMF, which
// re-reads the previous plane, needs the runs: dealt one block at a time it
// is 35 % slower than with one run per XCD, in runs of 16 6 % faster.
constexpr long LAUNCH_ORDER_CELLS = 1L << 21;
inline Shape flat_order(const bnz_ctx *c, Shape sh) {
  if (c->dims == 3 && !sh.tiled && sh.ty == 0 && c->ncell >= LAUNCH_ORDER_CELLS) sh.ty = -1;
  return sh;
}

struct Launch {
  bnz_ctx *c;
  int idx;
  hipEvent_t e0{}, e1{};
  Launch(bnz_ctx *ctx, const char *name) : c(ctx), idx(-1) {
    if (!c->timing) return;
    for (size_t q = 0; q < c->ktimes.size(); ++q)
      if (c->ktimes[q].name == name) idx = (int)q;
    if (idx < 0) {
      c->ktimes.push_back(KTime{});
      c->ktimes.back().name = name;
      idx = (int)c->ktimes.size() - 1;
    }
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, c->stream);
  }
  ~Launch() {
    if (idx < 0) return;
    (void)hipEventRecord(e1, c->stream);
    c->ktimes[idx].pending.emplace_back(e0, e1);
  }
};

inline void collect_timing(bnz_ctx *c) {
  for (auto &kt : c->ktimes) {
    for (auto &pr : kt.pending) {
      (void)hipEventSynchronize(pr.second);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, pr.first, pr.second);
      kt.total_ms += ms;
      kt.launches += 1;
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
    kt.pending.clear();
  }
}

inline Lims integr_lims(const bnz_ctx *c, int order) {
  Lims L;
  for (int a = 0; a < 3; ++a) {
    L.lo[a] = order + 1;
    L.hi[a] = c->nt[a] - 2 - order;
  }
  if (c->dims < 2) L.lo[1] = L.hi[1] = 0;
  if (c->dims < 3) L.lo[2] = L.hi[2] = 0;
  return L;
}

// Union of the x, y and z face ranges of a stage with limits L: x faces
// i in [il, iu+1], j in [jl-1, ju+1], k in [kl-1, ku+1]; y faces i in
// [il-1, iu+1], j in [jl, ju+1]; z faces k in [kl, ku+1] (godunov.pyx:66-206).
inline Box flux3_box(int dims, const Lims &L) {
  Box b;
  b.lo[0] = dims >= 2 ? L.lo[0] - 1 : L.lo[0];
  b.n[0] = L.hi[0] + 1 - b.lo[0] + 1;
  b.lo[1] = dims >= 2 ? L.lo[1] - 1 : 0;
  b.n[1] = dims >= 2 ? L.hi[1] + 1 - b.lo[1] + 1 : 1;
  b.lo[2] = dims >= 3 ? L.lo[2] - 1 : 0;
  b.n[2] = dims >= 3 ? L.hi[2] + 1 - b.lo[2] + 1 : 1;
  return b;
}

// Small 2D grids form the edge EMFs in the update launches (EmfFly2D,
// Emf32Fly2D), so their steps run no EMF launch: a launch there costs more
// than the four cells' redundant EMF arithmetic.  On 2D grids emf_chunk (no
// work shape there) picks it: 0 by size, 1 EMF launches, 2 or more in the
// updates.
constexpr long EMF_FLY_2D_CELLS = 1L << 18;
inline bool emf_fly(const bnz_ctx *c) {
  if (c->dims != 2 || c->p.ct_upwind) return false;
  if (c->emf_chunk == 0) return c->ncell <= EMF_FLY_2D_CELLS;
  return c->emf_chunk >= 2;
}

// Small 2D and 3D grids in the flat work order give each flux workgroup one
// face direction (k_flux3 SPLIT): the launch there is bound by one thread's
// chain of solves, not by the work.
constexpr long FLUX_SPLIT_CELLS = 1L << 18;
inline bool flux_split(const bnz_ctx *c, const Shape &sh) {
  return c->dims >= 2 && !sh.tiled && sh.ty <= 0 && c->ncell <= FLUX_SPLIT_CELLS;
}

inline Box active_box(const bnz_ctx *c) {
  Box b;
  b.lo[0] = c->i1; b.n[0] = c->i2 - c->i1 + 1;
  b.lo[1] = c->j1; b.n[1] = c->j2 - c->j1 + 1;
  b.lo[2] = c->k1; b.n[2] = c->k2 - c->k1 + 1;
  return b;
}

inline int check_launch(bnz_ctx *c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, BNZ_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return BNZ_OK;
}


// --------------------------------------------------------------------------
// Ghost fill of one grid (GridBC.apply_grid_bc, src/bc/grid_bc.pyx:323-339)
// as two host-built lists.  Each axis has a map from a ghost index to the
// active index it is filled from (periodic copy_layer_*, outflow prolong_*,
// reflect copy_reflect_layer_*: util_bc.pyx:9-90); the reference fills z,
// then y, then x over the full extent of the other axes, which composes the
// per-axis maps, so a ghost cell takes its value straight from one active
// cell.
//   pairs: ghost cells untouched by a reflecting side, (cell, source), all
//          variables copied as they are;
//   ops:   cells a reflecting side touches: one source for the cell
//          variables and one per face component (the face-normal component
//          is mirrored about the wall face, not about the wall cell), sign
//          flips (normal velocity or momentum and normal cell field flip,
//          the face-normal field flips across its own wall) and the zeroed
//          wall face (set_layer_*, util_bc.pyx:172-205).
// Sources are active cells that the fill never writes, so every entry runs
// independently.  Cell indices are offset by off (a box in a pool).


inline void build_ghosts(const int nt[3], int dims, int ng, const int bc[6], long off,
                         std::vector<int2> &pairs, std::vector<GhostOp> &ops) {
  std::vector<int> mc[3], mf[3];
  std::vector<char> fc[3], ff[3], zf[3], refl[3];
  for (int a = 0; a < 3; ++a) {
    const int N = nt[a];
    mc[a].resize(N);
    mf[a].resize(N);
    fc[a].assign(N, 0);
    ff[a].assign(N, 0);
    zf[a].assign(N, 0);
    refl[a].assign(N, 0);
    for (int i = 0; i < N; ++i) mc[a][i] = mf[a][i] = i;
    if (a >= dims) continue;
    const int a1 = ng, a2 = N - ng - 1;
    for (int side = 0; side < 2; ++side) {
      const int flag = bc[2 * a + side];
      if (flag < 0) continue;  // side left alone (one-side passes)
      for (int g = 0; g < ng; ++g) {
        const int idx = side == 0 ? a1 - ng + g : a2 + 1 + g;
        if (flag == BNZ_BC_PERIODIC) {
          mc[a][idx] = mf[a][idx] = side == 0 ? a2 - ng + 1 + g : a1 + g;
        } else if (flag == BNZ_BC_OUTFLOW) {  // D14: z1 copies k1
          mc[a][idx] = mf[a][idx] = side == 0 ? a1 : a2;
        } else {  // reflect: mirror about the wall
          mc[a][idx] = side == 0 ? 2 * a1 - 1 - idx : 2 * a2 + 1 - idx;
          fc[a][idx] = 1;
          refl[a][idx] = 1;
        }
      }
      if (flag == BNZ_BC_REFLECT) {  // face-normal component: wall face zeroed
        const int wall = side == 0 ? a1 : a2 + 1;
        zf[a][wall] = 1;
        refl[a][wall] = 1;
        mf[a][wall] = wall;
        if (side == 0) {
          for (int idx = 0; idx < a1; ++idx) {
            mf[a][idx] = 2 * a1 - idx;
            ff[a][idx] = 1;
          }
        } else {
          for (int idx = a2 + 2; idx < N; ++idx) {
            mf[a][idx] = 2 * a2 + 2 - idx;
            ff[a][idx] = 1;
          }
        }
      }
    }
  }
  auto cell = [&](int i, int j, int k) { return off + ((long)k * nt[1] + j) * nt[0] + i; };
  for (int k = 0; k < nt[2]; ++k)
    for (int j = 0; j < nt[1]; ++j)
      for (int i = 0; i < nt[0]; ++i) {
        const int id[3] = {i, j, k};
        const int si = mc[0][i], sj = mc[1][j], sk = mc[2][k];
        if (!refl[0][i] && !refl[1][j] && !refl[2][k]) {
          if (si == i && sj == j && sk == k) continue;
          pairs.push_back(make_int2((int)cell(i, j, k), (int)cell(si, sj, sk)));
          continue;
        }
        GhostOp o;
        o.dst = (int)cell(i, j, k);
        o.src = (int)cell(si, sj, sk);
        o.bits = 0;
        for (int a = 0; a < 3; ++a)
          if (fc[a][id[a]]) o.bits ^= (1u << (VX + a)) | (1u << (BX + a));
        for (int f = 0; f < 3; ++f) {
          int s3[3] = {si, sj, sk};
          s3[f] = mf[f][id[f]];
          o.srcf[f] = (int)cell(s3[0], s3[1], s3[2]);
          if (ff[f][id[f]]) o.bits |= 1u << (9 + f);
          if (zf[f][id[f]]) o.bits |= 1u << (12 + f);
        }
        if (o.src == o.dst && (o.bits & 0x1ffu) == 0) o.bits |= 1u << 15;
        ops.push_back(o);
      }
}

// FP32-compensated step (bnz_gpu32.hip)
int f32_alloc(bnz_ctx *c);
void f32_free(bnz_ctx *c);
int f32_from64(bnz_ctx *c);   // state from the FP64 prim/face arrays (all cells)
int f32_to64(bnz_ctx *c);     // FP64 prim/face arrays from the state
void f32_ghost(bnz_ctx *c);
// FP32C state of rows user fixed values (ufix) into a new buffer *out
int f32_fixed_rows(bnz_ctx *c, const double *ufix, int rows, float **out);
void f32_dt(bnz_ctx *c, int use_fixed, double fixed_dt);
void f32_step_body(bnz_ctx *c);
// shared launches (bnz_gpu.hip)
void defer_dt_final(bnz_ctx *c, int npart, int use_fixed, double fixed_dt);
void launch_time_advance(bnz_ctx *c);
// The pending dt reduction for a step's first flux launch (npart 0 if none),
// or launched on its own (flush_dt) where dt is needed first.
inline DtFold take_dt(bnz_ctx *c) {
  DtFold f{};
  if (c->pdt_pending) f = c->pdt;
  c->pdt_pending = false;
  return f;
}
void flush_dt(bnz_ctx *c);
void launch_ghost64(bnz_ctx *c);
void launch_ghost_step64(bnz_ctx *c);  // a step's fill: also t += dt, step += 1
// the FP64 step compiled with contraction (bnz_gpu64f.hip)
void fma_step_body(bnz_ctx *c);

}  // namespace bnz
