/*
 * bnzgpu.h -- C ABI of the MI355X Van-Leer MHD step (libbnzgpu.so).
 *
 * Drop-in boundary for the path that sergekomarov/bonzo's
 * src/integration/vanleer.pyx:59-179 (advance loop) and the kernels below it
 * implement.  Plain C types only; host arrays are C-contiguous doubles in the
 * reference's layout [var][k][j][i] (x fastest) over the full grid including
 * ghosts (src/coordinates/grid.pyx:28-52):
 *   prim : 9 rows  rho, vx, vy, vz, p, passive scalar, Bx, By, Bz (cell)
 *   cons : 9 rows  rho, mx, my, mz, E, passive scalar, Bx, By, Bz (cell)
 *   bfld : 3 rows  face-normal Bx, By, Bz at the lower face of each cell
 * Every call returns BNZ_OK (0) or a negative BNZ_E* code; bnz_last_error()
 * gives the message.  A context owns its device buffers and one HIP stream;
 * calls are stream-ordered, a context is not thread-safe, and host-visible
 * results are synchronised only by download, new_dt, advance, get_time and
 * diag.
 */
#ifndef BNZGPU_H
#define BNZGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bnz_ctx bnz_ctx;

enum {
  BNZ_OK = 0,
  BNZ_EINVAL = -1,       /* bad argument or configuration                   */
  BNZ_ENOMEM = -2,       /* device allocation failed                        */
  BNZ_EHIP = -3,         /* HIP runtime error                               */
  BNZ_EUNSUPPORTED = -4, /* selection the device path does not implement    */
  BNZ_ESTATE = -5,       /* call out of order (e.g. step before upload)     */
  BNZ_ERANGE = -6        /* size beyond a fixed maximum of the library      */
};

/* Riemann solvers: numbering of RSolver, src/integration/integrator.pxd:59-65 */
enum { BNZ_RS_HLL = 0, BNZ_RS_HLLC = 1, BNZ_RS_HLLD = 2, BNZ_RS_HLLA = 3, BNZ_RS_HLLT = 4 };
/* Reconstruction: numbering of Reconstr, src/integration/integrator.pxd:67-71 */
enum { BNZ_RCN_CONST = 0, BNZ_RCN_LINEAR = 1, BNZ_RCN_WENO = 2, BNZ_RCN_PARAB = 3 };
/* Boundary flags: src/bc/grid_bc.pyx:45-50 (0 periodic, 1 outflow, 2 reflect, 3 user) */
enum { BNZ_BC_PERIODIC = 0, BNZ_BC_OUTFLOW = 1, BNZ_BC_REFLECT = 2, BNZ_BC_USER = 3 };
/* Arithmetic of the step: FP64 throughout (parity path: IEEE, no
   contraction, bit for bit the CPU restatement of the reference); FP32
   fluxes, reconstruction and EMFs with an FP32 hi + lo conserved state
   updated in FP64 (pressure recovery, CT increments and dt in FP64); or the
   FP64 step with a*b+c fused into FMAs (single grids only: within 1e-12
   relative L1 of the parity path, as the reference's own fast-math build
   is, setup.py:436-451). */
enum { BNZ_PREC_FP64 = 0, BNZ_PREC_FP32C = 1, BNZ_PREC_FP64_FMA = 2 };

typedef struct {
  int nx, ny, nz;       /* active cells (ny=nz=1 in 1D, nz=1 in 2D), [computation] Nx,Ny,Nz */
  int rsolver;          /* BNZ_RS_*,  [computation] rsolver                */
  int reconstr;         /* BNZ_RCN_*, [computation] reconstr              */
  int bc[6];            /* x1,x2,y1,y2,z1,z2: [physics] bc_x1 ... bc_z2;
                           BNZ_BC_USER sides need bnz_set_user_bc            */
  double lmin[3];       /* [computation] xmin,ymin,zmin                    */
  double lmax[3];       /* [computation] xmax,ymax,zmax                    */
  double gam;           /* [physics] gam                                   */
  double cour;          /* [computation] cour                              */
  int dt_fms_ref;       /* 1: dt with the x fast speed on every axis, as
                           src/integration/new_dt.pyx:76-91 does; 0: per axis */
  int precision;        /* BNZ_PREC_*                                      */
  int device;           /* HIP device ordinal                              */
  int char_proj;        /* [computation] charact_proj: 1 limits the PLM
                           slopes in characteristic variables of each cell's
                           state (src/coordinates/reconstr_c.c:64-80,
                           src/mhd/eos_c.h:100-280); FP64 single grids only */
  int hydro;            /* 1: the reference's hydro build (setup.py without
                           -mf, MFIELD=0): rsolver BNZ_RS_HLL or BNZ_RS_HLLC
                           (fluxes_c.c:230-327) with the sound speed in their
                           wave estimates and in the dt (eos_c.h:523-527).
                           The state keeps the field slots, which must be
                           zero (bnz_upload checks) and stay exactly zero.
                           FP64 single grids only; 0 is the MHD build */
  int ct_upwind;        /* corner EMF of the CT update: 0 the reference's
                           average interp_ee1 (src/integration/ct.pyx:153-215),
                           1 the upwind interp_ee2 (ct.pyx:229-358, which the
                           reference leaves commented out at
                           integrator.pyx:106-108), weights from the face mass
                           fluxes and the step's dt; FP64 single grids only */
  int tintegr;          /* [computation] tintegr: 0 vl (vanleer.pyx), 1 rk3,
                           the three-stage SSP Runge-Kutta of
                           src/integration/_rk3.pyx:20-175 with ng = 9 and no
                           ghost fill between its stages (grid.pyx:142-143);
                           FP64 single grids only */
} bnz_params;

typedef struct {
  double mass, mom[3], energy, scalar; /* sums over active cells times cell volume */
  double divb_max;                     /* max |div B| over active cells (code units) */
  double bmax;                         /* max |B_cell| over active cells             */
  double pmin, rhomin;                 /* failure detection: minima over active cells */
  long nonfinite;                      /* active cells with a non-finite prim value  */
  long hlld_guards;                    /* FP32 HLLD faces whose star denominators
                                          were guarded (cumulative; 0 in FP64)      */
  long npneg, nrhoneg;                 /* active cells with p <= 0, with rho <= 0     */
  long bad_cell;                       /* failure location: the lowest linear index
                                          (k*Ntot_y + j)*Ntot_x + i of an active cell
                                          that is non-finite or has p <= 0 or rho <= 0,
                                          -1 if there is none                         */
  int bad_ijk[3];                      /* its i, j, k over the full grid (ghosts
                                          included); -1 if none                      */
  int bad_kind;                        /* what is wrong there: bit 0 a non-finite
                                          prim value, bit 1 p <= 0, bit 2 rho <= 0  */
  int step;                            /* steps taken when the diagnostics ran       */
  double time;                         /* and the time                               */
} bnz_diag_out;

/* Replaces BnzGrid/BnzIntegr construction (src/coordinates/grid.pyx:59-82,
   src/integration/integrator.pyx:23-28, init_method :205-291). */
int bnz_create(const bnz_params *p, bnz_ctx **out);
int bnz_destroy(bnz_ctx *ctx);
const char *bnz_last_error(const bnz_ctx *ctx);
/* Ntot per axis and ng (src/coordinates/grid.pyx:129-161) */
int bnz_get_shape(const bnz_ctx *ctx, int *ntot3, int *ng);

/* Host <-> device copies of the grid arrays (GridData, grid.pyx:28-52).
   upload waits for the context's queued work; the state then needs
   bnz_init_state (or bnz_load_state) before the next step. */
int bnz_upload(bnz_ctx *ctx, const double *prim, const double *bfld);
int bnz_download(bnz_ctx *ctx, double *prim, double *bfld, double *cons);

/* init_sim after set_problem: ghost fill, interp_bc, ghost fill of cell B,
   prim2cons (src/integration/vanleer.pyx:38-48) */
int bnz_init_state(bnz_ctx *ctx);
/* GridBC.apply_grid_bc over all NVAR (src/bc/grid_bc.pyx:323-339).  An
   FP32C context fills the hi words of the conserved state, the FP32 prim and
   the faces; the lo words of ghost cells are never read by the step and stay
   undefined (also in bnz_save_state blobs). */
int bnz_fill_ghosts(bnz_ctx *ctx);
/* new_dt (src/integration/new_dt.pyx:18-135): cour / max (|v|+c_f)/dx */
int bnz_new_dt(bnz_ctx *ctx, double *dt);
/* One VL step with a caller-supplied dt: predictor, corrector, ghost fill,
   prim2cons (src/integration/vanleer.pyx:100-163).  dt = 0 changes nothing. */
int bnz_vl_step(bnz_ctx *ctx, double dt);
/* vanleer.advance (src/integration/vanleer.pyx:87-167): steps with the
   device-computed CFL dt, clipped to tmax, until t >= tmax or max_steps
   (max_steps <= 0: no limit).  *nsteps gets the number taken.  Steps are
   issued in chunks; one that would start at t >= tmax does nothing. */
int bnz_advance(bnz_ctx *ctx, double tmax, int max_steps, int *nsteps);
/* nsteps steps with the device CFL dt and no tmax clip, fully
   stream-ordered (no host synchronisation). */
int bnz_run_steps(bnz_ctx *ctx, int nsteps);
int bnz_get_time(bnz_ctx *ctx, double *time, int *step, double *last_dt);
int bnz_set_time(bnz_ctx *ctx, double time, int step);
/* Conservation, div B and failure diagnostics (src/diagnostics.pyx:18-125
   prints the energies; the div B and positivity checks, and the location of
   the first failing cell with the step it was found at, have no reference
   counterpart). */
int bnz_diag(bnz_ctx *ctx, bnz_diag_out *out);
/* One history row of get_hst_vars_grid (src/io/output.pyx:20-138, the
   columns write_history appends to out/hst.txt, :221-243): step, time, then
   cell means over the active cells of rho, kinetic, thermal, magnetic and
   total energy density, mx, my, mz, Bx, By, Bz, and the rms deviations
   sqrt(|<B^2> - <B>^2|) of Bx, By, Bz.  16 doubles. */
int bnz_history(bnz_ctx *ctx, double *out16);
int bnz_synchronize(bnz_ctx *ctx);

/* Exact state for checkpoint/resume (the reference writes restart files of
   the grid data, src/io/restart.pyx:23-86): a header (time, step, last dt,
   the CFL maximum pending for the next step) and the arrays the next step
   reads -- FP64: prim and face B; FP32C: conserved hi and lo words, FP32
   prim and face B -- so that a context that loads it continues bit for bit
   like the one that saved it.  load_state needs a context of the same
   shape and precision and replaces upload + init_state. */
int bnz_state_size(const bnz_ctx *ctx, long *bytes);
int bnz_save_state(bnz_ctx *ctx, void *host, long bytes);
int bnz_load_state(bnz_ctx *ctx, const void *host, long bytes);

/* Work shapes of the 3D Godunov-flux, edge-EMF and stage-update kernels:
   0 one thread per face/edge/cell in x-fastest order (the default; the
   flux, EMF and update launches of 3D grids of at least 2^21 cells with
   ghosts then take their workgroups in runs of 16 dealt round-robin over
   the XCDs),
   1 tiles of 64 x 4 cells ordered along z (FP64; FP32C runs it flat),
   n > 1 slab columns of n rows walked plane by plane.  On 2D FP64 grids emf_chunk
   places the edge EMFs instead: 0 (the default) forms them inside the
   update launches up to 2^18 cells with ghosts and in launches of their own
   beyond, 1 always in launches of their own, 2 or more always inside the
   updates.  Results do not depend on either choice.  Chosen by the
   autotuner (bonzo_amd/autotune.py, 3D); no counterpart in the reference. */
int bnz_set_tuning(bnz_ctx *ctx, int flux_chunk, int emf_chunk, int update_chunk);

/* User boundary (flag 3) of one side.  The reference leaves such a side's
   ghost fill to code in the problem file (set_bc_grid_ptrs_user,
   src/problem/problem.pxd:13; the hook grid_bc.pyx:266-268; the one use,
   st2d.pyx:198-372: fixed inflow states on the x sides, rows copied with an
   x shift on the y sides).  Here the caller describes the fill once, for
   every cell of the side's slab: the ng ghost layers along the axis (side 0
   the lowest indices, 1 the highest) times the full extent, ghosts
   included, of the other axes, x fastest.
     src[q] >= 0: cell q takes every variable (prim and face fields) of that
                  cell of the full grid, (k * Ntot_y + j) * Ntot_x + i;
     src[q] == -1: it takes fixed[12 q .. 12 q + 11] = rho, vx, vy, vz, p,
                  scalar, bx, by, bz (cell-centred) and the x, y, z face
                  fields (fixed may be NULL without such entries);
     src[q] == -2: it is left alone.
   No source may be a cell the same side writes.  With a user side on a grid
   the fill runs one side after the other in the reference's order (z1, z2,
   y1, y2, x1, x2), each side reading what the earlier ones wrote, as the
   reference's per-side functions do.  Every ghost cell must end up with a
   value that comes from active cells or fixed values through earlier sides:
   a value read from a ghost cell that a later side fills (what that cell
   held between fills) is refused, as is a ghost cell no side writes
   (BNZ_EINVAL once the last flag-3 side is set; a refused side leaves the
   boundaries as they were).  Every flag-3 side must be set before
   bnz_init_state (BNZ_ESTATE otherwise); setting one again takes effect at
   the next fill.  Single grids, FP64 and FP32C. */
int bnz_set_user_bc(bnz_ctx *ctx, int axis, int side, const int *src, const double *fixed);

/* Step graphs in bnz_advance: on = 1 replays full chunks of eight steps
   from a captured hipGraph, on = 0 (the default) launches every kernel from
   the host.  On MI355X a replay measured slower than direct launches on 2D
   and 3D grids (a step there is bound by the device's per-kernel time, not
   by host launch cost) and up to 14 % faster on FP32C grids of a few
   thousand cells.  Results do not depend on the setting.  No counterpart in
   the reference. */
int bnz_set_step_graph(bnz_ctx *ctx, int on);

/* Kernel timing on the context's stream: when on, every launch is bracketed
   by HIP events and its duration accumulated per kernel name; switching it
   on clears the table. */
int bnz_set_kernel_timing(bnz_ctx *ctx, int on);
int bnz_kernel_timing(bnz_ctx *ctx, int idx, char *name, int name_len, double *total_ms,
                      long *launches);

/* Riemann solver on n faces, device pointers, layout [9][n] in the local
   frame of the face (slot 6 = normal field, overwritten by bx), on the
   default stream; synchronous.  Same seam as the pencil solvers of
   src/mhd/fluxes_c.h:3-21 (hll_flux, hlld_flux).  BNZ_RS_HLLC is the hydro
   build's HLLC (fluxes_c.c:230-327) on states with zero field. */
int bnz_gpu_flux_pencils(int rsolver, const double *wl, const double *wr, const double *bx,
                         double *flux, long n, double gam);

/* The reference's Riemann-solver seam itself: the signature of RSolverFunc
   (src/integration/integrator.pxd:76-79) and of hll_flux / hlld_flux /
   hllt_flux (src/mhd/fluxes_c.h:3-21), so that a maintainer can assign them
   to integr.rsolver_func in the reference's CPU driver.  Host row pointers:
   flux[n], wl[n], wr[n] for the 9 slots of the sweep's local frame and bx,
   indexed by face, faces i1..i2 inclusive.  Each call stages the rows
   through device memory and runs the solver on the GPU (synchronous; the
   PCIe round trip per pencil makes this a compatibility seam, not the fast
   path, which is bnz_vl_step).  A face whose solver matches no region keeps
   its flux row, as in the reference; the seam has no status, so a failure
   prints to stderr and leaves every flux row as it was.  Thread-safe: the
   staging memory is per host thread. */
void bnz_hll_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                  double gam);
void bnz_hlld_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                   double gam);
void bnz_hllt_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                   double gam);
/* The hydro build's hllc_flux (fluxes_c.c:230-327, bound in the build
   without -mf, integrator.pyx:259-264): the same seam with that build's 6
   rows (rho, vx, vy, vz, p, scalar; NMODE = 6) and bx unused (the hydro
   driver passes no field row, godunov.pyx:111-112); may be NULL. */
void bnz_hllc_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                   double gam);

/* No counterpart in the reference.  The Riemann-solver seam above for
   callers that cannot allocate memory (tight per-pencil loops, harnesses
   that forbid allocation): rsolver is BNZ_RS_HLL, BNZ_RS_HLLD or
   BNZ_RS_HLLT, the rows and faces are those of bnz_hll_flux & co., and the
   results are theirs bit for bit.  It allocates nothing: the rows are staged
   through a fixed host array and a fixed device array of
   28 x BNZ_NOALLOC_MAX_FACES doubles (3.7 MB) that each GPU it runs on holds
   until the process ends.  At most BNZ_NOALLOC_MAX_FACES faces per call
   (i2 - i1 + 1); a longer pencil returns BNZ_ERANGE and leaves the rows as
   they were.  Runs on the current device, synchronous; calls that overlap
   in one process race on the staging arrays, so callers serialise them.
   Returns BNZ_OK or a BNZ_E* code. */
#define BNZ_NOALLOC_MAX_FACES 16384
int bnz_rsolver_rows_noalloc(int rsolver, double **flux, double **wl, double **wr, double *bx,
                             int i1, int i2, double gam);

/* ------------------------------------------------------------------------
 * Batches of independent boxes of uneven size (parameter sweeps).  Each box
 * is the grid that one bnz_create with the same bnz_params would build (its
 * own sizes, extent, boundary flags and t_max); all boxes advance together,
 * one launch per stage over every box of a dimensionality (1D boxes run each
 * stage as one fused flux-and-update pass, and a batch of 1D boxes only each
 * whole step as one pass), with a per-box CFL dt, time and
 * "running" flag on the device, so that each box reproduces the single-grid
 * run of vanleer.advance (src/integration/vanleer.pyx:87-167) to its own
 * t_max bit for bit.  The reference has no batched form: a sweep there is one
 * process per parameter set.
 */
typedef struct bnz_batch bnz_batch;

typedef struct {
  int nx, ny, nz;   /* active cells, as bnz_params                         */
  int bc[6];        /* BNZ_BC_PERIODIC, _OUTFLOW or _REFLECT per side       */
  double lmin[3];   /* extent                                              */
  double lmax[3];
  double tmax;      /* [time] tmax of this box                             */
} bnz_box;

/* Common selection from p (rsolver, reconstr, gam, cour, dt_fms_ref,
   precision, device; its sizes, extent and flags are ignored). */
int bnz_batch_create(const bnz_params *p, const bnz_box *boxes, int nbox, bnz_batch **out);
int bnz_batch_destroy(bnz_batch *b);
const char *bnz_batch_last_error(const bnz_batch *b);
/* Ntot per axis of one box; its host arrays are [var][Ntot_z][Ntot_y][Ntot_x] */
int bnz_batch_box_shape(const bnz_batch *b, int box, int *ntot3);
int bnz_batch_upload(bnz_batch *b, int box, const double *prim, const double *bfld);
int bnz_batch_download(bnz_batch *b, int box, double *prim, double *bfld);
/* init_sim for every box and the first CFL maxima (vanleer.pyx:38-48) */
int bnz_batch_init_state(bnz_batch *b);
/* Steps until every box has reached its t_max, or max_steps (> 0) step
   launches.  use_graph: replay a hipGraph holding two captured steps (steps
   are then launched in pairs).  Steps go out in chunks (four pairs, or one
   eager step) and the host reads a chunk's count of running boxes while
   the next chunk runs, so up to one chunk may run after the last box
   finished (a box at its t_max idles).  *nsteps gets the number of step
   launches; bnz_batch_get_times gives the steps each box took. */
int bnz_batch_advance(bnz_batch *b, int max_steps, int use_graph, int *nsteps);
/* Workgroups of the stage launches: at most flux_blocks (2D/3D fluxes),
   emf_blocks (EMFs) and update_blocks (2D/3D updates and the fused 1D
   stages); each workgroup walks the work blocks of its launch that many
   apart.  0 restores the default (16384).  Results do not depend on it;
   bonzo_amd.autotune.autotune_batch picks it on the device.  Drops the
   captured graph.  No reference counterpart (the reference has no batch). */
int bnz_batch_set_tuning(bnz_batch *b, int flux_blocks, int emf_blocks, int update_blocks);
/* Per-box time and step count (arrays of nbox) */
int bnz_batch_get_times(bnz_batch *b, double *t, int *steps);
int bnz_batch_synchronize(bnz_batch *b);

#ifdef __cplusplus
}
#endif
#endif
