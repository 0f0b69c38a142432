/*
 * bnz_oracle.h -- CPU oracle for the Van-Leer MHD step (TEST INFRASTRUCTURE).
 *
 * This is the checker, never the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It restates, in plain C99 and
 * with the reference's operation order, the algorithm of
 *   src/mhd/eos_c.h, src/mhd/fluxes_c.c, src/coordinates/reconstr_c.{c,h},
 *   src/integration/{godunov,ct,new_dt,integrator,vanleer}.pyx,
 *   src/bc/{grid_bc,grid_bc_funcs,util_bc}.pyx, src/coordinates/{grid,coord,
 *   cartesian}.pyx
 * of sergekomarov/bonzo.  Array layout is the reference's: C-contiguous
 * [var][k][j][i], x fastest, ng ghost cells per active axis.
 *
 * Every function cites the reference file:line it follows.  Behavioural
 * decisions where the reference is broken are marked D<n>.
 */
#ifndef BNZ_ORACLE_H
#define BNZ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_NMODE = 9, OR_NVAR = 12 };

/* Riemann solver ids; same numbering as the reference's RSolver enum
   (src/integration/integrator.pxd:59-65). */
enum { OR_RS_HLL = 0, OR_RS_HLLC = 1, OR_RS_HLLD = 2, OR_RS_HLLA = 3, OR_RS_HLLT = 4 };

/* ---- point functions (src/mhd/eos_c.h) ---- */
void   or_cons2prim(double *w, const double *u, double gam);
void   or_prim2cons(double *u, const double *w, double gam);
void   or_primcons2flux(double *f, const double *w, const double *u, double gam);
double or_fms(const double *w, double bx, double gam);
double or_fms_hydro(const double *w, double gam);
double or_vl_lim(double a, double b);
/* characteristic projection (src/mhd/eos_c.h:100-280): in place on vc, in the
   eigensystem of state w */
void   or_prim2char(double *vc, const double *w, double gam);
void   or_char2prim(double *vc, const double *w, double gam);

/* ---- pencil ABI (src/mhd/fluxes_c.h:3-21; integrator.pxd:76-79) ---- */
void or_hll_flux (double **flux, double **wl, double **wr, double *bx, int i1, int i2, double gam);
void or_hllt_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2, double gam);
void or_hlld_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2, double gam);
/* hydro build (MFIELD=0), 9-slot rows with zero field */
void or_hll_flux_hydro(double **flux, double **wl, double **wr, double *bx, int i1, int i2,
                       double gam);
void or_hllc_flux(double **flux, double **wl, double **wr, double *bx, int i1, int i2, double gam);

/* ---- reconstruction (src/coordinates/reconstr_c.h:4-12), uniform grids ---- */
void or_reconstr_const (double **wl, double **wr, double ***w, double ***scr, int ax, int i1, int i2);
void or_reconstr_linear(double **wl, double **wr, double ***w, double ***scr, int ax, int i1, int i2);
void or_reconstr_linear_cp(double **wl, double **wr, double ***w, double ***scr, int ax, int i1,
                           int i2, int char_proj, double gam);

/* ---- whole-grid simulation object ---- */
typedef struct or_sim or_sim;

/* bc flags per the reference: 0 periodic, 1 outflow, 2 reflect (src/bc/grid_bc.pyx:45-50) */
or_sim *or_create(int nx, int ny, int nz, int reconstr_order,
                  const double *lmin, const double *lmax,
                  double gam, double cour, int rsolver, const int *bc_flags,
                  int nthreads);
or_sim *or_create_ex(int nx, int ny, int nz, int reconstr_order, const double *lmin,
                     const double *lmax, double gam, double cour, int rsolver,
                     const int *bc_flags, int nthreads, int tintegr); /* 1: RK3, ng 9 */
void    or_destroy(or_sim *s);
int     or_ntot(const or_sim *s, int axis);
int     or_ng(const or_sim *s);
double *or_prim(or_sim *s);
double *or_cons(or_sim *s);
double *or_bfld(or_sim *s);
double *or_coord(or_sim *s, int which, int axis); /* 0 lf,1 lv,2 dlf,3 dlf_inv */
void    or_set_dt_fms_ref(or_sim *s, int on);
/* PLM in characteristic variables (charact_proj=1, src/coordinates/reconstr_c.c:64-80) */
void    or_set_char_proj(or_sim *s, int on);
void    or_set_hydro(or_sim *s, int on);
void    or_set_ct_upwind(or_sim *s, int on);
/* user boundary (flag 3) of one side, the layout of bnz_set_user_bc */
int     or_set_user_bc(or_sim *s, int axis, int side, const int *src, const double *fixed);
double *or_flux(or_sim *s, int axis);   /* [9][cells] face fluxes (tests)  */
double *or_edge_emf(or_sim *s);         /* [3][cells] edge EMFs (tests)    */
void    or_corner_emf(or_sim *s, int upwind, double dt);
void    or_apply_grid_bc(or_sim *s);
void    or_init(or_sim *s);
double  or_new_dt(or_sim *s);
void    or_vl_step(or_sim *s, double dt);
void    or_rk3_step(or_sim *s, double dt);
int     or_advance(or_sim *s, double tmax, int max_steps);
int     or_advance_steps(or_sim *s, int nsteps);
double  or_time(const or_sim *s);
int     or_step(const or_sim *s);
double  or_last_dt(const or_sim *s);
/* history row of get_hst_vars_grid (src/io/output.pyx:20-138): step, time,
   cell means of rho, Ek, Et, Em, E, mx, my, mz, Bx, By, Bz, rms Bx, By, Bz */
void    or_history(or_sim *s, double *out16);

#ifdef __cplusplus
}
#endif
#endif
