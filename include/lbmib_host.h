/*
 * lbmib_host.h -- the reference's per-step C API on top of the MI355X engine
 * (liblbmib_host.so).
 *
 * Same names, argument types and call order as cube_pthread/lbm_ib.c: a
 * caller written against the reference (GV/LV structs, do_thread-style loop
 * with a barrier after each call) links this library instead.  Every
 * per-step routine takes the reference's LV; when several threads call it,
 * only tid 0 enqueues device work and the others return at once, so the
 * caller's barriers keep working.  The fluid state lives on the GPU; the
 * fiber arrays and, for dumps, the fluid mirror are refreshed on the host by
 * lbmib_sync_to_host().
 *
 * Errors: the reference exit(1)s on a fiber leaving the grid
 * (lbm_ib.c:1458-1466); here the first device error is stored in
 * gv->last_error and returned by lbmib_host_status().
 */
#ifndef LBMIB_HOST_H
#define LBMIB_HOST_H

#include <pthread.h>
#include <stdbool.h>

#include "lbmib.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Fibernode, lbm_ib.c:54-59 (96 B, same field order) */
typedef struct fiber_node_t {
  double x, y, z;
  double bend_force_x, bend_force_y, bend_force_z;
  double stretch_force_x, stretch_force_y, stretch_force_z;
  double elastic_force_x, elastic_force_y, elastic_force_z;
} Fibernode;

/* Fiber, Fibersheet, Fibershape, lbm_ib.c:62-83 */
typedef struct fiber_t {
  Fibernode *nodes;
  int num_nodes;
} Fiber;

typedef struct fiber_sheet_t {
  Fiber *fibers;
  double width, height;
  int num_cols, num_rows;
  double x_orig, y_orig, z_orig;
} Fibersheet;

typedef struct fiber_shape_t {
  Fibersheet *sheets;
  int num_sheets;
} Fibershape;

/* Fluidnode, lbm_ib.c:88-97 (824 B, same field order): the element type of
 * the inlet/outlet surfaces.  The engine keeps the fluid state on the
 * device; init_df_inout fills df_inout of both surfaces on the host as the
 * reference does (the equilibrium feq(rho_l, (u_l, 0, 0)), lbm_ib.c:1610-1670). */
typedef struct fluid_node_t {
  double vel_x, vel_y, vel_z;
  double rho, pressure;
  double df1[19];
  double df2[19];
  double dfeq[19];
  double df_inout[2][19];
  double elastic_force_x, elastic_force_y, elastic_force_z;
} Fluidnode;

/* Fluidsurface, lbm_ib.c:100-103: y_dim * z_dim nodes, index y * z_dim + z */
typedef struct fluid_surface_t {
  Fluidnode *nodes;
} Fluidsurface;

/* Sub_Fluidgrid, lbm_ib.c:108-119: one cube_size^3 cube of the grid.  Here
 * the cube descriptors carry the coordinates of the reference's cube
 * decomposition (for callers that partition host work with cube2thread);
 * nodes stays NULL: the populations live on the device. */
typedef struct sub_fluid_grid_t {
  int sub_coordx_start, sub_coordx_end;
  int sub_coordy_start, sub_coordy_end;
  int sub_coordz_start, sub_coordz_end;
  Fluidnode *nodes;
  bool isboundary;
  int grid_dim_x;
  int grid_dim_y;
  int grid_dim_z;
} Sub_Fluidgrid;

/* Fluidgrid, lbm_ib.c:124-132 (the reference's fields first).  The
 * populations live on the device; the host keeps the dimensions, the
 * inlet/outlet surfaces, the cube descriptors and, after
 * lbmib_sync_to_host(), a mirror of the post-stream state in x-fastest
 * arrays (index x + x_dim*(y + y_dim*z)). */
typedef struct fluid_grid_t {
  int x_dim, y_dim, z_dim;
  Fluidsurface *inlet;
  Fluidsurface *outlet;
  Sub_Fluidgrid *sub_fluid_grid;  /* (x_dim/cube)(y_dim/cube)(z_dim/cube) cubes, BI*ncy*ncz + BJ*ncz + BK */
  double *df;     /* [19][n]: df1 == df2 after replace_old_DF */
  double *dfeq;   /* [19][n]: equilibrium of the last collision */
  double *rho;    /* [n] */
  double *vel;    /* [3][n] */
  double *force;  /* [3][n] elastic_force_{x,y,z} */
} Fluidgrid;

/* GV, lbm_ib.c:135-161, plus the device context and the physics switches
 * that the reference hard-codes. */
typedef struct gv_t {
  Fibershape *fiber_shape;
  Fluidgrid *fluid_grid;
  double tau, nu_l, u_l, rho_l, L_l, M_l, g_l, Ks_l, Kb_l, Kst_l;
  double Re, cs_l, Ma, Kn, Kshat, Ksthat, Kbhat, Mhat, Fr;
  int dt, time, TIME_STOP, dump;
  double c[19][3];
  int ib, ie, je, jb, ke, kb;
  pthread_t *threads;            /* set by the caller, as the reference's main() does */
  pthread_mutex_t *lock_Fluid;   /* set by the caller; the engine takes no host locks */
  int total_threads;
  int cube_size;
  int num_cubes_x, num_cubes_y, num_cubes_z;
  int P, Q, R;
  pthread_barrier_t barr;        /* do_thread waits on it after every routine when total_threads > 1 */
  /* -- MI355X engine -- */
  lbmib_ctx *dev;
  int custom_physics; /* 1: init_gv keeps tau,u_l,rho_l,Ks_l,Kb_l and the fields below */
  int bc_x, bc_y, bc_z;
  int precision;      /* LBMIB_F64 / F32 / F16 / BF16 */
  int quirks;
  int shift_all_sheets;
  int last_error;
  int quiet;
  int dump_mask;      /* do_thread dumps: 1 binary fiber, 2 fluid cubes, 4 FieldView, 8 text fiber */
} *GV;

/* LV, lbm_ib.c:164-167 */
typedef struct lv_t {
  int tid;
  GV gv;
} *LV;

/* construction: lbm_ib.c:705, :866, :924 */
void init_gv(GV gv, Fibershape *fiber_shape, Fluidgrid *fluid_grid, int cube_size);
Fibershape *gen_fiber_shape(GV gv, double width, double height, int num_cols, int num_rows, double x0, double y0,
                            double z0);
Fluidgrid *gen_fluid_grid(GV gv, int elem_x, int elem_y, int elem_z, int sub_fluid_grid_dim_k);
/* Places sheet s with its own geometry and U(-jitter, jitter) node jitter
 * drawn from SplitMix64(seed) in (row, col, xyz) order.  Not in the
 * reference (its gen_fiber_shape gives every sheet the same geometry). */
void gen_fiber_sheet_jittered(GV gv, int s, double width, double height, int num_cols, int num_rows, double x0,
                              double y0, double z0, double jitter, unsigned long long seed);

/* mappings, lbm_ib.c:3572-3602 (kept for callers that partition host work) */
int cube2thread(int BI, int BJ, int BK, int num_cubes_x, int num_cubes_y, int num_cubes_z, int P, int Q, int R);
int fiber2thread(int fiber_row, int num_fibers, int num_threads);

/* IB routines, lbm_ib.c:213-216 and the mpi/ twins */
void compute_bendingforce(LV lv);
void compute_stretchingforce(LV lv);
void compute_elasticforce(LV lv);
void get_influentialdomain_fluid_grid_and_SpreadForce(LV lv);
void fiber_SpreadForce(LV lv);
void fluid_get_SpreadForce(LV lv);
void fluid_SpreadVelocity(LV lv);
void fiber_get_SpreadVelocity(LV lv);

/* LBM routines, lbm_ib.c:219-228 */
void init_eqlbrmdistrfuncDF0(GV gv);
void init_df_inout(GV gv);
void init_df1(GV gv);
void compute_eqlbrmdistrfuncDF1(LV lv);
void stream_distrfunc(LV lv);
void bounceback_rigidwalls(LV lv);
void compute_rho_and_u(LV lv);
void moveFiberSheet(LV lv);
void copy_inout_to_df2(LV lv);
void periodicBC(LV lv);
void replace_old_DF(LV lv);

/* the step loop, lbm_ib.c:3087-3570: called by all total_threads threads
 * of the caller (each with its own LV) or by one; with total_threads > 1 every
 * thread waits on gv->barr (initialised by the caller for total_threads)
 * after each routine, as the reference does; thread 0 enqueues the device
 * work, writes the dumps (at time == 1 and every gv->dump steps when dumps
 * are enabled) and advances gv->time */
void *do_thread(void *v);

/* dump writers, lbm_ib.c:234-692 (fluid values from the host mirror of
 * lbmib_sync_to_host(gv, 1)) */
void print_fluid_sub_grid(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y, int end_z,
                          int cube_size);                                  /* lbm_ib.c:234-289 */
void print_fluid_cube(GV gv, int BI_start, int BJ_start, int BK_start, int BI_end, int BJ_end, int BK_end,
                      int cube_size);                                      /* lbm_ib.c:292-329 */
void save_fluid_sub_grid(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y, int end_z,
                         char fName[]);                                    /* lbm_ib.c:332-392 */
void save_fluid_sub_grid_Vel(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y, int end_z,
                             char fName[]);                                /* lbm_ib.c:395-432 */
void save_fluid_sub_grid_vel_use_fv(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y,
                                    int end_z, char fName[]);              /* lbm_ib.c:436-507 */
void save_fluid_sub_grid_cord_use_fv(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y,
                                     int end_z, char fName[]);             /* lbm_ib.c:510-548 */
void print_fiber_sub_grid(GV gv, int start_y, int start_z, int end_y, int end_z); /* lbm_ib.c:551-572 */
void save_fiber_sub_grid(GV gv, int start_y, int start_z, int end_y, int end_z, char fName[]);
void save_fiber_sub_grid_binary(GV gv, int start_y, int start_z, int end_y, int end_z, char fName[]);
void save_fiber_sub_grid_cord_ef_use_fv(GV gv, int start_y, int start_z, int end_y, int end_z, char grid_fname[],
                                        char elastic_force_fname[]);       /* lbm_ib.c:650-692 */
double get_cur_time(void);

/* host mirror and status */
int lbmib_sync_to_host(GV gv, int with_fluid);
int lbmib_host_status(GV gv);
void lbmib_host_free(GV gv);
/* dump switches for do_thread: 1 = binary fiber dump, 2 = fluid cube dump,
 * 8 = text fiber dump (lbm_ib.c:560-601),
 * 4 = FieldView files (the reference's USE_FV, lbm_ib.c:30) */
void lbmib_host_set_dumps(GV gv, int mask);

#ifdef __cplusplus
}
#endif
#endif /* LBMIB_HOST_H */
