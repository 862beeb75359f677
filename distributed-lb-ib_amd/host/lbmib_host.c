/*
 * lbmib_host.c -- the reference's per-step API (include/lbmib_host.h) over
 * the MI355X engine's C-ABI (include/lbmib.h).
 *
 * This is the drop-in layer: the routine names, the GV/LV arguments, the
 * call order of do_thread and the dump formats follow
 * cube_pthread/lbm_ib.c; the work itself is enqueued on the GPU by
 * liblbmib.so.  Only thread 0 of a multi-threaded caller enqueues.
 */
#define _POSIX_C_SOURCE 200809L
#include "../../include/lbmib_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static const int HCX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int HCY[19] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, -1, 1, 1, -1};
static const int HCZ[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0};

static double host_feq(int q, double rho, double vx, double vy, double vz);

static void note(GV gv, int rc) {
  if (rc && !gv->last_error) gv->last_error = rc;
}

double get_cur_time(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + tv.tv_usec / 1000000.0;
}

/* ---------------------------------------------------------- construction */

Fibershape *gen_fiber_shape(GV gv, double w, double h, int num_cols, int num_rows, double x_orig, double y_orig,
                            double z_orig) {
  /* lbm_ib.c:866-921: every sheet gets this geometry, normal to x, rows
   * along y and columns along z */
  Fibershape *shape = gv->fiber_shape;
  if (!shape) {
    shape = (Fibershape *)calloc(1, sizeof(Fibershape));
    shape->num_sheets = 1;
    shape->sheets = (Fibersheet *)calloc(1, sizeof(Fibersheet));
    gv->fiber_shape = shape;
  }
  for (int s = 0; s < shape->num_sheets; ++s)
    gen_fiber_sheet_jittered(gv, s, w, h, num_cols, num_rows, x_orig, y_orig, z_orig, 0.0, 0);
  return shape;
}

static unsigned long long splitmix64_next(unsigned long long *state) {
  unsigned long long z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void gen_fiber_sheet_jittered(GV gv, int s, double w, double h, int num_cols, int num_rows, double x_orig,
                              double y_orig, double z_orig, double jitter, unsigned long long seed) {
  Fibersheet *sheet = gv->fiber_shape->sheets + s;
  double zdist = w / (num_cols - 1);
  double ydist = h / (num_rows - 1);
  sheet->num_cols = num_cols;
  sheet->num_rows = num_rows;
  sheet->width = w;
  sheet->height = h;
  sheet->x_orig = x_orig;
  sheet->y_orig = y_orig;
  sheet->z_orig = z_orig;
  sheet->fibers = (Fiber *)calloc(num_rows, sizeof(Fiber));
  unsigned long long st = seed;
  for (int row = 0; row < num_rows; ++row) {
    Fiber *fiber = sheet->fibers + row;
    fiber->num_nodes = num_cols;
    fiber->nodes = (Fibernode *)calloc(num_cols, sizeof(Fibernode));
    for (int col = 0; col < num_cols; ++col) {
      Fibernode *nd = fiber->nodes + col;
      nd->x = x_orig;
      nd->y = ydist * row + y_orig;
      nd->z = zdist * col + z_orig;
      if (jitter != 0.0) {
        double *p[3] = {&nd->x, &nd->y, &nd->z};
        for (int a = 0; a < 3; ++a) {
          double u = (double)(splitmix64_next(&st) >> 11) * (1.0 / 9007199254740992.0);
          *p[a] += jitter * (2.0 * u - 1.0);
        }
      }
    }
  }
}

Fluidgrid *gen_fluid_grid(GV gv, int dim_x, int dim_y, int dim_z, int cube_size) {
  /* lbm_ib.c:924-958: the grid, its inlet/outlet surfaces and the cube
   * descriptors; the populations are allocated on the device by init_gv,
   * the host mirror on demand by lbmib_sync_to_host */
  Fluidgrid *g = (Fluidgrid *)calloc(1, sizeof(Fluidgrid));
  g->x_dim = dim_x;
  g->y_dim = dim_y;
  g->z_dim = dim_z;
  g->inlet = (Fluidsurface *)calloc(1, sizeof(Fluidsurface));
  g->outlet = (Fluidsurface *)calloc(1, sizeof(Fluidsurface));
  g->inlet->nodes = (Fluidnode *)calloc((size_t)dim_y * dim_z, sizeof(Fluidnode));
  g->outlet->nodes = (Fluidnode *)calloc((size_t)dim_y * dim_z, sizeof(Fluidnode));
  if (cube_size > 0) {
    int ncx = dim_x / cube_size, ncy = dim_y / cube_size, ncz = dim_z / cube_size;
    g->sub_fluid_grid = (Sub_Fluidgrid *)calloc((size_t)ncx * ncy * ncz, sizeof(Sub_Fluidgrid));
    for (int bi = 0; bi < ncx; ++bi)
      for (int bj = 0; bj < ncy; ++bj)
        for (int bk = 0; bk < ncz; ++bk) {
          Sub_Fluidgrid *c = g->sub_fluid_grid + ((size_t)bi * ncy + bj) * ncz + bk;
          c->sub_coordx_start = bi * cube_size; c->sub_coordx_end = (bi + 1) * cube_size - 1;
          c->sub_coordy_start = bj * cube_size; c->sub_coordy_end = (bj + 1) * cube_size - 1;
          c->sub_coordz_start = bk * cube_size; c->sub_coordz_end = (bk + 1) * cube_size - 1;
          c->nodes = NULL;
          c->isboundary = bi == 0 || bj == 0 || bk == 0 || bi == ncx - 1 || bj == ncy - 1 || bk == ncz - 1;
          c->grid_dim_x = c->grid_dim_y = c->grid_dim_z = cube_size;
        }
  }
  gv->fluid_grid = g;
  return g;
}

static void set_lattice(GV gv) {
  for (int q = 0; q < 19; ++q) {
    gv->c[q][0] = HCX[q];
    gv->c[q][1] = HCY[q];
    gv->c[q][2] = HCZ[q];
  }
}

void init_gv(GV gv, Fibershape *fiber_shape, Fluidgrid *fluid_grid, int cube_size) {
  if (fiber_shape == NULL) { fprintf(stderr, "Error(%s): Null fiber shape!\n", __func__); gv->last_error = LBMIB_ERR_ARG; return; }
  if (fluid_grid == NULL) { fprintf(stderr, "Error(%s): Null fluid grid!\n", __func__); gv->last_error = LBMIB_ERR_ARG; return; }
  gv->fiber_shape = fiber_shape;
  gv->fluid_grid = fluid_grid;
  gv->cube_size = cube_size;
  int dim_x = fluid_grid->x_dim, dim_y = fluid_grid->y_dim, dim_z = fluid_grid->z_dim;
  if (cube_size > 0) {
    gv->num_cubes_x = dim_x / cube_size;
    gv->num_cubes_y = dim_y / cube_size;
    gv->num_cubes_z = dim_z / cube_size;
  }
  gv->ib = 2; gv->ie = dim_x - 3; gv->jb = 2; gv->je = dim_y - 3; gv->kb = 2; gv->ke = dim_z - 3;
  if (!gv->custom_physics) {
    /* lbm_ib.c:750-773 */
    gv->Re = 1.5e2;
    gv->rho_l = 1.0e0;
    gv->u_l = 0.001;
    gv->L_l = 2.0e1;
    gv->nu_l = gv->u_l * gv->L_l / gv->Re;
    gv->tau = 0.5 * (6.0 * gv->nu_l + 1.0);
    gv->Fr = 1.0e10;
    gv->Kbhat = 0.005;
    gv->Kshat = 2.0e1;
    gv->Ksthat = 0.0e0;
    gv->g_l = 0.0;
    gv->Kb_l = gv->Kbhat * gv->rho_l * gv->u_l * gv->u_l * pow(gv->L_l, 4);
    gv->Ks_l = gv->Kshat * gv->rho_l * gv->u_l * gv->u_l * gv->L_l;
    gv->bc_x = LBMIB_BC_INOUT;
    gv->bc_y = LBMIB_BC_WALL;
    gv->bc_z = LBMIB_BC_WALL;
    gv->quirks = 1;
    gv->shift_all_sheets = 0;
  }
  gv->cs_l = 1.0 / (sqrt(3));
  set_lattice(gv);
  gv->dt = 1;
  gv->time = gv->dt;

  /* the one-step pre-shift X += u_l dt, lbm_ib.c:857-862 (sheets[0] only in
   * the reference) */
  int nsh = fiber_shape->num_sheets;
  for (int s = 0; s < (gv->shift_all_sheets ? nsh : (nsh > 0 ? 1 : 0)); ++s) {
    Fibersheet *sh = fiber_shape->sheets + s;
    for (int i = 0; i < sh->num_rows; i++)
      for (int j = 0; j < sh->num_cols; j++) sh->fibers[i].nodes[j].x += gv->u_l * gv->dt;
  }

  lbmib_fluid_params fp;
  memset(&fp, 0, sizeof(fp));
  fp.nx = dim_x; fp.ny = dim_y; fp.nz = dim_z;
  fp.bc_x = gv->bc_x; fp.bc_y = gv->bc_y; fp.bc_z = gv->bc_z;
  fp.precision = gv->precision;
  fp.quirks = gv->quirks;
  fp.tau = gv->tau;
  fp.rho0 = gv->rho_l;
  fp.u0[0] = gv->u_l;
  fp.g = gv->g_l;
  lbmib_sheet_params sp[16];
  if (nsh > 16) { gv->last_error = LBMIB_ERR_ARG; return; }
  for (int s = 0; s < nsh; ++s) {
    Fibersheet *sh = fiber_shape->sheets + s;
    sp[s].rows = sh->num_rows; sp[s].cols = sh->num_cols;
    sp[s].width = sh->width; sp[s].height = sh->height;
    sp[s].ks = gv->Ks_l; sp[s].kb = gv->Kb_l;
  }
  int rc = lbmib_create(&fp, nsh, sp, &gv->dev);
  if (rc) { note(gv, rc); gv->dev = NULL; return; }
  for (int s = 0; s < nsh; ++s) {
    Fibersheet *sh = fiber_shape->sheets + s;
    size_t m = (size_t)sh->num_rows * sh->num_cols;
    double *xyz = (double *)malloc(sizeof(double) * 3 * m);
    for (int i = 0; i < sh->num_rows; ++i)
      for (int j = 0; j < sh->num_cols; ++j) {
        Fibernode *nd = sh->fibers[i].nodes + j;
        size_t k = (size_t)i * sh->num_cols + j;
        xyz[3 * k] = nd->x; xyz[3 * k + 1] = nd->y; xyz[3 * k + 2] = nd->z;
      }
    note(gv, lbmib_set_fiber_positions(gv->dev, s, xyz));
    free(xyz);
  }
}

void lbmib_host_free(GV gv) {
  if (gv->dev) lbmib_destroy(gv->dev);
  gv->dev = NULL;
  if (gv->fluid_grid) {
    /* every pointer is cleared after its free, so a caller that frees
     * sub_fluid_grid itself afterwards, as the reference's main() does
     * (lbm_ib.c:3949), frees NULL */
    Fluidgrid *g = gv->fluid_grid;
    free(g->df); free(g->dfeq); free(g->rho); free(g->vel); free(g->force);
    g->df = g->dfeq = g->rho = g->vel = g->force = NULL;
    Fluidsurface *sf[2] = {g->inlet, g->outlet};
    for (int k = 0; k < 2; ++k)
      if (sf[k]) { free(sf[k]->nodes); free(sf[k]); }
    g->inlet = g->outlet = NULL;
    free(g->sub_fluid_grid);
    g->sub_fluid_grid = NULL;
  }
}

void lbmib_host_set_dumps(GV gv, int mask) { gv->dump_mask = mask; }

/* ------------------------------------------------------------- mappings */

int fiber2thread(int fiber_row_i, int num_fibers, int num_threads) {
  /* lbm_ib.c:3572-3583 */
  int block = num_fibers / num_threads;
  if (num_fibers % num_threads != 0 && block * num_threads < num_fibers) block = block + 1;
  return fiber_row_i / block;
}

int cube2thread(int BI, int BJ, int BK, int num_cubes_x, int num_cubes_y, int num_cubes_z, int P, int Q, int R) {
  /* lbm_ib.c:3585-3602 */
  int bx = num_cubes_x / P, by = num_cubes_y / Q, bz = num_cubes_z / R;
  return (BI / bx) * Q * R + (BJ / by) * R + (BK / bz);
}

/* -------------------------------------------------------- per-step calls */

#define TID0(lv)          \
  if ((lv)->tid != 0) return; \
  GV gv = (lv)->gv;       \
  if (!gv->dev) return;

void compute_bendingforce(LV lv) { TID0(lv); note(gv, lbmib_compute_bendingforce(gv->dev)); }
void compute_stretchingforce(LV lv) { TID0(lv); note(gv, lbmib_compute_stretchingforce(gv->dev)); }
void compute_elasticforce(LV lv) { TID0(lv); note(gv, lbmib_compute_elasticforce(gv->dev)); }
void get_influentialdomain_fluid_grid_and_SpreadForce(LV lv) {
  TID0(lv);
  note(gv, lbmib_get_influentialdomain_fluid_grid_and_SpreadForce(gv->dev));
}
void fiber_SpreadForce(LV lv) { TID0(lv); note(gv, lbmib_fiber_SpreadForce(gv->dev)); }
void fluid_get_SpreadForce(LV lv) { TID0(lv); note(gv, lbmib_fluid_get_SpreadForce(gv->dev)); }
void fluid_SpreadVelocity(LV lv) { TID0(lv); note(gv, lbmib_fluid_SpreadVelocity(gv->dev)); }
void fiber_get_SpreadVelocity(LV lv) { TID0(lv); note(gv, lbmib_fiber_get_SpreadVelocity(gv->dev)); }
void compute_eqlbrmdistrfuncDF1(LV lv) { TID0(lv); note(gv, lbmib_compute_eqlbrmdistrfuncDF1(gv->dev)); }
void stream_distrfunc(LV lv) { TID0(lv); note(gv, lbmib_stream_distrfunc(gv->dev)); }
void bounceback_rigidwalls(LV lv) { TID0(lv); note(gv, lbmib_bounceback_rigidwalls(gv->dev)); }
void compute_rho_and_u(LV lv) { TID0(lv); note(gv, lbmib_compute_rho_and_u(gv->dev)); }
void moveFiberSheet(LV lv) { TID0(lv); note(gv, lbmib_moveFiberSheet(gv->dev)); }
void copy_inout_to_df2(LV lv) { TID0(lv); note(gv, lbmib_copy_inout_to_df2(gv->dev)); }
void periodicBC(LV lv) { TID0(lv); note(gv, lbmib_periodicBC(gv->dev)); }
void replace_old_DF(LV lv) { TID0(lv); note(gv, lbmib_replace_old_DF(gv->dev)); }

void init_eqlbrmdistrfuncDF0(GV gv) { if (gv->dev) note(gv, lbmib_init_eqlbrmdistrfuncDF0(gv->dev)); }
void init_df1(GV gv) { if (gv->dev) note(gv, lbmib_init_df1(gv->dev)); }
void init_df_inout(GV gv) {
  if (!gv->dev) return;
  note(gv, lbmib_init_df_inout(gv->dev));
  /* the host tables of lbm_ib.c:1610-1670: the initial equilibrium at the
   * inlet and (sampled at x = N-1) the outlet, both feq(rho_l, (u_l, 0, 0)) */
  Fluidgrid *g = gv->fluid_grid;
  if (!g || !g->inlet || !g->outlet || !g->inlet->nodes || !g->outlet->nodes) return;
  double tab[19];
  for (int q = 0; q < 19; ++q) tab[q] = host_feq(q, gv->rho_l, gv->u_l, 0.0, 0.0);
  size_t m = (size_t)g->y_dim * g->z_dim;
  for (size_t k = 0; k < m; ++k)
    for (int q = 0; q < 19; ++q) {
      g->inlet->nodes[k].df_inout[0][q] = tab[q];
      g->outlet->nodes[k].df_inout[1][q] = tab[q];
    }
}

/* ----------------------------------------------------------- host mirror */

int lbmib_host_status(GV gv) {
  if (gv->dev) note(gv, lbmib_status(gv->dev));
  return gv->last_error;
}

int lbmib_sync_to_host(GV gv, int with_fluid) {
  if (!gv->dev) return LBMIB_ERR_STATE;
  Fibershape *shape = gv->fiber_shape;
  for (int s = 0; s < shape->num_sheets; ++s) {
    Fibersheet *sh = shape->sheets + s;
    size_t m = (size_t)sh->num_rows * sh->num_cols;
    double *buf = (double *)malloc(sizeof(double) * 12 * m);
    int rc = lbmib_get_fibers(gv->dev, s, buf);
    if (rc) { free(buf); note(gv, rc); return rc; }
    for (int i = 0; i < sh->num_rows; ++i)
      memcpy(sh->fibers[i].nodes, buf + (size_t)12 * i * sh->num_cols, sizeof(Fibernode) * sh->num_cols);
    free(buf);
  }
  if (with_fluid) {
    Fluidgrid *g = gv->fluid_grid;
    size_t n = (size_t)g->x_dim * g->y_dim * g->z_dim;
    if (!g->df) g->df = (double *)malloc(sizeof(double) * 19 * n);
    if (!g->rho) g->rho = (double *)malloc(sizeof(double) * n);
    if (!g->vel) g->vel = (double *)malloc(sizeof(double) * 3 * n);
    if (!g->force) g->force = (double *)malloc(sizeof(double) * 3 * n);
    int rc = lbmib_get_df(gv->dev, g->df);
    if (!rc) rc = lbmib_get_macroscopic(gv->dev, g->rho, g->vel);
    if (!rc) rc = lbmib_get_force(gv->dev, g->force);
    if (rc) { note(gv, rc); return rc; }
  }
  return lbmib_host_status(gv);
}

/* equilibrium of init_eqlbrmdistrfuncDF0 / compute_eqlbrmdistrfuncDF1 */
static double host_feq(int q, double rho, double vx, double vy, double vz) {
  double usq = vx * vx + vy * vy + vz * vz;
  if (q == 0) return 1.0 / 3.0 * rho * (1.0 - 1.5 * usq);
  double cu = HCX[q] * vx + HCY[q] * vy + HCZ[q] * vz;
  double w = q <= 6 ? 1.0 / 18.0 : 1.0 / 36.0;
  return w * rho * (1.0 + 3.0 * cu + 4.5 * (pow(cu, 2)) - 1.5 * usq);
}

/* dfeq of a step is feq of the rho/u the collision used (those of the
 * previous step); the host captures them before compute_eqlbrmdistrfuncDF1
 * of a dump step. */
static int capture_dfeq(GV gv) {
  Fluidgrid *g = gv->fluid_grid;
  size_t n = (size_t)g->x_dim * g->y_dim * g->z_dim;
  double *rho = (double *)malloc(sizeof(double) * n), *vel = (double *)malloc(sizeof(double) * 3 * n);
  int rc = lbmib_get_macroscopic(gv->dev, rho, vel);
  if (!rc) {
    if (!g->dfeq) g->dfeq = (double *)malloc(sizeof(double) * 19 * n);
    for (size_t c = 0; c < n; ++c)
      for (int q = 0; q < 19; ++q) g->dfeq[q * n + c] = host_feq(q, rho[c], vel[c], vel[n + c], vel[2 * n + c]);
  }
  free(rho);
  free(vel);
  note(gv, rc);
  return rc;
}

/* ------------------------------------------------------------ dump files */

void save_fiber_sub_grid(GV gv, int start_y, int start_z, int end_y, int end_z, char fName[]) {
  /* text format of lbm_ib.c:560-601 (sheet 0) */
  FILE *f = fopen(fName, "w");
  if (!f) return;
  Fiber *fa = gv->fiber_shape->sheets[0].fibers;
  fprintf(f, "Fiber_sheets[%d] (i,j): {cord_x, cord_y, cord_z} || SF_x, SF_y, SF_z || BF_x, BF_y, BF_z || EF_x, EF_y, EF_z\n", 0);
  for (int i = start_y; i <= end_y; ++i)
    for (int j = start_z; j <= end_z; ++j) {
      Fibernode *nd = fa[i].nodes + j;
      fprintf(f, "(%2d,%2d):{%.24f,%.24f,%.24f} || %.24f,%.24f,%.24f || %.24f,%.24f,%.24f || %.24f,%.24f,%.24f\n", i, j,
              nd->x, nd->y, nd->z, nd->stretch_force_x, nd->stretch_force_y, nd->stretch_force_z, nd->bend_force_x,
              nd->bend_force_y, nd->bend_force_z, nd->elastic_force_x, nd->elastic_force_y, nd->elastic_force_z);
    }
  fclose(f);
}

void save_fiber_sub_grid_binary(GV gv, int start_y, int start_z, int end_y, int end_z, char fName[]) {
  /* {int32 i, int32 j, Fibernode} records, lbm_ib.c:604-646, for every
   * sheet in order (the reference sizes its buffer for one sheet only) */
  FILE *f = fopen(fName, "wb");
  if (!f) return;
  for (int s = 0; s < gv->fiber_shape->num_sheets; ++s) {
    Fiber *fa = gv->fiber_shape->sheets[s].fibers;
    for (int i = start_y; i <= end_y; ++i)
      for (int j = start_z; j <= end_z; ++j) {
        int ij[2] = {i, j};
        fwrite(ij, sizeof(int), 2, f);
        fwrite(fa[i].nodes + j, sizeof(Fibernode), 1, f);
      }
  }
  fclose(f);
}

void save_fluid_sub_grid(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y, int end_z,
                         char fName[]) {
  /* %.24f cube dump of lbm_ib.c:332-392: whole cubes covering the box, z
   * fastest inside a cube */
  FILE *f = fopen(fName, "w");
  if (!f) return;
  Fluidgrid *g = gv->fluid_grid;
  const int cs = gv->cube_size > 0 ? gv->cube_size : 1;
  const size_t n = (size_t)g->x_dim * g->y_dim * g->z_dim;
  for (int BI = start_x / cs; BI <= end_x / cs; ++BI)
    for (int BJ = start_y / cs; BJ <= end_y / cs; ++BJ)
      for (int BK = start_z / cs; BK <= end_z / cs; ++BK) {
        fprintf(f, "(BI,BJ,BK): {vel_x, vel_y, vel_z} || {G0, DF1, DF2}|| rho || {ElasticF_x, y, z}\n");
        for (int li = 0; li < cs; li++)
          for (int lj = 0; lj < cs; lj++)
            for (int lk = 0; lk < cs; lk++) {
              fprintf(f, "For cube <%d,%d,%d> Rank %d\n", BI, BJ, BK, 0);
              int X = BI * cs + li, Y = BJ * cs + lj, Z = BK * cs + lk;
              if (X >= g->x_dim || Y >= g->y_dim || Z >= g->z_dim) continue;
              size_t c = (size_t)X + (size_t)g->x_dim * ((size_t)Y + (size_t)g->y_dim * Z);
              for (int q = 0; q < 19; ++q) {
                double eq = g->dfeq ? g->dfeq[q * n + c] : 0.0;
                fprintf(f, "(%d,%d,%d, %2d):{%.24f,%.24f,%.24f} || {%.24f,%.24f,%.24f} || %.24f || {%.24f,%.24f,%.24f}\n",
                        X, Y, Z, q, g->vel[c], g->vel[n + c], g->vel[2 * n + c], eq, g->df[q * n + c],
                        g->df[q * n + c], g->rho[c], g->force[c], g->force[n + c], g->force[2 * n + c]);
              }
            }
      }
  fclose(f);
}

/* Host mirror value helpers: the fluid arrays are x-fastest planes. */
static size_t mirror_cell(const Fluidgrid *g, int X, int Y, int Z) {
  return (size_t)X + (size_t)g->x_dim * ((size_t)Y + (size_t)g->y_dim * Z);
}

/* Whole cubes BI..BI_end x BJ..BJ_end x BK..BK_end (closed), every node of
 * each with its cube-local coordinates and 19 populations, to stdout, in the
 * layout of print_fluid_sub_grid / print_fluid_cube (lbm_ib.c:234-330;
 * `tag` is the per-node line of each). */
static void print_cubes(GV gv, int BI0, int BJ0, int BK0, int BI1, int BJ1, int BK1, int cs, const char *tag) {
  Fluidgrid *g = gv->fluid_grid;
  const size_t n = (size_t)g->x_dim * g->y_dim * g->z_dim;
  if (cs <= 0) cs = 1;
  printf("(BI,BJ,BK): {vel_x, vel_y, vel_z} || {G0, DF1, DF2}|| rho || {ElasticF_x, y, z}\n");
  for (int BI = BI0; BI <= BI1; ++BI)
    for (int BJ = BJ0; BJ <= BJ1; ++BJ)
      for (int BK = BK0; BK <= BK1; ++BK)
        for (int li = 0; li < cs; li++)
          for (int lj = 0; lj < cs; lj++)
            for (int lk = 0; lk < cs; lk++) {
              printf(tag, BI, BJ, BK);
              const int X = BI * cs + li, Y = BJ * cs + lj, Z = BK * cs + lk;
              if (X >= g->x_dim || Y >= g->y_dim || Z >= g->z_dim) continue;
              const size_t c = mirror_cell(g, X, Y, Z);
              for (int q = 0; q < 19; ++q)
                printf("(%d, %d, %d, %d):{%.12f, %.12f, %.12f} || {%.12f, %.12f, %.12f} || %.12f || {%.12f, %.12f, %.12f}\n",
                       li, lj, lk, q, g->vel[c], g->vel[n + c], g->vel[2 * n + c], g->dfeq ? g->dfeq[q * n + c] : 0.0,
                       g->df[q * n + c], g->df[q * n + c], g->rho[c], g->force[c], g->force[n + c],
                       g->force[2 * n + c]);
            }
  printf("\n");
}

void print_fluid_sub_grid(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y, int end_z,
                          int cube_size) {
  const int cs = cube_size > 0 ? cube_size : 1;
  print_cubes(gv, start_x / cs, start_y / cs, start_z / cs, end_x / cs, end_y / cs, end_z / cs, cs,
              "For cube <%d, %d, %d>\n");
}

void print_fluid_cube(GV gv, int BI_start, int BJ_start, int BK_start, int BI_end, int BJ_end, int BK_end,
                      int cube_size) {
  print_cubes(gv, BI_start, BJ_start, BK_start, BI_end, BJ_end, BK_end, cube_size, "FOr cube <%d,%d,%d>\n");
}

void save_fluid_sub_grid_Vel(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y, int end_z,
                             char fName[]) {
  /* "u,v,w" lines over [start, end) with X slowest and Z fastest,
   * lbm_ib.c:395-432 (there the cube counts are left unset, :400; here the
   * mirror is indexed directly) */
  FILE *f = fopen(fName, "w");
  if (!f) return;
  Fluidgrid *g = gv->fluid_grid;
  const size_t n = (size_t)g->x_dim * g->y_dim * g->z_dim;
  for (int X = start_x; X < end_x; X++)
    for (int Y = start_y; Y < end_y; Y++)
      for (int Z = start_z; Z < end_z; Z++) {
        const size_t c = mirror_cell(g, X, Y, Z);
        fprintf(f, "%f,%f,%f\n", g->vel[c], g->vel[n + c], g->vel[2 * n + c]);
      }
  fclose(f);
}

void save_fluid_sub_grid_vel_use_fv(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y,
                                    int end_z, char fName[]) {
  /* FieldView function file, lbm_ib.c:436-507: "nx,ny,nz,3", then all u,
   * all v, all w over [start, end) with X fastest, each value "%f," */
  FILE *f = fopen(fName, "w");
  if (!f) return;
  Fluidgrid *g = gv->fluid_grid;
  const size_t n = (size_t)g->x_dim * g->y_dim * g->z_dim;
  fprintf(f, "%d,%d,%d,3\n", end_x - start_x, end_y - start_y, end_z - start_z);
  for (int comp = 0; comp < 3; ++comp)
    for (int Z = start_z; Z < end_z; Z++)
      for (int Y = start_y; Y < end_y; Y++)
        for (int X = start_x; X < end_x; X++) fprintf(f, "%f,", g->vel[comp * n + mirror_cell(g, X, Y, Z)]);
  fclose(f);
}

void save_fluid_sub_grid_cord_use_fv(GV gv, int start_x, int start_y, int start_z, int end_x, int end_y,
                                     int end_z, char fName[]) {
  /* FieldView grid file, lbm_ib.c:510-548: "nx,ny,nz", then the X, Y and Z
   * coordinates of every node over [start, end) with X fastest */
  (void)gv;
  FILE *f = fopen(fName, "wb");
  if (!f) return;
  fprintf(f, "%d,%d,%d\n", end_x - start_x, end_y - start_y, end_z - start_z);
  for (int axis = 0; axis < 3; ++axis)
    for (int Z = start_z; Z < end_z; Z++)
      for (int Y = start_y; Y < end_y; Y++)
        for (int X = start_x; X < end_x; X++) fprintf(f, "%d,", axis == 0 ? X : (axis == 1 ? Y : Z));
  fclose(f);
}

void print_fiber_sub_grid(GV gv, int start_y, int start_z, int end_y, int end_z) {
  /* sheet 0, rows start_y..end_y and columns start_z..end_z (closed), to
   * stdout, lbm_ib.c:551-572 */
  Fiber *fa = gv->fiber_shape->sheets[0].fibers;
  printf("(i,j): {cord_x, cord_y, cord_z} || SF_x,SF_y, SF_z || BF_X, BF_y, BF_z\n");
  for (int i = start_y; i <= end_y; ++i) {
    for (int j = start_z; j <= end_z; ++j) {
      Fibernode *nd = fa[i].nodes + j;
      printf("(%d, %d):{%f, %f, %f} || %.12f, %.12f, %.12f || %.12f, %.12f, %.12f\n", i, j, nd->x, nd->y, nd->z,
             nd->stretch_force_x, nd->stretch_force_y, nd->stretch_force_z, nd->bend_force_x, nd->bend_force_y,
             nd->bend_force_z);
    }
    printf("\n");
  }
}

void save_fiber_sub_grid_cord_ef_use_fv(GV gv, int start_y, int start_z, int end_y, int end_z, char grid_fname[],
                                        char elastic_force_fname[]) {
  /* FieldView grid ("rows,cols,1") and function ("rows,cols,1,3") files of
   * sheet 0: all x then y then z (resp. elastic force components) with the
   * row index fastest, lbm_ib.c:650-692 */
  FILE *gf = fopen(grid_fname, "w");
  FILE *ef = fopen(elastic_force_fname, "w");
  if (!gf || !ef) {
    if (gf) fclose(gf);
    if (ef) fclose(ef);
    return;
  }
  Fiber *fa = gv->fiber_shape->sheets[0].fibers;
  fprintf(gf, "%d,%d,1\n", end_y - start_y + 1, end_z - start_z + 1);
  fprintf(ef, "%d,%d,1,3\n", end_y - start_y + 1, end_z - start_z + 1);
  for (int comp = 0; comp < 3; ++comp)
    for (int j = start_z; j <= end_z; ++j)
      for (int i = start_y; i <= end_y; ++i) {
        Fibernode *nd = fa[i].nodes + j;
        fprintf(gf, "%f,", comp == 0 ? nd->x : (comp == 1 ? nd->y : nd->z));
        fprintf(ef, "%f,", comp == 0 ? nd->elastic_force_x : (comp == 1 ? nd->elastic_force_y : nd->elastic_force_z));
      }
  fclose(gf);
  fclose(ef);
}

/* --------------------------------------------------------------- the loop */

/* The barrier after every routine of lbm_ib.c:3124-3563, when several
 * threads run the loop; a single caller needs none. */
static void step_barrier(GV gv) {
  if (gv->total_threads > 1) pthread_barrier_wait(&gv->barr);
}

void *do_thread(void *v) {
  LV lv = (LV)v;
  GV gv = lv->gv;
  char fname[128], fname1[128];
  if ((gv->dump_mask & 4) && lv->tid == 0) {
    /* FieldView variable names and fluid grid, written once (lbm_ib.c:3100-3112) */
    FILE *nf = fopen("FV_nameFluid.nam", "w");
    if (nf) { fprintf(nf, "u\nv\nw"); fclose(nf); }
    nf = fopen("FV_nameFiber.nam", "w");
    if (nf) { fprintf(nf, "Afx\nAfy\nAfz"); fclose(nf); }
    Fluidgrid *g = gv->fluid_grid;
    snprintf(fname, sizeof fname, "FV_gridFluid.xyz");
    save_fluid_sub_grid_cord_use_fv(gv, 0, 0, 0, g->x_dim, g->y_dim, g->z_dim, fname);
  }
  step_barrier(gv);
  while (gv->time <= gv->TIME_STOP) {
    const int dumping = gv->dump_mask && gv->dump > 0 && ((gv->time % gv->dump == 0) || gv->time == 1);
    /* the collision of this step uses the equilibrium of the rho/u that the
     * previous step's compute_rho_and_u left (with the previous step's force,
     * lbm_ib.c:2521-2523, :1731ff): captured before this step's spreading */
    if (dumping && (gv->dump_mask & 2) && lv->tid == 0) capture_dfeq(gv);
    compute_bendingforce(lv);
    step_barrier(gv);
    compute_stretchingforce(lv);
    step_barrier(gv);
    compute_elasticforce(lv);
    step_barrier(gv);
    get_influentialdomain_fluid_grid_and_SpreadForce(lv);
    step_barrier(gv);
    compute_eqlbrmdistrfuncDF1(lv);
    step_barrier(gv);
    stream_distrfunc(lv);
    step_barrier(gv);
    bounceback_rigidwalls(lv);
    step_barrier(gv);
    compute_rho_and_u(lv);
    step_barrier(gv);
    moveFiberSheet(lv);
    step_barrier(gv);
    copy_inout_to_df2(lv);
    step_barrier(gv);
    replace_old_DF(lv);
    step_barrier(gv);
    /* periodicBC(lv) is not called by the reference either (lbm_ib.c:3484) */
    if (dumping && lv->tid == 0) {
      lbmib_sync_to_host(gv, (gv->dump_mask & 6) != 0);
      if (gv->dump_mask & 2) {
        snprintf(fname, sizeof fname, "Fluid%d_dump_step%d.dat", 0, gv->time);
        save_fluid_sub_grid(gv, 19, 20, 10, 22, 25, 33, fname);  /* the box of lbm_ib.c:3511 */
      }
      if (gv->dump_mask & 1) {
        snprintf(fname, sizeof fname, "Fiber%d_dump_step%d.bin", 1, gv->time);
        Fibersheet *sh = gv->fiber_shape->sheets;
        save_fiber_sub_grid_binary(gv, 0, 0, sh->num_rows - 1, sh->num_cols - 1, fname);
      }
      if (gv->dump_mask & 8) {  /* the text fiber dump, lbm_ib.c:3529-3532 (USE_CATALYST) */
        snprintf(fname, sizeof fname, "Fiber%d_dump_step%d.dat", 1, gv->time);
        Fibersheet *sh = gv->fiber_shape->sheets;
        save_fiber_sub_grid(gv, 0, 0, sh->num_rows - 1, sh->num_cols - 1, fname);
      }
      if (gv->dump_mask & 4) {  /* FieldView, lbm_ib.c:3514-3517, :3531-3535 */
        Fluidgrid *g = gv->fluid_grid;
        snprintf(fname, sizeof fname, "FV_fluid_vel_%d.f", gv->time);
        save_fluid_sub_grid_vel_use_fv(gv, 0, 0, 0, g->x_dim, g->y_dim, g->z_dim, fname);
        Fibersheet *sh = gv->fiber_shape->sheets;
        snprintf(fname, sizeof fname, "FV_gridFiber%d.xyz", gv->time);
        snprintf(fname1, sizeof fname1, "FV_fiber_elastic_force_%d.f", gv->time);
        save_fiber_sub_grid_cord_ef_use_fv(gv, 0, 0, sh->num_rows - 1, sh->num_cols - 1, fname, fname1);
      }
    }
    step_barrier(gv);
    if (lv->tid == 0 && !gv->quiet) {
      printf("End of time step %d\n", gv->time);
      fflush(stdout);
    }
    /* every thread reads the error thread 0 noted before the last barrier */
    const int stop = gv->last_error != 0;
    step_barrier(gv);
    if (lv->tid == 0) gv->time += gv->dt;
    step_barrier(gv);
    if (stop) break;
  }
  if (lv->tid == 0) lbmib_host_status(gv);
  return NULL;
}
