/*
 * lbmib_cli.c -- command-line driver of the MI355X LBM-IB engine.
 *
 * Accepts the reference's flags with the reference's meaning (main,
 * lbm_ib.c:3611-3955), including its x/z swap of -fluid_grid_xyz
 * (lbm_ib.c:3792-3794, the grid's x is the third number) and its sanity
 * checks (lbm_ib.c:3815-3833).  -thread_per_task_xyz starts P*Q*R threads
 * on do_thread as the reference does; thread 0 enqueues the GPU work.  Extensions (not in the reference) set the
 * physics the reference hard-codes:
 *   -tau T -u_in U -ks K -kb K -bc X Y Z (INOUT|WALL|PERIODIC)
 *   -precision f64|f32|f16|bf16  -jitter A -seed S  -no_quirks  -shift_all
 *   -dump_fiber -dump_fluid -dump_fv -dump_fiber_text  -quiet
 * (the reference's dumps are compile-time switches: DUMP_VERIFY and USE_FV,
 * lbm_ib.c:30, :3504-3538).
 * Prints "TOTAL TIME TAKEN IN Seconds:%f" like the reference, then MLUPS.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lbmib_host.h"

static void needs_argument(int i, int argc, const char *flag, int n) {
  if (i + n >= argc) {
    fprintf(stderr, "error: Flag \"%s\" requires an argument\n", flag);
    exit(2);
  }
}

static int parse_bc(const char *s) {
  if (!strcmp(s, "INOUT")) return LBMIB_BC_INOUT;
  if (!strcmp(s, "WALL")) return LBMIB_BC_WALL;
  if (!strcmp(s, "PERIODIC")) return LBMIB_BC_PERIODIC;
  fprintf(stderr, "error: unknown boundary kind \"%s\"\n", s);
  exit(2);
}

static int positive_int(const char *flag, const char *v) {
  int x = atoi(v);
  if (x <= 0) {
    fprintf(stderr, "error: Invalid flag \"%s %d\" must be > 0\n", flag, x);
    exit(2);
  }
  return x;
}

static double positive_double(const char *flag, const char *v) {
  double x = atof(v);
  if (x <= 0) {
    fprintf(stderr, "error: Invalid flag \"%s %f\" must be > 0\n", flag, x);
    exit(2);
  }
  return x;
}

int main(int argc, char *argv[]) {
  GV gv = (GV)calloc(1, sizeof(*gv));
  int fx = 0, fy = 0, fz = 0;
  double tau = 0, u_in = 0, ks = -1, kb = -1, jitter = 0;
  unsigned long long seed = 0;
  int bc[3] = {-1, -1, -1}, custom = 0, dumps = 0;
  gv->precision = LBMIB_F64;
  gv->quirks = 1;

  for (int i = 1; i < argc; i++) {
    const char *a = argv[i];
    if (!strcmp(a, "-steps")) { needs_argument(i, argc, a, 1); gv->TIME_STOP = positive_int(a, argv[++i]); }
    else if (!strcmp(a, "-dump")) { needs_argument(i, argc, a, 1); gv->dump = positive_int(a, argv[++i]); }
    else if (!strcmp(a, "-fluid_grid_xyz")) {
      needs_argument(i, argc, a, 3);
      fx = positive_int("-fluid_grid_x", argv[++i]);
      fy = positive_int("-fluid_grid_y", argv[++i]);
      fz = positive_int("-fluid_grid_z", argv[++i]);
    } else if (!strcmp(a, "-cube_size")) { needs_argument(i, argc, a, 1); gv->cube_size = positive_int(a, argv[++i]); }
    else if (!strcmp(a, "-thread_per_task_xyz")) {
      needs_argument(i, argc, a, 3);
      gv->P = positive_int("-thread_per_task_x", argv[++i]);
      gv->Q = positive_int("-thread_per_task_y", argv[++i]);
      gv->R = positive_int("-thread_per_task_z", argv[++i]);
    } else if (!strcmp(a, "-num_fibersht")) {
      needs_argument(i, argc, a, 1);
      int n = positive_int(a, argv[++i]);
      gv->fiber_shape = (Fibershape *)calloc(1, sizeof(Fibershape));
      gv->fiber_shape->num_sheets = n;
      gv->fiber_shape->sheets = (Fibersheet *)calloc(n, sizeof(Fibersheet));
    } else if (!strcmp(a, "-fibersht_width_height") || !strcmp(a, "-fibersht_row_clmn") ||
               !strcmp(a, "-fibersht_xyz_0")) {
      if (!gv->fiber_shape) { fprintf(stderr, "error: %s before -num_fibersht\n", a); return 2; }
      for (int s = 0; s < gv->fiber_shape->num_sheets; ++s) {
        Fibersheet *sh = gv->fiber_shape->sheets + s;
        if (!strcmp(a, "-fibersht_width_height")) {
          needs_argument(i, argc, a, 2);
          sh->width = positive_double("-fibersht_width", argv[++i]);
          sh->height = positive_double("-fibersht_height", argv[++i]);
        } else if (!strcmp(a, "-fibersht_row_clmn")) {
          needs_argument(i, argc, a, 2);
          sh->num_rows = positive_int("-fibersht_row", argv[++i]);
          sh->num_cols = positive_int("-fibersht_clmn", argv[++i]);
        } else {
          needs_argument(i, argc, a, 3);
          sh->x_orig = positive_double("-fibersht_x0", argv[++i]);
          sh->y_orig = positive_double("-fibersht_y0", argv[++i]);
          sh->z_orig = positive_double("-fibersht_z0", argv[++i]);
        }
      }
    }
    /* extensions */
    else if (!strcmp(a, "-tau")) { needs_argument(i, argc, a, 1); tau = atof(argv[++i]); custom = 1; }
    else if (!strcmp(a, "-u_in")) { needs_argument(i, argc, a, 1); u_in = atof(argv[++i]); custom = 1; }
    else if (!strcmp(a, "-ks")) { needs_argument(i, argc, a, 1); ks = atof(argv[++i]); custom = 1; }
    else if (!strcmp(a, "-kb")) { needs_argument(i, argc, a, 1); kb = atof(argv[++i]); custom = 1; }
    else if (!strcmp(a, "-bc")) {
      needs_argument(i, argc, a, 3);
      for (int k = 0; k < 3; ++k) bc[k] = parse_bc(argv[++i]);
      custom = 1;
    } else if (!strcmp(a, "-precision")) {
      needs_argument(i, argc, a, 1);
      const char *p = argv[++i];
      gv->precision = !strcmp(p, "f32") ? LBMIB_F32 : !strcmp(p, "f16") ? LBMIB_F16
                      : !strcmp(p, "bf16") ? LBMIB_BF16 : LBMIB_F64;
    } else if (!strcmp(a, "-jitter")) { needs_argument(i, argc, a, 1); jitter = atof(argv[++i]); }
    else if (!strcmp(a, "-seed")) { needs_argument(i, argc, a, 1); seed = strtoull(argv[++i], NULL, 10); }
    else if (!strcmp(a, "-no_quirks")) { gv->quirks = 0; custom = 1; }
    else if (!strcmp(a, "-shift_all")) { gv->shift_all_sheets = 1; }
    else if (!strcmp(a, "-dump_fiber")) dumps |= 1;
    else if (!strcmp(a, "-dump_fluid")) dumps |= 2;
    else if (!strcmp(a, "-dump_fv")) dumps |= 4;
    else if (!strcmp(a, "-dump_fiber_text")) dumps |= 8;
    else if (!strcmp(a, "-quiet")) gv->quiet = 1;
    else { fprintf(stderr, "error: unknown flag \"%s\"\n", a); return 2; }
  }
  if (!gv->fiber_shape || fx <= 0 || gv->cube_size <= 0 || gv->P <= 0 || gv->TIME_STOP <= 0) {
    fprintf(stderr, "usage: %s -steps N -dump N -fluid_grid_xyz Z Y X -cube_size K -thread_per_task_xyz P Q R "
                    "-num_fibersht S -fibersht_width_height W H -fibersht_row_clmn R C -fibersht_xyz_0 X0 Y0 Z0\n",
            argv[0]);
    return 2;
  }
  if (gv->dump <= 0) gv->dump = gv->TIME_STOP + 1;

  Fibersheet *s0 = gv->fiber_shape->sheets;
  const int fluidgrid_z = fx, fluidgrid_y = fy, fluidgrid_x = fz; /* lbm_ib.c:3792-3794 */
  const int total_threads = gv->P * gv->Q * gv->R;
  const int k = gv->cube_size;
  printf("***********IB Simulation on MI355X (LBM-IB engine) starts************\n");
  printf("elem_z:%d , elem_y :%d, elem_x: %d\n", fluidgrid_z, fluidgrid_y, fluidgrid_x);
  if (s0->num_rows % 2 != 0 || total_threads % 2 != 0) {
    fprintf(stderr, "Check #Fiber_row ==:%d and total threads ==:%d :Should be multiple of 2", s0->num_rows,
            total_threads);
    return 1;
  }
  if ((fluidgrid_z / k) % gv->R != 0) { fprintf(stderr, "Check Fluid_elem_z== :%d and  R==:%d :Should be divisible \n", fluidgrid_z, gv->R); return 1; }
  if ((fluidgrid_y / k) % gv->Q != 0) { fprintf(stderr, "Check Fluid_elem_y== :%d and  Q==:%d :Should be divisible \n", fluidgrid_y, gv->Q); return 1; }
  if ((fluidgrid_x / k) % gv->P != 0) { fprintf(stderr, "Check Fluid_elem_x== :%d and  P==:%d :Should be divisible \n", fluidgrid_x, gv->P); return 1; }
  gv->total_threads = total_threads;

  if (jitter != 0.0) {
    for (int s = 0; s < gv->fiber_shape->num_sheets; ++s) {
      Fibersheet *sh = gv->fiber_shape->sheets + s;
      gen_fiber_sheet_jittered(gv, s, sh->width, sh->height, sh->num_cols, sh->num_rows, sh->x_orig, sh->y_orig,
                               sh->z_orig, jitter, seed + (unsigned long long)s);
    }
  } else {
    gen_fiber_shape(gv, s0->width, s0->height, s0->num_cols, s0->num_rows, s0->x_orig, s0->y_orig, s0->z_orig);
  }
  Fluidgrid *grid = gen_fluid_grid(gv, fluidgrid_x, fluidgrid_y, fluidgrid_z, k);

  if (custom) {
    double ref_tau, ref_u, ref_ks, ref_kb;
    { /* the reference values for anything not given */
      double u = 0.001, L = 20.0, Re = 150.0;
      ref_u = u;
      ref_tau = 0.5 * (6.0 * (u * L / Re) + 1.0);
      ref_kb = 0.005 * 1.0 * u * u * L * L * L * L;
      ref_ks = 20.0 * 1.0 * u * u * L;
    }
    gv->custom_physics = 1;
    gv->tau = tau > 0 ? tau : ref_tau;
    gv->u_l = u_in != 0 ? u_in : ref_u;
    gv->rho_l = 1.0;
    gv->Ks_l = ks >= 0 ? ks : ref_ks;
    gv->Kb_l = kb >= 0 ? kb : ref_kb;
    gv->g_l = 0.0;
    gv->bc_x = bc[0] >= 0 ? bc[0] : LBMIB_BC_INOUT;
    gv->bc_y = bc[1] >= 0 ? bc[1] : LBMIB_BC_WALL;
    gv->bc_z = bc[2] >= 0 ? bc[2] : LBMIB_BC_WALL;
  }
  init_gv(gv, gv->fiber_shape, grid, k);
  if (gv->last_error) { fprintf(stderr, "init failed: %d\n", gv->last_error); return 1; }
  init_eqlbrmdistrfuncDF0(gv);
  init_df1(gv);
  init_df_inout(gv);
  lbmib_host_set_dumps(gv, dumps);

  /* P*Q*R threads run do_thread with a barrier after every routine, as the
   * reference's main() starts them (lbm_ib.c:3844-3867, :3910-3928); thread 0
   * enqueues the device work */
  gv->lock_Fluid = (pthread_mutex_t *)malloc(sizeof(*gv->lock_Fluid) * total_threads);
  for (int i = 0; i < total_threads; ++i)
    if (pthread_mutex_init(&gv->lock_Fluid[i], NULL)) { fprintf(stderr, "Unable to initialize a mutex\n"); return 1; }
  if (pthread_barrier_init(&gv->barr, NULL, total_threads)) { fprintf(stderr, "Could not create a barrier\n"); return 1; }
  pthread_t *threads = (pthread_t *)malloc(sizeof(pthread_t) * total_threads);
  struct lv_t *lvs = (struct lv_t *)malloc(sizeof(struct lv_t) * total_threads);
  gv->threads = threads;
  lbmib_sync_to_host(gv, 0);
  double t0 = get_cur_time();
  for (int i = 0; i < total_threads; ++i) {
    lvs[i].tid = i;
    lvs[i].gv = gv;
    if (pthread_create(threads + i, NULL, do_thread, lvs + i)) {
      perror("Thread creation do_thread\n");
      return 1;
    }
  }
  for (int i = 0; i < total_threads; ++i) pthread_join(threads[i], NULL);
  int rc = lbmib_host_status(gv);
  double t1 = get_cur_time();
  pthread_barrier_destroy(&gv->barr);
  for (int i = 0; i < total_threads; ++i) pthread_mutex_destroy(&gv->lock_Fluid[i]);
  free(gv->lock_Fluid);
  gv->lock_Fluid = NULL;
  free(threads);
  free(lvs);
  gv->threads = NULL;
  printf("***********IB Simulation on MI355X (LBM-IB engine) ends************\n\n\n");
  if (rc) {
    fprintf(stderr, "error: device status %d\n", rc);
    lbmib_host_free(gv);
    return 1;
  }
  printf("TOTAL TIME TAKEN IN Seconds:%f\n", t1 - t0);
  const double cells = (double)fluidgrid_x * fluidgrid_y * fluidgrid_z;
  printf("MLUPS:%f\n", cells * gv->TIME_STOP / (t1 - t0) / 1e6);
  lbmib_host_free(gv);
  return 0;
}
