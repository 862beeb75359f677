/*
 * host_pthread_caller.c -- a caller written against the reference's API
 * (include/lbmib_host.h), the way cube_pthread/lbm_ib.c's main() drives its
 * routines (lbm_ib.c:3836-3955): it fills GV, builds the sheet and the grid,
 * allocates one mutex per thread in gv->lock_Fluid, initialises gv->barr for
 * P*Q*R threads, runs do_thread on every thread with its own LV and joins
 * them.  Used by tests/test_gpu_parity.py to check that a multi-threaded
 * caller reproduces the reference's binary fiber dumps.
 *
 * usage: host_pthread_caller N STEPS DUMP P Q R ROWS COLS WIDTH HEIGHT X0 Y0 Z0
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>

#include "lbmib_host.h"

int main(int argc, char **argv) {
  if (argc != 14) {
    fprintf(stderr, "usage: %s N STEPS DUMP P Q R ROWS COLS WIDTH HEIGHT X0 Y0 Z0\n", argv[0]);
    return 2;
  }
  const int n = atoi(argv[1]), steps = atoi(argv[2]), dump = atoi(argv[3]);
  const int P = atoi(argv[4]), Q = atoi(argv[5]), R = atoi(argv[6]);
  const int rows = atoi(argv[7]), cols = atoi(argv[8]);
  const double width = atof(argv[9]), height = atof(argv[10]);
  const double x0 = atof(argv[11]), y0 = atof(argv[12]), z0 = atof(argv[13]);
  const int total_threads = P * Q * R, cube = 4;

  GV gv = (GV)calloc(1, sizeof(*gv));
  gv->TIME_STOP = steps;
  gv->dump = dump;
  gv->P = P; gv->Q = Q; gv->R = R;
  gv->quiet = 1;
  gen_fiber_shape(gv, width, height, cols, rows, x0, y0, z0);
  Fluidgrid *grid = gen_fluid_grid(gv, n, n, n, cube);
  init_gv(gv, gv->fiber_shape, grid, cube);
  if (gv->last_error) { fprintf(stderr, "init_gv failed: %d\n", gv->last_error); return 1; }

  gv->lock_Fluid = (pthread_mutex_t *)malloc(sizeof(*gv->lock_Fluid) * total_threads);
  for (int i = 0; i < total_threads; ++i)
    if (pthread_mutex_init(&gv->lock_Fluid[i], NULL)) { fprintf(stderr, "mutex init\n"); return 1; }
  if (pthread_barrier_init(&gv->barr, NULL, total_threads)) { fprintf(stderr, "barrier init\n"); return 1; }
  gv->total_threads = total_threads;

  init_eqlbrmdistrfuncDF0(gv);
  init_df1(gv);
  init_df_inout(gv);
  lbmib_host_set_dumps(gv, 1);
  /* the inlet table holds the initial equilibrium, as init_df_inout leaves it */
  const Fibersheet *sh = gv->fiber_shape->sheets;
  printf("sheet %dx%d, inlet df_inout[0][0] %.17g, cubes %d, cube 0 x [%d, %d]\n", sh->num_rows, sh->num_cols,
         gv->fluid_grid->inlet->nodes[0].df_inout[0][0], gv->num_cubes_x * gv->num_cubes_y * gv->num_cubes_z,
         gv->fluid_grid->sub_fluid_grid[0].sub_coordx_start, gv->fluid_grid->sub_fluid_grid[0].sub_coordx_end);

  pthread_t *threads = (pthread_t *)malloc(sizeof(pthread_t) * total_threads);
  struct lv_t *lvs = (struct lv_t *)malloc(sizeof(struct lv_t) * total_threads);
  gv->threads = threads;
  const double t0 = get_cur_time();
  for (int i = 0; i < total_threads; ++i) {
    lvs[i].tid = i;
    lvs[i].gv = gv;
    if (pthread_create(threads + i, NULL, do_thread, lvs + i)) { perror("pthread_create"); return 1; }
  }
  for (int i = 0; i < total_threads; ++i) pthread_join(threads[i], NULL);
  const double t1 = get_cur_time();
  const int rc = lbmib_host_status(gv);
  printf("time %d, status %d\n", gv->time, rc);
  printf("TOTAL TIME TAKEN IN Seconds:%f\n", t1 - t0);

  lbmib_host_free(gv);
  free(gv->fluid_grid->sub_fluid_grid); /* NULL after lbmib_host_free */
  free(gv->fluid_grid);
  pthread_barrier_destroy(&gv->barr);
  for (int i = 0; i < total_threads; ++i) pthread_mutex_destroy(&gv->lock_Fluid[i]);
  free(gv->lock_Fluid);
  free(threads);
  free(lvs);
  free(gv);
  return rc ? 1 : 0;
}
