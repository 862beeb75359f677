/* Calls the host layer's dump writers (distributed-lb-ib_amd/host/
 * lbmib_host.c) on a small hand-made grid and sheet, with no device, for
 * tests/test_host_writers.py.  Fluid mirror values: u = X + 10 Y + 100 Z,
 * v = -u / 4, w = u / 8, rho = 1 + u / 1000; sheet node (i, j): position
 * (i + 0.25, j + 0.5, i * j), elastic force (i, -j, 0.5), stretch (1, 2, 3),
 * bend (4, 5, 6). */
#include <stdlib.h>
#include <string.h>

#include "lbmib_host.h"

int main(void) {
  Fluidgrid g;
  memset(&g, 0, sizeof g);
  g.x_dim = 5; g.y_dim = 4; g.z_dim = 3;
  const size_t n = 60;
  g.vel = (double *)calloc(3 * n, sizeof(double));
  g.rho = (double *)calloc(n, sizeof(double));
  g.force = (double *)calloc(3 * n, sizeof(double));
  g.df = (double *)calloc(19 * n, sizeof(double));
  for (int Z = 0; Z < 3; ++Z)
    for (int Y = 0; Y < 4; ++Y)
      for (int X = 0; X < 5; ++X) {
        const size_t c = X + 5 * (Y + 4 * Z);
        const double u = X + 10.0 * Y + 100.0 * Z;
        g.vel[c] = u; g.vel[n + c] = -u / 4; g.vel[2 * n + c] = u / 8;
        g.rho[c] = 1 + u / 1000;
        for (int q = 0; q < 19; ++q) g.df[q * n + c] = q + u / 1000;
      }
  Fibernode nodes[3][4];
  Fiber fibers[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 4; ++j) {
      Fibernode *nd = &nodes[i][j];
      nd->x = i + 0.25; nd->y = j + 0.5; nd->z = i * j;
      nd->elastic_force_x = i; nd->elastic_force_y = -j; nd->elastic_force_z = 0.5;
      nd->stretch_force_x = 1; nd->stretch_force_y = 2; nd->stretch_force_z = 3;
      nd->bend_force_x = 4; nd->bend_force_y = 5; nd->bend_force_z = 6;
    }
    fibers[i].nodes = nodes[i];
    fibers[i].num_nodes = 4;
  }
  Fibersheet sheet;
  memset(&sheet, 0, sizeof sheet);
  sheet.fibers = fibers; sheet.num_rows = 3; sheet.num_cols = 4;
  Fibershape shape = {&sheet, 1};
  struct gv_t gvs;
  memset(&gvs, 0, sizeof gvs);
  GV gv = &gvs;
  gv->fluid_grid = &g;
  gv->fiber_shape = &shape;
  gv->cube_size = 2;
  save_fluid_sub_grid_vel_use_fv(gv, 0, 0, 0, 5, 4, 3, "fv_vel.f");
  save_fluid_sub_grid_cord_use_fv(gv, 0, 0, 0, 5, 4, 3, "fv_grid.xyz");
  save_fluid_sub_grid_Vel(gv, 1, 1, 1, 4, 3, 3, "vel.txt");
  save_fiber_sub_grid_cord_ef_use_fv(gv, 0, 0, 2, 3, "fv_fiber.xyz", "fv_fiber_ef.f");
  print_fiber_sub_grid(gv, 0, 1, 1, 2);
  print_fluid_sub_grid(gv, 0, 0, 0, 1, 1, 1, 2);
  free(g.vel); free(g.rho); free(g.force); free(g.df);
  return 0;
}
