/* Test infrastructure only: drives the FP64 oracle (lbmib_oracle.c) through
 * whole steps on small grids, so that tests/test_sanitizers.py can run it
 * built with AddressSanitizer and UndefinedBehaviorSanitizer.  Covers every
 * boundary kind, two sheets and the fiber-leaves-the-grid error path.
 * Exit status 0 when every case behaves as expected. */
#include <stdio.h>
#include <stdlib.h>

typedef struct orc orc_t;
orc_t *orc_create(int nx, int ny, int nz, int bcx, int bcy, int bcz, double tau, double rho0, double ux0,
                  double uy0, double uz0, int quirks);
void orc_destroy(orc_t *o);
int orc_add_sheet(orc_t *o, int rows, int cols, double width, double height, double ks, double kb,
                  const double *xyz);
void orc_init(orc_t *o);
void orc_preshift(orc_t *o, int all_sheets);
int orc_step(orc_t *o, int nsteps);
int orc_error(const orc_t *o);

/* A flat rows x cols sheet, unit spacing, normal to x, centred at c. */
static double *flat_sheet(int rows, int cols, double cx, double cy, double cz) {
  double *xyz = (double *)malloc(sizeof(double) * 3 * rows * cols);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      double *p = xyz + 3 * (i * cols + j);
      p[0] = cx + 0.01 * ((i + j) % 3);
      p[1] = cy - 0.5 * (rows - 1) + i;
      p[2] = cz - 0.5 * (cols - 1) + j;
    }
  return xyz;
}

static int run(int bcx, int bcy, int bcz, int nsheets, double sheet_x, int steps, int want_error) {
  orc_t *o = orc_create(20, 18, 16, bcx, bcy, bcz, 0.8, 1.0, 0.02, 0.0, 0.0, 1);
  if (!o) return 1;
  for (int s = 0; s < nsheets; ++s) {
    double *xyz = flat_sheet(6, 5, sheet_x + 5.0 * s, 9.0, 8.0);
    if (orc_add_sheet(o, 6, 5, 5.0, 4.0, 1e-3, 1e-3, xyz) != 0) return 1;
    free(xyz);
  }
  orc_init(o);
  orc_preshift(o, 1);
  const int rc = orc_step(o, steps);
  const int err = orc_error(o);
  orc_destroy(o);
  if (want_error) return (rc != 0 || err != 0) ? 0 : 1;
  return (rc == 0 && err == 0) ? 0 : 1;
}

int main(void) {
  const int bcs[4][3] = {{0, 1, 1}, {0, 2, 1}, {2, 2, 2}, {1, 1, 1}};
  int fails = 0;
  for (int k = 0; k < 4; ++k) {
    fails += run(bcs[k][0], bcs[k][1], bcs[k][2], 0, 6.0, 4, 0);
    fails += run(bcs[k][0], bcs[k][1], bcs[k][2], 2, 6.0, 4, 0);
  }
  fails += run(0, 1, 1, 1, 0.5, 2, 1); /* support starts left of the grid */
  printf("sanitize_check: %d failing case(s)\n", fails);
  return fails ? 1 : 0;
}
