/*
 * lbmib_oracle.c -- sequential FP64 CPU restatement of the LBM-IB time step.
 *
 * THIS FILE IS TEST INFRASTRUCTURE.  It is the oracle that the HIP product
 * path is checked against: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  No product code links or calls it.
 *
 * It plays the role of the "cube_sequential" code that the reference README
 * lists (README.md:3-4) but does not ship: a single-threaded restatement of
 * cube_pthread/lbm_ib.c with the same operation order, so that in reference
 * mode it reproduces the compiled reference bit for bit (gcc, no FMA
 * contraction, glibc libm).  Every routine cites the reference lines it
 * follows.  The grid is stored as plain x-fastest arrays instead of the
 * reference's 824-byte AoS cubes (lbm_ib.c:89-97); the cube blocking only
 * changes memory order, never arithmetic.
 *
 * Generalisations beyond the reference (needed by the benchmark workloads,
 * which the unmodified reference cannot express):
 *   - per-axis boundary kind: x in {INOUT, WALL, PERIODIC}, y/z in
 *     {WALL, PERIODIC}.  INOUT/WALL reproduce lbm_ib.c exactly; PERIODIC is a
 *     real wrap over the whole axis (the reference periodicBC, lbm_ib.c:2970,
 *     is never called and writes only ghost planes);
 *   - any number of sheets (the reference only touches sheets[0]);
 *   - the two force quirks of the reference are switchable (quirks=1 keeps
 *     them).
 *
 * Pinning (what checks each function, so a dropped term / wrong sign fails):
 *   orc_init_feq / collide   : Sum feq = rho, Sum c feq = rho u; Guo term
 *                              moments; fixed point f=feq,F=0; fixtures.
 *   stream / bounceback      : brute-force permutation test on 8^3, fixtures.
 *   rho_u                    : numpy moments of random f, fixtures.
 *   fiber forces             : hand stencils on 6x6, zero on flat sheet,
 *                              Newton's 3rd law (quirks off), fixtures.
 *   spread / move            : partition of unity, force conservation,
 *                              uniform-u interpolation, fixtures.
 *   step order (lagged u, frozen planes, outlet table, pre-shift): fixtures
 *                              from the compiled reference (tests/golden/).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979 /* truncated constant, lbm_ib.c:33 */

enum { ORC_BC_INOUT = 0, ORC_BC_WALL = 1, ORC_BC_PERIODIC = 2 };
enum { ORC_OK = 0, ORC_ERR_FIBER_OOB = 1, ORC_ERR_ALLOC = 2, ORC_ERR_ARG = 3 };

/* D3Q19 velocity table, lbm_ib.c:777-833 (same order as mpi/init_gv.c:25-35) */
static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CY[19] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, -1, 1, 1, -1};
static const int CZ[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0};
/* weight denominators: 1/3, 1/18 (i=1..6), 1/36 (i=7..18), lbm_ib.c:1731-1818 */
static const double WDEN[19] = {3.0, 18.0, 18.0, 18.0, 18.0, 18.0, 18.0, 36.0, 36.0, 36.0,
                                36.0, 36.0, 36.0, 36.0, 36.0, 36.0, 36.0, 36.0, 36.0};

/* Fibernode layout of lbm_ib.c:54-59 (12 doubles, 96 B) */
enum { FN_X = 0, FN_Y, FN_Z, FN_BX, FN_BY, FN_BZ, FN_SX, FN_SY, FN_SZ, FN_EX, FN_EY, FN_EZ, FN_N };

typedef struct {
  int rows, cols;        /* rows along y, cols along z (lbm_ib.c:895-906) */
  double width, height;  /* ds1 = width/(cols-1), ds2 = height/(rows-1) */
  double ks, kb;         /* Ks_l, Kb_l */
  double *node;          /* rows*cols*FN_N */
} orc_sheet;

typedef struct {
  int nx, ny, nz;
  long n;
  int bc[3];
  double tau, rho0, u0[3], g;
  int dt;
  int quirks;
  double cs, cs2, cs4;
  double *df1, *df2, *dfeq; /* [19][n] */
  double *rho;              /* [n] */
  double *vel;              /* [3][n] */
  double *force;            /* [3][n] */
  double inlet[19], outlet[19];
  int nsheets;
  orc_sheet *sheets;
  int time;
  int err;
} orc_t;

static inline long IDX(const orc_t *o, int x, int y, int z) {
  return (long)x + (long)o->nx * ((long)y + (long)o->ny * (long)z);
}

static int dimof(const orc_t *o, int a) { return a == 0 ? o->nx : (a == 1 ? o->ny : o->nz); }

/* Index ranges, lbm_ib.c:743-748 and Appendix A.3 of the reference layout:
 * active (collide/stream source) [2,N-3]; moments [3,N-4] along an inlet/
 * outlet axis, [2,N-3] along a wall axis; swap [1,N-2].  Periodic: whole axis. */
static void active_range(const orc_t *o, int a, int *lo, int *hi) {
  int n = dimof(o, a);
  if (o->bc[a] == ORC_BC_PERIODIC) { *lo = 0; *hi = n - 1; }
  else { *lo = 2; *hi = n - 3; }
}
static void moment_range(const orc_t *o, int a, int *lo, int *hi) {
  int n = dimof(o, a);
  if (o->bc[a] == ORC_BC_PERIODIC) { *lo = 0; *hi = n - 1; }
  else if (o->bc[a] == ORC_BC_INOUT) { *lo = 3; *hi = n - 4; }
  else { *lo = 2; *hi = n - 3; }
}
static void swap_range(const orc_t *o, int a, int *lo, int *hi) {
  int n = dimof(o, a);
  if (o->bc[a] == ORC_BC_PERIODIC) { *lo = 0; *hi = n - 1; }
  else { *lo = 1; *hi = n - 2; }
}

/* Equilibrium, lbm_ib.c:1561-1594 (init) == lbm_ib.c:1732-1737/1757-1768/1788-1799
 * (collision): same expression and evaluation order. */
static double feq_q(int q, double rho, double vx, double vy, double vz) {
  double usq = vx * vx + vy * vy + vz * vz;
  if (q == 0) return 1.0 / 3.0 * rho * (1.0 - 1.5 * usq);
  double cu = CX[q] * vx + CY[q] * vy + CZ[q] * vz;
  double w = (q <= 6) ? 1.0 / 18.0 : 1.0 / 36.0;
  return w * rho * (1.0 + 3.0 * cu + 4.5 * (pow(cu, 2)) - 1.5 * usq);
}

/* ------------------------------------------------------------------ setup */

orc_t *orc_create(int nx, int ny, int nz, int bcx, int bcy, int bcz, double tau, double rho0,
                  double ux0, double uy0, double uz0, int quirks) {
  if (nx < 6 || ny < 6 || nz < 6) return NULL;
  if (bcx < 0 || bcx > 2 || bcy < 1 || bcy > 2 || bcz < 1 || bcz > 2) return NULL;
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  if (!o) return NULL;
  o->nx = nx; o->ny = ny; o->nz = nz;
  o->n = (long)nx * ny * nz;
  o->bc[0] = bcx; o->bc[1] = bcy; o->bc[2] = bcz;
  o->tau = tau; o->rho0 = rho0;
  o->u0[0] = ux0; o->u0[1] = uy0; o->u0[2] = uz0;
  o->g = 0.0;      /* lbm_ib.c:769 */
  o->dt = 1;       /* lbm_ib.c:852 */
  o->quirks = quirks;
  o->cs = 1.0 / (sqrt(3)); /* lbm_ib.c:765 */
  o->cs2 = o->cs * o->cs;
  o->cs4 = pow(o->cs, 4);
  o->df1 = (double *)calloc((size_t)19 * o->n, sizeof(double));
  o->df2 = (double *)calloc((size_t)19 * o->n, sizeof(double));
  o->dfeq = (double *)calloc((size_t)19 * o->n, sizeof(double));
  o->rho = (double *)calloc((size_t)o->n, sizeof(double));
  o->vel = (double *)calloc((size_t)3 * o->n, sizeof(double));
  o->force = (double *)calloc((size_t)3 * o->n, sizeof(double));
  if (!o->df1 || !o->df2 || !o->dfeq || !o->rho || !o->vel || !o->force) { o->err = ORC_ERR_ALLOC; }
  o->time = 1; /* lbm_ib.c:853 */
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  for (int s = 0; s < o->nsheets; ++s) free(o->sheets[s].node);
  free(o->sheets);
  free(o->df1); free(o->df2); free(o->dfeq);
  free(o->rho); free(o->vel); free(o->force);
  free(o);
}

/* Adds one sheet; xyz holds rows*cols*3 positions in (row, col, xyz) order. */
int orc_add_sheet(orc_t *o, int rows, int cols, double width, double height, double ks, double kb,
                  const double *xyz) {
  if (rows < 4 || cols < 4) return ORC_ERR_ARG;
  orc_sheet *ns = (orc_sheet *)realloc(o->sheets, sizeof(orc_sheet) * (o->nsheets + 1));
  if (!ns) return ORC_ERR_ALLOC;
  o->sheets = ns;
  orc_sheet *sh = &o->sheets[o->nsheets];
  sh->rows = rows; sh->cols = cols; sh->width = width; sh->height = height;
  sh->ks = ks; sh->kb = kb;
  sh->node = (double *)calloc((size_t)rows * cols * FN_N, sizeof(double));
  if (!sh->node) return ORC_ERR_ALLOC;
  for (long k = 0; k < (long)rows * cols; ++k) {
    sh->node[k * FN_N + FN_X] = xyz[3 * k + 0];
    sh->node[k * FN_N + FN_Y] = xyz[3 * k + 1];
    sh->node[k * FN_N + FN_Z] = xyz[3 * k + 2];
  }
  o->nsheets++;
  return ORC_OK;
}

/* init_gv density/velocity, lbm_ib.c:837-849; init_eqlbrmdistrfuncDF0,
 * lbm_ib.c:1528-1606; init_df1 (df1 = dfeq on [2,N-3]^3), lbm_ib.c:2725-2767;
 * init_df_inout (inlet = dfeq at x=ib, outlet = dfeq at x=N-1),
 * lbm_ib.c:1610-1670. */
void orc_init(orc_t *o) {
  long n = o->n;
  for (long c = 0; c < n; ++c) {
    o->rho[c] = o->rho0;
    o->vel[c] = o->u0[0];
    o->vel[n + c] = o->u0[1];
    o->vel[2 * n + c] = o->u0[2];
  }
  for (int q = 0; q < 19; ++q) {
    double v = feq_q(q, o->rho0, o->u0[0], o->u0[1], o->u0[2]);
    for (long c = 0; c < n; ++c) o->dfeq[q * n + c] = v;
  }
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) active_range(o, a, &lo[a], &hi[a]);
  memset(o->df1, 0, sizeof(double) * 19 * n);
  memset(o->df2, 0, sizeof(double) * 19 * n);
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        long c = IDX(o, x, y, z);
        for (int q = 0; q < 19; ++q) o->df1[q * n + c] = o->dfeq[q * n + c];
      }
  /* tables: dfeq is uniform at init, so the sampled planes are equal */
  long cin = IDX(o, 2, 2, 2), cout = IDX(o, o->nx - 1, 2, 2);
  for (int q = 0; q < 19; ++q) {
    o->inlet[q] = o->dfeq[q * n + cin];
    o->outlet[q] = o->dfeq[q * n + cout];
  }
  memset(o->force, 0, sizeof(double) * 3 * n);
  o->time = 1;
  o->err = ORC_OK;
}

/* init_gv's one-step sheet shift X += u_l*dt, lbm_ib.c:857-862.  The reference
 * shifts sheets[0] only; all_sheets=1 shifts every sheet. */
void orc_preshift(orc_t *o, int all_sheets) {
  int ns = all_sheets ? o->nsheets : (o->nsheets > 0 ? 1 : 0);
  for (int s = 0; s < ns; ++s) {
    orc_sheet *sh = &o->sheets[s];
    for (long k = 0; k < (long)sh->rows * sh->cols; ++k) sh->node[k * FN_N + FN_X] += o->u0[0] * o->dt;
  }
}

/* ---------------------------------------------------------- fiber forces */

#define NODE(sh, i, j) ((sh)->node + ((long)(i) * (sh)->cols + (j)) * FN_N)

/* compute_bendingforce, lbm_ib.c:961-1152.  Row pass '=' then column pass
 * '+=' with k = Kb/ds2^3 in both passes (lbm_ib.c:996, :1065).  quirks=1
 * keeps the row-pass x-component slip at j=cols-2 (lbm_ib.c:1031-1035). */
static void bending_sheet(orc_t *o, orc_sheet *sh) {
  int R = sh->rows, C = sh->cols;
  double ds2 = sh->height / (R - 1);
  double bc = sh->kb / pow(ds2, 3);
  for (int i = 0; i < R; ++i) {
    for (int d = 0; d < 3; ++d) {
      double *n0 = NODE(sh, i, 0), *n1 = NODE(sh, i, 1), *n2 = NODE(sh, i, 2), *n3 = NODE(sh, i, 3);
      n0[FN_BX + d] = bc * (-n2[d] + 2 * n1[d] - n0[d]);
      n1[FN_BX + d] = bc * (-n3[d] + 4 * n2[d] - 5 * n1[d] + 2 * n0[d]);
      for (int j = 2; j < C - 2; ++j) {
        double *p2 = NODE(sh, i, j + 2), *p1 = NODE(sh, i, j + 1), *p0 = NODE(sh, i, j);
        double *m1 = NODE(sh, i, j - 1), *m2 = NODE(sh, i, j - 2);
        p0[FN_BX + d] = bc * (-p2[d] + 4.0 * p1[d] - 6.0 * p0[d] + 4.0 * m1[d] - m2[d]);
      }
      int j = C - 2;
      double *p1 = NODE(sh, i, j + 1), *p0 = NODE(sh, i, j), *m1 = NODE(sh, i, j - 1), *m2 = NODE(sh, i, j - 2);
      double last = (o->quirks && d == 0) ? p0[d] : m2[d];
      p0[FN_BX + d] = bc * (2 * p1[d] - 5 * p0[d] + 4 * m1[d] - last);
      j = C - 1;
      double *q0 = NODE(sh, i, j), *q1 = NODE(sh, i, j - 1), *q2 = NODE(sh, i, j - 2);
      q0[FN_BX + d] = bc * (-q0[d] + 2 * q1[d] - q2[d]);
    }
  }
  bc = sh->kb / pow(ds2, 3);
  for (int j = 0; j < C; ++j) {
    for (int d = 0; d < 3; ++d) {
      double *f0 = NODE(sh, 0, j), *f1 = NODE(sh, 1, j), *f2 = NODE(sh, 2, j), *f3 = NODE(sh, 3, j);
      f0[FN_BX + d] += bc * (-f2[d] + 2 * (f1[d]) - (f0[d]));
      f1[FN_BX + d] += bc * (-f3[d] + 4 * (f2[d]) - 5 * (f1[d]) + 2 * (f0[d]));
      for (int i = 2; i < R - 2; ++i) {
        double *p2 = NODE(sh, i + 2, j), *p1 = NODE(sh, i + 1, j), *p0 = NODE(sh, i, j);
        double *m1 = NODE(sh, i - 1, j), *m2 = NODE(sh, i - 2, j);
        p0[FN_BX + d] += bc * (-p2[d] + 4.0 * (p1[d]) - 6.0 * (p0[d]) + 4.0 * (m1[d]) - (m2[d]));
      }
      int i = R - 2;
      double *p1 = NODE(sh, i + 1, j), *p0 = NODE(sh, i, j), *m1 = NODE(sh, i - 1, j), *m2 = NODE(sh, i - 2, j);
      p0[FN_BX + d] += bc * (2 * p1[d] - 5 * (p0[d]) + 4 * (m1[d]) - (m2[d]));
      i = R - 1;
      double *q0 = NODE(sh, i, j), *q1 = NODE(sh, i - 1, j), *q2 = NODE(sh, i - 2, j);
      q0[FN_BX + d] += bc * (-(q0[d]) + 2 * (q1[d]) - (q2[d]));
    }
  }
}

/* compute_stretchingforce, lbm_ib.c:1154-1282.  k = Ks/ds1^2 along rows,
 * Ks/ds2^2 along columns.  quirks=1 keeps the reference's tangent at the last
 * node of a row / column, which points away from its neighbour
 * (lbm_ib.c:1215-1217, :1269-1271); quirks=0 points it at the neighbour. */
static void stretch_pair(const double *a, const double *b, double ds, double *t, double *dist) {
  /* t = b - a */
  t[0] = b[0] - a[0];
  t[1] = b[1] - a[1];
  t[2] = b[2] - a[2];
  *dist = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  (void)ds;
}

static void stretching_sheet(orc_t *o, orc_sheet *sh) {
  int R = sh->rows, C = sh->cols;
  double ds1 = sh->width / (C - 1);
  double ds2 = sh->height / (R - 1);
  double sc = sh->ks / (ds1 * ds1);
  double tr[3], tl[3], dr, dl;
  for (int i = 0; i < R; ++i) {
    for (int j = 1; j < C - 1; ++j) {
      double *p = NODE(sh, i, j);
      stretch_pair(p, NODE(sh, i, j + 1), ds1, tr, &dr);
      stretch_pair(p, NODE(sh, i, j - 1), ds1, tl, &dl);
      for (int d = 0; d < 3; ++d)
        p[FN_SX + d] = sc * ((dr - ds1) * (tr[d] / dr) + (dl - ds1) * (tl[d] / dl));
    }
    double *p = NODE(sh, i, 0);
    stretch_pair(p, NODE(sh, i, 1), ds1, tr, &dr);
    for (int d = 0; d < 3; ++d) p[FN_SX + d] = sc * ((dr - ds1) * (tr[d] / dr));
    p = NODE(sh, i, C - 1);
    if (o->quirks) stretch_pair(NODE(sh, i, C - 2), p, ds1, tl, &dl);
    else stretch_pair(p, NODE(sh, i, C - 2), ds1, tl, &dl);
    for (int d = 0; d < 3; ++d) p[FN_SX + d] = sc * ((dl - ds1) * (tl[d] / dl));
  }
  sc = sh->ks / (ds2 * ds2);
  for (int j = 0; j < C; ++j) {
    for (int i = 1; i < R - 1; ++i) {
      double *p = NODE(sh, i, j);
      stretch_pair(p, NODE(sh, i + 1, j), ds2, tr, &dr);
      stretch_pair(p, NODE(sh, i - 1, j), ds2, tl, &dl);
      for (int d = 0; d < 3; ++d)
        p[FN_SX + d] += sc * ((dr - ds2) * (tr[d] / dr) + (dl - ds2) * (tl[d] / dl));
    }
    double *p = NODE(sh, 0, j);
    stretch_pair(p, NODE(sh, 1, j), ds2, tr, &dr);
    for (int d = 0; d < 3; ++d) p[FN_SX + d] += sc * ((dr - ds2) * (tr[d] / dr));
    p = NODE(sh, R - 1, j);
    if (o->quirks) stretch_pair(NODE(sh, R - 2, j), p, ds2, tl, &dl);
    else stretch_pair(p, NODE(sh, R - 2, j), ds2, tl, &dl);
    for (int d = 0; d < 3; ++d) p[FN_SX + d] += sc * ((dl - ds2) * (tl[d] / dl));
  }
}

/* compute_elasticforce, lbm_ib.c:1285-1318 */
static void elastic_sheet(orc_sheet *sh) {
  for (long k = 0; k < (long)sh->rows * sh->cols; ++k) {
    double *p = sh->node + k * FN_N;
    for (int d = 0; d < 3; ++d) p[FN_EX + d] = p[FN_SX + d] + p[FN_BX + d];
  }
}

void orc_bending(orc_t *o) { for (int s = 0; s < o->nsheets; ++s) bending_sheet(o, &o->sheets[s]); }
void orc_stretching(orc_t *o) { for (int s = 0; s < o->nsheets; ++s) stretching_sheet(o, &o->sheets[s]); }
void orc_elastic(orc_t *o) { for (int s = 0; s < o->nsheets; ++s) elastic_sheet(&o->sheets[s]); }

/* --------------------------------------------------------------- IB coupling */

static int in_grid(const orc_t *o, int x, int y, int z) {
  return x >= 0 && x < o->nx && y >= 0 && y < o->ny && z >= 0 && z < o->nz;
}

/* get_influentialdomain_fluid_grid_and_SpreadForce, lbm_ib.c:1323-1525:
 * zero the force field, then for every fiber node (sheet, row, col order)
 * add F * cf (1+cos(pi rx/2))(1+cos(pi ry/2))(1+cos(pi rz/2)) to the 4x4x4
 * nodes starting at floor(X-2)+1.  Out-of-grid support is an error
 * (lbm_ib.c:1458-1466). */
void orc_spread(orc_t *o) {
  const double dx = 1.0, dy = 1.0, dz = 1.0;
  const double c_x = ORC_PI / (2.0 * dx), c_y = ORC_PI / (2.0 * dy), c_z = ORC_PI / (2.0 * dz);
  const double cf = 1.0 / (64.0 * dx * dy * dz);
  long n = o->n;
  memset(o->force, 0, sizeof(double) * 3 * n);
  for (int s = 0; s < o->nsheets; ++s) {
    orc_sheet *sh = &o->sheets[s];
    for (int i = 0; i < sh->rows; ++i)
      for (int j = 0; j < sh->cols; ++j) {
        double *p = NODE(sh, i, j);
        int istart = floor(p[FN_X] / dx - 2) + 1, istop = istart + 3;
        int jstart = floor(p[FN_Y] / dy - 2) + 1, jstop = jstart + 3;
        int kstart = floor(p[FN_Z] / dz - 2) + 1, kstop = kstart + 3;
        for (int ii = istart; ii <= istop; ii++)
          for (int jj = jstart; jj <= jstop; jj++)
            for (int kk = kstart; kk <= kstop; kk++) {
              double rx = dx * ii - p[FN_X];
              double ry = dy * jj - p[FN_Y];
              double rz = dz * kk - p[FN_Z];
              double tmp = cf * (1.0e0 + cos(c_x * rx)) * (1.0e0 + cos(c_y * ry)) * (1.0e0 + cos(c_z * rz));
              if (!in_grid(o, ii, jj, kk)) { o->err = ORC_ERR_FIBER_OOB; return; }
              long c = IDX(o, ii, jj, kk);
              double ex = p[FN_EX] * tmp, ey = p[FN_EY] * tmp, ez = p[FN_EZ] * tmp;
              o->force[c] += ex;
              o->force[n + c] += ey;
              o->force[2 * n + c] += ez;
            }
      }
  }
}

/* moveFiberSheet, lbm_ib.c:2540-2723: u(X) = Sum u(node) delta over the same
 * 4x4x4 support (product order vel*(1+cos)*(1+cos)*(1+cos)*c_64,
 * lbm_ib.c:2670-2673), then X += dt*u(X) (lbm_ib.c:2697-2699). */
void orc_move(orc_t *o) {
  const double dx = 1.0, dy = 1.0, dz = 1.0;
  const double dx1 = 1.0 / dx, dy1 = 1.0 / dy, dz1 = 1.0 / dz;
  const double c_x = ORC_PI / (2.0 * dx), c_y = ORC_PI / (2.0 * dy), c_z = ORC_PI / (2.0 * dz);
  const double c_64 = 1.0 / (6.4e1);
  long n = o->n;
  for (int s = 0; s < o->nsheets; ++s) {
    orc_sheet *sh = &o->sheets[s];
    for (int jf = 0; jf < sh->rows; jf++)
      for (int kf = 0; kf < sh->cols; kf++) {
        double *p = NODE(sh, jf, kf);
        double s1 = 0, s2 = 0, s3 = 0;
        int istart = floor(p[FN_X] * dx1 - 2) + 1;
        int jstart = floor(p[FN_Y] * dy1 - 2) + 1;
        int kstart = floor(p[FN_Z] * dz1 - 2) + 1;
        for (int i2 = 0; i2 <= 3; i2++) {
          int gi = istart + i2;
          double rx = dx * gi - p[FN_X];
          for (int j2 = 0; j2 <= 3; j2++) {
            int gj = jstart + j2;
            double ry = dy * gj - p[FN_Y];
            for (int k2 = 0; k2 <= 3; k2++) {
              int gk = kstart + k2;
              double rz = dz * gk - p[FN_Z];
              if (!in_grid(o, gi, gj, gk)) { o->err = ORC_ERR_FIBER_OOB; return; }
              long c = IDX(o, gi, gj, gk);
              s1 += o->vel[c] * (1.0 + cos(c_x * rx)) * (1.0 + cos(c_y * ry)) * (1.0 + cos(c_z * rz)) * c_64;
              s2 += o->vel[n + c] * (1.0 + cos(c_x * rx)) * (1.0 + cos(c_y * ry)) * (1.0 + cos(c_z * rz)) * c_64;
              s3 += o->vel[2 * n + c] * (1.0 + cos(c_x * rx)) * (1.0 + cos(c_y * ry)) * (1.0 + cos(c_z * rz)) * c_64;
            }
          }
        }
        p[FN_X] += o->dt * s1;
        p[FN_Y] += o->dt * s2;
        p[FN_Z] += o->dt * s3;
      }
  }
}

/* ------------------------------------------------------------------- LBM */

/* compute_eqlbrmdistrfuncDF1, lbm_ib.c:1673-1825: BGK with Guo forcing over
 * the active box, using the stored (lagged) rho/vel and the current force. */
void orc_collide(orc_t *o) {
  long n = o->n;
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) active_range(o, a, &lo[a], &hi[a]);
  const double omc = 1.0 - 1.0 / o->tau, om = 1.0 / o->tau;
  double pref[19];
  for (int q = 0; q < 19; ++q) pref[q] = o->dt * (1.0 - 1.0 / (2.0 * o->tau)) / WDEN[q];
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        long c = IDX(o, x, y, z);
        double rho = o->rho[c], vx = o->vel[c], vy = o->vel[n + c], vz = o->vel[2 * n + c];
        double fx = o->force[c], fy = o->force[n + c], fz = o->force[2 * n + c];
        for (int q = 0; q < 19; ++q) {
          double eq = feq_q(q, rho, vx, vy, vz);
          o->dfeq[q * n + c] = eq;
          double a = (CX[q] - vx) * fx + (CY[q] - vy) * fy + (CZ[q] - vz) * (fz + rho * o->g);
          double cu = CX[q] * vx + CY[q] * vy + CZ[q] * vz;
          double cf = CX[q] * fx + CY[q] * fy + CZ[q] * (fz + rho * o->g);
          o->df1[q * n + c] = o->df1[q * n + c] * omc + om * eq + pref[q] * (a / o->cs2 + cu / o->cs4 * cf);
        }
      }
}

/* stream_distrfunc, lbm_ib.c:1827-2254: push df2[x+c] = df1[x] from every
 * active source (periodic axes wrap). */
void orc_stream(orc_t *o) {
  long n = o->n;
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) active_range(o, a, &lo[a], &hi[a]);
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        long c = IDX(o, x, y, z);
        for (int q = 0; q < 19; ++q) {
          int dx = x + CX[q], dy = y + CY[q], dz = z + CZ[q];
          if (o->bc[0] == ORC_BC_PERIODIC) dx = (dx + o->nx) % o->nx;
          if (o->bc[1] == ORC_BC_PERIODIC) dy = (dy + o->ny) % o->ny;
          if (o->bc[2] == ORC_BC_PERIODIC) dz = (dz + o->nz) % o->nz;
          o->df2[q * n + IDX(o, dx, dy, dz)] = o->df1[q * n + c];
        }
      }
}

/* bounceback_rigidwalls, lbm_ib.c:2257-2437: on each wall plane, every
 * population whose source lies beyond the wall takes the post-collision
 * opposite population of the same node.  Order: z low, z high, y low,
 * y high (reference), then x low/high (wall x, not in the reference).
 * In-plane ranges are the moment ranges of the other axes (x in [3,N-4],
 * wall axes [2,N-3], lbm_ib.c:2301-2305). */
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

static void bb_plane(orc_t *o, int axis, int plane, int sign) {
  long n = o->n;
  int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
  int lo1, hi1, lo2, hi2;
  moment_range(o, a1, &lo1, &hi1);
  moment_range(o, a2, &lo2, &hi2);
  for (int u = lo1; u <= hi1; ++u)
    for (int v = lo2; v <= hi2; ++v) {
      int xyz[3];
      xyz[axis] = plane; xyz[a1] = u; xyz[a2] = v;
      long c = IDX(o, xyz[0], xyz[1], xyz[2]);
      for (int q = 0; q < 19; ++q) {
        int cq = axis == 0 ? CX[q] : (axis == 1 ? CY[q] : CZ[q]);
        if (cq == sign) o->df2[q * n + c] = o->df1[OPP[q] * n + c];
      }
    }
}

void orc_bounceback(orc_t *o) {
  if (o->bc[2] == ORC_BC_WALL) { bb_plane(o, 2, 2, +1); bb_plane(o, 2, o->nz - 3, -1); }
  if (o->bc[1] == ORC_BC_WALL) { bb_plane(o, 1, 2, +1); bb_plane(o, 1, o->ny - 3, -1); }
  if (o->bc[0] == ORC_BC_WALL) { bb_plane(o, 0, 2, +1); bb_plane(o, 0, o->nx - 3, -1); }
}

/* compute_rho_and_u, lbm_ib.c:2440-2538: rho = Sum df2, u = (Sum c df2 +
 * dt/2 F)/rho over the moment box; elsewhere rho/vel keep their values. */
void orc_rho_u(orc_t *o) {
  long n = o->n;
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) moment_range(o, a, &lo[a], &hi[a]);
  for (int z = lo[2]; z <= hi[2]; ++z)
    for (int y = lo[1]; y <= hi[1]; ++y)
      for (int x = lo[0]; x <= hi[0]; ++x) {
        long c = IDX(o, x, y, z);
        double s1 = o->df2[c];
        double s2 = CX[0] * o->df2[c], s3 = CY[0] * o->df2[c], s4 = CZ[0] * o->df2[c];
        for (int q = 1; q <= 18; ++q) {
          double f = o->df2[q * n + c];
          s1 += f;
          s2 += CX[q] * f;
          s3 += CY[q] * f;
          s4 += CZ[q] * f;
        }
        o->rho[c] = s1;
        o->vel[c] = (s2 + 0.5 * o->dt * o->force[c]) / s1;
        o->vel[n + c] = (s3 + 0.5 * o->dt * o->force[n + c]) / s1;
        o->vel[2 * n + c] = (s4 + 0.5 * o->dt * (o->force[2 * n + c] + o->rho[c] * o->g)) / s1;
      }
}

/* copy_inout_to_df2, lbm_ib.c:2769-2907: x=1,2 <- inlet, x=N-3,N-2 <- outlet
 * for y,z in the swap ranges. */
void orc_inout(orc_t *o) {
  if (o->bc[0] != ORC_BC_INOUT) return;
  long n = o->n;
  int ylo, yhi, zlo, zhi;
  swap_range(o, 1, &ylo, &yhi);
  swap_range(o, 2, &zlo, &zhi);
  const int xs[4] = {2, 1, o->nx - 3, o->nx - 2};
  for (int k = 0; k < 4; ++k) {
    const double *tab = k < 2 ? o->inlet : o->outlet;
    for (int z = zlo; z <= zhi; ++z)
      for (int y = ylo; y <= yhi; ++y) {
        long c = IDX(o, xs[k], y, z);
        for (int q = 0; q < 19; ++q) o->df2[q * n + c] = tab[q];
      }
  }
}

/* replace_old_DF, lbm_ib.c:2909-2968: df1 <- df2 over the swap box. */
void orc_swap(orc_t *o) {
  long n = o->n;
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) swap_range(o, a, &lo[a], &hi[a]);
  for (int q = 0; q < 19; ++q)
    for (int z = lo[2]; z <= hi[2]; ++z)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int x = lo[0]; x <= hi[0]; ++x) {
          long c = IDX(o, x, y, z);
          o->df1[q * n + c] = o->df2[q * n + c];
        }
}

/* One time step in do_thread order, lbm_ib.c:3115-3563 (periodicBC at
 * lbm_ib.c:3484 is commented out in the reference and is not called). */
int orc_step(orc_t *o, int nsteps) {
  for (int t = 0; t < nsteps; ++t) {
    if (o->nsheets > 0) {
      orc_bending(o);
      orc_stretching(o);
      orc_elastic(o);
      orc_spread(o);
      if (o->err) return o->err;
    }
    orc_collide(o);
    orc_stream(o);
    orc_bounceback(o);
    orc_rho_u(o);
    if (o->nsheets > 0) {
      orc_move(o);
      if (o->err) return o->err;
    }
    orc_inout(o);
    orc_swap(o);
    o->time += o->dt;
  }
  return o->err;
}

/* ---------------------------------------------------------- accessors */
long orc_ncells(const orc_t *o) { return o->n; }
double *orc_df1(orc_t *o) { return o->df1; }
double *orc_df2(orc_t *o) { return o->df2; }
double *orc_dfeq(orc_t *o) { return o->dfeq; }
double *orc_rho(orc_t *o) { return o->rho; }
double *orc_vel(orc_t *o) { return o->vel; }
double *orc_force(orc_t *o) { return o->force; }
double *orc_inlet(orc_t *o) { return o->inlet; }
double *orc_outlet(orc_t *o) { return o->outlet; }
double *orc_sheet_nodes(orc_t *o, int s) { return (s >= 0 && s < o->nsheets) ? o->sheets[s].node : NULL; }
int orc_time(const orc_t *o) { return o->time; }
int orc_error(const orc_t *o) { return o->err; }
double orc_cs2(const orc_t *o) { return o->cs2; }
double orc_cs4(const orc_t *o) { return o->cs4; }

/* Reference-mode constants of init_gv, lbm_ib.c:750-773, evaluated in the
 * reference's order: out = {tau, u_l, rho_l, Ks_l, Kb_l}. */
void orc_reference_constants(double *out) {
  double Re = 1.5e2, rho_l = 1.0e0, u_l = 0.001, L_l = 2.0e1;
  double nu_l = u_l * L_l / Re;
  double tau = 0.5 * (6.0 * nu_l + 1.0);
  double Kbhat = 0.005, Kshat = 2.0e1;
  double Kb_l = Kbhat * rho_l * u_l * u_l * pow(L_l, 4);
  double Ks_l = Kshat * rho_l * u_l * u_l * L_l;
  out[0] = tau; out[1] = u_l; out[2] = rho_l; out[3] = Ks_l; out[4] = Kb_l;
}
