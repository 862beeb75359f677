// lbmib_api.hip -- the C-ABI of liblbmib.so (declared in include/lbmib.h).
//
// One context owns one device, one HIP stream (default priority) and all
// device buffers of a simulation:
//   f[2]      population ping-pong, g.pelems elements each (planes or rows layout)
//   F[2]      force field ping-pong (step parity), dense [3][n], valid only
//             inside that step's support boxes
//   node arrays (FP64 positions/forces, support starts, delta factors)
//   bins      (y,z) 4x4-column bins of fiber node ids for the gather
// Per-step entry points only enqueue; downloads synchronise.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "lbmib_coarse.h"
#include "lbmib_ctx.h"
#include "lbmib_internal.h"

#define HIPCHK(x)                               \
  do {                                          \
    if ((x) != hipSuccess) return LBMIB_ERR_HIP; \
  } while (0)
#define RCHK(x)             \
  do {                      \
    int _r = (x);           \
    if (_r) return _r;      \
  } while (0)

static double host_feq(int q, double rho, double vx, double vy, double vz) {
  // the expression of init_eqlbrmdistrfuncDF0, lbm_ib.c:1561-1594
  double usq = vx * vx + vy * vy + vz * vz;
  if (q == 0) return 1.0 / 3.0 * rho * (1.0 - 1.5 * usq);
  double cu = kCX[q] * vx + kCY[q] * vy + kCZ[q] * vz;
  double w = (q <= 6) ? 1.0 / 18.0 : 1.0 / 36.0;
  return w * rho * (1.0 + 3.0 * cu + 4.5 * (pow(cu, 2.0)) - 1.5 * usq);
}

static void axis_ranges(int bc, int n, int *lo, int *hi, int *mlo, int *mhi) {
  if (bc == LBMIB_BC_PERIODIC) { *lo = 0; *hi = n - 1; *mlo = 0; *mhi = n - 1; }
  else if (bc == LBMIB_BC_INOUT) { *lo = 2; *hi = n - 3; *mlo = 3; *mhi = n - 4; }
  else { *lo = 2; *hi = n - 3; *mlo = 2; *mhi = n - 3; }
}

static int map_device_err(int bits) {
  if (bits & DERR_FIBER_OOB) return LBMIB_ERR_FIBER_OOB;
  if (bits & DERR_BIN_OVERFLOW) return LBMIB_ERR_BIN_OVERFLOW;
  return LBMIB_OK;
}

static size_t elem_size(int prec) {
  switch (prec) {
    case LBMIB_F64: return 8;
    case LBMIB_F32: return 4;
    case LBMIB_F16:
    case LBMIB_BF16: return 2;
  }
  return 0;
}

extern "C" int lbmib_destroy(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto e : ctx->evs) (void)hipEventDestroy(e);
  if (ctx->stream_b) (void)hipStreamSynchronize(ctx->stream_b);
  for (hipEvent_t e : {ctx->evFib, ctx->evBulk, ctx->evJoin})
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream_b) (void)hipStreamDestroy(ctx->stream_b);
  void *ptrs[] = {ctx->f[0], ctx->f[1], ctx->F[0], ctx->F[1], ctx->ib.pos, ctx->ib.frc, ctx->ib.start,
                  ctx->ib.wgt, ctx->ib.bin_count, ctx->ib.bin_list, ctx->ib.box, ctx->ib.done, ctx->ib.ub, ctx->d_sheets,
                  ctx->d_err, ctx->d_partials, ctx->d_tot, ctx->d_mat, ctx->d_tab, ctx->d_rowtab, ctx->d_rect, ctx->d_ref,
                  ctx->f_spare, ctx->pos_save, ctx->F_save, ctx->box_save, ctx->d_raster, ctx->d_cpart};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LBMIB_OK;
}

namespace lbmib {
int ensure_cpart(::lbmib_ctx *ctx, int F) {
  const size_t need = coarse_part_bytes(ctx->g, F);
  if (ctx->cpart_bytes >= need) return LBMIB_OK;
  if (ctx->d_cpart) (void)hipFree(ctx->d_cpart);
  ctx->d_cpart = nullptr;
  ctx->cpart_bytes = 0;
  if (hipMalloc((void **)&ctx->d_cpart, need) != hipSuccess) return LBMIB_ERR_ALLOC;
  ctx->cpart_bytes = need;
  return LBMIB_OK;
}

// Bytes of one step's stores in the rows the sheets' support boxes meet
// (estimated from the sheets' extents).
double box_row_bytes(const lbmib_ctx *ctx) {
  double rows = 0;
  for (int s = 0; s < ctx->nsheets; ++s) rows += box_rows_estimate(ctx->sd[s]);
  const double nrows = (double)ctx->g.ny * ctx->g.nz;
  return (rows < nrows ? rows : nrows) * 19.0 * ctx->g.nx * (double)ctx->esz;
}

int box_rows_auto(const lbmib_ctx *ctx) {
  return ctx->nsheets > 0 && box_row_bytes(ctx) <= kBoxRowCacheBytes ? 1 : 0;
}
}  // namespace lbmib

extern "C" int lbmib_create(const lbmib_fluid_params *fluid, int nsheets, const lbmib_sheet_params *sheets,
                            lbmib_ctx **out) {
  if (!fluid || !out || nsheets < 0 || nsheets > kMaxSheets || (nsheets > 0 && !sheets)) return LBMIB_ERR_ARG;
  *out = nullptr;
  const lbmib_fluid_params &p = *fluid;
  if (p.nx < 8 || p.ny < 8 || p.nz < 8 || p.nx > 4096 || p.ny > 65535 || p.nz > 65535) return LBMIB_ERR_ARG;
  if (p.bc_x < 0 || p.bc_x > 2 || p.bc_y < 1 || p.bc_y > 2 || p.bc_z < 1 || p.bc_z > 2) return LBMIB_ERR_ARG;
  if (!(p.tau > 0.5) || !(p.rho0 > 0)) return LBMIB_ERR_ARG;
  if (elem_size(p.precision) == 0) return LBMIB_ERR_ARG;
  if (p.dev_ref != LBMIB_DEV_EQ0 && p.dev_ref != LBMIB_DEV_REST) return LBMIB_ERR_ARG;
  if (p.layout != LBMIB_LAYOUT_PLANES && p.layout != LBMIB_LAYOUT_ROWS) return LBMIB_ERR_ARG;
  const long n = (long)p.nx * p.ny * p.nz;
  if (n > (1L << 31) - 1) return LBMIB_ERR_ARG;

  lbmib_ctx *ctx = new lbmib_ctx();
  ctx->fp = p;
  ctx->nsheets = nsheets;
  int off = 0;
  for (int s = 0; s < nsheets; ++s) {
    const lbmib_sheet_params &q = sheets[s];
    if (q.rows < 4 || q.cols < 4 || !(q.width > 0) || !(q.height > 0)) { delete ctx; return LBMIB_ERR_ARG; }
    ctx->sp[s] = q;
    SheetDesc &d = ctx->sd[s];
    d.rows = q.rows; d.cols = q.cols; d.off = off; d.pad = 0;
    d.ds1 = q.width / (q.cols - 1);   // lbm_ib.c:994
    d.ds2 = q.height / (q.rows - 1);  // lbm_ib.c:995
    d.bend_c = q.kb / pow(d.ds2, 3.0);
    d.str_row_c = q.ks / (d.ds1 * d.ds1);
    d.str_col_c = q.ks / (d.ds2 * d.ds2);
    off += q.rows * q.cols;
  }
  ctx->nnodes = off;

  Geometry &g = ctx->g;
  g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
  g.bc[0] = p.bc_x; g.bc[1] = p.bc_y; g.bc[2] = p.bc_z;
  const int dims[3] = {p.nx, p.ny, p.nz};
  for (int a = 0; a < 3; ++a) axis_ranges(g.bc[a], dims[a], &g.lo[a], &g.hi[a], &g.mlo[a], &g.mhi[a]);
  g.n = n;
  long pad = 256;
  if (const char *e = getenv("LBMIB_PLANE_PAD")) pad = atol(e);
  // the pad keeps the x-shifted reads of the last row's edge cells inside
  // the array (the pair kernel on a periodic x axis reads one element past
  // the last row), and the buffer descriptors do not clamp: never below 128
  if (pad < 128) pad = 128;
  pad = (pad + 127) / 128 * 128;
  if (p.layout == LBMIB_LAYOUT_ROWS) {
    g.pq = p.nx;
    g.prow = 19L * p.nx;
    g.pelems = 19 * n + pad;
  } else {
    g.pq = (n + 127) / 128 * 128 + pad;
    g.prow = p.nx;
    g.pelems = 19 * g.pq;
  }

  // BGK/Guo constants as init_gv and compute_eqlbrmdistrfuncDF1 form them
  FluidConsts<double> &c = ctx->c;
  const double tau = p.tau;
  c.omc = 1.0 - 1.0 / tau;
  c.om = 1.0 / tau;
  for (int q = 0; q < 19; ++q) {
    c.pref[q] = 1 * (1.0 - 1.0 / (2.0 * tau)) / (double)kWDEN[q];
    c.pw[q] = (1.0 / kWDEN[q]) * (1.0 - 1.0 / (2.0 * tau));
  }
  const double cs = 1.0 / (sqrt(3.0));  // lbm_ib.c:765
  c.cs2 = cs * cs;
  c.cs4 = pow(cs, 4.0);
  c.g = p.g;
  c.rho0 = p.rho0; c.u0x = p.u0[0]; c.u0y = p.u0[1]; c.u0z = p.u0[2];
  // the deviation forms store f - feq(rhor, ur): the initial flow, or the
  // rest state (r = w)
  const bool dev = p.precision != LBMIB_F64, eq0 = p.dev_ref == LBMIB_DEV_EQ0;
  c.rhor = eq0 ? p.rho0 : 1.0;
  c.urx = eq0 ? p.u0[0] : 0.0;  // the inflow component
  for (int q = 0; q < 19; ++q) {
    // the inlet/outlet tables are the initial equilibrium feq(rho0, u0)
    const double feq = host_feq(q, p.rho0, p.u0[0], p.u0[1], p.u0[2]);
    c.ref[q] = dev ? host_feq(q, c.rhor, c.urx, 0.0, 0.0) : 0.0;
    c.tab_in[q] = c.tab_out[q] = feq - c.ref[q];
  }

  ctx->esz = elem_size(p.precision);
  ctx->fsz = p.precision == LBMIB_F64 ? 8 : 4;

  int rc = LBMIB_OK;
  auto alloc = [&](void **ptr, size_t bytes) {
    if (rc) return;
    if (hipMalloc(ptr, bytes) != hipSuccess) { *ptr = nullptr; rc = LBMIB_ERR_ALLOC; }
  };
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) rc = LBMIB_ERR_HIP;
  if (!rc && hipStreamCreateWithFlags(&ctx->stream_b, hipStreamNonBlocking) != hipSuccess) rc = LBMIB_ERR_HIP;
  for (hipEvent_t *e : {&ctx->evFib, &ctx->evBulk, &ctx->evJoin})
    if (!rc && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) rc = LBMIB_ERR_HIP;
  {
    int dev = 0;
    if (!rc && (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess))
      rc = LBMIB_ERR_HIP;
  }
  const size_t fbytes = (size_t)g.pelems * ctx->esz;
  alloc(&ctx->f[0], fbytes);
  alloc(&ctx->f[1], fbytes);
  alloc((void **)&ctx->d_err, sizeof(int));
  ctx->nparts = totals_blocks(g);
  alloc((void **)&ctx->d_partials, sizeof(double) * 4 * (size_t)ctx->nparts);
  alloc((void **)&ctx->d_tot, sizeof(double) * 4);
  alloc((void **)&ctx->ib.box, sizeof(int) * 2 * kMaxSheets * 6);
  alloc(&ctx->d_tab, sizeof(double) * 19);
  alloc((void **)&ctx->d_rect, sizeof(int) * 2 * kRectBufInts);
  if (!rc && hipMemset(ctx->d_rect, 0, sizeof(int) * 2 * kRectBufInts) != hipSuccess) rc = LBMIB_ERR_HIP;
  ctx->ib.band = ctx->d_rect;
  alloc((void **)&ctx->d_ref, sizeof(double) * 19);
  if (!rc) {
    rc = upload_fluid_tables(p.precision, c, ctx->d_tab, ctx->d_ref);
    c.dtab = ctx->d_tab;
    c.dref = ctx->d_ref;
  }
  if (!rc && row_path_buffered(g, ctx->esz)) {
    std::vector<uint32_t> tab;
    if (row_source_table(g, ctx->esz, tab) == LBMIB_OK) {
      alloc((void **)&ctx->d_rowtab, sizeof(uint32_t) * tab.size());
      if (!rc && hipMemcpy(ctx->d_rowtab, tab.data(), sizeof(uint32_t) * tab.size(), hipMemcpyHostToDevice) !=
                     hipSuccess)
        rc = LBMIB_ERR_HIP;
    }
  }

  IBState &ib = ctx->ib;
  ib.nsheets = nsheets;
  ctx->kv.box_rows_cached = box_rows_auto(ctx);
  ctx->move_tiles = move_tile_count(ctx->sd, nsheets);
  ib.nnodes = ctx->nnodes;
  ib.quirks = p.quirks;
  ib.nby = (p.ny + 3) / 4;
  ib.nbz = (p.nz + 3) / 4;
  // bin capacity: a 4x4 column holds at most (4/ds1+1)(4/ds2+1) nodes of a
  // flat sheet; (4/ds1+2)(4/ds2+2) per sheet with a factor 1.5 of headroom
  // for deformation (unit spacing: 64 for one sheet, 2.6x the flat maximum).
  // Overflow is reported (LBMIB_ERR_BIN_OVERFLOW).  The spreading kernel's
  // LDS grows with it, and with it fewer tiles are resident per CU.
  double dens = 0;
  for (int s = 0; s < nsheets; ++s) dens += (4.0 / ctx->sd[s].ds1 + 2.0) * (4.0 / ctx->sd[s].ds2 + 2.0);
  int cap = 32;
  while (cap < 1.5 * dens && cap < 1024) cap <<= 1;
  ib.cap = cap;
  if (ctx->nnodes > 0) {
    const size_t M = ctx->nnodes;
    alloc(&ctx->F[0], (size_t)3 * n * ctx->fsz);
    alloc(&ctx->F[1], (size_t)3 * n * ctx->fsz);
    alloc((void **)&ib.pos, sizeof(double) * 3 * M);
    alloc((void **)&ib.frc, sizeof(double) * NF_COUNT * M);
    alloc((void **)&ib.start, sizeof(int) * 3 * M);
    alloc((void **)&ib.wgt, sizeof(double) * 12 * M);
    alloc((void **)&ib.bin_count, sizeof(int) * (size_t)ib.nby * ib.nbz);
    alloc((void **)&ib.bin_list, sizeof(int4) * (size_t)ib.nby * ib.nbz * cap);
    alloc((void **)&ctx->d_sheets, sizeof(SheetDesc) * kMaxSheets);
    alloc((void **)&ib.ub, sizeof(double) * 3 * n);
    alloc((void **)&ib.done, sizeof(unsigned));
  }
  ib.sheets = ctx->d_sheets;
  ib.err = ctx->d_err;
  if (!rc && ctx->nnodes > 0) {
    if (hipMemcpy(ctx->d_sheets, ctx->sd, sizeof(SheetDesc) * nsheets, hipMemcpyHostToDevice) != hipSuccess)
      rc = LBMIB_ERR_HIP;
    if (!rc && hipMemset(ib.pos, 0, sizeof(double) * 3 * ctx->nnodes) != hipSuccess) rc = LBMIB_ERR_HIP;
    if (!rc && hipMemset(ib.frc, 0, sizeof(double) * NF_COUNT * ctx->nnodes) != hipSuccess) rc = LBMIB_ERR_HIP;
    if (!rc && hipMemset(ib.done, 0, sizeof(unsigned)) != hipSuccess) rc = LBMIB_ERR_HIP;
  }
  if (!rc && hipMemset(ctx->d_err, 0, sizeof(int)) != hipSuccess) rc = LBMIB_ERR_HIP;
  if (!rc) rc = launch_ib_reset(ib, ctx->nnodes > 0, 3, ctx->stream);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = LBMIB_ERR_HIP;
  if (rc) {
    lbmib_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return LBMIB_OK;
}

extern "C" int lbmib_set_fiber_positions(lbmib_ctx *ctx, int sheet, const double *xyz) {
  if (!ctx || !xyz || sheet < 0 || sheet >= ctx->nsheets) return LBMIB_ERR_ARG;
  const SheetDesc &d = ctx->sd[sheet];
  const int m = d.rows * d.cols, M = ctx->nnodes;
  std::vector<double> soa((size_t)3 * m);
  for (int k = 0; k < m; ++k)
    for (int a = 0; a < 3; ++a) soa[(size_t)a * m + k] = xyz[3 * (size_t)k + a];
  double ext[4] = {1e300, -1e300, 1e300, -1e300};
  for (int k = 0; k < m; ++k) {
    ext[0] = fmin(ext[0], xyz[3 * (size_t)k + 1]);
    ext[1] = fmax(ext[1], xyz[3 * (size_t)k + 1]);
    ext[2] = fmin(ext[2], xyz[3 * (size_t)k + 2]);
    ext[3] = fmax(ext[3], xyz[3 * (size_t)k + 2]);
  }
  for (int a = 0; a < 4; ++a) ctx->hyz[sheet][a] = ext[a];
  for (int a = 0; a < 3; ++a)
    HIPCHK(hipMemcpyAsync(ctx->ib.pos + (size_t)a * M + d.off, soa.data() + (size_t)a * m, sizeof(double) * m,
                          hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return LBMIB_OK;
}

// ------------------------------------------------------------------ init

extern "C" int lbmib_init_eqlbrmdistrfuncDF0(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  // rho = rho0 and u = u0 everywhere (lbm_ib.c:837-849); the equilibrium is
  // uniform, so it is held as the 19-entry table in the context.
  ctx->init_bits |= 1;
  ctx->time = 1;
  return LBMIB_OK;
}

extern "C" int lbmib_init_df1(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  if (!(ctx->init_bits & 1)) return LBMIB_ERR_STATE;
  RCHK(launch_fill_init(ctx->fp.precision, ctx->g, ctx->f[0], ctx->c, ctx->stream));
  ctx->cur = 0;
  ctx->own = 1;
  ctx->pending = 0;
  ctx->band_done = 0;
  ctx->spread_slot = -1;
  RCHK(launch_ib_reset(ctx->ib, ctx->nnodes > 0, 3, ctx->stream));
  ctx->ib_clean = 1;
  ctx->init_bits |= 2;
  return LBMIB_OK;
}

extern "C" int lbmib_init_df_inout(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  if (!(ctx->init_bits & 1)) return LBMIB_ERR_STATE;
  // the inlet/outlet tables (feq(rho0,u0) sampled at x = 2 and x = N-1,
  // lbm_ib.c:1632-1668) are the context's tab_in/tab_out
  ctx->init_bits |= 4;
  return LBMIB_OK;
}

static int ready(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  if (ctx->init_bits != 7) return LBMIB_ERR_STATE;
  return LBMIB_OK;
}

// ----------------------------------------------------------- IB pieces

static int slot_now(const lbmib_ctx *ctx) { return ctx->time & 1; }

static int ensure_ib_clean(lbmib_ctx *ctx) {
  if (!ctx->ib_clean) {
    RCHK(launch_ib_reset(ctx->ib, 1, 1 << slot_now(ctx), ctx->stream));
    ctx->ib_clean = 1;
  }
  return LBMIB_OK;
}

extern "C" int lbmib_compute_bendingforce(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  return launch_fiber(ctx->ib, ctx->g, FK_BEND, slot_now(ctx), ctx->stream);
}
extern "C" int lbmib_compute_stretchingforce(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  return launch_fiber(ctx->ib, ctx->g, FK_STRETCH, slot_now(ctx), ctx->stream);
}
extern "C" int lbmib_compute_elasticforce(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  return launch_fiber(ctx->ib, ctx->g, FK_ELASTIC, slot_now(ctx), ctx->stream);
}
extern "C" int lbmib_compute_fiber_forces(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  return launch_fiber(ctx->ib, ctx->g, FK_BEND | FK_STRETCH | FK_ELASTIC, slot_now(ctx), ctx->stream);
}

extern "C" int lbmib_fiber_SpreadForce(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  RCHK(ensure_ib_clean(ctx));
  ctx->ib_clean = 0;
  return launch_fiber(ctx->ib, ctx->g, FK_PREP, slot_now(ctx), ctx->stream);
}

extern "C" int lbmib_fluid_get_SpreadForce(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  if (ctx->nnodes == 0) return LBMIB_OK;
  const int s = slot_now(ctx);
  RCHK(launch_gather(ctx->fp.precision, ctx->ib, ctx->g, s, ctx->F[s], ctx->stream));
  ctx->spread_slot = s;
  return LBMIB_OK;
}

extern "C" int lbmib_get_influentialdomain_fluid_grid_and_SpreadForce(lbmib_ctx *ctx) {
  RCHK(lbmib_fiber_SpreadForce(ctx));
  return lbmib_fluid_get_SpreadForce(ctx);
}

// ------------------------------------------------------------ fluid step

extern "C" int lbmib_compute_eqlbrmdistrfuncDF1(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  if (ctx->pending) return LBMIB_ERR_STATE;
  const int s = slot_now(ctx);
  const int *box = ctx->ib.box;
  RCHK(launch_collide_stream(ctx->fp.precision, ctx->g, ctx->f[ctx->cur], ctx->f[1 - ctx->cur], ctx->own, ctx->c,
                             ctx->nsheets, box + s * kMaxSheets * 6, box + (1 - s) * kMaxSheets * 6, ctx->F[s],
                             ctx->F[1 - s], ctx->d_rowtab, ctx->kv, ctx->stream));
  ctx->pending = 1;
  ctx->band_done = 0;
  return LBMIB_OK;
}

extern "C" int lbmib_stream_distrfunc(lbmib_ctx *ctx) { return ready(ctx); }
extern "C" int lbmib_bounceback_rigidwalls(lbmib_ctx *ctx) { return ready(ctx); }
extern "C" int lbmib_copy_inout_to_df2(lbmib_ctx *ctx) { return ready(ctx); }
extern "C" int lbmib_periodicBC(lbmib_ctx *ctx) { return ready(ctx); }

// rho/u are evaluated inside the interpolation kernel on exactly the cells
// it reads (the fiber supports); full fields come from
// lbmib_get_macroscopic.  Kept for drop-in compatibility.
extern "C" int lbmib_compute_rho_and_u(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  if (!ctx->pending) return LBMIB_ERR_STATE;
  return LBMIB_OK;
}

extern "C" int lbmib_fluid_SpreadVelocity(lbmib_ctx *ctx) { return lbmib_compute_rho_and_u(ctx); }

extern "C" int lbmib_fiber_get_SpreadVelocity(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  if (!ctx->pending) return LBMIB_ERR_STATE;
  if (ctx->nnodes == 0 || ctx->spread_slot != slot_now(ctx)) return LBMIB_OK;
  const int s = slot_now(ctx);
  RCHK(launch_move(ctx->fp.precision, ctx->ib, ctx->g, ctx->f[1 - ctx->cur], ctx->F[s], ctx->c, 1 - s, ctx->move_tiles, ctx->stream));
  ctx->ib_clean = 1;  // k_move clears the bins and the next step's box slot
  return LBMIB_OK;
}

extern "C" int lbmib_moveFiberSheet(lbmib_ctx *ctx) { return lbmib_fiber_get_SpreadVelocity(ctx); }

extern "C" int lbmib_replace_old_DF(lbmib_ctx *ctx) {
  RCHK(ready(ctx));
  if (!ctx->pending) return LBMIB_ERR_STATE;
  ctx->cur = 1 - ctx->cur;
  ctx->own = 0;
  ctx->pending = 0;
  ctx->band_done = 0;
  if (!ctx->ib_clean) {
    // a spread without a move: clear before the next step's binning
    RCHK(launch_ib_reset(ctx->ib, 1, 1 << ((ctx->time + 1) & 1), ctx->stream));
    ctx->ib_clean = 1;
  }
  ctx->time += 1;
  return LBMIB_OK;
}

// cstored (lbmib_step_coarse): the fluid launch also writes the coarse
// stored-moment planes of the new state at factor cf.
static int one_step(lbmib_ctx *ctx, hipEvent_t e0, hipEvent_t e1, int cf = 0, float *cstored = nullptr) {
  const int s = slot_now(ctx);
  const int prec = ctx->fp.precision;
  if (ctx->nnodes > 0) {
    RCHK(ensure_ib_clean(ctx));
    RCHK(launch_fiber(ctx->ib, ctx->g, FK_ALL, s, ctx->stream));
    RCHK(launch_gather(prec, ctx->ib, ctx->g, s, ctx->F[s], ctx->stream));
    ctx->spread_slot = s;
  }
  const int *box = ctx->ib.box;
  if (e0) HIPCHK(hipEventRecord(e0, ctx->stream));
  if (cstored)
    RCHK(launch_collide_coarse(prec, ctx->g, ctx->f[ctx->cur], ctx->f[1 - ctx->cur], ctx->own, ctx->c, ctx->nsheets,
                               box + s * kMaxSheets * 6, box + (1 - s) * kMaxSheets * 6, ctx->F[s], ctx->F[1 - s],
                               ctx->d_rowtab, ctx->kv, cf, ctx->d_cpart, cstored, ctx->stream));
  else
    RCHK(launch_collide_stream(prec, ctx->g, ctx->f[ctx->cur], ctx->f[1 - ctx->cur], ctx->own, ctx->c,
                               ctx->nsheets, box + s * kMaxSheets * 6, box + (1 - s) * kMaxSheets * 6, ctx->F[s],
                               ctx->F[1 - s], ctx->d_rowtab, ctx->kv, ctx->stream));
  if (e1) HIPCHK(hipEventRecord(e1, ctx->stream));
  if (ctx->nnodes > 0) {
    RCHK(launch_move(prec, ctx->ib, ctx->g, ctx->f[1 - ctx->cur], ctx->F[s], ctx->c, 1 - s, ctx->move_tiles, ctx->stream));
    ctx->ib_clean = 1;
  }
  ctx->cur = 1 - ctx->cur;
  ctx->own = 0;
  ctx->pending = 0;
  ctx->band_done = 0;
  ctx->time += 1;
  return LBMIB_OK;
}

// ------------------------------------------------------------ split steps
//
// With sheets present, a split step updates the band -- the cells of the
// sheets' support boxes of this and the previous step, widened by one cell,
// whose boxes the step's fiber kernel stores -- and the bulk, every other
// cell, in two launches that read the same populations and write disjoint
// cells (lbmib_fluid.hip, "split steps"):
//   stream   : wait fiber(t) | bulk(t)
//   stream_b : fiber(t), spreading(t) | wait bulk(t-1) | band(t) | band
//              velocity + move(t) | fiber(t+1) | spreading(t+1)
// The bulk needs no force, only the band boxes and the previous step's
// populations (the previous band precedes fiber(t) on stream_b); the band
// velocity reads the new populations only on support cells and their
// neighbours, which the band holds.  So the main stream runs bulk after bulk
// while the band and the whole IB chain run beside it.  Results are
// bit-identical to unsplit steps: every cell does the same arithmetic.

// Split steps need sheets and the buffer-addressed row kernels, and run only
// when asked for (mode 1): unsplit steps measured as fast or faster on every
// workload once the IB chain was short, because the IB kernels run several
// times slower beside a bulk launch that saturates HBM.
static bool split_enabled(const lbmib_ctx *ctx) {
  if (ctx->nnodes == 0 || !ctx->d_rowtab) return false;
  return ctx->overlap_req == 1;
}

// Grid of the band update (256 threads a block): about one thread per band
// cell, from the sheets' y/z extent at upload and a support about 8 cells
// thick in x (the kernel strides over any further cells).
static int band_blocks(const lbmib_ctx *ctx) {
  double cells = 0;
  for (int s = 0; s < ctx->nsheets; ++s)
    cells += 8.0 * (ctx->hyz[s][1] - ctx->hyz[s][0] + 8.0) * (ctx->hyz[s][3] - ctx->hyz[s][2] + 8.0);
  return (int)std::min(std::max(cells / 256.0, 64.0), 16.0 * ctx->ncu);
}

static int split_step(lbmib_ctx *ctx, hipEvent_t e0, hipEvent_t e1, bool more) {
  const int s = slot_now(ctx), prec = ctx->fp.precision;
  const int *box = ctx->ib.box;
  int *band = ctx->d_rect + s * kRectBufInts;  // written by fiber(t)
  if (ctx->split_run == 0) {
    // stream_b starts after everything already queued on the main stream,
    // which also stands for the previous step's bulk
    HIPCHK(hipEventRecord(ctx->evJoin, ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->stream_b, ctx->evJoin, 0));
    HIPCHK(hipEventRecord(ctx->evBulk, ctx->stream));
  }
  if (!ctx->ib_ready) {
    if (!ctx->ib_clean) {
      RCHK(launch_ib_reset(ctx->ib, 1, 1 << s, ctx->stream_b));
      ctx->ib_clean = 1;
    }
    RCHK(launch_fiber(ctx->ib, ctx->g, FK_ALL, s, ctx->stream_b, true));
    HIPCHK(hipEventRecord(ctx->evFib, ctx->stream_b));
    RCHK(launch_gather(prec, ctx->ib, ctx->g, s, ctx->F[s], ctx->stream_b));
  }
  ctx->spread_slot = s;
  // main stream: the bulk, after fiber(t) (band boxes; the previous band
  // precedes it on stream_b); no sheets: its cells carry no force
  HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->evFib, 0));
  if (e0) HIPCHK(hipEventRecord(e0, ctx->stream));
  RCHK(launch_collide_stream(prec, ctx->g, ctx->f[ctx->cur], ctx->f[1 - ctx->cur], ctx->own, ctx->c, 0,
                             box + s * kMaxSheets * 6, box + (1 - s) * kMaxSheets * 6, ctx->F[s], ctx->F[1 - s],
                             ctx->d_rowtab, ctx->kv, ctx->stream, 2, band, 0));
  if (e1) HIPCHK(hipEventRecord(e1, ctx->stream));
  // second stream: the band after this step's spreading (stream order) and
  // the previous bulk, which read the buffer the band writes
  HIPCHK(hipStreamWaitEvent(ctx->stream_b, ctx->evBulk, 0));
  HIPCHK(hipEventRecord(ctx->evBulk, ctx->stream));
  RCHK(launch_collide_stream(prec, ctx->g, ctx->f[ctx->cur], ctx->f[1 - ctx->cur], ctx->own, ctx->c, ctx->nsheets,
                             box + s * kMaxSheets * 6, box + (1 - s) * kMaxSheets * 6, ctx->F[s], ctx->F[1 - s],
                             ctx->d_rowtab, ctx->kv, ctx->stream_b, 1, band, band_blocks(ctx)));
  // band velocity and move, then the next step's fiber forces and spreading
  // (slot 1 - s, reset by the move)
  RCHK(launch_move(prec, ctx->ib, ctx->g, ctx->f[1 - ctx->cur], ctx->F[s], ctx->c, 1 - s, ctx->move_tiles, ctx->stream_b));
  ctx->ib_clean = 1;
  ctx->ib_ready = 0;
  if (more) {
    RCHK(launch_fiber(ctx->ib, ctx->g, FK_ALL, 1 - s, ctx->stream_b, true));
    HIPCHK(hipEventRecord(ctx->evFib, ctx->stream_b));
    RCHK(launch_gather(prec, ctx->ib, ctx->g, 1 - s, ctx->F[1 - s], ctx->stream_b));
    ctx->ib_ready = 1;
  }
  ctx->cur = 1 - ctx->cur;
  ctx->own = 0;
  ctx->pending = 0;
  ctx->band_done = 0;
  ctx->time += 1;
  ctx->split_run += 1;
  return LBMIB_OK;
}

// Joins the second stream back into the main one after a run of split steps.
static int split_join(lbmib_ctx *ctx) {
  if (ctx->split_run == 0) return LBMIB_OK;
  HIPCHK(hipEventRecord(ctx->evJoin, ctx->stream_b));
  HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->evJoin, 0));
  ctx->split_run = 0;
  return LBMIB_OK;
}

extern "C" int lbmib_set_overlap(lbmib_ctx *ctx, int mode) {
  if (!ctx || mode < -1 || mode > 1) return LBMIB_ERR_ARG;
  RCHK(split_join(ctx));
  ctx->overlap_req = mode;
  return LBMIB_OK;
}

extern "C" int lbmib_timed_kernel_cells(lbmib_ctx *ctx, double *cells) {
  if (!ctx || !cells) return LBMIB_ERR_ARG;
  *cells = (double)ctx->g.n;  // split steps time the band and bulk updates together
  return LBMIB_OK;
}

extern "C" int lbmib_step(lbmib_ctx *ctx, int nsteps) {
  RCHK(ready(ctx));
  if (ctx->pending || nsteps < 0) return LBMIB_ERR_STATE;
  if (split_enabled(ctx)) {
    for (int k = 0; k < nsteps; ++k) RCHK(split_step(ctx, nullptr, nullptr, k + 1 < nsteps));
    return split_join(ctx);
  }
  for (int k = 0; k < nsteps; ++k) RCHK(one_step(ctx, nullptr, nullptr));
  return LBMIB_OK;
}

extern "C" int lbmib_step_coarse(lbmib_ctx *ctx, int factor, float *d_stored) {
  RCHK(ready(ctx));
  if (!d_stored || (factor != 1 && factor != 2 && factor != 4)) return LBMIB_ERR_ARG;
  const Geometry &g = ctx->g;
  if (g.nx % 64 || g.ny % 4 || g.nz % 4) return LBMIB_ERR_ARG;
  if (ctx->pending) return LBMIB_ERR_STATE;
  if (split_enabled(ctx) || !collide_coarse_ok(g, ctx->esz, ctx->d_rowtab != nullptr, factor)) {
    // the same result in two passes: a plain step, then the read
    RCHK(lbmib_step(ctx, 1));
    return lbmib_coarse_state(ctx, factor, nullptr, d_stored);
  }
  RCHK(ensure_cpart(ctx, factor));
  return one_step(ctx, nullptr, nullptr, factor, d_stored);
}

// --------------------------------------------------------------- queries

static int poll_status(lbmib_ctx *ctx) {
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream_b));
  int bits = 0;
  HIPCHK(hipMemcpy(&bits, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost));
  const int e = map_device_err(bits);
  if (e && !ctx->latched) ctx->latched = e;
  return ctx->latched;
}

extern "C" int lbmib_sync(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  return poll_status(ctx);
}

extern "C" int lbmib_status(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  return poll_status(ctx);
}

extern "C" int lbmib_get_time(lbmib_ctx *ctx, int *t) {
  if (!ctx || !t) return LBMIB_ERR_ARG;
  *t = ctx->time;
  return LBMIB_OK;
}

extern "C" int lbmib_get_fibers(lbmib_ctx *ctx, int sheet, double *out) {
  if (!ctx || !out || sheet < 0 || sheet >= ctx->nsheets) return LBMIB_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const SheetDesc &d = ctx->sd[sheet];
  const int m = d.rows * d.cols, M = ctx->nnodes;
  std::vector<double> pos((size_t)3 * m), frc((size_t)NF_COUNT * m);
  for (int a = 0; a < 3; ++a)
    HIPCHK(hipMemcpy(pos.data() + (size_t)a * m, ctx->ib.pos + (size_t)a * M + d.off, sizeof(double) * m,
                     hipMemcpyDeviceToHost));
  for (int a = 0; a < NF_COUNT; ++a)
    HIPCHK(hipMemcpy(frc.data() + (size_t)a * m, ctx->ib.frc + (size_t)a * M + d.off, sizeof(double) * m,
                     hipMemcpyDeviceToHost));
  for (int k = 0; k < m; ++k) {
    double *r = out + (size_t)12 * k;
    for (int a = 0; a < 3; ++a) r[a] = pos[(size_t)a * m + k];
    for (int a = 0; a < NF_COUNT; ++a) r[3 + a] = frc[(size_t)a * m + k];
  }
  return LBMIB_OK;
}

static int ensure_mat(lbmib_ctx *ctx, size_t bytes) {
  if (ctx->mat_bytes >= bytes) return LBMIB_OK;
  if (ctx->d_mat) (void)hipFree(ctx->d_mat);
  ctx->d_mat = nullptr;
  ctx->mat_bytes = 0;
  if (hipMalloc((void **)&ctx->d_mat, bytes) != hipSuccess) return LBMIB_ERR_ALLOC;
  ctx->mat_bytes = bytes;
  return LBMIB_OK;
}

// slot/box/F describing the force of the last completed step
static void last_force(lbmib_ctx *ctx, const int **box, const void **F, int *nsh) {
  const int s = ctx->spread_slot;
  if (s < 0 || ctx->nnodes == 0) {
    *box = ctx->ib.box;  // both slots empty before the first spread
    *F = nullptr;
    *nsh = 0;
    return;
  }
  *box = ctx->ib.box + s * kMaxSheets * 6;
  *F = ctx->F[s];
  *nsh = ctx->nsheets;
}

extern "C" int lbmib_get_macroscopic(lbmib_ctx *ctx, double *rho, double *vel) {
  RCHK(ready(ctx));
  if (ctx->pending || (!rho && !vel)) return LBMIB_ERR_STATE;
  const long n = ctx->g.n;
  RCHK(ensure_mat(ctx, sizeof(double) * 4 * (size_t)n));
  const int *box;
  const void *F;
  int nsh;
  last_force(ctx, &box, &F, &nsh);
  RCHK(launch_materialize(ctx->fp.precision, ctx->g, ctx->f[ctx->cur], ctx->own, ctx->c, nsh, box, F, nullptr,
                          ctx->d_mat, ctx->d_mat + n, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (rho) HIPCHK(hipMemcpy(rho, ctx->d_mat, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (vel) HIPCHK(hipMemcpy(vel, ctx->d_mat + n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  return LBMIB_OK;
}

extern "C" int lbmib_get_df(lbmib_ctx *ctx, double *df) {
  RCHK(ready(ctx));
  if (ctx->pending || !df) return LBMIB_ERR_STATE;
  const long n = ctx->g.n;
  RCHK(ensure_mat(ctx, sizeof(double) * 19 * (size_t)n));
  RCHK(launch_materialize(ctx->fp.precision, ctx->g, ctx->f[ctx->cur], ctx->own, ctx->c, 0, ctx->ib.box, nullptr,
                          ctx->d_mat, nullptr, nullptr, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(df, ctx->d_mat, sizeof(double) * 19 * n, hipMemcpyDeviceToHost));
  return LBMIB_OK;
}

extern "C" int lbmib_get_force(lbmib_ctx *ctx, double *force) {
  RCHK(ready(ctx));
  if (!force) return LBMIB_ERR_ARG;
  const long n = ctx->g.n;
  memset(force, 0, sizeof(double) * 3 * n);
  if (ctx->nnodes == 0 || ctx->spread_slot < 0) return LBMIB_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int s = ctx->spread_slot;
  int box[kMaxSheets * 6];
  HIPCHK(hipMemcpy(box, ctx->ib.box + s * kMaxSheets * 6, sizeof(int) * 6 * ctx->nsheets, hipMemcpyDeviceToHost));
  std::vector<char> raw((size_t)3 * n * ctx->fsz);
  HIPCHK(hipMemcpy(raw.data(), ctx->F[s], raw.size(), hipMemcpyDeviceToHost));
  for (int sh = 0; sh < ctx->nsheets; ++sh) {
    const int *b = box + 6 * sh;
    if (b[0] > b[1]) continue;
    for (int z = b[4]; z <= b[5]; ++z)
      for (int y = b[2]; y <= b[3]; ++y)
        for (int x = b[0]; x <= b[1]; ++x) {
          const long c0 = (long)x + (long)ctx->g.nx * ((long)y + (long)ctx->g.ny * z);
          for (int a = 0; a < 3; ++a) {
            const size_t k = (size_t)a * n + c0;
            force[k] = ctx->fsz == 8 ? ((const double *)raw.data())[k] : (double)((const float *)raw.data())[k];
          }
        }
  }
  return LBMIB_OK;
}

extern "C" int lbmib_get_totals(lbmib_ctx *ctx, double *out4) {
  RCHK(ready(ctx));
  if (ctx->pending || !out4) return LBMIB_ERR_STATE;
  RCHK(launch_totals(ctx->fp.precision, ctx->g, ctx->f[ctx->cur], ctx->own, ctx->c, ctx->d_partials, ctx->nparts,
                     ctx->d_tot, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(out4, ctx->d_tot, sizeof(double) * 4, hipMemcpyDeviceToHost));
  for (int k = 0; k < 4; ++k)
    if (!isfinite(out4[k])) {
      if (!ctx->latched) ctx->latched = LBMIB_ERR_NONFINITE;
      return LBMIB_ERR_NONFINITE;
    }
  return LBMIB_OK;
}

extern "C" int lbmib_get_band_box(lbmib_ctx *ctx, int sheet, int *box6) {
  if (!ctx || !box6 || sheet < 0 || sheet >= ctx->nsheets) return LBMIB_ERR_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const int s = ctx->spread_slot < 0 ? 0 : ctx->spread_slot;
  HIPCHK(hipMemcpy(box6, ctx->ib.box + (s * kMaxSheets + sheet) * 6, sizeof(int) * 6, hipMemcpyDeviceToHost));
  return LBMIB_OK;
}

// ----------------------------------------------------------- measurement

extern "C" int lbmib_time_steps(lbmib_ctx *ctx, int nwarm, int nsteps, double *ms_total, double *ms_main_mean) {
  RCHK(ready(ctx));
  if (ctx->pending || nsteps <= 0 || nwarm < 0) return LBMIB_ERR_STATE;
  const size_t need = 2 + 2 * (size_t)nsteps;
  while (ctx->evs.size() < need) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    ctx->evs.push_back(e);
  }
  const bool split = split_enabled(ctx);
  auto step_once = [&](hipEvent_t a, hipEvent_t b, bool more) {
    return split ? split_step(ctx, a, b, more) : one_step(ctx, a, b);
  };
  for (int k = 0; k < nwarm; ++k) RCHK(step_once(nullptr, nullptr, k + 1 < nwarm));
  RCHK(split_join(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipEventRecord(ctx->evs[0], ctx->stream));
  for (int k = 0; k < nsteps; ++k) RCHK(step_once(nullptr, nullptr, k + 1 < nsteps));
  RCHK(split_join(ctx));
  HIPCHK(hipEventRecord(ctx->evs[1], ctx->stream));
  HIPCHK(hipEventSynchronize(ctx->evs[1]));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->evs[0], ctx->evs[1]));
  if (ms_total) *ms_total = ms;
  if (ms_main_mean) {
    for (int k = 0; k < nsteps; ++k) RCHK(step_once(ctx->evs[2 + 2 * k], ctx->evs[3 + 2 * k], k + 1 < nsteps));
    RCHK(split_join(ctx));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double acc = 0;
    for (int k = 0; k < nsteps; ++k) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, ctx->evs[2 + 2 * k], ctx->evs[3 + 2 * k]));
      acc += t;
    }
    *ms_main_mean = acc / nsteps;
  }
  return poll_status(ctx);
}

extern "C" int lbmib_set_box_row_stores(lbmib_ctx *ctx, int cached) {
  if (!ctx || cached < -1 || cached > 1) return LBMIB_ERR_ARG;
  ctx->kv.box_rows_cached = cached < 0 ? box_rows_auto(ctx) : cached;
  return LBMIB_OK;
}

extern "C" int lbmib_set_kernel_variant(lbmib_ctx *ctx, int cells_per_thread, int nontemporal, int blocks_per_cu) {
  if (!ctx || cells_per_thread < 0 || cells_per_thread > 8 || nontemporal < -1 || nontemporal > 3 ||
      blocks_per_cu < 0 || blocks_per_cu > 64)
    return LBMIB_ERR_ARG;
  const bool pairs = cells_per_thread == 2 && ctx->d_rowtab && pair_path_ok(ctx->g);
  const bool rows2 = cells_per_thread == 3 && ctx->d_rowtab && rows2_path_ok(ctx->g);
  if (cells_per_thread == 3 && !rows2) return LBMIB_ERR_ARG;
  if (cells_per_thread > 1 && !pairs && !rows2) {
    // the vector-per-lane kernel: only where its width fits the precision
    // and the row, one block per row, cached loads
    if (collide_vec_width(ctx->fp.precision, ctx->g.nx, cells_per_thread) != cells_per_thread) return LBMIB_ERR_ARG;
    if (blocks_per_cu > 0 || (nontemporal > 0 && (nontemporal & 2))) return LBMIB_ERR_ARG;
  }
  ctx->kv.cells_per_thread = cells_per_thread;
  ctx->kv.nt_mask = nontemporal;
  ctx->kv.blocks = blocks_per_cu * ctx->ncu;
  return LBMIB_OK;
}

extern "C" int lbmib_bytes_per_cell(lbmib_ctx *ctx, double *bytes) {
  if (!ctx || !bytes) return LBMIB_ERR_ARG;
  *bytes = 2.0 * 19.0 * (double)ctx->esz;
  return LBMIB_OK;
}

// ------------------------------------------------------- copy bandwidth

typedef unsigned int probe_u32x4 __attribute__((ext_vector_type(4)));

// Each block copies 4 x 256 consecutive 16-byte elements; every lane issues
// its four loads before its stores, so that four reads per lane are in flight.
__global__ void __launch_bounds__(256) k_copy16(const probe_u32x4 *__restrict__ a, probe_u32x4 *__restrict__ b,
                                                long n) {
  const long i0 = (long)blockIdx.x * 1024 + threadIdx.x;
  probe_u32x4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (i0 + 256 * k < n) v[k] = __builtin_nontemporal_load(a + i0 + 256 * k);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (i0 + 256 * k < n) __builtin_nontemporal_store(v[k], b + i0 + 256 * k);
}

extern "C" int lbmib_probe_copy_bandwidth(double bytes, int reps, double *gbps) {
  if (!gbps || bytes < 1e6 || reps < 1) return LBMIB_ERR_ARG;
  const long n = (long)(bytes / 2 / sizeof(probe_u32x4));
  if ((n + 1023) / 1024 > 0x7fffffffL) return LBMIB_ERR_ARG;
  probe_u32x4 *a = nullptr, *b = nullptr;
  if (hipMalloc((void **)&a, n * sizeof(probe_u32x4)) != hipSuccess) return LBMIB_ERR_ALLOC;
  if (hipMalloc((void **)&b, n * sizeof(probe_u32x4)) != hipSuccess) { (void)hipFree(a); return LBMIB_ERR_ALLOC; }
  int rc = LBMIB_OK;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = LBMIB_ERR_HIP;
  double best = 0;
  if (!rc) {
    (void)hipMemset(a, 1, n * sizeof(probe_u32x4));
    const long blocks = (n + 1023) / 1024;
    for (int r = 0; r < reps + 1 && !rc; ++r) {
      (void)hipEventRecord(e0, 0);
      k_copy16<<<(unsigned)blocks, 256, 0, 0>>>(a, b, n);
      (void)hipEventRecord(e1, 0);
      if (hipEventSynchronize(e1) != hipSuccess) { rc = LBMIB_ERR_HIP; break; }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      const double g = 2.0 * n * sizeof(probe_u32x4) / (ms * 1e-3) / 1e9;
      if (r > 0 && g > best) best = g;  // the first launch warms up
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  *gbps = best;
  return rc;
}
