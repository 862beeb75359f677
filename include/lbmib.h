/*
 * lbmib.h -- C-ABI of the MI355X LBM-IB engine (liblbmib.so).
 *
 * Plain C types only: an opaque context, ints, doubles and host pointers.
 * Every per-step entry point enqueues HIP kernels on the context's stream and
 * returns immediately; results become visible through the get/sync calls.
 *
 * Each declaration names the reference routine it replaces
 * (cube_pthread/lbm_ib.c unless a path is given).  The reference passes an
 * `LV = {int tid; GV gv}` to every per-step routine (lbm_ib.c:164-167) and
 * calls it from all P*Q*R threads; here one host thread calls each entry
 * point once with the context that replaces GV.  The host layer in
 * include/lbmib_host.h keeps the reference's own names and LV signature on
 * top of these calls.
 *
 * Errors: the reference calls exit(1) when a fiber's support leaves the grid
 * (lbm_ib.c:1458-1466, 2624-2632) and abort() on bad flags.  Here every entry
 * point returns an LBMIB_* status; device-side conditions are latched in a
 * device flag and reported by lbmib_status() and by every synchronising call.
 */
#ifndef LBMIB_H
#define LBMIB_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define LBMIB_OK 0
#define LBMIB_ERR_ARG 1          /* invalid argument / configuration */
#define LBMIB_ERR_HIP 2          /* HIP runtime call failed */
#define LBMIB_ERR_ALLOC 3        /* device or host allocation failed */
#define LBMIB_ERR_FIBER_OOB 4    /* fiber support left the grid (lbm_ib.c:1458-1466) */
#define LBMIB_ERR_BIN_OVERFLOW 5 /* more fiber nodes in one spread bin than its capacity */
#define LBMIB_ERR_NONFINITE 6    /* NaN/Inf in the mass/momentum watchdog */
#define LBMIB_ERR_STATE 7        /* call out of order (e.g. step before init) */

/* ---- boundary kinds per axis ---- */
#define LBMIB_BC_INOUT 0    /* x only: fixed-equilibrium planes x=1,2 and N-3,N-2 (lbm_ib.c:2769-2907) */
#define LBMIB_BC_WALL 1     /* half-way bounce-back at planes 2 and N-3 (lbm_ib.c:2257-2437) */
#define LBMIB_BC_PERIODIC 2 /* wrap over the whole axis */

/* ---- population storage ---- */
#define LBMIB_F64 0  /* f in FP64, FP64 arithmetic in the reference's operation order */
#define LBMIB_F32 1  /* f - r_i in FP32, FP64 collision */
#define LBMIB_F16 2  /* (f - r_i) * 2^8 in FP16, FP32 collision, FP64 reductions */
#define LBMIB_BF16 3 /* (f - r_i) * 2^8 in BF16, FP32 collision, FP64 reductions */

/* Reference populations r_i of the deviation forms (lbmib_fluid_params.dev_ref) */
#define LBMIB_DEV_EQ0 0  /* r_i = feq_i(rho0, u0), the initial flow (default) */
#define LBMIB_DEV_REST 1 /* r_i = w_i, the rest state */

/* Population layout in HBM (lbmib_fluid_params.layout); results are
 * identical in both.  The rows layout keeps the loads and stores of a row's
 * update in a few contiguous blocks and streams ~3% faster. */
#define LBMIB_LAYOUT_ROWS 0   /* f[z][y][q][x]: the 19 populations of each x row side by side (default) */
#define LBMIB_LAYOUT_PLANES 1 /* f[q][z][y][x]: 19 planes of the whole grid */

typedef struct lbmib_ctx lbmib_ctx;

/* Fluid configuration.  The reference hard-codes these in init_gv
 * (lbm_ib.c:750-773): tau = 0.5004, rho_l = 1, u_l = (0.001, 0, 0), g = 0. */
typedef struct {
  int nx, ny, nz;      /* grid; x is the inlet->outlet axis and the sheet normal */
  int bc_x, bc_y, bc_z;
  int precision;       /* LBMIB_F64 / F32 / F16 / BF16 */
  int quirks;          /* 1: reproduce the reference's two fiber-force slips */
  double tau;
  double rho0;
  double u0[3];        /* initial velocity; inlet/outlet tables are feq(rho0, u0) */
  double g;            /* gravity along z (lbm_ib.c:769 sets 0) */
  int dev_ref;         /* LBMIB_DEV_EQ0 / LBMIB_DEV_REST; ignored by LBMIB_F64 */
  int layout;          /* LBMIB_LAYOUT_ROWS (0) / LBMIB_LAYOUT_PLANES */
} lbmib_fluid_params;

/* One fiber sheet (Fibersheet, lbm_ib.c:68-76): rows along y, cols along z.
 * Spacings are ds1 = width/(cols-1), ds2 = height/(rows-1)
 * (lbm_ib.c:994-995). ks/kb are Ks_l and Kb_l. */
typedef struct {
  int rows, cols;
  double width, height;
  double ks, kb;
} lbmib_sheet_params;

/* ---- construction (gen_fluid_grid lbm_ib.c:924, gen_fiber_shape :866, init_gv :705) ---- */
int lbmib_create(const lbmib_fluid_params *fluid, int nsheets, const lbmib_sheet_params *sheets,
                 lbmib_ctx **out);
int lbmib_destroy(lbmib_ctx *ctx);
/* Uploads sheet positions, rows*cols*3 doubles in (row, col, xyz) order.
 * Replaces the coordinate fill of gen_fiber_shape (lbm_ib.c:899-906). */
int lbmib_set_fiber_positions(lbmib_ctx *ctx, int sheet, const double *xyz);

/* ---- initialisation ---- */
int lbmib_init_eqlbrmdistrfuncDF0(lbmib_ctx *ctx); /* lbm_ib.c:1528-1606 (rho/u of init_gv :837-849) */
int lbmib_init_df1(lbmib_ctx *ctx);                 /* lbm_ib.c:2725-2767 */
int lbmib_init_df_inout(lbmib_ctx *ctx);            /* lbm_ib.c:1610-1670 */

/* ---- fiber forces (one kernel per sheet) ---- */
int lbmib_compute_bendingforce(lbmib_ctx *ctx);    /* lbm_ib.c:961-1152 */
int lbmib_compute_stretchingforce(lbmib_ctx *ctx); /* lbm_ib.c:1154-1282 */
int lbmib_compute_elasticforce(lbmib_ctx *ctx);    /* lbm_ib.c:1285-1318 */
int lbmib_compute_fiber_forces(lbmib_ctx *ctx);    /* the three above in one launch */

/* ---- force spreading: deterministic cell-owned gather ---- */
int lbmib_fiber_SpreadForce(lbmib_ctx *ctx);      /* mpi/fiber_SpreadForce.c:64-283: bin nodes, support boxes */
int lbmib_fluid_get_SpreadForce(lbmib_ctx *ctx);  /* mpi/fluid_get_SpreadForce.c:28-189: cell-owned apply */
int lbmib_get_influentialdomain_fluid_grid_and_SpreadForce(lbmib_ctx *ctx); /* lbm_ib.c:1323-1525 */

/* ---- fluid step ----
 * compute_eqlbrmdistrfuncDF1 launches the fused kernel: it pulls the previous
 * step's post-collision populations (streaming, bounce-back and inlet/outlet
 * of the previous step), then collides with the lagged rho/u and the current
 * force.  The separate stream/bounce-back/inout/periodic calls are therefore
 * no-ops kept for drop-in compatibility. */
int lbmib_compute_eqlbrmdistrfuncDF1(lbmib_ctx *ctx); /* lbm_ib.c:1673-1825 (+ :1827, :2257, :2769) */
int lbmib_stream_distrfunc(lbmib_ctx *ctx);           /* lbm_ib.c:1827-2254: fused, no-op */
int lbmib_bounceback_rigidwalls(lbmib_ctx *ctx);      /* lbm_ib.c:2257-2437: fused, no-op */
int lbmib_copy_inout_to_df2(lbmib_ctx *ctx);          /* lbm_ib.c:2769-2907: fused, no-op */
int lbmib_periodicBC(lbmib_ctx *ctx);                 /* lbm_ib.c:2970-3083: wrap is fused; no-op */
int lbmib_compute_rho_and_u(lbmib_ctx *ctx);          /* lbm_ib.c:2440-2538 over the sheets' support band */
int lbmib_fluid_SpreadVelocity(lbmib_ctx *ctx);       /* mpi/fluid_SpreadVelocity.c:25-163 (= band rho/u) */
int lbmib_fiber_get_SpreadVelocity(lbmib_ctx *ctx);   /* mpi/fiber_get_SpreadVelocity.c:26-276: interpolate + move */
int lbmib_moveFiberSheet(lbmib_ctx *ctx);             /* lbm_ib.c:2540-2723 (uses the band rho/u) */
int lbmib_replace_old_DF(lbmib_ctx *ctx);             /* lbm_ib.c:2909-2968: buffer swap, time += dt */

/* ---- whole steps: the do_thread loop body, lbm_ib.c:3115-3563 ---- */
int lbmib_step(lbmib_ctx *ctx, int nsteps);
/* Split steps for lbmib_step / lbmib_time_steps (not in the reference): with
 * sheets present, each step updates the cells near the sheets' support (the
 * band) on a second stream, which then moves the sheets and computes the
 * next step's fiber forces and spreading, while the main stream updates all
 * other cells -- the IB chain overlaps the bulk of the fluid update.
 * Results are bit-identical to unsplit steps.  mode: -1 automatic (the
 * default: unsplit, which measured as fast or faster on every workload here),
 * 0 never, 1 always (with sheets and the buffer-addressed row kernels). */
int lbmib_set_overlap(lbmib_ctx *ctx, int mode);
/* Cells updated by the kernels lbmib_time_steps times per step (all cells;
 * with split steps the timed launch is the bulk, all cells but the band's,
 * under 1% of a large grid). */
int lbmib_timed_kernel_cells(lbmib_ctx *ctx, double *cells);

/* ---- queries (synchronising) ---- */
int lbmib_sync(lbmib_ctx *ctx);
int lbmib_status(lbmib_ctx *ctx);
int lbmib_get_time(lbmib_ctx *ctx, int *time);
/* Fibernode records (lbm_ib.c:54-59): rows*cols*12 doubles
 * {x,y,z, bend xyz, stretch xyz, elastic xyz}. */
int lbmib_get_fibers(lbmib_ctx *ctx, int sheet, double *out);
/* Full-field rho (n) and velocity (3n) after the last step, with the
 * reference's frozen planes (x = 2, N-3 keep their initial values). */
int lbmib_get_macroscopic(lbmib_ctx *ctx, double *rho, double *vel);
/* Post-stream populations df2 (19n, = df1 after replace_old_DF) as f,
 * on active cells; zero on dead/ghost cells. */
int lbmib_get_df(lbmib_ctx *ctx, double *df);
/* Force field of the last spread (3n), zero outside the sheets' band. */
int lbmib_get_force(lbmib_ctx *ctx, double *force);
/* FP64 deterministic totals over active cells of the post-stream state:
 * out = {mass, momentum x, y, z}.  Non-finite totals latch LBMIB_ERR_NONFINITE. */
int lbmib_get_totals(lbmib_ctx *ctx, double *out4);
/* Support box of the last spread for one sheet: {x0,x1,y0,y1,z0,z1}. */
int lbmib_get_band_box(lbmib_ctx *ctx, int sheet, int *box6);

/* ---- measurement ----
 * Runs nsteps full steps after nwarm untimed ones and reports, from HIP
 * events on the context's stream: total milliseconds, and the mean duration
 * of the fused collide-stream kernel (timed in a separate pass of the same
 * length with one event pair per launch). */
int lbmib_time_steps(lbmib_ctx *ctx, int nwarm, int nsteps, double *ms_total, double *ms_main_mean);
/* Launch shape of the fused collide-stream kernel.  cells_per_thread:
 *   0 automatic (2 for the 16-bit forms where it applies, else 1);
 *   1 one cell per lane, one block per (y, z) row (or a persistent grid);
 *   2 two x-adjacent cells per lane (needs nx even; wall/periodic x edges
 *     then run in a small extra kernel);
 *   3 two rows (y, y+1) per block, one cell of each per lane (needs ny even);
 *   4, 8 x-vectors per lane (16 bytes per lane at most; one block per row,
 *     cached loads).
 * nontemporal: bit 0 stores, bit 1 loads; -1 automatic (both).
 * blocks_per_cu > 0: a persistent grid of that many blocks per compute unit
 * walking the rows (0 = one block per row or row pair).  Not in the
 * reference: the knobs bench.py's autotuning sweep sets.  Every variant
 * computes bit-identical results. */
int lbmib_set_kernel_variant(lbmib_ctx *ctx, int cells_per_thread, int nontemporal, int blocks_per_cu);
/* With nontemporal stores (the default), the fused kernel can store the rows
 * that meet a sheet's support box (widened by one cell in y and z) with the
 * default cache policy (cached = 1) instead (0).  -1, the default at
 * create: 1 when those rows' stores, estimated from the sheets' extents,
 * fit MI355X's 256 MB memory-side cache, else 0 -- measured at 256^3, F32
 * 4.6% and BF16 1.4% faster with them cached (216 / 108 MB), F64 and the
 * larger workloads 6-8% slower.  Not in the reference; results are the same
 * bits. */
int lbmib_set_box_row_stores(lbmib_ctx *ctx, int cached);
/* Bytes per cell update of the population storage (2 * 19 * sizeof). */
int lbmib_bytes_per_cell(lbmib_ctx *ctx, double *bytes);
/* Device copy bandwidth on the current device: a 16-byte-per-lane streaming
 * copy of `bytes` (read + write counted), best of `reps` launches timed with
 * HIP events.  Not in the reference: the measured ceiling that the fused
 * kernel's achieved bandwidth is compared against. */
int lbmib_probe_copy_bandwidth(double bytes, int reps, double *gbps);

/* ---- batched ensembles (not in the reference) ----
 * n independent simulations -- each with its own grid, fluid constants and
 * sheets -- stepped together: each kernel of a step (fiber forces,
 * spreading, fused collide-stream, band velocity, move) runs as one launch
 * over all instances, whose blocks map to (instance, block of that
 * instance).  Each instance is an lbmib_ctx (lbmib_batch_instance) for
 * positions, readbacks and status, and its results are those of lbmib_step
 * on it alone, bit for bit.  All instances share the precision and need an
 * inlet/outlet x axis and nx <= 1024.  Not in the reference: the ensemble
 * form of the workload. */
typedef struct lbmib_batch lbmib_batch;
/* sheets: the sheets of every instance in order, nsheets[i] of instance i. */
int lbmib_batch_create(int n, const lbmib_fluid_params *fluids, const int *nsheets,
                       const lbmib_sheet_params *sheets, lbmib_batch **out);
int lbmib_batch_destroy(lbmib_batch *b); /* destroys the instances too */
int lbmib_batch_instance(lbmib_batch *b, int i, lbmib_ctx **ctx);
/* the three init calls of every instance (after their fiber positions) */
int lbmib_batch_init(lbmib_batch *b);
int lbmib_batch_step(lbmib_batch *b, int nsteps);
/* Tile shape of the batched fused launch: 64-lane waves per block (one row,
 * or 64-cell chunk of a row, per wave; a block holds whole rows) and the
 * nontemporal mask as in lbmib_set_kernel_variant.  bench.py autotunes it. */
int lbmib_batch_set_tile(lbmib_batch *b, int waves_per_block, int nontemporal);
/* Rows per lane of the batched fused launch: 1 (default) or 2, where each
 * lane updates one cell of two neighbouring rows (y, y + 1) of one z plane,
 * which doubles the bytes a wave keeps in flight; 2 needs every instance's
 * ny even (else LBMIB_ERR_ARG).  Results are the same bits either way.
 * bench.py autotunes it with the tile. */
int lbmib_batch_set_rows_per_lane(lbmib_batch *b, int rows);
/* lbmib_set_box_row_stores for the batched fused launch; -1 (the default
 * at create) caches the box rows when all instances' together fit. */
int lbmib_batch_set_box_row_stores(lbmib_batch *b, int cached);
/* on = 1 (default): two groups of instances of about equal cell count, each
 * on its own stream, whose fused launches alternate so that one group's IB
 * kernels run beside the other group's fluid update; on = 0: one group.
 * Results are the same bits either way. */
int lbmib_batch_set_overlap(lbmib_batch *b, int on);
/* Grid cells and fiber nodes over all instances. */
int lbmib_batch_cells(lbmib_batch *b, double *cells, int *nodes);
/* First non-zero status of the instances (synchronising). */
int lbmib_batch_status(lbmib_batch *b);
/* nwarm untimed batched steps, then nsteps timed with HIP events (total
 * milliseconds), then, if ms_fused_mean is given, nsteps more with an event
 * pair around each batched fused launch (the mean per step of their sum over
 * the groups, in milliseconds). */
int lbmib_batch_time_steps(lbmib_batch *b, int nwarm, int nsteps, double *ms_total, double *ms_fused_mean);

/* ---- coupling of the learned surrogate (not in the reference) ----
 * The reference steps every time step (do_thread, lbm_ib.c:3115-3563).  These
 * calls let a caller read the coarse-grained state, replace K steps by a
 * predicted change of it (include/lbmib_surrogate.h trains and runs the
 * network), and undo the jump when the plain solver's re-verification
 * rejects it.  Coarse fields are device FP32 planes [c][nz/f][ny/f][nx/f];
 * f in {1, 2, 4}; nx a multiple of 64, ny and nz of 4. */
/* c = 0..3: rho, ux, uy, uz of lbmib_get_macroscopic (compute_rho_and_u,
 * lbm_ib.c:2440-2538) averaged over each f^3 block. */
int lbmib_coarse_moments(lbmib_ctx *ctx, int factor, float *d_out);
/* lbmib_coarse_moments into d_moments and, in the same pass over the
 * populations (d_stored nullable), the means of the stored post-collision
 * populations' own moments: rho and u = sum c f / rho on the updated cells
 * (the active box without the frozen inlet/outlet planes), rho0/u0 on the
 * other cells -- the convention of lbmib_jump's d_after.  With d_moments
 * null only the stored planes are computed, in a cheaper pass that reads
 * each cell's own populations and no neighbour's (same bits). */
int lbmib_coarse_state(lbmib_ctx *ctx, int factor, float *d_moments, float *d_stored);
/* One step (lbmib_step(ctx, 1)) whose fluid launch also reports the stored
 * moments of the state it writes, from the values it stores, into d_stored:
 * the populations of lbmib_step(ctx, 1) and the planes of a following
 * lbmib_coarse_state(ctx, factor, NULL, d_stored), bit for bit, without that
 * read's pass over the populations.  Where that launch does not apply (an x
 * axis other than inlet/outlet, rows wider than 1024 cells, a population
 * array over 4 GiB) it makes exactly those two calls.  The stored planes of
 * every call come from per-row partial sums (16 bytes per coarse x cell of
 * each row), which the context allocates on first use. */
int lbmib_step_coarse(lbmib_ctx *ctx, int factor, float *d_stored);
/* c = 0..3: number of fiber nodes in each coarse cell and the sums of their
 * elastic forces (compute_elasticforce, lbm_ib.c:1285-1318); deterministic. */
int lbmib_coarse_sheets(lbmib_ctx *ctx, int factor, float *d_out);
/* Advances the simulation by ksteps (even, >= 2) steps without stepping:
 * d_coarse holds coarse moments of the current state (either plane set of
 * lbmib_coarse_state; its velocity planes move the fiber nodes), d_delta the
 * predicted change of those four planes over ksteps.  Every updated cell
 * keeps its non-equilibrium part and takes the equilibrium of its moments
 * plus the change interpolated trilinearly at the cell; fiber nodes move by
 * ksteps times the mean of the coarse velocity before and after the jump at
 * the node.  d_after (nullable) receives, from the same pass, the four
 * coarse planes of the jumped state in lbmib_coarse_state's stored
 * convention.  The pre-jump state is kept until lbmib_jump_commit or
 * lbmib_jump_revert; steps may run in between (the re-verification).  Needs
 * unsplit steps and at least one completed step. */
int lbmib_jump(lbmib_ctx *ctx, int factor, const float *d_coarse, const float *d_delta, int ksteps,
               float *d_after);
/* The context's HIP stream (hipStream_t as void *), on which every entry
 * point enqueues; work of other libraries placed on it is ordered with the
 * simulation's (e.g. the surrogate's kernels between lbmib_coarse_moments
 * and lbmib_jump). */
int lbmib_get_stream(lbmib_ctx *ctx, void **stream);
/* Accepts the last jump (drops its checkpoint). */
int lbmib_jump_commit(lbmib_ctx *ctx);
/* Restores the state and time of before the last jump, whatever steps ran
 * since; the next lbmib_step continues the plain trajectory bit for bit. */
int lbmib_jump_revert(lbmib_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* LBMIB_H */
