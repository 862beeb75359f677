// lbmib_batch.hip -- batched ensembles (lbmib_batch_* of include/lbmib.h).
//
// An ensemble is a set of independent simulations, each a full lbmib_ctx
// with its own grid, fluid constants, sheets and buffers.  A batched step
// runs every kernel of the step once for a whole group of instances: the
// launch's blocks are mapped to (instance, block within the instance) by a
// table per kernel, so 64 instances of uneven size cost five launches per
// step and group instead of 320.  The instances step in lockstep and their
// results are those of lbmib_step on each alone, bit for bit: the batched
// kernels run the same per-cell and per-node code.
//
// Groups (lbmib_batch_set_overlap).  The instances are split into two groups
// of about equal cell count, each with its own descriptors, block tables and
// stream.  The fused fluid launches alternate between the groups -- A(t),
// B(t), A(t+1), ... -- through one event per group, and each group's
// latency-bound IB kernels (band velocity and move of step t, fiber forces
// and spreading of step t+1) run on its own stream beside the other group's
// fluid launch, which keeps HBM busy.  Each instance's own kernel sequence is
// unchanged.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "lbmib_ctx.h"
#include "lbmib_internal.h"

namespace {

enum { kMapFused = 0, kMapFiber, kMapGather, kMapBandU, kMapMove, kMapCount };
constexpr int kMaxGroups = 2;

struct BatchGroup {
  std::vector<int> members;  // instance indices, in group-local order
  BatchDesc *d_desc = nullptr;
  int *d_map[kMapCount] = {};
  int nblk[kMapCount] = {};
  int mcap_max = 0;              // largest gather candidate capacity of the group
  hipStream_t stream = nullptr;  // the batch stream for group 0
  hipEvent_t ev_fluid = nullptr; // last fused launch of the group done
};

}  // namespace

struct lbmib_batch {
  int n = 0;
  int prec = 0;
  std::vector<lbmib_ctx *> inst;
  std::vector<BatchDesc> desc;  // host copies, by instance
  BatchGroup grp[kMaxGroups];
  int ngroups = 0;
  int overlap = 1;    // two groups when n > 1
  int wpb = 0;        // waves per block of the fused launch (tile shape)
  int wpb_min = 0;    // widest instance row in waves
  int nt_mask = -1;   // nontemporal loads/stores of the fused launch
  int rows = 1;       // rows per lane of the fused launch (2: row pairs, every ny even)
  int box_rows = 0;   // rows meeting a support box stored with the default cache policy (set at create)
  int nnodes = 0;     // fiber nodes over all instances
  double cells = 0;   // grid cells over all instances
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t ev = nullptr, ev2 = nullptr, ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> ev_inst;
};

#define HIPCHK(x)                                \
  do {                                           \
    if ((x) != hipSuccess) return LBMIB_ERR_HIP; \
  } while (0)
#define RCHK(x)        \
  do {                 \
    int _r = (x);      \
    if (_r) return _r; \
  } while (0)

// The batch's box rows are cached when all instances' together fit.
static int batch_box_rows_auto(const lbmib_batch *b) {
  double bytes = 0;
  for (const lbmib_ctx *c : b->inst) bytes += box_row_bytes(c);
  return bytes > 0 && bytes <= kBoxRowCacheBytes ? 1 : 0;
}

static int waves_per_row(const lbmib_ctx *c) { return (c->g.nx + 63) / 64; }

static void free_group(BatchGroup &g) {
  for (int k = 0; k < kMapCount; ++k) {
    if (g.d_map[k]) (void)hipFree(g.d_map[k]);
    g.d_map[k] = nullptr;
    g.nblk[k] = 0;
  }
  if (g.d_desc) (void)hipFree(g.d_desc);
  g.d_desc = nullptr;
  g.members.clear();
  g.mcap_max = 0;
}

// Block tables of every batched launch of group g for the current tile shape.
static int build_maps(lbmib_batch *b, BatchGroup &g) {
  std::vector<int> maps[kMapCount];
  for (int li = 0; li < (int)g.members.size(); ++li) {
    const lbmib_ctx *c = b->inst[g.members[li]];
    const int rpb = b->wpb / waves_per_row(c);  // rows (row pairs when b->rows == 2) per block
    const int nrows = c->g.ny * c->g.nz / b->rows;
    const int M = c->nnodes;
    const int counts[kMapCount] = {(nrows + rpb - 1) / rpb, M ? fiber_blocks(M) : 0, M ? gather_blocks(M) : 0,
                                   M ? band_u_blocks(M) : 0, M ? move_blocks(M) : 0};
    for (int k = 0; k < kMapCount; ++k) {
      if (counts[k] >= (1 << kBatchShift)) return LBMIB_ERR_ARG;
      for (int lb = 0; lb < counts[k]; ++lb) maps[k].push_back(li << kBatchShift | lb);
    }
  }
  for (int k = 0; k < kMapCount; ++k) {
    if (g.d_map[k]) (void)hipFree(g.d_map[k]);
    g.d_map[k] = nullptr;
    g.nblk[k] = (int)maps[k].size();
    if (maps[k].empty()) continue;
    if (hipMalloc((void **)&g.d_map[k], sizeof(int) * maps[k].size()) != hipSuccess) return LBMIB_ERR_ALLOC;
    HIPCHK(hipMemcpy(g.d_map[k], maps[k].data(), sizeof(int) * maps[k].size(), hipMemcpyHostToDevice));
  }
  return LBMIB_OK;
}

// Splits the instances into groups (largest first, each to the group with
// fewer cells) and uploads each group's descriptors and block tables.
static int build_groups(lbmib_batch *b) {
  RCHK(b->stream ? (hipStreamSynchronize(b->stream) == hipSuccess ? 0 : LBMIB_ERR_HIP) : 0);
  RCHK(b->stream2 ? (hipStreamSynchronize(b->stream2) == hipSuccess ? 0 : LBMIB_ERR_HIP) : 0);
  for (BatchGroup &g : b->grp) free_group(g);
  b->ngroups = b->overlap && b->n > 1 ? 2 : 1;
  std::vector<int> order(b->n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [b](int i, int j) { return b->inst[i]->g.n > b->inst[j]->g.n; });
  double load[kMaxGroups] = {0, 0};
  for (int i : order) {
    const int k = b->ngroups == 1 || load[0] <= load[1] ? 0 : 1;
    b->grp[k].members.push_back(i);
    load[k] += (double)b->inst[i]->g.n;
  }
  for (int k = 0; k < b->ngroups; ++k) {
    BatchGroup &g = b->grp[k];
    std::sort(g.members.begin(), g.members.end());
    std::vector<BatchDesc> d;
    for (int i : g.members) {
      d.push_back(b->desc[i]);
      g.mcap_max = std::max(g.mcap_max, b->desc[i].mcap);
    }
    if (hipMalloc((void **)&g.d_desc, sizeof(BatchDesc) * d.size()) != hipSuccess) return LBMIB_ERR_ALLOC;
    HIPCHK(hipMemcpy(g.d_desc, d.data(), sizeof(BatchDesc) * d.size(), hipMemcpyHostToDevice));
    g.stream = k == 0 ? b->stream : b->stream2;
    RCHK(build_maps(b, g));
  }
  return LBMIB_OK;
}

extern "C" int lbmib_batch_destroy(lbmib_batch *b) {
  if (!b) return LBMIB_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  for (lbmib_ctx *c : b->inst)
    if (c) (void)lbmib_destroy(c);
  for (BatchGroup &g : b->grp) {
    free_group(g);
    if (g.ev_fluid) (void)hipEventDestroy(g.ev_fluid);
  }
  for (hipEvent_t e : {b->ev, b->ev2, b->ev0, b->ev1})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->ev_inst)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t st : {b->stream, b->stream2})
    if (st) (void)hipStreamDestroy(st);
  delete b;
  return LBMIB_OK;
}

extern "C" int lbmib_batch_create(int n, const lbmib_fluid_params *fluids, const int *nsheets,
                                  const lbmib_sheet_params *sheets, lbmib_batch **out) {
  if (!out || n <= 0 || n >= (1 << (31 - kBatchShift)) || !fluids || !nsheets) return LBMIB_ERR_ARG;
  *out = nullptr;
  lbmib_batch *b = new lbmib_batch();
  b->n = n;
  b->prec = fluids[0].precision;
  int rc = LBMIB_OK, off = 0;
  for (int i = 0; i < n && !rc; ++i) {
    // the batched fused kernel has no x-edge pass: inlet/outlet x only
    if (fluids[i].precision != b->prec || fluids[i].bc_x != LBMIB_BC_INOUT || nsheets[i] < 0) rc = LBMIB_ERR_ARG;
    lbmib_ctx *c = nullptr;
    if (!rc) rc = lbmib_create(fluids + i, nsheets[i], nsheets[i] ? sheets + off : nullptr, &c);
    if (c) b->inst.push_back(c);
    off += nsheets[i];
    if (!rc && !c->d_rowtab) rc = LBMIB_ERR_ARG;  // needs the buffer-addressed row path
  }
  if (rc) {
    lbmib_batch_destroy(b);
    return rc;
  }
  b->desc.resize(n);
  for (int i = 0; i < n; ++i) {
    lbmib_ctx *c = b->inst[i];
    BatchDesc &d = b->desc[i];
    memset(&d, 0, sizeof(d));
    d.g = c->g;
    d.ib = c->ib;
    d.ib.band = nullptr;  // no split steps in a batch
    d.f[0] = c->f[0];
    d.f[1] = c->f[1];
    d.F[0] = c->F[0];
    d.F[1] = c->F[1];
    d.rowtab = c->d_rowtab;
    d.nsheets = c->nsheets;
    d.mcap = gather_mcap(c->ib.cap);
    d.wpr = waves_per_row(c);
    d.nrows = c->g.ny * c->g.nz;
    d.u0[0] = c->c.u0x;
    d.u0[1] = c->c.u0y;
    d.u0[2] = c->c.u0z;
    d.gz = c->c.g;
    d.ref = c->d_ref;
    kconsts_bytes(b->prec, c->c, d.kc);
    b->wpb_min = std::max(b->wpb_min, d.wpr);
    b->nnodes += c->nnodes;
    b->cells += (double)c->g.n;
  }
  b->wpb = b->wpb_min;
  for (hipStream_t *st : {&b->stream, &b->stream2})
    if (!rc && hipStreamCreateWithFlags(st, hipStreamNonBlocking) != hipSuccess) rc = LBMIB_ERR_HIP;
  for (hipEvent_t *e : {&b->ev, &b->ev2, &b->ev0, &b->ev1})
    if (!rc && hipEventCreate(e) != hipSuccess) rc = LBMIB_ERR_HIP;
  for (BatchGroup &g : b->grp)
    if (!rc && hipEventCreateWithFlags(&g.ev_fluid, hipEventDisableTiming) != hipSuccess) rc = LBMIB_ERR_HIP;
  b->ev_inst.assign(n, nullptr);
  for (int i = 0; i < n && !rc; ++i)
    if (hipEventCreateWithFlags(&b->ev_inst[i], hipEventDisableTiming) != hipSuccess) rc = LBMIB_ERR_HIP;
  if (!rc) {
    b->box_rows = batch_box_rows_auto(b);
    rc = build_groups(b);
  }
  if (rc) {
    lbmib_batch_destroy(b);
    return rc;
  }
  *out = b;
  return LBMIB_OK;
}

extern "C" int lbmib_batch_instance(lbmib_batch *b, int i, lbmib_ctx **ctx) {
  if (!b || !ctx || i < 0 || i >= b->n) return LBMIB_ERR_ARG;
  *ctx = b->inst[i];
  return LBMIB_OK;
}

extern "C" int lbmib_batch_init(lbmib_batch *b) {
  if (!b) return LBMIB_ERR_ARG;
  for (lbmib_ctx *c : b->inst) {
    RCHK(lbmib_init_eqlbrmdistrfuncDF0(c));
    RCHK(lbmib_init_df1(c));
    RCHK(lbmib_init_df_inout(c));
  }
  for (lbmib_ctx *c : b->inst) RCHK(lbmib_sync(c));
  return LBMIB_OK;
}

extern "C" int lbmib_batch_set_tile(lbmib_batch *b, int waves_per_block, int nontemporal) {
  if (!b || waves_per_block < 1 || waves_per_block > 8 || nontemporal < -1 || nontemporal > 3) return LBMIB_ERR_ARG;
  for (lbmib_ctx *c : b->inst)
    if (waves_per_row(c) > waves_per_block) return LBMIB_ERR_ARG;  // a block must hold a whole row
  // the tables of launches still queued must stay valid
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipStreamSynchronize(b->stream2));
  b->wpb = waves_per_block;
  b->nt_mask = nontemporal;
  for (int k = 0; k < b->ngroups; ++k) RCHK(build_maps(b, b->grp[k]));
  return LBMIB_OK;
}

extern "C" int lbmib_batch_set_rows_per_lane(lbmib_batch *b, int rows) {
  if (!b || (rows != 1 && rows != 2)) return LBMIB_ERR_ARG;
  if (rows == 2)
    for (lbmib_ctx *c : b->inst)
      if (c->g.ny % 2) return LBMIB_ERR_ARG;  // row pairs within one z plane
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipStreamSynchronize(b->stream2));
  b->rows = rows;
  for (int k = 0; k < b->ngroups; ++k) RCHK(build_maps(b, b->grp[k]));
  return LBMIB_OK;
}

extern "C" int lbmib_batch_set_box_row_stores(lbmib_batch *b, int cached) {
  if (!b || cached < -1 || cached > 1) return LBMIB_ERR_ARG;
  b->box_rows = cached < 0 ? batch_box_rows_auto(b) : cached;
  return LBMIB_OK;
}

extern "C" int lbmib_batch_set_overlap(lbmib_batch *b, int on) {
  if (!b || on < 0 || on > 1) return LBMIB_ERR_ARG;
  b->overlap = on;
  return build_groups(b);
}

extern "C" int lbmib_batch_cells(lbmib_batch *b, double *cells, int *nodes) {
  if (!b) return LBMIB_ERR_ARG;
  if (cells) *cells = b->cells;
  if (nodes) *nodes = b->nnodes;
  return LBMIB_OK;
}

// Every instance must be at the same step, ready for a whole step (no
// per-step entry point half done).
static int batch_state_ok(const lbmib_batch *b) {
  const lbmib_ctx *c0 = b->inst[0];
  for (const lbmib_ctx *c : b->inst) {
    if (c->init_bits != 7 || c->pending || !c->ib_clean || c->split_run || c->ib_ready) return LBMIB_ERR_STATE;
    if (c->time != c0->time || c->cur != c0->cur || c->own != c0->own) return LBMIB_ERR_STATE;
  }
  return LBMIB_OK;
}

// e0/e1 bracket the whole run; with kev, events kev[2(kG + g)], kev[2(kG + g) + 1]
// bracket the fused launch of group g in step k (G groups).
static int batch_steps(lbmib_batch *b, int nsteps, hipEvent_t e0, hipEvent_t e1, const hipEvent_t *kev = nullptr) {
  RCHK(batch_state_ok(b));
  const int G = b->ngroups;
  // the batch streams start after the instances' own queued work
  for (int i = 0; i < b->n; ++i) {
    HIPCHK(hipEventRecord(b->ev_inst[i], b->inst[i]->stream));
    for (int k = 0; k < G; ++k) HIPCHK(hipStreamWaitEvent(b->grp[k].stream, b->ev_inst[i], 0));
  }
  lbmib_ctx *c0 = b->inst[0];
  int cur = c0->cur, own = c0->own, time = c0->time, slot = 0;
  if (e0) {
    HIPCHK(hipEventRecord(e0, b->stream));
    if (G > 1) HIPCHK(hipStreamWaitEvent(b->stream2, e0, 0));
  }
  for (int k = 0; k < nsteps; ++k) {
    slot = time & 1;
    for (int gi = 0; gi < G; ++gi) {
      BatchGroup &g = b->grp[gi];
      hipStream_t st = g.stream;
      RCHK(launch_fiber_batch(g.d_desc, g.d_map[kMapFiber], g.nblk[kMapFiber], slot, st));
      RCHK(launch_gather_batch(b->prec, g.d_desc, g.d_map[kMapGather], g.nblk[kMapGather], slot, g.mcap_max, st));
      // fused launches alternate: A(t) after B(t-1), B(t) after A(t)
      if (G > 1 && (gi > 0 || k > 0)) HIPCHK(hipStreamWaitEvent(st, b->grp[(gi + G - 1) % G].ev_fluid, 0));
      if (kev) HIPCHK(hipEventRecord(kev[2 * (k * G + gi)], st));
      RCHK(launch_collide_batch(b->prec, g.d_desc, g.d_map[kMapFused], g.nblk[kMapFused], b->wpb, cur, own, slot,
                                b->nt_mask, b->rows, b->box_rows, st));
      if (kev) HIPCHK(hipEventRecord(kev[2 * (k * G + gi) + 1], st));
      if (G > 1) HIPCHK(hipEventRecord(g.ev_fluid, st));
      RCHK(launch_band_u_batch(b->prec, g.d_desc, g.d_map[kMapBandU], g.nblk[kMapBandU], 1 - cur, slot, st));
      RCHK(launch_move_batch(g.d_desc, g.d_map[kMapMove], g.nblk[kMapMove], 1 - slot, st));
    }
    cur = 1 - cur;
    own = 0;
    ++time;
  }
  if (G > 1) {
    HIPCHK(hipEventRecord(b->ev2, b->stream2));
    HIPCHK(hipStreamWaitEvent(b->stream, b->ev2, 0));
  }
  if (e1) HIPCHK(hipEventRecord(e1, b->stream));
  // the instances' own streams (readbacks, status) follow the batch
  HIPCHK(hipEventRecord(b->ev, b->stream));
  for (lbmib_ctx *c : b->inst) {
    HIPCHK(hipStreamWaitEvent(c->stream, b->ev, 0));
    c->cur = cur;
    c->own = own;
    c->time = time;
    c->pending = 0;
    c->band_done = 0;
    c->ib_clean = 1;
    if (nsteps > 0 && c->nnodes > 0) c->spread_slot = slot;
  }
  return LBMIB_OK;
}

extern "C" int lbmib_batch_step(lbmib_batch *b, int nsteps) {
  if (!b || nsteps < 0) return LBMIB_ERR_ARG;
  return batch_steps(b, nsteps, nullptr, nullptr);
}

extern "C" int lbmib_batch_status(lbmib_batch *b) {
  if (!b) return LBMIB_ERR_ARG;
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipStreamSynchronize(b->stream2));
  for (lbmib_ctx *c : b->inst) {
    const int rc = lbmib_status(c);
    if (rc) return rc;
  }
  return LBMIB_OK;
}

extern "C" int lbmib_batch_time_steps(lbmib_batch *b, int nwarm, int nsteps, double *ms_total,
                                      double *ms_fused_mean) {
  if (!b || nwarm < 0 || nsteps <= 0) return LBMIB_ERR_ARG;
  RCHK(batch_steps(b, nwarm, nullptr, nullptr));
  RCHK(batch_steps(b, nsteps, b->ev0, b->ev1));
  HIPCHK(hipEventSynchronize(b->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (ms_total) *ms_total = ms;
  if (ms_fused_mean) {
    // a second pass of the same length with an event pair per fused launch;
    // a step's fused time is the sum over the groups (their fused launches
    // run one after the other)
    const int G = b->ngroups;
    std::vector<hipEvent_t> kev(2 * (size_t)nsteps * G, nullptr);
    int rc = LBMIB_OK;
    for (auto &e : kev)
      if (!rc && hipEventCreate(&e) != hipSuccess) rc = LBMIB_ERR_HIP;
    if (!rc) rc = batch_steps(b, nsteps, nullptr, nullptr, kev.data());
    double acc = 0;
    if (!rc && hipStreamSynchronize(b->stream) != hipSuccess) rc = LBMIB_ERR_HIP;
    for (int k = 0; k < nsteps * G && !rc; ++k) {
      float t = 0;
      if (hipEventElapsedTime(&t, kev[2 * k], kev[2 * k + 1]) != hipSuccess) rc = LBMIB_ERR_HIP;
      acc += t;
    }
    for (auto e : kev)
      if (e) (void)hipEventDestroy(e);
    RCHK(rc);
    *ms_fused_mean = acc / nsteps;
  }
  return lbmib_batch_status(b);
}
