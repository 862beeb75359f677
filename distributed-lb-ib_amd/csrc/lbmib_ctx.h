// lbmib_ctx.h -- the engine context behind the opaque lbmib_ctx of the
// C-ABI, shared by the translation units that implement the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "lbmib_internal.h"

using namespace lbmib;

struct lbmib_ctx {
  lbmib_fluid_params fp;
  int nsheets = 0;
  lbmib_sheet_params sp[kMaxSheets];
  SheetDesc sd[kMaxSheets];
  int nnodes = 0;
  int move_tiles = 0;  // workgroups of k_move_tiles (move_tile_count)
  Geometry g;
  FluidConsts<double> c;
  size_t esz = 0, fsz = 0;  // population / force element sizes
  void *f[2] = {nullptr, nullptr};
  int cur = 0;
  int own = 1;       // f[cur] holds the initial f (not a post-collision state)
  int pending = 0;   // collide done, new state in f[1-cur], swap pending
  int band_done = 0; // band moments of the pending state computed
  int ib_clean = 1;  // bins empty and this step's box slot reset
  int spread_slot = -1;  // slot of the last completed gather
  void *F[2] = {nullptr, nullptr};
  IBState ib;
  SheetDesc *d_sheets = nullptr;
  int *d_err = nullptr;
  double *d_partials = nullptr;
  int nparts = 0;
  double *d_tot = nullptr;
  double *d_mat = nullptr;   // materialisation scratch (lazy)
  void *d_tab = nullptr;     // frozen-plane populations in kernel units (19 values)
  double *d_ref = nullptr;   // reference populations of the deviation forms
  uint32_t *d_rowtab = nullptr;  // row source table of the buffer-addressed fused kernel
  size_t mat_bytes = 0;
  hipStream_t stream = nullptr;
  int time = 1;
  int init_bits = 0;  // 1: DF0, 2: df1, 4: inout
  int latched = LBMIB_OK;
  KernelVariant kv = {0, -1, 0, 0};  // fused-kernel launch shape (box_rows_cached set at create)
  int ncu = 0;                    // compute units of the device
  // split steps (lbmib_set_overlap): the bulk on `stream`; the band, band
  // velocity, move and the next step's fiber forces and spreading on
  // `stream_b`, beside the bulk
  int overlap_req = -1;           // -1 automatic (unsplit), 0 off, 1 on
  hipStream_t stream_b = nullptr;
  hipEvent_t evFib = nullptr, evBulk = nullptr, evJoin = nullptr;  // split steps
  int split_run = 0;              // steps taken in the current run of split steps
  int ib_ready = 0;               // fiber forces and spreading of the current step done (stream_b)
  int *d_rect = nullptr;          // [2][kRectBufInts] band boxes of split steps by step parity (IBState::band)
  double hyz[kMaxSheets][4] = {};      // host y/z extent of each sheet's nodes (from the uploaded positions)
  std::vector<hipEvent_t> evs;
  // surrogate jumps (lbmib_jump.hip), allocated on first use: the population
  // buffer a jump writes into (afterwards the pre-jump checkpoint), the
  // pre-jump node positions, the previous step's force slot inside its boxes
  // and those boxes, and the sheet raster's fixed-point accumulators
  void *f_spare = nullptr;
  double *pos_save = nullptr;
  void *F_save = nullptr;
  int *box_save = nullptr;
  unsigned long long *d_raster = nullptr;
  long raster_cells = 0;
  float *d_cpart = nullptr;  // per-row partial sums of the stored planes (lbmib_coarse.h)
  size_t cpart_bytes = 0;
  int jump_active = 0;      // a jump awaits lbmib_jump_commit / lbmib_jump_revert
  int jump_time = 0;        // time and spread slot before the jump
  int jump_spread_slot = -1;
};

namespace lbmib {
// The context's partial-sum buffer for stored planes at factor F (lbmib_coarse.h),
// grown on demand.
int ensure_cpart(::lbmib_ctx *ctx, int F);
}  // namespace lbmib
