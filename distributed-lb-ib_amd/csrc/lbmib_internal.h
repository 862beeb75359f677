// lbmib_internal.h -- device-side types shared by the HIP translation units of
// liblbmib.so.  Nothing here is part of the public C-ABI (include/lbmib.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/lbmib.h"

namespace lbmib {

// D3Q19 lattice in the reference's direction order (lbm_ib.c:777-833).
constexpr int kCX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int kCY[19] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, -1, 1, 1, -1};
constexpr int kCZ[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0};
// pairs (2k-1, 2k) are opposite (Appendix of the layout: lbm_ib.c:2310-2423)
constexpr int kOPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
// 1/w_q denominators (lbm_ib.c:1742, 1773, 1804)
constexpr int kWDEN[19] = {3, 18, 18, 18, 18, 18, 18, 36, 36, 36,
                                                36, 36, 36, 36, 36, 36, 36, 36, 36};
// weight class of each direction: 0 rest (1/3), 1 axis (1/18), 2 diagonal (1/36)
constexpr int kCLS[19] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2};
constexpr int kWDEN_CLS[3] = {3, 18, 36};

// The reference's truncated pi (lbm_ib.c:33).
constexpr double kRefPi = 3.14159265358979;

constexpr int kMaxSheets = 16;
constexpr int kDevScale16 = 256;  // 16-bit deviation storage holds (f - r) * 2^8

// Fibernode-equivalent SoA offsets inside the node force array.
enum { NF_BX = 0, NF_BY, NF_BZ, NF_SX, NF_SY, NF_SZ, NF_EX, NF_EY, NF_EZ, NF_COUNT };

// Fiber kernel task bits.
enum { FK_BEND = 1, FK_STRETCH = 2, FK_ELASTIC = 4, FK_PREP = 8, FK_ALL = 15 };

// Device error bits (combined with atomicOr, mapped to LBMIB_ERR_* on the host).
enum { DERR_FIBER_OOB = 1, DERR_BIN_OVERFLOW = 2 };

struct SheetDesc {
  int rows, cols, off;   // node range [off, off + rows*cols)
  int pad;
  double ds1, ds2;       // spacings along cols (z) and rows (y), lbm_ib.c:994-995
  double bend_c;         // Kb / ds2^3 (both passes, lbm_ib.c:996, :1065)
  double str_row_c;      // Ks / ds1^2 (lbm_ib.c:1171)
  double str_col_c;      // Ks / ds2^2 (lbm_ib.c:1235)
};

// Host copy of the fluid constants in FP64 (kernels receive KConsts).
template <class T>
struct FluidConsts {
  T omc, om;          // 1 - 1/tau, 1/tau
  T pref[19];         // dt (1 - 1/(2 tau)) / wden_q      (F64 exact form)
  T pw[19];           // w_q (1 - 1/(2 tau))               (deviation form)
  T cs2, cs4;         // cs^2, cs^4 as the reference computes them
  T g;
  T rho0, u0x, u0y, u0z;
  T rhor, urx;                // deviation forms: reference state, ref = feq(rhor, (urx, 0, 0))
  T tab_in[19], tab_out[19];  // f (F64) or f - ref (deviation forms)
  T ref[19];                  // deviation forms: reference populations; F64: 0
  const void *dtab;           // device copy of tab_in in kernel units (see KConsts::tab)
  const double *dref;         // device copy of ref
};

// Constants of the fluid kernels, passed by value, in kernel units: the F64
// form works on f; the deviation forms on h = s (f - r) with s = 1 (F32) or
// 2^8 (16-bit storage), the stored value itself, where r_q = feq_q(rhor, ur)
// is a fixed reference equilibrium with ur = (urx, 0, 0) along the inflow
// axis: by default the initial flow (rhor, urx) = (rho0, u0x), so the stored
// values are the departure from it; (1, 0) gives r_q = w_q.  With ur along x
// every reference term depends on c_x and the weight class only:
//   r_q / w_q = rhor + kr[|c_x|] + c_x orr[class] / (s om w_q)
template <class T>
struct KConsts {
  T omc, om;          // 1 - 1/tau, 1/tau
  T wa[3];            // deviation forms, per weight class: s om w
  T wb[3];            // deviation forms: s w (1 - 1/(2 tau)); F64: dt (1 - 1/(2 tau)) / wden
  T cs2, cs4, g;
  T rho0, u0x, u0y, u0z;
  T rhor, d0, mrx;    // reference density; rho0 - rhor; rhor urx
  T kr[2];            // rhor (4.5 (c_x urx)^2 - 1.5 urx^2) for c_x = 0, +-1
  T orr[3];           // per weight class: 3 wa rhor urx (odd part of s om r, c_x = +1)
  T br[3];            // per weight class: 6 s w rhor urx; s (r_opp - r_q) = -c_x br
  const T *tab;       // [19] populations of the frozen inlet/outlet planes, kernel units
  const double *ref;  // [19] reference populations r_q (deviation forms)
};

// The bounce-back correction constant br of weight class k (kernel units).
inline double host_bounce_corr(const FluidConsts<double> &c, int k, double s) {
  return k == 0 ? 0.0 : 6.0 * s * c.rhor * c.urx / (k == 1 ? 18.0 : 36.0);
}

struct Geometry {
  int nx, ny, nz;
  int bc[3];
  int lo[3], hi[3];    // active box per axis
  int mlo[3], mhi[3];  // moment box per axis
  long n;              // cells
  // population arrays: population q of cell (x, y, z) is element
  // q * pq + x + prow * (y + ny * z) (pop_index); pelems elements in all
  //   planes (LBMIB_LAYOUT_PLANES): pq = plane pitch (>= n), prow = nx
  //   rows (LBMIB_LAYOUT_ROWS):     pq = nx, prow = 19 nx
  long pq, prow, pelems;
};

// Element of population 0 of cell (x, y, z) in a population array.
__host__ __device__ inline long pop_index(const Geometry &g, int x, int y, int z) {
  return (long)x + g.prow * ((long)y + (long)g.ny * z);
}

// IB device state pointers, passed by value to the IB kernels.
struct IBState {
  int nsheets, nnodes;
  const SheetDesc *sheets;
  double *pos;       // [3][M]
  double *frc;       // [9][M]
  int *start;        // [3][M]; start[0] < 0 marks an invalid node
  double *wgt;       // [12][M]: 1+cos for x(4), y(4), z(4)
  int *bin_count;    // [nbins]
  int4 *bin_list;    // [nbins][cap] (node id, support start x, y, z)
  int nby, nbz, cap;
  int *box;          // [2][kMaxSheets][6]
  unsigned *done;    // blocks of k_fiber finished (reset by the last one)
  int *err;
  int quirks;
  double *ub;        // [3][n] fluid velocity on the support boxes (k_band_u -> k_move)
  int *band;         // [2][kRectBufInts] band boxes of split steps by step parity (written by k_fiber)
};

// host-side launchers (defined in lbmib_fluid.hip / lbmib_ib.hip)

// k_fiber: four lanes per node, kFiberThreads per block; each block folds its
// nodes' support boxes into the box slot with atomic min/max.
constexpr int kFiberThreads = 256;  // four lanes per node: 64 nodes per block
__host__ __device__ inline int fiber_blocks(int nnodes) { return (4 * nnodes + kFiberThreads - 1) / kFiberThreads; }
// Grids of the other IB kernels for a sheet set of nnodes nodes (the kernels
// stride over any further work).
__host__ __device__ inline int gather_blocks(int nnodes) {
  // about one 4x4-bin tile per 16 nodes at unit spacing; a multiple of 8
  // (the XCD count) for k_gather's tile mapping
  const int b = (nnodes / 8 + 7) / 8 * 8;
  return b < 16 ? 16 : (b > 4096 ? 4096 : b);
}
__host__ __device__ inline int band_u_blocks(int nnodes) {
  // a flat rows x cols sheet at unit spacing has a 5 x (rows+4) x (cols+4)
  // support box: at most 7 cells per node from 20 x 20 nodes up, so one pass
  const int b = (7 * nnodes + 255) / 256;
  return b < 16 ? 16 : (b > 8192 ? 8192 : b);
}
__host__ __device__ inline int move_blocks(int nnodes) { return (nnodes + 3) / 4; }  // one wave per node
// k_move_tiles: one workgroup per tile of kMvNodes x kMvNodes nodes, the
// tiles of all sheets sheet after sheet, row-major within a sheet.
constexpr int kMvNodes = 4;
inline int move_tile_count(const SheetDesc *sd, int nsheets) {
  int t = 0;
  for (int s = 0; s < nsheets; ++s)
    t += ((sd[s].rows + kMvNodes - 1) / kMvNodes) * ((sd[s].cols + kMvNodes - 1) / kMvNodes);
  return t;
}

// One instance of a batched ensemble, as the batched kernels read it.
struct BatchDesc {
  Geometry g;
  IBState ib;
  const void *f[2];         // population buffers (ping-pong)
  void *F[2];               // force fields by step parity
  const uint32_t *rowtab;   // row source table (row_source_table)
  int nsheets, mcap;        // sheets; gather candidate capacity
  int wpr;                  // waves per row of the fused launch (ceil(nx / 64))
  int nrows;                // ny * nz
  double u0[3], gz;         // init_gv velocity, gravity
  const double *ref;        // reference populations of the deviation forms
  alignas(16) unsigned char kc[256];  // KConsts<T> of the batch's precision
};
// Block map of a batched launch: one int per block, instance << kBatchShift |
// block index within the instance.
constexpr int kBatchShift = 20;
double box_row_bytes(const ::lbmib_ctx *ctx);  // lbmib_api.hip
int box_rows_auto(const ::lbmib_ctx *ctx);

int launch_collide_batch(int prec, const BatchDesc *d, const int *map, int nblocks, int wpb, int cur, int own,
                         int slot, int nt_mask, int rows, int box_rows_cached, hipStream_t s);
int launch_fiber_batch(const BatchDesc *d, const int *map, int nblocks, int slot, hipStream_t s);
int launch_gather_batch(int prec, const BatchDesc *d, const int *map, int nblocks, int slot, int mcap_max,
                        hipStream_t s);
int launch_band_u_batch(int prec, const BatchDesc *d, const int *map, int nblocks, int cur_new, int slot,
                        hipStream_t s);
int launch_move_batch(const BatchDesc *d, const int *map, int nblocks, int reset_slot, hipStream_t s);
int gather_mcap(int cap);
int kconsts_bytes(int prec, const FluidConsts<double> &c, unsigned char out[256]);
// Copies tab_in (19), the pair constants KConsts::pq (20) -- in the kernels'
// units and arithmetic type -- and ref (19 doubles) to the device; d_tab has
// room for 19 doubles, d_pq for 20, d_ref for 19.
int upload_fluid_tables(int prec, const FluidConsts<double> &c, void *d_tab, double *d_ref);

int launch_fill_init(int prec, const Geometry &g, void *f, const FluidConsts<double> &c, hipStream_t s);
// Launch shape of the fused kernel (all forms compute identical bits).
// Bytes of one step's population stores that the fused kernel may keep in
// the default cache policy (the rows meeting the support boxes): about
// MI355X's 256 MB memory-side Infinity Cache.  Measured at 256^3: caching
// 216 MB (F32) or 108 MB (BF16) of box rows made the step 4.6% / 1.4%
// faster, 433 MB (F64) or more (two 128^2 sheets, the 64-instance
// ensemble) 6-8% slower.
constexpr double kBoxRowCacheBytes = 256.0 * (1 << 20);

// Rows of a grid that one sheet's support boxes meet, estimated from the
// sheet's extent (rows along y, columns along z, as the reference lays
// sheets out): the extent plus the 4-point support and the pulls' halo.
inline double box_rows_estimate(const SheetDesc &sd) {
  const double h = sd.ds2 * (sd.rows - 1), w = sd.ds1 * (sd.cols - 1);
  return (h + 7.0) * (w + 7.0);
}

struct KernelVariant {
  int cells_per_thread;  // 0 automatic, 1, 2 (x pairs), 3 (two rows per block), 4, 8 (x vectors)
  int nt_mask;           // -1 automatic; bit 0 nontemporal stores, bit 1 nontemporal loads
  int blocks;            // > 0: persistent grid of this many blocks walking rows (one cell per thread)
  int box_rows_cached;   // with nontemporal stores: rows meeting a support box stored with the default policy
                         // (1), or not (0); -1 automatic (box_rows_auto)
};

// rowtab: the row source table (row_source_table) on the device, or null for
// the general path.  sel: which cells to update -- 0 all, 1 the band of a
// split step (the cells of the band boxes in rect_buf, which the step's
// fiber kernel wrote; a grid of band_blocks blocks of 256), 2 all others;
// 1 and 2 need rowtab and rect_buf.
int launch_collide_stream(int prec, const Geometry &g, const void *fin, void *fout, int own_mode,
                          const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                          const void *Fcur, const void *Fprev, const uint32_t *rowtab, const KernelVariant &kv,
                          hipStream_t s, int sel = 0, int *rect_buf = nullptr, int band_blocks = 0);
constexpr int kRectBufInts = 1 + 6 * kMaxSheets;
bool row_path_buffered(const Geometry &g, size_t es);
// A fused step that also writes the coarse stored-moment planes of the state
// it writes (lbmib_step_coarse): whether it applies, and its launch
// (LBMIB_ERR_ARG where it does not apply).
bool collide_coarse_ok(const Geometry &g, size_t es, bool rowtab, int F);
int launch_collide_coarse(int prec, const Geometry &g, const void *fin, void *fout, int own_mode,
                          const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                          const void *Fcur, const void *Fprev, const uint32_t *rowtab, const KernelVariant &kv,
                          int F, float *part, float *stored, hipStream_t s);
bool pair_path_ok(const Geometry &g);
bool rows2_path_ok(const Geometry &g);
int row_source_table(const Geometry &g, size_t es, std::vector<uint32_t> &tab);
int collide_vec_width(int prec, int nx, int vec_req);
int launch_materialize(int prec, const Geometry &g, const void *f, int own_mode, const FluidConsts<double> &c,
                       int nsheets, const int *box, const void *F, double *df, double *rho, double *vel,
                       hipStream_t s);
int launch_totals(int prec, const Geometry &g, const void *f, int own_mode, const FluidConsts<double> &c,
                  double *partials, int nblocks_max, double *out4, hipStream_t s);
int totals_blocks(const Geometry &g);

// band: also derive the band boxes of split steps (IBState::band) in the
// launch's last block
int launch_fiber(const IBState &ib, const Geometry &g, int mask, int slot, hipStream_t s, bool band = false);
int launch_gather(int prec, const IBState &ib, const Geometry &g, int slot, void *F, hipStream_t s);
int launch_move(int prec, const IBState &ib, const Geometry &g, const void *f, const void *F,
                const FluidConsts<double> &c, int reset_slot, int ntiles, hipStream_t s);
int launch_ib_reset(const IBState &ib, int reset_bins, int reset_slot_mask, hipStream_t s);

}  // namespace lbmib
