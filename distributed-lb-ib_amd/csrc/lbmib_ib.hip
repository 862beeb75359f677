// lbmib_ib.hip -- immersed-boundary kernels for gfx950.
//
// Per step (the order of do_thread, lbm_ib.c:3124-3412):
//   k_fiber        four lanes per fiber node: bending, stretching and
//                  elastic force (FP64) on one, the node's 4-point-delta
//                  support start floor(X-2)+1 and 1+cos(pi r/2) factors on
//                  one lane per axis; its (y,z) bin and the sheet's support
//                  box.  One launch covers every sheet.
//   k_gather       cell-owned force spreading: each workgroup owns a 4x4
//                  (y,z) column tile of a sheet's support box, loads the
//                  candidate node ids of the 2x2 bins that can reach it into
//                  LDS, rank-sorts them, and every cell sums F_node * delta
//                  over the candidates in ascending (sheet, row, col) order --
//                  the order of the sequential spread (lbm_ib.c:1420-1495).
//                  No float atomics; the result is run-to-run identical.
//   k_move_tiles   one workgroup per 4x4-node tile of a sheet: rho/u of the
//                  new populations (compute_rho_and_u, lbm_ib.c:2440-2538)
//                  once per cell of the tile's support region into LDS, then
//                  u(X) = Sum u delta over each node's 64 support cells in the
//                  reference's order (lbm_ib.c:2670-2673) and X += dt u(X),
//                  FP64; also clears the bins and the box slot that the next
//                  step will fill.  The batched ensemble uses the two-kernel
//                  form of the same arithmetic: k_band_u (rho/u on each
//                  support-box cell) and k_move (one wave per node).
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdlib.h>

#include "lbmib_device.h"

namespace lbmib {

__device__ __forceinline__ int find_sheet(const IBState &ib, int n) {
  int s = 0;
  for (int k = 1; k < ib.nsheets; ++k)
    if (n >= ib.sheets[k].off) s = k;
  return s;
}

// --------------------------------------------------------------- k_fiber

// The stencils below read a node's line neighbourhood from registers:
// W[k + 2] = coordinate of line element j + k (k = -2..2; entries outside
// the line are zero and never used).

// Row stencil of compute_bendingforce (lbm_ib.c:1004-1054), one component.
__device__ __forceinline__ double bend_w(const double W[5], int j, int C, double bc, bool slip) {
  if (j == 0) return bc * (-W[4] + 2 * W[3] - W[2]);
  if (j == 1) return bc * (-W[4] + 4 * W[3] - 5 * W[2] + 2 * W[1]);
  if (j == C - 2) return bc * (2 * W[3] - 5 * W[2] + 4 * W[1] - (slip ? W[2] : W[0]));
  if (j == C - 1) return bc * (-W[2] + 2 * W[1] - W[0]);
  return bc * (-W[4] + 4.0 * W[3] - 6.0 * W[2] + 4.0 * W[1] - W[0]);
}

// Spring term of compute_stretchingforce for one line (lbm_ib.c:1188-1221,
// :1236-1276): the three components of the force on element j.
__device__ __forceinline__ void stretch_w(const double W[3][5], int j, int C, double ds, double sc, bool quirk,
                                          double out[3]) {
  if (j > 0 && j < C - 1) {
    double tr[3], tl[3];
    for (int d = 0; d < 3; ++d) {
      tr[d] = W[d][3] - W[d][2];
      tl[d] = W[d][1] - W[d][2];
    }
    const double dr = sqrt(tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2]);
    const double dl = sqrt(tl[0] * tl[0] + tl[1] * tl[1] + tl[2] * tl[2]);
    for (int d = 0; d < 3; ++d) out[d] = sc * ((dr - ds) * (tr[d] / dr) + (dl - ds) * (tl[d] / dl));
  } else if (j == 0) {
    double tr[3];
    for (int d = 0; d < 3; ++d) tr[d] = W[d][3] - W[d][2];
    const double dr = sqrt(tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2]);
    for (int d = 0; d < 3; ++d) out[d] = sc * ((dr - ds) * (tr[d] / dr));
  } else {
    // last node: the reference's tangent points away from the neighbour
    // (lbm_ib.c:1215-1217, :1269-1271); quirk=0 points it at the neighbour.
    double tl[3];
    for (int d = 0; d < 3; ++d) tl[d] = quirk ? W[d][2] - W[d][1] : W[d][1] - W[d][2];
    const double dl = sqrt(tl[0] * tl[0] + tl[1] * tl[1] + tl[2] * tl[2]);
    for (int d = 0; d < 3; ++d) out[d] = sc * ((dl - ds) * (tl[d] / dl));
  }
}

// Warp-level min/max so that only one lane per wave touches the shared
// support box (same-address atomics from every node serialise in L2).
__device__ __forceinline__ int wave_min(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// Four lanes per node (a quad of consecutive lanes): lane 0 of the quad
// computes the node forces, lane 1 + d the support start and the four delta
// factors along axis d, so that the FP64 force and cosine chains run side
// by side.  The quad exchanges the starts by lane shuffles for the node's
// validity, bin entry and support box.  Every value is the same expression
// as for one lane per node.
__device__ __forceinline__ void fiber_body(const IBState &ib, const Geometry &g, int mask, int slot, int bid,
                                           int nblk) {
  const int t = bid * blockDim.x + threadIdx.x;
  const int n = t >> 2, role = t & 3;
  const bool live = n < ib.nnodes;
  const int M = ib.nnodes;
  const double *__restrict__ pos = ib.pos;
  double *__restrict__ frc = ib.frc;
  int s = 0;
  bool ok = false;
  int st[3] = {INT_MAX, INT_MAX, INT_MAX};
  if (live) s = find_sheet(ib, n);
  if (live && role == 0 && (mask & (FK_BEND | FK_STRETCH | FK_ELASTIC))) {
    const SheetDesc sd = ib.sheets[s];
    const int k = n - sd.off, i = k / sd.cols, j = k % sd.cols;
    const bool quirk = ib.quirks != 0;
    // the node's row (along cols) and column (along rows) neighbourhoods,
    // loaded once
    double Wr[3][5], Wc[3][5];
    for (int d = 0; d < 3; ++d) {
      const double *line = pos + (long)d * M + sd.off;
      for (int o = -2; o <= 2; ++o) {
        const int jj = j + o, ii = i + o;
        Wr[d][o + 2] = (jj >= 0 && jj < sd.cols) ? line[(long)i * sd.cols + jj] : 0.0;
        Wc[d][o + 2] = (ii >= 0 && ii < sd.rows) ? line[(long)ii * sd.cols + j] : 0.0;
      }
    }
    double bend[3] = {0, 0, 0}, str[3] = {0, 0, 0};
    if (mask & (FK_BEND | FK_ELASTIC)) {
      for (int d = 0; d < 3; ++d) {
        const double r = bend_w(Wr[d], j, sd.cols, sd.bend_c, quirk && d == 0);
        const double c = bend_w(Wc[d], i, sd.rows, sd.bend_c, false);
        bend[d] = r + c;  // row pass '=', column pass '+=' (lbm_ib.c:1070)
      }
    }
    if (mask & (FK_STRETCH | FK_ELASTIC)) {
      double r[3], c[3];
      stretch_w(Wr, j, sd.cols, sd.ds1, sd.str_row_c, quirk, r);
      stretch_w(Wc, i, sd.rows, sd.ds2, sd.str_col_c, quirk, c);
      for (int d = 0; d < 3; ++d) str[d] = r[d] + c[d];
    }
    if (mask & FK_BEND)
      for (int d = 0; d < 3; ++d) frc[(NF_BX + d) * (long)M + n] = bend[d];
    if (mask & FK_STRETCH)
      for (int d = 0; d < 3; ++d) frc[(NF_SX + d) * (long)M + n] = str[d];
    if (mask & FK_ELASTIC) {
      // elastic = stretch + bend of the stored node forces (lbm_ib.c:1300-1315)
      for (int d = 0; d < 3; ++d) {
        const double sv = (mask & FK_STRETCH) ? str[d] : frc[(NF_SX + d) * (long)M + n];
        const double bv = (mask & FK_BEND) ? bend[d] : frc[(NF_BX + d) * (long)M + n];
        frc[(NF_EX + d) * (long)M + n] = sv + bv;
      }
    }
  }
  if (mask & FK_PREP) {
    // lane 1 + d: support start floor(X - 2) + 1 along axis d
    // (lbm_ib.c:1424-1429) and whether it lies on the grid
    const int d = role - 1;
    int sdv = INT_MAX, okd = 1;
    double Xd = 0.0;
    if (live && role > 0) {
      Xd = pos[(long)d * M + n];
      const int dim = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
      // tested in FP64 before the conversion: a NaN or a value beyond the
      // int range (which the conversion would turn into 0 or INT_MAX) fails
      const double sdd = floor(Xd / 1.0 - 2) + 1;
      okd = sdd >= 0.0 && sdd + 3.0 < (double)dim;
      if (okd) sdv = (int)sdd;
    }
    const int q0 = (threadIdx.x & 63) & ~3;  // the quad's lane 0 in the wave
    for (int a = 0; a < 3; ++a) st[a] = __shfl(sdv, q0 + 1 + a, 64);
    const int okq = __shfl(okd, q0 + 1, 64) & __shfl(okd, q0 + 2, 64) & __shfl(okd, q0 + 3, 64);
    ok = live && okq;
    if (live && !ok) {
      // the reference exits here (lbm_ib.c:1458-1466); latch the error and
      // leave the node out of spreading and moving.
      if (role == 0) {
        atomicOr(ib.err, DERR_FIBER_OOB);
        ib.start[n] = -1;
      }
    } else if (live && role > 0) {
      // 1 + cos(pi r / 2) at r = (start + a) - X, a = 0..3, each cosine
      // evaluated as the reference does (lbm_ib.c:1456, :2670-2673)
      const double cpi = kRefPi / (2.0 * 1.0);
      ib.start[d * (long)M + n] = sdv;
      for (int a = 0; a < 4; ++a) {
        const double r = 1.0 * (sdv + a) - Xd;
        ib.wgt[(4 * d + a) * (long)M + n] = 1.0e0 + cos(cpi * r);
      }
    } else if (live) {
      const int b = (st[1] >> 2) * ib.nbz + (st[2] >> 2);
      const int slot_in_bin = atomicAdd(&ib.bin_count[b], 1);
      if (slot_in_bin < ib.cap) ib.bin_list[(long)b * ib.cap + slot_in_bin] = make_int4(n, st[0], st[1], st[2]);
      else atomicOr(ib.err, DERR_BIN_OVERFLOW);
    }
  }
  if (!(mask & FK_PREP)) return;  // uniform over the grid
  // Support box of each sheet: wave min/max, LDS combine per block, then one
  // global atomic min/max per block and coordinate into the box slot (reset
  // to empty by the previous move); min/max are order-independent, so the
  // box is deterministic.  The last block to finish derives the band boxes.
  __shared__ int sbox[kMaxSheets * 6];
  __shared__ int s_last;
  for (int t = threadIdx.x; t < kMaxSheets * 6; t += blockDim.x) sbox[t] = (t & 1) ? INT_MIN : INT_MAX;
  __syncthreads();
  const int s_lo = __shfl(s, 0, 64);
  const bool uniform = __all(!live || s == s_lo);
  const int lo[3] = {ok ? st[0] : INT_MAX, ok ? st[1] : INT_MAX, ok ? st[2] : INT_MAX};
  const int hi[3] = {ok ? st[0] + 3 : INT_MIN, ok ? st[1] + 3 : INT_MIN, ok ? st[2] + 3 : INT_MIN};
  if (uniform) {
    int rlo[3], rhi[3];
    for (int d = 0; d < 3; ++d) {
      rlo[d] = wave_min(lo[d]);
      rhi[d] = wave_max(hi[d]);
    }
    if ((threadIdx.x & 63) == 0 && rlo[0] != INT_MAX)
      for (int d = 0; d < 3; ++d) {
        atomicMin(&sbox[s_lo * 6 + 2 * d], rlo[d]);
        atomicMax(&sbox[s_lo * 6 + 2 * d + 1], rhi[d]);
      }
  } else if (ok) {
    for (int d = 0; d < 3; ++d) {  // a wave across a sheet seam
      atomicMin(&sbox[s * 6 + 2 * d], lo[d]);
      atomicMax(&sbox[s * 6 + 2 * d + 1], hi[d]);
    }
  }
  __syncthreads();
  int *box = ib.box + slot * kMaxSheets * 6;
  const int nk = ib.nsheets * 6;
  for (int t = threadIdx.x; t < nk; t += blockDim.x) {
    if (t & 1) {
      if (sbox[t] != INT_MIN) atomicMax(box + t, sbox[t]);
    } else if (sbox[t] != INT_MAX) {
      atomicMin(box + t, sbox[t]);
    }
  }
  if (!ib.band) return;  // no band boxes wanted (unsplit steps)
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(ib.done, 1u) == (unsigned)(nblk - 1);
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  for (int k = threadIdx.x; k < nk; k += blockDim.x)
    sbox[k] = __hip_atomic_load(box + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  // the band boxes of this step for split steps: this step's boxes and the
  // previous step's (the other slot, not written by this launch)
  if (threadIdx.x == 0) {
    int *rb = ib.band + slot * kRectBufInts;
    const int *prev = ib.box + (1 - slot) * kMaxSheets * 6;
    int n = 0;
    for (int s = 0; s < ib.nsheets; ++s) {
      int b1[6], b0[6], r[6];
      for (int k = 0; k < 6; ++k) { b1[k] = sbox[s * 6 + k]; b0[k] = prev[s * 6 + k]; }
      if (!band_box_of(g, b1, b0, r)) continue;
      for (int k = 0; k < 6; ++k) rb[1 + 6 * n + k] = r[k];
      ++n;
    }
    rb[0] = n;
  }
  if (threadIdx.x == 0) atomicExch(ib.done, 0u);
}

__global__ void __launch_bounds__(kFiberThreads) k_fiber(IBState ib, Geometry g, int mask, int slot) {
  fiber_body(ib, g, mask, slot, blockIdx.x, gridDim.x);
}

// -------------------------------------------------------------- k_gather

constexpr int kGatherThreads = 128;
constexpr int kTileBins = 1;              // tile side in 4-cell bins
constexpr int kTile = 4 * kTileBins;      // tile side in cells (y and z)
constexpr int kCandBins = kTileBins + 1;  // candidate bins per axis

// Enumerates (sheet, tile) work items of the support boxes in a fixed order.
// Returns false when w is past the last item.
__device__ __forceinline__ bool tile_of(const IBState &ib, int slot, int w, int &s, int &ty, int &tz) {
  for (s = 0; s < ib.nsheets; ++s) {
    const int *b = ib.box + (slot * kMaxSheets + s) * 6;
    if (ldc(b) > ldc(b + 1)) continue;  // empty (all nodes invalid)
    const int ty0 = ldc(b + 2) / kTile, ty1 = ldc(b + 3) / kTile, tz0 = ldc(b + 4) / kTile, tz1 = ldc(b + 5) / kTile;
    const int nt = (ty1 - ty0 + 1) * (tz1 - tz0 + 1);
    if (w < nt) {
      ty = ty0 + w / (tz1 - tz0 + 1);
      tz = tz0 + w % (tz1 - tz0 + 1);
      return true;
    }
    w -= nt;
  }
  return false;
}

constexpr int kGatherMaxCand = 384;    // candidates whose support meets the tile: 13.5 KB of LDS
constexpr int kGatherStage = 128;      // candidates whose delta factors / forces are staged: 15 KB

// LDS of a gather block for mcap candidates (the layout of gather_body).
static size_t gather_shm_bytes(int mcap) {
  return sizeof(double) * 15 * (size_t)kGatherStage + sizeof(int) * (9 * (size_t)mcap + 4);
}

template <class FT>
__device__ __forceinline__ void gather_body(const IBState &ib, const Geometry &g, int slot, FT *__restrict__ F, int mcap,
                                            int bid, int nblk) {
  // LDS: delta factors and forces of the staged candidates, the sorted
  // candidates' support starts and ids (x, y, z, id), the raw candidates'
  // starts and the raw candidate ids (padded to a multiple of 4 for the
  // 16-byte reads of the rank sort)
  extern __shared__ __attribute__((aligned(16))) double smem_d[];
  double *sw = smem_d;                      // [kGatherStage][12] delta factors
  double *se = sw + kGatherStage * 12;      // [kGatherStage][3] elastic force
  int4 *sst = reinterpret_cast<int4 *>(se + kGatherStage * 3);
  int4 *rst = sst + mcap;  // support starts of the raw candidates
  int *raw = reinterpret_cast<int *>(rst + mcap);
  __shared__ int s_tot;
  const int M = ib.nnodes;
  const double cf = 1.0 / (64.0 * 1.0 * 1.0 * 1.0);  // lbm_ib.c:1357
  // Workgroups are dealt to the 8 XCDs in turn: virtual index v = bid + k
  // nblk (nblk a multiple of 8) runs on XCD v % 8 and takes tile (v % 8) per
  // + v / 8, so that each XCD spreads a contiguous band of tiles and the
  // candidates that neighbouring tiles share are read through its own L2.
  int ntiles = 0;
  for (int s = 0; s < ib.nsheets; ++s) {
    const int *b = ib.box + (slot * kMaxSheets + s) * 6;
    if (ldc(b) > ldc(b + 1)) continue;
    ntiles += (ldc(b + 3) / kTile - ldc(b + 2) / kTile + 1) * (ldc(b + 5) / kTile - ldc(b + 4) / kTile + 1);
  }
  const int per = (ntiles + 7) / 8;
  for (int v = bid; v < 8 * per; v += nblk) {
    const int w = (v % 8) * per + v / 8;
    int s, ty, tz;
    if (!tile_of(ib, slot, w, s, ty, tz)) continue;
    const int *b = ib.box + (slot * kMaxSheets + s) * 6;
    const int x0 = b[0], x1 = b[1], bybox0 = b[2], bybox1 = b[3], bzbox0 = b[4], bzbox1 = b[5];
    // a node with support start j0 reaches rows j0..j0+3, so tile rows
    // kTile*ty .. kTile*ty+kTile-1 need j0 in [kTile*ty-3, kTile*ty+kTile-1]:
    // bins kTileBins*ty-1 .. kTileBins*ty+kTileBins-1.  Of those, only the
    // nodes whose x support [i0, i0+3] meets the box's x range reach a cell
    // of the tile (the others, e.g. of another sheet in the same (y, z)
    // bins, add nothing); they are gathered in any order and rank-sorted
    // below, so each cell still sums its contributions in node order.
    // The bins' counts are loaded together, then every lane takes candidates
    // of the concatenated bin lists, so that the list reads of all bins are
    // in flight at once.
    const int by0 = kTileBins * ty - 1, bz0 = kTileBins * tz - 1;
    constexpr int kNB = kCandBins * kCandBins;
    int bins[kNB], cnt[kNB];
#pragma unroll
    for (int k = 0; k < kNB; ++k) {
      const int by = by0 + k / kCandBins, bz = bz0 + k % kCandBins;
      const bool in = by >= 0 && bz >= 0 && by < ib.nby && bz < ib.nbz;
      bins[k] = in ? by * ib.nbz + bz : 0;
      cnt[k] = in ? ldc(ib.bin_count + bins[k]) : 0;
    }
    int seen = 0;
#pragma unroll
    for (int k = 0; k < kNB; ++k) {
      const int c = min(cnt[k], ib.cap);
      cnt[k] = seen;  // becomes the bin's offset in the concatenation
      seen += c;
    }
    if (threadIdx.x == 0) s_tot = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < seen; e += blockDim.x) {
      int bin = bins[0], off = 0;  // the last bin whose offset is <= e (empty bins share the next offset)
#pragma unroll
      for (int j = 1; j < kNB; ++j)
        if (e >= cnt[j]) { bin = bins[j]; off = cnt[j]; }
      const int4 q = ib.bin_list[(long)bin * ib.cap + (e - off)];  // id, support start
      if (q.y <= x1 && q.y + 3 >= x0) {
        const int k = atomicAdd(&s_tot, 1);
        if (k < mcap) {  // LDS holds up to mcap candidates; more is reported
          raw[k] = q.x;
          rst[k] = make_int4(q.y, q.z, q.w, q.x);
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_tot > mcap) atomicOr(ib.err, DERR_BIN_OVERFLOW);
    const int tot = min(s_tot, mcap), tot4 = (tot + 3) & ~3;
    const bool staged = tot <= kGatherStage;
    if ((int)threadIdx.x < tot4 - tot) {
      // padding up to a multiple of 4: an id above every id for the rank
      // sort; a support start far from every cell and zero factors and
      // forces for the unrolled sum
      const int k = tot + threadIdx.x;
      raw[k] = INT_MAX;
      if (staged) {
        sst[k] = make_int4(-(1 << 28), -(1 << 28), -(1 << 28), 0);
        for (int a = 0; a < 12; ++a) sw[k * 12 + a] = 0.0;
        for (int a = 0; a < 3; ++a) se[k * 3 + a] = 0.0;
      }
    }
    __syncthreads();
    // rank sort (node ids are distinct): ascending (sheet, row, col) order
    const int4 *raw4 = reinterpret_cast<const int4 *>(raw);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int v = raw[e];
      int r = 0;
      for (int k = 0; k < tot4 / 4; ++k) {
        const int4 q = raw4[k];
        r += (q.x < v) + (q.y < v) + (q.z < v) + (q.w < v);
      }
      sst[r] = rst[e];  // starts and id
      if (staged) {
        for (int a = 0; a < 12; ++a) sw[r * 12 + a] = ib.wgt[a * (long)M + v];
        for (int a = 0; a < 3; ++a) se[r * 3 + a] = ib.frc[(NF_EX + a) * (long)M + v];
      }
    }
    __syncthreads();
    const int y0 = max(kTile * ty, bybox0), y1 = min(kTile * ty + kTile - 1, bybox1);
    const int z0 = max(kTile * tz, bzbox0), z1 = min(kTile * tz + kTile - 1, bzbox1);
    const int nxs = x1 - x0 + 1, nys = y1 - y0 + 1, nzs = z1 - z0 + 1;
    const int ncell = nxs * nys * nzs;
    for (int e = threadIdx.x; e < ncell; e += blockDim.x) {
      const int x = x0 + e % nxs, y = y0 + (e / nxs) % nys, z = z0 + e / (nxs * nys);
      double ax = 0.0, ay = 0.0, az = 0.0;
      if (staged) {
        // branch-free over the candidates, four at a time, so that their LDS
        // reads are all in flight together (the empty asm keeps the force
        // reads unconditional): a candidate whose support misses the cell
        // adds fma(0, w, a) = a (w is a finite factor read at an in-range
        // index; a sum that starts at +0 is never -0), so each cell still
        // adds exactly its own contributions in node order
        for (int k = 0; k < tot4; k += 4) {
          int4 st[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) st[u] = sst[k + u];
          bool hit[4];
          double wa[4], wb[4], wc[4], ex[4], ey[4], ez[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int ix = x - st[u].x, iy = y - st[u].y, iz = z - st[u].z;
            hit[u] = (unsigned)ix <= 3u && (unsigned)iy <= 3u && (unsigned)iz <= 3u;
            const double *w = sw + (k + u) * 12;
            wa[u] = w[ix & 3];
            wb[u] = w[4 + (iy & 3)];
            wc[u] = w[8 + (iz & 3)];
            ex[u] = se[(k + u) * 3];
            ey[u] = se[(k + u) * 3 + 1];
            ez[u] = se[(k + u) * 3 + 2];
          }
          asm volatile("" : "+v"(ex[0]), "+v"(ey[0]), "+v"(ez[0]), "+v"(ex[1]), "+v"(ey[1]), "+v"(ez[1]),
                       "+v"(ex[2]), "+v"(ey[2]), "+v"(ez[2]), "+v"(ex[3]), "+v"(ey[3]), "+v"(ez[3]));
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double wgt = cf * wa[u] * wb[u] * wc[u];
            ax = fma(hit[u] ? ex[u] : 0.0, wgt, ax);
            ay = fma(hit[u] ? ey[u] : 0.0, wgt, ay);
            az = fma(hit[u] ? ez[u] : 0.0, wgt, az);
          }
        }
      } else {
        for (int k = 0; k < tot; ++k) {
          const int4 st = sst[k];
          const int ix = x - st.x, iy = y - st.y, iz = z - st.z;
          if ((unsigned)ix > 3u || (unsigned)iy > 3u || (unsigned)iz > 3u) continue;
          const int n = st.w;
          const double wgt =
              cf * ib.wgt[ix * (long)M + n] * ib.wgt[(4 + iy) * (long)M + n] * ib.wgt[(8 + iz) * (long)M + n];
          ax = fma(ib.frc[NF_EX * (long)M + n], wgt, ax);
          ay = fma(ib.frc[NF_EY * (long)M + n], wgt, ay);
          az = fma(ib.frc[NF_EZ * (long)M + n], wgt, az);
        }
      }
      const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
      F[c0] = (FT)ax;
      F[g.n + c0] = (FT)ay;
      F[2 * g.n + c0] = (FT)az;
    }
    __syncthreads();
  }
}

template <class FT>
__global__ void __launch_bounds__(kGatherThreads)
    k_gather(IBState ib, Geometry g, int slot, FT *__restrict__ F, int mcap) {
  gather_body<FT>(ib, g, slot, F, mcap, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------------------- k_move
//
// Two kernels.  k_band_u evaluates compute_rho_and_u (lbm_ib.c:2498-2525) on
// every cell of each sheet's support box from the new populations and the
// current force -- once per cell, where a node-centred evaluation would
// redo each cell for the ~16 nodes whose support covers it -- into the u
// field ub (cells outside the moment box keep init_gv's velocity).  k_move
// then runs one wave per fiber node, one lane per cell of its 4x4x4 support
// (lanes ordered along x for coalescing), weights u, and lane 0..2 sum the 64
// products in the reference's (i2, j2, k2) order (lbm_ib.c:2614-2682) before
// X += dt u(X).

// compute_rho_and_u (lbm_ib.c:2498-2525) at cell (x, y, z) from the new
// populations and the current force; cells outside the moment box keep
// init_gv's velocity (u0x, u0y, u0z).
template <int P>
__device__ __forceinline__ void cell_u(const Geometry &g, const typename Traits<P>::S *__restrict__ f,
                                       const typename Traits<P>::FT *__restrict__ F, double u0x, double u0y,
                                       double u0z, double gz, const double *__restrict__ ref, typename Traits<P>::T b1,
                                       typename Traits<P>::T b2, int x, int y, int z, double u[3]) {
  const bool mom =
      x >= g.mlo[0] && x <= g.mhi[0] && y >= g.mlo[1] && y <= g.mhi[1] && z >= g.mlo[2] && z <= g.mhi[2];
  u[0] = u0x;
  u[1] = u0y;
  u[2] = u0z;
  if (mom) {
    const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
    typename Traits<P>::T v[19];
    pull<P>(g, f, x, y, z, b1, b2, v);
    double fv[19];
    for (int q = 0; q < 19; ++q) fv[q] = Traits<P>::dev ? (double)v[q] + ref[q] : (double)v[q];
    double s1 = fv[0], s2 = 0, s3 = 0, s4 = 0;
    for (int q = 1; q < 19; ++q) {
      s1 += fv[q];
      if (kCX[q]) s2 = kCX[q] > 0 ? s2 + fv[q] : s2 - fv[q];
      if (kCY[q]) s3 = kCY[q] > 0 ? s3 + fv[q] : s3 - fv[q];
      if (kCZ[q]) s4 = kCZ[q] > 0 ? s4 + fv[q] : s4 - fv[q];
    }
    const double fx = F[c0], fy = F[g.n + c0], fz = F[2 * g.n + c0];
    u[0] = (s2 + 0.5 * fx) / s1;
    u[1] = (s3 + 0.5 * fy) / s1;
    u[2] = (s4 + 0.5 * (fz + s1 * gz)) / s1;
  }
}

template <int P>
__device__ __forceinline__ void band_u_body(const IBState &ib, const Geometry &g, const typename Traits<P>::S *__restrict__ f,
                                            const typename Traits<P>::FT *__restrict__ F, double u0x, double u0y,
                                            double u0z, double gz, const double *__restrict__ ref,
                                            typename Traits<P>::T b1, typename Traits<P>::T b2, int slot, int bid,
                                            int nblk) {
  const long tid = (long)bid * blockDim.x + threadIdx.x, stride = (long)nblk * blockDim.x;
  // the cells of all sheets' boxes in one index range, so that one pass of
  // the grid covers every sheet (a cell in two boxes gets the same value
  // twice)
  long tot = 0;
  for (int s = 0; s < ib.nsheets; ++s) {
    const int *b = ib.box + (slot * kMaxSheets + s) * 6;
    if (b[0] <= b[1]) tot += (long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1);
  }
  for (long e0 = tid; e0 < tot; e0 += stride) {
    long e = e0;
    int s = 0;
    for (;; ++s) {
      const int *b = ib.box + (slot * kMaxSheets + s) * 6;
      const long cnt = b[0] <= b[1] ? (long)(b[1] - b[0] + 1) * (b[3] - b[2] + 1) * (b[5] - b[4] + 1) : 0;
      if (e < cnt || s == ib.nsheets - 1) break;
      e -= cnt;
    }
    {
      const int *b = ib.box + (slot * kMaxSheets + s) * 6;
      const int x0 = b[0], y0 = b[2], z0 = b[4];
      const int nxs = b[1] - x0 + 1, nys = b[3] - y0 + 1;
      const int x = x0 + (int)(e % nxs), y = y0 + (int)((e / nxs) % nys), z = z0 + (int)(e / ((long)nxs * nys));
      const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
      double u[3];
      cell_u<P>(g, f, F, u0x, u0y, u0z, gz, ref, b1, b2, x, y, z, u);
      ib.ub[c0] = u[0];
      ib.ub[g.n + c0] = u[1];
      ib.ub[2 * g.n + c0] = u[2];
    }
  }
}

template <int P>
__global__ void __launch_bounds__(256)
    k_band_u(IBState ib, Geometry g, const typename Traits<P>::S *__restrict__ f,
             const typename Traits<P>::FT *__restrict__ F, double u0x, double u0y, double u0z, double gz,
             const double *__restrict__ ref, typename Traits<P>::T b1, typename Traits<P>::T b2, int slot) {
  band_u_body<P>(ib, g, f, F, u0x, u0y, u0z, gz, ref, b1, b2, slot, blockIdx.x, gridDim.x);
}

constexpr int kMoveThreads = 256;

__device__ __forceinline__ void move_body(const IBState &ib, const Geometry &g, int reset_slot, int bid, int nblk) {
  __shared__ double prod[kMoveThreads / 64][3][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = bid * (kMoveThreads / 64) + wv;
  const int M = ib.nnodes;
  // the node's support start, delta factors and (lanes 0..2) position are
  // read together, at a clamped index for the waves past the last node;
  // only the u reads depend on them
  const int nn = min(n, M - 1);
  const int i2 = lane & 3, j2 = (lane >> 2) & 3, k2 = lane >> 4;
  const int sx = ib.start[nn], sy = ib.start[M + nn], sz = ib.start[2 * M + nn];
  const double wx = ib.wgt[i2 * (long)M + nn], wy = ib.wgt[(4 + j2) * (long)M + nn],
               wz = ib.wgt[(8 + k2) * (long)M + nn];
  const double X = ib.pos[(long)min(lane, 2) * M + nn];
  const bool live = n < M && sx >= 0;  // start -1: support off the grid (error latched)
  if (live) {
    // lanes run along x (consecutive addresses); the product goes to its
    // slot in the reference's (i2, j2, k2) loop order
    const int slot_ref = i2 * 16 + j2 * 4 + k2;
    const int x = sx + i2, y = sy + j2, z = sz + k2;
    const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
    const double ux = ib.ub[c0], uy = ib.ub[g.n + c0], uz = ib.ub[2 * g.n + c0];
    const double c_64 = 1.0 / (6.4e1);
    prod[wv][0][slot_ref] = ux * wx * wy * wz * c_64;
    prod[wv][1][slot_ref] = uy * wx * wy * wz * c_64;
    prod[wv][2][slot_ref] = uz * wx * wy * wz * c_64;
  }
  __syncthreads();
  if (live && lane < 3) {
    double acc = 0;
    for (int l = 0; l < 64; ++l) acc += prod[wv][lane][l];
    ib.pos[(long)lane * M + n] = X + 1 * acc;  // X += dt * u(X), lbm_ib.c:2697-2699
  }
  // clear the bins for the next spread and the box slot it will fill
  const int tid = bid * blockDim.x + threadIdx.x;
  const int nbins = ib.nby * ib.nbz;
  for (int b = tid; b < nbins; b += nblk * blockDim.x) ib.bin_count[b] = 0;
  if (tid < ib.nsheets * 6) {
    const int s = tid / 6, k = tid % 6;
    ib.box[(reset_slot * kMaxSheets + s) * 6 + k] = (k & 1) ? INT_MIN : INT_MAX;
  }
}

__global__ void __launch_bounds__(kMoveThreads) k_move(IBState ib, Geometry g, int reset_slot) {
  move_body(ib, g, reset_slot, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------------- k_move_tiles
//
// The band velocity and the move of the single-instance steps in one launch.
// One workgroup per tile of 4x4 nodes of a sheet (rows x cols).  The tile's
// 48 (node, component) lanes, all in wave 0, load their node's support
// start, delta factors and position; the wave reduces the tile's support
// bounding box; the workgroup evaluates cell_u once per cell of that box
// into LDS (about 5x7x7 cells for a sheet across the flow), in x slabs when
// the box is wider than LDS holds; each lane then sums its node's 64
// products u wx wy wz / 64 in the reference's (i2, j2, k2) order
// (lbm_ib.c:2614-2682) -- i2, the x offset, is the outer index, so slabs of
// increasing x keep that order -- and sets X += dt u(X) (lbm_ib.c:2697-2699).
// A box whose (y, z) face alone exceeds LDS (a sheet stretched far out of
// shape) is evaluated per lane from global memory instead.  Same expressions
// and summation order as k_band_u + k_move, so the same bits.  Every
// workgroup also clears its share of the bins for the next spread, and
// workgroup 0 the box slot the next step fills.

constexpr int kMvThreads = 256;
constexpr int kMvCells = 1024;  // cells of u per slab: 24 KB
static_assert(3 * kMvNodes * kMvNodes <= 64, "the tile's lanes fit in one wave");

template <int P>
__global__ void __launch_bounds__(kMvThreads)
    k_move_tiles(IBState ib, Geometry g, const typename Traits<P>::S *__restrict__ f,
                 const typename Traits<P>::FT *__restrict__ F, double u0x, double u0y, double u0z, double gz,
                 const double *__restrict__ ref, typename Traits<P>::T b1, typename Traits<P>::T b2, int reset_slot) {
  __shared__ double su[3][kMvCells];
  __shared__ int s_box[6];
  const int tid = threadIdx.x;
  const int M = ib.nnodes;
  {
    // clear the bins for the next spread and the box slot it will fill
    const int nbins = ib.nby * ib.nbz;
    for (int k = blockIdx.x * kMvThreads + tid; k < nbins; k += gridDim.x * kMvThreads) ib.bin_count[k] = 0;
    if (blockIdx.x == 0 && tid < ib.nsheets * 6)
      ib.box[(reset_slot * kMaxSheets + tid / 6) * 6 + tid % 6] = (tid & 1) ? INT_MIN : INT_MAX;
  }
  // this workgroup's tile: workgroups are dealt to the 8 XCDs in turn, so
  // workgroup b takes tile (b % 8) * per + b / 8, a contiguous band of tiles
  // per XCD, whose overlapping support regions then meet in that XCD's L2
  const int per = (gridDim.x + 7) / 8;
  int s = 0, w = (blockIdx.x % 8) * per + blockIdx.x / 8, tcols = 1;
  for (;; ++s) {
    if (s == ib.nsheets) return;
    const SheetDesc &sd = ib.sheets[s];
    tcols = (sd.cols + kMvNodes - 1) / kMvNodes;
    const int nt = ((sd.rows + kMvNodes - 1) / kMvNodes) * tcols;
    if (w < nt) break;
    w -= nt;
  }
  const SheetDesc &sd = ib.sheets[s];
  const int pr = tid / 3, comp = tid % 3;
  const int i = (w / tcols) * kMvNodes + pr / kMvNodes, j = (w % tcols) * kMvNodes + pr % kMvNodes;
  const int n = sd.off + i * sd.cols + j;
  bool act = tid < 3 * kMvNodes * kMvNodes && i < sd.rows && j < sd.cols;
  int st[3] = {0, 0, 0};
  double wx[4], wy[4], wz[4], X = 0.0;
  if (act) {
    st[0] = ib.start[n];
    st[1] = ib.start[M + n];
    st[2] = ib.start[2 * M + n];
    for (int a = 0; a < 4; ++a) {
      wx[a] = ib.wgt[a * (long)M + n];
      wy[a] = ib.wgt[(4 + a) * (long)M + n];
      wz[a] = ib.wgt[(8 + a) * (long)M + n];
    }
    X = ib.pos[(long)comp * M + n];
    act = st[0] >= 0;  // start -1: support off the grid (error latched), not moved
  }
  if (tid < 64) {
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = wave_min(act ? st[d] : INT_MAX);
      hi[d] = wave_max(act ? st[d] + 3 : INT_MIN);
    }
    if (tid == 0)
      for (int d = 0; d < 3; ++d) {
        s_box[2 * d] = lo[d];
        s_box[2 * d + 1] = hi[d];
      }
  }
  __syncthreads();
  const int xlo = s_box[0], xhi = s_box[1], ylo = s_box[2], zlo = s_box[4];
  if (xlo > xhi) return;  // no node of the tile is moved
  const int nyr = s_box[3] - ylo + 1, nzr = s_box[5] - zlo + 1;
  const double c_64 = 1.0 / (6.4e1);
  double acc = 0;
  if (nyr * nzr <= kMvCells) {
    const int W = min(xhi - xlo + 1, kMvCells / (nyr * nzr));
    for (int xa = xlo; xa <= xhi; xa += W) {
      const int nxw = min(W, xhi - xa + 1), ncell = nxw * nyr * nzr;
      if (xa != xlo) __syncthreads();  // the previous slab's readers are done
      for (int c = tid; c < ncell; c += kMvThreads) {
        const int x = xa + c % nxw, y = ylo + (c / nxw) % nyr, z = zlo + c / (nxw * nyr);
        double u[3];
        cell_u<P>(g, f, F, u0x, u0y, u0z, gz, ref, b1, b2, x, y, z, u);
        su[0][c] = u[0];
        su[1][c] = u[1];
        su[2][c] = u[2];
      }
      __syncthreads();
      if (act) {
        for (int i2 = 0; i2 < 4; ++i2) {
          const int xr = st[0] + i2 - xa;
          if (xr < 0 || xr >= nxw) continue;
          for (int j2 = 0; j2 < 4; ++j2)
            for (int k2 = 0; k2 < 4; ++k2) {
              const double u = su[comp][xr + nxw * ((st[1] + j2 - ylo) + nyr * (st[2] + k2 - zlo))];
              acc += u * wx[i2] * wy[j2] * wz[k2] * c_64;
            }
        }
      }
    }
  } else if (act) {
    for (int i2 = 0; i2 < 4; ++i2)
      for (int j2 = 0; j2 < 4; ++j2)
        for (int k2 = 0; k2 < 4; ++k2) {
          double u[3];
          cell_u<P>(g, f, F, u0x, u0y, u0z, gz, ref, b1, b2, st[0] + i2, st[1] + j2, st[2] + k2, u);
          acc += u[comp] * wx[i2] * wy[j2] * wz[k2] * c_64;
        }
  }
  if (act) ib.pos[(long)comp * M + n] = X + 1 * acc;  // X += dt * u(X)
}

__global__ void k_ib_reset(IBState ib, int reset_bins, int slot_mask) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (reset_bins) {
    const int nbins = ib.nby * ib.nbz;
    for (int b = tid; b < nbins; b += gridDim.x * blockDim.x) ib.bin_count[b] = 0;
  }
  for (int sl = 0; sl < 2; ++sl) {
    if (!(slot_mask & (1 << sl))) continue;
    if (tid < kMaxSheets * 6) {
      const int k = tid % 6;
      ib.box[sl * kMaxSheets * 6 + tid] = (k & 1) ? INT_MIN : INT_MAX;
    }
  }
}

// ------------------------------------------------------ batched ensemble
//
// The same kernels over every instance of an ensemble in one launch: block b
// serves instance map[b] >> kBatchShift as its block map[b] & mask of that
// instance's own grid (fiber_blocks, gather_blocks, band_u_blocks,
// move_blocks of its node count).

__device__ __forceinline__ void batch_block(const int *map, int &inst, int &lb) {
  const int m = map[blockIdx.x];
  inst = m >> kBatchShift;
  lb = m & ((1 << kBatchShift) - 1);
}

// The descriptor is read-only during a launch: its IB state and geometry are
// copied into registers with scalar loads (as kernel arguments would be).
__global__ void __launch_bounds__(kFiberThreads) k_fiber_batch(const BatchDesc *__restrict__ descs,
                                                               const int *__restrict__ map, int slot) {
  int inst, lb;
  batch_block(map, inst, lb);
  const BatchDesc *dp = descs + inst;
  const IBState ib = ldc_obj(&dp->ib);
  fiber_body(ib, ldc_obj(&dp->g), FK_ALL, slot, lb, fiber_blocks(ib.nnodes));
}

template <class FT>
__global__ void __launch_bounds__(kGatherThreads) k_gather_batch(const BatchDesc *__restrict__ descs,
                                                                 const int *__restrict__ map, int slot) {
  int inst, lb;
  batch_block(map, inst, lb);
  const BatchDesc *dp = descs + inst;
  const IBState ib = ldc_obj(&dp->ib);
  gather_body<FT>(ib, ldc_obj(&dp->g), slot, (FT *)ldc(&dp->F[slot]), ldc(&dp->mcap), lb, gather_blocks(ib.nnodes));
}

template <int P>
__global__ void __launch_bounds__(256) k_band_u_batch(const BatchDesc *__restrict__ descs,
                                                      const int *__restrict__ map, int cur_new, int slot) {
  using T = typename Traits<P>::T;
  int inst, lb;
  batch_block(map, inst, lb);
  const BatchDesc *dp = descs + inst;
  const IBState ib = ldc_obj(&dp->ib);
  const KConsts<T> *kc = reinterpret_cast<const KConsts<T> *>(dp->kc);
  band_u_body<P>(ib, ldc_obj(&dp->g), (const typename Traits<P>::S *)ldc(&dp->f[cur_new]),
                 (const typename Traits<P>::FT *)ldc(&dp->F[slot]), ldc(&dp->u0[0]), ldc(&dp->u0[1]),
                 ldc(&dp->u0[2]), ldc(&dp->gz), ldc(&dp->ref), ldc(&kc->br[1]), ldc(&kc->br[2]), slot, lb,
                 band_u_blocks(ib.nnodes));
}

__global__ void __launch_bounds__(kMoveThreads) k_move_batch(const BatchDesc *__restrict__ descs,
                                                             const int *__restrict__ map, int reset_slot) {
  int inst, lb;
  batch_block(map, inst, lb);
  const BatchDesc *dp = descs + inst;
  const IBState ib = ldc_obj(&dp->ib);
  move_body(ib, ldc_obj(&dp->g), reset_slot, lb, move_blocks(ib.nnodes));
}

// ------------------------------------------------------------- launchers

static inline int grid_for(long n, int b) { return (int)((n + b - 1) / b); }

int launch_fiber(const IBState &ib, const Geometry &g, int mask, int slot, hipStream_t s, bool band) {
  if (ib.nnodes <= 0) return LBMIB_OK;
  IBState k = ib;
  if (!band) k.band = nullptr;
  k_fiber<<<fiber_blocks(ib.nnodes), kFiberThreads, 0, s>>>(k, g, mask, slot);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int launch_gather(int prec, const IBState &ib, const Geometry &g, int slot, void *F, hipStream_t s) {
  if (ib.nnodes <= 0) return LBMIB_OK;
  const int mcap = gather_mcap(ib.cap);
  const size_t shm = gather_shm_bytes(mcap);
  const int blocks = gather_blocks(ib.nnodes);
  if (prec == LBMIB_F64) k_gather<double><<<blocks, kGatherThreads, shm, s>>>(ib, g, slot, (double *)F, mcap);
  else k_gather<float><<<blocks, kGatherThreads, shm, s>>>(ib, g, slot, (float *)F, mcap);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

// LBMIB_MOVE_KERNELS=2 selects the two-kernel move (k_band_u + k_move, the
// batched path's kernels) for the single-instance steps, to compare the two.
static bool move_two_kernels() {
  static const bool two = [] {
    const char *v = getenv("LBMIB_MOVE_KERNELS");
    return v && v[0] == '2';
  }();
  return two;
}

int launch_move(int prec, const IBState &ib, const Geometry &g, const void *f, const void *F,
                const FluidConsts<double> &c, int reset_slot, int ntiles, hipStream_t s) {
  if (ib.nnodes <= 0) return LBMIB_OK;
  const int slot = 1 - reset_slot;  // the box of this step's spread
  // bounce-back correction of the deviation forms, scale 1 (F32) or 2^8
  const double s16 = kDevScale16;
  const double d1 = host_bounce_corr(c, 1, 1.0), d2 = host_bounce_corr(c, 2, 1.0);
  const float h1 = (float)host_bounce_corr(c, 1, s16), h2 = (float)host_bounce_corr(c, 2, s16);
  if (!move_two_kernels()) {
    const int mb = (ntiles + 7) / 8 * 8;  // whole rounds over the 8 XCDs
    switch (prec) {
      case LBMIB_F64:
        k_move_tiles<LBMIB_F64><<<mb, kMvThreads, 0, s>>>(ib, g, (const double *)f, (const double *)F, c.u0x, c.u0y,
                                                          c.u0z, c.g, c.dref, 0.0, 0.0, reset_slot);
        break;
      case LBMIB_F32:
        k_move_tiles<LBMIB_F32><<<mb, kMvThreads, 0, s>>>(ib, g, (const float *)f, (const float *)F, c.u0x, c.u0y,
                                                          c.u0z, c.g, c.dref, d1, d2, reset_slot);
        break;
      case LBMIB_F16:
        k_move_tiles<LBMIB_F16><<<mb, kMvThreads, 0, s>>>(ib, g, (const half_t *)f, (const float *)F, c.u0x, c.u0y,
                                                          c.u0z, c.g, c.dref, h1, h2, reset_slot);
        break;
      case LBMIB_BF16:
        k_move_tiles<LBMIB_BF16><<<mb, kMvThreads, 0, s>>>(ib, g, (const uint16_t *)f, (const float *)F, c.u0x,
                                                           c.u0y, c.u0z, c.g, c.dref, h1, h2, reset_slot);
        break;
      default:
        return LBMIB_ERR_ARG;
    }
    return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
  }
  const int ublocks = band_u_blocks(ib.nnodes);
  switch (prec) {
    case LBMIB_F64:
      k_band_u<LBMIB_F64><<<ublocks, 256, 0, s>>>(ib, g, (const double *)f, (const double *)F, c.u0x, c.u0y, c.u0z,
                                                  c.g, c.dref, 0.0, 0.0, slot);
      break;
    case LBMIB_F32:
      k_band_u<LBMIB_F32><<<ublocks, 256, 0, s>>>(ib, g, (const float *)f, (const float *)F, c.u0x, c.u0y, c.u0z,
                                                  c.g, c.dref, d1, d2, slot);
      break;
    case LBMIB_F16:
      k_band_u<LBMIB_F16><<<ublocks, 256, 0, s>>>(ib, g, (const half_t *)f, (const float *)F, c.u0x, c.u0y, c.u0z,
                                                  c.g, c.dref, h1, h2, slot);
      break;
    case LBMIB_BF16:
      k_band_u<LBMIB_BF16><<<ublocks, 256, 0, s>>>(ib, g, (const uint16_t *)f, (const float *)F, c.u0x, c.u0y,
                                                   c.u0z, c.g, c.dref, h1, h2, slot);
      break;
    default:
      return LBMIB_ERR_ARG;
  }
  k_move<<<move_blocks(ib.nnodes), kMoveThreads, 0, s>>>(ib, g, reset_slot);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int launch_ib_reset(const IBState &ib, int reset_bins, int reset_slot_mask, hipStream_t s) {
  const int nbins = ib.nby * ib.nbz;
  const long work = nbins > kMaxSheets * 6 ? nbins : kMaxSheets * 6;
  k_ib_reset<<<grid_for(work, 256), 256, 0, s>>>(ib, reset_bins, reset_slot_mask);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int gather_mcap(int cap) { return min(kCandBins * kCandBins * cap, kGatherMaxCand); }

int launch_fiber_batch(const BatchDesc *d, const int *map, int nblocks, int slot, hipStream_t s) {
  if (nblocks <= 0) return LBMIB_OK;
  k_fiber_batch<<<nblocks, kFiberThreads, 0, s>>>(d, map, slot);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}
int launch_gather_batch(int prec, const BatchDesc *d, const int *map, int nblocks, int slot, int mcap_max,
                        hipStream_t s) {
  if (nblocks <= 0) return LBMIB_OK;
  const size_t shm = gather_shm_bytes(mcap_max);
  if (prec == LBMIB_F64) k_gather_batch<double><<<nblocks, kGatherThreads, shm, s>>>(d, map, slot);
  else k_gather_batch<float><<<nblocks, kGatherThreads, shm, s>>>(d, map, slot);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}
int launch_band_u_batch(int prec, const BatchDesc *d, const int *map, int nblocks, int cur_new, int slot,
                        hipStream_t s) {
  if (nblocks <= 0) return LBMIB_OK;
  switch (prec) {
    case LBMIB_F64: k_band_u_batch<LBMIB_F64><<<nblocks, 256, 0, s>>>(d, map, cur_new, slot); break;
    case LBMIB_F32: k_band_u_batch<LBMIB_F32><<<nblocks, 256, 0, s>>>(d, map, cur_new, slot); break;
    case LBMIB_F16: k_band_u_batch<LBMIB_F16><<<nblocks, 256, 0, s>>>(d, map, cur_new, slot); break;
    case LBMIB_BF16: k_band_u_batch<LBMIB_BF16><<<nblocks, 256, 0, s>>>(d, map, cur_new, slot); break;
    default: return LBMIB_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}
int launch_move_batch(const BatchDesc *d, const int *map, int nblocks, int reset_slot, hipStream_t s) {
  if (nblocks <= 0) return LBMIB_OK;
  k_move_batch<<<nblocks, kMoveThreads, 0, s>>>(d, map, reset_slot);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

}  // namespace lbmib
