// lbmib_fluid.hip -- fused LBM kernels for gfx950.
//
// Layout in HBM: population q of cell (x, y, z) at f[q * pq + x + prow *
// (y + ny * z)], x fastest, in one of two layouts (Geometry::pq, prow):
//   planes  one plane of the whole grid per population (pq >= nx ny nz,
//           prow = nx): f[q][z][y][x]
//   rows    the 19 populations of each x row side by side (pq = nx,
//           prow = 19 nx): f[z][y][q][x], so that the loads and stores of a
//           row's update fall in a few contiguous blocks of memory
// with two such buffers (ping-pong).  The stored state between steps is the
// post-collision population f* of the last step.  One launch of
// k_collide_stream per step pulls f*(x - c_q) from the previous step's buffer
// (this is stream_distrfunc, bounceback_rigidwalls and copy_inout_to_df2 of
// the previous step, lbm_ib.c:1827-2437, 2769-2907), recomputes the lagged
// rho/u from the pulled populations plus dt/2 of the previous force
// (compute_rho_and_u, lbm_ib.c:2440-2538), and collides with the current
// force (compute_eqlbrmdistrfuncDF1, lbm_ib.c:1673-1825).  Every population
// is read once and written once per step: 2*19*sizeof(storage) bytes per
// cell, the HBM roofline of the path.
//
// Storage forms:
//   F64       f itself; arithmetic in the reference's operation order with no
//             FMA contraction (the library is built with -ffp-contract=off),
//             so results track the FP64 oracle to a few ulp.
//   F32       h = f - r_q in FP32, r_q = feq_q(rho0, (u0x, 0, 0)) by default
//             (or w_q, LBMIB_DEV_REST); collision in FP64 on the stored
//             deviations, which keeps the small velocity (|u| ~ 1e-2) out of
//             the rounding of the O(0.3) rest population.
//   F16/BF16  h = 2^8 (f - r_q) in 16 bits; collision in FP32.
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "lbmib_coarse.h"
#include "lbmib_device.h"

namespace lbmib {

// Host FP64 constants -> kernel constants of precision P (kernel units).
template <int P>
static KConsts<typename Traits<P>::T> kconsts(const FluidConsts<double> &d) {
  using T = typename Traits<P>::T;
  KConsts<T> c;
  const double s = Traits<P>::scale;
  c.omc = (T)d.omc;
  c.om = (T)d.om;
  for (int k = 0; k < 3; ++k) {
    const double w = 1.0 / kWDEN_CLS[k];
    const int q = k == 0 ? 0 : (k == 1 ? 1 : 7);  // a direction of class k
    c.wa[k] = Traits<P>::dev ? (T)(s * d.om * w) : (T)0;
    c.wb[k] = Traits<P>::dev ? (T)(s * d.pw[q]) : (T)d.pref[q];
  }
  c.cs2 = (T)d.cs2;
  c.cs4 = (T)d.cs4;
  c.g = (T)d.g;
  c.rho0 = (T)d.rho0;
  c.u0x = (T)d.u0x;
  c.u0y = (T)d.u0y;
  c.u0z = (T)d.u0z;
  c.rhor = (T)d.rhor;
  c.d0 = (T)(d.rho0 - d.rhor);
  c.mrx = (T)(d.rhor * d.urx);
  const double u2 = d.urx * d.urx;
  c.kr[0] = (T)(d.rhor * (-1.5 * u2));
  c.kr[1] = (T)(d.rhor * (4.5 * u2 - 1.5 * u2));
  for (int k = 0; k < 3; ++k) {
    const double w = 1.0 / kWDEN_CLS[k];
    c.orr[k] = Traits<P>::dev && k ? (T)(3.0 * (s * d.om * w) * d.rhor * d.urx) : (T)0;
    c.br[k] = Traits<P>::dev ? (T)host_bounce_corr(d, k, s) : (T)0;
  }
  c.tab = (const T *)d.dtab;
  c.ref = d.dref;
  return c;
}

// Equilibrium + BGK + Guo forcing in the reference's operation order
// (lbm_ib.c:1731-1818).  Used by the F64 form.  pref_q depends on the
// weight class only.
__device__ __forceinline__ double collide_exact(int q, double fq, double rho, double vx, double vy, double vz,
                                                double fx, double fy, double fz, const KConsts<double> &c) {
  const double cx = kCX[q], cy = kCY[q], cz = kCZ[q];
  const double usq = vx * vx + vy * vy + vz * vz;
  const double cu = cx * vx + cy * vy + cz * vz;
  double eq;
  if (q == 0) eq = 1.0 / 3.0 * rho * (1.0 - 1.5 * usq);
  else if (q <= 6) eq = 1.0 / 18.0 * rho * (1.0 + 3.0 * cu + 4.5 * (cu * cu) - 1.5 * usq);
  else eq = 1.0 / 36.0 * rho * (1.0 + 3.0 * cu + 4.5 * (cu * cu) - 1.5 * usq);
  const double fzg = fz + rho * c.g;
  const double a = (cx - vx) * fx + (cy - vy) * fy + (cz - vz) * fzg;
  const double cf = cx * fx + cy * fy + cz * fzg;
  return fq * c.omc + c.om * eq + c.wb[kCLS[q]] * (a / c.cs2 + cu / c.cs4 * cf);
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// c_q . (a, b, c) with only the non-zero lattice components (q is a
// compile-time constant after unrolling).
template <class T>
__device__ __forceinline__ T cdot(int q, T a, T b, T c) {
  T r = 0;
  bool first = true;
  if (kCX[q]) { r = kCX[q] > 0 ? a : -a; first = false; }
  if (kCY[q]) { const T t = kCY[q] > 0 ? b : -b; r = first ? t : r + t; first = false; }
  if (kCZ[q]) { const T t = kCZ[q] > 0 ? c : -c; r = first ? t : r + t; }
  return r;
}

// Lagged rho/u of compute_rho_and_u (lbm_ib.c:2498-2525) from the pulled
// populations (kernel units) and the previous step's force.  d = rho - 1.
template <int P>
__device__ __forceinline__ void lagged_moments(const typename Traits<P>::T v[19], typename Traits<P>::T fpx,
                                               typename Traits<P>::T fpy, typename Traits<P>::T fpz,
                                               const KConsts<typename Traits<P>::T> &c, typename Traits<P>::T &d,
                                               typename Traits<P>::T &rho, typename Traits<P>::T &ux,
                                               typename Traits<P>::T &uy, typename Traits<P>::T &uz) {
  if constexpr (!Traits<P>::dev) {
    // rho = Sum df2, u = (Sum c df2 + dt/2 F)/rho in the reference's order
    double s1 = v[0], s2 = 0.0, s3 = 0.0, s4 = 0.0;
#pragma unroll
    for (int q = 1; q < 19; ++q) {
      s1 += v[q];
      if (kCX[q]) s2 = kCX[q] > 0 ? s2 + v[q] : s2 - v[q];
      if (kCY[q]) s3 = kCY[q] > 0 ? s3 + v[q] : s3 - v[q];
      if (kCZ[q]) s4 = kCZ[q] > 0 ? s4 + v[q] : s4 - v[q];
    }
    rho = s1;
    d = s1 - 1.0;
    ux = (s2 + 0.5 * fpx) / s1;
    uy = (s3 + 0.5 * fpy) / s1;
    uz = (s4 + 0.5 * (fpz + rho * c.g)) / s1;
  } else {
    // sums of h = s (f - w); rho = 1 + Sum(h)/s
    using T = typename Traits<P>::T;
    T D = v[0], MX = v[1], MY = v[6], MZ = v[3];
#pragma unroll
    for (int q = 1; q < 19; ++q) {
      D += v[q];
      if (kCX[q] && q != 1) MX = kCX[q] > 0 ? MX + v[q] : MX - v[q];
      if (kCY[q] && q != 6) MY = kCY[q] > 0 ? MY + v[q] : MY - v[q];
      if (kCZ[q] && q != 3) MZ = kCZ[q] > 0 ? MZ + v[q] : MZ - v[q];
    }
    // h = s (f - r): rho = rhor + Sum(h)/s, rho u = rhor ur + Sum(c h)/s + F/2
    constexpr T is = (T)Traits<P>::inv_scale;
    d = D * is;  // rho - rhor
    rho = c.rhor + d;
    const T ir = (T)1 / rho;  // correctly rounded
    ux = fma_t(MX, is, fma_t((T)0.5, fpx, c.mrx)) * ir;
    uy = fma_t(MY, is, (T)0.5 * fpy) * ir;
    uz = fma_t(MZ, is, (T)0.5 * fma_t(rho, c.g, fpz)) * ir;
  }
}

// BGK + Guo collision of one cell, in place, in kernel units.
template <int P>
__device__ __forceinline__ void collide_cell(typename Traits<P>::T v[19], typename Traits<P>::T d,
                                             typename Traits<P>::T rho, typename Traits<P>::T ux,
                                             typename Traits<P>::T uy, typename Traits<P>::T uz,
                                             typename Traits<P>::T fx, typename Traits<P>::T fy,
                                             typename Traits<P>::T fz, const KConsts<typename Traits<P>::T> &c) {
  if constexpr (!Traits<P>::dev) {
#pragma unroll
    for (int q = 0; q < 19; ++q) v[q] = collide_exact(q, v[q], rho, ux, uy, uz, fx, fy, fz, c);
  } else {
    // Deviation form on h_q = s (f_q - r_q), r = feq(rhor, ur), d = rho - rhor:
    //   h*_q = (1 - 1/tau) h_q + s/tau (feq_q - r_q) + s S_q
    //   feq_q - r_q = w_q (rho (1 + 3 cu + 4.5 cu^2 - 1.5 u^2)
    //                      - rhor (1 + 3 cur + 4.5 cur^2 - 1.5 ur^2))
    //   S_q = w_q (1 - 1/(2 tau)) (3 (c_q - u).F + 9 (c_q.u)(c_q.F))
    // Opposite directions share the even part E and flip the odd part O:
    //   h*_q = omc h_q + E + O,  h*_opp = omc h_opp + E - O
    //   E = wa (d - k_p + rho (4.5 cu^2 - 1.5 u^2)) + wb (9 cu cF - 3 u.F)
    //   O = 3 wa rho cu - o_p + 3 wb cF
    // with k_p = rhor (4.5 cur^2 - 1.5 ur^2), o_p = 3 wa rhor cur, which for
    // ur = (urx, 0, 0) are kr[|c_x|] and c_x orr[class].
    using T = typename Traits<P>::T;
    const T fzg = fma_t(rho, c.g, fz);
    const T usq15 = (T)1.5 * fma_t(ux, ux, fma_t(uy, uy, uz * uz));
    const T uF3 = (T)3 * fma_t(ux, fx, fma_t(uy, fy, uz * fzg));
    const T r45 = (T)4.5 * rho;
    v[0] = fma_t(c.omc, v[0], fma_t(c.wa[0], fma_t(-rho, usq15, d - c.kr[0]), -c.wb[0] * uF3));
    const T ar3[3] = {(T)0, (T)3 * c.wa[1] * rho, (T)3 * c.wa[2] * rho};
    const T wb3[3] = {(T)0, (T)3 * c.wb[1], (T)3 * c.wb[2]};
#pragma unroll
    for (int q = 1; q < 19; q += 2) {
      const int k = kCLS[q];
      const T cu = cdot<T>(q, ux, uy, uz);
      const T cF = cdot<T>(q, fx, fy, fzg);
      const T base = fma_t(r45, cu * cu, fma_t(-rho, usq15, d - c.kr[kCX[q] != 0]));
      const T s9 = fma_t((T)9 * cu, cF, -uF3);
      const T E = fma_t(c.wa[k], base, c.wb[k] * s9);
      const T o_p = kCX[q] > 0 ? c.orr[k] : (kCX[q] < 0 ? -c.orr[k] : (T)0);
      const T O = fma_t(ar3[k], cu, fma_t(wb3[k], cF, -o_p));
      v[q] = fma_t(c.omc, v[q], E + O);
      v[q + 1] = fma_t(c.omc, v[q + 1], E - O);
    }
  }
}

// Arriving populations (kernel units) and lagged moments of one active cell
// by the general rules: inlet/outlet table on the frozen planes, own
// populations on the first step, pull otherwise.
template <int P, bool NTL>
__device__ __forceinline__ void arrive_general(const Geometry &g, const typename Traits<P>::S *__restrict__ fin,
                                               int own, const KConsts<typename Traits<P>::T> &c, int nsheets,
                                               const int *__restrict__ box_prev,
                                               const typename Traits<P>::FT *__restrict__ Fprev, int x, int y,
                                               int z, long c0, typename Traits<P>::T v[19],
                                               typename Traits<P>::T &d, typename Traits<P>::T &rho,
                                               typename Traits<P>::T &ux, typename Traits<P>::T &uy,
                                               typename Traits<P>::T &uz) {
  using T = typename Traits<P>::T;
  const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
  if (frozen || own) {
    // planes x = 2 and N-3 hold the inlet/outlet tables after
    // copy_inout_to_df2, and their rho/vel are never updated (the moment box
    // starts at x = 3, lbm_ib.c:2482-2483); on the first step every cell
    // keeps init_gv's rho/u (lbm_ib.c:837-849).
    if (frozen) {
#pragma unroll
      for (int q = 0; q < 19; ++q) v[q] = ldc(c.tab + q);
    } else {
#pragma unroll
      for (int q = 0; q < 19; ++q) v[q] = ld_h<P>(fin + q * g.pq + pop_index(g, x, y, z));
    }
    rho = c.rho0; d = c.d0; ux = c.u0x; uy = c.u0y; uz = c.u0z;
    return;
  }
  // the force is loaded before the box test (the field exists whenever
  // there are sheets), so that its loads overlap the pulls; it counts only
  // inside a support box
  T fpx = 0, fpy = 0, fpz = 0;
  if (nsheets > 0) {
    const T lx = (T)Fprev[c0], ly = (T)Fprev[g.n + c0], lz = (T)Fprev[2 * g.n + c0];
    const bool in = in_any_box(box_prev, nsheets, x, y, z);
    fpx = in ? lx : (T)0; fpy = in ? ly : (T)0; fpz = in ? lz : (T)0;
  }
  pull<P, NTL, true>(g, fin, x, y, z, c.br[1], c.br[2], v);
  lagged_moments<P>(v, fpx, fpy, fpz, c, d, rho, ux, uy, uz);
}

// The whole update of one active cell by the general rules.
template <int P, bool NT, bool NTL = false>
__device__ __forceinline__ void update_cell(const Geometry &g, const typename Traits<P>::S *__restrict__ fin,
                                            typename Traits<P>::S *__restrict__ fout, int own,
                                            const KConsts<typename Traits<P>::T> &c, int nsheets,
                                            const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                                            const typename Traits<P>::FT *__restrict__ Fcur,
                                            const typename Traits<P>::FT *__restrict__ Fprev, int x, int y, int z) {
  using T = typename Traits<P>::T;
  if (x < g.lo[0] || x > g.hi[0] || y < g.lo[1] || y > g.hi[1] || z < g.lo[2] || z > g.hi[2]) return;
  const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
  T v[19];
  T d, rho, ux, uy, uz;
  T fx = 0, fy = 0, fz = 0;
  if (nsheets > 0) {  // loaded early, as in arrive_general
    const T lx = (T)Fcur[c0], ly = (T)Fcur[g.n + c0], lz = (T)Fcur[2 * g.n + c0];
    const bool in = in_any_box(box_cur, nsheets, x, y, z);
    fx = in ? lx : (T)0; fy = in ? ly : (T)0; fz = in ? lz : (T)0;
  }
  arrive_general<P, NTL>(g, fin, own, c, nsheets, box_prev, Fprev, x, y, z, c0, v, d, rho, ux, uy, uz);
  collide_cell<P>(v, d, rho, ux, uy, uz, fx, fy, fz, c);
  const long p0 = pop_index(g, x, y, z);
#pragma unroll
  for (int q = 0; q < 19; ++q) st_h<P, NT>(fout + q * g.pq + p0, v[q]);
}

// Byte offsets, in a population array, of the 19 sources of cell (0, y, z)
// of every row (y, z) -- the pull of stream_distrfunc with the y/z wrap and
// half-way bounce-back of the previous step -- for the buffer-addressed row
// path; the lane adds x * sizeof(S).  Entry 19 of a row is the mask of its
// bounced-back directions (bit q).  Built once per context on the host.
constexpr int kRowTabStride = 20;

int row_source_table(const Geometry &g, size_t es, std::vector<uint32_t> &tab) {
  const long nrows = (long)g.ny * g.nz;
  tab.assign((size_t)nrows * kRowTabStride, 0u);
  const uint64_t pitch_b = (uint64_t)g.pq * es, sy_b = (uint64_t)g.prow * es;
  for (long row = 0; row < nrows; ++row) {
    const int y = (int)(row % g.ny), z = (int)(row / g.ny);
    if (y < g.lo[1] || y > g.hi[1] || z < g.lo[2] || z > g.hi[2]) continue;  // dead row
    uint32_t mask = 0;
    for (int q = 0; q < 19; ++q) {
      int py = y - kCY[q], pz = z - kCZ[q];
      bool bb = false;
      if (g.bc[1] == LBMIB_BC_PERIODIC) py = (py + g.ny) % g.ny;
      else bb |= py < g.lo[1] || py > g.hi[1];
      if (g.bc[2] == LBMIB_BC_PERIODIC) pz = (pz + g.nz) % g.nz;
      else bb |= pz < g.lo[2] || pz > g.hi[2];
      const int64_t off = bb ? (int64_t)(kOPP[q] * pitch_b + row * sy_b)
                             : (int64_t)(q * pitch_b + (uint64_t)(py + (long)g.ny * pz) * sy_b) - kCX[q] * (int64_t)es;
      if (off < 0 || off > 0xFFFFFFFFll) return LBMIB_ERR_ARG;
      tab[(size_t)row * kRowTabStride + q] = (uint32_t)off;
      mask |= (uint32_t)bb << q;
    }
    tab[(size_t)row * kRowTabStride + 19] = mask;
  }
  return LBMIB_OK;
}

// ---------------------------------------------------------- split steps
//
// A split step (lbmib_set_overlap) updates the band -- per sheet, the box
// spanned by the support boxes of this and the previous step (all cells with
// a force in either), widened by one cell on each side for the pulls of the
// band velocity (band_box_of) -- and the bulk, every other cell, in two
// launches that read the same populations and write disjoint cells, so they
// may run side by side.  The step's fiber kernel stores the band boxes
// ({n, (x0, x1, y0, y1, z0, z1) x n}, one buffer per step parity);
// k_collide_band updates the band's cells one by one by the general rules,
// and the bulk row kernels skip them.  The bulk reads neither the force
// field nor the support boxes (its cells have no force), so the IB chain may
// rewrite them while it runs.  A support box moves by less than a cell per
// step, so the next step's box lies inside this step's band box.
enum { kRowsAll = 0, kRowsBand = 1, kRowsBulk = 2 };

// (the band boxes are written by the step's fiber kernel, before both launches)
__device__ __forceinline__ bool in_band_boxes(const int *rect_buf, int x, int y, int z) {
  bool in = false;
  const int n = min(ldc(rect_buf), kMaxSheets);
  for (int k = 0; k < n; ++k) {
    const int *r = rect_buf + 1 + 6 * k;
    in |= x >= ldc(r) && x <= ldc(r + 1) && y >= ldc(r + 2) && y <= ldc(r + 3) && z >= ldc(r + 4) &&
          z <= ldc(r + 5);
  }
  return in;
}

// Whether row (y, z) meets any band box (wave-uniform pre-test of the bulk).
__device__ __forceinline__ bool row_meets_band(const int *rect_buf, int y, int z) {
  bool in = false;
  const int n = min(ldc(rect_buf), kMaxSheets);
  for (int k = 0; k < n; ++k) {
    const int *r = rect_buf + 1 + 6 * k;
    in |= y >= ldc(r + 2) && y <= ldc(r + 3) && z >= ldc(r + 4) && z <= ldc(r + 5);
  }
  return in;
}

// Next row of this block's walk over all rows; false when the walk is over.
// `it` starts at blockIdx.x.
__device__ __forceinline__ bool next_row(const Geometry &g, long it, int &row) {
  if (it >= (long)g.ny * g.nz) return false;
  row = (int)it;
  return true;
}

// Deviation forms: a population reflected at a y/z wall (bit q of the row's
// mask) arrives as h* of the opposite direction, stored relative to r_opp;
// add s (r_opp - r_q) = -c_x br (non-zero for c_x != 0 only).  The mask is wave-uniform; rows away from the walls
// skip the loop.
template <int P>
__device__ __forceinline__ void bounce_fix(uint32_t mask, typename Traits<P>::T b1, typename Traits<P>::T b2,
                                           typename Traits<P>::T v[19]) {
  if constexpr (Traits<P>::dev) {
    if (mask) {
#pragma unroll
      for (int q = 0; q < 19; ++q)
        if (kCX[q] != 0 && ((mask >> q) & 1u)) v[q] -= kCX[q] * (kCLS[q] == 1 ? b1 : b2);
    }
  }
}

// A sink that takes nothing (row_cell's default).
struct NoSink {
  template <class V>
  __device__ __forceinline__ void operator()(const V *) const {}
};

// One lane of the buffer-addressed row path: the update of cell (x, row) by
// a wave whose lanes hold consecutive x of the same row (SGPR row offsets).
// Dead lanes of an active row store zeros (whole cache lines); x edges of a
// wall or periodic x axis are left to k_collide_xedges.  sink(v) sees the
// post-collision values (kernel units, before the stores round them) of an
// updated cell: not on dead rows or lanes, x edges or frozen planes.
template <int P, bool NT, bool NTL, class Sink = NoSink>
__device__ __forceinline__ void row_cell(const Geometry &g, __amdgpu_buffer_rsrc_t rin, __amdgpu_buffer_rsrc_t rout,
                                         int own, const KConsts<typename Traits<P>::T> &c, int nsheets,
                                         const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                                         const typename Traits<P>::FT *__restrict__ Fcur,
                                         const typename Traits<P>::FT *__restrict__ Fprev,
                                         const uint32_t *__restrict__ rowtab, int row, int x,
                                         const int *__restrict__ band = nullptr, int box_rows_cached = 0,
                                         Sink sink = Sink()) {
  using T = typename Traits<P>::T;
  using S = typename Traits<P>::S;
  constexpr uint32_t es = sizeof(S);
  const uint32_t pitch_b = (uint32_t)g.pq * es;   // between the populations of a cell
  const uint32_t sy_b = (uint32_t)g.prow * es;    // between rows (y, z) and (y + 1, z)
  // bulk of a split step: the band launch has stored the band's cells
  if (band && row_meets_band(band, row % g.ny, row / g.ny) && in_band_boxes(band, x, row % g.ny, row / g.ny)) return;
  const bool edge = x == g.lo[0] || x == g.hi[0];
  // frozen inlet/outlet planes; x edges of wall/periodic x go to k_collide_xedges
  const bool frozen = edge && g.bc[0] == LBMIB_BC_INOUT;
  // Filler lanes (dead x, and the x edges left to k_collide_xedges) store
  // zeros, so that every store instruction writes whole cache lines of the
  // row: partially written lines cost HBM write bandwidth.  Nothing reads
  // a dead cell; k_collide_xedges rewrites its edge cells afterwards.
  const bool filler = x < g.lo[0] || x > g.hi[0] || (edge && !frozen);
  const uint32_t vo = (uint32_t)x * es;
  const int y = row % g.ny, z = row / g.ny;
  if (y < g.lo[1] || y > g.hi[1] || z < g.lo[2] || z > g.hi[2]) return;  // dead row
  const uint32_t row_b = (uint32_t)row * sy_b;  // byte offset of (0, y, z) in a plane
  const long c0 = (long)x + (long)g.nx * row;
  T v[19];
  if (filler) {
#pragma unroll
    for (int q = 0; q < 19; ++q) v[q] = (T)0;
  } else {
    T d, rho, ux, uy, uz;
    if (frozen) {
#pragma unroll
      for (int q = 0; q < 19; ++q) v[q] = ldc(c.tab + q);
      rho = c.rho0; d = c.d0; ux = c.u0x; uy = c.u0y; uz = c.u0z;
    } else {
      // per-row source offsets (row_source_table): wrap and bounce-back
      // of the y/z axes and the x shift folded in
      if (own) {
        uint32_t qb = row_b;  // first step: the cell's own populations
#pragma unroll
        for (int q = 0; q < 19; ++q) {
          v[q] = bld_h<P, NTL>(rin, vo, qb);
          qb += pitch_b;
        }
      } else {
        const uint32_t *rt = rowtab + (long)row * kRowTabStride;
        uint32_t so[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) so[q] = ldc(rt + q);
#pragma unroll
        for (int q = 0; q < 19; ++q) v[q] = bld_h<P, NTL>(rin, vo, so[q]);
        bounce_fix<P>(ldc(rt + 19), c.br[1], c.br[2], v);
      }
      if (own) {
        rho = c.rho0; d = c.d0; ux = c.u0x; uy = c.u0y; uz = c.u0z;  // lbm_ib.c:837-849
      } else {
        T fpx = 0, fpy = 0, fpz = 0;
        if (in_any_box(box_prev, nsheets, x, y, z)) {
          fpx = (T)Fprev[c0]; fpy = (T)Fprev[g.n + c0]; fpz = (T)Fprev[2 * g.n + c0];
        }
        lagged_moments<P>(v, fpx, fpy, fpz, c, d, rho, ux, uy, uz);
      }
    }
    T fx = 0, fy = 0, fz = 0;
    if (in_any_box(box_cur, nsheets, x, y, z)) {
      fx = (T)Fcur[c0]; fy = (T)Fcur[g.n + c0]; fz = (T)Fcur[2 * g.n + c0];
    }
    collide_cell<P>(v, d, rho, ux, uy, uz, fx, fy, fz, c);
    if (!frozen) sink(v);
  }
  uint32_t sb = row_b;
  if (NT && box_rows_cached && row_near_box(box_cur, nsheets, y, z)) {
    // the rows meeting a support box: the default cache policy (as in
    // k_collide_stream_pairs)
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      bst_h<P, false>(rout, vo, sb, v[q]);
      sb += pitch_b;
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    bst_h<P, NT>(rout, vo, sb, v[q]);
    sb += pitch_b;
  }
}

// Row form: blocks walk whole (y, z) rows with a stride (one block per row
// when the grid has as many blocks as rows).
//
// BUF (buffer addressing, needs a population array under 2^32 bytes): for each
// population the source row -- wrapped on periodic axes, or the own row's
// opposite population where the source lies beyond a y/z wall (half-way
// bounce-back) -- is the same for the whole row, so its byte offset is a
// wave-uniform (SGPR) value and the 19 loads and 19 stores of a lane share one
// VGPR offset, x * sizeof(S) (one cell per thread, blockDim.x >= nx).  Lanes
// on the frozen inlet/outlet planes take the table; the two x-edge cells of a
// row on a wall or periodic x axis are left to k_collide_xedges.  Without BUF
// every cell takes update_cell.
template <int P, bool NT, bool NTL, bool BUF>
__global__ void __launch_bounds__(1024)
    k_collide_stream_rows(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                          typename Traits<P>::S *__restrict__ fout, int own, KConsts<typename Traits<P>::T> c,
                          int nsheets, const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                          const typename Traits<P>::FT *__restrict__ Fcur,
                          const typename Traits<P>::FT *__restrict__ Fprev, const uint32_t *__restrict__ rowtab,
                          const int *__restrict__ band, int box_rows_cached) {
  if constexpr (!BUF) {
    int row;
    for (long it = blockIdx.x; next_row(g, it, row); it += gridDim.x) {
      const int y = row % g.ny, z = row / g.ny;
      for (int x = threadIdx.x; x < g.nx; x += blockDim.x)
        if (!band || !in_band_boxes(band, x, y, z))
          update_cell<P, NT, NTL>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, x, y, z);
    }
  } else {
    // one cell per thread: blockDim.x >= nx (checked by the launcher)
    const __amdgpu_buffer_rsrc_t rin = make_rsrc(fin), rout = make_rsrc(fout);
    const int x = threadIdx.x;
    if (x >= g.nx) return;
    int row;
    for (long it = blockIdx.x; next_row(g, it, row); it += gridDim.x)
      row_cell<P, NT, NTL>(g, rin, rout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, row, x, band,
                           box_rows_cached);
  }
}

// Two consecutive cells of a row per lane (the buffer-addressed path with
// loads and stores twice as wide): lane t updates x = 2t and 2t + 1.  A pair
// load of plane q reads the two sources (x - c_x, x + 1 - c_x), which are
// adjacent; for c_x != 0 it is one element off the pair alignment, which the
// memory system handles as an unaligned access.  Requires nx even, the
// active x range [lo, hi] with lo even and hi odd (so no pair straddles it),
// and blockDim.x >= nx / 2 (nx <= 1024).  Edge pairs of a wall or periodic x axis compute
// their edge cell with the in-row neighbour, and k_collide_xedges, launched
// after this kernel on the same stream, rewrites that cell.
template <int P>
struct PairBits;
template <>
struct PairBits<LBMIB_F64> { typedef unsigned int type __attribute__((ext_vector_type(4))); };
template <>
struct PairBits<LBMIB_F32> { typedef unsigned int type __attribute__((ext_vector_type(2))); };
template <>
struct PairBits<LBMIB_F16> { typedef unsigned int type; };
template <>
struct PairBits<LBMIB_BF16> { typedef unsigned int type; };

template <int P, bool NTL>
__device__ __forceinline__ void bld_pair(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                         typename Traits<P>::T &a, typename Traits<P>::T &b) {
  using S = typename Traits<P>::S;
  constexpr int aux = NTL ? 2 : 0;
  struct Two { S s[2]; };
  Two t;
  if constexpr (P == LBMIB_F64) t = __builtin_bit_cast(Two, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, aux));
  else if constexpr (P == LBMIB_F32) t = __builtin_bit_cast(Two, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, aux));
  else t = __builtin_bit_cast(Two, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, aux));
  a = from_bits<P>(t.s[0]);
  b = from_bits<P>(t.s[1]);
}
template <int P, bool NT>
__device__ __forceinline__ void bst_pair(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                         typename Traits<P>::T a, typename Traits<P>::T b) {
  using S = typename Traits<P>::S;
  constexpr int aux = NT ? 2 : 0;
  struct Two { S s[2]; };
  Two t;
  t.s[0] = to_bits<P>(a);
  t.s[1] = to_bits<P>(b);
  const typename PairBits<P>::type u = __builtin_bit_cast(typename PairBits<P>::type, t);
  if constexpr (P == LBMIB_F64) __builtin_amdgcn_raw_buffer_store_b128(u, r, voff, soff, aux);
  else if constexpr (P == LBMIB_F32) __builtin_amdgcn_raw_buffer_store_b64(u, r, voff, soff, aux);
  else __builtin_amdgcn_raw_buffer_store_b32(u, r, voff, soff, aux);
}

template <int P, bool NT, bool NTL>
__global__ void __launch_bounds__(512)
    k_collide_stream_pairs(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                           typename Traits<P>::S *__restrict__ fout, int own, KConsts<typename Traits<P>::T> c,
                           int nsheets, const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                           const typename Traits<P>::FT *__restrict__ Fcur,
                           const typename Traits<P>::FT *__restrict__ Fprev, const uint32_t *__restrict__ rowtab,
                           const int *__restrict__ band, int box_rows_cached) {
  using T = typename Traits<P>::T;
  using S = typename Traits<P>::S;
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(fin), rout = make_rsrc(fout);
  constexpr uint32_t es = sizeof(S);
  const uint32_t pitch_b = (uint32_t)g.pq * es;   // between the populations of a cell
  const uint32_t sy_b = (uint32_t)g.prow * es;    // between rows (y, z) and (y + 1, z)
  const int x0 = 2 * (int)threadIdx.x;
  if (x0 >= g.nx) return;
  // dead pairs store zeros so that the row's cache lines are written whole
  const bool filler = x0 + 1 < g.lo[0] || x0 > g.hi[0];
  const bool inout_x = g.bc[0] == LBMIB_BC_INOUT;
  const bool frozen[2] = {inout_x && x0 == g.lo[0], inout_x && x0 + 1 == g.hi[0]};
  const uint32_t vo = (uint32_t)x0 * es;
  int row;
  for (long it = blockIdx.x; next_row(g, it, row); it += gridDim.x) {
    const int y = row % g.ny, z = row / g.ny;
    if (y < g.lo[1] || y > g.hi[1] || z < g.lo[2] || z > g.hi[2]) continue;  // dead row
    // bulk of a split step: the band launch stores the band's cells (whole
    // x pairs)
    if (band && row_meets_band(band, y, z) && in_band_boxes(band, x0, y, z)) continue;
    const uint32_t row_b = (uint32_t)row * sy_b;
    const long c0 = (long)x0 + (long)g.nx * row;
    T v[2][19];
    if (filler) {
#pragma unroll
      for (int q = 0; q < 19; ++q) v[0][q] = v[1][q] = (T)0;
    } else {
      T d[2], rho[2], ux[2], uy[2], uz[2];
      if (own) {
        uint32_t qb = row_b;
#pragma unroll
        for (int q = 0; q < 19; ++q) {
          bld_pair<P, NTL>(rin, vo, qb, v[0][q], v[1][q]);
          qb += pitch_b;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          rho[k] = c.rho0; d[k] = c.d0; ux[k] = c.u0x; uy[k] = c.u0y; uz[k] = c.u0z;
        }
      } else {
        const uint32_t *rt = rowtab + (long)row * kRowTabStride;
        uint32_t so[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) so[q] = ldc(rt + q);
#pragma unroll
        for (int q = 0; q < 19; ++q) bld_pair<P, NTL>(rin, vo, so[q], v[0][q], v[1][q]);
        const uint32_t bbm = ldc(rt + 19);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          bounce_fix<P>(bbm, c.br[1], c.br[2], v[k]);
          T fpx = 0, fpy = 0, fpz = 0;
          if (in_any_box(box_prev, nsheets, x0 + k, y, z)) {
            fpx = (T)Fprev[c0 + k]; fpy = (T)Fprev[g.n + c0 + k]; fpz = (T)Fprev[2 * g.n + c0 + k];
          }
          lagged_moments<P>(v[k], fpx, fpy, fpz, c, d[k], rho[k], ux[k], uy[k], uz[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (frozen[k]) {
          // inlet/outlet plane: the table, with init_gv's rho/u (lbm_ib.c:2482-2483)
#pragma unroll
          for (int q = 0; q < 19; ++q) v[k][q] = ldc(c.tab + q);
          rho[k] = c.rho0; d[k] = c.d0; ux[k] = c.u0x; uy[k] = c.u0y; uz[k] = c.u0z;
        }
        T fx = 0, fy = 0, fz = 0;
        if (in_any_box(box_cur, nsheets, x0 + k, y, z)) {
          fx = (T)Fcur[c0 + k]; fy = (T)Fcur[g.n + c0 + k]; fz = (T)Fcur[2 * g.n + c0 + k];
        }
        collide_cell<P>(v[k], d[k], rho[k], ux[k], uy[k], uz[k], fx, fy, fz, c);
      }
    }
    uint32_t sb = row_b;
    if (NT && box_rows_cached && row_near_box(box_cur, nsheets, y, z)) {
      // rows meeting a support box (widened by the pulls' one cell) are
      // stored with the default cache policy: measured 6% faster for the
      // whole F32 launch than nontemporal stores on every row
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        bst_pair<P, false>(rout, vo, sb, v[0][q], v[1][q]);
        sb += pitch_b;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        bst_pair<P, NT>(rout, vo, sb, v[0][q], v[1][q]);
        sb += pitch_b;
      }
    }
  }
}

// One lane of the two-row path: cells (x, row0) and (x, row0 + 1), rows
// (y0, z) and (y0 + 1, z) of one z plane (row0 even, ny even), with two
// independent sets of aligned buffer loads, which doubles the bytes each wave
// keeps in flight.  Filler lanes (dead x, and the x edges of a wall or
// periodic x axis, left to k_collide_xedges) store zeros so that the rows'
// cache lines are written whole.
template <int P, bool NT, bool NTL>
__device__ __forceinline__ void row_pair_cell(const Geometry &g, __amdgpu_buffer_rsrc_t rin,
                                              __amdgpu_buffer_rsrc_t rout, int own,
                                              const KConsts<typename Traits<P>::T> &c, int nsheets,
                                              const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                                              const typename Traits<P>::FT *__restrict__ Fcur,
                                              const typename Traits<P>::FT *__restrict__ Fprev,
                                              const uint32_t *__restrict__ rowtab, int row0, int x,
                                              int box_rows_cached = 0) {
  using T = typename Traits<P>::T;
  using S = typename Traits<P>::S;
  constexpr uint32_t es = sizeof(S);
  const uint32_t pitch_b = (uint32_t)g.pq * es;   // between the populations of a cell
  const uint32_t sy_b = (uint32_t)g.prow * es;    // between rows (y, z) and (y + 1, z)
  const bool edge = x == g.lo[0] || x == g.hi[0];
  const bool frozen = edge && g.bc[0] == LBMIB_BC_INOUT;
  const bool filler = x < g.lo[0] || x > g.hi[0] || (edge && !frozen);
  const uint32_t vo = (uint32_t)x * es;
  const int y0 = row0 % g.ny, z = row0 / g.ny;  // rows row0, row0 + 1 = (y0, z), (y0 + 1, z)
  if (z < g.lo[2] || z > g.hi[2]) return;  // dead rows
  bool live[2];
  T v[2][19];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int y = y0 + k;
    live[k] = y >= g.lo[1] && y <= g.hi[1];
  }
  if (filler) {
#pragma unroll
    for (int q = 0; q < 19; ++q) v[0][q] = v[1][q] = (T)0;
  } else {
    T d[2], rho[2], ux[2], uy[2], uz[2];
    if (own) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        uint32_t qb = (uint32_t)(row0 + k) * sy_b;
#pragma unroll
        for (int q = 0; q < 19; ++q) {
          v[k][q] = bld_h<P, NTL>(rin, vo, qb);
          qb += pitch_b;
        }
        rho[k] = c.rho0; d[k] = c.d0; ux[k] = c.u0x; uy[k] = c.u0y; uz[k] = c.u0z;
      }
    } else {
      const uint32_t *rt = rowtab + (long)row0 * kRowTabStride;
      uint32_t so[2][19];
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int q = 0; q < 19; ++q) so[k][q] = ldc(rt + k * kRowTabStride + q);
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int q = 0; q < 19; ++q) v[k][q] = bld_h<P, NTL>(rin, vo, so[k][q]);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        bounce_fix<P>(ldc(rt + k * kRowTabStride + 19), c.br[1], c.br[2], v[k]);
        const long ck = (long)x + (long)g.nx * (row0 + k);
        T fpx = 0, fpy = 0, fpz = 0;
        if (in_any_box(box_prev, nsheets, x, y0 + k, z)) {
          fpx = (T)Fprev[ck]; fpy = (T)Fprev[g.n + ck]; fpz = (T)Fprev[2 * g.n + ck];
        }
        lagged_moments<P>(v[k], fpx, fpy, fpz, c, d[k], rho[k], ux[k], uy[k], uz[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const long ck = (long)x + (long)g.nx * (row0 + k);
      if (frozen) {
        // inlet/outlet plane: the table, with init_gv's rho/u (lbm_ib.c:2482-2483)
#pragma unroll
        for (int q = 0; q < 19; ++q) v[k][q] = ldc(c.tab + q);
        rho[k] = c.rho0; d[k] = c.d0; ux[k] = c.u0x; uy[k] = c.u0y; uz[k] = c.u0z;
      }
      T fx = 0, fy = 0, fz = 0;
      if (in_any_box(box_cur, nsheets, x, y0 + k, z)) {
        fx = (T)Fcur[ck]; fy = (T)Fcur[g.n + ck]; fz = (T)Fcur[2 * g.n + ck];
      }
      collide_cell<P>(v[k], d[k], rho[k], ux[k], uy[k], uz[k], fx, fy, fz, c);
    }
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (!live[k]) continue;  // a dead row of a wall y axis
    uint32_t sb = (uint32_t)(row0 + k) * sy_b;
    if (NT && box_rows_cached && row_near_box(box_cur, nsheets, y0 + k, z)) {
      // the rows meeting a support box: the default cache policy (as in
      // k_collide_stream_pairs)
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        bst_h<P, false>(rout, vo, sb, v[k][q]);
        sb += pitch_b;
      }
      continue;
    }
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      bst_h<P, NT>(rout, vo, sb, v[k][q]);
      sb += pitch_b;
    }
  }
}

// Two rows per block (y and y + 1 of one z plane): lane x updates (x, y, z)
// and (x, y + 1, z) (row_pair_cell), which doubles the bytes each wave keeps
// in flight without the unaligned pair loads of k_collide_stream_pairs.
// Requires ny even and blockDim.x >= nx; x-edge cells of a wall or periodic x
// axis are left to k_collide_xedges.
template <int P, bool NT, bool NTL>
__global__ void __launch_bounds__(512)
    k_collide_stream_rows2(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                           typename Traits<P>::S *__restrict__ fout, int own, KConsts<typename Traits<P>::T> c,
                           int nsheets, const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                           const typename Traits<P>::FT *__restrict__ Fcur,
                           const typename Traits<P>::FT *__restrict__ Fprev, const uint32_t *__restrict__ rowtab,
                           int box_rows_cached) {
  const int npairs = g.ny * g.nz / 2;
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(fin), rout = make_rsrc(fout);
  const int x = threadIdx.x;
  if (x >= g.nx) return;
  for (int rp = blockIdx.x; rp < npairs; rp += gridDim.x)
    row_pair_cell<P, NT, NTL>(g, rin, rout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, 2 * rp, x,
                              box_rows_cached);
}

// Batched ensemble: every instance's rows in one launch.  Block b serves
// instance map[b] >> kBatchShift; each wave of it one row (or one 64-lane x
// chunk of a row wider than 64: wpr waves per row), so the row offsets stay
// wave-uniform for any mix of grid sizes; wpb waves per block is the tile
// shape of the launch.  Instances need an inlet/outlet x axis (no x-edge
// kernel).
template <int P, bool NT, bool NTL>
__global__ void __launch_bounds__(512)
    k_collide_batch(const BatchDesc *__restrict__ descs, const int *__restrict__ map, int wpb, int cur, int own,
                    int slot, int box_rows_cached) {
  using T = typename Traits<P>::T;
  using FT = typename Traits<P>::FT;
  const int m = map[blockIdx.x];
  const int inst = m >> kBatchShift, lb = m & ((1 << kBatchShift) - 1);
  // The descriptor is read-only during the launch: its wave-uniform fields
  // come in with scalar loads, as kernel arguments do.
  const BatchDesc *dp = descs + inst;
  const int wpr = ldc(&dp->wpr), nrows = ldc(&dp->nrows);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int rpb = wpb / wpr;
  if (wave >= rpb * wpr) return;
  const int row = lb * rpb + wave / wpr;
  const int x = (wave % wpr) * 64 + lane;
  const Geometry g = ldc_obj(&dp->g);
  if (row >= nrows || x >= g.nx) return;
  const KConsts<T> c = ldc_obj(reinterpret_cast<const KConsts<T> *>(dp->kc));
  const int *box = ldc(&dp->ib.box);
  row_cell<P, NT, NTL>(g, make_rsrc(ldc(&dp->f[cur])), make_rsrc(ldc(&dp->f[1 - cur])), own, c, ldc(&dp->nsheets),
                       box + slot * kMaxSheets * 6, box + (1 - slot) * kMaxSheets * 6,
                       (const FT *)ldc(&dp->F[slot]), (const FT *)ldc(&dp->F[1 - slot]), ldc(&dp->rowtab), row, x,
                       nullptr, box_rows_cached);
}

// Two rows per lane: each wave serves one 64-lane x chunk of a row pair
// (row0, row0 + 1) of one z plane (row_pair_cell), so that a wave keeps
// twice the bytes in flight; rpb = wpb / wpr row pairs per block.  Every
// instance's ny is even.
template <int P, bool NT, bool NTL>
__global__ void __launch_bounds__(512)
    k_collide_batch2(const BatchDesc *__restrict__ descs, const int *__restrict__ map, int wpb, int cur, int own,
                     int slot, int box_rows_cached) {
  using T = typename Traits<P>::T;
  using FT = typename Traits<P>::FT;
  const int m = map[blockIdx.x];
  const int inst = m >> kBatchShift, lb = m & ((1 << kBatchShift) - 1);
  const BatchDesc *dp = descs + inst;
  const int wpr = ldc(&dp->wpr), nrows = ldc(&dp->nrows);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int rpb = wpb / wpr;
  if (wave >= rpb * wpr) return;
  const int rp = lb * rpb + wave / wpr;
  const int x = (wave % wpr) * 64 + lane;
  const Geometry g = ldc_obj(&dp->g);
  if (2 * rp >= nrows || x >= g.nx) return;
  const KConsts<T> c = ldc_obj(reinterpret_cast<const KConsts<T> *>(dp->kc));
  const int *box = ldc(&dp->ib.box);
  row_pair_cell<P, NT, NTL>(g, make_rsrc(ldc(&dp->f[cur])), make_rsrc(ldc(&dp->f[1 - cur])), own, c,
                            ldc(&dp->nsheets), box + slot * kMaxSheets * 6, box + (1 - slot) * kMaxSheets * 6,
                            (const FT *)ldc(&dp->F[slot]), (const FT *)ldc(&dp->F[1 - slot]), ldc(&dp->rowtab),
                            2 * rp, x, box_rows_cached);
}

// The band of a split step: every cell of the band boxes (written by the
// step's fiber kernel) by the general rules, each once (a cell in several
// boxes goes with the first).  Threads stride over the cells of all boxes.
template <int P, bool NT, bool NTL>
__global__ void __launch_bounds__(256)
    k_collide_band(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                   typename Traits<P>::S *__restrict__ fout, int own, KConsts<typename Traits<P>::T> c, int nsheets,
                   const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                   const typename Traits<P>::FT *__restrict__ Fcur, const typename Traits<P>::FT *__restrict__ Fprev,
                   const int *__restrict__ band) {
  const int nb = min(ldc(band), kMaxSheets);
  long total = 0;
  for (int k = 0; k < nb; ++k) {
    const int *r = band + 1 + 6 * k;
    total += (long)(ldc(r + 1) - ldc(r) + 1) * (ldc(r + 3) - ldc(r + 2) + 1) * (ldc(r + 5) - ldc(r + 4) + 1);
  }
  for (long it = (long)blockIdx.x * blockDim.x + threadIdx.x; it < total; it += (long)gridDim.x * blockDim.x) {
    long rest = it;
    int k = 0;
    int r[6];
    for (; k < nb; ++k) {
      for (int a = 0; a < 6; ++a) r[a] = ldc(band + 1 + 6 * k + a);
      const long vol = (long)(r[1] - r[0] + 1) * (r[3] - r[2] + 1) * (r[5] - r[4] + 1);
      if (rest < vol) break;
      rest -= vol;
    }
    if (k == nb) break;  // not reached: it < total
    // within one box (a part of the grid: under 2^31 cells) in 32 bits
    const int w = r[1] - r[0] + 1, h = r[3] - r[2] + 1, e = (int)rest;
    const int x = r[0] + e % w, y = r[2] + (e / w) % h, z = r[4] + e / (w * h);
    bool dup = false;
    for (int t = 0; t < k; ++t) {
      int q[6];
      for (int a = 0; a < 6; ++a) q[a] = ldc(band + 1 + 6 * t + a);
      dup |= in_box3(q, x, y, z);
    }
    if (!dup) update_cell<P, NT, NTL>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, x, y, z);
  }
}

// The x-edge cells (x = lo, hi) of every active row when x is a wall or a
// periodic axis, by the general rules; completes the buffer-addressed rows.
template <int P, bool NT, bool NTL>
__global__ void __launch_bounds__(256)
    k_collide_xedges(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                     typename Traits<P>::S *__restrict__ fout, int own, KConsts<typename Traits<P>::T> c,
                     int nsheets, const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                     const typename Traits<P>::FT *__restrict__ Fcur,
                     const typename Traits<P>::FT *__restrict__ Fprev, const int *__restrict__ band) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2L * g.ny * g.nz) return;
  const long row = e >> 1;
  const int x = (e & 1) ? g.hi[0] : g.lo[0];
  if ((e & 1) && g.lo[0] == g.hi[0]) return;  // one-cell-wide axis: done by the even item
  if (band && in_band_boxes(band, x, (int)(row % g.ny), (int)(row / g.ny))) return;  // stored by the band launch
  update_cell<P, NT, NTL>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, x, (int)(row % g.ny),
                          (int)(row / g.ny));
}

// V consecutive cells of one x-row per thread: the 9 populations with
// c_x = 0 are read with aligned V-wide loads, the 10 x-shifted ones with
// one-element-offset V-wide loads (dword aligned), and all 19 results are
// written with aligned V-wide stores, so each wave moves 64*V*sizeof(S)
// contiguous bytes per memory instruction.  Threads whose cells touch a
// boundary (frozen planes, walls, periodic wrap, the first step) fall back
// to the per-cell path.  Requires nx % V == 0.
template <class S, int V>
struct alignas(sizeof(S) * V) VecT {
  S a[V];
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Aligned store of a VecT as native 16- or 8-byte vectors (nontemporal when
// NT: the populations are not re-read before the next step).
template <bool NT, class VT>
__device__ __forceinline__ void store_vec(VT *dst, const VT &t) {
  constexpr int B = sizeof(VT);
  if constexpr (B % 16 == 0) {
    u32x4 tmp[B / 16];
    __builtin_memcpy(tmp, &t, B);
    u32x4 *d = reinterpret_cast<u32x4 *>(dst);
#pragma unroll
    for (int i = 0; i < B / 16; ++i) {
      if constexpr (NT) __builtin_nontemporal_store(tmp[i], d + i);
      else d[i] = tmp[i];
    }
  } else if constexpr (B == 8) {
    u32x2 tmp;
    __builtin_memcpy(&tmp, &t, 8);
    if constexpr (NT) __builtin_nontemporal_store(tmp, reinterpret_cast<u32x2 *>(dst));
    else *reinterpret_cast<u32x2 *>(dst) = tmp;
  } else {
    *dst = t;
  }
}

template <int P, int V, bool NT>
__global__ void __launch_bounds__(256)
    k_collide_stream_vec(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                         typename Traits<P>::S *__restrict__ fout, int own, KConsts<typename Traits<P>::T> c,
                         int nsheets, const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                         const typename Traits<P>::FT *__restrict__ Fcur,
                         const typename Traits<P>::FT *__restrict__ Fprev) {
  using T = typename Traits<P>::T;
  using S = typename Traits<P>::S;
  const int x0 = threadIdx.x * V;
  const int y = blockIdx.x * blockDim.y + threadIdx.y, z = blockIdx.y;
  if (y >= g.ny || x0 >= g.nx) return;
  const bool fast = !own && x0 - 1 >= g.lo[0] && x0 + V <= g.hi[0] && y > g.lo[1] && y < g.hi[1] && z > g.lo[2] &&
                    z < g.hi[2];
  if (!fast) {
#pragma unroll 1
    for (int k = 0; k < V; ++k)
      update_cell<P, NT>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, x0 + k, y, z);
    return;
  }
  const long c0 = (long)x0 + (long)g.nx * ((long)y + (long)g.ny * z);
  const long p0 = pop_index(g, x0, y, z);
  const long sy = g.prow, sz = g.prow * g.ny;
  T v[V][19];
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const S *p = fin + q * g.pq + p0 - (kCX[q] + kCY[q] * sy + kCZ[q] * sz);
    VecT<S, V> t;
    if (kCX[q] == 0) t = *reinterpret_cast<const VecT<S, V> *>(p);
    else __builtin_memcpy(&t, p, sizeof(t));
#pragma unroll
    for (int k = 0; k < V; ++k) v[k][q] = from_bits<P>(t.a[k]);
  }
  const bool inprev = nsheets > 0 && any_box_overlaps(box_prev, nsheets, x0, x0 + V - 1, y, z);
  const bool incur = nsheets > 0 && any_box_overlaps(box_cur, nsheets, x0, x0 + V - 1, y, z);
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const int x = x0 + k;
    T fpx = 0, fpy = 0, fpz = 0, fx = 0, fy = 0, fz = 0;
    if (inprev && in_any_box(box_prev, nsheets, x, y, z)) {
      fpx = (T)Fprev[c0 + k]; fpy = (T)Fprev[g.n + c0 + k]; fpz = (T)Fprev[2 * g.n + c0 + k];
    }
    if (incur && in_any_box(box_cur, nsheets, x, y, z)) {
      fx = (T)Fcur[c0 + k]; fy = (T)Fcur[g.n + c0 + k]; fz = (T)Fcur[2 * g.n + c0 + k];
    }
    T d, rho, ux, uy, uz;
    lagged_moments<P>(v[k], fpx, fpy, fpz, c, d, rho, ux, uy, uz);
    collide_cell<P>(v[k], d, rho, ux, uy, uz, fx, fy, fz, c);
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    VecT<S, V> t;
#pragma unroll
    for (int k = 0; k < V; ++k) t.a[k] = to_bits<P>(v[k][q]);
    store_vec<NT>(reinterpret_cast<VecT<S, V> *>(fout + q * g.pq + p0), t);
  }
}

// Fills the active box with the equilibrium (init_eqlbrmdistrfuncDF0 +
// init_df1, lbm_ib.c:1528-1606, 2725-2767); other cells are zero.
template <int P>
__global__ void k_fill_init(Geometry g, typename Traits<P>::S *__restrict__ f, KConsts<typename Traits<P>::T> c) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= g.nx) return;
  const long p0 = pop_index(g, x, y, z);
  const bool act = x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2];
#pragma unroll
  for (int q = 0; q < 19; ++q) st_h<P>(f + q * g.pq + p0, act ? c.tab[q] : (typename Traits<P>::T)0);
}

// Materialises df2 (= df1 after the swap), rho and vel of the last completed
// step for every cell, in FP64 and in the reference's conventions.
template <int P>
__global__ void k_materialize(Geometry g, const typename Traits<P>::S *__restrict__ f, int own,
                              KConsts<typename Traits<P>::T> c, int nsheets, const int *__restrict__ box,
                              const typename Traits<P>::FT *__restrict__ F, double *__restrict__ df,
                              double *__restrict__ rho_out, double *__restrict__ vel) {
  using T = typename Traits<P>::T;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= g.nx) return;
  const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
  const bool act = x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2];
  const bool mom = x >= g.mlo[0] && x <= g.mhi[0] && y >= g.mlo[1] && y <= g.mhi[1] && z >= g.mlo[2] && z <= g.mhi[2];
  double fv[19];
  if (act) {
    T v[19];
    const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
    if (frozen) {
      for (int q = 0; q < 19; ++q) v[q] = c.tab[q] * Traits<P>::inv_scale;
    } else if (own) {
      for (int q = 0; q < 19; ++q) v[q] = ld<P>(f + q * g.pq + pop_index(g, x, y, z));
    } else {
      pull<P>(g, f, x, y, z, c.br[1], c.br[2], v);
    }
    for (int q = 0; q < 19; ++q) fv[q] = Traits<P>::dev ? (double)v[q] + c.ref[q] : (double)v[q];
  } else {
    for (int q = 0; q < 19; ++q) fv[q] = 0.0;
  }
  if (df)
    for (int q = 0; q < 19; ++q) df[q * g.n + c0] = fv[q];
  double r = c.rho0, ux = c.u0x, uy = c.u0y, uz = c.u0z;
  if (mom && !own) {
    double fx = 0, fy = 0, fz = 0;
    if (in_any_box(box, nsheets, x, y, z)) { fx = F[c0]; fy = F[g.n + c0]; fz = F[2 * g.n + c0]; }
    double s1 = fv[0], s2 = 0, s3 = 0, s4 = 0;
    for (int q = 1; q < 19; ++q) {
      s1 += fv[q];
      if (kCX[q]) s2 = kCX[q] > 0 ? s2 + fv[q] : s2 - fv[q];
      if (kCY[q]) s3 = kCY[q] > 0 ? s3 + fv[q] : s3 - fv[q];
      if (kCZ[q]) s4 = kCZ[q] > 0 ? s4 + fv[q] : s4 - fv[q];
    }
    r = s1;
    ux = (s2 + 0.5 * fx) / s1;
    uy = (s3 + 0.5 * fy) / s1;
    uz = (s4 + 0.5 * (fz + r * (double)c.g)) / s1;
  }
  if (rho_out) rho_out[c0] = r;
  if (vel) {
    vel[c0] = ux;
    vel[g.n + c0] = uy;
    vel[2 * g.n + c0] = uz;
  }
}

// Deterministic FP64 totals: fixed-shape block tree, then one block sums the
// partials in index order.
constexpr int kTotThreads = 256;
constexpr int kTotRowsPerBlock = 16;

template <int P>
__global__ void __launch_bounds__(kTotThreads)
    k_totals_partial(Geometry g, const typename Traits<P>::S *__restrict__ f, int own,
                     KConsts<typename Traits<P>::T> c, double *__restrict__ partials) {
  using T = typename Traits<P>::T;
  __shared__ double red[4][kTotThreads];
  double acc[4] = {0, 0, 0, 0};
  const long rows = (long)g.ny * g.nz;
  for (int r = 0; r < kTotRowsPerBlock; ++r) {
    const long row = (long)blockIdx.x * kTotRowsPerBlock + r;
    if (row >= rows) break;
    const int y = (int)(row % g.ny), z = (int)(row / g.ny);
    for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
      if (x < g.lo[0] || x > g.hi[0] || y < g.lo[1] || y > g.hi[1] || z < g.lo[2] || z > g.hi[2]) continue;
      T v[19];
      const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
      if (frozen) {
        for (int q = 0; q < 19; ++q) v[q] = c.tab[q] * Traits<P>::inv_scale;
      } else if (own) {
        for (int q = 0; q < 19; ++q) v[q] = ld<P>(f + q * g.pq + pop_index(g, x, y, z));
      } else {
        pull<P>(g, f, x, y, z, c.br[1], c.br[2], v);
      }
      for (int q = 0; q < 19; ++q) {
        const double fq = Traits<P>::dev ? (double)v[q] + c.ref[q] : (double)v[q];
        acc[0] += fq;
        acc[1] += kCX[q] * fq;
        acc[2] += kCY[q] * fq;
        acc[3] += kCZ[q] * fq;
      }
    }
  }
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = kTotThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) partials[(long)blockIdx.x * 4 + k] = red[k][0];
}

__global__ void __launch_bounds__(kTotThreads) k_totals_final(const double *__restrict__ partials, int nb,
                                                               double *__restrict__ out4) {
  __shared__ double red[4][kTotThreads];
  double acc[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int k = 0; k < 4; ++k) acc[k] += partials[(long)b * 4 + k];
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = kTotThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) out4[k] = red[k][0];
}

// ------------------------------------------------------------ launchers

// Row-aligned blocks: one block covers a contiguous run of x in one (y, z)
// row, so the one-element x shifts of the pull stay inside the block's own
// cache lines.
static inline dim3 row_block(const Geometry &g) {
  int bx;

  if (g.nx % 64 == 0 && g.nx >= 256 && g.nx <= 1024) bx = g.nx;
  else if (g.nx > 1024) bx = 256;
  else bx = ((g.nx + 63) / 64) * 64;
  return dim3(bx, 1, 1);
}
static inline dim3 row_grid(const Geometry &g, dim3 b) { return dim3((g.nx + b.x - 1) / b.x, g.ny, g.nz); }

#define LBMIB_DISPATCH(prec, CALL)                   \
  switch (prec) {                                    \
    case LBMIB_F64: { constexpr int P = LBMIB_F64; CALL; break; }  \
    case LBMIB_F32: { constexpr int P = LBMIB_F32; CALL; break; }  \
    case LBMIB_F16: { constexpr int P = LBMIB_F16; CALL; break; }  \
    case LBMIB_BF16: { constexpr int P = LBMIB_BF16; CALL; break; } \
    default: return LBMIB_ERR_ARG;                   \
  }

int launch_fill_init(int prec, const Geometry &g, void *f, const FluidConsts<double> &c, hipStream_t s) {
  dim3 b = row_block(g), gr = row_grid(g, b);
  LBMIB_DISPATCH(prec, ({
    k_fill_init<P><<<gr, b, 0, s>>>(g, (typename Traits<P>::S *)f, kconsts<P>(c));
  }));
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

// Cells per thread of the fused kernel: at most 16 bytes per lane (wider
// vectors push the kernel past 256 VGPRs).  Falls back to one cell per
// thread when nx is not a multiple of the width or a row would need more
// than 256 threads.
int collide_vec_width(int prec, int nx, int vec_req) {
  int v = vec_req;
  if (v <= 0) v = 1;
  const int max_v = prec == LBMIB_F64 ? 2 : (prec == LBMIB_F32 ? 4 : 8);
  const int min_v = (prec == LBMIB_F16 || prec == LBMIB_BF16) ? 4 : 2;
  if (v == 1 || v > max_v || v < min_v || (v & (v - 1))) return 1;
  if (nx % v != 0 || nx / v > 256) return 1;
  return v;
}

template <int P, int V, bool NT>
static void launch_vec(const Geometry &g, const void *fin, void *fout, int own, const FluidConsts<double> &c,
                       int nsheets, const int *box_cur, const int *box_prev, const void *Fcur, const void *Fprev,
                       hipStream_t s) {
  using FT = typename Traits<P>::FT;
  using S = typename Traits<P>::S;
  const int bx = g.nx / V;
  const int by = bx >= 256 ? 1 : 256 / bx;
  dim3 b(bx, by, 1), gr((g.ny + by - 1) / by, g.nz, 1);
  k_collide_stream_vec<P, V, NT><<<gr, b, 0, s>>>(g, (const S *)fin, (S *)fout, own, kconsts<P>(c), nsheets,
                                                  box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev);
}

// Whether the buffer-addressed row path applies: every byte offset of a
// population array fits 32 bits and one block covers a row.
bool row_path_buffered(const Geometry &g, size_t es) {
  return (unsigned long long)g.pelems * es <= 0xFFFFFFFFull && (int)row_block(g).x >= g.nx;
}

template <int P, bool NT, bool NTL>
static void launch_rows2(int nblocks, const Geometry &g, const void *fin, void *fout, int own,
                         const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                         const void *Fcur, const void *Fprev, const uint32_t *rowtab, int box_rows_cached,
                         hipStream_t s) {
  using FT = typename Traits<P>::FT;
  using S = typename Traits<P>::S;
  const int npairs = g.ny * g.nz / 2;
  const dim3 b((unsigned)((g.nx + 63) / 64 * 64));
  const dim3 gr(nblocks < npairs ? nblocks : npairs);
  k_collide_stream_rows2<P, NT, NTL><<<gr, b, 0, s>>>(g, (const S *)fin, (S *)fout, own, kconsts<P>(c), nsheets,
                                                      box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev, rowtab,
                                                      box_rows_cached);
  if (g.bc[0] != LBMIB_BC_INOUT) {  // whole-grid steps only
    const long items = 2L * g.ny * g.nz;
    k_collide_xedges<P, NT, NTL><<<(unsigned)((items + 255) / 256), 256, 0, s>>>(
        g, (const S *)fin, (S *)fout, own, kconsts<P>(c), nsheets, box_cur, box_prev, (const FT *)Fcur,
        (const FT *)Fprev, nullptr);
  }
}

// Whether the two-rows-per-block path applies (see k_collide_stream_rows2).
bool rows2_path_ok(const Geometry &g) { return g.ny % 2 == 0 && g.nx <= 512; }

// Whether the two-cells-per-lane row path applies (see k_collide_stream_pairs).
bool pair_path_ok(const Geometry &g) {
  return g.nx % 2 == 0 && g.lo[0] % 2 == 0 && g.hi[0] % 2 == 1 && g.nx / 2 <= 512;
}

template <int P, bool NT, bool NTL>
static void launch_pairs(int nblocks, const Geometry &g, const void *fin, void *fout, int own,
                         const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                         const void *Fcur, const void *Fprev, const uint32_t *rowtab, const int *band,
                         int box_rows_cached, hipStream_t s) {
  using FT = typename Traits<P>::FT;
  using S = typename Traits<P>::S;
  const int nrows = g.ny * g.nz;
  const dim3 b((unsigned)(((g.nx / 2) + 63) / 64 * 64));
  const dim3 gr(nblocks < nrows ? nblocks : nrows);
  k_collide_stream_pairs<P, NT, NTL><<<gr, b, 0, s>>>(g, (const S *)fin, (S *)fout, own, kconsts<P>(c), nsheets,
                                                      box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev, rowtab,
                                                      band, box_rows_cached);
  if (g.bc[0] != LBMIB_BC_INOUT) {
    const long items = 2L * g.ny * g.nz;
    k_collide_xedges<P, NT, NTL><<<(unsigned)((items + 255) / 256), 256, 0, s>>>(
        g, (const S *)fin, (S *)fout, own, kconsts<P>(c), nsheets, box_cur, box_prev, (const FT *)Fcur,
        (const FT *)Fprev, band);
  }
}

template <int P, bool NT, bool NTL>
static void launch_rows(int nblocks, const Geometry &g, const void *fin, void *fout, int own,
                        const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                        const void *Fcur, const void *Fprev, const uint32_t *rowtab, const int *band,
                        int box_rows_cached, hipStream_t s) {
  using FT = typename Traits<P>::FT;
  using S = typename Traits<P>::S;
  const dim3 b = row_block(g);
  const int nrows = g.ny * g.nz;
  const dim3 gr(nblocks < nrows ? nblocks : nrows);
  if (rowtab && row_path_buffered(g, sizeof(S))) {
    k_collide_stream_rows<P, NT, NTL, true><<<gr, b, 0, s>>>(g, (const S *)fin, (S *)fout, own, kconsts<P>(c), nsheets,
                                                             box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev,
                                                             rowtab, band, box_rows_cached);
    if (g.bc[0] != LBMIB_BC_INOUT) {
      const long items = 2L * g.ny * g.nz;
      k_collide_xedges<P, NT, NTL><<<(unsigned)((items + 255) / 256), 256, 0, s>>>(
          g, (const S *)fin, (S *)fout, own, kconsts<P>(c), nsheets, box_cur, box_prev, (const FT *)Fcur,
          (const FT *)Fprev, band);
    }
  } else
    k_collide_stream_rows<P, NT, NTL, false><<<gr, b, 0, s>>>(g, (const S *)fin, (S *)fout, own, kconsts<P>(c),
                                                              nsheets, box_cur, box_prev, (const FT *)Fcur,
                                                              (const FT *)Fprev, nullptr, band, box_rows_cached);
}

template <int P, bool NT>
static void launch_any(int v, const Geometry &g, const void *fin, void *fout, int own, const FluidConsts<double> &c,
                       int nsheets, const int *box_cur, const int *box_prev, const void *Fcur, const void *Fprev,
                       hipStream_t s) {
  if constexpr (P == LBMIB_F64) {
    if (v == 2) return launch_vec<P, 2, NT>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, s);
  } else if constexpr (P == LBMIB_F32) {
    if (v == 2) return launch_vec<P, 2, NT>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, s);
    if (v == 4) return launch_vec<P, 4, NT>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, s);
  } else {
    if (v == 4) return launch_vec<P, 4, NT>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, s);
    if (v == 8) return launch_vec<P, 8, NT>(g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, s);
  }
  launch_rows<P, NT, false>(g.ny * g.nz, g, fin, fout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, nullptr,
                            nullptr, 0, s);
}

int launch_collide_stream(int prec, const Geometry &g, const void *fin, void *fout, int own_mode,
                          const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                          const void *Fcur, const void *Fprev, const uint32_t *rowtab, const KernelVariant &kv,
                          hipStream_t s, int sel, int *rect_buf, int band_blocks) {
  if (sel != kRowsAll && (!rowtab || !rect_buf)) return LBMIB_ERR_ARG;
  // nontemporal stores and loads by default: the populations are touched
  // once per step, and keeping them out of the caches was fastest measured
  const int nt = kv.nt_mask < 0 ? 1 : (kv.nt_mask & 1);
  const int ntl = kv.nt_mask < 0 ? 1 : ((kv.nt_mask >> 1) & 1);
  // automatic shape: x pairs for the 16-bit forms (4-byte lanes), one cell
  // per lane otherwise -- the fastest measured at 256^3
  if (sel == kRowsBand) {
    // the band of a split step, cell by cell (band_blocks blocks of 256)
    LBMIB_DISPATCH(prec, ({
      using S = typename Traits<P>::S;
      using FT = typename Traits<P>::FT;
      const dim3 gr((unsigned)(band_blocks > 0 ? band_blocks : 1));
      if (nt && ntl) k_collide_band<P, true, true><<<gr, 256, 0, s>>>(g, (const S *)fin, (S *)fout, own_mode, kconsts<P>(c), nsheets, box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev, rect_buf);
      else if (nt) k_collide_band<P, true, false><<<gr, 256, 0, s>>>(g, (const S *)fin, (S *)fout, own_mode, kconsts<P>(c), nsheets, box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev, rect_buf);
      else if (ntl) k_collide_band<P, false, true><<<gr, 256, 0, s>>>(g, (const S *)fin, (S *)fout, own_mode, kconsts<P>(c), nsheets, box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev, rect_buf);
      else k_collide_band<P, false, false><<<gr, 256, 0, s>>>(g, (const S *)fin, (S *)fout, own_mode, kconsts<P>(c), nsheets, box_cur, box_prev, (const FT *)Fcur, (const FT *)Fprev, rect_buf);
    }));
    return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
  }
  // the bulk of a split step skips the band's cells
  const int *band = sel == kRowsBulk ? rect_buf : nullptr;
  int cpt = kv.cells_per_thread;
  if (cpt <= 0) cpt = ((prec == LBMIB_F16 || prec == LBMIB_BF16) && rowtab && pair_path_ok(g)) ? 2 : 1;
  if (cpt == 3 && sel != kRowsAll) cpt = 1;  // split steps use the row and pair kernels
  if (cpt == 3 && rowtab && rows2_path_ok(g)) {
    const int nb = kv.blocks > 0 ? kv.blocks : g.ny * g.nz / 2;
    LBMIB_DISPATCH(prec, ({
      if (nt && ntl) launch_rows2<P, true, true>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, kv.box_rows_cached, s);
      else if (nt) launch_rows2<P, true, false>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, kv.box_rows_cached, s);
      else if (ntl) launch_rows2<P, false, true>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, kv.box_rows_cached, s);
      else launch_rows2<P, false, false>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, kv.box_rows_cached, s);
    }));
    return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
  }
  if (cpt == 2 && rowtab && pair_path_ok(g)) {
    const int nb = kv.blocks > 0 ? kv.blocks : g.ny * g.nz;
    LBMIB_DISPATCH(prec, ({
      if (nt && ntl) launch_pairs<P, true, true>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
      else if (nt) launch_pairs<P, true, false>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
      else if (ntl) launch_pairs<P, false, true>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
      else launch_pairs<P, false, false>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
    }));
    return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
  }
  const int v = sel != kRowsAll ? 1 : collide_vec_width(prec, g.nx, cpt);
  if (v == 1) {
    const int nb = kv.blocks > 0 ? kv.blocks : g.ny * g.nz;
    LBMIB_DISPATCH(prec, ({
      if (nt && ntl) launch_rows<P, true, true>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
      else if (nt) launch_rows<P, true, false>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
      else if (ntl) launch_rows<P, false, true>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
      else launch_rows<P, false, false>(nb, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, band, kv.box_rows_cached, s);
    }));
    return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
  }
  LBMIB_DISPATCH(prec, ({
    if (nt) launch_any<P, true>(v, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, s);
    else launch_any<P, false>(v, g, fin, fout, own_mode, c, nsheets, box_cur, box_prev, Fcur, Fprev, s);
  }));
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

// ------------------------------------------------ step with coarse moments
//
// One step's fluid update in the row form (one block per (y, z) row, one
// lane per x, every cell updated as row_cell updates it) that also reports
// the stored moments of the state it writes, taken from the post-collision
// values the stores round, as per-row partial sums (lbmib_coarse.h; cells
// that are not updated report rho0/u0 like the stored read).  The planes are
// lbmib_coarse_state's stored planes of the new state, bit for bit.  Needs
// the buffer-addressed row path with a block as wide as the row, and an
// inlet/outlet x axis (the x edges of other axes are updated by a second
// launch, after the sums).
template <int P, bool NT, bool NTL, int F>
__global__ void __launch_bounds__(1024)
    k_collide_stream_rows_coarse(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                                 typename Traits<P>::S *__restrict__ fout, int own, KConsts<typename Traits<P>::T> c,
                                 int nsheets, const int *__restrict__ box_cur, const int *__restrict__ box_prev,
                                 const typename Traits<P>::FT *__restrict__ Fcur,
                                 const typename Traits<P>::FT *__restrict__ Fprev,
                                 const uint32_t *__restrict__ rowtab, int box_rows_cached, double rhor, double mrx,
                                 double rho0, double u0x, double u0y, double u0z, float *__restrict__ part) {
  using T = typename Traits<P>::T;
  using S = typename Traits<P>::S;
  using A = JumpT<P>;
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(fin), rout = make_rsrc(fout);
  const int x = threadIdx.x;  // blockDim.x == nx, a multiple of 64
  const long ncx = g.nx / F;
  int row;
  for (long it = blockIdx.x; next_row(g, it, row); it += gridDim.x) {
    A sr = (A)rho0, sx = (A)u0x, sy = (A)u0y, sz = (A)u0z;
    row_cell<P, NT, NTL>(g, rin, rout, own, c, nsheets, box_cur, box_prev, Fcur, Fprev, rowtab, row, x, nullptr,
                         box_rows_cached, [&](const T *v) {
                           S h[19];
#pragma unroll
                           for (int q = 0; q < 19; ++q) h[q] = to_bits<P>(v[q]);
                           stored_moments<P>(h, rhor, mrx, sr, sx, sy, sz);
                         });
    float m[4] = {(float)sr, (float)sx, (float)sy, (float)sz};
    row_partial<F>(m, x, row, ncx, part);
  }
}

bool collide_coarse_ok(const Geometry &g, size_t es, bool rowtab, int F) {
  return rowtab && row_path_buffered(g, es) && (int)row_block(g).x == g.nx && g.bc[0] == LBMIB_BC_INOUT &&
         (F == 1 || F == 2 || F == 4) && g.nx % 64 == 0 && g.ny % 4 == 0 && g.nz % 4 == 0;
}

int launch_collide_coarse(int prec, const Geometry &g, const void *fin, void *fout, int own_mode,
                          const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                          const void *Fcur, const void *Fprev, const uint32_t *rowtab, const KernelVariant &kv,
                          int F, float *part, float *stored, hipStream_t s) {
  if (!part || !stored) return LBMIB_ERR_ARG;
  const int nt = kv.nt_mask < 0 ? 1 : (kv.nt_mask & 1);
  const int ntl = kv.nt_mask < 0 ? 1 : ((kv.nt_mask >> 1) & 1);
  const int nrows = g.ny * g.nz;
  const int nb = kv.blocks > 0 && kv.blocks < nrows ? kv.blocks : nrows;
  const double mrx = c.rhor * c.urx;
  LBMIB_DISPATCH(prec, ({
    using S = typename Traits<P>::S;
    using FT = typename Traits<P>::FT;
    if (!collide_coarse_ok(g, sizeof(S), rowtab != nullptr, F)) return LBMIB_ERR_ARG;
    auto go = [&](auto kern) {
      kern<<<nb, g.nx, 0, s>>>(g, (const S *)fin, (S *)fout, own_mode, kconsts<P>(c), nsheets, box_cur, box_prev,
                               (const FT *)Fcur, (const FT *)Fprev, rowtab, kv.box_rows_cached, c.rhor, mrx, c.rho0,
                               c.u0x, c.u0y, c.u0z, part);
    };
    auto pick = [&](auto nt_c, auto ntl_c) {
      constexpr bool NTc = decltype(nt_c)::value, NTLc = decltype(ntl_c)::value;
      if (F == 1) go(k_collide_stream_rows_coarse<P, NTc, NTLc, 1>);
      else if (F == 2) go(k_collide_stream_rows_coarse<P, NTc, NTLc, 2>);
      else go(k_collide_stream_rows_coarse<P, NTc, NTLc, 4>);
    };
    if (nt && ntl) pick(std::true_type{}, std::true_type{});
    else if (nt) pick(std::true_type{}, std::false_type{});
    else if (ntl) pick(std::false_type{}, std::true_type{});
    else pick(std::false_type{}, std::false_type{});
  }));
  if (hipGetLastError() != hipSuccess) return LBMIB_ERR_HIP;
  return launch_coarse_rows(g, F, part, stored, s) == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int launch_materialize(int prec, const Geometry &g, const void *f, int own_mode, const FluidConsts<double> &c,
                       int nsheets, const int *box, const void *F, double *df, double *rho, double *vel,
                       hipStream_t s) {
  dim3 b = row_block(g), gr = row_grid(g, b);
  LBMIB_DISPATCH(prec, ({
    using FT = typename Traits<P>::FT;
    k_materialize<P><<<gr, b, 0, s>>>(g, (const typename Traits<P>::S *)f, own_mode, kconsts<P>(c), nsheets,
                                       box, (const FT *)F, df, rho, vel);
  }));
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int totals_blocks(const Geometry &g) {
  const long rows = (long)g.ny * g.nz;
  return (int)((rows + kTotRowsPerBlock - 1) / kTotRowsPerBlock);
}

int launch_totals(int prec, const Geometry &g, const void *f, int own_mode, const FluidConsts<double> &c,
                  double *partials, int nblocks_max, double *out4, hipStream_t s) {
  const int nb = totals_blocks(g);
  if (nb > nblocks_max) return LBMIB_ERR_ARG;
  LBMIB_DISPATCH(prec, ({
    k_totals_partial<P><<<nb, kTotThreads, 0, s>>>(g, (const typename Traits<P>::S *)f, own_mode,
                                                    kconsts<P>(c), partials);
  }));
  k_totals_final<<<1, kTotThreads, 0, s>>>(partials, nb, out4);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int kconsts_bytes(int prec, const FluidConsts<double> &c, unsigned char out[256]) {
  memset(out, 0, 256);
  LBMIB_DISPATCH(prec, ({
    const KConsts<typename Traits<P>::T> k = kconsts<P>(c);
    static_assert(sizeof(k) <= 256, "KConsts fits the batch descriptor");
    memcpy(out, &k, sizeof(k));
  }));
  return LBMIB_OK;
}

int upload_fluid_tables(int prec, const FluidConsts<double> &c, void *d_tab, double *d_ref) {
  LBMIB_DISPATCH(prec, ({
    using TT = typename Traits<P>::T;
    TT t[19];
    // kernel units: f (F64), s (f - ref) (deviation forms; tab_in holds f - ref)
    for (int q = 0; q < 19; ++q) t[q] = (TT)(Traits<P>::scale * c.tab_in[q]);
    if (hipMemcpy(d_tab, t, sizeof(t), hipMemcpyHostToDevice) != hipSuccess) return LBMIB_ERR_HIP;
    if (hipMemcpy(d_ref, c.ref, sizeof(double) * 19, hipMemcpyHostToDevice) != hipSuccess) return LBMIB_ERR_HIP;
  }));
  return LBMIB_OK;
}

int launch_collide_batch(int prec, const BatchDesc *d, const int *map, int nblocks, int wpb, int cur, int own,
                         int slot, int nt_mask, int rows, int box_rows_cached, hipStream_t s) {
  if (nblocks <= 0) return LBMIB_OK;
  if (wpb < 1 || wpb > 8 || (rows != 1 && rows != 2)) return LBMIB_ERR_ARG;
  const int nt = nt_mask < 0 ? 1 : (nt_mask & 1);
  const int ntl = nt_mask < 0 ? 1 : ((nt_mask >> 1) & 1);
  const dim3 b(64 * wpb);
  if (rows == 2) {
    LBMIB_DISPATCH(prec, ({
      if (nt && ntl) k_collide_batch2<P, true, true><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
      else if (nt) k_collide_batch2<P, true, false><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
      else if (ntl) k_collide_batch2<P, false, true><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
      else k_collide_batch2<P, false, false><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
    }));
    return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
  }
  LBMIB_DISPATCH(prec, ({
    if (nt && ntl) k_collide_batch<P, true, true><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
    else if (nt) k_collide_batch<P, true, false><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
    else if (ntl) k_collide_batch<P, false, true><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
    else k_collide_batch<P, false, false><<<nblocks, b, 0, s>>>(d, map, wpb, cur, own, slot, box_rows_cached);
  }));
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

}  // namespace lbmib
