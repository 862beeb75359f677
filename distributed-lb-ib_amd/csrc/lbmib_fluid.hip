// lbmib_fluid.hip -- fused LBM kernels for gfx950.
//
// Layout in HBM: structure-of-arrays, one plane per population q, x fastest:
//   f[q * pitch + x + nx * (y + ny * z)]
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
//   F32       g = f - w_q in FP32; collision in FP32 on deviations, which
//             keeps the small velocity (|u| ~ 1e-2) out of the rounding of
//             the O(0.3) rest population.
//   F16/BF16  g * 2^8 in 16 bits; collision in FP32.
#include <hip/hip_runtime.h>

#include "lbmib_internal.h"

namespace lbmib {

typedef _Float16 half_t;

template <int P>
struct Traits;
template <>
struct Traits<LBMIB_F64> {
  using S = double;
  using T = double;
  using FT = double;
  static constexpr bool dev = false;
};
template <>
struct Traits<LBMIB_F32> {
  using S = float;
  using T = float;
  using FT = float;
  static constexpr bool dev = true;
};
template <>
struct Traits<LBMIB_F16> {
  using S = half_t;
  using T = float;
  using FT = float;
  static constexpr bool dev = true;
};
template <>
struct Traits<LBMIB_BF16> {
  using S = uint16_t;
  using T = float;
  using FT = float;
  static constexpr bool dev = true;
};

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (values are finite)
  return (uint16_t)(u >> 16);
}

template <int P>
__device__ __forceinline__ typename Traits<P>::T ld(const typename Traits<P>::S *p) {
  if constexpr (P == LBMIB_F16) return (float)(*p) * (1.0f / kDevScale16);
  else if constexpr (P == LBMIB_BF16) return bf16_to_f32(*p) * (1.0f / kDevScale16);
  else return *p;
}
template <int P>
__device__ __forceinline__ void st(typename Traits<P>::S *p, typename Traits<P>::T v) {
  if constexpr (P == LBMIB_F16) *p = (half_t)(v * (float)kDevScale16);
  else if constexpr (P == LBMIB_BF16) *p = f32_to_bf16(v * (float)kDevScale16);
  else *p = v;
}

// Host double constants -> kernel compute type.
template <class T>
static FluidConsts<T> cast_consts(const FluidConsts<double> &d) {
  FluidConsts<T> c;
  c.omc = (T)d.omc;
  c.om = (T)d.om;
  for (int q = 0; q < 19; ++q) {
    c.pref[q] = (T)d.pref[q];
    c.pw[q] = (T)d.pw[q];
    c.tab_in[q] = (T)d.tab_in[q];
    c.tab_out[q] = (T)d.tab_out[q];
  }
  c.cs2 = (T)d.cs2;
  c.cs4 = (T)d.cs4;
  c.g = (T)d.g;
  c.rho0 = (T)d.rho0;
  c.u0x = (T)d.u0x;
  c.u0y = (T)d.u0y;
  c.u0z = (T)d.u0z;
  return c;
}

__device__ __forceinline__ bool in_box6(const int *b, int x, int y, int z) {
  return x >= b[0] && x <= b[1] && y >= b[2] && y <= b[3] && z >= b[4] && z <= b[5];
}

__device__ __forceinline__ bool in_any_box(const int *boxes, int nsheets, int x, int y, int z) {
  bool in = false;
  for (int s = 0; s < nsheets; ++s) in |= in_box6(boxes + 6 * s, x, y, z);
  return in;
}

// Pulls the 19 populations arriving at (x,y,z) from the post-collision buffer.
// Wall axes: a source outside the active box is replaced by the opposite
// population of the same node (half-way bounce-back, lbm_ib.c:2310-2423).
// Periodic axes wrap.  Interior cells take the branch-free path.
template <int P>
__device__ __forceinline__ void pull(const Geometry &g, const typename Traits<P>::S *__restrict__ f, int x, int y,
                                     int z, long c, typename Traits<P>::T v[19]) {
  const bool interior = x > g.lo[0] && x < g.hi[0] && y > g.lo[1] && y < g.hi[1] && z > g.lo[2] && z < g.hi[2];
  const long sx = 1, sy = g.nx, sz = (long)g.nx * g.ny;
  if (interior) {
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const long off = kCX[q] * sx + kCY[q] * sy + kCZ[q] * sz;
      v[q] = ld<P>(f + q * g.pitch + c - off);
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    int px = x - kCX[q], py = y - kCY[q], pz = z - kCZ[q];
    bool bb = false;
    if (g.bc[0] == LBMIB_BC_PERIODIC) px = px < 0 ? px + g.nx : (px >= g.nx ? px - g.nx : px);
    else bb |= (px < g.lo[0] || px > g.hi[0]);
    if (g.bc[1] == LBMIB_BC_PERIODIC) py = py < 0 ? py + g.ny : (py >= g.ny ? py - g.ny : py);
    else bb |= (py < g.lo[1] || py > g.hi[1]);
    if (g.bc[2] == LBMIB_BC_PERIODIC) pz = pz < 0 ? pz + g.nz : (pz >= g.nz ? pz - g.nz : pz);
    else bb |= (pz < g.lo[2] || pz > g.hi[2]);
    const long src = bb ? c + (long)kOPP[q] * g.pitch - (long)q * g.pitch
                        : (long)px + sy * py + sz * pz;
    v[q] = ld<P>(f + q * g.pitch + src);
  }
}

// Equilibrium + BGK + Guo forcing in the reference's operation order
// (lbm_ib.c:1731-1818).  Used by the F64 form.
__device__ __forceinline__ double collide_exact(int q, double fq, double rho, double vx, double vy, double vz,
                                                double fx, double fy, double fz, const FluidConsts<double> &c) {
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
  return fq * c.omc + c.om * eq + c.pref[q] * (a / c.cs2 + cu / c.cs4 * cf);
}

template <int P>
__global__ void __launch_bounds__(1024)
    k_collide_stream(Geometry g, const typename Traits<P>::S *__restrict__ fin, typename Traits<P>::S *__restrict__ fout,
                     int own, FluidConsts<typename Traits<P>::T> c, int nsheets, const int *__restrict__ box_cur,
                     const int *__restrict__ box_prev, const typename Traits<P>::FT *__restrict__ Fcur,
                     const typename Traits<P>::FT *__restrict__ Fprev) {
  using T = typename Traits<P>::T;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x < g.lo[0] || x > g.hi[0] || y < g.lo[1] || y > g.hi[1] || z < g.lo[2] || z > g.hi[2]) return;
  const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);

  T v[19];
  T rho, ux, uy, uz;
  const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
  if (frozen) {
    // planes x = 2 and N-3 hold the inlet/outlet tables after
    // copy_inout_to_df2, and their rho/vel are never updated (the moment box
    // starts at x = 3, lbm_ib.c:2482-2483).
    const T *tab = (x == g.lo[0]) ? c.tab_in : c.tab_out;
#pragma unroll
    for (int q = 0; q < 19; ++q) v[q] = tab[q];
    rho = c.rho0; ux = c.u0x; uy = c.u0y; uz = c.u0z;
  } else {
    if (own) {
#pragma unroll
      for (int q = 0; q < 19; ++q) v[q] = ld<P>(fin + q * g.pitch + c0);
      rho = c.rho0; ux = c.u0x; uy = c.u0y; uz = c.u0z;  // init_gv values, lbm_ib.c:837-849
    } else {
      pull<P>(g, fin, x, y, z, c0, v);
      T fpx = 0, fpy = 0, fpz = 0;
      if (in_any_box(box_prev, nsheets, x, y, z)) {
        fpx = (T)Fprev[c0]; fpy = (T)Fprev[g.n + c0]; fpz = (T)Fprev[2 * g.n + c0];
      }
      if constexpr (!Traits<P>::dev) {
        // rho = Sum df2, u = (Sum c df2 + dt/2 F)/rho (lbm_ib.c:2498-2525)
        double s1 = v[0], s2 = 0.0, s3 = 0.0, s4 = 0.0;
#pragma unroll
        for (int q = 1; q < 19; ++q) {
          s1 += v[q];
          if (kCX[q]) s2 = kCX[q] > 0 ? s2 + v[q] : s2 - v[q];
          if (kCY[q]) s3 = kCY[q] > 0 ? s3 + v[q] : s3 - v[q];
          if (kCZ[q]) s4 = kCZ[q] > 0 ? s4 + v[q] : s4 - v[q];
        }
        rho = s1;
        ux = (s2 + 0.5 * fpx) / s1;
        uy = (s3 + 0.5 * fpy) / s1;
        uz = (s4 + 0.5 * (fpz + rho * c.g)) / s1;
      } else {
        T d = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
        for (int q = 0; q < 19; ++q) {
          d += v[q];
          if (kCX[q]) mx = kCX[q] > 0 ? mx + v[q] : mx - v[q];
          if (kCY[q]) my = kCY[q] > 0 ? my + v[q] : my - v[q];
          if (kCZ[q]) mz = kCZ[q] > 0 ? mz + v[q] : mz - v[q];
        }
        rho = (T)1 + d;
        const T ir = (T)1 / rho;
        ux = (mx + (T)0.5 * fpx) * ir;
        uy = (my + (T)0.5 * fpy) * ir;
        uz = (mz + (T)0.5 * (fpz + rho * c.g)) * ir;
      }
    }
  }

  T fx = 0, fy = 0, fz = 0;
  if (in_any_box(box_cur, nsheets, x, y, z)) {
    fx = (T)Fcur[c0]; fy = (T)Fcur[g.n + c0]; fz = (T)Fcur[2 * g.n + c0];
  }

  if constexpr (!Traits<P>::dev) {
#pragma unroll
    for (int q = 0; q < 19; ++q)
      st<P>(fout + q * g.pitch + c0, collide_exact(q, v[q], rho, ux, uy, uz, fx, fy, fz, c));
  } else {
    // Deviation form: with g_q = f_q - w_q and rho = 1 + Sum g,
    //   feq_q - w_q = w_q (drho + rho (3 cu + 4.5 cu^2 - 1.5 u^2))
    //   g*_q = (1 - 1/tau) g_q + (1/tau)(feq_q - w_q) + S_q,
    // S_q = w_q (1 - 1/(2 tau)) (3 (c_q - u).F + 9 (c_q.u)(c_q.F)).
    const T drho = rho - (T)1;
    const T usq15 = (T)1.5 * (ux * ux + uy * uy + uz * uz);
    const T fzg = fz + rho * c.g;
    const T uF = ux * fx + uy * fy + uz * fzg;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const T cu = (T)kCX[q] * ux + (T)kCY[q] * uy + (T)kCZ[q] * uz;
      const T cF = (T)kCX[q] * fx + (T)kCY[q] * fy + (T)kCZ[q] * fzg;
      const T w = (T)1 / (T)kWDEN[q];
      const T geq = w * (drho + rho * ((T)3 * cu + (T)4.5 * cu * cu - usq15));
      const T S = c.pw[q] * ((T)3 * (cF - uF) + (T)9 * cu * cF);
      st<P>(fout + q * g.pitch + c0, c.omc * v[q] + c.om * geq + S);
    }
  }
}

// Fills the active box with the equilibrium (init_eqlbrmdistrfuncDF0 +
// init_df1, lbm_ib.c:1528-1606, 2725-2767); other cells are zero.
template <int P>
__global__ void k_fill_init(Geometry g, typename Traits<P>::S *__restrict__ f, FluidConsts<typename Traits<P>::T> c) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= g.nx) return;
  const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
  const bool act = x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2];
#pragma unroll
  for (int q = 0; q < 19; ++q) st<P>(f + q * g.pitch + c0, act ? c.tab_in[q] : (typename Traits<P>::T)0);
}

// Materialises df2 (= df1 after the swap), rho and vel of the last completed
// step for every cell, in FP64 and in the reference's conventions.
template <int P>
__global__ void k_materialize(Geometry g, const typename Traits<P>::S *__restrict__ f, int own,
                              FluidConsts<typename Traits<P>::T> c, int nsheets, const int *__restrict__ box,
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
      const T *tab = (x == g.lo[0]) ? c.tab_in : c.tab_out;
      for (int q = 0; q < 19; ++q) v[q] = tab[q];
    } else if (own) {
      for (int q = 0; q < 19; ++q) v[q] = ld<P>(f + q * g.pitch + c0);
    } else {
      pull<P>(g, f, x, y, z, c0, v);
    }
    for (int q = 0; q < 19; ++q) fv[q] = Traits<P>::dev ? (double)v[q] + 1.0 / kWDEN[q] : (double)v[q];
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
                     FluidConsts<typename Traits<P>::T> c, double *__restrict__ partials) {
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
      const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
      T v[19];
      const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
      if (frozen) {
        const T *tab = (x == g.lo[0]) ? c.tab_in : c.tab_out;
        for (int q = 0; q < 19; ++q) v[q] = tab[q];
      } else if (own) {
        for (int q = 0; q < 19; ++q) v[q] = ld<P>(f + q * g.pitch + c0);
      } else {
        pull<P>(g, f, x, y, z, c0, v);
      }
      for (int q = 0; q < 19; ++q) {
        const double fq = Traits<P>::dev ? (double)v[q] + 1.0 / kWDEN[q] : (double)v[q];
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
    using TT = typename Traits<P>::T;
    k_fill_init<P><<<gr, b, 0, s>>>(g, (typename Traits<P>::S *)f, cast_consts<TT>(c));
  }));
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int launch_collide_stream(int prec, const Geometry &g, const void *fin, void *fout, int own_mode,
                          const FluidConsts<double> &c, int nsheets, const int *box_cur, const int *box_prev,
                          const void *Fcur, const void *Fprev, hipStream_t s) {
  dim3 b = row_block(g), gr = row_grid(g, b);
  LBMIB_DISPATCH(prec, ({
    using TT = typename Traits<P>::T;
    using FT = typename Traits<P>::FT;
    k_collide_stream<P><<<gr, b, 0, s>>>(g, (const typename Traits<P>::S *)fin, (typename Traits<P>::S *)fout,
                                          own_mode, cast_consts<TT>(c), nsheets, box_cur, box_prev,
                                          (const FT *)Fcur, (const FT *)Fprev);
  }));
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

int launch_materialize(int prec, const Geometry &g, const void *f, int own_mode, const FluidConsts<double> &c,
                       int nsheets, const int *box, const void *F, double *df, double *rho, double *vel,
                       hipStream_t s) {
  dim3 b = row_block(g), gr = row_grid(g, b);
  LBMIB_DISPATCH(prec, ({
    using TT = typename Traits<P>::T;
    using FT = typename Traits<P>::FT;
    k_materialize<P><<<gr, b, 0, s>>>(g, (const typename Traits<P>::S *)f, own_mode, cast_consts<TT>(c), nsheets,
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
    using TT = typename Traits<P>::T;
    k_totals_partial<P><<<nb, kTotThreads, 0, s>>>(g, (const typename Traits<P>::S *)f, own_mode,
                                                    cast_consts<TT>(c), partials);
  }));
  k_totals_final<<<1, kTotThreads, 0, s>>>(partials, nb, out4);
  return hipGetLastError() == hipSuccess ? LBMIB_OK : LBMIB_ERR_HIP;
}

}  // namespace lbmib
