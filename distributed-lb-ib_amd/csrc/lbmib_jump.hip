// lbmib_jump.hip -- coupling of the learned surrogate to a running simulation
// (the lbmib_coarse_* and lbmib_jump_* entry points of include/lbmib.h).
//
// Not in the reference: cube_pthread/lbm_ib.c steps every time step
// (do_thread, lbm_ib.c:3115-3563).  Here a caller can read the coarse-grained
// state (the rho/u of compute_rho_and_u, lbm_ib.c:2440-2538, averaged over
// f^3 blocks, and a raster of the sheets), replace K steps by a predicted
// change of it, and undo that jump when the plain solver's re-verification
// rejects it.
//
// A jump keeps the populations' non-equilibrium part and shifts each cell's
// equilibrium to the predicted moments: h' = h + s (feq(rho + drho, u + du)
// - feq(rho, u)), with (drho, du) interpolated trilinearly from the coarse
// prediction; the sheets move by K times the mean of the coarse velocity
// before and after the jump at each node.  The checkpoint costs no copy of
// the populations: the jump writes into a third buffer and the pre-jump
// buffer is set aside until the jump is committed or reverted.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <type_traits>

#include "lbmib_coarse.h"
#include "lbmib_ctx.h"
#include "lbmib_device.h"

namespace lbmib {
namespace {

#define JHIPCHK(x)                               \
  do {                                           \
    if ((x) != hipSuccess) return LBMIB_ERR_HIP; \
  } while (0)
#define JRCHK(x)        \
  do {                  \
    int _r = (x);       \
    if (_r) return _r;  \
  } while (0)

// A stored population as stored (no conversion), buffer-addressed: the
// per-lane byte offset voff and the wave-uniform soff, both below 2^32.
template <int P>
__device__ __forceinline__ typename Traits<P>::S bld_raw(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  using S = typename Traits<P>::S;
  if constexpr (P == LBMIB_F64) return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else if constexpr (P == LBMIB_F32) return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else return __builtin_bit_cast(S, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
}
template <int P>
__device__ __forceinline__ void bst_raw(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                        typename Traits<P>::S s) {
  if constexpr (P == LBMIB_F64) {
    typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, s), r, voff, soff, 0);
  } else if constexpr (P == LBMIB_F32) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, s), r, voff, soff, 0);
  } else {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, s), r, voff, soff, 0);
  }
}

// Per fine cell, the moments of lbmib_get_macroscopic (post-stream
// populations, u = (sum c f + F/2) / rho, the initial rho0/u0 outside the
// moment box), block-averaged into `out`.
template <int P, int F, bool BUF>
__global__ void __launch_bounds__(256)
    k_coarse_moments(Geometry g, const typename Traits<P>::S *__restrict__ f, int own,
                     const typename Traits<P>::T *__restrict__ tab, double rhor, double mrx,
                     typename Traits<P>::T b1, typename Traits<P>::T b2, double rho0, double u0x, double u0y,
                     double u0z, double gz, int nsheets, const int *__restrict__ box,
                     const typename Traits<P>::FT *__restrict__ Fc, float *__restrict__ out) {
  using S = typename Traits<P>::S;
  using A = JumpT<P>;
  __shared__ double red[4][4][64];
  const Tile tl = tile_of(g, F);
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x = tl.x * 64 + tx, y = tl.y * 4 + ty;
  const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
  const bool row_act = x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1];
  const bool row_mom = x >= g.mlo[0] && x <= g.mhi[0] && y >= g.mlo[1] && y <= g.mhi[1];
  const bool row_int = x > g.lo[0] && x < g.hi[0] && y > g.lo[1] && y < g.hi[1];
  // which cells pull here: interior cells load their pulled populations up
  // front; the others pull at their turn through the boundary rules
  auto act_at = [&](int z) { return row_act && z >= g.lo[2] && z <= g.hi[2]; };
  auto pulls = [&](int z) { return act_at(z) && row_mom && z >= g.mlo[2] && z <= g.mhi[2] && !own && !frozen; };
  auto fast = [&](int z) { return row_int && z > g.lo[2] && z < g.hi[2]; };
  S pv[19];
  // (BUF: the whole population array lies within 2^32 bytes, so the loads
  // take a 32-bit lane offset and a uniform one instead of 64-bit addresses)
  const __amdgpu_buffer_rsrc_t rf = make_rsrc(f);
  auto fetch = [&](int z, S(&pp)[19]) {
    const long c = pop_index(g, x, y, z);
    const long sy = g.prow, sz = g.prow * g.ny;
    if (pulls(z) && fast(z)) {
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        // the source element: the neighbour's population q, inside the array
        // for an interior cell (the lane part of the offset carries the
        // neighbour's position, which may precede this cell's)
        const long src = c - (kCX[q] + kCY[q] * sy + kCZ[q] * sz);
        if constexpr (BUF) pp[q] = bld_raw<P>(rf, (uint32_t)(src * sizeof(S)), (uint32_t)(q * g.pq * sizeof(S)));
        else pp[q] = f[q * g.pq + src];
      }
    }
  };
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
  for (int k = 0; k < F; ++k) {
    const int z = tl.z * F + k;
    fetch(z, pv);
    A r = (A)rho0, ux = (A)u0x, uy = (A)u0y, uz = (A)u0z;
    if (act_at(z) && row_mom && z >= g.mlo[2] && z <= g.mhi[2] && !own) {
      const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
      A fx = 0, fy = 0, fz = 0;
      if (in_any_box(box, nsheets, x, y, z)) { fx = (A)Fc[c0]; fy = (A)Fc[g.n + c0]; fz = (A)Fc[2 * g.n + c0]; }
      A s1 = 0, s2 = 0, s3 = 0, s4 = 0;
      auto add = [&](int q, A fq) {  // fq: f, or the deviation f - r
        s1 += fq;
        s2 = add_c(s2, kCX[q], fq);
        s3 = add_c(s3, kCY[q], fq);
        s4 = add_c(s4, kCZ[q], fq);
      };
      if (frozen) {
#pragma unroll
        for (int q = 0; q < 19; ++q) add(q, (A)(tab[q] * Traits<P>::inv_scale));
      } else if (fast(z)) {
#pragma unroll
        for (int q = 0; q < 19; ++q) add(q, (A)(from_bits<P>(pv[q]) * Traits<P>::inv_scale));
      } else {  // boundary cells: one population at a time, in the same order
        const long c = pop_index(g, x, y, z);
#pragma unroll 1
        for (int q = 0; q < 19; ++q) add(q, (A)pull_one<P>(g, f, x, y, z, c, q, b1, b2));
      }
      if (Traits<P>::dev) {  // sum r = rho_r, sum c r = rho_r u_r e_x
        s1 += (A)rhor;
        s2 += (A)mrx;
      }
      r = s1;
      ux = (s2 + (A)0.5 * fx) / s1;
      uy = (s3 + (A)0.5 * fy) / s1;
      uz = (s4 + (A)0.5 * (fz + r * (A)gz)) / s1;
    }
    acc[0] += (double)r;
    acc[1] += (double)ux;
    acc[2] += (double)uy;
    acc[3] += (double)uz;
  }
  block_mean_store<F>(acc, red, g, tl, out);
}

// The stored moments of every cell -- rho and u = sum c f / rho of its own
// populations on the updated cells (the active box without the frozen
// inlet/outlet planes), rho0/u0 elsewhere; the convention lbmib_jump reports
// a jumped state in -- as per-row partial sums (lbmib_coarse.h): one block
// per (y, z) row, blockDim.x a multiple of 64 dividing into nx in x steps.
template <int P, int F, bool BUF>
__global__ void __launch_bounds__(1024)
    k_coarse_stored_rows(Geometry g, const typename Traits<P>::S *__restrict__ f, double rhor, double mrx,
                         double rho0, double u0x, double u0y, double u0z, float *__restrict__ part) {
  using S = typename Traits<P>::S;
  using A = JumpT<P>;
  const __amdgpu_buffer_rsrc_t rf = make_rsrc(f);
  const long nrows = (long)g.ny * g.nz, ncx = g.nx / F;
  for (long row = blockIdx.x; row < nrows; row += gridDim.x) {
    const int y = (int)(row % g.ny), z = (int)(row / g.ny);
    const bool live = y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2];
    for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
      const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
      A sr = (A)rho0, sx = (A)u0x, sy = (A)u0y, sz = (A)u0z;
      if (live && x >= g.lo[0] && x <= g.hi[0] && !frozen) {
        const long c = pop_index(g, x, y, z);
        S h[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) {
          if constexpr (BUF) h[q] = bld_raw<P>(rf, (uint32_t)(c * sizeof(S)), (uint32_t)(q * g.pq * sizeof(S)));
          else h[q] = f[q * g.pq + c];
        }
        stored_moments<P>(h, rhor, mrx, sr, sx, sy, sz);
      }
      float m[4] = {(float)sr, (float)sx, (float)sy, (float)sz};
      row_partial<F>(m, x, row, ncx, part);
    }
  }
}

// trilinear interpolation of coarse plane `a` at fine position (px, py, pz)
__device__ __forceinline__ double interp(const float *__restrict__ cf, int a, long nc, int ncx, int ncy, int ncz,
                                         int F, double px, double py, double pz) {
  const double h = 0.5 * (F - 1);
  const double cx = (px - h) / F, cy = (py - h) / F, cz = (pz - h) / F;
  const double fx = floor(cx), fy = floor(cy), fz = floor(cz);
  const double wx = cx - fx, wy = cy - fy, wz = cz - fz;
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
    const int jx = min(max(ix + dx, 0), ncx - 1), jy = min(max(iy + dy, 0), ncy - 1), jz = min(max(iz + dz, 0), ncz - 1);
    const double w = (dx ? wx : 1.0 - wx) * (dy ? wy : 1.0 - wy) * (dz ? wz : 1.0 - wz);
    s += w * (double)cf[a * nc + (long)jx + (long)ncx * ((long)jy + (long)ncy * jz)];
  }
  return s;
}

// feq_q(rho', u') - feq_q(rho, u) of init_eqlbrmdistrfuncDF0's expression
// (lbm_ib.c:1561-1594) in difference form, u' = u + du, rho' = rho + drho:
//   w [drho (1 + 3 cu + 4.5 cu^2 - 1.5 u^2)
//      + rho' (3 c.du + 4.5 c.du (cu + cu') - 1.5 du.(u + u'))]
template <class T>
__device__ __forceinline__ T dfeq_q(int q, T rho, T drho, T ux, T uy, T uz, T dx, T dy, T dz) {
  const T cu = add_c(add_c(add_c((T)0, kCX[q], ux), kCY[q], uy), kCZ[q], uz);  // c.u
  const T cd = add_c(add_c(add_c((T)0, kCX[q], dx), kCY[q], dy), kCZ[q], dz);  // c.du
  const T usq = ux * ux + uy * uy + uz * uz;
  const T ud = dx * (ux + ux + dx) + dy * (uy + uy + dy) + dz * (uz + uz + dz);  // du.(u + u')
  const T a = (T)1 + (T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * usq;
  const T b = (T)3 * cd + (T)4.5 * cd * (cu + cu + cd) - (T)1.5 * ud;
  return (drho * a + (rho + drho) * b) / (T)kWDEN[q];
}

// One block per (y, z) row, in x chunks of blockDim.x cells (a multiple of
// 64 that divides nx, at most 512): the updated cells -- the active box
// without the frozen inlet/outlet planes -- take the equilibrium shift;
// every other element is copied unchanged.  With `part`, the jumped state's
// stored moments (rho + drho, u + du exactly, by the moments of feq; rho0/u0
// on the other cells) go to per-row partial sums (lbmib_coarse.h), the
// stored planes' convention.
template <int P, int F, bool BUF>
__global__ void __launch_bounds__(512)
    k_jump_apply_rows(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                      typename Traits<P>::S *__restrict__ fout, double rhor, double mrx,
                      const float *__restrict__ delta, double rho0, double u0x, double u0y, double u0z,
                      float *__restrict__ part) {
  using T = JumpT<P>;
  using S = typename Traits<P>::S;
  // the coarse change along the chunk: its coarse cells and one more on each
  // side in x, the two coarse rows around the row in y and in z (clamped at
  // the grid's edges), 4 planes
  constexpr int kSX = 512 / F + 2;
  __shared__ float sd[4][2][2][kSX];
  const int SX = (int)blockDim.x / F + 2;
  const int ncx = g.nx / F, ncy = g.ny / F, ncz = g.nz / F;
  const long nc = (long)ncx * ncy * ncz, nrows = (long)g.ny * g.nz;
  // (BUF: both population arrays lie within 2^32 bytes; loads and stores
  // take a 32-bit lane offset and a uniform one)
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(fin), rout = make_rsrc(fout);
  auto load = [&](long p, int q) -> S {
    if constexpr (BUF) return bld_raw<P>(rin, (uint32_t)(p * sizeof(S)), (uint32_t)(q * g.pq * sizeof(S)));
    else return fin[q * g.pq + p];
  };
  auto store = [&](long p, int q, S v) {
    if constexpr (BUF) bst_raw<P>(rout, (uint32_t)(p * sizeof(S)), (uint32_t)(q * g.pq * sizeof(S)), v);
    else fout[q * g.pq + p] = v;
  };
  const double hc = 0.5 * (F - 1);
  for (long row = blockIdx.x; row < nrows; row += gridDim.x) {
    const int y = (int)(row % g.ny), z = (int)(row / g.ny);
    const bool live = y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2];
    const double fy = floor((y - hc) / F), fz = floor((z - hc) / F);
    const double wy = (y - hc) / F - fy, wz = (z - hc) / F - fz;
    // the y-z factors of the trilinear weights (exact in FP32)
    const float wyz[2][2] = {{(float)((1.0 - wy) * (1.0 - wz)), (float)(wy * (1.0 - wz))},
                             {(float)((1.0 - wy) * wz), (float)(wy * wz)}};
    for (int xb = 0; xb < g.nx; xb += blockDim.x) {
      const int cx0 = xb / F - 1;
      __syncthreads();  // the previous chunk's readers are done
      for (int e = threadIdx.x; e < 16 * SX; e += blockDim.x) {
        const int i = e % SX, dy = (e / SX) & 1, dz = (e / (2 * SX)) & 1, a = e / (4 * SX);
        const int jx = min(max(cx0 + i, 0), ncx - 1), jy = min(max((int)fy + dy, 0), ncy - 1);
        const int jz = min(max((int)fz + dz, 0), ncz - 1);
        sd[a][dz][dy][i] = delta[a * nc + (long)jx + (long)ncx * ((long)jy + (long)ncy * jz)];
      }
      __syncthreads();
      const int x = xb + (int)threadIdx.x;
      const long p0 = pop_index(g, x, y, z);
      S h[19];
#pragma unroll
      for (int q = 0; q < 19; ++q) h[q] = load(p0, q);
      const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
      float m[4] = {(float)(T)rho0, (float)(T)u0x, (float)(T)u0y, (float)(T)u0z};
      if (!live || x < g.lo[0] || x > g.hi[0] || frozen) {
#pragma unroll
        for (int q = 0; q < 19; ++q) store(p0, q, h[q]);  // unchanged
      } else {
        T rho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int q = 0; q < 19; ++q) {
          const T fq = (T)from_bits<P>(h[q]) * (T)Traits<P>::inv_scale;  // f, or the deviation f - r
          rho += fq;
          jx = add_c(jx, kCX[q], fq);
          jy = add_c(jy, kCY[q], fq);
          jz = add_c(jz, kCZ[q], fq);
        }
        if (Traits<P>::dev) {  // sum r = rho_r, sum c r = rho_r u_r e_x
          rho += (T)rhor;
          jx += (T)mrx;
        }
        const T ux = jx / rho, uy = jy / rho, uz = jz / rho;
        // trilinear weights once for the four planes (the clamping of
        // interp(): a cell left of the first coarse centre takes the edge
        // value, which the clamped staging reproduces)
        T d0, d1, d2, d3;
        if constexpr (P == LBMIB_F64) {
          const double fx = floor((x - hc) / F);
          const double wx = (x - hc) / F - fx;
          const int ix = (int)fx - cx0;
          double dd[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const double w = (dx ? wx : 1.0 - wx) * (dy ? wy : 1.0 - wy) * (dz ? wz : 1.0 - wz);
#pragma unroll
            for (int a = 0; a < 4; ++a) dd[a] += w * (double)sd[a][dz][dy][ix + dx];
          }
          d0 = dd[0], d1 = dd[1], d2 = dd[2], d3 = dd[3];
#pragma unroll
          for (int q = 0; q < 19; ++q) {
            const T d = dfeq_q<T>(q, rho, d0, ux, uy, uz, d1, d2, d3);
            store(p0, q, to_bits<P>(from_bits<P>(h[q]) + (typename Traits<P>::T)(d * (T)Traits<P>::scale)));
          }
        } else {
          // FP32 throughout: the weights are multiples of 1 / (2F) with a
          // few significant bits, exact in FP32, as are their products
          const float v = ((float)x - (float)hc) * (1.0f / F), fx = floorf(v), wx = v - fx;
          const int ix = (int)fx - cx0;
          float dd[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int c = 0; c < 8; ++c) {
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const float w = (dx ? wx : 1.0f - wx) * wyz[dz][dy];
#pragma unroll
            for (int a = 0; a < 4; ++a) dd[a] = fmaf(w, sd[a][dz][dy][ix + dx], dd[a]);
          }
          d0 = dd[0], d1 = dd[1], d2 = dd[2], d3 = dd[3];
          // the shift s w_q [drho a_q + rho' b_q] with the per-cell terms
          // taken out:
          //   drho a_q + rho' b_q = A0 + 3 p_q + 4.5 r_q,
          //   A0 = drho (1 - 1.5 u^2) - 1.5 rho' du.(u + u'),
          //   p_q = drho cu + rho' cd,  r_q = drho cu^2 + rho' cd (2 cu + cd);
          // the opposite direction (c -> -c) has the same r and -p, so each
          // pair (2k - 1, 2k) is evaluated once
          const float rp = rho + d0;
          const float usq = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
          const float ud = fmaf(d1, ux + ux + d1, fmaf(d2, uy + uy + d2, d3 * (uz + uz + d3)));
          const float A0 = fmaf(-1.5f, rp * ud, d0 * fmaf(-1.5f, usq, 1.0f));
          constexpr float sc = Traits<P>::scale;
          const float wA[3] = {A0 * (sc / 3.0f), A0 * (sc / 18.0f), A0 * (sc / 36.0f)};
          store(p0, 0, to_bits<P>((typename Traits<P>::T)((float)from_bits<P>(h[0]) + wA[0])));
#pragma unroll
          for (int q = 1; q < 19; q += 2) {
            const float cu = add_c(add_c(add_c(0.0f, kCX[q], ux), kCY[q], uy), kCZ[q], uz);
            const float cd = add_c(add_c(add_c(0.0f, kCX[q], d1), kCY[q], d2), kCZ[q], d3);
            const float dcu = d0 * cu;
            const float pq = fmaf(rp, cd, dcu);
            const float r = fmaf(rp * cd, fmaf(2.0f, cu, cd), dcu * cu);
            const float wq = sc / (float)kWDEN[q];  // folded: q is unrolled
            const float base = fmaf(4.5f * wq, r, wA[kCLS[q]]);
            const float s1 = fmaf(3.0f * wq, pq, base), s2 = fmaf(-3.0f * wq, pq, base);
            store(p0, q, to_bits<P>((typename Traits<P>::T)((float)from_bits<P>(h[q]) + s1)));
            store(p0, q + 1, to_bits<P>((typename Traits<P>::T)((float)from_bits<P>(h[q + 1]) + s2)));
          }
        }
        m[0] = (float)(rho + d0);
        m[1] = (float)(ux + d1);
        m[2] = (float)(uy + d2);
        m[3] = (float)(uz + d3);
      }
      if (part) row_partial<F>(m, x, row, ncx, part);
    }
  }
}

// X += K (u(X) + u'(X)) / 2 with u the coarse velocity before the jump and
// u' = u + du after it.
__global__ void k_jump_fibers(double *__restrict__ pos, int M, const float *__restrict__ coarse,
                              const float *__restrict__ delta, int F, int ncx, int ncy, int ncz, double K) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const long nc = (long)ncx * ncy * ncz;
  const double px = pos[i], py = pos[M + i], pz = pos[2 * (long)M + i];
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double u0 = interp(coarse, 1 + a, nc, ncx, ncy, ncz, F, px, py, pz);
    const double du = interp(delta, 1 + a, nc, ncx, ncy, ncz, F, px, py, pz);
    d[a] = K * (u0 + 0.5 * du);
  }
  pos[i] = px + d[0];
  pos[M + i] = py + d[1];
  pos[2 * (long)M + i] = pz + d[2];
}

// copies the F slot values inside the support boxes `box` (device) from src to dst
template <class FT>
__global__ void k_copy_boxes(const FT *__restrict__ src, FT *__restrict__ dst, const int *__restrict__ box, int nsheets,
                             Geometry g) {
  for (int s = 0; s < nsheets; ++s) {
    const int *b = box + 6 * s;
    const int x0 = ldc(b), x1 = ldc(b + 1), y0 = ldc(b + 2), y1 = ldc(b + 3), z0 = ldc(b + 4), z1 = ldc(b + 5);
    if (x0 > x1 || y0 > y1 || z0 > z1) continue;
    const long bx = x1 - x0 + 1, by = y1 - y0 + 1, bz = z1 - z0 + 1;
    const long cells = bx * by * bz;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < cells; e += (long)gridDim.x * blockDim.x) {
      const int x = x0 + (int)(e % bx), y = y0 + (int)((e / bx) % by), z = z0 + (int)(e / (bx * by));
      const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
      for (int a = 0; a < 3; ++a) dst[a * g.n + c0] = src[a * g.n + c0];
    }
  }
}

// ------------------------------------------------------------- sheet raster
//
// Per coarse cell: node count and the sums of the nodes' elastic forces,
// accumulated in 64-bit fixed point (2^40 units per unit force), so that the
// order in which the atomics land does not change the result.
constexpr double kRasterScale = 1099511627776.0;  // 2^40

__global__ void k_raster_zero(unsigned long long *__restrict__ acc, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) acc[e] = 0ull;
}

__global__ void k_raster_nodes(const double *__restrict__ pos, const double *__restrict__ frc, int M, int F, int ncx,
                               int ncy, int ncz, unsigned long long *__restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const long nc = (long)ncx * ncy * ncz;
  const int cx = min(max((int)floor((pos[i] + 0.5) / F), 0), ncx - 1);
  const int cy = min(max((int)floor((pos[M + i] + 0.5) / F), 0), ncy - 1);
  const int cz = min(max((int)floor((pos[2 * (long)M + i] + 0.5) / F), 0), ncz - 1);
  const long cc = (long)cx + (long)ncx * ((long)cy + (long)ncy * cz);
  atomicAdd(acc + cc, 1ull);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const long long v = llrint(frc[(NF_EX + a) * (long)M + i] * kRasterScale);
    atomicAdd(acc + (1 + a) * nc + cc, (unsigned long long)v);
  }
}

__global__ void k_raster_final(const unsigned long long *__restrict__ acc, float *__restrict__ out, long nc) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < 4 * nc; e += (long)gridDim.x * blockDim.x) {
    const long long v = (long long)acc[e];
    out[e] = e < nc ? (float)v : (float)((double)v / kRasterScale);
  }
}

unsigned grid_of(long n, long cap = 16384) {
  long b = (n + 255) / 256;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Threads of a row kernel's block: the largest multiple of 64 up to cap that
// divides nx (nx is a multiple of 64), so that every chunk of a row is whole.
inline int row_threads(const Geometry &g, int cap) {
  int t = cap;
  while (g.nx % t) t -= 64;
  return t;
}

template <int P>
int coarse_launch(lbmib_ctx *ctx, int F, float *out, float *stored) {
  using T = typename Traits<P>::T;
  const Geometry &g = ctx->g;
  const double s = Traits<P>::scale;
  const T b1 = Traits<P>::dev ? (T)host_bounce_corr(ctx->c, 1, s) : (T)0;
  const T b2 = Traits<P>::dev ? (T)host_bounce_corr(ctx->c, 2, s) : (T)0;
  const int sl = ctx->spread_slot;
  const bool force = sl >= 0 && ctx->nnodes > 0;
  const int *box = ctx->ib.box + (force ? sl : 0) * kMaxSheets * 6;
  const typename Traits<P>::FT *Fc = force ? (const typename Traits<P>::FT *)ctx->F[sl] : nullptr;
  const int nsh = force ? ctx->nsheets : 0;
  const int grid = (g.nx / 64) * (g.ny / 4) * (g.nz / F);  // tiles, dealt by tile_of
  const auto *fp = (const typename Traits<P>::S *)ctx->f[ctx->cur];
  const T *tab = (const T *)ctx->c.dtab;
  const FluidConsts<double> &c = ctx->c;
  const bool buf = (unsigned long long)g.pelems * sizeof(typename Traits<P>::S) <= 0xFFFFFFFFull;
  if (out) {  // the pulled moments, on the coarse tiles
    auto go = [&](auto kern) {
      kern<<<grid, 256, 0, ctx->stream>>>(g, fp, ctx->own, tab, c.rhor, c.rhor * c.urx, b1, b2, c.rho0, c.u0x,
                                          c.u0y, c.u0z, c.g, nsh, box, Fc, out);
    };
    switch (F) {
      case 1: buf ? go(k_coarse_moments<P, 1, true>) : go(k_coarse_moments<P, 1, false>); break;
      case 2: buf ? go(k_coarse_moments<P, 2, true>) : go(k_coarse_moments<P, 2, false>); break;
      default: buf ? go(k_coarse_moments<P, 4, true>) : go(k_coarse_moments<P, 4, false>); break;
    }
    JHIPCHK(hipGetLastError());
  }
  if (stored) {  // the stored moments, by rows
    JRCHK(ensure_cpart(ctx, F));
    const int rows = g.ny * g.nz;
    auto go = [&](auto kern) {
      kern<<<rows, row_threads(g, 1024), 0, ctx->stream>>>(g, fp, c.rhor, c.rhor * c.urx, c.rho0, c.u0x, c.u0y,
                                                            c.u0z, ctx->d_cpart);
    };
    switch (F) {
      case 1: buf ? go(k_coarse_stored_rows<P, 1, true>) : go(k_coarse_stored_rows<P, 1, false>); break;
      case 2: buf ? go(k_coarse_stored_rows<P, 2, true>) : go(k_coarse_stored_rows<P, 2, false>); break;
      default: buf ? go(k_coarse_stored_rows<P, 4, true>) : go(k_coarse_stored_rows<P, 4, false>); break;
    }
    JHIPCHK(hipGetLastError());
    JHIPCHK(launch_coarse_rows(g, F, ctx->d_cpart, stored, ctx->stream));
  }
  return LBMIB_OK;
}

template <int P>
int apply_launch(lbmib_ctx *ctx, const void *fin, void *fout, const float *delta, int F, float *after) {
  const Geometry &g = ctx->g;
  const FluidConsts<double> &c = ctx->c;
  if (after) JRCHK(ensure_cpart(ctx, F));
  float *part = after ? ctx->d_cpart : nullptr;
  const auto *in = (const typename Traits<P>::S *)fin;
  auto *out = (typename Traits<P>::S *)fout;
  const bool buf = (unsigned long long)g.pelems * sizeof(typename Traits<P>::S) <= 0xFFFFFFFFull;
  const int rows = g.ny * g.nz;
  auto go = [&](auto kern) {
    kern<<<rows, row_threads(g, 512), 0, ctx->stream>>>(g, in, out, c.rhor, c.rhor * c.urx, delta, c.rho0, c.u0x,
                                                         c.u0y, c.u0z, part);
  };
  switch (F) {
    case 1: buf ? go(k_jump_apply_rows<P, 1, true>) : go(k_jump_apply_rows<P, 1, false>); break;
    case 2: buf ? go(k_jump_apply_rows<P, 2, true>) : go(k_jump_apply_rows<P, 2, false>); break;
    default: buf ? go(k_jump_apply_rows<P, 4, true>) : go(k_jump_apply_rows<P, 4, false>); break;
  }
  JHIPCHK(hipGetLastError());
  if (after) JHIPCHK(launch_coarse_rows(g, F, part, after, ctx->stream));
  return LBMIB_OK;
}

int check_coarse(const lbmib_ctx *ctx, int F) {
  if (F != 1 && F != 2 && F != 4) return LBMIB_ERR_ARG;
  const Geometry &g = ctx->g;
  if (g.nx % 64 || g.ny % 4 || g.nz % 4) return LBMIB_ERR_ARG;
  return LBMIB_OK;
}

int ready_for_jump(const lbmib_ctx *ctx) {
  if (ctx->init_bits != 7) return LBMIB_ERR_STATE;
  if (ctx->pending || ctx->own || ctx->split_run || ctx->ib_ready) return LBMIB_ERR_STATE;
  return LBMIB_OK;
}

}  // namespace
}  // namespace lbmib

using namespace lbmib;

extern "C" int lbmib_coarse_state(lbmib_ctx *ctx, int factor, float *d_moments, float *d_stored) {
  if (!ctx || (!d_moments && !d_stored)) return LBMIB_ERR_ARG;
  JRCHK(check_coarse(ctx, factor));
  if (ctx->init_bits != 7 || ctx->pending) return LBMIB_ERR_STATE;
  switch (ctx->fp.precision) {
    case LBMIB_F64: return coarse_launch<LBMIB_F64>(ctx, factor, d_moments, d_stored);
    case LBMIB_F32: return coarse_launch<LBMIB_F32>(ctx, factor, d_moments, d_stored);
    case LBMIB_F16: return coarse_launch<LBMIB_F16>(ctx, factor, d_moments, d_stored);
    default: return coarse_launch<LBMIB_BF16>(ctx, factor, d_moments, d_stored);
  }
}

extern "C" int lbmib_coarse_moments(lbmib_ctx *ctx, int factor, float *d_out) {
  return lbmib_coarse_state(ctx, factor, d_out, nullptr);
}

extern "C" int lbmib_coarse_sheets(lbmib_ctx *ctx, int factor, float *d_out) {
  if (!ctx || !d_out) return LBMIB_ERR_ARG;
  JRCHK(check_coarse(ctx, factor));
  const Geometry &g = ctx->g;
  const long nc = (long)(g.nx / factor) * (g.ny / factor) * (g.nz / factor);
  if (!ctx->d_raster || ctx->raster_cells != nc) {
    if (ctx->d_raster) (void)hipFree(ctx->d_raster);
    ctx->d_raster = nullptr;
    if (hipMalloc((void **)&ctx->d_raster, sizeof(unsigned long long) * 4 * nc) != hipSuccess) return LBMIB_ERR_ALLOC;
    ctx->raster_cells = nc;
  }
  k_raster_zero<<<grid_of(4 * nc), 256, 0, ctx->stream>>>(ctx->d_raster, 4 * nc);
  JHIPCHK(hipGetLastError());
  if (ctx->nnodes > 0) {
    k_raster_nodes<<<grid_of(ctx->nnodes, 1L << 20), 256, 0, ctx->stream>>>(
        ctx->ib.pos, ctx->ib.frc, ctx->nnodes, factor, g.nx / factor, g.ny / factor, g.nz / factor, ctx->d_raster);
    JHIPCHK(hipGetLastError());
  }
  k_raster_final<<<grid_of(4 * nc), 256, 0, ctx->stream>>>(ctx->d_raster, d_out, nc);
  JHIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_jump(lbmib_ctx *ctx, int factor, const float *d_coarse, const float *d_delta, int ksteps,
                          float *d_after) {
  if (!ctx || !d_coarse || !d_delta || ksteps < 2 || (ksteps & 1)) return LBMIB_ERR_ARG;
  JRCHK(check_coarse(ctx, factor));
  JRCHK(ready_for_jump(ctx));
  if (ctx->jump_active) return LBMIB_ERR_STATE;
  const Geometry &g = ctx->g;
  const size_t fbytes = (size_t)g.pelems * ctx->esz;
  if (!ctx->f_spare && hipMalloc(&ctx->f_spare, fbytes) != hipSuccess) {
    ctx->f_spare = nullptr;
    return LBMIB_ERR_ALLOC;
  }
  const int M = ctx->nnodes;
  if (M > 0) {
    if (!ctx->pos_save && hipMalloc((void **)&ctx->pos_save, sizeof(double) * 3 * M) != hipSuccess) {
      ctx->pos_save = nullptr;
      return LBMIB_ERR_ALLOC;
    }
    if (!ctx->F_save && hipMalloc(&ctx->F_save, (size_t)3 * g.n * ctx->fsz) != hipSuccess) {
      ctx->F_save = nullptr;
      return LBMIB_ERR_ALLOC;
    }
    if (!ctx->box_save && hipMalloc((void **)&ctx->box_save, sizeof(int) * kMaxSheets * 6) != hipSuccess) {
      ctx->box_save = nullptr;
      return LBMIB_ERR_ALLOC;
    }
    // the previous step's force slot, which the first step after the jump
    // reads (u = (sum c f + F/2) / rho) and a later one may overwrite
    const int pslot = (ctx->time - 1) & 1;
    const int *pbox = ctx->ib.box + pslot * kMaxSheets * 6;
    JHIPCHK(hipMemcpyAsync(ctx->box_save, pbox, sizeof(int) * kMaxSheets * 6, hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->fsz == 8)
      k_copy_boxes<double><<<256, 256, 0, ctx->stream>>>((const double *)ctx->F[pslot], (double *)ctx->F_save, pbox,
                                                         ctx->nsheets, g);
    else
      k_copy_boxes<float><<<256, 256, 0, ctx->stream>>>((const float *)ctx->F[pslot], (float *)ctx->F_save, pbox,
                                                        ctx->nsheets, g);
    JHIPCHK(hipGetLastError());
    JHIPCHK(hipMemcpyAsync(ctx->pos_save, ctx->ib.pos, sizeof(double) * 3 * M, hipMemcpyDeviceToDevice, ctx->stream));
  }
  // the pre-jump populations go aside; the jumped state is written into the
  // spare buffer, which becomes the current one
  void *old = ctx->f[ctx->cur];
  ctx->f[ctx->cur] = ctx->f_spare;
  ctx->f_spare = old;
  int rc;
  switch (ctx->fp.precision) {
    case LBMIB_F64: rc = apply_launch<LBMIB_F64>(ctx, old, ctx->f[ctx->cur], d_delta, factor, d_after); break;
    case LBMIB_F32: rc = apply_launch<LBMIB_F32>(ctx, old, ctx->f[ctx->cur], d_delta, factor, d_after); break;
    case LBMIB_F16: rc = apply_launch<LBMIB_F16>(ctx, old, ctx->f[ctx->cur], d_delta, factor, d_after); break;
    default: rc = apply_launch<LBMIB_BF16>(ctx, old, ctx->f[ctx->cur], d_delta, factor, d_after); break;
  }
  if (rc) return rc;
  if (M > 0) {
    k_jump_fibers<<<(M + 255) / 256, 256, 0, ctx->stream>>>(ctx->ib.pos, M, d_coarse, d_delta, factor, g.nx / factor,
                                                             g.ny / factor, g.nz / factor, (double)ksteps);
    JHIPCHK(hipGetLastError());
  }
  ctx->jump_active = 1;
  ctx->jump_time = ctx->time;
  ctx->jump_spread_slot = ctx->spread_slot;
  ctx->time += ksteps;
  return LBMIB_OK;
}

extern "C" int lbmib_get_stream(lbmib_ctx *ctx, void **stream) {
  if (!ctx || !stream) return LBMIB_ERR_ARG;
  *stream = (void *)ctx->stream;
  return LBMIB_OK;
}

extern "C" int lbmib_jump_commit(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  if (!ctx->jump_active) return LBMIB_ERR_STATE;
  ctx->jump_active = 0;
  return LBMIB_OK;
}

extern "C" int lbmib_jump_revert(lbmib_ctx *ctx) {
  if (!ctx) return LBMIB_ERR_ARG;
  if (!ctx->jump_active || ctx->pending || ctx->split_run || ctx->ib_ready) return LBMIB_ERR_STATE;
  void *cur = ctx->f[ctx->cur];
  ctx->f[ctx->cur] = ctx->f_spare;
  ctx->f_spare = cur;
  ctx->time = ctx->jump_time;
  ctx->spread_slot = ctx->jump_spread_slot;
  const int M = ctx->nnodes;
  if (M > 0) {
    const Geometry &g = ctx->g;
    const int pslot = (ctx->time - 1) & 1;
    int *pbox = ctx->ib.box + pslot * kMaxSheets * 6;
    JHIPCHK(hipMemcpyAsync(pbox, ctx->box_save, sizeof(int) * kMaxSheets * 6, hipMemcpyDeviceToDevice, ctx->stream));
    if (ctx->fsz == 8)
      k_copy_boxes<double><<<256, 256, 0, ctx->stream>>>((const double *)ctx->F_save, (double *)ctx->F[pslot],
                                                         ctx->box_save, ctx->nsheets, g);
    else
      k_copy_boxes<float><<<256, 256, 0, ctx->stream>>>((const float *)ctx->F_save, (float *)ctx->F[pslot],
                                                        ctx->box_save, ctx->nsheets, g);
    JHIPCHK(hipGetLastError());
    JHIPCHK(hipMemcpyAsync(ctx->ib.pos, ctx->pos_save, sizeof(double) * 3 * M, hipMemcpyDeviceToDevice, ctx->stream));
    // the step at the restored time bins into a clean slot
    JRCHK(launch_ib_reset(ctx->ib, 1, 1 << (ctx->time & 1), ctx->stream));
    ctx->ib_clean = 1;
  }
  ctx->jump_active = 0;
  return LBMIB_OK;
}
