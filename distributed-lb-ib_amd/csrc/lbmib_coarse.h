// lbmib_coarse.h -- coarse-grained moments shared by the coupling kernels
// (lbmib_jump.hip) and the fused step that reports them (lbmib_fluid.hip):
// the tile of the pulled read and its block means, one cell's stored
// moments, and the per-row partial sums every stored plane set comes from.
#pragma once

#include <type_traits>

#include "lbmib_device.h"

namespace lbmib {

// Arithmetic of the coupling kernels: FP64 for FP64 storage; FP32 on the
// deviations h = f - r for the deviation forms (their moments are rho_r +
// sum h and rho_r u_r + sum c h, so nothing cancels against the O(0.3)
// reference populations).
template <int P>
using JumpT = typename std::conditional<P == LBMIB_F64, double, float>::type;

// s + c v for a lattice component c in {-1, 0, 1}.  c v is exact and adding
// a zero term changes no finite sum, so this keeps the bits of s + (T)c * v
// while sparing the multiply and add that IEEE semantics keep for 0 * v.
template <class T>
__device__ __forceinline__ T add_c(T s, int c, T v) {
  return c > 0 ? s + v : (c < 0 ? s - v : s);
}

// ------------------------------------------------------------ coarse moments
//
// Block: 64 x 4 fine cells of one (x, y) tile; each thread walks the F cells
// of its column in z, so a block covers (64/F) x (4/F) coarse cells of one
// coarse z-plane.  The block sums F^3 fine values per coarse cell in a fixed
// order (lane shuffles in x, LDS in y) and writes their mean.
//
// The launch is 1-D over the (nx/64) x (ny/4) x (nz/F) tiles.  Workgroup b
// runs on XCD b % 8, so when the tile count is a multiple of 8 workgroup b
// takes tile (b % 8) * (count / 8) + b / 8: each XCD sweeps one contiguous
// slab of z-planes, and the rows that a tile pulls from its y and z
// neighbours are in that XCD's own L2.
struct Tile {
  int x, y, z;
};
__device__ __forceinline__ Tile tile_of(const Geometry &g, int F) {
  const int gx = g.nx / 64, gy = g.ny / 4, n = gx * gy * (g.nz / F);
  const int b = blockIdx.x;
  const int t = (n & 7) ? b : (b & 7) * (n >> 3) + (b >> 3);
  return {t % gx, (t / gx) % gy, t / (gx * gy)};
}

// acc: this thread's sums over its F cells in z; out: 4 planes of nc cells.
// The block's first 256 threads hold the tile's columns; threads beyond
// them (a block of one thread per cell) only meet the barriers.
template <int F>
__device__ __forceinline__ void block_mean_store(double acc[4], double (*red)[4][64], const Geometry &g, Tile tl,
                                                 float *__restrict__ out) {
  const int tx = threadIdx.x & 63, ty = (threadIdx.x >> 6) & 3;
  const bool lead = threadIdx.x < 256;
#pragma unroll
  for (int m = 1; m < F; m <<= 1)
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[a] += __shfl_xor(acc[a], m);
  if (lead)
#pragma unroll
    for (int a = 0; a < 4; ++a) red[a][ty][tx] = acc[a];
  __syncthreads();
  if (lead && ty % F == 0 && tx % F == 0) {
    const int x = tl.x * 64 + tx, y = tl.y * 4 + ty;
    const int ncx = g.nx / F, ncy = g.ny / F, ncz = g.nz / F;
    const long nc = (long)ncx * ncy * ncz;
    const long cc = (long)(x / F) + (long)ncx * ((long)(y / F) + (long)ncy * tl.z);
    const double inv = 1.0 / (F * F * F);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double t = 0.0;
      for (int k = 0; k < F; ++k) t += red[a][ty + k][tx];
      out[a * nc + cc] = (float)(t * inv);
    }
  }
  __syncthreads();
}

// One cell's stored moments from its 19 values as stored: rho and
// u = sum c f / rho (kernel units scaled back; the deviation forms add the
// reference's rho_r and rho_r u_r), in the coarse reads' arithmetic.
template <int P>
__device__ __forceinline__ void stored_moments(const typename Traits<P>::S h[19], double rhor, double mrx,
                                               JumpT<P> &rho, JumpT<P> &ux, JumpT<P> &uy, JumpT<P> &uz) {
  using A = JumpT<P>;
  A s1 = 0, s2 = 0, s3 = 0, s4 = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const A fq = (A)(from_bits<P>(h[q]) * Traits<P>::inv_scale);
    s1 += fq;
    s2 = add_c(s2, kCX[q], fq);
    s3 = add_c(s3, kCY[q], fq);
    s4 = add_c(s4, kCZ[q], fq);
  }
  if (Traits<P>::dev) {  // sum r = rho_r, sum c r = rho_r u_r e_x
    s1 += (A)rhor;
    s2 += (A)mrx;
  }
  rho = s1;
  ux = s2 / s1;
  uy = s3 / s1;
  uz = s4 / s1;
}

// ------------------------------------------------- stored planes by rows
//
// Every stored-moment plane set of the library -- lbmib_coarse_state's
// d_stored, lbmib_step_coarse's planes and lbmib_jump's d_after -- comes
// from per-row partial sums in one fixed order, so that equal cell values
// give equal planes: a row pass leaves, for each (y, z) row and each group
// of F cells along x, the FP32 sums of the group's cell values (rho, ux, uy,
// uz; lane shuffles) in part[row][x / F][4]; k_coarse_rows then adds, per
// coarse cell, the F x F rows of its block (over z within each y, then over
// y, in FP64) and writes the mean.  The F lanes of a group must all be
// active (F divides 64, and a block holds whole waves of one row).
template <int F>
__device__ __forceinline__ void row_partial(float m[4], int x, long row, long ncx, float *__restrict__ part) {
#pragma unroll
  for (int d = 1; d < F; d <<= 1)
#pragma unroll
    for (int a = 0; a < 4; ++a) m[a] += __shfl_xor(m[a], d);
  if (x % F == 0) *(float4 *)(part + (row * ncx + x / F) * 4) = make_float4(m[0], m[1], m[2], m[3]);
}

template <int F>
__global__ void __launch_bounds__(256) k_coarse_rows(const Geometry g, const float *__restrict__ part,
                                                     float *__restrict__ out) {
  // (the host checks that the coarse cells and the partials' float4 count
  // fit 32-bit indices; the F x F loads are all issued before the sums)
  const int ncx = g.nx / F, ncy = g.ny / F, ncz = g.nz / F;
  const int nc = ncx * ncy * ncz;
  const int cc = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (cc >= nc) return;
  const int cx = cc % ncx, cy = (cc / ncx) % ncy, cz = cc / (ncx * ncy);
  double t[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < F; ++j) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < F; ++k) {
      const int row = (cy * F + j) + g.ny * (cz * F + k);
      const float4 p = *(const float4 *)(part + ((long)row * ncx + cx) * 4);
      acc[0] += (double)p.x;
      acc[1] += (double)p.y;
      acc[2] += (double)p.z;
      acc[3] += (double)p.w;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) t[a] += acc[a];
  }
  const double inv = 1.0 / (F * F * F);
#pragma unroll
  for (int a = 0; a < 4; ++a) out[(long)a * nc + cc] = (float)(t[a] * inv);
}

// bytes of the partial sums at factor F; the reduction's launch
inline size_t coarse_part_bytes(const Geometry &g, int F) {
  return sizeof(float) * 4 * (size_t)(g.nx / F) * g.ny * g.nz;
}
inline hipError_t launch_coarse_rows(const Geometry &g, int F, const float *part, float *out, hipStream_t s) {
  const long nc = (long)(g.nx / F) * (g.ny / F) * (g.nz / F);
  if ((long)g.ny * g.nz >= (1L << 31) || nc >= (1L << 31)) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)((nc + 255) / 256);
  if (F == 1) k_coarse_rows<1><<<blocks, 256, 0, s>>>(g, part, out);
  else if (F == 2) k_coarse_rows<2><<<blocks, 256, 0, s>>>(g, part, out);
  else k_coarse_rows<4><<<blocks, 256, 0, s>>>(g, part, out);
  return hipGetLastError();
}

}  // namespace lbmib
