// lbmib_device.h -- device helpers shared by the HIP translation units:
// population storage forms, loads/stores, the support-box test and the
// pull of the 19 arriving populations with the boundary rules.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

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

template <int P, bool NTL = false>
__device__ __forceinline__ typename Traits<P>::T ld(const typename Traits<P>::S *p) {
  typename Traits<P>::S s;
  if constexpr (NTL) s = __builtin_nontemporal_load(p);
  else s = *p;
  if constexpr (P == LBMIB_F16) return (float)s * (1.0f / kDevScale16);
  else if constexpr (P == LBMIB_BF16) return bf16_to_f32(s) * (1.0f / kDevScale16);
  else return s;
}
template <int P, bool NT = false>
__device__ __forceinline__ void st(typename Traits<P>::S *p, typename Traits<P>::T v) {
  typename Traits<P>::S s;
  if constexpr (P == LBMIB_F16) s = (half_t)(v * (float)kDevScale16);
  else if constexpr (P == LBMIB_BF16) s = f32_to_bf16(v * (float)kDevScale16);
  else s = v;
  if constexpr (NT) __builtin_nontemporal_store(s, p);
  else *p = s;
}

__device__ __forceinline__ bool in_box6(const int *b, int x, int y, int z) {
  return x >= b[0] && x <= b[1] && y >= b[2] && y <= b[3] && z >= b[4] && z <= b[5];
}

// Any box overlapping the x-run [x0, x1] of row (y, z).
__device__ __forceinline__ bool any_box_overlaps(const int *boxes, int nsheets, int x0, int x1, int y, int z) {
  bool in = false;
  for (int s = 0; s < nsheets; ++s) {
    const int *b = boxes + 6 * s;
    in |= b[0] <= x1 && b[1] >= x0 && y >= b[2] && y <= b[3] && z >= b[4] && z <= b[5];
  }
  return in;
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
template <int P, bool NTL = false>
__device__ __forceinline__ void pull(const Geometry &g, const typename Traits<P>::S *__restrict__ f, int x, int y,
                                     int z, long c, typename Traits<P>::T v[19]) {
  const bool interior = x > g.lo[0] && x < g.hi[0] && y > g.lo[1] && y < g.hi[1] && z > g.lo[2] && z < g.hi[2];
  const long sx = 1, sy = g.nx, sz = (long)g.nx * g.ny;
  if (interior) {
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const long off = kCX[q] * sx + kCY[q] * sy + kCZ[q] * sz;
      v[q] = ld<P, NTL>(f + q * g.pitch + c - off);
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
    v[q] = ld<P, NTL>(f + q * g.pitch + src);
  }
}

}  // namespace lbmib
