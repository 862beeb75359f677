// lbmib_device.h -- device helpers shared by the HIP translation units:
// population storage forms, loads/stores, the support-box test and the
// pull of the 19 arriving populations with the boundary rules.
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "lbmib_internal.h"

namespace lbmib {

typedef _Float16 half_t;

template <int P>
struct Traits;
// S: stored type, T: arithmetic type, FT: force field type; dev: stored as
// deviations f - r from reference populations r; scale: stored value =
// scale * (f - r).
template <>
struct Traits<LBMIB_F64> {
  using S = double;
  using T = double;
  using FT = double;
  static constexpr bool dev = false;
  static constexpr float scale = 1.0f, inv_scale = 1.0f;
};
template <>
struct Traits<LBMIB_F32> {
  using S = float;
  using T = double;  // FP32 storage, FP64 collision (no rounding drift beyond the stores)
  using FT = float;
  static constexpr bool dev = true;
  static constexpr float scale = 1.0f, inv_scale = 1.0f;
};
template <>
struct Traits<LBMIB_F16> {
  using S = half_t;
  using T = float;
  using FT = float;
  static constexpr bool dev = true;
  static constexpr float scale = (float)kDevScale16, inv_scale = 1.0f / kDevScale16;
};
template <>
struct Traits<LBMIB_BF16> {
  using S = uint16_t;
  using T = float;
  using FT = float;
  static constexpr bool dev = true;
  static constexpr float scale = (float)kDevScale16, inv_scale = 1.0f / kDevScale16;
};

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ uint16_t f32_to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// Stored value <-> kernel units (h = scale (f - r), or f for F64).
template <int P>
__device__ __forceinline__ typename Traits<P>::T from_bits(typename Traits<P>::S s) {
  if constexpr (P == LBMIB_F16) return (float)s;
  else if constexpr (P == LBMIB_BF16) return bf16_to_f32(s);
  else return s;
}
// The 16-bit forms round the FP32 value h once more to 16 bits.  The empty
// asm pins h as an FP32 register value, so the compiler cannot fold the
// conversion into the FMA that produced h (v_fma_mix rounds the exact
// product-sum to 16 bits directly): every kernel shape then stores the same
// bits.
template <int P>
__device__ __forceinline__ typename Traits<P>::S to_bits(typename Traits<P>::T h) {
  if constexpr (P == LBMIB_F16 || P == LBMIB_BF16) {
    __asm__ volatile("" : "+v"(h));
    if constexpr (P == LBMIB_F16) return (half_t)h;
    else return f32_to_bf16(h);
  } else {
    return h;
  }
}
template <int P, bool NTL = false>
__device__ __forceinline__ typename Traits<P>::T ld_h(const typename Traits<P>::S *p) {
  if constexpr (NTL) return from_bits<P>(__builtin_nontemporal_load(p));
  else return from_bits<P>(*p);
}
template <int P, bool NT = false>
__device__ __forceinline__ void st_h(typename Traits<P>::S *p, typename Traits<P>::T h) {
  if constexpr (NT) __builtin_nontemporal_store(to_bits<P>(h), p);
  else *p = to_bits<P>(h);
}

// Buffer-addressed forms: a 128-bit resource over the whole population array,
// a per-lane byte offset (VGPR) and a wave-uniform byte offset (SGPR), so one
// VGPR offset serves all 19 planes.  Offsets must stay below 2^32.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, -1, 0x00020000);
}
template <int P, bool NTL>
__device__ __forceinline__ typename Traits<P>::T bld_h(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  constexpr int aux = NTL ? 2 : 0;  // nt
  if constexpr (P == LBMIB_F64) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, aux));
  } else if constexpr (P == LBMIB_F32) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, aux));
  } else {
    const uint16_t b = __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, aux);
    return from_bits<P>(__builtin_bit_cast(typename Traits<P>::S, b));
  }
}
template <int P, bool NT>
__device__ __forceinline__ void bst_h(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                      typename Traits<P>::T h) {
  constexpr int aux = NT ? 2 : 0;
  const typename Traits<P>::S s = to_bits<P>(h);
  if constexpr (P == LBMIB_F64) {
    typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, s), r, voff, soff, aux);
  } else if constexpr (P == LBMIB_F32) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, s), r, voff, soff, aux);
  } else {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, s), r, voff, soff, aux);
  }
}

// Deviation f - r (or f for F64) of a stored population.
template <int P, bool NTL = false>
__device__ __forceinline__ typename Traits<P>::T ld(const typename Traits<P>::S *p) {
  return ld_h<P, NTL>(p) * Traits<P>::inv_scale;
}
template <int P, bool NT = false>
__device__ __forceinline__ void st(typename Traits<P>::S *p, typename Traits<P>::T v) {
  st_h<P, NT>(p, v * Traits<P>::scale);
}

// A load of memory that no store of the running kernel touches, through the
// constant address space: a wave-uniform address then compiles to a scalar
// load (as for kernel arguments), not to a vector load of one address.
template <class X>
__device__ __forceinline__ X ldc(const X *p) {
  return *(const __attribute__((address_space(4))) X *)p;
}
// Copies a trivially copyable object from such memory into registers, one
// dword at a time (sizeof(X) % 4 == 0).
template <class X>
__device__ __forceinline__ X ldc_obj(const X *p) {
  static_assert(sizeof(X) % 4 == 0, "dword-sized object");
  X out;
  uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
  const uint32_t *src = reinterpret_cast<const uint32_t *>(p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(X) / 4); ++i) dst[i] = ldc(src + i);
  return out;
}

// Support boxes are read-only in every kernel that tests them.
__device__ __forceinline__ bool in_box6(const int *b, int x, int y, int z) {
  return x >= ldc(b) && x <= ldc(b + 1) && y >= ldc(b + 2) && y <= ldc(b + 3) && z >= ldc(b + 4) && z <= ldc(b + 5);
}

// Any box overlapping the x-run [x0, x1] of row (y, z).
__device__ __forceinline__ bool any_box_overlaps(const int *boxes, int nsheets, int x0, int x1, int y, int z) {
  bool in = false;
  for (int s = 0; s < nsheets; ++s) {
    const int *b = boxes + 6 * s;
    in |= ldc(b) <= x1 && ldc(b + 1) >= x0 && y >= ldc(b + 2) && y <= ldc(b + 3) && z >= ldc(b + 4) &&
          z <= ldc(b + 5);
  }
  return in;
}

// Band box of a split step for one sheet (see the split steps in
// lbmib_fluid.hip): the union of its support boxes b1 (this step) and b0
// (the previous step), widened by one cell on each side, in whole x pairs;
// on a periodic axis a box touching an end takes the whole axis; clamped to
// the grid.  False when neither box is valid.
__device__ __forceinline__ bool band_box_of(const Geometry &g, const int b1[6], const int b0[6], int r[6]) {
  const bool v1 = b1[0] <= b1[1], v0 = b0[0] <= b0[1];
  if (!v1 && !v0) return false;
  const int dims[3] = {g.nx, g.ny, g.nz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int lo = INT_MAX, hi = INT_MIN;
    if (v1) { lo = min(lo, b1[2 * a]); hi = max(hi, b1[2 * a + 1]); }
    if (v0) { lo = min(lo, b0[2 * a]); hi = max(hi, b0[2 * a + 1]); }
    lo -= 1;
    hi += 1;
    // whole x pairs (even x0, odd x1), so that no lane of the two-cell row
    // kernel straddles the band
    if (a == 0) { lo -= lo & 1; hi += 1 - (hi & 1); }
    if (g.bc[a] == LBMIB_BC_PERIODIC && (lo < 0 || hi >= dims[a])) { lo = 0; hi = dims[a] - 1; }
    r[2 * a] = max(lo, 0);
    r[2 * a + 1] = min(hi, dims[a] - 1);
  }
  return r[0] <= r[1] && r[2] <= r[3] && r[4] <= r[5];
}

// Whether row (y, z) lies in any support box widened by one cell in y and z
// (wave-uniform).
__device__ __forceinline__ bool row_near_box(const int *boxes, int nsheets, int y, int z) {
  bool in = false;
  for (int s = 0; s < nsheets; ++s) {
    const int *b = boxes + 6 * s;
    in |= y >= ldc(b + 2) - 1 && y <= ldc(b + 3) + 1 && z >= ldc(b + 4) - 1 && z <= ldc(b + 5) + 1;
  }
  return in;
}

__device__ __forceinline__ bool in_box3(const int *r, int x, int y, int z) {
  return x >= r[0] && x <= r[1] && y >= r[2] && y <= r[3] && z >= r[4] && z <= r[5];
}

__device__ __forceinline__ bool in_any_box(const int *boxes, int nsheets, int x, int y, int z) {
  bool in = false;
  for (int s = 0; s < nsheets; ++s) in |= in_box6(boxes + 6 * s, x, y, z);
  return in;
}

// Population q arriving at (x,y,z) (pop_index c) through the boundary rules
// of pull() below.
template <int P, bool NTL = false, bool RAW = false>
__device__ __forceinline__ typename Traits<P>::T pull_one(const Geometry &g, const typename Traits<P>::S *__restrict__ f,
                                                          int x, int y, int z, long c, int q,
                                                          typename Traits<P>::T b1, typename Traits<P>::T b2) {
  const long sy = g.prow, sz = g.prow * g.ny;
  int px = x - kCX[q], py = y - kCY[q], pz = z - kCZ[q];
  bool bb = false;
  if (g.bc[0] == LBMIB_BC_PERIODIC) px = px < 0 ? px + g.nx : (px >= g.nx ? px - g.nx : px);
  else bb |= (px < g.lo[0] || px > g.hi[0]);
  if (g.bc[1] == LBMIB_BC_PERIODIC) py = py < 0 ? py + g.ny : (py >= g.ny ? py - g.ny : py);
  else bb |= (py < g.lo[1] || py > g.hi[1]);
  if (g.bc[2] == LBMIB_BC_PERIODIC) pz = pz < 0 ? pz + g.nz : (pz >= g.nz ? pz - g.nz : pz);
  else bb |= (pz < g.lo[2] || pz > g.hi[2]);
  const long src = bb ? c + (long)kOPP[q] * g.pq - (long)q * g.pq : (long)px + sy * py + sz * pz;
  typename Traits<P>::T v = RAW ? ld_h<P, NTL>(f + q * g.pq + src) : ld<P, NTL>(f + q * g.pq + src);
  if constexpr (Traits<P>::dev) {
    if (kCX[q] != 0 && bb) {
      const typename Traits<P>::T corr = -kCX[q] * (kCLS[q] == 1 ? b1 : b2);
      v += RAW ? corr : corr * Traits<P>::inv_scale;
    }
  }
  return v;
}

// Pulls the 19 populations arriving at (x,y,z) from the post-collision buffer.
// Wall axes: a source outside the active box is replaced by the opposite
// population of the same node (half-way bounce-back, lbm_ib.c:2310-2423).
// Periodic axes wrap.  Interior cells take the branch-free path.
// RAW: values in kernel units (ld_h) instead of deviations (ld).  b1, b2:
// the deviation forms' bounce-back correction s (r_opp - r_q) = -c_x b_class
// (kernel units), since a reflected population is stored relative to r_opp.
template <int P, bool NTL = false, bool RAW = false>
__device__ __forceinline__ void pull(const Geometry &g, const typename Traits<P>::S *__restrict__ f, int x, int y,
                                     int z, typename Traits<P>::T b1, typename Traits<P>::T b2,
                                     typename Traits<P>::T v[19]) {
  const bool interior = x > g.lo[0] && x < g.hi[0] && y > g.lo[1] && y < g.hi[1] && z > g.lo[2] && z < g.hi[2];
  const long c = pop_index(g, x, y, z);
  const long sx = 1, sy = g.prow, sz = g.prow * g.ny;
  if (interior) {
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const long off = kCX[q] * sx + kCY[q] * sy + kCZ[q] * sz;
      v[q] = RAW ? ld_h<P, NTL>(f + q * g.pq + c - off) : ld<P, NTL>(f + q * g.pq + c - off);
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < 19; ++q) v[q] = pull_one<P, NTL, RAW>(g, f, x, y, z, c, q, b1, b2);
}

}  // namespace lbmib
