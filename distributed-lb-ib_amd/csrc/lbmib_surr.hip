// lbmib_surr.hip -- kernels of the learned surrogate (include/lbmib_surrogate.h):
// 3x3x3 convolutions on channels-last BF16 tensors with BF16 MFMA and FP32
// accumulation, their data and weight gradients, 2x pooling / upsampling, the
// loss, the optimiser and the packing of field planes.
//
// Convolutions are implicit GEMMs over voxels v, output channels n and the
// reduction index k = 27 taps x Cin input channels (tap-major, channel
// fastest, the weight layout [Cout][27][Cin]):
//   y[v][n] = sum_k W[n][k] X(v, k),   X(v, t*Cin + ci) = x[v + o(t)][ci]
// with o(t) the neighbour offset of tap t and zero outside the volume.  One
// v_mfma_f32_16x16x32_bf16 multiplies a 16 x 32 weight fragment (A: row n =
// lane & 15, k = 8 (lane >> 4) .. + 8) by a 32 x 16 input fragment (B:
// column v = lane & 15, the same 8 k), so that every lane fetches 8
// consecutive channels of one voxel (16 bytes) and of one weight row; the
// 16 x 16 result holds, per lane, 4 consecutive output channels of one voxel
// (rows 4 (lane >> 4) + r), stored as one 8-byte write.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/lbmib.h"
#include "../../include/lbmib_surrogate.h"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

#define HIPCHK(x)                                   \
  do {                                              \
    if ((x) != hipSuccess) return LBMIB_ERR_HIP;    \
  } while (0)

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// round to nearest even (finite inputs; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
  return __builtin_bit_cast(bf16x8, *(const u32x4 *)p);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u32x4{0u, 0u, 0u, 0u}); }

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__host__ __device__ inline int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) ++r;
  return r;
}

// ---------------------------------------------------------------- forward
//
// Block: 4 waves stacked along the voxels; a wave owns WM x 16 voxels and
// WN x 16 output channels (WM x WN accumulator fragments).  The reduction
// runs over chunks of 32 k; the fragments of chunk c + 1 are loaded while
// chunk c multiplies.  Operands come straight from global memory (L1/L2):
// every voxel's 8-channel pieces are 16-byte aligned loads.
//
// SPLIT: the reduction is split over gridDim.z (chunks_per_split chunks
// each) and the block writes its FP32 partial sums to part[z][v][n] instead
// of the epilogue; k_conv_reduce then sums the partials in order of z.
template <int WM, int WN, bool SPLIT>
__global__ void __launch_bounds__(256) k_conv_fwd(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias, const uint16_t *__restrict__ mask,
                                                  uint16_t *__restrict__ y, int D, int H, int W, long M, int Cin,
                                                  int lgCin, int Cout, int K, int relu, float *__restrict__ part,
                                                  int chunks_per_split) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const long mw = ((long)blockIdx.x * 4 + wave) * (WM * 16);
  const int n0 = blockIdx.y * (WN * 16);
  const long HW = (long)H * W;
  int vz[WM], vy[WM], vx[WM];
  bool vin[WM];
#pragma unroll
  for (int i = 0; i < WM; ++i) {
    const long m = mw + i * 16 + col;
    vin[i] = m < M;
    long r = vin[i] ? m : 0;
    vx[i] = (int)(r % W);
    r /= W;
    vy[i] = (int)(r % H);
    r /= H;
    vz[i] = (int)(r % D);
  }
  f32x4 acc[WN][WM];
#pragma unroll
  for (int j = 0; j < WN; ++j)
#pragma unroll
    for (int i = 0; i < WM; ++i) acc[j][i] = f32x4{0.f, 0.f, 0.f, 0.f};

  int c_begin = 0, c_end = (K + 31) >> 5;
  if (SPLIT) {
    c_begin = blockIdx.z * chunks_per_split;
    c_end = min(c_end, c_begin + chunks_per_split);
  }
  // The weight tile of a chunk (WN x 16 output channels x 32 k) is the same
  // for the block's four waves: it is loaded once per block, 16 bytes per
  // thread, into LDS (double-buffered, rows padded to 40 elements so that the
  // fragment reads of a wave spread over the banks); each wave reads its A
  // fragments from there.  The input fragments differ per wave and come
  // straight from L1/L2.
  constexpr int kPitch = 40;
  constexpr int kPieces = WN * 16 * 4;  // 16-byte pieces of a weight tile
  __shared__ __attribute__((aligned(16))) uint16_t sw[2][WN * 16 * kPitch];
  const int tid = threadIdx.x;
  const int prow = tid >> 2, pk = (tid & 3) << 3;  // this thread's piece: row, k offset
  u32x4 wpiece = {0u, 0u, 0u, 0u};
  auto load_w = [&](int c) {
    if (tid < kPieces) {
      const int k = (c << 5) + pk, n = n0 + prow;
      wpiece = (k < K && n < Cout) ? *(const u32x4 *)(w + (long)n * K + k) : u32x4{0u, 0u, 0u, 0u};
    }
  };
  auto store_w = [&](int buf) {
    if (tid < kPieces) *(u32x4 *)(&sw[buf][prow * kPitch + pk]) = wpiece;
  };
  bf16x8 b[WM];
  auto load_x = [&](int c) {
    const int k = (c << 5) + (kq << 3);
    const bool kin = k < K;
    const int tap = k >> lgCin, ci = k & (Cin - 1);
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int zz = vz[i] + dz, yy = vy[i] + dy, xx = vx[i] + dx;
      const bool ok = kin && vin[i] && (unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H &&
                      (unsigned)xx < (unsigned)W;
      const long m = mw + i * 16 + col;
      b[i] = ok ? ld8(x + (m + dz * HW + (long)dy * W + dx) * Cin + ci) : zero8();
    }
  };
  if (c_begin < c_end) {
    load_w(c_begin);
    load_x(c_begin);
    store_w(0);
  }
  __syncthreads();
  for (int c = c_begin; c < c_end; ++c) {
    const int buf = (c - c_begin) & 1;
    bf16x8 cb[WM];
#pragma unroll
    for (int i = 0; i < WM; ++i) cb[i] = b[i];
    if (c + 1 < c_end) {
      load_w(c + 1);
      load_x(c + 1);
    }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const bf16x8 ca = __builtin_bit_cast(bf16x8, *(const u32x4 *)(&sw[buf][(j * 16 + col) * kPitch + (kq << 3)]));
#pragma unroll
      for (int i = 0; i < WM; ++i) acc[j][i] = mfma16(ca, cb[i], acc[j][i]);
    }
    if (c + 1 < c_end) store_w(buf ^ 1);
    __syncthreads();
  }
  // epilogue: lane holds channels n0 + 16 j + 4 kq + r of voxel mw + 16 i + col
  if (SPLIT) {
    float *dst = part + (long)blockIdx.z * M * Cout;
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int n = n0 + j * 16 + 4 * kq;
      if (n >= Cout) continue;
#pragma unroll
      for (int i = 0; i < WM; ++i) {
        if (!vin[i]) continue;
        *(f32x4 *)(dst + (mw + i * 16 + col) * Cout + n) = acc[j][i];
      }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const int n = n0 + j * 16 + 4 * kq;
    if (n >= Cout) continue;
    float bs[4] = {0.f, 0.f, 0.f, 0.f};
    if (bias)
#pragma unroll
      for (int r = 0; r < 4; ++r) bs[r] = bias[n + r];
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      if (!vin[i]) continue;
      const long m = mw + i * 16 + col;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        v[r] = acc[j][i][r] + bs[r];
        if (relu) v[r] = fmaxf(v[r], 0.f);
      }
      if (mask) {
        const u32x2 mk = *(const u32x2 *)(mask + m * Cout + n);
        const uint16_t mb[4] = {(uint16_t)(mk[0] & 0xffff), (uint16_t)(mk[0] >> 16), (uint16_t)(mk[1] & 0xffff),
                                (uint16_t)(mk[1] >> 16)};
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (!(bf2f(mb[r]) > 0.f)) v[r] = 0.f;
      }
      u32x2 o;
      o[0] = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
      o[1] = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
      *(u32x2 *)(y + m * Cout + n) = o;
    }
  }
}

// Large-grid forward convolution with the input staged in LDS.  A block owns
// 256 consecutive voxels in linear order -- R = 256 / W whole x-rows of one
// z-plane of one sample (W divides 256, R divides H, W <= 64) -- and WN x 16
// output channels; its 4 waves own 64 voxels each (4 fragments of 16).  Per
// block of 8 input channels it stages in LDS the input box around the tile,
// (W + 2) x (R + 2) x 3 voxels with zeros outside the volume (16 bytes a
// voxel), and the weights of its output channels for all 27 taps, so each
// input voxel is fetched from memory once per block and channel block
// instead of once per tap, and the 7 chunks of the channel block (four taps
// x 8 channels each: slot 8 kq + e of chunk p, kq = lane >> 4, is tap
// 4p + kq, channel c0 + e, in both operands; a 28th tap is zero) read both
// operands from LDS without a barrier.  The next channel block's box and
// weights are loaded into registers while the current one multiplies
// (buffer loads: a piece with nothing to load points past the end of its
// tensor and reads zeros, so the loads carry no branch and no early wait).
constexpr int kBoxPieces = 5;  // ceil(box voxels / 256) for W = 16, 32, 64
constexpr int kWRow = 28 * 8 + 8;  // weight row pitch: 8 rows' 16-byte reads on distinct banks
__host__ __device__ inline int box_voxels(int W) { return (W + 2) * (256 / W + 2) * 3; }
inline bool box_conv_ok(int H, int W, long M, int Cin, int Cout) {
  return W >= 16 && W <= 64 && 256 % W == 0 && H % (256 / W) == 0 && M % 256 == 0 &&
         M * Cin * 2 < (1L << 31) && (long)Cout * 27 * Cin * 2 < (1L << 30);
}

template <int WN>
__global__ void __launch_bounds__(256) k_conv_fwd_box(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                                                      const float *__restrict__ bias,
                                                      const uint16_t *__restrict__ mask, uint16_t *__restrict__ y,
                                                      int D, int H, int W, long M, int Cin, int Cout, int relu) {
  constexpr int WM = 4, kWPieces = (WN * 16 * 28 + 255) / 256;
  // [box voxel][8 channels], then the weights [WN x 16 rows][kWRow]
  extern __shared__ __attribute__((aligned(16))) uint16_t xs[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const long m0 = (long)blockIdx.x * 256;
  const int n0 = blockIdx.y * (WN * 16);
  const int R = 256 / W, BX = W + 2, BY = R + 2, NB = BX * BY * 3;
  const long HW = (long)H * W, plane = m0 / HW;  // the tile's (sample, z) plane
  const int zt = (int)(plane % D), y0 = (int)((m0 % HW) / W);
  const int K = 27 * Cin;
  uint16_t *ws = xs + NB * 8;
  int vb[WM];  // box index of each fragment's voxel (box coordinates x + 1, row + 1, 1)
#pragma unroll
  for (int i = 0; i < WM; ++i) {
    const int lm = wave * 64 + i * 16 + col;
    vb[i] = (BY + lm / W + 1) * BX + lm % W + 1;
  }
  f32x4 acc[WN][WM];
#pragma unroll
  for (int j = 0; j < WN; ++j)
#pragma unroll
    for (int i = 0; i < WM; ++i) acc[j][i] = f32x4{0.f, 0.f, 0.f, 0.f};
  const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(x), (short)0,
                                                                      (int)(M * Cin * 2), 0x00020000);
  const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(w), (short)0,
                                                                      Cout * K * 2, 0x00020000);
  constexpr uint32_t kNone = 0x80000000u;  // past the end of both resources: loads zeros
  // this thread's pieces: byte offsets relative to the channel block (or kNone)
  uint32_t xo[kBoxPieces], wo[kWPieces];
#pragma unroll
  for (int r = 0; r < kBoxPieces; ++r) {
    const int vi = tid + 256 * r;
    const int bx = vi % BX, by = (vi / BX) % BY, bz = vi / (BX * BY);
    const int gx = bx - 1, gy = y0 + by - 1, gz = zt + bz - 1;
    const bool in = vi < NB && (unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D;
    xo[r] = in ? (uint32_t)((((plane - zt + gz) * HW + (long)gy * W + gx) * Cin) * 2) : kNone;
  }
#pragma unroll
  for (int r = 0; r < kWPieces; ++r) {
    const int e = tid + 256 * r, row = e / 28, tap = e % 28, n = n0 + row;
    wo[r] = (row < WN * 16 && tap < 27 && n < Cout) ? (uint32_t)((n * K + tap * Cin) * 2) : kNone;
  }
  u32x4 px[kBoxPieces], pw[kWPieces];
  auto fetch = [&](int c0) {  // the channel block's pieces into registers
    const uint32_t cb = (uint32_t)(c0 * 2);
#pragma unroll
    for (int r = 0; r < kBoxPieces; ++r)
      px[r] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rx, xo[r] | 0u, cb, 0));
#pragma unroll
    for (int r = 0; r < kWPieces; ++r)
      pw[r] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rw, wo[r], cb, 0));
  };
  auto put = [&]() {  // registers into LDS
#pragma unroll
    for (int r = 0; r < kBoxPieces; ++r) {
      const int vi = tid + 256 * r;
      if (vi < NB) *(u32x4 *)(xs + vi * 8) = px[r];
    }
#pragma unroll
    for (int r = 0; r < kWPieces; ++r) {
      const int e = tid + 256 * r;
      if (e < WN * 16 * 28) *(u32x4 *)(ws + (e / 28) * kWRow + (e % 28) * 8) = pw[r];
    }
  };
  fetch(0);
  for (int c0 = 0; c0 < Cin; c0 += 8) {
    __syncthreads();  // the previous channel block's readers are done
    put();
    __syncthreads();
    if (c0 + 8 < Cin) fetch(c0 + 8);  // in flight while this block multiplies
#pragma unroll 2
    for (int p = 0; p < 7; ++p) {
      const int tap = 4 * p + kq;
      bf16x8 b[WM];
      if (tap < 27) {
        const int off = ((tap / 9 - 1) * BY + (tap / 3) % 3 - 1) * BX + tap % 3 - 1;
#pragma unroll
        for (int i = 0; i < WM; ++i) b[i] = __builtin_bit_cast(bf16x8, *(const u32x4 *)(xs + (vb[i] + off) * 8));
      } else {
#pragma unroll
        for (int i = 0; i < WM; ++i) b[i] = zero8();
      }
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u32x4 *)(ws + (j * 16 + col) * kWRow + 32 * p + 8 * kq));
#pragma unroll
        for (int i = 0; i < WM; ++i) acc[j][i] = mfma16(a, b[i], acc[j][i]);
      }
    }
  }
  // epilogue: lane holds channels n0 + 16 j + 4 kq + r of voxel m0 + 64 wave + 16 i + col
  const long mw = m0 + wave * 64;
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const int n = n0 + j * 16 + 4 * kq;
    if (n >= Cout) continue;
    float bs[4] = {0.f, 0.f, 0.f, 0.f};
    if (bias)
#pragma unroll
      for (int r = 0; r < 4; ++r) bs[r] = bias[n + r];
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const long m = mw + i * 16 + col;
      float v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        v[r] = acc[j][i][r] + bs[r];
        if (relu) v[r] = fmaxf(v[r], 0.f);
      }
      if (mask) {
        const u32x2 mk = *(const u32x2 *)(mask + m * Cout + n);
        const uint16_t mb[4] = {(uint16_t)(mk[0] & 0xffff), (uint16_t)(mk[0] >> 16), (uint16_t)(mk[1] & 0xffff),
                                (uint16_t)(mk[1] >> 16)};
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (!(bf2f(mb[r]) > 0.f)) v[r] = 0.f;
      }
      u32x2 o;
      o[0] = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
      o[1] = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
      *(u32x2 *)(y + m * Cout + n) = o;
    }
  }
}

template <int WN>
int launch_fwd_box(const void *x, const void *w, const float *bias, const void *mask, void *y, int D, int H, int W,
                   long M, int Cin, int Cout, int relu, hipStream_t s) {
  const dim3 grid((unsigned)(M / 256), (unsigned)((Cout + WN * 16 - 1) / (WN * 16)));
  const size_t shm = sizeof(uint16_t) * (8 * box_voxels(W) + WN * 16 * kWRow);
  k_conv_fwd_box<WN><<<grid, 256, shm, s>>>((const uint16_t *)x, (const uint16_t *)w, bias, (const uint16_t *)mask,
                                            (uint16_t *)y, D, H, W, M, Cin, Cout, relu);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

// Small-grid forward convolution: the KW waves of a block all own the same
// WM x 16 voxels x WN x 16 output channels and split the reduction between
// them -- wave w multiplies chunks w, w + KW, w + 2 KW, ... of 32 k.  Both
// operands come straight from global memory (L1/L2) into a ring of P
// register stages, so that P chunks' loads are in flight while one
// multiplies, and the loop has no barrier.  The loads are raw buffer loads
// whose resources end at the tensors' ends: a lane with nothing to load
// (a neighbour outside the volume, k past 27 Cin) points its offset past
// the end and the hardware returns zeros, so the loads carry no branch and
// the waits stay per stage.  At the end the KW partial sums
// meet in LDS and are added in wave order (a fixed order: the result does
// not depend on timing), then bias, activation and mask as in k_conv_fwd.
// One launch per layer, where a split over blocks needs a second pass over
// FP32 partials in memory.
template <int WM, int WN, int KW, int P>
__global__ void __launch_bounds__(KW * 64) k_conv_fwd_kw(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                                                         const float *__restrict__ bias,
                                                         const uint16_t *__restrict__ mask, uint16_t *__restrict__ y,
                                                         int D, int H, int W, long M, int Cin, int lgCin, int Cout,
                                                         int K, int relu) {
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, kq = lane >> 4;
  const long mb = (long)blockIdx.x * (WM * 16);
  const int n0 = blockIdx.y * (WN * 16);
  const long HW = (long)H * W;
  int vz[WM], vy[WM], vx[WM];
  bool vin[WM];
#pragma unroll
  for (int i = 0; i < WM; ++i) {
    const long m = mb + i * 16 + col;
    vin[i] = m < M;
    long r = vin[i] ? m : 0;
    vx[i] = (int)(r % W);
    r /= W;
    vy[i] = (int)(r % H);
    r /= H;
    vz[i] = (int)(r % D);
  }
  f32x4 acc[WN][WM];
#pragma unroll
  for (int j = 0; j < WN; ++j)
#pragma unroll
    for (int i = 0; i < WM; ++i) acc[j][i] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nch = (K + 31) >> 5;
  const int nw = nch > wave ? (nch - wave + KW - 1) / KW : 0;  // this wave's chunks
  const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(w), (short)0,
                                                                      Cout * K * 2, 0x00020000);
  const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t *>(x), (short)0,
                                                                      (int)(M * Cin * 2), 0x00020000);
  constexpr uint32_t kNone = 0x80000000u;  // past the end of both resources: loads zeros
  bf16x8 a[P][WN], b[P][WM];
  // stage s <- the wave's t-th chunk; slot 8 kq + e of a fragment is k = 32 c + 8 kq + e
  auto load = [&](int s, int t) {
    const int k = ((wave + t * KW) << 5) + (kq << 3);
    const bool kin = k < K;
    // k past 27 Cin needs no test on this side: the input slot is zero
    // there (the weights are finite), and past the end of w the load is zero
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const uint32_t off = (uint32_t)(((n0 + j * 16 + col) * K + k) * 2);
      a[s][j] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rw, off, 0, 0));
    }
    const int tap = k >> lgCin, ci = k & (Cin - 1);
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
      const int zz = vz[i] + dz, yy = vy[i] + dy, xx = vx[i] + dx;
      const bool ok = kin && vin[i] && (unsigned)zz < (unsigned)D && (unsigned)yy < (unsigned)H &&
                      (unsigned)xx < (unsigned)W;
      const int m = (int)mb + i * 16 + col;
      const uint32_t off = ok ? (uint32_t)(((m + dz * (int)HW + dy * W + dx) * Cin + ci) * 2) : kNone;
      b[s][i] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rx, off, 0, 0));
    }
  };
#pragma unroll
  for (int s = 0; s < P; ++s) load(s, s);
  for (int t0 = 0; t0 < nw; t0 += P) {
#pragma unroll
    for (int s = 0; s < P; ++s) {
      if (t0 + s < nw) {
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
          for (int i = 0; i < WM; ++i) acc[j][i] = mfma16(a[s][j], b[s][i], acc[j][i]);
      }
      load(s, t0 + s + P);
    }
  }
  // the waves' partial sums, fragment f = j WM + i, lane l
  __shared__ f32x4 red[KW][WN * WM][64];
#pragma unroll
  for (int j = 0; j < WN; ++j)
#pragma unroll
    for (int i = 0; i < WM; ++i) red[wave][j * WM + i][lane] = acc[j][i];
  __syncthreads();
  for (int e = threadIdx.x; e < WN * WM * 64; e += KW * 64) {
    const int f = e >> 6, l = e & 63;
    const int j = f / WM, i = f % WM;
    f32x4 sum = red[0][f][l];
#pragma unroll
    for (int v = 1; v < KW; ++v) sum += red[v][f][l];
    // lane l of fragment (j, i): channels n0 + 16 j + 4 (l >> 4) + r of voxel mb + 16 i + (l & 15)
    const int n = n0 + j * 16 + 4 * (l >> 4);
    const long m = mb + i * 16 + (l & 15);
    if (n >= Cout || m >= M) continue;
    float v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[r] = sum[r] + (bias ? bias[n + r] : 0.f);
      if (relu) v[r] = fmaxf(v[r], 0.f);
    }
    if (mask) {
      const u32x2 mk = *(const u32x2 *)(mask + m * Cout + n);
      const uint16_t mbits[4] = {(uint16_t)(mk[0] & 0xffff), (uint16_t)(mk[0] >> 16), (uint16_t)(mk[1] & 0xffff),
                                 (uint16_t)(mk[1] >> 16)};
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (!(bf2f(mbits[r]) > 0.f)) v[r] = 0.f;
    }
    u32x2 o;
    o[0] = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
    o[1] = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
    *(u32x2 *)(y + m * Cout + n) = o;
  }
}

template <int WM, int WN, int KW, int P>
int launch_fwd_kw(const void *x, const void *w, const float *bias, const void *mask, void *y, int D, int H, int W,
                  long M, int Cin, int Cout, int relu, hipStream_t s) {
  const dim3 grid((unsigned)((M + WM * 16 - 1) / (WM * 16)), (unsigned)((Cout + WN * 16 - 1) / (WN * 16)));
  k_conv_fwd_kw<WM, WN, KW, P><<<grid, KW * 64, 0, s>>>((const uint16_t *)x, (const uint16_t *)w, bias,
                                                        (const uint16_t *)mask, (uint16_t *)y, D, H, W, M, Cin,
                                                        ilog2(Cin), Cout, 27 * Cin, relu);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

// Small samples (8^3, 16^3: the z-planes a block of 256 voxels spans, one
// more on each side, padded by one cell in x and y, hold at most kVolMax
// voxels): a block owns up to 256 voxels of one sample (4 waves x 4
// fragments of 16), 16 output channels, and one slice of 16 input channels
// of the reduction (gridDim.z slices).  It stages that padded slab of the
// slice in LDS (zeros outside the volume; rows of 24 elements, so that 8
// lanes' 16-byte reads fall on distinct banks) and the slice's weights for
// all 27 taps (rows of 28 x 16 + 8 elements), then multiplies its 14 chunks from LDS with no
// barrier: slot 8 kq + e of chunk p is tap 2p + (kq >> 1), channel
// c0 + 8 (kq & 1) + e, in both operands (a 28th tap is zero).  The FP32
// partial of the slice goes to part[slice][v][n]; k_conv_reduce adds the
// slices in order with bias, activation and mask.  Each input and weight
// element is then read once per block, where the split over chunks reads
// the weights once per 64 voxels and the inputs once per tap.
constexpr int kVolMax = 1024, kVolPitch = 24, kVolWRow = 28 * 16 + 8;

__global__ void __launch_bounds__(256) k_conv_fwd_vol(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                                                      float *__restrict__ part, int D, int H, int W, long M, int Cin,
                                                      int Cout) {
  extern __shared__ __attribute__((aligned(16))) uint16_t ws[];  // [16 rows][kVolWRow], then [slab voxel][24]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int S = D * H * W, parts = (S + 255) / 256;
  const int smp = blockIdx.x / parts, v0 = (blockIdx.x % parts) * 256;
  const int n0 = blockIdx.y * 16, c0 = blockIdx.z * 16, K = 27 * Cin;
  // the z-planes of the block's voxels, and one more on each side
  const int zlo = v0 / (H * W), zhi = min(v0 + 255, S - 1) / (H * W);
  const int PX = W + 2, PY = H + 2, NP = PX * PY * (zhi - zlo + 3);
  uint16_t *xs = ws + 16 * kVolWRow;
  const long base = (long)smp * S;
  for (int e = tid; e < 2 * NP; e += 256) {
    const int vi = e >> 1, half = e & 1;
    const int gx = vi % PX - 1, gy = (vi / PX) % PY - 1, gz = zlo - 1 + vi / (PX * PY);
    u32x4 v = {0u, 0u, 0u, 0u};
    if ((unsigned)gx < (unsigned)W && (unsigned)gy < (unsigned)H && (unsigned)gz < (unsigned)D)
      v = *(const u32x4 *)(x + (base + ((long)gz * H + gy) * W + gx) * Cin + c0 + 8 * half);
    *(u32x4 *)(xs + vi * kVolPitch + 8 * half) = v;
  }
  for (int e = tid; e < 16 * 56; e += 256) {
    const int row = e / 56, pc = e % 56, tap = pc >> 1, half = pc & 1, n = n0 + row;
    u32x4 v = {0u, 0u, 0u, 0u};
    if (tap < 27 && n < Cout) v = *(const u32x4 *)(w + (long)n * K + tap * Cin + c0 + 8 * half);
    *(u32x4 *)(ws + row * kVolWRow + tap * 16 + 8 * half) = v;
  }
  __syncthreads();
  int vb[4], lv[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    lv[i] = v0 + wave * 64 + i * 16 + col;
    const int l = lv[i] < S ? lv[i] : v0;  // past the end: any voxel of the slab
    const int vx = l % W, vy = (l / W) % H, vz = l / (W * H);
    vb[i] = ((vz - zlo + 1) * PY + vy + 1) * PX + vx + 1;
  }
  f32x4 acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
  for (int p = 0; p < 14; ++p) {
    const int tap = 2 * p + (kq >> 1);
    bf16x8 b[4];
    if (tap < 27) {
      const int off = ((tap / 9 - 1) * PY + (tap / 3) % 3 - 1) * PX + tap % 3 - 1;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        b[i] = __builtin_bit_cast(bf16x8, *(const u32x4 *)(xs + (vb[i] + off) * kVolPitch + 8 * (kq & 1)));
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) b[i] = zero8();
    }
    const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u32x4 *)(ws + col * kVolWRow + 32 * p + 8 * kq));
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = mfma16(a, b[i], acc[i]);
  }
  // lane holds channels n0 + 4 kq + r of voxel lv[i]
  float *dst = part + (long)blockIdx.z * M * Cout;
  const int n = n0 + 4 * kq;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (lv[i] < S && n < Cout) *(f32x4 *)(dst + (base + lv[i]) * Cout + n) = acc[i];
}

// y[v][n] = act(sum over z < ns of part[z][v][n] + bias[n]) [* (mask > 0)]:
// one thread per (voxel, 4 channels)
__global__ void k_conv_reduce(const float *__restrict__ part, int ns, const float *__restrict__ bias,
                              const uint16_t *__restrict__ mask, uint16_t *__restrict__ y, long M, int Cout, int relu) {
  const long nq = M * (Cout / 4);
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < nq; e += (long)gridDim.x * blockDim.x) {
    const long off = e * 4;
    f32x4 s = {0.f, 0.f, 0.f, 0.f};
    for (int z = 0; z < ns; ++z) s += *(const f32x4 *)(part + (long)z * M * Cout + off);
    const int n = (int)(off % Cout);
    float v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      v[r] = s[r] + (bias ? bias[n + r] : 0.f);
      if (relu) v[r] = fmaxf(v[r], 0.f);
    }
    if (mask) {
      const u32x2 mk = *(const u32x2 *)(mask + off);
      const uint16_t mb[4] = {(uint16_t)(mk[0] & 0xffff), (uint16_t)(mk[0] >> 16), (uint16_t)(mk[1] & 0xffff),
                              (uint16_t)(mk[1] >> 16)};
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (!(bf2f(mb[r]) > 0.f)) v[r] = 0.f;
    }
    u32x2 o;
    o[0] = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
    o[1] = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
    *(u32x2 *)(y + off) = o;
  }
}

int workspace(void **out);

// ksplit > 1: split reduction through the library workspace
template <int WM, int WN>
int launch_fwd(const void *x, const void *w, const float *bias, const void *mask, void *y, int D, int H, int W, long M,
               int Cin, int Cout, int relu, int ksplit, hipStream_t s) {
  const int K = 27 * Cin;
  const long bm = 4L * WM * 16;
  const int nch = (K + 31) / 32;
  dim3 grid((unsigned)((M + bm - 1) / bm), (unsigned)((Cout + WN * 16 - 1) / (WN * 16)));
  if (ksplit <= 1) {
    k_conv_fwd<WM, WN, false><<<grid, 256, 0, s>>>((const uint16_t *)x, (const uint16_t *)w, bias,
                                                    (const uint16_t *)mask, (uint16_t *)y, D, H, W, M, Cin,
                                                    ilog2(Cin), Cout, K, relu, nullptr, nch);
    HIPCHK(hipGetLastError());
    return LBMIB_OK;
  }
  void *ws = nullptr;
  const int rc = workspace(&ws);
  if (rc) return rc;
  const int cps = (nch + ksplit - 1) / ksplit;
  grid.z = (nch + cps - 1) / cps;
  k_conv_fwd<WM, WN, true><<<grid, 256, 0, s>>>((const uint16_t *)x, (const uint16_t *)w, nullptr, nullptr, nullptr,
                                                 D, H, W, M, Cin, ilog2(Cin), Cout, K, relu, (float *)ws, cps);
  HIPCHK(hipGetLastError());
  const long nq = M * (Cout / 4);
  long nb = (nq + 255) / 256;
  if (nb > 4096) nb = 4096;
  k_conv_reduce<<<(unsigned)nb, 256, 0, s>>>((const float *)ws, (int)grid.z, bias, (const uint16_t *)mask,
                                              (uint16_t *)y, M, Cout, relu);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

// the largest padded slab over a sample's blocks of 256 voxels
long vol_slab_voxels(int D, int H, int W) {
  const long S = (long)D * H * W, HW = (long)H * W;
  long span = 0;
  for (long v0 = 0; v0 < S; v0 += 256) {
    const long v1 = v0 + 255 < S - 1 ? v0 + 255 : S - 1;
    const long sp = v1 / HW - v0 / HW + 1;
    if (sp > span) span = sp;
  }
  return (long)(W + 2) * (H + 2) * (span + 2);
}

// Taken where it measured faster than the other small-grid forms: samples
// of at most 512 voxels, or larger ones with at least 128 input channels
// (16^3, 128 -> 64: 17.7 against 22.2 us; with 64 input channels the
// two-wave form is faster: 13.9 against 16.6 us at 64 -> 128), and the
// slices' partials at most 8 MB (beyond, the reduction's pass over them
// cost more than the form saves: 16 MB at batch 4 on 8^3, 128 -> 256,
// 25.5 against 23.6 us split over chunks; 32 MB at 256 -> 256, 49.5 against
// 32.5 us).
bool vol_conv_ok(int D, int H, int W, long M, int Cin, int Cout) {
  const long S = (long)D * H * W;
  return (S <= 512 || (S <= (1L << 20) && Cin >= 128)) && vol_slab_voxels(D, H, W) <= kVolMax && Cin % 16 == 0 &&
         (size_t)(Cin / 16) * M * Cout * sizeof(float) <= (8u << 20);
}

int launch_fwd_vol(const void *x, const void *w, const float *bias, const void *mask, void *y, int D, int H, int W,
                   long M, int Cin, int Cout, int relu, hipStream_t s) {
  void *ws = nullptr;
  const int rc = workspace(&ws);
  if (rc) return rc;
  const int S = D * H * W;
  const dim3 grid((unsigned)(M / S * ((S + 255) / 256)), (unsigned)(Cout / 16), (unsigned)(Cin / 16));
  const size_t shm = sizeof(uint16_t) * ((size_t)vol_slab_voxels(D, H, W) * kVolPitch + 16 * kVolWRow);
  k_conv_fwd_vol<<<grid, 256, shm, s>>>((const uint16_t *)x, (const uint16_t *)w, (float *)ws, D, H, W, M, Cin, Cout);
  HIPCHK(hipGetLastError());
  const long nq = M * (Cout / 4);
  long nb = (nq + 255) / 256;
  if (nb > 4096) nb = 4096;
  k_conv_reduce<<<(unsigned)nb, 256, 0, s>>>((const float *)ws, (int)grid.z, bias, (const uint16_t *)mask,
                                              (uint16_t *)y, M, Cout, relu);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

bool shape_ok(int B, int D, int H, int W, int Cin, int Cout) {
  if (B < 1 || D < 1 || H < 1 || W < 1) return false;
  if (Cin < 16 || Cin > 256 || (Cin & (Cin - 1))) return false;
  if (Cout < 16 || Cout > 512 || Cout % 16) return false;
  const long M = (long)B * D * H * W;
  return M * (long)(Cin > Cout ? Cin : Cout) < (1L << 40);
}

int conv_fwd(const void *x, const void *w, const float *bias, const void *mask, void *y, int B, int D, int H, int W,
             int Cin, int Cout, int relu, hipStream_t s) {
  if (!x || !w || !y || !shape_ok(B, D, H, W, Cin, Cout)) return LBMIB_ERR_ARG;
  const long M = (long)B * D * H * W;
  // wave tiles: 4 x 16 voxels when the grid is large, fewer when the
  // voxel count alone cannot fill the device; 64 channels per block
  // (32 / 16 for the narrow layers)
  const int wn = Cout % 64 == 0 ? 4 : (Cout % 32 == 0 ? 2 : 1);
  const long blocks64 = (M + 255) / 256 * (Cout / (16 * wn));
  if (blocks64 >= 512) {
    if (box_conv_ok(H, W, M, Cin, Cout)) {  // the input box staged in LDS
      if (wn == 4) return launch_fwd_box<4>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
      if (wn == 2) return launch_fwd_box<2>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
      return launch_fwd_box<1>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
    }
    if (wn == 4) return launch_fwd<4, 4>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, 1, s);
    if (wn == 2) return launch_fwd<4, 2>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, 1, s);
    return launch_fwd<4, 1>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, 1, s);
  }
  // mid-size grids with wide inputs: the input box with 16-channel tiles
  // (32^3 at batch 1, 64 -> 64 channels: 29.1 against 35.4 us split)
  if (M >= 32768 && Cin >= 64 && box_conv_ok(H, W, M, Cin, Cout))
    return launch_fwd_box<1>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
  // tiny grids: the whole padded sample in LDS, the reduction split by input slices
  if (vol_conv_ok(D, H, W, M, Cin, Cout))
    return launch_fwd_vol(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
  // small grids whose weights stay modest: the two waves of a block split
  // the reduction of one 32-voxel tile, one launch and no partial sums in
  // memory; the weights are read once per tile, at most 128 MB in all
  const double kw_bytes = (double)((M + 31) / 32) * Cout * 27 * Cin * 2;
  if (M >= 4096 && kw_bytes <= 128e6 && M * Cin * 2 < (1L << 31)) {
    if (wn == 4) return launch_fwd_kw<2, 4, 2, 2>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
    if (wn == 2) return launch_fwd_kw<2, 2, 2, 2>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
    return launch_fwd_kw<2, 1, 2, 2>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, s);
  }
  // small grids: split the reduction so that about 1024 blocks run, each
  // with at least 4 chunks of 32, within the workspace
  const long blocks = (M + 63) / 64 * (Cout / (16 * wn));
  const int nch = (27 * Cin + 31) / 32;
  long ks = (1024 + blocks - 1) / blocks;
  if (ks > nch / 4) ks = nch / 4;
  const long cap = (long)(LBMIB_SG_WORKSPACE_BYTES / ((size_t)M * Cout * sizeof(float)));
  if (ks > cap) ks = cap;
  if (ks < 1) ks = 1;
  if (wn == 4) return launch_fwd<1, 4>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, (int)ks, s);
  if (wn == 2) return launch_fwd<1, 2>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, (int)ks, s);
  return launch_fwd<1, 1>(x, w, bias, mask, y, D, H, W, M, Cin, Cout, relu, (int)ks, s);
}

// wt[ci][26 - t][co] = w[co][t][ci]
__global__ void k_flip(const uint16_t *__restrict__ w, uint16_t *__restrict__ wt, int Cin, int Cout) {
  const long n = (long)Cout * 27 * Cin;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int ci = (int)(e % Cin);
    const long r = e / Cin;
    const int t = (int)(r % 27), co = (int)(r / 27);
    wt[((long)ci * 27 + (26 - t)) * Cout + co] = w[e];
  }
}

// ----------------------------------------------------------- weight grad
//
// dw[co][t][ci] = sum_v dy[v][co] x[v + o(t)][ci]: per tap a GEMM whose
// reduction runs over voxels.  A block owns a 32 x 32 (co, ci) tile (16 for
// 16-channel tensors) and all 27 taps, spread over its 4 waves (taps
// w, w + 4, ...: 7 accumulator tiles per wave), and a contiguous range of
// 32-voxel chunks (the last one zero-filled past the end).  Per chunk it stages dy[32 voxels][co tile] and, for
// every tap, x[neighbour][ci tile] in LDS with natural (channel-fastest)
// 16-byte pieces; the MFMA operands need 8 consecutive voxels per lane, so
// lanes gather them with 2-byte LDS reads.  A row pitch of 36 elements
// (72 bytes) puts the four 8-voxel groups of a fragment on distinct banks.
constexpr int kWgPitch = 36;
constexpr int kWgChunk = 32;
constexpr int kWgBox = kWgChunk + 2;  // a chunk's voxels and one more on each side

__global__ void __launch_bounds__(256) k_conv_wgrad(const uint16_t *__restrict__ x, const uint16_t *__restrict__ dy,
                                                    float *__restrict__ part, int D, int H, int W, long M, int Cin,
                                                    int Cout, int ci_tiles, int chunks_per_split, long nchunks) {
  // the chunk's neighbourhood: for each (dz, dy), interval j = 3 (dz + 1) +
  // dy + 1 of kWgBox linear voxel indices starting one before the chunk's
  // first voxel shifted by dz HW + dy W; tap (dz, dy, dx) of voxel p is slot
  // p + 1 + dx of interval j wherever that neighbour lies in the volume
  __shared__ uint16_t sx[9 * kWgBox * kWgPitch];
  __shared__ uint16_t sdy[kWgChunk * kWgPitch];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col = lane & 15, kq = lane >> 4;
  const int ci0 = (blockIdx.x % ci_tiles) * 32, co0 = (blockIdx.x / ci_tiles) * 32;
  const int nci = Cin - ci0 >= 32 ? 2 : 1;  // 16-channel fragments in the tile
  const int nco = Cout - co0 >= 32 ? 2 : 1;
  const long HW = (long)H * W;
  const int K = 27 * Cin;
  f32x4 acc[7][2][2];
#pragma unroll
  for (int tt = 0; tt < 7; ++tt)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[tt][a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  const long c_begin = (long)blockIdx.y * chunks_per_split;
  long c_end = c_begin + chunks_per_split;
  if (c_end > nchunks) c_end = nchunks;
  const int lgxp = nci == 2 ? 2 : 1;  // log2 of the 8-channel pieces per voxel in the ci tile
  const int npx = 9 * kWgBox << lgxp;
  const int npy = kWgChunk * nco * 2;
  // the chunk's voxels: x, y, z within the sample, and 1 if the voxel exists
  __shared__ int svc[kWgChunk][4];
  constexpr int kXRounds = (9 * kWgBox * 4 + 255) / 256;  // x pieces per thread at most
  for (long c = c_begin; c < c_end; ++c) {
    const long v0 = c * kWgChunk;
    if (tid < kWgChunk) {
      const long v = v0 + tid;
      int cv[4] = {0, 0, 0, 0};
      if (v < M) {
        long rr = v;
        cv[0] = (int)(rr % W);
        rr /= W;
        cv[1] = (int)(rr % H);
        rr /= H;
        cv[2] = (int)(rr % D);
        cv[3] = 1;
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) svc[tid][a] = cv[a];
    }
    // every piece's load is issued before any is stored: an unconditional
    // load from a valid address (the tensor's start for an index outside
    // [0, M)), the value zeroed after it
    u32x4 px[kXRounds];
    bool pok[kXRounds];
#pragma unroll
    for (int r = 0; r < kXRounds; ++r) {
      const int e = tid + 256 * r;
      const uint16_t *src = x;
      bool ok = false;
      if (e < npx) {
        const int q = e & ((1 << lgxp) - 1), rest = e >> lgxp, sl = rest % kWgBox, j = rest / kWgBox;
        const long li = v0 - 1 + sl + (long)(j / 3 - 1) * HW + (long)(j % 3 - 1) * W;
        ok = li >= 0 && li < M;
        if (ok) src = x + li * Cin + ci0 + q * 8;
      }
      px[r] = *(const u32x4 *)src;
      pok[r] = ok;
    }
#pragma unroll
    for (int r = 0; r < kXRounds; ++r) {
      const int e = tid + 256 * r;
      if (e < npx) {
        const int q = e & ((1 << lgxp) - 1), rest = e >> lgxp;
        const u32x4 val = pok[r] ? px[r] : u32x4{0u, 0u, 0u, 0u};
        uint16_t *dst = sx + rest * kWgPitch + q * 8;
        *(u32x2 *)dst = u32x2{val[0], val[1]};
        *(u32x2 *)(dst + 4) = u32x2{val[2], val[3]};
      }
    }
    u32x4 py = {0u, 0u, 0u, 0u};
    bool yok = false;
    {
      const int q = tid % (nco * 2), p = tid / (nco * 2);
      const uint16_t *src = dy;
      if (tid < npy && v0 + p < M) {
        src = dy + (v0 + p) * Cout + co0 + q * 8;
        yok = true;
      }
      py = *(const u32x4 *)src;
    }
    if (tid < npy) {
      const int q = tid % (nco * 2), p = tid / (nco * 2);
      const u32x4 val = yok ? py : u32x4{0u, 0u, 0u, 0u};
      uint16_t *dst = sdy + p * kWgPitch + q * 8;
      *(u32x2 *)dst = u32x2{val[0], val[1]};
      *(u32x2 *)(dst + 4) = u32x2{val[2], val[3]};
    }
    __syncthreads();
    // which of this lane's voxels p = 8 kq + e have their neighbour at
    // offset d - 1 along x, y, z inside the volume (bit e); mz also carries
    // whether the voxel exists
    uint32_t mx[3] = {0u, 0u, 0u}, my[3] = {0u, 0u, 0u}, mz[3] = {0u, 0u, 0u};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int *cv = svc[kq * 8 + e];
      const int vx = cv[0], vy = cv[1], vz = cv[2], ex = cv[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        mx[d] |= (uint32_t)((unsigned)(vx + d - 1) < (unsigned)W) << e;
        my[d] |= (uint32_t)((unsigned)(vy + d - 1) < (unsigned)H) << e;
        mz[d] |= (uint32_t)(ex && (unsigned)(vz + d - 1) < (unsigned)D) << e;
      }
    }
    // A: dy^T fragments (row co = 16 a + col, voxels 8 kq .. + 8)
    bf16x8 fa[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      uint16_t h[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) h[e] = sdy[(kq * 8 + e) * kWgPitch + a * 16 + col];
      u32x4 pk;
#pragma unroll
      for (int e = 0; e < 4; ++e) pk[e] = (uint32_t)h[2 * e] | ((uint32_t)h[2 * e + 1] << 16);
      fa[a] = __builtin_bit_cast(bf16x8, pk);
    }
#pragma unroll
    for (int tt = 0; tt < 7; ++tt) {
      const int t = wave + 4 * tt;
      if (t >= 27) break;
      const int tz = t / 9, ty = (t / 3) % 3, tx = t % 3;
      const uint32_t vm = mx[tx] & my[ty] & mz[tz];
      const uint16_t *row = sx + ((tz * 3 + ty) * kWgBox + kq * 8 + tx) * kWgPitch;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (b >= nci) break;
        uint16_t h[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const uint16_t v = row[e * kWgPitch + b * 16 + col];
          h[e] = (vm >> e) & 1u ? v : (uint16_t)0;
        }
        u32x4 pk;
#pragma unroll
        for (int e = 0; e < 4; ++e) pk[e] = (uint32_t)h[2 * e] | ((uint32_t)h[2 * e + 1] << 16);
        const bf16x8 fb = __builtin_bit_cast(bf16x8, pk);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          if (a >= nco) break;
          acc[tt][a][b] = mfma16(fa[a], fb, acc[tt][a][b]);
        }
      }
    }
    __syncthreads();
  }
  // partials: lane holds (co = co0 + 16 a + 4 kq + r, ci = ci0 + 16 b + col)
  float *dst = part + (long)blockIdx.y * Cout * K;
#pragma unroll
  for (int tt = 0; tt < 7; ++tt) {
    const int t = wave + 4 * tt;
    if (t >= 27) break;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      if (a >= nco) break;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (b >= nci) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = co0 + a * 16 + 4 * kq + r, ci = ci0 + b * 16 + col;
          dst[(long)co * K + t * Cin + ci] = acc[tt][a][b][r];
        }
      }
    }
  }
}

// out[e] = sum over s < ns of part[s][e], in order of s
// out[e] = sum over k < ns of part[k][e], in a fixed order.  A block covers
// 64 consecutive e (one per lane, so every wave reads 256 contiguous bytes
// per k) and its 4 waves split k: wave g sums k = g, g + 8, g + 16, ... and
// k = g + 4, g + 12, ... in two running sums; the four waves' sums are then
// added through LDS as (w0 + w1) + (w2 + w3).  Launch with kSumBlock
// threads and ceil(n / 64) blocks.
constexpr int kSumBlock = 256;
__global__ void __launch_bounds__(kSumBlock) k_sum_parts(const float *__restrict__ part, float *__restrict__ out,
                                                         long n, int ns) {
  __shared__ float red[4][64];
  const int el = threadIdx.x & 63, g = threadIdx.x >> 6;
  const long e = (long)blockIdx.x * 64 + el;
  float s0 = 0.f, s1 = 0.f;
  if (e < n) {
    int k = g;
    for (; k + 4 < ns; k += 8) {
      s0 += part[(long)k * n + e];
      s1 += part[(long)(k + 4) * n + e];
    }
    if (k < ns) s0 += part[(long)k * n + e];
  }
  red[g][el] = s0 + s1;
  __syncthreads();
  if (g == 0 && e < n) out[e] = (red[0][el] + red[1][el]) + (red[2][el] + red[3][el]);
}

// bias gradient partials: part[blk][co] = sum of dy[v][co] over the block's
// voxel range.  A thread sums 8 channels of every G-th voxel (16-byte
// loads; G = 256 / (C / 8) voxel lanes), then the G lanes of each channel
// group are added through LDS in lane order.  C is a multiple of 16 and at
// most 512.
__global__ void __launch_bounds__(256) k_colsum(const uint16_t *__restrict__ dy, float *__restrict__ part, long M,
                                                int C, long per_block) {
  __shared__ float red[256][9];
  const int ng = C / 8, G = 256 / ng;  // channel groups, voxel lanes
  const int tid = threadIdx.x, cg = tid % ng, vl = tid / ng;
  const long v0 = (long)blockIdx.x * per_block;
  long v1 = v0 + per_block;
  if (v1 > M) v1 = M;
  for (int c0 = 0; c0 < C; c0 += 8 * ng) {  // one pass: 8 ng = C
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (vl < G)
      for (long v = v0 + vl; v < v1; v += G) {
        const u32x4 q = *(const u32x4 *)(dy + v * C + c0 + 8 * cg);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          s[2 * r] += bf2f((uint16_t)(q[r] & 0xffff));
          s[2 * r + 1] += bf2f((uint16_t)(q[r] >> 16));
        }
      }
#pragma unroll
    for (int r = 0; r < 8; ++r) red[tid][r] = s[r];
    __syncthreads();
    for (int ch = tid; ch < 8 * ng; ch += 256) {
      const int g = ch / 8, r = ch % 8;
      float t = 0.f;
      for (int k = 0; k < G; ++k) t += red[k * ng + g][r];
      part[(long)blockIdx.x * C + c0 + ch] = t;
    }
    __syncthreads();
  }
}

// Split of the voxel reduction: chunks per split, chosen from the shape only
// (so every entry point that shares it gives the same bits), capped so that
// the partials fit in `cap_bytes`.
struct WgradPlan {
  int ci_tiles, co_tiles, splits, cps;
  long nchunks;
  size_t part_bytes;
};

WgradPlan wgrad_plan(long M, int Cin, int Cout, size_t cap_bytes) {
  WgradPlan p;
  p.ci_tiles = (Cin + 31) / 32;
  p.co_tiles = (Cout + 31) / 32;
  p.nchunks = (M + kWgChunk - 1) / kWgChunk;
  const int tiles = p.ci_tiles * p.co_tiles;
  const size_t per_split = (size_t)Cout * 27 * Cin * sizeof(float);
  long want = (512 + tiles - 1) / tiles;  // about 512 blocks: one round at 2 per CU
  if (want > p.nchunks) want = p.nchunks;
  long cap = (long)(cap_bytes / per_split);
  if (want > cap) want = cap;
  if (want < 1) want = 1;
  p.cps = (int)((p.nchunks + want - 1) / want);
  p.splits = (int)((p.nchunks + p.cps - 1) / p.cps);
  p.part_bytes = per_split * p.splits;
  return p;
}

constexpr int kColsumBlocks = 1024;

int wgrad_with(const void *x, const void *dy, float *dw, float *db, int B, int D, int H, int W, int Cin, int Cout,
               void *ws, size_t ws_bytes, hipStream_t s) {
  const long M = (long)B * D * H * W;
  const WgradPlan p = wgrad_plan(M, Cin, Cout, ws_bytes);
  if (p.part_bytes > ws_bytes || p.nchunks < 1) return LBMIB_ERR_ARG;
  float *part = (float *)ws;
  k_conv_wgrad<<<dim3(p.ci_tiles * p.co_tiles, p.splits), 256, 0, s>>>(
      (const uint16_t *)x, (const uint16_t *)dy, part, D, H, W, M, Cin, Cout, p.ci_tiles, p.cps, p.nchunks);
  HIPCHK(hipGetLastError());
  const long n = (long)Cout * 27 * Cin;
  k_sum_parts<<<(unsigned)((n + 63) / 64), kSumBlock, 0, s>>>(part, dw, n, p.splits);
  HIPCHK(hipGetLastError());
  if (db) {
    if ((size_t)kColsumBlocks * Cout * sizeof(float) > ws_bytes) return LBMIB_ERR_ARG;
    const long per = (M + kColsumBlocks - 1) / kColsumBlocks;
    const int nb = (int)((M + per - 1) / per);
    k_colsum<<<nb, 256, 0, s>>>((const uint16_t *)dy, part, M, Cout, per);
    HIPCHK(hipGetLastError());
    k_sum_parts<<<(unsigned)((Cout + 63) / 64), kSumBlock, 0, s>>>(part, db, Cout, nb);
    HIPCHK(hipGetLastError());
  }
  return LBMIB_OK;
}

// The library's workspace: one buffer per device, allocated on first use.
constexpr int kMaxDevices = 64;
void *g_ws[kMaxDevices] = {};
std::mutex g_ws_mu;

// Scratch of lbmib_sg_conv3d_bwd_weight_noalloc: a file-scope device array
// that every GPU the library runs on holds until the process ends.  In host
// code its name is not a device address; each device's own copy is reached
// through the address hipGetSymbolAddress returns with that device current,
// kept per device in g_scratch.
static __device__ __align__(256) float scratch[LBMIB_SG_NOALLOC_SCRATCH_BYTES / sizeof(float)];
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_ptr(void **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return LBMIB_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LBMIB_ERR_HIP;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return LBMIB_OK;
}

int workspace(void **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return LBMIB_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  if (!g_ws[dev]) {
    void *p = nullptr;
    if (hipMalloc(&p, LBMIB_SG_WORKSPACE_BYTES) != hipSuccess) return LBMIB_ERR_ALLOC;
    g_ws[dev] = p;
  }
  *out = g_ws[dev];
  return LBMIB_OK;
}

// lbmib_sg_sqdiff's block partials (two doubles for each of at most
// kSqBlocks blocks): a buffer of their own, one per device, so that the
// sums a caller takes on one stream never share memory with a network
// running on another stream through the workspace.
constexpr int kSqBlocks = 1024;
void *g_sq[kMaxDevices] = {};

int sq_partials(void **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return LBMIB_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  if (!g_sq[dev]) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(double) * 2 * kSqBlocks) != hipSuccess) return LBMIB_ERR_ALLOC;
    g_sq[dev] = p;
  }
  *out = g_sq[dev];
  return LBMIB_OK;
}

// ------------------------------------------------------------ elementwise
//
// One thread per (voxel, 8-channel piece); C is a multiple of 16.

__global__ void k_pool2(const uint16_t *__restrict__ x, uint16_t *__restrict__ y, int D, int H, int W, int C, long n) {
  const int P = C / 8;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int q = (int)(e % P);
    long r = e / P;
    const int ox = (int)(r % W);
    r /= W;
    const int oy = (int)(r % H);
    r /= H;
    const int oz = (int)(r % D);
    const long b = r / D;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const long iz = 2 * oz + (c >> 2), iy = 2 * oy + ((c >> 1) & 1), ix = 2 * ox + (c & 1);
      const long v = ((b * 2 * D + iz) * 2 * H + iy) * 2 * W + ix;
      const u32x4 val = *(const u32x4 *)(x + v * C + q * 8);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        acc[2 * k] += bf2f((uint16_t)(val[k] & 0xffff));
        acc[2 * k + 1] += bf2f((uint16_t)(val[k] >> 16));
      }
    }
    u32x4 o;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      o[k] = (uint32_t)f2bf(acc[2 * k] * 0.125f) | ((uint32_t)f2bf(acc[2 * k + 1] * 0.125f) << 16);
    *(u32x4 *)(y + e * 8) = o;
  }
}

__device__ __forceinline__ void unpack8(u32x4 v, float f[8]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[2 * k] = bf2f((uint16_t)(v[k] & 0xffff));
    f[2 * k + 1] = bf2f((uint16_t)(v[k] >> 16));
  }
}
__device__ __forceinline__ u32x4 pack8(const float f[8]) {
  u32x4 o;
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
  return o;
}

// voxel index of the parent (D, H, W fine dims) of fine voxel v
__device__ __forceinline__ long parent_of(long v, int D, int H, int W) {
  const int x = (int)(v % W);
  long r = v / W;
  const int y = (int)(r % H);
  r /= H;
  const int z = (int)(r % D);
  const long b = r / D;
  return ((b * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2;
}

__global__ void k_up2_add(const uint16_t *__restrict__ c, const uint16_t *__restrict__ e, uint16_t *__restrict__ s,
                          int D, int H, int W, int C, long n) {
  const int P = C / 8;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int q = (int)(i % P);
    const long v = i / P;
    const long u = parent_of(v, D, H, W);
    float a[8], b[8];
    unpack8(*(const u32x4 *)(c + u * C + q * 8), a);
    if (e) {
      unpack8(*(const u32x4 *)(e + i * 8), b);
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] += b[k];
    }
    *(u32x4 *)(s + i * 8) = pack8(a);
  }
}

__global__ void k_pool2_bwd(const uint16_t *__restrict__ a, const uint16_t *__restrict__ g,
                            const uint16_t *__restrict__ mask, uint16_t *__restrict__ out, int D, int H, int W, int C,
                            long n) {
  const int P = C / 8;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int q = (int)(i % P);
    const long v = i / P;
    const long u = parent_of(v, D, H, W);
    float f[8], t[8];
    unpack8(*(const u32x4 *)(g + u * C + q * 8), f);
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] *= 0.125f;
    if (a) {
      unpack8(*(const u32x4 *)(a + i * 8), t);
#pragma unroll
      for (int k = 0; k < 8; ++k) f[k] += t[k];
    }
    if (mask) {
      unpack8(*(const u32x4 *)(mask + i * 8), t);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (!(t[k] > 0.f)) f[k] = 0.f;
    }
    *(u32x4 *)(out + i * 8) = pack8(f);
  }
}

__global__ void k_up2_bwd(const uint16_t *__restrict__ ds, const uint16_t *__restrict__ mask,
                          uint16_t *__restrict__ out, int D, int H, int W, int C, long n) {
  const int P = C / 8;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int q = (int)(i % P);
    long r = i / P;
    const int ox = (int)(r % W);
    r /= W;
    const int oy = (int)(r % H);
    r /= H;
    const int oz = (int)(r % D);
    const long b = r / D;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, t[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const long iz = 2 * oz + (c >> 2), iy = 2 * oy + ((c >> 1) & 1), ix = 2 * ox + (c & 1);
      const long v = ((b * 2 * D + iz) * 2 * H + iy) * 2 * W + ix;
      unpack8(*(const u32x4 *)(ds + v * C + q * 8), t);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += t[k];
    }
    if (mask) {
      unpack8(*(const u32x4 *)(mask + i * 8), t);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (!(t[k] > 0.f)) acc[k] = 0.f;
    }
    *(u32x4 *)(out + i * 8) = pack8(acc);
  }
}

// loss partials (per block) and the gradient
__global__ void __launch_bounds__(256) k_mse(const uint16_t *__restrict__ out, const float *__restrict__ tgt,
                                             uint16_t *__restrict__ dout, float *__restrict__ part, long M, int C,
                                             int nl, float gscale) {
  __shared__ float red[256];
  float s = 0.f;
  const long n = M * C;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int c = (int)(e % C);
    const long v = e / C;
    float gval = 0.f;
    if (c < nl) {
      const float d = bf2f(out[e]) - tgt[v * nl + c];
      s += d * d;
      gval = gscale * d;
    }
    dout[e] = f2bf(gval);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_final_sum(const float *__restrict__ part, int n, float scale,
                                                   float *__restrict__ out) {
  __shared__ float r[256];
  float s = 0.f;
  for (int k = threadIdx.x; k < n; k += 256) s += part[k];
  r[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = r[0] * scale;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       uint16_t *__restrict__ w16, long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float gr = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gr;
    const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    const float pn = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    p[e] = pn;
    if (w16) w16[e] = f2bf(pn);
  }
}

__global__ void k_cast(const float *__restrict__ src, uint16_t *__restrict__ dst, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    dst[e] = f2bf(src[e]);
}

struct Norm16 {
  float a[16], b[16];
};

__global__ void k_pack(const float *__restrict__ f, int nf, Norm16 nm, int nconst, uint16_t *__restrict__ x, long M,
                       int C) {
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < M; v += (long)gridDim.x * blockDim.x) {
    for (int c0 = 0; c0 < C; c0 += 8) {
      float t[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int c = c0 + k;
        float val = 0.f;
        if (c < nf)
          val = (f[(long)c * M + v] - nm.a[c]) * nm.b[c];
        else if (c < nf + nconst)
          val = nm.a[c];
        t[k] = val;
      }
      *(u32x4 *)(x + v * C + c0) = pack8(t);
    }
  }
}

__global__ void k_target(const float *__restrict__ f0, const float *__restrict__ f1, int nl, Norm16 nm,
                         float *__restrict__ tgt, long M) {
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < M; v += (long)gridDim.x * blockDim.x)
    for (int c = 0; c < nl; ++c) {
      const long i = (long)c * M + v;
      tgt[v * nl + c] = ((f1[i] - f0[i]) - nm.a[c]) * nm.b[c];
    }
}

__global__ void k_unpack(const uint16_t *__restrict__ out, int C, int nl, Norm16 nm, float *__restrict__ d, long M) {
  for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < M; v += (long)gridDim.x * blockDim.x)
    for (int c = 0; c < nl; ++c) d[(long)c * M + v] = bf2f(out[v * C + c]) * nm.b[c] + nm.a[c];
}

__global__ void __launch_bounds__(256) k_sqdiff(const float *__restrict__ a, const float *__restrict__ b, long n,
                                                double *__restrict__ part) {
  __shared__ double r0[256], r1[256];
  double s0 = 0.0, s1 = 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const double av = a[e];
    if (b) {
      const double bv = b[e];
      s0 += (av - bv) * (av - bv);
      s1 += bv * bv;
    } else {
      s0 += av * av;
    }
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r0[threadIdx.x] += r0[threadIdx.x + w];
      r1[threadIdx.x] += r1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r0[0];
    part[2 * blockIdx.x + 1] = r1[0];
  }
}

// one block of 256: strided partial sums, then a fixed tree (deterministic)
__global__ void __launch_bounds__(256) k_sqdiff_final(const double *__restrict__ part, int nb,
                                                      double *__restrict__ out2) {
  __shared__ double r0[256], r1[256];
  double s0 = 0.0, s1 = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) {
    s0 += part[2 * k];
    s1 += part[2 * k + 1];
  }
  r0[threadIdx.x] = s0;
  r1[threadIdx.x] = s1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r0[threadIdx.x] += r0[threadIdx.x + w];
      r1[threadIdx.x] += r1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out2[0] = r0[0];
    out2[1] = r1[0];
  }
}

unsigned grid_for(long n, int per_block = 256, long cap = 8192) {
  long g = (n + per_block - 1) / per_block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}

bool norm_from(const double *a, const double *b, int n, Norm16 &nm) {
  if (n > 16) return false;
  for (int k = 0; k < 16; ++k) {
    nm.a[k] = (k < n && a) ? (float)a[k] : 0.f;
    nm.b[k] = (k < n && b) ? (float)b[k] : 0.f;
  }
  return true;
}

}  // namespace

// ------------------------------------------------------------------ C-ABI

extern "C" int lbmib_sg_conv3d_fwd(const void *x, const void *w, const float *bias, const void *mask, void *y, int B,
                                   int D, int H, int W, int Cin, int Cout, int relu, void *stream) {
  return conv_fwd(x, w, bias, mask, y, B, D, H, W, Cin, Cout, relu, (hipStream_t)stream);
}

extern "C" int lbmib_sg_conv3d_bwd_data(const void *dy, const void *w, void *wt, const void *mask, void *dx, int B,
                                        int D, int H, int W, int Cin, int Cout, void *stream) {
  if (!dy || !w || !wt || !dx) return LBMIB_ERR_ARG;
  // the data gradient convolves dy (Cout channels) with wt into Cin channels
  if (!shape_ok(B, D, H, W, Cout, Cin) || Cin % 16) return LBMIB_ERR_ARG;
  if (Cout & (Cout - 1)) return LBMIB_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  const long n = (long)Cout * 27 * Cin;
  k_flip<<<grid_for(n), 256, 0, s>>>((const uint16_t *)w, (uint16_t *)wt, Cin, Cout);
  HIPCHK(hipGetLastError());
  return conv_fwd(dy, wt, nullptr, mask, dx, B, D, H, W, Cout, Cin, 0, s);
}

extern "C" int lbmib_sg_conv3d_bwd_weight(const void *x, const void *dy, float *dw, float *db, int B, int D, int H,
                                          int W, int Cin, int Cout, void *stream) {
  if (!x || !dy || !dw || !shape_ok(B, D, H, W, Cin, Cout)) return LBMIB_ERR_ARG;
  void *ws = nullptr;
  const int rc = workspace(&ws);
  if (rc) return rc;
  return wgrad_with(x, dy, dw, db, B, D, H, W, Cin, Cout, ws, LBMIB_SG_WORKSPACE_BYTES, (hipStream_t)stream);
}

extern "C" int lbmib_sg_conv3d_bwd_weight_noalloc(const void *x, const void *dy, float *dw, float *db, int B, int D,
                                                  int H, int W, int Cin, int Cout, void *stream) {
  if (!x || !dy || !dw || !shape_ok(B, D, H, W, Cin, Cout)) return LBMIB_ERR_ARG;
  void *ws = nullptr;
  const int rc = scratch_ptr(&ws);
  if (rc) return rc;
  // the same split plan as lbmib_sg_conv3d_bwd_weight (same cap), so the
  // partials and their fixed-order sums are the same bits; every partial is
  // written before it is read, so nothing needs clearing between calls
  return wgrad_with(x, dy, dw, db, B, D, H, W, Cin, Cout, ws, LBMIB_SG_NOALLOC_SCRATCH_BYTES, (hipStream_t)stream);
}

extern "C" int lbmib_sg_pool2(const void *x, void *y, int B, int D, int H, int W, int C, void *stream) {
  if (!x || !y || B < 1 || D < 1 || H < 1 || W < 1 || C < 8 || C % 8) return LBMIB_ERR_ARG;
  const long n = (long)B * D * H * W * (C / 8);
  k_pool2<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const uint16_t *)x, (uint16_t *)y, D, H, W, C, n);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_up2_add(const void *c, const void *e, void *s, int B, int D, int H, int W, int C,
                                void *stream) {
  if (!c || !s || B < 1 || D < 2 || H < 2 || W < 2 || (D | H | W) & 1 || C < 8 || C % 8) return LBMIB_ERR_ARG;
  const long n = (long)B * D * H * W * (C / 8);
  k_up2_add<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const uint16_t *)c, (const uint16_t *)e, (uint16_t *)s, D,
                                                          H, W, C, n);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_pool2_bwd(const void *a, const void *g, const void *mask, void *out, int B, int D, int H, int W,
                                  int C, void *stream) {
  if (!g || !out || B < 1 || D < 2 || H < 2 || W < 2 || (D | H | W) & 1 || C < 8 || C % 8) return LBMIB_ERR_ARG;
  const long n = (long)B * D * H * W * (C / 8);
  k_pool2_bwd<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const uint16_t *)a, (const uint16_t *)g,
                                                            (const uint16_t *)mask, (uint16_t *)out, D, H, W, C, n);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_up2_bwd(const void *ds, const void *mask, void *out, int B, int D, int H, int W, int C,
                                void *stream) {
  if (!ds || !out || B < 1 || D < 1 || H < 1 || W < 1 || C < 8 || C % 8) return LBMIB_ERR_ARG;
  const long n = (long)B * D * H * W * (C / 8);
  k_up2_bwd<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const uint16_t *)ds, (const uint16_t *)mask,
                                                          (uint16_t *)out, D, H, W, C, n);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_mse(const void *out, const float *tgt, void *dout, float *loss, long M, int C, int nl,
                            void *stream) {
  if (!out || !tgt || !dout || !loss || M < 1 || C < 1 || nl < 1 || nl > C) return LBMIB_ERR_ARG;
  void *ws = nullptr;
  const int rc = workspace(&ws);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = grid_for(M * C, 256, 1024);
  const float denom = (float)((double)M * nl);
  k_mse<<<nb, 256, 0, s>>>((const uint16_t *)out, tgt, (uint16_t *)dout, (float *)ws, M, C, nl, 2.0f / denom);
  HIPCHK(hipGetLastError());
  k_final_sum<<<1, 256, 0, s>>>((const float *)ws, (int)nb, 1.0f / denom, loss);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_adam(float *p, const float *g, float *m, float *v, void *w16, long n, float lr, float beta1,
                             float beta2, float eps, int t, void *stream) {
  if (!p || !g || !m || !v || n < 1 || t < 1) return LBMIB_ERR_ARG;
  const float c1 = 1.f - powf(beta1, (float)t), c2 = 1.f - powf(beta2, (float)t);
  k_adam<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(p, g, m, v, (uint16_t *)w16, n, lr, beta1, beta2, eps, c1, c2);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_cast_bf16(const float *src, void *dst, long n, void *stream) {
  if (!src || !dst || n < 1) return LBMIB_ERR_ARG;
  k_cast<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(src, (uint16_t *)dst, n);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_pack(const float *f, int nf, const double *mean, const double *istd, int nconst,
                             const double *cval, void *x, long M, int C, void *stream) {
  if (!x || M < 1 || C % 8 || nf < 0 || nconst < 0 || nf + nconst > C || nf + nconst > 16 || (nf && !f))
    return LBMIB_ERR_ARG;
  Norm16 nm;
  for (int k = 0; k < 16; ++k) nm.a[k] = nm.b[k] = 0.f;
  for (int k = 0; k < nf; ++k) {
    nm.a[k] = mean ? (float)mean[k] : 0.f;
    nm.b[k] = istd ? (float)istd[k] : 1.f;
  }
  for (int k = 0; k < nconst; ++k) nm.a[nf + k] = cval ? (float)cval[k] : 0.f;
  k_pack<<<grid_for(M), 256, 0, (hipStream_t)stream>>>(f, nf, nm, nconst, (uint16_t *)x, M, C);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_target(const float *f0, const float *f1, int nl, const double *mean, const double *istd,
                               float *tgt, long M, void *stream) {
  Norm16 nm;
  if (!f0 || !f1 || !tgt || nl < 1 || M < 1 || !norm_from(mean, istd, nl, nm)) return LBMIB_ERR_ARG;
  k_target<<<grid_for(M), 256, 0, (hipStream_t)stream>>>(f0, f1, nl, nm, tgt, M);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_unpack(const void *out, int C, int nl, const double *mean, const double *std, float *d, long M,
                               void *stream) {
  Norm16 nm;
  if (!out || !d || nl < 1 || nl > C || M < 1 || !norm_from(mean, std, nl, nm)) return LBMIB_ERR_ARG;
  k_unpack<<<grid_for(M), 256, 0, (hipStream_t)stream>>>((const uint16_t *)out, C, nl, nm, d, M);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}

extern "C" int lbmib_sg_sqdiff(const float *a, const float *b, long n, double *out2, void *stream) {
  if (!a || !out2 || n < 1) return LBMIB_ERR_ARG;
  void *ws = nullptr;
  const int rc = sq_partials(&ws);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const unsigned nb = grid_for(n, 256, kSqBlocks);
  k_sqdiff<<<nb, 256, 0, s>>>(a, b, n, (double *)ws);
  HIPCHK(hipGetLastError());
  k_sqdiff_final<<<1, 256, 0, s>>>((const double *)ws, (int)nb, out2);
  HIPCHK(hipGetLastError());
  return LBMIB_OK;
}
