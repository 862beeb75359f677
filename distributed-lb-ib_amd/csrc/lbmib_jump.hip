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

// ------------------------------------------------------------ coarse moments
//
// Block: 64 x 4 fine cells of one (x, y) tile; each thread walks the F cells
// of its column in z, so a block covers (64/F) x (4/F) coarse cells of one
// coarse z-plane.  Per fine cell the moments are those of
// lbmib_get_macroscopic (post-stream populations, u = (sum c f + F/2) / rho,
// the initial rho0/u0 outside the moment box); the block sums F^3 of them
// in a fixed order (lane shuffles in x, LDS in y).
template <int P, int F>
__global__ void __launch_bounds__(256)
    k_coarse_moments(Geometry g, const typename Traits<P>::S *__restrict__ f, int own,
                     const typename Traits<P>::T *__restrict__ tab, const double *__restrict__ ref,
                     typename Traits<P>::T b1, typename Traits<P>::T b2, double rho0, double u0x, double u0y,
                     double u0z, double gz, int nsheets, const int *__restrict__ box,
                     const typename Traits<P>::FT *__restrict__ Fc, float *__restrict__ out) {
  using T = typename Traits<P>::T;
  __shared__ double red[4][4][64];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x = blockIdx.x * 64 + tx, y = blockIdx.y * 4 + ty;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < F; ++k) {
    const int z = blockIdx.z * F + k;
    const bool act = x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2];
    const bool mom =
        x >= g.mlo[0] && x <= g.mhi[0] && y >= g.mlo[1] && y <= g.mhi[1] && z >= g.mlo[2] && z <= g.mhi[2];
    double r = rho0, ux = u0x, uy = u0y, uz = u0z;
    if (act && mom && !own) {
      T v[19];
      const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
      if (frozen) {
        for (int q = 0; q < 19; ++q) v[q] = tab[q] * Traits<P>::inv_scale;
      } else {
        pull<P>(g, f, x, y, z, b1, b2, v);
      }
      const long c0 = (long)x + (long)g.nx * ((long)y + (long)g.ny * z);
      double fx = 0, fy = 0, fz = 0;
      if (in_any_box(box, nsheets, x, y, z)) { fx = Fc[c0]; fy = Fc[g.n + c0]; fz = Fc[2 * g.n + c0]; }
      double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
#pragma unroll
      for (int q = 0; q < 19; ++q) {
        const double fq = Traits<P>::dev ? (double)v[q] + ref[q] : (double)v[q];
        s1 += fq;
        s2 += kCX[q] * fq;
        s3 += kCY[q] * fq;
        s4 += kCZ[q] * fq;
      }
      r = s1;
      ux = (s2 + 0.5 * fx) / s1;
      uy = (s3 + 0.5 * fy) / s1;
      uz = (s4 + 0.5 * (fz + r * gz)) / s1;
    }
    acc[0] += r;
    acc[1] += ux;
    acc[2] += uy;
    acc[3] += uz;
  }
#pragma unroll
  for (int m = 1; m < F; m <<= 1)
#pragma unroll
    for (int a = 0; a < 4; ++a) acc[a] += __shfl_xor(acc[a], m);
#pragma unroll
  for (int a = 0; a < 4; ++a) red[a][ty][tx] = acc[a];
  __syncthreads();
  if (ty % F == 0 && tx % F == 0) {
    const int ncx = g.nx / F, ncy = g.ny / F, ncz = g.nz / F;
    const long nc = (long)ncx * ncy * ncz;
    const long cc = (long)(x / F) + (long)ncx * ((long)(y / F) + (long)ncy * blockIdx.z);
    const double inv = 1.0 / (F * F * F);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      double s = 0.0;
      for (int k = 0; k < F; ++k) s += red[a][ty + k][tx];
      out[a * nc + cc] = (float)(s * inv);
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

// feq_q(rho, u) of init_eqlbrmdistrfuncDF0 (lbm_ib.c:1561-1594)
template <class T>
__device__ __forceinline__ T feq_q(int q, T rho, T ux, T uy, T uz) {
  const T usq = ux * ux + uy * uy + uz * uz;
  const T cu = (T)kCX[q] * ux + (T)kCY[q] * uy + (T)kCZ[q] * uz;
  const T w = (T)1 / (T)kWDEN[q];
  return w * rho * ((T)1 + (T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * usq);
}

// One cell per thread (x across the block, y and z from the grid): the
// updated cells -- the active box without the frozen inlet/outlet planes --
// take the equilibrium shift; every other element is copied unchanged.
template <int P>
__global__ void k_jump_apply(Geometry g, const typename Traits<P>::S *__restrict__ fin,
                             typename Traits<P>::S *__restrict__ fout, const double *__restrict__ ref,
                             const float *__restrict__ delta, int F) {
  using T = typename Traits<P>::T;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= g.nx) return;
  const long p0 = pop_index(g, x, y, z);
  const bool act = x >= g.lo[0] && x <= g.hi[0] && y >= g.lo[1] && y <= g.hi[1] && z >= g.lo[2] && z <= g.hi[2];
  const bool frozen = g.bc[0] == LBMIB_BC_INOUT && (x == g.lo[0] || x == g.hi[0]);
  if (!act || frozen) {
#pragma unroll
    for (int q = 0; q < 19; ++q) fout[q * g.pq + p0] = fin[q * g.pq + p0];
    return;
  }
  T h[19];
  T rho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    h[q] = ld_h<P>(fin + q * g.pq + p0);
    const T fq = Traits<P>::dev ? h[q] * (T)Traits<P>::inv_scale + (T)ref[q] : h[q];
    rho += fq;
    jx += (T)kCX[q] * fq;
    jy += (T)kCY[q] * fq;
    jz += (T)kCZ[q] * fq;
  }
  const T ux = jx / rho, uy = jy / rho, uz = jz / rho;
  const int ncx = g.nx / F, ncy = g.ny / F, ncz = g.nz / F;
  const long nc = (long)ncx * ncy * ncz;
  const T rn = rho + (T)interp(delta, 0, nc, ncx, ncy, ncz, F, x, y, z);
  const T vx = ux + (T)interp(delta, 1, nc, ncx, ncy, ncz, F, x, y, z);
  const T vy = uy + (T)interp(delta, 2, nc, ncx, ncy, ncz, F, x, y, z);
  const T vz = uz + (T)interp(delta, 3, nc, ncx, ncy, ncz, F, x, y, z);
#pragma unroll
  for (int q = 0; q < 19; ++q) {
    const T d = feq_q<T>(q, rn, vx, vy, vz) - feq_q<T>(q, rho, ux, uy, uz);
    st_h<P>(fout + q * g.pq + p0, h[q] + d * (T)Traits<P>::scale);
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

template <int P>
int coarse_launch(lbmib_ctx *ctx, int F, float *out) {
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
  dim3 grid(g.nx / 64, g.ny / 4, g.nz / F);
  const auto *fp = (const typename Traits<P>::S *)ctx->f[ctx->cur];
  const T *tab = (const T *)ctx->c.dtab;
  const FluidConsts<double> &c = ctx->c;
  switch (F) {
    case 1: k_coarse_moments<P, 1><<<grid, 256, 0, ctx->stream>>>(g, fp, ctx->own, tab, ctx->d_ref, b1, b2, c.rho0, c.u0x, c.u0y, c.u0z, c.g, nsh, box, Fc, out); break;
    case 2: k_coarse_moments<P, 2><<<grid, 256, 0, ctx->stream>>>(g, fp, ctx->own, tab, ctx->d_ref, b1, b2, c.rho0, c.u0x, c.u0y, c.u0z, c.g, nsh, box, Fc, out); break;
    default: k_coarse_moments<P, 4><<<grid, 256, 0, ctx->stream>>>(g, fp, ctx->own, tab, ctx->d_ref, b1, b2, c.rho0, c.u0x, c.u0y, c.u0z, c.g, nsh, box, Fc, out); break;
  }
  JHIPCHK(hipGetLastError());
  return LBMIB_OK;
}

template <int P>
int apply_launch(lbmib_ctx *ctx, const void *fin, void *fout, const float *delta, int F) {
  const Geometry &g = ctx->g;
  const int bx = g.nx >= 256 ? 256 : 64;
  dim3 grid((g.nx + bx - 1) / bx, g.ny, g.nz);
  k_jump_apply<P><<<grid, bx, 0, ctx->stream>>>(g, (const typename Traits<P>::S *)fin, (typename Traits<P>::S *)fout,
                                                 ctx->d_ref, delta, F);
  JHIPCHK(hipGetLastError());
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

extern "C" int lbmib_coarse_moments(lbmib_ctx *ctx, int factor, float *d_out) {
  if (!ctx || !d_out) return LBMIB_ERR_ARG;
  JRCHK(check_coarse(ctx, factor));
  if (ctx->init_bits != 7 || ctx->pending) return LBMIB_ERR_STATE;
  switch (ctx->fp.precision) {
    case LBMIB_F64: return coarse_launch<LBMIB_F64>(ctx, factor, d_out);
    case LBMIB_F32: return coarse_launch<LBMIB_F32>(ctx, factor, d_out);
    case LBMIB_F16: return coarse_launch<LBMIB_F16>(ctx, factor, d_out);
    default: return coarse_launch<LBMIB_BF16>(ctx, factor, d_out);
  }
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

extern "C" int lbmib_jump(lbmib_ctx *ctx, int factor, const float *d_coarse, const float *d_delta, int ksteps) {
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
    case LBMIB_F64: rc = apply_launch<LBMIB_F64>(ctx, old, ctx->f[ctx->cur], d_delta, factor); break;
    case LBMIB_F32: rc = apply_launch<LBMIB_F32>(ctx, old, ctx->f[ctx->cur], d_delta, factor); break;
    case LBMIB_F16: rc = apply_launch<LBMIB_F16>(ctx, old, ctx->f[ctx->cur], d_delta, factor); break;
    default: rc = apply_launch<LBMIB_BF16>(ctx, old, ctx->f[ctx->cur], d_delta, factor); break;
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
