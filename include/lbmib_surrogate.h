/*
 * lbmib_surrogate.h -- C-ABI of the learned surrogate's kernels (liblbmib.so).
 *
 * Not in the reference: cube_pthread/lbm_ib.c has no learned model.  These
 * kernels train and run a 3-D convolutional encoder-decoder that jumps K
 * time steps of the LBM-IB loop (do_thread, lbm_ib.c:3115-3563) on the
 * coarse-grained fluid state of compute_rho_and_u (lbm_ib.c:2440-2538) and
 * the sheet state of moveFiberSheet (lbm_ib.c:2540-2723); the engine's own
 * entry points (lbmib_coarse_moments, lbmib_jump_*, include/lbmib.h) take the
 * state out of and back into a running simulation.
 *
 * Conventions:
 *   - every tensor argument is a DEVICE pointer; sizes are plain ints;
 *   - activations are BF16, channels-last: [B][D][H][W][C] ("voxel-major"),
 *     C a multiple of 16;
 *   - convolution weights are BF16 [Cout][27][Cin], tap = 9(dz+1) + 3(dy+1)
 *     + (dx+1), zero padding ("same" 3x3x3, stride 1); master weights,
 *     gradients and optimiser moments are FP32 of the same layout;
 *   - `stream` is a hipStream_t passed as void* (NULL: the null stream);
 *     every call only enqueues work on it;
 *   - return values are LBMIB_* codes of include/lbmib.h.
 *
 * Arithmetic: the convolutions multiply BF16 operands on the MFMA units
 * (v_mfma_f32_16x16x32_bf16) and accumulate in FP32; outputs are rounded
 * to BF16 once (activations) or kept in FP32 (weight gradients).
 */
#ifndef LBMIB_SURROGATE_H
#define LBMIB_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Forward convolution: y = act(conv(x, w) + bias) [* (mask > 0)].
 * relu: 1 applies max(0, .); mask (nullable) has the shape of y.  bias is
 * FP32 [Cout] or NULL.  Cin in {16, 32, 64, 128, 256}, Cout a multiple of 16.
 * On some small grids the reduction is split into FP32 partials in the
 * library workspace (below): the calls that use it -- this one,
 * lbmib_sg_conv3d_bwd_data, lbmib_sg_conv3d_bwd_weight, lbmib_sg_mse -- must
 * not overlap on one device (one stream, or ordered). */
int lbmib_sg_conv3d_fwd(const void *x, const void *w, const float *bias, const void *mask, void *y, int B, int D,
                        int H, int W, int Cin, int Cout, int relu, void *stream);

/* Data gradient: dx = conv^T(dy, w) [* (mask > 0)], where w is the forward
 * weight [Cout][27][Cin]; wt is a [Cin][27][Cout] buffer the call fills with
 * the flipped, transposed weight (wt[ci][26 - t][co] = w[co][t][ci]) and then
 * convolves dy with.  mask (nullable): the forward input when it is a ReLU
 * output, so that dx is the gradient at the previous layer's pre-activation. */
int lbmib_sg_conv3d_bwd_data(const void *dy, const void *w, void *wt, const void *mask, void *dx, int B, int D, int H,
                             int W, int Cin, int Cout, void *stream);

/* Weight and bias gradients: dw[co][t][ci] = sum_v dy[v][co] x[v + o(t)][ci],
 * db[co] = sum_v dy[v][co] (db nullable), FP32, overwritten.  The sum over
 * voxels is split over blocks into FP32 partials in a workspace the library
 * allocates once per device (LBMIB_SG_WORKSPACE_BYTES), then reduced in a
 * fixed order: results are deterministic.  Calls on one device share that
 * workspace and must not overlap (one stream, or ordered). */
int lbmib_sg_conv3d_bwd_weight(const void *x, const void *dy, float *dw, float *db, int B, int D, int H, int W,
                               int Cin, int Cout, void *stream);
#define LBMIB_SG_WORKSPACE_BYTES (32u << 20)

/* The same weight and bias gradients, bit for bit, with no allocation of
 * any kind.  No counterpart in the reference: this entry point is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight training loops, harnesses that forbid hipMalloc).  Its FP32 partials
 * live in a fixed device array of LBMIB_SG_NOALLOC_SCRATCH_BYTES (32 MiB)
 * that each GPU the library runs on holds until the process ends; shapes
 * whose partials need more (Cout * 27 * Cin * 4 bytes above that) return
 * LBMIB_ERR_ARG.  All calls on one GPU share that array: calls that overlap
 * on one GPU (different streams without ordering) race.  The caller passes
 * every input and output. */
int lbmib_sg_conv3d_bwd_weight_noalloc(const void *x, const void *dy, float *dw, float *db, int B, int D, int H,
                                       int W, int Cin, int Cout, void *stream);
#define LBMIB_SG_NOALLOC_SCRATCH_BYTES (32u << 20)

/* 2x2x2 average pooling: x [B][2D][2H][2W][C] -> y [B][D][H][W][C]. */
int lbmib_sg_pool2(const void *x, void *y, int B, int D, int H, int W, int C, void *stream);
/* Nearest 2x upsampling plus skip: s[v] = c[parent(v)] + e[v] (e nullable);
 * s, e: [B][D][H][W][C] (fine), c: [B][D/2][H/2][W/2][C]. */
int lbmib_sg_up2_add(const void *c, const void *e, void *s, int B, int D, int H, int W, int C, void *stream);
/* Backward of pooling, with a skip gradient and a ReLU mask:
 * out[v] = (a[v] + g[parent(v)] / 8) * (mask[v] > 0); a, mask nullable;
 * out, a, mask fine [B][D][H][W][C], g coarse. */
int lbmib_sg_pool2_bwd(const void *a, const void *g, const void *mask, void *out, int B, int D, int H, int W, int C,
                       void *stream);
/* Backward of upsampling, with a ReLU mask: out[u] = (sum over the 8
 * children v of u of ds[v]) * (mask[u] > 0); mask nullable; out, mask coarse
 * [B][D][H][W][C], ds fine. */
int lbmib_sg_up2_bwd(const void *ds, const void *mask, void *out, int B, int D, int H, int W, int C, void *stream);

/* Mean-squared error over the first nl channels: loss (FP32, one value) =
 * mean over v, c < nl of (out[v][c] - tgt[v][c])^2; dout = its gradient
 * (zero on channels >= nl).  out, dout BF16 [M][C]; tgt FP32 [M][nl]. */
int lbmib_sg_mse(const void *out, const float *tgt, void *dout, float *loss, long M, int C, int nl, void *stream);

/* Adam step on n FP32 parameters (p, g, m, v), t = step count from 1; writes
 * the BF16 copy w16 (nullable) that the convolutions read. */
int lbmib_sg_adam(float *p, const float *g, float *m, float *v, void *w16, long n, float lr, float beta1, float beta2,
                  float eps, int t, void *stream);

/* BF16 copy (round to nearest even) of n FP32 values. */
int lbmib_sg_cast_bf16(const float *src, void *dst, long n, void *stream);

/* Network input from field planes: x[v][c] = (f[c][v] - mean[c]) * istd[c]
 * for c < nf (f FP32 [nf][M]), cval[c - nf] for nf <= c < nf + nconst, zero
 * up to C.  mean/istd/cval are host arrays. */
int lbmib_sg_pack(const float *f, int nf, const double *mean, const double *istd, int nconst, const double *cval,
                  void *x, long M, int C, void *stream);
/* Training target: tgt[v][c] = ((f1[c][v] - f0[c][v]) - mean[c]) * istd[c],
 * c < nl; f0, f1 FP32 [nl][M] (field planes at t and t + K). */
int lbmib_sg_target(const float *f0, const float *f1, int nl, const double *mean, const double *istd, float *tgt,
                    long M, void *stream);
/* Network output to field planes: d[c][v] = out[v][c] * std[c] + mean[c],
 * c < nl (d FP32 [nl][M]). */
int lbmib_sg_unpack(const void *out, int C, int nl, const double *mean, const double *std, float *d, long M,
                    void *stream);
/* FP64 sums {sum (a - b)^2, sum b^2} over n FP32 values (b nullable: then
 * {sum a^2, 0}), written to out2 (device, 2 doubles); deterministic.  Its
 * partial sums have a buffer of their own, not the workspace, so it may run
 * beside the convolutions on another stream (calls of it must not overlap
 * each other). */
int lbmib_sg_sqdiff(const float *a, const float *b, long n, double *out2, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LBMIB_SURROGATE_H */
