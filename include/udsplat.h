/*
 * udsplat.h -- C ABI of the MI355X (gfx950) splatting hot path.
 *
 * Every entry point takes raw device pointers, plain sizes and the HIP stream
 * to launch on, and returns a status code (UDE_OK = 0, negative on error).
 * Outputs are caller-allocated device buffers; scratch memory is a
 * caller-provided workspace whose size is given by the matching
 * *_workspace_bytes query.  No entry point allocates device memory or
 * synchronises the host.
 *
 * Tensor layouts are those of the reference op surface:
 *   slot tensors (indices, weights)  (bn, W, H, k), pixel (x=i, y=j) at j + H (i + W b)
 *   points / normals                 (bn, N, 3) fp32 camera space
 *   std-devs                         (sdBatch, N, 2), sdBatch in {1, bn}
 * Index value -1 marks an empty slot.
 */
#ifndef UDSPLAT_H
#define UDSPLAT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* a HIP stream (hipStream_t) */
typedef struct ihipStream_t* ude_stream_t;

#define UDE_OK 0
#define UDE_ERR_ARG (-1)
#define UDE_ERR_LAUNCH (-2)
#define UDE_ERR_WORKSPACE (-3)
#define UDE_ERR_SCRATCH (-4)     /* the allocation-free entry point's sizes exceed its static scratch */

/* Size of the static device scratch of ude_splat_backward_static (64 MiB per device). */
#define UDE_STATIC_SCRATCH_BYTES (64u << 20)

/* ---- sampling ---------------------------------------------------------- */

/* Workspace for ude_sample (one view's tree, reused across views). */
size_t ude_sample_workspace_bytes(int N, int W, int H, int k);

/* sampling.sample -- replaces cudaSample / batchSample
 * (FMMKdTreeSampling/sampling.cpp:48-94, samplingCuda.cu:197-321).
 * out: int32 (bn, W, H, k).  seed keys the Philox4x32-10 uniforms (the
 * reference seeds cuRAND with time(NULL)).  *errFlag (device int, caller
 * zeroed) receives a non-zero bit if a descent exceeded 63 levels. */
int ude_sample(int bn, int N, int W, int H, int k, double f, double smoothing, double precision,
               const float* pts, const float* nrm, const float* sd, int sdBatch,
               uint64_t seed, int32_t* out, int* errFlag, void* ws, size_t wsBytes,
               ude_stream_t stream);

/* Workspace for ude_fmm_weights (one view's tree, reused across views). */
size_t ude_fmm_weights_workspace_bytes(int N, int W, int H);

/* sampling.approximateFMMWeights -- replaces cudaApproxWeights / batchApproxWeights
 * (FMMKdTreeSampling/sampling.cpp:136-175, samplingCuda.cu:398-493): the
 * fixed-order (degree 6) Hermite estimate of sum_i exp(-|q - s_i|^2 / 2h^2) /
 * (2 pi sqrt(h)) at every pixel q.  out: float (bn, W, H), pixel (i, j) at
 * j + H i (the reference writes j + W i, identical for W == H). */
int ude_fmm_weights(int bn, int N, int W, int H, double f, double smoothing, double precision,
                    const float* pts, const float* nrm, const float* sd, int sdBatch,
                    float* out, void* ws, size_t wsBytes, ude_stream_t stream);

/* Workspace for ude_sample_poisson over bn views with every view's tree
 * resident (bn per-view blocks); a workspace of one block (bn = 1) also
 * works, view after view. */
size_t ude_poisson_workspace_bytes(int bn, int N, int W, int H);

/* Poisson-rate sampling with inclusion probabilities (no counterpart in the
 * reference, whose plain-Poisson kernel poissonRadixTree.cuh:70-200 is dead
 * code with unknown inclusion probabilities).  At pixel q every culled splat i
 * is included independently with the defensive-mixture probability
 *   pi_i = min(1, r p_i),  p_i = (g0_i(q) / max(D0(q), c0) + g1_i(q) / max(D1(q), c1)) / 2,
 * r = rate * tileScale, g_b = c_b exp(-|q - s_i|^2 / 2 h_b^2), c_b = 1 / (2 pi
 * sqrt(h_b)), h_0 the tree's bandwidth, h_1 = max(h_0, 99.9th percentile of
 * the visible splats' sqrt(lambda_max(V_i))), D_b the fixed-order density of h_b
 * (nothing where D0 < precision; pairs with p_i < 1e-4 are outside the support,
 * which bounds 1 / pi by 1e4 / r).  tileScale (optional, (bn, ceil(W/16),
 * ceil(H/16))) multiplies the rate per 16x16 tile.  Outputs (bn, W, H, kcap):
 * splat ids (-1 padded) and 1 / pi_i (0 padded), so sum_s w_s a_s / pi_s is a
 * Horvitz-Thompson estimate of sum_i w_i a_i.  densOut (optional, (bn, 2, W, H))
 * receives D0 and D1.  A pixel with more than kcap inclusions sets bit 4 of
 * *errFlag; a walk longer than the tree (not expected) stops and sets bit 8; a
 * tile whose rate times multiplier is negative or not finite draws nothing and
 * sets bit 16.  seedDev (optional, device memory): the Philox key is read there
 * at run time instead of from seed, so a captured HIP graph draws fresh
 * samples when the caller advances it between replays.  With a workspace of
 * ude_poisson_workspace_bytes(bn, ...) bytes all views' trees are built
 * first and each per-pixel pass (both densities, the thinning) is one launch
 * over the whole batch. */
int ude_sample_poisson(int bn, int N, int W, int H, int kcap, double f, double smoothing, double precision,
                       double rate, const float* tileScale, const float* pts, const float* nrm, const float* sd,
                       int sdBatch, uint64_t seed, const uint64_t* seedDev, int32_t* outIds, float* outIpi,
                       float* densOut, int* errFlag, void* ws, size_t wsBytes, ude_stream_t stream);

/* Proposal of ude_sample_poisson (no counterpart in the reference).  mode 0:
 * the two density bands above.  mode 1: the wide band plus each splat's own
 * footprint, p_i = g1_i / (2 N1) + f_i / (2 kw N1), f_i(q) = min(1, w_i(q),
 * E(|q - s_i|)) with w_i the splat's EWA weight (FP32 screen covariance and
 * Jacobian, amp exp(-(u^2 + v^2) / 2)), E(d) = exp(-max(0, d - 2)^2 / 2h_1^2) a
 * cap common to all splats (a node's bound is E at its box distance),
 * N1 = max(D1, c1), and
 * kw = footprintScale, or, when footprintScale <= 0, the visible splats' mean
 * footprint mass (integral of min(w, 1)) over h_1^1.5, which gives each half
 * about r / 2 expected samples.  Applies to later calls in the process (one
 * setting, not per thread).  ude_get_poisson_proposal returns the mode and
 * writes the scale. */
int ude_set_poisson_proposal(int mode, double footprintScale);
int ude_get_poisson_proposal(double* footprintScale);

/* Order of the fixed-order Hermite expansions behind ude_fmm_weights and the
 * densities of ude_sample_poisson (no counterpart in the reference, whose
 * order is a compile-time constant, hermitefmm.cuh:11): 4 .. 8, default 6.
 * A node carries an expansion when its order-p truncation bound meets the
 * precision, so every order keeps the stated error bound; lower orders accept
 * fewer, smaller nodes.  Process-wide; read when a call starts. */
int ude_set_fixed_degree(int p);
int ude_get_fixed_degree(void);

/* Launch shape of the per-pixel passes (no counterpart in the reference):
 * kernel 0 the fixed-order density evaluation, 1 the Poisson-rate thinning.
 * One wavefront per tileW x tileH pixel tile (tileW tileH = 64, powers of
 * two), wavesPerBlock in {1, 2, 4, 8} per workgroup, xcdOrder 1 to give each
 * of the 8 XCDs a contiguous band of tiles.  Results do not depend on it.
 * Process-wide; default 8 x 8, 4, 0.  get writes the four values to cfg. */
int ude_set_tile_config(int kernel, int tileW, int tileH, int wavesPerBlock, int xcdOrder);
int ude_get_tile_config(int kernel, int* cfg);

/* Features of 16x16 pixel tiles for the rate network (no counterpart in the
 * reference).  dens: (bn, 2, W, H) densities D0, D1 as returned by
 * ude_sample_poisson (densOut); gimg (optional): (bn, W, H, 3) image gradient.
 * out: (bn, ceil(W/16), ceil(H/16), 32): [0] log(1 + visible splat centres),
 * [1] [2] mean / std of their depth, [3] [4] mean / max of their bandwidth,
 * [5..8] mean / max / min / std of log(1 + D0) over the tile's pixels,
 * [9..12] the same for D1, [13] fraction of pixels with D0 >= precision,
 * [14..16] mean / max / std of |gimg|, [17] log(1 + sum |gimg|^2), [18..25]
 * per quadrant mean log(1 + D0) and mean |gimg|, [26] [27] tile centre / (W, H),
 * [28] log(1 + sum D0), [29] log(1 + sum |gimg| D0), [30] pixel coverage,
 * [31] log(1 + count coverage). */
size_t ude_tile_features_workspace_bytes(int W, int H);
int ude_tile_features(int bn, int N, int W, int H, double f, double smoothing, double precision, const float* pts,
                      const float* nrm, const float* sd, int sdBatch, const float* dens, const float* gimg,
                      float* out, void* ws, size_t wsBytes, ude_stream_t stream);

/* ---- per-tile rate network (no counterpart in the reference) ----------- */

/* A 32 -> 64 -> 64 -> 1 perceptron (ReLU) from 32 features of a 16x16 pixel
 * tile to the log of its Poisson-rate multiplier; bfloat16 matrix-core
 * products with FP32 accumulation.  Parameters: FP32 block of
 * ude_mlp_param_count() floats (W1 [32][64], b1, W2 [64][64], b2, W3 [64], b3). */
int ude_mlp_param_count(void);
size_t ude_mlp_workspace_bytes(int T);

/* y[t] = network(X[t]) for T rows of 32 features. */
int ude_mlp_forward(int T, const float* X, const float* params, float* y, void* ws, size_t wsBytes,
                    ude_stream_t stream);

/* Weighted mean-square loss sum_t w_t (y_t - target_t)^2 / sum_t w_t (weight
 * optional) and its gradient with respect to the parameter block (grad, fully
 * written); *loss receives the loss. */
int ude_mlp_loss_grad(int T, const float* X, const float* params, const float* target, const float* weight,
                      float* grad, float* loss, void* ws, size_t wsBytes, ude_stream_t stream);

/* One Adam step (bias-corrected, step >= 1) on n FP32 parameters. */
int ude_adam_step(int n, float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                  float eps, int step, ude_stream_t stream);

/* Profiling aid (no counterpart in the reference): between begin and end,
 * every launch of one sampling kernel is bracketed by HIP events on its
 * stream -- kernel 0: the per-pixel FGT walk, capacity descents and keyed
 * draws of ude_sample (one pair of launches per view, bracketed together);
 * 1: the fixed-order density evaluation (one launch per batch of views for
 * both bands in ude_sample_poisson, one per view in ude_fmm_weights); 2: the
 * Poisson-rate thinning pass of ude_sample_poisson.  end waits for the
 * recorded events and returns the summed kernel time and the number of
 * launches.  Host-thread-unsafe; meant for benchmarks. */
int ude_profile_sampler_begin(int kernel);
int ude_profile_sampler_end(double* totalMs, int* launches);

/* sampling.covariance -- replaces cudaCovariance (sampling.cpp:96-134,
 * samplingCuda.cu:323-396).  out: (bn, N, 2, 2). */
int ude_covariance(int bn, int N, int W, int H, double f, double smoothing, const float* pts,
                   const float* nrm, const float* sd, int sdBatch, float* out, ude_stream_t stream);

/* ---- splatting --------------------------------------------------------- */

size_t ude_splat_workspace_bytes(int bn, int N);

/* splatting.forward -- replaces cudaSplattingForward (splatting.cpp:43-82,
 * splattingCuda.cu:18-101).  out: (bn, W, H, k) = min(w, 1), 0 for empty slots. */
int ude_splat_forward(int bn, int N, int W, int H, int k, double f, double smoothing,
                      const int32_t* idx, const float* pts, const float* nrm, const float* sd, int sdBatch,
                      float* out, void* ws, size_t wsBytes, ude_stream_t stream);

/* splatting.backward -- replaces cudaSplattingBackward (splatting.cpp:84-129,
 * splattingCuda.cu:114-234).  gPts, gNrm: (bn, N, 3), fully written.
 * gSd: (sdBatch, N, 2); fully written if sdBatch == bn, accumulated into
 * (caller zeroed) if sdBatch == 1 < bn. */
int ude_splat_backward(int bn, int N, int W, int H, int k, double f, double smoothing,
                       const float* gradIn, const int32_t* idx, const float* pts, const float* nrm,
                       const float* sd, int sdBatch, float* gPts, float* gNrm, float* gSd,
                       void* ws, size_t wsBytes, ude_stream_t stream);

/* splatting.backward followed by the post-processing of the reference's
 * EWASplatting.backward -- replaces splatting/extensions/__init__.py:210-225
 * (the cudaExtensions call, conditionalNormalization of the point and std-dev
 * gradients, utils.py:147-150, and their NaN -> 0), fused into the splat pass:
 * gPts and gSd rows become g / max(|g| + 1e-4, 1) with NaN entries set to 0;
 * gNrm stays raw (the reference returns no normal gradient).  Per-view
 * std-devs only (sdBatch = bn: a shared row is a sum over views, which the
 * post-processing must follow).  gPts, gNrm (bn, N, 3) and gSd (bn, N, 2) fully
 * written. */
int ude_splat_backward_normalized(int bn, int N, int W, int H, int k, double f, double smoothing,
                                  const float* gradIn, const int32_t* idx, const float* pts, const float* nrm,
                                  const float* sd, float* gPts, float* gNrm, float* gSd, void* ws, size_t wsBytes,
                                  ude_stream_t stream);
int ude_splat_backward_normalized_f64(int bn, int N, int W, int H, int k, double f, double smoothing,
                                      const double* gradIn, const int32_t* idx, const double* pts,
                                      const double* nrm, const double* sd, double* gPts, double* gNrm,
                                      double* gSd, void* ws, size_t wsBytes, ude_stream_t stream);

/* No counterpart in the reference: ude_splat_backward for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and outputs as
 * ude_splat_backward without the workspace: the kernels run on a static device
 * array of UDE_STATIC_SCRATCH_BYTES per device (its address taken once per device,
 * on the first call with that device current).  Needs
 * ude_splat_workspace_bytes(bn, N) <= UDE_STATIC_SCRATCH_BYTES (bn * N <= about
 * 1M splat-views), else returns UDE_ERR_SCRATCH.  Allocates nothing.  Every call on
 * one device shares that device's array: calls that overlap on one device (other
 * streams, other host threads) race; calls ordered on one stream do not. */
int ude_splat_backward_static(int bn, int N, int W, int H, int k, double f, double smoothing,
                              const float* gradIn, const int32_t* idx, const float* pts, const float* nrm,
                              const float* sd, int sdBatch, float* gPts, float* gNrm, float* gSd,
                              ude_stream_t stream);

/* ---- filtering (z-sort + alpha compositing) ----------------------------- */

/* filtering.forward -- replaces cudaAlphaBlendForward (zSorting.cpp:24-32,
 * zSortingCuda.cu:64-123).  pts/nrm: (pb, N, 3), pb in {1, bn}.  k <= 1024. */
int ude_filter_forward(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const float* wts,
                       const float* pts, const float* nrm, int pb, float* out, ude_stream_t stream);

/* filtering.backward -- replaces cudaAlphaBlendBackward (zSorting.cpp:34-42,
 * zSortingCuda.cu:180-243).  out: gradient w.r.t. the weights (bn, W, H, k). */
int ude_filter_backward(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const float* wts,
                        const float* gin, const float* pts, const float* nrm, int pb, float* out,
                        ude_stream_t stream);

/* filtering.backward(..., exact=True): the exact derivative of the composite.  The
 * reference drops a slot's own term where |w| <= 1e-4 and its occlusion term where
 * |1 - w| <= 1e-4 (alphacompositingGrad.cuh:7-26: it divides by w and by 1 - w);
 * the adjoint here divides by neither, so this variant keeps every term.  Same
 * arguments as ude_filter_backward.  Used by the raw (unclipped) gradient mode. */
int ude_filter_backward_exact(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const float* wts,
                              const float* gin, const float* pts, const float* nrm, int pb, float* out,
                              ude_stream_t stream);
int ude_filter_backward_exact_f64(int bn, int N, int W, int H, int k, double f, const int32_t* idx,
                                  const double* wts, const double* gin, const double* pts, const double* nrm, int pb,
                                  double* out, ude_stream_t stream);

/* ---- transform ---------------------------------------------------------- */

/* transform.{affine,linear}Forward -- replaces cudaTransformForward
 * (transform.cpp:34-60, transformCuda.cu:39-72).  M: (bn, 4, 4); v: (vb, N, 3). */
int ude_transform_forward(int affine, int bn, int vb, int N, const float* M, const float* v, float* out,
                          ude_stream_t stream);

/* transform.linearForwardFacing -- no counterpart in the reference's extension:
 * linearForward of normals plus, in the same pass, the camera-facing copy the
 * reference's renderer makes with two torch.where calls (render.py:200-203):
 * facing = -out where out.z > 0, else out. */
int ude_transform_forward_facing(int bn, int vb, int N, const float* M, const float* v, float* out, float* facing,
                                 ude_stream_t stream);

/* transform.{affine,linear}Backward -- replaces cudaTransformBackward
 * (transform.cpp:62-85, transformCuda.cu:119-156).  gM (bn,4,4) and, when
 * vb == 1 < bn, gV (1,N,3) are accumulated into (caller zeroed). */
int ude_transform_backward(int affine, int bn, int vb, int N, const float* g, const float* M, const float* v,
                           float* gM, float* gV, ude_stream_t stream);

/* ---- interpolation ------------------------------------------------------ */

/* interpolation.forward -- replaces cudaInterpolationForward
 * (interpolation.cpp:31-58, interpolationCuda.cu:5-81).  attr: (ab, N, C). */
int ude_interp_forward(int bn, int W, int H, int k, int ab, int N, int C, const int32_t* idx, const float* wts,
                       const float* attr, float* out, ude_stream_t stream);

/* interpolation.backward -- replaces cudaInterpolationBackward
 * (interpolation.cpp:60-96, interpolationCuda.cu:83-167).  gW fully written,
 * gA accumulated into (caller zeroed). */
int ude_interp_backward(int bn, int W, int H, int k, int ab, int N, int C, const float* g, const int32_t* idx,
                        const float* wts, const float* attr, float* gW, float* gA, ude_stream_t stream);

/* ---- Horvitz-Thompson image estimate ------------------------------------ */

/* No reference counterpart: splatting.forward, the 1 / pi weighting and
 * interpolation.forward of a three-channel attribute fused per pixel, and their
 * adjoints fused per sample, for the unbiased renderer over samplePoisson's
 * output.  idx / ipi: (bn, W, H, k) with each pixel's samples first, then -1
 * (ids) padding; a pixel's slots after its first -1 are not read.  attr:
 * (ab, N, 3), ab in {1, bn}; pts / nrm (bn, N, 3); sd (sdBatch, N, 2).
 * The image equals the unfused ops' bit for bit; the gradients equal theirs up
 * to the order of atomic sums.  No per-slot tensor is stored. */
size_t ude_ht_workspace_bytes(int bn, int N);

/* out: (bn, W, H, 3) = sum_s min(w_s, 1) ipi_s attr[idx_s]. */
int ude_ht_forward(int bn, int N, int W, int H, int k, double f, double smoothing, const int32_t* idx,
                   const float* ipi, const float* pts, const float* nrm, const float* sd, int sdBatch, int ab,
                   const float* attr, float* out, void* ws, size_t wsBytes, ude_stream_t stream);

/* Adjoint of ude_ht_forward for gradImage (bn, W, H, 3).  gPts, gNrm fully
 * written; gSd as in ude_splat_backward; gAttr fully written if ab == bn,
 * accumulated into (caller zeroed) if ab == 1 < bn.  prepared != 0: ws is the
 * workspace ude_ht_forward used for the same splats, and its per-splat records
 * are reused instead of recomputed. */
int ude_ht_backward(int bn, int N, int W, int H, int k, double f, double smoothing, const float* gradImage,
                    const int32_t* idx, const float* ipi, const float* pts, const float* nrm, const float* sd,
                    int sdBatch, int ab, const float* attr, float* gPts, float* gNrm, float* gSd, float* gAttr,
                    int prepared, void* ws, size_t wsBytes, ude_stream_t stream);

/* ---- float64 and half variants ------------------------------------------ */

/* The reference dispatches these ops on float and double (AT_DISPATCH_FLOATING_TYPES:
 * splattingCuda.cu:84,214; transformCuda.cu:53,135; interpolationCuda.cu:68,153;
 * shadingCuda.cu:63,153,225,315), filtering also on half (zSortingCuda.cu:100,220).
 * Each _f64 entry point is its float counterpart above with every real tensor in
 * double (FP64 arithmetic and FP64 atomics); the _f16 filtering entry points take
 * half tensors and compute in float.  Arguments and outputs as the float ones. */
size_t ude_splat_workspace_bytes_f64(int bn, int N);
int ude_splat_forward_f64(int bn, int N, int W, int H, int k, double f, double smoothing,
                          const int32_t* idx, const double* pts, const double* nrm, const double* sd, int sdBatch,
                          double* out, void* ws, size_t wsBytes, ude_stream_t stream);
int ude_splat_backward_f64(int bn, int N, int W, int H, int k, double f, double smoothing,
                           const double* gradIn, const int32_t* idx, const double* pts, const double* nrm,
                           const double* sd, int sdBatch, double* gPts, double* gNrm, double* gSd,
                           void* ws, size_t wsBytes, ude_stream_t stream);
int ude_filter_forward_f64(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const double* wts,
                           const double* pts, const double* nrm, int pb, double* out, ude_stream_t stream);
int ude_filter_backward_f64(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const double* wts,
                            const double* gin, const double* pts, const double* nrm, int pb, double* out,
                            ude_stream_t stream);
/* half tensors (IEEE binary16) passed as void* */
int ude_filter_forward_f16(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const void* wts,
                           const void* pts, const void* nrm, int pb, void* out, ude_stream_t stream);
int ude_filter_backward_f16(int bn, int N, int W, int H, int k, double f, const int32_t* idx, const void* wts,
                            const void* gin, const void* pts, const void* nrm, int pb, void* out,
                            ude_stream_t stream);
int ude_transform_forward_f64(int affine, int bn, int vb, int N, const double* M, const double* v, double* out,
                              ude_stream_t stream);
int ude_transform_backward_f64(int affine, int bn, int vb, int N, const double* g, const double* M,
                               const double* v, double* gM, double* gV, ude_stream_t stream);
int ude_interp_forward_f64(int bn, int W, int H, int k, int ab, int N, int C, const int32_t* idx,
                           const double* wts, const double* attr, double* out, ude_stream_t stream);
int ude_interp_backward_f64(int bn, int W, int H, int k, int ab, int N, int C, const double* g,
                            const int32_t* idx, const double* wts, const double* attr, double* gW, double* gA,
                            ude_stream_t stream);
int ude_sh_forward_f64(int bn, int W, int H, int cb, const double* coef, const double* nrm, const double* diff,
                       double* out, ude_stream_t stream);
int ude_sh_backward_f64(int bn, int W, int H, int cb, const double* g, const double* coef, const double* nrm,
                        const double* diff, double* gCoef, double* gN, double* gD, ude_stream_t stream);
int ude_lambert_forward_f64(int bn, int W, int H, int L, int ib, const double* dirs, const double* ints,
                            const double* nrm, const double* diff, double* out, ude_stream_t stream);
int ude_lambert_backward_f64(int bn, int W, int H, int L, int ib, const double* g, const double* dirs,
                             const double* ints, const double* nrm, const double* diff, double* gDir,
                             double* gInt, double* gN, double* gD, ude_stream_t stream);

/* ---- pixelgather ------------------------------------------------------- */

/* pixelgather.scatter -- replaces scatterCuda (splatting/extensions/pixelGather.cpp:10-13,
 * pixelGatherCuda.cu:7-74): out (bn, W, H, k, C)[b, i, j, s, c] = values[b, idx, c]
 * for idx = indices[b, i, j, s] >= 0, else 0 (every element written).  values
 * (bn, N, C).  float, double and (as void*) half, as the reference dispatches. */
int ude_pixel_scatter(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const float* values,
                      float* out, ude_stream_t stream);
int ude_pixel_scatter_f64(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const double* values,
                          double* out, ude_stream_t stream);
int ude_pixel_scatter_f16(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const void* values,
                          void* out, ude_stream_t stream);

/* pixelgather.gather -- replaces gatherCuda (pixelGather.cpp:15-18,
 * pixelGatherCuda.cu:76-152): out (bn, N, C)[b, idx, c] += values[b, i, j, s, c]
 * over the slots with idx = indices[b, i, j, s] >= 0 (caller zeroed; device
 * atomics, so the summation order is not fixed).  float and double. */
int ude_pixel_gather(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const float* values, float* out,
                     ude_stream_t stream);
int ude_pixel_gather_f64(int bn, int W, int H, int k, int N, int C, const int32_t* idx, const double* values,
                         double* out, ude_stream_t stream);

/* ---- nearestsamples ------------------------------------------------------ */

/* Workspace of ude_nearest_samples: N * K 64-bit words. */
size_t ude_nearest_workspace_bytes(int N, int K);

/* nearestsamples.nearest -- replaces cudaFastNearestSamples (nearestSamples.cpp:30-75,
 * nearestSamplesCuda.cu:290-564).  idx / wts (bn, W, H, k); pts (1, N, 3) (the
 * reference also assumes one point set for all views).  Each pixel's surface
 * splat is its first slot of largest weight; out (N, K) int32 holds, for every
 * splat, the K nearest distinct splats sampled at the pixels it is the surface
 * of, by float32 squared distance (ties by index), the splat itself excluded, -1
 * padded.  The rows do not depend on the order the pixels run in.  float and
 * double (the double variant orders by the squared distance rounded to float). */
int ude_nearest_samples(int bn, int W, int H, int k, int N, int K, const int32_t* idx, const float* wts,
                        const float* pts, int32_t* out, void* ws, size_t wsBytes, ude_stream_t stream);
int ude_nearest_samples_f64(int bn, int W, int H, int k, int N, int K, const int32_t* idx, const double* wts,
                            const double* pts, int32_t* out, void* ws, size_t wsBytes, ude_stream_t stream);

/* ---- composited Horvitz-Thompson image estimate ------------------------- */

/* No reference counterpart: the Poisson-rate counterpart of splatting.forward ->
 * filtering.forward -> interpolation.forward (splattingCuda.cu:18-101,
 * zSortingCuda.cu:11-123, interpolationCuda.cu:5-46) and of their backward ops
 * (splattingCuda.cu:114-234, zSortingCuda.cu:126-243, interpolationCuda.cu:83-167).
 * Per pixel, with alpha_s = min(w_s, 1) ipi_s and the samples ordered by the
 * ray depth filtering uses (alphacompositing.cuh:6-54, stable, invalid depth
 * 10000):  image = sum_s alpha_s attr_s prod_{t in front of s} (1 - alpha_t).
 * With independent inclusions (samplePoisson) it is an unbiased estimate of the
 * image composited over all splats of the sampling support.  Inputs as for
 * ude_ht_forward; k <= 112. */
size_t ude_htc_workspace_bytes(int bn, int N, int W, int H, int k);

/* out: (bn, W, H, 3), the composited estimate. */
int ude_htc_forward(int bn, int N, int W, int H, int k, double f, double smoothing, const int32_t* idx,
                    const float* ipi, const float* pts, const float* nrm, const float* sd, int sdBatch, int ab,
                    const float* attr, float* out, void* ws, size_t wsBytes, ude_stream_t stream);

/* Exact adjoint of ude_htc_forward (the depth order held fixed, as the
 * reference's filtering backward holds it): outputs as ude_ht_backward.
 * prepared != 0: ws is the workspace ude_htc_forward used for the same splats. */
int ude_htc_backward(int bn, int N, int W, int H, int k, double f, double smoothing, const float* gradImage,
                     const int32_t* idx, const float* ipi, const float* pts, const float* nrm, const float* sd,
                     int sdBatch, int ab, const float* attr, float* gPts, float* gNrm, float* gSd, float* gAttr,
                     int prepared, void* ws, size_t wsBytes, ude_stream_t stream);

/* ---- L2 image loss -------------------------------------------------------- */

/* No reference counterpart (its scripts use torch ops): grad = 2 (image - target)
 * and sums[0] = sum (image - target)^2 over n floats, accumulated in FP64 in a
 * fixed order; sums holds 1 + 1024 doubles (the rest is scratch).  Records no
 * memset (usable under graph capture). */
int ude_l2_grad(long n, const float* image, const float* target, float* grad, double* sums, ude_stream_t stream);

/* ---- shading ------------------------------------------------------------ */

/* shading.forwardSh -- replaces cudaShadingShForward (shading.cpp:141-168,
 * shadingCuda.cu:177-242).  coef: (cb, 3, 9); normals/diffuse (bn, W, H, 3). */
int ude_sh_forward(int bn, int W, int H, int cb, const float* coef, const float* nrm, const float* diff,
                   float* out, ude_stream_t stream);

/* shading.backwardSh -- replaces cudaShadingShBackward (shading.cpp:170-200,
 * shadingCuda.cu:244-330).  gCoef accumulated into (caller zeroed). */
int ude_sh_backward(int bn, int W, int H, int cb, const float* g, const float* coef, const float* nrm,
                    const float* diff, float* gCoef, float* gN, float* gD, ude_stream_t stream);

/* shading.forward (Lambert) -- replaces cudaShadingForward (shading.cpp:51-88,
 * shadingCuda.cu:12-81).  dirs (bn, L, 3); ints (ib, L, 3). */
int ude_lambert_forward(int bn, int W, int H, int L, int ib, const float* dirs, const float* ints,
                        const float* nrm, const float* diff, float* out, ude_stream_t stream);

/* shading.backward (Lambert) -- replaces cudaShadingBackward (shading.cpp:90-139,
 * shadingCuda.cu:83-175).  gDir, gInt accumulated into (caller zeroed). */
int ude_lambert_backward(int bn, int W, int H, int L, int ib, const float* g, const float* dirs,
                         const float* ints, const float* nrm, const float* diff, float* gDir, float* gInt,
                         float* gN, float* gD, ude_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* UDSPLAT_H */
