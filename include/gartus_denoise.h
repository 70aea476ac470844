/*
 * gartus_denoise.h — C ABI of the learned denoiser.
 *
 * No counterpart in the reference: gartus renders plain Monte Carlo images.
 * It returns the per-pixel averages that RayCamera::render_world*_hdr_image
 * and render_world*_denoising_aovs (src/graphics/camera.rs:4175-4336,
 * 3064-3116) produce. The denoiser takes exactly those outputs:
 *   - a low-spp beauty image;
 *   - first-hit albedo;
 *   - normal AOVs.
 * It returns an estimate of the converged image. The plain path tracer stays
 * as the fallback, and as the reference the denoiser's error is measured against.
 *
 * Network: a convolutional encoder-decoder (U-Net).
 *   - `levels` resolutions, with channels[l] features at level l.
 *   - Each level has two 3x3 convolutions with ReLU.
 *   - Between levels: 2x2 max pooling down, nearest 2x upsampling up, and a
 *     skip concatenation.
 *   - Output: a final 3x3 convolution to three channels.
 *   - Input: log1p(beauty), albedo, normal, padded to 16 channels. With
 *     beauty_halves, the beauty is two independent renders of half the
 *     samples each, a and b: the network sees log1p of their mean where it
 *     would see log1p(beauty), and log1p(a) - log1p(b) clipped to [-2, 2]
 *     in channels 9-11. A
 *     caustic shows in both halves, a single rare sample in one, which is
 *     how the network tells a converging feature from a firefly.
 *   - Output z: radiance = expm1(max(z, 0)).
 *   - Weights and activations are bf16 with fp32 accumulation, on the matrix
 *     cores. Master weights, Adam state, loss and gradient reductions are fp32.
 *
 * Training loss, per pixel and channel, averaged (gt_denoiser_config.loss):
 *   - GT_DENOISE_LOSS_REL_L2 (the default): relative L2,
 *     (y - t)^2 / (sg(y)^2 + 0.01) with y = expm1(z) for z >= 0 (z below 0)
 *     and sg(y) = max(y, 0) held constant in the gradient. Over noisy
 *     targets its gradient vanishes at the targets' mean, so the network
 *     learns the expected radiance (Monte Carlo targets are unbiased, not
 *     symmetric: their median lies below their mean);
 *   - GT_DENOISE_LOSS_L1_TONEMAP: L1 between Reinhard-tone-mapped prediction
 *     and target, T(x) = x / (1 + x); the prediction's tone map is
 *     1 - exp(-z) for z >= 0, and z below 0. It learns the targets' median in
 *     tone-mapped space, darker than their mean where the noise is skewed.
 *
 * Images are fp32, [n][h][w][3], row-major, in device memory on the
 * context's device. h and w may be any size: the network pads to a multiple
 * of 2^(levels-1) by replicating the edge. All work runs on the context's
 * stream. Every call returns when its results are complete.
 */
#ifndef GARTUS_DENOISE_H
#define GARTUS_DENOISE_H

#include <stddef.h>
#include <stdint.h>

#include "gartus_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gt_denoiser gt_denoiser;

enum { GT_DENOISE_LOSS_REL_L2 = 0, GT_DENOISE_LOSS_L1_TONEMAP = 1 };

typedef struct gt_denoiser_config {
  uint32_t levels;        /* 2..6 */
  uint32_t channels[6];   /* features per level, multiples of 16, at most 128 */
  uint64_t seed;          /* He-normal initialisation (SplitMix64 + Box-Muller) */
  uint32_t loss;          /* GT_DENOISE_LOSS_*: the training loss */
  uint32_t beauty_halves; /* 0: d_beauty is one image [n][h][w][3]; 1: two half-sample images
                             [2][n][h][w][3] (a, then b) */
} gt_denoiser_config;

/* The standard network: 5 levels, 32-48-64-96-128 channels (1.02 M parameters),
 * relative-L2 loss. */
void gt_denoiser_default_config(gt_denoiser_config* cfg);

int gt_denoiser_create(gt_context* ctx, const gt_denoiser_config* cfg, gt_denoiser** out);
int gt_denoiser_destroy(gt_denoiser* dn);

/* Parameters in a fixed order: for each convolution (encoder levels 0..L-1,
 * two each; decoder levels L-2..0, two each; the output convolution), its
 * weights [cout][9][cin] then its bias [cout]. Counts include the zero
 * padding of the 16-channel input and output. */
int gt_denoiser_param_count(const gt_denoiser* dn, uint64_t* count);
int gt_denoiser_get_params(gt_denoiser* dn, float* host_params);
int gt_denoiser_set_params(gt_denoiser* dn, const float* host_params);

/* Denoises n images: d_out = network(d_beauty, d_albedo, d_normal). */
int gt_denoiser_infer(gt_denoiser* dn, uint32_t n, uint32_t h, uint32_t w, const float* d_beauty,
                      const float* d_albedo, const float* d_normal, float* d_out);

/* Loss and its gradient for a batch, without updating: host_grads (nullable)
 * receives d loss / d parameters in the parameter order. */
int gt_denoiser_loss_and_gradients(gt_denoiser* dn, uint32_t n, uint32_t h, uint32_t w,
                                   const float* d_beauty, const float* d_albedo, const float* d_normal,
                                   const float* d_target, float* out_loss, float* host_grads);

/* One training step: forward, loss, backward, then an Adam update with
 * learning rate lr (beta1 0.9, beta2 0.999, eps 1e-8). */
int gt_denoiser_train_step(gt_denoiser* dn, uint32_t n, uint32_t h, uint32_t w, const float* d_beauty,
                           const float* d_albedo, const float* d_normal, const float* d_target, float lr,
                           float* out_loss);

#ifdef __cplusplus
}
#endif
#endif /* GARTUS_DENOISE_H */
