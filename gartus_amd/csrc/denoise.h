// denoise.h — launch interface of the denoiser's kernels (denoise.hip).
//
// Tensors are NHWC with bf16 elements (uint16_t storage) and a channel count
// that is a multiple of 16; P = N*H*W pixels. Convolutions are 3x3, stride 1,
// zero padding 1, with weights stored [Cout][9][Cin] (tap = 3*(dy+1)+(dx+1))
// and, for the data gradient, transposed as [Cin][9][Cout].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gartus_denoise.h"

namespace gdn {

struct Dims {
  int n, h, w;
  int pixels() const { return n * h * w; }
};

// y = act(conv(x) + bias); y is bf16, or fp32 when out_f32. A bf16 output can
// also be written 2x2 max-pooled to `pooled` (rows and columns in pairs; an
// odd last row or column is dropped).
void conv_fwd(hipStream_t s, const uint16_t* x, int cin, const uint16_t* w, const float* bias, void* y, int cout,
              Dims d, bool relu, bool out_f32, uint16_t* pooled = nullptr);
// The network output convolution (cin channels -> 16, no ReLU) written straight as radiance
// expm1(max(z, 0)) of channels 0..2, fp32 [n][h][w][3] for the h x w image inside d; the same
// values as conv_fwd (out_f32) followed by unpack_output. False, and nothing launched, for shapes
// the fused kernel does not take.
bool conv_fwd_radiance(hipStream_t s, const uint16_t* x, int cin, const uint16_t* w, const float* bias, float* out,
                       int h, int w_img, Dims d);
// y = act(conv(cat(up2(a), b)) + bias) without materialising the concatenation: a has ca channels
// at half resolution (nearest 2x up), b has cb channels at d.
void conv_fwd_upcat(hipStream_t s, const uint16_t* a, int ca, const uint16_t* b, int cb, const uint16_t* w,
                    const float* bias, uint16_t* y, int cout, Dims d, bool relu);
// dx = conv^T(dy * [y > 0]) (mask only when y != null), then dx * [dx_mask > 0] when dx_mask != null
// (the ReLU of the layer that produced the convolution's input, applied where its gradient is made).
void conv_dgrad(hipStream_t s, const uint16_t* dy, const uint16_t* y, int cout, const uint16_t* wt, uint16_t* dx,
                int cin, Dims d, const uint16_t* dx_mask = nullptr);
// dW[co][t][ci] = sum_p dy[p][co] x[p + off_t][ci], dB[co] = sum_p dy[p][co] (dy already masked by the
// ReLU of the forward output where it was made); summed in a fixed order (per-split partials in
// `scratch`, then reduced).
// With xa, x is the decoder's concatenation [up2(xa) (ca channels), x (cin - ca channels)] read in
// place, as conv_fwd_upcat reads it.
void conv_wgrad(hipStream_t s, const uint16_t* x, int cin, const uint16_t* dy, int cout, Dims d, float* dw, float* db,
                float* scratch, size_t scratch_floats, const uint16_t* xa = nullptr, int ca = 0);
size_t wgrad_scratch_floats(int cin, int cout, Dims d);

// dx = route dy to each window's first maximum, plus add[:, add_off : add_off + c] when add != null
// (add has add_c channels at the full resolution); with relu_mask, zero where x is not positive.
void pool_bwd(hipStream_t s, const uint16_t* x, const uint16_t* dy, const uint16_t* add, int add_c, int add_off,
              uint16_t* dx, int c, Dims d, bool relu_mask = false);
// da = sum over the 4 children of dout[:, 0:ca] (dout has ca + cb channels), zero where
// da_mask (same layout as da) is not positive when da_mask != null.
void upcat_bwd(hipStream_t s, const uint16_t* dout, int ca, int cb, uint16_t* da, Dims d,
               const uint16_t* da_mask = nullptr);

// x (16 channels, padded dims dp) from fp32 images [n][h][w][3]: log1p(max(beauty, 0)), albedo,
// normal, zeros; pixels beyond (h, w) replicate the edge.
// With beauty_b (two half-sample images a = beauty, b = beauty_b): log1p of their mean, and
// log1p(a) - log1p(b) in channels 9..11.
void pack_input(hipStream_t s, const float* beauty, const float* beauty_b, const float* albedo, const float* normal,
                int h, int w, uint16_t* x, Dims dp);
// Training loss, mean over n*h*w*3 elements; dz (bf16, 16 channels) gets d loss / d z, zero in
// padding; block partial sums go to partial[0..blocks). kind (GT_DENOISE_LOSS_*):
//   relative L2: (y - t)^2 / (max(y, 0)^2 + 0.01), y = expm1(z) for z >= 0 and z below, the
//     denominator held constant in the gradient;
//   L1 on Reinhard-tone-mapped radiance: |T(z) - t/(1+t)|, T(z) = 1 - exp(-z) for z >= 0 and z below.
int loss_blocks(Dims dp);
void loss(hipStream_t s, int kind, const float* z, const float* target, int h, int w, uint16_t* dz, float* partial,
          Dims dp);
// radiance = expm1(max(z, 0)) for the valid region, fp32 [n][h][w][3].
void unpack_output(hipStream_t s, const float* z, float* out, int h, int w, Dims dp);

// A convolution's weights [w_off, w_off + cout*9*cin) in the parameter vector, with their
// bf16 copies.
struct AdamSeg {
  size_t w_off;
  int cin, cout;
  uint16_t* w16;
  uint16_t* wt16;
};
// Adam on fp32 master weights, over the whole parameter vector p[0, n) in one launch; the
// weights of each of the nseg segments (segs in device memory) also get their bf16 copies
// (w16 [cout][9][cin] and wt16 [cin][9][cout]).
void adam_all(hipStream_t s, float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
              float eps, float c1, float c2, const AdamSeg* segs, int nseg);
// bf16 copies of fp32 weights (after loading parameters).
void to_bf16(hipStream_t s, const float* p, uint16_t* w16, uint16_t* wt16, int cin, int cout);

}  // namespace gdn
