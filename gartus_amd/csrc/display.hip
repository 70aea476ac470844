// display.hip — the output boundary: linear RGB images to the display and
// file encodings of src/graphics/display.rs:164-468 and colors.rs:388-422,
// one thread per pixel. The arithmetic is the reference's, in FP64 with no
// contraction, so the bytes equal the reference's for the same input values.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/gartus_hip.h"
#include "status.h"

namespace {

struct ToneArgs {
  double exposure, inv_gamma;
  uint32_t op;
  uint32_t gamma2;  // |gamma - 2| <= f64::EPSILON: the sqrt form (display.rs:269)
};

// Rust's `f64 as u8`: truncation toward zero, saturating, NaN -> 0.
__device__ inline uint32_t as_u8(double x) {
  return (isnan(x) || x <= 0.0) ? 0u : (x >= 255.0 ? 255u : uint32_t(x));
}
// f64::clamp: a NaN stays NaN.
__device__ inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// ToneMap::map_color for one channel (display.rs:264-277), with
// ToneMappingOperator::map_component (:173-189).
__device__ inline uint32_t tone_channel(double c, const ToneArgs& t) {
#pragma clang fp contract(off)
  double v = c * t.exposure;
  v = isfinite(v) ? fmax(v, 0.0) : 0.0;
  if (t.op == GT_TONE_REINHARD) {
    v = v / (1.0 + v);
  } else if (t.op == GT_TONE_ACES) {
    v = clampd(v * (2.51 * v + 0.03) / (v * (2.43 * v + 0.59) + 0.14), 0.0, 1.0);
  }
  double g;
  if (t.gamma2) g = isnan(v) ? 0.0 : (v > 0.0 ? sqrt(v) : 0.0);  // linear_to_gamma_component
  else g = v <= 0.0 ? 0.0 : pow(v, t.inv_gamma);
  return as_u8(256.0 * clampd(g, 0.0, 0.999));
}

// Rgb::from_raw_linear_color for one channel (colors.rs:412-422).
__device__ inline uint32_t raw_channel(double c) {
#pragma clang fp contract(off)
  return as_u8(255.999 * clampd(isnan(c) ? 0.0 : c, 0.0, 1.0));
}

// encode_rgbe (display.rs:444-468) as bytes r | g << 8 | b << 16 | e << 24.
// The shared exponent is floor(log2(max)) + 1 as the reference computes it;
// its 2^-e is exact for e < 1024 and, as f64::powi evaluates it (the
// reciprocal of an overflowed 2^e), zero from 1024 on.
__device__ inline uint32_t rgbe(double r, double g, double b) {
#pragma clang fp contract(off)
  r = isfinite(r) ? fmax(r, 0.0) : 0.0;
  g = isfinite(g) ? fmax(g, 0.0) : 0.0;
  b = isfinite(b) ? fmax(b, 0.0) : 0.0;
  const double m = fmax(fmax(r, g), b);
  if (m <= 1.0e-32) return 0u;
  const int e = int(floor(log2(m))) + 1;
  const double scale = e >= 1024 ? 0.0 : ldexp(256.0, -e);
  const uint32_t eb = uint32_t(min(max(e + 128, 0), 255));
  return as_u8(clampd(r * scale, 0.0, 255.0)) | (as_u8(clampd(g * scale, 0.0, 255.0)) << 8) |
         (as_u8(clampd(b * scale, 0.0, 255.0)) << 16) | (eb << 24);
}

template <typename T, int FMT>
__global__ __launch_bounds__(256) void k_encode(const T* __restrict__ src, uint32_t w, uint32_t h, ToneArgs t,
                                                unsigned char* __restrict__ dst) {
  const size_t n = size_t(w) * h;
  for (size_t p = blockIdx.x * size_t(blockDim.x) + threadIdx.x; p < n; p += size_t(gridDim.x) * blockDim.x) {
    if constexpr (FMT == GT_ENCODE_PFM) {
      // output row y holds image row h - 1 - y (display.rs:371)
      const size_t y = p / w, x = p - y * w;
      const T* s = src + ((size_t(h) - 1 - y) * w + x) * 3;
      float* o = reinterpret_cast<float*>(dst) + p * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double v = double(s[k]);
        o[k] = isfinite(v) ? float(v) : 0.0f;
      }
    } else {
      const T* s = src + p * 3;
      const double c0 = double(s[0]), c1 = double(s[1]), c2 = double(s[2]);
      if constexpr (FMT == GT_ENCODE_RGBE) {
        const uint32_t v = rgbe(c0, c1, c2);
        unsigned char* o = dst + p * 4;
        o[0] = v & 255u;
        o[1] = (v >> 8) & 255u;
        o[2] = (v >> 16) & 255u;
        o[3] = v >> 24;
      } else {
        unsigned char* o = dst + p * 3;
        if constexpr (FMT == GT_ENCODE_CANVAS) {
          o[0] = tone_channel(c0, t);
          o[1] = tone_channel(c1, t);
          o[2] = tone_channel(c2, t);
        } else {
          o[0] = raw_channel(c0);
          o[1] = raw_channel(c1);
          o[2] = raw_channel(c2);
        }
      }
    }
  }
}

template <typename T>
hipError_t launch(int format, const void* src, uint32_t w, uint32_t h, const ToneArgs& t, void* dst,
                  hipStream_t s) {
  const size_t n = size_t(w) * h;
  const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, size_t(1) << 20));
  const T* in = static_cast<const T*>(src);
  unsigned char* out = static_cast<unsigned char*>(dst);
  switch (format) {
    case GT_ENCODE_CANVAS: hipLaunchKernelGGL((k_encode<T, GT_ENCODE_CANVAS>), grid, 256, 0, s, in, w, h, t, out); break;
    case GT_ENCODE_RAW: hipLaunchKernelGGL((k_encode<T, GT_ENCODE_RAW>), grid, 256, 0, s, in, w, h, t, out); break;
    case GT_ENCODE_RGBE: hipLaunchKernelGGL((k_encode<T, GT_ENCODE_RGBE>), grid, 256, 0, s, in, w, h, t, out); break;
    default: hipLaunchKernelGGL((k_encode<T, GT_ENCODE_PFM>), grid, 256, 0, s, in, w, h, t, out); break;
  }
  return hipGetLastError();
}

std::string header_text(int format, uint32_t w, uint32_t h) {
  char buf[96];
  if (format == GT_ENCODE_PFM) std::snprintf(buf, sizeof buf, "PF\n%u %u\n-1.0\n", w, h);
  else if (format == GT_ENCODE_RGBE) std::snprintf(buf, sizeof buf, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
  else buf[0] = 0;
  return buf;
}

bool known_format(int f) { return f >= GT_ENCODE_CANVAS && f <= GT_ENCODE_PFM; }

}  // namespace

extern "C" {

uint64_t gt_encoded_size(int format, uint32_t width, uint32_t height) {
  const uint64_t n = uint64_t(width) * height;
  switch (format) {
    case GT_ENCODE_CANVAS:
    case GT_ENCODE_RAW: return 3 * n;
    case GT_ENCODE_RGBE: return 4 * n;
    case GT_ENCODE_PFM: return 12 * n;
    default: return 0;
  }
}

int gt_encoded_header(int format, uint32_t width, uint32_t height, char* buf, size_t cap, size_t* out_len) {
  if (!known_format(format)) return gt::set_error(GT_EINVAL, "unknown encoding format");
  const std::string s = header_text(format, width, height);
  if (buf && cap) std::memcpy(buf, s.data(), std::min(cap, s.size()));
  if (out_len) *out_len = s.size();
  return GT_OK;
}

int gt_encode_image(gt_context* ctx, const void* src, int src_is_double, int src_on_device, uint32_t width,
                    uint32_t height, int format, const gt_tone_map* tone_map, void* dst, int dst_on_device) {
  if (!ctx) return gt::set_error(GT_EINVAL, "context is null");
  if (!known_format(format)) return gt::set_error(GT_EINVAL, "unknown encoding format");
  gt_tone_map tm{1.0, 2.0, GT_TONE_NONE, 0};
  if (tone_map) tm = *tone_map;
  if (!(std::isfinite(tm.exposure) && tm.exposure >= 0.0))
    return gt::set_error(GT_EINVAL, "tone-map exposure must be finite and non-negative");
  if (!(std::isfinite(tm.gamma) && tm.gamma > 0.0))
    return gt::set_error(GT_EINVAL, "tone-map gamma must be positive and finite");
  if (tm.op > GT_TONE_ACES) return gt::set_error(GT_EINVAL, "unknown tone-mapping operator");
  const size_t n = size_t(width) * height;
  if (n == 0) return GT_OK;
  if (n > (size_t(1) << 36)) return gt::set_error(GT_EINVAL, "image too large");
  if (!src || !dst) return gt::set_error(GT_EINVAL, "image buffer is null");
  const size_t elem = src_is_double ? sizeof(double) : sizeof(float);
  if (reinterpret_cast<uintptr_t>(src) % elem) return gt::set_error(GT_EINVAL, "source is not aligned to its element");
  if (format == GT_ENCODE_PFM && reinterpret_cast<uintptr_t>(dst) % sizeof(float))
    return gt::set_error(GT_EINVAL, "PFM destination must be 4-byte aligned");
  const size_t in_bytes = n * 3 * elem, out_bytes = gt_encoded_size(format, width, height);
  GT_CHECK_HIP(hipSetDevice(gt::context_device(ctx)));
  hipStream_t s = gt::context_stream(ctx);
  const void* din = src;
  void* dout = dst;
  if (!src_on_device) {
    void* p = nullptr;
    if (int rc = gt::context_scratch(ctx, 0, in_bytes, &p)) return rc;
    GT_CHECK_HIP(hipMemcpyAsync(p, src, in_bytes, hipMemcpyHostToDevice, s));
    din = p;
  }
  if (!dst_on_device) {
    if (int rc = gt::context_scratch(ctx, 1, out_bytes, &dout)) return rc;
  }
  ToneArgs t{tm.exposure, 1.0 / tm.gamma, tm.op, std::fabs(tm.gamma - 2.0) <= 2.220446049250313e-16 ? 1u : 0u};
  GT_CHECK_HIP(src_is_double ? launch<double>(format, din, width, height, t, dout, s)
                             : launch<float>(format, din, width, height, t, dout, s));
  if (!dst_on_device) GT_CHECK_HIP(hipMemcpyAsync(dst, dout, out_bytes, hipMemcpyDeviceToHost, s));
  GT_CHECK_HIP(hipStreamSynchronize(s));
  return GT_OK;
}

}  // extern "C"
