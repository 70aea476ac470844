// denoise_api.cpp — the learned denoiser's network runtime and the C ABI of
// include/gartus_denoise.h: parameter arena, activation and gradient
// buffers, the forward and backward sequences of the U-Net over the kernels
// of denoise.hip, and Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gartus_denoise.h"
#include "denoise.h"
#include "status.h"

namespace {

template <typename T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
};

constexpr int kIn = 16;  // padded input channels (9 used)
constexpr int kOut = 16; // padded output channels (3 used)
constexpr float kBeta1 = 0.9f, kBeta2 = 0.999f, kAdamEps = 1e-8f;

struct Conv {
  int cin = 0, cout = 0;
  bool relu = true;
  size_t w_off = 0, b_off = 0;  // into the parameter arena
  Buf<uint16_t> w16, wt16;      // bf16 copies: [cout][9][cin] and [cin][9][cout]
  size_t nw() const { return size_t(cout) * 9 * cin; }
};

// Per-level activations and their gradients (level l has dims d[l]).
struct Level {
  Buf<uint16_t> in;         // pooled input (l > 0)
  Buf<uint16_t> enc_a, enc; // encoder convolutions
  Buf<uint16_t> dec_a, dec; // decoder convolutions (l < L-1)
  Buf<uint16_t> g_in, g_enc_a, g_enc, g_cat, g_dec_a, g_dec;
};

}  // namespace

struct gt_denoiser {
  gt_context* ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  int levels = 0;
  int ch[6] = {0};
  int loss_kind = GT_DENOISE_LOSS_REL_L2;
  bool beauty_halves = false;
  std::vector<std::unique_ptr<Conv>> convs;  // enc l: 2l, 2l+1; dec l: 2L + 2(L-2-l), +1; output: last
  size_t nparams = 0;
  Buf<float> params, grads, adam_m, adam_v;
  Buf<gdn::AdamSeg> adam_segs;  // every convolution's weights and bf16 copies, for the one Adam launch
  uint64_t step = 0;
  // shape-dependent buffers
  gdn::Dims shape{0, 0, 0};  // padded level-0 dims the buffers were sized for
  std::vector<std::unique_ptr<Level>> lv;
  Buf<uint16_t> x0, dz;
  Buf<float> z, scratch, loss_partial;
  std::vector<float> host_loss;

  Conv& enc_conv(int l, int k) { return *convs[2 * l + k]; }
  Conv& dec_conv(int l, int k) { return *convs[2 * levels + 2 * (levels - 2 - l) + k]; }
  Conv& out_conv() { return *convs.back(); }
};

namespace {

int bad(int code, const std::string& m) { return gt::set_error(code, m); }

// He-normal initialisation from SplitMix64 + Box-Muller, in parameter order.
void init_params(gt_denoiser* dn, std::vector<float>& host, uint64_t seed) {
  uint64_t state = seed;
  auto next = [&state]() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto uniform = [&next]() { return (double(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
  host.assign(dn->nparams, 0.0f);
  for (size_t c = 0; c < dn->convs.size(); ++c) {
    Conv& cv = *dn->convs[c];
    const bool first = c == 0, last = c + 1 == dn->convs.size();
    const int real_in = first ? 9 : cv.cin;
    const int real_out = last ? 3 : cv.cout;
    const double sd = std::sqrt(2.0 / (9.0 * real_in));
    for (int co = 0; co < cv.cout; ++co)
      for (int t = 0; t < 9; ++t)
        for (int ci = 0; ci < cv.cin; ++ci) {
          const double u1 = uniform(), u2 = uniform();
          const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
          const bool live = ci < real_in && co < real_out;
          host[cv.w_off + (size_t(co) * 9 + t) * cv.cin + ci] = live ? float(g * sd * (last ? 0.1 : 1.0)) : 0.0f;
        }
  }
}

int upload_params(gt_denoiser* dn, const float* host) {
  GT_CHECK_HIP(hipMemcpyAsync(dn->params.p, host, dn->nparams * sizeof(float), hipMemcpyHostToDevice, dn->stream));
  for (auto& c : dn->convs) gdn::to_bf16(dn->stream, dn->params.p + c->w_off, c->w16.p, c->wt16.p, c->cin, c->cout);
  GT_CHECK_HIP(hipGetLastError());
  GT_CHECK_HIP(hipStreamSynchronize(dn->stream));
  return GT_OK;
}

int ensure_shape(gt_denoiser* dn, gdn::Dims d0) {
  if (d0.n == dn->shape.n && d0.h == dn->shape.h && d0.w == dn->shape.w) return GT_OK;
  const int L = dn->levels;
  dn->lv.clear();
  size_t scratch = 0;
  for (int l = 0; l < L; ++l) {
    auto lv = std::make_unique<Level>();
    gdn::Dims d{d0.n, d0.h >> l, d0.w >> l};
    const size_t P = size_t(d.pixels());
    const int c = dn->ch[l];
    if (l > 0) {
      GT_CHECK_HIP(lv->in.ensure(P * dn->ch[l - 1]));
      GT_CHECK_HIP(lv->g_in.ensure(P * dn->ch[l - 1]));
    }
    GT_CHECK_HIP(lv->enc_a.ensure(P * c));
    GT_CHECK_HIP(lv->enc.ensure(P * c));
    GT_CHECK_HIP(lv->g_enc_a.ensure(P * c));
    GT_CHECK_HIP(lv->g_enc.ensure(P * c));
    if (l < L - 1) {
      const int cc = dn->ch[l + 1] + c;
      GT_CHECK_HIP(lv->g_cat.ensure(P * cc));
      GT_CHECK_HIP(lv->dec_a.ensure(P * c));
      GT_CHECK_HIP(lv->dec.ensure(P * c));
      GT_CHECK_HIP(lv->g_dec_a.ensure(P * c));
      GT_CHECK_HIP(lv->g_dec.ensure(P * c));
      scratch = std::max(scratch, gdn::wgrad_scratch_floats(cc, c, d));
      scratch = std::max(scratch, gdn::wgrad_scratch_floats(c, c, d));
    }
    scratch = std::max(scratch, gdn::wgrad_scratch_floats(l == 0 ? kIn : dn->ch[l - 1], c, d));
    scratch = std::max(scratch, gdn::wgrad_scratch_floats(c, c, d));
    dn->lv.push_back(std::move(lv));
  }
  scratch = std::max(scratch, gdn::wgrad_scratch_floats(dn->ch[0], kOut, d0));
  const size_t P0 = size_t(d0.pixels());
  GT_CHECK_HIP(dn->x0.ensure(P0 * kIn));
  GT_CHECK_HIP(dn->dz.ensure(P0 * kOut));
  GT_CHECK_HIP(dn->z.ensure(P0 * kOut));
  GT_CHECK_HIP(dn->scratch.ensure(scratch));
  GT_CHECK_HIP(dn->loss_partial.ensure(size_t(gdn::loss_blocks(d0))));
  dn->shape = d0;
  return GT_OK;
}

gdn::Dims level_dims(const gt_denoiser* dn, int l) { return {dn->shape.n, dn->shape.h >> l, dn->shape.w >> l}; }

int prepare(gt_denoiser* dn, uint32_t n, uint32_t h, uint32_t w, const float* b, const float* a, const float* nr) {
  if (!dn) return bad(GT_EINVAL, "denoiser is null");
  if (!b || !a || !nr) return bad(GT_EINVAL, "null image");
  if (n == 0 || h == 0 || w == 0) return bad(GT_EINVAL, "empty batch");
  const uint32_t m = 1u << (dn->levels - 1);
  const uint64_t hp = (uint64_t(h) + m - 1) / m * m, wp = (uint64_t(w) + m - 1) / m * m;
  if (uint64_t(n) * hp * wp > (uint64_t(1) << 28)) return bad(GT_EINVAL, "batch too large (2^28 padded pixels)");
  GT_CHECK_HIP(hipSetDevice(dn->device));
  int rc = ensure_shape(dn, gdn::Dims{int(n), int(hp), int(wp)});
  if (rc != GT_OK) return rc;
  const float* b2 = dn->beauty_halves ? b + size_t(n) * h * w * 3 : nullptr;
  gdn::pack_input(dn->stream, b, b2, a, nr, int(h), int(w), dn->x0.p, dn->shape);
  return GT_OK;
}

// The decoder's concatenations are read in place from their two halves (here
// and by the weight gradient), never written. With radiance, the output
// convolution writes the h x w radiance there directly when it can (true)
// instead of z (false).
bool forward(gt_denoiser* dn, float* radiance = nullptr, int h = 0, int w = 0) {
  hipStream_t s = dn->stream;
  const int L = dn->levels;
  for (int l = 0; l < L; ++l) {
    Level& v = *dn->lv[l];
    const gdn::Dims d = level_dims(dn, l);
    const uint16_t* in = l == 0 ? dn->x0.p : v.in.p;
    Conv& c0 = dn->enc_conv(l, 0);
    Conv& c1 = dn->enc_conv(l, 1);
    gdn::conv_fwd(s, in, c0.cin, c0.w16.p, dn->params.p + c0.b_off, v.enc_a.p, c0.cout, d, true, false);
    // the second convolution also writes the next level's input, pooled
    gdn::conv_fwd(s, v.enc_a.p, c1.cin, c1.w16.p, dn->params.p + c1.b_off, v.enc.p, c1.cout, d, true, false,
                  l + 1 < L ? dn->lv[l + 1]->in.p : nullptr);
  }
  for (int l = L - 2; l >= 0; --l) {
    Level& v = *dn->lv[l];
    const gdn::Dims d = level_dims(dn, l);
    const uint16_t* up = l == L - 2 ? dn->lv[L - 1]->enc.p : dn->lv[l + 1]->dec.p;
    Conv& c0 = dn->dec_conv(l, 0);
    Conv& c1 = dn->dec_conv(l, 1);
    gdn::conv_fwd_upcat(s, up, dn->ch[l + 1], v.enc.p, dn->ch[l], c0.w16.p, dn->params.p + c0.b_off, v.dec_a.p,
                        c0.cout, d, true);
    gdn::conv_fwd(s, v.dec_a.p, c1.cin, c1.w16.p, dn->params.p + c1.b_off, v.dec.p, c1.cout, d, true, false);
  }
  Conv& co = dn->out_conv();
  if (radiance && gdn::conv_fwd_radiance(s, dn->lv[0]->dec.p, co.cin, co.w16.p, dn->params.p + co.b_off, radiance, h,
                                         w, dn->shape))
    return true;
  gdn::conv_fwd(s, dn->lv[0]->dec.p, co.cin, co.w16.p, dn->params.p + co.b_off, dn->z.p, co.cout, dn->shape, false,
                true);
  return false;
}

// Gradients of every parameter into dn->grads, from dn->dz.
void backward(gt_denoiser* dn) {
  hipStream_t s = dn->stream;
  const int L = dn->levels;
  float* G = dn->grads.p;
  float* S = dn->scratch.p;
  const size_t SN = dn->scratch.n;
  auto wgrad = [&](Conv& c, const uint16_t* x, const uint16_t* dy, gdn::Dims d, const uint16_t* xa = nullptr,
                   int ca = 0) {
    gdn::conv_wgrad(s, x, c.cin, dy, c.cout, d, G + c.w_off, G + c.b_off, S, SN, xa, ca);
  };
  // Each gradient is masked by the ReLU of the tensor it belongs to where it is made (the data
  // gradient, upsample or pooling backward that writes it), so the weight and data gradients
  // that read it need no mask.
  // output convolution (no ReLU)
  Conv& co = dn->out_conv();
  wgrad(co, dn->lv[0]->dec.p, dn->dz.p, dn->shape);
  gdn::conv_dgrad(s, dn->dz.p, nullptr, co.cout, co.wt16.p, dn->lv[0]->g_dec.p, co.cin, dn->shape, dn->lv[0]->dec.p);
  // decoder, nearest the output first
  for (int l = 0; l <= L - 2; ++l) {
    Level& v = *dn->lv[l];
    const gdn::Dims d = level_dims(dn, l);
    Conv& c1 = dn->dec_conv(l, 1);
    Conv& c0 = dn->dec_conv(l, 0);
    wgrad(c1, v.dec_a.p, v.g_dec.p, d);
    gdn::conv_dgrad(s, v.g_dec.p, nullptr, c1.cout, c1.wt16.p, v.g_dec_a.p, c1.cin, d, v.dec_a.p);
    const uint16_t* up = l == L - 2 ? dn->lv[L - 1]->enc.p : dn->lv[l + 1]->dec.p;
    wgrad(c0, v.enc.p, v.g_dec_a.p, d, up, dn->ch[l + 1]);
    gdn::conv_dgrad(s, v.g_dec_a.p, nullptr, c0.cout, c0.wt16.p, v.g_cat.p, c0.cin, d);
    uint16_t* g_up = l == L - 2 ? dn->lv[L - 1]->g_enc.p : dn->lv[l + 1]->g_dec.p;
    gdn::upcat_bwd(s, v.g_cat.p, dn->ch[l + 1], dn->ch[l], g_up, d, up);
  }
  // encoder, deepest first
  for (int l = L - 1; l >= 0; --l) {
    Level& v = *dn->lv[l];
    const gdn::Dims d = level_dims(dn, l);
    if (l < L - 1)  // pooled path plus the skip half of the concatenation
      gdn::pool_bwd(s, v.enc.p, dn->lv[l + 1]->g_in.p, v.g_cat.p, dn->ch[l + 1] + dn->ch[l], dn->ch[l + 1],
                    v.g_enc.p, dn->ch[l], d, true);
    Conv& c1 = dn->enc_conv(l, 1);
    Conv& c0 = dn->enc_conv(l, 0);
    wgrad(c1, v.enc_a.p, v.g_enc.p, d);
    gdn::conv_dgrad(s, v.g_enc.p, nullptr, c1.cout, c1.wt16.p, v.g_enc_a.p, c1.cin, d, v.enc_a.p);
    const uint16_t* in = l == 0 ? dn->x0.p : v.in.p;
    wgrad(c0, in, v.g_enc_a.p, d);
    if (l > 0) gdn::conv_dgrad(s, v.g_enc_a.p, nullptr, c0.cout, c0.wt16.p, v.g_in.p, c0.cin, d);
  }
}

int loss_into(gt_denoiser* dn, const float* target, uint32_t h, uint32_t w, float* out_loss) {
  gdn::loss(dn->stream, dn->loss_kind, dn->z.p, target, int(h), int(w), dn->dz.p, dn->loss_partial.p, dn->shape);
  const int nb = gdn::loss_blocks(dn->shape);
  dn->host_loss.resize(size_t(nb));
  GT_CHECK_HIP(hipMemcpyAsync(dn->host_loss.data(), dn->loss_partial.p, size_t(nb) * sizeof(float),
                              hipMemcpyDeviceToHost, dn->stream));
  GT_CHECK_HIP(hipStreamSynchronize(dn->stream));
  double sum = 0.0;
  for (float v : dn->host_loss) sum += v;
  if (out_loss) *out_loss = float(sum / (double(dn->shape.n) * h * w * 3));
  return GT_OK;
}

bool supported(int cin, int cout) {
  return cin % 16 == 0 && cout % 16 == 0 && cin > 0 && cout > 0 && cin <= 256 && cout <= 128;
}

}  // namespace

extern "C" {

void gt_denoiser_default_config(gt_denoiser_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof *cfg);
  cfg->levels = 5;
  const uint32_t ch[5] = {32, 48, 64, 96, 128};
  for (int i = 0; i < 5; ++i) cfg->channels[i] = ch[i];
  cfg->seed = 1;
  cfg->loss = GT_DENOISE_LOSS_REL_L2;
}

int gt_denoiser_create(gt_context* ctx, const gt_denoiser_config* cfg, gt_denoiser** out) {
  if (!ctx || !cfg || !out) return bad(GT_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->levels < 2 || cfg->levels > 6) return bad(GT_EINVAL, "levels must be 2..6");
  for (uint32_t l = 0; l < cfg->levels; ++l)
    if (cfg->channels[l] == 0 || cfg->channels[l] % 16 || cfg->channels[l] > 128)
      return bad(GT_EINVAL, "channels must be multiples of 16 in 16..128");
  if (cfg->loss != GT_DENOISE_LOSS_REL_L2 && cfg->loss != GT_DENOISE_LOSS_L1_TONEMAP)
    return bad(GT_EINVAL, "unknown loss");
  if (cfg->beauty_halves > 1) return bad(GT_EINVAL, "beauty_halves must be 0 or 1");
  std::unique_ptr<gt_denoiser> dn(new gt_denoiser());
  dn->loss_kind = int(cfg->loss);
  dn->beauty_halves = cfg->beauty_halves != 0;
  dn->ctx = ctx;
  dn->device = gt::context_device(ctx);
  dn->stream = gt::context_stream(ctx);
  dn->levels = int(cfg->levels);
  for (int l = 0; l < dn->levels; ++l) dn->ch[l] = int(cfg->channels[l]);
  const int L = dn->levels;
  auto add = [&](int cin, int cout, bool relu) {
    auto c = std::make_unique<Conv>();
    c->cin = cin;
    c->cout = cout;
    c->relu = relu;
    dn->convs.push_back(std::move(c));
  };
  for (int l = 0; l < L; ++l) {
    add(l == 0 ? kIn : dn->ch[l - 1], dn->ch[l], true);
    add(dn->ch[l], dn->ch[l], true);
  }
  for (int l = L - 2; l >= 0; --l) {
    add(dn->ch[l + 1] + dn->ch[l], dn->ch[l], true);
    add(dn->ch[l], dn->ch[l], true);
  }
  add(dn->ch[0], kOut, false);
  size_t off = 0;
  for (auto& c : dn->convs) {
    if (!supported(c->cin, c->cout)) return bad(GT_EUNSUPPORTED, "unsupported convolution shape");
    c->w_off = off;
    off += c->nw();
    c->b_off = off;
    off += size_t(c->cout);
  }
  dn->nparams = off;
  GT_CHECK_HIP(hipSetDevice(dn->device));
  GT_CHECK_HIP(dn->params.ensure(off));
  GT_CHECK_HIP(dn->grads.ensure(off));
  GT_CHECK_HIP(dn->adam_m.ensure(off));
  GT_CHECK_HIP(dn->adam_v.ensure(off));
  GT_CHECK_HIP(hipMemsetAsync(dn->adam_m.p, 0, off * sizeof(float), dn->stream));
  GT_CHECK_HIP(hipMemsetAsync(dn->adam_v.p, 0, off * sizeof(float), dn->stream));
  std::vector<gdn::AdamSeg> segs;
  for (auto& c : dn->convs) {
    GT_CHECK_HIP(c->w16.ensure(c->nw()));
    GT_CHECK_HIP(c->wt16.ensure(c->nw()));
    segs.push_back(gdn::AdamSeg{c->w_off, c->cin, c->cout, c->w16.p, c->wt16.p});
  }
  GT_CHECK_HIP(dn->adam_segs.ensure(segs.size()));
  GT_CHECK_HIP(hipMemcpy(dn->adam_segs.p, segs.data(), segs.size() * sizeof(gdn::AdamSeg), hipMemcpyHostToDevice));
  std::vector<float> host;
  init_params(dn.get(), host, cfg->seed);
  int rc = upload_params(dn.get(), host.data());
  if (rc != GT_OK) return rc;
  *out = dn.release();
  return GT_OK;
}

int gt_denoiser_destroy(gt_denoiser* dn) {
  if (!dn) return GT_OK;
  (void)hipSetDevice(dn->device);
  (void)hipStreamSynchronize(dn->stream);
  delete dn;
  return GT_OK;
}

int gt_denoiser_param_count(const gt_denoiser* dn, uint64_t* count) {
  if (!dn || !count) return bad(GT_EINVAL, "null argument");
  *count = dn->nparams;
  return GT_OK;
}

int gt_denoiser_get_params(gt_denoiser* dn, float* host) {
  if (!dn || !host) return bad(GT_EINVAL, "null argument");
  GT_CHECK_HIP(hipSetDevice(dn->device));
  GT_CHECK_HIP(hipMemcpyAsync(host, dn->params.p, dn->nparams * sizeof(float), hipMemcpyDeviceToHost, dn->stream));
  GT_CHECK_HIP(hipStreamSynchronize(dn->stream));
  return GT_OK;
}

int gt_denoiser_set_params(gt_denoiser* dn, const float* host) {
  if (!dn || !host) return bad(GT_EINVAL, "null argument");
  GT_CHECK_HIP(hipSetDevice(dn->device));
  return upload_params(dn, host);
}

int gt_denoiser_infer(gt_denoiser* dn, uint32_t n, uint32_t h, uint32_t w, const float* b, const float* a,
                      const float* nr, float* out) {
  if (!out) return bad(GT_EINVAL, "null output");
  int rc = prepare(dn, n, h, w, b, a, nr);
  if (rc != GT_OK) return rc;
  if (!forward(dn, out, int(h), int(w)))
    gdn::unpack_output(dn->stream, dn->z.p, out, int(h), int(w), dn->shape);
  GT_CHECK_HIP(hipGetLastError());
  GT_CHECK_HIP(hipStreamSynchronize(dn->stream));
  return GT_OK;
}

int gt_denoiser_loss_and_gradients(gt_denoiser* dn, uint32_t n, uint32_t h, uint32_t w, const float* b,
                                   const float* a, const float* nr, const float* target, float* out_loss,
                                   float* host_grads) {
  if (!target) return bad(GT_EINVAL, "null target");
  int rc = prepare(dn, n, h, w, b, a, nr);
  if (rc != GT_OK) return rc;
  forward(dn);
  if ((rc = loss_into(dn, target, h, w, out_loss)) != GT_OK) return rc;
  backward(dn);
  GT_CHECK_HIP(hipGetLastError());
  if (host_grads)
    GT_CHECK_HIP(hipMemcpyAsync(host_grads, dn->grads.p, dn->nparams * sizeof(float), hipMemcpyDeviceToHost,
                                dn->stream));
  GT_CHECK_HIP(hipStreamSynchronize(dn->stream));
  return GT_OK;
}

int gt_denoiser_train_step(gt_denoiser* dn, uint32_t n, uint32_t h, uint32_t w, const float* b, const float* a,
                           const float* nr, const float* target, float lr, float* out_loss) {
  if (!target) return bad(GT_EINVAL, "null target");
  if (!(std::isfinite(lr) && lr > 0.0f)) return bad(GT_EINVAL, "learning rate must be positive and finite");
  int rc = prepare(dn, n, h, w, b, a, nr);
  if (rc != GT_OK) return rc;
  forward(dn);
  if ((rc = loss_into(dn, target, h, w, out_loss)) != GT_OK) return rc;
  backward(dn);
  dn->step += 1;
  const float c1 = float(1.0 - std::pow(double(kBeta1), double(dn->step)));
  const float c2 = float(1.0 - std::pow(double(kBeta2), double(dn->step)));
  gdn::adam_all(dn->stream, dn->params.p, dn->grads.p, dn->adam_m.p, dn->adam_v.p, dn->nparams, lr, kBeta1, kBeta2,
                kAdamEps, c1, c2, dn->adam_segs.p, int(dn->convs.size()));
  GT_CHECK_HIP(hipGetLastError());
  GT_CHECK_HIP(hipStreamSynchronize(dn->stream));
  return GT_OK;
}

}  // extern "C"
