// api.cpp — the C ABI of include/gartus_hip.h.
//
// Host side of the MI355X path tracer: objects mirror RayScene,
// SamplingTargetList, EnvironmentLight and RayCamera; gt_render_* drives the
// wavefront kernels of kernels.hip in batches of whole-image samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gartus_hip.h"
#include "bvh_build.h"
#include "kernels.h"
#include "scene_build.h"
#include "status.h"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define GT_HIP(expr)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail(GT_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr double kEps = std::numeric_limits<double>::epsilon();

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
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

}  // namespace

struct gt_context {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t cus = 0;
  gt_tuning tuning{};
  // wavefront state: per-path records and path-id queues
  DevBuf<float4> ray_o, ray_d, thr, s_d[2], s_w[2], radiance;
  DevBuf<float2> hits;
  DevBuf<uint32_t> active[2], cat[gtk::kNumCategories], shadow[2];
  DevBuf<uint32_t> ctr;
  DevBuf<unsigned long long> tally;
  DevBuf<uint2> gstack;
  DevBuf<double4> acc;
  DevBuf<float> staging;           // device results of the host-output entry points
  DevBuf<unsigned char> scratch[2];  // staging of the other translation units (gt::context_scratch)
  // adaptive sampling: per-pixel Welford state, stop flags, active pixel lists
  DevBuf<double4> ad_mean, ad_m2;
  DevBuf<uint32_t> ad_done, ad_map[2], ad_count;
  DevBuf<uint32_t> band_map;  // progressive renders: the pixels of the current band
  DevBuf<gtk::Frame> frame_table;  // batched renders
  DevBuf<gtk::Seg> seg_table;
  size_t cap_paths = 0;
  uint32_t region = 0;  // queue entries per slice
  uint32_t gstride = 0;
  uint32_t* host_ctr = nullptr;  // pinned copy of the counters, one block per batch
  size_t host_ctr_words = 0;
  std::vector<hipEvent_t> events;
  // the captured launch sequence of the last render, and what it was captured for;
  // seen_key: the last render's key, so a sequence is captured only when it repeats
  hipGraphExec_t graph_exec = nullptr;
  std::vector<unsigned char> graph_key, seen_key;
  ~gt_context() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (host_ctr) (void)hipHostFree(host_ctr);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct gt_scene {
  gt_context* ctx = nullptr;
  std::vector<gt::HostPrim> prims;
  std::vector<gt::HostMat> mats;
  bool built = false;
  uint32_t leaf_size = 4;
  gt::BinaryBvh bvh;
  gt::PackedScene packed;
  DevBuf<GNode2> nodes2;
  DevBuf<GNode4> nodes4;
  DevBuf<GPrim> dprims;
  DevBuf<uint32_t> prim_ids;
  DevBuf<GMat> dmats;
  std::vector<float4> vnormals;  // mesh triangles' vertex normals, 3 per triangle (HostPrim::smooth - 1)
  DevBuf<float4> dvnormals;
  uint64_t device_bytes = 0;
  uint64_t n_nodes2 = 0, n_nodes4 = 0;  // node records the traversal may reference
  bool on_device = false;               // built by gt_scene_build_device
  gbvh::Result device_build{};
};

struct gt_lights {
  gt_context* ctx = nullptr;
  std::vector<gt::HostPrim> targets;
  std::vector<double> weights;
  double total = 0.0;
  bool dirty = true;
  DevBuf<GLight> dev;
};

struct gt_env {
  gt_context* ctx = nullptr;
  uint32_t kind = 1;  // 1 constant, 2 texture
  double color[3] = {0, 0, 0};
  uint32_t w = 0, h = 0;
  bool uniform = true;
  DevBuf<float4> tex;  // (r, g, b, solid-angle pdf) per texel
  DevBuf<uint2> alias;  // (keep threshold, alias texel) per texel
};

namespace {

bool finite3(double r, double g, double b) { return std::isfinite(r) && std::isfinite(g) && std::isfinite(b); }

int upload_materials(gt_scene* s, size_t* bytes) {
  std::vector<GMat> gm(s->mats.size());
  for (size_t i = 0; i < s->mats.size(); ++i) {
    const gt::HostMat& m = s->mats[i];
    GMat& g = gm[i];
    std::memset(&g, 0, sizeof g);
    for (int c = 0; c < 3; ++c) g.color[c] = float(m.color[c]);
    g.kind = uint32_t(m.kind);
    g.param = float(m.param);
    double lum = 0.2126 * m.color[0] + 0.7152 * m.color[1] + 0.0722 * m.color[2];  // material.rs:1724-1726
    g.f0 = float(std::min(std::max(lum, 0.0), 1.0));
  }
  GT_HIP(s->dmats.ensure(gm.size()));
  if (!gm.empty())
    GT_HIP(hipMemcpyAsync(s->dmats.p, gm.data(), gm.size() * sizeof(GMat), hipMemcpyHostToDevice, s->ctx->stream));
  GT_HIP(hipStreamSynchronize(s->ctx->stream));
  *bytes = gm.size() * sizeof(GMat);
  if (!s->vnormals.empty()) {
    GT_HIP(s->dvnormals.ensure(s->vnormals.size()));
    GT_HIP(hipMemcpyAsync(s->dvnormals.p, s->vnormals.data(), s->vnormals.size() * sizeof(float4),
                          hipMemcpyHostToDevice, s->ctx->stream));
    GT_HIP(hipStreamSynchronize(s->ctx->stream));
    *bytes += s->vnormals.size() * sizeof(float4);
  }
  return GT_OK;
}

int upload_scene(gt_scene* s) {
  gt::PackedScene& pk = s->packed;
  GT_HIP(hipSetDevice(s->ctx->device));
  GT_HIP(s->nodes2.ensure(pk.nodes2.size()));
  GT_HIP(s->nodes4.ensure(pk.nodes4.size()));
  GT_HIP(s->dprims.ensure(pk.prims.size()));
  GT_HIP(s->prim_ids.ensure(pk.prim_ids.size()));
  hipStream_t st = s->ctx->stream;
  if (!pk.nodes2.empty()) GT_HIP(hipMemcpyAsync(s->nodes2.p, pk.nodes2.data(), pk.nodes2.size() * sizeof(GNode2), hipMemcpyHostToDevice, st));
  if (!pk.nodes4.empty()) GT_HIP(hipMemcpyAsync(s->nodes4.p, pk.nodes4.data(), pk.nodes4.size() * sizeof(GNode4), hipMemcpyHostToDevice, st));
  if (!pk.prims.empty()) GT_HIP(hipMemcpyAsync(s->dprims.p, pk.prims.data(), pk.prims.size() * sizeof(GPrim), hipMemcpyHostToDevice, st));
  if (!pk.prim_ids.empty()) GT_HIP(hipMemcpyAsync(s->prim_ids.p, pk.prim_ids.data(), pk.prim_ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  GT_HIP(hipStreamSynchronize(st));
  size_t mat_bytes = 0;
  if (int rc = upload_materials(s, &mat_bytes)) return rc;
  s->device_bytes = pk.nodes2.size() * sizeof(GNode2) + pk.nodes4.size() * sizeof(GNode4) +
                    pk.prims.size() * sizeof(GPrim) + pk.prim_ids.size() * 4 + mat_bytes;
  s->n_nodes2 = pk.nodes2.size();
  s->n_nodes4 = pk.nodes4.size();
  s->on_device = false;
  return GT_OK;
}

int build_scene(gt_scene* s, uint32_t leaf_size) {
  s->leaf_size = leaf_size == 0 ? 1 : leaf_size;
  for (gt::HostPrim& p : s->prims) p.mat_kind = uint32_t(s->mats[p.material].kind);
  s->packed = gt::PackedScene();
  if (!s->prims.empty()) {
    gt::build_binary_bvh(s->prims, s->leaf_size, &s->bvh);
    gt::pack_scene(s->prims, s->bvh, &s->packed);
  } else {
    s->bvh = gt::BinaryBvh();
    // empty scene: an inverted root box that every slab test rejects
    for (int a = 0; a < 3; ++a) { s->packed.root_lo[a] = 1.0f; s->packed.root_hi[a] = -1.0f; }
    s->packed.root_ref2 = s->packed.root_ref4 = GT_LEAF_BIT;
  }
  int rc = upload_scene(s);
  if (rc != GT_OK) return rc;
  s->built = true;
  return GT_OK;
}

int build_scene_device(gt_scene* s, uint32_t leaf_size) {
  s->leaf_size = leaf_size == 0 ? 1 : leaf_size;
  const size_t n = s->prims.size();
  if (n == 0) return build_scene(s, leaf_size);  // the empty scene needs no device work
  if (n >= (size_t(1) << 30)) return fail(GT_EINVAL, "too many primitives for a device build (2^30)");
  for (gt::HostPrim& p : s->prims) p.mat_kind = uint32_t(s->mats[p.material].kind);
  GT_HIP(hipSetDevice(s->ctx->device));
  const size_t m = std::max<size_t>(n - 1, 1);
  GT_HIP(s->nodes2.ensure(m));
  GT_HIP(s->nodes4.ensure(m));
  GT_HIP(s->dprims.ensure(n));
  GT_HIP(s->prim_ids.ensure(n));
  DevBuf<unsigned char> scratch;
  GT_HIP(scratch.ensure(gbvh::scratch_bytes(uint32_t(n))));
  gbvh::Result r{};
  GT_HIP(gbvh::build(s->ctx->stream, s->prims.data(), uint32_t(n), s->leaf_size, scratch.p, s->dprims.p,
                     s->prim_ids.p, s->nodes2.p, s->nodes4.p, &r));
  size_t mat_bytes = 0;
  if (int rc = upload_materials(s, &mat_bytes)) return rc;
  s->bvh = gt::BinaryBvh();
  s->packed = gt::PackedScene();
  s->packed.root_ref2 = s->packed.root_ref4 = r.root_ref;
  for (int a = 0; a < 3; ++a) { s->packed.root_lo[a] = r.root_lo[a]; s->packed.root_hi[a] = r.root_hi[a]; }
  s->packed.depth2 = r.depth;
  s->packed.depth4 = r.depth;  // a wide level spans one or two binary levels
  const bool root_leaf = r.root_ref & GT_LEAF_BIT;
  s->n_nodes2 = s->n_nodes4 = root_leaf ? 0 : m;
  s->device_bytes = m * (sizeof(GNode2) + sizeof(GNode4)) + n * (sizeof(GPrim) + 4) + mat_bytes;
  s->device_build = r;
  s->on_device = true;
  s->built = true;
  return GT_OK;
}

int ensure_built(gt_scene* s) {
  if (s->built) return GT_OK;
  return build_scene(s, 4);  // RayScene::cached_bvh builds with the default options (scene.rs:1567-1573)
}

// Vose's alias method (Vose, "A linear algorithm for generating random numbers
// with a given distribution", IEEE TSE 1991) over weights w[0, n): slot i keeps
// itself when the low 32 bits of its draw are below thr[i] and otherwise
// yields alias[i]. Small and large slots are paired from the back of two
// index-ordered work lists, so the table is a fixed function of the weights.
uint32_t alias_threshold(double p) {
  if (!(p > 0.0)) return 0u;
  if (p >= 1.0) return 0xffffffffu;
  double t = std::floor(p * 4294967296.0);
  return t >= 4294967295.0 ? 0xffffffffu : uint32_t(t);
}
void build_alias(const std::vector<double>& w, double total, std::vector<uint32_t>* thr,
                 std::vector<uint32_t>* alias) {
  const size_t n = w.size();
  std::vector<double> p(n);
  std::vector<uint32_t> small, large;
  thr->assign(n, 0xffffffffu);
  alias->resize(n);
  for (size_t i = 0; i < n; ++i) {
    p[i] = (w[i] * double(n)) / total;
    (*alias)[i] = uint32_t(i);
    if (p[i] < 1.0) small.push_back(uint32_t(i));
    else large.push_back(uint32_t(i));
  }
  while (!small.empty() && !large.empty()) {
    uint32_t sm = small.back();
    small.pop_back();
    uint32_t lg = large.back();
    (*thr)[sm] = alias_threshold(p[sm]);
    (*alias)[sm] = lg;
    p[lg] = (p[lg] + p[sm]) - 1.0;
    if (p[lg] < 1.0) {
      large.pop_back();
      small.push_back(lg);
    }
  }
}

GLight pack_light(const gt::HostPrim& p, double weight, double total) {
  GLight g;
  std::memset(&g, 0, sizeof g);
  const double* q = p.q;
  g.b[3] = float(weight / total);
  g.w[2] = float(weight);
  g.c[3] = float(p.kind);
  switch (p.kind) {
    case GK_SPHERE:
      for (int i = 0; i < 3; ++i) g.a[i] = float(q[i]);
      g.a[3] = float(std::fmax(q[3], 0.0));
      break;
    case GK_MOVING_SPHERE:
      for (int i = 0; i < 3; ++i) { g.a[i] = float(q[i]); g.b[i] = float(q[3 + i]); }
      g.a[3] = float(std::fmax(q[6], 0.0));
      break;
    case GK_TRIANGLE: {
      double e1[3], e2[3];
      for (int i = 0; i < 3; ++i) { e1[i] = q[3 + i] - q[i]; e2[i] = q[6 + i] - q[i]; }
      double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int i = 0; i < 3; ++i) {
        g.a[i] = float(q[i]); g.b[i] = float(e1[i]); g.c[i] = float(e2[i]);
        g.n[i] = len < kEps ? 0.0f : float(n[i] / len);
      }
      g.n[3] = float(0.5 * len);  // triangle_pdf_value (object.rs:1088-1099)
      break;
    }
    default: {
      gt::HostPrim hp = p;
      GPrim gp = gt::pack_prim(hp, false);
      for (int i = 0; i < 3; ++i) { g.a[i] = gp.p0[i]; g.b[i] = gp.p1[i]; g.c[i] = gp.p2[i]; g.n[i] = gp.p3[i]; }
      g.w[0] = gp.p1[3];                 // 1 / |u x v|
      g.w[1] = gp.p2[3];                 // |u x v|^2
      g.w[3] = gp.p3[3];                 // plane d
      // quad_pdf_value area = |u x v| (object.rs:1075-1086)
      const double* u = q + 3;
      const double* v = q + 6;
      double nn[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
      g.n[3] = float(std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]));
      break;
    }
  }
  return g;
}

int upload_lights(gt_lights* l) {
  if (!l->dirty) return GT_OK;
  std::vector<GLight> v(l->targets.size());
  for (size_t i = 0; i < v.size(); ++i) v[i] = pack_light(l->targets[i], l->weights[i], l->total);
  GT_HIP(l->dev.ensure(v.size()));
  if (!v.empty()) {
    GT_HIP(hipMemcpyAsync(l->dev.p, v.data(), v.size() * sizeof(GLight), hipMemcpyHostToDevice, l->ctx->stream));
    GT_HIP(hipStreamSynchronize(l->ctx->stream));
  }
  l->dirty = false;
  return GT_OK;
}

// RayCamera::initialized + validate_view (camera.rs:1068-1201)
int make_camera(const gt_camera* c, GCamera* g) {
  if (!c) return fail(GT_EINVAL, "camera is null");
  if (c->image_width == 0) return fail(GT_EINVAL, "image width must be positive");
  if (!(std::isfinite(c->aspect_ratio) && c->aspect_ratio > 0.0)) return fail(GT_EINVAL, "aspect ratio must be positive and finite");
  if (!(std::isfinite(c->vertical_fov) && c->vertical_fov > 0.0 && c->vertical_fov < 180.0))
    return fail(GT_EINVAL, "vertical field of view must be finite and in 0..180 degrees");
  double from[3], at[3], up[3];
  for (int i = 0; i < 3; ++i) { from[i] = c->lookfrom[i]; at[i] = c->lookat[i]; up[i] = c->view_up[i]; }
  double wv[3] = {from[0] - at[0], from[1] - at[1], from[2] - at[2]};
  if (!(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2] > kEps)) return fail(GT_EINVAL, "lookfrom and lookat must be distinct");
  if (!(up[0] * up[0] + up[1] * up[1] + up[2] * up[2] > kEps)) return fail(GT_EINVAL, "view-up vector must be nonzero");
  auto norm = [](const double* v, double* o) {
    double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l < kEps) { o[0] = o[1] = o[2] = 0; return; }
    for (int i = 0; i < 3; ++i) o[i] = v[i] / l;
  };
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
  };
  double fwd_raw[3] = {at[0] - from[0], at[1] - from[1], at[2] - from[2]}, fwd[3], right_raw[3], right[3], upv[3];
  norm(fwd_raw, fwd);
  cross(fwd, up, right_raw);
  norm(right_raw, right);
  double fl = fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2];
  double rl = right[0] * right[0] + right[1] * right[1] + right[2] * right[2];
  if (fl <= kEps || rl <= kEps) return fail(GT_EINVAL, "view-up vector must not be parallel to the viewing direction");
  cross(right, fwd, upv);
  if (!(std::isfinite(c->defocus_angle) && c->defocus_angle >= 0.0 && c->defocus_angle < 180.0))
    return fail(GT_EINVAL, "defocus angle must be finite and in 0..180 degrees");
  if (!(std::isfinite(c->focus_distance) && c->focus_distance > 0.0)) return fail(GT_EINVAL, "focus distance must be positive and finite");
  if (!(std::isfinite(c->shutter_start) && std::isfinite(c->shutter_end) && c->shutter_start <= c->shutter_end))
    return fail(GT_EINVAL, "shutter interval must be finite and ordered");
  int nbg = c->background_kind == GT_BACKGROUND_VERTICAL_GRADIENT ? 9 : 3;
  for (int i = 0; i < nbg; ++i)
    if (!std::isfinite(c->background[i])) return fail(GT_EINVAL, "background color components must be finite");
  if (c->flags & ~uint32_t(GT_FLAG_GGX_REJECT_BELOW_HORIZON)) return fail(GT_EINVAL, "unknown camera flags");
  if (c->adaptive_error_threshold != 0.0 &&
      !(std::isfinite(c->adaptive_error_threshold) && c->adaptive_error_threshold > 0.0))
    return fail(GT_EINVAL, "adaptive sampling error threshold must be positive and finite");
  if (c->rng_mode != GT_RNG_COUNTER)
    return fail(GT_EUNSUPPORTED, "the GPU path draws per-(pixel, sample, bounce) streams only (rng_mode GT_RNG_COUNTER)");
  std::memset(g, 0, sizeof *g);
  g->width = c->image_width;
  g->height = gt_camera_image_height(c);
  double theta = c->vertical_fov * 3.14159265358979323846 / 180.0;
  double h = std::tan(theta * 0.5);
  double vh = 2.0 * h * c->focus_distance;
  double vw = vh * (double(g->width) / double(g->height));
  double w[3] = {-fwd[0], -fwd[1], -fwd[2]};
  double vu[3], vv[3], du[3], dv[3], ul[3], p00[3];
  for (int i = 0; i < 3; ++i) {
    vu[i] = vw * right[i];
    vv[i] = vh * -upv[i];
    du[i] = vu[i] / double(g->width);
    dv[i] = vv[i] / double(g->height);
  }
  for (int i = 0; i < 3; ++i) ul[i] = from[i] - c->focus_distance * w[i] - vu[i] / 2.0 - vv[i] / 2.0;
  for (int i = 0; i < 3; ++i) p00[i] = ul[i] + 0.5 * (du[i] + dv[i]);
  double dr = c->focus_distance * std::tan(c->defocus_angle * 0.5 * 3.14159265358979323846 / 180.0);
  for (int i = 0; i < 3; ++i) {
    g->center[i] = float(from[i]);
    g->pixel00[i] = float(p00[i]);
    g->du[i] = float(du[i]);
    g->dv[i] = float(dv[i]);
    g->disk_u[i] = float(dr * right[i]);
    g->disk_v[i] = float(dr * upv[i]);
  }
  g->shutter_start = float(c->shutter_start);
  g->shutter_end = float(c->shutter_end);
  g->spp = gt_camera_effective_spp(c);
  g->mode = c->pixel_sample_mode == GT_PIXEL_RANDOM ? 0 : 1;
  g->grid = 1;
  if (c->pixel_sample_mode == GT_PIXEL_STRATIFIED)
    g->grid = std::max<uint32_t>(uint32_t(std::sqrt(double(std::max<uint32_t>(c->samples_per_pixel, 1)))), 1u);
  else if (c->pixel_sample_mode == GT_PIXEL_STRATIFIED_GRID)
    g->grid = std::max<uint32_t>(c->stratified_grid_width, 1u);
  g->defocus = c->defocus_angle <= 0.0 ? 0u : 1u;
  g->seed = c->rng_seed;
  g->max_depth = c->max_depth;
  g->rr_min_depth = c->russian_roulette_min_depth;
  g->background_kind = c->background_kind;
  g->flags = c->flags;
  for (int i = 0; i < 9; ++i) g->background[i] = float(c->background[i]);
  return GT_OK;
}

gtd::DScene device_scene(const gt_scene* s, uint32_t width) {
  gtd::DScene d;
  d.nodes2 = s->nodes2.p;
  d.nodes4 = s->nodes4.p;
  d.prims = s->dprims.p;
  d.vnormals = s->vnormals.empty() ? nullptr : s->dvnormals.p;
  d.prim_ids = s->prim_ids.p;
  d.mats = s->dmats.p;
  for (int a = 0; a < 3; ++a) { d.root_lo[a] = s->packed.root_lo[a]; d.root_hi[a] = s->packed.root_hi[a]; }
  bool wide = width == 4 && s->n_nodes4 != 0;
  d.width = wide ? 4 : 2;
  d.root_ref = wide ? s->packed.root_ref4 : s->packed.root_ref2;
  if (s->n_nodes2 == 0) { d.width = 2; d.root_ref = s->packed.root_ref2; }
  return d;
}

uint64_t stack_depth(const gt_scene* s, uint32_t width) {
  uint64_t d = width == 4 ? s->packed.depth4 * 3 : s->packed.depth2;  // wide nodes push up to 3 per level
  return d + 2;
}

// LDS traversal-stack entries per lane (gt_tuning.stack_lds); 8 measured
// fastest on the 1080p GGX/environment frame (occupancy against spills).
constexpr uint32_t kDefaultStackLds = 8;
uint32_t stack_cap(const gt_context* ctx) {
  return ctx->tuning.stack_lds ? ctx->tuning.stack_lds : kDefaultStackLds;
}

// Persistent grid: blocks_per_cu per CU, or as many as the occupancy of the
// traversal kernels allows (their LDS stack is the binding limit).
uint32_t grid_for(gt_context* ctx) {
  uint32_t bpc = ctx->tuning.blocks_per_cu;
  if (bpc == 0) {
    uint32_t b = ctx->tuning.block_size;
    size_t lds = gtk::stack_lds_bytes(b, stack_cap(ctx));
    bpc = std::max<uint32_t>(1, std::min(gtk::occupancy_blocks(0, b, lds), gtk::occupancy_blocks(2, b, lds)));
  }
  return std::max<uint32_t>(1, ctx->cus * bpc);
}

int ensure_queues(gt_context* ctx, size_t paths, uint32_t max_depth) {
  if (paths > ctx->cap_paths) {
    const size_t region = (paths + gtk::kSlices - 1) / gtk::kSlices;
    const size_t lists = region * gtk::kSlices;
    GT_HIP(ctx->ray_o.ensure(paths)); GT_HIP(ctx->ray_d.ensure(paths)); GT_HIP(ctx->thr.ensure(paths));
    for (int k = 0; k < 2; ++k) {
      GT_HIP(ctx->s_d[k].ensure(paths)); GT_HIP(ctx->s_w[k].ensure(paths));
      GT_HIP(ctx->active[k].ensure(lists)); GT_HIP(ctx->shadow[k].ensure(lists));
    }
    GT_HIP(ctx->radiance.ensure(paths));
    GT_HIP(ctx->hits.ensure(paths));
    for (int c = gtk::CAT_EMIT; c < gtk::kNumCategories; ++c) GT_HIP(ctx->cat[c].ensure(lists));  // no miss list
    ctx->cap_paths = paths;
    ctx->region = uint32_t(region);
  }
  GT_HIP(ctx->ctr.ensure(size_t(max_depth + 1) * gtk::kBounceWords));
  GT_HIP(ctx->tally.ensure(gtk::tally_words(max_depth)));
  return GT_OK;
}

int ensure_gstack(gt_context* ctx, uint64_t depth_needed, uint32_t cap) {
  uint32_t grid = grid_for(ctx);
  uint32_t stride = grid * ctx->tuning.block_size;
  uint64_t levels = depth_needed > cap ? depth_needed - cap : 1;
  GT_HIP(ctx->gstack.ensure(size_t(levels) * stride));
  ctx->gstride = stride;
  return GT_OK;
}

gtk::Queues queues_of(gt_context* ctx) {
  gtk::Queues q;
  q.ray_o = ctx->ray_o.p; q.ray_d = ctx->ray_d.p; q.thr = ctx->thr.p;
  for (int k = 0; k < 2; ++k) {
    q.s_d[k] = ctx->s_d[k].p; q.s_w[k] = ctx->s_w[k].p;
    q.active[k] = ctx->active[k].p; q.shadow[k] = ctx->shadow[k].p;
  }
  q.hits = ctx->hits.p;
  for (int c = 0; c < gtk::kNumCategories; ++c) q.cat[c] = ctx->cat[c].p;
  q.radiance = ctx->radiance.p;
  q.region = ctx->region;
  q.ctr = ctx->ctr.p;
  q.tally = ctx->tally.p;
  q.gstack = ctx->gstack.p;
  q.gstride = ctx->gstride;
  q.stack_cap = stack_cap(ctx);
  return q;
}

int make_frame(gt_context* ctx, const gt_scene* scene, const gt_camera* camera, const gt_lights* lights,
               const gt_env* env, gtk::Frame* f) {
  int rc = make_camera(camera, &f->cam);
  if (rc != GT_OK) return rc;
  f->scene = device_scene(scene, ctx->tuning.bvh_width);
  f->lights.l = lights ? lights->dev.p : nullptr;
  f->lights.n = lights ? uint32_t(lights->targets.size()) : 0u;
  f->lights.total = lights ? float(lights->total) : 0.0f;
  std::memset(&f->env, 0, sizeof f->env);
  if (env) {
    f->env.kind = env->kind;
    for (int i = 0; i < 3; ++i) f->env.color[i] = float(env->color[i]);
    f->env.w = env->w; f->env.h = env->h;
    f->env.uniform = env->uniform ? 1u : 0u;
    f->env.tex = env->tex.p; f->env.alias = env->alias.p;
  }
  f->nee = camera->direct_lighting_mode == GT_LIGHTING_NEE ? 1u : 0u;
  f->has_lights = lights ? 1u : 0u;
  f->light_weight = float(std::min(std::max(camera->light_pdf_weight, 0.0), 1.0));
  f->npix = f->cam.width * f->cam.height;
  f->nact = f->npix;
  f->pixmap = nullptr;
  return GT_OK;
}

int check_objects(gt_context* ctx, const gt_scene* scene, const gt_lights* lights, const gt_env* env) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  if (!scene) return fail(GT_EINVAL, "scene is null");
  if (scene->ctx != ctx || (lights && lights->ctx != ctx) || (env && env->ctx != ctx))
    return fail(GT_EINVAL, "scene, lights and environment must belong to the rendering context");
  if (!std::isfinite(lights ? lights->total : 0.0)) return fail(GT_EINVAL, "light weights must be finite");
  return GT_OK;
}

}  // namespace

namespace gt {
int set_error(int code, const std::string& msg) { return fail(code, msg); }
int context_device(const gt_context* ctx) { return ctx->device; }
hipStream_t context_stream(const gt_context* ctx) { return ctx->stream; }
int context_scratch(gt_context* ctx, int slot, size_t bytes, void** out) {
  if (slot < 0 || slot > 1) return fail(GT_EINVAL, "scratch slot must be 0 or 1");
  GT_HIP(hipSetDevice(ctx->device));
  GT_HIP(ctx->scratch[slot].ensure(bytes));
  *out = ctx->scratch[slot].p;
  return GT_OK;
}
}  // namespace gt

namespace {

// One frame of a render: what the kernels read plus what the host plans with.
struct FrameJob {
  gtk::Frame f{};
  uint32_t spp = 0, depth = 0;
  uint32_t sample_base = 0;             // first sample index rendered (adaptive rounds)
  const gtk::Adapt* adapt = nullptr;    // adaptive round: Welford update instead of accumulation
  bool conn_l = false, conn_e = false;  // light / environment shadow rays
  uint64_t stack_need = 0;
  void* d_out = nullptr;                // finalize target (float or double W*H*3), or null
};

// Splits the frames' samples into batches of at most `cap` paths: each batch
// is a list of segments (whole-image sample ranges), at most one per frame,
// and a frame's segments come in sample order.
std::vector<std::vector<gtk::Seg>> plan_batches(const std::vector<FrameJob>& jobs, uint64_t cap, bool single) {
  std::vector<std::vector<gtk::Seg>> batches;
  if (single) {  // equal batches of k whole-image samples
    const FrameJob& j = jobs[0];
    const uint32_t npix = j.f.nact;
    uint32_t kmax = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(cap / npix, j.spp)));
    uint32_t nb = (j.spp + kmax - 1) / kmax, k = (j.spp + nb - 1) / nb;
    for (uint32_t b = 0; b < nb; ++b) {
      gtk::Seg sg{};
      sg.sample0 = j.sample_base + b * k;
      sg.nsamples = std::min(k, j.spp - b * k);
      batches.push_back({sg});
    }
    return batches;
  }
  std::vector<gtk::Seg> cur;
  uint64_t used = 0, pix = 0;
  for (uint32_t i = 0; i < jobs.size(); ++i) {
    const uint32_t npix = jobs[i].f.nact;
    uint32_t s0 = 0;
    while (s0 < jobs[i].spp) {
      uint64_t fit = (cap - used) / npix;
      if (fit == 0) {
        batches.push_back(cur);
        cur.clear();
        used = pix = 0;
        continue;
      }
      gtk::Seg sg{};
      sg.path0 = uint32_t(used);
      sg.frame = i;
      sg.sample0 = jobs[i].sample_base + s0;
      sg.nsamples = uint32_t(std::min<uint64_t>(fit, jobs[i].spp - s0));
      sg.pix0 = uint32_t(pix);
      cur.push_back(sg);
      used += uint64_t(sg.nsamples) * npix;
      pix += npix;
      s0 += sg.nsamples;
    }
  }
  if (!cur.empty()) batches.push_back(cur);
  return batches;
}

// The wavefront engine: renders every job's samples in batches and writes
// each job's finalized image. `multi` selects the frame-table kernels.
int run_jobs(gt_context* ctx, std::vector<FrameJob>& jobs, bool multi, int out_is_double, gt_render_stats* stats,
             uint64_t max_paths_override = 0) {
  uint64_t cap = max_paths_override ? max_paths_override
                                    : (ctx->tuning.max_paths ? ctx->tuning.max_paths : (uint64_t(1) << 25));
  cap = std::min<uint64_t>(cap, uint64_t(1) << 31);
  uint64_t acc_total = 0, stack_need = 0;
  uint32_t depth_all = 0;
  const bool adaptive = !multi && jobs.size() == 1 && jobs[0].adapt != nullptr;
  for (FrameJob& j : jobs) {
    if (uint64_t(j.f.npix) >= (uint64_t(1) << 31)) return fail(GT_EINVAL, "image too large (2^31 pixels)");
    if (j.depth > 4096) return fail(GT_EINVAL, "max_depth above 4096 is not supported");
    cap = std::max<uint64_t>(cap, j.f.nact);
    acc_total += j.f.npix;
    depth_all = std::max(depth_all, j.depth);
    stack_need = std::max(stack_need, j.stack_need);
  }
  if (acc_total >= (uint64_t(1) << 31)) return fail(GT_EINVAL, "batch too large (2^31 pixels in total)");
  std::vector<std::vector<gtk::Seg>> batches = plan_batches(jobs, cap, !multi);
  uint64_t max_batch = 0;
  for (auto& b : batches) {
    uint64_t t = 0;
    for (auto& sg : b) t += uint64_t(sg.nsamples) * jobs[sg.frame].f.nact;
    max_batch = std::max(max_batch, t);
  }
  int rc;
  if ((rc = ensure_queues(ctx, size_t(max_batch), depth_all)) != GT_OK) return rc;
  if ((rc = ensure_gstack(ctx, stack_need, stack_cap(ctx))) != GT_OK) return rc;
  GT_HIP(ctx->acc.ensure(acc_total));
  uint64_t off = 0;
  for (FrameJob& j : jobs) {
    j.f.acc = ctx->acc.p + off;
    off += j.f.npix;
  }
  size_t seg_total = 0;
  for (auto& b : batches) seg_total += b.size();
  if (multi) {
    std::vector<gtk::Frame> frames(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) frames[i] = jobs[i].f;
    std::vector<gtk::Seg> segs;
    for (auto& b : batches) segs.insert(segs.end(), b.begin(), b.end());
    GT_HIP(ctx->frame_table.ensure(frames.size()));
    GT_HIP(ctx->seg_table.ensure(segs.size()));
    GT_HIP(hipMemcpyAsync(ctx->frame_table.p, frames.data(), frames.size() * sizeof(gtk::Frame),
                          hipMemcpyHostToDevice, ctx->stream));
    GT_HIP(hipMemcpyAsync(ctx->seg_table.p, segs.data(), segs.size() * sizeof(gtk::Seg), hipMemcpyHostToDevice,
                          ctx->stream));
    GT_HIP(hipStreamSynchronize(ctx->stream));  // the host tables go out of scope
  }
  gtk::Queues q = queues_of(ctx);
  gtk::Launch L{grid_for(ctx), ctx->tuning.block_size, ctx->stream};
  hipStream_t st = ctx->stream;
  const size_t nbatch = batches.size();
  const size_t ctr_words = size_t(depth_all + 1) * gtk::kBounceWords;
  const size_t tally_words = gtk::tally_words(depth_all);
  const size_t host_words = ctr_words + 2 * tally_words;  // counters, then the tallies as u32 pairs
  const bool want_stats = stats != nullptr;
  const bool count = want_stats && ctx->tuning.count_traversal != 0;
  size_t bounces_total = 0;
  std::vector<uint32_t> batch_depth(nbatch, 0);
  std::vector<char> batch_l(nbatch, 0), batch_e(nbatch, 0);
  for (size_t b = 0; b < nbatch; ++b) {
    for (auto& sg : batches[b]) {
      const FrameJob& j = jobs[multi ? sg.frame : 0];
      batch_depth[b] = std::max(batch_depth[b], j.depth);
      batch_l[b] |= j.conn_l;
      batch_e[b] |= j.conn_e;
    }
    bounces_total += batch_depth[b];
  }
  size_t nev = want_stats ? 2 + 5 * bounces_total : 0;
  while (ctx->events.size() < nev) {
    hipEvent_t e;
    GT_HIP(hipEventCreate(&e));
    ctx->events.push_back(e);
  }
  if (want_stats && ctx->host_ctr_words < host_words * nbatch) {
    if (ctx->host_ctr) GT_HIP(hipHostFree(ctx->host_ctr));
    ctx->host_ctr = nullptr;
    GT_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_ctr), host_words * nbatch * sizeof(uint32_t)));
    ctx->host_ctr_words = host_words * nbatch;
  }
  // events per bounce: [e0] extend [e1][e2] shade [e3] connect [e4]
  size_t ev = 0;
  auto enqueue = [&]() -> int {
    if (want_stats) GT_HIP(hipEventRecord(ctx->events[ev++], st));
    if (!adaptive) gtk::zero_acc(L, ctx->acc.p, acc_total);
    size_t seg_off = 0;
    for (size_t b = 0; b < nbatch; ++b) {
      gtk::Work w{};
      w.multi = multi;
      w.bt.nseg = uint32_t(batches[b].size());
      uint64_t total = 0, pix = 0;
      for (auto& sg : batches[b]) {
        total += uint64_t(sg.nsamples) * jobs[multi ? sg.frame : 0].f.nact;
        pix += jobs[multi ? sg.frame : 0].f.nact;
      }
      w.bt.total = uint32_t(total);
      w.bt.npix_total = uint32_t(pix);
      if (multi) {
        w.bt.frames = ctx->frame_table.p;
        w.bt.segs = ctx->seg_table.p + seg_off;
      } else {
        w.f = jobs[0].f;
        w.bt.seg0 = batches[b][0];
      }
      seg_off += batches[b].size();
      const uint32_t depth = batch_depth[b];
      gtk::reset_counters(L, q, depth_all, w.bt.total);
      if (depth == 0) gtk::extend(L, w, q, 0, false);  // no bounce: every path is black, nothing is traced
      for (uint32_t bounce = 0; bounce < depth; ++bounce) {
        if (want_stats) GT_HIP(hipEventRecord(ctx->events[ev++], st));
        gtk::extend(L, w, q, bounce, count);
        if (want_stats) {
          GT_HIP(hipEventRecord(ctx->events[ev++], st));
          GT_HIP(hipEventRecord(ctx->events[ev++], st));
        }
        gtk::shade(L, w, q, bounce);
        if (want_stats) GT_HIP(hipEventRecord(ctx->events[ev++], st));
        if (batch_l[b]) gtk::connect(L, w, q, bounce, 0, count);
        if (batch_e[b]) gtk::connect(L, w, q, bounce, 1, count);
        if (want_stats) GT_HIP(hipEventRecord(ctx->events[ev++], st));
      }
      if (adaptive) gtk::adapt(L, w, q, *jobs[0].adapt);
      else gtk::accumulate(L, w, q);
      if (want_stats) {
        uint32_t* h = ctx->host_ctr + host_words * b;
        GT_HIP(hipMemcpyAsync(h, q.ctr, ctr_words * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        GT_HIP(hipMemcpyAsync(h + ctr_words, q.tally, tally_words * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
      }
    }
    for (FrameJob& j : jobs)
      if (j.d_out && !adaptive) gtk::finalize(L, j.f.acc, j.f.npix, j.d_out, out_is_double);
    if (want_stats) GT_HIP(hipEventRecord(ctx->events[ev++], st));
    return GT_OK;
  };
  if (ctx->tuning.use_graph && !want_stats && !adaptive) {
    // The whole launch sequence is captured once and replayed while nothing
    // it was captured with changes: frames (with their accumulators),
    // segments, queue and table addresses, grid, outputs.
    std::vector<unsigned char> key;
    auto put = [&key](const void* p, size_t n) {
      const unsigned char* c = static_cast<const unsigned char*>(p);
      key.insert(key.end(), c, c + n);
    };
    for (const FrameJob& j : jobs) {
      put(&j.f, sizeof j.f);
      put(&j.d_out, sizeof j.d_out);
    }
    for (auto& b : batches) put(b.data(), b.size() * sizeof(gtk::Seg));
    put(&q, sizeof q);
    put(&L, sizeof L);
    const void* tables[2] = {ctx->frame_table.p, ctx->seg_table.p};
    put(tables, sizeof tables);
    const uint32_t flags[4] = {multi ? 1u : 0u, uint32_t(out_is_double), depth_all, count ? 1u : 0u};
    put(flags, sizeof flags);
    const bool cached = ctx->graph_exec && key == ctx->graph_key;
    if (!cached && key != ctx->seen_key) {
      // first sight of this sequence: run it eagerly (a one-off render would pay for a capture
      // and an instantiation it never replays) and capture it if the next render repeats it
      ctx->seen_key = std::move(key);
      int erc = enqueue();
      if (erc != GT_OK) return erc;
    } else if (!cached) {
      if (ctx->graph_exec) GT_HIP(hipGraphExecDestroy(ctx->graph_exec));
      ctx->graph_exec = nullptr;
      ctx->graph_key.clear();
      hipGraph_t graph = nullptr;
      GT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int erc = enqueue();
      hipError_t ce = hipStreamEndCapture(st, &graph);
      if (erc != GT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return erc;
      }
      GT_HIP(ce);
      hipError_t ie = hipGraphInstantiate(&ctx->graph_exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      GT_HIP(ie);
      ctx->graph_key = std::move(key);
      GT_HIP(hipGraphLaunch(ctx->graph_exec, st));
    } else {
      GT_HIP(hipGraphLaunch(ctx->graph_exec, st));
    }
  } else {
    int erc = enqueue();
    if (erc != GT_OK) return erc;
  }
  GT_HIP(hipGetLastError());
  if (want_stats) {
    GT_HIP(hipStreamSynchronize(st));
    std::memset(stats, 0, sizeof *stats);
    for (const FrameJob& j : jobs) stats->camera_paths += uint64_t(j.f.nact) * j.spp;
    for (size_t b = 0; b < nbatch; ++b) {
      const uint32_t* c = ctx->host_ctr + host_words * b;
      std::vector<uint64_t> tally(tally_words);
      std::memcpy(tally.data(), c + ctr_words, tally_words * sizeof(uint64_t));
      if (!adaptive) {  // every sample of the batch but the dropped ones (adaptive: counted at the end)
        uint64_t paths = 0;
        for (auto& sg : batches[b]) paths += uint64_t(sg.nsamples) * jobs[multi ? sg.frame : 0].f.nact;
        stats->accepted_samples += paths - tally[0];
      }
      for (uint32_t bounce = 0; bounce < batch_depth[b]; ++bounce) {
        for (uint32_t sl = 0; sl < gtk::kSlices; ++sl) {
          stats->extension_rays += c[gtk::ctr_at(bounce, gtk::S_QCOUNT, sl)];
          if (batch_l[b]) stats->shadow_rays += c[gtk::ctr_at(bounce, gtk::S_SL_COUNT, sl)];
          if (batch_e[b]) stats->shadow_rays += c[gtk::ctr_at(bounce, gtk::S_SE_COUNT, sl)];
        }
        if (count) {
          const uint64_t* v = tally.data() + 1 + 4 * bounce;
          stats->ext_node_fetches += v[0]; stats->ext_prim_tests += v[1];
          stats->shadow_node_fetches += v[2]; stats->shadow_prim_tests += v[3];
        }
      }
    }
    float ms = 0;
    GT_HIP(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[ev - 1]));
    stats->device_ms = ms;
    size_t base = 1;
    for (size_t i = 0; i < bounces_total; ++i) {
      float a = 0, sh = 0, cn = 0;
      GT_HIP(hipEventElapsedTime(&a, ctx->events[base], ctx->events[base + 1]));
      GT_HIP(hipEventElapsedTime(&sh, ctx->events[base + 2], ctx->events[base + 3]));
      GT_HIP(hipEventElapsedTime(&cn, ctx->events[base + 3], ctx->events[base + 4]));
      stats->extend_ms += a; stats->shade_ms += sh; stats->connect_ms += cn;
      stats->extend_launches++;
      base += 5;
    }
  }
  return GT_OK;
}

int prepare_job(gt_context* ctx, const gt_scene* cscene, const gt_camera* camera, const gt_lights* clights,
                const gt_env* env, FrameJob* j) {
  int rc = check_objects(ctx, cscene, clights, env);
  if (rc != GT_OK) return rc;
  if (!camera) return fail(GT_EINVAL, "camera is null");
  gt_scene* scene = const_cast<gt_scene*>(cscene);
  gt_lights* lights = const_cast<gt_lights*>(clights);
  if ((rc = ensure_built(scene)) != GT_OK) return rc;
  if (lights && (rc = upload_lights(lights)) != GT_OK) return rc;
  if ((rc = make_frame(ctx, scene, camera, lights, env, &j->f)) != GT_OK) return rc;
  j->spp = j->f.cam.spp;
  j->depth = j->f.cam.max_depth;
  j->conn_l = j->f.nee && lights != nullptr;
  j->conn_e = j->f.nee && env != nullptr;
  j->stack_need = stack_depth(scene, j->f.scene.width);
  return GT_OK;
}

bool adaptive_camera(const gt_camera* c) {
  return c->adaptive_error_threshold != 0.0 && c->pixel_sample_mode == GT_PIXEL_RANDOM;
}

void add_stats(gt_render_stats* t, const gt_render_stats& r) {
  t->camera_paths += r.camera_paths;
  t->extension_rays += r.extension_rays;
  t->shadow_rays += r.shadow_rays;
  t->accepted_samples += r.accepted_samples;
  t->device_ms += r.device_ms;
  t->extend_ms += r.extend_ms;
  t->extend_launches += r.extend_launches;
  t->shade_ms += r.shade_ms;
  t->connect_ms += r.connect_ms;
  t->ext_node_fetches += r.ext_node_fetches;
  t->ext_prim_tests += r.ext_prim_tests;
  t->shadow_node_fetches += r.shadow_node_fetches;
  t->shadow_prim_tests += r.shadow_prim_tests;
}

// render_world_pixel_adaptive (camera.rs:2937-2970) for every pixel, in
// rounds: every pixel takes its first min samples, then the pixels that have
// not stopped take further rounds (half the samples taken so far, at least
// min) until none is left or max samples are taken. The Welford update runs
// over each round's samples in sample order and stops a pixel exactly where
// the sequential loop would; samples of a round after a pixel's stop are
// traced but unused. Paths depend only on (pixel, sample), so the image does
// not depend on the rounds.
constexpr uint64_t kMinRoundPaths = uint64_t(1) << 22;  // adaptive rounds: enough paths to fill one MI355X

int render_adaptive(gt_context* ctx, FrameJob job, const gt_camera* cam, int out_is_double,
                    gt_render_stats* stats, const uint32_t* first_map = nullptr, uint32_t first_n = 0) {
  const uint32_t npix = job.f.npix;
  const uint32_t mn = std::max<uint32_t>(cam->adaptive_min_samples, 1);
  const uint32_t mx = std::max(cam->adaptive_max_samples, mn);
  GT_HIP(ctx->ad_mean.ensure(npix));
  GT_HIP(ctx->ad_m2.ensure(npix));
  GT_HIP(ctx->ad_done.ensure(npix));
  GT_HIP(ctx->ad_map[0].ensure(npix));
  GT_HIP(ctx->ad_map[1].ensure(npix));
  GT_HIP(ctx->ad_count.ensure(1));
  hipStream_t st = ctx->stream;
  GT_HIP(hipMemsetAsync(ctx->ad_mean.p, 0, size_t(npix) * sizeof(double4), st));
  GT_HIP(hipMemsetAsync(ctx->ad_m2.p, 0, size_t(npix) * sizeof(double4), st));
  GT_HIP(hipMemsetAsync(ctx->ad_done.p, 0, size_t(npix) * sizeof(uint32_t), st));
  gtk::Adapt ad{ctx->ad_mean.p, ctx->ad_m2.p, ctx->ad_done.p, mn, cam->adaptive_error_threshold};
  gtk::Launch L{grid_for(ctx), ctx->tuning.block_size, st};
  if (stats) std::memset(stats, 0, sizeof *stats);
  const uint32_t* map = first_map;
  uint32_t nact = first_map ? first_n : npix, s = 0, k = std::min(mn, mx);
  int cur = 0;
  for (;;) {
    std::vector<FrameJob> jobs(1, job);
    jobs[0].f.pixmap = map;
    jobs[0].f.nact = nact;
    jobs[0].sample_base = s;
    jobs[0].spp = k;
    jobs[0].adapt = &ad;
    jobs[0].d_out = nullptr;
    gt_render_stats rs{};
    int rc = run_jobs(ctx, jobs, false, out_is_double, stats ? &rs : nullptr);
    if (rc != GT_OK) return rc;
    if (stats) add_stats(stats, rs);
    s += k;
    if (s >= mx) break;
    GT_HIP(hipMemsetAsync(ctx->ad_count.p, 0, sizeof(uint32_t), st));
    gtk::adapt_compact(L, map, nact, ctx->ad_done.p, ctx->ad_map[cur].p, ctx->ad_count.p);
    GT_HIP(hipGetLastError());
    uint32_t left = 0;
    GT_HIP(hipMemcpyAsync(&left, ctx->ad_count.p, sizeof left, hipMemcpyDeviceToHost, st));
    GT_HIP(hipStreamSynchronize(st));
    if (left == 0) break;
    nact = left;
    map = ctx->ad_map[cur].p;
    cur ^= 1;
    k = std::min(mx - s, std::max(mn, s / 2));
    // A round of few paths leaves most of the chip idle for its fixed cost of ~30 launches: it
    // takes at least kMinRoundPaths paths (or as many as the first round), even if that traces
    // samples a pixel will not use. Each pixel still stops at the same sample.
    const uint64_t floor_paths = std::min<uint64_t>(kMinRoundPaths, uint64_t(npix) * mn);
    const uint64_t want = (floor_paths + nact - 1) / nact;
    k = uint32_t(std::min<uint64_t>(mx - s, std::max<uint64_t>(k, want)));
  }
  if (job.d_out)
    gtk::adapt_finalize(L, ctx->ad_mean.p, first_map, first_map ? first_n : npix, job.d_out, out_is_double);
  if (stats) {  // the samples the pixels used
    GT_HIP(hipMemsetAsync(ctx->tally.p, 0, sizeof(unsigned long long), st));
    gtk::adapt_count(L, ctx->ad_mean.p, first_map, first_map ? first_n : npix, ctx->tally.p);
    unsigned long long used = 0;
    GT_HIP(hipMemcpyAsync(&used, ctx->tally.p, sizeof used, hipMemcpyDeviceToHost, st));
    GT_HIP(hipStreamSynchronize(st));
    stats->accepted_samples = used;
  }
  GT_HIP(hipGetLastError());
  return GT_OK;
}

int render_impl(gt_context* ctx, const gt_scene* scene, const gt_camera* camera, const gt_lights* lights,
                const gt_env* env, void* d_out, int out_is_double, float* d_albedo, float* d_normal,
                gt_render_stats* stats) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  GT_HIP(hipSetDevice(ctx->device));
  std::vector<FrameJob> jobs(1);
  int rc = prepare_job(ctx, scene, camera, lights, env, &jobs[0]);
  if (rc != GT_OK) return rc;
  jobs[0].d_out = d_out;
  if (adaptive_camera(camera)) rc = render_adaptive(ctx, jobs[0], camera, out_is_double, stats);
  else rc = run_jobs(ctx, jobs, false, out_is_double, stats);
  if (rc != GT_OK) return rc;
  if (d_albedo) {
    gtk::Launch L{grid_for(ctx), ctx->tuning.block_size, ctx->stream};
    gtk::aov(L, jobs[0].f, queues_of(ctx), d_albedo, d_normal);
    GT_HIP(hipGetLastError());
  }
  return GT_OK;
}

}  // namespace

extern "C" {

const char* gt_last_error(void) { return g_error.c_str(); }
const char* gt_version(void) { return "gartus-mi355x 0.1 (gfx950 wavefront path tracer)"; }

void gt_camera_default(gt_camera* c) {  // camera.rs:1034-1064
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->image_width = 100;
  c->aspect_ratio = 1.0;
  c->samples_per_pixel = 1;
  c->pixel_sample_mode = GT_PIXEL_RANDOM;
  c->stratified_grid_width = 1;
  c->max_depth = 10;
  c->russian_roulette_min_depth = 5;
  c->rng_seed = 1;
  c->vertical_fov = 90.0;
  c->lookfrom[0] = 0; c->lookfrom[1] = 0; c->lookfrom[2] = 0;
  c->lookat[0] = 0; c->lookat[1] = 0; c->lookat[2] = -1;
  c->view_up[0] = 0; c->view_up[1] = 1; c->view_up[2] = 0;
  c->defocus_angle = 0.0;
  c->focus_distance = 1.0;
  c->shutter_start = 0.0;
  c->shutter_end = 1.0;
  c->background_kind = GT_BACKGROUND_CONSTANT;
  c->background[0] = 0.70; c->background[1] = 0.80; c->background[2] = 1.00;
  c->direct_lighting_mode = GT_LIGHTING_CONTINUATION;
  c->light_pdf_weight = 0.5;
  c->rng_mode = GT_RNG_COUNTER;
}

uint32_t gt_camera_image_height(const gt_camera* c) {  // camera.rs:1071
  if (!c || c->image_width == 0 || !(c->aspect_ratio > 0.0)) return 1;
  double h = double(c->image_width) / c->aspect_ratio;
  uint32_t hh = h >= 4294967295.0 ? 0xffffffffu : (h > 0.0 ? uint32_t(h) : 0u);
  return std::max<uint32_t>(hh, 1u);
}

uint32_t gt_camera_effective_spp(const gt_camera* c) {  // camera.rs:1547-1557
  uint32_t s = std::max<uint32_t>(c->samples_per_pixel, 1u);
  if (c->pixel_sample_mode == GT_PIXEL_RANDOM) return s;
  if (c->pixel_sample_mode == GT_PIXEL_STRATIFIED) {
    uint32_t g = std::max<uint32_t>(uint32_t(std::sqrt(double(s))), 1u);
    return g * g;
  }
  uint32_t g = std::max<uint32_t>(c->stratified_grid_width, 1u);
  return g * g;
}

int gt_context_create(int device, gt_context** out) {
  if (!out) return fail(GT_EINVAL, "out is null");
  *out = nullptr;
  int n = 0;
  GT_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(GT_EINVAL, "no such HIP device");
  std::unique_ptr<gt_context> ctx(new gt_context());
  ctx->device = device;
  GT_HIP(hipSetDevice(device));
  GT_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  // Set-up of the no-allocation entry point: the address of this device's
  // copy of its scratch array, looked up once per device. A failure is kept
  // and reported by that entry point; the others do not need it.
  (void)gtk::noalloc_scratch_init(device);
  hipDeviceProp_t props;
  GT_HIP(hipGetDeviceProperties(&props, device));
  ctx->cus = uint32_t(props.multiProcessorCount);
  ctx->tuning.block_size = 128;
  ctx->tuning.blocks_per_cu = 0;
  ctx->tuning.max_paths = 0;
  ctx->tuning.bvh_width = 2;
  ctx->tuning.use_graph = 1;
  *out = ctx.release();
  return GT_OK;
}

int gt_context_destroy(gt_context* ctx) {
  if (!ctx) return GT_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return GT_OK;
}

int gt_context_synchronize(gt_context* ctx) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  GT_HIP(hipStreamSynchronize(ctx->stream));
  return GT_OK;
}

void* gt_context_stream(gt_context* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int gt_device_malloc(gt_context* ctx, size_t bytes, void** out) {
  if (!ctx || !out) return fail(GT_EINVAL, "null argument");
  GT_HIP(hipSetDevice(ctx->device));
  GT_HIP(hipMalloc(out, std::max<size_t>(bytes, 1)));
  return GT_OK;
}

int gt_device_free(gt_context* ctx, void* ptr) {
  if (!ctx) return fail(GT_EINVAL, "null argument");
  GT_HIP(hipSetDevice(ctx->device));
  GT_HIP(hipStreamSynchronize(ctx->stream));
  if (ptr) GT_HIP(hipFree(ptr));
  return GT_OK;
}

int gt_memcpy_to_host(gt_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(GT_EINVAL, "null argument");
  GT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  GT_HIP(hipStreamSynchronize(ctx->stream));
  return GT_OK;
}

int gt_memcpy_to_device(gt_context* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return fail(GT_EINVAL, "null argument");
  GT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  GT_HIP(hipStreamSynchronize(ctx->stream));
  return GT_OK;
}

int gt_context_get_tuning(gt_context* ctx, gt_tuning* out) {
  if (!ctx || !out) return fail(GT_EINVAL, "null argument");
  *out = ctx->tuning;
  return GT_OK;
}

int gt_context_set_tuning(gt_context* ctx, const gt_tuning* t) {
  if (!ctx || !t) return fail(GT_EINVAL, "null argument");
  if (t->block_size < 64 || t->block_size > 1024 || t->block_size % 64 != 0)
    return fail(GT_EINVAL, "block_size must be a multiple of 64 in 64..1024");
  if (t->bvh_width != 2 && t->bvh_width != 4) return fail(GT_EINVAL, "bvh_width must be 2 or 4");
  if (t->stack_lds != 0 && (t->stack_lds < 4 || t->stack_lds > 32))
    return fail(GT_EINVAL, "stack_lds must be 0 or 4..32");
  size_t lds = gtk::stack_lds_bytes(t->block_size, t->stack_lds ? t->stack_lds : kDefaultStackLds);
  if (lds > 64 * 1024) return fail(GT_EINVAL, "block_size too large for the LDS traversal stack");
  ctx->tuning = *t;
  return GT_OK;
}

// ---- scene ----
int gt_scene_create(gt_context* ctx, gt_scene** out) {
  if (!ctx || !out) return fail(GT_EINVAL, "null argument");
  gt_scene* s = new gt_scene();
  s->ctx = ctx;
  *out = s;
  return GT_OK;
}
int gt_scene_destroy(gt_scene* s) {
  if (!s) return GT_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  delete s;
  return GT_OK;
}

static int add_material(gt_scene* s, int kind, double r, double g, double b, double param, uint32_t* id) {
  if (!s || !id) return fail(GT_EINVAL, "null argument");
  if (s->mats.size() >= GT_MATERIAL_MASK) return fail(GT_EINVAL, "too many materials");
  gt::HostMat m;
  m.kind = kind;
  m.color[0] = r; m.color[1] = g; m.color[2] = b;
  m.param = param;
  *id = uint32_t(s->mats.size());
  s->mats.push_back(m);
  s->built = false;
  return GT_OK;
}
int gt_scene_add_lambertian(gt_scene* s, double r, double g, double b, uint32_t* id) {
  if (!finite3(r, g, b)) return fail(GT_EINVAL, "Lambertian albedo must be finite");  // material.rs:363
  return add_material(s, GM_LAMBERTIAN, r, g, b, 0.0, id);
}
int gt_scene_add_diffuse_light(gt_scene* s, double r, double g, double b, uint32_t* id) {
  if (!finite3(r, g, b)) return fail(GT_EINVAL, "diffuse light color must be finite");  // material.rs:946
  return add_material(s, GM_DIFFUSE_LIGHT, r, g, b, 0.0, id);
}
int gt_scene_add_ggx(gt_scene* s, double r, double g, double b, double roughness, uint32_t* id) {
  if (!finite3(r, g, b) || !std::isfinite(roughness))
    return fail(GT_EINVAL, "GGX microfacet material values must be finite");  // material.rs:1387
  return add_material(s, GM_GGX, r, g, b, std::min(std::max(roughness, 0.0), 1.0), id);
}
int gt_scene_add_metal(gt_scene* s, double r, double g, double b, double fuzz, uint32_t* id) {
  if (!finite3(r, g, b) || !std::isfinite(fuzz)) return fail(GT_EINVAL, "metal material values must be finite");  // material.rs:1968
  return add_material(s, GM_METAL, r, g, b, std::min(std::max(fuzz, 0.0), 1.0), id);
}
int gt_scene_add_isotropic(gt_scene* s, double r, double g, double b, uint32_t* id) {
  if (!finite3(r, g, b)) return fail(GT_EINVAL, "isotropic albedo must be finite");  // material.rs:1062
  return add_material(s, GM_ISOTROPIC, r, g, b, 0.0, id);
}
int gt_scene_add_henyey_greenstein(gt_scene* s, double r, double g, double b, double anisotropy, uint32_t* id) {
  if (!finite3(r, g, b)) return fail(GT_EINVAL, "Henyey-Greenstein albedo must be finite");  // material.rs:1195
  if (!(std::isfinite(anisotropy) && std::fabs(anisotropy) < 1.0))  // validate_henyey_greenstein_g :1328-1333
    return fail(GT_EINVAL, "Henyey-Greenstein anisotropy must be finite and in (-1, 1)");
  return add_material(s, GM_HENYEY_GREENSTEIN, r, g, b, anisotropy, id);
}
int gt_scene_add_dielectric(gt_scene* s, double ior, uint32_t* id) {
  if (!std::isfinite(ior)) return fail(GT_EINVAL, "refraction index must be finite");
  return add_material(s, GM_DIELECTRIC, 1.0, 1.0, 1.0, ior, id);
}

static int add_prims(gt_scene* s, int kind, int stride, const double* data, size_t n, const uint32_t* mids) {
  if (!s) return fail(GT_EINVAL, "scene is null");
  if (n == 0) return GT_OK;
  if (!data || !mids) return fail(GT_EINVAL, "null primitive data");
  for (size_t i = 0; i < n; ++i)
    if (mids[i] >= s->mats.size()) return fail(GT_EINVAL, "ray scene material id out of bounds");  // scene.rs:1158-1161
  if (s->prims.size() + n >= (size_t(1) << 31)) return fail(GT_EINVAL, "too many primitives");
  s->prims.reserve(s->prims.size() + n);
  for (size_t i = 0; i < n; ++i) {
    gt::HostPrim p;
    std::memset(&p, 0, sizeof p);
    p.kind = kind;
    for (int j = 0; j < stride; ++j) p.q[j] = data[size_t(stride) * i + j];
    if (kind == GK_SPHERE) p.q[3] = std::fmax(p.q[3], 0.0);          // geometry.rs:445-450
    if (kind == GK_MOVING_SPHERE) p.q[6] = std::fmax(p.q[6], 0.0);   // geometry.rs:498-504
    p.material = mids[i];
    p.mat_kind = 0;
    s->prims.push_back(p);
  }
  s->built = false;  // invalidate_bvh (scene.rs:1575-1577)
  return GT_OK;
}
int gt_scene_add_spheres(gt_scene* s, const double* d, size_t n, const uint32_t* m) { return add_prims(s, GK_SPHERE, 4, d, n, m); }
int gt_scene_add_moving_spheres(gt_scene* s, const double* d, size_t n, const uint32_t* m) { return add_prims(s, GK_MOVING_SPHERE, 7, d, n, m); }
int gt_scene_add_triangles(gt_scene* s, const double* d, size_t n, const uint32_t* m) { return add_prims(s, GK_TRIANGLE, 9, d, n, m); }
int gt_scene_add_quads(gt_scene* s, const double* d, size_t n, const uint32_t* m) { return add_prims(s, GK_QUAD, 9, d, n, m); }
int gt_scene_add_mesh_triangles(gt_scene* s, const double* positions, const double* normals, size_t n,
                                const uint32_t* m) {
  if (!s) return fail(GT_EINVAL, "scene is null");
  if (n == 0) return GT_OK;
  if (!normals) return fail(GT_EINVAL, "null vertex normals");
  if (s->vnormals.size() / 3 + n >= (size_t(1) << 31)) return fail(GT_EINVAL, "too many mesh triangles");
  const size_t first = s->prims.size();
  int rc = add_prims(s, GK_TRIANGLE, 9, positions, n, m);
  if (rc != GT_OK) return rc;
  for (size_t i = 0; i < n; ++i) {
    s->prims[first + i].smooth = uint32_t(s->vnormals.size() / 3 + 1);
    for (int k = 0; k < 3; ++k) {
      const double* v = normals + 9 * i + 3 * k;
      s->vnormals.push_back(make_float4(float(v[0]), float(v[1]), float(v[2]), 0.0f));
    }
  }
  return GT_OK;
}

int gt_scene_build(gt_scene* s, uint32_t leaf_size) {
  if (!s) return fail(GT_EINVAL, "scene is null");
  return build_scene(s, leaf_size);
}

int gt_scene_build_device(gt_scene* s, uint32_t leaf_size) {
  if (!s) return fail(GT_EINVAL, "scene is null");
  return build_scene_device(s, leaf_size);
}

int gt_scene_get_info(const gt_scene* s, gt_scene_info* out) {
  if (!s || !out) return fail(GT_EINVAL, "null argument");
  std::memset(out, 0, sizeof *out);
  out->primitive_count = s->prims.size();
  out->material_count = s->mats.size();
  if (s->built && s->on_device) {
    const gbvh::Result& r = s->device_build;
    out->bvh_node_count = r.internal_nodes + r.leaves;
    out->bvh_depth = r.depth;
    out->device_bytes = s->device_bytes;
    for (int a = 0; a < 3; ++a) { out->bounds_min[a] = r.root_lo[a]; out->bounds_max[a] = r.root_hi[a]; }
    out->build_ms = r.ms;
    out->built_on_device = 1;
  } else if (s->built && !s->bvh.nodes.empty()) {
    out->bvh_node_count = s->bvh.nodes.size();
    out->bvh_depth = s->bvh.depth;
    out->device_bytes = s->device_bytes;
    for (int a = 0; a < 3; ++a) { out->bounds_min[a] = s->bvh.nodes[0].bounds.lo[a]; out->bounds_max[a] = s->bvh.nodes[0].bounds.hi[a]; }
  }
  return GT_OK;
}

// ---- intersect ----
static_assert(sizeof(gt_ray) == 2 * sizeof(float4) && sizeof(gt_hit) == 2 * sizeof(float4),
              "rays and hits are two float4 records on the device");

static void copy_stats(const unsigned long long v[7], gt_traversal_stats* stats) {
  stats->rays = v[0]; stats->node_bounds_tests = v[1]; stats->node_bounds_hits = v[2];
  stats->leaf_visits = v[3]; stats->primitive_candidates = v[4]; stats->leaf_hit_results = v[5];
  stats->max_stack_depth = v[6];
}

int gt_intersect(gt_context* ctx, const gt_scene* cs, const gt_ray* rays, size_t n, float t_min, gt_hit* out,
                 gt_traversal_stats* stats) {
  int rc = check_objects(ctx, cs, nullptr, nullptr);
  if (rc != GT_OK) return rc;
  if (n == 0) { if (stats) std::memset(stats, 0, sizeof *stats); return GT_OK; }
  if (!rays || !out) return fail(GT_EINVAL, "null rays or hits");
  gt_scene* s = const_cast<gt_scene*>(cs);
  GT_HIP(hipSetDevice(ctx->device));
  if ((rc = ensure_built(s)) != GT_OK) return rc;
  gtk::Frame f;
  std::memset(&f, 0, sizeof f);
  f.scene = device_scene(s, ctx->tuning.bvh_width);
  if ((rc = ensure_gstack(ctx, stack_depth(s, f.scene.width), stack_cap(ctx))) != GT_OK) return rc;
  gtk::Queues q = queues_of(ctx);
  gtk::Launch L{grid_for(ctx), ctx->tuning.block_size, ctx->stream};
  const size_t chunk = size_t(1) << 22;
  const size_t m = std::min(n, chunk);
  DevBuf<float4> dr, dh;
  DevBuf<unsigned long long> dstats;
  GT_HIP(dr.ensure(2 * m));
  GT_HIP(dh.ensure(2 * m));
  if (stats) {
    GT_HIP(dstats.ensure(8));
    gtk::zero_words(ctx->stream, dstats.p, 8);
  }
  for (size_t off = 0; off < n; off += chunk) {
    const size_t c = std::min(chunk, n - off);
    GT_HIP(hipMemcpyAsync(dr.p, rays + off, c * sizeof(gt_ray), hipMemcpyHostToDevice, ctx->stream));
    gtk::intersect(L, f, q, dr.p, c, t_min, dh.p, stats ? dstats.p : nullptr);
    GT_HIP(hipGetLastError());
    GT_HIP(hipMemcpyAsync(out + off, dh.p, c * sizeof(gt_hit), hipMemcpyDeviceToHost, ctx->stream));
    GT_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (stats) {
    unsigned long long v[8];
    GT_HIP(hipMemcpy(v, dstats.p, sizeof v, hipMemcpyDeviceToHost));
    copy_stats(v, stats);
  }
  return GT_OK;
}

int gt_intersect_device_noalloc(gt_context* ctx, const gt_scene* s, const gt_ray* d_rays, size_t n, float t_min,
                                gt_hit* d_hits, gt_traversal_stats* d_stats) {
  int rc = check_objects(ctx, s, nullptr, nullptr);
  if (rc != GT_OK) return rc;
  if (!s->built) return fail(GT_ESTATE, "scene is not built: gt_intersect_device_noalloc does not build or "
                                        "upload (call gt_scene_build or gt_scene_build_device first)");
  if (n != 0 && (!d_rays || !d_hits)) return fail(GT_EINVAL, "null rays or hits");
  GT_HIP(hipSetDevice(ctx->device));
  uint2* scratch = nullptr;
  const hipError_t se = gtk::noalloc_scratch(ctx->device, &scratch);
  if (se != hipSuccess || !scratch)
    return fail(GT_EDEVICE, std::string("no-allocation scratch unavailable on this device: ") + hipGetErrorString(se));
  gtk::Frame f;
  std::memset(&f, 0, sizeof f);
  f.scene = device_scene(s, ctx->tuning.bvh_width);
  // Overflow levels of the per-lane stack beyond the LDS entries; the grid
  // shrinks until lanes * levels fits the scratch, down to one block per CU.
  const uint32_t cap = stack_cap(ctx), block = ctx->tuning.block_size;
  const uint64_t need = stack_depth(s, f.scene.width);
  const uint64_t levels = need > cap ? need - cap : 1;
  const uint64_t fit_blocks = gtk::kNoAllocStackEntries / (levels * block);
  if (fit_blocks < ctx->cus) {
    const uint64_t max_levels = gtk::kNoAllocStackEntries / (uint64_t(ctx->cus) * block);
    return fail(GT_EINVAL, "BVH too deep for the no-allocation scratch: traversal needs " + std::to_string(need) +
                               " stack entries per ray, at most " + std::to_string(max_levels + cap) +
                               " fit at block_size " + std::to_string(block) + " (use gt_intersect)");
  }
  gtk::Queues q{};
  q.gstack = scratch;
  q.stack_cap = cap;
  const uint32_t grid = uint32_t(std::min<uint64_t>(grid_for(ctx), fit_blocks));
  q.gstride = grid * block;
  gtk::Launch L{grid, block, ctx->stream};
  unsigned long long* st = reinterpret_cast<unsigned long long*>(d_stats);
  if (st) gtk::zero_words(ctx->stream, st, 7);
  if (n != 0) {
    gtk::intersect(L, f, q, reinterpret_cast<const float4*>(d_rays), n, t_min, reinterpret_cast<float4*>(d_hits), st);
  }
  GT_HIP(hipGetLastError());
  return GT_OK;
}

// ---- lights ----
int gt_lights_create(gt_context* ctx, gt_lights** out) {
  if (!ctx || !out) return fail(GT_EINVAL, "null argument");
  gt_lights* l = new gt_lights();
  l->ctx = ctx;
  *out = l;
  return GT_OK;
}
int gt_lights_destroy(gt_lights* l) {
  if (!l) return GT_OK;
  (void)hipStreamSynchronize(l->ctx->stream);
  delete l;
  return GT_OK;
}
static int add_light(gt_lights* l, int kind, const double* q, int nq, double weight) {
  if (!l) return fail(GT_EINVAL, "lights is null");
  if (!(std::isfinite(weight) && weight > 0.0)) return fail(GT_EINVAL, "sampling target weight must be finite and positive");  // scene.rs:529-537
  double total = l->total + weight;
  if (!std::isfinite(total)) return fail(GT_EINVAL, "sampling target total weight must remain finite");
  gt::HostPrim p;
  std::memset(&p, 0, sizeof p);
  p.kind = kind;
  for (int i = 0; i < nq; ++i) p.q[i] = q[i];
  if (kind == GK_SPHERE) p.q[3] = std::fmax(p.q[3], 0.0);
  if (kind == GK_MOVING_SPHERE) p.q[6] = std::fmax(p.q[6], 0.0);
  l->total = total;
  l->targets.push_back(p);
  l->weights.push_back(weight);
  l->dirty = true;
  return GT_OK;
}
int gt_lights_add_sphere(gt_lights* l, const double c[3], double r, double w) {
  double q[4] = {c[0], c[1], c[2], r};
  return add_light(l, GK_SPHERE, q, 4, w);
}
int gt_lights_add_moving_sphere(gt_lights* l, const double c0[3], const double c1[3], double r, double w) {
  double q[7] = {c0[0], c0[1], c0[2], c1[0], c1[1], c1[2], r};
  return add_light(l, GK_MOVING_SPHERE, q, 7, w);
}
int gt_lights_add_quad(gt_lights* l, const double c[3], const double u[3], const double v[3], double w) {
  double q[9] = {c[0], c[1], c[2], u[0], u[1], u[2], v[0], v[1], v[2]};
  return add_light(l, GK_QUAD, q, 9, w);
}
int gt_lights_add_triangle(gt_lights* l, const double p0[3], const double p1[3], const double p2[3], double w) {
  double q[9] = {p0[0], p0[1], p0[2], p1[0], p1[1], p1[2], p2[0], p2[1], p2[2]};
  return add_light(l, GK_TRIANGLE, q, 9, w);
}
size_t gt_lights_len(const gt_lights* l) { return l ? l->targets.size() : 0; }

int gt_scene_emissive_targets(const gt_scene* s, gt_lights** out) {  // scene.rs:1403-1423
  if (!s || !out) return fail(GT_EINVAL, "null argument");
  gt_lights* l = nullptr;
  int rc = gt_lights_create(s->ctx, &l);
  if (rc != GT_OK) return rc;
  for (const gt::HostPrim& p : s->prims) {
    if (s->mats[p.material].kind != GM_DIFFUSE_LIGHT) continue;
    int nq = p.kind == GK_SPHERE ? 4 : (p.kind == GK_MOVING_SPHERE ? 7 : 9);
    rc = add_light(l, p.kind, p.q, nq, 1.0);
    if (rc != GT_OK) { delete l; return rc; }
  }
  *out = l;
  return GT_OK;
}

// ---- environment ----
int gt_env_create_constant(gt_context* ctx, double r, double g, double b, gt_env** out) {
  if (!ctx || !out) return fail(GT_EINVAL, "null argument");
  if (!finite3(r, g, b)) return fail(GT_EINVAL, "environment colour must be finite");
  gt_env* e = new gt_env();
  e->ctx = ctx;
  e->kind = 1;
  e->color[0] = r; e->color[1] = g; e->color[2] = b;
  e->uniform = true;
  *out = e;
  return GT_OK;
}

int gt_env_create_rgb32f(gt_context* ctx, const float* rgb, uint32_t w, uint32_t h, gt_env** out) {
  if (!ctx || !out || !rgb) return fail(GT_EINVAL, "null argument");
  if (w == 0 || h == 0) return fail(GT_EINVAL, "environment map must be non-empty");
  if (uint64_t(w) * h >= (uint64_t(1) << 31)) return fail(GT_EINVAL, "environment map too large");
  std::unique_ptr<gt_env> e(new gt_env());
  e->ctx = ctx;
  e->kind = 2;
  e->w = w; e->h = h;
  const size_t n = size_t(w) * h;
  std::vector<float4> tex(n);
  std::vector<double> weight(n);
  double total = 0.0;
  const double pi = 3.14159265358979323846;
  // rebuild_distribution (environment.rs:158-184) over linear FP32 texels
  for (uint32_t y = 0; y < h; ++y) {
    double theta = pi * ((double(y) + 0.5) / double(h));
    double st = std::fmax(std::sin(theta), 0.0);
    for (uint32_t x = 0; x < w; ++x) {
      size_t i = size_t(y) * w + x;
      float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
      if (!std::isfinite(r) || !std::isfinite(g) || !std::isfinite(b)) return fail(GT_EINVAL, "environment texels must be finite");
      tex[i] = make_float4(r, g, b, 0.0f);
      double wt = std::fmax(0.2126 * r + 0.7152 * g + 0.0722 * b, 0.0) * st;
      weight[i] = wt;
      total += wt;
    }
  }
  e->uniform = total <= kEps;
  for (uint32_t y = 0; y < h; ++y) {
    double t0 = pi * (double(y) / double(h)), t1 = pi * (double(y + 1) / double(h));
    double sa = (2.0 * pi / double(w)) * std::fmax(std::cos(t0) - std::cos(t1), 0.0);  // :186-190
    for (uint32_t x = 0; x < w; ++x) {
      size_t i = size_t(y) * w + x;
      tex[i].w = (e->uniform || sa <= kEps) ? 0.0f : float((weight[i] / total) / sa);
    }
  }
  std::vector<uint2> alias(n);
  if (!e->uniform) {
    std::vector<uint32_t> thr, idx;
    build_alias(weight, total, &thr, &idx);
    for (size_t i = 0; i < n; ++i) alias[i] = make_uint2(thr[i], idx[i]);
  }
  GT_HIP(hipSetDevice(ctx->device));
  GT_HIP(e->tex.ensure(n)); GT_HIP(e->alias.ensure(n));
  GT_HIP(hipMemcpy(e->tex.p, tex.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  GT_HIP(hipMemcpy(e->alias.p, alias.data(), n * sizeof(uint2), hipMemcpyHostToDevice));
  *out = e.release();
  return GT_OK;
}

int gt_env_destroy(gt_env* e) {
  if (!e) return GT_OK;
  (void)hipStreamSynchronize(e->ctx->stream);
  delete e;
  return GT_OK;
}

// ---- rendering ----
int gt_render_hdr_device(gt_context* ctx, const gt_scene* scene, const gt_camera* camera, const gt_lights* lights,
                         const gt_env* env, void* d_out, int out_is_double, gt_render_stats* stats) {
  if (!d_out) return fail(GT_EINVAL, "output is null");
  return render_impl(ctx, scene, camera, lights, env, d_out, out_is_double, nullptr, nullptr, stats);
}

int gt_render_hdr(gt_context* ctx, const gt_scene* scene, const gt_camera* camera, const gt_lights* lights,
                  const gt_env* env, float* out_rgb, gt_render_stats* stats) {
  if (!out_rgb) return fail(GT_EINVAL, "output is null");
  if (!ctx || !camera) return fail(GT_EINVAL, "null argument");
  GT_HIP(hipSetDevice(ctx->device));
  size_t npix = size_t(camera->image_width) * gt_camera_image_height(camera);
  DevBuf<float>& d = ctx->staging;
  GT_HIP(d.ensure(3 * npix));
  int rc = render_impl(ctx, scene, camera, lights, env, d.p, 0, nullptr, nullptr, stats);
  if (rc != GT_OK) return rc;
  GT_HIP(hipMemcpyAsync(out_rgb, d.p, 3 * npix * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  GT_HIP(hipStreamSynchronize(ctx->stream));
  return GT_OK;
}

int gt_render_aovs_device(gt_context* ctx, const gt_scene* scene, const gt_camera* camera, const gt_lights* lights,
                          const gt_env* env, float* d_beauty, float* d_albedo, float* d_normal,
                          gt_render_stats* stats) {
  if (!d_beauty || !d_albedo || !d_normal) return fail(GT_EINVAL, "output is null");
  return render_impl(ctx, scene, camera, lights, env, d_beauty, 0, d_albedo, d_normal, stats);
}

int gt_render_aovs(gt_context* ctx, const gt_scene* scene, const gt_camera* camera, const gt_lights* lights,
                   const gt_env* env, float* beauty, float* albedo, float* normal, gt_render_stats* stats) {
  if (!beauty || !albedo || !normal) return fail(GT_EINVAL, "output is null");
  if (!ctx || !camera) return fail(GT_EINVAL, "null argument");
  GT_HIP(hipSetDevice(ctx->device));
  size_t npix = size_t(camera->image_width) * gt_camera_image_height(camera);
  DevBuf<float>& d = ctx->staging;
  GT_HIP(d.ensure(9 * npix));
  int rc = render_impl(ctx, scene, camera, lights, env, d.p, 0, d.p + 3 * npix, d.p + 6 * npix, stats);
  if (rc != GT_OK) return rc;
  GT_HIP(hipMemcpyAsync(beauty, d.p, 3 * npix * 4, hipMemcpyDeviceToHost, ctx->stream));
  GT_HIP(hipMemcpyAsync(albedo, d.p + 3 * npix, 3 * npix * 4, hipMemcpyDeviceToHost, ctx->stream));
  GT_HIP(hipMemcpyAsync(normal, d.p + 6 * npix, 3 * npix * 4, hipMemcpyDeviceToHost, ctx->stream));
  GT_HIP(hipStreamSynchronize(ctx->stream));
  return GT_OK;
}

int gt_render_batch_device(gt_context* ctx, size_t n, const gt_scene* const* scenes, const gt_camera* cameras,
                           const gt_lights* const* lights, const gt_env* const* envs, void* const* d_out,
                           int out_is_double, gt_render_stats* stats) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  if (n == 0) {
    if (stats) std::memset(stats, 0, sizeof *stats);
    return GT_OK;
  }
  if (!scenes || !cameras || !d_out) return fail(GT_EINVAL, "null argument");
  if (n > (size_t(1) << 20)) return fail(GT_EINVAL, "too many frames in one batch");
  GT_HIP(hipSetDevice(ctx->device));
  std::vector<FrameJob> jobs(n);
  for (size_t i = 0; i < n; ++i) {
    if (adaptive_camera(&cameras[i]))
      return fail(GT_EUNSUPPORTED, "adaptive sampling renders one frame at a time (gt_render_hdr*)");
    int rc = prepare_job(ctx, scenes[i], &cameras[i], lights ? lights[i] : nullptr, envs ? envs[i] : nullptr, &jobs[i]);
    if (rc != GT_OK) return rc;
    jobs[i].d_out = d_out[i];
  }
  return run_jobs(ctx, jobs, true, out_is_double, stats);
}

int gt_render_batch(gt_context* ctx, size_t n, const gt_scene* const* scenes, const gt_camera* cameras,
                    const gt_lights* const* lights, const gt_env* const* envs, float* const* out_rgb,
                    gt_render_stats* stats) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  if (n && (!cameras || !out_rgb)) return fail(GT_EINVAL, "null argument");
  GT_HIP(hipSetDevice(ctx->device));
  size_t total = 0;
  std::vector<size_t> offs(n);
  for (size_t i = 0; i < n; ++i) {
    offs[i] = total;
    const uint64_t px = uint64_t(cameras[i].image_width) * gt_camera_image_height(&cameras[i]);
    if (px == 0 || px >= (uint64_t(1) << 31)) return fail(GT_EINVAL, "image size out of range");
    total += size_t(px) * 3;
  }
  DevBuf<float>& tmp = ctx->staging;
  GT_HIP(tmp.ensure(std::max<size_t>(total, 1)));
  std::vector<void*> d_out(n);
  for (size_t i = 0; i < n; ++i) d_out[i] = tmp.p + offs[i];
  int rc = gt_render_batch_device(ctx, n, scenes, cameras, lights, envs, d_out.data(), 0, stats);
  if (rc != GT_OK) return rc;
  GT_HIP(hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < n; ++i) {
    const size_t len = (i + 1 < n ? offs[i + 1] : total) - offs[i];
    GT_HIP(hipMemcpy(out_rgb[i], tmp.p + offs[i], len * sizeof(float), hipMemcpyDeviceToHost));
  }
  return GT_OK;
}

int gt_render_progressive(gt_context* ctx, const gt_scene* scene, const gt_camera* camera, const gt_lights* lights,
                          const gt_env* env, uint32_t tile_size, gt_progress_fn progress, void* user,
                          uint8_t* out_canvas) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  if (!out_canvas) return fail(GT_EINVAL, "null canvas");
  GT_HIP(hipSetDevice(ctx->device));
  FrameJob job;
  int rc = prepare_job(ctx, scene, camera, lights, env, &job);
  if (rc != GT_OK) return rc;
  const uint32_t W = job.f.cam.width, H = job.f.cam.height, npix = job.f.npix;
  const uint32_t t = tile_size ? tile_size : 16;
  const uint32_t tiles_x = (W + t - 1) / t, tiles_y = (H + t - 1) / t;
  const uint64_t total_tiles = uint64_t(tiles_x) * tiles_y;
  // tile rows per band: about half a queue of paths, at least one row
  const uint64_t cap = ctx->tuning.max_paths ? ctx->tuning.max_paths : (uint64_t(1) << 25);
  const bool adaptive = adaptive_camera(camera);
  const uint64_t spp = adaptive ? std::max<uint32_t>(camera->adaptive_min_samples, 1) : std::max<uint32_t>(job.spp, 1);
  const uint64_t row_paths = uint64_t(W) * t * spp;
  const uint32_t band_rows = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(tiles_y, cap / 2 / row_paths)));
  std::memset(out_canvas, 0, size_t(npix) * 3);
  GT_HIP(ctx->staging.ensure(size_t(npix) * 3));
  GT_HIP(ctx->band_map.ensure(npix));
  gtk::Launch L{grid_for(ctx), ctx->tuning.block_size, ctx->stream};
  std::vector<uint32_t> ids;
  uint64_t done = 0;
  for (uint32_t ty0 = 0; ty0 < tiles_y; ty0 += band_rows) {
    const uint32_t ty1 = std::min(tiles_y, ty0 + band_rows);
    const uint32_t y0 = ty0 * t, y1 = std::min(H, ty1 * t);
    const uint32_t n = (y1 - y0) * W;
    ids.resize(n);
    for (uint32_t i = 0; i < n; ++i) ids[i] = y0 * W + i;
    GT_HIP(hipMemcpyAsync(ctx->band_map.p, ids.data(), size_t(n) * sizeof(uint32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    FrameJob j = job;
    j.d_out = ctx->staging.p;
    if (adaptive) {
      if ((rc = render_adaptive(ctx, j, camera, 0, nullptr, ctx->band_map.p, n)) != GT_OK) return rc;
    } else {
      j.f.pixmap = ctx->band_map.p;
      j.f.nact = n;
      j.d_out = nullptr;
      std::vector<FrameJob> jobs(1, j);
      if ((rc = run_jobs(ctx, jobs, false, 0, nullptr)) != GT_OK) return rc;
      gtk::finalize_map(L, jobs[0].f.acc, ctx->band_map.p, n, ctx->staging.p, 0);
      GT_HIP(hipGetLastError());
    }
    // the band's display rows (Rgb::from_linear_color) into the host canvas
    if ((rc = gt_encode_image(ctx, ctx->staging.p + size_t(y0) * W * 3, 0, 1, W, y1 - y0, GT_ENCODE_CANVAS,
                              nullptr, out_canvas + size_t(y0) * W * 3, 0)) != GT_OK)
      return rc;
    for (uint32_t ty = ty0; ty < ty1; ++ty)
      for (uint32_t tx = 0; tx < tiles_x; ++tx) {
        gt_render_progress u{};
        u.tile.x = tx * t;
        u.tile.y = ty * t;
        u.tile.width = std::min(W, (tx + 1) * t) - tx * t;
        u.tile.height = std::min(H, (ty + 1) * t) - ty * t;
        u.completed_tiles = ++done;
        u.total_tiles = total_tiles;
        u.image_width = W;
        u.image_height = H;
        u.pixels = out_canvas;
        if (progress) {
          const int r = progress(user, &u);
          if (r != 0) return fail(GT_ECANCELLED, "progressive render stopped by its callback (" + std::to_string(r) + ")");
        }
      }
  }
  return GT_OK;
}

int gt_render_pixels_device(gt_context* ctx, const gt_scene* scene, const gt_camera* camera,
                            const gt_lights* lights, const gt_env* env, const uint32_t* pixels, size_t n,
                            void* d_out, int out_is_double, gt_render_stats* stats) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  GT_HIP(hipSetDevice(ctx->device));
  FrameJob job;
  int rc = prepare_job(ctx, scene, camera, lights, env, &job);
  if (rc != GT_OK) return rc;
  if (stats) std::memset(stats, 0, sizeof *stats);
  if (n == 0) return GT_OK;
  if (!pixels || !d_out) return fail(GT_EINVAL, "null pixel list or output");
  const uint32_t npix = job.f.npix;
  if (n > npix) return fail(GT_EINVAL, "more pixels listed than the image has");
  // every listed pixel owns one accumulator: indices in range, each at most once
  std::vector<uint8_t> seen(npix, 0);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t px = pixels[i];
    if (px >= npix) return fail(GT_EINVAL, "pixel index " + std::to_string(px) + " outside the image");
    if (seen[px]) return fail(GT_EINVAL, "pixel index " + std::to_string(px) + " listed twice");
    seen[px] = 1;
  }
  GT_HIP(ctx->band_map.ensure(n));
  GT_HIP(hipMemcpyAsync(ctx->band_map.p, pixels, n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
  GT_HIP(hipStreamSynchronize(ctx->stream));
  job.d_out = d_out;
  if (adaptive_camera(camera))
    return render_adaptive(ctx, job, camera, out_is_double, stats, ctx->band_map.p, uint32_t(n));
  job.f.pixmap = ctx->band_map.p;
  job.f.nact = uint32_t(n);
  job.d_out = nullptr;
  std::vector<FrameJob> jobs(1, job);
  if ((rc = run_jobs(ctx, jobs, false, out_is_double, stats)) != GT_OK) return rc;
  gtk::Launch L{grid_for(ctx), ctx->tuning.block_size, ctx->stream};
  gtk::finalize_map(L, jobs[0].f.acc, ctx->band_map.p, uint32_t(n), d_out, out_is_double);
  GT_HIP(hipGetLastError());
  return GT_OK;
}

int gt_autotune(gt_context* ctx, size_t n, const gt_scene* const* scenes, const gt_camera* cameras,
                const gt_lights* const* lights, const gt_env* const* envs, uint32_t sample_spp,
                gt_tuning* out_best, double* out_best_ms) {
  if (!ctx) return fail(GT_EINVAL, "context is null");
  if (n == 0 || !scenes || !cameras) return fail(GT_EINVAL, "nothing to tune on");
  GT_HIP(hipSetDevice(ctx->device));
  const gt_tuning start = ctx->tuning;
  // the workload: every frame at its own spp (or sample_spp random samples), no output
  std::vector<gt_camera> cams(cameras, cameras + n);
  uint64_t paths = 0;
  for (gt_camera& c : cams) {
    if (sample_spp) {
      c.samples_per_pixel = sample_spp;
      c.pixel_sample_mode = GT_PIXEL_RANDOM;
    }
    paths += uint64_t(c.image_width) * gt_camera_image_height(&c) * gt_camera_effective_spp(&c);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  GT_HIP(hipEventCreate(&e0));
  GT_HIP(hipEventCreate(&e1));
  struct Ev {
    hipEvent_t a, b;
    ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  } guard{e0, e1};
  // time one candidate: a warm run (allocations, code load), then a timed one
  auto measure = [&](const gt_tuning& t, double* ms_out) -> int {
    int rc = gt_context_set_tuning(ctx, &t);
    if (rc != GT_OK) return rc;
    std::vector<FrameJob> jobs(n);
    for (size_t i = 0; i < n; ++i) {
      rc = prepare_job(ctx, scenes[i], &cams[i], lights ? lights[i] : nullptr, envs ? envs[i] : nullptr, &jobs[i]);
      if (rc != GT_OK) return rc;
    }
    const bool multi = n > 1;
    if ((rc = run_jobs(ctx, jobs, multi, 0, nullptr)) != GT_OK) return rc;
    GT_HIP(hipEventRecord(e0, ctx->stream));
    if ((rc = run_jobs(ctx, jobs, multi, 0, nullptr)) != GT_OK) return rc;
    GT_HIP(hipEventRecord(e1, ctx->stream));
    GT_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GT_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_out = ms;
    return GT_OK;
  };
  gt_tuning best = start;
  best.count_traversal = 0;
  best.use_graph = 0;
  best.max_paths = uint64_t(1) << 25;
  double best_ms = 1e300;
  // stage 1: threads per block x BVH width x LDS stack depth at the default queue
  const uint32_t blocks[] = {64, 128, 256};
  const uint32_t widths[] = {2, 4};
  const uint32_t stacks[] = {8, 16};
  const gt_tuning base = best;
  for (uint32_t width : widths)
    for (uint32_t block : blocks)
      for (uint32_t stack : stacks) {
        gt_tuning t = base;
        t.block_size = block;
        t.bvh_width = width;
        t.stack_lds = stack;
        double ms = 0;
        int rc = measure(t, &ms);
        if (rc == GT_EINVAL) continue;  // shape not allowed (LDS)
        if (rc != GT_OK) {
          ctx->tuning = start;
          return rc;
        }
        if (ms < best_ms) { best_ms = ms; best = t; }
      }
  // stage 2: queue capacity for the winner (capacities beyond twice the workload add nothing)
  const gt_tuning winner = best;
  for (uint64_t cap : {uint64_t(1) << 23, uint64_t(1) << 24, uint64_t(1) << 26}) {
    if (cap > 2 * std::max<uint64_t>(paths, 1)) continue;
    gt_tuning t = winner;
    t.max_paths = cap;
    double ms = 0;
    int rc = measure(t, &ms);
    if (rc == GT_EINVAL || rc == GT_ENOMEM) continue;
    if (rc != GT_OK) {
      ctx->tuning = start;
      return rc;
    }
    if (ms < best_ms) { best_ms = ms; best = t; }
  }
  best.use_graph = start.use_graph;
  best.count_traversal = start.count_traversal;
  ctx->tuning = best;
  if (out_best) *out_best = best;
  if (out_best_ms) *out_best_ms = best_ms;
  return GT_OK;
}

}  // extern "C"
