// gartus.hpp — header-only C++ host API over the C ABI in gartus_hip.h.
//
// Mirrors the reference's Rust host types for the path-tracing hot path:
//   RayMaterial      src/graphics/raytracing/material.rs:2355 (the five GPU variants)
//   RayScene         src/graphics/raytracing/scene.rs:1102-1578
//   SamplingTargetList / WeightedSamplingTargetList  scene.rs:363-609
//   EnvironmentLight src/graphics/raytracing/environment.rs:22-195 (+ FP32 HDR maps)
//   RayCamera        src/graphics/camera.rs:612-1557 (builder subset used by renders)
//   PathTracer       src/graphics/raytracing/renderer.rs:55-886
//   HdrImage, ToneMap src/graphics/display.rs:164-468 (encodings on the device)
//   DenoisingAovs    src/graphics/camera.rs:491-526
// Names and argument meaning follow the reference. Where the reference
// panics on invalid input (assert!), these types throw gartus::Panic.
#pragma once

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gartus_hip.h"

namespace gartus {

struct Panic : std::runtime_error {
  int code;
  Panic(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != GT_OK) throw Panic(rc, gt_last_error());
}

using Vec3 = std::array<double, 3>;

// ---- context -------------------------------------------------------------
class Context {
 public:
  explicit Context(int device = 0) {
    gt_context* c = nullptr;
    check(gt_context_create(device, &c));
    ctx_.reset(c, [](gt_context* p) { gt_context_destroy(p); });
  }
  // No device: a placeholder that only compares and copies.
  explicit Context(std::nullptr_t) {}
  gt_context* get() const { return ctx_.get(); }
  bool valid() const { return ctx_ != nullptr; }
  gt_tuning tuning() const {
    gt_tuning t{};
    check(gt_context_get_tuning(ctx_.get(), &t));
    return t;
  }
  void set_tuning(const gt_tuning& t) { check(gt_context_set_tuning(ctx_.get(), &t)); }

 private:
  std::shared_ptr<gt_context> ctx_;
};

// ---- materials (RayMaterial) ----------------------------------------------
struct RayMaterial {
  enum Kind { Lambertian, DiffuseLight, GgxMicrofacet, Metal, Dielectric } kind;
  Vec3 color{0, 0, 0};
  double param = 0.0;  // roughness, fuzz or refraction index

  static RayMaterial lambertian(Vec3 albedo) { return {Lambertian, albedo, 0.0}; }
  static RayMaterial diffuse_light(Vec3 emit) { return {DiffuseLight, emit, 0.0}; }
  static RayMaterial ggx_microfacet(Vec3 specular, double roughness) { return {GgxMicrofacet, specular, roughness}; }
  static RayMaterial metal(Vec3 albedo, double fuzz) { return {Metal, albedo, fuzz}; }
  static RayMaterial dielectric(double refraction_index) { return {Dielectric, {1, 1, 1}, refraction_index}; }
  bool is_emissive() const { return kind == DiffuseLight; }
};

using MaterialId = uint32_t;
class SamplingTargetList;

// ---- RayScene ---------------------------------------------------------------
// Materials and primitives live in the device-side scene; the BVH is built on
// first render (RayScene's lazy OnceLock, scene.rs:1567-1573) and rebuilt
// after any mutation (scene.rs:1575-1577).
class RayScene {
 public:
  explicit RayScene(const Context& ctx) : ctx_(ctx) {
    gt_scene* s = nullptr;
    check(gt_scene_create(ctx_.get(), &s));
    scene_.reset(s, [](gt_scene* p) { gt_scene_destroy(p); });
  }
  MaterialId add_material(const RayMaterial& m) {
    uint32_t id = 0;
    const Vec3& c = m.color;
    switch (m.kind) {
      case RayMaterial::Lambertian: check(gt_scene_add_lambertian(s(), c[0], c[1], c[2], &id)); break;
      case RayMaterial::DiffuseLight: check(gt_scene_add_diffuse_light(s(), c[0], c[1], c[2], &id)); break;
      case RayMaterial::GgxMicrofacet: check(gt_scene_add_ggx(s(), c[0], c[1], c[2], m.param, &id)); break;
      case RayMaterial::Metal: check(gt_scene_add_metal(s(), c[0], c[1], c[2], m.param, &id)); break;
      case RayMaterial::Dielectric: check(gt_scene_add_dielectric(s(), m.param, &id)); break;
    }
    return id;
  }
  void add_sphere(Vec3 center, double radius, MaterialId m) {
    double d[4] = {center[0], center[1], center[2], radius};
    check(gt_scene_add_spheres(s(), d, 1, &m));
  }
  void add_moving_sphere(Vec3 c0, Vec3 c1, double radius, MaterialId m) {
    double d[7] = {c0[0], c0[1], c0[2], c1[0], c1[1], c1[2], radius};
    check(gt_scene_add_moving_spheres(s(), d, 1, &m));
  }
  void add_triangle(Vec3 p0, Vec3 p1, Vec3 p2, MaterialId m) {
    double d[9] = {p0[0], p0[1], p0[2], p1[0], p1[1], p1[2], p2[0], p2[1], p2[2]};
    check(gt_scene_add_triangles(s(), d, 1, &m));
  }
  void add_quad(Vec3 corner, Vec3 u, Vec3 v, MaterialId m) {
    double d[9] = {corner[0], corner[1], corner[2], u[0], u[1], u[2], v[0], v[1], v[2]};
    check(gt_scene_add_quads(s(), d, 1, &m));
  }
  // Bulk forms: `data` packs n primitives in the layouts of gartus_hip.h.
  void add_triangles(const std::vector<double>& data, const std::vector<MaterialId>& mats) {
    check(gt_scene_add_triangles(s(), data.data(), mats.size(), mats.data()));
  }
  void add_spheres(const std::vector<double>& data, const std::vector<MaterialId>& mats) {
    check(gt_scene_add_spheres(s(), data.data(), mats.size(), mats.data()));
  }
  MaterialId add_sphere_with_material(Vec3 center, double radius, const RayMaterial& m) {
    MaterialId id = add_material(m);
    add_sphere(center, radius, id);
    return id;
  }
  MaterialId add_quad_with_material(Vec3 corner, Vec3 u, Vec3 v, const RayMaterial& m) {
    MaterialId id = add_material(m);
    add_quad(corner, u, v, id);
    return id;
  }
  void build_bvh() { build_bvh_with_options(4); }  // BvhBuildOptions::default (bvh.rs:67-71)
  void build_bvh_with_options(uint32_t leaf_size) { check(gt_scene_build(s(), leaf_size)); }
  RayScene& with_bvh() { build_bvh(); return *this; }
  gt_scene_info info() const {
    gt_scene_info i{};
    check(gt_scene_get_info(s(), &i));
    return i;
  }
  size_t len() const { return size_t(info().primitive_count); }
  bool is_empty() const { return len() == 0; }
  size_t material_count() const { return size_t(info().material_count); }
  size_t bvh_node_count() const { return size_t(info().bvh_node_count); }
  inline SamplingTargetList emissive_targets() const;

  gt_scene* s() const { return scene_.get(); }
  const Context& context() const { return ctx_; }

 private:
  Context ctx_;
  std::shared_ptr<gt_scene> scene_;
};

// ---- light targets ------------------------------------------------------------
class SamplingTargetList {
 public:
  explicit SamplingTargetList(const Context& ctx) {
    gt_lights* l = nullptr;
    check(gt_lights_create(ctx.get(), &l));
    lights_.reset(l, [](gt_lights* p) { gt_lights_destroy(p); });
  }
  explicit SamplingTargetList(gt_lights* owned) { lights_.reset(owned, [](gt_lights* p) { gt_lights_destroy(p); }); }
  void add_sphere_weighted(Vec3 c, double r, double w) { check(gt_lights_add_sphere(l(), c.data(), r, w)); }
  void add_moving_sphere_weighted(Vec3 c0, Vec3 c1, double r, double w) {
    check(gt_lights_add_moving_sphere(l(), c0.data(), c1.data(), r, w));
  }
  void add_quad_weighted(Vec3 q, Vec3 u, Vec3 v, double w) { check(gt_lights_add_quad(l(), q.data(), u.data(), v.data(), w)); }
  void add_triangle_weighted(Vec3 a, Vec3 b, Vec3 c, double w) {
    check(gt_lights_add_triangle(l(), a.data(), b.data(), c.data(), w));
  }
  void add_sphere(Vec3 c, double r) { add_sphere_weighted(c, r, 1.0); }
  void add_quad(Vec3 q, Vec3 u, Vec3 v) { add_quad_weighted(q, u, v, 1.0); }
  void add_triangle(Vec3 a, Vec3 b, Vec3 c) { add_triangle_weighted(a, b, c, 1.0); }
  size_t len() const { return gt_lights_len(l()); }
  bool is_empty() const { return len() == 0; }
  gt_lights* l() const { return lights_.get(); }

 private:
  std::shared_ptr<gt_lights> lights_;
};
using WeightedSamplingTargetList = SamplingTargetList;

inline SamplingTargetList RayScene::emissive_targets() const {
  gt_lights* l = nullptr;
  check(gt_scene_emissive_targets(s(), &l));
  return SamplingTargetList(l);
}

// ---- environment --------------------------------------------------------------
class EnvironmentLight {
 public:
  static EnvironmentLight constant(const Context& ctx, Vec3 c) {
    gt_env* e = nullptr;
    check(gt_env_create_constant(ctx.get(), c[0], c[1], c[2], &e));
    return EnvironmentLight(e);
  }
  // Linear RGB texels, row 0 at the +y pole, width*height*3 floats.
  static EnvironmentLight from_rgb32f(const Context& ctx, const std::vector<float>& rgb, uint32_t w, uint32_t h) {
    if (rgb.size() != size_t(w) * h * 3) throw Panic(GT_EINVAL, "environment texel count must match dimensions");
    gt_env* e = nullptr;
    check(gt_env_create_rgb32f(ctx.get(), rgb.data(), w, h, &e));
    return EnvironmentLight(e);
  }
  gt_env* e() const { return env_.get(); }

 private:
  explicit EnvironmentLight(gt_env* e) { env_.reset(e, [](gt_env* p) { gt_env_destroy(p); }); }
  std::shared_ptr<gt_env> env_;
};

// ---- camera ---------------------------------------------------------------------
enum class PixelSampleMode : uint32_t { Random = GT_PIXEL_RANDOM, Stratified = GT_PIXEL_STRATIFIED };
enum class DirectLightingMode : uint32_t { Continuation = GT_LIGHTING_CONTINUATION, NextEventEstimation = GT_LIGHTING_NEE };

// Builder over gt_camera; validation happens at render, as in the reference's
// initialize() (camera.rs:1068-1130, validate_view :1161-1201).
class RayCamera {
 public:
  static RayCamera make(uint32_t image_width, double aspect_ratio) {
    RayCamera c;
    gt_camera_default(&c.c_);
    c.c_.image_width = image_width;
    c.c_.aspect_ratio = aspect_ratio;
    if (image_width == 0) throw Panic(GT_EINVAL, "image width must be greater than zero");
    return c;
  }
  RayCamera with_samples_per_pixel(uint32_t n) const { auto c = *this; c.c_.samples_per_pixel = n ? n : 1; return c; }
  // with_adaptive_sampling (camera.rs:1244-1264; AdaptiveSampling::new :377-393)
  RayCamera with_adaptive_sampling(uint32_t min_samples, uint32_t max_samples, double error_threshold) const {
    if (!(std::isfinite(error_threshold) && error_threshold > 0.0))
      throw Panic(GT_EINVAL, "adaptive sampling error threshold must be positive and finite");
    auto c = *this;
    c.c_.adaptive_min_samples = std::max<uint32_t>(min_samples, 1);
    c.c_.adaptive_max_samples = std::max(max_samples, c.c_.adaptive_min_samples);
    c.c_.adaptive_error_threshold = error_threshold;
    c.c_.samples_per_pixel = c.c_.adaptive_max_samples;
    return c;
  }
  RayCamera without_adaptive_sampling() const {
    auto c = *this;
    c.c_.adaptive_min_samples = c.c_.adaptive_max_samples = 0;
    c.c_.adaptive_error_threshold = 0.0;
    return c;
  }
  RayCamera with_pixel_sample_mode(PixelSampleMode m) const { auto c = *this; c.c_.pixel_sample_mode = uint32_t(m); return c; }
  RayCamera with_stratified_sampling() const { return with_pixel_sample_mode(PixelSampleMode::Stratified); }
  RayCamera with_stratified_grid_width(uint32_t g) const {
    auto c = *this;
    c.c_.pixel_sample_mode = GT_PIXEL_STRATIFIED_GRID;
    c.c_.stratified_grid_width = g ? g : 1;
    return c;
  }
  RayCamera with_max_depth(uint32_t d) const { auto c = *this; c.c_.max_depth = d; return c; }
  RayCamera with_russian_roulette_min_depth(uint32_t d) const { auto c = *this; c.c_.russian_roulette_min_depth = int32_t(d); return c; }
  RayCamera without_russian_roulette() const { auto c = *this; c.c_.russian_roulette_min_depth = -1; return c; }
  RayCamera with_rng_seed(uint64_t s) const { auto c = *this; c.c_.rng_seed = s; return c; }
  RayCamera with_vertical_fov(double f) const { auto c = *this; c.c_.vertical_fov = f; return c; }
  RayCamera with_look_at(Vec3 from, Vec3 at) const {
    auto c = *this;
    for (int i = 0; i < 3; ++i) { c.c_.lookfrom[i] = from[i]; c.c_.lookat[i] = at[i]; }
    return c;
  }
  RayCamera with_view_up(Vec3 up) const { auto c = *this; for (int i = 0; i < 3; ++i) c.c_.view_up[i] = up[i]; return c; }
  RayCamera with_defocus_angle(double a) const { auto c = *this; c.c_.defocus_angle = a; return c; }
  RayCamera with_focus_distance(double d) const { auto c = *this; c.c_.focus_distance = d; return c; }
  RayCamera with_shutter_interval(double s, double e) const { auto c = *this; c.c_.shutter_start = s; c.c_.shutter_end = e; return c; }
  RayCamera with_background(Vec3 bg) const {
    auto c = *this;
    c.c_.background_kind = GT_BACKGROUND_CONSTANT;
    for (int i = 0; i < 3; ++i) c.c_.background[i] = bg[i];
    return c;
  }
  RayCamera with_vertical_gradient_background(Vec3 nadir, Vec3 horizon, Vec3 zenith) const {
    auto c = *this;
    c.c_.background_kind = GT_BACKGROUND_VERTICAL_GRADIENT;
    for (int i = 0; i < 3; ++i) {
      c.c_.background[i] = nadir[i];
      c.c_.background[3 + i] = horizon[i];
      c.c_.background[6 + i] = zenith[i];
    }
    return c;
  }
  RayCamera with_direct_lighting_mode(DirectLightingMode m) const { auto c = *this; c.c_.direct_lighting_mode = uint32_t(m); return c; }
  RayCamera with_light_pdf_weight(double w) const { auto c = *this; c.c_.light_pdf_weight = w; return c; }

  uint32_t image_width() const { return c_.image_width; }
  uint32_t image_height() const { return gt_camera_image_height(&c_); }
  uint32_t samples_per_pixel() const { return c_.samples_per_pixel; }
  uint32_t effective_samples_per_pixel() const { return gt_camera_effective_spp(&c_); }
  const gt_camera& raw() const { return c_; }

 private:
  gt_camera c_{};
};

// ---- images ---------------------------------------------------------------------
enum class ToneMappingOperator { None, Reinhard, Aces };

// ToneMap (display.rs:191-278): exposure, operator, gamma; map_color runs on
// the device (gt_encode_image, GT_ENCODE_CANVAS).
struct ToneMap {
  double exposure = 1.0;
  ToneMappingOperator op = ToneMappingOperator::None;
  double gamma = 2.0;
  ToneMap with_exposure(double e) const {
    if (!(std::isfinite(e) && e >= 0.0)) throw Panic(GT_EINVAL, "tone-map exposure must be finite and non-negative");
    auto t = *this; t.exposure = e; return t;
  }
  ToneMap with_operator(ToneMappingOperator o) const { auto t = *this; t.op = o; return t; }
  ToneMap with_gamma(double g) const {
    if (!(std::isfinite(g) && g > 0.0)) throw Panic(GT_EINVAL, "tone-map gamma must be positive and finite");
    auto t = *this; t.gamma = g; return t;
  }
  gt_tone_map raw() const { return gt_tone_map{exposure, gamma, uint32_t(op), 0}; }
};

// The header text gt_encoded_header gives for a format.
inline std::string encoded_header(int format, uint32_t width, uint32_t height) {
  char buf[128];
  size_t n = 0;
  check(gt_encoded_header(format, width, height, buf, sizeof buf, &n));
  return std::string(buf, std::min(n, sizeof buf));
}

// HdrImage (display.rs:282-441): linear RGB, row-major, top row first. The
// conversions and file bodies are computed on the device of `context` (the
// rendering context for rendered images).
struct HdrImage {
  uint32_t width = 0, height = 0;
  std::vector<float> pixels;  // width * height * 3
  Context context{nullptr};

  HdrImage() = default;
  HdrImage(const Context& ctx, uint32_t w, uint32_t h, std::vector<float> px)
      : width(w), height(h), pixels(std::move(px)), context(ctx) {
    if (pixels.size() != size_t(w) * h * 3) throw Panic(GT_EINVAL, "HDR image pixel count must match dimensions");
  }

  std::vector<uint8_t> to_canvas(const ToneMap& tm = ToneMap()) const { return encode<uint8_t>(GT_ENCODE_CANVAS, tm); }
  std::vector<uint8_t> to_canvas_tone_mapped(const ToneMap& tm) const { return to_canvas(tm); }
  // Rgb::from_raw_linear_color per pixel (colors.rs:412-422): AOV canvases.
  std::vector<uint8_t> to_raw_canvas() const { return encode<uint8_t>(GT_ENCODE_RAW, ToneMap()); }
  std::vector<uint8_t> rgbe() const { return encode<uint8_t>(GT_ENCODE_RGBE, ToneMap()); }
  // write_pfm (display.rs:365-387): little-endian f32, bottom row first,
  // non-finite values written as 0.
  void save_pfm(const std::string& path) const { write(path, GT_ENCODE_PFM, encode<float>(GT_ENCODE_PFM, ToneMap())); }
  // write_radiance_hdr (display.rs:397-409): flat RGBE, top row first.
  void save_radiance_hdr(const std::string& path) const { write(path, GT_ENCODE_RGBE, rgbe()); }
  // save_tone_mapped (display.rs:437-440) as a binary PPM (the reference's
  // Canvas::write_binary_ppm, display.rs:1744-1754).
  void save_tone_mapped(const std::string& path, const ToneMap& tm) const {
    const std::string head = "P6\n" + std::to_string(width) + " " + std::to_string(height) + "\n255\n";
    write_bytes(path, head, to_canvas(tm));
  }
  // save_extension (display.rs:412-431): .pfm, .hdr/.rgbe keep HDR data;
  // .ppm is the default canvas. Other raster formats are the reference's
  // Canvas encoders and are not provided here.
  void save_extension(const std::string& path) const {
    const size_t dot = path.rfind('.');
    std::string ext = dot == std::string::npos ? "" : path.substr(dot + 1);
    for (char& c : ext) c = char(std::tolower(static_cast<unsigned char>(c)));
    if (ext == "pfm") save_pfm(path);
    else if (ext == "hdr" || ext == "rgbe") save_radiance_hdr(path);
    else if (ext == "ppm") save_tone_mapped(path, ToneMap());
    else throw Panic(GT_EINVAL, "unsupported image extension: " + path);
  }

 private:
  template <class T>
  std::vector<T> encode(int format, const ToneMap& tm) const {
    if (!context.valid()) throw Panic(GT_EINVAL, "HdrImage has no device context");
    std::vector<T> out(gt_encoded_size(format, width, height) / sizeof(T));
    const gt_tone_map t = tm.raw();
    check(gt_encode_image(context.get(), pixels.data(), 0, 0, width, height, format, &t, out.data(), 0));
    return out;
  }
  template <class T>
  void write(const std::string& path, int format, const std::vector<T>& body) const {
    write_bytes(path, encoded_header(format, width, height), body);
  }
  template <class T>
  static void write_bytes(const std::string& path, const std::string& head, const std::vector<T>& body) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open " + path);
    std::fwrite(head.data(), 1, head.size(), f);
    std::fwrite(body.data(), sizeof(T), body.size(), f);
    if (std::fclose(f) != 0) throw std::runtime_error("cannot write " + path);
  }
};

// DenoisingAovs (camera.rs:491-526), linear forms.
struct DenoisingAovs {
  HdrImage beauty, albedo, normal;
};

// ---- PathTracer -----------------------------------------------------------------
class PathTracer {
 public:
  explicit PathTracer(const RayCamera& camera) : camera_(camera) {}
  static PathTracer make(const RayCamera& camera) { return PathTracer(camera); }
  PathTracer with_camera(const RayCamera& camera) const { return PathTracer(camera); }
  const RayCamera& camera() const { return camera_; }

  // renderer.rs:112-119 / :193-201
  HdrImage render_hdr_image(const RayScene& world) const { return render(world, camera_, nullptr, nullptr); }
  HdrImage render_ray_scene_hdr_image(const RayScene& scene) const { return render_hdr_image(scene); }
  std::vector<uint8_t> render(const RayScene& world) const { return render_hdr_image(world).to_canvas(); }
  // render_progressive (renderer.rs:130-144; camera.rs:3172-3294): the display
  // canvas, with progress(update) after each tile on this thread. progress
  // returns false to stop; the render then throws Panic(GT_ECANCELLED).
  using Progress = std::function<bool(const gt_render_progress&)>;
  std::vector<uint8_t> render_progressive(const RayScene& world, const Progress& progress,
                                          uint32_t tile_size = 0) const {
    return progressive(world, camera_, nullptr, progress, tile_size);
  }
  std::vector<uint8_t> render_with_lights_progressive(const RayScene& world, const SamplingTargetList& lights,
                                                      const Progress& progress, uint32_t tile_size = 0) const {
    return progressive(world, camera_, &lights, progress, tile_size);
  }
  // renderer.rs:391-407: lights force next-event estimation only through
  // the light-connection entry points; this one keeps the camera's strategy.
  HdrImage render_with_lights_hdr_image(const RayScene& world, const SamplingTargetList& lights) const {
    return render(world, camera_, &lights, nullptr);
  }
  // light_connection_camera (camera.rs:4167-4169)
  HdrImage render_with_light_connections_hdr_image(const RayScene& world, const SamplingTargetList& lights) const {
    return render(world, nee(camera_), &lights, nullptr);
  }
  // environment_importance_camera (camera.rs:4171-4173): environment renders use NEE
  HdrImage render_with_environment_hdr_image(const RayScene& world, const EnvironmentLight& env) const {
    return render(world, nee(camera_), nullptr, &env);
  }
  HdrImage render_with_lights_and_environment_hdr_image(const RayScene& world, const SamplingTargetList& lights,
                                                        const EnvironmentLight& env) const {
    return render(world, nee(camera_), &lights, &env);
  }
  // renderer.rs:121-129 / :217-224
  DenoisingAovs render_denoising_aovs(const RayScene& world) const { return aovs(world, camera_, nullptr); }
  DenoisingAovs render_with_lights_denoising_aovs(const RayScene& world, const SamplingTargetList& lights) const {
    return aovs(world, camera_, &lights);
  }

 private:
  static RayCamera nee(const RayCamera& c) { return c.with_direct_lighting_mode(DirectLightingMode::NextEventEstimation); }
  static HdrImage blank(const RayScene& world, const RayCamera& c) {
    return HdrImage(world.context(), c.image_width(), c.image_height(),
                    std::vector<float>(size_t(c.image_width()) * c.image_height() * 3, 0.0f));
  }
  static HdrImage render(const RayScene& world, const RayCamera& cam, const SamplingTargetList* lights,
                         const EnvironmentLight* env) {
    HdrImage img = blank(world, cam);
    check(gt_render_hdr(world.context().get(), world.s(), &cam.raw(), lights ? lights->l() : nullptr,
                        env ? env->e() : nullptr, img.pixels.data(), nullptr));
    return img;
  }
  static std::vector<uint8_t> progressive(const RayScene& world, const RayCamera& cam,
                                          const SamplingTargetList* lights, const Progress& progress,
                                          uint32_t tile_size) {
    std::vector<uint8_t> canvas(size_t(cam.image_width()) * cam.image_height() * 3);
    struct Hook {
      const Progress* fn;
      static int call(void* user, const gt_render_progress* u) {
        const Progress& f = *static_cast<Hook*>(user)->fn;
        return (!f || f(*u)) ? 0 : 1;
      }
    } hook{&progress};
    check(gt_render_progressive(world.context().get(), world.s(), &cam.raw(), lights ? lights->l() : nullptr,
                                nullptr, tile_size, &Hook::call, &hook, canvas.data()));
    return canvas;
  }
  static DenoisingAovs aovs(const RayScene& world, const RayCamera& cam, const SamplingTargetList* lights) {
    DenoisingAovs a{blank(world, cam), blank(world, cam), blank(world, cam)};
    check(gt_render_aovs(world.context().get(), world.s(), &cam.raw(), lights ? lights->l() : nullptr, nullptr,
                         a.beauty.pixels.data(), a.albedo.pixels.data(), a.normal.pixels.data(), nullptr));
    return a;
  }
  RayCamera camera_;
};

}  // namespace gartus
