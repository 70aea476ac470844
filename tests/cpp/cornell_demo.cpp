// cornell_demo.cpp — the C++ host API (gartus.hpp) used the way the
// reference's own examples use RayScene / RayCamera / PathTracer.
//
//   cornell_demo host            camera, tone-map and header checks that need no GPU
//   cornell_demo render OUT.pfm  Cornell box (scenes.rs:226-294) with light
//                                connections, 64x64, 2x2 stratified, depth 8
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "gartus.hpp"

using gartus::Vec3;

namespace {

int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    ++failures;
  }
}

// rotate_y_point / rotate_y_vector (instance.rs:379-385)
Vec3 rotate_y(const Vec3& p, double s, double c) { return {c * p[0] + s * p[2], p[1], -s * p[0] + c * p[2]}; }

// box_object (object.rs:930-976): six (corner, u, v) from the min and max corners.
void add_box(gartus::RayScene& scene, const Vec3& size, double angle_deg, const Vec3& offset, gartus::MaterialId m) {
  const double rad = angle_deg * M_PI / 180.0;
  const double s = std::sin(rad), c = std::cos(rad);
  const Vec3 mn{0.0, 0.0, 0.0}, mx = size;
  const Vec3 dx{mx[0] - mn[0], 0.0, 0.0}, dy{0.0, mx[1] - mn[1], 0.0}, dz{0.0, 0.0, mx[2] - mn[2]};
  auto neg = [](const Vec3& v) { return Vec3{-v[0], -v[1], -v[2]}; };
  const Vec3 q[6][3] = {
      {{mn[0], mn[1], mx[2]}, dx, dy},      {{mx[0], mn[1], mx[2]}, neg(dz), dy},
      {{mx[0], mn[1], mn[2]}, neg(dx), dy}, {{mn[0], mn[1], mn[2]}, dz, dy},
      {{mn[0], mx[1], mx[2]}, dx, neg(dz)}, {{mn[0], mn[1], mn[2]}, dx, dz},
  };
  for (const auto& f : q) {
    Vec3 rc = rotate_y(f[0], s, c);
    scene.add_quad({rc[0] + offset[0], rc[1] + offset[1], rc[2] + offset[2]}, rotate_y(f[1], s, c),
                   rotate_y(f[2], s, c), m);
  }
}

int host_checks() {
  using gartus::RayCamera;
  RayCamera cam = RayCamera::make(400, 16.0 / 9.0);
  expect(cam.image_height() == 225, "400 x 16/9 gives height 225 (camera.rs:1071)");
  expect(RayCamera::make(1920, 16.0 / 9.0).image_height() == 1080, "1920 x 16/9 gives 1080");
  expect(cam.samples_per_pixel() == 1, "default spp is 1 (camera.rs:1043)");
  expect(cam.with_stratified_grid_width(8).effective_samples_per_pixel() == 64, "grid 8 takes 64 samples");
  expect(cam.raw().max_depth == 10 && cam.raw().russian_roulette_min_depth == 5 && cam.raw().rng_seed == 1,
         "defaults: depth 10, roulette from 5, seed 1 (camera.rs:1041-1064)");
  bool threw = false;
  try {
    RayCamera::make(0, 1.0);
  } catch (const gartus::Panic& p) {
    threw = p.code == GT_EINVAL;
  }
  expect(threw, "zero image width panics (camera.rs:1035)");
  threw = false;
  try {
    gartus::ToneMap().with_gamma(0.0);
  } catch (const gartus::Panic& p) {
    threw = p.code == GT_EINVAL;
  }
  expect(threw, "gamma 0 panics (display.rs:251-257)");
  expect(gartus::encoded_header(GT_ENCODE_PFM, 2, 1) == "PF\n2 1\n-1.0\n", "PFM header (display.rs:366-369)");
  expect(gartus::encoded_header(GT_ENCODE_RGBE, 3, 2) == "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 3\n",
         "RGBE header (display.rs:398-401)");
  expect(gt_encoded_size(GT_ENCODE_RGBE, 3, 2) == 24, "RGBE body size");
  threw = false;
  try {
    gartus::HdrImage img;
    img.width = img.height = 1;
    img.pixels = {1, 1, 1};
    img.to_canvas();
  } catch (const gartus::Panic& p) {
    threw = p.code == GT_EINVAL;
  }
  expect(threw, "an image without a device context cannot encode");
  std::printf("host checks: %d failure(s)\n", failures);
  return failures == 0 ? 0 : 1;
}

int render(const char* out) {
  gartus::Context ctx;
  gartus::RayScene scene(ctx);
  using gartus::RayMaterial;
  auto green = scene.add_material(RayMaterial::lambertian({0.12, 0.45, 0.15}));
  auto red = scene.add_material(RayMaterial::lambertian({0.65, 0.05, 0.05}));
  auto light = scene.add_material(RayMaterial::diffuse_light({15.0, 15.0, 15.0}));
  auto white = scene.add_material(RayMaterial::lambertian({0.73, 0.73, 0.73}));
  const Vec3 lc{343.0, 554.0, 332.0}, lu{-130.0, 0.0, 0.0}, lv{0.0, 0.0, -105.0};
  scene.add_quad({555.0, 0.0, 0.0}, {0.0, 555.0, 0.0}, {0.0, 0.0, 555.0}, green);
  scene.add_quad({0.0, 0.0, 0.0}, {0.0, 555.0, 0.0}, {0.0, 0.0, 555.0}, red);
  scene.add_quad(lc, lu, lv, light);
  scene.add_quad({0.0, 0.0, 0.0}, {555.0, 0.0, 0.0}, {0.0, 0.0, 555.0}, white);
  scene.add_quad({555.0, 555.0, 555.0}, {-555.0, 0.0, 0.0}, {0.0, 0.0, -555.0}, white);
  scene.add_quad({0.0, 0.0, 555.0}, {555.0, 0.0, 0.0}, {0.0, 555.0, 0.0}, white);
  add_box(scene, {165.0, 330.0, 165.0}, 15.0, {265.0, 0.0, 295.0}, white);
  add_box(scene, {165.0, 165.0, 165.0}, -18.0, {130.0, 0.0, 65.0}, white);
  scene.build_bvh();
  gartus::SamplingTargetList lights(ctx);
  lights.add_quad(lc, lu, lv);
  auto cam = gartus::RayCamera::make(64, 1.0)
                 .with_stratified_grid_width(2)
                 .with_max_depth(8)
                 .with_vertical_fov(40.0)
                 .with_look_at({278.0, 278.0, -800.0}, {278.0, 278.0, 0.0})
                 .with_background({0.0, 0.0, 0.0});
  gartus::HdrImage img = gartus::PathTracer(cam).render_with_light_connections_hdr_image(scene, lights);
  img.save_pfm(out);
  // the device encoders on a small image of known values
  gartus::HdrImage k(ctx, 2, 1, {4.0f, 1.0f, 0.25f, 1.0f, 0.5f, NAN});
  std::vector<uint8_t> c = k.to_canvas();
  expect(c[0] == 255 && c[1] == 255 && c[2] == 128 && c[5] == 0, "gamma-2 canvas (display.rs:2158-2163)");
  c = k.to_canvas(gartus::ToneMap().with_operator(gartus::ToneMappingOperator::Reinhard).with_gamma(1.0));
  expect(c[0] == 204 && c[1] == 128 && c[2] == 51, "Reinhard canvas (display.rs:2164-2169)");
  std::vector<uint8_t> e = k.rgbe();
  expect(e[4] == 128 && e[5] == 64 && e[6] == 0 && e[7] == 129, "RGBE of (1, 0.5, NaN)");
  const std::string tmp = std::string(out) + ".check.pfm";
  k.save_pfm(tmp);
  std::FILE* f = std::fopen(tmp.c_str(), "rb");
  char head[32] = {0};
  size_t n = std::fread(head, 1, 12, f);
  float px[6];
  size_t m = std::fread(px, sizeof(float), 6, f);
  std::fclose(f);
  std::remove(tmp.c_str());
  expect(n == 12 && std::strncmp(head, "PF\n2 1\n-1.0\n", 12) == 0, "PFM header");
  expect(m == 6 && px[0] == 4.0f && px[5] == 0.0f, "PFM body, non-finite written as 0");
  // progressive: 64x64 in tiles of 16 reports 16 tiles in order; false stops it
  {
    uint64_t seen = 0;
    bool in_order = true;
    std::vector<uint8_t> prog = gartus::PathTracer(cam).render_with_lights_progressive(
        scene, lights, [&](const gt_render_progress& u) {
          in_order = in_order && u.completed_tiles == seen + 1;
          seen = u.completed_tiles;
          return true;
        }, 16);
    (void)prog;
    expect(seen == 16 && in_order, "16 tiles of 16 reported in order");
    bool threw_cancel = false;
    try {
      gartus::PathTracer(cam).render_with_lights_progressive(scene, lights,
                                                             [](const gt_render_progress&) { return false; });
    } catch (const gartus::Panic& p) {
      threw_cancel = p.code == GT_ECANCELLED;
    }
    expect(threw_cancel, "a callback returning false stops the render");
  }
  if (failures) return 1;
  std::printf("rendered %ux%u, %zu primitives, %zu lights -> %s\n", img.width, img.height, scene.len(), lights.len(),
              out);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::string(argv[1]) == "host") return host_checks();
    if (argc >= 3 && std::string(argv[1]) == "render") return render(argv[2]);
  } catch (const gartus::Panic& p) {
    std::fprintf(stderr, "gartus panic (%d): %s\n", p.code, p.what());
    return 2;
  }
  std::fprintf(stderr, "usage: %s host | render OUT.pfm\n", argv[0]);
  return 64;
}
