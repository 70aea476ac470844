// cpu_ref.cpp — CPU oracle for the gartus path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, in f64 C++, the reference
// algorithm of gartus's PathTracer over a RayScene so that the HIP path can
// be checked against it. Only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg may load it, and only as the checker. The product never
// links it.
//
// Every function cites the reference file:line it follows (paths are relative
// to the reference crate root). Arithmetic keeps the reference's operation
// order; build with -ffp-contract=off so no FMA contraction changes rounding.
// Rust's powi(n) with a constant exponent lowers to square-and-multiply
// (x^5 = x * ((x*x)*(x*x))); it is written out that way here.
//
// Two RNG modes (gt_camera.rng_mode):
//   GT_RNG_PIXEL_STREAM — the reference's one sequential SplitMix64 stream per
//     pixel, seeded by pixel_seed(seed, x, y) (camera.rs:2669) and consumed by
//     all samples and bounces in order.
//   GT_RNG_COUNTER — the stream used by the GPU: one SplitMix64 stream per
//     (pixel, sample, dimension), dimension 0 = camera ray, 1+b = bounce b,
//     seeded by stream_seed() below. Draws inside a dimension follow the
//     reference order exactly.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "../include/gartus_hip.h"

namespace cr {

constexpr double EPS = std::numeric_limits<double>::epsilon();  // f64::EPSILON
constexpr double PI = 3.14159265358979323846;
constexpr double TAU = 6.28318530717958647692;
constexpr double INF = std::numeric_limits<double>::infinity();
constexpr double SHADOW_ACNE_EPSILON = 0.001;       // raytracing.rs:146
constexpr double SHADOW_ACNE_PDF_EPSILON = 0.001;   // object.rs:1163
constexpr double TRIANGLE_BOUNDS_EPSILON = 1e-10;   // geometry.rs:9
constexpr double TRIANGLE_HIT_EPSILON = 1e-10;      // geometry.rs:10
constexpr double QUAD_BOUNDS_EPSILON = 1e-4;        // geometry.rs:11
constexpr double QUAD_HIT_EPSILON = 1e-8;           // geometry.rs:12
constexpr double GGX_MIN_ALPHA = 1.0e-4;            // sampling.rs:12
constexpr double RR_MIN = 0.05, RR_MAX = 0.95;      // camera.rs:31-32
constexpr int BVH_BUCKETS = 12;                     // bvh.rs:4

// Rust f64::clamp: NaN passes through.
static inline double clampd(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
// Rust's `f64 as u8`: truncation toward zero, saturating at 0 and 255, NaN -> 0.
static inline uint8_t as_u8(double x) { return std::isnan(x) || x <= 0.0 ? 0 : (x >= 255.0 ? 255 : uint8_t(x)); }
// f64::powi as Rust lowers it (the llvm.powi intrinsic, compiler-rt's
// __powidf2): binary powering of |b|, then a reciprocal for b < 0, so
// 2^-1024 evaluates as 1 / inf = 0.
static inline double powi_f64(double a, int b) {
  const bool recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}
// Rust f64::min / max: a NaN operand yields the other operand.
static inline double fminr(double a, double b) { return std::fmin(a, b); }
static inline double fmaxr(double a, double b) { return std::fmax(a, b); }
static inline double pow5(double x) { double x2 = x * x; double x4 = x2 * x2; return x * x4; }
static inline double rem_euclid(double x, double m) {  // f64::rem_euclid
  double r = std::fmod(x, m);
  return r < 0.0 ? r + std::fabs(m) : r;
}
static inline double degrees_to_radians(double d) { return d * PI / 180.0; }  // raytracing.rs:150-152

// ---- vectors (src/gmath/vector.rs) ----------------------------------------
struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }  // Vector / f64
// LinearRgb / f64 (colors.rs:185-190) multiplies by the reciprocal.
static inline V3 color_div(V3 a, double s) { const double r = 1.0 / s; return {r * a.x, r * a.y, r * a.z}; }
static inline V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }  // component_mul
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // vector.rs:133-135
static inline V3 cross(V3 a, V3 b) {  // vector.rs:121-129
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline double length_squared(V3 a) { return dot(a, a); }
static inline double length(V3 a) { return std::sqrt(length_squared(a)); }
static inline V3 normalized(V3 a) {  // vector.rs:151-158
  double len = length(a);
  if (len < EPS) return V3();
  return a / len;
}
static inline V3 reflected(V3 d, V3 n) { return d - 2.0 * dot(d, n) * n; }  // vector.rs:162-164
static inline V3 refracted(V3 d, V3 n, double eta) {  // vector.rs:168-173
  double cos_theta = fminr(dot(-d, n), 1.0);
  V3 perp = eta * (d + cos_theta * n);
  V3 par = -std::sqrt(std::fabs(1.0 - length_squared(perp))) * n;
  return perp + par;
}
static inline bool finite3(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
static inline double max_component(V3 c) { return fmaxr(fmaxr(c.x, c.y), c.z); }  // colors.rs:126-128

struct Ray {
  V3 o, d;
  double time = 0;
  V3 at(double t) const { return o + t * d; }  // ray.rs:112-114
};

// ---- RNG (src/gmath/random.rs:12-152) --------------------------------------
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t state;
  explicit Rng(uint64_t s = 1) : state(s) {}
  uint64_t next_u64() {  // random.rs:23-29
    state += 0x9e3779b97f4a7c15ULL;
    return mix64(state);
  }
  double random_double() { return double(next_u64() >> 11) * (1.0 / double(1ULL << 53)); }  // :56-59
  double random_range(double lo, double hi) { return lo + (hi - lo) * random_double(); }  // :63-65
  V3 unit_vector_spherical() {  // :107-115
    double r1 = random_double(), r2 = random_double();
    double z = 1.0 - 2.0 * r2;
    double radius = std::sqrt(1.0 - z * z);
    double phi = TAU * r1;
    return {std::cos(phi) * radius, std::sin(phi) * radius, z};
  }
  V3 cosine_direction() {  // :119-126
    double r1 = random_double(), r2 = random_double();
    double phi = TAU * r1;
    double radius = std::sqrt(r2);
    return {std::cos(phi) * radius, std::sin(phi) * radius, std::sqrt(1.0 - r2)};
  }
  V3 in_unit_disk() {  // :141-152
    for (;;) {
      double a = random_range(-1.0, 1.0);
      double b = random_range(-1.0, 1.0);
      V3 p(a, b, 0.0);
      if (length_squared(p) < 1.0) return p;
    }
  }
};

// camera.rs:1720-1727
static inline uint64_t pixel_seed(uint64_t seed, uint32_t x, uint32_t y) {
  uint64_t z = seed ^ (uint64_t(x) * 0x9E3779B97F4A7C15ULL) ^ (uint64_t(y) * 0xD1B54A32D192ED03ULL);
  return mix64(z);
}
// Counter-mode stream key for (pixel, sample, dimension). Defined by this
// build (the reference has one sequential stream per pixel); the same
// expression runs on the GPU.
static inline uint64_t stream_seed(uint64_t pixel_key, uint32_t sample, uint32_t dim) {
  uint64_t z = pixel_key ^ (uint64_t(sample + 1) * 0xD6E8FEB86659FD93ULL) ^
               (uint64_t(dim + 1) * 0xA0761D6478BD642FULL);
  return mix64(z);
}

struct PathRng {
  int mode;
  uint64_t pixel_key;
  Rng rng;
  PathRng(int m, uint64_t key) : mode(m), pixel_key(key), rng(key) {}
  void begin(uint32_t sample, uint32_t dim) {
    if (mode == GT_RNG_COUNTER) rng.state = stream_seed(pixel_key, sample, dim);
  }
};

// ---- Bounds3 (src/gmath/polygon_matrix.rs:19-221) ---------------------------
struct Aabb {
  V3 lo, hi;
  static Aabb from_points(V3 a, V3 b) {  // :28-33
    return {V3(fminr(a.x, b.x), fminr(a.y, b.y), fminr(a.z, b.z)),
            V3(fmaxr(a.x, b.x), fmaxr(a.y, b.y), fmaxr(a.z, b.z))};
  }
  Aabb unite(const Aabb& o) const {  // :36-50
    return {V3(fminr(lo.x, o.lo.x), fminr(lo.y, o.lo.y), fminr(lo.z, o.lo.z)),
            V3(fmaxr(hi.x, o.hi.x), fmaxr(hi.y, o.hi.y), fmaxr(hi.z, o.hi.z))};
  }
  Aabb union_point(V3 p) const { return unite(from_points(p, p)); }
  Aabb padded(double e) const {  // :84-106
    e = std::fabs(e);
    double half = 0.5 * e;
    Aabb r = *this;
    if (std::fabs(r.hi.x - r.lo.x) < e) { r.lo.x -= half; r.hi.x += half; }
    if (std::fabs(r.hi.y - r.lo.y) < e) { r.lo.y -= half; r.hi.y += half; }
    if (std::fabs(r.hi.z - r.lo.z) < e) { r.lo.z -= half; r.hi.z += half; }
    return r;
  }
  V3 centroid() const {  // :115-121
    return {0.5 * (lo.x + hi.x), 0.5 * (lo.y + hi.y), 0.5 * (lo.z + hi.z)};
  }
  double surface_area() const {  // :124-129
    double dx = fmaxr(hi.x - lo.x, 0.0), dy = fmaxr(hi.y - lo.y, 0.0), dz = fmaxr(hi.z - lo.z, 0.0);
    return 2.0 * (dx * dy + dx * dz + dy * dz);
  }
  int largest_axis() const {  // :133-144
    double dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    if (dx >= dy && dx >= dz) return 0;
    if (dy >= dz) return 1;
    return 2;
  }
};

// ---- primitives (src/gmath/geometry.rs, src/graphics/raytracing/object.rs) -
enum { K_SPHERE = 0, K_MOVING_SPHERE = 1, K_TRIANGLE = 2, K_QUAD = 3 };
// cr_scene_add_prims kind of a mesh triangle with vertex normals (MeshTriangle,
// mesh.rs:24-151): stored as a K_TRIANGLE with `smooth` set.
enum { K_MESH_TRIANGLE = 4 };

struct Prim {
  int kind = K_SPHERE;
  V3 a, b, c;        // sphere: a=center; moving: a=c0, b=c1; triangle: a,b,c; quad: a=corner, b=u, c=v
  double radius = 0;
  V3 qn, qw;         // quad normal and w (geometry.rs:293-312)
  double qd = 0;     // quad plane d
  bool smooth = false;  // mesh triangle: vertex normals na, nb, nc (mesh.rs:24-151)
  V3 na, nb, nc;
  uint32_t material = 0;
};

static Prim make_sphere(V3 c, double r) {
  Prim p; p.kind = K_SPHERE; p.a = c; p.radius = fmaxr(r, 0.0); return p;  // geometry.rs:445-450
}
static Prim make_moving_sphere(V3 c0, V3 c1, double r) {
  Prim p; p.kind = K_MOVING_SPHERE; p.a = c0; p.b = c1; p.radius = fmaxr(r, 0.0); return p;  // :498-504
}
static Prim make_triangle(V3 a, V3 b, V3 c) {
  Prim p; p.kind = K_TRIANGLE; p.a = a; p.b = b; p.c = c; return p;
}
static Prim make_quad(V3 corner, V3 u, V3 v) {  // geometry.rs:293-312
  Prim p; p.kind = K_QUAD; p.a = corner; p.b = u; p.c = v;
  V3 n = cross(u, v);
  p.qn = normalized(n);
  p.qd = dot(p.qn, corner);
  double nls = length_squared(n);
  p.qw = nls <= EPS ? V3() : n / nls;
  return p;
}
static inline V3 moving_center(const Prim& p, double time) { return p.a + time * (p.b - p.a); }  // geometry.rs:507-509

struct Surf {
  V3 point, normal;
  V3 shading;        // HitRecord::shading_normal: the normal unless a mesh sets it (object.rs:470-479)
  double t = 0, u = 0, v = 0;
  bool front = false;
};
// object.rs:127-149
static inline Surf with_uv(const Ray& r, V3 p, V3 outward, double t, double u, double v) {
  Surf s;
  s.point = p; s.t = t; s.u = u; s.v = v;
  s.front = dot(r.d, outward) < 0.0;
  s.normal = s.front ? outward : -outward;
  s.shading = s.normal;
  return s;
}

// object.rs:986-1013 (ray_t.surrounds = strict interval)
static bool hit_sphere_in_interval(V3 center, double radius, const Ray& r, double tmin, double tmax, double* root) {
  V3 oc = center - r.o;
  double a = length_squared(r.d);
  if (a <= EPS) return false;
  double h = dot(r.d, oc);
  double c = length_squared(oc) - radius * radius;
  double disc = h * h - a * c;
  if (disc < 0.0) return false;
  double sq = std::sqrt(disc);
  double first = (h - sq) / a;
  if (tmin < first && first < tmax) { *root = first; return true; }
  double second = (h + sq) / a;
  if (tmin < second && second < tmax) { *root = second; return true; }
  return false;
}
static inline V3 sphere_outward(V3 center, double radius, V3 p) {  // geometry.rs:476-483
  V3 out = p - center;
  if (std::fabs(radius) <= EPS) return normalized(out);
  return out / radius;
}
static inline void sphere_uv(V3 n, double* u, double* v) {  // object.rs:1038-1042
  double theta = std::acos(-n.y);
  double phi = std::atan2(-n.z, n.x) + PI;
  *u = phi / (2.0 * PI);
  *v = theta / PI;
}
struct TriHit { double t, u, v; };
static bool triangle_hit_ray(const Prim& p, const Ray& r, double tmin, double tmax, TriHit* out) {  // geometry.rs:239-265
  V3 e1 = p.b - p.a, e2 = p.c - p.a;
  V3 pvec = cross(r.d, e2);
  double det = dot(e1, pvec);
  if (std::fabs(det) < TRIANGLE_HIT_EPSILON) return false;
  double inv_det = 1.0 / det;
  V3 tvec = r.o - p.a;
  double u = dot(tvec, pvec) * inv_det;
  if (!(u >= 0.0 && u <= 1.0)) return false;
  V3 qvec = cross(tvec, e1);
  double v = dot(r.d, qvec) * inv_det;
  if (v < 0.0 || u + v > 1.0) return false;
  double t = dot(e2, qvec) * inv_det;
  if (!(tmin < t && t < tmax)) return false;
  *out = {t, u, v};
  return true;
}
static bool quad_hit_ray(const Prim& p, const Ray& r, double tmin, double tmax, TriHit* out) {  // geometry.rs:394-415
  if (length_squared(cross(p.b, p.c)) <= EPS) return false;
  double denom = dot(p.qn, r.d);
  if (std::fabs(denom) < QUAD_HIT_EPSILON) return false;
  double t = (p.qd - dot(p.qn, r.o)) / denom;
  if (!(tmin < t && t < tmax)) return false;
  V3 planar = r.at(t) - p.a;  // plane_coordinates :384-390
  double a = dot(p.qw, cross(planar, p.c));
  double b = dot(p.qw, cross(p.b, planar));
  if (!(a >= 0.0 && a <= 1.0) || !(b >= 0.0 && b <= 1.0)) return false;
  *out = {t, a, b};
  return true;
}
// Intersect::intersect per geometry (object.rs:163-234) and RayGeometry (:338-346)
static bool prim_intersect(const Prim& p, const Ray& r, double tmin, double tmax, Surf* s) {
  switch (p.kind) {
    case K_SPHERE:
    case K_MOVING_SPHERE: {
      V3 center = p.kind == K_SPHERE ? p.a : moving_center(p, r.time);
      double root;
      if (!hit_sphere_in_interval(center, p.radius, r, tmin, tmax, &root)) return false;
      V3 point = r.at(root);
      V3 outward = sphere_outward(center, p.radius, point);
      double u, v;
      sphere_uv(outward, &u, &v);
      *s = with_uv(r, point, outward, root, u, v);
      return true;
    }
    case K_TRIANGLE: {
      TriHit h;
      if (!triangle_hit_ray(p, r, tmin, tmax, &h)) return false;
      V3 normal = normalized(cross(p.b - p.a, p.c - p.a));  // geometry.rs:176-185
      if (length_squared(normal) <= EPS) return false;
      *s = with_uv(r, r.at(h.t), normal, h.t, h.u, h.v);
      if (p.smooth) {  // MeshTriangle::hit / shading_normal (mesh.rs:53-119), set_shading_normal (object.rs:470-479)
        const double w = 1.0 - h.u - h.v;
        V3 sn = (w * p.na) + (h.u * p.nb) + (h.v * p.nc);
        if (length_squared(sn) > EPS) {
          sn = normalized(sn);
          s->shading = s->front ? sn : -sn;
        }
      }
      return true;
    }
    default: {
      TriHit h;
      if (!quad_hit_ray(p, r, tmin, tmax, &h)) return false;
      V3 normal = p.qn;
      if (length_squared(normal) <= EPS) return false;
      *s = with_uv(r, r.at(h.t), normal, h.t, h.u, h.v);
      return true;
    }
  }
}
static Aabb prim_bounds(const Prim& p) {
  switch (p.kind) {
    case K_SPHERE: {  // object.rs:1015-1022
      V3 rv(p.radius, p.radius, p.radius);
      return Aabb::from_points(p.a - rv, p.a + rv);
    }
    case K_MOVING_SPHERE: {  // object.rs:1024-1036
      V3 rv(p.radius, p.radius, p.radius);
      return Aabb::from_points(p.a - rv, p.a + rv).unite(Aabb::from_points(p.b - rv, p.b + rv));
    }
    case K_TRIANGLE:  // geometry.rs:197-203
      return Aabb::from_points(p.a, p.a).union_point(p.b).union_point(p.c).padded(TRIANGLE_BOUNDS_EPSILON);
    default: {  // geometry.rs:339-362
      V3 p0 = p.a, p1 = p.a + p.b, p2 = p.a + p.c, p3 = p.a + p.b + p.c;
      return Aabb::from_points(p0, p0).union_point(p1).union_point(p2).union_point(p3).padded(QUAD_BOUNDS_EPSILON);
    }
  }
}

// ---- ONB and PDFs (geometry.rs:122-137, sampling.rs) -------------------------
struct Onb {
  V3 u, v, w;
  V3 local(V3 l) const { return l.x * u + l.y * v + l.z * w; }  // geometry.rs:159-161
};
static bool onb_from_w(V3 axis, Onb* out) {  // geometry.rs:122-137
  V3 w = normalized(axis);
  if (length_squared(w) <= EPS) return false;
  V3 a = std::fabs(w.x) > 0.9 ? V3(0.0, 1.0, 0.0) : V3(1.0, 0.0, 0.0);
  V3 v = normalized(cross(w, a));
  V3 u = cross(w, v);
  *out = {u, v, w};
  return true;
}
static double ggx_alpha(double r) { double c = clampd(r, 0.0, 1.0); return fmaxr(c * c, GGX_MIN_ALPHA); }  // sampling.rs:310-312
static double ggx_D(double ndh, double roughness) {  // sampling.rs:229-241
  if (!std::isfinite(ndh) || ndh <= 0.0 || !std::isfinite(roughness)) return 0.0;
  double alpha = ggx_alpha(roughness);
  double a2 = alpha * alpha;
  double cos2 = ndh * ndh;
  double den = cos2 * (a2 - 1.0) + 1.0;
  if (den <= EPS) return 0.0;
  return a2 / (PI * den * den);
}
static double smith_g1(double ndd, double roughness) {  // sampling.rs:314-322
  if (!std::isfinite(ndd) || ndd <= 0.0 || !std::isfinite(roughness)) return 0.0;
  double alpha = ggx_alpha(roughness);
  double a2 = alpha * alpha;
  double cos2 = ndd * ndd;
  return 2.0 * ndd / (ndd + std::sqrt(a2 + (1.0 - a2) * cos2));
}
static double schlick(double cosine, double f0) {  // sampling.rs:255-261
  if (!std::isfinite(cosine) || !std::isfinite(f0)) return 0.0;
  f0 = clampd(f0, 0.0, 1.0);
  return f0 + (1.0 - f0) * pow5(1.0 - clampd(cosine, 0.0, 1.0));
}

// HenyeyGreensteinPdf::phase_value (sampling.rs:126-136)
static double hg_phase(double cosine, double g) {
  if (!std::isfinite(cosine) || !std::isfinite(g) || std::fabs(g) >= 1.0) return 0.0;
  double den = 1.0 + g * g - 2.0 * g * clampd(cosine, -1.0, 1.0);
  if (den <= EPS) return 0.0;
  return (1.0 - g * g) / (4.0 * PI * std::pow(den, 1.5));
}
struct MatPdf {  // MaterialPdf::{Cosine, GgxReflection, Sphere, HenyeyGreenstein} (pdf.rs:46-104)
  int kind = 0;  // 0 cosine, 1 ggx, 2 sphere, 3 henyey-greenstein
  Onb basis;
  V3 outgoing;
  double roughness = 0, alpha = 0;
  double g = 0;
  double value(V3 d) const {
    if (kind == 2) return 1.0 / (4.0 * PI);  // SpherePdf (sampling.rs:33-36)
    if (kind == 3) {  // sampling.rs:139-147
      double len = length(d);
      if (len <= EPS) return 0.0;
      return hg_phase(dot(d, basis.w) / len, g);
    }
    if (kind == 0) {  // sampling.rs:65-77
      double len = length(d);
      if (len <= EPS) return 0.0;
      double c = dot(d, basis.w) / len;
      return c <= 0.0 ? 0.0 : c / PI;
    }
    V3 in = normalized(d);  // sampling.rs:269-281
    if (length_squared(in) <= EPS || dot(in, basis.w) <= 0.0) return 0.0;
    V3 half = normalized(in + outgoing);
    double odh = dot(outgoing, half);
    if (length_squared(half) <= EPS || odh <= EPS) return 0.0;
    double ndh = dot(half, basis.w);  // half_vector_pdf :263-266
    return ggx_D(ndh, roughness) * fmaxr(ndh, 0.0) / (4.0 * std::fabs(odh));
  }
  // *below (when given) reports a GGX reflection below the surface, where
  // the reference returns the outgoing direction (sampling.rs:313-318).
  V3 generate(Rng& rng, bool* below = nullptr) const {
    if (kind == 2) return rng.unit_vector_spherical();  // SpherePdf (sampling.rs:38-40)
    if (kind == 3) {  // sampling.rs:149-165
      double r1 = rng.random_double(), r2 = rng.random_double();
      double ct;
      if (std::fabs(g) < 1e-3) {
        ct = 1.0 - 2.0 * r1;
      } else {
        double term = (1.0 - g * g) / (1.0 - g + 2.0 * g * r1);
        ct = clampd((1.0 + g * g - term * term) / (2.0 * g), -1.0, 1.0);
      }
      double st = std::sqrt(fmaxr(1.0 - ct * ct, 0.0));
      double phi = 2.0 * PI * r2;
      return basis.local(V3(std::cos(phi) * st, std::sin(phi) * st, ct));
    }
    if (kind == 0) return basis.local(rng.cosine_direction());  // sampling.rs:79-81
    double r1 = rng.random_double(), r2 = rng.random_double();  // sampling.rs:283-307
    double phi = 2.0 * PI * r1;
    double a2 = alpha * alpha;
    double cr2 = fminr(r2, 1.0 - EPS);
    double tan2 = a2 * cr2 / (1.0 - cr2);
    double ct = 1.0 / std::sqrt(1.0 + tan2);
    double st = std::sqrt(fmaxr(1.0 - ct * ct, 0.0));
    V3 half = basis.local(V3(std::cos(phi) * st, std::sin(phi) * st, ct));
    V3 refl = reflected(-outgoing, normalized(half));
    const bool up = dot(refl, basis.w) > 0.0;
    if (below) *below = !up;
    return up ? refl : outgoing;
  }
};
static bool ggx_pdf_new(V3 normal, V3 outgoing, double roughness, MatPdf* out) {  // sampling.rs:183-202
  MatPdf p;
  p.kind = 1;
  if (!onb_from_w(normal, &p.basis)) return false;
  V3 o = normalized(outgoing);
  if (length_squared(o) <= EPS || dot(o, p.basis.w) <= 0.0) return false;
  p.outgoing = o;
  p.roughness = clampd(roughness, 0.0, 1.0);
  p.alpha = ggx_alpha(p.roughness);
  *out = p;
  return true;
}
// material.rs:1686-1722
static double ggx_weight(V3 normal, V3 outgoing, V3 incoming, double roughness, double f0) {
  normal = normalized(normal);
  outgoing = normalized(outgoing);
  incoming = normalized(incoming);
  double ndi = dot(normal, incoming), ndo = dot(normal, outgoing);
  if (ndi <= 0.0 || ndo <= 0.0) return 0.0;
  V3 half = normalized(incoming + outgoing);
  if (length_squared(half) <= EPS) return 0.0;
  double ndh = dot(normal, half), odh = dot(outgoing, half);
  if (ndh <= 0.0 || odh <= 0.0) return 0.0;
  double D = ggx_D(ndh, roughness);
  double G = smith_g1(ndi, roughness) * smith_g1(ndo, roughness);
  f0 = clampd(f0, 0.0, 1.0);
  if (f0 <= EPS) return 0.0;
  double F = schlick(odh, f0) / f0;
  return D * G * F / (4.0 * ndo);
}
static double specular_luminance(V3 c) { return clampd(0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z, 0.0, 1.0); }  // :1724-1726

// ---- materials (src/graphics/raytracing/material.rs) -------------------------
enum { M_LAMBERTIAN = 0, M_DIFFUSE_LIGHT = 1, M_GGX = 2, M_METAL = 3, M_DIELECTRIC = 4, M_ISOTROPIC = 5,
       M_HENYEY_GREENSTEIN = 6 };
struct Mat {
  int kind = M_LAMBERTIAN;
  V3 color;
  double param = 0;  // ggx roughness, metal fuzz, dielectric ior, henyey-greenstein g
};

struct Hit {
  Surf s;
  int64_t prim = -1;
  const Mat* mat = nullptr;
};

struct Scatter {
  int kind = 0;  // 0 absorbed, 1 specular, 2 scattering
  Ray ray;
  V3 attenuation;
  MatPdf pdf;
};

static Scatter mat_scatter(const Mat& m, const Ray& in, const Hit& h, Rng& rng) {  // dispatch :2646-2661
  Scatter sc;
  const V3 shading = h.s.shading;  // the geometric normal unless a mesh triangle interpolates one
  switch (m.kind) {
    case M_LAMBERTIAN: {  // :508-520
      MatPdf p;
      p.kind = 0;
      if (!onb_from_w(shading, &p.basis)) return sc;
      sc.kind = 2; sc.attenuation = m.color; sc.pdf = p;
      return sc;
    }
    case M_GGX: {  // :1583-1594
      V3 outgoing = -normalized(in.d);
      MatPdf p;
      if (!ggx_pdf_new(shading, outgoing, m.param, &p)) return sc;
      sc.kind = 2; sc.attenuation = m.color; sc.pdf = p;
      return sc;
    }
    case M_METAL: {  // :2087-2107
      V3 refl = reflected(normalized(in.d), shading);
      V3 dir = refl + m.param * rng.unit_vector_spherical();
      if (dot(dir, h.s.normal) <= 0.0) return sc;
      sc.kind = 1; sc.ray = {h.s.point, dir, in.time}; sc.attenuation = m.color;
      return sc;
    }
    case M_DIELECTRIC: {  // :2321-2350
      double ratio = h.s.front ? 1.0 / m.param : m.param;
      V3 unit = normalized(in.d);
      double cos_theta = fminr(dot(-unit, h.s.normal), 1.0);
      double sin_theta = std::sqrt(1.0 - cos_theta * cos_theta);
      bool cannot = ratio * sin_theta > 1.0;
      V3 dir;
      double r0 = (1.0 - ratio) / (1.0 + ratio);  // reflectance :2214-2218
      r0 = r0 * r0;
      if (cannot || r0 + (1.0 - r0) * pow5(1.0 - cos_theta) > rng.random_double())
        dir = reflected(unit, h.s.normal);
      else
        dir = refracted(unit, h.s.normal, ratio);
      sc.kind = 1; sc.ray = {h.s.point, dir, in.time}; sc.attenuation = V3(1.0, 1.0, 1.0);
      return sc;
    }
    case M_ISOTROPIC: {  // :1123-1137
      sc.kind = 2; sc.attenuation = m.color; sc.pdf.kind = 2;
      return sc;
    }
    case M_HENYEY_GREENSTEIN: {  // :1270-1284, HenyeyGreensteinPdf::new (sampling.rs:96-112)
      MatPdf p;
      p.kind = 3;
      p.g = m.param;
      if (!onb_from_w(in.d, &p.basis)) return sc;
      sc.kind = 2; sc.attenuation = m.color; sc.pdf = p;
      return sc;
    }
    default:
      return sc;  // DiffuseLight does not scatter
  }
}
static double mat_scattering_pdf(const Mat& m, const Ray& in, const Hit& h, V3 scattered) {
  if (m.kind == M_LAMBERTIAN) {  // :522-529
    double c = dot(h.s.shading, normalized(scattered));
    return c <= 0.0 ? 0.0 : c / PI;
  }
  if (m.kind == M_GGX) {  // :1596-1612
    V3 outgoing = -normalized(in.d);
    V3 incoming = normalized(scattered);
    if (dot(incoming, h.s.normal) <= 0.0 || dot(outgoing, h.s.normal) <= 0.0) return 0.0;
    return ggx_weight(h.s.shading, outgoing, incoming, m.param, specular_luminance(m.color));
  }
  if (m.kind == M_ISOTROPIC) return 1.0 / (4.0 * PI);  // :1139-1141
  if (m.kind == M_HENYEY_GREENSTEIN) {  // :1286-1291
    MatPdf p;
    p.kind = 3;
    p.g = m.param;
    if (!onb_from_w(in.d, &p.basis)) return 0.0;
    return p.value(scattered);
  }
  return 0.0;
}
static V3 mat_emitted(const Mat& m, const Hit& h) {  // :1009-1022
  if (m.kind == M_DIFFUSE_LIGHT && h.s.front) return m.color;
  return V3();
}
static V3 mat_denoise_albedo(const Mat& m) {  // :534, :1024, :1618, :2109, :2295
  if (m.kind == M_DIELECTRIC) return V3(1.0, 1.0, 1.0);
  return m.color;
}

// ---- BVH (src/graphics/raytracing/bvh.rs) ------------------------------------
struct Stats {
  uint64_t rays = 0, node_tests = 0, node_hits = 0, leaf_visits = 0, candidates = 0, leaf_hits = 0, max_stack = 0;
};
struct Node {
  Aabb bounds;
  bool leaf = true;
  uint64_t a = 0, b = 0;  // leaf: first,count; internal: left,right
};
struct Traversal {  // bvh.rs:92-147
  double o[3], inv[3];
  bool par[3];
  explicit Traversal(const Ray& r) {
    double d[3] = {r.d.x, r.d.y, r.d.z};
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z;
    for (int i = 0; i < 3; ++i) {
      par[i] = std::fabs(d[i]) <= EPS;
      inv[i] = par[i] ? 0.0 : 1.0 / d[i];
    }
  }
  bool hit_bounds(const Aabb& b, double tmin, double tmax, double* entry_out) const {
    double lo[3] = {b.lo.x, b.lo.y, b.lo.z}, hi[3] = {b.hi.x, b.hi.y, b.hi.z};
    double entry = tmin, exit = tmax;
    for (int a = 0; a < 3; ++a) {
      if (par[a]) {
        if (o[a] < lo[a] || o[a] > hi[a]) return false;
        continue;
      }
      double t0 = (lo[a] - o[a]) * inv[a];
      double t1 = (hi[a] - o[a]) * inv[a];
      if (inv[a] < 0.0) std::swap(t0, t1);
      entry = fmaxr(entry, t0);
      exit = fminr(exit, t1);
      if (exit <= entry) return false;
    }
    *entry_out = entry;
    return true;
  }
};

struct Bvh {
  std::vector<Node> nodes;
  std::vector<uint64_t> indices;
  std::vector<Aabb> pb;
  std::vector<V3> pc;
  size_t leaf_size = 4;

  static int bucket_index(double offset) {  // bvh.rs:499-502
    double f = offset * double(BVH_BUCKETS);
    size_t b = (f != f || f <= 0.0) ? 0 : (f >= 1e18 ? SIZE_MAX : size_t(f));
    return int(std::min<size_t>(b, BVH_BUCKETS - 1));
  }
  static double axis_of(V3 p, int a) { return p[a]; }

  size_t midpoint_split(uint64_t* idx, size_t n, int axis) {  // bvh.rs:508-522
    std::stable_sort(idx, idx + n, [&](uint64_t l, uint64_t r) {
      double la = axis_of(pc[l], axis), ra = axis_of(pc[r], axis);
      if (la < ra) return true;
      if (la > ra) return false;
      return l < r;
    });
    return n / 2;
  }
  // bvh.rs:430-496; returns 0 for "no split"
  size_t split(uint64_t* idx, size_t n) {
    if (n <= leaf_size) return 0;
    Aabb cb = Aabb::from_points(pc[idx[0]], pc[idx[0]]);
    for (size_t i = 1; i < n; ++i) cb = cb.union_point(pc[idx[i]]);
    int axis = cb.largest_axis();
    double cmin = cb.lo[axis];
    double extent = cb.hi[axis] - cmin;
    if (extent <= EPS) return midpoint_split(idx, n, axis);
    size_t cnt[BVH_BUCKETS] = {0};
    Aabb bb[BVH_BUCKETS];
    bool has[BVH_BUCKETS] = {false};
    for (size_t i = 0; i < n; ++i) {
      uint64_t k = idx[i];
      int b = bucket_index((axis_of(pc[k], axis) - cmin) / extent);
      cnt[b]++;
      bb[b] = has[b] ? bb[b].unite(pb[k]) : pb[k];
      has[b] = true;
    }
    int best = 0;
    double best_cost = INF;
    for (int s = 0; s < BVH_BUCKETS - 1; ++s) {
      size_t lc = 0, rc = 0;
      Aabb lb, rb;
      bool lh = false, rh = false;
      for (int i = 0; i <= s; ++i) {  // merge_buckets :524-535
        lc += cnt[i];
        if (has[i]) { lb = lh ? lb.unite(bb[i]) : bb[i]; lh = true; }
      }
      for (int i = s + 1; i < BVH_BUCKETS; ++i) {
        rc += cnt[i];
        if (has[i]) { rb = rh ? rb.unite(bb[i]) : bb[i]; rh = true; }
      }
      if (lc == 0 || rc == 0) continue;
      double cost = lb.surface_area() * double(lc) + rb.surface_area() * double(rc);
      if (cost < best_cost) { best_cost = cost; best = s; }
    }
    if (!std::isfinite(best_cost)) return midpoint_split(idx, n, axis);
    size_t left = 0;
    for (size_t next = 0; next < n; ++next) {
      uint64_t k = idx[next];
      int b = bucket_index((axis_of(pc[k], axis) - cmin) / extent);
      if (b <= best) { std::swap(idx[left], idx[next]); ++left; }
    }
    if (left == 0 || left == n) return midpoint_split(idx, n, axis);
    return left;
  }
  size_t build_range(size_t first, size_t count) {  // bvh.rs:257-288
    Aabb b = pb[indices[first]];
    for (size_t i = 1; i < count; ++i) b = b.unite(pb[indices[first + i]]);
    size_t ni = nodes.size();
    Node nd; nd.bounds = b; nd.leaf = true; nd.a = first; nd.b = count;
    nodes.push_back(nd);
    size_t lc = split(indices.data() + first, count);
    if (lc != 0) {
      size_t l = build_range(first, lc);
      size_t r = build_range(first + lc, count - lc);
      nodes[ni].leaf = false; nodes[ni].a = l; nodes[ni].b = r;
    }
    return ni;
  }
  bool build(const std::vector<Prim>& prims, size_t leaf) {  // bvh.rs:233-247
    nodes.clear(); indices.clear();
    if (prims.empty()) return false;
    leaf_size = leaf == 0 ? 1 : leaf;
    pb.resize(prims.size()); pc.resize(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) { pb[i] = prim_bounds(prims[i]); pc[i] = pb[i].centroid(); }
    indices.resize(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) indices[i] = i;
    nodes.reserve(prims.size() * 2);
    build_range(0, prims.size());
    return true;
  }
  size_t depth(size_t n = 0) const {
    if (nodes[n].leaf) return 1;
    return 1 + std::max(depth(nodes[n].a), depth(nodes[n].b));
  }
};

// ---- scene (src/graphics/raytracing/scene.rs) ---------------------------------
struct Scene {
  std::vector<Mat> mats;
  std::vector<Prim> prims;
  Bvh bvh;
  bool built = false;

  // hit_ray_scene_indices (scene.rs:849-874) with the primitive id exposed
  bool hit_indices(const uint64_t* idx, size_t n, const Ray& r, double tmin, double tmax, Hit* out) const {
    double closest = tmax;
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
      uint64_t k = idx[i];
      Surf s;
      if (prim_intersect(prims[k], r, tmin, closest, &s)) {
        closest = s.t;
        out->s = s; out->prim = int64_t(k); out->mat = &mats[prims[k].material];
        any = true;
      }
    }
    return any;
  }
  bool hit_bruteforce(const Ray& r, double tmin, double tmax, Hit* out) const {  // scene.rs:1459-1467
    double closest = tmax;
    bool any = false;
    for (size_t k = 0; k < prims.size(); ++k) {
      Surf s;
      if (prim_intersect(prims[k], r, tmin, closest, &s)) {
        closest = s.t;
        out->s = s; out->prim = int64_t(k); out->mat = &mats[prims[k].material];
        any = true;
      }
    }
    return any;
  }
  // FlatBvh::hit_with_stats (bvh.rs:306-399); LIFO order equals the
  // reference's inline-array + overflow stack (:173-220).
  bool hit(const Ray& r, double tmin, double tmax, Hit* out, Stats* st = nullptr) const {
    if (!built || bvh.nodes.empty()) return hit_bruteforce(r, tmin, tmax, out);
    Stats local;
    Stats& s = st ? *st : local;
    Traversal tr(r);
    s.rays++;
    s.node_tests++;
    double root_entry;
    if (!tr.hit_bounds(bvh.nodes[0].bounds, tmin, tmax, &root_entry)) return false;
    s.node_hits++;
    struct E { uint64_t node; double entry; };
    std::vector<E> stack;
    stack.reserve(64);
    stack.push_back({0, root_entry});
    s.max_stack = std::max<uint64_t>(s.max_stack, stack.size());
    double closest = tmax;
    bool any = false;
    while (!stack.empty()) {
      E e = stack.back();
      stack.pop_back();
      if (e.entry >= closest) continue;
      const Node& nd = bvh.nodes[e.node];
      if (nd.leaf) {
        s.leaf_visits++;
        s.candidates += nd.b;
        Hit h;
        if (hit_indices(bvh.indices.data() + nd.a, nd.b, r, tmin, closest, &h)) {
          s.leaf_hits++;
          closest = h.s.t;
          *out = h;
          any = true;
        }
      } else {
        s.node_tests += 2;
        double le = 0.0, re = 0.0;
        bool lh = tr.hit_bounds(bvh.nodes[nd.a].bounds, tmin, closest, &le);
        bool rh = tr.hit_bounds(bvh.nodes[nd.b].bounds, tmin, closest, &re);
        s.node_hits += uint64_t(lh) + uint64_t(rh);
        if (lh && rh && re < le) {
          stack.push_back({nd.a, le});
          stack.push_back({nd.b, re});
        } else if (lh && rh) {
          stack.push_back({nd.b, re});
          stack.push_back({nd.a, le});
        } else if (lh) {
          stack.push_back({nd.a, le});
        } else if (rh) {
          stack.push_back({nd.b, re});
        }
        s.max_stack = std::max<uint64_t>(s.max_stack, stack.size());
      }
    }
    return any;
  }
};

// ---- light targets (scene.rs:363-609, object.rs:288-311,1048-1161) -----------
struct Target {
  Prim g;
  double weight = 1.0;
};
struct Lights {
  std::vector<Target> targets;
  double total = 0.0;
  void add(const Prim& g, double w) { total = total + w; targets.push_back({g, w}); }
};

static double area_pdf_value(V3 dir, double t, V3 normal, double area) {  // object.rs:1101-1113
  if (area <= EPS) return 0.0;
  double dist2 = length_squared(t * dir);
  double cosine = std::fabs(dot(normalized(dir), normal));
  if (cosine <= EPS) return 0.0;
  return dist2 / (cosine * area);
}
static double sphere_pdf_value(V3 center, double radius, V3 origin, V3 dir) {  // object.rs:1048-1073
  Ray r{origin, dir, 0.0};
  double root;
  if (!hit_sphere_in_interval(center, radius, r, SHADOW_ACNE_PDF_EPSILON, INF, &root)) return 0.0;
  V3 cd = center - origin;
  double d2 = length_squared(cd);
  double r2 = radius * radius;
  if (d2 <= r2) return 1.0 / (4.0 * PI);
  double cos_max = std::sqrt(1.0 - r2 / d2);
  double sa = 2.0 * PI * (1.0 - cos_max);
  return sa <= EPS ? 0.0 : 1.0 / sa;
}
static double geometry_pdf_value(const Prim& g, V3 origin, double time, V3 dir) {  // object.rs:288-298
  switch (g.kind) {
    case K_SPHERE: return sphere_pdf_value(g.a, g.radius, origin, dir);
    case K_MOVING_SPHERE: return sphere_pdf_value(moving_center(g, time), g.radius, origin, dir);
    case K_TRIANGLE: {  // :1088-1099
      TriHit h;
      if (!triangle_hit_ray(g, Ray{origin, dir, 0.0}, SHADOW_ACNE_PDF_EPSILON, INF, &h)) return 0.0;
      double area = 0.5 * std::sqrt(length_squared(cross(g.b - g.a, g.c - g.a)));
      return area_pdf_value(dir, h.t, normalized(cross(g.b - g.a, g.c - g.a)), area);
    }
    default: {  // :1075-1086
      TriHit h;
      if (!quad_hit_ray(g, Ray{origin, dir, 0.0}, SHADOW_ACNE_PDF_EPSILON, INF, &h)) return 0.0;
      double area = std::sqrt(length_squared(cross(g.b, g.c)));
      return area_pdf_value(dir, h.t, g.qn, area);
    }
  }
}
static V3 random_to_sphere(double radius, double d2, Rng& rng) {  // object.rs:1154-1161
  double r1 = rng.random_double(), r2 = rng.random_double();
  double z = 1.0 + r2 * (std::sqrt(1.0 - radius * radius / d2) - 1.0);
  double phi = 2.0 * PI * r1;
  double rz = std::sqrt(1.0 - z * z);
  return {std::cos(phi) * rz, std::sin(phi) * rz, z};
}
static V3 random_direction_to_sphere(V3 center, double radius, V3 origin, Rng& rng) {  // :1115-1130
  V3 dir = center - origin;
  double d2 = length_squared(dir);
  double r2 = radius * radius;
  if (d2 <= r2 || d2 <= EPS) return rng.unit_vector_spherical();
  V3 local = random_to_sphere(radius, d2, rng);
  Onb b;
  return onb_from_w(dir, &b) ? b.local(local) : local;
}
static V3 geometry_random_direction(const Prim& g, V3 origin, double time, Rng& rng) {  // object.rs:300-311
  switch (g.kind) {
    case K_SPHERE: return random_direction_to_sphere(g.a, g.radius, origin, rng);
    case K_MOVING_SPHERE: return random_direction_to_sphere(moving_center(g, time), g.radius, origin, rng);
    case K_TRIANGLE: {  // :1138-1152
      double a = rng.random_double();
      double b = rng.random_double();
      if (a + b > 1.0) { a = 1.0 - a; b = 1.0 - b; }
      V3 p = g.a + a * (g.b - g.a) + b * (g.c - g.a);
      return p - origin;
    }
    default: {  // :1132-1136
      double a = rng.random_double();
      V3 p0 = g.a + a * g.b;
      double b = rng.random_double();
      V3 p = p0 + b * g.c;
      return p - origin;
    }
  }
}
static double lights_pdf_value(const Lights& L, V3 origin, double time, V3 dir) {  // scene.rs:571-582
  if (L.targets.empty() || L.total <= 0.0) return 0.0;
  double sum = 0.0;
  for (const Target& t : L.targets) sum += t.weight / L.total * geometry_pdf_value(t.g, origin, time, dir);
  return sum;
}
static V3 lights_random_direction(const Lights& L, V3 origin, double time, Rng& rng) {  // scene.rs:584-600
  if (L.targets.empty() || L.total <= 0.0) return V3(1.0, 0.0, 0.0);
  double pick = rng.random_range(0.0, L.total);
  for (const Target& t : L.targets) {
    pick -= t.weight;
    if (pick <= 0.0) return geometry_random_direction(t.g, origin, time, rng);
  }
  return geometry_random_direction(L.targets.back().g, origin, time, rng);
}

// ---- environment (src/graphics/raytracing/environment.rs) ---------------------
// Vose's alias method (Vose, "A linear algorithm for generating random numbers
// with a given distribution", IEEE TSE 1991) over weights w[0, n): slot i keeps
// itself when the low 32 bits of its draw are below thr[i] and otherwise
// yields alias[i]. The GPU samples environment texels this way in the
// per-(pixel, sample, bounce) stream mode (one 64-bit draw, integer
// arithmetic); this restatement builds the same table in the same order so
// the oracle makes the same choices. The reference's CDF search stays for
// the reference stream.
static uint32_t alias_threshold(double p) {
  if (!(p > 0.0)) return 0u;
  if (p >= 1.0) return 0xffffffffu;
  double t = std::floor(p * 4294967296.0);
  return t >= 4294967295.0 ? 0xffffffffu : uint32_t(t);
}
static void build_alias(const std::vector<double>& w, double total, std::vector<uint32_t>* thr,
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

struct Env {
  bool constant = true;
  V3 color;
  bool rgb8 = false;             // reference 8-bit canvas form
  uint32_t w = 0, h = 0;
  std::vector<float> rgb32;      // linear RGB texels (FP32 HDR form)
  std::vector<uint8_t> rgb8px;   // 8-bit texels (reference form)
  std::vector<double> weights, cdf;
  std::vector<uint32_t> alias_thr, alias_idx;  // counter-stream texel choice (build_alias)
  double total = 0.0;

  V3 texel_linear(uint32_t x, uint32_t y) const {
    size_t i = size_t(y) * w + x;
    if (rgb8) {  // LinearRgb::from_rgb_srgb (colors.rs:78-84)
      auto dec = [](uint8_t c) { double v = double(c) / 255.0; return v * v; };
      return {dec(rgb8px[3 * i]), dec(rgb8px[3 * i + 1]), dec(rgb8px[3 * i + 2])};
    }
    return {double(rgb32[3 * i]), double(rgb32[3 * i + 1]), double(rgb32[3 * i + 2])};
  }
  void rebuild() {  // environment.rs:158-184
    weights.clear(); cdf.clear(); total = 0.0;
    if (w == 0 || h == 0) return;
    weights.reserve(size_t(w) * h); cdf.reserve(size_t(w) * h);
    for (uint32_t y = 0; y < h; ++y) {
      double theta = PI * ((double(y) + 0.5) / double(h));
      double st = fmaxr(std::sin(theta), 0.0);
      for (uint32_t x = 0; x < w; ++x) {
        V3 c = texel_linear(x, y);
        double wgt = fmaxr(0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z, 0.0) * st;
        total += wgt;
        weights.push_back(wgt);
        cdf.push_back(total);
      }
    }
    if (total > EPS) build_alias(weights, total, &alias_thr, &alias_idx);
  }
  // linear_axis (texture.rs:695-711); bilinear (texture.rs:681-692)
  static void axis_repeat(double value, uint32_t size, double* coord, uint32_t* lo, uint32_t* hi) {
    double c = rem_euclid(value, 1.0) * double(size);
    double fl = std::floor(c);
    uint32_t l = fl <= 0.0 ? 0u : (fl >= 4294967295.0 ? 0xffffffffu : uint32_t(fl));
    l = std::min(l, size - 1);
    *coord = c; *lo = l; *hi = (l + 1 == size) ? 0 : l + 1;
  }
  static void axis_clamp(double value, uint32_t size, double* coord, uint32_t* lo, uint32_t* hi) {
    double c = clampd(value, 0.0, 1.0) * double(size - 1);
    double fl = std::floor(c);
    uint32_t l = fl <= 0.0 ? 0u : uint32_t(fl);
    *coord = c; *lo = l; *hi = std::min(l + 1, size - 1);
  }
  V3 radiance(V3 dir) const {  // environment.rs:98-107
    if (constant) return color;
    V3 d = normalized(dir);  // direction_to_latlong_uv :193-198
    double phi = rem_euclid(std::atan2(d.z, d.x), TAU);
    double theta = std::acos(clampd(d.y, -1.0, 1.0));
    double u = phi / TAU, v = 1.0 - theta / PI;
    if (!std::isfinite(u) || !std::isfinite(v)) return V3();
    double cx, cy; uint32_t x0, x1, y0, y1;
    axis_repeat(u, w, &cx, &x0, &x1);
    axis_clamp(1.0 - v, h, &cy, &y0, &y1);
    double tx = cx - double(x0), ty = cy - double(y0);
    if (rgb8) {  // Rgb::lerp rounds to 8 bits per step (colors.rs:525-537)
      auto px = [&](uint32_t x, uint32_t y, int ch) { return rgb8px[3 * (size_t(y) * w + x) + ch]; };
      auto lerp8 = [](int a, int b, double t) {
        t = clampd(t, 0.0, 1.0);
        double r = std::round(double(a) + double(b - a) * t);
        return int(r < 0 ? 0 : (r > 255 ? 255 : r));
      };
      V3 out;
      double* o = &out.x;
      for (int ch = 0; ch < 3; ++ch) {
        int top = lerp8(px(x0, y0, ch), px(x1, y0, ch), tx);
        int bot = lerp8(px(x0, y1, ch), px(x1, y1, ch), tx);
        int c = lerp8(top, bot, ty);
        double vv = double(c) / 255.0;
        o[ch] = vv * vv;
      }
      return out;
    }
    V3 a = texel_linear(x0, y0), b = texel_linear(x1, y0), c = texel_linear(x0, y1), e = texel_linear(x1, y1);
    V3 top = a + (b - a) * tx;
    V3 bot = c + (e - c) * tx;
    return top + (bot - top) * ty;
  }
  double pixel_solid_angle(uint32_t y) const {  // :186-190
    double t0 = PI * (double(y) / double(h)), t1 = PI * (double(y + 1) / double(h));
    return (TAU / double(w)) * fmaxr(std::cos(t0) - std::cos(t1), 0.0);
  }
  double pdf_value(V3 dir) const {  // :136-156
    if (total <= EPS || weights.empty()) return 1.0 / (4.0 * PI);
    V3 d = normalized(dir);
    double theta = std::acos(clampd(d.y, -1.0, 1.0));
    double phi = rem_euclid(std::atan2(d.z, d.x), TAU);
    double fx = std::floor((phi / TAU) * double(w)), fy = std::floor((theta / PI) * double(h));
    size_t x = fx > 0.0 ? size_t(fx) : 0, y = fy > 0.0 ? size_t(fy) : 0;  // `as usize` saturates
    x = std::min<size_t>(x, w - 1); y = std::min<size_t>(y, h - 1);
    double prob = weights[y * w + x] / total;
    double sa = pixel_solid_angle(uint32_t(y));
    return sa <= EPS ? 0.0 : prob / sa;
  }
  // alias: the texel from one 64-bit draw through the alias table (counter
  // streams, the GPU's choice); otherwise the reference's CDF search.
  void sample_direction(Rng& rng, V3* dir, double* pdf, bool alias = false) const {  // :109-132
    if (total <= EPS || weights.empty()) {
      *dir = rng.unit_vector_spherical();
      *pdf = 1.0 / (4.0 * PI);
      return;
    }
    size_t index;
    if (alias) {
      const uint64_t x = rng.next_u64(), n = weights.size();
      const unsigned __int128 m = (unsigned __int128)x * n;
      const size_t slot = size_t(uint64_t(m >> 64));
      const uint32_t frac = uint32_t(uint64_t(m) >> 32);
      index = frac < alias_thr[slot] ? slot : alias_idx[slot];
    } else {
      double target = rng.random_double() * total;
      index = size_t(std::lower_bound(cdf.begin(), cdf.end(), target) - cdf.begin());  // partition_point(< target)
      index = std::min(index, weights.size() - 1);
    }
    size_t x = index % w, y = index / w;
    double u = (double(x) + rng.random_double()) / double(w);
    double t0 = PI * (double(y) / double(h)), t1 = PI * (double(y + 1) / double(h));
    double c0 = std::cos(t0), c1 = std::cos(t1);
    double ct = c0 + rng.random_double() * (c1 - c0);
    double st = std::sqrt(fmaxr(1.0 - ct * ct, 0.0));
    double phi = TAU * u;
    *dir = V3(st * std::cos(phi), ct, st * std::sin(phi));
    *pdf = pdf_value(*dir);
  }
};

// ---- camera (src/graphics/camera.rs, RayCamera half) --------------------------
struct Camera {
  gt_camera p;
  uint32_t width = 0, height = 0;
  V3 center, pixel00, du, dv, disk_u, disk_v;
  uint32_t spp = 1, grid = 1;

  bool init(const gt_camera& c, const char** err) {  // initialized :1068-1130, validate_view :1161-1201
    p = c;
    if (c.image_width == 0) { *err = "image width must be positive"; return false; }
    if (!(std::isfinite(c.aspect_ratio) && c.aspect_ratio > 0.0)) { *err = "aspect ratio must be positive and finite"; return false; }
    if (!(std::isfinite(c.vertical_fov) && c.vertical_fov > 0.0 && c.vertical_fov < 180.0)) { *err = "vertical field of view must be finite and in 0..180 degrees"; return false; }
    V3 from(c.lookfrom[0], c.lookfrom[1], c.lookfrom[2]), at(c.lookat[0], c.lookat[1], c.lookat[2]), up(c.view_up[0], c.view_up[1], c.view_up[2]);
    if (!(length_squared(from - at) > EPS)) { *err = "lookfrom and lookat must be distinct"; return false; }
    if (!(length_squared(up) > EPS)) { *err = "view-up vector must be nonzero"; return false; }
    V3 forward = normalized(at - from);  // CameraBasis::looking_at geometry.rs:22-36
    V3 right = normalized(cross(forward, up));
    if (length_squared(forward) <= EPS || length_squared(right) <= EPS) { *err = "view-up vector must not be parallel to the viewing direction"; return false; }
    V3 upv = cross(right, forward);
    if (!(std::isfinite(c.defocus_angle) && c.defocus_angle >= 0.0 && c.defocus_angle < 180.0)) { *err = "defocus angle must be finite and in 0..180 degrees"; return false; }
    if (!(std::isfinite(c.focus_distance) && c.focus_distance > 0.0)) { *err = "focus distance must be positive and finite"; return false; }
    if (!(std::isfinite(c.shutter_start) && std::isfinite(c.shutter_end) && c.shutter_start <= c.shutter_end)) { *err = "shutter interval must be finite and ordered"; return false; }
    int nbg = c.background_kind == GT_BACKGROUND_VERTICAL_GRADIENT ? 9 : 3;
    for (int i = 0; i < nbg; ++i) if (!std::isfinite(c.background[i])) { *err = "background color components must be finite"; return false; }
    if (c.adaptive_error_threshold != 0.0 && !(std::isfinite(c.adaptive_error_threshold) && c.adaptive_error_threshold > 0.0)) { *err = "adaptive sampling error threshold must be positive and finite"; return false; }
    width = c.image_width;
    double hh = double(width) / c.aspect_ratio;
    height = std::max<uint32_t>(hh >= 4294967295.0 ? 0xffffffffu : uint32_t(hh), 1u);
    double theta = degrees_to_radians(c.vertical_fov);
    double h = std::tan(theta * 0.5);
    double vh = 2.0 * h * c.focus_distance;
    double vw = vh * (double(width) / double(height));
    center = from;
    V3 w = -forward, u = right, v = upv;
    V3 vu = vw * u, vv = vh * -v;
    du = vu / double(width);
    dv = vv / double(height);
    V3 upper_left = center - c.focus_distance * w - vu / 2.0 - vv / 2.0;
    pixel00 = upper_left + 0.5 * (du + dv);
    double dr = c.focus_distance * std::tan(degrees_to_radians(c.defocus_angle * 0.5));
    disk_u = dr * u; disk_v = dr * v;
    uint32_t s = std::max<uint32_t>(c.samples_per_pixel, 1u);
    if (c.pixel_sample_mode == GT_PIXEL_RANDOM) { spp = s; grid = 1; }
    else if (c.pixel_sample_mode == GT_PIXEL_STRATIFIED) {  // :1547-1557, :1710-1713
      grid = std::max<uint32_t>(uint32_t(std::sqrt(double(s))), 1u); spp = grid * grid;
    } else { grid = std::max<uint32_t>(c.stratified_grid_width, 1u); spp = grid * grid; }
    return true;
  }
  Ray make_ray(uint32_t x, uint32_t y, double ox, double oy, Rng& rng) const {  // :1658-1692
    V3 ps = pixel00 + (double(x) + ox) * du + (double(y) + oy) * dv;
    V3 origin = center;
    if (!(p.defocus_angle <= 0.0)) {  // defocus_disk_sample :1715-1718
      V3 d = rng.in_unit_disk();
      origin = center + d.x * disk_u + d.y * disk_v;
    }
    double time = rng.random_range(p.shutter_start, p.shutter_end);
    return {origin, ps - origin, time};
  }
  Ray sample_ray(uint32_t x, uint32_t y, uint32_t s, Rng& rng) const {
    if (p.pixel_sample_mode == GT_PIXEL_RANDOM) {  // sample_square :1694-1696
      double ox = rng.random_double() - 0.5;
      double oy = rng.random_double() - 0.5;
      return make_ray(x, y, ox, oy, rng);
    }
    uint32_t sx = s % grid, sy = s / grid;  // sample_square_stratified :1698-1708
    double rec = 1.0 / double(grid);
    double ox = ((double(sx) + rng.random_double()) * rec) - 0.5;
    double oy = ((double(sy) + rng.random_double()) * rec) - 0.5;
    return make_ray(x, y, ox, oy, rng);
  }
  V3 background(V3 dir) const {  // RayBackground::radiance camera.rs:297-313
    if (p.background_kind == GT_BACKGROUND_VERTICAL_GRADIENT) {
      V3 nadir(p.background[0], p.background[1], p.background[2]);
      V3 horizon(p.background[3], p.background[4], p.background[5]);
      V3 zenith(p.background[6], p.background[7], p.background[8]);
      double t = 0.5 * (normalized(dir).y + 1.0);
      if (t < 0.5) { double bl = 2.0 * t; return nadir * (1.0 - bl) + horizon * bl; }
      double bl = 2.0 * (t - 0.5);
      return horizon * (1.0 - bl) + zenith * bl;
    }
    return {p.background[0], p.background[1], p.background[2]};
  }
};

struct Counters {
  uint64_t paths = 0, ext_rays = 0, shadow_rays = 0, accepted = 0;
};

struct Ctx {
  const Scene* scene;
  const Lights* lights;
  const Env* env;
  const Camera* cam;
  bool nee;
  double light_weight;
};

static double power_heuristic(double a, double b) {  // camera.rs:2486-2496
  if (!std::isfinite(a) || a <= EPS) return 0.0;
  if (!std::isfinite(b) || b <= EPS) return 1.0;
  double s = a * a, o = b * b;
  return s / (s + o);
}
static bool rr_survives(uint32_t bounce, int32_t min_depth, V3* att, Rng& rng) {  // camera.rs:2498-2523
  if (min_depth < 0 || bounce < uint32_t(min_depth)) return true;
  double m = max_component(*att);
  if (!std::isfinite(m) || m <= EPS) return false;
  double p = clampd(m, RR_MIN, RR_MAX);
  if (rng.random_double() >= p) return false;
  *att = color_div(*att, p);
  return true;
}

// estimate_direct_lighting camera.rs:2225-2262
static V3 estimate_direct(const Ctx& c, const Ray& in, const Hit& h, V3 satt, Rng& rng, Counters& k) {
  V3 dir = lights_random_direction(*c.lights, h.s.point, in.time, rng);
  double lpdf = lights_pdf_value(*c.lights, h.s.point, in.time, dir);
  if (!std::isfinite(lpdf) || lpdf <= EPS) return V3();
  Ray shadow{h.s.point, dir, in.time};
  double spdf = mat_scattering_pdf(*h.mat, in, h, dir);
  if (!std::isfinite(spdf) || spdf <= 0.0) return V3();
  k.shadow_rays++;
  Hit lh;
  if (!c.scene->hit(shadow, SHADOW_ACNE_EPSILON, INF, &lh)) return V3();
  V3 em = mat_emitted(*lh.mat, lh);
  if (!finite3(em)) return V3();
  return mul(satt, em) * (spdf / lpdf);
}
// estimate_environment_lighting camera.rs:2264-2297
static V3 estimate_env(const Ctx& c, const Ray& in, const Hit& h, V3 satt, Rng& rng, Counters& k) {
  V3 dir; double epdf;
  c.env->sample_direction(rng, &dir, &epdf, c.cam->p.rng_mode == GT_RNG_COUNTER);
  if (!std::isfinite(epdf) || epdf <= EPS) return V3();
  Ray shadow{h.s.point, dir, in.time};
  double spdf = mat_scattering_pdf(*h.mat, in, h, dir);
  if (!std::isfinite(spdf) || spdf <= 0.0) return V3();
  double mis = power_heuristic(epdf, spdf);
  k.shadow_rays++;
  Hit oh;
  if (c.scene->hit(shadow, SHADOW_ACNE_EPSILON, INF, &oh)) return V3();
  return mul(satt, c.env->radiance(dir)) * (mis * spdf / epdf);
}

// RayCamera::ray_color camera.rs:1730-1862
static V3 ray_color(const Ctx& c, Ray ray, uint32_t sample, PathRng& pr, Counters& k) {
  V3 att(1.0, 1.0, 1.0), color;
  bool allow_emitted = true;
  double miss_w = 1.0;
  const uint32_t depth = c.cam->p.max_depth;
  const bool has_lights = c.lights != nullptr;
  const bool has_env = c.env != nullptr;
  for (uint32_t b = 0; b < depth; ++b) {
    pr.begin(sample, 1 + b);
    Rng& rng = pr.rng;
    k.ext_rays++;
    Hit rec;
    if (!c.scene->hit(ray, SHADOW_ACNE_EPSILON, INF, &rec)) {
      V3 miss = (has_env ? c.env->radiance(ray.d) : c.cam->background(ray.d)) * miss_w;
      return color + mul(att, miss);
    }
    if (allow_emitted) color = color + mul(att, mat_emitted(*rec.mat, rec));
    Scatter sc = mat_scatter(*rec.mat, ray, rec, rng);
    if (sc.kind == 0) return color;
    if (sc.kind == 1) {
      att = mul(att, sc.attenuation);
      ray = sc.ray;
      allow_emitted = true;
      miss_w = 1.0;
      if (!rr_survives(b, c.cam->p.russian_roulette_min_depth, &att, rng)) return color;
      continue;
    }
    if (c.nee) {
      if (has_lights) color = color + mul(att, estimate_direct(c, ray, rec, sc.attenuation, rng, k));
      if (has_env) color = color + mul(att, estimate_env(c, ray, rec, sc.attenuation, rng, k));
    }
    // SamplingStrategy::continuation_sample camera.rs:165-195
    V3 dir; double cpdf;
    bool suppress, weight_env;
    bool below = false;
    if (!c.nee && has_lights) {
      double wgt = c.light_weight;  // MixturePdf::weighted clamps (sampling.rs:355-365)
      if (rng.random_double() < wgt) dir = lights_random_direction(*c.lights, rec.s.point, ray.time, rng);
      else dir = sc.pdf.generate(rng, &below);
      cpdf = wgt * lights_pdf_value(*c.lights, rec.s.point, ray.time, dir) + (1.0 - wgt) * sc.pdf.value(dir);
      suppress = false; weight_env = false;
    } else {
      dir = sc.pdf.generate(rng, &below);
      cpdf = sc.pdf.value(dir);
      suppress = c.nee && has_lights;
      weight_env = c.nee && has_env;
    }
    // opt-in: end the path instead of the reference's retro-reflection fallback
    if (below && (c.cam->p.flags & GT_FLAG_GGX_REJECT_BELOW_HORIZON)) return color;
    if (!std::isfinite(cpdf) || cpdf <= EPS) return color;
    Ray scattered{rec.s.point, dir, ray.time};
    double spdf = mat_scattering_pdf(*rec.mat, ray, rec, dir);
    if (!std::isfinite(spdf) || spdf <= 0.0) return color;
    double next_w = weight_env ? power_heuristic(cpdf, c.env->pdf_value(dir)) : 1.0;
    att = mul(att, sc.attenuation * (spdf / cpdf));
    ray = scattered;
    allow_emitted = !suppress;
    miss_w = next_w;
    if (!rr_survives(b, c.cam->p.russian_roulette_min_depth, &att, rng)) return color;
  }
  return color;
}

// adaptive_error (camera.rs:2972-2982): the largest channel's standard error
// of the mean; infinite below two accepted samples.
static double adaptive_error(V3 m2, uint32_t n) {
  if (n <= 1) return INF;
  V3 variance = color_div(m2, double(n - 1));
  return std::sqrt(fmaxr(max_component(color_div(variance, double(n))), 0.0));
}

// The Welford state of render_world_pixel_adaptive (camera.rs:2944-2966).
struct Welford {
  V3 mean, m2;
  uint32_t n = 0;
  // Adds a finite sample; true when the pixel stops.
  bool add(V3 v, uint32_t min_samples, double threshold) {
    ++n;
    const double count = double(n);
    V3 delta = v - mean;
    mean = mean + color_div(delta, count);
    V3 delta2 = v - mean;
    m2 = m2 + mul(delta, delta2);
    return n >= min_samples && adaptive_error(m2, n) < threshold;
  }
};

static bool adaptive_on(const gt_camera& p) {
  return p.adaptive_error_threshold != 0.0 && p.pixel_sample_mode == GT_PIXEL_RANDOM;
}

// render_world_linear_pixel_with_context camera.rs:2660-2723
static V3 render_pixel(const Ctx& c, uint32_t x, uint32_t y, Counters& k) {
  const Camera& cam = *c.cam;
  PathRng pr(int(cam.p.rng_mode), pixel_seed(cam.p.rng_seed, x, y));
  if (adaptive_on(cam.p)) {  // render_world_pixel_adaptive :2937-2970
    const uint32_t mn = std::max<uint32_t>(cam.p.adaptive_min_samples, 1);
    const uint32_t mx = std::max(cam.p.adaptive_max_samples, mn);
    Welford w;
    for (uint32_t s = 0; s < mx; ++s) {
      pr.begin(s, 0);
      Ray r = cam.sample_ray(x, y, s, pr.rng);
      k.paths++;
      V3 v = ray_color(c, r, s, pr, k);
      if (!finite3(v)) continue;
      if (w.add(v, mn, cam.p.adaptive_error_threshold)) break;
    }
    k.accepted += w.n;
    return w.mean;
  }
  V3 sum;
  uint32_t accepted = 0;
  for (uint32_t s = 0; s < cam.spp; ++s) {
    pr.begin(s, 0);
    Ray r = cam.sample_ray(x, y, s, pr.rng);
    k.paths++;
    V3 v = ray_color(c, r, s, pr, k);
    if (finite3(v)) { sum = sum + v; ++accepted; }  // add_finite_sample :2984-2991
  }
  k.accepted += accepted;
  return accepted == 0 ? V3() : color_div(sum, double(accepted));  // :2993-2999
}
// render_denoising_aov_linear_pixel camera.rs:3064-3116 (albedo and normal in one pass)
static void render_aov_pixel(const Ctx& c, uint32_t x, uint32_t y, V3* albedo, V3* normal) {
  const Camera& cam = *c.cam;
  PathRng pa(int(cam.p.rng_mode), pixel_seed(cam.p.rng_seed, x, y));
  PathRng pn(int(cam.p.rng_mode), pixel_seed(cam.p.rng_seed, x, y));
  V3 sa, sn;
  for (uint32_t s = 0; s < cam.spp; ++s) {
    pa.begin(s, 0);
    pn.begin(s, 0);
    Ray ra = cam.sample_ray(x, y, s, pa.rng);
    Ray rn = cam.sample_ray(x, y, s, pn.rng);
    Hit h;
    if (c.scene->hit(ra, SHADOW_ACNE_EPSILON, INF, &h)) sa = sa + mat_denoise_albedo(*h.mat);
    if (c.scene->hit(rn, SHADOW_ACNE_EPSILON, INF, &h)) sn = sn + 0.5 * (h.s.shading + V3(1.0, 1.0, 1.0));
  }
  *albedo = color_div(sa, double(cam.spp));  // camera.rs:3095
  *normal = color_div(sn, double(cam.spp));
}

}  // namespace cr

// =============================================================================
// C ABI for the Python test harness (ctypes). Names are cr_*.
// =============================================================================
using namespace cr;

static thread_local const char* g_err = "";

extern "C" {

const char* cr_last_error(void) { return g_err; }

void* cr_scene_new(void) { return new Scene(); }
void cr_scene_free(void* s) { delete static_cast<Scene*>(s); }

int cr_scene_add_material(void* sp, int kind, double r, double g, double b, double param, uint32_t* id) {
  Scene* s = static_cast<Scene*>(sp);
  Mat m;
  m.kind = kind;
  m.color = V3(r, g, b);
  m.param = param;
  if (kind != M_DIELECTRIC && !finite3(m.color)) { g_err = "material colour must be finite"; return GT_EINVAL; }
  if (kind == M_GGX) {
    if (!std::isfinite(param)) { g_err = "GGX roughness must be finite"; return GT_EINVAL; }
    m.param = clampd(param, 0.0, 1.0);  // material.rs:1383-1396
  } else if (kind == M_METAL) {
    if (!std::isfinite(param)) { g_err = "metal fuzz must be finite"; return GT_EINVAL; }
    m.param = clampd(param, 0.0, 1.0);  // material.rs:1962-1975
  } else if (kind == M_DIELECTRIC) {
    m.color = V3(1.0, 1.0, 1.0);
  } else if (kind == M_HENYEY_GREENSTEIN) {  // validate_henyey_greenstein_g (material.rs:1328-1333)
    if (!(std::isfinite(param) && std::fabs(param) < 1.0)) {
      g_err = "Henyey-Greenstein anisotropy must be finite and in (-1, 1)";
      return GT_EINVAL;
    }
  } else if (kind != M_LAMBERTIAN && kind != M_DIFFUSE_LIGHT && kind != M_GGX && kind != M_METAL &&
             kind != M_ISOTROPIC) {
    g_err = "unknown material kind";
    return GT_EINVAL;
  }
  *id = uint32_t(s->mats.size());
  s->mats.push_back(m);
  return GT_OK;
}

int cr_scene_add_prims(void* sp, int kind, const double* d, size_t n, const uint32_t* mids) {
  Scene* s = static_cast<Scene*>(sp);
  for (size_t i = 0; i < n; ++i)
    if (mids[i] >= s->mats.size()) { g_err = "ray scene material id out of bounds"; return GT_EINVAL; }
  for (size_t i = 0; i < n; ++i) {
    Prim p;
    if (kind == K_SPHERE) p = make_sphere(V3(d[4 * i], d[4 * i + 1], d[4 * i + 2]), d[4 * i + 3]);
    else if (kind == K_MOVING_SPHERE) {
      const double* q = d + 7 * i;
      p = make_moving_sphere(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), q[6]);
    } else if (kind == K_MESH_TRIANGLE) {  // positions (9) then vertex normals (9)
      const double* q = d + 18 * i;
      p = make_triangle(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]));
      p.smooth = true;
      p.na = V3(q[9], q[10], q[11]);
      p.nb = V3(q[12], q[13], q[14]);
      p.nc = V3(q[15], q[16], q[17]);
    } else {
      const double* q = d + 9 * i;
      if (kind == K_TRIANGLE) p = make_triangle(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]));
      else p = make_quad(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]));
    }
    p.material = mids[i];
    s->prims.push_back(p);
  }
  s->built = false;
  return GT_OK;
}

int cr_scene_build(void* sp, uint32_t leaf_size) {
  Scene* s = static_cast<Scene*>(sp);
  s->built = s->bvh.build(s->prims, leaf_size);
  return GT_OK;
}
uint64_t cr_scene_node_count(void* sp) { Scene* s = static_cast<Scene*>(sp); return s->built ? s->bvh.nodes.size() : 0; }
uint64_t cr_scene_depth(void* sp) { Scene* s = static_cast<Scene*>(sp); return s->built ? s->bvh.depth() : 0; }
// Node dump: bounds (6 per node), leaf flag, a, b; indices permutation.
int cr_scene_bvh_dump(void* sp, double* bounds, uint8_t* leaf, uint64_t* a, uint64_t* b, uint64_t* indices) {
  Scene* s = static_cast<Scene*>(sp);
  if (!s->built) return GT_ESTATE;
  for (size_t i = 0; i < s->bvh.nodes.size(); ++i) {
    const Node& nd = s->bvh.nodes[i];
    double* o = bounds + 6 * i;
    o[0] = nd.bounds.lo.x; o[1] = nd.bounds.lo.y; o[2] = nd.bounds.lo.z;
    o[3] = nd.bounds.hi.x; o[4] = nd.bounds.hi.y; o[5] = nd.bounds.hi.z;
    leaf[i] = nd.leaf; a[i] = nd.a; b[i] = nd.b;
  }
  for (size_t i = 0; i < s->bvh.indices.size(); ++i) indices[i] = s->bvh.indices[i];
  return GT_OK;
}

// rays: 8 doubles each (o.xyz, time, d.xyz, tmax)
int cr_intersect(void* sp, const double* rays, size_t n, double tmin, int use_bvh, int64_t* prim, double* tuv,
                 double* normal, uint8_t* front, uint64_t* stats7) {
  Scene* s = static_cast<Scene*>(sp);
  Stats st;
  for (size_t i = 0; i < n; ++i) {
    const double* q = rays + 8 * i;
    Ray r{V3(q[0], q[1], q[2]), V3(q[4], q[5], q[6]), q[3]};
    Hit h;
    bool ok = use_bvh ? s->hit(r, tmin, q[7], &h, &st) : s->hit_bruteforce(r, tmin, q[7], &h);
    prim[i] = ok ? h.prim : -1;
    if (tuv) { tuv[3 * i] = ok ? h.s.t : 0; tuv[3 * i + 1] = ok ? h.s.u : 0; tuv[3 * i + 2] = ok ? h.s.v : 0; }
    if (normal) { normal[3 * i] = h.s.normal.x; normal[3 * i + 1] = h.s.normal.y; normal[3 * i + 2] = h.s.normal.z; }
    if (front) front[i] = ok && h.s.front;
  }
  if (stats7) {
    uint64_t v[7] = {st.rays, st.node_tests, st.node_hits, st.leaf_visits, st.candidates, st.leaf_hits, st.max_stack};
    std::memcpy(stats7, v, sizeof v);
  }
  return GT_OK;
}

void* cr_lights_new(void) { return new Lights(); }
void cr_lights_free(void* l) { delete static_cast<Lights*>(l); }
int cr_lights_add(void* lp, int kind, const double* q, double weight) {
  Lights* L = static_cast<Lights*>(lp);
  if (!(std::isfinite(weight) && weight > 0.0)) { g_err = "sampling target weight must be finite and positive"; return GT_EINVAL; }
  if (!std::isfinite(L->total + weight)) { g_err = "sampling target total weight must remain finite"; return GT_EINVAL; }
  Prim p;
  if (kind == K_SPHERE) p = make_sphere(V3(q[0], q[1], q[2]), q[3]);
  else if (kind == K_MOVING_SPHERE) p = make_moving_sphere(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), q[6]);
  else if (kind == K_TRIANGLE) p = make_triangle(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]));
  else p = make_quad(V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), V3(q[6], q[7], q[8]));
  L->add(p, weight);
  return GT_OK;
}
size_t cr_lights_len(void* lp) { return static_cast<Lights*>(lp)->targets.size(); }
void* cr_scene_emissive_targets(void* sp) {  // scene.rs:1403-1423
  Scene* s = static_cast<Scene*>(sp);
  Lights* L = new Lights();
  for (const Prim& p : s->prims)
    if (s->mats[p.material].kind == M_DIFFUSE_LIGHT) L->add(p, 1.0);
  return L;
}
double cr_lights_pdf_value(void* lp, const double* origin, double time, const double* dir) {
  return lights_pdf_value(*static_cast<Lights*>(lp), V3(origin[0], origin[1], origin[2]), time, V3(dir[0], dir[1], dir[2]));
}
void cr_lights_random_direction(void* lp, const double* origin, double time, uint64_t seed, size_t n, double* out) {
  Rng rng(seed);
  for (size_t i = 0; i < n; ++i) {
    V3 d = lights_random_direction(*static_cast<Lights*>(lp), V3(origin[0], origin[1], origin[2]), time, rng);
    out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
  }
}

void* cr_env_constant(double r, double g, double b) {
  Env* e = new Env();
  e->constant = true;
  e->color = V3(r, g, b);
  return e;
}
void* cr_env_rgb32f(const float* rgb, uint32_t w, uint32_t h) {
  Env* e = new Env();
  e->constant = false; e->rgb8 = false; e->w = w; e->h = h;
  e->rgb32.assign(rgb, rgb + size_t(w) * h * 3);
  e->rebuild();
  return e;
}
void* cr_env_rgb8(const uint8_t* rgb, uint32_t w, uint32_t h) {  // EnvironmentLight::from_canvas
  Env* e = new Env();
  e->constant = false; e->rgb8 = true; e->w = w; e->h = h;
  e->rgb8px.assign(rgb, rgb + size_t(w) * h * 3);
  e->rebuild();
  return e;
}
void cr_env_free(void* e) { delete static_cast<Env*>(e); }
void cr_env_sample(void* ep, uint64_t seed, size_t n, double* dirs, double* pdfs) {
  Rng rng(seed);
  for (size_t i = 0; i < n; ++i) {
    V3 d; double p;
    static_cast<Env*>(ep)->sample_direction(rng, &d, &p);
    dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z; pdfs[i] = p;
  }
}
void cr_env_sample_alias(void* ep, uint64_t seed, size_t n, double* dirs, double* pdfs) {
  Rng rng(seed);
  for (size_t i = 0; i < n; ++i) {
    V3 d; double p;
    static_cast<Env*>(ep)->sample_direction(rng, &d, &p, true);
    dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z; pdfs[i] = p;
  }
}
// The alias table (n entries each; 0 when the map has no weight).
size_t cr_env_alias_table(void* ep, uint32_t* thr, uint32_t* alias) {
  const Env* e = static_cast<Env*>(ep);
  if (thr) std::copy(e->alias_thr.begin(), e->alias_thr.end(), thr);
  if (alias) std::copy(e->alias_idx.begin(), e->alias_idx.end(), alias);
  return e->alias_thr.size();
}
void cr_env_eval(void* ep, const double* dirs, size_t n, double* radiance, double* pdf) {
  Env* e = static_cast<Env*>(ep);
  for (size_t i = 0; i < n; ++i) {
    V3 d(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
    V3 r = e->radiance(d);
    if (radiance) { radiance[3 * i] = r.x; radiance[3 * i + 1] = r.y; radiance[3 * i + 2] = r.z; }
    if (pdf) pdf[i] = e->pdf_value(d);
  }
}

// Unit hooks for known-answer tests.
uint64_t cr_pixel_seed(uint64_t seed, uint32_t x, uint32_t y) { return pixel_seed(seed, x, y); }
uint64_t cr_stream_seed(uint64_t key, uint32_t sample, uint32_t dim) { return stream_seed(key, sample, dim); }
void cr_rng_u64(uint64_t seed, size_t n, uint64_t* out) { Rng r(seed); for (size_t i = 0; i < n; ++i) out[i] = r.next_u64(); }
void cr_rng_double(uint64_t seed, size_t n, double* out) { Rng r(seed); for (size_t i = 0; i < n; ++i) out[i] = r.random_double(); }
void cr_rng_vectors(uint64_t seed, int which, size_t n, double* out) {
  Rng r(seed);
  for (size_t i = 0; i < n; ++i) {
    V3 v = which == 0 ? r.unit_vector_spherical() : (which == 1 ? r.cosine_direction() : r.in_unit_disk());
    out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
  }
}
double cr_cosine_pdf_value(const double* n, const double* d) {
  MatPdf p; p.kind = 0;
  if (!onb_from_w(V3(n[0], n[1], n[2]), &p.basis)) return -1.0;
  return p.value(V3(d[0], d[1], d[2]));
}
int cr_ggx_pdf(const double* n, const double* o, double roughness, const double* d, uint64_t seed, size_t ngen,
               double* value, double* gen) {
  MatPdf p;
  if (!ggx_pdf_new(V3(n[0], n[1], n[2]), V3(o[0], o[1], o[2]), roughness, &p)) return GT_EINVAL;
  if (value) *value = p.value(V3(d[0], d[1], d[2]));
  Rng rng(seed);
  for (size_t i = 0; i < ngen; ++i) {
    V3 g = p.generate(rng);
    gen[3 * i] = g.x; gen[3 * i + 1] = g.y; gen[3 * i + 2] = g.z;
  }
  return GT_OK;
}
double cr_ggx_D(double ndh, double r) { return ggx_D(ndh, r); }
double cr_smith_g1(double ndd, double r) { return smith_g1(ndd, r); }
double cr_schlick(double c, double f0) { return schlick(c, f0); }
double cr_ggx_weight(const double* n, const double* o, const double* i, double r, double f0) {
  return ggx_weight(V3(n[0], n[1], n[2]), V3(o[0], o[1], o[2]), V3(i[0], i[1], i[2]), r, f0);
}
double cr_power_heuristic(double a, double b) { return power_heuristic(a, b); }
// RayTraversal::hit_bounds (bvh.rs:119-146): returns 1 and the entry t, or 0.
int cr_hit_bounds(const double* box6, const double* ray7, double tmin, double tmax, double* entry) {
  Ray r{V3(ray7[0], ray7[1], ray7[2]), V3(ray7[4], ray7[5], ray7[6]), ray7[3]};
  Aabb b{V3(box6[0], box6[1], box6[2]), V3(box6[3], box6[4], box6[5])};
  Traversal tr(r);
  return tr.hit_bounds(b, tmin, tmax, entry) ? 1 : 0;
}
// Russian roulette on one attenuation (camera.rs:2498-2523). Returns survival, updates att.
int cr_russian_roulette(uint32_t bounce, int32_t min_depth, double* att, uint64_t seed) {
  Rng rng(seed);
  V3 a(att[0], att[1], att[2]);
  bool ok = rr_survives(bounce, min_depth, &a, rng);
  att[0] = a.x; att[1] = a.y; att[2] = a.z;
  return ok;
}
int cr_camera_frame(const gt_camera* cam, double* out15, uint32_t* height) {
  Camera c;
  const char* err = "";
  if (!c.init(*cam, &err)) { g_err = err; return GT_EINVAL; }
  V3 v[5] = {c.center, c.pixel00, c.du, c.dv, c.disk_u};
  for (int i = 0; i < 5; ++i) { out15[3 * i] = v[i].x; out15[3 * i + 1] = v[i].y; out15[3 * i + 2] = v[i].z; }
  *height = c.height;
  return GT_OK;
}
int cr_camera_ray(const gt_camera* cam, uint32_t x, uint32_t y, uint32_t sample, double* out7) {
  Camera c;
  const char* err = "";
  if (!c.init(*cam, &err)) { g_err = err; return GT_EINVAL; }
  PathRng pr(int(cam->rng_mode), pixel_seed(cam->rng_seed, x, y));
  pr.begin(sample, 0);
  Ray r = c.sample_ray(x, y, sample, pr.rng);
  double v[7] = {r.o.x, r.o.y, r.o.z, r.time, r.d.x, r.d.y, r.d.z};
  std::memcpy(out7, v, sizeof v);
  return GT_OK;
}

// One material interaction at a hit, for known-answer tests of scatter and
// scattering_pdf. hit: point(3), outward normal(3), t; returns scatter kind.
int cr_material_scatter(void* sp, uint32_t mid, const double* ray7, const double* hit7, uint64_t seed,
                        double* out_dir, double* out_att, double* out_pdf_of_dir) {
  Scene* s = static_cast<Scene*>(sp);
  Ray in{V3(ray7[0], ray7[1], ray7[2]), V3(ray7[4], ray7[5], ray7[6]), ray7[3]};
  Hit h;
  h.s = with_uv(in, V3(hit7[0], hit7[1], hit7[2]), V3(hit7[3], hit7[4], hit7[5]), hit7[6], 0, 0);
  h.mat = &s->mats[mid];
  Rng rng(seed);
  Scatter sc = mat_scatter(*h.mat, in, h, rng);
  if (sc.kind == 1) {
    out_dir[0] = sc.ray.d.x; out_dir[1] = sc.ray.d.y; out_dir[2] = sc.ray.d.z;
  } else if (sc.kind == 2) {
    V3 d = sc.pdf.generate(rng);
    out_dir[0] = d.x; out_dir[1] = d.y; out_dir[2] = d.z;
    *out_pdf_of_dir = sc.pdf.value(d);
  }
  if (sc.kind) { out_att[0] = sc.attenuation.x; out_att[1] = sc.attenuation.y; out_att[2] = sc.attenuation.z; }
  return sc.kind;
}
double cr_material_scattering_pdf(void* sp, uint32_t mid, const double* ray7, const double* hit7, const double* dir) {
  Scene* s = static_cast<Scene*>(sp);
  Ray in{V3(ray7[0], ray7[1], ray7[2]), V3(ray7[4], ray7[5], ray7[6]), ray7[3]};
  Hit h;
  h.s = with_uv(in, V3(hit7[0], hit7[1], hit7[2]), V3(hit7[3], hit7[4], hit7[5]), hit7[6], 0, 0);
  h.mat = &s->mats[mid];
  return mat_scattering_pdf(*h.mat, in, h, V3(dir[0], dir[1], dir[2]));
}

// Render a crop [x0, x0+w) x [y0, y0+h) of the camera's image with the
// reference's 16x16 tiles on `threads` workers (render_values_tiled,
// camera.rs:3134-3170). out_rgb: w*h*3 doubles. counters: paths, extension
// rays, shadow rays, accepted samples.
static int render_common(void* sp, const gt_camera* camp, void* lp, void* ep, int threads, uint32_t x0, uint32_t y0,
                         uint32_t w, uint32_t h, double* out, double* albedo, double* normal, uint64_t* counters) {
  Scene* s = static_cast<Scene*>(sp);
  Camera cam;
  const char* err = "";
  if (!cam.init(*camp, &err)) { g_err = err; return GT_EINVAL; }
  if (w == 0) w = cam.width - x0;
  if (h == 0) h = cam.height - y0;
  if (x0 + w > cam.width || y0 + h > cam.height) { g_err = "crop outside the image"; return GT_EINVAL; }
  if (!s->built && !s->prims.empty()) s->built = s->bvh.build(s->prims, 4);  // cached_bvh (scene.rs:1567-1573)
  Lights* L = static_cast<Lights*>(lp);
  Env* E = static_cast<Env*>(ep);
  Ctx c{s, L, E, &cam,
        camp->direct_lighting_mode == GT_LIGHTING_NEE, clampd(camp->light_pdf_weight, 0.0, 1.0)};
  const uint32_t tile = 16;
  uint32_t tx = (w + tile - 1) / tile, ty = (h + tile - 1) / tile;
  std::atomic<uint32_t> next{0};
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  std::vector<Counters> ks(threads);
  auto worker = [&](int id) {
    for (;;) {
      uint32_t t = next.fetch_add(1);
      if (t >= tx * ty) break;
      uint32_t bx = (t % tx) * tile, by = (t / tx) * tile;
      for (uint32_t yy = by; yy < std::min(by + tile, h); ++yy)
        for (uint32_t xx = bx; xx < std::min(bx + tile, w); ++xx) {
          size_t i = size_t(yy) * w + xx;
          if (out) {
            V3 v = render_pixel(c, x0 + xx, y0 + yy, ks[id]);
            out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
          }
          if (albedo) {
            V3 a, n;
            render_aov_pixel(c, x0 + xx, y0 + yy, &a, &n);
            albedo[3 * i] = a.x; albedo[3 * i + 1] = a.y; albedo[3 * i + 2] = a.z;
            normal[3 * i] = n.x; normal[3 * i + 1] = n.y; normal[3 * i + 2] = n.z;
          }
        }
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < threads; ++i) pool.emplace_back(worker, i);
  worker(0);
  for (auto& t : pool) t.join();
  if (counters) {
    Counters tot;
    for (auto& k : ks) { tot.paths += k.paths; tot.ext_rays += k.ext_rays; tot.shadow_rays += k.shadow_rays; tot.accepted += k.accepted; }
    counters[0] = tot.paths; counters[1] = tot.ext_rays; counters[2] = tot.shadow_rays; counters[3] = tot.accepted;
  }
  return GT_OK;
}
int cr_render(void* sp, const gt_camera* cam, void* lp, void* ep, int threads, uint32_t x0, uint32_t y0, uint32_t w,
              uint32_t h, double* out_rgb, uint64_t* counters) {
  return render_common(sp, cam, lp, ep, threads, x0, y0, w, h, out_rgb, nullptr, nullptr, counters);
}
int cr_render_aovs(void* sp, const gt_camera* cam, void* lp, void* ep, int threads, uint32_t x0, uint32_t y0,
                   uint32_t w, uint32_t h, double* beauty, double* albedo, double* normal, uint64_t* counters) {
  return render_common(sp, cam, lp, ep, threads, x0, y0, w, h, beauty, albedo, normal, counters);
}

// ToneMap::map_color (display.rs:264-277) and Rgb::from_linear_color (colors.rs:401-409)
void cr_tone_map(const double* rgb, size_t n, double exposure, int op, double gamma, uint8_t* out) {
  for (size_t i = 0; i < 3 * n; ++i) {
    double v = rgb[i] * exposure;
    v = std::isfinite(v) ? fmaxr(v, 0.0) : 0.0;  // map_component :173-189
    if (op == 1) v = v / (1.0 + v);
    else if (op == 2) v = clampd(v * (2.51 * v + 0.03) / (v * (2.43 * v + 0.59) + 0.14), 0.0, 1.0);
    double g;
    if (std::fabs(gamma - 2.0) <= EPS) g = std::isnan(v) ? 0.0 : (v > 0.0 ? std::sqrt(v) : 0.0);
    else g = v <= 0.0 ? 0.0 : std::pow(v, 1.0 / gamma);
    out[i] = as_u8(256.0 * clampd(g, 0.0, 0.999));
  }
}

// render_world_pixel_adaptive's loop (camera.rs:2944-2966) over a given
// sequence of n samples (rgb triples; a non-finite triple is skipped), with
// max_samples = n: the mean, the number of samples taken and accepted.
void cr_adaptive_run(const double* samples, uint32_t n, uint32_t min_samples, double threshold, double* mean,
                     uint32_t* taken, uint32_t* accepted) {
  Welford w;
  uint32_t t = 0;
  for (uint32_t s = 0; s < n; ++s) {
    ++t;
    V3 v(samples[3 * s], samples[3 * s + 1], samples[3 * s + 2]);
    if (!finite3(v)) continue;
    if (w.add(v, std::max<uint32_t>(min_samples, 1), threshold)) break;
  }
  mean[0] = w.mean.x; mean[1] = w.mean.y; mean[2] = w.mean.z;
  *taken = t;
  *accepted = w.n;
}

double cr_adaptive_error(const double* m2, uint32_t n) { return adaptive_error(V3(m2[0], m2[1], m2[2]), n); }

// Rgb::from_raw_linear_color (colors.rs:412-422): the AOV canvases
// (camera.rs:4212-4243). NaN -> 0, clamp to [0, 1], times 255.999.
void cr_encode_raw(const double* rgb, size_t n, uint8_t* out) {
  for (size_t i = 0; i < 3 * n; ++i) out[i] = as_u8(255.999 * clampd(std::isnan(rgb[i]) ? 0.0 : rgb[i], 0.0, 1.0));
}

// encode_rgbe (display.rs:444-468), 4 bytes per pixel in storage order
// (write_radiance_hdr, display.rs:397-409).
void cr_encode_rgbe(const double* rgb, size_t n, uint8_t* out) {
  for (size_t p = 0; p < n; ++p) {
    double c[3];
    for (int k = 0; k < 3; ++k) c[k] = std::isfinite(rgb[3 * p + k]) ? fmaxr(rgb[3 * p + k], 0.0) : 0.0;
    const double m = fmaxr(fmaxr(c[0], c[1]), c[2]);
    uint8_t* o = out + 4 * p;
    if (m <= 1.0e-32) {
      o[0] = o[1] = o[2] = o[3] = 0;
      continue;
    }
    const int e = int(std::floor(std::log2(m))) + 1;
    const double scale = 256.0 * powi_f64(2.0, -e);
    for (int k = 0; k < 3; ++k) o[k] = as_u8(clampd(c[k] * scale, 0.0, 255.0));
    o[3] = uint8_t(std::min(std::max(e + 128, 0), 255));
  }
}

// write_pfm's body (display.rs:371-386): rows bottom first, f32 little
// endian, a non-finite value written as 0 (`component as f32` otherwise).
void cr_encode_pfm(const double* rgb, uint32_t w, uint32_t h, float* out) {
  for (uint32_t y = 0; y < h; ++y)
    for (size_t i = 0; i < size_t(w) * 3; ++i) {
      const double v = rgb[size_t(h - 1 - y) * w * 3 + i];
      out[size_t(y) * w * 3 + i] = std::isfinite(v) ? float(v) : 0.0f;
    }
}

}  // extern "C"
