// device.h — device-side math, RNG, BSDFs, light and environment sampling
// for the wavefront path tracer (gfx950, FP32).
//
// Arithmetic follows the reference (cited per function) in FP32. Constants
// that the reference compares against f64::EPSILON keep that value.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace gtd {

constexpr float EPS = 2.220446049250313e-16f;  // f64::EPSILON as a float
constexpr float PI = 3.14159265358979323846f;
constexpr float TAU = 6.28318530717958647692f;
constexpr float INF = __builtin_huge_valf();
constexpr float SHADOW_ACNE = 0.001f;          // raytracing.rs:146, object.rs:1163

struct f3 {
  float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len2(f3 a) { return dot(a, a); }
__device__ __forceinline__ f3 normalized(f3 a) {  // vector.rs:151-158
  float l = sqrtf(len2(a));
  if (l < EPS) return {0.f, 0.f, 0.f};
  return a / l;
}
__device__ __forceinline__ f3 reflected(f3 d, f3 n) { return d - 2.0f * dot(d, n) * n; }  // vector.rs:162
__device__ __forceinline__ f3 refracted(f3 d, f3 n, float eta) {  // vector.rs:168-173
  float c = fminf(dot(-d, n), 1.0f);
  f3 perp = eta * (d + c * n);
  f3 par = -sqrtf(fabsf(1.0f - len2(perp))) * n;
  return perp + par;
}
__device__ __forceinline__ bool finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ float pow5(float x) { float x2 = x * x; float x4 = x2 * x2; return x * x4; }
__device__ __forceinline__ f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }

// ---- RNG: SplitMix64 (random.rs:23-29) -------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t pixel_seed(uint64_t seed, uint32_t x, uint32_t y) {  // camera.rs:1720-1727
  return mix64(seed ^ (uint64_t(x) * 0x9E3779B97F4A7C15ULL) ^ (uint64_t(y) * 0xD1B54A32D192ED03ULL));
}
// Per-(pixel, sample, dimension) stream key; dimension 0 = camera, 1+b = bounce b.
__device__ __forceinline__ uint64_t stream_seed(uint64_t key, uint32_t sample, uint32_t dim) {
  return mix64(key ^ (uint64_t(sample + 1) * 0xD6E8FEB86659FD93ULL) ^ (uint64_t(dim + 1) * 0xA0761D6478BD642FULL));
}
struct Rng {
  uint64_t s;
  // random_double (random.rs:56-59) at float precision: the top 24 bits.
  __device__ __forceinline__ float next() {
    s += 0x9e3779b97f4a7c15ULL;
    return float(uint32_t(mix64(s) >> 40)) * 5.9604644775390625e-8f;
  }
  __device__ __forceinline__ uint64_t next_u64() {  // next_u64 (random.rs:23-29)
    s += 0x9e3779b97f4a7c15ULL;
    return mix64(s);
  }
  __device__ __forceinline__ float range(float lo, float hi) { return lo + (hi - lo) * next(); }
  __device__ __forceinline__ f3 unit_vector_spherical() {  // random.rs:107-115
    float r1 = next(), r2 = next();
    float z = 1.0f - 2.0f * r2;
    float radius = sqrtf(1.0f - z * z);
    float phi = TAU * r1;
    float s, c;
    sincosf(phi, &s, &c);
    return {c * radius, s * radius, z};
  }
  __device__ __forceinline__ f3 cosine_direction() {  // random.rs:119-126
    float r1 = next(), r2 = next();
    float phi = TAU * r1;
    float radius = sqrtf(r2);
    float s, c;
    sincosf(phi, &s, &c);
    return {c * radius, s * radius, sqrtf(1.0f - r2)};
  }
  __device__ __forceinline__ f3 in_unit_disk() {  // random.rs:141-152
    for (;;) {
      float a = range(-1.0f, 1.0f);
      float b = range(-1.0f, 1.0f);
      if (a * a + b * b < 1.0f) return {a, b, 0.0f};
    }
  }
};

// ---- orthonormal basis and PDFs (geometry.rs:122-137, sampling.rs) ----------
struct Onb {
  f3 u, v, w;
  __device__ __forceinline__ f3 local(f3 l) const { return l.x * u + l.y * v + l.z * w; }
};
__device__ __forceinline__ bool onb_from_w(f3 axis, Onb* o) {
  f3 w = normalized(axis);
  if (len2(w) <= EPS) return false;
  f3 a = fabsf(w.x) > 0.9f ? mk(0.f, 1.f, 0.f) : mk(1.f, 0.f, 0.f);
  f3 v = normalized(cross(w, a));
  o->u = cross(w, v);
  o->v = v;
  o->w = w;
  return true;
}
__device__ __forceinline__ float ggx_alpha(float r) {  // sampling.rs:310-312
  float c = clampf(r, 0.f, 1.f);
  return fmaxf(c * c, 1.0e-4f);
}
// sampling.rs:229-241. The reference's denominator cos^2 (a^2 - 1) + 1 is
// evaluated as sin^2 + a^2 cos^2 with sin^2 = |n x h|^2 from the vectors:
// the same value for unit n and h, but free of the cancellation that makes
// the FP32 form useless near the peak of a smooth lobe (at roughness 0.02,
// a^2 = 1.6e-7 is a couple of ulps of 1).
__device__ __forceinline__ float ggx_D(float ndh, float sin2, float roughness) {
  if (!isfinite(ndh) || ndh <= 0.f) return 0.f;
  float alpha = ggx_alpha(roughness);
  float a2 = alpha * alpha;
  float den = sin2 + a2 * (ndh * ndh);
  if (den <= EPS) return 0.f;
  return a2 / (PI * den * den);
}
__device__ __forceinline__ float smith_g1(float ndd, float roughness) {  // sampling.rs:314-322
  if (!isfinite(ndd) || ndd <= 0.f) return 0.f;
  float alpha = ggx_alpha(roughness);
  float a2 = alpha * alpha;
  return 2.0f * ndd / (ndd + sqrtf(a2 + (1.0f - a2) * ndd * ndd));
}
__device__ __forceinline__ float schlick(float c, float f0) {  // sampling.rs:255-261
  f0 = clampf(f0, 0.f, 1.f);
  return f0 + (1.0f - f0) * pow5(1.0f - clampf(c, 0.f, 1.f));
}
// HenyeyGreensteinPdf::phase_value (sampling.rs:126-136)
__device__ __forceinline__ float hg_phase(float cosine, float g) {
  if (!isfinite(cosine) || !isfinite(g) || fabsf(g) >= 1.0f) return 0.f;
  const float den = 1.0f + g * g - 2.0f * g * clampf(cosine, -1.f, 1.f);
  if (den <= EPS) return 0.f;
  return (1.0f - g * g) / (4.0f * PI * (den * sqrtf(den)));
}
// The material pdfs (MaterialPdf, pdf.rs:46-104): CosinePdf (sampling.rs:48-84),
// GgxReflectionPdf (:174-308), SpherePdf (:31-41), HenyeyGreensteinPdf (:91-166).
enum { PDF_COSINE = 0, PDF_GGX = 1, PDF_SPHERE = 2, PDF_HG = 3 };
struct MatPdf {
  Onb basis;
  f3 outgoing;
  float roughness, alpha;
  float g;
  uint32_t kind;
  __device__ float value(f3 d) const {
    if (kind == PDF_SPHERE) return 1.0f / (4.0f * PI);
    if (kind == PDF_HG) {
      const float l = sqrtf(len2(d));
      if (l <= EPS) return 0.f;
      return hg_phase(dot(d, basis.w) / l, g);
    }
    if (kind == PDF_COSINE) {
      float l = sqrtf(len2(d));
      if (l <= EPS) return 0.f;
      float c = dot(d, basis.w) / l;
      return c <= 0.f ? 0.f : c / PI;
    }
    f3 in = normalized(d);
    if (len2(in) <= EPS || dot(in, basis.w) <= 0.f) return 0.f;
    f3 half = normalized(in + outgoing);
    float odh = dot(outgoing, half);
    if (len2(half) <= EPS || odh <= EPS) return 0.f;
    float ndh = dot(half, basis.w);
    return ggx_D(ndh, len2(cross(basis.w, half)), roughness) * fmaxf(ndh, 0.f) / (4.0f * fabsf(odh));
  }
  // *below (when given) reports a GGX reflection below the surface, where
  // the reference falls back to the outgoing direction.
  __device__ f3 generate(Rng& rng, bool* below = nullptr) const {
    if (kind == PDF_SPHERE) return rng.unit_vector_spherical();
    if (kind == PDF_HG) {  // sampling.rs:149-165
      const float r1 = rng.next(), r2 = rng.next();
      float ct;
      if (fabsf(g) < 1e-3f) {
        ct = 1.0f - 2.0f * r1;
      } else {
        const float term = (1.0f - g * g) / (1.0f - g + 2.0f * g * r1);
        ct = clampf((1.0f + g * g - term * term) / (2.0f * g), -1.f, 1.f);
      }
      const float st = sqrtf(fmaxf(1.0f - ct * ct, 0.f));
      float sn, cs;
      sincosf(2.0f * PI * r2, &sn, &cs);
      return basis.local(mk(cs * st, sn * st, ct));
    }
    if (kind == PDF_COSINE) return basis.local(rng.cosine_direction());
    float r1 = rng.next(), r2 = rng.next();
    float phi = 2.0f * PI * r1;
    float a2 = alpha * alpha;
    float cr2 = fminf(r2, 0.99999994f);
    float tan2 = a2 * cr2 / (1.0f - cr2);
    float ct = 1.0f / sqrtf(1.0f + tan2);
    // sin from tan: the reference's sqrt(1 - cos^2) cancels in FP32 when
    // tan^2 is a few ulps of 1, which would quantise a smooth lobe's angles
    float st = sqrtf(tan2) * ct;
    float s, c;
    sincosf(phi, &s, &c);
    f3 half = basis.local(mk(c * st, s * st, ct));
    f3 refl = reflected(-outgoing, normalized(half));
    const bool up = dot(refl, basis.w) > 0.f;
    if (below) *below = !up;
    return up ? refl : outgoing;
  }
};
// ggx_reflection_scattering_weight (material.rs:1686-1722)
__device__ __forceinline__ float ggx_weight(f3 n, f3 o, f3 i, float roughness, float f0) {
  n = normalized(n); o = normalized(o); i = normalized(i);
  float ndi = dot(n, i), ndo = dot(n, o);
  if (ndi <= 0.f || ndo <= 0.f) return 0.f;
  f3 half = normalized(i + o);
  if (len2(half) <= EPS) return 0.f;
  float ndh = dot(n, half), odh = dot(o, half);
  if (ndh <= 0.f || odh <= 0.f) return 0.f;
  float D = ggx_D(ndh, len2(cross(n, half)), roughness);
  float G = smith_g1(ndi, roughness) * smith_g1(ndo, roughness);
  f0 = clampf(f0, 0.f, 1.f);
  if (f0 <= EPS) return 0.f;
  float F = schlick(odh, f0) / f0;
  return D * G * F / (4.0f * ndo);
}

// ---- primitive intersection (geometry.rs, object.rs) -------------------------
struct PrimHit {
  float t, u, v;
};
__device__ __forceinline__ uint32_t meta_of(const float4& p0) { return __float_as_uint(p0.w); }

// hit_sphere_in_interval (object.rs:986-1013). The discriminant uses the
// perpendicular-distance form a*(r^2 - |oc - (h/a) d|^2), algebraically equal
// to h^2 - a*c and free of its cancellation in FP32.
//
// from_surface: the ray starts on this sphere (its origin is a hit on it).
// In exact arithmetic one root is then 0 and the other 2h/a, so the sphere
// is met again only by a ray heading inward, at the far root. FP32 cannot
// place the origin on a large sphere within t_min (on the 1000-radius
// ground sphere a coordinate's ulp is 6e-5 and the near root's error grows
// as 1/cos at grazing angles), so the near root is skipped by this rule, not
// compared with t_min: the decision the reference's f64 arithmetic makes.
__device__ __forceinline__ bool sphere_hit(f3 c, float r, f3 o, f3 d, float tmin, float tmax, float* root,
                                           bool from_surface = false) {
  f3 oc = c - o;
  float a = len2(d);
  if (a <= EPS) return false;
  float h = dot(d, oc);
  if (from_surface && h <= 0.f) return false;
  f3 perp = oc - (h / a) * d;
  float disc = a * (r * r - len2(perp));
  if (disc < 0.f) return false;
  float sq = sqrtf(disc);
  float first = (h - sq) / a;
  if (!from_surface && tmin < first && first < tmax) { *root = first; return true; }
  float second = (h + sq) / a;
  if (tmin < second && second < tmax) { *root = second; return true; }
  return false;
}
// TriangleGeometry::hit_ray (geometry.rs:239-265)
__device__ __forceinline__ bool tri_hit(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float tmin, float tmax, PrimHit* h) {
  f3 pvec = cross(d, e2);
  float det = dot(e1, pvec);
  if (fabsf(det) < 1e-10f) return false;
  float inv = 1.0f / det;
  f3 tvec = o - v0;
  float u = dot(tvec, pvec) * inv;
  if (!(u >= 0.f && u <= 1.f)) return false;
  f3 qvec = cross(tvec, e1);
  float v = dot(d, qvec) * inv;
  if (v < 0.f || u + v > 1.f) return false;
  float t = dot(e2, qvec) * inv;
  if (!(tmin < t && t < tmax)) return false;
  h->t = t; h->u = u; h->v = v;
  return true;
}
// QuadGeometry::hit_ray (geometry.rs:394-415); w = normal / |u x v|.
__device__ __forceinline__ bool quad_hit(f3 corner, f3 u, f3 v, f3 n, float inv_len, float nls, float plane_d,
                                         f3 o, f3 d, float tmin, float tmax, PrimHit* h) {
  if (nls <= EPS) return false;
  float denom = dot(n, d);
  if (fabsf(denom) < 1e-8f) return false;
  float t = (plane_d - dot(n, o)) / denom;
  if (!(tmin < t && t < tmax)) return false;
  f3 pp = (o + t * d) - corner;
  f3 w = n * inv_len;
  float a = dot(w, cross(pp, v));
  float b = dot(w, cross(u, pp));
  if (!(a >= 0.f && a <= 1.f) || !(b >= 0.f && b <= 1.f)) return false;
  h->t = t; h->u = a; h->v = b;
  return true;
}
// One primitive record against a ray interval; returns the root only.
// from_surface: the ray starts on this primitive. A ray leaving a plane
// cannot meet it again (the reference's t is 0 up to rounding, below
// t_min); a sphere is met again only at its far root (sphere_hit).
__device__ __forceinline__ bool prim_test(const GPrim* __restrict__ prims, uint32_t k, f3 o, f3 d, float time,
                                          float tmin, float tmax, PrimHit* h, uint32_t* meta_out,
                                          bool from_surface) {
  const float4* rec = reinterpret_cast<const float4*>(prims + k);
  float4 p0 = rec[0];
  uint32_t meta = meta_of(p0);
  *meta_out = meta;
  uint32_t kind = (meta >> GT_KIND_SHIFT) & GT_KIND_MASK;
  if (from_surface && (kind == GK_TRIANGLE || kind == GK_QUAD)) return false;
  float4 p1 = rec[1];
  if (kind == GK_TRIANGLE) {
    float4 p2 = rec[2];
    return tri_hit(mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), o, d, tmin, tmax, h);
  }
  if (kind == GK_QUAD) {
    float4 p2 = rec[2];
    float4 p3 = rec[3];
    return quad_hit(mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), mk(p3.x, p3.y, p3.z), p1.w,
                    p2.w, p3.w, o, d, tmin, tmax, h);
  }
  f3 c = mk(p0.x, p0.y, p0.z);
  float r = p1.x;
  if (kind == GK_MOVING_SPHERE) {
    c = c + time * (mk(p1.x, p1.y, p1.z) - c);
    r = p1.w;
  }
  float root;
  if (!sphere_hit(c, r, o, d, tmin, tmax, &root, from_surface)) return false;
  h->t = root; h->u = 0.f; h->v = 0.f;
  return true;
}
// ---- FP64 refinement of the chosen hit ---------------------------------------
// The FP32 traversal decides which primitive a ray hits. Its distance, and
// the surface coordinates, are then recomputed once per ray in FP64 with the
// reference's own expressions (contraction off, as the reference evaluates
// them): near the ray origin FP32 cannot resolve the distance to a surface
// (a ray 5 cm above the 1000-radius ground sphere loses 1e-3 of its t), FP64
// from the same FP32 inputs reproduces the reference's t to FP32 rounding.
// If FP64 rejects the primitive (an edge or silhouette case where the two
// precisions decide differently) the FP32 values stand.
struct d3 {
  double x, y, z;
};
__device__ __forceinline__ d3 dd(float4 v) { return {double(v.x), double(v.y), double(v.z)}; }
__device__ __forceinline__ d3 dd(f3 v) { return {double(v.x), double(v.y), double(v.z)}; }
__device__ __forceinline__ double ddot(d3 a, d3 b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ d3 dsub(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 dcross(d3 a, d3 b) {
#pragma clang fp contract(off)
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ bool refine_hit(const GPrim* __restrict__ prims, uint32_t k, f3 of, f3 df, float time,
                                        float tmin, PrimHit* h) {
#pragma clang fp contract(off)
  const float4* rec = reinterpret_cast<const float4*>(prims + k);
  const float4 p0 = rec[0], p1 = rec[1];
  const uint32_t kind = (meta_of(p0) >> GT_KIND_SHIFT) & GT_KIND_MASK;
  const d3 o = dd(of), d = dd(df);
  const double lo = double(tmin);
  if (kind == GK_TRIANGLE) {  // geometry.rs:239-265
    const float4 p2 = rec[2];
    const d3 v0 = dd(p0), e1 = dd(p1), e2 = dd(p2);
    const d3 pv = dcross(d, e2);
    const double det = ddot(e1, pv);
    if (fabs(det) < 1e-10) return false;
    const double inv = 1.0 / det;
    const d3 tv = dsub(o, v0);
    const double u = ddot(tv, pv) * inv;
    if (!(u >= 0.0 && u <= 1.0)) return false;
    const d3 qv = dcross(tv, e1);
    const double v = ddot(d, qv) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    const double t = ddot(e2, qv) * inv;
    if (!(lo < t)) return false;
    h->t = float(t); h->u = float(u); h->v = float(v);
    return true;
  }
  if (kind == GK_QUAD) {  // QuadGeometry::new and hit_ray (geometry.rs:293-312,394-415)
    const float4 p2 = rec[2];
    const d3 c = dd(p0), u = dd(p1), v = dd(p2);
    const d3 n = dcross(u, v);
    const double nls = ddot(n, n);
    if (nls <= 2.220446049250313e-16) return false;
    const double len = sqrt(nls);
    const d3 nn = {n.x / len, n.y / len, n.z / len};
    const double pd = ddot(nn, c);
    const double denom = ddot(nn, d);
    if (fabs(denom) < 1e-8) return false;
    const double t = (pd - ddot(nn, o)) / denom;
    if (!(lo < t)) return false;
    const d3 w = {n.x / nls, n.y / nls, n.z / nls};
    const d3 pl = dsub({o.x + t * d.x, o.y + t * d.y, o.z + t * d.z}, c);
    const double a = ddot(w, dcross(pl, v)), b = ddot(w, dcross(u, pl));
    if (!(a >= 0.0 && a <= 1.0) || !(b >= 0.0 && b <= 1.0)) return false;
    h->t = float(t); h->u = float(a); h->v = float(b);
    return true;
  }
  // hit_sphere_in_interval (object.rs:986-1013): of the roots above t_min,
  // the one the FP32 test chose (the nearer to its distance)
  d3 c = dd(p0);
  double r = double(p1.x);
  if (kind == GK_MOVING_SPHERE) {  // geometry.rs:507-509
    const d3 c1 = dd(p1);
    const double tm = double(time);
    c = {c.x + tm * (c1.x - c.x), c.y + tm * (c1.y - c.y), c.z + tm * (c1.z - c.z)};
    r = double(p1.w);
  }
  const d3 oc = dsub(c, o);
  const double a = ddot(d, d);
  if (a <= 2.220446049250313e-16) return false;
  const double hh = ddot(d, oc);
  const double cc = ddot(oc, oc) - r * r;
  const double disc = hh * hh - a * cc;
  if (disc < 0.0) return false;
  const double sq = sqrt(disc);
  const double first = (hh - sq) / a, second = (hh + sq) / a;
  const double t32 = double(h->t);
  double t;
  if (lo < first && (!(lo < second) || fabs(first - t32) <= fabs(second - t32))) t = first;
  else if (lo < second) t = second;
  else return false;
  h->t = float(t);
  return true;
}

// Outward normal at a hit (object.rs:163-234); for spheres (p - c)/r, or the
// normalised offset when |r| <= EPS (geometry.rs:476-483).
__device__ __forceinline__ f3 prim_outward_normal(const GPrim* __restrict__ prims, uint32_t k, f3 p, float time,
                                                  uint32_t* material, uint32_t* kind_out) {
  const float4* rec = reinterpret_cast<const float4*>(prims + k);
  float4 p0 = rec[0];
  uint32_t meta = meta_of(p0);
  *material = meta & GT_MATERIAL_MASK;
  uint32_t kind = (meta >> GT_KIND_SHIFT) & GT_KIND_MASK;
  *kind_out = kind;
  if (kind == GK_TRIANGLE || kind == GK_QUAD) {
    float4 p3 = rec[3];
    return mk(p3.x, p3.y, p3.z);
  }
  float4 p1 = rec[1];
  f3 c = mk(p0.x, p0.y, p0.z);
  float r = p1.x;
  if (kind == GK_MOVING_SPHERE) {
    c = c + time * (mk(p1.x, p1.y, p1.z) - c);
    r = p1.w;
  }
  f3 out = p - c;
  if (fabsf(r) <= EPS) return normalized(out);
  return out / r;
}

// Shading normal of a mesh triangle (MeshTriangle::shading_normal,
// mesh.rs:105-119, oriented by set_shading_normal, object.rs:470-479): the
// vertex normals interpolated at barycentrics (u, v) of v1 and v2, normalized
// and flipped with the face; n itself (the oriented geometric normal) for a
// flat triangle, any other primitive, or a degenerate interpolation.
__device__ __forceinline__ f3 shading_normal(const GPrim* __restrict__ prims, const float4* __restrict__ vn,
                                             uint32_t k, float u, float v, bool front, f3 n) {
  const float4 p1 = reinterpret_cast<const float4*>(prims + k)[1];
  const uint32_t kind = (meta_of(reinterpret_cast<const float4*>(prims + k)[0]) >> GT_KIND_SHIFT) & GT_KIND_MASK;
  const uint32_t smooth = __float_as_uint(p1.w);
  if (kind != GK_TRIANGLE || smooth == 0u) return n;
  const float4* q = vn + 3 * size_t(smooth - 1u);
  const float4 a = q[0], b = q[1], c = q[2];
  const float w = 1.0f - u - v;
  const f3 s = mk(w * a.x + u * b.x + v * c.x, w * a.y + u * b.y + v * c.y, w * a.z + u * b.z + v * c.z);
  if (!(len2(s) > EPS)) return n;
  const f3 sn = normalized(s);
  return front ? sn : -sn;
}
// Barycentric coordinates (of v1, v2) of a point p in a triangle record, for
// a hit whose FP64 refinement was rejected (an edge case).
__device__ __forceinline__ void tri_barycentric(const GPrim* __restrict__ prims, uint32_t k, f3 p, float* u,
                                                float* v) {
  const float4* rec = reinterpret_cast<const float4*>(prims + k);
  const float4 p0 = rec[0], p1 = rec[1], p2 = rec[2];
  const f3 e1 = mk(p1.x, p1.y, p1.z), e2 = mk(p2.x, p2.y, p2.z), q = p - mk(p0.x, p0.y, p0.z);
  const float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2), d20 = dot(q, e1), d21 = dot(q, e2);
  const float den = d00 * d11 - d01 * d01;
  if (!(fabsf(den) > 0.f)) { *u = 0.f; *v = 0.f; return; }
  *u = (d11 * d20 - d01 * d21) / den;
  *v = (d00 * d21 - d01 * d20) / den;
}

// ---- scene and traversal ---------------------------------------------------
struct DScene {
  const GNode2* __restrict__ nodes2;
  const GNode4* __restrict__ nodes4;
  const GPrim* __restrict__ prims;
  const float4* __restrict__ vnormals;  // mesh triangles' vertex normals, 3 per triangle
  const uint32_t* __restrict__ prim_ids;
  const GMat* __restrict__ mats;
  float root_lo[3], root_hi[3];
  uint32_t root_ref;
  uint32_t width;  // 2 or 4
};

struct Stats {
  uint32_t node_tests, node_hits, leaf_visits, candidates, leaf_hits, max_stack;
};

// Per-lane traversal stack: the first `cap` entries in LDS (lane-interleaved),
// the rest in a per-lane slice of global memory.
struct Stack {
  uint2* lds;
  uint32_t stride;
  uint2* glob;
  uint32_t gstride;
  int cap;
  int n;
  __device__ __forceinline__ void push(uint32_t ref, float t) {
    uint2 e = make_uint2(ref, __float_as_uint(t));
    if (n < cap) lds[n * stride] = e;
    else glob[size_t(n - cap) * gstride] = e;
    ++n;
  }
  __device__ __forceinline__ uint2 pop() {
    --n;
    return n < cap ? lds[n * stride] : glob[size_t(n - cap) * gstride];
  }
};

// Precomputed slab data (RayTraversal::new, bvh.rs:98-117). A parallel axis
// (|d| <= EPS) gets inv = +inf, which with IEEE maxNum/minNum reproduces the
// reference's origin-in-slab rule: (lo-o)*inf is -inf, +inf or NaN (ignored).
struct RayPre {
  f3 o;
  float inv[3];
  bool neg[3];
};
__device__ __forceinline__ RayPre ray_pre(f3 o, f3 d) {
  RayPre r;
  r.o = o;
  float dd[3] = {d.x, d.y, d.z};
  for (int a = 0; a < 3; ++a) {
    bool par = fabsf(dd[a]) <= EPS;
    r.inv[a] = par ? INF : 1.0f / dd[a];
    r.neg[a] = r.inv[a] < 0.f;
  }
  return r;
}
// Slab test, conservative in FP32 (the far bound is widened by a few ulp);
// returns the entry distance or -1 for a miss.
__device__ __forceinline__ bool slab(const RayPre& r, float lox, float hix, float loy, float hiy, float loz, float hiz,
                                     float tmin, float tmax, float* entry) {
  float nx = r.neg[0] ? hix : lox, fx = r.neg[0] ? lox : hix;
  float ny = r.neg[1] ? hiy : loy, fy = r.neg[1] ? loy : hiy;
  float nz = r.neg[2] ? hiz : loz, fz = r.neg[2] ? loz : hiz;
  float t0x = (nx - r.o.x) * r.inv[0], t1x = (fx - r.o.x) * r.inv[0];
  float t0y = (ny - r.o.y) * r.inv[1], t1y = (fy - r.o.y) * r.inv[1];
  float t0z = (nz - r.o.z) * r.inv[2], t1z = (fz - r.o.z) * r.inv[2];
  float en = fmaxf(fmaxf(tmin, t0x), fmaxf(t0y, t0z));
  float ex = fminf(fminf(tmax, t1x), fminf(t1y, t1z));
  *entry = en;
  return ex * 1.0000004f >= en;
}

// Closest hit (ANY = false) or any hit (ANY = true) over the BVH2 layout,
// walking the reference's order: both children tested against the current
// closest, nearer first (left on ties), entries pruned when entry >= closest
// (bvh.rs:306-399). Leaves test primitives in permutation order with the
// interval shrinking on each accepted hit (scene.rs:849-874).
template <bool ANY, bool STATS>
__device__ __forceinline__ bool trace2(const DScene& sc, f3 o, f3 d, float time, float tmin, float tmax,
                                       uint32_t self, Stack& st, float* t_out, uint32_t* prim_out, PrimHit* hit_out,
                                       Stats* stats, uint32_t* meta_out = nullptr) {
  RayPre rp = ray_pre(o, d);
  float closest = tmax;
  bool any = false;
  uint32_t best = 0, best_meta = 0;
  PrimHit bh{0.f, 0.f, 0.f};
  if (STATS) { stats->node_tests += 1; }
  float root_entry;
  if (!slab(rp, sc.root_lo[0], sc.root_hi[0], sc.root_lo[1], sc.root_hi[1], sc.root_lo[2], sc.root_hi[2], tmin,
            tmax, &root_entry))
    return false;
  if (STATS) { stats->node_hits += 1; stats->max_stack = max(stats->max_stack, 1u); }
  uint32_t cur = sc.root_ref;
  st.n = 0;
  if (ANY && !STATS) {
    // Any-hit rays walk while-while with speculation (Aila and Laine,
    // "Understanding the efficiency of ray traversal on GPUs", HPG 2009): a
    // ray that reaches a leaf parks it and walks on through inner nodes until
    // every ray of the wave has a leaf parked, then the wave tests leaves
    // together. An any-hit query never narrows its interval before it
    // returns, so it visits the same nodes in either loop and its answer is
    // the same. (For closest hits the walk would cull with a stale closest hit
    // and could reach leaves that the reference's pruning skips.)
    constexpr uint32_t kNone = 0xffffffffu;  // no node: has the leaf bit, so it ends the inner loop
    uint32_t leaf = kNone;
    if (cur & GT_LEAF_BIT) { leaf = cur; cur = kNone; }
    auto next = [&]() -> uint32_t {  // pop, skipping entries that cannot beat the closest hit
      while (st.n > 0) {
        const uint2 e = st.pop();
        if (__uint_as_float(e.y) < closest) return e.x;
      }
      return kNone;
    };
    while (cur != kNone || leaf != kNone) {
      while (!(cur & GT_LEAF_BIT)) {
        const float4* nd = reinterpret_cast<const float4*>(sc.nodes2 + cur);
        const float4 a = nd[0], b = nd[1], c = nd[2];
        const uint4 refs = *reinterpret_cast<const uint4*>(nd + 3);
        float le, re;
        const bool lh = slab(rp, a.x, a.y, a.z, a.w, c.x, c.y, tmin, closest, &le);
        const bool rh = slab(rp, b.x, b.y, b.z, b.w, c.z, c.w, tmin, closest, &re);
        if (lh && rh) {
          const bool right_first = re < le;
          st.push(right_first ? refs.x : refs.y, right_first ? le : re);
          cur = right_first ? refs.y : refs.x;
        } else if (lh) {
          cur = refs.x;
        } else if (rh) {
          cur = refs.y;
        } else {
          cur = next();
        }
        if (cur != kNone && (cur & GT_LEAF_BIT) && leaf == kNone) {
          leaf = cur;
          cur = next();
        }
        if (!__any(leaf == kNone)) break;  // every ray still walking has a leaf parked
      }
      while (leaf != kNone) {
        for (uint32_t k = leaf & ~GT_LEAF_BIT;; ++k) {
          PrimHit h;
          uint32_t meta;
          if (prim_test(sc.prims, k, o, d, time, tmin, closest, &h, &meta, k == self)) {
            closest = h.t;
            best = k;
            best_meta = meta;
            bh = h;
            any = true;
            if (ANY) return true;
          }
          if (meta & GT_LAST_BIT) break;
        }
        if (cur != kNone && (cur & GT_LEAF_BIT)) {
          leaf = cur;
          cur = next();
        } else {
          leaf = kNone;
        }
      }
    }
    if (any) {
      *t_out = closest;
      *prim_out = best;
      if (hit_out) *hit_out = bh;
      if (meta_out) *meta_out = best_meta;
    }
    return any;
  }
  for (;;) {
    if (cur & GT_LEAF_BIT) {
      uint32_t k = cur & ~GT_LEAF_BIT;
      bool leaf_hit = false;
      if (STATS) stats->leaf_visits += 1;
      for (;;) {
        PrimHit h;
        uint32_t meta;
        if (STATS) stats->candidates += 1;
        if (prim_test(sc.prims, k, o, d, time, tmin, closest, &h, &meta, k == self)) {
          closest = h.t;
          best = k;
          best_meta = meta;
          bh = h;
          any = true;
          leaf_hit = true;
          if (ANY) return true;
        }
        if (meta & GT_LAST_BIT) break;
        ++k;
      }
      if (STATS && leaf_hit) stats->leaf_hits += 1;
    } else {
      const float4* nd = reinterpret_cast<const float4*>(sc.nodes2 + cur);
      float4 a = nd[0], b = nd[1], c = nd[2];
      uint4 refs = *reinterpret_cast<const uint4*>(nd + 3);
      float le, re;
      bool lh = slab(rp, a.x, a.y, a.z, a.w, c.x, c.y, tmin, closest, &le);
      bool rh = slab(rp, b.x, b.y, b.z, b.w, c.z, c.w, tmin, closest, &re);
      if (STATS) {
        stats->node_tests += 2;
        stats->node_hits += uint32_t(lh) + uint32_t(rh);
        uint32_t len = uint32_t(st.n) + uint32_t(lh) + uint32_t(rh);
        stats->max_stack = max(stats->max_stack, len);
      }
      if (lh && rh) {
        bool right_first = re < le;
        uint32_t nearer = right_first ? refs.y : refs.x;
        uint32_t farther = right_first ? refs.x : refs.y;
        st.push(farther, right_first ? le : re);
        cur = nearer;
        continue;
      }
      if (lh) { cur = refs.x; continue; }
      if (rh) { cur = refs.y; continue; }
    }
    // pop, skipping entries that cannot beat the current closest hit
    bool found = false;
    while (st.n > 0) {
      uint2 e = st.pop();
      if (__uint_as_float(e.y) >= closest) continue;
      cur = e.x;
      found = true;
      break;
    }
    if (!found) break;
  }
  if (any) {
    *t_out = closest;
    *prim_out = best;
    if (hit_out) *hit_out = bh;
    if (meta_out) *meta_out = best_meta;
  }
  return any;
}

// Four-wide variant: children of a wide node are visited nearest-first.
template <bool ANY, bool STATS>
__device__ __forceinline__ bool trace4(const DScene& sc, f3 o, f3 d, float time, float tmin, float tmax,
                                       uint32_t self, Stack& st, float* t_out, uint32_t* prim_out, PrimHit* hit_out,
                                       Stats* stats, uint32_t* meta_out = nullptr) {
  RayPre rp = ray_pre(o, d);
  float closest = tmax;
  bool any = false;
  uint32_t best = 0, best_meta = 0;
  PrimHit bh{0.f, 0.f, 0.f};
  float root_entry;
  if (STATS) stats->node_tests += 1;
  if (!slab(rp, sc.root_lo[0], sc.root_hi[0], sc.root_lo[1], sc.root_hi[1], sc.root_lo[2], sc.root_hi[2], tmin,
            tmax, &root_entry))
    return false;
  if (STATS) { stats->node_hits += 1; stats->max_stack = max(stats->max_stack, 1u); }
  uint32_t cur = sc.root_ref;
  st.n = 0;
  for (;;) {
    if (cur & GT_LEAF_BIT) {
      uint32_t k = cur & ~GT_LEAF_BIT;
      bool leaf_hit = false;
      if (STATS) stats->leaf_visits += 1;
      for (;;) {
        PrimHit h;
        uint32_t meta;
        if (STATS) stats->candidates += 1;
        if (prim_test(sc.prims, k, o, d, time, tmin, closest, &h, &meta, k == self)) {
          closest = h.t; best = k; best_meta = meta; bh = h; any = true; leaf_hit = true;
          if (ANY) return true;
        }
        if (meta & GT_LAST_BIT) break;
        ++k;
      }
      if (STATS && leaf_hit) stats->leaf_hits += 1;
    } else {
      const float4* nd = reinterpret_cast<const float4*>(sc.nodes4 + cur);
      float4 lox = nd[0], hix = nd[1], loy = nd[2], hiy = nd[3], loz = nd[4], hiz = nd[5];
      uint4 ch = *reinterpret_cast<const uint4*>(nd + 6);
      float e[4];
      bool h[4];
      h[0] = slab(rp, lox.x, hix.x, loy.x, hiy.x, loz.x, hiz.x, tmin, closest, &e[0]) && ch.x != GT_EMPTY_CHILD;
      h[1] = slab(rp, lox.y, hix.y, loy.y, hiy.y, loz.y, hiz.y, tmin, closest, &e[1]) && ch.y != GT_EMPTY_CHILD;
      h[2] = slab(rp, lox.z, hix.z, loy.z, hiy.z, loz.z, hiz.z, tmin, closest, &e[2]) && ch.z != GT_EMPTY_CHILD;
      h[3] = slab(rp, lox.w, hix.w, loy.w, hiy.w, loz.w, hiz.w, tmin, closest, &e[3]) && ch.w != GT_EMPTY_CHILD;
      uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
      if (STATS) {
        stats->node_tests += 4;
        stats->node_hits += uint32_t(h[0]) + uint32_t(h[1]) + uint32_t(h[2]) + uint32_t(h[3]);
      }
      // order hit children by entry (stable: lower slot first on ties), push
      // all but the nearest in far-to-near order
      int idx[4];
      int nh = 0;
      for (int k = 0; k < 4; ++k) if (h[k]) idx[nh++] = k;
      for (int i = 1; i < nh; ++i) {
        int v = idx[i];
        int j = i - 1;
        while (j >= 0 && e[idx[j]] > e[v]) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = v;
      }
      if (STATS) stats->max_stack = max(stats->max_stack, uint32_t(st.n + nh));
      if (nh > 0) {
        for (int i = nh - 1; i >= 1; --i) st.push(c[idx[i]], e[idx[i]]);
        cur = c[idx[0]];
        continue;
      }
    }
    bool found = false;
    while (st.n > 0) {
      uint2 en = st.pop();
      if (__uint_as_float(en.y) >= closest) continue;
      cur = en.x;
      found = true;
      break;
    }
    if (!found) break;
  }
  if (any) {
    *t_out = closest;
    *prim_out = best;
    if (hit_out) *hit_out = bh;
    if (meta_out) *meta_out = best_meta;
  }
  return any;
}

// self: the primitive (leaf order) the ray starts on, kNoSelf for a ray from
// a camera or a caller; prim_test decides it by rule, not by t_min.
constexpr uint32_t kNoSelf = 0xffffffffu;
template <bool ANY, bool STATS>
__device__ __forceinline__ bool trace(const DScene& sc, f3 o, f3 d, float time, float tmin, float tmax, uint32_t self,
                                      Stack& st, float* t_out, uint32_t* prim_out, PrimHit* hit_out, Stats* stats,
                                      uint32_t* meta_out = nullptr) {
  if (sc.width == 4)
    return trace4<ANY, STATS>(sc, o, d, time, tmin, tmax, self, st, t_out, prim_out, hit_out, stats, meta_out);
  return trace2<ANY, STATS>(sc, o, d, time, tmin, tmax, self, st, t_out, prim_out, hit_out, stats, meta_out);
}

// ---- light targets (scene.rs:363-609, object.rs:1048-1161) --------------------
struct DLights {
  const GLight* __restrict__ l;
  uint32_t n;
  float total;
};
__device__ __forceinline__ float area_pdf(f3 dir, float t, f3 n, float area) {  // object.rs:1101-1113
  if (area <= EPS) return 0.f;
  float d2 = len2(t * dir);
  float cosine = fabsf(dot(normalized(dir), n));
  if (cosine <= EPS) return 0.f;
  return d2 / (cosine * area);
}
__device__ __forceinline__ float light_pdf_one(const GLight& L, f3 origin, float time, f3 dir) {
  uint32_t kind = uint32_t(L.c[3]);
  if (kind == GK_SPHERE || kind == GK_MOVING_SPHERE) {  // object.rs:1048-1073
    f3 c = ld3(L.a);
    if (kind == GK_MOVING_SPHERE) c = c + time * (ld3(L.b) - c);
    float r = L.a[3];
    float root;
    if (!sphere_hit(c, r, origin, dir, SHADOW_ACNE, INF, &root)) return 0.f;
    float d2 = len2(c - origin);
    float r2 = r * r;
    if (d2 <= r2) return 1.0f / (4.0f * PI);
    float cos_max = sqrtf(1.0f - r2 / d2);
    float sa = 2.0f * PI * (1.0f - cos_max);
    return sa <= EPS ? 0.f : 1.0f / sa;
  }
  PrimHit h;
  if (kind == GK_TRIANGLE) {
    if (!tri_hit(ld3(L.a), ld3(L.b), ld3(L.c), origin, dir, SHADOW_ACNE, INF, &h)) return 0.f;
  } else {
    if (!quad_hit(ld3(L.a), ld3(L.b), ld3(L.c), ld3(L.n), L.w[0], L.w[1], L.w[3], origin, dir, SHADOW_ACNE, INF, &h))
      return 0.f;
  }
  return area_pdf(dir, h.t, ld3(L.n), L.n[3]);
}
__device__ __forceinline__ float lights_pdf(const DLights& ls, f3 origin, float time, f3 dir) {  // scene.rs:571-582
  if (ls.n == 0 || ls.total <= 0.f) return 0.f;
  float sum = 0.f;
  for (uint32_t i = 0; i < ls.n; ++i) sum += ls.l[i].b[3] * light_pdf_one(ls.l[i], origin, time, dir);
  return sum;
}
__device__ __forceinline__ f3 light_random_one(const GLight& L, f3 origin, float time, Rng& rng) {
  uint32_t kind = uint32_t(L.c[3]);
  if (kind == GK_SPHERE || kind == GK_MOVING_SPHERE) {  // object.rs:1115-1130
    f3 c = ld3(L.a);
    if (kind == GK_MOVING_SPHERE) c = c + time * (ld3(L.b) - c);
    float r = L.a[3];
    f3 dir = c - origin;
    float d2 = len2(dir);
    float r2 = r * r;
    if (d2 <= r2 || d2 <= EPS) return rng.unit_vector_spherical();
    float r1 = rng.next(), rr2 = rng.next();  // random_to_sphere :1154-1161
    float z = 1.0f + rr2 * (sqrtf(1.0f - r2 / d2) - 1.0f);
    float phi = 2.0f * PI * r1;
    float rz = sqrtf(1.0f - z * z);
    float s, cc;
    sincosf(phi, &s, &cc);
    f3 local = mk(cc * rz, s * rz, z);
    Onb b;
    return onb_from_w(dir, &b) ? b.local(local) : local;
  }
  if (kind == GK_TRIANGLE) {  // :1138-1152
    float a = rng.next();
    float b = rng.next();
    if (a + b > 1.0f) { a = 1.0f - a; b = 1.0f - b; }
    f3 p = ld3(L.a) + a * ld3(L.b) + b * ld3(L.c);
    return p - origin;
  }
  float a = rng.next();  // quad :1132-1136
  f3 p0 = ld3(L.a) + a * ld3(L.b);
  float b = rng.next();
  return (p0 + b * ld3(L.c)) - origin;
}
__device__ __forceinline__ f3 lights_random(const DLights& ls, f3 origin, float time, Rng& rng) {  // scene.rs:584-600
  if (ls.n == 0 || ls.total <= 0.f) return mk(1.f, 0.f, 0.f);
  float pick = rng.range(0.f, ls.total);
  for (uint32_t i = 0; i < ls.n; ++i) {
    pick -= ls.l[i].w[2];  // raw weight
    if (pick <= 0.f) return light_random_one(ls.l[i], origin, time, rng);
  }
  return light_random_one(ls.l[ls.n - 1], origin, time, rng);
}

// ---- environment (environment.rs) ---------------------------------------------
struct DEnv {
  uint32_t kind;        // 0 none, 1 constant, 2 texture
  uint32_t w, h;
  uint32_t uniform;     // total weight <= EPS: uniform sphere sampling
  float color[3];
  // Linear RGB texels, row-major, row 0 at +y; .w holds the texel's
  // solid-angle pdf (:136-156), so a pdf lookup shares the cache line of the
  // radiance lookup around the same direction.
  const float4* __restrict__ tex;
  // Alias table over the texels (Vose's method): slot i keeps texel i when
  // the low 32 bits of its draw are below .x and otherwise yields texel .y.
  // It samples the reference's texel distribution (rebuild_distribution,
  // :158-184) with one draw and two dependent loads instead of a search.
  const uint2* __restrict__ alias;
};
__device__ __forceinline__ float power_heuristic(float a, float b) {  // camera.rs:2486-2496
  if (!isfinite(a) || a <= EPS) return 0.f;
  if (!isfinite(b) || b <= EPS) return 1.0f;
  float s = a * a, o = b * b;
  return s / (s + o);
}
__device__ __forceinline__ float rem_euclid(float x, float m) {
  float r = fmodf(x, m);
  return r < 0.f ? r + fabsf(m) : r;
}
__device__ __forceinline__ f3 env_radiance_uv(const DEnv& e, float u, float v);
__device__ __forceinline__ f3 env_radiance(const DEnv& e, f3 dir) {  // :98-107
  if (e.kind == 1) return ld3(e.color);
  f3 d = normalized(dir);
  float phi = rem_euclid(atan2f(d.z, d.x), TAU);
  float theta = acosf(clampf(d.y, -1.f, 1.f));
  float u = phi / TAU, v = 1.0f - theta / PI;
  if (!isfinite(u) || !isfinite(v)) return mk(0.f, 0.f, 0.f);
  return env_radiance_uv(e, u, v);
}
// pdf_value (:136-156) of a direction at lat-long angles (phi, theta): the
// solid-angle pdf of the texel that holds it.
__device__ __forceinline__ float env_pdf_at(const DEnv& e, float phi, float theta) {
  if (e.kind != 2 || e.uniform) return 1.0f / (4.0f * PI);
  float fx = floorf((phi / TAU) * float(e.w)), fy = floorf((theta / PI) * float(e.h));
  uint32_t x = fx > 0.f ? min(uint32_t(fx), e.w - 1) : 0u;
  uint32_t y = fy > 0.f ? min(uint32_t(fy), e.h - 1) : 0u;
  return e.tex[size_t(y) * e.w + x].w;
}
// env_radiance(dir) for a path that missed. On entry *mw is the path's MIS
// weight, or -cpdf when the weight is the environment's power heuristic
// power_heuristic(cpdf, pdf_value(dir)); on return it is the weight. The
// direction maps to lat-long coordinates once for both lookups.
__device__ __forceinline__ f3 env_radiance_weighted(const DEnv& e, f3 dir, float* mw) {
  if (e.kind != 2) {
    if (*mw < 0.f) *mw = power_heuristic(-*mw, env_pdf_at(e, 0.f, 0.f));
    return ld3(e.color);
  }
  f3 d = normalized(dir);
  float phi = rem_euclid(atan2f(d.z, d.x), TAU);
  float theta = acosf(clampf(d.y, -1.f, 1.f));
  if (*mw < 0.f) *mw = power_heuristic(-*mw, env_pdf_at(e, phi, theta));
  float u = phi / TAU, v = 1.0f - theta / PI;
  if (!isfinite(u) || !isfinite(v)) return mk(0.f, 0.f, 0.f);
  return env_radiance_uv(e, u, v);
}
// The sampled texel's radiance and pdf. The reference evaluates
// pdf_value(direction) and radiance(direction) from the returned direction
// (environment.rs:130-131); the direction lies in texel (x, y) by
// construction, so the texel's pdf and the lat-long coordinates
// u = (x + r) / w, theta = acos(cos theta) are used directly.
struct EnvSample {
  f3 dir;
  float pdf;
  float u, cos_theta;
  bool texel;
};
__device__ __forceinline__ f3 env_radiance_uv(const DEnv& e, float u, float v) {  // texture.rs:681-711
  float cx = rem_euclid(u, 1.0f) * float(e.w);
  uint32_t x0 = min(uint32_t(floorf(cx)), e.w - 1);
  uint32_t x1 = (x0 + 1 == e.w) ? 0 : x0 + 1;
  float cy = clampf(1.0f - v, 0.f, 1.f) * float(e.h - 1);
  uint32_t y0 = min(uint32_t(floorf(cy)), e.h - 1);
  uint32_t y1 = min(y0 + 1, e.h - 1);
  float tx = cx - float(x0), ty = cy - float(y0);
  float4 a = e.tex[size_t(y0) * e.w + x0], b = e.tex[size_t(y0) * e.w + x1];
  float4 c = e.tex[size_t(y1) * e.w + x0], f = e.tex[size_t(y1) * e.w + x1];
  f3 top = mk(a.x + (b.x - a.x) * tx, a.y + (b.y - a.y) * tx, a.z + (b.z - a.z) * tx);
  f3 bot = mk(c.x + (f.x - c.x) * tx, c.y + (f.y - c.y) * tx, c.z + (f.z - c.z) * tx);
  return top + (bot - top) * ty;
}
__device__ __forceinline__ EnvSample env_sample(const DEnv& e, Rng& rng) {  // :109-132
  EnvSample r;
  if (e.kind != 2 || e.uniform) {
    r.dir = rng.unit_vector_spherical();
    r.pdf = 1.0f / (4.0f * PI);
    r.texel = false;
    return r;
  }
  // The texel: slot = floor(draw * n) from the high bits of draw * n, the
  // keep-or-alias test from its low bits (one 64-bit draw, integer math).
  const uint64_t draw = rng.next_u64();
  const uint64_t n = uint64_t(e.w) * e.h;
  const uint32_t slot = uint32_t(__umul64hi(draw, n));
  const uint32_t frac = uint32_t((draw * n) >> 32);
  const uint2 rec = e.alias[slot];
  const uint32_t texel = frac < rec.x ? slot : rec.y;
  const uint32_t x = texel % e.w, y = texel / e.w;
  float u = (float(x) + rng.next()) / float(e.w);
  float t0 = PI * (float(y) / float(e.h)), t1 = PI * (float(y + 1) / float(e.h));
  float c0 = cosf(t0), c1 = cosf(t1);
  float ct = c0 + rng.next() * (c1 - c0);
  float st = sqrtf(fmaxf(1.0f - ct * ct, 0.f));
  float phi = TAU * u;
  float s, c;
  sincosf(phi, &s, &c);
  r.dir = mk(st * c, ct, st * s);
  r.pdf = e.tex[size_t(y) * e.w + x].w;
  r.u = u;
  r.cos_theta = ct;
  r.texel = true;
  return r;
}
__device__ __forceinline__ f3 env_sample_radiance(const DEnv& e, const EnvSample& s) {
  if (e.kind == 1) return ld3(e.color);
  if (!s.texel) return env_radiance(e, s.dir);
  float theta = acosf(clampf(s.cos_theta, -1.f, 1.f));
  return env_radiance_uv(e, s.u, 1.0f - theta / PI);
}


// ---- camera rays (camera.rs:1658-1718) ------------------------------------------
__device__ __forceinline__ void camera_ray(const GCamera& cam, uint32_t x, uint32_t y, uint32_t s, Rng& rng, f3* o,
                                           f3* d, float* time) {
  float ox, oy;
  if (cam.mode == 0) {
    ox = rng.next() - 0.5f;
    oy = rng.next() - 0.5f;
  } else {
    uint32_t sx = s % cam.grid, sy = s / cam.grid;
    float rec = 1.0f / float(cam.grid);
    ox = ((float(sx) + rng.next()) * rec) - 0.5f;
    oy = ((float(sy) + rng.next()) * rec) - 0.5f;
  }
  f3 ps = ld3(cam.pixel00) + (float(x) + ox) * ld3(cam.du) + (float(y) + oy) * ld3(cam.dv);
  f3 origin = ld3(cam.center);
  if (cam.defocus) {
    f3 dk = rng.in_unit_disk();
    origin = origin + dk.x * ld3(cam.disk_u) + dk.y * ld3(cam.disk_v);
  }
  *time = rng.range(cam.shutter_start, cam.shutter_end);
  *o = origin;
  *d = ps - origin;
}
__device__ __forceinline__ f3 background(const GCamera& cam, f3 dir) {  // camera.rs:297-313
  if (cam.background_kind == 1) {
    f3 nadir = ld3(cam.background), horizon = ld3(cam.background + 3), zenith = ld3(cam.background + 6);
    float t = 0.5f * (normalized(dir).y + 1.0f);
    if (t < 0.5f) { float bl = 2.0f * t; return nadir * (1.0f - bl) + horizon * bl; }
    float bl = 2.0f * (t - 0.5f);
    return horizon * (1.0f - bl) + zenith * bl;
  }
  return ld3(cam.background);
}

}  // namespace gtd
