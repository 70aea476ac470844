// kernels.hip — the wavefront path tracer for gfx950 (CDNA4).
//
// One batch of paths (a set of samples for every pixel) runs as
//   reset -> { extend(b) -> shade(b, every category) -> connect(b, light)
//   -> connect(b, env) } for b < max_depth -> accumulate
// where extend(0) generates the camera rays itself
// over path state in HBM. Every stage is a persistent kernel: a fixed grid
// whose waves pull chunks of work from a device-side cursor, so the host never
// reads a queue length and the loop needs no synchronisation.
//
// Path state sits at the path's index and is updated in place; queues are
// lists of path ids. A wave appends everything it produced from one chunk
// with a single atomic per queue: same-address atomics are served at the
// memory side across the chip's eight XCDs, and one per wave per item cost
// more than the traversal itself.
//
// The per-path integrator is RayCamera::ray_color (camera.rs:1730-1862), split
// at its two traversals: the closest hit of the path ray (extend) and the
// shadow rays of next-event estimation (connect). Radiance accumulates per
// path in the order the reference adds it (emission, light NEE, environment
// NEE, next bounce), so a path's result does not depend on queue order.
#include "kernels.h"

#include <algorithm>
#include <mutex>

#include "../../include/gartus_hip.h"

using namespace gtd;

// Traversal-stack overflow of gt_intersect_device_noalloc, which allocates
// nothing: kNoAllocStackEntries 8-byte entries (16 MiB) per GPU, held until
// the process ends. Every call on a GPU shares that GPU's copy, so calls that
// overlap on one GPU race.
static __device__ __align__(256) uint2 scratch[gtk::kNoAllocStackEntries];

namespace gtk {

namespace {

// Items per lane per cursor fetch (128 per wave). On one MI355X, 2 against 4:
// 1080p GGX/environment frame 14.9 -> 14.8 ms, Cornell 4.1 -> 3.9 ms, 64-scene
// batch 325 -> 318 ms; 1 was no better than 4.
constexpr uint32_t kChunk = 2;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned long long lanes_below() {
  return lane_id() == 0 ? 0ull : ((~0ull) >> (64 - lane_id()));
}

// Claims per_lane * 64 consecutive work items for the calling wave.
__device__ __forceinline__ uint32_t wave_fetch(uint32_t* cursor, uint32_t per_lane) {
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(cursor, per_lane * 64u);
  return __shfl(base, 0);
}

// Appends the ids of the items of one chunk whose bit is set, one atomic per
// wave: bit k of `bits` marks item k of this lane, `ids[k]` its path id.
__device__ __forceinline__ void wave_append_chunk(uint32_t* counter, uint32_t* list, uint32_t bits,
                                                  const uint32_t (&ids)[kChunk]) {
  uint32_t total = 0;
#pragma unroll
  for (uint32_t k = 0; k < kChunk; ++k) total += uint32_t(__popcll(__ballot((bits >> k) & 1u)));
  if (total == 0) return;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(counter, total);
  base = __shfl(base, 0);
  const unsigned long long below = lanes_below();
#pragma unroll
  for (uint32_t k = 0; k < kChunk; ++k) {
    unsigned long long m = __ballot((bits >> k) & 1u);
    if ((bits >> k) & 1u) list[base + uint32_t(__popcll(m & below))] = ids[k];
    base += uint32_t(__popcll(m));
  }
}

__device__ __forceinline__ Stack make_stack(const Queues& q, uint2* smem) {
  Stack st;
  st.lds = smem + threadIdx.x;
  st.stride = blockDim.x;
  st.glob = q.gstack + (blockIdx.x * blockDim.x + threadIdx.x);
  st.gstride = q.gstride;
  st.cap = int(q.stack_cap);
  st.n = 0;
  return st;
}

// Pixel and sample of a path: a segment's paths are sample-major,
// path = path0 + (sample - sample0) * nact + idx, pixel = pixmap[idx] (or idx).
__device__ __forceinline__ void path_coords(const Frame& f, const Seg& sg, uint32_t path, uint32_t* x, uint32_t* y,
                                            uint32_t* sample) {
  const uint32_t local = path - sg.path0;
  const uint32_t idx = local % f.nact;
  const uint32_t pixel = f.pixmap ? f.pixmap[idx] : idx;
  *sample = sg.sample0 + local / f.nact;
  *x = pixel % f.cam.width;
  *y = pixel / f.cam.width;
}

// Index of the segment holding `key` in a table sorted by `field`
// (the last entry whose field is <= key).
template <uint32_t Seg::*FIELD>
__device__ __forceinline__ uint32_t find_seg(const Batch& bt, uint32_t key) {
  uint32_t lo = 0, n = bt.nseg;
  while (n > 1) {
    const uint32_t half = n >> 1;
    if (bt.segs[lo + half].*FIELD <= key) lo += half;
    n -= half;
  }
  return lo;
}

// Runs body(frame, segment) for a lane's path with the frame wave-uniform.
// A single-frame render passes its frame and segment as kernel arguments. A
// batched render looks the segment up and loops once per distinct segment
// among the wave's lanes (a "waterfall"), so the frame is always read at a
// uniform address; segments are contiguous in path id, so a wave normally
// sees one.
template <bool MULTI, uint32_t Seg::*FIELD = &Seg::path0, typename Body>
__device__ __forceinline__ void with_frame(const Frame& f0, const Batch& bt, bool valid, uint32_t key, Body&& body) {
  if constexpr (!MULTI) {
    if (valid) body(f0, bt.seg0);
  } else {
    const uint32_t mine = valid ? find_seg<FIELD>(bt, key) : 0u;
    bool pending = valid;
    while (__ballot(pending)) {
      if (pending) {
        const uint32_t u = __builtin_amdgcn_readfirstlane(mine);
        if (mine == u) {
          const Seg sg = bt.segs[u];
          body(bt.frames[sg.frame], sg);
          pending = false;
        }
      }
    }
  }
}

// The direction's fourth word: bit 31 = the ray was scattered (emission it
// meets is not counted: NEE took it), bits 0-30 = 1 + the primitive the ray
// leaves (leaf order), 0 for a camera ray.
__device__ __forceinline__ uint32_t ray_word(bool scattered, uint32_t self) {
  return (scattered ? 0x80000000u : 0u) | ((self + 1u) & 0x7fffffffu);
}
__device__ __forceinline__ uint32_t ray_self(float w) { return (__float_as_uint(w) & 0x7fffffffu) - 1u; }

// Russian roulette (camera.rs:2498-2523)
__device__ __forceinline__ bool rr_survives(uint32_t bounce, int32_t min_depth, f3* T, Rng& rng) {
  if (min_depth < 0 || bounce < uint32_t(min_depth)) return true;
  float m = fmaxf(fmaxf(T->x, T->y), T->z);
  if (!isfinite(m) || m <= EPS) return false;
  float p = clampf(m, 0.05f, 0.95f);
  if (rng.next() >= p) return false;
  const float inv = 1.0f / p;  // LinearRgb / f64 multiplies by the reciprocal (colors.rs:185-190)
  *T = mk(inv * T->x, inv * T->y, inv * T->z);
  return true;
}

// Material::scattering_pdf (material.rs:522-529, :1139-1141, :1286-1291,
// :1596-1612; the others return 0). n is the oriented geometric normal, ns
// the shading normal (equal to n except on mesh triangles).
__device__ __forceinline__ float scattering_pdf(const GMat& m, f3 d_in, f3 n, f3 ns, f3 dir) {
  if (m.kind == GM_LAMBERTIAN) {
    float c = dot(ns, normalized(dir));
    return c <= 0.f ? 0.f : c / PI;
  }
  if (m.kind == GM_GGX) {
    f3 outgoing = -normalized(d_in);
    f3 incoming = normalized(dir);
    if (dot(incoming, n) <= 0.f || dot(outgoing, n) <= 0.f) return 0.f;
    return ggx_weight(ns, outgoing, incoming, m.param, m.f0);
  }
  if (m.kind == GM_ISOTROPIC) return 1.0f / (4.0f * PI);
  if (m.kind == GM_HENYEY_GREENSTEIN) {  // HenyeyGreensteinPdf::new(ray_in.direction, g).value
    const f3 w = normalized(d_in);
    const float l = sqrtf(len2(dir));
    if (len2(w) <= EPS || l <= EPS) return 0.f;
    return hg_phase(dot(dir, w) / l, m.param);
  }
  return 0.f;
}

// Clears the counters of bounces 0..depth (one word per 128-byte line) and
// the tallies.
// Clears the counters, except that slice s of bounce 0 queues the camera
// paths [total * s / kSlices, total * (s + 1) / kSlices).
__global__ void k_reset(uint32_t* ctr, uint32_t ncounters, unsigned long long* tally, uint32_t ntally,
                        uint32_t total) {
  static_assert(S_QCOUNT == 0 && S_BASE == 1, "bounce 0 starts with the queue count and base of each slice");
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ncounters; i += gridDim.x * blockDim.x) {
    uint32_t v = 0;
    if (i < 2 * kSlices) {  // ctr_at(0, S_QCOUNT or S_BASE, s)
      const uint32_t s = i % kSlices;
      const uint32_t lo = uint32_t(uint64_t(total) * s / kSlices), hi = uint32_t(uint64_t(total) * (s + 1) / kSlices);
      v = i / kSlices == S_BASE ? lo : hi - lo;
    }
    ctr[size_t(i) * kLineWords] = v;
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntally; i += gridDim.x * blockDim.x) tally[i] = 0;
}

// Slices a block works on: b % G, b % G + G, ... with G = min(grid, kSlices),
// so every slice has at least one block whatever the grid.
__device__ __forceinline__ uint32_t slice_step() { return gridDim.x < kSlices ? gridDim.x : kSlices; }

// Items per lane per claim on a slice queue of `count` items: kChunk (one
// atomic per kChunk * 64 items) while the queue holds two such claims for
// every wave serving the slice, else one, so that a short queue (late
// bounces, few samples) spreads over all the waves instead of a few that each
// trace several items one after another.
__device__ __forceinline__ uint32_t items_per_lane(uint32_t count) {
  const uint32_t waves = (gridDim.x / slice_step()) * (blockDim.x / 64u);
  return count >= waves * (2u * kChunk * 64u) ? kChunk : 1u;
}

// Wave-reduced 64-bit add of per-lane counts (one atomic per wave).
__device__ __forceinline__ void wave_count(unsigned long long* dst, uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if (lane_id() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}
__device__ __forceinline__ unsigned long long* count_slot(const Queues& q, uint32_t bounce, int k) {
  return q.tally + 1 + 4 * bounce + k;
}
__device__ __forceinline__ uint32_t node_fetches(const Stats& s, uint32_t width) {
  // every internal node fetched tests `width` child boxes; the root box test is not a fetch
  return (s.node_tests - 1) / width;
}

// Camera rays for samples [sample0, sample0 + nsamples) of every pixel
// (render_world_linear_pixel_with_context camera.rs:2660-2723; rays :1658-1718).
// Closest hit of every active path; each hit is binned by the shading
// category of its material; a miss is shaded on the spot. At bounce 0
// (CAMERA) the kernel generates the camera rays itself (camera.rs:1658-1718)
// and stores a path's state only if it hit something: a path that misses
// ends here with the background as its radiance.
__device__ __forceinline__ uint32_t category_of(uint32_t meta) {
  switch ((meta >> GT_MATKIND_SHIFT) & 7u) {
    case GM_DIFFUSE_LIGHT: return CAT_EMIT;
    case GM_LAMBERTIAN:
    case GM_ISOTROPIC:
    case GM_HENYEY_GREENSTEIN: return CAT_LAMBERT;  // the scattering materials without a GGX lobe
    case GM_GGX: return CAT_GGX;
    default: return CAT_SPECULAR;
  }
}

template <bool COUNT, bool MULTI, bool CAMERA>
// Traversal kernels run at 8 waves per SIMD (at most 64 VGPRs, a few spilled
// values): more waves hide more of the dependent node-fetch latency. On one
// MI355X this took the 720p terrain's extend from 14.6 to 13.8 ms per frame
// and the Cornell frame from 5.0 to 4.8 ms; a 5-wave target for the shade
// kernels was slower.
__global__ __attribute__((amdgpu_waves_per_eu(8))) void k_extend(Frame f0, Batch bt, Queues q, uint32_t bounce) {
  extern __shared__ uint2 smem[];
  Stack st = make_stack(q, smem);
  for (uint32_t s = blockIdx.x % slice_step(); s < kSlices; s += slice_step()) {
  const uint32_t count = q.ctr[ctr_at(bounce, S_QCOUNT, s)];
  if (count == 0) continue;
  const uint32_t first = CAMERA ? q.ctr[ctr_at(0, S_BASE, s)] : 0;
  const uint32_t* __restrict__ active = q.active[bounce & 1] + size_t(s) * q.region;
  const uint32_t per_lane = items_per_lane(count);
  for (;;) {
    uint32_t base = wave_fetch(q.ctr + ctr_at(bounce, S_EXT_WORK, s), per_lane);
    if (base >= count) break;
    uint32_t nodes = 0, prims = 0;
    uint32_t ids[kChunk], cats[kChunk];
#pragma unroll
    for (uint32_t k = 0; k < kChunk; ++k) {
      uint32_t j = base + k * 64u + lane_id();
      cats[k] = kNumCategories;
      ids[k] = 0;
      const bool valid = k < per_lane && j < count;
      const uint32_t p = !valid ? 0u : CAMERA ? first + j : active[j];
      ids[k] = p;
      with_frame<MULTI>(f0, bt, valid, p, [&](const Frame& f, const Seg& sg) {
        float4 o4, d4;
        if (CAMERA) {
          if (f.cam.max_depth == 0) {  // ray_color with no bounce is black
            q.radiance[p] = make_float4(0.f, 0.f, 0.f, 0.f);
            return;
          }
          uint32_t x, y, sample;
          path_coords(f, sg, p, &x, &y, &sample);
          Rng rng{stream_seed(pixel_seed(f.cam.seed, x, y), sample, 0)};
          f3 o, d;
          float time;
          camera_ray(f.cam, x, y, sample, rng, &o, &d, &time);
          o4 = make_float4(o.x, o.y, o.z, time);
          d4 = make_float4(d.x, d.y, d.z, 0.f);
          q.ray_o[p] = o4;  // stored before the trace: holding the ray across it costs registers
          q.ray_d[p] = d4;
        } else {
          o4 = q.ray_o[p];
          d4 = q.ray_d[p];
        }
        float t;
        uint32_t prim, meta = 0;
        Stats s{0, 0, 0, 0, 0, 0};
        bool hit = trace<false, COUNT>(f.scene, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), o4.w, SHADOW_ACNE, INF,
                                       CAMERA ? kNoSelf : ray_self(d4.w), st, &t, &prim, nullptr, &s, &meta);
        if (hit) {
          q.hits[p] = make_float2(t, __uint_as_float(prim));
          // a camera path has no radiance yet; its throughput and MIS weight are 1, which shade
          // assumes at bounce 0 without reading them
          if (CAMERA) q.radiance[p] = make_float4(0.f, 0.f, 0.f, 0.f);
          cats[k] = category_of(meta);
        } else {
          // a miss ends the path here (camera.rs:1747-1750): background or environment, times the
          // MIS weight; shading it in place saves a queue and a launch per bounce
          const float4 t4 = CAMERA ? make_float4(1.f, 1.f, 1.f, 1.f) : q.thr[p];
          const f3 d = mk(d4.x, d4.y, d4.z);
          float mw = t4.w;
          const f3 rad = f.env.kind != 0 ? env_radiance_weighted(f.env, d, &mw) : background(f.cam, d);
          const f3 miss = rad * mw;
          const float4 L = CAMERA ? make_float4(0.f, 0.f, 0.f, 0.f) : q.radiance[p];
          q.radiance[p] = make_float4(L.x + t4.x * miss.x, L.y + t4.y * miss.y, L.z + t4.z * miss.z, 0.f);
        }
        if (COUNT) { nodes += node_fetches(s, f.scene.width); prims += s.candidates; }
      });
    }
#pragma unroll
    for (uint32_t c = CAT_EMIT; c < kNumCategories; ++c) {  // misses were shaded above
      uint32_t bits = 0;
#pragma unroll
      for (uint32_t k = 0; k < kChunk; ++k) bits |= uint32_t(cats[k] == c) << k;
      wave_append_chunk(q.ctr + ctr_at(bounce, S_CAT_COUNT + c, s), q.cat[c] + size_t(s) * q.region, bits, ids);
    }
    if (COUNT) {
      wave_count(count_slot(q, bounce, 0), nodes);
      wave_count(count_slot(q, bounce, 1), prims);
    }
  }
  }
}

// ---- shading: one bounce of ray_color (camera.rs:1742-1858), by category --
struct Item {
  uint32_t path;
  bool after_scatter;
  f3 o, d, T;
  float time, t;
  uint32_t prim;
};
// CAMERA: the hits of camera rays made by k_extend<.., CAMERA = true>, whose
// paths start with throughput 1 and MIS weight 1 and never store them.
template <bool CAMERA>
__device__ __forceinline__ Item load_item(const Queues& q, uint32_t p) {
  Item it;
  float4 o4 = q.ray_o[p], d4 = q.ray_d[p], t4 = CAMERA ? make_float4(1.f, 1.f, 1.f, 1.f) : q.thr[p];
  float2 h = q.hits[p];
  it.path = p;
  it.after_scatter = (__float_as_uint(d4.w) >> 31) != 0;
  it.o = mk(o4.x, o4.y, o4.z);
  it.time = o4.w;
  it.d = mk(d4.x, d4.y, d4.z);
  it.T = mk(t4.x, t4.y, t4.z);
  it.t = h.x;
  it.prim = __float_as_uint(h.y);
  return it;
}

// Outputs of one shaded path, written in place at the path's index; the
// returned bits say which queues the path joins (1 continuation, 2 light
// shadow ray, 4 environment shadow ray).
enum { OUT_EXT = 1, OUT_SL = 2, OUT_SE = 4 };

// The continuation's direction and throughput; its origin is the vertex,
// written by the caller.
__device__ __forceinline__ uint32_t continue_dir(const Queues& q, uint32_t p, f3 dir, bool scattered, uint32_t self,
                                                 f3 T, float mw) {
  q.ray_d[p] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(ray_word(scattered, self)));
  q.thr[p] = make_float4(T.x, T.y, T.z, mw);
  return OUT_EXT;
}
__device__ __forceinline__ uint32_t continue_path(const Queues& q, uint32_t p, f3 pt, float time, f3 dir,
                                                  bool scattered, uint32_t self, f3 T, float mw) {
  q.ray_o[p] = make_float4(pt.x, pt.y, pt.z, time);
  return continue_dir(q, p, dir, scattered, self, T, mw);
}

// Lambertian (material.rs:508-532), Isotropic (:1123-1141), Henyey-Greenstein
// (:1270-1291) or GGX (:1583-1616) vertex: light NEE, environment NEE, then
// the continuation sample (camera.rs:1779-1857). n is the oriented geometric
// normal, ns the shading normal.
template <bool GGX>
__device__ __forceinline__ uint32_t scatter_vertex(const Frame& f, const Queues& q, const Item& it, const GMat& m,
                                                   f3 p, f3 n, f3 ns, Rng& rng, bool last, uint32_t bounce) {
  const bool has_env = f.env.kind != 0;
  uint32_t out = 0;
  MatPdf pdf;
  pdf.kind = GGX ? PDF_GGX : m.kind == GM_ISOTROPIC ? PDF_SPHERE : m.kind == GM_HENYEY_GREENSTEIN ? PDF_HG : PDF_COSINE;
  bool ok = true;
  if (pdf.kind == PDF_HG) {  // HenyeyGreensteinPdf::new (sampling.rs:96-112): basis around the incident direction
    pdf.g = m.param;
    ok = onb_from_w(it.d, &pdf.basis);
  } else if (pdf.kind != PDF_SPHERE) {
    ok = onb_from_w(ns, &pdf.basis);
  }
  if (GGX) {  // GgxReflectionPdf::new (sampling.rs:183-202)
    f3 outgoing = normalized(-normalized(it.d));
    ok = ok && len2(outgoing) > EPS && dot(outgoing, pdf.basis.w) > 0.f;
    pdf.outgoing = outgoing;
    pdf.roughness = m.param;
    pdf.alpha = ggx_alpha(m.param);
  }
  if (!ok) return 0;  // scatter returns None: the path ends
  const f3 color = ld3(m.color);
  f3 T = it.T;
  if (f.nee) {
    if (f.has_lights) {  // estimate_direct_lighting camera.rs:2225-2262
      f3 ldir = lights_random(f.lights, p, it.time, rng);
      float lpdf = lights_pdf(f.lights, p, it.time, ldir);
      if (isfinite(lpdf) && lpdf > EPS) {
        float spdf = scattering_pdf(m, it.d, n, ns, ldir);
        if (isfinite(spdf) && spdf > 0.f) {
          f3 w = mul(T, color) * (spdf / lpdf);
          q.s_d[0][it.path] = make_float4(ldir.x, ldir.y, ldir.z, __uint_as_float(it.prim));
          q.s_w[0][it.path] = make_float4(w.x, w.y, w.z, 0.f);
          out |= OUT_SL;
        }
      }
    }
    if (has_env) {  // estimate_environment_lighting camera.rs:2264-2297
      EnvSample es = env_sample(f.env, rng);
      if (isfinite(es.pdf) && es.pdf > EPS) {
        float spdf = scattering_pdf(m, it.d, n, ns, es.dir);
        if (isfinite(spdf) && spdf > 0.f) {
          float mis = power_heuristic(es.pdf, spdf);
          f3 w = mul(T, mul(color, env_sample_radiance(f.env, es)) * (mis * spdf / es.pdf));
          q.s_d[1][it.path] = make_float4(es.dir.x, es.dir.y, es.dir.z, __uint_as_float(it.prim));
          q.s_w[1][it.path] = make_float4(w.x, w.y, w.z, 0.f);
          out |= OUT_SE;
        }
      }
    }
  }
  if (last) return out;  // the bounce loop ends here; the continuation is never traced
  // SamplingStrategy::continuation_sample camera.rs:165-195
  f3 dir;
  float cpdf;
  bool suppress, weight_env;
  bool below = false;
  if (!f.nee && f.has_lights) {
    float wgt = f.light_weight;
    if (rng.next() < wgt) dir = lights_random(f.lights, p, it.time, rng);
    else dir = pdf.generate(rng, &below);
    cpdf = wgt * lights_pdf(f.lights, p, it.time, dir) + (1.0f - wgt) * pdf.value(dir);
    suppress = false;
    weight_env = false;
  } else {
    dir = pdf.generate(rng, &below);
    cpdf = pdf.value(dir);
    suppress = f.nee && f.has_lights;
    weight_env = f.nee && has_env;
  }
  if (below && (f.cam.flags & GT_FLAG_GGX_REJECT_BELOW_HORIZON)) return out;
  if (!(isfinite(cpdf) && cpdf > EPS)) return out;
  float spdf = scattering_pdf(m, it.d, n, ns, dir);
  if (!(isfinite(spdf) && spdf > 0.f)) return out;
  // The environment's MIS weight power_heuristic(cpdf, pdf_value(dir)) applies
  // only if the continuation misses: it is stored as -cpdf (cpdf > EPS here)
  // and evaluated by the miss shading in extend.
  float next_w = weight_env ? -cpdf : 1.0f;
  T = mul(T, color * (spdf / cpdf));
  if (rr_survives(bounce, f.cam.rr_min_depth, &T, rng)) out |= continue_dir(q, it.path, dir, suppress, it.prim, T, next_w);
  return out;
}
// The vertex also becomes the origin of its shadow rays and continuation: one
// write serves all three (connect reads it before the next extend does).
template <bool GGX>
__device__ __forceinline__ uint32_t shade_scatter(const Frame& f, const Queues& q, const Item& it, const GMat& m, f3 p,
                                                  f3 n, f3 ns, Rng& rng, bool last, uint32_t bounce) {
  const uint32_t out = scatter_vertex<GGX>(f, q, it, m, p, n, ns, rng, last, bounce);
  if (out) q.ray_o[it.path] = make_float4(p.x, p.y, p.z, it.time);
  return out;
}

// Metal (material.rs:2087-2107: reflection about the shading normal, absorbed
// below the geometric one) or Dielectric (:2321-2350, geometric normal) vertex.
__device__ __forceinline__ uint32_t shade_specular(const Frame& f, const Queues& q, const Item& it, const GMat& m, f3 p,
                                                   f3 n, f3 ns, bool front, Rng& rng, bool last, uint32_t bounce) {
  f3 T = it.T;
  f3 dir;
  if (m.kind == GM_METAL) {
    f3 refl = reflected(normalized(it.d), ns);
    dir = refl + m.param * rng.unit_vector_spherical();
    if (dot(dir, n) <= 0.f) return 0;  // absorbed
    T = mul(T, ld3(m.color));
  } else {
    float ratio = front ? 1.0f / m.param : m.param;
    f3 unit = normalized(it.d);
    float ct = fminf(dot(-unit, n), 1.0f);
    float stt = sqrtf(1.0f - ct * ct);
    bool cannot = ratio * stt > 1.0f;
    float r0 = (1.0f - ratio) / (1.0f + ratio);
    r0 = r0 * r0;
    bool refl = cannot || (r0 + (1.0f - r0) * pow5(1.0f - ct)) > rng.next();
    dir = refl ? reflected(unit, n) : refracted(unit, n, ratio);
  }
  if (!last && rr_survives(bounce, f.cam.rr_min_depth, &T, rng))
    return continue_path(q, it.path, p, it.time, dir, false, it.prim, T, 1.0f);
  return 0;
}

// Shades the hits of one category, all slices.
template <int CAT, bool MULTI, bool CAMERA>
__device__ __forceinline__ void shade_category(const Frame& f0, const Batch& bt, const Queues& q, uint32_t bounce) {
  constexpr bool kScatters = CAT == CAT_LAMBERT || CAT == CAT_GGX || CAT == CAT_SPECULAR;
  for (uint32_t s = blockIdx.x % slice_step(); s < kSlices; s += slice_step()) {
  const uint32_t count = q.ctr[ctr_at(bounce, S_CAT_COUNT + CAT, s)];
  if (count == 0) continue;
  const size_t region = size_t(s) * q.region;
  const uint32_t* __restrict__ list = q.cat[CAT] + region;
  const uint32_t per_lane = items_per_lane(count);
  for (;;) {
    uint32_t base = wave_fetch(q.ctr + ctr_at(bounce, S_CAT_WORK + CAT, s), per_lane);
    if (base >= count) break;
    uint32_t ids[kChunk], outs[kChunk];
#pragma unroll
    for (uint32_t k = 0; k < kChunk; ++k) {
      const uint32_t j = base + k * 64u + lane_id();
      const bool valid = k < per_lane && j < count;
      const uint32_t p = valid ? list[j] : 0u;
      ids[k] = p;
      outs[k] = 0;
      with_frame<MULTI>(f0, bt, valid, p, [&](const Frame& f, const Seg& sg) {
        Item it = load_item<CAMERA>(q, p);
        PrimHit ph{it.t, 0.f, 0.f};  // the chosen primitive's distance in FP64 (traversal decides in FP32)
        const bool refined = refine_hit(f.scene.prims, it.prim, it.o, it.d, it.time, SHADOW_ACNE, &ph);
        it.t = ph.t;
        f3 pt = it.o + it.t * it.d;
        uint32_t mid, kind;
        f3 outward = prim_outward_normal(f.scene.prims, it.prim, pt, it.time, &mid, &kind);
        bool front = dot(it.d, outward) < 0.f;  // object.rs:127-149
        f3 n = front ? outward : -outward;
        f3 ns = n;
        if (kind == GK_TRIANGLE) {  // a mesh triangle interpolates its vertex normals
          if (!refined) tri_barycentric(f.scene.prims, it.prim, pt, &ph.u, &ph.v);
          ns = shading_normal(f.scene.prims, f.scene.vnormals, it.prim, ph.u, ph.v, front, n);
        }
        const GMat m = f.scene.mats[mid];
        if (CAT == CAT_EMIT) {  // emitted_at camera.rs:1753-1755, DiffuseLight material.rs:1009-1022
          if (!it.after_scatter && front) {
            float4 L = q.radiance[p];
            q.radiance[p] = make_float4(L.x + it.T.x * m.color[0], L.y + it.T.y * m.color[1],
                                        L.z + it.T.z * m.color[2], 0.f);
          }
          return;
        }
        uint32_t x, y, sample;
        path_coords(f, sg, p, &x, &y, &sample);
        Rng rng{stream_seed(pixel_seed(f.cam.seed, x, y), sample, 1 + bounce)};
        const bool last = bounce + 1 >= f.cam.max_depth;  // the camera's final bounce
        if (CAT == CAT_SPECULAR) outs[k] = shade_specular(f, q, it, m, pt, n, ns, front, rng, last, bounce);
        else outs[k] = shade_scatter<CAT == CAT_GGX>(f, q, it, m, pt, n, ns, rng, last, bounce);
      });
    }
    if (kScatters) {
      uint32_t bits_ext = 0, bits_sl = 0, bits_se = 0;
#pragma unroll
      for (uint32_t k = 0; k < kChunk; ++k) {
        bits_ext |= uint32_t((outs[k] & OUT_EXT) != 0) << k;
        bits_sl |= uint32_t((outs[k] & OUT_SL) != 0) << k;
        bits_se |= uint32_t((outs[k] & OUT_SE) != 0) << k;
      }
      wave_append_chunk(q.ctr + ctr_at(bounce + 1, S_QCOUNT, s), q.active[(bounce + 1) & 1] + region, bits_ext, ids);
      wave_append_chunk(q.ctr + ctr_at(bounce, S_SL_COUNT, s), q.shadow[0] + region, bits_sl, ids);
      wave_append_chunk(q.ctr + ctr_at(bounce, S_SE_COUNT, s), q.shadow[1] + region, bits_se, ids);
    }
  }
  }
}

// One launch per bounce drains the category queues in turn, so a wave that
// finds one queue empty moves on to the next without a launch boundary.
// CAMERA matches the extend launch of the same bounce (camera_bounce).
template <bool MULTI, bool CAMERA>
__global__ void k_shade(Frame f0, Batch bt, Queues q, uint32_t bounce) {
  shade_category<CAT_EMIT, MULTI, CAMERA>(f0, bt, q, bounce);
  shade_category<CAT_LAMBERT, MULTI, CAMERA>(f0, bt, q, bounce);
  shade_category<CAT_GGX, MULTI, CAMERA>(f0, bt, q, bounce);
  shade_category<CAT_SPECULAR, MULTI, CAMERA>(f0, bt, q, bounce);
}

// Shadow rays of one kind. Light targets: closest hit, then the emission of
// whatever was hit (camera.rs:2248-2261). Environment: any hit occludes
// (camera.rs:2284-2293; "is some hit" equals "is some closest hit").
template <bool COUNT, bool MULTI>
__global__ __attribute__((amdgpu_waves_per_eu(8))) void k_connect(Frame f0, Batch bt, Queues q, uint32_t bounce,
                                                                   uint32_t kind) {
  extern __shared__ uint2 smem[];
  Stack st = make_stack(q, smem);
  const float4* __restrict__ sd = q.s_d[kind];
  const float4* __restrict__ sw = q.s_w[kind];
  for (uint32_t s = blockIdx.x % slice_step(); s < kSlices; s += slice_step()) {
  const uint32_t count = q.ctr[ctr_at(bounce, kind == 0 ? S_SL_COUNT : S_SE_COUNT, s)];
  if (count == 0) continue;
  uint32_t* cursor = q.ctr + ctr_at(bounce, kind == 0 ? S_SL_WORK : S_SE_WORK, s);
  const uint32_t* __restrict__ list = q.shadow[kind] + size_t(s) * q.region;
  const uint32_t per_lane = items_per_lane(count);
  for (;;) {
    uint32_t base = wave_fetch(cursor, per_lane);
    if (base >= count) break;
    uint32_t nodes = 0, prims = 0;
    for (uint32_t k = 0; k < per_lane; ++k) {
      const uint32_t j = base + k * 64u + lane_id();
      const bool valid = j < count;
      const uint32_t p = valid ? list[j] : 0u;
      with_frame<MULTI>(f0, bt, valid, p, [&](const Frame& f, const Seg&) {
      Stats s{0, 0, 0, 0, 0, 0};
      const float4 o4 = q.ray_o[p], d4 = sd[p];  // the origin is the shading vertex
      f3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
      const uint32_t self = __float_as_uint(d4.w);  // the shading vertex's primitive
      float t;
      uint32_t prim;
      f3 add = mk(0.f, 0.f, 0.f);
      bool contributes = false;
      if (kind == 0) {
        if (trace<false, COUNT>(f.scene, o, d, o4.w, SHADOW_ACNE, INF, self, st, &t, &prim, nullptr, &s)) {
          f3 pt = o + t * d;
          uint32_t mid, pk;
          f3 outward = prim_outward_normal(f.scene.prims, prim, pt, o4.w, &mid, &pk);
          const GMat& m = f.scene.mats[mid];
          if (m.kind == GM_DIFFUSE_LIGHT && dot(d, outward) < 0.f) {
            f3 em = ld3(m.color);
            if (finite3(em)) { add = em; contributes = true; }
          }
        }
      } else {
        if (!trace<true, COUNT>(f.scene, o, d, o4.w, SHADOW_ACNE, INF, self, st, &t, &prim, nullptr, &s)) {
          add = mk(1.f, 1.f, 1.f);
          contributes = true;
        }
      }
      if (contributes) {  // the weight is read only for a ray that lands
        const float4 w4 = sw[p];
        add = mul(mk(w4.x, w4.y, w4.z), add);
        float4 L = q.radiance[p];
        q.radiance[p] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, 0.f);
      }
      if (COUNT) { nodes += node_fetches(s, f.scene.width); prims += s.candidates; }
      });
    }
    if (COUNT) {
      wave_count(count_slot(q, bounce, 2), nodes);
      wave_count(count_slot(q, bounce, 3), prims);
    }
  }
  }
}

// Adds each finished sample of the batch to its pixel in sample order,
// dropping non-finite samples (add_finite_sample camera.rs:2984-2991).
// Work item w is pixel w - pix0 of the segment holding w; a frame has at most
// one segment per batch, so each accumulator is updated by one thread.
template <bool MULTI>
__global__ void k_accumulate(Frame f0, Batch bt, Queues q) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t rounds = (bt.npix_total + stride - 1) / stride;  // wave-uniform trip count
  uint32_t dropped = 0;  // non-finite samples, counted (rare) rather than the accepted ones (every sample)
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t w = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    with_frame<MULTI, &Seg::pix0>(f0, bt, w < bt.npix_total, w, [&](const Frame& f, const Seg& sg) {
      const uint32_t px = w - sg.pix0;
      const uint32_t pixel = f.pixmap ? f.pixmap[px] : px;
      double4 a = f.acc[pixel];
      for (uint32_t s = 0; s < sg.nsamples; ++s) {
        float4 L = q.radiance[sg.path0 + size_t(s) * f.nact + px];
        if (isfinite(L.x) && isfinite(L.y) && isfinite(L.z)) {
          a.x += double(L.x); a.y += double(L.y); a.z += double(L.z); a.w += 1.0;
        } else {
          ++dropped;
        }
      }
      f.acc[pixel] = a;
    });
  }
  if (__ballot(dropped != 0u) == 0ull) return;  // the usual case: no atomic at all
  for (int off = 32; off > 0; off >>= 1) dropped += __shfl_down(dropped, off);
  if (lane_id() == 0) atomicAdd(q.tally, (unsigned long long)dropped);
}

// adaptive_error (camera.rs:2972-2982): the largest channel's standard error
// of the mean, infinite below two samples.
__device__ __forceinline__ double adaptive_error(double4 m2, uint32_t n) {
#pragma clang fp contract(off)
  if (n <= 1) return __builtin_inf();
  // (m2 / (n - 1)) / n, each LinearRgb / f64 a multiply by the reciprocal (colors.rs:185-190)
  const double c = 1.0 / double(n - 1), k = 1.0 / double(n);
  const double vx = k * (c * m2.x), vy = k * (c * m2.y), vz = k * (c * m2.z);
  return sqrt(fmax(fmax(fmax(vx, vy), vz), 0.0));
}

// render_world_pixel_adaptive (camera.rs:2937-2970) over the samples of one
// batch: each covered pixel that has not stopped takes them in sample order,
// skipping non-finite ones, and stops at the first accepted count >= min
// whose error is below the threshold; later samples of the batch are unused.
__global__ void k_adapt(Frame f, Batch bt, Queues q, Adapt ad) {
#pragma clang fp contract(off)
  const Seg& sg = bt.seg0;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t rounds = (f.nact + stride - 1) / stride;  // wave-uniform trip count
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t w = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= f.nact) continue;
    const uint32_t pixel = f.pixmap ? f.pixmap[w] : w;
    if (ad.done[pixel]) continue;
    double4 mean = ad.mean[pixel], m2 = ad.m2[pixel];
    uint32_t n = uint32_t(mean.w);
    bool stop = false;
    for (uint32_t s = 0; s < sg.nsamples && !stop; ++s) {
      const float4 L = q.radiance[sg.path0 + size_t(s) * f.nact + w];
      if (!(isfinite(L.x) && isfinite(L.y) && isfinite(L.z))) continue;
      ++n;
      const double inv = 1.0 / double(n);  // delta / count as LinearRgb / f64 (colors.rs:185-190)
      const double dx = double(L.x) - mean.x, dy = double(L.y) - mean.y, dz = double(L.z) - mean.z;
      mean.x += inv * dx;
      mean.y += inv * dy;
      mean.z += inv * dz;
      m2.x += dx * (double(L.x) - mean.x);
      m2.y += dy * (double(L.y) - mean.y);
      m2.z += dz * (double(L.z) - mean.z);
      stop = n >= ad.min_samples && adaptive_error(m2, n) < ad.threshold;
    }
    mean.w = double(n);
    ad.mean[pixel] = mean;
    ad.m2[pixel] = m2;
    if (stop) ad.done[pixel] = 1u;
  }
}

// Samples the adaptive update used, summed over the pixels of map (or
// 0..n-1): the accepted count each pixel's running mean carries.
__global__ void k_adapt_count(const double4* mean, const uint32_t* map, uint32_t n, unsigned long long* out) {
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    c += uint32_t(mean[map ? map[i] : i].w);
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
  if (lane_id() == 0 && c) atomicAdd(out, (unsigned long long)c);
}

__global__ void k_adapt_compact(const uint32_t* map, uint32_t n, const uint32_t* done, uint32_t* out,
                                uint32_t* count) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t rounds = (n + stride - 1) / stride;  // wave-uniform trip count
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t w = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t pixel = 0;
    bool keep = false;
    if (w < n) {
      pixel = map ? map[w] : w;
      keep = done[pixel] == 0u;
    }
    const unsigned long long bal = __ballot(keep);
    if (bal == 0ull) continue;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(count, uint32_t(__popcll(bal)));
    base = __shfl(base, 0);
    if (keep) out[base + uint32_t(__popcll(bal & ((1ull << lane_id()) - 1ull)))] = pixel;
  }
}

__global__ void k_adapt_finalize(const double4* mean, const uint32_t* map, uint32_t n, void* out,
                                 int out_is_double) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t px = map ? map[i] : i;
    const double4 m = mean[px];
    if (out_is_double) {
      double* o = static_cast<double*>(out) + 3 * size_t(px);
      o[0] = m.x; o[1] = m.y; o[2] = m.z;
    } else {
      float* o = static_cast<float*>(out) + 3 * size_t(px);
      o[0] = float(m.x); o[1] = float(m.y); o[2] = float(m.z);
    }
  }
}

// Clears accumulators with a kernel (no memset node, so the sequence can be
// captured in a graph and replayed).
__global__ void k_zero(double4* acc, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    acc[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// average_accepted_samples (camera.rs:2993-2999) for the n pixels of map
// (pixels 0..n-1 without one).
__global__ void k_finalize(const double4* acc, const uint32_t* map, uint32_t n, void* out, int out_is_double) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t px = map ? map[i] : i;
    double4 a = acc[px];
    double r = 0, g = 0, b = 0;
    if (a.w > 0.0) {  // LinearRgb / f64 multiplies by the reciprocal (colors.rs:185-190)
      const double inv = 1.0 / a.w;
      r = inv * a.x; g = inv * a.y; b = inv * a.z;
    }
    if (out_is_double) {
      double* o = static_cast<double*>(out) + 3 * size_t(px);
      o[0] = r; o[1] = g; o[2] = b;
    } else {
      float* o = static_cast<float*>(out) + 3 * size_t(px);
      o[0] = float(r); o[1] = float(g); o[2] = float(b);
    }
  }
}

// First-hit AOVs (render_denoising_aov_linear_pixel camera.rs:3064-3116):
// albedo = denoise_albedo, normal = 0.5 (n + 1); 0 on a miss.
__global__ void k_aov(Frame f, Queues q, float* albedo, float* normal) {
  extern __shared__ uint2 smem[];
  Stack st = make_stack(q, smem);
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < f.npix; px += gridDim.x * blockDim.x) {
    uint32_t x = px % f.cam.width, y = px / f.cam.width;
    uint64_t key = pixel_seed(f.cam.seed, x, y);
    f3 sa = mk(0.f, 0.f, 0.f), sn = mk(0.f, 0.f, 0.f);
    for (uint32_t s = 0; s < f.cam.spp; ++s) {
      Rng rng{stream_seed(key, s, 0)};
      f3 o, d;
      float time, t;
      uint32_t prim;
      camera_ray(f.cam, x, y, s, rng, &o, &d, &time);
      PrimHit ph{0.f, 0.f, 0.f};
      if (trace<false, false>(f.scene, o, d, time, SHADOW_ACNE, INF, kNoSelf, st, &t, &prim, &ph, nullptr)) {
        refine_hit(f.scene.prims, prim, o, d, time, SHADOW_ACNE, &ph);
        t = ph.t;
        f3 p = o + t * d;
        uint32_t mid, kind;
        f3 outward = prim_outward_normal(f.scene.prims, prim, p, time, &mid, &kind);
        const bool front = dot(d, outward) < 0.f;
        f3 n = front ? outward : -outward;
        n = shading_normal(f.scene.prims, f.scene.vnormals, prim, ph.u, ph.v, front, n);  // 0.5 (shading_normal + 1)
        const GMat& m = f.scene.mats[mid];
        sa = sa + (m.kind == GM_DIELECTRIC ? mk(1.f, 1.f, 1.f) : ld3(m.color));
        sn = sn + 0.5f * (n + mk(1.f, 1.f, 1.f));
      }
    }
    float inv = float(f.cam.spp);
    albedo[3 * size_t(px)] = sa.x / inv; albedo[3 * size_t(px) + 1] = sa.y / inv; albedo[3 * size_t(px) + 2] = sa.z / inv;
    normal[3 * size_t(px)] = sn.x / inv; normal[3 * size_t(px) + 1] = sn.y / inv; normal[3 * size_t(px) + 2] = sn.z / inv;
  }
}

// Closest-hit queries with the hit record of HitRecord::from_surface
// (object.rs:435-450) and optional BvhTraversalStats (bvh.rs:9-52). A ray
// and a hit are two float4 each (gt_ray, gt_hit); the statistics are summed
// per lane, then once per wave.
template <bool STATS>
__global__ void k_intersect(Frame f, Queues q, const float4* __restrict__ rays, size_t n, float tmin,
                            float4* __restrict__ hits, unsigned long long* stats) {
  extern __shared__ uint2 smem[];
  Stack st = make_stack(q, smem);
  unsigned long long sums[6] = {0, 0, 0, 0, 0, 0};
  uint32_t deepest = 0;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    float4 o4 = rays[2 * i], d4 = rays[2 * i + 1];
    f3 o = mk(o4.x, o4.y, o4.z), d = mk(d4.x, d4.y, d4.z);
    float t = 0.f;
    uint32_t prim = 0;
    PrimHit ph{0.f, 0.f, 0.f};
    Stats s{0, 0, 0, 0, 0, 0};
    bool hit = trace<false, STATS>(f.scene, o, d, o4.w, tmin, d4.w, kNoSelf, st, &t, &prim, &ph, &s);
    if (hit) {
      refine_hit(f.scene.prims, prim, o, d, o4.w, tmin, &ph);  // the chosen primitive's t, u, v in FP64
      t = ph.t;
      f3 p = o + t * d;
      uint32_t mid, kind;
      f3 outward = prim_outward_normal(f.scene.prims, prim, p, o4.w, &mid, &kind);
      bool front = dot(d, outward) < 0.f;
      f3 nn = front ? outward : -outward;
      float u = ph.u, v = ph.v;
      if (kind == GK_SPHERE || kind == GK_MOVING_SPHERE) {  // sphere_uv object.rs:1038-1042
        float theta = acosf(-outward.y);
        float phi = atan2f(-outward.z, outward.x) + PI;
        u = phi / (2.0f * PI);
        v = theta / PI;
      }
      hits[2 * i] = make_float4(t, u, v, __int_as_float(int32_t(f.scene.prim_ids[prim])));
      hits[2 * i + 1] = make_float4(nn.x, nn.y, nn.z, __uint_as_float(front ? 1u : 0u));
    } else {
      hits[2 * i] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
      hits[2 * i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (STATS) {
      sums[0] += 1;
      sums[1] += s.node_tests;
      sums[2] += s.node_hits;
      sums[3] += s.leaf_visits;
      sums[4] += s.candidates;
      sums[5] += s.leaf_hits;
      deepest = max(deepest, s.max_stack);
    }
  }
  if (STATS) {  // every lane of the wave is back here: reduce, then one atomic per word
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      unsigned long long v = sums[k];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
      if (lane_id() == 0 && v) atomicAdd(&stats[k], v);
    }
    for (int off = 32; off > 0; off >>= 1) deepest = max(deepest, uint32_t(__shfl_down(deepest, off)));
    if (lane_id() == 0) atomicMax(&stats[6], (unsigned long long)deepest);
  }
}

__global__ void k_zero_words(unsigned long long* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

}  // namespace

uint32_t occupancy_blocks(int kernel, uint32_t block, size_t lds) {
  int nb = 0;
  hipError_t e = hipSuccess;
  if (kernel == 0) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_extend<false, false, false>, int(block), lds);
  else if (kernel == 1) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_shade<false, false>, int(block), 0);
  else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_connect<false, false>, int(block), lds);
  if (e != hipSuccess || nb <= 0) return 1;
  return uint32_t(nb);
}

// The one place that decides which bounce's extend generates camera rays;
// the shade launch of that bounce takes the same template choice, so the
// camera paths' unstored throughput and the generated path ids stay paired.
static inline bool camera_bounce(uint32_t bounce) { return bounce == 0; }

void reset_counters(const Launch& l, const Queues& q, uint32_t depth, uint32_t total) {
  const uint32_t n = (depth + 1) * kSlots * kSlices;
  hipLaunchKernelGGL(k_reset, dim3((n + 255) / 256), dim3(256), 0, l.stream, q.ctr, n, q.tally,
                     uint32_t(tally_words(depth)), total);
}
void zero_acc(const Launch& l, double4* acc, size_t n) {
  hipLaunchKernelGGL(k_zero, dim3(l.grid), dim3(l.block), 0, l.stream, acc, n);
}
void extend(const Launch& l, const Work& w, const Queues& q, uint32_t bounce, bool count) {
  const size_t lds = stack_lds_bytes(l.block, q.stack_cap);
#define GT_EXTEND(C, M, K) \
  hipLaunchKernelGGL((k_extend<C, M, K>), dim3(l.grid), dim3(l.block), lds, l.stream, w.f, w.bt, q, bounce)
#define GT_EXTEND_CM(K)                                                  \
  if (w.multi) {                                                         \
    if (count) GT_EXTEND(true, true, K); else GT_EXTEND(false, true, K); \
  } else {                                                               \
    if (count) GT_EXTEND(true, false, K); else GT_EXTEND(false, false, K); \
  }
  if (camera_bounce(bounce)) {
    GT_EXTEND_CM(true)
  } else {
    GT_EXTEND_CM(false)
  }
#undef GT_EXTEND_CM
#undef GT_EXTEND
}
void shade(const Launch& l, const Work& w, const Queues& q, uint32_t bounce) {
#define GT_SHADE(M, K) \
  hipLaunchKernelGGL((k_shade<M, K>), dim3(l.grid), dim3(l.block), 0, l.stream, w.f, w.bt, q, bounce)
  if (camera_bounce(bounce)) {
    if (w.multi) GT_SHADE(true, true); else GT_SHADE(false, true);
  } else {
    if (w.multi) GT_SHADE(true, false); else GT_SHADE(false, false);
  }
#undef GT_SHADE
}
void connect(const Launch& l, const Work& w, const Queues& q, uint32_t bounce, uint32_t kind, bool count) {
  const size_t lds = stack_lds_bytes(l.block, q.stack_cap);
#define GT_CONNECT(C, M) \
  hipLaunchKernelGGL((k_connect<C, M>), dim3(l.grid), dim3(l.block), lds, l.stream, w.f, w.bt, q, bounce, kind)
  if (w.multi) {
    if (count) GT_CONNECT(true, true); else GT_CONNECT(false, true);
  } else {
    if (count) GT_CONNECT(true, false); else GT_CONNECT(false, false);
  }
#undef GT_CONNECT
}
// One thread per pixel (not the persistent grid): each pixel's loop over its
// samples is a chain of dependent loads, so more threads hide more of it.
static inline unsigned pixel_grid(uint32_t items, uint32_t block) {
  return std::max(1u, (items + block - 1) / block);
}

void accumulate(const Launch& l, const Work& w, const Queues& q) {
  const unsigned g = pixel_grid(w.bt.npix_total, l.block);
  if (w.multi) hipLaunchKernelGGL(k_accumulate<true>, dim3(g), dim3(l.block), 0, l.stream, w.f, w.bt, q);
  else hipLaunchKernelGGL(k_accumulate<false>, dim3(g), dim3(l.block), 0, l.stream, w.f, w.bt, q);
}
void adapt(const Launch& l, const Work& w, const Queues& q, const Adapt& a) {
  hipLaunchKernelGGL(k_adapt, dim3(pixel_grid(w.f.nact, l.block)), dim3(l.block), 0, l.stream, w.f, w.bt, q, a);
}

void adapt_compact(const Launch& l, const uint32_t* map, uint32_t n, const uint32_t* done, uint32_t* out,
                   uint32_t* count) {
  hipLaunchKernelGGL(k_adapt_compact, dim3(l.grid), dim3(l.block), 0, l.stream, map, n, done, out, count);
}

void adapt_count(const Launch& l, const double4* mean, const uint32_t* map, uint32_t n, unsigned long long* out) {
  hipLaunchKernelGGL(k_adapt_count, dim3(l.grid), dim3(l.block), 0, l.stream, mean, map, n, out);
}

void adapt_finalize(const Launch& l, const double4* mean, const uint32_t* map, uint32_t n, void* out,
                    int out_is_double) {
  hipLaunchKernelGGL(k_adapt_finalize, dim3(l.grid), dim3(l.block), 0, l.stream, mean, map, n, out, out_is_double);
}

void finalize(const Launch& l, const double4* acc, uint32_t npix, void* out, int out_is_double) {
  hipLaunchKernelGGL(k_finalize, dim3(l.grid), dim3(l.block), 0, l.stream, acc, nullptr, npix, out, out_is_double);
}

void finalize_map(const Launch& l, const double4* acc, const uint32_t* map, uint32_t n, void* out,
                  int out_is_double) {
  hipLaunchKernelGGL(k_finalize, dim3(l.grid), dim3(l.block), 0, l.stream, acc, map, n, out, out_is_double);
}
void aov(const Launch& l, const Frame& f, const Queues& q, float* albedo, float* normal) {
  hipLaunchKernelGGL(k_aov, dim3(l.grid), dim3(l.block), stack_lds_bytes(l.block, q.stack_cap), l.stream, f, q,
                     albedo, normal);
}
void intersect(const Launch& l, const Frame& f, const Queues& q, const float4* rays, size_t n, float tmin,
               float4* hits, unsigned long long* stats) {
  if (stats)
    hipLaunchKernelGGL(k_intersect<true>, dim3(l.grid), dim3(l.block), stack_lds_bytes(l.block, q.stack_cap),
                       l.stream, f, q, rays, n, tmin, hits, stats);
  else
    hipLaunchKernelGGL(k_intersect<false>, dim3(l.grid), dim3(l.block), stack_lds_bytes(l.block, q.stack_cap),
                       l.stream, f, q, rays, n, tmin, hits, stats);
}
void zero_words(hipStream_t stream, unsigned long long* p, uint32_t n) {
  hipLaunchKernelGGL(k_zero_words, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
uint2* g_scratch_ptr[kMaxDevices];
}  // namespace

// The array's name in host code is a stand-in for the symbol, not a device
// address; each device's copy is reached only through the address looked up
// here with that device current.
hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? static_cast<uint2*>(p) : nullptr;
    g_scratch_done[device] = true;
  }
  return g_scratch_err[device];
}
hipError_t noalloc_scratch(int device, uint2** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[device]) return hipErrorNotInitialized;
  *out = g_scratch_ptr[device];
  return g_scratch_err[device];
}

}  // namespace gtk
