// kernels.h — host-side launch interface of the wavefront kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"

namespace gtk {

// Shading categories: extend bins every hit by the BSDF it needs, so each
// shade kernel runs one BSDF without divergence across material types.
// Misses are shaded by extend itself; CAT_MISS keeps its counter slot but has
// no list and no kernel.
enum { CAT_MISS = 0, CAT_EMIT = 1, CAT_LAMBERT = 2, CAT_GGX = 3, CAT_SPECULAR = 4, kNumCategories = 5 };

// Work is split into kSlices slices, one per XCD: the dispatcher hands
// workgroups to the eight XCDs round-robin, so block b works on slice
// b % kSlices. Each slice owns region s of every queue (entries
// [s * region, (s + 1) * region)) and its own counters, each on its own
// 128-byte line: same-address device atomics are served one at a time at
// the memory side, so one address per slice lets eight proceed at once.
// A path never leaves the slice it started in.
constexpr uint32_t kSlices = 8;
constexpr uint32_t kLineWords = 32;
enum : uint32_t {
  S_QCOUNT = 0,   // extension rays queued for bounce b
  S_BASE,         // bounce 0: first path id of the slice
  S_EXT_WORK,     // work cursor of extend(b)
  S_SL_COUNT,     // light shadow rays queued at bounce b
  S_SE_COUNT,     // environment shadow rays queued at bounce b
  S_SL_WORK,
  S_SE_WORK,
  S_CAT_COUNT,    // hits binned by shading category (kNumCategories slots)
  S_CAT_WORK = S_CAT_COUNT + kNumCategories,
  kSlots = S_CAT_WORK + kNumCategories
};
constexpr uint32_t kBounceWords = kSlots * kSlices * kLineWords;
__host__ __device__ constexpr uint32_t ctr_at(uint32_t bounce, uint32_t slot, uint32_t slice) {
  return ((bounce * kSlots + slot) * kSlices + slice) * kLineWords;
}
// 64-bit tallies: [0] the batch's non-finite (dropped) samples (not counted
// in adaptive rounds); with counting,
// [1 + 4 b + k]: k = 0 extension node fetches, 1 extension prim tests,
// 2 shadow node fetches, 3 shadow prim tests at bounce b.
inline size_t tally_words(uint32_t depth) { return 1 + 4 * size_t(depth); }

// Path state lives at the path's index (path = sample-in-batch * npix +
// pixel) and is updated in place; every queue is a list of 4-byte path ids.
// Each wave reserves its share of a queue with one atomic per work chunk.
struct Queues {
  float4* ray_o;    // origin, time
  float4* ray_d;    // direction, flags (bit 31: last event was a scattering event)
  float4* thr;      // throughput, miss-radiance weight
  float2* hits;     // t, reordered primitive index (0xffffffff = miss)
  float4* s_d[2];   // shadow rays: [0] light targets, [1] environment: direction, and the bits of the
                    // primitive the ray leaves (leaf order); origin ray_o
  float4* s_w[2];   // weight (x, y, z, unused)
  float4* radiance; // per-path radiance L
  uint32_t* active[2];            // paths to extend at bounce b, parity b & 1 (bounce 0: all paths)
  uint32_t* cat[kNumCategories];  // paths hit this bounce, per shading category
  uint32_t* shadow[2];            // paths with a light / environment shadow ray this bounce
  uint32_t region;                // queue entries per slice
  uint32_t* ctr;                  // per-bounce counters, ctr_at()
  unsigned long long* tally;
  uint2* gstack;    // traversal-stack overflow, gstride entries per level
  uint32_t gstride;
  uint32_t stack_cap;  // LDS entries per lane
};

// Everything one render of one scene reads: scene, lights, environment,
// camera, and the FP64 per-pixel accumulator it writes.
struct Frame {
  gtd::DScene scene;
  gtd::DLights lights;
  gtd::DEnv env;
  GCamera cam;
  uint32_t nee;            // direct lighting by next-event estimation
  uint32_t has_lights;
  float light_weight;      // mixture weight in continuation mode
  uint32_t npix;
  double4* acc;            // per-pixel sums (rgb, accepted count)
  // Pixels this render covers: paths of a sample are pixel indices
  // 0..nact-1, mapped through pixmap when it is set (adaptive rounds render
  // only the pixels still sampling); nact = npix and no map otherwise.
  const uint32_t* pixmap;
  uint32_t nact;
};

// Per-pixel state of adaptive sampling (render_world_pixel_adaptive,
// camera.rs:2937-2970), kept across rounds.
struct Adapt {
  double4* mean;        // running mean (rgb) and accepted sample count (w)
  double4* m2;          // sums of squared deviations (rgb)
  uint32_t* done;       // 1 once the pixel has stopped
  uint32_t min_samples;
  double threshold;
};

// A batch of paths is a list of segments: segment i renders samples
// [sample0, sample0 + nsamples) of every pixel of frame `frame` as paths
// [path0, path0 + nsamples * npix), sample-major. Its pixels are accumulate
// work items [pix0, pix0 + npix). Segments are sorted by path0 and pix0, and
// a frame has at most one segment per batch.
struct Seg {
  uint32_t path0, frame, sample0, nsamples, pix0;
  uint32_t pad[3];
};
struct Batch {
  const Frame* frames;  // batched renders: frame table in device memory
  const Seg* segs;      // batched renders: segment table in device memory
  uint32_t nseg;
  uint32_t total;       // paths in the batch
  uint32_t npix_total;  // accumulate work items
  Seg seg0;             // single-frame renders: the one segment (frame 0)
};

struct Launch {
  uint32_t grid, block;
  hipStream_t stream;
};

// What the wavefront kernels of one batch read: for a single-frame render the
// frame itself (multi = false, passed by value); for a batched render the
// frame and segment tables in bt (multi = true).
struct Work {
  Frame f;
  Batch bt;
  bool multi;
};

void reset_counters(const Launch& l, const Queues& q, uint32_t depth, uint32_t total);
void zero_acc(const Launch& l, double4* acc, size_t n);
void extend(const Launch& l, const Work& w, const Queues& q, uint32_t bounce, bool count);
// One bounce of shading: one launch drains the category queues in turn.
void shade(const Launch& l, const Work& w, const Queues& q, uint32_t bounce);
void connect(const Launch& l, const Work& w, const Queues& q, uint32_t bounce, uint32_t kind, bool count);
void accumulate(const Launch& l, const Work& w, const Queues& q);
// Adaptive sampling (single-frame renders): the Welford update of every
// covered pixel over this batch's samples, in sample order, up to its stop.
void adapt(const Launch& l, const Work& w, const Queues& q, const Adapt& a);
// The covered pixels (map, or 0..n-1 without one) that have not stopped,
// appended to out in any order; *count must be 0 on entry.
void adapt_compact(const Launch& l, const uint32_t* map, uint32_t n, const uint32_t* done, uint32_t* out,
                   uint32_t* count);
// *out += the accepted samples of the n pixels of map (0..n-1 without one).
void adapt_count(const Launch& l, const double4* mean, const uint32_t* map, uint32_t n, unsigned long long* out);
// out[pixel] = the running mean (float or double rgb) for the n pixels of
// map (pixels 0..n-1 without one).
void adapt_finalize(const Launch& l, const double4* mean, const uint32_t* map, uint32_t n, void* out,
                    int out_is_double);
void finalize(const Launch& l, const double4* acc, uint32_t npix, void* out, int out_is_double);
// finalize for the n pixels of map only.
void finalize_map(const Launch& l, const double4* acc, const uint32_t* map, uint32_t n, void* out,
                  int out_is_double);
void aov(const Launch& l, const Frame& f, const Queues& q, float* albedo, float* normal);

// Closest-hit queries: rays are gt_ray records (origin, time; direction,
// t_max), hits gt_hit records (t, u, v, prim_id; normal, front_face), both
// two float4 per entry. stats, when set, receives the BvhTraversalStats sums
// (6 words) and the largest stack depth (1 word), added to what it holds.
void intersect(const Launch& l, const Frame& f, const Queues& q, const float4* rays, size_t n, float tmin,
               float4* hits, unsigned long long* stats);
// Sets n 64-bit words to zero with a kernel (no memset node under capture).
void zero_words(hipStream_t stream, unsigned long long* p, uint32_t n);

// Scratch of the no-allocation closest-hit entry point: the traversal-stack
// overflow lives in a file-scope device array of kNoAllocStackEntries
// entries, one copy per GPU. noalloc_scratch_init looks its address up on
// the current device (once per device; call it with `device` current);
// noalloc_scratch returns that address afterwards.
constexpr size_t kNoAllocStackEntries = size_t(1) << 21;  // 16 MiB of 8-byte entries
hipError_t noalloc_scratch_init(int device);
hipError_t noalloc_scratch(int device, uint2** out);

// Resident blocks per CU of a kernel: 0 extend, 1 shade (GGX), 2 connect.
uint32_t occupancy_blocks(int kernel, uint32_t block, size_t lds_bytes);

// Bytes of LDS per block for the traversal stack.
inline size_t stack_lds_bytes(uint32_t block, uint32_t cap) { return size_t(block) * cap * sizeof(uint2); }

}  // namespace gtk
