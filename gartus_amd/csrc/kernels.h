// kernels.h — host-side launch interface of the wavefront kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"

namespace gtk {

// Counters per bounce b live at ctr[b * kCtrStride + ...].
constexpr int kCtrStride = 32;
enum {
  C_QCOUNT = 0,      // extension rays queued for bounce b
  C_EXT_WORK = 1,    // work cursor of extend(b)
  C_SHADE_WORK = 2,  // work cursor of shade(b)
  C_SL_COUNT = 3,    // light shadow rays queued at bounce b
  C_SE_COUNT = 4,    // environment shadow rays queued at bounce b
  C_SL_WORK = 5,
  C_SE_WORK = 6,
  C_ACCEPTED = 7,    // finite samples accumulated in this batch (bounce 0 block)
  // with counting: u64 totals at 64-bit word (b * kCtrStride + C_COUNT64) / 2 + k,
  // k = 0 extension node fetches, 1 extension prim tests, 2 shadow node
  // fetches, 3 shadow prim tests
  C_COUNT64 = 8,
  C_CAT_COUNT = 16,  // hits binned by shading category (kNumCategories slots)
  C_CAT_WORK = 24,   // work cursors of the category shade kernels
};


// Shading categories: extend bins every hit by the BSDF it needs, so each
// shade kernel runs one BSDF without divergence across material types.
enum { CAT_MISS = 0, CAT_EMIT = 1, CAT_LAMBERT = 2, CAT_GGX = 3, CAT_SPECULAR = 4, kNumCategories = 5 };

// Path state lives at the path's index (path = sample-in-batch * npix +
// pixel) and is updated in place; every queue is a list of 4-byte path ids.
// Each wave reserves its share of a queue with one atomic per work chunk.
struct Queues {
  float4* ray_o;    // origin, time
  float4* ray_d;    // direction, flags (bit 31: last event was a scattering event)
  float4* thr;      // throughput, miss-radiance weight
  float2* hits;     // t, reordered primitive index (0xffffffff = miss)
  float4* s_o[2];   // shadow rays: [0] light targets, [1] environment
  float4* s_d[2];
  float4* s_w[2];   // contribution weight
  float4* radiance; // per-path radiance L
  uint32_t* active[2];            // paths to extend at bounce b, parity b & 1 (bounce 0: all paths)
  uint32_t* cat[kNumCategories];  // paths hit this bounce, per shading category
  uint32_t* shadow[2];            // paths with a light / environment shadow ray this bounce
  uint32_t* ctr;
  uint2* gstack;    // traversal-stack overflow, gstride entries per level
  uint32_t gstride;
  uint32_t stack_cap;  // LDS entries per lane
};

struct Frame {
  gtd::DScene scene;
  gtd::DLights lights;
  gtd::DEnv env;
  GCamera cam;
  uint32_t nee;            // direct lighting by next-event estimation
  uint32_t has_lights;
  float light_weight;      // mixture weight in continuation mode
  uint32_t npix;
};

struct Launch {
  uint32_t grid, block;
  hipStream_t stream;
};

void reset_counters(const Launch& l, uint32_t* ctr, uint32_t n);
void raygen(const Launch& l, const Frame& f, const Queues& q, uint32_t sample0, uint32_t nsamples);
void extend(const Launch& l, const Frame& f, const Queues& q, uint32_t bounce, bool count);
// One bounce of shading: one kernel per category (empty categories exit at once).
void shade(const Launch& l, const Frame& f, const Queues& q, uint32_t bounce, uint32_t sample0, uint32_t last);
void connect(const Launch& l, const Frame& f, const Queues& q, uint32_t bounce, uint32_t kind, bool count);
void accumulate(const Launch& l, const Frame& f, const Queues& q, double4* acc, uint32_t nsamples);
void finalize(const Launch& l, const double4* acc, uint32_t npix, void* out, int out_is_double);
void aov(const Launch& l, const Frame& f, const Queues& q, float* albedo, float* normal);

struct IntersectOut {
  float* t;
  float* u;
  float* v;
  int32_t* prim;
  float* normal;
  uint32_t* front;
  unsigned long long* stats;  // 6 sums + max stack (7 words), or null
};
void intersect(const Launch& l, const Frame& f, const Queues& q, const float4* rays, size_t n, float tmin,
               IntersectOut out);

// Resident blocks per CU of a kernel: 0 extend, 1 shade (GGX), 2 connect.
uint32_t occupancy_blocks(int kernel, uint32_t block, size_t lds_bytes);

// Bytes of LDS per block for the traversal stack.
inline size_t stack_lds_bytes(uint32_t block, uint32_t cap) { return size_t(block) * cap * sizeof(uint2); }

}  // namespace gtk
