// bvh_build.h — scene compile on the device: a Karras LBVH over the
// primitives, emitted in the traversal's node layout (layout.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "layout.h"
#include "scene_build.h"

namespace gbvh {

struct Result {
  uint32_t root_ref;          // node 0, or GT_LEAF_BIT | 0 when the whole scene is one leaf
  float root_lo[3], root_hi[3];
  uint64_t depth;             // binary levels, root = 1 (bounds the traversal stack)
  uint64_t internal_nodes;    // internal nodes emitted (those inside a leaf are not)
  uint64_t leaves;            // leaves emitted
  double ms;                  // device time of the build, HIP events
};

// Device memory the build needs besides its outputs, for n primitives.
size_t scratch_bytes(uint32_t n);

// Builds the tree for n primitives (host records, copied to the device) on
// stream s. Outputs, each sized by the caller: prims[n] (reordered records,
// last-in-leaf bits set), ids[n] (reordered index -> primitive id),
// nodes2[max(n - 1, 1)], nodes4[max(n - 1, 1)]. Returns a hipError_t.
hipError_t build(hipStream_t s, const gt::HostPrim* host_prims, uint32_t n, uint32_t leaf_size, void* scratch,
                 GPrim* prims, uint32_t* ids, GNode2* nodes2, GNode4* nodes4, Result* out);

}  // namespace gbvh
