// scene_build.h — host-side scene compile: primitive bounds, the SAH binary
// BVH (the reference's split rules, bvh.rs:233-535), and packing into the
// GPU layout of layout.h.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "layout.h"

namespace gt {

struct HostPrim {
  int kind;          // GK_*
  double q[9];       // sphere: c, r | moving: c0, c1, r | triangle: p0, p1, p2 | quad: corner, u, v
  uint32_t material;
  uint32_t mat_kind;   // GM_* of the material, packed into the record's meta word
  uint32_t smooth;     // triangles: 0 flat, 1 + i for a mesh triangle with vertex normals i (3 entries each)
};

struct HostMat {
  int kind;          // GM_*
  double color[3];
  double param;
};

struct Box {
  double lo[3], hi[3];
};

// One node of the reference's FlatBvh (pre-order numbering, bvh.rs:149-165).
struct BinNode {
  Box bounds;
  bool leaf;
  uint64_t a, b;     // leaf: first, count ; internal: left, right
};

struct BinaryBvh {
  std::vector<BinNode> nodes;
  std::vector<uint64_t> indices;   // leaf ranges index this permutation
  uint64_t depth = 0;
};

Box prim_bounds(const HostPrim& p);

// Builds the tree exactly as FlatBvh::build does (same buckets, costs,
// partition order and midpoint fallback). leaf_size 0 is clamped to 1.
bool build_binary_bvh(const std::vector<HostPrim>& prims, size_t leaf_size, BinaryBvh* out);

struct PackedScene {
  std::vector<GNode2> nodes2;
  std::vector<GNode4> nodes4;
  std::vector<GPrim> prims;          // reordered by the BVH permutation
  std::vector<uint32_t> prim_ids;    // reordered index -> original primitive id
  uint32_t root_ref2 = 0, root_ref4 = 0;
  float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
  uint64_t depth2 = 0, depth4 = 0;
};

void pack_scene(const std::vector<HostPrim>& prims, const BinaryBvh& bvh, PackedScene* out);
GPrim pack_prim(const HostPrim& p, bool last);

}  // namespace gt
