// Test shim: exposes the product's host BVH builder (gartus_amd/csrc/
// scene_build.cpp) through a C ABI so a CPU test can compare its tree with
// the oracle's. Built by tests/test_scene_build.py with g++.
#include <cstring>
#include <vector>

#include "../gartus_amd/csrc/scene_build.h"

static gt::BinaryBvh g_bvh;
static gt::PackedScene g_packed;

extern "C" uint64_t shim_build(size_t n, const int* kinds, const double* q9, uint32_t leaf_size) {
  std::vector<gt::HostPrim> prims(n);
  for (size_t i = 0; i < n; ++i) {
    prims[i].kind = kinds[i];
    std::memcpy(prims[i].q, q9 + 9 * i, 9 * sizeof(double));
    prims[i].material = 0;
    prims[i].mat_kind = 0;
  }
  gt::build_binary_bvh(prims, leaf_size, &g_bvh);
  gt::pack_scene(prims, g_bvh, &g_packed);
  return g_bvh.nodes.size();
}

extern "C" void shim_dump(double* bounds, uint8_t* leaf, uint64_t* a, uint64_t* b, uint64_t* indices,
                          uint64_t* depth, uint64_t* packed_nodes2, uint64_t* packed_nodes4) {
  for (size_t i = 0; i < g_bvh.nodes.size(); ++i) {
    const gt::BinNode& nd = g_bvh.nodes[i];
    for (int k = 0; k < 3; ++k) { bounds[6 * i + k] = nd.bounds.lo[k]; bounds[6 * i + 3 + k] = nd.bounds.hi[k]; }
    leaf[i] = nd.leaf; a[i] = nd.a; b[i] = nd.b;
  }
  for (size_t i = 0; i < g_bvh.indices.size(); ++i) indices[i] = g_bvh.indices[i];
  *depth = g_bvh.depth;
  *packed_nodes2 = g_packed.nodes2.size();
  *packed_nodes4 = g_packed.nodes4.size();
}
