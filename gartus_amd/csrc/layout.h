// layout.h — data layout of a compiled scene in HBM, shared by the host
// builder and the HIP kernels.
//
// BVH: the reference's binary SAH tree (bvh.rs:233-288) stored as 64-byte
// nodes that hold both children's boxes ("child boxes in the parent"), so one
// node fetch feeds two slab tests. Only internal nodes get a record; a child
// reference with GT_LEAF_BIT set names the first primitive of a leaf, whose
// records are contiguous (the reference's index permutation) and whose last
// record carries GT_LAST_BIT.
//
// Primitive records are 64 bytes, one cache-line-aligned record per
// primitive, type-tagged in the first word:
//   sphere         p0 = (center, meta)  p1 = (r, 0, 0, 0)
//   moving sphere  p0 = (c0, meta)      p1 = (c1, r)
//   triangle       p0 = (v0, meta)      p1 = (v1-v0, s) p2 = (v2-v0, 0) p3 = (unit normal, 0)
//                  s: the bits of 0 for a flat triangle, of 1 + i for a mesh triangle
//                  whose vertex normals are entries 3i..3i+2 of the scene's normal array
//   quad           p0 = (corner, meta)  p1 = (u, 1/|u x v|) p2 = (v, 0) p3 = (unit normal, plane d)
// meta = material id (25 bits) | material kind << 25 (3 bits) | primitive
// kind << 28 (3 bits) | last-in-leaf << 31. The material kind lets the
// traversal kernel sort hits by BSDF without another load.
#pragma once
#include <stdint.h>

#define GT_LEAF_BIT 0x80000000u
#define GT_LAST_BIT 0x80000000u
#define GT_KIND_SHIFT 28
#define GT_KIND_MASK 0x7u
#define GT_MATKIND_SHIFT 25
#define GT_MATERIAL_MASK 0x01ffffffu

enum { GK_SPHERE = 0, GK_MOVING_SPHERE = 1, GK_TRIANGLE = 2, GK_QUAD = 3 };
enum { GM_LAMBERTIAN = 0, GM_DIFFUSE_LIGHT = 1, GM_GGX = 2, GM_METAL = 3, GM_DIELECTRIC = 4, GM_ISOTROPIC = 5,
       GM_HENYEY_GREENSTEIN = 6 };

struct alignas(16) GNode2 {
  float lx[2], ly[2];  // left child x/y slabs (lo, hi)
  float rx[2], ry[2];  // right child
  float lz[2], rz[2];
  uint32_t left, right, pad0, pad1;
};
static_assert(sizeof(GNode2) == 64, "BVH2 node must be 64 bytes");

// Four-wide node built by collapsing two levels of the binary tree:
// children in the reference's traversal order (left subtree first).
struct alignas(16) GNode4 {
  float lox[4], hix[4];
  float loy[4], hiy[4];
  float loz[4], hiz[4];
  uint32_t child[4];   // GT_LEAF_BIT | first prim, node index, or GT_EMPTY_CHILD
  uint32_t pad[4];
};
static_assert(sizeof(GNode4) == 128, "BVH4 node must be 128 bytes");
#define GT_EMPTY_CHILD 0xffffffffu

struct alignas(16) GPrim {
  float p0[4], p1[4], p2[4], p3[4];
};
static_assert(sizeof(GPrim) == 64, "primitive record must be 64 bytes");

struct alignas(16) GMat {
  float color[3];
  uint32_t kind;
  float param;     // ggx roughness (clamped), metal fuzz (clamped), dielectric ior, henyey-greenstein g
  float f0;        // ggx: specular_luminance(color)
  float pad[2];
};

// Light target (WeightedSamplingTargetList entry, scene.rs:458-609).
struct alignas(16) GLight {
  float a[4];      // sphere: center, r | moving: c0, r | triangle: p0 | quad: corner ; a[3] = radius
  float b[4];      // moving: c1 | triangle: e1 | quad: u ; b[3] = weight / total
  float c[4];      // triangle: e2 | quad: v ; c[3] = kind (as float)
  float n[4];      // triangle/quad unit normal, n[3] = area
  float w[4];      // quad: w = n / |n|^2 ; w[3] = plane d
};

// Camera frame computed on the host in f64 (camera.rs:1068-1130).
struct GCamera {
  float center[3], pixel00[3], du[3], dv[3], disk_u[3], disk_v[3];
  float shutter_start, shutter_end;
  uint32_t width, height, spp, grid, mode;  // mode = pixel_sample_mode
  uint32_t defocus;
  uint64_t seed;
  uint32_t max_depth;
  int32_t rr_min_depth;
  uint32_t background_kind;
  float background[9];
  uint32_t flags;  // GT_FLAG_* of gt_camera
};
