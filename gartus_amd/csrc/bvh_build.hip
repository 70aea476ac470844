// bvh_build.hip — the scene compile on the device (gfx950).
//
// A linear BVH after Karras, "Maximizing Parallelism in the Construction of
// BVHs, Octrees, and k-d Trees" (HPG 2012):
//   1. pack every primitive into its 64-byte record and its FP64 box (the
//      host compile's formulas, so records and boxes are bit-identical),
//      rounded outward to FP32; reduce the centroid bounds;
//   2. 63-bit Morton codes of the centroids (21 bits per axis of one cubic
//      grid over the centroid bounds), radix-sorted with the primitive index
//      as payload;
//   3. one internal node per adjacent pair of sorted keys: its range and
//      split from the longest common prefixes (equal keys break ties by
//      index), children and parent links;
//   4. boxes bottom up: every leaf walks towards the root and the second of
//      two arrivals at a node unites both children, counts the emitted nodes
//      of its subtree, and continues (an acquire-release counter per node, no
//      waiting);
//   5. subtrees of at most leaf_size primitives become leaves (a contiguous
//      range of sorted records, the last flagged); every larger internal node
//      writes its GNode2 record and its GNode4 record (its grandchildren,
//      left-left first) at its depth-first preorder position among the
//      emitted nodes, so a node's nearer-left child usually follows it in
//      memory (the emitted-node counts of step 4 give each node its position
//      from its root path alone);
//   6. records reordered by the sort, the tree depth from leaf-to-root walks.
// The tree differs from the reference's SAH tree (bvh.rs:233-288); closest
// hits do not depend on the tree, traversal counts do.
#include "bvh_build.h"

#include <hipcub/hipcub.hpp>

#include <cmath>

namespace gbvh {
namespace {

constexpr uint32_t kNone = 0xffffffffu;

struct Box {
  double lo[3], hi[3];
};

__device__ inline Box point_box(double x, double y, double z) { return Box{{x, y, z}, {x, y, z}}; }
__device__ inline Box unite(const Box& a, const Box& b) {
  Box r;
  for (int i = 0; i < 3; ++i) { r.lo[i] = fmin(a.lo[i], b.lo[i]); r.hi[i] = fmax(a.hi[i], b.hi[i]); }
  return r;
}
__device__ inline Box padded(Box b, double e) {  // polygon_matrix.rs:84-106
  e = fabs(e);
  const double half = 0.5 * e;
  for (int i = 0; i < 3; ++i)
    if (fabs(b.hi[i] - b.lo[i]) < e) { b.lo[i] -= half; b.hi[i] += half; }
  return b;
}

// The host compile's prim_bounds (object.rs:1015-1036, geometry.rs:197-203,
// 339-362), in FP64 with contraction off.
__device__ Box prim_box(const gt::HostPrim& p) {
#pragma clang fp contract(off)
  const double* q = p.q;
  switch (p.kind) {
    case GK_SPHERE: {
      const double r = q[3];
      return Box{{q[0] - r, q[1] - r, q[2] - r}, {q[0] + r, q[1] + r, q[2] + r}};
    }
    case GK_MOVING_SPHERE: {
      const double r = q[6];
      Box a{{q[0] - r, q[1] - r, q[2] - r}, {q[0] + r, q[1] + r, q[2] + r}};
      Box b{{q[3] - r, q[4] - r, q[5] - r}, {q[3] + r, q[4] + r, q[5] + r}};
      Box x, y;
      for (int i = 0; i < 3; ++i) {
        x.lo[i] = fmin(a.lo[i], a.hi[i]); x.hi[i] = fmax(a.lo[i], a.hi[i]);
        y.lo[i] = fmin(b.lo[i], b.hi[i]); y.hi[i] = fmax(b.lo[i], b.hi[i]);
      }
      return unite(x, y);
    }
    case GK_TRIANGLE: {
      Box b = unite(unite(point_box(q[0], q[1], q[2]), point_box(q[3], q[4], q[5])), point_box(q[6], q[7], q[8]));
      return padded(b, 1e-10);
    }
    default: {
      Box b = point_box(q[0], q[1], q[2]);
      b = unite(b, point_box(q[0] + q[3], q[1] + q[4], q[2] + q[5]));
      b = unite(b, point_box(q[0] + q[6], q[1] + q[7], q[2] + q[8]));
      b = unite(b, point_box(q[0] + q[3] + q[6], q[1] + q[4] + q[7], q[2] + q[5] + q[8]));
      return padded(b, 1e-4);
    }
  }
}

// The host compile's pack_prim (layout.h record formats), last bit clear.
__device__ GPrim pack(const gt::HostPrim& p) {
#pragma clang fp contract(off)
  GPrim g;
  for (int i = 0; i < 4; ++i) g.p0[i] = g.p1[i] = g.p2[i] = g.p3[i] = 0.0f;
  const double* q = p.q;
  const uint32_t meta = (p.material & GT_MATERIAL_MASK) | ((p.mat_kind & 7u) << GT_MATKIND_SHIFT) |
                        (uint32_t(p.kind) << GT_KIND_SHIFT);
  g.p0[3] = __uint_as_float(meta);
  constexpr double eps = 2.220446049250313e-16;
  switch (p.kind) {
    case GK_SPHERE:
      for (int i = 0; i < 3; ++i) g.p0[i] = float(q[i]);
      g.p1[0] = float(fmax(q[3], 0.0));
      break;
    case GK_MOVING_SPHERE:
      for (int i = 0; i < 3; ++i) { g.p0[i] = float(q[i]); g.p1[i] = float(q[3 + i]); }
      g.p1[3] = float(fmax(q[6], 0.0));
      break;
    case GK_TRIANGLE: {
      double e1[3], e2[3];
      for (int i = 0; i < 3; ++i) { e1[i] = q[3 + i] - q[i]; e2[i] = q[6 + i] - q[i]; }
      const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
      const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int i = 0; i < 3; ++i) {
        g.p0[i] = float(q[i]); g.p1[i] = float(e1[i]); g.p2[i] = float(e2[i]);
        g.p3[i] = len < eps ? 0.0f : float(n[i] / len);
      }
      g.p1[3] = __uint_as_float(p.smooth);
      break;
    }
    default: {
      const double* u = q + 3;
      const double* v = q + 6;
      const double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
      const double nls = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
      const double len = sqrt(nls);
      double nn[3] = {0.0, 0.0, 0.0};
      if (!(len < eps)) for (int i = 0; i < 3; ++i) nn[i] = n[i] / len;
      const double d = nn[0] * q[0] + nn[1] * q[1] + nn[2] * q[2];
      for (int i = 0; i < 3; ++i) {
        g.p0[i] = float(q[i]); g.p1[i] = float(u[i]); g.p2[i] = float(v[i]); g.p3[i] = float(nn[i]);
      }
      g.p1[3] = nls <= eps ? 0.0f : float(1.0 / len);
      g.p2[3] = float(nls);
      g.p3[3] = float(d);
      break;
    }
  }
  return g;
}

__device__ inline float round_down(double v) {
  float f = float(v);
  if (double(f) > v) f = nextafterf(f, -INFINITY);
  return f;
}
__device__ inline float round_up(double v) {
  float f = float(v);
  if (double(f) < v) f = nextafterf(f, INFINITY);
  return f;
}

// Order-preserving map of doubles to unsigned 64-bit integers (for atomics).
__device__ inline unsigned long long dkey(double d) {
  const unsigned long long b = __double_as_longlong(d);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double dval(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double(b);
}

__global__ __launch_bounds__(256) void k_prep(const gt::HostPrim* __restrict__ in, uint32_t n,
                                              GPrim* __restrict__ packed, float4* __restrict__ plo,
                                              float4* __restrict__ phi, double* __restrict__ cen,
                                              unsigned long long* __restrict__ cbounds) {
#pragma clang fp contract(off)
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const gt::HostPrim p = in[i];
    packed[i] = pack(p);
    const Box b = prim_box(p);
    plo[i] = make_float4(round_down(b.lo[0]), round_down(b.lo[1]), round_down(b.lo[2]), 0.0f);
    phi[i] = make_float4(round_up(b.hi[0]), round_up(b.hi[1]), round_up(b.hi[2]), 0.0f);
    for (int a = 0; a < 3; ++a) {
      const double c = 0.5 * (b.lo[a] + b.hi[a]);  // Bounds3::centroid
      cen[3 * size_t(i) + a] = c;
      mn[a] = fmin(mn[a], c);
      mx[a] = fmax(mx[a], c);
    }
  }
  for (int a = 0; a < 3; ++a)
    for (int off = 32; off > 0; off >>= 1) {
      mn[a] = fmin(mn[a], __shfl_xor(mn[a], off));
      mx[a] = fmax(mx[a], __shfl_xor(mx[a], off));
    }
  if ((threadIdx.x & 63) == 0) {
    for (int a = 0; a < 3; ++a) {
      if (mn[a] <= mx[a]) {
        atomicMin(cbounds + a, dkey(mn[a]));
        atomicMax(cbounds + 3 + a, dkey(mx[a]));
      }
    }
  }
}

__device__ inline unsigned long long spread21(unsigned long long x) {  // 21 bits -> every third of 63
  x &= 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ __launch_bounds__(256) void k_morton(const double* __restrict__ cen, uint32_t n,
                                                const unsigned long long* __restrict__ cbounds,
                                                unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals) {
#pragma clang fp contract(off)
  // one cubic grid over the centroid bounds: a flat axis gets only the bits its extent needs,
  // so a height field is ordered by its two long axes first
  double lo[3], ext = 0.0;
  for (int a = 0; a < 3; ++a) {
    lo[a] = dval(cbounds[a]);
    ext = fmax(ext, dval(cbounds[3 + a]) - lo[a]);
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    unsigned long long code = 0;
    for (int a = 0; a < 3; ++a) {
      double t = ext > 0.0 ? (cen[3 * size_t(i) + a] - lo[a]) / ext : 0.0;
      t = fmin(fmax(t * 2097152.0, 0.0), 2097151.0);
      code |= spread21((unsigned long long)t) << (2 - a);
    }
    keys[i] = code;
    vals[i] = i;
  }
}

// Longest common prefix of sorted keys i and j, with the index breaking ties
// between equal keys; -1 outside [0, n).
__device__ inline int delta(const unsigned long long* __restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const unsigned long long a = k[i], b = k[j];
  if (a == b) return 64 + __clz(uint32_t(i) ^ uint32_t(j));
  return __clzll(a ^ b);
}

__global__ __launch_bounds__(256) void k_hierarchy(const unsigned long long* __restrict__ keys, int n,
                                                   uint2* __restrict__ child, uint2* __restrict__ range,
                                                   uint32_t* __restrict__ parent_leaf,
                                                   uint32_t* __restrict__ parent_int) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
    const int d = delta(keys, n, i, i + 1) > delta(keys, n, i, i - 1) ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int first = min(i, j), last = max(i, j);
    const bool lleaf = first == gamma, rleaf = last == gamma + 1;
    child[i] = make_uint2(uint32_t(gamma) | (lleaf ? GT_LEAF_BIT : 0u), uint32_t(gamma + 1) | (rleaf ? GT_LEAF_BIT : 0u));
    range[i] = make_uint2(uint32_t(first), uint32_t(last));
    if (lleaf) parent_leaf[gamma] = uint32_t(i);
    else parent_int[gamma] = uint32_t(i);
    if (rleaf) parent_leaf[gamma + 1] = uint32_t(i);
    else parent_int[gamma + 1] = uint32_t(i);
  }
}

__device__ inline void child_box(uint32_t c, const uint32_t* order, const float4* plo, const float4* phi,
                                 const float4* nlo, const float4* nhi, float4* lo, float4* hi) {
  if (c & GT_LEAF_BIT) {
    const uint32_t p = order[c & ~GT_LEAF_BIT];
    *lo = plo[p];
    *hi = phi[p];
  } else {
    *lo = nlo[c];
    *hi = nhi[c];
  }
}

__global__ __launch_bounds__(256) void k_bounds(uint32_t n, const uint32_t* __restrict__ order,
                                                const float4* __restrict__ plo, const float4* __restrict__ phi,
                                                const uint32_t* __restrict__ parent_leaf,
                                                const uint32_t* __restrict__ parent_int, const uint2* __restrict__ child,
                                                const uint2* __restrict__ range, uint32_t leaf_size,
                                                float4* nlo, float4* nhi, uint32_t* icount, uint32_t* arrivals) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    uint32_t node = parent_leaf[k];
    while (node != kNone) {
      // the first of the two children to arrive stops; the second sees both boxes
      const uint32_t seen = __hip_atomic_fetch_add(arrivals + node, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (seen == 0u) break;
      const uint2 c = child[node];
      float4 al, ah, bl, bh;
      child_box(c.x, order, plo, phi, nlo, nhi, &al, &ah);
      child_box(c.y, order, plo, phi, nlo, nhi, &bl, &bh);
      nlo[node] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
      nhi[node] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
      // emitted internal nodes in the subtree: this one unless it falls inside a leaf
      const uint2 rg = range[node];
      icount[node] = uint32_t(rg.y - rg.x + 1u > leaf_size) + ((c.x & GT_LEAF_BIT) ? 0u : icount[c.x]) +
                     ((c.y & GT_LEAF_BIT) ? 0u : icount[c.y]);
      node = parent_int[node];
    }
  }
}

// Preorder position of every emitted internal node among the emitted nodes:
// from the root, each step to a left child adds 1, each step to a right child
// adds 1 plus the emitted nodes of the left sibling's subtree.
__global__ __launch_bounds__(256) void k_preorder(uint32_t n, uint32_t leaf_size, const uint2* __restrict__ range,
                                                  const uint2* __restrict__ child,
                                                  const uint32_t* __restrict__ parent_int,
                                                  const uint32_t* __restrict__ icount, uint32_t* __restrict__ pre) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n; i += gridDim.x * blockDim.x) {
    const uint2 rg = range[i];
    if (rg.y - rg.x + 1u <= leaf_size) continue;  // inside a leaf: never referenced
    uint32_t pos = 0;
    for (uint32_t c = i, p = parent_int[i]; p != kNone; c = p, p = parent_int[p]) {
      const uint2 pc = child[p];
      pos += 1u;
      if (pc.y == c) pos += (pc.x & GT_LEAF_BIT) ? 0u : icount[pc.x];
    }
    pre[i] = pos;
  }
}

__global__ __launch_bounds__(256) void k_depth(uint32_t n, const uint32_t* __restrict__ parent_leaf,
                                               const uint32_t* __restrict__ parent_int, uint32_t* depth) {
  uint32_t best = 0;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    uint32_t d = 1;
    for (uint32_t node = parent_leaf[k]; node != kNone; node = parent_int[node]) ++d;
    best = max(best, d);
  }
  for (int off = 32; off > 0; off >>= 1) best = max(best, uint32_t(__shfl_xor(int(best), off)));
  if ((threadIdx.x & 63) == 0) atomicMax(depth, best);
}

// A child reference as the traversal sees it: a leaf (a range of at most
// leaf_size records, the first named) or an internal node index; also its
// box and the last record of its range.
struct Ref {
  uint32_t ref, last, node;  // node: the internal node's build index (for its children)
  float4 lo, hi;
  bool leaf;
};
__device__ inline Ref make_ref(uint32_t c, uint32_t leaf_size, const uint2* range, const uint32_t* order,
                               const float4* plo, const float4* phi, const float4* nlo, const float4* nhi,
                               const uint32_t* pre) {
  Ref r;
  child_box(c, order, plo, phi, nlo, nhi, &r.lo, &r.hi);
  if (c & GT_LEAF_BIT) {
    r.leaf = true;
    r.ref = c;
    r.node = 0;
    r.last = c & ~GT_LEAF_BIT;
    return r;
  }
  const uint2 rg = range[c];
  r.leaf = rg.y - rg.x + 1u <= leaf_size;
  r.node = c;
  r.ref = r.leaf ? (GT_LEAF_BIT | rg.x) : pre[c];
  r.last = rg.y;
  return r;
}

__global__ __launch_bounds__(256) void k_emit(uint32_t n, uint32_t leaf_size, const uint2* __restrict__ child,
                                              const uint2* __restrict__ range, const uint32_t* __restrict__ order,
                                              const float4* __restrict__ plo, const float4* __restrict__ phi,
                                              const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                              const uint32_t* __restrict__ pre, GNode2* __restrict__ nodes2,
                                              GNode4* __restrict__ nodes4, uint8_t* __restrict__ last,
                                              unsigned long long* counts) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n == 1 || n <= leaf_size)) last[n - 1] = 1;  // the root is a leaf
  uint32_t emitted = 0, leaves = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n; i += gridDim.x * blockDim.x) {
    const uint2 rg = range[i];
    if (rg.y - rg.x + 1u <= leaf_size) continue;  // inside a leaf
    ++emitted;
    const uint2 c = child[i];
    const Ref L = make_ref(c.x, leaf_size, range, order, plo, phi, nlo, nhi, pre);
    const Ref R = make_ref(c.y, leaf_size, range, order, plo, phi, nlo, nhi, pre);
    if (L.leaf) { last[L.last] = 1; ++leaves; }
    if (R.leaf) { last[R.last] = 1; ++leaves; }
    GNode2 g;
    g.lx[0] = L.lo.x; g.lx[1] = L.hi.x; g.ly[0] = L.lo.y; g.ly[1] = L.hi.y; g.lz[0] = L.lo.z; g.lz[1] = L.hi.z;
    g.rx[0] = R.lo.x; g.rx[1] = R.hi.x; g.ry[0] = R.lo.y; g.ry[1] = R.hi.y; g.rz[0] = R.lo.z; g.rz[1] = R.hi.z;
    g.left = L.ref;
    g.right = R.ref;
    g.pad0 = g.pad1 = 0;
    const uint32_t at = pre[i];
    nodes2[at] = g;
    // the four-wide record: grandchildren through internal children, left subtree first
    Ref kids[4];
    int nk = 0;
    const Ref* sides[2] = {&L, &R};
    for (int side = 0; side < 2; ++side) {
      const Ref* r = sides[side];
      if (r->leaf) {
        kids[nk++] = *r;
      } else {
        const uint2 cc = child[r->node];
        kids[nk++] = make_ref(cc.x, leaf_size, range, order, plo, phi, nlo, nhi, pre);
        kids[nk++] = make_ref(cc.y, leaf_size, range, order, plo, phi, nlo, nhi, pre);
      }
    }
    GNode4 w;
    for (int k = 0; k < 4; ++k) {
      w.pad[k] = 0;
      if (k >= nk) {
        w.child[k] = GT_EMPTY_CHILD;
        w.lox[k] = w.loy[k] = w.loz[k] = 1.0f;
        w.hix[k] = w.hiy[k] = w.hiz[k] = -1.0f;
        continue;
      }
      w.lox[k] = kids[k].lo.x; w.loy[k] = kids[k].lo.y; w.loz[k] = kids[k].lo.z;
      w.hix[k] = kids[k].hi.x; w.hiy[k] = kids[k].hi.y; w.hiz[k] = kids[k].hi.z;
      w.child[k] = kids[k].ref;
    }
    nodes4[at] = w;
  }
  for (int off = 32; off > 0; off >>= 1) {
    emitted += __shfl_xor(emitted, off);
    leaves += __shfl_xor(leaves, off);
  }
  if ((threadIdx.x & 63) == 0 && (emitted | leaves)) {
    atomicAdd(counts, (unsigned long long)emitted);
    atomicAdd(counts + 1, (unsigned long long)leaves);
  }
}

__global__ __launch_bounds__(256) void k_reorder(uint32_t n, const uint32_t* __restrict__ order,
                                                 const GPrim* __restrict__ packed, const uint8_t* __restrict__ last,
                                                 GPrim* __restrict__ prims, uint32_t* __restrict__ ids) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t id = order[k];
    GPrim g = packed[id];
    if (last[k]) g.p0[3] = __uint_as_float(__float_as_uint(g.p0[3]) | GT_LAST_BIT);
    prims[k] = g;
    ids[k] = id;
  }
}

__global__ void k_init(unsigned long long* cbounds, uint32_t* depth, unsigned long long* counts) {
  if (threadIdx.x < 3) {
    cbounds[threadIdx.x] = ~0ull;    // min slots
    cbounds[3 + threadIdx.x] = 0ull;  // max slots
  }
  if (threadIdx.x == 0) { *depth = 0; counts[0] = counts[1] = 0; }
}

// Scratch layout, each part 256-byte aligned.
struct Layout {
  size_t staged, packed, plo, phi, cen, keys, keys2, vals, vals2, child, range, pleaf, pint, nlo, nhi, arrivals,
      icount, pre, last, cbounds, depth, counts, sort_tmp, sort_bytes, total;
};
inline size_t up(size_t x) { return (x + 255) & ~size_t(255); }

Layout layout(uint32_t n) {
  Layout L{};
  const size_t m = n > 1 ? n - 1 : 1;
  size_t o = 0;
  auto take = [&o](size_t bytes) { size_t at = o; o += up(bytes); return at; };
  L.staged = take(sizeof(gt::HostPrim) * size_t(n));
  L.packed = take(sizeof(GPrim) * size_t(n));
  L.plo = take(sizeof(float4) * size_t(n));
  L.phi = take(sizeof(float4) * size_t(n));
  L.cen = take(sizeof(double) * 3 * size_t(n));
  L.keys = take(8 * size_t(n));
  L.keys2 = take(8 * size_t(n));
  L.vals = take(4 * size_t(n));
  L.vals2 = take(4 * size_t(n));
  L.child = take(sizeof(uint2) * m);
  L.range = take(sizeof(uint2) * m);
  L.pleaf = take(4 * size_t(n));
  L.pint = take(4 * m);
  L.nlo = take(sizeof(float4) * m);
  L.nhi = take(sizeof(float4) * m);
  L.arrivals = take(4 * m);
  L.icount = take(4 * m);
  L.pre = take(4 * m);
  L.last = take(size_t(n));
  L.cbounds = take(6 * 8);
  L.depth = take(4);
  L.counts = take(16);
  size_t sort_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, static_cast<unsigned long long*>(nullptr),
                                           static_cast<unsigned long long*>(nullptr), static_cast<uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), int(n), 0, 63);
  L.sort_bytes = sort_bytes;
  L.sort_tmp = take(sort_bytes);
  L.total = o;
  return L;
}

unsigned grid_for(size_t n) { return unsigned(std::min<size_t>((n + 255) / 256, 65536)); }

}  // namespace

size_t scratch_bytes(uint32_t n) { return layout(n).total; }

hipError_t build(hipStream_t s, const gt::HostPrim* host_prims, uint32_t n, uint32_t leaf_size, void* scratch,
                 GPrim* prims, uint32_t* ids, GNode2* nodes2, GNode4* nodes4, Result* out) {
  if (n == 0 || leaf_size == 0) return hipErrorInvalidValue;
  const Layout L = layout(n);
  char* base = static_cast<char*>(scratch);
  auto at = [base](size_t off) { return static_cast<void*>(base + off); };
  GPrim* packed = static_cast<GPrim*>(at(L.packed));
  float4* plo = static_cast<float4*>(at(L.plo));
  float4* phi = static_cast<float4*>(at(L.phi));
  double* cen = static_cast<double*>(at(L.cen));
  auto* keys = static_cast<unsigned long long*>(at(L.keys));
  auto* keys2 = static_cast<unsigned long long*>(at(L.keys2));
  uint32_t* vals = static_cast<uint32_t*>(at(L.vals));
  uint32_t* order = static_cast<uint32_t*>(at(L.vals2));
  uint2* child = static_cast<uint2*>(at(L.child));
  uint2* range = static_cast<uint2*>(at(L.range));
  uint32_t* pleaf = static_cast<uint32_t*>(at(L.pleaf));
  uint32_t* pint = static_cast<uint32_t*>(at(L.pint));
  float4* nlo = static_cast<float4*>(at(L.nlo));
  float4* nhi = static_cast<float4*>(at(L.nhi));
  uint32_t* arrivals = static_cast<uint32_t*>(at(L.arrivals));
  uint32_t* icount = static_cast<uint32_t*>(at(L.icount));
  uint32_t* pre = static_cast<uint32_t*>(at(L.pre));
  uint8_t* last = static_cast<uint8_t*>(at(L.last));
  auto* cbounds = static_cast<unsigned long long*>(at(L.cbounds));
  uint32_t* depth = static_cast<uint32_t*>(at(L.depth));
  auto* counts = static_cast<unsigned long long*>(at(L.counts));
  const size_t m = n > 1 ? n - 1 : 1;

  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  auto done = [&](hipError_t r) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return r;
  };
  if (e != hipSuccess) return done(e);
  gt::HostPrim* staged = static_cast<gt::HostPrim*>(at(L.staged));
  if ((e = hipMemcpyAsync(staged, host_prims, sizeof(gt::HostPrim) * size_t(n), hipMemcpyHostToDevice, s)) !=
      hipSuccess)
    return done(e);
  if ((e = hipEventRecord(e0, s)) != hipSuccess) return done(e);
  hipLaunchKernelGGL(k_init, dim3(1), dim3(64), 0, s, cbounds, depth, counts);
  if ((e = hipMemsetAsync(pleaf, 0xff, 4 * size_t(n), s)) != hipSuccess) return done(e);
  if ((e = hipMemsetAsync(pint, 0xff, 4 * m, s)) != hipSuccess) return done(e);
  if ((e = hipMemsetAsync(arrivals, 0, 4 * m, s)) != hipSuccess) return done(e);
  if ((e = hipMemsetAsync(last, 0, size_t(n), s)) != hipSuccess) return done(e);
  hipLaunchKernelGGL(k_prep, dim3(grid_for(n)), dim3(256), 0, s, staged, n, packed, plo, phi, cen, cbounds);
  hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(256), 0, s, cen, n, cbounds, keys, vals);
  size_t sort_bytes = L.sort_bytes;
  if ((e = hipcub::DeviceRadixSort::SortPairs(at(L.sort_tmp), sort_bytes, keys, keys2, vals, order, int(n), 0, 63,
                                              s)) != hipSuccess)
    return done(e);
  if (n > 1) {
    hipLaunchKernelGGL(k_hierarchy, dim3(grid_for(n - 1)), dim3(256), 0, s, keys2, int(n), child, range, pleaf, pint);
    hipLaunchKernelGGL(k_bounds, dim3(grid_for(n)), dim3(256), 0, s, n, order, plo, phi, pleaf, pint, child, range,
                       leaf_size, nlo, nhi, icount, arrivals);
    hipLaunchKernelGGL(k_preorder, dim3(grid_for(n - 1)), dim3(256), 0, s, n, leaf_size, range, child, pint, icount,
                       pre);
    hipLaunchKernelGGL(k_depth, dim3(grid_for(n)), dim3(256), 0, s, n, pleaf, pint, depth);
  }
  hipLaunchKernelGGL(k_emit, dim3(std::max(1u, grid_for(n))), dim3(256), 0, s, n, leaf_size, child, range, order, plo,
                     phi, nlo, nhi, pre, nodes2, nodes4, last, counts);
  hipLaunchKernelGGL(k_reorder, dim3(grid_for(n)), dim3(256), 0, s, n, order, packed, last, prims, ids);
  if ((e = hipGetLastError()) != hipSuccess) return done(e);
  if ((e = hipEventRecord(e1, s)) != hipSuccess) return done(e);
  // root box, depth, counts
  float4 rb[2];
  uint32_t dep = 0;
  unsigned long long cnt[2] = {0, 0};
  const float4* rsrc_lo = n > 1 ? nlo : plo;
  const float4* rsrc_hi = n > 1 ? nhi : phi;
  if ((e = hipMemcpyAsync(&rb[0], rsrc_lo, sizeof(float4), hipMemcpyDeviceToHost, s)) != hipSuccess) return done(e);
  if ((e = hipMemcpyAsync(&rb[1], rsrc_hi, sizeof(float4), hipMemcpyDeviceToHost, s)) != hipSuccess) return done(e);
  if ((e = hipMemcpyAsync(&dep, depth, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return done(e);
  if ((e = hipMemcpyAsync(cnt, counts, 16, hipMemcpyDeviceToHost, s)) != hipSuccess) return done(e);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return done(e);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  out->ms = ms;
  out->root_ref = (n == 1 || n <= leaf_size) ? GT_LEAF_BIT : 0u;
  out->root_lo[0] = rb[0].x; out->root_lo[1] = rb[0].y; out->root_lo[2] = rb[0].z;
  out->root_hi[0] = rb[1].x; out->root_hi[1] = rb[1].y; out->root_hi[2] = rb[1].z;
  out->depth = n > 1 ? dep : 1;
  out->internal_nodes = cnt[0];
  out->leaves = out->root_ref == GT_LEAF_BIT ? 1 : cnt[1];
  return done(hipSuccess);
}

}  // namespace gbvh
