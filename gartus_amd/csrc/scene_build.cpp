// scene_build.cpp — host scene compile (see scene_build.h).
//
// The binary tree must be the reference's tree node for node: its traversal
// order decides which primitive wins an exact tie in t, and the parity tests
// require the GPU's closest-hit ids to equal the reference's. The build is
// iterative (an explicit range stack, left subtree first) so deep trees cannot
// overflow the host stack; numbering is the reference's pre-order.
#include "scene_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace gt {

namespace {
constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr int kBuckets = 12;  // bvh.rs:4

inline double dmin(double a, double b) { return std::fmin(a, b); }
inline double dmax(double a, double b) { return std::fmax(a, b); }

inline Box box_of_point(const double* p) { return Box{{p[0], p[1], p[2]}, {p[0], p[1], p[2]}}; }
inline Box unite(const Box& a, const Box& b) {
  Box r;
  for (int i = 0; i < 3; ++i) { r.lo[i] = dmin(a.lo[i], b.lo[i]); r.hi[i] = dmax(a.hi[i], b.hi[i]); }
  return r;
}
inline Box padded(Box b, double e) {  // polygon_matrix.rs:84-106
  e = std::fabs(e);
  double half = 0.5 * e;
  for (int i = 0; i < 3; ++i)
    if (std::fabs(b.hi[i] - b.lo[i]) < e) { b.lo[i] -= half; b.hi[i] += half; }
  return b;
}
inline double surface_area(const Box& b) {  // polygon_matrix.rs:124-129
  double dx = dmax(b.hi[0] - b.lo[0], 0.0), dy = dmax(b.hi[1] - b.lo[1], 0.0), dz = dmax(b.hi[2] - b.lo[2], 0.0);
  return 2.0 * (dx * dy + dx * dz + dy * dz);
}
inline int largest_axis(const Box& b) {  // polygon_matrix.rs:133-144
  double dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
  if (dx >= dy && dx >= dz) return 0;
  return dy >= dz ? 1 : 2;
}
inline int bucket_of(double offset) {  // bvh.rs:499-502 (`as usize` saturates)
  double f = offset * double(kBuckets);
  size_t b = (f != f || f <= 0.0) ? 0 : (f >= 1e18 ? SIZE_MAX : size_t(f));
  return int(std::min<size_t>(b, kBuckets - 1));
}

struct Builder {
  const std::vector<Box>& pb;
  const std::vector<double>& cx;  // centroids, 3 per primitive
  size_t leaf_size;
  std::vector<uint64_t>& idx;

  double cen(uint64_t k, int a) const { return cx[3 * k + a]; }

  size_t midpoint(uint64_t* v, size_t n, int axis) const {  // bvh.rs:508-522
    std::stable_sort(v, v + n, [&](uint64_t l, uint64_t r) {
      double la = cen(l, axis), ra = cen(r, axis);
      if (la < ra) return true;
      if (la > ra) return false;
      return l < r;
    });
    return n / 2;
  }

  // split_bvh_indices (bvh.rs:430-496); 0 = leaf.
  size_t split(uint64_t* v, size_t n) const {
    if (n <= leaf_size) return 0;
    Box cb = box_of_point(&cx[3 * v[0]]);
    for (size_t i = 1; i < n; ++i) cb = unite(cb, box_of_point(&cx[3 * v[i]]));
    int axis = largest_axis(cb);
    double cmin = cb.lo[axis];
    double extent = cb.hi[axis] - cmin;
    if (extent <= kEps) return midpoint(v, n, axis);
    size_t cnt[kBuckets] = {0};
    Box bb[kBuckets];
    for (size_t i = 0; i < n; ++i) {
      uint64_t k = v[i];
      int b = bucket_of((cen(k, axis) - cmin) / extent);
      bb[b] = cnt[b] ? unite(bb[b], pb[k]) : pb[k];
      cnt[b]++;
    }
    // Prefix and suffix merges give the same unions as merge_buckets' folds
    // (min/max are exact, so association does not change the result).
    size_t lc[kBuckets], rc[kBuckets];
    Box lb[kBuckets], rb[kBuckets];
    bool lh[kBuckets], rh[kBuckets];
    size_t c = 0; bool h = false; Box acc{};
    for (int i = 0; i < kBuckets; ++i) {
      c += cnt[i];
      if (cnt[i]) { acc = h ? unite(acc, bb[i]) : bb[i]; h = true; }
      lc[i] = c; lb[i] = acc; lh[i] = h;
    }
    c = 0; h = false;
    for (int i = kBuckets - 1; i >= 0; --i) {
      c += cnt[i];
      if (cnt[i]) { acc = h ? unite(acc, bb[i]) : bb[i]; h = true; }
      rc[i] = c; rb[i] = acc; rh[i] = h;
    }
    int best = 0;
    double best_cost = std::numeric_limits<double>::infinity();
    for (int s = 0; s < kBuckets - 1; ++s) {
      if (lc[s] == 0 || rc[s + 1] == 0 || !lh[s] || !rh[s + 1]) continue;
      double cost = surface_area(lb[s]) * double(lc[s]) + surface_area(rb[s + 1]) * double(rc[s + 1]);
      if (cost < best_cost) { best_cost = cost; best = s; }
    }
    if (!std::isfinite(best_cost)) return midpoint(v, n, axis);
    size_t left = 0;
    for (size_t next = 0; next < n; ++next) {
      if (bucket_of((cen(v[next], axis) - cmin) / extent) <= best) { std::swap(v[left], v[next]); ++left; }
    }
    if (left == 0 || left == n) return midpoint(v, n, axis);
    return left;
  }
};
}  // namespace

Box prim_bounds(const HostPrim& p) {
  const double* q = p.q;
  switch (p.kind) {
    case GK_SPHERE: {  // object.rs:1015-1022
      double r = q[3];
      return Box{{q[0] - r, q[1] - r, q[2] - r}, {q[0] + r, q[1] + r, q[2] + r}};
    }
    case GK_MOVING_SPHERE: {  // object.rs:1024-1036
      double r = q[6];
      Box a{{q[0] - r, q[1] - r, q[2] - r}, {q[0] + r, q[1] + r, q[2] + r}};
      Box b{{q[3] - r, q[4] - r, q[5] - r}, {q[3] + r, q[4] + r, q[5] + r}};
      Box x, y;
      for (int i = 0; i < 3; ++i) {
        x.lo[i] = dmin(a.lo[i], a.hi[i]); x.hi[i] = dmax(a.lo[i], a.hi[i]);
        y.lo[i] = dmin(b.lo[i], b.hi[i]); y.hi[i] = dmax(b.lo[i], b.hi[i]);
      }
      return unite(x, y);
    }
    case GK_TRIANGLE: {  // geometry.rs:197-203
      Box b = box_of_point(q);
      b = unite(b, box_of_point(q + 3));
      b = unite(b, box_of_point(q + 6));
      return padded(b, 1e-10);
    }
    default: {  // geometry.rs:339-362: corner, corner+u, corner+v, corner+u+v
      double p1[3], p2[3], p3[3];
      for (int i = 0; i < 3; ++i) { p1[i] = q[i] + q[3 + i]; p2[i] = q[i] + q[6 + i]; p3[i] = q[i] + q[3 + i] + q[6 + i]; }
      Box b = box_of_point(q);
      b = unite(b, box_of_point(p1));
      b = unite(b, box_of_point(p2));
      b = unite(b, box_of_point(p3));
      return padded(b, 1e-4);
    }
  }
}

bool build_binary_bvh(const std::vector<HostPrim>& prims, size_t leaf_size, BinaryBvh* out) {
  out->nodes.clear();
  out->indices.clear();
  out->depth = 0;
  if (prims.empty()) return false;
  if (leaf_size == 0) leaf_size = 1;
  const size_t n = prims.size();
  std::vector<Box> pb(n);
  std::vector<double> cx(3 * n);
  for (size_t i = 0; i < n; ++i) {
    pb[i] = prim_bounds(prims[i]);
    for (int a = 0; a < 3; ++a) cx[3 * i + a] = 0.5 * (pb[i].lo[a] + pb[i].hi[a]);  // Bounds3::centroid
  }
  out->indices.resize(n);
  for (size_t i = 0; i < n; ++i) out->indices[i] = i;
  out->nodes.reserve(2 * n);
  Builder b{pb, cx, leaf_size, out->indices};

  struct Task { uint64_t first, count, parent, depth; int side; };  // side: -1 root, 0 left, 1 right
  std::vector<Task> stack;
  stack.push_back({0, n, 0, 1, -1});
  while (!stack.empty()) {
    Task t = stack.back();
    stack.pop_back();
    Box bounds = pb[out->indices[t.first]];
    for (uint64_t i = 1; i < t.count; ++i) bounds = unite(bounds, pb[out->indices[t.first + i]]);
    uint64_t ni = out->nodes.size();
    out->nodes.push_back(BinNode{bounds, true, t.first, t.count});
    if (t.side == 0) out->nodes[t.parent].a = ni;
    else if (t.side == 1) out->nodes[t.parent].b = ni;
    out->depth = std::max(out->depth, t.depth);
    size_t lc = b.split(out->indices.data() + t.first, t.count);
    if (lc != 0) {
      out->nodes[ni].leaf = false;
      // pre-order: the left subtree is numbered before the right one
      stack.push_back({t.first + lc, t.count - lc, ni, t.depth + 1, 1});
      stack.push_back({t.first, lc, ni, t.depth + 1, 0});
    }
  }
  return true;
}

namespace {
inline float round_down(double v) {
  float f = float(v);
  if (double(f) > v) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}
inline float round_up(double v) {
  float f = float(v);
  if (double(f) < v) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}
inline uint32_t as_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float from_bits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
}  // namespace

GPrim pack_prim(const HostPrim& p, bool last) {
  GPrim g;
  std::memset(&g, 0, sizeof g);
  const double* q = p.q;
  uint32_t meta = (p.material & GT_MATERIAL_MASK) | ((p.mat_kind & 7u) << GT_MATKIND_SHIFT) |
                  (uint32_t(p.kind) << GT_KIND_SHIFT) | (last ? GT_LAST_BIT : 0u);
  g.p0[3] = from_bits(meta);
  switch (p.kind) {
    case GK_SPHERE:
      for (int i = 0; i < 3; ++i) g.p0[i] = float(q[i]);
      g.p1[0] = float(std::fmax(q[3], 0.0));  // geometry.rs:445-450
      break;
    case GK_MOVING_SPHERE:
      for (int i = 0; i < 3; ++i) { g.p0[i] = float(q[i]); g.p1[i] = float(q[3 + i]); }
      g.p1[3] = float(std::fmax(q[6], 0.0));
      break;
    case GK_TRIANGLE: {
      double e1[3], e2[3];
      for (int i = 0; i < 3; ++i) { e1[i] = q[3 + i] - q[i]; e2[i] = q[6 + i] - q[i]; }
      double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int i = 0; i < 3; ++i) {
        g.p0[i] = float(q[i]); g.p1[i] = float(e1[i]); g.p2[i] = float(e2[i]);
        g.p3[i] = len < std::numeric_limits<double>::epsilon() ? 0.0f : float(n[i] / len);
      }
      g.p1[3] = from_bits(p.smooth);
      break;
    }
    default: {  // QuadGeometry::new (geometry.rs:293-312)
      const double* u = q + 3;
      const double* v = q + 6;
      double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
      double nls = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
      double len = std::sqrt(nls);
      double nn[3] = {0, 0, 0};
      if (!(len < std::numeric_limits<double>::epsilon())) for (int i = 0; i < 3; ++i) nn[i] = n[i] / len;
      double d = nn[0] * q[0] + nn[1] * q[1] + nn[2] * q[2];
      for (int i = 0; i < 3; ++i) {
        g.p0[i] = float(q[i]); g.p1[i] = float(u[i]); g.p2[i] = float(v[i]); g.p3[i] = float(nn[i]);
      }
      g.p1[3] = nls <= std::numeric_limits<double>::epsilon() ? 0.0f : float(1.0 / len);
      g.p2[3] = float(nls);
      g.p3[3] = float(d);
      break;
    }
  }
  return g;
}

void pack_scene(const std::vector<HostPrim>& prims, const BinaryBvh& bvh, PackedScene* out) {
  const size_t n = prims.size();
  out->prims.resize(n);
  out->prim_ids.resize(n);
  // Reorder primitives by the permutation; mark each leaf's last record.
  std::vector<uint8_t> last(n, 0);
  for (const BinNode& nd : bvh.nodes)
    if (nd.leaf) last[nd.a + nd.b - 1] = 1;
  for (size_t k = 0; k < n; ++k) {
    uint64_t id = bvh.indices[k];
    out->prims[k] = pack_prim(prims[id], last[k] != 0);
    out->prim_ids[k] = uint32_t(id);
  }
  // Internal nodes get GPU records in pre-order.
  std::vector<uint32_t> gidx(bvh.nodes.size(), 0);
  uint32_t count = 0;
  for (size_t i = 0; i < bvh.nodes.size(); ++i)
    if (!bvh.nodes[i].leaf) gidx[i] = count++;
  auto ref = [&](uint64_t i) -> uint32_t {
    const BinNode& c = bvh.nodes[i];
    return c.leaf ? (GT_LEAF_BIT | uint32_t(c.a)) : gidx[i];
  };
  out->nodes2.resize(count);
  for (size_t i = 0; i < bvh.nodes.size(); ++i) {
    const BinNode& nd = bvh.nodes[i];
    if (nd.leaf) continue;
    GNode2& g = out->nodes2[gidx[i]];
    const Box& L = bvh.nodes[nd.a].bounds;
    const Box& R = bvh.nodes[nd.b].bounds;
    g.lx[0] = round_down(L.lo[0]); g.lx[1] = round_up(L.hi[0]);
    g.ly[0] = round_down(L.lo[1]); g.ly[1] = round_up(L.hi[1]);
    g.lz[0] = round_down(L.lo[2]); g.lz[1] = round_up(L.hi[2]);
    g.rx[0] = round_down(R.lo[0]); g.rx[1] = round_up(R.hi[0]);
    g.ry[0] = round_down(R.lo[1]); g.ry[1] = round_up(R.hi[1]);
    g.rz[0] = round_down(R.lo[2]); g.rz[1] = round_up(R.hi[2]);
    g.left = ref(nd.a);
    g.right = ref(nd.b);
    g.pad0 = g.pad1 = 0;
  }
  const Box& root = bvh.nodes[0].bounds;
  for (int a = 0; a < 3; ++a) { out->root_lo[a] = round_down(root.lo[a]); out->root_hi[a] = round_up(root.hi[a]); }
  out->root_ref2 = ref(0);
  out->depth2 = bvh.depth;

  // BVH4: each wide node holds the (up to four) grandchildren of a binary
  // node, listed left-left, left-right, right-left, right-right, so a
  // near-first walk of the wide node visits children in the binary order.
  out->nodes4.clear();
  struct W { uint64_t bin; uint32_t slot; };
  std::vector<uint32_t> wide_of(bvh.nodes.size(), GT_EMPTY_CHILD);
  std::vector<uint64_t> order;  // binary internal nodes that become wide nodes, pre-order
  std::vector<uint64_t> work;
  if (!bvh.nodes[0].leaf) work.push_back(0);
  while (!work.empty()) {
    uint64_t i = work.back();
    work.pop_back();
    wide_of[i] = uint32_t(order.size());
    order.push_back(i);
    const BinNode& nd = bvh.nodes[i];
    uint64_t kids[4];
    int nk = 0;
    for (uint64_t c : {nd.a, nd.b}) {
      if (bvh.nodes[c].leaf) kids[nk++] = c;
      else { kids[nk++] = bvh.nodes[c].a; kids[nk++] = bvh.nodes[c].b; }
    }
    for (int k = nk - 1; k >= 0; --k)
      if (!bvh.nodes[kids[k]].leaf) work.push_back(kids[k]);
  }
  out->nodes4.resize(order.size());
  uint64_t depth4 = 0;
  std::vector<uint64_t> wdepth(order.size(), 1);
  for (size_t w = 0; w < order.size(); ++w) {
    const BinNode& nd = bvh.nodes[order[w]];
    GNode4& g = out->nodes4[w];
    std::memset(&g, 0, sizeof g);
    uint64_t kids[4];
    int nk = 0;
    for (uint64_t c : {nd.a, nd.b}) {
      if (bvh.nodes[c].leaf) kids[nk++] = c;
      else { kids[nk++] = bvh.nodes[c].a; kids[nk++] = bvh.nodes[c].b; }
    }
    for (int k = 0; k < 4; ++k) {
      if (k >= nk) {
        g.child[k] = GT_EMPTY_CHILD;
        g.lox[k] = g.loy[k] = g.loz[k] = 1.0f;
        g.hix[k] = g.hiy[k] = g.hiz[k] = -1.0f;
        continue;
      }
      const BinNode& c = bvh.nodes[kids[k]];
      g.lox[k] = round_down(c.bounds.lo[0]); g.hix[k] = round_up(c.bounds.hi[0]);
      g.loy[k] = round_down(c.bounds.lo[1]); g.hiy[k] = round_up(c.bounds.hi[1]);
      g.loz[k] = round_down(c.bounds.lo[2]); g.hiz[k] = round_up(c.bounds.hi[2]);
      g.child[k] = c.leaf ? (GT_LEAF_BIT | uint32_t(c.a)) : wide_of[kids[k]];
      if (!c.leaf) { wdepth[wide_of[kids[k]]] = wdepth[w] + 1; depth4 = std::max(depth4, wdepth[w] + 1); }
    }
    depth4 = std::max(depth4, wdepth[w]);
  }
  out->root_ref4 = bvh.nodes[0].leaf ? (GT_LEAF_BIT | uint32_t(bvh.nodes[0].a)) : 0u;
  out->depth4 = depth4 + 1;
  (void)as_bits;
}

}  // namespace gt
