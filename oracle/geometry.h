// Oracle geometry: math, watertight triangle test, binned-SAH QBVH build,
// traversal, two-level closest hit and the shadow rule, restated from the
// reference (R/ = crates/renderer/src).  TEST INFRASTRUCTURE ONLY.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace orc {

// The oracle's arithmetic type.  float: the reference computes in f32, and
// this is the instantiation every parity check uses (liboracle.so).  With
// ORC_FP64 (liboracle64.so) every value is widened to double, the same
// algorithm in FP64, which serves as the gradients' FP64 reference; the
// sampler's outputs stay the f32 values of the reference either way.
#ifdef ORC_FP64
using real = double;
#else
using real = float;
#endif

constexpr real kInf = std::numeric_limits<real>::infinity();

// ---------------------------------------------------------------------------
// glam::Vec3 semantics (glam 0.32, scalar Vec3): component-wise f32 ops,
// dot = x*x' + y*y' + z*z' left to right, min/max = f32::min/max (NaN-ignoring).
struct V3 {
  real x, y, z;
};
inline V3 v3(real x, real y, real z) { return V3{x, y, z}; }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 mul(real s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline real rmin(real a, real b) { return std::fmin(a, b); }  // Rust f32::min
inline real rmax(real a, real b) { return std::fmax(a, b); }  // Rust f32::max
inline V3 vmin(V3 a, V3 b) { return v3(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return v3(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)); }
inline V3 vabs(V3 a) { return v3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)); }
inline real max_element(V3 a) { return rmax(a.x, rmax(a.y, a.z)); }
inline real comp(V3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// Mat4 column-major (glam layout: x_axis, y_axis, z_axis, w_axis).
// transform_point3: res = x*px; res = y*py + res; res = z*pz + res; res = w + res.
inline V3 transform_point3(const real* m, V3 p) {
  real rx = m[0] * p.x, ry = m[1] * p.x, rz = m[2] * p.x;
  rx = m[4] * p.y + rx; ry = m[5] * p.y + ry; rz = m[6] * p.y + rz;
  rx = m[8] * p.z + rx; ry = m[9] * p.z + ry; rz = m[10] * p.z + rz;
  rx = m[12] + rx; ry = m[13] + ry; rz = m[14] + rz;
  return v3(rx, ry, rz);
}
inline V3 transform_vector3(const real* m, V3 p) {
  real rx = m[0] * p.x, ry = m[1] * p.x, rz = m[2] * p.x;
  rx = m[4] * p.y + rx; ry = m[5] * p.y + ry; rz = m[6] * p.y + rz;
  rx = m[8] * p.z + rx; ry = m[9] * p.z + ry; rz = m[10] * p.z + rz;
  return v3(rx, ry, rz);
}

// ---------------------------------------------------------------------------
// R/math/mod.rs:136-161
constexpr real kMachineEpsilon = std::numeric_limits<real>::epsilon() * 0.5f;

inline real difference_of_products(real a, real b, real c, real d) {
  real cd = c * d;
  real dop = std::fma(a, b, -cd);
  real err = std::fma(-c, d, cd);
  return dop + err;
}
inline real gamma_n(int n) {
  real nf = (real)n;
  return (nf * kMachineEpsilon) / (1.0f - nf * kMachineEpsilon);
}
inline int max_component_index(V3 v) {
  if (v.x > v.y) return v.x > v.z ? 0 : 2;
  return v.y > v.z ? 1 : 2;
}
inline V3 permute(V3 v, int x, int y, int z) {
  real a[3] = {v.x, v.y, v.z};
  return v3(a[x], a[y], a[z]);
}

struct TriHit {
  real t, b0, b1, b2;
};

// R/math/ray.rs:107-209 (pbrt-v4 watertight test, fp64 fallback on e==0).
bool intersect_triangle(V3 ro, V3 rd, real t_max, V3 v0, V3 v1, V3 v2, TriHit* out) {
  {
    V3 c = cross(sub(v2, v0), sub(v1, v0));
    if (dot(c, c) == 0.0f) return false;
  }
  V3 p0t = sub(v0, ro), p1t = sub(v1, ro), p2t = sub(v2, ro);
  int kz = max_component_index(vabs(rd));
  int kx = kz + 1;
  if (kx == 3) kx = 0;
  int ky = kx + 1;
  if (ky == 3) ky = 0;
  V3 d = permute(rd, kx, ky, kz);
  if (d.z == 0.0f) return false;
  p0t = permute(p0t, kx, ky, kz);
  p1t = permute(p1t, kx, ky, kz);
  p2t = permute(p2t, kx, ky, kz);
  real sx = -d.x / d.z;
  real sy = -d.y / d.z;
  real sz = 1.0f / d.z;
  p0t.x += sx * p0t.z;
  p0t.y += sy * p0t.z;
  p1t.x += sx * p1t.z;
  p1t.y += sy * p1t.z;
  p2t.x += sx * p2t.z;
  p2t.y += sy * p2t.z;
  real e0 = difference_of_products(p1t.x, p2t.y, p1t.y, p2t.x);
  real e1 = difference_of_products(p2t.x, p0t.y, p2t.y, p0t.x);
  real e2 = difference_of_products(p0t.x, p1t.y, p0t.y, p1t.x);
  if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
    e0 = (real)((double)p2t.y * (double)p1t.x - (double)p2t.x * (double)p1t.y);
    e1 = (real)((double)p0t.y * (double)p2t.x - (double)p0t.x * (double)p2t.y);
    e2 = (real)((double)p1t.y * (double)p0t.x - (double)p1t.x * (double)p0t.y);
  }
  if ((e0 < 0.0f || e1 < 0.0f || e2 < 0.0f) && (e0 > 0.0f || e1 > 0.0f || e2 > 0.0f)) return false;
  real det = e0 + e1 + e2;
  if (det == 0.0f) return false;
  p0t.z *= sz;
  p1t.z *= sz;
  p2t.z *= sz;
  real t_scaled = e0 * p0t.z + e1 * p1t.z + e2 * p2t.z;
  if ((det < 0.0f && (t_scaled >= 0.0f || t_scaled < t_max * det)) ||
      (det > 0.0f && (t_scaled <= 0.0f || t_scaled > t_max * det)))
    return false;
  real inv_det = 1.0f / det;
  real b0 = e0 * inv_det, b1 = e1 * inv_det, b2 = e2 * inv_det;
  real t = t_scaled * inv_det;
  real max_zt = max_element(vabs(v3(p0t.z, p1t.z, p2t.z)));
  real delta_z = gamma_n(3) * max_zt;
  real max_xt = max_element(vabs(v3(p0t.x, p1t.x, p2t.x)));
  real max_yt = max_element(vabs(v3(p0t.y, p1t.y, p2t.y)));
  real delta_x = gamma_n(5) * (max_xt + max_zt);
  real delta_y = gamma_n(5) * (max_yt + max_zt);
  real delta_e = 2.0f * (gamma_n(2) * max_xt * max_yt + delta_y * max_xt + delta_x * max_yt);
  real max_e = max_element(vabs(v3(e0, e1, e2)));
  real delta_t =
      3.0f * (gamma_n(3) * max_e * max_zt + delta_e * max_zt + delta_z * max_e) * std::fabs(inv_det);
  if (t <= delta_t) return false;
  out->t = t;
  out->b0 = b0;
  out->b1 = b1;
  out->b2 = b2;
  return true;
}

// ---------------------------------------------------------------------------
// R/scene/mesh.rs:11-42 Bounds
struct Bounds {
  V3 mn, mx;
};
const Bounds kEmptyBounds = {{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}};
inline V3 center(const Bounds& b) { return mul(0.5f, add(b.mn, b.mx)); }
inline V3 extent(const Bounds& b) { return sub(b.mx, b.mn); }
inline real surface_area(const Bounds& b) {
  V3 e = vmax(extent(b), v3(0.f, 0.f, 0.f));
  return 2.0f * (e.x * e.y + e.x * e.z + e.y * e.z);
}
inline Bounds unite(const Bounds& a, const Bounds& b) { return Bounds{vmin(a.mn, b.mn), vmax(a.mx, b.mx)}; }

// ---------------------------------------------------------------------------
// R/qbvh.rs:7-63 constants, child encoding, node.
constexpr int kBinCount = 16;
constexpr size_t kTargetLeafPrims = 4;
constexpr size_t kMaxLeafPrims = 16;
constexpr real kCostTraverse = 1.0f;
constexpr real kCostIntersect = 1.0f;
constexpr uint32_t kLeafBit = 1u << 31;
constexpr uint32_t kCountShift = 24;
constexpr uint32_t kCountMask = 0x7F;
constexpr uint32_t kOffsetMask = (1u << 24) - 1;
constexpr uint32_t kEmptyChild = kLeafBit;

inline uint32_t encode_leaf(uint32_t offset, uint32_t count) { return kLeafBit | (count << kCountShift) | offset; }

struct Node {  // QbvhNode, R/qbvh.rs:54-64 (SoA over 4 children)
  real min_x[4], min_y[4], min_z[4], max_x[4], max_y[4], max_z[4];
  uint32_t children[4];
  uint8_t valid_mask;
};
Node empty_node() {  // R/qbvh.rs:79-90
  Node n;
  for (int i = 0; i < 4; ++i) {
    n.min_x[i] = n.min_y[i] = n.min_z[i] = kInf;
    n.max_x[i] = n.max_y[i] = n.max_z[i] = -kInf;
    n.children[i] = kEmptyChild;
  }
  n.valid_mask = 0;
  return n;
}
void set_child(Node& n, int slot, uint32_t ref, const Bounds& b) {  // R/qbvh.rs:92-115
  n.min_x[slot] = b.mn.x;
  n.min_y[slot] = b.mn.y;
  n.min_z[slot] = b.mn.z;
  n.max_x[slot] = b.mx.x;
  n.max_y[slot] = b.mx.y;
  n.max_z[slot] = b.mx.z;
  n.children[slot] = ref;
  n.valid_mask |= (uint8_t)(1u << slot);
}

struct Qbvh {
  std::vector<Node> nodes;
  std::vector<uint32_t> prim;  // primitive_indices
};

// ---------------------------------------------------------------------------
// Binned SAH build, R/qbvh.rs:136-516.  The reference parallelises with rayon
// but combines halves in a fixed order and merges bins exactly (min/max and
// integer counts), so this sequential restatement produces the same tree.
struct BPrim {
  uint32_t index;
  Bounds bounds;
  V3 centroid;
};

struct BuildNode {
  bool leaf;
  Bounds bounds;
  uint32_t count;
  std::vector<BuildNode> children;
};

Bounds bounds_of(const BPrim* p, size_t n) {  // R/qbvh.rs:484-490
  Bounds b = kEmptyBounds;
  for (size_t i = 0; i < n; ++i) b = unite(b, p[i].bounds);
  return b;
}
Bounds centroid_bounds_of(const BPrim* p, size_t n) {  // R/qbvh.rs:492-501
  Bounds b = kEmptyBounds;
  for (size_t i = 0; i < n; ++i) b = unite(b, Bounds{p[i].centroid, p[i].centroid});
  return b;
}

struct Bin {
  Bounds bounds;
  uint32_t count;
};
struct BinSplit {
  int axis;
  real cmin, cmax;
  int bin_split;
  real cost;
};

size_t bin_index_for_centroid(real c, real cmin, real inv) {  // R/qbvh.rs:396-400
  real raw = (c - cmin) * inv;
  // f32::clamp(0, 15): NaN stays NaN; `as usize` maps NaN to 0 and saturates.
  real cl = raw;
  if (cl < 0.0f) cl = 0.0f;
  if (cl > (real)(kBinCount - 1)) cl = (real)(kBinCount - 1);
  if (std::isnan(cl)) return 0;
  return (size_t)cl;
}

void bin_primitives(const BPrim* p, size_t n, int axis, real cmin, real cmax, Bin* bins) {  // :376-394
  for (int i = 0; i < kBinCount; ++i) bins[i] = Bin{kEmptyBounds, 0};
  real ext = rmax(cmax - cmin, 1.0e-30f);
  real inv = (real)kBinCount / ext;
  for (size_t i = 0; i < n; ++i) {
    size_t bi = bin_index_for_centroid(comp(p[i].centroid, axis), cmin, inv);
    bins[bi].bounds = unite(bins[bi].bounds, p[i].bounds);
    bins[bi].count += 1;
  }
}

bool evaluate_bins(const Bin* bins, real parent_area, int* split_out, real* cost_out) {  // :402-448
  Bounds lb[kBinCount], rb[kBinCount];
  uint32_t lc[kBinCount], rc[kBinCount];
  Bounds acc = kEmptyBounds;
  uint32_t accc = 0;
  for (int i = 0; i < kBinCount; ++i) {
    acc = unite(acc, bins[i].bounds);
    accc += bins[i].count;
    lb[i] = acc;
    lc[i] = accc;
  }
  acc = kEmptyBounds;
  accc = 0;
  for (int i = kBinCount - 1; i >= 0; --i) {
    acc = unite(acc, bins[i].bounds);
    accc += bins[i].count;
    rb[i] = acc;
    rc[i] = accc;
  }
  (void)rc;
  uint32_t total = lc[kBinCount - 1];
  if (total <= 1) return false;
  real best_cost = kInf;
  int best = -1;
  for (int s = 0; s < kBinCount - 1; ++s) {
    uint32_t nl = lc[s];
    uint32_t nr = total - nl;
    if (nl == 0 || nr == 0) continue;
    real al = surface_area(lb[s]);
    real ar = surface_area(rb[s + 1]);
    real cost = kCostTraverse + kCostIntersect * (al * (real)nl + ar * (real)nr) / parent_area;
    if (cost < best_cost) {
      best_cost = cost;
      best = s;
    }
  }
  if (best < 0) return false;
  *split_out = best;
  *cost_out = best_cost;
  return true;
}

bool find_best_binary_split(const BPrim* p, size_t n, const Bounds& parent, BinSplit* out) {  // :318-354
  if (n <= 1) return false;
  Bounds cb = centroid_bounds_of(p, n);
  real parent_area = rmax(surface_area(parent), 1.0e-30f);
  bool have = false;
  for (int axis = 0; axis < 3; ++axis) {
    real cmin = comp(cb.mn, axis), cmax = comp(cb.mx, axis);
    if (cmax - cmin <= 0.0f) continue;
    Bin bins[kBinCount];
    bin_primitives(p, n, axis, cmin, cmax, bins);
    int s;
    real cost;
    if (!evaluate_bins(bins, parent_area, &s, &cost)) continue;
    if (!have || cost < out->cost) {
      *out = BinSplit{axis, cmin, cmax, s, cost};
      have = true;
    }
  }
  return have;
}

inline real leaf_cost(uint32_t n) { return kCostIntersect * (real)n; }  // :450-452

size_t partition_in_place(BPrim* p, size_t n, const BinSplit& s) {  // :454-474
  real ext = rmax(s.cmax - s.cmin, 1.0e-30f);
  real inv = (real)kBinCount / ext;
  size_t left = 0, right = n;
  while (left < right) {
    size_t bin = bin_index_for_centroid(comp(p[left].centroid, s.axis), s.cmin, inv);
    if ((int)bin <= s.bin_split) {
      left += 1;
    } else {
      right -= 1;
      std::swap(p[left], p[right]);
    }
  }
  return left;
}

BuildNode make_leaf(const Bounds& b, uint32_t count) {
  BuildNode n;
  n.leaf = true;
  n.bounds = b;
  n.count = count;
  return n;
}

BuildNode build_node(BPrim* p, size_t n);

std::vector<BuildNode> split_into_children(BPrim* p, size_t n) {  // :222-258
  std::vector<BuildNode> out;
  if (n == 0) return out;
  Bounds b = bounds_of(p, n);
  if (n <= kTargetLeafPrims) {
    out.push_back(build_node(p, n));
    return out;
  }
  BinSplit s;
  if (!find_best_binary_split(p, n, b, &s)) {
    out.push_back(build_node(p, n));
    return out;
  }
  if (s.cost >= leaf_cost((uint32_t)n) && n <= kMaxLeafPrims) {
    out.push_back(build_node(p, n));
    return out;
  }
  size_t mid = partition_in_place(p, n, s);
  if (mid == 0 || mid == n) {
    out.push_back(build_node(p, n));
    return out;
  }
  out.push_back(build_node(p, mid));
  out.push_back(build_node(p + mid, n - mid));
  return out;
}

BuildNode make_leaf_or_force_split(BPrim* p, size_t n, const Bounds& b) {  // :260-295
  if (n <= kMaxLeafPrims) return make_leaf(b, (uint32_t)n);
  size_t mid = n / 2;
  V3 e = extent(b);
  int axis = (e.x >= e.y && e.x >= e.z) ? 0 : (e.y >= e.z ? 1 : 2);
  // sort_unstable_by on centroid[axis].  This branch is reached only when the
  // SAH finds no split, i.e. every centroid coincides, so every key compares
  // Equal; Rust's ipnsort (edition 2024 std) detects the whole slice as one
  // non-descending run and leaves it untouched.  A stable sort reproduces that.
  std::stable_sort(p, p + n, [axis](const BPrim& a, const BPrim& c) {
    return comp(a.centroid, axis) < comp(c.centroid, axis);
  });
  BuildNode node;
  node.leaf = false;
  node.bounds = b;
  node.count = 0;
  auto l = split_into_children(p, mid);
  auto r = split_into_children(p + mid, n - mid);
  for (auto& c : l) node.children.push_back(std::move(c));
  for (auto& c : r) node.children.push_back(std::move(c));
  return node;
}

BuildNode build_node(BPrim* p, size_t n) {  // :180-216
  Bounds b = bounds_of(p, n);
  if (n <= kTargetLeafPrims) return make_leaf(b, (uint32_t)n);
  BinSplit s;
  if (!find_best_binary_split(p, n, b, &s)) return make_leaf_or_force_split(p, n, b);
  if (s.cost >= leaf_cost((uint32_t)n) && n <= kMaxLeafPrims) return make_leaf(b, (uint32_t)n);
  size_t mid = partition_in_place(p, n, s);
  if (mid == 0 || mid == n) return make_leaf_or_force_split(p, n, b);
  BuildNode node;
  node.leaf = false;
  node.bounds = b;
  node.count = 0;
  auto l = split_into_children(p, mid);
  auto r = split_into_children(p + mid, n - mid);
  for (auto& c : l) node.children.push_back(std::move(c));
  for (auto& c : r) node.children.push_back(std::move(c));
  return node;
}

// R/qbvh.rs:476-482,  flatten_recursive :484-505 (pre-order, node slot reserved first)
std::pair<uint32_t, uint32_t> flatten(const BuildNode& bn, uint32_t off, std::vector<Node>& nodes) {
  if (bn.leaf) return {encode_leaf(off, bn.count), bn.count};
  uint32_t idx = (uint32_t)nodes.size();
  nodes.push_back(empty_node());
  Node nn = empty_node();
  uint32_t cur = off;
  for (size_t i = 0; i < bn.children.size(); ++i) {
    auto rc = flatten(bn.children[i], cur, nodes);
    set_child(nn, (int)i, rc.first, bn.children[i].bounds);
    cur += rc.second;
  }
  nodes[idx] = nn;
  return {idx, cur - off};
}

// R/qbvh.rs:144-166
bool build_qbvh(const Bounds* pb, size_t n, Qbvh* out) {
  if (n == 0) return false;
  std::vector<BPrim> prims(n);
  for (size_t i = 0; i < n; ++i) prims[i] = BPrim{(uint32_t)i, pb[i], center(pb[i])};
  BuildNode root = build_node(prims.data(), n);
  out->prim.resize(n);
  for (size_t i = 0; i < n; ++i) out->prim[i] = prims[i].index;
  out->nodes.clear();
  flatten(root, 0, out->nodes);
  return true;
}

// ---------------------------------------------------------------------------
// Traversal, R/qbvh.rs:540-675.
inline real safe_inv(real d) { return d == 0.0f ? kInf : 1.0f / d; }  // :574-576
// wide::f32x4::fast_min / fast_max = SSE minps/maxps: (a<b)?a:b and (a>b)?a:b.
inline real fmin_sse(real a, real b) { return a < b ? a : b; }
inline real fmax_sse(real a, real b) { return a > b ? a : b; }

struct TraversalStats {
  uint64_t nodes = 0, tris = 0, insts = 0;
};

// intersect_4_aabbs, R/qbvh.rs:578-607
uint32_t intersect4(const Node& nd, V3 o, V3 inv, real t_max, real tmin_out[4]) {
  uint32_t bits = 0;
  for (int i = 0; i < 4; ++i) {
    real t1x = (nd.min_x[i] - o.x) * inv.x, t2x = (nd.max_x[i] - o.x) * inv.x;
    real t1y = (nd.min_y[i] - o.y) * inv.y, t2y = (nd.max_y[i] - o.y) * inv.y;
    real t1z = (nd.min_z[i] - o.z) * inv.z, t2z = (nd.max_z[i] - o.z) * inv.z;
    real tminx = fmin_sse(t1x, t2x), tmaxx = fmax_sse(t1x, t2x);
    real tminy = fmin_sse(t1y, t2y), tmaxy = fmax_sse(t1y, t2y);
    real tminz = fmin_sse(t1z, t2z), tmaxz = fmax_sse(t1z, t2z);
    real tmin = fmax_sse(fmax_sse(fmax_sse(tminx, tminy), tminz), 0.0f);
    real tmax = fmin_sse(fmin_sse(fmin_sse(tmaxx, tmaxy), tmaxz), t_max);
    if (tmin <= tmax) bits |= 1u << i;  // simd_le: ordered <=
    tmin_out[i] = tmin;
  }
  return bits & (uint32_t)nd.valid_mask;
}

constexpr int kStackDepth = 128;

// traverse_qbvh, R/qbvh.rs:611-675.  leaf(offset, count, t_max) -> new t_max.
template <class Leaf>
real traverse_qbvh(const Qbvh& q, V3 o, V3 d, real t_max0, Leaf&& leaf, TraversalStats* st) {
  if (q.nodes.empty()) {
    if (q.prim.empty()) return t_max0;
    return leaf(0u, (uint32_t)q.prim.size(), t_max0);
  }
  V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  real closest = t_max0;
  uint32_t stack[kStackDepth];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    uint32_t ni = stack[--sp];
    const Node& nd = q.nodes[ni];
    real tmin[4];
    if (st) st->nodes++;
    uint32_t bits = intersect4(nd, o, inv, closest, tmin);
    if (bits == 0) continue;
    real ht[4];
    uint32_t hc[4];
    int cnt = 0;
    for (int s = 0; s < 4; ++s) {
      if (!(bits & (1u << s))) continue;
      ht[cnt] = tmin[s];
      hc[cnt] = nd.children[s];
      cnt++;
    }
    for (int i = 1; i < cnt; ++i) {
      int j = i;
      while (j > 0 && ht[j - 1] > ht[j]) {
        std::swap(ht[j - 1], ht[j]);
        std::swap(hc[j - 1], hc[j]);
        j--;
      }
    }
    for (int k = cnt - 1; k >= 0; --k) {
      uint32_t c = hc[k];
      if (c & kLeafBit) {
        uint32_t count = (c >> kCountShift) & kCountMask;
        if (count == 0) continue;
        closest = leaf(c & kOffsetMask, count, closest);
      } else {
        if (sp >= kStackDepth) return closest;  // debug_assert in the reference
        stack[sp++] = c;
      }
    }
  }
  return closest;
}

// ---------------------------------------------------------------------------
// Scene, R/scene/mod.rs:62-98 (geometry only).
struct Mesh {
  std::vector<real> pos;     // 3 per vertex
  std::vector<uint32_t> idx;  // 3 per triangle
  Bounds bounds;
  Qbvh qbvh;
  bool has_qbvh = false;
  V3 vert(uint32_t i) const { return v3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]); }
  size_t tri_count() const { return idx.size() / 3; }
};
struct Instance {
  uint32_t mesh;
  real l2w[16], w2l[16];
  Bounds world_bounds;
};
struct Scene {
  std::vector<Mesh> meshes;
  std::vector<Instance> inst;
  Qbvh tlas;
  bool built = false;
};

Bounds transform_bounds(const Bounds& b, const real* m) {  // R/scene/mod.rs:869-893
  V3 c[8] = {v3(b.mn.x, b.mn.y, b.mn.z), v3(b.mn.x, b.mn.y, b.mx.z), v3(b.mn.x, b.mx.y, b.mn.z),
             v3(b.mn.x, b.mx.y, b.mx.z), v3(b.mx.x, b.mn.y, b.mn.z), v3(b.mx.x, b.mn.y, b.mx.z),
             v3(b.mx.x, b.mx.y, b.mn.z), v3(b.mx.x, b.mx.y, b.mx.z)};
  V3 f = transform_point3(m, c[0]);
  Bounds r{f, f};
  for (int i = 1; i < 8; ++i) {
    V3 p = transform_point3(m, c[i]);
    r.mn = vmin(r.mn, p);
    r.mx = vmax(r.mx, p);
  }
  return r;
}

struct SceneHit {
  uint32_t inst, tri;
  real t, b0, b1, b2;
};

// closest_mesh_hit_with_filter, R/scene/mod.rs:686-723 (no alpha test)
bool closest_mesh_hit(const Mesh& m, V3 o, V3 d, real t_max, SceneHit* best, TraversalStats* st) {
  bool found = false;
  traverse_qbvh(
      m.qbvh, o, d, t_max,
      [&](uint32_t off, uint32_t count, real cur) {
        real tm = cur;
        for (uint32_t k = off; k < off + count; ++k) {
          uint32_t ti = m.qbvh.prim[k];
          V3 a = m.vert(m.idx[3 * ti]), b = m.vert(m.idx[3 * ti + 1]), c = m.vert(m.idx[3 * ti + 2]);
          TriHit h;
          if (st) st->tris++;
          if (intersect_triangle(o, d, tm, a, b, c, &h)) {
            tm = h.t;
            best->tri = ti;
            best->t = h.t;
            best->b0 = h.b0;
            best->b1 = h.b1;
            best->b2 = h.b2;
            found = true;
          }
        }
        return tm;
      },
      st);
  return found;
}

// Scene::closest_hit, R/scene/mod.rs:340-395.  t_max0 = +inf reproduces the
// reference exactly; a finite value is this library's extension.
bool closest_hit(const Scene& s, V3 o, V3 d, real t_max0, SceneHit* out, TraversalStats* st) {
  if (s.inst.empty()) return false;
  bool found = false;
  traverse_qbvh(
      s.tlas, o, d, t_max0,
      [&](uint32_t off, uint32_t count, real cur) {
        real tm = cur;
        for (uint32_t k = off; k < off + count; ++k) {
          uint32_t ii = s.tlas.prim[k];
          const Instance& in = s.inst[ii];
          V3 lo = transform_point3(in.w2l, o);
          V3 ld = transform_vector3(in.w2l, d);
          if (st) st->insts++;
          SceneHit h;
          if (closest_mesh_hit(s.meshes[in.mesh], lo, ld, tm, &h, st)) {
            tm = h.t;
            h.inst = ii;
            *out = h;
            found = true;
          }
        }
        return tm;
      },
      st);
  return found;
}

// unoccluded_ray, R/integrator/mod.rs:239-265 (the shadow ray is built by the
// caller with spawn_ray, R/integrator/mod.rs:52-85).
//   target given      -> unoccluded iff no hit, or the closest hit is the target
//   distance infinite -> unoccluded iff no hit
//   otherwise         -> unoccluded iff no hit, or t >= distance * (1 - 1e-3)
constexpr real kShadowTolerance = 1.0e-3f;
bool unoccluded_ray(const Scene& s, V3 o, V3 d, real distance, int64_t target_inst, int64_t target_tri,
                    TraversalStats* st) {
  SceneHit h;
  if (!closest_hit(s, o, d, kInf, &h, st)) return true;
  if (target_inst >= 0) return h.inst == (uint32_t)target_inst && h.tri == (uint32_t)target_tri;
  if (std::isinf(distance)) return false;
  return h.t >= distance * (1.0f - kShadowTolerance);
}

template <class F>
void parallel_for(size_t n, int threads, F&& f) {
  if (threads <= 1 || n < 2) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  // chunks of up to 256 items, at least eight per thread (a small crop of
  // heavy pixels still spreads over every thread)
  const size_t chunk = std::max<size_t>(1, std::min<size_t>(256, n / (8 * (size_t)threads)));
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        size_t b = next.fetch_add(chunk);
        if (b >= n) break;
        size_t e = b + chunk < n ? b + chunk : n;
        for (size_t i = b; i < e; ++i) f(i);
      }
    });
  for (auto& th : pool) th.join();
}

}  // namespace orc
