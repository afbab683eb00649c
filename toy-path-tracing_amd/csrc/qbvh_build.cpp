// Host QBVH builder: binned-SAH 4-wide BVH, the build of R/qbvh.rs:144-516
// (R/ = crates/renderer/src/).
//
// The node array is the interface between the renderer and the GPU kernels,
// so the tree must be exactly the reference's: same split decisions, same
// in-place partition order, same pre-order flattening.  Float expressions
// keep the reference's evaluation order (no contraction: built with
// -ffp-contract=off).  Subtrees above a size threshold are built on worker
// threads; each task owns a disjoint primitive range and its result is
// spliced in a fixed order, so the output does not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <future>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "tpt_internal.h"

namespace tpt {
namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr int kBins = 16;            // BIN_COUNT
constexpr size_t kLeafTarget = 4;    // TARGET_LEAF_PRIMS
constexpr size_t kLeafMax = 16;      // MAX_LEAF_PRIMS
constexpr size_t kTaskGrain = 16384; // subtrees at least this large become tasks

struct Box {
  float lo[3], hi[3];
};
inline Box empty_box() { return Box{{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}}; }
// glam Vec3::min / max are f32::min / max per component (NaN-ignoring).
inline void grow(Box& a, const Box& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::fmin(a.lo[k], b.lo[k]);
    a.hi[k] = std::fmax(a.hi[k], b.hi[k]);
  }
}
inline void grow_point(Box& a, const float* p) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = std::fmin(a.lo[k], p[k]);
    a.hi[k] = std::fmax(a.hi[k], p[k]);
  }
}
// Bounds::surface_area, R/scene/mesh.rs:32-35
inline float area(const Box& b) {
  float ex = std::fmax(b.hi[0] - b.lo[0], 0.0f);
  float ey = std::fmax(b.hi[1] - b.lo[1], 0.0f);
  float ez = std::fmax(b.hi[2] - b.lo[2], 0.0f);
  return 2.0f * (ex * ey + ex * ez + ey * ez);
}

struct Prim {
  Box box;
  float c[3];  // Bounds::center = 0.5 * (min + max)
  uint32_t index;
};

// Build tree: children are 2..4 entries; a leaf covers `count` primitives.
struct BNode {
  Box box;
  uint32_t count = 0;  // leaf primitive count (leaf iff nchild == 0)
  int nchild = 0;
  std::unique_ptr<BNode> child[4];
};

inline size_t bin_of(float c, float cmin, float inv) {  // bin_index_for_centroid, R/qbvh.rs:396-400
  float raw = (c - cmin) * inv;
  if (raw < 0.0f) raw = 0.0f;
  if (raw > (float)(kBins - 1)) raw = (float)(kBins - 1);
  if (raw != raw) return 0;  // NaN as usize == 0
  return (size_t)raw;
}

struct Split {
  int axis;
  float cmin, cmax;
  int bin;
  float cost;
};

Box range_box(const Prim* p, size_t n) {
  Box b = empty_box();
  for (size_t i = 0; i < n; ++i) grow(b, p[i].box);
  return b;
}

// find_best_binary_split + bin_primitives + evaluate_bins, R/qbvh.rs:318-448.
bool best_split(const Prim* p, size_t n, const Box& parent, Split* out) {
  if (n <= 1) return false;
  Box cb = empty_box();
  for (size_t i = 0; i < n; ++i) grow_point(cb, p[i].c);
  const float parent_area = std::fmax(area(parent), 1.0e-30f);
  bool found = false;
  for (int axis = 0; axis < 3; ++axis) {
    const float cmin = cb.lo[axis], cmax = cb.hi[axis];
    if (cmax - cmin <= 0.0f) continue;
    const float inv = (float)kBins / std::fmax(cmax - cmin, 1.0e-30f);
    Box bb[kBins];
    uint32_t bc[kBins];
    for (int i = 0; i < kBins; ++i) {
      bb[i] = empty_box();
      bc[i] = 0;
    }
    for (size_t i = 0; i < n; ++i) {
      size_t b = bin_of(p[i].c[axis], cmin, inv);
      grow(bb[b], p[i].box);
      bc[b]++;
    }
    // prefix (left) sweep and suffix (right) sweep
    Box left[kBins], right[kBins];
    uint32_t lcount[kBins];
    Box acc = empty_box();
    uint32_t accn = 0;
    for (int i = 0; i < kBins; ++i) {
      grow(acc, bb[i]);
      accn += bc[i];
      left[i] = acc;
      lcount[i] = accn;
    }
    acc = empty_box();
    for (int i = kBins - 1; i >= 0; --i) {
      grow(acc, bb[i]);
      right[i] = acc;
    }
    const uint32_t total = lcount[kBins - 1];
    if (total <= 1) continue;
    float best = kInf;
    int best_bin = -1;
    for (int s = 0; s < kBins - 1; ++s) {
      const uint32_t nl = lcount[s], nr = total - nl;
      if (nl == 0 || nr == 0) continue;
      // COST_TRAVERSE + COST_INTERSECT * (al*nl + ar*nr) / parent_area, both costs 1
      const float cost = 1.0f + 1.0f * (area(left[s]) * (float)nl + area(right[s + 1]) * (float)nr) / parent_area;
      if (cost < best) {
        best = cost;
        best_bin = s;
      }
    }
    if (best_bin < 0) continue;
    if (!found || best < out->cost) {
      *out = Split{axis, cmin, cmax, best_bin, best};
      found = true;
    }
  }
  return found;
}

// partition_in_place, R/qbvh.rs:454-474 (swap-from-the-right, order matters).
size_t partition(Prim* p, size_t n, const Split& s) {
  const float inv = (float)kBins / std::fmax(s.cmax - s.cmin, 1.0e-30f);
  size_t l = 0, r = n;
  while (l < r) {
    if ((int)bin_of(p[l].c[s.axis], s.cmin, inv) <= s.bin) {
      ++l;
    } else {
      --r;
      std::swap(p[l], p[r]);
    }
  }
  return l;
}

class Builder {
 public:
  explicit Builder(int threads) : budget_(threads > 1 ? threads - 1 : 0) {}

  std::unique_ptr<BNode> node(Prim* p, size_t n);  // build_qbvh_node, R/qbvh.rs:180-216

 private:
  // split_into_children, R/qbvh.rs:222-258: one or two subtrees of this range.
  int halves(Prim* p, size_t n, std::unique_ptr<BNode>* out);
  std::unique_ptr<BNode> forced(Prim* p, size_t n, const Box& b);  // make_leaf_or_force_split
  // Runs two independent builds, on a worker thread when the range is large.
  template <class F, class G>
  void pair(size_t n, F&& f, G&& g) {
    if (n >= kTaskGrain && take_thread()) {
      auto fut = std::async(std::launch::async, std::forward<F>(f));
      g();
      fut.get();
      give_thread();
    } else {
      f();
      g();
    }
  }
  bool take_thread() {
    int b = budget_.load();
    while (b > 0)
      if (budget_.compare_exchange_weak(b, b - 1)) return true;
    return false;
  }
  void give_thread() { budget_.fetch_add(1); }
  std::atomic<int> budget_;
};

std::unique_ptr<BNode> make_leaf(const Box& b, size_t n) {
  auto l = std::make_unique<BNode>();
  l->box = b;
  l->count = (uint32_t)n;
  return l;
}

std::unique_ptr<BNode> Builder::node(Prim* p, size_t n) {
  const Box b = range_box(p, n);
  if (n <= kLeafTarget) return make_leaf(b, n);
  Split s;
  if (!best_split(p, n, b, &s)) return forced(p, n, b);
  if (s.cost >= (float)n && n <= kLeafMax) return make_leaf(b, n);  // leaf_cost(n) = n
  const size_t mid = partition(p, n, s);
  if (mid == 0 || mid == n) return forced(p, n, b);
  auto in = std::make_unique<BNode>();
  in->box = b;
  std::unique_ptr<BNode> lc[2], rc[2];
  int nl = 0, nr = 0;
  pair(n, [&] { nl = halves(p, mid, lc); }, [&] { nr = halves(p + mid, n - mid, rc); });
  for (int i = 0; i < nl; ++i) in->child[in->nchild++] = std::move(lc[i]);
  for (int i = 0; i < nr; ++i) in->child[in->nchild++] = std::move(rc[i]);
  return in;
}

int Builder::halves(Prim* p, size_t n, std::unique_ptr<BNode>* out) {
  if (n == 0) return 0;
  const Box b = range_box(p, n);
  if (n <= kLeafTarget) {
    out[0] = node(p, n);
    return 1;
  }
  Split s;
  if (!best_split(p, n, b, &s) || (s.cost >= (float)n && n <= kLeafMax)) {
    out[0] = node(p, n);
    return 1;
  }
  const size_t mid = partition(p, n, s);
  if (mid == 0 || mid == n) {
    out[0] = node(p, n);
    return 1;
  }
  pair(n, [&] { out[0] = node(p, mid); }, [&] { out[1] = node(p + mid, n - mid); });
  return 2;
}

std::unique_ptr<BNode> Builder::forced(Prim* p, size_t n, const Box& b) {
  if (n <= kLeafMax) return make_leaf(b, n);
  const size_t mid = n / 2;
  const float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
  const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
  // The SAH finds no split only when all centroids coincide, so all keys are
  // equal and Rust's sort_unstable_by leaves the slice as it is (one sorted
  // run); a stable sort reproduces that and any sorted input order.
  std::stable_sort(p, p + n, [axis](const Prim& a, const Prim& c) { return a.c[axis] < c.c[axis]; });
  auto in = std::make_unique<BNode>();
  in->box = b;
  std::unique_ptr<BNode> lc[2], rc[2];
  int nl = 0, nr = 0;
  pair(n, [&] { nl = halves(p, mid, lc); }, [&] { nr = halves(p + mid, n - mid, rc); });
  for (int i = 0; i < nl; ++i) in->child[in->nchild++] = std::move(lc[i]);
  for (int i = 0; i < nr; ++i) in->child[in->nchild++] = std::move(rc[i]);
  return in;
}

// flatten_recursive, R/qbvh.rs:484-505: pre-order, the parent slot is
// reserved before its children, leaf offsets run in child order.
uint32_t flatten(const BNode& b, uint32_t offset, std::vector<TptQbvhNode>& nodes, uint32_t* ref, bool* overflow) {
  if (b.nchild == 0) {
    if (b.count > kChildCountMask || offset > kChildOffsetMask) *overflow = true;
    *ref = kChildLeafBit | (b.count << kChildCountShift) | (offset & kChildOffsetMask);
    return b.count;
  }
  const uint32_t idx = (uint32_t)nodes.size();
  nodes.push_back(empty_qbvh_node());
  TptQbvhNode nn = empty_qbvh_node();
  uint32_t cur = offset;
  for (int i = 0; i < b.nchild; ++i) {
    uint32_t cref = 0;
    cur += flatten(*b.child[i], cur, nodes, &cref, overflow);
    const Box& cb = b.child[i]->box;
    nn.min_x[i] = cb.lo[0];
    nn.min_y[i] = cb.lo[1];
    nn.min_z[i] = cb.lo[2];
    nn.max_x[i] = cb.hi[0];
    nn.max_y[i] = cb.hi[1];
    nn.max_z[i] = cb.hi[2];
    nn.children[i] = cref;
    nn.valid_mask |= (uint8_t)(1u << i);
  }
  nodes[idx] = nn;
  *ref = idx;
  return cur - offset;
}

}  // namespace

TptQbvhNode empty_qbvh_node() {  // QbvhNode::empty, R/qbvh.rs:79-90
  TptQbvhNode n;
  std::memset(&n, 0, sizeof(n));
  for (int i = 0; i < 4; ++i) {
    n.min_x[i] = n.min_y[i] = n.min_z[i] = kInf;
    n.max_x[i] = n.max_y[i] = n.max_z[i] = -kInf;
    n.children[i] = kChildLeafBit;  // EMPTY_CHILD
  }
  return n;
}

int build_qbvh(const float* bmin, const float* bmax, size_t n, int threads, HostQbvh* out) {
  out->nodes.clear();
  out->prim.clear();
  if (n == 0) return 0;
  if (n > 0xFFFFFFFFull) return -1;
  std::vector<Prim> prims(n);
  for (size_t i = 0; i < n; ++i) {
    Prim& p = prims[i];
    for (int k = 0; k < 3; ++k) {
      p.box.lo[k] = bmin[3 * i + k];
      p.box.hi[k] = bmax[3 * i + k];
      p.c[k] = 0.5f * (p.box.lo[k] + p.box.hi[k]);
    }
    p.index = (uint32_t)i;
  }
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  Builder builder(threads);
  std::unique_ptr<BNode> root = builder.node(prims.data(), n);
  out->prim.resize(n);
  for (size_t i = 0; i < n; ++i) out->prim[i] = prims[i].index;
  bool overflow = false;
  uint32_t ref = 0;
  flatten(*root, 0, out->nodes, &ref, &overflow);
  return overflow ? -2 : 0;
}

}  // namespace tpt
