// QBVH build on the GPU: the binned-SAH 4-wide build of R/qbvh.rs:144-516
// (R/ = crates/renderer/src/) producing the same node array and primitive
// order as the reference's build (and the host builder, qbvh_build.cpp).
//
// The reference builds top-down: build_qbvh_node(range) (a node) and
// split_into_children(range) (one or two children of a node) each take the
// range's box, find the best binned split (16 bins per axis, SAH), and either
// stop (a leaf, or a forced median split when no split exists) or partition
// the range in place with the swap-from-the-right loop of
// partition_in_place.  Every decision depends only on the range's contents,
// and every value it uses is either a min / max (order-free) or computed by
// one thread with the reference's expression order, so the tree is the
// reference's whatever the order of the parallel reductions.
//
// The partition's permutation is computed in closed form.  With mid = the
// number of primitives that go left, r = the number of right-going
// primitives in [begin, mid) (= the number of left-going ones in
// [mid, end)), p_j the position of the j-th right-going primitive from the
// front and q_j that of the j-th left-going one from the back (q_0 = end),
// the loop leaves: a left-going primitive in the front where it is; the j-th
// left-going one from the back at p_j; the j-th right-going one from the
// front at q_{j-1} - 1; a right-going primitive at mid at q_r - 1; every
// other right-going primitive in the back one place lower.
//
// Phases:
//   1  ranges of more than kSmall primitives, level by level: per range the
//      box, centroid box and 3 x 16 bins are reduced over chunks of
//      kChunk primitives (LDS then global atomics on order-preserving
//      integer images of the floats), one thread per range decides, and the
//      partition runs over the chunks (rank scans, then the scatter);
//   2  every range of at most kSmall primitives is finished by one block,
//      depth first with an LDS stack;
// in both, one wave evaluates a range's 3 x 15 split candidates at once.
//   3  flattening: the reference numbers interior nodes in pre-order, which
//      is the order of (first primitive, -primitive count): node X's index
//      is the number of interior nodes starting before X plus the number of
//      its ancestors starting where X starts.  Leaves are encoded with their
//      first primitive and count.
#include <algorithm>
#include <cstring>
#include <vector>

#include "tpt_device.h"

namespace tpt {
namespace {

constexpr int kBins = 16;              // BIN_COUNT
constexpr uint32_t kLeafTarget = 4;    // TARGET_LEAF_PRIMS
constexpr uint32_t kLeafMax = 16;      // MAX_LEAF_PRIMS
constexpr uint32_t kSmall = 512;       // ranges up to this size are built by one block
constexpr uint32_t kChunk = 2048;      // primitives per block in phase 1
constexpr int kThreads = 256;
constexpr int kStack = 64;             // phase-2 stack (smaller range first: depth <= 2 log2(kSmall) + 2)
constexpr uint32_t kBinWords = 3 * kBins * 7;  // per axis and bin: lo xyz, hi xyz (ordered), count
constexpr uint32_t kStatWords = 12 + kBinWords;

// order-preserving integer image of a float (min / max as unsigned compares)
__device__ inline uint32_t f2o(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float o2f(uint32_t o) {
  const uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  return __uint_as_float(u);
}
constexpr uint32_t kOrdPosInf = 0xFF800000u;  // f2o(+inf): the empty box's lo
constexpr uint32_t kOrdNegInf = 0x007FFFFFu;  // f2o(-inf): the empty box's hi

struct alignas(16) Prim {
  float lo[3], hi[3], c[3];  // c = 0.5 * (min + max) (Bounds::center)
  uint32_t index;
  uint32_t pad[2];
};
static_assert(sizeof(Prim) == 48, "Prim layout");

enum Kind : uint32_t { kNode = 0, kHalves = 1 };
// A range to process: build_qbvh_node (kNode: its node links to `parent` at
// child slot `link`) or split_into_children (kHalves: its one or two nodes
// link to `parent` at slots 2 * link + k).
struct Task {
  uint32_t begin, count, kind;
  int32_t parent;
  uint32_t link, pad;
};
// Build node: the box of its range, the range, and up to four children by
// slot (-1: none).
struct BNode {
  float lo[3], hi[3];
  uint32_t begin, count;
  int32_t parent;
  uint32_t leaf;
  int32_t child[4];
};
enum Outcome : uint32_t { kLeaf = 0, kSplit = 1, kForced = 2 };
struct Decision {
  uint32_t outcome, axis, bin, nl;  // nl: the primitives going left
  float cmin, inv;                  // the split axis's centroid minimum and bin scale
  uint32_t reorder;                 // a forced split whose sort would reorder the range
  uint32_t pad;
};

// ---- per-range statistics (box, centroid box, bins) ------------------------------

__device__ inline void stats_init(uint32_t* s, int t, int nt) {
  for (int i = t; i < (int)kStatWords; i += nt) {
    if (i < 12) {
      s[i] = (i % 6) < 3 ? kOrdPosInf : kOrdNegInf;
    } else {
      const int w = (i - 12) % 7;
      s[i] = w < 3 ? kOrdPosInf : (w < 6 ? kOrdNegInf : 0u);
    }
  }
}

// bin_index_for_centroid, R/qbvh.rs:406-410 (NaN as usize is 0)
__device__ inline int bin_of(float c, float cmin, float inv) {
  float raw = (c - cmin) * inv;
  if (raw != raw) return 0;
  if (raw < 0.0f) raw = 0.0f;
  if (raw > (float)(kBins - 1)) raw = (float)(kBins - 1);
  return (int)raw;
}

struct FBox {
  float lo[3], hi[3];
};
__device__ inline FBox fbox(const uint32_t* o) {
  FBox b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = o2f(o[k]);
    b.hi[k] = o2f(o[3 + k]);
  }
  return b;
}
// Vec3::min / max per component (the boxes hold no NaN)
__device__ inline void grow(FBox& a, const FBox& b) {
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = fminf(a.lo[k], b.lo[k]);
    a.hi[k] = fmaxf(a.hi[k], b.hi[k]);
  }
}
// Bounds::surface_area, R/scene/mesh.rs:32-35
__device__ inline float area(const FBox& b) {
  const float ex = fmaxf(b.hi[0] - b.lo[0], 0.0f);
  const float ey = fmaxf(b.hi[1] - b.lo[1], 0.0f);
  const float ez = fmaxf(b.hi[2] - b.lo[2], 0.0f);
  return 2.0f * (ex * ey + ex * ez + ey * ez);
}

// The centroid-box axis range and bin scale of an axis (bins exist only for
// an axis of positive extent).
__device__ inline bool axis_scale(const uint32_t* s, int axis, float* cmin, float* inv) {
  const float lo = o2f(s[6 + axis]), hi = o2f(s[9 + axis]);
  if (!(hi - lo > 0.0f)) return false;
  *cmin = lo;
  *inv = (float)kBins / fmaxf(hi - lo, 1.0e-30f);
  return true;
}

// The range's box and centroid box over primitives [i0, n) in steps of
// `step`: per thread, then per wave by shuffles, then one set of LDS
// atomics per wave into s[0..12).
__device__ inline void range_boxes(uint32_t* s, const Prim* prims, uint32_t i0, uint32_t n, uint32_t step) {
  uint32_t v[12];
  for (int k = 0; k < 12; ++k) v[k] = (k % 6) < 3 ? kOrdPosInf : kOrdNegInf;
  for (uint32_t i = i0; i < n; i += step) {
    const Prim p = prims[i];
    for (int k = 0; k < 3; ++k) {
      v[k] = min(v[k], f2o(p.lo[k]));
      v[3 + k] = max(v[3 + k], f2o(p.hi[k]));
      v[6 + k] = min(v[6 + k], f2o(p.c[k]));
      v[9 + k] = max(v[9 + k], f2o(p.c[k]));
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 12; ++k) {
      const uint32_t o = __shfl_xor(v[k], off);
      v[k] = (k % 6) < 3 ? min(v[k], o) : max(v[k], o);
    }
  if ((threadIdx.x & 63u) == 0)
    for (int k = 0; k < 12; ++k) {
      if ((k % 6) < 3) atomicMin(&s[k], v[k]);
      else atomicMax(&s[k], v[k]);
    }
}

// Accumulates one primitive into the bins of the live axes with atomics.
__device__ inline void stats_add_bins(uint32_t* bins, const Prim& p, const float* cmin, const float* inv,
                                      const bool* live) {
  for (int axis = 0; axis < 3; ++axis) {
    if (!live[axis]) continue;
    uint32_t* b = bins + (axis * kBins + bin_of(p.c[axis], cmin[axis], inv[axis])) * 7;
    for (int k = 0; k < 3; ++k) {
      atomicMin(&b[k], f2o(p.lo[k]));
      atomicMax(&b[3 + k], f2o(p.hi[k]));
    }
    atomicAdd(&b[6], 1u);
  }
}

// find_best_binary_split + evaluate_bins (R/qbvh.rs:312-459) and the decisions
// of build_qbvh_node / split_into_children / make_leaf_or_force_split
// (:186-310), which agree for a range: a leaf at <= 4 primitives; without a
// split a leaf at <= 16 else the forced median split; a leaf when the split
// costs at least the leaf and the range has <= 16 primitives; else the split.
//
// Called by a whole wave with the same statistics: lane axis * 16 + split
// evaluates one candidate (its left boxes grown over bins 0..split in
// order, its right boxes over 15..split+1, as the reference's sweeps grow
// them, so the cost is the reference's value), and the wave takes the
// smallest cost, ties to the lowest (axis, split): the reference keeps the
// first strictly smaller cost in that order.
__device__ Decision decide(const uint32_t* s, uint32_t n) {
  Decision d{};
  d.outcome = kLeaf;
  if (n <= kLeafTarget) return d;
  const FBox parent = fbox(s);
  const float parent_area = fmaxf(area(parent), 1.0e-30f);
  const int lane = (int)(threadIdx.x & 63u);
  const int axis = lane >> 4, sp = lane & 15;
  float cost = INFINITY, cmin = 0.0f, inv = 0.0f;
  int who = 64;  // no candidate
  uint32_t nl = 0;
  if (axis < 3 && sp < kBins - 1 && axis_scale(s, axis, &cmin, &inv)) {
    const uint32_t* bins = s + 12 + axis * kBins * 7;
    FBox l{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}}, r = l;
    uint32_t total = 0;
    for (int i = 0; i < kBins; ++i) {
      total += bins[i * 7 + 6];
      if (i <= sp) {
        grow(l, fbox(bins + i * 7));
        nl += bins[i * 7 + 6];
      }
    }
    for (int i = kBins - 1; i > sp; --i) grow(r, fbox(bins + i * 7));
    const uint32_t nr = total - nl;
    if (total > 1 && nl != 0 && nr != 0) {
      cost = 1.0f + 1.0f * (area(l) * (float)nl + area(r) * (float)nr) / parent_area;
      who = lane;
    }
  }
  float best = cost;
  int best_lane = who;
  for (int off = 32; off > 0; off >>= 1) {
    const float oc = __shfl_xor(best, off);
    const int ol = __shfl_xor(best_lane, off);
    if (ol < 64 && (best_lane == 64 || oc < best || (oc == best && ol < best_lane))) {
      best = oc;
      best_lane = ol;
    }
  }
  const bool found = best_lane < 64;
  const float best_cost = best;
  if (found) {
    d.axis = (uint32_t)(best_lane >> 4);
    d.bin = (uint32_t)(best_lane & 15);
    d.nl = (uint32_t)__shfl((int)nl, best_lane);
    d.cmin = __shfl(cmin, best_lane);
    d.inv = __shfl(inv, best_lane);
  }
  if (!found) {
    d.outcome = n <= kLeafMax ? kLeaf : kForced;
    if (d.outcome == kForced) {
      // make_leaf_or_force_split sorts by the centroid on the box's longest
      // axis; the split search fails only when the centroids coincide, and
      // then the sort leaves the order as it is.  Otherwise report it.
      const float ex = parent.hi[0] - parent.lo[0], ey = parent.hi[1] - parent.lo[1],
                  ez = parent.hi[2] - parent.lo[2];
      const int ax = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
      d.axis = ax;
      d.reorder = s[6 + ax] == s[9 + ax] ? 0u : 1u;
    }
    return d;
  }
  if (best_cost >= (float)n && n <= kLeafMax) return d;  // leaf_cost(n) = n
  d.outcome = kSplit;
  return d;
}

// What a decided range creates: its build node `id` (a leaf, or an
// interior node for a forced split or a kNode split; none for a kHalves
// split, whose two halves become kNode tasks) linked into its parent, and
// the tasks to run next (up to two).
struct Emit {
  Task t[2];
  int n;
};
__device__ inline void link_child(BNode* nodes, int32_t parent, uint32_t slot, int32_t id) {
  if (parent >= 0) nodes[parent].child[slot] = id;
}
__device__ inline uint32_t nodes_needed(const Task& t, const Decision& d) {
  return (d.outcome == kSplit && t.kind == kHalves) ? 0u : 1u;
}
__device__ Emit emit(const Task& t, const Decision& d, const uint32_t* s, BNode* nodes, int32_t id) {
  Emit e{};
  const uint32_t slot = t.kind == kNode ? t.link : 2u * t.link;
  if (d.outcome == kSplit && t.kind == kHalves) {  // [node(left), node(right)]
    e.t[0] = Task{t.begin, d.nl, kNode, t.parent, 2u * t.link, 0};
    e.t[1] = Task{t.begin + d.nl, t.count - d.nl, kNode, t.parent, 2u * t.link + 1, 0};
    e.n = 2;
    return e;
  }
  BNode& b = nodes[id];
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = o2f(s[k]);
    b.hi[k] = o2f(s[3 + k]);
  }
  b.begin = t.begin;
  b.count = t.count;
  b.parent = t.parent;
  b.leaf = d.outcome == kLeaf ? 1u : 0u;
  for (int k = 0; k < 4; ++k) b.child[k] = -1;
  link_child(nodes, t.parent, slot, id);
  if (d.outcome == kLeaf) return e;
  // an interior node: split_into_children of its two halves
  const uint32_t mid = d.outcome == kSplit ? d.nl : t.count / 2;
  e.t[0] = Task{t.begin, mid, kHalves, id, 0, 0};
  e.t[1] = Task{t.begin + mid, t.count - mid, kHalves, id, 1, 0};
  e.n = 2;
  return e;
}

// Destination of the primitive at g (0-based in its range) under the
// partition (see the file comment).  pos_r[k] is the position of the k-th
// (0-based, from the front) right-going primitive in front of mid, pos_l[k]
// that of the k-th (0-based, from the front) left-going primitive from mid
// on, so p_j = pos_r[j - 1] and q_j = pos_l[r - j].  rank: the primitive's
// k in the array of its kind.
__device__ inline uint32_t partition_dest(bool left, uint32_t g, uint32_t mid, uint32_t n, uint32_t r,
                                          uint32_t rank, const uint32_t* pos_r, const uint32_t* pos_l) {
  if (g < mid) {
    if (left) return g;
    return (rank == 0 ? n : pos_l[r - rank]) - 1;  // j = rank + 1 goes to q_{j-1} - 1
  }
  if (left) return pos_r[r - 1 - rank];  // j = r - rank (from the back) goes to p_j
  if (g == mid) return (r == 0 ? n : pos_l[0]) - 1;  // q_r - 1
  return g - 1;
}

// ---- phase 0: primitives ---------------------------------------------------------------

__global__ void init_prims(const float* __restrict__ bmin, const float* __restrict__ bmax, uint32_t n,
                           Prim* __restrict__ prims) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Prim p;
  for (int k = 0; k < 3; ++k) {
    p.lo[k] = bmin[3 * (size_t)i + k];
    p.hi[k] = bmax[3 * (size_t)i + k];
    p.c[k] = 0.5f * (p.lo[k] + p.hi[k]);
  }
  p.index = i;
  p.pad[0] = p.pad[1] = 0;
  prims[i] = p;
}

// ---- phase 1: ranges of more than kSmall primitives, by chunks ----------------------

struct Chunk {
  uint32_t task, begin, count, pad;  // begin: absolute primitive index
};

__global__ void stats_reset(uint32_t* __restrict__ stats, uint32_t tasks) {
  for (uint32_t t = blockIdx.x; t < tasks; t += gridDim.x) stats_init(stats + (size_t)t * kStatWords, threadIdx.x, blockDim.x);
}

__global__ __launch_bounds__(kThreads) void chunk_boxes(const Chunk* __restrict__ chunks, const Prim* __restrict__ prims,
                                                        uint32_t* __restrict__ stats) {
  __shared__ uint32_t s[12];
  const Chunk c = chunks[blockIdx.x];
  if (threadIdx.x < 12) s[threadIdx.x] = (threadIdx.x % 6) < 3 ? kOrdPosInf : kOrdNegInf;
  __syncthreads();
  range_boxes(s, prims + c.begin, threadIdx.x, c.count, kThreads);
  __syncthreads();
  if (threadIdx.x < 12) {
    uint32_t* g = stats + (size_t)c.task * kStatWords;
    if ((threadIdx.x % 6) < 3) atomicMin(&g[threadIdx.x], s[threadIdx.x]);
    else atomicMax(&g[threadIdx.x], s[threadIdx.x]);
  }
}

__global__ __launch_bounds__(kThreads) void chunk_bins(const Chunk* __restrict__ chunks, const Prim* __restrict__ prims,
                                                       uint32_t* __restrict__ stats) {
  __shared__ uint32_t s[kStatWords];
  const Chunk c = chunks[blockIdx.x];
  uint32_t* g = stats + (size_t)c.task * kStatWords;
  stats_init(s, threadIdx.x, kThreads);  // empty bins
  __syncthreads();
  if (threadIdx.x < 12) s[threadIdx.x] = g[threadIdx.x];  // the range's box and centroid box
  __syncthreads();
  float cmin[3], inv[3];
  bool live[3];
  for (int a = 0; a < 3; ++a) live[a] = axis_scale(s, a, &cmin[a], &inv[a]);
  for (uint32_t i = threadIdx.x; i < c.count; i += kThreads) stats_add_bins(s + 12, prims[c.begin + i], cmin, inv, live);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kBinWords; i += kThreads) {
    const uint32_t w = i % 7, v = s[12 + i];
    uint32_t* dst = g + 12 + i;
    if (w < 3) {
      if (v != kOrdPosInf) atomicMin(dst, v);
    } else if (w < 6) {
      if (v != kOrdNegInf) atomicMax(dst, v);
    } else if (v) {
      atomicAdd(dst, v);
    }
  }
}

// one wave per range (decide is a wave operation)
__global__ __launch_bounds__(kThreads) void tasks_decide(const Task* __restrict__ tasks, uint32_t count,
                                                         const uint32_t* __restrict__ stats, Decision* __restrict__ dec,
                                                         uint32_t* __restrict__ flags) {
  const uint32_t t = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (t >= count) return;  // wave-uniform
  const Decision d = decide(stats + (size_t)t * kStatWords, tasks[t].count);
  if ((threadIdx.x & 63u) == 0) {
    if (d.outcome == kForced && d.reorder) atomicOr(flags, 1u);
    dec[t] = d;
  }
}

// per chunk: the right-going primitives in front of mid and the left-going
// ones from mid on (counts[2 * chunk + 0 / 1])
__global__ __launch_bounds__(kThreads) void chunk_counts(const Chunk* __restrict__ chunks, const Task* __restrict__ tasks,
                                                         const Decision* __restrict__ dec, const Prim* __restrict__ prims,
                                                         uint32_t* __restrict__ counts) {
  __shared__ uint32_t s[2];
  const Chunk c = chunks[blockIdx.x];
  const Decision d = dec[c.task];
  if (threadIdx.x < 2) s[threadIdx.x] = 0;
  __syncthreads();
  if (d.outcome == kSplit) {
    const uint32_t mid = tasks[c.task].begin + d.nl;
    uint32_t fr = 0, bl = 0;
    for (uint32_t i = threadIdx.x; i < c.count; i += kThreads) {
      const uint32_t g = c.begin + i;
      const bool left = bin_of(prims[g].c[d.axis], d.cmin, d.inv) <= (int)d.bin;
      fr += (g < mid && !left) ? 1u : 0u;
      bl += (g >= mid && left) ? 1u : 0u;
    }
    atomicAdd(&s[0], fr);
    atomicAdd(&s[1], bl);
  }
  __syncthreads();
  if (threadIdx.x < 2) counts[2 * blockIdx.x + threadIdx.x] = s[threadIdx.x];
}

// per task: exclusive prefix of its chunks' counts (in place) and r
__global__ void tasks_scan(const Task* __restrict__ tasks, uint32_t count, const uint32_t* __restrict__ first_chunk,
                           uint32_t* __restrict__ counts, uint32_t* __restrict__ r_out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t a = 0, b = 0;
  for (uint32_t c = first_chunk[t]; c < first_chunk[t + 1]; ++c) {
    const uint32_t x = counts[2 * c], y = counts[2 * c + 1];
    counts[2 * c] = a;
    counts[2 * c + 1] = b;
    a += x;
    b += y;
  }
  r_out[t] = a;  // = b
}

// Block-wide exclusive scan of two counters per thread (kThreads threads):
// wave scans by shuffles, the waves' totals through LDS (2 * waves words).
__device__ inline void block_scan2(uint32_t* lds, uint32_t& a, uint32_t& b) {
  constexpr int kWaves = kThreads / 64;
  const int lane = (int)(threadIdx.x & 63u), w = (int)(threadIdx.x >> 6);
  uint32_t ia = a, ib = b;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t xa = __shfl_up(ia, off), xb = __shfl_up(ib, off);
    if (lane >= off) {
      ia += xa;
      ib += xb;
    }
  }
  if (lane == 63) {
    lds[w] = ia;
    lds[kWaves + w] = ib;
  }
  __syncthreads();
  uint32_t oa = 0, ob = 0;
  for (int k = 0; k < w; ++k) {
    oa += lds[k];
    ob += lds[kWaves + k];
  }
  __syncthreads();
  a = oa + ia - a;
  b = ob + ib - b;
}

// Each thread takes a contiguous run of the chunk; ranks go to rank[] and
// the right-going front / left-going back positions (range-relative) to
// pos_r / pos_l at the range's begin + rank.
__global__ __launch_bounds__(kThreads) void chunk_ranks(const Chunk* __restrict__ chunks, const Task* __restrict__ tasks,
                                                        const Decision* __restrict__ dec, const Prim* __restrict__ prims,
                                                        const uint32_t* __restrict__ counts, uint32_t* __restrict__ rank,
                                                        uint32_t* __restrict__ pos_r, uint32_t* __restrict__ pos_l) {
  __shared__ uint32_t lds[2 * kThreads];
  const Chunk c = chunks[blockIdx.x];
  const Decision d = dec[c.task];
  if (d.outcome != kSplit) return;  // uniform per block
  const Task tk = tasks[c.task];
  const uint32_t mid = tk.begin + d.nl;
  const uint32_t per = (c.count + kThreads - 1) / kThreads;
  const uint32_t i0 = min(c.count, threadIdx.x * per), i1 = min(c.count, i0 + per);
  uint32_t a = 0, b = 0;
  for (uint32_t i = i0; i < i1; ++i) {
    const uint32_t g = c.begin + i;
    const bool left = bin_of(prims[g].c[d.axis], d.cmin, d.inv) <= (int)d.bin;
    a += (g < mid && !left) ? 1u : 0u;
    b += (g >= mid && left) ? 1u : 0u;
  }
  block_scan2(lds, a, b);
  a += counts[2 * blockIdx.x];
  b += counts[2 * blockIdx.x + 1];
  for (uint32_t i = i0; i < i1; ++i) {
    const uint32_t g = c.begin + i;
    const bool left = bin_of(prims[g].c[d.axis], d.cmin, d.inv) <= (int)d.bin;
    if (g < mid && !left) {
      rank[g] = a;
      pos_r[tk.begin + a] = g - tk.begin;
      ++a;
    } else if (g >= mid && left) {
      rank[g] = b;
      pos_l[tk.begin + b] = g - tk.begin;
      ++b;
    }
  }
}

// pos_l holds the left-going back positions in front-to-back order; the
// closed form counts them from the back, so index r - 1 - j for the j-th
// (0-based) from the back.
__global__ __launch_bounds__(kThreads) void chunk_scatter(const Chunk* __restrict__ chunks, const Task* __restrict__ tasks,
                                                          const Decision* __restrict__ dec, const uint32_t* __restrict__ r_of,
                                                          const Prim* __restrict__ prims, const uint32_t* __restrict__ rank,
                                                          const uint32_t* __restrict__ pos_r, const uint32_t* __restrict__ pos_l,
                                                          Prim* __restrict__ tmp, uint32_t* __restrict__ flags) {
  const Chunk c = chunks[blockIdx.x];
  const Decision d = dec[c.task];
  if (d.outcome != kSplit) return;
  const Task tk = tasks[c.task];
  const uint32_t r = r_of[c.task];
  const uint32_t* pr = pos_r + tk.begin;
  const uint32_t* pl = pos_l + tk.begin;
  for (uint32_t i = threadIdx.x; i < c.count; i += kThreads) {
    const uint32_t ga = c.begin + i, g = ga - tk.begin;
    const Prim p = prims[ga];
    const bool left = bin_of(p.c[d.axis], d.cmin, d.inv) <= (int)d.bin;
    // rank[] is written only for the right-going front and left-going back primitives
    const bool ranked = (g < d.nl) != left;
    const uint32_t dst = partition_dest(left, g, d.nl, tk.count, r, ranked ? rank[ga] : 0u, pr, pl);
    if (dst < tk.count) tmp[tk.begin + dst] = p;
    else atomicOr(flags, 8u);
  }
}

__global__ __launch_bounds__(kThreads) void chunk_copy(const Chunk* __restrict__ chunks, const Decision* __restrict__ dec,
                                                       const Prim* __restrict__ tmp, Prim* __restrict__ prims) {
  const Chunk c = chunks[blockIdx.x];
  if (dec[c.task].outcome != kSplit) return;
  for (uint32_t i = threadIdx.x; i < c.count; i += kThreads) prims[c.begin + i] = tmp[c.begin + i];
}

struct BuildCounters {
  uint32_t nodes, large, small, pad;
};

// caps: build nodes, large tasks, small tasks (exceeding one sets flag 16)
struct Caps {
  uint32_t nodes, large, small, prims;
};
__global__ void tasks_emit(const Task* __restrict__ tasks, uint32_t count, const Decision* __restrict__ dec,
                           const uint32_t* __restrict__ stats, BNode* __restrict__ nodes, BuildCounters* __restrict__ ctr,
                           Task* __restrict__ large, Task* __restrict__ small, Caps caps, uint32_t* __restrict__ flags) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const Task tk = tasks[t];
  const Decision d = dec[t];
  int32_t id = -1;
  if (nodes_needed(tk, d)) {
    const uint32_t k = atomicAdd(&ctr->nodes, 1u);
    if (k >= caps.nodes) {
      atomicOr(flags, 16u);
      return;
    }
    id = (int32_t)k;
  }
  const Emit e = emit(tk, d, stats + (size_t)t * kStatWords, nodes, id);
  for (int k = 0; k < e.n; ++k) {
    const bool big = e.t[k].count > kSmall;
    const uint32_t j = atomicAdd(big ? &ctr->large : &ctr->small, 1u);
    if (j >= (big ? caps.large : caps.small)) {
      atomicOr(flags, 16u);
      continue;
    }
    (big ? large : small)[j] = e.t[k];
  }
}

// ---- phase 2: one block finishes a range of at most kSmall primitives ---------------

__global__ __launch_bounds__(kThreads) void subtree(const Task* __restrict__ roots, Prim* __restrict__ prims,
                                                    Prim* __restrict__ tmp, BNode* __restrict__ nodes,
                                                    BuildCounters* __restrict__ ctr, Caps caps,
                                                    uint32_t* __restrict__ flags) {
  __shared__ uint32_t s[kStatWords];
  __shared__ uint32_t scan[2 * kThreads];
  __shared__ uint32_t pos_r[kSmall], pos_l[kSmall];
  __shared__ Task stack[kStack];
  __shared__ Task cur;
  __shared__ Decision dcur;
  __shared__ int depth;
  __shared__ uint32_t r_tot;
  if (threadIdx.x == 0) {
    stack[0] = roots[blockIdx.x];
    depth = 1;
  }
  __syncthreads();
  while (true) {
    if (threadIdx.x == 0) cur = stack[--depth];
    stats_init(s, threadIdx.x, kThreads);
    __syncthreads();
    const Task tk = cur;
    range_boxes(s, prims + tk.begin, threadIdx.x, tk.count, kThreads);
    __syncthreads();
    if (tk.count > kLeafTarget) {
      float cmin[3], inv[3];
      bool live[3];
      for (int a = 0; a < 3; ++a) live[a] = axis_scale(s, a, &cmin[a], &inv[a]);
      for (uint32_t i = threadIdx.x; i < tk.count; i += kThreads)
        stats_add_bins(s + 12, prims[tk.begin + i], cmin, inv, live);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const Decision dd = decide(s, tk.count);
      if (threadIdx.x == 0) {
        dcur = dd;
        if (dd.outcome == kForced && dd.reorder) atomicOr(flags, 1u);
      }
    }
    __syncthreads();
    const Decision d = dcur;
    if (d.outcome == kSplit) {
      // each thread takes a contiguous run: its counts, their block scan,
      // the positions in rank order, then the closed-form scatter
      const uint32_t n = tk.count, per = (n + kThreads - 1) / kThreads;
      const uint32_t i0 = min(n, threadIdx.x * per), i1 = min(n, i0 + per);
      uint32_t a = 0, b = 0;
      for (uint32_t i = i0; i < i1; ++i) {
        const bool left = bin_of(prims[tk.begin + i].c[d.axis], d.cmin, d.inv) <= (int)d.bin;
        a += (i < d.nl && !left) ? 1u : 0u;
        b += (i >= d.nl && left) ? 1u : 0u;
      }
      const uint32_t a_run = a;
      block_scan2(scan, a, b);
      const uint32_t a0 = a, b0 = b;
      if (threadIdx.x == kThreads - 1) r_tot = a + a_run;
      for (uint32_t i = i0; i < i1; ++i) {
        const bool left = bin_of(prims[tk.begin + i].c[d.axis], d.cmin, d.inv) <= (int)d.bin;
        if (i < d.nl && !left) pos_r[a++] = i;
        else if (i >= d.nl && left) pos_l[b++] = i;
      }
      __syncthreads();
      const uint32_t r = r_tot;
      a = a0;
      b = b0;
      for (uint32_t i = i0; i < i1; ++i) {
        const Prim p = prims[tk.begin + i];
        const bool left = bin_of(p.c[d.axis], d.cmin, d.inv) <= (int)d.bin;
        uint32_t rk = 0;
        if (i < d.nl && !left) rk = a++;
        else if (i >= d.nl && left) rk = b++;
        const uint32_t dst = partition_dest(left, i, d.nl, n, r, rk, pos_r, pos_l);
        if (dst < n) tmp[tk.begin + dst] = p;
        else atomicOr(flags, 8u);
      }
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < n; i += kThreads) prims[tk.begin + i] = tmp[tk.begin + i];
    }
    if (threadIdx.x == 0) {
      int32_t id = -1;
      bool room = true;
      if (nodes_needed(tk, d)) {
        const uint32_t k = atomicAdd(&ctr->nodes, 1u);
        room = k < caps.nodes;
        id = (int32_t)k;
        if (!room) atomicOr(flags, 16u);
      }
      const Emit e = room ? emit(tk, d, s, nodes, id) : Emit{};
      // the larger range first, so the smaller is finished first (bounded stack)
      if (e.n == 2) {
        const int big = e.t[1].count > e.t[0].count ? 1 : 0;
        if (depth + 2 <= kStack) {
          stack[depth++] = e.t[big];
          stack[depth++] = e.t[1 - big];
        } else {
          atomicOr(flags, 2u);
        }
      }
    }
    __syncthreads();
    if (depth == 0) break;
    __syncthreads();
  }
}

// ---- phase 3: flattening -------------------------------------------------------------

__global__ void count_starts(const BNode* __restrict__ nodes, uint32_t count, uint32_t* __restrict__ starts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count && !nodes[i].leaf) atomicAdd(&starts[nodes[i].begin], 1u);
}

// exclusive scan in place over n words, in blocks of kScanBlock
constexpr uint32_t kScanBlock = 2048;
__global__ __launch_bounds__(kThreads) void scan_blocks(uint32_t* __restrict__ v, uint32_t n, uint32_t* __restrict__ sums) {
  __shared__ uint32_t lds[2 * kThreads];
  const uint32_t base = blockIdx.x * kScanBlock, per = kScanBlock / kThreads;
  const uint32_t i0 = base + threadIdx.x * per;
  uint32_t a = 0, b = 0;
  for (uint32_t i = i0; i < i0 + per && i < n; ++i) a += v[i];
  block_scan2(lds, a, b);
  for (uint32_t i = i0; i < i0 + per && i < n; ++i) {
    const uint32_t x = v[i];
    v[i] = a;
    a += x;
  }
  if (threadIdx.x == kThreads - 1) sums[blockIdx.x] = a;  // the block's total
}
__global__ void scan_sums(uint32_t* __restrict__ sums, uint32_t blocks) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t acc = 0;
  for (uint32_t b = 0; b < blocks; ++b) {
    const uint32_t x = sums[b];
    sums[b] = acc;
    acc += x;
  }
}
__global__ void scan_add(uint32_t* __restrict__ v, uint32_t n, const uint32_t* __restrict__ sums) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] += sums[i / kScanBlock];
}

__global__ void preorder(const BNode* __restrict__ nodes, uint32_t count, const uint32_t* __restrict__ starts,
                         uint32_t* __restrict__ order) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count || nodes[i].leaf) return;
  const uint32_t b = nodes[i].begin;
  uint32_t k = starts[b];
  for (int32_t p = nodes[i].parent; p >= 0 && nodes[p].begin == b; p = nodes[p].parent) ++k;
  order[i] = k;
}

__global__ void write_nodes(const BNode* __restrict__ nodes, uint32_t count, const uint32_t* __restrict__ order,
                            uint32_t interior, TptQbvhNode* __restrict__ out, uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count || nodes[i].leaf) return;
  if (order[i] >= interior) {
    atomicOr(flags, 8u);
    return;
  }
  TptQbvhNode q;
  for (int k = 0; k < 4; ++k) {  // QbvhNode::empty, R/qbvh.rs:79-90
    q.min_x[k] = q.min_y[k] = q.min_z[k] = INFINITY;
    q.max_x[k] = q.max_y[k] = q.max_z[k] = -INFINITY;
    q.children[k] = kChildLeafBit;
  }
  q.valid_mask = 0;
  for (int k = 0; k < (int)sizeof(q.pad); ++k) q.pad[k] = 0;
  int slot = 0;
  for (int k = 0; k < 4; ++k) {
    const int32_t c = nodes[i].child[k];
    if (c < 0) continue;
    const BNode& cn = nodes[c];
    q.min_x[slot] = cn.lo[0];
    q.min_y[slot] = cn.lo[1];
    q.min_z[slot] = cn.lo[2];
    q.max_x[slot] = cn.hi[0];
    q.max_y[slot] = cn.hi[1];
    q.max_z[slot] = cn.hi[2];
    if (cn.leaf) {
      if (cn.count > kChildCountMask || cn.begin > kChildOffsetMask) atomicOr(flags, 4u);
      q.children[slot] = kChildLeafBit | (cn.count << kChildCountShift) | (cn.begin & kChildOffsetMask);
    } else {
      q.children[slot] = order[c];
    }
    q.valid_mask |= (uint8_t)(1u << slot);
    ++slot;
  }
  out[order[i]] = q;
}

__global__ void prim_order(const Prim* __restrict__ prims, uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = prims[i].index;
}

inline uint32_t blocks_for(size_t n, uint32_t per) { return (uint32_t)((n + per - 1) / per); }

}  // namespace

// Returns 0, or -2 when a leaf exceeds the child encoding, -3 for a forced
// split the GPU build does not reorder (non-coincident centroids), -4 for a
// device error (tpt::set_error has the message).
int build_qbvh_device(TptDevice* d, const float* bmin, const float* bmax, size_t n, bool device_pointers,
                      HostQbvh* out) {
  out->nodes.clear();
  out->prim.clear();
  if (n == 0) return 0;
  if (n > 0x7FFFFFFFull) return -1;
  hipStream_t st = d->stream;
  const uint32_t N = (uint32_t)n;
  const size_t cap_large = n / kSmall + 16;
  const size_t cap_chunks = n / kChunk + cap_large + 16;
  const size_t cap_nodes = 2 * n + 16;
  const size_t cap_scan = blocks_for(n + 1, kScanBlock) + 1;
  // one allocation, carved
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~size_t(255);
    return o;
  };
  const size_t o_prims = carve(sizeof(Prim) * n), o_tmp = carve(sizeof(Prim) * n);
  const size_t o_bounds = device_pointers ? 0 : carve(24 * n);
  const size_t o_rank = carve(4 * n), o_posr = carve(4 * n), o_posl = carve(4 * n);
  const size_t o_starts = carve(4 * (n + 1)), o_sums = carve(4 * cap_scan);
  const size_t o_nodes = carve(sizeof(BNode) * cap_nodes), o_order = carve(4 * cap_nodes);
  const size_t o_large0 = carve(sizeof(Task) * cap_large), o_large1 = carve(sizeof(Task) * cap_large);
  const size_t o_small = carve(sizeof(Task) * (n + 1));
  const size_t o_stats = carve(4ull * kStatWords * cap_large), o_dec = carve(sizeof(Decision) * cap_large);
  const size_t o_chunks = carve(sizeof(Chunk) * cap_chunks), o_counts = carve(8 * cap_chunks);
  const size_t o_first = carve(4 * (cap_large + 1)), o_r = carve(4 * cap_large);
  const size_t o_ctr = carve(sizeof(BuildCounters)), o_flags = carve(4);
  const size_t o_qnodes = carve(sizeof(TptQbvhNode) * (n + 1)), o_pidx = carve(4 * n);
  // the device's build workspace, kept (grow-only) for the next build
  if (!d->qbvh_buf.ensure(off)) return -4;
  char* base = static_cast<char*>(d->qbvh_buf.ptr);
  auto at = [&](size_t o) { return static_cast<void*>(base + o); };
  Prim* prims = (Prim*)at(o_prims);
  Prim* tmp = (Prim*)at(o_tmp);
  uint32_t* rank = (uint32_t*)at(o_rank);
  uint32_t* pos_r = (uint32_t*)at(o_posr);
  uint32_t* pos_l = (uint32_t*)at(o_posl);
  uint32_t* starts = (uint32_t*)at(o_starts);
  uint32_t* sums = (uint32_t*)at(o_sums);
  BNode* nodes = (BNode*)at(o_nodes);
  uint32_t* order = (uint32_t*)at(o_order);
  Task* large[2] = {(Task*)at(o_large0), (Task*)at(o_large1)};
  Task* small = (Task*)at(o_small);
  uint32_t* stats = (uint32_t*)at(o_stats);
  Decision* dec = (Decision*)at(o_dec);
  Chunk* chunks = (Chunk*)at(o_chunks);
  uint32_t* counts = (uint32_t*)at(o_counts);
  uint32_t* first = (uint32_t*)at(o_first);
  uint32_t* r_of = (uint32_t*)at(o_r);
  BuildCounters* ctr = (BuildCounters*)at(o_ctr);
  uint32_t* flags = (uint32_t*)at(o_flags);
  TptQbvhNode* qnodes = (TptQbvhNode*)at(o_qnodes);
  uint32_t* pidx = (uint32_t*)at(o_pidx);

  const Caps caps{(uint32_t)cap_nodes, (uint32_t)cap_large, N + 1, N};
  const float* dmin = bmin;
  const float* dmax = bmax;
  if (!device_pointers) {
    float* b = (float*)at(o_bounds);
    if (!hip_ok(hipMemcpyAsync(b, bmin, 12 * n, hipMemcpyHostToDevice, st), "hipMemcpyAsync(bounds)") ||
        !hip_ok(hipMemcpyAsync(b + 3 * n, bmax, 12 * n, hipMemcpyHostToDevice, st), "hipMemcpyAsync(bounds)"))
      return -4;
    dmin = b;
    dmax = b + 3 * n;
  }
  hipLaunchKernelGGL(init_prims, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, dmin, dmax, N, prims);
  if (!hip_ok(hipMemsetAsync(ctr, 0, sizeof(BuildCounters), st), "hipMemsetAsync") ||
      !hip_ok(hipMemsetAsync(flags, 0, 4, st), "hipMemsetAsync"))
    return -4;

  // phase 1
  Task root{0, N, kNode, -1, 0, 0};
  uint32_t n_large = 0, n_small = 0;
  if (N > kSmall) {
    if (!hip_ok(hipMemcpyAsync(large[0], &root, sizeof(Task), hipMemcpyHostToDevice, st), "hipMemcpyAsync")) return -4;
    n_large = 1;
  } else {
    if (!hip_ok(hipMemcpyAsync(small, &root, sizeof(Task), hipMemcpyHostToDevice, st), "hipMemcpyAsync")) return -4;
    n_small = 1;
    BuildCounters c0{0, 0, 1, 0};
    if (!hip_ok(hipMemcpyAsync(ctr, &c0, sizeof(c0), hipMemcpyHostToDevice, st), "hipMemcpyAsync")) return -4;
  }
  std::vector<Task> host_tasks;
  std::vector<Chunk> host_chunks;
  std::vector<uint32_t> host_first;
  int cur = 0;
  while (n_large) {
    host_tasks.resize(n_large);
    if (!hip_ok(hipMemcpyAsync(host_tasks.data(), large[cur], sizeof(Task) * n_large, hipMemcpyDeviceToHost, st),
                "hipMemcpyAsync") ||
        !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
      return -4;
    host_chunks.clear();
    host_first.assign(n_large + 1, 0);
    for (uint32_t t = 0; t < n_large; ++t) {
      host_first[t] = (uint32_t)host_chunks.size();
      for (uint32_t b = 0; b < host_tasks[t].count; b += kChunk)
        host_chunks.push_back(Chunk{t, host_tasks[t].begin + b, std::min(kChunk, host_tasks[t].count - b), 0});
    }
    host_first[n_large] = (uint32_t)host_chunks.size();
    const uint32_t nc = (uint32_t)host_chunks.size();
    if (nc > cap_chunks) {
      set_error("build_qbvh_device: chunk table overflow");
      return -4;
    }
    if (!hip_ok(hipMemcpyAsync(chunks, host_chunks.data(), sizeof(Chunk) * nc, hipMemcpyHostToDevice, st),
                "hipMemcpyAsync") ||
        !hip_ok(hipMemcpyAsync(first, host_first.data(), 4 * (n_large + 1), hipMemcpyHostToDevice, st),
                "hipMemcpyAsync"))
      return -4;
    Task* tasks = large[cur];
    hipLaunchKernelGGL(stats_reset, dim3(std::min<uint32_t>(n_large, 4096)), dim3(kThreads), 0, st, stats, n_large);
    hipLaunchKernelGGL(chunk_boxes, dim3(nc), dim3(kThreads), 0, st, chunks, prims, stats);
    hipLaunchKernelGGL(chunk_bins, dim3(nc), dim3(kThreads), 0, st, chunks, prims, stats);
    hipLaunchKernelGGL(tasks_decide, dim3(blocks_for(n_large, kThreads / 64)), dim3(kThreads), 0, st, tasks, n_large,
                       stats, dec, flags);
    hipLaunchKernelGGL(chunk_counts, dim3(nc), dim3(kThreads), 0, st, chunks, tasks, dec, prims, counts);
    hipLaunchKernelGGL(tasks_scan, dim3(blocks_for(n_large, 64)), dim3(64), 0, st, tasks, n_large, first, counts, r_of);
    hipLaunchKernelGGL(chunk_ranks, dim3(nc), dim3(kThreads), 0, st, chunks, tasks, dec, prims, counts, rank, pos_r,
                       pos_l);
    hipLaunchKernelGGL(chunk_scatter, dim3(nc), dim3(kThreads), 0, st, chunks, tasks, dec, r_of, prims, rank, pos_r,
                       pos_l, tmp, flags);
    hipLaunchKernelGGL(chunk_copy, dim3(nc), dim3(kThreads), 0, st, chunks, dec, tmp, prims);
    // next level: reset the large counter, emit
    if (!hip_ok(hipMemsetAsync(&ctr->large, 0, 4, st), "hipMemsetAsync")) return -4;
    hipLaunchKernelGGL(tasks_emit, dim3(blocks_for(n_large, 64)), dim3(64), 0, st, tasks, n_large, dec, stats, nodes,
                       ctr, large[1 - cur], small, caps, flags);
    BuildCounters c;
    if (!hip_ok(hipMemcpyAsync(&c, ctr, sizeof(c), hipMemcpyDeviceToHost, st), "hipMemcpyAsync") ||
        !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
      return -4;
    if (c.large > cap_large || c.small > n + 1) {
      set_error("build_qbvh_device: task table overflow");
      return -4;
    }
    n_large = c.large;
    n_small = c.small;
    cur = 1 - cur;
  }
  // phase 2
  if (n_small)
    hipLaunchKernelGGL(subtree, dim3(n_small), dim3(kThreads), 0, st, small, prims, tmp, nodes, ctr, caps, flags);
  BuildCounters c;
  uint32_t fl = 0;
  if (!hip_ok(hipMemcpyAsync(&c, ctr, sizeof(c), hipMemcpyDeviceToHost, st), "hipMemcpyAsync") ||
      !hip_ok(hipMemcpyAsync(&fl, flags, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync") ||
      !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
    return -4;
  if (fl & 26u) {
    set_error((fl & 2u) ? "build_qbvh_device: subtree stack overflow"
                        : (fl & 16u) ? "build_qbvh_device: table overflow" : "build_qbvh_device: partition error");
    return -4;
  }
  if (fl & 1u) return -3;
  const uint32_t n_nodes = c.nodes;
  if (n_nodes > cap_nodes) {
    set_error("build_qbvh_device: node table overflow");
    return -4;
  }
  // phase 3
  if (!hip_ok(hipMemsetAsync(starts, 0, 4 * (n + 1), st), "hipMemsetAsync")) return -4;
  hipLaunchKernelGGL(count_starts, dim3(blocks_for(n_nodes, kThreads)), dim3(kThreads), 0, st, nodes, n_nodes, starts);
  const uint32_t sb = blocks_for(n + 1, kScanBlock);
  hipLaunchKernelGGL(scan_blocks, dim3(sb), dim3(kThreads), 0, st, starts, N + 1, sums);
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(64), 0, st, sums, sb);
  hipLaunchKernelGGL(scan_add, dim3(blocks_for(n + 1, kThreads)), dim3(kThreads), 0, st, starts, N + 1, sums);
  hipLaunchKernelGGL(preorder, dim3(blocks_for(n_nodes, kThreads)), dim3(kThreads), 0, st, nodes, n_nodes, starts, order);
  uint32_t interior = 0;
  if (!hip_ok(hipMemcpyAsync(&interior, starts + n, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync") ||
      !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
    return -4;
  if (interior > n + 1) {
    set_error("build_qbvh_device: interior node count");
    return -4;
  }
  hipLaunchKernelGGL(write_nodes, dim3(blocks_for(n_nodes, kThreads)), dim3(kThreads), 0, st, nodes, n_nodes, order,
                     interior, qnodes, flags);
  hipLaunchKernelGGL(prim_order, dim3(blocks_for(n, kThreads)), dim3(kThreads), 0, st, prims, N, pidx);
  out->nodes.resize(interior);
  out->prim.resize(n);
  if ((interior && !hip_ok(hipMemcpyAsync(out->nodes.data(), qnodes, sizeof(TptQbvhNode) * interior,
                                          hipMemcpyDeviceToHost, st),
                           "hipMemcpyAsync")) ||
      !hip_ok(hipMemcpyAsync(out->prim.data(), pidx, 4 * n, hipMemcpyDeviceToHost, st), "hipMemcpyAsync") ||
      !hip_ok(hipMemcpyAsync(&fl, flags, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync") ||
      !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
    return -4;
  if (fl & 8u) {
    set_error("build_qbvh_device: node order error");
    return -4;
  }
  if (fl & 4u) return -2;
  return 0;
}

}  // namespace tpt
