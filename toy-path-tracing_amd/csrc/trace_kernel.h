// The persistent trace kernel (closest hit, R/scene/mod.rs:340-395 over
// R/qbvh.rs:611-675 and R/math/ray.rs:107-209; shadow visibility,
// R/integrator/mod.rs:239-265), shared by the C ABI's trace entry points.
#pragma once

#include <hip/hip_runtime.h>

#include "tpt_device.h"
#include "tpt_internal.h"
#include "traverse.h"

namespace tpt {

// Work distribution.  Rays are split into `regions` contiguous ranges (8 by
// default: one per XCD (workgroups are dispatched round-robin over the 8 XCDs, so block b runs
// on XCD b % 8): neighbouring rays stay on one XCD's L2.  A wave refills its
// idle lanes from its range with one atomic per refill; once its range is
// exhausted it moves on to the next XCD's range.  The last wave to finish
// resets the counters, so consecutive launches on one stream need no clear.

struct Query {
  // closest: hits; any: target, distance, out
  TptHit* hits;
  const uint32_t* target;
  const float* distance;
  uint8_t* out;
  uint32_t* counts;
  uint32_t* status;
  WorkCounters* work;
  const uint32_t* n_dev;  // when set, the ray count is read here (device-resident queue length)
  uint32_t n;
  uint32_t refill_below;  // refill when fewer lanes than this are busy
  uint32_t grab;          // ray ids taken per work-counter atomic (at least the idle lanes)
  uint32_t regions;       // work regions, 1..kMaxRegions
  uint32_t lds_depth;     // LDS stack entries per lane (the launch's dynamic LDS is lds_depth * kBlock words)
};

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

template <bool kQ8, int kMode, bool kExact, bool kCount, bool kTwoLevel, class Stack>
__device__ __forceinline__ void begin_ray(const SceneView& sc, const TptRay* rays, const Query& q, uint32_t rid,
                                          Lane& L, Stack& st) {
  F3 o, d;
  float tmax;
  load_ray(rays, rid, &o, &d, &tmax);
  if (kMode == kClosest) {
    lane_begin<kTwoLevel>(sc, L, st, o, d, tmax, __builtin_inff(), TPT_INVALID_INDEX, TPT_INVALID_INDEX);
    return;
  }
  // unoccluded_ray (R/integrator/mod.rs:239-265)
  uint32_t ti = TPT_INVALID_INDEX, tt = TPT_INVALID_INDEX;
  if (q.target) {
    ti = q.target[2 * rid];
    tt = q.target[2 * rid + 1];
    if (ti >= sc.instance_count) ti = TPT_INVALID_INDEX;
  }
  if (kExact) {  // the reference's own form: a full closest-hit query
    lane_begin<kTwoLevel>(sc, L, st, o, d, __builtin_inff(), __builtin_inff(), TPT_INVALID_INDEX, TPT_INVALID_INDEX);
    // the closest-hit traversal never reads skip_*: they keep the target
    // for the decision in finish_ray
    L.skip_inst = ti;
    L.skip_tri = tt;
    return;
  }
  const float dist = q.distance ? q.distance[rid] : __builtin_inff();
  float bound = __builtin_inff(), strict = __builtin_inff();
  if (ti != TPT_INVALID_INDEX) {
    // The target's own t bounds the query: any other hit nearer than it
    // makes the target not the closest hit.
    const GpuInstance in = sc.inst[ti];
    if (tt < in.tri_count) {
      Ray3 lr;
      to_local(in, o, d, &lr);
      const GpuTri* tp = sc.tris + in.tri_base + sc.tri_rank[in.tri_base + tt];
      TriResult h;
      if (!(tp->prim & kTriDegenerate) &&
          intersect_triangle(lr, __builtin_inff(), f3(tp->v[0], tp->v[1], tp->v[2]),
                             f3(tp->v[3], tp->v[4], tp->v[5]), f3(tp->v[6], tp->v[7], tp->v[8]), &h))
        bound = h.t;
    }
  } else if (!__builtin_isinf(dist)) {
    bound = dist * (1.0f - 1.0e-3f);  // SHADOW_TOLERANCE, R/integrator/mod.rs:50
    strict = bound;                   // occluded iff some hit has t < bound
  }
  lane_begin<kTwoLevel>(sc, L, st, o, d, bound, strict, ti, tt);
}

template <int kMode, bool kExact, bool kCount>
__device__ __forceinline__ void finish_ray(const Query& q, uint32_t rid, const Lane& L, const Counters& cnt) {
  if (kMode == kClosest) {
    TptHit h;
    h.instance = L.inst;
    h.triangle = L.tri;
    h.t = L.t;
    h.b[0] = L.b0;
    h.b[1] = L.b1;
    h.b[2] = L.b2;
    if (q.hits) q.hits[rid] = h;
  } else {
    uint8_t unocc;
    if (kExact) {
      const float dist = q.distance ? q.distance[rid] : __builtin_inff();
      if (!L.hit)
        unocc = 1;
      else if (L.skip_inst != TPT_INVALID_INDEX)
        unocc = (L.inst == L.skip_inst && L.tri == L.skip_tri) ? 1 : 0;
      else if (__builtin_isinf(dist))
        unocc = 0;
      else
        unocc = L.t >= dist * (1.0f - 1.0e-3f) ? 1 : 0;
    } else {
      unocc = L.hit ? 0 : 1;
    }
    if (q.out) q.out[rid] = unocc;
  }
  if (kCount) {
    q.counts[3 * rid] = cnt.nodes;
    q.counts[3 * rid + 1] = cnt.tris;
    q.counts[3 * rid + 2] = cnt.insts;
  }
}

// Persistent trace kernel: every wave loops, refilling idle lanes with new
// rays, until all rays are done.  Exits are wave-uniform.
// kLds: the LDS stack depth as a constant with a static LDS array (the
// default depth), or 0 for the tuned depth q.lds_depth in LDS sized at
// launch.
template <bool kQ8, int kMode, bool kExact, bool kCount, bool kTwoLevel, int kLds>
__global__ __launch_bounds__(kBlock) void trace_kernel(SceneView sc, const TptRay* __restrict__ rays, Query q) {
  uint32_t* lds;
  if constexpr (kLds > 0) {
    __shared__ uint32_t static_lds[kLds * kBlock];
    lds = static_lds;
  } else {
    extern __shared__ uint32_t dynamic_lds[];  // q.lds_depth * kBlock entries
    lds = dynamic_lds;
  }
  uint32_t spill[kStackDepth - (kLds > 0 ? kLds : kMinLdsDepth)];
  LaneStackT<kLds> st;
  st.lds = lds + threadIdx.x;
  st.spill = spill;
  st.sp = 0;
  st.lds_depth = kLds > 0 ? kLds : (int)q.lds_depth;
  st.overflow = false;
  Lane L;
  Counters cnt = {0, 0, 0};
  const uint32_t lane = lane_id();
  const uint32_t n = q.n_dev ? *q.n_dev : q.n;
  const uint32_t chunk = (n + q.regions - 1) / q.regions;
  uint32_t region = blockIdx.x % q.regions;
  uint32_t regions_left = q.regions;
  bool active = false;
  bool overflow = false;
  uint32_t rid = 0;
  uint32_t res_lo = 0, res_hi = 0;  // wave-uniform reserve: ray ids taken, not yet started
  for (;;) {
    uint64_t busy = __ballot(active);
    const uint32_t rem = res_hi - res_lo;
    if ((regions_left > 0 || rem > 0) && (uint32_t)__popcll(busy) < q.refill_below) {
      const uint64_t idle = ~busy;
      const uint32_t want = (uint32_t)__popcll(idle);
      uint32_t nb = 0, ne = 0;  // newly taken ids [nb, ne)
      const bool take = rem < want && regions_left > 0;
      if (take) {
        const uint32_t lo = (uint32_t)region * chunk;
        const uint32_t hi = min(n, lo + chunk);
        const uint32_t grab = max(want - rem, q.grab);
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(&q.work->next[region].v, grab);
        b = __shfl(b, 0);
        const uint64_t end = (uint64_t)lo + b + grab;
        nb = (uint32_t)min<uint64_t>((uint64_t)lo + b, hi);
        ne = (uint32_t)min<uint64_t>(end, hi);
        if (end >= hi) {  // this range is used up: move to the next XCD's
          region = region + 1 == q.regions ? 0u : region + 1;
          regions_left--;
        }
      }
      if (!active) {
        // lanes of rank < rem start reserved ids, the next ones new ids
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
        bool got = false;
        if (rank < rem) {
          rid = res_lo + rank;
          got = true;
        } else if (rank - rem < ne - nb) {
          rid = nb + (rank - rem);
          got = true;
        }
        if (got) {
          begin_ray<kQ8, kMode, kExact, kCount, kTwoLevel>(sc, rays, q, rid, L, st);
          if (kCount) cnt = Counters{0, 0, 0};
          active = true;
        }
      }
      const uint32_t k1 = min(want, rem);
      res_lo += k1;
      if (take) {
        res_lo = nb + min(want - k1, ne - nb);
        res_hi = ne;
      }
      busy = __ballot(active);
    }
    if (busy == 0) {
      if (regions_left == 0 && res_hi == res_lo) break;
      continue;
    }
    if (active) {
      // the exact shadow rule needs the full closest-hit traversal
      if (lane_step<kQ8, (kExact ? kClosest : kMode), kCount, kTwoLevel>(sc, L, st, cnt)) {
        finish_ray<kMode, kExact, kCount>(q, rid, L, cnt);
        overflow |= st.overflow;
        active = false;
      }
    }
  }
  if (__ballot(overflow) != 0 && lane == 0) atomicOr(q.status, 1u);
  // the last wave to finish resets the work counters for the next launch
  if (lane == 0) {
    __threadfence();
    const uint32_t waves = gridDim.x * (kBlock / 64);
    if (atomicAdd(&q.work->done_waves.v, 1u) == waves - 1) {
      for (uint32_t r = 0; r < q.regions; ++r) atomicExch(&q.work->next[r].v, 0u);
      atomicExch(&q.work->done_waves.v, 0u);
    }
  }
}

}  // namespace tpt
