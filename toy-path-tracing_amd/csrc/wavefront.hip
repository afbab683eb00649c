// Wavefront path tracer: the estimator of render.hip (R/main.rs:98-126 with
// R/integrator/mis.rs:19-139) reorganised into small kernels over queues of
// paths, for many frames of different resolution in one call.
//
// A pass traces samples [k*S, (k+1)*S) of every pixel of every frame: P =
// pixels * S paths with their state in structure-of-arrays buffers.  Per
// bounce b:
//   trace     closest hit of the ray queue (the persistent trace kernel;
//             queue length read on the device)
//   shade     emission at the hit (MIS against the previous vertex), then for
//             b < max_depth: light-tree NEE -> shadow queue, material sample,
//             Russian roulette -> next ray queue (wave-compacted appends)
//   shadow    the shadow queue through the any-hit trace kernel: the
//             target's own t bounds the query and the first other hit ends
//             it; with TPT_FLAG_ANY_EXACT the reference's rule literally
//             (closest hit, then the target test: unoccluded_ray)
//   add       unoccluded NEE contributions into the path radiance
// then `fold` adds the pass's samples to each pixel's running mean in sample
// order (color += (s - color) / n, R/main.rs:116-117).  The arithmetic per
// path is the megakernel's, in the same order, so with the exact shadow rule
// the image is the megakernel's bit for bit (the early-exit rule differs
// only when another triangle is hit at exactly the light sample's t); only
// the schedule differs: lanes are refilled with live paths at every bounce.
//
// Every launch of a pass reads its counts from device memory, so the pass is
// captured once as a HIP graph and replayed for every pass (no host round
// trip between bounces).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "shading.h"
#include "tpt_device.h"

namespace tpt {
namespace wf {

struct FrameDev {
  pt::CamView cam;
  uint32_t w, h, spp, max_depth;
  smp::ZSobolFrame zf;
  uint32_t pix_begin;  // first global pixel of this frame
  uint32_t pad;
  float* out;          // device rgb, w * h * 3
};

// Queue lengths, one per 256-byte line: every block appends with a
// device-scope atomic, and counters sharing a line serialise on it.
struct alignas(256) CountLine {
  uint32_t v;
  uint32_t pad[63];
};
struct Counters {
  CountLine rays[2];  // ping-pong ray-queue lengths
  CountLine shadow;   // shadow-queue length
  CountLine pass;     // current pass
};

struct State {
  float* thr;       // 3 per path
  float* rad;       // 3 per path
  float* prev_p;    // 3 per path: previous shading vertex position
  float* prev_ns;   // 3 per path: its shading normal
  float* prev_pdf;  // BSDF pdf of the sample that left it
  uint32_t* prev_info;  // material index | delta flag (bit 31)
  // the rest of its light-tree query (TreeQuery): 4 per path (diffuse and
  // glossy weights, glossy roughness); with the OpenPBR proxy policy also the
  // frame tangent and the view direction (3 each; null otherwise), written
  // when the glossy weight is positive
  float* prev_q;
  float* prev_t;
  float* prev_wo;
  // textured scenes: the current ray's differential (12 arrays of one float
  // per path: rx origin, ry origin, rx direction, ry direction), whether it
  // has one, and its cone (2 arrays: width, spread); null otherwise
  float* diff;
  uint32_t* dflag;
  float* cone;
};

struct Queue {
  TptRay* rays;
  uint32_t* path;
};

struct ShadowQ {
  TptRay* rays;
  uint32_t* target;  // (instance, triangle) per entry
  float* value;      // 3 per entry: throughput x unoccluded contribution
  uint32_t* path;
  uint8_t* unocc;
};

struct Pass {
  const FrameDev* frames;
  uint32_t frame_count;
  uint32_t spp_slots;     // S: samples per pixel per pass
  uint32_t total_pixels;  // over all frames
  uint32_t paths;         // total_pixels * S
};

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Block-aggregated append to up to two queues: every thread of the block
// calls it (block-uniform control flow); threads with want_a (want_b) get
// consecutive slots of queue a (b).  One global atomic per queue per block
// call: the queue counters are single words updated by every block on all
// XCDs, and appends of one atomic per wave serialise on them.  kTile: the
// block's threads (the queue tile: one queue entry per thread).
struct AppendSlots {
  uint32_t a, b;
};
template <uint32_t kTile>
__device__ __forceinline__ AppendSlots block_append(uint32_t* ca, bool want_a, uint32_t* cb, bool want_b) {
  constexpr uint32_t kWaves = kTile / 64;
  __shared__ uint32_t cnt[2][kWaves];
  __shared__ uint32_t base[2];
  const uint64_t ma = __ballot(want_a), mb = __ballot(want_b);
  const uint32_t wave = threadIdx.x / 64, lane = lane_id();
  if (lane == 0) {
    cnt[0][wave] = (uint32_t)__popcll(ma);
    cnt[1][wave] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ta = 0, tb = 0;
    for (uint32_t w = 0; w < kWaves; ++w) {
      ta += cnt[0][w];
      tb += cnt[1][w];
    }
    base[0] = ta ? atomicAdd(ca, ta) : 0u;
    base[1] = tb ? atomicAdd(cb, tb) : 0u;
  }
  __syncthreads();
  AppendSlots r{base[0], base[1]};
  for (uint32_t w = 0; w < wave; ++w) {
    r.a += cnt[0][w];
    r.b += cnt[1][w];
  }
  r.a += __builtin_amdgcn_mbcnt_hi((uint32_t)(ma >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ma, 0u));
  r.b += __builtin_amdgcn_mbcnt_hi((uint32_t)(mb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mb, 0u));
  __syncthreads();  // cnt and base are rewritten by the next call
  return r;
}

__device__ __forceinline__ uint32_t frame_of(const Pass& ps, uint32_t gp) {
  uint32_t f = 0;
  while (f + 1 < ps.frame_count && gp >= ps.frames[f + 1].pix_begin) ++f;
  return f;
}

__device__ __forceinline__ void store3(float* a, uint32_t i, F3 v) {
  a[3 * i] = v.x;
  a[3 * i + 1] = v.y;
  a[3 * i + 2] = v.z;
}
__device__ __forceinline__ F3 load3f(const float* a, uint32_t i) { return f3(a[3 * i], a[3 * i + 1], a[3 * i + 2]); }

__device__ __forceinline__ void put_ray(TptRay* r, uint32_t slot, F3 o, F3 d) {
  TptRay v;
  v.o[0] = o.x;
  v.o[1] = o.y;
  v.o[2] = o.z;
  v.tmax = __builtin_inff();
  v.d[0] = d.x;
  v.d[1] = d.y;
  v.d[2] = d.z;
  v.pad = 0.0f;
  r[slot] = v;
}

// Resets one ray-queue length and the shadow-queue length.
__global__ void wf_reset(Counters* c, uint32_t which) {
  if (threadIdx.x == 0) {
    c->rays[which].v = 0;
    c->shadow.v = 0;
  }
}

__global__ void wf_next_pass(Counters* c) {
  if (threadIdx.x == 0) c->pass.v += 1;
}

// Digit-permutation rows (smp::zsobol_index) of dimensions 0..dims-1.
__global__ void wf_ztab(uint8_t* __restrict__ tab, uint32_t dims) {
  const uint32_t n = dims * smp::kZTabPrefixes;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t dim = i / smp::kZTabPrefixes, v = i % smp::kZTabPrefixes;
    tab[i] = (uint8_t)smp::digit_perm_byte(v, 0x55555555ull * dim);
  }
}

// Row `dim` of the digit-permutation table into LDS: copied when the table
// holds it, computed otherwise.  All threads of the block call it.
__device__ __forceinline__ void stage_row(uint8_t* lds, const uint8_t* __restrict__ tab, uint32_t tab_dims,
                                          uint32_t dim) {
  if (dim < tab_dims) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(tab + (size_t)dim * smp::kZTabPrefixes);
    uint32_t* dst = reinterpret_cast<uint32_t*>(lds);
    for (uint32_t i = threadIdx.x; i < smp::kZTabPrefixes / 4; i += blockDim.x) dst[i] = src[i];
  } else {
    for (uint32_t v = threadIdx.x; v < smp::kZTabPrefixes; v += blockDim.x)
      lds[v] = (uint8_t)smp::digit_perm_byte(v, 0x55555555ull * dim);
  }
}

// The per-path ray differential and cone of textured scenes (SoA).
__device__ __forceinline__ void store_diff(const State& st, uint32_t P, uint32_t path, const pt::RayDiff& r) {
  st.dflag[path] = r.has ? 1u : 0u;
  if (r.has) {
    const F3 v[4] = {r.rxo, r.ryo, r.rxd, r.ryd};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      st.diff[(size_t)(3 * k) * P + path] = v[k].x;
      st.diff[(size_t)(3 * k + 1) * P + path] = v[k].y;
      st.diff[(size_t)(3 * k + 2) * P + path] = v[k].z;
    }
  }
  st.cone[path] = r.cw;
  st.cone[(size_t)P + path] = r.cs;
}
__device__ __forceinline__ void load_diff(const State& st, uint32_t P, uint32_t path, pt::RayDiff* r) {
  r->has = st.dflag[path] != 0u;
  if (r->has) {
    F3 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      v[k] = f3(st.diff[(size_t)(3 * k) * P + path], st.diff[(size_t)(3 * k + 1) * P + path],
                st.diff[(size_t)(3 * k + 2) * P + path]);
    r->rxo = v[0];
    r->ryo = v[1];
    r->rxd = v[2];
    r->ryd = v[3];
  }
  r->cw = st.cone[path];
  r->cs = st.cone[(size_t)P + path];
}

// Camera rays of the pass's samples (PathSampler camera dimension,
// PinholeCamera::generate_ray_differential), state init, queue 0 append.
template <bool kTex, uint32_t kTile>
__global__ __launch_bounds__(kTile) void wf_generate(Pass ps, State st, Queue q, Counters* c,
                                                            const uint8_t* __restrict__ ztab, uint32_t ztab_dims) {
  // One path slot per thread (the host launches ceil(paths / block) blocks):
  // no thread carries state from one slot to another.
  const uint32_t base = blockIdx.x * blockDim.x;
  if (base >= ps.paths) return;  // uniform per block, before any barrier
  __shared__ __align__(16) uint8_t zrow[smp::kZTabPrefixes];
  stage_row(zrow, ztab, ztab_dims, smp::kCameraDim);
  __syncthreads();
  const uint32_t k = c->pass.v;
  {
    const uint32_t path = base + threadIdx.x;
    bool want = false;
    F3 o = f3(0.0f, 0.0f, 0.0f), d = o;
    if (path < ps.paths) {
      const uint32_t gp = path / ps.spp_slots, s = path % ps.spp_slots;
      const FrameDev& fr = ps.frames[frame_of(ps, gp)];
      const uint32_t si = k * ps.spp_slots + s;
      store3(st.rad, path, f3(0.0f, 0.0f, 0.0f));
      if (si < fr.spp) {
        const uint32_t lp = gp - fr.pix_begin, x = lp % fr.w, y = lp / fr.w;
        const smp::PathSampler smpl{fr.zf, x, y, si};
        o = fr.cam.eye;
        if (kTex) {
          const pt::RayDiff r = pt::camera_ray_diff(fr.cam, smpl, x, y, fr.w, fr.h, fr.spp, zrow);
          d = r.d;
          store_diff(st, ps.paths, path, r);
        } else {
          d = pt::camera_dir(fr.cam, smpl, x, y, fr.w, fr.h, zrow);
        }
        store3(st.thr, path, f3(1.0f, 1.0f, 1.0f));
        want = true;
      }
    }
    const uint32_t slot = block_append<kTile>(&c->rays[0].v, want, &c->rays[0].v, false).a;
    if (want) {
      put_ray(q.rays, slot, o, d);
      q.path[slot] = path;
    }
  }
}

// ---- material sort of a bounce's hit queue ----------------------------------
// Counting sort by the hit's material index / group (misses in the last
// bin): wf_sort_count counts the keys, wf_sort_scatter gives every block a
// range per key and writes the queue entries' indices in key order.  Entries
// of one key keep no particular order (blocks reserve their ranges
// atomically); the shade results of a path do not depend on the order paths
// are shaded in.  group 1 sorts by material; larger groups give fewer bins
// (less sorting work, more divergence left in a wave).
constexpr uint32_t kMaxSortBins = 256;
struct SortBuf {
  uint32_t* hist;    // kMaxSortBins counts
  uint32_t* cursor;  // kMaxSortBins reservation cursors
  uint32_t* order;   // queue indices in material order (P)
};
__device__ __forceinline__ uint32_t sort_key(const TptHit& h, const GpuShade* inst, uint32_t bins, uint32_t group) {
  return h.instance == TPT_INVALID_INDEX ? bins - 1u : inst[h.instance].material / group;
}
__global__ __launch_bounds__(256) void wf_sort_zero(SortBuf sb) {
  sb.hist[threadIdx.x] = 0u;
  sb.cursor[threadIdx.x] = 0u;
}
__global__ __launch_bounds__(256) void wf_sort_count(const TptHit* __restrict__ hits, const GpuShade* inst,
                                                     const Counters* c, uint32_t in_idx, uint32_t bins, uint32_t group,
                                                     SortBuf sb) {
  __shared__ uint32_t cnt[kMaxSortBins];
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t n = c->rays[in_idx].v;
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) atomicAdd(&cnt[sort_key(hits[i], inst, bins, group)], 1u);
  __syncthreads();
  if (threadIdx.x < bins && cnt[threadIdx.x]) atomicAdd(&sb.hist[threadIdx.x], cnt[threadIdx.x]);
}
__global__ __launch_bounds__(256) void wf_sort_scatter(const TptHit* __restrict__ hits, const GpuShade* inst,
                                                       const Counters* c, uint32_t in_idx, uint32_t bins,
                                                       uint32_t group, SortBuf sb) {
  __shared__ uint32_t cnt[kMaxSortBins], first[kMaxSortBins];
  const uint32_t n = c->rays[in_idx].v;
  if (blockIdx.x * 256u >= n) return;  // uniform per block
  cnt[threadIdx.x] = 0u;
  // exclusive prefix of the key counts (256 entries, one per thread)
  first[threadIdx.x] = threadIdx.x < bins ? sb.hist[threadIdx.x] : 0u;
  __syncthreads();
  for (uint32_t d = 1; d < kMaxSortBins; d <<= 1) {
    const uint32_t v = threadIdx.x >= d ? first[threadIdx.x - d] : 0u;
    __syncthreads();
    first[threadIdx.x] += v;
    __syncthreads();
  }
  const uint32_t excl = first[threadIdx.x] - (threadIdx.x < bins ? sb.hist[threadIdx.x] : 0u);
  __syncthreads();
  first[threadIdx.x] = excl;
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  uint32_t key = 0, rank = 0;
  if (i < n) {
    key = sort_key(hits[i], inst, bins, group);
    rank = atomicAdd(&cnt[key], 1u);
  }
  __syncthreads();
  if (threadIdx.x < bins && cnt[threadIdx.x]) first[threadIdx.x] += atomicAdd(&sb.cursor[threadIdx.x], cnt[threadIdx.x]);
  __syncthreads();
  if (i < n) sb.order[first[key] + rank] = i;
}

// One bounce of mis::trace_radiance for every path of the input queue.
// With kTex the ray differentials and cones follow the paths and the
// materials' textures are applied at every vertex.  order: the queue indices
// in material order (nullptr: queue order).
// kFilm: some material has a thin film (the film-free instantiation leaves
// the film's interference code out)
template <bool kTex, uint32_t kTile, bool kFilm>
__global__ __launch_bounds__(kTile) void wf_shade(Pass ps, pt::ShadeView sv, State st, Queue qin,
                                                const TptHit* __restrict__ hits, Queue qout, ShadowQ sq,
                                                Counters* c, uint32_t in_idx, uint32_t bounce,
                                                const uint8_t* __restrict__ ztab, uint32_t ztab_dims,
                                                const uint32_t* __restrict__ order) {
  // One queue slot per thread (the host launches ceil(paths / block) blocks
  // and the queue never holds more than `paths` entries): no thread carries
  // state from one slot to another.  Blocks past the queue length leave at
  // once (uniform per block, before any barrier).
  const uint32_t n = c->rays[in_idx].v;
  const uint32_t base = blockIdx.x * blockDim.x;
  if (base >= n) return;
  // digit-permutation rows of this bounce's sampler dimensions (vertex base +0, +1, +3, +5, +7, +13)
  __shared__ __align__(16) uint8_t zrows[6][smp::kZTabPrefixes];
  {
    const uint32_t bd = smp::kVertexBase + bounce * smp::kVertexStride;
    const uint32_t off[6] = {0, 1, 3, 5, 7, 13};
    for (int r = 0; r < 6; ++r) stage_row(zrows[r], ztab, ztab_dims, bd + off[r]);
    __syncthreads();
  }
  const uint32_t k = c->pass.v;
  {
    const uint32_t slot = base + threadIdx.x;
    bool want_ray = false, want_shadow = false;
    uint32_t path = 0;
    F3 no = f3(0.0f, 0.0f, 0.0f), nd = no;
    pt::NeeSample ns;
    F3 nee_value = no;
    if (slot < n) {
      const uint32_t e = order ? order[slot] : slot;
      path = qin.path[e];
      const TptHit h = hits[e];
      if (h.instance == TPT_INVALID_INDEX && sv.env.px) {  // a miss ends the path: environment emission
        const TptRay r = qin.rays[e];
        const F3 dir = f3(r.d[0], r.d[1], r.d[2]);
        F3 rad = load3f(st.rad, path);
        if (bounce == 0) {
          rad = pt::env_le(sv, dir);
        } else {
          const uint32_t info = st.prev_info[path];
          rad = pt::add(rad, pt::emitted_infinite(sv, load3f(st.thr, path), st.prev_pdf[path], (info >> 31) != 0, dir));
        }
        store3(st.rad, path, rad);
      }
      if (h.instance != TPT_INVALID_INDEX) {
        const uint32_t gp = path / ps.spp_slots, s = path % ps.spp_slots;
        const FrameDev& fr = ps.frames[frame_of(ps, gp)];
        const uint32_t lp = gp - fr.pix_begin;
        const smp::PathSampler smpl{fr.zf, lp % fr.w, lp / fr.w, k * ps.spp_slots + s};
        const TptRay r = qin.rays[e];
        const F3 dir = f3(r.d[0], r.d[1], r.d[2]);
        pt::RayDiff in;
        if (kTex) {
          in.o = f3(r.o[0], r.o[1], r.o[2]);
          in.d = dir;
          load_diff(st, ps.paths, path, &in);
        }
        pt::VtxTex tx;
        const pt::Vtx vtx = pt::shading_vertex_t<kTex>(sv, h.instance, h.triangle, h.b[0], h.b[1], h.b[2], dir, &in, &tx);
        const pt::VtxMat<kTex> vm(sv, vtx.material, &tx);
        const TptMaterial& m = vm.get();
        F3 rad = load3f(st.rad, path);
        F3 thr = load3f(st.thr, path);
        if (bounce == 0) {
          F3 le;
          if (pt::mat_le(m, vtx, &le)) rad = pt::add(rad, le);
        } else {
          const uint32_t info = st.prev_info[path];
          pt::TreeQuery pq;
          const float4 q4 = reinterpret_cast<const float4*>(st.prev_q)[path];
          pq.rho_d = q4.x;
          pq.rho_g = q4.y;
          pq.ax = q4.z;
          pq.ay = q4.w;
          pq.p = load3f(st.prev_p, path);
          pq.n = load3f(st.prev_ns, path);
          if (pq.rho_g > 0.0f) {
            pq.t = load3f(st.prev_t, path);
            pq.wo = load3f(st.prev_wo, path);
          } else {
            pq.t = pq.wo = f3(0.0f, 0.0f, 0.0f);
          }
          rad = pt::add(rad, pt::emitted(sv, thr, st.prev_pdf[path], (info >> 31) != 0, pq, vtx, m));
        }
        if (bounce < fr.max_depth) {
          const uint32_t bd = smp::kVertexBase + bounce * smp::kVertexStride;
          const pt::TreeQuery tq = pt::tree_query(sv.nee_policy, m, vtx);
          if (m.kind != TPT_MATERIAL_EMISSIVE) {
            float us[2];
            smpl.s2(bd + 3, us, zrows[2]);
            constexpr int kMode = kFilm ? 0 : pt::kShadeNoFilm;
            if (pt::nee_sample<kTex, kMode>(sv, m, vtx, tq, smpl.s1(bd + 0, zrows[0]), smpl.s1(bd + 1, zrows[1]), us[0], us[1],
                                     &ns)) {
              want_shadow = true;
              nee_value = pt::hadamard(thr, ns.value);
            }
          }
          pt::Sample smp_;
          F3 wi;
          float ud[2];
          smpl.s2(bd + 7, ud, zrows[4]);
          if (pt::mat_sample<kFilm ? 0 : pt::kShadeNoFilm>(sv.tab, m, vtx, smpl.s1(bd + 5, zrows[3]), ud[0], ud[1], &smp_, &wi)) {
            thr = pt::hadamard(thr, smp_.weight);
            bool live = true;
            if (bounce + 1 >= 4) {
              const float qs = pt::clampf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 0.05f, 1.0f);
              if (smpl.s1(bd + 13, zrows[5]) > qs)
                live = false;
              else
                thr = pt::divs(thr, qs);
            }
            if (live) {
              want_ray = true;
              no = pt::offset_origin(vtx.p, vtx.ng, wi);
              nd = wi;
              if (kTex) store_diff(st, ps.paths, path, pt::spawn_scattered(in, h.t, vtx, tx, smp_, no, wi));
              store3(st.thr, path, thr);
              store3(st.prev_p, path, vtx.p);
              store3(st.prev_ns, path, vtx.ns);
              reinterpret_cast<float4*>(st.prev_q)[path] = make_float4(tq.rho_d, tq.rho_g, tq.ax, tq.ay);
              if (tq.rho_g > 0.0f) {
                store3(st.prev_t, path, tq.t);
                store3(st.prev_wo, path, tq.wo);
              }
              st.prev_pdf[path] = smp_.pdf;
              st.prev_info[path] = vtx.material | (smp_.delta ? 0x80000000u : 0u);
            }
          }
        }
        store3(st.rad, path, rad);
      }
    }
    const AppendSlots slots = block_append<kTile>(&c->rays[in_idx ^ 1u].v, want_ray, &c->shadow.v, want_shadow);
    const uint32_t rs = slots.a, ss = slots.b;
    if (want_ray) {
      put_ray(qout.rays, rs, no, nd);
      qout.path[rs] = path;
    }
    if (want_shadow) {
      put_ray(sq.rays, ss, ns.origin, ns.dir);
      sq.target[2 * ss] = ns.inst;
      sq.target[2 * ss + 1] = ns.tri;
      store3(sq.value, ss, nee_value);
      sq.path[ss] = path;
    }
  }
}

// Unoccluded NEE contributions into the path radiance (one per path per bounce).
__global__ __launch_bounds__(256) void wf_shadow_add(State st, ShadowQ sq, const Counters* c) {
  const uint32_t n = c->shadow.v;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!sq.unocc[i]) continue;
    const uint32_t path = sq.path[i];
    store3(st.rad, path, pt::add(load3f(st.rad, path), load3f(sq.value, i)));
  }
}

// Running mean over the pass's samples, in sample order.
__global__ __launch_bounds__(256) void wf_fold(Pass ps, State st, float* __restrict__ color, const Counters* c) {
  const uint32_t k = c->pass.v;
  for (uint32_t gp = blockIdx.x * blockDim.x + threadIdx.x; gp < ps.total_pixels; gp += gridDim.x * blockDim.x) {
    const FrameDev& fr = ps.frames[frame_of(ps, gp)];
    F3 col = load3f(color, gp);
    for (uint32_t s = 0; s < ps.spp_slots; ++s) {
      const uint32_t si = k * ps.spp_slots + s;
      if (si >= fr.spp) break;
      const F3 v = load3f(st.rad, gp * ps.spp_slots + s);
      col = pt::add(col, pt::divs(pt::sub(v, col), (float)(si + 1)));
    }
    store3(color, gp, col);
  }
}

__global__ __launch_bounds__(256) void wf_output(Pass ps, const float* __restrict__ color) {
  for (uint32_t gp = blockIdx.x * blockDim.x + threadIdx.x; gp < ps.total_pixels; gp += gridDim.x * blockDim.x) {
    const FrameDev& fr = ps.frames[frame_of(ps, gp)];
    const uint32_t lp = gp - fr.pix_begin;
    fr.out[3 * lp] = color[3 * gp] * fr.cam.exposure;
    fr.out[3 * lp + 1] = color[3 * gp + 1] * fr.cam.exposure;
    fr.out[3 * lp + 2] = color[3 * gp + 2] * fr.cam.exposure;
  }
}

__global__ void wf_zero(float* a, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0.0f;
}
__global__ void wf_zero_u32(uint32_t* a) {
  if (threadIdx.x == 0) a[0] = 0u;
}
__global__ void wf_zero_counters(Counters* c) {
  if (threadIdx.x == 0) {
    c->rays[0].v = 0u;
    c->rays[1].v = 0u;
    c->shadow.v = 0u;
    c->pass.v = 0u;
  }
}

}  // namespace wf
}  // namespace tpt

using namespace tpt;

namespace {

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

template <typename K>
unsigned resident_blocks(TptDevice* d, K kern, size_t work, unsigned block = 256) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), (int)block, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const size_t cap = (size_t)per_cu * (size_t)(d->cu_count > 0 ? d->cu_count : 1);
  const size_t want = (work + block - 1) / block;
  return (unsigned)std::max<size_t>(1, std::min(cap, want));
}

struct Plan {
  wf::Pass ps;
  wf::State st;
  wf::Queue q[2];
  wf::ShadowQ sq;
  TptHit* hits;
  uint8_t* ztab;      // digit-permutation rows
  uint32_t ztab_dims;
  wf::Counters* c;
  float* color;
  uint32_t max_depth;
  bool q8;
  bool exact_shadow;
  bool tex;  // the scene has textures: ray differentials and textured materials
  bool film;  // some material has a thin film
  wf::SortBuf sort;
  uint32_t sort_bins;   // ceil(material count / group) + 1 (misses); 0: shade in queue order
  uint32_t sort_group;  // materials per sort key
  uint32_t tile;        // threads per generate / shade block (one queue entry each): 128, 256 or 512
};

// The kernels of one pass, all reading their counts on the device; kTile:
// the generate / shade blocks' threads.
template <uint32_t kTile>
bool enqueue_pass_tile(TptScene* s, const Plan& p, const pt::ShadeView& sv, hipStream_t st) {
  TptDevice* d = s->dev;
  auto gen = p.tex ? wf::wf_generate<true, kTile> : wf::wf_generate<false, kTile>;
  auto shade = p.film ? (p.tex ? wf::wf_shade<true, kTile, true> : wf::wf_shade<false, kTile, true>)
                      : (p.tex ? wf::wf_shade<true, kTile, false> : wf::wf_shade<false, kTile, false>);
  // One thread per path slot: generate and shade take exactly one slot per
  // thread (they have no slot loop).  An earlier form looped over slots with
  // the shade grid capped at the resident capacity; there the untextured
  // build gave wrong OpenPBR paths on a thread's second slot, differently
  // from run to run (the cause was not found), so no launch may hand a
  // thread more than one slot.  Blocks past the queue length exit at once.
  const unsigned g_gen = (unsigned)((p.ps.paths + kTile - 1) / kTile);
  const unsigned g_shade = g_gen;
  if ((uint64_t)g_gen * kTile < p.ps.paths) {
    set_error("tpt_render_batch: generate / shade grid smaller than the paths in flight");
    return false;
  }
  const unsigned g_add = resident_blocks(d, wf::wf_shadow_add, p.ps.paths);
  const unsigned g_fold = resident_blocks(d, wf::wf_fold, p.ps.total_pixels);
  hipLaunchKernelGGL(wf::wf_reset, dim3(1), dim3(64), 0, st, p.c, 0u);
  hipLaunchKernelGGL(gen, dim3(g_gen), dim3(kTile), 0, st, p.ps, p.st, p.q[0], p.c,
                     (const uint8_t*)p.ztab, p.ztab_dims);
  for (uint32_t b = 0; b <= p.max_depth; ++b) {
    const uint32_t a = b & 1u;
    if (!launch_trace_queue(s, 0, p.q8, false, p.q[a].rays, &p.c->rays[a].v, p.ps.paths, p.hits, nullptr, nullptr, st))
      return false;
    hipLaunchKernelGGL(wf::wf_reset, dim3(1), dim3(64), 0, st, p.c, a ^ 1u);
    const uint32_t* order = nullptr;
    if (p.sort_bins) {
      const unsigned g_sort = (unsigned)((p.ps.paths + 255u) / 256u);
      hipLaunchKernelGGL(wf::wf_sort_zero, dim3(1), dim3(256), 0, st, p.sort);
      hipLaunchKernelGGL(wf::wf_sort_count, dim3(g_sort), dim3(256), 0, st, (const TptHit*)p.hits, sv.inst,
                         (const wf::Counters*)p.c, a, p.sort_bins, p.sort_group, p.sort);
      hipLaunchKernelGGL(wf::wf_sort_scatter, dim3(g_sort), dim3(256), 0, st, (const TptHit*)p.hits, sv.inst,
                         (const wf::Counters*)p.c, a, p.sort_bins, p.sort_group, p.sort);
      order = p.sort.order;
    }
    hipLaunchKernelGGL(shade, dim3(g_shade), dim3(kTile), 0, st, p.ps, sv, p.st, p.q[a], (const TptHit*)p.hits,
                       p.q[a ^ 1u], p.sq, p.c, a, b, (const uint8_t*)p.ztab, p.ztab_dims, order);
    if (b < p.max_depth && s->host_cats.count) {  // the light sampler has a category: NEE can run
      if (!launch_trace_queue(s, 1, p.q8, p.exact_shadow, p.sq.rays, &p.c->shadow.v, p.ps.paths, nullptr, p.sq.target,
                              p.sq.unocc, st))
        return false;
      hipLaunchKernelGGL(wf::wf_shadow_add, dim3(g_add), dim3(256), 0, st, p.st, p.sq, (const wf::Counters*)p.c);
    }
  }
  hipLaunchKernelGGL(wf::wf_fold, dim3(g_fold), dim3(256), 0, st, p.ps, p.st, p.color, (const wf::Counters*)p.c);
  hipLaunchKernelGGL(wf::wf_next_pass, dim3(1), dim3(64), 0, st, p.c);
  return hip_ok(hipGetLastError(), "wavefront launch");
}
bool enqueue_pass(TptScene* s, const Plan& p, const pt::ShadeView& sv, hipStream_t st) {
  if (p.tile == 128) return enqueue_pass_tile<128>(s, p, sv, st);
  if (p.tile == 512) return enqueue_pass_tile<512>(s, p, sv, st);
  return enqueue_pass_tile<256>(s, p, sv, st);
}

}  // namespace

extern "C" int32_t tpt_render_batch(TptScene* s, const TptFrameDesc* frames, uint32_t frame_count, uint32_t flags) {
  try {
    if (!s || (frame_count && !frames)) {
      set_error("tpt_render_batch: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    uint64_t total = 0;
    uint32_t max_spp = 0, max_depth = 0;
    for (uint32_t f = 0; f < frame_count; ++f) {
      const TptRenderDesc& r = frames[f].render;
      if (!frames[f].rgb || r.samples_per_pixel == 0) {
        set_error("tpt_render_batch: frame " + std::to_string(f) + ": null rgb or samples_per_pixel == 0");
        return TPT_ERROR_BAD_ARGUMENT;
      }
      total += (uint64_t)r.width * r.height;
      max_spp = std::max(max_spp, r.samples_per_pixel);
      max_depth = std::max(max_depth, r.max_depth);
    }
    if (total == 0) return TPT_OK;
    if (total > 0x3FFFFFFFull) {
      set_error("tpt_render_batch: too many pixels for one call");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    bool film = true;
    if (!render_materials_ok(s, "tpt_render_batch", &film)) return TPT_ERROR_BAD_ARGUMENT;
    TptDevice* d = s->dev;
    if (!hip_ok(hipSetDevice(d->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    const hipStream_t st = d->stream;
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    // S samples per pixel per pass, P = pixels * S paths in flight
    const uint64_t cap = std::max<uint64_t>(d->wf_paths, total);
    const uint32_t S = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(max_spp, cap / total));
    const uint64_t P = total * S;
    const uint32_t passes = (max_spp + S - 1) / S;
    // one allocation: counters | frames | color | staged outputs | path state | queues | hits | shadow queue
    size_t off = 0;
    auto take = [&](size_t bytes) {
      const size_t o = off;
      off += align256(bytes);
      return o;
    };
    const size_t o_cnt = take(sizeof(wf::Counters));
    const size_t o_frames = take(frame_count * sizeof(wf::FrameDev));
    const size_t o_color = take(12 * total);
    const size_t o_out = dev_ptr ? 0 : take(12 * total);
    const size_t o_thr = take(12 * P), o_rad = take(12 * P), o_pp = take(12 * P), o_pn = take(12 * P);
    const size_t o_pdf = take(4 * P), o_info = take(4 * P), o_pq = take(16 * P);
    const bool proxy = s->nee_policy == TPT_NEE_OPENPBR_PROXY;
    const size_t o_pt = proxy ? take(12 * P) : 0, o_pwo = proxy ? take(12 * P) : 0;
    const size_t o_q0r = take(sizeof(TptRay) * P), o_q0p = take(4 * P);
    const size_t o_q1r = take(sizeof(TptRay) * P), o_q1p = take(4 * P);
    const size_t o_hits = take(sizeof(TptHit) * P);
    const size_t o_sr = take(sizeof(TptRay) * P), o_st = take(8 * P), o_sv = take(12 * P), o_sp = take(4 * P);
    const size_t o_su = take(P);
    const bool tex = s->texture_count > 0;
    const size_t o_diff = tex ? take(12 * 4 * P) : 0, o_dflag = tex ? take(4 * P) : 0, o_cone = tex ? take(8 * P) : 0;
    const uint32_t sort_group = d->wf_sort;
    const uint32_t sort_keys = sort_group ? (s->material_count + sort_group - 1) / sort_group + 1 : 0;
    const uint32_t sort_bins = sort_keys <= wf::kMaxSortBins ? sort_keys : 0;
    const size_t o_sort = sort_bins ? take(4 * (2 * wf::kMaxSortBins + P)) : 0;
    // digit-permutation rows of every sampler dimension a path of up to 256 bounces reads
    // (deeper bounces compute their rows in the kernel)
    const uint32_t ztab_dims = smp::kVertexBase + smp::kVertexStride * (std::min<uint32_t>(max_depth, 256u) + 1);
    const size_t o_ztab = take((size_t)ztab_dims * smp::kZTabPrefixes);
    if (!d->wf_buf.ensure(off)) return TPT_ERROR_DEVICE;
    char* base = (char*)d->wf_buf.ptr;
    // frame table
    std::vector<wf::FrameDev> fd(frame_count);
    uint32_t pix = 0;
    size_t out_off = 0;
    for (uint32_t f = 0; f < frame_count; ++f) {
      const TptRenderDesc& r = frames[f].render;
      wf::FrameDev& x = fd[f];
      const TptCamera& c = r.camera;
      x.cam.eye = F3{c.eye[0], c.eye[1], c.eye[2]};
      x.cam.fwd = F3{c.forward[0], c.forward[1], c.forward[2]};
      x.cam.right = F3{c.right[0], c.right[1], c.right[2]};
      x.cam.up = F3{c.image_up[0], c.image_up[1], c.image_up[2]};
      x.cam.tan_half = c.tan_half_fov_y;
      x.cam.exposure = c.exposure;
      x.cam.spread = c.pixel_spread;
      x.w = r.width;
      x.h = r.height;
      x.spp = r.samples_per_pixel;
      x.max_depth = r.max_depth;
      x.zf = smp::zsobol_frame(r.samples_per_pixel, r.width, r.height);
      x.pix_begin = pix;
      x.pad = 0;
      x.out = dev_ptr ? frames[f].rgb : (float*)(base + o_out + out_off);
      pix += r.width * r.height;
      out_off += 12 * (size_t)r.width * r.height;
    }
    Plan p;
    p.ps.frames = (const wf::FrameDev*)(base + o_frames);
    p.ps.frame_count = frame_count;
    p.ps.spp_slots = S;
    p.ps.total_pixels = (uint32_t)total;
    p.ps.paths = (uint32_t)P;
    p.st = wf::State{(float*)(base + o_thr), (float*)(base + o_rad), (float*)(base + o_pp), (float*)(base + o_pn),
                     (float*)(base + o_pdf), (uint32_t*)(base + o_info), (float*)(base + o_pq),
                     proxy ? (float*)(base + o_pt) : nullptr, proxy ? (float*)(base + o_pwo) : nullptr,
                     tex ? (float*)(base + o_diff) : nullptr, tex ? (uint32_t*)(base + o_dflag) : nullptr,
                     tex ? (float*)(base + o_cone) : nullptr};
    p.q[0] = wf::Queue{(TptRay*)(base + o_q0r), (uint32_t*)(base + o_q0p)};
    p.q[1] = wf::Queue{(TptRay*)(base + o_q1r), (uint32_t*)(base + o_q1p)};
    p.sq = wf::ShadowQ{(TptRay*)(base + o_sr), (uint32_t*)(base + o_st), (float*)(base + o_sv),
                       (uint32_t*)(base + o_sp), (uint8_t*)(base + o_su)};
    p.hits = (TptHit*)(base + o_hits);
    p.ztab = (uint8_t*)(base + o_ztab);
    p.ztab_dims = ztab_dims;
    p.c = (wf::Counters*)(base + o_cnt);
    p.color = (float*)(base + o_color);
    p.max_depth = max_depth;
    p.q8 = flags & TPT_FLAG_QUANTIZED;
    p.exact_shadow = flags & TPT_FLAG_ANY_EXACT;
    p.tex = tex;
    p.film = film;
    p.sort_bins = sort_bins;
    p.sort_group = sort_group ? sort_group : 1u;
    p.tile = d->wf_tile;
    p.sort = sort_bins ? wf::SortBuf{(uint32_t*)(base + o_sort), (uint32_t*)(base + o_sort) + wf::kMaxSortBins,
                                     (uint32_t*)(base + o_sort) + 2 * wf::kMaxSortBins}
                       : wf::SortBuf{nullptr, nullptr, nullptr};
    const pt::ShadeView sv = pt::shade_view_of(s);
    if (!hip_ok(hipMemcpyAsync((void*)p.ps.frames, fd.data(), fd.size() * sizeof(wf::FrameDev), hipMemcpyHostToDevice,
                               st),
                "hipMemcpyAsync(frames)"))
      return TPT_ERROR_DEVICE;
    hipLaunchKernelGGL(wf::wf_zero_counters, dim3(1), dim3(64), 0, st, p.c);
    hipLaunchKernelGGL(wf::wf_ztab, dim3(resident_blocks(d, wf::wf_ztab, (size_t)ztab_dims * smp::kZTabPrefixes)),
                       dim3(256), 0, st, p.ztab, ztab_dims);
    hipLaunchKernelGGL(wf::wf_zero, dim3(resident_blocks(d, wf::wf_zero, 3 * total)), dim3(256), 0, st, p.color,
                       (size_t)(3 * total));
    if (!dev_ptr) hipLaunchKernelGGL(wf::wf_zero_u32, dim3(1), dim3(64), 0, st, d->status);
    bool graphed = false;
    if (d->wf_graph && passes > 1) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const bool ok = enqueue_pass(s, p, sv, st);
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (ok && e == hipSuccess && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
          graphed = true;
          for (uint32_t k = 0; k < passes && graphed; ++k) graphed = hipGraphLaunch(exec, st) == hipSuccess;
          if (!graphed) {
            set_error("tpt_render_batch: hipGraphLaunch failed");
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            return TPT_ERROR_DEVICE;
          }
        }
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
      }
      (void)hipGetLastError();  // a refused capture leaves an error behind; the plain launches follow
    }
    if (!graphed)
      for (uint32_t k = 0; k < passes; ++k)
        if (!enqueue_pass(s, p, sv, st)) return TPT_ERROR_DEVICE;
    hipLaunchKernelGGL(wf::wf_output, dim3(resident_blocks(d, wf::wf_output, total)), dim3(256), 0, st, p.ps,
                       (const float*)p.color);
    if (!hip_ok(hipGetLastError(), "wf_output launch")) return TPT_ERROR_DEVICE;
    if (!dev_ptr) {
      uint32_t status = 0;
      for (uint32_t f = 0; f < frame_count; ++f)
        if (!hip_ok(hipMemcpyAsync(frames[f].rgb, fd[f].out, 12 * (size_t)fd[f].w * fd[f].h, hipMemcpyDeviceToHost, st),
                    "hipMemcpyAsync(rgb)"))
          return TPT_ERROR_DEVICE;
      if (!hip_ok(hipMemcpyAsync(&status, d->status, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(status)") ||
          !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return TPT_ERROR_DEVICE;
      if (status & 1u) {
        set_error("tpt_render_batch: traversal stack overflow (128 entries)");
        return TPT_ERROR_OVERFLOW;
      }
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_render_batch: ") + e.what());
    return TPT_ERROR;
  }
}

namespace {

// Time one batched render (device outputs) on the device stream, ms.
bool time_batch(TptScene* s, std::vector<TptFrameDesc>& frames, uint32_t flags, int reps, float* ms) {
  TptDevice* d = s->dev;
  hipEvent_t a = nullptr, b = nullptr;
  if (!hip_ok(hipEventCreate(&a), "hipEventCreate") || !hip_ok(hipEventCreate(&b), "hipEventCreate")) return false;
  bool ok = tpt_render_batch(s, frames.data(), (uint32_t)frames.size(), flags | TPT_FLAG_DEVICE_POINTERS) == TPT_OK;
  ok = ok && hip_ok(hipEventRecord(a, d->stream), "hipEventRecord");
  for (int r = 0; ok && r < reps; ++r)
    ok = tpt_render_batch(s, frames.data(), (uint32_t)frames.size(), flags | TPT_FLAG_DEVICE_POINTERS) == TPT_OK;
  ok = ok && hip_ok(hipEventRecord(b, d->stream), "hipEventRecord") &&
       hip_ok(hipEventSynchronize(b), "hipEventSynchronize") && hip_ok(hipEventElapsedTime(ms, a, b), "hipEventElapsedTime");
  if (ok) *ms /= (float)reps;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ok;
}

}  // namespace

// A candidate replaces the current setting only when it is at least 1% faster:
// run-to-run noise of a few tenths of a percent must not walk the setting away
// from the starting one.
constexpr float kTuneMargin = 0.99f;
// Autotuner of the wavefront path tracer: coordinate descent over paths in
// flight, resident trace blocks per CU, the trace kernels' refill threshold,
// their traversal stack's LDS depth, the shade queue's tile (threads per
// generate / shade block) and the material sort's granularity (off, or
// materials per sort key), each candidate timed (HIP events) on the
// caller's frames at tune_spp samples per pixel; the fastest setting is left
// on the device.  A paths-in-flight candidate larger than the tuning batch's
// path count runs the batch in one pass, like the batch size itself, so it
// is not tried; sort granularities whose key count exceeds the sort's 256
// bins are not tried.
extern "C" int32_t tpt_autotune_render(TptScene* s, const TptFrameDesc* frames, uint32_t frame_count, uint32_t flags,
                                       uint32_t tune_spp, TptTuneResult* out) {
  try {
    if (!s || !frames || frame_count == 0 || !out) {
      set_error("tpt_autotune_render: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    TptDevice* d = s->dev;
    if (!hip_ok(hipSetDevice(d->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    size_t pixels = 0;
    for (uint32_t f = 0; f < frame_count; ++f) pixels += (size_t)frames[f].render.width * frames[f].render.height;
    DeviceBuffer img;
    if (!img.ensure(std::max<size_t>(12 * pixels, 256))) return TPT_ERROR_DEVICE;
    std::vector<TptFrameDesc> fr(frames, frames + frame_count);
    size_t off = 0;
    for (auto& f : fr) {
      f.rgb = (float*)img.ptr + 3 * off;
      off += (size_t)f.render.width * f.render.height;
      if (tune_spp) f.render.samples_per_pixel = std::min(f.render.samples_per_pixel, tune_spp);
    }
    uint32_t trials = 0;
    float best = 0.0f;
    if (!time_batch(s, fr, flags, 2, &best)) return TPT_ERROR_DEVICE;
    ++trials;
    const float base_ms = best;
    auto tryv = [&](uint32_t* knob, uint32_t v) -> int32_t {
      const uint32_t old = *knob;
      *knob = v;
      float ms = 0.0f;
      if (!time_batch(s, fr, flags, 2, &ms)) {
        *knob = old;
        return TPT_ERROR_DEVICE;
      }
      ++trials;
      if (ms < best * kTuneMargin) {
        best = ms;
      } else {
        *knob = old;
      }
      return TPT_OK;
    };
    const uint32_t paths[] = {1u << 19, 1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24};
    const uint32_t bpcs[] = {0u, 2u, 3u, 4u, 6u, 8u};
    const uint32_t refills[] = {8u, 16u, 24u, 32u, 48u};
    const uint32_t depths[] = {8u, 12u, 16u, 24u, 32u};
    const uint32_t tiles[] = {128u, 256u, 512u};
    const uint32_t groups[] = {0u, 1u, 2u, 4u, 8u};
    size_t batch_paths = 0;
    for (const auto& f : fr) batch_paths += (size_t)f.render.width * f.render.height * f.render.samples_per_pixel;
    for (uint32_t v : paths)
      if (v != d->wf_paths && v <= batch_paths && tryv(&d->wf_paths, v) != TPT_OK) return TPT_ERROR_DEVICE;
    for (uint32_t v : bpcs)
      if (v != d->blocks_per_cu && tryv(&d->blocks_per_cu, v) != TPT_OK) return TPT_ERROR_DEVICE;
    for (uint32_t v : refills)
      if (v != d->refill_below && tryv(&d->refill_below, v) != TPT_OK) return TPT_ERROR_DEVICE;
    for (uint32_t v : depths)
      if (v != d->lds_depth && tryv(&d->lds_depth, v) != TPT_OK) return TPT_ERROR_DEVICE;
    for (uint32_t v : tiles)
      if (v != d->wf_tile && tryv(&d->wf_tile, v) != TPT_OK) return TPT_ERROR_DEVICE;
    for (uint32_t v : groups) {
      const uint32_t keys = v ? (s->material_count + v - 1) / v + 1 : 0;
      if (v != d->wf_sort && keys <= wf::kMaxSortBins && tryv(&d->wf_sort, v) != TPT_OK) return TPT_ERROR_DEVICE;
    }
    out->wf_paths = d->wf_paths;
    out->blocks_per_cu = d->blocks_per_cu;
    out->refill_below = d->refill_below;
    out->trials = trials;
    out->wf_sort = d->wf_sort;
    out->ms_default = base_ms;
    out->ms_best = best;
    out->lds_depth = d->lds_depth;
    out->wf_tile = d->wf_tile;
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_autotune_render: ") + e.what());
    return TPT_ERROR;
  }
}
