// Plain FP32 megakernel path tracer: the render loop of R/main.rs:98-126 with
// the unidirectional MIS integrator of R/integrator/mis.rs:19-139, over the
// materials tpt_render accepts (include/tpt.h).  One thread per pixel runs its
// samples in order, so the reference's running mean `color += (s - color) / n`
// is reproduced exactly.
//
// Per-sample work: PathSampler (bit-exact ZSobol), pinhole ray, two-level
// closest hit (the trace kernels' traversal), shading vertex
// (R/scene/mod.rs:471-544), next-event estimation through the SG light tree
// with a shadow ray and balance-heuristic MIS (R/integrator/mis.rs:141-210),
// material sample (OpenPBR R/material/open_pbr.rs:489-548; NormalizedLambert;
// Mirror), Russian roulette, emission found by BSDF sampling with its MIS
// weight (R/integrator/mis.rs:212-234).
//
// OpenPBR subset: specular weight 0 makes the modulated specular eta exactly
// 1, so f_avg(eta) = 0 (R/bsdf/open_pbr.rs:2127-2135) and every term that
// reads the reference's directional-albedo and energy-compensation tables is
// multiplied by zero; metal, coat, fuzz, transmission, thin film and thin
// walls are zero.  Light NEE under the reference policy gives no sample for
// OpenPBR shading points (no light-tree query, no environment), so emission
// found by BSDF sampling from them has MIS weight 1; NormalizedLambert
// surfaces build a diffuse light-tree query (R/material/normalized_lambert.rs:
// 201-218), so for them NEE runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

#include "sampler.h"
#include "tpt_device.h"
#include "tpt_internal.h"
#include "traverse.h"

#include "shading.h"

namespace tpt {
namespace pt {

// closest hit of one ray with the trace kernels' lane traversal
template <bool kTwoLevel>
__device__ __forceinline__ bool closest(const SceneView& sc, F3 o, F3 d, Lane& L, LaneStack& st, uint32_t* inst,
                                        uint32_t* tri, float* t, float* b0, float* b1, float* b2) {
  lane_begin<kTwoLevel>(sc, L, st, o, d, __builtin_inff(), __builtin_inff(), TPT_INVALID_INDEX, TPT_INVALID_INDEX);
  Counters cnt = {0, 0, 0};
  while (!lane_step<false, kClosest, false, kTwoLevel>(sc, L, st, cnt)) {
  }
  *inst = L.inst;
  *tri = L.tri;
  *t = L.t;
  *b0 = L.b0;
  *b1 = L.b1;
  *b2 = L.b2;
  return L.hit;
}

// direct_light_mis_contribution with the shadow ray traced by the lane
// traversal (closest hit, then the target test).
template <bool kTwoLevel, bool kTex, int kMode>
__device__ __forceinline__ F3 direct_light(const SceneView& sc, const ShadeView& sv, const TptMaterial& m,
                                           const Vtx& vtx, const TreeQuery& tq, float u_root, float u_tree,
                                           float us0, float us1, Lane& L, LaneStack& st, F3* dir = nullptr) {
  NeeSample ns;
  if (!nee_sample<kTex, kMode>(sv, m, vtx, tq, u_root, u_tree, us0, us1, &ns)) return f3(0.0f, 0.0f, 0.0f);
  if (dir) *dir = ns.dir;
  uint32_t hi, ht;
  float tt, hb0, hb1, hb2;
  if (closest<kTwoLevel>(sc, ns.origin, ns.dir, L, st, &hi, &ht, &tt, &hb0, &hb1, &hb2) &&
      !(hi == ns.inst && ht == ns.tri))
    return f3(0.0f, 0.0f, 0.0f);
  return ns.value;
}

// Path-replay adjoint state of one sample: the path's radiance from the
// primal pass, dLoss/dL of this sample (dLoss/dI_pixel * exposure / spp), the
// block's gradient accumulators (3 per material, LDS) and the texel gradient
// array (laid out like the texels; null when not requested).
template <typename Acc, typename TexAcc = float>
struct Adjoint {
  F3 total;
  F3 w;
  Acc* grad;        // rows of `rows` gradients per material
  TexAcc* texgrad;  // laid out like the texels (every level)
  uint32_t rows;    // 3: albedo / base colour; 5 (kParamRows): + specular_roughness, base_metalness
};

// mis::trace_radiance (R/integrator/mis.rs:19-139) for one camera ray.  With
// kTex the ray's differential and cone follow the path
// (spawn_scattered_ray) and the materials' textures are applied at every
// vertex.  With
// kAdj, the replay pass of path-replay backpropagation: at every
// NormalizedLambert vertex, every contribution collected from its NEE
// onward carries the vertex's albedo exactly once (NEE's f = rho / pi, the
// BSDF sample's weight = rho; the light-tree pmf, MIS weights and pdfs do not
// depend on rho), so dL/drho_c += (L_total - L_before)_c / rho_c.  Russian
// roulette probabilities are held fixed (detached).  At an OpenPBR vertex
// whose BSDF is affine in the base colour c (base_color_linear), channel k of
// a contribution that went through the BSDF value f(w) changes by
// (df_k/dc_k)(w) / f_k(w) per unit of c_k: the vertex's NEE contribution at
// the light direction and everything collected after it (the BSDF-sampled
// continuation) at the sampled direction, sampling and MIS weights held
// fixed; c = parameter * texture value, so dL/dparameter = dL/dc * texture
// value and dL/dtexel = dL/dc * parameter * d(lookup)/d(texel).
template <bool kTwoLevel, bool kAdj, bool kTex, int kMode, typename Acc, typename TexAcc = float>
__device__ __forceinline__ F3 trace_path(const SceneView& sc, const ShadeView& sv, const RayDiff& camera_ray,
                                         const smp::PathSampler& ps, uint32_t max_depth, Lane& L, LaneStack& st,
                                         bool& overflow, const Adjoint<Acc, TexAcc>& adj) {
  constexpr bool kFilm = (kMode & kShadeNoFilm) == 0;
  F3 radiance = f3(0.0f, 0.0f, 0.0f), thr = f3(1.0f, 1.0f, 1.0f);
  uint32_t hi, ht;
  float tt, hb0, hb1, hb2;
  if (!closest<kTwoLevel>(sc, camera_ray.o, camera_ray.d, L, st, &hi, &ht, &tt, &hb0, &hb1, &hb2)) {
    overflow |= st.overflow;
    return env_le(sv, camera_ray.d);  // infinite_light_le of the camera ray
  }
  overflow |= st.overflow;
  RayDiff ray = camera_ray;
  float hit_t = tt;
  VtxTex tx;
  Vtx vtx = shading_vertex_t<kTex>(sv, hi, ht, hb0, hb1, hb2, ray.d, &ray, &tx);
  VtxMat<kTex> vm(sv, vtx.material, &tx);
  F3 le;
  if (mat_le(vm.get(), vtx, &le)) radiance = add(radiance, le);
  for (uint32_t depth = 0; depth < max_depth; ++depth) {
    const uint32_t base = smp::kVertexBase + depth * smp::kVertexStride;
    const TptMaterial& m = vm.get();
    if (kAdj && m.kind == TPT_MATERIAL_NORMALIZED_LAMBERT) {
      const F3 rest = sub(adj.total, radiance);
      Acc* g = adj.grad + adj.rows * vtx.material;
      const float* rho = m.u.lambert.rho;
      if (rho[0] != 0.0f) atomicAdd(&g[0], (Acc)(adj.w.x * rest.x / rho[0]));
      if (rho[1] != 0.0f) atomicAdd(&g[1], (Acc)(adj.w.y * rest.y / rho[1]));
      if (rho[2] != 0.0f) atomicAdd(&g[2], (Acc)(adj.w.z * rest.z / rho[2]));
    }
    F3 nee_c = f3(0.0f, 0.0f, 0.0f), nee_dir = nee_c;
    const TreeQuery tq = tree_query(sv.nee_policy, m, vtx);
    if (m.kind != TPT_MATERIAL_EMISSIVE) {  // should_sample_direct_light: not a pure emitter
      float us[2];
      ps.s2(base + 3, us);
      const F3 dl = direct_light<kTwoLevel, kTex, kMode>(sc, sv, m, vtx, tq, ps.s1(base + 0), ps.s1(base + 1), us[0], us[1],
                                                  L, st, kAdj ? &nee_dir : nullptr);
      overflow |= st.overflow;
      nee_c = hadamard(thr, dl);
      radiance = add(radiance, nee_c);
    }
    Sample s;
    F3 wi;
    float ud[2];
    ps.s2(base + 7, ud);
    const bool sampled = mat_sample<kMode>(sv.tab, m, vtx, ps.s1(base + 5), ud[0], ud[1], &s, &wi);
    const bool lin = kAdj && base_color_linear(m);
    if (kAdj && (lin || (adj.rows == kParamRows && m.kind == TPT_MATERIAL_OPEN_PBR))) {
      const F3 rest = sub(adj.total, radiance);  // collected after this vertex's NEE
      // a smooth lobe's mirror sample (delta) carries the lobe's weight
      // numerator, a function of every parameter but an affine base colour
      const bool use_nee = len2(nee_c) > 0.0f, use_smp = sampled && !s.delta;
      const bool use_dlt = sampled && s.delta && m.kind == TPT_MATERIAL_OPEN_PBR && adj.rows == kParamRows;
      const F3 wn = normalize_or_zero(to_frame(vtx.frame, nee_dir));
      F3 g = f3(0.0f, 0.0f, 0.0f);
      const TptOpenPbr& p0 = sv.mats[vtx.material].u.open_pbr;  // the parameters, untextured
      Acc* gm = adj.grad + adj.rows * vtx.material;
      if (adj.rows == kParamRows) {
        // every parameter by central differences of the BSDF: dL/dtheta_k =
        // sum over channels of w * (NEE * r_k(light dir) + rest * r_k(sampled dir))
        F3 rn[kParamRows], rs[kParamRows];
        const F3 zero = f3(0.0f, 0.0f, 0.0f);
        for (int k = 0; k < kParamRows; ++k) rn[k] = rs[k] = zero;
        const int first = lin ? 3 : 0;  // an affine base colour takes the exact ratio below
        // one rebuild of the BSDF per parameter and side serves both the light
        // direction and the sampled direction
        const bool thin_lobe = use_smp && s.trans && p0.thin_walled;  // the thin-wall BTDF's own estimate
        if (use_nee || use_smp)
          bsdf_param_ratios2<kFilm>(sv.tab, m, vtx, use_nee, wn, rn, use_smp, s.wi, rs, first, thin_lobe);
        if (use_dlt) bsdf_delta_param_ratios<kFilm>(sv.tab, m, vtx, s.lobe, rs, first);
        float gk[kParamRows];
        for (int k = 0; k < kParamRows; ++k)
          gk[k] = dot(adj.w, add(hadamard(nee_c, rn[k]), hadamard(rest, rs[k])));
        if (lin) {  // the base colour exactly (the ratio is diagonal and exact for affine BSDFs)
          if (use_nee) g = hadamard(nee_c, base_color_ratio<kFilm>(sv.tab, m, vtx, wn));
          if (use_smp) g = add(g, hadamard(rest, base_color_ratio<kFilm>(sv.tab, m, vtx, s.wi)));
          g = hadamard(g, adj.w);
        } else {
          g = f3(gk[0], gk[1], gk[2]);
        }
        // roughness and metalness: the effective value is the parameter
        // times the texture's; textured slots get no row
        if (p0.specular_roughness_texture == TPT_INVALID_INDEX && gk[3] != 0.0f) atomicAdd(&gm[3], (Acc)gk[3]);
        if (p0.base_metalness_texture == TPT_INVALID_INDEX && gk[4] != 0.0f) atomicAdd(&gm[4], (Acc)gk[4]);
      } else {
        if (use_nee) g = hadamard(nee_c, base_color_ratio<kFilm>(sv.tab, m, vtx, wn));
        if (use_smp) g = add(g, hadamard(rest, base_color_ratio<kFilm>(sv.tab, m, vtx, s.wi)));
        g = hadamard(g, adj.w);  // dL/dc of this sample
      }
      F3 tv = f3(1.0f, 1.0f, 1.0f);
      if (kTex && p0.base_color_texture != TPT_INVALID_INDEX) {
        const tex::V2 uv = tex::v2(tx.u, tx.v), dx = tex::v2(tx.dudx, tx.dvdx), dy = tex::v2(tx.dudy, tx.dvdy);
        const pt::GpuTexture& t = sv.tex.tex[p0.base_color_texture];
        tv = tex::filtered<3>(sv.tex, t, uv, dx, dy);
        if (adj.texgrad)
          tex::scatter_filtered<3, TexAcc>(sv.tex, t, uv, dx, dy, hadamard(g, rgb3(p0.base_color)), adj.texgrad);
      }
      const F3 gp = hadamard(g, tv);
      if (gp.x != 0.0f) atomicAdd(&gm[0], (Acc)gp.x);
      if (gp.y != 0.0f) atomicAdd(&gm[1], (Acc)gp.y);
      if (gp.z != 0.0f) atomicAdd(&gm[2], (Acc)gp.z);
    }
    if (!sampled) break;
    thr = hadamard(thr, s.weight);
    if (depth + 1 >= 4) {
      const float q = clampf(fmaxf(thr.x, fmaxf(thr.y, thr.z)), 0.05f, 1.0f);
      if (ps.s1(base + 13) > q) break;
      thr = divs(thr, q);
    }
    const F3 no = offset_origin(vtx.p, vtx.ng, wi);
    if (!closest<kTwoLevel>(sc, no, wi, L, st, &hi, &ht, &tt, &hb0, &hb1, &hb2)) {
      overflow |= st.overflow;
      radiance = add(radiance, emitted_infinite(sv, thr, s.pdf, s.delta, wi));
      break;
    }
    overflow |= st.overflow;
    RayDiff next;
    if (kTex) next = spawn_scattered(ray, hit_t, vtx, tx, s, no, wi);
    VtxTex ntx;
    const Vtx nv = shading_vertex_t<kTex>(sv, hi, ht, hb0, hb1, hb2, wi, &next, &ntx);
    const VtxMat<kTex> nvm(sv, nv.material, &ntx);
    radiance = add(radiance, emitted(sv, thr, s.pdf, s.delta, tq, nv, nvm.get()));
    vtx = nv;
    vm = nvm;
    if (kTex) {
      tx = ntx;
      ray = next;
      hit_t = tt;
    }
  }
  return radiance;
}

// The camera ray of a pixel sample: with its differential and cone when the
// path tracer carries them.
template <bool kTex>
__device__ __forceinline__ RayDiff camera_ray(const CamView& cam, const smp::PathSampler& ps, uint32_t x, uint32_t y,
                                              uint32_t w, uint32_t h, uint32_t spp) {
  if (kTex) return camera_ray_diff(cam, ps, x, y, w, h, spp);
  RayDiff r;
  r.o = cam.eye;
  r.d = camera_dir(cam, ps, x, y, w, h);
  return r;
}

// The one-thread-per-pixel kernels keep a whole path in registers: without a
// bound their textured / two-level instantiations take 256 VGPRs plus a few
// AGPRs, one wave per SIMD.  Two waves per SIMD caps them at 256 registers
// (a handful of spills, measured with -Rpass-analysis=kernel-resource-usage)
// and doubles the waves that hide traversal latency; three waves would spill
// 90-250 registers per lane.
constexpr int kPathWaves = 2;

// kFilm: some material has a thin film (the film-free instantiation leaves
// the film's interference code out)
template <bool kTwoLevel, bool kTex, bool kFilm>
__global__ __launch_bounds__(kBlock, kPathWaves) void render_kernel(SceneView sc, ShadeView sv, CamView cam, uint32_t w,
                                                        uint32_t h, uint32_t spp, uint32_t max_depth,
                                                        float* __restrict__ rgb, uint32_t* __restrict__ status) {
  __shared__ uint32_t lds[kLdsDepth * kBlock];
  uint32_t spill[kStackDepth - kLdsDepth];
  const uint32_t pix = blockIdx.x * kBlock + threadIdx.x;
  if (pix >= w * h) return;
  LaneStack st;
  st.lds = lds + threadIdx.x;
  st.spill = spill;
  st.sp = 0;
  st.lds_depth = kLdsDepth;
  st.overflow = false;
  Lane L;
  bool overflow = false;
  const uint32_t x = pix % w, y = pix / w;
  const smp::ZSobolFrame zf = smp::zsobol_frame(spp, w, h);
  const Adjoint<float> none{};
  F3 color = f3(0.0f, 0.0f, 0.0f);
  for (uint32_t si = 0; si < spp; ++si) {
    const smp::PathSampler ps{zf, x, y, si};
    const RayDiff ray = camera_ray<kTex>(cam, ps, x, y, w, h, spp);
    constexpr int kMode = kShadeOutline | (kFilm ? 0 : kShadeNoFilm);
    const F3 radiance = trace_path<kTwoLevel, false, kTex, kMode>(sc, sv, ray, ps, max_depth, L, st, overflow, none);
    const float cnt = (float)(si + 1);
    color = add(color, divs(sub(radiance, color), cnt));
  }
  float* o = rgb + 3 * (size_t)pix;
  o[0] = color.x * cam.exposure;
  o[1] = color.y * cam.exposure;
  o[2] = color.z * cam.exposure;
  if (overflow) atomicOr(status, 1u);
}

// Path-replay adjoint of the render loop: per pixel sample a primal pass
// (its radiance) and a replay pass that accumulates dLoss/d(albedo) into
// the block's LDS sums; each block writes its sums to `partials`.
template <bool kTwoLevel, bool kTex, bool kFilm, typename Acc>
__global__ __launch_bounds__(kBlock, kPathWaves) void adjoint_kernel(SceneView sc, ShadeView sv, CamView cam, uint32_t w,
                                                         uint32_t h, uint32_t spp, uint32_t max_depth,
                                                         const float* __restrict__ dimg, Acc* __restrict__ partials,
                                                         uint32_t n_grad, uint32_t rows, float* __restrict__ texgrad,
                                                         uint32_t* __restrict__ status) {
  __shared__ uint32_t lds[kLdsDepth * kBlock];
  extern __shared__ __align__(16) unsigned char dyn[];
  Acc* grad = reinterpret_cast<Acc*>(dyn);
  for (uint32_t i = threadIdx.x; i < n_grad; i += kBlock) grad[i] = (Acc)0;
  __syncthreads();
  uint32_t spill[kStackDepth - kLdsDepth];
  const uint32_t pix = blockIdx.x * kBlock + threadIdx.x;
  if (pix < w * h) {
    LaneStack st;
    st.lds = lds + threadIdx.x;
    st.spill = spill;
    st.sp = 0;
    st.lds_depth = kLdsDepth;
    st.overflow = false;
    Lane L;
    bool overflow = false;
    const uint32_t x = pix % w, y = pix / w;
    const smp::ZSobolFrame zf = smp::zsobol_frame(spp, w, h);
    const float scale = cam.exposure / (float)spp;  // I = exposure * mean over samples
    Adjoint<Acc> adj;
    adj.w = f3(dimg[3 * pix] * scale, dimg[3 * pix + 1] * scale, dimg[3 * pix + 2] * scale);
    adj.grad = grad;
    adj.texgrad = texgrad;
    adj.rows = rows;
    const Adjoint<Acc> none{};
    // a pixel whose image gradient is zero contributes nothing (a loss over
    // a crop or a mask replays only the pixels it reads)
    const uint32_t n_samples = (adj.w.x != 0.0f || adj.w.y != 0.0f || adj.w.z != 0.0f) ? spp : 0u;
    for (uint32_t si = 0; si < n_samples; ++si) {
      const smp::PathSampler ps{zf, x, y, si};
      const RayDiff ray = camera_ray<kTex>(cam, ps, x, y, w, h, spp);
      constexpr int kMode = kShadeOutline | (kFilm ? 0 : kShadeNoFilm);
      adj.total = trace_path<kTwoLevel, false, kTex, kMode>(sc, sv, ray, ps, max_depth, L, st, overflow, none);
      (void)trace_path<kTwoLevel, true, kTex, kMode>(sc, sv, ray, ps, max_depth, L, st, overflow, adj);
    }
    if (overflow) atomicOr(status, 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n_grad; i += kBlock) partials[(size_t)blockIdx.x * n_grad + i] = grad[i];
}

// Sums the per-block partial gradients in FP64.
template <typename Acc>
__global__ void adjoint_reduce(const Acc* __restrict__ partials, uint32_t blocks, uint32_t n_grad,
                               float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_grad) return;
  double s = 0.0;
  for (uint32_t b = 0; b < blocks; ++b) s += (double)partials[(size_t)b * n_grad + i];
  out[i] = (float)s;
}


// ---- batched adjoint over many views ----------------------------------------------
// One view of a batched adjoint: camera, size, samples per pixel, depth, its
// sampler frame, its first pixel in the batch and dLoss/d(image).
struct AdjFrame {
  CamView cam;
  uint32_t w, h, spp, max_depth;
  smp::ZSobolFrame zf;
  uint32_t pix_begin;
  uint32_t pad;
  const float* dimg;
};

// The frame of global pixel gp (frames in pixel order; a binary search).
__device__ __forceinline__ uint32_t adj_frame_of(const AdjFrame* fr, uint32_t n, uint32_t gp) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) / 2;
    if (fr[mid].pix_begin <= gp) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// TwoSum (Knuth): s + e == a + b exactly, s = fl(a + b).
__device__ __forceinline__ float two_sum(float a, float b, float* e) {
  const float s = a + b;
  const float bp = s - a;
  *e = (a - (s - bp)) + (b - bp);
  return s;
}

// Path-replay adjoint of many views in one launch: one thread per pixel of
// every view (the pixel loop of R/main.rs:98-126 per view), a primal pass
// and a replay pass per pixel sample as adjoint_kernel.  The block's
// gradient rows accumulate per sample index in an FP32 LDS epoch buffer
// (LDS atomics over at most one sample of the block's pixels); after every
// sample index the epoch is folded into the block's compensated FP32 sums
// (sum, compensation) with TwoSum, so the long accumulation over samples
// carries no rounding drift; each block writes sum and compensation, which a
// second kernel adds in FP64.  kFp64: FP64 LDS sums and FP64 texel
// gradients instead (the reference accumulation of the error report).
template <bool kTwoLevel, bool kTex, bool kFilm, bool kFp64>
__global__ __launch_bounds__(kBlock, kPathWaves) void adjoint_batch_kernel(SceneView sc, ShadeView sv,
                                                               const AdjFrame* __restrict__ frames, uint32_t nframes,
                                                               uint32_t total, uint32_t max_spp,
                                                               void* __restrict__ partials, uint32_t n_grad,
                                                               void* __restrict__ texgrad,
                                                               uint32_t* __restrict__ status) {
  using Acc = typename std::conditional<kFp64, double, float>::type;
  __shared__ uint32_t lds[kLdsDepth * kBlock];
  extern __shared__ __align__(16) unsigned char dyn[];
  Acc* epoch = reinterpret_cast<Acc*>(dyn);                  // n_grad
  float* sum = reinterpret_cast<float*>(epoch + n_grad);     // n_grad (FP32 mode)
  float* comp = sum + n_grad;                                // n_grad (FP32 mode)
  for (uint32_t i = threadIdx.x; i < n_grad; i += kBlock) {
    epoch[i] = (Acc)0;
    if (!kFp64) sum[i] = comp[i] = 0.0f;
  }
  __syncthreads();
  uint32_t spill[kStackDepth - kLdsDepth];
  const uint32_t gp = blockIdx.x * kBlock + threadIdx.x;
  const bool live = gp < total;
  LaneStack st;
  st.lds = lds + threadIdx.x;
  st.spill = spill;
  st.sp = 0;
  st.lds_depth = kLdsDepth;
  st.overflow = false;
  Lane L;
  bool overflow = false;
  AdjFrame fr;
  uint32_t x = 0, y = 0;
  Adjoint<Acc, Acc> adj;
  adj.grad = epoch;
  adj.texgrad = reinterpret_cast<Acc*>(texgrad);
  adj.rows = kParamRows;
  if (live) {
    fr = frames[adj_frame_of(frames, nframes, gp)];
    const uint32_t lp = gp - fr.pix_begin;
    x = lp % fr.w;
    y = lp / fr.w;
    const float scale = fr.cam.exposure / (float)fr.spp;
    adj.w = f3(fr.dimg[3 * lp] * scale, fr.dimg[3 * lp + 1] * scale, fr.dimg[3 * lp + 2] * scale);
  }
  const Adjoint<Acc, Acc> none{};
  // a pixel whose image gradient is zero contributes nothing
  const bool replay = live && (adj.w.x != 0.0f || adj.w.y != 0.0f || adj.w.z != 0.0f);
  for (uint32_t si = 0; si < max_spp; ++si) {
    if (replay && si < fr.spp) {
      const smp::PathSampler ps{fr.zf, x, y, si};
      const RayDiff ray = camera_ray<kTex>(fr.cam, ps, x, y, fr.w, fr.h, fr.spp);
      constexpr int kMode = kShadeOutline | (kFilm ? 0 : kShadeNoFilm);
      adj.total = trace_path<kTwoLevel, false, kTex, kMode>(sc, sv, ray, ps, fr.max_depth, L, st, overflow, none);
      (void)trace_path<kTwoLevel, true, kTex, kMode>(sc, sv, ray, ps, fr.max_depth, L, st, overflow, adj);
    }
    if (!kFp64) {  // fold the sample's epoch into the compensated sums
      __syncthreads();
      for (uint32_t i = threadIdx.x; i < n_grad; i += kBlock) {
        const float v = (float)epoch[i];
        if (v != 0.0f) {
          float e;
          sum[i] = two_sum(sum[i], v, &e);
          comp[i] += e;
          epoch[i] = (Acc)0;
        }
      }
      __syncthreads();
    }
  }
  if (overflow) atomicOr(status, 1u);
  __syncthreads();
  if (kFp64) {
    double* out = reinterpret_cast<double*>(partials) + (size_t)blockIdx.x * n_grad;
    for (uint32_t i = threadIdx.x; i < n_grad; i += kBlock) out[i] = (double)epoch[i];
  } else {
    float* out = reinterpret_cast<float*>(partials) + (size_t)blockIdx.x * 2 * n_grad;
    for (uint32_t i = threadIdx.x; i < n_grad; i += kBlock) {
      out[2 * i] = sum[i];
      out[2 * i + 1] = comp[i];
    }
  }
}

// Adds the blocks' partial sums in FP64: (sum, compensation) pairs, or FP64
// partials.
template <bool kFp64>
__global__ void adjoint_batch_reduce(const void* __restrict__ partials, uint32_t blocks, uint32_t n_grad,
                                     float* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_grad) return;
  double s = 0.0;
  if (kFp64) {
    const double* p = reinterpret_cast<const double*>(partials);
    for (uint32_t b = 0; b < blocks; ++b) s += p[(size_t)b * n_grad + i];
  } else {
    const float* p = reinterpret_cast<const float*>(partials);
    for (uint32_t b = 0; b < blocks; ++b) {
      const float* q = p + (size_t)b * 2 * n_grad + 2 * i;
      s += (double)q[0] + (double)q[1];
    }
  }
  out[i] = (float)s;
}

}  // namespace pt
}  // namespace tpt

using namespace tpt;

namespace {

// The materials the path tracer implements: OpenPBR without dispersion,
// NormalizedLambert, Emissive and Mirror.
bool supported(const TptMaterial& mat, std::string* why) {
  if (mat.kind == TPT_MATERIAL_NORMALIZED_LAMBERT || mat.kind == TPT_MATERIAL_EMISSIVE ||
      mat.kind == TPT_MATERIAL_MIRROR)
    return true;
  if (mat.kind != TPT_MATERIAL_OPEN_PBR) {
    *why = "unknown material kind " + std::to_string(mat.kind);
    return false;
  }
  const TptOpenPbr& m = mat.u.open_pbr;
  if (m.transmission_dispersion_scale > 0.0f) {
    *why = "transmission_dispersion_scale must be 0";
    return false;
  }
  return true;
}

// The inputs of a material's GGX tables (see upload_ggx_tables).
bool same_ggx_inputs(const TptMaterial& a, const TptMaterial& b) {
  if (a.kind != TPT_MATERIAL_OPEN_PBR) return true;
  const TptOpenPbr &x = a.u.open_pbr, &y = b.u.open_pbr;
  return x.specular_ior == y.specular_ior && x.specular_weight == y.specular_weight &&
         x.coat_weight == y.coat_weight && x.coat_ior == y.coat_ior && x.base_metalness == y.base_metalness &&
         x.base_metalness_texture == y.base_metalness_texture;
}

// Material BSDF queries in the shading frame (tpt_material_bsdf).
__global__ __launch_bounds__(256) void bsdf_query_kernel(pt::Tables tb, const TptMaterial* __restrict__ mats,
                                                         uint32_t material_count, const TptBsdfQuery* __restrict__ q,
                                                         size_t n, TptBsdfResult* __restrict__ out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const TptBsdfQuery a = q[i];
    TptBsdfResult r;
    for (int k = 0; k < 3; ++k) r.f[k] = r.weight[k] = r.wi[k] = 0.0f;
    r.pdf = r.sample_pdf = 0.0f;
    r.flags = 0;
    if (a.material < material_count) {
      const TptMaterial& m = mats[a.material];
      const F3 wo = f3(a.wo[0], a.wo[1], a.wo[2]), wi = f3(a.wi[0], a.wi[1], a.wi[2]);
      pt::Sample smp{};
      bool ok = false;
      F3 f = f3(0.0f, 0.0f, 0.0f);
      if (m.kind == TPT_MATERIAL_OPEN_PBR) {
        pt::Bsdf b;
        pt::make_bsdf(tb, a.material, m.u.open_pbr, a.front_face != 0, wo, &b);
        f = pt::bsdf_eval(b, wo, wi);
        r.pdf = pt::bsdf_pdf(b, wo, wi);
        ok = pt::bsdf_sample(b, wo, a.u[0], a.u[1], a.u[2], &smp);
      } else if (m.kind == TPT_MATERIAL_NORMALIZED_LAMBERT) {  // R/bsdf/normalized_lambert.rs:35-90
        if (wo.z > 0.0f && wi.z > 0.0f) {
          f = pt::divs(pt::rgb3(m.u.lambert.rho), pt::kPi);
          r.pdf = fmaxf(wi.z, 0.0f) / pt::kPi;
        }
        if (wo.z > 0.0f) {
          smp.wi = pt::cosine_hemisphere(a.u[1], a.u[2]);
          smp.pdf = fmaxf(smp.wi.z, 0.0f) / pt::kPi;
          smp.weight = pt::rgb3(m.u.lambert.rho);
          smp.delta = false;
          ok = smp.pdf > 0.0f;
        }
      } else if (m.kind == TPT_MATERIAL_MIRROR && wo.z > 0.0f) {  // R/bsdf/mirror.rs:23-39
        smp.wi = pt::normalize_or_zero(f3(-wo.x, -wo.y, wo.z));
        smp.pdf = 1.0f;
        smp.weight = pt::rgb3(m.u.mirror.color);
        smp.delta = true;
        ok = true;
      }
      r.f[0] = f.x;
      r.f[1] = f.y;
      r.f[2] = f.z;
      if (ok) {
        r.weight[0] = smp.weight.x;
        r.weight[1] = smp.weight.y;
        r.weight[2] = smp.weight.z;
        r.wi[0] = smp.wi.x;
        r.wi[1] = smp.wi.y;
        r.wi[2] = smp.wi.z;
        r.sample_pdf = smp.pdf;
        r.flags = 1u | (smp.delta ? 2u : 0u) | (smp.trans ? 4u : 0u);
      }
    }
    out[i] = r;
  }
}

}  // namespace

namespace tpt {
bool render_materials_ok(TptScene* s, const char* fn, bool* film) {
  std::vector<TptMaterial> mats(s->material_count);
  std::vector<GpuShade> shade(s->instance_count);
  if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice")) return false;
  if (s->material_count &&
      !hip_ok(hipMemcpy(mats.data(), s->mats, mats.size() * sizeof(TptMaterial), hipMemcpyDeviceToHost),
              "hipMemcpy(materials)"))
    return false;
  if (s->instance_count &&
      !hip_ok(hipMemcpy(shade.data(), s->shade, shade.size() * sizeof(GpuShade), hipMemcpyDeviceToHost),
              "hipMemcpy(instances)"))
    return false;
  if (film) {  // thin_film_thickness_nm_at is non-zero (NaN counts: the film path handles every value)
    *film = false;
    for (const TptMaterial& m : mats)
      *film |= m.kind == TPT_MATERIAL_OPEN_PBR && !(m.u.open_pbr.thin_film_weight <= 0.0f) &&
               !(m.u.open_pbr.thin_film_thickness <= 0.0f);
  }
  for (uint32_t i = 0; i < s->instance_count; ++i) {
    if (shade[i].material >= s->material_count) {
      set_error(std::string(fn) + ": instance " + std::to_string(i) + " has no material");
      return false;
    }
    std::string why;
    if (!supported(mats[shade[i].material], &why)) {
      set_error(std::string(fn) + ": material " + std::to_string(shade[i].material) + " not supported: " + why);
      return false;
    }
  }
  return true;
}
}  // namespace tpt

extern "C" int32_t tpt_render(TptScene* s, const TptRenderDesc* desc, float* rgb, uint32_t flags) {
  try {
    if (!s || !desc || !rgb) {
      set_error("tpt_render: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    const uint32_t w = desc->width, h = desc->height;
    if (w == 0 || h == 0) return TPT_OK;
    if ((uint64_t)w * h > 0x7FFFFFFFull || desc->samples_per_pixel == 0) {
      set_error("tpt_render: frame too large or samples_per_pixel == 0");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    bool film = true;
    if (!render_materials_ok(s, "tpt_render", &film)) return TPT_ERROR_BAD_ARGUMENT;
    TptDevice* d = s->dev;
    pt::CamView cam;
    const TptCamera& c = desc->camera;
    cam.eye = F3{c.eye[0], c.eye[1], c.eye[2]};
    cam.fwd = F3{c.forward[0], c.forward[1], c.forward[2]};
    cam.right = F3{c.right[0], c.right[1], c.right[2]};
    cam.up = F3{c.image_up[0], c.image_up[1], c.image_up[2]};
    cam.tan_half = c.tan_half_fov_y;
    cam.exposure = c.exposure;
    cam.spread = c.pixel_spread;
    const pt::ShadeView sv = pt::shade_view_of(s);
    const SceneView v = view_of(s);
    const size_t n = (size_t)w * h;
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    float* drgb = rgb;
    DeviceBuffer tmp;
    if (!dev_ptr) {
      if (!tmp.ensure(n * 12)) return TPT_ERROR_DEVICE;
      drgb = (float*)tmp.ptr;
      if (!hip_ok(hipMemsetAsync(d->status, 0, 4, d->stream), "hipMemsetAsync(status)")) return TPT_ERROR_DEVICE;
    }
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
    const bool two = !(s->instance_count == 1 && s->tlas_node_count == 0);
    const bool tex = s->texture_count > 0;
    auto kernel = film ? (two ? (tex ? pt::render_kernel<true, true, true> : pt::render_kernel<true, false, true>)
                              : (tex ? pt::render_kernel<false, true, true> : pt::render_kernel<false, false, true>))
                       : (two ? (tex ? pt::render_kernel<true, true, false> : pt::render_kernel<true, false, false>)
                              : (tex ? pt::render_kernel<false, true, false> : pt::render_kernel<false, false, false>));
    hipLaunchKernelGGL(kernel, grid, block, 0, d->stream, v, sv, cam, w, h, desc->samples_per_pixel, desc->max_depth,
                       drgb, d->status);
    if (!hip_ok(hipGetLastError(), "render_kernel launch")) return TPT_ERROR_DEVICE;
    if (!dev_ptr) {
      uint32_t status = 0;
      if (!hip_ok(hipMemcpyAsync(rgb, drgb, n * 12, hipMemcpyDeviceToHost, d->stream), "hipMemcpyAsync(rgb)") ||
          !hip_ok(hipMemcpyAsync(&status, d->status, 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpyAsync") ||
          !hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize"))
        return TPT_ERROR_DEVICE;
      if (status & 1u) {
        set_error("tpt_render: traversal stack overflow (128 entries)");
        return TPT_ERROR_OVERFLOW;
      }
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(e.what());
    return TPT_ERROR;
  }
}

namespace {

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

// The emission parameters of a material (what the light tree is built from).
bool same_emission(const TptMaterial& a, const TptMaterial& b) {
  if (a.kind != b.kind) return false;
  if (a.kind == TPT_MATERIAL_EMISSIVE)
    return std::memcmp(&a.u.emissive, &b.u.emissive, sizeof(a.u.emissive)) == 0;
  if (a.kind == TPT_MATERIAL_OPEN_PBR)
    return a.u.open_pbr.emission_luminance == b.u.open_pbr.emission_luminance &&
           std::memcmp(a.u.open_pbr.emission_color, b.u.open_pbr.emission_color, 12) == 0 &&
           a.u.open_pbr.emission_color_texture == b.u.open_pbr.emission_color_texture;
  return true;
}

pt::CamView cam_view(const TptCamera& c) {
  pt::CamView cam;
  cam.eye = F3{c.eye[0], c.eye[1], c.eye[2]};
  cam.fwd = F3{c.forward[0], c.forward[1], c.forward[2]};
  cam.right = F3{c.right[0], c.right[1], c.right[2]};
  cam.up = F3{c.image_up[0], c.image_up[1], c.image_up[2]};
  cam.tan_half = c.tan_half_fov_y;
  cam.exposure = c.exposure;
  cam.spread = c.pixel_spread;
  return cam;
}

template <typename Acc>
int32_t adjoint_impl(TptScene* s, const TptRenderDesc* desc, const float* dimg, float* d_albedo, float* d_texels,
                     bool dev_ptr, uint32_t rows, bool film) {
  TptDevice* d = s->dev;
  const uint32_t w = desc->width, h = desc->height;
  const size_t n = (size_t)w * h;
  const uint32_t n_grad = rows * s->material_count;
  const unsigned blocks = (unsigned)((n + kBlock - 1) / kBlock);
  const size_t nb_part = align_up((size_t)blocks * n_grad * sizeof(Acc));
  const size_t nb_img = dev_ptr ? 0 : align_up(n * 12);
  const size_t nb_out = dev_ptr ? 0 : align_up((size_t)n_grad * 4);
  // texel gradients: an array laid out like the texels, and (host outputs)
  // staging for the packed level-0 gradients
  const size_t n_tg = d_texels ? texel_grad_floats(s) : 0;
  const size_t nb_tbuf = d_texels ? align_up(texel_floats(s) * 4) : 0;
  const size_t nb_tout = (d_texels && !dev_ptr) ? align_up(n_tg * 4) : 0;
  DeviceBuffer tmp;
  if (!tmp.ensure(nb_part + nb_img + nb_out + nb_tbuf + nb_tout)) return TPT_ERROR_DEVICE;
  Acc* partials = (Acc*)tmp.ptr;
  const float* ddimg = dimg;
  float* dgrad = d_albedo;
  float* tbuf = d_texels ? (float*)((char*)tmp.ptr + nb_part + nb_img + nb_out) : nullptr;
  float* tout = d_texels ? (dev_ptr ? d_texels : (float*)((char*)tmp.ptr + nb_part + nb_img + nb_out + nb_tbuf)) : nullptr;
  if (!dev_ptr) {
    ddimg = (const float*)((char*)tmp.ptr + nb_part);
    dgrad = (float*)((char*)tmp.ptr + nb_part + nb_img);
    if (!hip_ok(hipMemcpyAsync((void*)ddimg, dimg, n * 12, hipMemcpyHostToDevice, d->stream), "hipMemcpyAsync(dimg)"))
      return TPT_ERROR_DEVICE;
    if (!hip_ok(hipMemsetAsync(d->status, 0, 4, d->stream), "hipMemsetAsync(status)")) return TPT_ERROR_DEVICE;
  }
  if (tbuf && !hip_ok(hipMemsetAsync(tbuf, 0, nb_tbuf, d->stream), "hipMemsetAsync(texel gradients)"))
    return TPT_ERROR_DEVICE;
  const pt::CamView cam = cam_view(desc->camera);
  const pt::ShadeView sv = pt::shade_view_of(s);
  const SceneView v = view_of(s);
  const bool two = !(s->instance_count == 1 && s->tlas_node_count == 0);
  const size_t lds_bytes = (size_t)n_grad * sizeof(Acc);
  const bool tex = s->texture_count > 0;
  auto kernel = film ? (two ? (tex ? pt::adjoint_kernel<true, true, true, Acc> : pt::adjoint_kernel<true, false, true, Acc>)
                            : (tex ? pt::adjoint_kernel<false, true, true, Acc> : pt::adjoint_kernel<false, false, true, Acc>))
                     : (two ? (tex ? pt::adjoint_kernel<true, true, false, Acc> : pt::adjoint_kernel<true, false, false, Acc>)
                            : (tex ? pt::adjoint_kernel<false, true, false, Acc>
                                   : pt::adjoint_kernel<false, false, false, Acc>));
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), lds_bytes, d->stream, v, sv, cam, w, h,
                     desc->samples_per_pixel, desc->max_depth, ddimg, partials, n_grad, rows, tbuf, d->status);
  hipLaunchKernelGGL((pt::adjoint_reduce<Acc>), dim3((n_grad + 255) / 256), dim3(256), 0, d->stream,
                     (const Acc*)partials, blocks, n_grad, dgrad);
  if (!hip_ok(hipGetLastError(), "adjoint launch")) return TPT_ERROR_DEVICE;
  if (tbuf && !texel_grads_to_level0(s, tbuf, tout, d->stream)) return TPT_ERROR_DEVICE;
  // the scratch buffer is released when this call returns: wait for the kernels
  uint32_t status = 0;
  if (!dev_ptr &&
      (!hip_ok(hipMemcpyAsync(d_albedo, dgrad, (size_t)n_grad * 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpyAsync") ||
       (d_texels && n_tg &&
        !hip_ok(hipMemcpyAsync(d_texels, tout, n_tg * 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpyAsync")) ||
       !hip_ok(hipMemcpyAsync(&status, d->status, 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpyAsync")))
    return TPT_ERROR_DEVICE;
  if (!hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize")) return TPT_ERROR_DEVICE;
  if (status & 1u) {
    set_error("tpt_render_adjoint: traversal stack overflow (128 entries)");
    return TPT_ERROR_OVERFLOW;
  }
  return TPT_OK;
}

}  // namespace

namespace {
int32_t render_adjoint(TptScene* s, const TptRenderDesc* desc, const float* dloss_dimage, float* d_albedo,
                       float* d_texels, uint32_t flags, uint32_t rows = 3) {
  try {
    if (!s || !desc || !dloss_dimage || !d_albedo) {
      set_error("tpt_render_adjoint: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    const uint32_t w = desc->width, h = desc->height;
    if ((uint64_t)w * h > 0x7FFFFFFFull || desc->samples_per_pixel == 0) {
      set_error("tpt_render_adjoint: frame too large or samples_per_pixel == 0");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    const uint32_t max_mats = rows == 3 ? 4096u : 2048u;  // the block's gradient rows live in LDS
    if (s->material_count == 0 || s->material_count > max_mats) {
      set_error("tpt_render_adjoint: needs 1.." + std::to_string(max_mats) + " materials");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    bool film = true;
    if (!render_materials_ok(s, "tpt_render_adjoint", &film)) return TPT_ERROR_BAD_ARGUMENT;
    if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    if (rows == pt::kParamRows && !s->cond_tables) {
      // the metalness derivative reads the conductor tables even where a
      // material's metalness is 0 (tables are otherwise built only for
      // materials with a non-zero metalness)
      s->force_cond = true;
      if (!upload_ggx_tables(s, s->host_mats)) return TPT_ERROR_DEVICE;
    }
    const size_t n_tg = d_texels ? texel_grad_floats(s) : 0;
    if ((uint64_t)w * h == 0) {
      if (flags & TPT_FLAG_DEVICE_POINTERS)
        return hip_ok(hipMemsetAsync(d_albedo, 0, 4 * rows * (size_t)s->material_count, s->dev->stream),
                      "hipMemsetAsync") &&
                       (!n_tg || hip_ok(hipMemsetAsync(d_texels, 0, 4 * n_tg, s->dev->stream), "hipMemsetAsync"))
                   ? TPT_OK
                   : TPT_ERROR_DEVICE;
      std::memset(d_albedo, 0, 4 * rows * (size_t)s->material_count);
      if (n_tg) std::memset(d_texels, 0, 4 * n_tg);
      return TPT_OK;
    }
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    return (flags & TPT_FLAG_GRAD_FP64)
               ? adjoint_impl<double>(s, desc, dloss_dimage, d_albedo, d_texels, dev_ptr, rows, film)
               : adjoint_impl<float>(s, desc, dloss_dimage, d_albedo, d_texels, dev_ptr, rows, film);
  } catch (const std::exception& e) {
    set_error(std::string("tpt_render_adjoint: ") + e.what());
    return TPT_ERROR;
  }
}
}  // namespace

extern "C" int32_t tpt_render_adjoint(TptScene* s, const TptRenderDesc* desc, const float* dloss_dimage,
                                      float* d_albedo, uint32_t flags) {
  return render_adjoint(s, desc, dloss_dimage, d_albedo, nullptr, flags);
}

extern "C" int32_t tpt_render_adjoint_texels(TptScene* s, const TptRenderDesc* desc, const float* dloss_dimage,
                                             float* d_albedo, float* d_texels, uint32_t flags) {
  if (s && !d_texels) {
    set_error("tpt_render_adjoint_texels: null texel gradient output");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  return render_adjoint(s, desc, dloss_dimage, d_albedo, d_texels, flags);
}

extern "C" int32_t tpt_render_adjoint_params(TptScene* s, const TptRenderDesc* desc, const float* dloss_dimage,
                                             float* d_params, float* d_texels, uint32_t flags) {
  return render_adjoint(s, desc, dloss_dimage, d_params, d_texels, flags, pt::kParamRows);
}

namespace {
// The batched adjoint (tpt_render_adjoint_batch), kFp64: FP64 block sums and
// texel gradients.
template <bool kFp64>
int32_t adjoint_batch_impl(TptScene* s, const TptAdjointFrameDesc* frames, uint32_t nframes, float* d_params,
                           float* d_texels, bool dev_ptr, bool film) {
  using Acc = typename std::conditional<kFp64, double, float>::type;
  TptDevice* d = s->dev;
  const uint32_t n_grad = pt::kParamRows * s->material_count;
  std::vector<pt::AdjFrame> host(nframes);
  uint64_t total = 0;
  uint32_t max_spp = 0;
  size_t img_floats = 0;
  for (uint32_t f = 0; f < nframes; ++f) {
    const TptRenderDesc& r = frames[f].render;
    pt::AdjFrame& a = host[f];
    a.cam = cam_view(r.camera);
    a.w = r.width;
    a.h = r.height;
    a.spp = r.samples_per_pixel;
    a.max_depth = r.max_depth;
    a.zf = smp::zsobol_frame(r.samples_per_pixel, r.width, r.height);
    a.pix_begin = (uint32_t)total;
    a.pad = 0;
    a.dimg = frames[f].dloss_dimage;
    total += (uint64_t)r.width * r.height;
    img_floats += (size_t)r.width * r.height * 3;
    max_spp = std::max(max_spp, r.samples_per_pixel);
  }
  const unsigned blocks = (unsigned)((total + kBlock - 1) / kBlock);
  const size_t nb_frames = align_up(nframes * sizeof(pt::AdjFrame));
  const size_t nb_part = align_up((size_t)blocks * n_grad * (kFp64 ? 8 : 8));  // FP64 partials or (sum, comp) pairs
  const size_t nb_img = dev_ptr ? 0 : align_up(img_floats * 4);
  const size_t nb_out = dev_ptr ? 0 : align_up((size_t)n_grad * 4);
  const size_t n_tg = d_texels ? texel_grad_floats(s) : 0;
  const size_t nb_tbuf = d_texels ? align_up(texel_floats(s) * sizeof(Acc)) : 0;
  const size_t nb_tout = (d_texels && !dev_ptr) ? align_up(n_tg * 4) : 0;
  DeviceBuffer tmp;
  if (!tmp.ensure(nb_frames + nb_part + nb_img + nb_out + nb_tbuf + nb_tout)) return TPT_ERROR_DEVICE;
  char* base = (char*)tmp.ptr;
  pt::AdjFrame* dframes = (pt::AdjFrame*)base;
  void* partials = base + nb_frames;
  float* dimg_stage = (float*)(base + nb_frames + nb_part);
  float* dgrad = dev_ptr ? d_params : (float*)(base + nb_frames + nb_part + nb_img);
  Acc* tbuf = d_texels ? (Acc*)(base + nb_frames + nb_part + nb_img + nb_out) : nullptr;
  float* tout = d_texels ? (dev_ptr ? d_texels : (float*)(base + nb_frames + nb_part + nb_img + nb_out + nb_tbuf))
                         : nullptr;
  hipStream_t st = d->stream;
  if (!dev_ptr) {  // stage every view's dLoss/d(image)
    size_t o = 0;
    for (uint32_t f = 0; f < nframes; ++f) {
      const size_t nf = (size_t)host[f].w * host[f].h * 3;
      if (!hip_ok(hipMemcpyAsync(dimg_stage + o, frames[f].dloss_dimage, nf * 4, hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync(dloss)"))
        return TPT_ERROR_DEVICE;
      host[f].dimg = dimg_stage + o;
      o += nf;
    }
  }
  if (!hip_ok(hipMemcpyAsync(dframes, host.data(), nframes * sizeof(pt::AdjFrame), hipMemcpyHostToDevice, st),
              "hipMemcpyAsync(frames)") ||
      !hip_ok(hipMemsetAsync(d->status, 0, 4, st), "hipMemsetAsync(status)"))
    return TPT_ERROR_DEVICE;
  if (tbuf && !hip_ok(hipMemsetAsync(tbuf, 0, nb_tbuf, st), "hipMemsetAsync(texel gradients)")) return TPT_ERROR_DEVICE;
  const pt::ShadeView sv = pt::shade_view_of(s);
  const SceneView v = view_of(s);
  const bool two = !(s->instance_count == 1 && s->tlas_node_count == 0);
  const bool tex = s->texture_count > 0;
  const size_t lds_bytes = (size_t)n_grad * (kFp64 ? sizeof(double) : 3 * sizeof(float));
  auto kernel =
      film ? (two ? (tex ? pt::adjoint_batch_kernel<true, true, true, kFp64> : pt::adjoint_batch_kernel<true, false, true, kFp64>)
                  : (tex ? pt::adjoint_batch_kernel<false, true, true, kFp64>
                         : pt::adjoint_batch_kernel<false, false, true, kFp64>))
           : (two ? (tex ? pt::adjoint_batch_kernel<true, true, false, kFp64>
                         : pt::adjoint_batch_kernel<true, false, false, kFp64>)
                  : (tex ? pt::adjoint_batch_kernel<false, true, false, kFp64>
                         : pt::adjoint_batch_kernel<false, false, false, kFp64>));
  hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), lds_bytes, st, v, sv, (const pt::AdjFrame*)dframes, nframes,
                     (uint32_t)total, max_spp, partials, n_grad, (void*)tbuf, d->status);
  hipLaunchKernelGGL((pt::adjoint_batch_reduce<kFp64>), dim3((n_grad + 255) / 256), dim3(256), 0, st,
                     (const void*)partials, blocks, n_grad, dgrad);
  if (!hip_ok(hipGetLastError(), "adjoint batch launch")) return TPT_ERROR_DEVICE;
  if (tbuf && !texel_grads_to_level0<Acc>(s, tbuf, tout, st)) return TPT_ERROR_DEVICE;
  uint32_t status = 0;
  if (!dev_ptr &&
      (!hip_ok(hipMemcpyAsync(d_params, dgrad, (size_t)n_grad * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync") ||
       (d_texels && n_tg &&
        !hip_ok(hipMemcpyAsync(d_texels, tout, n_tg * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync"))))
    return TPT_ERROR_DEVICE;
  if (!hip_ok(hipMemcpyAsync(&status, d->status, 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(status)") ||
      !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))  // the scratch is released on return
    return TPT_ERROR_DEVICE;
  if (status & 1u) {
    set_error("tpt_render_adjoint_batch: traversal stack overflow (128 entries)");
    return TPT_ERROR_OVERFLOW;
  }
  return TPT_OK;
}
}  // namespace

extern "C" int32_t tpt_render_adjoint_batch(TptScene* s, const TptAdjointFrameDesc* frames, uint32_t frame_count,
                                            float* d_params, float* d_texels, uint32_t flags) {
  try {
    if (!s || !frames || !frame_count || !d_params) {
      set_error("tpt_render_adjoint_batch: null argument or no frame");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    uint64_t total = 0;
    for (uint32_t f = 0; f < frame_count; ++f) {
      const TptRenderDesc& r = frames[f].render;
      if (!frames[f].dloss_dimage || r.width == 0 || r.height == 0 || r.samples_per_pixel == 0) {
        set_error("tpt_render_adjoint_batch: frame " + std::to_string(f) +
                  " has a null dloss_dimage, an empty size or samples_per_pixel 0");
        return TPT_ERROR_BAD_ARGUMENT;
      }
      total += (uint64_t)r.width * r.height;
    }
    if (total > 0x7FFFFF00ull) {
      set_error("tpt_render_adjoint_batch: more than 2^31 pixels in one batch");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    // the block's rows live in LDS next to the traversal stacks: epoch,
    // sum and compensation (FP32) or FP64 sums, 5 rows per material
    if (s->material_count == 0 || s->material_count > 1024) {
      set_error("tpt_render_adjoint_batch: needs 1..1024 materials");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    bool film = true;
    if (!render_materials_ok(s, "tpt_render_adjoint_batch", &film)) return TPT_ERROR_BAD_ARGUMENT;
    if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    if (!s->cond_tables) {  // the metalness derivative reads the conductor tables
      s->force_cond = true;
      if (!upload_ggx_tables(s, s->host_mats)) return TPT_ERROR_DEVICE;
    }
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    return (flags & TPT_FLAG_GRAD_FP64)
               ? adjoint_batch_impl<true>(s, frames, frame_count, d_params, d_texels, dev_ptr, film)
               : adjoint_batch_impl<false>(s, frames, frame_count, d_params, d_texels, dev_ptr, film);
  } catch (const std::exception& e) {
    set_error(std::string("tpt_render_adjoint_batch: ") + e.what());
    return TPT_ERROR;
  }
}

extern "C" int32_t tpt_scene_set_materials(TptScene* s, const TptMaterial* materials, uint32_t count) {
  try {
    if (!s || (count && !materials)) {
      set_error("tpt_scene_set_materials: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (count != s->material_count) {
      set_error("tpt_scene_set_materials: the material count must stay " + std::to_string(s->material_count));
      return TPT_ERROR_BAD_ARGUMENT;
    }
    for (uint32_t i = 0; i < count; ++i)
      if (!same_emission(materials[i], s->host_mats[i])) {
        set_error("tpt_scene_set_materials: material " + std::to_string(i) +
                  " changes kind or emission (the light tree is built from emission: create a new scene)");
        return TPT_ERROR_BAD_ARGUMENT;
      }
    for (uint32_t i = 0; i < count; ++i) {
      const std::string terr = material_texture_error(materials[i], i, s->host_textures);
      if (!terr.empty()) {
        set_error("tpt_scene_set_materials: " + terr);
        return TPT_ERROR_BAD_ARGUMENT;
      }
    }
    if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    if (count && !hip_ok(hipMemcpyAsync(s->mats, materials, count * sizeof(TptMaterial), hipMemcpyHostToDevice,
                                        s->dev->stream),
                         "hipMemcpyAsync(materials)"))
      return TPT_ERROR_DEVICE;
    if (!hip_ok(hipStreamSynchronize(s->dev->stream), "hipStreamSynchronize")) return TPT_ERROR_DEVICE;
    bool tables = false;
    for (uint32_t i = 0; i < count; ++i) tables |= !same_ggx_inputs(materials[i], s->host_mats[i]);
    s->host_mats.assign(materials, materials + count);
    if (tables && !upload_ggx_tables(s, s->host_mats)) return TPT_ERROR_DEVICE;
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_scene_set_materials: ") + e.what());
    return TPT_ERROR;
  }
}

extern "C" int32_t tpt_scene_set_nee_policy(TptScene* s, uint32_t policy) {
  if (!s) {
    set_error("tpt_scene_set_nee_policy: null scene");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  if (policy != TPT_NEE_REFERENCE && policy != TPT_NEE_OPENPBR_PROXY) {
    set_error("tpt_scene_set_nee_policy: unknown policy " + std::to_string(policy));
    return TPT_ERROR_BAD_ARGUMENT;
  }
  s->nee_policy = policy;
  return TPT_OK;
}

extern "C" int32_t tpt_material_bsdf(TptScene* s, const TptBsdfQuery* queries, size_t n, TptBsdfResult* out,
                                     uint32_t flags) {
  try {
    if (!s || (n && (!queries || !out))) {
      set_error("tpt_material_bsdf: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (n == 0) return TPT_OK;
    if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    const hipStream_t st = s->dev->stream;
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    DeviceBuffer qb, rb;
    const TptBsdfQuery* dq = queries;
    TptBsdfResult* dr = out;
    if (!dev_ptr) {
      if (!qb.ensure(n * sizeof(TptBsdfQuery)) || !rb.ensure(n * sizeof(TptBsdfResult))) return TPT_ERROR_DEVICE;
      if (!hip_ok(hipMemcpyAsync(qb.ptr, queries, n * sizeof(TptBsdfQuery), hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync(queries)"))
        return TPT_ERROR_DEVICE;
      dq = (const TptBsdfQuery*)qb.ptr;
      dr = (TptBsdfResult*)rb.ptr;
    }
    const pt::ShadeView sv = pt::shade_view_of(s);
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(bsdf_query_kernel, dim3(blocks), dim3(256), 0, st, sv.tab, (const TptMaterial*)s->mats,
                       s->material_count, dq, n, dr);
    if (!hip_ok(hipGetLastError(), "bsdf_query_kernel launch")) return TPT_ERROR_DEVICE;
    if (!dev_ptr) {
      if (!hip_ok(hipMemcpyAsync(out, dr, n * sizeof(TptBsdfResult), hipMemcpyDeviceToHost, st),
                  "hipMemcpyAsync(results)") ||
          !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return TPT_ERROR_DEVICE;
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_material_bsdf: ") + e.what());
    return TPT_ERROR;
  }
}
