// libtpt: MI355X (gfx950) ray queries over the reference QBVH.
//
// Kernels: closest hit (R/scene/mod.rs:340-395 over R/qbvh.rs:611-675 and
// R/math/ray.rs:107-209) and shadow visibility (R/integrator/mod.rs:239-265),
// each over FP32 or 8-bit quantised nodes.  One ray per lane, wave64.
//
// Bit-exactness with the reference CPU code rests on:
//  * the same float expression trees (built with -ffp-contract=off; fma only
//    where the reference calls mul_add; IEEE division and square root);
//  * the slab test's compare-select min/max (x86 minps/maxps semantics of
//    wide::f32x4::fast_min/fast_max), written as explicit selects;
//  * the reference visit order: children sorted near->far by (tmin, slot),
//    leaves handled far->near before the next pop, interiors pushed far->near.
//
// Traversal stack: a per-lane LDS stack of kLdsDepth entries laid out
// [depth][thread] (bank = thread, conflict-free), continued in a per-lane
// private array up to the reference's 128 entries (R/qbvh.rs:609).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "environment.h"
#include "ggx_tables.h"
#include "q8_encode.h"
#include "tpt_device.h"
#include "tpt_internal.h"
#include "trace_kernel.h"
#include "traverse.h"

namespace tpt {

// ---------------------------------------------------------------------------
// host side

namespace {
thread_local std::string g_last_error = "";
}  // namespace

void set_error(const std::string& s) { g_last_error = s; }

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

}  // namespace tpt


struct TptQbvh {
  tpt::HostQbvh q;
};

namespace tpt {
namespace {

// The 8-bit copies of n FP32 nodes, encoded on the device (q8_encode.h).
__global__ void encode_q8_kernel(const GpuNode* __restrict__ nodes, uint32_t n, GpuNodeQ8* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) quantize_node_impl(nodes[i], &out[i]);
}
bool encode_q8(TptDevice* d, const GpuNode* nodes, size_t n, GpuNodeQ8** dst, size_t* bytes) {
  *dst = nullptr;
  const size_t nb = (n ? n : 1) * sizeof(GpuNodeQ8);
  if (!hip_ok(hipMalloc((void**)dst, nb), "hipMalloc")) return false;
  *bytes += nb;
  if (!n) return true;
  hipLaunchKernelGGL(encode_q8_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, d->stream, nodes, (uint32_t)n,
                     *dst);
  return hip_ok(hipGetLastError(), "encode_q8_kernel") && hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
}

template <class T>
bool upload(const std::vector<T>& v, T** dst, size_t* bytes) {
  *dst = nullptr;
  const size_t nb = (v.empty() ? 1 : v.size()) * sizeof(T);
  if (!hip_ok(hipMalloc((void**)dst, nb), "hipMalloc")) return false;
  *bytes += nb;
  if (!v.empty() && !hip_ok(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy"))
    return false;
  return true;
}

void free_scene(TptScene* s) {
  if (!s) return;
  (void)hipFree(s->nodes);
  (void)hipFree(s->qnodes);
  (void)hipFree(s->tris);
  (void)hipFree(s->tri_rank);
  (void)hipFree(s->inst);
  (void)hipFree(s->tlas);
  (void)hipFree(s->qtlas);
  (void)hipFree(s->tlas_prim);
  (void)hipFree(s->stage);
  (void)hipFree(s->shade);
  (void)hipFree(s->vpos);
  (void)hipFree(s->vnrm);
  (void)hipFree(s->vuv);
  (void)hipFree(s->vidx);
  (void)hipFree(s->mats);
  (void)hipFree(s->lights);
  (void)hipFree(s->tri_leaf);
  (void)hipFree(s->env_px);
  (void)hipFree(s->env_pyr);
  (void)hipFree(s->env_levels);
  (void)hipFree(s->ggx_albedo);
  (void)hipFree(s->ggx_mat);
  (void)hipFree(s->ggx_cond);
  (void)hipFree(s->ggx_diel);
  (void)hipFree(s->light_cats);
  (void)hipFree(s->tex_texels);
  (void)hipFree(s->tex_levels);
  (void)hipFree(s->textures);
  delete s;
}

}  // namespace

// GGX tables of the scene's OpenPBR materials (see TptScene): every
// distinct specular IOR and, for coated materials, coat IOR gets its
// directional-albedo table; the energy-compensation tables are uploaded
// when some material reads them with a non-zero weight (metalness > 0 for
// the conductor's; a specular IOR other than 1 or a coat for the
// dielectric's) and as zeros otherwise.
// LightSampler::build_from_scene (R/light/mod.rs:133-191): the entries with a
// positive weight in the reference's order, the running total in f32 and the
// CDF scaled by 1 / total.
bool build_light_cats(TptScene* s, const TptSceneDesc& d) {
  pt::LightCats& t = s->host_cats;
  t = pt::LightCats{};
  if (d.directional_count > pt::kMaxDirectional) {
    set_error("tpt_scene_create: at most " + std::to_string(pt::kMaxDirectional) + " directional lights");
    return false;
  }
  if (d.directional_count && !d.directional_lights) {
    set_error("tpt_scene_create: directional_count > 0 with a null directional light table");
    return false;
  }
  auto add = [&](uint32_t kind, float w) {
    t.kind[t.count] = kind;
    t.weight[t.count] = w;
    t.count++;
  };
  if (s->w_tree > 0.0f) add(0, s->w_tree);
  if (s->w_env > 0.0f) add(1, s->w_env);
  for (uint32_t i = 0; i < d.directional_count; ++i) {
    const TptDirectionalLight& l = d.directional_lights[i];
    for (int c = 0; c < 3; ++c) {
      t.dir[i][c] = l.direction[c];
      t.radiance[i][c] = l.color[c] * l.intensity;  // sample_li: color * intensity
    }
    const float lum = 0.2126f * l.color[0] + 0.7152f * l.color[1] + 0.0722f * l.color[2];  // sg::luminance
    const float w = std::fmax(lum * l.intensity, 0.0f);
    if (w > 0.0f) add(2 + i, w);
  }
  float total = 0.0f;
  t.cdf[0] = 0.0f;
  for (uint32_t k = 0; k < t.count; ++k) {
    total += t.weight[k];
    t.cdf[k + 1] = total;
  }
  if (total > 0.0f) {
    const float inv = 1.0f / total;
    for (uint32_t k = 0; k <= t.count; ++k) t.cdf[k] *= inv;
  }
  t.total = total;
  std::vector<pt::LightCats> one(1, t);
  return upload(one, &s->light_cats, &s->bytes);
}

bool upload_ggx_tables(TptScene* s, const std::vector<TptMaterial>& mats) {
  std::vector<float> albedo(ggx::kAlbedoLen, 0.0f);
  std::map<uint32_t, uint32_t> index;
  auto table_of = [&](float eta) {
    const float e = (std::isfinite(eta) && eta > 0.0f) ? eta : 1.5f;
    uint32_t key;
    std::memcpy(&key, &e, 4);
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    const std::vector<float>& t = ggx::albedo_table(e);
    const uint32_t id = (uint32_t)(albedo.size() / ggx::kAlbedoLen);
    albedo.insert(albedo.end(), t.begin(), t.end());
    index.emplace(key, id);
    return id;
  };
  std::vector<uint32_t> mt(2 * std::max<size_t>(1, mats.size()), 0u);
  bool need_cond = false, need_diel = false;
  for (size_t i = 0; i < mats.size(); ++i) {
    if (mats[i].kind != TPT_MATERIAL_OPEN_PBR) continue;
    const TptOpenPbr& m = mats[i].u.open_pbr;
    const float eta = ggx::modulated_eta(std::fmax(m.specular_ior, 1.0e-4f), std::fmax(m.specular_weight, 0.0f));
    mt[2 * i] = table_of(eta);
    const float cw = m.coat_weight < 0.0f ? 0.0f : (m.coat_weight > 1.0f ? 1.0f : m.coat_weight);
    if (cw > 0.0f) {
      mt[2 * i + 1] = table_of(std::fmax(m.coat_ior, 1.0e-4f));
      need_diel = true;
    }
    if (eta != 1.0f) need_diel = true;
    // metalness clamp(base_metalness * texel, 0, 1) can be positive only
    // for a non-zero parameter
    if (m.base_metalness > 0.0f || (m.base_metalness_texture != TPT_INVALID_INDEX && m.base_metalness != 0.0f))
      need_cond = true;
  }
  const size_t ncond = ggx::kCondCos * ggx::kCondRough + ggx::kCondRough;
  const size_t ndiel = ggx::kDielCos * ggx::kDielRough * ggx::kDielEta + ggx::kDielRough * ggx::kDielEta;
  need_cond = need_cond || s->force_cond;
  s->cond_tables = need_cond;
  const std::vector<float> cond = need_cond ? ggx::conductor_table() : std::vector<float>(ncond, 0.0f);
  const std::vector<float> diel = need_diel ? ggx::dielectric_table() : std::vector<float>(ndiel, 0.0f);
  (void)hipFree(s->ggx_albedo);
  (void)hipFree(s->ggx_mat);
  (void)hipFree(s->ggx_cond);
  (void)hipFree(s->ggx_diel);
  s->ggx_albedo = nullptr;
  s->ggx_mat = nullptr;
  s->ggx_cond = nullptr;
  s->ggx_diel = nullptr;
  return upload(albedo, &s->ggx_albedo, &s->bytes) && upload(mt, &s->ggx_mat, &s->bytes) &&
         upload(cond, &s->ggx_cond, &s->bytes) && upload(diel, &s->ggx_diel, &s->bytes);
}

namespace {

// Upper bound of the traversal stack: along any root-to-node path, a visit
// leaves (interior children - 1) entries below the next pop.
int stack_bound(const std::vector<TptQbvhNode>& nodes, uint32_t root, int depth_guard) {
  if (nodes.empty() || depth_guard > 256) return 0;
  const TptQbvhNode& n = nodes[root];
  int interior = 0, best = 0;
  for (int i = 0; i < 4; ++i) {
    if (!(n.valid_mask & (1u << i))) continue;
    const uint32_t c = n.children[i];
    if (c & kChildLeafBit) continue;
    if (c >= nodes.size()) return 1 << 20;
    interior++;
    best = std::max(best, stack_bound(nodes, c, depth_guard + 1));
  }
  return interior + best;
}

bool ensure_stage(TptScene* s, size_t nb) {
  if (s->stage_bytes >= nb) return true;
  (void)hipFree(s->stage);
  s->stage = nullptr;
  s->stage_bytes = 0;
  if (!hip_ok(hipMalloc(&s->stage, nb), "hipMalloc(stage)")) return false;
  s->stage_bytes = nb;
  return true;
}

bool read_status(TptDevice* d, uint32_t* v) {
  return hip_ok(hipMemcpyAsync(v, d->status, 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpyAsync(status)") &&
         hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
}

}  // namespace
}  // namespace tpt

using namespace tpt;

extern "C" {

const char* tpt_last_error(void) { return g_last_error.c_str(); }

TptDevice* tpt_device_create(int32_t device_index) {
  try {
    int count = 0;
    if (!hip_ok(hipGetDeviceCount(&count), "hipGetDeviceCount")) return nullptr;
    if (device_index < 0 || device_index >= count) {
      set_error("tpt_device_create: device index out of range");
      return nullptr;
    }
    if (!hip_ok(hipSetDevice(device_index), "hipSetDevice")) return nullptr;
    TptDevice* d = new TptDevice();
    d->index = device_index;
    if (!hip_ok(hipStreamCreateWithFlags(&d->own, hipStreamNonBlocking), "hipStreamCreate")) {
      delete d;
      return nullptr;
    }
    d->stream = d->own;
    hipDeviceProp_t prop;
    if (!hip_ok(hipGetDeviceProperties(&prop, device_index), "hipGetDeviceProperties")) {
      (void)hipStreamDestroy(d->own);
      delete d;
      return nullptr;
    }
    d->cu_count = prop.multiProcessorCount;
    if (!hip_ok(hipMalloc((void**)&d->status, 4), "hipMalloc(status)") ||
        !hip_ok(hipMemset(d->status, 0, 4), "hipMemset(status)") ||
        !hip_ok(hipMalloc((void**)&d->work, sizeof(WorkCounters)), "hipMalloc(work)") ||
        !hip_ok(hipMemset(d->work, 0, sizeof(WorkCounters)), "hipMemset(work)") || !noalloc_setup(d)) {
      (void)hipFree(d->status);
      (void)hipFree(d->work);
      (void)hipStreamDestroy(d->own);
      delete d;
      return nullptr;
    }
    return d;
  } catch (const std::exception& e) {
    set_error(e.what());
    return nullptr;
  }
}

void tpt_device_release(TptDevice* d) {
  if (!d) return;
  (void)hipSetDevice(d->index);
  (void)hipStreamSynchronize(d->stream);
  (void)hipFree(d->status);
  (void)hipFree(d->work);
  (void)hipStreamDestroy(d->own);
  delete d;
}

int32_t tpt_device_set_param(TptDevice* d, const char* name, int64_t value) {
  if (!d || !name) {
    set_error("tpt_device_set_param: null argument");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  const std::string n(name);
  if (n == "regions" && value >= 1 && value <= kMaxRegions) {
    d->regions = (uint32_t)value;
    return TPT_OK;
  }
  if (n == "refill_grab" && value >= 0 && value <= 65536) {
    d->refill_grab = (uint32_t)value;
    return TPT_OK;
  }
  if (n == "refill_below" && value >= 1 && value <= 64) {
    d->refill_below = (uint32_t)value;
  } else if (n == "blocks_per_cu" && value >= 0 && value <= 64) {
    d->blocks_per_cu = (uint32_t)value;
  } else if (n == "wf_paths" && value >= 65536 && value <= (1 << 26)) {
    d->wf_paths = (uint32_t)value;
  } else if (n == "wf_graph" && (value == 0 || value == 1)) {
    d->wf_graph = (uint32_t)value;
  } else if (n == "lds_depth" && value >= kMinLdsDepth && value <= kMaxLdsDepth) {
    d->lds_depth = (uint32_t)value;
  } else if (n == "wf_tile" && (value == 128 || value == 256 || value == 512)) {
    d->wf_tile = (uint32_t)value;
  } else if (n == "wf_sort" && value >= 0 && value <= 64) {
    d->wf_sort = (uint32_t)value;
  } else {
    set_error("tpt_device_set_param: unknown parameter or value out of range: " + n);
    return TPT_ERROR_BAD_ARGUMENT;
  }
  return TPT_OK;
}

int32_t tpt_device_set_stream(TptDevice* d, void* stream) {
  if (!d) {
    set_error("tpt_device_set_stream: null device");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  d->stream = stream ? (hipStream_t)stream : d->own;
  return TPT_OK;
}

int32_t tpt_device_synchronize(TptDevice* d) {
  if (!d) {
    set_error("tpt_device_synchronize: null device");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  return hip_ok(hipStreamSynchronize(d->stream), "hipStreamSynchronize") ? TPT_OK : TPT_ERROR_DEVICE;
}

TptQbvh* tpt_qbvh_build(const float* bmin, const float* bmax, size_t n, int32_t threads) {
  try {
    if (n == 0) {
      set_error("tpt_qbvh_build: no primitives (build_qbvh returns None)");
      return nullptr;
    }
    if (!bmin || !bmax) {
      set_error("tpt_qbvh_build: null bounds");
      return nullptr;
    }
    TptQbvh* q = new TptQbvh();
    const int rc = build_qbvh(bmin, bmax, n, threads, &q->q);
    if (rc != 0) {
      set_error(rc == -2 ? "tpt_qbvh_build: leaf offset exceeds the 24-bit child encoding"
                         : "tpt_qbvh_build: bad input");
      delete q;
      return nullptr;
    }
    return q;
  } catch (const std::exception& e) {
    set_error(e.what());
    return nullptr;
  }
}

TptQbvh* tpt_qbvh_build_device(TptDevice* d, const float* bmin, const float* bmax, size_t n, uint32_t flags) {
  try {
    if (!d) {
      set_error("tpt_qbvh_build_device: null device");
      return nullptr;
    }
    if (n == 0) {
      set_error("tpt_qbvh_build_device: no primitives (build_qbvh returns None)");
      return nullptr;
    }
    if (!bmin || !bmax) {
      set_error("tpt_qbvh_build_device: null bounds");
      return nullptr;
    }
    if (!hip_ok(hipSetDevice(d->index), "hipSetDevice")) return nullptr;
    TptQbvh* q = new TptQbvh();
    const int rc = build_qbvh_device(d, bmin, bmax, n, (flags & TPT_FLAG_DEVICE_POINTERS) != 0, &q->q);
    if (rc != 0) {
      if (rc == -2) set_error("tpt_qbvh_build_device: leaf offset exceeds the 24-bit child encoding");
      else if (rc == -3)
        set_error("tpt_qbvh_build_device: a forced median split over non-coincident centroids (the reference "
                  "sorts them; use tpt_qbvh_build)");
      else if (rc == -1) set_error("tpt_qbvh_build_device: more than 2^31 - 1 primitives");
      delete q;
      return nullptr;
    }
    return q;
  } catch (const std::exception& e) {
    set_error(e.what());
    return nullptr;
  }
}

void tpt_qbvh_release(TptQbvh* q) { delete q; }

int32_t tpt_scene_q8_nodes(const TptScene* s, uint32_t tlas, void* out, size_t capacity, size_t* count) {
  if (!s || !count) {
    set_error("tpt_scene_q8_nodes: null argument");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  const size_t n = tlas ? s->tlas_node_count : s->node_count;
  *count = n;
  if (!out) return TPT_OK;
  if (capacity < n) {
    set_error("tpt_scene_q8_nodes: capacity below the node count");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  if (!n) return TPT_OK;
  if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice") ||
      !hip_ok(hipStreamSynchronize(s->dev->stream), "hipStreamSynchronize") ||
      !hip_ok(hipMemcpy(out, tlas ? s->qtlas : s->qnodes, n * sizeof(GpuNodeQ8), hipMemcpyDeviceToHost), "hipMemcpy"))
    return TPT_ERROR_DEVICE;
  return TPT_OK;
}
size_t tpt_qbvh_node_count(const TptQbvh* q) { return q ? q->q.nodes.size() : 0; }
size_t tpt_qbvh_primitive_count(const TptQbvh* q) { return q ? q->q.prim.size() : 0; }
int32_t tpt_qbvh_export(const TptQbvh* q, TptQbvhNode* nodes, uint32_t* prim) {
  if (!q || (!nodes && !q->q.nodes.empty()) || !prim) {
    set_error("tpt_qbvh_export: null argument");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  if (!q->q.nodes.empty()) std::memcpy(nodes, q->q.nodes.data(), q->q.nodes.size() * sizeof(TptQbvhNode));
  std::memcpy(prim, q->q.prim.data(), q->q.prim.size() * 4);
  return TPT_OK;
}

int32_t tpt_qbvh_quantize(const TptQbvhNode* nodes, size_t n, void* out) {
  if ((!nodes || !out) && n) {
    set_error("tpt_qbvh_quantize: null argument");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  GpuNodeQ8* q = (GpuNodeQ8*)out;
  for (size_t i = 0; i < n; ++i) quantize_node(to_gpu_node(nodes[i], 0), &q[i]);
  return TPT_OK;
}

TptScene* tpt_scene_create(TptDevice* dev, const TptSceneDesc* desc) {
  TptScene* s = nullptr;
  try {
    if (!dev || !desc) {
      set_error("tpt_scene_create: null argument");
      return nullptr;
    }
    if (desc->instance_count && (!desc->instances || !desc->tlas_primitive_indices)) {
      set_error("tpt_scene_create: instances or tlas_primitive_indices missing");
      return nullptr;
    }
    if (!hip_ok(hipSetDevice(dev->index), "hipSetDevice")) return nullptr;
    std::vector<GpuNode> nodes;
    std::vector<GpuTri> tris;
    std::vector<uint32_t> rank;
    std::vector<uint32_t> mesh_node_base(desc->mesh_count), mesh_tri_base(desc->mesh_count);
    int blas_stack = 0;
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
      const TptMeshDesc& md = desc->meshes[m];
      if ((md.triangle_count && (!md.indices || !md.positions || !md.primitive_indices)) ||
          (md.node_count && !md.nodes)) {
        set_error("tpt_scene_create: mesh " + std::to_string(m) + " has null arrays");
        return nullptr;
      }
      if (md.triangle_count > kChildOffsetMask + 1u) {
        set_error("tpt_scene_create: mesh exceeds 2^24 triangles (24-bit leaf offset, R/qbvh.rs:15-18)");
        return nullptr;
      }
      const uint32_t nb = (uint32_t)nodes.size();
      mesh_node_base[m] = nb;
      mesh_tri_base[m] = (uint32_t)tris.size();
      std::vector<TptQbvhNode> mn(md.nodes, md.nodes + md.node_count);
      for (uint32_t k = 0; k < md.node_count; ++k) {
        for (int c = 0; c < 4; ++c) {
          const uint32_t ch = mn[k].children[c];
          if (!(mn[k].valid_mask & (1u << c))) continue;
          if (ch & kChildLeafBit) {
            const uint32_t off = ch & kChildOffsetMask, cnt = (ch >> kChildCountShift) & kChildCountMask;
            if (off + cnt > md.triangle_count) {
              set_error("tpt_scene_create: leaf range outside the mesh's triangles");
              return nullptr;
            }
          } else if (ch >= md.node_count) {
            set_error("tpt_scene_create: interior child index out of range");
            return nullptr;
          }
        }
        nodes.push_back(to_gpu_node(mn[k], nb));
      }
      blas_stack = std::max(blas_stack, stack_bound(mn, 0, 0) + 1);
      const size_t tb = tris.size();
      tris.resize(tb + md.triangle_count);
      rank.resize(tb + md.triangle_count);
      for (uint32_t k = 0; k < md.triangle_count; ++k) {
        const uint32_t ti = md.primitive_indices[k];
        if (ti >= md.triangle_count) {
          set_error("tpt_scene_create: primitive index out of range");
          return nullptr;
        }
        GpuTri& g = tris[tb + k];
        std::memset(&g, 0, sizeof(g));
        for (int v = 0; v < 3; ++v) {
          const uint32_t vi = md.indices[3 * ti + v];
          if (vi >= md.vertex_count) {
            set_error("tpt_scene_create: vertex index out of range");
            return nullptr;
          }
          g.v[3 * v + 0] = md.positions[3 * vi + 0];
          g.v[3 * v + 1] = md.positions[3 * vi + 1];
          g.v[3 * v + 2] = md.positions[3 * vi + 2];
        }
        {  // (v2 - v0) x (v1 - v0) has zero length: the triangle is never hit
          const float ax = g.v[6] - g.v[0], ay = g.v[7] - g.v[1], az = g.v[8] - g.v[2];
          const float bx = g.v[3] - g.v[0], by = g.v[4] - g.v[1], bz = g.v[5] - g.v[2];
          const float cx = ay * bz - by * az, cy = az * bx - bz * ax, cz = ax * by - bx * ay;
          g.prim = ti | ((cx * cx + cy * cy + cz * cz == 0.0f) ? kTriDegenerate : 0u);
        }
        rank[tb + ti] = k;
      }
    }
    std::vector<GpuInstance> inst(desc->instance_count);
    for (uint32_t i = 0; i < desc->instance_count; ++i) {
      const TptInstanceDesc& id = desc->instances[i];
      if (id.mesh >= desc->mesh_count) {
        set_error("tpt_scene_create: instance mesh index out of range");
        return nullptr;
      }
      GpuInstance& g = inst[i];
      const float* w = id.world_to_local;  // column-major: column c = w[4c..4c+2]
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) g.m[3 * c + r] = w[4 * c + r];
      g.node_base = mesh_node_base[id.mesh];
      g.node_count = desc->meshes[id.mesh].node_count;
      g.tri_base = mesh_tri_base[id.mesh];
      g.tri_count = desc->meshes[id.mesh].triangle_count;
    }
    std::vector<GpuNode> tlas;
    std::vector<TptQbvhNode> tn(desc->tlas_nodes, desc->tlas_nodes + desc->tlas_node_count);
    for (uint32_t k = 0; k < desc->tlas_node_count; ++k) tlas.push_back(to_gpu_node(tn[k], 0));
    std::vector<uint32_t> tprim(desc->tlas_primitive_indices, desc->tlas_primitive_indices + desc->instance_count);
    for (uint32_t v : tprim)
      if (v >= desc->instance_count) {
        set_error("tpt_scene_create: TLAS primitive index out of range");
        return nullptr;
      }
    const int tlas_stack = stack_bound(tn, 0, 0) + 1;
    if (tlas_stack + blas_stack > kStackDepth) {
      set_error("tpt_scene_create: tree needs more than 128 traversal stack entries (R/qbvh.rs:609)");
      return nullptr;
    }
    // shading data for the path tracer: vertex arrays in original order,
    // per-instance transforms and materials (R/scene/mod.rs:62-70, 426-446)
    std::vector<float> vpos, vnrm, vuv;
    std::vector<uint32_t> vidx;
    std::vector<uint32_t> mesh_vbase(desc->mesh_count), mesh_ibase(desc->mesh_count), mesh_flags(desc->mesh_count);
    for (uint32_t m = 0; m < desc->mesh_count; ++m) {
      const TptMeshDesc& md = desc->meshes[m];
      mesh_vbase[m] = (uint32_t)(vpos.size() / 3);
      mesh_ibase[m] = (uint32_t)vidx.size();
      mesh_flags[m] = (md.normals ? 1u : 0u) | (md.uvs ? 2u : 0u);
      vpos.insert(vpos.end(), md.positions, md.positions + 3 * (size_t)md.vertex_count);
      if (md.normals)
        vnrm.insert(vnrm.end(), md.normals, md.normals + 3 * (size_t)md.vertex_count);
      else
        vnrm.resize(vnrm.size() + 3 * (size_t)md.vertex_count, 0.0f);
      if (md.uvs)
        vuv.insert(vuv.end(), md.uvs, md.uvs + 2 * (size_t)md.vertex_count);
      else
        vuv.resize(vuv.size() + 2 * (size_t)md.vertex_count, 0.0f);
      vidx.insert(vidx.end(), md.indices, md.indices + 3 * (size_t)md.triangle_count);
    }
    std::vector<GpuShade> shade(desc->instance_count);
    for (uint32_t i = 0; i < desc->instance_count; ++i) {
      const TptInstanceDesc& id = desc->instances[i];
      GpuShade& g = shade[i];
      std::memset(&g, 0, sizeof(g));
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 3; ++r) g.l2w[3 * c + r] = id.local_to_world[4 * c + r];
      // normal_to_world = Mat3::from_mat4(world_to_local.transpose()):
      // its column c is row c of world_to_local's upper 3x3
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) g.nrm[3 * c + r] = id.world_to_local[4 * r + c];
      g.material = id.material;
      g.vbase = mesh_vbase[id.mesh];
      g.ibase = mesh_ibase[id.mesh];
      g.flags = mesh_flags[id.mesh];
    }
    std::vector<TptMaterial> mats(desc->materials ? desc->materials : nullptr,
                                  desc->materials ? desc->materials + desc->material_count : nullptr);
    // SG light tree over the emissive triangles (Scene::build_light_tree)
    std::vector<TptLightNode> lnodes;
    std::vector<uint32_t> light_base, tri_leaf;
    std::string lerr;
    if (!build_light_tree(*desc, &lnodes, &light_base, &tri_leaf, &lerr)) {
      set_error("tpt_scene_create: " + lerr);
      return nullptr;
    }
    for (uint32_t i = 0; i < desc->instance_count; ++i) shade[i].light_base = light_base[i];
    std::vector<GpuLightNode> glights(lnodes.size());
    for (size_t k = 0; k < lnodes.size(); ++k) {
      const TptLightNode& a = lnodes[k];
      GpuLightNode& g = glights[k];
      std::memset(&g, 0, sizeof(g));
      for (int c = 0; c < 3; ++c) {
        g.mu[c] = a.mu[c];
        g.nu[c] = a.nu[c];
      }
      g.flux = a.flux;
      g.lambda = a.lambda;
      g.sigma_s2 = a.sigma_s2;
      g.bound_var = 0.5f * a.radius * a.radius;
      g.sg_norm = sg_norm_host(a.lambda);
      g.area = a.area;
      g.left = a.left;
      g.right = a.right;
      g.parent = a.parent;
      g.instance = a.instance;
      g.triangle = a.triangle;
    }
    s = new TptScene();
    s->dev = dev;
    s->instance_count = desc->instance_count;
    s->tlas_node_count = desc->tlas_node_count;
    s->material_count = desc->materials ? desc->material_count : 0;
    s->light_count = (uint32_t)lnodes.size();
    s->w_tree = lnodes.empty() ? 0.0f : std::fmax(lnodes[0].flux, 0.0f);
    s->light_nodes = std::move(lnodes);
    s->host_mats = mats;
    s->node_count = (uint32_t)nodes.size();
    // the 8-bit nodes are encoded on the device from the uploaded FP32 nodes
    if (!upload(nodes, &s->nodes, &s->bytes) || !encode_q8(dev, s->nodes, nodes.size(), &s->qnodes, &s->bytes) ||
        !upload(tris, &s->tris, &s->bytes) || !upload(rank, &s->tri_rank, &s->bytes) ||
        !upload(inst, &s->inst, &s->bytes) || !upload(tlas, &s->tlas, &s->bytes) ||
        !encode_q8(dev, s->tlas, tlas.size(), &s->qtlas, &s->bytes) || !upload(tprim, &s->tlas_prim, &s->bytes) ||
        !upload(shade, &s->shade, &s->bytes) || !upload(vpos, &s->vpos, &s->bytes) ||
        !upload(vnrm, &s->vnrm, &s->bytes) || !upload(vuv, &s->vuv, &s->bytes) ||
        !upload(vidx, &s->vidx, &s->bytes) || !upload(mats, &s->mats, &s->bytes) ||
        !upload(glights, &s->lights, &s->bytes) || !upload(tri_leaf, &s->tri_leaf, &s->bytes)) {
      free_scene(s);
      return nullptr;
    }
    if (!upload_ggx_tables(s, mats)) {
      free_scene(s);
      return nullptr;
    }
    if (!upload_textures(s, *desc)) {
      free_scene(s);
      return nullptr;
    }
    for (uint32_t i = 0; i < (uint32_t)mats.size(); ++i) {
      const std::string terr = material_texture_error(mats[i], i, s->host_textures);
      if (!terr.empty()) {
        set_error("tpt_scene_create: " + terr);
        free_scene(s);
        return nullptr;
      }
    }
    if (desc->env_pixels) {  // environment light: texels, compensated pyramid, level table
      HostEnv he;
      std::string eerr;
      if (!build_environment(desc->env_pixels, desc->env_width, desc->env_height, desc->env_scale,
                             desc->env_rotate_y, &he, &eerr)) {
        set_error("tpt_scene_create: " + eerr);
        free_scene(s);
        return nullptr;
      }
      if (!upload(he.px, &s->env_px, &s->bytes) || !upload(he.pyr, &s->env_pyr, &s->bytes) ||
          !upload(he.levels, &s->env_levels, &s->bytes)) {
        free_scene(s);
        return nullptr;
      }
      pt::EnvView& e = s->env;
      e.px = s->env_px;
      e.pyr = s->env_pyr;
      e.levels = s->env_levels;
      e.w = he.w;
      e.h = he.h;
      e.level_count = he.level_count;
      e.scale = he.scale;
      e.total_mis = he.total_mis;
      for (int k = 0; k < 9; ++k) {
        e.rot[k] = he.rot[k];
        e.inv_rot[k] = he.inv_rot[k];
      }
      s->w_env = std::fmax(he.power, 0.0f);
    }
    if (!build_light_cats(s, *desc)) {
      free_scene(s);
      return nullptr;
    }
    return s;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_scene_create: ") + e.what());
    free_scene(s);
    return nullptr;
  }
}

void tpt_scene_release(TptScene* s) {
  if (!s) return;
  (void)hipSetDevice(s->dev->index);
  (void)hipStreamSynchronize(s->dev->stream);
  free_scene(s);
}

size_t tpt_scene_device_bytes(const TptScene* s) { return s ? s->bytes : 0; }

int32_t tpt_scene_light_tree(const TptScene* s, TptLightNode* out, uint32_t capacity, uint32_t* count) {
  if (!s || !count || (capacity && !out)) {
    set_error("tpt_scene_light_tree: null argument");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  *count = (uint32_t)s->light_nodes.size();
  const uint32_t n = std::min(capacity, *count);
  if (n) std::memcpy(out, s->light_nodes.data(), n * sizeof(TptLightNode));
  return TPT_OK;
}

}  // extern "C"

namespace tpt {
namespace {

// Kernel variant for (query, node format, exact shadow rule, counters,
// levels, the default LDS stack depth as a constant).
const void* pick_kernel(int query, bool q8, bool exact, bool cnt, bool two, bool fixed_lds) {
#define TPT_K(Q8, M, E, C, T)                                                         \
  (fixed_lds ? reinterpret_cast<const void*>(&trace_kernel<Q8, M, E, C, T, kLdsDepth>) \
             : reinterpret_cast<const void*>(&trace_kernel<Q8, M, E, C, T, 0>))
#define TPT_BY_LEVEL(Q8, M, E, C) (two ? TPT_K(Q8, M, E, C, true) : TPT_K(Q8, M, E, C, false))
#define TPT_BY_COUNT(Q8, M, E) (cnt ? TPT_BY_LEVEL(Q8, M, E, true) : TPT_BY_LEVEL(Q8, M, E, false))
  if (query == 0) return q8 ? TPT_BY_COUNT(true, kClosest, false) : TPT_BY_COUNT(false, kClosest, false);
  if (exact) return q8 ? TPT_BY_COUNT(true, kAny, true) : TPT_BY_COUNT(false, kAny, true);
  return q8 ? TPT_BY_COUNT(true, kAny, false) : TPT_BY_COUNT(false, kAny, false);
#undef TPT_BY_COUNT
#undef TPT_BY_LEVEL
#undef TPT_K
}

// Persistent grid: enough resident blocks to fill every CU (or the tuned
// blocks per CU), never more blocks than the rays need.  lds: the launch's
// dynamic LDS bytes (the traversal stack).
unsigned grid_blocks(TptDevice* d, const void* kern, size_t n, size_t lds) {
  static std::mutex mu;
  static std::vector<std::pair<std::pair<const void*, size_t>, int>> cache;
  int per_cu = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    for (auto& e : cache)
      if (e.first.first == kern && e.first.second == lds) per_cu = e.second;
    if (per_cu == 0) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
      cache.emplace_back(std::make_pair(kern, lds), per_cu);
    }
  }
  if (d->blocks_per_cu && (int)d->blocks_per_cu < per_cu) per_cu = (int)d->blocks_per_cu;
  const size_t want = (n + kBlock - 1) / kBlock;
  const size_t cap = (size_t)per_cu * (size_t)(d->cu_count > 0 ? d->cu_count : 1);
  return (unsigned)(want < cap ? want : cap);
}

// One implementation behind tpt_trace_closest / tpt_trace_any /
// tpt_trace_counts.  query 0: closest (hits), 1: shadow (unoccluded).
int32_t trace_impl(TptScene* s, int query, const TptRay* rays, size_t n, const uint32_t* target,
                   const float* distance, TptHit* hits, uint8_t* unocc, uint32_t* counts, uint32_t flags,
                   const char* fn) {
  if (!s) {
    set_error(std::string(fn) + ": null scene");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  if (n == 0) return TPT_OK;
  if (!rays || (query == 0 && !hits && !counts) || (query == 1 && !unocc && !counts)) {
    set_error(std::string(fn) + ": null rays or output");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  if (n > 0x7FFFFFFFull) {
    set_error(std::string(fn) + ": too many rays for one call");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  TptDevice* d = s->dev;
  if (!hip_ok(hipSetDevice(d->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
  const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
  const bool q8 = flags & TPT_FLAG_QUANTIZED;
  const bool exact = flags & TPT_FLAG_ANY_EXACT;
  const TptRay* drays = rays;
  const uint32_t* dtarget = target;
  const float* ddist = distance;
  TptHit* dhits = hits;
  uint8_t* dout = unocc;
  uint32_t* dcounts = counts;
  if (!dev_ptr) {
    // one staging block: rays | target | distance | output | counts
    const size_t nb_r = n * sizeof(TptRay);
    const size_t nb_t = (query == 1 && target) ? n * 8 : 0;
    const size_t nb_d = (query == 1 && distance) ? n * 4 : 0;
    const size_t nb_o = ((query == 0 ? n * sizeof(TptHit) : n) + 15) & ~size_t(15);
    const size_t nb_c = counts ? n * 12 : 0;
    if (!ensure_stage(s, nb_r + nb_t + nb_d + nb_o + nb_c)) return TPT_ERROR_DEVICE;
    char* p = (char*)s->stage;
    drays = (const TptRay*)p;
    dtarget = nb_t ? (const uint32_t*)(p + nb_r) : nullptr;
    ddist = nb_d ? (const float*)(p + nb_r + nb_t) : nullptr;
    char* po = p + nb_r + nb_t + nb_d;
    dhits = query == 0 ? (TptHit*)po : nullptr;
    dout = query == 1 ? (uint8_t*)po : nullptr;
    dcounts = counts ? (uint32_t*)(po + nb_o) : nullptr;
    if (!hip_ok(hipMemcpyAsync((void*)drays, rays, nb_r, hipMemcpyHostToDevice, d->stream), "hipMemcpyAsync(rays)"))
      return TPT_ERROR_DEVICE;
    if (nb_t && !hip_ok(hipMemcpyAsync((void*)dtarget, target, nb_t, hipMemcpyHostToDevice, d->stream),
                        "hipMemcpyAsync(target)"))
      return TPT_ERROR_DEVICE;
    if (nb_d && !hip_ok(hipMemcpyAsync((void*)ddist, distance, nb_d, hipMemcpyHostToDevice, d->stream),
                        "hipMemcpyAsync(distance)"))
      return TPT_ERROR_DEVICE;
    if (!hip_ok(hipMemsetAsync(d->status, 0, 4, d->stream), "hipMemsetAsync(status)")) return TPT_ERROR_DEVICE;
  }
  // Outputs the caller did not ask for still need a destination on device.
  if (query == 0 && !dhits) {
    if (!ensure_stage(s, n * sizeof(TptHit))) return TPT_ERROR_DEVICE;
    dhits = (TptHit*)s->stage;
  }
  if (query == 1 && !dout) {
    if (!ensure_stage(s, n)) return TPT_ERROR_DEVICE;
    dout = (uint8_t*)s->stage;
  }
  const SceneView v = view_of(s);
  Query q;
  q.hits = dhits;
  q.target = dtarget;
  q.distance = ddist;
  q.out = dout;
  q.counts = dcounts;
  q.status = d->status;
  q.work = d->work;
  q.n_dev = nullptr;
  q.n = (uint32_t)n;
  q.refill_below = d->refill_below;
  q.grab = d->refill_grab;
  q.regions = d->regions;
  q.lds_depth = d->lds_depth;
  const bool fixed_lds = q.lds_depth == (uint32_t)kLdsDepth;  // static LDS, no launch-sized LDS
  const size_t lds = fixed_lds ? 0 : (size_t)q.lds_depth * kBlock * 4;
  const bool two = !(s->instance_count == 1 && s->tlas_node_count == 0);
  const bool cnt = dcounts != nullptr;
  const void* kern = pick_kernel(query, q8, exact, cnt, two, fixed_lds);
  const unsigned blocks = grid_blocks(d, kern, n, lds);
  void* args[] = {(void*)&v, (void*)&drays, (void*)&q};
  if (!hip_ok(hipLaunchKernel(kern, dim3(blocks), dim3(kBlock), args, lds, d->stream), "hipLaunchKernel"))
    return TPT_ERROR_DEVICE;
  hipStream_t st = d->stream;
  if (!hip_ok(hipGetLastError(), "kernel launch")) return TPT_ERROR_DEVICE;
  if (!dev_ptr) {
    if (query == 0 && hits &&
        !hip_ok(hipMemcpyAsync(hits, dhits, n * sizeof(TptHit), hipMemcpyDeviceToHost, st), "hipMemcpyAsync(hits)"))
      return TPT_ERROR_DEVICE;
    if (query == 1 && unocc && !hip_ok(hipMemcpyAsync(unocc, dout, n, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(out)"))
      return TPT_ERROR_DEVICE;
    if (counts && !hip_ok(hipMemcpyAsync(counts, dcounts, n * 12, hipMemcpyDeviceToHost, st), "hipMemcpyAsync(counts)"))
      return TPT_ERROR_DEVICE;
    uint32_t status = 0;
    if (!read_status(d, &status)) return TPT_ERROR_DEVICE;
    if (status & 1u) {
      set_error(std::string(fn) + ": traversal stack overflow (128 entries)");
      return TPT_ERROR_OVERFLOW;
    }
  }
  return TPT_OK;
}

}  // namespace

// Wavefront queues: a persistent trace kernel whose ray count is read on the
// device, so a whole bounce loop can be enqueued (or captured in a graph)
// without a host round trip.  The grid is sized for n_max rays.
bool launch_trace_queue(TptScene* s, int query, bool q8, bool exact, const TptRay* rays, const uint32_t* n_dev,
                        uint32_t n_max, TptHit* hits, const uint32_t* target, uint8_t* out, hipStream_t st) {
  TptDevice* d = s->dev;
  const SceneView v = view_of(s);
  Query q;
  q.hits = hits;
  q.target = target;
  q.distance = nullptr;
  q.out = out;
  q.counts = nullptr;
  q.status = d->status;
  q.work = d->work;
  q.n_dev = n_dev;
  q.n = n_max;
  q.refill_below = d->refill_below;
  q.grab = d->refill_grab;
  q.regions = d->regions;
  q.lds_depth = d->lds_depth;
  const bool fixed_lds = q.lds_depth == (uint32_t)kLdsDepth;  // static LDS, no launch-sized LDS
  const size_t lds = fixed_lds ? 0 : (size_t)q.lds_depth * kBlock * 4;
  const bool two = !(s->instance_count == 1 && s->tlas_node_count == 0);
  const void* kern = pick_kernel(query, q8, exact, false, two, fixed_lds);
  const unsigned blocks = grid_blocks(d, kern, n_max, lds);
  void* args[] = {(void*)&v, (void*)&rays, (void*)&q};
  return hip_ok(hipLaunchKernel(kern, dim3(blocks), dim3(kBlock), args, lds, st), "hipLaunchKernel(trace queue)");
}

}  // namespace tpt

extern "C" {

int32_t tpt_trace_closest(TptScene* s, const TptRay* rays, size_t n, TptHit* hits, uint32_t flags) {
  try {
    if (!hits && n) {
      set_error("tpt_trace_closest: null hits");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    return trace_impl(s, 0, rays, n, nullptr, nullptr, hits, nullptr, nullptr, flags, "tpt_trace_closest");
  } catch (const std::exception& e) {
    set_error(e.what());
    return TPT_ERROR;
  }
}

int32_t tpt_trace_any(TptScene* s, const TptRay* rays, size_t n, const uint32_t* target, const float* distance,
                      uint8_t* out, uint32_t flags) {
  try {
    if (!out && n) {
      set_error("tpt_trace_any: null output");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    return trace_impl(s, 1, rays, n, target, distance, nullptr, out, nullptr, flags, "tpt_trace_any");
  } catch (const std::exception& e) {
    set_error(e.what());
    return TPT_ERROR;
  }
}

int32_t tpt_trace_counts(TptScene* s, uint32_t query, const TptRay* rays, size_t n, const uint32_t* target,
                         const float* distance, uint32_t* counts, uint32_t flags) {
  try {
    if (query > 1 || (!counts && n)) {
      set_error("tpt_trace_counts: query must be 0 or 1 and counts non-null");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    return trace_impl(s, (int)query, rays, n, target, distance, nullptr, nullptr, counts, flags, "tpt_trace_counts");
  } catch (const std::exception& e) {
    set_error(e.what());
    return TPT_ERROR;
  }
}

}  // extern "C"
