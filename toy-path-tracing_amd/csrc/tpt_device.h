// Device handle and host-side helpers shared by the libtpt compilation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include <vector>

#include "light_tree.h"
#include "tpt_internal.h"
#include "traverse.h"

namespace tpt {

// Thread-local last error (shim.h convention of the reference's C shim).
void set_error(const std::string& s);
// true on hipSuccess; otherwise records "<what>: <hip error>" and returns false.
bool hip_ok(hipError_t e, const char* what);

// Ray counters of the persistent trace kernels: one per XCD range plus the
// count of finished waves; the last wave resets them.
constexpr int kMaxRegions = 64;  // work regions of the persistent trace kernel
// One counter per 256-byte line: every wave of every XCD updates these with
// device-scope atomics, and counters sharing a line serialise on it.
struct alignas(256) WorkLine {
  uint32_t v;
  uint32_t pad[63];
};
struct WorkCounters {
  WorkLine next[kMaxRegions];
  WorkLine done_waves;
};

// A device allocation owned by the current host scope.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool ensure(size_t nb) {
    if (bytes >= nb) return true;
    release();
    if (!hip_ok(hipMalloc(&ptr, nb), "hipMalloc")) {
      ptr = nullptr;
      return false;
    }
    bytes = nb;
    return true;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  ~DeviceBuffer() { release(); }
};

}  // namespace tpt

namespace tpt {
namespace pt {
// Environment light on the device: RGB texels, the MIS-compensated weight
// pyramid (levels coarsest first, table of offset / width / height), radiance
// scale and the rotation about +Y (R/light/environment.rs:9-20).
struct EnvView {
  const float* px;  // nullptr: no environment light
  const float* pyr;
  const uint32_t* levels;
  uint32_t w, h, level_count;
  float scale, total_mis;
  float rot[9], inv_rot[9];
};
// Textures on the device: every mip level of every texture in one texel
// array (RGB texels padded to float4 so one texel is one 16-byte load; scalar
// texels one float), a level table (offset in floats, size) and per texture
// its first level, level count, channels and level-0 maximum
// (R/material/texture.rs:53-67).
struct GpuTexLevel {
  uint32_t offset, width, height, pow2;  // pow2: bit 0 width, bit 1 height is a power of two
};
struct GpuTexture {
  uint32_t first_level, levels, channels;
  float max_value;
};
struct TexView {
  const float* texels;
  const GpuTexLevel* levels;
  const GpuTexture* tex;
  uint32_t count;
};
// The light sampler's categories (LightSampler::build_from_scene,
// R/light/mod.rs:133-191): the light tree (kind 0, weight root flux), the
// environment (kind 1, total power) and each directional light i (kind 2 + i,
// luminance(color) * intensity), in that order and only those with a positive
// weight; the normalised CDF and the total weight, as the reference forms
// them; and the directional lights (direction, color * intensity).
constexpr uint32_t kMaxDirectional = 16;
constexpr uint32_t kMaxLightCats = 2 + kMaxDirectional;
struct LightCats {
  uint32_t count;
  float total;
  uint32_t kind[kMaxLightCats];
  float weight[kMaxLightCats];
  float cdf[kMaxLightCats + 1];
  float dir[kMaxDirectional][3];
  float radiance[kMaxDirectional][3];
};
}  // namespace pt
}  // namespace tpt

struct TptDevice {
  int index = 0;
  int cu_count = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  uint32_t* status = nullptr;         // device word: bit 0 = traversal stack overflow
  tpt::WorkCounters* work = nullptr;  // persistent-kernel ray counters (self-resetting)
  // Defaults from a sweep on MI355X over the 2M-triangle traversal workload
  // (refill 16..62 x blocks/CU 2..8, conservative any-hit slab test): the
  // kernels are bound by VALU issue with divergent lanes, so every resident
  // wave the occupancy allows helps, and refilling a wave once fewer than 40
  // of its 64 lanes are busy keeps more lanes working (shadow rays 1.107 ->
  // 0.948 ms against 4 blocks per CU and a threshold of 24).
  uint32_t refill_below = 40;         // refill a wave when fewer lanes are busy
  uint32_t refill_grab = 0;           // ray ids a wave takes per work-counter atomic (at least its idle lanes)
  uint32_t regions = 8;               // work regions (multiple of the 8 XCDs keeps block b on region b's XCD)
  uint32_t blocks_per_cu = 0;         // resident blocks per CU (0: occupancy limit)
  uint32_t lds_depth = 16;            // traversal stack entries per lane kept in LDS (the rest in scratch)
  // wavefront path tracer: paths in flight per pass, graph capture of a pass,
  // the shade queue's tile (threads per generate / shade block, one queue
  // entry each), the material sort's granularity (0: shade in queue order;
  // g: the hits sorted by material index / g, misses last), and its buffers
  // (grow-only)
  uint32_t wf_paths = 1u << 21;
  uint32_t wf_graph = 1;
  uint32_t wf_tile = 256;
  uint32_t wf_sort = 0;
  tpt::DeviceBuffer wf_buf;
  // tpt_qbvh_build_device's workspace (grow-only: about 420 bytes per primitive)
  tpt::DeviceBuffer qbvh_buf;
  // tpt_trace_noalloc: this device's address of the file-scope scratch array
  // (looked up once by noalloc_setup) and the resident blocks per CU of its
  // kernel variants
  void* noalloc_scratch = nullptr;
  int noalloc_per_cu[12] = {};
};

namespace tpt {
// The QBVH build on the device (qbvh_build_gpu.hip): the host builder's
// output from bounds in host or device memory.  Returns 0, -1 (too many
// primitives), -2 (child encoding overflow), -3 (a forced split it does not
// reorder) or -4 (device error, set_error holds it).
int build_qbvh_device(TptDevice* d, const float* bmin, const float* bmax, size_t n, bool device_pointers,
                      HostQbvh* out);
// Looks up the allocation-free entry point's scratch on the current device
// (tpt_device_create calls it once per device handle).
bool noalloc_setup(TptDevice* d);
}  // namespace tpt

// A scene resident in device memory (tpt_scene_create).
struct TptScene {
  TptDevice* dev = nullptr;
  // traversal
  tpt::GpuNode* nodes = nullptr;
  tpt::GpuNodeQ8* qnodes = nullptr;
  tpt::GpuTri* tris = nullptr;
  uint32_t* tri_rank = nullptr;
  tpt::GpuInstance* inst = nullptr;
  tpt::GpuNode* tlas = nullptr;
  tpt::GpuNodeQ8* qtlas = nullptr;
  uint32_t* tlas_prim = nullptr;
  uint32_t instance_count = 0;
  uint32_t tlas_node_count = 0;
  uint32_t node_count = 0;  // BLAS nodes of every mesh (nodes / qnodes entries)
  // shading (path tracer): per-instance records, mesh vertex arrays in
  // original order, indices, materials
  tpt::GpuShade* shade = nullptr;
  float* vpos = nullptr;
  float* vnrm = nullptr;
  float* vuv = nullptr;
  uint32_t* vidx = nullptr;
  TptMaterial* mats = nullptr;
  uint32_t material_count = 0;
  // SG light tree over the emissive triangles and the triangle -> leaf table
  tpt::GpuLightNode* lights = nullptr;
  uint32_t light_count = 0;
  uint32_t* tri_leaf = nullptr;
  std::vector<TptLightNode> light_nodes;  // host copy (tpt_scene_light_tree)
  std::vector<TptMaterial> host_mats;     // host copy of the material table
  // light-sampler category weights and the environment light
  float w_tree = 0.0f, w_env = 0.0f;
  uint32_t nee_policy = TPT_NEE_REFERENCE;  // tpt_scene_set_nee_policy
  tpt::pt::LightCats* light_cats = nullptr;  // device copy of host_cats
  tpt::pt::LightCats host_cats{};
  float* env_px = nullptr;
  float* env_pyr = nullptr;
  uint32_t* env_levels = nullptr;
  tpt::pt::EnvView env{};
  // GGX tables of the OpenPBR materials (ggx_tables.h): directional-albedo
  // tables (table 0 all zero), per material its spec and coat table, the
  // conductor and dielectric energy-compensation tables (zero when unused)
  float* ggx_albedo = nullptr;
  uint32_t* ggx_mat = nullptr;
  float* ggx_cond = nullptr;
  bool cond_tables = false;  // ggx_cond holds the conductor tables (not zeros)
  bool force_cond = false;   // build them whatever the materials (the parameter adjoint asked)
  float* ggx_diel = nullptr;
  // textures (TptSceneDesc::textures): texels of every level, level table,
  // per-texture records; host copies of the two tables
  float* tex_texels = nullptr;
  tpt::pt::GpuTexLevel* tex_levels = nullptr;
  tpt::pt::GpuTexture* textures = nullptr;
  uint32_t texture_count = 0;
  std::vector<tpt::pt::GpuTexLevel> host_tex_levels;
  std::vector<tpt::pt::GpuTexture> host_textures;
  size_t bytes = 0;
  // staging for host-pointer calls (grow-only)
  void* stage = nullptr;
  size_t stage_bytes = 0;
};

namespace tpt {
bool launch_trace_queue(TptScene* s, int query, bool q8, bool exact, const TptRay* rays, const uint32_t* n_dev,
                        uint32_t n_max, TptHit* hits, const uint32_t* target, uint8_t* out, hipStream_t st);
// true if every instance's material is one the path tracers implement;
// otherwise records "<fn>: ..." and returns false.
// Builds and uploads the GGX tables of the scene's materials (trace.hip).
bool upload_ggx_tables(TptScene* s, const std::vector<TptMaterial>& mats);
// Copies the scene description's textures to the device and builds their
// mip pyramids there (texture.hip); records "<fn>: ..." on failure.
bool upload_textures(TptScene* s, const TptSceneDesc& desc);
// Level-0 maximum of a texture description (Texture::max_value): the fold
// of f32::max over every channel, from 0.
float texture_max_value(const TptTextureDesc& t);
// Floats of the scene's texel array (every level; RGB texels 4 floats).
size_t texel_floats(const TptScene* s);
// Floats of the level-0 texels of every texture packed (channels per texel).
size_t texel_grad_floats(const TptScene* s);
// Turns a gradient array laid out like the texels (every level) into
// level-0 gradients (the adjoint of the pyramid build, coarsest level first;
// modifies g) and packs them into out, texture after texture; enqueued on st.
template <typename T>
bool texel_grads_to_level0(TptScene* s, T* g, float* out, hipStream_t st);
// Checks the materials' texture indices against the scene's texture table
// (colour slots need RGB, scalar slots one channel); "" when valid.
std::string material_texture_error(const TptMaterial& m, uint32_t index, const std::vector<pt::GpuTexture>& tex);
// film (optional): set to whether some material has a thin film
bool render_materials_ok(TptScene* s, const char* fn, bool* film = nullptr);

inline SceneView view_of(const TptScene* s) {
  SceneView v;
  v.nodes = s->nodes;
  v.qnodes = s->qnodes;
  v.tris = s->tris;
  v.tri_rank = s->tri_rank;
  v.inst = s->inst;
  v.tlas = s->tlas;
  v.qtlas = s->qtlas;
  v.tlas_prim = s->tlas_prim;
  v.instance_count = s->instance_count;
  v.tlas_node_count = s->tlas_node_count;
  return v;
}
}  // namespace tpt
