// Textures: upload of the scene description's images, the box-filtered mip
// pyramid built on the device (build_mip_levels, R/material/texture.rs:433-466),
// the material texture-slot checks and the lookup entry points
// (tpt_texture_sample / tpt_texture_info / tpt_texture_level).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tpt.h"
#include "texture.h"
#include "tpt_device.h"

namespace tpt {
namespace {

using pt::GpuTexLevel;
using pt::GpuTexture;
using pt::TexView;

// One level of the pyramid from the one below: texel (x, y) is the mean of
// the wrapped 2x2 block at (2x, 2y), summed in the reference's order
// ((p00 + p10) + p01) + p11, times 0.25.  RGB texels are float4 (w stays 0).
template <int C>
__global__ __launch_bounds__(256) void mip_kernel(float* __restrict__ texels, GpuTexLevel src, GpuTexLevel dst) {
  const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), y = blockIdx.y * 16u + (threadIdx.x >> 4);
  if (x >= dst.width || y >= dst.height) return;
  const uint32_t sx0 = tex::wrap_index((int)(2u * x), src.width, src.pow2 & 1u);
  const uint32_t sx1 = tex::wrap_index((int)(2u * x + 1u), src.width, src.pow2 & 1u);
  const uint32_t sy0 = tex::wrap_index((int)(2u * y), src.height, (src.pow2 >> 1) & 1u);
  const uint32_t sy1 = tex::wrap_index((int)(2u * y + 1u), src.height, (src.pow2 >> 1) & 1u);
  const size_t o = (size_t)y * dst.width + x;
  if (C == 3) {
    const float4* s = reinterpret_cast<const float4*>(texels + src.offset);
    const float4 a = s[(size_t)sy0 * src.width + sx0], b = s[(size_t)sy0 * src.width + sx1];
    const float4 c = s[(size_t)sy1 * src.width + sx0], d = s[(size_t)sy1 * src.width + sx1];
    float4 r;
    r.x = (((a.x + b.x) + c.x) + d.x) * 0.25f;
    r.y = (((a.y + b.y) + c.y) + d.y) * 0.25f;
    r.z = (((a.z + b.z) + c.z) + d.z) * 0.25f;
    r.w = 0.0f;
    reinterpret_cast<float4*>(texels + dst.offset)[o] = r;
  } else {
    const float* s = texels + src.offset;
    const float a = s[(size_t)sy0 * src.width + sx0], b = s[(size_t)sy0 * src.width + sx1];
    const float c = s[(size_t)sy1 * src.width + sx0], d = s[(size_t)sy1 * src.width + sx1];
    texels[dst.offset + o] = (((a + b) + c) + d) * 0.25f;
  }
}

template <int C>
__global__ __launch_bounds__(256) void texture_query_kernel(TexView tv, uint32_t ti, uint32_t mode,
                                                            const TptTextureQuery* __restrict__ q, size_t n,
                                                            float* __restrict__ out) {
  const GpuTexture t = tv.tex[ti];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const TptTextureQuery a = q[i];
    const tex::V2 uv = tex::v2(a.uv[0], a.uv[1]), dx = tex::v2(a.dstdx[0], a.dstdx[1]),
                  dy = tex::v2(a.dstdy[0], a.dstdy[1]);
    F3 r;
    switch (mode) {
      case TPT_TEXTURE_BILINEAR: r = tex::bilerp<C>(tv, t, 0, uv); break;
      case TPT_TEXTURE_NEAREST: r = tex::nearest<C>(tv, t, uv); break;
      case TPT_TEXTURE_TRILINEAR: r = tex::trilinear<C>(tv, t, uv, dx, dy); break;
      case TPT_TEXTURE_MIP_BILINEAR: r = tex::mip_bilinear<C>(tv, t, uv, dx, dy); break;
      default: r = tex::filtered<C>(tv, t, uv, dx, dy); break;
    }
    if (C == 3) {
      out[3 * i] = r.x;
      out[3 * i + 1] = r.y;
      out[3 * i + 2] = r.z;
    } else {
      out[i] = r.x;
    }
  }
}

// Adjoint of one pyramid level: each level-L gradient flows to its four
// wrapped level-(L-1) texels with weight 0.25 (children shared by two parents
// under wrap addressing get both).
template <int C, typename T>
__global__ __launch_bounds__(256) void mip_adjoint_kernel(T* __restrict__ g, GpuTexLevel src, GpuTexLevel dst) {
  const uint32_t x = blockIdx.x * 16u + (threadIdx.x & 15u), y = blockIdx.y * 16u + (threadIdx.x >> 4);
  if (x >= dst.width || y >= dst.height) return;
  const uint32_t xs[2] = {tex::wrap_index((int)(2u * x), src.width, src.pow2 & 1u),
                          tex::wrap_index((int)(2u * x + 1u), src.width, src.pow2 & 1u)};
  const uint32_t ys[2] = {tex::wrap_index((int)(2u * y), src.height, (src.pow2 >> 1) & 1u),
                          tex::wrap_index((int)(2u * y + 1u), src.height, (src.pow2 >> 1) & 1u)};
  const size_t o = (size_t)y * dst.width + x;
  for (int c = 0; c < C; ++c) {
    const T v = (T)0.25 * g[dst.offset + (C == 3 ? 4 * o + c : o)];
    if (v == (T)0) continue;
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        const size_t k = (size_t)ys[j] * src.width + xs[i];
        atomicAdd(&g[src.offset + (C == 3 ? 4 * k + c : k)], v);
      }
  }
}
// Level 0 of a gradient array (texel layout) packed to channels floats per texel.
template <typename T>
__global__ __launch_bounds__(256) void pack_level_kernel(const T* __restrict__ g, GpuTexLevel L, uint32_t channels,
                                                         float* __restrict__ out) {
  const size_t n = (size_t)L.width * L.height;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    if (channels == 3) {
      out[3 * i] = (float)g[L.offset + 4 * i];
      out[3 * i + 1] = (float)g[L.offset + 4 * i + 1];
      out[3 * i + 2] = (float)g[L.offset + 4 * i + 2];
    } else {
      out[i] = (float)g[L.offset + i];
    }
  }
}

bool is_pow2(uint32_t v) { return v && !(v & (v - 1u)); }

// Levels 1.. of one texture from its level 0 (enqueued on st).
bool build_pyramid(float* texels, const GpuTexLevel* levels, const GpuTexture& t, hipStream_t st) {
  for (uint32_t l = 1; l < t.levels; ++l) {
    const GpuTexLevel &src = levels[l - 1], &dst = levels[l];
    const dim3 grid((dst.width + 15u) / 16u, (dst.height + 15u) / 16u);
    if (t.channels == 3)
      hipLaunchKernelGGL(mip_kernel<3>, grid, dim3(256), 0, st, texels, src, dst);
    else
      hipLaunchKernelGGL(mip_kernel<1>, grid, dim3(256), 0, st, texels, src, dst);
  }
  return hip_ok(hipGetLastError(), "mip_kernel launch");
}

// RGB texels (3 floats) into the float4 layout of a level.
__global__ __launch_bounds__(256) void unpack_rgb_kernel(const float* __restrict__ src, size_t n,
                                                         float* __restrict__ dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float4 v;
    v.x = src[3 * i];
    v.y = src[3 * i + 1];
    v.z = src[3 * i + 2];
    v.w = 0.0f;
    reinterpret_cast<float4*>(dst)[i] = v;
  }
}

TexView tex_view_of(const TptScene* s) {
  TexView v;
  v.texels = s->tex_texels;
  v.levels = s->tex_levels;
  v.tex = s->textures;
  v.count = s->texture_count;
  return v;
}

}  // namespace

float texture_max_value(const TptTextureDesc& t) {
  float m = 0.0f;
  const size_t n = (size_t)t.width * t.height;
  if (!t.pixels) return m;
  for (size_t i = 0; i < n; ++i) {
    float v;
    if (t.channels == 3) {  // Vec3::max_element
      v = std::fmax(t.pixels[3 * i], std::fmax(t.pixels[3 * i + 1], t.pixels[3 * i + 2]));
    } else {
      v = t.pixels[i];
    }
    m = std::fmax(m, v);  // f32::max ignores NaN
  }
  return m;
}

std::string material_texture_error(const TptMaterial& m, uint32_t index, const std::vector<GpuTexture>& tex) {
  if (m.kind != TPT_MATERIAL_OPEN_PBR) return "";
  const TptOpenPbr& p = m.u.open_pbr;
  const struct {
    uint32_t slot;
    uint32_t channels;
    const char* name;
  } slots[6] = {{p.base_color_texture, 3, "base_color_texture"},
                {p.base_metalness_texture, 1, "base_metalness_texture"},
                {p.specular_roughness_texture, 1, "specular_roughness_texture"},
                {p.fuzz_weight_texture, 1, "fuzz_weight_texture"},
                {p.fuzz_roughness_texture, 1, "fuzz_roughness_texture"},
                {p.emission_color_texture, 3, "emission_color_texture"}};
  for (const auto& sl : slots) {
    if (sl.slot == TPT_INVALID_INDEX) continue;
    const std::string where = "material " + std::to_string(index) + " " + sl.name + " " + std::to_string(sl.slot);
    if (sl.slot >= tex.size()) return where + " is not a texture of the scene";
    if (tex[sl.slot].channels != sl.channels)
      return where + (sl.channels == 3 ? " must be an RGB texture" : " must be a scalar texture");
  }
  return "";
}

bool upload_textures(TptScene* s, const TptSceneDesc& d) {
  s->texture_count = 0;
  if (!d.texture_count) return true;
  if (!d.textures) {
    set_error("tpt_scene_create: texture_count > 0 with a null texture table");
    return false;
  }
  std::vector<GpuTexLevel> levels;
  std::vector<GpuTexture> texs(d.texture_count);
  uint64_t total = 0;
  for (uint32_t i = 0; i < d.texture_count; ++i) {
    const TptTextureDesc& t = d.textures[i];
    if (!t.pixels || !t.width || !t.height || (t.channels != 1 && t.channels != 3)) {
      set_error("tpt_scene_create: texture " + std::to_string(i) +
                " needs pixels, a non-zero size and 1 or 3 channels");
      return false;
    }
    GpuTexture& g = texs[i];
    g.first_level = (uint32_t)levels.size();
    g.channels = t.channels;
    g.max_value = texture_max_value(t);
    const uint32_t stride = t.channels == 3 ? 4u : 1u;
    uint32_t w = t.width, h = t.height;
    for (;;) {  // div_ceil halving down to 1 x 1 (build_mip_levels)
      GpuTexLevel L;
      L.offset = (uint32_t)total;
      L.width = w;
      L.height = h;
      L.pow2 = (is_pow2(w) ? 1u : 0u) | (is_pow2(h) ? 2u : 0u);
      levels.push_back(L);
      total += ((uint64_t)w * h * stride + 3u) & ~(uint64_t)3u;
      if (total >= 0xFFFFFFF0ull) {
        set_error("tpt_scene_create: textures exceed 2^32 floats");
        return false;
      }
      if (w <= 1 && h <= 1) break;
      w = std::max((w + 1u) / 2u, 1u);
      h = std::max((h + 1u) / 2u, 1u);
    }
    g.levels = (uint32_t)levels.size() - g.first_level;
  }
  if (!hip_ok(hipMalloc((void**)&s->tex_texels, std::max<uint64_t>(total, 4) * sizeof(float)), "hipMalloc(texels)"))
    return false;
  s->bytes += total * sizeof(float);
  const hipStream_t st = s->dev->stream;
  std::vector<float> rgba;
  for (uint32_t i = 0; i < d.texture_count; ++i) {
    const TptTextureDesc& t = d.textures[i];
    const GpuTexLevel& L0 = levels[texs[i].first_level];
    const size_t n = (size_t)t.width * t.height;
    if (t.channels == 3) {
      rgba.assign(4 * n, 0.0f);
      for (size_t k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c) rgba[4 * k + c] = t.pixels[3 * k + c];
      if (!hip_ok(hipMemcpy(s->tex_texels + L0.offset, rgba.data(), 4 * n * sizeof(float), hipMemcpyHostToDevice),
                  "hipMemcpy(texels)"))
        return false;
    } else if (!hip_ok(hipMemcpy(s->tex_texels + L0.offset, t.pixels, n * sizeof(float), hipMemcpyHostToDevice),
                       "hipMemcpy(texels)")) {
      return false;
    }
    if (!build_pyramid(s->tex_texels, levels.data() + texs[i].first_level, texs[i], st)) return false;
  }
  if (!hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize(mip)")) return false;
  const size_t nl = levels.size() * sizeof(GpuTexLevel), nt = texs.size() * sizeof(GpuTexture);
  if (!hip_ok(hipMalloc((void**)&s->tex_levels, nl), "hipMalloc(levels)") ||
      !hip_ok(hipMalloc((void**)&s->textures, nt), "hipMalloc(textures)") ||
      !hip_ok(hipMemcpy(s->tex_levels, levels.data(), nl, hipMemcpyHostToDevice), "hipMemcpy(levels)") ||
      !hip_ok(hipMemcpy(s->textures, texs.data(), nt, hipMemcpyHostToDevice), "hipMemcpy(textures)"))
    return false;
  s->bytes += nl + nt;
  s->texture_count = d.texture_count;
  s->host_tex_levels = std::move(levels);
  s->host_textures = std::move(texs);
  return true;
}

size_t texel_floats(const TptScene* s) {
  if (!s->texture_count) return 0;
  const GpuTexLevel& last = s->host_tex_levels.back();
  const GpuTexture& t = s->host_textures.back();
  return last.offset + (size_t)last.width * last.height * (t.channels == 3 ? 4u : 1u);
}

size_t texel_grad_floats(const TptScene* s) {
  size_t n = 0;
  for (uint32_t i = 0; i < s->texture_count; ++i) {
    const GpuTexLevel& L = s->host_tex_levels[s->host_textures[i].first_level];
    n += (size_t)L.width * L.height * s->host_textures[i].channels;
  }
  return n;
}

template <typename T>
bool texel_grads_to_level0(TptScene* s, T* g, float* out, hipStream_t st) {
  size_t o = 0;
  for (uint32_t i = 0; i < s->texture_count; ++i) {
    const GpuTexture& t = s->host_textures[i];
    for (uint32_t l = t.levels - 1; l >= 1; --l) {
      const GpuTexLevel &dst = s->host_tex_levels[t.first_level + l], &src = s->host_tex_levels[t.first_level + l - 1];
      const dim3 grid((dst.width + 15u) / 16u, (dst.height + 15u) / 16u);
      if (t.channels == 3)
        hipLaunchKernelGGL((mip_adjoint_kernel<3, T>), grid, dim3(256), 0, st, g, src, dst);
      else
        hipLaunchKernelGGL((mip_adjoint_kernel<1, T>), grid, dim3(256), 0, st, g, src, dst);
    }
    const GpuTexLevel& L0 = s->host_tex_levels[t.first_level];
    const size_t n = (size_t)L0.width * L0.height;
    hipLaunchKernelGGL(pack_level_kernel<T>, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, st,
                       (const T*)g, L0, t.channels, out + o);
    o += n * t.channels;
  }
  return hip_ok(hipGetLastError(), "texel gradient launch");
}
template bool texel_grads_to_level0<float>(TptScene*, float*, float*, hipStream_t);
template bool texel_grads_to_level0<double>(TptScene*, double*, float*, hipStream_t);

}  // namespace tpt

using namespace tpt;

extern "C" {

int32_t tpt_texture_info(const TptScene* s, uint32_t texture, uint32_t* levels, uint32_t* channels,
                         float* max_value) {
  if (!s) {
    set_error("tpt_texture_info: null scene");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  if (texture >= s->texture_count) {
    set_error("tpt_texture_info: texture index out of range");
    return TPT_ERROR_BAD_ARGUMENT;
  }
  const pt::GpuTexture& t = s->host_textures[texture];
  if (levels) *levels = t.levels;
  if (channels) *channels = t.channels;
  if (max_value) *max_value = t.max_value;
  return TPT_OK;
}

int32_t tpt_texture_level(TptScene* s, uint32_t texture, uint32_t level, float* out, size_t capacity,
                          uint32_t* width, uint32_t* height) {
  try {
    if (!s || texture >= (s ? s->texture_count : 0)) {
      set_error("tpt_texture_level: null scene or texture index out of range");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    const pt::GpuTexture& t = s->host_textures[texture];
    if (level >= t.levels) {
      set_error("tpt_texture_level: level out of range");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    const pt::GpuTexLevel& L = s->host_tex_levels[t.first_level + level];
    if (width) *width = L.width;
    if (height) *height = L.height;
    if (!out) return TPT_OK;
    const size_t n = (size_t)L.width * L.height;
    if (capacity < n * t.channels) {
      set_error("tpt_texture_level: output capacity below width * height * channels");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice") ||
        !hip_ok(hipStreamSynchronize(s->dev->stream), "hipStreamSynchronize"))
      return TPT_ERROR_DEVICE;
    if (t.channels == 3) {
      std::vector<float> rgba(4 * n);
      if (!hip_ok(hipMemcpy(rgba.data(), s->tex_texels + L.offset, 4 * n * sizeof(float), hipMemcpyDeviceToHost),
                  "hipMemcpy(level)"))
        return TPT_ERROR_DEVICE;
      for (size_t k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c) out[3 * k + c] = rgba[4 * k + c];
    } else if (!hip_ok(hipMemcpy(out, s->tex_texels + L.offset, n * sizeof(float), hipMemcpyDeviceToHost),
                       "hipMemcpy(level)")) {
      return TPT_ERROR_DEVICE;
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_texture_level: ") + e.what());
    return TPT_ERROR;
  }
}

int32_t tpt_scene_set_texture(TptScene* s, uint32_t texture, const float* pixels, uint32_t flags) {
  try {
    if (!s || !pixels) {
      set_error("tpt_scene_set_texture: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (texture >= s->texture_count) {
      set_error("tpt_scene_set_texture: texture index out of range");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    for (uint32_t i = 0; i < (uint32_t)s->host_mats.size(); ++i) {
      const TptMaterial& m = s->host_mats[i];
      if (m.kind == TPT_MATERIAL_OPEN_PBR && m.u.open_pbr.emission_color_texture == texture) {
        set_error("tpt_scene_set_texture: texture " + std::to_string(texture) + " is the emission texture of material " +
                  std::to_string(i) + " (the light tree is built from its maximum: create a new scene)");
        return TPT_ERROR_BAD_ARGUMENT;
      }
    }
    if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    const hipStream_t st = s->dev->stream;
    const GpuTexture& t = s->host_textures[texture];
    const GpuTexLevel& L0 = s->host_tex_levels[t.first_level];
    const size_t n = (size_t)L0.width * L0.height;
    float* dst = s->tex_texels + L0.offset;
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    DeviceBuffer stage;
    const float* src = pixels;
    if (!dev_ptr) {
      if (!stage.ensure(n * t.channels * sizeof(float)) ||
          !hip_ok(hipMemcpyAsync(stage.ptr, pixels, n * t.channels * sizeof(float), hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync(pixels)"))
        return TPT_ERROR_DEVICE;
      src = (const float*)stage.ptr;
    }
    if (t.channels == 3)
      hipLaunchKernelGGL(unpack_rgb_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, st,
                         src, n, dst);
    else if (!hip_ok(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, st), "hipMemcpyAsync(level)"))
      return TPT_ERROR_DEVICE;
    if (!build_pyramid(s->tex_texels, s->host_tex_levels.data() + t.first_level, t, st) ||
        !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
      return TPT_ERROR_DEVICE;
    // the level-0 maximum (Texture::max_value) of the new pixels
    std::vector<float> host(n * t.channels);
    if (!hip_ok(hipMemcpy(host.data(), src, host.size() * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy(pixels)"))
      return TPT_ERROR_DEVICE;
    TptTextureDesc td{host.data(), L0.width, L0.height, t.channels, 0};
    s->host_textures[texture].max_value = texture_max_value(td);
    return hip_ok(hipMemcpy(&s->textures[texture], &s->host_textures[texture], sizeof(GpuTexture),
                            hipMemcpyHostToDevice),
                  "hipMemcpy(texture record)")
               ? TPT_OK
               : TPT_ERROR_DEVICE;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_scene_set_texture: ") + e.what());
    return TPT_ERROR;
  }
}

int32_t tpt_texture_sample(TptScene* s, uint32_t texture, uint32_t mode, const TptTextureQuery* queries, size_t n,
                           float* out, uint32_t flags) {
  try {
    if (!s || (n && (!queries || !out))) {
      set_error("tpt_texture_sample: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (texture >= s->texture_count) {
      set_error("tpt_texture_sample: texture index out of range");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (mode > TPT_TEXTURE_MIP_BILINEAR) {
      set_error("tpt_texture_sample: unknown mode");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    if (n == 0) return TPT_OK;
    if (!hip_ok(hipSetDevice(s->dev->index), "hipSetDevice")) return TPT_ERROR_DEVICE;
    const hipStream_t st = s->dev->stream;
    const uint32_t ch = s->host_textures[texture].channels;
    const bool dev_ptr = flags & TPT_FLAG_DEVICE_POINTERS;
    DeviceBuffer qb, rb;
    const TptTextureQuery* dq = queries;
    float* dr = out;
    if (!dev_ptr) {
      if (!qb.ensure(n * sizeof(TptTextureQuery)) || !rb.ensure(n * ch * sizeof(float))) return TPT_ERROR_DEVICE;
      if (!hip_ok(hipMemcpyAsync(qb.ptr, queries, n * sizeof(TptTextureQuery), hipMemcpyHostToDevice, st),
                  "hipMemcpyAsync(queries)"))
        return TPT_ERROR_DEVICE;
      dq = (const TptTextureQuery*)qb.ptr;
      dr = (float*)rb.ptr;
    }
    const TexView tv = tex_view_of(s);
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    if (ch == 3)
      hipLaunchKernelGGL(texture_query_kernel<3>, dim3(blocks), dim3(256), 0, st, tv, texture, mode, dq, n, dr);
    else
      hipLaunchKernelGGL(texture_query_kernel<1>, dim3(blocks), dim3(256), 0, st, tv, texture, mode, dq, n, dr);
    if (!hip_ok(hipGetLastError(), "texture_query_kernel launch")) return TPT_ERROR_DEVICE;
    if (!dev_ptr) {
      if (!hip_ok(hipMemcpyAsync(out, dr, n * ch * sizeof(float), hipMemcpyDeviceToHost, st), "hipMemcpyAsync(out)") ||
          !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return TPT_ERROR_DEVICE;
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_texture_sample: ") + e.what());
    return TPT_ERROR;
  }
}

}  // extern "C"
