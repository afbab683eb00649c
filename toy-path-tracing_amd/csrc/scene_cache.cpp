// The scene description as an on-disk cache: tpt_scene_desc_save writes a
// TptSceneDesc (meshes with their built QBVHs, instances, TLAS, materials,
// environment, textures, directional lights) to one file, and
// tpt_scene_desc_load reads it back into a descriptor that tpt_scene_create
// accepts, so a scene is generated and built once and then uploaded from the
// file.  No reference counterpart (the reference rebuilds its scene on every
// run, R/main.rs:60-97).
//
// File layout (little-endian, every value and array padded to 8 bytes):
//   header   magic "TPTSCENE", format version, the sizes of the records it
//            stores (so a build with other record layouts refuses the file),
//            payload bytes, checksum of the payload
//   payload  the descriptor's counts and scalars, then per mesh its counts,
//            presence flags and arrays, then the instance, TLAS, material,
//            environment, texture and directional-light arrays
// The checksum is FNV-1a 64 over the payload's 64-bit words in four
// interleaved lanes (word i feeds lane i mod 4), the lanes folded in order
// with the word count.  Saving streams the arrays to the file (no copy of
// the scene in memory); a loaded descriptor owns one allocation holding the
// TptSceneDesc, its mesh / texture tables and the payload read in place.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "tpt_device.h"

namespace {

constexpr char kMagic[8] = {'T', 'P', 'T', 'S', 'C', 'E', 'N', 'E'};
constexpr uint32_t kVersion = 1;
constexpr uint64_t kFnvBasis = 1469598103934665603ull, kFnvPrime = 1099511628211ull;

struct Header {
  char magic[8];
  uint32_t version;
  uint32_t sizes[6];  // TptMeshDesc, TptInstanceDesc, TptQbvhNode, TptMaterial, TptTextureDesc, TptDirectionalLight
  uint32_t pad;
  uint64_t payload_bytes;
  uint64_t checksum;
};
static_assert(sizeof(Header) == 56, "cache header layout");

void record_sizes(uint32_t* s) {
  s[0] = sizeof(TptMeshDesc);
  s[1] = sizeof(TptInstanceDesc);
  s[2] = sizeof(TptQbvhNode);
  s[3] = sizeof(TptMaterial);
  s[4] = sizeof(TptTextureDesc);
  s[5] = sizeof(TptDirectionalLight);
}

size_t align8(size_t n) { return (n + 7) & ~size_t(7); }

// The payload checksum, fed whole 64-bit words.
struct Checksum {
  uint64_t lane[4] = {kFnvBasis, kFnvBasis ^ 1u, kFnvBasis ^ 2u, kFnvBasis ^ 3u};
  uint64_t words = 0;
  void add(const unsigned char* p, size_t n_words) {
    for (size_t i = 0; i < n_words; ++i, ++words) {
      uint64_t w;
      std::memcpy(&w, p + 8 * i, 8);
      uint64_t& h = lane[words & 3];
      h = (h ^ w) * kFnvPrime;
    }
  }
  uint64_t value() const {
    uint64_t h = kFnvBasis;
    for (uint64_t l : lane) h = (h ^ l) * kFnvPrime;
    return (h ^ words) * kFnvPrime;
  }
};

// Payload writer: streams values and arrays to the file, each padded to 8
// bytes, and keeps the checksum and the byte count.
struct Writer {
  FILE* f;
  Checksum sum;
  uint64_t bytes = 0;
  bool ok = true;
  void raw(const void* p, size_t n) {
    if (!ok || n == 0) return;
    const unsigned char* c = static_cast<const unsigned char*>(p);
    const size_t whole = n & ~size_t(7);
    sum.add(c, whole / 8);
    ok = std::fwrite(c, 1, n, f) == n;
    if (n != whole) {  // the last partial word, zero padded
      unsigned char tail[8] = {};
      std::memcpy(tail, c + whole, n - whole);
      sum.add(tail, 1);
      const size_t pad = 8 - (n - whole);
      ok = ok && std::fwrite(tail + (n - whole), 1, pad, f) == pad;
    }
    bytes += align8(n);
  }
  void u32(uint32_t v) { raw(&v, 4); }
  void f32(float v) { raw(&v, 4); }
};

// Payload reader over the loaded payload: every read checks the bounds.
struct Reader {
  unsigned char* p;
  size_t n, at = 0;
  bool ok = true;
  void* take(size_t bytes) {
    if (!ok || bytes > n - at) {
      ok = false;
      return nullptr;
    }
    void* r = p + at;
    at = align8(at + bytes);
    if (at > n) at = n;
    return r;
  }
  // an array of a * b * c bytes (a product that overflows fails the read)
  void* take(uint64_t a, uint64_t b, uint64_t c = 1) {
    uint64_t ab, abc;
    if (__builtin_mul_overflow(a, b, &ab) || __builtin_mul_overflow(ab, c, &abc) || abc > SIZE_MAX) {
      ok = false;
      return nullptr;
    }
    return take((size_t)abc);
  }
  uint32_t u32() {
    uint32_t v = 0;
    if (void* q = take(4)) std::memcpy(&v, q, 4);
    return v;
  }
  float f32() {
    float v = 0.0f;
    if (void* q = take(4)) std::memcpy(&v, q, 4);
    return v;
  }
};

bool check_desc(const TptSceneDesc* d, std::string* why) {
  if (d->mesh_count && !d->meshes) *why = "meshes is NULL";
  else if (d->instance_count && !d->instances) *why = "instances is NULL";
  else if (d->tlas_node_count && !d->tlas_nodes) *why = "tlas_nodes is NULL";
  else if (d->instance_count && !d->tlas_primitive_indices) *why = "tlas_primitive_indices is NULL";
  else if (d->material_count && !d->materials) *why = "materials is NULL";
  else if (d->texture_count && !d->textures) *why = "textures is NULL";
  else if (d->directional_count && !d->directional_lights) *why = "directional_lights is NULL";
  else if (d->env_pixels && (d->env_width == 0 || d->env_height == 0)) *why = "environment without a size";
  else {
    for (uint32_t m = 0; m < d->mesh_count; ++m) {
      const TptMeshDesc& md = d->meshes[m];
      if ((md.vertex_count && !md.positions) || (md.triangle_count && (!md.indices || !md.primitive_indices)) ||
          (md.node_count && !md.nodes)) {
        *why = "mesh " + std::to_string(m) + " has a NULL array";
        return false;
      }
    }
    for (uint32_t t = 0; t < d->texture_count; ++t) {
      const TptTextureDesc& td = d->textures[t];
      if (!td.pixels || (td.channels != 1 && td.channels != 3)) {
        *why = "texture " + std::to_string(t) + " has no pixels or a channel count not 1 or 3";
        return false;
      }
    }
    return true;
  }
  return false;
}

void write_payload(const TptSceneDesc* d, Writer& w) {
  w.u32(d->mesh_count);
  w.u32(d->instance_count);
  w.u32(d->tlas_node_count);
  w.u32(d->material_count);
  w.u32(d->texture_count);
  w.u32(d->directional_count);
  w.u32(d->env_pixels ? 1u : 0u);
  w.u32(d->env_width);
  w.u32(d->env_height);
  w.f32(d->env_scale);
  w.f32(d->env_rotate_y);
  for (uint32_t m = 0; m < d->mesh_count; ++m) {
    const TptMeshDesc& md = d->meshes[m];
    w.u32(md.vertex_count);
    w.u32(md.triangle_count);
    w.u32(md.node_count);
    w.u32(md.normals ? 1u : 0u);
    w.u32(md.uvs ? 1u : 0u);
    w.raw(md.positions, 12ull * md.vertex_count);
    w.raw(md.indices, 12ull * md.triangle_count);
    w.raw(md.nodes, sizeof(TptQbvhNode) * (size_t)md.node_count);
    w.raw(md.primitive_indices, 4ull * md.triangle_count);
    if (md.normals) w.raw(md.normals, 12ull * md.vertex_count);
    if (md.uvs) w.raw(md.uvs, 8ull * md.vertex_count);
  }
  w.raw(d->instances, sizeof(TptInstanceDesc) * (size_t)d->instance_count);
  w.raw(d->tlas_nodes, sizeof(TptQbvhNode) * (size_t)d->tlas_node_count);
  w.raw(d->tlas_primitive_indices, 4ull * d->instance_count);
  w.raw(d->materials, sizeof(TptMaterial) * (size_t)d->material_count);
  if (d->env_pixels) w.raw(d->env_pixels, 12ull * d->env_width * d->env_height);
  for (uint32_t t = 0; t < d->texture_count; ++t) {
    const TptTextureDesc& td = d->textures[t];
    w.u32(td.width);
    w.u32(td.height);
    w.u32(td.channels);
    w.raw(td.pixels, 4ull * td.width * td.height * td.channels);
  }
  w.raw(d->directional_lights, sizeof(TptDirectionalLight) * (size_t)d->directional_count);
}

}  // namespace

using namespace tpt;

extern "C" int32_t tpt_scene_desc_save(const TptSceneDesc* d, const char* path) {
  try {
    if (!d || !path) {
      set_error("tpt_scene_desc_save: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    std::string why;
    if (!check_desc(d, &why)) {
      set_error("tpt_scene_desc_save: " + why);
      return TPT_ERROR_BAD_ARGUMENT;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
      set_error(std::string("tpt_scene_desc_save: cannot open ") + path);
      return TPT_ERROR;
    }
    // a header of zeros first (a file cut short here is no cache), the
    // payload, then the header with the payload's size and checksum
    Header h{};
    Writer w{f};
    w.ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    write_payload(d, w);
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    record_sizes(h.sizes);
    h.payload_bytes = w.bytes;
    h.checksum = w.sum.value();
    const bool ok = w.ok && std::fseek(f, 0, SEEK_SET) == 0 && std::fwrite(&h, sizeof(h), 1, f) == 1;
    if (std::fclose(f) != 0 || !ok) {
      std::remove(path);
      set_error(std::string("tpt_scene_desc_save: write failed: ") + path);
      return TPT_ERROR;
    }
    return TPT_OK;
  } catch (const std::exception& e) {
    set_error(std::string("tpt_scene_desc_save: ") + e.what());
    return TPT_ERROR;
  }
}

extern "C" TptSceneDesc* tpt_scene_desc_load(const char* path) {
  FILE* f = nullptr;
  unsigned char* block = nullptr;
  auto fail = [&](const std::string& why) -> TptSceneDesc* {
    if (f) std::fclose(f);
    ::operator delete(block);
    set_error(std::string("tpt_scene_desc_load: ") + (path ? path : "(null)") + ": " + why);
    return nullptr;
  };
  try {
    if (!path) return fail("null path");
    f = std::fopen(path, "rb");
    if (!f) {
      set_error(std::string("tpt_scene_desc_load: cannot open ") + path);
      return nullptr;
    }
    Header h{};
    uint32_t sizes[6];
    record_sizes(sizes);
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0) return fail("not a scene cache");
    if (h.version != kVersion) return fail("format version " + std::to_string(h.version) + " (expected 1)");
    if (std::memcmp(h.sizes, sizes, sizeof(sizes)) != 0) return fail("written with other record layouts");
    // the payload must be the rest of the file
    const long start = std::ftell(f);
    if (start < 0 || std::fseek(f, 0, SEEK_END) != 0) return fail("cannot size the file");
    const long end = std::ftell(f);
    if (end < start || (uint64_t)(end - start) != h.payload_bytes || (h.payload_bytes & 7) != 0 ||
        h.payload_bytes < 48 || std::fseek(f, start, SEEK_SET) != 0)
      return fail("truncated or corrupt (payload size)");
    // the six counts lead the payload: they size the descriptor's tables
    unsigned char lead[48];
    if (std::fread(lead, 1, sizeof(lead), f) != sizeof(lead) || std::fseek(f, start, SEEK_SET) != 0)
      return fail("truncated or corrupt (payload size)");
    uint32_t counts[6];
    for (int i = 0; i < 6; ++i) std::memcpy(&counts[i], lead + 8 * i, 4);
    const uint32_t mesh_count = counts[0], texture_count = counts[4];
    if (mesh_count > h.payload_bytes || texture_count > h.payload_bytes) return fail("counts beyond the file");
    const size_t tables = align8(sizeof(TptSceneDesc)) + align8(sizeof(TptMeshDesc) * (size_t)mesh_count) +
                          align8(sizeof(TptTextureDesc) * (size_t)texture_count);
    block = static_cast<unsigned char*>(::operator new(tables + h.payload_bytes, std::nothrow));
    if (!block) return fail("out of host memory");
    unsigned char* payload = block + tables;
    if (std::fread(payload, 1, h.payload_bytes, f) != h.payload_bytes) return fail("truncated (read)");
    std::fclose(f);
    f = nullptr;
    Checksum sum;
    sum.add(payload, h.payload_bytes / 8);
    if (sum.value() != h.checksum) return fail("truncated or corrupt (checksum)");

    TptSceneDesc* d = reinterpret_cast<TptSceneDesc*>(block);
    std::memset(d, 0, sizeof(*d));
    TptMeshDesc* meshes = reinterpret_cast<TptMeshDesc*>(block + align8(sizeof(TptSceneDesc)));
    TptTextureDesc* texs = reinterpret_cast<TptTextureDesc*>(block + align8(sizeof(TptSceneDesc)) +
                                                             align8(sizeof(TptMeshDesc) * (size_t)mesh_count));
    Reader r{payload, (size_t)h.payload_bytes};
    d->mesh_count = r.u32();
    d->instance_count = r.u32();
    d->tlas_node_count = r.u32();
    d->material_count = r.u32();
    d->texture_count = r.u32();
    d->directional_count = r.u32();
    const bool has_env = r.u32() != 0;
    d->env_width = r.u32();
    d->env_height = r.u32();
    d->env_scale = r.f32();
    d->env_rotate_y = r.f32();
    d->meshes = mesh_count ? meshes : nullptr;
    for (uint32_t m = 0; m < mesh_count && r.ok; ++m) {
      TptMeshDesc& md = meshes[m];
      std::memset(&md, 0, sizeof(md));
      md.vertex_count = r.u32();
      md.triangle_count = r.u32();
      md.node_count = r.u32();
      const bool nrm = r.u32() != 0, uv = r.u32() != 0;
      md.positions = static_cast<const float*>(r.take(12, md.vertex_count));
      md.indices = static_cast<const uint32_t*>(r.take(12, md.triangle_count));
      md.nodes = md.node_count ? static_cast<const TptQbvhNode*>(r.take(sizeof(TptQbvhNode), md.node_count)) : nullptr;
      md.primitive_indices = static_cast<const uint32_t*>(r.take(4, md.triangle_count));
      md.normals = nrm ? static_cast<const float*>(r.take(12, md.vertex_count)) : nullptr;
      md.uvs = uv ? static_cast<const float*>(r.take(8, md.vertex_count)) : nullptr;
    }
    d->instances = static_cast<const TptInstanceDesc*>(r.take(sizeof(TptInstanceDesc), d->instance_count));
    d->tlas_nodes =
        d->tlas_node_count ? static_cast<const TptQbvhNode*>(r.take(sizeof(TptQbvhNode), d->tlas_node_count)) : nullptr;
    d->tlas_primitive_indices = static_cast<const uint32_t*>(r.take(4, d->instance_count));
    d->materials = static_cast<const TptMaterial*>(r.take(sizeof(TptMaterial), d->material_count));
    d->env_pixels = has_env ? static_cast<const float*>(r.take(12, d->env_width, d->env_height)) : nullptr;
    d->textures = texture_count ? texs : nullptr;
    for (uint32_t t = 0; t < texture_count && r.ok; ++t) {
      TptTextureDesc& td = texs[t];
      std::memset(&td, 0, sizeof(td));
      td.width = r.u32();
      td.height = r.u32();
      td.channels = r.u32();
      td.pixels = static_cast<const float*>(r.take(4ull * td.channels, td.width, td.height));
    }
    d->directional_lights =
        d->directional_count
            ? static_cast<const TptDirectionalLight*>(r.take(sizeof(TptDirectionalLight), d->directional_count))
            : nullptr;
    if (!r.ok || r.at != r.n || d->mesh_count != mesh_count || d->texture_count != texture_count)
      return fail("arrays do not match the payload");
    return d;
  } catch (const std::exception& e) {
    return fail(e.what());
  }
}

extern "C" void tpt_scene_desc_free(TptSceneDesc* desc) { ::operator delete(desc); }
