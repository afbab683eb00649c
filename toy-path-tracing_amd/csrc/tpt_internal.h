// Internal declarations shared by the host builder, the scene upload and the
// kernels of libtpt.  Not part of the ABI (include/tpt.h is).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/tpt.h"

namespace tpt {

// Child reference encoding, R/qbvh.rs:15-51: leaf bit | 7-bit count | 24-bit
// offset, or an interior node index; EMPTY_CHILD = leaf bit with count 0.
constexpr uint32_t kChildLeafBit = 1u << 31;
constexpr uint32_t kChildCountShift = 24;
constexpr uint32_t kChildCountMask = 0x7Fu;
constexpr uint32_t kChildOffsetMask = (1u << 24) - 1;

struct HostQbvh {
  std::vector<TptQbvhNode> nodes;
  std::vector<uint32_t> prim;  // primitive_indices
};

TptQbvhNode empty_qbvh_node();
// Returns 0 on success, -1 on bad input, -2 when a leaf exceeds the child
// encoding (more than 2^24-1 primitive offset or 127 primitives).
int build_qbvh(const float* bmin, const float* bmax, size_t n, int threads, HostQbvh* out);

// ---- device layouts --------------------------------------------------------

// FP32 node: 4 child boxes as SoA float4 rows + child refs.  128 bytes, so a
// node is one aligned 128-byte line; the kernel reads the first 112 bytes.
struct alignas(16) GpuNode {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  uint32_t child[4];
  uint32_t pad[4];
};
static_assert(sizeof(GpuNode) == 128, "GpuNode must be 128 bytes");

// 8-bit quantised node, 64 bytes.  Child box corner = origin + q * 2^e per
// axis, evaluated in FP32 on the device exactly as on the host encoder; the
// encoder picks q so that the decoded box contains the FP32 box.
// flags bit 0: node not quantisable (non-finite bounds) -> use the FP32 node.
struct alignas(16) GpuNodeQ8 {
  float origin[3];
  int8_t exp[3];
  uint8_t flags;
  uint8_t q[24];  // lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4]
  uint32_t pad[2];
  uint32_t child[4];  // 16-byte aligned at offset 48
};
static_assert(sizeof(GpuNodeQ8) == 64, "GpuNodeQ8 must be 64 bytes");

// Leaf-ordered triangle record: the three vertex positions of
// primitive_indices[k] and that triangle index.  48 bytes, 40 read.
// Leaf-ordered triangle record: vertex positions and the triangle id, whose
// bit 31 flags a zero-area triangle (never hit, R/math/ray.rs:114-116).
constexpr uint32_t kTriIdMask = 0x7FFFFFFFu;
constexpr uint32_t kTriDegenerate = 0x80000000u;
struct alignas(16) GpuTri {
  float v[9];
  uint32_t prim;
  uint32_t pad[2];
};
static_assert(sizeof(GpuTri) == 48, "GpuTri must be 48 bytes");

// Instance: world_to_local (columns 0..3, xyz rows) and where its mesh lives.
struct alignas(16) GpuInstance {
  float m[12];
  uint32_t node_base;  // first node of the mesh's BLAS in the node array
  uint32_t node_count; // 0: root leaf over all triangles
  uint32_t tri_base;   // first leaf-ordered triangle of the mesh
  uint32_t tri_count;
};
static_assert(sizeof(GpuInstance) == 64, "GpuInstance must be 64 bytes");

// Per-instance shading record: local_to_world (columns 0..3, xyz rows),
// normal_to_world = the transpose of world_to_local's upper 3x3
// (Mat3::from_mat4(world_to_local.transpose()), R/scene/mod.rs:226), the
// material, and where the mesh's vertices / indices start.
struct alignas(16) GpuShade {
  float l2w[12];
  float nrm[9];
  uint32_t material;
  uint32_t vbase;
  uint32_t ibase;
  uint32_t flags;  // bit 0: vertex normals present, bit 1: uvs present
  uint32_t light_base;  // offset of this instance's triangles in the triangle -> light-leaf table, or ~0
  uint32_t pad[6];
};
static_assert(sizeof(GpuShade) == 128, "GpuShade must be 128 bytes");

// Host-side quantiser used by the scene upload (exposed for tests).
void quantize_node(const GpuNode& n, GpuNodeQ8* q);
GpuNode to_gpu_node(const TptQbvhNode& n, uint32_t interior_base);

}  // namespace tpt
