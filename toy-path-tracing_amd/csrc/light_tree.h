// Host light-tree build (light_tree.cpp) and the device light-tree record.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tpt.h"

namespace tpt {

// Builds the scene's SG light tree over its emissive triangles.  out: nodes
// in depth-first pre-order (root 0); light_base[instance]: offset of the
// instance's triangles in tri_leaf, or 0xFFFFFFFF for a non-emissive
// instance; tri_leaf: leaf node of each emissive triangle.
bool build_light_tree(const TptSceneDesc& desc, std::vector<TptLightNode>* out, std::vector<uint32_t>* light_base,
                      std::vector<uint32_t>* tri_leaf, std::string* err);

// Device light-tree node: the fields the stochastic descent reads, with the
// two per-node constants of sg_light_for_node (R/light_tree/lobe.rs:233-274)
// evaluated on the host: bound_var = 0.5 r^2 and sg_norm =
// max(sg_integral(lambda), 1e-30).
struct alignas(16) GpuLightNode {
  float mu[3];
  float flux;
  float nu[3];
  float lambda;
  float sigma_s2;
  float bound_var;
  float sg_norm;
  float area;
  uint32_t left, right, parent, instance;
  uint32_t triangle, pad[3];
};
static_assert(sizeof(GpuLightNode) == 80, "GpuLightNode must be 80 bytes");

// max(sg_integral(lambda), 1e-30) with sg_integral = 4 pi expm1_over_x(-2 lambda)
// (R/math/sg.rs:59-69, 124-126), host libm.
float sg_norm_host(float lambda);

}  // namespace tpt
