// Host-side conversion of the reference QBVH layout into the device layouts:
// FP32 GpuNode (interior indices rebased into one node array per scene) and
// the 8-bit quantised GpuNodeQ8 with conservative outward rounding (the
// encoder in q8_encode.h, which the scene upload runs on the device).
#include <cmath>
#include <cstring>
#include <limits>

#include "q8_encode.h"
#include "tpt_internal.h"

namespace tpt {

GpuNode to_gpu_node(const TptQbvhNode& n, uint32_t interior_base) {
  GpuNode g;
  std::memset(&g, 0, sizeof(g));
  for (int i = 0; i < 4; ++i) {
    g.lox[i] = n.min_x[i];
    g.hix[i] = n.max_x[i];
    g.loy[i] = n.min_y[i];
    g.hiy[i] = n.max_y[i];
    g.loz[i] = n.min_z[i];
    g.hiz[i] = n.max_z[i];
    uint32_t c = n.children[i];
    // A slot outside valid_mask never reaches the leaf callback
    // (R/qbvh.rs:602 masks it); it is stored as EMPTY_CHILD.
    if (!(n.valid_mask & (1u << i))) c = kChildLeafBit;
    if (!(c & kChildLeafBit)) c += interior_base;
    g.child[i] = c;
  }
  return g;
}

void quantize_node(const GpuNode& n, GpuNodeQ8* out) { quantize_node_impl(n, out); }

}  // namespace tpt
