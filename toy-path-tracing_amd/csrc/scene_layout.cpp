// Host-side conversion of the reference QBVH layout into the device layouts:
// FP32 GpuNode (interior indices rebased into one node array per scene) and
// the 8-bit quantised GpuNodeQ8 with conservative outward rounding.
#include <cmath>
#include <cstring>
#include <limits>

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

namespace {

// Decoded corner, bit-identical on host and device: q * 2^e is exact for
// q <= 255 and e in [-126, 120], and the single addition rounds to nearest
// (the device's fused multiply-add of the same exact product rounds once too).
inline float decode(float origin, int q, int e) {
  const float s = std::ldexp(1.0f, e);
  return origin + (float)q * s;
}

}  // namespace

void quantize_node(const GpuNode& n, GpuNodeQ8* out) {
  std::memset(out, 0, sizeof(*out));
  for (int i = 0; i < 4; ++i) out->child[i] = n.child[i];
  const float* lo[3] = {n.lox, n.loy, n.loz};
  const float* hi[3] = {n.hix, n.hiy, n.hiz};
  bool any = false;
  for (int i = 0; i < 4; ++i) {
    if (n.child[i] == kChildLeafBit) continue;
    any = true;
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(lo[k][i]) || !std::isfinite(hi[k][i])) {
        out->flags = 1;
        return;
      }
  }
  for (int k = 0; k < 3; ++k) {
    float o = std::numeric_limits<float>::infinity();
    float top = -std::numeric_limits<float>::infinity();
    for (int i = 0; i < 4; ++i) {
      if (n.child[i] == kChildLeafBit) continue;
      o = std::fmin(o, lo[k][i]);
      top = std::fmax(top, hi[k][i]);
    }
    if (!any) {
      o = 0.0f;
      top = 0.0f;
    }
    // Smallest scale 2^e with origin + 255 * 2^e >= top.
    const double span = (double)top - (double)o;
    int e = -126;
    if (span > 0.0) {
      int ex;
      std::frexp(span / 255.0, &ex);  // span/255 in [2^(ex-1), 2^ex)
      e = ex - 1;
      if (e < -126) e = -126;
    }
    while (decode(o, 255, e) < top && e < 120) ++e;
    if (decode(o, 255, e) < top) {
      out->flags = 1;
      return;
    }
    out->origin[k] = o;
    out->exp[k] = (int8_t)e;
    uint8_t* qlo = out->q + 8 * k;
    uint8_t* qhi = out->q + 8 * k + 4;
    const double s = std::ldexp(1.0, e);
    for (int i = 0; i < 4; ++i) {
      if (n.child[i] == kChildLeafBit) {
        qlo[i] = 255;
        qhi[i] = 0;
        continue;
      }
      // lo: largest q with decode(q) <= lo ; hi: smallest q with decode(q) >= hi.
      int ql = (int)std::floor(((double)lo[k][i] - (double)o) / s);
      ql = ql < 0 ? 0 : (ql > 255 ? 255 : ql);
      while (ql > 0 && decode(o, ql, e) > lo[k][i]) --ql;
      while (ql < 255 && decode(o, ql + 1, e) <= lo[k][i]) ++ql;
      int qh = (int)std::ceil(((double)hi[k][i] - (double)o) / s);
      qh = qh < 0 ? 0 : (qh > 255 ? 255 : qh);
      while (qh < 255 && decode(o, qh, e) < hi[k][i]) ++qh;
      while (qh > 0 && decode(o, qh - 1, e) >= hi[k][i]) --qh;
      qlo[i] = (uint8_t)ql;
      qhi[i] = (uint8_t)qh;
    }
  }
}

}  // namespace tpt
