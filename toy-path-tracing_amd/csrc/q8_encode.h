// The 8-bit node encoder, compiled for the host (tpt_qbvh_quantize) and the
// device (the scene upload's encode kernel) from this one source, so both
// produce the same bytes: every step is exact or one IEEE rounding (q * 2^e
// is exact for q <= 255 and e in [-126, 120]; double division and
// floor / ceil are correctly rounded on both).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "tpt_internal.h"

namespace tpt {

// Decoded corner: origin + q * 2^e, one rounding (as the traversal decodes).
__host__ __device__ inline float q8_decode(float origin, int q, int e) {
  const float s = ldexpf(1.0f, e);
  return origin + (float)q * s;
}

__host__ __device__ inline bool q8_finite(float x) { return x - x == 0.0f; }

// Conservative 8-bit encoding of a node's child boxes: per axis the origin
// is the children's minimum and 2^e the smallest power of two with
// origin + 255 * 2^e >= the children's maximum; each child's lo rounds down
// and hi rounds up to the grid, so the decoded box contains the FP32 box.
// Empty slots get the inverted box (q lo 255, hi 0).  A node with a
// non-finite child bound is flagged (flags bit 0) for the FP32 path.
__host__ __device__ inline void quantize_node_impl(const GpuNode& n, GpuNodeQ8* out) {
  for (int k = 0; k < 3; ++k) {
    out->origin[k] = 0.0f;
    out->exp[k] = 0;
  }
  out->flags = 0;
  for (int k = 0; k < 24; ++k) out->q[k] = 0;
  out->pad[0] = out->pad[1] = 0;
  for (int i = 0; i < 4; ++i) out->child[i] = n.child[i];
  const float* lo[3] = {n.lox, n.loy, n.loz};
  const float* hi[3] = {n.hix, n.hiy, n.hiz};
  bool any = false;
  for (int i = 0; i < 4; ++i) {
    if (n.child[i] == kChildLeafBit) continue;
    any = true;
    for (int k = 0; k < 3; ++k)
      if (!q8_finite(lo[k][i]) || !q8_finite(hi[k][i])) {
        out->flags = 1;
        return;
      }
  }
  for (int k = 0; k < 3; ++k) {
    float o = INFINITY;
    float top = -INFINITY;
    for (int i = 0; i < 4; ++i) {
      if (n.child[i] == kChildLeafBit) continue;
      o = fminf(o, lo[k][i]);
      top = fmaxf(top, hi[k][i]);
    }
    if (!any) {
      o = 0.0f;
      top = 0.0f;
    }
    // smallest scale 2^e with origin + 255 * 2^e >= top
    const double span = (double)top - (double)o;
    int e = -126;
    if (span > 0.0) {
      int ex;
      (void)frexp(span / 255.0, &ex);  // span / 255 in [2^(ex-1), 2^ex)
      e = ex - 1;
      if (e < -126) e = -126;
    }
    while (q8_decode(o, 255, e) < top && e < 120) ++e;
    if (q8_decode(o, 255, e) < top) {
      out->flags = 1;
      return;
    }
    out->origin[k] = o;
    out->exp[k] = (int8_t)e;
    uint8_t* qlo = out->q + 8 * k;
    uint8_t* qhi = out->q + 8 * k + 4;
    const double s = ldexp(1.0, e);
    for (int i = 0; i < 4; ++i) {
      if (n.child[i] == kChildLeafBit) {
        qlo[i] = 255;
        qhi[i] = 0;
        continue;
      }
      // lo: the largest q with decode(q) <= lo; hi: the smallest q with decode(q) >= hi
      int ql = (int)floor(((double)lo[k][i] - (double)o) / s);
      ql = ql < 0 ? 0 : (ql > 255 ? 255 : ql);
      while (ql > 0 && q8_decode(o, ql, e) > lo[k][i]) --ql;
      while (ql < 255 && q8_decode(o, ql + 1, e) <= lo[k][i]) ++ql;
      int qh = (int)ceil(((double)hi[k][i] - (double)o) / s);
      qh = qh < 0 ? 0 : (qh > 255 ? 255 : qh);
      while (qh < 255 && q8_decode(o, qh, e) < hi[k][i]) ++qh;
      while (qh > 0 && q8_decode(o, qh - 1, e) >= hi[k][i]) --qh;
      qlo[i] = (uint8_t)ql;
      qhi[i] = (uint8_t)qh;
    }
  }
}

}  // namespace tpt
