// Texture lookups on the device: the filtered, mip-mapped Texture of
// R/material/texture.rs (wrap addressing, bilinear, nearest, trilinear,
// mip-bilinear and the EWA filter of sample_filtered), over the level table
// of TexView.  The arithmetic keeps the reference's f32 expression order
// (the library is built without contraction), so every mode except the EWA
// weights (expf) and the level choice (log2f) is exact against a host
// evaluation of the same formulas.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tpt_device.h"
#include "traverse.h"

namespace tpt {
namespace tex {

using pt::GpuTexLevel;
using pt::GpuTexture;
using pt::TexView;

constexpr float kMaxAnisotropy = 8.0f;    // MAX_ANISOTROPY, R/material/texture.rs:86
constexpr float kMinFilterWidth = 1.0e-8f;  // MIN_FILTER_WIDTH, :87
constexpr float kEwaAlpha = 2.0f;          // EWA_ALPHA, :88
// EWA footprints span at most about 2 sqrt(16^2 + 16^2 + 1) + 1 = 47 texels
// per axis once the level is chosen (see ewa_level); a wider box arises only
// where |uv| * size exceeds about 2^30, where float spacing exceeds the filter
// radius, and those lookups take the level's bilinear value.
constexpr long long kMaxEwaSpan = 256;

struct V2 {
  float x, y;
};
__device__ __forceinline__ V2 v2(float x, float y) { return V2{x, y}; }
__device__ __forceinline__ bool finite2(V2 a) { return __builtin_isfinite(a.x) && __builtin_isfinite(a.y); }
__device__ __forceinline__ float len2(V2 a) { return a.x * a.x + a.y * a.y; }

// TexturePixel::lerp: self + (other - self) * t per channel (glam 0.32 Vec3::lerp)
__device__ __forceinline__ F3 lerp(F3 a, F3 b, float t) {
  return f3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}

// wrap_unit (:575-581): t.rem_euclid(1.0) for finite t, else 0.  fmodf is
// exact, so this equals Rust's f32 % followed by the sign fix.
__device__ __forceinline__ float wrap_unit(float t) {
  if (!__builtin_isfinite(t)) return 0.0f;
  const float r = fmodf(t, 1.0f);
  return r < 0.0f ? r + 1.0f : r;
}
// wrap_index (:583-589) for 32-bit indices: mask for a power-of-two size
// (two's complement, as (index as usize) & (size - 1)), else rem_euclid.
__device__ __forceinline__ uint32_t wrap_index(int i, uint32_t n, bool pow2) {
  if (pow2) return (uint32_t)i & (n - 1u);
  const int m = i % (int)n;
  return (uint32_t)(m < 0 ? m + (int)n : m);
}
__device__ __forceinline__ uint32_t wrap_index64(long long i, uint32_t n, bool pow2) {
  if (pow2) return (uint32_t)((unsigned long long)i & (unsigned long long)(n - 1u));
  const long long m = i % (long long)n;
  return (uint32_t)(m < 0 ? m + (long long)n : m);
}
__device__ __forceinline__ uint32_t wrap_any(int i, uint32_t n, bool pow2) { return wrap_index(i, n, pow2); }
__device__ __forceinline__ uint32_t wrap_any(long long i, uint32_t n, bool pow2) { return wrap_index64(i, n, pow2); }
// `f as isize` (saturating, NaN -> 0)
__device__ __forceinline__ long long to_isize(float f) {
  if (!(f == f)) return 0;
  if (f >= 9.2233720368547758e18f) return 0x7fffffffffffffffLL;
  if (f <= -9.2233720368547758e18f) return (long long)0x8000000000000000ULL;
  return (long long)f;
}

template <int C>
__device__ __forceinline__ F3 texel(const TexView& tv, const GpuTexLevel& L, uint32_t x, uint32_t y) {
  const size_t i = (size_t)y * L.width + x;
  if (C == 3) {
    const float4 v = *reinterpret_cast<const float4*>(tv.texels + L.offset + 4 * i);
    return f3(v.x, v.y, v.z);
  }
  return f3(tv.texels[L.offset + i], 0.0f, 0.0f);
}
// pixel_level_wrapped (:302-308); `level` already below the level count
template <int C>
__device__ __forceinline__ F3 pixel(const TexView& tv, const GpuTexLevel& L, int x, int y) {
  return texel<C>(tv, L, wrap_index(x, L.width, L.pow2 & 1u), wrap_index(y, L.height, (L.pow2 >> 1) & 1u));
}

__device__ __forceinline__ const GpuTexLevel& level_of(const TexView& tv, const GpuTexture& t, uint32_t level) {
  return tv.levels[t.first_level + (level < t.levels ? level : t.levels - 1u)];
}

// bilerp_level (:225-243)
template <int C>
__device__ __forceinline__ F3 bilerp(const TexView& tv, const GpuTexture& t, uint32_t level, V2 uv) {
  const GpuTexLevel& L = level_of(tv, t, level);
  const float u = wrap_unit(uv.x), v = wrap_unit(uv.y);
  const float x = u * (float)L.width - 0.5f, y = v * (float)L.height - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const int x0 = (int)fx, y0 = (int)fy;
  const float tx = x - (float)x0, ty = y - (float)y0;
  const F3 c00 = pixel<C>(tv, L, x0, y0), c10 = pixel<C>(tv, L, x0 + 1, y0);
  const F3 c01 = pixel<C>(tv, L, x0, y0 + 1), c11 = pixel<C>(tv, L, x0 + 1, y0 + 1);
  return lerp(lerp(c00, c10, tx), lerp(c01, c11, tx), ty);
}

// The EWA footprint sum over the lattice points s0..=s1 x t0..=t1, with
// 32-bit indices when every bound fits (I = int) and 64-bit ones otherwise.
template <int C, typename I>
__device__ __forceinline__ bool ewa_sum(const TexView& tv, const GpuTexLevel& L, V2 st, float a, float b, float c,
                                        I s0, I s1, I t0, I t1, F3* out) {
  const float cutoff = expf(-kEwaAlpha);
  const bool px2 = L.pow2 & 1u, py2 = (L.pow2 >> 1) & 1u;
  F3 sum = f3(0.0f, 0.0f, 0.0f);
  float wsum = 0.0f;
  uint32_t y = wrap_any(t0, L.height, py2);
  for (I it = t0; it <= t1; ++it) {
    const float tt = (float)it - st.y;
    uint32_t x = wrap_any(s0, L.width, px2);
    for (I is = s0; is <= s1; ++is) {
      const float ss = (float)is - st.x;
      const float r2 = a * ss * ss + b * ss * tt + c * tt * tt;
      if (!(r2 >= 1.0f)) {
        const float w = expf(-kEwaAlpha * r2) - cutoff;
        const F3 p = texel<C>(tv, L, x, y);
        sum = f3(sum.x + p.x * w, sum.y + p.y * w, sum.z + p.z * w);
        wsum += w;
      }
      x = x + 1u == L.width ? 0u : x + 1u;
    }
    y = y + 1u == L.height ? 0u : y + 1u;
  }
  if (!(wsum > 0.0f)) return false;
  const float r = 1.0f / wsum;
  *out = f3(sum.x * r, sum.y * r, sum.z * r);
  return true;
}

// ewa_level (:245-300).  The footprint loop runs over the integer lattice
// points inside the ellipse's bounding box; its half-widths are
// sqrt(|d_x|^2 + 1) texels in each axis (the +1 terms bound the normalised
// quadratic form), a few texels once the level is chosen from the shorter
// axis.  A box wider than kMaxEwaSpan (or with a NaN bound: non-finite uv)
// takes the level's bilinear value, which is also the reference's result
// for non-finite uv (its NaN weights fail the weight_sum > 0 test).
template <int C>
__device__ __noinline__ F3 ewa_level(const TexView& tv, const GpuTexture& t, uint32_t level, V2 uv, V2 d0, V2 d1) {
  if (level >= t.levels) {
    const GpuTexLevel& L = tv.levels[t.first_level + t.levels - 1u];
    return texel<C>(tv, L, 0, 0);
  }
  const GpuTexLevel& L = tv.levels[t.first_level + level];
  const float W = (float)L.width, H = (float)L.height;
  const V2 st = v2(uv.x * W - 0.5f, uv.y * H - 0.5f);
  d0 = v2(d0.x * W, d0.y * H);
  d1 = v2(d1.x * W, d1.y * H);
  float a = d0.y * d0.y + d1.y * d1.y + 1.0f;
  float b = -2.0f * (d0.x * d0.y + d1.x * d1.y);
  float c = d0.x * d0.x + d1.x * d1.x + 1.0f;
  const float inv_f = 1.0f / (a * c - 0.25f * b * b);
  if (!__builtin_isfinite(inv_f)) return bilerp<C>(tv, t, level, uv);
  a *= inv_f;
  b *= inv_f;
  c *= inv_f;
  const float det = -b * b + 4.0f * a * c;
  if (det <= 0.0f || !__builtin_isfinite(det)) return bilerp<C>(tv, t, level, uv);
  const float inv_det = 1.0f / det;
  const float u_sqrt = sqrtf(fmaxf(det * c, 0.0f)), v_sqrt = sqrtf(fmaxf(a * det, 0.0f));
  const float fs0 = ceilf(st.x - 2.0f * inv_det * u_sqrt), fs1 = floorf(st.x + 2.0f * inv_det * u_sqrt);
  const float ft0 = ceilf(st.y - 2.0f * inv_det * v_sqrt), ft1 = floorf(st.y + 2.0f * inv_det * v_sqrt);
  if (!(fs1 - fs0 <= (float)kMaxEwaSpan) || !(ft1 - ft0 <= (float)kMaxEwaSpan)) return bilerp<C>(tv, t, level, uv);
  F3 out;
  bool ok;
  const float lim = 1073741824.0f;  // 2^30: bounds and bound + 1 stay in int
  if (fabsf(fs0) < lim && fabsf(fs1) < lim && fabsf(ft0) < lim && fabsf(ft1) < lim)
    ok = ewa_sum<C, int>(tv, L, st, a, b, c, (int)fs0, (int)fs1, (int)ft0, (int)ft1, &out);
  else
    ok = ewa_sum<C, long long>(tv, L, st, a, b, c, to_isize(fs0), to_isize(fs1), to_isize(ft0), to_isize(ft1), &out);
  return ok ? out : bilerp<C>(tv, t, level, uv);
}

// sample_ewa / sample_filtered (:140-193).  Not inlined: a material reads
// up to six textures per vertex.
template <int C>
__device__ __noinline__ F3 filtered(const TexView& tv, const GpuTexture& t, V2 uv, V2 d0, V2 d1) {
  if (!finite2(d0) || !finite2(d1)) return bilerp<C>(tv, t, 0, uv);
  if (len2(d0) < len2(d1)) {
    const V2 k = d0;
    d0 = d1;
    d1 = k;
  }
  const float longer = sqrtf(len2(d0));
  float shorter = sqrtf(len2(d1));
  if (shorter > 0.0f && shorter * kMaxAnisotropy < longer) {
    const float scale = longer / (shorter * kMaxAnisotropy);
    d1 = v2(d1.x * scale, d1.y * scale);
    shorter *= scale;
  }
  if (shorter <= 0.0f) return bilerp<C>(tv, t, 0, uv);
  const float lod = fmaxf((float)(t.levels - 1u) + log2f(fmaxf(shorter, kMinFilterWidth)), 0.0f);
  const float fl = floorf(lod);
  const uint32_t level = fl >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)fl;
  const float tl = lod - fl;
  const F3 a = ewa_level<C>(tv, t, level, uv, d0, d1);
  const F3 b = ewa_level<C>(tv, t, level == 0xFFFFFFFFu ? level : level + 1u, uv, d0, d1);
  return lerp(a, b, tl);
}

// sample_trilinear (:200-223)
template <int C>
__device__ __forceinline__ F3 trilinear(const TexView& tv, const GpuTexture& t, V2 uv, V2 dx, V2 dy) {
  if (!finite2(dx) || !finite2(dy)) return bilerp<C>(tv, t, 0, uv);
  const float m = fmaxf(fmaxf(fmaxf(fabsf(dx.x), fabsf(dy.x)), fmaxf(fabsf(dx.y), fabsf(dy.y))), kMinFilterWidth);
  const float width = 2.0f * m;
  const float top = (float)(t.levels - 1u);
  const float lod = top + log2f(width);
  if (lod >= top) return texel<C>(tv, tv.levels[t.first_level + t.levels - 1u], 0, 0);
  const float fl = fmaxf(floorf(lod), 0.0f);
  const uint32_t level = (uint32_t)fl;
  if (level == 0) return bilerp<C>(tv, t, 0, uv);
  return lerp(bilerp<C>(tv, t, level, uv), bilerp<C>(tv, t, level + 1u, uv), lod - (float)level);
}

// sample_mip_bilinear (:147-159)
template <int C>
__device__ __forceinline__ F3 mip_bilinear(const TexView& tv, const GpuTexture& t, V2 uv, V2 dx, V2 dy) {
  if (!finite2(dx) || !finite2(dy)) return bilerp<C>(tv, t, 0, uv);
  const float m = fmaxf(fmaxf(fmaxf(fabsf(dx.x), fabsf(dy.x)), fmaxf(fabsf(dx.y), fabsf(dy.y))), kMinFilterWidth);
  const float lod = fmaxf((float)(t.levels - 1u) + log2f(m), 0.0f);
  const uint32_t level = lod >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)lod;
  return bilerp<C>(tv, t, level < t.levels - 1u ? level : t.levels - 1u, uv);
}

// sample_nearest (:134-142)
template <int C>
__device__ __forceinline__ F3 nearest(const TexView& tv, const GpuTexture& t, V2 uv) {
  const GpuTexLevel& L = tv.levels[t.first_level];
  const float u = wrap_unit(uv.x), v = wrap_unit(uv.y);
  const int x = (int)floorf(u * (float)L.width), y = (int)floorf(v * (float)L.height);
  return pixel<C>(tv, L, x, y);
}

// ---- adjoints of the lookups ------------------------------------------------------
// Each scatter adds g times d(lookup)/d(texel) to a gradient array laid out
// like the texels (same level offsets, RGB texels 4 floats apart), following
// the lookup's own branch decisions; global float atomics (texels are shared
// by many lookups).
template <int C, typename T = float>
__device__ __forceinline__ void scatter_texel(const GpuTexLevel& L, uint32_t x, uint32_t y, F3 g, float w,
                                              T* __restrict__ gbuf) {
  const size_t i = (size_t)y * L.width + x;
  if (C == 3) {
    T* a = gbuf + L.offset + 4 * i;
    atomicAdd(a, (T)(g.x * w));
    atomicAdd(a + 1, (T)(g.y * w));
    atomicAdd(a + 2, (T)(g.z * w));
  } else {
    atomicAdd(gbuf + L.offset + i, (T)(g.x * w));
  }
}
// bilerp_level's adjoint: the four texels with the bilinear weights
template <int C, typename T = float>
__device__ __forceinline__ void scatter_bilerp(const TexView& tv, const GpuTexture& t, uint32_t level, V2 uv, F3 g,
                                               T* __restrict__ gbuf) {
  const GpuTexLevel& L = level_of(tv, t, level);
  const float u = wrap_unit(uv.x), v = wrap_unit(uv.y);
  const float x = u * (float)L.width - 0.5f, y = v * (float)L.height - 0.5f;
  const int x0 = (int)floorf(x), y0 = (int)floorf(y);
  const float tx = x - (float)x0, ty = y - (float)y0;
  const bool px2 = L.pow2 & 1u, py2 = (L.pow2 >> 1) & 1u;
  const uint32_t xa = wrap_index(x0, L.width, px2), xb = wrap_index(x0 + 1, L.width, px2);
  const uint32_t ya = wrap_index(y0, L.height, py2), yb = wrap_index(y0 + 1, L.height, py2);
  scatter_texel<C, T>(L, xa, ya, g, (1.0f - tx) * (1.0f - ty), gbuf);
  scatter_texel<C, T>(L, xb, ya, g, tx * (1.0f - ty), gbuf);
  scatter_texel<C, T>(L, xa, yb, g, (1.0f - tx) * ty, gbuf);
  scatter_texel<C, T>(L, xb, yb, g, tx * ty, gbuf);
}
// ewa_level's adjoint: the same footprint, weights w / weight_sum (a first
// pass sums the weights, a second scatters)
template <int C, typename T = float>
__device__ __noinline__ void scatter_ewa_level(const TexView& tv, const GpuTexture& t, uint32_t level, V2 uv, V2 d0,
                                               V2 d1, F3 g, T* __restrict__ gbuf) {
  if (level >= t.levels) {
    scatter_texel<C, T>(tv.levels[t.first_level + t.levels - 1u], 0, 0, g, 1.0f, gbuf);
    return;
  }
  const GpuTexLevel& L = tv.levels[t.first_level + level];
  const float W = (float)L.width, H = (float)L.height;
  const V2 st = v2(uv.x * W - 0.5f, uv.y * H - 0.5f);
  d0 = v2(d0.x * W, d0.y * H);
  d1 = v2(d1.x * W, d1.y * H);
  float a = d0.y * d0.y + d1.y * d1.y + 1.0f;
  float b = -2.0f * (d0.x * d0.y + d1.x * d1.y);
  float c = d0.x * d0.x + d1.x * d1.x + 1.0f;
  const float inv_f = 1.0f / (a * c - 0.25f * b * b);
  if (!__builtin_isfinite(inv_f)) return scatter_bilerp<C, T>(tv, t, level, uv, g, gbuf);
  a *= inv_f;
  b *= inv_f;
  c *= inv_f;
  const float det = -b * b + 4.0f * a * c;
  if (det <= 0.0f || !__builtin_isfinite(det)) return scatter_bilerp<C, T>(tv, t, level, uv, g, gbuf);
  const float inv_det = 1.0f / det;
  const float u_sqrt = sqrtf(fmaxf(det * c, 0.0f)), v_sqrt = sqrtf(fmaxf(a * det, 0.0f));
  const float fs0 = ceilf(st.x - 2.0f * inv_det * u_sqrt), fs1 = floorf(st.x + 2.0f * inv_det * u_sqrt);
  const float ft0 = ceilf(st.y - 2.0f * inv_det * v_sqrt), ft1 = floorf(st.y + 2.0f * inv_det * v_sqrt);
  const float lim = 1073741824.0f;
  if (!(fs1 - fs0 <= (float)kMaxEwaSpan) || !(ft1 - ft0 <= (float)kMaxEwaSpan) || !(fabsf(fs0) < lim) ||
      !(fabsf(fs1) < lim) || !(fabsf(ft0) < lim) || !(fabsf(ft1) < lim))
    return scatter_bilerp<C, T>(tv, t, level, uv, g, gbuf);
  const int s0 = (int)fs0, s1 = (int)fs1, t0 = (int)ft0, t1 = (int)ft1;
  const float cutoff = expf(-kEwaAlpha);
  float wsum = 0.0f;
  for (int it = t0; it <= t1; ++it) {
    const float tt = (float)it - st.y;
    for (int is = s0; is <= s1; ++is) {
      const float ss = (float)is - st.x;
      const float r2 = a * ss * ss + b * ss * tt + c * tt * tt;
      if (!(r2 >= 1.0f)) wsum += expf(-kEwaAlpha * r2) - cutoff;
    }
  }
  if (!(wsum > 0.0f)) return scatter_bilerp<C, T>(tv, t, level, uv, g, gbuf);
  const float inv = 1.0f / wsum;
  const bool px2 = L.pow2 & 1u, py2 = (L.pow2 >> 1) & 1u;
  uint32_t y = wrap_index(t0, L.height, py2);
  for (int it = t0; it <= t1; ++it) {
    const float tt = (float)it - st.y;
    uint32_t x = wrap_index(s0, L.width, px2);
    for (int is = s0; is <= s1; ++is) {
      const float ss = (float)is - st.x;
      const float r2 = a * ss * ss + b * ss * tt + c * tt * tt;
      if (!(r2 >= 1.0f)) scatter_texel<C, T>(L, x, y, g, (expf(-kEwaAlpha * r2) - cutoff) * inv, gbuf);
      x = x + 1u == L.width ? 0u : x + 1u;
    }
    y = y + 1u == L.height ? 0u : y + 1u;
  }
}
// sample_filtered's adjoint: levels level and level + 1 with weights 1 - t and t
template <int C, typename T = float>
__device__ __noinline__ void scatter_filtered(const TexView& tv, const GpuTexture& t, V2 uv, V2 d0, V2 d1, F3 g,
                                              T* __restrict__ gbuf) {
  if (!finite2(d0) || !finite2(d1)) return scatter_bilerp<C, T>(tv, t, 0, uv, g, gbuf);
  if (len2(d0) < len2(d1)) {
    const V2 k = d0;
    d0 = d1;
    d1 = k;
  }
  const float longer = sqrtf(len2(d0));
  float shorter = sqrtf(len2(d1));
  if (shorter > 0.0f && shorter * kMaxAnisotropy < longer) {
    const float scale = longer / (shorter * kMaxAnisotropy);
    d1 = v2(d1.x * scale, d1.y * scale);
    shorter *= scale;
  }
  if (shorter <= 0.0f) return scatter_bilerp<C, T>(tv, t, 0, uv, g, gbuf);
  const float lod = fmaxf((float)(t.levels - 1u) + log2f(fmaxf(shorter, kMinFilterWidth)), 0.0f);
  const float fl = floorf(lod);
  const uint32_t level = fl >= 4294967295.0f ? 0xFFFFFFFFu : (uint32_t)fl;
  const float tl = lod - fl;
  // lerp(a, b, t) = a + (b - a) t: d/da = 1 - t, d/db = t
  scatter_ewa_level<C, T>(tv, t, level, uv, d0, d1, f3(g.x * (1.0f - tl), g.y * (1.0f - tl), g.z * (1.0f - tl)),
                          gbuf);
  if (tl != 0.0f)
    scatter_ewa_level<C, T>(tv, t, level == 0xFFFFFFFFu ? level : level + 1u, uv, d0, d1,
                            f3(g.x * tl, g.y * tl, g.z * tl), gbuf);
}

// Texture::sample_filtered of texture `ti` (RGB, or the scalar in .x)
__device__ __forceinline__ F3 sample_filtered(const TexView& tv, uint32_t ti, V2 uv, V2 dx, V2 dy) {
  const GpuTexture& t = tv.tex[ti];
  return t.channels == 3 ? filtered<3>(tv, t, uv, dx, dy) : filtered<1>(tv, t, uv, dx, dy);
}

}  // namespace tex
}  // namespace tpt
