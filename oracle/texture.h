// Oracle: the reference's filtered, mip-mapped Texture restated on the CPU.
// TEST INFRASTRUCTURE ONLY.
//
//   R/material/texture.rs:90-115   from_pixels, max_value
//   R/material/texture.rs:130-142  sample (bilinear level 0), sample_nearest
//   R/material/texture.rs:147-159  sample_mip_bilinear
//   R/material/texture.rs:160-193  sample_filtered = sample_ewa
//   R/material/texture.rs:200-223  sample_trilinear
//   R/material/texture.rs:225-243  bilerp_level
//   R/material/texture.rs:245-300  ewa_level
//   R/material/texture.rs:302-308  pixel_level_wrapped
//   R/material/texture.rs:433-473  build_mip_levels, pixel_wrapped_in_level
//   R/material/texture.rs:575-589  wrap_unit, wrap_index
//
// A pixel is 3 floats (Texture<Vec3>) or 1 (ScalarTexture); Vec3 arithmetic
// is per channel in f32 and TexturePixel::lerp is a + (b - a) * t (glam 0.32).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace orc {
namespace texture {

constexpr real kMaxAnisotropy = 8.0f;      // :86
constexpr real kMinFilterWidth = 1.0e-8f;  // :87
constexpr real kEwaAlpha = 2.0f;           // :88

struct Level {
  size_t width = 0, height = 0;
  std::vector<real> px;  // width * height * channels
};
struct Texture {
  int channels = 3;
  std::vector<Level> levels;
  real max_value = 0.0f;
};
struct Px {  // one pixel value (channels used: 1 or 3)
  real c[3] = {0.0f, 0.0f, 0.0f};
};

inline Px px_add(const Px& a, const Px& b) { return Px{{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
inline Px px_scale(const Px& a, real s) { return Px{{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
inline Px px_lerp(const Px& a, const Px& b, real t) {
  Px r;
  for (int k = 0; k < 3; ++k) r.c[k] = a.c[k] + (b.c[k] - a.c[k]) * t;
  return r;
}

// wrap_index (:583-589): mask for a power-of-two size ((index as usize) &
// (size - 1), two's complement), else rem_euclid.
inline size_t wrap_index(int64_t index, size_t size) {
  if (size && !(size & (size - 1))) return (size_t)((uint64_t)index & (uint64_t)(size - 1));
  int64_t m = index % (int64_t)size;
  if (m < 0) m += (int64_t)size;
  return (size_t)m;
}
// wrap_unit (:575-581): rem_euclid(1.0) of a finite value (f32 % is fmod), else 0.
inline real wrap_unit(real t) {
  if (!std::isfinite(t)) return 0.0f;
  real r = std::fmod(t, 1.0f);
  return r < 0.0f ? r + 1.0f : r;
}
// Rust's `f32 as isize`: truncation, saturating, NaN -> 0.
inline int64_t as_isize(real f) {
  if (std::isnan(f)) return 0;
  if (f >= 9.2233720368547758e18f) return INT64_MAX;
  if (f <= -9.2233720368547758e18f) return INT64_MIN;
  return (int64_t)f;
}
// Rust's `f32 as usize`: truncation, saturating, NaN and negatives -> 0.
inline uint64_t as_usize(real f) {
  if (std::isnan(f) || f <= 0.0f) return 0;
  if (f >= 1.8446744073709552e19f) return UINT64_MAX;
  return (uint64_t)f;
}

inline Px level_pixel(const Texture& t, const Level& l, int64_t x, int64_t y) {  // pixel_wrapped_in_level
  const size_t xi = wrap_index(x, l.width), yi = wrap_index(y, l.height);
  Px p;
  for (int k = 0; k < t.channels; ++k) p.c[k] = l.px[(yi * l.width + xi) * t.channels + k];
  return p;
}
inline size_t level_count(const Texture& t) { return t.levels.size(); }
inline Px pixel_level_wrapped(const Texture& t, size_t level, int64_t x, int64_t y) {  // :302-308
  return level_pixel(t, t.levels[std::min(level, level_count(t) - 1)], x, y);
}

// Texture::from_pixels (:90-115) with build_mip_levels (:433-466).
inline Texture from_pixels(size_t width, size_t height, int channels, const float* pixels) {
  Texture t;
  t.channels = channels;
  Level base;
  base.width = width;
  base.height = height;
  base.px.assign(pixels, pixels + width * height * channels);
  real mx = 0.0f;
  for (size_t i = 0; i < width * height; ++i) {
    real v = pixels[i * channels];
    if (channels == 3) v = std::fmax(pixels[3 * i], std::fmax(pixels[3 * i + 1], pixels[3 * i + 2]));
    mx = std::fmax(mx, v);
  }
  t.max_value = mx;
  t.levels.push_back(std::move(base));
  while (t.levels.back().width > 1 || t.levels.back().height > 1) {
    const Level& prev = t.levels.back();
    Level next;
    next.width = std::max<size_t>((prev.width + 1) / 2, 1);
    next.height = std::max<size_t>((prev.height + 1) / 2, 1);
    next.px.resize(next.width * next.height * channels);
    for (size_t y = 0; y < next.height; ++y)
      for (size_t x = 0; x < next.width; ++x) {
        const int64_t sx = 2 * (int64_t)x, sy = 2 * (int64_t)y;
        Px s = px_add(px_add(px_add(level_pixel(t, prev, sx, sy), level_pixel(t, prev, sx + 1, sy)),
                             level_pixel(t, prev, sx, sy + 1)),
                      level_pixel(t, prev, sx + 1, sy + 1));
        s = px_scale(s, 0.25f);
        for (int k = 0; k < channels; ++k) next.px[(y * next.width + x) * channels + k] = s.c[k];
      }
    t.levels.push_back(std::move(next));
  }
  return t;
}

// bilerp_level (:225-243)
inline Px bilerp_level(const Texture& t, size_t level, real u_in, real v_in) {
  level = std::min(level, level_count(t) - 1);
  const Level& l = t.levels[level];
  const real u = wrap_unit(u_in), v = wrap_unit(v_in);
  const real x = u * (real)l.width - 0.5f;
  const real y = v * (real)l.height - 0.5f;
  const int64_t x0 = as_isize(std::floor(x)), y0 = as_isize(std::floor(y));
  const real tx = x - (real)x0, ty = y - (real)y0;
  const Px c00 = level_pixel(t, l, x0, y0), c10 = level_pixel(t, l, x0 + 1, y0);
  const Px c01 = level_pixel(t, l, x0, y0 + 1), c11 = level_pixel(t, l, x0 + 1, y0 + 1);
  return px_lerp(px_lerp(c00, c10, tx), px_lerp(c01, c11, tx), ty);
}
inline Px sample(const Texture& t, real u, real v) { return bilerp_level(t, 0, u, v); }  // :130-132

// sample_nearest (:134-142)
inline Px sample_nearest(const Texture& t, real u_in, real v_in) {
  const Level& l = t.levels[0];
  const real u = wrap_unit(u_in), v = wrap_unit(v_in);
  return pixel_level_wrapped(t, 0, as_isize(std::floor(u * (real)l.width)), as_isize(std::floor(v * (real)l.height)));
}

// ewa_level (:245-300)
inline Px ewa_level(const Texture& t, size_t level, real u, real v, real d0x, real d0y, real d1x, real d1y) {
  if (level >= level_count(t)) return pixel_level_wrapped(t, level_count(t) - 1, 0, 0);
  const Level& l = t.levels[level];
  const real W = (real)l.width, H = (real)l.height;
  const real stx = u * W - 0.5f, sty = v * H - 0.5f;
  d0x *= W;
  d0y *= H;
  d1x *= W;
  d1y *= H;
  real a = d0y * d0y + d1y * d1y + 1.0f;
  real b = -2.0f * (d0x * d0y + d1x * d1y);
  real c = d0x * d0x + d1x * d1x + 1.0f;
  const real inv_f = 1.0f / (a * c - 0.25f * b * b);
  if (!std::isfinite(inv_f)) return bilerp_level(t, level, u, v);
  a *= inv_f;
  b *= inv_f;
  c *= inv_f;
  const real det = -b * b + 4.0f * a * c;
  if (det <= 0.0f || !std::isfinite(det)) return bilerp_level(t, level, u, v);
  const real inv_det = 1.0f / det;
  const real u_sqrt = std::sqrt(std::fmax(det * c, 0.0f));
  const real v_sqrt = std::sqrt(std::fmax(a * det, 0.0f));
  const int64_t s0 = as_isize(std::ceil(stx - 2.0f * inv_det * u_sqrt));
  const int64_t s1 = as_isize(std::floor(stx + 2.0f * inv_det * u_sqrt));
  const int64_t t0 = as_isize(std::ceil(sty - 2.0f * inv_det * v_sqrt));
  const int64_t t1 = as_isize(std::floor(sty + 2.0f * inv_det * v_sqrt));
  Px sum;
  real weight_sum = 0.0f;
  const real cutoff = std::exp(-kEwaAlpha);
  for (int64_t it = t0; it <= t1; ++it) {
    const real tt = (real)it - sty;
    for (int64_t is = s0; is <= s1; ++is) {
      const real ss = (real)is - stx;
      const real r2 = a * ss * ss + b * ss * tt + c * tt * tt;
      if (r2 >= 1.0f) continue;
      const real weight = std::exp(-kEwaAlpha * r2) - cutoff;
      sum = px_add(sum, px_scale(pixel_level_wrapped(t, level, is, it), weight));
      weight_sum += weight;
    }
  }
  if (weight_sum > 0.0f) return px_scale(sum, 1.0f / weight_sum);
  return bilerp_level(t, level, u, v);
}

// sample_ewa (:164-193), what sample_filtered calls
inline Px sample_filtered(const Texture& t, real u, real v, real d0x, real d0y, real d1x, real d1y) {
  if (!std::isfinite(d0x) || !std::isfinite(d0y) || !std::isfinite(d1x) || !std::isfinite(d1y)) return sample(t, u, v);
  if (d0x * d0x + d0y * d0y < d1x * d1x + d1y * d1y) {
    std::swap(d0x, d1x);
    std::swap(d0y, d1y);
  }
  const real longer = std::sqrt(d0x * d0x + d0y * d0y);
  real shorter = std::sqrt(d1x * d1x + d1y * d1y);
  if (shorter > 0.0f && shorter * kMaxAnisotropy < longer) {
    const real scale = longer / (shorter * kMaxAnisotropy);
    d1x *= scale;
    d1y *= scale;
    shorter *= scale;
  }
  if (shorter <= 0.0f) return sample(t, u, v);
  const real lod = std::fmax((real)(level_count(t) - 1) + std::log2(std::fmax(shorter, kMinFilterWidth)), 0.0f);
  const uint64_t level = as_usize(std::floor(lod));
  const real tl = lod - (real)level;
  const Px a = ewa_level(t, level, u, v, d0x, d0y, d1x, d1y);
  const Px b = ewa_level(t, level + 1, u, v, d0x, d0y, d1x, d1y);
  return px_lerp(a, b, tl);
}

// sample_trilinear (:200-223)
inline Px sample_trilinear(const Texture& t, real u, real v, real dxx, real dxy, real dyx, real dyy) {
  if (!std::isfinite(dxx) || !std::isfinite(dxy) || !std::isfinite(dyx) || !std::isfinite(dyy)) return sample(t, u, v);
  const real mx = std::fmax(std::fabs(dxx), std::fabs(dyx)), my = std::fmax(std::fabs(dxy), std::fabs(dyy));
  const real width = 2.0f * std::fmax(std::fmax(mx, my), kMinFilterWidth);
  const real top = (real)(level_count(t) - 1);
  const real lod = top + std::log2(width);
  if (lod >= top) return pixel_level_wrapped(t, level_count(t) - 1, 0, 0);
  const size_t level = (size_t)as_usize(std::fmax(std::floor(lod), 0.0f));
  if (level == 0) return bilerp_level(t, 0, u, v);
  return px_lerp(bilerp_level(t, level, u, v), bilerp_level(t, level + 1, u, v), lod - (real)level);
}

// sample_mip_bilinear (:147-159)
inline Px sample_mip_bilinear(const Texture& t, real u, real v, real dxx, real dxy, real dyx, real dyy) {
  if (!std::isfinite(dxx) || !std::isfinite(dxy) || !std::isfinite(dyx) || !std::isfinite(dyy))
    return bilerp_level(t, 0, u, v);
  const real mx = std::fmax(std::fabs(dxx), std::fabs(dyx)), my = std::fmax(std::fabs(dxy), std::fabs(dyy));
  const real width = std::fmax(std::fmax(mx, my), kMinFilterWidth);
  const real lod = std::fmax((real)(level_count(t) - 1) + std::log2(width), 0.0f);
  const size_t level = (size_t)std::min<uint64_t>(as_usize(lod), level_count(t) - 1);
  return bilerp_level(t, level, u, v);
}

}  // namespace texture
}  // namespace orc
