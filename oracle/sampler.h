// Oracle sampler and camera: ZSobol + scrambled Sobol (R/sampler/zsobol.rs),
// PCG (R/sampler/pcg.rs), AuxRng (R/sampler/aux_rng.rs), PathSampler
// (R/sampler/path_sampler.rs), tent filter (R/math/mod.rs:13-24), pinhole
// camera helpers (R/scene/camera.rs).  TEST INFRASTRUCTURE ONLY.
#pragma once

#include <cmath>
#include <cstdint>

namespace orc {

// ---- R/sampler/pcg.rs:1-13
inline uint32_t pcg_state_transition(uint32_t s) { return s * 747796405u + 2891336453u; }
inline uint32_t pcg_output(uint32_t s) {
  s ^= s >> (4u + (s >> 28));
  s *= 277803737u;
  return s ^ (s >> 22);
}
inline uint32_t pcg_hash(uint32_t k) { return pcg_output(pcg_state_transition(k)); }

// ---- R/sampler/zsobol.rs
inline uint32_t reverse32(uint32_t v) {
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i)
    if (v & (1u << i)) r |= 1u << (31 - i);
  return r;
}
inline uint16_t reverse16(uint16_t v) {
  uint16_t r = 0;
  for (int i = 0; i < 16; ++i)
    if (v & (1u << i)) r |= (uint16_t)(1u << (15 - i));
  return r;
}
inline uint32_t laine_karras(uint32_t v, uint32_t seed) {  // :9-16
  v ^= v * 0x3d20adeau;
  v += seed;
  v *= (seed >> 16) | 1u;
  v ^= v * 0x05526c56u;
  v ^= v * 0x53a22864u;
  return v;
}
inline uint32_t reverse_and_shuffle(uint32_t v, uint32_t seed) { return laine_karras(reverse32(v), seed); }
inline uint32_t rotate_bytes(uint32_t v, uint32_t dist) {
  uint32_t r = (dist * 8u) & 31u;
  return r ? (v >> r) | (v << (32u - r)) : v;
}
inline uint32_t scramble_and_reverse(uint32_t v, uint32_t seed) { return reverse32(laine_karras(v, seed)); }
inline float uint_to_float(uint32_t v) {  // :35-39
  uint32_t mask = v >> 24;
  uint32_t safe = v & ~mask;
  return (float)safe * (1.0f / 4294967296.0f);
}
inline uint64_t mix_bits(uint64_t v) {  // :41-48
  v ^= v >> 31;
  v *= 0x7fb5d329728ea185ull;
  v ^= v >> 27;
  v *= 0x81dadef4bc2dd44dull;
  v ^= v >> 33;
  return v;
}
inline uint64_t left_shift_2(uint64_t v) {  // :50-58
  v &= 0xffffffffull;
  v = (v ^ (v << 16)) & 0x0000ffff0000ffffull;
  v = (v ^ (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v ^ (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v ^ (v << 2)) & 0x3333333333333333ull;
  v = (v ^ (v << 1)) & 0x5555555555555555ull;
  return v;
}
inline uint64_t encode_morton_2(uint32_t x, uint32_t y) { return (left_shift_2(y) << 1) | left_shift_2(x); }
inline uint32_t log2_int(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }
inline uint32_t round_up_pow2(uint32_t v) {
  if (v < 1) v = 1;
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Sobol generator matrices of the first four dimensions (columns as 16-bit
// words, bit i of the index selects column i), R/sampler/zsobol.rs:87-165.
const uint16_t kSobol[4][16] = {
    {0x8000, 0x4000, 0x2000, 0x1000, 0x0800, 0x0400, 0x0200, 0x0100, 0x0080, 0x0040, 0x0020, 0x0010, 0x0008, 0x0004,
     0x0002, 0x0001},
    {0xffff, 0x5555, 0x3333, 0x1111, 0x0f0f, 0x0505, 0x0303, 0x0101, 0x00ff, 0x0055, 0x0033, 0x0011, 0x000f, 0x0005,
     0x0003, 0x0001},
    {0xaa09, 0x7706, 0x3903, 0x1601, 0x09aa, 0x0677, 0x0339, 0x0116, 0x00a3, 0x0071, 0x003a, 0x0017, 0x0009, 0x0006,
     0x0003, 0x0001},
    {0xa0c3, 0x4041, 0x302d, 0x101e, 0x0b67, 0x079a, 0x02a4, 0x011b, 0x00c9, 0x0045, 0x002e, 0x001f, 0x000a, 0x0004,
     0x0003, 0x0001},
};

inline uint16_t sobol_reversed_index(uint16_t index, int dim) {  // :71-177
  if (dim == 0) return reverse16(index);
  uint16_t s = 0;
  for (int i = 0; i < 16; ++i)
    if (index & (1u << i)) s ^= kSobol[dim][i];
  return s;
}

void shuffled_scrambled_sobol(uint32_t index, uint32_t seed, int depth, uint32_t* out) {  // :179-188
  uint32_t idx = reverse_and_shuffle(index, seed);
  for (int d = 0; d < depth; ++d) {
    uint16_t sob = sobol_reversed_index((uint16_t)(idx >> 16), d);
    out[d] = scramble_and_reverse((uint32_t)sob, rotate_bytes(seed, (uint32_t)d));
  }
}

const uint8_t kPerm[24][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 1, 3}, {0, 2, 3, 1}, {0, 3, 2, 1}, {0, 3, 1, 2},
                              {1, 0, 2, 3}, {1, 0, 3, 2}, {1, 2, 0, 3}, {1, 2, 3, 0}, {1, 3, 2, 0}, {1, 3, 0, 2},
                              {2, 1, 0, 3}, {2, 1, 3, 0}, {2, 0, 1, 3}, {2, 0, 3, 1}, {2, 3, 0, 1}, {2, 3, 1, 0},
                              {3, 1, 2, 0}, {3, 1, 0, 2}, {3, 2, 1, 0}, {3, 2, 0, 1}, {3, 0, 2, 1}, {3, 0, 1, 2}};

struct IndexSampler {  // :203-286
  uint32_t log2_spp, n_digits;
  IndexSampler(uint32_t spp, uint32_t rx, uint32_t ry) {
    log2_spp = log2_int(round_up_pow2(spp));
    uint32_t res = round_up_pow2(rx > ry ? rx : ry);
    uint32_t log4_spp = (log2_spp + 1) / 2;
    n_digits = log2_int(res) + log4_spp;
  }
  uint64_t sample_index(uint32_t px, uint32_t py, uint32_t si, uint32_t dim) const {
    uint64_t morton = (encode_morton_2(px, py) << log2_spp) | (uint64_t)si;
    uint64_t out = 0;
    bool pow2 = (log2_spp & 1) != 0;
    uint32_t last = pow2 ? 1 : 0;
    uint64_t dh = 0x55555555ull * (uint64_t)dim;
    for (int i = (int)n_digits - 1; i >= (int)last; --i) {
      uint32_t shift = 2u * (uint32_t)i - (pow2 ? 1u : 0u);
      uint32_t digit = (uint32_t)((morton >> shift) & 3u);
      uint64_t higher = morton >> (shift + 2);
      uint32_t p = (uint32_t)((mix_bits(higher ^ dh) >> 24) % 24);
      out |= (uint64_t)kPerm[p][digit] << shift;
    }
    if (pow2) {
      uint64_t digit = morton & 1u;
      out |= digit ^ (mix_bits((morton >> 1) ^ dh) & 1u);
    }
    return out;
  }
};

inline uint32_t fold(uint64_t v) { return (uint32_t)v ^ (uint32_t)(v >> 32); }

struct PathSampler {  // R/sampler/path_sampler.rs:55-158
  IndexSampler is;
  uint32_t px, py, si;
  static constexpr uint32_t kScrambleSeed = 0;
  PathSampler(uint32_t x, uint32_t y, uint32_t s, uint32_t spp, uint32_t rx, uint32_t ry)
      : is(spp, rx, ry), px(x), py(y), si(s) {}
  uint32_t index_for(uint32_t dim) const { return fold(is.sample_index(px, py, si, dim)); }
  uint32_t seed_for(uint32_t dim) const { return fold(mix_bits(((uint64_t)kScrambleSeed << 32) | dim)); }
  float s1(uint32_t dim) const {
    uint32_t v;
    shuffled_scrambled_sobol(index_for(dim), seed_for(dim), 1, &v);
    return uint_to_float(v);
  }
  void s2(uint32_t dim, float* o) const {
    uint32_t v[2];
    shuffled_scrambled_sobol(index_for(dim), seed_for(dim), 2, v);
    o[0] = uint_to_float(v[0]);
    o[1] = uint_to_float(v[1]);
  }
  uint32_t aux_seed(uint32_t dim) const {
    uint64_t idx = is.sample_index(px, py, si, dim);
    uint64_t bits = mix_bits(idx ^ ((uint64_t)dim << 32) ^ (uint64_t)kScrambleSeed);
    return pcg_hash(fold(bits));
  }
};

// sample_tent_1d, R/math/mod.rs:17-24
inline float tent(float u) {
  u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
  if (u < 0.5f) return std::sqrt(2.0f * u) - 1.0f;
  return 1.0f - std::sqrt(2.0f - 2.0f * u);
}

struct CF3 {
  float x, y, z;
};
inline CF3 add(CF3 a, CF3 b) { return CF3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline CF3 sub(CF3 a, CF3 b) { return CF3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline CF3 scale(float s, CF3 a) { return CF3{s * a.x, s * a.y, s * a.z}; }
inline float dot(CF3 a, CF3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline CF3 normalize(CF3 a) {  // glam: self * (1 / sqrt(dot))
  float r = 1.0f / std::sqrt(dot(a, a));
  return CF3{a.x * r, a.y * r, a.z * r};
}
inline CF3 normalize_or_zero(CF3 a) {
  float r = 1.0f / std::sqrt(dot(a, a));
  if (std::isfinite(r) && r > 0.0f) return CF3{a.x * r, a.y * r, a.z * r};
  return CF3{0.f, 0.f, 0.f};
}

}  // namespace orc
