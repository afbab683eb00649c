// Oracle: CPU restatement of the reference sampler and pinhole camera.
//
// TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
//   R/sampler/zsobol.rs        ZSobol index permutation + scrambled 4-D Sobol
//   R/sampler/pcg.rs           PCG hash
//   R/sampler/aux_rng.rs       AuxRng
//   R/sampler/path_sampler.rs  PathSampler dimension layout
//   R/scene/camera.rs          PinholeCamera::generate_ray_differential
//   R/math/mod.rs:13-24        sample_tent_1d / 2d
// Integer code is exact; the camera uses only +,-,*,/,sqrt (correctly
// rounded), so it is exact too.  The transcendental set-up values
// (tan(fov/2), atan of the pixel footprint) are inputs computed by the caller.
#include <cmath>
#include <cstdint>

namespace {

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

struct F3 {
  float x, y, z;
};
inline F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 scale(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 normalize(F3 a) {  // glam: self * (1 / sqrt(dot))
  float r = 1.0f / std::sqrt(dot(a, a));
  return {a.x * r, a.y * r, a.z * r};
}
inline F3 normalize_or_zero(F3 a) {
  float r = 1.0f / std::sqrt(dot(a, a));
  if (std::isfinite(r) && r > 0.0f) return {a.x * r, a.y * r, a.z * r};
  return {0.f, 0.f, 0.f};
}

}  // namespace

extern "C" {

uint32_t orc_pcg_hash(uint32_t k) { return pcg_hash(k); }
uint32_t orc_pcg_state_transition(uint32_t s) { return pcg_state_transition(s); }
uint32_t orc_pcg_output(uint32_t s) { return pcg_output(s); }
uint32_t orc_laine_karras(uint32_t v, uint32_t seed) { return laine_karras(v, seed); }
uint32_t orc_reverse_and_shuffle(uint32_t v, uint32_t seed) { return reverse_and_shuffle(v, seed); }
float orc_uint_to_float(uint32_t v) { return uint_to_float(v); }
uint64_t orc_encode_morton_2(uint32_t x, uint32_t y) { return encode_morton_2(x, y); }
uint64_t orc_mix_bits(uint64_t v) { return mix_bits(v); }
void orc_shuffled_scrambled_sobol(uint32_t index, uint32_t seed, int depth, uint32_t* out) {
  shuffled_scrambled_sobol(index, seed, depth, out);
}
uint64_t orc_zsobol_sample_index(uint32_t spp, uint32_t rx, uint32_t ry, uint32_t px, uint32_t py, uint32_t si,
                                 uint32_t dim) {
  return IndexSampler(spp, rx, ry).sample_index(px, py, si, dim);
}

// IndexSampler::sample_index for every pixel of (rx, ry) and every sample
// index < spp in one dimension; out[(y*rx + x)*spp + s].
void orc_zsobol_sample_indices(uint32_t spp, uint32_t rx, uint32_t ry, uint32_t dim, uint64_t* out) {
  IndexSampler is(spp, rx, ry);
  for (uint32_t y = 0; y < ry; ++y)
    for (uint32_t x = 0; x < rx; ++x)
      for (uint32_t s = 0; s < spp; ++s) out[((size_t)y * rx + x) * spp + s] = is.sample_index(x, y, s, dim);
}

// AuxRng (R/sampler/aux_rng.rs:14-30): n floats from a seed.
void orc_aux_rng(uint32_t seed, int n, float* out) {
  uint32_t s = pcg_state_transition(seed);
  for (int i = 0; i < n; ++i) {
    s = pcg_state_transition(s);
    out[i] = (float)(pcg_output(s) >> 8) * (1.0f / 16777216.0f);
  }
}

// PathSampler values for one (pixel, sample): out[0..2) camera sample,
// out_seed[0] initial aux seed; per depth d < ndepth, out[2 + 14*d + k] =
// u_category, u_tree, u_light_aux, u_surface.xy, u_lobe, u_layer, u_dir.xy,
// u_extra0..3, u_rr and out_seed[1 + d] = aux_rng_seed.
void orc_path_sampler(uint32_t rx, uint32_t ry, uint32_t spp, uint32_t px, uint32_t py, uint32_t si, int ndepth,
                      float* out, uint32_t* out_seed) {
  PathSampler ps(px, py, si, spp, rx, ry);
  ps.s2(0, out);                       // CAMERA_SAMPLE_DIM
  out_seed[0] = ps.aux_seed(0x80000000u);  // INITIAL_AUX_RNG_SEED_DIM
  for (int d = 0; d < ndepth; ++d) {
    uint32_t base = 2u + (uint32_t)d * 16u;
    float* o = out + 2 + 14 * d;
    o[0] = ps.s1(base + 0);
    o[1] = ps.s1(base + 1);
    o[2] = ps.s1(base + 2);
    ps.s2(base + 3, o + 3);
    o[5] = ps.s1(base + 5);
    o[6] = ps.s1(base + 6);
    ps.s2(base + 7, o + 7);
    o[9] = ps.s1(base + 9);
    o[10] = ps.s1(base + 10);
    o[11] = ps.s1(base + 11);
    o[12] = ps.s1(base + 12);
    o[13] = ps.s1(base + 13);
    out_seed[1 + d] = ps.aux_seed(base + 14);
  }
}

// PinholeCamera::generate_ray_differential (R/scene/camera.rs:48-78) for every
// pixel of a (rx, ry) image and one sample index.  cam: eye[3], forward[3],
// right[3], image_up[3], tan_half_fov_y, spread (= pixel_spread_angle, the
// atan of the footprint).  out per pixel (row-major): origin[3], dir[3],
// rx_dir[3], ry_dir[3], cone_spread (13 floats).
void orc_camera_rays(const float* cam, uint32_t rx, uint32_t ry, uint32_t spp, uint32_t si, float* out) {
  F3 eye{cam[0], cam[1], cam[2]}, fwd{cam[3], cam[4], cam[5]}, right{cam[6], cam[7], cam[8]},
      up{cam[9], cam[10], cam[11]};
  const float th = cam[12], spread = cam[13];
  const float rxf = (float)rx, ryf = (float)ry;
  const float aspect = rxf / ryf;
  auto dir_for = [&](float px, float py) {
    float sx = ((2.0f * px / rxf) - 1.0f) * aspect * th;
    float sy = (1.0f - (2.0f * py / ryf)) * th;
    return normalize(add(add(fwd, scale(sx, right)), scale(sy, up)));
  };
  const float sc = std::fmax(1.0f / std::sqrt((float)(spp < 1 ? 1 : spp)), 0.125f);  // differential_scale
  for (uint32_t y = 0; y < ry; ++y)
    for (uint32_t x = 0; x < rx; ++x) {
      PathSampler ps(x, y, si, spp, rx, ry);
      float u[2];
      ps.s2(0, u);
      const float px = ((float)x + 0.5f) + 0.5f * tent(u[0]);
      const float py = ((float)y + 0.5f) + 0.5f * tent(u[1]);
      F3 d = dir_for(px, py);
      F3 dxd = dir_for(px + 1.0f, py);
      F3 dyd = dir_for(px, py + 1.0f);
      F3 rxd = normalize_or_zero(add(d, scale(sc, sub(dxd, d))));
      F3 ryd = normalize_or_zero(add(d, scale(sc, sub(dyd, d))));
      float* o = out + 13 * ((size_t)y * rx + x);
      o[0] = eye.x; o[1] = eye.y; o[2] = eye.z;
      o[3] = d.x; o[4] = d.y; o[5] = d.z;
      o[6] = rxd.x; o[7] = rxd.y; o[8] = rxd.z;
      o[9] = ryd.x; o[10] = ryd.y; o[11] = ryd.z;
      o[12] = std::fmax(spread * sc, 0.0f);  // RayCone::new(0, spread * scale)
    }
}

}  // extern "C"
