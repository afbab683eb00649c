// Device sampler and pinhole camera: the reference's ZSobol sampler
// (R/sampler/zsobol.rs), PCG hash (R/sampler/pcg.rs), AuxRng
// (R/sampler/aux_rng.rs), PathSampler dimension layout
// (R/sampler/path_sampler.rs:61-158) and PinholeCamera ray generation
// (R/scene/camera.rs:48-96).  Integer code is exact; the camera uses only
// + - * / sqrt (correctly rounded with -fhip-fp32-correctly-rounded-divide-sqrt
// and -ffp-contract=off), so every value matches the reference bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tpt {
namespace smp {

__host__ __device__ __forceinline__ uint32_t pcg_step(uint32_t s) { return s * 747796405u + 2891336453u; }
__host__ __device__ __forceinline__ uint32_t pcg_out(uint32_t s) {
  s ^= s >> (4u + (s >> 28));
  s *= 277803737u;
  return s ^ (s >> 22);
}
__host__ __device__ __forceinline__ uint32_t pcg_hash(uint32_t k) { return pcg_out(pcg_step(k)); }

__device__ __forceinline__ uint32_t brev(uint32_t v) { return __builtin_bitreverse32(v); }

// Laine-Karras style permutation of the reversed index bits.
__device__ __forceinline__ uint32_t lk_permute(uint32_t v, uint32_t seed) {
  v ^= v * 0x3d20adeau;
  v += seed;
  v *= (seed >> 16) | 1u;
  v ^= v * 0x05526c56u;
  v ^= v * 0x53a22864u;
  return v;
}

__device__ __forceinline__ float u32_to_unit(uint32_t v) {
  // clears the low byte when the top byte is all ones: max = 1 - 2^-24
  return (float)(v & ~(v >> 24)) * (1.0f / 4294967296.0f);
}

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t v) {
  v ^= v >> 31;
  v *= 0x7fb5d329728ea185ull;
  v ^= v >> 27;
  v *= 0x81dadef4bc2dd44dull;
  v ^= v >> 33;
  return v;
}

__host__ __device__ __forceinline__ uint64_t spread_bits(uint32_t x) {
  uint64_t v = x;
  v = (v | (v << 16)) & 0x0000ffff0000ffffull;
  v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
  v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}
__host__ __device__ __forceinline__ uint64_t morton2(uint32_t x, uint32_t y) {
  return (spread_bits(y) << 1) | spread_bits(x);
}

// Columns of the generator matrices of Sobol dimensions 1..3 (dimension 0 is
// the bit reversal); bit i of the 16-bit index selects column i.
__constant__ static const uint16_t kSobolCols[3][16] = {
    {0xffff, 0x5555, 0x3333, 0x1111, 0x0f0f, 0x0505, 0x0303, 0x0101, 0x00ff, 0x0055, 0x0033, 0x0011, 0x000f, 0x0005,
     0x0003, 0x0001},
    {0xaa09, 0x7706, 0x3903, 0x1601, 0x09aa, 0x0677, 0x0339, 0x0116, 0x00a3, 0x0071, 0x003a, 0x0017, 0x0009, 0x0006,
     0x0003, 0x0001},
    {0xa0c3, 0x4041, 0x302d, 0x101e, 0x0b67, 0x079a, 0x02a4, 0x011b, 0x00c9, 0x0045, 0x002e, 0x001f, 0x000a, 0x0004,
     0x0003, 0x0001},
};

__device__ __forceinline__ uint32_t sobol16(uint32_t index16, int dim) {
  if (dim == 0) return brev(index16) >> 16;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < 16; ++i) s ^= ((index16 >> i) & 1u) ? (uint32_t)kSobolCols[dim - 1][i] : 0u;
  return s;
}

// Scrambled Sobol point, `depth` dimensions (<= 2 used here).
__device__ __forceinline__ void sobol_point(uint32_t index, uint32_t seed, int depth, uint32_t* out) {
  const uint32_t idx = lk_permute(brev(index), seed) >> 16;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    if (d >= depth) break;
    const uint32_t r = (8u * (uint32_t)d) & 31u;
    const uint32_t dseed = r ? (seed >> r) | (seed << (32u - r)) : seed;
    out[d] = brev(lk_permute(sobol16(idx, d), dseed));
  }
}

// (h >> 24) % 24 in 32-bit arithmetic: h >> 24 = xh * 2^32 + xl with
// xh < 256, and 2^32 = 16 (mod 24).
__host__ __device__ __forceinline__ uint32_t top40_mod24(uint64_t h) {
  const uint32_t xl = (uint32_t)(h >> 24), xh = (uint32_t)(h >> 56);
  return (xh * 16u + xl % 24u) % 24u;
}

// Per-frame constants of the ZSobol index: log2 of the pow2-rounded spp and
// the number of base-4 digits of the Morton index.
struct ZSobolFrame {
  uint32_t log2_spp;
  uint32_t n_digits;
};

__host__ __device__ inline uint32_t ilog2(uint32_t v) {
  uint32_t r = 0;
  while (v >>= 1) ++r;
  return r;
}
__host__ __device__ inline uint32_t pow2_ceil(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}
__host__ __device__ inline ZSobolFrame zsobol_frame(uint32_t spp, uint32_t width, uint32_t height) {
  ZSobolFrame f;
  f.log2_spp = ilog2(pow2_ceil(spp < 1 ? 1 : spp));
  const uint32_t res = pow2_ceil(width > height ? width : height);
  f.n_digits = ilog2(res) + (f.log2_spp + 1) / 2;
  return f;
}

// The 24 permutations of four base-4 digits, in the reference's table order
// (R/sampler/zsobol.rs), packed 8 bits per permutation (the image of digit d
// in bits 2d..2d+1), permutations 0-7, 8-15 and 16-23 in three 64-bit
// words: perm_byte(p) is a select and a shift, not a per-lane table load
// inside the digit loop.
__host__ __device__ __forceinline__ uint32_t perm_byte(uint32_t p) {
  const uint64_t w = p < 8u ? 0xb1e19c6c78d8b4e4ull : (p < 16u ? 0x72d236c68d2d39c9ull : 0x93634b1b87271e4eull);
  return (uint32_t)(w >> ((p & 7u) << 3)) & 0xffu;
}

// Permutation byte of one digit: hash of the higher digits (value `higher`)
// and the dimension hash dh.
__host__ __device__ __forceinline__ uint32_t digit_perm_byte(uint64_t higher, uint64_t dh) {
  return perm_byte(top40_mod24(mix64(higher ^ dh)));
}

// Digit-permutation rows.  For the top kZTabDigits base-4 digits of any
// ZSobol index the higher digits form a value below kZTabPrefixes, so their
// permutations depend on (dimension, value) only: row `dim` holds
// digit_perm_byte(v, 0x55555555 * dim) for every v < kZTabPrefixes, and a
// lookup replaces a 64-bit hash for those digits.
constexpr uint32_t kZTabDigits = 6;
constexpr uint32_t kZTabPrefixes = 1u << (2 * (kZTabDigits - 1));

// Sample index of (pixel, sample, dimension): Morton-ordered pixel + sample,
// each base-4 digit permuted by a hash of the higher digits and the dimension.
// row: the dimension's digit-permutation row, or null (every digit hashed).
__host__ __device__ __forceinline__ uint64_t zsobol_index(const ZSobolFrame& f, uint32_t px, uint32_t py, uint32_t si,
                                                          uint32_t dim, const uint8_t* row = nullptr) {
  const uint64_t m = (morton2(px, py) << f.log2_spp) | (uint64_t)si;
  const bool odd = (f.log2_spp & 1u) != 0;
  const uint32_t sub = odd ? 1u : 0u;
  const uint64_t dh = 0x55555555ull * (uint64_t)dim;
  const int last = odd ? 1 : 0;
  const int top = (int)f.n_digits - 1;
  int i = top;
  uint64_t out = 0;
  if (row) {
    // m < 2^(2 n_digits - sub), so digit i's higher digits are below 4^(top - i)
    if (top - (int)kZTabDigits + 1 >= last) {
      // the usual case, unrolled: the row reads issue together
      uint32_t b[kZTabDigits];
#pragma unroll
      for (int j = 0; j < (int)kZTabDigits; ++j) b[j] = row[(uint32_t)(m >> (2u * (uint32_t)(top - j) - sub + 2u))];
#pragma unroll
      for (int j = 0; j < (int)kZTabDigits; ++j) {
        const uint32_t sh = 2u * (uint32_t)(top - j) - sub;
        out |= (uint64_t)((b[j] >> (2u * (uint32_t)((m >> sh) & 3u))) & 3u) << sh;
      }
      i = top - (int)kZTabDigits;
    } else {
      for (; i >= last; --i) {
        const uint32_t sh = 2u * (uint32_t)i - sub;
        const uint32_t digit = (uint32_t)((m >> sh) & 3u);
        const uint32_t b = row[(uint32_t)(m >> (sh + 2))];
        out |= (uint64_t)((b >> (2u * digit)) & 3u) << sh;
      }
    }
  }
  for (; i >= last; --i) {
    const uint32_t sh = 2u * (uint32_t)i - sub;
    const uint32_t digit = (uint32_t)((m >> sh) & 3u);
    const uint32_t b = digit_perm_byte(m >> (sh + 2), dh);
    out |= (uint64_t)((b >> (2u * digit)) & 3u) << sh;
  }
  if (odd) out |= (m & 1u) ^ (mix64((m >> 1) ^ dh) & 1u);
  return out;
}

__host__ __device__ __forceinline__ uint32_t fold64(uint64_t v) { return (uint32_t)v ^ (uint32_t)(v >> 32); }

// PathSampler: one pixel sample's view of the sequence (scramble seed 0).
// row: the dimension's digit-permutation row (null: hash every digit).
struct PathSampler {
  ZSobolFrame f;
  uint32_t px, py, si;
  __device__ __forceinline__ uint32_t index(uint32_t dim, const uint8_t* row = nullptr) const {
    return fold64(zsobol_index(f, px, py, si, dim, row));
  }
  __device__ __forceinline__ static uint32_t seed(uint32_t dim) { return fold64(mix64((uint64_t)dim)); }
  __device__ __forceinline__ float s1(uint32_t dim, const uint8_t* row = nullptr) const {
    uint32_t v[2];
    sobol_point(index(dim, row), seed(dim), 1, v);
    return u32_to_unit(v[0]);
  }
  __device__ __forceinline__ void s2(uint32_t dim, float* o, const uint8_t* row = nullptr) const {
    uint32_t v[2];
    sobol_point(index(dim, row), seed(dim), 2, v);
    o[0] = u32_to_unit(v[0]);
    o[1] = u32_to_unit(v[1]);
  }
  __device__ __forceinline__ uint32_t aux_seed(uint32_t dim) const {
    const uint64_t idx = zsobol_index(f, px, py, si, dim);
    return pcg_hash(fold64(mix64(idx ^ ((uint64_t)dim << 32))));
  }
};

// Dimension layout of R/sampler/path_sampler.rs:61-78.
constexpr uint32_t kCameraDim = 0;
constexpr uint32_t kInitialAuxSeedDim = 0x80000000u;
constexpr uint32_t kVertexBase = 2;
constexpr uint32_t kVertexStride = 16;

// AuxRng: PCG stream with 24-bit float output.
struct AuxRng {
  uint32_t state;
  __device__ __forceinline__ explicit AuxRng(uint32_t seed) : state(pcg_step(seed)) {}
  __device__ __forceinline__ float next() {
    state = pcg_step(state);
    return (float)(pcg_out(state) >> 8) * (1.0f / 16777216.0f);
  }
};

// sqrtf lowers to the correctly rounded sequence (v_sqrt_f32 plus a residual
// fix-up); __fsqrt_rn lowers to the bare 1-ulp v_sqrt_f32 in this toolchain
// and must not be used where bit parity is required.
__device__ __forceinline__ float tent(float u) {
  u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
  if (u < 0.5f) return sqrtf(2.0f * u) - 1.0f;
  return 1.0f - sqrtf(2.0f - 2.0f * u);
}

}  // namespace smp
}  // namespace tpt
