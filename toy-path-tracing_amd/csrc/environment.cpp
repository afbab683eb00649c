// Host build of the environment light's sampling data
// (EnvironmentLight::from_pixels, R/light/environment.rs:47-83): texel
// weights luminance * sin(theta) and their MIS-compensated variant (minus the
// solid-angle mean luminance, R/light/environment.rs:362-408), each summed
// into a power-of-two padded pyramid (2x2 / 2x1 / 1x2 box sums up to a single
// total, :426-478).  The float sums follow the reference's loop order.
#include <cmath>
#include <string>
#include <vector>

#include "environment.h"

namespace tpt {
namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;

inline float lum_pos(const float* c) { return std::fmax(0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2], 0.0f); }
inline float row_sin(uint32_t j, uint32_t h) { return std::sin(((float)j + 0.5f) / (float)h * kPi); }
inline uint32_t pow2_at_least(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Pyramid of one weight image: levels coarsest (1x1) first, concatenated.
void pyramid(uint32_t w, uint32_t h, const std::vector<float>& wt, std::vector<float>* out,
             std::vector<uint32_t>* table, float* total) {
  const uint32_t pw = pow2_at_least(w), ph = pow2_at_least(h);
  std::vector<std::vector<float>> lv;
  std::vector<std::pair<uint32_t, uint32_t>> dims;
  std::vector<float> leaf((size_t)pw * ph, 0.0f);
  for (uint32_t j = 0; j < h; ++j)
    for (uint32_t i = 0; i < w; ++i) leaf[(size_t)j * pw + i] = std::fmax(wt[(size_t)j * w + i], 0.0f);
  lv.push_back(std::move(leaf));
  dims.emplace_back(pw, ph);
  while (dims.back().first > 1 || dims.back().second > 1) {
    const uint32_t fw = dims.back().first, fh = dims.back().second;
    const uint32_t si = fw > 1 ? 2 : 1, sj = fh > 1 ? 2 : 1;
    const uint32_t cw = fw / si, ch = fh / sj;
    std::vector<float> c((size_t)cw * ch, 0.0f);
    const std::vector<float>& f = lv.back();
    for (uint32_t j = 0; j < ch; ++j)
      for (uint32_t i = 0; i < cw; ++i) {
        float sum = 0.0f;
        for (uint32_t dj = 0; dj < sj; ++dj)
          for (uint32_t di = 0; di < si; ++di) sum += f[(size_t)(j * sj + dj) * fw + i * si + di];
        c[(size_t)j * cw + i] = sum;
      }
    lv.push_back(std::move(c));
    dims.emplace_back(cw, ch);
  }
  out->clear();
  table->clear();
  for (size_t k = lv.size(); k-- > 0;) {
    table->push_back((uint32_t)out->size());
    table->push_back(dims[k].first);
    table->push_back(dims[k].second);
    out->insert(out->end(), lv[k].begin(), lv[k].end());
  }
  *total = (*out)[0];
}

}  // namespace

bool build_environment(const float* px, uint32_t w, uint32_t h, float scale, float rotate_y, HostEnv* e,
                       std::string* err) {
  if (!px || w == 0 || h == 0) {
    *err = "environment map must be non-empty";
    return false;
  }
  e->w = w;
  e->h = h;
  e->scale = scale;
  e->px.assign(px, px + 3 * (size_t)w * h);
  const size_t n = (size_t)w * h;
  std::vector<float> plain(n), comp(n);
  for (uint32_t j = 0; j < h; ++j) {
    const float st = row_sin(j, h);
    for (uint32_t i = 0; i < w; ++i) plain[(size_t)j * w + i] = lum_pos(&px[3 * ((size_t)j * w + i)]) * st;
  }
  float weighted = 0.0f, weights = 0.0f;
  for (uint32_t j = 0; j < h; ++j) {
    const float st = row_sin(j, h);
    weights += st * (float)w;
    for (uint32_t i = 0; i < w; ++i) weighted += lum_pos(&px[3 * ((size_t)j * w + i)]) * st;
  }
  const float mean = weights > 0.0f ? weighted / weights : 0.0f;
  const float eps = std::fmax(std::fabs(mean), 1.0f) * 1.0e-6f;
  for (uint32_t j = 0; j < h; ++j) {
    const float st = row_sin(j, h);
    for (uint32_t i = 0; i < w; ++i) {
      const float c = lum_pos(&px[3 * ((size_t)j * w + i)]) - mean;
      comp[(size_t)j * w + i] = c > eps ? c * st : 0.0f;
    }
  }
  std::vector<float> plain_pyr;
  std::vector<uint32_t> plain_table;
  pyramid(w, h, plain, &plain_pyr, &plain_table, &e->total_plain);
  pyramid(w, h, comp, &e->pyr, &e->levels, &e->total_mis);
  e->level_count = (uint32_t)(e->levels.size() / 3);
  const float s = std::sin(rotate_y), c = std::cos(rotate_y);  // Mat3::from_rotation_y
  const float r[9] = {c, 0.0f, -s, 0.0f, 1.0f, 0.0f, s, 0.0f, c};
  const float si = std::sin(-rotate_y), ci = std::cos(-rotate_y);
  const float ri[9] = {ci, 0.0f, -si, 0.0f, 1.0f, 0.0f, si, 0.0f, ci};
  for (int k = 0; k < 9; ++k) {
    e->rot[k] = r[k];
    e->inv_rot[k] = ri[k];
  }
  // total_power (:129-132): light-sampler category weight
  const float texel = kTau * kPi / ((float)w * (float)h);
  e->power = e->total_plain * scale * texel;
  return true;
}

}  // namespace tpt
