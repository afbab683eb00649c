// Host builds of the GGX tables (ggx_tables.h).  The estimators follow the
// reference's (R/bsdf/directional_albedo.rs:135-168, 292-341, 700-829,
// 832-1023) sample for sample and in the same summation order, with the
// BSDF pieces they call: DielectricGgxBsdf reflection eval and rough
// sampling (R/bsdf/dielectric_ggx.rs:121-176, 501-587), GGX D / Lambda / G
// and VNDF sampling (R/bsdf/smith_ggx.rs), dielectric Fresnel
// (R/math/mod.rs:50-71).  Cells are independent, so they are built on
// several threads.
#include "ggx_tables.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace tpt {
namespace ggx {
namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;
constexpr float kMinAlpha = 1.0e-4f;
constexpr float kSmoothAlpha = 1.0e-3f;

struct V {
  float x, y, z;
};
inline V v3(float x, float y, float z) { return V{x, y, z}; }
inline V add(V a, V b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V scale(V a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return v3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
inline V normalize_or_zero(V a) {
  const float r = 1.0f / std::sqrt(dot(a, a));
  if (std::isfinite(r) && r > 0.0f) return scale(a, r);
  return v3(0.0f, 0.0f, 0.0f);
}

float fresnel_dielectric(float ci, float ei, float et) {
  ci = clampf(ci, -1.0f, 1.0f);
  if (ci < 0.0f) {
    std::swap(ei, et);
    ci = -ci;
  }
  const float si = std::sqrt(std::fmax(1.0f - ci * ci, 0.0f));
  const float st = (ei / et) * si;
  if (st >= 1.0f) return 1.0f;
  const float ct = std::sqrt(std::fmax(1.0f - st * st, 0.0f));
  const float rpar = (et * ci - ei * ct) / (et * ci + ei * ct);
  const float rper = (ei * ci - et * ct) / (ei * ci + et * ct);
  return 0.5f * (rpar * rpar + rper * rper);
}
bool half_vector(V wo, V wi, V* wm) {
  const V m = normalize_or_zero(add(wo, wi));
  if (dot(m, m) == 0.0f) return false;
  *wm = m.z < 0.0f ? v3(-m.x, -m.y, -m.z) : m;
  return true;
}
float ggx_d(V m, float ax, float ay) {
  if (m.z <= 0.0f) return 0.0f;
  const float t = m.x * m.x / (ax * ax) + m.y * m.y / (ay * ay) + m.z * m.z;
  const float den = kPi * ax * ay * t * t;
  return den <= 0.0f ? 0.0f : 1.0f / den;
}
float ggx_lambda(V w, float ax, float ay) {
  const float c = std::fabs(w.z);
  if (c <= 0.0f) return INFINITY;
  const float t = 1.0f + (ax * ax * w.x * w.x + ay * ay * w.y * w.y) / (c * c);
  return 0.5f * (-1.0f + std::sqrt(t));
}
float ggx_g1(V w, float ax, float ay) {
  const float l = ggx_lambda(w, ax, ay);
  return std::isfinite(l) ? 1.0f / (1.0f + l) : 0.0f;
}
float ggx_g2(V wo, V wi, float ax, float ay) {
  const float lo = ggx_lambda(wo, ax, ay), li = ggx_lambda(wi, ax, ay);
  if (!std::isfinite(lo) || !std::isfinite(li)) return 0.0f;
  return 1.0f / (1.0f + lo + li);
}
V reflect_local(V wo, V wm) { return normalize_or_zero(add(v3(-wo.x, -wo.y, -wo.z), scale(wm, 2.0f * dot(wo, wm)))); }
float pdf_wm_vndf(V wo, V wm, float ax, float ay) {
  if (!(wo.z > 0.0f) || wm.z <= 0.0f) return 0.0f;
  const float d = ggx_d(wm, ax, ay), g1 = ggx_g1(wo, ax, ay), c = std::fmax(dot(wo, wm), 0.0f);
  if (d <= 0.0f || g1 <= 0.0f || c <= 0.0f) return 0.0f;
  return d * g1 * c / wo.z;
}
bool sample_wm_vndf(V wo, float ax, float ay, float u0, float u1, V* out) {  // Heitz 2018
  if (wo.z <= 0.0f) return false;
  ax = std::fmax(ax, kMinAlpha);
  ay = std::fmax(ay, kMinAlpha);
  const V ws = normalize_or_zero(v3(ax * wo.x, ay * wo.y, wo.z));
  if (dot(ws, ws) == 0.0f) return false;
  const float l2 = ws.x * ws.x + ws.y * ws.y;
  V t1 = v3(1.0f, 0.0f, 0.0f);
  if (l2 > 0.0f) {
    const float l = std::sqrt(l2);
    t1 = v3(-ws.y / l, ws.x / l, 0.0f / l);
  }
  const V t2 = cross(ws, t1);
  const float r = std::sqrt(clampf(u0, 0.0f, 1.0f));
  const float phi = kTau * u1;
  const float p1 = r * std::cos(phi);
  float p2 = r * std::sin(phi);
  const float s = 0.5f * (1.0f + ws.z);
  p2 = (1.0f - s) * std::sqrt(std::fmax(1.0f - p1 * p1, 0.0f)) + s * p2;
  const V nh = add(add(scale(t1, p1), scale(t2, p2)), scale(ws, std::sqrt(std::fmax(1.0f - p1 * p1 - p2 * p2, 0.0f))));
  const V wm = normalize_or_zero(v3(ax * nh.x, ay * nh.y, std::fmax(nh.z, 0.0f)));
  if (dot(wm, wm) == 0.0f || wm.z <= 0.0f) return false;
  *out = wm;
  return true;
}
bool refract_about_wm(V wo, V wm, float eta, V* out) {
  const float co = dot(wo, wm);
  if (co <= 0.0f) return false;
  const float s2 = eta * eta * std::fmax(1.0f - co * co, 0.0f);
  if (s2 >= 1.0f) return false;
  const float ct = std::sqrt(std::fmax(1.0f - s2, 0.0f));
  const V wi = normalize_or_zero(add(scale(wo, -eta), scale(wm, eta * co - ct)));
  if (dot(wi, wi) == 0.0f) return false;
  *out = wi;
  return true;
}
float radical_inverse_vdc(uint32_t bits) {
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) r |= ((bits >> i) & 1u) << (31 - i);
  return (float)r * 2.3283064e-10f;
}
float radical_inverse_base3(uint32_t n) {
  float q = 0.0f, bk = 1.0f / 3.0f;
  while (n > 0) {
    q += (float)(n % 3u) * bk;
    n /= 3u;
    bk /= 3.0f;
  }
  return q;
}
inline float cell_center(uint32_t i, uint32_t res) { return ((float)i + 0.5f) / (float)res; }
inline float sanitize_eta(float eta) { return (std::isfinite(eta) && eta > 0.0f) ? eta : 1.5f; }

template <class F>
void parallel_for(uint32_t n, F f) {
  const unsigned th = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < th; ++t)
    pool.emplace_back([&, t] {
      for (uint32_t i = t; i < n; i += th) f(i);
    });
  for (auto& p : pool) p.join();
}

// reflection-only DielectricGgxBsdf eval, colour 1, front face, no compensation
float dielectric_reflect(float eta, float ax, float ay, V wo, V wi) {
  ax = std::fmax(ax, kMinAlpha);
  ay = std::fmax(ay, kMinAlpha);
  if (!(wo.z > 0.0f) || eta <= 0.0f) return 0.0f;
  if (std::fmax(ax, ay) < kSmoothAlpha) return 0.0f;
  if (!(wo.z * wi.z > 0.0f)) return 0.0f;
  V wm;
  if (!half_vector(wo, wi, &wm)) return 0.0f;
  const float cwm = dot(wo, wm);
  if (cwm <= 0.0f) return 0.0f;
  const float co = std::fabs(wo.z), ci = std::fabs(wi.z);
  if (co <= 0.0f || ci <= 0.0f) return 0.0f;
  const float d = ggx_d(wm, ax, ay);
  if (d <= 0.0f) return 0.0f;
  const float g = ggx_g2(wo, wi, ax, ay);
  if (g <= 0.0f) return 0.0f;
  return d * g * fresnel_dielectric(cwm, 1.0f, eta) / (4.0f * co * ci);
}

std::vector<float> build_albedo(float eta) {
  constexpr uint32_t kSamples = 64;
  V hs[kSamples];
  for (uint32_t i = 0; i < kSamples; ++i) {
    const float u = ((float)i + 0.5f) / (float)kSamples, w = radical_inverse_vdc(i);
    const float z = clampf(u, 0.0f, 1.0f), r = std::sqrt(std::fmax(1.0f - z * z, 0.0f)), phi = kTau * w;
    hs[i] = v3(r * std::cos(phi), r * std::sin(phi), z);
  }
  const float uniform_pdf = 1.0f / (2.0f * kPi);
  std::vector<float> out(kAlbedoLen, 0.0f);
  parallel_for(kAlbedoLen, [&](uint32_t idx) {
    const uint32_t an = idx % kAlbedoAniso;
    uint32_t rest = idx / kAlbedoAniso;
    const uint32_t ri = rest % kAlbedoRough;
    rest /= kAlbedoRough;
    const uint32_t pi = rest % kAlbedoPhi, ci = rest / kAlbedoPhi;
    const float sc = cell_center(ci, kAlbedoCos), ct = sc * sc;
    const float st = std::sqrt(std::fmax(1.0f - ct * ct, 0.0f));
    const float phi = cell_center(pi, kAlbedoPhi) * (kPi / 2.0f);
    const V wo = v3(st * std::cos(phi), st * std::sin(phi), ct);
    const float rough = clampf(cell_center(ri, kAlbedoRough), 0.0f, 1.0f);
    // slices 7 and 8 of the reference's 16 anisotropy cells: OpenPBR reads
    // the table at anisotropy 0, halfway between them
    const float aniso = clampf(2.0f * cell_center(an + 7, 16) - 1.0f, -1.0f, 1.0f);
    const float alpha = rough * rough, aspect = std::sqrt(1.0f - 0.9f * std::fabs(aniso));
    float ax = aniso >= 0.0f ? alpha / aspect : alpha * aspect;
    float ay = aniso >= 0.0f ? alpha * aspect : alpha / aspect;
    ax = clampf(ax, kMinAlpha, 1.0f);
    ay = clampf(ay, kMinAlpha, 1.0f);
    float val;
    if (wo.z <= 0.0f) {
      val = 0.0f;
    } else if (std::fmax(ax, ay) < kSmoothAlpha) {
      val = fresnel_dielectric(clampf(wo.z, 0.0f, 1.0f), 1.0f, eta);
    } else {
      float sum = 0.0f;
      for (uint32_t s = 0; s < kSamples; ++s)
        sum += dielectric_reflect(eta, ax, ay, wo, hs[s]) * std::fmax(hs[s].z, 0.0f) / uniform_pdf;
      val = clampf(sum / (float)kSamples, 0.0f, 1.0f);
    }
    out[idx] = val;
  });
  return out;
}

std::vector<float> build_conductor() {
  constexpr uint32_t kSamples = 256;
  float us[kSamples][2];
  for (uint32_t i = 0; i < kSamples; ++i) {
    us[i][0] = ((float)i + 0.5f) / (float)kSamples;
    us[i][1] = radical_inverse_vdc(i);
  }
  std::vector<float> out(kCondCos * kCondRough + kCondRough, 0.0f);
  parallel_for(kCondCos * kCondRough, [&](uint32_t idx) {
    const uint32_t ci = idx % kCondCos, ri = idx / kCondCos;
    const float sc = cell_center(ci, kCondCos), ct = sc * sc;
    const float rough = cell_center(ri, kCondRough);
    const float alpha = clampf(rough * rough, kMinAlpha, 1.0f);
    float val;
    if (ct <= 0.0f) {
      val = 0.0f;
    } else if (alpha < kSmoothAlpha) {
      val = 1.0f;
    } else {
      const float st = std::sqrt(std::fmax(1.0f - ct * ct, 0.0f));
      const V wo = v3(st, 0.0f, ct);
      const float g1 = ggx_g1(wo, alpha, alpha);
      if (g1 <= 0.0f) {
        val = 0.0f;
      } else {
        float sum = 0.0f;
        for (uint32_t s = 0; s < kSamples; ++s) {
          V wm;
          float term = 0.0f;
          if (sample_wm_vndf(wo, alpha, alpha, us[s][0], us[s][1], &wm) && dot(wo, wm) > 0.0f) {
            const V wi = reflect_local(wo, wm);
            if (wi.z > 0.0f) term = ggx_g2(wo, wi, alpha, alpha) / g1;
          }
          sum += term;
        }
        val = clampf(sum / (float)kSamples, 0.0f, 1.0f);
      }
    }
    out[idx] = val;
  });
  for (uint32_t r = 0; r < kCondRough; ++r) {
    float w = 0.0f;
    for (uint32_t c = 0; c < kCondCos; ++c) {
      const float sc = cell_center(c, kCondCos), ct = sc * sc;
      w += out[r * kCondCos + c] * ct * sc;
    }
    out[kCondCos * kCondRough + r] = clampf(4.0f / (float)kCondCos * w, 0.0f, 1.0f);
  }
  return out;
}

// DielectricGgxBsdf::sample_rough, front face, no compensation, radiance mode
bool dielectric_sample(float eta, float alpha, V wo, float uc, float u0, float u1, float* w, bool* trans) {
  const float eta_rel = 1.0f / eta;
  V wm;
  if (!sample_wm_vndf(wo, alpha, alpha, u0, u1, &wm)) return false;
  const float cwm = dot(wo, wm);
  if (cwm <= 0.0f) return false;
  const float f = fresnel_dielectric(cwm, 1.0f, eta);
  const float g1 = ggx_g1(wo, alpha, alpha);
  if (g1 <= 0.0f) return false;
  float pr = std::fmax(f, 0.0f), pt = std::fmax(1.0f - f, 0.0f);
  const float ps = pr + pt;
  if (ps <= 0.0f) return false;
  pr /= ps;
  pt /= ps;
  if (uc < pr) {
    const V wi = reflect_local(wo, wm);
    if (!(wi.z > 0.0f)) return false;
    const float g2 = ggx_g2(wo, wi, alpha, alpha), pw = pdf_wm_vndf(wo, wm, alpha, alpha);
    if (pw <= 0.0f) return false;
    if (pr * pw / (4.0f * cwm) <= 0.0f) return false;
    *w = f * g2 / (pr * g1);
    *trans = false;
    return true;
  }
  V wi;
  if (!refract_about_wm(wo, wm, eta_rel, &wi) || wi.z >= 0.0f) return false;
  const float cwi = dot(wi, wm);
  const float den = cwi + eta_rel * cwm;
  if (std::fabs(den) < 1.0e-6f) return false;
  const float g2 = ggx_g2(wo, wi, alpha, alpha), pw = pdf_wm_vndf(wo, wm, alpha, alpha);
  if (pw <= 0.0f) return false;
  if (pt * pw * std::fabs(cwi) / (den * den) <= 0.0f) return false;
  const float rs = 1.0f / (eta_rel * eta_rel);
  *w = 1.0f * (rs * (1.0f - f) * g2 / (pt * g1));
  *trans = true;
  return true;
}

float axis_to_eta(float t) {
  t = clampf(t, 0.0f, 1.0f);
  if (t >= 0.5f) return 1.0f + (t - 0.5f) * 2.0f * (3.0f - 1.0f);
  return 1.0f / (1.0f + (0.5f - t) * 2.0f * (3.0f - 1.0f));
}

std::vector<float> build_dielectric() {
  constexpr uint32_t kSamples = 4096;
  const uint32_t ne = kDielCos * kDielRough * kDielEta;
  std::vector<float> out(ne + kDielRough * kDielEta, 0.0f);
  parallel_for(ne, [&](uint32_t idx) {
    const uint32_t ci = idx % kDielCos;
    const uint32_t rest = idx / kDielCos;
    const uint32_t ri = rest % kDielRough, ei = rest / kDielRough;
    const float sc = cell_center(ci, kDielCos), ct = sc * sc;
    const float rough = cell_center(ri, kDielRough);
    const float eta = axis_to_eta(cell_center(ei, kDielEta));
    float val;
    if (ct <= 0.0f || eta <= 0.0f) {
      val = 0.0f;
    } else {
      const float alpha = clampf(rough * rough, kMinAlpha, 1.0f);
      if (alpha < kSmoothAlpha) {
        val = 1.0f;
      } else {
        const float st = std::sqrt(std::fmax(1.0f - ct * ct, 0.0f));
        const V wo = v3(st, 0.0f, ct);
        const float inv_eta_sq = 1.0f / (eta * eta);
        float sum = 0.0f;
        for (uint32_t s = 0; s < kSamples; ++s) {
          const float uc = ((float)s + 0.5f) / (float)kSamples;
          float w;
          bool tr;
          if (!dielectric_sample(eta, alpha, wo, uc, radical_inverse_vdc(s + 1), radical_inverse_base3(s + 1), &w,
                                 &tr))
            continue;
          if (!std::isfinite(w)) continue;
          sum += w * (tr ? inv_eta_sq : 1.0f);
        }
        val = clampf(sum / (float)kSamples, 0.0f, 1.0f);
      }
    }
    out[idx] = val;
  });
  for (uint32_t e = 0; e < kDielEta; ++e)
    for (uint32_t r = 0; r < kDielRough; ++r) {
      float w = 0.0f;
      for (uint32_t c = 0; c < kDielCos; ++c) {
        const float sc = cell_center(c, kDielCos), ct = sc * sc;
        w += out[(e * kDielRough + r) * kDielCos + c] * ct * sc;
      }
      out[ne + e * kDielRough + r] = clampf(4.0f / (float)kDielCos * w, 0.0f, 1.0f);
    }
  return out;
}

std::mutex& table_mutex() {
  static std::mutex m;
  return m;
}

}  // namespace

float modulated_eta(float eta, float w) {
  if (std::fabs(w - 1.0f) < 1.0e-6f) return eta;
  const float q = (eta - 1.0f) / (eta + 1.0f);
  const float f0 = q * q;
  const float sg = (eta - 1.0f) < 0.0f ? -1.0f : 1.0f;
  const float eps = std::sqrt(clampf(w * f0, 0.0f, 0.99999f)) * sg;
  return std::fmax((1.0f + eps) / std::fmax(1.0f - eps, 1.0e-6f), 1.0e-4f);
}

const std::vector<float>& albedo_table(float eta_in) {
  static std::map<uint32_t, std::unique_ptr<std::vector<float>>> cache;
  const float eta = sanitize_eta(eta_in);
  uint32_t key;
  std::memcpy(&key, &eta, 4);
  std::lock_guard<std::mutex> g(table_mutex());
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, std::make_unique<std::vector<float>>(build_albedo(eta))).first;
  return *it->second;
}
const std::vector<float>& conductor_table() {
  static const std::vector<float> t = build_conductor();
  return t;
}
const std::vector<float>& dielectric_table() {
  static const std::vector<float> t = build_dielectric();
  return t;
}

}  // namespace ggx
}  // namespace tpt
