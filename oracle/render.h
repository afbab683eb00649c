// Oracle: the unidirectional MIS path tracer of the reference for OpenPBR
// materials, restated line by line.  TEST INFRASTRUCTURE ONLY.
//
//   R/main.rs:98-126                     pixel loop, running mean, exposure
//   R/integrator/mis.rs:19-139, 212-257  trace_radiance, emission MIS
//   R/integrator/mod.rs:52-85            spawn_ray / offset_ray_origin
//   R/scene/mod.rs:426-544               shading vertex
//   R/math/orthonormal_basis.rs          ONB (Duff et al.)
//   R/material/open_pbr.rs:351-606       OpenPbrMaterial le / make_bsdf / sample / eval / pdf
//   R/material/mod.rs:33-79              shading-normal correction
//   R/bsdf/open_pbr.rs                   OpenPbrBsdf (layer weights, lobe choice, lobes)
//   R/bsdf/eon.rs                        EON diffuse
//   R/bsdf/smith_ggx.rs                  GGX D / Lambda / G2, bounded VNDF pdf
//
// Supported OpenPBR subset (the same one the GPU path supports): specular
// weight 0 (the modulated specular eta is then exactly 1, so f_avg = 0 and
// every LUT-dependent term is multiplied by zero: the reference's directional
// albedo and energy-compensation tables do not influence the result), no
// metal, coat, fuzz, transmission, thin film, thin walls, textures or normal
// maps; base weight / colour / diffuse roughness and emission are free.
// Light NEE under the reference policy samples no light-tree light for
// OpenPBR materials (R/material/open_pbr.rs:592-606 returns no light-tree
// query); the environment light, when present, is sampled for every
// material (its branch needs no query, R/light/mod.rs:289-331).
//
// The reference's NormalizedLambert, Emissive and Mirror materials
// (R/material/{normalized_lambert,emissive,mirror}.rs) are carried too:
// Lambert surfaces build a diffuse light-tree query, so next-event estimation
// through the SG light tree with the balance heuristic
// (R/integrator/mis.rs:141-234) is exercised on them.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "geometry.h"
#include "lights.h"
#include "sampler.h"

namespace orc {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTau = 6.28318530717958647692f;

// ---- glam-style helpers --------------------------------------------------
inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline V3 vmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 smul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 sdiv(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline float len_sq(V3 a) { return dot(a, a); }
inline V3 normalize3(V3 a) {  // v * (1 / |v|)
  float r = 1.0f / std::sqrt(dot(a, a));
  return v3(a.x * r, a.y * r, a.z * r);
}
inline V3 normalize_or_zero3(V3 a) {
  float r = 1.0f / std::sqrt(dot(a, a));
  if (std::isfinite(r) && r > 0.0f) return v3(a.x * r, a.y * r, a.z * r);
  return v3(0.f, 0.f, 0.f);
}
inline float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }  // f32::clamp
inline V3 face_forward(V3 n, V3 ref) { return dot(n, ref) < 0.0f ? neg(n) : n; }  // R/math/mod.rs:42-48
inline float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }  // R/math/sg.rs:314-316
inline V3 interp3(float b0, float b1, float b2, V3 a, V3 b, V3 c) {
  return add(add(mul(b0, a), mul(b1, b)), mul(b2, c));
}

// ---- OrthonormalBasis, R/math/orthonormal_basis.rs:11-71 -------------------
struct Onb {
  V3 t, b, n;
};
inline Onb onb_from_normal(V3 normal) {
  V3 n = len_sq(normal) > 0.0f ? normalize3(normal) : v3(0.f, 0.f, 1.f);
  float sign = std::copysign(1.0f, n.z);
  float a = -1.0f / (sign + n.z);
  float b = n.x * n.y * a;
  V3 t = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  V3 bt = v3(b, sign + n.y * n.y * a, -n.y);
  return Onb{t, bt, n};
}
inline Onb onb_from_normal_and_tangent(V3 normal, V3 hint) {
  V3 n = len_sq(normal) > 0.0f ? normalize3(normal) : v3(0.f, 0.f, 1.f);
  V3 t = normalize_or_zero3(sub(hint, mul(dot(hint, n), n)));
  if (len_sq(t) == 0.0f) return onb_from_normal(n);
  V3 b = normalize_or_zero3(cross(n, t));
  t = normalize_or_zero3(cross(b, n));
  return Onb{t, b, n};
}
inline V3 onb_to_world(const Onb& f, V3 l) { return normalize3(add(add(mul(l.x, f.t), mul(l.y, f.b)), mul(l.z, f.n))); }
inline V3 onb_to_local(const Onb& f, V3 w) { return v3(dot(w, f.t), dot(w, f.b), dot(w, f.n)); }

// ---- GGX, R/bsdf/smith_ggx.rs ------------------------------------------------
constexpr float kMinAlpha = 1.0e-4f;
constexpr float kSmoothAlpha = 1.0e-3f;
inline bool reflection_half_vector(V3 wo, V3 wi, V3* wm) {
  V3 m = normalize_or_zero3(add(wo, wi));
  if (len_sq(m) == 0.0f) return false;
  *wm = m.z < 0.0f ? neg(m) : m;
  return true;
}
inline float ggx_d(V3 m, float ax, float ay) {
  if (m.z <= 0.0f) return 0.0f;
  float term = m.x * m.x / (ax * ax) + m.y * m.y / (ay * ay) + m.z * m.z;
  float denom = kPi * ax * ay * term * term;
  return denom <= 0.0f ? 0.0f : 1.0f / denom;
}
inline float ggx_lambda(V3 w, float ax, float ay) {
  float c = std::fabs(w.z);
  if (c <= 0.0f) return INFINITY;
  float term = 1.0f + (ax * ax * w.x * w.x + ay * ay * w.y * w.y) / (c * c);
  return 0.5f * (-1.0f + std::sqrt(term));
}
inline float ggx_g2(V3 wo, V3 wi, float ax, float ay) {
  float lo = ggx_lambda(wo, ax, ay), li = ggx_lambda(wi, ax, ay);
  if (!std::isfinite(lo) || !std::isfinite(li)) return 0.0f;
  return 1.0f / (1.0f + lo + li);
}
inline float pdf_wm_bounded_vndf(V3 wo, V3 wm, float ax, float ay) {
  if (wo.z <= 0.0f || wm.z <= 0.0f) return 0.0f;
  ax = std::fmax(ax, kMinAlpha);
  ay = std::fmax(ay, kMinAlpha);
  float a = clampf(std::fmin(ax, ay), 0.0f, 1.0f);
  float s = 1.0f + std::sqrt(wo.x * wo.x + wo.y * wo.y);
  float a2 = a * a, s2 = s * s;
  float k = (1.0f - a2) * s2 / (s2 + a2 * wo.z * wo.z);
  float t = std::sqrt(ax * ax * wo.x * wo.x + ay * ay * wo.y * wo.y + wo.z * wo.z);
  float denom = k * wo.z + t;
  if (denom <= 0.0f) return 0.0f;
  return 2.0f * ggx_d(wm, ax, ay) * std::fmax(dot(wo, wm), 0.0f) / denom;
}

// ---- fresnel, R/math/mod.rs:50-72 -----------------------------------------------
inline float fresnel_dielectric(float ci, float ei, float et) {
  ci = clampf(ci, -1.0f, 1.0f);
  if (ci < 0.0f) {
    std::swap(ei, et);
    ci = -ci;
  }
  float si = std::sqrt(std::fmax(1.0f - ci * ci, 0.0f));
  float st = (ei / et) * si;
  if (st >= 1.0f) return 1.0f;
  float ct = std::sqrt(std::fmax(1.0f - st * st, 0.0f));
  float rpar = (et * ci - ei * ct) / (et * ci + ei * ct);
  float rper = (ei * ci - et * ct) / (ei * ci + et * ct);
  return 0.5f * (rpar * rpar + rper * rper);
}
// weighted_fresnel_ratio, R/bsdf/open_pbr.rs:2161-2172
inline float weighted_fresnel_ratio(float c, float eti, float etp) {
  eti = std::fmax(eti, 1.0e-4f);
  etp = std::fmax(etp, 1.0e-4f);
  if (etp >= 1.0f) return fresnel_dielectric(c, 1.0f, etp);
  float cc = clampf(c, 0.0f, 1.0f);
  float mu2 = 1.0f - (1.0f - cc * cc) / (eti * eti);
  if (mu2 <= 0.0f) return 1.0f;
  return fresnel_dielectric(std::sqrt(mu2), 1.0f, 1.0f / etp);
}
// f_avg_dielectric, R/bsdf/open_pbr.rs:2127-2135
inline float f_avg_dielectric(float eta) {
  if (eta >= 1.0f) return clampf((eta - 1.0f) / (4.08567f + 1.00071f * eta), 0.0f, 1.0f);
  float e = std::fmax(eta, 1.0e-6f);
  return clampf(0.997118f + 0.1014f * e - 0.965241f * e * e - 0.130607f * e * e * e, 0.0f, 1.0f);
}

// ---- EON, R/bsdf/eon.rs:12-197 -----------------------------------------------
constexpr float kRcpPi = 1.0f / kPi;
const float kC1Fon = 0.5f - 2.0f / (3.0f * kPi);
const float kC2Fon = 2.0f / 3.0f - 28.0f / (15.0f * kPi);
inline float e_fon_approx(float mu, float r) {
  float mc = 1.0f - mu;
  float g = mc * (0.05710853f + mc * (0.49188188f + mc * (-0.33218145f + mc * 0.071443f)));
  return (1.0f + r * g) / (1.0f + kC1Fon * r);
}
inline V3 f_eon(V3 rho, float r, V3 wo, V3 wi) {
  float mo = wo.z, mi = wi.z;
  float s = dot(wo, wi) - mi * mo;
  float sot = s > 0.0f ? s / std::fmax(mo, mi) : s;
  float af = 1.0f / (1.0f + kC1Fon * r);
  V3 fss = mul(kRcpPi * af * (1.0f + r * sot), rho);
  float efo = e_fon_approx(mo, r), efi = e_fon_approx(mi, r);
  float avg = af * (1.0f + kC2Fon * r);
  float oma = 1.0f - avg;
  V3 denom = sub(v3(1.f, 1.f, 1.f), smul(rho, oma));
  V3 rho_ms = v3(rho.x * rho.x * avg / denom.x, rho.y * rho.y * avg / denom.y, rho.z * rho.z * avg / denom.z);
  float scale = kRcpPi * std::fmax(1.0f - efo, 1.0e-7f) * std::fmax(1.0f - efi, 1.0e-7f) / std::fmax(oma, 1.0e-7f);
  return add(fss, smul(rho_ms, scale));
}
inline void ltc_coeffs(float mu, float r, float* a, float* b, float* c, float* d) {
  *a = 1.0f + r * (0.303392f + (-0.518982f + 0.111709f * mu) * mu + (-0.276266f + 0.335918f * mu) * r);
  *b = r * (-1.16407f + 1.15859f * mu + (0.150815f - 0.150105f * mu) * r) / (mu * mu * mu - 1.43545f);
  *c = 1.0f + r * (0.20013f + (-0.506373f + 0.261777f * mu) * mu);
  *d = r * (0.540852f + (-1.01625f + 0.475392f * mu) * mu) / (-1.0743f + (0.0725628f + mu) * mu);
}
inline void ltc_basis_x(V3 wo, float* x, float* y) {
  float l2 = wo.x * wo.x + wo.y * wo.y;
  if (l2 > 0.0f) {
    float inv = 1.0f / std::sqrt(l2);
    *x = wo.x * inv;
    *y = wo.y * inv;
  } else {
    *x = 1.0f;
    *y = 0.0f;
  }
}
inline V3 ltc_from(V3 wo, V3 v) {
  float xx, xy;
  ltc_basis_x(wo, &xx, &xy);
  float yx = -xy, yy = xx;
  return v3(xx * v.x + yx * v.y, xy * v.x + yy * v.y, v.z);
}
inline V3 ltc_to(V3 wo, V3 v) {
  float xx, xy;
  ltc_basis_x(wo, &xx, &xy);
  float yx = -xy, yy = xx;
  return v3(xx * v.x + xy * v.y, yx * v.x + yy * v.y, v.z);
}
inline V3 cltc_sample(V3 wo, float r, float u1, float u2, float* pdf) {
  float a, b, c, d;
  ltc_coeffs(wo.z, r, &a, &b, &c, &d);
  float radius = std::sqrt(u1);
  float phi = 2.0f * kPi * u2;
  float x0 = radius * std::cos(phi), y0 = radius * std::sin(phi);
  float vz = 1.0f / std::sqrt(d * d + 1.0f);
  float s = 0.5f * (1.0f + vz);
  float edge = std::sqrt(std::fmax(1.0f - y0 * y0, 0.0f));
  float x = -((1.0f - s) * edge + s * x0);
  float y = y0;
  float whz = std::sqrt(std::fmax(1.0f - (x * x + y * y), 0.0f));
  V3 wh = v3(x, y, whz);
  float pdf_wh = wh.z / (kPi * s);
  V3 wl = v3(a * wh.x + b * wh.z, c * wh.y, d * wh.x + wh.z);
  float len = std::sqrt(dot(wl, wl));
  float det = c * (a - b * d);
  *pdf = pdf_wh * (len * len * len) / det;
  return ltc_from(wo, sdiv(wl, len));
}
inline float cltc_pdf(V3 wo, V3 wi, float r) {
  V3 wl = ltc_to(wo, wi);
  float a, b, c, d;
  ltc_coeffs(wo.z, r, &a, &b, &c, &d);
  float det = c * (a - b * d);
  V3 wh = v3(c * (wl.x - b * wl.z), (a - b * d) * wl.y, -c * (d * wl.x - a * wl.z));
  float l2 = std::fmax(dot(wh, wh), 1.0e-30f);
  float vz = 1.0f / std::sqrt(d * d + 1.0f);
  float s = 0.5f * (1.0f + vz);
  float q = det / l2;
  return q * q * std::fmax(wh.z, 0.0f) / (kPi * s);
}
inline float p_uniform(float mu, float r) {
  return std::pow(r, 0.1f) * (0.162925f + (-0.372058f + (0.538233f - 0.290822f * mu) * mu) * mu);
}
inline V3 uniform_lobe_sample(float u1, float u2) {
  float st = std::sqrt(std::fmax(1.0f - u1 * u1, 0.0f));
  float phi = 2.0f * kPi * u2;
  return v3(st * std::cos(phi), st * std::sin(phi), u1);
}
inline V3 sample_eon(V3 wo, float r, float u1, float u2, float* pdf) {
  float pu = p_uniform(wo.z, r);
  float pc = 1.0f - pu;
  V3 wi;
  float pdf_c;
  if (u1 < pu) {
    float u1n = u1 / pu;
    wi = uniform_lobe_sample(u1n, u2);
    pdf_c = cltc_pdf(wo, wi, r);
  } else {
    float den = std::fmax(pc, 1.0e-30f);
    float u1n = (u1 - pu) / den;
    wi = cltc_sample(wo, r, u1n, u2, &pdf_c);
  }
  *pdf = pu * (1.0f / (2.0f * kPi)) + pc * pdf_c;
  return wi;
}
inline float pdf_eon(V3 wo, V3 wi, float r) {
  float pu = p_uniform(wo.z, r);
  return pu * (1.0f / (2.0f * kPi)) + (1.0f - pu) * cltc_pdf(wo, wi, r);
}

// ---- materials -------------------------------------------------------------------
// OpenPbrMaterial parameters (R/material/open_pbr.rs:23-82), the fields the
// supported subset reads; unsupported ones must be zero.
struct OpenPbr {
  float base_weight;
  V3 base_color;
  float base_diffuse_roughness;
  float base_metalness;
  float specular_weight;
  V3 specular_color;
  float specular_roughness;
  float specular_ior;
  float specular_roughness_anisotropy;
  float transmission_weight, coat_weight, fuzz_weight, thin_film_weight;
  float emission_luminance;
  V3 emission_color;
  uint32_t thin_walled;
};

// OpenPbrBsdfParams of the subset (R/material/open_pbr.rs:393-487).
struct PbrBsdf {
  V3 base_color;
  float base;
  V3 specular_color;
  float ax, ay;
  float specular_eta, transmission_eta;
  float diffuse_roughness;
  bool front_face;
};

inline float modulated_eta(float eta, float w) {  // R/material/open_pbr.rs:871-878
  if (std::fabs(w - 1.0f) < 1.0e-6f) return eta;
  float f0 = std::pow((eta - 1.0f) / (eta + 1.0f), 2.0f);
  float sg = (eta - 1.0f) > 0.0f ? 1.0f : ((eta - 1.0f) < 0.0f ? -1.0f : 1.0f);  // f32::signum (+0 -> +1)
  float eps = std::sqrt(clampf(w * f0, 0.0f, 0.99999f)) * sg;
  return std::fmax((1.0f + eps) / std::fmax(1.0f - eps, 1.0e-6f), 1.0e-4f);
}
inline void alpha_xy(float roughness, float anisotropy, float* ax, float* ay) {  // :861-869
  float r = clampf(roughness, 0.0f, 1.0f);
  float a = clampf(anisotropy, 0.0f, 1.0f);
  float alpha = r * r;
  float oma = 1.0f - a;
  float x = alpha * std::sqrt(2.0f / (1.0f + oma * oma));
  float y = oma * x;
  *ax = clampf(x, kMinAlpha, 1.0f);
  *ay = clampf(y, kMinAlpha, 1.0f);
}

inline PbrBsdf make_bsdf(const OpenPbr& m, bool front_face) {
  PbrBsdf p;
  p.base_color = m.base_color;  // * ONE (no texture)
  float rough = clampf(m.specular_roughness, 0.0f, 1.0f);  // coat = fuzz = 0: no roughening
  alpha_xy(rough, m.specular_roughness_anisotropy, &p.ax, &p.ay);
  p.specular_color = v3(clampf(m.specular_color.x, 0.f, 1.f), clampf(m.specular_color.y, 0.f, 1.f),
                        clampf(m.specular_color.z, 0.f, 1.f));
  p.specular_eta = modulated_eta(std::fmax(m.specular_ior, 1.0e-4f), std::fmax(m.specular_weight, 0.0f));
  p.transmission_eta = p.specular_eta;
  p.base = clampf(m.base_weight, 0.0f, 1.0f);
  p.diffuse_roughness = clampf(m.base_diffuse_roughness, 0.0f, 1.0f);
  p.front_face = front_face;
  return p;
}

struct LayerW {
  V3 spec_brdf, diff_brdf;
};
// layer_weights (R/bsdf/open_pbr.rs:333-390) with fuzz = coat = metal =
// transmission = subsurface = 0 and e_spec = 0 (eta 1).
inline LayerW layer_weights(const PbrBsdf& p) {
  V3 below_metal = v3(1.f, 1.f, 1.f);
  float e_spec = 0.0f;
  V3 spec_w = vmul(below_metal, p.specular_color);
  V3 leak = vmax(sub(v3(1.f, 1.f, 1.f), smul(p.specular_color, e_spec)), v3(0.f, 0.f, 0.f));
  V3 below_spec = vmul(below_metal, leak);
  V3 below_trans = smul(below_spec, 1.0f - 0.0f);
  V3 diff_w = vmul(smul(below_trans, (1.0f - 0.0f) * p.base), p.base_color);
  return LayerW{spec_w, diff_w};
}
struct Probs {
  float spec, diff, total;
};
inline Probs lobe_probs(const PbrBsdf& p) {  // :392-422
  LayerW w = layer_weights(p);
  float coat = 0.0f;
  float metal = luminance(v3(0.f, 0.f, 0.f));
  float spec = luminance(w.spec_brdf);
  float spec_btdf = 0.0f;
  float fuzz = std::fmax(luminance(v3(0.f, 0.f, 0.f)), 0.0f);
  float diff = std::fmax(luminance(w.diff_brdf), 0.0f);
  float diff_btdf = std::fmax(luminance(v3(0.f, 0.f, 0.f)), 0.0f);
  float total = coat + metal + spec + spec_btdf + fuzz + diff + diff_btdf;
  return Probs{spec, diff, total};
}
inline bool spec_smooth(const PbrBsdf& p) { return std::fmax(p.ax, p.ay) < kSmoothAlpha; }

// Spec lobe pieces at eta 1 (dielectric_ms_params / lobe weights :593-636):
// f_avg = 0, so ss = ms = 0, total = 1e-8 and the ms factor is 0.
inline float spec_eval_scalar(const PbrBsdf& p, V3 wo, V3 wi) {  // eval_spec_brdf :841-867
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  V3 wm;
  if (!reflection_half_vector(wo, wi, &wm)) return 0.0f;
  float d = ggx_d(wm, p.ax, p.ay);
  float g = ggx_g2(wo, wi, p.ax, p.ay);
  if (d <= 0.0f || g <= 0.0f) return 0.0f;
  float f = p.front_face ? weighted_fresnel_ratio(std::fabs(dot(wo, wm)), p.transmission_eta, p.specular_eta)
                         : weighted_fresnel_ratio(std::fabs(dot(wo, wm)), 1.0f / p.transmission_eta,
                                                  1.0f / p.specular_eta);
  float ss = f * (d * g / (4.0f * wo.z * wi.z));
  float ms = 0.0f * 0.0f;  // dielectric_reflection_ms_factor = 0 when f_avg = 0
  return ss + ms;
}
inline float spec_pdf(const PbrBsdf& p, V3 wo, V3 wi) {  // pdf_specular_brdf :1006-1022
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  float pdf_ss = 0.0f;
  V3 wm;
  if (reflection_half_vector(wo, wi, &wm)) {
    float pw = pdf_wm_bounded_vndf(wo, wm, p.ax, p.ay);
    float den = 4.0f * std::fabs(dot(wo, wm));
    pdf_ss = den <= 0.0f ? 0.0f : pw / den;
  }
  float pdf_ms = std::fmax(wi.z, 0.0f) / kPi;
  const float ss_w = 0.0f, ms_w = 0.0f, total = std::fmax(ss_w + ms_w, 1.0e-8f);
  return (ss_w * pdf_ss + ms_w * pdf_ms) / total;
}

// OpenPbrBsdf::eval (:224-259) / pdf (:261-293)
inline V3 bsdf_eval(const PbrBsdf& p, V3 wo, V3 wi) {
  if (!(wo.z > 0.0f)) return v3(0.f, 0.f, 0.f);
  LayerW w = layer_weights(p);
  V3 total = v3(0.f, 0.f, 0.f);
  if (len_sq(w.spec_brdf) > 0.0f && !spec_smooth(p)) total = add(total, smul(w.spec_brdf, spec_eval_scalar(p, wo, wi)));
  if (len_sq(w.diff_brdf) > 0.0f) {
    V3 f = (wo.z <= 0.0f || wi.z <= 0.0f) ? v3(0.f, 0.f, 0.f) : f_eon(v3(1.f, 1.f, 1.f), p.diffuse_roughness, wo, wi);
    total = add(total, vmul(w.diff_brdf, f));
  }
  return total;
}
inline float bsdf_pdf(const PbrBsdf& p, V3 wo, V3 wi) {
  if (!(wo.z > 0.0f)) return 0.0f;
  Probs pr = lobe_probs(p);
  if (pr.total <= 0.0f) return 0.0f;
  float pdf = 0.0f;
  if (pr.spec > 0.0f && !spec_smooth(p)) pdf += (pr.spec / pr.total) * spec_pdf(p, wo, wi);
  if (pr.diff > 0.0f) {
    float pd = (wo.z <= 0.0f || wi.z <= 0.0f) ? 0.0f : pdf_eon(wo, wi, p.diffuse_roughness);
    pdf += (pr.diff / pr.total) * pd;
  }
  return pdf;
}

struct BSample {
  V3 weight, wi;
  float pdf;
  bool delta;
};
// finalize_rough_sample (:1958-1999)
inline bool finalize(const PbrBsdf& p, V3 wo, V3 wi, BSample* out) {
  float pdf = bsdf_pdf(p, wo, wi);
  if (pdf <= 0.0f) return false;
  V3 f = bsdf_eval(p, wo, wi);
  if (len_sq(f) == 0.0f) return false;
  float c = std::fabs(wi.z);
  if (c <= 0.0f) return false;
  out->weight = smul(f, c / pdf);
  out->wi = wi;
  out->pdf = pdf;
  out->delta = false;
  return true;
}
// sample_cosine_weighted_hemisphere, R/math/mod.rs:26-36
inline V3 cosine_hemisphere(float u, float v) {
  float r = std::sqrt(clampf(u, 0.0f, 1.0f));
  float phi = kTau * v;
  float dx = r * std::cos(phi), dy = r * std::sin(phi);
  float z = std::sqrt(std::fmax(1.0f - (dx * dx + dy * dy), 0.0f));
  return v3(dx, dy, z);
}
// OpenPbrBsdf::sample (:295-331) with pick_lobe (:2016-2041)
inline bool bsdf_sample(const PbrBsdf& p, V3 wo, float u_lobe, float ux, float uy, BSample* out) {
  if (!(wo.z > 0.0f)) return false;
  Probs pr = lobe_probs(p);
  if (pr.total <= 0.0f) return false;
  float u = u_lobe * pr.total;
  // coat, metal: 0
  bool spec;
  if (u < 0.0f) return false;  // coat (prob 0): u < 0 never holds for u_lobe >= 0
  u -= 0.0f;
  if (u < 0.0f) return false;  // metal
  u -= 0.0f;
  if (u < pr.spec) {
    spec = true;
  } else {
    spec = false;  // spec_btdf, fuzz (prob 0) are skipped; diff_brdf else diff_btdf
    u -= pr.spec;
    u -= 0.0f;
    u -= 0.0f;
    if (!(u < pr.diff)) {
      // ChosenLobe::DiffBtdf: probability 0 -> p_lobe = 0 -> None
      return false;
    }
  }
  float p_lobe = (spec ? pr.spec : pr.diff) / pr.total;
  if (p_lobe <= 0.0f) return false;
  if (spec) {  // sample_spec_brdf :1589-1668
    if (spec_smooth(p)) {
      V3 wi = v3(-wo.x, -wo.y, wo.z);
      float f = p.front_face ? weighted_fresnel_ratio(std::fabs(wi.z), p.transmission_eta, p.specular_eta)
                             : weighted_fresnel_ratio(std::fabs(wi.z), 1.0f / p.transmission_eta,
                                                      1.0f / p.specular_eta);
      LayerW w = layer_weights(p);
      out->weight = sdiv(smul(w.spec_brdf, f), p_lobe);
      out->wi = wi;
      out->pdf = p_lobe;
      out->delta = true;
      return true;
    }
    const float ss_prob = 0.0f / 1.0e-8f;  // weights.ss / weights.total
    if (ux >= ss_prob) {
      float uu = ss_prob < 1.0f ? (ux - ss_prob) / (1.0f - ss_prob) : 0.0f;
      return finalize(p, wo, cosine_hemisphere(uu, uy), out);
    }
    return false;  // unreachable at eta 1 (ss probability 0)
  }
  // sample_diff_brdf :1936-1940 -> EonBsdf::sample (R/bsdf/eon.rs:40-58)
  if (wo.z <= 0.0f) return false;
  float pdf;
  V3 wi = sample_eon(wo, p.diffuse_roughness, ux, uy, &pdf);
  if (pdf <= 0.0f || wi.z <= 0.0f) return false;
  return finalize(p, wo, wi, out);
}

// ---- scene shading data ------------------------------------------------------------
// Material enum of the backend (R/material/mod.rs:100-115 variants carried).
enum MatKind : uint32_t { kMatOpenPbr = 0, kMatLambert = 1, kMatEmissive = 2, kMatMirror = 3 };
struct Material {
  uint32_t kind = kMatOpenPbr;
  OpenPbr pbr{};
  V3 color{0.f, 0.f, 0.f};  // Lambert rho, Emissive colour, Mirror colour
  float strength = 0.0f;    // Emissive
};

struct ShadeMesh {
  std::vector<float> nrm;  // 3 per vertex (may be empty: geometric normal)
  std::vector<float> uv;   // 2 per vertex (may be empty)
};
struct RenderScene {
  const Scene* geo = nullptr;
  std::vector<ShadeMesh> meshes;
  std::vector<uint32_t> inst_material;
  std::vector<Material> materials;
  // light tree over the emissive triangles (Scene::build_light_tree) and the
  // (instance, triangle) -> leaf node table
  lt::Tree tree;
  std::vector<uint32_t> leaf_base;  // per instance: offset into leaf_of, or lt::kInvalid
  std::vector<uint32_t> leaf_of;
  bool has_env = false;
  lt::env::Env env;
  uint32_t leaf_node(uint32_t inst, uint32_t tri) const {
    return leaf_base[inst] == lt::kInvalid ? lt::kInvalid : leaf_of[leaf_base[inst] + tri];
  }
};

struct Vertex {
  V3 p, ng, ns, wo;
  Onb frame;
  bool front_face;
  uint32_t material;
  uint32_t inst, tri;
};

inline V3 mat3_mul_normal(const float* w2l, V3 n) {
  // normal_to_world = Mat3::from_mat4(world_to_local.transpose()): column c is
  // row c of world_to_local's upper 3x3.  mul_vec3 = (x*cx + y*cy) + z*cz.
  V3 c0 = v3(w2l[0], w2l[4], w2l[8]), c1 = v3(w2l[1], w2l[5], w2l[9]), c2 = v3(w2l[2], w2l[6], w2l[10]);
  return add(add(mul(n.x, c0), mul(n.y, c1)), mul(n.z, c2));
}

// Scene::triangle_positions (R/scene/mod.rs:432-438): world positions.
inline void triangle_world(const RenderScene& rs, uint32_t inst, uint32_t tri, V3* p) {
  const Instance& in = rs.geo->inst[inst];
  const Mesh& m = rs.geo->meshes[in.mesh];
  for (int k = 0; k < 3; ++k) p[k] = transform_point3(in.l2w, m.vert(m.idx[3 * tri + k]));
}

// base_shading_vertex_from_triangle_sample_impl (R/scene/mod.rs:471-544),
// the fields the integrator reads; `dir` is the incident direction.
inline Vertex vertex_at(const RenderScene& rs, uint32_t inst, uint32_t tri, float b0, float b1, float b2, V3 dir) {
  const Instance& in = rs.geo->inst[inst];
  const Mesh& m = rs.geo->meshes[in.mesh];
  const ShadeMesh& sm = rs.meshes[in.mesh];
  uint32_t i0 = m.idx[3 * tri], i1 = m.idx[3 * tri + 1], i2 = m.idx[3 * tri + 2];
  V3 p0 = transform_point3(in.l2w, m.vert(i0)), p1 = transform_point3(in.l2w, m.vert(i1)),
     p2 = transform_point3(in.l2w, m.vert(i2));
  auto nrm = [&](uint32_t i) {
    V3 n = sm.nrm.empty() ? v3(0.f, 0.f, 0.f) : v3(sm.nrm[3 * i], sm.nrm[3 * i + 1], sm.nrm[3 * i + 2]);
    return normalize_or_zero3(mat3_mul_normal(in.w2l, n));
  };
  auto uvv = [&](uint32_t i, int k) { return sm.uv.empty() ? 0.0f : sm.uv[2 * i + k]; };
  V3 n0 = nrm(i0), n1 = nrm(i1), n2 = nrm(i2);
  Vertex v;
  v.p = interp3(b0, b1, b2, p0, p1, p2);
  V3 gn = normalize_or_zero3(cross(sub(p1, p0), sub(p2, p0)));
  V3 sn = normalize_or_zero3(interp3(b0, b1, b2, n0, n1, n2));
  V3 nd = neg(dir);
  v.front_face = dot(gn, nd) >= 0.0f;
  v.ng = face_forward(gn, nd);
  V3 raw_ns = len_sq(sn) > 0.0f ? sn : v.ng;
  v.ns = face_forward(raw_ns, v.ng);
  // compute_surface_partials (R/math/mod.rs:112-134) over world positions
  float du1 = uvv(i1, 0) - uvv(i0, 0), dv1 = uvv(i1, 1) - uvv(i0, 1);
  float du2 = uvv(i2, 0) - uvv(i0, 0), dv2 = uvv(i2, 1) - uvv(i0, 1);
  float det = du1 * dv2 - dv1 * du2;
  V3 dpdu;
  bool ok = false;
  if (!(std::fabs(det) <= 1.0e-8f)) {
    float inv = 1.0f / det;
    V3 dp1 = sub(p1, p0), dp2 = sub(p2, p0);
    dpdu = smul(sub(mul(dv2, dp1), mul(dv1, dp2)), inv);
    V3 dpdv = smul(add(mul(-du2, dp1), mul(du1, dp2)), inv);
    ok = !(len_sq(dpdu) == 0.0f || len_sq(dpdv) == 0.0f);
  }
  if (!ok) dpdu = onb_from_normal(v.ns).t;
  v.frame = onb_from_normal_and_tangent(v.ns, dpdu);
  v.wo = normalize_or_zero3(nd);
  v.material = rs.inst_material[inst];
  v.inst = inst;
  v.tri = tri;
  return v;
}
inline Vertex shading_vertex(const RenderScene& rs, const SceneHit& h, V3 dir) {
  return vertex_at(rs, h.inst, h.tri, h.b0, h.b1, h.b2, dir);
}

// OpenPbrMaterial::le (R/material/open_pbr.rs:351-375) with exit tint 1.
inline bool pbr_le(const OpenPbr& m, const Vertex& v, V3* le) {
  if (m.emission_luminance <= 0.0f) return false;
  if (!m.thin_walled && !v.front_face) return false;
  V3 wl = normalize_or_zero3(onb_to_local(v.frame, v.wo));
  float co = clampf(std::fabs(wl.z), 0.0f, 1.0f);
  V3 tint = co <= 0.0f ? v3(0.f, 0.f, 0.f) : v3(1.f, 1.f, 1.f);
  *le = vmul(smul(m.emission_color, std::fmax(m.emission_luminance, 0.0f)), tint);
  return true;
}

// OpenPbrMaterial::sample (R/material/open_pbr.rs:489-548) in radiance mode.
constexpr float kGeomCosEps = 1.0e-6f;  // R/material/mod.rs:31
inline bool pbr_sample(const OpenPbr& m, const Vertex& v, float u_lobe, float ux, float uy, BSample* out,
                       V3* wi_world) {
  if (dot(v.wo, v.ng) <= 0.0f) return false;
  V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
  if (wol.z <= 0.0f) return false;
  PbrBsdf p = make_bsdf(m, v.front_face);
  BSample s;
  if (!bsdf_sample(p, wol, u_lobe, ux, uy, &s)) return false;
  V3 wi = onb_to_world(v.frame, s.wi);
  if (dot(wi, v.ng) <= kGeomCosEps) return false;  // reflection lobes only in the subset
  *out = s;  // modified_bsdf_sample_weight leaves radiance-mode weights unchanged
  *wi_world = wi;
  return true;
}

// ---- Material dispatch (R/material/mod.rs:251-412) ------------------------------
inline bool mat_is_pure_emitter(const Material& m) { return m.kind == kMatEmissive; }
inline bool mat_may_emit(const Material& m) {
  if (m.kind == kMatEmissive) return true;
  if (m.kind == kMatOpenPbr) return m.pbr.emission_luminance > 0.0f;  // open_pbr.rs:377-379
  return false;
}
inline float mat_max_emission(const Material& m) {
  if (m.kind == kMatEmissive)  // emissive.rs max_emission, no texture
    return std::fmax(lt::vmax_elem(smul(m.color, m.strength)) * 1.0f, 0.0f);
  if (m.kind == kMatOpenPbr) {  // open_pbr.rs:381-391
    if (m.pbr.emission_luminance <= 0.0f) return 0.0f;
    return std::fmax(lt::vmax_elem(m.pbr.emission_color) * m.pbr.emission_luminance * 1.0f, 0.0f);
  }
  return 0.0f;
}
inline bool mat_le(const Material& m, const Vertex& v, V3* le) {
  if (m.kind == kMatEmissive) {  // color * strength * texture factor (1)
    *le = vmul(smul(m.color, m.strength), v3(1.f, 1.f, 1.f));
    return true;
  }
  if (m.kind == kMatOpenPbr) return pbr_le(m.pbr, v, le);
  return false;
}
// NormalizedLambertMaterial::sample (normalized_lambert.rs:120-158) and
// MirrorMaterial::sample (mirror.rs:112-143), radiance mode.
inline bool mat_sample(const Material& m, const Vertex& v, float u_lobe, float ux, float uy, BSample* out,
                       V3* wi_world) {
  if (m.kind == kMatOpenPbr) return pbr_sample(m.pbr, v, u_lobe, ux, uy, out, wi_world);
  if (m.kind == kMatEmissive) return false;
  V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
  if (wol.z <= 0.0f) return false;
  BSample s;
  if (m.kind == kMatLambert) {  // NormalizedLambertBsdf::sample (bsdf/normalized_lambert.rs:35-60)
    V3 wi = cosine_hemisphere(ux, uy);
    float pdf = std::fmax(wi.z, 0.0f) / kPi;
    if (pdf <= 0.0f) return false;
    s.weight = m.color;
    s.wi = wi;
    s.pdf = pdf;
    s.delta = false;
  } else {  // MirrorBsdf::sample (bsdf/mirror.rs:23-39)
    s.weight = m.color;
    s.wi = normalize_or_zero3(v3(-wol.x, -wol.y, wol.z));
    s.pdf = 1.0f;
    s.delta = true;
  }
  V3 wi = onb_to_world(v.frame, s.wi);
  if (dot(wi, v.ng) <= kGeomCosEps) return false;
  *out = s;
  *wi_world = wi;
  return true;
}
// modified_bsdf_eval (R/material/mod.rs:33-53), radiance mode
inline V3 shading_normal_correction(const Vertex& v, V3 wi, V3 f) {
  if (len_sq(f) == 0.0f) return v3(0.f, 0.f, 0.f);
  float cg = std::fabs(dot(wi, v.ng));
  if (cg <= kGeomCosEps) return v3(0.f, 0.f, 0.f);
  return smul(f, std::fabs(dot(wi, v.ns)) / cg);
}
// Material::eval / pdf (R/material/mod.rs:251-370) in radiance mode:
// NormalizedLambert (normalized_lambert.rs:160-193), OpenPBR
// (open_pbr.rs:549-590); Mirror and Emissive evaluate to zero.
inline V3 mat_eval(const Material& m, const Vertex& v, V3 wi) {
  const V3 zero = v3(0.f, 0.f, 0.f);
  if (m.kind == kMatLambert) {
    if (dot(v.wo, v.ng) <= 0.0f || dot(wi, v.ng) <= 0.0f) return zero;
    V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
    V3 wil = normalize_or_zero3(onb_to_local(v.frame, wi));
    V3 f = (wol.z <= 0.0f || wil.z <= 0.0f) ? zero : sdiv(m.color, kPi);
    return shading_normal_correction(v, wi, f);
  }
  if (m.kind == kMatOpenPbr) {
    if (dot(v.wo, v.ng) <= 0.0f) return zero;
    V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
    V3 wil = normalize_or_zero3(onb_to_local(v.frame, wi));
    if (wol.z <= 0.0f) return zero;
    return shading_normal_correction(v, wi, bsdf_eval(make_bsdf(m.pbr, v.front_face), wol, wil));
  }
  return zero;
}
inline float mat_pdf(const Material& m, const Vertex& v, V3 wi) {
  if (m.kind == kMatLambert) {
    if (dot(v.wo, v.ng) <= 0.0f || dot(wi, v.ng) <= 0.0f) return 0.0f;
    V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
    V3 wil = normalize_or_zero3(onb_to_local(v.frame, wi));
    if (wol.z <= 0.0f || wil.z <= 0.0f) return 0.0f;
    return std::fmax(wil.z, 0.0f) / kPi;
  }
  if (m.kind == kMatOpenPbr) {
    if (dot(v.wo, v.ng) <= 0.0f) return 0.0f;
    V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
    V3 wil = normalize_or_zero3(onb_to_local(v.frame, wi));
    if (wol.z <= 0.0f) return 0.0f;
    return bsdf_pdf(make_bsdf(m.pbr, v.front_face), wol, wil);
  }
  return 0.0f;
}
// light_tree_precompute: NormalizedLambert (normalized_lambert.rs:201-218)
// gives a diffuse query; every other material here gives none.
inline bool mat_query(const Material& m, const Vertex& v, lt::Query* q) {
  if (m.kind != kMatLambert) return false;
  float rho = lt::luminance(m.color);
  if (rho <= 0.0f) return false;
  *q = lt::Query{v.p, v.ns, rho};
  return true;
}

// Scene::register_area_light_triangles (R/scene/mod.rs:649-683) +
// build_light_tree (R/light_tree/builder.rs:58-121), triangle leaves.
inline void build_lights(RenderScene& rs) {
  std::vector<lt::Leaf> leaves;
  const Scene& g = *rs.geo;
  rs.leaf_base.assign(g.inst.size(), lt::kInvalid);
  rs.leaf_of.clear();
  for (uint32_t i = 0; i < (uint32_t)g.inst.size(); ++i) {
    const Material& m = rs.materials[rs.inst_material[i]];
    if (!mat_may_emit(m)) continue;
    float me = mat_max_emission(m);
    if (me <= 0.0f) continue;
    const Mesh& mesh = g.meshes[g.inst[i].mesh];
    const uint32_t ntri = (uint32_t)(mesh.idx.size() / 3);
    rs.leaf_base[i] = (uint32_t)rs.leaf_of.size();
    rs.leaf_of.resize(rs.leaf_of.size() + ntri, lt::kInvalid);
    for (uint32_t t = 0; t < ntri; ++t) {
      V3 p[3];
      triangle_world(rs, i, t, p);
      float area = 0.5f * lt::vlen(cross(sub(p[1], p[0]), sub(p[2], p[0])));
      if (area * me <= 0.0f) continue;
      lt::Leaf l;
      if (lt::triangle_leaf(i, t, p[0], p[1], p[2], me, &l)) leaves.push_back(l);
    }
  }
  rs.tree = lt::build_tree(leaves);
  for (uint32_t k = 0; k < (uint32_t)rs.tree.nodes.size(); ++k) {
    const lt::Node& n = rs.tree.nodes[k];
    if (n.leaf) rs.leaf_of[rs.leaf_base[n.inst] + n.tri] = k;
  }
}

inline float balance_heuristic(float a, float b) {  // R/math/mod.rs:163-171
  float sum = a + b;
  if (a <= 0.0f || sum <= 0.0f) return 0.0f;
  return a / sum;
}

constexpr float kRayEpsilon = 1.0e-4f;  // R/integrator/mod.rs:49
inline V3 offset_origin(V3 p, V3 ng, V3 dir) {
  V3 off = dot(dir, ng) >= 0.0f ? ng : neg(ng);
  return add(p, mul(kRayEpsilon, off));
}

// LightSampler (R/light/mod.rs:136-233): the tree category (weight = root
// flux) and the environment category (weight = total power), a normalised
// CDF over them.
struct Category {
  int kind;  // 0 tree, 1 environment
  float pmf;
};
inline void light_categories(const RenderScene& rs, float* w_tree, float* w_env, float* total) {
  *w_tree = std::fmax(rs.tree.root_flux(), 0.0f);
  *w_env = rs.has_env ? std::fmax(lt::env::total_power(rs.env), 0.0f) : 0.0f;
  *total = 0.0f;
  if (*w_tree > 0.0f) *total += *w_tree;
  if (*w_env > 0.0f) *total += *w_env;
}
inline bool sample_category(const RenderScene& rs, float u, Category* out) {
  float wt, we, total;
  light_categories(rs, &wt, &we, &total);
  float w[2];
  int kind[2];
  int n = 0;
  if (wt > 0.0f) { w[n] = wt; kind[n++] = 0; }
  if (we > 0.0f) { w[n] = we; kind[n++] = 1; }
  if (n == 0 || total <= 0.0f) return false;
  float cdf[3] = {0.0f, 0.0f, 0.0f};
  float acc = 0.0f;
  for (int k = 0; k < n; ++k) {
    acc += w[k];
    cdf[k + 1] = acc;
  }
  const float inv = 1.0f / acc;
  for (int k = 0; k <= n; ++k) cdf[k] *= inv;
  u = clampf(u, 0.0f, 1.0f);
  int pp = 0;
  while (pp <= n && cdf[pp] <= u) ++pp;  // partition_point(c <= u)
  int idx = pp > 0 ? pp - 1 : 0;
  if (idx > n - 1) idx = n - 1;
  out->kind = kind[idx];
  out->pmf = w[idx] / total;
  return true;
}
inline float category_pmf(const RenderScene& rs, int kind) {
  float wt, we, total;
  light_categories(rs, &wt, &we, &total);
  if (total <= 0.0f) return 0.0f;
  const float w = kind == 0 ? wt : we;
  return w > 0.0f ? w / total : 0.0f;
}

// direct_light_mis_contribution (R/integrator/mis.rs:141-210) with
// sample_light_mis_compensated_lazy (R/light/mod.rs:289-331),
// sample_li_for_triangle (R/light/area.rs:12-54) and the environment's
// MIS-compensated sample (R/light/environment.rs:260-276).
inline V3 direct_light(const RenderScene& rs, const Material& m, const Vertex& vtx, float u_root, float u_tree,
                       float u_aux, float us0, float us1) {
  (void)u_aux;
  const V3 zero = v3(0.f, 0.f, 0.f);
  Category cat;
  if (!sample_category(rs, u_root, &cat)) return zero;
  V3 le, wi;
  float li_pdf, selection;
  uint32_t t_inst = lt::kInvalid, t_tri = lt::kInvalid;
  if (cat.kind == 0) {
    lt::Query q;
    if (!mat_query(m, vtx, &q)) return zero;
    uint32_t leaf;
    float leaf_pmf;
    if (!lt::sample_tree(rs.tree, q, u_tree, &leaf, &leaf_pmf)) return zero;
    const lt::Node& ln = rs.tree.nodes[leaf];
    // sample_triangle_point (R/scene/mod.rs:616-633)
    V3 p[3];
    triangle_world(rs, ln.inst, ln.tri, p);
    float su0 = std::sqrt(clampf(us0, 0.0f, 1.0f));
    float u1 = clampf(us1, 0.0f, 1.0f);
    float b0 = 1.0f - su0, b1 = u1 * su0, b2 = (1.0f - u1) * su0;
    V3 lp = interp3(b0, b1, b2, p[0], p[1], p[2]);
    float area = 0.5f * lt::vlen(cross(sub(p[1], p[0]), sub(p[2], p[0])));
    float pdf_area = area > 0.0f ? 1.0f / area : 0.0f;
    if (pdf_area <= 0.0f) return zero;
    V3 to_light = sub(lp, vtx.p);
    float d2 = len_sq(to_light);
    if (d2 <= 0.0f) return zero;
    float dist = std::sqrt(d2);
    wi = sdiv(to_light, dist);
    const Material& lm = rs.materials[rs.inst_material[ln.inst]];
    if (!mat_may_emit(lm)) return zero;
    Vertex lv = vertex_at(rs, ln.inst, ln.tri, b0, b1, b2, wi);
    if (!mat_le(lm, lv, &le)) return zero;
    float cos_light = std::fmax(dot(lv.ng, neg(wi)), 0.0f);
    if (cos_light <= 0.0f) return zero;
    li_pdf = pdf_area * d2 / cos_light;
    selection = cat.pmf * leaf_pmf;
    t_inst = ln.inst;
    t_tri = ln.tri;
  } else {
    lt::env::Sample es;
    if (!lt::env::sample(rs.env, rs.env.dist_mis, us0, us1, &es)) return zero;
    if (es.pdf <= 0.0f) return zero;
    le = es.radiance;
    wi = es.dir;
    li_pdf = es.pdf;
    selection = cat.pmf;
  }
  if (li_pdf <= 0.0f || selection <= 0.0f) return zero;
  V3 f = mat_eval(m, vtx, wi);
  float bsdf_pdf = len_sq(f) == 0.0f ? 0.0f : mat_pdf(m, vtx, wi);
  if (len_sq(f) == 0.0f) return zero;
  float cos_surface = std::fabs(dot(vtx.ng, wi));
  if (cos_surface <= 0.0f) return zero;
  // unoccluded_ray (R/integrator/mod.rs:239-265): a light triangle target, or
  // an infinite light that any hit blocks
  SceneHit sh;
  if (closest_hit(*rs.geo, offset_origin(vtx.p, vtx.ng, wi), wi, INFINITY, &sh, nullptr) &&
      !(t_inst != lt::kInvalid && sh.inst == t_inst && sh.tri == t_tri))
    return zero;
  float pdf_total = selection * li_pdf;
  if (pdf_total <= 0.0f) return zero;
  float mis = balance_heuristic(pdf_total, bsdf_pdf);
  return smul(vmul(le, f), cos_surface * mis / pdf_total);
}

// pdf_for_triangle_hit (R/light/mod.rs:407-429) with
// Scene::area_light_pdf_solid_angle (R/scene/mod.rs:587-611).
inline float light_pdf_for_hit(const RenderScene& rs, const Material& shading_m, const Vertex& vtx, const Vertex& lv) {
  if (rs.tree.nodes.empty()) return 0.0f;
  lt::Query q;
  if (!mat_query(shading_m, vtx, &q)) return 0.0f;
  uint32_t leaf = rs.leaf_node(lv.inst, lv.tri);
  float leaf_pmf = lt::pdf_leaf(rs.tree, q, leaf);
  if (leaf_pmf <= 0.0f) return 0.0f;
  float cat_pmf = category_pmf(rs, 0);
  const float area = rs.tree.nodes[leaf].area;
  if (area <= 0.0f) return 0.0f;
  float pdf_area = 1.0f / area;
  V3 to_light = sub(lv.p, vtx.p);
  float d2 = len_sq(to_light);
  if (d2 <= 0.0f) return 0.0f;
  float dist = std::sqrt(d2);
  V3 wi = sdiv(to_light, dist);
  float cos_light = std::fmax(dot(lv.ng, neg(wi)), 0.0f);
  if (cos_light <= 0.0f) return 0.0f;
  return cat_pmf * leaf_pmf * (pdf_area * d2 / cos_light);
}

// emitted_radiance_from_bsdf_sample_area (R/integrator/mis.rs:212-234)
inline V3 emitted_from_bsdf_sample(const RenderScene& rs, V3 throughput, float bsdf_pdf, bool delta,
                                   const Material& shading_m, const Vertex& vtx, const Vertex& lv,
                                   const Material& light_m) {
  V3 le;
  if (!mat_le(light_m, lv, &le)) return v3(0.f, 0.f, 0.f);
  if (delta) return vmul(throughput, le);
  float light_pdf = light_pdf_for_hit(rs, shading_m, vtx, lv);
  float mis = balance_heuristic(bsdf_pdf, light_pdf);
  return smul(vmul(throughput, le), mis);
}

// infinite_light_le (R/light/environment.rs:305-311)
inline V3 env_le(const RenderScene& rs, V3 dir) {
  return rs.has_env ? lt::env::radiance(rs.env, dir) : v3(0.f, 0.f, 0.f);
}
// emitted_radiance_from_bsdf_sample_infinite (R/integrator/mis.rs:236-257)
// with pdf_for_environment_hit (R/light/mod.rs:431-443), MIS-compensated.
inline V3 emitted_infinite(const RenderScene& rs, V3 throughput, float bsdf_pdf, bool delta, V3 dir) {
  V3 le = env_le(rs, dir);
  if (le.x == 0.0f && le.y == 0.0f && le.z == 0.0f) return v3(0.f, 0.f, 0.f);
  if (delta) return vmul(throughput, le);
  float cat = category_pmf(rs, 1);
  float light_pdf = cat <= 0.0f ? 0.0f : cat * lt::env::pdf(rs.env, rs.env.dist_mis, dir);
  return smul(vmul(throughput, le), balance_heuristic(bsdf_pdf, light_pdf));
}

// mis::trace_radiance (R/integrator/mis.rs:19-139).
inline V3 trace_radiance(const RenderScene& rs, V3 o, V3 d, const PathSampler& ps, uint32_t max_depth) {
  V3 radiance = v3(0.f, 0.f, 0.f), throughput = v3(1.f, 1.f, 1.f);
  SceneHit h;
  if (!closest_hit(*rs.geo, o, d, INFINITY, &h, nullptr)) return env_le(rs, d);
  Vertex vtx = shading_vertex(rs, h, d);
  const Material* mat = &rs.materials[vtx.material];
  V3 le;
  if (mat_le(*mat, vtx, &le)) radiance = add(radiance, le);
  const uint32_t rr_start = 4;
  for (uint32_t depth = 0; depth < max_depth; ++depth) {
    const uint32_t base = 2u + depth * 16u;
    if (!mat_is_pure_emitter(*mat)) {
      float us[2];
      ps.s2(base + 3, us);
      V3 dl = direct_light(rs, *mat, vtx, ps.s1(base + 0), ps.s1(base + 1), ps.s1(base + 2), us[0], us[1]);
      radiance = add(radiance, vmul(throughput, dl));
    }
    BSample s;
    V3 wi;
    float udir[2];
    ps.s2(base + 7, udir);
    if (!mat_sample(*mat, vtx, ps.s1(base + 5), udir[0], udir[1], &s, &wi)) break;
    const bool delta = s.delta;
    throughput = vmul(throughput, s.weight);
    if (depth + 1 >= rr_start) {
      float q = clampf(rmax(throughput.x, rmax(throughput.y, throughput.z)), 0.05f, 1.0f);
      if (ps.s1(base + 13) > q) break;
      throughput = sdiv(throughput, q);
    }
    V3 no = offset_origin(vtx.p, vtx.ng, wi);
    SceneHit nh;
    if (!closest_hit(*rs.geo, no, wi, INFINITY, &nh, nullptr)) {
      radiance = add(radiance, emitted_infinite(rs, throughput, s.pdf, delta, wi));
      break;
    }
    Vertex nv = shading_vertex(rs, nh, wi);
    const Material* nm = &rs.materials[nv.material];
    radiance = add(radiance, emitted_from_bsdf_sample(rs, throughput, s.pdf, delta, *mat, vtx, nv, *nm));
    vtx = nv;
    mat = nm;
  }
  return radiance;
}

}  // namespace orc
