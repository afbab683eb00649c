// Oracle: BSDFs of the reference for the carried materials, restated line by
// line, with the GGX tables they read.  TEST INFRASTRUCTURE ONLY.
//
//   R/math/orthonormal_basis.rs          ONB (Duff et al.)
//   R/bsdf/smith_ggx.rs                  GGX D / Lambda / G1 / G2, VNDF and bounded VNDF
//   R/math/mod.rs:26-72                  cosine hemisphere, dielectric Fresnel
//   R/bsdf/eon.rs                        EON diffuse
//   R/bsdf/directional_albedo.rs         dielectric GGX directional albedo table, conductor and
//                                        dielectric energy-compensation tables (host builds, lookups)
//   R/bsdf/dielectric_ggx.rs:121-587     the dielectric GGX eval / rough sampling those builds use
//   R/bsdf/open_pbr.rs                   OpenPbrBsdf: coat, F82 metal, dielectric specular with
//                                        multiple-scattering compensation, fuzz (LTC), EON diffuse
//   R/material/open_pbr.rs:393-941       OpenPbrMaterial::make_bsdf, roughness coupling, coat darkening
//
// Carried OpenPBR parameters: base, metalness, specular (weight, colour,
// roughness, anisotropy, IOR), coat (weight, colour, roughness, anisotropy,
// IOR, darkening), fuzz (weight, colour, roughness), diffuse roughness and
// emission, the specular BTDF (rough and smooth, with its multiple-scattering
// compensation) and thin walls.  Dispersion, thin film, subsurface and normal
// maps are not carried (the backend rejects such materials).
//
// Transcendental functions (sin, cos, atan2, exp, log, pow) are the host
// libm's, as the reference's f32 methods are on Linux.  f32::powi of a
// run-time base is LLVM's binary-exponentiation expansion (powi_rt); of
// compile-time constants LLVM folds it through the double pow (powi_const).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

#include "geometry.h"

namespace orc {

constexpr real kPi = 3.14159265358979323846f;
constexpr real kTau = 6.28318530717958647692f;

// ---- glam-style helpers --------------------------------------------------
inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline V3 vmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 smul(V3 a, real s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 sdiv(V3 a, real s) { return v3(a.x / s, a.y / s, a.z / s); }
inline real len_sq(V3 a) { return dot(a, a); }
inline V3 normalize3(V3 a) {  // v * (1 / |v|)
  real r = 1.0f / std::sqrt(dot(a, a));
  return v3(a.x * r, a.y * r, a.z * r);
}
inline V3 normalize_or_zero3(V3 a) {
  real r = 1.0f / std::sqrt(dot(a, a));
  if (std::isfinite(r) && r > 0.0f) return v3(a.x * r, a.y * r, a.z * r);
  return v3(0.f, 0.f, 0.f);
}
inline real clampf(real x, real lo, real hi) { return x < lo ? lo : (x > hi ? hi : x); }  // f32::clamp
inline V3 face_forward(V3 n, V3 ref) { return dot(n, ref) < 0.0f ? neg(n) : n; }  // R/math/mod.rs:42-48
inline real luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }  // R/math/sg.rs:314-316
inline V3 interp3(real b0, real b1, real b2, V3 a, V3 b, V3 c) {
  return add(add(mul(b0, a), mul(b1, b)), mul(b2, c));
}

// ---- OrthonormalBasis, R/math/orthonormal_basis.rs:11-71 -------------------
struct Onb {
  V3 t, b, n;
};
inline Onb onb_from_normal(V3 normal) {
  V3 n = len_sq(normal) > 0.0f ? normalize3(normal) : v3(0.f, 0.f, 1.f);
  real sign = std::copysign(1.0f, n.z);
  real a = -1.0f / (sign + n.z);
  real b = n.x * n.y * a;
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
constexpr real kMinAlpha = 1.0e-4f;
constexpr real kSmoothAlpha = 1.0e-3f;
inline bool reflection_half_vector(V3 wo, V3 wi, V3* wm) {
  V3 m = normalize_or_zero3(add(wo, wi));
  if (len_sq(m) == 0.0f) return false;
  *wm = m.z < 0.0f ? neg(m) : m;
  return true;
}
inline real ggx_d(V3 m, real ax, real ay) {
  if (m.z <= 0.0f) return 0.0f;
  real term = m.x * m.x / (ax * ax) + m.y * m.y / (ay * ay) + m.z * m.z;
  real denom = kPi * ax * ay * term * term;
  return denom <= 0.0f ? 0.0f : 1.0f / denom;
}
inline real ggx_lambda(V3 w, real ax, real ay) {
  real c = std::fabs(w.z);
  if (c <= 0.0f) return INFINITY;
  real term = 1.0f + (ax * ax * w.x * w.x + ay * ay * w.y * w.y) / (c * c);
  return 0.5f * (-1.0f + std::sqrt(term));
}
inline real ggx_g2(V3 wo, V3 wi, real ax, real ay) {
  real lo = ggx_lambda(wo, ax, ay), li = ggx_lambda(wi, ax, ay);
  if (!std::isfinite(lo) || !std::isfinite(li)) return 0.0f;
  return 1.0f / (1.0f + lo + li);
}
inline real pdf_wm_bounded_vndf(V3 wo, V3 wm, real ax, real ay) {
  if (wo.z <= 0.0f || wm.z <= 0.0f) return 0.0f;
  ax = std::fmax(ax, kMinAlpha);
  ay = std::fmax(ay, kMinAlpha);
  real a = clampf(std::fmin(ax, ay), 0.0f, 1.0f);
  real s = 1.0f + std::sqrt(wo.x * wo.x + wo.y * wo.y);
  real a2 = a * a, s2 = s * s;
  real k = (1.0f - a2) * s2 / (s2 + a2 * wo.z * wo.z);
  real t = std::sqrt(ax * ax * wo.x * wo.x + ay * ay * wo.y * wo.y + wo.z * wo.z);
  real denom = k * wo.z + t;
  if (denom <= 0.0f) return 0.0f;
  return 2.0f * ggx_d(wm, ax, ay) * std::fmax(dot(wo, wm), 0.0f) / denom;
}

// ---- fresnel, R/math/mod.rs:50-72 -----------------------------------------------
inline real fresnel_dielectric(real ci, real ei, real et) {
  ci = clampf(ci, -1.0f, 1.0f);
  if (ci < 0.0f) {
    std::swap(ei, et);
    ci = -ci;
  }
  real si = std::sqrt(std::fmax(1.0f - ci * ci, 0.0f));
  real st = (ei / et) * si;
  if (st >= 1.0f) return 1.0f;
  real ct = std::sqrt(std::fmax(1.0f - st * st, 0.0f));
  real rpar = (et * ci - ei * ct) / (et * ci + ei * ct);
  real rper = (ei * ci - et * ct) / (ei * ci + et * ct);
  return 0.5f * (rpar * rpar + rper * rper);
}
// weighted_fresnel_ratio, R/bsdf/open_pbr.rs:2161-2172
inline real weighted_fresnel_ratio(real c, real eti, real etp) {
  eti = std::fmax(eti, 1.0e-4f);
  etp = std::fmax(etp, 1.0e-4f);
  if (etp >= 1.0f) return fresnel_dielectric(c, 1.0f, etp);
  real cc = clampf(c, 0.0f, 1.0f);
  real mu2 = 1.0f - (1.0f - cc * cc) / (eti * eti);
  if (mu2 <= 0.0f) return 1.0f;
  return fresnel_dielectric(std::sqrt(mu2), 1.0f, 1.0f / etp);
}
// f_avg_dielectric, R/bsdf/open_pbr.rs:2127-2135
inline real f_avg_dielectric(real eta) {
  if (eta >= 1.0f) return clampf((eta - 1.0f) / (4.08567f + 1.00071f * eta), 0.0f, 1.0f);
  real e = std::fmax(eta, 1.0e-6f);
  return clampf(0.997118f + 0.1014f * e - 0.965241f * e * e - 0.130607f * e * e * e, 0.0f, 1.0f);
}

// ---- EON, R/bsdf/eon.rs:12-197 -----------------------------------------------
constexpr real kRcpPi = 1.0f / kPi;
const real kC1Fon = 0.5f - 2.0f / (3.0f * kPi);
const real kC2Fon = 2.0f / 3.0f - 28.0f / (15.0f * kPi);
inline real e_fon_approx(real mu, real r) {
  real mc = 1.0f - mu;
  real g = mc * (0.05710853f + mc * (0.49188188f + mc * (-0.33218145f + mc * 0.071443f)));
  return (1.0f + r * g) / (1.0f + kC1Fon * r);
}
inline V3 f_eon(V3 rho, real r, V3 wo, V3 wi) {
  real mo = wo.z, mi = wi.z;
  real s = dot(wo, wi) - mi * mo;
  real sot = s > 0.0f ? s / std::fmax(mo, mi) : s;
  real af = 1.0f / (1.0f + kC1Fon * r);
  V3 fss = mul(kRcpPi * af * (1.0f + r * sot), rho);
  real efo = e_fon_approx(mo, r), efi = e_fon_approx(mi, r);
  real avg = af * (1.0f + kC2Fon * r);
  real oma = 1.0f - avg;
  V3 denom = sub(v3(1.f, 1.f, 1.f), smul(rho, oma));
  V3 rho_ms = v3(rho.x * rho.x * avg / denom.x, rho.y * rho.y * avg / denom.y, rho.z * rho.z * avg / denom.z);
  real scale = kRcpPi * std::fmax(1.0f - efo, 1.0e-7f) * std::fmax(1.0f - efi, 1.0e-7f) / std::fmax(oma, 1.0e-7f);
  return add(fss, smul(rho_ms, scale));
}
inline void ltc_coeffs(real mu, real r, real* a, real* b, real* c, real* d) {
  *a = 1.0f + r * (0.303392f + (-0.518982f + 0.111709f * mu) * mu + (-0.276266f + 0.335918f * mu) * r);
  *b = r * (-1.16407f + 1.15859f * mu + (0.150815f - 0.150105f * mu) * r) / (mu * mu * mu - 1.43545f);
  *c = 1.0f + r * (0.20013f + (-0.506373f + 0.261777f * mu) * mu);
  *d = r * (0.540852f + (-1.01625f + 0.475392f * mu) * mu) / (-1.0743f + (0.0725628f + mu) * mu);
}
inline void ltc_basis_x(V3 wo, real* x, real* y) {
  real l2 = wo.x * wo.x + wo.y * wo.y;
  if (l2 > 0.0f) {
    real inv = 1.0f / std::sqrt(l2);
    *x = wo.x * inv;
    *y = wo.y * inv;
  } else {
    *x = 1.0f;
    *y = 0.0f;
  }
}
inline V3 ltc_from(V3 wo, V3 v) {
  real xx, xy;
  ltc_basis_x(wo, &xx, &xy);
  real yx = -xy, yy = xx;
  return v3(xx * v.x + yx * v.y, xy * v.x + yy * v.y, v.z);
}
inline V3 ltc_to(V3 wo, V3 v) {
  real xx, xy;
  ltc_basis_x(wo, &xx, &xy);
  real yx = -xy, yy = xx;
  return v3(xx * v.x + xy * v.y, yx * v.x + yy * v.y, v.z);
}
inline V3 cltc_sample(V3 wo, real r, real u1, real u2, real* pdf) {
  real a, b, c, d;
  ltc_coeffs(wo.z, r, &a, &b, &c, &d);
  real radius = std::sqrt(u1);
  real phi = 2.0f * kPi * u2;
  real x0 = radius * std::cos(phi), y0 = radius * std::sin(phi);
  real vz = 1.0f / std::sqrt(d * d + 1.0f);
  real s = 0.5f * (1.0f + vz);
  real edge = std::sqrt(std::fmax(1.0f - y0 * y0, 0.0f));
  real x = -((1.0f - s) * edge + s * x0);
  real y = y0;
  real whz = std::sqrt(std::fmax(1.0f - (x * x + y * y), 0.0f));
  V3 wh = v3(x, y, whz);
  real pdf_wh = wh.z / (kPi * s);
  V3 wl = v3(a * wh.x + b * wh.z, c * wh.y, d * wh.x + wh.z);
  real len = std::sqrt(dot(wl, wl));
  real det = c * (a - b * d);
  *pdf = pdf_wh * (len * len * len) / det;
  return ltc_from(wo, sdiv(wl, len));
}
inline real cltc_pdf(V3 wo, V3 wi, real r) {
  V3 wl = ltc_to(wo, wi);
  real a, b, c, d;
  ltc_coeffs(wo.z, r, &a, &b, &c, &d);
  real det = c * (a - b * d);
  V3 wh = v3(c * (wl.x - b * wl.z), (a - b * d) * wl.y, -c * (d * wl.x - a * wl.z));
  real l2 = std::fmax(dot(wh, wh), 1.0e-30f);
  real vz = 1.0f / std::sqrt(d * d + 1.0f);
  real s = 0.5f * (1.0f + vz);
  real q = det / l2;
  return q * q * std::fmax(wh.z, 0.0f) / (kPi * s);
}
inline real p_uniform(real mu, real r) {
  return std::pow(r, 0.1f) * (0.162925f + (-0.372058f + (0.538233f - 0.290822f * mu) * mu) * mu);
}
inline V3 uniform_lobe_sample(real u1, real u2) {
  real st = std::sqrt(std::fmax(1.0f - u1 * u1, 0.0f));
  real phi = 2.0f * kPi * u2;
  return v3(st * std::cos(phi), st * std::sin(phi), u1);
}
inline V3 sample_eon(V3 wo, real r, real u1, real u2, real* pdf) {
  real pu = p_uniform(wo.z, r);
  real pc = 1.0f - pu;
  V3 wi;
  real pdf_c;
  if (u1 < pu) {
    real u1n = u1 / pu;
    wi = uniform_lobe_sample(u1n, u2);
    pdf_c = cltc_pdf(wo, wi, r);
  } else {
    real den = std::fmax(pc, 1.0e-30f);
    real u1n = (u1 - pu) / den;
    wi = cltc_sample(wo, r, u1n, u2, &pdf_c);
  }
  *pdf = pu * (1.0f / (2.0f * kPi)) + pc * pdf_c;
  return wi;
}
inline real pdf_eon(V3 wo, V3 wi, real r) {
  real pu = p_uniform(wo.z, r);
  return pu * (1.0f / (2.0f * kPi)) + (1.0f - pu) * cltc_pdf(wo, wi, r);
}


// ---- shared sampling helpers -----------------------------------------------------------
struct BSample {
  V3 weight, wi;
  real pdf;
  bool delta;
  bool glossy = false;  // BsdfFlags::GLOSSY: a rough coat, metal, specular, transmission or fuzz lobe sample
  bool transmission = false;  // BsdfFlags::TRANSMISSION
  real eta = 1.0f;           // BsdfSample::eta (relative IOR of a refraction, 1 otherwise)
  real cone_spread = 0.0f;  // MaterialSample::cone_spread (set by the material)
  int lobe = -1;            // OpenPBR: the picked lobe (bsdf_sample's numbering) and its probability
  real p_lobe = 0.0f;
};
// sample_cosine_weighted_hemisphere, R/math/mod.rs:26-36
inline V3 cosine_hemisphere(real u, real v) {
  real r = std::sqrt(clampf(u, 0.0f, 1.0f));
  real phi = kTau * v;
  real dx = r * std::cos(phi), dy = r * std::sin(phi);
  real z = std::sqrt(std::fmax(1.0f - (dx * dx + dy * dy), 0.0f));
  return v3(dx, dy, z);
}
inline real cosine_hemisphere_pdf(real c) { return std::fmax(c, 0.0f) / kPi; }  // R/math/mod.rs:38-40

// f32::powi with a run-time base: LLVM expands it by binary exponentiation
// (x^n = product of the squarings selected by the bits of n, lowest first).
inline real powi_rt(real x, int n) {
  real res = 0.0f, cur = x;
  bool have = false;
  while (n) {
    if (n & 1) {
      res = have ? res * cur : cur;
      have = true;
    }
    n >>= 1;
    if (n) cur = cur * cur;
  }
  return have ? res : 1.0f;
}
// f32::powi of compile-time constants, folded by LLVM through the double pow.
inline real powi_const(real x, int n) { return (real)std::pow((double)x, (double)n); }
inline V3 vec3_powf(V3 v, real e) {  // R/bsdf/open_pbr.rs:2179-2185
  return v3(std::pow(std::fmax(v.x, 0.0f), e), std::pow(std::fmax(v.y, 0.0f), e), std::pow(std::fmax(v.z, 0.0f), e));
}
inline V3 vclamp01(V3 v) { return v3(clampf(v.x, 0.f, 1.f), clampf(v.y, 0.f, 1.f), clampf(v.z, 0.f, 1.f)); }
inline V3 vsplat(real s) { return v3(s, s, s); }
inline V3 vlerp(V3 a, V3 b, real t) { return add(a, smul(sub(b, a), t)); }  // a + (b - a) * t
inline real lerpf(real a, real b, real t) { return a + (b - a) * t; }

// ---- GGX sampling, R/bsdf/smith_ggx.rs ---------------------------------------------------
inline real ggx_g1(V3 w, real ax, real ay) {  // :45-51
  real l = ggx_lambda(w, ax, ay);
  if (!std::isfinite(l)) return 0.0f;
  return 1.0f / (1.0f + l);
}
inline V3 reflect_local(V3 wo, V3 wm) {  // :12-14
  return normalize_or_zero3(add(neg(wo), smul(wm, 2.0f * dot(wo, wm))));
}
inline real pdf_wm_vndf(V3 wo, V3 wm, real ax, real ay) {  // :62-75
  if (!(wo.z > 0.0f) || wm.z <= 0.0f) return 0.0f;
  real d = ggx_d(wm, ax, ay), g1 = ggx_g1(wo, ax, ay), c = std::fmax(dot(wo, wm), 0.0f);
  if (d <= 0.0f || g1 <= 0.0f || c <= 0.0f) return 0.0f;
  return d * g1 * c / wo.z;
}
// Heitz 2018 VNDF sampling, :78-119
inline bool sample_wm_vndf(V3 wo, real ax, real ay, real u0, real u1, V3* out) {
  if (wo.z <= 0.0f) return false;
  ax = std::fmax(ax, kMinAlpha);
  ay = std::fmax(ay, kMinAlpha);
  V3 ws = normalize_or_zero3(v3(ax * wo.x, ay * wo.y, wo.z));
  if (len_sq(ws) == 0.0f) return false;
  real lensq = ws.x * ws.x + ws.y * ws.y;
  V3 t1 = lensq > 0.0f ? sdiv(v3(-ws.y, ws.x, 0.0f), std::sqrt(lensq)) : v3(1.f, 0.f, 0.f);
  V3 t2 = cross(ws, t1);
  real r = std::sqrt(clampf(u0, 0.0f, 1.0f));
  real phi = kTau * u1;
  real p1 = r * std::cos(phi);
  real p2 = r * std::sin(phi);
  real s = 0.5f * (1.0f + ws.z);
  p2 = (1.0f - s) * std::sqrt(std::fmax(1.0f - p1 * p1, 0.0f)) + s * p2;
  V3 nh = add(add(smul(t1, p1), smul(t2, p2)), smul(ws, std::sqrt(std::fmax(1.0f - p1 * p1 - p2 * p2, 0.0f))));
  V3 wm = normalize_or_zero3(v3(ax * nh.x, ay * nh.y, std::fmax(nh.z, 0.0f)));
  if (len_sq(wm) == 0.0f || wm.z <= 0.0f) return false;
  *out = wm;
  return true;
}
// bounded VNDF sampling (Eto and Tokuyoshi), :121-150
inline bool sample_wm_bounded_vndf(V3 wo, real ax, real ay, real u0, real u1, V3* out) {
  if (wo.z <= 0.0f) return false;
  ax = std::fmax(ax, kMinAlpha);
  ay = std::fmax(ay, kMinAlpha);
  V3 ws = normalize_or_zero3(v3(ax * wo.x, ay * wo.y, wo.z));
  if (len_sq(ws) == 0.0f) return false;
  real phi = kTau * u0;
  real a = clampf(std::fmin(ax, ay), 0.0f, 1.0f);
  real s = 1.0f + std::sqrt(wo.x * wo.x + wo.y * wo.y);
  real a2 = a * a, s2 = s * s;
  real k = (1.0f - a2) * s2 / (s2 + a2 * wo.z * wo.z);
  real lb = wo.z > 0.0f ? -k * ws.z : -ws.z;
  real z = std::fma(lb, u1, 1.0f - u1);  // f32::mul_add
  real st = std::sqrt(clampf(1.0f - z * z, 0.0f, 1.0f));
  V3 os = v3(st * std::cos(phi), st * std::sin(phi), z);
  V3 ms = add(ws, os);
  V3 wm = normalize_or_zero3(v3(ax * ms.x, ay * ms.y, ms.z));
  if (len_sq(wm) == 0.0f || wm.z <= 0.0f) return false;
  *out = wm;
  return true;
}
// refract_about_wm, R/bsdf/dielectric_ggx.rs:723-738
inline bool refract_about_wm(V3 wo, V3 wm, real eta, V3* out) {
  real co = dot(wo, wm);
  if (co <= 0.0f) return false;
  real s2t = eta * eta * std::fmax(1.0f - co * co, 0.0f);
  if (s2t >= 1.0f) return false;
  real ct = std::sqrt(std::fmax(1.0f - s2t, 0.0f));
  V3 wi = normalize_or_zero3(add(smul(wo, -eta), smul(wm, eta * co - ct)));
  if (len_sq(wi) == 0.0f) return false;
  *out = wi;
  return true;
}

// ---- LUTs, R/bsdf/directional_albedo.rs ------------------------------------------------
constexpr real kUniformHemiPdf = 1.0f / (2.0f * kPi);  // :30
inline real radical_inverse_vdc(uint32_t bits) {       // :343-345
  uint32_t r = 0;
  for (int i = 0; i < 32; ++i) r |= ((bits >> i) & 1u) << (31 - i);
  return (real)r * 2.3283064e-10f;
}
inline real radical_inverse_base3(uint32_t n) {  // :347-356
  real q = 0.0f, bk = 1.0f / 3.0f;
  while (n > 0) {
    q += (real)(n % 3u) * bk;
    n /= 3u;
    bk /= 3.0f;
  }
  return q;
}
inline real lut_cell_center(uint32_t i, uint32_t res) { return ((real)i + 0.5f) / (real)res; }  // :410-412
struct AxisLerp {
  uint32_t lo, hi;
  real t;
};
inline AxisLerp axis_lerp(real value, uint32_t res) {  // compute_lut_axis_lerp :388-408
  if (res == 1) return AxisLerp{0, 0, 0.0f};
  real pos = clampf(value, 0.0f, 1.0f) * (real)res - 0.5f;
  if (pos <= 0.0f) return AxisLerp{0, 0, 0.0f};
  uint32_t mx = res - 1;
  if (pos >= (real)mx) return AxisLerp{mx, mx, 0.0f};
  uint32_t lo = (uint32_t)std::floor(pos);
  return AxisLerp{lo, lo + 1, pos - (real)lo};
}
// alpha_xy_from_roughness of the LUT module (:358-370)
inline void lut_alpha_xy(real roughness, real anisotropy, real* ax, real* ay) {
  roughness = clampf(roughness, 0.0f, 1.0f);
  anisotropy = clampf(anisotropy, -1.0f, 1.0f);
  real alpha = roughness * roughness;
  real aspect = std::sqrt(1.0f - 0.9f * std::fabs(anisotropy));
  real x, y;
  if (anisotropy >= 0.0f) {
    x = alpha / aspect;
    y = alpha * aspect;
  } else {
    x = alpha * aspect;
    y = alpha / aspect;
  }
  *ax = clampf(x, kMinAlpha, 1.0f);
  *ay = clampf(y, kMinAlpha, 1.0f);
}
inline real sanitize_eta(real eta) { return (std::isfinite(eta) && eta > 0.0f) ? eta : 1.5f; }  // :372-378

// Reflection-only DielectricGgxBsdf::eval (R/bsdf/dielectric_ggx.rs:121-176),
// colour 1, front face, no energy compensation.
inline real dielectric_reflect_eval(real eta, real ax, real ay, V3 wo, V3 wi) {
  ax = std::fmax(ax, kMinAlpha);
  ay = std::fmax(ay, kMinAlpha);
  if (!(wo.z > 0.0f) || eta <= 0.0f) return 0.0f;
  if (std::fmax(ax, ay) < kSmoothAlpha) return 0.0f;
  if (!(wo.z * wi.z > 0.0f)) return 0.0f;
  V3 wm;
  if (!reflection_half_vector(wo, wi, &wm)) return 0.0f;
  real cwm = dot(wo, wm);
  if (cwm <= 0.0f) return 0.0f;
  real co = std::fabs(wo.z), ci = std::fabs(wi.z);
  if (co <= 0.0f || ci <= 0.0f) return 0.0f;
  real d = ggx_d(wm, ax, ay);
  if (d <= 0.0f) return 0.0f;
  real g = ggx_g2(wo, wi, ax, ay);
  if (g <= 0.0f) return 0.0f;
  real f = fresnel_dielectric(cwm, 1.0f, eta);
  return d * g * f / (4.0f * co * ci);
}

// DielectricGgxDirectionalAlbedoLut (:128-290): sqrt(cos) x phi x roughness x
// anisotropy = 64 x 16 x 64 x 16 cells.  OpenPBR looks it up at anisotropy 0
// only (R/bsdf/open_pbr.rs:431,446), which interpolates slices 7 and 8 with
// weight 1/2 each, so those two slices are built (kAlbedoAniso) and stored as
// slice 0 and 1.
constexpr uint32_t kAlbedoCos = 64, kAlbedoPhi = 16, kAlbedoRough = 64, kAlbedoAnisoRes = 16, kAlbedoAniso = 2;
constexpr uint32_t kAlbedoLen = kAlbedoCos * kAlbedoPhi * kAlbedoRough * kAlbedoAniso;
constexpr uint32_t kAlbedoSamples = 64;
struct AlbedoLut {
  real eta = 1.5f;
  std::vector<real> v;  // [((cos * 16 + phi) * 64 + rough) * 2 + (aniso - 7)]
};
template <class F>
inline void parallel_for(uint32_t n, F f) {
  unsigned th = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < th; ++t)
    pool.emplace_back([&, t] {
      for (uint32_t i = t; i < n; i += th) f(i);
    });
  for (auto& p : pool) p.join();
}
inline AlbedoLut build_albedo_lut(real eta_in) {
  AlbedoLut L;
  L.eta = sanitize_eta(eta_in);
  const real eta = L.eta;
  V3 hs[kAlbedoSamples];  // generate_uniform_hemisphere_samples :327-341
  for (uint32_t i = 0; i < kAlbedoSamples; ++i) {
    real u = ((real)i + 0.5f) / (real)kAlbedoSamples, vv = radical_inverse_vdc(i);
    real z = clampf(u, 0.0f, 1.0f), r = std::sqrt(std::fmax(1.0f - z * z, 0.0f)), phi = kTau * vv;
    hs[i] = v3(r * std::cos(phi), r * std::sin(phi), z);
  }
  L.v.assign(kAlbedoLen, 0.0f);
  parallel_for(kAlbedoLen, [&](uint32_t idx) {
    uint32_t an = idx % kAlbedoAniso, rest = idx / kAlbedoAniso;
    uint32_t ri = rest % kAlbedoRough;
    rest /= kAlbedoRough;
    uint32_t pi = rest % kAlbedoPhi, ci = rest / kAlbedoPhi;
    real sc = lut_cell_center(ci, kAlbedoCos);
    real ct = sc * sc, st = std::sqrt(std::fmax(1.0f - ct * ct, 0.0f));
    real phi = lut_cell_center(pi, kAlbedoPhi) * (kPi / 2.0f);
    V3 wo = v3(st * std::cos(phi), st * std::sin(phi), ct);
    real rough = lut_cell_center(ri, kAlbedoRough);
    real aniso = 2.0f * lut_cell_center(an + 7, kAlbedoAnisoRes) - 1.0f;
    // estimate_dielectric_ggx_directional_albedo :292-325
    real val;
    if (wo.z <= 0.0f) {
      val = 0.0f;
    } else {
      real ax, ay;
      lut_alpha_xy(rough, aniso, &ax, &ay);
      if (std::fmax(ax, ay) < kSmoothAlpha) {
        val = fresnel_dielectric(clampf(wo.z, 0.0f, 1.0f), 1.0f, eta);
      } else {
        real sum = 0.0f;
        for (uint32_t s = 0; s < kAlbedoSamples; ++s)
          sum += dielectric_reflect_eval(eta, ax, ay, wo, hs[s]) * std::fmax(hs[s].z, 0.0f) / kUniformHemiPdf;
        val = clampf(sum / (real)kAlbedoSamples, 0.0f, 1.0f);
      }
    }
    L.v[idx] = val;
  });
  return L;
}
inline real fold_phi(V3 w) {  // fold_phi_to_first_quadrant :380-386
  if (w.x == 0.0f && w.y == 0.0f) return 0.0f;
  return std::atan2(std::fabs(w.y), std::fabs(w.x));
}
// DielectricGgxDirectionalAlbedoLut::lookup (:170-226) at anisotropy 0.
inline real albedo_lookup(const AlbedoLut& L, V3 w, real roughness) {
  if (w.z <= 0.0f) return 0.0f;
  roughness = clampf(roughness, 0.0f, 1.0f);
  real ax, ay;
  lut_alpha_xy(roughness, 0.0f, &ax, &ay);
  if (std::fmax(ax, ay) < kSmoothAlpha) return fresnel_dielectric(clampf(w.z, 0.0f, 1.0f), 1.0f, L.eta);
  real sc = std::sqrt(clampf(w.z, 0.0f, 1.0f));
  real phi = fold_phi(w) / (kPi / 2.0f);
  AxisLerp a = axis_lerp(sc, kAlbedoCos), b = axis_lerp(phi, kAlbedoPhi), c = axis_lerp(roughness, kAlbedoRough);
  AxisLerp d = axis_lerp(0.5f, kAlbedoAnisoRes);  // slices 7, 8 -> stored 0, 1
  uint32_t ia[2] = {a.lo, a.hi}, ib[2] = {b.lo, b.hi}, ic[2] = {c.lo, c.hi}, id[2] = {d.lo - 7, d.hi - 7};
  real wa[2] = {1.0f - a.t, a.t}, wb[2] = {1.0f - b.t, b.t}, wc[2] = {1.0f - c.t, c.t}, wd[2] = {1.0f - d.t, d.t};
  real value = 0.0f;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k)
        for (int l = 0; l < 2; ++l)
          value += wa[i] * wb[j] * wc[k] * wd[l] *
                   L.v[((ia[i] * kAlbedoPhi + ib[j]) * kAlbedoRough + ic[k]) * kAlbedoAniso + id[l]];
  return clampf(value, 0.0f, 1.0f);
}

// ConductorGgxEnergyCompensationLut (:694-776): 64 x 64 directional albedo of
// a white-Fresnel GGX conductor, plus its cosine-weighted average.
constexpr uint32_t kCondCos = 64, kCondRough = 64, kCondSamples = 256;
struct CondLut {
  std::vector<real> e, eavg;
};
inline CondLut build_cond_lut() {
  CondLut L;
  L.e.assign(kCondCos * kCondRough, 0.0f);
  real us[kCondSamples][2];  // generate_stratified_unit_square_samples_ec :823-829
  for (uint32_t i = 0; i < kCondSamples; ++i) {
    us[i][0] = ((real)i + 0.5f) / (real)kCondSamples;
    us[i][1] = radical_inverse_vdc(i);
  }
  parallel_for(kCondCos * kCondRough, [&](uint32_t idx) {
    uint32_t ci = idx % kCondCos, ri = idx / kCondCos;
    real sc = lut_cell_center(ci, kCondCos);
    real ct = sc * sc;
    real rough = lut_cell_center(ri, kCondRough);
    // estimate_conductor_ggx_directional_albedo_white_fresnel :778-821
    real val;
    real alpha = clampf(rough * rough, kMinAlpha, 1.0f);
    if (ct <= 0.0f) {
      val = 0.0f;
    } else if (alpha < kSmoothAlpha) {
      val = 1.0f;
    } else {
      real st = std::sqrt(std::fmax(1.0f - ct * ct, 0.0f));
      V3 wo = v3(st, 0.0f, ct);
      real g1 = ggx_g1(wo, alpha, alpha);
      if (g1 <= 0.0f) {
        val = 0.0f;
      } else {
        real sum = 0.0f;
        for (uint32_t s = 0; s < kCondSamples; ++s) {
          V3 wm;
          real term = 0.0f;
          if (sample_wm_vndf(wo, alpha, alpha, us[s][0], us[s][1], &wm)) {
            real c = dot(wo, wm);
            if (c > 0.0f) {
              V3 wi = reflect_local(wo, wm);
              if (wi.z > 0.0f) term = ggx_g2(wo, wi, alpha, alpha) / g1;
            }
          }
          sum += term;
        }
        val = clampf(sum / (real)kCondSamples, 0.0f, 1.0f);
      }
    }
    L.e[idx] = val;
  });
  L.eavg.assign(kCondRough, 0.0f);
  for (uint32_t r = 0; r < kCondRough; ++r) {
    real w = 0.0f;
    for (uint32_t c = 0; c < kCondCos; ++c) {
      real sc = lut_cell_center(c, kCondCos), ct = sc * sc;
      w += L.e[r * kCondCos + c] * ct * sc;
    }
    L.eavg[r] = clampf(4.0f / (real)kCondCos * w, 0.0f, 1.0f);
  }
  return L;
}
inline real cond_lookup_e(const CondLut& L, real c, real r) {  // :738-754
  c = clampf(c, 0.0f, 1.0f);
  r = clampf(r, 0.0f, 1.0f);
  AxisLerp a = axis_lerp(std::sqrt(c), kCondCos), b = axis_lerp(r, kCondRough);
  uint32_t ia[2] = {a.lo, a.hi}, ib[2] = {b.lo, b.hi};
  real wa[2] = {1.0f - a.t, a.t}, wb[2] = {1.0f - b.t, b.t};
  real value = 0.0f;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) value += wa[i] * wb[j] * L.e[ib[j] * kCondCos + ia[i]];
  return clampf(value, 0.0f, 1.0f);
}
inline real cond_lookup_eavg(const CondLut& L, real r) {  // :756-762
  r = clampf(r, 0.0f, 1.0f);
  AxisLerp a = axis_lerp(r, kCondRough);
  return clampf((1.0f - a.t) * L.eavg[a.lo] + a.t * L.eavg[a.hi], 0.0f, 1.0f);
}

// DielectricGgxEnergyCompensationLut (:832-990): 32^3 full-sphere albedo of a
// rough dielectric over (sqrt(cos), roughness, eta axis), plus averages.
constexpr uint32_t kDielCos = 32, kDielRough = 32, kDielEta = 32, kDielSamples = 4096;
struct DielLut {
  std::vector<real> e, eavg;
};
inline real diel_axis_to_eta(real t) {  // :855-863
  t = clampf(t, 0.0f, 1.0f);
  if (t >= 0.5f) return 1.0f + (t - 0.5f) * 2.0f * (3.0f - 1.0f);
  real inv = 1.0f + (0.5f - t) * 2.0f * (3.0f - 1.0f);
  return 1.0f / inv;
}
inline real diel_eta_to_axis(real eta) {  // :845-853
  eta = clampf(eta, 1.0f / 3.0f, 3.0f);
  if (eta >= 1.0f) return 0.5f + 0.5f * ((eta - 1.0f) / (3.0f - 1.0f));
  real inv = 1.0f / eta;
  return 0.5f - 0.5f * ((inv - 1.0f) / (3.0f - 1.0f));
}
// DielectricGgxBsdf::sample_rough (R/bsdf/dielectric_ggx.rs:501-587), front
// face, colour 1, no energy compensation, radiance mode: the weight's x and
// whether the sample transmits.
inline bool dielectric_sample_rough(real eta, real alpha, V3 wo, real uc, real u0, real u1, real* w, bool* trans) {
  const real ei = 1.0f, et = eta, eta_rel = 1.0f / eta;
  V3 wm;
  if (!sample_wm_vndf(wo, alpha, alpha, u0, u1, &wm)) return false;
  real cwm = dot(wo, wm);
  if (cwm <= 0.0f) return false;
  real f = fresnel_dielectric(cwm, ei, et);
  real g1 = ggx_g1(wo, alpha, alpha);
  if (g1 <= 0.0f) return false;
  real pr = std::fmax(f, 0.0f), pt = std::fmax(1.0f - f, 0.0f);
  real psum = pr + pt;
  if (psum <= 0.0f) return false;
  pr /= psum;
  pt /= psum;
  if (uc < pr) {
    V3 wi = reflect_local(wo, wm);
    if (!(wi.z > 0.0f)) return false;
    real g2 = ggx_g2(wo, wi, alpha, alpha);
    real pw = pdf_wm_vndf(wo, wm, alpha, alpha);
    if (pw <= 0.0f) return false;
    real pdf = pr * pw / (4.0f * cwm);
    if (pdf <= 0.0f) return false;
    *w = f * g2 / (pr * g1);
    *trans = false;
    return true;
  }
  V3 wi;
  if (!refract_about_wm(wo, wm, eta_rel, &wi)) return false;
  if (wi.z >= 0.0f) return false;
  real cwi = dot(wi, wm);
  real den = cwi + eta_rel * cwm;
  if (std::fabs(den) < 1.0e-6f) return false;
  real g2 = ggx_g2(wo, wi, alpha, alpha);
  real pw = pdf_wm_vndf(wo, wm, alpha, alpha);
  if (pw <= 0.0f) return false;
  real pdf = pt * pw * std::fabs(cwi) / (den * den);
  if (pdf <= 0.0f) return false;
  real rs = 1.0f / (eta_rel * eta_rel);  // TransportMode::Radiance.transmission_scale
  *w = 1.0f * (rs * (1.0f - f) * g2 / (pt * g1));
  *trans = true;
  return true;
}
inline DielLut build_diel_lut() {
  DielLut L;
  L.e.assign(kDielCos * kDielRough * kDielEta, 0.0f);
  parallel_for(kDielCos * kDielRough * kDielEta, [&](uint32_t idx) {
    uint32_t ci = idx % kDielCos, rest = idx / kDielCos;
    uint32_t ri = rest % kDielRough, ei = rest / kDielRough;
    real sc = lut_cell_center(ci, kDielCos), ct = sc * sc;
    real rough = lut_cell_center(ri, kDielRough);
    real eta = diel_axis_to_eta(lut_cell_center(ei, kDielEta));
    // estimate_dielectric_ggx_directional_albedo_full_sphere :975-1023
    real val;
    if (ct <= 0.0f || eta <= 0.0f) {
      val = 0.0f;
    } else {
      real alpha = clampf(rough * rough, kMinAlpha, 1.0f);
      if (alpha < kSmoothAlpha) {
        val = 1.0f;
      } else {
        real st = std::sqrt(std::fmax(1.0f - ct * ct, 0.0f));
        V3 wo = v3(st, 0.0f, ct);
        real inv_eta_sq = 1.0f / (eta * eta);
        real sum = 0.0f;
        for (uint32_t s = 0; s < kDielSamples; ++s) {
          real uc = ((real)s + 0.5f) / (real)kDielSamples;
          real u = radical_inverse_vdc(s + 1), v = radical_inverse_base3(s + 1);
          real w;
          bool tr;
          if (!dielectric_sample_rough(eta, alpha, wo, uc, u, v, &w, &tr)) continue;
          if (!std::isfinite(w)) continue;
          sum += w * (tr ? inv_eta_sq : 1.0f);
        }
        val = clampf(sum / (real)kDielSamples, 0.0f, 1.0f);
      }
    }
    L.e[idx] = val;
  });
  L.eavg.assign(kDielRough * kDielEta, 0.0f);
  for (uint32_t e = 0; e < kDielEta; ++e)
    for (uint32_t r = 0; r < kDielRough; ++r) {
      real w = 0.0f;
      for (uint32_t c = 0; c < kDielCos; ++c) {
        real sc = lut_cell_center(c, kDielCos), ct = sc * sc;
        w += L.e[(e * kDielRough + r) * kDielCos + c] * ct * sc;
      }
      L.eavg[e * kDielRough + r] = clampf(4.0f / (real)kDielCos * w, 0.0f, 1.0f);
    }
  return L;
}
inline real diel_lookup_e(const DielLut& L, real c, real r, real eta) {  // :919-944
  c = clampf(c, 0.0f, 1.0f);
  r = clampf(r, 0.0f, 1.0f);
  AxisLerp a = axis_lerp(std::sqrt(c), kDielCos), b = axis_lerp(r, kDielRough), e = axis_lerp(diel_eta_to_axis(eta), kDielEta);
  uint32_t ie[2] = {e.lo, e.hi}, ib[2] = {b.lo, b.hi}, ia[2] = {a.lo, a.hi};
  real we[2] = {1.0f - e.t, e.t}, wb[2] = {1.0f - b.t, b.t}, wa[2] = {1.0f - a.t, a.t};
  real value = 0.0f;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) value += we[i] * wb[j] * wa[k] * L.e[(ie[i] * kDielRough + ib[j]) * kDielCos + ia[k]];
  return clampf(value, 0.0f, 1.0f);
}
inline real diel_lookup_eavg(const DielLut& L, real r, real eta) {  // :946-961
  r = clampf(r, 0.0f, 1.0f);
  AxisLerp b = axis_lerp(r, kDielRough), e = axis_lerp(diel_eta_to_axis(eta), kDielEta);
  uint32_t ie[2] = {e.lo, e.hi}, ib[2] = {b.lo, b.hi};
  real we[2] = {1.0f - e.t, e.t}, wb[2] = {1.0f - b.t, b.t};
  real value = 0.0f;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) value += we[i] * wb[j] * L.eavg[ie[i] * kDielRough + ib[j]];
  return clampf(value, 0.0f, 1.0f);
}

// The tables one material reads.  A table a material does not need is an
// all-zero stand-in that is only ever read with weight zero: the conductor
// table is read with a non-zero weight only for metalness > 0, the dielectric
// energy-compensation table only where f_avg(eta) > 0 (eta != 1) or for a
// coat, and the coat albedo table only for coat > 0.
struct Luts {
  const AlbedoLut* spec = nullptr;
  const AlbedoLut* coat = nullptr;
  const CondLut* cond = nullptr;
  const DielLut* diel = nullptr;
};
// Process-wide tables, built once (the reference caches them per process,
// R/bsdf/directional_albedo.rs:37-116).
const AlbedoLut& albedo_lut_for(real eta);
const CondLut& cond_lut();
const DielLut& diel_lut();
const AlbedoLut& zero_albedo_lut();
const CondLut& zero_cond_lut();
const DielLut& zero_diel_lut();

// ---- OpenPBR material, R/material/open_pbr.rs -----------------------------------------------
// OpenPbrMaterial parameters (:23-82) carried by the backend; dispersion,
// thin film, subsurface and normal maps are rejected.
constexpr uint32_t kNoTexture = 0xFFFFFFFFu;
struct OpenPbr {
  real base_weight;
  V3 base_color;
  real base_diffuse_roughness;
  real base_metalness;
  real specular_weight;
  V3 specular_color;
  real specular_roughness;
  real specular_ior;
  real specular_roughness_anisotropy;
  real transmission_weight, coat_weight, fuzz_weight, thin_film_weight;
  real emission_luminance;
  V3 emission_color;
  uint32_t thin_walled;
  V3 coat_color;
  real coat_roughness, coat_roughness_anisotropy, coat_ior, coat_darkening;
  V3 fuzz_color;
  real fuzz_roughness;
  // texture slots (indices into the scene's textures, kNoTexture for none):
  // base_color, base_metalness, specular_roughness, fuzz_weight,
  // fuzz_roughness, emission_color (R/material/open_pbr.rs:66-72)
  uint32_t tex[6] = {kNoTexture, kNoTexture, kNoTexture, kNoTexture, kNoTexture, kNoTexture};
  V3 transmission_color = {1.0f, 1.0f, 1.0f};
  real transmission_dispersion_scale = 0.0f, transmission_dispersion_abbe_number = 20.0f;
  real thin_film_thickness = 0.5f, thin_film_ior = 1.4f;
};
enum TexSlot { kTexBaseColor = 0, kTexMetalness, kTexSpecRough, kTexFuzzWeight, kTexFuzzRough, kTexEmission };

inline real modulated_eta(real eta, real w) {  // :871-878
  if (std::fabs(w - 1.0f) < 1.0e-6f) return eta;
  real f0 = powi_rt((eta - 1.0f) / (eta + 1.0f), 2);
  real sg = (eta - 1.0f) > 0.0f ? 1.0f : ((eta - 1.0f) < 0.0f ? -1.0f : 1.0f);  // f32::signum (+0 -> +1)
  real eps = std::sqrt(clampf(w * f0, 0.0f, 0.99999f)) * sg;
  return std::fmax((1.0f + eps) / std::fmax(1.0f - eps, 1.0e-6f), 1.0e-4f);
}
inline void alpha_xy(real roughness, real anisotropy, real* ax, real* ay) {  // :861-869
  real r = clampf(roughness, 0.0f, 1.0f);
  real a = clampf(anisotropy, 0.0f, 1.0f);
  real alpha = r * r;
  real oma = 1.0f - a;
  real x = alpha * std::sqrt(2.0f / (1.0f + oma * oma));
  real y = oma * x;
  *ax = clampf(x, kMinAlpha, 1.0f);
  *ay = clampf(y, kMinAlpha, 1.0f);
}
// zeltner_dir_albedo / LTC coefficients, R/bsdf/open_pbr.rs:2070-2091 (the
// material module's copy :890-898 is the same expression)
inline real zeltner_dir_albedo(real ndv, real rough) {
  real x = clampf(ndv, 0.0f, 1.0f), y = clampf(rough, 0.01f, 1.0f);
  real s = y * (0.0206607f + 1.58491f * y) / (0.0379424f + y * (1.32227f + y));
  real m = y * (-0.193854f + y * (-1.14885f + y * (1.7932f - 0.95943f * y * y))) / (0.046391f + y);
  real o = y * (0.000654023f + (-0.0207818f + 0.119681f * y) * y) / (1.26264f + y * (-1.92021f + y));
  real q = (x - m) / s;
  return std::exp(-0.5f * (q * q)) / (s * std::sqrt(2.0f * kPi)) + o;
}
inline real zeltner_a_inv(real ndv, real rough) {
  real x = clampf(ndv, 0.0f, 1.0f), y = clampf(rough, 0.01f, 1.0f);
  return (2.58126f * x + 0.813703f * y) * y / (1.0f + 0.310327f * x * x + 2.60994f * x * y);
}
inline real zeltner_b_inv(real ndv, real rough) {
  real x = clampf(ndv, 0.0f, 1.0f), y = clampf(rough, 0.01f, 1.0f);
  return std::sqrt(1.0f - x) * (y - 1.0f) * y * y * y / (0.0000254053f + 1.71228f * x - 1.71506f * x * y + 1.34174f * y * y);
}
inline real dielectric_fresnel_avg_ln(real eta) {  // dielectric_fresnel_avg :880-886
  eta = clampf(eta, 1.0f, 3.0f);
  return clampf(std::log(std::fmax((10893.0f * eta - 1438.2f) / (-774.4f * eta * eta + 10212.0f * eta + 1.0f), 1.0e-6f)),
                0.0f, 1.0f);
}

// OpenPbrBsdfParams (R/bsdf/open_pbr.rs:33-71) of the carried subset.
struct PbrBsdf {
  V3 base_color;
  real base, specular;
  V3 specular_color;
  real ax, ay, specular_eta, transmission_eta, metalness;
  real coat;
  V3 coat_color, coat_darkening;
  real cax, cay, coat_eta;
  real fuzz;
  V3 fuzz_color;
  real fuzz_roughness;
  real diffuse_roughness;
  real transmission;     // transmission_weight clamped (the BTDF reads the specular alphas and eta)
  V3 transmission_color;  // clamped to [0, 1]
  bool thin_walled;
  bool front_face;
  Luts luts;
  real spec_roughness;  // effective_specular_roughness (the sample's cone spread reads it)
  // thin film (OpenPbrBsdfParams :58-65): weight in [0, 1], thickness in nm
  // (0 without a film), film IOR; the metal's complex IOR
  // (artist_friendly_complex_ior of base_color and specular_color)
  real film_weight = 0.0f, film_thickness = 0.0f, film_ior = 1.4f;
  V3 metal_n = {0.0f, 0.0f, 0.0f}, metal_k = {0.0f, 0.0f, 0.0f};
};

// ---- thin film (R/bsdf/thin_film.rs, R/bsdf/conductor_complex.rs:132-147) -----------------
// artist_friendly_complex_ior (conductor_complex.rs:132-147): Gulbrandsen's
// n, k from the face (normal-incidence) and edge colours
inline void artist_friendly_complex_ior(V3 face, V3 edge, V3* n_out, V3* k_out) {
  const real c[3] = {face.x, face.y, face.z}, e[3] = {edge.x, edge.y, edge.z};
  real n[3], k[3];
  for (int i = 0; i < 3; ++i) {
    const real r = clampf(c[i], 0.0f, 0.999f), g = clampf(e[i], 0.0f, 1.0f);
    const real rs = std::sqrt(std::fmax(r, 0.0f));
    const real n_max = (1.0f + rs) / std::fmax(1.0f - rs, 1.0e-6f);
    const real n_min = (1.0f - r) / (1.0f + r);
    n[i] = g * n_min + (1.0f - g) * n_max;
    const real np1 = n[i] + 1.0f, nm1 = n[i] - 1.0f;
    const real k2 = std::fmax((np1 * np1 * r - nm1 * nm1) / std::fmax(1.0f - r, 1.0e-6f), 0.0f);
    k[i] = std::sqrt(std::fmax(k2, 0.0f));
  }
  *n_out = v3(n[0], n[1], n[2]);
  *k_out = v3(k[0], k[1], k[2]);
}
// fresnel_dielectric_polarized (thin_film.rs:143-156): |r_s|, |r_p| and
// their phases (pi where the amplitude is negative); total internal
// reflection gives (1, 1) with phase 0
inline void film_fresnel_dielectric(real cos_i, real eta_i, real eta_t, real r[2], real phi[2]) {
  cos_i = clampf(cos_i, 0.0f, 1.0f);
  const real sin_i_sq = std::fmax(1.0f - cos_i * cos_i, 0.0f);
  const real ratio = eta_i / eta_t;
  const real sin_t_sq = ratio * ratio * sin_i_sq;
  if (sin_t_sq >= 1.0f) {
    r[0] = r[1] = 1.0f;
    phi[0] = phi[1] = 0.0f;
    return;
  }
  const real cos_t = std::sqrt(std::fmax(1.0f - sin_t_sq, 0.0f));
  const real rs = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
  const real rp = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
  phi[0] = rs < 0.0f ? kPi : 0.0f;
  phi[1] = rp < 0.0f ? kPi : 0.0f;
  r[0] = std::fabs(rs);
  r[1] = std::fabs(rp);
}
// fresnel_complex_polarized (thin_film.rs:193-244): film (eta_i) into a
// base of complex IOR eta_t + i kappa_t: amplitudes and phases per
// polarisation
inline void film_fresnel_complex(real cos_i, real eta_i, real eta_t, real kappa_t, real* rs_out, real* rp_out,
                                 real* phs_out, real* php_out) {
  cos_i = clampf(cos_i, 0.0f, 1.0f);
  const real cos_i2 = cos_i * cos_i;
  const real sin_i2 = std::fmax(1.0f - cos_i2, 0.0f);
  const real eta_r = eta_t / std::fmax(eta_i, 1.0e-6f);
  const real kappa_r = kappa_t / std::fmax(eta_i, 1.0e-6f);
  const real eta_r2 = eta_r * eta_r, kappa_r2 = kappa_r * kappa_r;
  const real inner = eta_r2 - kappa_r2 - sin_i2;
  const real radicand = std::fmax(inner * inner + 4.0f * eta_r2 * kappa_r2, 0.0f);
  const real a2_plus_b2 = std::sqrt(radicand);
  const real a2 = std::fmax(0.5f * (a2_plus_b2 + inner), 0.0f);
  const real b2 = std::fmax(0.5f * (a2_plus_b2 - inner), 0.0f);
  const real a = std::sqrt(a2);
  const real t1 = a2_plus_b2 + cos_i2, t2 = 2.0f * a * cos_i;
  const real r_s2 = (t1 - t2) / std::fmax(t1 + t2, 1.0e-12f);
  const real t3 = a2_plus_b2 * cos_i2 + sin_i2 * sin_i2, t4 = t2 * sin_i2;
  const real r_p2 = r_s2 * (t3 - t4) / std::fmax(t3 + t4, 1.0e-12f);
  *rs_out = std::sqrt(std::fmax(r_s2, 0.0f));
  *rp_out = std::sqrt(std::fmax(r_p2, 0.0f));
  *phs_out = std::atan2(2.0f * std::sqrt(b2) * cos_i, cos_i2 - a2 - b2);
  if (r_p2 > 0.0f) {
    const real p_num = 2.0f * cos_i * (std::sqrt(b2) * (eta_r2 + kappa_r2) - 2.0f * a * eta_r * kappa_r);
    const real p_den = (eta_r2 + kappa_r2) * (eta_r2 + kappa_r2) * cos_i2 - (a2 + b2);
    *php_out = std::atan2(p_num, p_den);
  } else {
    *php_out = 0.0f;
  }
}
// eval_sensitivity (thin_film.rs:111-141): the Fourier transform of the
// CIE XYZ matching functions (Gaussian fits) at the optical path
// difference, with a phase shift per channel, in linear Rec.709
inline V3 film_sensitivity(real opd, V3 shift) {
  const real phase = 2.0f * kPi * opd * 1.0e-9f;
  const real val[3] = {5.4856e-13f, 4.4201e-13f, 5.2481e-13f};
  const real pos[3] = {1.6810e+06f, 1.7953e+06f, 2.2084e+06f};
  const real var[3] = {4.3278e+09f, 9.3046e+09f, 6.6121e+09f};
  const real sh[3] = {shift.x, shift.y, shift.z};
  real xyz[3];
  for (int i = 0; i < 3; ++i) {
    const real amp = std::sqrt(std::fmax(var[i] * (2.0f * kPi), 0.0f));
    xyz[i] = val[i] * amp * std::cos(pos[i] * phase + sh[i]) * std::exp(-phase * phase * var[i]);
  }
  const real extra_var = 4.5282e+09f;
  xyz[0] += 9.7470e-14f * std::sqrt(2.0f * kPi * extra_var) * std::cos(2.2399e+06f * phase + sh[0]) *
            std::exp(-extra_var * phase * phase);
  for (int i = 0; i < 3; ++i) xyz[i] /= 1.0685e-7f;
  // XYZ_TO_REC709 (thin_film.rs:5-9), columns applied in order
  const real m[3][3] = {{3.2404542f, -0.969266f, 0.0556434f},
                        {-1.5371385f, 1.8750108f, -0.2040259f},
                        {-0.4985314f, 0.041556f, 1.0572252f}};
  real out[3];
  for (int r = 0; r < 3; ++r) out[r] = m[0][r] * xyz[0] + m[1][r] * xyz[1] + m[2][r] * xyz[2];
  return v3(out[0], out[1], out[2]);
}
// eval_thin_film_inner (thin_film.rs:47-109): the film's reflectance for an
// incident cosine from the outside medium, the film IOR, the base's
// (complex) IOR per channel and the thickness in nm: the incoherent
// two-interface sum plus the first two interference orders, per
// polarisation, averaged and clamped to [0, 1].
inline V3 thin_film(real cos_theta_i, real eta_outside, real eta_film, V3 n_base, V3 k_base, real thickness_nm) {
  const real cos1 = clampf(cos_theta_i, 0.0f, 1.0f);
  eta_outside = std::fmax(eta_outside, 1.0e-3f);
  eta_film = std::fmax(eta_film, 1.0e-3f);
  const real eta_ratio = eta_outside / eta_film;
  const real sin1_sq = std::fmax(1.0f - cos1 * cos1, 0.0f);
  const real sin2_sq = eta_ratio * eta_ratio * sin1_sq;
  if (sin2_sq >= 1.0f) return vsplat(1.0f);
  const real cos2 = std::sqrt(std::fmax(1.0f - sin2_sq, 0.0f));
  real r12[2], phi12[2];
  film_fresnel_dielectric(cos1, eta_outside, eta_film, r12, phi12);
  const real phi21[2] = {kPi - phi12[0], kPi - phi12[1]};
  const real r12_sq[2] = {r12[0] * r12[0], r12[1] * r12[1]};
  const real t121_sq[2] = {std::fmax(1.0f - r12_sq[0], 0.0f), std::fmax(1.0f - r12_sq[1], 0.0f)};
  const real nb[3] = {n_base.x, n_base.y, n_base.z}, kb[3] = {k_base.x, k_base.y, k_base.z};
  real r23_sq[2][3], phi23[2][3];
  for (int c = 0; c < 3; ++c) {
    real rs, rp, ps, pp;
    film_fresnel_complex(cos2, eta_film, nb[c], kb[c], &rs, &rp, &ps, &pp);
    r23_sq[0][c] = rs * rs;
    r23_sq[1][c] = rp * rp;
    phi23[0][c] = ps;
    phi23[1][c] = pp;
  }
  const real opd = 2.0f * eta_film * thickness_nm * cos2;
  real res[2][3];
  for (int pol = 0; pol < 2; ++pol) {
    const V3 phi2 = v3(phi21[pol] + phi23[pol][0], phi21[pol] + phi23[pol][1], phi21[pol] + phi23[pol][2]);
    real i_c[3], cm[3], r_geo[3];
    for (int c = 0; c < 3; ++c) {
      r_geo[c] = std::sqrt(std::fmax(r12_sq[pol] * r23_sq[pol][c], 0.0f));
      i_c[c] = r12_sq[pol] + t121_sq[pol] * r23_sq[pol][c] / std::fmax(1.0f - r12_sq[pol] * r23_sq[pol][c], 1.0e-6f);
      cm[c] = t121_sq[pol] * r23_sq[pol][c];
    }
    for (int m = 1; m <= 2; ++m) {
      const V3 sens = film_sensitivity((real)m * opd, smul(phi2, (real)m));
      const real sv[3] = {sens.x, sens.y, sens.z};
      for (int c = 0; c < 3; ++c) {
        i_c[c] += 2.0f * cm[c] * sv[c];
        cm[c] *= r_geo[c];
      }
    }
    for (int c = 0; c < 3; ++c) res[pol][c] = i_c[c];
  }
  return v3(clampf(0.5f * (res[0][0] + res[1][0]), 0.0f, 1.0f), clampf(0.5f * (res[0][1] + res[1][1]), 0.0f, 1.0f),
            clampf(0.5f * (res[0][2] + res[1][2]), 0.0f, 1.0f));
}
// the film over a base seen through the outside medium or, with a coat,
// blended with the film under the coat by the coat weight (the common form
// of metal_fresnel :761-790, spec_brdf_fresnel :869-900 and
// spec_btdf_fresnel_with_eta :1120-1152)
inline V3 film_coated(const PbrBsdf& p, real c, V3 n_base, V3 k_base) {
  const V3 uncoated = thin_film(c, 1.0f, p.film_ior, n_base, k_base, p.film_thickness);
  if (p.coat <= 0.0f) return uncoated;
  const V3 coated = thin_film(c, p.coat_eta, p.film_ior, n_base, k_base, p.film_thickness);
  return vlerp(uncoated, coated, clampf(p.coat, 0.0f, 1.0f));
}
inline bool film_active(const PbrBsdf& p) { return p.film_weight > 0.0f && p.film_thickness > 0.0f; }

// apply_fuzz_roughening / effective_*_roughness (:736-778)
inline real fuzz_roughening(const OpenPbr& m, real r, real fw, real fr) {
  if (fw <= 0.0f) return r;
  V3 fc = vclamp01(m.fuzz_color);
  real e = zeltner_dir_albedo(1.0f, fr);
  real rf = clampf(luminance(smul(fc, e)), 0.0f, 1.0f);
  real rough = std::pow(std::fmin(powi_rt(r, 4) + 2.0f * powi_rt(rf, 4), 1.0f), 0.25f);
  return lerpf(r, rough, fw);
}
// coat_darkening_factor (:900-941); wo_local: the shading-frame outgoing direction.
inline V3 coat_darkening_factor(const OpenPbr& m, V3 base_color, V3 wo_local) {
  real cw = clampf(m.coat_weight, 0.0f, 1.0f), delta = clampf(m.coat_darkening, 0.0f, 1.0f);
  if (cw <= 0.0f || delta <= 0.0f) return vsplat(1.0f);
  real ec = std::fmax(m.coat_ior, 1.0f);
  real ef = dielectric_fresnel_avg_ln(ec);
  real kr = 1.0f - (1.0f - ef) / (ec * ec);
  real ks = fresnel_dielectric(clampf(std::fabs(wo_local.z), 0.0f, 1.0f), 1.0f, ec);
  real sior = std::fmax(m.specular_ior, 1.0e-4f);
  real eta_s = lerpf(sior, sior / ec, cw);
  real f0 = powi_rt((eta_s - 1.0f) / (eta_s + 1.0f), 2);
  real fs = clampf(std::fmax(m.specular_weight, 0.0f) * f0, 0.0f, 1.0f);
  real r_spec = clampf(m.specular_roughness * 1.0f, 0.0f, 1.0f);
  real metal = clampf(m.base_metalness * 1.0f, 0.0f, 1.0f);
  real rd = lerpf(1.0f, r_spec, fs);
  real rb = lerpf(rd, r_spec, metal);
  real k = clampf(lerpf(ks, kr, rb), 0.0f, 1.0f);
  V3 bw = smul(base_color, clampf(m.base_weight, 0.0f, 1.0f));
  V3 ssc = vsplat(0.8f);  // subsurface_color default, weight 0
  V3 e_dielec = vlerp(vlerp(bw, vclamp01(ssc), 0.0f), vsplat(1.0f - f0), clampf(m.transmission_weight, 0.0f, 1.0f));
  V3 e_metal = vclamp01(smul(bw, std::fmax(m.specular_weight, 0.0f)));
  V3 e_base = vclamp01(vlerp(e_dielec, e_metal, metal));
  V3 cc = vclamp01(m.coat_color);
  V3 kc = smul(cc, k);
  V3 den = vmax(sub(vsplat(1.0f), vmul(e_base, kc)), vsplat(1.0e-4f));
  real omk = std::fmax(1.0f - k, 0.0f);
  V3 df = v3(omk / den.x, omk / den.y, omk / den.z);
  return vclamp01(vlerp(vsplat(1.0f), df, cw * delta));
}
// OpenPbrMaterial::make_bsdf (:393-487)
inline PbrBsdf make_bsdf(const OpenPbr& m, bool front_face, V3 wo_local, const Luts& luts) {
  PbrBsdf p;
  V3 base_color = vmul(m.base_color, vsplat(1.0f));
  real metal = clampf(m.base_metalness * 1.0f, 0.0f, 1.0f);
  real fw = clampf(m.fuzz_weight * 1.0f, 0.0f, 1.0f);
  real fr = clampf(m.fuzz_roughness * 1.0f, 0.0f, 1.0f);
  real coat_r = fuzz_roughening(m, clampf(m.coat_roughness, 0.0f, 1.0f), fw, fr);
  real r_base = clampf(m.specular_roughness * 1.0f, 0.0f, 1.0f);
  real coat_rough = std::pow(std::fmin(powi_rt(r_base, 4) + 2.0f * powi_rt(coat_r, 4), 1.0f), 0.25f);
  real with_coat = lerpf(r_base, coat_rough, clampf(m.coat_weight, 0.0f, 1.0f));
  real spec_r = fuzz_roughening(m, with_coat, fw, fr);
  alpha_xy(spec_r, m.specular_roughness_anisotropy, &p.ax, &p.ay);
  alpha_xy(coat_r, m.coat_roughness_anisotropy, &p.cax, &p.cay);
  p.spec_roughness = spec_r;
  p.base_color = base_color;
  p.base = clampf(m.base_weight, 0.0f, 1.0f);
  p.specular = std::fmax(m.specular_weight, 0.0f);
  p.specular_color = vclamp01(m.specular_color);
  p.specular_eta = modulated_eta(std::fmax(m.specular_ior, 1.0e-4f), std::fmax(m.specular_weight, 0.0f));
  p.transmission_eta = p.specular_eta;
  p.metalness = metal;
  p.coat = clampf(m.coat_weight, 0.0f, 1.0f);
  p.coat_color = vclamp01(m.coat_color);
  p.coat_darkening = coat_darkening_factor(m, base_color, wo_local);
  p.coat_eta = std::fmax(m.coat_ior, 1.0e-4f);
  p.fuzz = fw;
  p.fuzz_color = vclamp01(m.fuzz_color);
  p.fuzz_roughness = fr;
  p.diffuse_roughness = clampf(m.base_diffuse_roughness, 0.0f, 1.0f);
  p.transmission = clampf(m.transmission_weight, 0.0f, 1.0f);
  p.transmission_color = vclamp01(m.transmission_color);
  p.thin_walled = m.thin_walled != 0;
  p.front_face = front_face;
  p.luts = luts;
  // thin_film_thickness_nm_at (R/material/open_pbr.rs:792-813), untextured
  p.film_weight = clampf(m.thin_film_weight, 0.0f, 1.0f);
  p.film_thickness = m.thin_film_weight <= 0.0f ? 0.0f : std::fmax(m.thin_film_thickness, 0.0f) * 1000.0f;
  p.film_ior = std::fmax(m.thin_film_ior, 1.0e-4f);
  artist_friendly_complex_ior(vclamp01(base_color), vclamp01(m.specular_color), &p.metal_n, &p.metal_k);
  return p;
}

// The tables a material reads (see Luts).
inline Luts luts_for(const OpenPbr& m) {
  Luts l;
  const real eta = modulated_eta(std::fmax(m.specular_ior, 1.0e-4f), std::fmax(m.specular_weight, 0.0f));
  const bool coat = clampf(m.coat_weight, 0.0f, 1.0f) > 0.0f;
  l.spec = &albedo_lut_for(eta);
  l.coat = coat ? &albedo_lut_for(std::fmax(m.coat_ior, 1.0e-4f)) : &zero_albedo_lut();
  // metalness clamp(base_metalness * texel, 0, 1) is positive only for a
  // non-zero parameter
  const bool metal = clampf(m.base_metalness, 0.0f, 1.0f) > 0.0f ||
                     (m.tex[kTexMetalness] != kNoTexture && m.base_metalness != 0.0f);
  l.cond = metal ? &cond_lut() : &zero_cond_lut();
  const bool trans = clampf(m.transmission_weight, 0.0f, 1.0f) > 0.0f;
  l.diel = (eta != 1.0f || coat || trans) ? &diel_lut() : &zero_diel_lut();
  return l;
}
// OpenPbrMaterial::emission_exit_tint (:608-638); wo_local normalised.
inline V3 emission_exit_tint(const OpenPbr& m, V3 wo_local) {
  real co = clampf(std::fabs(wo_local.z), 0.0f, 1.0f);
  if (co <= 0.0f) return vsplat(0.0f);
  V3 tint = vsplat(1.0f);
  real cw = clampf(m.coat_weight, 0.0f, 1.0f);
  if (cw > 0.0f) {
    real ec = std::fmax(m.coat_ior, 1.0e-4f);
    real mu = std::fmax(std::sqrt(std::fmax(1.0f - (1.0f - co * co) / (ec * ec), 0.0f)), 1.0e-4f);
    V3 tr = vec3_powf(vclamp01(m.coat_color), 1.0f / mu);
    tint = vmul(tint, vlerp(vsplat(1.0f), tr, cw));
  }
  real fw = clampf(m.fuzz_weight * 1.0f, 0.0f, 1.0f);
  if (fw > 0.0f) {
    real ef = zeltner_dir_albedo(co, clampf(m.fuzz_roughness * 1.0f, 0.0f, 1.0f));
    tint = vmul(tint, vsplat(lerpf(1.0f, std::fmax(1.0f - ef, 0.0f), fw)));
  }
  return tint;
}

// ---- OpenPbrBsdf, R/bsdf/open_pbr.rs ------------------------------------------------------
constexpr real kCos82 = 1.0f / 7.0f;    // :30
constexpr real kMsDenomEps = 1.0e-6f;  // :2125
inline V3 schlick_vec3(V3 f0, real c) {  // :2148-2151
  real om = 1.0f - clampf(c, 0.0f, 1.0f);
  real p5 = powi_rt(om, 5);
  return add(f0, smul(sub(vsplat(1.0f), f0), p5));
}
inline real refracted_cos(real mu, real eta) {  // refracted_cos_in_dielectric :2174-2177
  real s2 = std::fmax(1.0f - powi_rt(clampf(mu, 0.0f, 1.0f), 2), 0.0f) / std::fmax(eta * eta, kMsDenomEps);
  return std::sqrt(std::fmax(1.0f - s2, 0.0f));
}
inline real rough_proxy(real ax, real ay) { return clampf(std::sqrt(std::fmax(ax, ay)), 0.0f, 1.0f); }
inline bool coat_smooth(const PbrBsdf& p) { return p.coat <= 0.0f || std::fmax(p.cax, p.cay) < kSmoothAlpha; }
inline bool metal_smooth(const PbrBsdf& p) { return p.metalness <= 0.0f || std::fmax(p.ax, p.ay) < kSmoothAlpha; }
inline bool spec_smooth(const PbrBsdf& p) { return p.specular <= 0.0f || std::fmax(p.ax, p.ay) < kSmoothAlpha; }
// spec_btdf_is_smooth (:547-551); the transmission alphas are the specular
// ones (R/material/open_pbr.rs:451-452)
inline bool trans_smooth(const PbrBsdf& p) { return p.transmission <= 0.0f || std::fmax(p.ax, p.ay) < kSmoothAlpha; }

struct MsParams {
  real eta_o, eta_rel, req, e_avg_o, ratio_r, ome_o, ome_t;
};
inline MsParams ms_params(const PbrBsdf& p, real ax, real ay, real eta) {  // dielectric_ms_params :593-625
  MsParams s;
  s.eta_o = p.front_face ? std::fmax(eta, 1.0e-4f) : 1.0f / std::fmax(eta, 1.0e-4f);
  s.eta_rel = 1.0f / s.eta_o;
  s.req = std::pow(ax * ay, 0.25f);
  s.e_avg_o = diel_lookup_eavg(*p.luts.diel, s.req, s.eta_o);
  real e_avg_t = diel_lookup_eavg(*p.luts.diel, s.req, s.eta_rel);
  s.ome_o = std::fmax(1.0f - s.e_avg_o, kMsDenomEps);
  s.ome_t = std::fmax(1.0f - e_avg_t, kMsDenomEps);
  real fo = f_avg_dielectric(s.eta_o), ft = f_avg_dielectric(s.eta_rel);
  real a = (1.0f - fo) / s.ome_o;
  real b = (1.0f - ft) * s.eta_o * s.eta_o / s.ome_t;
  real x = a + b > kMsDenomEps ? b / (a + b) : 0.5f;
  s.ratio_r = clampf(1.0f - x * (1.0f - fo), 0.0f, 1.0f);
  return s;
}
inline real ms_factor(const MsParams& s) {  // dielectric_reflection_ms_factor :573-577
  real fa = clampf(f_avg_dielectric(s.eta_o), 0.0f, 1.0f);
  real den = 1.0f - fa * std::fmax(1.0f - s.e_avg_o, 0.0f);
  return clampf(fa * fa * s.e_avg_o / std::fmax(den, kMsDenomEps), 0.0f, 1.0f);
}
struct TwoLobe {
  real ss, ms, total;
};
inline TwoLobe refl_lobe_weights(const MsParams& s) {  // :627-636
  real fa = clampf(f_avg_dielectric(s.eta_o), 0.0f, 1.0f);
  real ss = clampf(fa * s.e_avg_o, 0.0f, 1.0f);
  real ms = std::fmax(1.0f - s.e_avg_o, 0.0f) * ms_factor(s);
  return TwoLobe{ss, ms, std::fmax(ss + ms, 1.0e-8f)};
}
inline TwoLobe cond_lobe_weights(const PbrBsdf& p) {  // :579-591
  real e = clampf(cond_lookup_eavg(*p.luts.cond, std::pow(p.ax * p.ay, 0.25f)), 0.0f, 1.0f);
  real ms = std::fmax(1.0f - e, 0.0f);
  return TwoLobe{e, ms, std::fmax(e + ms, 1.0e-8f)};
}
inline real albedo_with_ms(const PbrBsdf& p, V3 wo, real ax, real ay, real eta, real ss) {  // :464-481
  if (wo.z <= 0.0f || std::fmax(ax, ay) < kSmoothAlpha) return clampf(ss, 0.0f, 1.0f);
  MsParams s = ms_params(p, ax, ay, eta);
  real eo = diel_lookup_e(*p.luts.diel, wo.z, s.req, s.eta_o);
  return clampf(ss + ms_factor(s) * std::fmax(1.0f - eo, 0.0f), 0.0f, 1.0f);
}
inline real coat_albedo(const PbrBsdf& p, V3 wo) {  // lookup_coat_albedo :424-437
  if (wo.z <= 0.0f) return 0.0f;
  real ss = albedo_lookup(*p.luts.coat, wo, rough_proxy(p.cax, p.cay));
  return albedo_with_ms(p, wo, p.cax, p.cay, p.coat_eta, ss);
}
inline real spec_albedo(const PbrBsdf& p, V3 wo) {  // lookup_spec_albedo :439-454
  if (wo.z <= 0.0f) return 0.0f;
  real ss = albedo_lookup(*p.luts.spec, wo, rough_proxy(p.ax, p.ay));
  if (p.thin_walled) return ss;
  return albedo_with_ms(p, wo, p.ax, p.ay, p.specular_eta, ss);
}
inline real fuzz_albedo(const PbrBsdf& p, V3 wo) {  // lookup_fuzz_albedo :456-462
  if (p.fuzz <= 0.0f) return 0.0f;
  return zeltner_dir_albedo(clampf(wo.z, 0.0f, 1.0f), clampf(p.fuzz_roughness, 0.01f, 1.0f));
}
inline V3 coat_absorb_tint(const PbrBsdf& p, V3 wo, V3 wi) {  // :483-492
  if (wo.z <= 0.0f || wi.z <= 0.0f) return vsplat(1.0f);
  real mo = refracted_cos(wo.z, p.coat_eta), mi = refracted_cos(wi.z, p.coat_eta);
  real ex = 0.5f * (1.0f / std::fmax(mi, 1.0e-4f) + 1.0f / std::fmax(mo, 1.0e-4f));
  return vec3_powf(p.coat_color, ex);
}

struct LayerW {
  real coat_amp;
  V3 metal, spec_brdf, spec_btdf, fuzz, diff_brdf, diff_btdf;
};
// layer_weights (:333-390); subsurface is 0 (OpenPbrMaterial::make_bsdf
// forces it, R/material/open_pbr.rs:455), so diff_btdf is always zero.
inline LayerW layer_weights(const PbrBsdf& p, V3 wo, const V3* wi) {
  const V3 one = vsplat(1.0f);
  real e_fuzz = fuzz_albedo(p, wo);
  V3 fuzz_w = smul(p.fuzz_color, p.fuzz);
  real bfs = std::fmax(1.0f - p.fuzz * e_fuzz, 0.0f);
  V3 below_fuzz = vsplat(bfs);
  real coat_amp = p.coat;
  real e_coat = coat_amp > 0.0f ? coat_albedo(p, wo) : 0.0f;
  V3 under_coat = below_fuzz;
  if (coat_amp > 0.0f) {
    V3 tint = wi ? coat_absorb_tint(p, wo, *wi) : p.coat_color;
    V3 thr = vmul(vmul(p.coat_darkening, tint), vsplat(std::fmax(1.0f - e_coat, 0.0f)));
    under_coat = vmul(below_fuzz, add(smul(one, 1.0f - coat_amp), smul(thr, coat_amp)));
  }
  LayerW w;
  w.metal = smul(under_coat, p.metalness);
  V3 below_metal = smul(under_coat, 1.0f - p.metalness);
  real e_spec = spec_albedo(p, wo);
  w.spec_brdf = vmul(below_metal, p.specular_color);
  V3 leak = vmax(sub(one, smul(p.specular_color, e_spec)), vsplat(0.0f));
  V3 below_spec = vmul(below_metal, leak);
  const real trans = p.transmission;
  const V3 tint = p.thin_walled ? one : p.transmission_color;
  w.spec_btdf = vmul(smul(below_spec, trans), tint);
  V3 below_trans = smul(below_spec, 1.0f - trans);
  const real thin_factor = p.thin_walled ? 1.0f : 0.0f;
  w.diff_brdf = vmul(smul(below_trans, (1.0f - 0.0f) * p.base), p.base_color);
  w.diff_btdf = vmul(smul(below_trans, 0.0f * thin_factor), vsplat(0.8f));
  w.coat_amp = coat_amp * bfs;
  w.fuzz = fuzz_w;
  return w;
}
struct Probs {
  real coat, metal, spec, spec_btdf, fuzz, diff, diff_btdf, total;
};
inline Probs lobe_probs(const PbrBsdf& p, V3 wo) {  // :392-422
  LayerW w = layer_weights(p, wo, nullptr);
  Probs r;
  r.coat = p.coat > 0.0f ? w.coat_amp : 0.0f;
  r.metal = luminance(vmul(w.metal, add(vsplat(0.5f), smul(p.base_color, 0.5f))));
  r.spec = luminance(w.spec_brdf);
  r.spec_btdf = p.transmission > 0.0f ? std::fmax(luminance(w.spec_btdf), 0.0f) : 0.0f;
  r.fuzz = std::fmax(luminance(w.fuzz), 0.0f);
  r.diff = std::fmax(luminance(w.diff_brdf), 0.0f);
  r.diff_btdf = std::fmax(luminance(w.diff_btdf), 0.0f);
  r.total = r.coat + r.metal + r.spec + r.spec_btdf + r.fuzz + r.diff + r.diff_btdf;
  return r;
}

inline real eval_refl_ms(const PbrBsdf& p, V3 wo, V3 wi, real ax, real ay, real eta) {  // :648-670
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  MsParams s = ms_params(p, ax, ay, eta);
  real eo = diel_lookup_e(*p.luts.diel, wo.z, s.req, s.eta_o);
  real ei = diel_lookup_e(*p.luts.diel, wi.z, s.req, s.eta_o);
  return ms_factor(s) * (1.0f - eo) * (1.0f - ei) / (kPi * s.ome_o);
}
inline real eval_coat(const PbrBsdf& p, V3 wo, V3 wi) {  // :698-722 (one channel)
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  V3 wm;
  if (!reflection_half_vector(wo, wi, &wm)) return 0.0f;
  real d = ggx_d(wm, p.cax, p.cay), g = ggx_g2(wo, wi, p.cax, p.cay);
  if (d <= 0.0f || g <= 0.0f) return 0.0f;
  real f = fresnel_dielectric(std::fabs(dot(wo, wm)), 1.0f, p.coat_eta);
  real ss = d * g * f / (4.0f * wo.z * wi.z);
  return ss + eval_refl_ms(p, wo, wi, p.cax, p.cay, p.coat_eta);
}
inline real pdf_coat(const PbrBsdf& p, V3 wo, V3 wi) {  // :724-742
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  TwoLobe w = refl_lobe_weights(ms_params(p, p.cax, p.cay, p.coat_eta));
  real pms = cosine_hemisphere_pdf(wi.z);
  V3 wm;
  if (!reflection_half_vector(wo, wi, &wm)) return w.ms * pms / w.total;
  real pw = pdf_wm_bounded_vndf(wo, wm, p.cax, p.cay);
  real den = 4.0f * std::fabs(dot(wo, wm));
  if (den <= 0.0f) return w.ms * pms / w.total;
  return (w.ss * (pw / den) + w.ms * pms) / w.total;
}
inline V3 metal_f82_fresnel(const PbrBsdf& p, real c) {  // metal_f82_tint_fresnel :792-802
  V3 f0 = vmax(smul(p.base_color, p.base), vsplat(0.0f));
  V3 fs = schlick_vec3(f0, c), f82 = schlick_vec3(f0, kCos82);
  V3 ft = vmul(vclamp01(p.specular_color), f82);
  real cc = clampf(c, 0.0f, 1.0f);
  real den = kCos82 * powi_const(1.0f - kCos82, 6);
  real corr = cc * powi_rt(1.0f - cc, 6) / std::fmax(den, 1.0e-8f);
  V3 r = sub(fs, smul(sub(f82, ft), corr));
  return vclamp01(smul(r, p.specular));
}
// metal_fresnel (:761-790): the F82-tint Fresnel, blended with the film
// over the metal's complex IOR
inline V3 metal_fresnel(const PbrBsdf& p, real c) {
  const V3 no_film = metal_f82_fresnel(p, c);
  if (!film_active(p)) return no_film;
  return vlerp(no_film, film_coated(p, c, p.metal_n, p.metal_k), p.film_weight);
}
inline V3 metal_fresnel_avg(const PbrBsdf& p);
inline V3 metal_f82_fresnel_avg(const PbrBsdf& p) {  // metal_f82_tint_fresnel_avg :832-839
  V3 f0 = vmax(smul(p.base_color, p.base), vsplat(0.0f));
  V3 f82 = schlick_vec3(f0, kCos82);
  V3 ft = vmul(vclamp01(p.specular_color), f82);
  real den = kCos82 * powi_const(1.0f - kCos82, 6);
  V3 b = sdiv(sub(f82, ft), std::fmax(den, 1.0e-8f));
  V3 r = sub(add(f0, sdiv(sub(vsplat(1.0f), f0), 21.0f)), sdiv(b, 126.0f));
  return vclamp01(smul(r, p.specular));
}
// metal_fresnel_avg (:818-830): with a film, the cosine-weighted mean of
// metal_fresnel over 16 midpoints
inline V3 metal_fresnel_avg(const PbrBsdf& p) {
  if (p.film_weight <= 0.0f || p.film_thickness <= 0.0f) return metal_f82_fresnel_avg(p);
  V3 sum = vsplat(0.0f);
  for (int i = 0; i < 16; ++i) {
    const real mu = ((real)i + 0.5f) / 16.0f;
    sum = add(sum, smul(metal_fresnel(p, mu), 2.0f * mu));
  }
  return sdiv(sum, 16.0f);
}
inline V3 eval_metal(const PbrBsdf& p, V3 wo, V3 wi) {  // :744-759, 804-816
  const V3 zero = vsplat(0.0f);
  if (wo.z <= 0.0f || wi.z <= 0.0f) return zero;
  V3 wm;
  if (!reflection_half_vector(wo, wi, &wm)) return zero;
  real d = ggx_d(wm, p.ax, p.ay), g = ggx_g2(wo, wi, p.ax, p.ay);
  if (d <= 0.0f || g <= 0.0f) return zero;
  V3 f = metal_fresnel(p, std::fabs(dot(wo, wm)));
  V3 ss = smul(f, d * g / (4.0f * wo.z * wi.z));
  // eval_metal_ms
  real req = std::pow(p.ax * p.ay, 0.25f);
  real eo = cond_lookup_e(*p.luts.cond, wo.z, req), ei = cond_lookup_e(*p.luts.cond, wi.z, req);
  real ea = cond_lookup_eavg(*p.luts.cond, req);
  real omea = std::fmax(1.0f - ea, kMsDenomEps);
  V3 fa = metal_fresnel_avg(p);
  // compute_conductor_f_ms :2137-2146
  real om = std::fmax(1.0f - ea, 0.0f);
  V3 den = sub(vsplat(1.0f), smul(fa, om));
  V3 ds = v3(std::fmax(den.x, kMsDenomEps), std::fmax(den.y, kMsDenomEps), std::fmax(den.z, kMsDenomEps));
  V3 fms = smul(vmul(fa, fa), ea);
  fms = v3(fms.x / ds.x, fms.y / ds.y, fms.z / ds.z);
  return add(ss, smul(fms, (1.0f - eo) * (1.0f - ei) / (kPi * omea)));
}
inline real base_dielectric_fresnel(const PbrBsdf& p, real c, real ep, real epr) {  // :902-929
  real un = p.front_face ? weighted_fresnel_ratio(c, ep, epr) : weighted_fresnel_ratio(c, 1.0f / ep, 1.0f / epr);
  real coat = clampf(p.coat, 0.0f, 1.0f);
  if (coat <= 0.0f) return un;
  real ce = std::fmax(p.coat_eta, 1.0e-4f);
  real eb = ep / ce, ebp = epr / ce;
  real a, b;
  if (p.coat_eta > ep) {
    a = 1.0f / std::fmax(eb, 1.0e-4f);
    b = 1.0f / std::fmax(ebp, 1.0e-4f);
  } else {
    a = eb;
    b = ebp;
  }
  real co = p.front_face ? weighted_fresnel_ratio(c, a, b)
                          : weighted_fresnel_ratio(c, 1.0f / std::fmax(a, 1.0e-4f), 1.0f / std::fmax(b, 1.0e-4f));
  return lerpf(un, co, coat);
}
// spec_brdf_fresnel (:869-900): the base's dielectric Fresnel, blended with
// the film over a dielectric base of IOR specular_eta
inline V3 spec_brdf_fresnel(const PbrBsdf& p, real c) {
  const V3 no_film = vsplat(base_dielectric_fresnel(p, c, p.transmission_eta, p.specular_eta));
  if (!film_active(p)) return no_film;
  return vlerp(no_film, film_coated(p, c, vsplat(p.specular_eta), vsplat(0.0f)), p.film_weight);
}
// spec_btdf_fresnel_with_eta (:1120-1152): the same for the refraction
// Fresnel (no film on a thin wall)
inline V3 spec_btdf_fresnel(const PbrBsdf& p, real c, real eta_physical, real eta_prime) {
  const V3 no_film = vsplat(base_dielectric_fresnel(p, c, eta_physical, eta_prime));
  if (!film_active(p) || p.thin_walled) return no_film;
  return vlerp(no_film, film_coated(p, c, vsplat(eta_prime), vsplat(0.0f)), p.film_weight);
}
// thin_wall_coefficients (:931-959): the slab's total reflectance and
// transmittance for an outgoing cosine, with the interior absorption
// transmission_color^(1/cos_i) and the internal reflections summed.
inline V3 vdiv3(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline void thin_wall_coefficients(const PbrBsdf& p, real cos_theta_o, V3* refl, V3* trans) {
  const real eta = std::fmax(p.transmission_eta, 1.0e-4f), eta_prime = std::fmax(p.specular_eta, 1.0e-4f);
  const real cos_o = clampf(cos_theta_o, 0.0f, 1.0f);
  const real sin2_o = std::fmax(1.0f - cos_o * cos_o, 0.0f);
  const real sin2_i = clampf(sin2_o / (eta * eta), 0.0f, 1.0f);
  const real cos_i = std::fmax(std::sqrt(1.0f - sin2_i), 1.0e-4f);
  const V3 f_o = vclamp01(spec_brdf_fresnel(p, cos_o));
  // thin_wall_internal_fresnel (:961-979): the film seen from inside the
  // slab (outside medium eta, base 1)
  V3 f_i = vsplat(weighted_fresnel_ratio(cos_i, 1.0f / std::fmax(eta, 1.0e-4f), 1.0f / std::fmax(eta_prime, 1.0e-4f)));
  if (film_active(p))
    f_i = vlerp(f_i, thin_film(cos_i, std::fmax(eta, 1.0e-4f), p.film_ior, vsplat(1.0f), vsplat(0.0f), p.film_thickness),
                p.film_weight);
  f_i = vclamp01(f_i);
  const V3 a = vec3_powf(vclamp01(p.transmission_color), 1.0f / cos_i);
  const V3 a2 = vmul(a, a);
  const V3 fi2a2 = vmul(vmul(a2, f_i), f_i);
  const V3 denom = vmax(sub(vsplat(1.0f), fi2a2), vsplat(1.0e-6f));
  const V3 omfo = vmax(sub(vsplat(1.0f), f_o), vsplat(0.0f)), omfi = vmax(sub(vsplat(1.0f), f_i), vsplat(0.0f));
  const V3 r = add(f_o, vdiv3(vmul(vmul(vmul(omfo, omfi), f_i), a2), denom));
  const V3 t = vmul(vmul(vmul(omfo, omfi), a), add(vsplat(1.0f), vdiv3(fi2a2, denom)));
  *refl = vclamp01(r);
  *trans = vclamp01(t);
}
// thin_wall_transmission_alpha_xy (:553-563)
inline void thin_wall_alpha(const PbrBsdf& p, real* ax, real* ay) {
  const real eta = std::fmax(p.transmission_eta, 1.0f);
  const real scale = std::sqrt(std::fmax(3.7f * (eta - 1.0f) * powi_rt(eta - 0.5f, 2) / powi_rt(eta, 3), 0.0f));
  const real alpha_scale = scale * scale;
  *ax = clampf(p.ax * alpha_scale, kMinAlpha, 1.0f);
  *ay = clampf(p.ay * alpha_scale, kMinAlpha, 1.0f);
}
inline V3 eval_spec(const PbrBsdf& p, V3 wo, V3 wi) {  // eval_spec_brdf :841-867
  const V3 zero = vsplat(0.0f);
  if (wo.z <= 0.0f || wi.z <= 0.0f) return zero;
  V3 wm;
  if (!reflection_half_vector(wo, wi, &wm)) return zero;
  real d = ggx_d(wm, p.ax, p.ay), g = ggx_g2(wo, wi, p.ax, p.ay);
  if (d <= 0.0f || g <= 0.0f) return zero;
  if (p.thin_walled) {
    V3 r, t;
    thin_wall_coefficients(p, std::fabs(wo.z), &r, &t);
    return smul(r, d * g / (4.0f * wo.z * wi.z));
  }
  V3 f = spec_brdf_fresnel(p, std::fabs(dot(wo, wm)));
  V3 ss = smul(f, d * g / (4.0f * wo.z * wi.z));
  return add(ss, vsplat(eval_refl_ms(p, wo, wi, p.ax, p.ay, p.specular_eta)));
}
inline real pdf_spec_ss(const PbrBsdf& p, V3 wo, V3 wi) {  // :981-994
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  V3 wm;
  if (!reflection_half_vector(wo, wi, &wm)) return 0.0f;
  real pw = pdf_wm_bounded_vndf(wo, wm, p.ax, p.ay);
  real den = 4.0f * std::fabs(dot(wo, wm));
  if (den <= 0.0f) return 0.0f;
  return pw / den;
}
inline real pdf_metal(const PbrBsdf& p, V3 wo, V3 wi) {  // :996-1004
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  TwoLobe w = cond_lobe_weights(p);
  return (w.ss * pdf_spec_ss(p, wo, wi) + w.ms * cosine_hemisphere_pdf(wi.z)) / w.total;
}
inline real pdf_spec(const PbrBsdf& p, V3 wo, V3 wi) {  // pdf_specular_brdf :1006-1022
  if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
  if (p.thin_walled) return pdf_spec_ss(p, wo, wi);
  TwoLobe w = refl_lobe_weights(ms_params(p, p.ax, p.ay, p.specular_eta));
  return (w.ss * pdf_spec_ss(p, wo, wi) + w.ms * cosine_hemisphere_pdf(wi.z)) / w.total;
}

// ---- specular BTDF (:638-697, 1024-1222), radiance mode, no dispersion ------------------------
// transmission_eta_rel (:1245-1248): relative IOR across the interface, as
// seen from the outgoing side
inline real trans_eta_rel(const PbrBsdf& p) { return p.front_face ? 1.0f / p.transmission_eta : p.transmission_eta; }
inline TwoLobe trans_lobe_weights(const MsParams& s) {  // dielectric_transmission_lobe_weights :638-646
  real ss = clampf(s.e_avg_o, 0.0f, 1.0f);
  real ms = std::fmax(1.0f - s.e_avg_o, 0.0f) * clampf(1.0f - s.ratio_r, 0.0f, 1.0f);
  return TwoLobe{ss, ms, std::fmax(ss + ms, 1.0e-8f)};
}
// eval_dielectric_transmission_ms (:672-696)
inline real eval_trans_ms(const PbrBsdf& p, V3 wo, V3 wi, real ax, real ay, real eta) {
  if (wo.z <= 0.0f || wi.z >= 0.0f) return 0.0f;
  MsParams s = ms_params(p, ax, ay, eta);
  real eo = diel_lookup_e(*p.luts.diel, wo.z, s.req, s.eta_o);
  real ei = diel_lookup_e(*p.luts.diel, std::fabs(wi.z), s.req, s.eta_rel);
  real rs = 1.0f / (s.eta_rel * s.eta_rel);  // TransportMode::Radiance.transmission_scale
  return (1.0f - s.ratio_r) * (1.0f - eo) * (1.0f - ei) * rs / (kPi * s.ome_t);
}
// the generalised half vector of a refraction pair (:1081-1096, 1194-1209):
// wm = normalize(eta_rel wo + wi) flipped up; false where undefined
inline bool refraction_half_vector(V3 wo, V3 wi, real eta_rel, V3* wm, real* cos_wo_wm, real* cos_wi_wm, real* den) {
  V3 u = add(smul(wo, eta_rel), wi);
  if (len_sq(u) < 1.0e-12f) return false;
  V3 m = normalize3(u);
  if (m.z < 0.0f) m = neg(m);
  real co = dot(wo, m);
  if (co <= 0.0f) return false;
  real ci = dot(wi, m);
  real dn = ci + eta_rel * co;
  if (std::fabs(dn) < 1.0e-6f) return false;
  *wm = m;
  *cos_wo_wm = co;
  *cos_wi_wm = ci;
  *den = dn;
  return true;
}
// eval_spec_btdf_scalar_with_eta (:1074-1118): D G |cos_i cos_o| scale / (den^2 |wo.z| |wi.z|)
inline bool eval_btdf_scalar(const PbrBsdf& p, V3 wo, V3 wi, real eta_rel, real* scalar, real* cos_wo_wm) {
  V3 wm;
  real co, ci, dn;
  if (!refraction_half_vector(wo, wi, eta_rel, &wm, &co, &ci, &dn)) return false;
  real d = ggx_d(wm, p.ax, p.ay), g = ggx_g2(wo, wi, p.ax, p.ay);
  if (d <= 0.0f || g <= 0.0f) return false;
  real rs = 1.0f / (eta_rel * eta_rel);
  real sc = d * g * std::fabs(ci * co);
  real denom = dn * dn * std::fabs(wo.z) * std::fabs(wi.z);
  if (denom <= 0.0f) return false;
  *scalar = sc * rs / denom;
  *cos_wo_wm = co;
  return true;
}
// eval_thin_wall_spec_btdf (:1053-1072): the reflection lobe at the
// thin-wall alphas, mirrored through the surface, times the slab transmittance
inline V3 eval_thin_btdf(const PbrBsdf& p, V3 wo, V3 wi) {
  const V3 zero = vsplat(0.0f);
  if (wo.z <= 0.0f || wi.z >= 0.0f) return zero;
  V3 r, t;
  thin_wall_coefficients(p, std::fabs(wo.z), &r, &t);
  if (trans_smooth(p)) return zero;
  V3 wim = v3(wi.x, wi.y, -wi.z), wm;
  if (!reflection_half_vector(wo, wim, &wm)) return zero;
  real tax, tay;
  thin_wall_alpha(p, &tax, &tay);
  real d = ggx_d(wm, tax, tay), g = ggx_g2(wo, wim, tax, tay);
  if (d <= 0.0f || g <= 0.0f) return zero;
  return smul(t, d * g / (4.0f * wo.z * std::fabs(wi.z)));
}
inline V3 eval_spec_btdf(const PbrBsdf& p, V3 wo, V3 wi) {  // eval_spec_btdf :1024-1051
  if (p.thin_walled) return eval_thin_btdf(p, wo, wi);
  const V3 zero = vsplat(0.0f);
  if (wo.z <= 0.0f || wi.z >= 0.0f) return zero;
  real sc, co;
  if (!eval_btdf_scalar(p, wo, wi, trans_eta_rel(p), &sc, &co)) return zero;
  V3 f = spec_btdf_fresnel(p, co, p.transmission_eta, p.specular_eta);
  V3 ss = smul(vmax(sub(vsplat(1.0f), f), vsplat(0.0f)), sc);
  return add(ss, vsplat(eval_trans_ms(p, wo, wi, p.ax, p.ay, p.specular_eta)));
}
inline real pdf_btdf_ss(const PbrBsdf& p, V3 wo, V3 wi, real eta_rel) {  // pdf_specular_btdf_ss_with_eta :1193-1221
  V3 wm;
  real co, ci, dn;
  if (!refraction_half_vector(wo, wi, eta_rel, &wm, &co, &ci, &dn)) return 0.0f;
  real pw = pdf_wm_vndf(wo, wm, p.ax, p.ay);
  if (pw <= 0.0f) return 0.0f;
  return pw * std::fabs(ci) / (dn * dn);
}
inline real pdf_spec_btdf(const PbrBsdf& p, V3 wo, V3 wi) {  // pdf_specular_btdf :1154-1191
  if (p.thin_walled) {
    if (trans_smooth(p) || wo.z <= 0.0f || wi.z >= 0.0f) return 0.0f;
    V3 wim = v3(wi.x, wi.y, -wi.z), wm;
    if (!reflection_half_vector(wo, wim, &wm)) return 0.0f;
    real tax, tay;
    thin_wall_alpha(p, &tax, &tay);
    real pw = pdf_wm_bounded_vndf(wo, wm, tax, tay);
    real den = 4.0f * std::fabs(dot(wo, wm));
    return den > 0.0f ? pw / den : 0.0f;
  }
  if (wo.z <= 0.0f || wi.z >= 0.0f) return 0.0f;
  TwoLobe w = trans_lobe_weights(ms_params(p, p.ax, p.ay, p.transmission_eta));
  real pss = pdf_btdf_ss(p, wo, wi, trans_eta_rel(p));
  return (w.ss * pss + w.ms * cosine_hemisphere_pdf(-wi.z)) / w.total;
}
// OpenPbrFuzzBsdf (:148-205) with the LTC of Zeltner et al.
inline real phi_of(V3 v) {  // :2060-2063
  real ph = std::atan2(v.y, v.x);
  return ph < 0.0f ? ph + 2.0f * kPi : ph;
}
inline V3 rotate_z(V3 v, real angle) {  // :2065-2068
  real s = std::sin(angle), c = std::cos(angle);
  return v3(c * v.x - s * v.y, s * v.x + c * v.y, v.z);
}
inline real eval_ltc(V3 wi, real a, real b) {  // :2097-2110
  V3 w = v3(a * wi.x + b * wi.z, a * wi.y, wi.z);
  real l = std::sqrt(dot(w, w));
  if (l <= 0.0f) return 0.0f;
  w = sdiv(w, l);
  if (w.z <= 0.0f) return 0.0f;
  real det = a * a;
  real jac = det / (l * l * l);
  return (w.z / kPi) * jac;
}
inline real fuzz_eval(const PbrBsdf& p, V3 wo, V3 wi, real* pdf) {  // eval (one channel) and pdf
  real r = clampf(p.fuzz_roughness, 0.01f, 1.0f);
  if (wo.z <= 0.0f || wi.z <= 0.0f) {
    *pdf = 0.0f;
    return 0.0f;
  }
  V3 ws = rotate_z(wi, -phi_of(wo));
  real a = zeltner_a_inv(wo.z, r), b = zeltner_b_inv(wo.z, r);
  real v = eval_ltc(ws, a, b);
  *pdf = v;
  real rr = zeltner_dir_albedo(wo.z, r);
  return rr * v / std::fmax(wi.z, 1.0e-6f);
}
inline bool fuzz_sample(const PbrBsdf& p, V3 wo, real u0, real u1, V3* wi) {  // :184-200, sample_ltc :2112-2123
  real r = clampf(p.fuzz_roughness, 0.01f, 1.0f);
  if (wo.z <= 0.0f) return false;
  real a = zeltner_a_inv(wo.z, r), b = zeltner_b_inv(wo.z, r);
  real phi = 2.0f * kPi * u0;
  real ct = std::sqrt(u1), st = std::sqrt(std::fmax(1.0f - u1, 0.0f));
  V3 o = v3(std::cos(phi) * st, std::sin(phi) * st, ct);
  V3 ws = normalize_or_zero3(v3(o.x / a - o.z * b / a, o.y / a, o.z));
  if (ws.z <= 0.0f) return false;
  V3 w = rotate_z(ws, phi_of(wo));
  if (w.z <= 0.0f) return false;
  *wi = w;
  return true;
}

// OpenPbrBsdf::eval (:224-259) / pdf (:261-293), radiance mode.
inline V3 bsdf_eval(const PbrBsdf& p, V3 wo, V3 wi) {
  V3 total = vsplat(0.0f);
  if (!(wo.z > 0.0f)) return total;
  LayerW w = layer_weights(p, wo, &wi);
  if (w.coat_amp > 0.0f && !coat_smooth(p)) total = add(total, smul(vsplat(eval_coat(p, wo, wi)), w.coat_amp));
  if (len_sq(w.metal) > 0.0f && !metal_smooth(p)) total = add(total, vmul(w.metal, eval_metal(p, wo, wi)));
  if (len_sq(w.spec_brdf) > 0.0f && !spec_smooth(p)) total = add(total, vmul(w.spec_brdf, eval_spec(p, wo, wi)));
  if (len_sq(w.spec_btdf) > 0.0f && !trans_smooth(p)) total = add(total, vmul(w.spec_btdf, eval_spec_btdf(p, wo, wi)));
  if (len_sq(w.fuzz) > 0.0f) {
    real pd;
    total = add(total, vmul(w.fuzz, vsplat(fuzz_eval(p, wo, wi, &pd))));
  }
  if (len_sq(w.diff_brdf) > 0.0f) {
    V3 f = (wo.z <= 0.0f || wi.z <= 0.0f) ? vsplat(0.0f) : f_eon(vsplat(1.0f), p.diffuse_roughness, wo, wi);
    total = add(total, vmul(w.diff_brdf, f));
  }
  return total;
}
inline real bsdf_pdf(const PbrBsdf& p, V3 wo, V3 wi) {
  if (!(wo.z > 0.0f)) return 0.0f;
  Probs pr = lobe_probs(p, wo);
  if (pr.total <= 0.0f) return 0.0f;
  real pdf = 0.0f;
  if (pr.coat > 0.0f && !coat_smooth(p)) pdf += (pr.coat / pr.total) * pdf_coat(p, wo, wi);
  if (pr.metal > 0.0f && !metal_smooth(p)) pdf += (pr.metal / pr.total) * pdf_metal(p, wo, wi);
  if (pr.spec > 0.0f && !spec_smooth(p)) pdf += (pr.spec / pr.total) * pdf_spec(p, wo, wi);
  if (pr.spec_btdf > 0.0f && !trans_smooth(p)) pdf += (pr.spec_btdf / pr.total) * pdf_spec_btdf(p, wo, wi);
  if (pr.fuzz > 0.0f) {
    real pd;
    fuzz_eval(p, wo, wi, &pd);
    pdf += (pr.fuzz / pr.total) * pd;
  }
  if (pr.diff > 0.0f) {
    real pd = (wo.z <= 0.0f || wi.z <= 0.0f) ? 0.0f : pdf_eon(wo, wi, p.diffuse_roughness);
    pdf += (pr.diff / pr.total) * pd;
  }
  return pdf;
}
// finalize_rough_sample (:1958-1999)
inline bool finalize(const PbrBsdf& p, V3 wo, V3 wi, BSample* out) {
  real pdf = bsdf_pdf(p, wo, wi);
  if (pdf <= 0.0f) return false;
  V3 f = bsdf_eval(p, wo, wi);
  if (len_sq(f) == 0.0f) return false;
  real c = std::fabs(wi.z);
  if (c <= 0.0f) return false;
  out->weight = smul(f, c / pdf);
  out->wi = wi;
  out->pdf = pdf;
  out->delta = false;
  return true;
}
inline bool delta_sample(V3 weight, V3 wi, real p_lobe, BSample* out) {
  out->weight = weight;
  out->wi = wi;
  out->pdf = p_lobe;
  out->delta = true;
  return true;
}
// the two-lobe (single scattering by bounded VNDF, multiple scattering by
// cosine) sampling of sample_coat / sample_metal / sample_spec_brdf
inline bool sample_two_lobe(const PbrBsdf& p, V3 wo, real ax, real ay, TwoLobe w, real ux, real uy, BSample* out) {
  real ssp = w.ss / w.total;
  if (ux >= ssp) {
    real u = ssp < 1.0f ? (ux - ssp) / (1.0f - ssp) : 0.0f;
    return finalize(p, wo, cosine_hemisphere(u, uy), out);
  }
  real u = ssp > 0.0f ? ux / ssp : 0.0f;
  V3 wm;
  if (!sample_wm_bounded_vndf(wo, ax, ay, u, uy, &wm)) return false;
  V3 wi = reflect_local(wo, wm);
  if (wi.z <= 0.0f) return false;
  return finalize(p, wo, wi, out);
}
inline bool sample_lobe(const PbrBsdf& p, V3 wo, int lobe, real p_lobe, real ux, real uy, BSample* out);
// OpenPbrBsdf::sample (:295-331) with pick_lobe (:2016-2041)
inline bool bsdf_sample(const PbrBsdf& p, V3 wo, real u_lobe, real ux, real uy, BSample* out) {
  if (!(wo.z > 0.0f)) return false;
  Probs pr = lobe_probs(p, wo);
  if (pr.total <= 0.0f) return false;
  real u = u_lobe * pr.total;
  int lobe;  // 0 coat, 1 metal, 2 spec, 3 spec btdf, 4 fuzz, 5 diff, 6 diff btdf
  real plv;
  if (u < pr.coat) {
    lobe = 0;
    plv = pr.coat;
  } else if ((u -= pr.coat) < pr.metal) {
    lobe = 1;
    plv = pr.metal;
  } else if ((u -= pr.metal) < pr.spec) {
    lobe = 2;
    plv = pr.spec;
  } else if ((u -= pr.spec) < pr.spec_btdf) {
    lobe = 3;
    plv = pr.spec_btdf;
  } else if ((u -= pr.spec_btdf) < pr.fuzz) {
    lobe = 4;
    plv = pr.fuzz;
  } else if ((u -= pr.fuzz) < pr.diff) {
    lobe = 5;
    plv = pr.diff;
  } else {
    lobe = 6;
    plv = pr.diff_btdf;
  }
  real p_lobe = plv / pr.total;
  if (p_lobe <= 0.0f) return false;
  return sample_lobe(p, wo, lobe, p_lobe, ux, uy, out);
}
// The picked lobe's sample (the lobe branches of OpenPbrBsdf::sample, :295-331).
inline bool sample_lobe(const PbrBsdf& p, V3 wo, int lobe, real p_lobe, real ux, real uy, BSample* out) {
  out->lobe = lobe;
  out->p_lobe = p_lobe;
  out->glossy = lobe != 5 && lobe != 6;  // delta samples carry DELTA instead (read with !delta)
  out->transmission = false;
  out->eta = 1.0f;
  if (lobe == 0) {  // sample_coat :1467-1530
    if (coat_smooth(p)) {
      V3 wi = v3(-wo.x, -wo.y, wo.z);
      real f = fresnel_dielectric(std::fabs(wi.z), 1.0f, p.coat_eta);
      LayerW w = layer_weights(p, wo, &wi);
      return delta_sample(vsplat(w.coat_amp * f / p_lobe), wi, p_lobe, out);
    }
    return sample_two_lobe(p, wo, p.cax, p.cay, refl_lobe_weights(ms_params(p, p.cax, p.cay, p.coat_eta)), ux, uy, out);
  }
  if (lobe == 1) {  // sample_metal :1532-1587
    if (metal_smooth(p)) {
      V3 wi = v3(-wo.x, -wo.y, wo.z);
      V3 f = metal_fresnel(p, std::fabs(wi.z));
      LayerW w = layer_weights(p, wo, &wi);
      return delta_sample(sdiv(vmul(w.metal, f), p_lobe), wi, p_lobe, out);
    }
    return sample_two_lobe(p, wo, p.ax, p.ay, cond_lobe_weights(p), ux, uy, out);
  }
  if (lobe == 2) {  // sample_spec_brdf :1589-1668
    if (spec_smooth(p)) {
      V3 wi = v3(-wo.x, -wo.y, wo.z);
      V3 f;
      if (p.thin_walled) {
        V3 t;
        thin_wall_coefficients(p, std::fabs(wo.z), &f, &t);
      } else {
        f = spec_brdf_fresnel(p, std::fabs(wi.z));
      }
      LayerW w = layer_weights(p, wo, &wi);
      return delta_sample(sdiv(vmul(w.spec_brdf, f), p_lobe), wi, p_lobe, out);
    }
    if (p.thin_walled) {
      V3 wm;
      if (!sample_wm_bounded_vndf(wo, p.ax, p.ay, ux, uy, &wm)) return false;
      V3 wi = reflect_local(wo, wm);
      if (wi.z <= 0.0f) return false;
      return finalize(p, wo, wi, out);
    }
    return sample_two_lobe(p, wo, p.ax, p.ay, refl_lobe_weights(ms_params(p, p.ax, p.ay, p.specular_eta)), ux, uy, out);
  }
  if (lobe == 3) {  // sample_spec_btdf :1670-1837, no dispersion
    out->transmission = true;
    if (p.thin_walled) {
      V3 r, t;
      thin_wall_coefficients(p, std::fabs(wo.z), &r, &t);
      if (trans_smooth(p)) {  // pass-through
        V3 wi = neg(wo);
        LayerW w = layer_weights(p, wo, &wi);
        return delta_sample(sdiv(vmul(w.spec_btdf, t), p_lobe), wi, p_lobe, out);
      }
      real tax, tay;
      thin_wall_alpha(p, &tax, &tay);
      V3 wm;
      if (!sample_wm_bounded_vndf(wo, tax, tay, ux, uy, &wm)) return false;
      V3 wr = reflect_local(wo, wm);
      if (wr.z <= 0.0f) return false;
      V3 wi = v3(wr.x, wr.y, -wr.z);
      LayerW w = layer_weights(p, wo, &wi);
      real pdf_total = p_lobe * pdf_spec_btdf(p, wo, wi);
      if (pdf_total <= 0.0f) return false;
      V3 f = eval_thin_btdf(p, wo, wi);
      out->weight = sdiv(smul(vmul(w.spec_btdf, f), std::fabs(wi.z)), pdf_total);
      out->wi = wi;
      out->pdf = pdf_total;
      out->delta = false;
      return true;
    }
    const real eta_used = p.transmission_eta, eta_fresnel = p.specular_eta;
    const real eta_rel = p.front_face ? 1.0f / eta_used : eta_used;
    out->eta = eta_rel;
    if (trans_smooth(p)) {
      // refract (R/math/mod.rs:83-102)
      if (wo.z <= 0.0f) return false;
      real s2o = std::fmax(1.0f - wo.z * wo.z, 0.0f);
      real s2t = eta_rel * eta_rel * s2o;
      if (s2t >= 1.0f) return false;
      real ct = std::sqrt(std::fmax(1.0f - s2t, 0.0f));
      V3 wi = normalize_or_zero3(v3(-eta_rel * wo.x, -eta_rel * wo.y, -ct));
      if (len_sq(wi) == 0.0f) return false;
      real f = base_dielectric_fresnel(p, std::fabs(wo.z), eta_used, eta_fresnel);
      real scale = 1.0f / (eta_rel * eta_rel);
      LayerW w = layer_weights(p, wo, &wi);
      V3 num = smul(smul(vmul(w.spec_btdf, vsplat(1.0f)), 1.0f - f), scale);
      return delta_sample(sdiv(num, p_lobe), wi, p_lobe, out);
    }
    TwoLobe w = trans_lobe_weights(ms_params(p, p.ax, p.ay, eta_fresnel));
    real ssp = w.ss / w.total;
    if (ux >= ssp) {
      real uu = ssp < 1.0f ? (ux - ssp) / (1.0f - ssp) : 0.0f;
      V3 up = cosine_hemisphere(uu, uy);
      return finalize(p, wo, v3(up.x, up.y, -up.z), out);
    }
    real uu = ssp > 0.0f ? ux / ssp : 0.0f;
    V3 wm;
    if (!sample_wm_vndf(wo, p.ax, p.ay, uu, uy, &wm)) return false;
    real co = dot(wo, wm);
    if (co <= 0.0f) return false;
    V3 wi;
    if (!refract_about_wm(wo, wm, eta_rel, &wi)) return false;  // refract_about :2043-2058
    if (wi.z >= 0.0f) return false;
    real den = dot(wi, wm) + eta_rel * co;
    if (std::fabs(den) < 1.0e-6f) return false;
    return finalize(p, wo, wi, out);
  }
  if (lobe == 4) {  // sample_fuzz :1929-1934
    V3 wi;
    if (!fuzz_sample(p, wo, ux, uy, &wi)) return false;
    return finalize(p, wo, wi, out);
  }
  if (lobe == 5) {  // sample_diff_brdf :1936-1940 -> EonBsdf::sample (R/bsdf/eon.rs:40-58)
    real pdf;
    V3 wi = sample_eon(wo, p.diffuse_roughness, ux, uy, &pdf);
    if (pdf <= 0.0f || wi.z <= 0.0f) return false;
    return finalize(p, wo, wi, out);
  }
  return false;  // the diffuse BTDF: zero weight (subsurface 0), never picked with p_lobe > 0
}


// ---- table cache --------------------------------------------------------------------------
}  // namespace orc
#include <map>
#include <memory>
#include <mutex>
namespace orc {
inline std::mutex& lut_mutex() {
  static std::mutex m;
  return m;
}
inline const AlbedoLut& albedo_lut_for(real eta) {
  static std::map<real, std::unique_ptr<AlbedoLut>> cache;  // keyed by the (finite) eta value
  const real e = sanitize_eta(eta);
  const real key = e;
  std::lock_guard<std::mutex> g(lut_mutex());
  auto it = cache.find(key);
  if (it == cache.end()) it = cache.emplace(key, std::make_unique<AlbedoLut>(build_albedo_lut(e))).first;
  return *it->second;
}
inline const CondLut& cond_lut() {
  static const CondLut l = build_cond_lut();
  return l;
}
inline const DielLut& diel_lut() {
  static const DielLut l = build_diel_lut();
  return l;
}
inline const AlbedoLut& zero_albedo_lut() {
  static const AlbedoLut l = [] {
    AlbedoLut z;
    z.v.assign(kAlbedoLen, 0.0f);
    return z;
  }();
  return l;
}
inline const CondLut& zero_cond_lut() {
  static const CondLut l = [] {
    CondLut z;
    z.e.assign(kCondCos * kCondRough, 0.0f);
    z.eavg.assign(kCondRough, 0.0f);
    return z;
  }();
  return l;
}
inline const DielLut& zero_diel_lut() {
  static const DielLut l = [] {
    DielLut z;
    z.e.assign(kDielCos * kDielRough * kDielEta, 0.0f);
    z.eavg.assign(kDielRough * kDielEta, 0.0f);
    return z;
  }();
  return l;
}
}  // namespace orc
