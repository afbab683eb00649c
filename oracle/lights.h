// Oracle: spherical-Gaussian light tree and area-light sampling of the
// reference, restated line by line.  TEST INFRASTRUCTURE ONLY.
//
//   R/math/sg.rs:26-346               SG / vMF helpers (expm1_over_x, erf, products, integrals)
//   R/light_tree/builder.rs:58-765    binned SAOH build (triangle leaves)
//   R/light_tree/traversal.rs:26-160  stochastic descent, leaf pmf
//   R/light_tree/lobe.rs:112-163,233-436  make_glossy_lobe, sg_light_for_node, diffuse /
//                                     glossy importance, sggx_reflection_pdf, merge_glossy_roughness
//   R/light/mod.rs:136-233            LightSampler (tree category only)
//   R/light/area.rs:12-54             sample_li_for_triangle
//   R/scene/mod.rs:558-683            area-light table, pdfs, sample_triangle_point
//
// Scope: emissive-triangle leaves (point and spot lights are not part of
// the backend's scene description).  The builder sorts with a stable sort
// where the reference uses an unstable one; the two differ only when two
// leaf centroids tie exactly on the split axis.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "geometry.h"

namespace orc {
namespace lt {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kFltMin = 1.17549435e-38f;        // f32::MIN_POSITIVE
constexpr float kFltEps = 1.1920929e-7f;          // f32::EPSILON
constexpr float kSgSharpnessMax = 2199023255552.0f;  // 2^41
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

inline float fmaxr(float a, float b) { return std::fmax(a, b); }  // f32::max (NaN-ignoring)
inline float fminr(float a, float b) { return std::fmin(a, b); }
inline float clampr(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }  // f32::clamp
inline V3 vscale(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 vdiv(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline float vlen(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 vnormalize(V3 a) { return vscale(a, 1.0f / vlen(a)); }  // glam normalize
inline V3 vnormalize_or_zero(V3 a) {
  float r = 1.0f / vlen(a);
  return (std::isfinite(r) && r > 0.0f) ? vscale(a, r) : v3(0.f, 0.f, 0.f);
}
inline float vmax_elem(V3 a) { return fmaxr(a.x, fmaxr(a.y, a.z)); }
inline float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }  // R/math/sg.rs:314-316

// ---- SG math (R/math/sg.rs) ------------------------------------------------------------
inline float expm1_over_x(float x) {  // :59-69
  float u = std::exp(x);
  if (u == 1.0f) return 1.0f;
  float y = u - 1.0f;
  return std::fabs(x) < 1.0f ? y / std::log(u) : y / x;
}
inline float erf_approx(float x) {  // :74-107
  if (std::fabs(x) >= 4.0f) return std::copysign(1.0f, x);
  if (std::fabs(x) > 1.0f) {
    const float a1 = 1.6284595f, a2 = 9.156747e-1f, a3 = 1.5432939e-1f, a4 = -3.5175983e-2f, a5 = 5.6679556e-3f,
                a6 = -5.648746e-4f, a7 = 2.5890768e-5f;
    float a = std::fabs(x);
    float p = (((((((a7 * a + a6) * a + a5) * a + a4) * a + a3) * a + a2) * a + a1) * a);
    float y = 1.0f - std::exp2(-p);
    return std::copysign(y, x);
  }
  const float a1 = 1.12837917f, a2 = -3.76123e-1f, a3 = 1.1279922e-1f, a4 = -2.6703065e-2f, a5 = 4.9073556e-3f,
              a6 = -5.5885315e-4f;
  float x2 = x * x;
  return (((((a6 * x2 + a5) * x2 + a4) * x2 + a3) * x2 + a2) * x2 + a1) * x;
}
inline float erfc_approx(float x) { return 1.0f - erf_approx(x); }
inline float sg_integral(float s) { return 4.0f * kPiF * expm1_over_x(-2.0f * s); }  // :124-126

struct SgLobe {
  V3 axis;
  float sharpness, log_amplitude;
};
inline SgLobe sg_product(V3 a1, float s1, V3 a2, float s2) {  // :137-156
  V3 axis = add(vscale(a1, s1), vscale(a2, s2));
  float sharp = vlen(axis);
  V3 d = sub(a1, a2);
  float len2 = dot(d, d);
  float denom = fmaxr(sharp + s1 + s2, kFltMin);
  float log_amp = -s1 * s2 * len2 / denom;
  V3 n = sharp > kFltMin ? vdiv(axis, sharp) : v3(0.f, 0.f, 1.f);
  return SgLobe{n, sharp, log_amp};
}
inline float upper_cc_over_2pi(float s) {  // :236-252
  if (s <= 0.5f) {
    return (((((((-1.0f / 362880.0f) * s + 1.0f / 40320.0f) * s - 1.0f / 5040.0f) * s + 1.0f / 720.0f) * s -
              1.0f / 120.0f) * s + 1.0f / 24.0f) * s - 1.0f / 6.0f) * s + 0.5f;
  }
  return (std::expm1(-s) + s) / (s * s);
}
inline float lower_cc_over_2pi(float s) {  // :256-275
  float e = std::exp(-s);
  if (s <= 0.5f) {
    return e * (((((((((1.0f / 403200.0f) * s - 1.0f / 45360.0f) * s + 1.0f / 5760.0f) * s - 1.0f / 840.0f) * s +
                    1.0f / 144.0f) * s - 1.0f / 30.0f) * s + 1.0f / 8.0f) * s - 1.0f / 3.0f) * s + 0.5f);
  }
  return e * (-std::expm1(-s) - s * e) / (s * s);
}
inline float sg_cc_product_over_pi(float z, float s) {  // :282-306
  const float A = 2.736083f, B = 17.021297f, C = 4.0100827f, D = 15.219156f, E = 76.0879f;
  float t = s * std::sqrt(0.5f * ((s + A) * s + B) / (((s + C) * s + D) * s + E));
  float tz = t * z;
  const float inv_sqrt_pi = 0.5641896f;
  const float clamp_min = 0.5f * kFltEps;
  float lerp = clampr(0.5f * (z * erfc_approx(-tz) + erfc_approx(t)) -
                          0.5f * inv_sqrt_pi * std::exp(-tz * tz) * std::expm1(t * t * (z * z - 1.0f)) / fmaxr(t, kFltMin),
                      clamp_min, 1.0f);
  float lower = lower_cc_over_2pi(s);
  float upper = upper_cc_over_2pi(s);
  return 2.0f * (lower + (upper - lower) * lerp);
}
inline float vmf_axis_length_to_sharpness(float len) {  // :196-208
  float r = clampr(len, 0.0f, 1.0f);
  if (r >= 1.0f) return kSgSharpnessMax;
  float om = 1.0f - r * r;
  if (om <= 0.0f) return kSgSharpnessMax;
  float lambda = r * (3.0f - r * r) / om;
  return clampr(lambda, 0.0f, kSgSharpnessMax);
}
struct Merged {
  V3 mu;
  float sigma_s2;
  V3 nu_bar;
  float flux;
};
inline Merged merge_cluster_params(float fl, V3 mul_, float sl, V3 nul, float fr, V3 mur, float sr, V3 nur) {  // :318-346
  float total = fl + fr;
  if (total <= 0.0f) return Merged{v3(0.f, 0.f, 0.f), 0.0f, v3(0.f, 0.f, 0.f), 0.0f};
  float wl = fl / total, wr = fr / total;
  V3 mu = add(vscale(mul_, wl), vscale(mur, wr));
  V3 md = sub(mul_, mur);
  float s2 = wl * sl + wr * sr + wl * wr * dot(md, md);
  V3 nu = add(vscale(nul, wl), vscale(nur, wr));
  return Merged{mu, s2, nu, total};
}

// ---- tree ----------------------------------------------------------------------------
struct Box {
  V3 mn{INFINITY, INFINITY, INFINITY}, mx{-INFINITY, -INFINITY, -INFINITY};
  float area() const {  // Bounds::surface_area (R/scene/mesh.rs:31-34)
    V3 e = vmax(sub(mx, mn), v3(0.f, 0.f, 0.f));
    return 2.0f * (e.x * e.y + e.x * e.z + e.y * e.z);
  }
  Box join(const Box& o) const { return Box{vmin(mn, o.mn), vmax(mx, o.mx)}; }
};

struct Node {
  float flux;
  V3 mu;
  float sigma_s2;
  V3 nu;
  float lambda;
  Box aabb;
  float radius;
  V3 cone_axis;
  float cone_theta_o, cone_theta_e;
  bool leaf;
  uint32_t inst, tri;  // triangle leaf
  float area;          // triangle area (leaf)
  uint32_t left, right, parent;
};

struct Leaf {
  uint32_t inst, tri;
  float flux;
  V3 mu;
  float sigma_s2;
  V3 nu_bar;
  Box aabb;
  float radius;
  V3 cone_axis;
  float theta_o, theta_e;
  V3 centroid;
  float area;
};

struct Cone {  // OrientationCone (R/light_tree/builder.rs:240-338)
  V3 axis{0.f, 0.f, 1.f};
  float theta_o = 0.0f, theta_e = 0.0f;
  bool valid = false;
  static Cone of(const Leaf& l) { return Cone{l.cone_axis, l.theta_o, l.theta_e, true}; }
  float measure() const {
    if (!valid) return 0.0f;
    float to = clampr(theta_o, 0.0f, kPiF), te = clampr(theta_e, 0.0f, kPiF * 0.5f);
    float tw = fminr(to + te, kPiF);
    float co = std::cos(to), so = std::sin(to);
    return 2.0f * kPiF * (1.0f - co) +
           0.5f * kPiF * (2.0f * tw * so - std::cos(to - 2.0f * tw) - 2.0f * to * so + co);
  }
  static Cone merge(Cone a, Cone b) {
    if (!a.valid) return b;
    if (!b.valid) return a;
    float te = fmaxr(a.theta_e, b.theta_e);
    if (!(a.theta_o >= b.theta_o)) std::swap(a, b);
    float cd = clampr(dot(a.axis, b.axis), -1.0f, 1.0f);
    float td = std::acos(cd);
    if (fminr(td + b.theta_o, kPiF) <= a.theta_o) return Cone{a.axis, a.theta_o, te, true};
    float to = fminr((a.theta_o + td + b.theta_o) * 0.5f, kPiF);
    if (to >= kPiF - 1e-5f) return Cone{a.axis, kPiF, te, true};
    V3 ra = cross(a.axis, b.axis);
    V3 na;
    if (dot(ra, ra) > 1.0e-12f) {
      float tr = to - a.theta_o;
      V3 k = vnormalize(ra);
      float cr = std::cos(tr), sr = std::sin(tr);
      na = vnormalize_or_zero(add(add(vscale(a.axis, cr), vscale(cross(k, a.axis), sr)),
                                  vscale(vscale(k, dot(k, a.axis)), 1.0f - cr)));
    } else {
      na = a.axis;
    }
    return Cone{dot(na, na) > 0.0f ? na : a.axis, to, te, true};
  }
};

struct Cluster {  // AggregatedCluster (R/light_tree/builder.rs:350-437)
  float flux = 0.0f;
  V3 mu{0.f, 0.f, 0.f};
  float sigma_s2 = 0.0f;
  V3 nu_bar{0.f, 0.f, 0.f};
  Box aabb;
  float radius = 0.0f;
  Cone cone;
  void add_leaf(const Leaf& l) {
    float nf = flux + l.flux;
    if (nf <= 0.0f) return;
    if (flux <= 0.0f) {
      flux = l.flux;
      mu = l.mu;
      sigma_s2 = l.sigma_s2;
      nu_bar = l.nu_bar;
      aabb = l.aabb;
      cone = Cone::of(l);
      return;
    }
    Merged m = merge_cluster_params(flux, mu, sigma_s2, nu_bar, l.flux, l.mu, l.sigma_s2, l.nu_bar);
    flux = m.flux;
    mu = m.mu;
    sigma_s2 = m.sigma_s2;
    nu_bar = m.nu_bar;
    aabb = aabb.join(l.aabb);
    cone = Cone::merge(cone, Cone::of(l));
  }
  void add(const Cluster& o) {
    if (o.flux <= 0.0f) return;
    if (flux <= 0.0f) {
      *this = o;
      return;
    }
    Merged m = merge_cluster_params(flux, mu, sigma_s2, nu_bar, o.flux, o.mu, o.sigma_s2, o.nu_bar);
    flux = m.flux;
    mu = m.mu;
    sigma_s2 = m.sigma_s2;
    nu_bar = m.nu_bar;
    aabb = aabb.join(o.aabb);
    cone = Cone::merge(cone, o.cone);
  }
  void finish_radius() {
    if (flux <= 0.0f) {
      radius = 0.0f;
      return;
    }
    V3 a = sub(aabb.mx, mu), b = sub(aabb.mn, mu);
    V3 ext = vmax(v3(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)), v3(std::fabs(b.x), std::fabs(b.y), std::fabs(b.z)));
    radius = vlen(ext);
  }
};

inline float axis_of(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// triangle_leaf (R/light_tree/builder.rs:123-185); positions in world space.
inline bool triangle_leaf(uint32_t inst, uint32_t tri, V3 p0, V3 p1, V3 p2, float max_emission, Leaf* out) {
  V3 e1 = sub(p1, p0), e2 = sub(p2, p0);
  V3 gn = cross(e1, e2);
  float ad = vlen(gn);
  if (ad <= 0.0f) return false;
  V3 ng = vdiv(gn, ad);
  float area = 0.5f * ad;
  float phi = max_emission * area;
  if (phi <= 0.0f) return false;
  Leaf l;
  l.inst = inst;
  l.tri = tri;
  l.flux = phi;
  l.mu = vdiv(add(add(p0, p1), p2), 3.0f);
  l.sigma_s2 = (dot(e1, e1) + dot(e2, e2) - dot(e1, e2)) / 18.0f;
  l.radius = fmaxr(fmaxr(vlen(sub(p0, l.mu)), vlen(sub(p1, l.mu))), vlen(sub(p2, l.mu)));
  l.aabb = Box{vmin(vmin(p0, p1), p2), vmax(vmax(p0, p1), p2)};
  l.nu_bar = vscale(ng, 0.5f);
  l.cone_axis = ng;
  l.theta_o = 0.0f;
  l.theta_e = kPiF * 0.5f;
  l.centroid = l.mu;
  l.area = area;
  *out = l;
  return true;
}

inline size_t bin_of(float v) {  // `as usize` saturates; NaN -> 0
  if (!(v > 0.0f)) return 0;
  if (v >= 12.0f) return 11;
  return std::min<size_t>((size_t)v, 11);
}

// find_best_binned_split (R/light_tree/builder.rs:656-765).  Returns 0 for None.
inline size_t best_split(Leaf* lv, size_t n, const Cluster& total) {
  if (n < 2) return 0;
  V3 cmin{INFINITY, INFINITY, INFINITY}, cmax{-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; ++i) {
    cmin = vmin(cmin, lv[i].centroid);
    cmax = vmax(cmax, lv[i].centroid);
  }
  V3 ext = sub(cmax, cmin);
  if (vmax_elem(ext) <= 0.0f) return 0;
  float parea = fmaxr(total.aabb.area(), 1.0e-30f);
  float pcone = fmaxr(total.cone.measure(), 1.0e-30f);
  float best = INFINITY;
  int best_axis = 0;
  size_t best_at = 0;
  for (int a = 0; a < 3; ++a) {
    if (axis_of(ext, a) <= 0.0f) continue;
    Cluster bins[12];
    uint32_t cnt[12] = {0};
    float inv = 12.0f / axis_of(ext, a);
    float lo = axis_of(cmin, a);
    for (size_t i = 0; i < n; ++i) {
      size_t b = bin_of((axis_of(lv[i].centroid, a) - lo) * inv);
      bins[b].add_leaf(lv[i]);
      cnt[b]++;
    }
    Cluster pre[12], suf[12];
    uint32_t pc[12], sc[12];
    Cluster acc;
    uint32_t ac = 0;
    for (int i = 0; i < 12; ++i) {
      acc.add(bins[i]);
      ac += cnt[i];
      pre[i] = acc;
      pc[i] = ac;
    }
    Cluster sacc;
    uint32_t scn = 0;
    for (int i = 11; i >= 0; --i) {
      sacc.add(bins[i]);
      scn += cnt[i];
      suf[i] = sacc;
      sc[i] = scn;
    }
    for (int i = 0; i < 12; ++i) {
      pre[i].finish_radius();
      suf[i].finish_radius();
    }
    for (int s = 1; s < 12; ++s) {
      const Cluster& l = pre[s - 1];
      const Cluster& r = suf[s];
      if (pc[s - 1] == 0 || sc[s] == 0) continue;
      float cost = l.flux * l.aabb.area() * l.cone.measure() + r.flux * r.aabb.area() * r.cone.measure();
      if (!std::isfinite(cost)) continue;
      float norm = cost / (parea * pcone);
      if (norm < best) {
        best = norm;
        best_axis = a;
        best_at = pc[s - 1];
      }
    }
  }
  if (!std::isfinite(best)) return 0;
  std::stable_sort(lv, lv + n, [&](const Leaf& x, const Leaf& y) {
    return axis_of(x.centroid, best_axis) < axis_of(y.centroid, best_axis);
  });
  return std::min(std::max<size_t>(best_at, 1), n - 1);
}

inline uint32_t push_leaf(std::vector<Node>& nodes, const Leaf& l) {  // :591-617
  Node nd{};
  nd.flux = l.flux;
  nd.mu = l.mu;
  nd.sigma_s2 = l.sigma_s2;
  nd.lambda = vmf_axis_length_to_sharpness(vlen(l.nu_bar));
  nd.nu = dot(l.nu_bar, l.nu_bar) > 0.0f ? vnormalize(l.nu_bar) : v3(0.f, 0.f, 1.f);
  nd.aabb = l.aabb;
  nd.radius = l.radius;
  nd.cone_axis = l.cone_axis;
  nd.cone_theta_o = l.theta_o;
  nd.cone_theta_e = l.theta_e;
  nd.leaf = true;
  nd.inst = l.inst;
  nd.tri = l.tri;
  nd.area = l.area;
  nd.left = nd.right = nd.parent = kInvalid;
  nodes.push_back(nd);
  return (uint32_t)nodes.size() - 1;
}

inline uint32_t build_rec(Leaf* lv, size_t n, std::vector<Node>& nodes) {  // :439-560
  if (n <= 1) return push_leaf(nodes, lv[0]);
  Cluster total;
  for (size_t i = 0; i < n; ++i) total.add_leaf(lv[i]);
  total.finish_radius();
  size_t split = best_split(lv, n, total);
  if (split == 0) {
    V3 mn{INFINITY, INFINITY, INFINITY}, mx{-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i) {
      mn = vmin(mn, lv[i].centroid);
      mx = vmax(mx, lv[i].centroid);
    }
    V3 d = sub(mx, mn);
    int axis = (d.x >= d.y && d.x >= d.z) ? 0 : (d.y >= d.z ? 1 : 2);
    std::stable_sort(lv, lv + n,
                     [&](const Leaf& x, const Leaf& y) { return axis_of(x.centroid, axis) < axis_of(y.centroid, axis); });
    split = n / 2;
  }
  split = std::min(std::max<size_t>(split, 1), n - 1);
  uint32_t idx = (uint32_t)nodes.size();
  nodes.push_back(Node{});
  uint32_t l = build_rec(lv, split, nodes);
  uint32_t r = build_rec(lv + split, n - split, nodes);
  Node& nd = nodes[idx];
  nd.left = l;
  nd.right = r;
  nd.leaf = false;
  nd.inst = nd.tri = kInvalid;
  nd.area = 0.0f;
  nd.lambda = vmf_axis_length_to_sharpness(vlen(total.nu_bar));
  nd.nu = dot(total.nu_bar, total.nu_bar) > 0.0f ? vnormalize(total.nu_bar) : v3(0.f, 0.f, 1.f);
  nd.flux = total.flux;
  nd.mu = total.mu;
  nd.sigma_s2 = total.sigma_s2;
  nd.aabb = total.aabb;
  nd.radius = total.radius;
  nd.cone_axis = total.cone.axis;
  nd.cone_theta_o = total.cone.theta_o;
  nd.cone_theta_e = total.cone.theta_e;
  return idx;
}

struct Tree {
  std::vector<Node> nodes;
  uint32_t root = 0;
  float root_flux() const { return nodes.empty() ? 0.0f : nodes[root].flux; }
};

inline Tree build_tree(std::vector<Leaf> leaves) {  // build_light_tree (:58-97)
  Tree t;
  std::vector<Leaf> kept;
  for (const Leaf& l : leaves)
    if (l.flux > 0.0f) kept.push_back(l);
  if (kept.empty()) return t;
  t.nodes.reserve(2 * kept.size() - 1);
  t.root = build_rec(kept.data(), kept.size(), t.nodes);
  t.nodes[t.root].parent = kInvalid;
  for (size_t i = 0; i < t.nodes.size(); ++i) {
    if (t.nodes[i].leaf) continue;
    t.nodes[t.nodes[i].left].parent = (uint32_t)i;
    t.nodes[t.nodes[i].right].parent = (uint32_t)i;
  }
  return t;
}

// ---- importance and traversal ------------------------------------------------------------
// sg_light_for_node (R/light_tree/lobe.rs:233-274)
inline bool sg_light_for_node(V3 p, V3 n, const Node& nd, float* w, SgLobe* lobe) {
  if (nd.flux <= 0.0f) return false;
  V3 d = sub(nd.mu, p);
  float r2 = dot(d, d);
  if (r2 <= 0.0f) return false;
  float inv_r = 1.0f / std::sqrt(r2);
  float cos_below = fmaxr(-(dot(n, d)) * inv_r, 0.0f);
  float bound_var = 0.5f * nd.radius * nd.radius;
  float sigma2 = nd.sigma_s2 * (1.0f - cos_below) + bound_var * cos_below;
  float min_sigma2 = r2 / kSgSharpnessMax;
  if (sigma2 < min_sigma2) sigma2 = min_sigma2;
  V3 axis = vscale(d, inv_r);
  float sharp = clampr(r2 / sigma2, 0.0f, kSgSharpnessMax);
  *lobe = sg_product(vneg(nd.nu), nd.lambda, axis, sharp);
  float sg_norm = fmaxr(sg_integral(nd.lambda), 1.0e-30f);
  float w_amp = nd.flux / (2.0f * kPiF * sigma2 * sg_norm);
  *w = w_amp * std::exp(lobe->log_amplitude);
  return std::isfinite(*w) && *w > 0.0f;
}

// diffuse_importance (R/light_tree/lobe.rs:276-287)
inline float diffuse_importance(float rho, V3 n, float w, const SgLobe& lobe) {
  if (rho <= 0.0f || w <= 0.0f) return 0.0f;
  float c = clampr(dot(lobe.axis, n), -1.0f, 1.0f);
  float integral = sg_cc_product_over_pi(c, lobe.sharpness) * kPiF;
  return rho * w * integral;
}

// ---- glossy lobe (R/light_tree/lobe.rs:112-163, 290-341, 396-436) -----------------
// sg_normalized_hemispherical_integral (R/math/sg.rs:155-171)
inline float sg_normalized_hemispherical_integral(float cosine, float sharpness) {
  const float A = 0.65173288f, B = 1.341828f, C = 7.221669f;
  float steepness = sharpness * std::sqrt((0.5f * sharpness + A) / ((sharpness + B) * sharpness + C));
  cosine = clampr(cosine, -1.0f, 1.0f);
  float denom = erf_approx(steepness);
  float numer = erf_approx(steepness * cosine);
  float v = std::fabs(denom) > 0.0f ? 0.5f + 0.5f * (numer / denom) : 0.5f;
  return clampr(v, 0.0f, 1.0f);
}
// vmf_hemispherical_integral (R/math/sg.rs:189-194)
inline float vmf_hemispherical_integral(float cosine, float sharpness) {
  float v = sg_normalized_hemispherical_integral(cosine, sharpness);
  float e = std::exp(-sharpness);
  float num = e + (1.0f - e) * v;
  return num / (e + 1.0f);
}
// glam Mat2: two columns; a * b takes column j of b through a
struct M2 {
  float c0x, c0y, c1x, c1y;
};
inline M2 m2_mul(const M2& a, const M2& b) {
  return M2{a.c0x * b.c0x + a.c1x * b.c0y, a.c0y * b.c0x + a.c1y * b.c0y, a.c0x * b.c1x + a.c1x * b.c1y,
            a.c0y * b.c1x + a.c1y * b.c1y};
}
inline M2 m2_transpose(const M2& a) { return M2{a.c0x, a.c1x, a.c0y, a.c1y}; }
// OrthonormalBasis: tangent, bitangent, normal
struct Frame3 {
  V3 t, b, n;
};
inline V3 frame_to_local(const Frame3& f, V3 w) { return v3(dot(w, f.t), dot(w, f.b), dot(w, f.n)); }
struct GlossyLobe {  // GlossyLobePrecompute (R/light_tree/lobe.rs:36-46)
  float rho;
  V3 wi_ts;
  M2 jj;
  float det_jj4;
  float beta_x, beta_y;
  float alpha2_max;
  V3 xi_p;
  float kappa_p;
};
// make_glossy_lobe (R/light_tree/lobe.rs:112-163); false for None
inline bool make_glossy_lobe(float rho, const Frame3& frame, V3 wo_world, float alpha_x, float alpha_y,
                             GlossyLobe* g) {
  if (rho <= 0.0f) return false;
  V3 wi_ts = frame_to_local(frame, wo_world);
  if (wi_ts.z <= 0.0f) return false;
  float vlen = std::sqrt(wi_ts.x * wi_ts.x + wi_ts.y * wi_ts.y);
  float vx = vlen > 0.0f ? wi_ts.x / vlen : 1.0f, vy = vlen > 0.0f ? wi_ts.y / vlen : 0.0f;
  M2 r{vx, vy, -vy, vx};
  M2 d{0.5f, 0.0f, 0.0f, 0.5f / fmaxr(wi_ts.z, 1.0e-4f)};
  M2 jac = m2_mul(r, d);
  g->jj = m2_mul(jac, m2_transpose(jac));
  g->det_jj4 = 1.0f / (4.0f * wi_ts.z * wi_ts.z);
  float a2x = clampr(alpha_x * alpha_x, 0.0f, 1.0f - 1.0e-4f);
  float a2y = clampr(alpha_y * alpha_y, 0.0f, 1.0f - 1.0e-4f);
  g->beta_x = a2x / (1.0f - a2x);
  g->beta_y = a2y / (1.0f - a2y);
  g->alpha2_max = fmaxr(fmaxr(a2x, a2y), 1.0e-6f);
  V3 n = frame.n;
  g->xi_p = vnormalize_or_zero(add(vneg(wo_world), vscale(n, 2.0f * dot(wo_world, n))));
  g->kappa_p = clampr((1.0f - g->alpha2_max) / (2.0f * g->alpha2_max), 0.0f, kSgSharpnessMax);
  g->rho = rho;
  g->wi_ts = wi_ts;
  return true;
}
// sggx_reflection_pdf (R/light_tree/lobe.rs:396-419)
inline float sggx_reflection_pdf(V3 wi, V3 h, const M2& m) {
  float det = fmaxr(m.c0x * m.c1y - m.c0y * m.c1x, 1.0e-30f);
  M2 adj{m.c1y, -m.c0y, -m.c1x, m.c0x};
  // adj * m_xy
  float ax = adj.c0x * h.x + adj.c1x * h.y, ay = adj.c0y * h.x + adj.c1y * h.y;
  float length2 = (h.x * ax + h.y * ay) / det + h.z * h.z;
  float sggx = 1.0f / (kPiF * std::sqrt(det) * (length2 * length2));
  float inv_norm = std::sqrt(fmaxr(wi.x * wi.x * m.c0x + 2.0f * wi.x * wi.y * m.c0y + wi.y * wi.y * m.c1y + wi.z * wi.z,
                                   0.0f));
  return sggx / (4.0f * fmaxr(inv_norm, 1.0e-30f));
}
// glossy_importance (R/light_tree/lobe.rs:290-341)
inline float glossy_importance(const GlossyLobe& g, const Frame3& frame, V3 n, float w, const SgLobe& lobe) {
  if (g.rho <= 0.0f || w <= 0.0f) return 0.0f;
  float inv_kappa = 1.0f / fmaxr(lobe.sharpness, 1.0e-30f);
  M2 fp{g.beta_x + g.jj.c0x * inv_kappa, 0.0f + g.jj.c0y * inv_kappa, 0.0f + g.jj.c1x * inv_kappa,
        g.beta_y + g.jj.c1y * inv_kappa};
  float det = g.beta_x * g.beta_y + 2.0f * inv_kappa * (g.beta_x * g.jj.c1y + g.beta_y * g.jj.c0x) +
              inv_kappa * inv_kappa * g.det_jj4 / 4.0f;
  float tr = fp.c0x + fp.c1y;
  float denom = fmaxr(1.0f + tr + det, 1.0e-12f);
  M2 a_bar{(fp.c0x + det) / denom, fp.c0y / denom, fp.c1x / denom, (fp.c1y + det) / denom};
  V3 xi_ts = frame_to_local(frame, lobe.axis);
  V3 hu = add(g.wi_ts, xi_ts);
  if (dot(hu, hu) <= 1.0e-20f) return 0.0f;
  V3 h = vnormalize(hu);
  float lobe_value = sggx_reflection_pdf(g.wi_ts, h, a_bar);
  V3 prod = add(vscale(g.xi_p, g.kappa_p), vscale(lobe.axis, lobe.sharpness));
  float prod_sharp = vlen(prod);
  if (prod_sharp <= 0.0f) return 0.0f;
  V3 prod_dir = vdiv(prod, prod_sharp);
  float v_vis = vmf_hemispherical_integral(clampr(dot(prod_dir, n), -1.0f, 1.0f), prod_sharp);
  float sg_full = sg_integral(lobe.sharpness);
  return g.rho * w * v_vis * lobe_value * sg_full;
}
// merge_glossy_roughness (R/light_tree/lobe.rs:421-436)
inline void merge_glossy_roughness(float rho_a, float ax_a, float ay_a, float rho_b, float ax_b, float ay_b,
                                   float* ax, float* ay) {
  float total = fmaxr(rho_a + rho_b, 1.0e-30f);
  float wa = rho_a / total, wb = rho_b / total;
  float a2x = wa * ax_a * ax_a + wb * ax_b * ax_b;
  float a2y = wa * ay_a * ay_a + wb * ay_b * ay_b;
  *ax = std::sqrt(a2x);
  *ay = std::sqrt(a2y);
}

// A light-tree query (LightTreePrecompute, R/light_tree/lobe.rs:20-28): a
// diffuse lobe (rho > 0) and / or a glossy lobe (has_glossy) in `frame`.
struct Query {
  V3 p, n;
  float rho;
  bool has_glossy = false;
  Frame3 frame{};
  GlossyLobe glossy{};
};
// evaluate_node_importance (R/light_tree/importance.rs:41-48) with the
// material's light_tree_importance: the diffuse lobe alone
// (R/material/normalized_lambert.rs), or the lobes summed and clamped at
// zero (R/material/standard_surface.rs:700-717).
inline float node_importance(const Query& q, const Node& nd) {
  float w;
  SgLobe lobe;
  if (!sg_light_for_node(q.p, q.n, nd, &w, &lobe)) return 0.0f;
  float imp = diffuse_importance(q.rho, q.n, w, lobe);
  if (!q.has_glossy) return imp;
  return fmaxr(imp + glossy_importance(q.glossy, q.frame, q.n, w, lobe), 0.0f);
}

// sample_light_tree (R/light_tree/traversal.rs:26-92)
inline bool sample_tree(const Tree& t, const Query& q, float u, uint32_t* leaf_node, float* pmf_out) {
  if (t.nodes.empty()) return false;
  u = clampr(u, 0.0f, 1.0f - kFltEps);
  uint32_t ni = t.root;
  float pmf = 1.0f;
  for (;;) {
    const Node& nd = t.nodes[ni];
    if (nd.leaf) {
      *leaf_node = ni;
      *pmf_out = pmf;
      return true;
    }
    const Node& L = t.nodes[nd.left];
    const Node& R = t.nodes[nd.right];
    float il = node_importance(q, L), ir = node_importance(q, R);
    float total = il + ir;
    if (total <= 0.0f) {
      float lf = fmaxr(L.flux, 0.0f), rf = fmaxr(R.flux, 0.0f);
      float tf = lf + rf;
      if (tf <= 0.0f) return false;
      float pl = lf / tf;
      if (u < pl) {
        u /= fmaxr(pl, kFltMin);
        pmf *= pl;
        ni = nd.left;
      } else {
        u = (u - pl) / fmaxr(1.0f - pl, kFltMin);
        pmf *= 1.0f - pl;
        ni = nd.right;
      }
      continue;
    }
    float pl = il / total;
    if (u < pl) {
      u /= fmaxr(pl, kFltMin);
      pmf *= pl;
      ni = nd.left;
    } else {
      u = (u - pl) / fmaxr(1.0f - pl, kFltMin);
      pmf *= 1.0f - pl;
      ni = nd.right;
    }
    u = clampr(u, 0.0f, 1.0f - kFltEps);
  }
}

// pdf_for_leaf (R/light_tree/traversal.rs:97-140)
inline float pdf_leaf(const Tree& t, const Query& q, uint32_t leaf_node) {
  if (t.nodes.empty() || leaf_node == kInvalid) return 0.0f;
  if (!t.nodes[leaf_node].leaf) return 0.0f;
  float pmf = 1.0f;
  uint32_t child = leaf_node, parent = t.nodes[leaf_node].parent;
  while (parent != kInvalid) {
    const Node& pn = t.nodes[parent];
    const Node& L = t.nodes[pn.left];
    const Node& R = t.nodes[pn.right];
    float il = node_importance(q, L), ir = node_importance(q, R);
    float total = il + ir;
    float pl;
    if (total <= 0.0f) {
      float lf = fmaxr(L.flux, 0.0f), rf = fmaxr(R.flux, 0.0f);
      float tf = lf + rf;
      if (tf <= 0.0f) return 0.0f;
      pl = lf / tf;
    } else {
      pl = il / total;
    }
    pmf *= child == pn.left ? pl : 1.0f - pl;
    child = parent;
    parent = pn.parent;
  }
  return pmf;
}


// ---- environment light (R/light/environment.rs) ----------------------------------------
namespace env {

constexpr float kOneMinusEps = 1.0f - kFltEps;
constexpr float kTauF = 6.28318530717958647692f;

struct Level {
  size_t w, h;
  std::vector<float> wt;
};
struct Dist {  // HierarchicalDistribution: levels[0] is 1x1, last is the padded leaf
  std::vector<Level> levels;
  float total = 0.0f;
  float texel(size_t i, size_t j) const { return levels.back().wt[j * levels.back().w + i]; }
};

inline size_t next_pow2(size_t v) {
  size_t p = 1;
  while (p < v) p <<= 1;
  return p;
}
inline float row_sin_theta(size_t j, size_t h) { return std::sin(((float)j + 0.5f) / (float)h * kPiF); }
inline float positive_luminance(V3 c) { return fmaxr(lt::luminance(c), 0.0f); }

inline Dist build_distribution(size_t w, size_t h, const std::vector<float>& weights) {  // :426-478
  Dist d;
  const size_t pw = next_pow2(w), ph = next_pow2(h);
  Level leaf{pw, ph, std::vector<float>(pw * ph, 0.0f)};
  for (size_t j = 0; j < h; ++j)
    for (size_t i = 0; i < w; ++i) leaf.wt[j * pw + i] = fmaxr(weights[j * w + i], 0.0f);
  std::vector<Level> lv{leaf};
  while (lv.back().w > 1 || lv.back().h > 1) {
    const Level& f = lv.back();
    const size_t si = f.w > 1 ? 2 : 1, sj = f.h > 1 ? 2 : 1;
    Level c{f.w / si, f.h / sj, {}};
    c.wt.assign(c.w * c.h, 0.0f);
    for (size_t j = 0; j < c.h; ++j)
      for (size_t i = 0; i < c.w; ++i) {
        float sum = 0.0f;
        for (size_t dj = 0; dj < sj; ++dj)
          for (size_t di = 0; di < si; ++di) sum += f.wt[(j * sj + dj) * f.w + i * si + di];
        c.wt[j * c.w + i] = sum;
      }
    lv.push_back(c);
  }
  d.levels.assign(lv.rbegin(), lv.rend());
  d.total = d.levels[0].wt[0];
  return d;
}

struct Env {
  size_t w = 0, h = 0;
  float scale = 1.0f;
  float rot[9], inv_rot[9];  // Mat3 columns
  std::vector<V3> px;
  Dist dist, dist_mis;
};

inline void rot_y(float angle, float* m) {  // Mat3::from_rotation_y, columns
  const float s = std::sin(angle), c = std::cos(angle);
  const float v[9] = {c, 0.0f, -s, 0.0f, 1.0f, 0.0f, s, 0.0f, c};
  for (int k = 0; k < 9; ++k) m[k] = v[k];
}
inline V3 mat3_mul(const float* m, V3 v) {  // Mat3 * Vec3: (x*c0 + y*c1) + z*c2
  return add(add(vscale(v3(m[0], m[1], m[2]), v.x), vscale(v3(m[3], m[4], m[5]), v.y)), vscale(v3(m[6], m[7], m[8]), v.z));
}

inline Env make(size_t w, size_t h, std::vector<V3> px, float scale, float rotate_y) {  // from_pixels (:47-83)
  Env e;
  e.w = w;
  e.h = h;
  e.scale = scale;
  e.px = std::move(px);
  std::vector<float> wt(w * h), wm(w * h);
  for (size_t j = 0; j < h; ++j) {  // environment_weights (:362-371)
    const float st = row_sin_theta(j, h);
    for (size_t i = 0; i < w; ++i) wt[j * w + i] = positive_luminance(e.px[j * w + i]) * st;
  }
  float ws = 0.0f, wsum = 0.0f;  // solid_angle_mean_luminance (:393-408)
  for (size_t j = 0; j < h; ++j) {
    const float st = row_sin_theta(j, h);
    wsum += st * (float)w;
    for (size_t i = 0; i < w; ++i) ws += positive_luminance(e.px[j * w + i]) * st;
  }
  const float mean = wsum > 0.0f ? ws / wsum : 0.0f;
  const float eps = fmaxr(std::fabs(mean), 1.0f) * 1.0e-6f;  // mis_compensated_environment_weights (:373-391)
  for (size_t j = 0; j < h; ++j) {
    const float st = row_sin_theta(j, h);
    for (size_t i = 0; i < w; ++i) {
      const float c = positive_luminance(e.px[j * w + i]) - mean;
      wm[j * w + i] = c > eps ? c * st : 0.0f;
    }
  }
  e.dist = build_distribution(w, h, wt);
  e.dist_mis = build_distribution(w, h, wm);
  rot_y(rotate_y, e.rot);
  rot_y(-rotate_y, e.inv_rot);
  return e;
}

inline float total_power(const Env& e) {  // :129-132
  const float texel = kTauF * kPiF / ((float)e.w * (float)e.h);
  return e.dist.total * e.scale * texel;
}
inline size_t pixel_coord(float t, size_t n) {  // :326-330
  const float c = clampr(t, 0.0f, 1.0f);
  const float f = c * (float)n;
  size_t idx = f >= 1.8446744e19f ? ~size_t(0) : (size_t)f;
  return std::min(idx, n - 1);
}
inline V3 uv_to_direction(float u, float v) {  // :332-337
  const float phi = u * kTauF, theta = v * kPiF;
  const float st = std::sin(theta);
  return v3(st * std::sin(phi), std::cos(theta), st * std::cos(phi));
}
inline void direction_to_uv(V3 d, float* u, float* v) {  // :339-348
  const V3 dir = vnormalize_or_zero(d);
  const float y = clampr(dir.y, -1.0f, 1.0f);
  const float theta = std::acos(y);
  float phi = std::atan2(dir.x, dir.z);
  if (phi < 0.0f) phi += kTauF;
  *u = clampr(phi / kTauF, 0.0f, 1.0f);
  *v = clampr(theta / kPiF, 0.0f, 1.0f);
}
inline V3 radiance(const Env& e, V3 dir) {  // :134-140
  float u, v;
  direction_to_uv(mat3_mul(e.inv_rot, dir), &u, &v);
  const V3 p = e.px[pixel_coord(v, e.h) * e.w + pixel_coord(u, e.w)];
  return vscale(p, e.scale);
}
inline float pdf(const Env& e, const Dist& d, V3 dir) {  // pdf_with_distribution (:156-177)
  if (d.total <= 0.0f) return 0.0f;
  float u, v;
  direction_to_uv(mat3_mul(e.inv_rot, dir), &u, &v);
  const float st = std::sin(v * kPiF);
  if (st <= 0.0f) return 0.0f;
  const float wgt = d.texel(pixel_coord(u, e.w), pixel_coord(v, e.h));
  if (wgt <= 0.0f) return 0.0f;
  const float p2d = wgt / d.total * (float)e.w * (float)e.h;
  return p2d / (kTauF * kPiF * st);
}
// hsw_step (:480-530)
inline bool hsw_step(const Level& parent, const Level& child, size_t* i, size_t* j, float* u, float* v) {
  const size_t si = child.w / parent.w, sj = child.h / parent.h;
  size_t ci = *i * si, cj = *j * sj;
  if (sj == 2) {
    float top = 0.0f, bot = 0.0f;
    for (size_t di = 0; di < si; ++di) {
      top += child.wt[cj * child.w + ci + di];
      bot += child.wt[(cj + 1) * child.w + ci + di];
    }
    const float total = top + bot;
    if (total <= 0.0f) return false;
    const float pt = top / total;
    if (*v < pt) {
      *v = clampr(*v / pt, 0.0f, kOneMinusEps);
    } else {
      *v = clampr((*v - pt) / (1.0f - pt), 0.0f, kOneMinusEps);
      cj += 1;
    }
  }
  if (si == 2) {
    const float l = child.wt[cj * child.w + ci], r = child.wt[cj * child.w + ci + 1];
    const float total = l + r;
    if (total <= 0.0f) return false;
    const float pl = l / total;
    if (*u < pl) {
      *u = clampr(*u / pl, 0.0f, kOneMinusEps);
    } else {
      *u = clampr((*u - pl) / (1.0f - pl), 0.0f, kOneMinusEps);
      ci += 1;
    }
  }
  *i = ci;
  *j = cj;
  return true;
}
struct Sample {
  V3 dir, radiance;
  float pdf;
};
// sample_with_distribution (:179-234)
inline bool sample(const Env& e, const Dist& d, float us0, float us1, Sample* out) {
  if (d.total <= 0.0f) return false;
  float u = clampr(us0, 0.0f, 1.0f), v = clampr(us1, 0.0f, 1.0f);
  size_t i = 0, j = 0;
  for (size_t l = 1; l < d.levels.size(); ++l)
    if (!hsw_step(d.levels[l - 1], d.levels[l], &i, &j, &u, &v)) return false;
  if (i >= e.w || j >= e.h) return false;
  const float wgt = d.texel(i, j);
  if (wgt <= 0.0f) return false;
  const float uc = ((float)i + u) / (float)e.w, vc = ((float)j + v) / (float)e.h;
  const V3 local = uv_to_direction(clampr(uc, 0.0f, kOneMinusEps), clampr(vc, 0.0f, kOneMinusEps));
  out->dir = mat3_mul(e.rot, local);
  const float st = std::sin(vc * kPiF);
  if (st <= 0.0f) return false;
  const float p2d = wgt / d.total * (float)e.w * (float)e.h;
  out->pdf = p2d / (kTauF * kPiF * st);
  out->radiance = vscale(e.px[j * e.w + i], e.scale);
  return true;
}

}  // namespace env

}  // namespace lt
}  // namespace orc
