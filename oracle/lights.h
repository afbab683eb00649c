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

constexpr real kPiF = 3.14159265358979323846f;
constexpr real kFltMin = 1.17549435e-38f;        // f32::MIN_POSITIVE
constexpr real kFltEps = 1.1920929e-7f;          // f32::EPSILON
constexpr real kSgSharpnessMax = 2199023255552.0f;  // 2^41
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

inline real fmaxr(real a, real b) { return std::fmax(a, b); }  // f32::max (NaN-ignoring)
inline real fminr(real a, real b) { return std::fmin(a, b); }
inline real clampr(real x, real lo, real hi) { return x < lo ? lo : (x > hi ? hi : x); }  // f32::clamp
inline V3 vscale(V3 a, real s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 vdiv(V3 a, real s) { return v3(a.x / s, a.y / s, a.z / s); }
inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline real vlen(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 vnormalize(V3 a) { return vscale(a, 1.0f / vlen(a)); }  // glam normalize
inline V3 vnormalize_or_zero(V3 a) {
  real r = 1.0f / vlen(a);
  return (std::isfinite(r) && r > 0.0f) ? vscale(a, r) : v3(0.f, 0.f, 0.f);
}
inline real vmax_elem(V3 a) { return fmaxr(a.x, fmaxr(a.y, a.z)); }
inline real luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }  // R/math/sg.rs:314-316

// ---- SG math (R/math/sg.rs) ------------------------------------------------------------
inline real expm1_over_x(real x) {  // :59-69
  real u = std::exp(x);
  if (u == 1.0f) return 1.0f;
  real y = u - 1.0f;
  return std::fabs(x) < 1.0f ? y / std::log(u) : y / x;
}
inline real erf_approx(real x) {  // :74-107
  if (std::fabs(x) >= 4.0f) return std::copysign(1.0f, x);
  if (std::fabs(x) > 1.0f) {
    const real a1 = 1.6284595f, a2 = 9.156747e-1f, a3 = 1.5432939e-1f, a4 = -3.5175983e-2f, a5 = 5.6679556e-3f,
                a6 = -5.648746e-4f, a7 = 2.5890768e-5f;
    real a = std::fabs(x);
    real p = (((((((a7 * a + a6) * a + a5) * a + a4) * a + a3) * a + a2) * a + a1) * a);
    real y = 1.0f - std::exp2(-p);
    return std::copysign(y, x);
  }
  const real a1 = 1.12837917f, a2 = -3.76123e-1f, a3 = 1.1279922e-1f, a4 = -2.6703065e-2f, a5 = 4.9073556e-3f,
              a6 = -5.5885315e-4f;
  real x2 = x * x;
  return (((((a6 * x2 + a5) * x2 + a4) * x2 + a3) * x2 + a2) * x2 + a1) * x;
}
inline real erfc_approx(real x) { return 1.0f - erf_approx(x); }
inline real sg_integral(real s) { return 4.0f * kPiF * expm1_over_x(-2.0f * s); }  // :124-126

struct SgLobe {
  V3 axis;
  real sharpness, log_amplitude;
};
inline SgLobe sg_product(V3 a1, real s1, V3 a2, real s2) {  // :137-156
  V3 axis = add(vscale(a1, s1), vscale(a2, s2));
  real sharp = vlen(axis);
  V3 d = sub(a1, a2);
  real len2 = dot(d, d);
  real denom = fmaxr(sharp + s1 + s2, kFltMin);
  real log_amp = -s1 * s2 * len2 / denom;
  V3 n = sharp > kFltMin ? vdiv(axis, sharp) : v3(0.f, 0.f, 1.f);
  return SgLobe{n, sharp, log_amp};
}
inline real upper_cc_over_2pi(real s) {  // :236-252
  if (s <= 0.5f) {
    return (((((((-1.0f / 362880.0f) * s + 1.0f / 40320.0f) * s - 1.0f / 5040.0f) * s + 1.0f / 720.0f) * s -
              1.0f / 120.0f) * s + 1.0f / 24.0f) * s - 1.0f / 6.0f) * s + 0.5f;
  }
  return (std::expm1(-s) + s) / (s * s);
}
inline real lower_cc_over_2pi(real s) {  // :256-275
  real e = std::exp(-s);
  if (s <= 0.5f) {
    return e * (((((((((1.0f / 403200.0f) * s - 1.0f / 45360.0f) * s + 1.0f / 5760.0f) * s - 1.0f / 840.0f) * s +
                    1.0f / 144.0f) * s - 1.0f / 30.0f) * s + 1.0f / 8.0f) * s - 1.0f / 3.0f) * s + 0.5f);
  }
  return e * (-std::expm1(-s) - s * e) / (s * s);
}
inline real sg_cc_product_over_pi(real z, real s) {  // :282-306
  const real A = 2.736083f, B = 17.021297f, C = 4.0100827f, D = 15.219156f, E = 76.0879f;
  real t = s * std::sqrt(0.5f * ((s + A) * s + B) / (((s + C) * s + D) * s + E));
  real tz = t * z;
  const real inv_sqrt_pi = 0.5641896f;
  const real clamp_min = 0.5f * kFltEps;
  real lerp = clampr(0.5f * (z * erfc_approx(-tz) + erfc_approx(t)) -
                          0.5f * inv_sqrt_pi * std::exp(-tz * tz) * std::expm1(t * t * (z * z - 1.0f)) / fmaxr(t, kFltMin),
                      clamp_min, 1.0f);
  real lower = lower_cc_over_2pi(s);
  real upper = upper_cc_over_2pi(s);
  return 2.0f * (lower + (upper - lower) * lerp);
}
inline real vmf_axis_length_to_sharpness(real len) {  // :196-208
  real r = clampr(len, 0.0f, 1.0f);
  if (r >= 1.0f) return kSgSharpnessMax;
  real om = 1.0f - r * r;
  if (om <= 0.0f) return kSgSharpnessMax;
  real lambda = r * (3.0f - r * r) / om;
  return clampr(lambda, 0.0f, kSgSharpnessMax);
}
struct Merged {
  V3 mu;
  real sigma_s2;
  V3 nu_bar;
  real flux;
};
inline Merged merge_cluster_params(real fl, V3 mul_, real sl, V3 nul, real fr, V3 mur, real sr, V3 nur) {  // :318-346
  real total = fl + fr;
  if (total <= 0.0f) return Merged{v3(0.f, 0.f, 0.f), 0.0f, v3(0.f, 0.f, 0.f), 0.0f};
  real wl = fl / total, wr = fr / total;
  V3 mu = add(vscale(mul_, wl), vscale(mur, wr));
  V3 md = sub(mul_, mur);
  real s2 = wl * sl + wr * sr + wl * wr * dot(md, md);
  V3 nu = add(vscale(nul, wl), vscale(nur, wr));
  return Merged{mu, s2, nu, total};
}

// ---- tree ----------------------------------------------------------------------------
struct Box {
  V3 mn{INFINITY, INFINITY, INFINITY}, mx{-INFINITY, -INFINITY, -INFINITY};
  real area() const {  // Bounds::surface_area (R/scene/mesh.rs:31-34)
    V3 e = vmax(sub(mx, mn), v3(0.f, 0.f, 0.f));
    return 2.0f * (e.x * e.y + e.x * e.z + e.y * e.z);
  }
  Box join(const Box& o) const { return Box{vmin(mn, o.mn), vmax(mx, o.mx)}; }
};

struct Node {
  real flux;
  V3 mu;
  real sigma_s2;
  V3 nu;
  real lambda;
  Box aabb;
  real radius;
  V3 cone_axis;
  real cone_theta_o, cone_theta_e;
  bool leaf;
  uint32_t inst, tri;  // triangle leaf
  real area;          // triangle area (leaf)
  uint32_t left, right, parent;
};

struct Leaf {
  uint32_t inst, tri;
  real flux;
  V3 mu;
  real sigma_s2;
  V3 nu_bar;
  Box aabb;
  real radius;
  V3 cone_axis;
  real theta_o, theta_e;
  V3 centroid;
  real area;
};

struct Cone {  // OrientationCone (R/light_tree/builder.rs:240-338)
  V3 axis{0.f, 0.f, 1.f};
  real theta_o = 0.0f, theta_e = 0.0f;
  bool valid = false;
  static Cone of(const Leaf& l) { return Cone{l.cone_axis, l.theta_o, l.theta_e, true}; }
  real measure() const {
    if (!valid) return 0.0f;
    real to = clampr(theta_o, 0.0f, kPiF), te = clampr(theta_e, 0.0f, kPiF * 0.5f);
    real tw = fminr(to + te, kPiF);
    real co = std::cos(to), so = std::sin(to);
    return 2.0f * kPiF * (1.0f - co) +
           0.5f * kPiF * (2.0f * tw * so - std::cos(to - 2.0f * tw) - 2.0f * to * so + co);
  }
  static Cone merge(Cone a, Cone b) {
    if (!a.valid) return b;
    if (!b.valid) return a;
    real te = fmaxr(a.theta_e, b.theta_e);
    if (!(a.theta_o >= b.theta_o)) std::swap(a, b);
    real cd = clampr(dot(a.axis, b.axis), -1.0f, 1.0f);
    real td = std::acos(cd);
    if (fminr(td + b.theta_o, kPiF) <= a.theta_o) return Cone{a.axis, a.theta_o, te, true};
    real to = fminr((a.theta_o + td + b.theta_o) * 0.5f, kPiF);
    if (to >= kPiF - 1e-5f) return Cone{a.axis, kPiF, te, true};
    V3 ra = cross(a.axis, b.axis);
    V3 na;
    if (dot(ra, ra) > 1.0e-12f) {
      real tr = to - a.theta_o;
      V3 k = vnormalize(ra);
      real cr = std::cos(tr), sr = std::sin(tr);
      na = vnormalize_or_zero(add(add(vscale(a.axis, cr), vscale(cross(k, a.axis), sr)),
                                  vscale(vscale(k, dot(k, a.axis)), 1.0f - cr)));
    } else {
      na = a.axis;
    }
    return Cone{dot(na, na) > 0.0f ? na : a.axis, to, te, true};
  }
};

struct Cluster {  // AggregatedCluster (R/light_tree/builder.rs:350-437)
  real flux = 0.0f;
  V3 mu{0.f, 0.f, 0.f};
  real sigma_s2 = 0.0f;
  V3 nu_bar{0.f, 0.f, 0.f};
  Box aabb;
  real radius = 0.0f;
  Cone cone;
  void add_leaf(const Leaf& l) {
    real nf = flux + l.flux;
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

inline real axis_of(V3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// triangle_leaf (R/light_tree/builder.rs:123-185); positions in world space.
inline bool triangle_leaf(uint32_t inst, uint32_t tri, V3 p0, V3 p1, V3 p2, real max_emission, Leaf* out) {
  V3 e1 = sub(p1, p0), e2 = sub(p2, p0);
  V3 gn = cross(e1, e2);
  real ad = vlen(gn);
  if (ad <= 0.0f) return false;
  V3 ng = vdiv(gn, ad);
  real area = 0.5f * ad;
  real phi = max_emission * area;
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

inline size_t bin_of(real v) {  // `as usize` saturates; NaN -> 0
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
  real parea = fmaxr(total.aabb.area(), 1.0e-30f);
  real pcone = fmaxr(total.cone.measure(), 1.0e-30f);
  real best = INFINITY;
  int best_axis = 0;
  size_t best_at = 0;
  for (int a = 0; a < 3; ++a) {
    if (axis_of(ext, a) <= 0.0f) continue;
    Cluster bins[12];
    uint32_t cnt[12] = {0};
    real inv = 12.0f / axis_of(ext, a);
    real lo = axis_of(cmin, a);
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
      real cost = l.flux * l.aabb.area() * l.cone.measure() + r.flux * r.aabb.area() * r.cone.measure();
      if (!std::isfinite(cost)) continue;
      real norm = cost / (parea * pcone);
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
  real root_flux() const { return nodes.empty() ? 0.0f : nodes[root].flux; }
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
inline bool sg_light_for_node(V3 p, V3 n, const Node& nd, real* w, SgLobe* lobe) {
  if (nd.flux <= 0.0f) return false;
  V3 d = sub(nd.mu, p);
  real r2 = dot(d, d);
  if (r2 <= 0.0f) return false;
  real inv_r = 1.0f / std::sqrt(r2);
  real cos_below = fmaxr(-(dot(n, d)) * inv_r, 0.0f);
  real bound_var = 0.5f * nd.radius * nd.radius;
  real sigma2 = nd.sigma_s2 * (1.0f - cos_below) + bound_var * cos_below;
  real min_sigma2 = r2 / kSgSharpnessMax;
  if (sigma2 < min_sigma2) sigma2 = min_sigma2;
  V3 axis = vscale(d, inv_r);
  real sharp = clampr(r2 / sigma2, 0.0f, kSgSharpnessMax);
  *lobe = sg_product(vneg(nd.nu), nd.lambda, axis, sharp);
  real sg_norm = fmaxr(sg_integral(nd.lambda), 1.0e-30f);
  real w_amp = nd.flux / (2.0f * kPiF * sigma2 * sg_norm);
  *w = w_amp * std::exp(lobe->log_amplitude);
  return std::isfinite(*w) && *w > 0.0f;
}

// diffuse_importance (R/light_tree/lobe.rs:276-287)
inline real diffuse_importance(real rho, V3 n, real w, const SgLobe& lobe) {
  if (rho <= 0.0f || w <= 0.0f) return 0.0f;
  real c = clampr(dot(lobe.axis, n), -1.0f, 1.0f);
  real integral = sg_cc_product_over_pi(c, lobe.sharpness) * kPiF;
  return rho * w * integral;
}

// ---- glossy lobe (R/light_tree/lobe.rs:112-163, 290-341, 396-436) -----------------
// sg_normalized_hemispherical_integral (R/math/sg.rs:155-171)
inline real sg_normalized_hemispherical_integral(real cosine, real sharpness) {
  const real A = 0.65173288f, B = 1.341828f, C = 7.221669f;
  real steepness = sharpness * std::sqrt((0.5f * sharpness + A) / ((sharpness + B) * sharpness + C));
  cosine = clampr(cosine, -1.0f, 1.0f);
  real denom = erf_approx(steepness);
  real numer = erf_approx(steepness * cosine);
  real v = std::fabs(denom) > 0.0f ? 0.5f + 0.5f * (numer / denom) : 0.5f;
  return clampr(v, 0.0f, 1.0f);
}
// vmf_hemispherical_integral (R/math/sg.rs:189-194)
inline real vmf_hemispherical_integral(real cosine, real sharpness) {
  real v = sg_normalized_hemispherical_integral(cosine, sharpness);
  real e = std::exp(-sharpness);
  real num = e + (1.0f - e) * v;
  return num / (e + 1.0f);
}
// glam Mat2: two columns; a * b takes column j of b through a
struct M2 {
  real c0x, c0y, c1x, c1y;
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
  real rho;
  V3 wi_ts;
  M2 jj;
  real det_jj4;
  real beta_x, beta_y;
  real alpha2_max;
  V3 xi_p;
  real kappa_p;
};
// make_glossy_lobe (R/light_tree/lobe.rs:112-163); false for None
inline bool make_glossy_lobe(real rho, const Frame3& frame, V3 wo_world, real alpha_x, real alpha_y,
                             GlossyLobe* g) {
  if (rho <= 0.0f) return false;
  V3 wi_ts = frame_to_local(frame, wo_world);
  if (wi_ts.z <= 0.0f) return false;
  real vlen = std::sqrt(wi_ts.x * wi_ts.x + wi_ts.y * wi_ts.y);
  real vx = vlen > 0.0f ? wi_ts.x / vlen : 1.0f, vy = vlen > 0.0f ? wi_ts.y / vlen : 0.0f;
  M2 r{vx, vy, -vy, vx};
  M2 d{0.5f, 0.0f, 0.0f, 0.5f / fmaxr(wi_ts.z, 1.0e-4f)};
  M2 jac = m2_mul(r, d);
  g->jj = m2_mul(jac, m2_transpose(jac));
  g->det_jj4 = 1.0f / (4.0f * wi_ts.z * wi_ts.z);
  real a2x = clampr(alpha_x * alpha_x, 0.0f, 1.0f - 1.0e-4f);
  real a2y = clampr(alpha_y * alpha_y, 0.0f, 1.0f - 1.0e-4f);
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
inline real sggx_reflection_pdf(V3 wi, V3 h, const M2& m) {
  real det = fmaxr(m.c0x * m.c1y - m.c0y * m.c1x, 1.0e-30f);
  M2 adj{m.c1y, -m.c0y, -m.c1x, m.c0x};
  // adj * m_xy
  real ax = adj.c0x * h.x + adj.c1x * h.y, ay = adj.c0y * h.x + adj.c1y * h.y;
  real length2 = (h.x * ax + h.y * ay) / det + h.z * h.z;
  real sggx = 1.0f / (kPiF * std::sqrt(det) * (length2 * length2));
  real inv_norm = std::sqrt(fmaxr(wi.x * wi.x * m.c0x + 2.0f * wi.x * wi.y * m.c0y + wi.y * wi.y * m.c1y + wi.z * wi.z,
                                   0.0f));
  return sggx / (4.0f * fmaxr(inv_norm, 1.0e-30f));
}
// glossy_importance (R/light_tree/lobe.rs:290-341)
inline real glossy_importance(const GlossyLobe& g, const Frame3& frame, V3 n, real w, const SgLobe& lobe) {
  if (g.rho <= 0.0f || w <= 0.0f) return 0.0f;
  real inv_kappa = 1.0f / fmaxr(lobe.sharpness, 1.0e-30f);
  M2 fp{g.beta_x + g.jj.c0x * inv_kappa, 0.0f + g.jj.c0y * inv_kappa, 0.0f + g.jj.c1x * inv_kappa,
        g.beta_y + g.jj.c1y * inv_kappa};
  real det = g.beta_x * g.beta_y + 2.0f * inv_kappa * (g.beta_x * g.jj.c1y + g.beta_y * g.jj.c0x) +
              inv_kappa * inv_kappa * g.det_jj4 / 4.0f;
  real tr = fp.c0x + fp.c1y;
  real denom = fmaxr(1.0f + tr + det, 1.0e-12f);
  M2 a_bar{(fp.c0x + det) / denom, fp.c0y / denom, fp.c1x / denom, (fp.c1y + det) / denom};
  V3 xi_ts = frame_to_local(frame, lobe.axis);
  V3 hu = add(g.wi_ts, xi_ts);
  if (dot(hu, hu) <= 1.0e-20f) return 0.0f;
  V3 h = vnormalize(hu);
  real lobe_value = sggx_reflection_pdf(g.wi_ts, h, a_bar);
  V3 prod = add(vscale(g.xi_p, g.kappa_p), vscale(lobe.axis, lobe.sharpness));
  real prod_sharp = vlen(prod);
  if (prod_sharp <= 0.0f) return 0.0f;
  V3 prod_dir = vdiv(prod, prod_sharp);
  real v_vis = vmf_hemispherical_integral(clampr(dot(prod_dir, n), -1.0f, 1.0f), prod_sharp);
  real sg_full = sg_integral(lobe.sharpness);
  return g.rho * w * v_vis * lobe_value * sg_full;
}
// merge_glossy_roughness (R/light_tree/lobe.rs:421-436)
inline void merge_glossy_roughness(real rho_a, real ax_a, real ay_a, real rho_b, real ax_b, real ay_b,
                                   real* ax, real* ay) {
  real total = fmaxr(rho_a + rho_b, 1.0e-30f);
  real wa = rho_a / total, wb = rho_b / total;
  real a2x = wa * ax_a * ax_a + wb * ax_b * ax_b;
  real a2y = wa * ay_a * ay_a + wb * ay_b * ay_b;
  *ax = std::sqrt(a2x);
  *ay = std::sqrt(a2y);
}

// A light-tree query (LightTreePrecompute, R/light_tree/lobe.rs:20-28): a
// diffuse lobe (rho > 0) and / or a glossy lobe (has_glossy) in `frame`.
struct Query {
  V3 p, n;
  real rho;
  bool has_glossy = false;
  Frame3 frame{};
  GlossyLobe glossy{};
};
// evaluate_node_importance (R/light_tree/importance.rs:41-48) with the
// material's light_tree_importance: the diffuse lobe alone
// (R/material/normalized_lambert.rs), or the lobes summed and clamped at
// zero (R/material/standard_surface.rs:700-717).
inline real node_importance(const Query& q, const Node& nd) {
  real w;
  SgLobe lobe;
  if (!sg_light_for_node(q.p, q.n, nd, &w, &lobe)) return 0.0f;
  real imp = diffuse_importance(q.rho, q.n, w, lobe);
  if (!q.has_glossy) return imp;
  return fmaxr(imp + glossy_importance(q.glossy, q.frame, q.n, w, lobe), 0.0f);
}

// sample_light_tree (R/light_tree/traversal.rs:26-92)
inline bool sample_tree(const Tree& t, const Query& q, real u, uint32_t* leaf_node, real* pmf_out) {
  if (t.nodes.empty()) return false;
  u = clampr(u, 0.0f, 1.0f - kFltEps);
  uint32_t ni = t.root;
  real pmf = 1.0f;
  for (;;) {
    const Node& nd = t.nodes[ni];
    if (nd.leaf) {
      *leaf_node = ni;
      *pmf_out = pmf;
      return true;
    }
    const Node& L = t.nodes[nd.left];
    const Node& R = t.nodes[nd.right];
    real il = node_importance(q, L), ir = node_importance(q, R);
    real total = il + ir;
    if (total <= 0.0f) {
      real lf = fmaxr(L.flux, 0.0f), rf = fmaxr(R.flux, 0.0f);
      real tf = lf + rf;
      if (tf <= 0.0f) return false;
      real pl = lf / tf;
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
    real pl = il / total;
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
inline real pdf_leaf(const Tree& t, const Query& q, uint32_t leaf_node) {
  if (t.nodes.empty() || leaf_node == kInvalid) return 0.0f;
  if (!t.nodes[leaf_node].leaf) return 0.0f;
  real pmf = 1.0f;
  uint32_t child = leaf_node, parent = t.nodes[leaf_node].parent;
  while (parent != kInvalid) {
    const Node& pn = t.nodes[parent];
    const Node& L = t.nodes[pn.left];
    const Node& R = t.nodes[pn.right];
    real il = node_importance(q, L), ir = node_importance(q, R);
    real total = il + ir;
    real pl;
    if (total <= 0.0f) {
      real lf = fmaxr(L.flux, 0.0f), rf = fmaxr(R.flux, 0.0f);
      real tf = lf + rf;
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

constexpr real kOneMinusEps = 1.0f - kFltEps;
constexpr real kTauF = 6.28318530717958647692f;

struct Level {
  size_t w, h;
  std::vector<real> wt;
};
struct Dist {  // HierarchicalDistribution: levels[0] is 1x1, last is the padded leaf
  std::vector<Level> levels;
  real total = 0.0f;
  real texel(size_t i, size_t j) const { return levels.back().wt[j * levels.back().w + i]; }
};

inline size_t next_pow2(size_t v) {
  size_t p = 1;
  while (p < v) p <<= 1;
  return p;
}
inline real row_sin_theta(size_t j, size_t h) { return std::sin(((real)j + 0.5f) / (real)h * kPiF); }
inline real positive_luminance(V3 c) { return fmaxr(lt::luminance(c), 0.0f); }

inline Dist build_distribution(size_t w, size_t h, const std::vector<real>& weights) {  // :426-478
  Dist d;
  const size_t pw = next_pow2(w), ph = next_pow2(h);
  Level leaf{pw, ph, std::vector<real>(pw * ph, 0.0f)};
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
        real sum = 0.0f;
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
  real scale = 1.0f;
  real rot[9], inv_rot[9];  // Mat3 columns
  std::vector<V3> px;
  Dist dist, dist_mis;
};

inline void rot_y(real angle, real* m) {  // Mat3::from_rotation_y, columns
  const real s = std::sin(angle), c = std::cos(angle);
  const real v[9] = {c, 0.0f, -s, 0.0f, 1.0f, 0.0f, s, 0.0f, c};
  for (int k = 0; k < 9; ++k) m[k] = v[k];
}
inline V3 mat3_mul(const real* m, V3 v) {  // Mat3 * Vec3: (x*c0 + y*c1) + z*c2
  return add(add(vscale(v3(m[0], m[1], m[2]), v.x), vscale(v3(m[3], m[4], m[5]), v.y)), vscale(v3(m[6], m[7], m[8]), v.z));
}

inline Env make(size_t w, size_t h, std::vector<V3> px, real scale, real rotate_y) {  // from_pixels (:47-83)
  Env e;
  e.w = w;
  e.h = h;
  e.scale = scale;
  e.px = std::move(px);
  std::vector<real> wt(w * h), wm(w * h);
  for (size_t j = 0; j < h; ++j) {  // environment_weights (:362-371)
    const real st = row_sin_theta(j, h);
    for (size_t i = 0; i < w; ++i) wt[j * w + i] = positive_luminance(e.px[j * w + i]) * st;
  }
  real ws = 0.0f, wsum = 0.0f;  // solid_angle_mean_luminance (:393-408)
  for (size_t j = 0; j < h; ++j) {
    const real st = row_sin_theta(j, h);
    wsum += st * (real)w;
    for (size_t i = 0; i < w; ++i) ws += positive_luminance(e.px[j * w + i]) * st;
  }
  const real mean = wsum > 0.0f ? ws / wsum : 0.0f;
  const real eps = fmaxr(std::fabs(mean), 1.0f) * 1.0e-6f;  // mis_compensated_environment_weights (:373-391)
  for (size_t j = 0; j < h; ++j) {
    const real st = row_sin_theta(j, h);
    for (size_t i = 0; i < w; ++i) {
      const real c = positive_luminance(e.px[j * w + i]) - mean;
      wm[j * w + i] = c > eps ? c * st : 0.0f;
    }
  }
  e.dist = build_distribution(w, h, wt);
  e.dist_mis = build_distribution(w, h, wm);
  rot_y(rotate_y, e.rot);
  rot_y(-rotate_y, e.inv_rot);
  return e;
}

inline real total_power(const Env& e) {  // :129-132
  const real texel = kTauF * kPiF / ((real)e.w * (real)e.h);
  return e.dist.total * e.scale * texel;
}
inline size_t pixel_coord(real t, size_t n) {  // :326-330
  const real c = clampr(t, 0.0f, 1.0f);
  const real f = c * (real)n;
  size_t idx = f >= 1.8446744e19f ? ~size_t(0) : (size_t)f;
  return std::min(idx, n - 1);
}
inline V3 uv_to_direction(real u, real v) {  // :332-337
  const real phi = u * kTauF, theta = v * kPiF;
  const real st = std::sin(theta);
  return v3(st * std::sin(phi), std::cos(theta), st * std::cos(phi));
}
inline void direction_to_uv(V3 d, real* u, real* v) {  // :339-348
  const V3 dir = vnormalize_or_zero(d);
  const real y = clampr(dir.y, -1.0f, 1.0f);
  const real theta = std::acos(y);
  real phi = std::atan2(dir.x, dir.z);
  if (phi < 0.0f) phi += kTauF;
  *u = clampr(phi / kTauF, 0.0f, 1.0f);
  *v = clampr(theta / kPiF, 0.0f, 1.0f);
}
inline V3 radiance(const Env& e, V3 dir) {  // :134-140
  real u, v;
  direction_to_uv(mat3_mul(e.inv_rot, dir), &u, &v);
  const V3 p = e.px[pixel_coord(v, e.h) * e.w + pixel_coord(u, e.w)];
  return vscale(p, e.scale);
}
inline real pdf(const Env& e, const Dist& d, V3 dir) {  // pdf_with_distribution (:156-177)
  if (d.total <= 0.0f) return 0.0f;
  real u, v;
  direction_to_uv(mat3_mul(e.inv_rot, dir), &u, &v);
  const real st = std::sin(v * kPiF);
  if (st <= 0.0f) return 0.0f;
  const real wgt = d.texel(pixel_coord(u, e.w), pixel_coord(v, e.h));
  if (wgt <= 0.0f) return 0.0f;
  const real p2d = wgt / d.total * (real)e.w * (real)e.h;
  return p2d / (kTauF * kPiF * st);
}
// hsw_step (:480-530)
inline bool hsw_step(const Level& parent, const Level& child, size_t* i, size_t* j, real* u, real* v) {
  const size_t si = child.w / parent.w, sj = child.h / parent.h;
  size_t ci = *i * si, cj = *j * sj;
  if (sj == 2) {
    real top = 0.0f, bot = 0.0f;
    for (size_t di = 0; di < si; ++di) {
      top += child.wt[cj * child.w + ci + di];
      bot += child.wt[(cj + 1) * child.w + ci + di];
    }
    const real total = top + bot;
    if (total <= 0.0f) return false;
    const real pt = top / total;
    if (*v < pt) {
      *v = clampr(*v / pt, 0.0f, kOneMinusEps);
    } else {
      *v = clampr((*v - pt) / (1.0f - pt), 0.0f, kOneMinusEps);
      cj += 1;
    }
  }
  if (si == 2) {
    const real l = child.wt[cj * child.w + ci], r = child.wt[cj * child.w + ci + 1];
    const real total = l + r;
    if (total <= 0.0f) return false;
    const real pl = l / total;
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
  real pdf;
};
// sample_with_distribution (:179-234)
inline bool sample(const Env& e, const Dist& d, real us0, real us1, Sample* out) {
  if (d.total <= 0.0f) return false;
  real u = clampr(us0, 0.0f, 1.0f), v = clampr(us1, 0.0f, 1.0f);
  size_t i = 0, j = 0;
  for (size_t l = 1; l < d.levels.size(); ++l)
    if (!hsw_step(d.levels[l - 1], d.levels[l], &i, &j, &u, &v)) return false;
  if (i >= e.w || j >= e.h) return false;
  const real wgt = d.texel(i, j);
  if (wgt <= 0.0f) return false;
  const real uc = ((real)i + u) / (real)e.w, vc = ((real)j + v) / (real)e.h;
  const V3 local = uv_to_direction(clampr(uc, 0.0f, kOneMinusEps), clampr(vc, 0.0f, kOneMinusEps));
  out->dir = mat3_mul(e.rot, local);
  const real st = std::sin(vc * kPiF);
  if (st <= 0.0f) return false;
  const real p2d = wgt / d.total * (real)e.w * (real)e.h;
  out->pdf = p2d / (kTauF * kPiF * st);
  out->radiance = vscale(e.px[j * e.w + i], e.scale);
  return true;
}

}  // namespace env

}  // namespace lt
}  // namespace orc
