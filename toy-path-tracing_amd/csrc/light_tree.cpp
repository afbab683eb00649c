// Host build of the SG light tree over a scene's emissive triangles:
// Scene::register_area_light_triangles (R/scene/mod.rs:649-683, instance
// order) + build_light_tree (R/light_tree/builder.rs:58-765): binned SAOH
// over 12 centroid bins per axis, cost flux x box area x orientation-cone
// measure, bottom-up (flux, mu, sigma_s^2, nu_bar) aggregation.  The float
// expressions follow the reference's evaluation order so the tree is the
// reference's bit for bit; the build itself is iterative (explicit task
// stack, depth-first pre-order emission) instead of recursive.
//
// Leaf ranges are reordered with a stable sort where the reference uses an
// unstable one: the two orders differ only between leaves whose centroids
// tie exactly on the split axis.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "light_tree.h"
#include "tpt_device.h"

namespace tpt {
namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct V {
  float x, y, z;
};
inline V mk(float x, float y, float z) { return V{x, y, z}; }
inline V operator+(V a, V b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V operator-(V a, V b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V operator*(V a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline V operator/(V a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }
inline float length(V a) { return std::sqrt(dot(a, a)); }
inline float fmx(float a, float b) { return std::fmax(a, b); }  // f32::max
inline float fmn(float a, float b) { return std::fmin(a, b); }
inline V vmin(V a, V b) { return mk(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z)); }
inline V vmax(V a, V b) { return mk(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z)); }
inline float clampr(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float comp(V v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline V normalized(V a) { return a * (1.0f / length(a)); }
inline V normalized_or_zero(V a) {
  const float r = 1.0f / length(a);
  return (std::isfinite(r) && r > 0.0f) ? a * r : mk(0.0f, 0.0f, 0.0f);
}

// glam Mat4::transform_point3 (SSE column order) of a column-major matrix
inline V transform_point(const float* m, V p) {
  float rx = m[0] * p.x, ry = m[1] * p.x, rz = m[2] * p.x;
  rx = m[4] * p.y + rx;
  ry = m[5] * p.y + ry;
  rz = m[6] * p.y + rz;
  rx = m[8] * p.z + rx;
  ry = m[9] * p.z + ry;
  rz = m[10] * p.z + rz;
  return mk(m[12] + rx, m[13] + ry, m[14] + rz);
}

// vMF axis length -> sharpness (R/math/sg.rs:196-208)
inline float sharpness_of(float len) {
  const float cap = 2199023255552.0f;
  const float r = clampr(len, 0.0f, 1.0f);
  if (r >= 1.0f) return cap;
  const float om = 1.0f - r * r;
  if (om <= 0.0f) return cap;
  return clampr(r * (3.0f - r * r) / om, 0.0f, cap);
}

struct Box {
  V lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
  float surface() const {  // Bounds::surface_area (R/scene/mesh.rs:31-34)
    const V e = vmax(hi - lo, mk(0.0f, 0.0f, 0.0f));
    return 2.0f * (e.x * e.y + e.x * e.z + e.y * e.z);
  }
};
inline Box join(const Box& a, const Box& b) { return Box{vmin(a.lo, b.lo), vmax(a.hi, b.hi)}; }

// Bounding cone of emission normals (R/light_tree/builder.rs:240-338).
struct Cone {
  V axis{0.0f, 0.0f, 1.0f};
  float to = 0.0f, te = 0.0f;
  bool on = false;
};
inline float cone_measure(const Cone& c) {
  if (!c.on) return 0.0f;
  const float to = clampr(c.to, 0.0f, kPi), te = clampr(c.te, 0.0f, kPi * 0.5f);
  const float tw = fmn(to + te, kPi);
  const float co = std::cos(to), so = std::sin(to);
  return 2.0f * kPi * (1.0f - co) + 0.5f * kPi * (2.0f * tw * so - std::cos(to - 2.0f * tw) - 2.0f * to * so + co);
}
inline Cone cone_union(Cone a, Cone b) {
  if (!a.on) return b;
  if (!b.on) return a;
  const float te = fmx(a.te, b.te);
  const Cone& wide = a.to >= b.to ? a : b;
  const Cone& narrow = a.to >= b.to ? b : a;
  const float td = std::acos(clampr(dot(wide.axis, narrow.axis), -1.0f, 1.0f));
  if (fmn(td + narrow.to, kPi) <= wide.to) return Cone{wide.axis, wide.to, te, true};
  const float to = fmn((wide.to + td + narrow.to) * 0.5f, kPi);
  if (to >= kPi - 1e-5f) return Cone{wide.axis, kPi, te, true};
  const V rot = cross(wide.axis, narrow.axis);
  V axis = wide.axis;
  if (dot(rot, rot) > 1.0e-12f) {
    const float tr = to - wide.to;
    const V k = normalized(rot);
    const float cr = std::cos(tr), sr = std::sin(tr);
    axis = normalized_or_zero(wide.axis * cr + cross(k, wide.axis) * sr + (k * dot(k, wide.axis)) * (1.0f - cr));
  }
  return Cone{dot(axis, axis) > 0.0f ? axis : wide.axis, to, te, true};
}

struct Leaf {
  uint32_t inst, tri;
  float flux, sigma_s2, radius, area, theta_o, theta_e;
  V mu, nu_bar, cone_axis, centroid;
  Box box;
};

// Flux-weighted cluster moments (AggregatedCluster, R/light_tree/builder.rs:350-437;
// merge_cluster_params, R/math/sg.rs:318-346).
struct Moments {
  float flux = 0.0f, sigma_s2 = 0.0f, radius = 0.0f;
  V mu{0.0f, 0.0f, 0.0f}, nu_bar{0.0f, 0.0f, 0.0f};
  Box box;
  Cone cone;

  void absorb(float f, V m, float s2, V nb, const Box& b, const Cone& c) {
    if (f <= 0.0f) return;  // both for a leaf (flux + f <= 0) and an empty cluster
    if (flux <= 0.0f) {
      flux = f;
      mu = m;
      sigma_s2 = s2;
      nu_bar = nb;
      box = b;
      cone = c;
      return;
    }
    const float total = flux + f;
    const float wl = flux / total, wr = f / total;
    const V d = mu - m;
    const float s = wl * sigma_s2 + wr * s2 + wl * wr * dot(d, d);
    mu = mu * wl + m * wr;
    sigma_s2 = s;
    nu_bar = nu_bar * wl + nb * wr;
    flux = total;
    box = join(box, b);
    cone = cone_union(cone, c);
  }
  void absorb(const Leaf& l) { absorb(l.flux, l.mu, l.sigma_s2, l.nu_bar, l.box, Cone{l.cone_axis, l.theta_o, l.theta_e, true}); }
  void absorb(const Moments& o) { absorb(o.flux, o.mu, o.sigma_s2, o.nu_bar, o.box, o.cone); }
  void set_radius() {
    if (flux <= 0.0f) {
      radius = 0.0f;
      return;
    }
    const V a = box.hi - mu, b = box.lo - mu;
    radius = length(vmax(mk(std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)),
                         mk(std::fabs(b.x), std::fabs(b.y), std::fabs(b.z))));
  }
};

inline size_t bin_index(float v) {
  if (!(v > 0.0f)) return 0;
  if (v >= 11.0f) return 11;
  return (size_t)v;
}

// Best binned SAOH split of leaves[b, e): sorts the range on the chosen axis
// and returns the split offset from b, or 0 when no split is found.
size_t saoh_split(std::vector<Leaf>& leaves, size_t b, size_t e, const Moments& total) {
  const size_t n = e - b;
  if (n < 2) return 0;
  V cl{INFINITY, INFINITY, INFINITY}, ch{-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = b; i < e; ++i) {
    cl = vmin(cl, leaves[i].centroid);
    ch = vmax(ch, leaves[i].centroid);
  }
  const V ext = ch - cl;
  if (fmx(ext.x, fmx(ext.y, ext.z)) <= 0.0f) return 0;
  const float norm = fmx(total.box.surface(), 1.0e-30f) * fmx(cone_measure(total.cone), 1.0e-30f);
  float best = INFINITY;
  int best_axis = 0;
  size_t best_count = 0;
  for (int a = 0; a < 3; ++a) {
    if (comp(ext, a) <= 0.0f) continue;
    Moments bin[12];
    uint32_t cnt[12] = {0};
    const float scale = 12.0f / comp(ext, a), lo = comp(cl, a);
    for (size_t i = b; i < e; ++i) {
      const size_t k = bin_index((comp(leaves[i].centroid, a) - lo) * scale);
      bin[k].absorb(leaves[i]);
      ++cnt[k];
    }
    Moments left[12], right[12];
    uint32_t lc[12], rc[12];
    Moments acc;
    uint32_t c = 0;
    for (int k = 0; k < 12; ++k) {
      acc.absorb(bin[k]);
      c += cnt[k];
      left[k] = acc;
      lc[k] = c;
    }
    acc = Moments();
    c = 0;
    for (int k = 11; k >= 0; --k) {
      acc.absorb(bin[k]);
      c += cnt[k];
      right[k] = acc;
      rc[k] = c;
    }
    for (int s = 1; s < 12; ++s) {
      if (lc[s - 1] == 0 || rc[s] == 0) continue;
      const Moments& L = left[s - 1];
      const Moments& R = right[s];
      const float cost = L.flux * L.box.surface() * cone_measure(L.cone) + R.flux * R.box.surface() * cone_measure(R.cone);
      if (!std::isfinite(cost)) continue;
      const float v = cost / norm;
      if (v < best) {
        best = v;
        best_axis = a;
        best_count = lc[s - 1];
      }
    }
  }
  if (!std::isfinite(best)) return 0;
  std::stable_sort(leaves.begin() + b, leaves.begin() + e, [best_axis](const Leaf& p, const Leaf& q) {
    return comp(p.centroid, best_axis) < comp(q.centroid, best_axis);
  });
  return std::min(std::max<size_t>(best_count, 1), n - 1);
}

TptLightNode node_of(float flux, V mu, float s2, V nu_bar, float radius) {
  TptLightNode d;
  std::memset(&d, 0, sizeof(d));
  d.flux = flux;
  d.mu[0] = mu.x;
  d.mu[1] = mu.y;
  d.mu[2] = mu.z;
  d.sigma_s2 = s2;
  const V nu = dot(nu_bar, nu_bar) > 0.0f ? normalized(nu_bar) : mk(0.0f, 0.0f, 1.0f);
  d.nu[0] = nu.x;
  d.nu[1] = nu.y;
  d.nu[2] = nu.z;
  d.lambda = sharpness_of(length(nu_bar));
  d.radius = radius;
  d.left = d.right = d.parent = d.instance = d.triangle = kNone;
  return d;
}

}  // namespace

float sg_norm_host(float lambda) {
  const float x = -2.0f * lambda;
  const float u = std::exp(x);
  float e = 1.0f;
  if (u != 1.0f) {
    const float y = u - 1.0f;
    e = std::fabs(x) < 1.0f ? y / std::log(u) : y / x;
  }
  return std::fmax(4.0f * kPi * e, 1.0e-30f);
}

bool build_light_tree(const TptSceneDesc& desc, std::vector<TptLightNode>* out, std::vector<uint32_t>* light_base,
                      std::vector<uint32_t>* tri_leaf, std::string* err) {
  out->clear();
  tri_leaf->clear();
  light_base->assign(desc.instance_count, kNone);
  std::vector<Leaf> leaves;
  for (uint32_t i = 0; i < desc.instance_count; ++i) {
    const TptInstanceDesc& in = desc.instances[i];
    if (!desc.materials || in.material >= desc.material_count) continue;
    const TptMaterial& m = desc.materials[in.material];
    float emax = 0.0f;  // Material::may_emit / max_emission (R/material/mod.rs:372-412)
    if (m.kind == TPT_MATERIAL_EMISSIVE) {
      const V c = mk(m.u.emissive.color[0], m.u.emissive.color[1], m.u.emissive.color[2]) * m.u.emissive.strength;
      emax = fmx(fmx(c.x, fmx(c.y, c.z)) * 1.0f, 0.0f);
    } else if (m.kind == TPT_MATERIAL_OPEN_PBR && m.u.open_pbr.emission_luminance > 0.0f) {
      const float* ec = m.u.open_pbr.emission_color;
      // texture factor: the emission texture's level-0 maximum (open_pbr.rs:381-391)
      float tf = 1.0f;
      const uint32_t et = m.u.open_pbr.emission_color_texture;
      if (et != TPT_INVALID_INDEX) {
        if (et >= desc.texture_count || !desc.textures) {
          *err = "light tree: emission_color_texture index out of range";
          return false;
        }
        tf = texture_max_value(desc.textures[et]);
      }
      emax = fmx(fmx(ec[0], fmx(ec[1], ec[2])) * m.u.open_pbr.emission_luminance * tf, 0.0f);
    } else {
      continue;
    }
    if (emax <= 0.0f) continue;
    const TptMeshDesc& md = desc.meshes[in.mesh];
    (*light_base)[i] = (uint32_t)tri_leaf->size();
    tri_leaf->resize(tri_leaf->size() + md.triangle_count, kNone);
    for (uint32_t t = 0; t < md.triangle_count; ++t) {
      V p[3];
      for (int k = 0; k < 3; ++k) {
        const uint32_t vi = md.indices[3 * (size_t)t + k];
        if (vi >= md.vertex_count) {
          *err = "light tree: vertex index out of range";
          return false;
        }
        p[k] = transform_point(in.local_to_world,
                               mk(md.positions[3 * (size_t)vi], md.positions[3 * (size_t)vi + 1], md.positions[3 * (size_t)vi + 2]));
      }
      const V e1 = p[1] - p[0], e2 = p[2] - p[0];
      const V gn = cross(e1, e2);
      const float twice = length(gn);
      const float area = 0.5f * twice;
      if (!(area * emax > 0.0f) || twice <= 0.0f) continue;  // area-light weight, then triangle_leaf
      Leaf l;
      l.inst = i;
      l.tri = t;
      l.flux = emax * area;
      if (l.flux <= 0.0f) continue;
      l.area = area;
      l.mu = (p[0] + p[1] + p[2]) / 3.0f;
      l.sigma_s2 = (dot(e1, e1) + dot(e2, e2) - dot(e1, e2)) / 18.0f;
      l.radius = fmx(fmx(length(p[0] - l.mu), length(p[1] - l.mu)), length(p[2] - l.mu));
      l.box = Box{vmin(vmin(p[0], p[1]), p[2]), vmax(vmax(p[0], p[1]), p[2])};
      const V ng = gn / twice;
      l.nu_bar = ng * 0.5f;
      l.cone_axis = ng;
      l.theta_o = 0.0f;
      l.theta_e = kPi * 0.5f;
      l.centroid = l.mu;
      leaves.push_back(l);
    }
  }
  if (leaves.empty()) return true;
  out->reserve(2 * leaves.size() - 1);
  struct Task {
    size_t b, e;
    uint32_t parent;
    bool right;
  };
  std::vector<Task> todo{{0, leaves.size(), kNone, false}};
  while (!todo.empty()) {
    const Task tk = todo.back();
    todo.pop_back();
    const uint32_t idx = (uint32_t)out->size();
    if (tk.parent != kNone) {
      TptLightNode& p = (*out)[tk.parent];
      (tk.right ? p.right : p.left) = idx;
    }
    if (tk.e - tk.b == 1) {
      const Leaf& l = leaves[tk.b];
      TptLightNode d = node_of(l.flux, l.mu, l.sigma_s2, l.nu_bar, l.radius);
      d.parent = tk.parent;
      d.instance = l.inst;
      d.triangle = l.tri;
      d.area = l.area;
      d.leaf = 1;
      out->push_back(d);
      continue;
    }
    Moments total;
    for (size_t i = tk.b; i < tk.e; ++i) total.absorb(leaves[i]);
    total.set_radius();
    size_t split = saoh_split(leaves, tk.b, tk.e, total);
    if (split == 0) {  // median of the widest centroid extent
      V cl{INFINITY, INFINITY, INFINITY}, ch{-INFINITY, -INFINITY, -INFINITY};
      for (size_t i = tk.b; i < tk.e; ++i) {
        cl = vmin(cl, leaves[i].centroid);
        ch = vmax(ch, leaves[i].centroid);
      }
      const V d = ch - cl;
      const int axis = (d.x >= d.y && d.x >= d.z) ? 0 : (d.y >= d.z ? 1 : 2);
      std::stable_sort(leaves.begin() + tk.b, leaves.begin() + tk.e, [axis](const Leaf& p, const Leaf& q) {
        return comp(p.centroid, axis) < comp(q.centroid, axis);
      });
      split = (tk.e - tk.b) / 2;
    }
    split = std::min(std::max<size_t>(split, 1), tk.e - tk.b - 1);
    TptLightNode d = node_of(total.flux, total.mu, total.sigma_s2, total.nu_bar, total.radius);
    d.parent = tk.parent;
    out->push_back(d);
    todo.push_back(Task{tk.b + split, tk.e, idx, true});
    todo.push_back(Task{tk.b, tk.b + split, idx, false});
  }
  for (uint32_t k = 0; k < (uint32_t)out->size(); ++k) {
    const TptLightNode& d = (*out)[k];
    if (d.leaf) (*tri_leaf)[(*light_base)[d.instance] + d.triangle] = k;
  }
  return true;
}

}  // namespace tpt

extern "C" int32_t tpt_light_tree_build(const TptSceneDesc* desc, TptLightNode* out, uint32_t capacity,
                                        uint32_t* count) {
  try {
    if (!desc || !count || (capacity && !out) || (desc->instance_count && (!desc->instances || !desc->meshes))) {
      tpt::set_error("tpt_light_tree_build: null argument");
      return TPT_ERROR_BAD_ARGUMENT;
    }
    for (uint32_t i = 0; i < desc->instance_count; ++i)
      if (desc->instances[i].mesh >= desc->mesh_count) {
        tpt::set_error("tpt_light_tree_build: instance mesh index out of range");
        return TPT_ERROR_BAD_ARGUMENT;
      }
    std::vector<TptLightNode> nodes;
    std::vector<uint32_t> base, leaf;
    std::string err;
    if (!tpt::build_light_tree(*desc, &nodes, &base, &leaf, &err)) {
      tpt::set_error("tpt_light_tree_build: " + err);
      return TPT_ERROR_BAD_ARGUMENT;
    }
    *count = (uint32_t)nodes.size();
    const uint32_t n = std::min(capacity, *count);
    if (n) std::memcpy(out, nodes.data(), n * sizeof(TptLightNode));
    return TPT_OK;
  } catch (const std::exception& e) {
    tpt::set_error(std::string("tpt_light_tree_build: ") + e.what());
    return TPT_ERROR;
  }
}
