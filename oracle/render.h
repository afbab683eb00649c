// Oracle: the unidirectional MIS path tracer of the reference for OpenPBR
// materials, restated line by line.  TEST INFRASTRUCTURE ONLY.
//
//   R/main.rs:98-126                     pixel loop, running mean, exposure
//   R/integrator/mis.rs:19-139, 212-257  trace_radiance, emission MIS
//   R/integrator/mod.rs:52-85            spawn_ray / offset_ray_origin
//   R/scene/mod.rs:426-544               shading vertex
//   R/material/open_pbr.rs:351-606       OpenPbrMaterial le (with exit tint) / sample / eval / pdf
//   R/material/mod.rs:33-79              shading-normal correction
//   bsdf.h                               the BSDFs and GGX tables
//
// OpenPBR materials carry base, metal, specular, coat, fuzz and emission
// (bsdf.h lists what is not carried).
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

#include "bsdf.h"
#include "geometry.h"
#include "lights.h"
#include "sampler.h"
#include "texture.h"

namespace orc {

// ---- scene shading data ------------------------------------------------------------
// Material enum of the backend (R/material/mod.rs:100-115 variants carried).
enum MatKind : uint32_t { kMatOpenPbr = 0, kMatLambert = 1, kMatEmissive = 2, kMatMirror = 3 };
// next-event-estimation policy: the reference's (OpenPBR builds no light-tree
// query) or the OpenPBR proxy query (openpbr_proxy_query)
enum NeePolicy : uint32_t { kNeeReference = 0, kNeeOpenPbrProxy = 1 };
struct Material {
  uint32_t kind = kMatOpenPbr;
  OpenPbr pbr{};
  Luts luts;  // OpenPBR: the GGX tables it reads
  V3 color{0.f, 0.f, 0.f};  // Lambert rho, Emissive colour, Mirror colour
  real strength = 0.0f;    // Emissive
};

struct ShadeMesh {
  std::vector<real> nrm;  // 3 per vertex (may be empty: geometric normal)
  std::vector<real> uv;   // 2 per vertex (may be empty)
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
  std::vector<texture::Texture> textures;  // indexed by the OpenPBR texture slots
  // directional lights (R/light/directional.rs:15-32): direction of travel,
  // tint, perpendicular irradiance
  struct DirLight {
    V3 direction, color;
    real intensity;
  };
  std::vector<DirLight> dirs;
  uint32_t nee_policy = 0;  // kNeeReference or kNeeOpenPbrProxy
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
  // uv and its screen-space derivatives, the position and normal partials
  // (ShadingVertex, R/material/mod.rs:120-170)
  real u = 0.0f, v = 0.0f, dudx = 0.0f, dvdx = 0.0f, dudy = 0.0f, dvdy = 0.0f;
  V3 dpdx{0.f, 0.f, 0.f}, dpdy{0.f, 0.f, 0.f}, dndu{0.f, 0.f, 0.f}, dndv{0.f, 0.f, 0.f};
};

// Ray with its optional differential and its cone (R/math/ray.rs:8-80).
struct RayCone {
  real width = 0.0f, spread = 0.0f;
};
inline RayCone make_cone(real width, real spread) {  // RayCone::new
  return RayCone{std::fmax(width, 0.0f), std::fmax(spread, 0.0f)};
}
inline real cone_width_at(const RayCone& c, real t) {  // RayCone::width_at
  if (std::isfinite(t)) return std::fmax(c.width + c.spread * std::fmax(t, 0.0f), 0.0f);
  return c.width;
}
struct Ray {
  V3 o, d;
  bool has_diff = false;
  V3 rxo{0.f, 0.f, 0.f}, ryo{0.f, 0.f, 0.f}, rxd{0.f, 0.f, 0.f}, ryd{0.f, 0.f, 0.f};
  RayCone cone;
};

// compute_surface_partials (R/math/mod.rs:112-134) of three values over uvs.
inline bool surface_partials(V3 a0, V3 a1, V3 a2, const real* uv, V3* dadu, V3* dadv) {
  V3 d1 = sub(a1, a0), d2 = sub(a2, a0);
  real du1 = uv[2] - uv[0], dv1 = uv[3] - uv[1], du2 = uv[4] - uv[0], dv2 = uv[5] - uv[1];
  real det = du1 * dv2 - dv1 * du2;
  if (std::fabs(det) <= 1.0e-8f) return false;
  real inv = 1.0f / det;
  V3 a = smul(sub(mul(dv2, d1), mul(dv1, d2)), inv);
  V3 b = smul(add(mul(-du2, d1), mul(du1, d2)), inv);
  if (len_sq(a) == 0.0f || len_sq(b) == 0.0f) return false;
  *dadu = a;
  *dadv = b;
  return true;
}

inline real finite_clamped_derivative(real x) {  // R/scene/mod.rs:861-867
  return std::isfinite(x) ? clampf(x, -1.0e8f, 1.0e8f) : 0.0f;
}
inline bool v3_finite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

// differentials_from_dp (R/scene/mod.rs:787-819)
inline bool differentials_from_dp(Vertex& v, V3 dpdu, V3 dpdv, V3 dpdx, V3 dpdy) {
  real a00 = dot(dpdu, dpdu), a01 = dot(dpdu, dpdv), a11 = dot(dpdv, dpdv);
  real det = difference_of_products(a00, a11, a01, a01);
  if (std::fabs(det) <= 1.0e-12f || !std::isfinite(det)) return false;
  real inv = 1.0f / det;
  real b0x = dot(dpdu, dpdx), b1x = dot(dpdv, dpdx), b0y = dot(dpdu, dpdy), b1y = dot(dpdv, dpdy);
  v.dpdx = dpdx;
  v.dpdy = dpdy;
  v.dudx = finite_clamped_derivative(difference_of_products(a11, b0x, a01, b1x) * inv);
  v.dvdx = finite_clamped_derivative(difference_of_products(a00, b1x, a01, b0x) * inv);
  v.dudy = finite_clamped_derivative(difference_of_products(a11, b0y, a01, b1y) * inv);
  v.dvdy = finite_clamped_derivative(difference_of_products(a00, b1y, a01, b0y) * inv);
  return true;
}
// compute_shading_differentials (R/scene/mod.rs:748-785) with
// cone_shading_differentials (:821-849) and ray_parameter_to_point (:851-859);
// n is the face-forwarded geometric normal.  No ray: zero differentials.
inline void shading_differentials(Vertex& v, V3 n, V3 dpdu, V3 dpdv, const Ray* ray) {
  if (!ray) return;
  if (ray->has_diff) {
    real plane_d = -dot(n, v.p);
    real rxden = dot(n, ray->rxd), ryden = dot(n, ray->ryd);
    if (std::fabs(rxden) > 1.0e-8f && std::fabs(ryden) > 1.0e-8f) {
      real tx = (-dot(n, ray->rxo) - plane_d) / rxden;
      real ty = (-dot(n, ray->ryo) - plane_d) / ryden;
      V3 px = add(ray->rxo, mul(tx, ray->rxd)), py = add(ray->ryo, mul(ty, ray->ryd));
      V3 dpdx = sub(px, v.p), dpdy = sub(py, v.p);
      if (std::isfinite(tx) && std::isfinite(ty) && v3_finite(dpdx) && v3_finite(dpdy) &&
          differentials_from_dp(v, dpdu, dpdv, dpdx, dpdy))
        return;
    }
  }
  real dl2 = len_sq(ray->d);
  if (dl2 <= 0.0f) return;
  real t = dot(sub(v.p, ray->o), ray->d) / dl2;
  if (!std::isfinite(t)) return;
  real width = cone_width_at(ray->cone, t);
  if (width <= 0.0f) return;
  real cos_theta = std::fmax(std::fabs(dot(normalize_or_zero3(ray->d), n)), 1.0e-3f);
  real projected = width / cos_theta;
  real dudx = projected / std::fmax(std::sqrt(len_sq(dpdu)), 1.0e-6f);
  real dvdy = projected / std::fmax(std::sqrt(len_sq(dpdv)), 1.0e-6f);
  v.dpdx = mul(dudx, dpdu);
  v.dpdy = mul(dvdy, dpdv);
  v.dudx = dudx;
  v.dvdx = 0.0f;
  v.dudy = 0.0f;
  v.dvdy = dvdy;
}

inline V3 mat3_mul_normal(const real* w2l, V3 n) {
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
// the fields the integrator reads; `dir` is the incident direction, `ray`
// the ray that hit (null for a light sample's vertex: zero differentials).
inline Vertex vertex_at(const RenderScene& rs, uint32_t inst, uint32_t tri, real b0, real b1, real b2, V3 dir,
                        const Ray* ray = nullptr) {
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
  const real uvs[6] = {uvv(i0, 0), uvv(i0, 1), uvv(i1, 0), uvv(i1, 1), uvv(i2, 0), uvv(i2, 1)};
  // interpolate_vec2: (b0 uv0 + b1 uv1) + b2 uv2
  v.u = b0 * uvs[0] + b1 * uvs[2] + b2 * uvs[4];
  v.v = b0 * uvs[1] + b1 * uvs[3] + b2 * uvs[5];
  // normal partials, flipped with the shading normal
  if (!surface_partials(n0, n1, n2, uvs, &v.dndu, &v.dndv)) v.dndu = v.dndv = v3(0.f, 0.f, 0.f);
  if (dot(v.ns, raw_ns) < 0.0f) {
    v.dndu = neg(v.dndu);
    v.dndv = neg(v.dndv);
  }
  // compute_surface_partials (R/math/mod.rs:112-134) over world positions,
  // else the Duff frame's tangent and bitangent
  V3 dpdu, dpdv;
  if (!surface_partials(p0, p1, p2, uvs, &dpdu, &dpdv)) {
    Onb f = onb_from_normal(v.ns);
    dpdu = f.t;
    dpdv = f.b;
  }
  v.frame = onb_from_normal_and_tangent(v.ns, dpdu);
  if (len_sq(dpdu) == 0.0f) dpdu = v.frame.t;
  if (len_sq(dpdv) == 0.0f) dpdv = v.frame.b;
  shading_differentials(v, v.ng, dpdu, dpdv, ray);
  v.wo = normalize_or_zero3(nd);
  v.material = rs.inst_material[inst];
  v.inst = inst;
  v.tri = tri;
  return v;
}
inline Vertex shading_vertex(const RenderScene& rs, const SceneHit& h, const Ray& ray) {
  return vertex_at(rs, h.inst, h.tri, h.b0, h.b1, h.b2, ray.d, &ray);
}

// OpenPbrMaterial::*_at (R/material/open_pbr.rs:670-790, le :351-375): a
// textured parameter is the parameter times Texture::sample_filtered at the
// vertex's uv with derivatives uv_dx() = (dudx, dvdx), uv_dy() = (dudy, dvdy);
// the clamps stay where the parameters are read.
inline Material material_at(const RenderScene& rs, const Material& m, const Vertex& v) {
  if (m.kind != kMatOpenPbr) return m;
  Material r = m;
  const OpenPbr& p = m.pbr;
  auto fetch = [&](int slot) {
    return texture::sample_filtered(rs.textures[p.tex[slot]], v.u, v.v, v.dudx, v.dvdx, v.dudy, v.dvdy);
  };
  if (p.tex[kTexBaseColor] != kNoTexture) {
    texture::Px t = fetch(kTexBaseColor);
    r.pbr.base_color = vmul(p.base_color, v3(t.c[0], t.c[1], t.c[2]));
  }
  if (p.tex[kTexMetalness] != kNoTexture) r.pbr.base_metalness = p.base_metalness * fetch(kTexMetalness).c[0];
  if (p.tex[kTexSpecRough] != kNoTexture) r.pbr.specular_roughness = p.specular_roughness * fetch(kTexSpecRough).c[0];
  if (p.tex[kTexFuzzWeight] != kNoTexture) r.pbr.fuzz_weight = p.fuzz_weight * fetch(kTexFuzzWeight).c[0];
  if (p.tex[kTexFuzzRough] != kNoTexture) r.pbr.fuzz_roughness = p.fuzz_roughness * fetch(kTexFuzzRough).c[0];
  if (p.tex[kTexEmission] != kNoTexture) {
    texture::Px t = fetch(kTexEmission);
    r.pbr.emission_color = vmul(p.emission_color, v3(t.c[0], t.c[1], t.c[2]));
  }
  return r;
}

// OpenPbrMaterial::le (R/material/open_pbr.rs:351-375), no emission texture.
inline bool pbr_le(const OpenPbr& m, const Vertex& v, V3* le) {
  if (m.emission_luminance <= 0.0f) return false;
  if (!m.thin_walled && !v.front_face) return false;
  V3 wl = normalize_or_zero3(onb_to_local(v.frame, v.wo));
  *le = vmul(smul(vmul(m.emission_color, vsplat(1.0f)), std::fmax(m.emission_luminance, 0.0f)),
             emission_exit_tint(m, wl));
  return true;
}

// OpenPbrMaterial::sample (R/material/open_pbr.rs:489-548) in radiance mode.
constexpr real kGeomCosEps = 1.0e-6f;  // R/material/mod.rs:31
inline bool pbr_sample(const OpenPbr& m, const Luts& luts, const Vertex& v, real u_lobe, real ux, real uy, BSample* out,
                       V3* wi_world) {
  if (dot(v.wo, v.ng) <= 0.0f) return false;
  V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
  if (wol.z <= 0.0f) return false;
  PbrBsdf p = make_bsdf(m, v.front_face, wol, luts);
  BSample s;
  if (!bsdf_sample(p, wol, u_lobe, ux, uy, &s)) return false;
  V3 wi = onb_to_world(v.frame, s.wi);
  if (s.transmission) {  // a rough refraction must leave through the geometric surface
    if (dot(wi, v.ng) >= -kGeomCosEps && !s.delta) return false;
  } else if (dot(wi, v.ng) <= kGeomCosEps) {
    return false;
  }
  // cone_spread: 2 * effective specular roughness for GLOSSY samples, else
  // GLOSSY_DIFFUSE_CONE_SPREAD 0.5 (:509-522)
  s.cone_spread = (s.glossy && !s.delta) ? 2.0f * clampf(p.spec_roughness, 0.0f, 1.0f) : 0.5f;
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
inline real mat_max_emission(const RenderScene& rs, const Material& m) {
  if (m.kind == kMatEmissive)  // emissive.rs max_emission, no texture
    return std::fmax(lt::vmax_elem(smul(m.color, m.strength)) * 1.0f, 0.0f);
  if (m.kind == kMatOpenPbr) {  // open_pbr.rs:381-391: texture factor = the emission texture's max_value
    if (m.pbr.emission_luminance <= 0.0f) return 0.0f;
    const uint32_t et = m.pbr.tex[kTexEmission];
    const real tf = et != kNoTexture ? rs.textures[et].max_value : 1.0f;
    return std::fmax(lt::vmax_elem(m.pbr.emission_color) * m.pbr.emission_luminance * tf, 0.0f);
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
inline bool mat_sample(const Material& m, const Vertex& v, real u_lobe, real ux, real uy, BSample* out,
                       V3* wi_world) {
  if (m.kind == kMatOpenPbr) return pbr_sample(m.pbr, m.luts, v, u_lobe, ux, uy, out, wi_world);
  if (m.kind == kMatEmissive) return false;
  V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
  if (wol.z <= 0.0f) return false;
  BSample s;
  if (m.kind == kMatLambert) {  // NormalizedLambertBsdf::sample (bsdf/normalized_lambert.rs:35-60)
    V3 wi = cosine_hemisphere(ux, uy);
    real pdf = std::fmax(wi.z, 0.0f) / kPi;
    if (pdf <= 0.0f) return false;
    s.weight = m.color;
    s.wi = wi;
    s.pdf = pdf;
    s.delta = false;
    s.cone_spread = 0.5f;  // DIFFUSE_CONE_SPREAD, normalized_lambert.rs:16
  } else {  // MirrorBsdf::sample (bsdf/mirror.rs:23-39)
    s.weight = m.color;
    s.wi = normalize_or_zero3(v3(-wol.x, -wol.y, wol.z));
    s.pdf = 1.0f;
    s.delta = true;
    s.cone_spread = 0.0f;  // mirror.rs:134
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
  real cg = std::fabs(dot(wi, v.ng));
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
    return shading_normal_correction(v, wi, bsdf_eval(make_bsdf(m.pbr, v.front_face, wol, m.luts), wol, wil));
  }
  return zero;
}
inline real mat_pdf(const Material& m, const Vertex& v, V3 wi) {
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
    return bsdf_pdf(make_bsdf(m.pbr, v.front_face, wol, m.luts), wol, wil);
  }
  return 0.0f;
}
// The OpenPBR light-tree query of the proxy policy (an extension of this
// backend; the reference's OpenPBR builds none, R/material/open_pbr.rs:
// 592-606): StandardSurface::light_tree_precompute
// (R/material/standard_surface.rs:599-698) with OpenPBR's parameters.
//   under    = lerp(1, coat_color, coat_weight / 2)
//   diffuse  = (1 - metal) (1 - transmission_weight) base_weight lum(base_color under)
//   metal    = metal base_weight lum(base_color under)
//   specular = (1 - metal) specular_weight lum(specular_color under) F0(specular_ior)
//   coat     = coat_weight F0(coat_ior)
// metal and specular at the specular roughness, coat at the coat roughness
// (alpha_xy), merged in that order by merge_glossy_roughness into one glossy
// lobe; its frame is (t, normalize(n x t), n) with t the vertex frame's
// tangent and n the normalised shading normal.
inline real dielectric_f0(real ior) {
  real q = (ior - 1.0f) / (ior + 1.0f);
  return q * q;
}
inline bool openpbr_proxy_query(const OpenPbr& m, const Vertex& v, lt::Query* q) {
  real metal = clampf(m.base_metalness, 0.0f, 1.0f), base = clampf(m.base_weight, 0.0f, 1.0f);
  V3 under = vlerp(vsplat(1.0f), vclamp01(m.coat_color), clampf(m.coat_weight, 0.0f, 1.0f) * 0.5f);
  real lb = luminance(vmul(m.base_color, under));
  real rho_d = std::fmax((1.0f - metal) * (1.0f - clampf(m.transmission_weight, 0.0f, 1.0f)) * base * lb, 0.0f);
  real rho_metal = metal * base * lb;
  real rho_spec = (1.0f - metal) * std::fmax(m.specular_weight, 0.0f) * luminance(vmul(vclamp01(m.specular_color), under)) *
                   dielectric_f0(std::fmax(m.specular_ior, 1.0e-4f));
  real rho_coat = clampf(m.coat_weight, 0.0f, 1.0f) * dielectric_f0(std::fmax(m.coat_ior, 1.0e-4f));
  real sx, sy, cx, cy;
  alpha_xy(clampf(m.specular_roughness, 0.0f, 1.0f), m.specular_roughness_anisotropy, &sx, &sy);
  alpha_xy(clampf(m.coat_roughness, 0.0f, 1.0f), m.coat_roughness_anisotropy, &cx, &cy);
  real rho = 0.0f, ax = 1.0f, ay = 1.0f;
  const real rl[3] = {rho_metal, rho_spec, rho_coat}, xl[3] = {sx, sx, cx}, yl[3] = {sy, sy, cy};
  for (int i = 0; i < 3; ++i) {
    if (!(rl[i] > 0.0f)) continue;
    if (rho > 0.0f) {
      lt::merge_glossy_roughness(rho, ax, ay, rl[i], xl[i], yl[i], &ax, &ay);
    } else {
      ax = xl[i];
      ay = yl[i];
    }
    rho += rl[i];
  }
  *q = lt::Query{v.p, v.ns, rho_d};
  if (rho > 0.0f) {
    V3 fn = len_sq(v.ns) > 0.0f ? normalize3(v.ns) : v3(0.f, 0.f, 1.f);
    lt::Frame3 f{v.frame.t, normalize_or_zero3(cross(fn, v.frame.t)), fn};
    q->has_glossy = lt::make_glossy_lobe(rho, f, v.wo, ax, ay, &q->glossy);
    q->frame = f;
  }
  return q->rho > 0.0f || q->has_glossy;
}
// light_tree_precompute: NormalizedLambert (normalized_lambert.rs:201-218)
// gives a diffuse query; OpenPBR gives the proxy query under the proxy
// policy; every other material here gives none.
inline bool mat_query(const RenderScene& rs, const Material& m, const Vertex& v, lt::Query* q) {
  if (m.kind == kMatOpenPbr && rs.nee_policy == kNeeOpenPbrProxy) return openpbr_proxy_query(m.pbr, v, q);
  if (m.kind != kMatLambert) return false;
  real rho = lt::luminance(m.color);
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
    real me = mat_max_emission(rs, m);
    if (me <= 0.0f) continue;
    const Mesh& mesh = g.meshes[g.inst[i].mesh];
    const uint32_t ntri = (uint32_t)(mesh.idx.size() / 3);
    rs.leaf_base[i] = (uint32_t)rs.leaf_of.size();
    rs.leaf_of.resize(rs.leaf_of.size() + ntri, lt::kInvalid);
    for (uint32_t t = 0; t < ntri; ++t) {
      V3 p[3];
      triangle_world(rs, i, t, p);
      real area = 0.5f * lt::vlen(cross(sub(p[1], p[0]), sub(p[2], p[0])));
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

inline real balance_heuristic(real a, real b) {  // R/math/mod.rs:163-171
  real sum = a + b;
  if (a <= 0.0f || sum <= 0.0f) return 0.0f;
  return a / sum;
}

constexpr real kRayEpsilon = 1.0e-4f;  // R/integrator/mod.rs:49
inline V3 offset_origin(V3 p, V3 ng, V3 dir) {
  V3 off = dot(dir, ng) >= 0.0f ? ng : neg(ng);
  return add(p, mul(kRayEpsilon, off));
}

// LightSampler::build_from_scene (R/light/mod.rs:133-191): entries with a
// positive weight in the order tree (root flux), environment (total power),
// directional lights (luminance(color) * intensity); the running total and
// the CDF scaled by 1 / total.  kind: 0 tree, 1 environment, 2 + i
// directional light i.
struct Category {
  int kind = -1;
  real pmf = 0.0f;
};
struct LightSampler {
  std::vector<int> kind;
  std::vector<real> weight, cdf;
  real total = 0.0f;
};
inline LightSampler light_sampler(const RenderScene& rs) {
  LightSampler ls;
  real wt = std::fmax(rs.tree.root_flux(), 0.0f);
  if (rs.tree.nodes.empty()) wt = 0.0f;
  if (wt > 0.0f) { ls.kind.push_back(0); ls.weight.push_back(wt); }
  if (rs.has_env) {
    real we = std::fmax(lt::env::total_power(rs.env), 0.0f);
    if (we > 0.0f) { ls.kind.push_back(1); ls.weight.push_back(we); }
  }
  for (size_t i = 0; i < rs.dirs.size(); ++i) {
    real w = std::fmax(lt::luminance(rs.dirs[i].color) * rs.dirs[i].intensity, 0.0f);
    if (w > 0.0f) { ls.kind.push_back(2 + (int)i); ls.weight.push_back(w); }
  }
  ls.cdf.push_back(0.0f);
  real total = 0.0f;
  for (real w : ls.weight) {
    total += w;
    ls.cdf.push_back(total);
  }
  if (total > 0.0f) {
    real inv = 1.0f / total;
    for (real& c : ls.cdf) c *= inv;
  }
  ls.total = total;
  return ls;
}
// sample_category (R/light/mod.rs:197-213): partition_point(c <= u) - 1
inline bool sample_category(const RenderScene& rs, real u, Category* out) {
  LightSampler ls = light_sampler(rs);
  const int n = (int)ls.kind.size();
  if (n == 0 || ls.total <= 0.0f) return false;
  u = clampf(u, 0.0f, 1.0f);
  int pp = 0;
  while (pp <= n && ls.cdf[pp] <= u) ++pp;
  int idx = pp > 0 ? pp - 1 : 0;
  if (idx > n - 1) idx = n - 1;
  out->kind = ls.kind[idx];
  out->pmf = ls.weight[idx] / ls.total;
  return true;
}
inline real category_pmf(const RenderScene& rs, int kind) {  // :215-224
  LightSampler ls = light_sampler(rs);
  if (ls.total <= 0.0f) return 0.0f;
  for (size_t i = 0; i < ls.kind.size(); ++i)
    if (ls.kind[i] == kind) return ls.weight[i] / ls.total;
  return 0.0f;
}

// direct_light_mis_contribution (R/integrator/mis.rs:141-210) with
// sample_light_mis_compensated_lazy (R/light/mod.rs:289-331),
// sample_li_for_triangle (R/light/area.rs:12-54) and the environment's
// MIS-compensated sample (R/light/environment.rs:260-276).
// em (optional): the material whose BSDF value the contribution carries,
// every decision and density staying those of m (detached evaluation, see
// trace_radiance).
inline V3 direct_light(const RenderScene& rs, const Material& m, const Vertex& vtx, real u_root, real u_tree,
                       real u_aux, real us0, real us1, const Material* em = nullptr) {
  (void)u_aux;
  const V3 zero = v3(0.f, 0.f, 0.f);
  Category cat;
  if (!sample_category(rs, u_root, &cat)) return zero;
  V3 le, wi;
  real li_pdf, selection;
  uint32_t t_inst = lt::kInvalid, t_tri = lt::kInvalid;
  if (cat.kind == 0) {
    lt::Query q;
    if (!mat_query(rs, m, vtx, &q)) return zero;
    uint32_t leaf;
    real leaf_pmf;
    if (!lt::sample_tree(rs.tree, q, u_tree, &leaf, &leaf_pmf)) return zero;
    const lt::Node& ln = rs.tree.nodes[leaf];
    // sample_triangle_point (R/scene/mod.rs:616-633)
    V3 p[3];
    triangle_world(rs, ln.inst, ln.tri, p);
    real su0 = std::sqrt(clampf(us0, 0.0f, 1.0f));
    real u1 = clampf(us1, 0.0f, 1.0f);
    real b0 = 1.0f - su0, b1 = u1 * su0, b2 = (1.0f - u1) * su0;
    V3 lp = interp3(b0, b1, b2, p[0], p[1], p[2]);
    real area = 0.5f * lt::vlen(cross(sub(p[1], p[0]), sub(p[2], p[0])));
    real pdf_area = area > 0.0f ? 1.0f / area : 0.0f;
    if (pdf_area <= 0.0f) return zero;
    V3 to_light = sub(lp, vtx.p);
    real d2 = len_sq(to_light);
    if (d2 <= 0.0f) return zero;
    real dist = std::sqrt(d2);
    wi = sdiv(to_light, dist);
    const Material& lm0 = rs.materials[rs.inst_material[ln.inst]];
    if (!mat_may_emit(lm0)) return zero;
    Vertex lv = vertex_at(rs, ln.inst, ln.tri, b0, b1, b2, wi);
    const Material lm = material_at(rs, lm0, lv);
    if (!mat_le(lm, lv, &le)) return zero;
    real cos_light = std::fmax(dot(lv.ng, neg(wi)), 0.0f);
    if (cos_light <= 0.0f) return zero;
    li_pdf = pdf_area * d2 / cos_light;
    selection = cat.pmf * leaf_pmf;
    t_inst = ln.inst;
    t_tri = ln.tri;
  } else if (cat.kind == 1) {
    lt::env::Sample es;
    if (!lt::env::sample(rs.env, rs.env.dist_mis, us0, us1, &es)) return zero;
    if (es.pdf <= 0.0f) return zero;
    le = es.radiance;
    wi = es.dir;
    li_pdf = es.pdf;
    selection = cat.pmf;
  } else {  // directional::sample_li (R/light/directional.rs:34-43)
    const RenderScene::DirLight& dl = rs.dirs[cat.kind - 2];
    le = smul(dl.color, dl.intensity);
    wi = neg(dl.direction);
    li_pdf = 1.0f;
    selection = cat.pmf;
  }
  if (li_pdf <= 0.0f || selection <= 0.0f) return zero;
  V3 f = mat_eval(m, vtx, wi);
  real bsdf_pdf = len_sq(f) == 0.0f ? 0.0f : mat_pdf(m, vtx, wi);
  if (len_sq(f) == 0.0f) return zero;
  real cos_surface = std::fabs(dot(vtx.ng, wi));
  if (cos_surface <= 0.0f) return zero;
  // unoccluded_ray (R/integrator/mod.rs:239-265): a light triangle target, or
  // an infinite light that any hit blocks
  SceneHit sh;
  if (closest_hit(*rs.geo, offset_origin(vtx.p, vtx.ng, wi), wi, INFINITY, &sh, nullptr) &&
      !(t_inst != lt::kInvalid && sh.inst == t_inst && sh.tri == t_tri))
    return zero;
  real pdf_total = selection * li_pdf;
  if (pdf_total <= 0.0f) return zero;
  if (em) f = mat_eval(*em, vtx, wi);
  if (cat.kind >= 2) return smul(vmul(le, f), cos_surface / pdf_total);  // delta light: no MIS
  real mis = balance_heuristic(pdf_total, bsdf_pdf);
  return smul(vmul(le, f), cos_surface * mis / pdf_total);
}

// pdf_for_triangle_hit (R/light/mod.rs:407-429) with
// Scene::area_light_pdf_solid_angle (R/scene/mod.rs:587-611).
inline real light_pdf_for_hit(const RenderScene& rs, const Material& shading_m, const Vertex& vtx, const Vertex& lv) {
  if (rs.tree.nodes.empty()) return 0.0f;
  lt::Query q;
  if (!mat_query(rs, shading_m, vtx, &q)) return 0.0f;
  uint32_t leaf = rs.leaf_node(lv.inst, lv.tri);
  real leaf_pmf = lt::pdf_leaf(rs.tree, q, leaf);
  if (leaf_pmf <= 0.0f) return 0.0f;
  real cat_pmf = category_pmf(rs, 0);
  const real area = rs.tree.nodes[leaf].area;
  if (area <= 0.0f) return 0.0f;
  real pdf_area = 1.0f / area;
  V3 to_light = sub(lv.p, vtx.p);
  real d2 = len_sq(to_light);
  if (d2 <= 0.0f) return 0.0f;
  real dist = std::sqrt(d2);
  V3 wi = sdiv(to_light, dist);
  real cos_light = std::fmax(dot(lv.ng, neg(wi)), 0.0f);
  if (cos_light <= 0.0f) return 0.0f;
  return cat_pmf * leaf_pmf * (pdf_area * d2 / cos_light);
}

// emitted_radiance_from_bsdf_sample_area (R/integrator/mis.rs:212-234)
inline V3 emitted_from_bsdf_sample(const RenderScene& rs, V3 throughput, real bsdf_pdf, bool delta,
                                   const Material& shading_m, const Vertex& vtx, const Vertex& lv,
                                   const Material& light_m) {
  V3 le;
  if (!mat_le(light_m, lv, &le)) return v3(0.f, 0.f, 0.f);
  if (delta) return vmul(throughput, le);
  real light_pdf = light_pdf_for_hit(rs, shading_m, vtx, lv);
  real mis = balance_heuristic(bsdf_pdf, light_pdf);
  return smul(vmul(throughput, le), mis);
}

// infinite_light_le (R/light/environment.rs:305-311)
inline V3 env_le(const RenderScene& rs, V3 dir) {
  return rs.has_env ? lt::env::radiance(rs.env, dir) : v3(0.f, 0.f, 0.f);
}
// emitted_radiance_from_bsdf_sample_infinite (R/integrator/mis.rs:236-257)
// with pdf_for_environment_hit (R/light/mod.rs:431-443), MIS-compensated.
inline V3 emitted_infinite(const RenderScene& rs, V3 throughput, real bsdf_pdf, bool delta, V3 dir) {
  V3 le = env_le(rs, dir);
  if (le.x == 0.0f && le.y == 0.0f && le.z == 0.0f) return v3(0.f, 0.f, 0.f);
  if (delta) return vmul(throughput, le);
  real cat = category_pmf(rs, 1);
  real light_pdf = cat <= 0.0f ? 0.0f : cat * lt::env::pdf(rs.env, rs.env.dist_mis, dir);
  return smul(vmul(throughput, le), balance_heuristic(bsdf_pdf, light_pdf));
}

// specular_ray_differential (R/integrator/mod.rs:88-138) with
// reflected_differential_directions (:140-157) for reflections,
// transmitted_differential_directions (:159-206) for refractions and
// differential_is_reasonable (:208-219).
inline bool specular_differential(const Ray& in, const Vertex& v, V3 wi_world, const BSample& s, Ray* out) {
  if (!in.has_diff) return false;
  V3 wo = v.wo, wi = normalize_or_zero3(wi_world);
  V3 dndx = add(mul(v.dudx, v.dndu), mul(v.dvdx, v.dndv));
  V3 dndy = add(mul(v.dudy, v.dndu), mul(v.dvdy, v.dndv));
  V3 dwodx = sub(neg(in.rxd), wo), dwody = sub(neg(in.ryd), wo);
  V3 rxo = offset_origin(add(v.p, v.dpdx), v.ng, wi), ryo = offset_origin(add(v.p, v.dpdy), v.ng, wi);
  V3 n = v.ns;
  V3 rxd, ryd;
  if (!s.transmission) {
    real dwn_dx = dot(dwodx, n) + dot(wo, dndx), dwn_dy = dot(dwody, n) + dot(wo, dndy);
    real wn = dot(wo, n);
    rxd = add(sub(wi, dwodx), mul(2.0f, add(mul(wn, dndx), mul(dwn_dx, n))));
    ryd = add(sub(wi, dwody), mul(2.0f, add(mul(wn, dndy), mul(dwn_dy, n))));
  } else {
    const real eta = s.eta;
    if (std::fabs(eta - 1.0f) <= 1.0e-6f) {
      rxd = sub(wi, dwodx);
      ryd = sub(wi, dwody);
    } else {
      if (eta <= 0.0f) return false;
      if (dot(wo, n) < 0.0f) {
        n = neg(n);
        dndx = neg(dndx);
        dndy = neg(dndy);
      }
      real win = dot(wi, n);
      if (std::fabs(win) <= 1.0e-6f) return false;
      real dwn_dx = dot(dwodx, n) + dot(wo, dndx), dwn_dy = dot(dwody, n) + dot(wo, dndy);
      real wn = dot(wo, n);
      real mu = wn * eta - std::fabs(win);
      real dmu_scale = eta + eta * eta * wn / win;
      real dmudx = dwn_dx * dmu_scale, dmudy = dwn_dy * dmu_scale;
      rxd = add(add(sub(wi, mul(eta, dwodx)), mul(mu, dndx)), mul(dmudx, n));
      ryd = add(add(sub(wi, mul(eta, dwody)), mul(mu, dndy)), mul(dmudy, n));
    }
  }
  const real kMax = 1.0e16f;
  if (!v3_finite(rxo) || !v3_finite(ryo) || !v3_finite(rxd) || !v3_finite(ryd) || len_sq(rxo) > kMax ||
      len_sq(ryo) > kMax || len_sq(rxd) > kMax || len_sq(ryd) > kMax)
    return false;
  out->rxo = rxo;
  out->ryo = ryo;
  out->rxd = rxd;
  out->ryd = ryd;
  return true;
}
// spawn_scattered_ray (R/integrator/mod.rs:58-75)
inline Ray spawn_scattered(const Ray& in, real hit_t, const Vertex& v, const BSample& s, V3 wi) {
  Ray r;
  r.o = offset_origin(v.p, v.ng, wi);
  r.d = wi;
  r.cone = make_cone(cone_width_at(in.cone, hit_t), in.cone.spread + s.cone_spread);
  if (s.delta) r.has_diff = specular_differential(in, v, wi, s, &r);
  return r;
}

// The sampled direction's weight with the BSDF value taken from another
// material mb and the sampling density of the path's own sample s: mb's
// colour for NormalizedLambert and Mirror (whose weights are their colour);
// for OpenPBR f_b(wi) |cos| / s.pdf on the rough lobes (finalize's weight),
// the lobe-only value of a rough thin-walled transmission over s.pdf (that
// sample's weight), and on a delta lobe mb's numerator of the same lobe over
// s's lobe probability.  A delta lobe that is not delta under mb yields NaN.
inline V3 detached_weight(const Material& mb, const Vertex& v, const BSample& s) {
  if (mb.kind == kMatLambert || mb.kind == kMatMirror) return mb.color;
  if (mb.kind != kMatOpenPbr) return vsplat((real)NAN);
  V3 wol = normalize_or_zero3(onb_to_local(v.frame, v.wo));
  const PbrBsdf pb = make_bsdf(mb.pbr, v.front_face, wol, mb.luts);
  if (s.delta) {
    BSample t;
    if (!sample_lobe(pb, wol, s.lobe, s.p_lobe, 0.0f, 0.0f, &t) || !t.delta) return vsplat((real)NAN);
    return t.weight;
  }
  if (s.transmission && pb.thin_walled) {
    V3 wi = s.wi;
    LayerW w = layer_weights(pb, wol, &wi);
    return sdiv(smul(vmul(w.spec_btdf, eval_thin_btdf(pb, wol, s.wi)), std::fabs(s.wi.z)), s.pdf);
  }
  return smul(bsdf_eval(pb, wol, s.wi), std::fabs(s.wi.z) / s.pdf);
}

// mis::trace_radiance (R/integrator/mis.rs:19-139) for a camera ray with its
// differential and cone.
//
// ev (optional): detached evaluation, the estimator a path-replay adjoint
// differentiates.  ev is the same geometry with other materials and
// textures; every sampling decision, density, MIS weight and Russian-roulette
// probability is that of rs's path, and the BSDF values (the sampled
// directions' weights, next-event contributions) and the emission are ev's.
// With ev's materials equal to rs's the result is rs's radiance.  A central
// difference over ev's parameters of this estimator is the derivative of the
// path's contribution with the sampling held fixed.
inline V3 trace_radiance(const RenderScene& rs, const Ray& camera_ray, const PathSampler& ps, uint32_t max_depth,
                         const RenderScene* ev = nullptr) {
  V3 radiance = v3(0.f, 0.f, 0.f), throughput = v3(1.f, 1.f, 1.f);
  V3 value = throughput;  // ev's throughput (throughput itself without ev)
  SceneHit h;
  if (!closest_hit(*rs.geo, camera_ray.o, camera_ray.d, INFINITY, &h, nullptr)) return env_le(rs, camera_ray.d);
  Ray ray = camera_ray;
  real hit_t = h.t;
  Vertex vtx = shading_vertex(rs, h, ray);
  Material mat = material_at(rs, rs.materials[vtx.material], vtx);
  Material emat = ev ? material_at(*ev, ev->materials[vtx.material], vtx) : mat;
  V3 le;
  if (mat_le(emat, vtx, &le)) radiance = add(radiance, le);
  const uint32_t rr_start = 4;
  for (uint32_t depth = 0; depth < max_depth; ++depth) {
    const uint32_t base = 2u + depth * 16u;
    if (!mat_is_pure_emitter(mat)) {
      float us[2];
      ps.s2(base + 3, us);
      V3 dl = direct_light(rs, mat, vtx, ps.s1(base + 0), ps.s1(base + 1), ps.s1(base + 2), us[0], us[1],
                           ev ? &emat : nullptr);
      radiance = add(radiance, vmul(ev ? value : throughput, dl));
    }
    BSample s;
    V3 wi;
    float udir[2];
    ps.s2(base + 7, udir);
    if (!mat_sample(mat, vtx, ps.s1(base + 5), udir[0], udir[1], &s, &wi)) break;
    const bool delta = s.delta;
    throughput = vmul(throughput, s.weight);
    if (ev) value = vmul(value, detached_weight(emat, vtx, s));
    if (depth + 1 >= rr_start) {
      real q = clampf(rmax(throughput.x, rmax(throughput.y, throughput.z)), 0.05f, 1.0f);
      if (ps.s1(base + 13) > q) break;
      throughput = sdiv(throughput, q);
      value = sdiv(value, q);
    }
    const V3 carried = ev ? value : throughput;
    const Ray next = spawn_scattered(ray, hit_t, vtx, s, wi);
    SceneHit nh;
    if (!closest_hit(*rs.geo, next.o, next.d, INFINITY, &nh, nullptr)) {
      radiance = add(radiance, emitted_infinite(rs, carried, s.pdf, delta, wi));
      break;
    }
    Vertex nv = shading_vertex(rs, nh, next);
    const Material nm = material_at(rs, rs.materials[nv.material], nv);
    const Material enm = ev ? material_at(*ev, ev->materials[nv.material], nv) : nm;
    radiance = add(radiance, emitted_from_bsdf_sample(rs, carried, s.pdf, delta, mat, vtx, nv, enm));
    vtx = nv;
    mat = nm;
    emat = enm;
    ray = next;
    hit_t = nh.t;
  }
  return radiance;
}
// A ray without differential and with a zero cone (single-path test entry).
inline V3 trace_radiance(const RenderScene& rs, V3 o, V3 d, const PathSampler& ps, uint32_t max_depth) {
  Ray r;
  r.o = o;
  r.d = d;
  return trace_radiance(rs, r, ps, max_depth);
}

}  // namespace orc

