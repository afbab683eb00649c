// Oracle: CPU restatement of the reference renderer's ray-query path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU path is
// compared against; it is never linked into the product library, and only
// tests/, the smoke check and the bench's CPU-baseline leg may load it.
//
// Every function follows a reference function line by line (file:line cited
// on each); operation order, comparison direction and NaN behaviour are kept,
// because the GPU kernels are required to agree with it bit for bit.
//   R/ = crates/renderer/src/
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).
//
// Pinning (what checks each function, see tests/test_oracle_*.py):
//   intersect_triangle   known answers of R/math/ray.rs:229-321 + fp64 brute force
//   qbvh build           invariants of R/qbvh.rs:1128-1157 + brute force
//   traverse_qbvh        brute-force closest of R/qbvh.rs:856-984
//   scene closest_hit    known answers of R/scene/mod.rs:1028-1105 + brute force
//   unoccluded_ray       closed-form cases (hit target / blocker / miss) + brute force

#include "geometry.h"
#include "sampler.h"
#include "render.h"
#include "texture.h"

using namespace orc;


// ===========================================================================
// C interface used by the Python tests (ctypes).
extern "C" {

// Ray record shared with the product ABI: o.xyz, tmax, d.xyz, pad (32 B).
struct OrcRay {
  float o[3];
  float tmax;
  float d[3];
  float pad;
};
// Hit record: instance, triangle (~0u on miss), t, b0, b1, b2 (24 B).
struct OrcHit {
  uint32_t inst, tri;
  float t, b0, b1, b2;
};

int orc_intersect_triangle(const float* ray6, float t_max, const float* v9, float* out4) {
  TriHit h;
  bool ok = intersect_triangle(v3(ray6[0], ray6[1], ray6[2]), v3(ray6[3], ray6[4], ray6[5]), t_max,
                               v3(v9[0], v9[1], v9[2]), v3(v9[3], v9[4], v9[5]), v3(v9[6], v9[7], v9[8]), &h);
  if (ok) {
    out4[0] = h.t;
    out4[1] = h.b0;
    out4[2] = h.b1;
    out4[3] = h.b2;
  }
  return ok ? 1 : 0;
}

float orc_difference_of_products(float a, float b, float c, float d) { return difference_of_products(a, b, c, d); }
float orc_gamma(int n) { return gamma_n(n); }

// QBVH over primitive bounds (bmin/bmax: 3 floats each per primitive).
void* orc_qbvh_build(const float* bmin, const float* bmax, size_t n) {
  if (n == 0) return nullptr;
  std::vector<Bounds> b(n);
  for (size_t i = 0; i < n; ++i)
    b[i] = Bounds{v3(bmin[3 * i], bmin[3 * i + 1], bmin[3 * i + 2]), v3(bmax[3 * i], bmax[3 * i + 1], bmax[3 * i + 2])};
  Qbvh* q = new Qbvh();
  build_qbvh(b.data(), n, q);
  return q;
}
void orc_qbvh_free(void* h) { delete (Qbvh*)h; }
size_t orc_qbvh_node_count(void* h) { return ((Qbvh*)h)->nodes.size(); }
size_t orc_qbvh_prim_count(void* h) { return ((Qbvh*)h)->prim.size(); }
// bounds: 24 floats per node in reference order (min_x[4],min_y[4],min_z[4],max_x[4],max_y[4],max_z[4]).
void orc_qbvh_export(void* h, float* bounds, uint32_t* children, uint8_t* valid, uint32_t* prim) {
  Qbvh* q = (Qbvh*)h;
  for (size_t i = 0; i < q->nodes.size(); ++i) {
    const Node& n = q->nodes[i];
    for (int k = 0; k < 4; ++k) {
      bounds[24 * i + 0 + k] = n.min_x[k];
      bounds[24 * i + 4 + k] = n.min_y[k];
      bounds[24 * i + 8 + k] = n.min_z[k];
      bounds[24 * i + 12 + k] = n.max_x[k];
      bounds[24 * i + 16 + k] = n.max_y[k];
      bounds[24 * i + 20 + k] = n.max_z[k];
    }
    std::memcpy(children + 4 * i, n.children, 16);
    valid[i] = n.valid_mask;
  }
  std::memcpy(prim, q->prim.data(), 4 * q->prim.size());
}
// Leaf visits of traverse_qbvh with a pass-through callback: writes (offset,count)
// pairs in visit order, returns the number of leaf callbacks.
size_t orc_qbvh_visit_leaves(void* h, const float* ray6, float t_max, uint32_t* out, size_t cap) {
  Qbvh* q = (Qbvh*)h;
  size_t n = 0;
  traverse_qbvh(
      *q, v3(ray6[0], ray6[1], ray6[2]), v3(ray6[3], ray6[4], ray6[5]), t_max,
      [&](uint32_t off, uint32_t count, float tm) {
        if (n < cap) {
          out[2 * n] = off;
          out[2 * n + 1] = count;
        }
        n++;
        return tm;
      },
      nullptr);
  return n;
}

// The reference test helper qbvh_closest (R/qbvh.rs:883-900) over primitive
// boxes: closest box by slab_intersect_scalar (R/qbvh.rs:856-869, glam
// Vec3::min/max = f32::min/max), t_max shrinking as boxes are accepted.
// Returns the primitive index (or -1) and writes its entry t.
long orc_qbvh_closest_box(void* h, const float* bmin, const float* bmax, const float* ray6, float* t_out) {
  Qbvh* q = (Qbvh*)h;
  V3 o = v3(ray6[0], ray6[1], ray6[2]), d = v3(ray6[3], ray6[4], ray6[5]);
  V3 inv = v3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  auto slab = [&](uint32_t idx, float t_max, float* t) {
    V3 lo = v3(bmin[3 * idx], bmin[3 * idx + 1], bmin[3 * idx + 2]);
    V3 hi = v3(bmax[3 * idx], bmax[3 * idx + 1], bmax[3 * idx + 2]);
    V3 t1 = v3((lo.x - o.x) * inv.x, (lo.y - o.y) * inv.y, (lo.z - o.z) * inv.z);
    V3 t2 = v3((hi.x - o.x) * inv.x, (hi.y - o.y) * inv.y, (hi.z - o.z) * inv.z);
    V3 tn = vmin(t1, t2), tx = vmax(t1, t2);
    float tmin = rmax(rmax(rmax(tn.x, tn.y), tn.z), 0.0f);
    float tmax = rmin(rmin(rmin(tx.x, tx.y), tx.z), t_max);
    *t = tmin;
    return tmin <= tmax;
  };
  long best = -1;
  float best_t = kInf;
  traverse_qbvh(
      *q, o, d, kInf,
      [&](uint32_t off, uint32_t count, float tm) {
        float nt = tm;
        for (uint32_t k = off; k < off + count; ++k) {
          uint32_t idx = q->prim[k];
          float t;
          if (slab(idx, nt, &t) && (best < 0 || t < best_t)) {
            best = idx;
            best_t = t;
            nt = t;
          }
        }
        return nt;
      },
      nullptr);
  *t_out = best_t;
  return best;
}

// Scene: meshes are added first, then instances (column-major 4x4 matrices,
// local_to_world and its inverse world_to_local), then orc_scene_build.
void* orc_scene_new(void) { return new Scene(); }
void orc_scene_free(void* h) { delete (Scene*)h; }
int orc_scene_add_mesh(void* h, const float* pos, size_t nv, const uint32_t* idx, size_t nt) {
  Scene* s = (Scene*)h;
  if (nv == 0) return -1;
  Mesh m;
  m.pos.assign(pos, pos + 3 * nv);
  m.idx.assign(idx, idx + 3 * nt);
  // compute_bounds, R/scene/mesh.rs:668-680: min/max over vertex positions
  Bounds b{m.vert(0), m.vert(0)};
  for (size_t i = 1; i < nv; ++i) {
    b.mn = vmin(b.mn, m.vert((uint32_t)i));
    b.mx = vmax(b.mx, m.vert((uint32_t)i));
  }
  m.bounds = b;
  s->meshes.push_back(std::move(m));
  return (int)s->meshes.size() - 1;
}
int orc_scene_add_instance(void* h, uint32_t mesh, const float* l2w, const float* w2l) {
  Scene* s = (Scene*)h;
  if (mesh >= s->meshes.size()) return -1;
  Instance in;
  in.mesh = mesh;
  for (int k = 0; k < 16; ++k) {
    in.l2w[k] = l2w[k];
    in.w2l[k] = w2l[k];
  }
  in.world_bounds = transform_bounds(s->meshes[mesh].bounds, in.l2w);
  s->inst.push_back(in);
  s->built = false;
  return (int)s->inst.size() - 1;
}
// Scene::build_qbvh, R/scene/mod.rs:305-316: one BLAS per mesh, TLAS over instance world bounds.
void orc_scene_build(void* h) {
  Scene* s = (Scene*)h;
  for (auto& m : s->meshes) {
    size_t nt = m.tri_count();
    std::vector<Bounds> tb(nt);
    for (size_t t = 0; t < nt; ++t) {  // Mesh::triangle_bounds, R/scene/mesh.rs:104-111
      V3 a = m.vert(m.idx[3 * t]), b = m.vert(m.idx[3 * t + 1]), c = m.vert(m.idx[3 * t + 2]);
      tb[t] = Bounds{vmin(vmin(a, b), c), vmax(vmax(a, b), c)};
    }
    m.qbvh = Qbvh();
    m.has_qbvh = build_qbvh(tb.data(), nt, &m.qbvh);
  }
  std::vector<Bounds> ib(s->inst.size());
  for (size_t i = 0; i < s->inst.size(); ++i) ib[i] = s->inst[i].world_bounds;
  s->tlas = Qbvh();
  build_qbvh(ib.data(), ib.size(), &s->tlas);
  s->built = true;
}
int orc_scene_is_built(void* h) { return ((Scene*)h)->built ? 1 : 0; }
void* orc_scene_mesh_qbvh(void* h, uint32_t mesh) { return &((Scene*)h)->meshes[mesh].qbvh; }
void* orc_scene_tlas(void* h) { return &((Scene*)h)->tlas; }

// Closest hit for n rays; returns -1 when the QBVH is not built
// (ClosestHitError::QbvhNotBuilt, R/scene/mod.rs:100-113), else 0.
// stats (nullable): per-ray uint32 triples (nodes visited, triangles tested, instances entered).
int orc_closest_hit(void* h, const OrcRay* rays, size_t n, OrcHit* hits, uint32_t* stats, int threads) {
  Scene* s = (Scene*)h;
  if (!s->inst.empty() && !s->built) return -1;
  parallel_for(n, threads, [&](size_t i) {
    const OrcRay& r = rays[i];
    SceneHit sh;
    TraversalStats st;
    bool ok = closest_hit(*s, v3(r.o[0], r.o[1], r.o[2]), v3(r.d[0], r.d[1], r.d[2]), r.tmax, &sh,
                          stats ? &st : nullptr);
    if (ok) {
      hits[i] = OrcHit{sh.inst, sh.tri, sh.t, sh.b0, sh.b1, sh.b2};
    } else {
      hits[i] = OrcHit{~0u, ~0u, (float)kInf, 0.f, 0.f, 0.f};
    }
    if (stats) {
      stats[3 * i] = (uint32_t)st.nodes;
      stats[3 * i + 1] = (uint32_t)st.tris;
      stats[3 * i + 2] = (uint32_t)st.insts;
    }
  });
  return 0;
}

// unoccluded_ray for n shadow rays.  target (nullable): per ray (instance,
// triangle) with instance = ~0u for "no target"; distance (nullable = +inf).
// out: 1 = unoccluded.  stats as for orc_closest_hit.
int orc_unoccluded(void* h, const OrcRay* rays, size_t n, const uint32_t* target, const float* distance,
                   uint8_t* out, uint32_t* stats, int threads) {
  Scene* s = (Scene*)h;
  if (!s->inst.empty() && !s->built) return -1;
  parallel_for(n, threads, [&](size_t i) {
    const OrcRay& r = rays[i];
    int64_t ti = -1, tt = -1;
    if (target && target[2 * i] != ~0u) {
      ti = target[2 * i];
      tt = target[2 * i + 1];
    }
    float dist = distance ? distance[i] : kInf;
    TraversalStats st;
    out[i] = unoccluded_ray(*s, v3(r.o[0], r.o[1], r.o[2]), v3(r.d[0], r.d[1], r.d[2]), dist, ti, tt,
                            stats ? &st : nullptr)
                 ? 1
                 : 0;
    if (stats) {
      stats[3 * i] = (uint32_t)st.nodes;
      stats[3 * i + 1] = (uint32_t)st.tris;
      stats[3 * i + 2] = (uint32_t)st.insts;
    }
  });
  return 0;
}

}  // extern "C"

// ---- sampler and camera --------------------------------------------------

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
  CF3 eye{cam[0], cam[1], cam[2]}, fwd{cam[3], cam[4], cam[5]}, right{cam[6], cam[7], cam[8]},
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
      CF3 d = dir_for(px, py);
      CF3 dxd = dir_for(px + 1.0f, py);
      CF3 dyd = dir_for(px, py + 1.0f);
      CF3 rxd = normalize_or_zero(add(d, scale(sc, sub(dxd, d))));
      CF3 ryd = normalize_or_zero(add(d, scale(sc, sub(dyd, d))));
      float* o = out + 13 * ((size_t)y * rx + x);
      o[0] = eye.x; o[1] = eye.y; o[2] = eye.z;
      o[3] = d.x; o[4] = d.y; o[5] = d.z;
      o[6] = rxd.x; o[7] = rxd.y; o[8] = rxd.z;
      o[9] = ryd.x; o[10] = ryd.y; o[11] = ryd.z;
      o[12] = std::fmax(spread * sc, 0.0f);  // RayCone::new(0, spread * scale)
    }
}

}  // extern "C"

extern "C" {

// ---- path tracer ------------------------------------------------------------
// Material record as laid out by the product ABI (TptMaterial): kind, pad,
// then the kind's parameters (TptOpenPbr for OpenPBR).
struct OrcMaterial {
  uint32_t kind, pad;
  float raw[48];
};
static OpenPbr parse_open_pbr(const OrcMaterial* m);

// Shading data for an oracle scene: per mesh vertex normals / uvs (nullable),
// per instance material index, and the material table.
void* orc_render_scene_new(void* geo, uint32_t mesh_count) {
  RenderScene* r = new RenderScene();
  r->geo = (Scene*)geo;
  r->meshes.resize(mesh_count);
  r->inst_material.assign(((Scene*)geo)->inst.size(), 0);
  return r;
}
void orc_render_scene_free(void* h) { delete (RenderScene*)h; }
void orc_render_set_mesh(void* h, uint32_t mesh, const float* normals, const float* uvs, size_t nv) {
  RenderScene* r = (RenderScene*)h;
  if (normals) r->meshes[mesh].nrm.assign(normals, normals + 3 * nv);
  if (uvs) r->meshes[mesh].uv.assign(uvs, uvs + 2 * nv);
}
void orc_render_set_instance_material(void* h, uint32_t inst, uint32_t mat) {
  ((RenderScene*)h)->inst_material[inst] = mat;
}
// Appends a directional light (direction of travel, as DirectionalLight::new
// leaves it; tint; perpendicular irradiance).
void orc_render_add_directional(void* h, const float* dir, const float* color, float intensity) {
  RenderScene* r = (RenderScene*)h;
  r->dirs.push_back(RenderScene::DirLight{v3(dir[0], dir[1], dir[2]), v3(color[0], color[1], color[2]), intensity});
}
// Light-sampler categories: up to cap entries of (kind, pmf); returns the count.
uint32_t orc_render_light_categories(void* h, int* kind, float* pmf, uint32_t cap) {
  const RenderScene& rs = *(RenderScene*)h;
  LightSampler ls = light_sampler(rs);
  uint32_t n = 0;
  for (size_t i = 0; i < ls.kind.size() && n < cap; ++i, ++n) {
    kind[n] = ls.kind[i];
    pmf[n] = ls.total > 0.0f ? ls.weight[i] / ls.total : 0.0f;
  }
  return (uint32_t)ls.kind.size();
}
// Appends a texture to the scene's table (before orc_render_set_materials:
// the light tree reads the emission textures' maxima).
uint32_t orc_render_add_texture(void* h, uint32_t w, uint32_t hh, int channels, const float* px) {
  RenderScene* r = (RenderScene*)h;
  r->textures.push_back(texture::from_pixels(w, hh, channels, px));
  return (uint32_t)(r->textures.size() - 1);
}
// Sets the material table and builds the area-light table and light tree.
void orc_render_set_materials(void* h, const OrcMaterial* m, size_t n) {
  RenderScene* r = (RenderScene*)h;
  r->materials.assign(n, Material{});
  for (size_t i = 0; i < n; ++i) {
    Material& d = r->materials[i];
    const float* f = m[i].raw;
    d.kind = m[i].kind;
    if (d.kind == kMatOpenPbr) {
      d.pbr = parse_open_pbr(&m[i]);
      const OpenPbr& o = d.pbr;
      d.luts = luts_for(o);
    } else {
      d.color = v3(f[0], f[1], f[2]);
      d.strength = f[3];
    }
  }
  build_lights(*r);
}

// Environment light from RGB pixels (EnvironmentLight::from_pixels,
// R/light/environment.rs:47-83).
void orc_render_set_environment(void* h, const float* px, uint32_t w, uint32_t hh, float scale, float rotate_y) {
  RenderScene* r = (RenderScene*)h;
  std::vector<V3> p((size_t)w * hh);
  for (size_t i = 0; i < p.size(); ++i) p[i] = v3(px[3 * i], px[3 * i + 1], px[3 * i + 2]);
  r->env = lt::env::make(w, hh, std::move(p), scale, rotate_y);
  r->has_env = true;
}
// Environment queries for tests: pdf (mis 0/1), radiance, sample (returns
// 0 if none; out: dir[3], radiance[3], pdf), total power.
float orc_env_pdf(void* h, const float* d, int mis) {
  const RenderScene& r = *(RenderScene*)h;
  return lt::env::pdf(r.env, mis ? r.env.dist_mis : r.env.dist, v3(d[0], d[1], d[2]));
}
void orc_env_radiance(void* h, const float* d, float* out) {
  V3 l = lt::env::radiance(((RenderScene*)h)->env, v3(d[0], d[1], d[2]));
  out[0] = l.x;
  out[1] = l.y;
  out[2] = l.z;
}
int orc_env_sample(void* h, float u, float v, int mis, float* out) {
  const RenderScene& r = *(RenderScene*)h;
  lt::env::Sample s;
  if (!lt::env::sample(r.env, mis ? r.env.dist_mis : r.env.dist, u, v, &s)) return 0;
  const float o[7] = {s.dir.x, s.dir.y, s.dir.z, s.radiance.x, s.radiance.y, s.radiance.z, s.pdf};
  std::memcpy(out, o, sizeof(o));
  return 1;
}
float orc_env_total_power(void* h) { return lt::env::total_power(((RenderScene*)h)->env); }
void orc_env_uv(const float* d, float* uv) {
  real u, v;
  lt::env::direction_to_uv(v3(d[0], d[1], d[2]), &u, &v);
  uv[0] = u;
  uv[1] = v;
}
void orc_env_dir(float u, float v, float* d) {
  V3 r = lt::env::uv_to_direction(u, v);
  d[0] = r.x;
  d[1] = r.y;
  d[2] = r.z;
}
// emitted_radiance_from_bsdf_sample_infinite for tests
void orc_render_emitted_infinite(void* h, const float* thr, float bsdf_pdf, int delta, const float* dir, float* out) {
  V3 r = emitted_infinite(*(RenderScene*)h, v3(thr[0], thr[1], thr[2]), bsdf_pdf, delta != 0,
                          v3(dir[0], dir[1], dir[2]));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

// Light tree export, 20 floats per node: flux, mu[3], sigma_s2, nu[3],
// lambda, radius, area, then as u32 bits: left, right, parent, instance,
// triangle (kInvalid where absent), leaf flag, pad x3.  Returns the count.
uint32_t orc_render_light_tree(void* h, float* out, uint32_t cap) {
  const lt::Tree& t = ((RenderScene*)h)->tree;
  const uint32_t n = (uint32_t)t.nodes.size();
  for (uint32_t i = 0; i < n && i < cap; ++i) {
    const lt::Node& d = t.nodes[i];
    float* o = out + 20 * i;
    float v[11] = {d.flux, d.mu.x, d.mu.y, d.mu.z, d.sigma_s2, d.nu.x, d.nu.y, d.nu.z, d.lambda, d.radius, d.area};
    std::memcpy(o, v, sizeof(v));
    uint32_t u[9] = {d.left, d.right, d.parent, d.inst, d.tri, d.leaf ? 1u : 0u, 0u, 0u, 0u};
    std::memcpy(o + 11, u, sizeof(u));
  }
  return n;
}

// Light-tree queries of a diffuse shading point (p, n, rho): the stochastic
// descent (leaf node + pmf, false if none) and the reverse leaf pmf.
int orc_render_tree_sample(void* h, const float* p, const float* n, float rho, float u, uint32_t* leaf, float* pmf) {
  lt::Query q{v3(p[0], p[1], p[2]), v3(n[0], n[1], n[2]), rho};
  real pm = 0;
  const bool ok = lt::sample_tree(((RenderScene*)h)->tree, q, u, leaf, &pm);
  *pmf = (float)pm;
  return ok ? 1 : 0;
}
float orc_render_tree_pdf(void* h, const float* p, const float* n, float rho, uint32_t leaf) {
  lt::Query q{v3(p[0], p[1], p[2]), v3(n[0], n[1], n[2]), rho};
  return lt::pdf_leaf(((RenderScene*)h)->tree, q, leaf);
}

// NEE policy of the renderer (kNeeReference 0, kNeeOpenPbrProxy 1).
void orc_render_set_nee_policy(void* h, uint32_t policy) { ((RenderScene*)h)->nee_policy = policy; }
// The light-tree query of the shading vertex of triangle (inst, tri) at
// barycentrics b, seen along dir, under the scene's policy: the stochastic
// descent (returns 0 when the material builds no query or the descent gives
// none) and the reverse pmf of a leaf node (0 without a query).
int orc_render_vertex_tree_sample(void* h, uint32_t inst, uint32_t tri, const float* b, const float* dir, float u,
                                  uint32_t* leaf, float* pmf) {
  const RenderScene& rs = *(RenderScene*)h;
  Vertex v = vertex_at(rs, inst, tri, b[0], b[1], b[2], v3(dir[0], dir[1], dir[2]));
  lt::Query q;
  if (!mat_query(rs, material_at(rs, rs.materials[v.material], v), v, &q)) return 0;
  real pm = 0;
  const bool ok = lt::sample_tree(rs.tree, q, u, leaf, &pm);
  *pmf = (float)pm;
  return ok ? 1 : 0;
}
float orc_render_vertex_tree_pdf(void* h, uint32_t inst, uint32_t tri, const float* b, const float* dir,
                                 uint32_t leaf) {
  const RenderScene& rs = *(RenderScene*)h;
  Vertex v = vertex_at(rs, inst, tri, b[0], b[1], b[2], v3(dir[0], dir[1], dir[2]));
  lt::Query q;
  if (!mat_query(rs, material_at(rs, rs.materials[v.material], v), v, &q)) return 0.0f;
  return lt::pdf_leaf(rs.tree, q, leaf);
}

// direct_light_mis_contribution at the shading vertex of triangle
// (inst, tri) at barycentrics b[3], seen along `dir` (incident direction).
void orc_render_direct_light(void* h, uint32_t inst, uint32_t tri, const float* b, const float* dir, float u_root,
                             float u_tree, float u_aux, const float* us, float* out) {
  const RenderScene& rs = *(RenderScene*)h;
  Vertex v = vertex_at(rs, inst, tri, b[0], b[1], b[2], v3(dir[0], dir[1], dir[2]));
  V3 r = direct_light(rs, material_at(rs, rs.materials[v.material], v), v, u_root, u_tree, u_aux, us[0], us[1]);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
// emitted_radiance_from_bsdf_sample_area for the shading vertex (inst, tri,
// b, dir) and the light vertex (linst, ltri, lb, ldir).
void orc_render_emitted(void* h, uint32_t inst, uint32_t tri, const float* b, const float* dir, uint32_t linst,
                        uint32_t ltri, const float* lb, const float* ldir, const float* throughput, float bsdf_pdf,
                        int delta, float* out) {
  const RenderScene& rs = *(RenderScene*)h;
  Vertex v = vertex_at(rs, inst, tri, b[0], b[1], b[2], v3(dir[0], dir[1], dir[2]));
  Vertex lv = vertex_at(rs, linst, ltri, lb[0], lb[1], lb[2], v3(ldir[0], ldir[1], ldir[2]));
  V3 r = emitted_from_bsdf_sample(rs, v3(throughput[0], throughput[1], throughput[2]), bsdf_pdf, delta != 0,
                                  material_at(rs, rs.materials[v.material], v), v, lv,
                                  material_at(rs, rs.materials[lv.material], lv));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
// One path: trace_radiance from (o, d) with the PathSampler of pixel
// (px, py), sample si of an (rx, ry) frame at spp samples per pixel.
void orc_render_trace(void* h, const float* o, const float* d, uint32_t rx, uint32_t ry, uint32_t spp, uint32_t px,
                      uint32_t py, uint32_t si, uint32_t max_depth, float* out) {
  const RenderScene& rs = *(RenderScene*)h;
  PathSampler ps(px, py, si, spp, rx, ry);
  V3 r = trace_radiance(rs, v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]), ps, max_depth);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

// SG helpers of R/math/sg.rs by index (test access): 0 expm1_over_x(a),
// 1 erf(a), 2 sg_integral(a), 3 upper clamped-cosine / 2pi (a), 4 lower (a),
// 5 clamped-cosine product / pi (z=a, sharpness=b), 6 vmf axis length ->
// sharpness (a).
float orc_sg(int fn, float a, float b) {
  switch (fn) {
    case 0: return lt::expm1_over_x(a);
    case 1: return lt::erf_approx(a);
    case 2: return lt::sg_integral(a);
    case 3: return lt::upper_cc_over_2pi(a);
    case 4: return lt::lower_cc_over_2pi(a);
    case 5: return lt::sg_cc_product_over_pi(a, b);
    case 6: return lt::vmf_axis_length_to_sharpness(a);
    case 7: return lt::vmf_hemispherical_integral(a, b);
    case 8: return lt::sg_normalized_hemispherical_integral(a, b);
  }
  return NAN;
}
// Glossy-lobe pieces of R/light_tree/lobe.rs (test access).
// sggx_reflection_pdf(wi, h, m) with m = glam Mat2 columns (c0x, c0y, c1x, c1y).
float orc_sggx_reflection_pdf(const float* wi, const float* h, const float* m) {
  return lt::sggx_reflection_pdf(v3(wi[0], wi[1], wi[2]), v3(h[0], h[1], h[2]), lt::M2{m[0], m[1], m[2], m[3]});
}
// merge_glossy_roughness: out = (alpha_x, alpha_y)
void orc_merge_glossy_roughness(float ra, float axa, float aya, float rb, float axb, float ayb, float* out) {
  real ax, ay;
  lt::merge_glossy_roughness(ra, axa, aya, rb, axb, ayb, &ax, &ay);
  out[0] = ax;
  out[1] = ay;
}
// make_glossy_lobe(rho, frame = (t, b, n) in frame[9], wo, ax, ay), then
// glossy_importance against the SG light (w, axis, sharpness) with normal n;
// -1 when make_glossy_lobe gives None.
float orc_glossy_importance(float rho, const float* frame, const float* wo, float ax, float ay, const float* n,
                            float w, const float* axis, float sharpness) {
  lt::Frame3 f{v3(frame[0], frame[1], frame[2]), v3(frame[3], frame[4], frame[5]), v3(frame[6], frame[7], frame[8])};
  lt::GlossyLobe g;
  if (!lt::make_glossy_lobe(rho, f, v3(wo[0], wo[1], wo[2]), ax, ay, &g)) return -1.0f;
  lt::SgLobe l;
  l.axis = v3(axis[0], axis[1], axis[2]);
  l.sharpness = sharpness;
  l.log_amplitude = 0.0f;
  return lt::glossy_importance(g, f, v3(n[0], n[1], n[2]), w, l);
}
// sg_product: out = axis[3], sharpness, log_amplitude
// eval_thin_film_inner (R/bsdf/thin_film.rs:47-109): out = RGB reflectance
void orc_thin_film(float cos_theta_i, float eta_outside, float eta_film, const float* n, const float* k,
                   float thickness_nm, float* out) {
  V3 r = thin_film(cos_theta_i, eta_outside, eta_film, v3(n[0], n[1], n[2]), v3(k[0], k[1], k[2]), thickness_nm);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
// artist_friendly_complex_ior (R/bsdf/conductor_complex.rs:132-147): out = n[3], k[3]
void orc_artist_ior(const float* face, const float* edge, float* out) {
  V3 n, k;
  artist_friendly_complex_ior(v3(face[0], face[1], face[2]), v3(edge[0], edge[1], edge[2]), &n, &k);
  const real v[6] = {n.x, n.y, n.z, k.x, k.y, k.z};
  for (int i = 0; i < 6; ++i) out[i] = (float)v[i];
}
void orc_sg_product(const float* a1, float s1, const float* a2, float s2, float* out) {
  lt::SgLobe l = lt::sg_product(v3(a1[0], a1[1], a1[2]), s1, v3(a2[0], a2[1], a2[2]), s2);
  out[0] = l.axis.x;
  out[1] = l.axis.y;
  out[2] = l.axis.z;
  out[3] = l.sharpness;
  out[4] = l.log_amplitude;
}

}  // extern "C"
// The pixel loop of R/main.rs:98-126 (PathSampler, camera ray,
// trace_radiance, running mean, exposure) over pixels [x0, x1) x [y0, y1) of
// a w x hgt frame (every pixel's samples depend on the full frame size only),
// written row-major into a (y1-y0) x (x1-x0) crop of 3 values per pixel.
// cam: eye, forward, right, image_up, tan_half_fov_y, pixel_spread, exposure.
// ev (optional): a renderer over the same geometry whose materials and
// textures give the BSDF values (trace_radiance's detached evaluation).
template <class Out>
static void render_rect(void* h, const float* cam, uint32_t w, uint32_t hgt, uint32_t spp, uint32_t max_depth,
                        uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, Out* rgb, int threads,
                        const void* ev = nullptr) {
  const RenderScene& rs = *(RenderScene*)h;
  const RenderScene* evs = (const RenderScene*)ev;
  CF3 eye{cam[0], cam[1], cam[2]}, fwd{cam[3], cam[4], cam[5]}, right{cam[6], cam[7], cam[8]},
      up{cam[9], cam[10], cam[11]};
  const float th = cam[12], spread = cam[13], exposure = cam[14];
  const float rxf = (float)w, ryf = (float)hgt, aspect = rxf / ryf;
  const float sc = std::fmax(1.0f / std::sqrt((float)(spp < 1 ? 1 : spp)), 0.125f);  // differential_scale
  auto dir_for = [&](float px, float py) {
    float sx = ((2.0f * px / rxf) - 1.0f) * aspect * th;
    float sy = (1.0f - (2.0f * py / ryf)) * th;
    return normalize(add(add(fwd, scale(sx, right)), scale(sy, up)));
  };
  if (y1 > hgt) y1 = hgt;
  if (x1 > w) x1 = w;
  const uint32_t cw = x1 > x0 ? x1 - x0 : 0;
  const size_t n = (size_t)(y1 > y0 ? y1 - y0 : 0) * cw;
  parallel_for(n, threads, [&](size_t k) {
    const uint32_t x = x0 + (uint32_t)(k % cw), y = y0 + (uint32_t)(k / cw);
    V3 color = v3(0.f, 0.f, 0.f);
    for (uint32_t si = 0; si < spp; ++si) {
      PathSampler ps(x, y, si, spp, w, hgt);
      float u[2];
      ps.s2(0, u);
      const float px = ((float)x + 0.5f) + 0.5f * tent(u[0]);
      const float py = ((float)y + 0.5f) + 0.5f * tent(u[1]);
      // generate_ray_differential (R/scene/camera.rs:48-78)
      CF3 d = dir_for(px, py);
      CF3 rxd = normalize_or_zero(add(d, scale(sc, sub(dir_for(px + 1.0f, py), d))));
      CF3 ryd = normalize_or_zero(add(d, scale(sc, sub(dir_for(px, py + 1.0f), d))));
      Ray ray;
      ray.o = v3(eye.x, eye.y, eye.z);
      ray.d = v3(d.x, d.y, d.z);
      ray.has_diff = true;
      ray.rxo = ray.ryo = ray.o;
      ray.rxd = v3(rxd.x, rxd.y, rxd.z);
      ray.ryd = v3(ryd.x, ryd.y, ryd.z);
      ray.cone = make_cone(0.0f, spread * sc);
      V3 s = trace_radiance(rs, ray, ps, max_depth, evs);
      const float cnt = (float)(si + 1);
      color = add(color, sdiv(sub(s, color), cnt));
    }
    Out* o = rgb + 3 * ((size_t)(y - y0) * cw + (x - x0));
    o[0] = color.x * exposure;
    o[1] = color.y * exposure;
    o[2] = color.z * exposure;
  });
}
extern "C" {
void orc_render_rect(void* h, const float* cam, uint32_t w, uint32_t hgt, uint32_t spp, uint32_t max_depth,
                     uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, float* rgb, int threads) {
  render_rect(h, cam, w, hgt, spp, max_depth, x0, x1, y0, y1, rgb, threads);
}
// The same with the running means written unrounded (double): the FP64
// build's image for the gradient reference.
void orc_render_rect_f64(void* h, const float* cam, uint32_t w, uint32_t hgt, uint32_t spp, uint32_t max_depth,
                         uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, double* rgb, int threads) {
  render_rect(h, cam, w, hgt, spp, max_depth, x0, x1, y0, y1, rgb, threads);
}
// render_rect_f64 with detached evaluation: h's paths (every decision and
// density), ev's BSDF values and emission.  ev must hold the same geometry
// (the caller builds both renderers from one scene, with other materials
// or textures); a delta or thin-walled transmission sample of an OpenPBR
// material makes the pixel NaN.
void orc_render_rect_detached_f64(void* h, const void* ev, const float* cam, uint32_t w, uint32_t hgt, uint32_t spp,
                                  uint32_t max_depth, uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1,
                                  double* rgb, int threads) {
  render_rect(h, cam, w, hgt, spp, max_depth, x0, x1, y0, y1, rgb, threads, ev);
}

// Rows [y0, y1) of the frame into the full-frame buffer rgb (w * hgt * 3).
void orc_render(void* h, const float* cam, uint32_t w, uint32_t hgt, uint32_t spp, uint32_t max_depth, uint32_t y0,
                uint32_t y1, float* rgb, int threads) {
  if (y1 > hgt) y1 = hgt;
  if (y1 <= y0) return;
  orc_render_rect(h, cam, w, hgt, spp, max_depth, 0, w, y0, y1, rgb + 3 * (size_t)y0 * w, threads);
}


// ---- OpenPBR BSDF and GGX tables ---------------------------------------------------------
static OpenPbr parse_open_pbr(const OrcMaterial* m) {
  const float* f = m->raw;
  OpenPbr o{};
  o.base_weight = f[0];
  o.base_color = v3(f[1], f[2], f[3]);
  o.base_diffuse_roughness = f[4];
  o.base_metalness = f[5];
  o.specular_weight = f[6];
  o.specular_color = v3(f[7], f[8], f[9]);
  o.specular_roughness = f[10];
  o.specular_ior = f[11];
  o.specular_roughness_anisotropy = f[12];
  o.transmission_weight = f[13];
  o.coat_weight = f[14];
  o.fuzz_weight = f[15];
  o.thin_film_weight = f[16];
  o.emission_luminance = f[17];
  o.emission_color = v3(f[18], f[19], f[20]);
  uint32_t tw;
  std::memcpy(&tw, &f[21], 4);
  o.thin_walled = tw;
  o.coat_color = v3(f[22], f[23], f[24]);
  o.coat_roughness = f[25];
  o.coat_roughness_anisotropy = f[26];
  o.coat_ior = f[27];
  o.coat_darkening = f[28];
  o.fuzz_color = v3(f[29], f[30], f[31]);
  o.fuzz_roughness = f[32];
  std::memcpy(o.tex, &f[33], sizeof(o.tex));
  o.transmission_color = v3(f[39], f[40], f[41]);
  o.transmission_dispersion_scale = f[42];
  o.transmission_dispersion_abbe_number = f[43];
  o.thin_film_thickness = f[44];
  o.thin_film_ior = f[45];
  return o;
}
// Table sizes: albedo (per eta), conductor e / e_avg, dielectric e / e_avg.
void orc_lut_sizes(uint32_t* out) {
  out[0] = kAlbedoLen;
  out[1] = kCondCos * kCondRough;
  out[2] = kCondRough;
  out[3] = kDielCos * kDielRough * kDielEta;
  out[4] = kDielRough * kDielEta;
}
static void copy_out(const std::vector<real>& v, float* out) {
  for (size_t i = 0; i < v.size(); ++i) out[i] = (float)v[i];
}
void orc_lut_albedo(float eta, float* out) { copy_out(albedo_lut_for(eta).v, out); }
void orc_lut_cond(float* e, float* eavg) {
  const CondLut& l = cond_lut();
  copy_out(l.e, e);
  copy_out(l.eavg, eavg);
}
void orc_lut_diel(float* e, float* eavg) {
  const DielLut& l = diel_lut();
  copy_out(l.e, e);
  copy_out(l.eavg, eavg);
}
// which: 0 albedo(eta; w, roughness), 1 conductor e(cos, r), 2 conductor e_avg(r),
// 3 dielectric e(cos, r, eta), 4 dielectric e_avg(r, eta)
float orc_lut_lookup(int which, float eta, const float* w, float a, float b) {
  switch (which) {
    case 0: return albedo_lookup(albedo_lut_for(eta), v3(w[0], w[1], w[2]), a);
    case 1: return cond_lookup_e(cond_lut(), a, b);
    case 2: return cond_lookup_eavg(cond_lut(), a);
    case 3: return diel_lookup_e(diel_lut(), a, b, eta);
    default: return diel_lookup_eavg(diel_lut(), a, eta);
  }
}
// f (3 per direction, radiance mode, no shading-normal correction) and pdf of
// OpenPbrMaterial::make_bsdf(...).eval / pdf in the shading frame, for n
// incident directions.
void orc_openpbr_eval(const OrcMaterial* m, int front_face, const float* wo, const float* wi, size_t n, float* f,
                      float* pdf) {
  const OpenPbr o = parse_open_pbr(m);
  const V3 w = v3(wo[0], wo[1], wo[2]);
  const PbrBsdf p = make_bsdf(o, front_face != 0, w, luts_for(o));
  for (size_t i = 0; i < n; ++i) {
    const V3 d = v3(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
    const V3 e = bsdf_eval(p, w, d);
    f[3 * i] = e.x;
    f[3 * i + 1] = e.y;
    f[3 * i + 2] = e.z;
    pdf[i] = bsdf_pdf(p, w, d);
  }
}
// OpenPbrBsdf::sample for n (u_lobe, u0, u1) triples; out per sample:
// ok, weight[3], wi[3], pdf, delta, transmission, eta (11 floats).
void orc_openpbr_sample(const OrcMaterial* m, int front_face, const float* wo, const float* u, size_t n, float* out) {
  const OpenPbr o = parse_open_pbr(m);
  const V3 w = v3(wo[0], wo[1], wo[2]);
  const PbrBsdf p = make_bsdf(o, front_face != 0, w, luts_for(o));
  for (size_t i = 0; i < n; ++i) {
    BSample s{};
    const bool ok = bsdf_sample(p, w, u[3 * i], u[3 * i + 1], u[3 * i + 2], &s);
    float* r = out + 11 * i;
    r[0] = ok ? 1.0f : 0.0f;
    r[1] = ok ? s.weight.x : 0.0f;
    r[2] = ok ? s.weight.y : 0.0f;
    r[3] = ok ? s.weight.z : 0.0f;
    r[4] = ok ? s.wi.x : 0.0f;
    r[5] = ok ? s.wi.y : 0.0f;
    r[6] = ok ? s.wi.z : 0.0f;
    r[7] = ok ? s.pdf : 0.0f;
    r[8] = ok && s.delta ? 1.0f : 0.0f;
    r[9] = ok && s.transmission ? 1.0f : 0.0f;
    r[10] = ok ? s.eta : 0.0f;
  }
}
// Pieces of the BSDF the reference's own tests check: 0 metal Fresnel at cos
// (3), 1 metal Fresnel average (3), 2 emission exit tint at wo (3), 3 coat
// darkening factor (3), 4 dielectric ms params (eta_o, ratio_r, ome_o) for
// (alpha = a, eta = b), 5 / 6 thin-wall reflectance / transmittance at
// cos_o = a (3), 7 thin-wall transmission alphas (2).
void orc_openpbr_probe(const OrcMaterial* m, int which, const float* wo, float a, float b, float* out) {
  const OpenPbr o = parse_open_pbr(m);
  const V3 w = v3(wo[0], wo[1], wo[2]);
  const PbrBsdf p = make_bsdf(o, true, w, luts_for(o));
  V3 r = vsplat(0.0f);
  if (which == 0) r = metal_fresnel(p, a);
  if (which == 1) r = metal_fresnel_avg(p);
  if (which == 2) r = emission_exit_tint(o, w);
  if (which == 3) r = p.coat_darkening;
  if (which == 4) {
    PbrBsdf q = p;
    q.luts.diel = &diel_lut();
    const MsParams s = ms_params(q, a, a, b);
    r = v3(s.eta_o, s.ratio_r, s.ome_o);
  }
  if (which == 5 || which == 6) {
    V3 rr, tt;
    thin_wall_coefficients(p, a, &rr, &tt);
    r = which == 5 ? rr : tt;
  }
  if (which == 7) {
    real ax, ay;
    thin_wall_alpha(p, &ax, &ay);
    r = v3(ax, ay, 0.0f);
  }
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}


// ---- ray differentials and cones (test entry points) ------------------------------------
// Ray record of the two entries below, 21 floats: origin[3], dir[3], has
// differential (0 / 1), rx origin[3], ry origin[3], rx dir[3], ry dir[3],
// cone width, cone spread.
static Ray ray_from(const float* r) {
  Ray ray;
  ray.o = v3(r[0], r[1], r[2]);
  ray.d = v3(r[3], r[4], r[5]);
  ray.has_diff = r[6] != 0.0f;
  ray.rxo = v3(r[7], r[8], r[9]);
  ray.ryo = v3(r[10], r[11], r[12]);
  ray.rxd = v3(r[13], r[14], r[15]);
  ray.ryd = v3(r[16], r[17], r[18]);
  ray.cone = make_cone(r[19], r[20]);
  return ray;
}
static void ray_to(const Ray& ray, float* r) {
  const V3 v[6] = {ray.o, ray.d, ray.rxo, ray.ryo, ray.rxd, ray.ryd};
  r[0] = v[0].x; r[1] = v[0].y; r[2] = v[0].z;
  r[3] = v[1].x; r[4] = v[1].y; r[5] = v[1].z;
  r[6] = ray.has_diff ? 1.0f : 0.0f;
  for (int k = 0; k < 4; ++k) {
    r[7 + 3 * k] = v[2 + k].x;
    r[8 + 3 * k] = v[2 + k].y;
    r[9 + 3 * k] = v[2 + k].z;
  }
  r[19] = ray.cone.width;
  r[20] = ray.cone.spread;
}
// compute_shading_differentials (R/scene/mod.rs:748-785 with the cone
// fallback :821-859) at point p with face-forwarded geometric normal n and
// position partials dpdu, dpdv for a ray (21 floats).  out: dudx, dvdx,
// dudy, dvdy, dpdx[3], dpdy[3].
void orc_shading_differentials(const float* p, const float* n, const float* dpdu, const float* dpdv,
                               const float* ray, float* out) {
  Vertex v;
  v.p = v3(p[0], p[1], p[2]);
  const Ray r = ray_from(ray);
  shading_differentials(v, v3(n[0], n[1], n[2]), v3(dpdu[0], dpdu[1], dpdu[2]), v3(dpdv[0], dpdv[1], dpdv[2]), &r);
  out[0] = v.dudx; out[1] = v.dvdx; out[2] = v.dudy; out[3] = v.dvdy;
  out[4] = v.dpdx.x; out[5] = v.dpdx.y; out[6] = v.dpdx.z;
  out[7] = v.dpdy.x; out[8] = v.dpdy.y; out[9] = v.dpdy.z;
}
// spawn_scattered_ray (R/integrator/mod.rs:58-75) with the specular ray
// differential of a delta sample (:88-157).  vtx: p, ng, ns, wo, dpdx, dpdy,
// dndu, dndv (3 floats each), dudx, dvdx, dudy, dvdy; the sample's delta
// flag, cone spread and direction wi; out: the spawned ray (21 floats).
void orc_spawn_scattered(const float* ray, float hit_t, const float* vtx, int delta, float cone_spread,
                         const float* wi, float* out) {
  Vertex v;
  V3* f[8] = {&v.p, &v.ng, &v.ns, &v.wo, &v.dpdx, &v.dpdy, &v.dndu, &v.dndv};
  for (int k = 0; k < 8; ++k) *f[k] = v3(vtx[3 * k], vtx[3 * k + 1], vtx[3 * k + 2]);
  v.dudx = vtx[24]; v.dvdx = vtx[25]; v.dudy = vtx[26]; v.dvdy = vtx[27];
  BSample s;
  s.delta = delta != 0;
  s.cone_spread = cone_spread;
  ray_to(spawn_scattered(ray_from(ray), hit_t, v, s, v3(wi[0], wi[1], wi[2])), out);
}

}  // extern "C"

// ===========================================================================
// Textures (texture.h): the reference's Texture / ScalarTexture.
extern "C" {

void* orc_texture_new(uint32_t w, uint32_t h, int channels, const float* px) {
  return new texture::Texture(texture::from_pixels(w, h, channels, px));
}
void orc_texture_free(void* t) { delete (texture::Texture*)t; }
uint32_t orc_texture_levels(void* t) { return (uint32_t)((texture::Texture*)t)->levels.size(); }
float orc_texture_max_value(void* t) { return ((texture::Texture*)t)->max_value; }
void orc_texture_level_size(void* t, uint32_t level, uint32_t* w, uint32_t* h) {
  const texture::Level& l = ((texture::Texture*)t)->levels[level];
  *w = (uint32_t)l.width;
  *h = (uint32_t)l.height;
}
void orc_texture_level(void* t, uint32_t level, float* out) {
  const texture::Level& l = ((texture::Texture*)t)->levels[level];
  std::copy(l.px.begin(), l.px.end(), out);
}
// Lookups: q = (u, v, dstdx.x, dstdx.y, dstdy.x, dstdy.y) per query; mode as
// TPT_TEXTURE_*: 0 filtered (EWA), 1 bilinear, 2 nearest, 3 trilinear,
// 4 mip-bilinear.  out: channels floats per query.
void orc_texture_sample(void* th, int mode, const float* q, size_t n, float* out) {
  const texture::Texture& t = *(texture::Texture*)th;
  for (size_t i = 0; i < n; ++i) {
    const float* a = q + 6 * i;
    texture::Px p;
    switch (mode) {
      case 1: p = texture::sample(t, a[0], a[1]); break;
      case 2: p = texture::sample_nearest(t, a[0], a[1]); break;
      case 3: p = texture::sample_trilinear(t, a[0], a[1], a[2], a[3], a[4], a[5]); break;
      case 4: p = texture::sample_mip_bilinear(t, a[0], a[1], a[2], a[3], a[4], a[5]); break;
      default: p = texture::sample_filtered(t, a[0], a[1], a[2], a[3], a[4], a[5]); break;
    }
    for (int k = 0; k < t.channels; ++k) out[t.channels * i + k] = p.c[k];
  }
}

}  // extern "C"
