/*
 * tpt.h — C ABI of the MI355X ray-query backend (libtpt.so).
 *
 * The backend replaces the per-ray scene queries of the reference renderer
 * (crates/renderer/src = R/):
 *   traverse_qbvh                R/qbvh.rs:611-675
 *   Scene::closest_hit           R/scene/mod.rs:340-395
 *   closest_mesh_hit_with_filter R/scene/mod.rs:686-723
 *   intersect_triangle           R/math/ray.rs:107-209
 *   unoccluded_ray               R/integrator/mod.rs:239-265
 *   build_qbvh (host)            R/qbvh.rs:144-166
 *
 * Conventions follow the reference's own C shim (crates/ocio-sys/src/shim.h):
 * opaque handles; create functions return NULL on failure; status functions
 * return int32_t (0 = ok); a thread-local last-error string; no exception or
 * panic crosses the boundary.  Inputs are caller-owned and copied (scene) or
 * only read during the call; outputs are caller-allocated.
 */
#ifndef TPT_H
#define TPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TPT_OK 0
#define TPT_ERROR 1              /* generic error, see tpt_last_error()          */
#define TPT_ERROR_NOT_BUILT 2    /* ClosestHitError::QbvhNotBuilt, R/scene/mod.rs:100-113 */
#define TPT_ERROR_BAD_ARGUMENT 3 /* null pointer, index out of range, bad layout  */
#define TPT_ERROR_DEVICE 4       /* HIP runtime error                            */
#define TPT_ERROR_OVERFLOW 5     /* leaf offset/count beyond the 24/7-bit child encoding (R/qbvh.rs:15-33) */

/* ---- per-call flags ---------------------------------------------------- */
/* Pointers passed to a trace call are device (HBM) pointers; the call only
 * enqueues work on the device stream and returns without synchronising.
 * Without this flag pointers are host pointers and the call is synchronous. */
#define TPT_FLAG_DEVICE_POINTERS 0x1u
/* Traverse the 8-bit quantised copy of the QBVH nodes instead of the FP32
 * nodes.  Child boxes are rounded outward, so the set of candidate
 * triangles only grows; the closest hit is unchanged except on exact ties. */
#define TPT_FLAG_QUANTIZED 0x2u
/* tpt_trace_any and the wavefront's shadow rays: answer with a full
 * closest-hit query followed by the reference's decision rule (bit-exact
 * with unoccluded_ray, no early exit). */
#define TPT_FLAG_ANY_EXACT 0x4u
/* tpt_render_adjoint: accumulate the per-block gradient sums in FP64. */
#define TPT_FLAG_GRAD_FP64 0x8u

/* ---- records ------------------------------------------------------------ */
/* Ray: origin, t_max, direction (R/math/ray.rs:5-11 without differentials).
 * The reference always queries with t_max = +inf (R/scene/mod.rs:352). */
typedef struct TptRay {
  float o[3];
  float tmax;
  float d[3];
  float pad;
} TptRay; /* 32 bytes */

/* SceneHit (R/scene/mod.rs:74-79): TriangleRef = (instance, triangle),
 * t, barycentric (b0, b1, b2).  A miss has instance = triangle = 0xFFFFFFFF. */
typedef struct TptHit {
  uint32_t instance;
  uint32_t triangle;
  float t;
  float b[3];
} TptHit; /* 24 bytes */

#define TPT_INVALID_INDEX 0xFFFFFFFFu

/* QbvhNode in the reference layout (R/qbvh.rs:54-64): four child boxes as
 * structure-of-arrays, child references (leaf bit | 7-bit count | 24-bit
 * offset, or interior node index), and the valid-slot mask. */
typedef struct TptQbvhNode {
  float min_x[4], min_y[4], min_z[4];
  float max_x[4], max_y[4], max_z[4];
  uint32_t children[4];
  uint8_t valid_mask;
  uint8_t pad[15];
} TptQbvhNode; /* 128 bytes */

/* One mesh (R/scene/mesh.rs:51-57): vertex positions (xyz), triangle indices,
 * and its built QBVH (nodes may be empty: root-leaf case, R/qbvh.rs:615-620). */
typedef struct TptMeshDesc {
  const float* positions;           /* 3 * vertex_count */
  uint32_t vertex_count;
  uint32_t triangle_count;
  const uint32_t* indices;          /* 3 * triangle_count */
  const TptQbvhNode* nodes;         /* node_count entries, may be NULL if 0 */
  uint32_t node_count;
  uint32_t pad;
  const uint32_t* primitive_indices;/* triangle_count entries (Qbvh::primitive_indices) */
  const float* normals;             /* 3 * vertex_count, NULL: no shading normals (Vertex::normal) */
  const float* uvs;                 /* 2 * vertex_count, NULL: all zero (Vertex::uv) */
} TptMeshDesc;

/* Instance (R/scene/mod.rs:62-70): 4x4 matrices in glam's column-major order. */
typedef struct TptInstanceDesc {
  uint32_t mesh;
  uint32_t material;
  float local_to_world[16];
  float world_to_local[16];
} TptInstanceDesc;

/* OpenPbrMaterial parameters (R/material/open_pbr.rs:23-82) as the renderer
 * holds them after prepare_color (rendering-space colours): base, metal,
 * specular, transmission (the specular BTDF, rough or smooth, and thin
 * walls), coat, fuzz, thin film (thin_film_weight, thin_film_thickness,
 * thin_film_ior; no thickness texture) and emission.  Materials with
 * dispersion (transmission_dispersion_scale > 0) are rejected with
 * TPT_ERROR_BAD_ARGUMENT (normal maps, opacity, subsurface
 * and the interior medium are not part of the record; the reference
 * forces the last two off, validate_and_warn :303-321).  The fields after
 * emission_color_texture carry transmission_color (default 1),
 * transmission_dispersion_scale (0), transmission_dispersion_abbe_number
 * (20), thin_film_thickness (0.5, micrometres) and thin_film_ior (1.4).  The coat and fuzz
 * fields follow the reference's field order after thin_walled; the
 * reference's defaults (OpenPbrMaterial::new) are coat_color 1,
 * coat_roughness 0, coat_ior 1.6, coat_darkening 1, fuzz_color 1,
 * fuzz_roughness 0.5.  The *_texture fields index the scene's texture
 * table (TptSceneDesc::textures) or are TPT_INVALID_INDEX for none: the
 * reference's base_color_texture, base_metalness_texture,
 * specular_roughness_texture, fuzz_weight_texture, fuzz_roughness_texture
 * and emission_color_texture (R/material/open_pbr.rs:66-72), each
 * multiplying its parameter by the filtered texel (RGB textures for the
 * two colours, scalar textures for the rest; *_at, :670-790).  Carried
 * inside TptMaterial. */
typedef struct TptOpenPbr {
  float base_weight;
  float base_color[3];
  float base_diffuse_roughness;
  float base_metalness;
  float specular_weight;
  float specular_color[3];
  float specular_roughness;
  float specular_ior;
  float specular_roughness_anisotropy;
  float transmission_weight;
  float coat_weight;
  float fuzz_weight;
  float thin_film_weight;
  float emission_luminance;
  float emission_color[3];
  uint32_t thin_walled;
  float coat_color[3];
  float coat_roughness;
  float coat_roughness_anisotropy;
  float coat_ior;
  float coat_darkening;
  float fuzz_color[3];
  float fuzz_roughness;
  uint32_t base_color_texture;
  uint32_t base_metalness_texture;
  uint32_t specular_roughness_texture;
  uint32_t fuzz_weight_texture;
  uint32_t fuzz_roughness_texture;
  uint32_t emission_color_texture;
  float transmission_color[3];
  float transmission_dispersion_scale;
  float transmission_dispersion_abbe_number;
  float thin_film_thickness;
  float thin_film_ior;
  uint32_t pad[2];
} TptOpenPbr; /* 192 bytes */

/* One entry of the scene's material table: the Material enum of
 * R/material/mod.rs:100-115 for the variants the backend carries.
 *   TPT_MATERIAL_OPEN_PBR           open_pbr (TptOpenPbr)
 *   TPT_MATERIAL_NORMALIZED_LAMBERT rho[3]              (R/material/normalized_lambert.rs:18-37)
 *   TPT_MATERIAL_EMISSIVE           color[3], strength  (R/material/emissive.rs:9-26)
 *   TPT_MATERIAL_MIRROR             color[3]            (R/material/mirror.rs)
 * Normal maps and opacity are not part of the record; OpenPBR carries its
 * textures as indices into the scene's texture table. */
#define TPT_MATERIAL_OPEN_PBR 0u
#define TPT_MATERIAL_NORMALIZED_LAMBERT 1u
#define TPT_MATERIAL_EMISSIVE 2u
#define TPT_MATERIAL_MIRROR 3u
typedef struct TptMaterial {
  uint32_t kind;
  uint32_t pad;
  union {
    TptOpenPbr open_pbr;
    struct { float rho[3]; } lambert;
    struct { float color[3]; float strength; } emissive;
    struct { float color[3]; } mirror;
    float raw[48];
  } u;
} TptMaterial; /* 200 bytes */

/* A filtered, mip-mapped image (Texture<Vec3> or ScalarTexture,
 * R/material/texture.rs:53-126): width * height texels of `channels` floats
 * (3: RGB, 1: scalar), row-major, row 0 first; texel (x, y) covers
 * uv [x/w, (x+1)/w) x [y/h, (y+1)/h).  tpt_scene_create copies level 0 to
 * the device and builds the box-filtered pyramid there (build_mip_levels,
 * :433-466). */
typedef struct TptTextureDesc {
  const float* pixels;
  uint32_t width;
  uint32_t height;
  uint32_t channels;
  uint32_t pad;
} TptTextureDesc; /* 24 bytes */

/* A directional light as DirectionalLight::new leaves it
 * (R/light/directional.rs:15-32): the unit direction the light travels
 * along, its RGB tint and its perpendicular irradiance; sample_li gives
 * radiance color * intensity from -direction with pdf 1 (a delta light). */
typedef struct TptDirectionalLight {
  float direction[3];
  float intensity;
  float color[3];
  float pad;
} TptDirectionalLight; /* 32 bytes */

/* Scene after Scene::build_qbvh (R/scene/mod.rs:305-316). */
typedef struct TptSceneDesc {
  const TptMeshDesc* meshes;
  uint32_t mesh_count;
  uint32_t instance_count;
  const TptInstanceDesc* instances;
  const TptQbvhNode* tlas_nodes;          /* over instance world bounds */
  uint32_t tlas_node_count;
  uint32_t material_count;
  const uint32_t* tlas_primitive_indices; /* instance_count entries */
  const TptMaterial* materials;           /* material_count entries (TptInstanceDesc::material) */
  /* Environment light (EnvironmentLight::from_pixels, R/light/environment.rs:47-83):
   * a lat-long RGB map, env_width * env_height * 3 floats row-major (v = 0 is
   * +Y), radiance scale and rotation about +Y in radians; NULL: none. */
  const float* env_pixels;
  uint32_t env_width;
  uint32_t env_height;
  float env_scale;
  float env_rotate_y;
  /* Texture table indexed by the materials' *_texture fields (may be NULL
   * with texture_count 0). */
  const TptTextureDesc* textures;
  uint32_t texture_count;
  /* Directional lights (Scene::add_directional_light, R/scene/mod.rs:277-281),
   * each its own light-sampler category after the light tree and the
   * environment (LightSampler::build_from_scene, R/light/mod.rs:133-191). */
  uint32_t directional_count;
  const TptDirectionalLight* directional_lights;
} TptSceneDesc;

/* A node of the scene's SG light tree (R/light_tree/mod.rs:52-88), as
 * tpt_scene_create builds it from the emissive triangles
 * (Scene::build_light_tree, R/light_tree/builder.rs:58-121): depth-first
 * pre-order, root 0; left/right/parent TPT_INVALID_INDEX where absent;
 * instance/triangle name the leaf's triangle; area is the leaf triangle's
 * area (0 for interior nodes). */
typedef struct TptLightNode {
  float flux;
  float mu[3];
  float sigma_s2;
  float nu[3];
  float lambda;
  float radius;
  float area;
  uint32_t left, right, parent;
  uint32_t instance, triangle;
  uint32_t leaf;
  uint32_t pad[3];
} TptLightNode; /* 80 bytes */

/* Pinhole camera after PinholeCamera::new (R/scene/camera.rs:21-42): the
 * renderer's normalised basis and tan(fov_y / 2), plus pixel_spread =
 * pixel_spread_angle(resolution) (R/scene/camera.rs:89-96), both evaluated
 * with the host's libm so the device does no transcendental work. */
typedef struct TptCamera {
  float eye[3];
  float tan_half_fov_y;
  float forward[3];
  float pixel_spread;
  float right[3];
  float exposure;
  float image_up[3];
  float pad;
} TptCamera; /* 64 bytes */

typedef struct TptDevice TptDevice;
typedef struct TptScene TptScene;
typedef struct TptQbvh TptQbvh;

/* ---- errors -------------------------------------------------------------- */
/* Last error of the calling thread (shim.h:14-15 convention); never NULL. */
const char* tpt_last_error(void);

/* ---- device -------------------------------------------------------------- */
TptDevice* tpt_device_create(int32_t device_index);
void tpt_device_release(TptDevice* device);
/* Stream that trace calls are enqueued on (a hipStream_t; NULL = the
 * library's own stream).  Lets a caller time or order work with its events. */
int32_t tpt_device_set_stream(TptDevice* device, void* hip_stream);
int32_t tpt_device_synchronize(TptDevice* device);
/* Tuning knobs of the trace kernels (no reference counterpart; set by an
 * autotuner or the caller):
 *   "refill_below"  1..64  a persistent wave fetches new rays for its idle
 *                          lanes when fewer than this many lanes are busy
 *   "blocks_per_cu" 0..64  cap on resident 256-thread blocks per CU (0: the
 *                          occupancy limit)
 *   "lds_depth"     4..32  traversal stack entries per lane kept in LDS (the
 *                          rest, up to the reference's 128, in scratch;
 *                          default 16; tpt_trace_noalloc keeps 16)
 * and of the wavefront path tracer (tpt_render_batch):
 *   "wf_paths"  65536..2^26  paths in flight per pass (default 2^21)
 *   "wf_graph"  0 or 1       capture one pass as a HIP graph and replay it
 *   "wf_tile"   128, 256, 512  the shade queue's tile: threads per generate /
 *                            shade block, one queue entry each (default 256)
 *   "wf_sort"   0..64        shade each bounce's hits grouped by material:
 *                            0 in queue order, g sorted by material index / g
 *                            (a counting sort of the hit queue into at most
 *                            256 keys; the image does not change) */
int32_t tpt_device_set_param(TptDevice* device, const char* name, int64_t value);

/* ---- sampler and camera --------------------------------------------------- */
/* Primary rays of one sample index for every pixel of a width x height frame
 * (row-major): PathSampler::camera_sample (R/sampler/path_sampler.rs:89-91)
 * then PinholeCamera::generate_ray_differential (R/scene/camera.rs:48-78).
 * rays: width*height TptRay (t_max = +inf).  differentials (nullable): 7
 * floats per pixel, rx_direction[3], ry_direction[3], cone spread angle. */
int32_t tpt_camera_rays(TptDevice* device, const TptCamera* camera, uint32_t width, uint32_t height,
                        uint32_t samples_per_pixel, uint32_t sample_index, TptRay* rays, float* differentials,
                        uint32_t flags);
/* The integrator's random numbers for one sample index, every pixel:
 * PathSampler::camera_sample, initial_aux_rng_seed and path_vertex_randoms
 * (R/sampler/path_sampler.rs:89-131) for depths 0..depth_count-1.
 * values: per pixel 2 + 14*depth_count floats (camera u,v; per depth
 * u_category, u_tree, u_light_aux, u_surface.xy, u_lobe, u_layer, u_dir.xy,
 * u_extra0..3, u_rr); seeds: per pixel 1 + depth_count aux-RNG seeds. */
int32_t tpt_path_randoms(TptDevice* device, uint32_t width, uint32_t height, uint32_t samples_per_pixel,
                         uint32_t sample_index, uint32_t depth_count, float* values, uint32_t* seeds,
                         uint32_t flags);

/* ---- host QBVH build (R/qbvh.rs:144-516) -------------------------------- */
/* Binned-SAH 4-wide build over primitive bounds (3 floats per min / max).
 * Returns NULL when n == 0 (build_qbvh returns None) or on error.
 * threads <= 0 uses the hardware concurrency; the tree does not depend on it. */
TptQbvh* tpt_qbvh_build(const float* bounds_min, const float* bounds_max, size_t n, int32_t threads);

/* build_qbvh (R/qbvh.rs:144-516) on the device of `device`: the same nodes
 * and primitive order as tpt_qbvh_build (and the reference), from the
 * primitives' bounds (3 floats each) in host memory, or in device memory
 * with TPT_FLAG_DEVICE_POINTERS.  Ranges of more than 2048 primitives are
 * split level by level over blocks, smaller ones by one block each.  NULL
 * with tpt_last_error for no primitives, more than 2^31 - 1, a leaf beyond
 * the 24-bit child encoding, or a forced median split whose centroids do not
 * coincide (the reference then sorts them; the host build reproduces it). */
TptQbvh* tpt_qbvh_build_device(TptDevice* device, const float* bounds_min, const float* bounds_max, size_t n,
                               uint32_t flags);
void tpt_qbvh_release(TptQbvh* qbvh);
/* The scene's 8-bit quantised nodes as the device holds them (64 bytes each,
 * the layout tpt_qbvh_quantize writes; BLAS nodes of every mesh in mesh
 * order, interior children rebased into that array, or with tlas != 0 the
 * TLAS nodes): *count receives the node count, and with out non-NULL and
 * capacity >= *count the nodes are copied.  tpt_scene_create encodes them on
 * the device from the FP32 nodes with the encoder tpt_qbvh_quantize runs on
 * the host.  No reference counterpart (an inspection call). */
int32_t tpt_scene_q8_nodes(const TptScene* scene, uint32_t tlas, void* out, size_t capacity, size_t* count);
size_t tpt_qbvh_node_count(const TptQbvh* qbvh);
size_t tpt_qbvh_primitive_count(const TptQbvh* qbvh);
/* Copies the nodes (node_count entries) and primitive_indices out. */
int32_t tpt_qbvh_export(const TptQbvh* qbvh, TptQbvhNode* nodes, uint32_t* primitive_indices);

/* 8-bit quantised copy of n nodes (64 bytes each) as the device traverses
 * them: per axis an FP32 origin and a power-of-two scale, child corners
 * origin + q * 2^e rounded outward so each decoded box contains the FP32
 * box.  No reference counterpart (R/qbvh.rs stores FP32 boxes only); exposed
 * so a caller can cache or inspect the encoding. */
int32_t tpt_qbvh_quantize(const TptQbvhNode* nodes, size_t n, void* out_q8_nodes);

/* Host build of the SG light tree of a scene description (the same tree
 * tpt_scene_create builds; no device needed): Scene::build_light_tree
 * (R/scene/mod.rs:300-303, R/light_tree/builder.rs:58-121) over the emissive
 * triangles in instance order.  At most `capacity` nodes are written;
 * *count gets the total (0: no emissive triangle). */
int32_t tpt_light_tree_build(const TptSceneDesc* desc, TptLightNode* out, uint32_t capacity, uint32_t* count);

/* ---- scene ---------------------------------------------------------------- */
/* Copies the scene to the device and derives the GPU layouts (leaf-ordered
 * triangle records, FP32 and 8-bit quantised node arrays), and builds the
 * SG light tree over the emissive triangles as the reference's
 * Scene::build_light_tree does (R/scene/mod.rs:300-303; area-light table in
 * instance order, R/light_tree/builder.rs:58-121). */
TptScene* tpt_scene_create(TptDevice* device, const TptSceneDesc* desc);

/* The scene description as an on-disk cache (no reference counterpart; the
 * reference rebuilds its scene each run, R/main.rs:60-97): save writes every
 * array of desc (meshes with their built QBVHs, instances, TLAS, materials,
 * environment, textures, directional lights) to one file with a format
 * version, the record sizes and a checksum; load reads it back into a
 * descriptor that tpt_scene_create accepts, owning one allocation that
 * tpt_scene_desc_free releases.  Load returns NULL (tpt_last_error says
 * why) for a missing, truncated or corrupt file, another format version or
 * other record layouts. */
int32_t tpt_scene_desc_save(const TptSceneDesc* desc, const char* path);
TptSceneDesc* tpt_scene_desc_load(const char* path);
void tpt_scene_desc_free(TptSceneDesc* desc);
void tpt_scene_release(TptScene* scene);
/* Device bytes held by the scene (nodes, quantised nodes, triangles, instances). */
size_t tpt_scene_device_bytes(const TptScene* scene);
/* Copies the scene's light tree out (at most `capacity` nodes; *count gets
 * the total).  Replaces reading Scene::light_tree (R/light_tree/mod.rs:98-121). */
int32_t tpt_scene_light_tree(const TptScene* scene, TptLightNode* out, uint32_t capacity, uint32_t* count);

/* ---- queries -------------------------------------------------------------- */
/* Closest hit for n rays: Scene::closest_hit (R/scene/mod.rs:340-395) with the
 * per-ray t_max as the initial bound. */
int32_t tpt_trace_closest(TptScene* scene, const TptRay* rays, size_t n, TptHit* hits, uint32_t flags);

/* Shadow-ray visibility, unoccluded_ray (R/integrator/mod.rs:239-265) for rays
 * already offset by spawn_ray.  target: per ray (instance, triangle) of the
 * light triangle, or instance = TPT_INVALID_INDEX for none; NULL = no targets.
 * distance: per ray light distance (+inf for infinite lights); NULL = +inf.
 * out[i] = 1 if unoccluded.  Without TPT_FLAG_ANY_EXACT this is an early-exit
 * any-hit query whose answer differs from the reference only on exact ties. */
int32_t tpt_trace_any(TptScene* scene, const TptRay* rays, size_t n, const uint32_t* target,
                      const float* distance, uint8_t* unoccluded, uint32_t flags);

/* ---- path tracer ------------------------------------------------------------ */
typedef struct TptRenderDesc {
  TptCamera camera;
  uint32_t width;
  uint32_t height;
  uint32_t samples_per_pixel;
  uint32_t max_depth;
} TptRenderDesc;

/* The render loop of R/main.rs:98-126 with IntegratorKind::Mis
 * (R/integrator/mis.rs:19-139): for every pixel, samples 0..spp-1 through
 * PathSampler, generate_ray_differential and trace_radiance, accumulated as
 * the reference's running mean, times the camera exposure.  rgb: width *
 * height * 3 floats, row-major.  Light NEE follows the reference policy: the
 * light sampler picks the light-tree or the environment category by power
 * (LightSampler, R/light/mod.rs); a NormalizedLambert shading point samples the SG
 * light tree (shadow ray, balance heuristic); an OpenPBR shading point gets
 * no light-tree sample (R/light_tree/importance.rs:26-36,
 * R/material/open_pbr.rs:592-606) unless the scene's policy is
 * TPT_NEE_OPENPBR_PROXY (tpt_scene_set_nee_policy); every material gets
 * environment samples. */
int32_t tpt_render(TptScene* scene, const TptRenderDesc* desc, float* rgb, uint32_t flags);

/* Gradient of an image loss with respect to every NormalizedLambert
 * material's albedo (rho) and every OpenPBR material's base_color, by
 * path-replay backpropagation: each pixel sample's path is traced twice from
 * its sampler state (primal radiance, then a replay that attributes the
 * radiance collected from each vertex onward to that vertex's parameters);
 * no path is stored.  Sampling decisions, MIS weights and Russian roulette
 * probabilities are held fixed.  OpenPBR materials are differentiated where
 * their BSDF is affine in the base colour (base_metalness 0 and no coat
 * darkening: coat_weight or coat_darkening 0); other OpenPBR rows are zero.
 * dloss_dimage: width * height * 3 floats (dLoss/d of the tpt_render image,
 * exposure included); d_albedo: 3 * material_count floats (zero for other
 * materials).  Sums are FP32 per block (LDS), then FP64 over blocks;
 * TPT_FLAG_GRAD_FP64 makes the per-block sums FP64.  No reference
 * counterpart (the reference renderer is not differentiable). */
int32_t tpt_render_adjoint(TptScene* scene, const TptRenderDesc* desc, const float* dloss_dimage, float* d_albedo,
                           uint32_t flags);
/* As tpt_render_adjoint, and also the gradient with respect to the texels of
 * the OpenPBR base-colour textures: d_texels gets the level-0 texels of every
 * texture of the scene in table order (width * height * channels floats
 * each; zero for textures that no differentiated material reads as its base
 * colour).  The lookups' gradients are scattered through the EWA weights
 * into every mip level and then through the box filter down to level 0.  No
 * reference counterpart. */
int32_t tpt_render_adjoint_texels(TptScene* scene, const TptRenderDesc* desc, const float* dloss_dimage,
                                  float* d_albedo, float* d_texels, uint32_t flags);

/* Path-replay adjoint with respect to every OpenPBR material's base colour
 * (RGB), specular_roughness and base_metalness: d_params holds 5 floats per
 * material (material_count * 5; NormalizedLambert rows: albedo, 0, 0).  A
 * vertex's contributions are scaled by (df/dtheta) / f of its BSDF at the
 * light and sampled directions, df/dtheta by a central difference of step
 * 1e-3 of the BSDF rebuilt at theta +- step (one-sided at the [0, 1] ends of
 * roughness and metalness and at base colour 0); the base colour of a
 * material whose BSDF is affine in it takes the exact ratio as in
 * tpt_render_adjoint.  Sampling, Russian roulette and MIS weights are held
 * fixed.  A roughness or metalness slot with a texture gets no gradient (the
 * base-colour texture does, into d_texels as in tpt_render_adjoint_texels;
 * d_texels may be NULL).  At most 2048 materials.  The parameters of the
 * configuration's inverse-rendering workload (R/material/open_pbr.rs:23-896
 * parameter meaning); no reference counterpart. */
int32_t tpt_render_adjoint_params(TptScene* scene, const TptRenderDesc* desc, const float* dloss_dimage,
                                  float* d_params, float* d_texels, uint32_t flags);

/* One view of a batched adjoint: its render description (camera,
 * resolution, samples per pixel, depth) and dLoss/d(image) of that view
 * (width * height * 3 floats; host or device memory per the call's flags). */
typedef struct TptAdjointFrameDesc {
  TptRenderDesc render;
  const float* dloss_dimage;
} TptAdjointFrameDesc;

/* tpt_render_adjoint_params for many views of uneven resolution in one
 * launch: d_params (material_count * 5 floats) and d_texels (nullable, as in
 * tpt_render_adjoint_texels) receive the sum over the views of each view's
 * gradient.  Every pixel of every view is a thread; the block's gradient
 * rows accumulate per sample index in FP32 (LDS) and are folded into FP32
 * compensated sums (TwoSum) after every sample index, the blocks' sums and
 * compensations then added in FP64; texel gradients are FP32 atomics.  With
 * TPT_FLAG_GRAD_FP64 the block sums and texel gradients are FP64 instead (the
 * reference for the FP32 accumulation's error).  At most 1024 materials and
 * 2^31 pixels per batch.  No reference counterpart (the pixel loop of
 * R/main.rs:98-126 per view, differentiated). */
int32_t tpt_render_adjoint_batch(TptScene* scene, const TptAdjointFrameDesc* frames, uint32_t frame_count,
                                 float* d_params, float* d_texels, uint32_t flags);

/* Replaces the scene's material table (same count; kinds and emission must
 * stay the same, as the light tree is built from them): the parameter update
 * of an inverse-rendering loop.  Synchronous. */
int32_t tpt_scene_set_materials(TptScene* scene, const TptMaterial* materials, uint32_t count);

/* ---- next-event-estimation policy ------------------------------------------ */
/* What light-tree query an OpenPBR shading point builds
 * (LightTreePrecompute, R/light_tree/lobe.rs:20-64):
 *   TPT_NEE_REFERENCE      none, as the reference (R/material/open_pbr.rs:
 *                          592-606): no light-tree NEE sample, and emitters
 *                          found by BSDF sampling get MIS weight 1;
 *   TPT_NEE_OPENPBR_PROXY  a diffuse lobe and one merged glossy lobe, built
 *                          as StandardSurface's precompute builds its own
 *                          (R/material/standard_surface.rs:599-698) from the
 *                          OpenPBR parameters.  An extension: the estimator
 *                          stays unbiased (the NEE pmf and the reverse pmf of
 *                          BSDF-sampled hits use the same query); only its
 *                          variance changes.
 * NormalizedLambert builds its diffuse query under both. */
#define TPT_NEE_REFERENCE 0u
#define TPT_NEE_OPENPBR_PROXY 1u

/* Selects the scene's NEE policy for every later render, batch and adjoint
 * call (default TPT_NEE_REFERENCE).  No reference counterpart: the reference
 * has the first policy only. */
int32_t tpt_scene_set_nee_policy(TptScene* scene, uint32_t policy);

/* One frame (view) of a batched render: its own camera, resolution, samples
 * per pixel and depth, and its output (width * height * 3 floats; host or
 * device memory per the call's flags). */
typedef struct TptFrameDesc {
  TptRenderDesc render;
  float* rgb;
} TptFrameDesc;

/* The same estimator and output as tpt_render for every frame of a batch,
 * run as a wavefront path tracer: all frames' paths share one set of
 * launches (per bounce: a ray-queue trace, a shade kernel with light-tree
 * NEE appending to shadow and next-ray queues, a shadow-queue trace), so
 * frames of uneven resolution fill the GPU together.  Shadow rays use the
 * early-exit any-hit query (the light sample's own t bounds it); with
 * TPT_FLAG_ANY_EXACT they use the reference's closest-hit rule and each
 * frame's image is the one tpt_render gives for it, bit for bit, with FP32
 * nodes (the two rules differ only when another triangle is hit at exactly
 * the light sample's t).  TPT_FLAG_QUANTIZED selects 8-bit nodes for every
 * trace.  Tuned by
 * tpt_device_set_param "wf_paths" (paths in flight) and "wf_graph" (capture
 * a pass as a HIP graph).  No reference counterpart beyond the pixel loop of
 * R/main.rs:98-126, which it runs for many frames at once. */
int32_t tpt_render_batch(TptScene* scene, const TptFrameDesc* frames, uint32_t frame_count, uint32_t flags);

/* Autotuner of tpt_render_batch for a workload: coordinate descent over the
 * device's "wf_paths", "blocks_per_cu", "refill_below", "lds_depth" (the
 * traversal stack's entries in LDS), "wf_tile" (the shade queue's tile:
 * threads per generate / shade block) and "wf_sort" (the material sort's
 * granularity: 0 off, g materials per sort key) parameters, each
 * candidate timed with HIP events on the given frames (samples per pixel
 * capped at tune_spp, 0: as given; outputs go to an internal buffer).  A
 * candidate replaces the current setting only when at least 1% faster, and
 * paths-in-flight candidates above the tuning batch's path count are not
 * tried (give tune_spp enough samples that the batch holds the frames' pass
 * structure).  The fastest setting stays on the device.  No reference
 * counterpart. */
typedef struct TptTuneResult {
  uint32_t wf_paths;
  uint32_t blocks_per_cu;
  uint32_t refill_below;
  uint32_t trials;
  uint32_t wf_sort;
  float ms_default;  /* the starting setting */
  float ms_best;
  uint32_t lds_depth;
  uint32_t wf_tile;
} TptTuneResult;
int32_t tpt_autotune_render(TptScene* scene, const TptFrameDesc* frames, uint32_t frame_count, uint32_t flags,
                            uint32_t tune_spp, TptTuneResult* out);

/* ---- material BSDF queries ------------------------------------------------------- */
/* One query of a material's BSDF in the shading frame (z = shading normal):
 * the outgoing direction wo, an incident direction wi for eval and pdf, and
 * the randoms u = (u_lobe, u0, u1) for a sample (MaterialSampleRandoms'
 * u_lobe and u_dir); front_face as the shading vertex has it. */
typedef struct TptBsdfQuery {
  float wo[3];
  float wi[3];
  float u[3];
  uint32_t material;
  uint32_t front_face;
  uint32_t pad;
} TptBsdfQuery; /* 48 bytes */
typedef struct TptBsdfResult {
  float f[3];       /* eval(wo, wi), radiance mode, before the shading-normal correction */
  float pdf;        /* pdf(wo, wi) */
  float weight[3];  /* sample(wo, u): f |cos| / pdf, or the delta lobe's weight */
  float sample_pdf; /* its pdf (the lobe probability for a delta sample) */
  float wi[3];      /* its direction */
  uint32_t flags;   /* bit 0: a sample was produced; bit 1: it is a delta sample; bit 2: it refracts */
} TptBsdfResult; /* 48 bytes */
/* The BSDF the path tracer builds for a material (Material::eval / pdf /
 * sample, R/material/mod.rs:251-370, at the BSDF level: OpenPbrBsdf
 * R/bsdf/open_pbr.rs:224-331 after OpenPbrMaterial::make_bsdf
 * R/material/open_pbr.rs:393-487, NormalizedLambertBsdf, MirrorBsdf), on the
 * device, for n queries.  Host or device pointers per flags. */
int32_t tpt_material_bsdf(TptScene* scene, const TptBsdfQuery* queries, size_t n, TptBsdfResult* out, uint32_t flags);

/* ---- textures ------------------------------------------------------------------ */
/* One texture lookup: uv and the screen-space derivatives of uv (the
 * shading vertex's uv_dx() / uv_dy(), R/material/mod.rs:163-169). */
typedef struct TptTextureQuery {
  float uv[2];
  float dstdx[2];
  float dstdy[2];
} TptTextureQuery; /* 24 bytes */
/* Lookup modes, the Texture methods of R/material/texture.rs:
 *   TPT_TEXTURE_FILTERED     sample_filtered: EWA over the mip pyramid
 *                            (alpha 2, anisotropy at most 8), :140-193, :245-300
 *   TPT_TEXTURE_BILINEAR     sample: bilinear on level 0, :130-132, :225-243
 *   TPT_TEXTURE_NEAREST      sample_nearest, :134-142
 *   TPT_TEXTURE_TRILINEAR    sample_trilinear, :200-223
 *   TPT_TEXTURE_MIP_BILINEAR sample_mip_bilinear, :147-159
 * The renderer uses TPT_TEXTURE_FILTERED (the materials' *_at fetches). */
#define TPT_TEXTURE_FILTERED 0u
#define TPT_TEXTURE_BILINEAR 1u
#define TPT_TEXTURE_NEAREST 2u
#define TPT_TEXTURE_TRILINEAR 3u
#define TPT_TEXTURE_MIP_BILINEAR 4u
/* n lookups of one texture of the scene on the device in the given mode;
 * out gets `channels` floats per query.  Host or device pointers per flags. */
int32_t tpt_texture_sample(TptScene* scene, uint32_t texture, uint32_t mode, const TptTextureQuery* queries, size_t n,
                           float* out, uint32_t flags);
/* Shape of a texture's pyramid: level count (Texture::level_count), channel
 * count and the level-0 maximum (Texture::max_value, :121-124). */
int32_t tpt_texture_info(const TptScene* scene, uint32_t texture, uint32_t* levels, uint32_t* channels,
                         float* max_value);
/* Replaces a texture's level-0 texels (same size and channels; host or
 * device pointer per flags) and rebuilds its pyramid on the device: the
 * texel update of an inverse-rendering loop.  An emission texture is
 * rejected (the light tree is built from its maximum).  Synchronous.  No
 * reference counterpart (the reference's textures are immutable). */
int32_t tpt_scene_set_texture(TptScene* scene, uint32_t texture, const float* pixels, uint32_t flags);
/* Copies mip level `level` of a texture to host memory: width * height *
 * channels floats (capacity in floats; *width / *height get its size).
 * Replaces reading TextureLevel::pixels (R/material/texture.rs:62-67). */
int32_t tpt_texture_level(TptScene* scene, uint32_t texture, uint32_t level, float* out, size_t capacity,
                          uint32_t* width, uint32_t* height);

/* The traversal of tpt_trace_closest (query 0) or tpt_trace_any (query 1)
 * with per-ray work counters instead of results: counts[3*i+0] nodes
 * visited, [3*i+1] triangles tested, [3*i+2] instances entered.  For query 0
 * the counts equal those of the reference traversal (same visit order); they
 * give the algorithmic bytes per ray of the roofline.  No reference
 * counterpart (the reference keeps no counters). */
int32_t tpt_trace_counts(TptScene* scene, uint32_t query, const TptRay* rays, size_t n, const uint32_t* target,
                         const float* distance, uint32_t* counts, uint32_t flags);

/* ---- allocation-free queries ------------------------------------------------ */
/* tpt_trace_closest (query 0: hits) or tpt_trace_any (query 1: unoccluded,
 * with optional target / distance, TPT_FLAG_ANY_EXACT honoured) for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  No counterpart in the
 * reference.  Device pointers only (flags must hold TPT_FLAG_DEVICE_POINTERS;
 * TPT_FLAG_QUANTIZED selects 8-bit nodes).  The trace kernel's scratch (its
 * work counters and status word, 16,896 bytes) is a static device array that
 * each GPU holds until the process ends; the call clears it with a kernel,
 * launches the trace on the device's stream and returns without
 * synchronising or allocating.  At most 2^31 - 1 rays per call.  Calls that
 * overlap on one GPU share that array and race; a traversal-stack overflow
 * is not reported (as with device pointers in tpt_trace_closest). */
int32_t tpt_trace_noalloc(TptScene* scene, uint32_t query, const TptRay* rays, size_t n, const uint32_t* target,
                          const float* distance, TptHit* hits, uint8_t* unoccluded, uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* TPT_H */
