"""Host-side mirror of the reference scene API for the ray-query path.

Mirrors the geometry part of ``Scene`` (crates/renderer/src/scene/mod.rs):
``add_mesh``, ``add_instance`` (R/scene/mod.rs:219-252), ``build_qbvh``
(R/scene/mod.rs:305-316) and ``closest_hit`` (R/scene/mod.rs:340-395, here
batched), plus the shadow query ``unoccluded`` (R/integrator/mod.rs:239-265).
Queries run on the GPU through libtpt; the QBVH build runs in libtpt's host
builder (R/qbvh.rs:144-516).  Matrices are given in the usual math
convention (column vectors, ``p' = M @ p``) and handed to the ABI in glam's
column-major order.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from enum import Enum

import numpy as np

from . import _lib
from ._lib import HIT_DTYPE, INVALID_INDEX, NODE_DTYPE, Q8_NODE_DTYPE, RAY_DTYPE, check, lib, ptr


class ClosestHitError(Exception):
    """ClosestHitError (R/scene/mod.rs:100-113)."""

    class Kind(Enum):
        QbvhNotBuilt = "scene QBVH has not been built yet"

    def __init__(self, kind: "ClosestHitError.Kind"):
        super().__init__(kind.value)
        self.kind = kind


@dataclass
class Qbvh:
    """Qbvh (R/qbvh.rs:129-133): nodes in the reference layout + primitive_indices."""

    nodes: np.ndarray  # NODE_DTYPE
    primitive_indices: np.ndarray  # uint32


def build_qbvh(bounds_min: np.ndarray, bounds_max: np.ndarray, threads: int = 0,
               device: "Device | None" = None) -> Qbvh | None:
    """build_qbvh (R/qbvh.rs:144-166): None for no primitives.  With a
    device the build runs on that GPU (tpt_qbvh_build_device; the same tree)."""
    bmin = np.ascontiguousarray(bounds_min, dtype=np.float32).reshape(-1, 3)
    bmax = np.ascontiguousarray(bounds_max, dtype=np.float32).reshape(-1, 3)
    n = bmin.shape[0]
    if n == 0:
        return None
    if device is not None:
        h = lib().tpt_qbvh_build_device(device.handle, ptr(bmin), ptr(bmax), n, 0)
    else:
        h = lib().tpt_qbvh_build(ptr(bmin), ptr(bmax), n, threads)
    if not h:
        raise _lib.TptError(_lib.TPT_ERROR, _lib.last_error())
    try:
        nn = lib().tpt_qbvh_node_count(h)
        nodes = np.zeros(nn, dtype=NODE_DTYPE)
        prim = np.zeros(n, dtype=np.uint32)
        check(lib().tpt_qbvh_export(h, ptr(nodes) if nn else None, ptr(prim)))
    finally:
        lib().tpt_qbvh_release(h)
    return Qbvh(nodes, prim)


def quantize_nodes(nodes: np.ndarray) -> np.ndarray:
    """8-bit quantised copies of reference-layout nodes, as traversed on the GPU."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    out = np.zeros(len(nodes), dtype=Q8_NODE_DTYPE)
    if len(nodes):
        check(lib().tpt_qbvh_quantize(ptr(nodes), len(nodes), ptr(out)))
    return out


@dataclass
class Mesh:
    """Mesh geometry (R/scene/mesh.rs:51-57): positions (V,3) and indices (T,3)."""

    positions: np.ndarray
    indices: np.ndarray
    normals: np.ndarray | None = None
    uvs: np.ndarray | None = None
    qbvh: Qbvh | None = None

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 3)
        if len(self.positions) == 0:
            raise ValueError("mesh must contain at least one vertex")  # Mesh::new expect

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        """compute_bounds (R/scene/mesh.rs:668-680)."""
        return self.positions.min(axis=0), self.positions.max(axis=0)

    def triangle_count(self) -> int:
        return len(self.indices)

    def triangle_bounds(self) -> tuple[np.ndarray, np.ndarray]:
        """Mesh::triangle_bounds (R/scene/mesh.rs:104-111), all triangles."""
        v = self.positions[self.indices]  # (T,3,3)
        lo = np.fmin(np.fmin(v[:, 0], v[:, 1]), v[:, 2])
        hi = np.fmax(np.fmax(v[:, 0], v[:, 1]), v[:, 2])
        return lo, hi

    def build_qbvh(self, threads: int = 0, device: "Device | None" = None) -> None:
        """Mesh::build_qbvh (R/scene/mesh.rs:114-119); on `device` when given."""
        lo, hi = self.triangle_bounds()
        self.qbvh = build_qbvh(lo, hi, threads, device)


def glam_cols(m: np.ndarray) -> np.ndarray:
    """A 4x4 math-convention matrix as glam's column-major 16 floats."""
    return np.ascontiguousarray(np.asarray(m, dtype=np.float32).T).reshape(16)


def transform_point3(m: np.ndarray, p: np.ndarray) -> np.ndarray:
    """glam Mat4::transform_point3 in f32: ((c0*x + c1*y) + c2*z) + c3."""
    m = np.asarray(m, dtype=np.float32)
    p = np.asarray(p, dtype=np.float32)
    r = m[:3, 0] * p[..., 0:1]
    r = m[:3, 1] * p[..., 1:2] + r
    r = m[:3, 2] * p[..., 2:3] + r
    return m[:3, 3] + r


def transform_bounds(lo: np.ndarray, hi: np.ndarray, m: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """transform_bounds (R/scene/mod.rs:869-893): box of the 8 transformed corners."""
    corners = np.array(
        [[lo[0], lo[1], lo[2]], [lo[0], lo[1], hi[2]], [lo[0], hi[1], lo[2]], [lo[0], hi[1], hi[2]],
         [hi[0], lo[1], lo[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], lo[2]], [hi[0], hi[1], hi[2]]],
        dtype=np.float32,
    )
    t = transform_point3(m, corners)
    return np.fmin.reduce(t, axis=0), np.fmax.reduce(t, axis=0)


@dataclass
class Instance:
    """Instance (R/scene/mod.rs:62-70)."""

    mesh_index: int
    material_index: int
    local_to_world: np.ndarray
    world_to_local: np.ndarray
    world_bounds: tuple[np.ndarray, np.ndarray]


@dataclass
class Environment:
    """EnvironmentLight::from_pixels (R/light/environment.rs:47-83): a
    lat-long RGB map (height, width, 3), radiance scale and rotation about +Y
    (radians)."""

    pixels: np.ndarray
    scale: float = 1.0
    rotate_y: float = 0.0


@dataclass
class DirectionalLight:
    """DirectionalLight (R/light/directional.rs:15-32): the direction the
    light travels along (normalised as DirectionalLight::new does, in f32:
    v * (1 / sqrt(dot))), RGB tint and perpendicular irradiance."""

    direction: tuple
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0

    @staticmethod
    def new(direction, color, intensity: float) -> "DirectionalLight":
        d = np.asarray(direction, np.float32)
        dot = np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])
        dot = np.float32(dot + np.float32(d[2] * d[2]))
        inv = np.float32(np.float32(1.0) / np.sqrt(dot, dtype=np.float32))
        return DirectionalLight(tuple(float(np.float32(x * inv)) for x in d), tuple(float(c) for c in color),
                                float(intensity))


@dataclass
class Scene:
    """Geometry of the reference Scene (R/scene/mod.rs:81-98)."""

    meshes: list[Mesh] = field(default_factory=list)
    instances: list[Instance] = field(default_factory=list)
    qbvh: Qbvh | None = None
    built: bool = False
    materials: list = field(default_factory=list)  # tpt.material records, by material index
    environment: Environment | None = None  # Scene::set_environment_light
    textures: list = field(default_factory=list)  # tpt.texture.Texture, indexed by the materials' *_texture
    directional_lights: list = field(default_factory=list)  # DirectionalLight

    def add_directional_light(self, light: "DirectionalLight") -> int:
        """Scene::add_directional_light (R/scene/mod.rs:277-281) -> its index."""
        self.directional_lights.append(light)
        return len(self.directional_lights) - 1

    def add_texture(self, texture) -> int:
        """Adds a texture to the scene's table -> its index (what an OpenPbr
        *_texture field takes; the reference holds Arc<Texture> in the
        material, R/material/open_pbr.rs:66-72)."""
        self.textures.append(texture)
        return len(self.textures) - 1

    def set_environment_light(self, env: Environment) -> None:
        self.environment = Environment(np.ascontiguousarray(env.pixels, np.float32), float(env.scale),
                                       float(env.rotate_y))

    def add_material(self, material) -> int:
        """Scene::add_material (R/scene/mod.rs:143-217) for OpenPBR."""
        self.materials.append(material)
        return len(self.materials) - 1

    def add_mesh(self, mesh: Mesh) -> int:
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_instance(self, mesh_index: int, material_index: int, local_to_world: np.ndarray,
                     world_to_local: np.ndarray | None = None) -> int:
        """Scene::add_instance (R/scene/mod.rs:219-252).  The inverse is taken
        in f64 and rounded to f32 unless the caller passes it (the renderer's
        glam inverse crosses the ABI as data)."""
        l2w = np.asarray(local_to_world, dtype=np.float32)
        if world_to_local is None:
            world_to_local = np.linalg.inv(l2w.astype(np.float64)).astype(np.float32)
        lo, hi = self.meshes[mesh_index].bounds
        wb = transform_bounds(lo, hi, l2w)
        self.instances.append(Instance(mesh_index, material_index, l2w, np.asarray(world_to_local, np.float32), wb))
        self.qbvh = None
        self.built = False
        return len(self.instances) - 1

    def build_qbvh(self, threads: int = 0, device: "Device | None" = None) -> None:
        """Scene::build_qbvh (R/scene/mod.rs:305-316); the meshes' and the
        instances' QBVHs built on `device` when given (the same trees)."""
        for m in self.meshes:
            m.build_qbvh(threads, device)
        if self.instances:
            lo = np.stack([i.world_bounds[0] for i in self.instances])
            hi = np.stack([i.world_bounds[1] for i in self.instances])
            self.qbvh = build_qbvh(lo, hi, threads, device)
        else:
            self.qbvh = None
        self.built = True

    def upload(self, device: "Device") -> "DeviceScene":
        if not self.built:
            raise ClosestHitError(ClosestHitError.Kind.QbvhNotBuilt)
        return DeviceScene(device, self)

    def save_cache(self, path) -> None:
        """Writes the built scene (meshes with their QBVHs, instances, TLAS,
        materials, environment, textures, directional lights) to one file
        (tpt_scene_desc_save), which DeviceScene.from_cache uploads without a
        build."""
        if not self.built:
            raise ClosestHitError(ClosestHitError.Kind.QbvhNotBuilt)
        desc, keep = scene_desc(self)
        check(lib().tpt_scene_desc_save(ctypes.byref(desc), os.fsencode(os.fspath(path))))
        del keep


class Device:
    """A GPU (tpt_device_create); trace calls go on its stream."""

    def __init__(self, index: int = 0):
        self.handle = lib().tpt_device_create(index)
        if not self.handle:
            raise _lib.TptError(_lib.TPT_ERROR_DEVICE, _lib.last_error())
        self.index = index

    def set_stream(self, stream_ptr: int | None) -> None:
        check(lib().tpt_device_set_stream(self.handle, stream_ptr))

    def set_param(self, name: str, value: int) -> None:
        """Trace-kernel tuning knob (tpt_device_set_param)."""
        check(lib().tpt_device_set_param(self.handle, name.encode(), int(value)))

    def synchronize(self) -> None:
        check(lib().tpt_device_synchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            lib().tpt_device_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scene_desc(scene: Scene):
    """TptSceneDesc of a built scene -> (desc, keep): `keep` holds the arrays
    the descriptor points into and must outlive every call that reads it."""
    keep = []
    meshes = (_lib.MeshDesc * max(1, len(scene.meshes)))()
    for i, m in enumerate(scene.meshes):
        q = m.qbvh
        d = meshes[i]
        d.positions = ptr(m.positions)
        d.vertex_count = len(m.positions)
        d.triangle_count = len(m.indices)
        d.indices = ptr(m.indices)
        if q is not None:
            d.nodes = ptr(q.nodes) if len(q.nodes) else None
            d.node_count = len(q.nodes)
            d.primitive_indices = ptr(q.primitive_indices)
            keep.append(q)
        else:
            d.node_count = 0
            d.primitive_indices = None
        if m.normals is not None:
            m.normals = np.ascontiguousarray(m.normals, dtype=np.float32).reshape(-1, 3)
            d.normals = ptr(m.normals)
        if m.uvs is not None:
            m.uvs = np.ascontiguousarray(m.uvs, dtype=np.float32).reshape(-1, 2)
            d.uvs = ptr(m.uvs)
    insts = (_lib.InstanceDesc * max(1, len(scene.instances)))()
    for i, inst in enumerate(scene.instances):
        insts[i].mesh = inst.mesh_index
        insts[i].material = inst.material_index
        insts[i].local_to_world[:] = glam_cols(inst.local_to_world).tolist()
        insts[i].world_to_local[:] = glam_cols(inst.world_to_local).tolist()
    desc = _lib.SceneDesc()
    desc.meshes = ctypes.addressof(meshes)
    desc.mesh_count = len(scene.meshes)
    desc.instance_count = len(scene.instances)
    desc.instances = ctypes.addressof(insts)
    if scene.qbvh is not None:
        desc.tlas_nodes = ptr(scene.qbvh.nodes) if len(scene.qbvh.nodes) else None
        desc.tlas_node_count = len(scene.qbvh.nodes)
        desc.tlas_primitive_indices = ptr(scene.qbvh.primitive_indices)
    if scene.materials:
        from .material import records
        mats = records(scene.materials)
        keep.append(mats)
        desc.materials = ctypes.addressof(mats)
        desc.material_count = len(scene.materials)
    if scene.environment is not None:
        px = np.ascontiguousarray(scene.environment.pixels, np.float32)
        keep.append(px)
        desc.env_pixels = ptr(px)
        desc.env_height, desc.env_width = px.shape[0], px.shape[1]
        desc.env_scale = scene.environment.scale
        desc.env_rotate_y = scene.environment.rotate_y
    if scene.directional_lights:
        dl = np.zeros(len(scene.directional_lights), _lib.DIRECTIONAL_DTYPE)
        for i, l in enumerate(scene.directional_lights):
            dl[i]["direction"], dl[i]["color"], dl[i]["intensity"] = l.direction, l.color, l.intensity
        keep.append(dl)
        desc.directional_lights = dl.ctypes.data
        desc.directional_count = len(dl)
    if scene.textures:
        texs = (_lib.TextureDesc * len(scene.textures))()
        for i, t in enumerate(scene.textures):
            px = np.ascontiguousarray(t.pixels, np.float32)
            keep.append(px)
            texs[i].pixels = ptr(px)
            texs[i].height, texs[i].width = px.shape[0], px.shape[1]
            texs[i].channels = 3 if px.ndim == 3 else 1
        keep.append(texs)
        desc.textures = ctypes.addressof(texs)
        desc.texture_count = len(scene.textures)
    keep += [meshes, insts]
    return desc, keep


def build_light_tree(scene: Scene) -> np.ndarray:
    """Scene::build_light_tree on the host (tpt_light_tree_build): the SG light
    tree over the emissive triangles, LIGHT_NODE_DTYPE records."""
    desc, keep = scene_desc(scene)
    count = ctypes.c_uint32(0)
    check(lib().tpt_light_tree_build(ctypes.byref(desc), None, 0, ctypes.byref(count)))
    out = np.zeros(count.value, dtype=_lib.LIGHT_NODE_DTYPE)
    if len(out):
        check(lib().tpt_light_tree_build(ctypes.byref(desc), ptr(out), count.value, ctypes.byref(count)))
    del keep
    return out


class DeviceScene:
    """A scene resident in device memory (tpt_scene_create)."""

    def __init__(self, device: Device, scene: Scene):
        self.device = device
        desc, keep = scene_desc(scene)
        self.handle = lib().tpt_scene_create(device.handle, ctypes.byref(desc))
        if not self.handle:
            raise _lib.TptError(_lib.TPT_ERROR, _lib.last_error())
        self.instance_count = len(scene.instances)
        self.material_count = len(scene.materials)
        # the textures' (height, width[, channels]) shapes: the gradient outputs'
        self.texture_shapes = [tuple(t.pixels.shape) for t in scene.textures]

    @classmethod
    def from_cache(cls, device: Device, path) -> "DeviceScene":
        """Uploads a scene saved by Scene.save_cache (tpt_scene_desc_load +
        tpt_scene_create): no mesh or QBVH build, the file holds the built
        arrays.  Raises TptError for a missing, truncated or corrupt file or
        one written with another format version or record layout."""
        L = lib()
        d = L.tpt_scene_desc_load(os.fsencode(os.fspath(path)))
        if not d:
            raise _lib.TptError(_lib.TPT_ERROR, _lib.last_error())
        try:
            desc = _lib.SceneDesc.from_address(d)
            self = cls.__new__(cls)
            self.device = device
            self.handle = L.tpt_scene_create(device.handle, d)
            if not self.handle:
                raise _lib.TptError(_lib.TPT_ERROR, _lib.last_error())
            self.instance_count = desc.instance_count
            self.material_count = desc.material_count
            shapes = []
            if desc.texture_count:
                tex = (_lib.TextureDesc * desc.texture_count).from_address(desc.textures)
                for t in tex:
                    shapes.append((t.height, t.width) if t.channels == 1 else (t.height, t.width, t.channels))
            self.texture_shapes = shapes
        finally:
            L.tpt_scene_desc_free(d)
        return self

    @property
    def device_bytes(self) -> int:
        return lib().tpt_scene_device_bytes(self.handle)

    def light_tree(self) -> np.ndarray:
        """The scene's SG light tree (tpt_scene_light_tree) as LIGHT_NODE_DTYPE."""
        count = ctypes.c_uint32(0)
        check(lib().tpt_scene_light_tree(self.handle, None, 0, ctypes.byref(count)))
        out = np.zeros(count.value, dtype=_lib.LIGHT_NODE_DTYPE)
        check(lib().tpt_scene_light_tree(self.handle, ptr(out) if len(out) else None, count.value,
                                         ctypes.byref(count)))
        return out

    def close(self) -> None:
        if self.handle:
            lib().tpt_scene_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # host-array queries (synchronous) ------------------------------------------
    def closest_hit(self, rays: np.ndarray, quantized: bool = False) -> np.ndarray:
        """Scene::closest_hit for a batch of RAY_DTYPE rays -> HIT_DTYPE hits."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(len(rays), dtype=HIT_DTYPE)
        flags = _lib.TPT_FLAG_QUANTIZED if quantized else 0
        check(lib().tpt_trace_closest(self.handle, ptr(rays), len(rays), ptr(hits), flags))
        return hits

    def unoccluded(self, rays: np.ndarray, target: np.ndarray | None = None, distance: np.ndarray | None = None,
                   quantized: bool = False, exact: bool = False) -> np.ndarray:
        """unoccluded_ray for shadow rays -> bool array."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        out = np.zeros(len(rays), dtype=np.uint8)
        if target is not None:
            target = np.ascontiguousarray(target, dtype=np.uint32).reshape(-1, 2)
        if distance is not None:
            distance = np.ascontiguousarray(distance, dtype=np.float32)
        flags = (_lib.TPT_FLAG_QUANTIZED if quantized else 0) | (_lib.TPT_FLAG_ANY_EXACT if exact else 0)
        check(lib().tpt_trace_any(self.handle, ptr(rays), len(rays), ptr(target), ptr(distance), ptr(out), flags))
        return out.astype(bool)

    def counts(self, rays: np.ndarray, query: str = "closest", target: np.ndarray | None = None,
               distance: np.ndarray | None = None, quantized: bool = False, exact: bool = False) -> np.ndarray:
        """Per-ray (nodes visited, triangles tested, instances entered) of a query."""
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        out = np.zeros((len(rays), 3), dtype=np.uint32)
        if target is not None:
            target = np.ascontiguousarray(target, dtype=np.uint32).reshape(-1, 2)
        if distance is not None:
            distance = np.ascontiguousarray(distance, dtype=np.float32)
        flags = (_lib.TPT_FLAG_QUANTIZED if quantized else 0) | (_lib.TPT_FLAG_ANY_EXACT if exact else 0)
        q = {"closest": 0, "any": 1}[query]
        check(lib().tpt_trace_counts(self.handle, q, ptr(rays), len(rays), ptr(target), ptr(distance), ptr(out),
                                     flags))
        return out

    def render(self, camera, width: int, height: int, spp: int, max_depth: int) -> np.ndarray:
        """The reference render loop (R/main.rs:98-126) with the MIS integrator
        -> (height, width, 3) float32, exposure applied."""
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        out = np.zeros((height, width, 3), np.float32)
        check(lib().tpt_render(self.handle, ctypes.byref(d), ptr(out), 0))
        return out

    def render_device(self, camera, width: int, height: int, spp: int, max_depth: int, out_dev) -> None:
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        check(lib().tpt_render(self.handle, ctypes.byref(d), ptr(out_dev), _lib.TPT_FLAG_DEVICE_POINTERS))

    def set_materials(self, materials) -> None:
        """tpt_scene_set_materials: a new material table of the same kinds and
        emission (e.g. updated albedos of an inverse-rendering step)."""
        from .material import records
        recs = records(materials)
        check(lib().tpt_scene_set_materials(self.handle, ctypes.addressof(recs), len(materials)))

    NEE_POLICIES = {"reference": 0, "openpbr_proxy": 1}

    def set_nee_policy(self, policy: str) -> None:
        """tpt_scene_set_nee_policy: "reference" (OpenPBR shading points build
        no light-tree query, as the reference) or "openpbr_proxy" (they build
        a diffuse + merged-glossy query, so emissive triangles get light-tree
        NEE from OpenPBR surfaces).  Applies to every later render, batch and
        adjoint call on this scene."""
        check(lib().tpt_scene_set_nee_policy(self.handle, self.NEE_POLICIES[policy]))

    def render_adjoint(self, camera, width: int, height: int, spp: int, max_depth: int, dloss_dimage,
                       fp64: bool = False) -> np.ndarray:
        """Path-replay adjoint (tpt_render_adjoint): dLoss/d(albedo) of every
        NormalizedLambert material -> (material_count, 3) float32."""
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        g = np.ascontiguousarray(dloss_dimage, dtype=np.float32).reshape(height, width, 3)
        out = np.zeros((self.material_count, 3), np.float32)
        check(lib().tpt_render_adjoint(self.handle, ctypes.byref(d), ptr(g), ptr(out),
                                       _lib.TPT_FLAG_GRAD_FP64 if fp64 else 0))
        return out

    def render_adjoint_texels(self, camera, width: int, height: int, spp: int, max_depth: int, dloss_dimage,
                              fp64: bool = False):
        """tpt_render_adjoint_texels -> (d_albedo (material_count, 3), [per
        texture level-0 gradient, shaped like its pixels])."""
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        g = np.ascontiguousarray(dloss_dimage, dtype=np.float32).reshape(height, width, 3)
        out = np.zeros((self.material_count, 3), np.float32)
        shapes = self.texture_shapes
        flat = np.zeros(max(1, sum(int(np.prod(sh)) for sh in shapes)), np.float32)
        check(lib().tpt_render_adjoint_texels(self.handle, ctypes.byref(d), ptr(g), ptr(out), ptr(flat),
                                              _lib.TPT_FLAG_GRAD_FP64 if fp64 else 0))
        grads, o = [], 0
        for sh in shapes:
            k = int(np.prod(sh))
            grads.append(flat[o:o + k].reshape(sh).copy())
            o += k
        return out, grads

    def render_adjoint_params(self, camera, width: int, height: int, spp: int, max_depth: int, dloss_dimage,
                              fp64: bool = False):
        """tpt_render_adjoint_params -> (d_params (material_count, 5): base
        colour RGB, specular_roughness, base_metalness per material; [per
        texture level-0 gradient, shaped like its pixels])."""
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        g = np.ascontiguousarray(dloss_dimage, dtype=np.float32).reshape(height, width, 3)
        out = np.zeros((self.material_count, 5), np.float32)
        shapes = self.texture_shapes
        flat = np.zeros(max(1, sum(int(np.prod(sh)) for sh in shapes)), np.float32)
        check(lib().tpt_render_adjoint_params(self.handle, ctypes.byref(d), ptr(g), ptr(out),
                                              ptr(flat) if shapes else None,
                                              _lib.TPT_FLAG_GRAD_FP64 if fp64 else 0))
        grads, o = [], 0
        for sh in shapes:
            k = int(np.prod(sh))
            grads.append(flat[o:o + k].reshape(sh).copy())
            o += k
        return out, grads

    def render_adjoint_device(self, camera, width: int, height: int, spp: int, max_depth: int, dloss_dev,
                              d_albedo_dev, fp64: bool = False) -> None:
        """As render_adjoint with device buffers (torch tensors): dloss
        (h*w*3), d_albedo (material_count*3); asynchronous."""
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        flags = _lib.TPT_FLAG_DEVICE_POINTERS | (_lib.TPT_FLAG_GRAD_FP64 if fp64 else 0)
        check(lib().tpt_render_adjoint(self.handle, ctypes.byref(d), ptr(dloss_dev), ptr(d_albedo_dev), flags))

    def render_adjoint_texels_device(self, camera, width: int, height: int, spp: int, max_depth: int, dloss_dev,
                                     d_albedo_dev, d_texels_dev, fp64: bool = False) -> None:
        """As render_adjoint_texels with device buffers (torch tensors):
        dloss (h*w*3), d_albedo (material_count*3), d_texels (every
        texture's level-0 texels packed)."""
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        flags = _lib.TPT_FLAG_DEVICE_POINTERS | (_lib.TPT_FLAG_GRAD_FP64 if fp64 else 0)
        check(lib().tpt_render_adjoint_texels(self.handle, ctypes.byref(d), ptr(dloss_dev), ptr(d_albedo_dev),
                                              ptr(d_texels_dev), flags))

    def render_adjoint_params_device(self, camera, width: int, height: int, spp: int, max_depth: int, dloss_dev,
                                     d_params_dev, d_texels_dev=None, fp64: bool = False) -> None:
        """As render_adjoint_params with device buffers (torch tensors):
        dloss (h*w*3), d_params (material_count*5), d_texels (optional)."""
        d = _lib.RenderDesc()
        d.camera = camera
        d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
        flags = _lib.TPT_FLAG_DEVICE_POINTERS | (_lib.TPT_FLAG_GRAD_FP64 if fp64 else 0)
        check(lib().tpt_render_adjoint_params(self.handle, ctypes.byref(d), ptr(dloss_dev), ptr(d_params_dev),
                                              ptr(d_texels_dev) if d_texels_dev is not None else None, flags))

    def set_texture(self, texture: int, pixels) -> None:
        """tpt_scene_set_texture: new level-0 texels (same shape) for a texture."""
        px = np.ascontiguousarray(pixels, np.float32)
        if tuple(px.shape) != self.texture_shapes[texture]:
            raise ValueError("set_texture: the texture's shape must stay " + str(self.texture_shapes[texture]))
        check(lib().tpt_scene_set_texture(self.handle, texture, ptr(px), 0))

    def render_adjoint_batch(self, views, fp64: bool = False, texels: bool = True):
        """tpt_render_adjoint_batch over host arrays.  views: (camera, width,
        height, spp, max_depth, dloss_dimage) tuples -> (d_params
        (material_count, 5), [per texture level-0 gradient] or None), the sums
        over the views."""
        arr = (_lib.AdjointFrameDesc * len(views))()
        keep = []
        for i, (cam, w, h, spp, depth, dl) in enumerate(views):
            d = arr[i].render
            d.camera = cam
            d.width, d.height, d.samples_per_pixel, d.max_depth = w, h, spp, depth
            g = np.ascontiguousarray(dl, dtype=np.float32).reshape(h, w, 3)
            keep.append(g)
            arr[i].dloss_dimage = g.ctypes.data
        out = np.zeros((self.material_count, 5), np.float32)
        shapes = self.texture_shapes
        flat = np.zeros(max(1, sum(int(np.prod(sh)) for sh in shapes)), np.float32) if texels and shapes else None
        check(lib().tpt_render_adjoint_batch(self.handle, arr, len(views), ptr(out), ptr(flat),
                                             _lib.TPT_FLAG_GRAD_FP64 if fp64 else 0))
        if flat is None:
            return out, None
        grads, o = [], 0
        for sh in shapes:
            k = int(np.prod(sh))
            grads.append(flat[o:o + k].reshape(sh).copy())
            o += k
        return out, grads

    def render_adjoint_batch_device(self, views, d_params_dev, d_texels_dev=None, fp64: bool = False) -> None:
        """As render_adjoint_batch with device buffers (torch tensors): each
        view's dloss (h*w*3), d_params (material_count*5), d_texels
        (optional, every texture's level-0 texels packed)."""
        arr = (_lib.AdjointFrameDesc * len(views))()
        for i, (cam, w, h, spp, depth, dl) in enumerate(views):
            d = arr[i].render
            d.camera = cam
            d.width, d.height, d.samples_per_pixel, d.max_depth = w, h, spp, depth
            arr[i].dloss_dimage = ptr(dl)
        flags = _lib.TPT_FLAG_DEVICE_POINTERS | (_lib.TPT_FLAG_GRAD_FP64 if fp64 else 0)
        check(lib().tpt_render_adjoint_batch(self.handle, arr, len(views), ptr(d_params_dev), ptr(d_texels_dev),
                                             flags))

    def set_texture_device(self, texture: int, pixels_dev) -> None:
        """tpt_scene_set_texture from a device buffer (a torch tensor of the
        texture's level-0 shape); synchronous."""
        if tuple(pixels_dev.shape) != self.texture_shapes[texture]:
            raise ValueError("set_texture: the texture's shape must stay " + str(self.texture_shapes[texture]))
        check(lib().tpt_scene_set_texture(self.handle, texture, ptr(pixels_dev), _lib.TPT_FLAG_DEVICE_POINTERS))

    def render_batch(self, frames, quantized: bool = False, exact_shadow: bool = False) -> list:
        """Wavefront path tracer over a batch of frames (tpt_render_batch).
        frames: (camera, width, height, spp, max_depth) tuples -> one
        (height, width, 3) float32 image per frame, as render() gives it."""
        outs = [np.zeros((f[2], f[1], 3), np.float32) for f in frames]
        arr = (_lib.FrameDesc * max(1, len(frames)))()
        for i, (cam, w, h, spp, depth) in enumerate(frames):
            d = arr[i].render
            d.camera = cam
            d.width, d.height, d.samples_per_pixel, d.max_depth = w, h, spp, depth
            arr[i].rgb = ptr(outs[i])
        flags = (_lib.TPT_FLAG_QUANTIZED if quantized else 0) | (_lib.TPT_FLAG_ANY_EXACT if exact_shadow else 0)
        check(lib().tpt_render_batch(self.handle, arr, len(frames), flags))
        return outs

    def autotune(self, frames, quantized: bool = False, tune_spp: int = 32) -> dict:
        """tpt_autotune_render over these frames at up to tune_spp samples per
        pixel (paths-in-flight candidates above the tuning batch's path count
        are skipped, so the batch should hold a few passes); the chosen
        parameters stay on the device.  -> dict of the result fields."""
        arr = (_lib.FrameDesc * max(1, len(frames)))()
        for i, (cam, w, h, spp, depth) in enumerate(frames):
            d = arr[i].render
            d.camera = cam
            d.width, d.height, d.samples_per_pixel, d.max_depth = w, h, spp, depth
        res = _lib.TuneResult()
        flags = _lib.TPT_FLAG_QUANTIZED if quantized else 0
        check(lib().tpt_autotune_render(self.handle, arr, len(frames), flags, tune_spp, ctypes.byref(res)))
        return {name: getattr(res, name) for name, _ in _lib.TuneResult._fields_}

    def texture_info(self, texture: int) -> dict:
        """tpt_texture_info -> {levels, channels, max_value}."""
        lv, ch, mx = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_float(0.0)
        check(lib().tpt_texture_info(self.handle, texture, ctypes.byref(lv), ctypes.byref(ch), ctypes.byref(mx)))
        return {"levels": lv.value, "channels": ch.value, "max_value": mx.value}

    def texture_level(self, texture: int, level: int) -> np.ndarray:
        """tpt_texture_level: mip level `level` as (h, w, 3) or (h, w) float32."""
        w, h = ctypes.c_uint32(0), ctypes.c_uint32(0)
        check(lib().tpt_texture_level(self.handle, texture, level, None, 0, ctypes.byref(w), ctypes.byref(h)))
        ch = self.texture_info(texture)["channels"]
        out = np.zeros((h.value, w.value, ch) if ch == 3 else (h.value, w.value), np.float32)
        check(lib().tpt_texture_level(self.handle, texture, level, ptr(out), out.size, ctypes.byref(w),
                                      ctypes.byref(h)))
        return out

    def texture_sample(self, texture: int, uv, dstdx=None, dstdy=None, mode: int = 0) -> np.ndarray:
        """tpt_texture_sample: lookups of one texture at rows of uv with the
        uv derivatives dstdx / dstdy (zero when None) in `mode`
        (tpt._lib.TEXTURE_*) -> (n, channels) float32."""
        uv = np.asarray(uv, np.float32).reshape(-1, 2)
        q = np.zeros(len(uv), _lib.TEXTURE_QUERY_DTYPE)
        q["uv"] = uv
        if dstdx is not None:
            q["dstdx"] = np.asarray(dstdx, np.float32).reshape(-1, 2)
        if dstdy is not None:
            q["dstdy"] = np.asarray(dstdy, np.float32).reshape(-1, 2)
        ch = self.texture_info(texture)["channels"]
        out = np.zeros((len(uv), ch), np.float32)
        check(lib().tpt_texture_sample(self.handle, texture, mode, q.ctypes.data, len(uv), ptr(out), 0))
        return out

    def material_bsdf(self, material, wo, wi, u, front_face: bool = True) -> np.ndarray:
        """tpt_material_bsdf: the BSDF of material index `material` in the
        shading frame, for rows of wo, wi (eval / pdf) and u = (u_lobe, u0,
        u1) (sample) -> structured array of BSDF_RESULT_DTYPE."""
        wo, wi, u = (np.asarray(a, np.float32).reshape(-1, 3) for a in (wo, wi, u))
        n = max(len(wo), len(wi), len(u))
        q = np.zeros(n, _lib.BSDF_QUERY_DTYPE)
        q["wo"], q["wi"], q["u"] = wo, wi, u
        q["material"] = np.asarray(material, np.uint32)
        q["front_face"] = int(bool(front_face))
        out = np.zeros(n, _lib.BSDF_RESULT_DTYPE)
        check(lib().tpt_material_bsdf(self.handle, q.ctypes.data, n, out.ctypes.data, 0))
        return out

    def render_batch_device(self, frames, outs_dev, quantized: bool = False, exact_shadow: bool = False) -> None:
        """As render_batch into device buffers (torch tensors); asynchronous."""
        arr = (_lib.FrameDesc * max(1, len(frames)))()
        for i, (cam, w, h, spp, depth) in enumerate(frames):
            d = arr[i].render
            d.camera = cam
            d.width, d.height, d.samples_per_pixel, d.max_depth = w, h, spp, depth
            arr[i].rgb = ptr(outs_dev[i])
        flags = (_lib.TPT_FLAG_DEVICE_POINTERS | (_lib.TPT_FLAG_QUANTIZED if quantized else 0)
                 | (_lib.TPT_FLAG_ANY_EXACT if exact_shadow else 0))
        check(lib().tpt_render_batch(self.handle, arr, len(frames), flags))

    # device-pointer queries (asynchronous on the device stream) -----------------
    def closest_hit_device(self, rays_dev, hits_dev, n: int, quantized: bool = False) -> None:
        flags = _lib.TPT_FLAG_DEVICE_POINTERS | (_lib.TPT_FLAG_QUANTIZED if quantized else 0)
        check(lib().tpt_trace_closest(self.handle, ptr(rays_dev), n, ptr(hits_dev), flags))

    def unoccluded_device(self, rays_dev, target_dev, distance_dev, out_dev, n: int, quantized: bool = False,
                          exact: bool = False) -> None:
        flags = (_lib.TPT_FLAG_DEVICE_POINTERS | (_lib.TPT_FLAG_QUANTIZED if quantized else 0)
                 | (_lib.TPT_FLAG_ANY_EXACT if exact else 0))
        check(lib().tpt_trace_any(self.handle, ptr(rays_dev), n, ptr(target_dev), ptr(distance_dev), ptr(out_dev),
                                  flags))


def make_rays(origins: np.ndarray, directions: np.ndarray, tmax: float | np.ndarray = np.inf) -> np.ndarray:
    r = np.zeros(len(origins), dtype=RAY_DTYPE)
    r["o"] = origins
    r["d"] = directions
    r["tmax"] = tmax
    return r


__all__ = [
    "ClosestHitError", "Device", "DeviceScene", "Instance", "Mesh", "Qbvh", "Scene", "build_qbvh", "glam_cols",
    "make_rays", "quantize_nodes", "transform_bounds", "transform_point3", "HIT_DTYPE", "RAY_DTYPE",
    "NODE_DTYPE", "Q8_NODE_DTYPE", "INVALID_INDEX",
]
