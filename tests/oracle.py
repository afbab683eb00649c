"""ctypes wrapper of the oracle (oracle/liboracle.so), the CPU restatement of
the reference ray-query path.  Test infrastructure: used only as the checker.

liboracle64.so is the same restatement compiled with every arithmetic value
widened to double (ORC_FP64): the FP64 reference renderer of the gradient
checks (OracleScene / OracleRenderer with fp64=True).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
ORACLE_LIB64 = ORACLE_DIR / "liboracle64.so"

_libs = {}


def lib(fp64: bool = False) -> ctypes.CDLL:
    """The oracle library: the f32 restatement, or with fp64 its FP64 build."""
    if fp64 not in _libs:
        path = ORACLE_LIB64 if fp64 else ORACLE_LIB
        if not path.exists():
            subprocess.run(["make", "-C", os.fspath(ORACLE_DIR)], check=True, capture_output=True)
        L = ctypes.CDLL(os.fspath(path))
        vp, sz, f32p = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p
        L.orc_intersect_triangle.argtypes = [f32p, ctypes.c_float, f32p, f32p]
        L.orc_intersect_triangle.restype = ctypes.c_int
        L.orc_difference_of_products.argtypes = [ctypes.c_float] * 4
        L.orc_difference_of_products.restype = ctypes.c_float
        L.orc_gamma.argtypes = [ctypes.c_int]
        L.orc_gamma.restype = ctypes.c_float
        L.orc_qbvh_build.argtypes = [vp, vp, sz]
        L.orc_qbvh_build.restype = vp
        L.orc_qbvh_free.argtypes = [vp]
        L.orc_qbvh_node_count.argtypes = [vp]
        L.orc_qbvh_node_count.restype = sz
        L.orc_qbvh_prim_count.argtypes = [vp]
        L.orc_qbvh_prim_count.restype = sz
        L.orc_qbvh_export.argtypes = [vp, vp, vp, vp, vp]
        L.orc_qbvh_visit_leaves.argtypes = [vp, f32p, ctypes.c_float, vp, sz]
        L.orc_qbvh_visit_leaves.restype = sz
        L.orc_qbvh_closest_box.argtypes = [vp, vp, vp, f32p, f32p]
        L.orc_qbvh_closest_box.restype = ctypes.c_long
        L.orc_scene_new.restype = vp
        L.orc_scene_free.argtypes = [vp]
        L.orc_scene_add_mesh.argtypes = [vp, vp, sz, vp, sz]
        L.orc_scene_add_mesh.restype = ctypes.c_int
        L.orc_scene_add_instance.argtypes = [vp, ctypes.c_uint32, vp, vp]
        L.orc_scene_add_instance.restype = ctypes.c_int
        L.orc_scene_build.argtypes = [vp]
        L.orc_scene_is_built.argtypes = [vp]
        L.orc_scene_is_built.restype = ctypes.c_int
        L.orc_scene_mesh_qbvh.argtypes = [vp, ctypes.c_uint32]
        L.orc_scene_mesh_qbvh.restype = vp
        L.orc_scene_tlas.argtypes = [vp]
        L.orc_scene_tlas.restype = vp
        L.orc_render_scene_new.argtypes = [vp, ctypes.c_uint32]
        L.orc_render_scene_new.restype = vp
        L.orc_render_scene_free.argtypes = [vp]
        L.orc_render_set_mesh.argtypes = [vp, ctypes.c_uint32, vp, vp, sz]
        L.orc_render_set_instance_material.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_render_set_materials.argtypes = [vp, vp, sz]
        L.orc_render.argtypes = [vp, vp] + [ctypes.c_uint32] * 6 + [vp, ctypes.c_int]
        L.orc_render_rect.argtypes = [vp, vp] + [ctypes.c_uint32] * 8 + [vp, ctypes.c_int]
        u32, f32 = ctypes.c_uint32, ctypes.c_float
        L.orc_render_light_tree.argtypes = [vp, vp, u32]
        L.orc_render_light_tree.restype = u32
        L.orc_render_tree_sample.argtypes = [vp, vp, vp, f32, f32, vp, vp]
        L.orc_render_tree_sample.restype = ctypes.c_int
        L.orc_render_tree_pdf.argtypes = [vp, vp, vp, f32, u32]
        L.orc_render_tree_pdf.restype = f32
        L.orc_render_set_nee_policy.argtypes = [vp, u32]
        L.orc_render_set_nee_policy.restype = None
        L.orc_render_vertex_tree_sample.argtypes = [vp, u32, u32, vp, vp, f32, vp, vp]
        L.orc_render_vertex_tree_sample.restype = ctypes.c_int
        L.orc_render_vertex_tree_pdf.argtypes = [vp, u32, u32, vp, vp, u32]
        L.orc_render_vertex_tree_pdf.restype = f32
        L.orc_sggx_reflection_pdf.argtypes = [vp, vp, vp]
        L.orc_sggx_reflection_pdf.restype = f32
        L.orc_merge_glossy_roughness.argtypes = [f32, f32, f32, f32, f32, f32, vp]
        L.orc_merge_glossy_roughness.restype = None
        L.orc_glossy_importance.argtypes = [f32, vp, vp, f32, f32, vp, f32, vp, f32]
        L.orc_glossy_importance.restype = f32
        L.orc_render_direct_light.argtypes = [vp, u32, u32, vp, vp, f32, f32, f32, vp, vp]
        L.orc_render_emitted.argtypes = [vp, u32, u32, vp, vp, u32, u32, vp, vp, vp, f32, ctypes.c_int, vp]
        L.orc_render_trace.argtypes = [vp, vp, vp] + [u32] * 7 + [vp]
        L.orc_render_set_environment.argtypes = [vp, vp, u32, u32, f32, f32]
        L.orc_env_pdf.argtypes = [vp, vp, ctypes.c_int]
        L.orc_env_pdf.restype = f32
        L.orc_env_radiance.argtypes = [vp, vp, vp]
        L.orc_env_sample.argtypes = [vp, f32, f32, ctypes.c_int, vp]
        L.orc_env_sample.restype = ctypes.c_int
        L.orc_env_total_power.argtypes = [vp]
        L.orc_env_total_power.restype = f32
        L.orc_env_uv.argtypes = [vp, vp]
        L.orc_env_dir.argtypes = [f32, f32, vp]
        L.orc_render_emitted_infinite.argtypes = [vp, vp, f32, ctypes.c_int, vp, vp]
        L.orc_sg.argtypes = [ctypes.c_int, f32, f32]
        L.orc_sg.restype = f32
        L.orc_sg_product.argtypes = [vp, f32, vp, f32, vp]
        L.orc_thin_film.argtypes = [f32, f32, f32, vp, vp, f32, vp]
        L.orc_artist_ior.argtypes = [vp, vp, vp]
        L.orc_lut_sizes.argtypes = [vp]
        L.orc_lut_albedo.argtypes = [f32, vp]
        L.orc_lut_cond.argtypes = [vp, vp]
        L.orc_lut_diel.argtypes = [vp, vp]
        L.orc_lut_lookup.argtypes = [ctypes.c_int, f32, vp, f32, f32]
        L.orc_lut_lookup.restype = f32
        L.orc_openpbr_eval.argtypes = [vp, ctypes.c_int, vp, vp, sz, vp, vp]
        L.orc_openpbr_sample.argtypes = [vp, ctypes.c_int, vp, vp, sz, vp]
        L.orc_openpbr_probe.argtypes = [vp, ctypes.c_int, vp, f32, f32, vp]
        L.orc_render_add_directional.argtypes = [vp, vp, vp, f32]
        L.orc_render_light_categories.argtypes = [vp, vp, vp, u32]
        L.orc_render_light_categories.restype = u32
        L.orc_render_add_texture.argtypes = [vp, u32, u32, ctypes.c_int, vp]
        L.orc_render_add_texture.restype = u32
        L.orc_texture_new.argtypes = [u32, u32, ctypes.c_int, vp]
        L.orc_texture_new.restype = vp
        L.orc_texture_free.argtypes = [vp]
        L.orc_texture_levels.argtypes = [vp]
        L.orc_texture_levels.restype = u32
        L.orc_texture_max_value.argtypes = [vp]
        L.orc_texture_max_value.restype = f32
        L.orc_texture_level_size.argtypes = [vp, u32, vp, vp]
        L.orc_texture_level.argtypes = [vp, u32, vp]
        L.orc_texture_sample.argtypes = [vp, ctypes.c_int, vp, sz, vp]
        L.orc_closest_hit.argtypes = [vp, vp, sz, vp, vp, ctypes.c_int]
        L.orc_closest_hit.restype = ctypes.c_int
        L.orc_unoccluded.argtypes = [vp, vp, sz, vp, vp, vp, vp, ctypes.c_int]
        L.orc_unoccluded.restype = ctypes.c_int
        L.orc_render_rect_f64.argtypes = [vp, vp] + [ctypes.c_uint32] * 8 + [vp, ctypes.c_int]
        L.orc_render_rect_detached_f64.argtypes = [vp, vp, vp] + [ctypes.c_uint32] * 8 + [vp, ctypes.c_int]
        _libs[fp64] = L
    return _libs[fp64]


def env_uv(d):
    da, uv = np.asarray(d, np.float32), np.zeros(2, np.float32)
    lib().orc_env_uv(_p(da), _p(uv))
    return uv


def env_dir(u, v):
    out = np.zeros(3, np.float32)
    lib().orc_env_dir(u, v, _p(out))
    return out


def _p(a):
    return None if a is None else a.ctypes.data


# ---- GGX tables and the OpenPBR BSDF (oracle/bsdf.h) ------------------------------------------
def lut_sizes():
    out = np.zeros(5, np.uint32)
    lib().orc_lut_sizes(_p(out))
    return [int(x) for x in out]


def lut_albedo(eta):
    """Dielectric GGX directional albedo table at eta: (64, 16, 64, 2) =
    (sqrt cos, phi, roughness, anisotropy slices 7 and 8)."""
    out = np.zeros(lut_sizes()[0], np.float32)
    lib().orc_lut_albedo(eta, _p(out))
    return out.reshape(64, 16, 64, 2)


def lut_conductor():
    """(e (roughness 64, sqrt cos 64), e_avg (64))."""
    n = lut_sizes()
    e, ea = np.zeros(n[1], np.float32), np.zeros(n[2], np.float32)
    lib().orc_lut_cond(_p(e), _p(ea))
    return e.reshape(64, 64), ea


def lut_dielectric():
    """(e (eta 32, roughness 32, sqrt cos 32), e_avg (eta 32, roughness 32))."""
    n = lut_sizes()
    e, ea = np.zeros(n[3], np.float32), np.zeros(n[4], np.float32)
    lib().orc_lut_diel(_p(e), _p(ea))
    return e.reshape(32, 32, 32), ea.reshape(32, 32)


def lut_lookup(which, eta=1.5, w=(0.0, 0.0, 1.0), a=0.0, b=0.0):
    wa = np.asarray(w, np.float32)
    return lib().orc_lut_lookup(which, eta, _p(wa), a, b)


def _mat_buf(material):
    rec = material.record()
    buf = np.frombuffer(bytes(rec), np.uint8).copy()
    return buf


def openpbr_eval(material, wo, wi, front_face=True):
    """OpenPbrBsdf eval (n, 3) and pdf (n,) in the shading frame."""
    m = _mat_buf(material)
    woa = np.asarray(wo, np.float32)
    wia = np.ascontiguousarray(np.asarray(wi, np.float32).reshape(-1, 3))
    f, pdf = np.zeros((len(wia), 3), np.float32), np.zeros(len(wia), np.float32)
    lib().orc_openpbr_eval(_p(m), int(front_face), _p(woa), _p(wia), len(wia), _p(f), _p(pdf))
    return f, pdf


def openpbr_sample(material, wo, u, front_face=True):
    """OpenPbrBsdf sample for rows (u_lobe, u0, u1) -> dict of ok, weight,
    wi, pdf, delta, transmission, eta."""
    m = _mat_buf(material)
    woa = np.asarray(wo, np.float32)
    ua = np.ascontiguousarray(np.asarray(u, np.float32).reshape(-1, 3))
    out = np.zeros((len(ua), 11), np.float32)
    lib().orc_openpbr_sample(_p(m), int(front_face), _p(woa), _p(ua), len(ua), _p(out))
    return dict(ok=out[:, 0] > 0, weight=out[:, 1:4], wi=out[:, 4:7], pdf=out[:, 7], delta=out[:, 8] > 0,
                transmission=out[:, 9] > 0, eta=out[:, 10])


def openpbr_probe(material, which, wo=(0.0, 0.0, 1.0), a=0.0, b=0.0):
    m = _mat_buf(material)
    woa, out = np.asarray(wo, np.float32), np.zeros(3, np.float32)
    lib().orc_openpbr_probe(_p(m), which, _p(woa), a, b, _p(out))
    return out


# ---- textures (oracle/texture.h) --------------------------------------------------------------
class OracleTexture:
    """The reference's Texture / ScalarTexture on the CPU: pixels (h, w, 3) or
    (h, w) float32."""

    def __init__(self, pixels):
        px = np.ascontiguousarray(pixels, np.float32)
        self.channels = 3 if px.ndim == 3 else 1
        self._px = px
        self.h = lib().orc_texture_new(px.shape[1], px.shape[0], self.channels, _p(px))

    @property
    def levels(self) -> int:
        return lib().orc_texture_levels(self.h)

    @property
    def max_value(self) -> float:
        return lib().orc_texture_max_value(self.h)

    def level(self, i):
        w, h = ctypes.c_uint32(0), ctypes.c_uint32(0)
        lib().orc_texture_level_size(self.h, i, ctypes.byref(w), ctypes.byref(h))
        out = np.zeros((h.value, w.value, self.channels) if self.channels == 3 else (h.value, w.value), np.float32)
        lib().orc_texture_level(self.h, i, _p(out))
        return out

    def sample(self, uv, dstdx=None, dstdy=None, mode=0):
        """mode: 0 filtered (EWA), 1 bilinear, 2 nearest, 3 trilinear, 4 mip-bilinear."""
        uv = np.asarray(uv, np.float32).reshape(-1, 2)
        q = np.zeros((len(uv), 6), np.float32)
        q[:, 0:2] = uv
        if dstdx is not None:
            q[:, 2:4] = np.asarray(dstdx, np.float32).reshape(-1, 2)
        if dstdy is not None:
            q[:, 4:6] = np.asarray(dstdy, np.float32).reshape(-1, 2)
        out = np.zeros((len(uv), self.channels), np.float32)
        lib().orc_texture_sample(self.h, mode, _p(q), len(uv), _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_texture_free(self.h)


HIT_DTYPE = np.dtype([("instance", "<u4"), ("triangle", "<u4"), ("t", "<f4"), ("b", "<f4", 3)])
RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmax", "<f4"), ("d", "<f4", 3), ("pad", "<f4")])
NODE_DTYPE = np.dtype(
    [("min_x", "<f4", 4), ("min_y", "<f4", 4), ("min_z", "<f4", 4), ("max_x", "<f4", 4), ("max_y", "<f4", 4),
     ("max_z", "<f4", 4), ("children", "<u4", 4), ("valid_mask", "u1"), ("pad", "u1", 15)])


def intersect_triangle(origin, direction, t_max, v0, v1, v2):
    """intersect_triangle -> (t, b0, b1, b2) or None."""
    r = np.array(list(origin) + list(direction), dtype=np.float32)
    v = np.array(list(v0) + list(v1) + list(v2), dtype=np.float32)
    out = np.zeros(4, dtype=np.float32)
    ok = lib().orc_intersect_triangle(_p(r), ctypes.c_float(t_max), _p(v), _p(out))
    return (float(out[0]), float(out[1]), float(out[2]), float(out[3])) if ok else None


def _export(h) -> tuple[np.ndarray, np.ndarray]:
    L = lib()
    nn = L.orc_qbvh_node_count(h)
    npr = L.orc_qbvh_prim_count(h)
    bounds = np.zeros((nn, 24), np.float32)
    ch = np.zeros((nn, 4), np.uint32)
    valid = np.zeros(nn, np.uint8)
    prim = np.zeros(npr, np.uint32)
    L.orc_qbvh_export(h, _p(bounds), _p(ch), _p(valid), _p(prim))
    nodes = np.zeros(nn, dtype=NODE_DTYPE)
    for k, name in enumerate(["min_x", "min_y", "min_z", "max_x", "max_y", "max_z"]):
        nodes[name] = bounds[:, 4 * k:4 * k + 4]
    nodes["children"] = ch
    nodes["valid_mask"] = valid
    return nodes, prim


class OracleQbvh:
    def __init__(self, bmin, bmax):
        self.bmin = np.ascontiguousarray(bmin, np.float32).reshape(-1, 3)
        self.bmax = np.ascontiguousarray(bmax, np.float32).reshape(-1, 3)
        self.h = lib().orc_qbvh_build(_p(self.bmin), _p(self.bmax), len(self.bmin))
        self.nodes, self.prim = _export(self.h) if self.h else (np.zeros(0, NODE_DTYPE), np.zeros(0, np.uint32))

    def leaves(self, origin, direction, t_max=np.inf, cap=1 << 16):
        r = np.array(list(origin) + list(direction), np.float32)
        out = np.zeros((cap, 2), np.uint32)
        n = lib().orc_qbvh_visit_leaves(self.h, _p(r), ctypes.c_float(t_max), _p(out), cap)
        return out[:min(n, cap)]

    def closest_box(self, origin, direction):
        """The reference test helper qbvh_closest -> (index, t) or None."""
        r = np.array(list(origin) + list(direction), np.float32)
        t = np.zeros(1, np.float32)
        i = lib().orc_qbvh_closest_box(self.h, _p(self.bmin), _p(self.bmax), _p(r), _p(t))
        return None if i < 0 else (int(i), float(t[0]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_qbvh_free(self.h)


class OracleScene:
    """Oracle scene: meshes + instances (math-convention 4x4 matrices); fp64:
    in the FP64 build of the oracle."""

    def __init__(self, fp64: bool = False):
        self._L = lib(fp64)
        self.h = self._L.orc_scene_new()
        self._keep = []

    def add_mesh(self, positions, indices) -> int:
        p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        i = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        return self._L.orc_scene_add_mesh(self.h, _p(p), len(p), _p(i), len(i))

    def add_instance(self, mesh, local_to_world, world_to_local) -> int:
        l2w = np.ascontiguousarray(np.asarray(local_to_world, np.float32).T).reshape(16)
        w2l = np.ascontiguousarray(np.asarray(world_to_local, np.float32).T).reshape(16)
        return self._L.orc_scene_add_instance(self.h, mesh, _p(l2w), _p(w2l))

    def build(self):
        self._L.orc_scene_build(self.h)

    def mesh_qbvh(self, mesh: int):
        return _export(self._L.orc_scene_mesh_qbvh(self.h, mesh))

    def tlas(self):
        return _export(self._L.orc_scene_tlas(self.h))

    def closest_hit(self, rays, threads=8, stats=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        st = np.zeros((len(rays), 3), np.uint32) if stats else None
        rc = self._L.orc_closest_hit(self.h, _p(rays), len(rays), _p(hits), _p(st), threads)
        if rc != 0:
            raise RuntimeError("QbvhNotBuilt")
        return (hits, st) if stats else hits

    def unoccluded(self, rays, target=None, distance=None, threads=8, stats=False):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(len(rays), np.uint8)
        tg = None if target is None else np.ascontiguousarray(target, np.uint32).reshape(-1, 2)
        ds = None if distance is None else np.ascontiguousarray(distance, np.float32)
        st = np.zeros((len(rays), 3), np.uint32) if stats else None
        rc = self._L.orc_unoccluded(self.h, _p(rays), len(rays), _p(tg), _p(ds), _p(out), _p(st), threads)
        if rc != 0:
            raise RuntimeError("QbvhNotBuilt")
        return (out.astype(bool), st) if stats else out.astype(bool)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_scene_free(self.h)


def oracle_scene_from(scene, fp64: bool = False) -> OracleScene:
    """An oracle copy of a tpt.Scene (same f32 arrays and matrices), built."""
    o = OracleScene(fp64)
    for m in scene.meshes:
        o.add_mesh(m.positions, m.indices)
    for inst in scene.instances:
        o.add_instance(inst.mesh_index, inst.local_to_world, inst.world_to_local)
    o.build()
    return o


class OracleRenderer:
    """The oracle's MIS path tracer over a tpt.Scene and its materials
    (tpt.material records): light tree built from the emissive instances.
    fp64: the FP64 build (render_rect_f64 returns its image in float64)."""

    def __init__(self, scene, materials, nee_policy: str = "reference", fp64: bool = False):
        from tpt.material import records
        self._L = lib(fp64)
        self.fp64 = fp64
        self.geo = oracle_scene_from(scene, fp64)
        self.h = self._L.orc_render_scene_new(self.geo.h, len(scene.meshes))
        self._keep = []
        for i, m in enumerate(scene.meshes):
            n = None if m.normals is None else np.ascontiguousarray(m.normals, np.float32)
            uv = None if m.uvs is None else np.ascontiguousarray(m.uvs, np.float32)
            self._keep += [n, uv]
            self._L.orc_render_set_mesh(self.h, i, _p(n), _p(uv), len(m.positions))
        for i, inst in enumerate(scene.instances):
            self._L.orc_render_set_instance_material(self.h, i, inst.material_index)
        self.scene_meshes = scene.meshes
        for t in getattr(scene, "textures", []):  # before the materials: the light tree reads them
            px = np.ascontiguousarray(t.pixels, np.float32)
            self._L.orc_render_add_texture(self.h, px.shape[1], px.shape[0], 3 if px.ndim == 3 else 1, _p(px))
        self.recs = records(materials)
        self._L.orc_render_set_materials(self.h, ctypes.addressof(self.recs), len(materials))
        for dl in getattr(scene, "directional_lights", []):
            d, c = np.asarray(dl.direction, np.float32), np.asarray(dl.color, np.float32)
            self._L.orc_render_add_directional(self.h, _p(d), _p(c), dl.intensity)
        env = getattr(scene, "environment", None)
        if env is not None:
            px = np.ascontiguousarray(env.pixels, np.float32)
            self._L.orc_render_set_environment(self.h, _p(px), px.shape[1], px.shape[0], env.scale, env.rotate_y)
        self.set_nee_policy(nee_policy)

    def set_nee_policy(self, policy: str) -> None:
        """"reference" (OpenPBR builds no light-tree query) or
        "openpbr_proxy" (the OpenPBR proxy query)."""
        self._L.orc_render_set_nee_policy(self.h, {"reference": 0, "openpbr_proxy": 1}[policy])

    def vertex_tree_sample(self, inst, tri, bary, direction, u):
        """The stochastic descent with the query of a shading vertex under the
        renderer's policy: (leaf node, pmf) or None."""
        leaf, pmf = ctypes.c_uint32(0), ctypes.c_float(0)
        b, d = np.asarray(bary, np.float32), np.asarray(direction, np.float32)
        ok = self._L.orc_render_vertex_tree_sample(self.h, inst, tri, _p(b), _p(d), u, ctypes.byref(leaf),
                                                 ctypes.byref(pmf))
        return (leaf.value, pmf.value) if ok else None

    def vertex_tree_pdf(self, inst, tri, bary, direction, leaf):
        b, d = np.asarray(bary, np.float32), np.asarray(direction, np.float32)
        return self._L.orc_render_vertex_tree_pdf(self.h, inst, tri, _p(b), _p(d), leaf)

    def light_categories(self):
        """The light sampler's categories: [(kind, pmf)], kind 0 tree, 1
        environment, 2 + i directional light i."""
        kind, pmf = np.zeros(32, np.int32), np.zeros(32, np.float32)
        n = self._L.orc_render_light_categories(self.h, _p(kind), _p(pmf), 32)
        return [(int(kind[i]), float(pmf[i])) for i in range(min(n, 32))]

    # environment light queries (R/light/environment.rs)
    def env_pdf(self, d, mis=False):
        da = np.asarray(d, np.float32)
        return self._L.orc_env_pdf(self.h, _p(da), int(mis))

    def env_radiance(self, d):
        da, out = np.asarray(d, np.float32), np.zeros(3, np.float32)
        self._L.orc_env_radiance(self.h, _p(da), _p(out))
        return out

    def env_sample(self, u, v, mis=False):
        out = np.zeros(7, np.float32)
        ok = self._L.orc_env_sample(self.h, u, v, int(mis), _p(out))
        return (out[:3], out[3:6], float(out[6])) if ok else None

    def env_total_power(self):
        return self._L.orc_env_total_power(self.h)

    def emitted_infinite(self, throughput, bsdf_pdf, delta, direction):
        t, d, out = np.asarray(throughput, np.float32), np.asarray(direction, np.float32), np.zeros(3, np.float32)
        self._L.orc_render_emitted_infinite(self.h, _p(t), bsdf_pdf, int(delta), _p(d), _p(out))
        return out

    def light_tree(self):
        """Light-tree nodes: structured array (see orc_render_light_tree)."""
        cap = 2 * sum(len(m.indices) for m in self.scene_meshes) + 1
        raw = np.zeros((cap, 20), np.float32)
        n = self._L.orc_render_light_tree(self.h, _p(raw), cap)
        raw = raw[:n]
        u = raw.view(np.uint32)
        return dict(flux=raw[:, 0], mu=raw[:, 1:4], sigma_s2=raw[:, 4], nu=raw[:, 5:8], **{"lambda": raw[:, 8]},
                    radius=raw[:, 9], area=raw[:, 10], left=u[:, 11], right=u[:, 12], parent=u[:, 13],
                    instance=u[:, 14], triangle=u[:, 15], leaf=u[:, 16].astype(bool))

    # numpy temporaries are bound to names so they outlive the ctypes calls
    def tree_sample(self, p, n, rho, u):
        leaf = ctypes.c_uint32(0)
        pmf = ctypes.c_float(0)
        pa, na = np.asarray(p, np.float32), np.asarray(n, np.float32)
        ok = self._L.orc_render_tree_sample(self.h, _p(pa), _p(na), rho, u, ctypes.byref(leaf), ctypes.byref(pmf))
        return (leaf.value, pmf.value) if ok else None

    def tree_pdf(self, p, n, rho, leaf):
        pa, na = np.asarray(p, np.float32), np.asarray(n, np.float32)
        return self._L.orc_render_tree_pdf(self.h, _p(pa), _p(na), rho, leaf)

    def direct_light(self, inst, tri, bary, direction, u_root, u_tree, u_aux, us):
        out = np.zeros(3, np.float32)
        b, d, u = (np.asarray(a, np.float32) for a in (bary, direction, us))
        self._L.orc_render_direct_light(self.h, inst, tri, _p(b), _p(d), u_root, u_tree, u_aux, _p(u), _p(out))
        return out

    def emitted(self, vertex, light_vertex, throughput, bsdf_pdf, delta):
        """vertex / light_vertex: (instance, triangle, barycentrics, incident direction)."""
        (i, t, b, d), (li, lt_, lb, ld) = vertex, light_vertex
        b, d, lb, ld, th = (np.asarray(a, np.float32) for a in (b, d, lb, ld, throughput))
        out = np.zeros(3, np.float32)
        self._L.orc_render_emitted(self.h, i, t, _p(b), _p(d), li, lt_, _p(lb), _p(ld), _p(th), bsdf_pdf, int(delta),
                                 _p(out))
        return out

    def trace(self, origin, direction, frame, spp, pixel, sample_index, max_depth):
        out = np.zeros(3, np.float32)
        o, d = np.asarray(origin, np.float32), np.asarray(direction, np.float32)
        self._L.orc_render_trace(self.h, _p(o), _p(d), frame[0], frame[1], spp, pixel[0], pixel[1], sample_index,
                               max_depth, _p(out))
        return out

    def render(self, cam, width, height, spp, max_depth, rows=None, threads=8):
        """cam: tpt.camera.pinhole(...) record.  Returns (height, width, 3) f32."""
        c = np.array(list(cam.eye) + list(cam.forward) + list(cam.right) + list(cam.image_up)
                     + [cam.tan_half_fov_y, cam.pixel_spread, cam.exposure], np.float32)
        out = np.zeros((height, width, 3), np.float32)
        y0, y1 = rows if rows else (0, height)
        self._L.orc_render(self.h, _p(c), width, height, spp, max_depth, y0, y1, _p(out), threads)
        return out

    def render_rect(self, cam, width, height, spp, max_depth, x0, x1, y0, y1, threads=8):
        """Pixels [x0, x1) x [y0, y1) of a width x height frame -> (y1-y0, x1-x0, 3)."""
        c = np.array(list(cam.eye) + list(cam.forward) + list(cam.right) + list(cam.image_up)
                     + [cam.tan_half_fov_y, cam.pixel_spread, cam.exposure], np.float32)
        out = np.zeros((y1 - y0, x1 - x0, 3), np.float32)
        self._L.orc_render_rect(self.h, _p(c), width, height, spp, max_depth, x0, x1, y0, y1, _p(out), threads)
        return out

    def render_rect_f64(self, cam, width, height, spp, max_depth, x0, x1, y0, y1, threads=8):
        """render_rect with the pixels' running means returned unrounded
        (float64; the FP64 build accumulates in double)."""
        c = np.array(list(cam.eye) + list(cam.forward) + list(cam.right) + list(cam.image_up)
                     + [cam.tan_half_fov_y, cam.pixel_spread, cam.exposure], np.float32)
        out = np.zeros((y1 - y0, x1 - x0, 3), np.float64)
        self._L.orc_render_rect_f64(self.h, _p(c), width, height, spp, max_depth, x0, x1, y0, y1, _p(out), threads)
        return out

    def render_rect_detached_f64(self, evaluate, cam, width, height, spp, max_depth, x0, x1, y0, y1, threads=8):
        """render_rect_f64 with detached evaluation: this renderer's paths
        (sampling decisions, densities, MIS weights, roulette) with the BSDF
        values and emission of `evaluate`, an OracleRenderer of the same
        build over the same geometry with other materials or textures.  Its
        central differences over `evaluate`'s parameters are the derivative a
        path-replay adjoint computes (sampling held at this renderer's
        parameters)."""
        if evaluate.fp64 != self.fp64:
            raise ValueError("both renderers must come from the same oracle build")
        c = np.array(list(cam.eye) + list(cam.forward) + list(cam.right) + list(cam.image_up)
                     + [cam.tan_half_fov_y, cam.pixel_spread, cam.exposure], np.float32)
        out = np.zeros((max(0, y1 - y0), max(0, x1 - x0), 3), np.float64)
        self._L.orc_render_rect_detached_f64(self.h, evaluate.h, _p(c), width, height, spp, max_depth, x0, x1, y0, y1,
                                             _p(out), threads)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_render_scene_free(self.h)


def thin_film(cos_theta_i, eta_outside, eta_film, n_base, k_base, thickness_nm):
    """eval_thin_film_inner (R/bsdf/thin_film.rs:47-109) -> RGB reflectance."""
    n = np.asarray(n_base, np.float32).reshape(3)
    k = np.asarray(k_base, np.float32).reshape(3)
    out = np.zeros(3, np.float32)
    lib().orc_thin_film(cos_theta_i, eta_outside, eta_film, _p(n), _p(k), thickness_nm, _p(out))
    return out


def artist_ior(face, edge):
    """artist_friendly_complex_ior (R/bsdf/conductor_complex.rs:132-147) -> (n, k)."""
    f = np.asarray(face, np.float32).reshape(3)
    e = np.asarray(edge, np.float32).reshape(3)
    out = np.zeros(6, np.float32)
    lib().orc_artist_ior(_p(f), _p(e), _p(out))
    return out[:3], out[3:]
