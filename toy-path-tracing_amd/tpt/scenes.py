"""Seeded procedural scenes and ray sets for the ray-query workloads.

The reference scenes load Git-LFS assets that are not available, so every
workload here is generated from a seed (numpy PCG64), in float64 and rounded
to float32 once.  Generated data is an input: the GPU path and the oracle see
the same arrays.

Workloads
---------
* ``cornell``: the Cornell box of the plain-method configuration — five walls
  of [0,1]^3 open at z=0, two rotated boxes (0.3^3 at -18 deg, 0.3x0.6x0.3 at
  +15 deg) and a 0.2x0.2 ceiling quad light; 36 triangles in 6 instances.
* ``sphere_field``: the traversal microbenchmark mesh — level-4 icospheres
  (5120 triangles each), random centre / log-uniform radius / random rotation,
  per-vertex radial noise of 5% of the radius, flattened into one mesh.
  410 spheres give 2,099,200 triangles (the 2M-triangle workload; 4096
  spheres would exceed the 2^24 per-mesh leaf-offset cap of R/qbvh.rs:15-18).
* ``city``: the wavefront-batching workload -- a 256 x 256 grid of bevelled
  (octagonal) towers with log-normal heights (mu 2, sigma 0.6), 400,000
  scattered rocks (noisy icosahedra, 20 triangles each) and a textured
  ground, about 9.4 M triangles in eleven flattened meshes; 2048 emissive
  street-lamp triangles in the light tree and a 2048 x 1024 procedural sky with
  an analytic sun of 0.27 degree radius; eight OpenPBR variants; four
  cameras of uneven resolution (1920x1080, 1280x720, 640x360, 3840x2160).
* ``primary_rays`` / ``bounce_rays`` / ``shadow_rays``: 1920x1080 pinhole
  rays, cosine-hemisphere bounce rays from primary hits, and shadow rays to
  uniform points on 64 random "emissive" triangles.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .scene import Mesh, Scene, make_rays

RAY_EPSILON = np.float32(1.0e-4)  # R/integrator/mod.rs:49


# --------------------------------------------------------------------------- geometry
def icosphere(level: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere by repeated 4-way midpoint subdivision (float64, int64)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


@dataclass
class SphereField:
    mesh: Mesh
    centers: np.ndarray
    radii: np.ndarray


def sphere_field(n_spheres: int = 410, level: int = 4, seed: int = 0, radius=(1.0, 6.0), cube: float = 100.0,
                 noise: float = 0.05) -> SphereField:
    """The traversal microbenchmark mesh (one flattened mesh)."""
    rng = np.random.default_rng(seed)
    uv, uf = icosphere(level)
    radii = np.exp(rng.uniform(math.log(radius[0]), math.log(radius[1]), n_spheres))
    centers = rng.uniform(0.0, 1.0, (n_spheres, 3)) * (cube - 2 * radii[:, None]) + radii[:, None]
    rots = random_rotations(rng, n_spheres)
    disp = 1.0 + rng.uniform(-noise, noise, (n_spheres, len(uv)))
    local = uv[None, :, :] * disp[:, :, None]  # (n, V, 3)
    pos = np.einsum("nij,nvj->nvi", rots, local) * radii[:, None, None] + centers[:, None, :]
    idx = uf[None, :, :] + (np.arange(n_spheres) * len(uv))[:, None, None]
    mesh = Mesh(pos.reshape(-1, 3).astype(np.float32), idx.reshape(-1, 3).astype(np.uint32))
    return SphereField(mesh, centers, radii)


def single_mesh_scene(mesh: Mesh, build: bool = True, threads: int = 0) -> Scene:
    """A flattened scene: one mesh, one identity instance."""
    s = Scene()
    mi = s.add_mesh(mesh)
    s.add_instance(mi, 0, np.eye(4, dtype=np.float32))
    if build:
        s.build_qbvh(threads)
    return s


def _quad(p0, p1, p2, p3) -> tuple[np.ndarray, np.ndarray]:
    return np.array([p0, p1, p2, p3], dtype=np.float64), np.array([[0, 1, 2], [0, 2, 3]])


def _rot_y(deg: float) -> np.ndarray:
    a = math.radians(deg)
    m = np.eye(4)
    m[0, 0] = math.cos(a)
    m[0, 2] = math.sin(a)
    m[2, 0] = -math.sin(a)
    m[2, 2] = math.cos(a)
    return m


def _translate(x, y, z) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def _scale(x, y, z) -> np.ndarray:
    return np.diag([x, y, z, 1.0])


def flat_quads(quads) -> Mesh:
    """A mesh of planar quads, each with its own four vertices, the quad's
    normal (from its winding: (p1-p0) x (p2-p0)) and uvs (0,0),(1,0),(1,1),(0,1)."""
    pos, nrm, uv, idx = [], [], [], []
    for q in quads:
        q = np.asarray(q, np.float64)
        n = np.cross(q[1] - q[0], q[2] - q[0])
        n /= np.linalg.norm(n)
        b = len(pos)
        pos += list(q)
        nrm += [n] * 4
        uv += [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
        idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
    return Mesh(np.array(pos, np.float32), np.array(idx, np.uint32), normals=np.array(nrm, np.float32),
                uvs=np.array(uv, np.float32))


def unit_box() -> Mesh:
    """Box [-0.5,0.5]^3: 6 outward-facing quads, 12 triangles, flat normals."""
    h = 0.5
    quads = [
        [(-h, -h, -h), (-h, -h, h), (-h, h, h), (-h, h, -h)],  # -x
        [(h, -h, -h), (h, h, -h), (h, h, h), (h, -h, h)],  # +x
        [(-h, -h, -h), (h, -h, -h), (h, -h, h), (-h, -h, h)],  # -y
        [(-h, h, -h), (-h, h, h), (h, h, h), (h, h, -h)],  # +y
        [(-h, -h, -h), (-h, h, -h), (h, h, -h), (h, -h, -h)],  # -z
        [(-h, -h, h), (h, -h, h), (h, h, h), (-h, h, h)],  # +z
    ]
    return flat_quads(quads)


@dataclass
class CornellScene:
    scene: Scene
    materials: list  # tpt.material.OpenPbr per material index
    light_instance: int
    camera: dict


def cornell(build: bool = True, specular_weight: float = 0.0) -> CornellScene:
    """Cornell box of the plain-method configuration: 36 triangles in 6
    instances.  Materials: OpenPBR ``new(c).with_specular_weight(w)`` for
    white (0.8), red (0.8,0.1,0.1), green (0.1,0.8,0.1), with w = 0 (the
    configuration's "diffuse OpenPBR"; ``specular_weight=1`` gives the
    reference's OpenPbrMaterial::new defaults: roughness 0.3, IOR 1.5); the
    light is OpenPBR emission 15 with base and specular weight 0.  Walls face
    into the box."""
    from .material import OpenPbr

    s = Scene()
    white = s.add_mesh(flat_quads([
        [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)],  # floor y=0, normal +y
        [(0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)],  # ceiling y=1, normal -y
        [(0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)],  # back wall z=1, normal -z
    ]))
    red = s.add_mesh(flat_quads([[(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)]]))  # x=0, normal +x
    green = s.add_mesh(flat_quads([[(1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)]]))  # x=1, normal -x
    box = s.add_mesh(unit_box())
    light = s.add_mesh(flat_quads([[(0.4, 0.999, 0.4), (0.6, 0.999, 0.4), (0.6, 0.999, 0.6),
                                     (0.4, 0.999, 0.6)]]))  # normal -y
    eye4 = np.eye(4)
    s.add_instance(white, 0, eye4)
    s.add_instance(red, 1, eye4)
    s.add_instance(green, 2, eye4)
    short = _translate(0.68, 0.15, 0.36) @ _rot_y(-18.0) @ _scale(0.3, 0.3, 0.3)
    tall = _translate(0.33, 0.3, 0.64) @ _rot_y(15.0) @ _scale(0.3, 0.6, 0.3)
    s.add_instance(box, 0, short)
    s.add_instance(box, 0, tall)
    li = s.add_instance(light, 3, eye4)
    if build:
        s.build_qbvh()
    diffuse = lambda c: OpenPbr.new(c).with_(specular_weight=float(specular_weight))
    mats = [diffuse((0.8, 0.8, 0.8)), diffuse((0.8, 0.1, 0.1)), diffuse((0.1, 0.8, 0.1)),
            OpenPbr.new((0.0, 0.0, 0.0)).with_(base_weight=0.0, specular_weight=0.0, emission_luminance=15.0,
                                               emission_color=(1.0, 1.0, 1.0))]
    for m in mats:
        s.add_material(m)
    cam = dict(eye=(0.5, 0.5, -1.4), look_at=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), fov_y_deg=40.0)
    return CornellScene(s, mats, li, cam)


def cornell_lambert(build: bool = True) -> CornellScene:
    """The same Cornell geometry with the reference's NormalizedLambert walls
    (white 0.8, red, green) and an Emissive light of strength 15: these
    surfaces build a diffuse light-tree query, so next-event estimation
    through the light tree is active (it is not for OpenPBR surfaces)."""
    from .material import Emissive, NormalizedLambert

    cs = cornell(build=build)
    mats = [NormalizedLambert((0.8, 0.8, 0.8)), NormalizedLambert((0.8, 0.1, 0.1)),
            NormalizedLambert((0.1, 0.8, 0.1)), Emissive((1.0, 1.0, 1.0), 15.0)]
    cs.scene.materials[:] = mats
    return CornellScene(cs.scene, mats, cs.light_instance, cs.camera)


def procedural_sky(width: int = 2048, height: int = 1024, sun_elevation_deg: float = 35.0,
                   sun_azimuth_deg: float = 40.0, sun_radius_deg: float = 0.27, sun_radiance: float = 2.0e4,
                   zenith=(0.25, 0.45, 0.9), horizon=(0.9, 0.95, 1.0), ground=(0.25, 0.22, 0.2)) -> np.ndarray:
    """Lat-long HDR sky (height, width, 3) in the environment light's
    convention (v = 0 is +Y, phi = atan2(x, z)): a zenith-to-horizon gradient,
    a dim ground below the horizon and an analytic sun disk of the given
    angular radius.  The disk's texels get the sun radiance weighted by the
    fraction of the texel the disk covers (supersampled 4x4)."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * np.pi
    y = np.cos(theta)
    t = np.clip(y, 0.0, 1.0) ** 0.6
    sky = (1 - t)[:, None] * np.asarray(horizon) + t[:, None] * np.asarray(zenith)
    img = np.repeat(sky[:, None, :], width, axis=1)
    img[y < 0] = np.asarray(ground)
    el, az, rad = np.radians(sun_elevation_deg), np.radians(sun_azimuth_deg), np.radians(sun_radius_deg)
    sd = np.array([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)])
    # texels near the sun: supersample coverage
    th_s = np.arccos(sd[1])
    ph_s = np.arctan2(sd[0], sd[2]) % (2 * np.pi)
    j0, j1 = int(max(0, (th_s - 3 * rad) / np.pi * height)), int(min(height, (th_s + 3 * rad) / np.pi * height + 1))
    for j in range(j0, j1):
        for i in range(width):
            ph_c = (i + 0.5) / width * 2 * np.pi
            if abs((ph_c - ph_s + np.pi) % (2 * np.pi) - np.pi) * np.sin(max(th_s, 1e-3)) > 4 * rad + 2 * np.pi / width:
                continue
            ss = (np.arange(4) + 0.5) / 4
            tt = (j + ss[:, None]) / height * np.pi
            pp = (i + ss[None, :]) / width * 2 * np.pi
            d = np.stack([np.sin(tt) * np.sin(pp), np.cos(tt) * np.ones_like(pp), np.sin(tt) * np.cos(pp)], -1)
            cover = np.mean(d @ sd >= np.cos(rad))
            if cover > 0:
                img[j, i] = img[j, i] * (1 - cover) + sun_radiance * cover
    return img.astype(np.float32)


def outdoor(materials: str = "openpbr", sky_size=(256, 128), build: bool = True, sun=None, sky: bool = True) \
        -> CornellScene:
    """The Cornell boxes on a 20 x 20 floor under the procedural sky (no
    walls, no area light): light comes from the environment only, so
    environment NEE is what lights OpenPBR surfaces here.  sun: optional
    directional lights [(direction of travel, color, intensity)]; sky=False
    drops the environment."""
    from .material import NormalizedLambert, OpenPbr
    from .scene import Environment

    s = Scene()
    floor = s.add_mesh(flat_quads([[(-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10)]]))
    box = s.add_mesh(unit_box())
    s.add_instance(floor, 0, np.eye(4))
    s.add_instance(box, 1, _translate(0.68, 0.15, 0.36) @ _rot_y(-18.0) @ _scale(0.3, 0.3, 0.3))
    s.add_instance(box, 2, _translate(0.33, 0.3, 0.64) @ _rot_y(15.0) @ _scale(0.3, 0.6, 0.3))
    if build:
        s.build_qbvh()
    cols = [(0.8, 0.8, 0.8), (0.8, 0.1, 0.1), (0.1, 0.8, 0.1)]
    if materials == "openpbr":
        mats = [OpenPbr.new(c).with_(specular_weight=0.0) for c in cols]
    else:
        mats = [NormalizedLambert(c) for c in cols]
    for m in mats:
        s.add_material(m)
    if sky:
        s.set_environment_light(Environment(procedural_sky(*sky_size), 1.0, 0.3))
    from .scene import DirectionalLight
    for d, c, e in (sun or []):
        s.add_directional_light(DirectionalLight.new(d, c, e))
    cam = dict(eye=(0.5, 0.9, -1.6), look_at=(0.5, 0.3, 0.5), up=(0.0, 1.0, 0.0), fov_y_deg=40.0)
    return CornellScene(s, mats, -1, cam)


# --------------------------------------------------------------------------- rays
def primary_rays(width: int, height: int, eye, look_at, up, fov_y_deg: float) -> np.ndarray:
    """Pixel-centre pinhole rays (direction_for_sample, R/scene/camera.rs:80-87), row-major."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(look_at, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    iup = np.cross(right, fwd)
    iup /= np.linalg.norm(iup)
    th = math.tan(0.5 * math.radians(fov_y_deg))
    aspect = width / height
    px = np.arange(width, dtype=np.float64) + 0.5
    py = np.arange(height, dtype=np.float64) + 0.5
    sx = ((2.0 * px / width) - 1.0) * aspect * th
    sy = (1.0 - (2.0 * py / height)) * th
    d = fwd[None, None, :] + sx[None, :, None] * right[None, None, :] + sy[:, None, None] * iup[None, None, :]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    d = d.reshape(-1, 3)
    o = np.broadcast_to(eye, d.shape)
    return make_rays(o.astype(np.float32), d.astype(np.float32))


def triangle_normals(mesh: Mesh, tri: np.ndarray) -> np.ndarray:
    v = mesh.positions[mesh.indices[tri]].astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def spawn_origins(p: np.ndarray, ng: np.ndarray, wi: np.ndarray) -> np.ndarray:
    """offset_ray_origin (R/integrator/mod.rs:76-85) in f32."""
    p = p.astype(np.float32)
    ng = ng.astype(np.float32)
    wi = wi.astype(np.float32)
    dot = wi[:, 0] * ng[:, 0] + wi[:, 1] * ng[:, 1] + wi[:, 2] * ng[:, 2]
    off = np.where((dot >= 0)[:, None], ng, -ng)
    return p + RAY_EPSILON * off


def _hit_points(rays: np.ndarray, hits: np.ndarray, mesh: Mesh):
    ok = np.nonzero(hits["triangle"] != 0xFFFFFFFF)[0]
    if len(ok) == 0:
        raise ValueError("no primary ray hits the scene")
    o = rays["o"][ok]
    d = rays["d"][ok]
    t = hits["t"][ok].astype(np.float32)
    p = o + t[:, None] * d  # Ray::at
    ng = triangle_normals(mesh, hits["triangle"][ok])
    ng = np.where((np.sum(ng * d, axis=1) > 0)[:, None], -ng, ng)  # face the incoming ray
    return p, ng


def _onb(n: np.ndarray):
    a = np.where((np.abs(n[:, 0]) > 0.9)[:, None], np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0]))
    t = np.cross(n, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    return t, b


def bounce_rays(rays: np.ndarray, hits: np.ndarray, mesh: Mesh, n: int, seed: int = 1) -> np.ndarray:
    """n cosine-hemisphere rays about the (ray-facing) geometric normal at primary hits."""
    rng = np.random.default_rng(seed)
    p, ng = _hit_points(rays, hits, mesh)
    sel = np.arange(n) % len(p)
    p, ng = p[sel], ng[sel]
    u1, u2 = rng.uniform(size=n), rng.uniform(size=n)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    x, y = r * np.cos(phi), r * np.sin(phi)
    z = np.sqrt(np.maximum(0.0, 1 - u1))
    t, b = _onb(ng)
    wi = x[:, None] * t + y[:, None] * b + z[:, None] * ng
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    return make_rays(spawn_origins(p, ng, wi), wi.astype(np.float32))


@dataclass
class ShadowSet:
    rays: np.ndarray
    target: np.ndarray  # (n, 2) uint32: (instance, triangle)
    lights: np.ndarray  # triangle ids of the emissive set


def shadow_rays(rays: np.ndarray, hits: np.ndarray, mesh: Mesh, n: int, n_lights: int = 64,
                seed: int = 2) -> ShadowSet:
    """n shadow rays from primary hits to uniform points (sample_triangle_point,
    R/light/area.rs) on n_lights random triangles, targeting that triangle."""
    rng = np.random.default_rng(seed)
    p, ng = _hit_points(rays, hits, mesh)
    sel = np.arange(n) % len(p)
    p, ng = p[sel], ng[sel]
    lights = rng.choice(mesh.triangle_count(), size=n_lights, replace=False).astype(np.uint32)
    k = rng.integers(0, n_lights, size=n)
    tri = lights[k]
    v = mesh.positions[mesh.indices[tri]].astype(np.float64)
    u1, u2 = rng.uniform(size=n), rng.uniform(size=n)
    su = np.sqrt(u1)
    q = (1 - su)[:, None] * v[:, 0] + (u2 * su)[:, None] * v[:, 1] + ((1 - u2) * su)[:, None] * v[:, 2]
    wi = q - p
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    r = make_rays(spawn_origins(p, ng, wi), wi.astype(np.float32))
    target = np.stack([np.zeros(n, np.uint32), tri], 1).astype(np.uint32)
    return ShadowSet(r, target, lights)


# canonical camera of the traversal microbenchmark: outside the 100-unit cube
TRAVERSAL_CAMERA = dict(eye=(50.0, 50.0, -140.0), look_at=(50.0, 50.0, 50.0), up=(0.0, 1.0, 0.0), fov_y_deg=25.0)


def cornell_textured(build: bool = True) -> CornellScene:
    """The Cornell box with every texture path of the backend in use: a
    tiled RGB checkerboard (64 x 64, uv x 4: wrap addressing) as the floor's
    base colour with a scalar roughness map, a non-power-of-two RGB gradient
    (48 x 40) as the back wall's base colour with a metalness map, fuzz weight
    and roughness maps on the ceiling, and an RGB emission texture (8 x 8) on
    the light.  The short box is a Mirror (delta bounces carry specular ray
    differentials) and the tall box NormalizedLambert (next-event estimation
    through the light tree, whose flux reads the emission texture's maximum).
    Materials: 0 floor, 1 back wall, 2 ceiling, 3 red, 4 green, 5 mirror,
    6 Lambert, 7 light."""
    from .material import Mirror, NormalizedLambert, OpenPbr
    from .texture import Texture, checkerboard

    s = Scene()
    floor_m = flat_quads([[(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)]])
    floor_m.uvs = floor_m.uvs * np.float32(4.0)
    back_m = flat_quads([[(0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)]])
    ceil_m = flat_quads([[(0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]])
    red = flat_quads([[(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)]])
    green = flat_quads([[(1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)]])
    light = flat_quads([[(0.4, 0.999, 0.4), (0.6, 0.999, 0.4), (0.6, 0.999, 0.6), (0.4, 0.999, 0.6)]])
    meshes = [s.add_mesh(m) for m in (floor_m, back_m, ceil_m, red, green)]
    box = s.add_mesh(unit_box())
    lm = s.add_mesh(light)
    eye4 = np.eye(4)
    for mi, mesh in enumerate(meshes):
        s.add_instance(mesh, mi, eye4)
    s.add_instance(box, 5, _translate(0.68, 0.15, 0.36) @ _rot_y(-18.0) @ _scale(0.3, 0.3, 0.3))
    s.add_instance(box, 6, _translate(0.33, 0.3, 0.64) @ _rot_y(15.0) @ _scale(0.3, 0.6, 0.3))
    li = s.add_instance(lm, 7, eye4)
    rng = np.random.default_rng(11)
    t_checker = s.add_texture(checkerboard(64, 64, 8, (0.9, 0.85, 0.8), (0.15, 0.2, 0.3)))
    t_rough = s.add_texture(Texture(rng.uniform(0.2, 1.0, (32, 32)).astype(np.float32)))
    gy, gx = np.mgrid[0:40, 0:48].astype(np.float32)
    grad = np.stack([gx / 47.0, gy / 39.0, 0.5 + 0.5 * np.sin(gx * 0.3) * np.cos(gy * 0.2)], -1) * 0.8 + 0.1
    t_grad = s.add_texture(Texture(grad.astype(np.float32)))
    t_metal = s.add_texture(Texture((np.add.outer(np.arange(16), np.arange(16)) % 5 == 0).astype(np.float32)))
    t_emit = s.add_texture(Texture((0.8 * np.stack([np.linspace(0.6, 1.0, 64).reshape(8, 8), np.full((8, 8), 0.9),
                                                    np.linspace(1.0, 0.5, 64).reshape(8, 8)], -1)).astype(np.float32)))
    if build:
        s.build_qbvh()
    mats = [
        OpenPbr.new((1.0, 1.0, 1.0)).with_(base_color_texture=t_checker, specular_roughness=0.8,
                                           specular_roughness_texture=t_rough),
        OpenPbr.new((1.0, 1.0, 1.0)).with_(base_color_texture=t_grad, base_metalness=0.9, specular_roughness=0.25,
                                           base_metalness_texture=t_metal),
        OpenPbr.new((0.8, 0.8, 0.8)).with_(specular_weight=0.0, fuzz_weight=1.0, fuzz_roughness=0.9,
                                           fuzz_color=(0.9, 0.7, 0.6), fuzz_weight_texture=t_rough,
                                           fuzz_roughness_texture=t_metal),
        OpenPbr.new((0.8, 0.1, 0.1)).with_(specular_weight=0.0),
        OpenPbr.new((0.1, 0.8, 0.1)).with_(specular_weight=0.0),
        Mirror((0.9, 0.9, 0.9)),
        NormalizedLambert((0.7, 0.7, 0.7)),
        OpenPbr.new((0.0, 0.0, 0.0)).with_(base_weight=0.0, specular_weight=0.0, emission_luminance=15.0,
                                           emission_color=(1.0, 1.0, 1.0), emission_color_texture=t_emit),
    ]
    for m in mats:
        s.add_material(m)
    cam = dict(eye=(0.5, 0.5, -1.4), look_at=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), fov_y_deg=40.0)
    return CornellScene(s, mats, li, cam)


def cornell_glass(build: bool = True, textured: bool = False, light_size: float = 0.2,
                  light_radiance: float = 15.0) -> CornellScene:
    """The Cornell box with OpenPBR transmission: the short box is smooth
    glass (IOR 1.5: delta refraction and reflection, total internal
    reflection inside), the tall box rough tinted glass (roughness 0.35,
    IOR 1.33, transmission colour (0.9, 0.75, 0.5): the GGX BTDF with its
    multiple-scattering lobe), and a thin-walled frosted pane (roughness
    0.3, transmission weight 0.9 over a white base) hangs in front of the
    back wall.  The walls carry the reference's OpenPbrMaterial::new
    defaults (glossy); the light is OpenPBR emission 15.  textured: the
    floor's base colour is a tiled checkerboard (64 x 64, uv x 4), so the
    path tracers run their textured instantiation, whose ray differentials
    refract through the smooth glass.  Materials: 0 walls, 1 red, 2 green,
    3 smooth glass, 4 rough glass, 5 pane, 6 light, 7 floor.  light_size /
    light_radiance: the ceiling light's side and radiance (a large dim light
    keeps caustics through the glass from being rare fireflies, which the
    finite-difference checks of the adjoint need)."""
    from .material import OpenPbr
    from .texture import checkerboard

    s = Scene()
    floor_m = flat_quads([[(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)]])
    floor_m.uvs = floor_m.uvs * np.float32(4.0)
    walls = s.add_mesh(flat_quads([
        [(0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)],  # ceiling y=1, normal -y
        [(0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)],  # back wall z=1, normal -z
    ]))
    floor = s.add_mesh(floor_m)
    red = s.add_mesh(flat_quads([[(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)]]))
    green = s.add_mesh(flat_quads([[(1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)]]))
    box = s.add_mesh(unit_box())
    pane = s.add_mesh(flat_quads([[(0.55, 0.35, 0.9), (0.55, 0.8, 0.9), (0.95, 0.8, 0.9), (0.95, 0.35, 0.9)]]))
    lo, hi = 0.5 - 0.5 * light_size, 0.5 + 0.5 * light_size
    light = s.add_mesh(flat_quads([[(lo, 0.999, lo), (hi, 0.999, lo), (hi, 0.999, hi), (lo, 0.999, hi)]]))
    eye4 = np.eye(4)
    s.add_instance(walls, 0, eye4)
    s.add_instance(floor, 7, eye4)
    s.add_instance(red, 1, eye4)
    s.add_instance(green, 2, eye4)
    s.add_instance(box, 3, _translate(0.68, 0.15, 0.36) @ _rot_y(-18.0) @ _scale(0.3, 0.3, 0.3))
    s.add_instance(box, 4, _translate(0.33, 0.3, 0.64) @ _rot_y(15.0) @ _scale(0.3, 0.6, 0.3))
    s.add_instance(pane, 5, eye4)
    li = s.add_instance(light, 6, eye4)
    tex = None
    if textured:
        tex = s.add_texture(checkerboard(64, 64, 8, (0.9, 0.85, 0.8), (0.15, 0.2, 0.3)))
    if build:
        s.build_qbvh()
    mats = [
        OpenPbr.new((0.8, 0.8, 0.8)),
        OpenPbr.new((0.8, 0.1, 0.1)),
        OpenPbr.new((0.1, 0.8, 0.1)),
        OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=0.0, specular_ior=1.5),
        OpenPbr(base_weight=0.0, transmission_weight=1.0, specular_roughness=0.35, specular_ior=1.33,
                transmission_color=(0.9, 0.75, 0.5)),
        OpenPbr(base_color=(0.9, 0.9, 0.9), transmission_weight=0.9, specular_roughness=0.3, thin_walled=True),
        OpenPbr.new((0.0, 0.0, 0.0)).with_(base_weight=0.0, specular_weight=0.0,
                                           emission_luminance=float(light_radiance), emission_color=(1.0, 1.0, 1.0)),
        OpenPbr.new((1.0, 1.0, 1.0) if textured else (0.8, 0.8, 0.8)).with_(base_color_texture=tex),
    ]
    for m in mats:
        s.add_material(m)
    cam = dict(eye=(0.5, 0.5, -1.4), look_at=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), fov_y_deg=40.0)
    return CornellScene(s, mats, li, cam)


def cornell_film(build: bool = True, textured: bool = False, light_size: float = 0.2,
                 light_radiance: float = 15.0) -> CornellScene:
    """The Cornell box with OpenPBR thin films (R/bsdf/thin_film.rs: two-beam
    interference over the base, per polarisation): the tall box is rough
    metal (roughness 0.3, a 0.45 um film of IOR 1.45 over a copper-like base
    colour), the short box a smooth coated dielectric (coat 0.6, a 0.3 um
    film of IOR 1.3 under the coat: the delta lobes' Fresnel carries the
    film), and a thin-walled frosted pane with a 0.6 um film (the slab's
    internal reflections see the film from inside) hangs in front of the
    back wall.  Walls, light and floor as cornell_glass.  Materials: 0 walls,
    1 red, 2 green, 3 coated film box, 4 metal film box, 5 film pane, 6 light,
    7 floor."""
    cs = cornell_glass(build=build, textured=textured, light_size=light_size, light_radiance=light_radiance)
    from .material import OpenPbr
    mats = list(cs.materials)
    mats[3] = OpenPbr.new((0.7, 0.75, 0.8)).with_(specular_roughness=0.0, coat_weight=0.6, thin_film_weight=1.0,
                                                  thin_film_thickness=0.3, thin_film_ior=1.3)
    mats[4] = OpenPbr.new((0.95, 0.64, 0.54)).with_(base_metalness=1.0, specular_roughness=0.3,
                                                    thin_film_weight=1.0, thin_film_thickness=0.45,
                                                    thin_film_ior=1.45)
    mats[5] = mats[5].with_(thin_film_weight=0.8, thin_film_thickness=0.6, thin_film_ior=1.5)
    cs.scene.materials.clear()
    for m in mats:
        cs.scene.add_material(m)
    return CornellScene(cs.scene, mats, cs.light_instance, cs.camera)


def cornell_inverse(build: bool = True, seed: int = 0, tex_size: int = 512, glossy: bool = False) -> CornellScene:
    """The inverse-rendering Cornell box: the Cornell geometry split into 16
    OpenPBR materials (five walls, the six faces of the short box, five faces
    of the tall box) with seeded diffuse base colours, a tex_size² RGB
    base-colour texture (seeded smooth noise) on the back wall, and the
    OpenPBR emitter as material 16.  Every material is affine in its base
    colour (metalness 0, no coat), so the adjoint differentiates all 16 base
    colours and the texels.  glossy=True: the inverse-rendering
    configuration's parameter set instead -- every material also draws a
    specular_roughness in [0.2, 0.8] and a metalness in [0, 1] from the seed
    (specular weight 1), all three learnable (tpt_render_adjoint_params)."""
    from .material import OpenPbr
    from .texture import Texture

    rng = np.random.default_rng(seed)
    s = Scene()
    walls = [
        [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)],  # floor
        [(0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)],  # ceiling
        [(0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)],  # back
        [(0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)],  # left
        [(1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)],  # right
    ]
    mi = 0
    for q in walls:
        s.add_instance(s.add_mesh(flat_quads([q])), mi, np.eye(4))
        mi += 1
    box = unit_box()
    faces = [flat_quads([box.positions[4 * f:4 * f + 4].tolist()]) for f in range(6)]
    short = _translate(0.68, 0.15, 0.36) @ _rot_y(-18.0) @ _scale(0.3, 0.3, 0.3)
    tall = _translate(0.33, 0.3, 0.64) @ _rot_y(15.0) @ _scale(0.3, 0.6, 0.3)
    for f in range(6):
        s.add_instance(s.add_mesh(faces[f]), mi, short)
        mi += 1
    for f in (0, 1, 3, 4, 5):  # the tall box's bottom face (-y) is hidden
        s.add_instance(s.add_mesh(faces[f]), mi, tall)
        mi += 1
    li = s.add_instance(s.add_mesh(flat_quads([[(0.4, 0.999, 0.4), (0.6, 0.999, 0.4), (0.6, 0.999, 0.6),
                                                 (0.4, 0.999, 0.6)]])), mi, np.eye(4))
    # smooth seeded noise: a sum of a few random sinusoids per channel
    y, x = np.mgrid[0:tex_size, 0:tex_size].astype(np.float64) / tex_size
    px = np.zeros((tex_size, tex_size, 3))
    for c in range(3):
        for _ in range(4):
            fx, fy, ph = rng.uniform(1, 8), rng.uniform(1, 8), rng.uniform(0, 2 * np.pi)
            px[..., c] += np.sin(2 * np.pi * (fx * x + fy * y) + ph)
    tex = s.add_texture(Texture((0.55 + 0.1 * px).clip(0.05, 1.0).astype(np.float32)))
    if build:
        s.build_qbvh()
    mats = [OpenPbr.new(tuple(rng.uniform(0.2, 0.9, 3).tolist())).with_(specular_weight=0.0) for _ in range(mi)]
    if glossy:
        g = np.random.default_rng(seed + 1)
        mats = [m.with_(specular_weight=1.0, specular_roughness=float(g.uniform(0.2, 0.8)),
                        base_metalness=float(g.uniform(0.0, 1.0))) for m in mats]
    mats[2] = mats[2].with_(base_color_texture=tex)
    mats.append(OpenPbr.new((0.0, 0.0, 0.0)).with_(base_weight=0.0, specular_weight=0.0, emission_luminance=15.0,
                                                   emission_color=(1.0, 1.0, 1.0)))
    for m in mats:
        s.add_material(m)
    cam = dict(eye=(0.5, 0.5, -1.4), look_at=(0.5, 0.5, 0.5), up=(0.0, 1.0, 0.0), fov_y_deg=40.0)
    return CornellScene(s, mats, li, cam)


# --------------------------------------------------------------------------- city
@dataclass
class CityScene:
    scene: Scene
    materials: list
    cameras: list  # (camera kwargs, width, height) per frame
    emissive_triangles: int


CITY_FRAMES = ((1920, 1080), (1280, 720), (640, 360), (3840, 2160))


def _octagon_towers(rng, centers, half, chamfer, heights):
    """Bevelled towers: octagonal prisms (the footprint square with its
    vertical edges chamfered), 8 side quads + a 6-triangle roof, open at the
    ground: 22 triangles and 16 vertices per tower."""
    n = len(centers)
    a, c = half[:, None], chamfer[:, None]
    # footprint octagon, counter-clockwise seen from above (+y)
    ox = np.stack([a, a - c, -(a - c), -a, -a, -(a - c), a - c, a], 1)[..., 0]
    oz = np.stack([a - c, a, a, a - c, -(a - c), -a, -a, -(a - c)], 1)[..., 0]
    x = ox + centers[:, 0:1]
    z = oz + centers[:, 1:2]
    y0 = np.zeros_like(x)
    y1 = np.broadcast_to(heights[:, None], x.shape)
    bottom = np.stack([x, y0, z], -1)  # (n, 8, 3)
    top = np.stack([x, y1, z], -1)
    pos = np.concatenate([bottom, top], 1)  # vertices 0..7 bottom, 8..15 top
    k = np.arange(8)
    k1 = (k + 1) % 8
    # side quads, wound so that (p1 - p0) x (p2 - p0) points out of the tower
    sides = np.concatenate([np.stack([k, k + 8, k1 + 8], 1), np.stack([k, k1 + 8, k1], 1)], 0)
    roof = np.array([[8, 8 + j + 1, 8 + j] for j in range(1, 7)])
    tri = np.concatenate([sides, roof], 0)  # (22, 3)
    idx = tri[None, :, :] + (np.arange(n) * 16)[:, None, None]
    return pos.reshape(-1, 3), idx.reshape(-1, 3)


def _rocks(rng, n, extent, size=(0.15, 0.8), noise=0.25):
    """n noisy icosahedra (level-0 icospheres, 20 triangles), random
    rotation, log-uniform size, resting on the ground anywhere in the city."""
    uv, uf = icosphere(0)
    r = np.exp(rng.uniform(math.log(size[0]), math.log(size[1]), n))
    cxz = rng.uniform(-extent, extent, (n, 2))
    rots = random_rotations(rng, n)
    disp = 1.0 + rng.uniform(-noise, noise, (n, len(uv)))
    local = uv[None, :, :] * disp[:, :, None]
    pos = np.einsum("nij,nvj->nvi", rots, local) * r[:, None, None]
    pos[:, :, 0] += cxz[:, 0:1]
    pos[:, :, 1] += 0.5 * r[:, None]
    pos[:, :, 2] += cxz[:, 1:2]
    idx = uf[None, :, :] + (np.arange(n) * len(uv))[:, None, None]
    return pos.reshape(-1, 3), idx.reshape(-1, 3)


def city_texture(size: int = 512) -> np.ndarray:
    """A facade / paving texture (size x size RGB): a grid of lit and unlit
    window cells with mortar lines and seeded per-cell tints."""
    rng = np.random.default_rng(7)
    cells = 16
    y, x = np.mgrid[0:size, 0:size]
    cx, cy = x * cells // size, y * cells // size
    fx, fy = (x * cells % size) / size, (y * cells % size) / size
    tint = rng.uniform(0.3, 0.9, (cells, cells, 3))
    img = tint[cy, cx]
    mortar = (fx < 0.08) | (fy < 0.08)
    img[mortar] = (0.12, 0.12, 0.13)
    return np.ascontiguousarray(img.astype(np.float32))


def city(grid: int = 256, rocks: int = 400_000, seed: int = 0, cell: float = 4.0, lights: int = 2048,
         sky_size=(2048, 1024), tex_size: int = 512, build: bool = True, threads: int = 0) -> CityScene:
    """The seeded procedural city.  Towers on a grid x grid lattice of
    `cell`-unit plots (footprint half-size 0.25-0.42 cell, chamfer 15-30% of
    it, height exp(N(2, 0.6))), each tower given one of the eight variants;
    `rocks` sheen rocks; `lights` emissive triangles as lights / 2 street
    lamps (downward-facing 0.6 x 0.6 quads 4-6 units above random plot
    corners, luminance 20); a textured ground.  Flattened meshes, one per tower variant, plus the rocks, the
    ground and the emitters, each under an identity instance.

    Materials: 0 diffuse, 1-3 metal at specular_roughness 0.05 / 0.3 / 0.8,
    4 coat 0.5, 5 transmission with IOR 1.5 (smooth glass: transmission
    weight 1, no diffuse base), 6 sheen (fuzz), 7 base-colour textured,
    8 the lamp emitter.  Every surface is wound outward ((p1 - p0) x (p2 - p0)
    points out of the tower, up from the ground, down from a lamp)."""
    from .material import OpenPbr
    from .scene import Environment
    from .texture import Texture

    rng = np.random.default_rng(seed)
    n = grid * grid
    extent = 0.5 * grid * cell
    gi, gj = np.meshgrid(np.arange(grid), np.arange(grid), indexing="ij")
    centers = np.stack([(gi.ravel() + 0.5) * cell - extent, (gj.ravel() + 0.5) * cell - extent], 1)
    half = rng.uniform(0.25, 0.42, n) * cell
    chamfer = rng.uniform(0.15, 0.30, n) * half
    heights = np.exp(rng.normal(2.0, 0.6, n))
    variant = rng.integers(0, 8, n)
    s = Scene()
    tex = s.add_texture(Texture(city_texture(tex_size)))

    def add(pos, idx, mat, uvs=None):
        mesh = Mesh(pos.astype(np.float32), idx.astype(np.uint32), uvs=None if uvs is None else uvs.astype(np.float32))
        s.add_instance(s.add_mesh(mesh), mat, np.eye(4))

    for v in range(8):
        sel = np.nonzero(variant == v)[0]
        if len(sel) == 0:
            continue
        pos, idx = _octagon_towers(rng, centers[sel], half[sel], chamfer[sel], heights[sel])
        uvs = np.stack([(pos[:, 0] + pos[:, 2]) * 0.05, pos[:, 1] * 0.05], 1) if v == 7 else None
        add(pos, idx, v, uvs)
    if rocks:  # sheen rocks
        rp, ri = _rocks(rng, rocks, extent)
        add(rp, ri, 6)
    # ground: textured paving
    e = extent * 1.05
    gp = np.array([[-e, 0, -e], [-e, 0, e], [e, 0, e], [e, 0, -e]], np.float64)
    add(gp, np.array([[0, 1, 2], [0, 2, 3]]), 7, gp[:, [0, 2]] * 0.02)
    # street lamps: downward-facing emissive quads over random plot corners
    corners = (grid - 1) ** 2
    quads = min(lights // 2, corners)
    pick = rng.choice(corners, quads, replace=False)
    c = np.stack([(pick // (grid - 1) + 1) * cell - extent, (pick % (grid - 1) + 1) * cell - extent], 1)
    h = rng.uniform(4.0, 6.0, quads)
    q = 0.3
    lp = np.stack([np.stack([c[:, 0] - q, h, c[:, 1] - q], 1), np.stack([c[:, 0] + q, h, c[:, 1] - q], 1),
                   np.stack([c[:, 0] + q, h, c[:, 1] + q], 1), np.stack([c[:, 0] - q, h, c[:, 1] + q], 1)], 1)
    li = np.array([[0, 1, 2], [0, 2, 3]])[None] + (np.arange(quads) * 4)[:, None, None]
    add(lp.reshape(-1, 3), li.reshape(-1, 3), 8)
    if build:
        s.build_qbvh(threads)
    mats = [
        OpenPbr.new((0.7, 0.7, 0.68)).with_(specular_weight=0.0),
        OpenPbr.new((0.95, 0.8, 0.55)).with_(base_metalness=1.0, specular_roughness=0.05),
        OpenPbr.new((0.9, 0.9, 0.92)).with_(base_metalness=1.0, specular_roughness=0.3),
        OpenPbr.new((0.75, 0.45, 0.35)).with_(base_metalness=1.0, specular_roughness=0.8),
        OpenPbr.new((0.55, 0.08, 0.06)).with_(coat_weight=0.5),
        OpenPbr.new((0.8, 0.8, 0.8)).with_(base_weight=0.0, transmission_weight=1.0, specular_ior=1.5,
                                           specular_roughness=0.0),
        OpenPbr.new((0.35, 0.33, 0.3)).with_(specular_weight=0.0, fuzz_weight=1.0, fuzz_roughness=0.5,
                                             fuzz_color=(0.8, 0.75, 0.6)),
        OpenPbr.new((1.0, 1.0, 1.0)).with_(base_color_texture=tex, specular_roughness=0.5),
        OpenPbr.new((0.0, 0.0, 0.0)).with_(base_weight=0.0, specular_weight=0.0, emission_luminance=20.0,
                                           emission_color=(1.0, 0.85, 0.6)),
    ]
    for m in mats:
        s.add_material(m)
    s.set_environment_light(Environment(procedural_sky(*sky_size), 1.0, 0.0))
    views = [
        dict(eye=(-0.55 * extent, 0.12 * extent, -0.55 * extent), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
             fov_y_deg=45.0),
        # street level on the street line x = 0
        dict(eye=(0.0, 1.7, -2.0), look_at=(6.0, 8.0, 60.0), up=(0.0, 1.0, 0.0), fov_y_deg=60.0),
        dict(eye=(0.3 * extent, 0.3 * extent, 0.1 * extent), look_at=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
             fov_y_deg=35.0),
        dict(eye=(0.0, 0.06 * extent, -0.9 * extent), look_at=(0.0, 10.0, 0.0), up=(0.0, 1.0, 0.0), fov_y_deg=50.0),
    ]
    cams = [(v, w, h) for v, (w, h) in zip(views, CITY_FRAMES)]
    return CityScene(s, mats, cams, 2 * quads)


# ------------------------------------------------------------ city, inverse rendering
def perlin_texture(size: int, seed: int, cells: int = 8, octaves: int = 4) -> np.ndarray:
    """A seeded RGB gradient-noise (Perlin) texture, size x size, values in
    [0.1, 0.9]: per channel, `octaves` octaves of tileable gradient noise on a
    cells x cells lattice doubling per octave, quintic fade."""
    rng = np.random.default_rng(seed)
    img = np.zeros((size, size, 3))
    y, x = np.mgrid[0:size, 0:size].astype(np.float64) / size
    for ch in range(3):
        acc = np.zeros((size, size))
        amp, n = 1.0, cells
        for _ in range(octaves):
            ang = rng.uniform(0.0, 2.0 * np.pi, (n, n))
            gx, gy = np.cos(ang), np.sin(ang)
            fx, fy = x * n, y * n
            ix, iy = np.floor(fx).astype(int), np.floor(fy).astype(int)
            tx, ty = fx - ix, fy - iy

            def dot(ox, oy):
                i, j = (ix + ox) % n, (iy + oy) % n
                return gx[j, i] * (tx - ox) + gy[j, i] * (ty - oy)

            fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)
            u, v = fade(tx), fade(ty)
            n0 = dot(0, 0) * (1 - u) + dot(1, 0) * u
            n1 = dot(0, 1) * (1 - u) + dot(1, 1) * u
            acc += amp * (n0 * (1 - v) + n1 * v)
            amp *= 0.5
            n *= 2
        img[..., ch] = acc
    img = (img - img.min()) / max(img.max() - img.min(), 1e-9)
    return np.ascontiguousarray((0.1 + 0.8 * img).astype(np.float32))


@dataclass
class CityInverseScene:
    scene: Scene
    materials: list
    views: list  # (camera kwargs, width, height) per view
    learnable: int  # the first `learnable` materials are OpenPBR with learnable parameters


def city_inverse(grid: int = 256, rocks: int = 400_000, seed: int = 0, cell: float = 4.0, lights: int = 2048,
                 materials: int = 64, tex_size: int = 1024, textures: int = 4, views: int = 32,
                 view_sizes=(256, 1024), sky_size=(2048, 1024), build: bool = True, threads: int = 0) \
        -> CityInverseScene:
    """The batched inverse-rendering city: the city's towers (as in city())
    split into materials - 2 seeded groups, the rocks and the ground, each an
    OpenPBR material with seeded base colour (uniform 0.2-0.9), specular
    roughness (0.2-0.8) and metalness (0-1), all learnable; `textures`
    seeded Perlin-noise RGB textures of tex_size^2 FP32 texels as the base
    colours of the first textures - 1 tower groups and of the ground; the
    street lamps (one more material, the emitter) and the sky.  `views`
    cameras at seeded positions over the city looking at seeded points near
    its centre, square images whose side is log-uniform in view_sizes."""
    from .material import OpenPbr
    from .scene import Environment
    from .texture import Texture

    rng = np.random.default_rng(seed)
    n = grid * grid
    extent = 0.5 * grid * cell
    gi, gj = np.meshgrid(np.arange(grid), np.arange(grid), indexing="ij")
    centers = np.stack([(gi.ravel() + 0.5) * cell - extent, (gj.ravel() + 0.5) * cell - extent], 1)
    half = rng.uniform(0.25, 0.42, n) * cell
    chamfer = rng.uniform(0.15, 0.30, n) * half
    heights = np.exp(rng.normal(2.0, 0.6, n))
    groups = materials - 2
    group = rng.integers(0, groups, n)
    s = Scene()
    tex = [s.add_texture(Texture(perlin_texture(tex_size, seed + 100 + t))) for t in range(textures)]

    def add(pos, idx, mat, uvs=None):
        mesh = Mesh(pos.astype(np.float32), idx.astype(np.uint32), uvs=None if uvs is None else uvs.astype(np.float32))
        s.add_instance(s.add_mesh(mesh), mat, np.eye(4))

    textured = {g: tex[g] for g in range(min(textures - 1, groups))}
    for g in range(groups):
        sel = np.nonzero(group == g)[0]
        if len(sel) == 0:
            continue
        pos, idx = _octagon_towers(rng, centers[sel], half[sel], chamfer[sel], heights[sel])
        uvs = np.stack([(pos[:, 0] + pos[:, 2]) * 0.05, pos[:, 1] * 0.05], 1) if g in textured else None
        add(pos, idx, g, uvs)
    if rocks:
        rp, ri = _rocks(rng, rocks, extent)
        add(rp, ri, groups)
    e = extent * 1.05
    gp = np.array([[-e, 0, -e], [-e, 0, e], [e, 0, e], [e, 0, -e]], np.float64)
    add(gp, np.array([[0, 1, 2], [0, 2, 3]]), groups + 1, gp[:, [0, 2]] * 0.02)
    corners = (grid - 1) ** 2
    quads = min(lights // 2, corners)
    pick = rng.choice(corners, quads, replace=False)
    c = np.stack([(pick // (grid - 1) + 1) * cell - extent, (pick % (grid - 1) + 1) * cell - extent], 1)
    h = rng.uniform(4.0, 6.0, quads)
    q = 0.3
    lp = np.stack([np.stack([c[:, 0] - q, h, c[:, 1] - q], 1), np.stack([c[:, 0] + q, h, c[:, 1] - q], 1),
                   np.stack([c[:, 0] + q, h, c[:, 1] + q], 1), np.stack([c[:, 0] - q, h, c[:, 1] + q], 1)], 1)
    li = np.array([[0, 1, 2], [0, 2, 3]])[None] + (np.arange(quads) * 4)[:, None, None]
    add(lp.reshape(-1, 3), li.reshape(-1, 3), materials)
    if build:
        s.build_qbvh(threads)
    mr = np.random.default_rng(seed + 1)
    mats = []
    for i in range(materials):
        m = OpenPbr.new(tuple(mr.uniform(0.2, 0.9, 3).tolist())).with_(
            specular_roughness=float(mr.uniform(0.2, 0.8)), base_metalness=float(mr.uniform(0.0, 1.0)))
        if i in textured:
            m = m.with_(base_color_texture=textured[i])
        mats.append(m)
    mats[groups + 1] = mats[groups + 1].with_(base_color_texture=tex[textures - 1])
    mats.append(OpenPbr.new((0.0, 0.0, 0.0)).with_(base_weight=0.0, specular_weight=0.0, emission_luminance=20.0,
                                                   emission_color=(1.0, 0.85, 0.6)))
    for m in mats:
        s.add_material(m)
    s.set_environment_light(Environment(procedural_sky(*sky_size), 1.0, 0.0))
    vr = np.random.default_rng(seed + 2)
    cams = []
    for _ in range(views):
        side = int(round(np.exp(vr.uniform(np.log(view_sizes[0]), np.log(view_sizes[1])))))
        r = vr.uniform(0.1, 0.6) * extent
        a = vr.uniform(0.0, 2.0 * np.pi)
        eye = (float(r * np.cos(a)), float(vr.uniform(3.0, 0.25 * extent)), float(r * np.sin(a)))
        look = (float(vr.uniform(-0.1, 0.1) * extent), 0.0, float(vr.uniform(-0.1, 0.1) * extent))
        cams.append((dict(eye=eye, look_at=look, up=(0.0, 1.0, 0.0), fov_y_deg=float(vr.uniform(40.0, 70.0))),
                     side, side))
    return CityInverseScene(s, mats, cams, materials)
